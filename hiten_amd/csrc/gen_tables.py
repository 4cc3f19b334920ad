"""Generate ``hh_tables.h``: Runge-Kutta tableaux and pow lookup tables.

Run from the repository root::

    python hiten_amd/csrc/gen_tables.py

Every constant is written as a C99/C++17 hexadecimal floating literal, so the
value compiled into the HIP kernels and into the CPU oracle is bit-for-bit the
double that Python computed here.

* DOP853 (Dormand-Prince 8(5,3)) coefficients are taken from SciPy's
  ``scipy.integrate._ivp.dop853_coefficients``.  hiten's
  ``integrators/coefficients/dop853.py`` holds the same decimal literals and
  builds ``E3 = B - bhat`` with the same double arithmetic, so these arrays are
  the ones the reference integrates with.
* The 13-stage 8th-order tableau is the Prince-Dormand RK8(7)13M high-order
  formula.  The reference (``integrators/coefficients/rk8.py:4-21``) writes each
  coefficient as ``p.0/q.0``; ``float(Fraction(p, q))`` is the same correctly
  rounded quotient.
* The classical 4-stage RK4 tableau and the 7-stage tableau the reference
  names RK6 (``integrators/coefficients/rk4.py``, ``rk6.py``: its rows and
  weights are the Dormand-Prince 5(4) pair's 5th-order formula) are written
  there as ``p.0/q.0`` quotients too, generated here from the same rationals.
* pow tables: ``log(1 + i/128)`` and ``2**(j/128)`` as double-double pairs and
  the rounded reciprocals ``1/(1 + i/128)``, evaluated with mpmath at 60
  significant digits.
"""

from fractions import Fraction as Fr
import os

import mpmath
import numpy as np
from scipy.integrate._ivp import dop853_coefficients as dop

HERE = os.path.dirname(os.path.abspath(__file__))
OUT = os.path.join(HERE, "hh_tables.h")

F = Fr
# Prince & Dormand RK8(7)13M, 8th-order weights.  Rows are stages 0..12.
RK8_A_ROWS = [
    [],
    [F(1, 18)],
    [F(1, 48), F(1, 16)],
    [F(1, 32), 0, F(3, 32)],
    [F(5, 16), 0, F(-75, 64), F(75, 64)],
    [F(3, 80), 0, 0, F(3, 16), F(3, 20)],
    [F(29443841, 614563906), 0, 0, F(77736538, 692538347), F(-28693883, 1125000000),
     F(23124283, 1800000000)],
    [F(16016141, 946692911), 0, 0, F(61564180, 158732637), F(22789713, 633445777),
     F(545815736, 2771057229), F(-180193667, 1043307555)],
    [F(39632708, 573591083), 0, 0, F(-433636366, 683701615), F(-421739975, 2616292301),
     F(100302831, 723423059), F(790204164, 839813087), F(800635310, 3783071287)],
    [F(246121993, 1340847787), 0, 0, F(-37695042795, 15268766246), F(-309121744, 1061227803),
     F(-12992083, 490766935), F(6005943493, 2108947869), F(393006217, 1396673457),
     F(123872331, 1001029789)],
    [F(-1028468189, 846180014), 0, 0, F(8478235783, 508512852), F(1311729495, 1432422823),
     F(-10304129995, 1701304382), F(-48777925059, 3047939560), F(15336726248, 1032824649),
     F(-45442868181, 3398467696), F(3065993473, 597172653)],
    [F(185892177, 718116043), 0, 0, F(-3185094517, 667107341), F(-477755414, 1098053517),
     F(-703635378, 230739211), F(5731566787, 1027545527), F(5232866602, 850066563),
     F(-4093664535, 808688257), F(3962137247, 1805957418), F(65686358, 487910083)],
    [F(403863854, 491063109), 0, 0, F(-5068492393, 434740067), F(-411421997, 543043805),
     F(652783627, 914296604), F(11173962825, 925320556), F(-13158990841, 6184727034),
     F(3936647629, 1978049680), F(-160528059, 685178525), F(248638103, 1413531060)],
]
RK8_B = [F(14005451, 335480064), 0, 0, 0, 0, F(-59238493, 1068277825),
         F(181606767, 758867731), F(561292985, 797845732), F(-1041891430, 1371343529),
         F(760417239, 1151165299), F(118820643, 751138087), F(-528747749, 2220607170),
         F(1, 4)]
RK8_C = [0, F(1, 18), F(1, 12), F(1, 8), F(5, 16), F(3, 8), F(59, 400), F(93, 200),
         F(5490023248, 9719169821), F(13, 20), F(1201146811, 1299019798), 1, 1]


# classical RK4
RK4_A_ROWS = [[], [F(1, 2)], [0, F(1, 2)], [0, 0, 1]]
RK4_B = [F(1, 6), F(1, 3), F(1, 3), F(1, 6)]
RK4_C = [0, F(1, 2), F(1, 2), 1]

# Dormand-Prince 5(4): the seven stages and the 5th-order weights
RK6_A_ROWS = [
    [],
    [F(1, 5)],
    [F(3, 40), F(9, 40)],
    [F(44, 45), F(-56, 15), F(32, 9)],
    [F(19372, 6561), F(-25360, 2187), F(64448, 6561), F(-212, 729)],
    [F(9017, 3168), F(-355, 33), F(46732, 5247), F(49, 176), F(-5103, 18656)],
    [F(35, 384), 0, F(500, 1113), F(125, 192), F(-2187, 6784), F(11, 84)],
]
RK6_B = [F(35, 384), 0, F(500, 1113), F(125, 192), F(-2187, 6784), F(11, 84), 0]
RK6_C = [0, F(1, 5), F(3, 10), F(4, 5), F(8, 9), 1, 1]


def tableau_arrays(rows, b, c):
    s = len(b)
    A = np.zeros((s, s))
    for i, row in enumerate(rows):
        for j, v in enumerate(row):
            A[i, j] = float(v)
    return A, np.array([float(v) for v in b]), np.array([float(v) for v in c])


def rk8_arrays():
    s = 13
    A = np.zeros((s, s))
    for i, row in enumerate(RK8_A_ROWS):
        for j, v in enumerate(row):
            A[i, j] = float(v)
    B = np.array([float(v) for v in RK8_B])
    C = np.array([float(v) for v in RK8_C])
    return A, B, C


def hexd(x: float) -> str:
    x = float(x)
    if x == 0.0:
        return "0x0p+0" if np.copysign(1.0, x) > 0 else "-0x0p+0"
    return float.hex(x)


def emit_array(name, arr, ctype="double"):
    arr = np.asarray(arr, dtype=np.float64)
    dims = "".join(f"[{d}]" for d in arr.shape)

    def row(vals):
        return ",\n     ".join(", ".join(hexd(v) for v in vals[i:i + 4]) for i in range(0, len(vals), 4))

    if arr.ndim == 1:
        body = row(list(arr))
    else:
        body = ",\n    ".join("{" + row(list(r)) + "}" for r in arr)
    return f"HH_TABLE {ctype} {name}{dims} = {{\n    {body}\n}};\n"


def dd(v):
    hi = float(v)
    lo = float(v - mpmath.mpf(hi))
    return hi, lo


def pow_tables():
    mpmath.mp.dps = 60
    log_hi, log_lo, exp_hi, exp_lo, invc = [], [], [], [], []
    for i in range(128):
        invc.append(float(1 / (1 + mpmath.mpf(i) / 128)))
        h, l = dd(mpmath.log(1 + mpmath.mpf(i) / 128))
        log_hi.append(h)
        log_lo.append(l)
    for j in range(128):
        h, l = dd(mpmath.power(2, mpmath.mpf(j) / 128))
        exp_hi.append(h)
        exp_lo.append(l)
    ln2 = mpmath.log(2)

    def split(v, bits):
        # round v to `bits` significant bits so that small integer multiples are exact
        m, e = mpmath.frexp(v)
        hi = mpmath.ldexp(mpmath.nint(mpmath.ldexp(m, bits)), e - bits)
        return float(hi), float(v - hi)

    ln2_hi, ln2_lo = split(ln2, 42)
    l128_hi, l128_lo = split(ln2 / 128, 36)
    inv_l128 = float(128 / ln2)
    return dict(log_hi=log_hi, log_lo=log_lo, exp_hi=exp_hi, exp_lo=exp_lo, invc=invc,
                ln2_hi=ln2_hi, ln2_lo=ln2_lo, l128_hi=l128_hi, l128_lo=l128_lo,
                inv_l128=inv_l128)


def main():
    A8, B8, C8 = rk8_arrays()
    pt = pow_tables()
    parts = []
    parts.append("""/* Generated by hiten_amd/csrc/gen_tables.py -- do not edit by hand.
 *
 * Runge-Kutta tableaux (RK4, RK6 and RK8 fixed step, DOP853 adaptive with its
 * dense-output extension) and the lookup tables of the double-double pow used
 * by the step-size controller.  Included by both the HIP kernels and the CPU
 * oracle, so both integrate with bit-identical coefficients.
 *
 * HH_TABLE must be defined by the includer: `static const` on the host,
 * `static __device__ const` in device code.
 */
#ifndef HH_TABLES_H
#define HH_TABLES_H
#ifndef HH_TABLE
#error "define HH_TABLE before including hh_tables.h"
#endif
""")
    for name, (rows, b, c) in (("RK4", (RK4_A_ROWS, RK4_B, RK4_C)),
                               ("RK6", (RK6_A_ROWS, RK6_B, RK6_C))):
        A, B, C = tableau_arrays(rows, b, c)
        parts.append(f"#define HH_{name}_STAGES {len(B)}\n")
        parts.append(emit_array(f"HH_{name}_A", A))
        parts.append(emit_array(f"HH_{name}_B", B))
        parts.append(emit_array(f"HH_{name}_C", C))
    parts.append("#define HH_RK8_STAGES 13\n")
    parts.append(emit_array("HH_RK8_A", A8))
    parts.append(emit_array("HH_RK8_B", B8))
    parts.append(emit_array("HH_RK8_C", C8))
    parts.append("#define HH_DOP853_STAGES 12\n#define HH_DOP853_STAGES_EXT 16\n#define HH_DOP853_INTERP_POWER 7\n")
    parts.append(emit_array("HH_DOP_A", dop.A))
    parts.append(emit_array("HH_DOP_B", dop.B))
    parts.append(emit_array("HH_DOP_C", dop.C))
    parts.append(emit_array("HH_DOP_E3", dop.E3))
    parts.append(emit_array("HH_DOP_E5", dop.E5))
    parts.append(emit_array("HH_DOP_D", dop.D))
    parts.append(emit_array("HH_POW_LOG_HI", pt["log_hi"]))
    parts.append(emit_array("HH_POW_LOG_LO", pt["log_lo"]))
    parts.append(emit_array("HH_POW_INVC", pt["invc"]))
    parts.append(emit_array("HH_POW_EXP_HI", pt["exp_hi"]))
    parts.append(emit_array("HH_POW_EXP_LO", pt["exp_lo"]))
    for k in ("ln2_hi", "ln2_lo", "l128_hi", "l128_lo", "inv_l128"):
        parts.append(f"#define HH_POW_{k.upper()} ({hexd(pt[k])})\n")
    parts.append("#endif /* HH_TABLES_H */\n")
    with open(OUT, "w") as fh:
        fh.write("".join(parts))
    print("wrote", OUT)


if __name__ == "__main__":
    main()
