"""Oracle (test infrastructure only): the 2x2 dense step of the corrector as
NumPy's LAPACK forms it, restated operation by operation.

The reference's _solve_delta_dense (corrector/backends/base.py:187-245)
calls numpy.linalg.cond and numpy.linalg.solve.  NumPy's solve is LAPACK
dgesv from the OpenBLAS it ships; for a 2x2 system OpenBLAS runs getf2
(partial pivoting on the first entry of largest magnitude, the multiplier
scaled by the pivot's reciprocal, the Schur complement a11 - l * a01 from
its gemv) and getrs with one right-hand side as two triangular solves
(trsv: axpy updates, which the x86-64 kernels fuse, and divisions by the
diagonal).  ``lu_solve2`` is that sequence; it is pinned against
numpy.linalg.solve on random systems by tests/test_correct_device_cpu.py.
Pivots below DBL_MIN (subnormal) are not pinned: NumPy's results there
match neither LAPACK branch as written (no Jacobian of the corrector comes
near them).

``cond2`` is the 2-norm condition number from the largest singular value
and an accurately rounded determinant.  NumPy forms it from dgesdd, whose
smallest singular value carries an error of about eps * sigma_max, so the
two agree to about eps * cond relative; the corrector only compares it with
1e8, where that is 2e-8: a decision can differ only for a condition number
within that of the threshold.
"""
from __future__ import annotations

import ctypes
import math

_libm = ctypes.CDLL("libm.so.6")
_libm.fma.restype = ctypes.c_double
_libm.fma.argtypes = [ctypes.c_double] * 3


def fma(a: float, b: float, c: float) -> float:
    """a * b + c with one rounding (the C library's fma)."""
    return _libm.fma(a, b, c)


DBL_MIN = 2.2250738585072014e-308


def lu_solve2(A, b):
    """numpy.linalg.solve(A, b) for 2x2 A (row-major 4-sequence) and b (2,);
    None where LAPACK reports an exactly singular U."""
    a00, a01, a10, a11 = (float(v) for v in A)
    b0, b1 = float(b[0]), float(b[1])
    if abs(a10) > abs(a00):
        a00, a01, a10, a11, b0, b1 = a10, a11, a00, a01, b1, b0
    if a00 == 0.0:
        return None
    l = a10 * (1.0 / a00) if abs(a00) >= DBL_MIN else a10 / a00
    u11 = a11 - l * a01
    if u11 == 0.0:
        return None
    y1 = fma(-l, b0, b1)
    x1 = y1 / u11
    x0 = fma(-a01, x1, b0) / a00
    return x0, x1


def cond2(A) -> float:
    """numpy.linalg.cond(A) for a 2x2 A (row-major 4-sequence)."""
    a, b, c, d = (float(v) for v in A)
    if any(math.isnan(v) for v in (a, b, c, d)):
        return math.nan
    smax = 0.5 * (math.sqrt((a + d) ** 2 + (c - b) ** 2) + math.sqrt((a - d) ** 2 + (c + b) ** 2))
    w = b * c
    det = fma(a, d, -w) + fma(-b, c, w)
    if smax == 0.0 or det == 0.0:
        return math.inf
    return smax / (abs(det) / smax)
