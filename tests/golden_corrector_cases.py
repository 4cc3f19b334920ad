"""Correction problems of tests/golden/corrector_golden.npz (shared by the
script that writes the vectors and the tests that read them): Halo and
Lyapunov problems that converge, stop at max_attempts or fail the line
search, a far-away state, vertical orbits on the z = 0 section, and guesses
drawn as the serving stream draws them."""
import numpy as np

from hiten_amd import corrector as Cr, workloads as W

MU = W.MU_EM


def problems():
    xl1 = W.l1_position(MU)
    c2 = (1 - MU) / abs(xl1 + MU)**3 + MU / abs(xl1 - 1 + MU)**3
    wp = np.sqrt((2 - c2 + np.sqrt(9 * c2**2 - 8 * c2)) / 2)
    kap = (wp**2 + 1 + 2 * c2) / (2 * wp)

    def lyap(ax):
        return [xl1 - ax, 0, 0, 0, kap * ax * wp, 0]

    ps = [
        Cr.halo_problem(MU, [0.8234, 0, 0.0224, 0, 0.1343, 0], max_attempts=20),
        Cr.halo_problem(MU, [0.84, 0, 0.1, 0, 0.22, 0], max_attempts=20),
        Cr.halo_problem(MU, [0.82, 0, 0.05, 0, 0.15, 0], max_attempts=4),
        Cr.lyapunov_problem(MU, lyap(0.005), max_attempts=20),
        Cr.lyapunov_problem(MU, lyap(0.01), max_attempts=20),
        Cr.halo_problem(MU, [0.8234, 0, 0.0224, 0, 0.1343, 0], max_attempts=20,
                        alpha_reduction=0.7, armijo_c=0.3),
        Cr.halo_problem(MU, [3.0, 0, 0.0, 0, 0.0, 0.0], max_attempts=5),
    ]
    ps += [Cr.vertical_problem(MU, [0.84, 0, 0, 0, vy, vz], max_attempts=25)
           for vy, vz in ((0.006, 0.1387), (0.0, 0.1187), (0.02, 0.1687))]
    rng = np.random.default_rng(20261202)
    for _ in range(24):      # seeds drawn as the serving stream draws them
        ps.append(Cr.halo_problem(MU, [rng.uniform(0.80, 0.88), 0, rng.uniform(-0.2, 0.2), 0,
                                       rng.uniform(-0.3, 0.3), 0], max_attempts=10))
    return ps
