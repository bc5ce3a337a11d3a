"""Host composition of the Ermak-McCammon step from pieces already pinned to the reference.

Test infrastructure.  Each piece has its own pin in tests/test_oracle.py:
  D(r)      po.compute_D           bit-identical to the reference's compute_D
  L         LAPACK dpotrf through scipy.linalg.cholesky: the OpenBLAS that the reference's
            LAPACK build links (oracle/Makefile); equal to the reference's factor (NR
            choldc and LAPACK)
  F(r)      po.forces              the reference's compute_forces to 1e-13
The composition follows comp_step_ermak / compute_hydro_ermak (src/diff_alg/diff_comp.c:
152-179, diff_equation.c:216-264, diff_rand.c:64-85) with one addition of this backend:
the factor is refreshed only on steps whose index since the start is a multiple of R
(R = 1 is the reference's refactor on every step, diff_rand.c:76-82), while the D.F term
always uses D at the step's own positions.  `schedule` overrides the refactor steps, so a
test can show that a shifted schedule is told apart.
"""
from __future__ import annotations

import numpy as np

KB = 0.001986248532617   # kB in kcal/(mol K), src/data.h:224


def compose_steps(sysm, P, xi, R, schedule=None):
    """Positions (N, 4) after each of len(xi) steps, and the per-step displacements
    (Rf + Rs, 3N) they came from.  refactor steps: k % R == 0, or k in `schedule`."""
    import scipy.linalg as sl
    from oracle import pyoracle as po
    p = po.params_from(P)
    coord = np.array(sysm.coord, float).copy()
    F, _, _ = po.forces(sysm, p, coord)
    pre = P.dt / (KB * P.T)
    s = np.sqrt(2.0 * P.dt)
    n = 3 * sysm.N
    D = np.empty((n, n))
    L = None
    out, disp = [], []
    for k, x in enumerate(np.asarray(xi, float)):
        po.compute_D(coord, p, out=D)
        refactor = (k in schedule) if schedule is not None else (k % R == 0)
        if L is None or refactor:
            # D is symmetric: its transpose is the Fortran-ordered operand dpotrf takes
            L = np.tril(sl.cholesky(D.T, lower=True, overwrite_a=False, check_finite=False))
        d = pre * (D @ F) + L @ (s * x)
        coord[:, :3] = coord[:, :3] + d.reshape(-1, 3)
        F, _, _ = po.forces(sysm, p, coord)
        out.append(coord.copy())
        disp.append(d)
    return out, disp
