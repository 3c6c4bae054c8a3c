"""The SUPG time discretisation (MatrixMRhsAssemble_SUPG at the start,
MatricesAKRhsAssemble_SUPG every step, system matrix M + K + tau th1 A,
right-hand side tau th3 oldrhs + tau th4 rhs + (M + K - tau th2 A) oldsol) on
the GPU against the oracle's time loop.

The rotating-convection transient (eps = 1e-4, b = (-(y - 1/2), x - 1/2, 0.2),
a Gaussian blob as initial value) on 16^3 Q1 hexahedra with 4 GMG levels,
Crank-Nicolson with tau = 1e-2, 10 steps, V(2,2) SSOR with omega = 0.5
(the time-derivative matrix K of the SUPG form is as large as M and makes
M + K far from diagonally dominant: with omega = 1 the multicolour sweep
needs 180 FGMRES iterations per step, the oracle shows the same).  The
oracle sweeps in the device's multicolour order, so the FP64 iteration
counts agree step by step; the final solution agrees within 1e-10 with
that oracle and within 1e-9 with the oracle's lexicographic sweeps (the
reference's order), and the per-step error norms (against the blob) agree
within 1e-9.
"""
import numpy as np
import pytest

import orc
import parmoon_amd as pa
from parmoon_amd import prm as P

pytestmark = pytest.mark.gpu


def rel(a, b):
    return float(np.max(np.abs(a - b)) / max(np.max(np.abs(b)), 1e-300))


def supg_opt(precision=0):
    return P.solver_opt(solver=16, restart=50, maxit=200, res_norm_min=0.0, red_factor=1e-12, smoother=3,
                        omega=0.5, precision=precision)


@pytest.mark.parametrize("precision", [0, 1])
def test_tcd_supg_rotating_matches_oracle(ctx, precision):
    print(f"[case] tcd supg rotating 16^3 Q1 precision={precision}", flush=True)
    prm = P.rotating(seed=1, form=P.FORM_TCD_SUPG_AK)
    tau, steps = 1e-2, 10
    S = pa.System(ctx, 1, 4, 1)
    fine = orc.Space(4, 1)
    u0 = fine.interpolate(prm)
    S.tcd_init(prm, u0)
    td = [tau, .5, .5, .5, .5]
    opt = supg_opt(precision)
    errs, its = [], []
    t = 0.0
    for m in range(steps):
        t += tau
        p = prm.copy()
        p[P.TIME] = t
        info, rc = S.tcd_step(p, td, opt)
        assert rc == 0 and info[5] == 0, info
        errs.append(S.errors(p))
        its.append(int(info[0]))
    sol = S.solution()
    tdisc = [tau, tau * (steps + 0.5), .5, .5, .5, .5]
    Om = orc.System(1, 4, 1)
    Om.set_multicolour()
    om_sol, om_err = Om.tcd3d(prm, tdisc, supg_opt(), u0, max_steps=steps)
    assert len(om_err) == steps
    assert rel(sol, om_sol) < 1e-10
    if precision == 0:
        assert all(abs(a - b) <= 1 for a, b in zip(its, om_err[:, 5])), (its, om_err[:, 5])
    for (e0, e1), row in zip(errs, om_err):
        assert abs(e0 - row[1]) <= 1e-9 * row[1] and abs(e1 - row[2]) <= 1e-9 * row[2]
    Ol = orc.System(1, 4, 1)
    ol_sol, _ = Ol.tcd3d(prm, tdisc, supg_opt(), u0, max_steps=steps)
    assert rel(sol, ol_sol) < 1e-9
    # the SUPG solution differs from the Galerkin one on this mesh (cell
    # Peclet number ~ 300): the streamline terms are in effect
    G = pa.System(ctx, 1, 4, 1)
    pg = P.rotating(seed=1, form=P.FORM_TCD_A)
    G.tcd_init(pg, u0)
    t = 0.0
    for m in range(steps):
        t += tau
        p = pg.copy()
        p[P.TIME] = t
        G.tcd_step(p, td, opt)
    assert rel(G.solution(), sol) > 1e-3
    G.close()
    S.close()
