"""Time stepping (TSystemTCD3D) on the GPU against the reference's golden run log
and the oracle's time loop.

16^3 Q1 heat equation (Examples/TCD_3D/amc.h), Crank-Nicolson with
tau = 0.1082532 to T = 2.05681, 4 GMG levels: every L2 error of
ReadIn/tcd3d_GMG.out to the 7 printed digits, the accumulated L2(0,T;L2)
within 1.5e-9, and the final solution within 1e-9 (max norm, relative) of the
oracle's loop.  Errors of the GPU solutions are evaluated by the oracle's
quadrature on the host, and the device's own error norms (L2 and H1 semi)
agree with those to 1e-11.
"""
import json
import os

import numpy as np
import pytest

import orc
import parmoon_amd as pa
from parmoon_amd import prm as P

pytestmark = pytest.mark.gpu
HERE = os.path.dirname(os.path.abspath(__file__))


def rel(a, b):
    return float(np.max(np.abs(a - b)) / max(np.max(np.abs(b)), 1e-300))


@pytest.mark.parametrize("precision", [0, 1])
def test_golden_tcd3d_log_gpu(ctx, precision):
    print(f"[case] tcd3d golden log precision={precision}", flush=True)
    g = json.load(open(os.path.join(HERE, "golden", "tcd3d_gmg_log.json")))
    S = pa.System(ctx, 1, 4, 1)
    O = orc.System(1, 4, 1)
    fine = O.levels[-1]
    prm = P.heat_amc(0.0)
    u0 = fine.interpolate(prm)
    # device interpolation of the initial function == the oracle's nodal values
    S.tcd_init(prm, None)
    assert rel(S.solution(), u0) < 1e-14
    S.tcd_init(prm, u0)
    tau = g["tau"]
    td = [tau, .5, .5, .5, .5]
    opt = P.solver_opt(solver=16, restart=20, maxit=100, res_norm_min=1e-13, smoother=3, omega=1.0,
                       precision=precision)
    t, acc = 0.0, 0.0
    e_old = fine.errors(prm, u0)[0]
    for m, (tg, l2, l2l2) in enumerate(g["steps"]):
        t += tau
        p = prm.copy()
        p[P.TIME] = t
        info, rc = S.tcd_step(p, td, opt)
        assert rc == 0, info
        e_or = fine.errors(p, S.solution())
        e = e_or[0]
        # the device error norms (pmx_system_errors) agree with the oracle's quadrature
        e_dev = S.errors(p)
        assert abs(e_dev[0] - e_or[0]) <= 1e-11 * e_or[0] and abs(e_dev[1] - e_or[1]) <= 1e-11 * e_or[1]
        if m > 0:
            acc += (e * e + e_old * e_old) * tau / 2.0
        e_old = e
        assert abs(t - tg) < 1e-6
        assert float(f"{e:.6e}") == l2, (m, e, l2)
        if l2l2 is not None:
            assert abs(np.sqrt(acc) - l2l2) <= 1.5e-9
    o_opt = P.solver_opt(solver=16, restart=20, maxit=100, res_norm_min=1e-13, smoother=3, omega=1.0)
    o_sol, _ = O.tcd3d(prm, [tau, g["end_time"], .5, .5, .5, .5], o_opt, u0)
    assert rel(S.solution(), o_sol) < 1e-9
    S.close()
