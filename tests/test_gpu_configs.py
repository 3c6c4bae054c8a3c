"""The small benchmark configurations at full size, against the oracle.

32^3 Q1 SUPG (35,937 DOFs, eps 1e-3, b = (1,2,3)/sqrt(14), manufactured
sin-sin-sin solution, delta0 = 1), levels 8^3..32^3 with the 8^3 level solved
exactly: the plain FP64 form and the FP32-in-FP64 form both reach FP64
relative residual 1e-10 and agree with the oracle's FP64 solution of the
same system within 1e-8 (max norm, relative).  Iteration counts are printed
(the precision form's report).  With SUPG delta0 = 0.25 the SOR/SSOR and
Jacobi sweeps of this operator diverge (spectral radii > 1, see
tests/test_oracle.py), so delta0 = 1 is used.
"""
import numpy as np
import pytest

import orc
import parmoon_amd as pa
from parmoon_amd import prm as P

pytestmark = pytest.mark.gpu


def test_config1_plain_and_precision_forms(ctx):
    print("[case] 32^3 Q1 SUPG, FP64 and FP32-in-FP64", flush=True)
    prm = P.config1b()
    S = pa.System(ctx, 3, 3, 1)
    S.assemble(prm)
    O = orc.System(3, 3, 1)
    O.assemble(prm)
    rhs = O.rhs()
    fine = O.levels[-1]
    sol0 = np.zeros_like(rhs)
    sol0[fine.active:] = rhs[fine.active:]
    ref, _ = O.solve(P.solver_opt(solver=16, restart=20, maxit=400, res_norm_min=0.0, red_factor=1e-12,
                                  smoother=3, omega=1.0), rhs, sol0)
    res = {}
    for prec in (0, 1):
        opt = P.solver_opt(solver=16, restart=20, maxit=400, res_norm_min=0.0, red_factor=0.0, smoother=3,
                           omega=1.0, precision=prec, rel_target=1e-10)
        s, info, rc = S.solve(opt, rhs, sol0)
        assert rc == 0 and info[1] <= 1e-10 * info[2] * 1.0001, info
        err = float(np.max(np.abs(s - ref)) / np.max(np.abs(ref)))
        res[prec] = (int(info[0]), int(info[3]), err)
        assert err < 1e-8, (prec, err)
    print(f"iterations / outer / error vs oracle: FP64 {res[0]}, FP32-in-FP64 {res[1]}", flush=True)
    S.close()
