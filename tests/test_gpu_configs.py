"""The small benchmark configurations at full size, against the oracle.

32^3 Q1 SUPG (35,937 DOFs, eps 1e-3, b = (1,2,3)/sqrt(14), manufactured
sin-sin-sin solution, delta0 = 1), levels 8^3..32^3 with the 8^3 level solved
exactly: the plain FP64 form and the FP32-in-FP64 form both reach FP64
relative residual 1e-10 and agree with the oracle's FP64 solution of the
same system within 1e-8 (max norm, relative).  Iteration counts are printed
(the precision form's report).  With SUPG delta0 = 0.25 the SOR/SSOR and
Jacobi sweeps of this operator diverge (spectral radii > 1, see
tests/test_oracle.py), so delta0 = 1 is used for the solvable variant.

The configurations exactly as stated (delta0 = 0.25, V(2,2) damped Jacobi
omega = 0.8, FGMRES(20), 8^3 level solved exactly) are run for a fixed
number of iterations: the GPU iterate after 40 FGMRES iterations equals the
oracle's (Jacobi does not depend on the row order), and the residual
stagnates as the oracle's does (0.9969 of the initial one after 40 and after
400 iterations).
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


def config1_as_specified(precision, maxit):
    return P.solver_opt(solver=16, restart=20, maxit=maxit, res_norm_min=0.0, red_factor=0.0, smoother=1, omega=0.8,
                        pre=2, post=2, coarse=17, precision=precision)


def test_config1_and_3_as_specified(ctx):
    print("[case] 32^3 Q1 SUPG delta0 0.25, damped Jacobi, FGMRES(20): config 1 and 3 as stated", flush=True)
    prm = P.config1()
    S = pa.System(ctx, 3, 3, 1)
    S.assemble(prm)
    O = orc.System(3, 3, 1)
    O.assemble(prm)
    rhs = O.rhs()
    fine = O.levels[-1]
    E = O.entries()
    sol0 = np.zeros_like(rhs)
    sol0[fine.active:] = rhs[fine.active:]
    r0 = fine.defect(E, sol0, rhs)[1]
    o40, _ = O.solve(config1_as_specified(0, 40), rhs, sol0)
    ro = fine.defect(E, o40, rhs)[1] / r0
    out = {}
    for prec in (0, 1):
        s, info, rc = S.solve(config1_as_specified(prec, 40), rhs, sol0, check=False)
        assert rc == 4 and np.all(np.isfinite(s)), (prec, rc, info)  # iteration limit, no breakdown
        err = float(np.max(np.abs(s - o40)) / np.max(np.abs(o40)))
        rg = fine.defect(E, s, rhs)[1] / r0
        s400, info400, rc400 = S.solve(config1_as_specified(prec, 400), rhs, sol0, check=False)
        r400 = fine.defect(E, s400, rhs)[1] / r0
        out[prec] = dict(err40=err, res40=rg, res400=r400, fp64_fallback=int(info[6]))
        assert rc400 == 4 and 0.99 < r400 <= rg * (1 + (1e-6 if prec == 0 else 1e-3)), (prec, r400, rg)
        if prec == 0:
            assert err < 1e-10, err
            assert abs(rg - ro) < 1e-10, (rg, ro)
        else:
            # config 3: the FP32 inner solve against the FP64 oracle iterate
            assert err < 1e-3 and abs(rg - ro) < 1e-3, (err, rg, ro)
    print(f"oracle residual ratio after 40 iterations {ro:.6f}; GPU {out}", flush=True)
    S.close()
