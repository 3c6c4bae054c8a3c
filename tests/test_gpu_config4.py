"""The learned-smoother training set of the benchmark configurations, scaled
from 64^3 to 32^3 Q1 (35,937 DOFs, levels 4^3 .. 32^3): coefficient fields
with log-normal diffusion (Gaussian random field, correlation length 0.1,
sigma 1) and convection (1,2,3)/sqrt(14) scaled to cell Peclet numbers
log-uniform in [10, 1000], SUPG; (residual, exact error) pairs from the plain
cycle on the three finest levels (the homogeneous problem, whose exact error
is minus the iterate: loss 0, data 0), the 4-layer 16-channel network, Adam
with learning rate 1e-3 and batch 8.  Trained on four fields, the networks
solve two held-out fields to FP64 relative residual 1e-10 in fewer FGMRES
iterations than the FP32 form with SSOR post-smoothing, and a batch of 8
copies of one sample takes the same Adam steps as batch 1 (mean gradient)."""
import numpy as np
import pytest

import parmoon_amd as pa
from parmoon_amd import prm as P

pytestmark = pytest.mark.gpu

BASE = dict(solver=16, restart=20, maxit=400, res_norm_min=0.0, red_factor=0.0, smoother=3, omega=1.0, pre=2,
            post=2, rel_target=1e-10)


def test_batch_is_the_mean_gradient(ctx):
    print("[case] learned training batch 8 of one sample == batch 1", flush=True)
    S = pa.System(ctx, 1, 3, 1)
    S.assemble(P.lognormal(seed=5, nref=3, peclet=30.0))
    out = {}
    for B in (1, 8):
        S.learned_init(3)
        S.learned_train(P.solver_opt(**dict(BASE, precision=1)), samples=1, steps=5, lr=1e-3, seed=1, levels=1,
                        loss=0, data=0, batch=B)
        out[B] = S.learned_params()
    rel = float(np.max(np.abs(out[1] - out[8])) / np.max(np.abs(out[1])))
    print(f"max relative parameter difference {rel:.2e}", flush=True)
    assert rel < 1e-5
    S.close()


def test_training_set_as_specified(ctx):
    print("[case] config-4 training set at 32^3 Q1", flush=True)
    S = pa.System(ctx, 2, 4, 1)
    rng = np.random.default_rng(0)
    pe = 10 ** rng.uniform(1, 3, 6)
    S.learned_init(1)
    for k in range(4):
        S.assemble(P.lognormal(seed=k, nref=5, peclet=pe[k]))
        tinfo = S.learned_train(P.solver_opt(**dict(BASE, precision=1)), samples=8, steps=100, lr=1e-3, seed=k,
                                levels=3, loss=0, data=0, batch=8)
        assert np.isfinite(tinfo[1]) and (k > 0 or tinfo[1] < 0.5 * tinfo[0]), tinfo
    rows = []
    for k in (4, 5):
        S.assemble(P.lognormal(seed=k, nref=5, peclet=pe[k]))
        its = {}
        for name, kw in (("fp64", dict(precision=0)), ("fp32", dict(precision=1)),
                         ("learned", dict(precision=2, learned_levels=3))):
            info, rc = S.solve_resident(P.solver_opt(**dict(BASE, **kw)), check=False)
            assert rc == 0 and info[1] <= 1e-10 * info[2] * 1.0001, (name, rc, info)
            its[name] = int(info[0])
            if name == "learned":
                assert int(info[5]) == 0, info
        rows.append((round(float(pe[k]), 1), its))
        assert its["learned"] < its["fp32"], its
    print(f"held-out fields (Peclet, iterations): {rows}", flush=True)
    S.close()
