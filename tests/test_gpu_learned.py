"""Learned post-smoother on the GPU (libpmx through the C-ABI) against the oracle.

Forward: BF16 operands, FP32 accumulation on MFMA vs the oracle with the same
BF16 rounding and exact sums: relative 2-norm error <= 1e-3 (observed up to
2.9e-4 on the 33^3 lattice).  Gradients: the device also rounds the backward
flow to BF16, so <= 1e-2 per layer (relative 2-norm; observed up to 4.6e-3)
with cosine similarity >= 0.9999.  Adam: FP32 update of the same
gradient, 1e-6 relative.  Post-smoothing step vs the oracle's FP64 restatement
of it: update within 3e-2 (relative 2-norm), defect never larger than before.
"""
import os
import sys

import numpy as np
import pytest

import orc
import parmoon_amd as pa
from parmoon_amd import prm as P

sys.path.insert(0, os.path.join(os.path.dirname(__file__), "..", "oracle"))
import learned_ref as LR  # noqa: E402

pytestmark = pytest.mark.gpu


def nrel(a, b):
    return float(np.linalg.norm(a - b) / max(np.linalg.norm(b), 1e-300))


@pytest.fixture(scope="module")
def small(ctx):
    S = pa.System(ctx, 1, 2, 1)
    yield S
    S.close()


@pytest.mark.parametrize("M", [5, 9, 20, 33])
def test_forward_vs_oracle(small, M):
    print(f"[case] learned forward M={M}", flush=True)
    rng = np.random.default_rng(M)
    p = LR.init_params(M, scale_last=1.0).astype(np.float32)
    small.learned_params(p)
    u = rng.standard_normal((M, M, M, 16)).astype(np.float32)
    out = small.learned_forward(u)
    ref = LR.forward(p.astype(np.float64), u.astype(np.float64))
    print(f"forward rel err {nrel(out, ref):.3e}", flush=True)
    assert nrel(out, ref) < 1e-3, nrel(out, ref)


@pytest.mark.parametrize("M", [6, 17])
def test_gradient_vs_oracle(small, M):
    print(f"[case] learned gradient M={M}", flush=True)
    rng = np.random.default_rng(100 + M)
    p = LR.init_params(7 + M, scale_last=1.0).astype(np.float32)
    small.learned_params(p)
    u = rng.standard_normal((M, M, M, 16)).astype(np.float32)
    t = rng.standard_normal((M, M, M)).astype(np.float32)
    g, loss = small.learned_gradient(u, t)
    l_ref, g_ref = LR.backward(p.astype(np.float64), u.astype(np.float64), t.astype(np.float64))
    assert abs(loss - l_ref) < 1e-3 * l_ref
    offs = np.cumsum([0] + [LR.CH[l + 1] * 27 * LR.CH[l] + LR.CH[l + 1] for l in range(4)])
    for l in range(4):
        a, b = g[offs[l]:offs[l + 1]], g_ref[offs[l]:offs[l + 1]]
        cos = float(a @ b / (np.linalg.norm(a) * np.linalg.norm(b)))
        print(f"gradient layer {l} rel err {nrel(a, b):.3e} cos {cos:.6f}", flush=True)
        assert nrel(a, b) < 1e-2 and cos > 0.9999, (l, nrel(a, b), cos)


def test_adam_step(small):
    print("[case] learned adam", flush=True)
    M = 7
    rng = np.random.default_rng(11)
    p = LR.init_params(11).astype(np.float32)
    small.learned_init(5)
    small.learned_params(p)     # fresh moments come with init
    u = rng.standard_normal((M, M, M, 16)).astype(np.float32)
    t = rng.standard_normal((M, M, M)).astype(np.float32)
    g, _ = small.learned_gradient(u, t)
    small.learned_adam(1e-3)
    p1 = small.learned_params()
    ref, _, _ = LR.adam(p.astype(np.float64), g.astype(np.float64), 0.0, 0.0, 1, 1e-3)
    assert np.max(np.abs(p1 - ref)) < 1e-6 * max(1.0, np.max(np.abs(ref)))


def lattice_map(space, M):
    xyz = space.dof_coords()
    ijk = np.rint(xyz * (M - 1)).astype(int)
    return ijk[:, 0] + M * (ijk[:, 1] + M * ijk[:, 2])


def post_reference(O, params, x, f):
    """FP64 restatement of one learned post-smoothing step (DOF numbering)."""
    fine = O.levels[-1]
    E = O.entries()
    bi, gn, rp, kc = fine.arrays()
    N, na = fine.N, fine.active
    rows = np.repeat(np.arange(N), np.diff(rp))
    diag = E[(kc == rows)]
    d, _ = fine.defect(E, x, f)
    u = np.where(np.arange(N) < na, d / diag, 0.0)
    nrm, sq = np.linalg.norm(u), np.sqrt(na)
    M = round(N ** (1 / 3))
    lat = lattice_map(fine, M)
    ul = np.zeros(M ** 3)
    ul[lat] = u * sq / nrm
    A_lat, _ = lattice_operator(O, M)
    out = LR.forward(params.astype(np.float64), ul.reshape(M, M, M), feat=LR.features(A_lat, M)).ravel()
    c = np.where(np.arange(N) < na, out[lat] * nrm / sq, 0.0)
    Ac = fine.matvec(E, c)
    alpha = float(d @ Ac) / float(Ac @ Ac)
    return x + max(0.0, alpha) * c, d, alpha


@pytest.mark.parametrize("order", [1, 2])
def test_post_step_vs_oracle(ctx, order):
    print(f"[case] learned post step Q{order}", flush=True)
    prm = P.config1b() if order == 1 else P.config2(nref=3)
    S = pa.System(ctx, 1, 3, order)
    S.assemble(prm)
    O = orc.System(1, 3, order)
    O.assemble(prm)
    p = LR.init_params(21 + order).astype(np.float32)
    fine = O.levels[-1]
    rng = np.random.default_rng(order)
    x = rng.standard_normal(fine.N)
    f = O.rhs()
    x[fine.active:] = f[fine.active:]
    # the parameters and their copy with the output layer negated give
    # corrections of opposite sign: one takes a step, the other is clamped
    n3 = 27 * 16 + 1
    alphas = []
    for sign in (1.0, -1.0):
        q = p.copy()
        q[-n3:] *= sign
        S.learned_params(q)
        ref, d0, alpha = post_reference(O, q, x, f)
        got = S.learned_post(x, f)
        alphas.append(alpha)
        if alpha > 0:
            assert nrel(got - x, ref - x) < 3e-2, nrel(got - x, ref - x)
        else:
            assert np.max(np.abs(got - x)) <= 1e-6 * np.max(np.abs(x))  # only FP32 rounding of x
        d1, _ = fine.defect(O.entries(), got, f)
        assert np.linalg.norm(d1) <= np.linalg.norm(d0) * (1 + 1e-6)
    assert alphas[0] * alphas[1] < 0
    S.close()


def lattice_operator(O, M):
    """Oracle operator and 1/diag of the finest level in lattice order."""
    import scipy.sparse as sp
    fine = O.levels[-1]
    E = O.entries()
    bi, gn, rp, kc = fine.arrays()
    lat = lattice_map(fine, M)
    A = sp.csr_matrix((E, kc, rp), shape=(fine.N, fine.N)).tocoo()
    A_lat = sp.csr_matrix((A.data, (lat[A.row], lat[A.col])), shape=(fine.N, fine.N))
    interior = np.zeros(fine.N, bool)
    interior[lat[:fine.active]] = True
    dinv = np.where(interior, 1.0 / A_lat.diagonal(), 0.0)
    return A_lat, dinv


@pytest.mark.parametrize("order", [1, 2])
def test_residual_gradient_vs_oracle(ctx, order):
    print(f"[case] learned residual-loss gradient Q{order}", flush=True)
    prm = P.config1b() if order == 1 else P.config2(nref=3)
    S = pa.System(ctx, 1, 3, order)
    S.assemble(prm)
    O = orc.System(1, 3, order)
    O.assemble(prm)
    M = round(O.levels[-1].N ** (1 / 3))
    A, dinv = lattice_operator(O, M)
    p = LR.init_params(31 + order, scale_last=1.0).astype(np.float32)
    S.learned_params(p)
    u = np.random.default_rng(40 + order).standard_normal((M, M, M)).astype(np.float32)
    g, loss = S.learned_gradient_level(u)
    l_ref, g_ref = LR.backward(p.astype(np.float64), u.astype(np.float64),
                               head=LR.residual_head(u.astype(np.float64), A, dinv), feat=LR.features(A, M))
    assert abs(loss - l_ref) < 1e-3 * l_ref, (loss, l_ref)
    offs = np.cumsum([0] + [LR.CH[l + 1] * 27 * LR.CH[l] + LR.CH[l + 1] for l in range(4)])
    for l in range(4):
        a, b = g[offs[l]:offs[l + 1]], g_ref[offs[l]:offs[l + 1]]
        cos = float(a @ b / (np.linalg.norm(a) * np.linalg.norm(b)))
        assert nrel(a, b) < 3e-2 and cos > 0.999, (l, nrel(a, b), cos)
    S.close()


def test_training_gain_and_graph_replay(ctx):
    """What the trained network gains: on 16^3 Q2 SUPG (3 levels) the FGMRES
    solve with the learned post-smoother on the finest level needs fewer
    iterations than the same cycle without post-smoothing and than with the
    untrained network; training lowers the loss.  The learned cycle replayed
    from its HIP graph matches the eager cycle (also after changing the number
    of learned steps, which must rebuild the graph)."""
    print("[case] learned training gain", flush=True)
    S = pa.System(ctx, 2, 3, 2)
    S.assemble(P.config2(nref=4))
    base = dict(solver=16, restart=30, maxit=300, res_norm_min=0.0, red_factor=0.0, smoother=3, pre=2, post=2,
                omega=1.0, rel_target=1e-10)

    def its(**kw):
        o = dict(base)
        o.update(kw)
        info, rc = S.solve_resident(P.solver_opt(**o), check=False)
        assert rc == 0 and info[1] <= 1e-10 * info[2] * 1.0001, (rc, info)
        return int(info[0]), S.solution()

    n_nopost, _ = its(precision=1, post=0)
    S.learned_init(7)
    n_untrained, _ = its(precision=2, learned_levels=1)
    tinfo = S.learned_train(P.solver_opt(**dict(base, precision=1)), samples=16, steps=400, lr=2e-3, seed=1,
                            levels=1, data=1)
    assert tinfo[1] < 0.6 * tinfo[0], tinfo
    n_learned, x_graph = its(precision=2, learned_levels=1)
    print(f"iterations: no post-smoothing {n_nopost}, untrained {n_untrained}, trained {n_learned}", flush=True)
    assert n_learned < n_nopost and n_learned < n_untrained
    # graph replay == eager, for one and two learned steps
    for steps in (1, 2):
        n_g, x_g = its(precision=2, learned_levels=1, learned_steps=steps)
        n_e, x_e = its(precision=2, learned_levels=1, learned_steps=steps, eager=1)
        assert n_g == n_e and np.max(np.abs(x_g - x_e)) <= 1e-12 * np.max(np.abs(x_e))
    S.close()


def test_learned_coarse_levels(ctx):
    """Learned smoothing on the levels below the finest (the finest keeps SSOR),
    two network steps replacing both pre- and post-smoothing there: on 32^3 Q2
    SUPG (4 levels) training only the coarse levels leaves the finest level's
    network untouched and lowers their losses, the solve reaches 1e-10 without
    falling back (within 1.5x the plain FP32 iterations after this short
    training) and is close to the FP64 solution, and the graph replay matches
    the eager cycle."""
    print("[case] learned coarse levels", flush=True)
    S = pa.System(ctx, 2, 4, 2)
    S.assemble(P.config2(nref=5))
    base = dict(solver=16, restart=30, maxit=300, res_norm_min=0.0, red_factor=0.0, smoother=3, pre=2, post=2,
                omega=1.0, rel_target=1e-10)
    coarse = dict(precision=2, learned_levels=3, learned_skip=1, learned_steps=2, learned_pre=1)

    def run(**kw):
        o = dict(base)
        o.update(kw)
        info, rc = S.solve_resident(P.solver_opt(**o), check=False)
        assert rc == 0 and info[1] <= 1e-10 * info[2] * 1.0001, (rc, info)
        assert info[5] == 0, info  # no fallback to the plain smoother
        return int(info[0]), S.solution()

    n64, x64 = run(precision=0)
    n32, _ = run(precision=1)
    S.learned_init(5)
    fine_before = S.learned_params(level=-1)
    tinfo = S.learned_train(P.solver_opt(**dict(base, precision=1)), samples=16, steps=300, lr=2e-3, seed=2,
                            levels=3, data=1, skip=1)
    assert np.array_equal(S.learned_params(level=-1), fine_before)
    for j in (1, 2):  # levels L-1 and L-2 trained: the loss falls
        assert tinfo[5 + 2 * j] < 0.8 * tinfo[4 + 2 * j], tinfo
    n_l, x_l = run(**coarse)
    print(f"iterations: fp64 {n64}, fp32 {n32}, learned coarse levels {n_l}", flush=True)
    assert n_l <= 1.5 * n32
    assert np.max(np.abs(x_l - x64)) <= 1e-7 * np.max(np.abs(x64))
    n_e, x_e = run(eager=1, **coarse)
    assert n_e == n_l and np.max(np.abs(x_e - x_l)) <= 1e-12 * np.max(np.abs(x_l))
    S.close()


def test_fallback_to_the_plain_smoother(ctx):
    """The plain smoother takes over when learned cycles stop reducing the
    residual: with networks whose output is zero on every level above the
    coarsest (learned steps there change nothing: the step length of a zero
    correction is 0) and replacing both pre- and post-smoothing, the cycle is
    a bare coarse-grid correction; the FGMRES estimate stalls, the solve
    switches to SSOR (info[5] >= 1) and still reaches 1e-10."""
    print("[case] learned fallback", flush=True)
    S = pa.System(ctx, 1, 3, 2)
    S.assemble(P.config2(nref=3))
    zero = np.zeros(S.learned_nparams(), np.float32)
    for lvl in (1, 2):
        S.learned_params(zero, level=lvl)
    base = dict(solver=16, restart=30, maxit=300, res_norm_min=0.0, red_factor=0.0, smoother=3, pre=2, post=2,
                omega=1.0, rel_target=1e-10)
    info, rc = S.solve_resident(P.solver_opt(**dict(base, precision=2, learned_levels=2, learned_skip=0,
                                                    learned_pre=1, learned_steps=1)), check=False)
    print(f"fallback solve: rc {rc} iterations {int(info[0])} fallbacks {int(info[5])}", flush=True)
    assert rc == 0 and info[5] >= 1, info
    assert info[1] <= 1e-10 * info[2] * 1.0001
    x = S.solution()
    info64, rc64 = S.solve_resident(P.solver_opt(**dict(base, precision=0)))
    assert rc64 == 0
    x64 = S.solution()
    assert np.max(np.abs(x - x64)) <= 1e-7 * np.max(np.abs(x64))
    S.close()


def test_learned_coarse_zero_start_shortcut(ctx):
    """A coarse level entered after its Reset (x = 0) starts its learned
    smoothing from d = f instead of forming f - A x: the solve gives the same
    bits with the shortcut switched off (PMX_NO_XZERO=1)."""
    print("[case] learned x_zero shortcut", flush=True)
    S = pa.System(ctx, 1, 3, 2)
    S.assemble(P.config2(nref=3))
    S.learned_init(3)
    base = dict(solver=16, restart=30, maxit=300, res_norm_min=0.0, red_factor=0.0, smoother=3, pre=2, post=2,
                omega=1.0, rel_target=1e-10, precision=2, learned_levels=2, learned_skip=1, learned_pre=1,
                learned_steps=2)
    S.learned_train(P.solver_opt(**dict(base, precision=1)), samples=4, steps=50, lr=2e-3, seed=1, levels=2,
                    data=1, skip=1)
    info_a, rc_a = S.solve_resident(P.solver_opt(**base))
    xa = S.solution()
    os.environ["PMX_NO_XZERO"] = "1"
    try:
        info_b, rc_b = S.solve_resident(P.solver_opt(**base))
        xb = S.solution()
    finally:
        del os.environ["PMX_NO_XZERO"]
    assert rc_a == 0 and rc_b == 0
    assert info_a[0] == info_b[0] and np.array_equal(xa, xb)
    S.close()
