"""Pins of the learned-smoother oracle (CPU only).

backward vs central finite differences; conv vs torch's CPU conv3d (float64);
the input-gradient convolution as the exact adjoint of conv; BF16 rounding vs
torch's float32 -> bfloat16 conversion and hand-worked ties; Adam's first step.
"""
import os
import sys

import numpy as np
import pytest

sys.path.insert(0, os.path.join(os.path.dirname(__file__), "..", "oracle"))
import learned_ref as LR  # noqa: E402


def test_nparams_layout():
    assert LR.nparams() == 3 * (16 * 27 * 16 + 16) + (1 * 27 * 16 + 1) == 21217
    p = LR.init_params(3)
    assert np.array_equal(LR.pack(LR.unpack(p)), p)


def test_conv_matches_torch_conv3d():
    torch = pytest.importorskip("torch")
    rng = np.random.default_rng(1)
    M, ci, co = 7, 3, 5
    a = rng.standard_normal((M, M, M, ci))
    W = rng.standard_normal((co, 27, ci))
    b = rng.standard_normal(co)
    ours = LR.conv(a, W, b)
    # torch: input (N, C, D, H, W) = (1, ci, z, y, x); weight (co, ci, kz, ky, kx), kx fastest = tap order
    wt = torch.from_numpy(W.transpose(0, 2, 1).reshape(co, ci, 3, 3, 3).copy())
    x = torch.from_numpy(a.transpose(3, 0, 1, 2)[None].copy())
    ref = torch.nn.functional.conv3d(x, wt, torch.from_numpy(b), padding=1)[0].numpy().transpose(1, 2, 3, 0)
    assert np.max(np.abs(ours - ref)) < 1e-12 * np.max(np.abs(ref))


def test_input_grad_is_adjoint():
    rng = np.random.default_rng(2)
    M, ci, co = 6, 4, 3
    a = rng.standard_normal((M, M, M, ci))
    g = rng.standard_normal((M, M, M, co))
    W = rng.standard_normal((co, 27, ci))
    lhs = float((LR.conv(a, W) * g).sum())
    rhs = float((a * LR.conv_input_grad(g, W)).sum())
    assert abs(lhs - rhs) < 1e-11 * abs(lhs)
    dW, db = LR.conv_weight_grad(g, a)
    # <conv(a, W + h E), g> is linear in W: its derivative along E is <dW, E>
    E = rng.standard_normal(W.shape)
    assert abs(float((LR.conv(a, E) * g).sum()) - float((dW * E).sum())) < 1e-10 * abs(float((dW * E).sum()))
    assert np.allclose(db, g.reshape(-1, co).sum(0))


def test_backward_vs_finite_differences():
    rng = np.random.default_rng(3)
    M = 5
    p = LR.init_params(4, scale_last=1.0)
    u = rng.standard_normal((M, M, M, 16))
    t = rng.standard_normal((M, M, M))
    l0, g = LR.backward(p, u, t, bf16_emulation=False)
    assert abs(l0 - LR.loss(LR.forward(p, u, False), t)) < 1e-14 * l0
    offs = np.cumsum([0] + [LR.CH[l + 1] * 27 * LR.CH[l] + LR.CH[l + 1] for l in range(4)])
    idx = np.concatenate([rng.choice(np.arange(offs[l], offs[l + 1]), 12, replace=False) for l in range(4)]
                         + [offs[1:] - 1])  # includes one bias of every layer
    h = 1e-5
    for i in idx:
        pp, pm = p.copy(), p.copy()
        pp[i] += h
        pm[i] -= h
        fd = (LR.loss(LR.forward(pp, u, False), t) - LR.loss(LR.forward(pm, u, False), t)) / (2 * h)
        assert abs(fd - g[i]) <= 1e-6 * max(abs(fd), 1e-3 * np.max(np.abs(g))), (i, fd, g[i])


def test_bf16_rounding():
    assert LR.bf16(1.0) == 1.0
    assert LR.bf16(1 + 2.0 ** -8) == 1.0                    # tie -> even mantissa
    assert LR.bf16(1 + 3 * 2.0 ** -8) == 1 + 2.0 ** -6      # tie -> even (upwards)
    assert LR.bf16(1 + 2.0 ** -8 + 2.0 ** -12) == 1 + 2.0 ** -7
    assert LR.bf16(-2.5) == -2.5
    torch = pytest.importorskip("torch")
    x = np.random.default_rng(5).standard_normal(100000).astype(np.float32) * 10.0 ** np.random.default_rng(6).integers(-6, 6, 100000)
    ref = torch.from_numpy(x).to(torch.bfloat16).to(torch.float64).numpy()
    assert np.array_equal(LR.bf16(x), ref)


def test_adam_first_step():
    rng = np.random.default_rng(7)
    p = rng.standard_normal(50)
    g = rng.standard_normal(50)
    g[3] = 0.0
    p1, m, v = LR.adam(p, g, np.zeros(50), np.zeros(50), 1, 1e-3)
    step = p - p1
    nz = g != 0
    assert np.allclose(np.abs(step[nz]), 1e-3, rtol=1e-6)
    assert np.all(np.sign(step[nz]) == np.sign(g[nz]))
    assert step[3] == 0.0


def test_residual_head_vs_finite_differences():
    import scipy.sparse as sp
    rng = np.random.default_rng(8)
    M = 5
    N = M ** 3
    A = sp.random(N, N, density=0.05, random_state=9, format="csr") + sp.eye(N) * 4.0
    m = LR.interior_mask(M).ravel()
    dinv = np.where(m, 1.0 / A.diagonal(), 0.0)
    p = LR.init_params(10, scale_last=1.0)
    u = rng.standard_normal((M, M, M))
    feat = rng.standard_normal((M, M, M, 15))
    head = LR.residual_head(u, A, dinv)
    l0, g = LR.backward(p, u, bf16_emulation=False, head=head, feat=feat)
    offs = np.cumsum([0] + [LR.CH[l + 1] * 27 * LR.CH[l] + LR.CH[l + 1] for l in range(4)])
    idx = np.concatenate([rng.choice(np.arange(offs[l], offs[l + 1]), 8, replace=False) for l in range(4)])
    h = 1e-5
    for i in idx:
        pp, pm = p.copy(), p.copy()
        pp[i] += h
        pm[i] -= h
        fd = (head(LR.forward(pp, u, False, feat=feat))[0] - head(LR.forward(pm, u, False, feat=feat))[0]) / (2 * h)
        assert abs(fd - g[i]) <= 1e-6 * max(abs(fd), 1e-3 * np.max(np.abs(g))), (i, fd, g[i])
    # with a zero correction the loss is the scaled defect itself
    pz = p.copy()
    k = 0
    for l in range(4):
        n = LR.CH[l + 1] * 27 * LR.CH[l] + LR.CH[l + 1]
        if l == 3:
            pz[k:k + n] = 0.0
        k += n
    lz, _ = LR.backward(pz, u, bf16_emulation=False, head=head, feat=feat)
    assert abs(lz - float((u[1:-1, 1:-1, 1:-1] ** 2).sum()) / (M - 2) ** 3) < 1e-12


def test_features_of_a_stencil():
    """features() reads the couplings along the axes: check on a 13-point
    operator (first and second neighbours) with distinct known coefficients."""
    import scipy.sparse as sp
    M = 7
    N = M ** 3
    coef = {1: -1.5, -1: -0.5, M: -2.0, -M: -0.25, M * M: -3.0, -M * M: -0.75,
            2: 0.1, -2: 0.2, 2 * M: 0.3, -2 * M: 0.4, 2 * M * M: 0.5, -2 * M * M: 0.6}
    rows, cols, vals = [], [], []
    for i in range(N):
        rows.append(i); cols.append(i); vals.append(8.0)
        for off, c in coef.items():
            j = i + off
            if 0 <= j < N:
                rows.append(i); cols.append(j); vals.append(c)
    A = sp.csr_matrix((vals, (rows, cols)), shape=(N, N))
    f = LR.features(A, M)
    m = LR.interior_mask(M)
    m2 = np.zeros((M, M, M), bool)  # nodes whose second neighbours lie in the lattice
    m2[2:-2, 2:-2, 2:-2] = True
    for q, off in enumerate((1, -1, M, -M, M * M, -M * M)):
        assert np.allclose(f[..., q][m], coef[off] / 8.0)
        assert np.allclose(f[..., 6 + q][m2], coef[2 * off] / 8.0)
        assert np.all(f[..., q][~m] == 0) and np.all(f[..., 6 + q][~m] == 0)
    assert np.all(f[..., 12:] == 0)
