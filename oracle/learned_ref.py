"""CPU oracle of the learned post-smoother network (test infrastructure only).

Nothing in the product imports this file; the tests, the smoke check and the
CPU-baseline leg of the benchmark use it as the checker.  The network has no
counterpart in the reference (it replaces the post-sweeps of
TMGLevel3D::SOR/SSOR, src/FE/MGLevel3D.C:280-467), so there is no reference
algorithm to follow; the oracle states the network itself:

* lattice arrays are indexed [z, y, x] (x fastest, the device order), channels
  last: a layer input has shape (M, M, M, cin);
* tap t = (dx+1) + 3 (dy+1) + 9 (dz+1); nodes outside the lattice read as 0;
* layer l: pre = sum_t sum_c W[o, t, c] a[z+dz, y+dy, x+dx, c] + b[o];
  hidden layers a_{l+1} = GELU(pre) (exact erf form), the last layer is linear;
* channels 16 -> 16 -> 16 -> 16 -> 1; parameters packed per layer as
  W[cout][27][cin] followed by b[cout], FP32 on the device;
* input channels per node: the scaled defect u, then the couplings of the
  row along the axes divided by its diagonal, a(p,p+e)/a(p,p) for
  e = +x, -x, +y, -y, +z, -z, then a(p,p+2e)/a(p,p) for the same e (0 where
  the stencil has no such entry and on Dirichlet nodes), then three zeros
  (``features``);
* BF16 emulation (``bf16=True``): the input, every weight and every hidden
  activation is rounded to BF16 (round to nearest even) before it enters a
  convolution, the products are summed exactly (FP64 here, FP32 on MFMA);
* loss = sum over interior nodes (out - target)^2 / (M-2)^3, or the residual
  loss of ``residual_head`` (the scaled defect left by the correction).

Pins: ``backward`` (with either loss head) is checked against central finite
differences of the loss (FP64, no BF16 rounding), ``conv`` against torch's CPU conv3d in
float64 (an independent library routine), and ``adam`` against the closed-form
first step (|step| = lr for every parameter with a nonzero gradient).
"""
import numpy as np
from scipy.special import erf

TAPS = 27
CH = [16, 16, 16, 16, 1]


def offsets():
    """(dz, dy, dx) of tap t = (dx+1) + 3 (dy+1) + 9 (dz+1)."""
    return [(t // 9 - 1, (t // 3) % 3 - 1, t % 3 - 1) for t in range(TAPS)]


def nparams():
    return sum(CH[l + 1] * TAPS * CH[l] + CH[l + 1] for l in range(4))


def unpack(p):
    """Flat parameter vector -> [(W (cout, 27, cin), b (cout,))] per layer."""
    p = np.asarray(p, np.float64)
    out, k = [], 0
    for l in range(4):
        ci, co = CH[l], CH[l + 1]
        W = p[k:k + co * TAPS * ci].reshape(co, TAPS, ci)
        k += co * TAPS * ci
        b = p[k:k + co]
        k += co
        out.append((W, b))
    assert k == p.size
    return out


def pack(layers):
    return np.concatenate([np.concatenate([W.ravel(), b.ravel()]) for W, b in layers])


def init_params(seed=0, scale_last=0.1):
    """Uniform(-a, a), a = sqrt(6 / (27 cin)), biases 0 (any draw serves the tests)."""
    rng = np.random.default_rng(seed)
    layers = []
    for l in range(4):
        ci, co = CH[l], CH[l + 1]
        a = np.sqrt(6.0 / (TAPS * ci)) * (scale_last if l == 3 else 1.0)
        layers.append((rng.uniform(-a, a, (co, TAPS, ci)), rng.uniform(-0.1, 0.1, co)))
    return pack(layers)


def bf16(x):
    """Round to the nearest BF16 value (ties to even), returned as float64."""
    f = np.asarray(x, np.float32)
    u = f.view(np.uint32).astype(np.uint64)
    r = ((u + 0x7FFF + ((u >> 16) & 1)) >> 16) << 16
    return r.astype(np.uint32).view(np.float32).astype(np.float64)


def gelu(v):
    return 0.5 * v * (1.0 + erf(v / np.sqrt(2.0)))


def gelu_d(v):
    return 0.5 * (1.0 + erf(v / np.sqrt(2.0))) + v * np.exp(-0.5 * v * v) / np.sqrt(2.0 * np.pi)


def _pad(a):
    return np.pad(a, ((1, 1), (1, 1), (1, 1), (0, 0)))


def conv(a, W, b=None):
    """a (M, M, M, cin), W (cout, 27, cin) -> (M, M, M, cout), zero outside."""
    M = a.shape[0]
    ap = _pad(a)
    out = np.zeros(a.shape[:3] + (W.shape[0],))
    for t, (dz, dy, dx) in enumerate(offsets()):
        sl = ap[1 + dz:1 + dz + M, 1 + dy:1 + dy + M, 1 + dx:1 + dx + M]
        out += sl @ W[:, t, :].T
    if b is not None:
        out += b
    return out


def conv_input_grad(g, W):
    """Gradient of conv's input for the output gradient g (M, M, M, cout)."""
    M = g.shape[0]
    gp = _pad(g)
    out = np.zeros(g.shape[:3] + (W.shape[2],))
    for t, (dz, dy, dx) in enumerate(offsets()):
        sl = gp[1 - dz:1 - dz + M, 1 - dy:1 - dy + M, 1 - dx:1 - dx + M]
        out += sl @ W[:, t, :]
    return out


def conv_weight_grad(g, a):
    """dW[o, t, c] = sum_n g[n, o] a[n + off_t, c];  db[o] = sum_n g[n, o]."""
    M = a.shape[0]
    ap = _pad(a)
    co, ci = g.shape[3], a.shape[3]
    dW = np.zeros((co, TAPS, ci))
    g2 = g.reshape(-1, co)
    for t, (dz, dy, dx) in enumerate(offsets()):
        sl = ap[1 + dz:1 + dz + M, 1 + dy:1 + dy + M, 1 + dx:1 + dx + M].reshape(-1, ci)
        dW[:, t, :] = g2.T @ sl
    return dW, g2.sum(axis=0)


def features(A, M):
    """(M, M, M, 15) operator features of a lattice-ordered sparse matrix A:
    a(p, p + e) / a(p, p) and a(p, p + 2e) / a(p, p) for e = +x, -x, +y, -y,
    +z, -z on interior nodes (0 where A has no such entry), then three zero
    channels."""
    A = A.tocsr()
    N = M ** 3
    d = A.diagonal()
    m = interior_mask(M).ravel()
    out = np.zeros((N, 15))
    idx = np.nonzero(m)[0]
    for q, off in enumerate((1, -1, M, -M, M * M, -M * M)):
        out[idx, q] = np.asarray(A[idx, idx + off]).ravel() / d[idx]
        j = idx + 2 * off
        ok = (j >= 0) & (j < N)
        vals = np.zeros(idx.size)
        vals[ok] = np.asarray(A[idx[ok], j[ok]]).ravel()
        out[idx, 6 + q] = vals / d[idx]
    return out.reshape(M, M, M, 15)


def net_input(u, feat=None):
    """Stack the scaled defect (M, M, M) and the features (M, M, M, 15) into
    the (M, M, M, 16) network input (features 0 if not given)."""
    u = np.asarray(u, np.float64)
    if u.ndim == 4:
        return u
    f = np.zeros(u.shape + (CH[0] - 1,)) if feat is None else feat
    return np.concatenate([u[..., None], f], axis=-1)


def forward(p, u, bf16_emulation=True, keep=False, feat=None):
    """Network output (M, M, M) for the input u: (M, M, M, 16), or the scaled
    defect (M, M, M) with the features `feat` (M, M, M, 15) (0 if None)."""
    r = bf16 if bf16_emulation else (lambda z: np.asarray(z, np.float64))
    layers = unpack(p)
    a = r(net_input(u, feat))
    acts, pres = [a], []
    for l, (W, b) in enumerate(layers):
        pre = conv(a, r(W), b)
        if l < 3:
            pres.append(pre)
            a = r(gelu(pre))
            acts.append(a)
        else:
            out = pre[..., 0]
    if keep:
        return out, acts, pres
    return out


def interior_mask(M):
    m = np.zeros((M, M, M), bool)
    m[1:-1, 1:-1, 1:-1] = True
    return m


def loss(out, t):
    M = out.shape[0]
    d = np.where(interior_mask(M), out - t, 0.0)
    return float((d * d).sum() / (M - 2) ** 3)


def mse_head(t):
    """Squared-error loss against t over interior nodes: out -> (loss, dloss/dout)."""
    def head(out):
        M = out.shape[0]
        return loss(out, t), np.where(interior_mask(M), 2.0 * (out - t) / (M - 2) ** 3, 0.0)
    return head


def residual_head(u, A, dinv):
    """Residual loss of the smoother: c = out on interior nodes (0 on the
    boundary), r = u - dinv * (A c) on interior nodes, loss = sum r^2 / (M-2)^3.
    A: (M^3, M^3) sparse operator in lattice order; dinv: 1/diag, 0 on the boundary."""
    def head(out):
        M = out.shape[0]
        m = interior_mask(M).ravel()
        n = (M - 2) ** 3
        c = np.where(m, out.ravel(), 0.0)
        r = np.where(m, np.ravel(u) - dinv * (A @ c), 0.0)
        gac = -2.0 / n * dinv * r
        return float(r @ r / n), np.where(m, A.T @ gac, 0.0).reshape(out.shape)
    return head


def backward(p, u, t=None, bf16_emulation=True, head=None, feat=None):
    """(loss, gradient of the loss w.r.t. the packed parameters); the loss is
    mse_head(t) unless another head is given.

    With BF16 emulation the rounding steps are treated as the identity
    (the gradient of the emulated network in the straight-through sense)."""
    r = bf16 if bf16_emulation else (lambda z: np.asarray(z, np.float64))
    layers = unpack(p)
    out, acts, pres = forward(p, u, bf16_emulation, keep=True, feat=feat)
    lval, gout = (head or mse_head(t))(out)
    g = gout[..., None]
    grads = [None] * 4
    for l in range(3, -1, -1):
        W, b = layers[l]
        grads[l] = conv_weight_grad(g, acts[l])
        if l > 0:
            ga = conv_input_grad(g, r(W))
            g = ga * gelu_d(pres[l - 1])
    return lval, pack(grads)


def adam(p, g, m, v, step, lr, b1=0.9, b2=0.999, eps=1e-8):
    """One Adam step (step counts from 1); returns (p, m, v)."""
    m = b1 * m + (1 - b1) * g
    v = b2 * v + (1 - b2) * g * g
    mh = m / (1 - b1 ** step)
    vh = v / (1 - b2 ** step)
    return p - lr * mh / (np.sqrt(vh) + eps), m, v
