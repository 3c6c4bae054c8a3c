"""The drop-in surfaces beyond the built-in functors, on the GPU.

* pmx_system_assemble_cb: the caller's CoeffFct3D / BoundCondFunct3D /
  BoundValueFunct3D (here Python callables through ctypes) give the same
  entries and right-hand sides as the built-in functors of the same example
  (1e-14 of the row's largest entry), for Q1 and Q2, the 27- and 64-point
  rules, Galerkin and SUPG forms; a non-Dirichlet boundary is rejected.
* pmx_iterate (TItMethod::Iterate on the caller's matrix): with the system's
  own entries it reproduces pmx_system_solve; with another caller matrix (the
  same problem at 10x the diffusion) FGMRES preconditioned by the system's MG
  levels reaches the solution of the caller's matrix (scipy direct solve).
* the TSystemTCD3D phases: AssembleARhs / AssembleSystMat / Solve /
  RestoreMassMat give the bits of pmx_tcd_step; between the phases the
  system matrix equals M + K + tau th1 A of the matrices read back; the
  reference's order checks (restore before assembling again) hold.
"""
import numpy as np
import pytest
import scipy.sparse as sps
import scipy.sparse.linalg as spla

import orc
import parmoon_amd as pa
from parmoon_amd import prm as P

pytestmark = pytest.mark.gpu
PI = np.pi


def row_rel_err(E_ref, E, rowptr):
    rows = np.repeat(np.arange(len(rowptr) - 1), np.diff(rowptr))
    scale = np.maximum.reduceat(np.abs(E_ref), rowptr[:-1])
    return float(np.max(np.abs(E - E_ref) / scale[rows]))


def rel(a, b):
    return float(np.max(np.abs(a - b)) / max(np.max(np.abs(b)), 1e-300))


def sine_cd_callbacks(eps, b, c):
    """Examples/CD_3D/Laplace.h's callbacks generalised to constant b, c (the
    built-in PMX_EX_SINE_CD functor): u = sin sin sin, f = -eps Lap u + b.grad u + c u."""
    def coeffs(x, y, z):
        sx, sy, sz = np.sin(PI * x), np.sin(PI * y), np.sin(PI * z)
        cx, cy, cz = np.cos(PI * x), np.cos(PI * y), np.cos(PI * z)
        u = sx * sy * sz
        f = 3 * PI * PI * eps * u + PI * (b[0] * cx * sy * sz + b[1] * sx * cy * sz + b[2] * sx * sy * cz) + c * u
        n = len(x)
        return np.column_stack([np.full(n, eps), np.full(n, b[0]), np.full(n, b[1]), np.full(n, b[2]),
                                np.full(n, c), f])

    def cond(comp, x, y, z):
        return 0  # DIRICHLET

    def value(comp, x, y, z):
        return np.sin(PI * x) * np.sin(PI * y) * np.sin(PI * z)

    return coeffs, cond, value


CB_CASES = [("Q1-sdfem", 1, 3, P.FORM_SDFEM, False), ("Q1-galerkin", 1, 3, P.FORM_GALERKIN, False),
            ("Q2-sdfem-27pt", 2, 2, P.FORM_SDFEM, True), ("Q2-sdfem-64pt", 2, 2, P.FORM_SDFEM, False),
            ("Q2-supgAK", 2, 2, P.FORM_TCD_SUPG_AK, False)]


@pytest.mark.parametrize("name,order,nref,form,linear", CB_CASES, ids=[c[0] for c in CB_CASES])
def test_callbacks_match_builtin_functor(ctx, name, order, nref, form, linear):
    print(f"[case] callbacks {name}", flush=True)
    b, c = np.array([1.0, 2.0, 3.0]) / np.sqrt(14.0), 0.3
    prm = P.base(P.EX_SINE_CD, eps=1e-3, form=form, b=b, c=c, linear=linear)
    prm[P.TAU], prm[P.THETA1] = 1e-2, 0.5
    S = pa.System(ctx, nref - 1, 2, order)
    S.assemble(prm)
    E0, r0 = [S.entries(l) for l in (0, 1)], [S.rhs(l) for l in (0, 1)]
    K0 = S.entries(1, which=1) if form == P.FORM_TCD_SUPG_AK else None
    S.assemble_cb(prm, *sine_cd_callbacks(1e-3, b, c))
    for l in (0, 1):
        rp, _ = S.structure(l)
        assert row_rel_err(E0[l], S.entries(l), rp) < 1e-14, l
        assert rel(S.rhs(l), r0[l]) < 1e-14, l
    if K0 is not None:
        rp, _ = S.structure(1)
        assert row_rel_err(K0, S.entries(1, which=1), rp) < 1e-14
    S.close()


def test_callback_rejects_neumann(ctx):
    S = pa.System(ctx, 2, 1, 1)
    coeffs, _, value = sine_cd_callbacks(1.0, (0.0, 0.0, 0.0), 0.0)
    with pytest.raises(pa.PmxError):
        S.assemble_cb(P.base(P.EX_SINE_CD, eps=1.0), coeffs, lambda comp, x, y, z: 1 if comp == 3 else 0, value)
    S.close()


def test_iterate_on_caller_matrix(ctx):
    print("[case] pmx_iterate", flush=True)
    prm = P.config1b()
    S = pa.System(ctx, 2, 3, 1)  # 4^3 .. 16^3 Q1
    S.assemble(prm)
    rp, kc = S.structure()
    E = S.entries()
    rhs = S.rhs()
    na = int(S.dims()[1])
    sol0 = np.zeros_like(rhs)
    sol0[na:] = rhs[na:]
    opt = P.solver_opt(solver=16, restart=20, maxit=200, res_norm_min=0.0, red_factor=0.0, rel_target=1e-12,
                       smoother=3, omega=1.0)
    s_sys, _, rc = S.solve(opt, rhs, sol0)
    s_it, info, rc2 = S.iterate(rp, kc, E, opt, rhs, sol0)
    assert rc == 0 and rc2 == 0 and rel(s_it, s_sys) < 1e-10
    # another matrix: the same problem at 10x the diffusion, entries from the oracle
    prm2 = prm.copy()
    prm2[P.EPS] = 10 * prm[P.EPS]
    E2, r2 = orc.Space(4, 1).assemble(prm2)
    A2 = sps.csr_matrix((E2, kc, rp), shape=(len(rhs), len(rhs)))
    exact = spla.spsolve(A2.tocsc(), r2)
    s2 = np.zeros_like(r2)
    s2[na:] = r2[na:]
    for precision in (0, 1):
        x, info, rc = S.iterate(rp, kc, E2, P.solver_opt(**dict(solver=16, restart=20, maxit=300, res_norm_min=0.0,
                                                                  red_factor=0.0, rel_target=1e-12, smoother=3,
                                                                  omega=1.0, precision=precision)), r2, s2)
        assert rc == 0 and rel(x, exact) < 1e-9, (precision, rel(x, exact), info)
    S.close()


def test_tcd_phases(ctx):
    print("[case] TSystemTCD3D phases", flush=True)
    prm = P.rotating(seed=2, form=P.FORM_TCD_SUPG_AK)
    td = [1e-2, .5, .5, .5, .5]
    opt = P.solver_opt(solver=16, restart=50, maxit=200, res_norm_min=0.0, red_factor=0.0, rel_target=1e-12,
                       smoother=3, omega=0.5)
    A = pa.System(ctx, 1, 3, 1)
    B = pa.System(ctx, 1, 3, 1)
    for S in (A, B):
        S.tcd_init(prm, None)
    t = 0.0
    for m in range(3):
        t += td[0]
        p = prm.copy()
        p[P.TIME] = t
        A.tcd_step(p, td, opt)
        B.tcd_assemble_arhs(p, td)
        B.tcd_assemble_systmat()
        with pytest.raises(pa.PmxError):
            B.tcd_assemble_systmat()  # not restored yet
        rp, _ = B.structure()
        na = int(B.dims()[1])
        act = slice(0, int(rp[na]))  # the active rows (Dirichlet rows of every matrix are identity rows)
        sysm = B.entries()
        M, K, Ast = B.entries(which=2), B.entries(which=1), B.entries(which=3)
        assert row_rel_err((M + K + td[0] * td[1] * Ast)[act], sysm[act], rp[:na + 1]) < 1e-15
        B.tcd_solve(opt)
        B.tcd_restore_massmat()
        with pytest.raises(pa.PmxError):
            B.tcd_restore_massmat()  # nothing to restore
        assert np.array_equal(A.solution(), B.solution()), m
    A.close()
    B.close()


def test_caller_level_matrices(ctx):
    """pmx_level_set_entries: the operators of every MG level replaced by the
    caller's matrices (TMGLevel3D(level, A, ...) + TMultiGrid3D::AddLevel):
    a system assembled with the Galerkin Laplace form, then given the
    oracle's SUPG matrices of its levels, solves like the system assembled
    with SUPG on the device (FP64 and FP32 forms); a matrix of another
    pattern and a non-identity Dirichlet row are rejected."""
    print("[case] caller level matrices", flush=True)
    prm = P.config1b()
    O = orc.System(1, 3, 1)
    O.assemble(prm)
    S = pa.System(ctx, 1, 3, 1)
    S.assemble(P.base(P.EX_LAPLACE, eps=1.0))
    for lv in range(3):
        S.set_entries(O.entries(lv), lv)
    R = pa.System(ctx, 1, 3, 1)
    R.assemble(prm)
    rhs = O.rhs()
    fine = O.levels[-1]
    s0 = np.zeros_like(rhs)
    s0[fine.active:] = rhs[fine.active:]
    for prec in (0, 1):
        opt = P.solver_opt(solver=16, restart=20, maxit=200, res_norm_min=0.0, red_factor=0.0, smoother=3,
                           omega=1.0, precision=prec, rel_target=1e-11)
        a, ia, _ = S.solve(opt, rhs, s0)
        b, ib, _ = R.solve(opt, rhs, s0)
        assert int(ia[0]) == int(ib[0]) and np.max(np.abs(a - b)) <= 1e-10 * np.max(np.abs(b)), (prec, ia, ib)
    # the exported operator is the caller's matrix
    assert np.max(np.abs(S.entries(level=1) - O.entries(1))) == 0.0
    rp1, kc1 = S.structure(1)
    with pytest.raises(pa.PmxError):
        S.set_entries(O.entries(1), level=2, rowptr=rp1, kcol=kc1)
    bad = O.entries(2).copy()
    rp, _ = S.structure(2)
    bad[rp[fine.active]] = 2.0
    with pytest.raises(pa.PmxError):
        S.set_entries(bad, level=2)
    S.close()
    R.close()


def test_dispatch_timings(ctx):
    """pmx_time_kernel's per-dispatch timings (10 colour launch, 11 SpMV) and
    the calibration dispatches of the colour launch (12 empty, 13 streaming
    read of each colour's values) in every precision return positive
    durations; timing before assembly and unknown precisions are errors."""
    S = pa.System(ctx, 1, 3, 2)
    with pytest.raises(pa.PmxError):
        S.time_kernel(10, -1, 1, reps=1)
    S.assemble(P.config1b())
    for k in (10, 11, 12, 13):
        for prec in (0, 1, 3):
            if k == 11 and prec == 3:
                continue
            ms = S.time_kernel(k, -1, prec, reps=2)
            assert 0.0 < ms < 50.0, (k, prec, ms)
    with pytest.raises(pa.PmxError):
        S.time_kernel(12, -1, 2, reps=1)
    S.close()
