"""HIP path (libpmx through the C-ABI) against the CPU oracle on identical inputs.

Tolerances: integer structure bit-exact; assembled entries and rhs within
1e-12 of the row's largest oracle entry (entries that cancel to zero make a
pure per-entry relative bound meaningless); SpMV / defect / smoothers /
transfers within 1e-13 relative (norm-wise); converged solutions within 1e-10
relative in the max norm.
"""
import numpy as np
import pytest

import orc
import parmoon_amd as pa
from parmoon_amd import prm as P

pytestmark = pytest.mark.gpu


def row_rel_err(E_ref, E, rowptr):
    rows = np.repeat(np.arange(len(rowptr) - 1), np.diff(rowptr))
    scale = np.maximum.reduceat(np.abs(E_ref), rowptr[:-1])
    return float(np.max(np.abs(E - E_ref) / scale[rows]))


def rel(a, b):
    return float(np.max(np.abs(a - b)) / max(np.max(np.abs(b)), 1e-300))


CASES = [
    ("laplace-Q1-galerkin", 1, 3, lambda: P.base(P.EX_LAPLACE, eps=1e-2)),
    ("config1-Q1-sdfem", 1, 3, lambda: P.config1()),
    ("config1-Q1-galerkin", 1, 3, lambda: P.config1(form=P.FORM_GALERKIN)),
    ("heat-Q1-mass", 1, 3, lambda: P.heat_amc(0.3, form=P.FORM_MASS)),
    ("heat-Q1-tcdA", 1, 3, lambda: P.heat_amc(0.3)),
    ("curl-Q2-sdfem-27pt", 2, 2, lambda: P.config2(nref=2)),
    ("curl-Q2-sdfem-64pt", 2, 2, lambda: P.config2(nref=2, linear=False)),
    ("config1-Q2-galerkin", 2, 2, lambda: P.config1(form=P.FORM_GALERKIN)),
    ("config1-Q2-sdfem-inline", 2, 2, lambda: P.config1()),
    ("heat-Q2-mass", 2, 2, lambda: P.heat_amc(0.3, form=P.FORM_MASS)),
    ("heat-Q2-tcdA", 2, 2, lambda: P.heat_amc(0.3)),
    ("rotating-Q2-supgAK", 2, 2, lambda: dict_supg(P.rotating())),
    ("lognormal-Q1-sdfem", 1, 3, lambda: P.lognormal(nref=3, peclet=10.0)),
    ("rotating-Q1-supgAK", 1, 3, lambda: dict_supg(P.rotating())),
    ("config1-Q1-sdfem-type0", 1, 3, lambda: with_sdfem_type0(P.config1())),
    ("config1-Q2-sdfem-type0-inline", 2, 2, lambda: with_sdfem_type0(P.config1())),
    ("curl-Q2-sdfem-type0", 2, 2, lambda: with_sdfem_type0(P.config2(nref=2))),
    ("rotating-Q1-mass-supg", 1, 3, lambda: with_form(P.rotating(), P.FORM_MASS_SUPG)),
    ("config1-Q2-mass-supg-inline", 2, 2, lambda: with_form(P.config1(), P.FORM_MASS_SUPG)),
    ("curl-Q2-mass-supg", 2, 2, lambda: with_form(P.config2(nref=2), P.FORM_MASS_SUPG)),
]


def with_form(p, form):
    p = p.copy()
    p[P.FORM] = form
    return p


def with_sdfem_type0(p):
    """SDFEM_TYPE = 0 (ConvDiff.C:568-590) with delta1 != delta0 so both branches differ."""
    p = p.copy()
    p[P.SDFEM_TYPE] = 0
    p[P.DELTA1] = 0.6
    return p


def dict_supg(p):
    p = p.copy()
    p[P.FORM] = P.FORM_TCD_SUPG_AK
    p[P.TAU] = 1e-2
    p[P.THETA1] = 0.5
    return p


@pytest.mark.parametrize("name,order,nref,mk", CASES, ids=[c[0] for c in CASES])
def test_assembly_matches_oracle(ctx, name, order, nref, mk):
    prm = mk()
    print(f"[case] assembly {name} order={order} nref={nref}", flush=True)
    sp = orc.Space(nref, order)
    two = prm[P.FORM] == P.FORM_TCD_SUPG_AK
    out = sp.assemble(prm, two=two)
    S = pa.System(ctx, nref, 1, order)
    S.assemble(prm)
    rp, kc = S.structure()
    bi, gn, rp0, kc0 = sp.arrays()
    assert np.array_equal(rp, rp0) and np.array_equal(kc, kc0)
    E = S.entries()
    assert row_rel_err(out[0], E, rp) < 1e-12
    rhs = S.rhs()
    assert rel(rhs, out[1]) < 1e-12
    if two:
        assert row_rel_err(out[2], S.entries(which=1), rp) < 1e-12
    S.close()


RHS_CASES = [
    ("config1-Q1-rhs-supg", 1, 3, lambda: P.config1()),
    ("config1-Q2-rhs-supg-inline", 2, 2, lambda: P.config1()),
    ("curl-Q2-rhs-supg", 2, 2, lambda: P.config2(nref=2)),
    ("lognormal-Q1-rhs-supg", 1, 3, lambda: P.lognormal(nref=3, peclet=10.0)),
]


@pytest.mark.parametrize("name,order,nref,mk", RHS_CASES, ids=[c[0] for c in RHS_CASES])
@pytest.mark.parametrize("form", [P.FORM_RHS, P.FORM_RHS_SUPG])
def test_rhs_forms_match_oracle(ctx, name, order, nref, mk, form):
    """RhsAssemble / RhsAssemble_SUPG (TCD3D.C:374-446): the right-hand side of
    every level is re-assembled and the matrices stay as they were."""
    prm = mk()
    print(f"[case] rhs form {form} {name} order={order} nref={nref}", flush=True)
    S = pa.System(ctx, nref - 1, 2, order)
    S.assemble(prm)
    E_before = S.entries()
    prm_r = with_form(prm, form)
    S.assemble(prm_r)
    assert np.array_equal(S.entries(), E_before)
    for lvl, nr in ((0, nref - 1), (1, nref)):
        sp = orc.Space(nr, order)
        _, r_ref = sp.assemble(prm_r)
        assert rel(S.rhs(lvl), r_ref) < 1e-12, lvl
    S.close()


def test_jacobi_mg_laplace(ctx):
    """damped-Jacobi V(2,2) GMG-FGMRES on the Laplace example (Laplace.h)."""
    prm = P.base(P.EX_LAPLACE, eps=1.0)
    S = pa.System(ctx, 1, 3, 1)
    S.assemble(prm)
    O = orc.System(1, 3, 1)
    O.assemble(prm)
    rhs = O.rhs()
    ref = reference_solution(O, rhs, 1)
    fine = O.levels[-1]
    sol0 = np.zeros_like(rhs)
    sol0[fine.active:] = rhs[fine.active:]
    opt = P.solver_opt(solver=16, restart=20, maxit=100, res_norm_min=0.0, red_factor=1e-13, smoother=1, omega=0.8)
    s, info, rc = S.solve(opt, rhs, sol0)
    assert rc == 0 and rel(s, ref) < 1e-10
    S.close()


def test_numbering_through_device(ctx):
    for order, nref in [(1, 3), (2, 2)]:
        sp = orc.Space(nref, order)
        bi0, gn0, _, _ = sp.arrays()
        S = pa.System(ctx, nref, 1, order)
        bi, gn = S.numbering()
        assert np.array_equal(bi, bi0) and np.array_equal(gn, gn0)
        S.close()


@pytest.fixture(scope="module")
def q1sys(ctx):
    prm = P.config1b()
    S = pa.System(ctx, 1, 3, 1)     # 2^3 .. 8^3 Q1
    S.assemble(prm)
    O = orc.System(1, 3, 1)
    O.assemble(prm)
    yield S, O, prm
    S.close()


@pytest.fixture(scope="module")
def q2sys(ctx):
    prm = P.config2(nref=3)
    S = pa.System(ctx, 1, 3, 2)     # 2^3 .. 8^3 Q2
    S.assemble(prm)
    O = orc.System(1, 3, 2)
    O.assemble(prm)
    yield S, O, prm
    S.close()


@pytest.mark.parametrize("which", ["q1sys", "q2sys"])
def test_blas2_and_smoothers(request, which):
    S, O, prm = request.getfixturevalue(which)
    fine = O.levels[-1]
    E = O.entries()
    rng = np.random.default_rng(1)
    x = rng.standard_normal(fine.N)
    b = rng.standard_normal(fine.N)
    print(f"[case] blas2 {which}", flush=True)
    assert rel(S.matvec(x), fine.matvec(E, x)) < 1e-13
    assert rel(S.matvec(x, precision=1), fine.matvec(E, x)) < 1e-6
    r, res = S.defect(x, b)
    r0, res0 = fine.defect(E, x, b)
    assert rel(r, r0) < 1e-13 and abs(res - res0) < 1e-12 * res0
    # damped Jacobi sweep: identical update
    assert rel(S.smooth(1, 0.8, 1, x, b), fine.smooth(E, 1, 0.8, x, b)) < 1e-13
    # multicolour SSOR equals SSOR with the rows in colour order
    col = S.colours()
    order = np.argsort(col[:fine.active], kind="stable").astype(np.int32)
    for kind in (2, 3):
        g = S.smooth(kind, 1.0, 1, x, b)
        o = fine.smooth_ordered(E, kind, 1.0, x, b, order)
        assert rel(g, o) < 1e-12, kind
    # generic CSR SpMV on the reference arrays
    rp, kc = S.structure()
    assert rel(S.ctx.csr_matvec(rp, kc, E, x), fine.matvec(E, x)) < 1e-13


@pytest.mark.parametrize("which", ["q1sys", "q2sys"])
def test_transfers(request, which):
    S, O, prm = request.getfixturevalue(which)
    rng = np.random.default_rng(2)
    for fl in (1, 2):
        c, f = O.levels[fl - 1], O.levels[fl]
        xc = rng.standard_normal(c.N)
        rf = rng.standard_normal(f.N)
        print(f"[case] transfers {which} fine_level={fl}", flush=True)
        assert rel(S.prolongate(xc, fl), orc.prolongate(c, f, xc)) < 1e-13
        assert rel(S.restrict(rf, fl), orc.restrict(c, f, rf)) < 1e-13


def reference_solution(O, rhs, smoother):
    opt = P.solver_opt(solver=16, restart=30, maxit=300, res_norm_min=0.0, red_factor=1e-13,
                       smoother=smoother, omega=0.8 if smoother == 1 else 1.0)
    sol0 = np.zeros_like(rhs)
    fine = O.levels[-1]
    sol0[fine.active:] = rhs[fine.active:]
    s, info = O.solve(opt, rhs, sol0)
    return s


@pytest.mark.parametrize("which,smoother", [("q1sys", 3), ("q1sys", 2), ("q2sys", 3)])
@pytest.mark.parametrize("precision", [0, 1])
def test_fgmres_mg_solution(request, which, smoother, precision):
    S, O, prm = request.getfixturevalue(which)
    rhs = O.rhs()
    ref = reference_solution(O, rhs, smoother)
    fine = O.levels[-1]
    r0 = np.linalg.norm(fine.defect(O.entries(), np.where(np.arange(fine.N) < fine.active, 0.0, rhs), rhs)[0])
    opt = P.solver_opt(solver=16, restart=20, maxit=200, res_norm_min=1e-12 * r0, smoother=smoother,
                       omega=0.8 if smoother == 1 else 1.0, precision=precision)
    sol0 = np.zeros_like(rhs)
    sol0[fine.active:] = rhs[fine.active:]
    print(f"[case] fgmres {which} smoother={smoother} precision={precision}", flush=True)
    s, info, rc = S.solve(opt, rhs, sol0)
    assert rc == 0, info
    assert info[1] <= 1e-12 * r0 * 1.0001
    assert rel(s, ref) < 1e-10


@pytest.mark.parametrize("solver", [11, 13])
def test_other_krylov(q1sys, solver):
    S, O, prm = q1sys
    rhs = O.rhs()
    ref = reference_solution(O, rhs, 3)
    fine = O.levels[-1]
    sol0 = np.zeros_like(rhs)
    sol0[fine.active:] = rhs[fine.active:]
    opt = P.solver_opt(solver=solver, maxit=400, res_norm_min=0.0, red_factor=1e-13, smoother=3, omega=1.0)
    print(f"[case] krylov solver={solver}", flush=True)
    s, info, rc = S.solve(opt, rhs, sol0)
    assert rc == 0, info
    assert rel(s, ref) < 1e-9
    o, oinfo = O.solve(opt, rhs, sol0)
    assert rel(o, ref) < 1e-9


def test_graph_replays_match_eager(q2sys):
    """The MG cycle is replayed from a captured HIP graph: several replays,
    with eager work and synchronisations in between, give the eager result."""
    S, O, prm = q2sys
    rhs = O.rhs()
    fine = O.levels[-1]
    sol0 = np.zeros_like(rhs)
    sol0[fine.active:] = rhs[fine.active:]
    for precision in (0, 1):
        base = dict(solver=16, restart=20, maxit=200, res_norm_min=0.0, red_factor=0.0, rel_target=1e-10,
                    smoother=3, omega=1.0, precision=precision)
        eager, _, rc0 = S.solve(P.solver_opt(eager=1, **base), rhs, sol0)
        assert rc0 == 0
        for rep in range(3):
            g, _, rc = S.solve(P.solver_opt(**base), rhs, sol0)
            assert rc == 0
            assert np.array_equal(g, eager), (precision, rep)
            S.matvec(g)  # eager work between replays


WG_CASES = [(o, n, p, sm) for o, n in ((1, 4), (2, 4)) for p, sm in ((0, 2), (0, 3), (1, 3), (1, 2))]


@pytest.mark.parametrize("order,nref,precision,smoother", WG_CASES)
def test_wg_sweep_matches_launch_per_colour(ctx, order, nref, precision, smoother):
    """SOR and SSOR sweeps run in one workgroup (all colours, workgroup
    barriers between them) give the launch-per-colour iterate bit for bit:
    FGMRES-GMG solves with the single-workgroup sweep forced on every level
    against none, FP64 and FP32.  The finest levels here (17^3 Q1: 422 rows per
    colour, 33^3 Q2: 1103) need several rows per lane and mix residue classes
    of different stencil shape inside a wavefront."""
    print(f"[case] wg sweep order={order} nref={nref} precision={precision} smoother={smoother}", flush=True)
    prm = P.config2(nref=nref) if order == 2 else P.config1b()
    S = pa.System(ctx, nref - 2, 3, order)
    S.assemble(prm)
    rhs = S.rhs(-1)
    n, na = int(S.dims(-1)[0]), int(S.dims(-1)[1])
    sol0 = np.zeros(n)
    sol0[na:] = rhs[na:]
    # (the FP32 form with forward SOR on the Q2 SUPG problem needs more than
    # 60 iterations; it used to end in NaN when the iteration limit cut an
    # FGMRES cycle short, see test_fgmres_limit_inside_a_cycle)
    base = dict(solver=16, restart=20, maxit=200, res_norm_min=0.0, red_factor=0.0, rel_target=1e-10,
                smoother=smoother, omega=1.0, precision=precision)
    a, ia, rca = S.solve(P.solver_opt(wg_sweep_rows=-1, **base), rhs, sol0, check=False)
    b, ib, rcb = S.solve(P.solver_opt(wg_sweep_rows=1 << 30, **base), rhs, sol0, check=False)
    assert rca == 0 and rcb == 0, (precision, smoother, rca, rcb)
    assert ia[0] == ib[0], (precision, smoother, ia[0], ib[0])
    assert np.array_equal(a, b), (precision, smoother, float(np.max(np.abs(a - b))))
    S.close()


@pytest.mark.parametrize("precision", [0, 1])
def test_fgmres_limit_inside_a_cycle(q2sys, precision):
    """An iteration limit that ends an FGMRES cycle before its restart length
    updates the iterate with the columns the cycle filled: a finite iterate
    whose FP64 residual is below the initial one, for every limit (the FP32
    form's inner solves get the remaining budget, which ends cycles early)."""
    S, O, prm = q2sys
    rhs = O.rhs()
    fine = O.levels[-1]
    sol0 = np.zeros_like(rhs)
    sol0[fine.active:] = rhs[fine.active:]
    last = None
    for maxit in (1, 3, 7, 13, 20, 27):
        print(f"[case] fgmres limit {maxit} precision {precision}", flush=True)
        opt = P.solver_opt(solver=16, restart=20, maxit=maxit, res_norm_min=0.0, red_factor=0.0, rel_target=1e-10,
                           smoother=2, omega=1.0, precision=precision)
        s, info, rc = S.solve(opt, rhs, sol0, check=False)
        assert rc in (0, 4) and np.all(np.isfinite(s)), (maxit, rc, info)
        assert np.isfinite(info[1]) and info[1] < info[2], (maxit, info)
        r, res = S.defect(s, rhs)
        # (the FP64 form reports the Givens estimate, which follows the true
        # residual to rounding: 1e-14 of the initial residual near convergence)
        assert abs(res - info[1]) <= 1e-6 * res + 1e-14 * info[2], (maxit, res, info)
        last = res
    assert last < 1e-3 * info[2]


@pytest.mark.parametrize("cycle", [2, 0])
@pytest.mark.parametrize("which", ["q1sys", "q2sys"])
def test_w_and_f_cycles_match_oracle(request, which, cycle):
    """W (SC_MG_CYCLE_SCALAR = 2) and F (0) cycles, TMultiGrid3D::SetRecursion
    (MultiGrid3D.C:718-729) and the F-cycle's recursion reset (:620): one
    damped-Jacobi cycle (independent of the row order) from a random right-hand
    side equals the oracle's cycle to 1e-12 and differs from the V-cycle; on the
    Q1 system FGMRES with the cycle reaches the oracle's solution to 1e-10
    (damped Jacobi does not converge on the Q2 SUPG operator, in the oracle
    either)."""
    S, O, prm = request.getfixturevalue(which)
    fine = O.levels[-1]
    rng = np.random.default_rng(10 + cycle)
    rhs = rng.standard_normal(fine.N)
    sol0 = np.zeros(fine.N)
    sol0[fine.active:] = rhs[fine.active:]
    print(f"[case] cycle {cycle} {which}", flush=True)
    one = dict(solver=0, smoother=1, omega=0.8, pre=2, post=2, coarse=17)
    g, _, rc = S.solve(P.solver_opt(cycle=cycle, **one), rhs, sol0)
    o, _ = O.solve(P.solver_opt(cycle=cycle, **one), rhs, sol0)
    v, _ = O.solve(P.solver_opt(cycle=1, **one), rhs, sol0)
    assert rc == 0 and rel(g, o) < 1e-12
    assert rel(o, v) > 1e-6
    if which != "q1sys":
        return
    rhs = O.rhs()
    sol0[fine.active:] = rhs[fine.active:]
    sol0[:fine.active] = 0.0
    r0 = fine.defect(O.entries(), sol0, rhs)[1]
    full = P.solver_opt(solver=16, restart=20, maxit=200, res_norm_min=1e-12 * r0, red_factor=0.0, smoother=1,
                        omega=0.8, cycle=cycle)
    g, info, rc = S.solve(full, rhs, sol0)
    o, _ = O.solve(full, rhs, sol0)
    assert rc == 0 and rel(g, o) < 1e-10


def test_ilu_smoother_is_rejected(q1sys):
    """Smoother id 4 (ILU in MultiGrid3D.C's switches) is not built: the solve
    reports an invalid argument instead of substituting another smoother."""
    S, O, prm = q1sys
    rhs = O.rhs()
    for kw in (dict(smoother=4), dict(coarse=4)):
        _, _, rc = S.solve(P.solver_opt(**dict(dict(solver=16, maxit=10, smoother=3), **kw)), rhs,
                           np.zeros_like(rhs), check=False)
        assert rc == 1, kw


def test_coarse_inverse_follows_the_operator(ctx):
    """The coarsest level's inverse is rebuilt only when its operator changed
    bit for bit: re-assembling with other coefficients must give the solve of
    a fresh system with those coefficients, and going back must give the
    first solve again, bit for bit."""
    print("[case] coarse inverse cache", flush=True)
    pa_, pb_ = P.config2(nref=4), P.config2(nref=4)
    pb_[P.EPS] = 10 * pa_[P.EPS]
    opt = P.solver_opt(solver=16, restart=20, maxit=100, res_norm_min=0.0, red_factor=0.0, rel_target=1e-10,
                       smoother=3, omega=1.0)

    def solve(S, prm):
        S.assemble(prm)
        info, rc = S.solve_resident(opt)
        assert rc == 0, info
        return S.solution()

    S = pa.System(ctx, 2, 3, 2)
    xa = solve(S, pa_)
    xa2 = solve(S, pa_)
    xb = solve(S, pb_)
    xa3 = solve(S, pa_)
    F = pa.System(ctx, 2, 3, 2)
    xb_fresh = solve(F, pb_)
    assert np.array_equal(xa, xa2) and np.array_equal(xa, xa3)
    assert np.array_equal(xb, xb_fresh)
    assert not np.array_equal(xa, xb)
    S.close()
    F.close()


@pytest.mark.parametrize("order", [1, 2])
def test_device_error_norms(ctx, order):
    """L2 and H1-semi errors computed on the device (pmx_system_errors) agree
    with the oracle's quadrature of the same solution to 1e-11, for both rules,
    and fall at the element's rates on the Laplace example (u = sin sin sin):
    L2 O(h^{k+1}), H1 O(h^k)."""
    print(f"[case] device error norms order={order}", flush=True)
    errs = []
    for nref in (3, 4):
        for linear in (True, False):
            prm = P.base(P.EX_LAPLACE, eps=1.0, linear=linear)
            S = pa.System(ctx, nref - 2, 3, order)
            S.assemble(prm)
            opt = P.solver_opt(solver=16, restart=20, maxit=100, res_norm_min=0.0, red_factor=0.0,
                               rel_target=1e-12, smoother=3, omega=1.0)
            info, rc = S.solve_resident(opt)
            assert rc == 0, info
            O = orc.System(nref - 2, 3, order)
            e_or = O.levels[-1].errors(prm, S.solution())
            e_dev = S.errors(prm)
            assert abs(e_dev[0] - e_or[0]) <= 1e-11 * e_or[0], (e_dev, e_or)
            assert abs(e_dev[1] - e_or[1]) <= 1e-11 * e_or[1], (e_dev, e_or)
            if linear:
                errs.append(e_dev)
            S.close()
    r0, r1 = errs[0][0] / errs[1][0], errs[0][1] / errs[1][1]
    assert abs(np.log2(r0) - (order + 1)) < 0.25 and abs(np.log2(r1) - order) < 0.25, (r0, r1)


@pytest.mark.parametrize("which", ["q1sys", "q2sys"])
@pytest.mark.parametrize("mode", [1, 2])
def test_step_length_control_cycle(request, which, mode):
    """One damped-Jacobi V(2,2) cycle with MG step length control
    (TMGLevel3D::StepLengthControl) on the finest level (mode 1) or every
    level above the coarsest (mode 2), FP64, against the oracle's cycle
    (Jacobi sweeps do not depend on the row order)."""
    S, O, prm = request.getfixturevalue(which)
    print(f"[case] step length control {which} mode {mode}", flush=True)
    rhs = O.rhs()
    fine = O.levels[-1]
    x0 = np.zeros(fine.N)
    x0[fine.active:] = rhs[fine.active:]
    opt = P.solver_opt(solver=0, smoother=1, omega=0.8, pre=2, post=2)
    O.set_slc(mode)
    ref, _ = O.solve(opt, rhs, x0)
    O.set_slc(0)
    plain, _ = O.solve(opt, rhs, x0)
    S.set_step_length_control(mode)
    try:
        s, info, rc = S.solve(opt, rhs, x0)
    finally:
        S.set_step_length_control(0)
    assert rc == 0
    assert rel(s, ref) < 1e-12, rel(s, ref)
    assert rel(plain, ref) > 1e-6
    with pytest.raises(pa.PmxError):
        S.set_step_length_control(3)


@pytest.mark.parametrize("which", ["q1sys", "q2sys"])
@pytest.mark.parametrize("coarse", [1, 3, 6])
def test_iterative_coarse_solve(request, which, coarse):
    """Coarsest level smoothed until its defect drops by SC_COARSE_RED_FACTOR
    (0.1) or SC_COARSE_MAXIT (10) sweeps (TMultiGrid3D CoarseSmooth).  Damped
    Jacobi everywhere: one cycle equals the oracle's (the sweeps do not depend
    on the row order); SSOR on the coarsest level (id 3, or 6, which the
    switch's default turns into SSOR): the FGMRES solutions (to a relative
    residual of 1e-10) agree to 1e-8."""
    S, O, prm = request.getfixturevalue(which)
    print(f"[case] iterative coarse solve {which} coarse {coarse}", flush=True)
    rhs = O.rhs()
    fine = O.levels[-1]
    x0 = np.zeros(fine.N)
    x0[fine.active:] = rhs[fine.active:]
    if coarse == 1:
        opt = P.solver_opt(solver=0, smoother=1, omega=0.8, pre=2, post=2, coarse=1, coarse_sweeps=10)
        ref, _ = O.solve(opt, rhs, x0)
        exact, _ = O.solve(P.solver_opt(solver=0, smoother=1, omega=0.8, pre=2, post=2), rhs, x0)
        s, info, rc = S.solve(opt, rhs, x0)
        assert rc == 0 and rel(s, ref) < 1e-12, rel(s, ref)
        assert rel(exact, ref) > 1e-8
    else:
        opt = P.solver_opt(solver=16, restart=20, maxit=200, res_norm_min=0.0, red_factor=1e-10, smoother=3,
                           omega=1.0, coarse=coarse, coarse_sweeps=10)
        ref, _ = O.solve(opt, rhs, x0)
        s, info, rc = S.solve(opt, rhs, x0)
        assert rc == 0 and rel(s, ref) < 1e-8, rel(s, ref)


@pytest.mark.parametrize("which", ["q1sys", "q2sys"])
def test_reference_input_file_solver(request, which):
    """The solver settings of the reference's ReadIn/Input_CD3D.dat: fixed
    point iteration (SC_SOLVER_SCALAR 11) preconditioned by a V-cycle with
    smoother id 6 (the switch's default, SOR) and 5 pre- and post-sweeps,
    coarse id 6 (SSOR) with at most 10 sweeps and reduction 0.1; solved to a
    relative residual of 1e-10, the solution agrees with the oracle's run of
    the same settings."""
    S, O, prm = request.getfixturevalue(which)
    print(f"[case] Input_CD3D.dat solver settings {which}", flush=True)
    rhs = O.rhs()
    fine = O.levels[-1]
    x0 = np.zeros(fine.N)
    x0[fine.active:] = rhs[fine.active:]
    opt = P.solver_opt(solver=11, maxit=200, res_norm_min=0.0, red_factor=1e-10, smoother=6, pre=5, post=5,
                       omega=1.0, coarse=6, coarse_sweeps=10, coarse_red=0.1)
    ref, oinfo = O.solve(opt, rhs, x0)
    s, info, rc = S.solve(opt, rhs, x0)
    assert rc == 0 and oinfo[1] <= 1e-10 * oinfo[2]
    assert rel(s, ref) < 1e-8, rel(s, ref)
