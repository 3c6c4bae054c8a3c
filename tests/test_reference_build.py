"""The DOF numbering and CSR pattern against ParMooN's own code.

tests/golden/reference_numbering.npz holds BeginIndex, GlobalNumbers, RowPtr
and KCol that the reference's mesh / FE-space / structure sources computed
(built with g++ by oracle/ref.mk, written by tests/golden/
make_numbering_fixtures.py) for the unit cube refined 1-2 times (Q1, Q2) and
3 times (Q1).  The oracle's restatement and libpmx's host set-up must equal
them bit for bit.  Where the reference build and tree are present (this
container), larger spaces are compared live: 16^3 and 32^3 Q1, 8^3 Q2.
"""
import os
import sys

import numpy as np
import pytest

import orc
import parmoon_amd as pa

HERE = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, os.path.join(HERE, "golden"))
import make_numbering_fixtures as MF  # noqa: E402

FIX = np.load(os.path.join(HERE, "golden", "reference_numbering.npz"))
REF = "/root/reference"


def check_equal(nref, order, N, act, bi, gn, rp, kc):
    sp = orc.Space(nref, order)
    assert (sp.N, sp.active) == (N, act)
    o = sp.arrays()
    d, *h = pa.host_space(nref, order)
    for name, ref, got_o, got_h in zip(("BeginIndex", "GlobalNumbers", "RowPtr", "KCol"), (bi, gn, rp, kc), o, h):
        assert np.array_equal(got_o, ref), ("oracle", name, nref, order)
        assert np.array_equal(got_h, ref), ("libpmx", name, nref, order)


@pytest.mark.parametrize("nref,order", MF.CASES)
def test_numbering_equals_reference_fixture(nref, order):
    k = f"n{nref}_q{order}"
    N, act = (int(v) for v in FIX[k + "_dims"])
    check_equal(nref, order, N, act, FIX[k + "_begin_index"], FIX[k + "_global_numbers"], FIX[k + "_rowptr"],
                FIX[k + "_kcol"])


@pytest.mark.parametrize("nref,order", [(4, 1), (3, 2), (5, 1)])
def test_numbering_equals_reference_live(nref, order):
    if not (os.path.exists(MF.EXE) and os.path.isdir(os.path.join(REF, "data"))):
        pytest.skip("the reference build (oracle/_ref) or the reference tree is not here")
    N, act, bi, gn, rp, kc = MF.run_reference(nref, order, REF)
    if (nref, order) == (4, 1):
        assert (N, len(kc)) == (4913, 92663)  # the golden log's sizes (ReadIn/tcd3d_GMG.out:386,392-394)
    check_equal(nref, order, N, act, bi, gn, rp, kc)


# ------------------------------------------------- assembly against ParMooN's own
import make_assembly_fixtures as AF  # noqa: E402

AFIX = np.load(os.path.join(HERE, "golden", "reference_assembly.npz"))
ACASES = [(name, nref, order) for name, nref, order, _ in AF.cases()]
RHS_ONLY = (8, 7)  # RhsAssemble, RhsAssemble_SUPG: no matrix


def row_rel_err(E_ref, E, rowptr):
    rows = np.repeat(np.arange(len(rowptr) - 1), np.diff(rowptr))
    scale = np.maximum.reduceat(np.abs(E_ref), rowptr[:-1])
    return float(np.max(np.abs(E - E_ref) / np.maximum(scale[rows], 1e-300)))


def rel(a, b):
    return float(np.max(np.abs(a - b)) / max(np.max(np.abs(b)), 1e-300))


@pytest.mark.parametrize("name,nref,order", ACASES, ids=[c[0] for c in ACASES])
def test_oracle_assembly_equals_reference_fixture(name, nref, order):
    """The oracle's assembly (every form, both rules, SDFEM types 0 and 2,
    variable convection) against ParMooN's Assemble3D: 1e-13 of the row's
    largest entry, rhs 1e-13 of its largest entry."""
    prm = AFIX[name + "_prm"]
    sp = orc.Space(nref, order)
    _, _, rp, _ = sp.arrays()
    two = name + "_entries2" in AFIX
    out = sp.assemble(prm, two=two)
    if int(prm[7]) not in RHS_ONLY:
        assert row_rel_err(AFIX[name + "_entries"], out[0], rp) < 1e-13
    assert rel(out[1], AFIX[name + "_rhs"]) < 1e-13
    if two:
        assert row_rel_err(AFIX[name + "_entries2"], out[2], rp) < 1e-13


@pytest.mark.parametrize("nref,order,form", [(3, 1, 2), (2, 2, 2), (2, 2, 5), (3, 2, 4)])
def test_oracle_assembly_equals_reference_live(nref, order, form):
    if not (os.path.exists(AF.EXE) and os.path.isdir(os.path.join(REF, "data"))):
        pytest.skip("the reference build (oracle/_ref) or the reference tree is not here")
    from parmoon_amd import prm as P
    prm = P.config2(nref=nref, form=form, linear=False)
    prm[P.TAU], prm[P.THETA1] = 0.02, 0.5
    E, rhs, E2 = AF.run_reference(nref, order, prm, REF)
    sp = orc.Space(nref, order)
    _, _, rp, _ = sp.arrays()
    out = sp.assemble(prm, two=E2 is not None)
    assert row_rel_err(E, out[0], rp) < 1e-13 and rel(out[1], rhs) < 1e-13
    if E2 is not None:
        assert row_rel_err(E2, out[2], rp) < 1e-13


# ---------------------------------------------- solvers against ParMooN's own
import make_solve_fixtures as SF  # noqa: E402

SFIX = np.load(os.path.join(HERE, "golden", "reference_solve.npz"))
SCASES = [c[0] for c in SF.cases() if c[0] != "q1_restart_mid"]


def oracle_solve(name, fix=SFIX):
    nref0, nlev, order = (int(v) for v in fix[name + "_levels"])
    O = orc.System(nref0, nlev, order)
    O.assemble(fix[name + "_prm"])
    rhs = O.rhs()
    fine = O.levels[-1]
    s0 = np.zeros_like(rhs)
    s0[fine.active:] = rhs[fine.active:]
    x, info = O.solve(fix[name + "_opt"], rhs, s0)
    return x, rhs, info


@pytest.mark.parametrize("name", SCASES)
def test_oracle_solve_equals_reference_fixture(name):
    """The oracle's FGMRES / Richardson with its multigrid preconditioner
    (lexicographic SOR/SSOR/Jacobi, V/W/F cycles, exact or iterative coarse
    solve, GMG damping, restarts) against ParMooN's TFgmresIte /
    TFixedPointIte + TMultiGridScaIte + TMultiGrid3D on the same matrices:
    the iterate after the same number of iterations within 1e-12 (max norm,
    relative), the rhs within 1e-13."""
    x, rhs, _ = oracle_solve(name)
    ref = SFIX[name + "_sol"]
    assert rel(rhs, SFIX[name + "_rhs"]) < 1e-13
    assert rel(x, ref) < 1e-12, rel(x, ref)


@pytest.mark.parametrize("name", ["q1_ssor", "q1_jacobi_w", "q2_curl_jacobi"])
def test_oracle_solve_equals_reference_live(name):
    if not (os.path.exists(SF.EXE) and os.path.isdir(os.path.join(REF, "data"))):
        pytest.skip("the reference build (oracle/_ref) or the reference tree is not here")
    case = {c[0]: c for c in SF.cases()}[name]
    _, nref0, nlev, order, prm, opt = case
    x, _ = SF.run_reference(nref0, nlev, order, prm, opt, REF)
    y, _, _ = oracle_solve(name)
    assert rel(y, x) < 1e-12


def test_fgmres_limit_inside_a_cycle_deviates_from_reference_on_purpose():
    """When the iteration limit falls inside a restart cycle (after m < restart
    steps), the reference back-solves all restart columns
    (UpdateGmresIterate(..., restart-1, ...), src/FE/FgmresIte.C:570), the
    last restart - m of them left over from the previous cycle.  The oracle
    (and the device) back-solve the m columns of the cycle, the FGMRES
    minimiser.  Both agree up to the last full cycle (the q1_restart case);
    here the reference's iterate differs and has the larger residual."""
    name = "q1_restart_mid"
    x, rhs, _ = oracle_solve(name)
    nref0, nlev, order = (int(v) for v in SFIX[name + "_levels"])
    O = orc.System(nref0, nlev, order)
    O.assemble(SFIX[name + "_prm"])
    fine, E = O.levels[-1], O.entries()
    ref = SFIX[name + "_sol"]
    r_ref, r_orc = fine.defect(E, ref, rhs)[1], fine.defect(E, x, rhs)[1]
    assert rel(x, ref) > 1e-6 and r_orc < r_ref, (rel(x, ref), r_orc, r_ref)
