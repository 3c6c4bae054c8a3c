"""The C++ classes of include/pmx_parmoon.hpp (ParMooN's TFESpace3D /
TSquareStructure3D / TSystemCD3D / TSystemTCD3D names over the C-ABI), driven
from C++ programs written like the reference's drivers.

Host part (no GPU): numbering and CSR bit-exact with the oracle, errors
reported as pmx::Error.  GPU part: assembly, FGMRES solve and three
Crank-Nicolson steps through the classes against the oracle (entries and rhs
within 1e-12 of the row's largest entry, solutions within 1e-10 / 1e-9).
"""
import os
import subprocess
import tempfile

import numpy as np
import pytest

import orc
from parmoon_amd import prm as P

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
BUILD = os.path.join(ROOT, "tests", "cpp", "_build")


def build(name, with_oracle=False):
    """Compile tests/cpp/<name>.cpp with g++ (done by the package build; here
    only when the binary is missing)."""
    exe = os.path.join(BUILD, name)
    if os.path.exists(exe):
        return exe
    os.makedirs(BUILD, exist_ok=True)
    cmd = ["g++", "-O2", "-std=c++17", "-Wall", "-I" + os.path.join(ROOT, "include"),
           os.path.join(ROOT, "tests", "cpp", name + ".cpp"), "-o", exe,
           "-L" + os.path.join(ROOT, "parmoon_amd"), "-lpmx",
           "-Wl,-rpath," + os.path.join(ROOT, "parmoon_amd")]
    if with_oracle:
        cmd += ["-L" + os.path.join(ROOT, "oracle"), "-lorc", "-Wl,-rpath," + os.path.join(ROOT, "oracle")]
    subprocess.run(cmd, check=True)
    return exe


def test_shim_host():
    out = subprocess.run([build("shim_host", with_oracle=True)], capture_output=True, text=True, timeout=300)
    assert out.returncode == 0 and out.stdout.strip().endswith("OK"), out.stdout + out.stderr


@pytest.mark.gpu
def test_shim_solve_and_time_steps():
    print("[case] C++ shim assemble/solve/tcd", flush=True)
    d = tempfile.mkdtemp()
    prm = P.config1b()
    opt = P.solver_opt(solver=16, restart=20, maxit=200, res_norm_min=0.0, red_factor=1e-13, smoother=3,
                       omega=1.0)
    prm.tofile(os.path.join(d, "prm.bin"))
    opt.tofile(os.path.join(d, "opt.bin"))
    O = orc.System(1, 4, 1)
    fine = O.levels[-1]
    prm_t = P.heat_amc(0.0)
    u0 = fine.interpolate(prm_t)
    tau, steps = 0.1082532, 3
    opt_t = P.solver_opt(solver=16, restart=20, maxit=100, res_norm_min=1e-13, smoother=3, omega=1.0)
    prm_t.tofile(os.path.join(d, "prm_tcd.bin"))
    opt_t.tofile(os.path.join(d, "opt_tcd.bin"))
    np.array([tau, steps, .5, .5, .5, .5]).tofile(os.path.join(d, "td.bin"))
    u0.tofile(os.path.join(d, "u0.bin"))
    # the caller's MG level matrices: the oracle's SUPG matrices of 2^1..2^3 Q1
    G = orc.System(1, 3, 1)
    G.assemble(prm)
    for lv in range(3):
        G.entries(lv).tofile(os.path.join(d, f"mg_entries_{lv}.bin"))
    G.rhs().tofile(os.path.join(d, "mg_rhs.bin"))
    out = subprocess.run([build("shim_solve"), d], capture_output=True, text=True, timeout=300)
    print(out.stdout, out.stderr, flush=True)
    assert out.returncode == 0 and out.stdout.strip().endswith("OK")
    # CD3D against the oracle
    C = orc.System(1, 3, 1)
    C.assemble(prm)
    cf = C.levels[-1]
    E = C.entries()
    _, _, rp, _ = cf.arrays()
    ent = np.fromfile(os.path.join(d, "entries.bin"))
    rows = np.repeat(np.arange(cf.N), np.diff(rp))
    scale = np.maximum.reduceat(np.abs(E), rp[:-1])
    assert np.max(np.abs(ent - E) / scale[rows]) < 1e-12
    rhs = np.fromfile(os.path.join(d, "rhs.bin"))
    assert np.max(np.abs(rhs - C.rhs())) <= 1e-12 * np.max(np.abs(C.rhs()))
    sol0 = np.zeros(cf.N)
    sol0[cf.active:] = rhs[cf.active:]
    ref, _ = C.solve(P.solver_opt(solver=16, restart=30, maxit=300, res_norm_min=0.0, red_factor=1e-13,
                                  smoother=3, omega=1.0), C.rhs(), sol0)
    sol = np.fromfile(os.path.join(d, "sol.bin"))
    assert np.max(np.abs(sol - ref)) < 1e-10 * np.max(np.abs(ref))
    # Laplace.h's callbacks through TAssembleMat3D == the built-in functor
    e_cb = np.fromfile(os.path.join(d, "laplace_cb_entries.bin"))
    e_fn = np.fromfile(os.path.join(d, "laplace_fn_entries.bin"))
    r_cb = np.fromfile(os.path.join(d, "laplace_cb_rhs.bin"))
    r_fn = np.fromfile(os.path.join(d, "laplace_fn_rhs.bin"))
    _, _, rp2, _ = orc.Space(4, 2).arrays()
    rows2 = np.repeat(np.arange(len(rp2) - 1), np.diff(rp2))
    scale2 = np.maximum.reduceat(np.abs(e_fn), rp2[:-1])
    assert np.max(np.abs(e_cb - e_fn) / scale2[rows2]) <= 1e-14
    assert np.max(np.abs(r_cb - r_fn)) <= 1e-14 * np.max(np.abs(r_fn))
    # TMGLevel3D + TMultiGrid3D::AddLevel with the caller's matrices, Iterate on
    # the caller's finest matrix: the solution of that system (the device's own
    # operators were assembled with another form)
    gf = G.levels[-1]
    g0 = np.zeros(gf.N)
    g0[gf.active:] = G.rhs()[gf.active:]
    gref, _ = G.solve(P.solver_opt(solver=16, restart=30, maxit=300, res_norm_min=0.0, red_factor=1e-13,
                                   smoother=3, omega=1.0), G.rhs(), g0)
    mg_sol = np.fromfile(os.path.join(d, "mg_sol.bin"))
    assert np.max(np.abs(mg_sol - gref)) < 1e-10 * np.max(np.abs(gref))
    # TCD3D against the oracle's time loop
    o_sol, _ = O.tcd3d(prm_t, [tau, 1e9, .5, .5, .5, .5], opt_t, u0, max_steps=steps)
    s_t = np.fromfile(os.path.join(d, "sol_tcd.bin"))
    assert np.max(np.abs(s_t - o_sol)) < 1e-9 * np.max(np.abs(o_sol))
