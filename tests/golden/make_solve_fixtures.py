"""Writes tests/golden/reference_solve.npz from the reference's own solvers.

oracle/_ref/ref_solve (built by oracle/ref.mk from the reference sources;
only in a container that holds the reference tree) sets up levels, matrices
and solver objects as ParMooN's TSystemCD3D does and runs TFgmresIte /
TFixedPointIte preconditioned by TMultiGridScaIte (TMultiGrid3D::Cycle with
the lexicographic TMGLevel3D smoothers, Prolongate / DefectRestriction, the
LAPACK coarse solve) from a zero start.  For each case the fixture holds the
inputs (parameter block, solver options, levels) and the finest-level
iterate after the given number of iterations.

    python tests/golden/make_solve_fixtures.py [REF]
"""
import os
import subprocess
import sys
import tempfile

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
ROOT = os.path.dirname(os.path.dirname(HERE))
sys.path.insert(0, ROOT)
from parmoon_amd import prm as P  # noqa: E402

EXE = os.path.join(ROOT, "oracle", "_ref", "ref_solve")


def cases():
    """(name, nref0, nlev, order, parameter block, solver options)."""
    out = []
    supg = P.config1b()            # 8^3 Q1 levels of the config-1 problem, delta0 = 1
    curl = P.config2(nref=3, linear=False)  # the curl-Fourier SUPG problem on 8^3 Q2
    spec = P.config1()             # config 1 as stated: delta0 = 0.25

    def o(**kw):
        base = dict(solver=16, restart=20, maxit=7, res_norm_min=0.0, red_factor=0.0, smoother=3, omega=1.0)
        base.update(kw)
        return P.solver_opt(**base)
    out += [("q1_ssor", 1, 3, 1, supg, o()),
            ("q1_sor", 1, 3, 1, supg, o(smoother=2)),
            ("q1_jacobi", 1, 3, 1, supg, o(smoother=1, omega=0.8)),
            ("q1_jacobi_w", 1, 3, 1, supg, o(smoother=1, omega=0.8, cycle=2, maxit=5)),
            ("q1_jacobi_f", 1, 3, 1, supg, o(smoother=1, omega=0.8, cycle=0, maxit=5)),
            ("q1_ssor_coarse_ssor", 1, 3, 1, supg, o(coarse=3, coarse_sweeps=10, coarse_red=0.1, maxit=5)),
            ("q1_jacobi_damp", 1, 3, 1, supg, o(smoother=1, omega=0.8, gmg_damp=0.8, maxit=5)),
            ("q1_richardson", 1, 3, 1, supg, o(solver=11, maxit=6)),
            ("q1_restart", 1, 3, 1, supg, o(smoother=1, omega=0.8, restart=4, maxit=12)),
            # the limit inside a restart cycle: the reference back-solves the
            # full cycle width with the previous cycle's Hessenberg columns
            ("q1_restart_mid", 1, 3, 1, supg, o(smoother=1, omega=0.8, restart=4, maxit=9)),
            ("q1_converged", 1, 3, 1, supg, o(maxit=200, red_factor=1e-13)),
            ("q1_config1_as_stated", 1, 3, 1, spec, o(smoother=1, omega=0.8, maxit=40)),
            ("q2_curl_ssor", 1, 3, 2, curl, o(restart=30, maxit=6)),
            ("q2_curl_jacobi", 1, 3, 2, curl, o(restart=30, smoother=1, omega=0.8, maxit=6)),
            ("q2_curl_converged", 1, 3, 2, curl, o(restart=30, maxit=300, red_factor=1e-12))]
    return out


def run_reference(nref0, nlev, order, prm, opt, ref="/root/reference", stats=False, exe=EXE, mesh=None):
    """(finest solution, finest rhs[, {iterations, setup_s, assembly_s,
    solve_s}]) of the reference's solve; mesh: (PRM, GEO) files (default:
    the reference's data/Wuerfel.*)."""
    prm_file, geo_file = mesh or (os.path.join(ref, "data", "Wuerfel.PRM"), os.path.join(ref, "data", "Wuerfel.GEO"))
    with tempfile.TemporaryDirectory() as d:
        pf, of, out = (os.path.join(d, x) for x in ("prm.bin", "opt.bin", "out.bin"))
        np.asarray(prm, np.float64).tofile(pf)
        np.asarray(opt[:24], np.float64).tofile(of)
        subprocess.run([exe, prm_file, geo_file, str(nref0), str(nlev), str(order), pf, of, out], check=True,
                       capture_output=True, timeout=600)
        a = np.fromfile(out)
    N = int(a[0])
    x, b = a[1:1 + N].copy(), a[1 + N:1 + 2 * N].copy()
    if not stats:
        return x, b
    t = a[1 + 2 * N:]
    return x, b, {"iterations": int(t[0]), "setup_s": float(t[1]), "assembly_s": float(t[2]), "solve_s": float(t[3])}


def main():
    ref = sys.argv[1] if len(sys.argv) > 1 else "/root/reference"
    data = {}
    for name, nref0, nlev, order, prm, opt in cases():
        x, b = run_reference(nref0, nlev, order, prm, opt, ref)
        data[name + "_prm"] = prm
        data[name + "_opt"] = opt
        data[name + "_levels"] = np.array([nref0, nlev, order])
        data[name + "_sol"] = x
        data[name + "_rhs"] = b
        print(name, len(x), flush=True)
    np.savez_compressed(os.path.join(HERE, "reference_solve.npz"), **data)


if __name__ == "__main__":
    main()
