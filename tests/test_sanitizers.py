"""The CPU code under AddressSanitizer and UndefinedBehaviorSanitizer.

* the host set-up of libpmx (parmoon_amd/csrc/pmx_mesh.cpp: numbering, CSR
  pattern, stencil layout) compiled with g++ and the sanitizers into
  tests/cpp/mesh_check, which checks the structural invariants;
* the oracle (oracle/orc.cpp) built as oracle/liborc_san.so and driven from
  Python with the sanitizer runtime preloaded: assembly of every form on Q1
  and Q2, FGMRES / Richardson / BiCGStab with every smoother, W and F
  cycles, transfers, the SUPG time loop and the error norms.
Any sanitizer report fails the run (halt_on_error, no recovery).
"""
import os
import subprocess
import sys

import pytest

HERE = os.path.dirname(os.path.abspath(__file__))
ROOT = os.path.dirname(HERE)
SAN = ["-fsanitize=address,undefined", "-fno-sanitize-recover=undefined", "-fno-omit-frame-pointer"]

DRIVER = r"""
import numpy as np, orc
from parmoon_amd import prm as P
for order, nref in ((1, 2), (2, 1)):
    sp = orc.Space(nref, order)
    for form in (1, 2, 3, 4, 5, 6, 7, 8):
        prm = P.config1(form=form)
        prm[P.TAU], prm[P.THETA1] = 0.01, 0.5
        sp.assemble(prm, two=form == 5)
for order in (1, 2):
    S = orc.System(1, 3, order)
    S.assemble(P.config1b() if order == 1 else P.config2(nref=3))
    fine = S.levels[-1]
    rhs = S.rhs(); sol0 = np.zeros(fine.N); sol0[fine.active:] = rhs[fine.active:]
    for solver in (16, 11, 13, 0):
        for sm in (1, 2, 3):
            for cyc in (1, 2, 0):
                S.solve(P.solver_opt(solver=solver, restart=7, maxit=9, smoother=sm, cycle=cyc, coarse=17 if cyc else 3),
                        rhs, sol0)
    S.set_slc(2)
    S.solve(P.solver_opt(solver=16, maxit=5, smoother=1), rhs, sol0)
    S.set_multicolour()
    S.solve(P.solver_opt(solver=16, maxit=5, smoother=3), rhs, sol0)
    c, f = S.levels[0], S.levels[1]
    orc.restrict(c, f, orc.prolongate(c, f, np.ones(c.N)))
O = orc.System(1, 2, 1)
prm = P.rotating(seed=1, form=P.FORM_TCD_SUPG_AK)
u0 = O.levels[-1].interpolate(prm)
O.tcd3d(prm, [1e-2, 0.035, .5, .5, .5, .5], P.solver_opt(solver=16, maxit=20, smoother=3, omega=0.5), u0, max_steps=3)
print("oracle driver done")
"""


def asan_runtime():
    out = subprocess.run(["g++", "-print-file-name=libasan.so"], capture_output=True, text=True).stdout.strip()
    return out if os.path.isabs(out) and os.path.exists(out) else None


def test_mesh_setup_under_sanitizers(tmp_path):
    exe = str(tmp_path / "mesh_check")
    subprocess.run(["g++", "-O1", "-g", "-std=c++17", *SAN, os.path.join(HERE, "cpp", "mesh_check.cpp"),
                    os.path.join(ROOT, "parmoon_amd", "csrc", "pmx_mesh.cpp"), "-o", exe], check=True)
    r = subprocess.run([exe], capture_output=True, text=True,
                       env=dict(os.environ, ASAN_OPTIONS="halt_on_error=1", UBSAN_OPTIONS="halt_on_error=1"))
    assert r.returncode == 0, r.stdout + r.stderr
    assert "spaces checked" in r.stdout and "runtime error" not in r.stderr


def test_oracle_under_sanitizers():
    rt = asan_runtime()
    if rt is None:
        pytest.skip("no libasan runtime for g++ in this image")
    subprocess.run(["make", "-C", os.path.join(ROOT, "oracle"), "liborc_san.so"], check=True, capture_output=True)
    env = dict(os.environ, LD_PRELOAD=rt, PMX_ORACLE_LIB=os.path.join(ROOT, "oracle", "liborc_san.so"),
               ASAN_OPTIONS="detect_leaks=0:halt_on_error=1", UBSAN_OPTIONS="halt_on_error=1",
               PYTHONPATH=os.pathsep.join([ROOT, HERE]))
    r = subprocess.run([sys.executable, "-c", DRIVER], capture_output=True, text=True, env=env, timeout=600)
    assert r.returncode == 0 and "oracle driver done" in r.stdout, r.stdout[-3000:] + r.stderr[-3000:]
    assert "runtime error" not in r.stderr, r.stderr[-3000:]
