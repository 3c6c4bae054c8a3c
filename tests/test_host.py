"""Host-side parts of libpmx (no GPU): numbering, CSR structure, C-ABI exports."""
import ctypes
import os
import re

import numpy as np
import pytest

import orc
import parmoon_amd as pa

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))


@pytest.mark.parametrize("nref,order", [(1, 1), (2, 1), (3, 1), (4, 1), (5, 1), (1, 2), (2, 2), (3, 2)])
def test_numbering_and_structure_bit_exact(nref, order):
    """Closed-form lattice numbering == ConstructSpace restatement; CSR == TSquareStructure3D+Sort."""
    s = orc.Space(nref, order)
    bi, gn, rp, kc = s.arrays()
    d, bi2, gn2, rp2, kc2 = pa.host_space(nref, order)
    assert int(d[0]) == s.N and int(d[1]) == s.active and int(d[3]) == s.nnz
    assert np.array_equal(bi, bi2)
    assert np.array_equal(gn, gn2)
    assert np.array_equal(rp, rp2)
    assert np.array_equal(kc, kc2)


@pytest.mark.parametrize("nref,order,N,nnz", [(5, 1, 35937, 810503), (6, 1, 274625, 6775847),
                                               (4, 2, 35937, 1867013), (5, 2, 274625, 15837829)])
def test_sizes_of_the_workloads(nref, order, N, nnz):
    """N and nnz of the workload levels (closed-form counts of the CSR layout)."""
    d, *_ = pa.host_space(nref, order)
    assert int(d[0]) == N and int(d[3]) == nnz


def test_capi_exports_every_declared_symbol():
    src = open(os.path.join(ROOT, "include", "pmx.h")).read()
    names = re.findall(r"^\s*(?:int|const char\s*\*|void)\s*\*?\s*(pmx_\w+)\s*\(", src, re.M)
    assert len(names) >= 25
    lib = ctypes.CDLL(pa.LIB_PATH)
    missing = [n for n in names if not hasattr(lib, n)]
    assert not missing, missing


def test_ctx_without_gpu_fails_loudly():
    """No GPU here: creating a context must raise, never fall back."""
    import subprocess
    import sys
    code = ("import parmoon_amd as pa\n"
            "try:\n    pa.Context(0)\nexcept pa.PmxError as e:\n    print('raised', e)\n")
    out = subprocess.run([sys.executable, "-c", code], cwd=ROOT, capture_output=True, text=True, timeout=120)
    if "raised" not in out.stdout:
        pytest.skip("a GPU is visible to this process")
    assert "raised" in out.stdout
