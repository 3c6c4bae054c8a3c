"""pmx_defect_noalloc (the allocation-free variant of pmx_defect for callers
that cannot allocate device memory) against pmx_defect through the C-ABI.

Device vectors are allocated by the test through the HIP runtime that
libpmx itself uses (hipMalloc / hipMemcpy via ctypes).  Two calls in a row on
one system, the second on a smaller level: the defect vector is bit-identical
to pmx_defect's (the same row and boundary kernels on the same values), the
norm agrees to 1e-14 relative (its partial sums are formed in another
order), and the scratch left by the first call does not leak into the second.
A level beyond the scratch limit is rejected with PMX_ERR_ARG.
"""
import ctypes
import os

import numpy as np
import pytest

import orc
import parmoon_amd as pa
from parmoon_amd import prm as P

pytestmark = pytest.mark.gpu


def lattice_index(sp):
    M = sp.order * 2 ** sp.nref + 1
    ijk = np.rint(sp.dof_coords() * (M - 1)).astype(int)
    return ijk[:, 0] + M * (ijk[:, 1] + M * ijk[:, 2])


class Hip:
    """hipMalloc / hipMemcpy / hipFree of the HIP runtime libpmx links."""

    def __init__(self):
        rocm = os.environ.get("ROCM_PATH", "/opt/rocm")
        self.L = ctypes.CDLL(os.path.join(rocm, "lib", "libamdhip64.so"))
        self.L.hipMalloc.argtypes = [ctypes.POINTER(ctypes.c_void_p), ctypes.c_size_t]
        self.L.hipMemcpy.argtypes = [ctypes.c_void_p, ctypes.c_void_p, ctypes.c_size_t, ctypes.c_int]
        self.L.hipFree.argtypes = [ctypes.c_void_p]
        self.bufs = []

    def upload(self, a):
        a = np.ascontiguousarray(a, np.float64)
        p = ctypes.c_void_p()
        assert self.L.hipMalloc(ctypes.byref(p), a.nbytes) == 0
        self.bufs.append(p)
        assert self.L.hipMemcpy(p, a.ctypes.data_as(ctypes.c_void_p), a.nbytes, 1) == 0
        return p.value

    def download(self, ptr, n):
        out = np.empty(n, np.float64)
        assert self.L.hipMemcpy(out.ctypes.data_as(ctypes.c_void_p), ctypes.c_void_p(ptr), out.nbytes, 2) == 0
        return out

    def free(self):
        for p in self.bufs:
            self.L.hipFree(p)
        self.bufs = []


def test_defect_noalloc_matches_pmx_defect(ctx):
    S = pa.System(ctx, 1, 3, 2)
    S.assemble(P.config2(nref=3))
    hip = Hip()
    rng = np.random.default_rng(7)
    for level, nref in ((2, 3), (1, 2)):
        print(f"[case] defect_noalloc level {level}", flush=True)
        sp = orc.Space(nref, 2)
        N = sp.N
        lat = lattice_index(sp)
        x = rng.standard_normal(N)
        b = rng.standard_normal(N)
        r_ref, res_ref = S.defect(x, b, level=level)
        xl, bl = np.empty(N), np.empty(N)
        xl[lat], bl[lat] = x, b
        px, pb = hip.upload(xl), hip.upload(bl)
        pr, pres = hip.upload(np.full(N, np.nan)), hip.upload(np.full(1, np.nan))
        S.defect_noalloc(px, pb, pr, pres, level=level)
        ctx.sync()
        r = hip.download(pr, N)[lat]
        res = float(hip.download(pres, 1)[0])
        assert np.array_equal(r, r_ref), float(np.max(np.abs(r - r_ref)))
        assert abs(res - res_ref) <= 1e-14 * res_ref, (res, res_ref)
    hip.free()
    S.close()


def test_defect_noalloc_rejects_null(ctx):
    S = pa.System(ctx, 1, 2, 2)
    S.assemble(P.config2(nref=2))
    rc = pa.lib().pmx_defect_noalloc(S.h, 0, None, None, None, None)
    assert rc == 1  # NULL arguments
    S.close()
