"""The level set-up on the device (pmx_setup.hip): the DOF numbering that
pmx_system_create builds on the GPU (cell frames from the tree-order index,
first encounters by atomicMin, inner / Dirichlet numbers by two scans) and
the CSR structure built from it on the GPU (row lengths, a scan, columns
sorted by rank) equal the host restatement's (pmx_host_space_arrays) bit for
bit, for Q1 and Q2 up to 64^3 Q2 (2,146,689 DOFs, 130 M entries; the host
numbering and structure themselves equal ParMooN's,
tests/test_reference_build.py)."""
import time

import numpy as np
import pytest

import parmoon_amd as pa

pytestmark = pytest.mark.gpu


@pytest.mark.parametrize("nref0,nlev,order", [(1, 3, 1), (1, 4, 2), (3, 3, 1), (2, 5, 2)])
def test_device_numbering_equals_host(ctx, nref0, nlev, order):
    print(f"[case] device numbering nref {nref0}..{nref0 + nlev - 1} order {order}", flush=True)
    t0 = time.perf_counter()
    S = pa.System(ctx, nref0, nlev, order)
    t_setup = time.perf_counter() - t0
    for lv in range(nlev):
        nref = nref0 + lv
        d, bi, gn, rp, kc = pa.host_space(nref, order)
        bi_d, gn_d = S.numbering(lv)
        assert np.array_equal(bi_d, bi) and np.array_equal(gn_d, gn), (nref, order)
        t0 = time.perf_counter()
        rp_d, kc_d = S.structure(lv)
        t_csr = time.perf_counter() - t0
        assert np.array_equal(rp_d, rp) and np.array_equal(kc_d, kc), (nref, order)
        print(f"level {nref}: {len(kc)} entries, device structure with download {t_csr:.3f} s", flush=True)
        del rp, kc, rp_d, kc_d
    print(f"set-up of {nlev} levels: {t_setup:.3f} s", flush=True)
    S.close()
