"""Writes tests/golden/reference_numbering.npz from the reference's own code.

The reference's mesh, refinement, FE-space and structure sources are built
with g++ into oracle/_ref/ref_numbering (oracle/ref.mk; only in a container
that holds the reference tree).  For the unit cube (data/Wuerfel.PRM/.GEO)
refined 1 and 2 times with Q1 and Q2 elements, and 3 times with Q1, the
program writes ParMooN's BeginIndex, GlobalNumbers, RowPtr and KCol
(TFESpace3D with every boundary Dirichlet, TSquareStructure3D + Sort()).
These arrays are the fixture: data the reference computed, no source text.

    python tests/golden/make_numbering_fixtures.py [REF]
"""
import os
import subprocess
import sys
import tempfile

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
ROOT = os.path.dirname(os.path.dirname(HERE))
EXE = os.path.join(ROOT, "oracle", "_ref", "ref_numbering")
CASES = [(1, 1), (1, 2), (2, 1), (2, 2), (3, 1)]


def run_reference(nref, order, ref="/root/reference"):
    """(N, active, BeginIndex, GlobalNumbers, RowPtr, KCol) of ParMooN's own code."""
    with tempfile.TemporaryDirectory() as d:
        out = os.path.join(d, "n.bin")
        subprocess.run([EXE, os.path.join(ref, "data", "Wuerfel.PRM"), os.path.join(ref, "data", "Wuerfel.GEO"),
                        str(nref), str(order), out], check=True, capture_output=True, timeout=600)
        a = np.fromfile(out, np.int32)
    N, act, nc, nb, nnz = (int(v) for v in a[:5])
    o = 5
    bi = a[o:o + nc + 1]
    o += nc + 1
    gn = a[o:o + nc * nb]
    o += nc * nb
    rp = a[o:o + N + 1]
    o += N + 1
    kc = a[o:o + nnz]
    return N, act, bi.copy(), gn.copy(), rp.copy(), kc.copy()


def main():
    ref = sys.argv[1] if len(sys.argv) > 1 else "/root/reference"
    arrays = {}
    for nref, order in CASES:
        N, act, bi, gn, rp, kc = run_reference(nref, order, ref)
        key = f"n{nref}_q{order}"
        arrays[key + "_dims"] = np.array([N, act], np.int64)
        arrays[key + "_begin_index"] = bi
        arrays[key + "_global_numbers"] = gn
        arrays[key + "_rowptr"] = rp
        arrays[key + "_kcol"] = kc
        print(f"{key}: N {N} active {act} nnz {len(kc)}")
    np.savez_compressed(os.path.join(HERE, "reference_numbering.npz"), **arrays)


if __name__ == "__main__":
    main()
