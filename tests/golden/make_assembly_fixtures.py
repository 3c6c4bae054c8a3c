"""Writes tests/golden/reference_assembly.npz from the reference's own assembly.

oracle/_ref/ref_assemble (built by oracle/ref.mk from the reference sources;
only in a container that holds the reference tree) runs ParMooN's Assemble3D
with its local forms (src/FE/ConvDiff3D.C, src/FE/TCD3D.C) and
Compute_SDFEM_delta on the unit cube; the example callbacks are those of the
driver (constant-coefficient manufactured sine solution, and the seeded
curl-of-Fourier-modes convection of the 64^3 Q2 workload).  For each case
the fixture holds the parameter block (the input) and the reference's
Entries, Rhs and, for the two-matrix SUPG form, the second matrix.

    python tests/golden/make_assembly_fixtures.py [REF]
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

EXE = os.path.join(ROOT, "oracle", "_ref", "ref_assemble")


def cases():
    """(name, nref, order, parameter block) of the fixture."""
    b = np.array([1.0, 2.0, 3.0]) / np.sqrt(14.0)
    out = []
    for order, nref in ((1, 2), (2, 1)):
        for form, fname in ((P.FORM_GALERKIN, "galerkin"), (P.FORM_SDFEM, "sdfem"), (P.FORM_MASS, "mass"),
                            (P.FORM_TCD_A, "tcdA"), (P.FORM_TCD_SUPG_AK, "supgAK"), (P.FORM_MASS_SUPG, "mass_supg"),
                            (P.FORM_RHS, "rhs"), (P.FORM_RHS_SUPG, "rhs_supg")):
            p = P.config2(nref=nref, form=form, linear=False)
            p[P.TAU], p[P.THETA1] = 0.02, 0.5
            out.append((f"q{order}_n{nref}_curl_{fname}", nref, order, p))
        p = P.base(P.EX_SINE_CD, eps=1e-3, form=P.FORM_SDFEM, b=b, c=0.3, delta0=0.7)
        out.append((f"q{order}_n{nref}_sine_sdfem", nref, order, p))
        p = P.base(P.EX_SINE_CD, eps=1e-3, form=P.FORM_SDFEM, b=b, c=0.3, sdfem_type=0, delta0=0.7, delta1=1.3)
        out.append((f"q{order}_n{nref}_sine_sdfem_type0", nref, order, p))
        p = P.config2(nref=nref, form=P.FORM_SDFEM, linear=True)
        out.append((f"q{order}_n{nref}_curl_sdfem_linear", nref, order, p))
        # CELL_MEASURE: shortest edge, cube root of the measure, and the mesh
        # size in the convection direction (per cell, at its first point)
        for cm in (2, 3, 4):
            p = P.config2(nref=nref, form=P.FORM_SDFEM, linear=False)
            p[P.CELL_MEASURE] = cm
            out.append((f"q{order}_n{nref}_curl_sdfem_cm{cm}", nref, order, p))
        for form, fname in ((P.FORM_TCD_SUPG_AK, "supgAK"), (P.FORM_MASS_SUPG, "mass_supg"),
                            (P.FORM_RHS_SUPG, "rhs_supg")):
            p = P.config2(nref=nref, form=form, linear=False)
            p[P.TAU], p[P.THETA1], p[P.CELL_MEASURE] = 0.02, 0.5, 4
            out.append((f"q{order}_n{nref}_curl_{fname}_cm4", nref, order, p))
        p = P.base(P.EX_SINE_CD, eps=1e-3, form=P.FORM_SDFEM, b=b, c=0.3, delta0=0.7, cell_measure=4)
        out.append((f"q{order}_n{nref}_sine_sdfem_cm4", nref, order, p))
    return out


def run_reference(nref, order, prm, ref="/root/reference"):
    """(Entries, Rhs, second Entries or None) of ParMooN's Assemble3D."""
    with tempfile.TemporaryDirectory() as d:
        pf, out = os.path.join(d, "prm.bin"), os.path.join(d, "out.bin")
        np.asarray(prm, np.float64).tofile(pf)
        subprocess.run([EXE, os.path.join(ref, "data", "Wuerfel.PRM"), os.path.join(ref, "data", "Wuerfel.GEO"),
                        str(nref), str(order), pf, out], check=True, capture_output=True, timeout=600)
        a = np.fromfile(out)
    N, nnz = int(a[0]), int(a[1])
    E, rhs = a[2:2 + nnz], a[2 + nnz:2 + nnz + N]
    E2 = a[2 + nnz + N:2 + 2 * nnz + N] if len(a) > 2 + nnz + N else None
    return E.copy(), rhs.copy(), None if E2 is None else E2.copy()


def main():
    ref = sys.argv[1] if len(sys.argv) > 1 else "/root/reference"
    arrays = {}
    for name, nref, order, p in cases():
        E, rhs, E2 = run_reference(nref, order, p, ref)
        arrays[name + "_prm"] = p
        arrays[name + "_entries"] = E
        arrays[name + "_rhs"] = rhs
        if E2 is not None:
            arrays[name + "_entries2"] = E2
        print(name, len(E), len(rhs))
    np.savez_compressed(os.path.join(HERE, "reference_assembly.npz"), **arrays)


if __name__ == "__main__":
    main()
