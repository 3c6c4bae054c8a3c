"""Prints the reference's source files of its sequential 3D build (test
infrastructure, used by oracle/ref.mk): the file(GLOB_RECURSE ...) and
list(APPEND ...) entries of the CMakeLists.txt of src/General, Geometry,
Refinement, QuadFormulas and FE, outside the 2D branch of `if("${AParMooN_GEO}" STREQUAL "2D") ... elseif(3D)
... endif`, minus the files that need libraries this image lacks (UMFPACK,
TECIO).  Reads the files; runs nothing of the reference.

    python3 oracle/ref_sources.py REF
"""
import os
import re
import sys

DIRS = ["General", "Geometry", "Refinement", "QuadFormulas", "FE"]
NEEDS_MISSING_LIBRARIES = {"DirectSolver.C", "Output3D.C", "Output2D.C"}


def sources(ref):
    out = []
    for d in DIRS:
        branch = None  # None: common part, "2D" / "3D": inside the GEO branch
        for line in open(os.path.join(ref, "src", d, "CMakeLists.txt")):
            s = line.strip()
            if s.startswith("#"):
                continue
            g = re.match(r'file\(GLOB_RECURSE \w+ "\$\{PROJECT_SOURCE_DIR\}/src/(\w+)/\*\.C"\)', s)
            if g:  # every .C file of the directory
                base = os.path.join(ref, "src", g.group(1))
                out += sorted(os.path.join(base, f) for f in os.listdir(base)
                              if f.endswith(".C") and f not in NEEDS_MISSING_LIBRARIES)
                continue
            if re.match(r'if\("\$\{AParMooN_GEO\}" STREQUAL "2D"\)', s):
                branch = "2D"
            elif re.match(r'elseif\("\$\{AParMooN_GEO\}" STREQUAL "3D"\)', s):
                branch = "3D"
            elif s.startswith("endif") and "AParMooN_GEO" in s:
                branch = None
            m = re.search(r'list\(APPEND \w+ "\$\{PROJECT_SOURCE_DIR\}/(src/[^"]+\.C)"\)', s)
            if m and branch != "2D" and os.path.basename(m.group(1)) not in NEEDS_MISSING_LIBRARIES:
                out.append(os.path.join(ref, m.group(1)))
    return out


if __name__ == "__main__":
    print(" ".join(sources(sys.argv[1] if len(sys.argv) > 1 else "/root/reference")))
