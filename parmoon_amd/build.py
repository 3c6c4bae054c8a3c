"""Build libpmx.so (HIP, gfx950) in-tree.

    python -m parmoon_amd.build

compiles parmoon_amd/csrc/*.hip and *.cpp with hipcc into parmoon_amd/libpmx.so.
"""
import os
import subprocess
import sys

HERE = os.path.dirname(os.path.abspath(__file__))
CSRC = os.path.join(HERE, "csrc")
OUT = os.path.join(HERE, "libpmx.so")
ARCH = os.environ.get("PMX_OFFLOAD_ARCH", "gfx950")

SOURCES = ["pmx_mesh.cpp", "pmx_ops.hip", "pmx_capi.hip", "pmx_learned.hip", "pmx_setup.hip"]

FLAGS = [
    "-O3", "-std=c++17", "-fPIC",
    f"--offload-arch={ARCH}",
    # packed FP32 instructions are kept out of the generated code
    "-Xclang", "-target-feature", "-Xclang", "-packed-fp32-ops",
    # the first 16 dwords of kernel arguments arrive preloaded in SGPRs
    # (gfx950), one dependent load less at the start of every dispatch: the
    # many short launches of the multicolour sweeps gain 3-5 %
    "-mllvm", "-amdgpu-kernarg-preload-count=16",
    "-fopenmp",
    "-Wall", "-Wno-unused-function", "-Wno-unused-variable",
]


# per-source extras: the FP64 MFMA of the cell-matrix assembly keeps its
# loop-carried accumulators in VGPRs (without this the compiler copies them
# to and from AGPRs around every MFMA)
EXTRA = {"pmx_ops.hip": ["-mllvm", "-amdgpu-mfma-vgpr-form=1"]}


def build(verbose=True):
    objs = []
    bdir = os.path.join(HERE, "_build")
    os.makedirs(bdir, exist_ok=True)
    procs = []
    for s in SOURCES:
        src = os.path.join(CSRC, s)
        if not os.path.exists(src):
            continue
        obj = os.path.join(bdir, s + ".o")
        lang = ["-x", "hip"] if s.endswith(".hip") else []
        cmd = ["hipcc", *FLAGS, *EXTRA.get(s, []), *lang, "-c", src, "-o", obj]
        procs.append((s, cmd, subprocess.Popen(cmd, stdout=subprocess.PIPE, stderr=subprocess.STDOUT)))
        objs.append(obj)
    failed = False
    for s, cmd, p in procs:
        out, _ = p.communicate()
        txt = out.decode(errors="replace")
        # the host pass does not know the packed-fp32 feature; that warning is expected
        lines = [l for l in txt.splitlines() if "packed-fp32-ops" not in l and l.strip()]
        if p.returncode != 0:
            failed = True
            print(f"[build] FAILED {s}:\n" + "\n".join(lines), file=sys.stderr)
        elif verbose and lines:
            print(f"[build] {s}:\n" + "\n".join(lines[:40]))
    if failed:
        raise RuntimeError("libpmx build failed")
    tmp = OUT + ".tmp"
    cmd = ["hipcc", "-shared", "-fPIC", f"--offload-arch={ARCH}", "-fopenmp", *objs, "-o", tmp]
    subprocess.run(cmd, check=True)
    os.replace(tmp, OUT)
    if verbose:
        print(f"[build] wrote {OUT}")
    return OUT


if __name__ == "__main__":
    build()
