"""Builds libfvmi.so (gfx950) in-tree with hipcc.

    python -m finitevolumegpu_amd.build [--verbose] [--define NAME=VALUE ...] [--flag FLAG ...] [--out PATH]

Four translation units: the exact-arithmetic kernels (-ffp-contract=off, IEEE
division/square root), the fast kernels (contraction allowed, the ILP-maximising
scheduler; the Euler and the shallow-water kernels in units of their own, which
compile in parallel) and the C-ABI runtime.  Packed FP32 instructions are
disabled for all device code.
"""

import os
import shutil
import subprocess
import sys

HERE = os.path.dirname(os.path.abspath(__file__))
CSRC = os.path.join(HERE, "csrc")
LIB = os.path.join(HERE, "libfvmi.so")
ARCH = "gfx950"

COMMON = [
    "-O3", "-std=c++17", "-fPIC", f"--offload-arch={ARCH}",
    "-Xclang", "-target-feature", "-Xclang", "-packed-fp32-ops",
    "-fno-gpu-rdc", "-Wall", "-Wno-unused-function",
]

UNITS = [
    ("fv_kernels_exact.hip", ["-ffp-contract=off", "-fhip-fp32-correctly-rounded-divide-sqrt"]),
    # the Euler kernels: 1 % faster with the ILP-maximising scheduler than with
    # the default one on the EE2D 8192^2 substep
    ("fv_kernels_fast.hip", ["-ffp-contract=on", "-mllvm", "--amdgpu-sched-strategy=max-ilp"]),
    # the shallow-water kernels with the ILP-maximising scheduler: 4 % faster on the
    # KP07_dimsplit 4096^2 substep, 6 % on its ensemble, 2 % on WAF and LxF
    ("fv_kernels_fast_ilp.hip", ["-ffp-contract=on", "-mllvm", "--amdgpu-sched-strategy=max-ilp"]),
    ("fvmi_runtime.hip", ["-ffp-contract=off"]),
]


def hipcc():
    exe = shutil.which("hipcc") or "/opt/rocm/bin/hipcc"
    if not os.path.exists(exe):
        raise RuntimeError("hipcc not found")
    return exe


def build(verbose=False, defines=(), out=LIB, flags=()):
    """Compiles the three units and links ``out``.  ``defines``: extra
    ``NAME=VALUE`` macros (e.g. FVMI_ROW_PREFETCH=2) and ``flags``: extra
    compiler flags, for code-generation studies; such a build goes to its own
    object folder."""
    tag = "_".join(d.replace("=", "") for d in defines)
    if flags:
        tag += "_f%08x" % (abs(hash(tuple(flags))) & 0xFFFFFFFF)
    objdir = os.path.join(HERE, "_build" + ("_" + tag if tag else ""))
    os.makedirs(objdir, exist_ok=True)
    objs = []
    for src, extra in UNITS:
        obj = os.path.join(objdir, src.replace(".hip", ".o"))
        cmd = [hipcc()] + COMMON + extra + list(flags) + ["-D" + d for d in defines] + ["-c", os.path.join(CSRC, src),
                                                                                         "-o", obj]
        if verbose:
            print(" ".join(cmd), flush=True)
        subprocess.run(cmd, check=True)
        objs.append(obj)
    tmp = out + ".tmp"
    cmd = [hipcc(), f"--offload-arch={ARCH}", "-shared", "-fPIC", "-o", tmp] + objs
    if verbose:
        print(" ".join(cmd), flush=True)
    subprocess.run(cmd, check=True)
    os.replace(tmp, out)
    return out


if __name__ == "__main__":
    args = sys.argv[1:]
    defs = [args[i + 1] for i, x in enumerate(args) if x == "--define"]
    outs = [args[i + 1] for i, x in enumerate(args) if x == "--out"]
    flags = [args[i + 1] for i, x in enumerate(args) if x == "--flag"]
    print(build(verbose="--verbose" in args, defines=defs, out=os.path.abspath(outs[0]) if outs else LIB,
                flags=flags))
