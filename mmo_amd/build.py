"""Builds libmmo_ff.so (gfx950) in-tree with hipcc.

    python -m mmo_amd.build

Every HIP source is compiled with packed-FP32 instructions disabled
(``-Xclang -target-feature -Xclang -packed-fp32-ops``): the pair kernels are written
for unpacked v_fma_f32 / v_mul_f32 issue.
"""
import os
import shutil
import subprocess
import sys

HERE = os.path.dirname(os.path.abspath(__file__))
ROOT = os.path.dirname(HERE)
CSRC = os.path.join(HERE, "csrc")
SOURCES = ["api.hip", "ff_kernels.hip", "pose_kernels.hip", "mc_kernels.hip", "min_kernels.hip",
           "select_kernels.hip", "desolv_kernels.hip", "molio.cpp"]
LIB = os.path.join(HERE, "libmmo_ff.so")
ARCH = os.environ.get("MMO_OFFLOAD_ARCH", "gfx950")

HIPCC_FLAGS = [
    "-O3", "-std=c++17", "-fPIC", "--offload-arch=%s" % ARCH,
    "-Xclang", "-target-feature", "-Xclang", "-packed-fp32-ops",
    "-Wall", "-Wno-unused-result",
    "-I", os.path.join(ROOT, "include"), "-I", CSRC,
]
# The persistent chain and minimiser kernels run long per-frame control code around
# their energy sweeps.  Machine-level loop-invariant code motion hoists constant
# materialisations (the FP64 polynomial coefficients of exp and sincos, for instance) out
# of the frame loop into VGPRs, which then spill to scratch and are reloaded, one
# dependent scratch load after another, on every frame; without it these kernels fit
# their register budgets (k_mc direct form: 220 -> 20 bytes of scratch per lane).
PER_FILE_FLAGS = {
    "mc_kernels.hip": ["-mllvm", "-disable-machine-licm"],
    "min_kernels.hip": ["-mllvm", "-disable-machine-licm"],
}


def hipcc():
    path = shutil.which("hipcc") or "/opt/rocm/bin/hipcc"
    if not os.path.exists(path):
        raise RuntimeError("hipcc not found")
    return path


def build(verbose=False, jobs=None):
    """Compiles every source (HIP kernels, host C++) to an object and links libmmo_ff.so."""
    objdir = os.path.join(HERE, "_build")
    os.makedirs(objdir, exist_ok=True)
    procs = []
    objs = []
    for src in SOURCES:
        obj = os.path.join(objdir, os.path.splitext(src)[0] + ".o")
        objs.append(obj)
        cmd = [hipcc()] + HIPCC_FLAGS + PER_FILE_FLAGS.get(src, []) + ["-c", os.path.join(CSRC, src), "-o", obj]
        if verbose:
            print(" ".join(cmd), flush=True)
        procs.append((src, subprocess.Popen(cmd, stdout=subprocess.PIPE, stderr=subprocess.STDOUT)))
    failed = []
    for src, p in procs:
        out, _ = p.communicate()
        text = out.decode(errors="replace")
        # hipcc's host pass does not know the packed-fp32 feature; that note is harmless
        lines = [l for l in text.splitlines() if "packed-fp32-ops" not in l]
        if p.returncode != 0:
            failed.append(src)
            sys.stderr.write("\n".join(lines) + "\n")
        elif verbose and lines:
            print("\n".join(lines))
    if failed:
        raise RuntimeError("hipcc failed for %s" % ", ".join(failed))
    tmp = LIB + ".tmp"
    cmd = [hipcc(), "-shared", "-fPIC", "--offload-arch=%s" % ARCH, "-o", tmp] + objs
    subprocess.run(cmd, check=True)
    os.replace(tmp, LIB)
    return LIB


if __name__ == "__main__":
    print(build(verbose="-v" in sys.argv))
