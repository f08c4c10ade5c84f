"""Build the HIP library libmfhip.so in-tree for gfx950 with hipcc.

Packed FP32 instructions are disabled for all device code
(``-target-feature -packed-fp32-ops``); the kernels are written in scalar
FP32 and compiled for CDNA4 only.

Floating-point contraction is ``-ffp-contract=on``: a multiply and an add
fuse only where they share one source expression (or an explicit fmaf).
With the default cross-statement fusion the compiler picks which of two
products to fuse from the surrounding code, so the same transform inlined
into two kernels (the flat filter and the refine queue) could round
differently in the last bit; with this setting every kernel that runs the
transform produces identical bits, and the flat kernel's time is unchanged
within noise (measured on gfx950).

Matrix-core accumulators live in ordinary VGPRs
(``-mllvm -amdgpu-mfma-vgpr-form``): the coarse screen reads every
accumulator back into VALU arithmetic, and with the AGPR form each read is
one more VALU instruction (36 of ~200 per pair before the switch).  Only the
MFMA kernels are affected.
"""
import os
import shutil
import subprocess
import sys

HERE = os.path.dirname(os.path.abspath(__file__))
ROOT = os.path.dirname(HERE)
CSRC = os.path.join(HERE, "csrc")
LIB = os.path.join(HERE, "libmfhip.so")

SOURCES = ["mfh_api.hip", "mfh_flat.hip", "mfh_hier.hip", "mfh_series.hip", "mfh_coarse.hip", "mfh_big.hip", "mfh_queue.hip", "mfh_grad.hip",
           "mfh_probe.hip"]
ARCH = os.environ.get("MFH_OFFLOAD_ARCH", "gfx950")


def hipcc():
    for cand in (os.environ.get("HIPCC"), "/opt/rocm/bin/hipcc", shutil.which("hipcc")):
        if cand and os.path.exists(cand):
            return cand
    raise RuntimeError("hipcc not found; the HIP library cannot be built")


FLAGS = ["--offload-arch=%s" % ARCH, "-O3", "-std=c++17", "-fPIC", "-ffp-contract=on",
         "-Xclang", "-target-feature", "-Xclang", "-packed-fp32-ops",
         "-mllvm", "-amdgpu-mfma-vgpr-form",
         "-I", os.path.join(ROOT, "include"), "-I", CSRC]
OBJDIR = os.path.join(ROOT, "_build")


def command(out=LIB):
    """One hipcc line that compiles every source and links the library."""
    return [hipcc()] + FLAGS + ["-shared", "-o", out] + [os.path.join(CSRC, s) for s in SOURCES]


def _compile(src, obj):
    res = subprocess.run([hipcc()] + FLAGS + ["-c", "-o", obj, os.path.join(CSRC, src)],
                         capture_output=True, text=True)
    return src, res


def build(verbose=False, jobs=None):
    """Compile every source into libmfhip.so (always recompiles): one hipcc
    per translation unit in parallel, then one link."""
    from concurrent.futures import ThreadPoolExecutor
    os.makedirs(OBJDIR, exist_ok=True)
    objs = [os.path.join(OBJDIR, os.path.splitext(s)[0] + ".o") for s in SOURCES]
    if verbose:
        print(" ".join(command()), flush=True)
    jobs = jobs or min(len(SOURCES), os.cpu_count() or 1, 16)
    with ThreadPoolExecutor(max_workers=jobs) as pool:
        results = list(pool.map(_compile, SOURCES, objs))
    failed = [(src, res) for src, res in results if res.returncode != 0]
    for src, res in failed:
        sys.stderr.write("%s:\n%s%s" % (src, res.stdout, res.stderr))
    if failed:
        raise RuntimeError("hipcc failed on %s" % ", ".join(src for src, _ in failed))
    tmp = LIB + ".tmp%d" % os.getpid()
    res = subprocess.run([hipcc(), "--offload-arch=%s" % ARCH, "-shared", "-fPIC", "-o", tmp] + objs,
                         capture_output=True, text=True)
    if res.returncode != 0:
        sys.stderr.write(res.stdout + res.stderr)
        raise RuntimeError("hipcc link failed (exit %d)" % res.returncode)
    os.replace(tmp, LIB)
    return LIB


if __name__ == "__main__":
    print(build(verbose=True))
