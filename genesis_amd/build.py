"""Build libgenesis_rigid.so for gfx950 in-tree (genesis_amd/_build/).

The kernels are split over several translation units (csrc/*.hip) that compile in parallel, one hipcc process each,
then link into one shared library. Every unit includes the same device headers; a kernel is launched only from the
unit that defines it, through the launchers declared in csrc/rigid_tiles.h.
"""
import glob
import os
import subprocess
import sys
import time
from concurrent.futures import ThreadPoolExecutor

HERE = os.path.dirname(os.path.abspath(__file__))
ROOT = os.path.dirname(HERE)
CSRC = os.path.join(HERE, "csrc")
OUT_DIR = os.path.join(HERE, "_build")
OUT = os.path.join(OUT_DIR, "libgenesis_rigid.so")

HIPCC = os.environ.get("HIPCC", "/opt/rocm/bin/hipcc")
FLAGS = [
    "--offload-arch=gfx950", "-O3", "-std=c++17", "-fPIC",
    # keep arithmetic in source order (no implicit FMA) so the step tracks the reference's rounding
    "-ffp-contract=off",
    "-Xclang", "-target-feature", "-Xclang", "-packed-fp32-ops",
]


def sources():
    return sorted(glob.glob(os.path.join(CSRC, "*.hip")))


def _jobs():
    n = os.cpu_count() or 1
    return max(1, min(16, int(os.environ.get("MAX_JOBS", n)), n))


def build(verbose: bool = False, incremental: bool = False) -> str:
    """Compile every unit (or, incremental=True, only units older than their source or any header) and link."""
    os.makedirs(OUT_DIR, exist_ok=True)
    srcs = sources()
    inc = "-I" + os.path.join(ROOT, "include")
    headers = glob.glob(os.path.join(CSRC, "*.h")) + glob.glob(os.path.join(ROOT, "include", "*.h"))
    newest_header = max(os.path.getmtime(h) for h in headers)

    def compile_one(src):
        obj = os.path.join(OUT_DIR, os.path.basename(src)[:-4] + ".o")
        if incremental and os.path.exists(obj) and os.path.getmtime(obj) > max(os.path.getmtime(src), newest_header):
            return src, obj, subprocess.CompletedProcess([], 0, "", ""), 0.0
        cmd = [HIPCC, *FLAGS, inc, "-c", "-o", obj, src]
        t0 = time.time()
        r = subprocess.run(cmd, capture_output=True, text=True)
        return src, obj, r, time.time() - t0

    # the slowest units first, so the pool finishes as early as it can
    srcs.sort(key=lambda s: (0 if "fused" in s or "phase" in s else 1, s))
    with ThreadPoolExecutor(_jobs()) as ex:
        results = list(ex.map(compile_one, srcs))
    failed = [r for r in results if r[2].returncode != 0]
    for src, obj, r, dt in results:
        if verbose:
            print(f"{os.path.basename(src)}: {dt:.1f} s")
    if failed:
        for src, _, r, _ in failed:
            sys.stderr.write(f"---- {src}\n{r.stderr}")
        raise RuntimeError("hipcc failed")
    tmp = OUT + ".tmp"
    cmd = [HIPCC, "--offload-arch=gfx950", "-shared", "-fPIC", "-o", tmp, *[o for _, o, _, _ in results]]
    r = subprocess.run(cmd, capture_output=True, text=True)
    if r.returncode != 0:
        sys.stderr.write(r.stderr)
        raise RuntimeError("hipcc link failed")
    os.replace(tmp, OUT)
    return OUT


if __name__ == "__main__":
    print(build(verbose=True, incremental="--incremental" in sys.argv))
