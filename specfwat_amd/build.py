"""Build recipe of libspecfwat_hip.so (hipcc, gfx950, in-tree).

Every source is compiled to its own object in parallel (no device code
crosses translation units), then the objects are linked into one library.
"""
import os
import subprocess
import tempfile
from concurrent.futures import ThreadPoolExecutor

HERE = os.path.dirname(os.path.abspath(__file__))
SOURCES = ["prepare.hip", "forces.hip", "timeloop.hip", "brick.hip", "smooth_pde.hip", "batch.hip", "postproc.hip",
           "optimize.hip", "measure.hip"]
OUT = os.path.join(HERE, "libspecfwat_hip.so")

HIPCC_FLAGS = [
    "-O3", "-std=c++17", "--offload-arch=gfx950", "-fPIC",
    # packed fp32 ops stay off: every FP32 operation is a scalar VALU op
    "-Xclang", "-target-feature", "-Xclang", "-packed-fp32-ops",
]


def _run(cmd):
    r = subprocess.run(cmd, capture_output=True, text=True)
    if r.returncode != 0:
        raise RuntimeError("hipcc failed:\n" + " ".join(cmd) + "\n" + r.stdout + r.stderr)
    return r.stderr


def build(verbose=False, jobs=None):
    hipcc = os.environ.get("HIPCC", "/opt/rocm/bin/hipcc")
    jobs = jobs or min(len(SOURCES), max(1, min(8, os.cpu_count() or 1)))
    with tempfile.TemporaryDirectory(prefix="specfwat_build_") as tmpd:
        objs = [os.path.join(tmpd, os.path.splitext(s)[0] + ".o") for s in SOURCES]
        cmds = [[hipcc] + HIPCC_FLAGS + ["-c", "-o", o, os.path.join(HERE, "csrc", s)]
                for s, o in zip(SOURCES, objs)]
        with ThreadPoolExecutor(jobs) as ex:
            logs = list(ex.map(_run, cmds))
        tmp = OUT + ".tmp"
        logs.append(_run([hipcc, "--offload-arch=gfx950", "-shared", "-fPIC", "-o", tmp] + objs))
    if verbose:
        for t in logs:
            if t:
                print(t)
    os.replace(tmp, OUT)
    return OUT


if __name__ == "__main__":
    print(build(verbose=True))
