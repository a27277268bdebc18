"""Build recipe of libspecfwat_hip.so (hipcc, gfx950, in-tree)."""
import os
import subprocess

HERE = os.path.dirname(os.path.abspath(__file__))
SOURCES = ["prepare.hip", "forces.hip", "timeloop.hip", "brick.hip", "smooth_pde.hip", "batch.hip", "postproc.hip", "optimize.hip", "measure.hip"]
OUT = os.path.join(HERE, "libspecfwat_hip.so")

HIPCC_FLAGS = [
    "-O3", "-std=c++17", "--offload-arch=gfx950", "-fPIC", "-shared",
    # packed fp32 ops stay off: every FP32 operation is a scalar VALU op
    "-Xclang", "-target-feature", "-Xclang", "-packed-fp32-ops",
]


def build(verbose=False):
    hipcc = os.environ.get("HIPCC", "/opt/rocm/bin/hipcc")
    srcs = [os.path.join(HERE, "csrc", s) for s in SOURCES]
    tmp = OUT + ".tmp"
    cmd = [hipcc] + HIPCC_FLAGS + ["-o", tmp] + srcs
    r = subprocess.run(cmd, capture_output=True, text=True)
    if r.returncode != 0:
        raise RuntimeError("hipcc failed:\n" + r.stdout + r.stderr)
    if verbose and r.stderr:
        print(r.stderr)
    os.replace(tmp, OUT)
    return OUT


if __name__ == "__main__":
    print(build(verbose=True))
