"""Build the HIP shared library for gfx950 in-tree (jetshow_amd/libjetshow_amd.so).

Flags:
  -ffp-contract=off      no implicit multiply-add contraction: the FP64 path
                         must reproduce the CPU restatement bit for bit
  -Xclang -target-feature -Xclang -packed-fp32-ops
                         no packed FP32 instructions
"""
import os
import subprocess
import sys

HERE = os.path.dirname(os.path.abspath(__file__))
CSRC = os.path.join(HERE, "csrc")
OUT = os.path.join(HERE, "libjetshow_amd.so")

HIPCC = os.environ.get("HIPCC", "/opt/rocm/bin/hipcc")
ARCH = os.environ.get("JSX_OFFLOAD_ARCH", "gfx950")

FLAGS = [
    "--offload-arch=%s" % ARCH, "-O3", "-std=c++17", "-fPIC", "-shared",
    "-ffp-contract=off", "-fno-fast-math",
    "-Xclang", "-target-feature", "-Xclang", "-packed-fp32-ops",
    "-Wall", "-Wno-unused-function",
]


def sources():
    return [os.path.join(CSRC, "jsx_host.hip")]


def build(verbose=False, extra=()):
    cmd = [HIPCC] + FLAGS + list(extra) + ["-o", OUT] + sources()
    if verbose:
        print(" ".join(cmd), flush=True)
    r = subprocess.run(cmd, capture_output=True, text=True)
    msgs = [ln for ln in (r.stderr or "").splitlines() if "packed-fp32-ops' is not a recognized" not in ln]
    if r.returncode != 0:
        sys.stderr.write("\n".join(msgs) + "\n")
        raise RuntimeError("hipcc failed building %s" % OUT)
    if verbose and msgs:
        print("\n".join(msgs))
    return OUT


if __name__ == "__main__":
    print(build(verbose=True))
