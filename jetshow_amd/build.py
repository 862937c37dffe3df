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
    return [os.path.join(CSRC, "jsx_host.hip"), os.path.join(CSRC, "jsx_textio.cpp")]


CXX_DIR = os.path.join(HERE, "cxx")


def build_cxx(verbose=False):
    """The reference's class API (cxx/jetshow.cpp) and the pyjetshow module:
    host C++ only, linked against libjetshow_amd.so (rpath $ORIGIN)."""
    import sysconfig

    import pybind11
    ext = sysconfig.get_config_var("EXT_SUFFIX") or ".so"
    out = os.path.join(HERE, "pyjetshow" + ext)
    inc = [pybind11.get_include(), sysconfig.get_paths()["include"], CXX_DIR]
    cmd = ["g++", "-O2", "-std=c++17", "-fPIC", "-shared", "-Wall"] + ["-I" + i for i in inc] + [
        os.path.join(CXX_DIR, "jetshow.cpp"), os.path.join(CXX_DIR, "pyjetshow.cpp"),
        "-L" + HERE, "-ljetshow_amd", "-Wl,-rpath,$ORIGIN", "-o", out]
    if verbose:
        print(" ".join(cmd), flush=True)
    r = subprocess.run(cmd, capture_output=True, text=True)
    if r.returncode != 0:
        sys.stderr.write(r.stderr)
        raise RuntimeError("g++ failed building %s" % out)
    return out


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
    build_cxx(verbose)
    return OUT


if __name__ == "__main__":
    print(build(verbose=True))
