"""The C++ class API compiles the way the reference's drivers use it: every
field and geometry class with the reference's constructor arguments, the
bf / nf / vf / is_within virtuals, Jet's getters and an Observation set up
as src/main.cpp does (tests/cxx/class_api.cpp).  Compile and link only; the
GPU test runs the program."""
import os
import subprocess

HERE = os.path.dirname(os.path.abspath(__file__))


def test_class_api_program_compiles(tmp_path):
    lib = os.path.join(os.path.dirname(HERE), "jetshow_amd", "libjetshow_amd.so")
    assert os.path.exists(lib), "build the library first"
    out = str(tmp_path / "class_api")
    r = subprocess.run(["bash", os.path.join(HERE, "cxx", "build.sh"), out], capture_output=True, text=True)
    assert r.returncode == 0, r.stderr
    assert os.path.exists(out)
