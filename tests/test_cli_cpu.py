"""The cabanamd driver and the C++ host classes, without a GPU: option
handling (inputCL.cpp:69-258 error behaviour) and that cabanamd_amd.hpp
compiles against the C ABI and links."""
import os
import subprocess

import pytest

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
EXE = os.path.join(ROOT, "cabanamd_amd", "cabanamd")


def _run(*args):
    return subprocess.run([EXE, *args], capture_output=True, text=True,
                          timeout=60)


def test_help():
    r = _run("--help")
    assert r.returncode == 0
    assert "--input-lammps" in r.stdout


@pytest.mark.parametrize("args,msg", [
    (["--force-iteration", "NEIGH_BOTH"], "Unknown commandline option"),
    (["--device-type", "CUDA"], "HIP devices only"),
    (["--neigh-type", "TREE_2D"], "Unknown commandline option"),
    (["--bogus"], "Unknown commandline option: --bogus"),
    ([], "No input file"),
    (["-il", "in.lj", "--vacuum", "1.0"], "Vacuum multiplier must be bigger than 1.0"),
])
def test_bad_options(args, msg):
    r = _run(*args)
    assert r.returncode == 1
    assert msg in r.stderr


def test_no_device_is_an_error(tmp_path):
    deck = tmp_path / "in.lj"
    deck.write_text("units lj\n")
    err = tmp_path / "err.txt"
    r = _run("-il", str(deck), "-e", str(err))
    if r.returncode == 0:
        pytest.skip("a HIP device is visible")
    assert "no HIP device" in r.stderr
    assert "no HIP device" in err.read_text()


def test_cpp_classes_compile_and_link(tmp_path):
    src = tmp_path / "use.cpp"
    src.write_text(r'''
#include "cabanamd_amd.hpp"
#include <cstdio>
int main() {
    try {
        cabanamd_amd::System s(1);
        cabanamd_amd::NeighborVerlet n(2.8, false, 50);
        cabanamd_amd::ForceLJ f(&s);
        f.init_coeff({{"pair_coeff", "1", "1", "1.0", "1.0", "2.5"}});
        std::puts("device");
    } catch (const cabanamd_amd::Error &e) {
        std::printf("status %d: %s\n", (int)e.status, e.what());
        return e.status == CMD_ERR_NO_DEVICE ? 0 : 2;
    }
    return 0;
}
''')
    exe = tmp_path / "use"
    lib = os.path.join(ROOT, "cabanamd_amd")
    subprocess.run(["g++", "-std=c++17", "-Wall", "-Wextra", "-Werror",
                    "-I" + os.path.join(ROOT, "include"), str(src), "-o",
                    str(exe), "-L" + lib, "-lcabanamd_amd",
                    "-Wl,-rpath," + lib], check=True, timeout=120)
    r = subprocess.run([str(exe)], capture_output=True, text=True, timeout=60)
    assert r.returncode == 0, r.stdout + r.stderr
