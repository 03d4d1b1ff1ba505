"""The C and C++ consumers of the public headers (tests/consumer): built against the library by
__graft_entry__.build(); the CPU test checks that they exist and link, the GPU test runs them
(each exits 0 when every check passes, else with the number of the failed check)."""
import os
import subprocess

import pytest

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
BIN = os.path.join(ROOT, "tests", "consumer", "bin")
PROGRAMS = ("c_consumer", "cpp_consumer")


@pytest.mark.parametrize("name", PROGRAMS)
def test_consumer_is_built_and_linked(name):
    exe = os.path.join(BIN, name)
    assert os.path.exists(exe), f"{exe} is not built (run __graft_entry__.build())"
    deps = subprocess.run(["ldd", exe], capture_output=True, text=True).stdout
    assert "libbrook_amd.so" in deps and "not found" not in deps.split("libbrook_amd.so")[1].split("\n")[0]


def test_headers_compile_as_c_and_cpp(tmp_path):
    src_c = tmp_path / "h.c"
    src_c.write_text('#include "brook.h"\nint main(void) { return BROOK_CUDA_ERROR == BROOK_DEVICE_ERROR ? 0 : 1; }\n')
    src_cpp = tmp_path / "h.cpp"
    src_cpp.write_text('#include "brook.hpp"\nint main() { return brook::api::skeletonize_defaults().flags == '
                       '(BROOK_FIX_BRANCHING | BROOK_FIX_BORDERS) ? 0 : 1; }\n')
    inc = "-I" + os.path.join(ROOT, "include")
    subprocess.run(["gcc", "-std=c99", "-Wall", "-Werror", "-c", inc, str(src_c), "-o", str(tmp_path / "h.o")], check=True)
    subprocess.run(["g++", "-std=c++17", "-Wall", "-Werror", "-c", inc, str(src_cpp), "-o", str(tmp_path / "hpp.o")],
                   check=True)


@pytest.mark.gpu
@pytest.mark.parametrize("name", PROGRAMS)
def test_consumer_runs(name):
    print(f"running consumer {name}", flush=True)
    r = subprocess.run([os.path.join(BIN, name)], capture_output=True, text=True, timeout=120)
    assert r.returncode == 0, (r.returncode, r.stdout, r.stderr)
    assert "consumer ok" in r.stdout
