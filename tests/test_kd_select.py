"""The device kd build's selection (skwarka_amd/csrc/kernels/kd_select.hpp) is libstdc++'s
std::nth_element restated; tests/cpp/kd_select_check.cpp compiles it for the host and checks that
it leaves every element where libstdc++ does (std::nth_element, and std::__introselect with small
depth budgets for the heap-select fallback), and that the parallel form of one partition step used
by the device kernels (k-th left / k-th right candidate pairs) equals the sequential scan, on
random keys, heavy ties, sorted and reversed runs."""
import subprocess

from tests.cpp import build as cpp_build


def test_selection_matches_libstdcxx():
    exe = cpp_build.build_kd_select()
    res = subprocess.run([exe], capture_output=True, text=True, timeout=300)
    assert res.returncode == 0, res.stdout + res.stderr
    assert res.stdout.startswith("ok ")
