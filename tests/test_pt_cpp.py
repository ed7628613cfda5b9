"""The C++ pkg/pt API (include/trace_pt.hpp, trace_amd/libtrace_pt.so) through
its own test program, tests/cpp/pt_test (built by tests/cpp/build_tests.py,
which the package build runs).  The program's host cases (matrix builders and
camera against closed forms, the buffers' merge rules, the OBJ reader on the
reference's cube fixture, the no-device error) run on the CPU; its device cases
(LBVH, PHR at branching 2 and 4, the optimisers, a scene graph rendered with the
benchmark and default renderers, the heat map, error paths) run on the GPU, each
checked against the FP64 oracle."""
import os
import subprocess

import pytest

HERE = os.path.dirname(os.path.abspath(__file__))
ROOT = os.path.dirname(HERE)
BINARY = os.path.join(HERE, "cpp", "_build", "pt_test")


def _run(mode: str, timeout: int):
    env = dict(os.environ, PT_TEST_GOLDEN=os.path.join(HERE, "golden"))
    res = subprocess.run([BINARY, mode], cwd=ROOT, env=env, capture_output=True, text=True, timeout=timeout)
    print(res.stdout[-4000:], res.stderr[-2000:], flush=True)
    return res


def test_pt_host_cases():
    if not os.path.exists(BINARY):
        # the CPU suite may run before a package build: build the test program here
        import importlib.util

        spec = importlib.util.spec_from_file_location("pt_build_tests", os.path.join(HERE, "cpp", "build_tests.py"))
        mod = importlib.util.module_from_spec(spec)
        spec.loader.exec_module(mod)
        if not os.path.exists(os.path.join(ROOT, "trace_amd", "libtrace_pt.so")):
            pytest.skip("libtrace_pt.so is not built (run __graft_entry__.build())")
        mod.build(verbose=False)
    res = _run("--cpu", 120)
    assert res.returncode == 0, res.stdout
    assert "5 of 5 cases passed" in res.stdout


@pytest.mark.gpu
def test_pt_device_cases():
    assert os.path.exists(BINARY), "tests/cpp/_build/pt_test is not built (run __graft_entry__.build())"
    print("pt_test --gpu", flush=True)
    res = _run("--gpu", 180)
    assert res.returncode == 0, res.stdout
    assert "FAIL" not in res.stdout
