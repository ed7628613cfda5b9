"""The C++ pkg/pt API (include/trace_pt.hpp, trace_amd/libtrace_pt.so) through
its own test program, tests/cpp/pt_test (built by tests/cpp/build_tests.py,
which the package build runs).  The program's host cases (matrix builders and
camera against closed forms, the buffers' merge rules, the OBJ reader on the
reference's cube fixture, the no-device error) run on the CPU; its device cases
(LBVH, PHR at branching 2 and 4, the optimisers, a scene graph rendered with the
benchmark and default renderers, the heat map, error paths) run on the GPU, each
checked against the FP64 oracle.  A CPU check also reads the header's function
declarations and finds each defined in libtrace_pt.so."""
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


def _declared_functions(header: str):
    """(scope, name) of every function trace_pt.hpp declares without a body:
    namespace-level functions (scope "") and members of its classes."""
    import re

    text = re.sub(r"//[^\n]*", "", open(header).read())
    out, scopes, depth, stmt = set(), [], 0, ""
    i = 0
    while i < len(text):
        ch = text[i]
        if ch == "{":
            m = re.search(r"\b(?:class|struct)\s+(\w+)[^;()]*$", stmt)
            scopes.append(m.group(1) if m else None)
            depth += 1
            stmt = ""
        elif ch == "}":
            depth -= 1
            scopes.pop()
            stmt = ""
        elif ch == ";":
            s = " ".join(stmt.split())
            stmt = ""
            if "(" in s and not s.startswith(("using", "constexpr", "return")) and "=" not in s.split("(")[0]:
                if not re.search(r"=\s*(default|0|delete)\s*$", s):
                    name = re.findall(r"(~?\w+)\s*\(", s)[0]
                    cls = [c for c in scopes if c and c != "pt"]
                    if name != "operator" and (len(scopes) == 1 or (cls and len(scopes) == 2)):
                        out.add(("" if len(scopes) == 1 else cls[-1], name))
        else:
            stmt += ch
        i += 1
    return out


def test_pt_library_defines_every_declared_function():
    lib = os.path.join(ROOT, "trace_amd", "libtrace_pt.so")
    if not os.path.exists(lib):
        pytest.skip("libtrace_pt.so is not built (run __graft_entry__.build())")
    nm = subprocess.run(["nm", "-D", "--defined-only", "-C", lib], capture_output=True, text=True, check=True).stdout
    declared = _declared_functions(os.path.join(ROOT, "include", "trace_pt.hpp"))
    assert ("", "DefaultLBVH") in declared and ("PhrBuilder", "BuildFromAuxilary") in declared
    assert ("ImageRenderer", "RenderToBuffer") in declared and len(declared) > 50
    missing = [f"pt::{c + '::' if c else ''}{n}" for c, n in sorted(declared)
               if f" pt::{c + '::' if c else ''}{n}(" not in nm]
    assert not missing, missing
