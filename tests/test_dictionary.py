"""The runtime's dictionary coding on the CPU (no GPU code runs).

Equality predicates join on dictionary codes, so equal codes must mean
equal byte strings.  tests/native/dict_check.cpp includes the runtime
source and codes columns of several cardinalities twice on the worker
pool; it checks code equality against string equality, stability and the
code bound.
"""
import os
import shutil
import subprocess

import pytest

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
HIPCC = shutil.which("hipcc") or "/opt/rocm/bin/hipcc"


@pytest.mark.skipif(not os.path.exists(HIPCC), reason="hipcc not available")
def test_dictionary_codes_equal_iff_strings_equal(tmp_path):
    exe = str(tmp_path / "dict_check")
    src = os.path.join(ROOT, "tests", "native", "dict_check.cpp")
    # host-only program: the runtime source has no device code; the kernel
    # launchers it references are never called
    subprocess.run([HIPCC, "-O2", "-std=c++17", "--offload-arch=gfx950", src, "-o", exe,
                    "-L/opt/rocm/lib", "-lamdhip64", "-Wl,--unresolved-symbols=ignore-all"],
                   check=True, capture_output=True, timeout=600)
    out = subprocess.run([exe], capture_output=True, text=True, timeout=300)
    assert out.returncode == 0, out.stdout + out.stderr
    assert out.stdout.strip().endswith("ok"), out.stdout
