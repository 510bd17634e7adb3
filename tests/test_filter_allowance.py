"""The bound filter's allowance arithmetic, checked exhaustively on the CPU.

Every filter rejection is decided as "SAD > allowance", where the allowance
of a pair of cell lengths comes from hyperblocker_amd/csrc/hb_allow.h
(neg_allowance / allowances_inv: one reciprocal and a +-1 correction).
tests/native/allow_check.cpp compiles that same header and compares the
decision with the direct float test (float)t * (1/la + 1/lb) < 2 * coef
(the test test_host.test_filter_bound_never_drops_a_kept_pair proves
sound) for every pair of lengths 1..255, the bound coefficients c0 and c3
of every threshold 0.000, 0.001, ..., 1.000, and reciprocals perturbed by
up to 2^-22 relative (the device reciprocal is within 1 ulp).
"""
import os
import subprocess

import numpy as np

import hyperblocker_amd as hb

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))


def _coefficients():
    ks = set()
    for thr in np.arange(0.0, 1.0005, 0.001):
        c = hb.filter_coefs(float(thr))
        for q in (0, 3):
            ks.add(float(np.float32(2.0) * np.float32(c[q])))
    return sorted(ks)


def test_allowances_equal_the_float_test_exhaustively(tmp_path):
    exe = str(tmp_path / "allow_check")
    src = os.path.join(ROOT, "tests", "native", "allow_check.cpp")
    subprocess.run(["g++", "-O2", "-std=c++17", "-pthread", src, "-o", exe], check=True,
                   capture_output=True, timeout=300)
    ks = _coefficients()
    assert len(ks) > 1500
    out = subprocess.run([exe] + [repr(k) for k in ks], capture_output=True, text=True,
                         timeout=600)
    assert out.returncode == 0, out.stdout[-2000:] + out.stderr[-2000:]
    last = out.stdout.strip().splitlines()[-1].split()
    assert last[0] == "cases" and int(last[1]) == 255 * 255 * len(ks) * 5, last
    assert last[2] == "mismatches" and int(last[3]) == 0, out.stdout[-2000:]
