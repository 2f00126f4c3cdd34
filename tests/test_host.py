"""CPU tests of libmc.so without a GPU: every symbol declared in
include/mc_abi.h is exported, the host FP64 drift model matches the oracle
and numpy, and device entry points fail with an error code (not a crash)
when no GPU is present."""
import os
import re

import numpy as np
import pytest

import motioncorr1exp_amd as mc
from _oracle import Oracle

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))


def _declared():
    hdr = open(os.path.join(ROOT, "include", "mc_abi.h")).read()
    hdr = re.sub(r"/\*.*?\*/", "", hdr, flags=re.S)
    return sorted(set(re.findall(r"\b(mc_[a-z0-9_]+)\s*\(", hdr)))


def test_header_and_binding_agree():
    assert sorted(mc.ABI_SYMBOLS) == _declared()


def test_library_exports_every_declared_symbol():
    L = mc.lib()
    missing = [s for s in _declared() if not hasattr(L, s)]
    assert not missing, missing


def test_version_string():
    assert b"gfx950" in mc.lib().mc_version()


@pytest.mark.parametrize("F,fod", [(10, 4), (24, 11), (40, 2), (40, 19), (7, 0), (2, 0)])
def test_pair_list_matches_oracle(F, fod):
    assert np.array_equal(mc.pair_list(F, fod), Oracle("f64").pair_list(F, fod))


def test_lsq_normal_equations_match_householder_and_numpy():
    o = Oracle("f64")
    rng = np.random.default_rng(1)
    for F, fod in [(10, 4), (24, 11), (40, 2)]:
        pairs = mc.pair_list(F, fod)
        b = rng.standard_normal((len(pairs), 2)) * 3
        s = mc.lsq(pairs, F, b)
        assert np.abs(s - o.lsq(pairs, F, b)).max() < 1e-10
        A = np.zeros((len(pairs), F - 1))
        for q, (i, j) in enumerate(pairs):
            A[q, i:j] = 1
        assert np.abs(s - np.linalg.lstsq(A, b, rcond=None)[0]).max() < 1e-10


def test_threshold_matches_oracle():
    o = Oracle("f64")
    rng = np.random.default_rng(2)
    for F, fod, kit in [(10, 4, 1.0), (24, 11, 0.3), (40, 2, 0.5), (6, 0, 0.05)]:
        pairs = mc.pair_list(F, fod)
        ki = np.abs(rng.standard_normal(len(pairs))) * 0.6
        ki[rng.integers(0, len(pairs), 3)] = ki[0]  # ties resolve by row index
        assert np.array_equal(mc.threshold(pairs, F, ki, kit), o.threshold(pairs, F, ki, kit))


def test_singular_system_reports_error():
    pairs = np.array([[0, 1], [0, 1]], np.int32)  # frame 2 never constrained
    with pytest.raises(mc.MCError):
        mc.lsq(pairs, 3, np.zeros((2, 2)))


def test_device_calls_fail_cleanly_without_gpu():
    if os.path.exists("/dev/kfd"):
        pytest.skip("a GPU is present")
    L = mc.lib()
    h = mc.ctypes.c_void_p()
    assert L.mc_context_create(0, mc.ctypes.byref(h)) != mc.MC_OK
    assert L.mc_init_gpu(0) != mc.MC_OK
    assert L.mc_device_count() == 0


def test_default_params_follow_verify_para():
    p = mc.default_params(40, nEndSum2=50)
    assert p["nEnd"] == 39 and p["nEndSum"] == 39 and p["nEndSum2"] == 39
    p = mc.default_params(40, nStartSum=5, nEndSum2=3)
    assert p["nEndSum2"] == 0 and p["nStartSum"] == 5


def test_shipped_tuning_table_is_well_formed():
    """Every line of the shipped autotuner table names this device type, a
    supported geometry and launch shapes inside the ranges mc_tune_load
    accepts (it skips anything else)."""
    import json
    import os

    path = os.path.join(os.path.dirname(mc.__file__), "tune_mi355x.jsonl")
    # transform sizes: any even size from 64 to 8192 (Stockham plans for
    # 2^k * {1, 3, 5, 15}, chirp-z for the rest, such as the 3710 crop)
    sizes = set(range(64, 8193, 2))
    lines = open(path).read().splitlines()
    assert lines
    assert any(json.loads(line)["n"] == 3710 for line in lines)
    for line in lines:
        e = json.loads(line)
        assert e["arch"] == "gfx950" and e["cus"] > 0
        assert e["nu"] in sizes and e["n"] in sizes and e["n"] <= e["nu"]
        assert e["dtype"] in mc.MC_DTYPE.values() and 16 <= e["box"] <= 256
        assert e["precision"] in (0, 1)
        # row shape (bits 0-2), persistent column pass (bit 3), column radix
        # plan and LDS padding (bits 4-6, fewer than 6 plans)
        assert 1 <= (e["r2c"] & 7) <= 3 and 0 <= (e["r2c"] >> 4) < 6
        assert e["chunk"] in (1, 2, 4, 8)
        assert 0 <= e["cols"] <= 3 and 0 <= e["rows"] <= 2


@pytest.mark.parametrize("P", [31, 53])
def test_odd_dft_coefficients_match_libm(tmp_path, P):
    """The compile-time cos/sin tables of the direct odd-length DFT
    (OddTrig<P> in mc_device.h, used for the 31- and 53-point stages of the
    two-level transforms at 4092 and 3710) equal libm's double cos and sin
    of 2 pi m/P rounded to float, for every m."""
    import shutil
    import subprocess

    hipcc = shutil.which("hipcc") or "/opt/rocm/bin/hipcc"
    if not os.path.exists(hipcc):
        pytest.skip("hipcc not available")
    src = tmp_path / "trig.cpp"
    src.write_text(
        '#include "mc_device.h"\n#include <cstdio>\n'
        "int main() { constexpr mc::OddTrig<%d> T{};\n"
        '  for (int m = 0; m < %d; m++) printf("%%a %%a\\n", (double)T.c[m], (double)T.s[m]); }\n'
        % (P, P))
    exe = tmp_path / "trig"
    inc = os.path.join(ROOT, "motioncorr1exp_amd", "csrc")
    subprocess.run([hipcc, "-std=c++17", "-O1", "-I", inc, str(src), "-o", str(exe)],
                   check=True, capture_output=True, timeout=300)
    out = subprocess.run([str(exe)], check=True, capture_output=True, text=True).stdout.split()
    got = np.array([float.fromhex(v) for v in out]).reshape(P, 2)
    m = np.arange(P)
    want_c = np.cos(2 * np.pi * m / P).astype(np.float32).astype(np.float64)
    want_s = np.sin(2 * np.pi * m / P).astype(np.float32).astype(np.float64)
    assert np.array_equal(got[:, 0], want_c) and np.array_equal(got[:, 1], want_s)
