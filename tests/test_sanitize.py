"""CPU check of the host C/C++ code under AddressSanitizer and
UndefinedBehaviorSanitizer: the oracle (oracle/mc_oracle.c), the stack I/O
(host/mrcio.cpp, libmchost.so) and the CLI's option handling
(host/dosefgpu_driftcorr.cpp, dfalign.cpp) are rebuilt with
-fsanitize=address,undefined (`make sanitize` in oracle/ and
motioncorr1exp_amd/host/), then tests/test_oracle.py and tests/test_io.py
run against those builds in a child process with libasan preloaded, and
the CLI's argument errors are exercised.  Any sanitizer report fails the
test (halt_on_error, no recovery from undefined behaviour)."""
import os
import subprocess
import sys

import numpy as np
import pytest

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))


def _rt(name):
    p = subprocess.run(["gcc", "-print-file-name=" + name], capture_output=True, text=True).stdout.strip()
    return p if os.path.isabs(p) and os.path.exists(p) else None


@pytest.fixture(scope="module")
def san_env():
    asan, ubsan = _rt("libasan.so"), _rt("libubsan.so")
    if not asan or not ubsan:
        pytest.skip("gcc sanitizer runtimes not available")
    for d in ("oracle", os.path.join("motioncorr1exp_amd", "host")):
        r = subprocess.run(["make", "-s", "-C", os.path.join(ROOT, d), "sanitize"],
                           capture_output=True, text=True)
        assert r.returncode == 0, r.stdout[-2000:] + r.stderr[-2000:]
    env = dict(os.environ)
    # malloc_fill_byte=0: the reference's own readers (oracle/_ref, the
    # checker of test_io.py, not sanitized) leave members of MRCT/MRCZ
    # uninitialised (their constructors delegate to a temporary); ASan's
    # default 0xbe fill of new memory would crash them
    env.update(LD_PRELOAD=asan + " " + ubsan, MC_SANITIZED="1",
               ASAN_OPTIONS="detect_leaks=0:halt_on_error=1:abort_on_error=0:malloc_fill_byte=0",
               UBSAN_OPTIONS="halt_on_error=1:print_stacktrace=1", OMP_NUM_THREADS="4")
    return env


def _clean(out):
    assert "ERROR: AddressSanitizer" not in out and "runtime error:" not in out, out[-4000:]


def test_oracle_and_io_tests_under_sanitizers(san_env):
    r = subprocess.run([sys.executable, "-m", "pytest", "-q", "-x", "-p", "no:cacheprovider",
                        "-m", "not gpu", os.path.join(ROOT, "tests", "test_oracle.py"),
                        os.path.join(ROOT, "tests", "test_io.py")],
                       cwd=ROOT, env=san_env, capture_output=True, text=True, timeout=900)
    out = r.stdout + r.stderr
    print(out[-600:])
    _clean(out)
    assert r.returncode == 0, out[-4000:]


def _mrc(path, data):
    nz, ny, nx = data.shape
    h = np.zeros(256, np.int32)
    h[0:4] = (nx, ny, nz, 2)
    hb = bytearray(h.tobytes())
    hb[208:212] = b"MAP "
    open(path, "wb").write(bytes(hb) + data.astype(np.float32).tobytes())


@pytest.mark.parametrize("args,msg", [
    ([], "Not a stack or wrong file"),
    (["s.mrc", "-fod"], "Wrong number of arguments"),
    (["s.mrc", "-xyz", "1"], "Undefined option: -xyz"),
    (["s.mrc", "-fod", "5"], "Frame distance offset is too big"),
    (["s.mrc", "-scx", "0.5"], "Stretching factor along X-axis needs to be >=1.0."),
    (["s.mrc", "-pbx", "8"], "Peak search box (-pbx) must be at least 16"),
    (["s.mrc", "-prc", "2"], "Precision form (-prc) must be 0 or 1."),
    (["s.mrc", "-nes", "2", "-2es", "3"], "cannot be larger than or equal to"),
])
def test_cli_argument_errors_under_sanitizers(san_env, tmp_path, args, msg):
    cli = os.path.join(ROOT, "motioncorr1exp_amd", "_san", "dosefgpu_driftcorr")
    _mrc(tmp_path / "s.mrc", np.zeros((8, 64, 64), np.float32))
    argv = [cli] + (args if args else ["missing.mrc"])
    env = dict(san_env, MC_SANITIZED="1")
    r = subprocess.run(argv, cwd=tmp_path, env=env, capture_output=True, text=True, timeout=60)
    out = r.stdout + r.stderr
    _clean(out)
    assert msg in out and r.returncode != 0, out[-2000:]
