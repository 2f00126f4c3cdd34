"""GPU test of the drop-in CLI's log text and secondary outputs against the
reference's formats and the FP64 oracle, on one 544 x 512 x 10 float stack
(square crop 512) run with -itr 2 -srs 1 -ssc 1 -scc 1 -2es 6 -sub 1 -fsc 1.

Log (consumed by the reference's dosef_logviewer): every line of the
alignment part is parsed and printed again with the reference's printf
format (DFAlign.cu:818-1320), which must reproduce it character for
character, and the sections must come in the reference's order.  The
printed values are compared with the oracle: pair shifts and CC
(DFAlign.cu:903), both ki rounds (:923, :944, :926, :947), the align-to-sum
rounds (:976-1049), the shifts after iteration (:1058-1068), the final
shifts (:1078-1086) and the shifts applied in the sum (:1151, :1181).

Files: -sub names both sums <name>_ox%d_oy%d_dim%d.mrc (:1216-1217,
:1300-1301); -scc writes the pair box maps (:1097), compared with the
oracle's maps, and one iterCC%03d.mrc per round with one section per frame;
-ssc writes each corrected frame (:1154-1259), compared with the oracle's
spectrum of that frame moved by its shift and inverted; -2es writes the
second sum (:1198-1229), compared with the oracle's."""
import os
import re
import subprocess

import numpy as np
import pytest

from _oracle import Oracle
from motioncorr1exp_amd.synth import make_movie

pytestmark = pytest.mark.gpu
ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
CLI = os.path.join(ROOT, "motioncorr1exp_amd", "bin", "dosefgpu_driftcorr")
F, NX, NY, N, ITR, NES2 = 10, 544, 512, 512, 2, 6

# reference formats (DFAlign.cu), as printf strings
FMT = {
    "read": "......Read and sum frame #%03d   mean:%f",
    "pair": "......%03d Frame #%03d VS #%03d xy-shift: %8.4f %8.4f      CC:%f",
    "ki1": "......ki #%03d of Frame #%03d VS #%03d: %8.4f ",
    "ki2": "......ki #%03d of Frame #%03d VS #%03d: %8.4f ",
    "avgki": "................................Average ki: %8.4f ",
    "iter": "Iteration [%d/%d]",
    "add": "......Add Frame #%03d with xy shift: %8.4f %8.4f",
    "vssum": "......%03d Frame #%03d VS <sum> xy-shift: %8.4f %8.4f (%8.4f %8.4f) CC:%f",
    "avgcc": "Average CC : %8.4f",
    "rmsd": "RMSD to <sum> : %8.4f  (%8.4f for convergence)",
    "final_rmsd": "Final RMSD to <sum> : %8.4f at iteration [%d/%d]",
    "delta": "...Frame #%03d xy-shift %8.4f %8.4f (%8.4f %8.4f)",
    "shift": "......Shift of Frame #%03d : %8.4f %8.4f",
    "sumrange": "Sum Frame #%03d - #%03d",
}
NUM = r"(-?\d+(?:\.\d+)?)"


def _parse(kind, line):
    """Numbers of a line of the given kind (None if it is not of that kind);
    the line must equal the reference format applied to them."""
    fmt = FMT[kind]
    pat = re.escape(fmt)
    for spec in ("%03d", "%d", "%8.4f", "%f"):
        pat = pat.replace(re.escape(spec), r" *" + NUM)
    m = re.fullmatch(pat, line)
    if not m:
        return None
    vals = []
    specs = re.findall(r"%0?\d*\.?\d*[df]", fmt)
    for spec, g in zip(specs, m.groups()):
        vals.append(int(g) if spec.endswith("d") else float(g))
    assert fmt % tuple(vals) == line, (kind, line, fmt % tuple(vals))
    return vals


def _parse_all(kind, seq):
    vals = [_parse(kind, l) for l in seq]
    assert all(v is not None for v in vals), (kind, seq)
    return np.array(vals)


def _mrc(fn):
    raw = open(fn, "rb").read()
    h = np.frombuffer(raw[:1024], np.int32)
    nx, ny, nz, mode = h[:4]
    assert mode == 2
    return np.frombuffer(raw[1024:1024 + 4 * nx * ny * nz], np.float32).reshape(nz, ny, nx)


def _mrc_bytes(data):
    nz, ny, nx = data.shape
    h = np.zeros(256, np.int32)
    h[0:4] = (nx, ny, nz, 2)
    hb = bytearray(h.tobytes())
    hb[208:212] = b"MAP "
    return bytes(hb) + data.astype(np.float32).tobytes()


@pytest.fixture(scope="module")
def run(tmp_path_factory):
    if not os.path.exists(CLI):
        pytest.skip("CLI not built")
    d = tmp_path_factory.mktemp("cliout")
    frames, _ = make_movie(NX, NY, F, dose=1.5, step=0.8, seed=777)
    open(d / "s.mrc", "wb").write(_mrc_bytes(frames))
    r = subprocess.run([CLI, "s.mrc", "-itr", str(ITR), "-srs", "1", "-ssc", "1", "-scc", "1",
                        "-2es", str(NES2), "-sub", "1", "-fsc", "1", "-dsp", "0"],
                       cwd=d, capture_output=True, text=True, timeout=120)
    assert r.returncode == 0, r.stdout[-2000:] + r.stderr[-2000:]
    o = Oracle("f64").align_movie(frames, dict(iter=ITR, nEndSum2=NES2, nthreads=16),
                                  want_spectra=True, want_boxes=True)
    return d, open(d / "s_Log.txt").read().split("\n"), o


def test_log_lines_reprint_and_match_oracle(run):
    d, lines, o = run
    P = len(o["pairs"])
    i = next(k for k, l in enumerate(lines) if l == "Calculate relative drift between frames")
    # stage-1 lines before it
    reads = [v for v in (_parse("read", l) for l in lines[:i]) if v]
    assert [v[0] for v in reads] == list(range(F))
    i += 1
    pairs = _parse_all("pair", lines[i:i + P])
    i += P
    assert np.array_equal(pairs[:, 0], np.arange(P))
    assert np.array_equal(pairs[:, 1:3], o["pairs"])
    db = np.abs(pairs[:, 3:5] - o["pair_b"]).max()
    dcc = np.abs(pairs[:, 5] - o["pair_cc"]).max()
    assert lines[i] == "" and lines[i + 1] == "......ki: First round "
    i += 2
    k1 = _parse_all("ki1", lines[i:i + P])
    i += P
    avg1 = _parse("avgki", lines[i])
    assert np.array_equal(k1[:, 0], np.arange(P)) and np.array_equal(k1[:, 1:3], o["pairs"])
    dk1 = np.abs(k1[:, 3] - o["ki1"]).max()
    assert abs(avg1[0] - o["ki1"].mean()) <= 5e-5 + 1e-6
    assert lines[i + 1] == "" and lines[i + 2] == "......ki: Second round "
    i += 3
    ng = len(o["good"])
    k2 = _parse_all("ki2", lines[i:i + ng])
    i += ng
    assert np.array_equal(k2[:, 1:3], o["pairs"][o["good"]])
    dk2 = np.abs(k2[:, 3] - o["ki2"]).max()
    assert _parse("avgki", lines[i]) is not None and lines[i + 1] == ""
    i += 2
    assert lines[i] == "Begin Iterative Alignment... "
    i += 1
    dit = 0.0
    for it in range(ITR):
        assert _parse("iter", lines[i]) == [it, ITR]
        i += 1
        _parse_all("add", lines[i:i + F])
        i += F
        assert lines[i] == "Sum complete... "
        i += 1
        vs = _parse_all("vssum", lines[i:i + F])
        i += F
        dit = max(dit, float(np.abs(vs[:, 2:4] - o["iter_meas"][it]).max()))
        assert lines[i] == "Save CC map to: iterCC%03d.mrc" % it
        i += 1
        assert _parse("avgcc", lines[i]) is not None
        r = _parse("rmsd", lines[i + 1])
        assert abs(r[0] - np.sqrt(o["rmsd"][it])) <= 5e-5 + 1e-6
        i += 2
    fr = _parse("final_rmsd", lines[i])
    assert fr[1:] == [ITR - 1, ITR]
    assert lines[i + 1] == "" and lines[i + 2] == "Shifts (delta) after iterative aligment"
    i += 3
    delta = _parse_all("delta", lines[i:i + F])
    i += F
    pos = np.vstack([[0, 0], np.cumsum(o["final_shift"], 0)])
    ddel = np.abs(delta[:, 1:3] - pos).max()
    assert lines[i] == "Final shift:"
    i += 1
    fin = _parse_all("shift", lines[i:i + F])
    i += F
    dfin = np.abs(fin[:, 1:3] - pos).max()
    assert lines[i] == "Save CC map to: s_CCMap.mrc"
    assert lines[i + 1] == "" and _parse("sumrange", lines[i + 2]) == [0, F - 1]
    i += 3
    sums = _parse_all("add", lines[i:i + F])
    dsum = np.abs(sums[:, 1:3] - o["frame_shift"]).max()
    print("log vs oracle (printed values): pair shifts %.2g, CC %.2g, ki %.2g / %.2g, rounds "
          "%.2g, positions %.2g, final %.2g, summed shifts %.2g" %
          (db, dcc, dk1, dk2, dit, ddel, dfin, dsum), flush=True)
    tol = 0.01 + 5e-5  # shifts: 0.01 px plus half a unit of the printed digit
    assert db <= tol and dit <= tol and ddel <= tol and dfin <= tol and dsum <= tol
    assert dk1 <= 0.01 + 5e-5 and dk2 <= 0.01 + 5e-5
    assert dcc <= 1e-5 * max(1.0, np.abs(o["pair_cc"]).max())
    assert lines[-2] == "Done."


def test_sub_area_names_and_second_sum(run):
    d, _, o = run
    main = d / ("s_SumCorr_ox0_oy0_dim%d.mrc" % N)
    second = d / ("s_SumCorr2_ox0_oy0_dim%d.mrc" % N)
    assert main.exists() and second.exists()
    s = _mrc(main)[0]
    s2 = _mrc(second)[0]
    rel = np.linalg.norm(s - o["sum_image"]) / np.linalg.norm(o["sum_image"])
    rel2 = np.linalg.norm(s2 - o["sum2_image"]) / np.linalg.norm(o["sum2_image"])
    print("sub-area sums: sum rel L2 %.3g, second sum (frames 0..%d) %.3g" % (rel, NES2, rel2))
    assert rel < 1e-4 and rel2 < 1e-4


def test_cc_maps_match_oracle(run):
    d, _, o = run
    cc = _mrc(d / "s_CCMap.mrc")
    obm = o["boxmaps"]
    assert cc.shape == obm.shape
    scale = np.abs(obm).reshape(len(obm), -1).max(1)[:, None, None]
    err = float((np.abs(cc - obm) / scale).max())
    print("CC maps vs oracle: max error / map max %.3g" % err)
    assert err < 1e-5
    for it in range(ITR):
        m = _mrc(d / ("iterCC%03d.mrc" % it))
        assert m.shape == (F, 96, 96) and np.isfinite(m).all()


def test_corrected_stack_matches_oracle(run):
    d, _, o = run
    st = _mrc(d / "s_Corrected.mrc")
    assert st.shape == (F, N, N)
    spec = o["spectra"].view(np.complex128)
    ora = Oracle("f64")
    worst = 0.0
    for k in range(F):
        sx, sy = o["frame_shift"][k]
        ref = ora.c2r2d(ora.phase_shift(spec[k], sx, sy)) / (N * N)
        worst = max(worst, float(np.linalg.norm(st[k] - ref) / np.linalg.norm(ref)))
    print("corrected stack vs oracle: worst frame rel L2 %.3g" % worst)
    assert worst < 1e-4
