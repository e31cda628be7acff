"""The sdddst executable on the device against the oracle's stepping
driver: result configuration and log of a zero-stress run with the
reference's defaults, and a fixed-rate run with point defects (A and KASQR
scaled by the command line, project_parser.cpp:208-212)."""
import os
import subprocess

import numpy as np
import pytest

import oracle_lib as O
from sdddst_amd import configs
from sdddst_amd.simulation import read_log

pytestmark = pytest.mark.gpu
ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
EXE = os.path.join(ROOT, "sdddst_amd", "sdddst")


def _cli(*args):
    r = subprocess.run([EXE, "-c", *args], capture_output=True, text=True, timeout=240)
    assert r.returncode == 0, r.stderr
    return r


def test_cli_zero_stress_defaults_match_oracle(tmp_path):
    n = 128
    x, y, b = configs.dislocations(n, 21)
    d, out, log = tmp_path / "in.dconf", tmp_path / "out.dconf", tmp_path / "run.log"
    configs.write_dconf(str(d), x, y, b)
    _cli("--dislocation-configuration", str(d), "--result-dislocation-configuration", str(out),
         "--logfile-path", str(log), "--step-count-limit", "25", "--calculate-strain", "--field-mode", "exact",
         "--linear-solver", "dense-lu")
    # reference defaults: cutoff multiplier 1e20, precision 1e-5, h0 1e-6
    x0, y0, b0 = configs.read_dconf(str(d))
    o = O.OracleSim(x0, y0, b0, cutoff_multiplier=1e20, prec=1e-5, h0=1e-6, stress="zero")
    while o.succesful_steps < 25:
        o.step()
    xr, yr, br = configs.read_dconf(str(out))
    assert np.array_equal(yr, y0) and np.array_equal(br, b0)
    assert np.max(np.abs(xr - o.x())) < 1e-9
    lg, lo = read_log(str(log)), o.log()
    assert lg.shape == lo.shape
    cols = [0, 1, 2, 4, 5, 6, 7, 9, 10, 11]
    a, r = lg[:, cols], lo[:, cols]
    assert (np.isnan(a) == np.isnan(r)).all()
    fin = ~np.isnan(r)
    assert np.max(np.abs(a[fin] - r[fin]) / np.maximum(np.abs(r[fin]), 1e-300)) < 1e-6
    o.close()


def test_cli_point_defects_fixed_rate_match_oracle(tmp_path):
    n, pc = 96, 24
    x, y, b = configs.dislocations(n, 22)
    px, py = configs.point_defects(n, pc, 22)
    d, p, out = tmp_path / "in.dconf", tmp_path / "pd.dconf", tmp_path / "out.dconf"
    configs.write_dconf(str(d), x, y, b)
    with open(p, "w") as f:
        for u, v in zip(px, py):
            f.write("%.16e %.16e\n" % (u, v))
    _cli("--dislocation-configuration", str(d), "--point-defect-configuration", str(p),
         "--result-dislocation-configuration", str(out), "--step-count-limit", "8", "--cutoff-multiplier", "2",
         "--position-precision", "1e-6", "--fixed-rate-external-stress", "50", "--field-mode", "exact",
         "--linear-solver", "dense-lu")
    x0, y0, b0 = configs.read_dconf(str(d))
    p0 = np.loadtxt(str(p), ndmin=2)
    A = (1e-4 * 16.0) * (1.0 / np.sqrt(n))        # A *= 1./sqrt(dc)
    K = 1.65 * 1.65 * 1e6 / 256.0 * n
    o = O.OracleSim(x0, y0, b0, cutoff_multiplier=2.0, prec=1e-6, h0=1e-6, stress="fixed-rate", p1=50.0,
                    px=p0[:, 0].copy(), py=p0[:, 1].copy(), A=A, K=K)
    while o.succesful_steps < 8:
        o.step()
    xr, _, _ = configs.read_dconf(str(out))
    assert np.max(np.abs(xr - o.x())) < 1e-9
    o.close()


def test_cli_ev_analysis_series(tmp_path):
    # TimeSeriesProcessor::run (time_series_processor.cpp:21-93): one gzip
    # text record per snapshot -- time, eigenvalues, participation numbers,
    # s^2 / lambda, speeds, x, s, the first K eigenvectors -- with the
    # device decomposition's values at six significant digits
    import gzip

    import sdddst_amd as S

    n, pc, K = 64, 8, 3
    px, py = configs.point_defects(n, pc, 41)
    pfile, lst, out = tmp_path / "pd.fconf", tmp_path / "list.txt", tmp_path / "ev.gz"
    with open(pfile, "w") as f:
        for u, v in zip(px, py):
            f.write("%.16e %.16e\n" % (u, v))
    snaps, lines = [], []
    for s, t in enumerate([0.0, 1.5e-3]):
        x, y, b = configs.dislocations(n, 40 + s)
        d = tmp_path / ("s%d.dconf" % s)
        configs.write_dconf(str(d), x, y, b)
        lines.append("%r %s" % (t, d))
        snaps.append((t,) + configs.read_dconf(str(d)))
    lst.write_text("\n".join(lines) + "\n")
    r = _cli("--ev-analyzation", "--result-ev-file", str(out), "--file-list", str(lst), "--pd-configuration",
             str(pfile), "--fixed-rate-external-stress", "2", "--eigenvector-to-write", str(K))
    assert r.stdout.split() == ["0", "0.0015"]
    rows = [np.array(list(map(float, ln.split()))) for ln in gzip.open(out, "rt").read().splitlines()]
    assert len(rows) == 2
    for (t, x, y, b), row in zip(snaps, rows):
        assert len(row) == 1 + 6 * n + K * n and np.all(np.isfinite(row))
        assert row[0] == pytest.approx(t)
        with S.Context(n, pc=pc) as c:
            c.upload_state(0, x, y, b)
            c.upload_points(px, py)
            ev = c.ev_decompose(0)
        lam = ev["eigenvalues"]
        np.testing.assert_allclose(row[1:1 + n], lam, rtol=1e-5, atol=1e-6 * np.abs(lam).max())
        np.testing.assert_allclose(row[1 + n:1 + 2 * n], ev["pn"], rtol=1e-5)
        np.testing.assert_allclose(row[1 + 4 * n:1 + 5 * n], x, rtol=1e-5, atol=1e-6)
        vecs = row[1 + 6 * n:].reshape(K, n)
        np.testing.assert_allclose(np.abs(vecs), np.abs(ev["eigenvectors"][:K]), rtol=1e-4, atol=1e-6)
    # --calculate-ev-deriv (time_series_processor.cpp:92-111): the same
    # records with the eigenvalue derivatives appended, from the speeds under
    # the fixed-rate stress 2 t (point-defect constants scaled by the
    # snapshot's size, data_time_series.cpp) and sigma' = 2
    out2 = tmp_path / "ev_deriv.gz"
    r = _cli("--ev-analyzation", "--result-ev-file", str(out2), "--file-list", str(lst), "--pd-configuration",
             str(pfile), "--fixed-rate-external-stress", "2", "--eigenvector-to-write", str(K),
             "--calculate-ev-deriv")
    rows2 = [np.array(list(map(float, ln.split()))) for ln in gzip.open(out2, "rt").read().splitlines()]
    assert len(rows2) == 2
    for (t, x, y, b), row, row2 in zip(snaps, rows, rows2):
        assert len(row2) == len(row) + n and np.array_equal(row2[:len(row)], row)
        with S.Context(n, pc=pc, A=1e-4 * 16.0 / np.sqrt(n), KASQR=1.65 * 1.65 * 1e6 / 256.0 * n) as c:
            c.upload_state(0, x, y, b)
            c.upload_points(px, py)
            v = c.speeds(0, S.V_SCRATCH, sigma=2.0 * t, update_tol=False)
            np.testing.assert_allclose(row[1 + 3 * n:1 + 4 * n], v, rtol=1e-5, atol=1e-6 * np.abs(v).max())
            d = c.ev_decompose(0, speeds=v, stress_rate=2.0)["deriv"]
        np.testing.assert_allclose(row2[len(row):], d, rtol=1e-5, atol=1e-6 * np.abs(d).max())
