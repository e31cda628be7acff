"""The sdddst executable's command line (project_parser.cpp:29-299): option
names and defaults, the reference's error messages and exit statuses.  No
device is needed for any of these."""
import os
import re
import subprocess

import pytest

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
EXE = os.path.join(ROOT, "sdddst_amd", "sdddst")
REFERENCE_OPTIONS = [
    "simulation", "ev-analyzation", "dislocation-configuration", "result-dislocation-configuration",
    "point-defect-configuration", "logfile-path", "time-limit", "speed-limit", "step-count-limit",
    "strain-increase-limit", "stress-limit", "avalanche-detection-limit", "avalanche-speed-threshold",
    "initial-stepsize", "cutoff-multiplier", "max-stepsize", "calculate-strain", "calculate-order-parameter",
    "position-precision", "save-sub-configurations", "sub-configuration-delay",
    "sub-configuration-delay-during-avalanche", "change-cutoff-to-inf-under-threshold", "post-relax",
    "periodic-stress-field-analytic", "no-external-stress", "fixed-rate-external-stress", "spring-constant",
    "file-list", "pd-configuration", "result-ev-file", "calculate-ev-deriv", "eigenvector-to-write",
    "write-correlation-matrices", "help", "hide-copyright"]


def run(*args):
    if not os.path.exists(EXE):
        pytest.fail("sdddst_amd/sdddst is not built (run __graft_entry__.build())")
    return subprocess.run([EXE, *args], capture_output=True, text=True, timeout=60)


def test_help_lists_every_reference_option_with_defaults():
    r = run("--help")
    assert r.returncode == 0
    assert "SDDDST" in r.stdout                  # licence first, as the reference
    names = set(re.findall(r"--([a-z0-9-]+)", r.stdout))
    assert set(REFERENCE_OPTIONS) <= names
    for opt, default in [("initial-stepsize", "1e-6"), ("cutoff-multiplier", "1e20"),
                         ("position-precision", "1e-5"), ("avalanche-speed-threshold", "1e-3"),
                         ("sub-configuration-delay", "5"), ("sub-configuration-delay-during-avalanche", "1"),
                         ("post-relax", "0"), ("eigenvector-to-write", "10")]:
        assert re.search(rf"--{opt} arg \(={re.escape(default)}\)", r.stdout), opt


def test_hide_copyright():
    r = run("-c", "--help")
    assert r.returncode == 0 and "Copyright" not in r.stdout and "ABSOLUTELY" not in r.stdout


@pytest.mark.parametrize("args,msg,code", [
    ([], "dislocation-configuration is missing!", 255),
    (["--dislocation-configuration", "a"], "result-dislocation-configuration is missing!", 255),
    (["--no-such-option"], "unrecognised option '--no-such-option'", 255),
    (["--step-count-limit", "-3"], "is invalid", 255),
    (["--time-limit", "abc"], "is invalid", 255),
    (["--time-limit"], "is missing", 255),
    (["--help=1"], "does not take any arguments", 255),
    (["--time-limit", "1", "--time-limit", "2"], "more than once", 255),
    (["--ev-analyzation"], "result-ev-file is not defined! Exiting...", 1),
])
def test_errors_as_the_reference(args, msg, code):
    r = run("-c", *args)
    assert r.returncode == code
    assert msg in r.stderr


def test_conflicting_stress_protocols(tmp_path):
    d = tmp_path / "a.dconf"
    d.write_text("0.1 0.2 1\n-0.1 0.3 -1\n")
    r = run("-c", "--dislocation-configuration", str(d), "--result-dislocation-configuration", str(tmp_path / "o"),
            "--no-external-stress", "--fixed-rate-external-stress=1")
    assert r.returncode == 236        # exit(-20)
    assert "Only one stress protocol" in r.stderr


def test_ev_analysis_requires_fixed_rate_and_series(tmp_path):
    # project_parser.cpp:340-348: without a fixed-rate protocol the reference
    # aborts; without the snapshot list there is nothing to analyse
    r = run("-c", "--ev-analyzation", "--result-ev-file", str(tmp_path / "ev"))
    assert r.returncode == -6
    r = run("-c", "--ev-analyzation", "--result-ev-file", str(tmp_path / "ev"), "--fixed-rate-external-stress", "1")
    assert r.returncode == 1 and "--file-list and --pd-configuration are required" in r.stderr


def test_simulation_fails_loudly_without_device(tmp_path):
    try:
        import torch
        if torch.cuda.is_available():
            pytest.skip("a device is present")
    except ImportError:
        pass
    d = tmp_path / "a.dconf"
    d.write_text("0.1 0.2 1\n-0.1 0.3 -1\n")
    r = run("-c", "--dislocation-configuration", str(d), "--result-dislocation-configuration",
            str(tmp_path / "o.dconf"), "--step-count-limit", "1")
    assert r.returncode == 1 and r.stderr.startswith("sdddst: ")
    assert not (tmp_path / "o.dconf").exists()
