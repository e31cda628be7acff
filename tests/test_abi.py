"""The C ABI libraries load and export every entry point their headers
declare (CPU: no device calls)."""
import ctypes
import os
import re

import pytest

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))


def _declared(header):
    text = open(os.path.join(ROOT, header)).read()
    text = re.sub(r"/\*.*?\*/", "", text, flags=re.S)
    return sorted(set(re.findall(r"\b(sdddst_\w+)\s*\(", text)))


@pytest.mark.parametrize("header,lib", [("include/sdddst_amd.h", "sdddst_amd/libsdddst_amd.so"),
                                        ("include/sdddst/core_capi.h", "sdddst_amd/libsdddstCore.so")])
def test_exports(header, lib):
    path = os.path.join(ROOT, lib)
    if not os.path.exists(path):
        pytest.fail(f"{lib} is not built (run __graft_entry__.build())")
    so = ctypes.CDLL(path)
    names = _declared(header)
    assert len(names) > 10
    missing = [n for n in names if not hasattr(so, n)]
    assert not missing, missing


def test_python_binding_covers_abi():
    from sdddst_amd import exported_symbols
    declared = set(_declared("include/sdddst_amd.h"))
    assert declared == set(exported_symbols())


def test_context_creation_fails_loudly_without_device():
    import sdddst_amd as S
    try:
        import torch  # noqa: F401
        has_gpu = torch.cuda.is_available()
    except Exception:
        has_gpu = False
    if has_gpu:
        pytest.skip("a GPU is present")
    with pytest.raises(S.SdddstError):
        S.Context(16)


def test_bench_refuses_multi_gpu():
    import subprocess
    import sys
    r = subprocess.run([sys.executable, os.path.join(ROOT, "bench.py"), "--gpus", "2"], capture_output=True,
                       text=True, timeout=120)
    assert r.returncode == 2 and "one GPU" in r.stderr


def test_fp64_probe_rejects_bad_iteration_count():
    # checked before any device call
    from sdddst_amd._capi import SdddstError, fp64_probe
    with pytest.raises(SdddstError, match="iters"):
        fp64_probe(0)
