"""The C-ABI library loads and exports every symbol include/fvk.h declares
(no device calls)."""
import ctypes
import os
import re

from tests.helpers import ROOT, require_lib


def declared_symbols():
    src = open(os.path.join(ROOT, "include", "fvk.h")).read()
    src = re.sub(r"/\*.*?\*/", "", src, flags=re.S)
    return sorted(set(re.findall(r"\b(fvk_[a-z0-9_]+)\s*\(", src)))


def test_header_symbols_exported():
    fvk = require_lib()
    lib = ctypes.CDLL(fvk.LIB_PATH)
    names = declared_symbols()
    assert len(names) > 30
    missing = [n for n in names if not hasattr(lib, n)]
    assert not missing, missing


def test_error_reporting_without_device_calls():
    fvk = require_lib()
    import pytest
    with pytest.raises(fvk.FvkError) as e:
        fvk.Mesh.read_neu("/nonexistent/file.neu")
    assert e.value.status == -5
    with pytest.raises(fvk.FvkError):
        fvk.Mesh.box(0, 3, 3)


def test_bench_refuses_more_than_one_gpu():
    """bench.py measures one GPU: --gpus N > 1 exits with status 2 and says so
    before touching a device."""
    import subprocess
    import sys
    r = subprocess.run([sys.executable, os.path.join(ROOT, "bench.py"), "--gpus", "2"], capture_output=True,
                       text=True, timeout=120)
    assert r.returncode == 2
    assert "one GPU" in r.stderr
