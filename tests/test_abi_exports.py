"""The C-ABI libraries load (no GPU needed) and export every function their
headers declare."""
import ctypes
import os
import re
import subprocess

import pytest

from conftest import ROOT

HEADERS = {
    "include/lsqt_hip.h": "gpuqt_amd/liblsqt_hip.so",
    "include/lsqt_host.h": "gpuqt_amd/liblsqt_host.so",
}


def declared_functions(header):
    text = open(os.path.join(ROOT, header)).read()
    text = re.sub(r"/\*.*?\*/", "", text, flags=re.S)
    names = re.findall(r"^\s*(?:const\s+)?[A-Za-z_][\w\s\*]*?\b(lsqt_\w+)\s*\(", text, flags=re.M)
    return sorted(set(names))


@pytest.mark.parametrize("header,lib", list(HEADERS.items()))
def test_exports(header, lib):
    path = os.path.join(ROOT, lib)
    assert os.path.exists(path), f"{lib} not built"
    handle = ctypes.CDLL(path)
    names = declared_functions(header)
    assert len(names) >= 8
    missing = [n for n in names if not hasattr(handle, n)]
    assert not missing, f"{lib} lacks {missing}"


def test_python_binding_covers_header():
    from gpuqt_amd import hostlib, native
    assert set(declared_functions("include/lsqt_hip.h")) == set(native.SIGNATURES)
    assert set(declared_functions("include/lsqt_host.h")) == set(hostlib.SIGNATURES)


def test_no_gpu_errors_cleanly():
    """Without a GPU the device library reports an error instead of crashing."""
    from gpuqt_amd import native
    if native.device_count() > 0:
        pytest.skip("a GPU is visible")
    with pytest.raises(native.LsqtError):
        native.Context(0)


def test_strerror():
    from gpuqt_amd import native
    lib = native.load()
    assert lib.lsqt_strerror(0) == b"ok"
    assert b"invalid" in lib.lsqt_strerror(1)
    assert lib.lsqt_version().startswith(b"lsqt_hip")


def test_cli_usage():
    cli = os.path.join(ROOT, "gpuqt_amd", "lsqt_gpu")
    r = subprocess.run([cli], capture_output=True, text=True)
    assert r.returncode == 1 and "Usage" in r.stdout
    r = subprocess.run([cli, "/nonexistent/driver.txt"], capture_output=True, text=True)
    assert r.returncode == 1 and "Failed to open" in r.stdout


def test_cli_input_error_exit(tmp_path):
    import shutil
    from conftest import GOLDEN
    from gpuqt_amd import hostlib
    src = os.path.join(GOLDEN, "failures", "unknown_keyword")
    work = tmp_path / "bad"
    shutil.copytree(src, work)
    driver = tmp_path / "input.txt"
    driver.write_text(f"{work}\n")
    r = hostlib.run_cli(str(driver), timeout=120)
    assert r.returncode == 1
    assert "Error: Unknown identifier in" in r.stdout
