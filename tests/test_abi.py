"""The C-ABI library loads and exports every symbol include/beedrill_amd.h
declares (no compute calls: this runs without a GPU)."""
import ctypes
import os
import subprocess

import pytest

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
LIB = os.path.join(ROOT, "beedrill_amd", "libbeedrill_amd.so")


def test_library_exports_header_symbols():
    import beedrill_amd as bd
    assert os.path.exists(LIB), "build the library first (__graft_entry__.build())"
    L = ctypes.CDLL(LIB)
    names = bd.exported_symbols()
    assert len(names) >= 25
    missing = [n for n in names if not hasattr(L, n)]
    assert not missing, missing
    out = subprocess.run(["nm", "-D", "--defined-only", LIB], capture_output=True, text=True,
                         check=True).stdout
    exported = {line.split()[-1] for line in out.splitlines() if " T " in line}
    assert set(names) <= exported


def test_bindings_cover_every_symbol():
    import beedrill_amd as bd
    L = bd.lib()
    for n in bd.exported_symbols():
        assert getattr(L, n).argtypes is not None, n


def test_no_gpu_means_no_context():
    import beedrill_amd as bd
    if bd.device_count() > 0:
        pytest.skip("a GPU is visible")
    with pytest.raises(bd.BeedrillError):
        bd.Context(0)


@pytest.mark.parametrize("name", ["hybrid_bfs", "components"])
def test_cli_binary_built(name):
    exe = os.path.join(ROOT, "beedrill_amd", "bin", name)
    assert os.path.exists(exe), "CLI not built"
    r = subprocess.run([exe, "--help"], capture_output=True, text=True)
    assert r.returncode == 1 and "--graph_filename" in r.stdout
