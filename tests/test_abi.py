"""The C-ABI libraries load without a GPU and export every function the
headers in include/ declare."""
import ctypes
import re
from pathlib import Path

import pytest

ROOT = Path(__file__).resolve().parent.parent


def declared(header: str) -> list[str]:
    text = (ROOT / "include" / header).read_text()
    text = re.sub(r"/\*.*?\*/", "", text, flags=re.S)
    return sorted(set(re.findall(r"\b(pjh?_[a-z0-9_]+)\s*\(", text)))


@pytest.mark.parametrize("header,lib", [("prunejuice.h", "libprunejuice_hip.so"),
                                        ("prunejuice_host.h", "libpj_host.so")])
def test_every_declared_symbol_is_exported(pj, header, lib):
    path = ROOT / "pattenmatching_sharedmemory_amd" / "lib" / lib
    assert path.exists(), f"{lib} not built"
    pj.native.host()  # loads both, device library first
    so = ctypes.CDLL(str(path))
    names = declared(header)
    assert names, header
    missing = [n for n in names if not hasattr(so, n)]
    assert not missing, missing


def test_missing_library_fails_loudly(monkeypatch, tmp_path):
    import importlib

    import pattenmatching_sharedmemory_amd.native as native

    monkeypatch.setattr(native, "_hip", None)
    monkeypatch.setattr(native, "HIP_LIB", tmp_path / "absent.so")
    with pytest.raises(RuntimeError, match="HIP library missing"):
        native.hip()
    importlib.reload(native)


def test_cli_usage_without_arguments(pj):
    # pjh_main prints the usage and returns 255 when -i/-p/-o are missing
    assert pj.run_cli([]) == 255
    assert pj.run_cli(["-i", "x"]) == 255
