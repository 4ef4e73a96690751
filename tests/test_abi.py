"""The C ABI library: it loads, and it exports every entry point the header declares.
No compute calls here (these run without a GPU)."""
import ctypes
import pathlib
import re

import pytest

from embloks_amd import gpusim

ROOT = pathlib.Path(__file__).resolve().parent.parent


def declared():
    text = (ROOT / "include" / "embloks_hip.h").read_text()
    return sorted(set(re.findall(r"^\s*(?:const\s+)?[\w]+\s*\*?\s*(emb_\w+)\s*\(", text, re.M)))


def test_header_declarations_are_all_bound():
    names = declared()
    assert len(names) >= 20
    assert set(names) == set(gpusim.SIGNATURES), set(names) ^ set(gpusim.SIGNATURES)


def test_library_exports_every_declared_symbol():
    if not gpusim.LIB_PATH.exists():
        pytest.fail("libembloks_hip.so not built")
    lib = ctypes.CDLL(str(gpusim.LIB_PATH))
    for name in declared():
        assert hasattr(lib, name), name


def test_settings_defaults_match_python_mirror():
    lib = gpusim.lib()
    from embloks_amd import abi
    s = abi.EmbSettings()
    lib.emb_settings_default(ctypes.byref(s))
    want = abi.settings()
    for name, _ in abi.EmbSettings._fields_:
        assert getattr(s, name) == getattr(want, name), name


def test_footprint_grows_with_the_scene():
    from embloks_amd import abi
    s = abi.settings()
    small = gpusim.footprint(s, 64, 64, 1000)
    big = gpusim.footprint(s, 1024, 1024, 1_000_000)
    assert 0 < small < big
    assert gpusim.footprint(s, 0, 10, 0) < 0
