"""The C-ABI library loads and exports every symbol include/ declares."""
import pathlib
import re

from zoomy_hip import _lib

ROOT = pathlib.Path(__file__).resolve().parents[1]


def declared():
    names = set()
    for h in (ROOT / "include").glob("*.h"):
        txt = h.read_text()
        txt = re.sub(r"/\*.*?\*/", "", txt, flags=re.S)
        names |= set(re.findall(r"\b(zj_[a-z0-9_]+)\s*\(", txt))
    return sorted(names)


def test_library_exports_declared_symbols():
    L = _lib.lib()
    names = declared()
    assert len(names) >= 10
    for n in names:
        assert hasattr(L, n), n
    assert set(names) == set(_lib.exported_symbols())


def test_version_string():
    assert _lib.lib().zj_version().startswith(b"zoomy_hip")
