"""The C-ABI library builds, loads without a GPU and exports every declared entry point."""
import re
import os

import pytest

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))


def declared():
    hdr = open(os.path.join(ROOT, "include", "fereg_hip.h")).read()
    return sorted(set(re.findall(r"\b(fh_[a-z0-9_]+)\s*\(", hdr)))


def test_library_exports_every_declared_symbol():
    import fereg_amd
    L = fereg_amd.lib()
    names = declared()
    assert len(names) >= 30
    missing = [n for n in names if not hasattr(L, n)]
    assert not missing, missing


def test_python_mirror_declares_every_symbol():
    from fereg_amd import _lib
    assert sorted(n for n, _, _ in _lib.SIGNATURES) == declared()


def test_error_codes_match_the_reference():
    """error.rs:25-37: InvalidInput 198, Convergence 430, OutOfMemory 909, TooFewObs 2001,
    Interrupted 1, Internal 9999."""
    import fereg_amd as F
    assert (F.FH_INVALID, F.FH_CONVERGENCE, F.FH_OOM, F.FH_TOO_FEW, F.FH_INTERRUPTED, F.FH_INTERNAL) == \
        (198, 430, 909, 2001, 1, 9999)


def test_no_cpu_fallback(tmp_path, monkeypatch):
    """Without the built library the package raises instead of computing anything."""
    from fereg_amd import _lib
    monkeypatch.setattr(_lib, "library_path", str(tmp_path / "missing.so"))
    monkeypatch.setattr(_lib, "_lib", None)
    with pytest.raises(ImportError):
        _lib.lib()
