"""CPU-side checks of the drop-in boundary: the C-ABI library loads and
exports every entry point include/*.h declares (no GPU calls)."""
import glob
import os
import re
import subprocess

import pytest

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
HEADERS = sorted(glob.glob(os.path.join(ROOT, "include", "*.h")))


def declared():
    names = set()
    for h in HEADERS:
        src = re.sub(r"/\*.*?\*/", "", open(h).read(), flags=re.S)
        names |= set(re.findall(r"\b(efs_[a-z0-9_]+)\s*\(", src))
    return sorted(names)


def test_header_declares_entry_points():
    names = declared()
    for must in ("efs_create", "efs_init", "efs_iterate", "efs_get_fields", "efs_cfg_pml", "efs_cfg_nf2ff",
                 "efs_farfield", "efs_energy", "efs_batch_add", "efs_batch_iterate", "efs_batch_tune"):
        assert must in names


def test_library_exports_every_declared_symbol():
    from efs_amd import engine
    lib = engine.load_library()
    out = subprocess.run(["nm", "-D", "--defined-only", engine.LIB_PATH], capture_output=True, text=True).stdout
    exported = set(re.findall(r"\bT (efs_[a-z0-9_]+)", out))
    for name in declared():
        assert name in exported, name
        assert hasattr(lib, name)
    assert set(engine.SIGNATURES) == set(declared())


def test_version_string():
    from efs_amd import engine
    assert engine.load_library().efs_version().decode().startswith("efs-mi355x")


def test_missing_library_fails_loudly(tmp_path):
    from efs_amd import engine
    saved = engine._lib
    engine._lib = None
    try:
        with pytest.raises(RuntimeError):
            engine.load_library(str(tmp_path / "absent.so"))
    finally:
        engine._lib = saved
