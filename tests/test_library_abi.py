"""The C-ABI library loads and exports every entry point include/bug.h declares (CPU only)."""
import ctypes
import os
import re

from bondupdategalerkin_amd import _lib

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))


def declared_symbols():
    names = set()
    for fn in os.listdir(os.path.join(ROOT, "include")):
        if not fn.endswith(".h"):
            continue
        text = open(os.path.join(ROOT, "include", fn)).read()
        text = re.sub(r"/\*.*?\*/", "", text, flags=re.S)
        names |= set(re.findall(r"^\s*(?:const\s+)?\w+\s*\*?\s*(bug_\w+)\s*\(", text, flags=re.M))
    return names


def test_header_declares_entry_points():
    names = declared_symbols()
    assert "bug_apply_one_site" in names and "bug_ctx_create" in names


def test_library_exports_every_declared_symbol():
    handle = ctypes.CDLL(_lib.LIB_PATH)
    missing = [n for n in sorted(declared_symbols()) if not hasattr(handle, n)]
    assert not missing, f"symbols declared but not exported: {missing}"


def test_python_binding_covers_header():
    assert declared_symbols() <= set(_lib.SIGNATURES), \
        sorted(declared_symbols() - set(_lib.SIGNATURES))


def test_ctx_create_without_gpu_fails_cleanly():
    # no GPU in the build container: context creation must return an error code, not abort
    import torch

    if torch.cuda.is_available():
        return
    h = _lib.lib()
    out = ctypes.c_void_p()
    rc = h.bug_ctx_create(0, ctypes.byref(out))
    assert rc == _lib.BUG_EDEVICE and not out.value


def test_schedule_table_file_roundtrip(tmp_path):
    """The shipped gfx950 tuning table loads, saves and reloads without a device; a malformed
    file and an out-of-range mode are refused; the process-wide mode is restored."""
    _lib.autotune_clear()
    try:
        _lib.autotune_load(_lib.SCHEDULE_TABLE)
        n, _ = _lib.autotune_count()
        rows = [ln for ln in open(_lib.SCHEDULE_TABLE) if not ln.startswith("#")]
        assert n == len(rows) > 0
        out = str(tmp_path / "t.txt")
        _lib.autotune_save(out)
        _lib.autotune_clear()
        assert _lib.autotune_count()[0] == 0
        _lib.autotune_load(out)
        assert _lib.autotune_count()[0] == n
        for bad in ("0123 9 1\n", "0123 1 0\n", "xyz\n"):
            p = tmp_path / "bad.txt"
            p.write_text(bad)
            try:
                _lib.autotune_load(str(p))
            except ValueError:
                pass
            else:
                raise AssertionError(f"accepted {bad!r}")
        try:
            _lib.set_autotune(7)
        except ValueError:
            pass
        else:
            raise AssertionError("accepted autotune mode 7")
        _lib.set_autotune(_lib.AUTOTUNE_USE)
    finally:
        _lib.set_autotune(_lib.AUTOTUNE_OFF)
        _lib.autotune_clear()
