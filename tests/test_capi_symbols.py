"""The C ABI library loads and exports every symbol include/wos_amd.h declares."""
import ctypes
import os
import re

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))


def declared():
    src = open(os.path.join(ROOT, "include", "wos_amd.h")).read()
    src = re.sub(r"/\*.*?\*/", "", src, flags=re.S)
    return sorted(set(re.findall(r"\b(wos_[a-z0-9_]+)\s*\(", src)))


def test_header_declares_entry_points():
    names = declared()
    assert "wos_walk_reduce" in names and "wos_scene_create" in names and len(names) >= 18


def test_library_exports_every_declared_symbol():
    import wos_ext
    lib = ctypes.CDLL(wos_ext.library_path)
    missing = [n for n in declared() if not hasattr(lib, n)]
    assert not missing, missing
    assert set(declared()) <= set(wos_ext.EXPORTED)


def test_error_path_without_device_work():
    import wos_ext
    h = ctypes.c_void_p()
    rc = wos_ext.lib.wos_scene_create(4, None, 0, None, 0, None, 0, 0, None, 0, None, ctypes.byref(h))
    assert rc == -1
    assert b"dim must be 2 or 3" in wos_ext.lib.wos_last_error()


def test_tuning_validation_and_window_default():
    """wos_set_tuning touches no device: the window accepts 0 (per-dimension
    default) and rejects negatives; other knobs keep their ranges."""
    import wos_ext
    before = wos_ext.get_tuning()
    try:
        assert wos_ext.lib.wos_set_tuning(128, 8, 6, 0) == 0
        assert wos_ext.get_tuning() == (128, 8, 6, 0)
        assert wos_ext.lib.wos_set_tuning(128, 8, 6, -1) == -1
        assert b"window" in wos_ext.lib.wos_last_error()
        assert wos_ext.lib.wos_set_tuning(32, 8, 6, 8) == -1
        wos_ext.set_tuning(window=12)
        assert wos_ext.get_tuning()[3] == 12
    finally:
        wos_ext.set_tuning(*before)
