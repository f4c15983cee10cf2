"""The C-ABI library loads and exports every function include/*.h declares
(no GPU needed: nothing is called that touches a device)."""
import ctypes
import re
from pathlib import Path

import mrt_hip as M

ROOT = Path(__file__).resolve().parents[1]


def declared_functions():
    names = set()
    for h in (ROOT / "include").glob("*.h"):
        text = re.sub(r"/\*.*?\*/", "", h.read_text(), flags=re.S)
        # header-only helpers are not exported
        inline = set(re.findall(r"static\s+inline\s+[\w\s\*]*?\b(mrt_[a-z0-9_]+)\s*\(",
                                text))
        for m in re.finditer(r"\b(mrt_[a-z0-9_]+)\s*\(", text):
            if m.group(1) not in inline:
                names.add(m.group(1))
    return sorted(names)


def test_every_declared_symbol_is_exported():
    lib = ctypes.CDLL(str(M.LIB_PATH))
    names = declared_functions()
    assert len(names) > 40
    missing = [n for n in names if not hasattr(lib, n)]
    assert not missing, missing


def test_version_string():
    assert "gfx950" in M.version()


def test_init_rejects_bad_sizes():
    ctx = ctypes.c_void_p()
    assert M.mrt_init(0, 0, 10, ctypes.byref(ctx)) == M.MRT_E_INVALID
    assert M.mrt_init(0, 10, 0, ctypes.byref(ctx)) == M.MRT_E_INVALID


def test_null_context_is_an_error_not_a_crash():
    assert M.mrt_render(None, 0, 1, 1) == M.MRT_E_INVALID
    assert M.mrt_set_option(None, M.OPT_SEED, 1.0) == M.MRT_E_INVALID
    assert M.mrt_last_error(None) == b"null context"
