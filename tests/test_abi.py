"""The C-ABI library loads and exports every function include/*.h declares
(no compute calls: runs without a GPU)."""
import ctypes
import pathlib
import re

ROOT = pathlib.Path(__file__).resolve().parents[1]


def _declared():
    names = set()
    for h in (ROOT / "include").glob("*.h"):
        txt = re.sub(r"/\*.*?\*/", "", h.read_text(), flags=re.S)
        names |= set(re.findall(r"\b(?:int|const char\s*\*|void)\s+\*?\s*(ebg_\w+)\s*\(", txt))
    return names


def test_header_declares_entry_points():
    names = _declared()
    assert "ebg_rk4_advect" in names and "ebg_ctx_create" in names


def test_library_exports_every_declared_symbol():
    import earthbox_gpu
    lib = ctypes.CDLL(str(earthbox_gpu.LIB_PATH))
    missing = [n for n in sorted(_declared()) if not hasattr(lib, n)]
    assert not missing, missing


def test_python_binding_covers_header():
    from earthbox_gpu import _lib
    assert _declared() <= set(_lib.SIGNATURES), sorted(_declared() - set(_lib.SIGNATURES))
