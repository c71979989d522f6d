"""The C-ABI libraries load and export every entry point include/*.h declares
(CPU only: no compute call is made)."""
import ctypes as C
import re
from pathlib import Path

import pytest

ROOT = Path(__file__).resolve().parent.parent
DECL = re.compile(r"^\s*(?:const\s+)?[\w]+\s*\**\s*\b(ll_(?:gpu|host)_\w+)\s*\(", re.M)


def declared(header: str) -> list[str]:
    text = (ROOT / "include" / header).read_text()
    text = re.sub(r"/\*.*?\*/", "", text, flags=re.S)
    return sorted(set(DECL.findall(text)))


@pytest.mark.parametrize("header,lib", [("ll_gpu.h", "libll_gpu.so"),
                                        ("ll_host.h", "libll_host.so")])
def test_every_declared_symbol_is_exported(header, lib):
    names = declared(header)
    assert len(names) > 10
    so = C.CDLL(str(ROOT / "llama_amd" / lib))
    missing = [n for n in names if not hasattr(so, n)]
    assert not missing, f"{lib} lacks {missing}"


def test_python_bindings_cover_the_headers():
    from llama_amd import _native as N
    gpu = set(declared("ll_gpu.h"))
    host = set(declared("ll_host.h"))
    assert gpu <= set(N.GPU_SIGNATURES) | {"ll_gpu_stream"}
    assert host <= set(N.HOST_SIGNATURES)


def test_create_without_gpu_returns_status():
    """On a machine without a usable GPU the library reports, never aborts."""
    from llama_amd import _native as N
    h = C.c_void_p()
    rc = N.gpu().ll_gpu_create(0, C.byref(h))
    if rc == 0:
        N.gpu().ll_gpu_destroy(h)
    else:
        assert rc in (-1, -3)
        assert N.gpu().ll_gpu_last_error(None)  # NULL context still gives a message
    assert N.gpu().ll_gpu_create(-1, C.byref(h)) != 0
    assert N.gpu().ll_gpu_pagerank(None, -1, 1, 0.85, None, None) == -1


def test_no_cpu_fallback_in_product_package():
    """The product package never imports the oracle."""
    for p in (ROOT / "llama_amd").rglob("*.py"):
        src = p.read_text()
        assert "import oracle" not in src and "from oracle" not in src, p
