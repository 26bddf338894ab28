"""C-ABI checks that need no GPU: the library loads, exports every entry
point include/zpic_amd.h declares, and refuses to run without a device."""
import ctypes
import os
import re

import pytest

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
HEADER = os.path.join(ROOT, "include", "zpic_amd.h")
LIB = os.path.join(ROOT, "zpic-parallel_amd", "libzpic_amd.so")


def declared_functions():
    text = open(HEADER).read()
    text = re.sub(r"/\*.*?\*/", "", text, flags=re.S)
    text = re.sub(r"//[^\n]*", "", text)
    names = set(re.findall(r"\b(zpic_[a-z0-9_]+)\s*\(", text))
    return sorted(names)


def test_header_declares_entry_points():
    names = declared_functions()
    for must in ("zpic_push", "zpic_move_deposit", "zpic_push_move_deposit", "zpic_tile_sort",
                 "zpic_current_advance", "zpic_yee", "zpic_simulation_new", "zpic_simulation_advance"):
        assert must in names


@pytest.mark.skipif(not os.path.exists(LIB), reason="libzpic_amd.so not built")
def test_library_exports_every_declared_symbol():
    L = ctypes.CDLL(LIB)
    missing = [n for n in declared_functions() if not hasattr(L, n)]
    assert not missing, missing


@pytest.mark.skipif(not os.path.exists(LIB), reason="libzpic_amd.so not built")
def test_geometry_helper_matches_reference_layout():
    import pyzpic
    g = pyzpic.Geom()
    assert pyzpic.lib().zpic_geom_init(ctypes.byref(g), 8, 8, 16, 16, 1, 2) == 0
    # grid.h:77-92: ext = 1+16+2 = 19, tile_vol = roundup4(361) = 364, offset = 19+1
    assert (g.ext_x, g.ext_y, g.tile_vol, g.offset) == (19, 19, 364, 20)
    assert pyzpic.lib().zpic_geom_init(ctypes.byref(g), 0, 8, 16, 16, 1, 2) != 0
    assert b"invalid" in pyzpic.lib().zpic_last_error()


@pytest.mark.skipif(not os.path.exists(LIB), reason="libzpic_amd.so not built")
def test_stage_entry_points_reject_bad_arguments():
    import pyzpic
    L = pyzpic.lib()
    L.zpic_push.argtypes = [ctypes.c_void_p] * 4 + [ctypes.c_float, ctypes.c_void_p, ctypes.c_int, ctypes.c_void_p]
    assert L.zpic_push(None, None, None, None, 0.0, None, 0, None) == -1
    L.zpic_tile_sort.argtypes = [ctypes.c_void_p] * 3 + [ctypes.c_int] * 6 + [ctypes.c_void_p]
    assert L.zpic_tile_sort(None, None, None, 1, 1, 16, 16, 1, 1, None) == -1


def _has_gpu():
    try:
        import torch
        return torch.cuda.is_available()
    except Exception:
        return False


@pytest.mark.skipif(not os.path.exists(LIB) or _has_gpu(), reason="needs a GPU-less host")
def test_simulation_fails_loudly_without_gpu():
    import pyzpic
    with pytest.raises(pyzpic.ZpicError):
        pyzpic.Simulation((2, 2), (16, 16), (3.2, 3.2), 0.07)


@pytest.mark.skipif(not os.path.exists(LIB), reason="library not built")
def test_large_tiles_fit_lds():
    # host-side sizing: 32x32 tiles cannot hold 8 private J copies in 64 KiB
    import pyzpic
    vol32 = (35 * 35 + 3) & ~3
    t = pyzpic.fitted_threads_per_tile(512, vol32)
    assert 64 <= t < 512 and t % 64 == 0
    assert pyzpic.fitted_threads_per_tile(512, 364) == 512
    assert pyzpic.fitted_threads_per_tile(100, 364) == 0


@pytest.mark.skipif(not os.path.exists(LIB), reason="library not built")
def test_subsample_rejects_mismatched_tilings():
    # the tile tables of both species are sized by their own geometry: a
    # destination with fewer tiles is refused before any device access
    import pyzpic
    L = pyzpic.lib()
    gs, gd = pyzpic.Geom(), pyzpic.Geom()
    assert L.zpic_geom_init(ctypes.byref(gs), 4, 4, 16, 16, 0, 0) == 0
    assert L.zpic_geom_init(ctypes.byref(gd), 2, 4, 16, 16, 0, 0) == 0
    buf = (ctypes.c_byte * 64)()
    L.zpic_subsample.argtypes = [ctypes.c_void_p] * 4 + [ctypes.c_int, ctypes.c_int, ctypes.c_void_p]
    rc = L.zpic_subsample(ctypes.addressof(buf), ctypes.addressof(gs), ctypes.addressof(buf),
                          ctypes.addressof(gd), 4, 0, None)
    assert rc == -1  # ZPIC_EINVAL
    assert b"same tiling" in L.zpic_last_error()
