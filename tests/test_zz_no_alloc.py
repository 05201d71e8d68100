"""The allocation-free variant of redistribute_mass_area_weighted, through
the C ABI: chysx_cloth_redistribute_mass_area_weighted_noalloc stages its
vertex-to-triangle incidence in the library's fixed device scratch array
instead of allocating it.  It runs the same kernel in the same order, so its
inverse masses equal the allocating entry point's bitwise, on consecutive
calls whose second input is smaller than the first (the scratch keeps the
first call's contents beyond the second's).  Sizes beyond the scratch are
refused with CHYSX_EINVAL before any device work."""

import ctypes as C
import os

import numpy as np
import pytest

from chysx import scenes

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
LIB = os.path.join(ROOT, "ai-phyiscs_amd", "chysx", "libchysx_amd.so")
SCRATCH_BYTES = 64 << 20


def lib():
    L = C.CDLL(LIB)
    L.chysx_last_error.restype = C.c_char_p
    for name in ("chysx_cloth_redistribute_mass_area_weighted",
                 "chysx_cloth_redistribute_mass_area_weighted_noalloc"):
        getattr(L, name).argtypes = [C.c_void_p, C.c_float, C.c_size_t, C.c_int, C.c_size_t]
    L.chysx_cloth_set_external_buffers.argtypes = [C.c_void_p, C.c_size_t, C.c_size_t, C.c_int, C.c_size_t,
                                                   C.c_size_t]
    L.chysx_cloth_set_mesh.argtypes = [C.c_void_p, C.c_void_p, C.c_int]
    return L


def make_cloth(L, pos_ptr, n, tris):
    h = C.c_void_p()
    assert L.chysx_cloth_create(C.byref(h)) == 0
    assert L.chysx_cloth_set_external_buffers(h, pos_ptr, pos_ptr, n, 0, 0) == 0
    t = np.ascontiguousarray(tris, np.int32)
    assert L.chysx_cloth_set_mesh(h, t.ctypes.data, t.shape[0]) == 0
    return h


def test_noalloc_refuses_sizes_beyond_its_scratch():
    """(n + 1 + 6 T) * 4 bytes above 64 MiB: a 1200 x 1200 grid needs 74.7 MB.
    Refused before any device work (the position pointer is never read)."""
    L = lib()
    w = 1200
    _, tris = scenes.grid(w)
    n = w * w
    assert (n + 1 + 6 * tris.shape[0]) * 4 > SCRATCH_BYTES
    h = make_cloth(L, 4096, n, tris)
    rc = L.chysx_cloth_redistribute_mass_area_weighted_noalloc(h, 0.3, 4096, n, 0)
    assert rc == 1
    assert b"scratch" in L.chysx_last_error()
    assert L.chysx_cloth_destroy(h) == 0


@pytest.mark.gpu
def test_noalloc_matches_allocating_entry_point(cuda):
    import torch
    L = lib()
    rng = np.random.default_rng(17)
    for w in (300, 40):  # the second call's input is smaller
        pos, tris = scenes.grid(w, 1.0)
        pos = (pos + rng.uniform(-0.002, 0.002, pos.shape)).astype(np.float32)
        n = pos.shape[0]
        dpos = torch.as_tensor(pos, device="cuda:0").contiguous()
        h = make_cloth(L, dpos.data_ptr(), n, tris)
        ref = torch.full((n,), -1.0, dtype=torch.float32, device="cuda:0")
        got = torch.full((n,), -2.0, dtype=torch.float32, device="cuda:0")
        rc = L.chysx_cloth_redistribute_mass_area_weighted(h, 0.3, ref.data_ptr(), n, 0)
        assert rc == 0, L.chysx_last_error()
        rc = L.chysx_cloth_redistribute_mass_area_weighted_noalloc(h, 0.3, got.data_ptr(), n, 0)
        assert rc == 0, L.chysx_last_error()
        torch.cuda.synchronize()
        r, g = ref.cpu().numpy(), got.cpu().numpy()
        np.testing.assert_array_equal(g, r)
        expect = 1.0 / (0.3 * scenes.vertex_areas(pos, tris))
        np.testing.assert_allclose(g, expect, rtol=1e-5)
        assert L.chysx_cloth_destroy(h) == 0
