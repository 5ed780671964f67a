"""GPU: tl_field_summary_noalloc, the field summary that allocates nothing
(its partial sums live in a static device array of the library), against
tl_field_summary on the same fields: bit-identical results on two calls in a
row, the second on a smaller grid, so nothing of the first call's scratch
contents can leak into the second.

The test's own device buffers come from the HIP runtime the library itself
links (libamdhip64.so.7), through ctypes: a second HIP runtime in the process
(torch's) cannot start once the library's runtime holds the device."""
import ctypes as C

import numpy as np
import pytest

import tealeaf_hip as T
from tealeaf_hip import _lib, decks

pytestmark = pytest.mark.gpu

H2D, D2H = 1, 2  # hipMemcpyHostToDevice, hipMemcpyDeviceToHost


def _hip():
    T.load()  # the library's dependency is loaded first, so this is the same runtime
    h = C.CDLL("libamdhip64.so.7")
    h.hipMalloc.argtypes = [C.POINTER(C.c_void_p), C.c_size_t]
    h.hipFree.argtypes = [C.c_void_p]
    h.hipMemcpy.argtypes = [C.c_void_p, C.c_void_p, C.c_size_t, C.c_int]
    h.hipDeviceSynchronize.argtypes = []
    return h


def _case(n, ny, seed):
    text = decks.random_rect_deck(seed, n=n, steps=1, nrect=6).replace(f"y_cells={n}", f"y_cells={ny}")
    s = T.Settings.from_text(text)
    c = T.initialiseapp(s)
    T.diffuse(c, s, 1)  # energy and u after a solve, not only the initial state
    want = T.fieldsummary(c)
    fields = {f: np.ascontiguousarray(c.field(f)) for f in ("density", "energy", "u")}
    c.close()
    return s, want, fields


def test_field_summary_noalloc_matches_context_summary():
    L = T.load()
    hip = _hip()
    for n, ny, seed in ((300, 170, 5), (70, 41, 6)):  # the second call on a smaller grid
        print(f"case: field summary without allocation, {n}x{ny} seed {seed}", flush=True)
        s, want, fields = _case(n, ny, seed)
        dev = {}
        try:
            for k, v in fields.items():
                ptr = C.c_void_p()
                assert hip.hipMalloc(C.byref(ptr), v.nbytes) == 0
                dev[k] = ptr
                assert hip.hipMemcpy(ptr, v.ctypes.data, v.nbytes, H2D) == 0
            out = C.c_void_p()
            assert hip.hipMalloc(C.byref(out), 32) == 0
            dev["out"] = out
            rc = L.tl_field_summary_noalloc(dev["density"], dev["energy"], dev["u"], n, ny, s.halodepth, s.dx,
                                            s.dy, out, None)
            assert rc == _lib.TL_OK
            assert hip.hipDeviceSynchronize() == 0
            got = np.full(4, np.nan)
            assert hip.hipMemcpy(got.ctypes.data, out, 32, D2H) == 0
        finally:
            for ptr in dev.values():
                hip.hipFree(ptr)
        assert np.array_equal(got, want), (got, want)


def test_field_summary_noalloc_rejects_oversized_and_invalid():
    print("case: field summary without allocation, invalid arguments", flush=True)
    L = T.load()
    p = C.c_void_p(16)  # never dereferenced: the checks come first
    # 65536^2 needs 16384 summation blocks, beyond the static array
    assert L.tl_field_summary_noalloc(p, p, p, 65536, 65536, 2, 1.0, 1.0, p, None) == _lib.TL_E_INVALID
    assert L.tl_field_summary_noalloc(None, p, p, 10, 10, 2, 1.0, 1.0, p, None) == _lib.TL_E_INVALID
    assert L.tl_field_summary_noalloc(p, p, p, 0, 10, 2, 1.0, 1.0, p, None) == _lib.TL_E_INVALID
