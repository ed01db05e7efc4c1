"""The allocation-free sort entry (bk_sort_pairs_noalloc) against the workspace entry (bk_sort_pairs) through the
C ABI: bit-identical keys and permutations on two calls in a row, the second smaller (so stale scratch from the
first call cannot leak into it), both also equal to numpy's stable argsort; sizes above the static scratch are
rejected on the host before any GPU work."""
import ctypes

import numpy as np
import pytest

pytestmark = pytest.mark.gpu

torch = pytest.importorskip("torch")
if not torch.cuda.is_available():
    pytest.skip("needs a ROCm GPU", allow_module_level=True)

from baking_amd import _native  # noqa: E402


def _sort(lib, keys, values, key_bits, ws=None):
    n = len(keys)
    kin = torch.as_tensor(keys.view(np.int32), device="cuda")
    vin = None if values is None else torch.as_tensor(values, device="cuda")
    kout = torch.empty_like(kin)
    vout = torch.empty(n, dtype=torch.int32, device="cuda")
    vptr = None if vin is None else _native.ptr(vin)
    if ws is None:
        rc = lib.bk_sort_pairs_noalloc(_native.ptr(kin), vptr, _native.ptr(kout), _native.ptr(vout), n, key_bits,
                                       _native.stream_ptr())
        assert rc == 0, lib.bk_noalloc_last_error()
    else:
        _native.check(lib.bk_sort_pairs(ws, _native.ptr(kin), vptr, _native.ptr(kout), _native.ptr(vout), n, key_bits,
                                        _native.stream_ptr()), "bk_sort_pairs")
    torch.cuda.synchronize()
    return kout.cpu().numpy().view(np.uint32), vout.cpu().numpy()


@pytest.mark.parametrize("with_values", [False, True])
def test_noalloc_sort_equals_the_workspace_sort(with_values):
    lib = _native.load()
    cap = lib.bk_sort_pairs_noalloc_capacity()
    assert cap == 1 << 20
    ws = ctypes.c_void_p(lib.bk_workspace_create(cap, 8, 0, 0))
    assert ws.value
    try:
        rng = np.random.default_rng(31 + int(with_values))
        for n, key_bits in ((cap, 32), (300001, 24), (4099, 12)):  # largest first: later calls reuse stale scratch
            hi = (1 << key_bits) if key_bits < 32 else (1 << 32)
            keys = rng.integers(0, min(hi, max(n // 5, 2)), size=n, dtype=np.uint64).astype(np.uint32)
            if key_bits == 32:
                keys[::7] |= np.uint32(0x80000000)  # the top bit set on some keys
            values = rng.permutation(n).astype(np.int32) if with_values else None
            k_na, v_na = _sort(lib, keys, values, key_bits)
            k_ws, v_ws = _sort(lib, keys, values, key_bits, ws)
            np.testing.assert_array_equal(k_na, k_ws)
            np.testing.assert_array_equal(v_na, v_ws)
            order = np.argsort(keys, kind="stable")
            np.testing.assert_array_equal(k_na, keys[order])
            np.testing.assert_array_equal(v_na, order if values is None else values[order])
            print(f"n {n} key_bits {key_bits}: no-alloc sort equals the workspace sort", flush=True)
    finally:
        lib.bk_workspace_destroy(ws)


def test_noalloc_sort_rejects_sizes_beyond_its_scratch():
    lib = _native.load()
    cap = lib.bk_sort_pairs_noalloc_capacity()
    rc = lib.bk_sort_pairs_noalloc(None, None, None, None, cap + 1, 24, _native.stream_ptr())
    assert rc == -1
    assert b"exceed the static scratch" in lib.bk_noalloc_last_error()
    assert lib.bk_sort_pairs_noalloc(None, None, None, None, 0, 24, _native.stream_ptr()) == 0
