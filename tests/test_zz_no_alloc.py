"""The allocation-free CA-CFAR entry point (zkx_ca_cfar_f32_noalloc) against
zkx_ca_cfar_f32 through the C ABI: bit-identical thresholds and detections on
two calls in a row (the second on a smaller map, so stale scratch contents
would show), on both 2-D code paths (width a multiple of 4, and not) and in
1-D; maps beyond the fixed scratch are refused before any launch."""
import ctypes

import numpy as np
import pytest

torch = pytest.importorskip("torch")

pytestmark = pytest.mark.gpu


@pytest.fixture(scope="module")
def lib():
    from zkx_dev_amd._lib import lib as get
    assert torch.cuda.is_available()
    return get()


def _ptr(t):
    return ctypes.c_void_p(t.data_ptr())


def _cfar(lib, fn, x, guard, ref, pfa, with_ws=False):
    rank = x.dim()
    thr = torch.full(x.shape, -1.0, dtype=torch.float32, device="cuda")
    det = torch.full(x.shape, 7, dtype=torch.uint8, device="cuda")
    shape = (ctypes.c_int * rank)(*x.shape)
    g = (ctypes.c_int * rank)(*guard)
    r = (ctypes.c_int * rank)(*ref)
    st = ctypes.c_void_p(torch.cuda.current_stream().cuda_stream)
    if fn == "noalloc":
        rc = lib.zkx_ca_cfar_f32_noalloc(_ptr(x), rank, shape, g, r, pfa, _ptr(thr), _ptr(det), st)
    else:
        ws = torch.empty(x.numel(), dtype=torch.float32, device="cuda") if with_ws else None
        rc = lib.zkx_ca_cfar_f32(_ptr(x), rank, shape, g, r, pfa, _ptr(thr), _ptr(det),
                                 None if ws is None else _ptr(ws), st)
    torch.cuda.synchronize()
    assert rc == 0, lib.zkx_last_error_message()
    return thr, det


def _power(shape, seed):
    rng = np.random.default_rng(seed)
    p = rng.exponential(10.0, size=shape).astype(np.float32)
    flat = p.reshape(-1)
    flat[rng.integers(0, flat.size, 16)] *= 1e4  # strong cells
    return torch.from_numpy(p).cuda()


def test_two_calls_second_smaller(lib):
    cases = [((512, 8192), (2, 2), (8, 8), 1e-6, 1),   # 4M cells: the whole scratch, width % 4 == 0
             ((128, 1000), (1, 3), (4, 6), 1e-4, 2)]   # smaller, width % 4 != 0 (band kernel)
    assert lib.zkx_ca_cfar_noalloc_max_cells() == 512 * 8192
    for shape, guard, ref, pfa, seed in cases:
        print(f"noalloc 2-D {shape} guard {guard} ref {ref}", flush=True)
        x = _power(shape, seed)
        t0, d0 = _cfar(lib, "plain", x, guard, ref, pfa, with_ws=True)
        t1, d1 = _cfar(lib, "noalloc", x, guard, ref, pfa)
        assert torch.equal(t0.view(torch.int32), t1.view(torch.int32))
        assert torch.equal(d0, d1)


def test_one_dimensional_after_two_dimensional(lib):
    for n, seed in ((100000, 3), (4097, 4)):
        print(f"noalloc 1-D n={n}", flush=True)
        x = _power((n,), seed)
        t0, d0 = _cfar(lib, "plain", x, (3,), (16,), 1e-3)
        t1, d1 = _cfar(lib, "noalloc", x, (3,), (16,), 1e-3)
        assert torch.equal(t0.view(torch.int32), t1.view(torch.int32))
        assert torch.equal(d0, d1)


def test_oversized_map_is_refused(lib):
    x = torch.zeros((4096, 1025), dtype=torch.float32, device="cuda")
    thr = torch.empty_like(x)
    det = torch.empty(x.shape, dtype=torch.uint8, device="cuda")
    shape = (ctypes.c_int * 2)(4096, 1025)
    g = (ctypes.c_int * 2)(2, 2)
    r = (ctypes.c_int * 2)(8, 8)
    rc = lib.zkx_ca_cfar_f32_noalloc(_ptr(x), 2, shape, g, r, 1e-6, _ptr(thr), _ptr(det), None)
    assert rc == 1002  # INVALID_SHAPE
    assert b"scratch" in lib.zkx_last_error_message()
