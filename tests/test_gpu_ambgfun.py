"""ambgfun on the HIP library against the FP64 oracle.

The oracle's direct sums (radartools_typed.cpp:329-409) are pinned to the
reference's CPU digests and to its archived Task1 GPU ambiguity maps; here the
device results (FP32 arithmetic, as in the reference) must agree with them to
relative L2 < 1e-5 (the reference's own bounds are 2e-4 for the 2-D map and
1e-4 for the cuts).  Large maps are checked row by row against the FFT
formulation (oracle.ambgfun_fft64_row), which is pinned to the direct sums."""
import numpy as np
import pytest

torch = pytest.importorskip("torch")

from oracle import oracle as orc  # noqa: E402

pytestmark = pytest.mark.gpu


def _rand(n, seed, scale=1.0):
    rng = np.random.default_rng(seed)
    return ((rng.standard_normal(n) + 1j * rng.standard_normal(n)) * scale).astype(np.complex64)


def _dev(a):
    return torch.from_numpy(np.ascontiguousarray(a)).cuda()


@pytest.mark.parametrize("nx,ny,fs", [(1, None, 1.0), (2, None, 1.0), (5, None, 10.0), (16, None, 16.0),
                                      (50, None, 1000.0), (100, None, 2000.0), (128, None, 2.56e6),
                                      (50, 50, 1000.0), (37, 20, 1000.0), (12, 40, 100.0), (120, 90, 1.0)])
def test_2d_vs_oracle(nx, ny, fs):
    """cuSignal's test shapes (test_radartools_cpu.py:196-243), unequal
    lengths, and the degenerate 1- and 2-sample cases (nfreq 1, 4)."""
    from zkx_dev_amd import radar
    print(f"ambgfun 2d nx={nx} ny={ny}", flush=True)
    x = _rand(nx, nx)
    y = None if ny is None else _rand(ny, 7 + ny, 3.0)
    a = radar.ambgfun(_dev(x), fs, 1.0, y=None if y is None else _dev(y)).cpu().numpy()
    ref = orc.ambgfun(x, fs, y=y, output_fp32=False)
    assert a.dtype == np.float32 and a.shape == ref.shape
    assert orc.rel_l2(a, ref) < 1e-5
    a64 = radar.ambgfun(_dev(x), fs, 1.0, y=None if y is None else _dev(y), out_fp64=True).cpu().numpy()
    assert a64.dtype == np.float64 and np.array_equal(a64, a.astype(np.float64))


@pytest.mark.parametrize("nx,fs,tau", [(1, 1.0, 0.0), (50, 1000.0, 0.0), (100, 2000.0, 5.0), (64, 1000.0, 0.0123),
                                       (128, 2.56e6, 3.3e-5), (1000, 1e6, -2.5e-4), (4096, 1.0, 17.25)])
def test_delay_cut_vs_oracle(nx, fs, tau):
    from zkx_dev_amd import radar
    print(f"ambgfun delay nx={nx} tau={tau}", flush=True)
    x = _rand(nx, 100 + nx)
    a = radar.ambgfun(_dev(x), fs, 1.0, cut="delay", cutValue=tau).cpu().numpy()
    ref = orc.ambgfun_fft64(x.astype(np.complex128), fs, cut="delay", cut_value=tau) if nx > 200 else \
        orc.ambgfun(x, fs, cut="delay", cut_value=tau)
    assert a.dtype == np.float64 and a.shape == ref.shape
    assert orc.rel_l2(a, ref) < 1e-5


@pytest.mark.parametrize("nx,ny,fs,nu", [(1, None, 1.0, 0.0), (50, None, 1000.0, 0.0), (100, None, 2000.0, 10.0),
                                         (33, None, 500.0, 71.5), (20, 45, 100.0, 3.0), (4096, None, 1e6, 1234.5),
                                         (6000, None, 1.0, 0.01)])
def test_doppler_cut_vs_oracle(nx, ny, fs, nu):
    """Includes lengths whose 2-D map would exceed the LDS transform (the
    Doppler cut has no length limit)."""
    from zkx_dev_amd import radar
    print(f"ambgfun doppler nx={nx} ny={ny} nu={nu}", flush=True)
    x = _rand(nx, 200 + nx)
    y = None if ny is None else _rand(ny, 300 + ny)
    a = radar.ambgfun(_dev(x), fs, 1.0, y=None if y is None else _dev(y), cut="doppler", cutValue=nu).cpu().numpy()
    if nx > 200:
        ref = orc.ambgfun_fft64(x.astype(np.complex128), fs, cut="doppler", cut_value=nu)
    else:
        ref = orc.ambgfun(x, fs, y=y, cut="doppler", cut_value=nu)
    assert a.dtype == np.float64 and a.shape == ref.shape
    assert orc.rel_l2(a, ref) < 1e-5


def test_2d_largest_map_rows():
    """nx = ny = 4096: rows 8191, nfreq 8192 (the largest LDS transform).
    Sampled rows against the FFT formulation; peak and bound everywhere."""
    from zkx_dev_amd import radar
    print("ambgfun 2d nx=4096", flush=True)
    x = _rand(4096, 4096)
    a = radar.ambgfun(_dev(x), 1.0, 1.0)
    assert tuple(a.shape) == (8191, 8192)
    xn = x.astype(np.complex128) / np.linalg.norm(x.astype(np.complex128))
    rng = np.random.default_rng(1)
    for row in [0, 1, 4095, 8190] + rng.integers(0, 8191, 6).tolist():
        ref = orc.ambgfun_fft64_row(xn, xn, int(row), 8192)
        got = a[int(row)].cpu().numpy()
        assert orc.rel_l2(got, ref) < 1e-5, row
    assert float(a.max()) <= 1.0 + 1e-5
    assert abs(float(a[4095, 4096]) - 1.0) < 1e-5


def test_high_rank_2d_uses_first_extent():
    """ambgfun_first_extent: a [8, 3] input takes its first 8 samples, the norm
    covers all 24 (radartools_typed.h:17-20)."""
    from zkx_dev_amd import radar
    x = _rand(24, 9)
    a = radar.ambgfun(_dev(x.reshape(8, 3)), 1.0, 1.0).cpu().numpy()
    ref = orc.ambgfun(x, nx=8, output_fp32=False)
    assert a.shape == ref.shape == (15, 16)
    assert orc.rel_l2(a, ref) < 1e-5


def test_errors_and_nan():
    from zkx_dev_amd import radar
    x = _dev(_rand(8, 1))
    with pytest.raises(ValueError):
        radar.ambgfun(x, 0.0, 1.0, cut="delay")
    with pytest.raises(ValueError):
        radar.ambgfun(_dev(_rand(8, 1).reshape(4, 2)), 1.0, 1.0, cut="doppler")
    with pytest.raises(ValueError):
        radar.ambgfun(x, 1.0, 1.0, cut="range")
    z = radar.ambgfun(_dev(np.zeros(6, np.complex64)), 1.0, 1.0).cpu().numpy()
    assert np.all(np.isnan(z))


@pytest.mark.parametrize("out_fp64", [False, True])
def test_2d_beyond_one_lds_pass(out_fp64):
    """nx = 6000, ny = 3000: rows 8999, nfreq 16384, through the global FFT
    in row chunks; sampled rows (both ends, the zero-delay row, chunk
    seams) against the FP64 FFT formulation."""
    from zkx_dev_amd import radar
    print(f"ambgfun 2d nx=6000 ny=3000 fp64={out_fp64}", flush=True)
    x, y = _rand(6000, 61), _rand(3000, 62, 0.5)
    a = radar.ambgfun(_dev(x), 1.0, 1.0, y=_dev(y), out_fp64=out_fp64)
    assert tuple(a.shape) == (8999, 16384)
    assert a.dtype == (torch.float64 if out_fp64 else torch.float32)
    xn = x.astype(np.complex128) / np.linalg.norm(x.astype(np.complex128))
    yn = y.astype(np.complex128) / np.linalg.norm(y.astype(np.complex128))
    chunk = (256 << 20) // (8 * 16384)
    rng = np.random.default_rng(2)
    for row in [0, 1, 5999, 8998, chunk - 1, chunk, 2 * chunk] + rng.integers(0, 8999, 5).tolist():
        ref = orc.ambgfun_fft64_row(xn, yn, int(row), 16384)
        assert orc.rel_l2(a[int(row)].cpu().numpy(), ref) < 1e-5, row
    assert float(a.max()) <= 1.0 + 1e-5


def test_delay_cut_beyond_one_lds_pass():
    """nx = ny = 5000 (nfreq 16384): the three whole-length transforms."""
    from zkx_dev_amd import radar
    x = _rand(5000, 71)
    for tau in (0.0, 3.1e-4, -1.7e-3):
        got = radar.ambgfun(_dev(x), 2.0e5, 1.0, cut="delay", cutValue=tau).cpu().numpy()
        ref = orc.ambgfun_fft64(x, 2.0e5, cut="delay", cut_value=tau)
        assert got.shape == ref.shape == (16384,)
        assert orc.rel_l2(got, ref) < 1e-5, tau
