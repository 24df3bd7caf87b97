"""GPU parity of upfirdn / resample_poly / firfilter against the CPU oracle
(faithful FP32 path and FP64), real and complex64 lines, any axis."""
import numpy as np
import pytest

torch = pytest.importorskip("torch")

from oracle import oracle as orc  # noqa: E402

pytestmark = pytest.mark.gpu
TOL = 1e-5


@pytest.fixture(scope="module")
def F():
    from zkx_dev_amd import filtering
    assert torch.cuda.is_available()
    return filtering


def dev(a):
    return torch.from_numpy(np.ascontiguousarray(a)).cuda()


def per_component(fn, x):
    """The complex extension applies the real operator to re and im."""
    return fn(x.real) + 1j * fn(x.imag)


@pytest.mark.parametrize("up,down,K,shape,axis", [
    (1, 4, 255, (3, 5000), 1), (3, 2, 121, (2, 3000), -1), (2, 1, 5, (100,), 0),
    (2, 3, 9, (100, 100), 0), (3, 2, 7, (32, 32), 1), (1, 1, 33, (4, 777), 1),
    (4, 4, 13, (3, 513), 1), (5, 3, 40, (257,), 0), (7, 1, 3, (50,), 0), (1, 9, 64, (2, 999), 1)])
@pytest.mark.parametrize("cplx", [False, True])
def test_upfirdn(F, up, down, K, shape, axis, cplx):
    rng = np.random.default_rng(K * 31 + up * 7 + down)
    h = rng.standard_normal(K).astype(np.float32)
    x = rng.standard_normal(shape).astype(np.float32)
    if cplx:
        x = (x + 1j * rng.standard_normal(shape)).astype(np.complex64)
    y = F.upfirdn(dev(h), dev(x), up, down, axis).cpu().numpy()
    f = lambda v: orc.upfirdn(h, v, up, down, axis)  # noqa: E731
    ref = per_component(f, x) if cplx else f(x)
    assert y.shape == ref.shape
    assert orc.rel_l2(y, ref) < TOL
    ref64 = per_component(lambda v: orc.upfirdn(h, v, up, down, axis, fp64_acc=True), x) if cplx \
        else orc.upfirdn(h, x, up, down, axis, fp64_acc=True)
    assert orc.rel_l2(y, ref64) < TOL


def test_upfirdn_probe_cases(F):
    """The reference probe's deterministic inputs (seeded integer patterns)."""
    for rank, rows, cols, taps, axis, up, down, seed in [(1, 100, 1, 5, 0, 2, 1, 84410001),
                                                         (2, 100, 100, 9, 0, 2, 3, 84410003),
                                                         (2, 32, 32, 7, 1, 3, 2, 84410002)]:
        h = orc.upfirdn_probe_input(taps, seed + 1, "FP32")
        x = orc.upfirdn_probe_input(rows * (cols if rank == 2 else 1), seed, "FP32")
        shape = (rows,) if rank == 1 else (rows, cols)
        y = F.upfirdn(dev(h), dev(x.reshape(shape)), up, down, axis).cpu().numpy()
        ref = orc.upfirdn(h, x.reshape(shape), up, down, axis).astype(np.float32)
        assert np.array_equal(y, ref)  # small integer multiples of 1/4: exact


@pytest.mark.parametrize("up,down,K,shape,axis", [
    (3, 2, 121, (2, 4000), 1), (2, 3, 5, (100, 100), 0), (2, 1, 21, (10, 10), 1),
    (160, 147, 301, (3000,), 0), (1, 4, 41, (5, 2048), 1), (5, 2, 51, (2, 1111), -1)])
@pytest.mark.parametrize("cplx", [False, True])
def test_resample_poly_custom(F, up, down, K, shape, axis, cplx):
    rng = np.random.default_rng(up * 100 + down + K)
    taps = rng.standard_normal(K).astype(np.float32)
    x = rng.standard_normal(shape).astype(np.float32)
    if cplx:
        x = (x + 1j * rng.standard_normal(shape)).astype(np.complex64)
    y = F.resample_poly(dev(x), up, down, axis, coefficients=taps).cpu().numpy()
    f = lambda v: orc.resample_poly_f32(v, up, down, axis, taps=taps)[0]  # noqa: E731
    ref = per_component(f, x) if cplx else f(x)
    assert y.shape == ref.shape
    assert orc.rel_l2(y, ref) < TOL
    # device-resident taps are scaled on the device to the same float32 taps
    y_dev = F.resample_poly(dev(x), up, down, axis, coefficients=dev(taps)).cpu().numpy()
    assert np.array_equal(y_dev, y)


@pytest.mark.parametrize("up,down", [(2, 1), (3, 2), (1, 3), (4, 6)])
def test_resample_poly_default_design(F, up, down):
    rng = np.random.default_rng(up + 10 * down)
    x = rng.standard_normal((3, 500)).astype(np.float32)
    y = F.resample_poly(dev(x), up, down, axis=1)
    assert y.dtype == torch.float64  # FP64 storage, as the reference's default design
    ref, dt = orc.resample_poly_f32(x, up, down, axis=1)
    assert dt == "FP64"
    assert orc.rel_l2(y.cpu().numpy(), ref) < TOL
    np.testing.assert_allclose(F.design_resample_poly_coefficients(up, down),
                               orc.resample_design(up, down), rtol=1e-14, atol=1e-15)


def test_resample_poly_identity(F):
    x = dev(np.arange(12, dtype=np.float32).reshape(3, 4))
    assert torch.equal(F.resample_poly(x, 3, 3, axis=1), x)


@pytest.mark.parametrize("K,shape,axis", [(9, (10000,), 0), (255, (3, 4096), 1), (31, (64, 17), 0),
                                          (1, (100,), 0), (128, (2, 300), -1)])
def test_firfilter(F, K, shape, axis):
    rng = np.random.default_rng(K)
    b = rng.standard_normal(K).astype(np.float32)
    x = rng.uniform(-1, 1, shape).astype(np.float32)
    y = F.firfilter(dev(b), dev(x), axis).cpu().numpy()
    ref, _ = orc.firfilter(b, x, axis)
    assert orc.rel_l2(y, ref) < TOL


@pytest.mark.parametrize("K,shape,axis,zshape", [(9, (200,), 0, (8,)), (5, (4, 50), 1, (4, 4)),
                                                 (5, (4, 50), 1, (1, 4)), (7, (30, 3), 0, (6, 1))])
def test_firfilter_zi_zf(F, K, shape, axis, zshape):
    rng = np.random.default_rng(K + len(shape))
    b = rng.standard_normal(K).astype(np.float32)
    x = rng.standard_normal(shape).astype(np.float32)
    zi = rng.standard_normal(zshape).astype(np.float32)
    y, zf = F.firfilter(dev(b), dev(x), axis, zi=dev(zi))
    ref_y, ref_zf = orc.firfilter(b, x, axis, zi=zi)
    assert orc.rel_l2(y.cpu().numpy(), ref_y) < TOL
    assert orc.rel_l2(zf.cpu().numpy(), ref_zf) < TOL


def test_firfilter_probe_case(F):
    x = orc.firfilter_probe_signal(10000, 29, "FP32")
    b = np.array([4 if i == 4 else (1 if i % 2 else 0) for i in range(9)], np.float32)
    y = F.firfilter(dev(b), dev(x)).cpu().numpy()
    ref, _ = orc.firfilter(b, x)
    assert orc.rel_l2(y, ref) < 1e-6


def test_filter_validation(F):
    x = dev(np.ones(10, np.float32))
    with pytest.raises(ValueError):
        F.upfirdn(dev(np.zeros(0, np.float32)), x)
    with pytest.raises(ValueError):
        F.upfirdn(dev(np.ones(3, np.float32)), x, 0, 1)
    with pytest.raises(ValueError):
        F.upfirdn(dev(np.ones(3, np.float32)), x, 1, 1, axis=2)
    with pytest.raises(ValueError):
        F.resample_poly(x, 0, 2)
    with pytest.raises(ValueError):
        F.firfilter(dev(np.ones(3, np.float32)), x, axis=3)


@pytest.mark.parametrize("n1,n2,batch", [(1000, 64, 3), (50, 50, 1), (64, 1000, 2), (20000, 255, 2),
                                         (3, 7, 4), (100000, 4096, 1), (8193, 1, 1)])
@pytest.mark.parametrize("mode", ["full", "same", "valid"])
def test_fftconvolve(F, n1, n2, batch, mode):
    """Against scipy's FP64 fftconvolve (the library cuSignal's tests pin to)
    and a direct numpy convolution."""
    from scipy import signal
    rng = np.random.default_rng(n1 + n2)
    a = (rng.standard_normal((batch, n1)) + 1j * rng.standard_normal((batch, n1))).astype(np.complex64)
    b = (rng.standard_normal(n2) + 1j * rng.standard_normal(n2)).astype(np.complex64)
    y = F.fftconvolve(dev(a), dev(b), mode).cpu().numpy()
    ref = np.stack([signal.fftconvolve(a[i].astype(np.complex128), b.astype(np.complex128), mode)
                    for i in range(batch)])
    assert y.shape == ref.shape
    assert orc.rel_l2(y, ref) < TOL
    if n1 * n2 <= 1_000_000:
        direct = np.stack([np.convolve(a[i].astype(np.complex128), b.astype(np.complex128), mode)
                           for i in range(batch)])
        if mode == "full" or n1 >= n2:  # numpy's same/valid follow the longer input
            assert orc.rel_l2(y, direct) < TOL


@pytest.mark.parametrize("kind,K,down,shape,cplx", [
    ("upfirdn", 255, 4, (4, 60000), True), ("upfirdn", 200, 3, (2, 30001), True),
    ("upfirdn", 100, 1, (3, 20000), False), ("upfirdn", 255, 4, (2, 50000), False),
    ("upfirdn", 129, 2, (2, 40001), True), ("upfirdn", 129, 2, (3, 9000), False),
    ("upfirdn", 400, 8, (2, 70001), True), ("upfirdn", 400, 8, (1, 33000), False),
    ("upfirdn", 513, 8, (1, 40000), True), ("upfirdn", 1000, 2, (1, 50000), True),
    ("firfilter", 255, 1, (2, 100000), False), ("firfilter", 513, 1, (1, 4096), False),
    ("firfilter", 1000, 1, (1, 70000), False)])
def test_frequency_domain_fir_matches_direct_and_oracle(F, kind, K, down, shape, cplx):
    """Long up = 1 filters on contiguous lines run as frequency-domain
    overlap-save, in the phase form (variant 1: D phase transforms of N
    points per segment) and the segment form (variant 2, the default for
    D in {1, 2, 4, 8} with history up to 1/8 of a 4096-point segment: one
    4096-point transform and an inverse of 4096/D points on the folded
    spectrum; other shapes fall back to the phase form); both agree with
    the direct polyphase kernels (variant 0) and the oracle."""
    from zkx_dev_amd._lib import lib
    rng = np.random.default_rng(K + down)
    h = (rng.standard_normal(K) / np.sqrt(K)).astype(np.float32)
    x = rng.uniform(-1, 1, shape).astype(np.float32)
    if cplx:
        x = (x + 1j * rng.uniform(-1, 1, shape)).astype(np.complex64)

    def run():
        if kind == "firfilter":
            return F.firfilter(dev(h), dev(x), 1).cpu().numpy()
        return F.upfirdn(dev(h), dev(x), 1, down, 1).cpu().numpy()
    assert lib().zkx_get_kernel_variant(b"fir_fft") == 2
    y_seg = run()
    try:
        assert lib().zkx_set_kernel_variant(b"fir_fft", 1) == 0
        y_ph = run()
        assert lib().zkx_set_kernel_variant(b"fir_fft", 0) == 0
        y_dir = run()
    finally:
        lib().zkx_set_kernel_variant(b"fir_fft", 2)
    if kind == "firfilter":
        ref, _ = orc.firfilter(h, x, 1)
    else:
        f = lambda v: orc.upfirdn(h, v, 1, down, 1, fp64_acc=True)  # noqa: E731
        ref = per_component(f, x) if cplx else f(x)
    for y_fft in (y_seg, y_ph):
        assert y_fft.shape == y_dir.shape
        assert orc.rel_l2(y_fft, y_dir) < 2e-6
        assert orc.rel_l2(y_fft, ref) < TOL


@pytest.mark.parametrize("api,up,down,K,shape", [
    ("resample_poly", 3, 2, 121, (2, 50001)), ("resample_poly", 4, 2, 200, (1, 40000)),
    ("resample_poly", 2, 3, 300, (1, 30000)), ("upfirdn", 2, 1, 129, (2, 30000)),
    ("upfirdn", 4, 1, 255, (1, 20000)), ("upfirdn", 3, 4, 240, (1, 60000)),
    ("upfirdn", 3, 2, 255, (3, 12000)), ("resample_poly", 3, 2, 301, (1, 30000))])
@pytest.mark.parametrize("cplx", [False, True])
def test_segment_form_rate_changing_fir(F, api, up, down, K, shape, cplx):
    """Rate-changing filters on long contiguous lines (up/g in {2, 3, 4},
    down/g in {1, 2, 4}, phases of 64 taps and more) run as 4096-point
    overlap-save segments whose forward transform is shared by the output
    residues (variant 2, the default); other factors (2/3 here) and shorter
    phases (3/2 with 121 taps) keep the direct kernels.  Both agree with the direct polyphase kernels (variant 0)
    and the oracle."""
    from zkx_dev_amd._lib import lib
    rng = np.random.default_rng(up * 1000 + down * 10 + K)
    taps = (rng.standard_normal(K) / np.sqrt(K)).astype(np.float32)
    x = rng.uniform(-1, 1, shape).astype(np.float32)
    if cplx:
        x = (x + 1j * rng.uniform(-1, 1, shape)).astype(np.complex64)

    def run():
        if api == "resample_poly":
            return F.resample_poly(dev(x), up, down, 1, coefficients=taps).cpu().numpy()
        return F.upfirdn(dev(taps), dev(x), up, down, 1).cpu().numpy()
    assert lib().zkx_get_kernel_variant(b"fir_fft") == 2
    y_seg = run()
    try:
        assert lib().zkx_set_kernel_variant(b"fir_fft", 0) == 0
        y_dir = run()
    finally:
        lib().zkx_set_kernel_variant(b"fir_fft", 2)
    if api == "resample_poly":
        f = lambda v: orc.resample_poly_f32(v, up, down, 1, taps=taps)[0]  # noqa: E731
    else:
        f = lambda v: orc.upfirdn(taps, v, up, down, 1, fp64_acc=True)  # noqa: E731
    ref = per_component(f, x) if cplx else f(x)
    assert y_seg.shape == y_dir.shape == ref.shape
    assert orc.rel_l2(y_seg, y_dir) < 2e-6
    assert orc.rel_l2(y_seg, ref) < TOL


def test_frequency_domain_firfilter_zi_zf(F):
    rng = np.random.default_rng(5)
    K, n = 255, 40000
    b = (rng.standard_normal(K) / 16).astype(np.float32)
    x = rng.uniform(-1, 1, (2, n)).astype(np.float32)
    zi = rng.standard_normal((2, K - 1)).astype(np.float32)
    y, zf = F.firfilter(dev(b), dev(x), 1, zi=dev(zi))
    ref_y, ref_zf = orc.firfilter(b, x, 1, zi=zi)
    assert orc.rel_l2(y.cpu().numpy(), ref_y) < TOL
    assert orc.rel_l2(zf.cpu().numpy(), ref_zf) < TOL


@pytest.mark.parametrize("shape,axis,q,order,zero_phase", [
    ((5000,), -1, 4, None, True), ((5000,), -1, 4, None, False), ((3, 2001), 1, 3, None, True),
    ((777, 6), 0, 5, None, True), ((4, 333, 3), 1, 2, 30, False), ((2, 3, 4, 257), -1, 8, None, True),
    ((32, 32), 1, 4, 32, True), ((1 << 18,), 0, 13, None, True), ((10,), 0, 16, None, True)])
def test_decimate_default_design(F, shape, axis, q, order, zero_phase):
    """Designed filter (FP64 storage), any rank and axis, vs the FP64 oracle."""
    x = np.random.default_rng(q + len(shape)).standard_normal(shape).astype(np.float32)
    print(f"decimate {shape} axis={axis} q={q} order={order} zero_phase={zero_phase}", flush=True)
    y = F.decimate(dev(x), q, order, axis=axis, zero_phase=zero_phase).cpu().numpy()
    ref = orc.decimate(x, q, axis=axis, filter_order=order, zero_phase=zero_phase)
    assert y.dtype == np.float64 and y.shape == ref.shape
    assert orc.rel_l2(y, ref) < TOL


@pytest.mark.parametrize("n,q,zero_phase", [(100, 1, True), (100, 2, False), (1000, 5, True),
                                            (4099, 7, False), (3, 5, True), (1, 2, True)])
def test_decimate_custom_taps(F, n, q, zero_phase):
    """Custom coefficients: float32 result; q = 1 still filters."""
    x = np.random.default_rng(n).standard_normal((2, n)).astype(np.float32)
    h = np.array([1, 2, 1, -1, 1, 0.5, -0.25], np.float32)
    y = F.decimate(dev(x), q, dev(h), axis=1, zero_phase=zero_phase).cpu().numpy()
    ref = orc.decimate(x, q, axis=1, taps=h.astype(np.float64), zero_phase=zero_phase, fp32=False)
    assert y.dtype == np.float32 and y.shape == ref.shape
    assert orc.rel_l2(y, ref) < TOL


def test_decimate_probe_cases(F):
    """The reference probe's inputs and filters (32x32 q=4 order 32 zero phase;
    100x100 q=5 custom; 10x10 q=2 custom causal): the device result is within
    1e-6 of the oracle, which reproduces the reference's digests."""
    h = np.array([1, 2, 1, -1, 1], np.float32)
    for rows, cols, q, order, zp, custom, seed in ((32, 32, 4, 32, True, False, 83410002),
                                                   (100, 100, 5, None, True, True, 83410003),
                                                   (10, 10, 2, None, False, True, 83410001)):
        x = orc.decimate_probe_input(rows * cols, seed, "FP32").reshape(rows, cols)
        y = F.decimate(dev(x), q, dev(h) if custom else order, axis=1, zero_phase=zp).cpu().numpy()
        ref = orc.decimate(x.astype(np.float64), q, axis=1, filter_order=order, zero_phase=zp,
                           taps=h.astype(np.float64) if custom else None, fp32=custom)
        assert orc.rel_l2(y, ref) < 1e-6, (rows, cols, q)


def test_decimate_validation(F):
    x = dev(np.ones(16, np.float32))
    with pytest.raises(ValueError):
        F.decimate(x, 0)
    with pytest.raises(ValueError):
        F.decimate(x, 1)  # the default design's cutoff 1 is the Nyquist endpoint
    with pytest.raises(ValueError):
        F.decimate(x, 2, -1)


@pytest.mark.parametrize("n,nh,M", [(128, 32, 8), (1024, 64, 8), (16384, 256, 16), (1000, 3, 1), (999, 20, 2),
                                    (4096, 128, 4), (4096, 1024, 32), (3000, 60, 6), (10000, 384, 12),
                                    (65536, 1000, 100), (1 << 20, 8192, 1024), (5000, 0, 16), (7, 16, 8),
                                    (65536, 2048, 64), (70021, 200, 24), (40000, 1280, 128), (100000, 8192, 256),
                                    (3000, 100, 32)])
def test_channelize_poly(F, n, nh, M):
    """Fused forms (M in 1..16 with the DFT in registers; M in 32..256 with
    the FFT in LDS; partial last blocks) and the FIR + row FFT + transpose
    form (others) vs the FP64 oracle."""
    rng = np.random.default_rng(n + nh + M)
    x = rng.standard_normal(n).astype(np.float32)
    h = rng.standard_normal(nh).astype(np.float32)
    print(f"channelize n={n} taps={nh} M={M}", flush=True)
    y = F.channelize_poly(dev(x), dev(h), M).cpu().numpy()
    ref = orc.channelize_poly(x, h, M)
    assert y.shape == ref.shape
    if ref.size:
        assert orc.rel_l2(y, ref) < TOL


def test_channelize_poly_probe_cases(F):
    for n, nh, M, seed in ((128, 32, 8, 83010001), (1024, 64, 8, 83010002), (16384, 256, 16, 83010003)):
        x = orc.channelize_probe_input(n, seed, False, "FP32")
        h = orc.channelize_probe_input(nh, seed + 1, True, "FP32")
        y = F.channelize_poly(dev(x), dev(h), M).cpu().numpy()
        assert orc.rel_l2(y, orc.channelize_poly(x, h, M)) < 1e-6


def test_channelize_poly_validation(F):
    x = dev(np.ones(64, np.float32))
    with pytest.raises(ValueError):
        F.channelize_poly(x, dev(np.ones(8, np.float32)), 0)
    with pytest.raises(ValueError):
        F.channelize_poly(x, dev(np.ones(66, np.float32)), 2)  # 33 taps per phase
