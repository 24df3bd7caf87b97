"""GPU parity of the radar operators against the CPU oracle.

Bars: complex outputs within 1e-5 relative L2 of the FP64 oracle; the
integral-image CA-CFAR bit-identical to the reference restatement (thresholds
and masks); the windowed FP64 CA-CFAR's masks identical to the FP64 oracle's
except cells whose statistic lies within 1e-6 relative of the threshold
(counted, and required to be few).
"""
import numpy as np
import pytest

torch = pytest.importorskip("torch")

from oracle import oracle as orc  # noqa: E402

pytestmark = pytest.mark.gpu

TOL = 1e-5


@pytest.fixture(scope="module")
def R():
    from zkx_dev_amd import radar
    assert torch.cuda.is_available()
    return radar


def dev(a, dtype=None):
    t = torch.from_numpy(np.ascontiguousarray(a))
    if dtype is not None:
        t = t.to(dtype)
    return t.cuda()


def rng_c64(rng, shape, scale=1.0):
    return (scale * (rng.standard_normal(shape) + 1j * rng.standard_normal(shape))).astype(np.complex64)


@pytest.mark.parametrize("P,S,T,nfft", [(8, 256, 64, -1), (4, 256, 64, 512), (3, 300, 40, 256),
                                        (2, 4096, 256, -1), (5, 64, 100, 128), (2, 16, 5, 16),
                                        # other lengths: overlap-save segments over the period nfft
                                        (4, 300, 40, -1), (3, 100, 120, 77), (2, 1000, 64, 1500), (5, 50, 8, 1),
                                        (2, 20000, 256, -1), (3, 5, 3, 9)])
@pytest.mark.parametrize("window", ["none", "hann", "hamming"])
@pytest.mark.parametrize("normalize", [False, True])
def test_pulse_compression_vs_fp64(R, P, S, T, nfft, window, normalize):
    rng = np.random.default_rng(P * 1000 + S + T)
    x = rng_c64(rng, (P, S))
    h = rng_c64(rng, (T,))
    y = R.pulse_compression(dev(x), dev(h), nfft=nfft, window=window, normalize=normalize,
                            widen=False).cpu().numpy()
    ref = orc.pulse_compression(x, h, nfft=nfft, window=window, normalize=normalize, exact=True)
    assert y.shape == ref.shape
    assert orc.rel_l2(y, ref) < TOL


def test_pulse_compression_probe_inputs(R):
    """The reference probe's deterministic inputs (seeds 3/7, echo [64,256],
    template 64) against the faithful oracle."""
    x = orc.probe_complex_input(64 * 256, 3, "FP32").astype(np.float32)
    t = orc.probe_complex_input(64, 7, "FP32").astype(np.float32)
    xc = (x[0::2] + 1j * x[1::2]).astype(np.complex64).reshape(64, 256)
    tc = (t[0::2] + 1j * t[1::2]).astype(np.complex64)
    y = R.pulse_compression(dev(xc), dev(tc), nfft=256).cpu().numpy()
    ref = orc.pulse_compression(xc, tc, nfft=256)
    assert orc.rel_l2(y, ref) < TOL


def test_pulse_compression_widening(R):
    rng = np.random.default_rng(5)
    x, h = rng_c64(rng, (4, 128)), rng_c64(rng, (16,))
    y = R.pulse_compression(dev(x), dev(h), window="hann")
    assert y.dtype == torch.complex128
    y32 = R.pulse_compression(dev(x), dev(h), window="hann", widen=False)
    assert torch.equal(y.to(torch.complex64), y32)


def test_pulse_compression_delta_template_is_shift(R):
    """A unit delta at tap 0 returns the pulse itself (circular, nfft = S)."""
    rng = np.random.default_rng(9)
    x = rng_c64(rng, (4, 512))
    h = np.zeros(8, np.complex64)
    h[0] = 1.0
    y = R.pulse_compression(dev(x), dev(h)).cpu().numpy()
    assert orc.rel_l2(y, x) < 1e-6


@pytest.mark.parametrize("P,Rg,nfft", [(16, 64, -1), (32, 100, -1), (128, 256, -1), (128, 96, 256),
                                       (64, 33, 32), (512, 48, -1), (8, 5, 8), (2, 7, 4),
                                       # other lengths: FP32 row transforms after a corner turn
                                       (100, 64, -1), (48, 30, 60), (130, 20, 100), (3, 9, 5), (1000, 8, -1),
                                       (40, 6, 1), (5000, 4, -1), (300, 5, 5003), (20000, 3, 16384),
                                       (64, 7, 16384)])
@pytest.mark.parametrize("window", ["none", "hann", "hamming"])
def test_pulse_doppler_vs_fp64(R, P, Rg, nfft, window):
    rng = np.random.default_rng(P * 7 + Rg)
    x = rng_c64(rng, (P, Rg))
    y = R.pulse_doppler(dev(x), nfft=nfft, window=window, widen=False).cpu().numpy()
    ref = orc.pulse_doppler_fft64(x, nfft=nfft, window=window)
    assert y.shape == ref.shape
    assert orc.rel_l2(y, ref) < TOL


def test_pulse_doppler_explicit_window_other_length(R):
    rng = np.random.default_rng(77)
    P, Rg, nfft = 40, 16, 50
    x = rng_c64(rng, (P, Rg))
    w = rng.uniform(0.2, 1.0, P).astype(np.float32)
    y = R.pulse_doppler(dev(x), nfft=nfft, window="explicit_fp32", explicit_window=dev(w), widen=False)
    ref = np.fft.fft(x.astype(np.complex128) * w.astype(np.float64)[:, None], nfft, axis=0)
    assert orc.rel_l2(y.cpu().numpy(), ref) < TOL


def test_pulse_doppler_power_and_tone(R):
    """A pure tone at Doppler bin f0 (no window) gives a Kronecker delta of
    height P; the fused |Y|^2 equals |Y|^2 of the complex output."""
    P, Rg, f0 = 128, 64, 37
    p = np.arange(P)[:, None]
    x = np.broadcast_to(np.exp(2j * np.pi * f0 * p / P), (P, Rg)).astype(np.complex64)
    xt = dev(x)
    out = torch.empty((P, Rg), dtype=torch.complex64, device="cuda")
    pw = torch.empty((P, Rg), dtype=torch.float32, device="cuda")
    R.pulse_doppler(xt, out=out, power=pw)
    y = out.cpu().numpy()
    expect = np.zeros((P, Rg), np.complex128)
    expect[f0] = P
    assert np.max(np.abs(y - expect)) < 1e-3
    assert torch.allclose(pw, (out.abs() ** 2), rtol=1e-5, atol=1e-3)


def test_pulse_doppler_probe_inputs_faithful(R):
    x = orc.probe_complex_input(128 * 128, 11, "FP32").astype(np.float32)
    xc = (x[0::2] + 1j * x[1::2]).astype(np.complex64).reshape(128, 128)
    y = R.pulse_doppler(dev(xc), nfft=128).cpu().numpy()
    ref = orc.pulse_doppler(xc, nfft=128)  # faithful: float pi
    exact = orc.pulse_doppler(xc, nfft=128, exact=True)
    assert orc.rel_l2(y, exact) < TOL
    assert orc.rel_l2(y, ref) < 2e-5  # the reference's float-pi gap (~7e-6 published)


def _cfar_inputs(kind, shape, seed):
    rng = np.random.default_rng(seed)
    if kind == "exp":
        return rng.exponential(10.0, size=shape).astype(np.float32)
    if kind == "probe":
        return orc.probe_real_input(int(np.prod(shape)), 23, "FP32").reshape(shape)
    if kind == "ones":
        return np.ones(shape, np.float32)
    raise AssertionError(kind)


@pytest.mark.parametrize("kind,shape,guard,ref,pfa", [
    ("probe", (16, 16), (1, 1), (4, 4), 1e-3),
    ("probe", (128, 128), (1, 1), (4, 4), 1e-3),
    ("exp", (128, 4096), (2, 2), (8, 8), 1e-6),
    ("exp", (64, 1000), (1, 2), (2, 6), 1e-3),
    ("exp", (33, 70), (0, 3), (5, 1), 0.5),
    ("ones", (40, 40), (2, 2), (3, 3), 1e-3),
    ("exp", (130, 516), (1, 2), (3, 5), 1e-4),
    ("exp", (7, 256), (1, 1), (1, 2), 1e-2),
    ("probe", (256, 192), (2, 1), (4, 3), 1e-3),
    ("exp", (65, 68), (1, 1), (2, 2), 1e-3),
])
def test_ca_cfar_bit_identical(R, kind, shape, guard, ref, pfa):
    x = _cfar_inputs(kind, shape, sum(shape))
    thr, det = R.ca_cfar(dev(x), guard, ref, pfa)
    othr, odet = orc.ca_cfar(x, guard, ref, pfa)
    assert np.array_equal(thr.cpu().numpy().view(np.uint32), othr.view(np.uint32))
    assert np.array_equal(det.cpu().numpy(), odet)


@pytest.mark.parametrize("n,guard,ref", [(1000, 2, 5), (64, 0, 1), (4097, 3, 16)])
def test_ca_cfar_rank1_bit_identical(R, n, guard, ref):
    x = _cfar_inputs("exp", (n,), n)
    thr, det = R.ca_cfar(dev(x), (guard,), (ref,), 1e-3)
    othr, odet = orc.ca_cfar(x, (guard,), (ref,), 1e-3)
    assert np.array_equal(thr.cpu().numpy().view(np.uint32), othr.view(np.uint32))
    assert np.array_equal(det.cpu().numpy(), odet)


def test_ca_cfar_known_answers(R):
    """cuSignal's closed-form cases: all-ones input gives T = alpha in the
    interior and 0 on the border; isolated targets are detected."""
    x = np.ones((32, 32), np.float32)
    thr, det = R.ca_cfar(dev(x), (1, 1), (2, 2), 1e-3)
    thr = thr.cpu().numpy()
    N = orc.ca_cfar_count((1, 1), (2, 2))
    alpha = orc.cfar_alpha(1e-3, N)
    assert np.allclose(thr[3:-3, 3:-3], alpha, rtol=1e-5)
    assert np.all(thr[:3] == 0) and np.all(thr[:, :3] == 0)
    y = np.random.default_rng(0).exponential(1.0, (64, 64)).astype(np.float32)
    y[20, 30] = 1e3
    y[40, 12] = 1e3
    _, det = R.ca_cfar(dev(y), (2, 2), (4, 4), 1e-6)
    d = det.cpu().numpy()
    assert d[20, 30] == 1 and d[40, 12] == 1


def test_ca_cfar_validation(R):
    with pytest.raises(ValueError):
        R.ca_cfar(dev(np.ones((4, 4, 4), np.float32)), (1, 1, 1), (1, 1, 1))
    with pytest.raises(ValueError):
        R.ca_cfar(dev(np.ones((32, 32), np.float32)), (1, 1), (2, 2), pfa=1.5)
    with pytest.raises(ValueError):
        R.ca_cfar(dev(np.ones((8, 8), np.float32)), (2, 2), (2, 2))


@pytest.mark.parametrize("shape,guard,ref,pfa", [((128, 4096), (2, 2), (8, 8), 1e-6),
                                                 ((100, 300), (1, 2), (3, 4), 1e-3)])
def test_ca_cfar_windowed_vs_fp64(R, shape, guard, ref, pfa):
    x = _cfar_inputs("exp", shape, 7)
    thr, det = R.ca_cfar_windowed(dev(x), guard, ref, pfa)
    thr, det = thr.cpu().numpy(), det.cpu().numpy()
    exact = orc.ca_cfar_exact(x, guard, ref, pfa)
    assert np.allclose(thr, exact, rtol=1e-6, atol=0)
    exact_det = (x.astype(np.float64) > exact).astype(np.uint8)
    near = np.abs(x - exact) <= 1e-6 * np.abs(exact)
    mism = det != exact_det
    assert not np.any(mism & ~near)
    assert mism.sum() <= near.sum()


def test_cfar_alpha(R):
    assert R.cfar_alpha(1e-6, 416) == pytest.approx(orc.cfar_alpha(1e-6, 416), rel=0, abs=0)
    assert R.cfar_alpha(1.0, 128) == 0.0
    with pytest.raises(ValueError):
        R.cfar_alpha(0.0, 10)
    with pytest.raises(ValueError):
        R.cfar_alpha(0.5, 0)


def test_output_tensors_validated_before_launch(R):
    """Every caller-supplied output is checked before any kernel launch: a
    host tensor, a short one and an expanded (non-contiguous) one each raise
    ValueError."""
    from zkx_dev_amd import convolution, fft
    rng = np.random.default_rng(5)
    x = dev(rng_c64(rng, (16, 64)))
    spec = R.pc_template_spectrum(dev(rng_c64(rng, (8,))), 64)
    pw = dev(np.abs(rng_c64(rng, (32, 64))).astype(np.float32))
    a = torch.ones(100, device="cuda")
    bad_c64 = [torch.empty((16, 64), dtype=torch.complex64), torch.empty(16 * 64 - 1, dtype=torch.complex64,
               device="cuda"), torch.empty(1, dtype=torch.complex64, device="cuda").expand(16 * 64)]
    bad_f32 = [torch.empty((32, 64)), torch.empty(32 * 64 - 1, device="cuda"),
               torch.empty(1, device="cuda").expand(32 * 64)]
    for out in bad_c64:
        with pytest.raises(ValueError):
            R.pulse_compression_prepared(x, spec, out=out)
        with pytest.raises(ValueError):
            R.pulse_doppler(x, out=out)
        with pytest.raises(ValueError):
            fft.Plan(64, batch=16).execute(x.reshape(-1), out=out)
    for out in bad_f32:
        with pytest.raises(ValueError):
            R.ca_cfar(pw, (1, 1), (2, 2), 1e-3, threshold=out)
        with pytest.raises(ValueError):
            R.ca_cfar_windowed(pw, (1, 1), (2, 2), 1e-3, threshold=out)
    for out in (torch.empty(100), torch.empty(99, device="cuda"), torch.empty(1, device="cuda").expand(100)):
        with pytest.raises(ValueError):
            convolution.correlate(a, a[:1], "same", out=out)
    with pytest.raises(ValueError):
        R.doppler_cfar(dev(rng_c64(rng, (32, 128))), (2, 2), (8, 8), 1e-6,
                       power=torch.empty(32 * 128 - 1, device="cuda"))


@pytest.mark.parametrize("nfft", [-1, 200])
def test_pulse_compression_template_longer_than_pulse(R, nfft):
    """The reference's operator accepts a template longer than the pulse and
    correlates with its first nfft taps after windowing and normalising the
    whole template (radartools_typed.cpp:618-640)."""
    rng = np.random.default_rng(5)
    P, S, T = 8, 256, 300
    x = (rng.standard_normal((P, S)) + 1j * rng.standard_normal((P, S))).astype(np.complex64)
    h = (rng.standard_normal(T) + 1j * rng.standard_normal(T)).astype(np.complex64)
    y = R.pulse_compression(dev(x), dev(h), nfft=nfft, window="hann", normalize=True).cpu().numpy()
    ref = orc.pulse_compression_fft64(x, h, nfft=nfft, window="hann", normalize=True)
    assert y.shape == ref.shape
    assert orc.rel_l2(y, ref) < TOL


def test_injected_fft_failure_is_backend_error_and_writes_nothing(R):
    """The host-forced transform failure returns BACKEND_ERROR before any
    device work: the preallocated outputs keep their contents."""
    import ctypes
    from zkx_dev_amd._lib import BACKEND_ERROR, ZkxError, check, lib
    L = lib()
    P, S = 16, 128
    x = torch.randn(P, S, dtype=torch.complex64, device="cuda")
    h = torch.randn(16, dtype=torch.complex64, device="cuda")
    out = torch.full((P, S), 3.5 + 0j, dtype=torch.complex64, device="cuda")
    power = torch.full((P, S), -2.0, dtype=torch.float32, device="cuda")
    ptr = lambda t: ctypes.c_void_p(t.data_ptr())  # noqa: E731
    check(L.zkx_set_kernel_variant(b"inject_fft_failure", 1))
    try:
        pc = L.zkx_pulse_compression_c64(ptr(x), P, S, ptr(h), 16, -1, 0, 1, None, ptr(out), None)
        pd = L.zkx_pulse_doppler_c64(ptr(x), P, S, -1, 1, None, ptr(out), ptr(power), None)
        with pytest.raises(ZkxError) as e:
            R.pulse_doppler(x, window="hann")
    finally:
        check(L.zkx_set_kernel_variant(b"inject_fft_failure", 0))
    torch.cuda.synchronize()
    assert pc == BACKEND_ERROR and pd == BACKEND_ERROR and e.value.code == BACKEND_ERROR
    assert bool(torch.all(out == 3.5 + 0j)) and bool(torch.all(power == -2.0))
    # normal operation again
    y = R.pulse_doppler(x, window="hann", widen=False)
    assert orc.rel_l2(y.cpu().numpy(), orc.pulse_doppler_fft64(x.cpu().numpy(), window="hann")) < TOL
