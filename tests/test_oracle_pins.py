"""CPU pins of the oracle beyond the reference's recorded digests: library
routines (numpy.fft, scipy.signal), closed forms and brute force on tiny
inputs.  Each check is chosen so that a plausible slip in the oracle (a
conjugate on the wrong operand, a shifted index, a dropped window or
normalisation, an off-by-one border) fails it."""
import numpy as np
import pytest

from oracle import oracle as orc

scipy_signal = pytest.importorskip("scipy.signal")


def _cplx(rng, *shape):
    return (rng.standard_normal(shape) + 1j * rng.standard_normal(shape)).astype(np.complex64)


@pytest.mark.parametrize("P,S,T,nfft,window,normalize", [(3, 64, 16, -1, "none", False),
                                                         (2, 50, 7, 64, "hamming", True),
                                                         (4, 40, 40, 32, "hann", False)])
def test_pulse_compression_direct_vs_numpy(P, S, T, nfft, window, normalize):
    rng = np.random.default_rng(1)
    x, t = _cplx(rng, P, S), _cplx(rng, T)
    direct = orc.pulse_compression(x, t, nfft, window, normalize, exact=True)
    # circular cross-correlation with numpy: y[k] = sum_t x[(k+t) mod n] conj(h[t])
    n = S if nfft == -1 else nfft
    h = t.astype(np.complex128)
    if window == "hamming":
        h = h * np.hamming(T)
    elif window == "hann":
        h = h * np.hanning(T)
    if normalize:
        h = h / np.linalg.norm(h)
    h = h[:n]
    xp = np.zeros((P, n), np.complex128)
    xp[:, :min(S, n)] = x[:, :min(S, n)]
    ref = np.array([[np.sum(xp[p, (k + np.arange(h.size)) % n] * np.conj(h)) for k in range(n)]
                    for p in range(P)])
    assert orc.rel_l2(direct, ref) < 1e-12
    assert orc.rel_l2(orc.pulse_compression_fft64(x, t, nfft, window, normalize), ref) < 1e-12


def test_pulse_compression_delta_and_autocorrelation():
    rng = np.random.default_rng(2)
    x = _cplx(rng, 2, 32)
    delta = np.zeros(8, np.complex64)
    delta[3] = 1
    y = orc.pulse_compression(x, delta, exact=True)
    assert np.allclose(y, np.roll(x, -3, axis=1), atol=1e-6)
    # a template equal to a delayed copy of the echo peaks at that delay
    t = _cplx(rng, 16)
    echo = np.zeros((1, 64), np.complex64)
    echo[0, 20:36] = t
    z = orc.pulse_compression(echo, t, exact=True)
    assert int(np.argmax(np.abs(z[0]))) == 20
    assert np.isclose(z[0, 20], np.sum(np.abs(t.astype(np.complex128)) ** 2))


@pytest.mark.parametrize("P,R,nfft,window", [(8, 5, -1, "none"), (6, 3, 16, "hann"), (16, 2, 8, "hamming")])
def test_pulse_doppler_exact_vs_numpy(P, R, nfft, window):
    rng = np.random.default_rng(3)
    x = _cplx(rng, P, R)
    w = {"none": np.ones(P), "hann": np.hanning(P), "hamming": np.hamming(P)}[window]
    n = P if nfft == -1 else nfft
    ref = np.fft.fft(x.astype(np.complex128) * w[:, None], n, axis=0) if n >= P else \
        np.fft.fft((x.astype(np.complex128) * w[:, None])[:n], n, axis=0)
    assert orc.rel_l2(orc.pulse_doppler(x, nfft, window, exact=True), ref) < 1e-12
    assert orc.rel_l2(orc.pulse_doppler_fft64(x, nfft, window), ref) < 1e-12


def test_pulse_doppler_tone_is_delta():
    P, f0 = 32, 5
    x = np.exp(2j * np.pi * f0 * np.arange(P) / P)[:, None].astype(np.complex64)
    y = orc.pulse_doppler(x, exact=True)[:, 0]
    expect = np.zeros(P)
    expect[f0] = P
    assert np.allclose(np.abs(y), expect, atol=1e-4)


def test_faithful_doppler_within_float_pi_of_exact():
    rng = np.random.default_rng(4)
    x = _cplx(rng, 128, 4)
    e = orc.rel_l2(orc.pulse_doppler(x, window="hann"), orc.pulse_doppler(x, window="hann", exact=True))
    assert 1e-9 < e < 1e-5  # the float-pi phase error is visible but small


def test_windows_match_numpy():
    for n in (1, 2, 7, 128):
        assert np.allclose(orc.window("hann", n, exact=True), np.hanning(n), atol=1e-15)
        assert np.allclose(orc.window("hamming", n, exact=True), np.hamming(n), atol=1e-15)


def test_cfar_alpha_closed_form():
    for pfa, n in ((1e-6, 416), (1e-3, 10), (0.5, 1), (1.0, 128)):
        assert orc.cfar_alpha(pfa, n) == pytest.approx(n * (pfa ** (-1.0 / n) - 1.0), rel=1e-14)
    assert orc.ca_cfar_count((2, 2), (8, 8)) == 416
    assert orc.cfar_alpha(1e-6, 416) == pytest.approx(14.04748, abs=1e-5)


def _brute_cfar(x, g, r, pfa):
    rows, cols = x.shape
    og, orr = g[0] + r[0], g[1] + r[1]
    n = (2 * og + 1) * (2 * orr + 1) - (2 * g[0] + 1) * (2 * g[1] + 1)
    alpha = n * (pfa ** (-1.0 / n) - 1.0)
    thr = np.zeros(x.shape)
    xd = x.astype(np.float64)
    for i in range(og, rows - og):
        for j in range(orr, cols - orr):
            outer = xd[i - og:i + og + 1, j - orr:j + orr + 1].sum()
            inner = xd[i - g[0]:i + g[0] + 1, j - g[1]:j + g[1] + 1].sum()
            thr[i, j] = alpha / n * (outer - inner)
    return thr


@pytest.mark.parametrize("shape,g,r", [((24, 30), (2, 2), (3, 4)), ((13, 40), (0, 1), (2, 3))])
def test_ca_cfar_brute_force(shape, g, r):
    rng = np.random.default_rng(5)
    x = rng.exponential(1.0, shape).astype(np.float32)
    brute = _brute_cfar(x, g, r, 1e-4)
    assert np.allclose(orc.ca_cfar_exact(x, g, r, 1e-4), brute, rtol=1e-12, atol=0)
    thr, det = orc.ca_cfar(x, g, r, 1e-4)
    assert np.allclose(thr, brute, rtol=2e-5, atol=1e-5)  # FP32 integral image
    og, orr = g[0] + r[0], g[1] + r[1]
    border = np.ones(shape, bool)
    border[og:shape[0] - og, orr:shape[1] - orr] = False
    assert np.all(thr[border] == 0) and np.all(det[border] == 1)


def test_ca_cfar_known_answers():
    # cuSignal test_radartools: all-ones input -> alpha in the interior, 0 on the border
    x = np.ones((30, 40), np.float32)
    thr, det = orc.ca_cfar(x, (1, 1), (2, 2), 1e-3)
    n = orc.ca_cfar_count((1, 1), (2, 2))
    alpha = np.float32(n * (1e-3 ** (-1.0 / n) - 1.0))
    assert np.allclose(thr[3:-3, 3:-3], alpha, rtol=1e-6)
    assert not det[3:-3, 3:-3].any()
    y = np.ones((30, 40), np.float32)
    y[15, 20] = 1e3
    _, det = orc.ca_cfar(y, (1, 1), (2, 2), 1e-3)
    assert det[15, 20] == 1


@pytest.mark.parametrize("up,down,K,n", [(1, 4, 31, 200), (3, 2, 12, 101), (5, 1, 7, 33), (2, 7, 50, 300)])
def test_upfirdn_vs_scipy(up, down, K, n):
    rng = np.random.default_rng(6)
    h = rng.standard_normal(K).astype(np.float32)
    x = rng.standard_normal((2, n)).astype(np.float32)
    ref = scipy_signal.upfirdn(h.astype(np.float64), x.astype(np.float64), up, down, axis=1)
    got = orc.upfirdn(h, x, up, down, axis=1)
    assert got.shape == ref.shape
    assert np.allclose(got, ref, rtol=1e-5, atol=1e-5)


@pytest.mark.parametrize("up,down,K", [(3, 2, 121), (1, 4, 40), (5, 3, 33)])
def test_resample_poly_custom_vs_scipy(up, down, K):
    rng = np.random.default_rng(7)
    h = scipy_signal.firwin(K, 1.0 / max(up, down)).astype(np.float32)
    x = rng.standard_normal(1000).astype(np.float32)
    ref = scipy_signal.resample_poly(x.astype(np.float64), up, down, window=h.astype(np.float64))
    got, kind = orc.resample_poly_f32(x, up, down, taps=h)
    assert kind == "FP32" and got.shape == ref.shape
    assert orc.rel_l2(got, ref) < 1e-6


def test_resample_poly_default_design_vs_scipy():
    # the default design is scipy's kaiser(5) firwin up to the A&S I0 approximation
    up, down = 3, 2
    h = orc.resample_design(up, down)
    half = 10 * max(up, down)
    ref = scipy_signal.firwin(2 * half + 1, 1.0 / max(up, down), window=("kaiser", 5.0)) * up
    assert h.size == 2 * half + 1
    assert np.allclose(h, ref, rtol=0, atol=1e-6 * np.abs(ref).max())


@pytest.mark.parametrize("K,n", [(9, 100), (255, 1000)])
def test_firfilter_vs_scipy_lfilter(K, n):
    rng = np.random.default_rng(8)
    b = rng.standard_normal(K).astype(np.float32)
    x = rng.uniform(-1, 1, n).astype(np.float32)
    zi = rng.standard_normal(K - 1).astype(np.float32)
    ref_y, ref_zf = scipy_signal.lfilter(b.astype(np.float64), [1.0], x.astype(np.float64), zi=zi.astype(np.float64))
    y, zf = orc.firfilter(b, x, zi=zi)
    assert np.allclose(y, ref_y, rtol=1e-5, atol=1e-5)
    assert np.allclose(zf, ref_zf, rtol=1e-5, atol=1e-5)


def test_cfar_alpha_f32_device_form():
    # FP32 alpha: a float32 value within FP32 rounding of the FP64 closed form;
    # powf's result p^(-1/n) ~ 1 is rounded before 1 is subtracted, so the
    # bound scales with n * p^(-1/n)
    for pfa, n in ((1e-6, 416), (1e-3, 24), (0.5, 1), (1.0, 128)):
        a = orc.cfar_alpha_f32(pfa, n)
        assert isinstance(a, np.float32)
        exact = orc.cfar_alpha(float(np.float32(pfa)), n)
        bound = 4 * np.finfo(np.float32).eps * n * float(np.float32(pfa)) ** (-1.0 / n)
        assert abs(float(a) - exact) <= bound
    assert orc.cfar_alpha_f32(1.0, 24) == 0.0


@pytest.mark.parametrize("dtype", ["FP32", "FP16", "INT32", "INT16", "INT8"])
def test_scaled_custom_device_matches_cpu_form_below_65536(dtype):
    # below up = 65536 the device form is the CPU form (pinned by the golden
    # resample_poly digests) rounded to float
    vals = {"FP32": [0.1, -3.5, 1e-3], "FP16": [0.1, -3.5, 1e-3], "INT32": [7, -100000, 2 ** 30],
            "INT16": [7, -300, 30000], "INT8": [7, -100, 127]}[dtype]
    for up in (1, 3, 127, 128, 255, 256, 32767, 32768):
        for v in vals:
            assert orc.scaled_custom_device(v, up, dtype) == np.float32(orc._scaled_custom(v, up, dtype))


def test_scaled_custom_device_closed_forms():
    # integer products wrap in the narrowest type the rule picks
    assert orc.scaled_custom_device(100, 3, "INT8") == 44.0          # 300 mod 256
    assert orc.scaled_custom_device(100, 200, "INT8") == 20000.0     # int16, no wrap
    assert orc.scaled_custom_device(300, 200, "INT16") == -5536.0    # 60000 - 65536
    assert orc.scaled_custom_device(300, 40000, "INT16") == 12000000.0  # int32 above 32767
    assert orc.scaled_custom_device(2 ** 30, 4, "INT32") == 0.0      # 2^32 mod 2^32
    # half: the product is rounded to half for up <= 255 only
    assert orc.scaled_custom_device(0.1, 3, "FP16") == np.float32(np.float16(np.float32(np.float16(0.1)) * 3))
    assert orc.scaled_custom_device(0.1, 300, "FP16") == np.float32(np.float16(0.1)) * np.float32(300)
    # float: FP32 product at every up (the CPU form switches to FP64 at 65536)
    assert orc.scaled_custom_device(0.1, 70000, "FP32") == np.float32(0.1) * np.float32(70000)


def test_ca_cfar_integer_detection_rule():
    # the integer rule compares with floor(T) in the input type: with
    # all-zero rows around one cell the threshold there is 0 and x = 0 is not
    # detected; negative inputs against threshold 0 never are
    x = np.zeros((15, 15), np.int16)
    x[7, 7] = 1
    thr, det = orc.ca_cfar(x, (1, 1), (2, 2), 1e-3)
    assert thr[7, 7] == 0 and det[7, 7] == 1
    assert det[6, 7] == 0  # x = 0, T > 0 from the neighbour
    xf = x.astype(np.float32)
    thr_f, _ = orc.ca_cfar(xf, (1, 1), (2, 2), 1e-3)
    assert np.array_equal(thr, thr_f)  # thresholds come from the float samples
    # fractional thresholds: x > floor(T), e.g. T = 0.3 -> x = 1 > 0 detected,
    # while the float rule would also detect 1 > 0.3; T = 1.7 -> x = 1 not, x = 2 yes
    y = np.full((15, 15), 1, np.int8)
    thr_y, det_y = orc.ca_cfar(y, (1, 1), (2, 2), 0.5)
    t = float(thr_y[7, 7])
    assert 0 < t < 1 and det_y[7, 7] == 1
    z = np.full((15, 15), 2, np.int8)
    z[7, 7] = 1
    thr_z, det_z = orc.ca_cfar(z, (1, 1), (2, 2), 1e-3)
    assert thr_z[7, 7] > 1 and det_z[7, 7] == (1 > np.floor(thr_z[7, 7]))


# ---------------------------------------------------------------------------
# decimate (filtering_typed.cpp:318-336, 2229-2317)
# ---------------------------------------------------------------------------

@pytest.mark.parametrize("numtaps,cutoff", [(1, 0.5), (33, 0.25), (81, 0.25), (101, 0.2), (161, 0.125), (40, 0.3)])
def test_firwin_lowpass_hamming_vs_scipy(numtaps, cutoff):
    """Same float cutoff into scipy.signal.firwin (Hamming, pass_zero, scale)."""
    from scipy import signal
    c = float(np.float32(cutoff))
    h = orc.firwin_lowpass_hamming(numtaps, c)
    assert np.allclose(h, signal.firwin(numtaps, c), rtol=0, atol=1e-15)
    assert abs(h.sum() - 1.0) < 1e-14


def test_firwin_rejects_nyquist_cutoff():
    """decimate's default design at q = 1 asks for cutoff 1, which firwin rejects."""
    with pytest.raises(ValueError):
        orc.decimate_design(1)


@pytest.mark.parametrize("shape,axis,q,zero_phase", [((1000,), -1, 4, True), ((1000,), -1, 4, False),
                                                     ((7, 300), 1, 3, True), ((300, 5), 0, 5, True),
                                                     ((4, 37, 3), 1, 2, False), ((50,), 0, 8, True)])
def test_decimate_default_design_vs_scipy(shape, axis, q, zero_phase):
    """scipy.signal.decimate(ftype='fir') designs firwin(20q+1, 1/q, hamming)
    and runs resample_poly(x, 1, q) (zero phase) or upfirdn(...)[:ceil(n/q)]:
    the same definition, with a double cutoff (taps differ by ~1e-8)."""
    from scipy import signal
    x = np.random.default_rng(q + len(shape)).standard_normal(shape)
    y = orc.decimate(x, q, axis=axis, zero_phase=zero_phase)
    ref = signal.decimate(x, q, ftype="fir", axis=axis, zero_phase=zero_phase)
    assert y.shape == ref.shape
    assert np.allclose(y, ref, rtol=0, atol=1e-6 * np.abs(x).max())


@pytest.mark.parametrize("q,zero_phase", [(1, True), (2, True), (3, False), (7, True), (12, False)])
def test_decimate_custom_taps_vs_defining_filters(q, zero_phase):
    """Custom FIR: zero phase is resample_poly(x, 1, q, window=h), the causal
    form the first ceil(n/q) outputs of upfirdn(h, x, 1, q) (FP64 path)."""
    from scipy import signal
    x = np.random.default_rng(q).standard_normal(523)
    h = np.array([1.0, 2.0, 1.0, -1.0, 1.0, 0.5, -0.25])
    y = orc.decimate(x, q, taps=h, zero_phase=zero_phase, fp32=False)
    if zero_phase:
        ref = signal.resample_poly(x, 1, q, window=h) if q > 1 else np.convolve(x, h)[(h.size - 1) // 2:][:x.size]
    else:
        ref = signal.upfirdn(h, x, 1, q)[:-(-x.size // q)]
    assert np.allclose(y, ref, rtol=0, atol=1e-12)


def test_decimate_fp32_path_rounds_in_tap_order():
    """The FP32-output path sums float products in tap order: it equals a
    float32 loop and stays within 1e-6 of the FP64 path."""
    x = (np.arange(64) % 11 - 5).astype(np.float64) * 0.25
    h = np.array([1.0, 2.0, 1.0, -1.0, 1.0])
    y32 = orc.decimate(x, 3, taps=h, fp32=True)
    y64 = orc.decimate(x, 3, taps=h, fp32=False)
    assert np.allclose(y32, y64, rtol=0, atol=1e-6)
    half, pre = 2, 3 - 2 % 3
    start = (half + pre) // 3
    acc = np.zeros(y32.size, np.float32)
    for o in range(y32.size):
        s = np.float32(0)
        for t in range(h.size):
            src = (o + start) * 3 - pre - t
            if 0 <= src < x.size:
                s = np.float32(s + np.float32(h[t]) * np.float32(x[src]))
        acc[o] = s
    assert np.array_equal(acc.astype(np.float64), y32)


# ---------------------------------------------------------------------------
# channelize_poly (filtering_typed.cpp:443-466 and its CPU form)
# ---------------------------------------------------------------------------

def _channelize_fft_form(x, h, M):
    """An independent formulation: the per-phase FIR as array slicing and the
    +i DFT as M times numpy's inverse FFT over the phase axis."""
    T = h.size // M
    P = x.size // M
    xb = x[:P * M].reshape(P, M)[:, ::-1]          # xb[g, ph] = x[g*M + M-1-ph]
    hb = h[:T * M].reshape(T, M) if T else np.zeros((0, M))
    F = np.zeros((P, M))
    for t in range(T):
        F[t:, :] += xb[:P - t, :] * hb[t][None, :]
    return (M * np.fft.ifft(F, axis=1)).T          # [M, P]


@pytest.mark.parametrize("n,nh,M", [(128, 32, 8), (1000, 64, 8), (4096, 256, 16), (777, 30, 6), (96, 0, 4),
                                    (1024, 1024, 32), (50, 7, 1), (4099, 100, 20)])
def test_channelize_vs_fft_formulation(n, nh, M):
    rng = np.random.default_rng(n + M)
    x = rng.standard_normal(n)
    h = rng.standard_normal(nh)
    y = orc.channelize_poly(x, h, M)
    ref = _channelize_fft_form(x, h, M)
    assert y.shape == ref.shape == (M, n // M)
    scale = max(1.0, np.abs(ref).max())
    assert np.allclose(y, ref, rtol=0, atol=2e-6 * scale)


def test_channelize_single_tap_is_blockwise_dft():
    """With one tap per phase equal to 1, each output column is the +i DFT of
    the reversed input block."""
    M = 8
    x = np.arange(64, dtype=np.float64)
    y = orc.channelize_poly(x, np.ones(M), M)
    for p in range(8):
        blk = x[p * M:(p + 1) * M][::-1]
        assert np.allclose(y[:, p], M * np.fft.ifft(blk), atol=1e-4)


def test_channelize_rejects_long_filters():
    with pytest.raises(ValueError):
        orc.channelize_poly(np.ones(64), np.ones(33 * 2), 2)


@pytest.mark.parametrize("boundary", ["fill", "wrap", "symm"])
@pytest.mark.parametrize("mode", ["full", "same", "valid"])
@pytest.mark.parametrize("shape1,shape2", [((13, 17), (3, 5)), ((9, 6), (5, 3)), ((4, 5), (7, 9)), ((20, 11), (1, 1))])
def test_correlate2d_vs_scipy(boundary, mode, shape1, shape2):
    """correlate2d against scipy.signal.correlate2d (FP64) on random floats
    for every boundary (fill with a nonzero fill value) and mode; odd kernels,
    where scipy's 'same' centring is the reference's r2/2, c2/2."""
    from scipy import signal
    rng = np.random.default_rng(sum(shape1) * 7 + sum(shape2))
    a = rng.standard_normal(shape1).astype(np.float32)
    b = rng.standard_normal(shape2).astype(np.float32)
    if mode == "valid" and not (shape1[0] >= shape2[0] and shape1[1] >= shape2[1]) \
            and not (shape2[0] >= shape1[0] and shape2[1] >= shape1[1]):
        pytest.skip("valid needs one input to dominate")
    got = orc.correlate2d(a, b, mode, boundary, 0.75, "FP32")
    ref = signal.correlate2d(a.astype(np.float64), b.astype(np.float64), mode=mode, boundary=boundary,
                             fillvalue=0.75)
    assert got.shape == ref.shape
    assert orc.rel_l2(got, ref) < 1e-6


def test_correlate2d_integer_wrap_and_fill():
    """INT8 sums wrap modulo 2^8 (no saturation) and the integer fill value is
    truncated toward zero, then wrapped: brute force in Python integers."""
    a = np.full((3, 3), 100, np.int8)
    b = np.full((2, 2), 3, np.int8)
    got = orc.correlate2d(a, b, "full", "fill", -258.9, "INT8")
    fill = -258 % 256  # trunc(-258.9) = -258 -> 254 -> -2 as int8
    fill = fill - 256 if fill >= 128 else fill
    assert orc.correlate2d_fill(-258.9, "INT8") == fill
    exp = np.zeros((4, 4), np.int64)
    for r in range(4):
        for c in range(4):
            s = 0
            for kr in range(2):
                for kc in range(2):
                    ir, ic = r - 1 + kr, c - 1 + kc
                    x = 100 if 0 <= ir < 3 and 0 <= ic < 3 else fill
                    s += x * 3
            exp[r, c] = ((s + 128) % 256) - 128
    assert np.array_equal(got.astype(np.int64), exp)
    with pytest.raises(ValueError):
        orc.correlate2d(a, b, "same", "fill", float("nan"), "INT8")


# ---------------------------------------------------------------------------
# N-D fftconvolve and correlate
# ---------------------------------------------------------------------------
_FFTCONV_CASES = [
    # (shape1, shape2, axes)
    ((37,), (9,), None),
    ((5, 7), (3, 4), None),
    ((6, 20), (1, 5), [1]),          # kernel broadcast over rows
    ((6, 20), (6, 5), [-1]),         # equal non-axis extents
    ((1, 20), (4, 5), [1]),          # in1 broadcast
    ((4, 6, 9), (2, 6, 2), [0, 2]),
    ((4, 1, 9), (2, 6, 2), [0, 2]),
    ((3, 4, 5), (3, 4, 5), None),
    ((8, 3), (5, 3), [0]),
]


@pytest.mark.parametrize("s1,s2,axes", _FFTCONV_CASES)
@pytest.mark.parametrize("mode", ["full", "same", "valid"])
@pytest.mark.parametrize("cplx", [False, True])
def test_fftconvolve_nd_vs_scipy(s1, s2, axes, mode, cplx):
    """scipy.signal.fftconvolve has cuSignal's semantics wherever cuSignal's
    'valid' swap rule (every dimension, not only the convolved ones) does not
    refuse the shapes; the oracle must match it there."""
    rng = np.random.default_rng(sum(s1) * 7 + sum(s2))
    a = rng.standard_normal(s1)
    b = rng.standard_normal(s2)
    if cplx:
        a = a + 1j * rng.standard_normal(s1)
    d1, d2 = np.array(s1), np.array(s2)
    if mode == "valid" and not (np.all(d1 >= d2) or np.all(d2 >= d1)):
        with pytest.raises(ValueError):
            orc.fftconvolve_nd(a, b, mode, axes)
        return
    ref = scipy_signal.fftconvolve(a, b, mode, axes=axes)
    got = orc.fftconvolve_nd(a, b, mode, axes)
    assert got.shape == ref.shape
    assert np.iscomplexobj(got) == cplx
    assert np.allclose(got, ref, rtol=1e-10, atol=1e-10)


@pytest.mark.parametrize("s1,s2,axes", [((5,), (3,), None), ((3, 4), (2, 3), None), ((3, 5), (1, 2), [1]),
                                        ((2, 3, 4), (2, 1, 3), [0, 2]), ((4, 2), (2, 2), [0])])
@pytest.mark.parametrize("mode", ["full", "same", "valid"])
def test_fftconvolve_nd_brute_force(s1, s2, axes, mode):
    rng = np.random.default_rng(3)
    a = rng.standard_normal(s1) + 1j * rng.standard_normal(s1)
    b = rng.standard_normal(s2)
    assert np.allclose(orc.fftconvolve_nd(a, b, mode, axes), orc.fftconvolve_nd(a, b, mode, axes, direct=True),
                       rtol=1e-10, atol=1e-10)


def test_fftconvolve_nd_cusignal_specific_rules():
    """Where cuSignal differs from current scipy: 'valid' decides the swap
    over every dimension (convolution_utils.py:40-70), so broadcast
    dimensions where in1 is smaller refuse the call; scalars multiply; an
    empty input gives an empty result; axes are validated."""
    a, b = np.ones((1, 10)), np.ones((3, 4))
    scipy_signal.fftconvolve(a, b, "valid", axes=[1])  # accepted by scipy
    with pytest.raises(ValueError):
        orc.fftconvolve_nd(a, b, "valid", [1])
    assert orc.fftconvolve_nd(np.float64(3.0), np.float64(-2.0)) == -6.0
    assert orc.fftconvolve_nd(np.ones((0, 3)), np.ones((2, 2))).size == 0
    for bad in ([2], [0, 0], []):
        with pytest.raises(ValueError):
            orc.fftconvolve_nd(np.ones((3, 3)), np.ones((2, 2)), "full", bad)
    with pytest.raises(ValueError):
        orc.fftconvolve_nd(np.ones((3, 4)), np.ones((2, 3)), "full", [0])  # 4 vs 3 on a broadcast axis


@pytest.mark.parametrize("s1,s2", [((6, 9), (3, 4)), ((3, 4), (6, 9)), ((4, 5, 3), (2, 3, 2)), ((7, 2), (2, 2))])
@pytest.mark.parametrize("mode", ["full", "same", "valid"])
@pytest.mark.parametrize("dtype", ["FP32", "FP16", "INT32", "INT16", "INT8"])
def test_correlate_nd_vs_scipy(s1, s2, mode, dtype):
    """scipy.signal.correlate (direct sums) in FP64, cast as the FFT method
    stores; integer inputs are small so the sums are exact."""
    rng = np.random.default_rng(len(s1) * 13 + s2[0])
    if orc.INTEGRAL[dtype]:
        a = rng.integers(-9, 10, s1).astype(orc.NP_DTYPE[dtype])
        b = rng.integers(-9, 10, s2).astype(orc.NP_DTYPE[dtype])
    else:
        a = (rng.standard_normal(s1) / 4).astype(orc.NP_DTYPE[dtype])
        b = (rng.standard_normal(s2) / 4).astype(orc.NP_DTYPE[dtype])
    ref = scipy_signal.correlate(a.astype(np.float64), b.astype(np.float64), mode, method="direct")
    got = orc.correlate_nd(a, b, mode, dtype)
    assert got.shape == ref.shape and got.dtype == orc.NP_DTYPE[dtype]
    if orc.INTEGRAL[dtype]:
        bits = {"INT32": 32, "INT16": 16, "INT8": 8}[dtype]
        wrapped = (np.rint(ref).astype(np.int64) % (1 << bits)).astype(np.dtype(f"uint{bits}")).view(got.dtype)
        assert np.array_equal(got, wrapped)
    else:
        assert np.allclose(got.astype(np.float64), ref.astype(orc.NP_DTYPE[dtype]).astype(np.float64),
                           rtol=2e-3 if dtype == "FP16" else 1e-6, atol=2e-3 if dtype == "FP16" else 1e-6)


def test_correlate_nd_shape_rules():
    """convolution_typed.cpp:153-199: valid needs one input to dominate;
    scalars multiply (in FP32); an empty input gives an empty output."""
    with pytest.raises(ValueError):
        orc.correlate_nd(np.ones((3, 5)), np.ones((4, 2)), "valid")
    assert orc.correlate_nd(np.float32(1.5), np.float32(-2.0)) == np.float32(-3.0)
    assert orc.correlate_nd(np.ones((0, 3), np.float32), np.ones((2, 2), np.float32)).size == 0
    # a 2 x 2 known answer: full correlation of [[1,2],[3,4]] with [[1,0],[0,1]]
    got = orc.correlate_nd(np.array([[1, 2], [3, 4]], np.float32), np.array([[1, 0], [0, 1]], np.float32))
    assert np.allclose(got, np.array([[1, 2, 0], [3, 5, 2], [0, 3, 4]], np.float32), rtol=0, atol=1e-6)
