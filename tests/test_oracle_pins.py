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
