"""GPU checks of the fft_* backend ABI against numpy FP64 and the reference
backend's shift / frequency definitions."""
import numpy as np
import pytest

torch = pytest.importorskip("torch")

from oracle import oracle as orc  # noqa: E402

pytestmark = pytest.mark.gpu


@pytest.fixture(scope="module")
def Fm():
    from zkx_dev_amd import fft
    assert torch.cuda.is_available()
    return fft


@pytest.mark.parametrize("n", [2, 4, 8, 16, 32, 64, 128, 256, 512, 1024, 2048, 4096, 8192])
@pytest.mark.parametrize("inverse", [False, True])
def test_fft_vs_numpy(Fm, n, inverse):
    rng = np.random.default_rng(n)
    batch = max(1, 4096 // n)
    x = (rng.standard_normal((batch, n)) + 1j * rng.standard_normal((batch, n))).astype(np.complex64)
    y = Fm.fft(torch.from_numpy(x).cuda(), inverse=inverse).cpu().numpy()
    ref = np.fft.ifft(x.astype(np.complex128), axis=1) if inverse else np.fft.fft(x.astype(np.complex128), axis=1)
    assert orc.rel_l2(y, ref) < 1e-6


def test_fft_parseval_and_roundtrip(Fm):
    rng = np.random.default_rng(1)
    x = (rng.standard_normal((64, 4096)) + 1j * rng.standard_normal((64, 4096))).astype(np.complex64)
    xt = torch.from_numpy(x).cuda()
    X = Fm.fft(xt)
    e_t = float((xt.abs() ** 2).sum())
    e_f = float((X.abs() ** 2).sum()) / 4096
    assert abs(e_t - e_f) / e_t < 1e-5
    back = Fm.fft(X, inverse=True)
    assert orc.rel_l2(back.cpu().numpy(), x) < 1e-6


def test_fft_plan_errors(Fm):
    with pytest.raises(Fm.FFTError):
        Fm.Plan(0)
    with pytest.raises(Fm.FFTError):
        Fm.Plan((1 << 26) * 2)  # powers of two up to 2^26
    with pytest.raises(Fm.FFTError):
        Fm.Plan((1 << 25) + 1)  # other lengths up to 2^25
    with pytest.raises(Fm.FFTError):
        Fm.Plan(64, batch=0)
    with pytest.raises(Fm.FFTError):
        Fm.Plan(64, type=7)


@pytest.mark.parametrize("n", [1, 3, 5, 7, 12, 100, 127, 1000, 1500, 4095])
@pytest.mark.parametrize("inverse", [False, True])
def test_fft_bluestein_vs_numpy(Fm, n, inverse):
    """Lengths that are not powers of two (Bluestein), forward unnormalised,
    inverse 1/n, as the reference backend."""
    rng = np.random.default_rng(n + 7)
    batch = max(1, 2048 // n)
    x = (rng.standard_normal((batch, n)) + 1j * rng.standard_normal((batch, n))).astype(np.complex64)
    y = Fm.fft(torch.from_numpy(x).cuda(), inverse=inverse).cpu().numpy()
    ref = np.fft.ifft(x.astype(np.complex128), axis=1) if inverse else np.fft.fft(x.astype(np.complex128), axis=1)
    assert orc.rel_l2(y, ref) < 1e-5


@pytest.mark.parametrize("n,batch", [(16384, 3), (32768, 2), (1 << 16, 2), (1 << 20, 1), (1 << 22, 1),
                                     (4097, 3), (5000, 2), (8191, 2), (10007, 2), (65537, 1), (100000, 1),
                                     (1 << 21 | 1, 1)])
@pytest.mark.parametrize("inverse", [False, True])
def test_fft_long_lengths_vs_numpy(Fm, n, batch, inverse):
    """Lengths beyond one LDS pass: four-step for powers of two above 8192,
    global-pass Bluestein for other lengths above 4096 (the reference backend
    takes any length)."""
    print(f"fft n={n} batch={batch} inverse={inverse}", flush=True)
    rng = np.random.default_rng(n % 1000 + 3)
    x = (rng.standard_normal((batch, n)) + 1j * rng.standard_normal((batch, n))).astype(np.complex64)
    y = Fm.fft(torch.from_numpy(x).cuda(), inverse=inverse).cpu().numpy()
    ref = np.fft.ifft(x.astype(np.complex128), axis=1) if inverse else np.fft.fft(x.astype(np.complex128), axis=1)
    assert orc.rel_l2(y, ref) < 1e-5


@pytest.mark.parametrize("ny,nx", [(3, 10007), (20000, 5), (2, 16384)])
def test_fft2_long_axis_vs_numpy(Fm, ny, nx):
    rng = np.random.default_rng(ny + nx)
    x = (rng.standard_normal((ny, nx)) + 1j * rng.standard_normal((ny, nx))).astype(np.complex64)
    y = Fm.fft2(torch.from_numpy(x).cuda()).cpu().numpy()
    assert orc.rel_l2(y, np.fft.fft2(x.astype(np.complex128))) < 1e-5


@pytest.mark.parametrize("n", [8, 64, 4096, 6, 9, 100, 1001, 10007, 16384])
def test_rfft_and_irfft_reference_semantics(Fm, n):
    rng = np.random.default_rng(n)
    x = rng.standard_normal((3, n)).astype(np.float32)
    X = Fm.rfft(torch.from_numpy(x).cuda()).cpu().numpy()
    assert X.shape == (3, n // 2 + 1)
    assert orc.rel_l2(X, np.fft.rfft(x.astype(np.float64), axis=1)) < 1e-5
    # C2R: bins 0 .. n/2 given, bins 1 .. n/2+1-2 mirrored, the rest zero,
    # inverse scaled by 1/n (and once more by 1/n for non-power-of-two n)
    half = n // 2 + 1
    Xin = (rng.standard_normal((2, half)) + 1j * rng.standard_normal((2, half))).astype(np.complex64)
    full = np.zeros((2, n), np.complex128)
    full[:, :half] = Xin
    for m in range(1, half - 1):
        full[:, n - m] = np.conj(Xin[:, m])
    expect = np.fft.ifft(full, axis=1).real  # unnormalised inverse, then 1/n
    if n & (n - 1):
        expect = expect / n
    got = Fm.irfft(torch.from_numpy(Xin).cuda(), n).cpu().numpy()
    assert orc.rel_l2(got, expect) < 1e-5


@pytest.mark.parametrize("ny,nx", [(64, 128), (48, 100), (7, 9), (1, 256), (256, 3)])
@pytest.mark.parametrize("inverse", [False, True])
def test_fft2_vs_numpy(Fm, ny, nx, inverse):
    rng = np.random.default_rng(ny * 1000 + nx)
    x = (rng.standard_normal((ny, nx)) + 1j * rng.standard_normal((ny, nx))).astype(np.complex64)
    y = Fm.fft2(torch.from_numpy(x).cuda(), inverse=inverse).cpu().numpy()
    ref = np.fft.ifft2(x.astype(np.complex128)) if inverse else np.fft.fft2(x.astype(np.complex128))
    assert orc.rel_l2(y, ref) < 1e-5


@pytest.mark.parametrize("n", [7, 8, 101])
def test_shift_matches_reference_definition(Fm, n):
    """fft_thrust.cu fft_shift_kernel: out[i] = in[(i + n/2) % n] for both
    directions."""
    x = (np.arange(3 * n) + 1j * np.arange(3 * n)).astype(np.complex64).reshape(3, n)
    xt = torch.from_numpy(x).cuda()
    expect = x[:, (np.arange(n) + n // 2) % n]
    assert np.array_equal(Fm.fftshift(xt).cpu().numpy(), expect)
    assert np.array_equal(Fm.fftshift(xt, inverse=True).cpu().numpy(), expect)


@pytest.mark.parametrize("ny,nx", [(4, 6), (5, 7), (1, 9)])
def test_shift2d_matches_reference_definition(Fm, ny, nx):
    x = (np.arange(ny * nx) * (1 + 1j)).astype(np.complex64).reshape(ny, nx)
    xt = torch.from_numpy(x).cuda()
    fwd = x[((np.arange(ny) + ny // 2) % ny)][:, (np.arange(nx) + nx // 2) % nx]
    inv = x[((np.arange(ny) + (ny + 1) // 2) % ny)][:, (np.arange(nx) + (nx + 1) // 2) % nx]
    assert np.array_equal(Fm.fftshift2d(xt).cpu().numpy(), fwd)
    assert np.array_equal(Fm.fftshift2d(xt, inverse=True).cpu().numpy(), inv)
    if ny % 2 == 0 and nx % 2 == 0:
        assert np.array_equal(fwd, np.fft.fftshift(x))


@pytest.mark.parametrize("n,d", [(8, 1.0), (9, 0.5), (1, 2.0)])
def test_freqs(Fm, n, d):
    np.testing.assert_allclose(Fm.fftfreq(n, d).numpy(), np.fft.fftfreq(n, d), rtol=1e-6)
    np.testing.assert_allclose(Fm.fftfreq(n, d, rfft=True).numpy(), np.fft.rfftfreq(n, d), rtol=1e-6)
    np.testing.assert_allclose(Fm.fftfreq(n, d, device="cuda").cpu().numpy(), np.fft.fftfreq(n, d), rtol=1e-6)
