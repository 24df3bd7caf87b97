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
        Fm.Plan(1000)
    with pytest.raises(Fm.FFTError):
        Fm.Plan(16384)
    with pytest.raises(Fm.FFTError):
        Fm.Plan(64, batch=0)


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
