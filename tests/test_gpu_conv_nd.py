"""GPU parity of the N-D FFT convolutions against the FP64 oracle.

fftconvolve over N-D axes (cuSignal convolve.py:159-320) and correlate over
N-D inputs by FFT (convolution_typed.cu:150-241): complex64 / float32 results
within 1e-5 relative L2 of the FP64 oracle (FP32 transforms); integer
correlations of small integers exact after the half-to-even rounding;
float16 within half-precision rounding.  Cases cover broadcasting on
non-convolved axes, 'valid' operand swaps, padded axes beyond one LDS pass
(four-step) on contiguous and strided axes, the overlap-save row route, and
the reference's error rules.
"""
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


def _inputs(rng, s1, s2, c1, c2):
    a = rng.standard_normal(s1).astype(np.float32)
    b = rng.standard_normal(s2).astype(np.float32)
    if c1:
        a = (a + 1j * rng.standard_normal(s1)).astype(np.complex64)
    if c2:
        b = (b + 1j * rng.standard_normal(s2)).astype(np.complex64)
    return a, b


CASES = [
    ((37,), (9,), None),
    ((5, 7), (3, 4), None),
    ((6, 20), (1, 5), [1]),
    ((6, 20), (6, 5), [-1]),
    ((1, 20), (4, 5), [1]),
    ((4, 6, 9), (2, 6, 2), [0, 2]),
    ((4, 1, 9), (2, 6, 2), [0, 2]),
    ((3, 4, 5), (3, 4, 5), None),
    ((8, 3), (5, 3), [0]),
    ((120, 300), (17, 31), None),
    ((2, 3, 40, 50), (1, 3, 7, 9), [2, 3]),
]


@pytest.mark.parametrize("s1,s2,axes", CASES)
@pytest.mark.parametrize("mode", ["full", "same", "valid"])
@pytest.mark.parametrize("kinds", [(False, False), (True, False), (True, True)])
def test_fftconvolve_nd(F, s1, s2, axes, mode, kinds):
    rng = np.random.default_rng(sum(s1) + 3 * sum(s2))
    a, b = _inputs(rng, s1, s2, *kinds)
    print(f"fftconvolve {s1} {s2} axes={axes} {mode} complex={kinds}", flush=True)
    try:
        ref = orc.fftconvolve_nd(a, b, mode, axes)
    except ValueError:
        with pytest.raises(ValueError):
            F.fftconvolve(torch.from_numpy(a).cuda(), torch.from_numpy(b).cuda(), mode, axes)
        return
    y = F.fftconvolve(torch.from_numpy(a).cuda(), torch.from_numpy(b).cuda(), mode, axes)
    assert tuple(y.shape) == ref.shape
    assert y.dtype == (torch.complex64 if any(kinds) else torch.float32)
    assert orc.rel_l2(y.cpu().numpy(), ref) < TOL


@pytest.mark.parametrize("s1,s2,axes", [((3, 6000), (1, 3000), [1]),     # contiguous axis, four-step
                                        ((6000, 3), (3000, 1), [0]),     # strided axis, four-step
                                        ((2, 9000), (2, 700), None)])
def test_fftconvolve_nd_long_axes(F, s1, s2, axes):
    rng = np.random.default_rng(11)
    a, b = _inputs(rng, s1, s2, True, False)
    print(f"fftconvolve long {s1} {s2} axes={axes}", flush=True)
    for mode in ("full", "same"):
        y = F.fftconvolve(torch.from_numpy(a).cuda(), torch.from_numpy(b).cuda(), mode, axes)
        assert orc.rel_l2(y.cpu().numpy(), orc.fftconvolve_nd(a, b, mode, axes)) < TOL


@pytest.mark.parametrize("s1,s2,axes", [((1000,), (77,), None),             # one self-paired row
                                        ((300, 40), (40, 9), None),        # 512-point last pass, 64 row pairs
                                        ((130, 7), (127, 3), [0, 1]),      # 256-point last pass
                                        ((3, 700, 20), (3, 50, 5), [1, 2]),  # leading axis not negated
                                        ((5, 2000), (5, 100), [1]),        # every row its own partner
                                        ((2100, 6), (30, 6), [0])])        # 4096-point last pass
def test_fftconvolve_nd_fused_product(F, s1, s2, axes):
    """Two real operands whose lowest convolved axis pads to 256..4096 points
    take the fused last-axis transform + packed product + inverse kernel."""
    rng = np.random.default_rng(sum(s1) + 7 * sum(s2))
    a, b = _inputs(rng, s1, s2, False, False)
    print(f"fftconvolve fused {s1} {s2} axes={axes}", flush=True)
    for mode in ("full", "same", "valid"):
        y = F.fftconvolve(torch.from_numpy(a).cuda(), torch.from_numpy(b).cuda(), mode, axes)
        ref = orc.fftconvolve_nd(a, b, mode, axes)
        assert tuple(y.shape) == ref.shape and y.dtype == torch.float32
        assert orc.rel_l2(y.cpu().numpy(), ref) < TOL


def test_fftconvolve_nd_rows_route_matches_general(F):
    """A complex kernel row against every row of in1 takes the overlap-save
    route; its result agrees with the general engine (real part of a real
    kernel run through the N-D path) and the oracle."""
    rng = np.random.default_rng(5)
    a, b = _inputs(rng, (16, 5000), (1, 301), True, True)
    for mode in ("full", "same"):
        y = F.fftconvolve(torch.from_numpy(a).cuda(), torch.from_numpy(b).cuda(), mode, [1])
        assert orc.rel_l2(y.cpu().numpy(), orc.fftconvolve_nd(a, b, mode, [1])) < TOL


def test_fftconvolve_nd_scalars_empty_and_errors(F):
    from zkx_dev_amd._lib import ArgumentError
    one = torch.tensor(3.0, device="cuda")
    two = torch.tensor(-2.0, device="cuda")
    r = F.fftconvolve(one, two)
    assert r.dim() == 0 and abs(float(r) + 6.0) < 1e-6
    assert F.fftconvolve(torch.ones(0, 3, device="cuda"), torch.ones(2, 2, device="cuda")).numel() == 0
    x = torch.ones(3, 4, device="cuda")
    with pytest.raises(ArgumentError) as e:
        F.fftconvolve(x, torch.ones(2, 3, device="cuda"), "full", [0])  # 4 vs 3 on a broadcast axis
    assert e.value.code == 1002
    with pytest.raises(ArgumentError) as e:
        F.fftconvolve(torch.ones(1, 10, device="cuda"), torch.ones(3, 4, device="cuda"), "valid", [1])
    assert e.value.code == 1002
    for bad in ([2], [0, 0], []):
        with pytest.raises(ArgumentError) as e:
            F.fftconvolve(x, x, "full", bad)
        assert e.value.code == 1001
    with pytest.raises(ArgumentError):
        F.fftconvolve(x, torch.ones(2, device="cuda").double())


DTYPES = {"FP32": torch.float32, "FP16": torch.float16, "INT32": torch.int32, "INT16": torch.int16,
          "INT8": torch.int8}


@pytest.mark.parametrize("s1,s2", [((6, 9), (3, 4)), ((3, 4), (6, 9)), ((4, 5, 3), (2, 3, 2)), ((7, 2), (2, 2)),
                                   ((64, 200), (9, 33)), ((2, 3, 4, 5), (2, 2, 2, 2)),
                                   ((300, 20), (30, 5))])  # fused 512-point last pass
@pytest.mark.parametrize("mode", ["full", "same", "valid"])
@pytest.mark.parametrize("dtype", list(DTYPES))
def test_correlate_nd(s1, s2, mode, dtype):
    from zkx_dev_amd import convolution
    rng = np.random.default_rng(len(s1) * 5 + s2[0])
    if orc.INTEGRAL[dtype]:
        a = rng.integers(-9, 10, s1).astype(orc.NP_DTYPE[dtype])
        b = rng.integers(-9, 10, s2).astype(orc.NP_DTYPE[dtype])
    else:
        a = (rng.standard_normal(s1) / 4).astype(orc.NP_DTYPE[dtype])
        b = (rng.standard_normal(s2) / 4).astype(orc.NP_DTYPE[dtype])
    print(f"correlate_nd {s1} {s2} {mode} {dtype}", flush=True)
    ref = orc.correlate_nd(a, b, mode, dtype)
    y = convolution.correlate(torch.from_numpy(a).cuda(), torch.from_numpy(b).cuda(), mode).cpu().numpy()
    assert y.shape == ref.shape and y.dtype == ref.dtype
    if orc.INTEGRAL[dtype]:
        assert np.array_equal(y, ref)
    elif dtype == "FP16":
        assert np.allclose(y.astype(np.float64), ref.astype(np.float64), rtol=2e-3, atol=2e-3)
    else:
        assert orc.rel_l2(y, ref) < TOL


def test_correlate_nd_scalars_empty_and_errors():
    from zkx_dev_amd import convolution
    from zkx_dev_amd._lib import ArgumentError
    s = convolution.correlate(torch.tensor(1.5, device="cuda"), torch.tensor(-2.0, device="cuda"))
    assert s.dim() == 0 and float(s) == -3.0
    si = convolution.correlate(torch.tensor(100, dtype=torch.int8, device="cuda"),
                               torch.tensor(3, dtype=torch.int8, device="cuda"))
    assert int(si) == np.int8(np.uint8(300 % 256).view(np.int8))
    e0 = convolution.correlate(torch.ones(0, 3, device="cuda"), torch.ones(2, 2, device="cuda"))
    assert e0.numel() == 0
    with pytest.raises(ArgumentError) as e:
        convolution.correlate(torch.ones(3, 5, device="cuda"), torch.ones(4, 2, device="cuda"), "valid")
    assert e.value.code == 1002
    with pytest.raises(ArgumentError) as e:
        convolution.correlate(torch.ones(3, 5, device="cuda"), torch.ones(2, 2, device="cuda"), "full", "direct")
    assert e.value.code == 1001
    with pytest.raises(ArgumentError) as e:
        convolution.correlate(torch.ones(3, 5, device="cuda"), torch.ones(2, device="cuda"))
    assert e.value.code == 1002
