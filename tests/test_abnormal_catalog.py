"""The reference's abnormal-input catalog for the hot-path operators, run
against this library without a GPU.

tests/golden/abnormal_catalog_hot_ops.csv holds the 21 rows of
ZKX/test_all/operators/abnormal/operator_abnormal_case_catalog.csv for
firfilter, resample_poly, upfirdn, ca_cfar, cfar_alpha, pulse_compression and
pulse_doppler (copied by tests/golden/extract_abnormal_catalog.py).  Each case
below applies the row's mutation to a valid call of the product -- the C ABI
through ctypes, or the Python operator API where the rule lives in the host
wrapper -- and checks, as the reference's runner does
(operator_abnormal_runner.cpp:117-160, abnormal_runtime.h:40-80), that the
call failed, wrote no output and reported the catalog's ErrorCode.

How the failure rows are forced from the host:
  * "inject CUDA ... kernel failure" (CUDA_RUNTIME_ERROR): the reference forces
    a real runtime error from a call that cannot reach a device
    (operator_abnormal_validation.h:75-91).  Here the call is valid and no
    device exists, so it fails in the HIP runtime.  Those calls and the host
    buffers every case passes are only safe without a device, so the module
    is skipped wherever torch sees one.
  * "inject FFT execution failure" (BACKEND_ERROR): the library's
    "inject_fft_failure" switch makes the transform stage report a failed
    backend status before any device work, as the reference synthesizes one
    (operator_abnormal_validation.h:93-104).

Two rows name a code that the reference's operators themselves never raise:
its runner takes every validation row from a validator that maps the
abnormal type straight to the expected code
(operator_abnormal_validation.h:123-167) and never calls the operator.
  * pulse_compression with a template longer than the pulse: the reference's
    operator accepts it and uses the first nfft taps
    (radartools_typed.cpp:618-640, the tap loop stops at nfft).
  * resample_poly over a zero-length axis: the reference's layout accepts it
    and returns an empty output (filtering_typed.cpp:933-990).
For these two the test pins the operator behaviour (no validation error: the
first runs into the missing device, the second succeeds with nothing to
write) and checks that the catalog names exactly these two as synthetic-only.
"""
import csv
import ctypes
import os

import numpy as np
import pytest

torch = pytest.importorskip("torch")

from zkx_dev_amd import _lib  # noqa: E402

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
CATALOG = os.path.join(ROOT, "tests", "golden", "abnormal_catalog_hot_ops.csv")
CODES = {name: code for code, name in _lib.ERRORS.items()}
SYNTHETIC_ONLY = {
    "radartools.pulse_compression.invalid_shape.template_length_exceeds_samples_per_pulse",
    "filtering.resample_poly.invalid_shape.input_axis_length_is_zero",
}

pytestmark = pytest.mark.skipif(torch.cuda.device_count() > 0,
                                reason="forces failures with host buffers; only safe without a device")


def catalog():
    with open(CATALOG, newline="", encoding="utf-8") as fh:
        return list(csv.DictReader(fh))


class Host:
    """Host buffers standing in for device arrays: inputs of `n` float32
    values and an output pre-filled with a sentinel so that a write shows."""

    SENTINEL = np.float32(-7.25)

    def __init__(self, n=8192):
        self.x = np.linspace(-1, 1, n, dtype=np.float32)
        self.out = np.full(n, self.SENTINEL, np.float32)
        self.out2 = np.full(n, self.SENTINEL, np.float32)

    @staticmethod
    def p(a):
        return ctypes.c_void_p(a.ctypes.data)

    def untouched(self):
        return bool(np.all(self.out == self.SENTINEL) and np.all(self.out2 == self.SENTINEL))


def _python(call):
    """Code of the ArgumentError a Python API call raises (None if none)."""
    try:
        call()
    except _lib.ArgumentError as e:
        return e.code
    return None


def _case_firfilter(row, h):
    L = _lib.lib()
    t = row["abnormal_type"]
    if t == "invalid_shape":   # filter coefficients are empty
        return L.zkx_firfilter_f32(h.p(h.x), 0, h.p(h.x), 1, 256, 1, None, h.p(h.out), None, None)
    if t == "invalid_parameter":   # axis outside rank: a host-wrapper rule
        from zkx_dev_amd import filtering
        return _python(lambda: filtering._lines((4, 256), 2))
    return L.zkx_firfilter_f32(h.p(h.x), 8, h.p(h.x), 1, 256, 1, None, h.p(h.out), None, None)


def _case_resample_poly(row, h):
    L = _lib.lib()
    t = row["abnormal_type"]
    if t == "invalid_shape":   # zero-length axis: accepted, nothing written
        return L.zkx_resample_poly_f32(h.p(h.x), 8, h.p(h.x), 1, 0, 1, 3, 2, h.p(h.out), None)
    if t == "invalid_parameter":
        codes = {L.zkx_resample_poly_f32(h.p(h.x), 8, h.p(h.x), 1, 256, 1, u, d, h.p(h.out), None)
                 for u, d in ((0, 2), (3, 0))}
        assert len(codes) == 1
        return codes.pop()
    return L.zkx_resample_poly_f32(h.p(h.x), 8, h.p(h.x), 1, 256, 1, 3, 2, h.p(h.out), None)


def _case_upfirdn(row, h):
    L = _lib.lib()
    t = row["abnormal_type"]
    if t == "invalid_shape":   # empty taps
        return L.zkx_upfirdn_f32(h.p(h.x), 0, h.p(h.x), 1, 256, 1, 2, 1, h.p(h.out), None)
    if t == "invalid_parameter":
        codes = {L.zkx_upfirdn_f32(h.p(h.x), 8, h.p(h.x), 1, 256, 1, u, d, h.p(h.out), None)
                 for u, d in ((0, 1), (2, 0))}
        assert len(codes) == 1
        return codes.pop()
    return L.zkx_upfirdn_f32(h.p(h.x), 8, h.p(h.x), 1, 256, 1, 2, 1, h.p(h.out), None)


def _case_ca_cfar(row, h):
    L = _lib.lib()
    t = row["abnormal_type"]
    shape = (ctypes.c_int * 3)(16, 16, 16)
    guard = (ctypes.c_int * 3)(1, 1, 1)
    ref = (ctypes.c_int * 3)(2, 2, 2)
    if t == "invalid_shape":   # rank 3, through the C ABI and the Python API
        from zkx_dev_amd import radar
        py = _python(lambda: radar.ca_cfar(torch.zeros(4, 16, 16), (1, 1, 1), (2, 2, 2)))
        abi = L.zkx_ca_cfar_f32(h.p(h.x), 3, shape, guard, ref, 1e-3, h.p(h.out), h.p(h.out2), None, None)
        assert py == abi
        return abi
    if t == "invalid_parameter":   # pfa outside (0, 1]
        codes = {L.zkx_ca_cfar_f32(h.p(h.x), 2, shape, guard, ref, pfa, h.p(h.out), h.p(h.out2), None, None)
                 for pfa in (0.0, -0.5, 1.5, float("nan"))}
        assert len(codes) == 1
        return codes.pop()
    return L.zkx_ca_cfar_f32(h.p(h.x), 2, shape, guard, ref, 1e-3, h.p(h.out), h.p(h.out2), None, None)


def _case_cfar_alpha(row, h):
    from zkx_dev_amd import radar
    t = row["abnormal_type"]
    if t == "invalid_parameter":   # reference cells = 0, C ABI and Python
        a = ctypes.c_double(-1.0)
        abi = _lib.lib().zkx_cfar_alpha(1e-3, 0, ctypes.byref(a))
        assert a.value == -1.0
        assert _python(lambda: radar.cfar_alpha(1e-3, 0)) == abi
        return abi
    if t == "invalid_shape":   # reference count given as a vector
        return _python(lambda: radar.cfar_alpha(1e-3, np.array([8, 8])))
    if t == "invalid_dtype":   # complex business input
        codes = {_python(lambda: radar.cfar_alpha(np.complex64(1e-3), 16)),
                 _python(lambda: radar.cfar_alpha(torch.tensor(1e-3 + 0j), 16))}
        assert len(codes) == 1
        return codes.pop()
    raise AssertionError(t)


def _with_injected_fft_failure(call):
    L = _lib.lib()
    assert L.zkx_set_kernel_variant(b"inject_fft_failure", 1) == 0
    try:
        return call()
    finally:
        assert L.zkx_set_kernel_variant(b"inject_fft_failure", 0) == 0


def _case_pulse_compression(row, h):
    L = _lib.lib()
    t = row["abnormal_type"]
    P, S = 4, 256
    if t == "invalid_shape":   # template longer than the pulse: accepted
        return L.zkx_pulse_compression_c64(h.p(h.x), P, S, h.p(h.x), S + 64, -1, 0, 0, None, h.p(h.out), None)
    if t == "invalid_parameter":   # unknown window kind
        from zkx_dev_amd import radar
        py = _python(lambda: radar._window_code("unknown", None, 16, "pulse_compression"))
        abi = L.zkx_pulse_compression_c64(h.p(h.x), P, S, h.p(h.x), 16, -1, 0, 99, None, h.p(h.out), None)
        assert py == abi
        return abi
    return _with_injected_fft_failure(lambda: L.zkx_pulse_compression_c64(
        h.p(h.x), P, S, h.p(h.x), 16, -1, 0, 1, None, h.p(h.out), None))


def _case_pulse_doppler(row, h):
    L = _lib.lib()
    t = row["abnormal_type"]
    P, S = 16, 128
    if t == "invalid_shape":   # rank 1 and rank 3 inputs
        from zkx_dev_amd import radar
        codes = {_python(lambda: radar.pulse_doppler(torch.zeros(P * S, dtype=torch.complex64))),
                 _python(lambda: radar.pulse_doppler(torch.zeros(2, P, S, dtype=torch.complex64)))}
        assert len(codes) == 1
        return codes.pop()
    if t == "invalid_parameter":
        return L.zkx_pulse_doppler_c64(h.p(h.x), P, S, -1, 99, None, h.p(h.out), h.p(h.out2), None)
    return _with_injected_fft_failure(lambda: L.zkx_pulse_doppler_c64(
        h.p(h.x), P, S, -1, 1, None, h.p(h.out), h.p(h.out2), None))


CASES = {
    "firfilter": _case_firfilter, "resample_poly": _case_resample_poly, "upfirdn": _case_upfirdn,
    "ca_cfar": _case_ca_cfar, "cfar_alpha": _case_cfar_alpha,
    "pulse_compression": _case_pulse_compression, "pulse_doppler": _case_pulse_doppler,
}


def test_catalog_covers_the_hot_operators():
    rows = catalog()
    assert len(rows) == 21
    assert {r["operator_name"] for r in rows} == set(CASES)
    assert SYNTHETIC_ONLY <= {r["abnormal_case_id"] for r in rows}
    # every expected code is one the library reports
    assert {r["expected_code"] for r in rows} <= set(CODES)


@pytest.mark.parametrize("row", catalog(), ids=lambda r: r["abnormal_case_id"])
def test_abnormal_case(row):
    h = Host()
    code = CASES[row["operator_name"]](row, h)
    assert h.untouched(), "a rejected call wrote its output"
    if row["abnormal_case_id"] in SYNTHETIC_ONLY:
        # accepted by the operator, as the reference's operator accepts it
        assert code in (0, CODES["CUDA_RUNTIME_ERROR"]), code
        return
    assert code == CODES[row["expected_code"]], (code, row["expected_code"])


def test_injected_fft_failure_switch_resets():
    L = _lib.lib()
    assert L.zkx_get_kernel_variant(b"inject_fft_failure") == 0
    assert L.zkx_set_kernel_variant(b"inject_fft_failure", 2) == _lib.INVALID_ARGUMENT
    assert L.zkx_get_kernel_variant(b"inject_fft_failure") == 0


def test_length_limits_refused_before_device_work():
    """Transform lengths beyond the FFT's range are refused with their codes
    before anything touches the (here absent) device."""
    L = _lib.lib()
    h = Host(16)
    big = (1 << 25) + 1  # nx + ny - 1 > 2^26 for the 2-D / delay cuts
    for cut in (0, 1):
        rc = L.zkx_ambgfun_c64(h.p(h.x), big, big, h.p(h.x), big, big, cut, 1.0, 0.0, 1, h.p(h.out), None)
        assert rc == CODES["UNSUPPORTED_OPERATION"], (cut, rc)
    s1 = (ctypes.c_longlong * 1)(1 << 26)
    s2 = (ctypes.c_longlong * 1)(2)
    rc = L.zkx_fftconvolve_nd(h.p(h.x), s1, 0, h.p(h.x), s2, 0, 1, 1, 0, h.p(h.out), None)
    assert rc == CODES["INVALID_SHAPE"]
    d1 = (ctypes.c_int * 2)(1 << 16, 1 << 16)
    d2 = (ctypes.c_int * 2)(2, 2)
    rc = L.zkx_correlate_nd(h.p(h.x), d1, h.p(h.x), d2, 2, 0, 0, h.p(h.out), None)
    assert rc == CODES["INVALID_SHAPE"]
    assert h.untouched()
