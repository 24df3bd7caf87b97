"""Host-side mirror of the reference's radar operator API on torch device tensors.

Names, argument meaning, dtype rules and errors follow the reference's
cusignal_cpp radartools API (ZKX/cusignal_cpp/include/cusignal/operators/
radartools/radartools_typed.h): pulse_compression, pulse_doppler, cfar_alpha,
ca_cfar.  Every call goes through the C ABI of libzkx_radar.so; torch only
provides device memory and the current stream.  Shape/argument/dtype problems
raise ArgumentError, a ValueError (the reference throws std::invalid_argument)
carrying the ErrorCode (INVALID_SHAPE, INVALID_ARGUMENT, INVALID_DTYPE);
runtime failures raise ZkxError carrying the reference ErrorCode.
"""
from __future__ import annotations

import ctypes

import torch

from ._check import out_tensor
from ._lib import INVALID_ARGUMENT, INVALID_DTYPE, INVALID_SHAPE, ArgumentError, ZkxError, check, lib

WINDOWS = {"none": 0, "hann": 1, "hamming": 2, "explicit_fp32": 3, "explicit_fp64": 4}


def _stream(stream=None):
    s = stream if stream is not None else torch.cuda.current_stream()
    return ctypes.c_void_p(s.cuda_stream)


def _ptr(t):
    return None if t is None else ctypes.c_void_p(t.data_ptr())


def _device_tensor(t, dtype, name):
    if not isinstance(t, torch.Tensor) or not t.is_cuda:
        raise ArgumentError(INVALID_ARGUMENT, f"{name} must be a device tensor")
    if t.dtype != dtype:
        raise ArgumentError(INVALID_DTYPE, f"{name} must be {dtype}, got {t.dtype}")
    return t.contiguous()


def _window_code(window, explicit_window, length, name):
    if window not in WINDOWS:
        raise ArgumentError(INVALID_ARGUMENT, f"unsupported {name} window")
    code = WINDOWS[window]
    if code >= 3:
        if explicit_window is None or explicit_window.numel() != length:
            raise ArgumentError(INVALID_ARGUMENT, "explicit radar window size mismatch")
        explicit_window = _device_tensor(explicit_window, torch.float32, "explicit_window")
    elif explicit_window is not None:
        raise ArgumentError(INVALID_ARGUMENT, "explicit radar window requires explicit_fp32 or explicit_fp64 kind")
    return code, explicit_window


def _result_is_fp64(window):
    """radartools_typed.cpp:271-278 for float inputs: FP64 storage iff a window
    other than explicit_fp32 is applied."""
    return window not in ("none", "explicit_fp32")


def pulse_compression(x, template, *, normalize=False, nfft=-1, window="none",
                      explicit_window=None, widen=True, stream=None):
    """pulse_compression_device (radartools_typed.h:237-262).

    x: complex64 [num_pulses, samples_per_pulse]; template: complex64 [T].
    Returns [num_pulses, nfft]: complex128 storage when the reference would
    return ComplexFP64 (window != none) and widen=True, else complex64.
    """
    if isinstance(x, torch.Tensor) and x.dim() != 2:
        raise ArgumentError(INVALID_SHAPE, "pulse_compression shape is invalid")
    x = _device_tensor(x, torch.complex64, "x")
    template = _device_tensor(template, torch.complex64, "template")
    P, S = x.shape
    n = S if nfft == -1 else int(nfft)
    if n < 1 or template.numel() < 1:
        raise ArgumentError(INVALID_SHAPE, "pulse_compression shape is invalid")
    code, ew = _window_code(window, explicit_window, template.numel(), "pulse_compression")
    out = torch.empty((P, n), dtype=torch.complex64, device=x.device)
    check(lib().zkx_pulse_compression_c64(_ptr(x), P, S, _ptr(template), template.numel(), n,
                                          int(bool(normalize)), code, _ptr(ew), _ptr(out),
                                          _stream(stream)))
    if widen and _result_is_fp64(window):
        return widen_complex(out, stream)
    return out


def pc_template_spectrum(template, nfft, *, normalize=False, window="none", explicit_window=None,
                         stream=None):
    """Resident template spectrum H = FFT_nfft(window*h/||window*h||) for the
    prepared pulse-compression path (kept on the device between dwells)."""
    template = _device_tensor(template, torch.complex64, "template")
    code, ew = _window_code(window, explicit_window, template.numel(), "pulse_compression")
    H = torch.empty(int(nfft), dtype=torch.complex64, device=template.device)
    check(lib().zkx_pc_template_prepare(_ptr(template), template.numel(), int(nfft),
                                        int(bool(normalize)), code, _ptr(ew), _ptr(H),
                                        _stream(stream)))
    return H


def pulse_compression_prepared(x, spectrum, *, out=None, stream=None):
    """Pulse compression with a resident spectrum; x [..., P, S] -> [..., P, nfft]."""
    x = _device_tensor(x, torch.complex64, "x")
    spectrum = _device_tensor(spectrum, torch.complex64, "spectrum")
    n = spectrum.numel()
    S = x.shape[-1]
    rows = x.numel() // S
    shape = tuple(x.shape[:-1]) + (n,)
    if out is None:
        out = torch.empty(shape, dtype=torch.complex64, device=x.device)
    out_tensor(out, "out", x, torch.complex64, rows * n)
    check(lib().zkx_pulse_compression_prepared_c64(_ptr(x), rows, S, _ptr(spectrum), n, _ptr(out),
                                                   _stream(stream)))
    return out


def pulse_doppler(x, *, nfft=-1, window="none", explicit_window=None, out=None, power=None,
                  widen=True, stream=None):
    """pulse_doppler_device (radartools_typed.h:276-311).

    x: complex64 [num_pulses, samples_per_pulse] -> [nfft, samples_per_pulse].
    With `out` given (the preallocated overload) the FP32 result is written
    there and returned unwidened.  `power` (float32, same shape) optionally
    receives |Y|^2.
    """
    if isinstance(x, torch.Tensor) and x.dim() != 2:
        raise ArgumentError(INVALID_SHAPE, "pulse_doppler shape is invalid")
    x = _device_tensor(x, torch.complex64, "x")
    P, R = x.shape
    n = P if nfft == -1 else int(nfft)
    if n < 1:
        raise ArgumentError(INVALID_SHAPE, "pulse_doppler shape is invalid")
    code, ew = _window_code(window, explicit_window, P, "pulse_doppler")
    preallocated = out is not None
    if out is None:
        out = torch.empty((n, R), dtype=torch.complex64, device=x.device)
    elif out.numel() != n * R or out.dtype != torch.complex64:
        raise ArgumentError(INVALID_SHAPE, "pulse_doppler preallocated output shape mismatch")
    out_tensor(out, "out", x, torch.complex64, n * R)
    out_tensor(power, "power", x, torch.float32, n * R)
    check(lib().zkx_pulse_doppler_c64(_ptr(x), P, R, n, code, _ptr(ew), _ptr(out), _ptr(power),
                                      _stream(stream)))
    if not preallocated and widen and _result_is_fp64(window):
        return widen_complex(out, stream)
    return out


def doppler_power(x, *, nfft=-1, window="none", explicit_window=None, stream=None):
    """|pulse_doppler(x)|^2 as float32 without materialising the complex map."""
    x = _device_tensor(x, torch.complex64, "x")
    P, R = x.shape
    n = P if nfft == -1 else int(nfft)
    code, ew = _window_code(window, explicit_window, P, "pulse_doppler")
    power = torch.empty((n, R), dtype=torch.float32, device=x.device)
    check(lib().zkx_pulse_doppler_c64(_ptr(x), P, R, n, code, _ptr(ew), None, _ptr(power),
                                      _stream(stream)))
    return power


def doppler_dft_mfma(x, *, window="none", terms=3, out=None, power=None, stream=None):
    """Doppler DFT over pulses (nfft = P) on the matrix cores with split-BF16
    operands (terms=3) or plain BF16 (terms=1), FP32 accumulate.
    x: complex64 [P, R] or [C, P, R]; returns the complex map (and fills
    `power` with |Y|^2 when given).  P must be a multiple of 16."""
    x = _device_tensor(x, torch.complex64, "x")
    cube = x if x.dim() == 3 else x.unsqueeze(0)
    C, P, R = cube.shape
    codes = {"none": 0, "hann": 1, "hamming": 2}
    if window not in codes:
        raise ValueError("doppler_dft_mfma window must be none, hann or hamming")
    y = out if out is not None else torch.empty_like(cube)
    out_tensor(y, "out", cube, torch.complex64, cube.numel())
    out_tensor(power, "power", cube, torch.float32, cube.numel())
    check(lib().zkx_doppler_dft_mfma_c64(_ptr(cube), C, P, R, codes[window], int(terms), _ptr(y), _ptr(power),
                                         _stream(stream)))
    return y if x.dim() == 3 else y[0]


def _alpha_args(pfa, reference):
    """Scalar checks of cfar_alpha's inputs: a real scalar pfa (INVALID_DTYPE
    for complex or non-numeric input, the reference's typed entry points take
    only real business types) and a scalar reference-cell count
    (INVALID_SHAPE otherwise)."""
    import numpy as np
    for v, name in ((pfa, "pfa"), (reference, "reference")):
        if isinstance(v, torch.Tensor):
            if v.numel() != 1 or v.dim() > 1:
                raise ArgumentError(INVALID_SHAPE, f"cfar_alpha {name} must be a scalar")
            if v.is_complex() or v.dtype == torch.bool:
                raise ArgumentError(INVALID_DTYPE, f"cfar_alpha {name} must be real")
        elif np.ndim(v) != 0:
            raise ArgumentError(INVALID_SHAPE, f"cfar_alpha {name} must be a scalar")
        elif isinstance(v, (bool, np.bool_)) or not isinstance(v, (int, float, np.integer, np.floating)):
            raise ArgumentError(INVALID_DTYPE, f"cfar_alpha {name} must be real")
    if isinstance(reference, float) and not float(reference).is_integer():
        raise ArgumentError(INVALID_ARGUMENT, "cfar_alpha reference must be an integer count")
    ref = reference.item() if isinstance(reference, torch.Tensor) else reference
    p = pfa.item() if isinstance(pfa, torch.Tensor) else pfa
    return float(p), int(ref)


def cfar_alpha(pfa, reference):
    """cfar_alpha_typed_cpu (radartools_typed.cpp:592-603): N*(pfa^(-1/N)-1), FP64."""
    pfa, reference = _alpha_args(pfa, reference)
    out = ctypes.c_double()
    check(lib().zkx_cfar_alpha(pfa, reference, ctypes.byref(out)))
    return out.value


def cfar_alpha_device(pfa, reference, *, stream=None):
    """cfar_alpha_device<float> (radartools_typed.cpp:605-617): alpha computed in
    FP32 on the device (powf) from the pfa loaded to float, returned as a
    one-element float64 device tensor."""
    pfa, reference = _alpha_args(pfa, reference)
    p = float(torch.tensor(pfa, dtype=torch.float32))
    if reference < 1 or not (p > 0.0 and p <= 1.0):
        raise ArgumentError(INVALID_ARGUMENT, "cfar_alpha requires reference >= 1 and pfa in (0,1]")
    out = torch.empty(1, dtype=torch.float64, device=torch.device("cuda", torch.cuda.current_device()))
    check(lib().zkx_cfar_alpha_f32_device(ctypes.c_float(p), int(reference), _ptr(out), _stream(stream)))
    return out


AMBGFUN_CUTS = {"2d": 0, "delay": 1, "doppler": 2}


def ambgfun(x, fs, prf, y=None, cut="2d", cutValue=0.0, *, out_fp64=False, stream=None):
    """ambgfun_device (radartools_typed.h:134-139) with cusignal.ambgfun's
    call form (cusignal/radartools/radartools.py:146).

    x, y: complex64 device tensors; for a higher-rank x (2-D cut only) the
    first extent selects the samples and the norm covers the whole array.
    Returns the 2-D cut as float32 [Nx+Ny-1, Nfreq] (float64 when out_fp64,
    the storage the reference uses for integer inputs) and the delay /
    Doppler cuts as float64 [Nfreq] / [Nx+Ny-1].  prf does not enter the
    result, as in the reference.
    """
    del prf
    cut = str(cut).lower()
    if cut not in AMBGFUN_CUTS:
        raise ValueError("2d, delay, and doppler are the only cut values allowed")
    code = AMBGFUN_CUTS[cut]
    x = _device_tensor(x, torch.complex64, "x")
    if y is not None:
        y = _device_tensor(y, torch.complex64, "y")
    if x.numel() == 0 or (y is not None and y.numel() == 0):
        raise ValueError("ambgfun inputs must be nonempty")
    if code != 0 and (x.dim() > 1 or (y is not None and y.dim() > 1)):
        raise ValueError("ambgfun high-rank input is only defined for cut='2d'")
    if code != 0 and float(fs) == 0.0:
        raise ValueError("ambgfun fs must be nonzero for one-dimensional cuts")
    nx = x.shape[0] if x.dim() > 0 else 1
    ny = nx if y is None else (y.shape[0] if y.dim() > 0 else 1)
    n = lib().zkx_ambgfun_out_len(nx, ny, code)
    if n < 0:
        raise ValueError("ambgfun FFT length exceeds int capacity")
    rows = nx + ny - 1
    fp64 = code != 0 or out_fp64
    out = torch.empty(n, dtype=torch.float64 if fp64 else torch.float32, device=x.device)
    rc = lib().zkx_ambgfun_c64(_ptr(x), x.numel(), nx, _ptr(y), 0 if y is None else y.numel(), ny, code,
                               float(fs), float(cutValue), int(fp64), _ptr(out), _stream(stream))
    if rc in (1001, 1002):
        raise ValueError(lib().zkx_last_error_message().decode())
    check(rc)
    return out.view(rows, n // rows) if code == 0 else out


def ca_cfar(x, guard_cells, reference_cells, pfa=1e-3, *, threshold=None, detections=None,
            stream=None):
    """ca_cfar_device (radartools_typed.h:154-195): float32 map of rank 1 or 2.

    Returns (threshold float32, detections uint8), bit-identical to the
    reference's FP32 integral-image CFAR.
    """
    if isinstance(x, torch.Tensor) and x.dim() > 2:
        raise ArgumentError(INVALID_SHAPE, "ca_cfar supports only scalar, 1D, or 2D input")
    x = _device_tensor(x, torch.float32, "x")
    rank = x.dim()
    if rank == 0:
        thr = torch.zeros_like(x)
        return thr, (x > thr).to(torch.uint8)
    guard = list(guard_cells)
    ref = list(reference_cells)
    if len(guard) != rank or len(ref) != rank:
        raise ArgumentError(INVALID_ARGUMENT, "ca_cfar cell parameter rank mismatch")
    if threshold is None:
        threshold = torch.empty_like(x)
    if detections is None:
        detections = torch.empty(x.shape, dtype=torch.uint8, device=x.device)
    if threshold.numel() != x.numel() or detections.numel() != x.numel():
        raise ArgumentError(INVALID_SHAPE, "ca_cfar preallocated output shape must match input")
    out_tensor(threshold, "threshold", x, torch.float32, x.numel())
    out_tensor(detections, "detections", x, torch.uint8, x.numel())
    shape = (ctypes.c_int * rank)(*x.shape)
    g = (ctypes.c_int * rank)(*guard)
    r = (ctypes.c_int * rank)(*ref)
    code = lib().zkx_ca_cfar_f32(_ptr(x), rank, shape, g, r, float(pfa), _ptr(threshold),
                                 _ptr(detections), None, _stream(stream))
    check(code)
    return threshold, detections


def ca_cfar_windowed(x, guard_cells, reference_cells, pfa=1e-3, *, threshold=None,
                     detections=None, stream=None):
    """Rank-2 CA-CFAR with FP64 windowed box sums (the fused pipeline's CFAR)."""
    x = _device_tensor(x, torch.float32, "x")
    if x.dim() != 2:
        raise ArgumentError(INVALID_SHAPE, "ca_cfar_windowed needs a 2-D map")
    if threshold is None:
        threshold = torch.empty_like(x)
    if detections is None:
        detections = torch.empty(x.shape, dtype=torch.uint8, device=x.device)
    out_tensor(threshold, "threshold", x, torch.float32, x.numel())
    out_tensor(detections, "detections", x, torch.uint8, x.numel())
    code = lib().zkx_ca_cfar_window_f32(_ptr(x), x.shape[0], x.shape[1], guard_cells[0],
                                        guard_cells[1], reference_cells[0], reference_cells[1],
                                        float(pfa), _ptr(threshold), _ptr(detections),
                                        _stream(stream))
    check(code)
    return threshold, detections


def doppler_cfar(x, guard_cells=(2, 2), reference_cells=(8, 8), pfa=1e-6, *, window="hann",
                 threshold=None, detections=None, rd_map=None, power=None, stream=None):
    """Task1 steps 3-4 fused: Doppler FFT (nfft = num_pulses), |.|^2 and the
    FP64 windowed CA-CFAR for x [..., P, R] complex64.  Returns (threshold,
    detections) shaped like x."""
    x = _device_tensor(x, torch.complex64, "x")
    P, R = x.shape[-2], x.shape[-1]
    B = x.numel() // (P * R)
    if threshold is None:
        threshold = torch.empty(x.shape, dtype=torch.float32, device=x.device)
    if detections is None:
        detections = torch.empty(x.shape, dtype=torch.uint8, device=x.device)
    out_tensor(threshold, "threshold", x, torch.float32, x.numel())
    out_tensor(detections, "detections", x, torch.uint8, x.numel())
    out_tensor(rd_map, "rd_map", x, torch.complex64, x.numel())
    out_tensor(power, "power", x, torch.float32, x.numel())
    code = lib().zkx_doppler_cfar_c64(_ptr(x), B, P, R, WINDOWS[window], guard_cells[0], guard_cells[1],
                                      reference_cells[0], reference_cells[1], float(pfa), _ptr(threshold),
                                      _ptr(detections), _ptr(rd_map), _ptr(power), _stream(stream))
    check(code)
    return threshold, detections


def widen_complex(t, stream=None):
    out = torch.empty(t.shape, dtype=torch.complex128, device=t.device)
    check(lib().zkx_widen_c64_to_c128(_ptr(t), _ptr(out), t.numel(), _stream(stream)))
    return out


class DwellPipeline:
    """Task1 steps 2-4 for batches of same-shape dwells, template resident.

    run(x[B, P, R]) -> (threshold[B, P, R] float32, detections[B, P, R] uint8)
    cfar_mode 1 (default) fuses Doppler, |.|^2 and an FP64 windowed CFAR in
    one kernel; 0 reproduces the reference's FP32 integral-image thresholds
    bit for bit (separate Doppler and CFAR kernels).
    """

    def __init__(self, template, num_pulses, range_bins, *, normalize=True, pc_window="hamming",
                 doppler_window="hann", guard=(2, 2), reference=(8, 8), pfa=1e-6, cfar_mode=1,
                 max_batch=1, device="cuda"):
        self.P, self.R = int(num_pulses), int(range_bins)
        self.guard, self.reference, self.pfa = tuple(guard), tuple(reference), float(pfa)
        self.cfar_mode = int(cfar_mode)
        self.doppler_window = WINDOWS[doppler_window]
        t = template if isinstance(template, torch.Tensor) else torch.as_tensor(template)
        t = t.to(device=device, dtype=torch.complex64)
        self.spectrum = pc_template_spectrum(t, self.R, normalize=normalize, window=pc_window)
        self.max_batch = 0
        self.workspace = None
        self.reserve(max_batch)

    def reserve(self, batch):
        if batch > self.max_batch:
            nbytes = lib().zkx_dwell_workspace_bytes(batch, self.P, self.R)
            self.workspace = torch.empty(nbytes, dtype=torch.uint8, device=self.spectrum.device)
            self.max_batch = batch

    def run(self, x, *, threshold=None, detections=None, rd_map=None, stream=None):
        x = _device_tensor(x, torch.complex64, "x")
        if x.dim() == 2:
            x = x.unsqueeze(0)
        B, P, R = x.shape
        if (P, R) != (self.P, self.R):
            raise ValueError("dwell shape does not match the pipeline")
        self.reserve(B)
        if threshold is None:
            threshold = torch.empty((B, P, R), dtype=torch.float32, device=x.device)
        if detections is None:
            detections = torch.empty((B, P, R), dtype=torch.uint8, device=x.device)
        out_tensor(threshold, "threshold", x, torch.float32, x.numel())
        out_tensor(detections, "detections", x, torch.uint8, x.numel())
        out_tensor(rd_map, "rd_map", x, torch.complex64, x.numel())
        check(lib().zkx_dwell_batch_c64(
            _ptr(x), B, P, R, _ptr(self.spectrum), self.doppler_window, self.guard[0],
            self.guard[1], self.reference[0], self.reference[1], self.pfa, self.cfar_mode,
            _ptr(threshold), _ptr(detections), _ptr(rd_map), _ptr(self.workspace),
            _stream(stream)))
        return threshold, detections


__all__ = ["pulse_compression", "pc_template_spectrum", "pulse_compression_prepared",
           "pulse_doppler", "doppler_power", "cfar_alpha", "ca_cfar", "ca_cfar_windowed",
           "doppler_cfar", "widen_complex", "DwellPipeline", "ZkxError"]
