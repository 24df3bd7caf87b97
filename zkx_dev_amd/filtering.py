"""Host-side mirror of the reference's polyphase filtering API on torch device tensors.

Follows ZKX/cusignal_cpp/include/cusignal/operators/filtering/filtering_typed.h:
upfirdn (UpfirdnOptions{shape, axis=-1}), resample_poly (ResamplePolyOptions{shape,
axis=0, window_mode, kaiser_beta, design_window}, optional custom coefficients)
firfilter (FirfilterOptions{shape, axis=-1}, optional zi) and decimate
(DecimateOptions{shape, axis=-1, filter_order, zero_phase}).  float32 lines
are the reference's FP32 path; complex64 lines are this library's extension
(real taps applied per component).  All compute runs in libzkx_radar.so.
"""
from __future__ import annotations

import ctypes
import math

import numpy as np
import torch

from ._lib import INVALID_ARGUMENT, INVALID_DTYPE, INVALID_SHAPE, ArgumentError, check, lib
from .radar import _ptr, _stream

_DTYPE_F32 = 0  # ZKX_DTYPE_F32


def _lines(shape, axis):
    rank = len(shape)
    ax = axis + rank if axis < 0 else axis
    if ax < 0 or ax >= rank:
        raise ArgumentError(INVALID_ARGUMENT, "axis is outside rank")
    outer = int(np.prod(shape[:ax])) if ax > 0 else 1
    inner = int(np.prod(shape[ax + 1:])) if ax + 1 < rank else 1
    return outer, int(shape[ax]), inner, ax


def _signal(x, name="x"):
    if not isinstance(x, torch.Tensor) or not x.is_cuda:
        raise ArgumentError(INVALID_ARGUMENT, f"{name} must be a device tensor")
    if x.dtype not in (torch.float32, torch.complex64):
        raise ArgumentError(INVALID_DTYPE, f"{name} must be float32 or complex64")
    return x.contiguous()


def _taps(h):
    if not isinstance(h, torch.Tensor):
        h = torch.as_tensor(np.asarray(h, np.float32), device="cuda")
    return h.to(torch.float32).contiguous()


def upfirdn(h, x, up=1, down=1, axis=-1, *, stream=None):
    """upfirdn_device (filtering_typed.h:744-748): upsample by `up`, filter
    with h, downsample by `down`, along `axis` of a rank-1 or rank-2 signal."""
    x = _signal(x)
    h = _taps(h)
    if h.numel() == 0:
        raise ArgumentError(INVALID_SHAPE, "upfirdn coefficients")
    if up < 1 or down < 1:
        raise ArgumentError(INVALID_ARGUMENT, "upfirdn factors")
    if x.dim() < 1 or x.dim() > 2:
        raise ArgumentError(INVALID_SHAPE, "upfirdn rank")
    outer, n, inner, ax = _lines(x.shape, axis)
    out_n = lib().zkx_upfirdn_out_len(n, h.numel(), up, down)
    shape = list(x.shape)
    shape[ax] = out_n
    y = torch.empty(shape, dtype=x.dtype, device=x.device)
    fn = lib().zkx_upfirdn_c64 if x.is_complex() else lib().zkx_upfirdn_f32
    check(fn(_ptr(h), h.numel(), _ptr(x), outer, n, inner, up, down, _ptr(y), _stream(stream)))
    return y


def _bessel_i0_as(value):
    """Abramowitz-Stegun polynomial I0, the approximation the reference's
    default design uses (filtering_typed.cpp:159-174)."""
    a = abs(value)
    if a < 3.75:
        y = (value / 3.75) ** 2
        return 1.0 + y * (3.5156229 + y * (3.0899424 + y * (1.2067492 + y * (0.2659732 + y *
                          (0.0360768 + y * 0.0045813)))))
    y = 3.75 / a
    return math.exp(a) / math.sqrt(a) * (0.39894228 + y * (0.01328592 + y * (0.00225319 + y * (
        -0.00157565 + y * (0.00916281 + y * (-0.02057706 + y * (0.02635537 + y * (
            -0.01647633 + y * 0.00392377))))))))


def design_resample_poly_coefficients(up, down, window="kaiser", beta=5.0, design_window=None):
    """Default resample_poly filter (filtering_typed.cpp:176-218), FP64 on the
    host: 2*10*max(up,down)+1 taps of a windowed sinc at cutoff 1/max(up,down),
    normalised to unit sum and scaled by up (up/down gcd-reduced first)."""
    g = math.gcd(up, down)
    u, d = up // g, down // g
    half = 10 * max(u, d)
    count = 2 * half + 1
    idx = np.arange(count)
    if window == "kaiser":
        den = _bessel_i0_as(beta)
        w = np.array([_bessel_i0_as(beta * math.sqrt(max(0.0, 1.0 - ((i - half) / half) ** 2))) / den
                      for i in range(count)])
    elif window == "hamming":
        w = 0.54 - 0.46 * np.cos(6.283185307179586476925286766559 * idx / (count - 1))
    elif window == "explicit":
        w = np.asarray(design_window, np.float64)
        if w.size != count:
            raise ValueError("resample_poly design window shape")
    else:
        raise ValueError("resample_poly window mode")
    cutoff = 1.0 / max(u, d)
    off = idx - half
    with np.errstate(invalid="ignore", divide="ignore"):
        sinc = np.where(off == 0, cutoff, np.sin(np.pi * cutoff * off) / (np.pi * off))
    c = sinc * w
    return c / np.sum(c) * u


def resample_poly(x, up, down, axis=0, *, coefficients=None, window="kaiser", beta=5.0,
                  design_window=None, widen=True, stream=None):
    """resample_poly_device (filtering_typed.h:765-770).

    Custom coefficients are scaled by the gcd-reduced `up` in float32 as the
    reference does for FP32 input and give a float32 result.  Without them the
    FP64 default design is rounded to float32 taps and the result is FP64
    storage (float64 / complex128) when widen=True, as the reference returns.
    """
    x = _signal(x)
    if up < 1 or down < 1:
        raise ArgumentError(INVALID_ARGUMENT, "resample_poly factors")
    outer, n, inner, ax = _lines(x.shape, axis)
    g = math.gcd(up, down)
    u, d = up // g, down // g
    if u == 1 and d == 1:
        return x.clone()
    if x.dim() > 2:
        raise ArgumentError(INVALID_SHAPE, "resample_poly nonidentity rank")
    if coefficients is not None and isinstance(coefficients, torch.Tensor) and coefficients.is_cuda and u < 65536:
        # device-resident float32 taps are scaled by the library (c * up in float32)
        c = coefficients.to(torch.float32).contiguous()
        taps_d = torch.empty(c.numel(), dtype=torch.float32, device=x.device)
        check(lib().zkx_resample_poly_scale_taps(_ptr(c), _DTYPE_F32, c.numel(), u, _ptr(taps_d), _stream(stream)))
        fp64_out = False
    else:
        if coefficients is not None:
            t = np.asarray(coefficients.cpu() if isinstance(coefficients, torch.Tensor) else coefficients,
                           np.float32)
            taps = (t * np.float32(u)).astype(np.float32) if u < 65536 else (t.astype(np.float64) * u)
            fp64_out = False
        else:
            taps = design_resample_poly_coefficients(up, down, window, beta, design_window)
            fp64_out = True
        taps_d = torch.as_tensor(np.ascontiguousarray(taps, np.float32), device=x.device)
    out_n = (n * u + d - 1) // d
    shape = list(x.shape)
    shape[ax] = out_n
    y = torch.empty(shape, dtype=x.dtype, device=x.device)
    fn = lib().zkx_resample_poly_c64 if x.is_complex() else lib().zkx_resample_poly_f32
    check(fn(_ptr(taps_d), taps_d.numel(), _ptr(x), outer, n, inner, up, down, _ptr(y),
             _stream(stream)))
    if fp64_out and widen:
        return y.to(torch.complex128 if x.is_complex() else torch.float64)
    return y


def firfilter(b, x, axis=-1, zi=None, *, stream=None):
    """firfilter_device (filtering_typed.h:552-565): causal FIR y = b * x along
    `axis`, first N samples; with zi (broadcast to the state shape, axis
    length K-1) also returns zf.  float32 only, as the reference."""
    if not isinstance(x, torch.Tensor) or not x.is_cuda:
        raise ArgumentError(INVALID_ARGUMENT, "x must be a float32 device tensor")
    if x.dtype != torch.float32:
        raise ArgumentError(INVALID_DTYPE, "x must be a float32 device tensor")
    x = x.contiguous()
    b = _taps(b)
    if b.numel() == 0:
        raise ArgumentError(INVALID_SHAPE, "firfilter coefficients")
    outer, n, inner, ax = _lines(x.shape, axis)
    y = torch.empty_like(x)
    zf = zi_t = None
    if zi is not None:
        state_shape = list(x.shape)
        state_shape[ax] = b.numel() - 1
        zi_t = torch.as_tensor(zi, dtype=torch.float32, device=x.device)
        try:
            zi_t = zi_t.broadcast_to(state_shape).contiguous()
        except RuntimeError as e:
            raise ArgumentError(INVALID_SHAPE, "firfilter zi broadcast shape") from e
        zf = torch.empty(state_shape, dtype=torch.float32, device=x.device)
    check(lib().zkx_firfilter_f32(_ptr(b), b.numel(), _ptr(x), outer, n, inner, _ptr(zi_t), _ptr(y),
                                  _ptr(zf), _stream(stream)))
    return (y, zf) if zi is not None else y


def design_decimate_filter(q, order=None):
    """design_decimate_filter (filtering_typed.cpp:318-336): firwin with
    order + 1 taps (order 20*q by default), cutoff 1/q in float, Hamming
    window, unit DC gain; FP64 taps designed on the host by the library."""
    q = int(q)
    if q < 1:
        raise ValueError("decimate q")
    order = 20 * q if order is None else int(order)
    if order < 0:
        raise ValueError("decimate filter order")
    taps = np.empty(order + 1, np.float64)
    code = lib().zkx_firwin_lowpass_hamming(order + 1, float(np.float32(1.0) / np.float32(q)),
                                            taps.ctypes.data_as(ctypes.c_void_p))
    if code == 1001:
        raise ValueError("firwin cutoff")
    check(code)
    return taps


def decimate(x, q, n=None, axis=-1, zero_phase=True, *, widen=True, stream=None):
    """decimate_device (filtering_typed.h:702-706) in cusignal.decimate's call
    form (cusignal/filtering/resample.py:82): an FIR anti-aliasing filter, then
    every q-th sample along `axis` of a float32 array of any rank.

    n: None (filter order 20*q), an int (the order) or the FIR coefficients.
    The designed filter gives FP64 storage (float64 when widen=True), custom
    coefficients a float32 result, as the reference returns."""
    if not isinstance(x, torch.Tensor) or not x.is_cuda or x.dtype != torch.float32:
        raise ValueError("x must be a float32 device tensor")
    q = int(q)
    if q < 1:
        raise ValueError("decimate q")
    x = x.contiguous()
    if x.dim() == 0:
        raise ValueError("decimate input rank")
    outer, length, inner, ax = _lines(x.shape, axis)
    custom = n is not None and not isinstance(n, (int, np.integer))
    if custom:
        taps_d = _taps(n)
        if taps_d.numel() == 0:
            raise ValueError("decimate coefficients")
    else:
        taps_d = torch.as_tensor(design_decimate_filter(q, n).astype(np.float32), device=x.device)
    out_n = 0 if length == 0 else 1 + (length - 1) // q
    shape = list(x.shape)
    shape[ax] = out_n
    y = torch.empty(shape, dtype=torch.float32, device=x.device)
    code = lib().zkx_decimate_f32(_ptr(taps_d), taps_d.numel(), _ptr(x), outer, length, inner, q,
                                  int(bool(zero_phase)), _ptr(y), _stream(stream))
    if code in (1001, 1002):
        raise ValueError(lib().zkx_last_error_message().decode())
    check(code)
    if not custom and widen:
        return y.to(torch.float64)
    return y


def channelize_poly(x, h, n_chans, *, stream=None):
    """channelize_poly_device (filtering_typed.h:513-523) in cusignal's call
    form (cusignal/filtering/channelize_poly.py): polyphase filter bank of
    n_chans channels, complex64 [n_chans, len(x) // n_chans] channel-major.
    x and h are float32; for higher-rank arrays the first extent selects the
    samples (the reference's len(x) rule)."""
    if not isinstance(x, torch.Tensor) or not x.is_cuda or x.dtype != torch.float32:
        raise ValueError("x must be a float32 device tensor")
    n_chans = int(n_chans)
    if n_chans <= 0:
        raise ValueError("channelize_poly n_chans")
    x = x.contiguous()
    h = _taps(h)
    nx = x.shape[0] if x.dim() > 0 else 1
    nh = h.shape[0] if h.dim() > 0 else 1
    y = torch.empty((n_chans, nx // n_chans), dtype=torch.complex64, device=x.device)
    code = lib().zkx_channelize_poly_f32(_ptr(x), nx, _ptr(h), nh, n_chans, _ptr(y), _stream(stream))
    if code in (1001, 1002):
        raise ValueError(lib().zkx_last_error_message().decode())
    check(code)
    return y


__all__ = ["upfirdn", "resample_poly", "firfilter", "decimate", "channelize_poly",
           "design_resample_poly_coefficients", "design_decimate_filter"]


def _fft_axes(axes, ndim):
    """cuSignal's axes normalisation (fftpack_helper.py:279-360): None means
    every axis; negative axes count from the end; out of range, repeated or
    (when given) empty axes are errors."""
    if axes is None:
        return list(range(ndim))
    ax = np.atleast_1d(np.asarray(axes))
    if ax.ndim != 1:
        raise ArgumentError(INVALID_ARGUMENT, "when given, axes values must be a scalar or vector")
    if ax.size and not np.issubdtype(ax.dtype, np.integer):
        raise ArgumentError(INVALID_ARGUMENT, "when given, axes values must be integers")
    ax = [int(a) + ndim if int(a) < 0 else int(a) for a in ax]
    if any(a < 0 or a >= ndim for a in ax):
        raise ArgumentError(INVALID_ARGUMENT, "axes exceeds dimensionality of input")
    if len(set(ax)) != len(ax):
        raise ArgumentError(INVALID_ARGUMENT, "all axes must be unique")
    if not ax:
        raise ArgumentError(INVALID_ARGUMENT, "when provided, axes cannot be empty")
    return sorted(ax)


def fftconvolve(in1, in2, mode="full", axes=None, *, stream=None):
    """fftconvolve (cuSignal convolve.py:159-320): N-D linear convolution of
    two device tensors of one rank over `axes` (default all), the other axes
    broadcast (equal extents or 1); modes 'full' / 'same' / 'valid'.  float32
    and complex64 inputs; the result is float32 when both are real, complex64
    otherwise.

    Extension kept from the row form: a 1-D in2 against an in1 of higher
    rank (axes not given) convolves every last-axis row of in1 with in2 and
    returns complex64."""
    modes = {"full": 0, "same": 1, "valid": 2}
    if mode not in modes:
        raise ArgumentError(INVALID_ARGUMENT, "acceptable mode flags are 'valid', 'same', or 'full'")
    for t, name in ((in1, "in1"), (in2, "in2")):
        if not isinstance(t, torch.Tensor) or not t.is_cuda:
            raise ArgumentError(INVALID_ARGUMENT, f"{name} must be a device tensor")
        if t.dtype not in (torch.float32, torch.complex64):
            raise ArgumentError(INVALID_DTYPE, f"{name} must be float32 or complex64")
    if in2.dim() == 1 and in1.dim() > 1 and axes is None:
        return _fftconvolve_rows(in1, in2, modes[mode], stream)
    if in1.dim() != in2.dim():
        raise ArgumentError(INVALID_SHAPE, "in1 and in2 should have the same dimensionality")
    if in1.numel() == 0 or in2.numel() == 0:
        return torch.empty(0, dtype=torch.float32, device=in1.device)
    scalar = in1.dim() == 0
    a = (in1.reshape(1) if scalar else in1).contiguous()
    b = (in2.reshape(1) if scalar else in2).contiguous()
    rank = a.dim()
    if rank > 8:
        raise ArgumentError(INVALID_SHAPE, "fftconvolve supports at most 8 dimensions")
    ax = _fft_axes(axes, rank)
    mask = sum(1 << d for d in ax)
    complex_out = a.is_complex() or b.is_complex()
    s1 = (ctypes.c_longlong * rank)(*a.shape)
    s2 = (ctypes.c_longlong * rank)(*b.shape)
    shape = (ctypes.c_longlong * rank)()
    check(lib().zkx_fftconvolve_nd_shape(s1, s2, rank, mask, modes[mode], shape))
    out_shape = tuple(int(v) for v in shape)
    y = torch.empty(out_shape, dtype=torch.complex64 if complex_out else torch.float32, device=a.device)
    check(lib().zkx_fftconvolve_nd(_ptr(a), s1, int(a.is_complex()), _ptr(b), s2, int(b.is_complex()), rank,
                                   mask, modes[mode], _ptr(y), _stream(stream)))
    return y.reshape(()) if scalar else y


def _fftconvolve_rows(in1, in2, mode, stream):
    a = in1.to(torch.complex64).contiguous()
    b = in2.to(torch.complex64).contiguous()
    if a.numel() == 0 or b.numel() == 0:
        return torch.empty(0, dtype=torch.complex64, device=a.device)
    n1 = a.shape[-1]
    rows = a.numel() // n1
    out_len = lib().zkx_fftconvolve_out_len(n1, b.numel(), mode)
    y = torch.empty(tuple(a.shape[:-1]) + (out_len,), dtype=torch.complex64, device=a.device)
    check(lib().zkx_fftconvolve_c64(_ptr(a), rows, n1, _ptr(b), b.numel(), mode, _ptr(y), _stream(stream)))
    return y


__all__.append("fftconvolve")
