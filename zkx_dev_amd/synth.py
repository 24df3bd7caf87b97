"""Seeded synthetic radar dwells (host side, numpy).

A dwell is a complex64 cube x[P, R] of P pulses by R range samples:
    x[p, r] = sum_i A * s(r - d_i) * exp(2 pi i f_i p / P) + n[p, r]
with s the LFM chirp, d_i the target range bins, f_i their Doppler bins, and
n complex Gaussian noise of per-sample power A^2 / snr (per component
variance half that).  The chirp follows the reference's Task1 echo model
(ZKX/Task1/pipeline/src/steps/step1/step1.cu:27-37): phase pi*(B/tau)*t^2 with
t = n/fs - tau/2, tau = T/fs; here fs = 1 and B = bw_over_fs.
"""
from __future__ import annotations

from dataclasses import dataclass, field

import numpy as np


def lfm_chirp(taps: int, bw_over_fs: float = 0.5) -> np.ndarray:
    n = np.arange(taps, dtype=np.float64)
    tau = float(taps)
    t = n - 0.5 * tau
    phase = np.pi * (bw_over_fs / tau) * t * t
    return np.exp(1j * phase).astype(np.complex64)


@dataclass
class Dwell:
    x: np.ndarray               # complex64 [P, R]
    template: np.ndarray        # complex64 [T]
    range_bins: list = field(default_factory=list)
    doppler_bins: list = field(default_factory=list)


def make_dwell(P=128, R=4096, taps=256, n_targets=5, snr_db=-10.0, seed=20261203,
               amplitude=1.0, bw_over_fs=0.5, margin=10) -> Dwell:
    """One seeded dwell with point targets clear of the CFAR border."""
    rng = np.random.default_rng(seed)
    s = lfm_chirp(taps, bw_over_fs).astype(np.complex128)
    hi_r = max(margin + 1, R - margin - taps)
    rbins = sorted(int(v) for v in rng.choice(np.arange(margin, hi_r), size=n_targets, replace=False))
    hi_f = max(margin + 1, P - margin)
    fbins = [int(v) for v in rng.integers(margin, hi_f, size=n_targets)]
    x = np.zeros((P, R), np.complex128)
    p = np.arange(P)[:, None]
    for d, f in zip(rbins, fbins):
        m = min(taps, R - d)
        x[:, d:d + m] += amplitude * s[None, :m] * np.exp(2j * np.pi * f * p / P)
    noise_power = amplitude ** 2 / (10.0 ** (snr_db / 10.0))
    sigma = np.sqrt(noise_power / 2.0)
    x += sigma * (rng.standard_normal((P, R)) + 1j * rng.standard_normal((P, R)))
    return Dwell(x.astype(np.complex64), lfm_chirp(taps, bw_over_fs), rbins, fbins)


def make_dwell_batch(batch, P=128, R=4096, taps=256, n_targets=5, snr_db=-10.0, seed=20261203):
    """batch dwells of one shape sharing one template, seeds seed, seed+1, ..."""
    cubes = np.empty((batch, P, R), np.complex64)
    truth = []
    tmpl = None
    for b in range(batch):
        d = make_dwell(P, R, taps, n_targets, snr_db, seed + b)
        cubes[b] = d.x
        tmpl = d.template
        truth.append((d.range_bins, d.doppler_bins))
    return cubes, tmpl, truth


def complex_noise_tones(channels, n, tones=3, seed=20261204):
    """Polyphase test signal: CN(0,1) plus `tones` unit tones per channel at
    seeded frequencies in (0, 0.5) cycles/sample."""
    rng = np.random.default_rng(seed)
    t = np.arange(n, dtype=np.float64)
    out = np.empty((channels, n), np.complex64)
    for c in range(channels):
        sig = (rng.standard_normal(n) + 1j * rng.standard_normal(n)) / np.sqrt(2.0)
        for f in rng.uniform(0.0, 0.5, size=tones):
            sig += np.exp(2j * np.pi * f * t)
        out[c] = sig
    return out


@dataclass
class UnevenBatch:
    """Dwells of mixed shape packed back to back in one flat buffer."""
    shapes: list            # [(P, R), ...]
    taps: list              # template length per dwell
    templates: list         # complex64 numpy templates
    truth: list             # [(range_bins, doppler_bins), ...]
    x: object               # flat complex64 torch tensor on the device
    offsets: list           # element offset of each dwell in x


def uneven_shapes(n, seed, pulses=(32, 64, 128, 256), r_lo=1024, r_hi=16384, r_step=256,
                  t_lo=64, t_hi=1024):
    """Seeded shapes of the uneven batch: P uniform over `pulses`, R uniform
    over r_lo..r_hi in steps of r_step, T uniform in [t_lo, min(t_hi, R)]."""
    rng = np.random.default_rng(seed)
    out = []
    for _ in range(n):
        P = int(rng.choice(pulses))
        R = int(rng.integers(r_lo // r_step, r_hi // r_step + 1)) * r_step
        T = int(rng.integers(t_lo, min(t_hi, R) + 1))
        out.append((P, R, T))
    return out


def make_uneven_batch_device(specs, seed=20261206, snr_db=-10.0, targets=(1, 8), device="cuda",
                             amplitude=1.0, bw_over_fs=0.5, margin=10):
    """Dwells for specs [(P, R, T), ...] generated on the device (torch):
    complex Gaussian noise at the given per-sample SNR plus 1-8 seeded point
    targets per dwell (LFM echo at range bin d, Doppler tone at bin f)."""
    import torch
    rng = np.random.default_rng(seed)
    gen = torch.Generator(device=device)
    gen.manual_seed(int(seed))
    total = sum(P * R for P, R, _ in specs)
    x = torch.empty(total, dtype=torch.complex64, device=device)
    noise_power = amplitude ** 2 / (10.0 ** (snr_db / 10.0))
    sigma = float(np.sqrt(noise_power / 2.0))
    offsets, truth, templates = [], [], []
    off = 0
    for P, R, T in specs:
        view = x[off:off + P * R].view(P, R)
        re = torch.randn((P, R), generator=gen, device=device) * sigma
        im = torch.randn((P, R), generator=gen, device=device) * sigma
        view.copy_(torch.complex(re, im))
        s = lfm_chirp(T, bw_over_fs)
        k = int(rng.integers(targets[0], targets[1] + 1))
        hi_r = max(margin + 1, R - margin - T)
        rb = sorted(int(v) for v in rng.choice(np.arange(margin, hi_r), size=k, replace=hi_r - margin < k))
        hi_f = max(margin + 1, P - margin)
        fb = [int(v) for v in rng.integers(margin, hi_f, size=k)]
        p = torch.arange(P, device=device, dtype=torch.float64)[:, None]
        sd = torch.from_numpy(s.astype(np.complex128)).to(device)
        for d, f in zip(rb, fb):
            m = min(T, R - d)
            tone = torch.exp(2j * np.pi * f * p / P)
            view[:, d:d + m] += (amplitude * sd[None, :m] * tone).to(torch.complex64)
        offsets.append(off)
        truth.append((rb, fb))
        templates.append(s)
        off += P * R
    return UnevenBatch([(P, R) for P, R, _ in specs], [T for _, _, T in specs], templates, truth, x, offsets)
