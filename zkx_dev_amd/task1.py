"""The reference's Task1 radar pipeline on the HIP library.

Task1 (ZKX/Task1/pipeline) runs five steps on one dwell:
  1. echo source: LFM waveform, delayed Doppler-shifted echo of one target,
     uniform hash noise (steps/step1/step1.cu:11-118)
  2. pulse compression, normalised template, no window, nfft = samples per
     pulse (steps/step2/step2.cu:20-36)
  3. pulse-Doppler FFT over the pulses, Hamming window, into a resident
     output (steps/step3/step3.cu:17-35)
  4. power |Y|^2, cfar_alpha_device, 2-D CA-CFAR (steps/step4/step4.cu:35-84)
  5. ambiguity function of the waveform: 2-D, the zero-Doppler cut over
     delay and the zero-delay cut over Doppler (steps/step5/step5.cu:22-57)

`Task1Config` holds the fields of its configuration file
(include/task1/task1_config.h:16-104) with the same validation rules;
`echo_source` and `run` drive the C ABI of libzkx_radar.so.  Every step runs
on the device; nothing here computes samples on the host.
"""
from __future__ import annotations

import ctypes
from dataclasses import dataclass, fields

import torch

from . import radar
from ._lib import check, lib

FIELDS_INT = ("num_pulses", "samples_per_pulse", "pulse_samples", "target_delay_samples", "doppler_bin",
              "noise_seed", "cfar_guard_doppler", "cfar_guard_range", "cfar_reference_doppler",
              "cfar_reference_range", "ambiguity_nfreq")


class _EchoParams(ctypes.Structure):
    """zkx_echo_params (include/zkx_radar.h)."""
    _fields_ = [("num_pulses", ctypes.c_int), ("samples_per_pulse", ctypes.c_int),
                ("pulse_samples", ctypes.c_int), ("target_delay_samples", ctypes.c_int),
                ("sample_rate_hz", ctypes.c_float), ("prf_hz", ctypes.c_float),
                ("bandwidth_hz", ctypes.c_float), ("doppler_hz", ctypes.c_float),
                ("target_amplitude", ctypes.c_float), ("noise_std", ctypes.c_float),
                ("noise_seed", ctypes.c_uint32)]


def _f32(v):
    return float(torch.tensor(float(v), dtype=torch.float32))


def _pow2(n):
    return n > 0 and (n & (n - 1)) == 0


@dataclass
class Task1Config:
    """task1_config.h:16-34.  Float fields are held as FP32 numbers, as the
    reference stores them (carrier_frequency_hz stays double)."""
    num_pulses: int
    samples_per_pulse: int
    pulse_samples: int
    sample_rate_hz: float
    prf_hz: float
    bandwidth_hz: float
    carrier_frequency_hz: float
    target_delay_samples: int
    doppler_bin: int
    target_amplitude: float
    noise_std: float
    noise_seed: int
    pfa: float
    cfar_guard_doppler: int
    cfar_guard_range: int
    cfar_reference_doppler: int
    cfar_reference_range: int
    ambiguity_nfreq: int

    def __post_init__(self):
        for f in fields(self):
            v = getattr(self, f.name)
            if f.name in FIELDS_INT:
                setattr(self, f.name, int(v))
            elif f.name != "carrier_frequency_hz":
                setattr(self, f.name, _f32(v))
            else:
                setattr(self, f.name, float(v))

    @classmethod
    def from_mapping(cls, m):
        missing = [f.name for f in fields(cls) if f.name not in m]
        if missing:
            raise ValueError(f"Task1 configuration is missing {missing}")
        cfg = cls(**{f.name: m[f.name] for f in fields(cls)})
        cfg.validate()
        return cfg

    @classmethod
    def from_conf(cls, path):
        """key=value lines with '#' comments, as the reference's .conf files."""
        m = {}
        with open(path, encoding="utf-8") as fh:
            for line in fh:
                line = line.strip()
                if line and not line.startswith("#"):
                    k, v = line.split("=", 1)
                    m[k.strip()] = float(v) if any(c in v for c in ".eE") else int(v)
        return cls.from_mapping(m)

    def validate(self):
        """task1_config.h:36-68 (the element-count cap aside)."""
        if not (_pow2(self.num_pulses) and _pow2(self.samples_per_pulse)):
            raise ValueError("Task1 FFT dimensions must be powers of two")
        if self.pulse_samples < 1 or self.pulse_samples > self.samples_per_pulse:
            raise ValueError("Task1 pulse_samples must lie within the receive window")
        if self.target_delay_samples < 0 or self.target_delay_samples + self.pulse_samples > self.samples_per_pulse:
            raise ValueError("Task1 target delay moves the pulse out of the receive window")
        if abs(self.doppler_bin) >= self.num_pulses // 2:
            raise ValueError("Task1 doppler_bin is outside the unaliased range")
        if not (self.sample_rate_hz > 0 and self.prf_hz > 0 and self.bandwidth_hz > 0 and
                self.carrier_frequency_hz > 0 and self.target_amplitude > 0 and self.noise_std >= 0):
            raise ValueError("Task1 rate, amplitude or noise parameters violate physical limits")
        if not (0 < self.pfa < 1):
            raise ValueError("Task1 pfa must lie in (0,1)")
        if (self.cfar_guard_doppler < 0 or self.cfar_guard_range < 0 or self.cfar_reference_doppler < 1 or
                self.cfar_reference_range < 1):
            raise ValueError("Task1 CFAR guard or reference cells are invalid")
        nfreq = 1
        while nfreq < 2 * self.pulse_samples - 1:
            nfreq <<= 1
        if self.ambiguity_nfreq != nfreq:
            raise ValueError("Task1 ambiguity_nfreq must equal next_power_of_two(2*pulse_samples-1)")

    def doppler_hz(self):
        """task1_config.h:100-103: doppler_bin * prf_hz / num_pulses in FP32."""
        t = torch.tensor(float(self.doppler_bin), dtype=torch.float32) * torch.tensor(self.prf_hz, dtype=torch.float32)
        return float(t / torch.tensor(float(self.num_pulses), dtype=torch.float32))

    def reference_count(self):
        """Number of CFAR reference cells (step4.cu:41-44)."""
        od = self.cfar_guard_doppler + self.cfar_reference_doppler
        orr = self.cfar_guard_range + self.cfar_reference_range
        return (2 * od + 1) * (2 * orr + 1) - (2 * self.cfar_guard_doppler + 1) * (2 * self.cfar_guard_range + 1)


def echo_source(cfg: Task1Config, *, device=None, stream=None, doppler_hz=None, target_delay_samples=None,
                noise_seed=None):
    """Step 1 on the device.  Returns dict(waveform [T], noiseless, noise,
    echo [P, S]) of complex64 tensors.  The keyword overrides serve the
    reference's perturbation checks (step1_validation.cu:94-108)."""
    dev = torch.device("cuda", torch.cuda.current_device()) if device is None else torch.device(device)
    P, S, T = cfg.num_pulses, cfg.samples_per_pulse, cfg.pulse_samples
    p = _EchoParams(P, S, T, cfg.target_delay_samples if target_delay_samples is None else int(target_delay_samples),
                    cfg.sample_rate_hz, cfg.prf_hz, cfg.bandwidth_hz,
                    cfg.doppler_hz() if doppler_hz is None else _f32(doppler_hz),
                    cfg.target_amplitude, cfg.noise_std,
                    (cfg.noise_seed if noise_seed is None else int(noise_seed)) & 0xFFFFFFFF)
    out = {"waveform": torch.empty(T, dtype=torch.complex64, device=dev)}
    for k in ("noiseless", "noise", "echo"):
        out[k] = torch.empty((P, S), dtype=torch.complex64, device=dev)
    code = lib().zkx_echo_source_c64(ctypes.byref(p), radar._ptr(out["waveform"]), radar._ptr(out["noiseless"]),
                                     radar._ptr(out["noise"]), radar._ptr(out["echo"]), radar._stream(stream))
    if code in (1001, 1002):
        raise ValueError(lib().zkx_last_error_message().decode())
    check(code)
    return out


def run(cfg: Task1Config, *, steps=5, stream=None):
    """Task1 steps 1..`steps` on the device; returns a dict of device tensors
    named after the reference's pipeline state (task1_common.h)."""
    st = echo_source(cfg, stream=stream)
    P, S = cfg.num_pulses, cfg.samples_per_pulse
    if steps >= 2:
        st["compressed"] = radar.pulse_compression(st["echo"], st["waveform"], normalize=True, nfft=S,
                                                   window="none", stream=stream)
    if steps >= 3:
        rd = torch.empty((P, S), dtype=torch.complex64, device=st["echo"].device)
        power = torch.empty((P, S), dtype=torch.float32, device=rd.device)
        radar.pulse_doppler(st["compressed"], nfft=P, window="hamming", out=rd, power=power, stream=stream)
        st["range_doppler"] = rd
        st["power"] = power
    if steps >= 4:
        st["cfar_alpha"] = radar.cfar_alpha_device(cfg.pfa, cfg.reference_count(), stream=stream)
        st["cfar_threshold"], st["detections"] = radar.ca_cfar(
            st["power"], (cfg.cfar_guard_doppler, cfg.cfar_guard_range),
            (cfg.cfar_reference_doppler, cfg.cfar_reference_range), cfg.pfa, stream=stream)
    if steps >= 5:
        w = st["waveform"]
        st["ambiguity_2d"] = radar.ambgfun(w, cfg.sample_rate_hz, cfg.prf_hz, cut="2d", stream=stream)
        st["ambiguity_delay"] = radar.ambgfun(w, cfg.sample_rate_hz, cfg.prf_hz, cut="doppler", cutValue=0.0,
                                              stream=stream)
        st["ambiguity_doppler"] = radar.ambgfun(w, cfg.sample_rate_hz, cfg.prf_hz, cut="delay", cutValue=0.0,
                                                stream=stream)
    return st
