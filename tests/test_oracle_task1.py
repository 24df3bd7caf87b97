"""Pin the oracle to the reference's archived Task1 pipeline outputs.

golden/task1_s11_strong.npz holds the reference's own GPU outputs of its
Task1 pipeline for profile task1_s11_strong (128 pulses x 512 samples, 128-tap
LFM, one target), read from its visualisation export by
golden/extract_task1_evidence.py.  These tests check the oracle against them:
bit for bit where the computation is exact (hash noise, the FP32 integral-image
CA-CFAR on the archived power map, the power map's FMA form) and within the
reference's own accuracy thresholds elsewhere
(ZKX/test_all/config/task_benchmarks/accuracy_thresholds.json, stated per test).
The ambiguity-function oracle is further pinned by an independent FP64 FFT
formulation (cuSignal's) and by its closed-form properties.
"""
import json
import os

import numpy as np
import pytest

from oracle import oracle as orc

HERE = os.path.dirname(os.path.abspath(__file__))
FIX = np.load(os.path.join(HERE, "golden", "task1_s11_strong.npz"))
CFG = json.load(open(os.path.join(HERE, "golden", "task1_s11_strong.json"), encoding="utf-8"))["config"]


def _f32(v):
    return float(np.float32(v))


def _doppler_hz(c, bin_=None):
    b = c["doppler_bin"] if bin_ is None else bin_
    return float(np.float32(b) * np.float32(c["prf_hz"]) / np.float32(c["num_pulses"]))


def _echo(c, delay=None, doppler_hz=None, seed=None):
    return orc.echo_source(c["num_pulses"], c["samples_per_pulse"], c["pulse_samples"], c["sample_rate_hz"],
                           c["prf_hz"], c["bandwidth_hz"],
                           c["target_delay_samples"] if delay is None else delay,
                           _doppler_hz(c) if doppler_hz is None else doppler_hz,
                           c["target_amplitude"], c["noise_std"], c["noise_seed"] if seed is None else seed)


def test_fixture_shapes():
    P, S, T = CFG["num_pulses"], CFG["samples_per_pulse"], CFG["pulse_samples"]
    assert FIX["echo"].shape == (P, S) and FIX["waveform"].shape == (T,)
    assert FIX["ambiguity"].shape == (2 * T - 1, CFG["ambiguity_nfreq"])


def test_echo_noise_bit_exact():
    """The hash noise (step1.cu:11-25) is integer arithmetic plus one rounded
    division: the oracle reproduces the reference GPU's noise bit for bit."""
    _, _, noise, _ = _echo(CFG)
    assert np.array_equal(noise.view(np.uint64), FIX["noise"].view(np.uint64))


def test_echo_waveform_and_echo_within_reference_threshold():
    """accuracy_thresholds.json Task1.step1.*: relative L2 <= 2e-5."""
    w, noiseless, noise, echo = _echo(CFG)
    assert orc.rel_l2(w, FIX["waveform"]) < 2e-5
    assert orc.rel_l2(echo, FIX["echo"]) < 2e-5
    assert orc.rel_l2(noiseless, FIX["echo"] - FIX["noise"]) < 2e-5
    # the target occupies delay .. delay + T - 1 of every pulse, nothing else
    d, T = CFG["target_delay_samples"], CFG["pulse_samples"]
    live = np.abs(noiseless).max(axis=0) > 0
    assert live.sum() == T and live[d] and live[d + T - 1]


def test_echo_perturbations_change_output():
    """step1_comparison.cpp:37-49: delay + 3, Doppler + one bin and seed + 1
    each change the echo (max abs difference > 0.1, > 0.1, > 1e-3)."""
    base = _echo(CFG)[3]
    d3 = _echo(CFG, delay=CFG["target_delay_samples"] + 3)[3]
    fd = _doppler_hz(CFG) + _f32(CFG["prf_hz"]) / CFG["num_pulses"]
    dd = _echo(CFG, doppler_hz=fd)[3]
    ds = _echo(CFG, seed=CFG["noise_seed"] + 1)[3]
    assert np.abs(d3 - base).max() > 0.1
    assert np.abs(dd - base).max() > 0.1
    assert np.abs(ds - base).max() > 1e-3


def test_echo_noise_free_case_is_seed_invariant():
    c = dict(CFG, noise_std=0.0)
    a, b = _echo(c)[3], _echo(c, seed=c["noise_seed"] + 1)[3]
    assert np.array_equal(a, b)
    assert not np.any(_echo(c)[2])


def test_uniform_noise_range_and_moments():
    """u = ((h & 0xffff)/32767.5 - 1) * sigma lies in [-sigma, sigma]; over many
    indices its mean is ~0 and its variance ~sigma^2/3."""
    u = orc.echo_uniform_noise(np.arange(1 << 18, dtype=np.uint64), 12345, 1.0)
    assert u.min() >= -1.0 and u.max() <= 1.0
    assert abs(u.mean()) < 5e-3 and abs(u.var() - 1 / 3) < 5e-3


def test_pulse_compression_matches_reference_gpu():
    """Task1 step 2 (normalize, no window, nfft = S) on the archived echo vs the
    reference GPU's compressed output; threshold 1e-4 (Task1.step2)."""
    y = orc.pulse_compression(FIX["echo"], FIX["waveform"], nfft=CFG["samples_per_pulse"], normalize=True)
    assert orc.rel_l2(y, FIX["compressed"]) < 1e-5


def test_pulse_doppler_matches_reference_gpu():
    """Task1 step 3 (Hamming, nfft = P) on the archived compressed map vs the
    reference GPU's range-Doppler map; threshold 5e-4 (Task1.step3).  The GPU
    path is an FFT with true pi; the faithful CPU form keeps the float pi."""
    exact = orc.pulse_doppler(FIX["compressed"], nfft=CFG["num_pulses"], window="hamming", exact=True)
    faithful = orc.pulse_doppler(FIX["compressed"], nfft=CFG["num_pulses"], window="hamming")
    assert orc.rel_l2(exact, FIX["range_doppler"]) < 1e-6
    assert orc.rel_l2(faithful, FIX["range_doppler"]) < 5e-6


def test_power_is_fused_multiply_add_form():
    """The reference's power map (step4.cu:18) equals fma(re, re, im*im) in
    every cell, and differs from the unfused re*re + im*im in many."""
    rd = FIX["range_doppler"]
    re, im = rd.real.astype(np.float64), rd.imag.astype(np.float64)
    im2 = (rd.imag * rd.imag).astype(np.float64)
    fused = (re * re + im2).astype(np.float32)  # re*re is exact in FP64
    plain = rd.real * rd.real + rd.imag * rd.imag
    assert np.array_equal(fused, FIX["power"])
    assert np.count_nonzero(plain != FIX["power"]) > 1000


@pytest.mark.parametrize("pfa", [0.001, float(np.float32(0.001))])
def test_ca_cfar_bit_identical_to_reference_gpu(pfa):
    """Task1 step 4 on the archived power map: thresholds and detections equal
    the reference GPU's bit for bit (border cells threshold 0, detected)."""
    g, r = (CFG["cfar_guard_doppler"], CFG["cfar_guard_range"]), (CFG["cfar_reference_doppler"],
                                                                  CFG["cfar_reference_range"])
    thr, det = orc.ca_cfar(FIX["power"], g, r, pfa)
    assert np.array_equal(thr.view(np.uint32), FIX["threshold"].view(np.uint32))
    assert np.array_equal(det, FIX["detection"])


def test_ambgfun_2d_matches_reference_gpu():
    """Task1 step 5, 2-D cut of the archived waveform; threshold 2e-4."""
    a = orc.ambgfun(FIX["waveform"], CFG["sample_rate_hz"], CFG["prf_hz"])
    assert a.dtype == np.float32 and a.shape == FIX["ambiguity"].shape
    assert orc.rel_l2(a, FIX["ambiguity"]) < 1e-6


def test_ambgfun_cuts_match_reference_gpu():
    """Step 5's cuts: cut='doppler' at 0 Hz is the archived cut over delay
    (rows values) and cut='delay' at 0 s the cut over Doppler (nfreq values);
    threshold 1e-4 each."""
    over_delay = orc.ambgfun(FIX["waveform"], CFG["sample_rate_hz"], cut="doppler", cut_value=0.0)
    over_doppler = orc.ambgfun(FIX["waveform"], CFG["sample_rate_hz"], cut="delay", cut_value=0.0)
    assert orc.rel_l2(over_delay, FIX["ambiguity_delay"]) < 1e-6
    assert orc.rel_l2(over_doppler, FIX["ambiguity_doppler"]) < 1e-6


def _rand(n, seed):
    rng = np.random.default_rng(seed)
    return (rng.standard_normal(n) + 1j * rng.standard_normal(n)).astype(np.complex64)


@pytest.mark.parametrize("nx,ny,fs,cut,cv", [
    (50, None, 1000.0, "2d", 0.0), (100, None, 2000.0, "2d", 0.0), (50, 50, 1000.0, "2d", 0.0),
    (37, 20, 1000.0, "2d", 0.0), (12, 40, 100.0, "2d", 0.0),
    (50, None, 1000.0, "delay", 0.0), (100, None, 2000.0, "delay", 5.0), (64, None, 1000.0, "delay", 0.0123),
    (50, None, 1000.0, "doppler", 0.0), (100, None, 2000.0, "doppler", 10.0), (33, None, 500.0, "doppler", 71.5),
])
def test_ambgfun_matches_fft_formulation(nx, ny, fs, cut, cv):
    """The direct-sum restatement equals cuSignal's FFT formulation
    (radartools.py:207-234) evaluated with numpy FP64 FFTs, on the shapes of
    cuSignal's own ambgfun tests (test_radartools_cpu.py:196-243) and more."""
    x = _rand(nx, nx)
    y = None if ny is None else _rand(ny, 1000 + ny)
    a = orc.ambgfun(x, fs, y=y, cut=cut, cut_value=cv, output_fp32=False)
    b = orc.ambgfun_fft64(x.astype(np.complex128), fs, y=None if y is None else y.astype(np.complex128),
                          cut=cut, cut_value=cv)
    assert a.shape == b.shape
    assert orc.rel_l2(a, b) < 1e-12


def test_ambgfun_closed_forms():
    """|A| <= 1 everywhere (Cauchy-Schwarz on the normalised waveforms), the
    zero-delay zero-Doppler cell is exactly the peak 1 for y = x, the cuts
    through the origin agree with the 2-D map, and a single impulse has a
    flat unit zero-delay row."""
    x = _rand(40, 3)
    a = orc.ambgfun(x, output_fp32=False)
    rows, nfreq = a.shape
    assert a.max() <= 1 + 1e-12
    assert abs(a[40 - 1, nfreq // 2] - 1.0) < 1e-12 and np.argmax(a) == (40 - 1) * nfreq + nfreq // 2
    over_delay = orc.ambgfun(x, 10.0, cut="doppler", cut_value=0.0)
    over_doppler = orc.ambgfun(x, 10.0, cut="delay", cut_value=0.0)
    assert np.allclose(over_delay, a[:, nfreq // 2], rtol=0, atol=1e-12)
    assert np.allclose(over_doppler, a[40 - 1, :], rtol=0, atol=1e-12)
    imp = np.zeros(8, np.complex64)
    imp[3] = 2 - 1j
    ai = orc.ambgfun(imp, output_fp32=False)
    assert np.allclose(ai[7, :], 1.0, atol=1e-12) and np.allclose(np.delete(ai, 7, axis=0), 0.0, atol=1e-12)


def test_ambgfun_zero_input_is_nan_and_errors():
    assert np.all(np.isnan(orc.ambgfun(np.zeros(4, np.complex64), output_fp32=False)))
    with pytest.raises(ValueError):
        orc.ambgfun(_rand(4, 1), 0.0, cut="delay")
