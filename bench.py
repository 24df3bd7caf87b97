"""Radar dwell throughput on one MI355X.

Workload (the reference's plain single-dwell configuration, run in batches):
complex64 dwells of 128 pulses x 4096 range samples, a 256-tap LFM template
(BW/fs = 0.5) with a Hamming window and normalisation, a 128-point Doppler FFT
with a Hann window, |.|^2, and 2-D CA-CFAR with guard 2x2, reference 8x8,
pfa = 1e-6.  Inputs are seeded synthetic dwells (5 point targets at -10 dB
per-sample SNR) resident in HBM before timing starts.

A step is one pass of the fused pipeline over a batch of B dwells.  The JSON
line reports whole-job dwells/s of the batched pipeline (`value`), the plain
per-dwell drop-in (one operator call per step, as the reference's Task1 chain
makes them) beside it, p50/p99 latency, the roofline of the dominant kernel
measured with events on the launch stream, and the CPU oracle timed on the
host.

    python bench.py [--steps K] [--warmup W] [--batch B] [--gpus 1]
"""
from __future__ import annotations

import argparse
import ctypes
import json
import os
import sys
import time

import numpy as np

P, R, TAPS = 128, 4096, 256
GUARD, REF, PFA = (2, 2), (8, 8), 1e-6
HBM_PEAK = 8.0e12


def parse():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=200)
    ap.add_argument("--warmup", type=int, default=50)
    ap.add_argument("--batch", type=int, default=64)
    ap.add_argument("--cfar-mode", type=int, default=1)
    ap.add_argument("--plain-dwells", type=int, default=40)
    ap.add_argument("--no-cpu", action="store_true")
    ap.add_argument("--workload", choices=["dwell", "filters", "batched", "serve", "split", "task1", "probe"],
                    default="dwell")
    ap.add_argument("--dwells", type=int, default=256)
    ap.add_argument("--no-tune", action="store_true")
    ap.add_argument("--lanes", type=int, default=1,
                    help="streams that take consecutive batches in turn (own workspace and outputs each); "
                         "one lane runs each batch's two launches back to back, so the compressed cube is "
                         "still in the memory-side cache when the Doppler+CFAR launch reads it")
    ap.add_argument("--compare-lanes", type=int, default=3,
                    help="also time the dwell pipeline over this many lanes for comparison (0: skip)")
    ap.add_argument("--inputs", type=int, default=4,
                    help="distinct input batches the timed steps rotate through (more bytes than the "
                         "memory-side cache holds, so no step reads its input from cache)")
    ap.add_argument("--no-others", action="store_true", help="skip the short filter / split-BF16 measurements")
    ap.add_argument("--parse-counters", nargs="+", metavar="CSV_OR_OUT",
                    help="counter CSVs of rocprofv3 runs of --workload probe, then the output JSON")
    return ap.parse_args()


def percentile(v, q):
    return float(np.percentile(np.asarray(v), q))


def bench_filters(args):
    """The polyphase configurations: 16 channels x 2^20 complex64 (noise plus
    3 tones), resample_poly up=3 down=2 with a 121-tap Kaiser(beta=5) FIR as
    custom taps; upfirdn up=1 down=4 with a 255-tap lowpass; real firfilter
    on 2^24 uniform[-1,1] samples with the same 255 taps; decimate by 4 of
    those samples with its designed 81-tap filter (designed per call and as
    resident custom taps); channelize_poly of them into 16 and 64 channels
    (32 taps per phase); their direct correlate with the 255 taps; and
    fftconvolve ('same') of the 16 complex channels with a 1025-tap lowpass;
    correlate2d ('same', symm) of a 2048 x 2048 float32 image with 7 x 7 taps;
    the N-D fftconvolve of that image with a 63 x 63 kernel, the N-D call's
    rows form on the channels, and the N-D correlate of the image with a
    33 x 33 kernel.
    Reports per-call time, algorithmic bytes, the HBM fraction and the
    direct-form flop rate for each."""
    import torch
    from scipy import signal
    from zkx_dev_amd import convolution, filtering, synth

    dev = torch.device("cuda:0")
    x = torch.from_numpy(synth.complex_noise_tones(16, 1 << 20)).to(dev)
    kaiser = signal.firwin(121, 1.0 / 3.0, window=("kaiser", 5.0)).astype(np.float32)
    lp255 = signal.firwin(255, 0.25).astype(np.float32)
    xr = torch.from_numpy(np.random.default_rng(7).uniform(-1, 1, 1 << 24).astype(np.float32)).to(dev)
    h_k, h_lp = torch.from_numpy(kaiser).to(dev), torch.from_numpy(lp255).to(dev)
    h_dec = torch.from_numpy(filtering.design_decimate_filter(4).astype(np.float32)).to(dev)
    h_ch16 = torch.from_numpy(signal.firwin(512, 1.0 / 16).astype(np.float32)).to(dev)
    h_ch64 = torch.from_numpy(signal.firwin(2048, 1.0 / 64).astype(np.float32)).to(dev)
    h_1025 = torch.from_numpy(signal.firwin(1025, 0.1).astype(np.complex64)).to(dev)
    img = torch.from_numpy(np.random.default_rng(8).standard_normal((2048, 2048)).astype(np.float32)).to(dev)
    k7 = torch.from_numpy(np.random.default_rng(9).standard_normal((7, 7)).astype(np.float32)).to(dev)
    k63 = torch.from_numpy(np.random.default_rng(10).standard_normal((63, 63)).astype(np.float32)).to(dev)
    k33 = torch.from_numpy(np.random.default_rng(11).standard_normal((33, 33)).astype(np.float32)).to(dev)
    h_1025_row = h_1025.reshape(1, -1).contiguous()
    stream = torch.cuda.current_stream()
    out = {}
    cases = {
        "resample_poly_3_2_121": (lambda: filtering.resample_poly(x, 3, 2, axis=1, coefficients=h_k),
                                  16 * (1 << 20) * 8 + 16 * 1572864 * 8 + 121 * 4,
                                  16 * 1572864 * (121 / 3) * 4),
        "upfirdn_1_4_255": (lambda: filtering.upfirdn(h_lp, x, 1, 4, axis=1),
                            16 * (1 << 20) * 8 + 16 * 262208 * 8 + 255 * 4, 16 * 262208 * 255 * 4),
        "firfilter_255_f32": (lambda: filtering.firfilter(h_lp, xr),
                              (1 << 24) * 8 + 255 * 4, (1 << 24) * 255 * 2),
        # decimate by 4 with its designed 81-tap filter (the design and its
        # upload are part of the call, as in the reference), FP64 result
        "decimate_4_f32": (lambda: filtering.decimate(xr, 4),
                           (1 << 24) * 4 + (1 << 22) * 8 + 81 * 4, (1 << 22) * 81 * 2),
        "decimate_4_f32_custom_taps": (lambda: filtering.decimate(xr, 4, h_dec),
                                       (1 << 24) * 4 + (1 << 22) * 4 + 81 * 4, (1 << 22) * 81 * 2),
        # polyphase channelizer of the same 2^24 samples: 16 channels with 32
        # taps per phase (fused form) and 64 channels (FIR + row FFT + transpose)
        "channelize_poly_16ch_32taps": (lambda: filtering.channelize_poly(xr, h_ch16, 16),
                                        (1 << 24) * 4 + (1 << 24) * 8 + 512 * 4, (1 << 24) * 32 * 2),
        "channelize_poly_64ch_32taps": (lambda: filtering.channelize_poly(xr, h_ch64, 64),
                                        (1 << 24) * 4 + (1 << 24) * 8 + 2048 * 4, (1 << 24) * 32 * 2),
        # direct 1-D correlate of the 2^24 samples with the 255-tap lowpass ('same')
        "correlate_direct_same_255": (lambda: convolution.correlate(xr, h_lp, "same"),
                                      (1 << 24) * 8 + 255 * 4, (1 << 24) * 255 * 2),
        # overlap-save fftconvolve of 16 x 2^20 complex64 rows with 1025 taps
        # ('same'); flops counted as the direct form's complex MACs
        "fftconvolve_same_16x2^20_1025": (lambda: filtering.fftconvolve(x, h_1025, "same"),
                                          16 * (1 << 20) * 16 + 1025 * 8, 16 * (1 << 20) * 1025 * 8),
        # direct 2-D correlate of a 2048 x 2048 image with a 7 x 7 kernel
        # ('same', mirrored boundary)
        "correlate2d_same_symm_2048sq_7x7": (lambda: convolution.correlate2d(img, k7, "same", "symm"),
                                             2048 * 2048 * 8 + 49 * 4, 2048 * 2048 * 49 * 2),
        # N-D fftconvolve of the image with a 63 x 63 kernel over both axes
        # ('same', real result; padded to 2048 + 62 -> 4096 per axis); flops
        # counted as the direct form's MACs
        "fftconvolve_nd_same_2048sq_63x63": (lambda: filtering.fftconvolve(img, k63, "same"),
                                             2048 * 2048 * 8 + 63 * 63 * 4, 2048 * 2048 * 63 * 63 * 2),
        # the 16 complex channels against one 1025-tap row along axis 1 of a
        # [1, 1025] kernel ('full'): the overlap-save rows form of the N-D call
        "fftconvolve_nd_rows_16x2^20_1x1025": (lambda: filtering.fftconvolve(x, h_1025_row, "full", [1]),
                                               16 * (1 << 20) * 8 + 16 * ((1 << 20) + 1024) * 8 + 1025 * 8,
                                               16 * (1 << 20) * 1025 * 8),
        # N-D correlate by FFT of the image with a 33 x 33 kernel ('same')
        "correlate_nd_same_2048sq_33x33": (lambda: convolution.correlate(img, k33, "same"),
                                           2048 * 2048 * 8 + 33 * 33 * 4, 2048 * 2048 * 33 * 33 * 2),
    }
    for name, (fn, nbytes, flops) in cases.items():
        for _ in range(args.warmup):
            fn()
        torch.cuda.synchronize()
        a, b = torch.cuda.Event(enable_timing=True), torch.cuda.Event(enable_timing=True)
        a.record(stream)
        for _ in range(args.steps):
            fn()
        b.record(stream)
        torch.cuda.synchronize()
        t = a.elapsed_time(b) / args.steps * 1e-3
        out[name] = {"us": t * 1e6, "alg_bytes": nbytes, "hbm_frac": nbytes / t / HBM_PEAK,
                     "gflops": flops / t / 1e9}
    print(json.dumps({"metric": "polyphase filtering call time (us) at the polyphase configurations",
                      "workload": "filters", "results": out}))


def batched_summary(dwells, steps, warmup, tune=True):
    """The uneven batch configuration: `dwells` dwells with P uniform over
    {32, 64, 128, 256}, R uniform over 1024..16384 in steps of 256, LFM
    templates of 64..1024 taps (Hamming, normalised), 1-8 targets per dwell at
    -10 dB, Doppler Hann, CA-CFAR g2x2 r8x8 pfa 1e-6 (FP64 window sums), all
    in one engine run: two persistent launches from a HIP graph, FFT lengths
    and grids from the autotuner.  A step is one run over all dwells."""
    import torch
    from zkx_dev_amd import engine, synth

    dev = torch.device("cuda:0")
    specs = synth.uneven_shapes(dwells, seed=20261205)
    ub = synth.make_uneven_batch_device(specs, seed=20261206)
    eng = engine.DwellEngine(pc_window="hamming", normalize=True, doppler_window="hann", guard=GUARD,
                             reference=REF, pfa=PFA, lanes=1)
    tids = [eng.add_template(torch.from_numpy(t)) for t in ub.templates]
    descs, total = engine.make_descs(ub.shapes, tids)
    thr = torch.empty(total, dtype=torch.float32, device=dev)
    det = torch.empty(total, dtype=torch.uint8, device=dev)
    stream = torch.cuda.current_stream()
    t0 = time.perf_counter()
    if tune:
        eng.autotune(ub.x, descs, thr, det)
    tune_s = time.perf_counter() - t0
    for _ in range(warmup):
        eng.run(ub.x, descs, thr, det)
    torch.cuda.synchronize()
    evs = [(torch.cuda.Event(enable_timing=True), torch.cuda.Event(enable_timing=True)) for _ in range(steps)]
    for a, b in evs:
        a.record(stream)
        eng.run(ub.x, descs, thr, det)
        b.record(stream)
    torch.cuda.synchronize()
    lat = [a.elapsed_time(b) for a, b in evs]
    step_ms = float(np.mean(lat))
    n = len(specs)
    samples = sum(P * R for P, R in ub.shapes)
    alg = sum(13 * P * R + 8 * T for (P, R), T in zip(ub.shapes, ub.taps))
    # one dwell per engine call, the unbatched form of the same kernels
    one_lat = []
    for i in range(min(n, 64)):
        one = descs[i:i + 1].copy()
        a, b = torch.cuda.Event(enable_timing=True), torch.cuda.Event(enable_timing=True)
        a.record(stream)
        eng.run(ub.x, one, thr, det)
        b.record(stream)
        torch.cuda.synchronize()
        one_lat.append(a.elapsed_time(b))
    one_rate = 1e3 / float(np.mean(one_lat))
    value = n / (step_ms * 1e-3)
    out = {"value": value, "unit": "dwells/s", "steps": steps, "warmup": warmup, "ms_per_step": step_ms,
           "msamples_per_s": samples / (step_ms * 1e-3) / 1e6,
           "pipeline_hbm_frac": alg / (step_ms * 1e-3) / HBM_PEAK,
           "alg_bytes_per_step": alg, "samples_per_step": samples,
           "latency_ms": {"step_p50": percentile(lat, 50), "step_p99": percentile(lat, 99)},
           "one_dwell_per_call_dwells_per_s": one_rate, "speedup_vs_one_per_call": value / one_rate,
           "autotune_s": tune_s, "tuning_lines": len(eng.tuning().splitlines()),
           "workload": f"{n} uneven dwells, P in {{32..256}}, R in 1024..16384 step 256, "
                       "T 64..1024, Doppler Hann, CA-CFAR g2x2 r8x8 pfa 1e-6; HIP graph of two persistent kernels"}
    eng.close()
    del ub, thr, det
    return out


def bench_batched(args):
    """The uneven batch configuration as its own JSON line (batched_summary)."""
    r = batched_summary(args.dwells, args.steps, args.warmup, tune=not args.no_tune)
    print(json.dumps({
        "metric": "range-Doppler dwells/s + Msamples/s (HBM roofline %) and p99 latency",
        "workload": "batched uneven dwells", "n_gpus": 1, **{k: v for k, v in r.items() if k != "workload"},
        "dtype": "fp32", "data": "synthetic (seeded shapes, targets and noise, generated on the device)",
        "config": {"workload": r["workload"], "launch": "HIP graph of two persistent kernels"}}))


def serve_summary(rates=(1000, 5000, 20000), seconds=1.0,
                  sweep_rates=(20000, 50000, 80000, 100000, 120000, 160000, 200000, 250000), sweep_seconds=0.5,
                  plain_n=40):
    """The serving configuration: Poisson arrivals at 1k, 5k and 20k
    requests/s of dwells drawn equiprobably from 64x2048, 128x4096 and
    256x8192 complex64 (resident in a device pool), 4 resident chirp templates
    of 128..512 taps, admission under a 64 GB HBM budget, continuous batching
    on 3 lanes.  Beside it: a sweep of 128x4096-only load to find the rate the
    server sustains at p99 <= 2 ms, and the plain per-dwell drop-in (three
    operator calls per dwell) on the same shapes."""
    import torch
    from zkx_dev_amd import engine, radar, serve, synth

    dev = torch.device("cuda:0")
    shapes = [(64, 2048), (128, 4096), (256, 8192)]
    taps = [128, 256, 384, 512]
    per_shape = 8
    specs = [(p, r, taps[(i + k) % 4]) for i, (p, r) in enumerate(shapes) for k in range(per_shape)]
    ub = synth.make_uneven_batch_device(specs, seed=20261207)
    eng = engine.DwellEngine(guard=GUARD, reference=REF, pfa=PFA, lanes=3)
    tids = [eng.add_template(torch.from_numpy(synth.lfm_chirp(t))) for t in taps]
    srv = serve.Server(eng, ub.x, budget_bytes=64 << 30, lanes=3, max_batch=64)
    rng = np.random.default_rng(20261208)

    def trace(rate, n, shape_ids):
        arr = np.cumsum(rng.exponential(1e6 / rate, n))
        sh = rng.choice(shape_ids, n)
        k = rng.integers(0, per_shape, n)
        idx = [s_ * per_shape + kk for s_, kk in zip(sh, k)]
        return (arr, [ub.offsets[i] for i in idx], [specs[i][0] for i in idx], [specs[i][1] for i in idx],
                list(rng.integers(0, 4, n)))

    def run(rate, seconds, shape_ids):
        n = max(200, int(rate * seconds))
        arr, off, pp, rr, tt = trace(rate, n, shape_ids)
        lat = srv.replay(arr, off, pp, rr, [tids[t_] for t_ in tt])
        makespan = float(np.max(arr + lat)) * 1e-6
        return {"offered_per_s": rate, "requests": n, "achieved_per_s": n / makespan,
                "p50_ms": percentile(lat, 50) * 1e-3, "p99_ms": percentile(lat, 99) * 1e-3,
                "msamples_per_s": sum(p_ * r_ for p_, r_ in zip(pp, rr)) / makespan / 1e6}

    run(5000, 0.1, [0, 1, 2])  # warm-up: graphs and workspaces on every lane
    mixed = {f"{r // 1000}k": run(r, seconds, [0, 1, 2]) for r in rates}
    sweep = {f"{r // 1000}k": run(r, sweep_seconds, [1]) for r in sweep_rates}
    ok = [v["achieved_per_s"] for v in sweep.values() if v["p99_ms"] <= 2.0
          and v["achieved_per_s"] >= 0.95 * v["offered_per_s"]]
    sustained = max(ok) if ok else 0.0
    stats = srv.stats()
    srv.close()

    # plain per-dwell drop-in on the same shapes: the same baseline function
    # as the dwell bench (three operator calls per dwell)
    plain = {}
    for si, (p, r) in enumerate(shapes):
        i0 = si * per_shape
        xs = [ub.x[ub.offsets[i0 + k]:ub.offsets[i0 + k] + p * r].view(p, r) for k in range(per_shape)]
        tm = torch.from_numpy(synth.lfm_chirp(taps[0])).to(dev)
        plain[f"{p}x{r}"], _ = plain_dropin(radar, xs, tm, plain_n)
    plain_mixed = 1.0 / np.mean([1.0 / v for v in plain.values()])
    out = {"mixed_shapes": mixed, "sweep_128x4096": sweep,
           "sustained_128x4096_at_p99_le_2ms": sustained,
           "plain_dropin_dwells_per_s": plain, "plain_dropin_mixed_dwells_per_s": plain_mixed,
           "sustained_vs_plain_128x4096": sustained / plain["128x4096"] if plain.get("128x4096") else None,
           "server_stats": stats}
    eng.close()
    del ub
    return out


def bench_serve(args):
    """The serving configuration as its own JSON line (serve_summary)."""
    r = serve_summary()
    print(json.dumps({
        "metric": "range-Doppler dwells/s + Msamples/s (HBM roofline %) and p99 latency",
        "workload": "serving", "unit": "dwells/s", "n_gpus": 1, **r, "dtype": "fp32",
        "data": "synthetic dwells resident in a device pool (seeded targets and noise)",
        "config": {"workload": "Poisson dwell requests, shapes 64x2048 / 128x4096 / 256x8192, 4 templates "
                               "128..512 taps, 64 GB admission budget, 3 lanes, batches <= 64 dwells"}}))


def bench_split(args):
    """The split-precision configuration: 16 channels x 512 pulses x 8192 range
    complex64 (pulse-compressed data: noise plus seeded Doppler tones), the
    512-point Doppler DFT as a split-BF16 MFMA contraction (FP32 accumulate)
    beside the FP32 Stockham FFT, with the error of each against FP64 on one
    channel.  Algorithmic work of the contraction: 8 P^2 R C flops."""
    import torch
    from oracle import oracle as orc
    from zkx_dev_amd import radar

    C, Pp, Rr = 16, 512, 8192
    dev = torch.device("cuda:0")
    g = torch.Generator(device=dev)
    g.manual_seed(20261209)
    x = torch.complex(torch.randn((C, Pp, Rr), generator=g, device=dev),
                      torch.randn((C, Pp, Rr), generator=g, device=dev)) * (0.5 ** 0.5)
    p = torch.arange(Pp, device=dev, dtype=torch.float32)[:, None]
    for c in range(C):
        for f in (17 + c, 200 + 3 * c, 400 - c):
            x[c, :, (f * 37) % Rr] += torch.exp(2j * np.pi * f * p / Pp)[:, 0]
    y = torch.empty_like(x)
    pw = torch.empty((C, Pp, Rr), dtype=torch.float32, device=dev)
    stream = torch.cuda.current_stream()

    def timed(fn, reps):
        fn()
        torch.cuda.synchronize()
        a, b = torch.cuda.Event(enable_timing=True), torch.cuda.Event(enable_timing=True)
        a.record(stream)
        for _ in range(reps):
            fn()
        b.record(stream)
        torch.cuda.synchronize()
        return a.elapsed_time(b) / reps * 1e-3

    reps = max(3, args.steps // 5)
    t3 = timed(lambda: radar.doppler_dft_mfma(x, window="hann", terms=3, out=y, power=pw), reps)
    t1 = timed(lambda: radar.doppler_dft_mfma(x, window="hann", terms=1, out=y, power=pw), reps)
    fy = torch.empty((Pp, Rr), dtype=torch.complex64, device=dev)

    def fft_all():
        for c in range(C):
            radar.pulse_doppler(x[c], window="hann", widen=False, out=fy)
    tf = timed(fft_all, reps)
    # error on channel 0 against FP64
    ref = orc.pulse_doppler_fft64(x[0].cpu().numpy(), window="hann")
    e3 = orc.rel_l2(radar.doppler_dft_mfma(x[:1], window="hann", terms=3)[0].cpu().numpy(), ref)
    e1 = orc.rel_l2(radar.doppler_dft_mfma(x[:1], window="hann", terms=1)[0].cpu().numpy(), ref)
    radar.pulse_doppler(x[0], window="hann", widen=False, out=fy)
    ef = orc.rel_l2(fy.cpu().numpy(), ref)
    flops = 8.0 * Pp * Pp * Rr * C
    peak = 2.5e15
    print(json.dumps({
        "metric": "split-BF16 Doppler DFT time and error vs FP64", "workload": "split",
        "config": {"workload": "16 ch x 512 pulses x 8192 range complex64, 512-pt Doppler DFT, Hann"},
        "dtype": "bf16x3 (split) with fp32 accumulate",
        "mfma_split_ms": t3 * 1e3, "mfma_bf16_ms": t1 * 1e3, "fft_fp32_ms": tf * 1e3,
        "roofline": {"bound": "mfma", "achieved": flops / t3 / 1e12, "peak": peak / 1e12, "unit": "TFLOP/s",
                     "frac": flops / t3 / peak, "issued_tflops": 3 * flops / t3 / 1e12,
                     "algorithmic_flops": flops},
        "rel_l2_vs_fp64": {"split_bf16": e3, "bf16": e1, "fft_fp32": ef},
        "cube_bytes": C * Pp * Rr * 8}))


def bench_task1(args):
    """The reference's Task1 chain around the radar path: the echo source
    (step 1) at the dwell shape 128 x 4096 with a 256-sample LFM, the whole
    five-step Task1 dwell at its largest profile (128 x 512, 128-sample LFM),
    and the ambiguity function (step 5) at that profile and for a 4096-sample
    waveform (the largest 2-D map, 8191 x 8192).  Per-call device time with
    events on the launch stream; algorithmic bytes are the outputs written
    (the inputs are a few KB and stay in L2)."""
    import torch
    from zkx_dev_amd import radar, task1

    dev = torch.device("cuda:0")
    stream = torch.cuda.current_stream()

    def timed(fn, reps):
        for _ in range(max(3, args.warmup // 10)):
            fn()
        torch.cuda.synchronize()
        a, b = torch.cuda.Event(enable_timing=True), torch.cuda.Event(enable_timing=True)
        a.record(stream)
        for _ in range(reps):
            fn()
        b.record(stream)
        torch.cuda.synchronize()
        return a.elapsed_time(b) / reps * 1e-3

    reps = max(10, args.steps // 2)
    s11 = dict(num_pulses=128, samples_per_pulse=512, pulse_samples=128, sample_rate_hz=2.56e6, prf_hz=8000,
               bandwidth_hz=6.4e5, carrier_frequency_hz=1e10, target_delay_samples=171, doppler_bin=16,
               target_amplitude=1.5, noise_std=0.035, noise_seed=20260720, pfa=1e-3, cfar_guard_doppler=1,
               cfar_guard_range=2, cfar_reference_doppler=2, cfar_reference_range=6, ambiguity_nfreq=256)
    big = dict(s11, samples_per_pulse=R, pulse_samples=TAPS, target_delay_samples=1000, ambiguity_nfreq=512)
    cfg_big = task1.Task1Config.from_mapping(big)
    cfg_s11 = task1.Task1Config.from_mapping(s11)
    from zkx_dev_amd._lib import check, lib
    res = {}
    # the kernels through the C ABI into preallocated outputs (device time)
    n = cfg_big.num_pulses * cfg_big.samples_per_pulse
    bufs = [torch.empty(cfg_big.pulse_samples, dtype=torch.complex64, device=dev)] + \
        [torch.empty(n, dtype=torch.complex64, device=dev) for _ in range(3)]
    params = task1._EchoParams(cfg_big.num_pulses, cfg_big.samples_per_pulse, cfg_big.pulse_samples,
                               cfg_big.target_delay_samples, cfg_big.sample_rate_hz, cfg_big.prf_hz,
                               cfg_big.bandwidth_hz, cfg_big.doppler_hz(), cfg_big.target_amplitude,
                               cfg_big.noise_std, cfg_big.noise_seed)
    ptrs = [radar._ptr(b) for b in bufs]
    st_ptr = radar._stream()
    t = timed(lambda: check(lib().zkx_echo_source_c64(ctypes.byref(params), *ptrs, st_ptr)), reps)
    nbytes = 24 * n + 8 * cfg_big.pulse_samples
    res["echo_source_128x4096_all_outputs"] = {"us": t * 1e6, "alg_bytes": nbytes, "hbm_frac": nbytes / t / HBM_PEAK,
                                               "gsamples_per_s": n / t / 1e9}
    t = timed(lambda: check(lib().zkx_echo_source_c64(ctypes.byref(params), None, None, None, ptrs[3], st_ptr)),
              reps)
    res["echo_source_128x4096_echo_only"] = {"us": t * 1e6, "alg_bytes": 8 * n, "hbm_frac": 8 * n / t / HBM_PEAK}
    # the whole Task1 dwell through the Python host layer (launch-bound at this size)
    t = timed(lambda: task1.run(cfg_s11), max(5, reps // 4))
    res["task1_dwell_128x512_steps1to5_python_calls"] = {"us": t * 1e6, "dwells_per_s": 1.0 / t}
    w = task1.echo_source(cfg_s11)["waveform"]
    small = torch.empty((255, 256), dtype=torch.float32, device=dev)
    t = timed(lambda: check(lib().zkx_ambgfun_c64(radar._ptr(w), 128, 128, None, 0, 0, 0, 2.56e6, 0.0, 0,
                                                  radar._ptr(small), st_ptr)), reps)
    cells = 255 * 256
    res["ambgfun_2d_T128"] = {"us": t * 1e6, "cells": cells, "alg_bytes": 4 * cells,
                              "hbm_frac": 4 * cells / t / HBM_PEAK}
    rng = np.random.default_rng(3)
    x = torch.from_numpy((rng.standard_normal(4096) + 1j * rng.standard_normal(4096)).astype(np.complex64)).to(dev)
    out = torch.empty((8191, 8192), dtype=torch.float32, device=dev)

    def amb_big():
        check(lib().zkx_ambgfun_c64(radar._ptr(x), 4096, 4096, None, 0, 0, 0, 1.0, 0.0, 0, radar._ptr(out), st_ptr))
    t = timed(amb_big, reps)
    cells = 8191 * 8192
    flops = 8191 * 5.0 * 8192 * 13
    res["ambgfun_2d_T4096"] = {"us": t * 1e6, "cells": cells, "alg_bytes": 4 * cells,
                               "hbm_frac": 4 * cells / t / HBM_PEAK, "fft_tflops": flops / t / 1e12,
                               "fp32_valu_frac": flops / t / 78.6e12}
    cut_out = torch.empty(8192, dtype=torch.float64, device=dev)
    for name, code in (("delay", 1), ("doppler", 2)):
        t = timed(lambda: check(lib().zkx_ambgfun_c64(radar._ptr(x), 4096, 4096, None, 0, 0, code, 1.0, 0.0, 1,
                                                      radar._ptr(cut_out), st_ptr)), reps)
        res[f"ambgfun_{name}_cut_T4096"] = {"us": t * 1e6}
    print(json.dumps({"metric": "Task1 chain around the radar path: per-call device time", "workload": "task1",
                      "results": res}))


def cpu_info():
    """CPU model and the cores this process may use (for cpu_baseline)."""
    model = "unknown"
    try:
        with open("/proc/cpuinfo") as fh:
            for line in fh:
                if line.startswith("model name"):
                    model = line.split(":", 1)[1].strip()
                    break
    except OSError:
        pass
    try:
        avail = len(os.sched_getaffinity(0))
    except AttributeError:
        avail = os.cpu_count() or 1
    return model, avail, os.cpu_count() or avail


def _cpu_chain_worker(seeds):
    """One process of the all-core CPU baseline: the oracle chain on the given
    seeded dwells; returns (dwells, seconds) of pure CPU work."""
    sys.path.insert(0, os.path.dirname(os.path.abspath(__file__)))
    from oracle import oracle as orc
    from zkx_dev_amd import synth
    t = 0.0
    for sd in seeds:
        d = synth.make_dwell(P, R, TAPS, seed=sd)
        c0 = time.perf_counter()
        pc = orc.pulse_compression(d.x, d.template, window="hamming", normalize=True)
        rd = orc.pulse_doppler(pc.astype(np.complex64), window="hann")
        orc.ca_cfar((np.abs(rd) ** 2).astype(np.float32), GUARD, REF, PFA)
        t += time.perf_counter() - c0
    return len(seeds), t


def cpu_baseline(max_workers=16):
    """The reference's CPU chain (the oracle's restatement of its *_typed_cpu
    functions) on bounded samples of the dwell workload: one core for about
    10 s, then every available core (up to the box's CPU share) on 2 dwells
    each."""
    from concurrent.futures import ProcessPoolExecutor
    import multiprocessing as mp
    model, avail, total = cpu_info()
    done, secs = 0, 0.0
    while done < 3 or (secs < 10.0 and done < 30):
        n, t = _cpu_chain_worker([100 + done])
        done += n
        secs += t
    workers = max(1, min(avail, max_workers))
    seeds = [[1000 + 2 * w, 1001 + 2 * w] for w in range(workers)]
    t0 = time.perf_counter()
    with ProcessPoolExecutor(workers, mp_context=mp.get_context("spawn")) as ex:
        res = list(ex.map(_cpu_chain_worker, seeds))
    wall = time.perf_counter() - t0
    all_dwells = sum(r[0] for r in res)
    return {"value": done / secs, "unit": "dwells/s", "cores": 1, "kind": "port",
            "sample": f"{done} seeded dwells 128x4096 on one core (about 10 s): direct FP64 circular "
                      "correlation (256 taps), direct DFT Doppler, FP32 integral-image CFAR (the "
                      "reference *_typed_cpu chain restated in C; its window is computed once per "
                      "call where the reference evaluates cosf per term, so the port is slightly "
                      "faster than the reference's own code), input generation excluded",
            "seconds": secs, "cpu_model": model, "cpus_available": avail, "cpus_total": total,
            "all_cores": {"value": all_dwells / wall, "unit": "dwells/s", "cores": workers,
                          "sample": f"{all_dwells} dwells, {workers} processes x 2 dwells (wall clock, "
                                    "process start-up included)",
                          "seconds": wall}}


def plain_dropin(radar, x_dwells, template, n, warm=3):
    """The plain per-dwell drop-in: the reference Task1 chain's three operator
    calls per dwell (pulse_compression, Doppler |.|^2, ca_cfar), each call
    allocating its outputs and workspaces as the reference's do, waiting for
    each dwell.  Returns (dwells/s, per-dwell latencies in ms)."""
    import torch

    def once(xd):
        y = radar.pulse_compression(xd, template, normalize=True, window="hamming", widen=False)
        pwr = radar.doppler_power(y, window="hann")
        radar.ca_cfar(pwr, GUARD, REF, PFA)
    for i in range(warm):
        once(x_dwells[i % len(x_dwells)])
    torch.cuda.synchronize()
    lat = []
    t0 = time.perf_counter()
    for i in range(n):
        a = time.perf_counter()
        once(x_dwells[i % len(x_dwells)])
        torch.cuda.synchronize()
        lat.append((time.perf_counter() - a) * 1e3)
    return n / (time.perf_counter() - t0), lat


def other_configs(reps):
    """Short measurements of the other configurations, for the default line:
    the three polyphase calls, the split-BF16 Doppler DFT, the uneven batch
    through the engine and a short serving replay."""
    import torch
    from scipy import signal
    from zkx_dev_amd import filtering, radar, synth

    dev = torch.device("cuda:0")
    stream = torch.cuda.current_stream()

    def timed(fn):
        fn()
        torch.cuda.synchronize()
        a, b = torch.cuda.Event(enable_timing=True), torch.cuda.Event(enable_timing=True)
        a.record(stream)
        for _ in range(reps):
            fn()
        b.record(stream)
        torch.cuda.synchronize()
        return a.elapsed_time(b) / reps * 1e-3

    out = {}
    x = torch.from_numpy(synth.complex_noise_tones(16, 1 << 20)).to(dev)
    xr = torch.from_numpy(np.random.default_rng(7).uniform(-1, 1, 1 << 24).astype(np.float32)).to(dev)
    h_k = torch.from_numpy(signal.firwin(121, 1.0 / 3.0, window=("kaiser", 5.0)).astype(np.float32)).to(dev)
    h_lp = torch.from_numpy(signal.firwin(255, 0.25).astype(np.float32)).to(dev)
    cases = {
        "resample_poly_3_2_121_16x2^20_c64": (lambda: filtering.resample_poly(x, 3, 2, axis=1, coefficients=h_k),
                                              16 * (1 << 20) * 8 + 16 * 1572864 * 8 + 121 * 4),
        "upfirdn_1_4_255_16x2^20_c64": (lambda: filtering.upfirdn(h_lp, x, 1, 4, axis=1),
                                        16 * (1 << 20) * 8 + 16 * 262208 * 8 + 255 * 4),
        "firfilter_255_2^24_f32": (lambda: filtering.firfilter(h_lp, xr), (1 << 24) * 8 + 255 * 4),
    }
    for name, (fn, nbytes) in cases.items():
        t = timed(fn)
        out[name] = {"us": t * 1e6, "alg_bytes": nbytes, "hbm_frac": nbytes / t / HBM_PEAK}
    del x, xr
    # split-BF16 Doppler: 16 ch x 512 pulses x 8192 range, Hann, 3 bf16 products
    C, Pp, Rr = 16, 512, 8192
    g = torch.Generator(device=dev)
    g.manual_seed(20261209)
    cube = torch.complex(torch.randn((C, Pp, Rr), generator=g, device=dev),
                         torch.randn((C, Pp, Rr), generator=g, device=dev))
    y = torch.empty_like(cube)
    t3 = timed(lambda: radar.doppler_dft_mfma(cube, window="hann", terms=3, out=y))
    flops = 8.0 * Pp * Pp * Rr * C
    out["split_bf16_doppler_16x512x8192"] = {"us": t3 * 1e6, "algorithmic_tflops": flops / t3 / 1e12,
                                             "mfma_frac": flops / t3 / 2.5e15,
                                             "hbm_frac": 2 * C * Pp * Rr * 8 / t3 / HBM_PEAK}
    del cube, y
    torch.cuda.synchronize()
    # batching: the 256 uneven dwells through the engine (autotuned, HIP graph)
    b = batched_summary(256, steps=30, warmup=30, tune=True)
    out["batched_uneven_256"] = {k: b[k] for k in ("value", "unit", "ms_per_step", "msamples_per_s",
                                                    "pipeline_hbm_frac", "latency_ms",
                                                    "one_dwell_per_call_dwells_per_s", "autotune_s", "workload")}
    # serving: Poisson arrivals at 20k requests/s of mixed shapes, and the
    # 128x4096-only rate sustained at p99 <= 2 ms from a short sweep
    sv = serve_summary(rates=(20000,), seconds=0.5, sweep_rates=(80000, 120000, 160000, 200000), sweep_seconds=0.3,
                       plain_n=20)
    out["serving_poisson"] = {"mixed_20k": sv["mixed_shapes"]["20k"],
                              "sustained_128x4096_at_p99_le_2ms": sv["sustained_128x4096_at_p99_le_2ms"],
                              "sustained_vs_plain_128x4096": sv["sustained_vs_plain_128x4096"],
                              "plain_dropin_128x4096": sv["plain_dropin_dwells_per_s"].get("128x4096")}
    return out


def counters_db():
    path = os.path.join(os.path.dirname(os.path.abspath(__file__)), "profiles", "traffic.json")
    if not os.path.exists(path):
        return {}
    with open(path) as fh:
        return json.load(fh).get("kernels", {})


def probe(args):
    """The counter workload of profiles/traffic.json (run under rocprofv3
    --pmc, one counter set per run): a 256 MiB device copy (calibration),
    then 3 launches of each measured kernel: the two dwell kernels at batch
    64, the reference-identical CFAR of one batch, and the polyphase calls."""
    import torch
    from scipy import signal
    from zkx_dev_amd import filtering, radar, synth

    dev = torch.device("cuda:0")
    a = torch.empty(COPY_BYTES // 4, dtype=torch.float32, device=dev).uniform_()
    b = torch.empty_like(a)
    for _ in range(3):
        b.copy_(a)
    del a, b
    B = 64
    cubes, tmpl, _ = synth.make_dwell_batch(2, P, R, TAPS, seed=5)
    x = torch.from_numpy(np.concatenate([cubes] * (B // 2))).to(dev)
    pipe = radar.DwellPipeline(torch.from_numpy(tmpl).to(dev), P, R, max_batch=B)
    pipe0 = radar.DwellPipeline(torch.from_numpy(tmpl).to(dev), P, R, cfar_mode=0, max_batch=B)
    pc = torch.empty_like(x)
    for _ in range(3):
        radar.pulse_compression_prepared(x, pipe.spectrum, out=pc)
        radar.doppler_cfar(pc, window="hann")
        pipe0.run(x)
    xs = torch.from_numpy(synth.complex_noise_tones(16, 1 << 20)).to(dev)
    xr = torch.from_numpy(np.random.default_rng(7).uniform(-1, 1, 1 << 24).astype(np.float32)).to(dev)
    h_k = torch.from_numpy(signal.firwin(121, 1.0 / 3.0, window=("kaiser", 5.0)).astype(np.float32)).to(dev)
    h_lp = torch.from_numpy(signal.firwin(255, 0.25).astype(np.float32)).to(dev)
    h_1025 = torch.from_numpy(signal.firwin(1025, 0.1).astype(np.complex64)).to(dev)
    for _ in range(3):
        filtering.resample_poly(xs, 3, 2, axis=1, coefficients=h_k)
        filtering.upfirdn(h_lp, xs, 1, 4, axis=1)
        filtering.firfilter(h_lp, xr)
        filtering.fftconvolve(xs, h_1025, "same")
    torch.cuda.synchronize()
    del xs, xr, x, pc
    # the uneven batch (256 dwells) through the engine, direct launches with
    # the default FFT lengths and grids
    from zkx_dev_amd import engine
    specs = synth.uneven_shapes(256, seed=20261205)
    ub = synth.make_uneven_batch_device(specs, seed=20261206)
    eng = engine.DwellEngine(pc_window="hamming", normalize=True, doppler_window="hann", guard=GUARD,
                             reference=REF, pfa=PFA, lanes=1)
    tids = [eng.add_template(torch.from_numpy(t)) for t in ub.templates]
    descs, total = engine.make_descs(ub.shapes, tids)
    thr = torch.empty(total, dtype=torch.float32, device=dev)
    det = torch.empty(total, dtype=torch.uint8, device=dev)
    for _ in range(3):
        eng.run(ub.x, descs, thr, det)
    torch.cuda.synchronize()
    print("probe done", flush=True)


COPY_BYTES = 256 * 1024 * 1024
# kernels of profiles/traffic.json: name -> (substring of the traced name, algorithmic bytes per launch)
PROBED = {
    "pc_fused_kernel<4096>": ("pc_fused_kernel<4096>", 64 * P * R * 16 + R * 8),
    "doppler_cfar_strip_kernel<128, 2, 2, 8, 8>": ("doppler_cfar_strip_kernel<128", 64 * P * R * 13),
    "doppler_kernel<128, 0>": ("doppler_kernel<128, 0>", 64 * P * R * (8 + 4 + 4)),
    "cfar_rowscan_tile_kernel": ("cfar_rowscan_tile_kernel", 64 * P * R * 8),
    "cfar_threshold_rows_kernel": ("cfar_threshold_rows_kernel", 64 * P * R * 13),
    "polyphase_kernel (resample_poly 3/2, 121 taps)": ("polyphase_kernel", 16 * (1 << 20) * 8 + 16 * 1572864 * 8),
    "fir_seg_kernel<4, float2> (upfirdn 1/4 c64, 255 taps)": ("fir_seg_kernel<4",
                                                               16 * (1 << 20) * 8 + 16 * 262208 * 8),
    "fir_seg_kernel<1, float> (firfilter 2^24 f32, 255 taps)": ("fir_seg_kernel<1", (1 << 24) * 8),
    "ols_kernel (fftconvolve same, 1025 taps)": ("ols_kernel", 16 * (1 << 20) * 16),
    # the engine's kernels on the 256-dwell uneven batch (bytes: see uneven_alg_bytes)
    "pc_engine_kernel (256 uneven dwells)": ("pc_engine_kernel", "pc"),
    "dcs_engine_kernel (256 uneven dwells, P 32..128)": ("dcs_engine_kernel", "dcs"),
    "dcs256_engine_kernel (256 uneven dwells, P 256)": ("dcs256_engine_kernel", "dcs256"),
}


def uneven_alg_bytes(kind):
    """Algorithmic bytes of one engine run over the 256-dwell uneven batch:
    pulse compression 16 B per sample plus the template spectra, the strip
    Doppler+CFAR 13 B per sample of its dwells."""
    from zkx_dev_amd import synth
    specs = synth.uneven_shapes(256, seed=20261205)
    shapes = [(s[0], s[1]) for s in specs]
    if kind == "pc":
        return sum(16 * p * r for p, r in shapes) + sum(8 * s[2] for s in specs)
    if kind == "dcs":
        return sum(13 * p * r for p, r in shapes if 32 <= p <= 128)
    return sum(13 * p * r for p, r in shapes if p == 256)


def parse_counters(paths, out_json):
    """Per kernel of PROBED: mean FETCH_SIZE / WRITE_SIZE (KiB per dispatch,
    scaled by what the same counters read for the calibration copy) and mean
    SQ_INSTS_VALU per dispatch, from rocprofv3 counter_collection CSV files."""
    import csv
    from collections import defaultdict
    vals = defaultdict(lambda: defaultdict(list))  # counter -> kernel name -> per-dispatch values
    for path in paths:
        per = defaultdict(float)
        names = {}
        with open(path, newline="") as fh:
            for r in csv.DictReader(fh):
                key = (r["Counter_Name"], r["Dispatch_Id"])
                per[key] += float(r["Counter_Value"])
                names[key] = r["Kernel_Name"]
        for (cnt, did), v in per.items():
            vals[cnt][names[(cnt, did)]].append(v)

    def mean_for(cnt, needle):
        got = [v for name, vs in vals[cnt].items() if needle in name for v in vs]
        return sum(got) / len(got) if got else None
    # calibration: the 256 MiB copy is the largest copyBuffer dispatch (the
    # host-to-device input copies are smaller)
    def max_for(cnt, needle):
        got = [v for name, vs in vals[cnt].items() if needle in name for v in vs]
        return max(got) if got else None
    cf, cw = max_for("FETCH_SIZE", "copyBuffer"), max_for("WRITE_SIZE", "copyBuffer")
    kf = COPY_BYTES / (cf * 1024.0) if cf else None
    kw = COPY_BYTES / (cw * 1024.0) if cw else None
    out = {"calibration": {"copy_bytes": COPY_BYTES, "copy_fetch_kib": cf, "copy_write_kib": cw,
                           "fetch_factor": kf, "write_factor": kw},
           "kernels": {}}
    for key, (needle, alg) in PROBED.items():
        if isinstance(alg, str):
            alg = uneven_alg_bytes(alg)
        rec = {"batch": 64 if ("dwell" in key or "cfar" in key or "pc_" in key) and "engine" not in key else None}
        fk, wk, va = mean_for("FETCH_SIZE", needle), mean_for("WRITE_SIZE", needle), mean_for("SQ_INSTS_VALU", needle)
        if fk is not None and wk is not None and kf and kw:
            rec.update({"fetch_kib": fk, "write_kib": wk, "hbm_bytes_per_launch": fk * 1024.0 * kf + wk * 1024.0 * kw})
            if alg:
                rec["alg_bytes_per_launch"] = alg
                rec["traffic_over_alg"] = rec["hbm_bytes_per_launch"] / alg
        if va is not None:
            rec["valu_insts_per_launch"] = va
        if len(rec) > 1:
            out["kernels"][key] = rec
    with open(out_json, "w") as fh:
        json.dump(out, fh, indent=1)
    print(json.dumps(out, indent=1))


def main():
    args = parse()
    if args.parse_counters:
        parse_counters(args.parse_counters[:-1], args.parse_counters[-1])
        return
    if args.gpus != 1:
        print(json.dumps({"error": "this build measures one GPU only; run with --gpus 1",
                          "n_gpus": args.gpus}))
        sys.exit(2)
    if args.workload == "probe":
        probe(args)
        return
    if args.workload == "filters":
        bench_filters(args)
        return
    if args.workload == "task1":
        bench_task1(args)
        return
    if args.workload == "batched":
        bench_batched(args)
        return
    if args.workload == "serve":
        bench_serve(args)
        return
    if args.workload == "split":
        sys.path.insert(0, os.path.dirname(os.path.abspath(__file__)))
        bench_split(args)
        return
    import torch
    from zkx_dev_amd import radar, synth

    torch.cuda.init()
    dev = torch.device("cuda:0")
    B = args.batch
    cubes, tmpl, _ = synth.make_dwell_batch(B, P, R, TAPS, seed=20261203)
    x = torch.from_numpy(cubes).to(dev)
    t = torch.from_numpy(tmpl).to(dev)
    del cubes
    # the timed steps rotate through distinct input batches (the first one
    # rotated in phase by a unit complex factor per batch)
    xs = [x] + [x * complex(0.6, 0.8) ** (i + 1) for i in range(max(1, args.inputs) - 1)]
    lanes = max(1, args.lanes)
    pipes = [radar.DwellPipeline(t, P, R, normalize=True, pc_window="hamming", doppler_window="hann",
                                 guard=GUARD, reference=REF, pfa=PFA, cfar_mode=args.cfar_mode, max_batch=B)
             for _ in range(lanes)]
    outs = [(torch.empty((B, P, R), dtype=torch.float32, device=dev),
             torch.empty((B, P, R), dtype=torch.uint8, device=dev)) for _ in range(lanes)]
    thr, det = outs[0]
    stream = torch.cuda.current_stream()
    cells = B * P * R

    # ---- per-kernel timing of the stages (events on the launch stream); this
    # also brings the clocks up before the timed steps ------------------------
    pc_out = torch.empty((B, P, R), dtype=torch.complex64, device=dev)
    reps = max(20, args.steps // 2)

    def timed(fn, n=reps):
        fn()
        torch.cuda.synchronize()
        a, b = torch.cuda.Event(enable_timing=True), torch.cuda.Event(enable_timing=True)
        a.record(stream)
        for _ in range(n):
            fn()
        b.record(stream)
        torch.cuda.synchronize()
        return a.elapsed_time(b) / n * 1e-3  # seconds per launch

    for _ in range(3):
        timed(lambda: pipes[0].run(x, threshold=thr, detections=det))
    t_pc = timed(lambda: radar.pulse_compression_prepared(x, pipes[0].spectrum, out=pc_out))
    stages = {"pc_fused_kernel<4096>": (t_pc, cells * 16 + R * 8)}
    if args.cfar_mode == 1:
        t_dc = timed(lambda: radar.doppler_cfar(pc_out, GUARD, REF, PFA, window="hann", threshold=thr,
                                                detections=det))
        stages["doppler_cfar_strip_kernel<128, 2, 2, 8, 8>"] = (t_dc, cells * (8 + 4 + 1))
    pipe0 = radar.DwellPipeline(t, P, R, normalize=True, pc_window="hamming", doppler_window="hann",
                                guard=GUARD, reference=REF, pfa=PFA, cfar_mode=0, max_batch=B)
    t_mode0 = timed(lambda: pipe0.run(x, threshold=thr, detections=det))

    # ---- the headline CFAR against the reference's own CFAR ------------------
    # (the FP32 integral image, bit-identical to the reference restatement) on
    # the same power maps: masks must agree except cells whose statistic lies
    # within 1e-6 relative of the reference threshold, which are counted
    parity = None
    if args.cfar_mode == 1:
        pwm = torch.empty((B, P, R), dtype=torch.float32, device=dev)
        radar.pulse_compression_prepared(x, pipes[0].spectrum, out=pc_out)
        t1, d1 = radar.doppler_cfar(pc_out, GUARD, REF, PFA, window="hann", power=pwm)
        pipes[0].run(x, threshold=thr, detections=det)
        same = bool(torch.equal(d1, det)) and bool(torch.equal(t1.view(torch.int32), thr.view(torch.int32)))
        differ = excused = near = 0
        drift = 0.0
        for b in range(B):
            t0, d0 = radar.ca_cfar(pwm[b], GUARD, REF, PFA)
            dm = d0 != d1[b]
            nr = (pwm[b] - t0).abs() <= 1e-6 * t0.abs()
            differ += int(dm.sum())
            excused += int((dm & nr).sum())
            near += int(nr.sum())
            live = t0 != 0
            drift = max(drift, float(((t1[b] - t0).abs() / t0.abs().clamp_min(1e-30))[live].max()))
        parity = {"cells": cells, "differ": differ, "differ_within_1e-6": excused,
                  "differ_outside_1e-6": differ - excused, "near_threshold_cells": near,
                  "max_threshold_rel_gap": drift, "pipeline_equals_fused_kernel": same,
                  "reference": "FP32 integral-image CA-CFAR in the reference's order (ca_cfar, bit-identical "
                               "to the reference restatement) on the same power maps"}
        del pwm, t1, d1

    # ---- batched fused pipeline: the timed steps -----------------------------
    # Consecutive batches go to `lanes` streams in turn, each lane with its
    # own pipeline workspace and output buffers (every batch still runs the
    # whole chain).  With one lane (the default) each batch's Doppler+CFAR
    # reads the compressed cube its pulse compression has just written, much
    # of it from the memory-side cache; several lanes instead overlap one
    # batch's Doppler+CFAR with the next batch's pulse compression.
    lane_streams = [torch.cuda.Stream() for _ in range(lanes)] if lanes > 1 else [stream]

    def run_steps(nlanes, steps, ev=None):
        start = torch.cuda.Event()
        start.record(stream)
        for st in lane_streams[:nlanes]:
            st.wait_event(start)
        for s in range(steps):
            k = s % nlanes
            st = lane_streams[k] if nlanes > 1 else stream
            if ev is not None:
                ev[s][0].record(st)
            pipes[k].run(xs[s % len(xs)], threshold=outs[k][0], detections=outs[k][1], stream=st)
            if ev is not None:
                ev[s][1].record(st)
        for st in lane_streams[:nlanes]:
            stream.wait_stream(st)

    def timed_steps(nlanes):
        # throughput from the wall clock around K steps with no per-step
        # events (timing events between launches cost overlap), then the
        # per-batch latency from a separate pass that records them
        run_steps(nlanes, args.warmup)
        torch.cuda.synchronize()
        t0 = time.perf_counter()
        run_steps(nlanes, args.steps)
        torch.cuda.synchronize()
        wall_s = time.perf_counter() - t0
        ev = [(torch.cuda.Event(enable_timing=True), torch.cuda.Event(enable_timing=True))
              for _ in range(args.steps)]
        run_steps(nlanes, args.steps, ev)
        torch.cuda.synchronize()
        return wall_s, [a.elapsed_time(b) for a, b in ev]

    wall, step_ms = timed_steps(lanes)
    dwells_per_s = B * args.steps / wall
    # the other lane arrangement, for comparison (three lanes overlap one
    # batch's Doppler+CFAR with the next batch's pulse compression)
    alt_lanes = args.compare_lanes if args.compare_lanes != lanes else 0
    if alt_lanes > len(lane_streams):
        pipes += [radar.DwellPipeline(t, P, R, normalize=True, pc_window="hamming", doppler_window="hann",
                                      guard=GUARD, reference=REF, pfa=PFA, cfar_mode=args.cfar_mode, max_batch=B)
                  for _ in range(alt_lanes - len(pipes))]
        outs += [(torch.empty((B, P, R), dtype=torch.float32, device=dev),
                  torch.empty((B, P, R), dtype=torch.uint8, device=dev)) for _ in range(alt_lanes - len(outs))]
        lane_streams = [torch.cuda.Stream() for _ in range(alt_lanes)]
    alt_dps = None
    if alt_lanes > 0:
        wall_alt, _ = timed_steps(alt_lanes)
        alt_dps = B * args.steps / wall_alt
    print(f"# batched: {dwells_per_s:.1f} dwells/s over {lanes} lane(s) ({alt_dps} over {alt_lanes}), "
          f"step p50 {percentile(step_ms, 50):.3f} ms", file=sys.stderr, flush=True)

    # ---- plain per-dwell drop-in ---------------------------------------------
    plain_dps, lat = plain_dropin(radar, [x[i] for i in range(min(B, 8))], t, max(5, args.plain_dwells))

    # ---- roofline of the dominant kernel ---------------------------------------
    db = counters_db()
    stage_info = {}
    for name, (t_k, bytes_k) in stages.items():
        rec = db.get(name, {})
        valu = rec.get("valu_insts_per_launch") if rec.get("batch") == B else None
        stage_info[name] = {"us": t_k * 1e6, "alg_bytes": bytes_k, "hbm_frac": bytes_k / t_k / HBM_PEAK,
                            # wave64 VALU instructions x 4 cycles on 1024 SIMDs at 2.4 GHz
                            "valu_issue_frac": valu * 4 / (1024 * 2.4e9 * t_k) if valu else None,
                            "traffic": rec.get("hbm_bytes_per_launch") if rec.get("batch") == B else None}
    name, (t_k, bytes_k) = max(stages.items(), key=lambda kv: kv[1][0])
    achieved = bytes_k / t_k
    roofline = {"bound": "hbm", "kernel": name, "achieved": achieved / 1e9,
                "peak": HBM_PEAK / 1e9, "unit": "GB/s", "frac": achieved / HBM_PEAK,
                "traffic": stage_info[name]["traffic"], "valu_issue_frac": stage_info[name]["valu_issue_frac"],
                "avg_launch_us": t_k * 1e6, "alg_bytes_per_launch": bytes_k, "stages": stage_info}
    pipe_bytes = B * (13 * P * R + 8 * TAPS)
    step_s = wall / args.steps

    # ---- the other configurations, briefly -------------------------------------
    others = None if args.no_others else other_configs(max(5, min(20, args.steps // 10)))

    # ---- CPU baseline on the host -----------------------------------------------
    cpu = None if args.no_cpu else cpu_baseline()

    out = {
        "metric": "range-Doppler dwells/s + Msamples/s (HBM roofline %) and p99 latency",
        "value": dwells_per_s,
        "unit": "dwells/s",
        "n_gpus": 1,
        "steps": args.steps,
        "warmup": args.warmup,
        "ms_per_step": step_s * 1e3,
        "higher_is_better": True,
        "scaling": "weak",
        "vs_baseline": None,
        "dtype": "fp32",
        "data": "synthetic (seeded LFM point targets + complex Gaussian noise, -10 dB SNR)",
        "config": {"workload": "dwell 128x4096 c64, LFM 256 taps Hamming, Doppler 128 Hann, "
                               "CA-CFAR g2x2 r8x8 pfa 1e-6", "batch_dwells": B,
                   "cfar_mode": "reference-identical integral image" if args.cfar_mode == 0
                   else "FP64 windowed, fused with the Doppler FFT",
                   "lanes": lanes, "distinct_input_batches": len(xs)},
        "compare_lanes": {"lanes": alt_lanes, "dwells_per_s": alt_dps} if alt_lanes > 0 else None,
        "msamples_per_s": dwells_per_s * P * R / 1e6,
        "pipeline_hbm_frac": pipe_bytes / step_s / HBM_PEAK,
        "latency_ms": {"batch_p50": percentile(step_ms, 50), "batch_p99": percentile(step_ms, 99),
                       "plain_dwell_p50": percentile(lat, 50), "plain_dwell_p99": percentile(lat, 99)},
        "plain_dropin_dwells_per_s": plain_dps,
        "speedup_vs_plain": dwells_per_s / plain_dps,
        "reference_identical_cfar_dwells_per_s": B / t_mode0,
        "cfar_vs_reference": parity,
        "roofline": roofline,
        "other_configs": others,
        "cpu_baseline": cpu,
    }
    print(json.dumps(out))


if __name__ == "__main__":
    main()
