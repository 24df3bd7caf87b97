"""Collect the reference's archived Task1 pipeline outputs into a committed fixture.

The reference keeps, for its Task1 profile ``task1_s11_strong`` (128 pulses x 512
range samples, 128-sample LFM, one target), the outputs of every pipeline step as
CSV files written with 17 significant digits (its visualisation export), so each
FP32 value is recovered exactly.  This script reads those data files (never any
reference source) and writes ``task1_s11_strong.npz`` (arrays, loadable with
``numpy.load`` without pickle) and ``task1_s11_strong.json`` (the profile).

Files read, under ZKX/Task1/evidence/visualization/:
  task1_selected.conf              the profile (key=value)
  data/task1_step1_waveform.csv    LFM waveform
  data/task1_step1_echo.csv        noise, echo (noiseless = echo - noise)
  data/task1_step2_compressed.csv  pulse compression output
  data/task1_step3_range_doppler.csv  Doppler output
  data/task1_step4_cfar.csv        power, CFAR threshold, detections
  data/task1_step5_ambiguity.csv   2-D ambiguity function
  data/task1_step5_delay_cut.csv   zero-Doppler cut (over delay)
  data/task1_step5_doppler_cut.csv zero-delay cut (over Doppler)

Run in a container where /root/reference exists:
    python tests/golden/extract_task1_evidence.py
"""
import json
import os

import numpy as np

SRC = "/root/reference/ZKX/Task1/evidence/visualization"
HERE = os.path.dirname(os.path.abspath(__file__))
NAME = "task1_s11_strong"


def table(name):
    path = os.path.join(SRC, "data", name)
    with open(path, encoding="utf-8") as fh:
        header = fh.readline().strip().split(",")
    data = np.loadtxt(path, delimiter=",", skiprows=1, ndmin=2)
    return {h: data[:, i] for i, h in enumerate(header)}


def exact_f32(values):
    f = values.astype(np.float32)
    assert np.array_equal(f.astype(np.float64), values), "value is not an FP32 number"
    return f


def cplx(t, re, im):
    return (exact_f32(t[re]) + 1j * exact_f32(t[im])).astype(np.complex64)


def main():
    conf = {}
    with open(os.path.join(SRC, "task1_selected.conf"), encoding="utf-8") as fh:
        for line in fh:
            line = line.strip()
            if line and not line.startswith("#"):
                k, v = line.split("=", 1)
                conf[k] = float(v) if any(c in v for c in ".e") else int(v)
    P, S, T = conf["num_pulses"], conf["samples_per_pulse"], conf["pulse_samples"]
    w = table("task1_step1_waveform.csv")
    e = table("task1_step1_echo.csv")
    c2 = table("task1_step2_compressed.csv")
    c3 = table("task1_step3_range_doppler.csv")
    c4 = table("task1_step4_cfar.csv")
    a2 = table("task1_step5_ambiguity.csv")
    ad = table("task1_step5_delay_cut.csv")
    af = table("task1_step5_doppler_cut.csv")
    for t in (e, c2):
        assert np.array_equal(t["pulse"], np.repeat(np.arange(P), S))
    for t in (c3, c4):
        assert np.array_equal(t["doppler_index"], np.repeat(np.arange(P), S))
    rows = 2 * T - 1
    nfreq = conf["ambiguity_nfreq"]
    assert np.array_equal(a2["delay_index"], np.repeat(np.arange(rows), nfreq))
    arrays = {
        "waveform": cplx(w, "real", "imag"),
        "noise": cplx(e, "noise_real", "noise_imag").reshape(P, S),
        "echo": cplx(e, "echo_real", "echo_imag").reshape(P, S),
        "compressed": cplx(c2, "real", "imag").reshape(P, S),
        "range_doppler": cplx(c3, "real", "imag").reshape(P, S),
        "power": exact_f32(c4["power"]).reshape(P, S),
        "threshold": exact_f32(c4["threshold"]).reshape(P, S),
        "detection": c4["detection"].astype(np.uint8).reshape(P, S),
        "ambiguity": exact_f32(a2["value"]).reshape(rows, nfreq),
        "ambiguity_delay": exact_f32(ad["value"]),
        "ambiguity_doppler": exact_f32(af["value"]),
    }
    assert arrays["waveform"].shape == (T,)
    # The noiseless echo is not stored: the reference wrote echo = noiseless + noise
    # in FP32, so echo - noise recovers it to within one rounding of the echo.
    gap = (e["echo_real"] - e["noiseless_real"] - e["noise_real"]).max()
    assert abs(gap) < 1e-6
    np.savez_compressed(os.path.join(HERE, NAME + ".npz"), **arrays)
    with open(os.path.join(HERE, NAME + ".json"), "w", encoding="utf-8") as fh:
        json.dump({"profile": NAME, "config": conf,
                   "source": "ZKX/Task1/evidence/visualization (reference GPU pipeline output)"},
                  fh, indent=1, sort_keys=True)
    print({k: v.shape for k, v in arrays.items()})


if __name__ == "__main__":
    main()
