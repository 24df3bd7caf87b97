"""HBM traffic of the dominant kernels from rocprofv3 counters.

Two steps, each its own process:

  1. run the probe workload under rocprofv3 twice, once per counter:
       rocprofv3 --pmc FETCH_SIZE --output-format csv -d OUT/fetch -o run -- \\
           python3 -m zkx_dev_amd.traffic probe
       rocprofv3 --pmc WRITE_SIZE --output-format csv -d OUT/write -o run -- \\
           python3 -m zkx_dev_amd.traffic probe
  2. python3 -m zkx_dev_amd.traffic parse OUT/fetch/run_counter_collection.csv \\
           OUT/write/run_counter_collection.csv profiles/traffic.json

The probe runs a device-to-device copy of a known size (the calibration: its
FETCH_SIZE and WRITE_SIZE, in KiB per dispatch, are scaled to the bytes it
must move), then the pulse-compression and fused Doppler+CFAR kernels on a
batch of 64 dwells of 128 x 4096.  `parse` writes, per kernel, the counter
readings, the calibration factors and the scaled HBM bytes per launch.
"""
from __future__ import annotations

import csv
import json
import sys
from collections import defaultdict

BATCH, P, R = 64, 128, 4096
COPY_BYTES = 256 * 1024 * 1024
KERNELS = {"pc_fused_kernel<4096>": "pc_fused_kernel<4096>",
           "doppler_cfar_kernel<128, 2, 2, 8, 8>": "doppler_cfar_kernel<128, 2, 2, 8, 8>"}


def probe():
    import numpy as np
    import torch

    from . import radar, synth

    dev = "cuda"
    a = torch.empty(COPY_BYTES // 4, dtype=torch.float32, device=dev).uniform_()
    b = torch.empty_like(a)
    for _ in range(3):
        b.copy_(a)
    cubes, tmpl, _ = synth.make_dwell_batch(2, P, R, 256, seed=5)
    x = torch.from_numpy(np.concatenate([cubes] * (BATCH // 2))).to(dev)
    pipe = radar.DwellPipeline(torch.from_numpy(tmpl).to(dev), P, R, max_batch=BATCH)
    pc = torch.empty_like(x)
    for _ in range(3):
        radar.pulse_compression_prepared(x, pipe.spectrum, out=pc)
        radar.doppler_cfar(pc, window="hann")
    torch.cuda.synchronize()
    print("probe done", flush=True)


def _read(path, counter):
    rows = defaultdict(list)
    with open(path, newline="") as fh:
        for r in csv.DictReader(fh):
            if r["Counter_Name"] != counter:
                continue
            rows[(r["Kernel_Name"], r["Grid_Size"])].append(float(r["Counter_Value"]))
    return rows


def _match(rows, needle, grid=None):
    vals = []
    for (name, g), v in rows.items():
        if needle in name and (grid is None or g == grid):
            vals.extend(v)
    return sum(vals) / len(vals) if vals else None


def parse(fetch_csv, write_csv, out_json):
    f = _read(fetch_csv, "FETCH_SIZE")
    w = _read(write_csv, "WRITE_SIZE")
    copy_f = max((sum(v) / len(v) for (n, g), v in f.items() if "copyBuffer" in n), default=None)
    copy_w = max((sum(v) / len(v) for (n, g), v in w.items() if "copyBuffer" in n), default=None)
    if not copy_f or not copy_w:
        raise SystemExit("calibration copy not found in the counter files")
    kf = COPY_BYTES / (copy_f * 1024.0)
    kw = COPY_BYTES / (copy_w * 1024.0)
    out = {"calibration": {"copy_bytes": COPY_BYTES, "copy_fetch_kib": copy_f, "copy_write_kib": copy_w,
                           "fetch_factor": kf, "write_factor": kw},
           "kernels": {}}
    for key, needle in KERNELS.items():
        fk = _match(f, needle.split("<")[0])
        wk = _match(w, needle.split("<")[0])
        if fk is None or wk is None:
            continue
        out["kernels"][key] = {"batch": BATCH, "fetch_kib": fk, "write_kib": wk,
                               "hbm_bytes_per_launch": fk * 1024.0 * kf + wk * 1024.0 * kw}
    with open(out_json, "w") as fh:
        json.dump(out, fh, indent=1)
    print(json.dumps(out, indent=1))


if __name__ == "__main__":
    if len(sys.argv) >= 2 and sys.argv[1] == "probe":
        probe()
    elif len(sys.argv) == 5 and sys.argv[1] == "parse":
        parse(sys.argv[2], sys.argv[3], sys.argv[4])
    else:
        raise SystemExit(__doc__)
