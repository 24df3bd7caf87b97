"""Collect the reference's recorded operator-probe digests into a committed fixture.

The reference archives one ``<op>_probe_raw.json`` per formal probe run, next to a
``*full.log`` whose ``command=`` line holds the probe's CLI arguments (shapes,
seeds, dtype).  This script reads those data files (never any reference source),
keeps the fields that pin an oracle -- the input manifest, the sha256 of the CPU
reference output and of the GPU output -- and writes ``probe_digests.json``.

Probe programs that produced these files:
  radartools: ZKX/test_all/operators/shared/radartools/radartools_operator_probe.cu
  upfirdn:    ZKX/test_all/operators/cases/filtering/upfirdn/operator_probe.cu
  resample:   ZKX/test_all/operators/cases/filtering/resample_poly/operator_probe.cu
  firfilter:  ZKX/test_all/operators/shared/filtering/task2_filter_probe.cu
  decimate:   ZKX/test_all/operators/cases/filtering/decimate/operator_probe.cu
  channelize: ZKX/test_all/operators/cases/filtering/channelize_poly/operator_probe.cu
  correlate:  ZKX/test_all/operators/suites/task2_features/task2_features_probe.cu

Run in a container where /root/reference exists:
    python tests/golden/extract_probe_digests.py
"""
import glob
import json
import os
import shlex
import sys

REF = "/root/reference"
OPS = ["pulse_compression", "pulse_doppler", "ca_cfar", "cfar_alpha", "ambgfun",
       "upfirdn", "resample_poly", "firfilter", "decimate", "channelize_poly", "correlate", "correlate2d"]
KEEP = ["requested_dtype", "output_dtype", "typed_input_manifest",
        "actual_input_digest", "cpu_output_digest", "gpu_output_digest",
        "output_shape", "relative_l2", "relative_linf"]


def parse_command(log_dir):
    for log in sorted(glob.glob(os.path.join(log_dir, "*full.log"))):
        with open(log, encoding="utf-8", errors="replace") as fh:
            for line in fh:
                if line.startswith("command="):
                    toks = shlex.split(line[len("command="):])
                    args = {}
                    i = 1
                    while i < len(toks):
                        if toks[i].startswith("--") and i + 1 < len(toks):
                            args[toks[i][2:]] = toks[i + 1]
                            i += 2
                        else:
                            i += 1
                    for k in ("output", "warmup", "measured"):
                        args.pop(k, None)
                    return args
    return None


def main():
    cases = {}
    for root in ("plot_in_report", "plot_in_report_old"):
        for op in OPS:
            pattern = os.path.join(REF, root, "**", f"{op}_probe_raw.json")
            for path in sorted(glob.glob(pattern, recursive=True)):
                args = parse_command(os.path.dirname(path))
                if args is None:
                    continue
                with open(path, encoding="utf-8") as fh:
                    raw = json.load(fh)
                rec = {k: raw[k] for k in KEEP if k in raw}
                key = json.dumps([op, sorted(args.items())])
                entry = cases.setdefault(key, {"op": op, "args": args, "records": []})
                sig = (rec.get("actual_input_digest"), rec.get("cpu_output_digest"),
                       rec.get("gpu_output_digest"))
                if all((r["actual_input_digest"], r["cpu_output_digest"],
                        r["gpu_output_digest"]) != sig for r in entry["records"]):
                    rec["source"] = os.path.relpath(path, REF)
                    entry["records"].append(rec)
    out = sorted(cases.values(), key=lambda e: (e["op"], json.dumps(e["args"], sort_keys=True)))
    dest = os.path.join(os.path.dirname(os.path.abspath(__file__)), "probe_digests.json")
    with open(dest, "w", encoding="utf-8") as fh:
        json.dump(out, fh, indent=1, ensure_ascii=False, sort_keys=True)
    print(f"wrote {len(out)} cases to {dest}", file=sys.stderr)


if __name__ == "__main__":
    main()
