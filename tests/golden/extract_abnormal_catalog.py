"""Copy the reference's abnormal-input catalog rows of the hot-path operators
into a committed fixture.

The reference lists one expected ErrorCode per (operator, mutation) in the
data file ZKX/test_all/operators/abnormal/operator_abnormal_case_catalog.csv.
This script reads that file (data only, never any reference source), keeps
the rows of firfilter, resample_poly, upfirdn, ca_cfar, cfar_alpha,
pulse_compression and pulse_doppler, and the columns that say what a case
mutates and which code it expects, and writes abnormal_catalog_hot_ops.csv.

Run in a container where /root/reference exists:
    python tests/golden/extract_abnormal_catalog.py
"""
import csv
import os

REF = "/root/reference"
SRC = os.path.join(REF, "ZKX", "test_all", "operators", "abnormal", "operator_abnormal_case_catalog.csv")
OPS = ("firfilter", "resample_poly", "upfirdn", "ca_cfar", "cfar_alpha", "pulse_compression", "pulse_doppler")
KEEP = ("operator_name", "abnormal_case_id", "abnormal_type", "mutation", "expected_code", "applicable_backend")
OUT = os.path.join(os.path.dirname(os.path.abspath(__file__)), "abnormal_catalog_hot_ops.csv")


def main():
    with open(SRC, newline="", encoding="utf-8") as fh:
        rows = [r for r in csv.DictReader(fh) if r["operator_name"] in OPS]
    with open(OUT, "w", newline="", encoding="utf-8") as fh:
        w = csv.DictWriter(fh, fieldnames=KEEP, lineterminator="\n")
        w.writeheader()
        for r in rows:
            w.writerow({k: r[k] for k in KEEP})
    print(f"wrote {len(rows)} rows to {OUT}")


if __name__ == "__main__":
    main()
