"""Writes the known-answer fixture used by the oracle tests.

Reads the reference's own data file
``src/data/2018-09-01-2024-Bitfinex_Spot-4h.csv`` (columns: timestamp, open,
close, high, low, volume -- the order read_candles_from_csv expects,
src/utilities/data_loader.rs:181-206) and stores the price columns as exact
float64 values (Python's float() and Rust's str::parse::<f64> are both
correctly rounded, so the numbers are bit-identical to what the reference's
tests load).  Run once in a container where the reference checkout exists:

    python tests/golden/make_fixtures.py /root/reference
"""
import csv
import os
import sys

import numpy as np


def main(ref_root: str) -> None:
    src = os.path.join(ref_root, "src", "data", "2018-09-01-2024-Bitfinex_Spot-4h.csv")
    cols = {"timestamp": [], "open": [], "close": [], "high": [], "low": [], "volume": []}
    with open(src, newline="") as f:
        rd = csv.reader(f)
        next(rd)  # header row "0,1,2,3,4,5"
        for rec in rd:
            cols["timestamp"].append(int(rec[0]))
            cols["open"].append(float(rec[1]))
            cols["close"].append(float(rec[2]))
            cols["high"].append(float(rec[3]))
            cols["low"].append(float(rec[4]))
            cols["volume"].append(float(rec[5]))
    out = os.path.join(os.path.dirname(os.path.abspath(__file__)), "bitfinex_4h.npz")
    np.savez_compressed(out, **{k: np.asarray(v) for k, v in cols.items()})
    print(f"wrote {out}: {len(cols['close'])} candles")


if __name__ == "__main__":
    main(sys.argv[1] if len(sys.argv) > 1 else "/root/reference")
