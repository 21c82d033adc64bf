"""Writes tests/golden/oracle_hashes.json: SHA-256 of the oracle's output
(hist_p1, hist_p2, hist_dij, jets bytes) on every event of the reference's input
files (354_fastjet.dat, 808_cms.dat, multi-event.dat, copied here as input
fixtures) for kt / C/A / anti-kt at R = 0.4 and 1.0.  Run from the repo root:
    python tests/golden/make_golden.py
"""
import hashlib
import json
import os
import sys

import numpy as np

ROOT = os.path.dirname(os.path.dirname(os.path.dirname(os.path.abspath(__file__))))
sys.path.insert(0, ROOT)
from oracle import oracle as orc  # noqa: E402

HERE = os.path.dirname(os.path.abspath(__file__))


def main():
    out = {}
    for fname in ("354_fastjet.dat", "808_cms.dat", "multi-event.dat"):
        for idx, ev in enumerate(orc.read_events(os.path.join(HERE, fname))):
            for p in (1, 0, -1):
                for R in ("0.4", "1.0"):
                    h = hashlib.sha256()
                    for x in orc.n2plain(ev, p, float(R)):
                        h.update(np.ascontiguousarray(x).tobytes())
                    out[f"{fname}|{idx}|{p}|{R}"] = h.hexdigest()
    with open(os.path.join(HERE, "oracle_hashes.json"), "w") as fh:
        json.dump(out, fh, indent=1, sort_keys=True)
    print(len(out), "hashes")


if __name__ == "__main__":
    main()
