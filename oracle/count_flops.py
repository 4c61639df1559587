"""Exact algorithmic flop counts of a workload frame, from the counted-float
instantiation of the oracle (every +,-,*,/,sqrt,floor and transcendental
counts 1; comparisons, min/max/abs are free).

TEST INFRASTRUCTURE: run offline; its output is the per-frame figure used
for the tracer's roofline.  Usage:
    python oracle/count_flops.py kerr09_1080p_high 0 1 2
"""
import json
import os
import sys
import time

HERE = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, HERE)
sys.path.insert(0, os.path.join(os.path.dirname(HERE), "black-hole_amd"))

import bhtrace  # noqa: E402  (parameter presets only: host code, no GPU)
import pyoracle  # noqa: E402


def main():
    name = sys.argv[1]
    frames = [int(x) for x in sys.argv[2:]] or [0]
    O = pyoracle.Oracle()
    prec = bhtrace.get_precision()   # the library's default FP64 band rule (host setting)
    O.set_band(prec.fp64Band, prec.critBand, prec.polarBand)
    out = {}
    for f in frames:
        P = bhtrace.workload(name, f)
        Q, _ = bhtrace.sanitized(P)
        t = time.time()
        c = O.count_flops(Q)
        c["seconds"] = round(time.time() - t, 1)
        c["rays"] = Q.width * Q.height
        out[f"{name}/{f}"] = c
        print(json.dumps({f"{name}/{f}": c}), flush=True)
    return out


if __name__ == "__main__":
    main()
