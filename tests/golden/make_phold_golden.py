"""Writes tests/golden/phold_golden.json from the CPU oracle (oracle/liboracle.so).

Run once on the CPU:  python tests/golden/make_phold_golden.py [--big]
  default : the 1024-LP workload and the small parity workloads (seconds)
  --big   : also the 2^20-LP workload (1.5e9 events; minutes on one core)

Each record holds the workload, the committed-event total, the combined trace hash,
and for workloads with at most 4096 LPs the per-LP counters and hashes.
"""
import argparse
import json
import os
import sys

HERE = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, os.path.join(HERE, ".."))

from oracle_lib import Oracle  # noqa: E402

SMALL = [
    # name, L, k, T
    ("tiny_l4_k2_t5", 4, 2, 5.0),
    ("small_l64_k8_t20", 64, 8, 20.0),
    ("phold_1024_k16_t100", 1024, 16, 100.0),
    ("mid_l4096_k16_t20", 4096, 16, 20.0),
    ("mid_l65536_k32_t10", 65536, 32, 10.0),
]
BIG = [("phold_2p20_k32_t50", 1 << 20, 32, 50.0)]


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--big", action="store_true")
    ap.add_argument("--out", default=os.path.join(HERE, "phold_golden.json"))
    args = ap.parse_args()
    orc = Oracle()
    old = {}
    if os.path.exists(args.out):
        with open(args.out) as f:
            old = json.load(f)
    out = dict(old)
    for name, L, k, T in SMALL + (BIG if args.big else []):
        res, cnt, hsh = orc.phold("seq", L, k, T, seed=42)
        rec = {"num_lps": L, "init_per_lp": k, "end_time": T, "seed": 42, "lookahead": 0.1,
               "p_self": 0.5, "events": int(res.events), "trace_hash": "%016x" % res.trace_hash,
               "in_flight_end": int(res.in_flight_end), "monotone_ok": int(res.monotone_ok)}
        if L <= 4096:
            rec["lp_count"] = [int(x) for x in cnt]
            rec["lp_hash"] = ["%016x" % int(x) for x in hsh]
        out[name] = rec
        print(name, rec["events"], rec["trace_hash"], "%.1fs" % res.seconds, flush=True)
    with open(args.out, "w") as f:
        json.dump(out, f, indent=1, sort_keys=True)


if __name__ == "__main__":
    main()
