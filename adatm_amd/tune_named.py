"""Autotune the named workloads and record the choices that bench.py uses.

For every named workload (the 5-mode Zipf tensor at rank 32, the 4-mode
uniform tensor at rank 16, the 8-mode clustered tensor at rank 64) and every
plan and precision the benchmark sweeps, run adatm_autotune_tiles (median of
`--nrep` CP-ALS sweeps per candidate) `--runs` times and keep the fastest
result; write them to planner/tuned_tiles.json as
{"C<config>/R<rank>/plan<index>/<precision>": [red_tile, scat_tile, batch,
flags, ms]}.  Needs a GPU.

    python -m adatm_amd.tune_named [--configs 2,1,3] [--runs 2] [--nrep 5] [--out FILE]
"""
from __future__ import annotations

import argparse
import json
import os
import time

import adatm_amd as A

HERE = os.path.dirname(os.path.abspath(__file__))
TABLE = os.path.join(HERE, "planner", "tuned_tiles.json")
RANK = {1: 16, 2: 32, 3: 64}
PREC_NAME = {A.FP64: "fp64", A.FP32: "fp32", A.BF16ROWS: "bf16rows"}


def key(config: int, rank: int, plan_index: int, precision: int) -> str:
    return f"C{config}/R{rank}/plan{plan_index}/{PREC_NAME[precision]}"


def load_table(path: str = TABLE) -> dict:
    try:
        with open(path) as f:
            return json.load(f)
    except OSError:
        return {}


def main() -> int:
    ap = argparse.ArgumentParser()
    ap.add_argument("--configs", default="2,1,3")
    ap.add_argument("--runs", type=int, default=2)
    ap.add_argument("--nrep", type=int, default=5)
    ap.add_argument("--out", default=TABLE, help="table to update (default: the committed one)")
    args = ap.parse_args()
    table = load_table(args.out)
    model_path = os.path.join(HERE, "planner", "costmodel.bin")
    for cfg in (int(c) for c in args.configs.split(",")):
        R = RANK[cfg]
        t = A.synth_tensor(cfg)
        cands, ch = A.plan_analytic(t, R, device=0)
        plans = {ch: cands[ch]}
        if os.path.exists(model_path):
            model = A.Mlp(device=0, path=model_path)
            lc, lch, _ = A.plan_learned(t, R, model, device=0)
            model.close()
            plans[lch] = lc[lch]
        for idx, plan in plans.items():
            precs = [A.FP32, A.FP64] + ([A.BF16ROWS] if cfg == 3 else [])
            for prec in precs:
                best = None
                for _ in range(args.runs):
                    t0 = time.time()
                    tl, ms = A.autotune_tiles(t, R, plan, prec, 0, args.nrep)
                    print(f"{key(cfg, R, idx, prec)}: {tl} {ms:.3f} ms "
                          f"({time.time() - t0:.1f}s)", flush=True)
                    if best is None or ms < best[1]:
                        best = (tl, ms)
                table[key(cfg, R, idx, prec)] = list(best[0]) + [round(best[1], 4)]
                with open(args.out, "w") as f:
                    json.dump(table, f, indent=1, sort_keys=True)
        t.free()
    return 0


if __name__ == "__main__":
    raise SystemExit(main())
