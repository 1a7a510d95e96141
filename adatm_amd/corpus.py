"""Training corpus and training of the learned memoization planner.

The corpus is the set of seeded synthetic tensors of the planner workload:
order 3..8, dims log-uniform in [1e2, 1e6], nonzeros log-uniform in
[1e5, 1e8] (draw count capped so that the engine's memo and CSFs stay within
a fixed device-memory budget), per-mode coordinates uniform or Zipf with
s ~ U[0, 1.5], rank in {8, 16, 32, 64}.  Every candidate plan of every tensor
(adatm_plan_candidates, the reference's enumeration of adatm_bench.c:276-458)
is built on the GPU and one FP32 CP-ALS sweep, captured in a hipGraph, is
timed (median of 5 after 2 warm-ups).  The timings are the labels of the cost
model and, for held-out tensors, the exhaustive search its regret is scored
against.

    python -m adatm_amd.corpus gen --start 0 --end 200 --out corpus_0_200.npz
    python -m adatm_amd.corpus train --data corpus_*.npz --out costmodel.bin
    python -m adatm_amd.corpus eval --data corpus_*.npz --model costmodel.bin

Tensors k < 320 train the model, k >= 320 are held out.
"""
from __future__ import annotations

import argparse
import glob
import json
import math
import sys
import time

import numpy as np

import adatm_amd as A

CORPUS_SEED = 0xADA7A004
NTENSORS = 400
NTRAIN = 320
RANKS = (8, 16, 32, 64)
MEM_BUDGET = 48e9  # device bytes the memo + CSF estimate of a tensor may use
TARGET = "log10(ms) of one FP32 CP-ALS sweep (hipGraph), median of 5"


def corpus_spec(k: int):
    """(SynthSpec, rank) of corpus tensor k, a pure function of k."""
    rng = np.random.Generator(np.random.PCG64([CORPUS_SEED, k]))
    N = int(rng.integers(3, 9))
    dims = np.exp(rng.uniform(math.log(1e2), math.log(1e6), N)).astype(np.int64)
    draws = float(np.exp(rng.uniform(math.log(1e5), math.log(1e8))))
    zipf = bool(rng.integers(0, 2))
    s = rng.uniform(0.0, 1.5, N)
    R = int(RANKS[rng.integers(0, len(RANKS))])
    # memo rows and group CSFs grow ~ nnz * (R * 4 * (N - 2) + 8 N) bytes
    cap = MEM_BUDGET / (R * 4.0 * max(N - 2, 1) + 8.0 * N)
    draws = int(min(draws, cap))
    spec = A.make_spec(list(dims), draws, dist=A.DIST_ZIPF if zipf else A.DIST_UNIFORM,
                       zipf_s=list(s), values=A.VALS_U01, seed=CORPUS_SEED + k)
    return spec, R


def measure(k: int, nwarm: int = 2, nrep: int = 5, device: int = 0) -> dict:
    spec, R = corpus_spec(k)
    t0 = time.time()
    t = A.Tensor.synth(spec)
    t.canonicalize()
    cands, ch = A.plan_analytic(t, R, ref_quirks=True, device=device)
    feats = A.plan_features(t, R, cands, device=device)
    cx, _, ms = A.plan_exhaustive(t, R, A.FP32, device, nwarm, nrep)
    assert [p.groups(t.nmodes) for p in cx] == [p.groups(t.nmodes) for p in cands]
    rec = dict(k=k, N=t.nmodes, nnz=t.nnz, R=R, dims=list(t.dims), nc=len(cands),
               analytic=ch, ms=ms, feats=feats,
               pops=np.array([p.pops for p in cands], dtype=np.float64),
               secs=time.time() - t0)
    return rec


def save(recs: list[dict], path: str) -> None:
    rows = sum(r["nc"] for r in recs)
    tens = np.zeros(rows, dtype=np.int32)
    X = np.zeros((rows, A.MLP_IN), dtype=np.float32)
    ms = np.zeros(rows, dtype=np.float64)
    pops = np.zeros(rows, dtype=np.float64)
    cand = np.zeros(rows, dtype=np.int32)
    i = 0
    for r in recs:
        n = r["nc"]
        tens[i:i + n] = r["k"]
        X[i:i + n] = r["feats"]
        ms[i:i + n] = r["ms"]
        pops[i:i + n] = r["pops"]
        cand[i:i + n] = np.arange(n)
        i += n
    meta = np.array([[r["k"], r["N"], r["nnz"], r["R"], r["nc"], r["analytic"]]
                     for r in recs], dtype=np.int64)
    np.savez_compressed(path, tensor=tens, X=X, ms=ms, pops=pops, cand=cand, meta=meta)


def load(paths) -> dict:
    parts = [np.load(p) for p in paths]
    d = {key: np.concatenate([p[key] for p in parts]) for key in parts[0].files}
    order = np.lexsort((d["cand"], d["tensor"]))
    for key in ("tensor", "X", "ms", "pops", "cand"):
        d[key] = d[key][order]
    d["meta"] = d["meta"][np.argsort(d["meta"][:, 0], kind="stable")]
    return d


def regrets(d: dict, choose) -> np.ndarray:
    """Per-tensor regret t(chosen)/t(best) - 1; choose(k, rows) -> local index."""
    out = []
    for k in np.unique(d["tensor"]):
        rows = np.nonzero(d["tensor"] == k)[0]
        ms = d["ms"][rows]
        out.append(ms[choose(int(k), rows)] / ms.min() - 1.0)
    return np.array(out)


def analytic_choice(d: dict):
    an = {int(m[0]): int(m[5]) for m in d["meta"]}
    return lambda k, rows: an[k]


def model_choice(d: dict, pred: np.ndarray):
    return lambda k, rows: int(np.argmin(pred[rows]))


def split(d: dict, train: bool) -> dict:
    keep = (d["tensor"] < NTRAIN) if train else (d["tensor"] >= NTRAIN)
    out = {key: d[key][keep] for key in ("tensor", "X", "ms", "pops", "cand")}
    mk = (d["meta"][:, 0] < NTRAIN) if train else (d["meta"][:, 0] >= NTRAIN)
    out["meta"] = d["meta"][mk]
    return out


def train(d: dict, epochs: int = 800, batch: int = 64, lr: float = 3e-3,
          seed: int = 1, device: int = 0) -> "A.Mlp":
    tr = split(d, True)
    m = A.Mlp(device=device, seed=seed)
    m.train(tr["X"], np.log10(tr["ms"]).astype(np.float32), epochs, batch, lr, seed)
    return m


def summary(r: np.ndarray) -> dict:
    return dict(mean=float(r.mean()), median=float(np.median(r)), max=float(r.max()),
                n=int(r.size))


def main(argv=None) -> int:
    ap = argparse.ArgumentParser(prog="python -m adatm_amd.corpus")
    sub = ap.add_subparsers(dest="cmd", required=True)
    g = sub.add_parser("gen")
    g.add_argument("--start", type=int, default=0)
    g.add_argument("--end", type=int, default=NTENSORS)
    g.add_argument("--out", required=True)
    g.add_argument("--budget-s", type=float, default=1e9,
                   help="stop after this many seconds (the file holds what finished)")
    t = sub.add_parser("train")
    t.add_argument("--data", nargs="+", required=True)
    t.add_argument("--out", required=True)
    t.add_argument("--epochs", type=int, default=800,
                   help="800 (with lr 3e-3, batch 64) had the lowest mean regret, over "
                        "three seeds, on tensors 256..319 when training on 0..255")
    e = sub.add_parser("eval")
    e.add_argument("--data", nargs="+", required=True)
    e.add_argument("--model", required=True)
    a = ap.parse_args(argv)

    if a.cmd == "gen":
        recs, t0 = [], time.time()
        for k in range(a.start, a.end):
            r = measure(k)
            recs.append(r)
            print(f"[corpus] k={k} N={r['N']} nnz={r['nnz']} R={r['R']} nc={r['nc']} "
                  f"analytic={r['analytic']} best={int(np.argmin(r['ms']))} "
                  f"ms={np.round(r['ms'], 3).tolist()} ({r['secs']:.1f}s)", flush=True)
            save(recs, a.out)
            if time.time() - t0 > a.budget_s:
                print(f"[corpus] time budget reached after k={k}", flush=True)
                break
        return 0
    paths = sorted(set(sum((glob.glob(p) for p in a.data), [])))
    d = load(paths)
    if a.cmd == "train":
        m = train(d, epochs=a.epochs)
        m.save(a.out)
    else:
        m = A.Mlp(path=a.model)
    te = split(d, False)
    pred = m.predict(te["X"])
    res = dict(heldout_tensors=int(len(np.unique(te["tensor"]))),
               learned=summary(regrets(te, model_choice(te, pred))),
               analytic=summary(regrets(te, analytic_choice(te))))
    print(json.dumps(res), flush=True)
    return 0


if __name__ == "__main__":
    sys.exit(main())
