"""Training pipeline of the learned push/pull scheduler.

Traces come from the plain method itself: every (graph, source batch) is run once
with every hop forced to push and once forced to pull, with per-hop HIP-event
timing.  Both runs produce the same layers bit for bit, so hop i has the same
frontier features in both and two measured times -- one sample
(features -> [log2 ms push, log2 ms pull]).  The MLP is trained on the GPU by the
library's own kernels (vs_sched_train) and judged against the density heuristic
on held-out graphs.

CLI:  python -m vertexsurge_amd.sched_train --out vertexsurge_amd/sched_weights.npz
"""
from __future__ import annotations

import argparse
import json
import math
import sys
import time
from pathlib import Path

import numpy as np

from . import Graph, HostCSR, Scheduler, sample_sources, vexpand
from ._lib import MLP_IN, MLP_OUT

WEIGHTS = Path(__file__).resolve().parent / "sched_weights.npz"


def features(hop, rows, V, E, pop, act, actdeg, vis_frac) -> np.ndarray:
    """Same formula as sched_features() in vs_mlp.hip (float64 math, float32 out)."""
    rows = rows if rows else 1
    x = [math.log2(1.0 + pop) / 32.0, math.log2(1.0 + act) / 32.0,
         math.log2(1.0 + actdeg) / 40.0, hop / 8.0, pop / (rows * (V if V else 1)),
         math.log2(1.0 + E) / 40.0, math.log2(1.0 + V) / 32.0, vis_frac]
    return np.array(x, dtype=np.float32)


def trace_query(graph: Graph, V: int, E: int, sources, k_max: int, path="shortest"):
    """Per-hop samples of one query: (X [n, IN], Y [n, OUT], M [n, OUT])."""
    _, sp = vexpand(graph, sources, 1, k_max, path, sched="push", timing=True)
    _, sl = vexpand(graph, sources, 1, k_max, path, sched="pull", timing=True)
    if sp["k_actual_max"] != sl["k_actual_max"] or sp["te"] != sl["te"]:
        raise RuntimeError("push and pull runs disagree")
    X, Y, M = [], [], []
    rows = len(sources)
    for i in range(sp["k_actual_max"]):
        if sp["frontier_bits"][i] == 0 or sp["num_tiles"] != 1:
            continue
        tp, tl = sp["ms_hop"][i], sl["ms_hop"][i]
        if tp <= 0 or tl <= 0:
            continue
        X.append(features(i + 1, rows, V, E, sp["frontier_bits"][i], sp["frontier_act"][i],
                          sp["frontier_deg"][i], sp["vis_frac"][i]))
        y = np.zeros(MLP_OUT, dtype=np.float32)
        m = np.zeros(MLP_OUT, dtype=np.float32)
        y[0], y[1] = math.log2(tp), math.log2(tl)
        m[0] = m[1] = 1.0
        Y.append(y)
        M.append(m)
    return X, Y, M


def collect(graph_specs, source_sizes=(256, 1024, 4096), k_max=5, seed=1, log=print):
    X, Y, M = [], [], []
    for spec in graph_specs:
        t0 = time.time()
        csr = make_graph(spec)
        G = Graph.from_csr(csr)
        for qi, S in enumerate(spec.get("sources", source_sizes)):
            src = sample_sources(csr.V, min(S, csr.V), seed * 7919 + qi)
            x, y, m = trace_query(G, csr.V, csr.E, src, k_max)
            X += x
            Y += y
            M += m
        log(f"traced {spec}: {len(X)} samples so far ({time.time() - t0:.1f}s)")
        G.close()
        csr.close()
    return np.array(X), np.array(Y), np.array(M)


def make_graph(spec: dict) -> HostCSR:
    spec = {k: v for k, v in spec.items() if k != "sources"}
    if spec["kind"] == "rmat":
        a = spec.get("a", 0.57)
        bc = (1.0 - a - 0.05) / 2
        return HostCSR.rmat(spec["scale"], spec.get("ef", 16), a, bc, bc, spec.get("seed", 1))
    if spec["kind"] == "er":
        return HostCSR.erdos_renyi(spec["V"], spec["pairs"], spec.get("seed", 1))
    raise ValueError(spec)


def train(X, Y, M, epochs=1500, lr=3e-3, seed=1) -> Scheduler:
    s = Scheduler(seed=seed)
    curve = s.train(X, Y, M, epochs=epochs, lr=lr)
    s.curve = curve
    return s


def accuracy(s: Scheduler, X, Y) -> dict:
    """Fraction of samples where the MLP picks the faster option, and the time
    lost against the oracle choice, next to the same for always-push/pull."""
    P = s.predict(X)
    best = np.minimum(Y[:, 0], Y[:, 1])
    pick = np.where(P[:, 0] < P[:, 1], Y[:, 0], Y[:, 1])
    t = lambda v: float(np.sum(np.exp2(v)))
    return {"n": int(len(X)), "agree": float(np.mean(pick == best)),
            "ms_mlp": t(pick), "ms_best": t(best), "ms_push": t(Y[:, 0]),
            "ms_pull": t(Y[:, 1])}


def save(s: Scheduler, path: Path = WEIGHTS, meta: dict | None = None):
    np.savez(path, params=s.get_params(), meta=np.array(json.dumps(meta or {})))


def load(path: Path = WEIGHTS, seed=1) -> Scheduler:
    d = np.load(path, allow_pickle=False)
    s = Scheduler(seed=seed)
    s.set_params(d["params"])
    return s


TRAIN_SPECS = [{"kind": "rmat", "scale": sc, "ef": 16, "a": a, "seed": 100 + i}
               for i, (sc, a) in enumerate([(s, a) for s in (18, 19, 20, 21)
                                            for a in (0.45, 0.5, 0.55, 0.6)])]
HELDOUT_SPECS = [{"kind": "rmat", "scale": 22, "ef": 16, "a": 0.57, "seed": 7},
                 {"kind": "er", "V": 1 << 22, "pairs": 16 << 22, "seed": 7}]
LARGE_SPECS = [{"kind": "rmat", "scale": sc, "ef": 32, "a": a, "seed": 200 + i}
               for i, (sc, a) in enumerate([(s, a) for s in (20, 21, 22)
                                            for a in (0.45, 0.5, 0.55, 0.6)])]
LARGE_HELDOUT = [{"kind": "rmat", "scale": 23, "ef": 32, "a": 0.57, "seed": 77},
                 {"kind": "er", "V": 1 << 23, "pairs": 32 << 23, "seed": 77}]
# R-MAT scales 22-24 (edge factor 32) over a in [0.45, 0.60]
C5_SPECS = ([{"kind": "rmat", "scale": 22, "ef": 32, "a": a, "seed": 300 + i,
              "sources": (256, 1024, 4096)} for i, a in enumerate((0.45, 0.5, 0.55, 0.6))] +
            [{"kind": "rmat", "scale": 23, "ef": 32, "a": a, "seed": 310 + i,
              "sources": (256, 1024, 4096)} for i, a in enumerate((0.45, 0.55))] +
            [{"kind": "rmat", "scale": 24, "ef": 32, "a": a, "seed": 320 + i,
              "sources": (256, 1024, 4096)} for i, a in enumerate((0.5, 0.6))])
# held out: R-MAT scale 25 and an Erdos-Renyi graph with the same V and pairs
C5_HELDOUT = [{"kind": "rmat", "scale": 25, "ef": 32, "a": 0.57, "seed": 777},
              {"kind": "er", "V": 1 << 25, "pairs": 32 << 25, "seed": 777}]
PRESETS = {"small": (TRAIN_SPECS, HELDOUT_SPECS), "large": (LARGE_SPECS, LARGE_HELDOUT),
           "c5": (C5_SPECS, C5_HELDOUT)}


def evaluate(s: Scheduler, specs, sources=(1024, 4096), k_max=4, reps=3, log=print):
    """Held-out comparison on whole queries: device time with the learned
    scheduler vs the density heuristic (median of reps), bits compared."""
    out = []
    for spec in specs:
        csr = make_graph(spec)
        G = Graph.from_csr(csr)
        G.set_scheduler(s)
        for qi, S in enumerate(sources):
            src = sample_sources(csr.V, min(S, csr.V), 900 + qi)
            t = {}
            res = {}
            for sched in ("heuristic", "learned", "heuristic", "learned"):
                r, st = vexpand(G, src, 1, k_max, sched=sched)
                res[sched] = (r.row_counts(0), r.row_digests(0), st["te"])
                r.close()
            for sched in ("heuristic", "learned"):
                ms = []
                for _ in range(reps):
                    r, st = vexpand(G, src, 1, k_max, sched=sched)
                    ms.append(st["ms_total"])
                    r.close()
                t[sched] = float(np.median(ms))
            same = all(np.array_equal(a, b) for a, b in
                       zip(res["heuristic"][:2], res["learned"][:2])) and \
                res["heuristic"][2] == res["learned"][2]
            row = {"graph": spec, "sources": int(S), "ms_heuristic": t["heuristic"],
                   "ms_learned": t["learned"], "bit_exact": bool(same),
                   "te": int(res["learned"][2])}
            log(json.dumps(row))
            out.append(row)
        G.set_scheduler(None)
        G.close()
        csr.close()
    return out


def main(argv=None):
    ap = argparse.ArgumentParser()
    ap.add_argument("--out", default=str(WEIGHTS))
    ap.add_argument("--epochs", type=int, default=2000)
    ap.add_argument("--quick", action="store_true")
    ap.add_argument("--large", action="store_true", help="same as --preset large")
    ap.add_argument("--preset", choices=sorted(PRESETS), default="small",
                    help="small: R-MAT 18-21 / held-out 22 + ER; large: 20-22 / 23 + ER; "
                         "c5: 22-24 / held-out R-MAT 25 + an Erdos-Renyi graph of its size")
    ap.add_argument("--eval-only", action="store_true", help="evaluate the saved weights")
    ap.add_argument("--no-eval", action="store_true", help="train and save only")
    ap.add_argument("--traces", help="also save the traces (X, Y, M) to this .npz")
    args = ap.parse_args(argv)
    preset = "large" if args.large else args.preset
    specs, held = PRESETS[preset]
    log = lambda m: print(f"[sched_train] {m}", file=sys.stderr, flush=True)
    if args.eval_only:
        s = load(Path(args.out))
        rows = evaluate(s, held, log=log)
        print(json.dumps({"heldout": rows}))
        return
    if args.quick:
        specs = specs[::4]
    X, Y, M = collect(specs, log=log)
    if args.traces:
        np.savez(args.traces, X=X, Y=Y, M=M)
    log(f"training on {len(X)} samples")
    s = train(X, Y, M, epochs=args.epochs)
    log(f"loss {s.curve[0]:.4f} -> {s.curve[-1]:.4f}")
    acc = accuracy(s, X, Y)
    log(f"train accuracy {acc}")
    save(s, Path(args.out), {"samples": int(len(X)), "train_acc": acc,
                             "specs": specs, "epochs": args.epochs, "preset": preset})
    if args.no_eval:
        print(json.dumps({"samples": int(len(X)), "final_loss": float(s.curve[-1]),
                          "train": acc}))
        return
    rows = evaluate(s, held, log=log)
    print(json.dumps({"samples": int(len(X)), "final_loss": float(s.curve[-1]),
                      "train": acc, "heldout": rows}))


if __name__ == "__main__":
    main()
