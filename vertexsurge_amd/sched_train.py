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


def query_features(S, vol2, degsum, k_max, V, E) -> np.ndarray:
    """Same formula as query_features() in vs_mlp.hip: the query-level row of
    the tile-width decision (hop feature 0)."""
    x = [math.log2(1.0 + S) / 32.0, math.log2(1.0 + vol2) / 48.0,
         math.log2(1.0 + degsum) / 40.0, 0.0, 0.0, math.log2(1.0 + E) / 40.0,
         math.log2(1.0 + V) / 32.0, k_max / 8.0]
    return np.array(x, dtype=np.float32)


TW_SLOT = {16: 2, 32: 3, 64: 4}


def trace_tile_widths(graph: Graph, V: int, E: int, sources, k_max: int, reps: int = 2,
                      path="shortest"):
    """One query-level sample: the whole query's device time at every tile
    width it allows (16, 32, 64; median of reps), heuristic push/pull."""
    W = (len(sources) + 63) // 64
    widths = [tw for tw in (16, 32, 64) if tw <= max(1, 1 << (W - 1).bit_length())]
    if len(widths) < 2:
        return [], [], []
    y = np.zeros(MLP_OUT, dtype=np.float32)
    m = np.zeros(MLP_OUT, dtype=np.float32)
    st = None
    for tw in widths:
        ms = []
        for _ in range(reps):
            r, st = vexpand(graph, sources, 1, k_max, path, tile_words=tw, timing=True)
            r.close()
            ms.append(st["ms_total"])
        y[TW_SLOT[tw]] = math.log2(float(np.median(ms)))
        m[TW_SLOT[tw]] = 1.0
    x = query_features(len(sources), st["src_vol2"], st["src_degsum"], k_max, V, E)
    return [x], [y], [m]


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


def collect(graph_specs, source_sizes=(256, 1024, 4096), k_max=5, seed=1, log=print,
            tw_sizes=(2048, 4096, 8192), tw_hops=(3, 4, 5)):
    """Push/pull samples per (hop, tile) and tile-width samples per query."""
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
            log(f"  push/pull traces of {S} sources: {len(x)} samples")
        for qi, S in enumerate(spec.get("tw_sources", tw_sizes)):
            for k in tw_hops:
                src = sample_sources(csr.V, min(S, csr.V), seed * 104729 + qi * 31 + k)
                x, y, m = trace_tile_widths(G, csr.V, csr.E, src, k)
                X += x
                Y += y
                M += m
            log(f"  tile-width traces of {S} sources done")
        log(f"traced {spec}: {len(X)} samples so far ({time.time() - t0:.1f}s)")
        G.close()
        csr.close()
    return np.array(X), np.array(Y), np.array(M)


def make_graph(spec: dict) -> HostCSR:
    spec = {k: v for k, v in spec.items() if k not in ("sources", "tw_sources")}
    if spec["kind"] == "rmat":
        a = spec.get("a", 0.57)
        bc = (1.0 - a - 0.05) / 2
        return HostCSR.rmat(spec["scale"], spec.get("ef", 16), a, bc, bc, spec.get("seed", 1))
    if spec["kind"] == "er":
        return HostCSR.erdos_renyi(spec["V"], spec["pairs"], spec.get("seed", 1))
    raise ValueError(spec)


def nn_distance(A, B, leave_one_out=False):
    """Euclidean distance of every row of A to its nearest row of B."""
    A = np.asarray(A, dtype=np.float64)
    B = np.asarray(B, dtype=np.float64)
    d = np.sqrt(((A[:, None, :] - B[None, :, :]) ** 2).sum(-1))
    if leave_one_out:
        np.fill_diagonal(d, np.inf)
    return d.min(axis=1)


def domain_of(X, max_rows=1024):
    """The decision domain of a model trained on X: its samples (every k-th
    when there are more than max_rows) and the largest leave-one-out
    nearest-neighbour distance among them, so every training sample lies
    inside and a decision farther from all of them than the training samples
    are from each other falls back to the heuristic."""
    X = np.asarray(X, dtype=np.float32)
    rows = X[:: max(1, -(-len(X) // max_rows))]
    radius = float(nn_distance(rows, rows, leave_one_out=True).max()) if len(rows) > 1 else 0.0
    return rows, radius


def train(X, Y, M, epochs=1500, lr=3e-3, seed=1) -> Scheduler:
    s = Scheduler(seed=seed)
    curve = s.train(X, Y, M, epochs=epochs, lr=lr)
    s.curve = curve
    s.set_domain(*domain_of(X))
    return s


def in_domain(X, domain):
    if domain is None or len(domain[0]) == 0:
        return np.ones(len(X), dtype=bool)
    return nn_distance(X, domain[0]) <= domain[1] + 1e-6


def hop_rows(X):
    """Rows of per-hop push/pull samples (hop feature > 0)."""
    return X[:, 3] > 0


def accuracy(s: Scheduler, X, Y) -> dict:
    """Fraction of push/pull samples where the MLP picks the faster option,
    and the time lost against the oracle choice, next to the same for
    always-push/pull."""
    h = hop_rows(X)
    X, Y = X[h], Y[h]
    P = s.predict(X)
    best = np.minimum(Y[:, 0], Y[:, 1])
    pick = np.where(P[:, 0] < P[:, 1], Y[:, 0], Y[:, 1])
    t = lambda v: float(np.sum(np.exp2(v)))
    return {"n": int(len(X)), "agree": float(np.mean(pick == best)),
            "ms_mlp": t(pick), "ms_best": t(best), "ms_push": t(Y[:, 0]),
            "ms_pull": t(Y[:, 1])}


def decide_push_pull(P, X, E, alpha=16, domain=None):
    """The library's decision (sched_decide): the MLP's faster option unless
    the sample lies outside the decision domain or the two predictions are
    within 0.1 log2 units, then the density heuristic (push iff frontier
    out-edges * alpha < E).  0 = push, 1 = pull."""
    actdeg = np.exp2(X[:, 2].astype(np.float64) * 40.0) - 1.0
    heur = np.where(actdeg * alpha < E, 0, 1)
    mlp = np.where(P[:, 0] < P[:, 1], 0, 1)
    keep = (np.abs(P[:, 0] - P[:, 1]) < 0.1) | ~in_domain(X, domain)
    return np.where(keep, heur, mlp), heur


def decide_tile_width(P, M, fallback=64, X=None, domain=None):
    """The library's decision (sched_tile_words) over the widths a sample
    allows (M): the lowest prediction, unless the sample lies outside the
    decision domain or the pick is not 0.1 log2 units below the fallback's.
    Returns output slots."""
    slots = [TW_SLOT[tw] for tw in (16, 32, 64)]
    fb = np.array([TW_SLOT[min(fallback, max(tw for tw in (16, 32, 64) if m[TW_SLOT[tw]]))]
                   for m in M])
    pick = fb.copy()
    inside = in_domain(X, domain) if X is not None else np.ones(len(P), dtype=bool)
    for n in range(len(P)):
        if not inside[n]:
            continue
        allowed = [k for k in slots if M[n, k]]
        best = min(allowed, key=lambda k: P[n, k])
        if P[n, fb[n]] - P[n, best] >= 0.1:
            pick[n] = best
    return pick, fb


def tw_accuracy(s: Scheduler, X, Y, M) -> dict:
    """Tile-width samples: time of the learned choice against the fallback
    (the widest the query allows) and the best width."""
    q = ~hop_rows(X)
    if not q.any():
        return {"n": 0}
    X, Y, M = X[q], Y[q], M[q]
    pick, fb = decide_tile_width(s.predict(X), M, X=X, domain=s.domain)
    r = np.arange(len(X))
    t = lambda v: float(np.sum(np.exp2(v)))
    best = np.array([min(Y[n, k] for k in (2, 3, 4) if M[n, k]) for n in r])
    return {"n": int(len(X)), "ms_mlp": t(Y[r, pick]), "ms_fallback": t(Y[r, fb]),
            "ms_best": t(best), "changed": int(np.sum(pick != fb))}


def save(s: Scheduler, path: Path = WEIGHTS, meta: dict | None = None):
    rows, radius = s.domain if s.domain is not None else (np.zeros((0, MLP_IN), np.float32), 0.0)
    np.savez(path, params=s.get_params(), domain=rows, radius=np.array([radius], np.float32),
             meta=np.array(json.dumps(meta or {})))


def load(path: Path = WEIGHTS, seed=1) -> Scheduler:
    d = np.load(path, allow_pickle=False)
    s = Scheduler(seed=seed)
    s.set_params(d["params"])
    if "domain" in d.files:
        s.set_domain(d["domain"], float(d["radius"][0]))
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
              "sources": (256, 1024, 4096)} for i, a in enumerate((0.5, 0.6))] +
            [{"kind": "rmat", "scale": 25, "ef": 32, "a": 0.5, "seed": 330,
              "sources": (256, 1024), "tw_sources": (2048, 4096)}])
# held out: R-MAT scale 25 and an Erdos-Renyi graph with the same V and pairs
C5_HELDOUT = [{"kind": "rmat", "scale": 25, "ef": 32, "a": 0.57, "seed": 777},
              {"kind": "er", "V": 1 << 25, "pairs": 32 << 25, "seed": 777}]
PRESETS = {"small": (TRAIN_SPECS, HELDOUT_SPECS), "large": (LARGE_SPECS, LARGE_HELDOUT),
           "c5": (C5_SPECS, C5_HELDOUT)}


def heldout_traces(s: Scheduler, specs, source_sizes=(128, 1024, 4096), k_max=5,
                   tw_sizes=(2048, 4096, 8192), tw_hops=(3, 4, 5), log=print, save=None):
    """Both decisions judged on measured traces of held-out graphs: per graph,
    the summed time of the library's learned choices against the density
    heuristic (push/pull) and against the default width (tile width), all on
    the same measured table, so no timing noise separates them."""
    out = []
    for spec in specs:
        csr = make_graph(spec)
        G = Graph.from_csr(csr)
        X, Y, M = collect_graph(G, csr, source_sizes, k_max, tw_sizes, tw_hops, seed=500)
        G.close()
        if save:
            np.savez(f"{save}_{len(out)}.npz", X=X, Y=Y, M=M, E=np.array([csr.E]))
        P = s.predict(X)
        h = hop_rows(X)
        row = {"graph": spec}
        if h.any():
            pick, heur = decide_push_pull(P[h], X[h], csr.E, domain=s.domain)
            r = np.arange(int(h.sum()))
            t = lambda k: float(np.sum(np.exp2(Y[h][r, k])))
            row.update(hops=int(h.sum()), ms_heuristic=t(heur), ms_learned=t(pick),
                       ms_best=float(np.sum(np.exp2(Y[h][:, :2].min(axis=1)))),
                       changed=int(np.sum(pick != heur)))
        q = ~h
        if q.any():
            pick, fb = decide_tile_width(P[q], M[q], X=X[q], domain=s.domain)
            r = np.arange(int(q.sum()))
            row.update(queries=int(q.sum()), tw_ms_default=float(np.sum(np.exp2(Y[q][r, fb]))),
                       tw_ms_learned=float(np.sum(np.exp2(Y[q][r, pick]))),
                       tw_changed=int(np.sum(pick != fb)))
        log(json.dumps(row))
        out.append(row)
        csr.close()
    return out


def collect_graph(G, csr, source_sizes, k_max, tw_sizes, tw_hops, seed):
    X, Y, M = [], [], []
    for qi, S in enumerate(source_sizes):
        x, y, m = trace_query(G, csr.V, csr.E, sample_sources(csr.V, min(S, csr.V), seed + qi),
                              k_max)
        X += x
        Y += y
        M += m
        print(f"[sched_train]   held-out push/pull traces of {S} sources", file=sys.stderr,
              flush=True)
    for qi, S in enumerate(tw_sizes):
        for k in tw_hops:
            src = sample_sources(csr.V, min(S, csr.V), seed * 7 + qi * 31 + k)
            x, y, m = trace_tile_widths(G, csr.V, csr.E, src, k)
            X += x
            Y += y
            M += m
    return np.array(X), np.array(Y), np.array(M)


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
            tw_learned = None
            for sched in ("heuristic", "learned", "heuristic", "learned"):
                r, st = vexpand(G, src, 1, k_max, sched=sched)
                res[sched] = (r.row_counts(0), r.row_digests(0), st["te"])
                if sched == "learned":
                    tw_learned = st["tile_words"]
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
                   "te": int(res["learned"][2]), "tile_words_learned": tw_learned}
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
    ap.add_argument("--no-timed-eval", action="store_true",
                    help="judge on held-out traces only (skip the whole-query timings)")
    ap.add_argument("--traces", help="also save the traces (X, Y, M) to this .npz")
    ap.add_argument("--heldout-traces", help="save each held-out graph's traces to PREFIX_i.npz")
    args = ap.parse_args(argv)
    preset = "large" if args.large else args.preset
    specs, held = PRESETS[preset]
    log = lambda m: print(f"[sched_train] {m}", file=sys.stderr, flush=True)
    if args.eval_only:
        s = load(Path(args.out))
        traces = heldout_traces(s, held, log=log, save=args.heldout_traces)
        rows = [] if args.no_timed_eval else evaluate(s, held, log=log)
        print(json.dumps({"heldout_traces": traces, "heldout": rows}))
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
    acc["tile_width"] = tw_accuracy(s, X, Y, M)
    log(f"train accuracy {acc}")
    save(s, Path(args.out), {"samples": int(len(X)), "train_acc": acc,
                             "specs": specs, "epochs": args.epochs, "preset": preset})
    if args.no_eval:
        print(json.dumps({"samples": int(len(X)), "final_loss": float(s.curve[-1]),
                          "train": acc}))
        return
    traces = heldout_traces(s, held, log=log, save=args.heldout_traces)
    rows = [] if args.no_timed_eval else evaluate(s, held, log=log)
    print(json.dumps({"samples": int(len(X)), "final_loss": float(s.curve[-1]),
                      "train": acc, "heldout_traces": traces, "heldout": rows}))


if __name__ == "__main__":
    main()
