"""VExpand throughput benchmark (one MI355X).

Workload (default): symmetrized R-MAT scale 24, edge factor 32
(a=.57 b=.19 c=.19 d=.05, self-loops and duplicates removed), 4096 seeded
uniform sources, Shortest, hops 1..4, frontier / visited / result bit-matrices
resident in HBM.  A step is one full VExpand over that source batch; the CSR is
uploaded once before timing.  Prints one JSON line.

The CPU baseline is the reference algorithm restated in C++ (oracle B: 1024-row
stacks, destination-partitioned Hilbert edge blocks, OpenMP) timed on a bounded
sample of the same workload on this host, and checked bit-for-bit against the
GPU result for the same sample.
"""
from __future__ import annotations

import argparse
import json
import os
import sys
import time
from pathlib import Path

ROOT = Path(__file__).resolve().parent
sys.path.insert(0, str(ROOT))

GRAPH_SEED = 1
SOURCE_SEED = 20241203
HBM_PEAK_GBS = 8000.0


def log(msg: str) -> None:
    print(f"[bench] {msg}", file=sys.stderr, flush=True)


def main() -> int:
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=5)
    ap.add_argument("--warmup", type=int, default=2)
    ap.add_argument("--scale", type=int, default=24)
    ap.add_argument("--edge-factor", type=int, default=32)
    ap.add_argument("--sources", type=int, default=4096)
    ap.add_argument("--k-min", type=int, default=1)
    ap.add_argument("--k-max", type=int, default=4)
    ap.add_argument("--path", default="shortest", choices=["shortest", "any"])
    ap.add_argument("--sched", default="heuristic",
                    choices=["heuristic", "push", "pull", "learned"])
    ap.add_argument("--tile-words", type=int, default=0)
    ap.add_argument("--cpu-sources", type=int, default=1024)
    ap.add_argument("--no-cpu", action="store_true")
    ap.add_argument("--profile-only", action="store_true",
                    help="graph + warmup + steps only (for rocprofv3 runs)")
    args = ap.parse_args()

    if args.gpus != 1 or int(os.environ.get("WORLD_SIZE", "1")) > 1:
        print("bench.py measures one GPU: this build runs on a single MI355X and has no "
              "multi-GPU path (sources would shard across GPUs with a replicated CSR; see "
              "the design notes).", file=sys.stderr)
        return 2

    import numpy as np

    import vertexsurge_amd as vs

    if vs.device_count() < 1:
        log("no GPU visible")
        return 1
    t0 = time.time()
    log(f"generating R-MAT scale {args.scale} edge factor {args.edge_factor}")
    csr = vs.HostCSR.rmat(args.scale, args.edge_factor, 0.57, 0.19, 0.19, GRAPH_SEED)
    log(f"graph V={csr.V} E={csr.E} built in {time.time() - t0:.1f}s; uploading")
    G = vs.Graph.from_csr(csr)
    src = vs.sample_sources(csr.V, args.sources, SOURCE_SEED)
    kw = dict(path=args.path, sched=args.sched, tile_words=args.tile_words)

    te = None
    for i in range(args.warmup):
        res, st = vs.vexpand(G, src, args.k_min, args.k_max, **kw)
        te = st["te"]
        res.close()
        log(f"warmup {i}: {st['ms_total']:.2f} ms device, TE={te}, k_actual={st['k_actual_max']}")
    vs._lib.call("vs_device_synchronize")
    times, dev_ms = [], []
    for i in range(args.steps):
        t = time.perf_counter()
        res, st = vs.vexpand(G, src, args.k_min, args.k_max, **kw)
        vs._lib.call("vs_device_synchronize")
        times.append(time.perf_counter() - t)
        dev_ms.append(st["ms_total"])
        te = st["te"]
        res.close()
    vs._lib.call("vs_device_synchronize")
    ms_step = 1e3 * float(np.mean(times))
    gteps = te / (ms_step / 1e3) / 1e9
    log(f"timed {args.steps} steps: {ms_step:.2f} ms/step (device {np.mean(dev_ms):.2f} ms), "
        f"{gteps:.1f} GTEPS")
    if args.profile_only:
        return 0

    # per-kernel HIP-event timings for the roofline figure
    res, tst = vs.vexpand(G, src, args.k_min, args.k_max, timing=True, **kw)
    counts = res.row_counts(0)
    res.close()
    dom = "pull" if tst["ms_pull"] >= tst["ms_push"] else "push"
    n_l = tst[f"n_{dom}"] or 1
    bytes_per_launch = tst[f"bytes_{dom}"] / n_l
    ms_per_launch = tst[f"ms_{dom}"] / n_l
    achieved = bytes_per_launch / (ms_per_launch / 1e3) / 1e9 if ms_per_launch > 0 else 0.0
    log(f"dominant kernel {dom}: {n_l} launches, {ms_per_launch:.3f} ms avg, "
        f"{bytes_per_launch / 1e9:.3f} GB alg/launch -> {achieved:.1f} GB/s; "
        f"modes per hop {tst['mode_hop']} ms per hop {[round(x, 2) for x in tst['ms_hop']]}")

    cpu = None
    if not args.no_cpu:
        import oracle
        ncpu = min(args.cpu_sources, args.sources)
        sample = src[:ncpu]
        threads = int(os.environ.get("OMP_NUM_THREADS", os.cpu_count() or 1))
        log(f"CPU baseline: building destination-partitioned Hilbert blocks ({threads} threads)")
        tb = time.time()
        blocks = oracle.EdgeBlocks(csr.offsets, csr.adj, nblocks=1000)
        log(f"blocks built in {time.time() - tb:.1f}s; running oracle B on {ncpu} sources")
        dv, ka_cpu, _, cpu_ms = blocks.vexpand(sample, args.k_min, args.k_max, args.path)
        rs, sst = vs.vexpand(G, sample, args.k_min, args.k_max, **kw)
        gpu_dv = rs.export(0, vs.LAYOUT_DVSM)
        exact = bool(np.array_equal(gpu_dv, dv)) and ka_cpu == sst["k_actual_max"]
        rs.close()
        del gpu_dv, dv
        blocks.close()
        cpu_gteps = sst["te"] / (cpu_ms / 1e3) / 1e9
        log(f"CPU {cpu_ms / 1e3:.2f}s for {ncpu} sources (TE={sst['te']}): {cpu_gteps:.2f} GTEPS;"
            f" GPU==CPU bit-exact: {exact}")
        if not exact:
            log("PARITY FAILURE between GPU and CPU baseline on the sample")
        cpu = {"value": round(cpu_gteps, 3), "unit": "GTEPS", "cores": threads,
               "kind": "port",
               "sample": f"{ncpu} of {args.sources} sources, hops {args.k_min}..{args.k_max}, "
                         f"same graph; TE of the sample / CPU time",
               "bit_exact_vs_gpu": exact}

    out = {
        "metric": "VExpand GTEPS, bit-exact, R-MAT-24 4096 src; batched+learned vs plain GPU",
        "value": round(gteps, 2),
        "unit": "GTEPS",
        "n_gpus": 1,
        "steps": args.steps,
        "warmup": args.warmup,
        "ms_per_step": round(ms_step, 3),
        "higher_is_better": True,
        "scaling": "weak",
        "vs_baseline": None,
        "dtype": "u64",
        "data": "synthetic",
        "config": {
            "workload": f"rmat{args.scale}_ef{args.edge_factor}_s{args.sources}_"
                        f"hops{args.k_min}-{args.k_max}_{args.path}",
            "scale": args.scale, "edge_factor": args.edge_factor, "vertices": csr.V,
            "edges": csr.E, "sources": args.sources, "k_min": args.k_min,
            "k_max": args.k_max, "path": args.path, "sched": args.sched,
            "tile_words": tst["tile_words"], "te": te, "k_actual_max": tst["k_actual_max"],
            "reachable_pairs": int(counts.sum()),
        },
        "roofline": {
            "bound": "hbm", "kernel": f"k_{dom}", "achieved": round(achieved, 1),
            "peak": HBM_PEAK_GBS, "unit": "GB/s", "frac": round(achieved / HBM_PEAK_GBS, 4),
            "traffic": None, "launches": n_l, "avg_ms": round(ms_per_launch, 4),
            "alg_bytes_per_launch": int(bytes_per_launch),
        },
        "cpu_baseline": cpu,
    }
    print(json.dumps(out), flush=True)
    return 0


if __name__ == "__main__":
    sys.exit(main())
