"""Mining benchmark: extended embeddings/s and mining wall time, batched vs plain, 1 MI355X.

A *step* is one full mining run (gs_mine: seeds, every extension level, MNI supports,
canonicality checks, ordered report) over one synthetic graph already resident in HBM.
Prints ONE JSON line.

Workloads (seeded, synthetic, reference binary-CSR layout).  The R-MAT generator is
(a,b,c,d) = (.57,.19,.19,.05) with self-loops removed and pairs deduplicated; the support
threshold of a scaled-down graph is scaled with its vertex count.

  rmat18b   (default) the batched configuration's graph -- 100 vertex labels Zipf(1.0),
            4 edge labels -- at R-MAT scale 18, edge factor 8; min-support 500 / 4 = 125;
            k = 2, the deepest level the full-size configuration can mine (level 3 holds
            ~4e13 rows at scale 20)
  rmat18    the plain-path configuration's graph (20 labels Zipf(1.1), 2 edge labels) at
            scale 18, min-support 2000 / 4 = 500, k = 2
  rmat18d   the learned-predictor configuration's graph (50 labels Zipf(1.2), 4 edge
            labels, edge factor 16) at scale 18, min-support 5000 / 16 = 312, k = 2
  rmat14b3  the batched configuration's distributions at scale 14, min-support 8, k = 3
            (a level below the last, where the learned sizes apply)
  er10k     Erdos-Renyi 10,000 vertices / 40,000 edges, 5 uniform labels, 1 edge label,
            min-support 200, k = 4 (the CPU-runnable configuration)
  rmat20, rmat20b, rmat22   the full-size R-MAT configurations; their graphs are only used
            by --ingest here (an R-MAT scale-20 mining run once faulted the GPU with an
            earlier version of the extension kernels, and its cause has not been shown)
"""
from __future__ import annotations

import argparse
import json
import os
import sys
import time

ROOT = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, ROOT)

WORKLOADS = {
    "rmat18b": dict(kind="rmat", scale=18, edge_factor=8, n_vlabels=100, vlabel_zipf=1.0,
                    n_elabels=4, min_sup=125, k=2,
                    of="batched R-MAT scale 20, min-support 500 (k = 2: its deepest level)"),
    "rmat18": dict(kind="rmat", scale=18, edge_factor=8, n_vlabels=20, vlabel_zipf=1.1,
                   n_elabels=2, min_sup=500, k=2,
                   of="plain-path R-MAT scale 20, min-support 2000"),
    "rmat18d": dict(kind="rmat", scale=18, edge_factor=16, n_vlabels=50, vlabel_zipf=1.2,
                    n_elabels=4, min_sup=312, k=2,
                    of="predictor R-MAT scale 22, min-support 5000"),
    "rmat14b3": dict(kind="rmat", scale=14, edge_factor=8, n_vlabels=100, vlabel_zipf=1.0,
                     n_elabels=4, min_sup=8, k=3),
    "rmat20": dict(kind="rmat", scale=20, edge_factor=8, n_vlabels=20, vlabel_zipf=1.1,
                   n_elabels=2, min_sup=2000, k=2),
    "rmat20b": dict(kind="rmat", scale=20, edge_factor=8, n_vlabels=100, vlabel_zipf=1.0,
                    n_elabels=4, min_sup=500, k=2),
    "rmat22": dict(kind="rmat", scale=22, edge_factor=16, n_vlabels=50, vlabel_zipf=1.2,
                   n_elabels=4, min_sup=5000, k=2),
    "er10k": dict(kind="er", n=10000, m=40000, n_vlabels=5, vlabel_zipf=None, n_elabels=1,
                  min_sup=200, k=4),
}
HBM_PEAK_GBS = 8000.0
# HBM bytes per launch of the dominant kernel, measured by rocprofv3 FETCH_SIZE and
# WRITE_SIZE runs of this bench's default steps and scaled by a plain copy's readings;
# profiles/traffic.json holds them per workload (absent: traffic null)
TRAFFIC_FILE = os.path.join(ROOT, "profiles", "traffic.json")


def log(msg: str) -> None:
    print(f"[bench] {msg}", file=sys.stderr, flush=True)


def make_graph(w: dict):
    from gspan_amd.graph import synthetic
    if w["kind"] == "rmat":
        return synthetic("rmat", scale=w["scale"], edge_factor=w["edge_factor"],
                         n_vlabels=w["n_vlabels"], vlabel_zipf=w["vlabel_zipf"],
                         n_elabels=w["n_elabels"], seed=1)
    return synthetic("er", n=w["n"], m=w["m"], n_vlabels=w["n_vlabels"],
                     n_elabels=w["n_elabels"], seed=1)


def physical_cores() -> int:
    """Physical cores of this host (distinct (core, socket) pairs in lscpu's table)."""
    import subprocess
    try:
        out = subprocess.run(["lscpu", "-p=CORE,SOCKET"], capture_output=True, text=True,
                             check=True).stdout
        n = len({l for l in out.splitlines() if l and not l.startswith("#")})
        return n or (os.cpu_count() or 1)
    except Exception:
        return os.cpu_count() or 1


def cpu_baseline(g, w: dict, k: int, target_s: float) -> dict:
    """The oracle (CPU/OpenMP restatement of the reference's path: extend_omp, support,
    is_canonical, the mine() recursion) mining the workload in the reference's order, five
    runs each cut at the first pattern visited after target_s / 5 seconds (the whole run
    when it ends first); the median rate of the five.  Threads: every physical core, or the
    OMP_NUM_THREADS share of this host when that is set and smaller."""
    sys.path.insert(0, os.path.join(ROOT, "oracle"))
    import oracle
    phys = physical_cores()
    share = int(os.environ.get("OMP_NUM_THREADS", "0") or 0)
    threads = min(phys, share) if share > 0 else phys
    rates, secs, cut = [], [], False
    for _ in range(5):
        t0 = time.perf_counter()
        _, st = oracle.fsm(g, k, w["min_sup"], nthreads=threads, max_seconds=target_s / 5)
        dt = time.perf_counter() - t0
        rates.append(st["extended"] / dt)
        secs.append(dt)
        cut |= bool(st["truncated"])
    rates.sort()
    sample = (f"median of 5 full mining runs of the workload (k={k})" if not cut else
              f"median of 5 runs of mining the workload in the reference's order, each cut "
              f"after {target_s / 5:.0f} s")
    return {"value": rates[2], "unit": "extended embeddings/s", "cores": threads,
            "physical_cores": phys, "kind": "port", "sample": sample,
            "runs_per_s": rates, "seconds": sum(secs)}


def ingest_bench(args) -> int:
    """Text -> CSR of the workload graph: GPU (gs_text_parse) vs the reference's convert."""
    import subprocess
    import tempfile
    from gspan_amd import graph as G
    w = dict(WORKLOADS[args.workload])
    g = make_graph(w)
    with tempfile.TemporaryDirectory() as td:
        p = os.path.join(td, "g.txt")
        G.to_txt(g, p)
        size = os.path.getsize(p)
        log(f"text file {size / 1e6:.1f} MB, {g.nr_edges // 2} edges")
        G.from_txt_gpu(p)  # warmup: context, code objects
        times = []
        for _ in range(max(args.steps, 1)):
            t = time.perf_counter()
            gg = G.from_txt_gpu(p)
            times.append(time.perf_counter() - t)
        assert gg.nr_nodes == g.nr_nodes and gg.nr_edges == g.nr_edges
        np_ok = bool((gg.edges == g.edges).all() and (gg.offsets == g.offsets).all())
        ref = None
        conv = os.path.join(ROOT, "oracle", "_ref", "convert")
        if os.path.exists(conv):
            t = time.perf_counter()
            subprocess.run([conv, p, os.path.join(td, "g.bin")], check=True,
                           stdout=subprocess.DEVNULL)
            ref_s = time.perf_counter() - t
            rg = G.CSRGraph.load(os.path.join(td, "g.bin"))
            ref = {"seconds": ref_s, "MB_per_s": size / ref_s / 1e6, "cores": 1,
                   "kind": "reference", "same_csr": bool((rg.edges == gg.edges).all())}
    best = min(times)
    print(json.dumps({"metric": "text ingestion MB/s (file read, parse, CSR build, copy back)",
                      "workload": args.workload, "bytes": size, "gpu_seconds": best,
                      "gpu_MB_per_s": size / best / 1e6, "same_as_generator": np_ok,
                      "cpu_baseline": ref}), flush=True)
    return 0


def main() -> int:
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=20)
    # warmup runs also let the launch-shape tuner try each shape once per bucket
    ap.add_argument("--warmup", type=int, default=8)
    ap.add_argument("--workload", default="rmat18b", choices=sorted(WORKLOADS))
    ap.add_argument("--k", type=int, default=0)
    ap.add_argument("--min-sup", type=int, default=0)
    ap.add_argument("--tile", type=int, default=0,
                    help="fixed tile size of the extension kernels (turns the tuner off)")
    ap.add_argument("--timed-mode", default="batched", choices=("batched", "plain"),
                    help="method of the timed steps (the metric is quoted for batched)")
    ap.add_argument("--groups", default="auto", choices=("auto", "on", "off"),
                    help="last level by neighbour runs and row groups, or by every neighbour")
    ap.add_argument("--two-pass", action="store_true",
                    help="count every level before filling it (no single-pass last level)")
    ap.add_argument("--two-edge", default="auto", choices=("auto", "off"),
                    help="two-edge last level from per-vertex neighbour runs (simple graphs), "
                         "or by extending the parent rows")
    ap.add_argument("--sets", default="auto", choices=("auto", "bits", "bytes"),
                    help="MNI image sets of the timed method: byte per vertex when it fits, "
                         "bitmaps, or bytemaps only")
    ap.add_argument("--no-plain", action="store_true")
    ap.add_argument("--no-predicted", action="store_true")
    ap.add_argument("--no-cpu", action="store_true")
    ap.add_argument("--cpu-seconds", type=float, default=20.0)
    ap.add_argument("--ingest", action="store_true",
                    help="only time text ingestion (Graph::from_txt + CSRGraph(Graph&)) of the "
                         "workload's graph on the GPU, beside the reference's own convert "
                         "(oracle/_ref/convert, when built)")
    ap.add_argument("--copy-calibration", type=int, default=0, metavar="MIB",
                    help="only copy MIB MiB device-to-device 5 times (a plain copy kernel of "
                         "known size, to scale rocprofv3 FETCH_SIZE/WRITE_SIZE readings)")
    args = ap.parse_args()
    if args.copy_calibration:
        import torch
        n = args.copy_calibration << 20
        a = torch.ones(n, dtype=torch.uint8, device="cuda:0")
        b = torch.empty_like(a)
        for _ in range(5):
            b.copy_(a)
        torch.cuda.synchronize()
        print(json.dumps({"copy_bytes_per_dispatch": n, "dispatches": 5}), flush=True)
        return 0
    if args.gpus != 1:
        print(f"bench.py measures one GPU: every box here has a single MI355X, so "
              f"--gpus {args.gpus} is not run (the path would shard by DFS subtree).",
              file=sys.stderr)
        return 2

    from gspan_amd.miner import Miner

    if args.ingest:
        return ingest_bench(args)

    w = dict(WORKLOADS[args.workload])
    if w["kind"] == "rmat" and w["scale"] >= 20:
        print(f"bench.py: mining {args.workload} on the GPU is not run: an R-MAT scale-20 "
              f"run faulted the GPU with an earlier version of the extension kernels and the "
              f"cause has not been shown (--ingest times text ingestion of its graph)",
              file=sys.stderr)
        return 2
    k = args.k or w["k"]
    ms = args.min_sup or w["min_sup"]
    t0 = time.perf_counter()
    g = make_graph(w)
    gstats = g.stats()
    log(f"graph {args.workload}: {gstats} ({time.perf_counter() - t0:.1f}s)")

    t_create = time.perf_counter()
    with Miner(g, mode=args.timed_mode, timing=True, tile=args.tile, groups=args.groups,
               two_pass=args.two_pass, sets=args.sets, two_edge=args.two_edge) as mn:
        # gs_create: device copy and indexes of the graph (label-ordered adjacency, runs,
        # two-edge run masks and pairs), before any step
        create_ms = (time.perf_counter() - t_create) * 1e3
        for i in range(args.warmup):
            t = time.perf_counter()
            res = mn.mine_raw(k, ms)
            log(f"warmup {i}: {res[1].shape[0]} patterns, {time.perf_counter() - t:.3f}s")
        mn.reset_stats()
        times = []
        for i in range(args.steps):
            t = time.perf_counter()
            res = mn.mine_raw(k, ms)  # synchronous: returns after the last device copy
            times.append(time.perf_counter() - t)
            log(f"step {i}: {times[-1]:.3f}s")
        st = mn.stats()
        tuning = mn.tuner_report()
    pats = Miner.to_patterns(*res)
    steps = args.steps
    ext_per_step = st["extended"] / steps
    total_s = sum(times)
    value = st["extended"] / total_s

    # dominant extension kernel: per-launch algorithmic bytes / per-launch time, the time
    # from HIP events the library records around each launch on its own stream
    fill_ms, count_ms = st["t_fill_ms"], st["t_count_ms"]
    if st["two_edge_levels"] > 0 and st["two_edge_levels"] == st["fill_launches"]:
        kname, kms, kbytes, klaunch = "k_te_accum + k_te_keys (two-edge run pass)", fill_ms, \
            st["alg_bytes_fill"], st["fill_launches"]
    elif fill_ms >= count_ms:
        kname, kms, kbytes, klaunch = "k_items fill (MNI sets / child rows)", fill_ms, \
            st["alg_bytes_fill"], st["fill_launches"]
    else:
        kname, kms, kbytes, klaunch = "k_items count", count_ms, st["alg_bytes_count"], \
            st["count_launches"]
    achieved = (kbytes / klaunch) / (kms / klaunch * 1e-3) / 1e9 if kms > 0 else 0.0
    traffic = None
    if os.path.exists(TRAFFIC_FILE) and not args.k and not args.min_sup:
        for rec in json.load(open(TRAFFIC_FILE)).get(args.workload, []):
            if rec.get("kernel") == kname:
                traffic = rec["bytes_per_launch"]
    roofline = {"bound": "hbm", "achieved": round(achieved, 1), "peak": HBM_PEAK_GBS,
                "unit": "GB/s", "frac": round(achieved / HBM_PEAK_GBS, 4), "traffic": traffic,
                "kernel": kname, "launches_per_step": klaunch / steps,
                "avg_launch_ms": kms / max(klaunch, 1),
                "alg_bytes_per_launch": kbytes / max(klaunch, 1)}

    # the same batched method with the last level extended from the parent rows, as every
    # other level (the two-edge pass off)
    generic = None
    if st["two_edge_levels"] > 0:
        with Miner(g, mode=args.timed_mode, timing=True, tile=args.tile, groups=args.groups,
                   sets=args.sets, two_edge="off") as mo:
            for _ in range(3):
                mo.mine_raw(k, ms)
            mo.reset_stats()
            gt = []
            for _ in range(5):
                t = time.perf_counter()
                gres = mo.mine_raw(k, ms)
                gt.append(time.perf_counter() - t)
            gst = mo.stats()
        if Miner.to_patterns(*gres) != pats:
            log("ERROR: two-edge and generic last levels differ")
            return 1
        generic = {"value": gst["extended"] / sum(gt), "ms_per_step": sum(gt) / len(gt) * 1e3}
        log(f"two-edge pass off: {generic['ms_per_step']:.2f} ms/step")

    plain = None
    if not args.no_plain:
        with Miner(g, mode="plain", groups=args.groups) as mp:
            t = time.perf_counter()
            pres = mp.mine_raw(k, ms)
            pdt = time.perf_counter() - t
            pst = mp.stats()
        ppats = Miner.to_patterns(*pres)
        if ppats != pats:
            log("ERROR: plain and batched results differ")
            return 1
        plain = {"value": pst["extended"] / pdt, "ms_per_step": pdt * 1e3,
                 "batches": pst["batches"]}
        log(f"plain: {pdt:.3f}s")

    predicted = None
    if not args.no_predicted:
        # learned sizes: log exact counts of one two-pass run, train the predictor on them
        # with the library's own kernels, then mine with predicted sizes (exact fallback)
        with Miner(g, mode="batched", two_pass=True) as ml:
            ml.predictor_log(True, 1 << 20)
            ml.mine_raw(k, ms)
            samples = ml.predictor_samples()
            t = time.perf_counter()
            loss = ml.predictor_train(epochs=300, lr=3e-3)
            train_s = time.perf_counter() - t
            weights = ml.predictor_weights()
        with Miner(g, mode="predicted", timing=True, groups=args.groups) as mq:
            mq.set_predictor_weights(weights)
            for _ in range(args.warmup):
                mq.mine_raw(k, ms)
            mq.reset_stats()
            qt = []
            for _ in range(args.steps):
                t = time.perf_counter()
                qres = mq.mine_raw(k, ms)
                qt.append(time.perf_counter() - t)
            qst = mq.stats()
        if Miner.to_patterns(*qres) != pats:
            log("ERROR: predicted and batched results differ")
            return 1
        predicted = {"value": qst["extended"] / sum(qt), "ms_per_step": sum(qt) / len(qt) * 1e3,
                     "predicted_steps_per_step": qst["predicted_steps"] / args.steps,
                     "fallbacks_per_step": qst["fallbacks"] / args.steps,
                     "items_per_step": qst["items"] / args.steps,
                     "train_samples": samples, "train_loss": loss, "train_s": train_s}
        if k <= 2:
            predicted["note"] = ("learned sizes apply to levels below the last; k = 2 has none, "
                                 "so this run is the batched method, timed right after the "
                                 "predictor's training run (a batched miner made at that point "
                                 "runs as slowly)")
        log(f"predicted: {predicted['ms_per_step']:.2f} ms/step")

    cpu = None
    if not args.no_cpu:
        log("cpu baseline ...")
        cpu = cpu_baseline(g, w, k, args.cpu_seconds)
        log(f"cpu baseline: {cpu['value']:.3e} /s over {cpu['seconds']:.1f}s")

    out = {
        "metric": "extended embeddings/s + mining wall-time, batched vs plain, 1 MI355X",
        "value": value,
        "unit": "extended embeddings/s",
        "n_gpus": 1,
        "steps": steps,
        "warmup": args.warmup,
        "ms_per_step": total_s / steps * 1e3,
        "higher_is_better": True,
        "scaling": "weak",
        "vs_baseline": None,
        "dtype": "int32",
        "data": "synthetic (seeded generator, reference binary-CSR layout)",
        "config": {
            "workload": f"{args.workload}: "
                        + (f"R-MAT scale {w['scale']} ef {w['edge_factor']}, "
                           if w["kind"] == "rmat" else f"ER n={w['n']} m={w['m']}, ")
                        + f"{w['n_vlabels']} vlabels"
                        + (f" Zipf({w['vlabel_zipf']})" if w["vlabel_zipf"] else " uniform")
                        + f", {w['n_elabels']} elabels, min_sup {ms}, k {k}",
            "downsized_from": w.get("of"),
            "k": k, "min_support": ms, "mode": args.timed_mode, "mni": "gpu",
            "groups": args.groups, "two_pass": args.two_pass, "sets": args.sets,
            "two_edge": args.two_edge,
            "graph": gstats,
        },
        "patterns": len(pats),
        "extended_per_step": ext_per_step,
        "create_ms": round(create_ms, 1),
        "mining_wall_s": total_s / steps,
        "plain": plain,
        "batched_vs_plain": (value / plain["value"]) if plain else None,
        "two_edge_off": generic,
        "predicted": predicted,
        "predicted_vs_plain": (predicted["value"] / plain["value"]) if (plain and predicted) else None,
        "roofline": roofline,
        "autotune": tuning,
        "cpu_baseline": cpu,
        "peak_device_bytes": st["peak_device_bytes"],
        # host time per step: waiting at synchronisations, planning children, ordering reports
        "host_ms_per_step": {k: round(st[f"t_{k}_ms"] / steps, 3) for k in ("sync", "plan", "order")},
    }
    print(json.dumps(out), flush=True)
    return 0


if __name__ == "__main__":
    sys.exit(main())
