#!/usr/bin/env python3
"""Benchmark of the Louvain hot path on one MI355X.

Workload: RMAT scale 22 (4,194,304 vertices, edge factor 16, a/b/c/d =
.57/.19/.19/.05, seed 22), self-loops and duplicates removed, symmetrised,
fp32 weights on the 2^-24 grid in (0,1], seeded vertex permutation.

A step is one plain sweep of the level-0 phase: the moves of the 7 degree
groups with their applies, and the modularity (with its host readback, the
termination test of the reference's sweep loop).  Steps run through the
level-0 phase from singletons (4 sweeps on this graph) and start it again
when the phase ends; a restart (community state reset and the modularity
before the first sweep) is inside the timed step.  value = CSR entries per
second over the timed level-0 sweeps.  Also reported: time to converge (full
multi-level run, the handle's second: its first, which also allocates the
aggregation's buffers, is reported apart) with per-level sweep times, and the
CPU oracle timed on the
same level-0 phase on 1 core and on all host cores, with its time to converge
on RMAT-20.  The steps and the run are timed on
a handle without per-kernel events; the same warmup and timed steps are then
repeated on a handle that records HIP events around every kernel on its
stream, which give the per-kernel timings and the roofline of the dominant
kernel (ms_per_step_instrumented is that pass's step time).

With --level N the step is one sweep of level N (reached by plain phases
and aggregations) from that level's singleton state, value = that level's
CSR entries per second.  Also reported by default: the autotuner's choice,
the learned-pruned vs plain comparison and the 4,096-graph batch.

Prints ONE JSON line on stdout; progress goes to stderr.
"""
import argparse
import json
import os
import sys
import time

ROOT = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, os.path.join(ROOT, "gpu-louvain_amd"))

HBM_PEAK_GBS = 8000.0


def log(msg):
    print(msg, file=sys.stderr, flush=True)


def cpu_model():
    """The host CPU's model name (as lscpu prints it) and logical CPU count."""
    name = None
    try:
        with open("/proc/cpuinfo") as f:
            for line in f:
                if line.startswith("model name"):
                    name = line.split(":", 1)[1].strip()
                    break
    except OSError:
        pass
    return name, os.cpu_count()


def host_threads():
    """Host threads for the all-cores oracle run: OMP_NUM_THREADS when set
    (the GPU box's share), else every CPU."""
    env = os.environ.get("OMP_NUM_THREADS")
    return max(1, int(env)) if env and env.isdigit() else (os.cpu_count() or 1)


def cpu_baseline(L, csr, threads, converge_scale=20):
    """The oracle on the host, timed on the GPU step's own work: the level-0
    phase from singletons (its sweeps with the modularity after each, the
    termination test of the reference's loop), once on 1 core and once on
    `threads` cores (same results: decisions inside a degree group are
    independent); then the oracle's whole multi-level run of RMAT-`converge_scale`
    (seed = scale) on `threads` cores, time to converge."""
    sys.path.insert(0, os.path.join(ROOT, "oracle"))
    import numpy as np

    import oracle as O

    og = O.Graph.from_csr(csr)
    M = O.total_weight(og)
    legs = {}
    comms = {}
    for t in (1, threads):
        comm = np.arange(csr.V, dtype=np.int32)
        o2c = np.arange(csr.V, dtype=np.int32)
        t0 = time.perf_counter()
        _, sweeps = O.phase(og, comm, 1e-6, o2c, M, threads=t)
        dt = time.perf_counter() - t0
        legs[t] = {"value": csr.E * sweeps / dt, "seconds": dt, "sweeps": sweeps, "threads": t}
        comms[t] = comm
        log(f"cpu baseline: level-0 phase on {t} thread(s): {sweeps} sweeps in {dt:.2f} s, "
            f"{legs[t]['value']:.3e} edges/s")
    same = bool(np.array_equal(comms[1], comms[threads]))
    g2 = L.gen_rmat(converge_scale, 16, 0.57, 0.19, 0.19, seed=converge_scale)
    st, _, _ = O.run(O.Graph.from_csr(g2), 1e-6, threads=threads)
    log(f"cpu baseline: RMAT-{converge_scale} converged in {st['seconds']:.2f} s on {threads} threads "
        f"({st['levels']} levels, {st['sweeps']} sweeps, Q={st['modularity']:.6f})")
    model, ncpu = cpu_model()
    top = legs[threads]
    return {"value": top["value"], "unit": "edges/s", "cores": threads, "kind": "port",
            "sample": f"the oracle's level-0 phase of the same RMAT-{csr_scale(csr)} graph from singletons "
                      f"({top['sweeps']} sweeps with the modularity after each), the GPU step's work",
            "cores_1": legs[1], "cores_all": top, "identical_results": same,
            "cpu_model": model, "host_cpus": ncpu,
            "time_to_converge_ms": st["seconds"] * 1e3,
            "converge_workload": f"RMAT-{converge_scale} seed {converge_scale}, {threads} threads, "
                                 f"{st['levels']} levels, {st['sweeps']} sweeps, Q={st['modularity']:.6f}",
            "converge_graph": {"V": g2.V, "E": g2.E}}


def gpu_converge(L, scale):
    """The library's time to converge on RMAT-`scale` (seed = scale), the
    second of two runs."""
    g = L.gen_rmat(scale, 16, 0.57, 0.19, 0.19, seed=scale)
    with L.Graph(g) as G:
        G.run(1e-6)
        G.reset()
        st = G.run(1e-6)
    return {"time_to_converge_ms": st["algo_ms"], "modularity": st["modularity"], "levels": st["levels"],
            "sweeps": st["sweeps"], "V": g.V, "E": g.E}


def calibrate_copy():
    """A plain device-to-device copy of 2^28 fp32 (1 GiB read, 1 GiB written),
    three times, for scaling the FETCH_SIZE / WRITE_SIZE counters."""
    import torch

    n = 1 << 28
    x = torch.ones(n, dtype=torch.float32, device="cuda")
    y = torch.empty_like(x)
    for _ in range(3):
        y.copy_(x)
    torch.cuda.synchronize()
    print(json.dumps({"copy_bytes_read": 4 * n, "copy_bytes_written": 4 * n, "launches": 3}))


def pruned_comparison(L, csr, min_gain=1e-6):
    """Learned-pruned vs plain on the same graph: the model is trained on the
    device from plain-run traces of 8 seeded RMAT-18 graphs (seeds 1801-1808),
    its threshold set for 97% recall of movers on a held-out RMAT-18 (seed
    1809); training is reported apart from the run times."""
    import numpy as np

    import frontier as F

    t0 = time.perf_counter()
    train = [L.gen_rmat(18, 16, 0.57, 0.19, 0.19, seed=s) for s in range(1801, 1809)]
    X, y, tinfo = F.collect_traces(train, min_gain=min_gain)
    params, losses, info = F.train(X, y)
    val = L.gen_rmat(18, 16, 0.57, 0.19, 0.19, seed=1809)
    Xv, yv, _ = F.collect_traces([val], min_gain=min_gain, per_sweep=1 << 30, seed=1)
    thr, frac = F.recall_threshold(params, Xv, yv, recall=0.97)
    train_s = time.perf_counter() - t0
    log(f"model trained in {train_s:.1f}s: {tinfo['samples']} samples, loss {info['loss_first']:.4f} -> "
        f"{info['loss_last']:.4f}, threshold {thr:.3f} selects {frac:.3f}")
    out = {"training": {"graphs": "RMAT-18 seeds 1801-1808", "samples": tinfo["samples"],
                        "positives": tinfo["positives"], "trace_s": tinfo["seconds"], "train_s": info["seconds"],
                        "epochs": info["epochs"], "loss_first": info["loss_first"], "loss_last": info["loss_last"],
                        "threshold": thr, "val_selected": frac, "total_s": train_s}}
    with L.Graph(csr) as G:
        G.set_model(params)
        G.run(min_gain)   # warm
        runs = {"plain": None,
                "learned": L.frontier(kind=L.FRONTIER_MODEL, threshold=thr, full_every=4),
                "neighbor": L.frontier(kind=L.FRONTIER_NEIGHBOR, full_every=4),
                "random": L.frontier(kind=L.FRONTIER_RANDOM, random_p=min(1.0, frac), seed=7, full_every=4)}
        res = {}
        for name, fr in runs.items():
            G.reset()
            st = G.run(min_gain) if fr is None else G.run_pruned(min_gain, fr)
            res[name] = {"modularity": st["modularity"], "time_ms": st["algo_ms"], "sweeps": st["sweeps"],
                         "full_sweeps": st["full_sweeps"], "evaluated": st["evaluated"],
                         "levels": st["levels"], "sweep_ms": st["sweep_ms"], "aggregate_ms": st["aggregate_ms"]}
            log(f"{name}: Q={st['modularity']:.6f} {st['algo_ms']:.1f} ms, {st['sweeps']} sweeps "
                f"({st['full_sweeps']} full), {st['evaluated']} evaluations")
        qp, tp, ep = res["plain"]["modularity"], res["plain"]["time_ms"], res["plain"]["evaluated"]
        for name in ("learned", "neighbor", "random"):
            r = res[name]
            r["dQ_rel"] = (qp - r["modularity"]) / qp
            r["time_ratio"] = r["time_ms"] / tp
            r["evaluated_ratio"] = r["evaluated"] / max(ep, 1)
        out.update(res)
        out["target"] = {"dQ_rel_max": 0.005, "time_ratio_max": 0.5,
                         "met": bool(res["learned"]["dQ_rel"] <= 0.005 and res["learned"]["time_ratio"] <= 0.5)}
    out["selection"] = selection_quality(L, csr, params, thr)
    log(f"selection on the level-0 phase: model {out['selection']['model_selected']:.3f} of vertices, recall "
        f"{out['selection']['model_recall']:.3f}; rule {out['selection']['rule_selected']:.3f}, recall "
        f"{out['selection']['rule_recall']:.3f}")
    return out, (params, thr)


def selection_quality(L, csr, params, thr, sweeps=8):
    """The model's frontier against the neighbour rule on the same states:
    over the sweeps of the plain level-0 phase after the first, the share of
    vertices each selects and the share of that sweep's movers it catches."""
    import numpy as np

    ALL = L.frontier(kind=L.FRONTIER_ALL)
    tot = dict(moved=0, n=0, m_sel=0, m_hit=0, r_sel=0, r_hit=0)
    with L.Graph(csr) as G:
        G.sweep_pruned(ALL, True)
        for _ in range(sweeps):
            ms = G.move_state()
            feat = G.features()
            c0 = G.communities().copy()
            _, mv, _ = G.sweep_pruned(ALL, True)
            moved = G.communities() != c0
            p = L.mlp_forward(params, feat)
            has_edges = feat[:, 0] > 0
            msel = (p > thr) & has_edges
            rsel = (((ms["hist"] & 1) > 0) | (ms["nmv"] > 0)) & has_edges
            tot["moved"] += int(moved.sum())
            tot["n"] += int(has_edges.sum())
            tot["m_sel"] += int(msel.sum())
            tot["m_hit"] += int((msel & moved).sum())
            tot["r_sel"] += int(rsel.sum())
            tot["r_hit"] += int((rsel & moved).sum())
            if mv == 0:
                break
    n, m = max(tot["n"], 1), max(tot["moved"], 1)
    return {"model_selected": tot["m_sel"] / n, "model_recall": tot["m_hit"] / m,
            "rule_selected": tot["r_sel"] / n, "rule_recall": tot["r_hit"] / m, "movers": tot["moved"]}


def c5_comparison(L, min_gain=1e-6):
    """Plain vs learned-pruned on RMAT-24 (seed 24) with the model retrained
    on plain-run traces of 8 RMAT-20 graphs (seeds 2001-2008), threshold for
    97% recall on seed 2009; the neighbour rule beside them."""
    import frontier as F

    t0 = time.perf_counter()
    train = [L.gen_rmat(20, 16, 0.57, 0.19, 0.19, seed=s) for s in range(2001, 2009)]
    X, y, tinfo = F.collect_traces(train, min_gain=min_gain, per_sweep=4096)
    del train
    params, losses, info = F.train(X, y)
    Xv, yv, _ = F.collect_traces([L.gen_rmat(20, 16, 0.57, 0.19, 0.19, seed=2009)], min_gain=min_gain,
                                 per_sweep=1 << 16, seed=1)
    thr, frac = F.recall_threshold(params, Xv, yv, recall=0.97)
    train_s = time.perf_counter() - t0
    log(f"c5: model trained on RMAT-20 traces in {train_s:.1f}s ({tinfo['samples']} samples), threshold {thr:.3f}")
    g = L.gen_rmat(24, 16, 0.57, 0.19, 0.19, seed=24)
    out = {"workload": f"RMAT-24 seed 24 (V={g.V}, E={g.E})", "training": {
        "graphs": "RMAT-20 seeds 2001-2008", "samples": tinfo["samples"], "threshold": thr,
        "val_selected": frac, "total_s": train_s}}
    with L.Graph(g) as G:
        G.set_model(params)
        G.run(min_gain)  # warm
        runs = {"plain": None,
                "learned": L.frontier(kind=L.FRONTIER_MODEL, threshold=thr, full_every=4),
                "neighbor": L.frontier(kind=L.FRONTIER_NEIGHBOR, full_every=4)}
        for name, fr in runs.items():
            G.reset()
            st = G.run(min_gain) if fr is None else G.run_pruned(min_gain, fr)
            out[name] = {"modularity": st["modularity"], "time_ms": st["algo_ms"], "sweeps": st["sweeps"],
                         "full_sweeps": st["full_sweeps"], "evaluated": st["evaluated"], "levels": st["levels"]}
            log(f"c5 {name}: Q={st['modularity']:.6f} {st['algo_ms']:.1f} ms, {st['sweeps']} sweeps, "
                f"{st['evaluated']} evaluations")
    qp, tp, ep = out["plain"]["modularity"], out["plain"]["time_ms"], out["plain"]["evaluated"]
    for name in ("learned", "neighbor"):
        r = out[name]
        r["dQ_rel"] = (qp - r["modularity"]) / qp
        r["time_ratio"] = r["time_ms"] / tp
        r["evaluated_ratio"] = r["evaluated"] / max(ep, 1)
    out["selection"] = selection_quality(L, g, params, thr)
    return out


def batch_comparison(L, n_graphs=4096, seed=4, sample=64, min_gain=1e-6, model=None, tune=True):
    """4,096 independent planted-partition graphs solved as one batch: sizes
    log-uniform in [200, 50000], blocks uniform in 4..64, p_in for mean
    in-block degree 12, p_out for mean cross-block degree 2, unit weights.
    Compared with running the first `sample` graphs one by one (lv_run),
    whose results must equal the batch's.  With model = (params, threshold)
    the batch also runs learned-pruned (lv_batch_run_pruned, full sweep every
    4), checked the same way against lv_run_pruned on the sample."""
    import numpy as np

    rng = np.random.default_rng(seed)
    t0 = time.perf_counter()
    gs = []
    for _ in range(n_graphs):
        n = int(np.exp(rng.uniform(np.log(200), np.log(50000))))
        b = int(rng.integers(4, 65))
        p_in = min(1.0, 12.0 / max(n / b - 1.0, 1.0))
        p_out = 2.0 / max(n - n / b, 1.0)
        gs.append(L.gen_planted(n, b, p_in, p_out, seed=int(rng.integers(1 << 62)), unit_weights=True))
    gen_s = time.perf_counter() - t0
    V = sum(g.V for g in gs)
    E = sum(g.E for g in gs)
    log(f"batch: {n_graphs} planted graphs, V={V} E={E}, generated in {gen_s:.1f}s")
    tuned = None
    if tune:
        import autotune as T

        # the tuner times whole batched runs of a sample of the batch's graphs
        bt = T.tune_batch(gs[:256], cache_path=None)
        log(f"batch autotune: {bt['tried']} candidates on 256 graphs, {bt['default_ms']:.2f} -> "
            f"{bt['ms']:.2f} ms with {bt['options']}")
    with L.Batch(gs) as B:
        B.run(min_gain)  # warm
        per, tot = B.run(min_gain)
        comms = B.assignment()
        if tune:
            B.set_options(**bt["options"])
            B.run(min_gain)  # warm
            _, ttot = B.run(min_gain)
            tcomms = B.assignment()
            same = all(np.array_equal(a, b) for a, b in zip(comms, tcomms))
            tuned = {"options": bt["options"], "sample_default_ms": bt["default_ms"], "sample_tuned_ms": bt["ms"],
                     "candidates": bt["tried"], "default_ms": tot["algo_ms"], "tuned_ms": ttot["algo_ms"],
                     "gain": 1.0 - ttot["algo_ms"] / tot["algo_ms"], "identical_results": bool(same)}
            log(f"batch: tuned options {ttot['algo_ms']:.1f} ms vs {tot['algo_ms']:.1f} ms, identical {same}")
    log(f"batch: solved in {tot['algo_ms']:.1f} ms ({tot['levels']} levels, {tot['sweeps']} union sweeps)")
    single_ms = 0.0
    mismatches = 0
    for i in range(sample):
        with L.Graph(gs[i]) as G:
            G.run(min_gain)
            G.reset()
            st = G.run(min_gain)
            o2c, lv = G.assignment()
        single_ms += st["algo_ms"]
        if (abs(st["modularity"] - per[i]["modularity"]) > 1e-12 or not np.array_equal(o2c, comms[i])
                or lv != per[i]["final_V"]):
            mismatches += 1
    per_graph = single_ms / sample if sample else None
    est = per_graph * n_graphs if sample else None
    log(f"batch: one-by-one {per_graph or 0.0:.2f} ms per graph (first {sample}), "
        f"{est or 0.0:.0f} ms estimated for all; {mismatches} mismatches")
    learned = None
    if model is not None:
        params, thr = model
        fr = L.frontier(kind=L.FRONTIER_MODEL, threshold=thr, full_every=4)
        with L.Batch(gs) as B:
            B.set_model(params)
            B.run_pruned(min_gain, fr)  # warm
            lper, ltot = B.run_pruned(min_gain, fr)
            lcomms = B.assignment()
        lsingle_ms, lmis = 0.0, 0
        for i in range(sample):
            with L.Graph(gs[i]) as G:
                G.set_model(params)
                G.run_pruned(min_gain, fr)
                G.reset()
                st = G.run_pruned(min_gain, fr)
                o2c, lv = G.assignment()
            lsingle_ms += st["algo_ms"]
            if (abs(st["modularity"] - lper[i]["modularity"]) > 1e-12 or not np.array_equal(o2c, lcomms[i])
                    or lv != lper[i]["final_V"]):
                lmis += 1
        qp = np.array([r["modularity"] for r in per])
        ql = np.array([r["modularity"] for r in lper])
        dl = (qp - ql) / qp
        # the plain method's own spread under a change of move order: every
        # fourth graph renumbered by a seeded permutation, run plainly
        sys.path.insert(0, os.path.join(ROOT, "tests"))
        import graphs as TG

        idx = list(range(0, n_graphs, 4))
        perm = [L.CSR(*TG.relabelled(gs[i].rowptr, gs[i].col, gs[i].w, 500 + i)) for i in idx]
        with L.Batch(perm) as B:
            rper, _ = B.run(min_gain)
        qpi = qp[idx]
        dp = (qpi - np.array([r["modularity"] for r in rper])) / qpi

        def spread(d):
            return {"mean": float(d.mean()), "p90_abs": float(np.quantile(np.abs(d), 0.9)),
                    "p99": float(np.quantile(d, 0.99)), "max": float(d.max()), "min": float(d.min()),
                    "share_over_5pct": float(np.mean(d > 0.05))}

        learned = {"time_ms": ltot["algo_ms"], "time_ratio_vs_plain_batch": ltot["algo_ms"] / tot["algo_ms"],
                   "union_sweeps": ltot["sweeps"], "graph_sweeps": int(sum(r["sweeps"] for r in lper)),
                   "graph_full_sweeps": ltot["full_sweeps"], "evaluated": ltot["evaluated"],
                   "evaluated_plain": tot["evaluated"], "mean_modularity": float(ql.mean()),
                   "mean_dQ_rel": float(np.mean(dl)), "max_dQ_rel": float(np.max(dl)),
                   "dQ_rel": spread(dl), "relabelled_plain_dQ_rel": spread(dp),
                   "relabelled_graphs": len(idx),
                   "one_by_one_ms_per_graph": lsingle_ms / sample if sample else None,
                   "sample_mismatches": lmis}
        log(f"batch learned-pruned: {ltot['algo_ms']:.1f} ms ({learned['time_ratio_vs_plain_batch']:.3f}x plain "
            f"batch), mean dQ_rel {learned['mean_dQ_rel']:.4f}, {ltot['evaluated']} evaluations vs "
            f"{tot['evaluated']}; {lmis} mismatches vs one-by-one; per-graph dQ_rel {learned['dQ_rel']}, "
            f"relabelled plain {learned['relabelled_plain_dQ_rel']}")
    return {"learned_pruned": learned, "autotune": tuned, "graphs": n_graphs, "V": V, "E": E, "time_ms": tot["algo_ms"],
            "graphs_per_s": n_graphs / (tot["algo_ms"] * 1e-3), "edges_per_s": E / (tot["algo_ms"] * 1e-3),
            "union_levels": tot["levels"], "union_sweeps": tot["sweeps"],
            "aggregate_ms": tot["aggregate_ms"], "upload_ms": tot["memory_ms"],
            "mean_modularity": float(np.mean([r["modularity"] for r in per])),
            "one_by_one_ms_per_graph": per_graph, "one_by_one_estimate_ms": est,
            "speedup_vs_one_by_one": est / tot["algo_ms"] if est else None, "sample_checked": sample,
            "sample_mismatches": mismatches}


def traffic_for(kernel, scale):
    """Measured HBM bytes per launch of `kernel` on RMAT-`scale` from the
    committed rocprofv3 PMC summary (FETCH_SIZE + WRITE_SIZE, copy-calibrated),
    or None when no summary covers it."""
    path = os.path.join(ROOT, "profiles", f"pmc_traffic_rmat{scale}.md")
    if not os.path.exists(path):
        return None
    for line in open(path):
        # table rows: | kernel | fetch bytes | write bytes | traffic bytes | ...
        cells = [c.strip() for c in line.strip().strip("|").split("|")]
        if len(cells) >= 4 and cells[0] == f"`{kernel}`":
            try:
                return float(cells[3])
            except ValueError:
                return None
    return None


def io_comparison(L, csr, tmpdir=None):
    """Reading the graph from disk: the library's MatrixMarket reader on all
    host threads and on one thread (a child process with OMP_NUM_THREADS=1),
    and its binary CSR file, on the bench graph written as MatrixMarket."""
    import subprocess
    import tempfile

    import numpy as np

    d = tempfile.mkdtemp(dir=tmpdir)
    mtx, bin_ = os.path.join(d, "g.mtx"), os.path.join(d, "g.lvcsr")
    try:
        t = time.perf_counter()
        L.write_mtx(mtx, csr)
        write_s = time.perf_counter() - t
        L.write_csr(bin_, csr)
        t = time.perf_counter()
        a = L.read_mtx(mtx)
        par_s = time.perf_counter() - t
        t = time.perf_counter()
        b = L.read_csr(bin_)
        bin_s = time.perf_counter() - t
        code = ("import sys, time, hashlib; sys.path.insert(0, sys.argv[1]); import louvain_amd as L; "
                "t = time.perf_counter(); g = L.read_mtx(sys.argv[2]); dt = time.perf_counter() - t; "
                "h = hashlib.sha1(g.rowptr.tobytes() + g.col.tobytes() + g.w.tobytes()).hexdigest(); "
                "print(dt, h)")
        env = dict(os.environ, OMP_NUM_THREADS="1")
        r = subprocess.run([sys.executable, "-c", code, os.path.join(ROOT, "gpu-louvain_amd"), mtx],
                           env=env, capture_output=True, text=True, check=True)
        one_s, h1 = r.stdout.split()
        import hashlib

        hN = hashlib.sha1(a.rowptr.tobytes() + a.col.tobytes() + a.w.tobytes()).hexdigest()
        same = bool(h1 == hN and np.array_equal(b.col, csr.col) and np.array_equal(b.rowptr, csr.rowptr))
        out = {"mtx_bytes": os.path.getsize(mtx), "write_mtx_s": write_s, "read_mtx_s": par_s,
               "read_mtx_one_thread_s": float(one_s), "read_binary_csr_s": bin_s,
               "threads": int(os.environ.get("OMP_NUM_THREADS", os.cpu_count() or 1)),
               "identical_csr": same}
        log(f"io: {out['mtx_bytes'] / 1e9:.2f} GB MatrixMarket, read {par_s:.2f}s "
            f"({out['threads']} threads) vs {float(one_s):.2f}s (1 thread), binary CSR {bin_s:.2f}s, "
            f"identical {same}")
        return out
    finally:
        for f in (mtx, bin_):
            if os.path.exists(f):
                os.remove(f)
        os.rmdir(d)


def per_level(G, min_gain=1e-6):
    """The multi-level run sweep by sweep (the library's phase rule: a sweep
    that moves nothing, or gains less than min_gain or nothing, ends the
    phase), timed per sweep on the host clock around each synchronous sweep:
    per level its size, sweeps, time and edges/s, and the median level-0
    sweep."""
    import numpy as np

    G.reset()
    levels, lvl0 = [], []
    for lvl in range(64):
        V, E = G.level_size()
        q_prev = G.modularity()
        changed, ts = False, []
        while True:
            t = time.perf_counter()
            q, moved = G.sweep()
            ts.append((time.perf_counter() - t) * 1e3)
            gain = q - q_prev
            if gain > 0.0:
                changed = True
            q_prev = q
            if moved == 0 or not (gain > 0.0) or gain < min_gain:
                break
        if lvl == 0:
            lvl0 = ts
        ms = float(np.sum(ts))
        levels.append({"level": lvl, "V": V, "E": E, "sweeps": len(ts), "sweep_ms": ms,
                       "edges_per_s": E * len(ts) / (ms * 1e-3) if ms > 0 else None})
        if not changed:
            break
        G.aggregate()
    return levels, (float(np.median(lvl0)) if lvl0 else None)


def csr_scale(csr):
    return int(round(__import__("math").log2(max(csr.V, 1))))


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=20)
    ap.add_argument("--warmup", type=int, default=3)
    ap.add_argument("--scale", type=int, default=22)
    ap.add_argument("--seed", type=int, default=None)
    ap.add_argument("--cpu-threads", type=int, default=0,
                    help="threads of the all-cores oracle run (0 = OMP_NUM_THREADS or every CPU)")
    ap.add_argument("--no-cpu-baseline", action="store_true")
    ap.add_argument("--no-converge", action="store_true")
    ap.add_argument("--no-pruned", action="store_true")
    ap.add_argument("--no-batch", action="store_true")
    ap.add_argument("--no-tune", action="store_true")
    ap.add_argument("--batch-sample", type=int, default=64,
                    help="batch graphs also run one by one and compared (0 = none)")
    ap.add_argument("--level", type=int, default=0,
                    help="time sweeps of this level (reached by plain phases) from its singleton state")
    ap.add_argument("--c5", action="store_true",
                    help="also run plain vs learned-pruned on RMAT-24 with the model retrained on RMAT-20 traces")
    ap.add_argument("--io", action="store_true",
                    help="also time reading the graph from MatrixMarket and binary CSR files")
    ap.add_argument("--calibrate-copy", action="store_true",
                    help="run only a plain 1 GiB device copy (PMC byte-counter calibration)")
    args = ap.parse_args()
    if args.calibrate_copy:
        calibrate_copy()
        return
    if args.gpus != 1:
        print(f"bench.py: this build runs on one GPU only (asked for {args.gpus}); "
              f"multi-GPU sharding is a design note, not built", file=sys.stderr)
        sys.exit(2)

    import louvain_amd as L

    seed = args.seed if args.seed is not None else args.scale
    t0 = time.perf_counter()
    csr = L.gen_rmat(args.scale, 16, 0.57, 0.19, 0.19, seed=seed, unit_weights=False)
    log(f"generated RMAT-{args.scale}: V={csr.V} E={csr.E} maxdeg={int(csr.degrees().max())} "
        f"in {time.perf_counter() - t0:.1f}s")

    tuned = None
    if not args.no_tune:
        import autotune as T

        tuned = T.tune(csr, cache_path=None)
        log(f"autotune: {tuned['tried']} candidates, default {tuned['default_ms']:.3f} ms/sweep -> "
            f"{tuned['ms']:.3f} ms with {tuned['options']}")
    opts = tuned["options"] if tuned else {}

    def timed_pass(G, before_timed=None):
        """Warmup, then args.steps timed steps on handle G; returns
        (ms per step, Q after the last step, its movers, step edges)."""
        E_step = csr.E
        if args.level > 0:
            import numpy as np

            for _ in range(args.level):
                G.optimise(1e-6)
                G.aggregate()
            V_l, E_step = G.level_size()
            start = np.arange(V_l, dtype=np.int32)
            log(f"level {args.level}: V={V_l} E={E_step}")

            def step():
                G.set_communities(start)
                t = time.perf_counter()
                r = G.sweep()
                return r, time.perf_counter() - t
        else:
            phase = {"ended": True, "q": 0.0}

            def step():
                t = time.perf_counter()
                if phase["ended"]:
                    G.reset()
                    phase["q"] = G.modularity()
                q, moved = G.sweep()
                gain = q - phase["q"]
                phase["q"] = q
                phase["ended"] = moved == 0 or not (gain > 0.0) or gain < 1e-6
                return (q, moved), time.perf_counter() - t
        for i in range(args.warmup):
            step()
        if args.level == 0:
            phase["ended"] = True  # the timed steps start a fresh phase
        if before_timed:
            before_timed()
        q, moved, elapsed = 0.0, 0, 0.0
        for i in range(args.steps):
            (q, moved), dt = step()
            elapsed += dt
        return elapsed * 1e3 / max(args.steps, 1), q, moved, E_step

    # The throughput: a handle without per-kernel events (their records
    # between kernels would add host work and gaps to every step).
    G = L.Graph(csr, **opts)
    ms_per_step, q, moved, E_step = timed_pass(G)
    log(f"warmup {args.warmup} sweeps; timed {args.steps} sweeps: {ms_per_step:.3f} ms/sweep, "
        f"Q after sweep {q:.6f}, moved {moved}")

    converge = None
    if not args.no_converge:
        # the first run on the handle also allocates the aggregation's
        # buffers (reported apart); the time to converge is the next run's
        G.reset()
        st_first = G.run(1e-6)
        G.reset()
        st = G.run(1e-6)
        converge = {"time_to_converge_ms": st["algo_ms"], "first_run_ms": st_first["algo_ms"],
                    "modularity": st["modularity"],
                    "levels": st["levels"], "sweeps": st["sweeps"], "level0_sweeps": st["level0_sweeps"],
                    "level0_sweep_ms": st["level0_sweep_ms"], "sweep_ms": st["sweep_ms"],
                    "aggregate_ms": st["aggregate_ms"], "upload_ms": st["memory_ms"],
                    "edges_swept": st["edges_swept"],
                    "edges_per_s_in_sweeps": st["edges_swept"] / (st["sweep_ms"] * 1e-3) if st["sweep_ms"] else None}
        log(f"converged: Q={st['modularity']:.6f} in {st['algo_ms']:.1f} ms, {st['levels']} levels, "
            f"{st['sweeps']} sweeps")
        lv, med0 = per_level(G)
        converge["per_level"] = lv
        converge["level0_median_sweep_ms"] = med0
        converge["level0_median_edges_per_s"] = csr.E / (med0 * 1e-3) if med0 else None
        log("per level: " + ", ".join(f"L{x['level']} {x['sweeps']} sweeps {x['sweep_ms']:.1f} ms "
                                      f"({x['edges_per_s'] / 1e9:.1f} G edges/s)" for x in lv)
            + f"; level-0 median sweep {med0:.3f} ms")
    G.close()

    # Per-kernel durations: the same warmup and timed steps on a handle that
    # records HIP events around every kernel on its stream.
    G = L.Graph(csr, timing=True, **opts)
    ms_instrumented, _, _, _ = timed_pass(G, before_timed=G.kernel_timing_reset)
    log(f"instrumented pass: {ms_instrumented:.3f} ms/sweep")

    kernels = {}
    for kid, name in enumerate(L.KERNEL_NAMES):
        kt = G.kernel_timing(kid)
        if kt["launches"] == 0:
            continue
        avg_ms = kt["total_ms"] / kt["launches"]
        kernels[name] = {"launches": kt["launches"], "avg_ms": avg_ms,
                         "ms_per_sweep": kt["total_ms"] / args.steps,
                         "gbs": kt["bytes"] / (kt["total_ms"] * 1e-3) / 1e9 if kt["total_ms"] > 0 else None}
    details = {}
    for kid, name in enumerate(L.DETAIL_NAMES):
        kt = G.kernel_timing_detail(kid)
        if kt["launches"] == 0:
            continue
        details[name] = {"kernel": L.DETAIL_KERNELS[kid], "launches": kt["launches"],
                         "avg_ms": kt["total_ms"] / kt["launches"],
                         "ms_per_sweep": kt["total_ms"] / args.steps,
                         "bytes_per_launch": kt["bytes"] / kt["launches"],
                         "gbs": kt["bytes"] / (kt["total_ms"] * 1e-3) / 1e9 if kt["total_ms"] > 0 else None}
    # roofline of the single kernel with the most time among those whose
    # algorithmic bytes are counted
    counted = {k: v for k, v in details.items() if v["bytes_per_launch"] > 0}
    roofline = None
    if counted:
        dom = max(counted, key=lambda k: counted[k]["ms_per_sweep"])
        d = counted[dom]
        achieved = d["bytes_per_launch"] / (d["avg_ms"] * 1e-3) / 1e9
        roofline = {"kernel": d["kernel"], "bound": "hbm", "achieved": achieved, "peak": HBM_PEAK_GBS,
                    "unit": "GB/s", "frac": achieved / HBM_PEAK_GBS,
                    "traffic": traffic_for(d["kernel"], args.scale) if args.level == 0 else None,
                    "bytes_per_launch": d["bytes_per_launch"], "avg_launch_ms": d["avg_ms"],
                    "share_of_sweep": d["ms_per_sweep"] / ms_instrumented}
    if roofline is not None and args.level == 0:
        # the whole plain sweep against the same peak: 24 B per CSR entry and
        # 64 B per vertex of algorithmic bytes (moves 12E + 32V, apply 16V,
        # modularity 12E + 16V) over the uninstrumented step time
        sweep_bytes = 24.0 * csr.E + 64.0 * csr.V
        roofline["sweep_bytes"] = sweep_bytes
        roofline["sweep_achieved"] = sweep_bytes / (ms_per_step * 1e-3) / 1e9
        roofline["sweep_frac"] = roofline["sweep_achieved"] / HBM_PEAK_GBS

    G.close()

    pruned = model = None
    if not args.no_pruned:
        pruned, model = pruned_comparison(L, csr)
    batch = None
    if not args.no_batch:
        batch = batch_comparison(L, model=model, tune=not args.no_tune, sample=args.batch_sample)

    io = io_comparison(L, csr) if args.io else None
    c5 = c5_comparison(L) if args.c5 else None

    cpu = None
    if not args.no_cpu_baseline:
        threads = args.cpu_threads if args.cpu_threads > 0 else host_threads()
        cpu = cpu_baseline(L, csr, threads)
        cpu["gpu_time_to_converge_ms_same_graph"] = gpu_converge(L, 20)["time_to_converge_ms"]
        log(f"cpu baseline: {cpu['value']:.3e} edges/s on {threads} threads ({cpu['cpu_model']})")

    value = E_step / (ms_per_step * 1e-3)
    out = {
        "metric": "Louvain edges/s per sweep (plain, level 0) + time-to-converge",
        "value": value,
        "unit": "edges/s",
        "n_gpus": 1,
        "steps": args.steps,
        "warmup": args.warmup,
        "ms_per_step": ms_per_step,
        "ms_per_step_instrumented": ms_instrumented,
        "higher_is_better": True,
        "scaling": "weak",
        "vs_baseline": None,
        "dtype": "fp32 gains / fp64 sums / int32 indices",
        "data": "synthetic (seeded RMAT generator)",
        "config": {"workload": f"RMAT-{args.scale} ef16 a.57 b.19 c.19 seed {seed}, fp32 weights in (0,1]"
                               + (f", sweeps of level {args.level}" if args.level else ""),
                   "V": csr.V, "E": csr.E, "min_gain": 1e-6, "parallelism": "single GPU"},
        "roofline": roofline,
        "cpu_baseline": cpu,
        "kernels": kernels,
        "kernel_detail": details,
        "converge": converge,
        "pruned": pruned,
        "batch": batch,
        "autotune": tuned,
        "io": io,
        "c5": c5,
    }
    print(json.dumps(out))


if __name__ == "__main__":
    main()
