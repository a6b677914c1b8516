"""Benchmark: harmonic-mean GTEPS over 64 roots on RMAT scale 24, learned
direction policy vs Beamer (alpha 15, beta 18), one MI355X.

A search is one breadth-first search from one root (hybrid_bfs::run_beamer on
the device, clear() untimed), its time the device time from the start of the
search (the source pushed, the first direction chosen) to the end of its last
level (the reference times run_beamer alone, hybrid_bfs_main.cc:267-282).
TEPS = count_num_traversed_edges() / time (hybrid_bfs.cc:401-414).  The roots
are the reference's first 64 lcg(0) picks (hybrid_bfs_main.cc:136-144), every
pass.  A step of this benchmark is one pass of the reference's trial loop
(hybrid_bfs_main.cc:252-306) over those 64 roots; --steps passes are timed
after --warmup untimed ones, the two policies alternating pass by pass.  The
metric aggregates every timed search as aggregate_stats does
(combined.cc:108-146): harmonic mean, mean, sample stddev, min, max.  After
timing, every root is searched again with each policy and the result checked
(hybrid_bfs::check, as --check_results does, hybrid_bfs_main.cc:283-292).
The graph is generated on the GPU with the build's seeded RMAT generator and
is resident in HBM before timing starts.

Prints one JSON line.  Beside the metric it carries the step kernel's
roofline, the CPU baselines (one core and all cores), what a perfect choice
of one step's direction could gain over Beamer (direction_bound), the 64-root
batch and connected components (components::run) on the same resident
graph.
"""
import argparse
import json
import os
import re
import sys
import time

import numpy as np

ROOT = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, ROOT)

METRIC = "harmonic-mean GTEPS over 64 roots, RMAT s24; learned vs Beamer policy"
HBM_PEAK_GBS = 8000.0
MODEL = os.path.join(ROOT, "beedrill_amd", "models", "direction_mlp_32x32.f32")
TRAFFIC = os.path.join(ROOT, "profiles", "step_kernel_traffic_rmat_s24.json")
# The step kernel (both directions) as rocprofv3 names it:
# k_step<workgroup size, 2>.
STEP_KERNEL_RE = r"k_step<\d+, 2>"
# Connected components: the edge pass (pull or push), and its traffic file.
CC_KERNEL_RE = r"k_cc_(hook|push|tile)"
CC_TRAFFIC = os.path.join(ROOT, "profiles", "cc_edge_pass_traffic_rmat_s24.json")
COPY_BYTES = 1 << 30  # calibration copy: 1 GiB read + 1 GiB written


def log(msg):
    print(f"[bench] {msg}", file=sys.stderr, flush=True)


NROOTS = 64  # the metric's roots: the first 64 lcg(0) picks


def hmean(x):
    x = np.asarray(x, np.float64)
    return float(len(x) / np.sum(1.0 / x))


def aggregate(x):
    """aggregate_stats::compute (combined.cc:108-146): harmonic mean, mean,
    sample standard deviation (N - 1), min and max."""
    x = np.asarray(x, np.float64)
    return {"n": int(len(x)), "hmean": hmean(x), "mean": float(np.mean(x)),
            "stddev": float(np.std(x, ddof=1)) if len(x) > 1 else 0.0,
            "min": float(np.min(x)), "max": float(np.max(x))}


def search_pass(ctx, roots, policy):
    """One pass of the trial loop: (GTEPS per root, device ms per root)."""
    g, ms = [], []
    for r in roots:
        st = ctx.bfs(int(r), 15, 18, policy)
        g.append(st["traversed_edges"] / (st["ms"] * 1e-3) / 1e9)
        ms.append(st["ms"])
    return g, ms


def run_policies(ctx, roots, policies, passes, warmup):
    """warmup untimed passes, then `passes` timed passes over the same roots,
    the policies taking turns pass by pass (so drift over the run touches
    both alike)."""
    for _ in range(warmup):
        for pol in policies:
            search_pass(ctx, roots, pol)
    res = {pol: {"gteps": [], "ms": [], "pass_ms": [], "pass_hmean": []} for pol in policies}
    t0 = time.perf_counter()
    for _ in range(passes):
        for pol in policies:
            g, ms = search_pass(ctx, roots, pol)
            r = res[pol]
            r["gteps"] += g
            r["ms"] += ms
            r["pass_ms"].append(float(np.sum(ms)))
            r["pass_hmean"].append(hmean(g))
    wall = time.perf_counter() - t0
    out = {}
    for pol, r in res.items():
        agg = aggregate(r["gteps"])
        ms = np.asarray(r["ms"])
        out[pol] = {"hmean_gteps": agg["hmean"], "mean_gteps": agg["mean"],
                    "stddev_gteps": agg["stddev"], "min_gteps": agg["min"],
                    "max_gteps": agg["max"], "searches": agg["n"],
                    "mean_ms": float(ms.mean()),
                    "ms_per_pass": float(np.mean(r["pass_ms"])),
                    "sum_edges_over_sum_time_gteps": float(
                        np.sum(np.asarray(r["gteps"]) * ms) / ms.sum()),
                    "pass_hmean_min": float(np.min(r["pass_hmean"])),
                    "pass_hmean_max": float(np.max(r["pass_hmean"]))}
    return out, wall


def check_roots(ctx, roots, policies):
    """hybrid_bfs::check on every root, once per policy (untimed), as the
    reference's --check_results does after each trial
    (hybrid_bfs_main.cc:283-292)."""
    failed = []
    for pol in policies:
        for r in roots:
            ctx.bfs(int(r), 15, 18, pol)
            if not ctx.check(int(r)):
                failed.append([pol, int(r)])
    return {"roots": len(roots), "policies": list(policies), "failed": failed,
            "pass": not failed}


def direction_bound(ctx, roots, policy, budget_s):
    """How much any direction policy could gain on these roots: per root the
    device time of the search with Beamer's choices and with `policy` (best
    of 2 each), and of every single-step flip of Beamer's choices (one run
    each), over as many roots as fit in budget_s.  best = the fastest of
    them per root.  Flipping an early bottom-up step to top-down, or a top-
    down step to bottom-up, can cost tens of milliseconds: those are the
    slow searches of this leg."""
    t0 = time.perf_counter()
    rows = []
    for r in roots:
        r = int(r)
        st = ctx.bfs(r, 15, 18, "beamer")
        b = min(st["ms"], ctx.bfs(r, 15, 18, "beamer")["ms"])
        dirs = [x["dir"] for x in ctx.levels()]
        lt = min(ctx.bfs(r, 15, 18, policy)["ms"] for _ in range(2))
        flips = []
        for L in range(len(dirs)):
            sc = list(dirs)
            sc[L] = 1 - sc[L]
            ctx.set_script(np.array(sc, np.int32))
            flips.append(ctx.bfs(r, 15, 18, "script")["ms"])
        rows.append((st["traversed_edges"], b, lt, min([b, lt] + flips)))
        if time.perf_counter() - t0 > budget_s:
            break
    g = {k: hmean([e / (t * 1e-3) / 1e9 for e, *ts in rows for t in [ts[i]]])
         for i, k in enumerate(("beamer", "policy", "best"))}
    return {"roots": len(rows), "policy": policy,
            "beamer_gteps": round(g["beamer"], 3), "policy_gteps": round(g["policy"], 3),
            "best_single_flip_gteps": round(g["best"], 3),
            "policy_over_beamer": round(g["policy"] / g["beamer"], 4),
            "best_over_beamer": round(g["best"] / g["beamer"], 4),
            "seconds": round(time.perf_counter() - t0, 1)}


def bu_bytes(rec, nv, n_isolated):
    """Algorithmic bytes of one bottom-up step: visited-bitmap scan (V/8),
    both offsets of every unvisited vertex with edges (16 U), the neighbour ids
    read up to the first hit (4 E), one frontier bit per id read (E/8),
    the parent of each newly reached vertex (4 N) and the next-frontier
    bitmap (V/8)."""
    U = rec["unvisited"] - n_isolated
    E = rec["edges_examined"]
    return nv / 8 + 16 * U + 4 * E + E / 8 + 4 * rec["next"] + nv / 8


def td_bytes(rec):
    """Algorithmic bytes of one top-down step: frontier ids (4 |F|), their
    offsets (16 |F|), the neighbour ids (4 sum deg F), one visited probe per
    neighbour (4 sum deg F), and parent and queue entry of each newly
    reached vertex (8 |N|)."""
    F, D = rec["frontier"], rec["frontier_deg"]
    return 4 * F + 16 * F + 4 * D + 4 * D + 8 * rec["next"]


def roofline(ctx, roots, policy, nv, n_isolated):
    """The step kernel against the HBM roofline.  Searches run device-driven,
    as in the timed region, with HIP events on the library's stream spanning
    each chunk of back-to-back step launches (bd_tuning.time_steps = 4): the
    mean launch is the quantity rocprofv3 reports for the kernel.  achieved =
    the algorithmic bytes of all steps / the time of all launches (the
    launches queued after a search ended count, with no bytes).  bu_levels:
    a second pass timing every launch alone (time_steps = 3) gives the
    bottom-up steps' own rate; those per-launch events add a few
    microseconds each, so it is a lower bound."""
    t = ctx.get_tuning()
    tot_b = tot_ms = bu_b = bu_ms = 0.0
    launches = bu_n = 0
    try:
        ctx.set_tuning(time_steps=4)
        for r in roots:
            ctx.bfs(int(r), 15, 18, policy)
            lt = ctx.launch_times()
            launches += len(lt)
            tot_ms += float(lt.sum())
            for rec in ctx.levels():
                tot_b += bu_bytes(rec, nv, n_isolated) if rec["dir"] == 1 else td_bytes(rec)
        ctx.set_tuning(time_steps=3)
        for r in roots:
            ctx.bfs(int(r), 15, 18, policy)
            for rec in ctx.levels():  # rec["ms"]: its launches' times
                if rec["dir"] == 1:
                    bu_b += bu_bytes(rec, nv, n_isolated)
                    bu_ms += rec["ms"]
                    bu_n += 1
    finally:
        ctx.set_tuning(time_steps=t["time_steps"])
    achieved = tot_b / (tot_ms * 1e-3) / 1e9 if tot_ms > 0 else 0.0
    bu_ach = bu_b / (bu_ms * 1e-3) / 1e9 if bu_ms > 0 else 0.0
    traffic = None
    if os.path.exists(TRAFFIC):
        with open(TRAFFIC) as f:
            traffic = json.load(f).get("bytes_per_launch")
    return {"kernel": f"k_step<{t['block']}, 2>", "bound": "hbm", "achieved": round(achieved, 1),
            "peak": HBM_PEAK_GBS, "unit": "GB/s", "frac": round(achieved / HBM_PEAK_GBS, 4),
            "traffic": traffic, "launches": launches,
            "algorithmic_bytes_per_launch": round(tot_b / max(launches, 1)),
            "avg_launch_ms": round(tot_ms / max(launches, 1), 5),
            "bu_levels": {"achieved": round(bu_ach, 1), "frac": round(bu_ach / HBM_PEAK_GBS, 4),
                          "launches": bu_n, "avg_launch_ms": round(bu_ms / max(bu_n, 1), 5)}}


def copy_calibration(ctx):
    """A plain device copy of known size (the library's k_copy), for scaling
    the FETCH_SIZE and WRITE_SIZE counters when this leg runs under
    rocprofv3 --pmc; returns its best rate (read + written GB/s)."""
    return ctx.copy_bandwidth(COPY_BYTES, reps=3)


def summarize_pmc(fetch_csv, write_csv, out_path, kernel_re=STEP_KERNEL_RE):
    """Traffic per launch of the kernel matching kernel_re from two rocprofv3
    --pmc runs of a leg (FETCH_SIZE, then WRITE_SIZE; KiB per dispatch), each
    scaled by what the counter reads for the calibration copy of known
    size."""
    import csv

    def per_kernel(path, counter):
        rows = {}
        with open(path) as f:
            for row in csv.DictReader(f):
                if row.get("Counter_Name", counter) != counter:
                    continue
                name = row["Kernel_Name"]
                rows.setdefault(name, []).append(float(row["Counter_Value"]))
        return rows

    out = {}
    for counter, path, copy_bytes in (("FETCH_SIZE", fetch_csv, COPY_BYTES),
                                      ("WRITE_SIZE", write_csv, COPY_BYTES)):
        rows = per_kernel(path, counter)
        bu = [v for k, vs in rows.items() if re.search(kernel_re, k) for v in vs]
        cp = [v for k, vs in rows.items() if "k_copy" in k for v in vs]
        if not bu or not cp:
            raise SystemExit(f"{path}: no {kernel_re} or copy dispatches for {counter}")
        scale = copy_bytes / (float(np.median(cp)) * 1024.0)
        out[counter] = {"kib_per_launch": float(np.mean(bu)), "launches": len(bu),
                        "copy_kib": float(np.median(cp)), "scale": scale,
                        "bytes_per_launch": float(np.mean(bu)) * 1024.0 * scale}
    out["bytes_per_launch"] = round(out["FETCH_SIZE"]["bytes_per_launch"] +
                                    out["WRITE_SIZE"]["bytes_per_launch"])
    out["kernel"] = sorted(k for k in rows if re.search(kernel_re, k))
    with open(out_path, "w") as f:
        json.dump(out, f, indent=1)
    print(json.dumps(out))


def batched_roots(ctx, roots, sequential_ms):
    """The same 64 roots in one bit-parallel search, against one pass of the
    64 sequential searches (the mean over the timed passes of the sum of
    their device times, which write one parent array per root).  Timed twice:
    writing every root's depth row (64 rows of V entries) and without (the
    search alone); best of 3 each.  levels: push (0) or pull (1) per level,
    and each level's time from a host-driven run."""
    ctx.multi_root(roots[:8], want_depths=False)  # warm up
    ms_d = min(ctx.multi_root(roots, want_depths=True)[1] for _ in range(3))
    ms_s = min(ctx.multi_root(roots, want_depths=False)[1] for _ in range(3))
    _, _, levels = ctx.multi_root(roots, want_depths=False)
    t = ctx.get_tuning()
    try:
        ctx.set_tuning(level_driver=2)
        ctx.multi_root(roots, want_depths=False)
        per_level = [{"mode": "push" if r["dir"] == 0 else "pull", "frontier": r["frontier"],
                      "ms": round(r["ms"], 3)} for r in ctx.levels()]
    finally:
        ctx.set_tuning(level_driver=t["level_driver"])
    return {"roots": len(roots), "ms": round(ms_d, 4), "search_only_ms": round(ms_s, 4),
            "levels": int(levels), "sequential_ms": round(sequential_ms, 4),
            "speedup_vs_sequential": round(sequential_ms / ms_d, 3) if ms_d > 0 else None,
            "search_only_speedup": round(sequential_ms / ms_s, 3) if ms_s > 0 else None,
            "per_level_host_driven": per_level}


def hmean_policy(ctx, roots, policy):
    g = []
    for r in roots:
        st = ctx.bfs(int(r), 15, 18, policy)
        g.append(st["traversed_edges"] / (st["ms"] * 1e-3) / 1e9)
    return hmean(g)


def leg_batches(ctx, args):
    """Batched searches: (a) 64 roots of RMAT scale 23 in one bit-parallel
    search against the 64 searches one by one; (b) 256 RMAT graphs of uneven
    size (scale uniform in 10-16, edge factor uniform in 8-32, generator
    seeds 1000+g) searched from one root each at once, against one search
    per graph."""
    out = []
    ctx.generate_rmat(23, 16, seed=args.seed)
    roots = ctx.pick_roots(64)
    for r in roots[:2]:
        ctx.bfs(int(r))
    seq = sum(ctx.bfs(int(r))["ms"] for r in roots)
    ctx.multi_root(roots[:8], want_depths=False)
    _, ms, lv = ctx.multi_root(roots, want_depths=True)
    out.append({"batch": "64 roots, RMAT scale 23, bit-parallel", "ms": round(ms, 3),
                "sequential_ms": round(seq, 3), "speedup": round(seq / ms, 3), "levels": lv})
    log(f"64 roots s23: {ms:.2f} ms batched vs {seq:.2f} ms sequential")
    rng = np.random.default_rng(7)
    graphs, roots, seq_ms = [], [], 0.0
    for g in range(256):
        scale, ef = int(rng.integers(10, 17)), int(rng.integers(8, 33))
        ctx.generate_rmat(scale, ef, seed=1000 + g)
        off, cols = ctx.csr()
        _, _, m = ctx.info()
        root = int(ctx.pick_roots(1)[0])
        ctx.bfs(root)
        seq_ms += ctx.bfs(root)["ms"]
        graphs.append((off, cols, m))
        roots.append(root)
    vbase = ctx.upload_union(graphs)
    ctx.multi_graph(vbase, roots, want_depths=False)
    _, ms = ctx.multi_graph(vbase, roots, want_depths=True)
    edges = ctx.count_traversed()
    out.append({"batch": "256 RMAT graphs, scales 10-16, edge factors 8-32, one root each",
                "vertices": int(vbase[-1]), "edges": int(sum(m for _, _, m in graphs)),
                "ms": round(ms, 3), "sequential_ms": round(seq_ms, 3),
                "speedup": round(seq_ms / ms, 3), "gteps": round(edges / (ms * 1e-3) / 1e9, 2)})
    log(f"256 graphs: {ms:.2f} ms at once vs {seq_ms:.2f} ms one by one")
    return out


def leg_stress(ctx, args, have_model):
    """Same-size stress graphs with the learned policy judged against Beamer:
    RMAT scale 26, an Erdos-Renyi graph of the same size (uniform endpoints,
    16 * 2^26 edges) and the 4096 x 4096 four-neighbour lattice.  One root of
    each is checked in full (hybrid_bfs::check on the host)."""
    out = []
    for name, gen, nroots in (("RMAT scale 26", lambda: ctx.generate_rmat(26, 16, seed=args.seed), 16),
                              ("Erdos-Renyi, 2^26 vertices, 2^30 edges",
                               lambda: ctx.generate_er(1 << 26, 16 << 26, seed=args.seed), 16),
                              ("lattice 4096 x 4096", lambda: ctx.generate_lattice(4096, 4096), 2)):
        t0 = time.perf_counter()
        m = gen()
        build = time.perf_counter() - t0
        roots = ctx.pick_roots(nroots)
        ctx.bfs(int(roots[0]))
        rec = {"graph": name, "num_vertices": ctx.info()[0], "num_edges": m,
               "build_s": round(build, 2), "roots": nroots,
               "beamer_gteps": round(hmean_policy(ctx, roots, "beamer"), 3)}
        st = ctx.bfs(int(roots[0]))
        rec["levels"] = int(st["steps"])
        rec["check_root0"] = ctx.check(int(roots[0]))
        if have_model:
            rec["learned_gteps"] = round(hmean_policy(ctx, roots, "learned"), 3)
            rec["learned_vs_beamer"] = round(rec["learned_gteps"] / rec["beamer_gteps"], 4)
        log(json.dumps(rec))
        out.append(rec)
    return out


def leg_components(ctx, args):
    """components::run (components.cc:38-104) on the same RMAT graph: device
    time of each run (label initialisation, edge passes, shortcuts and the
    count; the reference times run() alone, components_main.cc:170-180),
    over args.cc_runs runs after one warm-up.  The roofline is the edge pass
    (k_cc_hook pulls every label -- after the first pass as the long lists'
    slices plus eight XCD-tiled k_cc_tile phases -- and k_cc_push offers the
    labels of the vertices that changed), timed per pass with HIP events on
    the library's stream; avg_launch_ms is that time over the passes'
    kernel launches.  Algorithmic bytes of a pull: the column indices (4 B
    per adjacency entry), the labels read (4 B each), the offsets and label
    of every vertex (12 B); of a push: the changed-vertex bitmap (V/8), the
    offsets and label of the changed vertices (12 B each, bounded by 12 B
    per vertex), their column indices and the labels they read (4 + 4 B per
    label read).  The reference's rate is reported too:
    num_iters * num_edges / time (components_main.cc:179)."""
    t0 = time.perf_counter()
    m = ctx.generate_rmat(args.scale, args.edge_factor, seed=args.seed)
    gen_s = time.perf_counter() - t0
    out = components_measure(ctx, args, m)
    out = {"leg": "components",
           "metric": f"components::run time, RMAT scale {args.scale}", **out}
    out["config"] = {"workload": f"RMAT scale {args.scale}, edge factor {args.edge_factor}, "
                                 f"A/B/C/D .57/.19/.19/.05, seed {args.seed}",
                     "num_vertices": ctx.info()[0], "num_edges": m, "nnz": ctx.info()[1],
                     "graph_build_s": round(gen_s, 2)}
    return out


def components_measure(ctx, args, m):
    """components::run on the resident graph (see leg_components)."""
    nv, nnz, _ = ctx.info()
    ctx.components()  # warm-up; also builds the per-graph slice table
    runs = [ctx.components() for _ in range(args.cc_runs)]
    ok = ctx.components_check()
    ms = [r["ms"] for r in runs]
    passes = sum(r["num_iters"] for r in runs)
    launches = sum(r["pass_launches"] for r in runs)
    pass_ms = sum(r["pass_ms"] for r in runs)
    alg = sum(r["edge_bytes"] for r in runs)
    achieved = alg / (pass_ms * 1e-3) / 1e9
    traffic = None
    if os.path.exists(CC_TRAFFIC):
        with open(CC_TRAFFIC) as f:
            traffic = json.load(f).get("bytes_per_launch")
    copy_gbs = copy_calibration(ctx)
    mean_ms = float(np.mean(ms))
    cpu = None
    if not args.no_cpu:
        import oracle
        off, adj = ctx.csr()
        t0 = time.perf_counter()
        lab, n, it = oracle.components(off, adj)
        dt = time.perf_counter() - t0
        assert n == runs[0]["num_components"], (n, runs[0]["num_components"])
        cpu = {"value": round(dt * 1e3, 1), "unit": "ms", "cores": 1, "kind": "port",
               "sample": f"one full serial oracle components::run on the same graph "
                         f"({it} passes), labels equal to the device's count"}
    return {
        "value": round(mean_ms, 3), "unit": "ms", "higher_is_better": False,
        "runs": args.cc_runs, "min_ms": round(min(ms), 3),
        "num_components": runs[0]["num_components"],
        "num_iters": [r["num_iters"] for r in runs],
        "reference_gteps": round(float(np.mean([r["num_iters"] * m / (r["ms"] * 1e-3) / 1e9
                                                 for r in runs])), 2),
        "check": ok, "dtype": "int32",
        "roofline": {"kernel": "k_cc_hook + k_cc_tile + k_cc_push", "bound": "hbm",
                     "achieved": round(achieved, 1),
                     "peak": HBM_PEAK_GBS, "unit": "GB/s",
                     "frac": round(achieved / HBM_PEAK_GBS, 4), "traffic": traffic,
                     "passes": passes, "launches": launches,
                     "algorithmic_bytes_per_launch": round(alg / launches),
                     "avg_launch_ms": round(pass_ms / launches, 4),
                     "labels_read_per_launch": round(sum(r["labels_read"] for r in runs) / launches)},
        "copy_gbs": round(copy_gbs, 1),
        # built once per graph by the warm-up run, outside `value`: slices of
        # the long lists and the short lists regrouped by column block
        "regroup_once": {"ms": round(runs[0]["prep_ms"], 3),
                         "bytes": int(runs[0]["prep_bytes"])},
        "cpu_baseline": cpu,
    }


def cpu_baseline(ctx, roots, budget_s):
    import oracle
    off, adj = ctx.csr()
    _, _, m = ctx.info()
    gteps, used = [], []
    t_all = time.perf_counter()
    for r in roots:
        t0 = time.perf_counter()
        parent = oracle.beamer_timed(off, adj, m, int(r))
        dt = time.perf_counter() - t0
        gteps.append(oracle.count_traversed(off, parent) / dt / 1e9)
        used.append(int(r))
        if time.perf_counter() - t_all > budget_s:
            break
    serial_s = time.perf_counter() - t_all
    # the same restatement on every host core (OpenMP), over the same roots
    pg, t_par = [], time.perf_counter()
    for r in roots:
        t0 = time.perf_counter()
        parent, _ = oracle.beamer_omp(off, adj, m, int(r))
        dt = time.perf_counter() - t0
        pg.append(oracle.count_traversed(off, parent) / dt / 1e9)
        if time.perf_counter() - t_par > budget_s:
            break
    threads = oracle.omp_threads()
    return {"value": round(hmean(gteps), 4), "unit": "GTEPS", "cores": 1, "kind": "port",
            "sample": f"{len(used)} of the same lcg(0) roots on the same graph, serial oracle "
                      f"run_beamer (harmonic mean), {serial_s:.1f} s",
            "parallel": {"value": round(hmean(pg), 4), "unit": "GTEPS", "cores": threads,
                         "kind": "port",
                         "sample": f"{len(pg)} of the same roots, oracle run_beamer on {threads} "
                                   f"OpenMP threads (harmonic mean), "
                                   f"{time.perf_counter() - t_par:.1f} s"}}


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=10,
                    help="timed passes of the trial loop over the 64 roots, per policy")
    ap.add_argument("--warmup", type=int, default=2, help="untimed passes before them")
    ap.add_argument("--scale", type=int, default=24)
    ap.add_argument("--edge-factor", type=int, default=16)
    ap.add_argument("--seed", type=int, default=0)
    ap.add_argument("--roofline-roots", type=int, default=8)
    ap.add_argument("--cpu-budget", type=float, default=20.0)
    ap.add_argument("--bound-budget", type=float, default=40.0,
                    help="seconds for the direction-choice bound (roots in order until spent)")
    ap.add_argument("--no-cpu", action="store_true")
    ap.add_argument("--model", default=MODEL, help="direction network parameters (float32)")
    ap.add_argument("--cc-runs", type=int, default=5, help="timed runs of --leg components")
    ap.add_argument("--leg", choices=("all", "roofline", "batches", "stress", "components"),
                    default="all",
                    help="roofline: only the roofline searches and the calibration copy (the "
                         "command profiled with rocprofv3); batches: 64 roots at scale 23 and "
                         "256 uneven graphs; stress: RMAT scale 26, Erdos-Renyi and lattice; "
                         "components: connected components on the same graph (one JSON line)")
    ap.add_argument("--summarize-pmc", nargs=2, metavar=("FETCH_CSV", "WRITE_CSV"),
                    help="write the traffic file of the step kernel (--leg roofline runs) or, "
                         "with --leg components, of the components edge pass, from two "
                         "rocprofv3 --pmc runs, and exit")
    args = ap.parse_args()
    if args.summarize_pmc:
        if args.leg == "components":
            summarize_pmc(args.summarize_pmc[0], args.summarize_pmc[1], CC_TRAFFIC, CC_KERNEL_RE)
        else:
            summarize_pmc(args.summarize_pmc[0], args.summarize_pmc[1], TRAFFIC)
        return
    if args.gpus != 1:
        print(f"bench.py runs on one GPU only; --gpus {args.gpus} is not supported "
              "(the search is not sharded across GPUs in this build)", file=sys.stderr)
        sys.exit(2)

    import beedrill_amd as bd
    ctx = bd.Context(0)
    if args.leg in ("batches", "stress"):
        have_model = os.path.exists(args.model)
        if have_model:
            ctx.mlp_set(np.fromfile(args.model, np.float32), 32, 32)
            meta = os.path.splitext(args.model)[0] + ".json"
            if os.path.exists(meta):
                with open(meta) as f:
                    mm = json.load(f).get("margin_milli")
                if mm:
                    ctx.set_tuning(learned_margin_milli=int(mm))
        res = leg_batches(ctx, args) if args.leg == "batches" else leg_stress(ctx, args, have_model)
        print(json.dumps({"leg": args.leg, "results": res}))
        return
    if args.leg == "components":
        print(json.dumps(leg_components(ctx, args)))
        return
    t0 = time.perf_counter()
    m = ctx.generate_rmat(args.scale, args.edge_factor, seed=args.seed)
    gen_s = time.perf_counter() - t0
    nv, nnz, _ = ctx.info()
    log(f"graph: scale {args.scale}, {nv} vertices, {m} edges, built in {gen_s:.2f} s")
    roots = ctx.pick_roots(NROOTS)
    # degree-0 vertices are never scanned by a bottom-up step
    off_h = np.empty(nv + 1, np.int64)
    ctx._ok(ctx._L.bd_graph_download(ctx._h, off_h.ctypes.data, None))
    n_isolated = int(np.sum(np.diff(off_h) == 0))
    have_model = os.path.exists(args.model)
    margin = None
    if have_model:
        ctx.mlp_set(np.fromfile(args.model, np.float32), 32, 32)
        meta = os.path.splitext(args.model)[0] + ".json"
        if os.path.exists(meta):
            with open(meta) as f:
                margin = json.load(f).get("margin_milli")
        if margin:
            ctx.set_tuning(learned_margin_milli=int(margin))
    policy = "learned" if have_model else "beamer"

    if args.leg == "roofline":
        # warm-up searches are timed the same way, so that every step launch
        # of this process is in the average (as in rocprofv3's)
        roof = roofline(ctx, list(roots[:args.warmup]) + list(roots[:args.roofline_roots]),
                        policy, nv, n_isolated)
        copy_gbs = copy_calibration(ctx)
        print(json.dumps({"leg": "roofline", "roofline": roof, "copy_gbs": round(copy_gbs, 1)}))
        return

    policies = ("beamer", "learned") if have_model else ("beamer",)
    results, timed_s = run_policies(ctx, roots, policies, args.steps, args.warmup)
    for pol in policies:
        r = results[pol]
        log(f"{pol}: {r['hmean_gteps']:.2f} GTEPS harmonic mean over {r['searches']} searches "
            f"(stddev {r['stddev_gteps']:.1f}, {r['ms_per_pass']:.3f} ms per pass of "
            f"{len(roots)} roots)")
    head = results[policy]
    checked = check_roots(ctx, roots, policies)
    log(f"check: {'PASS' if checked['pass'] else 'FAIL'} on {len(roots)} roots x "
        f"{len(policies)} policies")
    bound = direction_bound(ctx, roots, policy, args.bound_budget) if have_model else None
    if bound:
        log(f"direction bound over {bound['roots']} roots: {policy} {bound['policy_over_beamer']}x, "
            f"best single flip {bound['best_over_beamer']}x Beamer ({bound['seconds']} s)")
    batch = batched_roots(ctx, roots, results["beamer"]["ms_per_pass"])
    log(f"64 roots batched: {batch['ms']:.3f} ms with depth rows, {batch['search_only_ms']:.3f} ms "
        f"search alone, vs {batch['sequential_ms']:.3f} ms sequential; levels "
        f"{[(x['mode'], x['ms']) for x in batch['per_level_host_driven']]}")
    roof = roofline(ctx, roots[:args.roofline_roots], policy, nv, n_isolated)
    log(f"roofline: {roof['achieved']} GB/s over {roof['launches']} step launches; bottom-up "
        f"steps {roof['bu_levels']['achieved']} GB/s")
    cpu = None if args.no_cpu else cpu_baseline(ctx, roots, args.cpu_budget)
    cc = components_measure(ctx, args, m)
    log(f"components: {cc['value']} ms, {cc['num_components']} components, passes "
        f"{cc['num_iters']}")

    out = {
        "metric": METRIC,
        "value": round(head["hmean_gteps"], 3),
        "unit": "GTEPS",
        "n_gpus": 1,
        "steps": args.steps,
        "warmup": args.warmup,
        "ms_per_step": round(head["ms_per_pass"], 4),
        "higher_is_better": True,
        "scaling": "weak",
        "vs_baseline": None,
        "dtype": "int32",
        "data": "synthetic: seeded RMAT generated on the GPU (reference generator stream, "
                "build-defined relabel/shuffle)",
        "config": {"workload": f"RMAT scale {args.scale}, edge factor {args.edge_factor}, "
                               f"A/B/C/D .57/.19/.19/.05, seed {args.seed}, {len(roots)} lcg(0) "
                               f"roots per step (one pass of the trial loop), alpha 15 beta 18, "
                               f"policy {policy}",
                   "roots": len(roots), "searches_per_policy": head["searches"],
                   "learned_margin_milli": ctx.get_tuning()["learned_margin_milli"],
                   "num_vertices": nv, "num_edges": m, "nnz": nnz,
                   "graph_build_s": round(gen_s, 2)},
        "aggregate": {"hmean": round(head["hmean_gteps"], 3),
                      "mean": round(head["mean_gteps"], 3),
                      "stddev": round(head["stddev_gteps"], 3),
                      "min": round(head["min_gteps"], 3), "max": round(head["max_gteps"], 3),
                      "n": head["searches"]},
        "check": checked["pass"],
        "checked": checked,
        "timed_s": round(timed_s, 3),
        "roofline": roof,
        "cpu_baseline": cpu,
        "batched_64_roots": batch,
        "policies": results,
        "learned_vs_beamer": (round(results["learned"]["hmean_gteps"] /
                                    results["beamer"]["hmean_gteps"], 4) if have_model else None),
        "direction_bound": bound,
        "components": {"metric": "components::run time on the same graph", **cc},
    }
    print(json.dumps(out))


if __name__ == "__main__":
    main()
