"""Benchmark: site-likelihood evaluations per second on one MI355X.

Workload (DNA GTR+I+G4, 200 taxa x 100,000 sites, ~60k unique patterns, 2% gap /
ambiguity cells, p_inv 0.2, Gamma alpha 0.5, Exp(0.1) branch lengths, seeded):
one step = one full post-order traversal with the tree lnL, then one
optimizeAllBranches sweep of Newton-Raphson over all 397 branches, starting
from the same initial branch lengths every step.

Unit: a site-likelihood evaluation is one unique pattern's likelihood under
one whole tree; every tree-level evaluation (the traversal lnL, each Newton
derivative evaluation, each lnL-from-buffer) counts the P patterns once.
value = P * evaluations / second, FP64 form (the drop-in). The FP32-storage
form's rate and its |delta lnL| against FP64 are reported beside it.
"""
import argparse
import json
import os
import sys
import time

ROOT = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, ROOT)

HBM_PEAK = 8.0e12          # MI355X HBM3E, bytes/s
FP64_PEAK = 78.6e12        # dense FP64, flop/s


def log(msg):
    print(f"[bench] {msg}", file=sys.stderr, flush=True)


def make_workload(ia, args):
    return ia.Workload(seq_type=ia.DNA, ntaxa=args.taxa, nsite=args.sites, ncat=4, mean_len=0.1,
                       alpha=0.5, p_inv=0.2, gap_frac=0.02, seed=20261202 + 1)


def run_steps(ctx, tree, steps, profile=False):
    evals = 0
    lnls = []
    if profile:
        ctx.profile_start()
    ctx.synchronize()
    t0 = time.perf_counter()
    for _ in range(steps):
        tree.reset()
        lnl0 = tree.compute_likelihood()
        lnl, nd = tree.optimize_all_branches(iterations=1, tolerance=1e-3, max_nr=100)
        evals += 1 + 1 + nd + 1      # traversal lnL, sweep's initial lnL, derivatives, from-buffer
        lnls.append((lnl0, lnl))
    ctx.synchronize()
    elapsed = time.perf_counter() - t0
    prof = ctx.profile_stop() if profile else None
    return elapsed, evals, lnls, prof


def time_traversal(ctx, tree, reps):
    ctx.synchronize()
    t0 = time.perf_counter()
    for _ in range(reps):
        tree.reset()
        lnl = tree.compute_likelihood()
    ctx.synchronize()
    return (time.perf_counter() - t0) / reps, lnl


def cpu_baseline(ia, w, budget_s, threads):
    """Oracle (FP64 CPU restatement, AVX2/FMA, OpenMP) on the same step - the
    full traversal, then Newton on every branch in the same pre-order - from
    the same initial lengths, repeated until the time budget is spent (at
    least once; a step that outlasts the budget stops after the branch that
    crosses it)."""
    from oracle.binding import OracleTree
    planner = ia.PhyloTree(None, w)
    orc = OracleTree.from_workload(w, nthreads=threads)
    branches = planner.branches()
    evals = 0
    steps = 0
    nbr = 0
    t0 = time.perf_counter()
    while True:
        planner.reset()
        a0, b0 = branches[0]
        ops, node, dad, length = planner.plan(b0, a0)
        orc.likelihood_branch(ops, node, dad, length)
        evals += 1
        nbr = 0
        for a, b in branches:
            ops, node, dad, length = planner.plan(b, a)
            x, _, nd, st = orc.optimize_branch(ops, node, dad, length)
            evals += nd + 1
            planner.set_length(a, b, x)
            nbr += 1
            if steps == 0 and time.perf_counter() - t0 > budget_s:
                break
        steps += 1
        if time.perf_counter() - t0 > budget_s or nbr < len(branches):
            break
    elapsed = time.perf_counter() - t0
    if nbr == len(branches):
        what = f"{steps} full steps (traversal + Newton on all {len(branches)} branches)"
    else:
        what = f"full traversal + Newton on the first {nbr} of {len(branches)} branches"
    return dict(value=w.npatterns * evals / elapsed, unit="site-lik evals/s", cores=threads,
                kind="port",
                sample=f"{what}: {evals} tree evaluations x {w.npatterns} patterns, {elapsed:.1f} s")


def make_partitions(ia, nparts, seed=20261202 + 3):
    """Config 4: one 64-taxon tree (Exp mean 0.1), partition lengths
    lognormal(ln 800, 0.8) clipped to [100, 5000], 70% DNA GTR+G4 / 30%
    protein +G4, independently seeded parameters (alpha ~ U(0.2, 2))."""
    import numpy as np
    rng = np.random.default_rng(seed)
    base = ia.Workload(seq_type=ia.DNA, ntaxa=64, nsite=10, mean_len=0.1, seed=seed)
    edges = base.edges()
    parts = []
    for k in range(nparts):
        nsite = int(np.clip(round(rng.lognormal(np.log(800.0), 0.8)), 100, 5000))
        st = ia.PROTEIN if rng.random() < 0.3 else ia.DNA
        parts.append(ia.Workload(seq_type=st, ntaxa=64, nsite=nsite, ncat=4, alpha=-1.0,
                                 seed=seed * 1000 + k, edges=edges))
    return parts


def bench_partitions(ia, ctx, nparts, reps, cpu_seconds, threads):
    import numpy as np
    t0 = time.perf_counter()
    parts = make_partitions(ia, nparts)
    plans = []
    for w in parts:
        planner = ia.PhyloTree(None, w)
        a, b = planner.branches()[0]
        plans.append(planner.plan(a, b))
    total_p = sum(w.npatterns for w in parts)
    total_sites = sum(w.nsite for w in parts)
    log(f"partitions: {nparts}, {total_sites} sites, {total_p} patterns "
        f"({time.perf_counter() - t0:.1f} s to generate)")
    res = {"partitions": nparts, "sites": total_sites, "patterns": total_p}
    lnl = {}
    for prec, name in ((ia.PREC_F64, "fp64"), (ia.PREC_F32, "fp32")):
        insts = [ia.Instance.from_workload(ctx, w, precision=prec) for w in parts]
        jobs = [(inst, *plan) for inst, plan in zip(insts, plans)]
        if prec == ia.PREC_F64:
            # plain form: one computeLikelihoodBranch launch per partition
            for inst, ops, node, dad, length in jobs:
                inst.likelihood_branch(ops, node, dad, length)
            ctx.synchronize()
            t = time.perf_counter()
            for _ in range(reps):
                single = [inst.likelihood_branch(ops, node, dad, length)
                          for inst, ops, node, dad, length in jobs]
            ctx.synchronize()
            dt = (time.perf_counter() - t) / reps
            res["plain_fp64_per_partition_evals_per_s"] = total_p / dt
            res["plain_fp64_ms"] = 1e3 * dt
            lnl["plain"] = np.array(single)
        plan = ia.BatchPlan(ctx, jobs)              # marshalled once, as a C++ caller holds its items
        for _ in range(10):                         # the recurring shape's first calls each time
            plan.run()                              # one launch-shape candidate (8 at most)
        ctx.synchronize()
        t = time.perf_counter()
        for _ in range(reps):
            out = plan.run()
        ctx.synchronize()
        dt = (time.perf_counter() - t) / reps
        res[f"batched_{name}_evals_per_s"] = total_p / dt
        res[f"batched_{name}_ms"] = 1e3 * dt
        lnl[name] = out
        for inst in insts:
            inst.close()
    res["batched_fp64_vs_plain"] = res["batched_fp64_evals_per_s"] / res["plain_fp64_per_partition_evals_per_s"]
    res["batched_fp32_vs_plain_fp64"] = res["batched_fp32_evals_per_s"] / res["plain_fp64_per_partition_evals_per_s"]
    res["max_abs_dlnl_fp32_per_tree"] = float(np.max(np.abs(lnl["fp32"] - lnl["fp64"])))
    res["max_rel_dlnl_fp32_per_tree"] = float(np.max(np.abs(lnl["fp32"] - lnl["fp64"]) / np.abs(lnl["fp64"])))
    res["max_abs_dlnl_batched_vs_plain_fp64"] = float(np.max(np.abs(lnl["fp64"] - lnl["plain"])))
    if cpu_seconds > 0:
        res["cpu_baseline"] = cpu_partitions(ia, parts, plans, cpu_seconds, threads)
        res["batched_fp32_vs_cpu"] = res["batched_fp32_evals_per_s"] / res["cpu_baseline"]["value"]
    return res


def cpu_partitions(ia, parts, plans, budget_s, threads):
    """Oracle over partitions the way PhyloSuperTree runs them: partitions in
    parallel, one thread each, largest first; the whole set is evaluated
    repeatedly until the time budget is spent (oracle trees built beforehand,
    untimed, as the partitions' PhyloTrees exist before an evaluation)."""
    from concurrent.futures import ThreadPoolExecutor
    from oracle.binding import OracleTree
    order = sorted(range(len(parts)), key=lambda k: -parts[k].npatterns * parts[k].nstates ** 2)
    orcs = {k: OracleTree.from_workload(parts[k], nthreads=1) for k in order}

    def run(k):
        ops, node, dad, length = plans[k]
        orcs[k].likelihood_branch(ops, node, dad, length)
        return parts[k].npatterns

    total, rounds = 0, 0
    t0 = time.perf_counter()
    with ThreadPoolExecutor(max_workers=threads) as ex:
        while True:
            total += sum(ex.map(run, order))
            rounds += 1
            if time.perf_counter() - t0 > budget_s:
                break
    elapsed = time.perf_counter() - t0
    return dict(value=total / elapsed, unit="site-lik evals/s", cores=threads, kind="port",
                sample=f"{rounds} evaluations of all {len(parts)} partitions (full traversal + "
                       f"lnL each), {elapsed:.1f} s")


def bench_mfma(ia, ctx, reps=3):
    """Config 3 (protein LG-like+G4, 100 taxa x 20,000 sites, Exp(0.2)
    lengths, alpha 0.8) and a codon GY94+G4 case (64 taxa x 5,000 codons):
    full post-order traversals with the tree lnL on the matrix-core kernels,
    FP64 and FP32 storage, with |dlnL| of FP32 against FP64."""
    cases = [("protein_100x20k", ia.Workload(seq_type=ia.PROTEIN, ntaxa=100, nsite=20000, ncat=4,
                                               mean_len=0.2, alpha=0.8, seed=20261202 + 2)),
             ("codon_64x5k", ia.Workload(seq_type=ia.CODON, ntaxa=64, nsite=5000, ncat=4,
                                           mean_len=0.1, alpha=0.5, seed=77))]
    out = {}
    for name, w in cases:
        res = {"patterns": w.npatterns}
        lnl = {}
        for prec, pn in ((ia.PREC_F64, "fp64"), (ia.PREC_F32, "fp32")):
            tree = ia.PhyloTree(ctx, w, precision=prec)
            tree.compute_likelihood()
            ctx.profile_start()
            t = time.perf_counter()
            for _ in range(reps):
                tree.clear_all_partial()
                lnl[pn] = tree.compute_likelihood()
            ctx.synchronize()
            dt = (time.perf_counter() - t) / reps
            k = ctx.profile_stop()["traverse_mfma"]
            res[pn] = {"traversal_ms": 1e3 * dt,
                       "site_lik_evals_per_s": w.npatterns / dt,
                       "kernel_ms": k["total_ms"] / k["launches"],
                       "tflops": k["flops"] / k["total_ms"] * 1e-9,
                       "gbps": k["bytes"] / k["total_ms"] * 1e-6}
            tree.close()
        res["rel_dlnl_fp32"] = abs(lnl["fp32"] - lnl["fp64"]) / abs(lnl["fp64"])
        out[name] = res
    return out


def make_requests(ia, nreq, seed=20261202 + 4):
    """Config 5: per request N ~ log-uniform[8, 256] taxa and L ~
    log-uniform[500, 50,000] sites on a seeded Exp(0.1) tree; DNA GTR+G4
    (60%), protein +G4 (30%) or codon GY94 +G4 (10%, omega ~ Gamma(2, 0.25),
    kappa ~ U(1, 4), equal codon frequencies); op full lnL, single-branch
    Newton or one multinomial(L, freq/L) bootstrap reweighting, in equal
    shares."""
    import numpy as np
    from iqtree2_amd.serving import Request
    rng = np.random.default_rng(seed)
    reqs = []
    for k in range(nreq):
        ntaxa = int(round(np.exp(rng.uniform(np.log(8), np.log(256)))))
        nsite = int(round(np.exp(rng.uniform(np.log(500), np.log(50000)))))
        u = rng.random()
        st = ia.DNA if u < 0.6 else (ia.PROTEIN if u < 0.9 else ia.CODON)
        w = ia.Workload(seq_type=st, ntaxa=ntaxa, nsite=nsite, ncat=4, mean_len=0.1,
                        alpha=-1.0, seed=seed * 1000 + k)
        kind = ("lnl", "newton", "rell")[int(rng.integers(3))]
        weights = None
        if kind == "rell":
            f = w.get("freq")
            weights = rng.multinomial(int(f.sum()), f / f.sum())[None, :]
        branch = int(rng.integers(2 * ntaxa - 3))
        reqs.append(Request(w, kind, branch=branch, weights=weights))
    return reqs


def bench_serving(ia, ctx, nreq, seed=20261202 + 4):
    """Capacity with every request queued at once, then a Poisson stream at
    70% of it; latency = completion - submission on the server clock."""
    import threading
    import numpy as np
    from iqtree2_amd.serving import Server
    t0 = time.perf_counter()
    reqs = make_requests(ia, nreq, seed)
    kinds = {0: "lnl", 1: "newton", 2: "rell"}
    log(f"serving: {nreq} requests, {sum(r.npatterns for r in reqs)} patterns, largest "
        f"footprint {max(r.footprint() for r in reqs) / 2**30:.2f} GiB "
        f"({time.perf_counter() - t0:.1f} s to generate)")

    def evals(r, res):
        # tree-level evaluations: traversal lnL, or theta + Newton steps + final lnL
        if r.kind == 1:
            return r.npatterns * (res["n_derv"] + 2)
        return r.npatterns

    # warm-up: a few requests through a server (kernel loads, autotuning)
    with Server(ctx) as srv:
        for r in reqs[:8]:
            srv.wait(srv.submit(r))
    # capacity: all queued at once
    with Server(ctx) as srv:
        c0 = srv.clock()
        tickets = [srv.submit(r) for r in reqs]
        closed = [srv.wait(t) for t in tickets]
        cap_s = srv.clock() - c0
        cap_stats = srv.stats()
    capacity = nreq / cap_s
    lam = 0.7 * capacity
    gaps = np.random.default_rng(seed + 1).exponential(1.0 / lam, nreq)
    arrivals = np.cumsum(gaps)
    results = [None] * nreq
    with Server(ctx) as srv:
        tk = [None] * nreq
        start = time.perf_counter()

        def submitter():
            for k, at in enumerate(arrivals):
                delay = at - (time.perf_counter() - start)
                if delay > 0:
                    time.sleep(delay)
                tk[k] = srv.submit(reqs[k])

        th = threading.Thread(target=submitter)
        s0 = srv.clock()
        th.start()
        for k in range(nreq):
            while tk[k] is None:
                time.sleep(0.0005)
            results[k] = srv.wait(tk[k])
        th.join()
        span = srv.clock() - s0
        stats = srv.stats()
    lat = np.array([r["finish_s"] - r["submit_s"] for r in results]) * 1e3
    ok = all(r["status"] == 0 for r in results)
    same = all(a["lnl"] == b["lnl"] and a["length"] == b["length"] for a, b in zip(results, closed))
    total_evals = sum(evals(r, res) for r, res in zip(reqs, results))
    mix = {}
    for r in reqs:
        key = f"{ {4: 'dna', 20: 'protein', 61: 'codon'}[r.nstates]}_{kinds[r.kind]}"
        mix[key] = mix.get(key, 0) + 1
    by_model = {}
    for r, l in zip(reqs, lat):
        m = {4: "dna", 20: "protein", 61: "codon"}[r.nstates]
        by_model.setdefault(m, []).append(l)
    return {
        "requests": nreq,
        "arrival_rate_per_s": lam,
        "capacity_req_per_s": capacity,
        "capacity_site_lik_evals_per_s": sum(evals(r, res) for r, res in zip(reqs, closed)) / cap_s,
        "p50_ms": float(np.percentile(lat, 50)),
        "p99_ms": float(np.percentile(lat, 99)),
        "mean_ms": float(lat.mean()),
        "max_ms": float(lat.max()),
        "p50_ms_by_model": {m: float(np.percentile(v, 50)) for m, v in by_model.items()},
        "p99_ms_by_model": {m: float(np.percentile(v, 99)) for m, v in by_model.items()},
        "throughput_req_per_s": nreq / span,
        "site_lik_evals_per_s": total_evals / span,
        "all_ok": ok,
        "results_identical_to_closed_loop_run": same,
        "mix": mix,
        "max_taxa": max(r.ntaxa for r in reqs),
        "max_patterns": max(r.npatterns for r in reqs),
        "scheduler": {"iterations": stats["iterations"],
                      "traversal_batches": stats["traversal_batches"],
                      "derv_batches": stats["derv_batches"],
                      "peak_inflight": stats["peak_inflight"],
                      "peak_gib": stats["peak_bytes"] / 2**30,
                      "closed_loop_peak_inflight": cap_stats["peak_inflight"],
                      "closed_loop_s": {k: round(cap_stats[k], 4) for k in
                                        ("admit_s", "traverse_s", "derv_s", "retire_s")},
                      "closed_loop_span_s": cap_s,
                      "stream_s": {k: round(stats[k], 4) for k in
                                   ("admit_s", "traverse_s", "derv_s", "retire_s")}},
    }


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=3)
    ap.add_argument("--warmup", type=int, default=1)
    ap.add_argument("--taxa", type=int, default=200)
    ap.add_argument("--sites", type=int, default=100000)
    ap.add_argument("--cpu-seconds", type=float, default=15.0)
    ap.add_argument("--no-cpu", action="store_true")
    ap.add_argument("--no-fp32", action="store_true")
    ap.add_argument("--partitions", type=int, default=500)
    ap.add_argument("--no-partitions", action="store_true")
    ap.add_argument("--requests", type=int, default=160)
    ap.add_argument("--no-serving", action="store_true")
    ap.add_argument("--no-mfma", action="store_true")
    ap.add_argument("--mfma-only", action="store_true",
                    help="only the protein/codon traversal leg (for profiling its kernels)")
    args = ap.parse_args()
    if args.gpus != 1:
        print("bench.py measures one GPU: there is no multi-GPU path (the likelihood path "
              "would shard by partition or request, one instance set per GPU)", file=sys.stderr)
        sys.exit(2)

    import iqtree2_amd as ia
    if args.mfma_only:
        ctx = ia.Context(0)
        print(json.dumps({"mfma_traversals": bench_mfma(ia, ctx)}), flush=True)
        ctx.close()
        return
    t0 = time.perf_counter()
    w = make_workload(ia, args)
    log(f"workload: {w.ntaxa} taxa, {w.nsite} sites, {w.npatterns} patterns "
        f"({time.perf_counter() - t0:.1f} s to generate)")
    ctx = ia.Context(0)
    tree = ia.PhyloTree(ctx, w, precision=ia.PREC_F64)
    for i in range(args.warmup):
        run_steps(ctx, tree, 1)
        log(f"warmup step {i + 1} done")
    elapsed, evals, lnls, prof = run_steps(ctx, tree, args.steps, profile=True)
    P = w.npatterns
    value = P * evals / elapsed
    log(f"fp64: {args.steps} steps in {elapsed:.3f} s, {evals} tree evaluations, "
        f"lnL {lnls[-1][0]:.6f} -> {lnls[-1][1]:.6f}")
    trav_s, trav_lnl = time_traversal(ctx, tree, 5)

    # dominant kernel and its roofline; HBM traffic per launch from the
    # committed rocprofv3 PMC summary of the same kernel on this workload
    dom_name, dom = max(prof.items(), key=lambda kv: kv[1]["total_ms"])
    achieved = dom["bytes"] / (dom["total_ms"] * 1e-3)
    traffic, traffic_src = None, None
    tpath = os.path.join(ROOT, "profiles", "pmc_traffic.json")
    if os.path.exists(tpath) and (args.taxa, args.sites) == (200, 100000):
        with open(tpath) as f:
            entry = json.load(f).get(dom_name)
        if entry:
            traffic, traffic_src = entry["traffic_bytes"], entry["source"]
    kernels = {k: dict(launches=v["launches"], ms=round(v["total_ms"], 3),
                       avg_us=round(1e3 * v["total_ms"] / v["launches"], 2),
                       gbps=round(v["bytes"] / max(v["total_ms"], 1e-9) * 1e-6, 1))
               for k, v in prof.items()}
    log(f"kernels: {json.dumps(kernels)}")

    fp32 = None
    if not args.no_fp32:
        tree32 = ia.PhyloTree(ctx, w, precision=ia.PREC_F32)
        run_steps(ctx, tree32, 1)
        e32, ev32, l32, _ = run_steps(ctx, tree32, args.steps)
        t32_s, t32_lnl = time_traversal(ctx, tree32, 5)
        fp32 = dict(value=P * ev32 / e32, ms_per_step=1e3 * e32 / args.steps,
                    traversal_ms=1e3 * t32_s,
                    abs_dlnl_traversal=abs(t32_lnl - trav_lnl),
                    abs_dlnl_after_newton=abs(l32[-1][1] - lnls[-1][1]),
                    rel_dlnl_traversal=abs(t32_lnl - trav_lnl) / abs(trav_lnl))
        log(f"fp32: {json.dumps(fp32)}")
        tree32.close()

    parts = None
    if not args.no_partitions:
        parts = bench_partitions(ia, ctx, args.partitions, max(1, args.steps),
                                 0.0 if args.no_cpu else args.cpu_seconds,
                                 min(16, os.cpu_count() or 1))
        log(f"partitions: {json.dumps(parts)}")

    mfma = None
    if not args.no_mfma:
        mfma = bench_mfma(ia, ctx)
        log(f"mfma traversals: {json.dumps(mfma)}")

    serving = None
    if not args.no_serving:
        serving = bench_serving(ia, ctx, args.requests)
        log(f"serving: {json.dumps(serving)}")

    cpu = None
    if not args.no_cpu:
        threads = min(16, os.cpu_count() or 1)
        cpu = cpu_baseline(ia, w, args.cpu_seconds, threads)
        # beside it, one core on a shorter sample of the same step
        one = cpu_baseline(ia, w, max(1.0, args.cpu_seconds / 2), 1)
        cpu["one_core"] = {"value": one["value"], "unit": one["unit"], "sample": one["sample"]}
        log(f"cpu baseline: {json.dumps(cpu)}")

    out = {
        "metric": "site-lik evals/s FP64 vs batched FP32 at |dlnL|<=1e-6; p50/p99 req latency",
        "value": value,
        "unit": "site-lik evals/s",
        "n_gpus": 1,
        "steps": args.steps,
        "warmup": args.warmup,
        "ms_per_step": 1e3 * elapsed / args.steps,
        "higher_is_better": True,
        "scaling": "weak",
        "vs_baseline": None,
        "dtype": "f64",
        "data": "synthetic (seeded simulation down a random tree)",
        "config": {"workload": "DNA GTR+I+G4, 200 taxa x 100,000 sites: full post-order "
                               "traversal + Newton over all 397 branches",
                   "taxa": w.ntaxa, "sites": w.nsite, "patterns": P,
                   "tree_evals_per_step": evals / args.steps,
                   "traversal_ms": 1e3 * trav_s,
                   "traversal_site_lik_evals_per_s": P / trav_s},
        "roofline": {"bound": "hbm", "kernel": dom_name, "achieved": achieved / 1e9,
                     "peak": HBM_PEAK / 1e9, "unit": "GB/s", "frac": achieved / HBM_PEAK,
                     "traffic": traffic, "traffic_source": traffic_src,
                     "avg_launch_us": 1e3 * dom["total_ms"] / dom["launches"],
                     "bytes_per_launch": dom["bytes"] / dom["launches"]},
        "kernels": kernels,
        "fp32_storage": fp32,
        "partition_batch": parts,
        "mfma_traversals": mfma,
        "serving": serving,
        "cpu_baseline": cpu,
    }
    tree.close()
    ctx.close()
    print(json.dumps(out), flush=True)


if __name__ == "__main__":
    main()
