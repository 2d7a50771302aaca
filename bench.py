#!/usr/bin/env python3
"""k-way partitioning throughput on one MI355X.

A step is one full multilevel k-way partition (coarsening, initial
partition, uncoarsening with greedy refinement) of the synthetic graph, whose
CSR is already resident in HBM when the timed region starts.  Default
workload: random geometric graph, 2^20 points in the unit square, r =
sqrt(12/(pi n)), Morton-ordered ids, vertex weight 1, edge weights U{1..10},
k = 64, SHEM + greedy, ufactor 30, niter 10, ncuts 1, seed 1.

Prints one JSON line: graphs/s (value), per-pass edge rates, the roofline of
the gain-table kernel on the finest level (HIP events on the library's
stream), and a CPU baseline (the oracle restatement of the reference on the
same graph, one host core).
"""
import argparse
import ctypes as C
import json
import math
import os
import sys
import time

ROOT = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, ROOT)

HBM_PEAK_GBS = 8000.0


def log(msg):
    print(f"[bench] {msg}", file=sys.stderr, flush=True)


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=5)
    ap.add_argument("--warmup", type=int, default=2)
    ap.add_argument("--workload", default="rgg20",
                    help="rgg20 | grid256 | rmat22 | rggNN | rmatNN | batch1024 (uneven batch of that many graphs)")
    ap.add_argument("--serve-seconds", type=float, default=20.0, help="serve: length of the arrival stream")
    ap.add_argument("--serve-rate", type=float, default=200.0, help="serve: Poisson arrival rate (requests/s)")
    ap.add_argument("--serve-pool", type=int, default=48, help="serve: distinct graphs the requests draw from")
    ap.add_argument("--slots", type=int, default=3, help="serve: batched calls in flight at once")
    ap.add_argument("--serve-max-batch", type=int, default=0,
                    help="serve: vertices per batched call at most (0: the server's default)")
    ap.add_argument("--plain-sample", type=int, default=40, help="serve: requests timed on the plain path")
    ap.add_argument("--plain-stride", type=int, default=3,
                    help="batch workloads: the plain path is timed on every n-th graph (odd: grids and RGGs alike)")
    ap.add_argument("--seed", type=int, default=1)
    ap.add_argument("--cpu-budget-s", type=float, default=20.0)
    ap.add_argument("--no-cpu", action="store_true")
    ap.add_argument("--no-host-api", action="store_true",
                    help="skip timing the host-buffer C entry point (PCIe-inclusive rate)")
    ap.add_argument("--copy-bytes", type=int, default=0,
                    help="issue one device-to-device copy of this many bytes first (a kernel of known "
                         "traffic, to scale FETCH_SIZE/WRITE_SIZE counters)")
    args = ap.parse_args()
    if args.gpus != 1:
        print(json.dumps({"error": "this build runs on one GPU only; --gpus must be 1",
                          "n_gpus": args.gpus}), flush=True)
        sys.exit(2)

    import numpy as np
    import rmetis_amd as R
    from rmetis_amd import _lib, synth

    L = _lib.lib()
    if R.device_count() < 1:
        raise SystemExit("no HIP device")
    if args.workload.startswith("batch"):
        return bench_batch(args)
    if args.workload == "serve":
        return bench_serve(args)
    t0 = time.time()
    g, k = synth.config_graph(args.workload, seed=args.seed)
    log(f"graph {g} k={k} generated in {time.time() - t0:.1f}s")
    DA = _lib.DeviceArray
    d_xadj, d_adj = DA(g.xadj), DA(g.adjncy)
    d_vw = DA(g.vwgt) if g.vwgt is not None else None
    d_aw = DA(g.adjwgt) if g.adjwgt is not None else None
    d_part = DA(count=g.n, dtype=np.int32)
    sync = L.rmetis_hip_device_synchronize
    sync()
    if args.copy_bytes > 0:
        src = L.rmetis_hip_device_alloc(args.copy_bytes)
        dst = L.rmetis_hip_device_alloc(args.copy_bytes)
        if not src or not dst or L.rmetis_hip_memcpy(dst, src, args.copy_bytes, 3) != 1:
            raise RuntimeError(f"calibration copy failed: {_lib.last_error()}")
        sync()
        L.rmetis_hip_device_free(src)
        L.rmetis_hip_device_free(dst)
        log(f"calibration copy of {args.copy_bytes} bytes done")
    opts = R.Options(rtype=R.RType.Greedy, seed=args.seed, ncuts=1, niter=10, ufactor=30).to_c_array()
    stats = _lib.HipStats()
    obj = C.c_int32()

    def step(with_stats=False):
        rc = L.rmetis_hip_kway_device(g.n, g.ncon, d_xadj.ptr, d_adj.ptr,
                                      d_vw.ptr if d_vw is not None else None,
                                      d_aw.ptr if d_aw is not None else None, k, None, None,
                                      opts.ctypes.data_as(_lib.i32p), C.byref(obj), d_part.ptr,
                                      C.byref(stats) if with_stats else None)
        if rc != 1:
            raise RuntimeError(f"kway failed rc={rc}: {_lib.last_error()}")

    for i in range(args.warmup):
        step()
        log(f"warmup {i} cut={obj.value}")
    def prof(tag, size=-1):
        ln, ms, by = C.c_int64(), C.c_double(), C.c_double()
        L.rmetis_hip_profile_read(tag, size, C.byref(ln), C.byref(ms), C.byref(by))
        return ln.value, ms.value, by.value

    # Timed region: only the full gain launches of the finest level are
    # bracketed by events (the roofline's kernel); every other launch runs
    # as it does for a caller that profiles nothing.
    L.rmetis_hip_profile_filter(0, g.n)
    L.rmetis_hip_profile_enable(1)
    sync()
    t_start = time.perf_counter()
    for i in range(args.steps):
        step()
    sync()
    elapsed = time.perf_counter() - t_start
    log(f"timed {args.steps} steps in {elapsed:.3f}s cut={obj.value}")
    g_launch, g_ms, g_bytes = prof(0, g.n)

    # One more step, untimed, with every kernel family bracketed and the
    # per-level statistics (which synchronise once per level) taken.
    L.rmetis_hip_profile_filter(-1, -1)
    L.rmetis_hip_profile_enable(1)
    step(with_stats=True)
    sync()
    prof_tags = {"gain": 0, "gain_inc": 6, "compact": 1, "admit": 2, "match": 3, "contract": 4}
    kernel_ms = {name: round(prof(t)[1], 3) for name, t in prof_tags.items()}
    L.rmetis_hip_profile_enable(0)
    avg_ms = g_ms / max(g_launch, 1)
    bytes_per_launch = g_bytes / max(g_launch, 1)
    achieved = bytes_per_launch / (avg_ms * 1e-3) / 1e9 if avg_ms > 0 else 0.0

    part = d_part.numpy()
    pw = np.bincount(part, weights=(g.vwgt if g.vwgt is not None else None), minlength=k)
    total = float(g.vwgt.sum()) if g.vwgt is not None else float(g.n)
    maxw = math.floor(total / k * 1.03)
    violations = int((pw > maxw).sum())

    levels = []
    for i in range(stats.nlevels):
        n_i, m_i = stats.level_n[i], stats.level_m[i]
        cm, rm, ps = stats.coarsen_ms[i], stats.refine_ms[i], stats.passes[i]
        levels.append({"n": n_i, "m": m_i, "coarsen_ms": round(cm, 3), "refine_ms": round(rm, 3),
                       "passes": ps,
                       "coarsen_edges_per_s": round(m_i / (cm * 1e-3), 1) if cm > 0 else None,
                       "refine_edges_per_s_per_pass": round(m_i * ps / (rm * 1e-3), 1) if rm > 0 else None})

    # the same partition through the host-buffer drop-in entry point (copies
    # the CSR in and the partition out over PCIe on every call)
    host_api = None
    if not args.no_host_api:
        nv, nc, npart = C.c_int32(g.n), C.c_int32(g.ncon), C.c_int32(k)
        i32 = _lib.i32p
        hx, ha = g.xadj.ctypes.data_as(i32), g.adjncy.ctypes.data_as(i32)
        hvw = g.vwgt.ctypes.data_as(i32) if g.vwgt is not None else None
        haw = g.adjwgt.ctypes.data_as(i32) if g.adjwgt is not None else None
        hpart = np.zeros(g.n, np.int32)
        hobj = C.c_int32()
        reps = max(1, min(args.steps, 3))
        t1 = time.perf_counter()
        for _ in range(reps):
            rc = L.RMETIS_PartGraphKway(C.byref(nv), C.byref(nc), hx, ha, hvw, None, haw, C.byref(npart), None,
                                        None, opts.ctypes.data_as(i32), C.byref(hobj), hpart.ctypes.data_as(i32))
            if rc != 1:
                raise RuntimeError(f"RMETIS_PartGraphKway failed rc={rc}: {_lib.last_error()}")
        t_host = time.perf_counter() - t1
        host_api = {"value": round(reps / t_host, 4), "unit": "graphs/s", "calls": reps,
                    "note": "RMETIS_PartGraphKway on host buffers: CSR upload, validation and the "
                            "partition download included"}
        log(f"host API: {reps} calls in {t_host:.3f}s cut={hobj.value}")

    cpu = None
    if not args.no_cpu:
        sys.path.insert(0, os.path.join(ROOT, "oracle"))
        import pyoracle as O
        O.build()
        runs, t_cpu = 0, 0.0
        while t_cpu < args.cpu_budget_s and runs < 10:
            t1 = time.perf_counter()
            rc, cobj, _ = O.part_graph_kway(g.xadj, g.adjncy, k, vwgt=g.vwgt, adjwgt=g.adjwgt, ncon=g.ncon,
                                            seed=args.seed + runs, niter=10, ufactor=30)
            t_cpu += time.perf_counter() - t1
            runs += 1
            log(f"cpu oracle run {runs}: cut={cobj} ({t_cpu:.1f}s total)")
        cpu = {"value": round(runs / t_cpu, 4), "unit": "graphs/s", "cores": 1, "kind": "port",
               "sample": f"{runs} full k-way partitions of the same graph (seeds {args.seed}..{args.seed + runs - 1}) "
                         f"by the oracle's C++ restatement of the reference, one host core, {t_cpu:.1f}s",
               "cut": int(cobj)}

    # HBM bytes per launch of the same kernel from the committed rocprofv3
    # FETCH_SIZE / WRITE_SIZE passes (scaled by a copy kernel of known size)
    traffic, traffic_src = None, None
    tpath = os.path.join(ROOT, "profiles", f"k_gain_tile_{args.workload}.json")
    if os.path.exists(tpath):
        with open(tpath) as fh:
            tj = json.load(fh)
        if tj.get("level_n") == g.n:
            traffic, traffic_src = tj["traffic_bytes"], tj["summary"]

    value = args.steps / elapsed
    out = {
        "metric": "graphs/s (k-way partitions of one graph per second)",
        "value": round(value, 4),
        "unit": "graphs/s",
        "n_gpus": 1,
        "steps": args.steps,
        "warmup": args.warmup,
        "ms_per_step": round(1e3 * elapsed / args.steps, 3),
        "higher_is_better": True,
        "scaling": "weak",
        "vs_baseline": None,
        "dtype": "int32",
        "data": "synthetic",
        "config": {"workload": f"{args.workload}: {g.name} n={g.n} m={g.m} k={k}",
                   "ctype": "shem", "rtype": "greedy", "ufactor": 30, "niter": 10, "ncuts": 1,
                   "seed": args.seed, "parallelism": "single GPU"},
        "edge_cut": int(obj.value),
        "balance_violations": violations,
        "edges_per_s": round(g.m * value, 1),
        "kernel_ms_per_step": kernel_ms,
        "levels": levels,
        "roofline": {"bound": "hbm", "kernel": "k_gain (finest level)", "achieved": round(achieved, 1),
                     "peak": HBM_PEAK_GBS, "unit": "GB/s", "frac": round(achieved / HBM_PEAK_GBS, 4),
                     "traffic": traffic, "traffic_source": traffic_src, "bytes_per_launch": int(bytes_per_launch),
                     "avg_launch_us": round(avg_ms * 1e3, 2), "launches": g_launch},
        "cpu_baseline": cpu,
        "host_api": host_api,
    }
    print(json.dumps(out), flush=True)


def _prof(L, tag, size=-1):
    ln, ms, by = C.c_int64(), C.c_double(), C.c_double()
    L.rmetis_hip_profile_read(tag, size, C.byref(ln), C.byref(ms), C.byref(by))
    return ln.value, ms.value, by.value


def bench_batch(args):
    """The uneven batch: `count` graphs (n log-uniform in 2^10..2^18, grids
    and RGGs, vwgt U{1..4}, adjwgt U{1..10}, k from {2..64}), 8 attempts per
    graph (seeds s..s+7), all packed into one segmented CSR resident in HBM
    and partitioned by one rmetis_hip_kway_batch_device call per step.  The
    plain path (one rmetis_hip_kway_device call per graph, attempts one after
    another) is timed on every plain_stride-th graph of the same batch."""
    import numpy as np
    import rmetis_amd as R
    from rmetis_amd import _lib, synth

    L = _lib.lib()
    DA = _lib.DeviceArray
    i32 = _lib.i32p
    sync = L.rmetis_hip_device_synchronize
    count = int(args.workload[5:] or 1024)
    ncuts = 8
    t0 = time.time()
    workers = max(1, min(16, os.cpu_count() or 1))
    batch = synth.uneven_batch(count, args.seed, workers=workers)
    graphs = [g for g, _ in batch]
    ks = np.array([k for _, k in batch], np.int32)
    gvtx, xadj, adjncy, vwgt, adjwgt, ncon = R.pack_batch(graphs)
    n, m = int(gvtx[-1]), int(xadj[-1])
    log(f"batch of {count} graphs: n={n} m={m} generated and packed in {time.time() - t0:.1f}s ({workers} workers)")
    d_x, d_a, d_vw, d_aw = DA(xadj), DA(adjncy), DA(vwgt), DA(adjwgt)
    d_part = DA(count=n, dtype=np.int32)
    sync()
    opts = R.Options(rtype=R.RType.Greedy, seed=args.seed, ncuts=ncuts, niter=10, ufactor=30).to_c_array()
    obj = np.zeros(count, np.int32)
    stats = _lib.HipStats()

    def step(with_stats=False):
        rc = L.rmetis_hip_kway_batch_device(count, gvtx.ctypes.data_as(i32), ncon, d_x.ptr, d_a.ptr, d_vw.ptr,
                                            d_aw.ptr, ks.ctypes.data_as(i32), None, opts.ctypes.data_as(i32), None,
                                            obj.ctypes.data_as(i32), d_part.ptr,
                                            C.byref(stats) if with_stats else None)
        if rc != 1:
            raise RuntimeError(f"batch kway failed rc={rc}: {_lib.last_error()}")

    for i in range(args.warmup):
        t1 = time.perf_counter()
        step()
        log(f"warmup {i}: {time.perf_counter() - t1:.2f}s")
    # timed region: events only around the finest level's full gain launches
    L.rmetis_hip_profile_filter(0, n)
    L.rmetis_hip_profile_enable(1)
    sync()
    t_start = time.perf_counter()
    for i in range(args.steps):
        step()
        log(f"step {i} done")
    sync()
    elapsed = time.perf_counter() - t_start
    value = count * args.steps / elapsed
    log(f"batched: {args.steps} steps in {elapsed:.2f}s = {value:.2f} graphs/s")
    g_launch, g_ms, g_bytes = _prof(L, 0, n)
    # one untimed step with every kernel family timed and the level statistics
    L.rmetis_hip_profile_filter(-1, -1)
    L.rmetis_hip_profile_enable(1)
    step(with_stats=True)
    sync()
    prof_tags = {"gain": 0, "gain_inc": 6, "compact": 1, "admit": 2, "match": 3, "contract": 4}
    kernel_ms = {name: round(_prof(L, t)[1], 2) for name, t in prof_tags.items()}
    L.rmetis_hip_profile_enable(0)
    avg_ms = g_ms / max(g_launch, 1)
    bytes_per_launch = g_bytes / max(g_launch, 1)
    achieved = bytes_per_launch / (avg_ms * 1e-3) / 1e9 if avg_ms > 0 else 0.0
    part = d_part.numpy()
    bobj = obj.copy()
    # balance: every part of every graph within floor(total/k * 1.03)
    violations, violators = 0, []
    for i, (g, k) in enumerate(batch):
        p = part[gvtx[i]:gvtx[i + 1]]
        pw = np.bincount(p, weights=g.vwgt, minlength=k)
        lim = math.floor((1.0 / k) * float(g.vwgt.sum()) * (1.0 + 30 / 1000.0))
        bad = int((pw > lim).sum())
        if bad:
            violations += bad
            violators.append({"graph": i, "n": g.n, "k": int(k), "parts_over": bad, "max_pw": int(pw.max()),
                              "limit": lim})
    levels = [{"n": stats.level_n[i], "m": stats.level_m[i], "coarsen_ms": round(stats.coarsen_ms[i], 3),
               "refine_ms": round(stats.refine_ms[i], 3)} for i in range(stats.nlevels)]
    # the plain path on a sample of the same graphs, same options
    sample = list(range(0, count, max(1, args.plain_stride)))
    dev = []
    for i in sample:
        g = graphs[i]
        dev.append((DA(g.xadj), DA(g.adjncy), DA(g.vwgt), DA(g.adjwgt), DA(count=g.n, dtype=np.int32)))
    sync()
    pobj = C.c_int32()

    def plain(j):
        i = sample[j]
        g = graphs[i]
        dx, da, dvw, daw, dp = dev[j]
        rc = L.rmetis_hip_kway_device(g.n, 1, dx.ptr, da.ptr, dvw.ptr, daw.ptr, int(ks[i]), None, None,
                                      opts.ctypes.data_as(i32), C.byref(pobj), dp.ptr, None)
        if rc != 1:
            raise RuntimeError(f"plain kway failed rc={rc}: {_lib.last_error()}")
        return pobj.value

    plain(0)
    sync()
    mism = 0
    t1 = time.perf_counter()
    for j in range(len(sample)):
        if plain(j) != bobj[sample[j]]:
            mism += 1
    sync()
    t_plain = time.perf_counter() - t1
    plain_value = len(sample) / t_plain
    log(f"plain: {len(sample)} graphs in {t_plain:.2f}s = {plain_value:.2f} graphs/s, objval mismatches {mism}")
    cpu = None
    if not args.no_cpu:
        sys.path.insert(0, os.path.join(ROOT, "oracle"))
        import pyoracle as O
        O.build()
        runs, t_cpu = 0, 0.0
        for i in sample:
            if t_cpu >= args.cpu_budget_s:
                break
            g = graphs[i]
            t2 = time.perf_counter()
            O.part_graph_kway(g.xadj, g.adjncy, int(ks[i]), vwgt=g.vwgt, adjwgt=g.adjwgt, ncuts=ncuts,
                              seed=args.seed, niter=10, ufactor=30)
            t_cpu += time.perf_counter() - t2
            runs += 1
        cpu = {"value": round(runs / t_cpu, 4), "unit": "graphs/s", "cores": 1, "kind": "port",
               "sample": f"the first {runs} graphs of the plain-path sample, {ncuts} attempts each, by the "
                         f"oracle's C++ restatement of the reference on one host core, {t_cpu:.1f}s"}
        log(f"cpu oracle: {runs} graphs in {t_cpu:.1f}s")
    out = {
        "metric": "graphs/s (k-way partitions of the uneven batch, 8 attempts per graph)",
        "value": round(value, 3),
        "unit": "graphs/s",
        "n_gpus": 1,
        "steps": args.steps,
        "warmup": args.warmup,
        "ms_per_step": round(1e3 * elapsed / args.steps, 1),
        "higher_is_better": True,
        "scaling": "weak",
        "vs_baseline": None,
        "dtype": "int32",
        "data": "synthetic",
        "config": {"workload": f"{args.workload}: {count} graphs, n={n} m={m}, k in {{2..64}}",
                   "ctype": "shem", "rtype": "greedy", "ufactor": 30, "niter": 10, "ncuts": ncuts,
                   "seed": args.seed, "parallelism": "single GPU, one segmented CSR"},
        "plain": {"value": round(plain_value, 3), "unit": "graphs/s", "graphs": len(sample),
                  "objval_mismatches": mism,
                  "note": "rmetis_hip_kway_device per graph (device-resident CSR), same options"},
        "batched_vs_plain": round(value / plain_value, 2),
        "edge_cut_sum": int(bobj.sum()),
        "balance_violations": violations,
        "balance_violators": violators[:16],
        "edges_per_s": round(m * ncuts * value / count, 1),
        "kernel_ms_per_step": kernel_ms,
        "levels": levels,
        "roofline": {"bound": "hbm", "kernel": "k_gain (finest union level)", "achieved": round(achieved, 1),
                     "peak": HBM_PEAK_GBS, "unit": "GB/s", "frac": round(achieved / HBM_PEAK_GBS, 4),
                     "traffic": None, "bytes_per_launch": int(bytes_per_launch),
                     "avg_launch_us": round(avg_ms * 1e3, 2), "launches": g_launch},
        "cpu_baseline": cpu,
    }
    print(json.dumps(out), flush=True)


def bench_serve(args):
    """Open-loop serving: Poisson arrivals at --serve-rate for --serve-seconds;
    each request partitions a graph drawn from a pool of --serve-pool graphs
    built as the uneven batch's (n log-uniform in 2^12..2^22, ncon 2, vwgt
    U{1..8} per constraint, adjwgt U{1..10}; uploaded to HBM once) into k parts,
    k uniform over {8, 16, ..., 512}, ubvec (1.05, 1.05), with its own seed.
    Requests go to a Server by device pointer; latency is submit to result.
    The plain path (rmetis_hip_kway_device, one request at a time) is timed on
    a sample of the same requests."""
    import threading
    import queue
    import numpy as np
    import rmetis_amd as R
    from rmetis_amd import _lib, synth

    L = _lib.lib()
    DA = _lib.DeviceArray
    i32 = _lib.i32p
    t0 = time.time()
    workers = max(1, min(16, os.cpu_count() or 1))
    pool = synth.serve_pool(args.serve_pool, args.seed, workers=workers)
    dev = [(DA(g.xadj), DA(g.adjncy), DA(g.vwgt), DA(g.adjwgt)) for g in pool]
    L.rmetis_hip_device_synchronize()
    log(f"pool of {len(pool)} graphs (n {min(g.n for g in pool)}..{max(g.n for g in pool)}, "
        f"{sum(g.n for g in pool)} vertices) built and uploaded in {time.time() - t0:.1f}s")
    ub = np.array([1.05, 1.05], np.float32)
    opts = R.Options(rtype=R.RType.Greedy, ncuts=1, niter=10, ufactor=30)
    maxn = max(g.n for g in pool)
    ring = 256
    outs = [DA(count=maxn, dtype=np.int32) for _ in range(ring)]
    srv = R.Server(slots=args.slots, max_batch_vertices=args.serve_max_batch)
    # warm-up (untimed): every pool graph once, so arenas and staging are grown
    wids = [srv.submit_device(g.n, 2, *d, 8, outs[i % ring], ubvec=ub, options=opts, seed=1)
            for i, (g, d) in enumerate(zip(pool, dev)) if i < ring]
    for w in wids:
        srv.result(w)

    def stream(rate, seconds, seed):
        """Runs one Poisson stream through the server; returns its requests,
        per-request (total, queued, running) ms, batch sizes and objvals."""
        rng = np.random.default_rng(seed)
        reqs, t = [], 0.0
        while True:
            t += rng.exponential(1.0 / rate)
            if t > seconds:
                break
            gi = int(rng.integers(len(pool)))
            k = min(int(2 ** rng.integers(3, 10)), pool[gi].n)
            reqs.append((t, gi, k, len(reqs) + 1))
        log(f"{len(reqs)} requests over {seconds:.0f}s at {rate:.0f}/s, {args.slots} slots")
        ids = queue.Queue()
        freed = [threading.Semaphore(0) for _ in range(ring)]
        start = time.perf_counter() + 0.05

        def submitter():
            for i, (ta, gi, k, sd) in enumerate(reqs):
                if i >= ring:
                    freed[i % ring].acquire()              # request i - ring has been collected
                dt = start + ta - time.perf_counter()
                if dt > 0:
                    time.sleep(dt)
                g, d = pool[gi], dev[gi]
                ids.put(srv.submit_device(g.n, 2, *d, k, outs[i % ring], ubvec=ub, options=opts, seed=sd))

        th = threading.Thread(target=submitter, daemon=True)
        th.start()
        rows, objs = [], []
        done_t = start
        for i in range(len(reqs)):
            r = srv.result(ids.get())
            rows.append((r.total_ms, r.queued_ms, r.running_ms, r.batch_graphs))
            objs.append(r.objval)
            done_t = time.perf_counter()
            freed[i % ring].release()
            if i % 1000 == 999:
                log(f"{i + 1} served")
        th.join()
        return reqs, np.array(rows), objs, done_t - start

    b0 = srv.stats()["batches"]
    reqs, rows, objs, elapsed = stream(args.serve_rate, args.serve_seconds, args.seed + 1)
    nb = srv.stats()["batches"] - b0
    value = len(reqs) / elapsed
    lat = rows[:, 0]
    p50, p99 = float(np.percentile(lat, 50)), float(np.percentile(lat, 99))
    log(f"served {len(reqs)} in {elapsed:.2f}s = {value:.1f} graphs/s, p50 {p50:.1f} ms, p99 {p99:.1f} ms, "
        f"{nb} batches")
    # capacity: a stream at five times the rate saturates the server
    creqs, crows, _, celapsed = stream(5 * args.serve_rate, max(3.0, args.serve_seconds / 4), args.seed + 3)
    capacity = len(creqs) / celapsed
    log(f"capacity: {len(creqs)} in {celapsed:.2f}s = {capacity:.1f} graphs/s, "
        f"mean batch {crows[:, 3].mean():.1f} graphs")
    st = srv.stats()
    srv.close()
    # the plain path, one request at a time, on a sample of the same requests
    samp = np.random.default_rng(args.seed + 2).choice(len(reqs), size=min(args.plain_sample, len(reqs)),
                                                        replace=False)
    obj = C.c_int32()
    plain_ms, mism = [], 0
    for j in samp:
        _, gi, k, seed = reqs[j]
        g, d = pool[gi], dev[gi]
        o = R.Options(rtype=R.RType.Greedy, ncuts=1, niter=10, ufactor=30, seed=seed).to_c_array()
        t1 = time.perf_counter()
        rc = L.rmetis_hip_kway_device(g.n, 2, d[0].ptr, d[1].ptr, d[2].ptr, d[3].ptr, k, None,
                                      ub.ctypes.data_as(_lib.f32p), o.ctypes.data_as(i32), C.byref(obj),
                                      outs[0].ptr, None)
        if rc != 1:
            raise RuntimeError(f"plain kway failed: {_lib.last_error()}")
        plain_ms.append((time.perf_counter() - t1) * 1e3)
        mism += int(obj.value != objs[j])
    plain_ms = np.array(plain_ms)
    plain_value = 1e3 / plain_ms.mean()
    log(f"plain: {len(samp)} requests, {plain_value:.1f} graphs/s sequential, objval mismatches {mism}")
    out = {
        "metric": "graphs/s (served open-loop stream, continuous batching)",
        "value": round(value, 2),
        "unit": "graphs/s",
        "n_gpus": 1,
        "steps": len(reqs),
        "warmup": len(wids),
        "ms_per_step": round(1e3 * elapsed / len(reqs), 3),
        "higher_is_better": True,
        "scaling": "weak",
        "vs_baseline": None,
        "dtype": "int32",
        "data": "synthetic",
        "config": {"workload": f"serve: Poisson {args.serve_rate:.0f}/s for {args.serve_seconds:.0f}s, "
                               f"{len(pool)}-graph pool n 2^12..2^22 ncon 2, k 8..512",
                   "rtype": "greedy", "ubvec": [1.05, 1.05], "ncuts": 1, "slots": args.slots,
                   "parallelism": "single GPU, continuous batching"},
        "latency_ms": {"p50": round(p50, 2), "p99": round(p99, 2), "mean": round(float(lat.mean()), 2),
                       "queued_p50": round(float(np.percentile(rows[:, 1], 50)), 2),
                       "running_p50": round(float(np.percentile(rows[:, 2], 50)), 2)},
        "batches": nb, "mean_batch_graphs": round(float(rows[:, 3].mean()), 2),
        "max_inflight_batches": st["max_inflight_batches"],
        "capacity": {"value": round(capacity, 2), "unit": "graphs/s", "requests": len(creqs),
                     "rate": 5 * args.serve_rate, "mean_batch_graphs": round(float(crows[:, 3].mean()), 2),
                     "p50_ms": round(float(np.percentile(crows[:, 0], 50)), 2)},
        "offered_vertices_per_s": round(sum(pool[gi].n for _, gi, _, _ in reqs) / args.serve_seconds, 1),
        "plain": {"value": round(plain_value, 2), "unit": "graphs/s", "requests": len(samp),
                  "service_ms_p50": round(float(np.percentile(plain_ms, 50)), 2),
                  "service_ms_p99": round(float(np.percentile(plain_ms, 99)), 2), "objval_mismatches": mism,
                  "note": "rmetis_hip_kway_device one request at a time, same requests"},
        "capacity_vs_plain": round(capacity / plain_value, 2),
    }
    print(json.dumps(out), flush=True)


if __name__ == "__main__":
    main()
