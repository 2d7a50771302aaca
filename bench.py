#!/usr/bin/env python3
"""k-way partitioning throughput on one MI355X.

Default workload (the headline): the uneven batch of 1,024 graphs (n
log-uniform in 2^10..2^18, half 2D grids with a random aspect ratio, half
random geometric graphs of mean degree 8-16, vwgt U{1..4}, adjwgt U{1..10},
k from {2,4,8,16,32,64}), 8 attempts (seeds) per graph, all packed into one
segmented CSR resident in HBM.  A step is one batched call that partitions
every graph (multilevel k-way: SHEM coarsening, BFS growing of the coarsest
level, greedy refinement, ufactor 30, niter 10).  The same graphs through the
plain one-graph-per-call GPU path give `plain`, and their ratio
`batched_vs_plain`.

The JSON line also carries the roofline of the step's dominant kernel (HIP
events on the library's stream over the timed steps, algorithmic bytes per
launch), per-kernel rates at the finest level, the CPU baseline (the
oracle's restatement of the reference on a bounded sample, one core and a
pool over the host cores) and, as an extra key, the single-graph workload
(`rgg20`: 2^20-point random geometric graph, k = 64).

Other workloads: --workload rgg20 | grid256 | rmat22 | rggNN | rmatNN | batchN
| serve | siblings (recursive bisection and nested dissection, also an extra
key of the default run).
"""
import argparse
import ctypes as C
import json
import math
import os
import sys
import threading
import time

ROOT = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, ROOT)

HBM_PEAK_GBS = 8000.0
METRIC = "graphs/s + level edges/s (%HBM peak), batched vs plain GPU, cut <= ref+3%"
# profiler tags of the library (rmetis_hip_profile_read)
TAGS = {"gain": 0, "gain_inc": 6, "compact": 1, "admit": 2, "match": 3, "contract": 4}
# kernel names behind each tag with an algorithmic byte model
TAG_KERNELS = {"gain": "k_gain_tile (full launches)", "match": "k_suitor_init + k_suitor + k_suitor_pairs",
               "contract": "k_leader + scans + k_cmap + k_contract_tile/big/copy"}


def log(msg):
    print(f"[bench] {msg}", file=sys.stderr, flush=True)


def cpu_model():
    try:
        with open("/proc/cpuinfo") as fh:
            for line in fh:
                if line.startswith("model name"):
                    return line.split(":", 1)[1].strip()
    except OSError:
        pass
    return "unknown"


def host_threads():
    try:
        n = len(os.sched_getaffinity(0))
    except AttributeError:
        n = os.cpu_count() or 1
    return max(1, min(16, n))


def _prof(L, tag, size=-1):
    ln, ms, by = C.c_int64(), C.c_double(), C.c_double()
    L.rmetis_hip_profile_read(tag, size, C.byref(ln), C.byref(ms), C.byref(by))
    return ln.value, ms.value, by.value


def profiled_step(L, step, size):
    """One untimed step with every kernel family bracketed: per tag, the
    total ms over the step and (launches, ms, bytes) at the level of `size`
    vertices (a callable: read after the step)."""
    L.rmetis_hip_profile_filter(-1, -1)
    L.rmetis_hip_profile_enable(1)
    step(True)
    L.rmetis_hip_device_synchronize()
    if callable(size):
        size = size()
    total = {name: _prof(L, t)[1] for name, t in TAGS.items()}
    finest = {name: _prof(L, t, size) for name, t in TAGS.items()}
    L.rmetis_hip_profile_enable(0)
    return total, finest


def level_kernels(finest):
    out = {}
    for name, (ln, ms, by) in finest.items():
        if ln == 0:
            continue
        e = {"launches": ln, "avg_us": round(1e3 * ms / ln, 2)}
        if by > 0 and ms > 0:
            gbs = by / (ms * 1e-3) / 1e9
            e.update({"bytes_per_launch": int(by / ln), "achieved_gbs": round(gbs, 1),
                      "frac": round(gbs / HBM_PEAK_GBS, 4)})
        out[name] = e
    return out


def dominant(total):
    """The tag with the most kernel time among those with a byte model."""
    modelled = [t for t in ("gain", "match", "contract") if total.get(t, 0) > 0]
    return max(modelled, key=lambda t: total[t])


def timed_steps(L, step, steps, tag, size):
    """Times `steps` steps; only the launches of `tag` (at every level when
    size < 0) are bracketed by events.  Returns (elapsed s, launches, ms,
    bytes) of the bracketed launches."""
    L.rmetis_hip_profile_filter(TAGS[tag], size)
    L.rmetis_hip_profile_enable(1)
    L.rmetis_hip_device_synchronize()
    t0 = time.perf_counter()
    for i in range(steps):
        step(False)
    L.rmetis_hip_device_synchronize()
    elapsed = time.perf_counter() - t0
    ln, ms, by = _prof(L, TAGS[tag], size)
    L.rmetis_hip_profile_enable(0)
    L.rmetis_hip_profile_filter(-1, -1)
    return elapsed, ln, ms, by


def roofline(tag, ln, ms, by, workload):
    avg_ms = ms / max(ln, 1)
    per = by / max(ln, 1)
    achieved = per / (avg_ms * 1e-3) / 1e9 if avg_ms > 0 else 0.0
    traffic, src = None, None
    tpath = os.path.join(ROOT, "profiles", f"traffic_{workload}.json")
    if os.path.exists(tpath):
        # HBM bytes per launch of each kernel family at the finest level, from
        # two rocprofv3 --pmc runs (FETCH_SIZE, WRITE_SIZE) scaled by the same
        # counters on a device copy of known size
        with open(tpath) as fh:
            tj = json.load(fh)
        fam = tj.get(tag)
        if isinstance(fam, dict) and fam.get("traffic_bytes_per_launch"):
            traffic, src = fam["traffic_bytes_per_launch"], tj.get("summary")
    return {"bound": "hbm", "kernel": TAG_KERNELS.get(tag, tag), "tag": tag, "achieved": round(achieved, 1),
            "peak": HBM_PEAK_GBS, "unit": "GB/s", "frac": round(achieved / HBM_PEAK_GBS, 4), "traffic": traffic,
            "traffic_source": src, "bytes_per_launch": int(per), "avg_launch_us": round(avg_ms * 1e3, 2),
            "launches": ln}


def cpu_baseline(jobs, seconds_one, seconds_pool, what):
    """The oracle (the C++ restatement of the reference) on host cores: jobs
    run one after another on one core for about seconds_one, then on a pool
    of threads over the host cores for about seconds_pool (the oracle releases
    the GIL inside each call).  Returns the cpu_baseline object."""
    sys.path.insert(0, os.path.join(ROOT, "oracle"))
    import pyoracle as O
    O.build()
    runs1, t1 = 0, 0.0
    for j in jobs:
        if t1 >= seconds_one:
            break
        t = time.perf_counter()
        j(O)
        t1 += time.perf_counter() - t
        runs1 += 1
    threads = host_threads()
    done = [0]
    lock = threading.Lock()
    stop_at = time.perf_counter() + seconds_pool
    nxt = [0]

    def worker():
        while time.perf_counter() < stop_at:
            with lock:
                i = nxt[0]
                nxt[0] += 1
            jobs[i % len(jobs)](O)
            with lock:
                done[0] += 1

    t0 = time.perf_counter()
    th = [threading.Thread(target=worker) for _ in range(threads)]
    for t in th:
        t.start()
    for t in th:
        t.join()
    tp = time.perf_counter() - t0
    return {"value": round(done[0] / tp, 4), "unit": "graphs/s", "cores": threads, "kind": "port",
            "sample": f"{what}: {done[0]} calls on a pool of {threads} host threads in {tp:.1f}s "
                      f"(independent calls; the reference has no parallelism inside one call)",
            "single_core": {"value": round(runs1 / t1, 4) if t1 > 0 else None, "unit": "graphs/s", "cores": 1,
                            "calls": runs1, "seconds": round(t1, 1)},
            "cpu_model": cpu_model()}


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=5)
    ap.add_argument("--warmup", type=int, default=1)
    ap.add_argument("--workload", default="batch1024",
                    help="batch1024 (default: uneven batch of that many graphs) | rgg20 | grid256 | rmat22 | "
                         "rggNN | rmatNN | serve | siblings")
    ap.add_argument("--serve-seconds", type=float, default=60.0, help="serve: length of the arrival stream")
    ap.add_argument("--serve-rate", type=float, default=200.0, help="serve: Poisson arrival rate (requests/s)")
    ap.add_argument("--serve-pool", type=int, default=192, help="serve: distinct graphs the requests draw from")
    ap.add_argument("--slots", type=int, default=3, help="serve: batched calls in flight at once")
    ap.add_argument("--serve-max-batch", type=int, default=0,
                    help="serve: vertices per batched call at most (0: the server's default)")
    ap.add_argument("--plain-sample", type=int, default=40, help="serve: requests timed on the plain path")
    ap.add_argument("--plain-stride", type=int, default=3,
                    help="batch workloads: the plain path is timed on every n-th graph (odd: grids and RGGs alike)")
    ap.add_argument("--seed", type=int, default=1)
    ap.add_argument("--cpu-budget-s", type=float, default=12.0, help="seconds of the one-core CPU baseline")
    ap.add_argument("--cpu-pool-s", type=float, default=12.0, help="seconds of the all-cores CPU baseline")
    ap.add_argument("--no-cpu", action="store_true")
    ap.add_argument("--no-extra", action="store_true", help="batch workloads: skip the rgg20 extra key")
    ap.add_argument("--no-host-api", action="store_true",
                    help="skip timing the host-buffer C entry point (PCIe-inclusive rate)")
    ap.add_argument("--copy-bytes", type=int, default=0,
                    help="issue one device-to-device copy of this many bytes first (a kernel of known "
                         "traffic, to scale FETCH_SIZE/WRITE_SIZE counters)")
    ap.add_argument("--rtype", default="greedy", choices=["greedy", "fm"],
                    help="single-graph workloads: refinement (fm: the reference's default k-way FM)")
    ap.add_argument("--roofline-tag", default="",
                    help="kernel family of the roofline (gain|match|contract; default: the dominant one)")
    args = ap.parse_args()
    if args.gpus != 1:
        print(json.dumps({"error": "this build runs on one GPU only; --gpus must be 1 (the path would shard "
                                   "batches and attempts across GPUs with one all-reduce, not built)",
                          "n_gpus": args.gpus}), flush=True)
        sys.exit(2)

    import rmetis_amd as R
    from rmetis_amd import _lib

    L = _lib.lib()
    if R.device_count() < 1:
        raise SystemExit("no HIP device")
    if args.copy_bytes > 0:
        src = L.rmetis_hip_device_alloc(args.copy_bytes)
        dst = L.rmetis_hip_device_alloc(args.copy_bytes)
        if not src or not dst or L.rmetis_hip_memcpy(dst, src, args.copy_bytes, 3) != 1:
            raise RuntimeError(f"calibration copy failed: {_lib.last_error()}")
        L.rmetis_hip_device_synchronize()
        L.rmetis_hip_device_free(src)
        L.rmetis_hip_device_free(dst)
        log(f"calibration copy of {args.copy_bytes} bytes done")
    if args.workload.startswith("batch"):
        out = bench_batch(args)
    elif args.workload == "serve":
        out = bench_serve(args)
    elif args.workload == "siblings":
        out = bench_siblings(args)
    else:
        out = bench_single(args, args.workload)
    print(json.dumps(out), flush=True)


def bench_single(args, workload, steps=None, warmup=None, extra=False):
    """Full k-way partitions of one graph resident in HBM (one graph per
    call, ncuts 1, seed 1)."""
    import numpy as np
    import rmetis_amd as R
    from rmetis_amd import _lib, synth

    L = _lib.lib()
    steps = args.steps * (4 if extra else 1) if steps is None else steps
    warmup = args.warmup if warmup is None else warmup
    t0 = time.time()
    g, k = synth.config_graph(workload, seed=args.seed)
    log(f"graph {g} k={k} generated in {time.time() - t0:.1f}s")
    DA = _lib.DeviceArray
    d_xadj, d_adj = DA(g.xadj), DA(g.adjncy)
    d_vw = DA(g.vwgt) if g.vwgt is not None else None
    d_aw = DA(g.adjwgt) if g.adjwgt is not None else None
    d_part = DA(count=g.n, dtype=np.int32)
    L.rmetis_hip_device_synchronize()
    rtype = R.RType.Fm if (args.rtype == "fm" and not extra) else R.RType.Greedy
    opts = R.Options(rtype=rtype, seed=args.seed, ncuts=1, niter=10, ufactor=30).to_c_array()
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

    for i in range(warmup):
        step()
        log(f"warmup {i} cut={obj.value}")
    total, finest = profiled_step(L, step, g.n)
    tag = args.roofline_tag or dominant(total)
    # the roofline kernel at the finest level only: its launches there are
    # the full-size ones
    elapsed, ln, ms, by = timed_steps(L, step, steps, tag, g.n)
    log(f"timed {steps} steps in {elapsed:.3f}s cut={obj.value}")
    part = d_part.numpy()
    pw = np.bincount(part, weights=(g.vwgt if g.vwgt is not None else None), minlength=k)
    tot = float(g.vwgt.sum()) if g.vwgt is not None else float(g.n)
    violations = int((pw > math.floor(tot / k * 1.03)).sum())
    levels = []
    for i in range(stats.nlevels):
        n_i, m_i = stats.level_n[i], stats.level_m[i]
        cm, rm, ps = stats.coarsen_ms[i], stats.refine_ms[i], stats.passes[i]
        levels.append({"n": n_i, "m": m_i, "coarsen_ms": round(cm, 3), "refine_ms": round(rm, 3), "passes": ps,
                       "coarsen_edges_per_s": round(m_i / (cm * 1e-3), 1) if cm > 0 else None,
                       "refine_edges_per_s_per_pass": round(m_i * ps / (rm * 1e-3), 1) if rm > 0 else None})
    value = steps / elapsed
    out = {
        "metric": "graphs/s (k-way partitions of one graph per call)",
        "value": round(value, 4), "unit": "graphs/s", "n_gpus": 1, "steps": steps, "warmup": warmup,
        "ms_per_step": round(1e3 * elapsed / steps, 3), "higher_is_better": True, "scaling": "weak",
        "vs_baseline": None, "dtype": "int32", "data": "synthetic",
        "config": {"workload": f"{workload}: {g.name} n={g.n} m={g.m} k={k}", "ctype": "shem",
                   "rtype": "fm" if rtype == R.RType.Fm else "greedy",
                   "ufactor": 30, "niter": 10, "ncuts": 1, "seed": args.seed, "parallelism": "single GPU"},
        "edge_cut": int(obj.value), "balance_violations": violations, "edges_per_s": round(g.m * value, 1),
        "kernel_ms_per_step": {n: round(v, 3) for n, v in total.items()},
        "level_kernels_finest": level_kernels(finest),
        "levels": levels,
        "roofline": roofline(tag, ln, ms, by, workload),
    }
    if extra:
        return out
    if not args.no_host_api:
        # the same partition through the host-buffer drop-in entry point
        nv, nc, npart = C.c_int32(g.n), C.c_int32(g.ncon), C.c_int32(k)
        i32 = _lib.i32p
        hx, ha = g.xadj.ctypes.data_as(i32), g.adjncy.ctypes.data_as(i32)
        hvw = g.vwgt.ctypes.data_as(i32) if g.vwgt is not None else None
        haw = g.adjwgt.ctypes.data_as(i32) if g.adjwgt is not None else None
        hpart = np.zeros(g.n, np.int32)
        hobj = C.c_int32()
        reps = max(1, min(steps, 3))
        t1 = time.perf_counter()
        for _ in range(reps):
            rc = L.RMETIS_PartGraphKway(C.byref(nv), C.byref(nc), hx, ha, hvw, None, haw, C.byref(npart), None,
                                        None, opts.ctypes.data_as(i32), C.byref(hobj), hpart.ctypes.data_as(i32))
            if rc != 1:
                raise RuntimeError(f"RMETIS_PartGraphKway failed rc={rc}: {_lib.last_error()}")
        t_host = time.perf_counter() - t1
        out["host_api"] = {"value": round(reps / t_host, 4), "unit": "graphs/s", "calls": reps,
                           "note": "RMETIS_PartGraphKway on host buffers: CSR upload, validation and the "
                                   "partition download included"}
    if not args.no_cpu:
        kw = dict(vwgt=g.vwgt, adjwgt=g.adjwgt, ncon=g.ncon, niter=10, ufactor=30,
                  rtype_greedy=rtype == R.RType.Greedy)
        jobs = [lambda O, s=s: O.part_graph_kway(g.xadj, g.adjncy, k, seed=s, **kw) for s in range(1, 9)]
        out["cpu_baseline"] = cpu_baseline(jobs, args.cpu_budget_s, args.cpu_pool_s,
                                           f"full k-way partitions of the same graph, seeds 1..8")
    return out


def bench_batch(args):
    """The uneven batch (see the module docstring), one batched call per step."""
    import numpy as np
    import rmetis_amd as R
    from rmetis_amd import _lib, synth

    L = _lib.lib()
    DA = _lib.DeviceArray
    i32 = _lib.i32p
    count = int(args.workload[5:] or 1024)
    ncuts = 8
    t0 = time.time()
    workers = host_threads()
    batch = synth.uneven_batch(count, args.seed, workers=workers)
    graphs = [g for g, _ in batch]
    ks = np.array([k for _, k in batch], np.int32)
    gvtx, xadj, adjncy, vwgt, adjwgt, ncon = R.pack_batch(graphs)
    n, m = int(gvtx[-1]), int(xadj[-1])
    log(f"batch of {count} graphs: n={n} m={m} generated and packed in {time.time() - t0:.1f}s ({workers} workers)")
    d_x, d_a, d_vw, d_aw = DA(xadj), DA(adjncy), DA(vwgt), DA(adjwgt)
    d_part = DA(count=n, dtype=np.int32)
    L.rmetis_hip_device_synchronize()
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
    # the finest level of a pass is the union of the attempts it runs together
    total, finest = profiled_step(L, step, lambda: stats.level_n[0])
    packed = int(stats.attempts_per_call)
    fin_n = int(stats.level_n[0])
    tag = args.roofline_tag or dominant(total)
    log(f"kernel ms per step {dict((k_, round(v, 1)) for k_, v in total.items())}; roofline kernel: {tag}; "
        f"attempts per pass {packed} (finest union level n={fin_n})")
    elapsed, ln, ms, by = timed_steps(L, step, args.steps, tag, fin_n)
    value = count * args.steps / elapsed
    log(f"batched: {args.steps} steps in {elapsed:.2f}s = {value:.2f} graphs/s")
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
    # the plain path (one graph, one attempt after another per call) on a
    # sample of the same graphs, same options
    sample = list(range(0, count, max(1, args.plain_stride)))
    dev = [(DA(graphs[i].xadj), DA(graphs[i].adjncy), DA(graphs[i].vwgt), DA(graphs[i].adjwgt),
            DA(count=graphs[i].n, dtype=np.int32)) for i in sample]
    L.rmetis_hip_device_synchronize()
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
    L.rmetis_hip_device_synchronize()
    mism = 0
    t1 = time.perf_counter()
    for j in range(len(sample)):
        if plain(j) != bobj[sample[j]]:
            mism += 1
    L.rmetis_hip_device_synchronize()
    t_plain = time.perf_counter() - t1
    plain_value = len(sample) / t_plain
    log(f"plain: {len(sample)} graphs in {t_plain:.2f}s = {plain_value:.2f} graphs/s, objval mismatches {mism}")
    for d in dev:
        for a in d:
            a.free()
    out = {
        "metric": METRIC,
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
        "config": {"workload": f"uneven batch of {count} graphs (n log-uniform 2^10..2^18, grids and RGGs, "
                               f"vwgt U{{1..4}}, adjwgt U{{1..10}}, k in {{2..64}}), {ncuts} seeds each, one "
                               f"segmented CSR: n={n} m={m}",
                   "ctype": "shem", "rtype": "greedy", "ufactor": 30, "niter": 10, "ncuts": ncuts,
                   "attempts_per_pass": packed, "seed": args.seed,
                   "parallelism": "single GPU, one segmented CSR (attempts packed as copies of the batch)"},
        "plain": {"value": round(plain_value, 3), "unit": "graphs/s", "graphs": len(sample),
                  "objval_mismatches": mism,
                  "note": "rmetis_hip_kway_device per graph (device-resident CSR, the 8 attempts one after "
                          "another), same options, every plain_stride-th graph"},
        "batched_vs_plain": round(value / plain_value, 2),
        "edge_cut_sum": int(bobj.sum()),
        "balance_violations": violations,
        "balance_violators": violators[:16],
        "edges_per_s": round(m * ncuts * value / count, 1),
        "kernel_ms_per_step": {k_: round(v, 2) for k_, v in total.items()},
        "level_kernels_finest": level_kernels(finest),
        "levels": levels,
        "roofline": roofline(tag, ln, ms, by, args.workload),
    }
    if not args.no_cpu:
        def job(i):
            g = graphs[i]
            return lambda O: O.part_graph_kway(g.xadj, g.adjncy, int(ks[i]), vwgt=g.vwgt, adjwgt=g.adjwgt,
                                               ncuts=ncuts, seed=args.seed, niter=10, ufactor=30)
        out["cpu_baseline"] = cpu_baseline([job(i) for i in sample], args.cpu_budget_s, args.cpu_pool_s,
                                           f"graphs of the plain-path sample, {ncuts} attempts each")
        log(f"cpu baseline: {out['cpu_baseline']['value']} graphs/s on {out['cpu_baseline']['cores']} threads")
    if not args.no_extra:
        out["rgg20"] = bench_single(args, "rgg20", extra=True)
        out["siblings"] = bench_siblings(args, extra=True)
    return out


def bench_siblings(args, extra=False):
    """The C-ABI siblings on graphs resident in HBM: recursive bisection
    (rmetis_hip_recursive_device, rgg 2^17, k = 64) and nested dissection
    (rmetis_hip_nodend_device, 256 x 256 grid), options as RMETIS_* with the
    default array except seed (SHEM, grow, FM, ufactor 30), against the
    oracle's restatement of the reference on the same graphs and options."""
    import numpy as np
    import rmetis_amd as R
    from rmetis_amd import _lib, synth

    L = _lib.lib()
    DA = _lib.DeviceArray
    steps = max(2, args.steps if not extra else 3)
    opts = R.Options(seed=args.seed).to_c_array()
    out = {}
    rb_g, rb_k = synth.rgg(1 << 17, args.seed, vwgt_max=4), 64
    nd_g = synth.grid(256, 256)
    for name, g in (("recursive", rb_g), ("nodend", nd_g)):
        d_x, d_a = DA(g.xadj), DA(g.adjncy)
        d_vw = DA(g.vwgt) if g.vwgt is not None else None
        d_aw = DA(g.adjwgt) if g.adjwgt is not None else None
        d_o1, d_o2 = DA(count=g.n, dtype=np.int32), DA(count=g.n, dtype=np.int32)
        obj = C.c_int32()

        def step():
            if name == "recursive":
                rc = L.rmetis_hip_recursive_device(g.n, g.ncon, d_x.ptr, d_a.ptr,
                                                   d_vw.ptr if d_vw is not None else None,
                                                   d_aw.ptr if d_aw is not None else None, rb_k, None, None,
                                                   opts.ctypes.data_as(_lib.i32p), C.byref(obj), d_o1.ptr)
            else:
                rc = L.rmetis_hip_nodend_device(g.n, d_x.ptr, d_a.ptr, d_vw.ptr if d_vw is not None else None,
                                                opts.ctypes.data_as(_lib.i32p), d_o1.ptr, d_o2.ptr)
            if rc != 1:
                raise RuntimeError(f"{name} failed rc={rc}: {_lib.last_error()}")

        step()
        L.rmetis_hip_device_synchronize()
        t0 = time.perf_counter()
        for _ in range(steps):
            step()
        L.rmetis_hip_device_synchronize()
        el = time.perf_counter() - t0
        e = {"value": round(steps / el, 3), "unit": "graphs/s", "ms_per_call": round(1e3 * el / steps, 2),
             "steps": steps, "graph": f"{g.name} n={g.n} m={g.m}"}
        if name == "recursive":
            e["k"] = rb_k
            e["edge_cut"] = int(obj.value)
        else:
            perm = d_o1.numpy()
            e["perm_is_permutation"] = bool(np.array_equal(np.sort(perm), np.arange(g.n)))
        if not args.no_cpu:
            sys.path.insert(0, os.path.join(ROOT, "oracle"))
            import pyoracle as O
            O.build()
            t1, calls, cuts = 0.0, 0, []
            while calls < 4 and (calls == 0 or t1 < (4.0 if extra else args.cpu_budget_s)):
                t = time.perf_counter()
                if name == "recursive":
                    _, cut, _ = O.part_graph_recursive(g.xadj, g.adjncy, rb_k, vwgt=g.vwgt, adjwgt=g.adjwgt,
                                                    ncon=g.ncon, iptype=0, rtype_greedy=False,
                                                    seed=args.seed + calls, ufactor=30)
                    cuts.append(int(cut))
                else:
                    O.node_nd(g.xadj, g.adjncy, vwgt=g.vwgt, seed=args.seed + calls)
                t1 += time.perf_counter() - t
                calls += 1
            e["cpu_baseline"] = {"value": round(calls / t1, 4), "unit": "graphs/s", "cores": 1, "kind": "port",
                                 "sample": f"{calls} calls of the oracle on the same graph, seeds "
                                           f"{args.seed}..{args.seed + calls - 1}"}
            if cuts:
                e["oracle_edge_cut_geomean"] = round(float(np.exp(np.mean(np.log(cuts)))), 1)
        for a in (d_x, d_a, d_vw, d_aw, d_o1, d_o2):
            if a is not None:
                a.free()
        out[name] = e
        log(f"{name}: {e['value']} graphs/s on the GPU ({e['ms_per_call']} ms per call)")
    if extra:
        return out
    return {"metric": "graphs/s (recursive bisection and nested dissection calls)",
            "value": out["recursive"]["value"], "unit": "graphs/s", "n_gpus": 1, "steps": steps,
            "warmup": 1, "ms_per_step": out["recursive"]["ms_per_call"], "higher_is_better": True,
            "scaling": "weak", "vs_baseline": None, "dtype": "int32", "data": "synthetic",
            "config": {"workload": "siblings: recursive bisection of rgg 2^17 (k = 64), nested dissection of "
                                   "grid 256 x 256", "seed": args.seed},
            **out}


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
    return out


if __name__ == "__main__":
    main()
