"""Autotuner for the tuning table of librmetis_hip.so (tuning.hpp).

    python -m rmetis_amd.autotune            # measure and print the table
    python -m rmetis_amd.autotune --write    # ... and write rmetis_amd/tuning.json

Needs one HIP device.  Every entry of the table is exact -- partitions do not
depend on it -- so the tuner only times; it checks that every timed call
returns the same edge cut whatever the entry.

* Suitor matching shape per level bucket (level size x mean degree): every
  shape is forced for every level of full k-way partitions of graphs
  spanning the buckets (a grid, random geometric graphs of mean degree
  6-20, an R-MAT graph, and one batched call over an uneven batch whose
  union levels reach tens of millions of vertices); each level's matching
  time (the library's profiler, HIP events around the matching launches of
  that level) is added to its bucket's total for that shape, and each bucket
  takes the shape with the least total.  Buckets no level fell into keep
  their current shape.
* incremental_min_n and spec_min_per_part: whole partitions (greedy
  refinement) of the same graphs timed for each candidate value.
"""
import argparse
import ctypes as C
import json
import math
import os
import sys
import time

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
TABLE = os.path.join(HERE, "tuning.json")
SIZES = ["n<2^19", "2^19<=n<2^21", "2^21<=n<2^23", "n>=2^23"]
DEGREES = ["[0,6)", "[6,10)", "[10,16)", "[16,24)", "[24,48)", "[48,inf)"]
BUCKETS = [f"{a} deg{b}" for a in SIZES for b in DEGREES]
EDGES = [6, 10, 16, 24, 48]
NB = len(BUCKETS)
NT = NB + 2
SHAPES = ["1 lane x 4", "2 lanes x 4", "4 lanes x 2", "8 lanes x 1"]
MATCH_TAG = 3


def bucket(n, m):
    sb = 0
    while sb < len(SIZES) - 1 and n >= 1 << (19 + 2 * sb):
        sb += 1
    avg = m / n if n > 0 else 0.0
    db = 0
    while db < len(EDGES) and avg >= EDGES[db]:
        db += 1
    return sb * len(DEGREES) + db


def graphs(scale):
    from rmetis_amd import synth
    n = 1 << scale
    out = [(synth.grid(int(math.sqrt(n)), int(math.sqrt(n))), 8)]
    for d in (6.0, 9.0, 14.0, 20.0):
        out.append((synth.rgg(n, 7, r=math.sqrt(d / (math.pi * n)), vwgt_max=4, name=f"rgg{n}_d{d:g}"), 64))
    out.append((synth.rmat(scale - 2, 16, 5), 64))
    return out


class BatchRunner:
    """One batched call over an uneven batch (its union levels are large)."""

    def __init__(self, count, ncuts):
        import rmetis_amd as R
        from rmetis_amd import _lib, synth
        self.L, self._lib = _lib.lib(), _lib
        members = synth.uneven_batch(count, 3)
        graphs = [g for g, _ in members]
        self.ks = np.array([k for _, k in members], np.int32)
        gvtx, x, a, vw, aw, _ = R.pack_batch([R.Graph(g.n, g.ncon, g.xadj, g.adjncy, g.vwgt, g.adjwgt)
                                             for g in graphs])
        self.gvtx = np.asarray(gvtx, np.int32)
        self.count, self.n, self.m = count, int(self.gvtx[-1]), int(x[-1])
        DA = _lib.DeviceArray
        self.d = [DA(x), DA(a), DA(vw), DA(aw), DA(count=self.n, dtype=np.int32)]
        self.obj = np.zeros(count, np.int32)
        self.stats = _lib.HipStats()
        self.ncuts = ncuts
        self.name = f"batch of {count} uneven graphs, {ncuts} attempts (n={self.n}, m={self.m})"

    def run(self, opts_greedy):
        o = opts_greedy.copy()
        o[7] = self.ncuts
        d = self.d
        rc = self.L.rmetis_hip_kway_batch_device(self.count, self.gvtx.ctypes.data_as(self._lib.i32p), 1, d[0].ptr,
                                                 d[1].ptr, d[2].ptr, d[3].ptr,
                                                 self.ks.ctypes.data_as(self._lib.i32p), None,
                                                 o.ctypes.data_as(self._lib.i32p), None,
                                                 self.obj.ctypes.data_as(self._lib.i32p), d[4].ptr,
                                                 C.byref(self.stats))
        if rc != 1:
            raise RuntimeError(f"batch failed: {self._lib.last_error()}")
        return int(self.obj.astype(np.int64).sum())

    def free(self):
        for a in self.d:
            a.free()


class Runner:
    def __init__(self, g, k):
        from rmetis_amd import _lib
        self.L = _lib.lib()
        self._lib = _lib
        self.g, self.k = g, k
        DA = _lib.DeviceArray
        self.d = [DA(g.xadj), DA(g.adjncy), DA(g.vwgt) if g.vwgt is not None else None,
                  DA(g.adjwgt) if g.adjwgt is not None else None, DA(count=g.n, dtype=np.int32)]
        self.stats = _lib.HipStats()
        self.name = f"{g.name} (n={g.n}, mean degree {g.m / g.n:.1f}, k={k})"

    def run(self, opts):
        L, d, obj = self.L, self.d, C.c_int32()
        rc = L.rmetis_hip_kway_device(self.g.n, self.g.ncon, d[0].ptr, d[1].ptr,
                                      d[2].ptr if d[2] is not None else None,
                                      d[3].ptr if d[3] is not None else None, self.k, None, None,
                                      opts.ctypes.data_as(self._lib.i32p), C.byref(obj), d[4].ptr,
                                      C.byref(self.stats))
        if rc != 1:
            raise RuntimeError(f"kway failed: {self._lib.last_error()}")
        return obj.value

    def free(self):
        for a in self.d:
            if a is not None:
                a.free()


def get_table(L):
    v = (C.c_int32 * NT)()
    L.rmetis_hip_get_tuning(v, NT)
    return list(v)


def set_table(L, t):
    from rmetis_amd import _lib
    v = (C.c_int32 * NT)(*t)
    if L.rmetis_hip_set_tuning(v, NT) != 1:
        raise RuntimeError(_lib.last_error())


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--scale", type=int, default=19, help="graphs of 2^scale vertices")
    ap.add_argument("--reps", type=int, default=2, help="timed calls per setting")
    ap.add_argument("--batch", type=int, default=512, help="graphs of the uneven batch (0: none)")
    ap.add_argument("--write", action="store_true", help="write the table to rmetis_amd/tuning.json")
    args = ap.parse_args()
    import rmetis_amd as R
    from rmetis_amd import _lib
    if R.device_count() < 1:
        raise SystemExit("the autotuner needs a HIP device")
    L = _lib.lib()
    base = get_table(L)
    opts = R.Options(rtype=R.RType.Greedy, seed=1, ncuts=1, niter=10, ufactor=30).to_c_array()
    runners = []
    for g, k in graphs(args.scale):
        runners.append(Runner(g, k))
        print(f"[autotune] {runners[-1].name}", file=sys.stderr, flush=True)
    big = []
    if args.batch > 0:
        big.append(BatchRunner(args.batch, 4))
        print(f"[autotune] {big[-1].name}", file=sys.stderr, flush=True)

    # suitor shapes per bucket
    ms = [[0.0] * len(SHAPES) for _ in BUCKETS]
    seen = [0] * len(BUCKETS)
    cuts = {}
    for s in range(len(SHAPES)):
        set_table(L, [s] * NB + base[NB:])
        for r in runners + big:
            r.run(opts)                                         # warm-up
            L.rmetis_hip_profile_filter(-1, -1)
            L.rmetis_hip_profile_enable(1)
            for _ in range(args.reps):
                cut = r.run(opts)
            L.rmetis_hip_device_synchronize()
            cuts.setdefault(r.name, set()).add(cut)
            st = r.stats
            for i in range(st.nlevels):
                n_i, m_i = st.level_n[i], st.level_m[i]
                ln, t, by = C.c_int64(), C.c_double(), C.c_double()
                L.rmetis_hip_profile_read(MATCH_TAG, n_i, C.byref(ln), C.byref(t), C.byref(by))
                if ln.value > 0:
                    b = bucket(n_i, m_i)
                    ms[b][s] += t.value / args.reps
                    if s == 0:
                        seen[b] += 1
            L.rmetis_hip_profile_enable(0)
        print(f"[autotune] shape {SHAPES[s]} done", file=sys.stderr, flush=True)
    for name, c in cuts.items():
        if len(c) != 1:
            raise RuntimeError(f"{name}: cuts differ across suitor shapes {sorted(c)}")
    table = list(base)
    for b in range(len(BUCKETS)):
        if seen[b]:
            table[b] = int(np.argmin(ms[b]))

    def total_ms(t):
        set_table(L, t)
        tot = 0.0
        for r in runners:
            r.run(opts)
            L.rmetis_hip_device_synchronize()
            t0 = time.perf_counter()
            for _ in range(args.reps):
                r.run(opts)
            L.rmetis_hip_device_synchronize()
            tot += (time.perf_counter() - t0) * 1e3 / args.reps
        return tot

    inc = {}
    for v in (1 << 16, 1 << 17, 1 << 18, 1 << 19, 1 << 21):
        inc[v] = total_ms(table[:NB] + [v, table[NB + 1]])
    table[NB] = min(inc, key=inc.get)
    spec = {}
    for v in (8, 16, 24, 32, 64):
        spec[v] = total_ms(table[:NB + 1] + [v])
    table[NB + 1] = min(spec, key=spec.get)
    set_table(L, base)
    for r in runners + big:
        r.free()

    out = {
        "comment": "Written by python -m rmetis_amd.autotune on one MI355X; read by librmetis_hip.so on first use.",
        "suitor_shape_by_level_bucket": table[:NB],
        "incremental_min_n": table[NB],
        "spec_min_per_part": table[NB + 1],
        "measured": {
            "graphs": [r.name for r in runners + big],
            "buckets": BUCKETS,
            "shapes": SHAPES,
            "matching_ms_by_bucket_and_shape": [[round(x, 3) for x in row] for row in ms],
            "levels_per_bucket": seen,
            "partition_ms_by_incremental_min_n": {str(k): round(v, 2) for k, v in inc.items()},
            "partition_ms_by_spec_min_per_part": {str(k): round(v, 2) for k, v in spec.items()},
        },
    }
    text = json.dumps(out, indent=1)
    print(text)
    if args.write:
        with open(TABLE, "w") as fh:
            fh.write(text + "\n")
        print(f"[autotune] wrote {TABLE}", file=sys.stderr)


if __name__ == "__main__":
    main()
