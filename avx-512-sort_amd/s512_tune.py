#!/usr/bin/env python3
"""s512_tune.py -- autotuner of the sort library's kernel shapes (run on the GPU).

The library's kernels come in several shapes (s512_options in include/s512.h):
  count_variant     threads x keys per thread x keys classified together
  scatter_variant   scatter workgroup shape, stored bucket ids, write combining
  tile_variant      keys per thread of the bitonic tile sorts
  target_leaf       mean leaf (bitonic tile) size the partition aims for
  max_fanout_log2   cap on the partition fan-out per level
Which shape is fastest depends on the key type and the number of keys.  For
every workload kind (int32, double, int32 key/value, segmented int32,
segmented double) and size class this script runs a coordinate descent over
the knobs (two passes, each knob in turn over all its values with the others
held) on device-generated keys, timing every candidate with HIP events
(median of three sorts after one warm-up) and scoring it by the sum of its
times over several input distributions (uniform, normal, sorted, 1000
distinct values, as the kind allows).

Correctness does not depend on any knob: every candidate's output is compared
bit for bit with the output of the library's defaults, which is itself
checked against numpy's sort on the host (keys up to 2^24; larger classes by
non-decreasing order and a multiset hash of the full array).

Output: tune_table.json (the measurements) and csrc/s512_tune_table.inc (the
table the library compiles in as its defaults; rebuild afterwards with
`make -C avx-512-sort_amd`).

    python avx-512-sort_amd/s512_tune.py            # every class
    python avx-512-sort_amd/s512_tune.py --quick    # one pass, smaller sizes
"""
import argparse
import json
import os
import statistics
import sys
import time

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, HERE)

KNOBS = {
    "count_variant": [1, 2, 3, 4, 5],
    "scatter_variant": [1, 2, 3, 4, 5, 6, 7, 8, 9, 10, 11],
    "tile_variant": [1, 2, 3],
    "target_leaf": [512, 1024, 2048, 4096],
    "max_fanout_log2": [7, 8, 9],
}
START = {"count_variant": 3, "scatter_variant": 1, "tile_variant": 3, "target_leaf": 1024, "max_fanout_log2": 9}
KIND_ID = {"i32": 0, "f64": 1, "kv": 2, "seg_i32": 3, "seg_f64": 4}
SCATTER_SHAPES = 4  # include/s512.h: scatter_variant = 1 + shape + 4 * reclassify (1..8)
# 9, 10: the line-buffered scatter (table shape 4) with stored ids / classifying;
# 11: the full-line scatter (table shape 5, stored ids)
SCATTER_SPECIAL = {9: (4, 1), 10: (4, 0), 11: (5, 1)}


def scatter_columns(variant):
    """(scatter_shape, ids) table columns of a scatter_variant."""
    if variant in SCATTER_SPECIAL:
        return SCATTER_SPECIAL[variant]
    sv = variant - 1
    return sv % SCATTER_SHAPES, 1 if sv < SCATTER_SHAPES else 0

# (kind, lg_lo, lg_hi, representative n): n in [2^lg_lo, 2^lg_hi) uses the entry
CLASSES = [
    ("i32", 0, 15, 1 << 14), ("i32", 15, 19, 1 << 18), ("i32", 19, 23, 1 << 22), ("i32", 23, 27, 1 << 26),
    ("i32", 27, 64, 1 << 28),
    ("f64", 0, 14, 1 << 13), ("f64", 14, 19, 1 << 18), ("f64", 19, 23, 1 << 22), ("f64", 23, 27, 1 << 26),
    ("f64", 27, 64, 1 << 28),
    ("kv", 0, 20, 1 << 18), ("kv", 20, 25, 1 << 23), ("kv", 25, 64, 1 << 28),
    ("seg_i32", 0, 64, 200_000), ("seg_f64", 0, 64, 200_000),
]


def _hash(t):
    import torch
    x = t.view(torch.int64) if t.dtype == torch.float64 else t.to(torch.int64)
    x = x * -7046029254386353131
    x = x ^ (x >> 29)
    x = x * -4658895280553007687
    return int((x ^ (x >> 32)).sum().item())


class Workload:
    """Device input of one class, the call that sorts it, and its checked reference output."""

    def __init__(self, s512, kind, n, seed=0, dist="uniform"):
        import torch
        self.s512, self.kind = s512, kind
        dev = torch.device("cuda:0")
        self.vals = None
        self.off = None
        if kind in ("seg_i32", "seg_f64"):
            off = s512.lognormal_offsets(n, seed=seed)  # n segments, median 512, sigma 1.5
            self.off = torch.from_numpy(off.view(np.int64)).to(dev)
            nk = int(off[-1])
        else:
            nk = n
        dt = torch.float64 if kind in ("f64", "seg_f64") else torch.int32
        self.src = torch.empty(nk, dtype=dt, device=dev)
        s512.fill(self.src, dist, seed=seed + 1, param=1000)
        if kind == "kv":
            self.vsrc = torch.arange(nk, dtype=torch.int32, device=dev)
            self.vals = torch.empty_like(self.vsrc)
        self.work = torch.empty_like(self.src)
        self.nkeys = nk
        self.ref = None
        self.ref = self._reference()

    def run(self, opts):
        s512 = self.s512
        o = s512.Options()
        for k, v in opts.items():
            setattr(o, k, v)
        if self.kind == "kv":
            s512.kv_sort_ex(self.work, self.vals, o)
        elif self.off is not None:
            s512.segsort(self.work, self.off, options=o)
        else:
            s512.sort_ex(self.work, o)

    def time(self, opts, reps=3):
        import torch
        ts = []
        for r in range(reps + 1):
            self.work.copy_(self.src)
            if self.vals is not None:
                self.vals.copy_(self.vsrc)
            e0, e1 = torch.cuda.Event(enable_timing=True), torch.cuda.Event(enable_timing=True)
            e0.record()
            self.run(opts)
            e1.record()
            e1.synchronize()
            if r:
                ts.append(e0.elapsed_time(e1))
        self.check()
        return statistics.median(ts)

    def _reference(self):
        import torch
        self.time(dict(START), reps=1)
        out = self.work.clone()
        if self.nkeys <= (1 << 24):
            a = self.src.cpu().numpy()
            if self.off is not None:
                off = self.off.cpu().numpy()
                want = a.copy()
                for i in range(len(off) - 1):
                    want[off[i]:off[i + 1]] = np.sort(a[off[i]:off[i + 1]])
            else:
                want = np.sort(a)
            got = out.cpu().numpy()
            if not np.array_equal(got.view(np.uint8), want.view(np.uint8)):
                raise SystemExit(f"{self.kind} n={self.nkeys}: default output differs from numpy's sort")
        elif self.off is not None:
            # every segment non-decreasing, and the same per-segment multiset hash
            step = torch.ones(self.nkeys, dtype=torch.bool, device=out.device)
            starts = self.off[:-1]
            step[starts[starts < self.nkeys]] = False  # a segment's first key has no predecessor in it
            ok = bool(torch.all((out[1:] >= out[:-1]) | ~step[1:]).item())
            if not ok or not torch.equal(self._seg_hash(out), self._seg_hash(self.src)):
                raise SystemExit(f"{self.kind} n={self.nkeys}: default output not sorted per segment")
        else:
            if not bool(torch.all(out[1:] >= out[:-1]).item()) or _hash(out) != _hash(self.src):
                raise SystemExit(f"{self.kind} n={self.nkeys}: default output not a sorted permutation")
        if self.vals is not None:
            self.vref = self.vals.clone()
        return out

    def _seg_hash(self, t):
        import torch
        x = t.view(torch.int64) if t.dtype == torch.float64 else t.to(torch.int64)
        x = x * -7046029254386353131
        x = x ^ (x >> 29)
        c = torch.cat([torch.zeros(1, dtype=torch.int64, device=t.device), torch.cumsum(x, 0)])
        return c[self.off[1:]] - c[self.off[:-1]]

    def check(self):
        import torch
        if self.ref is None:
            return
        a, b = self.work, self.ref
        if a.dtype == torch.float64:
            a, b = a.view(torch.int64), b.view(torch.int64)
        if not torch.equal(a, b):
            raise SystemExit(f"{self.kind} n={self.nkeys}: a tuned configuration changed the sorted keys")
        if self.vals is not None:  # the values may differ among equal keys; their multiset per key may not
            pk = self.work.to(torch.int64) * (1 << 32) + self.vals.to(torch.int64)
            pr = self.ref.to(torch.int64) * (1 << 32) + self.vref.to(torch.int64)
            if not torch.equal(torch.sort(pk).values, torch.sort(pr).values):
                raise SystemExit(f"{self.kind} n={self.nkeys}: key/value pairs changed")


DISTS = {"f64": ["normal", "uniform", "sorted", "distinct"], "i32": ["uniform", "sorted", "distinct"],
         "kv": ["uniform", "distinct"], "seg_f64": ["normal"], "seg_i32": ["uniform"]}


class MultiWorkload:
    """The sum of the times of one class over several input distributions."""

    def __init__(self, s512, kind, n):
        self.parts = [Workload(s512, kind, n, dist=d) for d in DISTS[kind]]

    def time(self, opts):
        return sum(w.time(opts) for w in self.parts)


def tune_class(s512, kind, n, passes, log):
    w = MultiWorkload(s512, kind, n)
    best = dict(START)
    best_t = w.time(best)
    tried = {tuple(sorted(best.items())): best_t}
    for p in range(passes):
        for knob, values in KNOBS.items():
            for v in values:
                cand = dict(best, **{knob: v})
                key = tuple(sorted(cand.items()))
                if key not in tried:
                    tried[key] = w.time(cand)
                if tried[key] < best_t * 0.995:  # a clear win
                    best, best_t = cand, tried[key]
        log(f"  {kind} n={n} pass {p + 1}: {best_t:.3f} ms {best}")
    del w
    import torch
    torch.cuda.empty_cache()
    return best, best_t, len(tried)


def write_inc(rows, path):
    lines = ["// s512_tune_table.inc -- autotuned defaults (generated by s512_tune.py; do not edit by hand).",
             "// kind: 0 int32, 1 double, 2 int32 key/value, 3 segmented int32, 4 segmented double.",
             "// Columns: kind, lg_lo, lg_hi, count_shape, scatter_shape, ids, tile_plan, target_leaf, max_logf.",
             "static const TuneEntry kTune[] = {"]
    for r in rows:
        o = r["options"]
        shape, ids = scatter_columns(o["scatter_variant"])
        lines.append(f"    {{{KIND_ID[r['kind']]}, {r['lg_lo']}, {r['lg_hi']}, {o['count_variant'] - 1}, "
                     f"{shape}, {ids}, {o['tile_variant'] - 1}, "
                     f"{o['target_leaf']}, {o['max_fanout_log2']}}},  // n ~ {r['n']}: {r['ms']:.3f} ms summed")
    lines.append("};")
    with open(path, "w") as f:
        f.write("\n".join(lines) + "\n")


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--quick", action="store_true", help="one pass, sizes capped at 2^24")
    ap.add_argument("--out", default=os.path.join(HERE, "tune_table.json"))
    ap.add_argument("--inc", default=os.path.join(HERE, "csrc", "s512_tune_table.inc"))
    args = ap.parse_args()
    import torch
    import s512
    if not torch.cuda.is_available():
        raise SystemExit("the autotuner needs a GPU")
    passes = 1 if args.quick else 2
    rows = []
    t0 = time.time()
    log = lambda m: print(m, flush=True)  # noqa: E731
    for kind, lo, hi, n in CLASSES:
        if args.quick and kind not in ("seg_i32", "seg_f64"):
            n = min(n, 1 << 24)
        best, ms, ntried = tune_class(s512, kind, n, passes, log)
        rows.append({"kind": kind, "lg_lo": lo, "lg_hi": hi, "n": n, "ms": ms, "candidates": ntried,
                     "options": best})
    doc = {"device": torch.cuda.get_device_name(0), "passes": passes, "seconds": time.time() - t0,
           "distributions": DISTS,
           "method": "coordinate descent, median of 3 after 1 warm-up, HIP events; every candidate checked "
                     "bit for bit against the default output, itself checked against numpy",
           "classes": rows}
    with open(args.out, "w") as f:
        json.dump(doc, f, indent=1)
    write_inc(rows, args.inc)
    log(f"wrote {args.out} and {args.inc} ({time.time() - t0:.0f} s)")


if __name__ == "__main__":
    main()
