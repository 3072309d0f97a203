#!/usr/bin/env python3
"""bench.py -- sorted keys/s and % of HBM peak for 268,435,456 fp64 keys.

Workload: the plain sort path (Sort512::Sort / SortOmp, sort512.hpp:5960,
5993) on 268,435,456 fp64 keys in device memory, five synthetic
distributions generated on the GPU: uniform [0,1), standard normal, 1000
distinct values, sorted ascending, sorted descending.  One step sorts each of
the five arrays once (5 x 2^28 keys); the unsorted input is restored from a
pristine copy before every sort, outside the timed events.

Output: one JSON line.
  value         keys sorted / device time of the sorts (HIP events on the
                stream the library runs on)
  roofline      the kernel phase with the largest share of device time, timed
                per launch with HIP events inside the library
                (s512_options.profile); algorithmic bytes = one read and one
                write of every key (16 B per fp64 key) per scatter / tile /
                copy launch, one read (8 B) per count launch
  traffic       HBM bytes per launch of that kernel, from two rocprofv3 --pmc
                child runs of this script on the uniform input (FETCH_SIZE,
                then WRITE_SIZE), each counter scaled by what it reads for the
                run's own 2 GiB restore copy (a copy kernel of known size);
                measured_frac = traffic / launch time / peak
  learned       the learned partitioner against sampled splitters on the
                skewed mixture and on uniform keys (same input, both exact)
  cpu_baseline  the reference's SortOmpPartition, compiled from the reference
                headers into oracle/_ref when they were available at build
                time, on a bounded sample on this host's cores; if that
                library or AVX-512 is absent, the oracle's C restatement (one
                core) is timed instead

Every sorted array is verified at full size by size-independent properties:
non-decreasing order and an order-independent multiset hash equal to the
input's.
"""
import argparse
import csv
import glob
import json
import os
import shutil
import subprocess
import sys
import tempfile
import time

ROOT = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, os.path.join(ROOT, "avx-512-sort_amd"))

HBM_PEAK_GBS = 8000.0
DISTS = ["uniform", "normal", "distinct", "sorted", "reversed"]
KERNEL_OF_PHASE = {"count": "k_count", "scatter": "k_scatter", "tile": "k_tile_sort", "copy": "k_copy_back"}


def parse():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    # five timed steps after two warm-ups (each step sorts every distribution once)
    ap.add_argument("--steps", type=int, default=5)
    ap.add_argument("--warmup", type=int, default=2)
    ap.add_argument("--n", type=int, default=268_435_456)
    ap.add_argument("--dists", default=",".join(DISTS))
    ap.add_argument("--no-cpu-baseline", action="store_true")
    ap.add_argument("--no-traffic", action="store_true", help="skip the rocprofv3 --pmc child runs")
    ap.add_argument("--no-learned", action="store_true", help="skip the learned-vs-splitters comparison")
    ap.add_argument("--cpu-sample", type=int, default=1 << 28)  # the full 268,435,456-key uniform input
    ap.add_argument("--workload", default="plain", choices=list(WORKLOADS),
                    help="plain: the 256M fp64 headline (default); the others time the secondary workloads "
                         "(256M int32, host-pointer end to end, int64 keys and fp64 argsort, top-k, key/value Zipf, "
                         "segmented batches, learned 1G) "
                         "and print their own JSON line")
    return ap.parse_args()


def multiset_hash(t):
    """Order-independent hash of the bit patterns (sum of a 64-bit mix)."""
    import torch
    x = t.view(torch.int64) if t.dtype == torch.float64 else t.to(torch.int64)
    x = x * -7046029254386353131  # 0x9E3779B97F4A7C15 as int64 (wraps)
    x = x ^ (x >> 29)
    x = x * -4658895280553007687  # 0xBF58476D1CE4E5B9
    x = x ^ (x >> 32)
    return int(x.sum().item())


def is_sorted(t):
    import torch
    if t.numel() < 2:
        return True
    return bool(torch.all(t[1:] >= t[:-1]).item())


def cpu_baseline(sample_n):
    """Reference SortOmpPartition (or the oracle restatement) on the host."""
    import ctypes
    import numpy as np
    sys.path.insert(0, os.path.join(ROOT, "oracle"))
    import oracle
    a0 = oracle.gen_f64("uniform", sample_n, 0)
    ref = oracle.ref_lib() if oracle.cpu_has_avx512() else None
    best = None
    runs = 0
    t_budget = time.time()
    if ref is not None:
        kind, cores = "reference", int(ref.ref_max_threads())
        fn = ref.ref_sort_omp_partition_f64
        what = "Sort512::SortOmpPartition"
    else:
        kind, cores = "port", 1
        what = "oracle restatement of Sort512::Sort"
        sample_n = min(sample_n, 1 << 23)
        a0 = a0[:sample_n].copy()
        fn = None
    while runs < 3 and (runs == 0 or time.time() - t_budget < 20):
        a = a0.copy()
        t0 = time.perf_counter()
        if fn is not None:
            fn(ctypes.c_void_p(a.ctypes.data), a.size)
        else:
            oracle.sort(a)
        dt = time.perf_counter() - t0
        assert np.all(a[1:] >= a[:-1])
        best = dt if best is None else min(best, dt)
        runs += 1
    out = {"value": sample_n / best, "unit": "keys/s", "cores": cores, "kind": kind,
           "sample": f"{what}, {sample_n} fp64 keys uniform [0,1) (SplitMix64 seed 0), best of {runs}",
           "cpu": _cpu_model()}
    if ref is not None:
        # the reference's other entry points on the same input, one run each
        # (Sort on one core on a 2^23-key prefix, to keep the run bounded)
        extra = {}
        fns = [("Sort512::Sort (1 core)", "ref_sort_f64", 1 << 23)]
        if cores <= 128 and cores & (cores - 1) == 0:  # the merge variants need 2^k threads (sort512.hpp:6006)
            fns += [("Sort512::SortOmpMerge", "ref_sort_omp_merge_f64", sample_n),
                    ("Sort512::SortOmpMergeDeps", "ref_sort_omp_mergedeps_f64", sample_n),
                    ("Sort512::SortOmpParMerge", "ref_sort_omp_parmerge_f64", sample_n)]
        for name, fn_name, m in fns:
            a = a0[:m].copy()
            t0 = time.perf_counter()
            getattr(ref, fn_name)(ctypes.c_void_p(a.ctypes.data), a.size)
            dt = time.perf_counter() - t0
            if not np.all(a[1:] >= a[:-1]):
                raise RuntimeError(f"{name} did not sort")
            extra[name] = {"keys": m, "keys_per_s": m / dt}
        out["other_reference_functions"] = extra
    return out


def _cpu_model():
    try:
        with open("/proc/cpuinfo") as f:
            for line in f:
                if line.startswith("model name"):
                    return line.split(":", 1)[1].strip()
    except OSError:
        pass
    return None


# ------------------------------------------------------------------ traffic
def _pmc_run(counter, n, tmp):
    """One rocprofv3 --pmc child run of this script; returns rows of its CSV."""
    out = os.path.join(tmp, counter)
    cmd = ["rocprofv3", "--pmc", counter, "--output-format", "csv", "-d", out, "-o", "p", "--",
           sys.executable, os.path.abspath(__file__), "--steps", "1", "--warmup", "0", "--n", str(n),
           "--dists", "uniform", "--no-cpu-baseline", "--no-traffic", "--no-learned"]
    subprocess.run(cmd, check=True, timeout=240, stdout=subprocess.DEVNULL, stderr=subprocess.DEVNULL)
    files = glob.glob(os.path.join(out, "**", "*counter_collection.csv"), recursive=True)
    if not files:
        raise RuntimeError(f"rocprofv3 wrote no counter file for {counter}")
    with open(files[0]) as f:
        return list(csv.DictReader(f))


def measure_traffic(n, kernel):
    """HBM bytes per dispatch of `kernel` (mean over its dispatches), each
    counter scaled by its reading for the 2 GiB restore copy of the run."""
    if shutil.which("rocprofv3") is None:
        return None, "rocprofv3 not found"
    tmp = tempfile.mkdtemp(prefix="s512_pmc_")
    try:
        copy_bytes = n * 8
        res = {}
        for counter in ("FETCH_SIZE", "WRITE_SIZE"):
            rows = _pmc_run(counter, n, tmp)
            copies = [float(r["Counter_Value"]) for r in rows if "copyBuffer" in r["Kernel_Name"]]
            if not copies:
                raise RuntimeError("no copy kernel in the counter run")
            scale = copy_bytes / (max(copies) * 1024.0)  # the restore copy is the largest
            vals = [float(r["Counter_Value"]) * 1024.0 * scale for r in rows if kernel in r["Kernel_Name"]]
            if not vals:
                raise RuntimeError(f"no {kernel} dispatch in the counter run")
            res[counter] = (sum(vals) / len(vals), scale)
        traffic = res["FETCH_SIZE"][0] + res["WRITE_SIZE"][0]
        return traffic, {"fetch_bytes": res["FETCH_SIZE"][0], "write_bytes": res["WRITE_SIZE"][0],
                         "fetch_scale": res["FETCH_SIZE"][1], "write_scale": res["WRITE_SIZE"][1]}
    finally:
        shutil.rmtree(tmp, ignore_errors=True)


# ------------------------------------------------------------------ learned
def learned_comparison(s512, n, reps=3):
    """Sampled splitters vs the learned partitioner on the same input."""
    import torch
    dev = torch.device("cuda:0")
    src = torch.empty(n, dtype=torch.float64, device=dev)
    work = torch.empty_like(src)
    out = {}
    for kind in ("mixture", "uniform"):
        s512.fill(src, kind, seed=7)
        times = {"splitters": [], "learned": []}
        ref = None
        info = {}
        for rep in range(reps + 1):
            for name in ("splitters", "learned"):
                work.copy_(src)
                e0, e1 = torch.cuda.Event(enable_timing=True), torch.cuda.Event(enable_timing=True)
                e0.record()
                st = s512.sort_learned(work) if name == "learned" else s512.sort_ex(work)
                e1.record()
                e1.synchronize()
                if rep:
                    times[name].append(e0.elapsed_time(e1))
                if ref is None:
                    ref = work.clone()
                elif not torch.equal(work.view(torch.int64), ref.view(torch.int64)):
                    raise RuntimeError(f"learned and splitter sorts differ on {kind}")
                if name == "learned":
                    info = {"fallback": int(st.learned_fallback), "levels_by_model": int(st.learned_levels),
                            "max_bucket_ratio": round(float(st.learned_max_bucket_ratio), 3)}
        sp, le = min(times["splitters"]), min(times["learned"])
        out[kind] = {"splitters_ms": sp, "learned_ms": le, "speedup": sp / le, **info}
    del src, work
    torch.cuda.empty_cache()
    return out


# ------------------------------------------------------------------ secondary workloads
def _timed(run, restore, steps, warmup):
    """Device ms of each timed run() (HIP events on torch's current stream, on
    which the library also runs); restore() precedes every run, untimed."""
    import torch
    ms = []
    for i in range(warmup + steps):
        restore()
        e0, e1 = torch.cuda.Event(enable_timing=True), torch.cuda.Event(enable_timing=True)
        e0.record()
        run()
        e1.record()
        e1.synchronize()
        if i >= warmup:
            ms.append(e0.elapsed_time(e1))
    return ms


def _line(metric, unit, items, ms, steps, warmup, dtype, data, config, extra):
    value = items / (sum(ms) / len(ms) / 1e3)
    out = {"metric": metric, "value": value, "unit": unit, "n_gpus": 1, "steps": steps, "warmup": warmup,
           "ms_per_step": sum(ms) / len(ms), "higher_is_better": True, "scaling": "weak", "vs_baseline": None,
           "dtype": dtype, "data": data, "config": config}
    out.update(extra)
    return out


def bench_kv_zipf(args):
    """Sort512kv::Sort layout: 536,870,912 (int32 key, int32 value) pairs,
    keys Zipf s=1.1 over 2^24 ranks through a seeded permutation, value =
    original index.  Checked: keys non-decreasing, the pair multiset kept."""
    import torch
    import s512
    n = args.n if args.n != 268_435_456 else 536_870_912
    dev = torch.device("cuda:0")
    k0 = torch.empty(n, dtype=torch.int32, device=dev)
    s512.fill(k0, "zipf", seed=3, param=1 << 24)
    v0 = torch.arange(n, dtype=torch.int32, device=dev)
    k, v = torch.empty_like(k0), torch.empty_like(v0)
    pair = lambda a, b: a.to(torch.int64) * (1 << 32) + b.to(torch.int64)  # noqa: E731
    h0 = multiset_hash(pair(k0, v0))

    def restore():
        k.copy_(k0)
        v.copy_(v0)
    ms = _timed(lambda: s512.kv_sort_ex(k, v), restore, args.steps, args.warmup)
    ok = is_sorted(k) and multiset_hash(pair(k, v)) == h0
    distinct = int(torch.unique(k0).numel())
    return _line("sorted key/value pairs/s, 512M int32 pairs, Zipf keys", "pairs/s", n, ms, args.steps,
                 args.warmup, "int32", "synthetic (device-generated Zipf s=1.1 via rejection-inversion)",
                 {"workload": f"kv sort, {n} (int32 key, int32 value) pairs, keys Zipf s=1.1 over 2^24 ranks "
                              f"({distinct} distinct), value = index"},
                 {"verified": ok, "effective_hbm_frac": (16.0 * n / (sum(ms) / len(ms) / 1e3) / 1e9) / HBM_PEAK_GBS})


def _seg_hash(t, off):
    import torch
    x = t.view(torch.int64) if t.dtype == torch.float64 else t.to(torch.int64)
    x = x * -7046029254386353131
    x = x ^ (x >> 29)
    c = torch.cat([torch.zeros(1, dtype=torch.int64, device=t.device), torch.cumsum(x, 0)])
    return c[off[1:]] - c[off[:-1]]


def _seg_sorted(t, off):
    import torch
    cont = torch.ones(t.numel(), dtype=torch.bool, device=t.device)
    starts = off[:-1]
    cont[starts[starts < t.numel()]] = False  # a segment's first key has no predecessor in it
    return bool(torch.all((t[1:] >= t[:-1]) | ~cont[1:]).item())


def bench_segsort(args, dtype):
    """200,000 independent arrays in one captured launch (segmented sort plan),
    sizes lognormal (median 512, sigma 1.5) clipped to [1, 2^20]; int32
    uniform or fp64 normal keys.  Checked per segment: non-decreasing, same
    multiset hash."""
    import numpy as np
    import torch
    import s512
    dev = torch.device("cuda:0")
    off = s512.lognormal_offsets(200_000, seed=0)
    n = int(off[-1])
    toff = torch.from_numpy(off.view(np.int64)).to(dev)
    tdt = torch.int32 if dtype == "int32" else torch.float64
    k0 = torch.empty(n, dtype=tdt, device=dev)
    s512.fill(k0, "uniform" if dtype == "int32" else "normal", seed=1)
    k = torch.empty_like(k0)
    h0 = _seg_hash(k0, toff)
    plan = s512.SegsortPlan(k, toff)
    ms = _timed(plan.launch, lambda: k.copy_(k0), args.steps, args.warmup)
    ok = _seg_sorted(k, toff) and bool(torch.equal(_seg_hash(k, toff), h0))
    plan.close()
    kb = 4 if dtype == "int32" else 8
    return _line(f"sorted keys/s, segmented batch of 200,000 {dtype} arrays", "keys/s", n, ms, args.steps,
                 args.warmup, dtype, "synthetic (device-generated; sizes lognormal, numpy PCG64 seed 0)",
                 {"workload": f"segmented sort plan (hipGraph), 200000 segments, {n} {dtype} keys, sizes "
                              f"lognormal(median 512, sigma 1.5) clipped to [1, 2^20]"},
                 {"verified": ok, "effective_hbm_frac": (2.0 * kb * n / (sum(ms) / len(ms) / 1e3) / 1e9) / HBM_PEAK_GBS})


def bench_learned(args):
    """1,073,741,824 fp64 keys: the Gaussian mixture + lognormal tail and the
    piecewise-constant adversarial input, each sorted by the learned
    partitioner and by sampled splitters (same input, both verified)."""
    import torch
    import s512
    n = args.n if args.n != 268_435_456 else 1 << 30
    dev = torch.device("cuda:0")
    src = torch.empty(n, dtype=torch.float64, device=dev)
    work = torch.empty_like(src)
    res, ok = {}, True
    for kind in ("mixture", "piecewise"):
        s512.fill(src, kind, seed=7)
        h0 = multiset_hash(src)
        r = {}
        for name in ("splitters", "learned"):
            st = {}

            def run():
                st["s"] = s512.sort_learned(work) if name == "learned" else s512.sort_ex(work)
            ms = _timed(run, lambda: work.copy_(src), args.steps, args.warmup)
            ok = ok and is_sorted(work) and multiset_hash(work) == h0
            r[name] = {"ms": sum(ms) / len(ms), "keys_per_s": n / (sum(ms) / len(ms) / 1e3)}
            if name == "learned":
                d = st["s"].as_dict()
                r[name].update({"classifier": d["classifier"], "fallback": d["learned_fallback"],
                                "max_bucket_ratio": round(d["learned_max_bucket_ratio"], 3),
                                "model_loss": d["model_loss"]})
        r["learned_vs_splitters"] = r["splitters"]["ms"] / r["learned"]["ms"]
        res[kind] = r
    ms = [res["mixture"]["learned"]["ms"]]
    return _line("sorted keys/s, 1G fp64, learned partitioner", "keys/s", n, ms, args.steps, args.warmup, "f64",
                 "synthetic (device-generated mixture / piecewise)",
                 {"workload": f"learned-partition sort, {n} fp64 keys, 8-component Gaussian mixture + "
                              f"lognormal(0,2) tail; piecewise-constant adversarial input alongside"},
                 {"verified": ok, "learned": res})


def bench_plain_i32(args):
    """The plain sort on 268,435,456 int32 keys: full-range uniform, 1000
    distinct values, sorted, reversed.  Roofline of the dominant kernel as in
    the fp64 headline (8 B per key read + written per scatter launch)."""
    import torch
    import s512
    n = args.n
    dev = torch.device("cuda:0")
    dists = ["uniform", "distinct", "sorted", "reversed"]
    work = torch.empty(n, dtype=torch.int32, device=dev)
    src = torch.empty_like(work)
    per, ok = {}, True
    ph_ms, ph_l, ph_k = {}, {}, {}
    for i, d in enumerate(dists):
        s512.fill(src, d, seed=i, param=1000)
        h0 = multiset_hash(src)
        st = {}

        def run():
            st["s"] = s512.sort_ex(work, profile=True)
        ms = _timed(run, lambda: work.copy_(src), args.steps, args.warmup)
        ok = ok and is_sorted(work) and multiset_hash(work) == h0
        per[d] = sum(ms) / len(ms)
        for j, name in enumerate(s512.PHASES):
            ph_ms[name] = ph_ms.get(name, 0.0) + st["s"].kernel_ms[j]
            ph_l[name] = ph_l.get(name, 0) + st["s"].launches[j]
            ph_k[name] = ph_k.get(name, 0) + st["s"].keys[j]
    bpk = {"count": 4, "scatter": 8, "tile": 8, "copy": 8}
    dom = max((p for p in bpk if ph_l.get(p)), key=lambda p: ph_ms[p])
    per_launch_bytes = ph_k[dom] * bpk[dom] / ph_l[dom]
    per_launch_ms = ph_ms[dom] / ph_l[dom]
    achieved = per_launch_bytes / (per_launch_ms / 1e3) / 1e9
    total = sum(per.values())
    return _line("sorted keys/s + % HBM peak, 256M int32", "keys/s", n * len(dists), [total], args.steps,
                 args.warmup, "int32", "synthetic (device-generated SplitMix64; four distributions)",
                 {"workload": f"plain sort, {n} int32 keys x {len(dists)} distributions ({','.join(dists)})"},
                 {"verified": ok, "per_distribution_ms": per,
                  "roofline": {"bound": "hbm", "kernel": KERNEL_OF_PHASE[dom], "achieved": achieved,
                               "peak": HBM_PEAK_GBS, "unit": "GB/s", "frac": achieved / HBM_PEAK_GBS,
                               "traffic": None, "bytes_per_launch": per_launch_bytes,
                               "avg_launch_ms": per_launch_ms}})


def bench_host(args):
    """Sort512::Sort on a host array (the reference's own calling form):
    268,435,456 fp64 keys in ordinary pageable host memory, end to end (host
    to device, sort, device to host, through the pipelined staging), against
    the PCIe bound measured in the same run: a pinned-memory copy of the
    array to the device and back."""
    import numpy as np
    import torch
    import s512
    n = args.n
    a0 = np.random.default_rng(0).random(n)
    pin = torch.empty(n, dtype=torch.float64, pin_memory=True)
    dev = torch.empty(n, dtype=torch.float64, device="cuda:0")
    bound = []
    for _ in range(3):
        torch.cuda.synchronize()
        t0 = time.perf_counter()
        dev.copy_(pin)
        pin.copy_(dev)
        torch.cuda.synchronize()
        bound.append(time.perf_counter() - t0)
    del pin, dev
    torch.cuda.empty_cache()
    ms, ok = [], True
    for i in range(args.warmup + args.steps):
        a = a0.copy()
        t0 = time.perf_counter()
        s512.sort(a)
        dt = time.perf_counter() - t0
        if i >= args.warmup:
            ms.append(dt * 1e3)
        ok = ok and bool(np.all(a[1:] >= a[:-1]))
    pcie_ms = min(bound) * 1e3
    t = sum(ms) / len(ms)
    return _line("sorted keys/s end to end from host memory, 256M fp64", "keys/s", n, [t], args.steps, args.warmup,
                 "f64", "synthetic (numpy PCG64 uniform, pageable host memory)",
                 {"workload": f"host-pointer sort, {n} fp64 keys, pageable host array"},
                 {"verified": ok, "pcie_bound_ms": pcie_ms, "ratio_to_pcie_bound": t / pcie_ms})


def bench_wide(args):
    """268,435,456 int64 keys (full range) sorted, and argsort of 268,435,456
    fp64 keys (standard normal): int32 positions of the ascending order."""
    import torch
    import s512
    n = args.n
    dev = torch.device("cuda:0")
    g = torch.Generator(device=dev)
    g.manual_seed(5)
    k64 = torch.randint(-(1 << 63), (1 << 63) - 1, (n,), dtype=torch.int64, device=dev, generator=g)
    w = torch.empty_like(k64)
    h0 = multiset_hash(k64)
    ms_i64 = _timed(lambda: s512.sort(w), lambda: w.copy_(k64), args.steps, args.warmup)
    ok = is_sorted(w) and multiset_hash(w) == h0
    del k64, w
    f = torch.empty(n, dtype=torch.float64, device=dev)
    s512.fill(f, "normal", seed=1)
    idx = torch.empty(n, dtype=torch.int32, device=dev)
    ms_arg = _timed(lambda: s512.argsort(f, out=idx), lambda: None, args.steps, args.warmup)
    ok = ok and is_sorted(f[idx.long()])
    t_i64, t_arg = sum(ms_i64) / len(ms_i64), sum(ms_arg) / len(ms_arg)
    return _line("sorted keys/s, 256M int64 keys", "keys/s", n, ms_i64, args.steps, args.warmup, "int64",
                 "synthetic (torch.randint full int64 range; fp64 normal for argsort)",
                 {"workload": f"int64 key sort and fp64 argsort, {n} keys each"},
                 {"verified": ok, "int64_sort_ms": t_i64, "argsort_f64_ms": t_arg,
                  "argsort_f64_keys_per_s": n / (t_arg / 1e3)})


def bench_topk(args):
    """Top-k and selection on 268,435,456 fp64 keys (standard normal): the
    1,048,576 smallest keys in order, and the median; checked against the
    full sort's prefix."""
    import torch
    import s512
    n, k = args.n, 1 << 20
    dev = torch.device("cuda:0")
    a = torch.empty(n, dtype=torch.float64, device=dev)
    s512.fill(a, "normal", seed=1)
    out = torch.empty(k, dtype=torch.float64, device=dev)
    ms_topk = _timed(lambda: s512.topk(a, k, out=out), lambda: None, args.steps, args.warmup)
    med = {}
    ms_sel = _timed(lambda: med.update(v=s512.select(a, n // 2)), lambda: None, args.steps, args.warmup)
    full = a.clone()
    s512.sort(full)
    ok = bool(torch.equal(out.view(torch.int64), full[:k].view(torch.int64))) and med["v"] == float(full[n // 2])
    t = sum(ms_topk) / len(ms_topk)
    return _line("keys/s scanned, top-1M of 256M fp64", "keys/s", n, ms_topk, args.steps, args.warmup, "f64",
                 "synthetic (device-generated normal)",
                 {"workload": f"top-k, {n} fp64 keys, k = {k}; selection of the median"},
                 {"verified": ok, "topk_ms": t, "select_median_ms": sum(ms_sel) / len(ms_sel)})


WORKLOADS = {"plain": None, "plain-i32": bench_plain_i32, "host": bench_host, "wide": bench_wide,
             "topk": bench_topk, "kv-zipf": bench_kv_zipf, "segsort-i32": lambda a: bench_segsort(a, "int32"),
             "segsort-f64": lambda a: bench_segsort(a, "float64"), "learned-1g": bench_learned}


def main():
    args = parse()
    if args.gpus != 1:
        print(f"bench.py measures one GPU: every box in this setup has one GPU and the sort runs on one device "
              f"(asked for --gpus {args.gpus}).", file=sys.stderr)
        sys.exit(2)
    if args.workload != "plain":
        out = WORKLOADS[args.workload](args)
        print(json.dumps(out))
        sys.exit(0 if out.get("verified") else 1)
    # counter runs first, before this process touches the GPU
    traffic_runs = None
    if not args.no_traffic:
        try:
            traffic_runs = {"uniform_scatter": measure_traffic(args.n, KERNEL_OF_PHASE["scatter"])}
        except Exception as e:  # reported, not hidden
            traffic_runs = {"error": repr(e)}
    import torch
    import s512
    dev = torch.device("cuda:0")
    n = args.n
    dists = args.dists.split(",")
    pristine = {}
    for i, d in enumerate(dists):
        t = torch.empty(n, dtype=torch.float64, device=dev)
        s512.fill(t, d, seed=i, param=1000)
        pristine[d] = t
    work = torch.empty(n, dtype=torch.float64, device=dev)
    torch.cuda.synchronize()
    hashes = {d: multiset_hash(pristine[d]) for d in dists}

    ev0 = torch.cuda.Event(enable_timing=True)
    ev1 = torch.cuda.Event(enable_timing=True)
    per_dist_ms = {d: [] for d in dists}
    phase_ms, phase_launch, phase_keys = {}, {}, {}
    levels, checked = {}, {}
    torch.cuda.synchronize()
    wall0 = None
    for step in range(args.warmup + args.steps):
        timed = step >= args.warmup
        if timed and wall0 is None:
            torch.cuda.synchronize()
            wall0 = time.perf_counter()
        for d in dists:
            work.copy_(pristine[d])
            ev0.record()
            st = s512.sort_ex(work, profile=True)
            ev1.record()
            ev1.synchronize()
            if not timed:
                continue
            per_dist_ms[d].append(ev0.elapsed_time(ev1))
            levels[d] = st.levels
            for i, name in enumerate(s512.PHASES):
                phase_ms[name] = phase_ms.get(name, 0.0) + st.kernel_ms[i]
                phase_launch[name] = phase_launch.get(name, 0) + st.launches[i]
                phase_keys[name] = phase_keys.get(name, 0) + st.keys[i]
            if d not in checked:
                checked[d] = is_sorted(work) and multiset_hash(work) == hashes[d]
    torch.cuda.synchronize()
    wall = time.perf_counter() - wall0
    total_ms = sum(sum(v) for v in per_dist_ms.values())
    keys = n * len(dists) * args.steps
    value = keys / (total_ms / 1e3)
    key_bytes = 8
    bpk = {"count": key_bytes, "scatter": 2 * key_bytes, "tile": 2 * key_bytes, "copy": 2 * key_bytes}
    dom = max((p for p in bpk if phase_launch.get(p)), key=lambda p: phase_ms[p])
    per_launch_bytes = phase_keys[dom] * bpk[dom] / phase_launch[dom]
    per_launch_ms = phase_ms[dom] / phase_launch[dom]
    achieved = per_launch_bytes / (per_launch_ms / 1e3) / 1e9
    traffic, traffic_detail = None, None
    if traffic_runs and "uniform_scatter" in traffic_runs and dom == "scatter":
        traffic, traffic_detail = traffic_runs["uniform_scatter"]
    elif traffic_runs and "error" in traffic_runs:
        traffic_detail = traffic_runs["error"]
    for d in dists:  # free the inputs before the learned comparison
        pristine[d] = None
    del work
    torch.cuda.empty_cache()
    out = {
        "metric": "sorted keys/s + % HBM peak, 256M fp64; learned vs sampled splitters",
        "value": value,
        "unit": "keys/s",
        "n_gpus": 1,
        "steps": args.steps,
        "warmup": args.warmup,
        "ms_per_step": total_ms / args.steps,
        "higher_is_better": True,
        "scaling": "weak",
        "vs_baseline": None,
        "dtype": "f64",
        "data": "synthetic (device-generated SplitMix64 / Box-Muller; five distributions)",
        "config": {"workload": f"plain sort, {n} fp64 keys x {len(dists)} distributions ({','.join(dists)})",
                   "keys_per_step": n * len(dists), "classifier": "sampled splitters"},
        "roofline": {"bound": "hbm", "kernel": KERNEL_OF_PHASE[dom], "achieved": achieved, "peak": HBM_PEAK_GBS,
                     "unit": "GB/s", "frac": achieved / HBM_PEAK_GBS, "traffic": traffic,
                     "measured_frac": (traffic / (per_launch_ms / 1e3) / 1e9 / HBM_PEAK_GBS) if traffic else None,
                     "traffic_detail": traffic_detail, "bytes_per_launch": per_launch_bytes,
                     "avg_launch_ms": per_launch_ms},
        "effective_hbm_frac": (16.0 * keys / (total_ms / 1e3) / 1e9) / HBM_PEAK_GBS,
        "per_distribution_ms": {d: sum(v) / len(v) for d, v in per_dist_ms.items()},
        "per_distribution_keys_per_s": {d: n / (sum(v) / len(v) / 1e3) for d, v in per_dist_ms.items()},
        "levels": levels,
        "phase_ms_per_step": {p: phase_ms[p] / args.steps for p in phase_ms if phase_launch.get(p)},
        "verified": checked,
        "wall_s_timed_region": wall,
        "learned": None,
        "cpu_baseline": None,
    }
    if not args.no_learned:
        try:
            out["learned"] = learned_comparison(s512, n)
        except Exception as e:  # reported, not hidden
            out["learned"] = {"error": repr(e)}
    if not all(checked.values()):
        out["error"] = "verification failed"
    if not args.no_cpu_baseline:
        try:
            out["cpu_baseline"] = cpu_baseline(args.cpu_sample)
        except Exception as e:  # reported, not hidden
            out["cpu_baseline"] = {"error": repr(e)}
    print(json.dumps(out))
    sys.exit(0 if all(checked.values()) else 1)


if __name__ == "__main__":
    main()
