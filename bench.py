#!/usr/bin/env python3
"""bench.py -- sorted keys/s and % of HBM peak for 268,435,456 fp64 keys.

Workload: the plain sort path (Sort512::Sort / SortOmp, sort512.hpp:5960,
5993) on 268,435,456 fp64 keys in device memory, five synthetic
distributions generated on the GPU: uniform [0,1), standard normal, 1000
distinct values, sorted ascending, sorted descending.  One step sorts each of
the five arrays once (5 x 2^28 keys); the unsorted input is restored from a
pristine copy before every sort, outside the timed events.

Output: one JSON line.
  value          keys per step / the median step time (HIP events on the
                 stream the library runs on, around the production calls only;
                 median of --steps >= 5 steps after --warmup steps)
  roofline       the kernel phase with the largest share of device time,
                 timed per launch with HIP events inside the library
                 (s512_options.profile) in a second pass of the same steps;
                 algorithmic bytes = one read and one write of every key (16 B
                 per fp64 key) per scatter / tile / copy launch, one read (8 B)
                 per count launch
  traffic        HBM bytes per launch of that kernel, from two rocprofv3 --pmc
                 child runs of this script (FETCH_SIZE, then WRITE_SIZE), each
                 counter scaled by what it reads for a 2 GiB copy of known size
                 in the same run; measured_frac = traffic / launch time / peak
  whole_sort     FETCH_SIZE + WRITE_SIZE summed over every library kernel of
                 one uniform sort (same counter runs), over that sort's median
                 production time and the peak: the sort's measured HBM
                 fraction; algorithmic_frac = 16 B per key over the same time
  int32_256M     the plain sort on 268,435,456 int32 keys (uniform, 1000
                 distinct, sorted, reversed), with its own roofline and
                 whole-sort figures
  entry_points   Partition512 (one pivot), the SortOmpMerge-family merge
                 sort, SmallSort16V on a batch of 128-key arrays, Sort512kv
                 on 536,870,912 Zipf pairs, the segmented batches and the
                 learned partitioner on 2^30 keys, each timed as above with
                 its measured HBM bytes where the counter runs cover it
  learned        the learned partitioner against sampled splitters on the
                 skewed mixture and on uniform keys (same input, both exact)
  cpu_baseline   the reference's SortOmpPartition, compiled from the reference
                 headers into oracle/_ref when they were available at build
                 time, on this host's physical cores (OMP_PROC_BIND=spread,
                 OMP_PLACES=cores) in child processes, on a bounded sample;
                 beside it int32, 2^30 keys, the per-segment loop of config 4,
                 the other reference entry points and the inputs that make its
                 quicksort degrade (per-call time limit, DNF, power-law fit
                 over 2^18..2^24 keys); if that library or AVX-512 is absent,
                 the oracle's C restatement (one core) is timed instead

Every sorted array is verified at full size by size-independent properties:
non-decreasing order and an order-independent multiset hash equal to the
input's.
"""
import argparse
import csv
import glob
import json
import math
import os
import shutil
import statistics
import subprocess
import sys
import tempfile
import time

ROOT = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, os.path.join(ROOT, "avx-512-sort_amd"))

HBM_PEAK_GBS = 8000.0
DISTS = ["uniform", "normal", "distinct", "sorted", "reversed"]
KERNEL_OF_PHASE = {"count": "k_count", "scatter": "k_scatter", "tile": "k_tile_sort", "copy": "k_copy_back"}
CALIB_BYTES = 1 << 31  # the counter runs' calibration copy


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
    ap.add_argument("--no-extras", action="store_true", help="headline only: no int32 line, no other entry points")
    ap.add_argument("--cpu-sample", type=int, default=1 << 28)  # the full 268,435,456-key uniform input
    ap.add_argument("--cpu-budget", type=float, default=150.0,
                    help="seconds of CPU-baseline work beyond the headline sample (items past it are skipped)")
    ap.add_argument("--option", action="append", default=[], metavar="NAME=VALUE",
                    help="s512_options field for the headline sorts (repeatable), e.g. scatter_variant=11")
    ap.add_argument("--workload", default="plain", choices=list(WORKLOADS),
                    help="plain: the 256M fp64 headline (default); the others time one secondary workload "
                         "and print its own JSON line")
    # child modes (this script started by itself)
    ap.add_argument("--pmc-child", choices=["f64", "i32"], help=argparse.SUPPRESS)
    ap.add_argument("--cpu-child", help=argparse.SUPPRESS)
    ap.add_argument("--gen-child", help=argparse.SUPPRESS)
    return ap.parse_args()


def median(xs):
    return statistics.median(xs) if xs else None


_T0 = time.time()


def progress(msg):
    """One progress line on stderr (stdout carries only the JSON line)."""
    print(f"[bench {time.time() - _T0:6.1f}s] {msg}", file=sys.stderr, flush=True)


HASH_SLICE = 1 << 28  # keys per slice of the checks (bounds their temporaries)


def multiset_hash(t):
    """Order-independent hash of the bit patterns (sum of a 64-bit mix)."""
    import torch
    h = 0
    for i in range(0, t.numel(), HASH_SLICE):
        s = t[i:i + HASH_SLICE]
        x = s.view(torch.int64) if s.dtype == torch.float64 else s.to(torch.int64)
        x = x * -7046029254386353131  # 0x9E3779B97F4A7C15 as int64 (wraps)
        x = x ^ (x >> 29)
        x = x * -4658895280553007687  # 0xBF58476D1CE4E5B9
        x = x ^ (x >> 32)
        h += int(x.sum().item())
    return ((h + (1 << 63)) % (1 << 64)) - (1 << 63)  # the sum modulo 2^64, as one slice would wrap it


def is_sorted(t):
    import torch
    n = t.numel()
    for i in range(0, n - 1, HASH_SLICE):
        s = t[i:min(n, i + HASH_SLICE + 1)]  # slices overlap by one key
        if not bool(torch.all(s[1:] >= s[:-1]).item()):
            return False
    return True


# ------------------------------------------------------------------ host CPU
def host_cpu():
    """CPU model, the logical CPUs this process may run on, their physical
    cores, the cgroup CPU quota, and the OpenMP thread count to use: one per
    physical core in the affinity mask, capped by the quota and by an
    OMP_NUM_THREADS the environment already sets."""
    info = {}
    try:
        with open("/proc/cpuinfo") as f:
            block = f.read().split("\n\n")[0]
        for line in block.splitlines():
            k, _, v = line.partition(":")
            k, v = k.strip(), v.strip()
            if k in ("model name", "vendor_id", "cpu family", "model", "stepping", "cpu MHz"):
                info[k.replace(" ", "_")] = v
            elif k == "flags":
                info["avx512"] = sorted(f for f in v.split() if f.startswith("avx512"))
    except OSError:
        pass
    aff = sorted(os.sched_getaffinity(0))
    cores = set()
    for c in aff:
        try:
            with open(f"/sys/devices/system/cpu/cpu{c}/topology/core_id") as f:
                core = f.read().strip()
            with open(f"/sys/devices/system/cpu/cpu{c}/topology/physical_package_id") as f:
                pkg = f.read().strip()
            cores.add((pkg, core))
        except OSError:
            cores.add(("?", str(c)))
    quota = None
    try:
        with open("/sys/fs/cgroup/cpu.max") as f:
            q, p = f.read().split()
        if q != "max":
            quota = int(q) / int(p)
    except (OSError, ValueError):
        pass
    threads = len(cores)
    if quota:
        threads = min(threads, max(1, int(quota)))
    env_n = os.environ.get("OMP_NUM_THREADS")
    if env_n and env_n.isdigit():
        threads = min(threads, int(env_n))
    info.update({"logical_cpus_in_affinity": len(aff), "physical_cores_in_affinity": len(cores),
                 "cgroup_cpu_quota": quota, "omp_num_threads_env": env_n, "threads_used": threads,
                 "omp_proc_bind": "spread", "omp_places": "cores"})
    return info, threads


def _cpu_env(threads):
    env = dict(os.environ)
    env.update({"OMP_NUM_THREADS": str(threads), "OMP_PROC_BIND": "spread", "OMP_PLACES": "cores"})
    return env


def cpu_child(spec):
    """One CPU timing in a child process (started with the OpenMP binding
    variables set): builds the input, times the reference call with a
    monotonic clock around the call only (best and median of up to `reps`
    runs within `budget_s`), checks the output, prints one JSON line."""
    import ctypes
    import numpy as np
    sys.path.insert(0, os.path.join(ROOT, "oracle"))
    import oracle
    n = spec["n"]
    src = spec["input"]
    vals = None
    off = None
    if src.startswith("f64:"):
        a0 = oracle.gen_f64(src[4:], n, spec.get("seed", 0))
    elif src == "i32:uniform":
        a0 = oracle.gen_i32_full(n, spec.get("seed", 0))
    elif src.startswith("file:"):
        a0 = np.array(np.load(src[5:], mmap_mode="r")[:n])
    elif src.startswith("seg:"):
        off = np.load(spec["offsets"])
        n = int(off[-1])
        a0 = oracle.gen_i32_full(n, 6) if src == "seg:i32" else oracle.gen_f64("normal", n, 6)
    else:
        raise ValueError(src)
    if spec.get("kv"):
        vals = np.arange(n, dtype=np.int32)
    ref = oracle.ref_lib()
    use_ref = ref is not None and oracle.cpu_has_avx512()
    fn = spec["fn"]
    if not use_ref:
        if fn != "sort_omp_partition" or src not in ("f64:uniform",):
            print(json.dumps({"skipped": "reference library or AVX-512 absent"}))
            return
        n = min(n, 1 << 23)
        a0 = a0[:n].copy()
    suffix = "f64" if a0.dtype == np.float64 else "i32"
    times, ok = [], True
    t_start = time.time()
    for rep in range(spec.get("reps", 1)):
        if rep and time.time() - t_start > spec.get("budget_s", 20):
            break
        a = a0.copy()
        v = vals.copy() if vals is not None else None
        pa = ctypes.c_void_p(a.ctypes.data)
        t0 = time.perf_counter()
        if not use_ref:
            oracle.sort(a)
        elif fn == "kv_sort_omp_partition":
            ref.ref_kv_sort_omp_partition(pa, ctypes.c_void_p(v.ctypes.data), n)
        elif fn == "segsort_loop":
            getattr(ref, f"ref_segsort_loop_{suffix}")(pa, ctypes.c_void_p(off.ctypes.data), len(off) - 1,
                                                      int(spec.get("parallel", 0)))
        else:
            getattr(ref, f"ref_{fn}_{suffix}")(pa, n)
        times.append(time.perf_counter() - t0)
        if off is not None:
            cont = np.ones(n, dtype=bool)  # False where a segment starts
            cont[off[:-1][off[:-1] < n].astype(np.int64)] = False
            ok = ok and bool(np.all((a[1:] >= a[:-1]) | ~cont[1:]))
        else:
            ok = ok and bool(np.all(a[1:] >= a[:-1]))
    kind = "reference" if use_ref else "port"
    print(json.dumps({"n": n, "times_s": times, "best_s": min(times), "median_s": median(times), "ok": ok,
                      "kind": kind, "threads": int(ref.ref_max_threads()) if use_ref else 1}))


def gen_child(outdir):
    """Inputs only the GPU generator makes (Zipf keys of config 3, the
    piecewise-constant keys of config 5b), written as .npy files for the CPU
    child runs: the same seeded inputs the GPU workloads sort."""
    import numpy as np
    import torch
    import s512
    dev = torch.device("cuda:0")
    t = torch.empty(1 << 29, dtype=torch.int32, device=dev)
    s512.fill(t, "zipf", seed=3, param=1 << 24)
    np.save(os.path.join(outdir, "zipf.npy"), t.cpu().numpy())
    del t
    p = torch.empty(1 << 24, dtype=torch.float64, device=dev)
    s512.fill(p, "piecewise", seed=8)
    np.save(os.path.join(outdir, "piecewise.npy"), p.cpu().numpy())
    print("ok")


def _run_cpu(spec, env, timeout):
    cmd = [sys.executable, os.path.abspath(__file__), "--cpu-child", json.dumps(spec)]
    t0 = time.time()
    try:
        p = subprocess.run(cmd, env=env, timeout=timeout, capture_output=True, text=True)
    except subprocess.TimeoutExpired:
        return {"dnf": True, "limit_s": timeout, "n": spec["n"]}
    if p.returncode != 0:
        return {"error": (p.stderr or "").strip().splitlines()[-1:] or f"exit {p.returncode}"}
    out = json.loads(p.stdout.strip().splitlines()[-1])
    out["wall_s"] = round(time.time() - t0, 2)
    return out


def _power_law(points):
    """Least-squares fit of t = c * n^p on (n, t) pairs (log-log)."""
    pts = [(math.log(n), math.log(t)) for n, t in points if t and t > 0]
    if len(pts) < 2:
        return None
    mx = sum(x for x, _ in pts) / len(pts)
    my = sum(y for _, y in pts) / len(pts)
    sxx = sum((x - mx) ** 2 for x, _ in pts)
    if sxx == 0:
        return None
    p = sum((x - mx) * (y - my) for x, y in pts) / sxx
    return p, math.exp(my - p * mx)


def cpu_baseline(sample_n, budget_s):
    """Reference SortOmpPartition on this host (the reported baseline) plus
    the other CPU figures, each in a child process on the physical cores."""
    import numpy as np
    info, threads = host_cpu()
    env = _cpu_env(threads)
    progress(f"cpu: SortOmpPartition on {sample_n} fp64 keys, {threads} threads")
    head = _run_cpu({"fn": "sort_omp_partition", "input": "f64:uniform", "n": sample_n, "reps": 3,
                     "budget_s": 20}, env, 300)
    if "best_s" not in head:
        return {"error": head, "cpu": info}
    out = {"value": head["n"] / head["best_s"], "unit": "keys/s", "cores": head["threads"], "kind": head["kind"],
           "sample": (f"Sort512::SortOmpPartition, {head['n']} fp64 keys uniform [0,1) (SplitMix64 seed 0), "
                      f"best of {len(head['times_s'])} (median {head['median_s']:.3f} s)")
           if head["kind"] == "reference" else
           f"oracle restatement of Sort512::Sort, {head['n']} fp64 keys uniform, one core",
           "cpu": info, "verified": head["ok"]}
    if head["kind"] != "reference":
        return out
    t0 = time.time()
    extra = {}

    def left():
        return budget_s - (time.time() - t0)

    def item(name, spec, limit):
        if left() <= 0:
            extra[name] = {"skipped": "CPU budget of this run spent"}
            return None
        progress(f"cpu: {name}")
        r = _run_cpu(spec, env, min(limit, max(5.0, left())))
        if "best_s" in r:
            r["keys_per_s"] = r["n"] / r["best_s"]
        extra[name] = r
        return r

    item("int32 256M SortOmpPartition", {"fn": "sort_omp_partition", "input": "i32:uniform", "n": 1 << 28,
                                          "reps": 3, "budget_s": 10}, 120)
    item("2^30 fp64 SortOmpPartition (config 5 size, uniform)",
         {"fn": "sort_omp_partition", "input": "f64:uniform", "n": 1 << 30, "reps": 1}, 120)
    tmp = tempfile.mkdtemp(prefix="s512_cpu_", dir="/dev/shm" if os.path.isdir("/dev/shm") else None)
    try:
        import s512  # numpy-only helpers; the GPU is not touched here
        off = s512.lognormal_offsets(200_000, seed=0).astype(np.uint64)
        offp = os.path.join(tmp, "off.npy")
        np.save(offp, off)
        for dt in ("i32", "f64"):
            for par in (0, 1):
                item(f"config 4 Sort per segment, {dt}, {'omp dynamic' if par else '1 thread'}",
                     {"fn": "segsort_loop", "input": f"seg:{dt}", "offsets": offp, "n": int(off[-1]),
                      "parallel": par, "reps": 1}, 90)
        pow2 = threads & (threads - 1) == 0 and threads <= 128  # sort512.hpp:6006: 2^k threads
        for fn, label in (("sort", "Sort512::Sort (1 core), 2^23 keys"),
                          ("sort_omp_merge", "Sort512::SortOmpMerge"), ("sort_omp_mergedeps", "Sort512::SortOmpMergeDeps"),
                          ("sort_omp_parmerge", "Sort512::SortOmpParMerge")):
            if fn != "sort" and not pow2:
                extra[label] = {"skipped": f"needs a power-of-two thread count <= 128, have {threads}"}
                continue
            e = env if fn != "sort" else _cpu_env(1)
            n = (1 << 23) if fn == "sort" else sample_n
            if left() > 0:
                progress(f"cpu: {label}")
                r = _run_cpu({"fn": fn, "input": "f64:uniform", "n": n, "reps": 1}, e, 60)
                if "best_s" in r:
                    r["keys_per_s"] = r["n"] / r["best_s"]
                extra[label] = r
        # inputs that make the reference's quicksort degrade: the full size
        # under a time limit, then 2^18..2^24 keys for a power-law fit
        gen = subprocess.run([sys.executable, os.path.abspath(__file__), "--gen-child", tmp], capture_output=True,
                             text=True, timeout=180)
        have_files = gen.returncode == 0
        patho = {}
        cases = [("2c 1000 distinct fp64", {"input": "f64:distinct", "seed": 2}, 1 << 28, "sort_omp_partition"),
                 ("2d sorted fp64", {"input": "f64:sorted"}, 1 << 28, "sort_omp_partition"),
                 ("2e reversed fp64", {"input": "f64:reversed"}, 1 << 28, "sort_omp_partition")]
        if have_files:
            cases += [("3 kv Zipf int32 pairs", {"input": "file:" + os.path.join(tmp, "zipf.npy"), "kv": True},
                       1 << 29, "kv_sort_omp_partition"),
                      ("5b piecewise fp64", {"input": "file:" + os.path.join(tmp, "piecewise.npy")}, None,
                       "sort_omp_partition")]
        else:
            patho["generator"] = {"error": (gen.stderr or "").strip().splitlines()[-1:]}
        for name, base, full, fn in cases:
            res = {"function": "Sort512kv::SortOmpPartition" if fn.startswith("kv") else "Sort512::SortOmpPartition"}
            pts = []
            for lg in (18, 20, 22, 24):
                if left() <= 0:
                    res[f"2^{lg}"] = {"skipped": "CPU budget of this run spent"}
                    continue
                progress(f"cpu: {name}, 2^{lg} keys")
                r = _run_cpu(dict(base, fn=fn, n=1 << lg, reps=1), env, 12)
                res[f"2^{lg}"] = r.get("best_s", r)
                if "best_s" in r:
                    pts.append((1 << lg, r["best_s"]))
                else:
                    break
            fit = _power_law(pts)
            if fit:
                res["power_law"] = {"exponent": round(fit[0], 3), "coef_s": fit[1]}
            if full:
                # the full size under a limit of 40 s of sorting (plus the
                # input's generation); not started when the fit predicts more
                limit = 40.0
                pred = fit[1] * full ** fit[0] if fit else None
                if pred is not None:
                    res["predicted_full_s"] = pred
                if pred is not None and pred > limit:
                    res[f"full {full}"] = {"dnf": True, "limit_s": limit, "not_started": "the fit predicts more"}
                elif left() > 0:
                    progress(f"cpu: {name}, {full} keys")
                    r = _run_cpu(dict(base, fn=fn, n=full, reps=1), env, limit + 30)
                    res[f"full {full}"] = ({"dnf": True, "limit_s": limit} if r.get("dnf")
                                           else r.get("best_s", r))
            patho[name] = res
        extra["degrading inputs"] = patho
    finally:
        shutil.rmtree(tmp, ignore_errors=True)
    extra["budget_s"] = budget_s
    extra["elapsed_s"] = round(time.time() - t0, 1)
    out["other_reference_functions"] = extra
    return out


# ------------------------------------------------------------------ traffic
def pmc_child(kind, n):
    """The counter runs' program: a calibration copy of 2 GiB, then each
    operation once, every operation preceded by a one-workgroup marker fill so
    that the parent can tell which dispatches belong to which operation."""
    import numpy as np
    import torch
    import s512
    dev = torch.device("cuda:0")
    marker = torch.empty(64, dtype=torch.float64, device=dev)

    def mark():
        s512.fill(marker, "uniform", seed=99)
    c0 = torch.empty(CALIB_BYTES // 8, dtype=torch.float64, device=dev)
    c1 = torch.empty_like(c0)
    c0.zero_()
    mark()
    c1.copy_(c0)  # group "calib"
    del c0, c1
    if kind == "f64":
        src = torch.empty(n, dtype=torch.float64, device=dev)
        s512.fill(src, "uniform", seed=0)
        work = torch.empty_like(src)
        work.copy_(src)
        mark()
        s512.sort_ex(work)  # group "sort"
        work.copy_(src)
        mark()
        s512.partition(work, 0, n - 1, 0.5)  # group "partition"
        work.copy_(src)
        mark()
        s512.sort_merge(work)  # group "merge"
        work.copy_(src)
        off = torch.from_numpy(np.arange(0, n + 1, 128, dtype=np.int64)).to(dev)
        mark()
        s512.segsort(work, off)  # group "smallsort"
    else:
        src = torch.empty(n, dtype=torch.int32, device=dev)
        s512.fill(src, "uniform", seed=0)
        work = torch.empty_like(src)
        work.copy_(src)
        mark()
        s512.sort_ex(work)  # group "sort"
        del src, work
        k = torch.empty(1 << 29, dtype=torch.int32, device=dev)
        s512.fill(k, "zipf", seed=3, param=1 << 24)
        v = torch.arange(1 << 29, dtype=torch.int32, device=dev)
        mark()
        s512.kv_sort_ex(k, v)  # group "kv"
    mark()
    torch.cuda.synchronize()


PMC_GROUPS = {"f64": ["calib", "sort", "partition", "merge", "smallsort"], "i32": ["calib", "sort", "kv"]}


def _pmc_run(counter, kind, n, tmp):
    """One rocprofv3 --pmc child run; returns its rows in dispatch order."""
    out = os.path.join(tmp, f"{counter}_{kind}")
    cmd = ["rocprofv3", "--pmc", counter, "--output-format", "csv", "-d", out, "-o", "p", "--",
           sys.executable, os.path.abspath(__file__), "--pmc-child", kind, "--n", str(n)]
    subprocess.run(cmd, check=True, timeout=300, stdout=subprocess.DEVNULL, stderr=subprocess.DEVNULL)
    files = glob.glob(os.path.join(out, "**", "*counter_collection.csv"), recursive=True)
    if not files:
        raise RuntimeError(f"rocprofv3 wrote no counter file for {counter}")
    with open(files[0]) as f:
        rows = list(csv.DictReader(f))
    rows.sort(key=lambda r: int(r.get("Dispatch_Id") or 0))
    return rows


def _groups(rows, names):
    """Split dispatches at the one-workgroup marker fills into named groups."""
    groups, cur = {}, None
    idx = -1
    for r in rows:
        name = r["Kernel_Name"]
        if "k_fill" in name and int(r.get("Grid_Size") or 0) <= 1024:
            idx += 1
            cur = names[idx] if idx < len(names) else None
            if cur:
                groups[cur] = []
            continue
        if cur:
            groups[cur].append(r)
    return groups


def measure_traffic(n_f64, n_i32):
    """Per operation: HBM bytes read and written by the library's kernels
    (summed over its dispatches) and per kernel name (mean per dispatch), each
    counter scaled by its reading for the 2 GiB calibration copy."""
    if shutil.which("rocprofv3") is None:
        return {"error": "rocprofv3 not found"}
    tmp = tempfile.mkdtemp(prefix="s512_pmc_")
    res = {}
    try:
        for kind, n in (("f64", n_f64), ("i32", n_i32)):
            per = {}
            for counter in ("FETCH_SIZE", "WRITE_SIZE"):
                progress(f"counter run {counter} ({kind})")
                g = _groups(_pmc_run(counter, kind, n, tmp), PMC_GROUPS[kind])
                calib = [float(r["Counter_Value"]) for r in g.get("calib", []) if "copyBuffer" in r["Kernel_Name"]]
                if not calib:
                    raise RuntimeError("no calibration copy in the counter run")
                scale = CALIB_BYTES / (max(calib) * 1024.0)
                for op, rows in g.items():
                    if op == "calib":
                        continue
                    # the library's kernels, without the input generators
                    lib_rows = [r for r in rows if "s512::" in r["Kernel_Name"] and "k_fill" not in r["Kernel_Name"]]
                    d = per.setdefault(op, {"kernels": {}})
                    d[counter] = sum(float(r["Counter_Value"]) for r in lib_rows) * 1024.0 * scale
                    d[counter + "_scale"] = scale
                    for r in lib_rows:
                        kn = r["Kernel_Name"].split("(")[0].replace("void ", "")
                        kd = d["kernels"].setdefault(kn, {})
                        kd.setdefault(counter, []).append(float(r["Counter_Value"]) * 1024.0 * scale)
            for op, d in per.items():
                d["bytes"] = d.get("FETCH_SIZE", 0.0) + d.get("WRITE_SIZE", 0.0)
                for kn, kd in d["kernels"].items():
                    f, w = kd.get("FETCH_SIZE", []), kd.get("WRITE_SIZE", [])
                    d["kernels"][kn] = {"dispatches": len(f), "fetch_per_dispatch": sum(f) / max(1, len(f)),
                                        "write_per_dispatch": sum(w) / max(1, len(w))}
                res[f"{kind}_{op}"] = d
        return res
    finally:
        shutil.rmtree(tmp, ignore_errors=True)


def _kernel_traffic(traffic, op, prefix):
    """Mean per-dispatch bytes of the op's kernels whose name (after the
    namespace) starts with prefix, e.g. "k_scatter<" or "k_tile_sort<double"
    (the leaf tiles; the sample sorts run on the encoded unsigned keys)."""
    d = (traffic or {}).get(op)
    if not d:
        return None
    f = w = 0.0
    cnt = 0
    names = set()
    for kn, kd in d["kernels"].items():
        short = kn.split("s512::")[-1]
        # a kernel of a gated pair that returned at once moves no bytes: not a launch of the phase
        if kd["fetch_per_dispatch"] + kd["write_per_dispatch"] < 1e6:
            continue
        if short.startswith(prefix):
            f += kd["fetch_per_dispatch"] * kd["dispatches"]
            w += kd["write_per_dispatch"] * kd["dispatches"]
            cnt += kd["dispatches"]
            names.add(short.split("<")[0])
    if not cnt:
        return None
    return {"traffic": (f + w) / cnt, "fetch_bytes": f / cnt, "write_bytes": w / cnt, "dispatches": cnt,
            "kernels": sorted(names)}


# kernel-name prefixes of the sort's phases, per key type (the scatter phase
# runs k_scatter, or the line-buffered / full-line kernels the tuned defaults
# or options select)
_PHASE_KERNEL = {t: {"scatter": ("k_scatter<", "k_scatter_fl<", "k_scatter_lb<"), "count": "k_count<",
                     "tile": f"k_tile_sort<{t}", "copy": "k_copy_back"}
                 for t in ("double", "int")}


def _whole(traffic, op, ms, keys, bytes_per_key):
    d = (traffic or {}).get(op)
    out = {"ms": ms, "algorithmic_bytes": keys * bytes_per_key,
           "algorithmic_frac": keys * bytes_per_key / (ms / 1e3) / 1e9 / HBM_PEAK_GBS}
    if d:
        out.update({"hbm_bytes": d["bytes"], "fetch_bytes": d.get("FETCH_SIZE"), "write_bytes": d.get("WRITE_SIZE"),
                    "hbm_bytes_per_key": d["bytes"] / keys,
                    "measured_frac": d["bytes"] / (ms / 1e3) / 1e9 / HBM_PEAK_GBS})
    return out


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
                            "max_bucket_ratio": round(float(st.learned_max_bucket_ratio), 3),
                            "leaf_max_ratio": round(float(st.learned_leaf_max_ratio), 3)}
        sp, le = median(times["splitters"]), median(times["learned"])
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
    t = median(ms)
    out = {"metric": metric, "value": items / (t / 1e3), "unit": unit, "n_gpus": 1, "steps": steps, "warmup": warmup,
           "ms_per_step": t, "ms_mean": sum(ms) / len(ms), "higher_is_better": True, "scaling": "weak",
           "vs_baseline": None, "dtype": dtype, "data": data, "config": config}
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
                 {"verified": ok, "effective_hbm_frac": (16.0 * n / (median(ms) / 1e3) / 1e9) / HBM_PEAK_GBS})


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
                 {"verified": ok, "effective_hbm_frac": (2.0 * kb * n / (median(ms) / 1e3) / 1e9) / HBM_PEAK_GBS})


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
            r[name] = {"ms": median(ms), "keys_per_s": n / (median(ms) / 1e3)}
            if name == "learned":
                d = st["s"].as_dict()
                r[name].update({"classifier": d["classifier"], "fallback": d["learned_fallback"],
                                "max_bucket_ratio": round(d["learned_max_bucket_ratio"], 3),
                                "leaf_max_ratio": round(d["learned_leaf_max_ratio"], 3),
                                "model_loss": d["model_loss"]})
        r["learned_vs_splitters"] = r["splitters"]["ms"] / r["learned"]["ms"]
        res[kind] = r
    ms = [res["mixture"]["learned"]["ms"]]
    return _line("sorted keys/s, 1G fp64, learned partitioner", "keys/s", n, ms, args.steps, args.warmup, "f64",
                 "synthetic (device-generated mixture / piecewise)",
                 {"workload": f"learned-partition sort, {n} fp64 keys, 8-component Gaussian mixture + "
                              f"lognormal(0,2) tail; piecewise-constant adversarial input alongside"},
                 {"verified": ok, "learned": res})


def _profiled_phases(s512, work, restore, steps):
    """Per-phase kernel time, launches and keys over `steps` profiled sorts
    (library profile mode: HIP events around every kernel)."""
    ph_ms, ph_l, ph_k = {}, {}, {}
    for _ in range(steps):
        restore()
        st = s512.sort_ex(work, profile=True)
        for j, name in enumerate(s512.PHASES):
            ph_ms[name] = ph_ms.get(name, 0.0) + st.kernel_ms[j]
            ph_l[name] = ph_l.get(name, 0) + st.launches[j]
            ph_k[name] = ph_k.get(name, 0) + st.keys[j]
    return ph_ms, ph_l, ph_k


def _roofline(ph_ms, ph_l, ph_k, key_bytes, traffic_of=None):
    """Roofline of the phase with the most kernel time; traffic_of(phase)
    gives the counter runs' bytes per launch of that phase's kernel."""
    bpk = {"count": key_bytes, "scatter": 2 * key_bytes, "tile": 2 * key_bytes, "copy": 2 * key_bytes}
    dom = max((p for p in bpk if ph_l.get(p)), key=lambda p: ph_ms[p])
    traffic = traffic_of(dom) if traffic_of else None
    per_launch_bytes = ph_k[dom] * bpk[dom] / ph_l[dom]
    per_launch_ms = ph_ms[dom] / ph_l[dom]
    achieved = per_launch_bytes / (per_launch_ms / 1e3) / 1e9
    out = {"bound": "hbm", "kernel": KERNEL_OF_PHASE[dom], "achieved": achieved, "peak": HBM_PEAK_GBS, "unit": "GB/s",
           "frac": achieved / HBM_PEAK_GBS, "traffic": None, "bytes_per_launch": per_launch_bytes,
           "avg_launch_ms": per_launch_ms,
           "timing": "per-kernel HIP events on the library's stream, in a profiled pass after the timed steps"}
    if traffic:
        out["traffic"] = traffic["traffic"]
        out["traffic_detail"] = traffic
        out["measured_frac"] = traffic["traffic"] / (per_launch_ms / 1e3) / 1e9 / HBM_PEAK_GBS
        if len(traffic.get("kernels", ())) == 1:
            out["kernel_in_counter_run"] = traffic["kernels"][0]  # the kernel the counter runs saw in this phase
    if dom == "scatter":
        out["kernel"] = ("scatter phase: k_scatter_fl or k_scatter per level (a device-side gate picks one of the "
                         "pair from the level's runs; the other returns at once)")
    return out


def bench_plain_i32(args, traffic=None):
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
    steps_ms = [0.0] * args.steps
    ph_ms, ph_l, ph_k = {}, {}, {}
    for i, d in enumerate(dists):
        s512.fill(src, d, seed=i, param=1000)
        h0 = multiset_hash(src)
        ms = _timed(lambda: s512.sort_ex(work), lambda: work.copy_(src), args.steps, args.warmup)
        ok = ok and is_sorted(work) and multiset_hash(work) == h0
        per[d] = median(ms)
        for j in range(args.steps):
            steps_ms[j] += ms[j]
        a, b, c = _profiled_phases(s512, work, lambda: work.copy_(src), 1)
        for k in a:
            ph_ms[k] = ph_ms.get(k, 0.0) + a[k]
            ph_l[k] = ph_l.get(k, 0) + b[k]
            ph_k[k] = ph_k.get(k, 0) + c[k]
    roof = _roofline(ph_ms, ph_l, ph_k, 4, lambda ph: _kernel_traffic(traffic, "i32_sort", _PHASE_KERNEL["int"][ph]))
    del work, src
    torch.cuda.empty_cache()
    return _line("sorted keys/s + % HBM peak, 256M int32", "keys/s", n * len(dists), steps_ms, args.steps,
                 args.warmup, "int32", "synthetic (device-generated SplitMix64; four distributions)",
                 {"workload": f"plain sort, {n} int32 keys x {len(dists)} distributions ({','.join(dists)})"},
                 {"verified": ok, "per_distribution_ms": per, "roofline": roof,
                  "whole_sort_uniform": _whole(traffic, "i32_sort", per["uniform"], n, 8)})


def bench_large(args):
    """The plain sort past 2^32 keys: 4,294,967,296 + 4,099 int32 keys, uniform
    over the full range (pieces of 2^31 keys sorted by the partition levels,
    then merge-path rounds).  --n overrides the size."""
    import torch
    import s512
    n = args.n if args.n != 268_435_456 else (1 << 32) + 4099
    dev = torch.device("cuda:0")
    src = torch.empty(n, dtype=torch.int32, device=dev)
    work = torch.empty_like(src)
    s512.fill(src, "uniform", seed=3)
    h0 = multiset_hash(src)
    ms = _timed(lambda: s512.sort_ex(work), lambda: work.copy_(src), args.steps, args.warmup)
    ok = is_sorted(work) and multiset_hash(work) == h0
    del src, work
    torch.cuda.empty_cache()
    return _line("sorted keys/s, int32 past 2^32 keys", "keys/s", n, ms, args.steps, args.warmup, "int32",
                 "synthetic (device-generated SplitMix64, uniform)",
                 {"workload": f"plain sort, {n} int32 keys (uniform), pieces of 2^31 keys + merge-path rounds"},
                 {"verified": ok})


def bench_partition(args, traffic=None):
    """Partition512 (sort512.hpp:5821) on 268,435,456 fp64 keys uniform
    [0,1) in device memory, pivot 0.5: one read and one write of every key
    plus one copy of the smaller side.  Checked: the bound equals the count of
    keys <= pivot, the sides hold the right keys, the multiset is kept."""
    import torch
    import s512
    n = args.n
    dev = torch.device("cuda:0")
    src = torch.empty(n, dtype=torch.float64, device=dev)
    s512.fill(src, "uniform", seed=0)
    work = torch.empty_like(src)
    want = int((src <= 0.5).sum().item())
    h0 = multiset_hash(src)
    got = {}
    ms = _timed(lambda: got.update(b=s512.partition(work, 0, n - 1, 0.5)), lambda: work.copy_(src), args.steps,
                args.warmup)
    b = got["b"]
    ok = b == want and bool(torch.all(work[:b] <= 0.5).item()) and bool(torch.all(work[b:] > 0.5).item()) \
        and multiset_hash(work) == h0
    t = median(ms)
    roof = {"bound": "hbm", "kernel": "Partition512 (k_part_sample + k_part_pass + k_part_copyback, one call)",
            "achieved": 16.0 * n / (t / 1e3) / 1e9, "peak": HBM_PEAK_GBS, "unit": "GB/s",
            "frac": 16.0 * n / (t / 1e3) / 1e9 / HBM_PEAK_GBS, "traffic": None,
            "timing": "HIP events around the whole call (it reads the bound back)"}
    d = (traffic or {}).get("f64_partition")
    if d:
        roof["traffic"] = d["bytes"]
        roof["measured_frac"] = d["bytes"] / (t / 1e3) / 1e9 / HBM_PEAK_GBS
        roof["hbm_bytes_per_key"] = d["bytes"] / n
    del src, work
    torch.cuda.empty_cache()
    return _line("partitioned keys/s, Partition512, 256M fp64", "keys/s", n, ms, args.steps, args.warmup, "f64",
                 "synthetic (device-generated SplitMix64 uniform)",
                 {"workload": f"Partition512, {n} fp64 keys uniform [0,1), pivot 0.5, whole range"},
                 {"verified": ok, "bound": b, "roofline": roof})


def bench_merge(args, traffic=None):
    """The SortOmpMerge / SortOmpMergeDeps / SortOmpParMerge family
    (sort512.hpp:6006-6174) as the GPU merge sort: bitonic tiles, then
    merge-path rounds, on 268,435,456 fp64 keys uniform [0,1)."""
    import torch
    import s512
    n = args.n
    dev = torch.device("cuda:0")
    src = torch.empty(n, dtype=torch.float64, device=dev)
    s512.fill(src, "uniform", seed=0)
    work = torch.empty_like(src)
    h0 = multiset_hash(src)
    ms = _timed(lambda: s512.sort_merge(work), lambda: work.copy_(src), args.steps, args.warmup)
    ok = is_sorted(work) and multiset_hash(work) == h0
    t = median(ms)
    rounds = max(0, math.ceil(math.log2(n / 2048)))
    d = (traffic or {}).get("f64_merge")
    extra = {"verified": ok, "merge_rounds": rounds, "whole_sort": _whole(traffic, "f64_merge", t, n, 16)}
    mk = _kernel_traffic(traffic, "f64_merge", "k_merge_tile")
    if mk:
        extra["k_merge_tile_traffic_per_dispatch"] = mk
    del src, work, d
    torch.cuda.empty_cache()
    return _line("sorted keys/s, merge sort (SortOmpMerge family), 256M fp64", "keys/s", n, ms, args.steps,
                 args.warmup, "f64", "synthetic (device-generated SplitMix64 uniform)",
                 {"workload": f"merge sort, {n} fp64 keys uniform [0,1)"}, extra)


def bench_smallsort(args, traffic=None):
    """SmallSort16V (sort512.hpp:2255) on a batch: 2,097,152 arrays of 128
    fp64 keys (its largest fp64 length) sorted by the segmented sort, every
    array by one wave's register network.  Checked per array."""
    import numpy as np
    import torch
    import s512
    n = args.n
    dev = torch.device("cuda:0")
    off = torch.from_numpy(np.arange(0, n + 1, 128, dtype=np.int64)).to(dev)
    src = torch.empty(n, dtype=torch.float64, device=dev)
    s512.fill(src, "uniform", seed=0)
    work = torch.empty_like(src)
    h0 = _seg_hash(src, off)
    ms = _timed(lambda: s512.segsort(work, off), lambda: work.copy_(src), args.steps, args.warmup)
    ok = _seg_sorted(work, off) and bool(torch.equal(_seg_hash(work, off), h0))
    t = median(ms)
    del src, work, off
    torch.cuda.empty_cache()
    return _line("sorted keys/s, SmallSort16V batch of 128-key fp64 arrays", "keys/s", n, ms, args.steps,
                 args.warmup, "f64", "synthetic (device-generated SplitMix64 uniform)",
                 {"workload": f"{n // 128} arrays of 128 fp64 keys (SmallSort16V length limit), one call"},
                 {"verified": ok, "whole": _whole(traffic, "f64_smallsort", t, n, 16)})


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
    return _line("sorted keys/s end to end from host memory, 256M fp64", "keys/s", n, ms, args.steps, args.warmup,
                 "f64", "synthetic (numpy PCG64 uniform, pageable host memory)",
                 {"workload": f"host-pointer sort, {n} fp64 keys, pageable host array"},
                 {"verified": ok, "pcie_bound_ms": pcie_ms, "ratio_to_pcie_bound": median(ms) / pcie_ms})


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
    return _line("sorted keys/s, 256M int64 keys", "keys/s", n, ms_i64, args.steps, args.warmup, "int64",
                 "synthetic (torch.randint full int64 range; fp64 normal for argsort)",
                 {"workload": f"int64 key sort and fp64 argsort, {n} keys each"},
                 {"verified": ok, "int64_sort_ms": median(ms_i64), "argsort_f64_ms": median(ms_arg),
                  "argsort_f64_keys_per_s": n / (median(ms_arg) / 1e3)})


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
    return _line("keys/s scanned, top-1M of 256M fp64", "keys/s", n, ms_topk, args.steps, args.warmup, "f64",
                 "synthetic (device-generated normal)",
                 {"workload": f"top-k, {n} fp64 keys, k = {k}; selection of the median"},
                 {"verified": ok, "topk_ms": median(ms_topk), "select_median_ms": median(ms_sel)})


WORKLOADS = {"plain": None, "plain-i32": bench_plain_i32, "host": bench_host, "wide": bench_wide,
             "topk": bench_topk, "kv-zipf": bench_kv_zipf, "segsort-i32": lambda a: bench_segsort(a, "int32"),
             "segsort-f64": lambda a: bench_segsort(a, "float64"), "learned-1g": bench_learned,
             "partition": bench_partition, "merge": bench_merge, "smallsort": bench_smallsort,
             "large": bench_large}


def _summary(line):
    """The parts of a secondary line that the headline keeps."""
    keep = ("metric", "value", "unit", "ms_per_step", "ms_mean", "dtype", "config", "verified", "roofline", "whole",
            "whole_sort", "whole_sort_uniform", "per_distribution_ms", "bound", "merge_rounds", "learned",
            "effective_hbm_frac", "k_merge_tile_traffic_per_dispatch")
    return {k: v for k, v in line.items() if k in keep}


def main():
    args = parse()
    if args.pmc_child:
        pmc_child(args.pmc_child, args.n)
        return
    if args.cpu_child:
        cpu_child(json.loads(args.cpu_child))
        return
    if args.gen_child:
        gen_child(args.gen_child)
        return
    if args.gpus != 1:
        print(f"bench.py measures one GPU: every box in this setup has one GPU and the sort runs on one device "
              f"(asked for --gpus {args.gpus}).", file=sys.stderr)
        sys.exit(2)
    if args.workload != "plain":
        out = WORKLOADS[args.workload](args)
        print(json.dumps(out))
        sys.exit(0 if out.get("verified") else 1)
    # child processes first, before this process touches the GPU: the counter
    # runs, then the CPU baseline (which starts its children with the OpenMP
    # binding variables set)
    traffic = None
    if not args.no_traffic:
        try:
            traffic = measure_traffic(args.n, args.n)
        except Exception as e:  # reported, not hidden
            traffic = {"error": repr(e)}
    cpu = None
    if not args.no_cpu_baseline:
        try:
            cpu = cpu_baseline(args.cpu_sample, args.cpu_budget)
        except Exception as e:  # reported, not hidden
            cpu = {"error": repr(e)}
    import torch
    import s512
    progress("headline: timed steps")
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
    chosen = dict(kv.split("=", 1) for kv in args.option)

    def options():
        """Library defaults unless --option set a field."""
        o = s512.Options()
        for k, v in chosen.items():
            if not hasattr(o, k):
                raise SystemExit(f"--option: s512_options has no field {k!r}")
            setattr(o, k, int(v))
        return o

    # ---- timed steps: the production call, HIP events around each sort only
    ev0 = torch.cuda.Event(enable_timing=True)
    ev1 = torch.cuda.Event(enable_timing=True)
    per_dist_ms = {d: [] for d in dists}
    step_ms = []
    levels, checked = {}, {}
    torch.cuda.synchronize()
    wall0 = None
    for step in range(args.warmup + args.steps):
        timed = step >= args.warmup
        if timed and wall0 is None:
            torch.cuda.synchronize()
            wall0 = time.perf_counter()
        tot = 0.0
        for d in dists:
            work.copy_(pristine[d])
            ev0.record()
            st = s512.sort_ex(work, options())
            ev1.record()
            ev1.synchronize()
            if not timed:
                continue
            ms = ev0.elapsed_time(ev1)
            per_dist_ms[d].append(ms)
            tot += ms
            levels[d] = st.levels
            if d not in checked:
                checked[d] = is_sorted(work) and multiset_hash(work) == hashes[d]
        if timed:
            step_ms.append(tot)
    torch.cuda.synchronize()
    wall = time.perf_counter() - wall0
    # ---- profiled pass: per-kernel HIP events inside the library
    ph_ms, ph_l, ph_k = {}, {}, {}
    prof_ms = []
    for step in range(args.steps):
        for d in dists:
            work.copy_(pristine[d])
            ev0.record()
            st = s512.sort_ex(work, options(), profile=True)
            ev1.record()
            ev1.synchronize()
            prof_ms.append(ev0.elapsed_time(ev1))
            for i, name in enumerate(s512.PHASES):
                ph_ms[name] = ph_ms.get(name, 0.0) + st.kernel_ms[i]
                ph_l[name] = ph_l.get(name, 0) + st.launches[i]
                ph_k[name] = ph_k.get(name, 0) + st.keys[i]
    t_step = median(step_ms)
    keys_step = n * len(dists)
    value = keys_step / (t_step / 1e3)
    roof = _roofline(ph_ms, ph_l, ph_k, 8,
                     lambda ph: _kernel_traffic(traffic, "f64_sort", _PHASE_KERNEL["double"][ph]))
    for d in dists:  # free the inputs before the other workloads
        pristine[d] = None
    del work
    torch.cuda.empty_cache()
    uni = median(per_dist_ms["uniform"]) if "uniform" in per_dist_ms else None
    out = {
        "metric": "sorted keys/s + % HBM peak, 256M fp64; learned vs sampled splitters",
        "value": value,
        "unit": "keys/s",
        "n_gpus": 1,
        "steps": args.steps,
        "warmup": args.warmup,
        "ms_per_step": t_step,
        "ms_per_step_mean": sum(step_ms) / len(step_ms),
        "higher_is_better": True,
        "scaling": "weak",
        "vs_baseline": None,
        "dtype": "f64",
        "data": "synthetic (device-generated SplitMix64 / Box-Muller; five distributions)",
        "config": {"workload": f"plain sort, {n} fp64 keys x {len(dists)} distributions ({','.join(dists)})",
                   "keys_per_step": keys_step, "classifier": "sampled splitters",
                   "options": chosen or "library defaults"},
        "roofline": roof,
        "whole_sort": _whole(traffic, "f64_sort", uni, n, 16) if uni else None,
        "effective_hbm_frac": (16.0 * keys_step / (t_step / 1e3) / 1e9) / HBM_PEAK_GBS,
        "per_distribution_ms": {d: median(v) for d, v in per_dist_ms.items()},
        "per_distribution_keys_per_s": {d: n / (median(v) / 1e3) for d, v in per_dist_ms.items()},
        "levels": levels,
        "phase_ms_per_step": {p: ph_ms[p] / args.steps for p in ph_ms if ph_l.get(p)},
        "profiled_ms_per_step": sum(prof_ms) / args.steps,
        "verified": checked,
        "wall_s_timed_region": wall,
        "int32_256M": None,
        "entry_points": None,
        "learned": None,
        "cpu_baseline": cpu,
        "traffic_runs": {k: {kk: vv for kk, vv in v.items() if kk != "kernels"} if isinstance(v, dict) else v
                         for k, v in (traffic or {}).items()} if traffic else None,
    }
    if not args.no_extras:
        sub = argparse.Namespace(**vars(args))
        progress("int32 256M")
        try:
            out["int32_256M"] = _summary(bench_plain_i32(sub, traffic))
        except Exception as e:  # reported, not hidden
            out["int32_256M"] = {"error": repr(e)}
        eps = {}
        for name, fn in (("Partition512", bench_partition), ("SortOmpMerge family", bench_merge),
                         ("SmallSort16V batch", bench_smallsort)):
            progress(name)
            try:
                eps[name] = _summary(fn(sub, traffic))
            except Exception as e:  # reported, not hidden
                eps[name] = {"error": repr(e)}
        small = argparse.Namespace(**dict(vars(args), steps=max(5, args.steps), warmup=1))
        for name, fn in (("Sort512kv::Sort, config 3", bench_kv_zipf),
                         ("segmented int32, config 4", lambda a: bench_segsort(a, "int32")),
                         ("segmented fp64, config 4", lambda a: bench_segsort(a, "float64"))):
            progress(name)
            try:
                line = _summary(fn(small))
                if name.startswith("Sort512kv"):
                    d = (traffic or {}).get("i32_kv")
                    if d:
                        line["whole"] = _whole(traffic, "i32_kv", line["ms_per_step"], 1 << 29, 16)
                eps[name] = line
            except Exception as e:  # reported, not hidden
                eps[name] = {"error": repr(e)}
            torch.cuda.empty_cache()
        progress("learned partitioner, 2^30 keys")
        try:
            eps["learned partitioner, config 5"] = _summary(bench_learned(
                argparse.Namespace(**dict(vars(args), steps=3, warmup=1))))
        except Exception as e:  # reported, not hidden
            eps["learned partitioner, config 5"] = {"error": repr(e)}
        torch.cuda.empty_cache()
        out["entry_points"] = eps
    if not args.no_learned:
        progress("learned vs splitters, 256M")
        try:
            out["learned"] = learned_comparison(s512, n)
        except Exception as e:  # reported, not hidden
            out["learned"] = {"error": repr(e)}
    ok = all(checked.values())
    for sect in ((out["int32_256M"] or {}),) + tuple((out["entry_points"] or {}).values()):
        if isinstance(sect, dict) and sect.get("verified") is False:
            ok = False
    if not ok:
        out["error"] = "verification failed"
    print(json.dumps(out))
    sys.exit(0 if ok else 1)


if __name__ == "__main__":
    main()
