"""Benchmark of the NArray data-parallel path on one MI355X.

Headline workload (the metric's configuration): strided view gather and
scatter, Int32 8192x8192, values U[0, 2^31) from seed 1.  One step is
    g = x.view(1..3..-1, ..-2..).to_narr()          # strided gather
    z = NArray.fill([8192, 8192], 0)                 # zeroed destination
    z.mutable_view(1..3..-1, ..-2..)[..] = g         # scatter back
(the fill and the scatter run as one fused kernel).  `value` is algorithmic
HBM GB/s of the whole step with inputs resident in HBM:
    gather   8 B/element x 11,186,176 elements            =  89.49 MB
    fill+scatter: 4 B x 67,108,864 written + 4 B x 11,186,176 read = 313.18 MB
The other workloads (broadcast, fused ProcView reductions, heat stencil,
batched small arrays) are reported under "extra".

Usage: python bench.py [--gpus 1] [--steps K] [--warmup W] [--no-cpu] [--quick]
"""
import argparse
import ctypes
import json
import os
import sys
import time

ROOT = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, os.path.join(ROOT, "ph-core_amd"))

PEAK_HBM_GBS = 8000.0


def parse():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=50)
    ap.add_argument("--warmup", type=int, default=5)
    ap.add_argument("--no-cpu", action="store_true", help="skip the CPU legs and the C3 error check")
    ap.add_argument("--quick", action="store_true", help="headline workload only")
    ap.add_argument("--heat-steps", type=int, default=2000)
    ap.add_argument("--no-unbatched", action="store_true",
                    help="skip timing the batched workload as per-array launches (60000 kernels)")
    ap.add_argument("--only", type=lambda v: set(v.split(",")), default=None,
                    help="run only these extra workloads (c1,copy,c3,c4,c5,concat,json,path), for profiling")
    return ap.parse_args()


class Events:
    """HIP events on the library stream (where every kernel is launched)."""

    def __init__(self, phase):
        self.lib = ctypes.CDLL(os.path.join(phase.LIB_DIR, "libphhip.so"))
        self.lib.ph_event_elapsed.argtypes = [ctypes.POINTER(ctypes.c_float), ctypes.c_void_p, ctypes.c_void_p]
        self.lib.ph_event_record.argtypes = [ctypes.c_void_p]

    def new(self):
        e = ctypes.c_void_p()
        assert self.lib.ph_event_create(ctypes.byref(e)) == 0
        return e

    def record(self, e):
        assert self.lib.ph_event_record(e) == 0

    def ms(self, a, b):
        v = ctypes.c_float()
        assert self.lib.ph_event_elapsed(ctypes.byref(v), a, b) == 0
        return v.value

    def sync(self):
        assert self.lib.ph_sync() == 0


_BUSY = {}


def busy(phase):
    """Queues a few milliseconds of read-only device work (FP64 sums over a
    256 MiB Float32 array through ph_l2_diff, whose kernels no workload
    below uses) so that the launches queued after it start back to back: the
    events around them then time the kernels, not the host's launch gaps."""
    D = phase.DType
    if "buf" not in _BUSY:
        lib = ctypes.CDLL(os.path.join(phase.LIB_DIR, "libphhip.so"))
        lib.ph_l2_diff.argtypes = [ctypes.c_int32, ctypes.c_void_p, ctypes.c_int32, ctypes.c_void_p,
                                   ctypes.c_int64, ctypes.c_void_p]
        _BUSY["lib"] = lib
        _BUSY["buf"] = phase.NArray.random([1 << 26], D.F32, seed=100)
        _BUSY["out"] = phase.NArray.empty([2], D.F64)
    p, o = _BUSY["buf"].device_ptr, _BUSY["out"].device_ptr
    for _ in range(24):
        assert _BUSY["lib"].ph_l2_diff(1, p, 1, p, 1 << 26, o) == 0


def time_loop(ev, fn, reps, phase=None):
    """Average device time of `fn` over `reps` back-to-back calls (HIP events
    on the library stream).  With `phase`, device work is queued first so the
    host's launch overhead does not open gaps between the calls."""
    a, b = ev.new(), ev.new()
    ev.sync()
    if phase is not None:
        busy(phase)
    ev.record(a)
    for _ in range(reps):
        fn()
    ev.record(b)
    return ev.ms(a, b) / reps


def load_traffic():
    """Measured HBM bytes per launch of the dominant kernel (rocprofv3 PMC),
    if a profile summary with them has been committed."""
    p = os.path.join(ROOT, "profiles", "pmc_traffic.json")
    if os.path.exists(p):
        try:
            return json.load(open(p))
        except Exception:
            return None
    return None


def cpu_baseline(xn, lits_parsed, n, alg_bytes):
    """ph-core's own CPU path, restated: NArray#unsafe_fetch_chunk (buffered
    lex odometer), NArray.fill, NArray#unsafe_set_chunk — single thread, on a
    bounded sample (the full step repeated for about 10 s)."""
    sys.path.insert(0, os.path.join(ROOT, "tests"))
    import numpy as np
    import oracle_api as O
    r = O.region(lits_parsed, [n, n])
    reps = 0
    t0 = time.perf_counter()
    while True:
        g = O.gather(xn, r)
        z = np.zeros((n, n), dtype=np.int32)  # NArray.fill(shape, 0)
        O.scatter(z, r, g)
        reps += 1
        if time.perf_counter() - t0 > 10.0 or reps >= 20:
            break
    dt = (time.perf_counter() - t0) / reps
    return {"value": round(alg_bytes / dt / 1e9, 3), "unit": "GB/s", "cores": 1, "kind": "port",
            "sample": f"{reps} full steps (gather + fill + scatter of the 8192x8192 Int32 workload) "
                      f"with the oracle's restatement of ph-core's buffered odometer path; "
                      f"{dt * 1e3:.1f} ms/step"}


def _oracle():
    sys.path.insert(0, os.path.join(ROOT, "tests"))
    import oracle_api as O
    return O


def cpu_c1(a, b, c):
    """ph-core's (a + b.tile(1024, 1)) * c restated by the oracle, 1 core."""
    import numpy as np
    O = _oracle()
    bt = np.tile(b, (a.shape[0], 1))
    reps, t0 = 0, time.perf_counter()
    while reps < 5 or time.perf_counter() - t0 < 2.0:
        t, _ = O.eval_expr("(add x0 x1)", [a, bt])
        O.eval_expr("(mul x0 x1)", [t, c])
        reps += 1
        if reps >= 50:
            break
    dt = (time.perf_counter() - t0) / reps
    by = 24 * a.size + 8 * b.size
    return {"value": round(by / dt / 1e9, 3), "unit": "GB/s", "cores": 1, "kind": "port",
            "sample": f"{reps} evaluations of the two elementwise ops at 1024^2 (tile materialised), {dt * 1e3:.1f} ms each"}


def cpu_c3(x):
    """ph-core's ProcView block, then sum / max in full and per row / column
    (slices), restated by the oracle on the first 512 rows, 1 core."""
    import numpy as np
    from crystal import trace
    O = _oracle()
    rows = 512
    xs = np.ascontiguousarray(x[:rows])
    t0 = time.perf_counter()
    f, _ = O.eval_expr(trace(lambda v: (v + 4) ** 2 * 0.5 - v), [xs])
    O.sum_seq(f)
    O.max_seq(f)
    O.slices_reduce(f, 0, 0)
    O.slices_reduce(f, 0, 1)
    O.slices_reduce(f, 1, 0)
    O.slices_reduce(f, 1, 1)
    dt = time.perf_counter() - t0
    by = 4 * xs.size
    return {"value": round(by / dt / 1e9, 3), "unit": "GB/s (input bytes)", "cores": 1, "kind": "port",
            "sample": f"the first {rows} of 16384 rows: block evaluation + sequential sum, max and the four "
                      f"slice reductions, {dt * 1e3:.0f} ms; about {dt * 16384 / rows:.1f} s for the full array"}


def c3_errors(x, storage, res):
    """Relative error of the device's six-result pass against the same block
    evaluated by numpy in Crystal's types ((x+4)**2 in Float32, then * 0.5
    and - x in Float64) and summed in 80-bit long double, 1024 rows at a
    time; beside it, the distance to ph-core's own sums (Enumerable#sum:
    sequential Float64 additions in lexicographic order, by the oracle for
    the full and row sums, row after row for the column sums) and those
    sequential sums' own error; maxima are compared bit for bit."""
    import numpy as np
    O = _oracle()
    if storage == "BF16":
        x = (x.astype(np.uint32) << 16).view(np.float32)
    n, m = x.shape
    rows = np.zeros(n, np.longdouble)
    cols = np.zeros(m, np.longdouble)
    seq_rows = np.zeros(n)
    seq_cols = np.zeros(m)
    seq_full = 0.0
    rmax = np.empty(n)
    cmax = np.full(m, -np.inf)
    for r0 in range(0, n, 1024):
        xs = x[r0:r0 + 1024]
        t = xs + np.float32(4)
        f = (t * t).astype(np.float64) * 0.5 - xs.astype(np.float64)
        fl = f.astype(np.longdouble)
        rows[r0:r0 + 1024] = fl.sum(axis=1)
        cols += fl.sum(axis=0)
        rmax[r0:r0 + 1024] = f.max(axis=1)
        cmax = np.maximum(cmax, f.max(axis=0))
        seq_full = O.sum_seq_from(f, seq_full)
        seq_rows[r0:r0 + 1024] = O.slices_reduce(f, 0, 0)
        for row in f:
            seq_cols += row
    full = rows.sum()

    def rel(got, want):
        got = np.asarray(got, np.longdouble)
        want = np.asarray(want, np.longdouble)
        return float(np.max(np.abs(got - want) / np.abs(want)))

    return {"max_rel_err_vs_fp64": {"full_sum": rel(res["full_sum"], full), "row_sum": rel(res["row_sum"], rows),
                                    "col_sum": rel(res["col_sum"], cols)},
            "max_rel_dist_to_sequential_sum": {"full_sum": rel(res["full_sum"], seq_full),
                                               "row_sum": rel(res["row_sum"], seq_rows),
                                               "col_sum": rel(res["col_sum"], seq_cols)},
            "sequential_sum_rel_err_vs_fp64": {"full_sum": rel(seq_full, full), "row_sum": rel(seq_rows, rows),
                                               "col_sum": rel(seq_cols, cols)},
            "maxima_bit_exact": bool(res["full_max"] == rmax.max() and np.array_equal(res["row_max"], rmax)
                                     and np.array_equal(res["col_max"], cmax))}


def cpu_info():
    """The host the CPU legs ran on: CPU model and the cores this process may
    use (the legs themselves use one)."""
    model = "unknown"
    try:
        for line in open("/proc/cpuinfo"):
            if line.startswith("model name"):
                model = line.split(":", 1)[1].strip()
                break
    except OSError:
        pass
    try:
        nproc = len(os.sched_getaffinity(0))
    except AttributeError:
        nproc = os.cpu_count()
    return {"cpu_model": model, "nproc": nproc}


def stamp_cpu_legs(obj, info):
    """Adds the host's CPU model and nproc to every CPU leg (a dict with
    "cores" and "kind") inside the bench line."""
    if isinstance(obj, dict):
        if "cores" in obj and "kind" in obj:
            obj.update(info)
        for v in obj.values():
            stamp_cpu_legs(v, info)
    elif isinstance(obj, list):
        for v in obj:
            stamp_cpu_legs(v, info)


def cpu_c4(u0):
    """The heat step restated by the oracle on the full 8192^2 FP64 grid for
    2 steps, 1 core; Gcell/s."""
    O = _oracle()
    steps = 2
    t0 = time.perf_counter()
    O.heat2d(u0, 0.2, steps)
    dt = time.perf_counter() - t0
    return {"value": round(u0.size * steps / dt / 1e9, 4), "unit": "Gcell/s", "cores": 1, "kind": "port",
            "sample": f"{steps} FP64 steps of the full 8192^2 grid, {dt:.2f} s"}


def main():
    args = parse()
    if args.gpus != 1:
        print(json.dumps({"error": "this build measures one GPU; multi-GPU runs are not part of it",
                          "n_gpus": args.gpus}))
        sys.exit(2)
    import numpy as np
    import phase
    phase.init(0)
    ev = Events(phase)
    D = phase.DType

    # ------------------------------------------------------------ headline C2
    n = 8192
    lits = ["1..3..-1", "..-2.."]
    x = phase.NArray.random([n, n], D.I32, seed=1, dist="int31")
    view = x.view(*lits)
    ng = view.size  # 2731 * 4096
    gather_bytes = 8 * ng
    scatter_bytes = 4 * n * n + 4 * ng
    step_bytes = gather_bytes + scatter_bytes

    def step():
        g = view.to_narr()
        z = phase.NArray.fill([n, n], 0)
        z.mutable_view(*lits)[[]] = g
        return z

    for _ in range(args.warmup):
        step()
    ev.sync()
    t0 = time.perf_counter()
    a, b = ev.new(), ev.new()
    ev.record(a)
    for _ in range(args.steps):
        step()
    ev.record(b)
    ev.sync()
    wall = time.perf_counter() - t0
    ms_step = ev.ms(a, b) / args.steps
    value = step_bytes / (ms_step * 1e-3) / 1e9

    # per-kernel durations (HIP events around back-to-back launches)
    g = view.to_narr()
    ms_gather = time_loop(ev, lambda: view.to_narr(), 20, phase)
    # the gather re-reads one third of x (about 134 MB of lines), which the
    # 256 MB MALL may partly hold between back-to-back launches: also time
    # each launch alone after reading a 512 MiB array (a MALL flush; a
    # written one would leave dirty lines whose write-back the gather pays)
    flush = phase.NArray.fill([1 << 27], 0).materialize()
    fa, fb = ev.new(), ev.new()
    tot = 0.0
    for _ in range(10):
        flush.sum()
        ev.record(fa)
        view.to_narr()
        ev.record(fb)
        ev.sync()
        tot += ev.ms(fa, fb)
    ms_gather_flushed = tot / 10
    del flush

    # fill + scatter: 8 back-to-back launches, each into a fresh pending fill
    zs = [phase.NArray.fill([n, n], 0) for _ in range(8)]
    mvs = [z.mutable_view(*lits) for z in zs]
    aa, bb = ev.new(), ev.new()
    ev.sync()
    busy(phase)
    ev.record(aa)
    for mv in mvs:
        mv[[]] = g
    ev.record(bb)
    ms_fs = ev.ms(aa, bb) / len(zs)
    achieved_fs = scatter_bytes / (ms_fs * 1e-3) / 1e9
    achieved_g = gather_bytes / (ms_gather * 1e-3) / 1e9

    traffic = load_traffic()
    roofline = {"bound": "hbm", "kernel": "k_fill_scatter_vec<uint32_t,4> (fused NArray.fill + MutableView scatter)",
                "achieved": round(achieved_fs, 1), "peak": PEAK_HBM_GBS, "unit": "GB/s",
                "frac": round(achieved_fs / PEAK_HBM_GBS, 4),
                "traffic": traffic.get("k_fill_scatter") if traffic else None,
                "ms_per_launch": round(ms_fs, 4), "algorithmic_bytes": scatter_bytes}
    extra = {"gather": {"kernel": "k_copy_rows<uint32_t,8,int,DCONT> (strided gather)", "achieved_gbs": round(achieved_g, 1),
                        "frac": round(achieved_g / PEAK_HBM_GBS, 4), "ms_per_launch": round(ms_gather, 4),
                        "algorithmic_bytes": gather_bytes,
                        "traffic": traffic.get("k_copy_rows") if traffic else None,
                        "mall_flushed": {"ms_per_launch": round(ms_gather_flushed, 4),
                                         "achieved_gbs": round(gather_bytes / ms_gather_flushed / 1e6, 1),
                                         "note": "each launch alone, after 512 MiB read (event timing "
                                                 "includes one launch's ramp-up)"}}}
    del zs, mvs, g

    # plain contiguous copy of 256 MiB (Int32): the practical copy ceiling of
    # the device, and the known-size kernel that scales the PMC byte counters
    cx = phase.NArray.random([1 << 26], D.I32, seed=9, dist="int31")
    cx.clone()
    ms_copy = time_loop(ev, lambda: cx.clone(), 10, phase)
    extra["copy_256mib"] = {"kernel": "k_copy_rows (contiguous)", "bytes": 2 << 28, "ms_per_launch": round(ms_copy, 4),
                            "achieved_gbs": round((2 << 28) / ms_copy / 1e6, 1)}
    del cx

    if not args.quick:
        extra.update(other_workloads(phase, ev, args))

    # the same fraction against a measured copy peak (a 2 GiB device copy,
    # beyond the 256 MB MALL), beside the 8 TB/s nominal one
    cp = extra.get("copy_2gib", {}).get("achieved_gbs")
    if cp:
        roofline["copy_peak_gbs"] = cp
        roofline["frac_of_copy_peak"] = round(roofline["achieved"] / cp, 4)

    cpu = None
    if not args.no_cpu:
        sys.path.insert(0, os.path.join(ROOT, "tests"))
        from crystal import parse_literal
        cpu = cpu_baseline(x.to_numpy(), [parse_literal(s) for s in lits], n, step_bytes)

    line = {
        "metric": "HBM GB/s & % of peak; heat-stencil Gcell/s FP32 vs FP64 at 1e-5 rel err",
        "value": round(value, 1),
        "unit": "GB/s",
        "n_gpus": 1,
        "steps": args.steps,
        "warmup": args.warmup,
        "ms_per_step": round(ms_step, 4),
        "higher_is_better": True,
        "scaling": "weak",
        "vs_baseline": None,
        "dtype": "i32",
        "data": "synthetic (seeded splitmix64 on the device)",
        "config": {"workload": "strided view gather + scatter, Int32 8192x8192, view(1..3..-1, ..-2..) "
                               "-> [2731, 4096], written back into a zeroed array through a MutableView",
                   "step_bytes": step_bytes, "frac_of_peak": round(value / PEAK_HBM_GBS, 4),
                   "wall_s_timed": round(wall, 3)},
        "roofline": roofline,
        "cpu_baseline": cpu,
        "extra": extra,
    }
    stamp_cpu_legs(line, cpu_info())
    print(json.dumps(line))


def other_workloads(phase, ev, args):
    import numpy as np
    D = phase.DType
    out = {}
    want = lambda name: not args.only or name in args.only  # noqa: E731
    tr = load_traffic() or {}
    if want("c1"):
        # C1 broadcast at the roofline shape (16384^2 FP64; 1024^2 is the parity case)
        n = 16384
        a = phase.NArray.random([n, n], D.F64, seed=0, offset=0, a=-1.0, b=1.0)
        b = phase.NArray.random([1, n], D.F64, seed=0, offset=n * n, a=-1.0, b=1.0)
        c = phase.NArray.random([n, n], D.F64, seed=0, offset=n * n + n, a=-1.0, b=1.0)
        bt = b.tile(n, 1)
        by = 24 * n * n + 8 * n
        # the operator form as ph-core code writes it: a + b.tile is left
        # pending and fused into the `* c` kernel when the result is used
        fo = lambda: ((a + bt) * c).materialize()  # noqa: E731
        fo()
        mso = time_loop(ev, fo, 10, phase)
        # the same with the intermediate forced (two kernels and a temporary, as
        # ph-core's eager operators evaluate it)
        def f2():
            t = (a + bt).materialize()
            return (t * c).materialize()
        f2()
        ms2 = time_loop(ev, f2, 10, phase)
        # map_with: one kernel, same per-element operations and roundings
        fz = lambda: a.map_with(bt, c, block=lambda x, y, z: (x + y) * z).materialize()  # noqa: E731
        fz()
        msf = time_loop(ev, fz, 10, phase)
        out["c1_broadcast_f64_16384"] = {
            "operators": {"ms": round(mso, 4), "achieved_gbs": round(by / mso / 1e6, 1),
                          "frac": round(by / mso / 1e6 / PEAK_HBM_GBS, 4),
                          "note": "(a + b.tile) * c with ph-core's operators: the pending a + b.tile fuses into the "
                                  "product's kernel (one kernel)"},
            "two_kernels": {"ms": round(ms2, 4), "achieved_gbs": round(by / ms2 / 1e6, 1),
                            "note": "the same with the intermediate forced: two kernels and a temporary"},
            "fused": {"ms": round(msf, 4), "achieved_gbs": round(by / msf / 1e6, 1),
                      "frac": round(by / msf / 1e6 / PEAK_HBM_GBS, 4),
                      "note": "a.map_with(b.tile, c) { (x + y) * z }: one kernel, stride-0 broadcast"},
            "bytes": by}
        del a, b, c, bt
        out["c1_broadcast_f64_16384"]["fused"]["traffic"] = tr.get("ph_map_fused_c1")

    if want("c1"):
        # C1 at its parity shape, 1024^2 FP64 (25 MB: cache-resident, launch-bound)
        m = 1024
        a = phase.NArray.random([m, m], D.F64, seed=0, offset=0, a=-1.0, b=1.0)
        b = phase.NArray.random([1, m], D.F64, seed=0, offset=m * m, a=-1.0, b=1.0)
        c = phase.NArray.random([m, m], D.F64, seed=0, offset=m * m + m, a=-1.0, b=1.0)
        bt = b.tile(m, 1)
        f1 = lambda: ((a + bt) * c).materialize()  # noqa: E731
        f1()
        ms1 = time_loop(ev, f1, 50, phase)
        out["c1_broadcast_f64_1024"] = {"ms": round(ms1, 4), "bytes": 24 * m * m + 8 * m,
                                        "achieved_gbs": round((24 * m * m + 8 * m) / ms1 / 1e6, 1),
                                        "note": "(a + b.tile) * c with the operators (one fused kernel); inputs fit "
                                                "the 256 MB MALL"}
        if not args.no_cpu:
            out["c1_broadcast_f64_1024"]["cpu"] = cpu_c1(a.to_numpy(), b.to_numpy(), c.to_numpy())
        del a, b, c, bt

    if want("copy"):
        # the device's copy rate for a working set well beyond the MALL (2 GiB)
        big = phase.NArray.random([1 << 29], D.I32, seed=9, dist="int31")
        big.clone()
        msb = time_loop(ev, lambda: big.clone(), 5, phase)
        out["copy_2gib"] = {"kernel": "k_copy_rows (contiguous)", "bytes": 2 << 31, "ms": round(msb, 4),
                            "achieved_gbs": round((2 << 31) / msb / 1e6, 1),
                            "note": "practical HBM copy ceiling for multi-GB working sets"}
        # the runtime's own device-to-device copy of the same 2 GiB
        # (hipMemcpyAsync through ph_d2d), independent of this build's kernels
        dst = phase.NArray.empty([1 << 29], D.I32)
        d2d = ev.lib.ph_d2d
        d2d.argtypes = [ctypes.c_void_p, ctypes.c_void_p, ctypes.c_size_t]
        pd, ps = dst.device_ptr, big.device_ptr
        assert d2d(pd, ps, 1 << 31) == 0
        msd = time_loop(ev, lambda: d2d(pd, ps, 1 << 31), 5, phase)
        out["copy_2gib"]["hipmemcpy_d2d"] = {"ms": round(msd, 4), "achieved_gbs": round((2 << 31) / msd / 1e6, 1),
                                             "note": "hipMemcpyAsync device to device, 2 GiB"}
        del dst
        del big

    if want("c3"):
        # C3 fused ProcView + six reductions (F32 and BF16 storage), one pass
        n = 16384
        for st, esz in [("F32", 4), ("BF16", 2)]:
            x = phase.NArray.random([n, n], getattr(D, st), seed=2, dist="normal", a=0.0, b=1.0)
            pv = x.process(lambda v: (v + 4) ** 2 * 0.5 - v)
            pv.reduce2d_all_device()
            ms = time_loop(ev, lambda: pv.reduce2d_all_device(), 10, phase)
            by = esz * n * n
            xh = x.to_numpy()
            if not args.no_cpu and st == "F32":
                out["c3_cpu"] = cpu_c3(xh)
            out[f"c3_procview_reduce_{st.lower()}_16384"] = {
                "ms": round(ms, 4), "achieved_gbs": round(by / ms / 1e6, 1),
                "frac": round(by / ms / 1e6 / PEAK_HBM_GBS, 4), "bytes": by,
                "traffic": tr.get(f"ph_red_tile_{st.lower()}"),
                **({} if args.no_cpu else c3_errors(xh, st, pv.reduce2d_all())),
                "note": "one pass: sum and max in full, along slices(0) and slices(1); FP64 sums inside each tile, compensated FP32 pairs across tiles"}
            del xh
            del x, pv

    if want("c4"):
        # C4 heat 8192^2, r = 0.2, 16 Gaussian bumps (seed 3), FP32 and FP64
        n = 8192
        steps = args.heat_steps
        from phase.workloads import heat_initial_field
        u0 = heat_initial_field(n, seed=3)
        res = {}
        fields = {}
        # the same run with 1, 2 and 4 steps per pass for both precisions (one
        # step per pass is the reference's loop structure; 1 and 2 are
        # HBM-bound in both precisions), next to
        # the default schedule (12 steps per pass for FP32, 8 for FP64, where
        # both are limited by VALU: FP32 runs at FP64's rate without packed FP32
        # math)
        per_pass = {}
        for kk in ("1", "2", "4"):
            os.environ["PH_HEAT_K"] = kk
            same = {}
            for dt, npdt in [("F64", np.float64), ("F32", np.float32)]:
                u = phase.NArray.from_numpy(u0.astype(npdt))
                ev.sync()
                aa, bb = ev.new(), ev.new()
                ev.record(aa)
                phase.heat2d(u, 0.2, steps, every=100)
                ev.record(bb)
                same[dt] = round(n * n * steps / (ev.ms(aa, bb) * 1e-3) / 1e9, 2)
                del u
            per_pass[kk] = same
            del os.environ["PH_HEAT_K"]
        for dt, npdt in [("F64", np.float64), ("F32", np.float32)]:
            u = phase.NArray.from_numpy(u0.astype(npdt))
            ev.sync()
            aa, bb = ev.new(), ev.new()
            ev.record(aa)
            e = phase.heat2d(u, 0.2, steps, every=100)
            ev.record(bb)
            ms = ev.ms(aa, bb)
            res[dt] = {"ms_total": round(ms, 2), "gcells": round(n * n * steps / (ms * 1e-3) / 1e9, 2),
                       "energy_first_last": [e[0], e[-1]]}
            fields[dt] = u
        # relative L2 of the FP32 run against the FP64 run
        f32 = fields["F32"].to_numpy().astype(np.float64)
        f64 = fields["F64"].to_numpy()
        rel = float(np.sqrt(np.sum((f32 - f64) ** 2) / np.sum(f64 ** 2)))
        out["c4_heat_8192"] = {"steps": steps, "f64": res["F64"], "f32": res["F32"],
                               "f32_over_f64": round(res["F32"]["gcells"] / res["F64"]["gcells"], 3),
                               "four_steps_per_pass": {"f64_gcells": per_pass["4"]["F64"], "f32_gcells": per_pass["4"]["F32"],
                                                       "f32_over_f64": round(per_pass["4"]["F32"] / per_pass["4"]["F64"], 3)},
                               "two_steps_per_pass": {"f64_gcells": per_pass["2"]["F64"], "f32_gcells": per_pass["2"]["F32"],
                                                      "f32_over_f64": round(per_pass["2"]["F32"] / per_pass["2"]["F64"], 3)},
                               "one_step_per_pass": {"f64_gcells": per_pass["1"]["F64"], "f32_gcells": per_pass["1"]["F32"],
                                                     "f32_over_f64": round(per_pass["1"]["F32"] / per_pass["1"]["F64"], 3),
                                                     "frac_of_roofline_f64": round(per_pass["1"]["F64"] / 500.0, 3),
                                                     "frac_of_roofline_f32": round(per_pass["1"]["F32"] / 1000.0, 3)},
                               "rel_l2_f32_vs_f64": rel,
                               "roofline_gcells_f64": 500.0, "roofline_gcells_f32": 1000.0,
                               "traffic_per_launch_of_four_steps": {"f64": tr.get("k_heat_multi_f64"),
                                                                    "f32": tr.get("k_heat_multi_f32")}}
        del fields, u
        if not args.no_cpu:
            out["c4_heat_8192"]["cpu"] = cpu_c4(u0)
    if want("c5"):
        out["c5_batched_20000"] = batched_workload(phase, ev, args)

    if want("concat"):
        # NArray.concatenate of 4 x [4096, 4096] Float32 along axis 1 (one
        # strided copy per input into its slab of the [4096, 16384] result)
        parts = [phase.NArray.random([4096, 4096], D.F32, seed=40 + k) for k in range(4)]
        cat = lambda: phase.NArray.concatenate_all(*parts, axis=1)  # noqa: E731
        cat()
        ms = time_loop(ev, cat, 10, phase)
        by = 2 * 4 * 4096 * 4096 * 4
        out["concatenate_f32_4x4096x4096_axis1"] = {"ms": round(ms, 4), "bytes": by,
                                                    "achieved_gbs": round(by / ms / 1e6, 1),
                                                    "frac": round(by / ms / 1e6 / PEAK_HBM_GBS, 4)}
        if not args.no_cpu:
            O = _oracle()
            hp = [p.to_numpy() for p in parts]
            out["concatenate_f32_4x4096x4096_axis1"]["cpu"] = _cpu_leg(
                lambda: O.concatenate(hp, axis=1), by, "NArray.concatenate restated by the oracle", "full size")
            del hp
        del parts
    if want("json"):
        # JSON / YAML text of a 1024 x 1024 Float64 array (device -> host -> text
        # and back; host-bound)
        a = phase.NArray.random([1024, 1024], D.F64, seed=41, a=-1.0, b=1.0)
        t0 = time.perf_counter()
        j = a.to_json()
        t1 = time.perf_counter()
        phase.NArray.from_json(j, D.F64)
        t2 = time.perf_counter()
        out["json_f64_1024x1024"] = {"to_json_ms": round((t1 - t0) * 1e3, 1), "from_json_ms": round((t2 - t1) * 1e3, 1),
                                     "text_mb": round(len(j) / 1e6, 2), "note": "host formatting / parsing in consecutive chunks, up to 16 threads"}
    if want("path"):
        out["path_ops_i32_8192"] = path_ops(phase, ev, args)
    return out


def _cpu_leg(fn, by, what, size="full 8192^2"):
    """One timed call of the oracle's restatement (1 core); `by` are the
    algorithmic bytes of the sample."""
    t0 = time.perf_counter()
    fn()
    dt = time.perf_counter() - t0
    return {"value": round(by / dt / 1e9, 3), "unit": "GB/s", "cores": 1, "kind": "port",
            "sample": f"{what}, {size}, {dt * 1e3:.0f} ms"}


def path_ops(phase, ev, args):
    """The other operations of the path on Int32 8192 x 8192 arrays:
    transposed gather and scatter (permuted View / MutableView), mask
    assignment with a scalar and with a source array, elementwise equality,
    map_with_coord, a map over a transposed operand, and a recurrence whose
    source is a ProcView of the array it writes.  Bytes are algorithmic: each element read or written
    once, Bool masks and results 1 byte."""
    import numpy as np
    D = phase.DType
    n = 8192
    nn = n * n
    x = phase.NArray.random([n, n], D.I32, seed=1, dist="int31")
    y = phase.NArray.random([n, n], D.I32, seed=5, dist="int31")
    z = phase.NArray.random([n, n], D.I32, seed=6, dist="int31")
    # the mask is an input of the two mask assignments: computed once, here
    mask = x.process(lambda v: (v & 1) == 0).to_narr().materialize()
    res = {}

    def entry(name, kernel, fn, by, cpu=None):
        fn()
        ms = time_loop(ev, fn, 10, phase)
        e = {"kernel": kernel, "ms": round(ms, 4), "bytes": by, "achieved_gbs": round(by / ms / 1e6, 1),
             "frac": round(by / ms / 1e6 / PEAK_HBM_GBS, 4)}
        if cpu is not None and not args.no_cpu:
            e["cpu"] = cpu()
        res[name] = e

    O = _oracle() if not args.no_cpu else None
    xh = x.to_numpy() if O else None
    # the oracle's transposes walk the view element by element: its legs
    # take the first 1024 of the view's 8192 rows
    rows = 1024
    if O:
        from crystal import parse_literal
        r = O.region([parse_literal(f"0...{rows}"), parse_literal("..")], [n, n])
    entry("transpose_gather", "k_copy_tiled_v (LDS 128x128 tiles, 16-byte lanes)", lambda: x.view().permute(1, 0).to_narr(), 8 * nn,
          lambda: _cpu_leg(lambda: O.view_gather(xh, [("region", r), ("permute", [1, 0])], [rows, n]),
                           8 * rows * n, "view_gather of x.permute(1, 0)", f"its first {rows} rows"))
    zt = z.mutable_view().permute(1, 0)

    def scat():
        zt[[]] = y
    yh = y.to_numpy() if O else None
    entry("transpose_scatter", "k_copy_tiled_v (LDS 128x128 tiles, 16-byte lanes)", scat, 8 * nn,
          lambda: _cpu_leg(lambda: O.view_scatter(xh.copy(), [("permute", [1, 0])], r, yh[:rows]),
                           8 * rows * n, "view_scatter through z.mutable_view.permute(1, 0)",
                           f"its first {rows} rows"))
    mh = mask.to_numpy() if O else None

    def ms_scalar():
        z[mask] = 7
    entry("set_mask_scalar", "ph_map_kernel (select)", ms_scalar, 9 * nn,
          lambda: _cpu_leg(lambda: O.set_mask(xh.copy(), mh, 7), 9 * nn, "NArray#[]=(mask, 7)"))

    def ms_src():
        z[mask] = y
    entry("set_mask_source", "ph_map_kernel (select)", ms_src, 13 * nn,
          lambda: _cpu_leg(lambda: O.set_mask(xh.copy(), mh, yh), 13 * nn, "NArray#[]=(mask, y)"))
    entry("eq", "ph_map_kernel", lambda: x.eq(y).materialize(), 9 * nn,
          lambda: _cpu_leg(lambda: O.eval_expr("(eq x0 x1)", [xh, yh]), 9 * nn, "x.eq(y)"))

    def cpu_coord():
        from crystal import trace
        ci, cj = np.indices((n, n), dtype=np.int32)
        t0 = time.perf_counter()
        O.eval_expr(trace(lambda el, i, j: el.wadd(i).wsub(j), 3), [xh, ci, cj])
        dt = time.perf_counter() - t0
        return {"value": round(8 * nn / dt / 1e9, 3), "unit": "GB/s", "cores": 1, "kind": "port",
                "sample": f"map_with_coord (element and coordinate zipped), full 8192^2, {dt * 1e3:.0f} ms"}
    entry("map_with_coord", "ph_map_kernel (index instruction)",
          lambda: x.map_with_coord(lambda el, c: el.wadd(c[0]).wsub(c[1])).materialize(), 8 * nn, cpu_coord)
    xt = x.view().permute(1, 0)

    def cpu_tmap():
        from crystal import trace
        t0 = time.perf_counter()
        g = O.view_gather(xh, [("region", r), ("permute", [1, 0])], [rows, n])
        O.eval_expr(trace(lambda a, b: a.wadd(b), 2), [g, yh[:rows]])
        dt = time.perf_counter() - t0
        return {"value": round(12 * rows * n / dt / 1e9, 3), "unit": "GB/s", "cores": 1, "kind": "port",
                "sample": f"map_with over x.permute(1, 0) and y, its first {rows} rows, {dt * 1e3:.0f} ms"}
    entry("map_transposed", "ph_map_tiled (transposed operand staged through LDS)",
          lambda: xt.map_with(y, block=lambda a, b: a.wadd(b)).materialize(), 12 * nn, cpu_tmap)
    # an aliased source (a12): each row becomes the row above it plus one,
    # row 1 first, as ph-core's lazy loop gives; every column is a chain.
    # Row 0 is refilled first with a value that alternates from call to
    # call, or the array would already be the loop's result from the
    # second call on.
    zr = phase.NArray.random([n, n], D.I32, seed=7, a=-1000, b=1000)
    zrh = zr.to_numpy()[:rows].copy() if O else None
    turn = [0]

    def rec():
        turn[0] ^= 1
        zr[["0..0", ".."]] = turn[0]
        zr[["1..", ".."]] = zr.view("..-2", "..").process(lambda v: v + 1)

    def cpu_rec():
        from crystal import parse_literal, trace
        dst = O.region([parse_literal("1.."), parse_literal("..")], [rows, n])
        src = O.region([parse_literal("..-2"), parse_literal("..")], [rows, n])
        h = zrh.copy()
        t0 = time.perf_counter()
        O.set_chunk_lazy(h, [], dst, trace(lambda v: v + 1), [(h, [("region", src)])], src.shape)
        dt = time.perf_counter() - t0
        return {"value": round(8 * (rows - 1) * n / dt / 1e9, 3), "unit": "GB/s", "cores": 1, "kind": "port",
                "sample": f"set_chunk of a ProcView of the same array, its first {rows} rows, {dt * 1e3:.0f} ms"}
    entry("aliased_recurrence", "row 0 refilled; ordered map: classify, a walk per column, the write",
          rec, 8 * (n - 1) * n + 4 * n, cpu_rec)
    return res


def batched_workload(phase, ev, args):
    """20000 Float32 arrays (rank 1-3, sides log-uniform [8, 512], seed 4,
    arrays over 2^18 elements drawn again): per array y = x + s_k, x.sum and
    the step-2 slice copy, as one graph replay of the persistent kernel, as
    three launches per array, and on the CPU (the oracle's per-array loop)."""
    import numpy as np
    from phase.workloads import batch_stats, batched_shapes
    shapes, scalars, redraws = batched_shapes()
    st = batch_stats(shapes)
    ne, nz = st["elements"], st["slice_elements"]
    op_bytes = 12 * ne + 8 * nz        # add 8 B, sum 4 B per element; copy 8 B per slice element
    pass_bytes = 8 * ne + 4 * nz       # one fused pass: read x, write y and z
    b = phase.Batch(shapes, scalars.tolist())
    b.fill_uniform(4)
    chunk, grid, dyn = b.tune()
    b.run(0)
    ms_graph = time_loop(ev, lambda: b.run(0), 20, phase)
    ms_direct = time_loop(ev, lambda: b.run(1), 20, phase)
    ms_unbatched = wall_unbatched = None
    if not args.no_unbatched:
        b.run(2)
        ev.sync()
        t0 = time.perf_counter()
        ms_unbatched = time_loop(ev, lambda: b.run(2), 2)
        wall_unbatched = (time.perf_counter() - t0) / 2
    res = {
        "arrays": st["arrays"], "elements": ne, "slice_elements": nz, "redraws_over_cap": redraws,
        "cap_elements": 1 << 18, "log2_histogram": st["log2_histogram"],
        "schedule": {"chunk": chunk, "grid": grid, "queue": bool(dyn)},
        "graph_ms": round(ms_graph, 4), "direct_ms": round(ms_direct, 4),
        "unbatched_ms": round(ms_unbatched, 3) if ms_unbatched else None,
        "unbatched_wall_ms": round(wall_unbatched * 1e3, 3) if wall_unbatched else None,
        "speedup_vs_unbatched": round(ms_unbatched / ms_graph, 2) if ms_unbatched else None,
        "op_bytes": op_bytes, "op_gbs": round(op_bytes / ms_graph / 1e6, 1),
        "pass_bytes": pass_bytes, "pass_gbs": round(pass_bytes / ms_graph / 1e6, 1),
        "pass_frac": round(pass_bytes / ms_graph / 1e6 / PEAK_HBM_GBS, 4),
        "traffic": (load_traffic() or {}).get("k_batch"),
        "note": "op_bytes counts each op's own traffic (12 B/element + 8 B/slice element); the fused pass "
                "moves pass_bytes (8 B/element + 4 B/slice element), whose rate is the roofline figure",
    }
    # the same work as a general batch: per record a map program (v + s), a
    # reduction program (v, summed) and a region ("..2.." on every axis),
    # run by the kernel generated for those programs
    recs = [{"shape": list(sh), "map": 0, "red": 1, "ops": "sum", "region": ["..2.."] * len(sh),
             "scalar": float(scalars[k])} for k, sh in enumerate(shapes)]
    g = phase.GBatch(phase.DType.F32, phase.DType.F32, [lambda v, s: v + s, lambda v, s: v], recs)
    g.set_x(phase.NArray.random([g.x.size], phase.DType.F32, seed=4, a=-1.0, b=1.0))
    gchunk, ggrid, gdyn = g.tune()
    g.run(0)
    gms_graph = time_loop(ev, lambda: g.run(0), 20, phase)
    gms_direct = time_loop(ev, lambda: g.run(1), 20, phase)
    gms_unbatched = None
    if not args.no_unbatched:
        ev.sync()
        g.run(2)
        gms_unbatched = time_loop(ev, lambda: g.run(2), 2)
    res["general"] = {
        "kernel": "ph_gbatch (generated for the records' programs) + ph_gbatch_final",
        "schedule": {"chunk": gchunk, "grid": ggrid, "queue": bool(gdyn)},
        "graph_ms": round(gms_graph, 4), "direct_ms": round(gms_direct, 4),
        "unbatched_ms": round(gms_unbatched, 3) if gms_unbatched else None,
        "speedup_vs_unbatched": round(gms_unbatched / gms_graph, 2) if gms_unbatched else None,
        "pass_gbs": round(pass_bytes / gms_graph / 1e6, 1),
        "pass_frac": round(pass_bytes / gms_graph / 1e6 / PEAK_HBM_GBS, 4),
        "traffic": (load_traffic() or {}).get("ph_gbatch"),
        "note": "records carry their own programs and regions (ph_gbatch); unbatched launches the same "
                "kernel once per array",
    }
    del g
    if not args.no_cpu:
        sys.path.insert(0, os.path.join(ROOT, "tests"))
        import oracle_api as O
        x = b.x.to_numpy()
        xo, zo = np.array(b.offsets_x), np.array(b.offsets_z)
        k = 4000  # bounded sample: the first 4000 arrays
        sub = shapes[:k]
        end = int(xo[k])
        t0 = time.perf_counter()
        O.batch_loop(sub, scalars[:k], x[:end])
        dt = time.perf_counter() - t0
        sb = 12 * sum(int(np.prod(s)) for s in sub) + 8 * sum(int(np.prod([(d + 1) // 2 for d in s])) for s in sub)
        res["cpu"] = {"value": round(sb / dt / 1e9, 3), "unit": "GB/s (op bytes)", "cores": 1, "kind": "port",
                      "sample": f"first {k} arrays, oracle per-array loop (map +, Float32 sum, step-2 view "
                                f"gather), {dt * 1e3:.1f} ms",
                      "est_full_batch_ms": round(dt * op_bytes / sb * 1e3, 1)}
    del b
    return res


if __name__ == "__main__":
    main()
