"""Benchmark of the evolution hot path on one MI355X.

Metric: grid-point updates per second, FP64: the sum over levels and level
steps of Nrtotal*Nztotal divided by wall time, state resident in HBM.  A
"step" is one coarse Berger-Oliger step of the workload (all levels).

Default workload: the batched scalar-collapse family (64 scalar pulses on
grids of 64x128 to 256x512 with 2-4 levels, RK4, 1+log + Gamma-driver 2)
advanced as one batch; the same members advanced one at a time (the plain
path) are timed in the same run and reported beside it.  --workload brill
runs the Brill-wave configuration (5 box-in-box levels of 256x512, ICN) on
the plain path.  Synthetic data from the seeded par generator.

Prints one JSON line.  Only one GPU is supported: --gpus above 1 exits with
a message.
"""
import argparse
import json
import os
import sys
import time

ROOT = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, os.path.join(ROOT, "ollinaxis-bib_amd"))

METRIC = "grid-point updates/s (FP64 RK4) batched vs plain, 1 GPU; p50/p99 solve latency"
TUNE_STEPS = 16                # steps the schedule autotuner needs before it settles
HBM_PEAK_GBS = 8000.0          # MI355X HBM3E, GB/s
FP64_PEAK_TFLOPS = 78.6        # MI355X FP64 vector, TFLOP/s


def workload_text(name, seed):
    from oaxis import workloads
    if name == "brill":
        return workloads.brill(seed), "brill: BrillWave a~U[1,4] (seed %d), 5 levels of 256x512, ICN, " \
                                      "1+log + Gammadriver2, psi from the elliptic solve" % seed
    if name == "plain_small":
        return workloads.plain_small(seed), "plain_small: Schwarzschild M=1, 1 level 128x256, RK4, 1+log"
    if name == "plain_256x512":
        return workloads.plain_small(seed, 256, 512), "plain: Schwarzschild M=1, 1 level 256x512, RK4, 1+log"
    if name == "plain_512x1024":
        return workloads.plain_small(seed, 512, 1024), "plain: Schwarzschild M=1, 1 level 512x1024, RK4, 1+log"
    raise SystemExit("unknown workload %s" % name)


def bytes_per_update(params, nfields):
    """Algorithmic HBM bytes per grid-point update: RK4 136 N, ICN-3 64 N
    (reads of y, y_p, acc and writes of acc, y per substep, 8 B each)."""
    if params.integrator == 1:       # rk4
        return 136.0 * nfields, 4
    it = params.icniter
    return 8.0 * nfields * (2 + 3 * (it - 1)), it


def measured_traffic(workload, algorithmic_bytes):
    """HBM bytes per RHS launch for this workload: the copy-calibrated
    FETCH_SIZE + WRITE_SIZE per point measured for it (committed summary),
    relative to its algorithmic bytes, times the launch's algorithmic bytes;
    None when no measurement is committed."""
    path = os.path.join(ROOT, "profiles", "k_rhs_hbm_traffic.md")
    try:
        with open(path) as f:
            for line in f:
                parts = line.split()
                if len(parts) == 3 and parts[0] == "traffic_ratio" and parts[1] == workload:
                    return float(parts[2]) * algorithmic_bytes
    except OSError:
        pass
    return None


def fp64_roofline(workload, points_per_launch, avg_launch_ms):
    """FP64 rate of the RHS kernel: the counter-measured FP64 operations per
    point of one RHS evaluation for this workload (committed summary,
    SQ_INSTS_VALU_{ADD,MUL,FMA,TRANS}_F64) times the points one launch
    evaluates, over the launch's average duration; None without a
    measurement."""
    path = os.path.join(ROOT, "profiles", "k_rhs_fp64_valu.md")
    per_point = None
    try:
        with open(path) as f:
            for line in f:
                parts = line.split()
                if len(parts) == 3 and parts[0] == "flop_per_point" and parts[1] == workload:
                    per_point = float(parts[2])
    except OSError:
        pass
    if per_point is None or avg_launch_ms <= 0.0:
        return None
    tflops = per_point * points_per_launch / (avg_launch_ms * 1e-3) / 1e12
    return {"bound": "fp64", "achieved": tflops, "peak": FP64_PEAK_TFLOPS, "unit": "TFLOP/s",
            "frac": tflops / FP64_PEAK_TFLOPS, "flop_per_point": per_point}


def cpu_baseline(texts, min_seconds, threads=1, states=None):
    """The oracle (plain C restatement) on a bounded sample: coarse steps of
    the given configurations until min_seconds of wall time have passed.
    With threads > 1 each worker thread owns instances of its own (thread t
    takes texts[t % len(texts)]) and steps them independently; the oracle's
    calls release the GIL, so the threads run on separate cores, like
    independent runs of the reference, one per core.  states[k], when given,
    is the initial state of texts[k] ({(level, name): array}, e.g. the
    device's, so the oracle need not repeat the elliptic solve).  Returns
    (aggregate updates/s, seconds, coarse steps)."""
    import threading
    sys.path.insert(0, os.path.join(ROOT, "oracle"))
    import oaxis
    import pyoracle

    def make(k):
        orc = pyoracle.Oracle(oaxis.params_from_text(texts[k]))
        if states is not None:
            for (lev, name), a in states[k].items():
                orc.set(name, a, lev)
        else:
            orc.initial_data()
        orc.initial_aux()
        d = orc.dims(0)
        return orc, sum(2 ** l for l in range(orc.nlevels)) * d["Nr"] * d["Nz"]

    if threads <= 1:
        owned = [list(range(len(texts)))]
    else:
        owned = [[t % len(texts)] for t in range(threads)]
    orcs = [None] * len(owned)
    tally = [(0.0, 0)] * len(owned)
    span = [None] * len(owned)          # each worker's own start and end of stepping
    start = threading.Barrier(len(owned) + 1)

    errors = []

    def worker(t):
        try:
            orcs[t] = [make(k) for k in owned[t]]
        except Exception as e:          # let the other threads and the caller go on
            errors.append(e)
            start.abort()
            return
        try:
            start.wait()
        except threading.BrokenBarrierError:
            return
        t0 = time.perf_counter()
        updates, steps, k = 0.0, 0, 0
        while time.perf_counter() - t0 < min_seconds:
            orc, per_coarse = orcs[t][k % len(orcs[t])]
            orc.coarse_step()
            updates += per_coarse
            steps += 1
            k += 1
        tally[t] = (updates, steps)
        span[t] = (t0, time.perf_counter())

    pool = [threading.Thread(target=worker, args=(t,)) for t in range(len(owned))]
    for th in pool:
        th.start()
    try:
        start.wait()
    except threading.BrokenBarrierError:
        pass
    for th in pool:
        th.join()
    for mine in orcs:
        for orc, _ in mine or []:
            orc.close()
    if errors:
        raise errors[0]
    # from the first worker's start to the last one's end
    dt = max(e for _, e in span) - min(b for b, _ in span)
    return sum(u for u, _ in tally) / dt, dt, sum(n for _, n in tally)


def host_threads():
    """Worker threads for the all-core CPU baseline: the cores this process
    may run on, at most 16 (the host share of one GPU)."""
    try:
        n = len(os.sched_getaffinity(0))
    except AttributeError:
        n = os.cpu_count() or 1
    return max(1, min(16, n))


def cpu_baseline_record(texts, seconds, what, states=None):
    """cpu_baseline of the bench line: the oracle on all host cores (value),
    with the one-core figure beside it."""
    progress("CPU baseline: the oracle on one core, then on %d" % host_threads())
    v1, s1, n1 = cpu_baseline(texts, seconds, 1, states)
    nt = host_threads()
    vn, sn, nn = cpu_baseline(texts, seconds, nt, states)
    return {"value": vn, "unit": "grid-point updates/s", "cores": nt, "kind": "port",
            "sample": "%s on the oracle (C restatement), %d threads each stepping its own instance, "
                      "%d coarse steps in %.1f s (evolution only)" % (what, nt, nn, sn),
            "one_core": {"value": v1, "sample": "%d coarse steps, 1 thread, %.1f s" % (n1, s1)}}


def copy_probe(nbytes):
    """A plain device-to-device copy of nbytes, three times: the calibration
    kernel for the FETCH_SIZE / WRITE_SIZE counters.  Runs in a process of its
    own (torch's HIP runtime and the library's must not share one)."""
    import torch
    n = nbytes // 8
    x = torch.rand(n, dtype=torch.float64, device="cuda")
    y = torch.empty_like(x)
    for _ in range(3):
        y.copy_(x)
    torch.cuda.synchronize()
    print(json.dumps({"copy_probe_bytes": n * 8}))


def run_serving(args, n, steps_scale, max_grid, plain=False):
    """The serving mix through the continuous-batching loop: Poisson
    arrivals, admission by device footprint (each request's initial data and
    elliptic solve prepared on a worker thread while the batch steps),
    retirement at the final step, on horizon formation (searched every 16
    coarse steps) or on dispersal.  Returns the loop's summary and, with
    plain, the summary of the same requests served one spacetime at a time."""
    from oaxis import serve
    grids = tuple(g for g in ((64, 128), (128, 256), (256, 512), (512, 1024)) if g[0] <= max_grid)
    lo, hi = max(1, int(200 * steps_scale)), max(1, int(2000 * steps_scale))
    reqs = serve.serving_mix(args.seed, n, rate=args.serving_rate, grids=grids, steps=(lo, hi))
    srv = serve.Server(memory_budget=args.serving_budget_gb * 1e9, graphs=not args.no_graphs)

    def retired(rec):
        progress("  request %d (%s) retired: %s after %d steps, latency %.2f s"
                 % (rec.rid, rec.kind, rec.reason, rec.steps, rec.latency))

    records, summary = srv.run(reqs, progress=retired)
    summary["sample"] = ("%d requests of the serving mix (seed %d, Poisson rate %.3g/s, grids up to %dx%d, "
                         "1-6 levels, %d-%d coarse steps (the mix's 200-2000 scaled by %g), horizon search "
                         "every 16), memory budget %.0f GB"
                         % (n, args.seed, args.serving_rate, max_grid, 2 * max_grid, lo, hi, steps_scale,
                            args.serving_budget_gb))
    if not plain:
        return summary
    progress("serving mix: %.3g updates/s batched; the same requests one at a time" % summary["updates_per_s"])
    _, plain_summary = srv.run_plain(reqs, progress=retired)
    return summary, plain_summary


def horizon_probe(with_oracle):
    """One apparent-horizon search (shooting method) on Schwarzschild M=2,
    2 levels of 128x256: device time of oa_ahfind (second of two calls) and,
    with the CPU baseline, the oracle's search of the same state on one core
    (its time, and whether it took the same shots); then the spectral
    method's first and second search on that state, likewise."""
    import oaxis
    from oaxis import workloads
    text = workloads.schwarzschild(2.0, 128, 256, 1)
    p = oaxis.params_from_text(text)
    out = {"workload": "Schwarzschild M=2, 2 levels of 128x256, shooting method"}
    with oaxis.Spacetime(p) as st:
        st.initial_data()
        st.initial_aux()
        st.ahfind(0)
        t0 = time.perf_counter()
        r = st.ahfind(0)
        out.update(device_ms=(time.perf_counter() - t0) * 1e3, found=r["found"], shots=r["shots"],
                   mass=r["mass"])
        if with_oracle:
            sys.path.insert(0, os.path.join(ROOT, "oracle"))
            import pyoracle
            orc = pyoracle.Oracle(p)
            orc.initial_data()
            orc.initial_aux()
            t0 = time.perf_counter()
            ro = orc.ahfind(0)
            out["cpu_baseline"] = {"oracle_ms": (time.perf_counter() - t0) * 1e3, "cores": 1,
                                   "matches": bool(ro["shots"] == r["shots"] and ro["zinit"] == r["zinit"]
                                                   and ro["found"] == r["found"])}
            orc.close()
    # the spectral finder on the same state: the first search flows from a
    # large circle, the next one starts from the horizon it found
    ps = oaxis.params_from_text(text + "ahmethod = spectral\n")
    sp = {"workload": "the same state, spectral method (ahNmodes 10, ahepsilon_spectral 1e-6)"}
    with oaxis.Spacetime(ps) as st:
        st.initial_data()
        st.initial_aux()
        t0 = time.perf_counter()
        r = st.ahfind(0)
        sp.update(device_ms=(time.perf_counter() - t0) * 1e3, found=r["found"], iterations=r["iterations"],
                  mass=r["mass"])
        t0 = time.perf_counter()
        r2 = st.ahfind(0)
        sp.update(again_ms=(time.perf_counter() - t0) * 1e3, again_iterations=r2["iterations"])
        if with_oracle:
            import pyoracle
            orc = pyoracle.Oracle(ps)
            orc.initial_data()
            orc.initial_aux()
            t0 = time.perf_counter()
            ro = orc.ahfind(0)
            sp["cpu_baseline"] = {"oracle_ms": (time.perf_counter() - t0) * 1e3, "cores": 1,
                                  "matches": bool(ro["found"] == r["found"]
                                                  and abs(ro["iterations"] - r["iterations"]) <= 1
                                                  and abs(ro["mass"] - r["mass"]) <= 1e-9 * max(abs(ro["mass"]), 1.0))}
            orc.close()
    out["spectral"] = sp
    return out


def run_bisection(args):
    """Critical-collapse amplitude bisection through the serving loop, reduced
    to a size that runs in about a minute: central scalar pulses (s = 1,
    Gamma-driver, 2 levels of 32x32 with equatorial symmetry), A in [0.2, 0.8],
    --members amplitudes per round, horizon search every 2 coarse steps."""
    from oaxis.bisection import critical_bisection
    from oaxis import workloads
    res = critical_bisection(lambda a: workloads.collapse_pulse(a), 0.2, 0.8, members=args.members,
                             rounds=max(1, args.steps), steps=48, ahfind_every=2)
    print(json.dumps({"metric": METRIC, "value": res["point_updates"] / res["wall_s"], "unit": "grid-point updates/s",
                      "n_gpus": 1, "higher_is_better": True, "dtype": "f64",
                      "data": "synthetic (central scalar pulses; psi from the elliptic solve)",
                      "config": {"workload": "critical-collapse bisection (reduced): %d amplitudes per round, "
                                             "%d round(s), 2 levels of 32x32, eqsym" % (args.members, len(res["rounds"])),
                                 "bracket": [res["lo"], res["hi"]], "wall_s": res["wall_s"],
                                 "rounds": [{k: r[k] for k in ("amplitudes", "horizon", "steps", "wall_s", "p50_s",
                                                                "p99_s")} for r in res["rounds"]]}}))


def run_bisection_full(args):
    """The critical-collapse bisection at its stated size: --members (32)
    concurrent shell pulses (r0 = 5, s = 1) with A spread over [0.01, 0.1],
    each on 7 levels of 256x512 (dr0 = 0.1, no equatorial symmetry), 1+log +
    Gamma-driver, RK4, horizon search on the device every 16 coarse steps,
    --bisection-rounds rounds of --bisection-steps coarse steps each."""
    from oaxis.bisection import critical_bisection
    from oaxis import workloads
    members = args.members if args.members_given else 32
    res = critical_bisection(lambda a: workloads.bisection_pulse(a), 0.01, 0.1, members=members,
                             rounds=args.bisection_rounds, steps=args.bisection_steps, ahfind_every=16,
                             progress=progress)
    print(json.dumps({"metric": METRIC, "value": res["point_updates"] / res["wall_s"], "unit": "grid-point updates/s",
                      "n_gpus": 1, "higher_is_better": True, "dtype": "f64",
                      "data": "synthetic (shell scalar pulses r0=5, s=1; psi from the elliptic solve)",
                      "config": {"workload": "critical-collapse bisection: %d amplitudes per round in [0.01, 0.1], "
                                             "%d round(s) of up to %d coarse steps, 7 levels of 256x512, horizon "
                                             "search every 16 coarse steps"
                                             % (members, len(res["rounds"]), args.bisection_steps),
                                 "bracket": [res["lo"], res["hi"]], "wall_s": res["wall_s"],
                                 "point_updates": res["point_updates"],
                                 "rounds": [{k: r[k] for k in ("amplitudes", "horizon", "reasons", "steps", "wall_s",
                                                                "first_step_s", "p50_s", "p99_s")}
                                            for r in res["rounds"]],
                                 "evolution_updates_per_s": res["point_updates"] / sum(
                                     r["wall_s"] - (r["first_step_s"] or 0.0) for r in res["rounds"])}}))


def prepare_spacetimes(texts, threads=8, graphs=True):
    """Spacetimes with their initial data (and its elliptic solve) and the
    initial auxiliary pass, prepared on worker threads: each runs on its own
    stream, so the many small kernels of the solves overlap on the device.
    graphs=False launches everything kernel by kernel, the solve included."""
    import oaxis
    from concurrent.futures import ThreadPoolExecutor

    def make(text):
        st = oaxis.Spacetime(oaxis.params_from_text(text))
        if not graphs:
            st.set_graphs(False)
        st.initial_data()
        st.initial_aux()
        return st

    with ThreadPoolExecutor(max_workers=max(1, threads)) as pool:
        return list(pool.map(make, texts))


def initial_spacetimes(texts, args):
    """The prepared spacetimes of a workload.  With --initial-states DIR they
    are restored from the checkpoints of an earlier run in DIR (one file per
    member, same parameters), or prepared and then saved there: a profiled
    run can then start from the same states without running the initial
    data and its elliptic solve itself."""
    import oaxis
    d = args.initial_states
    paths = [os.path.join(d, "member_%03d.ckpt" % k) for k in range(len(texts))] if d else []
    if paths and all(os.path.exists(p) for p in paths):
        sts = [oaxis.load_checkpoint(p) for p in paths]
        for k, (st, text) in enumerate(zip(sts, texts)):
            if bytes(st.params) != bytes(oaxis.params_from_text(text)):
                raise SystemExit("%s holds another workload's member %d" % (d, k))
            if args.no_graphs:
                st.set_graphs(False)
        return sts, "restored from %s" % d
    sts = prepare_spacetimes(texts, threads=args.prepare_threads, graphs=not args.no_graphs)
    if paths:
        os.makedirs(d, exist_ok=True)
        for p, st in zip(paths, sts):
            oaxis.save_checkpoint(st, p)
    return sts, "prepared (initial data with the elliptic solve)"


def progress(msg):
    print("bench: %s" % msg, file=sys.stderr, flush=True)


def run_batched(args):
    """The batched scalar-collapse workload: members of the seeded family
    (grids 64x128..256x512, 2-4 levels, RK4, 1+log + Gamma-driver 2, real
    scalar field) advanced together through one batch, then the same members
    one at a time (the plain path) for the same number of coarse steps."""
    import oaxis
    from oaxis import workloads
    texts = [workloads.scalar_pulse(args.seed, i) for i in range(args.members)]
    t_prep = time.perf_counter()
    sts, how = initial_spacetimes(texts, args)
    t_prep = time.perf_counter() - t_prep
    progress("%d members %s in %.1f s" % (len(sts), how, t_prep))
    # the initial states of the CPU baseline's sample, from the device
    cpu_states = [{(lev, name): st.download(name, lev) for lev in range(st.nlevels)
                   for name in oaxis.state_fields(st.params)} for st in sts[:8]]
    nfields = workloads.evolved_count(True, True)
    b = oaxis.Batch(sts)
    if args.no_graphs:
        b.set_graphs(False)
    # untimed: the schedule autotuner's steps (it times each schedule over
    # steady-state steps), then the requested warm-up
    b.evolve(TUNE_STEPS)
    if args.warmup > 0:
        b.evolve(args.warmup)
    b.sync()
    u0 = b.point_updates()
    t0 = time.perf_counter()
    b.evolve(args.steps)
    b.sync()
    wall = time.perf_counter() - t0
    updates = b.point_updates() - u0
    launches, items = b.launch_stats()
    device_ms = b.last_elapsed_ms()
    schedule = b.schedule()
    # RHS launch durations over a second timed pass, launched one by one
    b.set_kernel_timing(True)
    b.evolve(args.steps)
    b.sync()
    rhs_ms, rhs_n, rhs_points = b.rhs_kernel_stats()
    b.close()
    # the plain path: every member alone, one after the other
    plain_updates = 0.0
    plain_wall = 0.0
    progress("batched: %.3g updates/s; timing the plain path" % (updates / wall))
    for st in sts:
        if args.no_graphs:
            st.set_graphs(False)
        st.evolve(TUNE_STEPS)          # untimed: the member's own schedule autotuning
        if args.warmup > 0:
            st.evolve(args.warmup)
        p0 = st.point_updates()
        st.sync()
        t1 = time.perf_counter()
        st.evolve(args.steps)
        st.sync()
        plain_wall += time.perf_counter() - t1
        plain_updates += st.point_updates() - p0
    levels = [st.nlevels for st in sts]
    grids = sorted(set((st.params.Nrtotal, st.params.Nztotal) for st in sts))
    for st in sts:
        st.close()
    value = updates / wall
    plain_value = plain_updates / plain_wall
    bpu = 136.0 * nfields                      # RK4
    rhs_bytes = bpu / 4.0 * rhs_points         # over all timed RHS launches
    achieved = rhs_bytes / (rhs_ms * 1e-3) / 1e9 if rhs_ms > 0 else 0.0
    rec = {
        "metric": METRIC,
        "value": value,
        "unit": "grid-point updates/s",
        "n_gpus": 1,
        "steps": args.steps,
        "warmup": args.warmup,
        "ms_per_step": wall / args.steps * 1e3,
        "higher_is_better": True,
        "scaling": "weak",
        "vs_baseline": None,
        "dtype": "f64",
        "data": "synthetic (seeded par generator, seed %d; psi from the elliptic solve)" % args.seed,
        "config": {"workload": "batched scalar collapse: %d scalar pulses A~U[.01,.05], r0~U[3,6], s~U[.5,1.5] "
                               "(seed %d), grids %s, 2-4 levels, RK4, 1+log + Gammadriver2"
                               % (args.members, args.seed, " ".join("%dx%d" % g for g in grids)),
                   "tuning_steps": TUNE_STEPS, "prepare_s": t_prep,
                   "path": "batched (one launch per kernel kind over all members)",
                   "members": args.members, "levels_per_member": levels, "evolved_fields": nfields,
                   "integrator": "rk4", "point_updates": updates, "device_ms": device_ms,
                   "launches_per_coarse_step": launches, "work_items_per_coarse_step": items,
                   "plain_value": plain_value, "plain_ms_per_step": plain_wall / args.steps * 1e3,
                   "batched_vs_plain": value / plain_value, "schedule": schedule},
        "roofline": {"bound": "hbm", "achieved": achieved, "peak": HBM_PEAK_GBS, "unit": "GB/s",
                     "frac": achieved / HBM_PEAK_GBS,
                     "traffic": measured_traffic(args.workload, rhs_bytes / rhs_n if rhs_n else 0.0),
                     "kernel": "k_rhs_table", "avg_launch_ms": rhs_ms / rhs_n if rhs_n else 0.0,
                     "launches": rhs_n,
                     "algorithmic_bytes_per_launch": rhs_bytes / rhs_n if rhs_n else 0.0},
        "fp64": fp64_roofline(args.workload, rhs_points / rhs_n if rhs_n else 0.0,
                              rhs_ms / rhs_n if rhs_n else 0.0),
    }
    progress("plain: %.3g updates/s; horizon probe" % plain_value)
    rec["horizon"] = horizon_probe(not args.no_cpu_baseline)
    if args.serving_requests > 0:
        progress("serving mix: %d requests" % args.serving_requests)
        sv, pl = run_serving(args, args.serving_requests, args.serving_steps_scale, 512, plain=True)
        rec["latency"] = {"p50_s": sv["p50_s"], "p99_s": sv["p99_s"], "requests": sv["requests"],
                          "updates_per_s": sv["updates_per_s"], "wall_s": sv["wall_s"],
                          "retired": sv["reasons"], "sample": sv["sample"],
                          "plain_updates_per_s": pl["updates_per_s"], "plain_wall_s": pl["wall_s"],
                          "batched_vs_plain": sv["updates_per_s"] / pl["updates_per_s"]}
    if not args.no_cpu_baseline:
        sample = texts[:8]
        rec["cpu_baseline"] = cpu_baseline_record(
            sample, args.cpu_seconds, "members 0-%d of the same workload (their initial states taken from the "
            "device)" % (len(sample) - 1), cpu_states)
    print(json.dumps(rec))


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=8, help="timed coarse steps")
    ap.add_argument("--warmup", type=int, default=2, help="untimed coarse steps")
    ap.add_argument("--workload", default="batched_scalar",
                    help="batched_scalar (default), serving, brill, plain_small, plain_256x512, plain_512x1024, "
                         "horizon (one apparent-horizon search only), bisection (reduced critical-collapse "
                         "bisection; --members per round, --steps rounds), bisection_full (32 members on 7 "
                         "levels of 256x512; --bisection-rounds, --bisection-steps)")
    ap.add_argument("--members", type=int, default=64, help="spacetimes of the batched workload")
    ap.add_argument("--seed", type=int, default=0)
    ap.add_argument("--cpu-seconds", type=float, default=10.0, help="CPU time of the baseline sample")
    ap.add_argument("--no-cpu-baseline", action="store_true")
    ap.add_argument("--serving-requests", type=int, default=24,
                    help="requests of the serving sample reported as latency (0: none)")
    ap.add_argument("--serving-steps-scale", type=float, default=0.1,
                    help="scale of the 200-2000 coarse steps per request in the serving sample")
    ap.add_argument("--serving-rate", type=float, default=20.0, help="Poisson arrival rate (1/s)")
    ap.add_argument("--serving-budget-gb", type=float, default=200.0, help="device memory for admitted requests")
    ap.add_argument("--no-graphs", action="store_true",
                    help="launch coarse steps and elliptic solves kernel by kernel")
    ap.add_argument("--prepare-threads", type=int, default=8,
                    help="host threads preparing the batched members' initial data")
    ap.add_argument("--initial-states", default="", metavar="DIR",
                    help="restore the prepared members from checkpoints in DIR, or save them there")
    ap.add_argument("--copy-probe", type=int, default=0, metavar="BYTES",
                    help="only run a plain device copy of BYTES (counter calibration)")
    ap.add_argument("--bisection-rounds", type=int, default=1, help="rounds of the full-size bisection")
    ap.add_argument("--bisection-steps", type=int, default=320,
                    help="coarse steps per evolution of the full-size bisection")
    args = ap.parse_args()
    args.members_given = any(a == "--members" or a.startswith("--members=") for a in sys.argv[1:])
    if args.copy_probe:
        copy_probe(args.copy_probe)
        return

    if args.gpus != 1 or int(os.environ.get("WORLD_SIZE", "1")) > 1:
        print("bench.py: the library runs on one GPU only (every level of a spacetime stays resident on "
              "one device; batched instances shard by HBM footprint); --gpus %d is not supported" % args.gpus,
              file=sys.stderr)
        sys.exit(2)

    if args.workload == "batched_scalar":
        run_batched(args)
        return
    if args.workload == "bisection":
        run_bisection(args)
        return
    if args.workload == "bisection_full":
        run_bisection_full(args)
        return
    if args.workload == "horizon":
        # the apparent-horizon search alone (for profiling the finder)
        print(json.dumps({"horizon": horizon_probe(not args.no_cpu_baseline)}))
        return
    if args.workload == "serving":
        # the full serving mix (grids up to 512x1024, 200-2000 steps scaled by
        # --serving-steps-scale), then the same requests on the plain path
        sv, pl = run_serving(args, args.serving_requests, args.serving_steps_scale, 512, plain=True)
        print(json.dumps({"metric": METRIC, "value": sv["updates_per_s"], "unit": "grid-point updates/s",
                          "n_gpus": 1, "higher_is_better": True, "dtype": "f64",
                          "data": "synthetic (seeded par generator, seed %d; psi from the elliptic solve)" % args.seed,
                          "config": {"workload": "serving mix", "sample": sv["sample"],
                                     "plain_value": pl["updates_per_s"], "plain_wall_s": pl["wall_s"],
                                     "batched_vs_plain": sv["updates_per_s"] / pl["updates_per_s"]},
                          "latency": {k: sv[k] for k in ("p50_s", "p99_s", "requests", "wall_s", "reasons")}}))
        return

    import oaxis
    from oaxis import workloads
    text, desc = workload_text(args.workload, args.seed)
    params = oaxis.params_from_text(text)
    shift = params.shift != oaxis.ENUMS["OA_SHIFT_NONE"]
    scalar = bool(params.mattertype & oaxis.ENUMS["OA_MATTER_SCALAR"])
    nfields = workloads.evolved_count(shift, scalar)

    (st,), _ = initial_spacetimes([text], args)
    cpu_state = {(lev, name): st.download(name, lev) for lev in range(st.nlevels)
                 for name in oaxis.state_fields(params)}
    st.evolve(TUNE_STEPS)              # untimed: the schedule autotuner's steps
    if args.warmup > 0:
        st.evolve(args.warmup)
    u0 = st.point_updates()
    st.sync()
    t0 = time.perf_counter()
    st.evolve(args.steps)
    st.sync()
    wall = time.perf_counter() - t0
    updates = st.point_updates() - u0
    device_ms = st.last_elapsed_ms()
    value = updates / wall
    schedule = st.schedule()
    # the RHS kernel's launch durations: HIP events around each launch over a
    # second timed pass of the same number of steps (launched one by one,
    # since events cannot bracket kernels inside a graph)
    st.set_kernel_timing(True)
    st.evolve(args.steps)
    st.sync()
    rhs_ms, rhs_n = st.rhs_kernel_stats()
    st.set_kernel_timing(False)

    bpu, niter = bytes_per_update(params, nfields)
    # algorithmic bytes per RHS launch: per-update bytes over the substeps of a
    # step, times the points one launch advances (all levels have the same size)
    points_per_launch = float(params.Nrtotal) * float(st.level_info(0)["Nz"])
    bytes_per_launch = bpu / niter * points_per_launch
    achieved = bytes_per_launch / (rhs_ms * 1e-3) / 1e9 if rhs_ms > 0 else 0.0

    rec = {
        "metric": METRIC,
        "value": value,
        "unit": "grid-point updates/s",
        "n_gpus": 1,
        "steps": args.steps,
        "warmup": args.warmup,
        "ms_per_step": wall / args.steps * 1e3,
        "higher_is_better": True,
        "scaling": "weak",
        "vs_baseline": None,
        "dtype": "f64",
        "data": "synthetic (seeded par generator, seed %d)" % args.seed,
        "config": {"workload": desc, "path": "plain (one spacetime)", "levels": st.nlevels, "tuning_steps": TUNE_STEPS,
                   "Nrtotal": params.Nrtotal, "Nztotal": params.Nztotal, "evolved_fields": nfields,
                   "integrator": "rk4" if params.integrator == 1 else "icn",
                   "point_updates": updates, "device_ms": device_ms, "schedule": schedule},
        "roofline": {"bound": "hbm", "achieved": achieved, "peak": HBM_PEAK_GBS, "unit": "GB/s",
                     "frac": achieved / HBM_PEAK_GBS, "traffic": measured_traffic(args.workload, bytes_per_launch),
                     "kernel": "k_rhs_table", "avg_launch_ms": rhs_ms, "launches": rhs_n,
                     "algorithmic_bytes_per_launch": bytes_per_launch},
        "fp64": fp64_roofline(args.workload, points_per_launch, rhs_ms),
    }
    st.close()
    if not args.no_cpu_baseline:
        rec["cpu_baseline"] = cpu_baseline_record([text], args.cpu_seconds, "the same workload (initial state "
                                                  "taken from the device)", [cpu_state])
    print(json.dumps(rec))


if __name__ == "__main__":
    main()
