"""Critical-collapse amplitude bisection through the serving loop.

The workload behind parameter scans of scalar-field collapse: a family of
initial data indexed by an amplitude A is either subcritical (the field
disperses, or nothing forms before the final time) or supercritical (an
apparent horizon forms).  Each round evolves ``members`` amplitudes spread
over the current bracket [lo, hi] as concurrent requests of one serving loop
(one batch: shared launches, horizon searches in lockstep), then keeps the
adjacent pair where the outcome changes from "no horizon" to "horizon".  The
bracket shrinks by a factor members - 1 per round.  The reference has no such
driver; its runs are one amplitude each.
"""
import time

import numpy as np


def critical_bisection(text_of, lo, hi, members=8, rounds=3, steps=200, ahfind_every=16,
                       memory_budget=200e9, tol=0.0, progress=None, prepare_threads=8):
    """Bisect the amplitude A in [lo, hi]: ``text_of(A)`` gives the par-file
    text of one evolution (it should set mattertype, the grid and the gauge;
    horizon searches every ``ahfind_every`` coarse steps are added here).
    progress(text), when given, hears of every ahfind_every coarse steps of a
    round and of its end.  Returns dict(lo, hi, rounds=[per-round records],
    wall_s, point_updates)."""
    from .serve import Request, Server
    history = []
    t0 = time.perf_counter()
    total_updates = 0.0
    for rnd in range(rounds):
        amps = np.linspace(lo, hi, members)
        reqs = [Request(q, text_of(float(a)) + "\nahfind = .true.\nahfind_every = %d\n" % ahfind_every, steps,
                        0.0, "A=%.10g" % a) for q, a in enumerate(amps)]
        step_cb = admit_cb = None
        if progress is not None:
            def step_cb(k, n, rnd=rnd):
                progress("round %d: coarse step %d, %d evolutions running" % (rnd, k, n))

            def admit_cb(req, n, rnd=rnd):
                progress("round %d: %s prepared (%d of %d)" % (rnd, req.kind, n, members))
        # every amplitude of a round starts together and the round advances
        # them in lockstep
        server = Server(memory_budget, check_every=ahfind_every, prepare_threads=prepare_threads, gang=True)
        records, summary = server.run(reqs, step_progress=step_cb, step_every=ahfind_every, admit_progress=admit_cb)
        outcome = [r.reason == "horizon" for r in records]
        total_updates += summary["point_updates"]
        if progress is not None:
            progress("round %d done in %.1f s: %d of %d formed a horizon"
                     % (rnd, summary["wall_s"], sum(r.reason == "horizon" for r in records), len(records)))
        history.append(dict(amplitudes=amps.tolist(), horizon=outcome,
                            steps=[r.steps for r in records], reasons=[r.reason for r in records],
                            masses=[r.horizon.get("mass", 0.0) if r.horizon else 0.0 for r in records],
                            wall_s=summary["wall_s"], p50_s=summary["p50_s"], p99_s=summary["p99_s"],
                            first_step_s=summary["first_step_s"]))
        # the first amplitude that forms a horizon, with every smaller one not
        # forming one, brackets the threshold
        try:
            k = outcome.index(True)
        except ValueError:
            break                        # no horizon anywhere: the threshold is above hi
        if k == 0:
            break                        # a horizon everywhere: the threshold is below lo
        lo, hi = float(amps[k - 1]), float(amps[k])
        if hi - lo <= tol:
            break
    return dict(lo=lo, hi=hi, rounds=history, wall_s=time.perf_counter() - t0, point_updates=total_updates)
