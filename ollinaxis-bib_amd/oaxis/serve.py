"""Serving loop: continuous batching of evolution requests on one device.

Requests (par-file text, number of coarse steps, arrival time) are admitted
as they arrive while their device footprint fits the memory budget and a
member slot is free.  Every admitted spacetime keeps its state on the device
and is advanced inside one Batch, so requests that arrive while others run
join the same launches.  A request retires when it has done its coarse
steps (the reference's Nt), when an apparent horizon forms, or when its
scalar field has dispersed (max |scalar_phi| on level 0 below
``disperse_frac`` of its initial value).

Horizons: a request with ``ahfind = .true.`` is searched as the reference
does (horizon_finder at t = 0 and every ``ahfind_every`` coarse steps once
t >= ``ahafter``, evolve.f90:251 and :429-432), all due searches of one step
in one lockstep batch (oa_batch_ahfind).  It retires ("horizon") when a
horizon is found that was not there at t = 0 (a puncture's horizon exists
from the start and does not end the request).  A request without ahfind
retires when its lapse collapses instead (checked every ``check_every``
coarse steps: min alpha on the finest level below ``alpha_collapse``, for
requests whose lapse starts above it).  The reference has no serving loop;
the retirement criteria other than the step count are defined here.

Admission prepares a request (device allocation, initial data with its
elliptic solve, the initial auxiliary pass) on a host worker thread and the
request's own stream, while the batch keeps stepping on its stream; the
prepared spacetime joins the batch at the next step boundary.

Completion latency is retirement time minus arrival time; run() returns one
record per request and the aggregate throughput.  run_plain() serves the
same requests one spacetime at a time, for comparison.
"""
import time
from concurrent.futures import ThreadPoolExecutor
from dataclasses import dataclass, field

import numpy as np


@dataclass
class Request:
    rid: int
    text: str                 # par-file text (reference keys)
    steps: int                # coarse steps to run (Nt)
    arrival: float = 0.0      # seconds after the start of run()
    kind: str = ""


@dataclass
class Record:
    rid: int
    kind: str
    arrival: float
    admitted: float = -1.0
    retired: float = -1.0
    reason: str = ""
    steps: int = 0
    point_updates: float = 0.0
    footprint: float = 0.0
    norms: dict = field(default_factory=dict)
    horizon: dict = field(default_factory=dict)   # the oa_ahfind result that retired it

    @property
    def latency(self):
        return self.retired - self.arrival


class Server:
    def __init__(self, memory_budget, max_members=256, check_every=16, alpha_collapse=0.05,
                 disperse_frac=0.01, on_retire=None, graphs=True, prepare_threads=4, gang=False):
        import oaxis
        self.oaxis = oaxis
        # gang: requests that have arrived join together, once all of them are
        # prepared (a parameter scan advances its members in lockstep instead
        # of admitting each as soon as its initial data is ready)
        self.gang = bool(gang)
        self.prepare_threads = max(1, int(prepare_threads))
        self.budget = float(memory_budget)
        self.max_members = int(max_members)
        self.check_every = int(check_every)
        self.alpha_collapse = alpha_collapse
        self.disperse_frac = disperse_frac
        self.on_retire = on_retire       # callback(request, spacetime, record) before the state is freed
        self.graphs = bool(graphs)       # coarse steps as captured graphs (False: kernel by kernel)

    def _prepare(self, req):
        """Device state of a request, ready to join the batch (runs on a
        worker thread: the library calls release the interpreter lock and use
        the spacetime's own stream)."""
        oaxis = self.oaxis
        p = oaxis.params_from_text(req.text)
        st = oaxis.Spacetime(p)
        try:
            if not self.graphs:
                st.set_graphs(False)      # the elliptic solve kernel by kernel too
            st.initial_data()
            st.initial_aux()
        except Exception:
            st.close()
            raise
        return p, st

    def _admit(self, req, now, batch, active, used, prepared=None):
        oaxis = self.oaxis
        if prepared is None:
            prepared = self._prepare(req)
        p, st = prepared
        fp = oaxis.footprint(p)
        rec = Record(req.rid, req.kind, req.arrival, admitted=now, footprint=fp)
        scalar = bool(p.mattertype & oaxis.ENUMS["OA_MATTER_SCALAR"])
        phi0 = st.norms0d("scalar_phi", 0)["nm1"] if scalar else 0.0
        ah = bool(p.ahfind)
        if ah:
            # the search at t = 0 (evolve.f90:251) runs with the other new
            # members' before the next step (see run); only a horizon that
            # forms later retires the request
            watch = True
        else:
            # only a collapse that forms during the evolution retires a
            # request (a puncture's lapse starts near zero)
            watch = st.norms0d("alpha", st.nlevels - 1)["min"] >= self.alpha_collapse
        batch.add(st)
        active.append(dict(req=req, st=st, rec=rec, steps=0, scalar=scalar, phi0=phi0, watch=watch, ah=ah,
                           ah_t0=ah and p.ahafter <= 0.0,
                           ah_every=max(1, int(p.ahfind_every)), ah_after=float(p.ahafter),
                           dt0=st.level_info(0)["dt"], u0=st.point_updates()))
        return fp

    def _initial_searches(self, batch, active):
        """The t = 0 horizon searches of the members admitted since the last
        step, as one lockstep batch: a horizon present at t = 0 does not
        retire its request."""
        new = [a for a in active if a.get("ah_t0")]
        if not new:
            return []
        failed = []
        try:
            found = batch.ahfind([(a["st"], 0) for a in new])
        except self.oaxis.OaxisError:
            found = None
        for q, a in enumerate(new):
            a["ah_t0"] = False
            if found is None:
                failed.append(a)
            elif found[q]["found"]:
                a["watch"] = False
        return failed

    def _retire(self, a, reason, now, batch):
        st, rec = a["st"], a["rec"]
        batch.remove(st)
        rec.retired = now
        rec.reason = reason
        rec.steps = a["steps"]
        rec.point_updates = st.point_updates() - a["u0"]
        if reason != "failed":
            try:
                rec.norms = {"alpha": st.norms0d("alpha", 0), "chi": st.norms0d("chi", 0)}
            except self.oaxis.OaxisError:
                rec.reason = "failed"
        if self.on_retire is not None:
            self.on_retire(a["req"], st, rec)
        st.close()
        return rec

    def _check(self, a):
        st = a["st"]
        lev = st.nlevels - 1
        try:
            if a["watch"] and not a["ah"] and st.norms0d("alpha", lev)["min"] < self.alpha_collapse:
                return "lapse_collapse"
            if a["scalar"] and a["phi0"] > 0.0:
                if st.norms0d("scalar_phi", 0)["nm1"] < self.disperse_frac * a["phi0"]:
                    return "dispersed"
        except self.oaxis.OaxisError:
            return "failed"          # a NaN in the state: the request retires with that status
        return None

    def run(self, requests, clock=time.perf_counter, progress=None, step_progress=None, step_every=16,
            admit_progress=None):
        """Serve ``requests`` (arrival times relative to the start); returns
        (records, summary).  progress(record) is called at each retirement,
        step_progress(coarse steps, active members) every step_every coarse
        steps of the batch, admit_progress(request, active members) when a
        prepared request joins the batch."""
        oaxis = self.oaxis
        pending = sorted(requests, key=lambda r: (r.arrival, r.rid))
        active, done = [], []
        preparing = []            # (request, admission time, footprint, future), in admission order
        used = 0.0
        batch = oaxis.Batch()
        # the membership changes too often for the schedule autotuner to
        # settle: a fixed schedule (graphs unless switched off; RHS blocks
        # renumbered per XCD)
        batch.set_graphs(self.graphs)
        pool = ThreadPoolExecutor(max_workers=self.prepare_threads)
        t0 = clock()
        total_updates = 0.0
        batch_steps = 0
        search = None             # (members, HorizonSearch) begun after the last step
        first_step = None         # time of the first coarse step
        try:
            while pending or active or preparing:
                now = clock() - t0
                # admit what has arrived, while memory and slots allow: the
                # request is prepared on a worker thread meanwhile
                while (pending and pending[0].arrival <= now
                       and len(active) + len(preparing) < self.max_members):
                    fp = oaxis.footprint(oaxis.params_from_text(pending[0].text))
                    if (active or preparing) and used + fp > self.budget:
                        break
                    if fp > self.budget:
                        raise MemoryError("request %d needs %.3g B, more than the budget" % (pending[0].rid, fp))
                    req = pending.pop(0)
                    used += fp
                    preparing.append((req, now, fp, pool.submit(self._prepare, req)))
                    now = clock() - t0
                # prepared requests join at this step boundary, in admission order
                while preparing and (preparing[0][3].done() or not active or self.gang):
                    req, t_adm, fp, fut = preparing.pop(0)
                    self._admit(req, t_adm, batch, active, used, prepared=fut.result())
                    if admit_progress is not None:
                        admit_progress(req, len(active))
                if not active:
                    if pending and not preparing:
                        time.sleep(max(0.0, min(0.01, pending[0].arrival - now)))
                    continue
                for a in self._initial_searches(batch, active):
                    rec = self._retire(a, "failed", clock() - t0, batch)
                    total_updates += rec.point_updates
                    used -= rec.footprint
                    active.remove(a)
                    done.append(rec)
                if not active:
                    continue
                if first_step is None:
                    first_step = clock() - t0
                batch.coarse_step()
                batch_steps += 1
                for a in active:
                    a["steps"] += 1
                if step_progress is not None and batch_steps % step_every == 0:
                    step_progress(batch_steps, len(active))
                # the searches begun after the previous step ran beside this one
                found = self._end_searches(search)
                search = None
                retire = self._due_retirements(batch, active, found)
                due = [a for a in active if self._search_due(a) and all(a is not r for r, _ in retire)]
                if due:
                    try:
                        search = (due, batch.ahfind_begin([(a["st"], 0) for a in due]))
                    except oaxis.OaxisError:
                        retire.extend((a, "failed") for a in due)
                if retire:
                    batch.sync()
                    now = clock() - t0
                    for a, reason in retire:
                        rec = self._retire(a, reason, now, batch)
                        total_updates += rec.point_updates
                        used -= rec.footprint
                        active.remove(a)
                        done.append(rec)
                        if progress is not None:
                            progress(rec)
        finally:
            if search is not None:
                try:
                    search[1].end()
                except Exception:
                    pass
            for _, _, _, fut in preparing:
                try:
                    fut.result()[1].close()
                except Exception:
                    pass
            pool.shutdown(wait=True)
            for a in active:
                a["st"].close()
            batch.close()
        records, summary = self._summary(done, clock() - t0, total_updates)
        summary["first_step_s"] = first_step
        return records, summary

    def _search_due(self, a):
        """A horizon search of member a is due after its current step."""
        return (a["ah"] and a["watch"] and a["steps"] % a["ah_every"] == 0
                and a["steps"] * a["dt0"] >= a["ah_after"])

    def _end_searches(self, search):
        """Results of the searches begun after the previous step (they ran
        beside this step): {id(member): "horizon" | "failed"} for the members
        whose search found a horizon or failed."""
        if search is None:
            return {}
        members, hs = search
        try:
            res = hs.end()
        except self.oaxis.OaxisError:
            return {id(a): "failed" for a in members}
        out = {}
        for a, r in zip(members, res):
            if r["found"]:
                a["rec"].horizon = r
                out[id(a)] = "horizon"
        return out

    def _due_retirements(self, batch, active, found):
        """Members to retire after a coarse step: a horizon found by the
        search of the state one step back (the search runs beside the next
        step, and the member retires after it), the final step, and a check
        every check_every steps (lapse collapse, dispersal, NaN)."""
        retire = []
        for a in active:
            if id(a) in found:
                retire.append((a, found[id(a)]))
                continue
            if a["steps"] >= a["req"].steps:
                retire.append((a, "final_time"))
                continue
            if self.check_every > 0 and a["steps"] % self.check_every == 0:
                reason = self._check(a)
                if reason:
                    retire.append((a, reason))
        return retire

    @staticmethod
    def _summary(done, wall, total_updates):
        lat = np.array([r.latency for r in done]) if done else np.zeros(1)
        summary = dict(requests=len(done), wall_s=wall, point_updates=total_updates,
                       updates_per_s=total_updates / wall if wall > 0 else 0.0,
                       p50_s=float(np.percentile(lat, 50)), p99_s=float(np.percentile(lat, 99)),
                       reasons={k: sum(1 for r in done if r.reason == k) for k in set(r.reason for r in done)})
        return sorted(done, key=lambda r: r.rid), summary

    def run_plain(self, requests, clock=time.perf_counter, progress=None):
        """The same requests one spacetime at a time (the plain path): each is
        prepared, evolved with the same retirement rules and searches, and
        retired before the next starts; arrival times are ignored (every
        request is taken as waiting from the start).  Returns (records,
        summary) like run()."""
        done = []
        total_updates = 0.0
        t0 = clock()
        for req in sorted(requests, key=lambda r: (r.arrival, r.rid)):
            batch = self.oaxis.Batch()
            batch.set_graphs(self.graphs)
            active = []
            try:
                self._admit(req, clock() - t0, batch, active, 0.0)
                batch.remove(active[0]["st"])      # the plain path: the spacetime alone
                a = active[0]
                st = a["st"]
                if a["ah_t0"]:
                    a["ah_t0"] = False
                    if st.ahfind(0)["found"]:
                        a["watch"] = False
                reason = None
                search = None
                while reason is None:
                    st.coarse_step()
                    a["steps"] += 1
                    reason = self._plain_retirement(a, search)
                    search = None
                    if reason is None and self._search_due(a):
                        try:
                            search = ([a], st.ahfind_begin(0))
                        except self.oaxis.OaxisError:
                            reason = "failed"
                rec = a["rec"]
                rec.retired = clock() - t0
                rec.arrival = 0.0
                rec.reason = reason
                rec.steps = a["steps"]
                rec.point_updates = st.point_updates() - a["u0"]
                total_updates += rec.point_updates
                done.append(rec)
                st.close()
                if progress is not None:
                    progress(rec)
            finally:
                batch.close()
        return self._summary(done, clock() - t0, total_updates)

    def _plain_retirement(self, a, search):
        """As _due_retirements for a spacetime served alone; search is the
        one begun after its previous step."""
        if search is not None:
            members, hs = search
            try:
                r = hs.end()
            except self.oaxis.OaxisError:
                return "failed"
            if r["found"]:
                a["rec"].horizon = r
                return "horizon"
        if a["steps"] >= a["req"].steps:
            return "final_time"
        if self.check_every > 0 and a["steps"] % self.check_every == 0:
            reason = self._check(a)
            if reason:
                return reason
        return None


def serving_mix(seed, n, rate, grids=((64, 128), (128, 256), (256, 512), (512, 1024)), max_levels=6,
                steps=(200, 2000), ahfind_every=16):
    """Seeded request stream of the serving mix: Poisson arrivals at ``rate``
    per second, 40 % Schwarzschild (M~U[.5,2]), 30 % Brill (a~U[1,4]), 30 %
    scalar pulse; grid, level count (1..max_levels) and coarse steps drawn
    uniformly.  Every request searches for horizons every ``ahfind_every``
    coarse steps (0: no horizon search)."""
    from . import workloads
    rng = np.random.default_rng(seed)
    t = 0.0
    out = []
    for k in range(n):
        t += rng.exponential(1.0 / rate)
        u = rng.uniform()
        Nr, Nz = grids[int(rng.integers(0, len(grids)))]
        nl0 = int(rng.integers(0, max_levels))
        nsteps = int(rng.integers(steps[0], steps[1] + 1))
        if u < 0.4:
            kind = "schwarzschild"
            text = workloads.schwarzschild(float(rng.uniform(0.5, 2.0)), Nr, Nz, nl0)
        elif u < 0.7:
            kind = "brill"
            text = workloads.brill(int(rng.integers(0, 2 ** 31)), Nr, Nz, nl0, dr0=25.6 / Nr)
        else:
            kind = "scalar_pulse"
            text = workloads.scalar_pulse(seed, k) + "\nNrtotal = %d\nNztotal = %d\nNl0 = %d\ndr0 = %r\ndz0 = %r\n" % (
                Nr, Nz, nl0, 12.0 / Nr, 12.0 / Nr)
        if ahfind_every > 0:
            text += "\nahfind = .true.\nahfind_every = %d\n" % ahfind_every
        out.append(Request(k, text, nsteps, t, kind))
    return out
