"""Device-resident serving loop for implicit HDG steps, with continuous batching.

Requests are independent HDG problems (a mesh, an order p, a state) that each
need one or more Newton iterations of an implicit (BDF1) or steady solve
(kernel/hdg_solve.m:103-170, kernel/hdg_solve_bdf.m:26-48).  The loop keeps
every admitted request's state in HBM and advances all of them together:

  * admission: a request that has arrived is admitted as soon as its HBM
    footprint fits the budget and a slot is free;
  * admission builds the request's mesh connectivity and initial trace on
    the device (hdg_amd.devmesh: hdgx_mesh_faces / _elcon / _f2f), on the
    clock;
  * elemental phase, one call in flight at a time: the requests that became
    ready while the previous call ran, of every order p, run as ONE
    mixed-order batched call
    (hdgx_elemental_batch_mixed: a batched launch per order present, back to
    back), every request with its own physics parameters;
  * trace phase: each request's face-block system, block ILU(0), GMRES and
    local update run on a worker stream (hdgx_trace_*, hdgx_gmres,
    hdgx_local_update), so the trace solves of several requests and the next
    elemental batch overlap on the device;
  * a request whose iterations are done frees its memory; one with more
    iterations rejoins the next elemental batch.

Latency of a request = completion time - arrival time (wall clock).
"""
import argparse
import json
import os
import queue
import sys
import threading
import time
from dataclasses import dataclass, field

import numpy as np

from .runtime import Context, Dev, HdgxError, _torch
from .trace import TraceSystem

PHYS_NS = "ns"
PHYS_EULER = "euler"


@dataclass
class RequestSpec:
    rid: int
    t_arrival: float
    n: int                 # cells per side: 2*n*n triangles
    p: int
    minf: float
    seed: int
    dt: float = 0.05
    physics: str = PHYS_NS
    iters: int = 1         # Newton iterations (1 = one BDF1 Newton step)

    @property
    def ne(self):
        return 2 * self.n * self.n


def make_workload(nreq=512, seed=2024, rate=50.0, ne_lo=500, ne_hi=50000, levels=16, physics=PHYS_NS,
                  cfl=10.0):
    """Poisson arrivals at `rate` per second; ne log-uniform in [ne_lo, ne_hi]
    snapped to `levels` log-spaced square meshes (2*n*n triangles), p uniform
    in 1..4, M uniform in [0.2, 0.7], all from the seed.  The BDF1 step of
    each request is dt = cfl * h / (|u| + c) with h = 1/n, |u| = 1 and the
    free-stream sound speed c = 1/M (an acoustic CFL number)."""
    rng = np.random.default_rng(np.random.SeedSequence(int(seed)))
    t = np.cumsum(rng.exponential(1.0 / rate, size=nreq))
    ne = np.exp(rng.uniform(np.log(ne_lo), np.log(ne_hi), size=nreq))
    lv = np.unique(np.round(np.geomspace(np.sqrt(ne_lo / 2), np.sqrt(ne_hi / 2), levels)).astype(int))
    n = lv[np.argmin(np.abs(np.log(lv[None, :]) - np.log(np.sqrt(ne / 2))[:, None]), axis=1)]
    p = rng.integers(1, 5, size=nreq)
    minf = rng.uniform(0.2, 0.7, size=nreq)
    seeds = rng.integers(0, 2 ** 31, size=nreq)
    dt = cfl / n / (1.0 + 1.0 / minf)
    return [RequestSpec(i, float(t[i]), int(n[i]), int(p[i]), float(minf[i]), int(seeds[i]), float(dt[i]), physics)
            for i in range(nreq)]


class PayloadFactory:
    """Request messages: what a client sends -- the mesh geometry (vertices,
    triangles and DG node coordinates, cached per (n, p) since the workload
    reuses 16 mesh sizes), the physics parameters and the state from the
    request's seed.  The connectivity (faces, trace numbering, face tables,
    boundary tags) and the initial trace are the server's work, built on the
    device on admission (hdg_amd.devmesh)."""

    def __init__(self):
        self._geom = {}
        self._master = {}

    def master(self, p):
        from .master import mkmaster
        if p not in self._master:
            self._master[p] = mkmaster(2, p)
        return self._master[p]

    def geometry(self, n, p):
        """(vertices (nv, 2), t (ne, 3), dgnodes (npv, 2, ne)) of the jittered
        2 n^2-triangle square (mesh/squaremesh.m nodes, mesh/mkmesh/createnodes.m)."""
        from .mesh import createnodes, jitter, squaremesh
        key = (n, p)
        if key not in self._geom:
            vert, t = squaremesh(n + 1, n + 1)
            vert = jitter(vert, 1.0 / n, 0.1, seed=1000 + n)
            self._geom[key] = (vert, t, createnodes(vert, t, self.master(p).plocvl))
        return self._geom[key]

    # the host-built mesh of the same geometry (hdg_amd/mesh.py), for checks
    def host_mesh(self, n, p):
        from .mesh import mkmesh_square
        return mkmesh_square(n, n, self.master(p), 0.1, seed=1000 + n, ncd_extra=0)

    def build(self, spec):
        from .app import euler_app, ns_app
        from .problems import euler_state, ns_state
        m = self.master(spec.p)
        vert, t, dgnodes = self.geometry(spec.n, spec.p)
        if spec.physics == PHYS_NS:
            app = ns_app(minf=spec.minf)
            udg = ns_state(dgnodes, 1.4, spec.minf, spec.seed)
        else:
            app = euler_app(2, minf=spec.minf)
            udg = euler_state(dgnodes, 2, 1.4, spec.minf, spec.seed)
        sh = np.zeros_like(udg)
        if spec.dt > 0:
            # BDF1 (hdg_solve_bdf.m:29-43 with torder 1): fc_u = 1/dt, SH = u_n/dt
            app.tdep = 1
            app.fc_u = 1.0 / spec.dt
            sh[:, :app.ncu, :] = udg[:, :app.ncu, :] / spec.dt
        return dict(master=m, app=app, vertices=vert, t=t, dgnodes=dgnodes, udg=np.asfortranarray(udg),
                    sh=np.asfortranarray(sh), dt=spec.dt)


def faces_of(spec):
    """Faces of the request's 2 n^2-triangle square: 3 n^2 + 2 n."""
    return 3 * spec.n * spec.n + 2 * spec.n


def element_matrix_bytes(master, app):
    """Bytes of one element's matrices in the elemental call's workspace (the
    reference layout: BD, F, GK, H, Ru, Rh and, with q, M, C, E, Rq)."""
    npv, ncu, nc, nd = master.npv, app.ncu, app.nc, master.nd
    ndf = master.npf * master.nfe
    nu, nh, ncol = npv * ncu, ncu * ndf, npv * nc
    w = nu * ncol + nu * nh + nh * ncol + nh * nh + nu + nh
    if app.localsolve:
        w += npv * npv + npv * npv * nd + npv * ndf * nd + npv * ncu * nd
    return 8 * w


def footprint_bytes(spec, master, app, nf):
    """HBM bytes a request holds while admitted."""
    npv, ncu, nc = master.npv, app.ncu, app.nc
    ndf = master.npf * master.nfe
    ne = spec.ne
    ncf = ncu * master.npf
    nbf = 2 * master.nfe - 1
    el = ne * (ncu * ndf * ncu * ndf + ncu * ndf + npv * nc + npv * nc * ncu * ndf)   # AE FE DUDG DUDG_DUH
    st = ne * (3 * npv * nc + npv * 3 + ncu * ndf) + 4 * ncf * nf                   # UDG UDG0 SH DUDGT dgnodes uhe, UH..
    tr = nf * (ncf * ncf * (nbf - 1) + ncf * ncf * nbf + 2 * ncf)                     # K LU F
    return 8 * (el + st + tr)


@dataclass
class Slot:
    spec: RequestSpec
    payload: dict
    bytes: int
    state: object = None
    it: int = 0
    t_admit: float = 0.0
    t_done: float = 0.0
    residuals: list = field(default_factory=list)
    gmres_iters: list = field(default_factory=list)
    views: tuple = None
    event: object = None
    admit_s: float = 0.0


class _ReqState:
    """Device arrays of an admitted request; its connectivity and initial
    trace are built on the device (hdg_amd.devmesh) with the mesh context
    `mctx`.  The elemental outputs are views of a batch launch's buffers."""

    def __init__(self, mctx, payload, stream):
        from . import devmesh
        torch = _torch()
        m, app = payload["master"], payload["app"]
        self.mesh, self.UH, self.ts = devmesh.build(mctx, payload["vertices"], payload["t"], m,
                                                    payload["dgnodes"], payload["udg"], app.ncu)
        mesh = self.mesh
        with torch.cuda.stream(stream):
            ts = self.ts
            self.dgnodes = mesh.dgnodes
            self.bf = mesh.bf
            self.UDG = Dev.from_numpy(payload["udg"])
            if app.tdep:
                # BDF1 source SH = u_n / dt of the state just sent (hdg_solve_bdf.m:29-43,
                # torder 1), formed on the device instead of sent: zero q columns
                npv, nc, ne = self.UDG.shape
                U = self.UDG.t.view(ne, nc, npv)
                sh = torch.zeros_like(U)
                sh[:, :app.ncu, :] = U[:, :app.ncu, :] / payload["dt"]
                self.SH = Dev(sh.flatten(), (npv, nc, ne))
            else:
                self.SH = Dev.from_numpy(payload["sh"])
            self.uhe = Dev.empty((app.ncu, m.npf * m.nfe, mesh.ne))
            self.K = Dev.empty((ts.ncf, ts.ncf, ts.nbf - 1, ts.nf))
            self.F = Dev.empty((ts.ncf, ts.nf))
            self.fst = Dev.empty((ts.nf,), "int32")
            self.LU = Dev.empty((ts.ncf, ts.ncf, ts.nbf, ts.nf))
            self.DUH = Dev.zeros((ts.ncf, ts.nf))
            self.DUDGT = Dev.empty((m.npv, app.nc, mesh.ne))


class Server:
    """Continuous-batching loop on one device.

    budget_bytes: HBM the admitted requests may hold together; slots: maximum
    number of admitted requests; workers: concurrent trace solves (each with
    its own hdgx context and stream).

    Batching window: a request that becomes ready (admitted, or back from its
    trace solve) waits for others until the oldest ready one has waited
    batch_window seconds, the ready set holds batch_fill_elems elements, or
    nothing that could join is in flight; then, once the previous elemental
    call has finished, the ready set goes into one mixed-order call.  The
    window bounds the latency batching adds per Newton step."""

    def __init__(self, device=0, budget_bytes=None, slots=64, workers=4, restart=20, gmres_tol=1e-6,
                 max_batch_elems=2_000_000, admitters=2, batch_window=0.005, batch_fill_elems=200_000):
        torch = _torch()
        self.device = device
        if budget_bytes is None:
            free, total = torch.cuda.mem_get_info(device)
            budget_bytes = int(0.6 * free)
            # the elemental context's workspace (24 GiB at most) comes out of it
            from .master import mkmaster
            from .app import ns_app
            per = element_matrix_bytes(mkmaster(2, 4), ns_app())
            budget_bytes -= min(24 << 30, max_batch_elems * per)
        self.budget = budget_bytes
        self.slots = slots
        self.nworkers = workers
        self.nadmit = admitters
        self.restart, self.gmres_tol = restart, gmres_tol
        self.max_batch_elems = max_batch_elems
        self.batch_window = batch_window
        self.batch_fill_elems = batch_fill_elems
        self.ctx = Context(device)
        self.main_stream = torch.cuda.current_stream(device)


    # ------------------------------------------------------------- workers
    def _admitter(self, aq, doneq):
        """Admission: a request's connectivity, initial trace and state on the
        device, on a context and stream of its own (the mesh entry points
        synchronise their stream, which must not wait for queued batches)."""
        torch = _torch()
        stream = torch.cuda.Stream(self.device)
        ctx = Context(self.device)
        ctx.bind_stream(stream)
        while True:
            slot = aq.get()
            if slot is None:
                return
            try:
                t = time.perf_counter()
                with torch.cuda.stream(stream):
                    slot.state = _ReqState(ctx, slot.payload, stream)
                stream.synchronize()
                slot.admit_s = time.perf_counter() - t
            except Exception as exc:
                doneq.put(("built", slot, exc))
                continue
            doneq.put(("built", slot, None))

    def _worker(self, wq, doneq):
        torch = _torch()
        stream = torch.cuda.Stream(self.device)
        ctx = Context(self.device)
        ctx.bind_stream(stream)
        while True:
            item = wq.get()
            if item is None:
                return
            slot = item
            try:
                with torch.cuda.stream(stream):
                    stream.wait_event(slot.event)
                    self._trace_step(ctx, slot)
                    stream.synchronize()
            except Exception as exc:      # surfaced by the scheduler
                doneq.put(("done", slot, exc))
                continue
            doneq.put(("done", slot, None))

    def _trace_step(self, ctx, slot):
        S = slot.state
        ts = S.ts
        ts.ctx = ctx
        AE, FE, DUDG, DUDG_DUH, est = slot.views
        ts.assemble(AE, FE, out=(S.K, S.F, S.fst))
        slot.residuals.append(ts.norm2(S.F))
        ts.bilu0_factor(S.K, S.LU)
        S.DUH.t.zero_()
        _, it, rr, fl = ts.gmres(S.K, S.F, S.DUH, LU=S.LU, restart=self.restart, tol=self.gmres_tol)
        slot.gmres_iters.append(it)
        m, app = slot.payload["master"], slot.payload["app"]
        ts.local_update(m.npv, app.nc, DUDG, DUDG_DUH, S.DUH, S.DUDGT)
        ts.newton_update(m.npv, app.nc, app.nc, 1.0, S.UDG, S.DUDGT, S.UH, S.DUH, S.UDG, S.UH)
        slot.it += 1

    # ------------------------------------------------------------ elemental
    def _elemental_batches(self, ready):
        """Ready requests of one physics, whatever their orders, go into one
        mixed-order batched call (hdgx_elemental_batch_mixed); very large sets
        are split so that one call's output buffers stay bounded."""
        groups = {}
        for s in ready:
            groups.setdefault(s.spec.physics, []).append(s)
        launched = []
        for phys, slots in groups.items():
            chunk, cur = [], 0
            for s in slots:
                if chunk and cur + s.spec.ne > self.max_batch_elems:
                    launched.append(self._launch_group(chunk))
                    chunk, cur = [], 0
                chunk.append(s)
                cur += s.spec.ne
            if chunk:
                launched.append(self._launch_group(chunk))
        return launched

    def _launch_group(self, slots):
        torch = _torch()
        sizes = []
        for s in slots:
            m, app = s.payload["master"], s.payload["app"]
            ndf = m.npf * m.nfe
            nh = app.ncu * ndf
            ne = s.spec.ne
            sizes.append((nh * nh * ne, nh * ne, m.npv * app.nc * ne, m.npv * app.nc * nh * ne, ne))
        tot = [sum(z[k] for z in sizes) for k in range(5)]
        AE = torch.empty(tot[0], dtype=torch.float64, device="cuda")
        FE = torch.empty(tot[1], dtype=torch.float64, device="cuda")
        DU = torch.empty(tot[2], dtype=torch.float64, device="cuda")
        DX = torch.empty(tot[3], dtype=torch.float64, device="cuda")
        ST = torch.empty(tot[4], dtype=torch.int32, device="cuda")
        segs, masters = [], []
        off = [0] * 5
        for s, z in zip(slots, sizes):
            S = s.state
            m, app = s.payload["master"], s.payload["app"]
            npv, ncu, nc, ndf, ne = m.npv, app.ncu, app.nc, m.npf * m.nfe, s.spec.ne
            S.ts.ctx = self.ctx
            S.ts.gather(S.UH, S.uhe)
            v = (Dev(AE[off[0]:off[0] + z[0]], (ncu, ndf, ncu, ndf, ne)),
                 Dev(FE[off[1]:off[1] + z[1]], (ncu, ndf, ne)),
                 Dev(DU[off[2]:off[2] + z[2]], (npv, nc, ne)),
                 Dev(DX[off[3]:off[3] + z[3]], (npv, nc, ncu, ndf, ne)),
                 Dev(ST[off[4]:off[4] + z[4]], (ne,)))
            off = [a + b for a, b in zip(off, z)]
            s.views = v
            segs.append((app, S.dgnodes, S.bf, S.UDG, S.uhe, S.SH, v))
            masters.append(m)
        self.ctx.elemental_batch_mixed(masters, segs)
        ev = torch.cuda.Event()
        ev.record(self.main_stream)
        for s in slots:
            s.event = ev
        return len(slots), tot[4], len({s.spec.p for s in slots})

    # ----------------------------------------------------------------- loop
    def run(self, specs, factory=None, realtime=True, log=None, keep_state=False):
        """Serve `specs` (sorted by arrival).  realtime=False admits every
        request at once (arrival times ignored).  Returns a stats dict and the
        finished slots (with their device state when keep_state)."""
        torch = _torch()
        factory = factory or PayloadFactory()
        pending = sorted(specs, key=lambda s: s.t_arrival)
        payloads = {s.rid: factory.build(s) for s in pending}      # request messages, prepared off the clock
        wq, aq, doneq = queue.Queue(), queue.Queue(), queue.Queue()
        threads = [threading.Thread(target=self._worker, args=(wq, doneq), daemon=True) for _ in range(self.nworkers)]
        admitters = [threading.Thread(target=self._admitter, args=(aq, doneq), daemon=True)
                     for _ in range(self.nadmit)]
        for t in threads + admitters:
            t.start()
        used = 0
        active = []
        ready = []              # waiting for an elemental evaluation
        ready_since = None      # when the oldest of them became ready
        finished = []
        nbatches = 0
        batched_elems = 0
        admit_s = 0.0           # device mesh builds and uploads on admission
        mixed_batches = 0       # elemental calls over requests of several orders
        last_call = None        # device event of the last elemental call
        t0 = time.perf_counter()
        clock = (lambda: time.perf_counter() - t0) if realtime else (lambda: float("inf"))
        inflight = 0            # admissions and trace steps not reported yet
        try:
            while pending or active:
                now = clock()
                # admission by arrival, slots and HBM budget; the device work of
                # an admission runs on an admission thread
                while pending and pending[0].t_arrival <= now and len(active) < self.slots:
                    spec = pending[0]
                    pl = payloads[spec.rid]
                    b = footprint_bytes(spec, pl["master"], pl["app"], faces_of(spec))
                    if used + b > self.budget and active:
                        break
                    pending.pop(0)
                    slot = Slot(spec, pl, b)
                    slot.t_admit = time.perf_counter() - t0
                    used += b
                    active.append(slot)
                    aq.put(slot)
                    inflight += 1
                # one elemental call at a time (iteration-level batching): the
                # ready set closes when its window ends, it is large enough, or
                # no admission or trace step could still join it; what becomes
                # ready while the previous call still runs joins the next one
                waited = time.perf_counter() - ready_since if ready else 0.0
                full = sum(s.spec.ne for s in ready) >= self.batch_fill_elems
                close = ready and (waited >= self.batch_window or full or inflight == 0)
                if close and (last_call is None or last_call.query()):
                    for n, ne, norders in self._elemental_batches(ready):
                        nbatches += 1
                        batched_elems += ne
                        mixed_batches += norders > 1
                    last_call = ready[-1].event
                    for s in ready:
                        wq.put(s)
                        inflight += 1
                    ready, ready_since = [], None
                # collect admissions and finished trace steps; wait when nothing
                # else can progress
                timeout = None
                if pending and len(active) < self.slots:
                    timeout = max(0.0, pending[0].t_arrival - clock()) if realtime else 0.0
                if inflight == 0 and timeout is None:
                    timeout = 0.0
                if ready:                   # the window's end, or the previous call: poll
                    left = max(self.batch_window - (time.perf_counter() - ready_since), 0.0) if not close else 0.0
                    wait = max(min(left, 2e-4) if not close else 2e-4, 1e-5)
                    timeout = wait if timeout is None else min(timeout, wait)
                try:
                    item = doneq.get(timeout=timeout) if (timeout is None or timeout > 0) else doneq.get_nowait()
                except queue.Empty:
                    continue
                while True:
                    kind, slot, exc = item
                    inflight -= 1
                    if exc is not None:
                        raise exc
                    if kind == "built" or slot.it < slot.spec.iters:
                        if kind == "built":
                            admit_s += slot.admit_s
                        if not ready:
                            ready_since = time.perf_counter()
                        ready.append(slot)
                    else:
                        slot.t_done = time.perf_counter() - t0
                        slot.views = None
                        if not keep_state:
                            slot.state = None
                        active.remove(slot)
                        used -= slot.bytes
                        finished.append(slot)
                        if log:
                            t_arr = slot.spec.t_arrival if realtime else 0.0
                            log(f"[serve] req {slot.spec.rid} p={slot.spec.p} ne={slot.spec.ne} done "
                                f"{(slot.t_done - t_arr) * 1e3:.1f} ms (waited {(slot.t_admit - t_arr) * 1e3:.1f} ms, "
                                f"gmres {slot.gmres_iters})")
                    try:
                        item = doneq.get_nowait()
                    except queue.Empty:
                        break
        finally:
            for _ in threads:
                wq.put(None)
            for _ in admitters:
                aq.put(None)
            for t in threads + admitters:
                t.join()
        makespan = time.perf_counter() - t0
        lat = np.array([(s.t_done - (s.spec.t_arrival if realtime else 0.0)) for s in finished])
        ne_total = sum(s.spec.ne * s.spec.iters for s in finished)
        stats = {
            "requests": len(finished),
            "p50_ms": float(np.percentile(lat, 50) * 1e3) if lat.size else None,
            "p99_ms": float(np.percentile(lat, 99) * 1e3) if lat.size else None,
            "mean_ms": float(lat.mean() * 1e3) if lat.size else None,
            "makespan_s": makespan,
            "requests_per_s": len(finished) / makespan if makespan > 0 else None,
            "elements_per_s": ne_total / makespan if makespan > 0 else None,
            "elemental_batches": nbatches,
            "mean_requests_per_batch": (sum(s.spec.iters for s in finished) / nbatches) if nbatches else None,
            "gmres_iters_mean": float(np.mean([np.mean(s.gmres_iters) for s in finished])) if finished else None,
            "admission_ms_mean": admit_s / len(finished) * 1e3 if finished else None,
            "mixed_order_batches": mixed_batches,
            "batch_window_ms": self.batch_window * 1e3,
        }
        return stats, finished


def main():
    ap = argparse.ArgumentParser(description="HDG serving loop (continuous batching) on one GPU")
    ap.add_argument("--requests", type=int, default=512)
    ap.add_argument("--seed", type=int, default=2024)
    ap.add_argument("--rate", type=float, default=50.0)
    ap.add_argument("--workers", type=int, default=4)
    ap.add_argument("--slots", type=int, default=64)
    ap.add_argument("--physics", default=PHYS_NS, choices=[PHYS_NS, PHYS_EULER])
    ap.add_argument("--batch-window", type=float, default=0.005, help="seconds a ready request waits for others")
    a = ap.parse_args()
    specs = make_workload(a.requests, a.seed, a.rate, physics=a.physics)
    srv = Server(workers=a.workers, slots=a.slots, batch_window=a.batch_window)
    stats, _ = srv.run(specs, log=lambda m: print(m, file=sys.stderr, flush=True))
    print(json.dumps(stats))


if __name__ == "__main__":
    main()
