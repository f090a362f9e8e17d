"""Benchmark: element Schur solves/s of the HDG elemental hot path on one MI355X.

Workload: 2D Navier-Stokes HDG elemental step (u and q, nc = 12) on a
jittered unit square of 2 x 128 x 128 = 32768 p = 3 triangles, Re 1e4,
Pr 0.72, M 0.3, 2% sinusoidal velocity perturbation, laplacian artificial
viscosity 1e-3 on 5% of the elements; inputs synthetic and seeded, resident in
HBM.  One step = one hdg_elemental pass over all elements (volint + faceint +
schur_uq): (dgnodes, UDG, UH, SH, bf) -> (AE, FE, DUDG, DUDG_DUH).

Reported: FP64 throughput (`value`) and the FP32 + FP64-refinement form beside
it, each with its normwise relative error against the CPU FP64 oracle run on
the same inputs; the roofline of the dominant kernel from HIP events on the
launch stream; the oracle's CPU throughput on this host.

    python bench.py [--gpus 1] [--steps K] [--warmup W]
"""
import argparse
import json
import os
import sys
import time

import numpy as np

ROOT = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, ROOT)

PEAK_FP64_TFLOPS = 78.6
PEAK_FP32_TFLOPS = 157.3
PEAK_HBM_GBS = 8000.0
METRIC = "element Schur solves/s FP64 vs FP32+IR @1e-10 rel err; p50/p99 request latency"


def log(msg):
    print(msg, file=sys.stderr, flush=True)


def relerr(x, ref):
    n = np.linalg.norm(ref.ravel())
    return float(np.linalg.norm((x - ref).ravel()) / (n if n > 0 else 1.0))


def time_steps(torch, stream, fn, steps, warmup):
    for _ in range(warmup):
        fn()
    torch.cuda.synchronize()
    t0 = time.perf_counter()
    ev0 = torch.cuda.Event(enable_timing=True)
    ev1 = torch.cuda.Event(enable_timing=True)
    ev0.record(stream)
    for _ in range(steps):
        fn()
    ev1.record(stream)
    torch.cuda.synchronize()
    wall = time.perf_counter() - t0
    return ev0.elapsed_time(ev1) / steps, wall


def kernel_times(torch, ctx, P, dev, out, reps):
    """Per-kernel durations inside hdgx_elemental: the library brackets its
    quadrature and elimination launches with HIP events on the launch stream
    (hdgx_set_timing), over `reps` calls."""
    ctx.elemental(P.master, P.app, *dev, out=out)
    torch.cuda.synchronize()
    ctx.set_timing(True)
    try:
        for _ in range(reps):
            ctx.elemental(P.master, P.app, *dev, out=out)
        tq, ts, n = ctx.kernel_times()
    finally:
        ctx.set_timing(False)
    return tq / n, ts / n


def measured_traffic(kernel):
    """HBM bytes per launch of `kernel` on this workload from the committed
    rocprofv3 FETCH_SIZE / WRITE_SIZE passes (scaled by the same counters'
    reading of a 1 GiB device copy), or None."""
    path = os.path.join(ROOT, "profiles", "elemental_ns2d_p3_2x128x128_traffic.json")
    if not os.path.exists(path):
        return None
    d = json.load(open(path))
    k = d.get("kernels", {}).get(kernel)
    return None if k is None else k["bytes_per_launch"]


def cpu_baseline(P, nthreads):
    """CPU rate of the oracle on this host (a -march=native build when it
    builds, else the portable one) and, from the portable build that every
    check is pinned on, the reference outputs and the FP64 noise floor."""
    from oracle import oracle as O
    try:
        O.native_lib()
        native, kind_build = True, "native"
    except Exception:
        native, kind_build = False, "portable"
    args = (P.master, P.app, P.mesh.dgnodes, P.mesh.bf, P.udg, P.uh, P.sh)
    O.elemental(*args, nthreads=nthreads, native=native)     # warm-up
    runs = []
    for _ in range(5):
        t0 = time.perf_counter()
        O.elemental(*args, nthreads=nthreads, native=native)
        runs.append(time.perf_counter() - t0)
    out = O.elemental(*args, nthreads=nthreads)
    t_all = float(np.median(runs))
    # single core on a 2048-element sample (the MEX element loop is single-threaded)
    ns = min(2048, P.ne)
    sample = (P.master, P.app, P.mesh.dgnodes[:, :, :ns], P.mesh.bf[:, :ns], P.udg[:, :, :ns],
              P.uh[:, :, :ns], P.sh[:, :, :ns])
    O.elemental(*sample, nthreads=1, native=native)
    r1 = []
    for _ in range(5):
        t0 = time.perf_counter()
        O.elemental(*sample, nthreads=1, native=native)
        r1.append(time.perf_counter() - t0)
    t_one = float(np.median(r1))
    # FP64 noise floor of this workload: the oracle's own response to random
    # +-1 ulp perturbations of every input (node coordinates, state, trace).
    # Two FP64 evaluations that round differently (e.g. the geometry and the
    # summation order) differ by about this much; DUDG, a Newton update built
    # from a near-cancelling residual, sits far above the other arrays.
    rng = np.random.default_rng(7)
    ulp = lambda a: np.asfortranarray(a * (1.0 + np.ldexp(1.0, -52) * rng.choice([-1.0, 1.0], size=a.shape)))
    dg1 = P.mesh.dgnodes.copy(order="F")
    nd = P.master.nd
    dg1[:, :nd, :] = ulp(dg1[:, :nd, :])
    pert = O.elemental(P.master, P.app, dg1, P.mesh.bf, ulp(P.udg), ulp(P.uh), P.sh, nthreads=nthreads)
    floor = {n: relerr(pert[i], out[i]) for i, n in enumerate(("AE", "FE", "DUDG", "DUDG_DUH"))}
    return out, P.ne / t_all, ns / t_one, kind_build, sum(runs) + sum(r1), floor


def residual_leg(torch, ctx, P, dev, steps, warmup):
    """The residual without Jacobians (hdgx_residual, kernel/hdg_residual.m) on
    the same elements; Ru/Rh/Rq checked against the element-matrix kernel's
    (same arithmetic)."""
    from hdg_amd.runtime import Dev
    npv, ncu, nd, ndf, ne = P.master.npv, P.app.ncu, P.master.nd, P.master.ndf, P.ne
    rout = (Dev.empty((npv, ncu, ne)), Dev.empty((ncu, ndf, ne)), Dev.empty((npv, ncu, nd, ne)))
    ms, _ = time_steps(torch, ctx.stream, lambda: ctx.residual(P.master, P.app, *dev, out=rout), steps, warmup)
    mats = ctx.elemental_matrices(P.master, P.app, *dev)
    torch.cuda.synchronize()
    err = max(relerr(rout[0].numpy(), mats["Ru"].numpy()), relerr(rout[1].numpy(), -mats["Rh"].numpy()),
              relerr(rout[2].numpy(), mats["Rq"].numpy()))
    del mats
    log(f"[bench] residual: {ms:.3f} ms/step, {ne / ms * 1e3:.4g} elements/s")
    return {"workload": "hdg_residual element loop (Ru, Rh, Rq; no Jacobians), same elements",
            "ms_per_step": ms, "value": ne / (ms * 1e-3), "unit": "elements/s",
            "max_rel_diff_vs_element_matrices": err}


def getq_leg(torch, ctx, P, dev, steps, warmup):
    """q recovery (hdgx_getq, kernel/getq.m) on the same elements: time, its
    HBM rate against the algorithmic bytes (node coordinates, u, gathered
    trace and SH_q in; q out), and the local residual Rq with the recovered q
    put into UDG, relative to Rq with q = 0 (zero up to rounding)."""
    from hdg_amd.runtime import Dev
    npv, ncu, nd, ne, ndf = P.master.npv, P.app.ncu, P.master.nd, P.ne, P.master.ndf
    qout = Dev.empty((npv, ncu * nd, ne))
    ms, _ = time_steps(torch, ctx.stream, lambda: ctx.getq(P.master, dev[0], dev[2], dev[3], sh=dev[4],
                                                           fc_q=P.app.fc_q, out=qout), steps, warmup)
    nbytes = 8 * (npv * nd + npv * ncu + ncu * ndf + 2 * npv * ncu * nd) * ne    # + SH_q in
    udg = Dev(dev[2].t.clone(), dev[2].shape)
    ctx.getq(P.master, dev[0], udg, dev[3], sh=dev[4], fc_q=P.app.fc_q, inplace=True)
    rq1 = ctx.residual(P.master, P.app, dev[0], dev[1], udg, dev[3], dev[4])[2]
    udg.t.view(ne, -1, npv)[:, ncu:, :] = 0.0
    rqz = ctx.residual(P.master, P.app, dev[0], dev[1], udg, dev[3], dev[4])[2]
    torch.cuda.synchronize()
    red = float(rq1.t.abs().max() / rqz.t.abs().max())
    del udg, rq1, rqz
    log(f"[bench] getq: {ms:.3f} ms/step, {ne / ms * 1e3:.4g} elements/s, |Rq(q)| / |Rq(0)| = {red:.2e}")
    return {"workload": "getq q recovery (q = (fc_q M)^-1 (C u - E uh)), same elements",
            "ms_per_step": ms, "value": ne / (ms * 1e-3), "unit": "elements/s",
            "hbm_gbs_algorithmic": nbytes / (ms * 1e-3) / 1e9, "bytes_per_element": nbytes // ne,
            "rq_over_rq_at_q0": red}


def mesh_leg(torch, ctx, n=256, porder=3, reps=3):
    """Mesh connectivity for a new 2 n^2-triangle p3 mesh: hdgx_mesh_faces,
    hdgx_mesh_elcon and hdgx_mesh_f2f (DG nodes already on the device) against
    the host construction of hdg_amd/mesh.py and hdg_amd/trace.py (mkt2f,
    hdgelcon, mkf2f), whose arrays they must equal."""
    from hdg_amd.master import mkmaster
    from hdg_amd.mesh import createnodes, hdgelcon, jitter, mkt2f, squaremesh
    from types import SimpleNamespace
    from hdg_amd.runtime import Dev
    from hdg_amd.trace import mkf2f
    m = mkmaster(2, porder)
    p, t = squaremesh(n + 1, n + 1)
    p = jitter(p, 1.0 / n, 0.2, seed=n)
    dg = Dev.from_numpy(createnodes(p, t, m.plocvl))
    ctx.mesh_faces(t)                                  # warm-up (module load)
    torch.cuda.synchronize()
    t0 = time.perf_counter()
    for _ in range(reps):
        f, t2f, nf = ctx.mesh_faces(t)
        el = ctx.mesh_elcon(dg, m.perm, f, t2f, nf)
        fconn, f2f = ctx.mesh_f2f(f, t2f, nf)
    torch.cuda.synchronize()
    dt = (time.perf_counter() - t0) / reps
    t0 = time.perf_counter()
    f_ref, t2f_ref = mkt2f(t, 2)
    el_ref, _ = hdgelcon(dg.numpy(), m.perm, f_ref, t2f_ref)
    f2f_ref = mkf2f(SimpleNamespace(f=f_ref, t2f=t2f_ref))
    dth = time.perf_counter() - t0
    same = bool(np.array_equal(f.numpy().T, f_ref) and np.array_equal(t2f.numpy().T, t2f_ref)
                and np.array_equal(el.numpy(), el_ref) and np.array_equal(f2f.numpy(), f2f_ref))
    ne = t.shape[0]
    log(f"[bench] mesh: {ne} elements, {nf} faces: device {dt * 1e3:.2f} ms, host {dth * 1e3:.1f} ms, equal {same}")
    return {"workload": f"mkt2f + elcon + mkf2f, 2x{n}x{n} triangles p{porder}", "elements": ne, "faces": nf,
            "ms_device": dt * 1e3, "ms_host_numpy": dth * 1e3, "equal_to_host": same,
            "note": "device time includes the vertex-table upload and the two host syncs"}


def host_leg(torch, ctx, P, reps=2):
    """hdgx_elemental on host (pageable numpy) arrays, HDGX_HOST_PTRS: the
    rate a caller handing over host buffers sees, transfers included."""
    import ctypes
    from hdg_amd.app import MasterC
    from hdg_amd.runtime import lib
    m, a = P.master, P.app
    npv, nc, ne, nh = m.npv, a.nc, P.ne, a.ncu * m.ndf
    z = lambda *sh: np.zeros(sh, order="F")
    out = [z(nh, nh, ne), z(nh, ne), z(npv, nc, ne), z(npv, nc, nh, ne), np.zeros(ne, dtype=np.int32)]
    ins = [np.asfortranarray(x) for x in (P.mesh.dgnodes, P.udg, P.uh, P.sh)]
    bf = np.asfortranarray(P.mesh.bf.astype(np.int32))
    mc, ca = MasterC(m), a.to_c()
    p = lambda x: x.ctypes.data_as(ctypes.c_void_p)
    call = lambda: lib().hdgx_elemental(ctx.h, ctypes.byref(mc.c), ctypes.byref(ca), ne, p(ins[0]), p(bf), p(ins[1]),
                                        p(ins[2]), p(ins[3]), *[p(o) for o in out])
    ctx.set_flags(2)                      # HDGX_HOST_PTRS
    try:
        ctx._check(call())
        t0 = time.perf_counter()
        for _ in range(reps):
            ctx._check(call())
        dt = (time.perf_counter() - t0) / reps
    finally:
        ctx.set_flags(0)
    nbytes = sum(x.nbytes for x in ins) + bf.nbytes + sum(o.nbytes for o in out)
    log(f"[bench] host buffers: {dt * 1e3:.1f} ms/step, {ne / dt:.4g} elements/s, {nbytes / dt / 1e9:.1f} GB/s over PCIe")
    return {"workload": "hdgx_elemental on pageable host arrays (HDGX_HOST_PTRS), same elements",
            "ms_per_step": dt * 1e3, "value": ne / dt, "unit": "elements/s", "bytes_per_step": nbytes,
            "host_link_gbs": nbytes / dt / 1e9}


def serving_leg(nreq, seed):
    """Poisson stream of BDF1 Navier-Stokes Newton steps through the
    continuous-batching loop: p50/p99 latency and throughput at 50 requests/s,
    then the same requests all at once (saturation), where the loop merges
    the requests of every order into shared elemental calls.  Each request's
    mesh connectivity and initial trace are built on the device on admission,
    inside its latency."""
    from hdg_amd.serve import PayloadFactory, Server, make_workload
    specs = make_workload(nreq, seed=seed, rate=50.0)
    fac = PayloadFactory()
    t0 = time.perf_counter()
    for sp in specs:                       # the clients' messages (geometry, state), before the clock starts
        fac.geometry(sp.n, sp.p)
    log(f"[bench] serving: {nreq} requests, messages ready in {time.perf_counter() - t0:.1f}s")
    # untimed warm-up: one request of each order through the loop first, so the
    # measured run does not include loading the kernels' code objects
    warm = []
    for p in (1, 2, 3, 4):
        sp = next((x for x in specs if x.p == p), None)
        if sp is not None:
            warm.append(sp)
    Server(workers=4, slots=64).run(warm, factory=fac, realtime=False)
    srv = Server(workers=4, slots=64)
    stats, _ = srv.run(specs, factory=fac)
    log(f"[bench] serving: p50 {stats['p50_ms']:.1f} ms, p99 {stats['p99_ms']:.1f} ms, "
        f"{stats['requests_per_s']:.1f} req/s, {stats['elements_per_s']:.3g} el/s, "
        f"{stats['mean_requests_per_batch']:.2f} requests per elemental call")
    stats["workload"] = (f"{nreq} Poisson(50/s) BDF1 NS Newton steps, 2n^2 triangles with ne log-uniform "
                         f"500..50000 (16 mesh sizes), p uniform 1..4, M uniform 0.2..0.7, dt at acoustic CFL 10, "
                         f"GMRES(20)+BILU0 tol 1e-6; connectivity built on the device on admission; "
                         f"after an untimed warm-up of one request per order")
    sat, _ = Server(workers=4, slots=64).run(specs, factory=fac, realtime=False)
    log(f"[bench] serving (all at once): {sat['requests_per_s']:.1f} req/s, {sat['elements_per_s']:.3g} el/s, "
        f"{sat['mean_requests_per_batch']:.2f} requests per elemental call, p50 {sat['p50_ms']:.0f} ms")
    stats["saturation"] = {k: sat[k] for k in ("requests", "requests_per_s", "elements_per_s", "elemental_batches",
                                               "mean_requests_per_batch", "mixed_order_batches", "p50_ms", "p99_ms",
                                               "makespan_s", "admission_ms_mean")}
    stats["saturation"]["workload"] = "the same requests admitted as soon as slots and HBM allow (no arrival clock)"
    return stats


def uneven_leg(torch, total, seed, steps):
    """Mixed-order batch: `total` 2D Euler triangles whose orders p = 1..4 are
    drawn 40/30/20/10% from the seed, as four jittered rectangles; launch plan
    chosen by the autotuner among per-order, concurrent and fused persistent
    launches (hipGraph replays); the FP32 + FP64-refinement form beside it."""
    from hdg_amd.autotune import time_plan, tune
    from hdg_amd.mesh import mkmesh_square
    from hdg_amd.problems import euler2d_square
    from hdg_amd.runtime import Context, Dev, PREC_FP32_IR
    rng = np.random.default_rng(np.random.SeedSequence([seed, 3]))
    counts = rng.multinomial(total, [0.4, 0.3, 0.2, 0.1])
    t0 = time.perf_counter()
    probs = []
    for p, n_p in zip((1, 2, 3, 4), counts):
        nx = max(1, int(round(np.sqrt(n_p / 2))))
        ny = max(1, int(round(n_p / (2 * nx))))
        P = euler2d_square(1, p, seed=seed + p, minf=0.5)
        m = P.master
        from hdg_amd.problems import _finish, euler_state
        mesh = mkmesh_square(nx, ny, m, 0.1, seed + p)
        udg = euler_state(mesh.dgnodes, 2, 1.4, 0.5, seed + 10 + p)
        probs.append(_finish(m, P.app, mesh, udg, f"euler2d_p{p}_{nx}x{ny}"))
    ne_tot = sum(P.ne for P in probs)
    log(f"[bench] uneven: {ne_tot} elements ({[P.ne for P in probs]}) built in {time.perf_counter() - t0:.1f}s")
    ctx = Context(0)
    masters, segs = [], []
    for P in probs:
        npv, ncu, nc, ndf, ne = P.master.npv, P.app.ncu, P.app.nc, P.master.ndf, P.ne
        dev = (Dev.from_numpy(P.mesh.dgnodes), Dev.from_numpy(P.mesh.bf, np.int32), Dev.from_numpy(P.udg),
               Dev.from_numpy(P.uh), Dev.from_numpy(P.sh))
        out = (Dev.empty((ncu, ndf, ncu, ndf, ne)), Dev.empty((ncu, ndf, ne)), Dev.empty((npv, nc, ne)),
               Dev.empty((npv, nc, ncu, ndf, ne)), Dev.empty((ne,), "int32"))
        masters.append(P.master)
        segs.append((P.app, *dev, out))
    plan, timings = tune(ctx, masters, segs, reps=3, log=log)
    ms = time_plan(ctx, plan, reps=steps)
    fp64 = [[o.numpy() for o in seg[6][:4]] for seg in segs]
    bad = int(sum(np.count_nonzero(seg[6][4].numpy()) for seg in segs))
    plan.close()
    # FP32 + refinement, one launch per order
    ctx.set_precision(PREC_FP32_IR)

    def step32():
        for P, seg in zip(probs, segs):
            ctx.elemental(P.master, P.app, *seg[1:6], out=seg[6])
    ms32, _ = time_steps(torch, ctx.stream, step32, steps, 1)
    errs = []
    for seg, ref in zip(segs, fp64):
        errs += [relerr(o.numpy(), r) for o, r in zip(seg[6][:4], ref)]
    ctx.set_precision(0)
    best = min(timings, key=timings.get)
    res = {"workload": f"euler2d_mixed_p1-4_{ne_tot}", "elements": ne_tot,
           "elements_per_order": {f"p{p}": int(P.ne) for p, P in zip((1, 2, 3, 4), probs)},
           "ms_per_step": ms, "value": ne_tot / (ms * 1e-3), "unit": "elements/s", "plan": best,
           "candidates_ms": timings, "status_nonzero": bad,
           "fp32_ir": {"ms_per_step": ms32, "value": ne_tot / (ms32 * 1e-3),
                       "max_rel_err_vs_fp64": max(errs)}}
    log(f"[bench] uneven: {best} {ms:.3f} ms/step, {ne_tot / ms * 1e3:.4g} elements/s; fp32_ir {ms32:.3f} ms")
    del segs
    torch.cuda.empty_cache()
    return res


def tets_leg(torch, n, seed, steps):
    """3D Euler p=2 on 6 n^3 jittered tetrahedra (u-only): elemental throughput
    in FP64 and FP32 + FP64 refinement, with the outputs resident in HBM.  The
    step runs in element-matrix workspace chunks; a sample of elements -- both
    sides of every chunk boundary and 64 drawn from the seed -- is checked
    against the oracle (FP64) and its floor, the oracle's response to +-1 ulp
    in node coordinates, state and trace.  On the far-field faces parallel to
    the free stream u.n is zero to rounding, where the Roe |A|'s derivative
    takes sign(u.n): coordinates rounded differently flip that sign and move
    H by about 1e-10 there, which the coordinate perturbation puts in the
    floor."""
    from hdg_amd.problems import euler3d_cube
    from hdg_amd.runtime import Context, Dev, PREC_FP32_IR, PREC_FP64
    from oracle import oracle as O
    t0 = time.perf_counter()
    P = euler3d_cube(n, 2, seed=seed)
    log(f"[bench] tets: {P.ne} elements built in {time.perf_counter() - t0:.1f}s")
    ctx = Context(0)
    dev = (Dev.from_numpy(P.mesh.dgnodes), Dev.from_numpy(P.mesh.bf, np.int32), Dev.from_numpy(P.udg),
           Dev.from_numpy(P.uh), None)
    npv, ncu, nc, ndf, ne = P.master.npv, P.app.ncu, P.app.nc, P.master.ndf, P.ne
    out = (Dev.empty((ncu, ndf, ncu, ndf, ne)), Dev.empty((ncu, ndf, ne)), Dev.empty((npv, nc, ne)),
           Dev.empty((npv, nc, ncu, ndf, ne)), Dev.empty((ne,), "int32"))
    # chunk boundaries of the default 24 GiB workspace (as hdgx_ws_chunk)
    per = 8 * (ncu * npv * npv * nc + ncu * npv * ncu * ndf + ncu * ndf * npv * nc + (ncu * ndf) ** 2
               + ncu * npv + ncu * ndf)
    cmax = max(1, min(ne, (24 << 30) // per))
    nch = -(-ne // cmax)
    chunk = -(-ne // nch)
    rng = np.random.default_rng(seed)
    idx = set(rng.choice(ne, size=min(64, ne), replace=False).tolist())
    for b in range(chunk, ne, chunk):
        idx.update({b - 2, b - 1, b, b + 1})
    idx.update({0, ne - 1})
    idx = np.array(sorted(i for i in idx if 0 <= i < ne))
    sl = lambda a: np.asfortranarray(a[..., idx])
    args = (P.master, P.app, sl(P.mesh.dgnodes), sl(P.mesh.bf), sl(P.udg), sl(P.uh), sl(P.sh))
    ref = O.elemental(*args)
    ulp = lambda a: np.asfortranarray(a * (1.0 + np.ldexp(1.0, -52) * rng.choice([-1.0, 1.0], size=a.shape)))
    dg1 = args[2].copy(order="F")
    dg1[:, :3, :] = ulp(dg1[:, :3, :])
    pert = O.elemental(P.master, P.app, dg1, args[3], ulp(args[4]), ulp(args[5]), args[6])
    names = ("AE", "FE", "DUDG", "DUDG_DUH")
    floor = {nm: relerr(pert[i], ref[i]) for i, nm in enumerate(names)}
    res = {"workload": f"euler3d_hdg_elemental_p2_6x{n}^3", "elements": ne, "unit": "elements/s",
           "workspace_chunks": nch, "chunk_elements": chunk,
           "oracle_sample": f"{idx.size} elements: both sides of each chunk boundary, the first, the last, 64 random",
           "sample_noise_floor": floor}
    for prec, name in ((PREC_FP64, "fp64"), (PREC_FP32_IR, "fp32_ir")):
        ctx.set_precision(prec)
        ms, wall = time_steps(torch, ctx.stream, lambda: ctx.elemental(P.master, P.app, *dev, out=out), steps, 1)
        sel = torch.from_numpy(idx).cuda()
        got = [o.t.view(ne, -1)[sel].cpu().numpy().T for o in out[:4]]      # (per element, sample)
        errs = {nm: relerr(g.reshape(r.shape, order="F"), r) for nm, g, r in zip(names, got, ref[:4])}
        bad = int(np.count_nonzero(out[4].numpy()))
        entry = {"ms_per_step": ms, "value": ne / (ms * 1e-3), "status_nonzero": bad, "rel_err_sample_vs_oracle": errs}
        if name == "fp64":
            res.update(entry)
        else:
            res["fp32_ir"] = entry
            res["fp32_ir"]["speedup_vs_fp64"] = res["ms_per_step"] / ms
        log(f"[bench] tets {name}: {ms:.2f} ms/step, {ne / ms * 1e3:.4g} elements/s, sample errors {errs}")
    ctx.set_precision(PREC_FP64)
    ctx.set_timing(True)
    ctx.elemental(P.master, P.app, *dev, out=out)
    tq, ts, nl = ctx.kernel_times()
    ctx.set_timing(False)
    res["kernels_ms"] = {"quad_kernel": tq, "elimination": ts, "launch_pairs": nl}
    del out, dev
    torch.cuda.empty_cache()
    return res


def euler_c1_leg(torch, seed, steps, warmup):
    """2D Euler p=2 on the jittered 2x32x32 square (2048 triangles, u-only,
    schur_u), FP64 and FP32 + refinement, each against the oracle."""
    from hdg_amd.problems import euler2d_square
    from hdg_amd.runtime import Context, Dev, PREC_FP32_IR, PREC_FP64
    from hdg_amd.workload import problem_work
    from oracle import oracle as O
    P = euler2d_square(32, 2, seed=seed)
    W = problem_work(P)
    ctx = Context(0)
    dev = (Dev.from_numpy(P.mesh.dgnodes), Dev.from_numpy(P.mesh.bf, np.int32), Dev.from_numpy(P.udg),
           Dev.from_numpy(P.uh), Dev.from_numpy(P.sh))
    npv, ncu, nc, ndf, ne = P.master.npv, P.app.ncu, P.app.nc, P.master.ndf, P.ne
    out = (Dev.empty((ncu, ndf, ncu, ndf, ne)), Dev.empty((ncu, ndf, ne)), Dev.empty((npv, nc, ne)),
           Dev.empty((npv, nc, ncu, ndf, ne)), Dev.empty((ne,), "int32"))
    ref = O.elemental(P.master, P.app, P.mesh.dgnodes, P.mesh.bf, P.udg, P.uh, P.sh)
    names = ("AE", "FE", "DUDG", "DUDG_DUH")
    res = {"workload": "euler2d_hdg_elemental_p2_2x32x32", "elements": ne, "unit": "elements/s",
           "bytes_per_element": W.bytes, "flops_per_element": W.flops}
    for prec, name in ((PREC_FP64, "fp64"), (PREC_FP32_IR, "fp32_ir")):
        ctx.set_precision(prec)
        ms, _ = time_steps(torch, ctx.stream, lambda: ctx.elemental(P.master, P.app, *dev, out=out), steps, warmup)
        errs = {nm: relerr(o.numpy(), r) for nm, o, r in zip(names, out[:4], ref[:4])}
        entry = {"ms_per_step": ms, "value": ne / (ms * 1e-3), "rel_err_vs_oracle": errs,
                 "hbm_gbs_algorithmic": W.bytes * ne / (ms * 1e-3) / 1e9}
        if name == "fp64":
            res.update(entry)
        else:
            res["fp32_ir"] = entry
        log(f"[bench] C1 {name}: {ms * 1e3:.1f} us/step, {ne / ms * 1e3:.4g} elements/s, errors {errs}")
    ctx.set_precision(PREC_FP64)
    return res


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=20)
    ap.add_argument("--warmup", type=int, default=3)
    ap.add_argument("--n", type=int, default=128, help="cells per side (2*n*n triangles)")
    ap.add_argument("--seed", type=int, default=2024)
    ap.add_argument("--serve-requests", type=int, default=512,
                    help="requests of the Poisson serving workload (0 skips it)")
    ap.add_argument("--tets", type=int, default=40, help="cube cells per side of the 3D Euler line (0 skips it)")
    ap.add_argument("--uneven", type=int, default=200000, help="elements of the mixed-order batch (0 skips it)")
    ap.add_argument("--copy-calibration", type=float, default=0.0,
                    help="only copy this many GiB device to device 3 times (calibrates the HBM byte counters)")
    a = ap.parse_args()
    if a.copy_calibration > 0:
        import torch
        n = int(a.copy_calibration * (1 << 30)) // 8
        src = torch.ones(n, dtype=torch.float64, device="cuda")
        dst = torch.empty_like(src)
        for _ in range(3):
            dst.copy_(src)
        torch.cuda.synchronize()
        print(json.dumps({"copy_bytes_read": n * 8, "copy_bytes_written": n * 8, "copies": 3}), flush=True)
        return
    if a.gpus != 1 or int(os.environ.get("WORLD_SIZE", "1")) != 1:
        print("bench.py measures one GPU: the elemental step needs no communication and every box here has "
              "one GPU; run with --gpus 1", file=sys.stderr)
        sys.exit(2)

    import torch
    from hdg_amd.problems import ns2d_square
    from hdg_amd.runtime import Context, Dev, PREC_FP32_IR, PREC_FP64
    from hdg_amd.workload import problem_work

    t0 = time.perf_counter()
    P = ns2d_square(a.n, 3, seed=a.seed)
    W = problem_work(P)
    log(f"[bench] workload ns2d p=3 ne={P.ne} built in {time.perf_counter() - t0:.1f}s")

    ctx = Context(0)
    stream = ctx.stream
    dev = (Dev.from_numpy(P.mesh.dgnodes), Dev.from_numpy(P.mesh.bf, np.int32), Dev.from_numpy(P.udg),
           Dev.from_numpy(P.uh), Dev.from_numpy(P.sh))
    npv, ncu, nc, ndf, ne = P.master.npv, P.app.ncu, P.app.nc, P.master.ndf, P.ne
    out = (Dev.empty((ncu, ndf, ncu, ndf, ne)), Dev.empty((ncu, ndf, ne)), Dev.empty((npv, nc, ne)),
           Dev.empty((npv, nc, ncu, ndf, ne)), Dev.empty((ne,), "int32"))

    def step():
        ctx.elemental(P.master, P.app, *dev, out=out)

    res = {}
    outputs = {}
    for prec, name in ((PREC_FP64, "fp64"), (PREC_FP32_IR, "fp32_ir")):
        ctx.set_precision(prec)
        ms, wall = time_steps(torch, stream, step, a.steps, a.warmup)
        res[name] = (ms, wall)
        outputs[name] = [o.numpy() for o in out]
        log(f"[bench] {name}: {ms:.3f} ms/step, {ne / ms * 1e3:.4g} elements/s (wall {wall:.2f}s)")
    ctx.set_precision(PREC_FP64)

    tq, ts = kernel_times(torch, ctx, P, dev, out, reps=5)
    log(f"[bench] kernels (fp64, events inside hdgx_elemental): quad {tq:.3f} ms, elimination {ts:.3f} ms")
    resid = residual_leg(torch, ctx, P, dev, a.steps, a.warmup)
    getq = getq_leg(torch, ctx, P, dev, a.steps, a.warmup)
    host = host_leg(torch, ctx, P)
    mesh = mesh_leg(torch, ctx)

    ncpu = len(os.sched_getaffinity(0)) if hasattr(os, "sched_getaffinity") else os.cpu_count()
    ref, cpu_rate, cpu_rate1, build_kind, cpu_secs, noise_floor = cpu_baseline(P, ncpu)
    log(f"[bench] cpu oracle: {cpu_rate:.4g} elements/s on {ncpu} threads, {cpu_rate1:.4g} on 1 ({cpu_secs:.1f}s)")

    names = ("AE", "FE", "DUDG", "DUDG_DUH")
    errs = {p: {n: relerr(outputs[p][i], ref[i]) for i, n in enumerate(names)} for p in outputs}
    status_bad = {p: int(np.count_nonzero(outputs[p][4])) for p in outputs}

    ms64 = res["fp64"][0]
    ms32 = res["fp32_ir"][0]
    value = ne / (ms64 * 1e-3)
    # dominant kernel roofline (FP64 element Schur + quadrature are compute bound at this size)
    if ts >= tq:
        kname, kflops, kms = "schur64_kernel<NS2D p3, tile layout>", W.schur_flops * ne, ts
    else:
        kname, kflops, kms = "quad_kernel<NS2D p3, tile layout>", W.quad_flops * ne, tq
    achieved = kflops / (kms * 1e-3) / 1e12
    line = {
        "metric": METRIC,
        "value": value,
        "unit": "elements/s",
        "n_gpus": 1,
        "steps": a.steps,
        "warmup": a.warmup,
        "ms_per_step": ms64,
        "higher_is_better": True,
        "scaling": "weak",
        "vs_baseline": None,
        "dtype": "f64",
        "data": "synthetic (seeded jittered mesh and flow state)",
        "config": {"workload": f"ns2d_hdg_elemental_p3_2x{a.n}x{a.n}", "elements": ne, "porder": 3,
                   "nc": nc, "method": "u,q (schur_uq)", "parallelism": "single GPU"},
        "fp32_ir": {"value": ne / (ms32 * 1e-3), "ms_per_step": ms32, "speedup_vs_fp64": ms64 / ms32,
                    "max_rel_err_vs_oracle": max(errs["fp32_ir"].values()), "rel_err": errs["fp32_ir"]},
        "fp64_rel_err_vs_oracle": errs["fp64"],
        "fp64_noise_floor": noise_floor,
        "within_noise_floor": {p: {n: errs[p][n] <= max(1e-12 if p == "fp64" else 1e-10, noise_floor[n])
                                   for n in names} for p in errs},
        "status_nonzero": status_bad,
        "step_rates": {"flops_tflops_fp64": W.flops * ne / (ms64 * 1e-3) / 1e12,
                       "hbm_gbs_algorithmic_fp64": W.bytes * ne / (ms64 * 1e-3) / 1e9,
                       "bytes_per_element": W.bytes, "flops_per_element": W.flops},
        "kernels_ms": {"quad_kernel": tq, "elimination_kernel": ts},
        "roofline": {"kernel": kname, "bound": "mfma", "achieved": achieved, "peak": PEAK_FP64_TFLOPS,
                     "unit": "TFLOP/s", "frac": achieved / PEAK_FP64_TFLOPS, "traffic": measured_traffic(kname),
                     "flops_per_element": (W.schur_flops if ts >= tq else W.quad_flops)},
        "cpu_baseline": {"value": cpu_rate, "unit": "elements/s", "cores": ncpu, "kind": "port",
                         "sample": f"full workload ({ne} elements), oracle FP64 C++ restatement, "
                                   f"{build_kind} build, OpenMP over elements, median of 5 after 1 warm-up",
                         "single_core_value": cpu_rate1, "single_core_sample": f"first {min(2048, ne)} elements"},
        "speedup_vs_cpu": value / cpu_rate,
        "residual": resid,
        "getq": getq,
        "host_buffers": host,
        "mesh_preprocess": mesh,
    }
    del out, dev
    torch.cuda.empty_cache()
    line["c1_euler2d_p2"] = euler_c1_leg(torch, a.seed, max(a.steps, 20), a.warmup)
    if a.serve_requests > 0:
        line["serving"] = serving_leg(a.serve_requests, a.seed)
    if a.uneven > 0:
        line["uneven_batch"] = uneven_leg(torch, a.uneven, a.seed, min(a.steps, 5))
    if a.tets > 0:
        line["tets"] = tets_leg(torch, a.tets, a.seed, min(a.steps, 5))
    print(json.dumps(line), flush=True)


if __name__ == "__main__":
    main()
