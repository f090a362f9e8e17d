"""Damped Newton solve of the HDG system on the device (kernel/hdg_solve.m:1-170
with app.denseblock = 1) and the BDF time step around it
(kernel/hdg_solve_bdf.m:26-48).

Every array stays in HBM: the elemental operator (hdgx_elemental), the
face-block trace system (hdgx_trace_assemble), GMRES(20) with block ILU(0)
(hdgx_gmres), the local update (hdgx_local_update) and the damped step
(hdgx_newton_update) run on the context's stream; the host only sees the
residual norms that steer the iteration.
"""
from dataclasses import dataclass, field

import numpy as np

from .runtime import Dev, HdgxError
from .trace import TraceSystem


@dataclass
class NewtonStats:
    residuals: list = field(default_factory=list)    # |F| after every accepted step (index 0: initial)
    gmres_iters: list = field(default_factory=list)
    alphas: list = field(default_factory=list)


class HdgState:
    """Device arrays of one HDG problem: mesh data, state and the buffers of
    one Newton iteration."""

    def __init__(self, ctx, master, app, mesh, UDG, UH, SH=None, with_bilu0=True):
        self.ctx, self.master, self.app, self.mesh = ctx, master, app, mesh
        npv, nc, ncu = master.npv, app.nc, app.ncu
        ne = mesh.ne
        ndf = master.npf * master.nfe
        self.npv, self.nc, self.ncu, self.ne, self.ndf = npv, nc, ncu, ne, ndf
        # rows of DUDG kept by the elemental step (hdg_solve.m:56-60); with
        # getdqdg = 0 q is recovered from (u, uh) by getq after every step;
        # any other value keeps no DUDG / DUDG_DUH and the local update solves
        # the element problems again (hdg_local.m, hdg_solve.m:122-124)
        self.getdqdg = app.getdqdg
        self.local = app.getdqdg not in (0, 1)
        self.nco = ncu if app.getdqdg == 0 else nc
        nco = self.nco
        self.ts = TraceSystem(ctx, mesh, ncu, with_bilu0=with_bilu0)
        self.dgnodes = Dev.from_numpy(mesh.dgnodes)
        self.bf = Dev.from_numpy(mesh.bf, np.int32)
        self.UDG = Dev.from_numpy(UDG)
        self.UH = Dev.from_numpy(UH)
        sh = np.zeros((npv, nc, ne)) if SH is None else np.array(SH, dtype=np.float64)
        sh[:, ncu:, :] = 0.0                             # app.wave = 0 (hdg_solve.m:63-64)
        self.SH = Dev.from_numpy(sh)
        self.uhe = Dev.empty((ncu, ndf, ne))
        self.AE = Dev.empty((ncu, ndf, ncu, ndf, ne))
        self.FE = Dev.empty((ncu, ndf, ne))
        self.DUDG = None if self.local else Dev.empty((npv, nco, ne))
        self.DUDG_DUH = None if self.local else Dev.empty((npv, nco, ncu, ndf, ne))
        self.est = Dev.empty((ne,), "int32")
        ts = self.ts
        self.K = Dev.empty((ts.ncf, ts.ncf, ts.nbf - 1, ts.nf))
        self.F = Dev.empty((ts.ncf, ts.nf))
        self.fst = Dev.empty((ts.nf,), "int32")
        self.LU = Dev.empty((ts.ncf, ts.ncf, ts.nbf, ts.nf)) if with_bilu0 else None
        self.DUH = Dev.zeros((ts.ncf, ts.nf))
        self.DUDGT = Dev.empty((npv, nco, ne))
        self.UDG0 = Dev.empty((npv, nc, ne))
        self.UH0 = Dev.empty((ts.ncf, ts.nf))

    # hdg_assemble.m:1-73 + mkDenseBlockSystem
    def assemble(self):
        ts = self.ts
        ts.gather(self.UH, self.uhe)
        self.ctx.elemental(self.master, self.app, self.dgnodes, self.bf, self.UDG, self.uhe, self.SH,
                           out=(self.AE, self.FE, self.DUDG, self.DUDG_DUH, self.est))
        ts.assemble(self.AE, self.FE, out=(self.K, self.F, self.fst))

    def residual(self):
        return self.ts.norm2(self.F)

    def trace_solve(self, restart=20, tol=1e-6, maxit=0, precond="bilu0"):
        """DUH = gmres(K, F, f2f) (hdg_solve.m:112-116), optionally block-ILU(0)
        preconditioned.  Returns (iterations, relative residual, flag)."""
        ts = self.ts
        LU = None
        if precond == "bilu0":
            if self.LU is None:
                raise HdgxError("state built without the block ILU(0) colouring")
            ts.bilu0_factor(self.K, self.LU)
            LU = self.LU
        self.DUH.t.zero_()
        _, it, rr, fl = ts.gmres(self.K, self.F, self.DUH, LU=LU, restart=restart, tol=tol, maxit=maxit)
        return it, rr, fl

    def local_update(self):
        if self.local:
            # uhe still holds the gathered trace of this iteration's state
            self.ctx.local_solve(self.master, self.app, self.dgnodes, self.bf, self.UDG, self.uhe, self.SH,
                                 self.ts.elcon, self.DUH, out=self.DUDGT)
        else:
            self.ts.local_update(self.npv, self.nco, self.DUDG, self.DUDG_DUH, self.DUH, self.DUDGT)

    def step(self, alpha, base=True):
        """UDG = UDG0 + alpha DUDGT, UH = UH0 + alpha DUH (hdg_solve.m:145-146);
        base=False updates the current state in place."""
        s0, h0 = (self.UDG0, self.UH0) if base else (self.UDG, self.UH)
        self.ts.newton_update(self.npv, self.nc, self.nco, alpha, s0, self.DUDGT, h0, self.DUH, self.UDG, self.UH)

    def getq(self):
        """UDG(:, ncu+1:end, :) = getq(master, mesh, UDG, UH, [], app.fc_q)
        (hdg_solve.m:130-131, 149-150; hdg_solve_dirk.m:89-90).  No-op without
        q columns (nc = ncu)."""
        if self.nc == self.ncu:
            return
        self.ts.gather(self.UH, self.uhe)
        self.ctx.getq(self.master, self.dgnodes, self.UDG, self.uhe, fc_q=self.app.fc_q, ncu=self.ncu, inplace=True)

    def set_source(self, SH):
        """SH for the next solve, q columns zeroed (hdg_solve.m:63-64, app.wave = 0).
        SH: a torch tensor or Dev of shape (npv, nc, ne)."""
        t = SH.t if isinstance(SH, Dev) else SH
        self.SH.t.copy_(t.reshape(-1))
        self.SH.t.view(self.ne, self.nc, self.npv)[:, self.ncu:, :] = 0.0

    def solve(self, **newton):
        """The damped Newton solve of hdg_solve.m on this state (hdg_solve below)."""
        return hdg_solve(self, **newton)

    def save(self):
        self.UDG0.t.copy_(self.UDG.t)
        self.UH0.t.copy_(self.UH.t)

    def host(self):
        return self.UDG.numpy(), self.UH.numpy().reshape(self.ncu, -1, order="F")


def hdg_solve(state, newton_tol=1e-9, maxiter=16, restart=20, gmres_tol=1e-6, precond="bilu0", min_alpha=1e-2):
    """Damped Newton iteration of kernel/hdg_solve.m:103-170 on a HdgState:
    solve, local update, halve alpha while the residual grows (down to 1e-2),
    stop at |F|/|F0| <= newton_tol or after maxiter steps."""
    st = NewtonStats()
    state.assemble()
    r0 = state.residual()
    st.residuals.append(r0)
    rel = 1.0
    it = 0
    if r0 == 0.0:
        return st
    while rel > newton_tol and it < maxiter:
        git, _, _ = state.trace_solve(restart=restart, tol=gmres_tol, precond=precond)
        st.gmres_iters.append(git)
        state.local_update()
        it += 1
        state.save()
        duh0 = st.residuals[-1]
        alpha = 1.0
        while True:
            state.step(alpha)
            if state.getdqdg == 0:
                state.getq()
            state.assemble()
            duh = state.residual()
            if duh > duh0:
                alpha /= 2.0
                if alpha < min_alpha:
                    break
            else:
                break
        st.alphas.append(alpha)
        st.residuals.append(duh)
        rel = duh / r0
    return st


def bdf1_step(state, dt, restart=20, gmres_tol=1e-6, precond="bilu0"):
    """One Newton iteration of a BDF1 step (hdg_solve_bdf.m:26-48 with
    torder = 1): fc_u = 1/dt, SH = UDG_n/dt, then solve, local update and the
    full step.  app.tdep and app.fc_u must already be set for dt.  Returns
    (gmres iterations, relative GMRES residual)."""
    state.assemble()
    it, rr, _ = state.trace_solve(restart=restart, tol=gmres_tol, precond=precond)
    state.local_update()
    state.step(1.0, base=False)
    return it, rr
