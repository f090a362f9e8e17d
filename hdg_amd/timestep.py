"""Time integration around the damped Newton solve: BDF steps
(kernel/hdg_solve_bdf.m) and DIRK steps (kernel/hdg_solve_dirk.m) on a
HdgState.

A step sets the time term of the elemental operator -- fc_u, fc_q, the
source SH and the time, which volint.m:180-221 turns into s += SH - fc_u u --
runs the Newton solve (hdg_solve.m) and combines whole state vectors.  The
state is used only through ``app``, ``UDG.t`` / ``UH.t`` (flat arrays),
``set_source``, ``getq`` and ``solve``, and the combinations need only +, *
and a copy, so the stage logic runs unchanged on numpy arrays with a scalar
stand-in for the solve.
"""
import math

import numpy as np


def bdfcoeff(p):
    """BDF coefficients a (hdg_solve_bdf.m:64-75): u'(t_{n+1}) ~
    (a0 u_{n+1} + a1 u_n + ... + ap u_{n+1-p}) / dt."""
    table = {1: [1.0, -1.0],
             2: [3.0 / 2, -4.0 / 2, 1.0 / 2],
             3: [11.0 / 6, -18.0 / 6, 9.0 / 6, -2.0 / 6],
             4: [25.0 / 12, -48.0 / 12, 36.0 / 12, -16.0 / 12, 3.0 / 12]}
    if p not in table:
        raise ValueError("BDF order must be 1..4")
    return np.array(table[p])


def dirk_tableau(q, p):
    """Butcher tableau (a, b, t) of the q-stage DIRK scheme of order p
    (hdg_solve_dirk.m:113-186)."""
    s2, s3, s6 = math.sqrt(2.0), math.sqrt(3.0), math.sqrt(6.0)
    if (q, p) == (1, 1):
        a, b, t = [[1.0]], [1.0], [1.0]
    elif (q, p) == (1, 2):
        a, b, t = [[0.5]], [1.0], [0.5]
    elif (q, p) == (2, 2):
        g = 1 - 0.5 * s2
        a, b, t = [[g, 0], [0.5 * s2, g]], [0.5 * s2, g], [g, 1.0]
    elif (q, p) == (2, 3):
        g = 0.5 + 0.5 / s3
        a, b, t = [[g, 0], [-1 / s3, g]], [0.5, 0.5], [g, 0.5 - 0.5 / s3]
    elif (q, p) == (3, 3):
        a1 = 0.4358665215
        t1 = (1 + a1) / 2
        b1 = -(6 * a1 ** 2 - 16 * a1 + 1) / 4
        b2 = (6 * a1 ** 2 - 20 * a1 + 5) / 4
        a, b, t = [[a1, 0, 0], [t1 - a1, a1, 0], [b1, b2, a1]], [b1, b2, a1], [a1, t1, 1.0]
    elif (q, p) == (3, 4):
        a1 = 2 * math.cos(math.pi / 18) / s3
        h = 0.5 * (1 + a1)
        a = [[h, 0, 0], [-0.5 * a1, h, 0], [1 + a1, -(1 + 2 * a1), h]]
        b = [1 / (6 * a1 ** 2), 1 - 1 / (3 * a1 ** 2), 1 / (6 * a1 ** 2)]
        t = [h, 0.5, 0.5 * (1 - a1)]
    elif (q, p) == (5, 5):
        a11 = (6 - s6) / 10
        a = [[a11, 0, 0, 0, 0],
             [(-6 + 5 * s6) / 14, a11, 0, 0, 0],
             [(888 + 607 * s6) / 2850, (126 - 161 * s6) / 1425, a11, 0, 0],
             [(3153 - 3082 * s6) / 14250, (3213 + 1148 * s6) / 28500, (-267 + 88 * s6) / 500, a11, 0],
             [(-32583 + 14638 * s6) / 71250, (-17199 + 364 * s6) / 142500, (1329 - 544 * s6) / 2500,
              (-96 + 131 * s6) / 625, a11]]
        b = [0.0, 0.0, 1.0 / 9, (16 - s6) / 36, (16 + s6) / 36]
        t = [a11, (6 + 9 * s6) / 35, 1.0, (4 - s6) / 10, (4 + s6) / 10]
    else:
        raise ValueError("invalid (stages, order) combination")
    return np.array(a, dtype=float), np.array(b, dtype=float), np.array(t, dtype=float)


def dirkcoeff(q, p):
    """(d, c, t) of hdg_solve_dirk.m:186-187: d = inv(a), c = b d."""
    a, b, t = dirk_tableau(q, p)
    d = np.linalg.inv(a)
    return d, b @ d, t


def _copy(x):
    return x.clone() if hasattr(x, "clone") else x.copy()


def _set_time_term(state, time, fc_u, fc_q):
    state.app.time = time
    state.app.fc_u = fc_u
    state.app.fc_q = fc_q
    state.app.tdep = 1


def hdg_solve_bdf(state, UDG0, time, dt, torder, **newton):
    """One BDF step of order torder (hdg_solve_bdf.m:1-39) from the current
    state (the initial guess): fc_u = a0/dt, fc_q = 1, SH = -(a1 UDG0[0] + ...
    + a_torder UDG0[torder-1])/dt, then the Newton solve.  UDG0 is the list of
    previous UDG vectors, most recent first (at least torder of them); the
    returned list is shifted by one with the new state in front.  Returns
    (Newton statistics, history)."""
    if len(UDG0) < torder:
        raise ValueError("BDF history shorter than the order")
    a = bdfcoeff(torder)
    SH = a[1] * UDG0[0]
    for i in range(2, torder + 1):
        SH = SH + a[i] * UDG0[i - 1]
    SH = -SH / dt
    _set_time_term(state, time, a[0] / dt, 1.0)
    state.set_source(SH)
    stats = state.solve(**newton)
    return stats, [_copy(state.UDG.t)] + list(UDG0[:-1])


def hdg_solve_dirk(state, time, dt, nstage, torder, **newton):
    """One DIRK step (hdg_solve_dirk.m:1-107): stage i solves
    d_ii U_i - SDG_i = R(U_i) at time + t_i dt with
    SDG_i = (d_i1 + ... + d_ii) U_n - (d_i1 U_1 + ... + d_i,i-1 U_i-1)
    (d = inv(a)/dt), each stage starting from U_n with q from getq; then
    U_n+1 = (1 - sum c) U_n + sum c_i U_i for both UDG and UH.

    hdg_solve_dirk.m spells SDG out for stages 1-3 (stage 3 rebuilds d_31 U_1
    from stage 2's SDG instead of keeping U_1); the general form above is the
    same for those and is also right for the 5-stage scheme, where the
    reference would reuse stage 3's SDG.  Returns the per-stage Newton
    statistics; the state holds U_n+1."""
    d, c, t = dirkcoeff(nstage, torder)
    d = d / dt
    U0, H0 = _copy(state.UDG.t), _copy(state.UH.t)
    w0 = 1.0 - float(c.sum())
    Un, Hn = w0 * U0, w0 * H0
    stages, stats = [], []
    for i in range(nstage):
        SDG = float(d[i, :i + 1].sum()) * U0
        for j in range(i):
            SDG = SDG - float(d[i, j]) * stages[j]
        state.UDG.t[...] = U0
        state.UH.t[...] = H0
        _set_time_term(state, time + dt * float(t[i]), float(d[i, i]), 1.0)
        state.getq()
        state.set_source(SDG)
        stats.append(state.solve(**newton))
        Ui = _copy(state.UDG.t)
        stages.append(Ui)
        Un = Un + float(c[i]) * Ui
        Hn = Hn + float(c[i]) * state.UH.t
    state.UDG.t[...] = Un
    state.UH.t[...] = Hn
    return stats
