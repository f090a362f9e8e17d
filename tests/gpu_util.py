"""Helpers for the GPU parity tests."""
import numpy as np


def relerr(x, ref):
    x = np.asarray(x, dtype=np.float64)
    ref = np.asarray(ref, dtype=np.float64)
    n = np.linalg.norm(ref.ravel())
    return np.linalg.norm((x - ref).ravel()) / (n if n > 0 else 1.0)


def perturb(P, seed, amp=0.02):
    rng = np.random.default_rng(seed)
    P.udg = np.asfortranarray(P.udg * (1 + amp * rng.standard_normal(P.udg.shape)))
    P.uh = np.asfortranarray(P.uh * (1 + amp * rng.standard_normal(P.uh.shape)))
    return P


def ulp_perturb(a, rng):
    """a with every entry moved by one unit in the last place, up or down at
    random: the rounding any FP64 evaluation may make in its first step."""
    return np.asfortranarray(a * (1.0 + rng.choice([-1.0, 1.0], size=a.shape) * 2.0 ** -52))


def local_system(P, m, e):
    """Element e's local system as the reference forms it (kernel/schur_primal.m,
    kernel/schur_uq.c): A [q; u] = R for the u and q unknowns (rows: the q
    equations M q - C u = Rq, then the u equations D u + B q = Ru), or
    BD u = Ru without q; F the trace columns.  m holds the element matrices
    (oracle or hdgx_elemental_matrices, as numpy arrays).  Returns (A, R, F)
    with the unknowns in DUDG's order [u rows, q rows]."""
    npv, ncu, nc = P.master.npv, P.app.ncu, P.app.nc
    ndf = P.master.npf * P.master.nfe
    nu, nh = npv * ncu, ncu * ndf
    BD = m["BD"][..., e].reshape(nu, npv * nc, order="F")
    F = m["F"][..., e].reshape(nu, nh, order="F")
    Ru = m["Ru"][..., e].reshape(nu, order="F")
    if not P.app.localsolve:
        return BD, Ru, F
    nd = P.master.nd
    M = np.kron(np.eye(ncu * nd), m["M"][:, :, e])
    C = np.vstack([np.kron(np.eye(ncu), m["C"][:, :, d, e]) for d in range(nd)])
    E = np.zeros((nu * nd, nh))
    for d in range(nd):
        for c in range(ncu):
            E[d * nu + c * npv:d * nu + (c + 1) * npv, c::ncu] = m["E"][:, :, d, e]
    Rq = m["Rq"][..., e].reshape(nu * nd, order="F")
    A = np.block([[-C, M], [BD[:, :nu], BD[:, nu:]]])
    return A, np.concatenate([Rq, Ru]), np.vstack([E, F])


def backward_error(P, m, dudg, dudg_duh):
    """Normwise backward error of the local solutions over the whole mesh,
    sqrt(sum_e |R_e - A_e x_e|^2) / (sqrt(sum_e (|A_e| |x_e|)^2) + |R|), for
    x = DUDG (right-hand side R) and for every DUDG_DUH column (right-hand
    side -F).  Returns (eta_dudg, eta_dudg_duh)."""
    num = den = rr = 0.0
    numD = denD = ff = 0.0
    for e in range(P.ne):
        A, R, Fl = local_system(P, m, e)
        nA = np.linalg.norm(A, 2)
        x = dudg[:, :, e].reshape(-1, order="F")
        r = R - A @ x
        num += r @ r
        den += (nA * np.linalg.norm(x)) ** 2
        rr += R @ R
        X = dudg_duh[..., e].reshape(A.shape[0], -1, order="F")
        rD = -Fl - A @ X
        numD += np.sum(rD * rD)
        denD += (nA * np.linalg.norm(X)) ** 2
        ff += np.sum(Fl * Fl)
    return np.sqrt(num) / (np.sqrt(den) + np.sqrt(rr)), np.sqrt(numD) / (np.sqrt(denD) + np.sqrt(ff))
