"""Pins the oracle's volint/faceint/Schur restatement without the reference:

* every element matrix is a derivative of an element residual:
  BD = -dRu/dUDG, F = -dRu/dUH, GK = -dRh/dUDG, H = -dRh/dUH and, with q,
  M = -dRq/dq, C = dRq/du, E = -dRq/duh (central differences on tiny meshes);
* a uniform free stream with uh = u has Ru = Rh = Rq = 0 (constant fluxes
  integrate exactly; fbou vanishes at u = uh = uinf);
* M = jac * (master mass) on straight elements;
* the Schur outputs solve the full element system assembled densely
  (numpy LU), and the schur_uq.c (Cholesky) and schur_primal.m (inv) forms
  agree."""
import numpy as np
import pytest

from hdg_amd.problems import euler2d_square, ns2d_square, euler3d_cube, gather_uh, inituhat, with_bdf1


def _mats(oracle, P, udg=None, uh=None):
    return oracle.elemental_matrices(P.master, P.app, P.mesh.dgnodes, P.mesh.bf,
                                     P.udg if udg is None else udg, P.uh if uh is None else uh, P.sh)


def _fd_cols(oracle, P, field, key, h=1e-6):
    """FD of residual `key` w.r.t. every local DOF of `field` ('udg' or 'uh'),
    perturbing the same local DOF in all elements at once."""
    base = getattr(P, field)
    nloc = base.shape[0] * base.shape[1]
    ne = base.shape[2]
    cols = []
    for k in range(nloc):
        i, j = k % base.shape[0], k // base.shape[0]
        s = h * max(1.0, np.abs(base[i, j]).max())
        xp, xm = base.copy(order="F"), base.copy(order="F")
        xp[i, j, :] += s
        xm[i, j, :] -= s
        rp = _mats(oracle, P, **{field: xp})[key]
        rm = _mats(oracle, P, **{field: xm})[key]
        cols.append(((rp - rm) / (2 * s)).reshape(-1, ne, order="F"))
    return np.stack(cols, axis=1)          # (nres, nloc, ne)


CASES = {
    # flow at an angle so that no boundary face sits on the |u.n| = 0 kink of
    # the far-field upwinding, where central differences are meaningless
    "euler2d_p2": lambda: euler2d_square(2, 2, seed=3, alpha=0.4),
    "euler2d_p3": lambda: euler2d_square(2, 3, seed=4, alpha=0.4),
    "ns2d_p2": lambda: ns2d_square(2, 2, seed=5, av_frac=0.5),
    "euler3d_p1": lambda: euler3d_cube(1, 1, seed=6, alpha=0.4),
    "euler3d_p2": lambda: euler3d_cube(1, 2, seed=7, alpha=0.4),
    # BDF1 time term: s += SH - fc_u u, s_udg -= fc_u (volint.m:180-221)
    "euler2d_p2_bdf1": lambda: with_bdf1(euler2d_square(2, 2, seed=3, alpha=0.4), 0.05, seed=1),
    "ns2d_p2_bdf1": lambda: with_bdf1(ns2d_square(2, 2, seed=5, av_frac=0.5), 0.05, seed=2),
}


def _tilt3d(P, w=0.3):
    """Give a 3D free stream a z velocity so no face has u.n = 0."""
    if P.master.nd == 3:
        r = P.udg[:, 0, :]
        P.udg[:, 4, :] += 0.5 * r * w * w
        P.udg[:, 3, :] += r * w
        uhg = inituhat(P.master, P.mesh.elcon, P.udg, P.app.ncu, P.mesh.nsiz)
        P.uh = gather_uh(uhg, P.mesh.elcon)
    return P


def _perturb(P, rng, amp=0.02):
    """Move off the smooth initial state so every flux term is exercised."""
    _tilt3d(P)
    P.udg = np.asfortranarray(P.udg * (1 + amp * rng.standard_normal(P.udg.shape)))
    P.uh = np.asfortranarray(P.uh * (1 + amp * rng.standard_normal(P.uh.shape)))
    return P


@pytest.mark.parametrize("case", list(CASES))
def test_matrices_are_residual_derivatives(oracle, case):
    P = _perturb(CASES[case](), np.random.default_rng(11))
    m = _mats(oracle, P)
    ne = P.ne
    npv, ncu, nc = P.master.npv, P.app.ncu, P.app.nc
    ndf = P.master.npf * P.master.nfe
    nu, nh = npv * ncu, ncu * ndf

    def rel(a, b):
        return np.abs(a - b).max() / max(np.abs(b).max(), 1e-300)

    dRu_dU = _fd_cols(oracle, P, "udg", "Ru")                    # (nu, npv*nc, ne)
    assert rel(-dRu_dU, m["BD"].reshape(nu, npv * nc, ne, order="F")) < 2e-7
    dRu_dH = _fd_cols(oracle, P, "uh", "Ru")                     # (nu, nh, ne)
    assert rel(-dRu_dH, m["F"].reshape(nu, nh, ne, order="F")) < 2e-7
    dRh_dU = _fd_cols(oracle, P, "udg", "Rh")
    assert rel(-dRh_dU, m["GK"].reshape(nh, npv * nc, ne, order="F")) < 2e-7
    dRh_dH = _fd_cols(oracle, P, "uh", "Rh")
    assert rel(-dRh_dH, m["H"].reshape(nh, nh, ne, order="F")) < 2e-7
    if P.app.localsolve:
        nd = P.master.nd
        dRq_dU = _fd_cols(oracle, P, "udg", "Rq")                # (npv*ncu*nd, npv*nc, ne)
        Mk = np.kron(np.eye(ncu * nd), np.ones((1, 1)))
        # dRq/dq(b, comp) = -M for matching q component, dRq/du = C
        for e in range(ne):
            Mfull = np.kron(np.eye(ncu * nd), m["M"][:, :, e])
            assert rel(-dRq_dU[:, nu:, e], Mfull) < 2e-7
            Cfull = np.vstack([np.kron(np.eye(ncu), m["C"][:, :, d, e]) for d in range(nd)])
            assert rel(dRq_dU[:, :nu, e], Cfull) < 2e-7
        dRq_dH = _fd_cols(oracle, P, "uh", "Rq")                 # (npv*ncu*nd, nh, ne)
        for e in range(ne):
            # uh index (c, b): column c + ncu*b ; E(a, b, d)
            Efull = np.zeros((npv * ncu * nd, nh))
            for d in range(nd):
                for c in range(ncu):
                    Efull[d * nu + c * npv:d * nu + (c + 1) * npv, c::ncu] = m["E"][:, :, d, e]
            assert rel(-dRq_dH[:, :, e], Efull) < 2e-7


@pytest.mark.parametrize("case", list(CASES))
def test_uniform_state_zero_residual(oracle, case):
    P = CASES[case]()
    ncu = P.app.ncu
    uinf = P.app.bcs[0][:ncu]
    udg = np.zeros_like(P.udg)
    udg[:, :ncu, :] = uinf[None, :, None]
    uh = np.zeros_like(P.uh)
    uh[:, :, :] = uinf[:, None, None]
    if P.app.tdep:
        # steady in time as well: u_n = u, so SH - fc_u u = 0
        sh = np.zeros_like(P.udg)
        sh[:, :ncu, :] = P.app.fc_u * uinf[None, :, None]
        P.sh = np.asfortranarray(sh)
    m = _mats(oracle, P, udg=udg, uh=uh)
    scale = np.abs(m["BD"]).max()
    assert np.abs(m["Ru"]).max() < 1e-12 * scale
    # Rh is the element's share of the trace conservation law: it vanishes on
    # boundary faces (fbou = 0 at u = uh = uinf) and after assembly over the
    # two elements of every interior face (opposite normals)
    npf, nfe = P.master.perm.shape
    Rh = m["Rh"]                                   # (ncu, npf*nfe, ne)
    bnd = (P.mesh.bf < 0)                          # (nfe, ne)
    mask = np.repeat(bnd, npf, axis=0)             # (npf*nfe, ne)
    assert np.abs(Rh[:, mask]).max() < 1e-12 * scale
    tot = np.zeros((ncu, P.mesh.nsiz))
    for c in range(ncu):
        np.add.at(tot[c], P.mesh.elcon.ravel(order="F"), Rh[c].ravel(order="F"))
    assert np.abs(tot).max() < 1e-12 * scale
    assert np.abs(Rh).max() > 1e-3 * scale         # and it is not trivially zero per element
    if P.app.localsolve:
        assert np.abs(m["Rq"]).max() < 1e-12 * scale


@pytest.mark.parametrize("case", ["ns2d_p2"])
def test_mass_matrix_closed_form(oracle, case):
    P = CASES[case]()
    m = _mats(oracle, P)
    mst = P.master
    Mhat = mst.shapvl[:, :, 0] @ np.diag(mst.gwvl) @ mst.shapvl[:, :, 0].T
    X = P.mesh.dgnodes
    for e in range(P.ne):
        v = P.mesh.p[P.mesh.t[e]]
        jac = np.linalg.det(np.stack([v[1] - v[0], v[2] - v[0]], 1))
        assert np.abs(m["M"][:, :, e] - jac * Mhat).max() < 1e-14


def _full_system_check(P, m, outs):
    """Assemble [[M,-C,E],[B,D,F],[G,K,H]] (or [[D,F],[G,H]]) and check that the
    local maps solve its element rows and AE/FE are its Schur complement."""
    dudg, dudg_duh, ae, fe = outs[:4]
    npv, ncu, nc = P.master.npv, P.app.ncu, P.app.nc
    ndf = P.master.npf * P.master.nfe
    nu, nh = npv * ncu, ncu * ndf
    rng = np.random.default_rng(0)
    errs = []
    for e in range(P.ne):
        BD = m["BD"][..., e].reshape(nu, npv * nc, order="F")
        F = m["F"][..., e].reshape(nu, nh, order="F")
        GK = m["GK"][..., e].reshape(nh, npv * nc, order="F")
        H = m["H"][..., e].reshape(nh, nh, order="F")
        Ru = m["Ru"][..., e].reshape(nu, order="F")
        Rh = m["Rh"][..., e].reshape(nh, order="F")
        if P.app.localsolve:
            nd = P.master.nd
            M = np.kron(np.eye(ncu * nd), m["M"][:, :, e])
            C = np.vstack([np.kron(np.eye(ncu), m["C"][:, :, d, e]) for d in range(nd)])
            E = np.zeros((nu * nd, nh))
            for d in range(nd):
                for c in range(ncu):
                    E[d * nu + c * npv:d * nu + (c + 1) * npv, c::ncu] = m["E"][:, :, d, e]
            Rq = m["Rq"][..., e].reshape(nu * nd, order="F")
            B = BD[:, nu:]
            D = BD[:, :nu]
            # unknown order: [u, q]; rows: q-eq then u-eq
            A = np.block([[-C, M], [D, B]])
            Fl = np.vstack([E, F])
            R = np.concatenate([Rq, Ru])
            G = np.hstack([GK[:, :nu], GK[:, nu:]])
        else:
            A = BD
            Fl = F
            R = Ru
            G = GK
        X = np.linalg.solve(A, np.hstack([R[:, None], -Fl]))
        du_ref = X[:, 0]
        dlu_ref = X[:, 1:]
        du = dudg[:, :, e].reshape(-1, order="F")
        dlu = dudg_duh[..., e].reshape(npv * nc, nh, order="F")
        AE = ae[..., e].reshape(nh, nh, order="F")
        FE = fe[..., e].reshape(nh, order="F")
        AE_ref = H + G @ dlu_ref
        FE_ref = Rh - G @ du_ref
        for a, b in ((du, du_ref), (dlu, dlu_ref), (AE, AE_ref), (FE, FE_ref)):
            errs.append(np.linalg.norm(a - b) / max(np.linalg.norm(b), 1e-300))
    return max(errs)


@pytest.mark.parametrize("case", list(CASES))
def test_schur_solves_full_element_system(oracle, case):
    P = _perturb(CASES[case](), np.random.default_rng(12))
    m = _mats(oracle, P)
    if P.app.localsolve:
        outs = oracle.schur_uq(m["M"], m["C"], m["E"], m["BD"], m["F"], m["GK"], m["H"], m["Rq"], m["Ru"], m["Rh"])
    else:
        outs = oracle.schur_u(m["BD"], m["F"], m["GK"], m["H"], m["Ru"], m["Rh"])
    assert np.all(outs[4] == 0)
    assert _full_system_check(P, m, outs) < 1e-11


def test_schur_uq_chol_equals_inv_form(oracle):
    P = _perturb(ns2d_square(3, 3, seed=9, av_frac=0.3), np.random.default_rng(13))
    m = _mats(oracle, P)
    a = oracle.schur_uq(*(m[k] for k in ("M", "C", "E", "BD", "F", "GK", "H", "Rq", "Ru", "Rh")))
    b = oracle.schur_uq(*(m[k] for k in ("M", "C", "E", "BD", "F", "GK", "H", "Rq", "Ru", "Rh")), variant="inv")
    for x, y in zip(a[:4], b[:4]):
        assert np.linalg.norm(x - y) / np.linalg.norm(y) < 1e-12


@pytest.mark.parametrize("case", ["euler2d_p2", "ns2d_p2", "euler3d_p2"])
def test_elemental_equals_matrices_then_schur(oracle, case):
    P = _perturb(CASES[case](), np.random.default_rng(14))
    m = _mats(oracle, P)
    if P.app.localsolve:
        ref = oracle.schur_uq(*(m[k] for k in ("M", "C", "E", "BD", "F", "GK", "H", "Rq", "Ru", "Rh")))
    else:
        ref = oracle.schur_u(m["BD"], m["F"], m["GK"], m["H"], m["Ru"], m["Rh"])
    ae, fe, dudg, dudg_duh, st = oracle.elemental(P.master, P.app, P.mesh.dgnodes, P.mesh.bf, P.udg, P.uh, P.sh)
    assert np.array_equal(ae, ref[2]) and np.array_equal(fe, ref[3])
    assert np.array_equal(dudg, ref[0]) and np.array_equal(dudg_duh, ref[1])


def test_singular_element_reports_status(oracle):
    """A collapsed element (all nodes equal) gives a zero-area mass and a
    singular D; the status word says so instead of silently returning garbage
    (the reference ignores LAPACK's info, kernel/schur_uq.c:69,136)."""
    P = ns2d_square(2, 2, seed=5)
    P.mesh.dgnodes[:, :2, 0] = P.mesh.dgnodes[0, :2, 0][None, :]
    ae, fe, dudg, dudg_duh, st = oracle.elemental(P.master, P.app, P.mesh.dgnodes, P.mesh.bf, P.udg, P.uh, P.sh)
    assert st[0] != 0 and np.all(st[1:] == 0)
