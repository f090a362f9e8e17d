"""CPU restatement of kernel/getq.m (TEST INFRASTRUCTURE ONLY).

Only tests/, __graft_entry__.smoke() and bench.py's cpu_baseline leg may
import this module.  It follows getq.m step by step, vectorised over the
elements: geometry with getq.m's own sign conventions (its volgeom and
facegeom, getq.m:131-173), the mass, convection and face matrices
(getq.m:78-101), q = Mi (E uh - C u) with Mi = inv(M)/fc_q (getq.m:112-119),
the SH shift (getq.m:125-127) and the nodal gradient qq (getq.m:55-76).

Pinned by tests/test_getq_oracle.py: on straight-sided elements with uh the
trace of u and u a polynomial of degree <= p, q and qq equal -grad u exactly;
and q zeroes the Rq of the C++ restatement of volint/faceint.
"""
import numpy as np


def _volgeom(J):
    """getq.m:131-160: jac and Xx = -adj(J), J[..., a, b] = dx_b/dxi_a."""
    nd = J.shape[-1]
    Xx = np.empty_like(J)
    if nd == 2:
        jac = J[..., 0, 0] * J[..., 1, 1] - J[..., 0, 1] * J[..., 1, 0]
        Xx[..., 0, 0] = -J[..., 1, 1]
        Xx[..., 1, 0] = J[..., 1, 0]
        Xx[..., 0, 1] = J[..., 0, 1]
        Xx[..., 1, 1] = -J[..., 0, 0]
    elif nd == 3:
        g = lambda a, b: J[..., a - 1, b - 1]
        jac = (g(1, 1) * g(2, 2) * g(3, 3) - g(1, 1) * g(3, 2) * g(2, 3) + g(2, 1) * g(3, 2) * g(1, 3)
               - g(2, 1) * g(1, 2) * g(3, 3) + g(3, 1) * g(1, 2) * g(2, 3) - g(3, 1) * g(2, 2) * g(1, 3))
        Xx[..., 0, 0] = g(2, 3) * g(3, 2) - g(2, 2) * g(3, 3)
        Xx[..., 1, 0] = g(2, 1) * g(3, 3) - g(2, 3) * g(3, 1)
        Xx[..., 2, 0] = g(2, 2) * g(3, 1) - g(2, 1) * g(3, 2)
        Xx[..., 0, 1] = g(1, 2) * g(3, 3) - g(1, 3) * g(3, 2)
        Xx[..., 1, 1] = g(1, 3) * g(3, 1) - g(1, 1) * g(3, 3)
        Xx[..., 2, 1] = g(1, 1) * g(3, 2) - g(1, 2) * g(3, 1)
        Xx[..., 0, 2] = g(1, 3) * g(2, 2) - g(1, 2) * g(2, 3)
        Xx[..., 1, 2] = g(1, 1) * g(2, 3) - g(1, 3) * g(2, 1)
        Xx[..., 2, 2] = g(1, 2) * g(2, 1) - g(1, 1) * g(2, 2)
    else:
        raise ValueError("dimension not implemented")
    return jac, Xx


def _facegeom(dpg):
    """getq.m:162-173: unit normal and surface Jacobian from dpg[..., d, k] =
    dx_d/ds_k."""
    nd = dpg.shape[-2]
    if nd == 2:
        jac = np.sqrt(dpg[..., 0, 0] ** 2 + dpg[..., 1, 0] ** 2)
        nlg = np.stack([dpg[..., 1, 0], -dpg[..., 0, 0]], axis=-1) / jac[..., None]
    else:
        d = lambda a, k: dpg[..., a - 1, k - 1]
        nlg = np.stack([d(2, 1) * d(3, 2) - d(3, 1) * d(2, 2),
                        d(3, 1) * d(1, 2) - d(1, 1) * d(3, 2),
                        d(1, 1) * d(2, 2) - d(2, 1) * d(1, 2)], axis=-1)
        jac = np.sqrt((nlg ** 2).sum(-1))
        nlg = nlg / jac[..., None]
    return nlg, jac


def getq(master, dgnodes, UDG, UHE, SH=None, fc_q=1.0, shapnv=None):
    """QDG (npv, ncu*nd, ne) [and qq when shapnv = mkshape(p, plocvl, plocvl)
    is given] of getq.m for element-gathered traces UHE (ncu, npf*nfe, ne)
    (getq.m:21-22 gathers UH(:, elcon) the same way)."""
    nd, npv, npf = master.nd, master.npv, master.npf
    ngv, ngf = master.ngv, master.ngf
    perm = np.asarray(master.perm)                      # (npf, nfe), 0-based
    nfe = perm.shape[1]
    ncu = UHE.shape[0]
    ne = UDG.shape[2]
    dg = np.asarray(dgnodes)[:, :nd, :]                 # getq.m:47
    svv = np.asarray(master.shapvgdotshapvl).reshape(npv, npv, ngv, nd + 1, order="F")   # (jj, ii, g, d)
    sff = np.asarray(master.shapfgdotshapfc).reshape(npf, npf, ngf, nd, order="F")[..., 0]   # (a, b, g)

    # volume geometry (getq.m:29, 52-60): Jg(g, a, b) = dx_b/dxi_a
    dshapvt = np.asarray(master.shapvt)[:, :, 1:]       # (g, p, a)
    Jg = np.einsum("gpa,pbe->egab", dshapvt, dg)
    jac, Xx = _volgeom(Jg)                              # (ne, g), (ne, g, nd, nd)

    # face geometry (getq.m:50, 77-81): dpg(g, j, d, k) = dx_d/ds_k on face j
    pn = dg[perm, :, :]                                 # (npf, nfe, nd, ne)
    dshapft = np.asarray(master.shapft)[:, :, 1:]       # (g, a, k)
    dpg = np.einsum("gak,ajde->ejgdk", dshapft, pn)
    nlg, jacf = _facegeom(dpg)                          # (ne, nfe, ngf, nd), (ne, nfe, ngf)

    # M, Mi (getq.m:83-85)
    M = np.einsum("jig,eg->eji", svv[..., 0], jac)
    Mi = np.linalg.inv(M) / fc_q

    # C(:, :, i) = sum_j shapvgdotshapvl(:, :, j+1) * Xx(:, i, j) (getq.m:88-94)
    C = np.einsum("abgd,egid->eabi", svv[..., 1:], Xx)

    # E(perm(:, j), (j-1)*npf+1:j*npf, i) = shapfgdotshapfc * (nlg_i .* -jacf) (getq.m:97-104)
    njc = nlg * (-jacf)[..., None]                      # (ne, nfe, ngf, nd)
    wrk = np.einsum("abg,ejgi->eabji", sff, njc)        # (ne, a, b, j, i)
    E = np.zeros((ne, npv, npf * nfe, nd))
    for j in range(nfe):
        E[:, perm[:, j], j * npf:(j + 1) * npf, :] = wrk[:, :, :, j, :]

    # q = Mi E uh - Mi C u (getq.m:112-119), columns c + ncu*i
    u = np.asarray(UDG)[:, :ncu, :]
    uh = np.asarray(UHE)                                # (ncu, ndf, ne)
    Cu = np.einsum("eabi,bce->eaci", C, u)
    Euh = np.einsum("eabi,cbe->eaci", E, uh)
    q = np.einsum("eab,ebci->eaci", Mi, Euh - Cu)       # (ne, npv, ncu, nd)
    QDG = np.asfortranarray(np.transpose(q, (1, 2, 3, 0)).reshape(npv, ncu * nd, ne, order="F"))
    if SH is not None and fc_q != 0:
        QDG = QDG + np.asarray(SH)[:, -ncu * nd:, :] / fc_q    # getq.m:125-127
    if shapnv is None:
        return QDG

    # qq: -J^-1 grad_xi u at the nodes (getq.m:55-76)
    dshapnv = np.asarray(shapnv)[:, :, 1:]              # (phi, node, a)
    Jn = np.einsum("pka,pbe->ekab", dshapnv, dg)
    jacn, Xxn = _volgeom(Jn)
    JnInv = Xxn / jacn[..., None, None]
    du = np.einsum("pka,pce->ekac", dshapnv, u)         # (ne, node, a, c)
    qq = np.einsum("ekra,ekac->ekcr", JnInv, du)        # (ne, node, c, r)
    qq = np.asfortranarray(np.transpose(qq, (1, 2, 3, 0)).reshape(npv, ncu * nd, ne, order="F"))
    return QDG, qq
