"""Algorithmic work of one element Schur solve (dgnodes, UDG, UH, SH, bf) ->
(AE, FE, DUDG, DUDG_DUH): the dense operation count of the reference's own
contractions (kernel/volint.m:258-263, kernel/faceint.m:306-350,
kernel/schur_u.c, kernel/schur_uq.c) and the compulsory FP64 HBM traffic.
Structure-aware savings in the kernels (block-diagonal H, face-sparse GK)
do not change these counts."""
from dataclasses import dataclass


@dataclass
class ElemWork:
    quad_flops: int
    schur_flops: int
    read_bytes: int
    write_bytes: int

    @property
    def flops(self):
        return self.quad_flops + self.schur_flops

    @property
    def bytes(self):
        return self.read_bytes + self.write_bytes


def element_work(nd, npv, npf, ngv, ngf, ncu, nc, uq):
    nfe = nd + 1
    ndf = npf * nfe
    nu, nh, ncol = npv * ncu, ncu * ndf, npv * nc
    quad = (2 * npv * npv * ngv * (nd + 1) * ncu * nc + 2 * npv * ngv * (nd + 1) * ncu
            + 4 * npf * npf * ngf * nfe * (ncu * nc + ncu * ncu) + 2 * npf * ngf * nfe * ncu)
    if uq:
        quad += 2 * npv * npv * ngv * nd * nd + 2 * npv * ngf * nfe * npf * nd
    schur = (2 * nu ** 3) // 3 + 2 * nu * nu * (1 + nh) + 2 * nh * nh * ncol + 2 * nh * ncol
    if uq:
        schur += (npv ** 3 // 3 + 2 * npv * npv * (ncu * nd + npv * nd + ndf * nd) + 2 * nu * npv * npv * ncu * nd
                  + 2 * nu * npv * ndf * ncu * nd + 2 * nu * npv * nd * ncu
                  + 2 * npv * npv * ncu * nd * (1 + ncu * ndf))
    # read: dgnodes (nd coordinate columns), UDG, gathered UH, elcon (int32), bf (int32)
    read = 8 * (npv * nd + npv * nc + ncu * ndf) + 4 * (ndf + nfe)
    # write: AE, FE, DUDG, DUDG_DUH
    write = 8 * (nh * nh + nh + ncol + ncol * nh)
    return ElemWork(quad, schur, read, write)


def problem_work(P):
    m = P.master
    return element_work(m.nd, m.npv, m.npf, m.ngv, m.ngf, P.app.ncu, P.app.nc, bool(P.app.localsolve))
