// Mixed-order persistent launches: one launch of the quadrature kernel and
// one of the elimination kernel cover the elements of every order p of one
// physics and dimension.  A fixed grid of workgroups strides over the batch
// (ordered by decreasing p, so every workgroup meets every order); each
// element dispatches to the body of its order, with the LDS of all orders
// overlaid in one buffer.
#include <hip/hip_runtime.h>

#include "hdgx_device.h"
#include "hdgx_kernels.h"

namespace hdgx {

template <int ND>
constexpr int family_maxp() { return ND == 2 ? 4 : 2; }

// LDS of all orders of a family overlaid (the members keep the LDS address
// space, so the bodies address them as in the single-order kernels)
template <int PH, int ND> union QuadLdsU;
template <int PH> union QuadLdsU<PH, 2> {
  QuadLds<Cfg<PH, 2, 1>> p1; QuadLds<Cfg<PH, 2, 2>> p2; QuadLds<Cfg<PH, 2, 3>> p3; QuadLds<Cfg<PH, 2, 4>> p4;
};
template <int PH> union QuadLdsU<PH, 3> { QuadLds<Cfg<PH, 3, 1>> p1; QuadLds<Cfg<PH, 3, 2>> p2; };
template <int PH, int ND> union SchurLdsU;
template <int PH> union SchurLdsU<PH, 2> {
  SchurLds<Cfg<PH, 2, 1>> p1; SchurLds<Cfg<PH, 2, 2>> p2; SchurLds<Cfg<PH, 2, 3>> p3; SchurLds<Cfg<PH, 2, 4>> p4;
};
template <int PH> union SchurLdsU<PH, 3> { SchurLds<Cfg<PH, 3, 1>> p1; SchurLds<Cfg<PH, 3, 2>> p2; };

template <class C>
HDGX_DEV void quad_mixed_elem(const MixedTable& T, const ElemSeg& sg, const DevApp& ap, int64_t e, int64_t wp,
                              QuadLds<C>& lds) {
  constexpr int NPV = C::NPV, NFE = C::NFE, NC = C::NC, NU = C::NU, NH = C::NH, NCOL = C::NCOL;
  constexpr int ND = C::ND, NDF = C::NDF, NCU = C::NCU, P = C::P;
  QuadArgs A;
  A.ne = 1;
  A.dgnodes = sg.dgnodes + e * NPV * ap.ncd;
  A.bf = sg.bf + e * NFE;
  A.udg = sg.udg + e * NPV * NC;
  A.uh = sg.uh + e * NH;
  A.sh = sg.sh ? sg.sh + e * NPV * NC : nullptr;
  A.BD = T.BD[P] + wp * NU * NCOL;
  A.F = T.F[P] + wp * NU * NH;
  A.GK = T.GK[P] + wp * NH * NCOL;
  A.H = T.H[P] + wp * NH * NH;
  A.Ru = T.Ru[P] + wp * NU;
  A.Rh = T.Rh[P] + wp * NH;
  if constexpr (C::UQ) {
    A.M = T.M[P] + wp * NPV * NPV;
    A.C = T.C[P] + wp * NPV * NPV * ND;
    A.E = T.E[P] + wp * NPV * NDF * ND;
    A.Rq = T.Rq[P] + wp * NPV * NCU * ND;
  } else {
    A.M = A.C = A.E = A.Rq = nullptr;
  }
  A.status = nullptr;
  quad_body<C>(A, T.masters[P], ap, 0, lds);
}

template <class C>
HDGX_DEV void schur_mixed_elem(const MixedTable& T, const ElemSeg& sg, int64_t e, int64_t wp, SchurLds<C>& lds) {
  constexpr int NPV = C::NPV, NU = C::NU, NH = C::NH, NCOL = C::NCOL, ND = C::ND, NDF = C::NDF, NCU = C::NCU;
  constexpr int P = C::P;
  SchurArgs A;
  A.ne = 1;
  A.BD = T.BD[P] + wp * NU * NCOL;
  A.F = T.F[P] + wp * NU * NH;
  A.GK = T.GK[P] + wp * NH * NCOL;
  A.H = T.H[P] + wp * NH * NH;
  A.Ru = T.Ru[P] + wp * NU;
  A.Rh = T.Rh[P] + wp * NH;
  if constexpr (C::UQ) {
    A.M = T.M[P] + wp * NPV * NPV;
    A.C = T.C[P] + wp * NPV * NPV * ND;
    A.E = T.E[P] + wp * NPV * NDF * ND;
    A.Rq = T.Rq[P] + wp * NPV * NCU * ND;
  } else {
    A.M = A.C = A.E = A.Rq = nullptr;
  }
  A.dudg = sg.dudg + e * NCOL;
  A.dudg_duh = sg.dudg_duh + e * NCOL * NH;
  A.ae = sg.ae + e * NH * NH;
  A.fe = sg.fe + e * NH;
  A.status = sg.status ? sg.status + e : nullptr;
  A.prec = 0;
  A.nco = C::NC;
  A.aek = nullptr;
  A.bd_scratch = 0;
  schur_mfma_body<C>(A, T.masters[P].perm, 0, lds);
}

// the elements of order P, strided over the persistent grid
template <int PH, int ND, int P>
HDGX_DEV void quad_mixed_order(const MixedTable* Tp, QuadLds<Cfg<PH, ND, P>>& lds) {
  const int64_t w0 = Tp->first[P], w1 = Tp->end[P];
  for (int64_t w = w0 + blockIdx.x; w < w1; w += gridDim.x) {
    const MixedTable& T = *Tp;
    const ElemSeg sg = uload_obj(T.segs + find_seg(T.segs, T.nseg, w));
    quad_mixed_elem<Cfg<PH, ND, P>>(T, sg, T.apps[sg.app], w - sg.e0, w - w0, lds);
    __syncthreads();
  }
}

template <int PH, int ND, int P>
HDGX_DEV void schur_mixed_order(const MixedTable* Tp, SchurLds<Cfg<PH, ND, P>>& lds) {
  const int64_t w0 = Tp->first[P], w1 = Tp->end[P];
  for (int64_t w = w0 + blockIdx.x; w < w1; w += gridDim.x) {
    const MixedTable& T = *Tp;
    const ElemSeg sg = uload_obj(T.segs + find_seg(T.segs, T.nseg, w));
    schur_mixed_elem<Cfg<PH, ND, P>>(T, sg, w - sg.e0, w - w0, lds);
    __syncthreads();
  }
}

// orders in decreasing p; batch elements of order p are first[p] .. end[p]-1
template <int PH, int ND>
__global__ __launch_bounds__(QUAD_THREADS) void quad_mixed_kernel(const MixedTable* Tp) {
  __shared__ QuadLdsU<PH, ND> lds;
  if constexpr (ND == 2) {
    quad_mixed_order<PH, ND, 4>(Tp, lds.p4);
    quad_mixed_order<PH, ND, 3>(Tp, lds.p3);
  }
  quad_mixed_order<PH, ND, 2>(Tp, lds.p2);
  quad_mixed_order<PH, ND, 1>(Tp, lds.p1);
}

template <int PH, int ND>
__global__ __launch_bounds__(64) __attribute__((amdgpu_waves_per_eu(1, 1))) void schur_mixed_kernel(const MixedTable* Tp) {
  __shared__ SchurLdsU<PH, ND> lds;
  if constexpr (ND == 2) {
    schur_mixed_order<PH, ND, 4>(Tp, lds.p4);
    schur_mixed_order<PH, ND, 3>(Tp, lds.p3);
  }
  schur_mixed_order<PH, ND, 2>(Tp, lds.p2);
  schur_mixed_order<PH, ND, 1>(Tp, lds.p1);
}

}  // namespace hdgx
