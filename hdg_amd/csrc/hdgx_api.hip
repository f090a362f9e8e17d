// C ABI of the hdgx library (include/hdgx.h): contexts, master-table upload,
// argument validation and dispatch to the per-(physics, nd, p) kernels.
#include <hip/hip_runtime.h>

#include <cstdint>
#include <cstring>
#include <map>
#include <memory>
#include <mutex>
#include <algorithm>
#include <cmath>
#include <functional>
#include <string>
#include <type_traits>
#include <vector>

#include "../../include/hdgx.h"
#include "hdgx_ctx.h"
#include "hdgx_kernels.h"

#include "hdgx_launch.h"
#include "hdgx_shapes.h"

using namespace hdgx;
using hdgx_detail::fail;

namespace {

// Element matrices pass between the quadrature and elimination kernels through
// a context workspace of at most ctx->ws_cap bytes (24 GiB by default, a
// twelfth of MI355X's 288 GB; hdgx_set_workspace_limit).  Larger batches run
// in chunks of nearly equal size: 384 000 p = 2 tetrahedra (231 KB of
// matrices each) in 4 chunks rather than 14 under a 6 GiB cap -- each chunk
// boundary drains the GPU (tetrahedra step 142 -> 132 ms).
int64_t ws_chunk(int64_t n, size_t bytes_per_element, size_t cap) {
  if (n <= 0) return 1;
  const int64_t cmax = std::max<int64_t>(1, std::min<int64_t>(n, (int64_t)(cap / bytes_per_element)));
  const int64_t nch = (n + cmax - 1) / cmax;
  return (n + nch - 1) / nch;
}

// FNV-1a over 8-byte words (then the tail bytes), with a shift mixing each
// step: the master-table cache key, computed on every call
uint64_t fnv(uint64_t h, const void* p, size_t n) {
  const unsigned char* c = static_cast<const unsigned char*>(p);
  size_t i = 0;
  for (; i + 8 <= n; i += 8) {
    uint64_t w;
    std::memcpy(&w, c + i, 8);
    h ^= w;
    h *= 1099511628211ull;
    h ^= h >> 29;
  }
  for (; i < n; ++i) { h ^= c[i]; h *= 1099511628211ull; }
  return h;
}

// The master tables packed for the device: one host image of every table plus
// the node-on-face lookup loc(v, j) derived from perm, the byte offset of each
// table in it, and a content hash (FNV-1a over the sizes and every table).
struct MasterImage {
  std::vector<char> host;
  size_t off[14];
  uint64_t hash;
};

// AE k-step table of the elimination for nc components per node: an AE row
// tile (16 rows of the (ncu, npf, nfe) trace rows) meets the faces of its
// rows, and only the GK columns of those faces' nodes are nonzero
// (kernel/faceint.m:354-358).  Layout: the column count of each tile, then
// each tile's 4*KSMAX columns in k-step order (node fastest, then component),
// 255 past the count.  KSMAX = ceil(max(npv, 2 npf) nc / 4), as in the kernel.
std::vector<unsigned char> ae_table(int npv, int npf, int nfe, int ncu, int nc, const int32_t* perm) {
  const int nh = ncu * npf * nfe, trh = (nh + 15) / 16;
  const int nnmax = std::max(npv, 2 * npf), ksmax = (nnmax * nc + 3) / 4;
  std::vector<unsigned char> t((size_t)trh * (1 + 4 * ksmax), 255);
  // column indices are bytes: orders past the built configurations (npv nc
  // above 254) get empty tables, which no kernel reads
  if (npv * nc > 254) {
    for (int rta = 0; rta < trh; ++rta) t[rta] = 0;
    return t;
  }
  for (int rta = 0; rta < trh; ++rta) {
    const int j0 = (16 * rta) / (ncu * npf);
    const int j1 = std::min(nfe - 1, (16 * rta + 15) / (ncu * npf));
    std::vector<int> nodes;
    for (int j = j0; j <= j1; ++j)
      for (int a = 0; a < npf; ++a) {
        const int v = perm[a + npf * j];
        if (std::find(nodes.begin(), nodes.end(), v) == nodes.end()) nodes.push_back(v);
      }
    const int nn = (int)nodes.size(), klen = nn * nc;
    t[rta] = (unsigned char)std::min(klen, 255);
    for (int idx = 0; idx < 4 * ksmax && idx < klen; ++idx)
      t[trh + (size_t)rta * 4 * ksmax + idx] = (unsigned char)(nodes[idx % nn] + npv * (idx / nn));
  }
  return t;
}

// B fragments of the quadrature's BD product (DevMaster::bdv, BdShape in
// hdgx_shapes.h): SV(r', (jj, ii)) over the volume rows r' < ngv (nd+1)
// (shapvgdotshapvl, util/preprocess.m:23-37, as volint.m:262 contracts it) and
// the face points r' = ngv (nd+1) + j ngf + g (shapfgdotshapfc(la, lb, g) with
// la, lb the positions of jj, ii on face j, faceint.m:316,326, when both lie on
// it), laid out as 64-lane vectors per (row tile R of pairs jj + npv ii, k-step
// s): lane l = SV(4 s + l / 16, 16 R + l % 16), zero past either size.
std::vector<double> bd_table(const hdgx_master* m, const std::vector<int32_t>& loc) {
  const int npv = m->npv, npf = m->npf, nfe = m->nfe, ngv = m->ngv, ngf = m->ngf, nd = m->nd;
  const int nvv = npv * npv, rtb = (nvv + 15) / 16, nr = ngv * (nd + 1), kt = nr + nfe * ngf, ksb = (kt + 3) / 4;
  std::vector<double> t((size_t)rtb * ksb * 64, 0.0);
  for (int R = 0; R < rtb; ++R)
    for (int s = 0; s < ksb; ++s)
      for (int l = 0; l < 64; ++l) {
        const int r = 4 * s + l / 16, col = 16 * R + l % 16;
        if (r >= kt || col >= nvv) continue;
        double v = 0.0;
        if (r < nr) {
          v = m->shapvgdotshapvl[col + (size_t)nvv * r];
        } else {
          const int tt = r - nr, j = tt / ngf, g = tt % ngf, jj = col % npv, ii = col / npv;
          const int la = loc[jj + (size_t)npv * j], lb = loc[ii + (size_t)npv * j];
          if (la >= 0 && lb >= 0) v = m->shapfgdotshapfc[la + npf * lb + (size_t)npf * npf * g];
        }
        t[((size_t)R * ksb + s) * 64 + l] = v;
      }
  return t;
}

// Checks a master's sizes and table pointers and hashes its content (the
// cache key of its device image).
int master_hash(hdgx_ctx* ctx, const hdgx_master* m, uint64_t* out) {
  if (!m || !m->shapvt || !m->shapvg || !m->shapvgdotshapvl || !m->shapft || !m->shapfg || !m->shapfgdotshapfc ||
      !m->shapmv || !m->shapmf || !m->perm || !m->permgeom)
    return fail(ctx, HDGX_ERR_ARG, "master: null table");
  const int nd = m->nd, npv = m->npv, npf = m->npf, nfe = m->nfe, ngv = m->ngv, ngf = m->ngf;
  if (nd < 2 || nd > 3 || npv < 1 || npf < 1 || nfe != nd + 1 || ngv < 1 || ngf < 1)
    return fail(ctx, HDGX_ERR_ARG, "master: bad sizes");
  const size_t n_vt = (size_t)ngv * npv * (nd + 1), n_vg = (size_t)npv * ngv * (nd + 1);
  const size_t n_vv = (size_t)npv * npv * ngv * (nd + 1), n_ft = (size_t)ngf * npf * nd;
  const size_t n_fg = (size_t)npf * ngf * nd, n_ff = (size_t)npf * npf * ngf * nd;
  const size_t n_mv = n_vt, n_mf = n_ft, n_perm = (size_t)npf * nfe;
  uint64_t h = 1469598103934665603ull;
  int dims[7] = {nd, m->porder, npv, npf, nfe, ngv, ngf};
  h = fnv(h, dims, sizeof(dims));
  h = fnv(h, m->shapvt, n_vt * 8); h = fnv(h, m->shapvg, n_vg * 8); h = fnv(h, m->shapvgdotshapvl, n_vv * 8);
  h = fnv(h, m->shapft, n_ft * 8); h = fnv(h, m->shapfg, n_fg * 8); h = fnv(h, m->shapfgdotshapfc, n_ff * 8);
  h = fnv(h, m->shapmv, n_mv * 8); h = fnv(h, m->shapmf, n_mf * 8);
  h = fnv(h, m->perm, n_perm * 4); h = fnv(h, m->permgeom, n_perm * 4);
  *out = h;
  return HDGX_OK;
}

int pack_master(hdgx_ctx* ctx, const hdgx_master* m, MasterImage* img) {
  int rc = master_hash(ctx, m, &img->hash);
  if (rc) return rc;
  const int nd = m->nd, npv = m->npv, npf = m->npf, nfe = m->nfe, ngv = m->ngv, ngf = m->ngf;
  const size_t n_vt = (size_t)ngv * npv * (nd + 1), n_vg = (size_t)npv * ngv * (nd + 1);
  const size_t n_vv = (size_t)npv * npv * ngv * (nd + 1), n_ft = (size_t)ngf * npf * nd;
  const size_t n_fg = (size_t)npf * ngf * nd, n_ff = (size_t)npf * npf * ngf * nd;
  const size_t n_mv = n_vt, n_mf = n_ft, n_perm = (size_t)npf * nfe, n_loc = (size_t)npv * nfe;
  std::vector<int32_t> loc(n_loc, -1);
  for (int j = 0; j < nfe; ++j)
    for (int a = 0; a < npf; ++a) {
      const int v = m->perm[a + npf * j];
      if (v < 0 || v >= npv) return fail(ctx, HDGX_ERR_ARG, "master: perm out of range");
      loc[v + (size_t)npv * j] = a;
    }
  // a face lists each of its nodes once: loc(perm(a, j), j) = a (the
  // quadrature kernels' face-block keys rely on it)
  for (int j = 0; j < nfe; ++j)
    for (int a = 0; a < npf; ++a)
      if (loc[m->perm[a + npf * j] + (size_t)npv * j] != a) return fail(ctx, HDGX_ERR_ARG, "master: perm repeats a node on a face");
  for (size_t i = 0; i < n_perm; ++i)
    if (m->permgeom[i] < 0 || m->permgeom[i] >= npv) return fail(ctx, HDGX_ERR_ARG, "master: permgeom out of range");
  const int ncu = nd + 2;
  std::vector<unsigned char> aek_u = ae_table(npv, npf, nfe, ncu, ncu, m->perm);
  std::vector<unsigned char> aek_uq = ae_table(npv, npf, nfe, ncu, ncu * (nd + 1), m->perm);
  // after each AE table the node-on-face table loc(v, j) as bytes (255: v is
  // not on face j), which the elimination reads with it to gather F's face
  // blocks
  for (size_t i = 0; i < n_loc; ++i) {
    const unsigned char b = loc[i] < 0 ? (unsigned char)255 : (unsigned char)loc[i];
    aek_u.push_back(b);
    aek_uq.push_back(b);
  }
  const std::vector<double> bdv = bd_table(m, loc);
  // every table starts on a 16-byte boundary
  img->host.clear();
  int k = 0;
  auto put = [&](const void* src, size_t n) {
    const size_t at = (img->host.size() + 15) / 16 * 16;
    img->host.resize(at + n, 0);
    std::memcpy(img->host.data() + at, src, n);
    img->off[k++] = at;
  };
  put(m->shapvt, n_vt * 8); put(m->shapvg, n_vg * 8); put(m->shapvgdotshapvl, n_vv * 8); put(m->shapft, n_ft * 8);
  put(m->shapfg, n_fg * 8); put(m->shapfgdotshapfc, n_ff * 8); put(m->shapmv, n_mv * 8); put(m->shapmf, n_mf * 8);
  put(m->perm, n_perm * 4); put(m->permgeom, n_perm * 4); put(loc.data(), n_loc * 4);
  put(aek_u.data(), aek_u.size()); put(aek_uq.data(), aek_uq.size());
  put(bdv.data(), bdv.size() * 8);
  return HDGX_OK;
}

// the device view of a master image copied to `base`
void bind_master(const hdgx_master* m, const MasterImage& img, char* base, MasterBuf* mb) {
  auto dp = [&](int i) { return base + img.off[i]; };
  mb->tab.shapvt = (const double*)dp(0); mb->tab.shapvg = (const double*)dp(1); mb->tab.svv = (const double*)dp(2);
  mb->tab.shapft = (const double*)dp(3); mb->tab.shapfg = (const double*)dp(4); mb->tab.sff = (const double*)dp(5);
  mb->tab.shapmv = (const double*)dp(6); mb->tab.shapmf = (const double*)dp(7);
  mb->tab.perm = (const int32_t*)dp(8); mb->tab.permgeom = (const int32_t*)dp(9); mb->tab.loc = (const int32_t*)dp(10);
  mb->tab.aek[0] = (const unsigned char*)dp(11); mb->tab.aek[1] = (const unsigned char*)dp(12);
  mb->tab.bdv = (const double*)dp(13);
  mb->nd = m->nd; mb->porder = m->porder; mb->npv = m->npv; mb->npf = m->npf; mb->nfe = m->nfe;
  mb->ngv = m->ngv; mb->ngf = m->ngf;
}

// Upload (once per distinct table content and context) the master tables.
int get_master(hdgx_ctx* ctx, const hdgx_master* m, MasterBuf** out) {
  uint64_t hash = 0;
  int rc = master_hash(ctx, m, &hash);
  if (rc) return rc;
  auto it = ctx->masters.find(hash);
  if (it != ctx->masters.end()) { *out = &it->second; return HDGX_OK; }
  MasterImage img;
  rc = pack_master(ctx, m, &img);
  if (rc) return rc;
  MasterBuf mb;
  HIP_TRY(ctx, hipSetDevice(ctx->device));
  HIP_TRY(ctx, hipMalloc(&mb.dev, img.host.size()));
  HIP_TRY(ctx, hipMemcpy(mb.dev, img.host.data(), img.host.size(), hipMemcpyHostToDevice));
  bind_master(m, img, static_cast<char*>(mb.dev), &mb);
  auto res = ctx->masters.emplace(img.hash, mb);
  *out = &res.first->second;
  return HDGX_OK;
}

// Physics plugins and boundary models the kernels implement.
int make_app(hdgx_ctx* ctx, const hdgx_master* m, const hdgx_app* a, DevApp* d) {
  if (!a) return fail(ctx, HDGX_ERR_ARG, "app: null");
  const int nd = m->nd;
  if (a->nd != nd) return fail(ctx, HDGX_ERR_ARG, "app.nd != master.nd");
  if (a->ncu != nd + 2 || a->nch != a->ncu) return fail(ctx, HDGX_ERR_ARG, "app: ncu must be nd+2");
  if (a->physics == HDGX_PHYS_EULER) {
    if (a->nc != a->ncu || a->localsolve != 0) return fail(ctx, HDGX_ERR_UNSUPPORTED, "Euler: u-only (localsolve 0, nc = ncu)");
    if (a->nparam < 3) return fail(ctx, HDGX_ERR_ARG, "Euler: param needs {gam, eps, ..., tau}");
  } else if (a->physics == HDGX_PHYS_NS) {
    if (nd != 2 || a->nc != 12 || a->localsolve != 1) return fail(ctx, HDGX_ERR_UNSUPPORTED, "NS: 2D with u,q (nc 12, localsolve 1)");
    if (a->nparam < 6) return fail(ctx, HDGX_ERR_ARG, "NS: param needs {gam, eps, Re, Pr, Minf, tau}");
  } else {
    return fail(ctx, HDGX_ERR_UNSUPPORTED, "unknown physics");
  }
  if (a->ncd < nd || a->ncd > nd + 1) return fail(ctx, HDGX_ERR_ARG, "app: ncd must be nd or nd+1");
  if (a->av_col >= a->ncd || (a->av_col >= 0 && a->av_col < nd)) return fail(ctx, HDGX_ERR_ARG, "app: av_col");
  if (a->nbc < 0 || a->nbc > 8) return fail(ctx, HDGX_ERR_ARG, "app: nbc");
  std::memset(d, 0, sizeof(*d));
  d->gam = a->param[0];
  d->eps = a->param[1];
  if (a->physics == HDGX_PHYS_NS) { d->Re = a->param[2]; d->Pr = a->param[3]; d->Minf = a->param[4]; }
  d->tau = a->param[a->nparam - 1];
  d->fc_u = a->fc_u;
  d->fc_q = a->fc_q;
  d->tdep = a->tdep;
  d->av_col = a->av_col;
  d->ncd = a->ncd;
  d->nbc = a->nbc;
  for (int i = 0; i < a->nbc; ++i) {
    const int ib = a->bcm[i];
    // far field everywhere; 2D Euler: wall 2, prescribed pressure 3 and 4,
    // inflow 5; NS: adiabatic wall 2, isothermal wall 3, prescribed pressure 4,
    // Dirichlet 6
    const bool euler2 = a->physics == HDGX_PHYS_EULER && nd == 2, ns = a->physics == HDGX_PHYS_NS;
    const bool ok = ib == 1 || (euler2 && ib >= 2 && ib <= 5) || (ns && ib >= 2 && ib <= 4) || (ns && ib == 6);
    if (!ok) return fail(ctx, HDGX_ERR_UNSUPPORTED, "boundary model " + std::to_string(ib) + " not implemented");
    d->bcm[i] = ib;
    for (int c = 0; c < a->ncu; ++c) d->bcs[i * 5 + c] = a->bcs[i * HDGX_MAX_PARAM + c];
  }
  if (a->localsolve && !(a->fc_q != 0.0)) return fail(ctx, HDGX_ERR_ARG, "app: fc_q must be nonzero");
  return HDGX_OK;
}

// ---------------------------------------------------------------- dispatch
template <class C>
int launch_quad(hdgx_ctx* ctx, const MasterBuf& mb, const DevApp& ap, const QuadArgs& qa, bool tile = false) {
  if (mb.npv != C::NPV || mb.npf != C::NPF || mb.ngv != C::NGV || mb.ngf != C::NGF)
    return fail(ctx, HDGX_ERR_UNSUPPORTED, "master quadrature does not match the compiled pgauss = 2p configuration");
  if (qa.ne == 0) return HDGX_OK;
  HIP_TRY(ctx, quad_launch<C>(ctx->stream, qa, mb.tab, ap, tile));
  return HDGX_OK;
}

template <class C>
int launch_resid(hdgx_ctx* ctx, const MasterBuf& mb, const DevApp& ap, const QuadArgs& qa) {
  if (mb.npv != C::NPV || mb.npf != C::NPF || mb.ngv != C::NGV || mb.ngf != C::NGF)
    return fail(ctx, HDGX_ERR_UNSUPPORTED, "master quadrature does not match the compiled pgauss = 2p configuration");
  if (qa.ne == 0) return HDGX_OK;
  HIP_TRY(ctx, resid_launch<C>(ctx->stream, qa, mb.tab, ap));
  return HDGX_OK;
}

template <class C>
int launch_getq(hdgx_ctx* ctx, const MasterBuf& mb, const GetqArgs& ga) {
  if (mb.npv != C::NPV || mb.npf != C::NPF || mb.ngv != C::NGV || mb.ngf != C::NGF)
    return fail(ctx, HDGX_ERR_UNSUPPORTED, "master quadrature does not match the compiled pgauss = 2p configuration");
  if (ga.ne == 0) return HDGX_OK;
  HIP_TRY(ctx, getq_launch<C>(ctx->stream, ga, mb.tab));
  return HDGX_OK;
}

template <class C>
int launch_schur(hdgx_ctx* ctx, const SchurArgs& sa, const int32_t* perm, bool tile = false) {
  if (sa.ne == 0) return HDGX_OK;
  HIP_TRY(ctx, schur_launch<C>(ctx->stream, sa, perm, tile));
  return HDGX_OK;
}

// The configurations built into the library.
template <class F>
int with_cfg(int phys, int nd, int p, F&& f) {
#define HDGX_CASE(PH, ND, P) if (phys == PH && nd == ND && p == P) return f(Cfg<PH, ND, P>{});
  HDGX_FOR_EACH_CFG(HDGX_CASE)
#undef HDGX_CASE
  return -1;
}

template <class F>
int with_family(int phys, int nd, F&& f) {
#define HDGX_FCASE(PH, ND) if (phys == PH && nd == ND) return f(std::integral_constant<int, PH>{}, std::integral_constant<int, ND>{});
  HDGX_FOR_EACH_FAMILY(HDGX_FCASE)
#undef HDGX_FCASE
  return -1;
}

// per-element words of the element-matrix workspace of configuration
// (phys, nd, p), the larger of its reference and tile layouts (hdgx_shapes.h);
// 0 when no kernel is built for it
size_t ws_words_of(int phys, int nd, int p) {
  size_t w = 0;
  with_cfg(phys, nd, p, [&](auto cfg) {
    w = ws_words<decltype(cfg)>();
    return 0;
  });
  return w;
}

int porder_of(int nd, int npv) {
  for (int p = 1; p <= 6; ++p) {
    int n = nd == 2 ? (p + 1) * (p + 2) / 2 : (p + 1) * (p + 2) * (p + 3) / 6;
    if (n == npv) return p;
  }
  return -1;
}

int ensure_ws(hdgx_ctx* ctx, size_t bytes) {
  if (ctx->ws_bytes >= bytes) return HDGX_OK;
  if (ctx->ws) HIP_TRY(ctx, hipFree(ctx->ws));
  ctx->ws = nullptr;
  ctx->ws_bytes = 0;
  HIP_TRY(ctx, hipMalloc(&ctx->ws, bytes));
  ctx->ws_bytes = bytes;
  return HDGX_OK;
}

// hdgx_elemental's element loop over chunks of `chunk` elements, the element
// matrices of each chunk in the workspace at ws (ws_words_of * chunk
// doubles): the quadrature kernel writes them, the elimination reads them.
// The FP64 eliminations of up to three row tiles take F, GK and H in the tile
// layout of hdgx_shapes.h.
int elemental_launches(hdgx_ctx* ctx, const MasterBuf& mb, const DevApp& ap, const hdgx_master* master,
                       const hdgx_app* app, int64_t ne, const double* dgnodes, const int32_t* bf, const double* udg,
                       const double* uh, const double* sh, double* ae, double* fe, double* dudg, double* dudg_duh,
                       int32_t* status, double* ws, int64_t chunk) {
  const int nco = app->getdqdg == 1 ? app->nc : (app->getdqdg == 0 ? app->ncu : 0);
  const int nd = master->nd, npv = master->npv, nc = app->nc;
  const size_t ldo = (size_t)npv * nco;
  const int ndf = master->npf * master->nfe, nh = app->ncu * ndf;
  const int p = porder_of(nd, npv);
  const bool uq = app->localsolve != 0;
  for (int64_t e0 = 0; e0 < ne; e0 += chunk) {
    const int64_t n = std::min(chunk, ne - e0);
    const int rc = with_cfg(app->physics, nd, p, [&](auto cfg) {
      using Cc = decltype(cfg);
      const bool tile = ctx->prec == HDGX_PREC_FP64 && WsLayout<Cc>::TILE_OK;
      const WsPtrs W = ws_carve<Cc>(ws, n, tile);
      QuadArgs qa{n, dgnodes + (size_t)e0 * npv * app->ncd, bf + (size_t)e0 * master->nfe,
                  udg + (size_t)e0 * npv * nc, uh + (size_t)e0 * nh, sh ? sh + (size_t)e0 * npv * nc : nullptr,
                  W.BD, W.F, W.GK, W.H, W.Ru, W.Rh, W.M, W.C, W.E, W.Rq, nullptr};
      SchurArgs sa{n, W.M, W.C, W.E, W.BD, W.F, W.GK, W.H, W.Rq, W.Ru, W.Rh, nco ? dudg + (size_t)e0 * ldo : nullptr,
                   nco ? dudg_duh + (size_t)e0 * ldo * nh : nullptr, ae + (size_t)e0 * nh * nh, fe + (size_t)e0 * nh,
                   status ? status + e0 : nullptr, ctx->prec, nco, mb.tab.aek[uq ? 1 : 0], 1};
      if (ctx->timing) HIP_TRY(ctx, hipEventRecord(ctx->tev[0], ctx->stream));
      int r = launch_quad<Cc>(ctx, mb, ap, qa, tile);
      if (r) return r;
      if (ctx->timing) HIP_TRY(ctx, hipEventRecord(ctx->tev[1], ctx->stream));
      r = launch_schur<Cc>(ctx, sa, mb.tab.perm, tile);
      if (r) return r;
      if (ctx->timing) {
        HIP_TRY(ctx, hipEventRecord(ctx->tev[2], ctx->stream));
        HIP_TRY(ctx, hipEventSynchronize(ctx->tev[2]));
        float q = 0.f, el = 0.f;
        HIP_TRY(ctx, hipEventElapsedTime(&q, ctx->tev[0], ctx->tev[1]));
        HIP_TRY(ctx, hipEventElapsedTime(&el, ctx->tev[1], ctx->tev[2]));
        ctx->tsum[0] += q;
        ctx->tsum[1] += el;
        ++ctx->tcalls;
      }
      return (int)HDGX_OK;
    });
    if (rc < 0) return fail(ctx, HDGX_ERR_UNSUPPORTED, "no kernel for this (physics, nd, p)");
    if (rc) return rc;
  }
  return HDGX_OK;
}

// HDGX_HOST_PTRS: an entry point's array arguments are host memory.  The
// arrays are given device copies in the context's staging buffer (inputs
// copied in, outputs copied back after the call, which then synchronises),
// and the call runs on those copies with the flag cleared.
class HostStage {
 public:
  explicit HostStage(hdgx_ctx* ctx) : ctx_(ctx) {}
  template <class T> void in(const T*& p, size_t n) { add(p, n * sizeof(T), true, false, [&p](void* d) { p = (const T*)d; }); }
  template <class T> void out(T*& p, size_t n) { add(p, n * sizeof(T), false, true, [&p](void* d) { p = (T*)d; }); }
  template <class T> void inout(T*& p, size_t n) { add(p, n * sizeof(T), true, true, [&p](void* d) { p = (T*)d; }); }

  // copies the inputs in and points every registered argument at its copy
  int bind() {
    hdgx_ctx* ctx = ctx_;
    HIP_TRY(ctx, hipSetDevice(ctx->device));
    if (total_ > ctx->hws_bytes) {
      if (ctx->hws) HIP_TRY(ctx, hipFree(ctx->hws));
      ctx->hws = nullptr;
      ctx->hws_bytes = 0;
      HIP_TRY(ctx, hipMalloc(&ctx->hws, total_));
      ctx->hws_bytes = total_;
    }
    char* base = static_cast<char*>(ctx->hws);
    for (Item& it : items_) {
      it.dev = base + it.off;
      if (it.in) HIP_TRY(ctx, hipMemcpyAsync(it.dev, it.host, it.bytes, hipMemcpyHostToDevice, ctx->stream));
      it.set(it.dev);
    }
    return HDGX_OK;
  }
  // copies the outputs back and waits for them
  int finish() {
    hdgx_ctx* ctx = ctx_;
    for (Item& it : items_)
      if (it.out) HIP_TRY(ctx, hipMemcpyAsync(const_cast<void*>(it.host), it.dev, it.bytes, hipMemcpyDeviceToHost, ctx->stream));
    HIP_TRY(ctx, hipStreamSynchronize(ctx->stream));
    return HDGX_OK;
  }

 private:
  struct Item {
    const void* host;
    size_t bytes, off;
    bool in, out;
    std::function<void(void*)> set;
    void* dev;
  };
  void add(const void* host, size_t bytes, bool in, bool out, std::function<void(void*)> set) {
    if (!host || bytes == 0) return;
    items_.push_back({host, bytes, total_, in, out, std::move(set), nullptr});
    total_ += (bytes + 255) / 256 * 256;
  }
  hdgx_ctx* ctx_;
  std::vector<Item> items_;
  size_t total_ = 0;
};

// runs `call` with HDGX_HOST_PTRS cleared, then restores the flags
template <class F>
int without_host_ptrs(hdgx_ctx* ctx, F&& call) {
  const int saved = ctx->flags;
  ctx->flags &= ~HDGX_HOST_PTRS;
  const int rc = call();
  ctx->flags = saved;
  return rc;
}

// ------------------------------------------- the allocation-free entry point
// hdgx_elemental_noalloc keeps its master tables and its element-matrix
// workspace in this file-scope device array instead of memory it allocates:
// the first NOALLOC_TABLE_BYTES hold the packed master tables, the rest the
// chunked workspace.  Every GPU the library runs on holds one copy for the
// life of the process; all calls on one GPU share it.
constexpr size_t NOALLOC_BYTES = (size_t)128 << 20;
constexpr size_t NOALLOC_TABLE_BYTES = (size_t)1 << 20;
static_assert(NOALLOC_TABLE_BYTES % 256 == 0, "the workspace part stays 256-byte aligned");

// Host code must not pass this name to a kernel or a copy: on the host it is
// a stand-in, not a device address.  Its address on each device is taken once
// with hipGetSymbolAddress (noalloc_setup) and only that is used.
static __device__ __align__(256) char scratch[NOALLOC_BYTES];

constexpr int NOALLOC_MAXDEV = 64;
struct NoallocDevice {
  std::mutex mu;
  int state = 0;             // 0: address not taken yet, 1: ptr valid, -1: hipGetSymbolAddress failed
  char* ptr = nullptr;       // scratch's device address on this device
  uint64_t table_hash = 0;   // content hash of the master tables now in the table area (0: none)
  MasterBuf mb;              // their device view
};
NoallocDevice g_noalloc[NOALLOC_MAXDEV];

// once per device, with that device current (from hdgx_create)
void noalloc_setup(int device) {
  if (device < 0 || device >= NOALLOC_MAXDEV) return;
  NoallocDevice& d = g_noalloc[device];
  std::lock_guard<std::mutex> lock(d.mu);
  if (d.state != 0) return;
  void* p = nullptr;
  const hipError_t e = hipGetSymbolAddress(&p, HIP_SYMBOL(scratch));
  if (e == hipSuccess && p) {
    d.ptr = static_cast<char*>(p);
    d.state = 1;
  } else {
    (void)hipGetLastError();
    d.state = -1;
  }
}

}  // namespace

extern "C" {

int hdgx_create(hdgx_ctx** ctx, int device) {
  if (!ctx) return HDGX_ERR_ARG;
  int n = 0;
  if (hipGetDeviceCount(&n) != hipSuccess || device < 0 || device >= n) return HDGX_ERR_DEVICE;
  if (hipSetDevice(device) != hipSuccess) return HDGX_ERR_DEVICE;
  noalloc_setup(device);
  auto* c = new hdgx_ctx();
  c->device = device;
  *ctx = c;
  return HDGX_OK;
}

int hdgx_destroy(hdgx_ctx* ctx) {
  if (!ctx) return HDGX_ERR_ARG;
  (void)hipSetDevice(ctx->device);
  for (auto& kv : ctx->masters) (void)hipFree(kv.second.dev);
  if (ctx->ws) (void)hipFree(ctx->ws);
  if (ctx->tws) (void)hipFree(ctx->tws);
  if (ctx->segdev) (void)hipFree(ctx->segdev);
  if (ctx->hws) (void)hipFree(ctx->hws);
  if (ctx->mws) (void)hipFree(ctx->mws);
  if (ctx->pinned) (void)hipHostFree(ctx->pinned);
  for (hipEvent_t ev : ctx->tev)
    if (ev) (void)hipEventDestroy(ev);
  delete ctx;
  return HDGX_OK;
}

int hdgx_set_stream(hdgx_ctx* ctx, void* stream) {
  if (!ctx) return HDGX_ERR_ARG;
  ctx->stream = static_cast<hipStream_t>(stream);
  return HDGX_OK;
}

int hdgx_set_precision(hdgx_ctx* ctx, int precision) {
  if (!ctx) return HDGX_ERR_ARG;
  if (precision != HDGX_PREC_FP64 && precision != HDGX_PREC_FP32_IR) return fail(ctx, HDGX_ERR_ARG, "precision");
  ctx->prec = precision;
  return HDGX_OK;
}

int hdgx_set_workspace_limit(hdgx_ctx* ctx, int64_t bytes) {
  if (!ctx) return HDGX_ERR_ARG;
  if (bytes < 0) return fail(ctx, HDGX_ERR_ARG, "workspace limit < 0");
  ctx->ws_cap = bytes == 0 ? ((size_t)24 << 30) : (size_t)bytes;
  if (ctx->ws && ctx->ws_bytes > ctx->ws_cap) {
    // work already queued on the context's stream may still read it
    HIP_TRY(ctx, hipSetDevice(ctx->device));
    HIP_TRY(ctx, hipStreamSynchronize(ctx->stream));
    HIP_TRY(ctx, hipFree(ctx->ws));
    ctx->ws = nullptr;
    ctx->ws_bytes = 0;
  }
  return HDGX_OK;
}

int hdgx_set_timing(hdgx_ctx* ctx, int on) {
  if (!ctx) return HDGX_ERR_ARG;
  HIP_TRY(ctx, hipSetDevice(ctx->device));
  for (int i = 0; i < 3 && on; ++i)
    if (!ctx->tev[i]) HIP_TRY(ctx, hipEventCreate(&ctx->tev[i]));
  ctx->timing = on != 0;
  ctx->tsum[0] = ctx->tsum[1] = 0.0;
  ctx->tcalls = 0;
  return HDGX_OK;
}

int hdgx_kernel_times(hdgx_ctx* ctx, double* quad_ms, double* elim_ms, int64_t* launches) {
  if (!ctx || !quad_ms || !elim_ms) return HDGX_ERR_ARG;
  *quad_ms = ctx->tsum[0];
  *elim_ms = ctx->tsum[1];
  if (launches) *launches = ctx->tcalls;
  return HDGX_OK;
}

int hdgx_set_flags(hdgx_ctx* ctx, int flags) {
  if (!ctx) return HDGX_ERR_ARG;
  if (flags & ~(HDGX_INPLACE | HDGX_HOST_PTRS)) return fail(ctx, HDGX_ERR_ARG, "flags");
  ctx->flags = flags;
  return HDGX_OK;
}

const char* hdgx_last_error(hdgx_ctx* ctx) { return ctx ? ctx->err.c_str() : "null context"; }

int hdgx_elemental_matrices(hdgx_ctx* ctx, const hdgx_master* master, const hdgx_app* app, int64_t ne,
                            const double* dgnodes, const int32_t* bf, const double* udg, const double* uh,
                            const double* sh, double* BD, double* F, double* GK, double* H, double* Ru,
                            double* Rh, double* M, double* C, double* E, double* Rq) {
  if (!ctx) return HDGX_ERR_ARG;
  if (ne < 0) return fail(ctx, HDGX_ERR_ARG, "ne < 0");
  if (ctx->flags & HDGX_HOST_PTRS) {
    if (!master || !app) return fail(ctx, HDGX_ERR_ARG, "null master/app");
    const size_t n = (size_t)ne, npv = master->npv, ndf = (size_t)master->npf * master->nfe, nd = master->nd;
    const size_t ncu = app->ncu, nc = app->nc, nh = ncu * ndf;
    HostStage st(ctx);
    st.in(dgnodes, npv * app->ncd * n); st.in(bf, master->nfe * n); st.in(udg, npv * nc * n); st.in(uh, nh * n);
    st.in(sh, npv * nc * n);
    st.out(BD, npv * ncu * npv * nc * n); st.out(F, npv * ncu * nh * n); st.out(GK, nh * npv * nc * n);
    st.out(H, nh * nh * n); st.out(Ru, npv * ncu * n); st.out(Rh, nh * n); st.out(M, npv * npv * n);
    st.out(C, npv * npv * nd * n); st.out(E, npv * ndf * nd * n); st.out(Rq, npv * ncu * nd * n);
    int rc = st.bind();
    if (!rc) rc = without_host_ptrs(ctx, [&] {
      return hdgx_elemental_matrices(ctx, master, app, ne, dgnodes, bf, udg, uh, sh, BD, F, GK, H, Ru, Rh, M, C, E, Rq);
    });
    return rc ? rc : st.finish();
  }
  HIP_TRY(ctx, hipSetDevice(ctx->device));
  MasterBuf* mb = nullptr;
  int rc = get_master(ctx, master, &mb);
  if (rc) return rc;
  DevApp ap;
  if ((rc = make_app(ctx, master, app, &ap))) return rc;
  if (ne > 0 && (!dgnodes || !bf || !udg || !uh || !BD || !F || !GK || !H || !Ru || !Rh))
    return fail(ctx, HDGX_ERR_ARG, "null array");
  if (ne > 0 && app->localsolve && (!M || !C || !E || !Rq)) return fail(ctx, HDGX_ERR_ARG, "localsolve needs M, C, E, Rq");
  if (ne > 0 && (app->localsolve || app->tdep) && !sh && app->tdep) return fail(ctx, HDGX_ERR_ARG, "tdep needs sh");
  QuadArgs qa{ne, dgnodes, bf, udg, uh, sh, BD, F, GK, H, Ru, Rh, M, C, E, Rq, nullptr};
  const int p = porder_of(master->nd, master->npv);
  rc = with_cfg(app->physics, master->nd, p, [&](auto cfg) { return launch_quad<decltype(cfg)>(ctx, *mb, ap, qa); });
  if (rc < 0) return fail(ctx, HDGX_ERR_UNSUPPORTED, "no kernel for this (physics, nd, p)");
  return rc;
}

int hdgx_residual(hdgx_ctx* ctx, const hdgx_master* master, const hdgx_app* app, int64_t ne, const double* dgnodes,
                  const int32_t* bf, const double* udg, const double* uh, const double* sh, double* Ru, double* Rh,
                  double* Rq) {
  if (!ctx) return HDGX_ERR_ARG;
  if (ne < 0) return fail(ctx, HDGX_ERR_ARG, "ne < 0");
  if (ctx->flags & HDGX_HOST_PTRS) {
    if (!master || !app) return fail(ctx, HDGX_ERR_ARG, "null master/app");
    const size_t n = (size_t)ne, npv = master->npv, ndf = (size_t)master->npf * master->nfe, nd = master->nd;
    const size_t ncu = app->ncu, nc = app->nc, nh = ncu * ndf;
    HostStage st(ctx);
    st.in(dgnodes, npv * app->ncd * n); st.in(bf, master->nfe * n); st.in(udg, npv * nc * n); st.in(uh, nh * n);
    st.in(sh, npv * nc * n);
    st.out(Ru, npv * ncu * n); st.out(Rh, nh * n); st.out(Rq, npv * ncu * nd * n);
    int rc = st.bind();
    if (!rc) rc = without_host_ptrs(ctx, [&] {
      return hdgx_residual(ctx, master, app, ne, dgnodes, bf, udg, uh, sh, Ru, Rh, Rq);
    });
    return rc ? rc : st.finish();
  }
  HIP_TRY(ctx, hipSetDevice(ctx->device));
  MasterBuf* mb = nullptr;
  int rc = get_master(ctx, master, &mb);
  if (rc) return rc;
  DevApp ap;
  if ((rc = make_app(ctx, master, app, &ap))) return rc;
  if (ne > 0 && (!dgnodes || !bf || !udg || !uh || !Ru || !Rh)) return fail(ctx, HDGX_ERR_ARG, "null array");
  if (ne > 0 && app->localsolve && !Rq) return fail(ctx, HDGX_ERR_ARG, "localsolve needs Rq");
  if (ne > 0 && app->tdep && !sh) return fail(ctx, HDGX_ERR_ARG, "tdep needs sh");
  QuadArgs qa{ne, dgnodes, bf, udg, uh, sh, nullptr, nullptr, nullptr, nullptr, Ru, Rh,
              nullptr, nullptr, nullptr, Rq, nullptr};
  const int p = porder_of(master->nd, master->npv);
  rc = with_cfg(app->physics, master->nd, p, [&](auto cfg) { return launch_resid<decltype(cfg)>(ctx, *mb, ap, qa); });
  if (rc < 0) return fail(ctx, HDGX_ERR_UNSUPPORTED, "no kernel for this (physics, nd, p)");
  return rc;
}

int hdgx_getq(hdgx_ctx* ctx, const hdgx_master* master, int ncu, int nc, int ncd, int64_t ne,
              const double* dgnodes, const double* udg, const double* uh, const double* sh, double fc_q,
              double* qdg, int64_t ldq, const double* shapnv, double* qq) {
  if (!ctx) return HDGX_ERR_ARG;
  if (!master) return fail(ctx, HDGX_ERR_ARG, "master: null");
  const int nd = master->nd, ncq = ncu * nd;
  if (ne < 0) return fail(ctx, HDGX_ERR_ARG, "ne < 0");
  if ((ctx->flags & HDGX_HOST_PTRS) && ne > 0 && ldq >= (int64_t)master->npv * ncq && ncu > 0 && nc > 0) {
    const size_t n = (size_t)ne, npv = master->npv, ndf = (size_t)master->npf * master->nfe;
    HostStage st(ctx);
    st.in(dgnodes, npv * ncd * n); st.in(udg, npv * nc * n); st.in(uh, ncu * ndf * n); st.in(sh, npv * nc * n);
    st.in(shapnv, npv * npv * (nd + 1));
    // the q blocks may sit inside udg (ldq = npv*nc): the whole strided range
    // goes in and out so that the columns between them are kept
    st.inout(qdg, (n - 1) * ldq + npv * ncq);
    st.out(qq, npv * ncq * n);
    int rc = st.bind();
    if (!rc) rc = without_host_ptrs(ctx, [&] {
      return hdgx_getq(ctx, master, ncu, nc, ncd, ne, dgnodes, udg, uh, sh, fc_q, qdg, ldq, shapnv, qq);
    });
    return rc ? rc : st.finish();
  }
  if (ncu != nd + 2) return fail(ctx, HDGX_ERR_UNSUPPORTED, "getq: ncu must be nd+2");
  if (nc < ncu) return fail(ctx, HDGX_ERR_ARG, "getq: nc < ncu");
  if (sh && nc < ncq) return fail(ctx, HDGX_ERR_ARG, "getq: sh needs nc >= ncu*nd");
  if (ncd < nd) return fail(ctx, HDGX_ERR_ARG, "getq: ncd < nd");
  if (ldq < (int64_t)master->npv * ncq) return fail(ctx, HDGX_ERR_ARG, "getq: ldq < npv*ncu*nd");
  if (!(fc_q != 0.0) || !std::isfinite(fc_q)) return fail(ctx, HDGX_ERR_ARG, "getq: fc_q must be finite and nonzero");
  if ((shapnv == nullptr) != (qq == nullptr)) return fail(ctx, HDGX_ERR_ARG, "getq: shapnv and qq go together");
  if (ne > 0 && (!dgnodes || !udg || !uh || !qdg)) return fail(ctx, HDGX_ERR_ARG, "null array");
  HIP_TRY(ctx, hipSetDevice(ctx->device));
  MasterBuf* mb = nullptr;
  int rc = get_master(ctx, master, &mb);
  if (rc) return rc;
  GetqArgs ga{ne, nc, ncd, fc_q, dgnodes, udg, uh, sh, qdg, ldq, shapnv, qq};
  const int p = porder_of(nd, master->npv);
  rc = with_cfg(HDGX_PHYS_EULER, nd, p, [&](auto cfg) { return launch_getq<decltype(cfg)>(ctx, *mb, ga); });
  if (rc < 0) return fail(ctx, HDGX_ERR_UNSUPPORTED, "getq: no kernel for this (nd, p)");
  return rc;
}

int hdgx_schur_u(hdgx_ctx* ctx, int npv, int ncu, int nc, int ndf, int64_t ne, const double* BD, const double* F,
                 const double* GK, const double* H, const double* Ru, const double* Rh, double* dudg,
                 double* dudg_duh, double* ae, double* fe, int32_t* status) {
  if (!ctx) return HDGX_ERR_ARG;
  if (ne < 0 || nc != ncu) return fail(ctx, HDGX_ERR_ARG, "schur_u: nc must equal ncu");
  if (ctx->flags & HDGX_HOST_PTRS) {
    const size_t n = (size_t)ne, nu = (size_t)npv * ncu, nh = (size_t)ncu * ndf, ncol = (size_t)npv * nc;
    const bool inplace = ctx->flags & HDGX_INPLACE;
    HostStage st(ctx);
    double* bd = const_cast<double*>(BD);
    if (inplace) st.inout(bd, nu * ncol * n); else st.in(BD, nu * ncol * n);
    st.in(F, nu * nh * n); st.in(GK, nh * ncol * n); st.in(H, nh * nh * n); st.in(Ru, nu * n); st.in(Rh, nh * n);
    st.out(dudg, ncol * n); st.out(dudg_duh, ncol * nh * n); st.out(ae, nh * nh * n); st.out(fe, nh * n);
    st.out(status, n);
    int rc = st.bind();
    if (!rc) rc = without_host_ptrs(ctx, [&] {
      return hdgx_schur_u(ctx, npv, ncu, nc, ndf, ne, inplace ? bd : BD, F, GK, H, Ru, Rh, dudg, dudg_duh, ae, fe,
                          status);
    });
    return rc ? rc : st.finish();
  }
  if (ne > 0 && (!BD || !F || !GK || !H || !Ru || !Rh || !dudg || !dudg_duh || !ae || !fe))
    return fail(ctx, HDGX_ERR_ARG, "null array");
  HIP_TRY(ctx, hipSetDevice(ctx->device));
  const int nd = ncu - 2;
  const int p = porder_of(nd, npv);
  SchurArgs sa{ne, nullptr, nullptr, nullptr, BD, F, GK, H, nullptr, Ru, Rh, dudg, dudg_duh, ae, fe, status, ctx->prec, nc};
  int rc = with_cfg(0, nd, p, [&](auto cfg) {
    using Cc = decltype(cfg);
    if (Cc::NDF != ndf) return fail(ctx, HDGX_ERR_UNSUPPORTED, "schur_u: ndf does not match a simplex of this order");
    return launch_schur<Cc>(ctx, sa, nullptr);
  });
  if (rc < 0) return fail(ctx, HDGX_ERR_UNSUPPORTED, "schur_u: no kernel for these sizes");
  if (rc == HDGX_OK && (ctx->flags & HDGX_INPLACE) && ne > 0) {
    // schur_u.c:67: BD -> its LU (the caller's buffer, declared const, is written)
    InplaceArgs ia{ne, npv, ncu, nd, ndf, 0, nullptr, nullptr, nullptr, const_cast<double*>(BD), nullptr, nullptr,
                   nullptr};
    HIP_TRY(ctx, inplace_launch(ctx->stream, ia));
  }
  return rc;
}

int hdgx_schur_uq(hdgx_ctx* ctx, int nd, int npv, int ncu, int nc, int ndf, int64_t ne, const double* M,
                  const double* C, const double* E, const double* BD, const double* F, const double* GK,
                  const double* H, const double* Rq, const double* Ru, const double* Rh, double* dudg,
                  double* dudg_duh, double* ae, double* fe, int32_t* status) {
  if (!ctx) return HDGX_ERR_ARG;
  if (ne < 0 || nc != ncu * (nd + 1)) return fail(ctx, HDGX_ERR_ARG, "schur_uq: nc must be ncu*(nd+1)");
  if (ctx->flags & HDGX_HOST_PTRS) {
    const size_t n = (size_t)ne, nu = (size_t)npv * ncu, nh = (size_t)ncu * ndf, ncol = (size_t)npv * nc;
    const bool inplace = ctx->flags & HDGX_INPLACE;
    HostStage st(ctx);
    // the arrays HDGX_INPLACE writes go both ways
    double *m = const_cast<double*>(M), *c = const_cast<double*>(C), *e = const_cast<double*>(E);
    double *bd = const_cast<double*>(BD), *f = const_cast<double*>(F), *rq = const_cast<double*>(Rq);
    double* ru = const_cast<double*>(Ru);
    const size_t sm = (size_t)npv * npv * n, sc = sm * nd, se = (size_t)npv * ndf * nd * n;
    const size_t sbd = nu * ncol * n, sf = nu * nh * n, srq = nu * nd * n, sru = nu * n;
    if (inplace) {
      st.inout(m, sm); st.inout(c, sc); st.inout(e, se); st.inout(bd, sbd); st.inout(f, sf); st.inout(rq, srq);
      st.inout(ru, sru);
    } else {
      st.in(M, sm); st.in(C, sc); st.in(E, se); st.in(BD, sbd); st.in(F, sf); st.in(Rq, srq); st.in(Ru, sru);
    }
    st.in(GK, nh * ncol * n); st.in(H, nh * nh * n); st.in(Rh, nh * n);
    st.out(dudg, ncol * n); st.out(dudg_duh, ncol * nh * n); st.out(ae, nh * nh * n); st.out(fe, nh * n);
    st.out(status, n);
    int rc = st.bind();
    if (!rc) rc = without_host_ptrs(ctx, [&] {
      return inplace ? hdgx_schur_uq(ctx, nd, npv, ncu, nc, ndf, ne, m, c, e, bd, f, GK, H, rq, ru, Rh, dudg, dudg_duh,
                                     ae, fe, status)
                     : hdgx_schur_uq(ctx, nd, npv, ncu, nc, ndf, ne, M, C, E, BD, F, GK, H, Rq, Ru, Rh, dudg, dudg_duh,
                                     ae, fe, status);
    });
    return rc ? rc : st.finish();
  }
  if (ne > 0 && (!M || !C || !E || !BD || !F || !GK || !H || !Rq || !Ru || !Rh || !dudg || !dudg_duh || !ae || !fe))
    return fail(ctx, HDGX_ERR_ARG, "null array");
  if (nd != 2 || ncu != 4) return fail(ctx, HDGX_ERR_UNSUPPORTED, "schur_uq: 2D with ncu = 4 is built");
  HIP_TRY(ctx, hipSetDevice(ctx->device));
  const int p = porder_of(nd, npv);
  SchurArgs sa{ne, M, C, E, BD, F, GK, H, Rq, Ru, Rh, dudg, dudg_duh, ae, fe, status, ctx->prec, nc};
  int rc = with_cfg(1, nd, p, [&](auto cfg) {
    using Cc = decltype(cfg);
    if (Cc::NDF != ndf) return fail(ctx, HDGX_ERR_UNSUPPORTED, "schur_uq: ndf does not match");
    return launch_schur<Cc>(ctx, sa, nullptr);
  });
  if (rc < 0) return fail(ctx, HDGX_ERR_UNSUPPORTED, "schur_uq: no kernel for these sizes");
  if (rc == HDGX_OK && (ctx->flags & HDGX_INPLACE) && ne > 0) {
    // schur_uq.c:localsolveuq's side effects on M, C, E, BD, F, Rq, Ru (the
    // caller's buffers, declared const, are written)
    InplaceArgs ia{ne, npv, ncu, nd, ndf, 1, const_cast<double*>(M), const_cast<double*>(C), const_cast<double*>(E),
                   const_cast<double*>(BD), const_cast<double*>(F), const_cast<double*>(Rq), const_cast<double*>(Ru)};
    HIP_TRY(ctx, inplace_launch(ctx->stream, ia));
  }
  return rc;
}

int hdgx_elemental(hdgx_ctx* ctx, const hdgx_master* master, const hdgx_app* app, int64_t ne,
                   const double* dgnodes, const int32_t* bf, const double* udg, const double* uh,
                   const double* sh, double* ae, double* fe, double* dudg, double* dudg_duh, int32_t* status) {
  if (!ctx) return HDGX_ERR_ARG;
  if (ne < 0) return fail(ctx, HDGX_ERR_ARG, "ne < 0");
  if (!master || !app) return fail(ctx, HDGX_ERR_ARG, "null master/app");
  if (ctx->flags & HDGX_HOST_PTRS) {
    const size_t n = (size_t)ne, npv = master->npv, ndf = (size_t)master->npf * master->nfe;
    const size_t ncu = app->ncu, nc = app->nc, nh = ncu * ndf;
    const size_t nco = app->getdqdg == 1 ? nc : (app->getdqdg == 0 ? ncu : 0);
    HostStage st(ctx);
    st.in(dgnodes, npv * app->ncd * n); st.in(bf, master->nfe * n); st.in(udg, npv * nc * n); st.in(uh, nh * n);
    st.in(sh, npv * nc * n);
    st.out(ae, nh * nh * n); st.out(fe, nh * n); st.out(dudg, npv * nco * n); st.out(dudg_duh, npv * nco * nh * n);
    st.out(status, n);
    int rc = st.bind();
    if (!rc) rc = without_host_ptrs(ctx, [&] {
      return hdgx_elemental(ctx, master, app, ne, dgnodes, bf, udg, uh, sh, ae, fe, dudg, dudg_duh, status);
    });
    return rc ? rc : st.finish();
  }
  // DUDG / DUDG_DUH keep nco of the nc rows (hdg_elemental.m:20-26, 146-147)
  const int nco = app->getdqdg == 1 ? app->nc : (app->getdqdg == 0 ? app->ncu : 0);
  if (ne > 0 && (!dgnodes || !bf || !udg || !uh || !ae || !fe || (nco > 0 && (!dudg || !dudg_duh))))
    return fail(ctx, HDGX_ERR_ARG, "null array");
  HIP_TRY(ctx, hipSetDevice(ctx->device));
  MasterBuf* mb = nullptr;
  int rc = get_master(ctx, master, &mb);
  if (rc) return rc;
  DevApp ap;
  if ((rc = make_app(ctx, master, app, &ap))) return rc;
  if (app->tdep && !sh) return fail(ctx, HDGX_ERR_ARG, "tdep needs sh");
  if (ne == 0) return HDGX_OK;
  // per-element workspace for the Schur inputs
  const size_t per = ws_words_of(app->physics, master->nd, porder_of(master->nd, master->npv));
  if (per == 0) return fail(ctx, HDGX_ERR_UNSUPPORTED, "no kernel for this (physics, nd, p)");
  const int64_t chunk = ws_chunk(ne, per * 8, ctx->ws_cap);
  if ((rc = ensure_ws(ctx, per * 8 * (size_t)chunk))) return rc;
  return elemental_launches(ctx, *mb, ap, master, app, ne, dgnodes, bf, udg, uh, sh, ae, fe, dudg, dudg_duh, status,
                            static_cast<double*>(ctx->ws), chunk);
}

int hdgx_local_solve(hdgx_ctx* ctx, const hdgx_master* master, const hdgx_app* app, int64_t ne,
                     const double* dgnodes, const int32_t* bf, const double* udg, const double* uh, const double* sh,
                     const int32_t* elcon, const double* DUH, double* dudgt, int32_t* status) {
  if (!ctx) return HDGX_ERR_ARG;
  if (ne < 0) return fail(ctx, HDGX_ERR_ARG, "ne < 0");
  if (!master || !app) return fail(ctx, HDGX_ERR_ARG, "null master/app");
  if (ctx->flags & HDGX_HOST_PTRS) return fail(ctx, HDGX_ERR_ARG, "local_solve: device arrays only");
  if (ne > 0 && (!dgnodes || !bf || !udg || !uh || !elcon || !DUH || !dudgt)) return fail(ctx, HDGX_ERR_ARG, "null array");
  HIP_TRY(ctx, hipSetDevice(ctx->device));
  MasterBuf* mb = nullptr;
  int rc = get_master(ctx, master, &mb);
  if (rc) return rc;
  // the element solves keep every row (schur_primal's dudg, dudg_duh)
  hdgx_app full = *app;
  full.getdqdg = 1;
  DevApp ap;
  if ((rc = make_app(ctx, master, &full, &ap))) return rc;
  if (app->tdep && !sh) return fail(ctx, HDGX_ERR_ARG, "tdep needs sh");
  if (ne == 0) return HDGX_OK;
  const size_t per = ws_words_of(app->physics, master->nd, porder_of(master->nd, master->npv));
  if (per == 0) return fail(ctx, HDGX_ERR_UNSUPPORTED, "no kernel for this (physics, nd, p)");
  const size_t npv = master->npv, nc = app->nc, ncu = app->ncu, nfe = master->nfe;
  const size_t ndf = (size_t)master->npf * nfe, nh = ncu * ndf;
  // per element: the element matrices, then AE, FE, DUDG, DUDG_DUH and a
  // status word, all dropped once DUDGT is formed
  const size_t outw = nh * nh + nh + npv * nc + npv * nc * nh + 1;
  const int64_t chunk = ws_chunk(ne, (per + outw) * 8, ctx->ws_cap);
  if ((rc = ensure_ws(ctx, (per + outw) * 8 * (size_t)chunk))) return rc;
  double* const w = static_cast<double*>(ctx->ws);
  double* const sAE = w + per * (size_t)chunk;
  double* const sFE = sAE + nh * nh * (size_t)chunk;
  double* const sDU = sFE + nh * (size_t)chunk;
  double* const sDX = sDU + npv * nc * (size_t)chunk;
  int32_t* const sST = reinterpret_cast<int32_t*>(sDX + npv * nc * nh * (size_t)chunk);
  for (int64_t e0 = 0; e0 < ne; e0 += chunk) {
    const int64_t n = std::min(chunk, ne - e0);
    rc = elemental_launches(ctx, *mb, ap, master, &full, n, dgnodes + (size_t)e0 * npv * app->ncd,
                            bf + (size_t)e0 * nfe, udg + (size_t)e0 * npv * nc, uh + (size_t)e0 * nh,
                            sh ? sh + (size_t)e0 * npv * nc : nullptr, sAE, sFE, sDU, sDX, status ? status + e0 : sST,
                            w, n);
    if (rc) return rc;
    rc = hdgx_local_update(ctx, (int)npv, (int)nc, (int)ncu, master->npf, master->nfe, n, elcon + (size_t)e0 * ndf,
                           sDU, sDX, DUH, dudgt + (size_t)e0 * npv * nc);
    if (rc) return rc;
  }
  return HDGX_OK;
}

int hdgx_elemental_noalloc(hdgx_ctx* ctx, const hdgx_master* master, const hdgx_app* app, int64_t ne,
                           const double* dgnodes, const int32_t* bf, const double* udg, const double* uh,
                           const double* sh, double* ae, double* fe, double* dudg, double* dudg_duh,
                           int32_t* status) {
  if (!ctx) return HDGX_ERR_ARG;
  if (ne < 0) return fail(ctx, HDGX_ERR_ARG, "ne < 0");
  if (!master || !app) return fail(ctx, HDGX_ERR_ARG, "null master/app");
  if (ctx->flags & HDGX_HOST_PTRS)
    return fail(ctx, HDGX_ERR_ARG, "elemental_noalloc: device arrays only (HDGX_HOST_PTRS stages through allocated memory)");
  const int nco = app->getdqdg == 1 ? app->nc : (app->getdqdg == 0 ? app->ncu : 0);
  if (ne > 0 && (!dgnodes || !bf || !udg || !uh || !ae || !fe || (nco > 0 && (!dudg || !dudg_duh))))
    return fail(ctx, HDGX_ERR_ARG, "null array");
  if (ctx->device < 0 || ctx->device >= NOALLOC_MAXDEV)
    return fail(ctx, HDGX_ERR_UNSUPPORTED, "elemental_noalloc: device ordinal past the scratch table");
  NoallocDevice& d = g_noalloc[ctx->device];
  std::lock_guard<std::mutex> lock(d.mu);
  if (d.state != 1) return fail(ctx, HDGX_ERR_DEVICE, "elemental_noalloc: hipGetSymbolAddress of the scratch array failed");
  HIP_TRY(ctx, hipSetDevice(ctx->device));
  MasterImage img;
  // the image is built only when the table area holds another master's
  uint64_t hash = 0;
  int rc = master_hash(ctx, master, &hash);
  if (rc) return rc;
  const bool fresh = d.table_hash != hash;
  if (fresh) {
    if ((rc = pack_master(ctx, master, &img))) return rc;
    if (img.host.size() > NOALLOC_TABLE_BYTES)
      return fail(ctx, HDGX_ERR_ARG, "elemental_noalloc: master tables exceed the 1 MiB table area");
  }
  DevApp ap;
  if ((rc = make_app(ctx, master, app, &ap))) return rc;
  if (app->tdep && !sh) return fail(ctx, HDGX_ERR_ARG, "tdep needs sh");
  const size_t per = 8 * ws_words_of(app->physics, master->nd, porder_of(master->nd, master->npv));
  if (per == 0) return fail(ctx, HDGX_ERR_UNSUPPORTED, "no kernel for this (physics, nd, p)");
  const size_t ws_bytes = NOALLOC_BYTES - NOALLOC_TABLE_BYTES;
  if (per > ws_bytes) return fail(ctx, HDGX_ERR_ARG, "elemental_noalloc: one element's matrices exceed the scratch");
  if (ne == 0) return HDGX_OK;
  if (fresh) {
    // stream-ordered after earlier calls on this stream; the host image is a
    // local, so the copy completes before the call goes on
    HIP_TRY(ctx, hipMemcpyAsync(d.ptr, img.host.data(), img.host.size(), hipMemcpyHostToDevice, ctx->stream));
    HIP_TRY(ctx, hipStreamSynchronize(ctx->stream));
    bind_master(master, img, d.ptr, &d.mb);
    d.table_hash = img.hash;
  }
  const int64_t chunk = ws_chunk(ne, per, ws_bytes);
  return elemental_launches(ctx, d.mb, ap, master, app, ne, dgnodes, bf, udg, uh, sh, ae, fe, dudg, dudg_duh, status,
                            reinterpret_cast<double*>(d.ptr + NOALLOC_TABLE_BYTES), chunk);
}

int hdgx_elemental_batch(hdgx_ctx* ctx, const hdgx_master* master, int nseg, const hdgx_segment* segs) {
  if (!ctx) return HDGX_ERR_ARG;
  if (nseg < 0 || (nseg > 0 && !segs) || !master) return fail(ctx, HDGX_ERR_ARG, "batch: arguments");
  if (nseg == 0) return HDGX_OK;
  HIP_TRY(ctx, hipSetDevice(ctx->device));
  MasterBuf* mb = nullptr;
  int rc = get_master(ctx, master, &mb);
  if (rc) return rc;
  const hdgx_app* a0 = segs[0].app;
  if (!a0) return fail(ctx, HDGX_ERR_ARG, "batch: null app");
  std::vector<ElemSeg> tab(nseg);
  std::vector<DevApp> apps(nseg);
  int64_t total = 0;
  for (int s = 0; s < nseg; ++s) {
    const hdgx_segment& g = segs[s];
    if (!g.app) return fail(ctx, HDGX_ERR_ARG, "batch: null app");
    const hdgx_app* a = g.app;
    if (a->physics != a0->physics || a->nd != a0->nd || a->ncu != a0->ncu || a->nc != a0->nc ||
        a->localsolve != a0->localsolve)
      return fail(ctx, HDGX_ERR_ARG, "batch: segments must share physics, nd, ncu, nc and localsolve");
    if ((rc = make_app(ctx, master, a, &apps[s]))) return rc;
    if (a->getdqdg != 1) return fail(ctx, HDGX_ERR_UNSUPPORTED, "batch: getdqdg = 1 (nco = nc) only");
    if (g.ne < 0) return fail(ctx, HDGX_ERR_ARG, "batch: ne < 0");
    if (g.ne > 0 && (!g.dgnodes || !g.bf || !g.udg || !g.uh || !g.ae || !g.fe || !g.dudg || !g.dudg_duh))
      return fail(ctx, HDGX_ERR_ARG, "batch: null array");
    if (g.ne > 0 && a->tdep && !g.sh) return fail(ctx, HDGX_ERR_ARG, "batch: tdep needs sh");
    tab[s] = ElemSeg{total, g.ne, g.dgnodes, g.bf, g.udg, g.uh, g.sh, g.ae, g.fe, g.dudg, g.dudg_duh, g.status, s, 0};
    total += g.ne;
  }
  if (total == 0) return HDGX_OK;
  const int nd = master->nd, npv = master->npv, ncu = a0->ncu, nc = a0->nc;
  const int ndf = master->npf * master->nfe, nu = npv * ncu, nh = ncu * ndf, ncol = npv * nc;
  const bool uq = a0->localsolve != 0;
  const int p = porder_of(nd, npv);
  const size_t tbytes = sizeof(ElemSeg) * nseg + sizeof(DevApp) * nseg;
  if (ctx->segdev_bytes < tbytes) {
    if (ctx->segdev) HIP_TRY(ctx, hipFree(ctx->segdev));
    ctx->segdev = nullptr;
    ctx->segdev_bytes = 0;
    HIP_TRY(ctx, hipMalloc(&ctx->segdev, tbytes));
    ctx->segdev_bytes = tbytes;
  }
  ElemSeg* dsegs = static_cast<ElemSeg*>(ctx->segdev);
  DevApp* dapps = reinterpret_cast<DevApp*>(dsegs + nseg);
  HIP_TRY(ctx, hipMemcpyAsync(dsegs, tab.data(), sizeof(ElemSeg) * nseg, hipMemcpyHostToDevice, ctx->stream));
  HIP_TRY(ctx, hipMemcpyAsync(dapps, apps.data(), sizeof(DevApp) * nseg, hipMemcpyHostToDevice, ctx->stream));
  const size_t per = ws_words_of(a0->physics, nd, p);
  if (per == 0) return fail(ctx, HDGX_ERR_UNSUPPORTED, "no kernel for this (physics, nd, p)");
  const int64_t chunk = ws_chunk(total, per * 8, ctx->ws_cap);
  if ((rc = ensure_ws(ctx, per * 8 * (size_t)chunk))) return rc;
  for (int64_t g0 = 0; g0 < total; g0 += chunk) {
    const int64_t n = std::min(chunk, total - g0);
    rc = with_cfg(a0->physics, nd, p, [&](auto cfg) {
      using Cc = decltype(cfg);
      if (mb->npv != Cc::NPV || mb->npf != Cc::NPF || mb->ngv != Cc::NGV || mb->ngf != Cc::NGF)
        return fail(ctx, HDGX_ERR_UNSUPPORTED, "master quadrature does not match the compiled configuration");
      // batched launches take the tile layout where it exists (seg_kernels.hip)
      const WsPtrs W = ws_carve<Cc>(static_cast<double*>(ctx->ws), n, WsLayout<Cc>::TILE_OK);
      SegTable t{};
      t.segs = dsegs;
      t.apps = dapps;
      t.nseg = nseg;
      t.g0 = g0;
      t.total = n;
      t.BD = W.BD; t.F = W.F; t.GK = W.GK; t.H = W.H; t.Ru = W.Ru; t.Rh = W.Rh;
      t.M = W.M; t.C = W.C; t.E = W.E; t.Rq = W.Rq;
      HIP_TRY(ctx, seg_launch<Cc>(ctx->stream, t, mb->tab, mb->tab.perm));
      return (int)HDGX_OK;
    });
    if (rc < 0) return fail(ctx, HDGX_ERR_UNSUPPORTED, "no kernel for this (physics, nd, p)");
    if (rc) return rc;
  }
  return HDGX_OK;
}

int hdgx_elemental_batch_mixed(hdgx_ctx* ctx, int nseg, const hdgx_master* const* masters,
                               const hdgx_segment* segs) {
  if (!ctx) return HDGX_ERR_ARG;
  if (nseg < 0 || (nseg > 0 && (!segs || !masters))) return fail(ctx, HDGX_ERR_ARG, "batch_mixed: arguments");
  if (nseg == 0) return HDGX_OK;
  // segments grouped by their master's order, each group in its given order
  std::vector<int> order(nseg);
  std::vector<int> pof(nseg);
  for (int s = 0; s < nseg; ++s) {
    if (!masters[s]) return fail(ctx, HDGX_ERR_ARG, "batch_mixed: null master");
    if (masters[s]->nd != masters[0]->nd) return fail(ctx, HDGX_ERR_ARG, "batch_mixed: segments must share nd");
    pof[s] = porder_of(masters[s]->nd, masters[s]->npv);
    if (pof[s] < 1) return fail(ctx, HDGX_ERR_UNSUPPORTED, "batch_mixed: order");
    order[s] = s;
  }
  std::stable_sort(order.begin(), order.end(), [&](int x, int y) { return pof[x] > pof[y]; });
  std::vector<hdgx_segment> group;
  for (int k = 0; k < nseg;) {
    const int p = pof[order[k]];
    const hdgx_master* m = masters[order[k]];
    group.clear();
    for (; k < nseg && pof[order[k]] == p; ++k) {
      if (masters[order[k]]->npv != m->npv || masters[order[k]]->npf != m->npf)
        return fail(ctx, HDGX_ERR_ARG, "batch_mixed: masters of one order differ");
      group.push_back(segs[order[k]]);
    }
    const int rc = hdgx_elemental_batch(ctx, m, (int)group.size(), group.data());
    if (rc) return rc;
  }
  return HDGX_OK;
}

}  // extern "C"

// ------------------------------------------------------------ launch plans
struct hdgx_plan {
  int mode = 0, grid = 0, phys = 0, nd = 0;
  int64_t total = 0;
  int64_t n[HDGX_MAXP + 1] = {};            // elements per order
  void* dev = nullptr;                      // tables
  void* ws = nullptr;                       // element matrices of all orders
  SegTable order_tab[HDGX_MAXP + 1] = {};   // per-order launches
  DevMaster tab[HDGX_MAXP + 1] = {};
  const int32_t* perm[HDGX_MAXP + 1] = {};
  MixedTable* mixed = nullptr;              // device copy (fused launch)
  hipStream_t side[HDGX_MAXP + 1] = {};
  hipEvent_t fork = nullptr, join[HDGX_MAXP + 1] = {};
  hipStream_t cap = nullptr;
  hipGraphExec_t exec = nullptr;
};

namespace {

// frees everything a plan holds (a partly built one too)
void plan_free(hdgx_plan* plan) {
  if (!plan) return;
  if (plan->exec) (void)hipGraphExecDestroy(plan->exec);
  if (plan->cap) (void)hipStreamDestroy(plan->cap);
  for (int p = 0; p <= HDGX_MAXP; ++p) {
    if (plan->side[p]) (void)hipStreamDestroy(plan->side[p]);
    if (plan->join[p]) (void)hipEventDestroy(plan->join[p]);
  }
  if (plan->fork) (void)hipEventDestroy(plan->fork);
  if (plan->ws) (void)hipFree(plan->ws);
  if (plan->dev) (void)hipFree(plan->dev);
  delete plan;
}

struct PlanDeleter {
  void operator()(hdgx_plan* p) const { plan_free(p); }
};

int plan_record(hdgx_ctx* ctx, hdgx_plan* P, hipStream_t s) {
  if (P->mode == HDGX_PLAN_FUSED) {
    return with_family(P->phys, P->nd, [&](auto ph, auto nd) {
      HIP_TRY(ctx, (mixed_launch<decltype(ph)::value, decltype(nd)::value>(s, P->mixed, P->total, P->grid)));
      return (int)HDGX_OK;
    });
  }
  const bool conc = P->mode == HDGX_PLAN_CONCURRENT;
  if (conc) HIP_TRY(ctx, hipEventRecord(P->fork, s));
  for (int p = HDGX_MAXP - 1; p >= 1; --p) {
    if (P->n[p] == 0) continue;
    hipStream_t sp = conc ? P->side[p] : s;
    if (conc) HIP_TRY(ctx, hipStreamWaitEvent(sp, P->fork, 0));
    const int rc = with_cfg(P->phys, P->nd, p, [&](auto cfg) {
      HIP_TRY(ctx, seg_launch<decltype(cfg)>(sp, P->order_tab[p], P->tab[p], P->perm[p]));
      return (int)HDGX_OK;
    });
    if (rc) return rc < 0 ? fail(ctx, HDGX_ERR_UNSUPPORTED, "plan: no kernel for this order") : rc;
    if (conc) {
      HIP_TRY(ctx, hipEventRecord(P->join[p], sp));
      HIP_TRY(ctx, hipStreamWaitEvent(s, P->join[p], 0));
    }
  }
  return HDGX_OK;
}

}  // namespace

extern "C" {

int hdgx_plan_create(hdgx_ctx* ctx, int nseg, const hdgx_master* const* masters, const hdgx_segment* segs, int mode,
                     int grid, hdgx_plan** out) {
  if (!ctx) return HDGX_ERR_ARG;
  if (!out || nseg < 1 || !masters || !segs) return fail(ctx, HDGX_ERR_ARG, "plan: arguments");
  if (mode != HDGX_PLAN_PER_ORDER && mode != HDGX_PLAN_FUSED && mode != HDGX_PLAN_CONCURRENT)
    return fail(ctx, HDGX_ERR_ARG, "plan: mode");
  HIP_TRY(ctx, hipSetDevice(ctx->device));
  const hdgx_app* a0 = segs[0].app;
  if (!a0 || !masters[0]) return fail(ctx, HDGX_ERR_ARG, "plan: null app or master");
  // every error return below frees what the plan holds by then
  std::unique_ptr<hdgx_plan, PlanDeleter> P(new hdgx_plan());
  P->mode = mode;
  P->phys = a0->physics;
  P->nd = masters[0]->nd;
  if (mode == HDGX_PLAN_FUSED && with_family(P->phys, P->nd, [](auto, auto) { return 0; }) < 0)
    return fail(ctx, HDGX_ERR_UNSUPPORTED, "plan: no fused kernel for this physics and dimension");
  // order the segments by decreasing p (stable)
  std::vector<int> order(nseg), pof(nseg);
  std::vector<DevApp> apps(nseg);
  int rc;
  for (int s = 0; s < nseg; ++s) {
    const hdgx_segment& g = segs[s];
    if (!masters[s] || !g.app) return fail(ctx, HDGX_ERR_ARG, "plan: null app or master");
    const hdgx_app* a = g.app;
    if (a->physics != P->phys || masters[s]->nd != P->nd || a->ncu != a0->ncu || a->nc != a0->nc ||
        a->localsolve != a0->localsolve)
      return fail(ctx, HDGX_ERR_ARG, "plan: segments must share physics, nd, ncu, nc and localsolve");
    const int p = porder_of(P->nd, masters[s]->npv);
    if (p < 1 || p >= HDGX_MAXP) return fail(ctx, HDGX_ERR_UNSUPPORTED, "plan: order");
    pof[s] = p;
    order[s] = s;
    if ((rc = make_app(ctx, masters[s], a, &apps[s]))) return rc;
    if (a->getdqdg != 1) return fail(ctx, HDGX_ERR_UNSUPPORTED, "plan: getdqdg = 1 (nco = nc) only");
    if (g.ne < 0) return fail(ctx, HDGX_ERR_ARG, "plan: ne < 0");
    if (g.ne > 0 && (!g.dgnodes || !g.bf || !g.udg || !g.uh || !g.ae || !g.fe || !g.dudg || !g.dudg_duh))
      return fail(ctx, HDGX_ERR_ARG, "plan: null array");
    if (g.ne > 0 && a->tdep && !g.sh) return fail(ctx, HDGX_ERR_ARG, "plan: tdep needs sh");
    P->n[p] += g.ne;
  }
  std::stable_sort(order.begin(), order.end(), [&](int x, int y) { return pof[x] > pof[y]; });
  int64_t first[HDGX_MAXP + 1] = {};
  {
    int64_t acc = 0;
    for (int p = HDGX_MAXP - 1; p >= 1; --p) { first[p] = acc; acc += P->n[p]; }
    P->total = acc;
    first[HDGX_MAXP] = acc;
    first[0] = acc;
  }
  // batch-global table (fused) and per-order tables (order-relative e0)
  std::vector<ElemSeg> glob(nseg), per(nseg);
  int64_t cur[HDGX_MAXP + 1] = {};
  int start[HDGX_MAXP + 1] = {}, count[HDGX_MAXP + 1] = {};
  for (int k = 0; k < nseg; ++k) {
    const int s = order[k], p = pof[s];
    const hdgx_segment& g = segs[s];
    if (count[p] == 0) start[p] = k;
    ++count[p];
    glob[k] = ElemSeg{first[p] + cur[p], g.ne, g.dgnodes, g.bf, g.udg, g.uh, g.sh, g.ae, g.fe, g.dudg, g.dudg_duh,
                      g.status, s, p};
    per[k] = glob[k];
    per[k].e0 = cur[p];
    cur[p] += g.ne;
  }
  // masters
  for (int p = 1; p < HDGX_MAXP; ++p) {
    if (P->n[p] == 0) continue;
    int s0 = -1;
    for (int s = 0; s < nseg; ++s) if (pof[s] == p) { s0 = s; break; }
    MasterBuf* mb = nullptr;
    if ((rc = get_master(ctx, masters[s0], &mb))) return rc;
    for (int s = 0; s < nseg; ++s)
      if (pof[s] == p && masters[s]->npv != mb->npv) return fail(ctx, HDGX_ERR_ARG, "plan: masters of one order differ");
    rc = with_cfg(P->phys, P->nd, p, [&](auto cfg) {
      using Cc = decltype(cfg);
      if (mb->npv != Cc::NPV || mb->npf != Cc::NPF || mb->ngv != Cc::NGV || mb->ngf != Cc::NGF)
        return fail(ctx, HDGX_ERR_UNSUPPORTED, "plan: master quadrature does not match the compiled configuration");
      return (int)HDGX_OK;
    });
    if (rc < 0) return fail(ctx, HDGX_ERR_UNSUPPORTED, "plan: no kernel for this order");
    if (rc) return rc;
    P->tab[p] = mb->tab;
    P->perm[p] = mb->tab.perm;
  }
  // device tables: glob | per | apps | mixed
  const size_t b1 = sizeof(ElemSeg) * nseg, b2 = sizeof(DevApp) * nseg, b3 = sizeof(MixedTable);
  HIP_TRY(ctx, hipMalloc(&P->dev, 2 * b1 + b2 + b3 + 64));
  char* d = static_cast<char*>(P->dev);
  ElemSeg* dglob = reinterpret_cast<ElemSeg*>(d);
  ElemSeg* dper = reinterpret_cast<ElemSeg*>(d + b1);
  DevApp* dapps = reinterpret_cast<DevApp*>(d + 2 * b1);
  P->mixed = reinterpret_cast<MixedTable*>(d + 2 * b1 + ((b2 + 15) / 16) * 16);
  // workspace, per order
  const hdgx_app* a = a0;
  const bool uq = a->localsolve != 0;
  size_t words = 0, off[HDGX_MAXP + 1] = {};
  for (int p = HDGX_MAXP - 1; p >= 1; --p) {
    if (P->n[p] == 0) continue;
    int s0 = 0;
    for (int s = 0; s < nseg; ++s) if (pof[s] == p) { s0 = s; break; }
    (void)s0;
    off[p] = words;
    const size_t w = ws_words_of(P->phys, P->nd, p);
    if (w == 0) return fail(ctx, HDGX_ERR_UNSUPPORTED, "plan: no kernel for this order");
    words += w * (size_t)P->n[p];
  }
  HIP_TRY(ctx, hipMalloc(&P->ws, std::max<size_t>(words, 1) * 8));
  MixedTable mt{};
  mt.segs = dglob;
  mt.apps = dapps;
  mt.nseg = nseg;
  mt.total = P->total;
  for (int p = 0; p <= HDGX_MAXP; ++p) {
    mt.first[p] = first[p];
    mt.end[p] = first[p] + (p < HDGX_MAXP ? P->n[p] : 0);
  }
  for (int p = HDGX_MAXP - 1; p >= 1; --p) {
    if (P->n[p] == 0) continue;
    int s0 = 0;
    for (int s = 0; s < nseg; ++s) if (pof[s] == p) { s0 = s; break; }
    (void)s0;
    const int64_t n = P->n[p];
    double* w = static_cast<double*>(P->ws) + off[p];
    SegTable& t = P->order_tab[p];
    t.segs = dper + start[p];
    t.apps = dapps;
    t.nseg = count[p];
    t.g0 = 0;
    t.total = n;
    // per-order batched launches take the tile layout where it exists, the
    // fused persistent kernels the reference layout
    const int rcw = with_cfg(P->phys, P->nd, p, [&](auto cfg) {
      using Cc = decltype(cfg);
      const WsPtrs W = ws_carve<Cc>(w, n, mode != HDGX_PLAN_FUSED && WsLayout<Cc>::TILE_OK);
      t.BD = W.BD; t.F = W.F; t.GK = W.GK; t.H = W.H; t.Ru = W.Ru; t.Rh = W.Rh;
      t.M = W.M; t.C = W.C; t.E = W.E; t.Rq = W.Rq;
      return 0;
    });
    if (rcw < 0) return fail(ctx, HDGX_ERR_UNSUPPORTED, "plan: no kernel for this order");
    mt.masters[p] = P->tab[p];
    mt.BD[p] = t.BD; mt.F[p] = t.F; mt.GK[p] = t.GK; mt.H[p] = t.H; mt.Ru[p] = t.Ru; mt.Rh[p] = t.Rh;
    mt.M[p] = t.M; mt.C[p] = t.C; mt.E[p] = t.E; mt.Rq[p] = t.Rq;
  }
  HIP_TRY(ctx, hipMemcpy(dglob, glob.data(), b1, hipMemcpyHostToDevice));
  HIP_TRY(ctx, hipMemcpy(dper, per.data(), b1, hipMemcpyHostToDevice));
  HIP_TRY(ctx, hipMemcpy(dapps, apps.data(), b2, hipMemcpyHostToDevice));
  HIP_TRY(ctx, hipMemcpy(P->mixed, &mt, b3, hipMemcpyHostToDevice));
  int ncu_dev = 256;
  (void)hipDeviceGetAttribute(&ncu_dev, hipDeviceAttributeMultiprocessorCount, ctx->device);
  P->grid = grid > 0 ? grid : 4 * ncu_dev;
  if (mode == HDGX_PLAN_CONCURRENT) {
    HIP_TRY(ctx, hipEventCreateWithFlags(&P->fork, hipEventDisableTiming));
    for (int p = 1; p < HDGX_MAXP; ++p) {
      if (P->n[p] == 0) continue;
      HIP_TRY(ctx, hipStreamCreateWithFlags(&P->side[p], hipStreamNonBlocking));
      HIP_TRY(ctx, hipEventCreateWithFlags(&P->join[p], hipEventDisableTiming));
    }
  }
  *out = P.release();
  return HDGX_OK;
}

int hdgx_plan_launch(hdgx_ctx* ctx, hdgx_plan* plan) {
  if (!ctx) return HDGX_ERR_ARG;
  if (!plan) return fail(ctx, HDGX_ERR_ARG, "plan: null");
  HIP_TRY(ctx, hipSetDevice(ctx->device));
  if (plan->exec) {
    HIP_TRY(ctx, hipGraphLaunch(plan->exec, ctx->stream));
    return HDGX_OK;
  }
  return plan_record(ctx, plan, ctx->stream);
}

int hdgx_plan_capture(hdgx_ctx* ctx, hdgx_plan* plan) {
  if (!ctx) return HDGX_ERR_ARG;
  if (!plan) return fail(ctx, HDGX_ERR_ARG, "plan: null");
  HIP_TRY(ctx, hipSetDevice(ctx->device));
  if (plan->exec) return HDGX_OK;
  if (!plan->cap) HIP_TRY(ctx, hipStreamCreateWithFlags(&plan->cap, hipStreamNonBlocking));
  HIP_TRY(ctx, hipStreamBeginCapture(plan->cap, hipStreamCaptureModeThreadLocal));
  const int rc = plan_record(ctx, plan, plan->cap);
  hipGraph_t graph = nullptr;
  const hipError_t e = hipStreamEndCapture(plan->cap, &graph);
  if (rc) {
    if (graph) (void)hipGraphDestroy(graph);
    return rc;
  }
  if (e != hipSuccess) return fail(ctx, HDGX_ERR_DEVICE, std::string("plan capture: ") + hipGetErrorString(e));
  const hipError_t ei = hipGraphInstantiate(&plan->exec, graph, nullptr, nullptr, 0);
  (void)hipGraphDestroy(graph);
  if (ei != hipSuccess) return fail(ctx, HDGX_ERR_DEVICE, std::string("plan instantiate: ") + hipGetErrorString(ei));
  return HDGX_OK;
}

int hdgx_plan_destroy(hdgx_ctx* ctx, hdgx_plan* plan) {
  if (!plan) return HDGX_OK;
  if (ctx) (void)hipSetDevice(ctx->device);
  (void)hipDeviceSynchronize();
  plan_free(plan);
  return HDGX_OK;
}

}  // extern "C"
