// Mesh connectivity on the device, for requests that bring a new mesh:
//   hdgx_mesh_faces  mesh/mkmesh/mkt2f.m -- faces numbered in order of first
//                    visit over (element, local face), the element->face map
//                    and the face rows [vertices as the first element sees
//                    them, first element, second element or -1];
//   hdgx_mesh_elcon  mesh/mkmesh/elconnectivities.m:147-212 -- face i owns
//                    trace nodes i*npf .. i*npf+npf-1 in its first element's
//                    order; the second element's face nodes are matched to
//                    them by coordinates;
//   hdgx_mesh_f2f    mesh/mkmesh/mkf2f.m:8-55 -- each face's neighbouring
//                    faces, the block pattern of the trace system.
// Faces are found without a sort: every (element, local face) inserts its
// sorted vertex tuple into an open-addressing hash table (linear probing,
// capacity a power of two at least twice the number of local faces), keeping
// the smallest visiting index and a count per key.  A local face whose index
// is its key's smallest is a face's first visit; an exclusive prefix sum of
// those flags in index order gives every face its number.
#include <hip/hip_runtime.h>

#include <cstdint>

#include "../../include/hdgx.h"
#include "hdgx_ctx.h"

using hdgx_detail::fail;

namespace {

constexpr unsigned long long EMPTY_KEY = ~0ull;
constexpr int SCAN_THREADS = 256;
constexpr int SCAN_ITEMS = 4;                       // items per thread of a scan block
constexpr int SCAN_BLOCK = SCAN_THREADS * SCAN_ITEMS;

// local face vertices (mesh/mkmesh/mkt2f.m's face tables, 0-based)
__constant__ int kTriFace[3][2] = {{1, 2}, {2, 0}, {0, 1}};
__constant__ int kTetFace[4][3] = {{1, 2, 3}, {0, 3, 2}, {0, 1, 3}, {0, 2, 1}};

__device__ __forceinline__ int face_vertex(int nd, int j, int a) { return nd == 2 ? kTriFace[j][a] : kTetFace[j][a]; }

// splitmix64 finaliser
__device__ __forceinline__ unsigned long long mix64(unsigned long long x) {
  x ^= x >> 30;
  x *= 0xbf58476d1ce4e5b9ull;
  x ^= x >> 27;
  x *= 0x94d049bb133111ebull;
  return x ^ (x >> 31);
}

unsigned grid_of(int64_t n, int threads) {
  const int64_t b = (n + threads - 1) / threads;
  return (unsigned)(b < 1 ? 1 : b);
}

// err bits: 1 vertex id out of range, 2 a face shared by more than two elements
__global__ void face_insert_kernel(int nd, int64_t ne, const int32_t* __restrict__ t, int32_t nvert,
                                   unsigned long long* keys, int32_t* first, int32_t* cnt, int32_t* slot,
                                   unsigned long long mask, int* err) {
  const int nfe = nd + 1, nvf = nd, nv = nd + 1;
  const int64_t q = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (q >= ne * nfe) return;
  const int64_t e = q / nfe;
  const int j = (int)(q % nfe);
  int v[3] = {0, 0, 0};
  for (int a = 0; a < nvf; ++a) v[a] = t[face_vertex(nd, j, a) + (int64_t)nv * e];
  for (int a = 0; a < nvf; ++a)
    if (v[a] < 0 || v[a] >= nvert) { atomicOr(err, 1); slot[q] = -1; return; }
  // sorted tuple -> key (3D: 21 bits a vertex, checked on the host)
  if (v[0] > v[1]) { const int s = v[0]; v[0] = v[1]; v[1] = s; }
  if (nvf == 3) {
    if (v[1] > v[2]) { const int s = v[1]; v[1] = v[2]; v[2] = s; }
    if (v[0] > v[1]) { const int s = v[0]; v[0] = v[1]; v[1] = s; }
  }
  const unsigned long long key =
      nvf == 2 ? ((unsigned long long)v[0] << 32) | (unsigned)v[1]
               : ((unsigned long long)v[0] << 42) | ((unsigned long long)v[1] << 21) | (unsigned long long)v[2];
  unsigned long long h = mix64(key) & mask;
  // the table holds at most half its capacity, so an empty slot is reached
  for (;;) {
    const unsigned long long prev = atomicCAS(&keys[h], EMPTY_KEY, key);
    if (prev == EMPTY_KEY || prev == key) break;
    h = (h + 1) & mask;
  }
  atomicMin(&first[h], (int32_t)q);
  if (atomicAdd(&cnt[h], 1) >= 2) atomicOr(err, 2);
  slot[q] = (int32_t)h;
}

__global__ void first_flag_kernel(int64_t nq, const int32_t* __restrict__ first, const int32_t* __restrict__ slot,
                                  int32_t* flag) {
  const int64_t q = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (q >= nq) return;
  const int32_t h = slot[q];
  flag[q] = (h >= 0 && first[h] == (int32_t)q) ? 1 : 0;
}

// exclusive scan of a block of SCAN_THREADS values in LDS
__device__ int32_t block_exclusive_scan(int32_t v, int32_t* s, int32_t* total) {
  const int tid = threadIdx.x;
  s[tid] = v;
  __syncthreads();
  for (int off = 1; off < SCAN_THREADS; off <<= 1) {
    const int32_t add = tid >= off ? s[tid - off] : 0;
    __syncthreads();
    s[tid] += add;
    __syncthreads();
  }
  const int32_t incl = s[tid];
  *total = s[SCAN_THREADS - 1];
  __syncthreads();
  return incl - v;
}

// per block of SCAN_BLOCK items: block-local exclusive prefix and block total
__global__ __launch_bounds__(SCAN_THREADS) void scan_blocks_kernel(int64_t n, const int32_t* __restrict__ in,
                                                                   int32_t* out, int32_t* bsum) {
  __shared__ int32_t s[SCAN_THREADS];
  const int64_t base = (int64_t)blockIdx.x * SCAN_BLOCK + (int64_t)threadIdx.x * SCAN_ITEMS;
  int32_t x[SCAN_ITEMS], sum = 0;
  for (int i = 0; i < SCAN_ITEMS; ++i) {
    x[i] = base + i < n ? in[base + i] : 0;
    sum += x[i];
  }
  int32_t total;
  int32_t run = block_exclusive_scan(sum, s, &total);
  for (int i = 0; i < SCAN_ITEMS; ++i) {
    if (base + i < n) out[base + i] = run;
    run += x[i];
  }
  if (threadIdx.x == 0) bsum[blockIdx.x] = total;
}

// one block: exclusive offsets of nb block totals, and the grand total
__global__ __launch_bounds__(SCAN_THREADS) void scan_sums_kernel(int64_t nb, int32_t* bsum, int32_t* grand) {
  __shared__ int32_t s[SCAN_THREADS];
  const int64_t per = (nb + SCAN_THREADS - 1) / SCAN_THREADS;
  const int64_t lo = (int64_t)threadIdx.x * per;
  int32_t sum = 0;
  for (int64_t i = lo; i < lo + per && i < nb; ++i) sum += bsum[i];
  int32_t total;
  int32_t run = block_exclusive_scan(sum, s, &total);
  for (int64_t i = lo; i < lo + per && i < nb; ++i) {
    const int32_t v = bsum[i];
    bsum[i] = run;
    run += v;
  }
  if (threadIdx.x == 0) *grand = total;
}

__global__ void add_offsets_kernel(int64_t n, int32_t* out, const int32_t* __restrict__ boff) {
  const int64_t q = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (q < n) out[q] += boff[q / SCAN_BLOCK];
}

// t2f and the face rows; f (nvf + 2, nf): vertices, first element, second or -1
__global__ void face_rows_kernel(int nd, int64_t ne, const int32_t* __restrict__ t, const int32_t* __restrict__ first,
                                 const int32_t* __restrict__ cnt, const int32_t* __restrict__ slot,
                                 const int32_t* __restrict__ prefix, int32_t* t2f, int32_t* f) {
  const int nfe = nd + 1, nvf = nd, nv = nd + 1, w = nd + 2;
  const int64_t q = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (q >= ne * nfe) return;
  const int32_t h = slot[q];
  if (h < 0) return;                               // vertex id out of range: the call fails
  const int32_t q1 = first[h];
  const int32_t id = prefix[q1];
  const int64_t e = q / nfe;
  const int j = (int)(q % nfe);
  t2f[q] = id;                                     // t2f (nfe, ne): q = j + nfe*e
  int32_t* row = f + (int64_t)w * id;
  if (q == q1) {
    for (int a = 0; a < nvf; ++a) row[a] = t[face_vertex(nd, j, a) + (int64_t)nv * e];
    row[nvf] = (int32_t)e;
    if (cnt[h] == 1) row[nvf + 1] = -1;
  } else {
    row[nvf + 1] = (int32_t)e;
  }
}

// one thread per (face node b, face i): elcon of both sides.  The element ids
// of f and the node ids of perm are the caller's: any outside [0, ne) or
// [0, npv) sets error bit 16 before it indexes anything.
__global__ void elcon_kernel(int ncmp, int npv, int ncd, int npf, int nfe, int64_t ne, int64_t nf, int nvf,
                             const double* __restrict__ dgnodes, const int32_t* __restrict__ perm,
                             const int32_t* __restrict__ f, const int32_t* __restrict__ t2f, double tol, int32_t* elcon,
                             int* err) {
  const int64_t idx = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (idx >= nf * npf) return;
  const int b = (int)(idx % npf);
  const int64_t i = idx / npf;
  const int32_t* row = f + (int64_t)(nvf + 2) * i;
  const int32_t e1 = row[nvf], e2 = row[nvf + 1];
  if (e1 < 0 || e1 >= ne || e2 >= ne) { atomicOr(err, 16); return; }
  for (int k = 0; k < npf * nfe; ++k)
    if (perm[k] < 0 || perm[k] >= npv) { atomicOr(err, 16); return; }
  int j1 = -1, j2 = -1;
  for (int j = 0; j < nfe; ++j) {
    if (t2f[j + (int64_t)nfe * e1] == i && j1 < 0) j1 = j;
    if (e2 >= 0 && t2f[j + (int64_t)nfe * e2] == i && j2 < 0) j2 = j;
  }
  if (j1 < 0 || (e2 >= 0 && j2 < 0)) { atomicOr(err, 4); return; }
  const int64_t base = (int64_t)npf * nfe;
  elcon[b + npf * j1 + base * e1] = (int32_t)(i * npf + b);
  if (e2 < 0) return;
  // node b of the second side: the first-side node at the least squared distance
  const double* X1 = dgnodes + (int64_t)npv * ncd * e1;
  const double* X2 = dgnodes + (int64_t)npv * ncd * e2;
  const int n2 = perm[b + npf * j2];
  double best = 0.0;
  int ab = -1;
  for (int a = 0; a < npf; ++a) {
    const int n1 = perm[a + npf * j1];
    double d = 0.0;
    for (int c = 0; c < ncmp; ++c) {
      const double dx = X2[n2 + npv * c] - X1[n1 + npv * c];
      d += dx * dx;
    }
    if (ab < 0 || d < best) { best = d; ab = a; }
  }
  if (best > tol) atomicOr(err, 8);
  elcon[b + npf * j2 + base * e2] = (int32_t)(i * npf + ab);
}

// one thread per face: fconn (4, nf) = [e1, j1, e2 or -1, j2 or 0] and the
// face-to-face table f2f (2 nfe - 1, nf): the face, the other faces of its
// first element, then of its second (ascending local index), -1 where none.
// Element ids of f outside [0, ne) set error bit 16 before any t2f read.
__global__ void f2f_kernel(int nd, int64_t ne, int64_t nf, const int32_t* __restrict__ f,
                           const int32_t* __restrict__ t2f, int32_t* fconn, int32_t* f2f, int* err) {
  const int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= nf) return;
  const int nvf = nd, nfe = nd + 1, nbf = 2 * nfe - 1;
  const int32_t* row = f + (int64_t)(nvf + 2) * i;
  const int32_t e1 = row[nvf];
  const int32_t e2 = row[nvf + 1] >= 0 ? row[nvf + 1] : -1;
  if (e1 < 0 || e1 >= ne || e2 >= ne) { atomicOr(err, 16); return; }
  int j1 = -1, j2 = -1;
  for (int j = 0; j < nfe; ++j) {
    if (j1 < 0 && t2f[j + (int64_t)nfe * e1] == i) j1 = j;
    if (e2 >= 0 && j2 < 0 && t2f[j + (int64_t)nfe * e2] == i) j2 = j;
  }
  if (j1 < 0 || (e2 >= 0 && j2 < 0)) { atomicOr(err, 4); return; }
  if (fconn) {
    fconn[4 * i] = e1;
    fconn[4 * i + 1] = j1;
    fconn[4 * i + 2] = e2;
    fconn[4 * i + 3] = e2 >= 0 ? j2 : 0;
  }
  int32_t* out = f2f + (int64_t)nbf * i;
  out[0] = (int32_t)i;
  for (int k = 1; k < nfe; ++k) {
    const int l1 = k - 1 < j1 ? k - 1 : k;
    out[k] = t2f[l1 + (int64_t)nfe * e1];
    const int l2 = k - 1 < j2 ? k - 1 : k;
    out[nfe - 1 + k] = e2 >= 0 ? t2f[l2 + (int64_t)nfe * e2] : -1;
  }
}

// CBSR pattern of the face system (mesh/mkmesh/block_crs.m via
// elcon2entcon.m on t2f'): row i is face i, then the faces sharing an element
// with it, ascending.  Two simplices share at most one face, so the f2f slots
// other than the face itself are distinct and the row is f2f's valid slots.
// One thread per face: the row length, then (after the scan) the row.
__global__ void bcrs_count_kernel(int nbf, int64_t nf, const int32_t* __restrict__ f2f, int32_t* cnt, int* err) {
  const int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= nf) return;
  const int32_t* row = f2f + (int64_t)nbf * i;
  int n = 1;
  if (row[0] != i) atomicOr(err, 1);
  for (int k = 1; k < nbf; ++k) {
    const int32_t v = row[k];
    if (v >= nf || v == i) atomicOr(err, 2);
    n += v >= 0;
  }
  cnt[i] = n;
}

__global__ void bcrs_fill_kernel(int nbf, int64_t nf, const int32_t* __restrict__ f2f,
                                 const int32_t* __restrict__ rowpts, int32_t* colind) {
  const int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= nf) return;
  const int32_t* row = f2f + (int64_t)nbf * i;
  int32_t v[8];                                    // nbf - 1 <= 6 (nd <= 3)
  int n = 0;
  for (int k = 1; k < nbf; ++k)
    if (row[k] >= 0) {
      // insertion into the ascending list
      int at = n++;
      while (at > 0 && v[at - 1] > row[k]) { v[at] = v[at - 1]; --at; }
      v[at] = row[k];
    }
  int32_t* out = colind + rowpts[i];
  out[0] = (int32_t)i;
  for (int k = 0; k < n; ++k) out[1 + k] = v[k];
}

// the context's mesh scratch, grown on demand and kept for later calls
int ensure_mws(hdgx_ctx* ctx, size_t bytes) {
  if (ctx->mws_bytes >= bytes) return HDGX_OK;
  if (ctx->mws) HIP_TRY(ctx, hipFree(ctx->mws));
  ctx->mws = nullptr;
  ctx->mws_bytes = 0;
  HIP_TRY(ctx, hipMalloc(&ctx->mws, bytes));
  ctx->mws_bytes = bytes;
  return HDGX_OK;
}

}  // namespace

int hdgx_mesh_faces(hdgx_ctx* ctx, int nd, int64_t ne, int32_t nvert, const int32_t* t, int32_t* t2f, int32_t* f,
                    int64_t* nf) {
  if (!ctx) return HDGX_ERR_ARG;
  if ((nd != 2 && nd != 3) || ne < 0 || nvert < 0 || !nf) return fail(ctx, HDGX_ERR_ARG, "mesh_faces: sizes");
  if (nd == 3 && nvert > (1 << 21)) return fail(ctx, HDGX_ERR_ARG, "mesh_faces: more than 2^21 vertices in 3D");
  const int nfe = nd + 1;
  const int64_t nq = ne * nfe;
  if (nq > INT32_MAX / 2) return fail(ctx, HDGX_ERR_ARG, "mesh_faces: too many elements");
  *nf = 0;
  if (ne == 0) return HDGX_OK;
  if (!t || !t2f || !f) return fail(ctx, HDGX_ERR_ARG, "mesh_faces: null array");
  HIP_TRY(ctx, hipSetDevice(ctx->device));
  unsigned long long cap = 1;
  while (cap < 2ull * (unsigned long long)nq) cap <<= 1;
  const int64_t nb = (nq + SCAN_BLOCK - 1) / SCAN_BLOCK;
  // err, grand | keys (cap) | first (cap) | cnt (cap) | slot (nq) | flag (nq) | prefix (nq) | bsum (nb)
  const size_t bytes = 16 + cap * 8 + cap * 4 * 2 + (size_t)nq * 4 * 3 + (size_t)nb * 4;
  int rc = ensure_mws(ctx, bytes);
  if (rc) return rc;
  char* p = static_cast<char*>(ctx->mws);
  auto* err = reinterpret_cast<int32_t*>(p);
  auto* grand = err + 1;
  auto* keys = reinterpret_cast<unsigned long long*>(p + 16);
  auto* first = reinterpret_cast<int32_t*>(p + 16 + cap * 8);
  auto* cnt = first + cap;
  auto* slot = cnt + cap;
  auto* flag = slot + nq;
  auto* prefix = flag + nq;
  auto* bsum = prefix + nq;
  hipStream_t s = ctx->stream;
  HIP_TRY(ctx, hipMemsetAsync(err, 0, 2 * sizeof(int32_t), s));
  HIP_TRY(ctx, hipMemsetAsync(keys, 0xff, cap * 8, s));
  HIP_TRY(ctx, hipMemsetAsync(first, 0x7f, cap * 4, s));   // 0x7f7f7f7f: above every index
  HIP_TRY(ctx, hipMemsetAsync(cnt, 0, cap * 4, s));
  hipLaunchKernelGGL(face_insert_kernel, dim3(grid_of(nq, 256)), dim3(256), 0, s, nd, ne, t, nvert, keys, first, cnt,
                     slot, (unsigned long long)(cap - 1), err);
  hipLaunchKernelGGL(first_flag_kernel, dim3(grid_of(nq, 256)), dim3(256), 0, s, nq, first, slot, flag);
  hipLaunchKernelGGL(scan_blocks_kernel, dim3((unsigned)nb), dim3(SCAN_THREADS), 0, s, nq, flag, prefix, bsum);
  hipLaunchKernelGGL(scan_sums_kernel, dim3(1), dim3(SCAN_THREADS), 0, s, nb, bsum, grand);
  hipLaunchKernelGGL(add_offsets_kernel, dim3(grid_of(nq, 256)), dim3(256), 0, s, nq, prefix, bsum);
  hipLaunchKernelGGL(face_rows_kernel, dim3(grid_of(nq, 256)), dim3(256), 0, s, nd, ne, t, first, cnt, slot, prefix,
                     t2f, f);
  HIP_TRY(ctx, hipGetLastError());
  // the error word and the face count in one read-back (a failed call leaves
  // t2f and f undefined)
  int32_t h2[2] = {0, 0};
  HIP_TRY(ctx, hipMemcpyAsync(h2, err, sizeof(h2), hipMemcpyDeviceToHost, s));
  HIP_TRY(ctx, hipStreamSynchronize(s));
  if (h2[0] & 1) return fail(ctx, HDGX_ERR_ARG, "mesh_faces: vertex id outside 0..nvert-1");
  if (h2[0] & 2) return fail(ctx, HDGX_ERR_ARG, "mesh_faces: a face shared by more than two elements");
  *nf = h2[1];
  return HDGX_OK;
}

int hdgx_mesh_elcon(hdgx_ctx* ctx, int nd, int npv, int ncd, int npf, int64_t ne, int64_t nf, const double* dgnodes,
                    const int32_t* perm, const int32_t* f, const int32_t* t2f, int32_t* elcon) {
  if (!ctx) return HDGX_ERR_ARG;
  if ((nd != 2 && nd != 3) || npv < 1 || ncd < nd || npf < 1 || ne < 0 || nf < 0)
    return fail(ctx, HDGX_ERR_ARG, "mesh_elcon: sizes");
  if (nf == 0) return HDGX_OK;
  if (!dgnodes || !perm || !f || !t2f || !elcon) return fail(ctx, HDGX_ERR_ARG, "mesh_elcon: null array");
  HIP_TRY(ctx, hipSetDevice(ctx->device));
  int rc = ensure_mws(ctx, 16);
  if (rc) return rc;
  int* err = static_cast<int*>(ctx->mws);
  hipStream_t s = ctx->stream;
  HIP_TRY(ctx, hipMemsetAsync(err, 0, sizeof(int), s));
  const int ncmp = ncd < 3 ? ncd : 3;              // coordinates compared
  hipLaunchKernelGGL(elcon_kernel, dim3(grid_of(nf * npf, 256)), dim3(256), 0, s, ncmp, npv, ncd, npf, nd + 1, ne, nf,
                     nd, dgnodes, perm, f, t2f, 1e-16, elcon, err);
  HIP_TRY(ctx, hipGetLastError());
  int herr = 0;
  HIP_TRY(ctx, hipMemcpyAsync(&herr, err, sizeof(int), hipMemcpyDeviceToHost, s));
  HIP_TRY(ctx, hipStreamSynchronize(s));
  if (herr & 16) return fail(ctx, HDGX_ERR_ARG, "mesh_elcon: an element id of f outside 0..ne-1 or a perm entry outside 0..npv-1");
  if (herr & 4) return fail(ctx, HDGX_ERR_ARG, "mesh_elcon: a face missing from its element's t2f");
  if (herr & 8) return fail(ctx, HDGX_ERR_ARG, "mesh_elcon: trace nodes do not match across a face");
  return HDGX_OK;
}

int hdgx_mesh_f2f(hdgx_ctx* ctx, int nd, int64_t ne, int64_t nf, const int32_t* f, const int32_t* t2f, int32_t* fconn,
                  int32_t* f2f) {
  if (!ctx) return HDGX_ERR_ARG;
  if ((nd != 2 && nd != 3) || ne < 0 || nf < 0) return fail(ctx, HDGX_ERR_ARG, "mesh_f2f: sizes");
  if (nf == 0) return HDGX_OK;
  if (!f || !t2f || !f2f) return fail(ctx, HDGX_ERR_ARG, "mesh_f2f: null array");
  HIP_TRY(ctx, hipSetDevice(ctx->device));
  int rc = ensure_mws(ctx, 16);
  if (rc) return rc;
  int* err = static_cast<int*>(ctx->mws);
  hipStream_t s = ctx->stream;
  HIP_TRY(ctx, hipMemsetAsync(err, 0, sizeof(int), s));
  hipLaunchKernelGGL(f2f_kernel, dim3(grid_of(nf, 256)), dim3(256), 0, s, nd, ne, nf, f, t2f, fconn, f2f, err);
  HIP_TRY(ctx, hipGetLastError());
  int herr = 0;
  HIP_TRY(ctx, hipMemcpyAsync(&herr, err, sizeof(int), hipMemcpyDeviceToHost, s));
  HIP_TRY(ctx, hipStreamSynchronize(s));
  if (herr & 16) return fail(ctx, HDGX_ERR_ARG, "mesh_f2f: an element id of f outside 0..ne-1");
  if (herr) return fail(ctx, HDGX_ERR_ARG, "mesh_f2f: a face missing from its element's t2f");
  return HDGX_OK;
}

int hdgx_mesh_block_crs(hdgx_ctx* ctx, int nd, int64_t nf, const int32_t* f2f, int32_t* rowpts, int32_t* colind,
                        int64_t* nblks) {
  if (!ctx) return HDGX_ERR_ARG;
  if ((nd != 2 && nd != 3) || nf < 0 || !nblks) return fail(ctx, HDGX_ERR_ARG, "mesh_block_crs: sizes");
  if (nf > INT32_MAX / (2 * nd + 1)) return fail(ctx, HDGX_ERR_ARG, "mesh_block_crs: too many faces");
  *nblks = 0;
  if (!rowpts) return fail(ctx, HDGX_ERR_ARG, "mesh_block_crs: null array");
  HIP_TRY(ctx, hipSetDevice(ctx->device));
  hipStream_t s = ctx->stream;
  if (nf == 0) {
    HIP_TRY(ctx, hipMemsetAsync(rowpts, 0, sizeof(int32_t), s));
    HIP_TRY(ctx, hipStreamSynchronize(s));
    return HDGX_OK;
  }
  if (!f2f || !colind) return fail(ctx, HDGX_ERR_ARG, "mesh_block_crs: null array");
  const int nbf = 2 * nd + 1;
  const int64_t nb = (nf + SCAN_BLOCK - 1) / SCAN_BLOCK;
  // err | cnt (nf) | bsum (nb)
  int rc = ensure_mws(ctx, 16 + (size_t)nf * 4 + (size_t)nb * 4);
  if (rc) return rc;
  char* p = static_cast<char*>(ctx->mws);
  auto* err = reinterpret_cast<int32_t*>(p);
  auto* cnt = reinterpret_cast<int32_t*>(p + 16);
  auto* bsum = cnt + nf;
  HIP_TRY(ctx, hipMemsetAsync(err, 0, sizeof(int32_t), s));
  hipLaunchKernelGGL(bcrs_count_kernel, dim3(grid_of(nf, 256)), dim3(256), 0, s, nbf, nf, f2f, cnt, err);
  hipLaunchKernelGGL(scan_blocks_kernel, dim3((unsigned)nb), dim3(SCAN_THREADS), 0, s, nf, cnt, rowpts, bsum);
  hipLaunchKernelGGL(scan_sums_kernel, dim3(1), dim3(SCAN_THREADS), 0, s, nb, bsum, rowpts + nf);
  hipLaunchKernelGGL(add_offsets_kernel, dim3(grid_of(nf, 256)), dim3(256), 0, s, nf, rowpts, bsum);
  HIP_TRY(ctx, hipGetLastError());
  // the row lengths are checked before colind is written
  int32_t h2[2] = {0, 0};
  HIP_TRY(ctx, hipMemcpyAsync(&h2[0], err, sizeof(int32_t), hipMemcpyDeviceToHost, s));
  HIP_TRY(ctx, hipMemcpyAsync(&h2[1], rowpts + nf, sizeof(int32_t), hipMemcpyDeviceToHost, s));
  HIP_TRY(ctx, hipStreamSynchronize(s));
  if (h2[0] & 1) return fail(ctx, HDGX_ERR_ARG, "mesh_block_crs: f2f(0, i) is not face i");
  if (h2[0] & 2) return fail(ctx, HDGX_ERR_ARG, "mesh_block_crs: an f2f entry at or past nf, or a face its own neighbour");
  hipLaunchKernelGGL(bcrs_fill_kernel, dim3(grid_of(nf, 256)), dim3(256), 0, s, nbf, nf, f2f, rowpts, colind);
  HIP_TRY(ctx, hipGetLastError());
  HIP_TRY(ctx, hipStreamSynchronize(s));
  *nblks = h2[1];
  return HDGX_OK;
}
