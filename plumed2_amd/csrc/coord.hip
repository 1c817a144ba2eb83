// COORDINATION on gfx950: GPU neighbour structure, pair kernel, reductions,
// and the C ABI of the coordination context.
//
// What each piece replaces in the reference:
//   rebuild (k_bin .. k_list)   NeighborList::update cells branch + LinkCells
//                               (NeighborList.cpp:168-236, LinkCells.cpp:99-239);
//                               the O(N^2) pair array is never materialised,
//                               the list is one of 8x8-atom cluster pairs.
//   k_pack                      ActionAtomistic::retrieveAtoms gather + the
//                               image bookkeeping of Pbc::distance
//   k_pairs                     CoordinationBase::calculate pair loop
//                               (CoordinationBase.cpp:178-207) with
//                               fixedRational<N>::calculateSqr
//                               (SwitchingFunction.cpp:203-215) inlined
//   k_finalize                  the OpenMP critical merge + comm.Sum
//                               (CoordinationBase.cpp:209-224)
//   k_allpairs / k_pairmode     the same loop in FP64 for inputs the cluster
//                               path does not cover (NOPBC, unbounded D_MAX,
//                               cells narrower than 2*cutoff, PAIR)
//
// Layout in HBM (per group of atoms, A and optionally B):
//   slots      8 consecutive slots = one cluster; clusters are runs of atoms
//              in columns of the fractional (a,b) plane sorted along c;
//              float4 {x,y,z local to the cluster origin, requested index}
//   cl_origin  double4 per cluster, FP64 origin fixed at the rebuild
//   cl_bb      2 x float4 per cluster (centre, half extents), every step
//   list       CSR over i clusters of Entry{code, delta}: code = j cluster |
//              image code << 26 | diagonal << 31; delta = o_j + T - o_i (FP32)
#include <hip/hip_runtime.h>

#include <algorithm>
#include <cfloat>
#include <cmath>
#include <cstring>
#include <string>
#include <vector>

#include "p2a_internal.h"
#include "switch_eval.h"

namespace p2a {

#define P2A_HIP(x)                                                                           \
  do {                                                                                       \
    hipError_t e_ = (x);                                                                     \
    if (e_ != hipSuccess) return fail(P2A_ERR_HIP, std::string(#x) + ": " + hipGetErrorString(e_)); \
  } while (0)

constexpr int CL = 8;           // atoms per cluster
constexpr int WAVE = 64;
constexpr int PAIR_BLOCK = 256; // 4 waves per workgroup

template <class T>
struct DBuf {
  T* p = nullptr;
  size_t n = 0;
  int reserve(size_t m) {
    if (m <= n && p) return P2A_OK;
    if (p) (void)hipFree(p);
    p = nullptr;
    n = 0;
    size_t want = std::max<size_t>(m, 1);
    if (hipMalloc(&p, want * sizeof(T)) != hipSuccess) {
      p = nullptr;
      return fail(P2A_ERR_NOMEM, "hipMalloc of " + std::to_string(want * sizeof(T)) + " bytes failed");
    }
    n = want;
    return P2A_OK;
  }
  ~DBuf() {
    if (p) (void)hipFree(p);
  }
};

struct Entry {
  int code;
  float dx, dy, dz;
};

struct BoxArg {
  double a[9];    // lattice vectors as rows
  double inv[9];  // inverse of a
};

struct SwF {  // FP32 constants of fixedRational<N>
  float dmax2, invr0_2, stretch, shift, cder;  // cder = -N*invr0_2*stretch
};

// ------------------------------------------------------------------------
// small device helpers
// ------------------------------------------------------------------------
template <int CTRL>
__device__ __forceinline__ float dpp_add(float v) {
  return v + __int_as_float(__builtin_amdgcn_mov_dpp(__float_as_int(v), CTRL, 0xF, 0xF, false));
}

// sum over each aligned group of 8 lanes (every lane of the group gets it)
__device__ __forceinline__ float sum8(float v) {
  v = dpp_add<0x141>(v);  // row_half_mirror: i <-> 7-i
  v = dpp_add<0x4E>(v);   // quad_perm [2,3,0,1]
  v = dpp_add<0xB1>(v);   // quad_perm [1,0,3,2]
  return v;
}

__device__ __forceinline__ double wave_sum(double v) {
  for (int off = 32; off > 0; off >>= 1) v += __shfl_xor(v, off, 64);
  return v;
}

__device__ __forceinline__ int floordiv(int a, int n) { return (a >= 0) ? a / n : -((-a + n - 1) / n); }

// ------------------------------------------------------------------------
// rebuild kernels
// ------------------------------------------------------------------------

// fractional coordinates, image, column of every atom of a group
__global__ void k_bin(int n, const double* __restrict__ pos, int off, BoxArg box, int ncx, int ncy,
                      int4* __restrict__ info, double* __restrict__ key, int* __restrict__ col_count) {
  int i = blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= n) return;
  const double* x = pos + 3 * (size_t)(off + i);
  double s[3];
  int m[3];
  for (int j = 0; j < 3; j++) {
    s[j] = x[0] * box.inv[0 * 3 + j] + x[1] * box.inv[1 * 3 + j] + x[2] * box.inv[2 * 3 + j];
    double fl = floor(s[j]);
    m[j] = (int)fl;
    s[j] -= fl;
  }
  int cx = min((int)(s[0] * ncx), ncx - 1);
  int cy = min((int)(s[1] * ncy), ncy - 1);
  cx = max(cx, 0);
  cy = max(cy, 0);
  int col = cx + ncx * cy;
  info[i] = make_int4(m[0], m[1], m[2], col);
  key[i] = s[2];
  atomicAdd(&col_count[col], 1);
}

// exclusive scans of column counts (atoms) and of ceil(count/8) (clusters),
// one workgroup; also the column of each cluster is derived later
__global__ void k_scan_cols(int ncol, const int* __restrict__ cnt, int* __restrict__ col_start,
                            int* __restrict__ col_cl, int* __restrict__ totals) {
  __shared__ int sa[1024], sc[1024];
  const int t = threadIdx.x, T = blockDim.x;
  int chunk = (ncol + T - 1) / T;
  int b = t * chunk, e = min(ncol, b + chunk);
  int la = 0, lc = 0;
  for (int i = b; i < e; i++) {
    la += cnt[i];
    lc += (cnt[i] + CL - 1) / CL;
  }
  sa[t] = la;
  sc[t] = lc;
  __syncthreads();
  for (int o = 1; o < T; o <<= 1) {
    int va = (t >= o) ? sa[t - o] : 0, vc = (t >= o) ? sc[t - o] : 0;
    __syncthreads();
    sa[t] += va;
    sc[t] += vc;
    __syncthreads();
  }
  int ra = sa[t] - la, rc = sc[t] - lc;
  for (int i = b; i < e; i++) {
    col_start[i] = ra;
    col_cl[i] = rc;
    ra += cnt[i];
    rc += (cnt[i] + CL - 1) / CL;
  }
  if (t == T - 1) {
    col_start[ncol] = sa[t];
    col_cl[ncol] = sc[t];
    totals[0] = sa[t];
    totals[1] = sc[t];
  }
}

// generic exclusive scan of n ints (one workgroup); totals[slot] = sum
__global__ void k_scan(int n, const int* __restrict__ in, int* __restrict__ out, int* __restrict__ totals, int slot,
                       int chunk64) {
  __shared__ int s[1024];
  const int t = threadIdx.x, T = blockDim.x;
  int chunk = (n + T - 1) / T;
  int b = t * chunk, e = min(n, b + chunk);
  int l = 0;
  for (int i = b; i < e; i++) l += chunk64 ? (in[i] + 63) / 64 : in[i];
  s[t] = l;
  __syncthreads();
  for (int o = 1; o < T; o <<= 1) {
    int v = (t >= o) ? s[t - o] : 0;
    __syncthreads();
    s[t] += v;
    __syncthreads();
  }
  int r = s[t] - l;
  for (int i = b; i < e; i++) {
    out[i] = r;
    r += chunk64 ? (in[i] + 63) / 64 : in[i];
  }
  if (t == T - 1) {
    out[n] = s[t];
    totals[slot] = s[t];
  }
}

__global__ void k_fill_int(int* p, int n, int v) {
  int i = blockIdx.x * blockDim.x + threadIdx.x;
  if (i < n) p[i] = v;
}

__global__ void k_fill_double(double* p, size_t n) {
  size_t i = (size_t)blockIdx.x * blockDim.x + threadIdx.x;
  size_t stride = (size_t)gridDim.x * blockDim.x;
  for (; i < n; i += stride) p[i] = 0.0;
}

__global__ void k_scatter(int n, const int4* __restrict__ info, const int* __restrict__ col_start,
                          int* __restrict__ fill, int* __restrict__ tmp) {
  int i = blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= n) return;
  int col = info[i].w;
  int k = atomicAdd(&fill[col], 1);
  tmp[col_start[col] + k] = i;
}

// rank of each atom inside its column by (key, index): deterministic order.
// One workgroup per column; the column's keys are staged in LDS when they fit.
constexpr int kRankLds = 2048;
__global__ __launch_bounds__(256) void k_rank(int ncol, const double* __restrict__ key,
                                              const int* __restrict__ col_start, const int* __restrict__ col_cl,
                                              const int* __restrict__ tmp, int* __restrict__ slot_loc) {
  __shared__ double sk[kRankLds];
  __shared__ int si[kRankLds];
  const int col = blockIdx.x;
  if (col >= ncol) return;
  const int b = col_start[col], n = col_start[col + 1] - b;
  const int base = col_cl[col] * CL;
  if (n <= kRankLds) {
    for (int t = threadIdx.x; t < n; t += blockDim.x) {
      int o = tmp[b + t];
      si[t] = o;
      sk[t] = key[o];
    }
    __syncthreads();
    for (int t = threadIdx.x; t < n; t += blockDim.x) {
      const double ka = sk[t];
      const int ia = si[t];
      int rank = 0;
      for (int u = 0; u < n; u++) {
        const double kb = sk[u];
        rank += (kb < ka) || (kb == ka && si[u] < ia);
      }
      slot_loc[base + rank] = ia;
    }
  } else {
    for (int t = threadIdx.x; t < n; t += blockDim.x) {
      const int ia = tmp[b + t];
      const double ka = key[ia];
      int rank = 0;
      for (int u = 0; u < n; u++) {
        const int o = tmp[b + u];
        const double kb = key[o];
        rank += (kb < ka) || (kb == ka && o < ia);
      }
      slot_loc[base + rank] = ia;
    }
  }
}

__global__ void k_cl_col(int ncol, const int* __restrict__ col_cl, int* __restrict__ cl_col) {
  int c = blockIdx.x * blockDim.x + threadIdx.x;
  if (c >= ncol) return;
  for (int k = col_cl[c]; k < col_cl[c + 1]; k++) cl_col[k] = c;
}

// FP64 origin (bounding-box centre at the rebuild) and rebuild bounding box
__global__ void k_origin(int ncl, const int* __restrict__ slot_loc, const int4* __restrict__ info,
                         const double* __restrict__ pos, int off, BoxArg box, double4* __restrict__ origin,
                         float4* __restrict__ bb0) {
  int c = blockIdx.x * blockDim.x + threadIdx.x;
  if (c >= ncl) return;
  double lo[3] = {DBL_MAX, DBL_MAX, DBL_MAX}, hi[3] = {-DBL_MAX, -DBL_MAX, -DBL_MAX};
  for (int k = 0; k < CL; k++) {
    int a = slot_loc[c * CL + k];
    if (a < 0) continue;
    int4 m = info[a];
    const double* x = pos + 3 * (size_t)(off + a);
    for (int d = 0; d < 3; d++) {
      double w = x[d] - (m.x * box.a[0 * 3 + d] + m.y * box.a[1 * 3 + d] + m.z * box.a[2 * 3 + d]);
      lo[d] = fmin(lo[d], w);
      hi[d] = fmax(hi[d], w);
    }
  }
  double cx = 0.5 * (lo[0] + hi[0]), cy = 0.5 * (lo[1] + hi[1]), cz = 0.5 * (lo[2] + hi[2]);
  origin[c] = make_double4(cx, cy, cz, 0.0);
  bb0[2 * c] = make_float4((float)cx, (float)cy, (float)cz, 0.f);
  bb0[2 * c + 1] = make_float4((float)(0.5 * (hi[0] - lo[0])), (float)(0.5 * (hi[1] - lo[1])),
                               (float)(0.5 * (hi[2] - lo[2])), 0.f);
}

// Entry code bits: j cluster (bits 0-29), "count" (bit 30: this entry adds
// to the value and the virial; each unordered pair of atoms is counted by
// exactly one entry), "diagonal" (bit 31: i cluster paired with itself
// without image, pairs with j == i are skipped and counted when j > i).
constexpr int kCjMask = 0x3FFFFFFF;
constexpr int kCountBit = 0x40000000;

struct ListArgs {
  int nclI, ncxI, ncyI, ncxJ, ncyJ;
  int mode;        // 0 one group, 1 two groups (counting side), 2 two groups (other side)
  float rlist2;
  float rb[2];     // rlist * |reciprocal row| for a, b
  float A[9];      // lattice vectors (FP32) for the search
  double Ad[9];    // lattice vectors (FP64) for the offsets
};

// cluster-pair search: one wave per i cluster.  FILL=0 counts, FILL=1 writes.
template <int FILL>
__global__ __launch_bounds__(256) void k_list(ListArgs L, const int* __restrict__ cl_colI,
                                              const float4* __restrict__ bbI, const float4* __restrict__ bbJ,
                                              const int* __restrict__ col_clJ, const double4* __restrict__ orgI,
                                              const double4* __restrict__ orgJ, int* __restrict__ count,
                                              const int* __restrict__ start, Entry* __restrict__ out) {
  const int lane = threadIdx.x & 63;
  const int ci = blockIdx.x * (blockDim.x / WAVE) + (threadIdx.x / WAVE);
  if (ci >= L.nclI) return;
  const int colI = cl_colI[ci];
  const int cxi = colI % L.ncxI, cyi = colI / L.ncxI;
  const float4 ic = bbI[2 * ci], ih = bbI[2 * ci + 1];
  const int x0 = (int)floorf(((float)cxi / L.ncxI - L.rb[0]) * L.ncxJ) - 1;
  const int x1 = (int)floorf(((float)(cxi + 1) / L.ncxI + L.rb[0]) * L.ncxJ) + 1;
  const int y0 = (int)floorf(((float)cyi / L.ncyI - L.rb[1]) * L.ncyJ) - 1;
  const int y1 = (int)floorf(((float)(cyi + 1) / L.ncyI + L.rb[1]) * L.ncyJ) + 1;
  // images -1, 0, +1 along a and b cover every candidate when the cell is
  // wider than 2*rlist; each unwrapped column is a distinct (column, image)
  const int xa = max(x0, -L.ncxJ), xb = min(x1, 2 * L.ncxJ - 1);
  const int ya = max(y0, -L.ncyJ), yb = min(y1, 2 * L.ncyJ - 1);
  double4 oi = orgI[ci];
  int n = 0;
  int base = FILL ? start[ci] : 0;
  for (int cy = ya; cy <= yb; cy++) {
    int my = floordiv(cy, L.ncyJ), cyw = cy - my * L.ncyJ;
    for (int cx = xa; cx <= xb; cx++) {
      int mx = floordiv(cx, L.ncxJ), cxw = cx - mx * L.ncxJ;
      int col = cxw + L.ncxJ * cyw;
      int c0 = col_clJ[col], c1 = col_clJ[col + 1];
      int ncand = (c1 - c0) * 3;
      for (int t0 = 0; t0 < ncand; t0 += WAVE) {
        int t = t0 + lane;
        bool pass = false;
        int cj = 0, mz = 0;
        if (t < ncand) {
          cj = c0 + t / 3;
          mz = t % 3 - 1;
          float Tx = mx * L.A[0] + my * L.A[3] + mz * L.A[6];
          float Ty = mx * L.A[1] + my * L.A[4] + mz * L.A[7];
          float Tz = mx * L.A[2] + my * L.A[5] + mz * L.A[8];
          float4 jc = bbJ[2 * cj], jh = bbJ[2 * cj + 1];
          float gx = fmaxf(fabsf(jc.x + Tx - ic.x) - ih.x - jh.x, 0.f);
          float gy = fmaxf(fabsf(jc.y + Ty - ic.y) - ih.y - jh.y, 0.f);
          float gz = fmaxf(fabsf(jc.z + Tz - ic.z) - ih.z - jh.z, 0.f);
          pass = gx * gx + gy * gy + gz * gz < L.rlist2;
        }
        unsigned long long bal = __ballot(pass);
        if (FILL && pass) {
          int k = __popcll(bal & ((1ull << lane) - 1));
          double4 oj = orgJ[cj];
          double Tx = mx * L.Ad[0] + my * L.Ad[3] + mz * L.Ad[6];
          double Ty = mx * L.Ad[1] + my * L.Ad[4] + mz * L.Ad[7];
          double Tz = mx * L.Ad[2] + my * L.Ad[5] + mz * L.Ad[8];
          bool zero = (mx == 0 && my == 0 && mz == 0);
          bool diag = L.mode == 0 && cj == ci && zero;
          bool positive = (mx > 0) || (mx == 0 && (my > 0 || (my == 0 && mz > 0)));
          bool counted = L.mode == 1 || (L.mode == 0 && (cj > ci || (cj == ci && positive)));
          Entry en;
          en.code = cj | (counted ? kCountBit : 0) | (diag ? (int)0x80000000u : 0);
          en.dx = (float)(oj.x + Tx - oi.x);
          en.dy = (float)(oj.y + Ty - oi.y);
          en.dz = (float)(oj.z + Tz - oi.z);
          out[base + n + k] = en;
        }
        n += __popcll(bal);
      }
    }
  }
  if (!FILL && lane == 0) count[ci] = n;
}

__global__ void k_items(int ncl, const int* __restrict__ count, const int* __restrict__ item_start,
                        int2* __restrict__ items) {
  int c = blockIdx.x * blockDim.x + threadIdx.x;
  if (c >= ncl) return;
  int nch = (count[c] + 63) / 64;
  for (int k = 0; k < nch; k++) items[item_start[c] + k] = make_int2(c, k);
}

// ------------------------------------------------------------------------
// per-step kernels
// ------------------------------------------------------------------------

// local FP32 coordinates of every slot; padding slots far away and distinct
__global__ void k_pack(int nslot, const int* __restrict__ slot_loc, const int4* __restrict__ info,
                       const double* __restrict__ pos, int off, BoxArg box, const double4* __restrict__ origin,
                       float4* __restrict__ out) {
  int s = blockIdx.x * blockDim.x + threadIdx.x;
  if (s >= nslot) return;
  int a = slot_loc[s];
  if (a < 0) {
    out[s] = make_float4(1.0e5f + 1.0e3f * (float)(s & 1023), 1.0e5f + 1.0e3f * (float)((s >> 10) & 1023),
                         1.0e5f + 1.0e3f * (float)(s >> 20), __int_as_float(-1));
    return;
  }
  int4 m = info[a];
  double4 o = origin[s / CL];
  const double* x = pos + 3 * (size_t)(off + a);
  double w[3];
  for (int d = 0; d < 3; d++) w[d] = x[d] - (m.x * box.a[0 * 3 + d] + m.y * box.a[1 * 3 + d] + m.z * box.a[2 * 3 + d]);
  out[s] = make_float4((float)(w[0] - o.x), (float)(w[1] - o.y), (float)(w[2] - o.z), __int_as_float(off + a));
}

// current bounding box of each cluster in its local frame
__global__ void k_bbox(int ncl, const float4* __restrict__ p, float4* __restrict__ bb) {
  int c = blockIdx.x * blockDim.x + threadIdx.x;
  if (c >= ncl) return;
  float lo[3] = {FLT_MAX, FLT_MAX, FLT_MAX}, hi[3] = {-FLT_MAX, -FLT_MAX, -FLT_MAX};
  for (int k = 0; k < CL; k++) {
    float4 q = p[c * CL + k];
    if (__float_as_int(q.w) < 0) continue;
    lo[0] = fminf(lo[0], q.x);
    hi[0] = fmaxf(hi[0], q.x);
    lo[1] = fminf(lo[1], q.y);
    hi[1] = fmaxf(hi[1], q.y);
    lo[2] = fminf(lo[2], q.z);
    hi[2] = fmaxf(hi[2], q.z);
  }
  bb[2 * c] = make_float4(0.5f * (lo[0] + hi[0]), 0.5f * (lo[1] + hi[1]), 0.5f * (lo[2] + hi[2]), 0.f);
  bb[2 * c + 1] = make_float4(0.5f * (hi[0] - lo[0]), 0.5f * (hi[1] - lo[1]), 0.5f * (hi[2] - lo[2]), 0.f);
}

struct PbcArg {
  int type;
  double box[9], inv[9], red[9], ired[9];
  int nshift[8];
  double shifts[8][kMaxShift][3];
};

// Tools::pbc (Tools.h:547-570)
__device__ __forceinline__ double pbc1(double x) {
  x += 100.0;
  return (x >= 0) ? x - (double)(int)(x + 0.5) : x - (double)(int)(x - 0.5);
}

// Pbc::distance (Pbc.cpp:362-415)
__device__ void pbc_distance(const PbcArg& P, const double* a, const double* b, double* d) {
  d[0] = b[0] - a[0];
  d[1] = b[1] - a[1];
  d[2] = b[2] - a[2];
  if (P.type == 1) {
    for (int i = 0; i < 3; i++) d[i] = pbc1(d[i] * P.inv[4 * i]) * P.box[4 * i];
  } else if (P.type == 2) {
    double s[3];
    for (int j = 0; j < 3; j++) s[j] = d[0] * P.ired[j] + d[1] * P.ired[3 + j] + d[2] * P.ired[6 + j];
    for (int j = 0; j < 3; j++) s[j] = pbc1(s[j]);
    for (int j = 0; j < 3; j++) d[j] = s[0] * P.red[j] + s[1] * P.red[3 + j] + s[2] * P.red[6 + j];
    if (fabs(s[0]) + fabs(s[1]) + fabs(s[2]) > 0.5) {
      int oct = (s[0] > 0 ? 4 : 0) + (s[1] > 0 ? 2 : 0) + (s[2] > 0 ? 1 : 0);
      double best[3] = {d[0], d[1], d[2]};
      double lbest = best[0] * best[0] + best[1] * best[1] + best[2] * best[2];
      for (int k = 0; k < P.nshift[oct]; k++) {
        double t0 = d[0] + P.shifts[oct][k][0], t1 = d[1] + P.shifts[oct][k][1], t2 = d[2] + P.shifts[oct][k][2];
        double lt = t0 * t0 + t1 * t1 + t2 * t2;
        if (lt < lbest) {
          lbest = lt;
          best[0] = t0;
          best[1] = t1;
          best[2] = t2;
        }
      }
      d[0] = best[0];
      d[1] = best[1];
      d[2] = best[2];
    }
  }
}

// s and dfunc of fixedRational<2P> on r^2 in FP32
template <int P>
__device__ __forceinline__ void sw_fixed(const SwF& f, float r2, float& s, float& df) {
  float x = r2 * f.invr0_2;
  float xp = 1.0f;
#pragma unroll
  for (int k = 1; k < P; k++) xp *= x;  // x^(P-1)
  float res = __builtin_amdgcn_rcpf(fmaf(xp, x, 1.0f));
  s = fmaf(res, f.stretch, f.shift);
  df = f.cder * xp * res * res;
}

struct PairArgs {
  int nitems;
  int twogroup;
  int offI, offJ;  // group offsets in the requested-atom list
  float dmax2, prune2;
  SwF f;
  p2a_switch sw;   // FP64 parameters (generic switch and the cut band)
  const PbcArg* pbc;  // device copy of the prepared cell (FP64 cut decisions)
};

// FP64 re-evaluation of the squared distance of one pair, used only when the
// FP32 value lies within a narrow band around D_MAX^2: the inclusive cut of
// the reference is then decided on the distance Pbc::distance itself would
// produce from the raw positions (including the rounding of Tools::pbc).
// The pair is oriented as the reference lists it (GROUPA atom first; lower
// index first inside one group) because the rounding of Tools::pbc is not
// symmetric under d -> -d.
__device__ double exact_r2(const PbcArg* __restrict__ P, const double* __restrict__ pos, int reqi, int reqj) {
  double d[3];
  const int a = min(reqi, reqj), b = max(reqi, reqj);
  pbc_distance(*P, pos + 3 * (size_t)a, pos + 3 * (size_t)b, d);
  return d[0] * d[0] + d[1] * d[1] + d[2] * d[2];
}

// One wave per work item (i cluster, chunk of <=64 list entries).
//  1. lane-parallel pruning of the chunk against the current bounding boxes;
//  2. the passing entries are compacted and each lane stages the 8 atoms of
//     its j cluster into LDS with the image offset already added (8
//     independent loads per lane);
//  3. dense loop over the staged clusters: lane l evaluates the pair
//     (i = l&7 of cluster ci, j = l>>3 of the staged cluster).
template <int P>
__device__ __forceinline__ float sw_fixed_value(const SwF& f, float r2) {
  float x = r2 * f.invr0_2;
  float xp = 1.0f;
#pragma unroll
  for (int k = 1; k < P; k++) xp *= x;
  return fmaf(__builtin_amdgcn_rcpf(fmaf(xp, x, 1.0f)), f.stretch, f.shift);
}

// VO (value only): no derivatives and no virial; entries that do not add to
// the value are skipped, so the loop visits each unordered pair once.
template <int P, bool GENERIC, bool TWO, bool VO>
__global__ __launch_bounds__(PAIR_BLOCK) void k_pairs(PairArgs A, const int2* __restrict__ items,
                                                       const int* __restrict__ estart,
                                                       const Entry* __restrict__ ents,
                                                       const float4* __restrict__ posI,
                                                       const float4* __restrict__ posJ,
                                                       const float4* __restrict__ bbI, const float4* __restrict__ bbJ,
                                                       const int* __restrict__ absI, const int* __restrict__ absJ,
                                                       const double* __restrict__ pos64,
                                                       double* __restrict__ fpart, double* __restrict__ partial) {
  constexpr int NW = PAIR_BLOCK / WAVE;
  __shared__ float4 sj[NW][WAVE][CL];
  __shared__ int scode[NW][WAVE];
  __shared__ int sabs[TWO ? NW : 1][TWO ? WAVE : 1][CL];
  const int lane = threadIdx.x & 63;
  const int w = threadIdx.x >> 6;
  const int item = blockIdx.x * NW + w;
  if (item >= A.nitems) return;  // uniform per wave; no workgroup barrier below
  const int2 it = items[item];
  const int ci = it.x;
  const int e0 = estart[ci] + it.y * 64;
  const int e1 = min(estart[ci + 1], e0 + 64);
  const int ii = lane & 7, jj = lane >> 3;
  const float4 pi = posI[ci * CL + ii];
  const int reqi = __float_as_int(pi.w);
  const int ai = TWO ? absI[ci * CL + ii] : 0;

  Entry my = {0, 0.f, 0.f, 0.f};
  bool pass = false;
  {
    int e = e0 + lane;
    if (e < e1) {
      my = ents[e];
      int cj = my.code & kCjMask;
      float4 ic = bbI[2 * ci], ih = bbI[2 * ci + 1];
      float4 jc = bbJ[2 * cj], jh = bbJ[2 * cj + 1];
      float gx = fmaxf(fabsf(jc.x + my.dx - ic.x) - ih.x - jh.x, 0.f);
      float gy = fmaxf(fabsf(jc.y + my.dy - ic.y) - ih.y - jh.y, 0.f);
      float gz = fmaxf(fabsf(jc.z + my.dz - ic.z) - ih.z - jh.z, 0.f);
      pass = gx * gx + gy * gy + gz * gz <= A.prune2;
    }
  }
  const unsigned long long mask = __ballot(pass);
  const int npass = __popcll(mask);
  if (pass) {
    const int k = __popcll(mask & ((1ull << lane) - 1));
    const int cj = my.code & kCjMask;
    float4 q[CL];
#pragma unroll
    for (int t = 0; t < CL; t++) q[t] = posJ[cj * CL + t];
#pragma unroll
    for (int t = 0; t < CL; t++) sj[w][k][t] = make_float4(q[t].x + my.dx, q[t].y + my.dy, q[t].z + my.dz, q[t].w);
    scode[w][k] = my.code;
    if constexpr (TWO) {
#pragma unroll
      for (int t = 0; t < CL; t++) sabs[w][k][t] = absJ[cj * CL + t];
    }
  }
  // the staging above is read by other lanes of this wave only
  __builtin_amdgcn_fence(__ATOMIC_SEQ_CST, "wavefront");
  __builtin_amdgcn_wave_barrier();

  float cv = 0.f, fx = 0.f, fy = 0.f, fz = 0.f;
  float vxx = 0.f, vxy = 0.f, vxz = 0.f, vyy = 0.f, vyz = 0.f, vzz = 0.f;
  int npair = 0;
  const float band = 4.0e-6f * A.dmax2;

  for (int k = 0; k < npass; k++) {
    const int code = scode[w][k];
    if constexpr (VO) {
      if (!(code & kCountBit) && code >= 0) continue;  // uniform
    }
    const float4 pj = sj[w][k][jj];
    const float dx = pj.x - pi.x;
    const float dy = pj.y - pi.y;
    const float dz = pj.z - pi.z;
    const float r2 = fmaf(dx, dx, fmaf(dy, dy, dz * dz));
    bool cut = r2 <= A.dmax2;
    if (__builtin_expect(__ballot(fabsf(r2 - A.dmax2) < band) != 0ull, 0)) {
      const int reqj = __float_as_int(pj.w);
      if (fabsf(r2 - A.dmax2) < band && reqi >= 0 && reqj >= 0) {
        double r2d = exact_r2(A.pbc, pos64, reqi, reqj);
        cut = switch_inside(A.sw, r2d);
      }
    }
    const bool diag = code < 0;
    if (diag) cut = cut && (jj != ii);  // an atom is not paired with itself
    if constexpr (TWO) cut = cut && (sabs[w][k][jj] != ai);
    if constexpr (VO) {
      float s;
      if constexpr (GENERIC) {
        double dfd;
        s = (float)switch_calc_sqr(A.sw, (double)r2, dfd);
      } else {
        s = sw_fixed_value<P>(A.f, r2);
      }
      const bool c2 = diag ? (cut && (jj > ii)) : cut;
      npair += __popcll(__ballot(c2));
      cv += c2 ? s : 0.f;
      continue;
    }
    float s, df;
    if constexpr (GENERIC) {
      double dfd;
      double sd = switch_calc_sqr(A.sw, (double)r2, dfd);
      s = (float)sd;
      df = (float)dfd;
    } else {
      sw_fixed<P>(A.f, r2, s, df);
    }
    df = cut ? df : 0.f;
    const float gx = df * dx, gy = df * dy, gz = df * dz;
    // derivative of atom i over every partner (the full list holds both
    // orientations of each cluster pair)
    fx += gx;
    fy += gy;
    fz += gz;
    if (code & kCountBit) {
      // value and virial once per unordered pair
      s = cut ? s : 0.f;
      npair += __popcll(__ballot(cut));
      cv += s;
      vxx = fmaf(gx, dx, vxx);
      vxy = fmaf(gx, dy, vxy);
      vxz = fmaf(gx, dz, vxz);
      vyy = fmaf(gy, dy, vyy);
      vyz = fmaf(gy, dz, vyz);
      vzz = fmaf(gz, dz, vzz);
    } else if (diag) {
      const bool c2 = cut && (jj > ii);
      const float wgt = c2 ? 1.f : 0.f;
      npair += __popcll(__ballot(c2));
      cv = fmaf(wgt, s, cv);
      const float hx = wgt * gx, hy = wgt * gy, hz = wgt * gz;
      vxx = fmaf(hx, dx, vxx);
      vxy = fmaf(hx, dy, vxy);
      vxz = fmaf(hx, dz, vxz);
      vyy = fmaf(hy, dy, vyy);
      vyz = fmaf(hy, dz, vyz);
      vzz = fmaf(hz, dz, vzz);
    }
  }
  // derivative of atom i: -sum over the 8 lanes sharing i (lanes ii + 8k);
  // chunks of one cluster are combined by k_gather in a fixed order
  double dfx = fx, dfy = fy, dfz = fz;
  for (int o = 8; o < 64; o <<= 1) {
    dfx += __shfl_xor(dfx, o, 64);
    dfy += __shfl_xor(dfy, o, 64);
    dfz += __shfl_xor(dfz, o, 64);
  }
  if (!VO && jj == 0) {
    double* fp = fpart + (size_t)item * (CL * 3) + ii * 3;
    fp[0] = -dfx;
    fp[1] = -dfy;
    fp[2] = -dfz;
  }
  double r[7] = {cv, vxx, vxy, vxz, vyy, vyz, vzz};
  for (int q = 0; q < 7; q++) r[q] = wave_sum(r[q]);
  if (lane == 0) {
    double* out = partial + (size_t)item * 8;
    for (int q = 0; q < 7; q++) out[q] = r[q];
    out[7] = (double)npair;
  }
}

// derivatives of the atoms of every i cluster: sum of its chunks' partials
__global__ void k_gather(int ncl, const float4* __restrict__ p, const int* __restrict__ item_start,
                         const double* __restrict__ fpart, double* __restrict__ deriv) {
  const int t = blockIdx.x * blockDim.x + threadIdx.x;
  if (t >= ncl * CL * 3) return;
  const int c = t / (CL * 3), r = t % (CL * 3);
  const int req = __float_as_int(p[c * CL + r / 3].w);
  if (req < 0) return;
  double sum = 0.0;
  for (int it = item_start[c]; it < item_start[c + 1]; it++) sum += fpart[(size_t)it * (CL * 3) + r];
  deriv[3 * (size_t)req + r % 3] = sum;
}

// deterministic sum of the per-item partials -> value, virial, pair count.
// Stage 1: 128 blocks reduce contiguous ranges; stage 2: one block.
constexpr int kRedBlocks = 128;
__global__ __launch_bounds__(256) void k_reduce1(int nitems, const double* __restrict__ partial,
                                                 double* __restrict__ out) {
  __shared__ double s[8][256];
  const int t = threadIdx.x;
  const int per = (nitems + gridDim.x - 1) / gridDim.x;
  const int b = blockIdx.x * per, e = min(nitems, b + per);
  double acc[8] = {0, 0, 0, 0, 0, 0, 0, 0};
  const double4* p4 = reinterpret_cast<const double4*>(partial);
  for (int i = b + t; i < e; i += 256) {
    double4 x = p4[2 * (size_t)i], y = p4[2 * (size_t)i + 1];
    acc[0] += x.x; acc[1] += x.y; acc[2] += x.z; acc[3] += x.w;
    acc[4] += y.x; acc[5] += y.y; acc[6] += y.z; acc[7] += y.w;
  }
  for (int q = 0; q < 8; q++) s[q][t] = acc[q];
  __syncthreads();
  for (int o = 128; o > 0; o >>= 1) {
    if (t < o)
      for (int q = 0; q < 8; q++) s[q][t] += s[q][t + o];
    __syncthreads();
  }
  if (t < 8) out[(size_t)blockIdx.x * 8 + t] = s[t][0];
}

__global__ void k_finalize(int nitems, const double* __restrict__ partial, double* __restrict__ value,
                           double* __restrict__ virial, double* __restrict__ stats_out) {
  __shared__ double s[8][128];
  const int t = threadIdx.x;
  double acc[8] = {0, 0, 0, 0, 0, 0, 0, 0};
  for (int i = t; i < nitems; i += 128)
    for (int q = 0; q < 8; q++) acc[q] += partial[(size_t)i * 8 + q];
  for (int q = 0; q < 8; q++) s[q][t] = acc[q];
  __syncthreads();
  for (int o = 64; o > 0; o >>= 1) {
    if (t < o)
      for (int q = 0; q < 8; q++) s[q][t] += s[q][t + o];
    __syncthreads();
  }
  if (t == 0) {
    if (value) *value = s[0][0];
    if (virial) {
      // virial = -sum dd (x) d (CoordinationBase.cpp:200-206), symmetric
      double xx = -s[1][0], xy = -s[2][0], xz = -s[3][0], yy = -s[4][0], yz = -s[5][0], zz = -s[6][0];
      virial[0] = xx; virial[1] = xy; virial[2] = xz;
      virial[3] = xy; virial[4] = yy; virial[5] = yz;
      virial[6] = xz; virial[7] = yz; virial[8] = zz;
    }
    stats_out[0] = s[7][0];
    stats_out[1] += s[7][0];  // running total since the last reset
  }
}

// sum of nitems partials into value / virial / stats in two stages
static void reduce_partials(hipStream_t s, int nitems, const double* partial, double* scratch, double* value,
                            double* virial, double* stats) {
  if (nitems > 4 * kRedBlocks) {
    hipLaunchKernelGGL(k_reduce1, dim3(kRedBlocks), dim3(256), 0, s, nitems, partial, scratch);
    hipLaunchKernelGGL(k_finalize, dim3(1), dim3(128), 0, s, kRedBlocks, scratch, value, virial, stats);
  } else {
    hipLaunchKernelGGL(k_finalize, dim3(1), dim3(128), 0, s, nitems, partial, value, virial, stats);
  }
}

// ------------------------------------------------------------------------
// FP64 fallback kernels
// ------------------------------------------------------------------------
// Every requested atom accumulates its own derivative from all its partners
// (no atomics); the value and the virial are taken once per pair.
__global__ void k_allpairs(int nA, int nB, int use_pbc, PbcArg P, p2a_switch sw, const double* __restrict__ pos,
                           const long long* __restrict__ absidx, double* __restrict__ deriv,
                           double* __restrict__ partial) {
  const int natoms = nA + nB;
  const int i = blockIdx.x * blockDim.x + threadIdx.x;
  double cv = 0, v[6] = {0, 0, 0, 0, 0, 0}, f[3] = {0, 0, 0};
  double np = 0;
  if (i < natoms) {
    const bool two = nB > 0;
    int j0, j1;
    if (two) {
      if (i < nA) { j0 = nA; j1 = natoms; } else { j0 = 0; j1 = nA; }
    } else {
      j0 = 0;
      j1 = nA;
    }
    const double* xi = pos + 3 * (size_t)i;
    for (int j = j0; j < j1; j++) {
      if (j == i || absidx[i] == absidx[j]) continue;
      const double* xj = pos + 3 * (size_t)j;
      // pair (p,q) is oriented p -> q with p in GROUPA (two groups) or p<q
      bool own = two ? (i < nA) : (i < j);
      double d[3];
      if (use_pbc) pbc_distance(P, xi, xj, d);
      else { d[0] = xj[0] - xi[0]; d[1] = xj[1] - xi[1]; d[2] = xj[2] - xi[2]; }
      double r2 = d[0] * d[0] + d[1] * d[1] + d[2] * d[2];
      double df;
      double s = switch_calc_sqr(sw, r2, df);
      // deriv[i] -= df*d for d = x_j - x_i, the same for either orientation
      f[0] -= df * d[0];
      f[1] -= df * d[1];
      f[2] -= df * d[2];
      if (own) {
        cv += s;
        if (switch_inside(sw, r2)) np += 1;
        v[0] -= df * d[0] * d[0];
        v[1] -= df * d[0] * d[1];
        v[2] -= df * d[0] * d[2];
        v[3] -= df * d[1] * d[1];
        v[4] -= df * d[1] * d[2];
        v[5] -= df * d[2] * d[2];
      }
    }
    deriv[3 * (size_t)i + 0] = f[0];
    deriv[3 * (size_t)i + 1] = f[1];
    deriv[3 * (size_t)i + 2] = f[2];
  }
  // block partials: cv, -virial terms are stored as +sum g (x) d for k_finalize
  double r[8] = {cv, -v[0], -v[1], -v[2], -v[3], -v[4], -v[5], np};
  const int lane = threadIdx.x & 63;
  for (int q = 0; q < 8; q++) r[q] = wave_sum(r[q]);
  __shared__ double red[4][8];
  if (lane == 0)
    for (int q = 0; q < 8; q++) red[threadIdx.x >> 6][q] = r[q];
  __syncthreads();
  if (threadIdx.x == 0) {
    for (int q = 0; q < 8; q++) {
      double t = 0;
      for (int w = 0; w < (int)(blockDim.x / 64); w++) t += red[w][q];
      partial[(size_t)blockIdx.x * 8 + q] = t;
    }
  }
}

// PAIR: the k-th atom of GROUPA with the k-th of GROUPB
__global__ void k_pairmode(int nA, int use_pbc, PbcArg P, p2a_switch sw, const double* __restrict__ pos,
                           const long long* __restrict__ absidx, double* __restrict__ deriv,
                           double* __restrict__ partial) {
  const int i = blockIdx.x * blockDim.x + threadIdx.x;
  double r[8] = {0, 0, 0, 0, 0, 0, 0, 0};
  if (i < nA) {
    const int j = nA + i;
    double* di = deriv + 3 * (size_t)i;
    double* dj = deriv + 3 * (size_t)j;
    di[0] = di[1] = di[2] = 0.0;
    dj[0] = dj[1] = dj[2] = 0.0;
    if (absidx[i] != absidx[j]) {
      const double* xi = pos + 3 * (size_t)i;
      const double* xj = pos + 3 * (size_t)j;
      double d[3];
      if (use_pbc) pbc_distance(P, xi, xj, d);
      else { d[0] = xj[0] - xi[0]; d[1] = xj[1] - xi[1]; d[2] = xj[2] - xi[2]; }
      double r2 = d[0] * d[0] + d[1] * d[1] + d[2] * d[2];
      double df;
      double s = switch_calc_sqr(sw, r2, df);
      for (int k = 0; k < 3; k++) {
        di[k] = -df * d[k];
        dj[k] = df * d[k];
      }
      r[0] = s;
      r[1] = df * d[0] * d[0];
      r[2] = df * d[0] * d[1];
      r[3] = df * d[0] * d[2];
      r[4] = df * d[1] * d[1];
      r[5] = df * d[1] * d[2];
      r[6] = df * d[2] * d[2];
      r[7] = switch_inside(sw, r2) ? 1.0 : 0.0;
    }
  }
  const int lane = threadIdx.x & 63;
  for (int q = 0; q < 8; q++) r[q] = wave_sum(r[q]);
  __shared__ double red[4][8];
  if (lane == 0)
    for (int q = 0; q < 8; q++) red[threadIdx.x >> 6][q] = r[q];
  __syncthreads();
  if (threadIdx.x == 0) {
    for (int q = 0; q < 8; q++) {
      double t = 0;
      for (int w = 0; w < (int)(blockDim.x / 64); w++) t += red[w][q];
      partial[(size_t)blockIdx.x * 8 + q] = t;
    }
  }
}

// ------------------------------------------------------------------------
// host context
// ------------------------------------------------------------------------
struct Dir {  // one cluster-pair list: i clusters of group I against group J
  int I = 0, J = 0;
  DBuf<int> ecount, estart, istart;
  DBuf<Entry> ents;
  DBuf<int2> items;
  DBuf<double> fpart;  // per item: derivative partials of the 8 i atoms
  int nitems = 0, item_offset = 0;
  int64_t nentries = 0;
};

struct Group {
  int off = 0, n = 0;
  int ncx = 1, ncy = 1, ncol = 1;
  int ncl = 0;  // clusters (device total copied at rebuild)
  DBuf<int4> info;
  DBuf<double> key;
  DBuf<int> col_count, col_start, col_cl, fill, tmp, slot_loc, cl_col, abs_slot;
  DBuf<double4> origin;
  DBuf<float4> bb0, bb, p;
};

}  // namespace p2a

struct p2a_coord {
  int device = 0;
  hipStream_t own_stream = nullptr;
  hipStream_t stream = nullptr;
  p2a_coord_spec spec{};
  std::vector<int64_t> abs;  // absolute index of each requested atom
  int natoms = 0;
  bool firsttime = true;
  int path = 0;
  double rlist = 0.0;
  double box_at_build[9] = {0};
  bool built = false;
  p2a::Group g[2];
  p2a::DBuf<int> totals;       // [4..7] atoms/clusters per group, [8..11] entries/items per direction
  p2a::Dir dir[2];
  int ndir = 0;
  int nitems = 0;
  int64_t nentries = 0;
  p2a::DBuf<double> partial, red, stats, pos_buf, deriv_buf, out_buf;
  p2a::DBuf<long long> abs_dev;
  int64_t rebuilds = 0;
  bool profiling = false;
  double prof_ms = 0.0;
  int64_t prof_launches = 0;
  std::vector<hipEvent_t> ev_pool;
  std::vector<std::pair<hipEvent_t, hipEvent_t>> pending;
  bool has_duplicates = false;  // an absolute index listed twice in one group
  p2a::PbcHost pbc;
  p2a::PbcArg pbc_stage;
  p2a::DBuf<p2a::PbcArg> pbc_dev;
  double pbc_box[9] = {0};
  bool pbc_valid = false;
};

namespace p2a {

static void box_arg(const double* b9, BoxArg& B) {
  PbcHost tmp;
  pbc_setup(b9, tmp);
  std::memcpy(B.a, b9, sizeof(B.a));
  std::memcpy(B.inv, tmp.inv, sizeof(B.inv));
}

static int grid1(size_t n, int b) { return (int)((n + b - 1) / b); }

static int launch_check(const char* what) {
  hipError_t e = hipGetLastError();
  if (e != hipSuccess) return fail(P2A_ERR_HIP, std::string(what) + ": " + hipGetErrorString(e));
  return P2A_OK;
}

// decides the evaluation path for this box; 0 = cluster kernel
static int choose_path(p2a_coord* c, const double* box9) {
  const p2a_switch& sw = c->spec.sw;
  if (c->spec.pair && c->spec.nb > 0) return 2;
  if (!c->spec.pbc || c->has_duplicates) return 1;
  PbcHost p;
  pbc_setup(box9, p);
  if (p.type == 0) return 1;
  if (!(sw.dmax < 1e30)) return 1;
  double rl = c->spec.nl_mode != P2A_NL_NONE ? std::max(c->spec.nl_cutoff, sw.dmax) : sw.dmax;
  double w[3];
  pbc_widths(box9, w);
  double wmin = std::min(w[0], std::min(w[1], w[2]));
  if (wmin <= 2.0 * rl * 1.0001 + 1e-6) return 1;
  return 0;
}

static int rebuild(p2a_coord* c, const double* box9, const double* pos_dev) {
  BoxArg B;
  box_arg(box9, B);
  double w[3];
  pbc_widths(box9, w);
  const p2a_switch& sw = c->spec.sw;
  // list radius: NL_CUTOFF when a neighbour list is requested (its candidates
  // must cover every pair that can come within D_MAX before the next rebuild),
  // D_MAX otherwise (the list is rebuilt every step)
  c->rlist = (c->spec.nl_mode != P2A_NL_NONE ? std::max(c->spec.nl_cutoff, sw.dmax) : sw.dmax) + 1e-4;
  const int ngroups = c->spec.nb > 0 ? 2 : 1;
  double vol = std::fabs(w[0] * w[1] * w[2]);
  for (int gi = 0; gi < ngroups; gi++) {
    Group& G = c->g[gi];
    G.off = gi == 0 ? 0 : (int)c->spec.na;
    G.n = gi == 0 ? (int)c->spec.na : (int)c->spec.nb;
    double rho = std::max(G.n, 1) / vol;
    double side = std::cbrt((double)CL / rho);
    G.ncx = std::max(1, (int)std::floor(w[0] / side));
    G.ncy = std::max(1, (int)std::floor(w[1] / side));
    // keep the column count bounded for very dilute groups
    while ((int64_t)G.ncx * G.ncy > 4 * (int64_t)G.n + 16) {
      G.ncx = std::max(1, G.ncx / 2);
      G.ncy = std::max(1, G.ncy / 2);
    }
    G.ncol = G.ncx * G.ncy;
    int maxcl = G.n / CL + G.ncol + 1;
    int rc;
    if ((rc = G.info.reserve(G.n)) || (rc = G.key.reserve(G.n)) || (rc = G.col_count.reserve(G.ncol + 1)) ||
        (rc = G.col_start.reserve(G.ncol + 1)) || (rc = G.col_cl.reserve(G.ncol + 1)) ||
        (rc = G.fill.reserve(G.ncol + 1)) || (rc = G.tmp.reserve(G.n)) || (rc = G.slot_loc.reserve((size_t)maxcl * CL)) ||
        (rc = G.cl_col.reserve(maxcl)) || (rc = G.origin.reserve(maxcl)) || (rc = G.bb0.reserve(2 * (size_t)maxcl)) ||
        (rc = G.bb.reserve(2 * (size_t)maxcl)) || (rc = G.p.reserve((size_t)maxcl * CL)) ||
        (rc = G.abs_slot.reserve((size_t)maxcl * CL)))
      return rc;
  }
  int rc;
  if ((rc = c->totals.reserve(16))) return rc;
  hipStream_t s = c->stream;
  for (int gi = 0; gi < ngroups; gi++) {
    Group& G = c->g[gi];
    hipLaunchKernelGGL(k_fill_int, dim3(grid1(G.ncol + 1, 256)), dim3(256), 0, s, G.col_count.p, G.ncol + 1, 0);
    hipLaunchKernelGGL(k_fill_int, dim3(grid1(G.ncol + 1, 256)), dim3(256), 0, s, G.fill.p, G.ncol + 1, 0);
    hipLaunchKernelGGL(k_fill_int, dim3(grid1(G.slot_loc.n, 256)), dim3(256), 0, s, G.slot_loc.p, (int)G.slot_loc.n, -1);
    if (G.n > 0)
      hipLaunchKernelGGL(k_bin, dim3(grid1(G.n, 256)), dim3(256), 0, s, G.n, pos_dev, G.off, B, G.ncx, G.ncy,
                         G.info.p, G.key.p, G.col_count.p);
    hipLaunchKernelGGL(k_scan_cols, dim3(1), dim3(1024), 0, s, G.ncol, G.col_count.p, G.col_start.p, G.col_cl.p,
                       c->totals.p + 4 + 2 * gi);
    if (G.n > 0) {
      hipLaunchKernelGGL(k_scatter, dim3(grid1(G.n, 256)), dim3(256), 0, s, G.n, G.info.p, G.col_start.p, G.fill.p,
                         G.tmp.p);
      hipLaunchKernelGGL(k_rank, dim3(G.ncol), dim3(256), 0, s, G.ncol, G.key.p, G.col_start.p, G.col_cl.p,
                         G.tmp.p, G.slot_loc.p);
    }
    hipLaunchKernelGGL(k_cl_col, dim3(grid1(G.ncol, 256)), dim3(256), 0, s, G.ncol, G.col_cl.p, G.cl_col.p);
    if ((rc = launch_check("rebuild: binning"))) return rc;
  }
  int hostTot[8];
  P2A_HIP(hipMemcpyAsync(hostTot, c->totals.p, sizeof(hostTot), hipMemcpyDeviceToHost, s));
  P2A_HIP(hipStreamSynchronize(s));
  for (int gi = 0; gi < ngroups; gi++) {
    Group& G = c->g[gi];
    G.ncl = hostTot[4 + 2 * gi + 1];
    if (G.ncl > 0)
      hipLaunchKernelGGL(k_origin, dim3(grid1(G.ncl, 128)), dim3(128), 0, s, G.ncl, G.slot_loc.p, G.info.p, pos_dev,
                         G.off, B, G.origin.p, G.bb0.p);
  }
  // full cluster-pair lists: one direction for a single group, A->B (the
  // counting side) and B->A for two groups
  c->ndir = ngroups;
  int64_t tot_entries = 0;
  const int wpb = 256 / WAVE;
  ListArgs Ls[2];
  for (int d = 0; d < c->ndir; d++) {
    Dir& D = c->dir[d];
    D.I = (ngroups == 2) ? d : 0;
    D.J = (ngroups == 2) ? 1 - d : 0;
    Group& GI = c->g[D.I];
    Group& GJ = c->g[D.J];
    ListArgs& L = Ls[d];
    L.nclI = GI.ncl;
    L.ncxI = GI.ncx;
    L.ncyI = GI.ncy;
    L.ncxJ = GJ.ncx;
    L.ncyJ = GJ.ncy;
    L.mode = (ngroups == 1) ? 0 : (d == 0 ? 1 : 2);
    L.rlist2 = (float)(c->rlist * c->rlist);
    L.rb[0] = (float)(c->rlist / w[0]);
    L.rb[1] = (float)(c->rlist / w[1]);
    for (int k = 0; k < 9; k++) {
      L.A[k] = (float)box9[k];
      L.Ad[k] = box9[k];
    }
    if ((rc = D.ecount.reserve(GI.ncl + 1)) || (rc = D.estart.reserve(GI.ncl + 1)) ||
        (rc = D.istart.reserve(GI.ncl + 1)))
      return rc;
    if (GI.ncl > 0) {
      hipLaunchKernelGGL(k_list<0>, dim3(grid1(GI.ncl, wpb)), dim3(256), 0, s, L, GI.cl_col.p, GI.bb0.p, GJ.bb0.p,
                         GJ.col_cl.p, GI.origin.p, GJ.origin.p, D.ecount.p, (const int*)nullptr, (Entry*)nullptr);
      hipLaunchKernelGGL(k_scan, dim3(1), dim3(1024), 0, s, GI.ncl, D.ecount.p, D.estart.p, c->totals.p, 8 + 2 * d, 0);
      hipLaunchKernelGGL(k_scan, dim3(1), dim3(1024), 0, s, GI.ncl, D.ecount.p, D.istart.p, c->totals.p, 9 + 2 * d, 1);
    } else {
      hipLaunchKernelGGL(k_fill_int, dim3(1), dim3(64), 0, s, c->totals.p + 8 + 2 * d, 2, 0);
    }
  }
  if ((rc = launch_check("rebuild: list count"))) return rc;
  int hostT[16];
  P2A_HIP(hipMemcpyAsync(hostT, c->totals.p, sizeof(hostT), hipMemcpyDeviceToHost, s));
  P2A_HIP(hipStreamSynchronize(s));
  c->nitems = 0;
  for (int d = 0; d < c->ndir; d++) {
    Dir& D = c->dir[d];
    Group& GI = c->g[D.I];
    Group& GJ = c->g[D.J];
    D.nentries = hostT[8 + 2 * d];
    D.nitems = hostT[9 + 2 * d];
    D.item_offset = c->nitems;
    c->nitems += D.nitems;
    tot_entries += D.nentries;
    if ((rc = D.ents.reserve(std::max<int64_t>(D.nentries, 1))) || (rc = D.items.reserve(std::max(D.nitems, 1))) ||
        (rc = D.fpart.reserve((size_t)CL * 3 * std::max(D.nitems, 1))))
      return rc;
    if (GI.ncl > 0) {
      hipLaunchKernelGGL(k_list<1>, dim3(grid1(GI.ncl, wpb)), dim3(256), 0, s, Ls[d], GI.cl_col.p, GI.bb0.p,
                         GJ.bb0.p, GJ.col_cl.p, GI.origin.p, GJ.origin.p, D.ecount.p, (const int*)D.estart.p,
                         D.ents.p);
      hipLaunchKernelGGL(k_items, dim3(grid1(GI.ncl, 256)), dim3(256), 0, s, GI.ncl, D.ecount.p, D.istart.p,
                         D.items.p);
    }
  }
  c->nentries = tot_entries;
  if ((rc = c->partial.reserve(8 * (size_t)std::max(c->nitems, 1)))) return rc;
  // absolute index per slot (two groups: pairs of one atom with itself are skipped)
  if (ngroups == 2) {
    std::vector<int> loc[2];
    for (int gi = 0; gi < 2; gi++) {
      Group& G = c->g[gi];
      std::vector<int> sl((size_t)G.ncl * CL);
      if (G.ncl > 0)
        P2A_HIP(hipMemcpyAsync(sl.data(), G.slot_loc.p, sl.size() * sizeof(int), hipMemcpyDeviceToHost, s));
      P2A_HIP(hipStreamSynchronize(s));
      for (auto& v : sl) v = (v >= 0) ? (int)c->abs[G.off + v] : -1 - (int)(&v - sl.data());
      if (G.ncl > 0)
        P2A_HIP(hipMemcpyAsync(G.abs_slot.p, sl.data(), sl.size() * sizeof(int), hipMemcpyHostToDevice, s));
      P2A_HIP(hipStreamSynchronize(s));
    }
  }
  if ((rc = launch_check("rebuild: list fill"))) return rc;
  std::memcpy(c->box_at_build, box9, sizeof(c->box_at_build));
  c->built = true;
  c->rebuilds++;
  return P2A_OK;
}

static void fill_pbc_arg(const PbcHost& h, PbcArg& P) {
  std::memset(&P, 0, sizeof(P));
  P.type = h.type;
  std::memcpy(P.box, h.box, sizeof(P.box));
  std::memcpy(P.inv, h.inv, sizeof(P.inv));
  std::memcpy(P.red, h.reduced, sizeof(P.red));
  std::memcpy(P.ired, h.invReduced, sizeof(P.ired));
  std::memcpy(P.nshift, h.nshift, sizeof(P.nshift));
  std::memcpy(P.shifts, h.shifts, sizeof(P.shifts));
}

// prepared cell for this box, on the host (fallback kernels take it by value)
// and on the device (the pair kernel reads it for its FP64 cut decisions)
static int upload_pbc(p2a_coord* c, const double* box9) {
  int rc;
  if (c->pbc_valid && std::memcmp(c->pbc_box, box9, sizeof(c->pbc_box)) == 0 && c->pbc_dev.p) return P2A_OK;
  if ((rc = c->pbc_dev.reserve(1))) return rc;
  // the previous upload may still read the staging copy
  P2A_HIP(hipStreamSynchronize(c->stream));
  pbc_setup(box9, c->pbc);
  std::memcpy(c->pbc_box, box9, sizeof(c->pbc_box));
  c->pbc_valid = true;
  fill_pbc_arg(c->pbc, c->pbc_stage);
  P2A_HIP(hipMemcpyAsync(c->pbc_dev.p, &c->pbc_stage, sizeof(PbcArg), hipMemcpyHostToDevice, c->stream));
  return P2A_OK;
}

static int run_pairs(p2a_coord* c, const double* box9, const double* pos_dev, double* value_dev, double* deriv_dev,
                     double* virial_dev, bool value_only) {
  int rc;
  hipStream_t s = c->stream;
  const int ngroups = c->spec.nb > 0 ? 2 : 1;
  BoxArg B;
  box_arg(box9, B);
  for (int gi = 0; gi < ngroups; gi++) {
    Group& G = c->g[gi];
    if (G.ncl == 0) continue;
    hipLaunchKernelGGL(k_pack, dim3(grid1((size_t)G.ncl * CL, 256)), dim3(256), 0, s, G.ncl * CL, G.slot_loc.p,
                       G.info.p, pos_dev, G.off, B, G.origin.p, G.p.p);
    hipLaunchKernelGGL(k_bbox, dim3(grid1(G.ncl, 256)), dim3(256), 0, s, G.ncl, G.p.p, G.bb.p);
  }
  const p2a_switch& sw = c->spec.sw;
  PairArgs A;
  std::memset(&A, 0, sizeof(A));
  A.twogroup = ngroups == 2;
  A.dmax2 = (float)sw.dmax_2;
  double pr = sw.dmax + 1e-4;
  A.prune2 = (float)(pr * pr);
  A.f.dmax2 = (float)sw.dmax_2;
  A.f.invr0_2 = (float)sw.invr0_2;
  A.f.stretch = (float)sw.stretch;
  A.f.shift = (float)sw.shift;
  A.f.cder = (float)(-(double)sw.nn * sw.invr0_2 * sw.stretch);
  A.sw = sw;
  if ((rc = upload_pbc(c, box9))) return rc;
  A.pbc = c->pbc_dev.p;
  hipEvent_t e0 = nullptr, e1 = nullptr;
  if (c->profiling) {
    if (c->ev_pool.size() < 2) {
      hipEvent_t a, b;
      P2A_HIP(hipEventCreate(&a));
      P2A_HIP(hipEventCreate(&b));
      c->ev_pool.push_back(a);
      c->ev_pool.push_back(b);
    }
    e1 = c->ev_pool.back();
    c->ev_pool.pop_back();
    e0 = c->ev_pool.back();
    c->ev_pool.pop_back();
    P2A_HIP(hipEventRecord(e0, s));
  }
  // value only: the counting direction (the first) holds every counted pair
  const int ndir_run = value_only ? 1 : c->ndir;
  for (int d = 0; d < ndir_run; d++) {
    Dir& D = c->dir[d];
    if (D.nitems == 0) continue;
    Group& GI = c->g[D.I];
    Group& GJ = c->g[D.J];
    A.nitems = D.nitems;
    dim3 grid(grid1(D.nitems, PAIR_BLOCK / WAVE)), block(PAIR_BLOCK);
    double* part = c->partial.p + 8 * (size_t)D.item_offset;
#define P2A_PAIRS_VO(P, GEN, TWO, VO)                                                                          \
  hipLaunchKernelGGL((k_pairs<P, GEN, TWO, VO>), grid, block, 0, s, A, D.items.p, D.estart.p, D.ents.p, GI.p.p, \
                     GJ.p.p, GI.bb.p, GJ.bb.p, GI.abs_slot.p, GJ.abs_slot.p, pos_dev, D.fpart.p, part)
#define P2A_PAIRS(P, GEN)                                    \
  do {                                                       \
    if (A.twogroup) {                                        \
      if (value_only) P2A_PAIRS_VO(P, GEN, true, true);      \
      else P2A_PAIRS_VO(P, GEN, true, false);                \
    } else {                                                 \
      if (value_only) P2A_PAIRS_VO(P, GEN, false, true);     \
      else P2A_PAIRS_VO(P, GEN, false, false);               \
    }                                                        \
  } while (0)
    if (sw.type == P2A_SW_RATIONAL_FIXED) {
      switch (sw.nn / 2) {
        case 1: P2A_PAIRS(1, false); break;
        case 2: P2A_PAIRS(2, false); break;
        case 3: P2A_PAIRS(3, false); break;
        case 4: P2A_PAIRS(4, false); break;
        case 5: P2A_PAIRS(5, false); break;
        case 6: P2A_PAIRS(6, false); break;
        default: P2A_PAIRS(1, true); break;
      }
    } else {
      P2A_PAIRS(1, true);
    }
#undef P2A_PAIRS
#undef P2A_PAIRS_VO
  }
  if (c->profiling) {
    P2A_HIP(hipEventRecord(e1, s));
    c->pending.emplace_back(e0, e1);
  }
  if ((rc = launch_check("pair kernel"))) return rc;
  for (int d = 0; d < c->ndir && !value_only; d++) {
    Dir& D = c->dir[d];
    Group& GI = c->g[D.I];
    if (GI.ncl == 0) continue;
    hipLaunchKernelGGL(k_gather, dim3(grid1((size_t)GI.ncl * CL * 3, 256)), dim3(256), 0, s, GI.ncl, GI.p.p,
                       D.istart.p, D.fpart.p, deriv_dev);
  }
  if ((rc = c->red.reserve(8 * kRedBlocks))) return rc;
  reduce_partials(s, value_only ? c->dir[0].nitems : c->nitems, c->partial.p, c->red.p, value_dev, virial_dev,
                  c->stats.p);
  return launch_check("finalize");
}

static int run_fallback(p2a_coord* c, const double* box9, const double* pos_dev, double* value_dev,
                        double* deriv_dev, double* virial_dev) {
  hipStream_t s = c->stream;
  int rc0;
  if ((rc0 = upload_pbc(c, box9))) return rc0;
  const PbcArg& P = c->pbc_stage;
  const int nA = (int)c->spec.na, nB = (int)c->spec.nb;
  int nblocks;
  int rc;
  if (c->path == 2) {
    nblocks = grid1(nA, 256);
    if ((rc = c->partial.reserve(8 * (size_t)std::max(nblocks, 1)))) return rc;
    hipLaunchKernelGGL(k_pairmode, dim3(std::max(nblocks, 1)), dim3(256), 0, s, nA, c->spec.pbc, P, c->spec.sw,
                       pos_dev, c->abs_dev.p, deriv_dev, c->partial.p);
  } else {
    nblocks = grid1(c->natoms, 256);
    if ((rc = c->partial.reserve(8 * (size_t)std::max(nblocks, 1)))) return rc;
    hipLaunchKernelGGL(k_allpairs, dim3(std::max(nblocks, 1)), dim3(256), 0, s, nA, nB, c->spec.pbc, P, c->spec.sw,
                       pos_dev, c->abs_dev.p, deriv_dev, c->partial.p);
  }
  if ((rc = launch_check("fp64 pair kernel"))) return rc;
  if ((rc = c->red.reserve(8 * kRedBlocks))) return rc;
  reduce_partials(s, std::max(nblocks, 1), c->partial.p, c->red.p, value_dev, virial_dev, c->stats.p);
  return launch_check("finalize");
}

static int calc_device(p2a_coord* c, int64_t step, const double* box9, const double* pos_dev, double* value_dev,
                       double* deriv_dev, double* virial_dev) {
  int rc;
  if (!box9) return fail(P2A_ERR_ARG, "box is required");
  if (!c->stats.p) {
    if ((rc = c->stats.reserve(8))) return rc;
    hipLaunchKernelGGL(k_fill_double, dim3(1), dim3(64), 0, c->stream, c->stats.p, (size_t)8);
  }
  // no derivative and no virial requested: the cheaper value-only kernel
  const bool value_only = (deriv_dev == nullptr && virial_dev == nullptr);
  double* vdev = value_dev;
  double* wdev = virial_dev;
  double* ddev = deriv_dev;
  if (!vdev || !wdev) {
    if ((rc = c->out_buf.reserve(16))) return rc;
    if (!vdev) vdev = c->out_buf.p;
    if (!wdev) wdev = c->out_buf.p + 1;
  }
  if (!ddev) {
    if ((rc = c->deriv_buf.reserve(3 * (size_t)std::max(c->natoms, 1)))) return rc;
    ddev = c->deriv_buf.p;
  }
  int path = choose_path(c, box9);
  if (path != c->path) c->built = false;
  c->path = path;
  if (path != 0) return run_fallback(c, box9, pos_dev, vdev, ddev, wdev);
  // NeighborList::prepare (NeighborList.cpp:433-456): rebuild on the first
  // step and every NL_STRIDE steps; without a neighbour list every step.
  bool need = !c->built || c->firsttime;
  if (c->spec.nl_mode == P2A_NL_NONE || c->spec.nl_stride <= 1) need = true;
  else if (step % c->spec.nl_stride == 0) need = true;
  // the images of the list belong to one cell: a new cell needs a new list
  if (!need && std::memcmp(c->box_at_build, box9, sizeof(c->box_at_build)) != 0) need = true;
  c->firsttime = false;
  if (need && (rc = rebuild(c, box9, pos_dev))) return rc;
  return run_pairs(c, box9, pos_dev, vdev, ddev, wdev, value_only);
}

static void collect_profile(p2a_coord* c) {
  for (auto& pr : c->pending) {
    float ms = 0.f;
    if (hipEventElapsedTime(&ms, pr.first, pr.second) == hipSuccess) {
      c->prof_ms += ms;
      c->prof_launches++;
    }
    c->ev_pool.push_back(pr.first);
    c->ev_pool.push_back(pr.second);
  }
  c->pending.clear();
}

}  // namespace p2a

using namespace p2a;

extern "C" {

int p2a_device_count(void) {
  int n = 0;
  if (hipGetDeviceCount(&n) != hipSuccess) return 0;
  return n;
}

int p2a_coord_create(p2a_coord** out, int device, const p2a_coord_spec* spec) {
  if (!out || !spec) return fail(P2A_ERR_ARG, "null argument");
  *out = nullptr;
  if (spec->na <= 0) return fail(P2A_ERR_ARG, "GROUPA should contain at least one atom");
  if (spec->pair && spec->na != spec->nb)
    return fail(P2A_ERR_ARG, "when using PAIR the two groups should have the same number of atoms");
  if (spec->nl_mode != P2A_NL_NONE && (spec->nl_cutoff <= 0.0 || spec->nl_stride <= 0))
    return fail(P2A_ERR_ARG, "NL_CUTOFF and NL_STRIDE should be explicitly specified and positive");
  if (spec->na + spec->nb > (int64_t)1 << 29) return fail(P2A_ERR_UNSUPPORTED, "more than 2^29 atoms");
  if (spec->nl_mode != P2A_NL_NONE && spec->sw.dmax > spec->nl_cutoff)
    return fail(P2A_ERR_UNSUPPORTED,
                "D_MAX larger than NL_CUTOFF: the stale neighbour list of the reference would drop pairs; "
                "use D_MAX <= NL_CUTOFF");
  int ndev = 0;
  if (hipGetDeviceCount(&ndev) != hipSuccess || ndev == 0) return fail(P2A_ERR_HIP, "no HIP device available");
  if (device < 0 || device >= ndev) return fail(P2A_ERR_ARG, "device index out of range");
  P2A_HIP(hipSetDevice(device));
  auto* c = new p2a_coord();
  c->device = device;
  c->spec = *spec;
  c->spec.groupa = nullptr;
  c->spec.groupb = nullptr;
  c->natoms = (int)(spec->na + spec->nb);
  c->abs.resize(c->natoms);
  for (int64_t i = 0; i < spec->na; i++) c->abs[i] = spec->groupa[i];
  for (int64_t i = 0; i < spec->nb; i++) c->abs[spec->na + i] = spec->groupb[i];
  for (int gi = 0; gi < 2; gi++) {
    std::vector<int64_t> v(gi == 0 ? spec->groupa : spec->groupb,
                           (gi == 0 ? spec->groupa : spec->groupb) + (gi == 0 ? spec->na : spec->nb));
    std::sort(v.begin(), v.end());
    if (std::adjacent_find(v.begin(), v.end()) != v.end()) c->has_duplicates = true;
  }
  if (hipStreamCreateWithFlags(&c->own_stream, hipStreamNonBlocking) != hipSuccess) {
    delete c;
    return fail(P2A_ERR_HIP, "hipStreamCreate failed");
  }
  c->stream = c->own_stream;
  int rc;
  if ((rc = c->abs_dev.reserve(c->natoms))) {
    p2a_coord_destroy(c);
    return rc;
  }
  std::vector<long long> ab(c->abs.begin(), c->abs.end());
  if (hipMemcpy(c->abs_dev.p, ab.data(), ab.size() * sizeof(long long), hipMemcpyHostToDevice) != hipSuccess) {
    p2a_coord_destroy(c);
    return fail(P2A_ERR_HIP, "hipMemcpy of atom indices failed");
  }
  *out = c;
  g_last_error.clear();
  return P2A_OK;
}

int p2a_coord_destroy(p2a_coord* c) {
  if (!c) return P2A_OK;
  (void)hipSetDevice(c->device);
  if (c->stream) (void)hipStreamSynchronize(c->stream);
  for (auto& pr : c->pending) {
    c->ev_pool.push_back(pr.first);
    c->ev_pool.push_back(pr.second);
  }
  for (auto e : c->ev_pool) (void)hipEventDestroy(e);
  if (c->own_stream) (void)hipStreamDestroy(c->own_stream);
  delete c;
  return P2A_OK;
}

int p2a_coord_set_stream(p2a_coord* c, void* stream) {
  if (!c) return fail(P2A_ERR_ARG, "null context");
  // NULL selects the default (null) stream, which orders with every blocking stream
  c->stream = (hipStream_t)stream;
  return P2A_OK;
}

int p2a_coord_calculate_device(p2a_coord* c, int64_t step, const double* box9, const double* pos_dev,
                               double* value_dev, double* deriv_dev, double* virial_dev) {
  if (!c || !pos_dev) return fail(P2A_ERR_ARG, "null argument");
  P2A_HIP(hipSetDevice(c->device));
  return calc_device(c, step, box9, pos_dev, value_dev, deriv_dev, virial_dev);
}

int p2a_coord_backward_device(p2a_coord* c, const double* box9, const double* pos_dev, double* deriv_dev,
                              double* virial_dev) {
  if (!c || !pos_dev || !box9 || !deriv_dev) return fail(P2A_ERR_ARG, "null argument");
  P2A_HIP(hipSetDevice(c->device));
  int rc;
  if ((rc = c->out_buf.reserve(16))) return rc;
  double* wdev = virial_dev ? virial_dev : c->out_buf.p + 1;
  if (c->path != 0 || choose_path(c, box9) != 0) return run_fallback(c, box9, pos_dev, c->out_buf.p, deriv_dev, wdev);
  if (!c->built || std::memcmp(c->box_at_build, box9, sizeof(c->box_at_build)) != 0)
    return fail(P2A_ERR_ARG, "backward needs the neighbour structure of a forward call on the same box");
  return run_pairs(c, box9, pos_dev, c->out_buf.p, deriv_dev, wdev, false);
}

int p2a_coord_calculate(p2a_coord* c, int64_t step, const double* box9, const double* pos, double* value,
                        double* deriv, double* virial) {
  if (!c || !pos) return fail(P2A_ERR_ARG, "null argument");
  P2A_HIP(hipSetDevice(c->device));
  int rc;
  const size_t n3 = 3 * (size_t)c->natoms;
  if ((rc = c->pos_buf.reserve(n3)) || (rc = c->deriv_buf.reserve(n3)) || (rc = c->out_buf.reserve(16))) return rc;
  P2A_HIP(hipMemcpyAsync(c->pos_buf.p, pos, n3 * sizeof(double), hipMemcpyHostToDevice, c->stream));
  if ((rc = calc_device(c, step, box9, c->pos_buf.p, c->out_buf.p, deriv ? c->deriv_buf.p : nullptr,
                        (deriv || virial) ? c->out_buf.p + 1 : nullptr)))
    return rc;
  double out[10];
  P2A_HIP(hipMemcpyAsync(out, c->out_buf.p, 10 * sizeof(double), hipMemcpyDeviceToHost, c->stream));
  if (deriv) P2A_HIP(hipMemcpyAsync(deriv, c->deriv_buf.p, n3 * sizeof(double), hipMemcpyDeviceToHost, c->stream));
  P2A_HIP(hipStreamSynchronize(c->stream));
  if (value) *value = out[0];
  if (virial) std::memcpy(virial, out + 1, 9 * sizeof(double));
  return P2A_OK;
}

int p2a_coord_get_stats(p2a_coord* c, p2a_coord_stats* st) {
  if (!c || !st) return fail(P2A_ERR_ARG, "null argument");
  P2A_HIP(hipSetDevice(c->device));
  P2A_HIP(hipStreamSynchronize(c->stream));
  std::memset(st, 0, sizeof(*st));
  st->natoms = c->natoms;
  if (c->stats.p) {
    double d[2];
    P2A_HIP(hipMemcpy(d, c->stats.p, 2 * sizeof(double), hipMemcpyDeviceToHost));
    st->contributing_pairs = (int64_t)d[0];
    st->total_contributing_pairs = (int64_t)d[1];
  }
  const int ngroups = c->spec.nb > 0 ? 2 : 1;
  st->clusters_i = c->g[0].ncl;
  st->clusters_j = c->g[ngroups - 1].ncl;
  st->list_entries = c->nentries;
  st->rebuilds = c->rebuilds;
  st->path = c->path;
  collect_profile(c);
  st->pair_kernel_ms = c->prof_ms;
  st->pair_kernel_launches = c->prof_launches;
  return P2A_OK;
}

int p2a_coord_set_profiling(p2a_coord* c, int on) {
  if (!c) return fail(P2A_ERR_ARG, "null context");
  P2A_HIP(hipSetDevice(c->device));
  P2A_HIP(hipStreamSynchronize(c->stream));
  collect_profile(c);
  c->profiling = on != 0;
  if (on) {
    c->prof_ms = 0.0;
    c->prof_launches = 0;
    if (c->stats.p) {
      hipLaunchKernelGGL(k_fill_double, dim3(1), dim3(64), 0, c->stream, c->stats.p + 1, (size_t)1);
      P2A_HIP(hipStreamSynchronize(c->stream));
    }
  }
  return P2A_OK;
}

}  // extern "C"
