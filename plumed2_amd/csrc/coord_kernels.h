// Device code of the COORDINATION path: neighbour-structure build, the
// cluster-pair kernel, reductions and the FP64 kernels for the inputs the
// cluster path does not cover.  Shared by the single-system context
// (coord.hip) and the multi-system batch (batch.hip); every kernel has
// internal linkage, so each translation unit instantiates what it launches.
//
// What each piece replaces in the reference:
//   k_bin .. k_list   NeighborList::update cells branch + LinkCells
//                     (NeighborList.cpp:168-236, LinkCells.cpp:99-239); the
//                     O(N^2) pair array is never materialised, the list is one
//                     of 8x8-atom cluster pairs
//   k_pack_bbox       ActionAtomistic::retrieveAtoms gather + the image
//                     bookkeeping of Pbc::distance
//   k_pairs           CoordinationBase::calculate pair loop
//                     (CoordinationBase.cpp:178-207) with
//                     fixedRational<N>::calculateSqr (SwitchingFunction.cpp:
//                     203-215) inlined
//   k_gather_reduce,  the OpenMP critical merge + comm.Sum
//   k_finalize,       (CoordinationBase.cpp:209-224)
//   k_sys_reduce
//   k_allpairs,       the same loop in FP64 for NOPBC, PAIR, unbounded D_MAX,
//   k_pairmode        cells narrower than 2*cutoff
//
// Several systems (frames, walkers) can share one set of arrays: atoms,
// columns and clusters of system s are contiguous, SysArg[s] holds its cell
// and its column grid, and per-atom / per-cluster system indices select it.
// A single system is the case nsys = 1 with null system-index arrays.
//
// Layout in HBM:
//   slots      8 consecutive slots = one cluster; clusters are runs of atoms
//              in columns of the fractional (a,b) plane sorted along c;
//              float4 {x,y,z local to the cluster origin, requested index}
//   cl_origin  double4 per cluster, FP64 origin fixed at the rebuild
//   cl_bb      2 x float4 per cluster (centre, half extents), every step
//   cl_kr      float2 per cluster: range of fractional c of its atoms
//   list       CSR over i clusters of Entry{code, delta}: code = j cluster |
//              image << 25 | count << 30 | diagonal << 31; delta =
//              o_j + T - o_i (FP32), T the image's lattice translation
#pragma once
#include <hip/hip_runtime.h>

#include <algorithm>
#include <cfloat>
#include <climits>
#include <cmath>
#include <cstring>
#include <string>

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
    // 1/8 headroom: list sizes drift between rebuilds, and every regrowth
    // is a hipFree (which waits for the device) plus a hipMalloc
    size_t want = std::max<size_t>(m + m / 8, 1);
    if (hipMalloc(&p, want * sizeof(T)) != hipSuccess) {
      p = nullptr;
      return fail(P2A_ERR_NOMEM, "hipMalloc of " + std::to_string(want * sizeof(T)) + " bytes failed");
    }
    n = want;
    return P2A_OK;
  }
  DBuf() = default;
  DBuf(const DBuf&) = delete;
  DBuf& operator=(const DBuf&) = delete;
  ~DBuf() {
    if (p) (void)hipFree(p);
  }
};

struct Entry {
  int code;
  float dx, dy, dz;
};

// lattice vectors of a cell (rows), by value
struct BoxArg {
  double a[9];
};

// cell and column grid of one system
struct SysArg {
  double a[9];    // lattice vectors as rows
  double inv[9];  // inverse of a
  float af[9];    // lattice vectors in FP32 (list search)
  float rb[3];    // list radius as a fraction of the cell along a, b, c (r * |reciprocal row|)
  int ncx, ncy;   // columns along a and b
  int colbase;    // first column of this system in the shared column arrays
  int pad_;
};

struct PbcArg {
  int type;
  double box[9], inv[9], red[9], ired[9];
  int nshift[8];
  double shifts[8][kMaxShift][3];
};

// Entry code bits: j cluster (bits 0-24), image code (bits 25-29:
// (mx+1) + 3(my+1) + 9(mz+1), the lattice translation of the j cluster),
// "count" (bit 30: this entry adds to the value and the virial; each
// unordered pair of atoms is counted by exactly one entry), "diagonal" (bit
// 31: i cluster paired with itself without image, pairs with j == i are
// skipped and counted when j > i).
constexpr int kCjMask = 0x1FFFFFF;
constexpr int kImgShift = 25;
constexpr int kCountBit = 0x40000000;
constexpr int kMaxClusters = kCjMask;

namespace {

// ------------------------------------------------------------------------
// small device helpers
// ------------------------------------------------------------------------
__device__ __forceinline__ double wave_sum(double v) {
  for (int off = 32; off > 0; off >>= 1) v += __shfl_xor(v, off, 64);
  return v;
}

// exclusive prefix sum over the 64 lanes; *total = sum of all lanes
__device__ __forceinline__ int wave_excl_scan(int v, int lane, int& total) {
  int x = v;
#pragma unroll
  for (int o = 1; o < 64; o <<= 1) {
    const int y = __shfl_up(x, o, 64);
    if (lane >= o) x += y;
  }
  total = __shfl(x, 63, 64);
  return x - v;
}

// exclusive prefix sums of a and c over the workgroup (at most 1024 threads):
// wave scans, then the wave totals from LDS (sw: 32 ints); ta, tc receive
// the workgroup totals
__device__ __forceinline__ void block_excl_scan2(int& a, int& c, int& ta, int& tc, int* sw) {
  const int lane = threadIdx.x & 63, w = threadIdx.x >> 6, nw = (blockDim.x + 63) >> 6;
  int wa, wc;
  const int ea = wave_excl_scan(a, lane, wa);
  const int ec = wave_excl_scan(c, lane, wc);
  if (lane == 0) {
    sw[w] = wa;
    sw[16 + w] = wc;
  }
  __syncthreads();
  int pa = 0, pc = 0;
  ta = 0;
  tc = 0;
  for (int q = 0; q < nw; q++) {
    const int xa = sw[q], xc = sw[16 + q];
    if (q < w) {
      pa += xa;
      pc += xc;
    }
    ta += xa;
    tc += xc;
  }
  a = pa + ea;
  c = pc + ec;
  __syncthreads();
}

__device__ __forceinline__ int floordiv(int a, int n) { return (a >= 0) ? a / n : -((-a + n - 1) / n); }

__device__ __forceinline__ int nsys_of(const int* idx, int i) { return idx ? idx[i] : 0; }

// Workgroups are dispatched to the 8 XCDs of the chip in turn (workgroup b to
// XCD b mod 8), and each XCD has its own L2.  This index puts the workgroups
// of one XCD on one contiguous run of indices, so that an XCD takes one
// contiguous share of the spatially sorted work items and its L2 holds the
// positions of one region instead of all of them.  A bijection on
// [0, gridDim.x) for any grid size.
__device__ __forceinline__ int xcd_block() {
  const int nb = gridDim.x, b = blockIdx.x;
  const int q = nb >> 3, r = nb & 7, x = b & 7;
  return x * q + min(x, r) + (b >> 3);
}

// ------------------------------------------------------------------------
// rebuild kernels
// ------------------------------------------------------------------------

// Column counting is privatised per workgroup: the workgroup's atoms fall in
// a narrow window of columns (consecutive atoms of one system), counted in
// LDS and added to the global counts once per (workgroup, column).  A window
// wider than kBinLds falls back to one global atomic per atom.
constexpr int kBinLds = 4096;

__device__ __forceinline__ void block_col_range(int col, int* sh, int& lo, int& hi) {
  // col < 0 marks a thread without an atom
  const int t = threadIdx.x;
  sh[t] = col < 0 ? INT_MAX : col;
  sh[256 + t] = col;
  __syncthreads();
  for (int o = 128; o > 0; o >>= 1) {
    if (t < o) {
      sh[t] = min(sh[t], sh[t + o]);
      sh[256 + t] = max(sh[256 + t], sh[256 + t + o]);
    }
    __syncthreads();
  }
  lo = sh[0];
  hi = sh[256];
  __syncthreads();
}

// fractional coordinates, image and column of every atom (atom i of the
// arrays is requested atom off + i); column counts
// mmax (optional): largest |lattice translation| of a wrapped atom, as an
// atomic maximum; mclear is zeroed for the next rebuild (two alternating slots).
// order (optional): the atoms in the slot order of the previous build (norder
// slots, -1 = padding), so that a workgroup's atoms share few columns and its
// counts stay in LDS; otherwise thread i takes atom i.
__global__ __launch_bounds__(256) void k_bin(int n, const double* __restrict__ pos, int off,
                                             const SysArg* __restrict__ sys, const int* __restrict__ atom_sys,
                                             int4* __restrict__ info, double* __restrict__ key,
                                             int* __restrict__ col_count, int* __restrict__ mmax = nullptr,
                                             int* __restrict__ mclear = nullptr, const int* __restrict__ order = nullptr,
                                             int norder = 0) {
  __shared__ int sh[512];
  __shared__ int hist[kBinLds];
  const int s = blockIdx.x * blockDim.x + threadIdx.x;
  const int i = order ? (s < norder ? order[s] : -1) : (s < n ? s : -1);
  int col = -1;
  int am = 0;
  if (i >= 0) {
    const SysArg& S = sys[nsys_of(atom_sys, i)];
    const double* x = pos + 3 * (size_t)(off + i);
    double sf[3];
    int m[3];
    for (int j = 0; j < 3; j++) {
      sf[j] = x[0] * S.inv[0 * 3 + j] + x[1] * S.inv[1 * 3 + j] + x[2] * S.inv[2 * 3 + j];
      double fl = floor(sf[j]);
      m[j] = (int)fl;
      sf[j] -= fl;
    }
    int cx = min((int)(sf[0] * S.ncx), S.ncx - 1);
    int cy = min((int)(sf[1] * S.ncy), S.ncy - 1);
    cx = max(cx, 0);
    cy = max(cy, 0);
    col = S.colbase + cx + S.ncx * cy;
    info[i] = make_int4(m[0], m[1], m[2], col);
    key[i] = sf[2];
    am = max(abs(m[0]), max(abs(m[1]), abs(m[2])));
  }
  if (mmax) {
    for (int o = 32; o > 0; o >>= 1) am = max(am, __shfl_xor(am, o, 64));
    if ((threadIdx.x & 63) == 0 && am > 0) atomicMax(mmax, am);
    if (s == 0) *mclear = 0;
  }
  int lo, hi;
  block_col_range(col, sh, lo, hi);
  if (hi < 0) return;  // no atom in this workgroup (uniform)
  if (hi - lo >= kBinLds) {
    if (col >= 0) atomicAdd(&col_count[col], 1);
    return;
  }
  const int range = hi - lo + 1;
  for (int j = threadIdx.x; j < range; j += blockDim.x) hist[j] = 0;
  __syncthreads();
  if (col >= 0) atomicAdd(&hist[col - lo], 1);
  __syncthreads();
  for (int j = threadIdx.x; j < range; j += blockDim.x)
    if (hist[j]) atomicAdd(&col_count[lo + j], hist[j]);
}

// Two exclusive scans of one count array in one pass over it (out0 of in[i],
// out1 of ceil(in[i]/chunk)), across many workgroups: per-block sums, a
// scan of the block sums, then per-block local scans with the block offsets.
// n is read from n_dev when given; out0[n], out1[n] and totals[slot0/1]
// receive the two totals.
constexpr int kScanBlock = 1024;  // elements per workgroup (256 threads x 4)

__device__ __forceinline__ int block_excl_scan256(int v, int* sh, int& total) {
  const int t = threadIdx.x;
  sh[t] = v;
  __syncthreads();
  for (int o = 1; o < 256; o <<= 1) {
    const int u = (t >= o) ? sh[t - o] : 0;
    __syncthreads();
    sh[t] += u;
    __syncthreads();
  }
  total = sh[255];
  const int r = sh[t] - v;
  __syncthreads();
  return r;
}

// the same two scans in one workgroup of 1024 threads (a contiguous run of at
// most 8 counts per thread), for counts that one workgroup scans faster than
// three launches
constexpr int kScanOneMax = 8192;
__global__ __launch_bounds__(1024) void k_scan2_one(int n, const int* __restrict__ n_dev, const int* __restrict__ in,
                                                    int chunk, int* __restrict__ out0, int* __restrict__ out1,
                                                    int* __restrict__ totals, int slot0, int slot1) {
  __shared__ int sw[32];
  if (n_dev) n = *n_dev;
  const int t = threadIdx.x, T = blockDim.x;
  const int per = (n + T - 1) / T;  // <= 8: n <= kScanOneMax
  const int b = min(n, t * per), e = min(n, b + per);
  int v[8];
  int la = 0, lc = 0;
#pragma unroll
  for (int k = 0; k < 8; k++) v[k] = (b + k < e) ? in[b + k] : 0;
#pragma unroll
  for (int k = 0; k < 8; k++) {
    la += v[k];
    lc += (v[k] + chunk - 1) / chunk;
  }
  int ra = la, rc = lc, ta, tc;
  block_excl_scan2(ra, rc, ta, tc, sw);
#pragma unroll
  for (int k = 0; k < 8; k++) {
    if (b + k < e) {
      out0[b + k] = ra;
      out1[b + k] = rc;
      ra += v[k];
      rc += (v[k] + chunk - 1) / chunk;
    }
  }
  if (t == 0) {
    out0[n] = ta;
    out1[n] = tc;
    totals[slot0] = ta;
    totals[slot1] = tc;
  }
}

__global__ __launch_bounds__(256) void k_scan2_sums(int n, const int* __restrict__ n_dev, const int* __restrict__ in,
                                                    int chunk, int* __restrict__ bsum) {
  __shared__ int sh[256];
  if (n_dev) n = *n_dev;
  const int base = blockIdx.x * kScanBlock + threadIdx.x * 4;
  int a = 0, c = 0;
  for (int k = 0; k < 4; k++) {
    const int i = base + k;
    if (i < n) {
      a += in[i];
      c += (in[i] + chunk - 1) / chunk;
    }
  }
  int ta, tc;
  block_excl_scan256(a, sh, ta);
  block_excl_scan256(c, sh, tc);
  if (threadIdx.x == 0) {
    bsum[2 * blockIdx.x] = ta;
    bsum[2 * blockIdx.x + 1] = tc;
  }
}

__global__ __launch_bounds__(1024) void k_scan2_blocks(int nb, int n, const int* __restrict__ n_dev,
                                                       int* __restrict__ bsum, int* __restrict__ out0,
                                                       int* __restrict__ out1, int* __restrict__ totals, int slot0,
                                                       int slot1) {
  __shared__ int sa[1024], sc[1024];
  if (n_dev) n = *n_dev;
  const int t = threadIdx.x, T = blockDim.x;
  const int per = (nb + T - 1) / T;
  const int b = t * per, e = min(nb, b + per);
  int la = 0, lc = 0;
  for (int i = b; i < e; i++) {
    la += bsum[2 * i];
    lc += bsum[2 * i + 1];
  }
  sa[t] = la;
  sc[t] = lc;
  __syncthreads();
  for (int o = 1; o < T; o <<= 1) {
    const int va = (t >= o) ? sa[t - o] : 0, vc = (t >= o) ? sc[t - o] : 0;
    __syncthreads();
    sa[t] += va;
    sc[t] += vc;
    __syncthreads();
  }
  int ra = sa[t] - la, rc = sc[t] - lc;
  for (int i = b; i < e; i++) {
    const int xa = bsum[2 * i], xc = bsum[2 * i + 1];
    bsum[2 * i] = ra;
    bsum[2 * i + 1] = rc;
    ra += xa;
    rc += xc;
  }
  if (t == T - 1) {
    out0[n] = sa[t];
    out1[n] = sc[t];
    totals[slot0] = sa[t];
    totals[slot1] = sc[t];
  }
}

__global__ __launch_bounds__(256) void k_scan2_apply(int n, const int* __restrict__ n_dev, const int* __restrict__ in,
                                                     int chunk, const int* __restrict__ bsum, int* __restrict__ out0,
                                                     int* __restrict__ out1) {
  __shared__ int sh[256];
  if (n_dev) n = *n_dev;
  if (blockIdx.x * kScanBlock >= n) return;  // uniform per workgroup
  const int base = blockIdx.x * kScanBlock + threadIdx.x * 4;
  int va[4], vc[4], a = 0, c = 0;
  for (int k = 0; k < 4; k++) {
    const int i = base + k;
    va[k] = (i < n) ? in[i] : 0;
    vc[k] = (i < n) ? (va[k] + chunk - 1) / chunk : 0;
    a += va[k];
    c += vc[k];
  }
  int ta, tc;
  int ra = block_excl_scan256(a, sh, ta) + bsum[2 * blockIdx.x];
  int rc = block_excl_scan256(c, sh, tc) + bsum[2 * blockIdx.x + 1];
  for (int k = 0; k < 4; k++) {
    const int i = base + k;
    if (i < n) {
      out0[i] = ra;
      out1[i] = rc;
    }
    ra += va[k];
    rc += vc[k];
  }
}

__global__ void k_fill_int(int* p, int n, int v) {
  int i = blockIdx.x * blockDim.x + threadIdx.x;
  if (i < n) p[i] = v;
}

// three fills in one launch (grid-stride over the longest)
__global__ void k_fill_int3(int* p0, int n0, int v0, int* p1, int n1, int v1, int* p2, int n2, int v2) {
  const int i0 = blockIdx.x * blockDim.x + threadIdx.x, st = gridDim.x * blockDim.x;
  for (int i = i0; i < n0; i += st) p0[i] = v0;
  for (int i = i0; i < n1; i += st) p1[i] = v1;
  for (int i = i0; i < n2; i += st) p2[i] = v2;
}

// flag[0] = 1 when some atom lies farther than sqrt(lim2) from its position
// at the last list build (a NaN coordinate counts as moved); flag[0] must be
// zero before the launch
__global__ __launch_bounds__(256) void k_moved(int n, const double* __restrict__ pos,
                                               const double* __restrict__ ref, double lim2, int* __restrict__ flag) {
  const int i = blockIdx.x * blockDim.x + threadIdx.x;
  bool m = false;
  if (i < n) {
    const double dx = pos[3 * (size_t)i] - ref[3 * (size_t)i];
    const double dy = pos[3 * (size_t)i + 1] - ref[3 * (size_t)i + 1];
    const double dz = pos[3 * (size_t)i + 2] - ref[3 * (size_t)i + 2];
    m = !(dx * dx + dy * dy + dz * dz <= lim2);
  }
  if (__ballot(m) != 0ull && (threadIdx.x & (WAVE - 1)) == 0) atomicOr(flag, 1);
}

__global__ void k_fill_double(double* p, size_t n) {
  size_t i = (size_t)blockIdx.x * blockDim.x + threadIdx.x;
  size_t stride = (size_t)gridDim.x * blockDim.x;
  for (; i < n; i += stride) p[i] = 0.0;
}

// atoms listed by column (any order inside a column; k_rank orders them);
// slots within a column are reserved once per (workgroup, column) as in
// k_bin, over the same atom order
__global__ __launch_bounds__(256) void k_scatter(int n, const int4* __restrict__ info,
                                                 const int* __restrict__ col_start, int* __restrict__ fill,
                                                 int* __restrict__ tmp, const int* __restrict__ order = nullptr,
                                                 int norder = 0) {
  __shared__ int sh[512];
  __shared__ int hist[kBinLds];
  const int s = blockIdx.x * blockDim.x + threadIdx.x;
  const int i = order ? (s < norder ? order[s] : -1) : (s < n ? s : -1);
  const int col = (i >= 0) ? info[i].w : -1;
  int lo, hi;
  block_col_range(col, sh, lo, hi);
  if (hi < 0) return;  // uniform
  if (hi - lo >= kBinLds) {
    if (col >= 0) {
      const int k = atomicAdd(&fill[col], 1);
      tmp[col_start[col] + k] = i;
    }
    return;
  }
  const int range = hi - lo + 1;
  for (int j = threadIdx.x; j < range; j += blockDim.x) hist[j] = 0;
  __syncthreads();
  const int local = (col >= 0) ? atomicAdd(&hist[col - lo], 1) : 0;
  __syncthreads();
  // hist[j] becomes this workgroup's first slot in column lo + j
  for (int j = threadIdx.x; j < range; j += blockDim.x)
    if (hist[j]) hist[j] = atomicAdd(&fill[lo + j], hist[j]);
  __syncthreads();
  if (col >= 0) tmp[col_start[col] + hist[col - lo] + local] = i;
}

// Slots of one column and the bounds of its clusters; one workgroup per column.
// (1) Rank of each atom inside its column by (key, index): a deterministic
//     order.  When the column fits in LDS its atoms are grouped into buckets
//     of key (a counting pass, a scan, a scatter), and each atom is ranked
//     against its own bucket only; larger columns count over the whole
//     column in global memory.
// (2) Per cluster (8 lanes, one per slot): FP64 origin (bounding-box centre
//     of the wrapped positions at the rebuild), FP32 bounding box, range of
//     fractional c (widened outwards so that the FP32 range contains the
//     FP64 keys) and system.  Padding slots come last in a cluster, so its
//     first slot always holds an atom.
// With pk, also what k_pack_bbox would write for these positions.
// Columns hold ~8 atoms per column width of height (a few hundred in cells up
// to ~30 column widths tall); kRankLds bounds the LDS of a workgroup to
// ~21 KB, so that 7 columns share a CU.
constexpr int kRankLds = 1024;
constexpr int kRankBuckets = 256;  // the workgroup size: one scan
__device__ __forceinline__ int rank_bucket(double k) {
  return max(0, min((int)(k * kRankBuckets), kRankBuckets - 1));
}
__global__ __launch_bounds__(256) void k_rank(int ncol, const double* __restrict__ key,
                                              const int* __restrict__ col_start, const int* __restrict__ col_cl,
                                              const int* __restrict__ tmp, int* __restrict__ slot_loc,
                                              const int4* __restrict__ info, const double* __restrict__ pos, int off,
                                              const SysArg* __restrict__ sys, const int* __restrict__ atom_sys,
                                              double4* __restrict__ origin, float4* __restrict__ bb0,
                                              float2* __restrict__ kr, int* __restrict__ cl_sys,
                                              int* __restrict__ cl_col, float4* __restrict__ pk,
                                              float4* __restrict__ pbb, double pscale) {
  __shared__ double sk[kRankLds];
  __shared__ int si[kRankLds];
  __shared__ int ss[kRankLds];
  __shared__ int sbk[kRankLds];  // local indices grouped by key bucket
  __shared__ int bcnt[kRankBuckets], bstart[kRankBuckets + 1];
  __shared__ int sw[32];
  const int col = blockIdx.x;
  if (col >= ncol) return;
  const int b = col_start[col], n = col_start[col + 1] - b;
  const int c0 = col_cl[col], ncl = col_cl[col + 1] - c0;
  const int base = c0 * CL;
  for (int q = threadIdx.x; q < ncl; q += blockDim.x) cl_col[c0 + q] = col;
  // padding slots at the end of the column's last cluster
  for (int t = n + (int)threadIdx.x; t < ncl * CL; t += blockDim.x) slot_loc[base + t] = -1;
  const bool lds = n <= kRankLds;
  if (lds) {
    // the keys lie in [0, 1): bucket floor(key * kRankBuckets) keeps their
    // order, so an atom's rank is the atoms of the lower buckets plus those of
    // its own bucket below it by (key, index)
    for (int q = threadIdx.x; q < kRankBuckets; q += blockDim.x) bcnt[q] = 0;
    for (int t = threadIdx.x; t < n; t += blockDim.x) {
      const int o = tmp[b + t];
      si[t] = o;
      sk[t] = key[o];
    }
    __syncthreads();
    for (int t = threadIdx.x; t < n; t += blockDim.x)
      atomicAdd(&bcnt[rank_bucket(sk[t])], 1);
    __syncthreads();
    {
      int v = threadIdx.x < kRankBuckets ? bcnt[threadIdx.x] : 0, dummy = 0, tot, tot2;
      block_excl_scan2(v, dummy, tot, tot2, sw);
      if (threadIdx.x < kRankBuckets) {
        bstart[threadIdx.x] = v;
        bcnt[threadIdx.x] = v;  // fill pointer of the bucket
      }
      if (threadIdx.x == 0) bstart[kRankBuckets] = tot;
    }
    __syncthreads();
    for (int t = threadIdx.x; t < n; t += blockDim.x)
      sbk[atomicAdd(&bcnt[rank_bucket(sk[t])], 1)] = t;
    __syncthreads();
    for (int t = threadIdx.x; t < n; t += blockDim.x) {
      const double ka = sk[t];
      const int ia = si[t];
      const int bk = rank_bucket(ka);
      const int p1 = bstart[bk + 1];
      int rank = bstart[bk];
      for (int p = bstart[bk]; p < p1; p++) {
        const int u = sbk[p];
        const double kb = sk[u];
        rank += (kb < ka) || (kb == ka && si[u] < ia);
      }
      slot_loc[base + rank] = ia;
      ss[rank] = ia;
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
  __syncthreads();
  // (2) the column's clusters; the loop bound is uniform over the workgroup.
  // With pk, also what k_pack_bbox would write for these positions: the
  // packed slots (pk) and the boxes of the packed clusters (pbb).
  for (int t0 = 0; t0 < ncl * CL; t0 += blockDim.x) {
    const int t = t0 + (int)threadIdx.x;
    int a = -1;
    if (t < n) a = lds ? ss[t] : slot_loc[base + t];
    double lo[3] = {DBL_MAX, DBL_MAX, DBL_MAX}, hi[3] = {-DBL_MAX, -DBL_MAX, -DBL_MAX};
    double w[3] = {0.0, 0.0, 0.0};
    double klo = DBL_MAX, khi = -DBL_MAX;
    int sy = 0;
    if (a >= 0) {
      sy = nsys_of(atom_sys, a);
      const SysArg& S = sys[sy];
      const int4 m = info[a];
      const double* x = pos + 3 * (size_t)(off + a);
      for (int d = 0; d < 3; d++) {
        w[d] = x[d] - (m.x * S.a[0 * 3 + d] + m.y * S.a[1 * 3 + d] + m.z * S.a[2 * 3 + d]);
        lo[d] = w[d];
        hi[d] = w[d];
      }
      klo = khi = key[a];
    }
    // the 8 lanes of a cluster are aligned in the wave
    for (int o = 1; o < CL; o <<= 1) {
      for (int d = 0; d < 3; d++) {
        lo[d] = fmin(lo[d], __shfl_xor(lo[d], o, 64));
        hi[d] = fmax(hi[d], __shfl_xor(hi[d], o, 64));
      }
      klo = fmin(klo, __shfl_xor(klo, o, 64));
      khi = fmax(khi, __shfl_xor(khi, o, 64));
    }
    const double cx = 0.5 * (lo[0] + hi[0]), cy = 0.5 * (lo[1] + hi[1]), cz = 0.5 * (lo[2] + hi[2]);
    const int c = c0 + t / CL;
    if ((t & (CL - 1)) == 0 && t < ncl * CL) {
      if (cl_sys) cl_sys[c] = sy;
      origin[c] = make_double4(cx, cy, cz, 0.0);
      bb0[2 * c] = make_float4((float)cx, (float)cy, (float)cz, 0.f);
      bb0[2 * c + 1] = make_float4((float)(0.5 * (hi[0] - lo[0])), (float)(0.5 * (hi[1] - lo[1])),
                                   (float)(0.5 * (hi[2] - lo[2])), 0.f);
      kr[c] = make_float2((float)klo - 1.0e-6f, (float)khi + 1.0e-6f);
    }
    if (pk) {
      const int sl = base + t;
      float4 q;
      float ql[3], qh[3];
      if (a < 0) {
        q = make_float4(1.0e5f + 1.0e3f * (float)(sl & 1023), 1.0e5f + 1.0e3f * (float)((sl >> 10) & 1023),
                        1.0e5f + 1.0e3f * (float)(sl >> 20), __int_as_float(-1));
        ql[0] = ql[1] = ql[2] = FLT_MAX;
        qh[0] = qh[1] = qh[2] = -FLT_MAX;
      } else {
        q = make_float4((float)((w[0] - cx) * pscale), (float)((w[1] - cy) * pscale), (float)((w[2] - cz) * pscale),
                        __int_as_float(off + a));
        ql[0] = qh[0] = q.x;
        ql[1] = qh[1] = q.y;
        ql[2] = qh[2] = q.z;
      }
      if (t < ncl * CL) pk[sl] = q;
      for (int o = 1; o < CL; o <<= 1) {
        for (int d = 0; d < 3; d++) {
          ql[d] = fminf(ql[d], __shfl_xor(ql[d], o, 64));
          qh[d] = fmaxf(qh[d], __shfl_xor(qh[d], o, 64));
        }
      }
      if ((t & (CL - 1)) == 0 && t < ncl * CL) {
        pbb[2 * c] = make_float4(0.5f * (ql[0] + qh[0]), 0.5f * (ql[1] + qh[1]), 0.5f * (ql[2] + qh[2]), 0.f);
        pbb[2 * c + 1] = make_float4(0.5f * (qh[0] - ql[0]), 0.5f * (qh[1] - ql[1]), 0.5f * (qh[2] - ql[2]), 0.f);
      }
    }
  }
}

// absolute atom index of every slot of a group (two groups: an atom listed in
// both is not paired with itself); padding slots get distinct negative tags
__global__ void k_abs_slot(int nslot, const int* __restrict__ ncl_dev, const int* __restrict__ slot_loc, int off,
                           const long long* __restrict__ absidx, int* __restrict__ abs_slot) {
  const int s = blockIdx.x * blockDim.x + threadIdx.x;
  if (ncl_dev) nslot = *ncl_dev * CL;
  if (s >= nslot) return;
  const int a = slot_loc[s];
  abs_slot[s] = (a >= 0) ? (int)absidx[off + a] : -1 - s;
}

// candidates kept per i cluster between the count and fill passes of the
// list search (j cluster in bits 0-25, image code (mx+1)+3(my+1)+9(mz+1) above)
constexpr int kEll = 384;
static_assert(kEll >= 2 * WAVE, "k_list_stream reads two candidates per lane ahead");
constexpr unsigned kEllCj = 0x3FFFFFFu;

// list entry of i cluster ci for j cluster cj seen through image (mx,my,mz)
__device__ __forceinline__ Entry make_entry(int mode, int ci, int cj, int mx, int my, int mz, const SysArg& SJ,
                                            const double4* __restrict__ orgJ, const double4& oi, double pscale) {
  const double4 oj = orgJ[cj];
  const double Tx = mx * SJ.a[0] + my * SJ.a[3] + mz * SJ.a[6];
  const double Ty = mx * SJ.a[1] + my * SJ.a[4] + mz * SJ.a[7];
  const double Tz = mx * SJ.a[2] + my * SJ.a[5] + mz * SJ.a[8];
  const bool zero = (mx == 0 && my == 0 && mz == 0);
  const bool diag = mode == 0 && cj == ci && zero;
  const bool positive = (mx > 0) || (mx == 0 && (my > 0 || (my == 0 && mz > 0)));
  const bool counted = mode == 1 || (mode == 0 && (cj > ci || (cj == ci && positive)));
  Entry en;
  en.code = cj | (((mx + 1) + 3 * (my + 1) + 9 * (mz + 1)) << kImgShift) | (counted ? kCountBit : 0) |
            (diag ? (int)0x80000000u : 0);
  en.dx = (float)((oj.x + Tx - oi.x) * pscale);
  en.dy = (float)((oj.y + Ty - oi.y) * pscale);
  en.dz = (float)((oj.z + Tz - oi.z) * pscale);
  return en;
}

// entry offsets of a kept list for a new cell: the j cluster's lattice
// translation (mx, my, mz) now uses the lattice vectors a (rows), as the
// packed positions do; one wave per i cluster, lanes over its entries
__global__ __launch_bounds__(256) void k_reimage(int ncl, const int* __restrict__ estart,
                                                 const int* __restrict__ ecount, int ent_cap,
                                                 const double4* __restrict__ orgI, const double4* __restrict__ orgJ,
                                                 BoxArg box, double pscale, Entry* __restrict__ ents) {
  const int lane = threadIdx.x & (WAVE - 1);
  const int ci = blockIdx.x * (blockDim.x / WAVE) + threadIdx.x / WAVE;
  if (ci >= ncl) return;
  const int e0 = estart[ci], n = ecount[ci];
  const double4 oi = orgI[ci];
  for (int k = lane; k < n; k += WAVE) {
    const int e = e0 + k;
    if (e >= ent_cap) break;
    Entry en = ents[e];
    const int cj = en.code & kCjMask;
    const int im = (en.code >> kImgShift) & 31;
    const int mx = im % 3 - 1, my = (im / 3) % 3 - 1, mz = im / 9 - 1;
    const double4 oj = orgJ[cj];
    en.dx = (float)((oj.x + (mx * box.a[0] + my * box.a[3] + mz * box.a[6]) - oi.x) * pscale);
    en.dy = (float)((oj.y + (mx * box.a[1] + my * box.a[4] + mz * box.a[7]) - oi.y) * pscale);
    en.dz = (float)((oj.z + (mx * box.a[2] + my * box.a[5] + mz * box.a[8]) - oi.z) * pscale);
    ents[e] = en;
  }
}

// one value written to device memory in stream order: small descriptors go
// to the device as kernel arguments, so the host never waits for a staging copy
template <class T>
__global__ void k_store(T v, T* __restrict__ dst) {
  if (blockIdx.x == 0 && threadIdx.x == 0) *dst = v;
}

struct ListArgs {
  int nclI;
  const int* nclI_dev;  // device count of i clusters (null: nclI)
  int mode;             // 0 one group, 1 two groups (counting side), 2 two groups (other side)
  int ent_cap;          // capacity of the entry array (fill pass)
  float rlist2;
  double pscale;        // scale of the entry offsets (pack_scale)
};

// Cluster-pair search: one wave per i cluster.  FILL=0 counts and keeps the
// first kEll passing candidates of the cluster (j cluster | image code) in
// `ell`; FILL=1 writes the entries, streaming them from `ell` when the
// cluster's count fits there and searching again otherwise; FILL=2 does only
// the second part, after k_list_stream has done the first.
// Candidates are the clusters of the (column, image) window around the i
// cluster whose range of fractional c overlaps the i cluster's range widened
// by the list radius: a pair closer than r differs by at most r*|c*| in
// fractional c, so the window is exact, and the clusters of a column are
// sorted along c, so each window is one binary-searched run of clusters.
// Each column lane writes its runs into the wave's candidate list in LDS
// (j cluster | image code, in column order), kCandCap at a time, and the
// lanes then test the listed candidates, two per lane and trip.
constexpr int kCandCap = 512;
template <int FILL>
__global__ __launch_bounds__(256) void k_list(ListArgs L, const int* __restrict__ cl_colI,
                                              const int* __restrict__ cl_sysI, const SysArg* __restrict__ sysI,
                                              const SysArg* __restrict__ sysJ, const float4* __restrict__ bbI,
                                              const float4* __restrict__ bbJ, const float2* __restrict__ krI,
                                              const float2* __restrict__ krJ, const int* __restrict__ col_clJ,
                                              const double4* __restrict__ orgI, const double4* __restrict__ orgJ,
                                              int* __restrict__ count, const int* __restrict__ start,
                                              Entry* __restrict__ out, unsigned* __restrict__ ell) {
  __shared__ unsigned s_cand[4][kCandCap];
  const int lane = threadIdx.x & 63;
  const int wv = threadIdx.x / WAVE;
  const int nclI = L.nclI_dev ? *L.nclI_dev : L.nclI;
  // FILL=0/1: cluster ci = the wave.  FILL=2: the wave's 64 clusters, of which
  // it searches those whose candidates did not all fit in ell (the others are
  // k_list_stream's), in order.  No workgroup barrier below.
  int ci = blockIdx.x * (blockDim.x / WAVE) + wv;
  unsigned long long pending = 0ull;
  if (FILL == 2) {
    const int c0 = ci * WAVE;
    pending = __ballot(c0 + lane < nclI && count[c0 + lane] > kEll);
    ci = c0;
  } else if (ci >= nclI) {
    return;
  }
  const int cfirst = ci;
  while (FILL != 2 || pending) {
  if (FILL == 2) {
    ci = cfirst + __builtin_ctzll(pending);
    pending &= pending - 1ull;
  }
  const int s = nsys_of(cl_sysI, ci);
  const SysArg& SI = sysI[s];
  const SysArg& SJ = sysJ[s];
  const int colI = cl_colI[ci] - SI.colbase;
  const int cxi = colI % SI.ncx, cyi = colI / SI.ncx;
  const float4 ic = bbI[2 * ci], ih = bbI[2 * ci + 1];
  const float2 ki = krI[ci];
  const int x0 = (int)floorf(((float)cxi / SI.ncx - SJ.rb[0]) * SJ.ncx) - 1;
  const int x1 = (int)floorf(((float)(cxi + 1) / SI.ncx + SJ.rb[0]) * SJ.ncx) + 1;
  const int y0 = (int)floorf(((float)cyi / SI.ncy - SJ.rb[1]) * SJ.ncy) - 1;
  const int y1 = (int)floorf(((float)(cyi + 1) / SI.ncy + SJ.rb[1]) * SJ.ncy) + 1;
  // images -1, 0, +1 along a and b cover every candidate when the cell is
  // wider than 2*rlist; each unwrapped column is a distinct (column, image)
  const int xa = max(x0, -SJ.ncx), xb = min(x1, 2 * SJ.ncx - 1);
  const int ya = max(y0, -SJ.ncy), yb = min(y1, 2 * SJ.ncy - 1);
  const int nx = xb - xa + 1, ny = yb - ya + 1;
  const int ncolw = nx * ny;
  const float zlo = ki.x - SJ.rb[2] - 1.0e-6f, zhi = ki.y + SJ.rb[2] + 1.0e-6f;
  const double4 oi = orgI[ci];
  int n = 0;
  const int base = FILL ? start[ci] : 0;
  if (FILL == 1) {
    const int nc = count[ci];
    if (nc <= kEll) {
      for (int e = lane; e < nc; e += WAVE) {
        const unsigned pk = ell[(size_t)ci * kEll + e];
        const int cj = (int)(pk & kEllCj), im = (int)(pk >> 26);
        if (base + e < L.ent_cap)
          out[base + e] = make_entry(L.mode, ci, cj, im % 3 - 1, (im / 3) % 3 - 1, im / 9 - 1, SJ, orgJ, oi, L.pscale);
      }
      return;
    }
  }
  // one lane per window column; it covers the images -1, 0, +1 along c, of
  // which only those whose shifted window meets [0, 1) are searched
  for (int t0 = 0; t0 < ncolw; t0 += WAVE) {
    const int t = t0 + lane;
    int f[3] = {0, 0, 0}, cnt[3] = {0, 0, 0}, img = 0;
    if (t < ncolw) {
      const int cxu = xa + t % nx, cyu = ya + t / nx;
      const int mx = floordiv(cxu, SJ.ncx), cxw = cxu - mx * SJ.ncx;
      const int my = floordiv(cyu, SJ.ncy), cyw = cyu - my * SJ.ncy;
      const int col = SJ.colbase + cxw + SJ.ncx * cyw;
      const int c0 = col_clJ[col], c1 = col_clJ[col + 1];
      img = (mx + 1) | ((my + 1) << 2);
#pragma unroll
      for (int m = 0; m < 3; m++) {
        const float lo = zlo - (float)(m - 1), hi = zhi - (float)(m - 1);
        if (hi >= 0.f && lo < 1.f && c1 > c0) {
          int a = c0, b = c1;
          while (a < b) {  // first cluster whose upper key reaches lo
            const int q = (a + b) >> 1;
            if (krJ[q].y < lo) a = q + 1;
            else b = q;
          }
          f[m] = a;
          b = c1;
          while (a < b) {  // first cluster whose lower key exceeds hi
            const int q = (a + b) >> 1;
            if (krJ[q].x <= hi) a = q + 1;
            else b = q;
          }
          cnt[m] = a - f[m];
        }
      }
    }
    int total;
    const int off = wave_excl_scan(cnt[0] + cnt[1] + cnt[2], lane, total);
    for (int c0 = 0; c0 < total; c0 += kCandCap) {
      const int nc = min(total - c0, kCandCap);
      // this lane's candidates with list index in [c0, c0 + nc)
      int k = off;
#pragma unroll
      for (int m = 0; m < 3; m++) {
        const unsigned code = (unsigned)(((img & 3) + 3 * ((img >> 2) & 3) + 9 * m) << 26);
        const int lo = max(k, c0), hi = min(k + cnt[m], c0 + nc);
        for (int u = lo; u < hi; u++) s_cand[wv][u - c0] = (unsigned)(f[m] + (u - k)) | code;
        k += cnt[m];
      }
      __builtin_amdgcn_fence(__ATOMIC_SEQ_CST, "wavefront");
      __builtin_amdgcn_wave_barrier();
      // two candidates per lane and trip: their box loads are in flight together
      for (int u0 = 0; u0 < nc; u0 += 2 * WAVE) {
        int cj[2] = {0, 0}, mx[2] = {0, 0}, my[2] = {0, 0}, mz[2] = {0, 0};
        float4 jc[2], jh[2];
  #pragma unroll
        for (int h = 0; h < 2; h++) {
          const int u = u0 + h * WAVE + lane;
          jc[h] = jh[h] = make_float4(0.f, 0.f, 0.f, 0.f);
          if (u < nc) {
            const unsigned pk = s_cand[wv][u];
            const int im = (int)(pk >> 26);
            cj[h] = (int)(pk & kEllCj);
            mx[h] = im % 3 - 1;
            my[h] = (im / 3) % 3 - 1;
            mz[h] = im / 9 - 1;
            jc[h] = bbJ[2 * cj[h]];
            jh[h] = bbJ[2 * cj[h] + 1];
          }
        }
  #pragma unroll
        for (int h = 0; h < 2; h++) {
          const int u = u0 + h * WAVE + lane;
          bool pass = false;
          if (u < nc) {
            const float Tx = mx[h] * SJ.af[0] + my[h] * SJ.af[3] + mz[h] * SJ.af[6];
            const float Ty = mx[h] * SJ.af[1] + my[h] * SJ.af[4] + mz[h] * SJ.af[7];
            const float Tz = mx[h] * SJ.af[2] + my[h] * SJ.af[5] + mz[h] * SJ.af[8];
            const float gx = fmaxf(fabsf(jc[h].x + Tx - ic.x) - ih.x - jh[h].x, 0.f);
            const float gy = fmaxf(fabsf(jc[h].y + Ty - ic.y) - ih.y - jh[h].y, 0.f);
            const float gz = fmaxf(fabsf(jc[h].z + Tz - ic.z) - ih.z - jh[h].z, 0.f);
            pass = gx * gx + gy * gy + gz * gz < L.rlist2;
          }
          const unsigned long long bal = __ballot(pass);
          if (pass) {
            const int k = n + __popcll(bal & ((1ull << lane) - 1));
            if (FILL) {
              if (base + k < L.ent_cap)
                out[base + k] = make_entry(L.mode, ci, cj[h], mx[h], my[h], mz[h], SJ, orgJ, oi, L.pscale);
            } else if (k < kEll) {
              ell[(size_t)ci * kEll + k] =
                  (unsigned)cj[h] | ((unsigned)((mx[h] + 1) + 3 * (my[h] + 1) + 9 * (mz[h] + 1)) << 26);
            }
          }
          n += __popcll(bal);
        }
      }
      // the next chunk overwrites the candidate list read above
      __builtin_amdgcn_fence(__ATOMIC_SEQ_CST, "wavefront");
      __builtin_amdgcn_wave_barrier();
    }
  }
  if (!FILL && lane == 0) count[ci] = n;
  if (FILL != 2) break;
  }
}

// Fill pass of the clusters whose candidates all fit in `ell` (count <=
// kEll): their entries in the order the count pass kept them, one wave per
// i cluster, two entries per lane and trip.  The search kernel's registers
// are not held here, so more waves are resident to cover the two dependent
// loads (kept candidate, then the j cluster's origin) of each entry.  Used
// with k_list<2> where the fill covers many clusters (the batch); a single
// system's list is filled by k_list<1> in one launch.
__global__ __launch_bounds__(256) void k_list_stream(ListArgs L, const int* __restrict__ cl_sysI,
                                                     const SysArg* __restrict__ sysJ,
                                                     const double4* __restrict__ orgI,
                                                     const double4* __restrict__ orgJ, const int* __restrict__ count,
                                                     const int* __restrict__ start, Entry* __restrict__ out,
                                                     const unsigned* __restrict__ ell) {
  const int lane = threadIdx.x & 63;
  const int ci = blockIdx.x * (blockDim.x / WAVE) + threadIdx.x / WAVE;
  const int nclI = L.nclI_dev ? *L.nclI_dev : L.nclI;
  if (ci >= nclI) return;
  // everything that depends on ci only, and the first two kept candidates of
  // each lane (ell has kEll >= 2 * WAVE slots per cluster, written or not),
  // is loaded before the count decides anything
  const unsigned* kept = ell + (size_t)ci * kEll;
  const unsigned p0 = kept[lane], p1 = kept[WAVE + lane];
  const int nc = count[ci];
  const int base = start[ci];
  const int sy = nsys_of(cl_sysI, ci);
  const double4 oi = orgI[ci];
  if (nc > kEll) return;
  const SysArg& SJ = sysJ[sy];
  for (int e0 = 0; e0 < nc; e0 += 2 * WAVE) {
    const int ea = e0 + lane, eb = e0 + WAVE + lane;
    const unsigned pa = ea < nc ? (e0 == 0 ? p0 : kept[ea]) : 0u, pb = eb < nc ? (e0 == 0 ? p1 : kept[eb]) : 0u;
    const int ja = (int)(pa & kEllCj), ia = (int)(pa >> 26);
    const int jb = (int)(pb & kEllCj), ib = (int)(pb >> 26);
    if (ea < nc && base + ea < L.ent_cap)
      out[base + ea] = make_entry(L.mode, ci, ja, ia % 3 - 1, (ia / 3) % 3 - 1, ia / 9 - 1, SJ, orgJ, oi, L.pscale);
    if (eb < nc && base + eb < L.ent_cap)
      out[base + eb] = make_entry(L.mode, ci, jb, ib % 3 - 1, (ib / 3) % 3 - 1, ib / 9 - 1, SJ, orgJ, oi, L.pscale);
  }
}

// work items {i cluster, first entry, end entry} of `chunk` entries each;
// writes past items_cap are dropped
__global__ void k_items(int ncl, const int* __restrict__ ncl_dev, const int* __restrict__ count,
                        const int* __restrict__ estart, const int* __restrict__ item_start, int4* __restrict__ items,
                        int chunk, int items_cap) {
  int c = blockIdx.x * blockDim.x + threadIdx.x;
  if (ncl_dev) ncl = *ncl_dev;
  if (c >= ncl) return;
  const int n = count[c], e = estart[c];
  int nch = (n + chunk - 1) / chunk;
  for (int k = 0; k < nch; k++) {
    const int it = item_start[c] + k;
    if (it < items_cap) items[it] = make_int4(c, e + k * chunk, e + min(n, (k + 1) * chunk), 0);
  }
}

// Balanced static schedule of the pair kernel: the items split into nwaves
// contiguous ranges of about equal estimated cost, an item costing its
// entries plus kItemCost (the staging and sums each item pays).  Item i
// (cost prefix P_i, total C) goes to wave floor(P_i * nwaves / C); one
// thread per item writes the first item of each wave it starts.
// wave_start has nwaves + 1 entries; nitems = *nitems_dev when given.
constexpr int kItemCost = 32;
__global__ void k_wave_sched(int nitems, const int* __restrict__ nitems_dev, int items_cap,
                             const int4* __restrict__ items, int nwaves, int* __restrict__ wave_start) {
  const int i = blockIdx.x * blockDim.x + threadIdx.x;
  if (nitems_dev) nitems = min(*nitems_dev, items_cap);
  if (nitems == 0) {
    for (int w = i; w <= nwaves; w += gridDim.x * blockDim.x) wave_start[w] = 0;
    return;
  }
  if (i >= nitems) return;
  const long long total = (long long)items[nitems - 1].z + (long long)kItemCost * nitems;
  auto wave_of = [&](int q) {
    const long long pre = (long long)items[q].y + (long long)kItemCost * q;
    return (int)min((long long)nwaves - 1, pre * nwaves / total);
  };
  const int wi = wave_of(i);
  const int wp = i == 0 ? -1 : wave_of(i - 1);
  for (int w = wp + 1; w <= wi; w++) wave_start[w] = i;
  if (i == nitems - 1)
    for (int w = wi + 1; w <= nwaves; w++) wave_start[w] = nitems;
}

// ------------------------------------------------------------------------
// per-step kernels
// ------------------------------------------------------------------------

// k_pack and k_bbox in one pass: one thread per slot, the 8 slots of a
// cluster are 8 adjacent lanes, which reduce the cluster's box by shuffles.
// Positions relative to the cluster origin, times pscale (pack_scale).
__global__ __launch_bounds__(256) void k_pack_bbox(int nslot, const int* __restrict__ ncl_dev,
                                                   const int* __restrict__ slot_loc, const int4* __restrict__ info,
                                                   const double* __restrict__ pos, int off,
                                                   const SysArg* __restrict__ sys, const int* __restrict__ atom_sys,
                                                   const double4* __restrict__ origin, float4* __restrict__ out,
                                                   float4* __restrict__ bb, double pscale) {
  const int s = blockIdx.x * blockDim.x + threadIdx.x;
  if (ncl_dev) nslot = *ncl_dev * CL;
  // whole groups of 8 lanes stay together (nslot is a multiple of 8)
  if (s >= nslot) return;
  const int a = slot_loc[s];
  float4 q;
  float lo[3], hi[3];
  if (a < 0) {
    q = make_float4(1.0e5f + 1.0e3f * (float)(s & 1023), 1.0e5f + 1.0e3f * (float)((s >> 10) & 1023),
                    1.0e5f + 1.0e3f * (float)(s >> 20), __int_as_float(-1));
    lo[0] = lo[1] = lo[2] = FLT_MAX;
    hi[0] = hi[1] = hi[2] = -FLT_MAX;
  } else {
    const SysArg& S = sys[nsys_of(atom_sys, a)];
    const int4 m = info[a];
    const double4 o = origin[s / CL];
    const double* x = pos + 3 * (size_t)(off + a);
    double w[3];
    for (int d = 0; d < 3; d++) w[d] = x[d] - (m.x * S.a[0 * 3 + d] + m.y * S.a[1 * 3 + d] + m.z * S.a[2 * 3 + d]);
    q = make_float4((float)((w[0] - o.x) * pscale), (float)((w[1] - o.y) * pscale), (float)((w[2] - o.z) * pscale),
                    __int_as_float(off + a));
    lo[0] = hi[0] = q.x;
    lo[1] = hi[1] = q.y;
    lo[2] = hi[2] = q.z;
  }
  out[s] = q;
  for (int o = 1; o < CL; o <<= 1) {
    for (int d = 0; d < 3; d++) {
      lo[d] = fminf(lo[d], __shfl_xor(lo[d], o, 64));
      hi[d] = fmaxf(hi[d], __shfl_xor(hi[d], o, 64));
    }
  }
  if ((s & (CL - 1)) == 0) {
    const int c = s / CL;
    bb[2 * c] = make_float4(0.5f * (lo[0] + hi[0]), 0.5f * (lo[1] + hi[1]), 0.5f * (lo[2] + hi[2]), 0.f);
    bb[2 * c + 1] = make_float4(0.5f * (hi[0] - lo[0]), 0.5f * (hi[1] - lo[1]), 0.5f * (hi[2] - lo[2]), 0.f);
  }
}

// The FP64 helpers below take the reference's rounding: no fused
// multiply-add contraction (the reference's host build has none).
#pragma clang fp contract(off)

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

#pragma clang fp contract(fast)

// ------------------------------------------------------------------------
// cross-lane sums on the VALU (no LDS): DPP rotations inside 16-lane rows,
// v_permlane16_swap / v_permlane32_swap across rows
// ------------------------------------------------------------------------
template <int CTRL>
__device__ __forceinline__ float dpp_f(float v) {
  return __int_as_float(__builtin_amdgcn_update_dpp(0, __float_as_int(v), CTRL, 0xF, 0xF, false));
}
template <int CTRL, int ROWS>
__device__ __forceinline__ int dpp_i(int v) {
  return __builtin_amdgcn_update_dpp(0, v, CTRL, ROWS, 0xF, false);
}
constexpr int kRor1 = 0x121, kRor2 = 0x122, kRor4 = 0x124, kRor8 = 0x128;
constexpr int kShr1 = 0x111, kShr2 = 0x112, kShr4 = 0x114, kShr8 = 0x118;
constexpr int kBcast15 = 0x142, kBcast31 = 0x143;

// v[l] + v[l ^ 16]: odd 16-lane rows of one copy swap with even rows of the other
__device__ __forceinline__ float xor16_sum(float v) {
  const auto p = __builtin_amdgcn_permlane16_swap(__float_as_uint(v), __float_as_uint(v), false, false);
  return __uint_as_float(p[0]) + __uint_as_float(p[1]);
}
// v[l] + v[l ^ 32]
__device__ __forceinline__ float xor32_sum(float v) {
  const auto p = __builtin_amdgcn_permlane32_swap(__float_as_uint(v), __float_as_uint(v), false, false);
  return __uint_as_float(p[0]) + __uint_as_float(p[1]);
}
// the same moves for a double (both halves take the same path)
template <int CTRL>
__device__ __forceinline__ double dpp_d(double v) {
  const long long b = __double_as_longlong(v);
  const int lo = __builtin_amdgcn_update_dpp(0, (int)(b & 0xffffffffll), CTRL, 0xF, 0xF, false);
  const int hi = __builtin_amdgcn_update_dpp(0, (int)(b >> 32), CTRL, 0xF, 0xF, false);
  return __longlong_as_double(((long long)hi << 32) | (unsigned int)lo);
}
__device__ __forceinline__ double xor16_sum_d(double v) {
  const long long b = __double_as_longlong(v);
  const auto pl = __builtin_amdgcn_permlane16_swap((unsigned)(b & 0xffffffffll), (unsigned)(b & 0xffffffffll), false,
                                                   false);
  const auto ph = __builtin_amdgcn_permlane16_swap((unsigned)(b >> 32), (unsigned)(b >> 32), false, false);
  const double x = __longlong_as_double(((long long)ph[0] << 32) | pl[0]);
  const double y = __longlong_as_double(((long long)ph[1] << 32) | pl[1]);
  return x + y;
}
__device__ __forceinline__ double xor32_sum_d(double v) {
  const long long b = __double_as_longlong(v);
  const auto pl = __builtin_amdgcn_permlane32_swap((unsigned)(b & 0xffffffffll), (unsigned)(b & 0xffffffffll), false,
                                                   false);
  const auto ph = __builtin_amdgcn_permlane32_swap((unsigned)(b >> 32), (unsigned)(b >> 32), false, false);
  const double x = __longlong_as_double(((long long)ph[0] << 32) | pl[0]);
  const double y = __longlong_as_double(((long long)ph[1] << 32) | pl[1]);
  return x + y;
}
// sum over the 64 lanes of a double, without LDS (the value of lane 0 is used)
__device__ __forceinline__ double wave_sum_d(double v) {
  v += dpp_d<kRor8>(v);
  v += dpp_d<kRor4>(v);
  v += dpp_d<kRor2>(v);
  v += dpp_d<kRor1>(v);
  return xor32_sum_d(xor16_sum_d(v));
}

// sum over the 8 lanes l, l^8, l^16, ..., l^56 (lane = ii + 8 jj: the sum over jj)
__device__ __forceinline__ float sum_over_jj(float v) {
  v += dpp_f<kRor8>(v);
  return xor32_sum(xor16_sum(v));
}
// sum over the 64 lanes, in every lane
__device__ __forceinline__ float wave_sum_f(float v) {
  v += dpp_f<kRor8>(v);
  v += dpp_f<kRor4>(v);
  v += dpp_f<kRor2>(v);
  v += dpp_f<kRor1>(v);
  return xor32_sum(xor16_sum(v));
}
// exclusive prefix sum over the 64 lanes; *total = sum of all lanes (uniform)
__device__ __forceinline__ int wave_excl_scan_dpp(int x, int& total) {
  int v = x;
  v += dpp_i<kShr1, 0xF>(v);
  v += dpp_i<kShr2, 0xF>(v);
  v += dpp_i<kShr4, 0xF>(v);
  v += dpp_i<kShr8, 0xF>(v);
  v += dpp_i<kBcast15, 0xA>(v);
  v += dpp_i<kBcast31, 0xC>(v);
  total = __builtin_amdgcn_readlane(v, 63);
  return v - x;
}

// number of set bits of m in the lanes below this one (v_mbcnt)
__device__ __forceinline__ int bits_below(unsigned long long m) {
  return (int)__builtin_amdgcn_mbcnt_hi((unsigned)(m >> 32), __builtin_amdgcn_mbcnt_lo((unsigned)m, 0u));
}

// ------------------------------------------------------------------------
// the pair kernel
// ------------------------------------------------------------------------

// FP32 switch value and derivative factor of fixedRational<2P> on r^2
// without its stretch: sv = 1/(1+x^P) and dfu = x^(P-1) sv^2 with x = r^2/r0^2.
// The stretched value is stretch*sv + shift and dfunc = -2P/r0^2 stretch dfu
// (SwitchingFunction.cpp:203-215); both factors are applied once per work
// item, in FP64 (PairArgs::sv_scale, sv_shift, df_scale, vir_scale).  The
// cluster positions of these switches are stored in units of R_0 (x = r^2).
template <int P>
__device__ __forceinline__ void sw_fixed(float invr0_2, float r2, float& sv, float& dfu) {
  const float x = r2 * invr0_2;
  float xp = 1.0f;
#pragma unroll
  for (int k = 1; k < P; k++) xp *= x;  // x^(P-1)
  const float res = __builtin_amdgcn_rcpf(fmaf(xp, x, 1.0f));
  sv = res;
  dfu = xp * res * res;
}

// x for a pair outside the cut: x^(P-1) stays finite and 1 + x^P overflows,
// so that sv = 1/(1+x^P) and dfu = x^(P-1) sv^2 are exactly 0 (x = 2^ceil(128/P))
template <int P>
__device__ __forceinline__ constexpr float sw_far() {
  return P == 1 ? __builtin_huge_valf()
                : (P == 2 ? 0x1p64f : (P == 3 ? 0x1p43f : (P == 4 ? 0x1p32f : (P == 5 ? 0x1p26f : 0x1p22f))));
}
static_assert(sw_far<6>() == 0x1p22f, "far values cover P <= 6");

template <int P>
__device__ __forceinline__ void sw_fixed_x(float x, float& sv, float& dfu) {
  float xp = 1.0f;
#pragma unroll
  for (int k = 1; k < P; k++) xp *= x;
  const float res = __builtin_amdgcn_rcpf(fmaf(xp, x, 1.0f));
  sv = res;
  dfu = xp * res * res;
}

template <int P>
__device__ __forceinline__ float sw_fixed_value_x(float x) {
  float xp = 1.0f;
#pragma unroll
  for (int k = 1; k < P; k++) xp *= x;
  return __builtin_amdgcn_rcpf(fmaf(xp, x, 1.0f));
}

template <int P>
__device__ __forceinline__ float sw_fixed_value(float invr0_2, float r2) {
  const float x = r2 * invr0_2;
  float xp = 1.0f;
#pragma unroll
  for (int k = 1; k < P; k++) xp *= x;
  return __builtin_amdgcn_rcpf(fmaf(xp, x, 1.0f));
}

// band records a wave holds in LDS at once: the band pairs of an item are
// decided in rounds of this many (one per lane)
constexpr int kBandSlots = 64;

struct PairArgs {
  int nitems;
  const int* nitems_dev;  // device item count (null: nitems)
  int items_cap, ent_cap; // capacities of the item and entry arrays
  int twogroup;
  float dmax2, prune2, band;
  float cut_lo, cut_hi;   // FP32 r^2 at or below cut_lo: inside; in (cut_lo, cut_hi): the band
  float contact2;         // FP32 r^2 of DBL_EPSILON (the contact-matrix rule)
  float invr0_2;          // FP32 constant of fixedRational (1: positions in units of R_0)
  double r2_real;         // FP32 r^2 -> squared distance (generic switches)
  double sv_scale, sv_shift, df_scale, vir_scale;  // per-item FP32 sums -> outputs
  p2a_switch sw;          // FP64 parameters (generic switch and the cut band)
  const PbcArg* pbc;      // device copies of the prepared cells (FP64 cut decisions)
  const int* cl_sys;      // system of each i cluster (null: one system)
  const double* pos64;    // FP64 positions of the requested atoms (band pairs)
  // Value, six virial products and pair count: with item_partials every item
  // writes its own row of ipart; otherwise every workgroup writes one row of
  // `partial` over all of its items.
  int item_partials;
  double* ipart;
  int* bflag;             // band pairs of each item from a wave's 65th item on
  // Items of each wave of the grid: [wave_start[g], wave_start[g + 1]) for
  // wave g = xcd_block() * 4 + wave (k_wave_sched); null: items g, g + waves, ...
  const int* wave_start;
  // when set and *skip != 0 (written by an earlier kernel on the stream), the
  // launch does nothing: a pass queued on a list that may be stale
  const int* skip;
};

// FP64 re-evaluation of the squared distance of one pair, used only when the
// FP32 value lies within a narrow band around D_MAX^2: the inclusive cut of
// the reference is then decided on the distance Pbc::distance itself would
// produce from the raw positions (including the rounding of Tools::pbc).
// The pair is oriented as the reference lists it (GROUPA atom first; lower
// index first inside one group) because the rounding of Tools::pbc is not
// symmetric under d -> -d.
__device__ double exact_r2(const PbcArg* __restrict__ P, const double* __restrict__ pos, int reqi, int reqj) {
#pragma clang fp contract(off)
  double d[3];
  const int a = min(reqi, reqj), b = max(reqi, reqj);
  pbc_distance(*P, pos + 3 * (size_t)a, pos + 3 * (size_t)b, d);
  return d[0] * d[0] + d[1] * d[1] + d[2] * d[2];
}

// a staged j atom in LDS: position in the i cluster's frame (12 B); its
// origin (entry << 3 | slot, -1 for padding) is kept in a separate short
struct P3 {
  float x, y, z;
};

// per-lane FP32 sums of one work item: sv (counted pairs), dfu * d (i
// derivative), dfu * d (x) d (virial, counted pairs)
struct PairAcc {
  float sv, fx, fy, fz, vxx, vxy, vxz, vyy, vyz, vzz;
};

// the wave's LDS areas for one work item
struct WaveLds {
  P3* sj;       // staged j atoms
  short* sm;    // their origin: entry << 3 | slot, -1 for padding
  int* sc;      // entry codes of the item
  int* sa;      // absolute index of each staged atom (two groups)
  int* sr;      // band records of the current round (kBandSlots)
};

// a band record: staged index (bits 0-9), i slot (10-12), kind (13-14)
__device__ __forceinline__ int band_record(int staged, int ii, int kind) {
  return staged | (ii << 10) | (kind << 13);
}

// FP32 switch terms of one pair on its FP32 r^2
template <int P, bool GENERIC>
__device__ __forceinline__ void pair_switch(const PairArgs& A, float r2, float& sv, float& dfu) {
  if constexpr (GENERIC) {
    double dfd;
    sv = (float)switch_calc_sqr(A.sw, (double)r2 * A.r2_real, dfd);
    dfu = (float)dfd;
  } else {
    sw_fixed<P>(A.invr0_2, r2, sv, dfu);
  }
}

// FP32 separation and r^2 of staged atom pj from i atom pi (the one sequence
// of operations for a pair in the FP32 pass and in its band decisions)
__device__ __forceinline__ float pair_d(const P3& pj, const float4& pi, float& dx, float& dy, float& dz) {
  dx = pj.x - pi.x;
  dy = pj.y - pi.y;
  dz = pj.z - pi.z;
  return fmaf(dx, dx, fmaf(dy, dy, dz * dz));
}

// whether lane (ii, jj) evaluates its pair in a group of kind KIND
template <bool TWO, bool VO, int KIND>
__device__ __forceinline__ bool pair_keep(const WaveLds& L, int g, int ii, int jj, int ai) {
  bool keep = true;
  if constexpr (KIND == 0) keep = jj != ii;  // an atom is not paired with itself
  if constexpr (TWO) keep = keep && (L.sa[g + jj] != ai);
  if constexpr (VO && KIND == 0) keep = keep && jj > ii;  // value only: each pair once
  return keep;
}

// One group of 8 staged j atoms against the 8 i atoms of the item: lane
// (ii, jj) takes i = ii, j = staged atom g + jj.  KIND is uniform over the
// group: 0 = the i cluster with itself (slot order), 1 = entries that count
// the pair (value, virial, derivative), 2 = derivative only.
// A pair whose FP32 r^2 lies in the band around D_MAX^2 (cut_lo < r2 <
// cut_hi) is left out here and counted in nrec (uniform); band_terms
// decides the item's band pairs on FP64 distances after the pass.
template <int P, bool GENERIC, bool TWO, bool VO, int KIND>
__device__ __forceinline__ void pair_term(const PairArgs& A, const P3& pj, const float4& pi, int ii, int jj, int g,
                                          int ai, const WaveLds& L, int& nrec, int& npair, PairAcc& acc) {
  float dx, dy, dz;
  const float r2 = pair_d(pj, pi, dx, dy, dz);
  const bool keep = pair_keep<TWO, VO, KIND>(L, g, ii, jj, ai);
  // inside D_MAX and outside the band: r2 <= cut_lo; in the band: cut_lo < r2 < cut_hi
  const bool cut = keep && r2 <= A.cut_lo;
  const unsigned long long bcut = __ballot(cut);
  nrec += __popcll(__ballot(keep && r2 < A.cut_hi) & ~bcut);
  // a pair outside the cut gets sv = dfu = 0 from x = sw_far (fixed P) or by selection
  float sv, dfu;
  if constexpr (GENERIC) {
    pair_switch<P, GENERIC>(A, r2, sv, dfu);
    sv = cut ? sv : 0.f;
    dfu = cut ? dfu : 0.f;
  } else {
    // positions are in units of R_0: x = r^2
    const float x = cut ? r2 : sw_far<P>();
    if constexpr (VO) {
      sv = sw_fixed_value_x<P>(x);
      dfu = 0.f;
    } else {
      sw_fixed_x<P>(x, sv, dfu);
    }
  }
  if constexpr (VO) {
    npair += __popcll(bcut);
    acc.sv += sv;
    return;
  }
  const float gx = dfu * dx, gy = dfu * dy, gz = dfu * dz;
  // derivative of atom i over every partner (the full list holds both
  // orientations of each cluster pair)
  acc.fx += gx;
  acc.fy += gy;
  acc.fz += gz;
  if constexpr (KIND == 1) {
    // value and virial once per unordered pair
    npair += __popcll(bcut);
    acc.sv += sv;
    acc.vxx = fmaf(gx, dx, acc.vxx);
    acc.vxy = fmaf(gx, dy, acc.vxy);
    acc.vxz = fmaf(gx, dz, acc.vxz);
    acc.vyy = fmaf(gy, dy, acc.vyy);
    acc.vyz = fmaf(gy, dz, acc.vyz);
    acc.vzz = fmaf(gz, dz, acc.vzz);
  } else if constexpr (KIND == 0) {
    const bool c2 = cut && (jj > ii);
    const float wgt = c2 ? 1.f : 0.f;
    npair += __popcll(__ballot(c2));
    acc.sv = fmaf(wgt, sv, acc.sv);
    const float hx = wgt * gx, hy = wgt * gy, hz = wgt * gz;
    acc.vxx = fmaf(hx, dx, acc.vxx);
    acc.vxy = fmaf(hx, dy, acc.vxy);
    acc.vxz = fmaf(hx, dz, acc.vxz);
    acc.vyy = fmaf(hy, dy, acc.vyy);
    acc.vyz = fmaf(hy, dz, acc.vyz);
    acc.vzz = fmaf(hz, dz, acc.vzz);
  }
}

// Groups [gb, ge) of one kind (uniform bounds, multiples of 8), two groups per
// trip; each lane's next staged atom is read one group ahead (the staging
// area has spare slots past the last group for that read).
template <int P, bool GENERIC, bool TWO, bool VO, int KIND>
__device__ __forceinline__ void pair_groups(const PairArgs& A, int gb, int ge, const float4& pi, int ii, int jj,
                                            int ai, const WaveLds& L, int& nrec, int& npair, PairAcc& acc) {
  int g = gb;
  P3 pa = L.sj[g + jj];
  for (; g + CL < ge; g += 2 * CL) {
    const P3 pb = L.sj[g + CL + jj];
    pair_term<P, GENERIC, TWO, VO, KIND>(A, pa, pi, ii, jj, g, ai, L, nrec, npair, acc);
    pa = L.sj[g + 2 * CL + jj];
    pair_term<P, GENERIC, TWO, VO, KIND>(A, pb, pi, ii, jj, g + CL, ai, L, nrec, npair, acc);
  }
  if (g < ge) pair_term<P, GENERIC, TWO, VO, KIND>(A, pa, pi, ii, jj, g, ai, L, nrec, npair, acc);
}

// The band test of pair_term again for one group, numbering the band pairs in
// program order: those numbered [r0, r0 + kBandSlots) become the wave's band
// records (one round of band_terms).
template <bool TWO, bool VO, int KIND>
__device__ __forceinline__ void band_collect(const PairArgs& A, const P3& pj, const float4& pi, int ii, int jj, int g,
                                             int ai, const WaveLds& L, int r0, int& cnt) {
  float dx, dy, dz;
  const float r2 = pair_d(pj, pi, dx, dy, dz);
  const bool keep = pair_keep<TWO, VO, KIND>(L, g, ii, jj, ai);
  const bool cut = keep && r2 <= A.cut_lo;
  const unsigned long long bm = __ballot(keep && r2 < A.cut_hi) & ~__ballot(cut);
  if (bm != 0ull) {
    if ((bm >> (ii + CL * jj)) & 1ull) {
      const int k = cnt + bits_below(bm) - r0;
      if (k >= 0 && k < kBandSlots) L.sr[k] = band_record(g + jj, ii, KIND);
    }
    cnt += __popcll(bm);
  }
}

// The band pairs of the staged item, in rounds of kBandSlots (record r in
// lane r, collected by band_collect; the first round also counts them): the
// inclusive cut decided on the reference's Pbc::distance
// (exact_r2), and for the pairs inside, their FP32 terms from the FP32
// pass's r^2, added in record order in FP64: the i atom derivatives to
// fadd[3] (lanes jj == 0, atom ii), the row entries to radd (lane q < 8:
// value, six virial products, pair count).
template <int P, bool GENERIC, bool TWO, bool VO>
__device__ __forceinline__ void band_terms(const PairArgs& A, const WaveLds& L, int ci,
                                           const float4* __restrict__ posI, const float4* __restrict__ posJ,
                                           const float4& pi_lane, int ii, int jj, int lane, int ai, int n0, int b1,
                                           int n1, int b2, int n2, double* fadd, double& radd) {
  const PbcArg* pbc = A.pbc + nsys_of(A.cl_sys, ci);
  int nrec = 0;
  for (int r0 = 0; r0 == 0 || r0 < nrec; r0 += kBandSlots) {
    {
      int cnt = 0;
      if (n0) band_collect<TWO, VO, 0>(A, L.sj[jj], pi_lane, ii, jj, 0, ai, L, r0, cnt);
      for (int g = b1; g < b1 + n1; g += CL) band_collect<TWO, VO, 1>(A, L.sj[g + jj], pi_lane, ii, jj, g, ai, L, r0, cnt);
      if constexpr (!VO)
        for (int g = b2; g < b2 + n2; g += CL)
          band_collect<TWO, VO, 2>(A, L.sj[g + jj], pi_lane, ii, jj, g, ai, L, r0, cnt);
      nrec = cnt;
      // the records are read by other lanes of this wave
      __builtin_amdgcn_fence(__ATOMIC_SEQ_CST, "wavefront");
      __builtin_amdgcn_wave_barrier();
    }
    const int nr = min(kBandSlots, nrec - r0);
    int rii = 0, cnt = 0;
    float sv = 0.f, g[3] = {0.f, 0.f, 0.f}, d[3] = {0.f, 0.f, 0.f};
    if (lane < nr) {
      const int rec = L.sr[lane];
      const int st = rec & 1023, kind = (rec >> 13) & 3;
      rii = (rec >> 10) & 7;
      const float4 pi = posI[ci * CL + rii];
      const float r2 = pair_d(L.sj[st], pi, d[0], d[1], d[2]);
      const int meta = L.sm[st];  // (entry << 3 | slot), -1 for padding
      const int reqi = __float_as_int(pi.w);
      const int reqj = meta >= 0 ? __float_as_int(posJ[(L.sc[meta >> 3] & kCjMask) * CL + (meta & 7)].w) : -1;
      if (reqi >= 0 && reqj >= 0 && switch_inside(A.sw, exact_r2(pbc, A.pos64, reqi, reqj))) {
        float dfu;
        pair_switch<P, GENERIC>(A, r2, sv, dfu);
        if constexpr (!VO) {
          g[0] = dfu * d[0];
          g[1] = dfu * d[1];
          g[2] = dfu * d[2];
        }
        // a self pair counts once: the staged index of a self entry is its slot
        cnt = (kind == 1 || (kind == 0 && st > rii)) ? 1 : 0;
      } else {
        sv = 0.f;
      }
    }
    for (int r = 0; r < nr; r++) {
      const int ri = __builtin_amdgcn_readlane(rii, r);
      const int rc = __builtin_amdgcn_readlane(cnt, r);
      const float rs = __int_as_float(__builtin_amdgcn_readlane(__float_as_int(sv), r));
      float rg[3], rd[3];
      for (int c = 0; c < 3; c++) {
        rg[c] = __int_as_float(__builtin_amdgcn_readlane(__float_as_int(g[c]), r));
        rd[c] = __int_as_float(__builtin_amdgcn_readlane(__float_as_int(d[c]), r));
      }
      if (!VO && jj == 0 && ii == ri)
        for (int c = 0; c < 3; c++) fadd[c] += -A.df_scale * (double)rg[c];
      if (rc && lane < 8) {
        if (lane == 0) radd += A.sv_scale * (double)rs + A.sv_shift;
        else if (lane == 7) radd += 1.0;
        else if (!VO) {
          const int a = lane <= 3 ? 0 : (lane <= 5 ? 1 : 2);
          const int b = lane <= 3 ? lane - 1 : (lane <= 5 ? lane - 3 : 2);
          radd += A.vir_scale * (double)(rg[a] * rd[b]);
        }
      }
    }
    // the next round overwrites the records
    __builtin_amdgcn_fence(__ATOMIC_SEQ_CST, "wavefront");
    __builtin_amdgcn_wave_barrier();
  }
}

// Row entry `lane` (< 8) of FP32 per-lane sums: value, six virial products,
// pair count (npair uniform); the wave sums are formed in every lane.
template <bool VO>
__device__ __forceinline__ double item_row(const PairArgs& A, const PairAcc& acc, int npair, int lane) {
  float tv[7];
  tv[0] = wave_sum_f(acc.sv);
  if constexpr (!VO) {
    tv[1] = wave_sum_f(acc.vxx);
    tv[2] = wave_sum_f(acc.vxy);
    tv[3] = wave_sum_f(acc.vxz);
    tv[4] = wave_sum_f(acc.vyy);
    tv[5] = wave_sum_f(acc.vyz);
    tv[6] = wave_sum_f(acc.vzz);
  } else {
    tv[1] = tv[2] = tv[3] = tv[4] = tv[5] = tv[6] = 0.f;
  }
  float tq = tv[0];
#pragma unroll
  for (int q = 1; q < 7; q++)
    if (lane == q) tq = tv[q];
  return lane == 0 ? A.sv_scale * (double)tq + A.sv_shift * (double)npair
                   : (lane < 7 ? A.vir_scale * (double)tq : (double)npair);
}

// Staging of one work item by one wave (i cluster ci, entries [e0, e1)):
//  1. lane-parallel pruning of the chunk's entries against the current
//     bounding boxes (D_MAX);
//  2. each passing lane loads the 8 atoms of its j cluster, adds the image
//     offset and keeps the atoms within D_MAX of the i cluster's box; the
//     kept atoms are staged in LDS, compacted and grouped by entry kind
//     (self / counting / derivative-only), each kind padded to a multiple
//     of 8 with far-away atoms, so that every group of 8 has one kind.
// Returns the group bounds (uniform): n0 (0 or 8), [b1, b1+n1), [b2, b2+n2).
constexpr int kFewEntries = 16;
template <bool TWO, bool VO, int CH>
__device__ __forceinline__ void stage_item(const PairArgs& A, const Entry& my_e, int lane, int e0, int e1,
                                           const float4& ic, const float4& ih, const float4* __restrict__ posJ,
                                           const float4* __restrict__ bbJ, const int* __restrict__ absJ,
                                           const WaveLds& L, int& n0, int& b1, int& n1, int& b2, int& n2) {
  // Two groups: an item of at most kFewEntries entries (the dense group's
  // side of a sparse group) stages each entry with four lanes of two atoms
  // each instead of one lane of eight; lanes are ordered by (entry, atom)
  // either way, so the staged order is the same.  Single groups keep the
  // one-lane form (their items are mostly full, and the branch costs them).
  const bool few = TWO && e1 - e0 <= kFewEntries;  // uniform over the wave
  const int ent = few ? (lane >> 2) : lane;  // this lane's entry of the item
  const int sub2 = few ? 2 * (lane & 3) : 0;  // its first atom (few)
  Entry my = {0, 0.f, 0.f, 0.f};
  bool pass = false;
  {
    Entry src = my_e;
    if (few) {
      src.code = __shfl(my_e.code, ent);
      src.dx = __shfl(my_e.dx, ent);
      src.dy = __shfl(my_e.dy, ent);
      src.dz = __shfl(my_e.dz, ent);
    }
    const int e = e0 + ent;
    if (ent < CH && e < e1) {
      my = src;
      if (!VO || (my.code & kCountBit) || my.code < 0) {
        const int cj = my.code & kCjMask;
        const float4 jc = bbJ[2 * cj], jh = bbJ[2 * cj + 1];
        const float gx = fmaxf(fabsf(jc.x + my.dx - ic.x) - ih.x - jh.x, 0.f);
        const float gy = fmaxf(fabsf(jc.y + my.dy - ic.y) - ih.y - jh.y, 0.f);
        const float gz = fmaxf(fabsf(jc.z + my.dz - ic.z) - ih.z - jh.z, 0.f);
        pass = gx * gx + gy * gy + gz * gz <= A.prune2;
      }
    }
  }
  float4 q[CL];
  int m8 = 0, kind = 3;
  if (pass && few) {
    const int cj = my.code & kCjMask;
#pragma unroll
    for (int u = 0; u < 2; u++) {
      const float4 v = posJ[cj * CL + sub2 + u];
      q[u] = make_float4(v.x + my.dx, v.y + my.dy, v.z + my.dz, 0.f);
    }
    if (my.code < 0) {
      kind = 0;
      m8 = 3 << sub2;
    } else {
      kind = (my.code & kCountBit) ? 1 : 2;
#pragma unroll
      for (int u = 0; u < 2; u++) {
        const float gx = fmaxf(fabsf(q[u].x - ic.x) - ih.x, 0.f);
        const float gy = fmaxf(fabsf(q[u].y - ic.y) - ih.y, 0.f);
        const float gz = fmaxf(fabsf(q[u].z - ic.z) - ih.z, 0.f);
        if (gx * gx + gy * gy + gz * gz <= A.prune2) m8 |= 1 << (sub2 + u);
      }
    }
  } else if (pass) {
    const int cj = my.code & kCjMask;
#pragma unroll
    for (int t = 0; t < CL; t++) {
      const float4 v = posJ[cj * CL + t];
      q[t] = make_float4(v.x + my.dx, v.y + my.dy, v.z + my.dz, 0.f);
    }
    if (my.code < 0) {
      kind = 0;
      m8 = 0xFF;  // the i cluster with itself keeps slot order
    } else {
      kind = (my.code & kCountBit) ? 1 : 2;
#pragma unroll
      for (int t = 0; t < CL; t++) {
        const float gx = fmaxf(fabsf(q[t].x - ic.x) - ih.x, 0.f);
        const float gy = fmaxf(fabsf(q[t].y - ic.y) - ih.y, 0.f);
        const float gz = fmaxf(fabsf(q[t].z - ic.z) - ih.z, 0.f);
        if (gx * gx + gy * gy + gz * gz <= A.prune2) m8 |= 1 << t;
      }
    }
  }
  const int cnt = __popc(m8);
  n0 = __ballot(kind == 0) ? CL : 0;  // at most one self entry per item
  int tot1, tot2 = 0;
  const int o1 = wave_excl_scan_dpp(kind == 1 ? cnt : 0, tot1);
  int o2 = 0;
  if constexpr (!VO) o2 = wave_excl_scan_dpp(kind == 2 ? cnt : 0, tot2);
  n1 = (tot1 + CL - 1) & ~(CL - 1);
  n2 = (tot2 + CL - 1) & ~(CL - 1);
  b1 = n0;
  b2 = n0 + n1;
  if (pass) {
    if (sub2 == 0) L.sc[ent] = my.code;
    const int cj = my.code & kCjMask;
    int p = kind == 0 ? sub2 : (kind == 1 ? b1 + o1 : b2 + o2);
    if (few) {
#pragma unroll
      for (int u = 0; u < 2; u++) {
        const int t = sub2 + u;
        if ((m8 >> t) & 1) {
          L.sj[p] = P3{q[u].x, q[u].y, q[u].z};
          L.sm[p] = (short)((ent << 3) | t);
          if constexpr (TWO) L.sa[p] = absJ[cj * CL + t];
          p++;
        }
      }
    } else {
#pragma unroll
      for (int t = 0; t < CL; t++) {
        if ((m8 >> t) & 1) {
          L.sj[p] = P3{q[t].x, q[t].y, q[t].z};
          L.sm[p] = (short)((lane << 3) | t);
          if constexpr (TWO) L.sa[p] = absJ[cj * CL + t];
          p++;
        }
      }
    }
  }
  // padding of each kind up to a multiple of 8: far away, never within D_MAX
  if (lane < CL) {
    const P3 far = {1.0e5f, 1.0e5f, 1.0e5f};
    if (tot1 + lane < n1) {
      L.sj[b1 + tot1 + lane] = far;
      L.sm[b1 + tot1 + lane] = -1;
      if constexpr (TWO) L.sa[b1 + tot1 + lane] = -1;
    }
    if (tot2 + lane < n2) {
      L.sj[b2 + tot2 + lane] = far;
      L.sm[b2 + tot2 + lane] = -1;
      if constexpr (TWO) L.sa[b2 + tot2 + lane] = -1;
    }
  }
  // the staging above is read by other lanes of this wave only
  __builtin_amdgcn_fence(__ATOMIC_SEQ_CST, "wavefront");
  __builtin_amdgcn_wave_barrier();
}

// staged atoms per wave: 64 entries x 8, padding of two kinds, and the
// one-group read-ahead (the largest group end is 520)
template <int CH>
constexpr int stage_cap() {
  return CH * CL + 2 * CL;
}

// FP32 pass: one wave per work item (i cluster, chunk of <= CH list
// entries); a grid smaller than the item count makes the waves loop over
// items.  Staging (stage_item), then a dense loop over the groups of 8
// staged atoms: lane l evaluates the pair (i = l&7 of cluster ci, j = staged
// atom 8g + (l>>3)).  VO (value only): no derivatives and no virial;
// derivative-only entries are skipped, so each unordered pair is visited
// once.  The pairs in the band around D_MAX^2 are decided at the end of
// their item on FP64 distances.
// Outputs: fpart[item][8][3], the derivative of each i atom over the item,
// and the value / virial / pair count rows described at PairArgs, all over
// the pairs outside the band plus the FP64 terms of the band pairs inside
// the cut (band_terms).  For the workgroup rows each lane keeps FP32 running
// sums over all of the wave's items; their wave sums are taken once at the
// end, converted to FP64 and added to the wave's band sums, and the waves'
// rows are added in wave order.
template <int P, bool GENERIC, bool TWO, bool VO, int CH>
__global__ __launch_bounds__(PAIR_BLOCK) __attribute__((amdgpu_waves_per_eu(4))) void k_pairs(
    PairArgs A, const int4* __restrict__ items, const Entry* __restrict__ ents, const float4* __restrict__ posI,
    const float4* __restrict__ posJ, const float4* __restrict__ bbI, const float4* __restrict__ bbJ,
    const int* __restrict__ absI, const int* __restrict__ absJ, double* __restrict__ fpart,
    double* __restrict__ partial) {
  constexpr int NW = PAIR_BLOCK / WAVE;
  constexpr int CAP = stage_cap<CH>();
  __shared__ P3 sj[NW][CAP];
  __shared__ short sm[NW][CAP];
  __shared__ int scode[NW][CH];
  __shared__ int sabs[TWO ? NW : 1][TWO ? CAP : 1];
  __shared__ int srec[NW][kBandSlots];
  __shared__ double stot[NW][8];
  if (A.skip && *A.skip) return;
  const int lane = threadIdx.x & 63;
  const int w = threadIdx.x >> 6;
  const int ii = lane & 7, jj = lane >> 3;
  const int nitems = A.nitems_dev ? min(*A.nitems_dev, A.items_cap) : A.nitems;
  const WaveLds L = {sj[w], sm[w], scode[w], TWO ? sabs[w] : nullptr, srec[w]};
  PairAcc acc = {0.f, 0.f, 0.f, 0.f, 0.f, 0.f, 0.f, 0.f, 0.f, 0.f};
  int ntot = 0;
  // lanes < 8: the wave's FP64 sums of the band pairs' row entries
  if (lane < 8) stot[w][lane] = 0.0;

  // the next item's descriptor and entries are loaded while the current item
  // computes (a wave handles several items when the grid is persistent)
  // the wave's items: a contiguous range of balanced cost, or every
  // (grid waves)-th item
  const int blk = xcd_block();
  const int gw = blk * NW + w;
  int kfirst = gw, kstride = gridDim.x * NW, kend = nitems;
  if (A.wave_start) {
    kfirst = A.wave_start[gw];
    kend = min(A.wave_start[gw + 1], nitems);
    kstride = 1;
  }
  int k = kfirst;
  int4 it_n = make_int4(0, 0, 0, 0);
  Entry my_n = {0, 0.f, 0.f, 0.f};
  if (k < kend) {
    it_n = items[k];
    const int e = it_n.y + lane;
    if (lane < CH && e < min(it_n.z, A.ent_cap)) my_n = ents[e];
  }
  // the wave's items with band pairs: bit j for its j-th item (j < 64)
  unsigned long long bandbits = 0ull;
  for (int j = 0; k < kend; k += kstride, j++) {
    const int item = k;
    const int4 it = it_n;
    const Entry my_e = my_n;
    const int kn = k + kstride;
    if (kn < kend) it_n = items[kn];
    const int ci = it.x;
    const int e1 = min(it.z, A.ent_cap);
    const float4 pi = posI[ci * CL + ii];
    const int ai = TWO ? absI[ci * CL + ii] : 0;
    const float4 ic = bbI[2 * ci], ih = bbI[2 * ci + 1];
    int n0, b1, n1, b2, n2;
    stage_item<TWO, VO, CH>(A, my_e, lane, it.y, e1, ic, ih, posJ, bbJ, absJ, L, n0, b1, n1, b2, n2);
    if (kn < kend) {
      const int e = it_n.y + lane;
      my_n = (lane < CH && e < min(it_n.z, A.ent_cap)) ? ents[e] : Entry{0, 0.f, 0.f, 0.f};
    }
    // group bounds are wave-uniform: scalar loops
    n1 = __builtin_amdgcn_readfirstlane(n1);
    n2 = __builtin_amdgcn_readfirstlane(n2);
    b1 = __builtin_amdgcn_readfirstlane(b1);
    b2 = __builtin_amdgcn_readfirstlane(b2);
    int nrec = 0, npair = 0;
    // the pair loops issue ahead of the other waves' staging and item ends
    // (measured: the pass takes ~3% less time than with one priority)
    __builtin_amdgcn_s_setprio(1);
    if (n0) pair_term<P, GENERIC, TWO, VO, 0>(A, L.sj[jj], pi, ii, jj, 0, ai, L, nrec, npair, acc);
    pair_groups<P, GENERIC, TWO, VO, 1>(A, b1, b1 + n1, pi, ii, jj, ai, L, nrec, npair, acc);
    if constexpr (!VO)
      pair_groups<P, GENERIC, TWO, VO, 2>(A, b2, b2 + n2, pi, ii, jj, ai, L, nrec, npair, acc);
    __builtin_amdgcn_s_setprio(0);
    if (j < 64) {
      if (nrec > 0) bandbits |= 1ull << j;
    } else if (lane == 0) {
      A.bflag[item] = nrec;
    }
    if constexpr (!VO) {
      const float fx = sum_over_jj(acc.fx), fy = sum_over_jj(acc.fy), fz = sum_over_jj(acc.fz);
      if (jj == 0) {
        double* fp = fpart + (size_t)item * (CL * 3) + ii * 3;
        fp[0] = -A.df_scale * (double)fx;
        fp[1] = -A.df_scale * (double)fy;
        fp[2] = -A.df_scale * (double)fz;
      }
      acc.fx = acc.fy = acc.fz = 0.f;
    }
    if (A.item_partials) {
      const double o = item_row<VO>(A, acc, npair, lane);
      if (lane < 8) A.ipart[(size_t)item * 8 + lane] = o;
      acc = PairAcc{0.f, 0.f, 0.f, 0.f, 0.f, 0.f, 0.f, 0.f, 0.f, 0.f};
    } else {
      // value and virial stay in the lanes' running sums over the wave's items
      ntot += npair;
    }
    // the next item overwrites the staging read above
    __builtin_amdgcn_fence(__ATOMIC_SEQ_CST, "wavefront");
    __builtin_amdgcn_wave_barrier();
  }
  // the band pairs of the wave's items (rare), item by item after the pass so
  // that their FP64 registers are not held during it: the item is staged
  // again and its band terms are added to its outputs in FP64 (each output
  // by the lane that wrote it above)
  for (int jb = 0, kb = kfirst; kb < kend; jb++, kb += kstride) {
    if (jb < 64) {
      // straight to the next item with band pairs
      const unsigned long long rest = bandbits >> jb;
      if (rest == 0ull) {
        jb = 63;
        kb = kfirst + 63 * kstride;
        continue;
      }
      const int skip = __builtin_ctzll(rest);
      jb += skip;
      kb += skip * kstride;
    } else if (__builtin_amdgcn_readfirstlane(A.bflag[kb]) == 0) {
      continue;
    }
    const int4 it = items[kb];
    const int ci = it.x;
    const int e1 = min(it.z, A.ent_cap);
    const float4 pi = posI[ci * CL + ii];
    const int ai = TWO ? absI[ci * CL + ii] : 0;
    const float4 ic = bbI[2 * ci], ih = bbI[2 * ci + 1];
    const int e = it.y + lane;
    const Entry my_e = (lane < CH && e < e1) ? ents[e] : Entry{0, 0.f, 0.f, 0.f};
    int n0, b1, n1, b2, n2;
    stage_item<TWO, VO, CH>(A, my_e, lane, it.y, e1, ic, ih, posJ, bbJ, absJ, L, n0, b1, n1, b2, n2);
    n1 = __builtin_amdgcn_readfirstlane(n1);
    n2 = __builtin_amdgcn_readfirstlane(n2);
    b1 = __builtin_amdgcn_readfirstlane(b1);
    b2 = __builtin_amdgcn_readfirstlane(b2);
    double fadd[3] = {0.0, 0.0, 0.0}, radd = 0.0;
    band_terms<P, GENERIC, TWO, VO>(A, L, ci, posI, posJ, pi, ii, jj, lane, ai, n0, b1, n1, b2, n2, fadd, radd);
    if constexpr (!VO) {
      if (jj == 0) {
        double* fp = fpart + (size_t)kb * (CL * 3) + ii * 3;
        fp[0] += fadd[0];
        fp[1] += fadd[1];
        fp[2] += fadd[2];
      }
    }
    if (lane < 8) {
      if (A.item_partials) A.ipart[(size_t)kb * 8 + lane] += radd;
      else stot[w][lane] += radd;
    }
    __builtin_amdgcn_fence(__ATOMIC_SEQ_CST, "wavefront");
    __builtin_amdgcn_wave_barrier();
  }
  if (!A.item_partials) {
    const double o = item_row<VO>(A, acc, ntot, lane);
    if (lane < 8) stot[w][lane] += o;
    __syncthreads();
    if (w == 0 && lane < 8) {
      double s = 0.0;
      for (int v = 0; v < NW; v++) s += stot[v][lane];
      partial[(size_t)blk * 8 + lane] = s;
    }
  }
}

// virial = -sum dd (x) d (CoordinationBase.cpp:200-206), symmetric
__device__ __forceinline__ void write_virial(double* virial, const double* s6) {
  double xx = -s6[0], xy = -s6[1], xz = -s6[2], yy = -s6[3], yz = -s6[4], zz = -s6[5];
  virial[0] = xx; virial[1] = xy; virial[2] = xz;
  virial[3] = xy; virial[4] = yy; virial[5] = yz;
  virial[6] = xz; virial[7] = yz; virial[8] = zz;
}

// Inputs of the final value / virial / pair-count sum: the workgroup rows of
// the pair kernels.
struct RedArgs {
  int nred;
  const double* partial;
  double* value;
  double* virial;
  double* stats;
  const int* skip;  // as PairArgs::skip
};

// After the pair kernels, in one launch.  With do_red, workgroup 0 forms the
// value, virial and pair count from the rows of R in a fixed order (each
// thread a fixed set of rows, then a fixed tree); the other ngather
// workgroups form the derivatives of the atoms of every i cluster (fixed-order
// sums of the cluster's item partials).  The reduction workgroup comes first
// so that it starts with the first wave of workgroups.
__global__ __launch_bounds__(256) void k_gather_reduce(int ngather, int ncl, const int* __restrict__ ncl_dev,
                                                       const float4* __restrict__ p,
                                                       const int* __restrict__ item_start, int items_cap,
                                                       const double* __restrict__ fpart, double* __restrict__ deriv,
                                                       int do_red, RedArgs R) {
  if (R.skip && *R.skip) return;
  const int t = threadIdx.x;
  const int red = do_red ? 1 : 0;
  if ((int)blockIdx.x >= red) {
    if (!deriv) return;
    const int g = (blockIdx.x - red) * blockDim.x + t;
    if (ncl_dev) ncl = *ncl_dev;
    if (g >= ncl * CL * 3) return;
    const int c = g / (CL * 3), r = g % (CL * 3);
    // every load that depends only on c goes out before the first wait
    const int i0 = item_start[c];
    const int i1 = min(item_start[c + 1], items_cap);
    const int req = __float_as_int(p[c * CL + r / 3].w);
    if (req < 0) return;
    double sum = 0.0;
    for (int it = i0; it < i1; it += 4) {
      double v[4];
#pragma unroll
      for (int u = 0; u < 4; u++) v[u] = it + u < i1 ? fpart[(size_t)(it + u) * (CL * 3) + r] : 0.0;
#pragma unroll
      for (int u = 0; u < 4; u++) sum += v[u];
    }
    deriv[3 * (size_t)req + r % 3] = sum;
    return;
  }
  __shared__ double s[8][4];
  double acc[8] = {0, 0, 0, 0, 0, 0, 0, 0};
  const double4* p4 = reinterpret_cast<const double4*>(R.partial);
  // rows t, t + 256, t + 512, ... of each thread, eight in flight at once
  for (int i0 = t; i0 < R.nred; i0 += 256 * 8) {
    double4 x[8], y[8];
#pragma unroll
    for (int u = 0; u < 8; u++) {
      const int i = i0 + u * 256;
      if (i < R.nred) {
        x[u] = p4[2 * (size_t)i];
        y[u] = p4[2 * (size_t)i + 1];
      } else {
        x[u] = make_double4(0.0, 0.0, 0.0, 0.0);
        y[u] = x[u];
      }
    }
#pragma unroll
    for (int u = 0; u < 8; u++) {
      acc[0] += x[u].x; acc[1] += x[u].y; acc[2] += x[u].z; acc[3] += x[u].w;
      acc[4] += y[u].x; acc[5] += y[u].y; acc[6] += y[u].z; acc[7] += y[u].w;
    }
  }
  // wave sums on the VALU, then the four waves in order
#pragma unroll
  for (int q = 0; q < 8; q++) acc[q] = wave_sum_d(acc[q]);
  if ((t & 63) == 0)
    for (int q = 0; q < 8; q++) s[q][t >> 6] = acc[q];
  __syncthreads();
  if (t == 0) {
    double r[8];
    for (int q = 0; q < 8; q++) r[q] = ((s[q][0] + s[q][1]) + s[q][2]) + s[q][3];
    if (R.value) *R.value = r[0];
    if (R.virial) {
      double v6[6] = {r[1], r[2], r[3], r[4], r[5], r[6]};
      write_virial(R.virial, v6);
    }
    if (R.stats) {
      R.stats[0] = r[7];
      R.stats[1] += r[7];  // running total since the last reset
    }
  }
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
      double v6[6] = {s[1][0], s[2][0], s[3][0], s[4][0], s[5][0], s[6][0]};
      write_virial(virial, v6);
    }
    if (stats_out) {
      stats_out[0] = s[7][0];
      stats_out[1] += s[7][0];  // running total since the last reset
    }
  }
}

// per-system sums of the item partials: one workgroup per system; the items
// of system s are those of its clusters, contiguous after its columns
__global__ __launch_bounds__(256) void k_sys_reduce(int nsys, const int* __restrict__ sys_colbase,
                                                    const int* __restrict__ col_cl,
                                                    const int* __restrict__ item_start, int items_cap,
                                                    const double* __restrict__ partial, double* __restrict__ values,
                                                    double* __restrict__ virials, double* __restrict__ npairs) {
  __shared__ double s[8][256];
  const int sy = blockIdx.x;
  if (sy >= nsys) return;
  const int t = threadIdx.x;
  const int c0 = col_cl[sys_colbase[sy]], c1 = col_cl[sys_colbase[sy + 1]];
  const int i0 = min(item_start[c0], items_cap), i1 = min(item_start[c1], items_cap);
  double acc[8] = {0, 0, 0, 0, 0, 0, 0, 0};
  for (int i = i0 + t; i < i1; i += 256)
    for (int q = 0; q < 8; q++) acc[q] += partial[(size_t)i * 8 + q];
  for (int q = 0; q < 8; q++) s[q][t] = acc[q];
  __syncthreads();
  for (int o = 128; o > 0; o >>= 1) {
    if (t < o)
      for (int q = 0; q < 8; q++) s[q][t] += s[q][t + o];
    __syncthreads();
  }
  if (t == 0) {
    values[sy] = s[0][0];
    double v6[6] = {s[1][0], s[2][0], s[3][0], s[4][0], s[5][0], s[6][0]};
    write_virial(virials + 9 * (size_t)sy, v6);
    npairs[sy] = s[7][0];
  }
}

// ------------------------------------------------------------------------
// FP64 fallback kernels
// ------------------------------------------------------------------------
// Every requested atom accumulates its own derivative from all its partners
// (no atomics); the value and the virial are taken once per pair.  The
// partners are every atom of the other group (two groups) or of the group,
// or, with a neighbour list (nbr_start non-null), the atoms nbr[nbr_start[i]
// .. nbr_start[i+1]) listed for atom i in both directions.
__global__ void k_allpairs(int nA, int nB, int use_pbc, PbcArg P, p2a_switch sw, const double* __restrict__ pos,
                           const long long* __restrict__ absidx, double* __restrict__ deriv,
                           double* __restrict__ partial, const int* __restrict__ nbr_start = nullptr,
                           const int* __restrict__ nbr = nullptr) {
#pragma clang fp contract(off)
  const int natoms = nA + nB;
  const int i = blockIdx.x * blockDim.x + threadIdx.x;
  double cv = 0, v[6] = {0, 0, 0, 0, 0, 0}, f[3] = {0, 0, 0};
  double np = 0;
  if (i < natoms) {
    const bool two = nB > 0;
    int j0, j1;
    if (nbr_start) {
      j0 = nbr_start[i];
      j1 = nbr_start[i + 1];
    } else if (two) {
      if (i < nA) { j0 = nA; j1 = natoms; } else { j0 = 0; j1 = nA; }
    } else {
      j0 = 0;
      j1 = nA;
    }
    const double* xi = pos + 3 * (size_t)i;
    for (int jj = j0; jj < j1; jj++) {
      const int j = nbr_start ? nbr[jj] : jj;
      // absidx null: the requested atoms are distinct atoms
      if (j == i || (absidx && absidx[i] == absidx[j])) continue;
      const double* xj = pos + 3 * (size_t)j;
      // the pair is measured once, as the reference lists it: p -> q with p
      // in GROUPA (two groups) or p < q, d = Pbc::distance(x_p, x_q); atom
      // i is p ("own") or q.  Measured from q, a pair at half a cell could
      // take another of its tied images.
      const bool own = two ? (i < nA) : (i < j);
      const double* xp = own ? xi : xj;
      const double* xq = own ? xj : xi;
      double d[3];
      if (use_pbc) pbc_distance(P, xp, xq, d);
      else { d[0] = xq[0] - xp[0]; d[1] = xq[1] - xp[1]; d[2] = xq[2] - xp[2]; }
      double r2 = d[0] * d[0] + d[1] * d[1] + d[2] * d[2];
      double df;
      double s = switch_calc_sqr(sw, r2, df);
      // CoordinationBase: deriv[p] -= df*d, deriv[q] += df*d
      const double g = own ? -df : df;
      f[0] += g * d[0];
      f[1] += g * d[1];
      f[2] += g * d[2];
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
// (with a neighbour list, only the pairs k whose nl_keep[k] is set)
__global__ void k_pairmode(int nA, int use_pbc, PbcArg P, p2a_switch sw, const double* __restrict__ pos,
                           const long long* __restrict__ absidx, double* __restrict__ deriv,
                           double* __restrict__ partial, const int* __restrict__ nl_keep = nullptr) {
#pragma clang fp contract(off)
  const int i = blockIdx.x * blockDim.x + threadIdx.x;
  double r[8] = {0, 0, 0, 0, 0, 0, 0, 0};
  if (i < nA) {
    const int j = nA + i;
    double* di = deriv + 3 * (size_t)i;
    double* dj = deriv + 3 * (size_t)j;
    di[0] = di[1] = di[2] = 0.0;
    dj[0] = dj[1] = dj[2] = 0.0;
    if (absidx[i] != absidx[j] && (!nl_keep || nl_keep[i])) {
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
// Neighbour lists of the FP64 path, with the candidate sets of the
// reference (NeighborList::update, NeighborList.cpp:177-308), stored as a
// symmetric CSR: atom i lists every partner j of a listed pair (i,j) or (j,i)
// ------------------------------------------------------------------------

// LinkCells grid of the cells branch: cell of an atom from its scaled
// coordinates (LinkCells::findMyCell, LinkCells.cpp:277-292); `origin` is
// subtracted first when the grid box was built around the atoms (no cell
// given, LinkCells.cpp:49-73)
struct CellGrid {
  double inv[9];  // inverse of the grid box (rows), as Pbc::setBox forms it
  double origin[3];
  int shift_origin;
  int nc[3];
  int use_pbc;  // stencil wraps around (LinkCells::addRequiredCells)
};

#pragma clang fp contract(off)
__device__ __forceinline__ void grid_cell(const CellGrid& G, const double* x, int* c) {
  double m[3] = {x[0], x[1], x[2]};
  if (G.shift_origin)
    for (int k = 0; k < 3; k++) m[k] = m[k] - G.origin[k];
  for (int j = 0; j < 3; j++) {
    // Pbc::realToScaled: matmul(transpose(invBox), d)
    const double f = m[0] * G.inv[j] + m[1] * G.inv[3 + j] + m[2] * G.inv[6 + j];
    int cj = (int)floor((pbc1(f) + 0.5) * G.nc[j]);
    c[j] = min(max(cj, 0), G.nc[j] - 1);
  }
}
#pragma clang fp contract(fast)

// cells a and b are within one stencil (LinkCells::addRequiredCells,
// LinkCells.cpp:183-239: cells c-1..c+1 along each direction, wrapped with
// pbc, every cell when there are fewer than three; without pbc, clipped)
__device__ __forceinline__ bool grid_adjacent(const CellGrid& G, const int* a, const int* b) {
  for (int k = 0; k < 3; k++) {
    const int n = G.nc[k];
    if (G.use_pbc) {
      if (n >= 3) {
        const int d = (b[k] - a[k] + n) % n;
        if (d != 0 && d != 1 && d != n - 1) return false;
      }
    } else if (abs(a[k] - b[k]) > 1) {
      return false;
    }
  }
  return true;
}

// pair (p,q) as the reference lists it: GROUPA atom first with two groups,
// the lower index first in one group
__device__ __forceinline__ bool nl_member(int mode, int p, int q, int use_pbc, const PbcArg& P, double cut2,
                                          const CellGrid& G, const double* __restrict__ pos,
                                          const int* __restrict__ cellof) {
#pragma clang fp contract(off)
  if (mode == P2A_NL_CELLS) {
    const int a[3] = {cellof[3 * p], cellof[3 * p + 1], cellof[3 * p + 2]};
    const int b[3] = {cellof[3 * q], cellof[3 * q + 1], cellof[3 * q + 2]};
    return grid_adjacent(G, a, b);
  }
  const double* xa = pos + 3 * (size_t)p;
  const double* xb = pos + 3 * (size_t)q;
  double d[3];
  if (use_pbc) pbc_distance(P, xa, xb, d);
  else { d[0] = xb[0] - xa[0]; d[1] = xb[1] - xa[1]; d[2] = xb[2] - xa[2]; }
  return d[0] * d[0] + d[1] * d[1] + d[2] * d[2] <= cut2;
}

__global__ void k_nl_cells(int natoms, CellGrid G, const double* __restrict__ pos, int* __restrict__ cellof) {
  const int i = blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= natoms) return;
  int c[3];
  grid_cell(G, pos + 3 * (size_t)i, c);
  cellof[3 * i] = c[0];
  cellof[3 * i + 1] = c[1];
  cellof[3 * i + 2] = c[2];
}

// FILL = 0: count[i] = listed partners of atom i; FILL = 1: write them, in
// index order, from start[i]
template <int FILL>
__global__ void k_nl_atoms(int nA, int nB, int mode, int use_pbc, PbcArg P, double cut2, CellGrid G,
                           const double* __restrict__ pos, const int* __restrict__ cellof,
                           int* __restrict__ count, const int* __restrict__ start, int* __restrict__ nbr) {
  const int natoms = nA + nB;
  const int i = blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= natoms) return;
  const bool two = nB > 0;
  const int j0 = two ? (i < nA ? nA : 0) : 0;
  const int j1 = two ? (i < nA ? natoms : nA) : nA;
  int n = 0;
  const int base = FILL ? start[i] : 0;
  for (int j = j0; j < j1; j++) {
    if (!two && j == i) continue;
    const int p = two ? (i < nA ? i : j) : min(i, j);
    const int q = two ? (i < nA ? j : i) : max(i, j);
    if (nl_member(mode, p, q, use_pbc, P, cut2, G, pos, cellof)) {
      if (FILL) nbr[base + n] = j;
      n++;
    }
  }
  if (!FILL) count[i] = n;
}

// PAIR with a list: keep[k] = pair k within NL_CUTOFF (classic branch)
__global__ void k_nl_pairs_keep(int nA, int use_pbc, PbcArg P, double cut2, const double* __restrict__ pos,
                                int* __restrict__ keep) {
  const int i = blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= nA) return;
  keep[i] = nl_member(P2A_NL_CLASSIC, i, nA + i, use_pbc, P, cut2, CellGrid{}, pos, nullptr) ? 1 : 0;
}

// exclusive scan of n counts in one workgroup of 1024 threads (contiguous
// chunks per thread); start[n] = total
__global__ __launch_bounds__(1024) void k_scan_small(int n, const int* __restrict__ count, int* __restrict__ start,
                                                     long long* __restrict__ total) {
  __shared__ long long part[1024];
  const int t = threadIdx.x;
  const int per = (n + 1023) / 1024;
  const int a = min(n, t * per), b = min(n, a + per);
  long long sum = 0;
  for (int k = a; k < b; k++) sum += count[k];
  part[t] = sum;
  __syncthreads();
  for (int o = 1; o < 1024; o <<= 1) {
    long long v = t >= o ? part[t - o] : 0;
    __syncthreads();
    part[t] += v;
    __syncthreads();
  }
  long long run = part[t] - sum;
  for (int k = a; k < b; k++) {
    start[k] = (int)run;
    run += count[k];
  }
  if (t == 1023) {
    start[n] = (int)part[1023];
    *total = part[1023];
  }
}

// ------------------------------------------------------------------------
// host helpers shared by the single-system and batch contexts
// ------------------------------------------------------------------------
inline int grid1(size_t n, int b) { return (int)((n + b - 1) / b); }

inline int launch_check(const char* what) {
  hipError_t e = hipGetLastError();
  if (e != hipSuccess) return fail(P2A_ERR_HIP, std::string(what) + ": " + hipGetErrorString(e));
  return P2A_OK;
}

inline void fill_pbc_arg(const PbcHost& h, PbcArg& P) {
  std::memset(&P, 0, sizeof(P));
  P.type = h.type;
  std::memcpy(P.box, h.box, sizeof(P.box));
  std::memcpy(P.inv, h.inv, sizeof(P.inv));
  std::memcpy(P.red, h.reduced, sizeof(P.red));
  std::memcpy(P.ired, h.invReduced, sizeof(P.ired));
  std::memcpy(P.nshift, h.nshift, sizeof(P.nshift));
  std::memcpy(P.shifts, h.shifts, sizeof(P.shifts));
}

// cell part of a SysArg (column grid and list radius are set by the caller)
inline void sys_cell(const double* box9, double rlist, SysArg& S) {
  PbcHost tmp;
  pbc_setup(box9, tmp);
  std::memcpy(S.a, box9, sizeof(S.a));
  std::memcpy(S.inv, tmp.inv, sizeof(S.inv));
  for (int k = 0; k < 9; k++) S.af[k] = (float)box9[k];
  double w[3];
  pbc_widths(box9, w);
  for (int k = 0; k < 3; k++) S.rb[k] = (float)(rlist / w[k]);
}

// columns of a group of n atoms in a cell: side ~ cbrt(8/rho), bounded count
inline void sys_columns(const double* box9, int n, SysArg& S) {
  double w[3];
  pbc_widths(box9, w);
  double vol = std::fabs(w[0] * w[1] * w[2]);
  double rho = std::max(n, 1) / vol;
  double side = std::cbrt((double)CL / rho);
  S.ncx = std::max(1, (int)std::floor(w[0] / side));
  S.ncy = std::max(1, (int)std::floor(w[1] / side));
  // keep the column count bounded for very dilute groups
  while ((int64_t)S.ncx * S.ncy > 4 * (int64_t)n + 16) {
    S.ncx = std::max(1, S.ncx / 2);
    S.ncy = std::max(1, S.ncy / 2);
  }
}

// the chunk size (entries per work item) and the grid of the pair kernel
struct PairTuning {
  int chunk = 64;        // 32 or 64
  // 0: one wave per item; > 0: persistent grid of that many blocks per CU.
  // The persistent grid also bounds the workgroup rows of the final sum.
  int blocks_per_cu = 4;  // the workgroups of the pair kernel one CU holds at once
};

// ------------------------------------------------------------------------
// COORDINATIONNUMBER rows on the cluster list
// ------------------------------------------------------------------------
// cn_i = sum over the environment atoms j of s(r_ij) (ContactMatrix row times
// ONES, symfunc/CoordinationNumbers.cpp:326-372; ContactMatrix::
// calculateWeight, adjmat/ContactMatrix.h:95-121, skips |r|^2 < DBL_EPSILON).
// In the full cluster list every row atom i meets all its partners as the i
// atom of its own items, so lane (ii, jj) sums the s of its pairs over every
// entry kind.  With weights w (wreq: the weight of each requested atom, 0 for
// atoms that are no row) the kernel also forms the vector-Jacobian product
// sum_k w_k dcn_k/dx: F = sum_k w_k cn_k = sum over pairs (w_i + w_j) s_ij,
// so atom i takes (w_i + w_j)(-dfunc d) from each partner and the box
// derivative takes (w_i + w_j) dfunc d (x) d once per unordered pair.
// Band pairs (r^2 within `band` of D_MAX^2) are decided on the spot on the
// reference's FP64 Pbc::distance.
// Outputs: rpart[item][8] (row sums of the i atoms), and with WT fpart
// [item][8][3] and per-workgroup rows (0, six box-derivative products, 0).
constexpr double kContactEps2 = 2.220446049250313e-16;  // DBL_EPSILON

template <int P, bool GENERIC, bool TWO, bool WT, int KIND>
__device__ __forceinline__ void row_term(const PairArgs& A, const P3& pj, const float4& pi, int ii, int jj, int g,
                                         int ai, float wi, const float* __restrict__ swt, const WaveLds& L,
                                         const float4* __restrict__ posJ, const PbcArg* __restrict__ pbc,
                                         PairAcc& acc, float& cnt) {
  const float dx = pj.x - pi.x;
  const float dy = pj.y - pi.y;
  const float dz = pj.z - pi.z;
  const float r2 = fmaf(dx, dx, fmaf(dy, dy, dz * dz));
  const float t = r2 - A.dmax2;
  bool keep = r2 >= A.contact2;
  if constexpr (KIND == 0) keep = keep && jj != ii;
  if constexpr (TWO) keep = keep && (L.sa[g + jj] != ai);
  bool cut = keep && t <= 0.f;
  if (keep && fabsf(t) < A.band) {
    const int meta = L.sm[g + jj];
    const int reqi = __float_as_int(pi.w);
    if (reqi >= 0 && meta >= 0) {
      const int cj = L.sc[meta >> 3] & kCjMask;
      const int reqj = __float_as_int(posJ[cj * CL + (meta & 7)].w);
      if (reqj >= 0) cut = switch_inside(A.sw, exact_r2(pbc, A.pos64, reqi, reqj));
    }
  }
  float sv, dfu;
  pair_switch<P, GENERIC>(A, r2, sv, dfu);
  acc.sv += cut ? sv : 0.f;
  cnt += cut ? 1.f : 0.f;
  if constexpr (WT) {
    const float wg = cut ? (wi + swt[g + jj]) * dfu : 0.f;
    const float gx = wg * dx, gy = wg * dy, gz = wg * dz;
    acc.fx += gx;
    acc.fy += gy;
    acc.fz += gz;
    bool once = KIND == 1;
    if constexpr (KIND == 0) once = jj > ii;
    if (KIND != 2) {
      const float h = once ? 1.f : 0.f;
      acc.vxx = fmaf(h * gx, dx, acc.vxx);
      acc.vxy = fmaf(h * gx, dy, acc.vxy);
      acc.vxz = fmaf(h * gx, dz, acc.vxz);
      acc.vyy = fmaf(h * gy, dy, acc.vyy);
      acc.vyz = fmaf(h * gy, dz, acc.vyz);
      acc.vzz = fmaf(h * gz, dz, acc.vzz);
    }
  }
}

template <int P, bool GENERIC, bool TWO, bool WT>
__global__ __launch_bounds__(PAIR_BLOCK) void k_rows(PairArgs A, const float* __restrict__ wreq,
                                                     const int4* __restrict__ items, const Entry* __restrict__ ents,
                                                     const float4* __restrict__ posI, const float4* __restrict__ posJ,
                                                     const float4* __restrict__ bbI, const float4* __restrict__ bbJ,
                                                     const int* __restrict__ absI, const int* __restrict__ absJ,
                                                     double* __restrict__ rpart, double* __restrict__ fpart,
                                                     double* __restrict__ partial) {
  constexpr int CH = 64;
  constexpr int NW = PAIR_BLOCK / WAVE;
  constexpr int CAP = stage_cap<CH>();
  __shared__ P3 sj[NW][CAP];
  __shared__ short sm[NW][CAP];
  __shared__ int scode[NW][CH];
  __shared__ int sabs[TWO ? NW : 1][TWO ? CAP : 1];
  __shared__ float swt[WT ? NW : 1][WT ? CAP : 1];
  __shared__ double stot[NW][8];
  const int lane = threadIdx.x & 63;
  const int w = threadIdx.x >> 6;
  const int ii = lane & 7, jj = lane >> 3;
  const int nitems = A.nitems_dev ? min(*A.nitems_dev, A.items_cap) : A.nitems;
  const WaveLds L = {sj[w], sm[w], scode[w], TWO ? sabs[w] : nullptr};
  if (lane < 8) stot[w][lane] = 0.0;
  for (int k = blockIdx.x * NW + w; k < nitems; k += gridDim.x * NW) {
    const int item = k;
    const int4 it = items[item];
    const int ci = it.x;
    const int e1 = min(it.z, A.ent_cap);
    const float4 pi = posI[ci * CL + ii];
    const int ai = TWO ? absI[ci * CL + ii] : 0;
    const float4 ic = bbI[2 * ci], ih = bbI[2 * ci + 1];
    const PbcArg* pbc = A.pbc + nsys_of(A.cl_sys, ci);
    Entry my_e = {0, 0.f, 0.f, 0.f};
    if (lane < CH && it.y + lane < e1) my_e = ents[it.y + lane];
    int n0, b1, n1, b2, n2;
    stage_item<TWO, false, CH>(A, my_e, lane, it.y, e1, ic, ih, posJ, bbJ, absJ, L, n0, b1, n1, b2, n2);
    n1 = __builtin_amdgcn_readfirstlane(n1);
    n2 = __builtin_amdgcn_readfirstlane(n2);
    b1 = __builtin_amdgcn_readfirstlane(b1);
    b2 = __builtin_amdgcn_readfirstlane(b2);
    float wi = 0.f;
    if constexpr (WT) {
      // weights of the staged atoms (0 for padding)
      const int nst = b2 + n2;
      for (int p = lane; p < nst; p += WAVE) {
        const int meta = L.sm[p];
        float wv = 0.f;
        if (meta >= 0) {
          const int cj = L.sc[meta >> 3] & kCjMask;
          const int req = __float_as_int(posJ[cj * CL + (meta & 7)].w);
          if (req >= 0) wv = wreq[req];
        }
        swt[w][p] = wv;
      }
      const int reqi = __float_as_int(pi.w);
      wi = reqi >= 0 ? wreq[reqi] : 0.f;
      __builtin_amdgcn_fence(__ATOMIC_SEQ_CST, "wavefront");
      __builtin_amdgcn_wave_barrier();
    }
    const float* sw = WT ? swt[w] : nullptr;
    PairAcc acc = {0.f, 0.f, 0.f, 0.f, 0.f, 0.f, 0.f, 0.f, 0.f, 0.f};
    float cnt = 0.f;
    if (n0) row_term<P, GENERIC, TWO, WT, 0>(A, L.sj[jj], pi, ii, jj, 0, ai, wi, sw, L, posJ, pbc, acc, cnt);
    for (int g = b1; g < b1 + n1; g += CL)
      row_term<P, GENERIC, TWO, WT, 1>(A, L.sj[g + jj], pi, ii, jj, g, ai, wi, sw, L, posJ, pbc, acc, cnt);
    for (int g = b2; g < b2 + n2; g += CL)
      row_term<P, GENERIC, TWO, WT, 2>(A, L.sj[g + jj], pi, ii, jj, g, ai, wi, sw, L, posJ, pbc, acc, cnt);
    // rows of the 8 i atoms: s = stretch * s~ + shift per contributing pair
    const float cs = sum_over_jj(acc.sv), cc = sum_over_jj(cnt);
    if (jj == 0) rpart[(size_t)item * CL + ii] = A.sv_scale * (double)cs + A.sv_shift * (double)cc;
    if constexpr (WT) {
      const float fx = sum_over_jj(acc.fx), fy = sum_over_jj(acc.fy), fz = sum_over_jj(acc.fz);
      if (jj == 0) {
        double* fp = fpart + (size_t)item * (CL * 3) + ii * 3;
        fp[0] = -A.df_scale * (double)fx;
        fp[1] = -A.df_scale * (double)fy;
        fp[2] = -A.df_scale * (double)fz;
      }
      float tv[6];
      tv[0] = wave_sum_f(acc.vxx);
      tv[1] = wave_sum_f(acc.vxy);
      tv[2] = wave_sum_f(acc.vxz);
      tv[3] = wave_sum_f(acc.vyy);
      tv[4] = wave_sum_f(acc.vyz);
      tv[5] = wave_sum_f(acc.vzz);
      if (lane >= 1 && lane < 7) {
        float tq = tv[0];
        for (int q = 1; q < 6; q++)
          if (lane == q + 1) tq = tv[q];
        stot[w][lane] += A.vir_scale * (double)tq;
      }
    }
    // the next item overwrites the staging read above
    __builtin_amdgcn_fence(__ATOMIC_SEQ_CST, "wavefront");
    __builtin_amdgcn_wave_barrier();
  }
  if constexpr (WT) {
    __syncthreads();
    if (w == 0 && lane < 8) {
      double s = 0.0;
      for (int v = 0; v < NW; v++) s += stot[v][lane];
      partial[(size_t)blockIdx.x * 8 + lane] = s;
    }
  }
}

// rows of the requested atoms 0..nrows-1 (the i group): the fixed-order sum of
// their cluster's item rows; atoms beyond nrows (the environment) are skipped
__global__ __launch_bounds__(256) void k_gather_rows(int nslot, const float4* __restrict__ p,
                                                     const int* __restrict__ item_start, int items_cap,
                                                     const double* __restrict__ rpart, int nrows,
                                                     double* __restrict__ rows) {
  const int sidx = blockIdx.x * blockDim.x + threadIdx.x;
  if (sidx >= nslot) return;
  const int c = sidx / CL, r = sidx % CL;
  const int i0 = item_start[c];
  const int i1 = min(item_start[c + 1], items_cap);
  const int req = __float_as_int(p[sidx].w);
  if (req < 0 || req >= nrows) return;
  double sum = 0.0;
  for (int it = i0; it < i1; it++) sum += rpart[(size_t)it * CL + r];
  rows[req] = sum;
}

// FP32 weight of every requested atom: w[k] for the first nrows, 0 after
__global__ void k_row_weights(int nrows, int natoms, const double* __restrict__ w, float* __restrict__ wreq) {
  const int k = blockIdx.x * blockDim.x + threadIdx.x;
  if (k < natoms) wreq[k] = k < nrows ? (float)w[k] : 0.f;
}

inline void pair_args_generic(PairArgs& A);

inline void launch_rows(const PairArgs& A, bool weighted, int grid, hipStream_t s, const float* wreq,
                        const int4* items, const Entry* ents, const float4* posI, const float4* posJ,
                        const float4* bbI, const float4* bbJ, const int* absI, const int* absJ, double* rpart,
                        double* fpart, double* partial) {
  const p2a_switch& sw = A.sw;
#define P2A_R(P, GEN, TWO, WT)                                                                                   \
  hipLaunchKernelGGL((k_rows<P, GEN, TWO, WT>), dim3(grid), dim3(PAIR_BLOCK), 0, s, A, wreq, items, ents, posI,   \
                     posJ, bbI, bbJ, absI, absJ, rpart, fpart, partial)
#define P2A_RSW(P, GEN)                                        \
  do {                                                         \
    if (A.twogroup) {                                          \
      if (weighted) P2A_R(P, GEN, true, true);                 \
      else P2A_R(P, GEN, true, false);                         \
    } else {                                                   \
      if (weighted) P2A_R(P, GEN, false, true);                \
      else P2A_R(P, GEN, false, false);                        \
    }                                                          \
  } while (0)
  if (sw.type == P2A_SW_RATIONAL_FIXED && sw.nn / 2 == 3) P2A_RSW(3, false);
  else if (sw.type == P2A_SW_RATIONAL_FIXED && sw.nn / 2 == 4) P2A_RSW(4, false);
  else {
    // the other switches run the generic evaluation
    PairArgs G = A;
    pair_args_generic(G);
    const PairArgs& A = G;
    P2A_RSW(1, true);
  }
#undef P2A_RSW
#undef P2A_R
}

// launch of the pair kernel for a switch, group mode and tuning
inline void launch_pairs(const PairArgs& A, int grid, hipStream_t s, bool value_only, int chunk, const int4* items,
                         const Entry* ents, const float4* posI, const float4* posJ, const float4* bbI,
                         const float4* bbJ, const int* absI, const int* absJ, double* fpart, double* partial) {
  const p2a_switch& sw = A.sw;
#define P2A_K(P, GEN, TWO, VO, CH)                                                                                \
  hipLaunchKernelGGL((k_pairs<P, GEN, TWO, VO, CH>), dim3(grid), dim3(PAIR_BLOCK), 0, s, A, items, ents, posI,      \
                     posJ, bbI, bbJ, absI, absJ, fpart, partial)
#define P2A_KCH(P, GEN, TWO, VO)          \
  do {                                    \
    if (chunk == 32) P2A_K(P, GEN, TWO, VO, 32); \
    else P2A_K(P, GEN, TWO, VO, 64);      \
  } while (0)
#define P2A_KSW(P, GEN)                                     \
  do {                                                      \
    if (A.twogroup) {                                       \
      if (value_only) P2A_KCH(P, GEN, true, true);          \
      else P2A_KCH(P, GEN, true, false);                    \
    } else {                                                \
      if (value_only) P2A_KCH(P, GEN, false, true);         \
      else P2A_KCH(P, GEN, false, false);                   \
    }                                                       \
  } while (0)
  if (sw.type == P2A_SW_RATIONAL_FIXED) {
    switch (sw.nn / 2) {
      case 1: P2A_KSW(1, false); break;
      case 2: P2A_KSW(2, false); break;
      case 3: P2A_KSW(3, false); break;
      case 4: P2A_KSW(4, false); break;
      case 5: P2A_KSW(5, false); break;
      case 6: P2A_KSW(6, false); break;
      default: P2A_KSW(1, true); break;
    }
  } else {
    P2A_KSW(1, true);
  }
#undef P2A_KSW
#undef P2A_KCH
#undef P2A_K
}

// entry starts (out0) and work-item starts (out1, chunks of `chunk`) of the
// n = *n_dev (or n) i clusters; bsum holds 2 * ceil(ncap / kScanBlock) ints
inline void launch_list_scans(hipStream_t s, int ncap, const int* n_dev, const int* count, int chunk, int* bsum,
                              int* out0, int* out1, int* totals, int slot0, int slot1) {
  if (ncap <= kScanOneMax) {
    hipLaunchKernelGGL(k_scan2_one, dim3(1), dim3(1024), 0, s, ncap, n_dev, count, chunk, out0, out1, totals, slot0,
                       slot1);
    return;
  }
  const int nb = std::max(1, grid1(ncap, kScanBlock));
  hipLaunchKernelGGL(k_scan2_sums, dim3(nb), dim3(256), 0, s, ncap, n_dev, count, chunk, bsum);
  hipLaunchKernelGGL(k_scan2_blocks, dim3(1), dim3(1024), 0, s, nb, ncap, n_dev, bsum, out0, out1, totals, slot0,
                     slot1);
  hipLaunchKernelGGL(k_scan2_apply, dim3(nb), dim3(256), 0, s, ncap, n_dev, count, chunk, (const int*)bsum, out0,
                     out1);
}

// Scale of the FP32 cluster geometry (packed positions, list offsets,
// boxes): 1/R_0 for the fixedRational switches of the pair kernel (x = r^2
// then needs no multiply), 1 otherwise.
inline bool pair_fixed(const p2a_switch& sw) {
  return sw.type == P2A_SW_RATIONAL_FIXED && sw.nn / 2 >= 1 && sw.nn / 2 <= 6;
}
inline double pack_scale(const p2a_switch& sw) { return pair_fixed(sw) ? std::sqrt(sw.invr0_2) : 1.0; }

// FP32 constants of the pair kernels, the cut band, and the FP64 factors
// that turn the per-item FP32 sums into outputs: for fixedRational<N> the
// kernels sum sv = 1/(1+x^(N/2)) and dfu = x^(N/2-1) sv^2, so that
// value = stretch sum(sv) + shift npair and dfunc = -N/r0^2 stretch dfu
// (SwitchingFunction.cpp:203-215); the generic switches sum s and dfunc.
// The FP32 separations are pack_scale times the real ones, hence the
// 1/pack_scale of the derivative factor and 1/pack_scale^2 of the virial's.
inline void pair_args_switch(const p2a_switch& sw, PairArgs& A) {
  const double k = pack_scale(sw), k2 = k * k;
  A.dmax2 = (float)(sw.dmax_2 * k2);
  const double pr = (sw.dmax + 1e-4) * k;
  A.prune2 = (float)(pr * pr);
  // 1.8x the worst-case FP32 error of r^2 from cluster-local coordinates
  A.band = 4.0e-6f * A.dmax2;
  // the band as two thresholds, each rounded outwards
  A.cut_lo = std::nextafter((float)((double)A.dmax2 - (double)A.band), 0.0f);
  A.cut_hi = std::nextafter((float)((double)A.dmax2 + (double)A.band), HUGE_VALF);
  A.contact2 = (float)(kContactEps2 * k2);
  const bool fixed = pair_fixed(sw);
  A.invr0_2 = fixed ? 1.0f : (float)sw.invr0_2;
  A.r2_real = 1.0 / k2;
  A.sv_scale = fixed ? sw.stretch : 1.0;
  A.sv_shift = fixed ? sw.shift : 0.0;
  const double dfs = fixed ? -(double)sw.nn * sw.invr0_2 * sw.stretch : 1.0;
  A.df_scale = dfs / k;
  A.vir_scale = dfs / k2;
  A.sw = sw;
}

// the factors of pair_args_switch for a kernel that evaluates the switch
// generically (switch_calc_sqr) on the same geometry
inline void pair_args_generic(PairArgs& A) {
  const double k2 = 1.0 / A.r2_real;
  A.sv_scale = 1.0;
  A.sv_shift = 0.0;
  A.df_scale = 1.0 / std::sqrt(k2);
  A.vir_scale = 1.0 / k2;
}

}  // namespace

}  // namespace p2a
