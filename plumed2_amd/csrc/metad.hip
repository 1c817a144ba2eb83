// METAD bias grid on gfx950: hill deposition, well-tempered heights, spline
// bias/force, the fused per-step METAD kernel and the force chain rule.
//
// What each piece replaces in the reference:
//   k_deposit_seq   MetaD::update -> getHeight -> addGaussian -> evaluate-
//                   GaussianAndDerivatives -> Grid::addValueAndDerivatives
//                   (src/bias/MetaD.cpp:2377-2470, 2280-2305, 1886-1937,
//                   2186-2278; src/tools/Grid.cpp:55-106, 904-910), looping
//                   over many hills inside one persistent workgroup because
//                   each well-tempered height depends on the grid left by the
//                   previous hill
//   k_add_hills     addGaussian for hills of known heights (HILLS restart,
//                   MetaD::readGaussians :1807-1829), all hills in parallel
//   spline()        GridBase::getValueAndDerivatives spline branch
//                   (src/tools/Grid.cpp:442-505) with getSplineNeighbors
//                   (:328-356)
//   k_metad_step    MetaD::calculate (:2315-2375) + MetaD::update on a
//                   device-resident CV
//   k_apply_force   ActionWithValue::checkForForces daxpy
//                   (src/core/ActionWithValue.cpp:337-413)
//
// Layout in HBM: value[node] and gradient[node*ncv + i] in FP64, nodes in the
// column-major order of GridBase::getIndex (first CV fastest).
#include <hip/hip_runtime.h>

#include <algorithm>
#include <climits>
#include <cmath>
#include <cstring>
#include <string>
#include <vector>

#include "p2a_internal.h"

// The grid arithmetic follows the reference's FP64 operations one rounding at
// a time: no fused multiply-add contraction.  It also makes every evaluation
// of a hill's Gaussian round the same way wherever it is inlined, which the
// run-wise deposition relies on (a corner's value for a height and the node's
// value on the grid receive bit-identical additions).
#pragma clang fp contract(off)

namespace p2a {

#define P2A_HIPM(x)                                                                                  \
  do {                                                                                               \
    hipError_t e_ = (x);                                                                             \
    if (e_ != hipSuccess) return fail(P2A_ERR_HIP, std::string(#x) + ": " + hipGetErrorString(e_)); \
  } while (0)

constexpr int kMaxCv = 4;
constexpr int kDepositThreads = 256;
constexpr int WAVE_M = 64;

struct GridArg {
  int ncv;
  int pbc[kMaxCv];
  unsigned nbin[kMaxCv];  // nodes per direction
  int nneigh[kMaxCv];     // support half-width in bins
  double min[kMaxCv], dx[kMaxCv];
  double period[kMaxCv], inv_period[kMaxCv];
  double invsigma[kMaxCv];
  double height0, biasf, kbt, stretchA, stretchB;
  double inv_wt;          // 1 / (kbt (biasf - 1)) for biasf > 1, else 1 / kbt
  int welltemp;
  int support;            // prod(2*nneigh+1)
  int may_alias;          // a periodic support wider than the grid visits a node twice
};

// Tools::pbc (Tools.h:547-570)
__device__ __forceinline__ double mpbc1(double x) {
  x += 100.0;
  return (x >= 0) ? x - (double)(int)(x + 0.5) : x - (double)(int)(x - 0.5);
}

// Value::difference (src/core/Value.h:405-416): d2 - d1, wrapped when periodic
__device__ __forceinline__ double vdiff(const GridArg& G, int i, double d1, double d2) {
  if (!G.pbc[i]) return d2 - d1;
  double s = (d2 - d1) * G.inv_period[i];
  s = mpbc1(s);
  return s * G.period[i];
}

// getIndices(x) + getIndex (Grid.cpp:108-132): -1 outside the grid
__device__ long long grid_index_of(const GridArg& G, const double* x, unsigned* ind) {
  for (int i = 0; i < G.ncv; i++) {
    double f = floor((x[i] - G.min[i]) / G.dx[i]);
    if (!(f >= 0.0) || f >= (double)G.nbin[i]) return -1;
    ind[i] = (unsigned)f;
  }
  long long index = ind[G.ncv - 1];
  for (int i = G.ncv - 1; i > 0; --i) index = index * G.nbin[i - 1] + ind[i - 1];
  return index;
}

__device__ __forceinline__ long long flat(const GridArg& G, const unsigned* ind) {
  long long index = ind[G.ncv - 1];
  for (int i = G.ncv - 1; i > 0; --i) index = index * G.nbin[i - 1] + ind[i - 1];
  return index;
}

// spline value and gradient at x; returns false outside the grid
__device__ bool spline(const GridArg& G, const double* __restrict__ val, const double* __restrict__ der,
                       const double* x, double& value, double* dv) {
  unsigned ind[kMaxCv];
  for (int i = 0; i < G.ncv; i++) dv[i] = 0.0;
  value = 0.0;
  if (grid_index_of(G, x, ind) < 0) return false;
  double xfloor[kMaxCv];
  for (int i = 0; i < G.ncv; i++) xfloor[i] = G.min[i] + (double)ind[i] * G.dx[i];
  const unsigned nn = 1u << G.ncv;
  for (unsigned ip = 0; ip < nn; ip++) {
    unsigned ni[kMaxCv];
    unsigned t = ip;
    int cnt = 0;
    for (int j = 0; j < G.ncv; j++) {
      unsigned i0 = t % 2 + ind[j];
      t /= 2;
      if (!G.pbc[j] && i0 == G.nbin[j]) continue;
      if (G.pbc[j] && i0 == G.nbin[j]) i0 = 0;
      ni[cnt++] = i0;
    }
    if (cnt != G.ncv) continue;
    const long long node = flat(G, ni);
    const double grid = val[node];
    double C[kMaxCv], D[kMaxCv], ff = 1.0;
    for (int j = 0; j < G.ncv; j++) {
      // ni[j] equals ind[j] or the next node (possibly wrapped)
      const int x0 = (ni[j] == ind[j]) ? 0 : 1;
      const double dx = G.dx[j];
      const double X = fabs((x[j] - xfloor[j]) / dx - (double)x0);
      const double X2 = X * X, X3 = X2 * X;
      const double yy = fabs(grid) < 0.0000001 ? 0.0 : -der[node * G.ncv + j] / grid;
      const double sg = x0 ? -1.0 : 1.0;
      C[j] = (1.0 - 3.0 * X2 + 2.0 * X3) - sg * yy * (X - 2.0 * X2 + X3) * dx;
      D[j] = (-6.0 * X + 6.0 * X2) - sg * yy * (1.0 - 4.0 * X + 3.0 * X2) * dx;
      D[j] *= sg / dx;
      ff *= C[j];
    }
    for (int j = 0; j < G.ncv; j++) {
      double fd = D[j];
      for (int i = 0; i < G.ncv; i++)
        if (i != j) fd *= C[i];
      dv[j] += grid * fd;
    }
    value += grid * ff;
  }
  return true;
}

// Gaussian of one hill at the node point xx (MetaD.cpp:2244-2275, diagonal),
// in two parts: the shape (scaled distances dp and exp(-dp^2/2); false
// beyond the cutoff, where the hill adds nothing), and its value and
// gradient for the height h.  gaussian() is the two together.
// NCV > 0: that many CVs, known at compile time (the loops unroll and the
// small arrays stay in registers); 0: G.ncv.  The arithmetic is the same.
template <int NCV = 0>
__device__ __forceinline__ bool gauss_shape(const GridArg& G, const double* xx, const double* c, double* dp,
                                            double& e) {
  const int ncv = NCV > 0 ? NCV : G.ncv;
  double dp2 = 0.0;
#pragma unroll
  for (int i = 0; i < (NCV > 0 ? NCV : kMaxCv); i++) {
    if (i >= ncv) break;
    dp[i] = vdiff(G, i, c[i], xx[i]) * G.invsigma[i];
    dp2 += dp[i] * dp[i];
  }
  dp2 *= 0.5;
  e = 0.0;
  if (!(dp2 < 6.25)) return false;
  e = exp(-dp2);
  return true;
}

template <int NCV = 0>
__device__ __forceinline__ double gauss_apply(const GridArg& G, const double* dp, double e, double h, double* d) {
  const int ncv = NCV > 0 ? NCV : G.ncv;
  double bias = h * e;
#pragma unroll
  for (int i = 0; i < (NCV > 0 ? NCV : kMaxCv); i++) {
    if (i >= ncv) break;
    d[i] = -bias * dp[i] * G.invsigma[i] * G.stretchA;
  }
  bias = G.stretchA * bias + h * G.stretchB;
  return bias;
}

template <int NCV = 0>
__device__ __forceinline__ double gaussian(const GridArg& G, const double* xx, const double* c, double h,
                                           double* d) {
  double dp[kMaxCv], e;
#pragma unroll
  for (int i = 0; i < (NCV > 0 ? NCV : kMaxCv); i++) {
    if (i >= (NCV > 0 ? NCV : G.ncv)) break;
    d[i] = 0.0;
  }
  if (!gauss_shape<NCV>(G, xx, c, dp, e)) return 0.0;
  return gauss_apply<NCV>(G, dp, e, h, d);
}

// node of support offset `s` around the bin `ind` (GridBase::getNeighbors,
// Grid.cpp:55-106); false when it falls outside a non-periodic direction
__device__ __forceinline__ bool support_node(const GridArg& G, const unsigned* ind, int s, unsigned* node_ind) {
  for (int i = 0; i < G.ncv; i++) {
    const int w = 2 * G.nneigh[i] + 1;
    const int o = s % w;
    s /= w;
    int i0 = o - G.nneigh[i] + (int)ind[i];
    if (!G.pbc[i] && (i0 < 0 || i0 >= (int)G.nbin[i])) return false;
    if (G.pbc[i] && i0 < 0) i0 = G.nbin[i] - (-i0) % G.nbin[i];
    if (G.pbc[i] && i0 >= (int)G.nbin[i]) i0 %= G.nbin[i];
    node_ind[i] = (unsigned)i0;
  }
  return true;
}

// add one hill of height h centred at c to the grid, all threads of the block
// (distinct nodes per thread unless may_alias, then thread 0 alone)
__device__ void add_hill_block(const GridArg& G, double* __restrict__ val, double* __restrict__ der,
                               const unsigned* ind, const double* c, double h) {
  const int t0 = G.may_alias ? 0 : (int)threadIdx.x;
  const int step = G.may_alias ? 1 : (int)blockDim.x;
  if (G.may_alias && threadIdx.x != 0) return;
  for (int s = t0; s < G.support; s += step) {
    unsigned ni[kMaxCv];
    if (!support_node(G, ind, s, ni)) continue;
    const long long node = flat(G, ni);
    double xx[kMaxCv], d[kMaxCv];
    for (int i = 0; i < G.ncv; i++) xx[i] = G.min[i] + (double)ni[i] * G.dx[i];
    const double b = gaussian(G, xx, c, h, d);
    val[node] += b;
    for (int i = 0; i < G.ncv; i++) der[node * G.ncv + i] += d[i];
  }
}

// MetaD::getHeight (MetaD.cpp:2280-2305): height0 exp(-V / (kbt (biasf - 1)));
// the division is a product with its reciprocal, formed once on the host, so
// that the sequential chain of heights carries no division (the heights move
// by about one rounding of the exponent's argument)
__device__ __forceinline__ double wt_height(const GridArg& G, double vbias) {
  if (!G.welltemp) return G.height0;
  return G.height0 * exp(-vbias * G.inv_wt);
}

// spline value (no gradient) at x from the corner ip of its cell: the term
// `value += grid * ff` of GridBase::getValueAndDerivatives for one ip
__device__ double spline_corner(const GridArg& G, const double* __restrict__ val, const double* __restrict__ der,
                                const double* x, const unsigned* ind, unsigned ip) {
  unsigned ni[kMaxCv];
  unsigned t = ip;
  int cnt = 0;
  for (int j = 0; j < G.ncv; j++) {
    unsigned i0 = t % 2 + ind[j];
    t /= 2;
    if (!G.pbc[j] && i0 == G.nbin[j]) continue;
    if (G.pbc[j] && i0 == G.nbin[j]) i0 = 0;
    ni[cnt++] = i0;
  }
  if (cnt != G.ncv) return 0.0;
  const long long node = flat(G, ni);
  const double grid = val[node];
  double ff = 1.0;
  for (int j = 0; j < G.ncv; j++) {
    const int x0 = (ni[j] == ind[j]) ? 0 : 1;
    const double dx = G.dx[j];
    const double xfloor = G.min[j] + (double)ind[j] * dx;
    const double X = fabs((x[j] - xfloor) / dx - (double)x0);
    const double X2 = X * X, X3 = X2 * X;
    const double yy = fabs(grid) < 0.0000001 ? 0.0 : -der[node * G.ncv + j] / grid;
    const double sg = x0 ? -1.0 : 1.0;
    ff *= (1.0 - 3.0 * X2 + 2.0 * X3) - sg * yy * (X - 2.0 * X2 + X3) * dx;
  }
  return grid * ff;
}

// value and gradient terms of the corner ip of the cell of x: the terms
// `value += grid * ff` and `dv[j] += grid * fd` of GridBase::getValueAndDerivatives
// (Grid.cpp:442-505) for one ip, with the arithmetic of spline() above; a
// corner past the last node of a non-periodic CV gives zeros
__device__ void spline_corner_grad(const GridArg& G, const double* __restrict__ val, const double* __restrict__ der,
                                   const double* x, const unsigned* ind, unsigned ip, double& value, double* dv) {
  value = 0.0;
#pragma unroll
  for (int j = 0; j < kMaxCv; j++) dv[j] = 0.0;
  unsigned ni[kMaxCv];
  unsigned t = ip;
  int cnt = 0;
  for (int j = 0; j < G.ncv; j++) {
    unsigned i0 = t % 2 + ind[j];
    t /= 2;
    if (!G.pbc[j] && i0 == G.nbin[j]) continue;
    if (G.pbc[j] && i0 == G.nbin[j]) i0 = 0;
    ni[cnt++] = i0;
  }
  if (cnt != G.ncv) return;
  const long long node = flat(G, ni);
  const double grid = val[node];
  double C[kMaxCv], D[kMaxCv], ff = 1.0;
#pragma unroll
  for (int j = 0; j < kMaxCv; j++) {
    C[j] = 1.0;
    D[j] = 0.0;
    if (j >= G.ncv) continue;
    const int x0 = (ni[j] == ind[j]) ? 0 : 1;
    const double dx = G.dx[j];
    const double xfloor = G.min[j] + (double)ind[j] * dx;
    const double X = fabs((x[j] - xfloor) / dx - (double)x0);
    const double X2 = X * X, X3 = X2 * X;
    const double yy = fabs(grid) < 0.0000001 ? 0.0 : -der[node * G.ncv + j] / grid;
    const double sg = x0 ? -1.0 : 1.0;
    C[j] = (1.0 - 3.0 * X2 + 2.0 * X3) - sg * yy * (X - 2.0 * X2 + X3) * dx;
    D[j] = (-6.0 * X + 6.0 * X2) - sg * yy * (1.0 - 4.0 * X + 3.0 * X2) * dx;
    D[j] *= sg / dx;
    ff *= C[j];
  }
#pragma unroll
  for (int j = 0; j < kMaxCv; j++) {
    if (j >= G.ncv) continue;
    double fd = D[j];
#pragma unroll
    for (int i = 0; i < kMaxCv; i++)
      if (i != j && i < G.ncv) fd *= C[i];
    dv[j] = grid * fd;
  }
  value = grid * ff;
}

// Sequential well-tempered deposition of many hills in one workgroup: each
// hill's height needs the bias of the grid left by the previous hill
// (MetaD::update -> getHeight).  Per hill, the first wave evaluates the
// 2^ncv spline corners in parallel and sums them in corner order; then the
// whole workgroup adds the hill, one support node per thread.
constexpr int kSeqThreads = 1024;
__global__ __launch_bounds__(kSeqThreads) void k_deposit_seq(GridArg G, double* __restrict__ val,
                                                             double* __restrict__ der, long long nhills,
                                                             const double* __restrict__ centers,
                                                             double* __restrict__ heights, int* __restrict__ err) {
  __shared__ double sh_h;
  __shared__ int sh_ok;
  __shared__ unsigned sh_ind[kMaxCv];
  __shared__ double sh_c[kMaxCv];
  const int t = threadIdx.x;
  const unsigned nn = 1u << G.ncv;
  for (long long k = 0; k < nhills; k++) {
    if (t < WAVE_M) {
      double c[kMaxCv];
      for (int i = 0; i < G.ncv; i++) c[i] = centers[k * G.ncv + i];
      unsigned ind[kMaxCv];
      const int ok = grid_index_of(G, c, ind) >= 0;
      double v = 0.0;
      if (ok && G.welltemp && (unsigned)t < nn) v = spline_corner(G, val, der, c, ind, (unsigned)t);
      // corner order 0, 1, 2, ... as the sequential loop adds them
      double tot = 0.0;
      for (unsigned q = 0; q < nn; q++) tot += __shfl(v, (int)q, WAVE_M);
      if (t == 0) {
        const double h = (ok && G.welltemp) ? wt_height(G, tot) : G.height0;
        if (!ok) atomicCAS(err, 0, (int)(1 + (k < 0x3fffffff ? k : 0x3fffffff)));
        for (int i = 0; i < G.ncv; i++) {
          sh_ind[i] = ind[i];
          sh_c[i] = c[i];
        }
        sh_h = h;
        sh_ok = ok;
        if (heights) heights[k] = ok ? h : 0.0;
      }
    }
    __syncthreads();
    if (!sh_ok) return;  // uniform: every thread reads the same flag
    unsigned ind[kMaxCv];
    double c[kMaxCv];
    for (int i = 0; i < G.ncv; i++) {
      ind[i] = sh_ind[i];
      c[i] = sh_c[i];
    }
    add_hill_block(G, val, der, ind, c, sh_h);
    __syncthreads();
  }
}

// Two non-periodic CVs: the same sequential deposition with the grid around
// the walk held in LDS (a tile of tw x th <= kTileNodes nodes, shaped like
// the hill support: value and the two gradient components), in runs of up
// to kRun consecutive hills that the tile holds (support and upper spline
// corner).  A run's heights are found by wave 0 alone (lane 4r+q holds
// corner q of hill r: hill k's height comes from its four corners, then
// every later corner in hill k's support adds hill k's Gaussian), from the
// shapes of each hill at the corners of the later hills of the run, which
// the whole workgroup evaluates beforehand.  Runs are pipelined: while waves
// 1-15 add the current run to the tile (one node per thread, the run's hills
// in order), wave 0 finds the heights of the next run, whose corners start
// from the tile as it was before the current run plus the current run's
// hills added in order; only a run that needs another tile (the old one is
// written back first) waits for the current run.  Every node, and every
// corner read for a height, receives the same FP64 additions in the same
// order as hill-by-hill deposition, so grids and heights are identical to it.
constexpr int kTileNodes = 5120;
constexpr int kRun = 16;  // hills per run: 4 corners each on one wave
constexpr int kFacLen = 160;  // tile width + height at most: factor tables

// The 2-CV kernel takes the Gaussian as a product of one factor per CV,
// exp(-dp0^2/2) exp(-dp1^2/2) (the same value as exp(-dp2) up to rounding),
// so that a run's deposition evaluates one exponential per hill and grid
// column or row; the support and the cutoff test are those of gauss_shape.
__device__ __forceinline__ double gfac(double x) { return exp(-0.5 * (x * x)); }
__device__ __forceinline__ double gscaled(const GridArg& G, int i, double c, double x) {
  return vdiff(G, i, c, x) * G.invsigma[i];
}
__device__ __forceinline__ bool gauss_shape_f2(const GridArg& G, const double* xx, const double* c, double* dp,
                                               double& e) {
  dp[0] = gscaled(G, 0, c[0], xx[0]);
  dp[1] = gscaled(G, 1, c[1], xx[1]);
  double dp2 = dp[0] * dp[0];
  dp2 += dp[1] * dp[1];
  dp2 *= 0.5;
  e = 0.0;
  if (!(dp2 < 6.25)) return false;
  e = gfac(dp[0]) * gfac(dp[1]);
  return true;
}

__device__ __forceinline__ double spline_corner_vals(const GridArg& G, const double* x, const unsigned* ind,
                                                     const unsigned* ni, double grid, const double* der) {
  double ff = 1.0;
  for (int j = 0; j < 2; j++) {
    const int x0 = (ni[j] == ind[j]) ? 0 : 1;
    const double dx = G.dx[j];
    const double xfloor = G.min[j] + (double)ind[j] * dx;
    const double X = fabs((x[j] - xfloor) / dx - (double)x0);
    const double X2 = X * X, X3 = X2 * X;
    const double yy = fabs(grid) < 0.0000001 ? 0.0 : -der[j] / grid;
    const double sg = x0 ? -1.0 : 1.0;
    ff *= (1.0 - 3.0 * X2 + 2.0 * X3) - sg * yy * (X - 2.0 * X2 + X3) * dx;
  }
  return grid * ff;
}

// value of lane `l` (uniform) of a double, through the scalar unit
__device__ __forceinline__ double readlane_d(double v, int l) {
  const long long b = __double_as_longlong(v);
  const int lo = __builtin_amdgcn_readlane((int)(b & 0xffffffffll), l);
  const int hi = __builtin_amdgcn_readlane((int)(b >> 32), l);
  return __longlong_as_double(((long long)hi << 32) | (unsigned int)lo);
}

__global__ __launch_bounds__(kSeqThreads) void k_deposit_seq_tile2(GridArg G, double* __restrict__ val,
                                                                   double* __restrict__ der, long long nhills,
                                                                   const double* __restrict__ centers,
                                                                   double* __restrict__ heights,
                                                                   int* __restrict__ err, int tw, int th) {
  __shared__ double tv[kTileNodes], tdx[kTileNodes], tdy[kTileNodes];
  // two runs: the current one (heights known, not yet on the tile) and the next
  __shared__ double sh_c[2][kRun][2], sh_h[2][kRun];
  __shared__ int sh_ind[2][kRun][2];
  // shape of hill j at corner lane l (e, negative: no contribution; scaled
  // distances): of the next run's hills (j < r), or of the current run's
  // hills at the next run's corners
  __shared__ double sh_e[kRun][4 * kRun];
  // the same for the current run's hills at the next run's corners
  __shared__ double sh_ex[kRun][4 * kRun];
  // per hill of the current run: its factor at each column of the run's
  // box, then at each row (sh_box: the box)
  __shared__ double sh_f[kRun][kFacLen];
  __shared__ int sh_box[4];
  // value and gradient at each corner of the next run before its own hills
  __shared__ double sh_y[4 * kRun][3];
  __shared__ int sh_nb, sh_fit, sh_nox, sh_noy;
  const int t = threadIdx.x;
  const int lane = t & (WAVE_M - 1);
  const bool wave0 = t < WAVE_M;
  const int nx = (int)G.nbin[0], ny = (int)G.nbin[1];
  const int nn0 = G.nneigh[0], nn1 = G.nneigh[1];
  int ox = -1, oy = -1;  // tile origin in nodes (-1: no tile loaded)
  auto flush = [&]() {
    if (ox < 0) return;
    for (int j = t; j < tw * th; j += blockDim.x) {
      const int lx = j % tw, ly = j / tw;
      const long long node = (long long)(oy + ly) * nx + (ox + lx);
      const int l = lx + tw * ly;
      val[node] = tv[l];
      der[2 * node] = tdx[l];
      der[2 * node + 1] = tdy[l];
    }
  };
  // wave 0: the run of hills from k0 into buffer b; sh_fit = it lies in the
  // current tile, else the tile centred on its first hill is sh_nox/noy.
  // Lane l's centre (hill k0 + l) comes in c0, c1 when `pre`.
  auto decide = [&](long long k0, int b, bool pre, double c0, double c1) {
    const long long k = k0 + lane;
    const bool valid = k < nhills;
    double c[2] = {0.0, 0.0};
    unsigned ind[2] = {0u, 0u};
    int ok = 0;
    if (valid) {
      if (pre) {
        c[0] = c0;
        c[1] = c1;
      } else {
        c[0] = centers[2 * k];
        c[1] = centers[2 * k + 1];
      }
      ok = grid_index_of(G, c, ind) >= 0;
    }
    // nodes the hill touches: support, and the cell's upper spline corner
    const int xl = max(0, (int)ind[0] - nn0), xh = min(nx - 1, (int)ind[0] + nn0 + 1);
    const int yl = max(0, (int)ind[1] - nn1), yh = min(ny - 1, (int)ind[1] + nn1 + 1);
    const int xl0 = __builtin_amdgcn_readfirstlane(xl), xh0 = __builtin_amdgcn_readfirstlane(xh);
    const int yl0 = __builtin_amdgcn_readfirstlane(yl), yh0 = __builtin_amdgcn_readfirstlane(yh);
    const int i00 = __builtin_amdgcn_readfirstlane((int)ind[0]);
    const int i01 = __builtin_amdgcn_readfirstlane((int)ind[1]);
    const bool fit0 = ox >= 0 && xl0 >= ox && xh0 < ox + tw && yl0 >= oy && yh0 < oy + th;
    // a tile centred on the hill holds it whenever nneigh + 2 <= tile / 2,
    // which the launch guarantees; checked anyway before any LDS index
    const int nox = max(0, min(nx - tw, i00 - tw / 2));
    const int noy = max(0, min(ny - th, i01 - th / 2));
    const int bx = fit0 ? ox : nox, by = fit0 ? oy : noy;
    const bool fits = lane < kRun && valid && ok && xl >= bx && xh < bx + tw && yl >= by && yh < by + th;
    const unsigned long long bad = __ballot(!fits);
    const int nb = min(kRun, __ffsll((long long)bad) - 1);
    if (lane < nb) {
      sh_c[b][lane][0] = c[0];
      sh_c[b][lane][1] = c[1];
      sh_ind[b][lane][0] = (int)ind[0];
      sh_ind[b][lane][1] = (int)ind[1];
    }
    if (lane == 0) {
      sh_nb = nb;
      sh_fit = fit0;
      sh_nox = nox;
      sh_noy = noy;
      // a first hill that no tile can hold lies outside the grid
      if (nb == 0 && k0 < nhills) atomicCAS(err, 0, (int)(1 + (k0 < 0x3fffffff ? k0 : 0x3fffffff)));
    }
  };
  // the shapes of the hills j < nsrc of run bs at the corners of the hills
  // r < ndst of run bd (only j < r when both are the same run)
  auto shapes = [&](int bs, int nsrc, int bd, int ndst, double (*out)[4 * kRun]) {
    for (int p = t; p < kRun * 4 * kRun; p += blockDim.x) {
      const int j = p / (4 * kRun), l = p % (4 * kRun);
      const int r = l >> 2, q = l & 3;
      double e = -1.0, dp[2] = {0.0, 0.0};
      if (j < nsrc && r < ndst && (bs != bd || j < r)) {
        const unsigned n0 = (unsigned)sh_ind[bd][r][0] + (q & 1), n1 = (unsigned)sh_ind[bd][r][1] + (q >> 1);
        if (n0 < G.nbin[0] && n1 < G.nbin[1] && abs((int)n0 - sh_ind[bs][j][0]) <= nn0 &&
            abs((int)n1 - sh_ind[bs][j][1]) <= nn1) {
          const double xx[2] = {G.min[0] + (double)n0 * G.dx[0], G.min[1] + (double)n1 * G.dx[1]};
          const double cj[2] = {sh_c[bs][j][0], sh_c[bs][j][1]};
          double ej;
          if (gauss_shape_f2(G, xx, cj, dp, ej)) e = ej;
        }
      }
      out[j][l] = e;
    }
  };
  // threads t < 4 kRun: value and gradient at corner t of run b, from the
  // tile, or from the grid in memory for a node outside the tile
  auto snapshot = [&](int b, int nb) {
    if (t < 4 * kRun) {
      const int r = t >> 2, q = t & 3;
      double y0 = 0.0, y1 = 0.0, y2 = 0.0;
      if (r < nb) {
        const unsigned n0 = (unsigned)sh_ind[b][r][0] + (q & 1), n1 = (unsigned)sh_ind[b][r][1] + (q >> 1);
        if (n0 < G.nbin[0] && n1 < G.nbin[1]) {
          if ((int)n0 >= ox && (int)n0 < ox + tw && (int)n1 >= oy && (int)n1 < oy + th) {
            const int l = (int)(n0 - ox) + tw * (int)(n1 - oy);
            y0 = tv[l];
            y1 = tdx[l];
            y2 = tdy[l];
          } else {
            const long long node = (long long)n1 * nx + n0;
            y0 = val[node];
            y1 = der[2 * node];
            y2 = der[2 * node + 1];
          }
        }
      }
      sh_y[t][0] = y0;
      sh_y[t][1] = y1;
      sh_y[t][2] = y2;
    }
  };
  // wave 0: the heights of run b (nb hills from hill kfirst), from sh_y plus
  // the terms of the nprev hills of run 1 - b (sh_ex, added in order), and
  // the run's own shapes
  auto chain = [&](int b, int nb, long long kfirst, int nprev) {
    const int r = lane >> 2, q = lane & 3;
    const bool mine = r < nb;
    double c[2] = {0.0, 0.0};
    unsigned ind[2] = {0u, 0u}, ni[2] = {0u, 0u};
    bool has = false;
    double acc0 = 0.0, acc1 = 0.0, acc2 = 0.0;
    if (mine) {
      c[0] = sh_c[b][r][0];
      c[1] = sh_c[b][r][1];
      ind[0] = (unsigned)sh_ind[b][r][0];
      ind[1] = (unsigned)sh_ind[b][r][1];
      // corner q of the cell (Grid.cpp:442-505): absent past the last node
      ni[0] = ind[0] + (q & 1);
      ni[1] = ind[1] + (q >> 1);
      has = ni[0] < G.nbin[0] && ni[1] < G.nbin[1];
      if (has) {
        acc0 = sh_y[lane][0];
        acc1 = sh_y[lane][1];
        acc2 = sh_y[lane][2];
        const int pb = 1 - b;
        for (int j0 = 0; j0 < nprev; j0 += 4) {
          double tb[4], td0[4], td1[4];
          bool add[4];
#pragma unroll
          for (int u = 0; u < 4; u++) {
            const int j = min(j0 + u, nprev - 1);
            const double e = sh_ex[j][lane];
            add[u] = j0 + u < nprev && e >= 0.0;
            const double dp[2] = {gscaled(G, 0, sh_c[pb][j][0], G.min[0] + (double)ni[0] * G.dx[0]),
                                  gscaled(G, 1, sh_c[pb][j][1], G.min[1] + (double)ni[1] * G.dx[1])};
            double d[2];
            tb[u] = gauss_apply<2>(G, dp, e, sh_h[pb][j], d);
            td0[u] = d[0];
            td1[u] = d[1];
          }
#pragma unroll
          for (int u = 0; u < 4; u++) {
            if (add[u]) {
              acc0 += tb[u];
              acc1 += td0[u];
              acc2 += td1[u];
            }
          }
        }
      }
    }
    // the terms of spline_corner_vals that depend on the position only
    double X[2] = {0.0, 0.0}, X2[2] = {0.0, 0.0}, X3[2] = {0.0, 0.0}, sg[2] = {1.0, 1.0};
#pragma unroll
    for (int j = 0; j < 2; j++) {
      const int x0 = (ni[j] == ind[j]) ? 0 : 1;
      const double dx = G.dx[j];
      const double xfloor = G.min[j] + (double)ind[j] * dx;
      X[j] = fabs((c[j] - xfloor) / dx - (double)x0);
      X2[j] = X[j] * X[j];
      X3[j] = X2[j] * X[j];
      sg[j] = x0 ? -1.0 : 1.0;
    }
    // this lane's corner coordinates (the grid node of corner q of hill r)
    const double xn0 = G.min[0] + (double)ni[0] * G.dx[0], xn1 = G.min[1] + (double)ni[1] * G.dx[1];
    for (int k = 0; k < nb; k++) {
      // hill k's shape at this lane's corner does not depend on hill k's
      // height: read and scaled before the height is found
      const double ek = sh_e[k][lane];
      const double dpk[2] = {gscaled(G, 0, sh_c[b][k][0], xn0), gscaled(G, 1, sh_c[b][k][1], xn1)};
      // corner term of this lane (used by the lanes of hill k only): the
      // arithmetic of spline_corner_vals
      double v = 0.0;
      if (G.welltemp && has && r == k) {
        const double grid = acc0;
        const double dg[2] = {acc1, acc2};
        double ff = 1.0;
#pragma unroll
        for (int j = 0; j < 2; j++) {
          const double yy = fabs(grid) < 0.0000001 ? 0.0 : -dg[j] / grid;
          ff *= (1.0 - 3.0 * X2[j] + 2.0 * X3[j]) - sg[j] * yy * (X[j] - 2.0 * X2[j] + X3[j]) * G.dx[j];
        }
        v = grid * ff;
      }
      // the four corners of hill k in corner order, on every lane
      double tot = 0.0;
#pragma unroll
      for (int qq = 0; qq < 4; qq++) tot += readlane_d(v, 4 * k + qq);
      const double h = G.welltemp ? wt_height(G, tot) : G.height0;
      if (lane == 4 * k) {
        sh_h[b][k] = h;
        if (heights) heights[kfirst + k] = h;
      }
      if (r > k && mine) {
        if (ek >= 0.0) {
          double d[2];
          const double bb = gauss_apply<2>(G, dpk, ek, h, d);
          acc0 += bb;
          acc1 += d[0];
          acc2 += d[1];
        }
      }
    }
  };
  // run b (nb hills) onto the tile: nodes t0, t0 + step, ..., hills in order
  // the box and the factor tables of run b (all threads; a barrier follows)
  auto factors = [&](int b, int nb) {
    // nodes the run deposits on: ind -/+ nneigh, inside the grid
    int bx0 = INT_MAX, bx1 = -1, by0 = INT_MAX, by1 = -1;
    for (int k = 0; k < nb; k++) {
      bx0 = min(bx0, max(0, sh_ind[b][k][0] - nn0));
      bx1 = max(bx1, min(nx - 1, sh_ind[b][k][0] + nn0));
      by0 = min(by0, max(0, sh_ind[b][k][1] - nn1));
      by1 = max(by1, min(ny - 1, sh_ind[b][k][1] + nn1));
    }
    const int bw = bx1 - bx0 + 1, bh = by1 - by0 + 1;  // bw + bh <= tw + th <= kFacLen
    if (t == 0) {
      sh_box[0] = bx0;
      sh_box[1] = by0;
      sh_box[2] = bw;
      sh_box[3] = bh;
    }
    for (int p = t; p < nb * (bw + bh); p += blockDim.x) {
      const int k = p / (bw + bh), m = p % (bw + bh);
      const int i = m < bw ? 0 : 1;
      const int n = i == 0 ? bx0 + m : by0 + (m - bw);
      sh_f[k][m] = gfac(gscaled(G, i, sh_c[b][k][i], G.min[i] + (double)n * G.dx[i]));
    }
  };
  // run b (nb hills) onto the tile: nodes t0, t0 + step, ..., hills in order
  // (factors() of the run first)
  auto deposit = [&](int b, int nb, int t0, int step) {
    const int bx0 = sh_box[0], by0 = sh_box[1], bw = sh_box[2], bh = sh_box[3];
    for (int j = t0; j < bw * bh; j += step) {
      const int cx = j % bw, cy = j / bw;
      const int n0 = bx0 + cx, n1 = by0 + cy;
      const int l = (n0 - ox) + tw * (n1 - oy);
      const double xx[2] = {G.min[0] + (double)n0 * G.dx[0], G.min[1] + (double)n1 * G.dx[1]};
      double v = tv[l], gx = tdx[l], gy = tdy[l];
      for (int k = 0; k < nb; k++) {
        if (abs(n0 - sh_ind[b][k][0]) <= nn0 && abs(n1 - sh_ind[b][k][1]) <= nn1) {
          // gauss_shape_f2 with the factors from the tables, then gauss_apply
          double dp[2];
          dp[0] = gscaled(G, 0, sh_c[b][k][0], xx[0]);
          dp[1] = gscaled(G, 1, sh_c[b][k][1], xx[1]);
          double dp2 = dp[0] * dp[0];
          dp2 += dp[1] * dp[1];
          dp2 *= 0.5;
          double bb = 0.0, d[2] = {0.0, 0.0};
          if (dp2 < 6.25) bb = gauss_apply<2>(G, dp, sh_f[k][cx] * sh_f[k][bw + cy], sh_h[b][k], d);
          v += bb;
          gx += d[0];
          gy += d[1];
        }
      }
      tv[l] = v;
      tdx[l] = gx;
      tdy[l] = gy;
    }
  };

  // the old tile written back, the tile at sh_nox/noy read
  auto reload = [&]() {
    flush();
    // the tile written above may overlap the one read below
    __threadfence();
    __syncthreads();
    ox = sh_nox;
    oy = sh_noy;
    for (int j = t; j < tw * th; j += blockDim.x) {
      const int lx = j % tw, ly = j / tw;
      const long long node = (long long)(oy + ly) * nx + (ox + lx);
      const int l = lx + tw * ly;
      tv[l] = val[node];
      tdx[l] = der[2 * node];
      tdy[l] = der[2 * node + 1];
    }
    __syncthreads();
  };

  long long k0 = 0;  // first hill not yet in a run
  int cb = 0;        // buffer of the current run
  int nbc = 0;       // hills of the current run (heights known, not on the tile)
  // wave 0: the centres of the hills the next run decision starts with,
  // loaded while the heights of the next run are found
  bool pre = false;
  double pc0 = 0.0, pc1 = 0.0;
  auto prefetch = [&](long long kn) {
    pre = true;
    const long long k = kn + lane;
    if (lane < kRun && k < nhills) {
      pc0 = centers[2 * k];
      pc1 = centers[2 * k + 1];
    }
  };
  while (true) {
    if (wave0) decide(k0, 1 - cb, pre, pc0, pc1);
    __syncthreads();
    const int nbn = sh_nb;  // uniform; 0: no hill left, or the next is outside the grid
    const bool fit = sh_fit;
    if (nbc > 0 && nbn > 0) {
      // the next run's corners: the grid before the current run (tile, or
      // memory outside it, which the current run does not touch), then the
      // current run's hills in order (at the start of the chain); the next
      // run's own shapes; the current run's factors
      shapes(cb, nbc, 1 - cb, nbn, sh_ex);
      shapes(1 - cb, nbn, 1 - cb, nbn, sh_e);
      snapshot(1 - cb, nbn);
      factors(cb, nbc);
      __syncthreads();
      // wave 0: the next run's heights; the other waves: the current run
      if (wave0) {
        prefetch(k0 + nbn);
        chain(1 - cb, nbn, k0, nbc);
      } else {
        deposit(cb, nbc, t - WAVE_M, (int)blockDim.x - WAVE_M);
      }
      __syncthreads();
      // the next run lies in another tile
      if (!fit) reload();
    } else {
      // the first run, the end, or a hill outside the grid: in sequence
      if (nbc > 0) {
        factors(cb, nbc);
        __syncthreads();
        deposit(cb, nbc, t, blockDim.x);
        __syncthreads();
      }
      if (nbn == 0) break;
      if (!fit) reload();
      shapes(1 - cb, nbn, 1 - cb, nbn, sh_e);
      snapshot(1 - cb, nbn);
      __syncthreads();
      if (wave0) {
        prefetch(k0 + nbn);
        chain(1 - cb, nbn, k0, 0);
      }
      __syncthreads();
    }
    cb = 1 - cb;
    nbc = nbn;
    k0 += nbn;
  }
  flush();
}

// hills of given heights, in parallel (one block per hill)
__global__ __launch_bounds__(kDepositThreads) void k_add_hills(GridArg G, double* __restrict__ val,
                                                                double* __restrict__ der, long long nhills,
                                                                const double* __restrict__ centers,
                                                                const double* __restrict__ heights,
                                                                int* __restrict__ err) {
  const long long k = blockIdx.x;
  if (k >= nhills) return;
  double c[kMaxCv];
  unsigned ind[kMaxCv];
  for (int i = 0; i < G.ncv; i++) c[i] = centers[k * G.ncv + i];
  if (grid_index_of(G, c, ind) < 0) {
    if (threadIdx.x == 0) atomicCAS(err, 0, 1);
    return;
  }
  const double h = heights[k];
  for (int s = threadIdx.x; s < G.support; s += blockDim.x) {
    unsigned ni[kMaxCv];
    if (!support_node(G, ind, s, ni)) continue;
    const long long node = flat(G, ni);
    double xx[kMaxCv], d[kMaxCv];
    for (int i = 0; i < G.ncv; i++) xx[i] = G.min[i] + (double)ni[i] * G.dx[i];
    const double b = gaussian(G, xx, c, h, d);
    unsafeAtomicAdd(&val[node], b);
    for (int i = 0; i < G.ncv; i++) unsafeAtomicAdd(&der[node * G.ncv + i], d[i]);
  }
}

__global__ void k_bias(GridArg G, const double* __restrict__ val, const double* __restrict__ der, long long n,
                       const double* __restrict__ cv, double* __restrict__ bias, double* __restrict__ dout,
                       int* __restrict__ err) {
  const long long p = (long long)blockIdx.x * blockDim.x + threadIdx.x;
  if (p >= n) return;
  double x[kMaxCv], v, dv[kMaxCv];
  for (int i = 0; i < G.ncv; i++) x[i] = cv[p * G.ncv + i];
  if (!spline(G, val, der, x, v, dv)) {
    atomicCAS(err, 0, 1);
    v = 0.0;
  }
  bias[p] = v;
  if (dout)
    for (int i = 0; i < G.ncv; i++) dout[p * G.ncv + i] = dv[i];
}

// MetaD::calculate + update for one step, one workgroup
__global__ __launch_bounds__(kDepositThreads) void k_metad_step(GridArg G, double* __restrict__ val,
                                                                 double* __restrict__ der,
                                                                 const double* __restrict__ cv, int deposit,
                                                                 double* __restrict__ bias,
                                                                 double* __restrict__ force,
                                                                 double* __restrict__ height,
                                                                 int* __restrict__ err) {
  __shared__ double sh_h;
  __shared__ int sh_ok;
  __shared__ unsigned sh_ind[kMaxCv];
  __shared__ double sh_c[kMaxCv];
  if (threadIdx.x < WAVE_M) {
    // lane t < 2^ncv evaluates corner t; the corners are summed in corner
    // order, as spline() adds them
    const int t = threadIdx.x;
    double c[kMaxCv] = {0.0, 0.0, 0.0, 0.0};
    unsigned ind[kMaxCv] = {0u, 0u, 0u, 0u};
    for (int i = 0; i < G.ncv; i++) c[i] = cv[i];
    int ok = grid_index_of(G, c, ind) >= 0;
    const unsigned nn = 1u << G.ncv;
    double cvv = 0.0, cdv[kMaxCv];
    if (ok && (unsigned)t < nn) {
      spline_corner_grad(G, val, der, c, ind, (unsigned)t, cvv, cdv);
    } else {
#pragma unroll
      for (int j = 0; j < kMaxCv; j++) cdv[j] = 0.0;
    }
    double v = 0.0, dv[kMaxCv] = {0.0, 0.0, 0.0, 0.0};
    for (unsigned q = 0; q < nn; q++) {
      v += __shfl(cvv, (int)q, WAVE_M);
#pragma unroll
      for (int j = 0; j < kMaxCv; j++) dv[j] += __shfl(cdv[j], (int)q, WAVE_M);
    }
    if (t == 0) {
      if (!ok) atomicCAS(err, 0, 1);
      // MetaD.cpp:2348-2350: no bias (and no force) when BIASFACTOR=1
      const bool nobias = (G.biasf == 1.0);
      if (bias) *bias = nobias ? 0.0 : v;
      if (force)
        for (int i = 0; i < G.ncv; i++) force[i] = nobias ? 0.0 : -dv[i];
      // getHeight evaluates the bias of the grid before this hill: the same v
      const double h = wt_height(G, v);
      for (int i = 0; i < G.ncv; i++) {
        sh_ind[i] = ind[i];
        sh_c[i] = c[i];
      }
      sh_h = h;
      sh_ok = ok && deposit;
      if (height) *height = (ok && deposit) ? h : 0.0;
    }
  }
  __syncthreads();
  if (!sh_ok) return;
  unsigned ind[kMaxCv];
  double c[kMaxCv];
  for (int i = 0; i < G.ncv; i++) {
    ind[i] = sh_ind[i];
    c[i] = sh_c[i];
  }
  add_hill_block(G, val, der, ind, c, sh_h);
}

__global__ void k_apply_force(long long natoms, const double* __restrict__ f, const double* __restrict__ deriv,
                              const double* __restrict__ vcv, double* __restrict__ forces,
                              double* __restrict__ virial) {
  const long long i = (long long)blockIdx.x * blockDim.x + threadIdx.x;
  const double fv = *f;
  if (i < 3 * natoms) forces[i] += fv * deriv[i];
  if (virial && vcv && i < 9) virial[i] += fv * vcv[i];
}

}  // namespace p2a

struct p2a_metad {
  int device = 0;
  hipStream_t own_stream = nullptr, stream = nullptr;
  p2a_metad_spec spec{};
  p2a::GridArg G{};
  size_t size = 0;
  double* val = nullptr;
  double* der = nullptr;
  int* err = nullptr;
  double* scratch = nullptr;  // staging for host-array entry points
  size_t scratch_n = 0;
  bool first_step = true;
};

namespace p2a {

static int metad_scratch(p2a_metad* m, size_t n) {
  if (n <= m->scratch_n) return P2A_OK;
  if (m->scratch) (void)hipFree(m->scratch);
  m->scratch = nullptr;
  m->scratch_n = 0;
  if (hipMalloc(&m->scratch, n * sizeof(double)) != hipSuccess) return fail(P2A_ERR_NOMEM, "hipMalloc failed");
  m->scratch_n = n;
  return P2A_OK;
}

static int metad_read_err(p2a_metad* m, const char* what) {
  int e = 0;
  P2A_HIPM(hipMemcpyAsync(&e, m->err, sizeof(int), hipMemcpyDeviceToHost, m->stream));
  P2A_HIPM(hipStreamSynchronize(m->stream));
  if (e) {
    P2A_HIPM(hipMemsetAsync(m->err, 0, sizeof(int), m->stream));
    P2A_HIPM(hipStreamSynchronize(m->stream));
    return fail(P2A_ERR_GRID, std::string(what) + ": Looking for a value outside the grid");
  }
  return P2A_OK;
}

}  // namespace p2a

using namespace p2a;

extern "C" {

int p2a_metad_create(p2a_metad** out, int device, const p2a_metad_spec* spec) {
  if (!out || !spec) return fail(P2A_ERR_ARG, "null argument");
  *out = nullptr;
  const int ncv = spec->ncv;
  if (ncv < 1 || ncv > kMaxCv) return fail(P2A_ERR_UNSUPPORTED, "METAD grid supports 1 to 4 CVs");
  if (spec->pace <= 0) return fail(P2A_ERR_ARG, "PACE should be positive");
  if (spec->biasf < 1.0 && spec->biasf != -1.0) return fail(P2A_ERR_ARG, "well tempered bias factor is nonsensical");
  if (spec->biasf >= 1.0 && spec->kbt <= 0.0)
    return fail(P2A_ERR_ARG,
                "Unless the MD engine passes the temperature to plumed, with well-tempered metad you must specify it using TEMP");
  GridArg G;
  std::memset(&G, 0, sizeof(G));
  G.ncv = ncv;
  size_t size = 1;
  int64_t support = 1;
  for (int i = 0; i < ncv; i++) {
    if (!(spec->gmax[i] > spec->gmin[i])) return fail(P2A_ERR_ARG, "maximum in grid must be larger than minimum");
    if (spec->nbin[i] == 0) return fail(P2A_ERR_ARG, "number of grid points must be greater than zero");
    if (!(spec->sigma[i] > 0.0)) return fail(P2A_ERR_ARG, "SIGMA should be positive");
    G.pbc[i] = spec->periodic[i] != 0;
    G.min[i] = spec->gmin[i];
    G.dx[i] = (spec->gmax[i] - spec->gmin[i]) / (double)spec->nbin[i];
    G.nbin[i] = spec->nbin[i] + (G.pbc[i] ? 0 : 1);
    G.period[i] = spec->gmax[i] - spec->gmin[i];
    G.inv_period[i] = 1.0 / G.period[i];
    G.invsigma[i] = 1.0 / spec->sigma[i];
    // MetaD::getGaussianSupport (MetaD.cpp:1939-1994)
    G.nneigh[i] = (int)std::ceil(std::sqrt(2.0 * 6.25) * spec->sigma[i] / G.dx[i]);
    if (G.pbc[i] && 2 * G.nneigh[i] + 1 > (int)G.nbin[i]) G.may_alias = 1;
    size *= G.nbin[i];
    support *= 2 * G.nneigh[i] + 1;
  }
  if (support > (1 << 30)) return fail(P2A_ERR_UNSUPPORTED, "Gaussian support too large for the grid spacing");
  G.support = (int)support;
  G.height0 = spec->height;
  G.biasf = spec->biasf;
  G.kbt = spec->kbt;
  G.welltemp = spec->biasf >= 1.0;
  G.inv_wt = spec->biasf > 1.0 ? 1.0 / (spec->kbt * (spec->biasf - 1.0)) : (spec->kbt != 0.0 ? 1.0 / spec->kbt : 0.0);
  // dp2cutoff stretch (Tools.h:71-79, MetaD.cpp:762-765)
  G.stretchA = spec->nostretch ? 1.0 : 1.00193418799744762399;
  G.stretchB = spec->nostretch ? 0.0 : -.00193418799744762399;
  int ndev = 0;
  if (hipGetDeviceCount(&ndev) != hipSuccess || ndev == 0) return fail(P2A_ERR_HIP, "no HIP device available");
  if (device < 0 || device >= ndev) return fail(P2A_ERR_ARG, "device index out of range");
  P2A_HIPM(hipSetDevice(device));
  auto* m = new p2a_metad();
  m->device = device;
  m->spec = *spec;
  m->G = G;
  m->size = size;
  bool ok = hipStreamCreateWithFlags(&m->own_stream, hipStreamNonBlocking) == hipSuccess &&
            hipMalloc(&m->val, size * sizeof(double)) == hipSuccess &&
            hipMalloc(&m->der, size * ncv * sizeof(double)) == hipSuccess &&
            hipMalloc(&m->err, sizeof(int)) == hipSuccess;
  if (!ok) {
    p2a_metad_destroy(m);
    return fail(P2A_ERR_NOMEM, "device allocation of the METAD grid failed");
  }
  m->stream = m->own_stream;
  if (hipMemsetAsync(m->val, 0, size * sizeof(double), m->stream) != hipSuccess ||
      hipMemsetAsync(m->der, 0, size * ncv * sizeof(double), m->stream) != hipSuccess ||
      hipMemsetAsync(m->err, 0, sizeof(int), m->stream) != hipSuccess ||
      hipStreamSynchronize(m->stream) != hipSuccess) {
    p2a_metad_destroy(m);
    return fail(P2A_ERR_HIP, "clearing the METAD grid failed");
  }
  *out = m;
  g_last_error.clear();
  return P2A_OK;
}

int p2a_metad_destroy(p2a_metad* m) {
  if (!m) return P2A_OK;
  (void)hipSetDevice(m->device);
  if (m->stream) (void)hipStreamSynchronize(m->stream);
  if (m->val) (void)hipFree(m->val);
  if (m->der) (void)hipFree(m->der);
  if (m->err) (void)hipFree(m->err);
  if (m->scratch) (void)hipFree(m->scratch);
  if (m->own_stream) (void)hipStreamDestroy(m->own_stream);
  delete m;
  return P2A_OK;
}

int p2a_metad_set_stream(p2a_metad* m, void* stream) {
  if (!m) return fail(P2A_ERR_ARG, "null context");
  // NULL selects the default (null) stream, which orders with every blocking stream
  m->stream = (hipStream_t)stream;
  return P2A_OK;
}

int64_t p2a_metad_grid_size(const p2a_metad* m) { return m ? (int64_t)m->size : 0; }

int p2a_metad_deposit_device(p2a_metad* m, int64_t nhills, const double* centers_dev, double* heights_dev) {
  if (!m || (nhills > 0 && !centers_dev)) return fail(P2A_ERR_ARG, "null argument");
  if (nhills <= 0) return P2A_OK;
  P2A_HIPM(hipSetDevice(m->device));
  const GridArg& G = m->G;
  // LDS tile of the 2-CV kernel: shaped like the hill support (a walk
  // crosses it in a number of hills about equal along both CVs), at least
  // two supports plus the spline corner wide, within the grid
  int tw = 0, th = 0;
  bool tile2 = G.ncv == 2 && !G.pbc[0] && !G.pbc[1] && !G.may_alias;
  if (tile2) {
    const int nx = (int)G.nbin[0], ny = (int)G.nbin[1];
    const int need0 = 2 * G.nneigh[0] + 4, need1 = 2 * G.nneigh[1] + 4;
    const double a0 = 2.0 * G.nneigh[0] + 1.0, a1 = 2.0 * G.nneigh[1] + 1.0;
    tw = std::min(nx, std::max(need0, (int)std::floor(std::sqrt(kTileNodes * a0 / a1))));
    th = std::min(ny, kTileNodes / std::max(tw, 1));
    if (th < need1) {
      th = std::min(ny, need1);
      tw = std::min(nx, kTileNodes / th);
    }
    tile2 = tw >= std::min(nx, need0) && th >= std::min(ny, need1) && tw * th <= kTileNodes && tw + th <= kFacLen &&
            need0 <= nx && need1 <= ny;
  }
  if (tile2)
    hipLaunchKernelGGL(k_deposit_seq_tile2, dim3(1), dim3(kSeqThreads), 0, m->stream, m->G, m->val, m->der,
                       (long long)nhills, centers_dev, heights_dev, m->err, tw, th);
  else
    hipLaunchKernelGGL(k_deposit_seq, dim3(1), dim3(kSeqThreads), 0, m->stream, m->G, m->val, m->der,
                       (long long)nhills, centers_dev, heights_dev, m->err);
  P2A_HIPM(hipGetLastError());
  return P2A_OK;
}

int p2a_metad_deposit(p2a_metad* m, int64_t nhills, const double* centers, double* heights) {
  if (!m || (nhills > 0 && !centers)) return fail(P2A_ERR_ARG, "null argument");
  if (nhills <= 0) return P2A_OK;
  P2A_HIPM(hipSetDevice(m->device));
  const int ncv = m->G.ncv;
  int rc;
  if ((rc = metad_scratch(m, (size_t)nhills * (ncv + 1)))) return rc;
  double* c_dev = m->scratch;
  double* h_dev = m->scratch + (size_t)nhills * ncv;
  P2A_HIPM(hipMemcpyAsync(c_dev, centers, (size_t)nhills * ncv * sizeof(double), hipMemcpyHostToDevice, m->stream));
  if ((rc = p2a_metad_deposit_device(m, nhills, c_dev, h_dev))) return rc;
  if (heights)
    P2A_HIPM(hipMemcpyAsync(heights, h_dev, (size_t)nhills * sizeof(double), hipMemcpyDeviceToHost, m->stream));
  return metad_read_err(m, "METAD deposition");
}

int p2a_metad_add_hills(p2a_metad* m, int64_t nhills, const double* centers, const double* heights) {
  if (!m || (nhills > 0 && (!centers || !heights))) return fail(P2A_ERR_ARG, "null argument");
  if (nhills <= 0) return P2A_OK;
  P2A_HIPM(hipSetDevice(m->device));
  const int ncv = m->G.ncv;
  int rc;
  if ((rc = metad_scratch(m, (size_t)nhills * (ncv + 1)))) return rc;
  double* c_dev = m->scratch;
  double* h_dev = m->scratch + (size_t)nhills * ncv;
  P2A_HIPM(hipMemcpyAsync(c_dev, centers, (size_t)nhills * ncv * sizeof(double), hipMemcpyHostToDevice, m->stream));
  P2A_HIPM(hipMemcpyAsync(h_dev, heights, (size_t)nhills * sizeof(double), hipMemcpyHostToDevice, m->stream));
  hipLaunchKernelGGL(k_add_hills, dim3((unsigned)nhills), dim3(kDepositThreads), 0, m->stream, m->G, m->val, m->der,
                     (long long)nhills, c_dev, h_dev, m->err);
  P2A_HIPM(hipGetLastError());
  return metad_read_err(m, "METAD hills");
}

int p2a_metad_bias_device(p2a_metad* m, int64_t n, const double* cv_dev, double* bias_dev, double* der_dev) {
  if (!m || (n > 0 && (!cv_dev || !bias_dev))) return fail(P2A_ERR_ARG, "null argument");
  if (n <= 0) return P2A_OK;
  P2A_HIPM(hipSetDevice(m->device));
  hipLaunchKernelGGL(k_bias, dim3((unsigned)((n + 255) / 256)), dim3(256), 0, m->stream, m->G, m->val, m->der,
                     (long long)n, cv_dev, bias_dev, der_dev, m->err);
  P2A_HIPM(hipGetLastError());
  return P2A_OK;
}

int p2a_metad_bias(p2a_metad* m, int64_t n, const double* cv, double* bias, double* der) {
  if (!m || (n > 0 && (!cv || !bias))) return fail(P2A_ERR_ARG, "null argument");
  if (n <= 0) return P2A_OK;
  P2A_HIPM(hipSetDevice(m->device));
  const int ncv = m->G.ncv;
  int rc;
  if ((rc = metad_scratch(m, (size_t)n * (2 * ncv + 1)))) return rc;
  double* c_dev = m->scratch;
  double* b_dev = c_dev + (size_t)n * ncv;
  double* d_dev = b_dev + n;
  P2A_HIPM(hipMemcpyAsync(c_dev, cv, (size_t)n * ncv * sizeof(double), hipMemcpyHostToDevice, m->stream));
  if ((rc = p2a_metad_bias_device(m, n, c_dev, b_dev, d_dev))) return rc;
  P2A_HIPM(hipMemcpyAsync(bias, b_dev, (size_t)n * sizeof(double), hipMemcpyDeviceToHost, m->stream));
  if (der) P2A_HIPM(hipMemcpyAsync(der, d_dev, (size_t)n * ncv * sizeof(double), hipMemcpyDeviceToHost, m->stream));
  return metad_read_err(m, "METAD bias");
}

int p2a_metad_step_device(p2a_metad* m, int64_t step, const double* cv_dev, double* bias_dev, double* force_dev,
                          double* height_dev) {
  if (!m || !cv_dev) return fail(P2A_ERR_ARG, "null argument");
  P2A_HIPM(hipSetDevice(m->device));
  // MetaD::update (MetaD.cpp:2379-2385): a hill every PACE steps, never on
  // the first step seen
  int deposit = 0;
  if (step % m->spec.pace == 0 && !m->first_step) deposit = 1;
  else m->first_step = false;
  hipLaunchKernelGGL(k_metad_step, dim3(1), dim3(kDepositThreads), 0, m->stream, m->G, m->val, m->der, cv_dev,
                     deposit, bias_dev, force_dev, height_dev, m->err);
  P2A_HIPM(hipGetLastError());
  return P2A_OK;
}

int p2a_metad_check(p2a_metad* m) {
  if (!m) return fail(P2A_ERR_ARG, "null context");
  P2A_HIPM(hipSetDevice(m->device));
  return metad_read_err(m, "METAD");
}

int p2a_metad_get_grid(p2a_metad* m, double* values, double* derivs) {
  if (!m) return fail(P2A_ERR_ARG, "null context");
  P2A_HIPM(hipSetDevice(m->device));
  if (values) P2A_HIPM(hipMemcpyAsync(values, m->val, m->size * sizeof(double), hipMemcpyDeviceToHost, m->stream));
  if (derivs)
    P2A_HIPM(hipMemcpyAsync(derivs, m->der, m->size * m->G.ncv * sizeof(double), hipMemcpyDeviceToHost, m->stream));
  P2A_HIPM(hipStreamSynchronize(m->stream));
  return P2A_OK;
}

int p2a_metad_set_grid(p2a_metad* m, const double* values, const double* derivs) {
  if (!m) return fail(P2A_ERR_ARG, "null context");
  P2A_HIPM(hipSetDevice(m->device));
  if (values) P2A_HIPM(hipMemcpyAsync(m->val, values, m->size * sizeof(double), hipMemcpyHostToDevice, m->stream));
  if (derivs)
    P2A_HIPM(hipMemcpyAsync(m->der, derivs, m->size * m->G.ncv * sizeof(double), hipMemcpyHostToDevice, m->stream));
  P2A_HIPM(hipStreamSynchronize(m->stream));
  return P2A_OK;
}

int p2a_apply_force_device(void* stream, int64_t natoms, const double* f_dev, const double* deriv_dev,
                           const double* virial_cv_dev, double* forces_dev, double* virial_dev) {
  if (!f_dev || !deriv_dev || !forces_dev) return fail(P2A_ERR_ARG, "null argument");
  long long n = 3 * (long long)natoms;
  long long nt = n > 9 ? n : 9;
  hipLaunchKernelGGL(k_apply_force, dim3((unsigned)((nt + 255) / 256)), dim3(256), 0, (hipStream_t)stream,
                     (long long)natoms, f_dev, deriv_dev, virial_cv_dev, forces_dev, virial_dev);
  P2A_HIPM(hipGetLastError());
  return P2A_OK;
}

}  // extern "C"
