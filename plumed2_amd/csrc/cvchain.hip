// DISTANCE between the centres of mass of two atom groups on gfx950, and the
// chain rule that puts a bias force on such a CV back onto the atoms.
//
// What it replaces in the reference (the second CV of a METAD on
// COORDINATION and a distance between two centres):
//   c1: COM ATOMS=...   Center::calculate (src/vatom/Center.cpp:319-445) on
//   c2: COM ATOMS=...   the group made whole by ActionAtomistic::makeWhole
//                       (src/core/ActionAtomistic.cpp:587-607, no MOLINFO)
//   d: DISTANCE ATOMS=c1,c2   Distance::calculate / calculateCV
//                       (src/colvar/Distance.h:141-250): the two centres made
//                       whole, |c2 - c1|, derivatives +-d/|d|, box derivative
//                       by Colvar::setBoxDerivativesNoPbc
//   apply               ActionWithVirtualAtom::apply (src/core/
//                       ActionWithVirtualAtom.cpp:68-124) + ActionWithValue::
//                       checkForForces: F_i += f * w_i * (+-d/|d|); the centres
//                       are linear in the atoms and add no box term
//                       (ActionWithVirtualAtom.h:54-63)
//
// makeWhole is a chain: atom j+1 is placed at x'_j + Pbc::distance(x'_j,
// x_j+1).  Its image differs from the raw position by a lattice vector, and
// Pbc::distance gives the same vector for inputs that differ by a lattice
// vector, so x'_j = x_j + sum_{k<j} t_k with t_k the lattice vector that
// Pbc::distance adds to x_k+1 - x_k.  Each t_k is computed independently (in
// integer lattice coordinates) and the chain becomes an integer prefix sum:
// one workgroup, no sequential walk.  The centre is then a fixed-order FP64
// sum; it differs from the reference's running sum by rounding only.
#include <hip/hip_runtime.h>

#include <algorithm>
#include <cstring>
#include <string>
#include <vector>

#include "coord_kernels.h"

namespace p2a {

constexpr int kCdThreads = 1024;

// integer lattice coordinates of the vector Pbc::distance adds to b - a
__device__ __forceinline__ int3 link_shift(const PbcArg& P, const double* a, const double* b) {
  double d[3];
  pbc_distance(P, a, b, d);
  const double t[3] = {d[0] - (b[0] - a[0]), d[1] - (b[1] - a[1]), d[2] - (b[2] - a[2])};
  // t = n . box (box rows = lattice vectors)  =>  n = t . inv(box)
  int3 n;
  n.x = (int)rint(t[0] * P.inv[0] + t[1] * P.inv[3] + t[2] * P.inv[6]);
  n.y = (int)rint(t[0] * P.inv[1] + t[1] * P.inv[4] + t[2] * P.inv[7]);
  n.z = (int)rint(t[0] * P.inv[2] + t[1] * P.inv[5] + t[2] * P.inv[8]);
  return n;
}

// inclusive scan of int3 over the workgroup
__device__ __forceinline__ int3 block_scan3(int3 v, int* sx, int* sy, int* sz) {
  const int t = threadIdx.x;
  sx[t] = v.x;
  sy[t] = v.y;
  sz[t] = v.z;
  __syncthreads();
  for (int o = 1; o < (int)blockDim.x; o <<= 1) {
    int ax = 0, ay = 0, az = 0;
    if (t >= o) {
      ax = sx[t - o];
      ay = sy[t - o];
      az = sz[t - o];
    }
    __syncthreads();
    sx[t] += ax;
    sy[t] += ay;
    sz[t] += az;
    __syncthreads();
  }
  const int3 r = make_int3(sx[t], sy[t], sz[t]);
  __syncthreads();
  return r;
}

// FP64 sum over the workgroup in a fixed order (shuffles, then waves in order)
__device__ __forceinline__ double block_sum(double v, double* sh) {
  for (int o = 32; o > 0; o >>= 1) v += __shfl_xor(v, o, 64);
  const int lane = threadIdx.x & 63, w = threadIdx.x >> 6;
  if (lane == 0) sh[w] = v;
  __syncthreads();
  double s = 0.0;
  for (int k = 0; k < (int)(blockDim.x >> 6); k++) s += sh[k];
  __syncthreads();
  return s;
}

// centre of group g (whole chain when pbc); w[k] = m_k / M is written for
// every listed atom
__device__ void center_of(const PbcArg& P, int pbc, const double* __restrict__ pos, int n,
                          const long long* __restrict__ idx, const double* __restrict__ m, double* __restrict__ w,
                          double c[3], int* sx, int* sy, int* sz, double* sh) {
  const int t = threadIdx.x;
  double msum = 0.0;
  for (int k0 = 0; k0 < n; k0 += blockDim.x) {
    const int k = k0 + t;
    msum += block_sum(k < n ? (m ? m[k] : 1.0) : 0.0, sh);
  }
  double acc[3] = {0.0, 0.0, 0.0};
  int3 carry = make_int3(0, 0, 0);
  for (int k0 = 0; k0 < n; k0 += blockDim.x) {
    const int k = k0 + t;
    // shift of the link (k-1, k); the first atom of the group has none
    int3 s = make_int3(0, 0, 0);
    if (pbc && k < n && k > 0) s = link_shift(P, pos + 3 * idx[k - 1], pos + 3 * idx[k]);
    const int3 inc = block_scan3(s, sx, sy, sz);
    const int3 N = make_int3(carry.x + inc.x, carry.y + inc.y, carry.z + inc.z);
    double x[3] = {0.0, 0.0, 0.0}, wk = 0.0;
    if (k < n) {
      const double* xr = pos + 3 * idx[k];
      for (int q = 0; q < 3; q++)
        x[q] = xr[q] + ((double)N.x * P.box[q] + (double)N.y * P.box[3 + q] + (double)N.z * P.box[6 + q]);
      wk = (m ? m[k] : 1.0) / msum;
      w[k] = wk;
    }
    for (int q = 0; q < 3; q++) acc[q] += block_sum(wk * x[q], sh);
    // the last thread's inclusive sum carries into the next chunk
    if (t == (int)blockDim.x - 1) {
      sx[0] = N.x;
      sy[0] = N.y;
      sz[0] = N.z;
    }
    __syncthreads();
    carry = make_int3(sx[0], sy[0], sz[0]);
    __syncthreads();
  }
  c[0] = acc[0];
  c[1] = acc[1];
  c[2] = acc[2];
}

__global__ __launch_bounds__(kCdThreads) void k_center_distance(PbcArg P, int pbc, const double* __restrict__ pos,
                                                                int n1, const long long* __restrict__ idx1, int n2,
                                                                const long long* __restrict__ idx2,
                                                                const double* __restrict__ masses,
                                                                double* __restrict__ w, double* __restrict__ value,
                                                                double* __restrict__ deriv,
                                                                double* __restrict__ virial) {
  __shared__ int sx[kCdThreads], sy[kCdThreads], sz[kCdThreads];
  __shared__ double sh[kCdThreads / 64];
  __shared__ double g[6];
  double c1[3], c2[3];
  center_of(P, pbc, pos, n1, idx1, masses, w, c1, sx, sy, sz, sh);
  center_of(P, pbc, pos, n2, idx2, masses ? masses + n1 : nullptr, w + n1, c2, sx, sy, sz, sh);
  if (threadIdx.x == 0) {
    // DISTANCE makes its two atoms whole, then d = c2' - c1 (Distance.h:150-152, 205)
    if (pbc) {
      double d[3];
      pbc_distance(P, c1, c2, d);
      for (int q = 0; q < 3; q++) c2[q] = c1[q] + d[q];
    }
    const double d[3] = {c2[0] - c1[0], c2[1] - c1[1], c2[2] - c1[2]};
    const double v = sqrt(d[0] * d[0] + d[1] * d[1] + d[2] * d[2]);
    const double inv = 1.0 / v;
    for (int q = 0; q < 3; q++) {
      g[q] = -inv * d[q];
      g[3 + q] = inv * d[q];
    }
    *value = v;
    for (int a = 0; a < 3; a++)
      for (int b = 0; b < 3; b++) virial[3 * a + b] = -(c1[a] * g[b]) - c2[a] * g[3 + b];
  }
  __syncthreads();
  for (int k = threadIdx.x; k < n1 + n2; k += blockDim.x) {
    const double* gg = k < n1 ? g : g + 3;
    const double wk = w[k];
    for (int q = 0; q < 3; q++) deriv[3 * (size_t)k + q] = wk * gg[q];
  }
}

// forces[3*idx[k] + q] += f * deriv[3k + q] for the listed atoms, group 1 then
// group 2 (each group's atoms are distinct, so each phase writes distinct
// entries); virial += f * virial_cv
__global__ __launch_bounds__(kCdThreads) void k_apply_force_listed(int n1, const long long* __restrict__ idx1, int n2,
                                                                   const long long* __restrict__ idx2,
                                                                   const double* __restrict__ f,
                                                                   const double* __restrict__ deriv,
                                                                   const double* __restrict__ vcv,
                                                                   double* __restrict__ forces,
                                                                   double* __restrict__ virial) {
  const double fv = *f;
  for (int k = threadIdx.x; k < n1; k += blockDim.x)
    for (int q = 0; q < 3; q++) forces[3 * idx1[k] + q] += fv * deriv[3 * (size_t)k + q];
  __syncthreads();
  for (int k = threadIdx.x; k < n2; k += blockDim.x)
    for (int q = 0; q < 3; q++) forces[3 * idx2[k] + q] += fv * deriv[3 * (size_t)(n1 + k) + q];
  if (virial && vcv && threadIdx.x < 9) virial[threadIdx.x] += fv * vcv[threadIdx.x];
}

}  // namespace p2a

struct p2a_cdist {
  int device = 0;
  int64_t natoms = 0;
  int n1 = 0, n2 = 0;
  int pbc = 1;
  p2a::DBuf<long long> idx;  // group 1 then group 2
  p2a::DBuf<double> masses, w;
  bool has_masses = false;
};

using namespace p2a;

extern "C" {

int p2a_cdist_create(p2a_cdist** out, int device, int64_t natoms, int64_t n1, const int64_t* group1, int64_t n2,
                     const int64_t* group2, const double* masses, int use_pbc) {
  if (!out || (n1 > 0 && !group1) || (n2 > 0 && !group2)) return fail(P2A_ERR_ARG, "null argument");
  *out = nullptr;
  if (n1 <= 0 || n2 <= 0) return fail(P2A_ERR_ARG, "at least one atom should be specified");
  if (n1 + n2 > (1 << 26) || natoms > ((int64_t)1 << 40)) return fail(P2A_ERR_UNSUPPORTED, "groups too large");
  for (int g = 0; g < 2; g++) {
    const int64_t* a = g ? group2 : group1;
    const int64_t n = g ? n2 : n1;
    std::vector<int64_t> v(a, a + n);
    for (int64_t x : v)
      if (x < 0 || x >= natoms) return fail(P2A_ERR_ARG, "atom index outside the positions array");
    std::sort(v.begin(), v.end());
    if (std::adjacent_find(v.begin(), v.end()) != v.end())
      return fail(P2A_ERR_ARG, "an atom is listed twice in one group");
  }
  if (masses)
    for (int64_t k = 0; k < n1 + n2; k++)
      if (!(masses[k] > 0.0)) return fail(P2A_ERR_ARG, "masses must be positive");
  int ndev = 0;
  if (hipGetDeviceCount(&ndev) != hipSuccess || ndev == 0) return fail(P2A_ERR_HIP, "no HIP device available");
  if (device < 0 || device >= ndev) return fail(P2A_ERR_ARG, "device index out of range");
  P2A_HIP(hipSetDevice(device));
  auto* c = new p2a_cdist();
  c->device = device;
  c->natoms = natoms;
  c->n1 = (int)n1;
  c->n2 = (int)n2;
  c->pbc = use_pbc ? 1 : 0;
  std::vector<long long> all(group1, group1 + n1);
  all.insert(all.end(), group2, group2 + n2);
  int rc;
  if ((rc = c->idx.reserve(all.size())) || (rc = c->w.reserve(all.size())) ||
      (masses && (rc = c->masses.reserve(all.size())))) {
    delete c;
    return rc;
  }
  if (hipMemcpy(c->idx.p, all.data(), all.size() * sizeof(long long), hipMemcpyHostToDevice) != hipSuccess ||
      (masses && hipMemcpy(c->masses.p, masses, all.size() * sizeof(double), hipMemcpyHostToDevice) != hipSuccess)) {
    delete c;
    return fail(P2A_ERR_HIP, "hipMemcpy of the groups failed");
  }
  c->has_masses = masses != nullptr;
  *out = c;
  g_last_error.clear();
  return P2A_OK;
}

int p2a_cdist_destroy(p2a_cdist* c) {
  if (!c) return P2A_OK;
  (void)hipSetDevice(c->device);
  (void)hipDeviceSynchronize();
  delete c;
  return P2A_OK;
}

int p2a_cdist_calculate_device(p2a_cdist* c, void* stream, const double* box9, const double* pos_dev,
                               double* value_dev, double* deriv_dev, double* virial_dev) {
  if (!c || !pos_dev || !value_dev || !deriv_dev || !virial_dev) return fail(P2A_ERR_ARG, "null argument");
  if (c->pbc && !box9) return fail(P2A_ERR_ARG, "box is required with periodic boundaries");
  P2A_HIP(hipSetDevice(c->device));
  PbcArg P;
  std::memset(&P, 0, sizeof(P));
  if (c->pbc) {
    PbcHost h;
    pbc_setup(box9, h);
    if (h.type == 0) return fail(P2A_ERR_ARG, "the box is singular");
    fill_pbc_arg(h, P);
  }
  hipLaunchKernelGGL(k_center_distance, dim3(1), dim3(kCdThreads), 0, (hipStream_t)stream, P, c->pbc, pos_dev,
                     c->n1, (const long long*)c->idx.p, c->n2, (const long long*)c->idx.p + c->n1,
                     (const double*)(c->has_masses ? c->masses.p : nullptr), c->w.p, value_dev, deriv_dev,
                     virial_dev);
  return launch_check("centre distance");
}

int p2a_cdist_apply_force_device(p2a_cdist* c, void* stream, const double* f_dev, const double* deriv_dev,
                                 const double* virial_cv_dev, double* forces_dev, double* virial_dev) {
  if (!c || !f_dev || !deriv_dev || !forces_dev) return fail(P2A_ERR_ARG, "null argument");
  P2A_HIP(hipSetDevice(c->device));
  hipLaunchKernelGGL(k_apply_force_listed, dim3(1), dim3(kCdThreads), 0, (hipStream_t)stream, c->n1,
                     (const long long*)c->idx.p, c->n2, (const long long*)c->idx.p + c->n1, f_dev, deriv_dev,
                     virial_cv_dev, forces_dev, virial_dev);
  return launch_check("centre distance forces");
}

}  // extern "C"
