// Batched COORDINATION on gfx950: many independent single-group systems
// (frames or walkers with their own atom counts and periodic cells) in one
// set of arrays, so that every kernel of a step covers all systems.
//
// The reference evaluates such a batch as one COORDINATION action per system
// (CoordinationBase::prepare + calculate, CoordinationBase.cpp:137-232, with
// NLISTCELLS and NL_STRIDE=1) and sums nothing across systems; here atoms,
// columns and clusters of system s are contiguous, SysArg[s] carries its cell
// and column grid, and the per-system value / virial / pair count come out
// of one reduction kernel with one workgroup per system.
//
// A step never waits on the host between kernels: cluster counts and list
// sizes stay on the device, the list arrays have host-chosen capacities and
// the kernels drop what does not fit.  The host checks the device totals at
// the end of the step; when a capacity was exceeded it grows the arrays and
// repeats the step.  That makes the step capturable as one HIP graph.
#include <hip/hip_runtime.h>

#include <algorithm>
#include <cmath>
#include <cstring>
#include <string>
#include <vector>

#include "coord_kernels.h"

struct p2a_batch {
  int device = 0;
  int cu_count = 256;
  hipStream_t own_stream = nullptr;
  hipStream_t stream = nullptr;
  int nsys = 0;
  int64_t natoms = 0;
  std::vector<int64_t> nat, atom_base;
  p2a_switch sw{};
  double rlist = 0.0;
  bool boxes_set = false;
  // host descriptors
  std::vector<p2a::SysArg> sys_h;
  std::vector<p2a::PbcArg> pbc_h;
  std::vector<int> colbase_h;
  int ncol_tot = 0;
  int cl_cap = 0;
  int ent_cap = 0, items_cap = 0;
  int64_t generation = 0;  // bumped whenever a device array moves
  p2a::PairTuning tune;
  // device arrays
  p2a::DBuf<p2a::SysArg> sys;
  p2a::DBuf<p2a::PbcArg> pbc;
  p2a::DBuf<int> atom_sys, sys_colbase;
  p2a::DBuf<int4> info;
  p2a::DBuf<double> key;
  p2a::DBuf<int> col_count, col_start, col_cl, fill, tmp, slot_loc, cl_col, cl_sys;
  p2a::DBuf<double4> origin;
  p2a::DBuf<float4> bb0, bb, p;
  p2a::DBuf<float2> kr;
  p2a::DBuf<int> ecount, estart, istart, totals, scan_bsum, bflag, wave_start;
  p2a::DBuf<unsigned> ell;  // candidates the list count pass keeps for the fill pass
  p2a::DBuf<p2a::Entry> ents;
  p2a::DBuf<int4> items;
  p2a::DBuf<double> fpart, partial, npairs, out_values, out_virials, out_deriv;
  int* totals_host = nullptr;  // pinned
  // graph
  bool use_graph = false;
  hipGraphExec_t gexec = nullptr;
  const void* gkey[4] = {nullptr, nullptr, nullptr, nullptr};
  int64_t ggen = -1;
  int gchunk = 0, gbpc = -1;
  bool last_graph = false;
  int regrowths = 0;
  hipEvent_t ev0 = nullptr, ev1 = nullptr;
  double last_ms = 0.0;
};

namespace p2a {

static int batch_reserve_lists(p2a_batch* b, int ent_cap, int items_cap) {
  int rc;
  if ((rc = b->ents.reserve(std::max(ent_cap, 1))) || (rc = b->items.reserve(std::max(items_cap, 1))) ||
      (rc = b->fpart.reserve((size_t)CL * 3 * std::max(items_cap, 1))) ||
      (rc = b->partial.reserve(8 * (size_t)std::max(items_cap, 1))) ||
      (rc = b->bflag.reserve(std::max(items_cap, 1))) ||
      // the pair kernel's schedule: at most one wave per item (grid1(items_cap, 4) workgroups)
      (rc = b->wave_start.reserve((size_t)grid1(std::max(items_cap, 1), PAIR_BLOCK / WAVE) * (PAIR_BLOCK / WAVE) + 1)))
    return rc;
  b->ent_cap = ent_cap;
  b->items_cap = items_cap;
  b->generation++;
  return P2A_OK;
}

// the launches of one step, in stream order; nothing here waits on the host
static int batch_step(p2a_batch* b, const double* pos, double* values, double* deriv, double* virials) {
  hipStream_t s = b->stream;
  const int N = (int)b->natoms;
  const int ncol = b->ncol_tot;
  const int clc = b->cl_cap;
  const size_t nslot = (size_t)clc * CL;
  const int chunk = b->tune.chunk;
  const int* ncl_dev = b->totals.p + 1;
  hipLaunchKernelGGL(k_fill_int3, dim3(std::max(1, std::min(grid1(nslot, 256), 4096))), dim3(256), 0, s,
                     b->col_count.p, ncol + 1, 0, b->fill.p, ncol + 1, 0, b->slot_loc.p, (int)nslot, -1);
  hipLaunchKernelGGL(k_bin, dim3(grid1(N, 256)), dim3(256), 0, s, N, pos, 0, (const SysArg*)b->sys.p,
                     (const int*)b->atom_sys.p, b->info.p, b->key.p, b->col_count.p);
  // atom and cluster starts of the columns; totals[0], [1]
  launch_list_scans(s, ncol, nullptr, b->col_count.p, CL, b->scan_bsum.p, b->col_start.p, b->col_cl.p, b->totals.p, 0,
                    1);
  hipLaunchKernelGGL(k_scatter, dim3(grid1(N, 256)), dim3(256), 0, s, N, b->info.p, b->col_start.p, b->fill.p,
                     b->tmp.p);
  hipLaunchKernelGGL(k_rank, dim3(ncol), dim3(256), 0, s, ncol, b->key.p, b->col_start.p, b->col_cl.p, b->tmp.p,
                     b->slot_loc.p, b->info.p, pos, 0, (const SysArg*)b->sys.p, (const int*)b->atom_sys.p,
                     b->origin.p, b->bb0.p, b->kr.p, b->cl_sys.p, b->cl_col.p, b->p.p, b->bb.p, pack_scale(b->sw));
  ListArgs L;
  L.nclI = clc;
  L.nclI_dev = ncl_dev;
  L.mode = 0;
  L.ent_cap = b->ent_cap;
  L.rlist2 = (float)(b->rlist * b->rlist);
  L.pscale = pack_scale(b->sw);
  const int wpb = 256 / WAVE;
  hipLaunchKernelGGL(k_list<0>, dim3(grid1(clc, wpb)), dim3(256), 0, s, L, b->cl_col.p, (const int*)b->cl_sys.p,
                     (const SysArg*)b->sys.p, (const SysArg*)b->sys.p, b->bb0.p, b->bb0.p, b->kr.p, b->kr.p,
                     b->col_cl.p, b->origin.p, b->origin.p, b->ecount.p, (const int*)nullptr, (Entry*)nullptr, b->ell.p);
  launch_list_scans(s, clc, ncl_dev, b->ecount.p, chunk, b->scan_bsum.p, b->estart.p, b->istart.p, b->totals.p, 2, 3);
  hipLaunchKernelGGL(k_list_stream, dim3(grid1(clc, wpb)), dim3(256), 0, s, L, (const int*)b->cl_sys.p,
                     (const SysArg*)b->sys.p, b->origin.p, b->origin.p, b->ecount.p, (const int*)b->estart.p, b->ents.p,
                     b->ell.p);
  // k_list<2>: 64 clusters per wave
  hipLaunchKernelGGL(k_list<2>, dim3(grid1(clc, wpb * WAVE)), dim3(256), 0, s, L, b->cl_col.p, (const int*)b->cl_sys.p,
                     (const SysArg*)b->sys.p, (const SysArg*)b->sys.p, b->bb0.p, b->bb0.p, b->kr.p, b->kr.p,
                     b->col_cl.p, b->origin.p, b->origin.p, b->ecount.p, (const int*)b->estart.p, b->ents.p, b->ell.p);
  hipLaunchKernelGGL(k_items, dim3(grid1(clc, 256)), dim3(256), 0, s, clc, ncl_dev, b->ecount.p, b->estart.p,
                     b->istart.p, b->items.p, chunk, b->items_cap);
  // k_rank has packed the positions of this step (the list is built at every step)
  int rc;
  if ((rc = launch_check("batch: list build"))) return rc;
  PairArgs A;
  std::memset(&A, 0, sizeof(A));
  pair_args_switch(b->sw, A);
  A.nitems = 0;
  A.nitems_dev = b->totals.p + 3;
  A.items_cap = b->items_cap;
  A.ent_cap = b->ent_cap;
  A.twogroup = 0;
  A.pbc = b->pbc.p;
  A.cl_sys = b->cl_sys.p;
  A.pos64 = pos;
  A.item_partials = 1;  // per-system sums need the items' own partials
  A.ipart = b->partial.p;
  A.bflag = b->bflag.p;
  int grid = grid1(std::max(b->items_cap, 1), PAIR_BLOCK / WAVE);
  A.wave_start = nullptr;
  if (b->tune.blocks_per_cu > 0) grid = std::min(grid, b->tune.blocks_per_cu * b->cu_count);
  const int nw = grid * (PAIR_BLOCK / WAVE);
  // contiguous item ranges of balanced cost per wave of the persistent grid
  // (the first step runs before the list arrays exist: no items, no schedule)
  if (b->tune.blocks_per_cu > 0 && b->items_cap > 0 && b->wave_start.p && b->wave_start.n >= (size_t)nw + 1) {
    hipLaunchKernelGGL(k_wave_sched, dim3(grid1(std::max(b->items_cap, 1), 256)), dim3(256), 0, s, 0,
                       (const int*)A.nitems_dev, b->items_cap, (const int4*)b->items.p, nw, b->wave_start.p);
    A.wave_start = b->wave_start.p;
  }
  launch_pairs(A, grid, s, false, chunk, b->items.p, b->ents.p, b->p.p, b->p.p, b->bb.p, b->bb.p, (const int*)nullptr,
               (const int*)nullptr, b->fpart.p, b->partial.p);
  if ((rc = launch_check("batch: pair kernel"))) return rc;
  hipLaunchKernelGGL(k_gather_reduce, dim3(grid1(nslot * 3, 256)), dim3(256), 0, s, grid1(nslot * 3, 256), clc,
                     ncl_dev, b->p.p, b->istart.p, b->items_cap, b->fpart.p, deriv, 0, RedArgs{});
  hipLaunchKernelGGL(k_sys_reduce, dim3(b->nsys), dim3(256), 0, s, b->nsys, b->sys_colbase.p, b->col_cl.p,
                     b->istart.p, b->items_cap, b->partial.p, values, virials, b->npairs.p);
  return launch_check("batch: reductions");
}

static void drop_graph(p2a_batch* b) {
  if (b->gexec) (void)hipGraphExecDestroy(b->gexec);
  b->gexec = nullptr;
  b->ggen = -1;
}

// capture the step into a graph unless the current one matches
static int ensure_graph(p2a_batch* b, const double* pos, double* values, double* deriv, double* virials) {
  const void* key[4] = {pos, values, deriv, virials};
  if (b->gexec && b->ggen == b->generation && b->gchunk == b->tune.chunk && b->gbpc == b->tune.blocks_per_cu &&
      std::memcmp(key, b->gkey, sizeof(key)) == 0)
    return P2A_OK;
  drop_graph(b);
  hipGraph_t g = nullptr;
  P2A_HIP(hipStreamBeginCapture(b->stream, hipStreamCaptureModeThreadLocal));
  int rc = batch_step(b, pos, values, deriv, virials);
  hipError_t e = hipStreamEndCapture(b->stream, &g);
  if (rc) {
    if (g) (void)hipGraphDestroy(g);
    return rc;
  }
  if (e != hipSuccess) return fail(P2A_ERR_HIP, std::string("hipStreamEndCapture: ") + hipGetErrorString(e));
  e = hipGraphInstantiate(&b->gexec, g, nullptr, nullptr, 0);
  (void)hipGraphDestroy(g);
  if (e != hipSuccess) {
    b->gexec = nullptr;
    return fail(P2A_ERR_HIP, std::string("hipGraphInstantiate: ") + hipGetErrorString(e));
  }
  std::memcpy(b->gkey, key, sizeof(key));
  b->ggen = b->generation;
  b->gchunk = b->tune.chunk;
  b->gbpc = b->tune.blocks_per_cu;
  return P2A_OK;
}

static int batch_run(p2a_batch* b, const double* pos, double* values, double* deriv, double* virials,
                     bool allow_graph) {
  int rc;
  for (int attempt = 0; attempt < 4; attempt++) {
    const bool graph = allow_graph && b->use_graph && b->stream != nullptr;
    P2A_HIP(hipEventRecord(b->ev0, b->stream));
    if (graph) {
      if ((rc = ensure_graph(b, pos, values, deriv, virials))) return rc;
      P2A_HIP(hipGraphLaunch(b->gexec, b->stream));
    } else {
      if ((rc = batch_step(b, pos, values, deriv, virials))) return rc;
    }
    P2A_HIP(hipEventRecord(b->ev1, b->stream));
    P2A_HIP(hipMemcpyAsync(b->totals_host, b->totals.p, 4 * sizeof(int), hipMemcpyDeviceToHost, b->stream));
    P2A_HIP(hipStreamSynchronize(b->stream));
    b->last_graph = graph;
    const int ent = b->totals_host[2], its = b->totals_host[3];
    if (ent <= b->ent_cap && its <= b->items_cap) {
      float ms = 0.f;
      if (hipEventElapsedTime(&ms, b->ev0, b->ev1) == hipSuccess) b->last_ms = ms;
      return P2A_OK;
    }
    // a list array was too small: grow with headroom and repeat the step
    const int64_t ne = (int64_t)ent + ent / 8 + 4096, ni = (int64_t)its + its / 8 + 1024;
    if (ne > INT32_MAX || ni > INT32_MAX) return fail(P2A_ERR_UNSUPPORTED, "batch neighbour list exceeds 2^31 entries");
    if ((rc = batch_reserve_lists(b, (int)std::max<int64_t>(ne, b->ent_cap), (int)std::max<int64_t>(ni, b->items_cap))))
      return rc;
    b->regrowths++;
  }
  return fail(P2A_ERR_HIP, "batch: list capacity did not settle");
}

}  // namespace p2a

using namespace p2a;

extern "C" {

int p2a_batch_create(p2a_batch** out, int device, int nsys, const int64_t* natoms, const p2a_switch* sw,
                     double rlist) {
  if (!out || !natoms || !sw) return fail(P2A_ERR_ARG, "null argument");
  *out = nullptr;
  if (nsys <= 0) return fail(P2A_ERR_ARG, "a batch needs at least one system");
  int64_t tot = 0;
  for (int k = 0; k < nsys; k++) {
    if (natoms[k] <= 0) return fail(P2A_ERR_ARG, "system " + std::to_string(k) + " has no atoms");
    tot += natoms[k];
  }
  if (tot > ((int64_t)1 << 29)) return fail(P2A_ERR_UNSUPPORTED, "more than 2^29 atoms in one batch");
  if (!(sw->dmax < 1e30)) return fail(P2A_ERR_UNSUPPORTED, "a batch needs a finite D_MAX");
  if (rlist < sw->dmax)
    return fail(P2A_ERR_UNSUPPORTED, "list radius smaller than D_MAX: pairs within D_MAX would be dropped");
  int ndev = 0;
  if (hipGetDeviceCount(&ndev) != hipSuccess || ndev == 0) return fail(P2A_ERR_HIP, "no HIP device available");
  if (device < 0 || device >= ndev) return fail(P2A_ERR_ARG, "device index out of range");
  P2A_HIP(hipSetDevice(device));
  auto* b = new p2a_batch();
  b->device = device;
  b->nsys = nsys;
  b->natoms = tot;
  b->nat.assign(natoms, natoms + nsys);
  b->atom_base.resize(nsys + 1);
  b->atom_base[0] = 0;
  for (int k = 0; k < nsys; k++) b->atom_base[k + 1] = b->atom_base[k] + natoms[k];
  b->sw = *sw;
  b->rlist = rlist + 1e-4;
  int ncu = 0;
  if (hipDeviceGetAttribute(&ncu, hipDeviceAttributeMultiprocessorCount, device) == hipSuccess && ncu > 0)
    b->cu_count = ncu;
  int rc = P2A_OK;
  if (hipStreamCreateWithFlags(&b->own_stream, hipStreamNonBlocking) != hipSuccess ||
      hipEventCreate(&b->ev0) != hipSuccess || hipEventCreate(&b->ev1) != hipSuccess ||
      hipHostMalloc((void**)&b->totals_host, 16 * sizeof(int), hipHostMallocDefault) != hipSuccess) {
    p2a_batch_destroy(b);
    return fail(P2A_ERR_HIP, "batch: stream/event/pinned allocation failed");
  }
  b->stream = b->own_stream;
  std::vector<int> as((size_t)tot);
  for (int k = 0; k < nsys; k++)
    for (int64_t i = b->atom_base[k]; i < b->atom_base[k + 1]; i++) as[i] = k;
  if ((rc = b->atom_sys.reserve(tot)) || (rc = b->totals.reserve(16)) || (rc = b->npairs.reserve(nsys)) ||
      (rc = b->info.reserve(tot)) || (rc = b->key.reserve(tot)) || (rc = b->tmp.reserve(tot))) {
    p2a_batch_destroy(b);
    return rc;
  }
  if (hipMemcpy(b->atom_sys.p, as.data(), as.size() * sizeof(int), hipMemcpyHostToDevice) != hipSuccess) {
    p2a_batch_destroy(b);
    return fail(P2A_ERR_HIP, "hipMemcpy of the atom systems failed");
  }
  *out = b;
  g_last_error.clear();
  return P2A_OK;
}

int p2a_batch_destroy(p2a_batch* b) {
  if (!b) return P2A_OK;
  (void)hipSetDevice(b->device);
  if (b->stream) (void)hipStreamSynchronize(b->stream);
  drop_graph(b);
  if (b->ev0) (void)hipEventDestroy(b->ev0);
  if (b->ev1) (void)hipEventDestroy(b->ev1);
  if (b->totals_host) (void)hipHostFree(b->totals_host);
  if (b->own_stream) (void)hipStreamDestroy(b->own_stream);
  delete b;
  return P2A_OK;
}

int p2a_batch_set_stream(p2a_batch* b, void* stream) {
  if (!b) return fail(P2A_ERR_ARG, "null batch");
  P2A_HIP(hipSetDevice(b->device));
  if (b->stream) P2A_HIP(hipStreamSynchronize(b->stream));
  b->stream = (hipStream_t)stream;
  drop_graph(b);
  return P2A_OK;
}

int p2a_batch_set_boxes(p2a_batch* b, const double* boxes) {
  if (!b || !boxes) return fail(P2A_ERR_ARG, "null argument");
  P2A_HIP(hipSetDevice(b->device));
  std::vector<SysArg> S(b->nsys);
  std::vector<PbcArg> Pb(b->nsys);
  std::vector<int> colbase(b->nsys + 1);
  int64_t ncol = 0, clc = 0;
  for (int k = 0; k < b->nsys; k++) {
    const double* bx = boxes + 9 * (size_t)k;
    PbcHost ph;
    pbc_setup(bx, ph);
    if (ph.type == 0) return fail(P2A_ERR_ARG, "system " + std::to_string(k) + ": degenerate cell");
    double w[3];
    pbc_widths(bx, w);
    const double wmin = std::min(w[0], std::min(w[1], w[2]));
    if (wmin <= 2.0 * b->rlist * 1.0001 + 1e-6)
      return fail(P2A_ERR_UNSUPPORTED, "system " + std::to_string(k) + ": cell width " + std::to_string(wmin) +
                                           " nm is not larger than twice the list radius");
    sys_cell(bx, b->rlist, S[k]);
    sys_columns(bx, (int)b->nat[k], S[k]);
    S[k].colbase = (int)ncol;
    colbase[k] = (int)ncol;
    ncol += (int64_t)S[k].ncx * S[k].ncy;
    clc += b->nat[k] / CL + (int64_t)S[k].ncx * S[k].ncy + 1;
    fill_pbc_arg(ph, Pb[k]);
  }
  colbase[b->nsys] = (int)ncol;
  if (ncol > INT32_MAX / 2 || clc * CL > INT32_MAX / 2 || clc > kMaxClusters)
    return fail(P2A_ERR_UNSUPPORTED, "batch too large");
  if (b->stream) P2A_HIP(hipStreamSynchronize(b->stream));
  int rc;
  if ((rc = b->sys.reserve(b->nsys)) || (rc = b->pbc.reserve(b->nsys)) || (rc = b->sys_colbase.reserve(b->nsys + 1)) ||
      (rc = b->col_count.reserve(ncol + 1)) || (rc = b->col_start.reserve(ncol + 1)) ||
      (rc = b->col_cl.reserve(ncol + 1)) || (rc = b->fill.reserve(ncol + 1)) ||
      (rc = b->slot_loc.reserve((size_t)clc * CL)) || (rc = b->cl_col.reserve(clc)) || (rc = b->cl_sys.reserve(clc)) ||
      (rc = b->origin.reserve(clc)) || (rc = b->bb0.reserve(2 * (size_t)clc)) || (rc = b->bb.reserve(2 * (size_t)clc)) ||
      (rc = b->p.reserve((size_t)clc * CL)) || (rc = b->kr.reserve(clc)) || (rc = b->ecount.reserve(clc + 1)) ||
      (rc = b->estart.reserve(clc + 1)) || (rc = b->istart.reserve(clc + 1)) ||
      (rc = b->scan_bsum.reserve(2 * (size_t)grid1(std::max<int64_t>(clc, (int64_t)ncol), kScanBlock) + 2)) ||
      (rc = b->ell.reserve((size_t)clc * kEll)))
    return rc;
  P2A_HIP(hipMemcpy(b->sys.p, S.data(), S.size() * sizeof(SysArg), hipMemcpyHostToDevice));
  P2A_HIP(hipMemcpy(b->pbc.p, Pb.data(), Pb.size() * sizeof(PbcArg), hipMemcpyHostToDevice));
  P2A_HIP(hipMemcpy(b->sys_colbase.p, colbase.data(), colbase.size() * sizeof(int), hipMemcpyHostToDevice));
  b->sys_h = S;
  b->pbc_h = Pb;
  b->colbase_h = colbase;
  b->ncol_tot = (int)ncol;
  b->cl_cap = (int)clc;
  b->boxes_set = true;
  b->generation++;  // the graph holds the old grid sizes
  return P2A_OK;
}

int p2a_batch_calculate_device(p2a_batch* b, const double* pos_dev, double* values_dev, double* deriv_dev,
                               double* virials_dev) {
  if (!b || !pos_dev || !values_dev) return fail(P2A_ERR_ARG, "null argument");
  if (!b->boxes_set) return fail(P2A_ERR_ARG, "p2a_batch_set_boxes must be called first");
  if ((deriv_dev == nullptr) != (virials_dev == nullptr))
    return fail(P2A_ERR_ARG, "derivatives and virials are requested together");
  P2A_HIP(hipSetDevice(b->device));
  int rc;
  if (!deriv_dev) {
    if ((rc = b->out_deriv.reserve(3 * (size_t)b->natoms)) || (rc = b->out_virials.reserve(9 * (size_t)b->nsys)))
      return rc;
    deriv_dev = b->out_deriv.p;
    virials_dev = b->out_virials.p;
  }
  return batch_run(b, pos_dev, values_dev, deriv_dev, virials_dev, true);
}

int p2a_batch_get_pair_counts(p2a_batch* b, int64_t* out) {
  if (!b || !out) return fail(P2A_ERR_ARG, "null argument");
  P2A_HIP(hipSetDevice(b->device));
  if (b->stream) P2A_HIP(hipStreamSynchronize(b->stream));
  std::vector<double> d(b->nsys);
  P2A_HIP(hipMemcpy(d.data(), b->npairs.p, d.size() * sizeof(double), hipMemcpyDeviceToHost));
  for (int k = 0; k < b->nsys; k++) out[k] = (int64_t)d[k];
  return P2A_OK;
}

int p2a_batch_set_graph(p2a_batch* b, int on) {
  if (!b) return fail(P2A_ERR_ARG, "null batch");
  b->use_graph = on != 0;
  if (!b->use_graph) drop_graph(b);
  return P2A_OK;
}

int p2a_batch_set_tuning(p2a_batch* b, int chunk, int blocks_per_cu) {
  if (!b) return fail(P2A_ERR_ARG, "null batch");
  if (chunk != 32 && chunk != 64) return fail(P2A_ERR_ARG, "chunk must be 32 or 64");
  if (blocks_per_cu < 0 || blocks_per_cu > 64) return fail(P2A_ERR_ARG, "blocks_per_cu must be in [0, 64]");
  b->tune.chunk = chunk;
  b->tune.blocks_per_cu = blocks_per_cu;
  return P2A_OK;
}

int p2a_batch_autotune(p2a_batch* b, const double* pos_dev, int reps, int* chunk_out, int* bpc_out) {
  if (!b || !pos_dev) return fail(P2A_ERR_ARG, "null argument");
  if (!b->boxes_set) return fail(P2A_ERR_ARG, "p2a_batch_set_boxes must be called first");
  P2A_HIP(hipSetDevice(b->device));
  reps = std::max(1, std::min(reps, 50));
  int rc;
  if ((rc = b->out_values.reserve(b->nsys)) || (rc = b->out_deriv.reserve(3 * (size_t)b->natoms)) ||
      (rc = b->out_virials.reserve(9 * (size_t)b->nsys)))
    return rc;
  // candidates: entries per wave x persistent grid (0 = one wave per item)
  const int chunks[2] = {64, 32};
  const int grids[4] = {0, 2, 4, 8};
  const PairTuning saved = b->tune;
  PairTuning best = saved;
  double best_ms = 1e300;
  for (int ch : chunks) {
    for (int bpc : grids) {
      b->tune.chunk = ch;
      b->tune.blocks_per_cu = bpc;
      // one untimed step settles the list capacities for this chunk
      if ((rc = batch_run(b, pos_dev, b->out_values.p, b->out_deriv.p, b->out_virials.p, false))) {
        b->tune = saved;
        return rc;
      }
      double tot = 0.0;
      for (int r = 0; r < reps; r++) {
        if ((rc = batch_run(b, pos_dev, b->out_values.p, b->out_deriv.p, b->out_virials.p, false))) {
          b->tune = saved;
          return rc;
        }
        tot += b->last_ms;
      }
      if (tot / reps < best_ms) {
        best_ms = tot / reps;
        best = b->tune;
      }
    }
  }
  b->tune = best;
  drop_graph(b);
  if (chunk_out) *chunk_out = best.chunk;
  if (bpc_out) *bpc_out = best.blocks_per_cu;
  return P2A_OK;
}

int p2a_batch_get_stats(p2a_batch* b, p2a_batch_stats* st) {
  if (!b || !st) return fail(P2A_ERR_ARG, "null argument");
  P2A_HIP(hipSetDevice(b->device));
  if (b->stream) P2A_HIP(hipStreamSynchronize(b->stream));
  std::memset(st, 0, sizeof(*st));
  st->nsys = b->nsys;
  st->natoms = b->natoms;
  if (b->totals.p) {
    int t[4];
    P2A_HIP(hipMemcpy(t, b->totals.p, sizeof(t), hipMemcpyDeviceToHost));
    st->clusters = t[1];
    st->list_entries = t[2];
    st->items = t[3];
  }
  if (b->npairs.p) {
    std::vector<double> d(b->nsys);
    P2A_HIP(hipMemcpy(d.data(), b->npairs.p, d.size() * sizeof(double), hipMemcpyDeviceToHost));
    double s = 0;
    for (double v : d) s += v;
    st->contributing_pairs = (int64_t)s;
  }
  st->chunk = b->tune.chunk;
  st->blocks_per_cu = b->tune.blocks_per_cu;
  st->graph = b->last_graph ? 1 : 0;
  st->regrowths = b->regrowths;
  st->last_step_ms = b->last_ms;
  return P2A_OK;
}

}  // extern "C"
