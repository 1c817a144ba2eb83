// COORDINATION on gfx950: the single-system context and its C ABI.
//
// One context = one COORDINATION action (CoordinationBase.cpp:42-131): its
// groups, switch and neighbour-list settings, the device arrays of the
// cluster-pair structure, and a HIP stream.  The kernels are in
// coord_kernels.h; this file sizes the arrays, decides the rebuild steps
// (NeighborList::prepare, NeighborList.cpp:433-456) and launches.
#include <hip/hip_runtime.h>

#include <algorithm>
#include <cmath>
#include <cstring>
#include <mutex>
#include <string>
#include <vector>

#include "coord_kernels.h"

namespace p2a {

struct Dir {  // one cluster-pair list: i clusters of group I against group J
  int I = 0, J = 0;
  DBuf<int> ecount, estart, istart;
  DBuf<Entry> ents;
  DBuf<int4> items;
  DBuf<double> fpart;  // per item: derivative partials of the 8 i atoms
  DBuf<double> rpart;  // per item: per-atom coordination-number rows of the 8 i atoms
  DBuf<int> bflag;     // per item: band pairs found by the pair kernel
  DBuf<int> wave_start;  // the pair kernel's balanced schedule (k_wave_sched)
  int sched_waves = -1;  // the grid it was made for
  int64_t sched_gen = -1;  // and the list build
  DBuf<unsigned> ell;  // candidates the list count pass keeps for the fill pass
  int nitems = 0, item_offset = 0;
  int64_t nentries = 0;
};

struct Group {
  int off = 0, n = 0;
  SysArg S{};   // cell and column grid of this group
  SysArg S_dev{};  // the copy the device holds
  bool S_valid = false;
  int maxcl = 0;   // cluster capacity: n/8 + one partial cluster per column
  int ncol = 1;
  int ncl = 0;  // clusters (device total copied at rebuild)
  // slot_loc of the last build (order_nslot slots at order_ptr) orders the
  // atoms of the next one; 0 = no such order
  int order_nslot = 0;
  const int* order_ptr = nullptr;
  DBuf<SysArg> sys;
  DBuf<int4> info;
  DBuf<double> key;
  DBuf<int> col_count, col_start, col_cl, fill, tmp, slot_loc, cl_col, abs_slot;
  DBuf<double4> origin;
  DBuf<float4> bb0, bb, p;
  DBuf<float2> kr;
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
  double box_at_build[9] = {0};    // cell the list entries are imaged for
  double box_at_rebuild[9] = {0};  // cell of the last rebuild (wrapping integers, images)
  int max_wrap = 0;                // largest |lattice translation| of a wrapped atom at that rebuild
  bool totals_clear = false;
  bool built = false;
  p2a::Group g[2];
  p2a::DBuf<int> totals;       // [4..7] atoms/clusters per group, [8..11] entries/items per direction
  p2a::Dir dir[2];
  int ndir = 0;
  int nitems = 0;
  int64_t nentries = 0;
  p2a::PairTuning tune;
  int cu_count = 256;
  p2a::DBuf<double> partial, red, stats, pos_buf, deriv_buf, out_buf;
  p2a::DBuf<int> scan_bsum;   // block sums of the list scans
  p2a::DBuf<long long> abs_dev;
  int64_t rebuilds = 0;
  bool profiling = false;
  int prof_every = 1;       // events around every prof_every-th pair pass
  int64_t prof_calls = 0;
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
  double path_box[9] = {0};  // box of the last path decision
  int path_cached = -1;
  // without a neighbour list: list built at D_MAX + skin and kept while no
  // atom has moved more than skin/2 (skin_active: 0 when the cell is too
  // narrow for the larger radius)
  double skin = 0.0, skin_active = 0.0;
  // positions packed ahead by the skin test, for the list of build packed_gen
  const double* packed_pos = nullptr;
  int64_t packed_gen = -1;
  p2a::DBuf<double> pos_ref;  // positions at the last build
  p2a::DBuf<int> moved;
  hipEvent_t moved_ev = nullptr;  // the moved flag has reached host_ro[32]
  int* host_ro = nullptr;  // pinned host words for the few device-to-host reads
  // neighbour list of the FP64 path: partners of every atom (symmetric CSR),
  // or one flag per pair with PAIR; cell of every atom for NLISTCELLS
  p2a::DBuf<int> nl_count, nl_start, nl_nbr, nl_cell;
  p2a::DBuf<long long> nl_total;
  bool nl_valid = false;
  // per-atom coordination numbers: weights per requested atom; row and
  // environment indices for the FP64 path
  p2a::DBuf<float> wreq;
  p2a::DBuf<int> cn_rows, cn_env;
};

namespace p2a {

static int choose_path_uncached(p2a_coord* c, const double* box9);

// decides the evaluation path for this box; 0 = cluster kernel
static int choose_path(p2a_coord* c, const double* box9) {
  if (c->path_cached >= 0 && std::memcmp(c->path_box, box9, sizeof(c->path_box)) == 0) return c->path_cached;
  c->path_cached = choose_path_uncached(c, box9);
  std::memcpy(c->path_box, box9, sizeof(c->path_box));
  return c->path_cached;
}

static int choose_path_uncached(p2a_coord* c, const double* box9) {
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
  c->skin_active = 0.0;
  if (c->spec.nl_mode == P2A_NL_NONE && c->skin > 0.0 && wmin > 2.0 * (sw.dmax + c->skin) * 1.0001 + 1e-6)
    c->skin_active = c->skin;
  return 0;
}

static int rebuild(p2a_coord* c, const double* box9, const double* pos_dev) {
  const p2a_switch& sw = c->spec.sw;
  // list radius: NL_CUTOFF when a neighbour list is requested (its candidates
  // must cover every pair that can come within D_MAX before the next rebuild),
  // D_MAX + skin otherwise (skin 0: the list is rebuilt every step)
  c->rlist = (c->spec.nl_mode != P2A_NL_NONE ? std::max(c->spec.nl_cutoff, sw.dmax) : sw.dmax + c->skin_active) + 1e-4;
  const int ngroups = c->spec.nb > 0 ? 2 : 1;
  hipStream_t s = c->stream;
  int rc;
  for (int gi = 0; gi < ngroups; gi++) {
    Group& G = c->g[gi];
    G.off = gi == 0 ? 0 : (int)c->spec.na;
    G.n = gi == 0 ? (int)c->spec.na : (int)c->spec.nb;
    sys_cell(box9, c->rlist, G.S);
    sys_columns(box9, G.n, G.S);
    G.S.colbase = 0;
    G.ncol = G.S.ncx * G.S.ncy;
    int64_t maxcl64 = (int64_t)G.n / CL + G.ncol + 1;
    if (maxcl64 > kMaxClusters) return fail(P2A_ERR_UNSUPPORTED, "too many atom clusters for the list encoding");
    const int maxcl = (int)maxcl64;
    G.maxcl = maxcl;
    if ((rc = G.sys.reserve(1)) || (rc = G.info.reserve(G.n)) || (rc = G.key.reserve(G.n)) ||
        (rc = G.col_count.reserve(G.ncol + 1)) || (rc = G.col_start.reserve(G.ncol + 1)) ||
        (rc = G.col_cl.reserve(G.ncol + 1)) || (rc = G.fill.reserve(G.ncol + 1)) || (rc = G.tmp.reserve(G.n)) ||
        (rc = G.slot_loc.reserve((size_t)maxcl * CL)) || (rc = G.cl_col.reserve(maxcl)) ||
        (rc = G.origin.reserve(maxcl)) || (rc = G.bb0.reserve(2 * (size_t)maxcl)) ||
        (rc = G.bb.reserve(2 * (size_t)maxcl)) || (rc = G.p.reserve((size_t)maxcl * CL)) ||
        (rc = G.kr.reserve(maxcl)) || (rc = G.abs_slot.reserve((size_t)maxcl * CL)))
      return rc;
    // the descriptor goes to the device in stream order when it changed
    if (!G.S_valid || std::memcmp(&G.S, &G.S_dev, sizeof(SysArg)) != 0) {
      hipLaunchKernelGGL(k_store<SysArg>, dim3(1), dim3(64), 0, s, G.S, G.sys.p);
      G.S_dev = G.S;
      G.S_valid = true;
    }
  }
  if ((rc = c->totals.reserve(16))) return rc;
  if (!c->totals_clear) {
    hipLaunchKernelGGL(k_fill_int, dim3(1), dim3(64), 0, s, c->totals.p, 16, 0);
    c->totals_clear = true;
  }
  // two alternating slots for the largest wrap of this rebuild
  const int wslot = 12 + (int)(c->rebuilds & 1);
  for (int gi = 0; gi < ngroups; gi++) {
    Group& G = c->g[gi];
    // the atoms are binned in the slot order of the last build when it is
    // still in slot_loc (atoms move little between builds, so a workgroup's
    // column counts stay in LDS); k_rank then writes every slot of the new
    // clusters, padding included
    const bool use_order = G.order_nslot > 0 && G.order_ptr == G.slot_loc.p;
    const int norder = use_order ? G.order_nslot : 0;
    const int* order = use_order ? G.slot_loc.p : nullptr;
    const int nthr = use_order ? norder : G.n;
    G.order_nslot = 0;
    hipLaunchKernelGGL(k_fill_int3, dim3(std::max(1, std::min(grid1(G.slot_loc.n, 256), 1024))), dim3(256), 0, s,
                       G.col_count.p, G.ncol + 1, 0, G.fill.p, G.ncol + 1, 0, G.slot_loc.p,
                       use_order ? 0 : (int)G.slot_loc.n, -1);
    if (G.n > 0)
      hipLaunchKernelGGL(k_bin, dim3(grid1(nthr, 256)), dim3(256), 0, s, G.n, pos_dev, G.off, (const SysArg*)G.sys.p,
                         (const int*)nullptr, G.info.p, G.key.p, G.col_count.p, c->totals.p + wslot,
                         c->totals.p + (25 - wslot), order, norder);
    // atom and cluster starts of the columns; totals[4 + 2 gi], [5 + 2 gi]
    if ((rc = c->scan_bsum.reserve(2 * (size_t)grid1(std::max(G.ncol, 1), kScanBlock) + 2))) return rc;
    launch_list_scans(s, G.ncol, nullptr, G.col_count.p, CL, c->scan_bsum.p, G.col_start.p, G.col_cl.p, c->totals.p,
                      4 + 2 * gi, 5 + 2 * gi);
    if (G.n > 0) {
      hipLaunchKernelGGL(k_scatter, dim3(grid1(nthr, 256)), dim3(256), 0, s, G.n, G.info.p, G.col_start.p, G.fill.p,
                         G.tmp.p, order, norder);
      hipLaunchKernelGGL(k_rank, dim3(G.ncol), dim3(256), 0, s, G.ncol, G.key.p, G.col_start.p, G.col_cl.p,
                         G.tmp.p, G.slot_loc.p, G.info.p, pos_dev, G.off, (const SysArg*)G.sys.p,
                         (const int*)nullptr, G.origin.p, G.bb0.p, G.kr.p, (int*)nullptr, G.cl_col.p, G.p.p, G.bb.p,
                         pack_scale(c->spec.sw));
    }
    if ((rc = launch_check("rebuild: binning"))) return rc;
  }
  // cluster counts stay on the device (totals[5 + 2 g]) until the one host
  // read below; the kernels in between are sized by the cluster capacity
  // full cluster-pair lists: one direction for a single group, A->B (the
  // counting side) and B->A for two groups
  c->ndir = ngroups;
  int64_t tot_entries = 0;
  const int wpb = 256 / WAVE;
  const int chunk = c->tune.chunk;
  ListArgs Ls[2];
  for (int d = 0; d < c->ndir; d++) {
    Dir& D = c->dir[d];
    D.I = (ngroups == 2) ? d : 0;
    D.J = (ngroups == 2) ? 1 - d : 0;
    Group& GI = c->g[D.I];
    Group& GJ = c->g[D.J];
    ListArgs& L = Ls[d];
    const int* ncl_dev = c->totals.p + 5 + 2 * D.I;
    L.nclI = GI.maxcl;
    L.nclI_dev = ncl_dev;
    L.mode = (ngroups == 1) ? 0 : (d == 0 ? 1 : 2);
    L.ent_cap = 0;
    L.rlist2 = (float)(c->rlist * c->rlist);
    L.pscale = pack_scale(c->spec.sw);
    if ((rc = D.ecount.reserve(GI.maxcl + 1)) || (rc = D.estart.reserve(GI.maxcl + 1)) ||
        (rc = D.istart.reserve(GI.maxcl + 1)) || (rc = c->scan_bsum.reserve(2 * (size_t)grid1(GI.maxcl, kScanBlock) + 2)) ||
        (rc = D.ell.reserve((size_t)GI.maxcl * kEll)))
      return rc;
    hipLaunchKernelGGL(k_list<0>, dim3(grid1(GI.maxcl, wpb)), dim3(256), 0, s, L, GI.cl_col.p, (const int*)nullptr,
                       (const SysArg*)GI.sys.p, (const SysArg*)GJ.sys.p, GI.bb0.p, GJ.bb0.p, GI.kr.p, GJ.kr.p,
                       GJ.col_cl.p, GI.origin.p, GJ.origin.p, D.ecount.p, (const int*)nullptr, (Entry*)nullptr, D.ell.p);
    launch_list_scans(s, GI.maxcl, ncl_dev, D.ecount.p, chunk, c->scan_bsum.p, D.estart.p, D.istart.p, c->totals.p,
                      8 + 2 * d, 9 + 2 * d);
  }
  if ((rc = launch_check("rebuild: list count"))) return rc;
  int* hostT = c->host_ro;
  P2A_HIP(hipMemcpyAsync(hostT, c->totals.p, 16 * sizeof(int), hipMemcpyDeviceToHost, s));
  P2A_HIP(hipStreamSynchronize(s));
  for (int gi = 0; gi < ngroups; gi++) {
    Group& G = c->g[gi];
    G.ncl = hostT[5 + 2 * gi];
    G.order_nslot = G.ncl * CL;
    G.order_ptr = G.slot_loc.p;
  }
  c->max_wrap = hostT[wslot];
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
        (rc = D.fpart.reserve((size_t)CL * 3 * std::max(D.nitems, 1))) ||
        (rc = D.bflag.reserve(std::max(D.nitems, 1))))
      return rc;
    Ls[d].ent_cap = (int)D.ents.n;
    if (GI.ncl > 0) {
      hipLaunchKernelGGL(k_list<1>, dim3(grid1(GI.ncl, wpb)), dim3(256), 0, s, Ls[d], GI.cl_col.p,
                         (const int*)nullptr, (const SysArg*)GI.sys.p, (const SysArg*)GJ.sys.p, GI.bb0.p, GJ.bb0.p,
                         GI.kr.p, GJ.kr.p, GJ.col_cl.p, GI.origin.p, GJ.origin.p, D.ecount.p,
                         (const int*)D.estart.p, D.ents.p, D.ell.p);
      hipLaunchKernelGGL(k_items, dim3(grid1(GI.ncl, 256)), dim3(256), 0, s, GI.ncl, (const int*)nullptr,
                         D.ecount.p, D.estart.p, D.istart.p, D.items.p, chunk, (int)D.items.n);
    }
  }
  c->nentries = tot_entries;
  // absolute index per slot (two groups: pairs of one atom with itself are skipped)
  if (ngroups == 2) {
    for (int gi = 0; gi < 2; gi++) {
      Group& G = c->g[gi];
      if (G.ncl > 0)
        hipLaunchKernelGGL(k_abs_slot, dim3(grid1((size_t)G.ncl * CL, 256)), dim3(256), 0, s, G.ncl * CL,
                           (const int*)nullptr, G.slot_loc.p, G.off, (const long long*)c->abs_dev.p, G.abs_slot.p);
    }
  }
  if ((rc = launch_check("rebuild: list fill"))) return rc;
  if (c->skin_active > 0.0) {
    const size_t n3 = 3 * (size_t)c->natoms;
    if ((rc = c->pos_ref.reserve(n3))) return rc;
    P2A_HIP(hipMemcpyAsync(c->pos_ref.p, pos_dev, n3 * sizeof(double), hipMemcpyDeviceToDevice, s));
  }
  std::memcpy(c->box_at_build, box9, sizeof(c->box_at_build));
  std::memcpy(c->box_at_rebuild, box9, sizeof(c->box_at_rebuild));
  c->built = true;
  c->rebuilds++;
  // k_rank has packed these positions for the new list
  c->packed_pos = pos_dev;
  c->packed_gen = c->rebuilds;
  return P2A_OK;
}

// prepared cell for this box, on the host (fallback kernels take it by value)
// and on the device (the pair kernel reads it for its FP64 cut decisions);
// the device copy is written by a kernel that takes it as an argument, so
// a new box costs no host wait
static int upload_pbc(p2a_coord* c, const double* box9) {
  int rc;
  if (c->pbc_valid && std::memcmp(c->pbc_box, box9, sizeof(c->pbc_box)) == 0 && c->pbc_dev.p) return P2A_OK;
  if ((rc = c->pbc_dev.reserve(1))) return rc;
  pbc_setup(box9, c->pbc);
  std::memcpy(c->pbc_box, box9, sizeof(c->pbc_box));
  c->pbc_valid = true;
  fill_pbc_arg(c->pbc, c->pbc_stage);
  hipLaunchKernelGGL(k_store<PbcArg>, dim3(1), dim3(64), 0, c->stream, c->pbc_stage, c->pbc_dev.p);
  return launch_check("cell upload");
}

// A kept neighbour list under a new cell (NLIST / NLISTCELLS between two
// rebuilds, as NeighborList keeps its pair indices across box changes,
// NeighborList.cpp:433-456): the packed positions and the entry offsets take
// the lattice vectors of the new cell with the integer images chosen at the
// rebuild.  For a pair closer than D_MAX that image is the minimum image of
// the new cell, since every cell width stays above twice the list radius
// (choose_path), which is what Pbc::distance finds for the reference.
// Whether the kept list, imaged for box9, still holds every pair of the
// reference's list that comes within D_MAX.  A listed pair is measured as
// x_j - x_i + k.a with integers k of at most 2 max_wrap + 1 per component
// (wrapping of both atoms plus the cluster image); it is the minimum image of
// the new cell whenever it is shorter than D_MAX (the cell stays wider than
// twice the list radius), and a pair whose minimum image now lies within
// D_MAX was within D_MAX + 2 displacement + |k.(a' - a)| at the rebuild.
// The cell change may use a quarter of NL_CUTOFF - D_MAX for the last term;
// beyond that the list is rebuilt.
static bool reimage_exact(const p2a_coord* c, const double* box9) {
  const double margin = c->spec.nl_cutoff - c->spec.sw.dmax;
  if (!(margin > 0.0)) return false;
  double drift = 0.0;
  for (int r = 0; r < 3; r++) {
    double q = 0.0;
    for (int k = 0; k < 3; k++) {
      const double t = box9[3 * r + k] - c->box_at_rebuild[3 * r + k];
      q += t * t;
    }
    drift += std::sqrt(q);
  }
  return (2.0 * c->max_wrap + 1.0) * drift <= 0.25 * margin;
}

static int reimage(p2a_coord* c, const double* box9) {
  hipStream_t s = c->stream;
  const int ngroups = c->spec.nb > 0 ? 2 : 1;
  BoxArg B;
  std::memcpy(B.a, box9, sizeof(B.a));
  for (int gi = 0; gi < ngroups; gi++) {
    Group& G = c->g[gi];
    SysArg S = G.S;  // the cell grid of the rebuild stays; only the lattice changes
    sys_cell(box9, c->rlist, S);
    hipLaunchKernelGGL(k_store<SysArg>, dim3(1), dim3(64), 0, s, S, G.sys.p);
    G.S_dev = S;
  }
  const int wpb = 256 / WAVE;
  for (int d = 0; d < c->ndir; d++) {
    Dir& D = c->dir[d];
    Group& GI = c->g[D.I];
    Group& GJ = c->g[D.J];
    if (GI.ncl == 0) continue;
    hipLaunchKernelGGL(k_reimage, dim3(grid1(GI.ncl, wpb)), dim3(256), 0, s, GI.ncl, (const int*)D.estart.p,
                       (const int*)D.ecount.p, (int)D.ents.n, (const double4*)GI.origin.p,
                       (const double4*)GJ.origin.p, B, pack_scale(c->spec.sw), D.ents.p);
  }
  std::memcpy(c->box_at_build, box9, sizeof(c->box_at_build));
  return launch_check("reimage");
}

// value / virial / pair count from nred rows of c->partial, in one workgroup
// and a fixed order; with gather_dir >= 0 the same launch also forms that
// direction's derivatives
static int reduce_partials(p2a_coord* c, int nred, double* value_dev, double* virial_dev,
                           int gather_dir = -1, double* deriv_dev = nullptr, bool with_stats = true,
                           const int* skip = nullptr) {
  hipStream_t s = c->stream;
  int ngather = 0, ncl = 0;
  const float4* gp = nullptr;
  const int* istart = nullptr;
  int items_cap = 0;
  const double* fpart = nullptr;
  if (gather_dir >= 0) {
    Dir& D = c->dir[gather_dir];
    Group& GI = c->g[D.I];
    ncl = GI.ncl;
    ngather = grid1((size_t)ncl * CL * 3, 256);
    gp = GI.p.p;
    istart = D.istart.p;
    items_cap = (int)D.items.n;
    fpart = D.fpart.p;
  }
  RedArgs R;
  std::memset(&R, 0, sizeof(R));
  R.nred = nred;
  R.partial = c->partial.p;
  R.value = value_dev;
  R.virial = virial_dev;
  R.stats = with_stats ? c->stats.p : nullptr;
  R.skip = skip;
  hipLaunchKernelGGL(k_gather_reduce, dim3(ngather + 1), dim3(256), 0, s, ngather, ncl, (const int*)nullptr, gp,
                     istart, items_cap, fpart, deriv_dev, 1, R);
  return P2A_OK;
}

// FP32 cluster positions and boxes of this step's positions, unless the
// skin test already packed them for the list in use
static void pack_positions(p2a_coord* c, const double* pos_dev) {
  const bool ahead = c->packed_pos == pos_dev && c->packed_gen == c->rebuilds;
  c->packed_pos = nullptr;
  if (ahead) return;
  const int ngroups = c->spec.nb > 0 ? 2 : 1;
  for (int gi = 0; gi < ngroups; gi++) {
    Group& G = c->g[gi];
    if (G.ncl == 0) continue;
    hipLaunchKernelGGL(k_pack_bbox, dim3(grid1((size_t)G.ncl * CL, 256)), dim3(256), 0, c->stream, G.ncl * CL,
                       (const int*)nullptr, G.slot_loc.p, G.info.p, pos_dev, G.off, (const SysArg*)G.sys.p,
                       (const int*)nullptr, G.origin.p, G.p.p, G.bb.p, pack_scale(c->spec.sw));
  }
}

// skip: a device flag; when it reads nonzero at run time the pair and
// reduction kernels do nothing (a pass queued before the skin test's answer)
static int run_pairs(p2a_coord* c, const double* box9, const double* pos_dev, double* value_dev, double* deriv_dev,
                     double* virial_dev, bool value_only, const int* skip = nullptr) {
  int rc;
  hipStream_t s = c->stream;
  const int ngroups = c->spec.nb > 0 ? 2 : 1;
  pack_positions(c, pos_dev);
  PairArgs A;
  std::memset(&A, 0, sizeof(A));
  A.twogroup = ngroups == 2;
  pair_args_switch(c->spec.sw, A);
  if ((rc = upload_pbc(c, box9))) return rc;
  A.pbc = c->pbc_dev.p;
  A.cl_sys = nullptr;
  A.nitems_dev = nullptr;
  A.pos64 = pos_dev;
  A.skip = skip;
  A.item_partials = 0;  // value / virial / pair count per workgroup
  // value only: the counting direction (the first) holds every counted pair
  const int ndir_run = value_only ? 1 : c->ndir;
  int grids[2] = {0, 0}, nred = 0;
  for (int d = 0; d < ndir_run; d++) {
    Dir& D = c->dir[d];
    if (D.nitems == 0) continue;
    int grid = grid1(D.nitems, PAIR_BLOCK / WAVE);
    if (c->tune.blocks_per_cu > 0) grid = std::min(grid, c->tune.blocks_per_cu * c->cu_count);
    grids[d] = grid;
    nred += grid;
  }
  if ((rc = c->partial.reserve(8 * (size_t)std::max(nred, 1)))) return rc;
  hipEvent_t e0 = nullptr, e1 = nullptr;
  const bool rec = c->profiling && (c->prof_calls++ % c->prof_every) == 0;
  if (rec) {
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
  int off = 0;
  for (int d = 0; d < ndir_run; d++) {
    Dir& D = c->dir[d];
    if (grids[d] == 0) continue;
    Group& GI = c->g[D.I];
    Group& GJ = c->g[D.J];
    A.nitems = D.nitems;
    A.items_cap = (int)D.items.n;
    A.ent_cap = (int)D.ents.n;
    A.bflag = D.bflag.p;
    A.wave_start = nullptr;
    if (c->tune.blocks_per_cu > 0) {
      // a persistent grid takes contiguous item ranges of balanced cost
      const int nw = grids[d] * (PAIR_BLOCK / WAVE);
      if (D.sched_waves != nw || D.sched_gen != c->rebuilds) {
        if ((rc = D.wave_start.reserve((size_t)nw + 1))) return rc;
        hipLaunchKernelGGL(k_wave_sched, dim3(grid1(std::max(D.nitems, 1), 256)), dim3(256), 0, s, D.nitems,
                           (const int*)nullptr, (int)D.items.n, (const int4*)D.items.p, nw, D.wave_start.p);
        D.sched_waves = nw;
        D.sched_gen = c->rebuilds;
      }
      A.wave_start = D.wave_start.p;
    }
    launch_pairs(A, grids[d], s, value_only, c->tune.chunk, D.items.p, D.ents.p, GI.p.p, GJ.p.p, GI.bb.p, GJ.bb.p,
                 GI.abs_slot.p, GJ.abs_slot.p, D.fpart.p, c->partial.p + 8 * (size_t)off);
    off += grids[d];
  }
  if (rec) {
    P2A_HIP(hipEventRecord(e1, s));
    c->pending.emplace_back(e0, e1);
  }
  if ((rc = launch_check("pair kernel"))) return rc;
  // derivatives of the last direction go with the reduction
  const int last = value_only ? -1 : c->ndir - 1;
  for (int d = 0; d < last; d++) {
    Dir& D = c->dir[d];
    Group& GI = c->g[D.I];
    if (GI.ncl == 0) continue;
    const int ng = grid1((size_t)GI.ncl * CL * 3, 256);
    RedArgs R;
    std::memset(&R, 0, sizeof(R));
    R.skip = skip;
    hipLaunchKernelGGL(k_gather_reduce, dim3(ng), dim3(256), 0, s, ng, GI.ncl, (const int*)nullptr, GI.p.p,
                       D.istart.p, (int)D.items.n, D.fpart.p, deriv_dev, 0, R);
  }
  if ((rc = reduce_partials(c, nred, value_dev, virial_dev, last, deriv_dev, true, skip))) return rc;
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
  if ((rc = reduce_partials(c, std::max(nblocks, 1), value_dev, virial_dev))) return rc;
  return launch_check("finalize");
}

// LinkCells grid of the reference's cells branch for this cell and these
// positions (LinkCells::setupCells / createCells, LinkCells.cpp:49-122): the
// given cell, or, when the box is all zero, an orthorhombic box of whole
// NL_CUTOFF multiples around the atoms (whose extent needs the positions on
// the host: this rare case waits for a copy)
static int cell_grid(p2a_coord* c, const double* box9, const double* pos_dev, CellGrid& G) {
  std::memset(&G, 0, sizeof(G));
  const double cut = c->spec.nl_cutoff;
  double box[9];
  bool zero = true;
  for (int k = 0; k < 9; k++) zero = zero && box9[k] == 0.0;
  if (zero) {
    const size_t n3 = 3 * (size_t)c->natoms;
    std::vector<double> x(n3);
    P2A_HIP(hipMemcpyAsync(x.data(), pos_dev, n3 * sizeof(double), hipMemcpyDeviceToHost, c->stream));
    P2A_HIP(hipStreamSynchronize(c->stream));
    std::memset(box, 0, sizeof(box));
    for (int k = 0; k < 3; k++) {
      double lo = x[k], hi = x[k];
      for (int i = 1; i < c->natoms; i++) {
        if (x[3 * (size_t)i + k] > hi) hi = x[3 * (size_t)i + k];
        if (x[3 * (size_t)i + k] < lo) lo = x[3 * (size_t)i + k];
      }
      box[4 * k] = cut < std::sqrt(DBL_MAX) ? cut * (1 + std::ceil((hi - lo) / cut)) : hi - lo + 1;
      G.origin[k] = (lo + hi) / 2;
    }
    G.shift_origin = 1;
  } else {
    std::memcpy(box, box9, sizeof(box));
  }
  PbcHost h;
  pbc_setup(box, h);
  if (h.type == 0 && !zero) return fail(P2A_ERR_ARG, "cell lists cannot be built for a box with null volume");
  std::memcpy(G.inv, h.inv, sizeof(G.inv));
  for (int k = 0; k < 3; k++) {
    // reciprocal row k = column k of the inverse
    const double len = std::sqrt(h.inv[k] * h.inv[k] + h.inv[3 + k] * h.inv[3 + k] + h.inv[6 + k] * h.inv[6 + k]);
    const int n = (int)std::floor(1.0 / len / cut);
    G.nc[k] = n == 0 ? 1 : n;
  }
  G.use_pbc = c->spec.pbc ? 1 : 0;
  return P2A_OK;
}

// The FP64 path with NLIST / NLISTCELLS: the reference's candidate pairs are
// listed on the first step and every NL_STRIDE steps (NeighborList::prepare,
// NeighborList.cpp:433-456) and evaluated, with the current positions and
// cell, until the next rebuild.  Classic: pairs with |d|^2 <= NL_CUTOFF^2
// from Pbc::distance (NeighborList.cpp:237-308); cells: the pairs of atoms
// whose link cells are within one stencil (NeighborList.cpp:177-236).
// build: list the pairs at these positions first; value_dev null: no evaluation
static int run_fallback_nl(p2a_coord* c, bool build, const double* box9, const double* pos_dev,
                           double* value_dev, double* deriv_dev, double* virial_dev) {
  hipStream_t s = c->stream;
  int rc;
  if ((rc = upload_pbc(c, box9))) return rc;
  const PbcArg& P = c->pbc_stage;
  const int nA = (int)c->spec.na, nB = (int)c->spec.nb;
  const int natoms = c->natoms;
  const double cut2 = c->spec.nl_cutoff * c->spec.nl_cutoff;
  const int use_pbc = c->spec.pbc ? 1 : 0;
  if (build || !c->nl_valid) {
    if (c->spec.pair) {
      if ((rc = c->nl_count.reserve(std::max(nA, 1)))) return rc;
      hipLaunchKernelGGL(k_nl_pairs_keep, dim3(grid1(std::max(nA, 1), 256)), dim3(256), 0, s, nA, use_pbc, P, cut2,
                         pos_dev, c->nl_count.p);
    } else {
      CellGrid G;
      std::memset(&G, 0, sizeof(G));
      const int mode = c->spec.nl_mode;
      if ((rc = c->nl_count.reserve(natoms)) || (rc = c->nl_start.reserve((size_t)natoms + 1)) ||
          (rc = c->nl_total.reserve(1)))
        return rc;
      if (mode == P2A_NL_CELLS) {
        if ((rc = cell_grid(c, box9, pos_dev, G))) return rc;
        if ((rc = c->nl_cell.reserve(3 * (size_t)natoms))) return rc;
        hipLaunchKernelGGL(k_nl_cells, dim3(grid1(natoms, 256)), dim3(256), 0, s, natoms, G, pos_dev, c->nl_cell.p);
      }
      hipLaunchKernelGGL(k_nl_atoms<0>, dim3(grid1(natoms, 256)), dim3(256), 0, s, nA, nB, mode, use_pbc, P, cut2, G,
                         pos_dev, (const int*)c->nl_cell.p, c->nl_count.p, (const int*)nullptr, (int*)nullptr);
      hipLaunchKernelGGL(k_scan_small, dim3(1), dim3(1024), 0, s, natoms, (const int*)c->nl_count.p, c->nl_start.p,
                         c->nl_total.p);
      long long total = 0;
      P2A_HIP(hipMemcpyAsync(c->host_ro + 40, c->nl_total.p, sizeof(long long), hipMemcpyDeviceToHost, s));
      P2A_HIP(hipStreamSynchronize(s));
      std::memcpy(&total, c->host_ro + 40, sizeof(total));
      if (total > INT32_MAX) return fail(P2A_ERR_UNSUPPORTED, "neighbour list too long for the FP64 path");
      if ((rc = c->nl_nbr.reserve(std::max<long long>(total, 1)))) return rc;
      hipLaunchKernelGGL(k_nl_atoms<1>, dim3(grid1(natoms, 256)), dim3(256), 0, s, nA, nB, mode, use_pbc, P, cut2, G,
                         pos_dev, (const int*)c->nl_cell.p, (int*)nullptr, (const int*)c->nl_start.p, c->nl_nbr.p);
    }
    if ((rc = launch_check("FP64 neighbour list"))) return rc;
    c->nl_valid = true;
    c->rebuilds++;
  }
  if (!value_dev) return P2A_OK;
  int nblocks;
  if (c->spec.pair) {
    nblocks = grid1(nA, 256);
    if ((rc = c->partial.reserve(8 * (size_t)std::max(nblocks, 1)))) return rc;
    hipLaunchKernelGGL(k_pairmode, dim3(std::max(nblocks, 1)), dim3(256), 0, s, nA, c->spec.pbc, P, c->spec.sw,
                       pos_dev, c->abs_dev.p, deriv_dev, c->partial.p, (const int*)c->nl_count.p);
  } else {
    nblocks = grid1(natoms, 256);
    if ((rc = c->partial.reserve(8 * (size_t)std::max(nblocks, 1)))) return rc;
    hipLaunchKernelGGL(k_allpairs, dim3(std::max(nblocks, 1)), dim3(256), 0, s, nA, nB, c->spec.pbc, P, c->spec.sw,
                       pos_dev, c->abs_dev.p, deriv_dev, c->partial.p, (const int*)c->nl_start.p,
                       (const int*)c->nl_nbr.p);
  }
  if ((rc = launch_check("fp64 pair kernel"))) return rc;
  if ((rc = reduce_partials(c, std::max(nblocks, 1), value_dev, virial_dev))) return rc;
  return launch_check("finalize");
}

// COORDINATIONNUMBER rows of GROUPA over the environment (GROUPA itself, or
// GROUPB) on the cluster list, and with weights their vector-Jacobian
// product and box derivative
static int run_rows(p2a_coord* c, const double* box9, const double* pos_dev, double* rows_dev,
                    const double* weights_dev, double* vjp_dev, double* virial_dev) {
  int rc;
  hipStream_t s = c->stream;
  const int ngroups = c->spec.nb > 0 ? 2 : 1;
  pack_positions(c, pos_dev);
  PairArgs A;
  std::memset(&A, 0, sizeof(A));
  A.twogroup = ngroups == 2;
  pair_args_switch(c->spec.sw, A);
  if ((rc = upload_pbc(c, box9))) return rc;
  A.pbc = c->pbc_dev.p;
  A.pos64 = pos_dev;
  const bool weighted = weights_dev != nullptr;
  if (weighted) {
    if ((rc = c->wreq.reserve(std::max(c->natoms, 1)))) return rc;
    hipLaunchKernelGGL(k_row_weights, dim3(grid1(c->natoms, 256)), dim3(256), 0, s, (int)c->spec.na, c->natoms,
                       weights_dev, c->wreq.p);
  }
  // the rows come from the first direction (i = GROUPA); the weighted
  // product needs every direction for the derivatives of all atoms
  const int ndir_run = weighted ? c->ndir : 1;
  int grids[2] = {0, 0}, nred = 0;
  for (int d = 0; d < ndir_run; d++) {
    Dir& D = c->dir[d];
    if (D.nitems == 0) continue;
    int grid = grid1(D.nitems, PAIR_BLOCK / WAVE);
    if (c->tune.blocks_per_cu > 0) grid = std::min(grid, c->tune.blocks_per_cu * c->cu_count);
    grids[d] = grid;
    nred += grid;
  }
  if ((rc = c->partial.reserve(8 * (size_t)std::max(nred, 1)))) return rc;
  int off = 0;
  for (int d = 0; d < ndir_run; d++) {
    Dir& D = c->dir[d];
    if (grids[d] == 0) continue;
    Group& GI = c->g[D.I];
    Group& GJ = c->g[D.J];
    if ((rc = D.rpart.reserve((size_t)CL * std::max(D.nitems, 1)))) return rc;
    A.nitems = D.nitems;
    A.items_cap = (int)D.items.n;
    A.ent_cap = (int)D.ents.n;
    launch_rows(A, weighted, grids[d], s, c->wreq.p, D.items.p, D.ents.p, GI.p.p, GJ.p.p, GI.bb.p, GJ.bb.p,
                GI.abs_slot.p, GJ.abs_slot.p, D.rpart.p, D.fpart.p, c->partial.p + 8 * (size_t)off);
    off += grids[d];
  }
  if ((rc = launch_check("row kernel"))) return rc;
  {
    Dir& D = c->dir[0];
    Group& GI = c->g[D.I];
    if (GI.ncl > 0 && D.nitems > 0)
      hipLaunchKernelGGL(k_gather_rows, dim3(grid1((size_t)GI.ncl * CL, 256)), dim3(256), 0, s, GI.ncl * CL, GI.p.p,
                         (const int*)D.istart.p, (int)D.items.n, (const double*)D.rpart.p, (int)c->spec.na,
                         rows_dev);
    else
      hipLaunchKernelGGL(k_fill_double, dim3(grid1((size_t)c->spec.na, 256)), dim3(256), 0, s, rows_dev,
                         (size_t)c->spec.na);
  }
  if (!weighted) return launch_check("rows");
  if ((rc = c->out_buf.reserve(16))) return rc;
  const int last = c->ndir - 1;
  for (int d = 0; d < last; d++) {
    Dir& D = c->dir[d];
    Group& GI = c->g[D.I];
    if (GI.ncl == 0) continue;
    const int ng = grid1((size_t)GI.ncl * CL * 3, 256);
    RedArgs R;
    std::memset(&R, 0, sizeof(R));
    hipLaunchKernelGGL(k_gather_reduce, dim3(ng), dim3(256), 0, s, ng, GI.ncl, (const int*)nullptr, GI.p.p,
                       D.istart.p, (int)D.items.n, D.fpart.p, vjp_dev, 0, R);
  }
  double* wdev = virial_dev ? virial_dev : c->out_buf.p + 1;
  if ((rc = reduce_partials(c, nred, c->out_buf.p, wdev, last, vjp_dev, false))) return rc;
  return launch_check("row derivatives");
}

// The cluster list for this step: NeighborList::prepare (NeighborList.cpp:
// 433-456) rebuilds on the first step and every NL_STRIDE steps.  Without a
// neighbour list every step, or, with a skin, when some atom has moved more
// than skin/2 since the build: every pair within D_MAX now was within D_MAX +
// skin then, so the pairs evaluated are those of a list rebuilt now.
// The skin test on the device: moved = 1 when some atom has moved more than
// skin/2 since the build; the flag is copied to host_ro[32] and moved_ev
// recorded after the copy.
static int launch_moved_test(p2a_coord* c, const double* pos_dev) {
  int rc;
  if ((rc = c->moved.reserve(1))) return rc;
  if (!c->moved_ev) P2A_HIP(hipEventCreateWithFlags(&c->moved_ev, hipEventDisableTiming));
  const double half = 0.5 * c->skin_active;
  hipLaunchKernelGGL(k_fill_int, dim3(1), dim3(64), 0, c->stream, c->moved.p, 1, 0);
  hipLaunchKernelGGL(k_moved, dim3(grid1((size_t)c->natoms, 256)), dim3(256), 0, c->stream, c->natoms, pos_dev,
                     (const double*)c->pos_ref.p, half * half, c->moved.p);
  P2A_HIP(hipMemcpyAsync(c->host_ro + 32, c->moved.p, sizeof(int), hipMemcpyDeviceToHost, c->stream));
  P2A_HIP(hipEventRecord(c->moved_ev, c->stream));
  return launch_check("skin test");
}

// A list that the skin test may keep: no neighbour list, a skin in use, a
// list built on an earlier call in this same cell.
static bool skin_kept_candidate(p2a_coord* c, const double* box9) {
  return c->spec.nl_mode == P2A_NL_NONE && c->skin_active > 0.0 && c->built && !c->firsttime &&
         std::memcmp(c->box_at_build, box9, sizeof(c->box_at_build)) == 0;
}

static int prepare_cluster_list(p2a_coord* c, int64_t step, const double* box9, const double* pos_dev) {
  int rc;
  bool need = !c->built || c->firsttime;
  const bool new_cell = !need && std::memcmp(c->box_at_build, box9, sizeof(c->box_at_build)) != 0;
  if (c->spec.nl_mode == P2A_NL_NONE) {
    // the skin test measures motion in one cell: a new cell rebuilds
    if (new_cell) need = true;
    if (!need && c->skin_active > 0.0) {
      if ((rc = launch_moved_test(c, pos_dev))) return rc;
      // the positions are packed for the kept list while the host waits for
      // the answer (a rebuild packs them again)
      pack_positions(c, pos_dev);
      c->packed_pos = pos_dev;
      c->packed_gen = c->rebuilds;
      P2A_HIP(hipEventSynchronize(c->moved_ev));
      need = c->host_ro[32] != 0;
    } else {
      need = true;
    }
  } else if (c->spec.nl_stride <= 1 || step % c->spec.nl_stride == 0) {
    need = true;
  }
  c->firsttime = false;
  if (!need && new_cell && !reimage_exact(c, box9)) need = true;
  if (need) {
    if ((rc = rebuild(c, box9, pos_dev))) return rc;
  } else if (new_cell) {
    if ((rc = reimage(c, box9))) return rc;
  }
  return P2A_OK;
}

static int calc_device(p2a_coord* c, int64_t step, const double* box9, const double* pos_dev, double* value_dev,
                       double* deriv_dev, double* virial_dev) {
  int rc;
  if (!box9) return fail(P2A_ERR_ARG, "box is required");
  // positions packed ahead count only within the call that packed them (the
  // same address may hold other positions at a later call)
  c->packed_pos = nullptr;
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
  if (path != c->path) {
    // the list of the other path is not carried over: rebuild now
    c->built = false;
    c->nl_valid = false;
  }
  c->path = path;
  if (path != 0) {
    if (c->spec.nl_mode != P2A_NL_NONE) {
      const bool build = c->spec.nl_stride <= 1 || step % c->spec.nl_stride == 0;
      return run_fallback_nl(c, build, box9, pos_dev, vdev, ddev, wdev);
    }
    return run_fallback(c, box9, pos_dev, vdev, ddev, wdev);
  }
  if (skin_kept_candidate(c, box9)) {
    // The pass on the kept list is queued behind the displacement test and
    // does nothing on the device if the test finds an atom that moved too far;
    // the host waits only for the test's answer, while the pass runs.
    if ((rc = launch_moved_test(c, pos_dev))) return rc;
    if ((rc = run_pairs(c, box9, pos_dev, vdev, ddev, wdev, value_only, c->moved.p))) return rc;
    P2A_HIP(hipEventSynchronize(c->moved_ev));
    if (c->host_ro[32] == 0) return P2A_OK;
    if ((rc = rebuild(c, box9, pos_dev))) return rc;
    return run_pairs(c, box9, pos_dev, vdev, ddev, wdev, value_only);
  }
  if ((rc = prepare_cluster_list(c, step, box9, pos_dev))) return rc;
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

// COORDINATIONNUMBER rows: cn[k] = sum over env atoms j of the contact
// matrix element of (rows[k], j), ContactMatrix::calculateWeight
// (src/adjmat/ContactMatrix.h:95-121): nothing when |r|^2 < DBL_EPSILON (the
// diagonal, or two atoms at one place), s(r) zeroed below DBL_EPSILON.  One
// thread per row, FP64, environment in list order.  An index outside
// [0, natoms) makes that row NaN instead of being read.
__global__ __launch_bounds__(256) void k_cn_rows(int natoms, int nrows, const int* __restrict__ rows, int nenv,
                                                 const int* __restrict__ env, int use_pbc, PbcArg P, p2a_switch sw,
                                                 const double* __restrict__ pos, double* __restrict__ cn) {
  const int k = blockIdx.x * blockDim.x + threadIdx.x;
  if (k >= nrows) return;
  const int i = rows[k];
  if (i < 0 || i >= natoms) {
    cn[k] = __builtin_nan("");
    return;
  }
  const double* xi = pos + 3 * (size_t)i;
  double sum = 0.0;
  for (int q = 0; q < nenv; q++) {
    const int j = env[q];
    if (j < 0 || j >= natoms) {
      sum = __builtin_nan("");
      break;
    }
    const double* xj = pos + 3 * (size_t)j;
    double d[3];
    if (use_pbc) pbc_distance(P, xi, xj, d);
    else { d[0] = xj[0] - xi[0]; d[1] = xj[1] - xi[1]; d[2] = xj[2] - xi[2]; }
    const double mod2 = d[0] * d[0] + d[1] * d[1] + d[2] * d[2];
    if (mod2 < DBL_EPSILON) continue;
    double df;
    const double v = switch_calc_sqr(sw, mod2, df);
    if (v < DBL_EPSILON) continue;
    sum += v;
  }
  cn[k] = sum;
}

// Scratch of the allocation-free entry point: the per-workgroup partials of
// k_allpairs that k_finalize sums (8 doubles per 256-atom workgroup).  Each
// device holds its own copy for the life of the process; its address is
// looked up once per device and kept in g_scratch_dev.
constexpr int kNoAllocMaxAtoms = 1 << 18;
constexpr int kNoAllocBlocks = kNoAllocMaxAtoms / 256;
static __device__ __align__(256) double scratch[8 * kNoAllocBlocks];

constexpr int kMaxDevices = 64;
static std::mutex g_scratch_mu;
static double* g_scratch_dev[kMaxDevices] = {};

// device address of `scratch` on the current device (hipGetSymbolAddress,
// once per device); the host-side name is not a device address
static int noalloc_scratch(int device, double** out) {
  if (device < 0 || device >= kMaxDevices) return fail(P2A_ERR_ARG, "device index out of range");
  std::lock_guard<std::mutex> lk(g_scratch_mu);
  if (!g_scratch_dev[device]) {
    void* ptr = nullptr;
    hipError_t e = hipGetSymbolAddress(&ptr, HIP_SYMBOL(scratch));
    if (e != hipSuccess || !ptr)
      return fail(P2A_ERR_HIP, std::string("hipGetSymbolAddress(scratch): ") + hipGetErrorString(e));
    g_scratch_dev[device] = static_cast<double*>(ptr);
  }
  *out = g_scratch_dev[device];
  return P2A_OK;
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
  if (spec->pair && spec->nl_mode == P2A_NL_CELLS)
    return fail(P2A_ERR_ARG, "Cell list should not be active with a Pair NL");
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
  int ncu = 0;
  if (hipDeviceGetAttribute(&ncu, hipDeviceAttributeMultiprocessorCount, device) == hipSuccess && ncu > 0)
    c->cu_count = ncu;
  if (hipStreamCreateWithFlags(&c->own_stream, hipStreamNonBlocking) != hipSuccess) {
    delete c;
    return fail(P2A_ERR_HIP, "hipStreamCreate failed");
  }
  c->stream = c->own_stream;
  if (hipHostMalloc((void**)&c->host_ro, 64 * sizeof(int), hipHostMallocDefault) != hipSuccess) {
    c->host_ro = nullptr;
    p2a_coord_destroy(c);
    return fail(P2A_ERR_HIP, "hipHostMalloc failed");
  }
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
  if (c->moved_ev) (void)hipEventDestroy(c->moved_ev);
  if (c->own_stream) (void)hipStreamDestroy(c->own_stream);
  if (c->host_ro) (void)hipHostFree(c->host_ro);
  delete c;
  return P2A_OK;
}

int p2a_coord_set_stream(p2a_coord* c, void* stream) {
  if (!c) return fail(P2A_ERR_ARG, "null context");
  // NULL selects the default (null) stream, which orders with every blocking stream
  c->stream = (hipStream_t)stream;
  return P2A_OK;
}

int p2a_coord_set_tuning(p2a_coord* c, int chunk, int blocks_per_cu) {
  if (!c) return fail(P2A_ERR_ARG, "null context");
  if (chunk != 32 && chunk != 64) return fail(P2A_ERR_ARG, "chunk must be 32 or 64");
  if (blocks_per_cu < 0 || blocks_per_cu > 64) return fail(P2A_ERR_ARG, "blocks_per_cu must be in [0, 64]");
  // the work items depend on the chunk: the next call rebuilds them
  if (chunk != c->tune.chunk) c->built = false;
  c->tune.chunk = chunk;
  c->tune.blocks_per_cu = blocks_per_cu;
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
  c->packed_pos = nullptr;  // packed ahead by an earlier call: pack again
  int rc;
  if ((rc = c->out_buf.reserve(16))) return rc;
  double* wdev = virial_dev ? virial_dev : c->out_buf.p + 1;
  if (c->path != 0 || choose_path(c, box9) != 0) {
    // the pairs of the last listing (a neighbour list is kept as the forward left it)
    if (c->spec.nl_mode != P2A_NL_NONE) return run_fallback_nl(c, false, box9, pos_dev, c->out_buf.p, deriv_dev, wdev);
    return run_fallback(c, box9, pos_dev, c->out_buf.p, deriv_dev, wdev);
  }
  if (!c->built || std::memcmp(c->box_at_build, box9, sizeof(c->box_at_build)) != 0)
    return fail(P2A_ERR_ARG, "backward needs the neighbour structure of a forward call on the same box");
  return run_pairs(c, box9, pos_dev, c->out_buf.p, deriv_dev, wdev, false);
}

int p2a_coord_calculate_noalloc(void* stream, const p2a_switch* sw, const double* box9, int use_pbc, int64_t na,
                                int64_t nb, const int64_t* absidx_dev, const double* pos_dev, double* value_dev,
                                double* deriv_dev, double* virial_dev) {
  if (!sw || !pos_dev || !value_dev || !deriv_dev || !virial_dev) return fail(P2A_ERR_ARG, "null argument");
  if (na <= 0) return fail(P2A_ERR_ARG, "GROUPA should contain at least one atom");
  if (nb < 0) return fail(P2A_ERR_ARG, "GROUPB cannot have a negative size");
  if (na + nb > kNoAllocMaxAtoms)
    return fail(P2A_ERR_ARG, "p2a_coord_calculate_noalloc handles at most " + std::to_string(kNoAllocMaxAtoms) +
                                 " atoms (its scratch holds 8 doubles per 256 atoms)");
  if (use_pbc && !box9) return fail(P2A_ERR_ARG, "box is required with periodic boundaries");
  int dev = 0;
  P2A_HIP(hipGetDevice(&dev));
  if (stream) {
    int sdev = -1;
    P2A_HIP(hipStreamGetDevice((hipStream_t)stream, &sdev));
    if (sdev != dev) return fail(P2A_ERR_ARG, "the stream belongs to another device than the current one");
  }
  double* part = nullptr;
  int rc;
  if ((rc = noalloc_scratch(dev, &part))) return rc;
  PbcArg P;
  std::memset(&P, 0, sizeof(P));
  if (use_pbc) {
    PbcHost h;
    pbc_setup(box9, h);
    if (h.type == 0) return fail(P2A_ERR_ARG, "the box is singular");
    fill_pbc_arg(h, P);
  }
  const int natoms = (int)(na + nb);
  const int nblocks = grid1(natoms, 256);
  hipStream_t s = (hipStream_t)stream;
  // every workgroup writes its 8 partials, so nothing in scratch needs clearing
  hipLaunchKernelGGL(k_allpairs, dim3(nblocks), dim3(256), 0, s, (int)na, (int)nb, use_pbc ? 1 : 0, P, *sw, pos_dev,
                     (const long long*)absidx_dev, deriv_dev, part);
  hipLaunchKernelGGL(k_finalize, dim3(1), dim3(128), 0, s, nblocks, (const double*)part, value_dev, virial_dev,
                     (double*)nullptr);
  return launch_check("p2a_coord_calculate_noalloc");
}

int p2a_coord_list_generation(p2a_coord* c, int64_t* gen) {
  if (!c || !gen) return fail(P2A_ERR_ARG, "null argument");
  // the host counts the rebuilds it enqueues: no synchronisation
  *gen = c->rebuilds;
  return P2A_OK;
}

int p2a_coord_rebuild_device(p2a_coord* c, const double* box9, const double* pos_dev) {
  if (!c || !box9 || !pos_dev) return fail(P2A_ERR_ARG, "null argument");
  P2A_HIP(hipSetDevice(c->device));
  int rc;
  if (!c->stats.p) {
    if ((rc = c->stats.reserve(8))) return rc;
    hipLaunchKernelGGL(k_fill_double, dim3(1), dim3(64), 0, c->stream, c->stats.p, (size_t)8);
  }
  const int path = choose_path(c, box9);
  if (path != c->path) {
    c->built = false;
    c->nl_valid = false;
  }
  c->path = path;
  if (path != 0) {
    // without a neighbour list the FP64 kernels keep no pair structure
    if (c->spec.nl_mode == P2A_NL_NONE) return P2A_OK;
    return run_fallback_nl(c, true, box9, pos_dev, nullptr, nullptr, nullptr);
  }
  c->firsttime = false;
  return rebuild(c, box9, pos_dev);
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

int p2a_coord_set_skin(p2a_coord* c, double skin) {
  if (!c) return fail(P2A_ERR_ARG, "null argument");
  if (!(skin >= 0.0) || !std::isfinite(skin)) return fail(P2A_ERR_ARG, "the skin must be a finite distance >= 0");
  c->skin = skin;
  c->path_cached = -1;  // the radius decides whether the cell is wide enough
  c->built = false;
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
  if (on < 0) return fail(P2A_ERR_ARG, "the profiling interval must be >= 0");
  c->profiling = on != 0;
  if (on) {
    c->prof_every = on;
    c->prof_calls = 0;
    c->prof_ms = 0.0;
    c->prof_launches = 0;
    if (c->stats.p) {
      hipLaunchKernelGGL(k_fill_double, dim3(1), dim3(64), 0, c->stream, c->stats.p + 1, (size_t)1);
      P2A_HIP(hipStreamSynchronize(c->stream));
    }
  }
  return P2A_OK;
}

int p2a_coord_rows_device(p2a_coord* c, int64_t step, const double* box9, const double* pos_dev, double* rows_dev,
                          const double* weights_dev, double* vjp_dev, double* virial_dev) {
  if (!c || !box9 || !pos_dev || !rows_dev) return fail(P2A_ERR_ARG, "null argument");
  if (weights_dev && !vjp_dev) return fail(P2A_ERR_ARG, "weights need an output for their derivatives");
  if (c->spec.pair) return fail(P2A_ERR_UNSUPPORTED, "PAIR gives no per-atom rows");
  P2A_HIP(hipSetDevice(c->device));
  c->packed_pos = nullptr;  // packed ahead by an earlier call: pack again
  int rc;
  if (!c->stats.p) {
    if ((rc = c->stats.reserve(8))) return rc;
    hipLaunchKernelGGL(k_fill_double, dim3(1), dim3(64), 0, c->stream, c->stats.p, (size_t)8);
  }
  const int path = choose_path(c, box9);
  if (path != c->path) {
    c->built = false;
    c->nl_valid = false;
  }
  c->path = path;
  if (path != 0) {
    if (weights_dev)
      return fail(P2A_ERR_UNSUPPORTED,
                  "derivatives of per-atom coordination numbers need the cluster path (a periodic cell wider than "
                  "twice the list radius, distinct atoms)");
    // every environment atom of every row, in FP64
    const int na = (int)c->spec.na, nb = (int)c->spec.nb;
    const int nenv = nb > 0 ? nb : na;
    if ((rc = c->cn_rows.reserve(na)) || (rc = c->cn_env.reserve(nenv))) return rc;
    std::vector<int> r(na), e(nenv);
    for (int k = 0; k < na; k++) r[k] = k;
    for (int k = 0; k < nenv; k++) e[k] = nb > 0 ? na + k : k;
    P2A_HIP(hipMemcpyAsync(c->cn_rows.p, r.data(), na * sizeof(int), hipMemcpyHostToDevice, c->stream));
    P2A_HIP(hipMemcpyAsync(c->cn_env.p, e.data(), nenv * sizeof(int), hipMemcpyHostToDevice, c->stream));
    if ((rc = upload_pbc(c, box9))) return rc;
    hipLaunchKernelGGL(k_cn_rows, dim3(grid1((size_t)na, 256)), dim3(256), 0, c->stream, c->natoms, na,
                       (const int*)c->cn_rows.p, nenv, (const int*)c->cn_env.p, c->spec.pbc ? 1 : 0, c->pbc_stage,
                       c->spec.sw, pos_dev, rows_dev);
    // the host vectors are read by the copies above before they return
    P2A_HIP(hipStreamSynchronize(c->stream));
    return launch_check("coordination numbers");
  }
  if ((rc = prepare_cluster_list(c, step, box9, pos_dev))) return rc;
  return run_rows(c, box9, pos_dev, rows_dev, weights_dev, vjp_dev, virial_dev);
}

int p2a_coordination_numbers_device(void* stream, const p2a_switch* sw, const double* box9, int use_pbc,
                                    int64_t natoms, int64_t nrows, const int* rows_dev, int64_t nenv,
                                    const int* env_dev, const double* pos_dev, double* cn_dev) {
  if (!sw) return fail(P2A_ERR_ARG, "null argument");
  if (natoms < 0 || nrows < 0 || nenv < 0 || natoms > INT32_MAX || nrows > INT32_MAX || nenv > INT32_MAX)
    return fail(P2A_ERR_ARG, "sizes must lie in [0, 2^31)");
  // no rows: nothing to write (an empty output may have a null address)
  if (nrows == 0) return P2A_OK;
  if (!pos_dev || !cn_dev || !rows_dev || (nenv > 0 && !env_dev)) return fail(P2A_ERR_ARG, "null argument");
  if (use_pbc && !box9) return fail(P2A_ERR_ARG, "box is required with periodic boundaries");
  PbcArg P;
  std::memset(&P, 0, sizeof(P));
  if (use_pbc) {
    PbcHost h;
    pbc_setup(box9, h);
    fill_pbc_arg(h, P);
  }
  hipLaunchKernelGGL(k_cn_rows, dim3(grid1((size_t)nrows, 256)), dim3(256), 0, (hipStream_t)stream, (int)natoms,
                     (int)nrows, rows_dev, (int)nenv, env_dev, use_pbc ? 1 : 0, P, *sw, pos_dev, cn_dev);
  return launch_check("coordination numbers");
}

}  // extern "C"
