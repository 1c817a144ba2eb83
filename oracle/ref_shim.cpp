// C entry points over the reference's own tools classes, compiled from the
// reference sources where they lie (see oracle/Makefile).  TEST INFRASTRUCTURE
// ONLY: the tests use this library to pin the FP64 oracle; no product code
// loads it.  The pair loop in ref_coordination restates
// src/colvar/CoordinationBase.cpp:142-232 around the reference NeighborList,
// Pbc and SwitchingFunction objects (the Action/Colvar layer needs the whole
// PLUMED kernel and is not built here).
#include "tools/SwitchingFunction.h"
#include "tools/Pbc.h"
#include "tools/Tensor.h"
#include "tools/NeighborList.h"
#include "tools/Communicator.h"
#include "tools/AtomNumber.h"
#include "tools/Grid.h"
#include <cstdint>
#include <cstring>
#include <string>
#include <vector>
#include <memory>
#ifdef _OPENMP
#include <omp.h>
#endif

using namespace PLMD;

#define REF_EXPORT extern "C" __attribute__((visibility("default")))

static thread_local std::string g_err;

REF_EXPORT const char* ref_last_error() { return g_err.c_str(); }

REF_EXPORT void* ref_switch_create(const char* def) {
  auto* sf = new SwitchingFunction();
  std::string err;
  try {
    sf->set(std::string(def), err);
  } catch (std::exception& e) {
    err = e.what();
  }
  g_err = err;
  if (!err.empty()) { delete sf; return nullptr; }
  return sf;
}

REF_EXPORT void* ref_switch_create_nnmm(int nn, int mm, double r0, double d0) {
  auto* sf = new SwitchingFunction();
  sf->set(nn, mm, r0, d0);
  return sf;
}

REF_EXPORT void ref_switch_destroy(void* h) { delete static_cast<SwitchingFunction*>(h); }

REF_EXPORT void ref_switch_eval_sqr(void* h, int n, const double* d2, double* val, double* dfunc) {
  auto* sf = static_cast<SwitchingFunction*>(h);
  for (int i = 0; i < n; i++) val[i] = sf->calculateSqr(d2[i], dfunc[i]);
}

REF_EXPORT void ref_switch_eval(void* h, int n, const double* d, double* val, double* dfunc) {
  auto* sf = static_cast<SwitchingFunction*>(h);
  for (int i = 0; i < n; i++) val[i] = sf->calculate(d[i], dfunc[i]);
}

REF_EXPORT double ref_switch_dmax(void* h) { return static_cast<SwitchingFunction*>(h)->get_dmax(); }

static Tensor box_from(const double* b9) {
  return Tensor(b9[0], b9[1], b9[2], b9[3], b9[4], b9[5], b9[6], b9[7], b9[8]);
}

REF_EXPORT void ref_pbc_distance(const double* b9, int n, const double* v1, const double* v2, double* out) {
  Pbc pbc;
  pbc.setBox(box_from(b9));
  for (int i = 0; i < n; i++) {
    Vector d = pbc.distance(Vector(v1[3 * i], v1[3 * i + 1], v1[3 * i + 2]),
                            Vector(v2[3 * i], v2[3 * i + 1], v2[3 * i + 2]));
    out[3 * i] = d[0]; out[3 * i + 1] = d[1]; out[3 * i + 2] = d[2];
  }
}

// nlmode: 0 = no neighbour list, 1 = NLIST (classic), 2 = NLISTCELLS.
// pos holds the requested atoms (GROUPA then GROUPB); group atoms are numbered
// by absidx.  Returns the number of pairs in the list, or -1 on error.
REF_EXPORT int64_t ref_coordination(void* sfh, const double* b9, int pbc_on, const double* pos,
                                    const int64_t* absidx, int nA, int nB, int dopair,
                                    int nlmode, double nl_cut,
                                    double* value, double* deriv, double* virial) {
  try {
    auto* sf = static_cast<SwitchingFunction*>(sfh);
    Pbc pbc;
    pbc.setBox(box_from(b9));
    Communicator comm;
    std::vector<AtomNumber> ga(nA), gb(nB);
    for (int i = 0; i < nA; i++) ga[i].setIndex(absidx[i]);
    for (int i = 0; i < nB; i++) gb[i].setIndex(absidx[nA + i]);
    std::unique_ptr<NeighborList> nl;
    bool serial = true;
    if (nB > 0) {
      if (nlmode) nl = std::make_unique<NeighborList>(ga, gb, serial, dopair, pbc_on, pbc, comm, nl_cut, 1, nlmode == 2);
      else nl = std::make_unique<NeighborList>(ga, gb, serial, dopair, pbc_on, pbc, comm);
    } else {
      if (nlmode) nl = std::make_unique<NeighborList>(ga, serial, pbc_on, pbc, comm, nl_cut, 1, nlmode == 2);
      else nl = std::make_unique<NeighborList>(ga, serial, pbc_on, pbc, comm);
    }
    const int nat = nA + nB;
    std::vector<Vector> x(nat);
    for (int i = 0; i < nat; i++) x[i] = Vector(pos[3 * i], pos[3 * i + 1], pos[3 * i + 2]);
    if (nlmode) nl->update(x);
    double ncoord = 0.0;
    Tensor vir;
    std::vector<Vector> der(nat);
    const unsigned nn = nl->size();
    for (unsigned k = 0; k < nn; ++k) {
      auto pr = nl->getClosePair(k);
      unsigned i0 = pr.first, i1 = pr.second;
      if (absidx[i0] == absidx[i1]) continue;
      Vector d = pbc_on ? pbc.distance(x[i0], x[i1]) : delta(x[i0], x[i1]);
      double dfunc = 0.;
      ncoord += sf->calculateSqr(d.modulo2(), dfunc);
      Vector dd(dfunc * d);
      Tensor vv(dd, d);
      der[i0] -= dd;
      der[i1] += dd;
      vir -= vv;
    }
    *value = ncoord;
    for (int i = 0; i < nat; i++) for (int k = 0; k < 3; k++) deriv[3 * i + k] = der[i][k];
    for (int a = 0; a < 3; a++) for (int b = 0; b < 3; b++) virial[3 * a + b] = vir[a][b];
    return nn;
  } catch (std::exception& e) {
    g_err = e.what();
    return -1;
  }
}

// ---- Grid (src/tools/Grid.h) with spline and derivatives -------------------
REF_EXPORT void* ref_grid_create(int dim, const double* gmin, const double* gmax, const unsigned* nbin, const int* pbc) {
  std::vector<std::string> names, smin, smax, pmin, pmax;
  std::vector<unsigned> nb;
  std::vector<bool> isp;
  for (int i = 0; i < dim; i++) {
    names.push_back("x" + std::to_string(i));
    char b1[64], b2[64];
    std::snprintf(b1, sizeof b1, "%.17g", gmin[i]);
    std::snprintf(b2, sizeof b2, "%.17g", gmax[i]);
    smin.push_back(b1); smax.push_back(b2);
    pmin.push_back(b1); pmax.push_back(b2);
    nb.push_back(nbin[i]);
    isp.push_back(pbc[i] != 0);
  }
  return new Grid("bias", names, smin, smax, nb, true, true, isp, pmin, pmax);
}
REF_EXPORT void ref_grid_destroy(void* h) { delete static_cast<Grid*>(h); }
REF_EXPORT uint64_t ref_grid_size(void* h) { return static_cast<Grid*>(h)->getSize(); }
REF_EXPORT void ref_grid_set(void* h, uint64_t index, double v, const double* der) {
  auto* g = static_cast<Grid*>(h);
  std::vector<double> d(der, der + g->getDimension());
  double old;
  std::vector<double> oldd(g->getDimension());
  old = g->getValueAndDerivatives(index, oldd);
  for (unsigned i = 0; i < d.size(); i++) d[i] -= oldd[i];
  g->addValueAndDerivatives(index, v - old, d);
}
REF_EXPORT int ref_grid_spline(void* h, const double* x, double* value, double* der) {
  try {
    auto* g = static_cast<Grid*>(h);
    std::vector<double> xx(x, x + g->getDimension()), dd(g->getDimension());
    *value = g->getValueAndDerivatives(xx, dd);
    for (unsigned i = 0; i < dd.size(); i++) der[i] = dd[i];
    return 0;
  } catch (std::exception& e) {
    g_err = e.what();
    return -1;
  }
}
// neighbours of the node containing x with nneigh bins per direction
// (GridBase::getNeighbors); returns the count, writes at most cap indices
REF_EXPORT int64_t ref_grid_neighbors(void* h, const double* x, const unsigned* nneigh, uint64_t* out, int64_t cap) {
  try {
    auto* g = static_cast<Grid*>(h);
    std::vector<double> xx(x, x + g->getDimension());
    std::vector<unsigned> nn(nneigh, nneigh + g->getDimension());
    auto v = g->getNeighbors(xx, nn);
    for (int64_t i = 0; i < (int64_t)v.size() && i < cap; i++) out[i] = v[i];
    return (int64_t)v.size();
  } catch (std::exception& e) {
    g_err = e.what();
    return -1;
  }
}
REF_EXPORT void ref_grid_point(void* h, uint64_t index, double* x) {
  auto* g = static_cast<Grid*>(h);
  std::vector<double> xx(g->getDimension());
  g->getPoint(index, xx);
  for (unsigned i = 0; i < xx.size(); i++) x[i] = xx[i];
}

// ---- timing harness for the CPU baseline ------------------------------------
// A persistent reference NeighborList (cells or classic) so that the rebuild
// (NeighborList::update) and the per-step pair loop can be timed separately;
// the loop follows CoordinationBase::calculate (:162-217) with per-thread
// private derivatives merged in a critical section.
struct RefCoordCtx {
  SwitchingFunction* sf;
  Pbc pbc;
  Communicator comm;
  std::vector<AtomNumber> ga, gb;
  std::unique_ptr<NeighborList> nl;
  std::vector<int64_t> absidx;
  int nat;
};

REF_EXPORT void* ref_coordctx_create(void* sfh, const double* b9, const int64_t* absidx, int nA, int nB,
                                     int nlmode, double nl_cut) {
  try {
    auto* c = new RefCoordCtx();
    c->sf = static_cast<SwitchingFunction*>(sfh);
    c->pbc.setBox(box_from(b9));
    c->ga.resize(nA);
    c->gb.resize(nB);
    for (int i = 0; i < nA; i++) c->ga[i].setIndex(absidx[i]);
    for (int i = 0; i < nB; i++) c->gb[i].setIndex(absidx[nA + i]);
    c->absidx.assign(absidx, absidx + nA + nB);
    c->nat = nA + nB;
    if (nB > 0) c->nl = std::make_unique<NeighborList>(c->ga, c->gb, false, false, true, c->pbc, c->comm, nl_cut, 10, nlmode == 2);
    else c->nl = std::make_unique<NeighborList>(c->ga, false, true, c->pbc, c->comm, nl_cut, 10, nlmode == 2);
    return c;
  } catch (std::exception& e) {
    g_err = e.what();
    return nullptr;
  }
}

REF_EXPORT void ref_coordctx_destroy(void* h) { delete static_cast<RefCoordCtx*>(h); }

REF_EXPORT int64_t ref_coordctx_update(void* h, const double* pos) {
  auto* c = static_cast<RefCoordCtx*>(h);
  std::vector<Vector> x(c->nat);
  for (int i = 0; i < c->nat; i++) x[i] = Vector(pos[3 * i], pos[3 * i + 1], pos[3 * i + 2]);
  c->nl->update(x);
  return c->nl->size();
}

// NeighborList::getReducedAtomList (NeighborList.cpp:330-355): the remap of
// the pair indices onto the sorted list of requested atoms that
// CoordinationBase::prepare requests after each rebuild; returns the number
// of requested atoms
REF_EXPORT int64_t ref_coordctx_reduce(void* h) {
  auto* c = static_cast<RefCoordCtx*>(h);
  return (int64_t)c->nl->getReducedAtomList().size();
}

REF_EXPORT int64_t ref_coordctx_calc(void* h, const double* pos, int nthreads, double* value, double* deriv,
                                     double* virial) {
  auto* c = static_cast<RefCoordCtx*>(h);
  const int nat = c->nat;
  std::vector<Vector> x(nat);
  for (int i = 0; i < nat; i++) x[i] = Vector(pos[3 * i], pos[3 * i + 1], pos[3 * i + 2]);
  double ncoord = 0.;
  Tensor vir;
  std::vector<Vector> der(nat);
  const unsigned nn = c->nl->size();
  int64_t ncontrib = 0;
  const double dmax2 = c->sf->get_dmax2();
#pragma omp parallel num_threads(nthreads)
  {
    std::vector<Vector> od(nat);
    Tensor ov;
    int64_t oc = 0;
#pragma omp for reduction(+ : ncoord) nowait
    for (unsigned k = 0; k < nn; ++k) {
      auto pr = c->nl->getClosePair(k);
      unsigned i0 = pr.first, i1 = pr.second;
      if (c->absidx[i0] == c->absidx[i1]) continue;
      Vector d = c->pbc.distance(x[i0], x[i1]);
      double dfunc = 0.;
      double m2 = d.modulo2();
      ncoord += c->sf->calculateSqr(m2, dfunc);
      if (m2 <= dmax2) oc++;
      Vector dd(dfunc * d);
      Tensor vv(dd, d);
      od[i0] -= dd;
      od[i1] += dd;
      ov -= vv;
    }
#pragma omp critical
    {
      for (int i = 0; i < nat; i++) der[i] += od[i];
      vir += ov;
      ncontrib += oc;
    }
  }
  *value = ncoord;
  if (deriv)
    for (int i = 0; i < nat; i++) for (int k = 0; k < 3; k++) deriv[3 * i + k] = der[i][k];
  if (virial)
    for (int a = 0; a < 3; a++) for (int b = 0; b < 3; b++) virial[3 * a + b] = vir[a][b];
  return ncontrib;
}
