// Host preparation of the periodic cell for the FP64 all-pairs kernel:
// Pbc::setBox (src/tools/Pbc.cpp:165-212), LatticeReduction::reduceFast
// (src/tools/LatticeReduction.cpp:146-190) and Pbc::buildShifts
// (src/tools/Pbc.cpp:59-135).  Runs once per box change.
#include <cmath>
#include <cstring>
#include <utility>

#include "p2a_internal.h"

namespace p2a {

namespace {
struct V3 {
  double x[3];
};
inline double dot(const V3& a, const V3& b) { return a.x[0] * b.x[0] + a.x[1] * b.x[1] + a.x[2] * b.x[2]; }
inline double mod2(const V3& a) { return dot(a, a); }

double det3(const double* m) {
  return m[0] * m[4] * m[8] + m[1] * m[5] * m[6] + m[2] * m[3] * m[7] - m[0] * m[5] * m[7] - m[1] * m[3] * m[8] -
         m[2] * m[4] * m[6];
}

void inv3(const double* m, double* out) {
  double id = 1.0 / det3(m);
  for (int i = 0; i < 3; i++)
    for (int j = 0; j < 3; j++) {
      int i1 = (j + 1) % 3, i2 = (j + 2) % 3, j1 = (i + 1) % 3, j2 = (i + 2) % 3;
      out[3 * i + j] = (m[3 * i1 + j1] * m[3 * i2 + j2] - m[3 * i1 + j2] * m[3 * i2 + j1]) * id;
    }
}

void lr_sort(V3 v[3]) {
  double m[3] = {mod2(v[0]), mod2(v[1]), mod2(v[2])};
  for (int i = 0; i < 3; i++)
    for (int j = i + 1; j < 3; j++)
      if (m[i] > m[j]) {
        std::swap(v[i], v[j]);
        std::swap(m[i], m[j]);
      }
}

// Gauss reduction of a pair (LatticeReduction::reduce(a,b))
void lr_reduce2(V3& a, V3& b) {
  const double onePlusEpsilon = 1.0 + 1e-14;
  double ma = mod2(a), mb = mod2(b);
  for (unsigned it = 0; it < 1000000u; it++) {
    if (mb > ma) {
      std::swap(a, b);
      std::swap(ma, mb);
    }
    double f = std::floor(dot(a, b) / mb + 0.5);
    for (int k = 0; k < 3; k++) a.x[k] -= b.x[k] * f;
    ma = mod2(a);
    if (mb <= ma * onePlusEpsilon) break;
  }
  std::swap(a, b);
}

void lr_reduce_fast(double* t) {
  const double onePlusEpsilon = 1.0 + 1e-14;
  V3 v[3];
  for (int i = 0; i < 3; i++)
    for (int k = 0; k < 3; k++) v[i].x[k] = t[3 * i + k];
  for (unsigned it = 0; it < 1000000u; it++) {
    lr_sort(v);
    lr_reduce2(v[0], v[1]);
    double b11 = mod2(v[0]), b22 = mod2(v[1]), b12 = dot(v[0], v[1]), b13 = dot(v[0], v[2]), b23 = dot(v[1], v[2]);
    double z = b11 * b22 - b12 * b12;
    double y2 = -(b11 * b23 - b12 * b13) / z;
    double y1 = -(b22 * b13 - b12 * b23) / z;
    int x1min = (int)std::floor(y1), x1max = x1min + 1, x2min = (int)std::floor(y2), x2max = x2min + 1;
    bool first = true;
    double mbest = 0;
    V3 best = {{0, 0, 0}};
    for (int x1 = x1min; x1 <= x1max; x1++)
      for (int x2 = x2min; x2 <= x2max; x2++) {
        V3 trial;
        for (int k = 0; k < 3; k++) trial.x[k] = v[2].x[k] + x2 * v[1].x[k] + x1 * v[0].x[k];
        double mt = mod2(trial);
        if (first || mt < mbest) {
          mbest = mt;
          best = trial;
          first = false;
        }
      }
    if (mod2(best) * onePlusEpsilon >= mod2(v[2])) break;
    v[2] = best;
  }
  lr_sort(v);
  for (int i = 0; i < 3; i++)
    for (int k = 0; k < 3; k++) t[3 * i + k] = v[i].x[k];
}

void build_shifts(PbcHost& p) {
  const double small = 1e-28;
  const double* R = p.reduced;
  for (int o = 0; o < 8; o++) p.nshift[o] = 0;
  for (int l = -1; l <= 1; l++)
    for (int m = -1; m <= 1; m++)
      for (int n = -1; n <= 1; n++) {
        int ish[3] = {l, m, n};
        double ds[3] = {(double)l, (double)m, (double)n};
        int count = (l != 0) + (m != 0) + (n != 0);
        if (count == 0 || count == 3) continue;
        double tmp[3], cosdir[3];
        for (int i = 0; i < 3; i++) tmp[i] = R[0 * 3 + i] * ds[0] + R[1 * 3 + i] * ds[1] + R[2 * 3 + i] * ds[2];
        for (int i = 0; i < 3; i++) cosdir[i] = R[i * 3 + 0] * tmp[0] + R[i * 3 + 1] * tmp[1] + R[i * 3 + 2] * tmp[2];
        double dp = ds[0] * cosdir[0] + ds[1] * cosdir[1] + ds[2] * cosdir[2];
        double ref = (ds[0] * ds[0] + ds[1] * ds[1] + ds[2] * ds[2]) *
                     (cosdir[0] * cosdir[0] + cosdir[1] * cosdir[1] + cosdir[2] * cosdir[2]);
        if (std::fabs(ref - dp * dp) < small) continue;
        for (int i = 0; i < 2; i++)
          for (int j = 0; j < 2; j++)
            for (int k = 0; k < 2; k++) {
              int block[3] = {2 * i - 1, 2 * j - 1, 2 * k - 1};
              bool skip = false;
              for (int s = 0; s < 3; s++)
                if (ish[s] * block[s] > 0) skip = true;
              if (skip) continue;
              skip = true;
              for (int s = 0; s < 3; s++)
                if (((1 - ish[s] * ish[s]) * block[s]) * cosdir[s] < -small) skip = false;
              if (skip) continue;
              int oct = i * 4 + j * 2 + k;
              int ns = p.nshift[oct];
              if (ns >= kMaxShift) continue;
              for (int c = 0; c < 3; c++) p.shifts[oct][ns][c] = R[0 * 3 + c] * ds[0] + R[1 * 3 + c] * ds[1] + R[2 * 3 + c] * ds[2];
              p.nshift[oct] = ns + 1;
            }
      }
}
}  // namespace

void pbc_setup(const double* b9, PbcHost& p) {
  p = PbcHost();
  std::memcpy(p.box, b9, sizeof(p.box));
  const double eps = 1e-28;
  double d = det3(p.box);
  if (d * d < eps) {
    p.type = 0;
    return;
  }
  const double* B = p.box;
  bool cxy = B[1] * B[1] < eps && B[3] * B[3] < eps;
  bool cxz = B[2] * B[2] < eps && B[6] * B[6] < eps;
  bool cyz = B[5] * B[5] < eps && B[7] * B[7] < eps;
  inv3(p.box, p.inv);
  std::memcpy(p.reduced, p.box, sizeof(p.box));
  if (cxy && cxz && cyz) {
    p.type = 1;
    inv3(p.reduced, p.invReduced);
  } else {
    p.type = 2;
    lr_reduce_fast(p.reduced);
    inv3(p.reduced, p.invReduced);
    build_shifts(p);
  }
}

void pbc_widths(const double* b9, double w[3]) {
  double inv[9];
  inv3(b9, inv);
  for (int k = 0; k < 3; k++) {
    // reciprocal row k = column k of the inverse (LinkCells.cpp:99-122)
    double r = std::sqrt(inv[0 * 3 + k] * inv[0 * 3 + k] + inv[1 * 3 + k] * inv[1 * 3 + k] + inv[2 * 3 + k] * inv[2 * 3 + k]);
    w[k] = 1.0 / r;
  }
}

}  // namespace p2a
