/*
 * FP64 CPU oracle for the COORDINATION + METAD hot path.
 *
 * TEST INFRASTRUCTURE ONLY.  This file restates, in plain C, the reference
 * algorithms that the HIP library replaces, so that the tests, the smoke check
 * and the CPU-baseline leg of the benchmark have an independent checker.  No
 * product code path links or calls it.
 *
 * Every function cites the reference source it follows (paths relative to the
 * plumed2 source tree).  What pins each part (compiled reference subset,
 * regtest golden files, closed forms) is listed in the test files.
 */
#define _GNU_SOURCE
#include <math.h>
#include <float.h>
#include <stdint.h>
#include <stdlib.h>
#include <string.h>

#define OR_EXPORT __attribute__((visibility("default")))

/* ------------------------------------------------------------------------ */
/* Switching functions: src/tools/SwitchingFunction.{h,cpp}                  */
/* ------------------------------------------------------------------------ */

enum {
  OSW_RATIONAL_FIXED = 0, /* fixedRational<N>            SwitchingFunction.cpp:185-216 */
  OSW_RATIONAL = 1,       /* rational<fast?,simplified?> SwitchingFunction.cpp:220-304 */
  OSW_EXP = 2,            /* exponentialSwitch           :375-387 */
  OSW_GAUSSIAN = 3,       /* gaussianSwitch              :389-400 */
  OSW_FASTGAUSSIAN = 4,   /* fastgaussianSwitch          :402-431 */
  OSW_SMAP = 5,           /* smapSwitch                  :433-455 */
  OSW_CUBIC = 6,          /* cubicSwitch                 :457-469 */
  OSW_TANH = 7,           /* tanhSwitch                  :471-486 */
  OSW_COSINUS = 8,        /* cosinusSwitch               :488-507 */
  OSW_Q = 9               /* nativeqSwitch               :509-548 */
};

typedef struct {
  int type;
  int fast;        /* rational: N,M even and d0==0 -> works on r^2      */
  int simplified;  /* rational: 2N==M                                    */
  int nn, mm;      /* rational powers (fixed: nn = N)                    */
  int nnf, mmf;    /* N/2, M/2                                            */
  int a, b;        /* smap                                               */
  double d0, dmax, dmax_2, invr0, invr0_2, stretch, shift;
  double preRes, preDfunc, preSecDev, preDfuncF, preSecDevF;
  double c, d;     /* smap                                               */
  double beta, lambda, ref; /* Q                                         */
} osw_t;

/* Tools::fastpow (src/tools/Tools.h:582-597), square-and-multiply */
static double o_fastpow(double base, int exp) {
  if (exp < 0) { exp = -exp; base = 1.0 / base; }
  double result = 1.0;
  while (exp) {
    if (exp & 1) result *= base;
    exp >>= 1;
    base *= base;
  }
  return result;
}

/* Data::init (SwitchingFunction.cpp:84-94) */
static void o_data_init(osw_t *s, double d0, double dmax, double r0) {
  s->d0 = d0;
  s->dmax = dmax;
  s->dmax_2 = (dmax < sqrt(DBL_MAX)) ? dmax * dmax : DBL_MAX;
  s->invr0 = 1.0 / r0;
  s->invr0_2 = s->invr0 * s->invr0;
  s->stretch = 1.0;
  s->shift = 0.0;
}

/* rational<...>::doRational (SwitchingFunction.cpp:239-281) */
static void o_do_rational(const osw_t *s, double rdist, double secDev, int N, int M,
                          double dfunc0, double result0, double *res, double *df) {
  const double moreThanOne = 1.0 + 5.0e10 * DBL_EPSILON;
  const double lessThanOne = 1.0 - 5.0e10 * DBL_EPSILON;
  if (s->simplified) {
    double rN = o_fastpow(rdist, N - 1);
    double r = 1.0 / (1.0 + rN * rdist);
    *res = r;
    *df = -N * rN * r * r;
  } else if (!((rdist > lessThanOne) && (rdist < moreThanOne))) {
    double rN = o_fastpow(rdist, N - 1);
    double rM = o_fastpow(rdist, M - 1);
    double num = 1.0 - rN * rdist;
    double iden = 1.0 / (1.0 - rM * rdist);
    double r = num * iden;
    *res = r;
    *df = ((M * r * rM) - (N * rN)) * iden;
  } else {
    double x = rdist - 1.0;
    *res = result0 + x * (dfunc0 + 0.5 * x * secDev);
    *df = dfunc0 + x * secDev;
  }
}

/* the per-type "function(rdist)" of each switch container (value, d/d rdist) */
static void o_function(const osw_t *s, double rdist, double *f, double *df) {
  switch (s->type) {
  case OSW_RATIONAL_FIXED: { /* doRational<N>(rdist) :192-197 */
    double rN = o_fastpow(rdist, s->nn - 1);
    double r = 1.0 / (1.0 + rN * rdist);
    *f = r; *df = -s->nn * rN * r * r; return; }
  case OSW_RATIONAL:
    o_do_rational(s, rdist, s->preSecDev, s->nn, s->mm, s->preDfunc, s->preRes, f, df);
    return;
  case OSW_EXP: { double r = exp(-rdist); *f = r; *df = -r; return; }
  case OSW_GAUSSIAN:
  case OSW_FASTGAUSSIAN: { double r = exp(-0.5 * rdist * rdist); *f = r; *df = -rdist * r; return; }
  case OSW_SMAP: {
    double sx = s->c * o_fastpow(rdist, s->a);
    double r = pow(1.0 + sx, s->d);
    *f = r; *df = -s->b * sx / rdist * r / (1.0 + sx); return; }
  case OSW_CUBIC: {
    double t1 = rdist - 1.0, t2 = 1.0 + 2.0 * rdist;
    *f = t1 * t1 * t2; *df = 2 * t1 * t2 + 2 * t1 * t1; return; }
  case OSW_TANH: { double t = tanh(rdist); *f = 1.0 - t; *df = t * t - 1.0; return; }
  case OSW_COSINUS: {
    *f = 0.0; *df = 0.0;
    if (rdist <= 1.0) {
      double rp = rdist * M_PI;
      *f = 0.5 * (cos(rp) + 1.0); *df = -0.5 * M_PI * sin(rp);
    }
    return; }
  default: *f = 0.0; *df = 0.0; return;
  }
}

/* baseSwitch::calculate (SwitchingFunction.cpp:133-184) + applystretch (:124-130);
 * nativeqSwitch overrides calculate (:521-548). dfunc = (1/r) ds/dr. */
OR_EXPORT double osw_calculate(const osw_t *s, double distance, double *dfunc) {
  if (s->type == OSW_Q) {
    double res = 0.0, df = 0.0;
    if (distance <= s->dmax) {
      res = 1.0;
      if (distance > s->d0) {
        double rdist = s->beta * (distance - s->lambda * s->ref);
        double e = exp(rdist);
        res = 1.0 / (1.0 + e);
        df = -s->beta / (e + 2.0 + 1.0 / e) / distance;
        df *= s->stretch;
      }
      res = res * s->stretch + s->shift;
    }
    *dfunc = df;
    return res;
  }
  double rdist = (distance - s->d0) * s->invr0;
  if (distance > s->dmax) { *dfunc = 0.0; return 0.0; }
  if (!(rdist > 0.0)) { *dfunc = 0.0; return s->stretch + s->shift; }
  double f, df;
  o_function(s, rdist, &f, &df);
  f = f * s->stretch + s->shift;
  df *= s->stretch;
  df *= s->invr0;
  df /= distance;
  *dfunc = df;
  return f;
}

/* calculateSqr of each container: fixedRational :203-215, fast rational :283-298,
 * fastgaussian :411-431, everything else sqrt -> calculate (:180-183). */
OR_EXPORT double osw_calculate_sqr(const osw_t *s, double d2, double *dfunc) {
  if (s->type == OSW_RATIONAL_FIXED) {
    double result = 0.0, df = 0.0;
    if (d2 <= s->dmax_2) {
      double rdist = d2 * s->invr0_2;
      int P = s->nn / 2;
      double rN = o_fastpow(rdist, P - 1);
      result = 1.0 / (1.0 + rN * rdist);
      df = -P * rN * result * result;
      df *= 2 * s->invr0_2;
      result = result * s->stretch + s->shift;
      df *= s->stretch;
    }
    *dfunc = df;
    return result;
  }
  if (s->type == OSW_RATIONAL && s->fast) {
    double result = 0.0, df = 0.0;
    if (d2 <= s->dmax_2) {
      double rdist = d2 * s->invr0_2;
      o_do_rational(s, rdist, s->preSecDevF, s->nnf, s->mmf, s->preDfuncF, s->preRes, &result, &df);
      df *= 2 * s->invr0_2;
      result = result * s->stretch + s->shift;
      df *= s->stretch;
    }
    *dfunc = df;
    return result;
  }
  if (s->type == OSW_FASTGAUSSIAN) {
    double result = 0.0, df = 0.0;
    if (d2 < s->dmax_2) {
      result = 1.0;
      if (d2 > 0.0) {
        result = exp(-0.5 * d2);
        df = -result;
        result = result * s->stretch + s->shift;
        df *= s->stretch;
      }
    }
    *dfunc = df;
    return result;
  }
  return osw_calculate(s, sqrt(d2), dfunc);
}

/* SwitchInterface::setupStretch (SwitchingFunction.cpp:63-72): uses the
 * non-squared calculate at 0 and at dmax */
static void o_setup_stretch(osw_t *s) {
  if (s->dmax < DBL_MAX) {
    double dd;
    s->stretch = 1.0;
    s->shift = 0.0;
    double s0 = osw_calculate(s, 0.0, &dd);
    double sd = osw_calculate(s, s->dmax, &dd);
    s->stretch = 1.0 / (s0 - sd);
    s->shift = -sd * s->stretch;
  }
}

/* rationalFactory (SwitchingFunction.cpp:307-347) + rational::init (:226-251) */
static void o_rational_init(osw_t *s, double d0, double dmax, double r0, int N, int M) {
  memset(s, 0, sizeof(*s));
  o_data_init(s, d0, dmax, r0);
  int fast = (N % 2 == 0) && (M % 2 == 0) && d0 == 0.0;
  if (((2 * N) == M || M == 0) && fast && N <= 12 && N > 0) {
    s->type = OSW_RATIONAL_FIXED;
    s->nn = N;
    s->mm = 2 * N;
    return;
  }
  s->type = OSW_RATIONAL;
  s->nn = N;
  s->mm = (M == 0) ? N * 2 : M;
  s->preRes = (double)s->nn / s->mm;
  s->preDfunc = 0.5 * s->nn * (s->nn - s->mm) / (double)s->mm;
  s->preSecDev = (s->nn * (s->mm * s->mm - 3.0 * s->mm * (-1 + s->nn) + s->nn * (-3 + 2 * s->nn))) / (6.0 * s->mm);
  s->nnf = s->nn / 2;
  s->mmf = s->mm / 2;
  s->preDfuncF = 0.5 * s->nnf * (s->nnf - s->mmf) / (double)s->mmf;
  s->preSecDevF = (s->nnf * (s->mmf * s->mmf - 3.0 * s->mmf * (-1 + s->nnf) + s->nnf * (-3 + 2 * s->nnf))) / (6.0 * s->mmf);
  s->fast = fast;
  s->simplified = (2 * N == M || M == 0);
}

/*
 * One constructor for every switch type, mirroring SwitchingFunction::set(string)
 * (SwitchingFunction.cpp:1055-1159) after keyword parsing: the caller passes the
 * parsed numbers.  dmax<0 means "D_MAX not given".  stretch is applied when
 * dmax is given and NOSTRETCH is absent (:1156-1158).
 */
OR_EXPORT int osw_init(osw_t *s, int type, double d0, double dmax, double r0, int nn, int mm,
                       int a, int b, double beta, double lambda, double ref, int nostretch) {
  if (dmax < 0) dmax = DBL_MAX;
  memset(s, 0, sizeof(*s));
  switch (type) {
  case OSW_RATIONAL_FIXED:
  case OSW_RATIONAL:
    o_rational_init(s, d0, dmax, r0, nn, mm);
    break;
  case OSW_EXP: case OSW_TANH: case OSW_COSINUS:
    o_data_init(s, d0, dmax, r0); s->type = type; break;
  case OSW_GAUSSIAN:
  case OSW_FASTGAUSSIAN:
    if (r0 == 1.0 && d0 == 0.0) { o_data_init(s, 0.0, dmax, 1.0); s->type = OSW_FASTGAUSSIAN; }
    else { o_data_init(s, d0, dmax, r0); s->type = OSW_GAUSSIAN; }
    break;
  case OSW_SMAP:
    o_data_init(s, d0, dmax, r0); s->type = OSW_SMAP;
    s->a = a; s->b = b;
    s->c = pow(2., (double)a / (double)b) - 1.0;
    s->d = -(double)b / (double)a;
    break;
  case OSW_CUBIC: /* cubic uses d0 and dmax only: r0 = dmax-d0 (:459-462) */
    o_data_init(s, d0, dmax, dmax - d0); s->type = OSW_CUBIC; break;
  case OSW_Q:
    o_data_init(s, d0, dmax, r0); s->type = OSW_Q;
    s->beta = beta; s->lambda = lambda; s->ref = ref; break;
  default: return -1;
  }
  if (!nostretch && dmax != DBL_MAX) o_setup_stretch(s);
  return 0;
}

/* SwitchingFunction::set(nn,mm,r0,d0) (SwitchingFunction.cpp:1176-1184):
 * automatic dmax where s = 1e-5, always stretched */
OR_EXPORT int osw_init_nnmm(osw_t *s, int nn, int mm, double r0, double d0) {
  if (mm == 0) mm = 2 * nn;
  double dmax = d0 + r0 * pow(0.00001, 1. / (nn - mm));
  o_rational_init(s, d0, dmax, r0, nn, mm);
  o_setup_stretch(s);
  return 0;
}

OR_EXPORT size_t osw_sizeof(void) { return sizeof(osw_t); }
OR_EXPORT double osw_get(const osw_t *s, int what) {
  switch (what) {
  case 0: return s->dmax; case 1: return s->stretch; case 2: return s->shift;
  case 3: return 1.0 / s->invr0; case 4: return s->d0; case 5: return (double)s->type;
  default: return NAN;
  }
}

/* ------------------------------------------------------------------------ */
/* Pbc: src/tools/Pbc.cpp, src/tools/LatticeReduction.cpp, Tools::pbc        */
/* ------------------------------------------------------------------------ */

#define OPBC_MAXSHIFT 32
typedef struct {
  int type; /* 0 unset, 1 orthorhombic, 2 generic */
  double box[3][3], invBox[3][3], reduced[3][3], invReduced[3][3];
  int nshift[2][2][2];
  double shifts[2][2][2][OPBC_MAXSHIFT][3];
} opbc_t;

/* Tools::pbc, offset-100 branch (src/tools/Tools.h:547-570) */
static double o_pbc1(double x) {
  const double offset = 100.0;
  x += offset;
  if (x >= 0) return x - (int)(x + 0.5);
  return x - (int)(x - 0.5);
}

static double o_det(const double m[3][3]) {
  return m[0][0] * m[1][1] * m[2][2] + m[0][1] * m[1][2] * m[2][0] + m[0][2] * m[1][0] * m[2][1]
       - m[0][0] * m[1][2] * m[2][1] - m[0][1] * m[1][0] * m[2][2] - m[0][2] * m[1][1] * m[2][0];
}

/* TensorGeneric<3,3>::inverse (src/tools/Tensor.h): cofactor / determinant */
static void o_inverse(const double m[3][3], double inv[3][3]) {
  double det = o_det(m);
  double id = 1.0 / det;
  for (int i = 0; i < 3; i++)
    for (int j = 0; j < 3; j++) {
      int i1 = (j + 1) % 3, i2 = (j + 2) % 3, j1 = (i + 1) % 3, j2 = (i + 2) % 3;
      inv[i][j] = (m[i1][j1] * m[i2][j2] - m[i1][j2] * m[i2][j1]) * id;
    }
}

static double o_mod2(const double v[3]) { return v[0] * v[0] + v[1] * v[1] + v[2] * v[2]; }
static double o_dot(const double a[3], const double b[3]) { return a[0] * b[0] + a[1] * b[1] + a[2] * b[2]; }

/* LatticeReduction::sort (LatticeReduction.cpp:31-58) */
static void o_lr_sort(double v[3][3]) {
  double m[3] = {o_mod2(v[0]), o_mod2(v[1]), o_mod2(v[2])};
  for (int i = 0; i < 3; i++)
    for (int j = i + 1; j < 3; j++)
      if (m[i] > m[j]) {
        double t[3]; memcpy(t, v[i], sizeof t); memcpy(v[i], v[j], sizeof t); memcpy(v[j], t, sizeof t);
        double tm = m[i]; m[i] = m[j]; m[j] = tm;
      }
}

/* LatticeReduction::reduce(a,b) (LatticeReduction.cpp:60-84) */
static void o_lr_reduce2(double a[3], double b[3]) {
  const double onePlusEpsilon = 1.0 + 1e-14;
  double ma = o_mod2(a), mb = o_mod2(b);
  for (unsigned counter = 0; counter < 1000000u; counter++) {
    if (mb > ma) {
      double t[3]; memcpy(t, a, sizeof t); memcpy(a, b, sizeof t); memcpy(b, t, sizeof t);
      double tm = ma; ma = mb; mb = tm;
    }
    double f = floor(o_dot(a, b) / mb + 0.5);
    for (int k = 0; k < 3; k++) a[k] -= b[k] * f;
    ma = o_mod2(a);
    if (mb <= ma * onePlusEpsilon) break;
  }
  double t[3]; memcpy(t, a, sizeof t); memcpy(a, b, sizeof t); memcpy(b, t, sizeof t);
}

/* LatticeReduction::reduceFast (LatticeReduction.cpp:146-190) */
static void o_lr_reduce_fast(double t[3][3]) {
  const double onePlusEpsilon = 1.0 + 1e-14;
  double v[3][3];
  memcpy(v, t, sizeof v);
  for (unsigned counter = 0; counter < 1000000u; counter++) {
    o_lr_sort(v);
    o_lr_reduce2(v[0], v[1]);
    double b11 = o_mod2(v[0]), b22 = o_mod2(v[1]), b12 = o_dot(v[0], v[1]);
    double b13 = o_dot(v[0], v[2]), b23 = o_dot(v[1], v[2]);
    double z = b11 * b22 - b12 * b12;
    double y2 = -(b11 * b23 - b12 * b13) / z;
    double y1 = -(b22 * b13 - b12 * b23) / z;
    int x1min = (int)floor(y1), x1max = x1min + 1, x2min = (int)floor(y2), x2max = x2min + 1;
    int first = 1;
    double mbest = 0, best[3] = {0, 0, 0};
    for (int x1 = x1min; x1 <= x1max; x1++)
      for (int x2 = x2min; x2 <= x2max; x2++) {
        double trial[3];
        for (int k = 0; k < 3; k++) trial[k] = v[2][k] + x2 * v[1][k] + x1 * v[0][k];
        double mt = o_mod2(trial);
        if (first || mt < mbest) { mbest = mt; memcpy(best, trial, sizeof best); first = 0; }
      }
    if (o_mod2(best) * onePlusEpsilon >= o_mod2(v[2])) break;
    memcpy(v[2], best, sizeof best);
  }
  o_lr_sort(v);
  memcpy(t, v, sizeof v);
}

/* Pbc::buildShifts (Pbc.cpp:59-135) */
static void o_build_shifts(opbc_t *p) {
  const double small = 1e-28;
  memset(p->nshift, 0, sizeof p->nshift);
  for (int l = -1; l <= 1; l++)
    for (int m = -1; m <= 1; m++)
      for (int n = -1; n <= 1; n++) {
        int ishift[3] = {l, m, n};
        double dshift[3] = {l, m, n};
        int count = (l != 0) + (m != 0) + (n != 0);
        if (count == 0 || count == 3) continue;
        /* cosdir = reduced * reduced^T * dshift */
        double tmp[3], cosdir[3];
        for (int i = 0; i < 3; i++) tmp[i] = p->reduced[0][i] * dshift[0] + p->reduced[1][i] * dshift[1] + p->reduced[2][i] * dshift[2];
        for (int i = 0; i < 3; i++) cosdir[i] = p->reduced[i][0] * tmp[0] + p->reduced[i][1] * tmp[1] + p->reduced[i][2] * tmp[2];
        double dp = o_dot(dshift, cosdir);
        double ref = o_mod2(dshift) * o_mod2(cosdir);
        if (fabs(ref - dp * dp) < small) continue;
        for (int i = 0; i < 2; i++)
          for (int j = 0; j < 2; j++)
            for (int k = 0; k < 2; k++) {
              int block[3] = {2 * i - 1, 2 * j - 1, 2 * k - 1};
              int skip = 0;
              for (int s = 0; s < 3; s++) if (ishift[s] * block[s] > 0) skip = 1;
              if (skip) continue;
              skip = 1;
              for (int s = 0; s < 3; s++)
                if (((1 - ishift[s] * ishift[s]) * block[s]) * cosdir[s] < -small) skip = 0;
              if (skip) continue;
              int ns = p->nshift[i][j][k]++;
              /* transpose(reduced) * dshift */
              for (int c = 0; c < 3; c++)
                p->shifts[i][j][k][ns][c] = p->reduced[0][c] * dshift[0] + p->reduced[1][c] * dshift[1] + p->reduced[2][c] * dshift[2];
            }
      }
}

/* Pbc::setBox (Pbc.cpp:165-212); box rows are the lattice vectors */
OR_EXPORT void opbc_set_box(opbc_t *p, const double *b9) {
  memset(p, 0, sizeof(*p));
  for (int i = 0; i < 3; i++) for (int j = 0; j < 3; j++) p->box[i][j] = b9[3 * i + j];
  const double eps = 1e-28;
  double det = o_det(p->box);
  if (det * det < eps) { p->type = 0; return; }
  int cxy = p->box[0][1] * p->box[0][1] < eps && p->box[1][0] * p->box[1][0] < eps;
  int cxz = p->box[0][2] * p->box[0][2] < eps && p->box[2][0] * p->box[2][0] < eps;
  int cyz = p->box[1][2] * p->box[1][2] < eps && p->box[2][1] * p->box[2][1] < eps;
  o_inverse(p->box, p->invBox);
  if (cxy && cxz && cyz) {
    p->type = 1;
    memcpy(p->reduced, p->box, sizeof p->box);
    o_inverse(p->reduced, p->invReduced);
  } else {
    p->type = 2;
    memcpy(p->reduced, p->box, sizeof p->box);
    o_lr_reduce_fast(p->reduced);
    o_inverse(p->reduced, p->invReduced);
    o_build_shifts(p);
  }
}

/* Pbc::distance (Pbc.cpp:362-415): minimum-image v2-v1 */
OR_EXPORT void opbc_distance(const opbc_t *p, const double *v1, const double *v2, double *out) {
  double d[3] = {v2[0] - v1[0], v2[1] - v1[1], v2[2] - v1[2]};
  if (p->type == 1) {
    for (int i = 0; i < 3; i++) d[i] = o_pbc1(d[i] * p->invBox[i][i]) * p->box[i][i];
  } else if (p->type == 2) {
    double s[3];
    for (int j = 0; j < 3; j++) s[j] = d[0] * p->invReduced[0][j] + d[1] * p->invReduced[1][j] + d[2] * p->invReduced[2][j];
    for (int j = 0; j < 3; j++) s[j] = o_pbc1(s[j]);
    for (int j = 0; j < 3; j++) d[j] = s[0] * p->reduced[0][j] + s[1] * p->reduced[1][j] + s[2] * p->reduced[2][j];
    if (fabs(s[0]) + fabs(s[1]) + fabs(s[2]) > 0.5) {
      int i0 = s[0] > 0 ? 1 : 0, i1 = s[1] > 0 ? 1 : 0, i2 = s[2] > 0 ? 1 : 0;
      double best[3] = {d[0], d[1], d[2]};
      double lbest = o_mod2(best);
      for (int k = 0; k < p->nshift[i0][i1][i2]; k++) {
        const double *sh = p->shifts[i0][i1][i2][k];
        double t[3] = {d[0] + sh[0], d[1] + sh[1], d[2] + sh[2]};
        double lt = o_mod2(t);
        if (lt < lbest) { lbest = lt; memcpy(best, t, sizeof best); }
      }
      memcpy(d, best, sizeof best);
    }
  }
  out[0] = d[0]; out[1] = d[1]; out[2] = d[2];
}

OR_EXPORT size_t opbc_sizeof(void) { return sizeof(opbc_t); }
OR_EXPORT void opbc_get_reduced(const opbc_t *p, double *out9) {
  for (int i = 0; i < 3; i++) for (int j = 0; j < 3; j++) out9[3 * i + j] = p->reduced[i][j];
}
OR_EXPORT int opbc_type(const opbc_t *p) { return p->type; }

/* ------------------------------------------------------------------------ */
/* Neighbour lists: src/tools/NeighborList.cpp, src/tools/LinkCells.cpp      */
/* ------------------------------------------------------------------------ */

/* LinkCells::createCells (LinkCells.cpp:99-122): cells per lattice direction
 * from the reciprocal-row lengths (valid for triclinic boxes). */
OR_EXPORT void ocells_count(const opbc_t *p, double cutoff, int *nc) {
  for (int k = 0; k < 3; k++) {
    /* reciprocal = transpose(invBox): row k = column k of invBox */
    double r[3] = {p->invBox[0][k], p->invBox[1][k], p->invBox[2][k]};
    double len = sqrt(o_mod2(r));
    int n = (int)floor(1.0 / len / cutoff);
    nc[k] = n == 0 ? 1 : n;
  }
}

/* LinkCells::findMyCell (LinkCells.cpp:277-292) with pbc */
static int o_find_cell(const opbc_t *p, const int *nc, const double *x) {
  int c[3];
  for (int j = 0; j < 3; j++) {
    double f = x[0] * p->invBox[0][j] + x[1] * p->invBox[1][j] + x[2] * p->invBox[2][j];
    c[j] = (int)floor((o_pbc1(f) + 0.5) * nc[j]);
    if (c[j] >= nc[j]) c[j] = nc[j] - 1; /* reference asserts; cannot happen for finite input */
    if (c[j] < 0) c[j] = 0;
  }
  return c[0] + nc[0] * (c[1] + nc[1] * c[2]);
}

/* LinkCells::addRequiredCells (LinkCells.cpp:183-239), pbc on */
static int o_required_cells(const int *nc, int cell, int *out) {
  int cc[3] = {cell % nc[0], (cell / nc[0]) % nc[1], cell / (nc[0] * nc[1])};
  int lo[3], hi[3];
  for (int k = 0; k < 3; k++) {
    lo[k] = cc[k] + ((nc[k] < 2) ? 0 : -1);
    hi[k] = cc[k] + ((nc[k] < 3) ? 1 : 2);
  }
  int n = 0;
  for (int x = lo[0]; x < hi[0]; x++)
    for (int y = lo[1]; y < hi[1]; y++)
      for (int z = lo[2]; z < hi[2]; z++) {
        int xv = (x < 0) ? nc[0] - 1 : x % nc[0];
        int yv = (y < 0) ? nc[1] - 1 : y % nc[1];
        int zv = (z < 0) ? nc[2] - 1 : z % nc[2];
        int box = xv + nc[0] * (yv + nc[1] * zv);
        int seen = 0;
        for (int k = 0; k < n; k++) if (out[k] == box) { seen = 1; break; }
        if (!seen) out[n++] = box;
      }
  return n;
}

/*
 * NeighborList::update, cells branch (NeighborList.cpp:177-236): all pairs of
 * atoms in required cells, single list keeps B>A.  Positions are the full
 * requested list (A then B).  Writes at most cap pairs to pairs[2k],[2k+1];
 * returns the total count (which may exceed cap; call again with more room).
 */
OR_EXPORT int64_t onl_cells(const opbc_t *p, double cutoff, const double *pos, int nA, int nB,
                            int32_t *pairs, int64_t cap) {
  int nc[3];
  ocells_count(p, cutoff, nc);
  int ncell = nc[0] * nc[1] * nc[2];
  int natoms = nA + nB;
  int *cellof = (int *)malloc(sizeof(int) * (natoms > 0 ? natoms : 1));
  int two = nB > 0;
  for (int i = 0; i < natoms; i++) cellof[i] = o_find_cell(p, nc, pos + 3 * i);
  /* counting sort per list, in atom order (LinkCells::resetCollection :138-181) */
  int *startA = (int *)calloc(ncell + 1, sizeof(int)), *startB = (int *)calloc(ncell + 1, sizeof(int));
  int *listA = (int *)malloc(sizeof(int) * (nA > 0 ? nA : 1)), *listB = (int *)malloc(sizeof(int) * (nB > 0 ? nB : 1));
  for (int i = 0; i < nA; i++) startA[cellof[i] + 1]++;
  for (int i = 0; i < nB; i++) startB[cellof[nA + i] + 1]++;
  for (int c = 0; c < ncell; c++) { startA[c + 1] += startA[c]; startB[c + 1] += startB[c]; }
  int *fillA = (int *)calloc(ncell, sizeof(int)), *fillB = (int *)calloc(ncell, sizeof(int));
  for (int i = 0; i < nA; i++) listA[startA[cellof[i]] + fillA[cellof[i]]++] = i;
  for (int i = 0; i < nB; i++) listB[startB[cellof[nA + i]] + fillB[cellof[nA + i]]++] = nA + i;
  int64_t np = 0;
  int req[27];
  for (int c = 0; c < ncell; c++) {
    if (startA[c + 1] == startA[c]) continue;
    int nr = o_required_cells(nc, c, req);
    for (int ia = startA[c]; ia < startA[c + 1]; ia++) {
      int A = listA[ia];
      for (int r = 0; r < nr; r++) {
        int cb = req[r];
        if (two) {
          for (int ib = startB[cb]; ib < startB[cb + 1]; ib++) {
            if (np < cap) { pairs[2 * np] = A; pairs[2 * np + 1] = listB[ib]; }
            np++;
          }
        } else {
          for (int ib = startA[cb]; ib < startA[cb + 1]; ib++) {
            int B = listA[ib];
            if (B > A) {
              if (np < cap) { pairs[2 * np] = A; pairs[2 * np + 1] = B; }
              np++;
            }
          }
        }
      }
    }
  }
  free(cellof); free(startA); free(startB); free(listA); free(listB); free(fillA); free(fillB);
  return np;
}

/* NeighborList::update, classic branch (NeighborList.cpp:237-308):
 * pairs with |d|^2 <= cutoff^2 (pbc via Pbc::distance when pbc is on) */
OR_EXPORT int64_t onl_classic(const opbc_t *p, int pbc, double cutoff, const double *pos, int nA, int nB,
                              int dopair, int32_t *pairs, int64_t cap) {
  double d2 = cutoff * cutoff;
  int64_t np = 0;
  double d[3];
  if (nB > 0 && dopair) {
    for (int i = 0; i < nA; i++) {
      const double *a = pos + 3 * i, *b = pos + 3 * (nA + i);
      if (pbc) opbc_distance(p, a, b, d); else { d[0] = b[0] - a[0]; d[1] = b[1] - a[1]; d[2] = b[2] - a[2]; }
      if (o_mod2(d) <= d2) { if (np < cap) { pairs[2 * np] = i; pairs[2 * np + 1] = nA + i; } np++; }
    }
    return np;
  }
  for (int i = 0; i < nA; i++) {
    int j0 = nB > 0 ? nA : i + 1, j1 = nB > 0 ? nA + nB : nA;
    for (int j = j0; j < j1; j++) {
      const double *a = pos + 3 * i, *b = pos + 3 * j;
      if (pbc) opbc_distance(p, a, b, d); else { d[0] = b[0] - a[0]; d[1] = b[1] - a[1]; d[2] = b[2] - a[2]; }
      if (o_mod2(d) <= d2) { if (np < cap) { pairs[2 * np] = i; pairs[2 * np + 1] = j; } np++; }
    }
  }
  return np;
}

/* ------------------------------------------------------------------------ */
/* COORDINATION: src/colvar/CoordinationBase.cpp:142-232                     */
/* ------------------------------------------------------------------------ */

static void o_pair_term(const osw_t *s, const opbc_t *p, int pbc, const double *pos, int i0, int i1,
                        double *ncoord, double *deriv, double *virial, int64_t *ncontrib) {
  double d[3];
  const double *a = pos + 3 * i0, *b = pos + 3 * i1;
  if (pbc) opbc_distance(p, a, b, d); else { d[0] = b[0] - a[0]; d[1] = b[1] - a[1]; d[2] = b[2] - a[2]; }
  double dfunc = 0.0;
  double r2 = o_mod2(d);
  *ncoord += osw_calculate_sqr(s, r2, &dfunc);
  if (r2 <= s->dmax_2) (*ncontrib)++;
  double dd[3] = {dfunc * d[0], dfunc * d[1], dfunc * d[2]};
  for (int k = 0; k < 3; k++) { deriv[3 * i0 + k] -= dd[k]; deriv[3 * i1 + k] += dd[k]; }
  for (int x = 0; x < 3; x++) for (int y = 0; y < 3; y++) virial[3 * x + y] -= dd[x] * d[y];
}

/*
 * CoordinationBase::calculate.  pos: requested atoms (GROUPA then GROUPB),
 * absidx: their absolute indices (pairs with equal absolute index are skipped,
 * :183).  If pairs==NULL the pair set is the full list of
 * NeighborList::initialize (:133-140): TwoList i*nB+j, SingleList i<j,
 * Pair i<->nA+i.  Outputs: value, deriv[3*(nA+nB)], virial[9] (box derivative,
 * -sum dd (x) d).  Returns the number of contributing pairs (r^2 <= dmax^2).
 */
OR_EXPORT int64_t ocoord_calculate(const osw_t *s, const opbc_t *p, int pbc, const double *pos,
                                   const int64_t *absidx, int nA, int nB, int dopair,
                                   const int32_t *pairs, int64_t npairs,
                                   double *value, double *deriv, double *virial) {
  int natoms = nA + nB;
  memset(deriv, 0, sizeof(double) * 3 * natoms);
  memset(virial, 0, sizeof(double) * 9);
  double ncoord = 0.0;
  int64_t nc = 0;
  if (pairs) {
    for (int64_t k = 0; k < npairs; k++) {
      int i0 = pairs[2 * k], i1 = pairs[2 * k + 1];
      if (absidx[i0] == absidx[i1]) continue;
      o_pair_term(s, p, pbc, pos, i0, i1, &ncoord, deriv, virial, &nc);
    }
  } else if (nB > 0 && dopair) {
    for (int i = 0; i < nA; i++) {
      if (absidx[i] == absidx[nA + i]) continue;
      o_pair_term(s, p, pbc, pos, i, nA + i, &ncoord, deriv, virial, &nc);
    }
  } else if (nB > 0) {
    for (int i = 0; i < nA; i++)
      for (int j = nA; j < natoms; j++) {
        if (absidx[i] == absidx[j]) continue;
        o_pair_term(s, p, pbc, pos, i, j, &ncoord, deriv, virial, &nc);
      }
  } else {
    for (int i = 0; i < nA; i++)
      for (int j = i + 1; j < nA; j++) {
        if (absidx[i] == absidx[j]) continue;
        o_pair_term(s, p, pbc, pos, i, j, &ncoord, deriv, virial, &nc);
      }
  }
  *value = ncoord;
  return nc;
}

/* ------------------------------------------------------------------------ */
/* COORDINATIONNUMBER: src/symfunc/CoordinationNumbers.cpp:326-372, a        */
/* CONTACT_MATRIX (src/adjmat/ContactMatrix.h:95-121) times a vector of ONES  */
/* ------------------------------------------------------------------------ */

/*
 * cn[k] = sum over env atoms j of the matrix element of (rows[k], j):
 * ContactMatrix::calculateWeight on pos = Pbc::distance(x_i, x_j)
 * (AdjacencyMatrixBase): skipped when |pos|^2 < DBL_EPSILON (an atom is not
 * bonded to itself, nor to another atom at the same place), s = calculateSqr,
 * zeroed when s < DBL_EPSILON.  With w != NULL the vector-Jacobian product
 * deriv[3*natoms] = sum_k w[k] d cn[k] / d x and the box derivative
 * virial[9] = -sum_k w[k] sum_j dfunc pos (x) pos (the Colvar layout of
 * CoordinationBase.cpp:200-206) are accumulated too.  Indices outside
 * [0, natoms) make the row NaN.
 */
OR_EXPORT void ocn_rows(const osw_t *s, const opbc_t *p, int pbc, const double *pos, int natoms, int nrows,
                        const int32_t *rows, int nenv, const int32_t *env, const double *w, double *cn,
                        double *deriv, double *virial) {
  if (w) {
    memset(deriv, 0, sizeof(double) * 3 * (size_t)natoms);
    memset(virial, 0, sizeof(double) * 9);
  }
  for (int k = 0; k < nrows; k++) {
    const int i = rows[k];
    if (i < 0 || i >= natoms) { cn[k] = NAN; continue; }
    double sum = 0.0;
    for (int q = 0; q < nenv; q++) {
      const int j = env[q];
      if (j < 0 || j >= natoms) { sum = NAN; break; }
      double d[3];
      const double *a = pos + 3 * (size_t)i, *b = pos + 3 * (size_t)j;
      if (pbc) opbc_distance(p, a, b, d); else { d[0] = b[0] - a[0]; d[1] = b[1] - a[1]; d[2] = b[2] - a[2]; }
      const double mod2 = o_mod2(d);
      if (mod2 < DBL_EPSILON) continue;
      double dfunc = 0.0;
      const double v = osw_calculate_sqr(s, mod2, &dfunc);
      if (v < DBL_EPSILON) continue;
      sum += v;
      if (w) {
        const double f = w[k] * dfunc;
        for (int c = 0; c < 3; c++) {
          deriv[3 * (size_t)i + c] -= f * d[c];
          deriv[3 * (size_t)j + c] += f * d[c];
        }
        for (int x = 0; x < 3; x++) for (int y = 0; y < 3; y++) virial[3 * x + y] -= f * d[x] * d[y];
      }
    }
    cn[k] = sum;
  }
}

/* ------------------------------------------------------------------------ */
/* DISTANCE between two centres: COM (src/vatom/Center.cpp:236-445) of each  */
/* group, then DISTANCE (src/colvar/Distance.h:141-250) between them          */
/* ------------------------------------------------------------------------ */

/* one centre: ActionAtomistic::makeWhole (src/core/ActionAtomistic.cpp:
 * 587-607, no MOLINFO) chains each atom to the previous one through
 * Pbc::distance, then pos += w_i x_i in atom order with w_i = m_i / M (COM,
 * Center.cpp:333-367) */
static void o_center(const opbc_t *p, int pbc, const double *pos, int n, const int32_t *idx, const double *m,
                     double *w, double c[3]) {
  double *x = (double *)malloc(sizeof(double) * 3 * (size_t)(n > 0 ? n : 1));
  for (int k = 0; k < n; k++) for (int q = 0; q < 3; q++) x[3 * k + q] = pos[3 * (size_t)idx[k] + q];
  if (pbc)
    for (int j = 0; j + 1 < n; j++) {
      double d[3];
      opbc_distance(p, x + 3 * j, x + 3 * (j + 1), d);
      for (int q = 0; q < 3; q++) x[3 * (j + 1) + q] = x[3 * j + q] + d[q];
    }
  double mass = 0.0;
  for (int k = 0; k < n; k++) mass += m ? m[k] : 1.0;
  for (int k = 0; k < n; k++) w[k] = (m ? m[k] : 1.0) / mass;
  c[0] = c[1] = c[2] = 0.0;
  for (int k = 0; k < n; k++) for (int q = 0; q < 3; q++) c[q] += w[k] * x[3 * k + q];
  free(x);
}

/*
 * c1: COM ATOMS=idx1, c2: COM ATOMS=idx2, d: DISTANCE ATOMS=c1,c2.
 * masses: [n1+n2] (group 1 then group 2) or NULL for equal masses.  With pbc
 * both centres come from whole groups and the distance is made whole as
 * well (Distance.h:150-152).  Outputs: the value; deriv[(n1+n2)*3], the
 * derivative with respect to each listed atom (the chain rule through the
 * centres' derivatives w_i * identity); virial[9], the box derivative
 * -sum_v c_v (x) dd/dc_v of DISTANCE (Colvar::setBoxDerivativesNoPbc; the
 * centres, linear in the atoms, add none, ActionWithVirtualAtom.h:54-63).
 * An atom listed in both groups collects both terms in its two entries.
 */
OR_EXPORT double ocenter_distance(const opbc_t *p, int pbc, const double *pos, int n1, const int32_t *idx1, int n2,
                                  const int32_t *idx2, const double *masses, double *deriv, double *virial,
                                  double *c1out, double *c2out) {
  double *w = (double *)malloc(sizeof(double) * (size_t)(n1 + n2 > 0 ? n1 + n2 : 1));
  double c1[3], c2[3];
  o_center(p, pbc, pos, n1, idx1, masses, w, c1);
  o_center(p, pbc, pos, n2, idx2, masses ? masses + n1 : NULL, w + n1, c2);
  if (pbc) {
    double d[3];
    opbc_distance(p, c1, c2, d);
    for (int q = 0; q < 3; q++) c2[q] = c1[q] + d[q];
  }
  double d[3] = {c2[0] - c1[0], c2[1] - c1[1], c2[2] - c1[2]};
  const double value = sqrt(o_mod2(d));
  const double inv = 1.0 / value;
  double g0[3], g1[3];
  for (int q = 0; q < 3; q++) { g0[q] = -inv * d[q]; g1[q] = inv * d[q]; }
  for (int k = 0; k < n1; k++) for (int q = 0; q < 3; q++) deriv[3 * k + q] = w[k] * g0[q];
  for (int k = 0; k < n2; k++) for (int q = 0; q < 3; q++) deriv[3 * (n1 + k) + q] = w[n1 + k] * g1[q];
  for (int a = 0; a < 3; a++)
    for (int b = 0; b < 3; b++) virial[3 * a + b] = -(c1[a] * g0[b]) - c2[a] * g1[b];
  if (c1out) memcpy(c1out, c1, sizeof c1);
  if (c2out) memcpy(c2out, c2, sizeof c2);
  free(w);
  return value;
}

/* ------------------------------------------------------------------------ */
/* METAD grid: src/tools/Grid.cpp, src/bias/MetaD.cpp                        */
/* ------------------------------------------------------------------------ */

#define OG_MAXDIM 4
typedef struct {
  int dim;
  int pbc[OG_MAXDIM];
  unsigned nbin[OG_MAXDIM]; /* number of nodes (nbin+1 when not periodic) */
  double min[OG_MAXDIM], max[OG_MAXDIM], dx[OG_MAXDIM];
  size_t size;
  double *grid; /* value per node */
  double *der;  /* gradient per node, interleaved [node*dim + i] (Grid.cpp:904-910) */
} ogrid_t;

/* GridBase::Init (Grid.cpp:239-286) */
OR_EXPORT int ogrid_init(ogrid_t *g, int dim, const double *gmin, const double *gmax, const unsigned *nbin, const int *pbc) {
  if (dim < 1 || dim > OG_MAXDIM) return -1;
  memset(g, 0, sizeof(*g));
  g->dim = dim;
  g->size = 1;
  for (int i = 0; i < dim; i++) {
    g->pbc[i] = pbc[i];
    g->min[i] = gmin[i];
    g->max[i] = gmax[i];
    g->nbin[i] = nbin[i];
    g->dx[i] = (g->max[i] - g->min[i]) / (double)nbin[i];
    if (!pbc[i]) { g->max[i] += g->dx[i]; g->nbin[i] += 1; }
    g->size *= g->nbin[i];
  }
  g->grid = (double *)calloc(g->size, sizeof(double));
  g->der = (double *)calloc(g->size * dim, sizeof(double));
  return 0;
}
OR_EXPORT void ogrid_free(ogrid_t *g) { free(g->grid); free(g->der); g->grid = g->der = NULL; }
OR_EXPORT size_t ogrid_sizeof(void) { return sizeof(ogrid_t); }
OR_EXPORT size_t ogrid_size(const ogrid_t *g) { return g->size; }
OR_EXPORT double *ogrid_values(ogrid_t *g) { return g->grid; }
OR_EXPORT double *ogrid_derivs(ogrid_t *g) { return g->der; }

/* column-major flat index (Grid.cpp:108-118); returns -1 outside the grid */
static int64_t o_gindex(const ogrid_t *g, const unsigned *ind) {
  for (int i = 0; i < g->dim; i++) if (ind[i] >= g->nbin[i]) return -1;
  int64_t index = ind[g->dim - 1];
  for (int i = g->dim - 1; i > 0; --i) index = index * g->nbin[i - 1] + ind[i - 1];
  return index;
}
static void o_gindices(const ogrid_t *g, int64_t index, unsigned *ind) {
  for (int i = 0; i < g->dim - 1; i++) { ind[i] = (unsigned)(index % g->nbin[i]); index /= g->nbin[i]; }
  ind[g->dim - 1] = (unsigned)index;
}
/* getIndices(x) (Grid.cpp:126-132): floor((x-min)/dx) */
static void o_gindices_x(const ogrid_t *g, const double *x, unsigned *ind) {
  for (int i = 0; i < g->dim; i++) ind[i] = (unsigned)floor((x[i] - g->min[i]) / g->dx[i]);
}

/* Gaussian value/derivative at cv (MetaD.cpp:2186-2278, diagonal hill, no
 * INTERVAL).  der[] is accumulated into (der -= ...). */
/* Value::difference (src/core/Value.h:405-416): d2-d1, wrapped by Tools::pbc
 * when the argument is periodic (period > 0) */
static double o_vdiff(double period, double d1, double d2) {
  if (!(period > 0.0)) return d2 - d1;
  double inv = 1.0 / period;
  double s = (d2 - d1) * inv;
  s = o_pbc1(s);
  return s * period;
}

static double o_gauss(int ncv, const double *cv, const double *center, const double *sigma, double height,
                      double stretchA, double stretchB, double *der, const double *period) {
  double dp[OG_MAXDIM], inv[OG_MAXDIM], dp2 = 0.0, bias = 0.0;
  for (int i = 0; i < ncv; i++) {
    inv[i] = fabs(sigma[i]) > 1.e-20 ? 1.0 / sigma[i] : 0.0;
    dp[i] = o_vdiff(period ? period[i] : 0.0, center[i], cv[i]) * inv[i];
    dp2 += dp[i] * dp[i];
  }
  dp2 *= 0.5;
  if (dp2 < 6.25) {
    bias = height * exp(-dp2);
    for (int i = 0; i < ncv; i++) der[i] -= bias * dp[i] * inv[i] * stretchA;
    bias = stretchA * bias + height * stretchB;
  }
  return bias;
}

/* MetaD::addGaussian grid branch (MetaD.cpp:1886-1905) with
 * getGaussianSupport (:1939-1994) and GridBase::getNeighbors (Grid.cpp:55-106).
 * Returns the number of nodes touched, or -1 if the centre is outside the grid. */
OR_EXPORT int64_t ogrid_add_gaussian(ogrid_t *g, const double *center, const double *sigma, double height,
                                     double stretchA, double stretchB) {
  int dim = g->dim;
  unsigned nneigh[OG_MAXDIM], ind[OG_MAXDIM], small_bin[OG_MAXDIM], small[OG_MAXDIM], tmp[OG_MAXDIM];
  for (int i = 0; i < dim; i++) {
    double cutoff = sqrt(2.0 * 6.25) * sigma[i];
    nneigh[i] = (unsigned)ceil(cutoff / g->dx[i]);
  }
  o_gindices_x(g, center, ind);
  if (o_gindex(g, ind) < 0) return -1;
  unsigned small_nbin = 1;
  for (int j = 0; j < dim; j++) { small_bin[j] = 2 * nneigh[j] + 1; small_nbin *= small_bin[j]; small[j] = 0; }
  int64_t touched = 0;
  for (unsigned index = 0; index < small_nbin; ++index) {
    int ll = 0;
    for (int i = 0; i < dim; i++) {
      int i0 = (int)small[i] - (int)nneigh[i] + (int)ind[i];
      if (!g->pbc[i] && i0 < 0) continue;
      if (!g->pbc[i] && i0 >= (int)g->nbin[i]) continue;
      if (g->pbc[i] && i0 < 0) i0 = g->nbin[i] - (-i0) % g->nbin[i];
      if (g->pbc[i] && i0 >= (int)g->nbin[i]) i0 %= g->nbin[i];
      tmp[ll++] = (unsigned)i0;
    }
    if (ll == dim) {
      int64_t node = o_gindex(g, tmp);
      double xx[OG_MAXDIM], der[OG_MAXDIM] = {0, 0, 0, 0};
      for (int i = 0; i < dim; i++) xx[i] = g->min[i] + (double)tmp[i] * g->dx[i];
      double period[OG_MAXDIM];
      for (int i = 0; i < dim; i++) period[i] = g->pbc[i] ? g->max[i] - g->min[i] : 0.0;
      double b = o_gauss(dim, xx, center, sigma, height, stretchA, stretchB, der, period);
      g->grid[node] += b;
      for (int i = 0; i < dim; i++) g->der[node * dim + i] += der[i];
      touched++;
    }
    small[0]++;
    for (int j = 0; j < dim - 1; j++)
      if (small[j] == small_bin[j]) { small[j] = 0; small[j + 1]++; }
  }
  return touched;
}

/* GridBase::getValueAndDerivatives(x) spline branch (Grid.cpp:442-505),
 * with getSplineNeighbors (:328-356).  Returns 0, or -1 outside the grid. */
OR_EXPORT int ogrid_spline(const ogrid_t *g, const double *x, double *value, double *der) {
  int dim = g->dim;
  unsigned ind[OG_MAXDIM], nind[OG_MAXDIM];
  double xfloor[OG_MAXDIM], C[OG_MAXDIM], D[OG_MAXDIM], fd[OG_MAXDIM];
  double val = 0.0;
  for (int i = 0; i < dim; i++) der[i] = 0.0;
  o_gindices_x(g, x, ind);
  if (o_gindex(g, ind) < 0) return -1;
  for (int i = 0; i < dim; i++) xfloor[i] = g->min[i] + (double)ind[i] * g->dx[i];
  unsigned nneigh = 1u << dim;
  for (unsigned ip = 0; ip < nneigh; ip++) {
    unsigned t = ip, cnt = 0, ni[OG_MAXDIM];
    for (int j = 0; j < dim; j++) {
      unsigned i0 = t % 2 + ind[j];
      t /= 2;
      if (!g->pbc[j] && i0 == g->nbin[j]) continue;
      if (g->pbc[j] && i0 == g->nbin[j]) i0 = 0;
      ni[cnt++] = i0;
    }
    if ((int)cnt != dim) continue;
    int64_t node = o_gindex(g, ni);
    double grid = g->grid[node];
    o_gindices(g, node, nind);
    double ff = 1.0;
    for (int j = 0; j < dim; j++) {
      int x0 = (nind[j] == ind[j]) ? 0 : 1;
      double dx = g->dx[j];
      double X = fabs((x[j] - xfloor[j]) / dx - (double)x0);
      double X2 = X * X, X3 = X2 * X;
      double yy = fabs(grid) < 0.0000001 ? 0.0 : -g->der[node * dim + j] / grid;
      double sg = x0 ? -1.0 : 1.0;
      C[j] = (1.0 - 3.0 * X2 + 2.0 * X3) - sg * yy * (X - 2.0 * X2 + X3) * dx;
      D[j] = (-6.0 * X + 6.0 * X2) - sg * yy * (1.0 - 4.0 * X + 3.0 * X2) * dx;
      D[j] *= sg / dx;
      ff *= C[j];
    }
    for (int j = 0; j < dim; j++) {
      fd[j] = D[j];
      for (int i = 0; i < dim; i++) if (i != j) fd[j] *= C[i];
    }
    val += grid * ff;
    for (int j = 0; j < dim; j++) der[j] += grid * fd[j];
  }
  *value = val;
  return 0;
}

/*
 * Well-tempered deposition of a sequence of hills (MetaD::update :2377-2470 +
 * getHeight :2280-2305 with the spline getBias): for each hill k, h_k =
 * height0*exp(-V(c_k)/(kbt*(biasf-1))) (biasf>1) and the Gaussian is added to
 * the grid.  heights_out[k] receives h_k (the stored Gaussian height; HILLS
 * files print h_k*biasf/(biasf-1), MetaD.cpp:1874-1879).  biasf<=1 deposits
 * height0.  Returns 0, or -1-k if hill k lies outside the grid.
 */
OR_EXPORT int ogrid_deposit_wt(ogrid_t *g, int nhills, const double *centers, const double *sigma, double height0,
                               double biasf, double kbt, double stretchA, double stretchB, double *heights_out) {
  int dim = g->dim;
  for (int k = 0; k < nhills; k++) {
    const double *c = centers + (size_t)k * dim;
    double h = height0;
    if (biasf > 1.0) {
      double v, der[OG_MAXDIM];
      if (ogrid_spline(g, c, &v, der) < 0) return -1 - k;
      h = height0 * exp(-v / (kbt * (biasf - 1.0)));
    }
    if (heights_out) heights_out[k] = h;
    if (ogrid_add_gaussian(g, c, sigma, h, stretchA, stretchB) < 0) return -1 - k;
  }
  return 0;
}

/* Hills-list (no GRID) bias: sum of Gaussians (MetaD.cpp:2031-2097) */
OR_EXPORT double ometad_bias_hills(int ncv, int nhills, const double *centers, const double *sigma,
                                   const double *heights, const double *cv, double stretchA, double stretchB,
                                   double *der) {
  double bias = 0.0;
  for (int i = 0; i < ncv; i++) der[i] = 0.0;
  for (int k = 0; k < nhills; k++)
    bias += o_gauss(ncv, cv, centers + (size_t)k * ncv, sigma, heights[k], stretchA, stretchB, der, NULL);
  return bias;
}
