/*
 * plumed2_amd: MI355X (gfx950) implementation of PLUMED's COORDINATION
 * collective variable and of the METAD bias grid that consumes it.
 *
 * C ABI.  Plain pointers and sizes, integer status codes, no exceptions.
 * Every context owns its device buffers and one HIP stream (or uses the stream
 * the caller sets); contexts are independent, so several PLUMED instances in
 * different threads may each own theirs.  Errors set a thread-local message
 * readable with p2a_last_error().
 *
 * Units and layouts follow the reference:
 *   positions      AoS double [natoms][3], the requested atoms in the order
 *                  GROUPA then GROUPB (CoordinationBase.cpp:101
 *                  requestAtoms(nl->getFullAtomList()))
 *   box            double[9], row i = lattice vector i (Pbc::setBox,
 *                  Pbc.cpp:165)
 *   derivatives    double [natoms][3], d CV / d x  (Colvar.h:73-103 layout
 *                  3*i+k)
 *   virial         double[9], the box derivative  -sum dd (x) d, row-major
 *                  (CoordinationBase.cpp:200-206, setBoxDerivatives)
 */
#ifndef PLUMED2_AMD_H
#define PLUMED2_AMD_H

#include <stddef.h>
#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

#if defined(__GNUC__)
#pragma GCC visibility push(default)
#endif

#define P2A_OK 0
#define P2A_ERR_ARG (-1)      /* invalid argument or input syntax          */
#define P2A_ERR_HIP (-2)      /* HIP runtime error                          */
#define P2A_ERR_NOMEM (-3)    /* device allocation failed                   */
#define P2A_ERR_UNSUPPORTED (-4) /* valid input the library does not handle */
#define P2A_ERR_GRID (-5)     /* lookup outside the METAD grid (Grid.cpp:108-118) */

/* Library version string. */
const char *p2a_version(void);
/* Message of the last failing call on this thread ("" if none). */
const char *p2a_last_error(void);
/* Number of visible HIP devices (0 when there is no GPU). */
int p2a_device_count(void);

/* ------------------------------------------------------------------------
 * Switching function
 * ------------------------------------------------------------------------ */

/* switch types (SwitchingFunction.h:36-57) */
enum {
  P2A_SW_RATIONAL_FIXED = 0, /* fixedRational<N>, 2N==M, N,M even, D_0==0, N<=12 */
  P2A_SW_RATIONAL = 1,       /* general rational (fast/simplified flags)         */
  P2A_SW_EXP = 2,
  P2A_SW_GAUSSIAN = 3,
  P2A_SW_FASTGAUSSIAN = 4,   /* GAUSSIAN with R_0=1, D_0=0                       */
  P2A_SW_SMAP = 5,
  P2A_SW_CUBIC = 6,
  P2A_SW_TANH = 7,
  P2A_SW_COSINUS = 8,
  P2A_SW_Q = 9
};

/* Fully resolved switching-function parameters (SwitchingFunction.h:58-94
 * switchContainers::Data plus the dispatch decision of rationalFactory). */
typedef struct p2a_switch {
  int32_t type;
  int32_t fast, simplified;
  int32_t nn, mm, nnf, mmf;
  int32_t a, b;
  int32_t pad_;
  double d0, dmax, dmax_2, invr0, invr0_2, stretch, shift;
  double preRes, preDfunc, preSecDev, preDfuncF, preSecDevF;
  double c, d;
  double beta, lambda, ref;
} p2a_switch;

/* Parse the body of a SWITCH={...} keyword.  Replaces
 * SwitchingFunction::set(const std::string&, std::string&)
 * (src/tools/SwitchingFunction.cpp:1055-1159): same keywords (R_0 D_0 D_MAX
 * NN MM A B BETA LAMBDA REF NOSTRETCH STRETCH), same defaults, same error
 * texts ("could not parse D_MAX", "R_0 is required for RATIONAL", "found the
 * following rogue keywords in switching function input : ...").  CUSTOM /
 * MATHEVAL (lepton) returns P2A_ERR_UNSUPPORTED. */
int p2a_switch_parse(const char *definition, p2a_switch *out);

/* NN/MM/R_0/D_0 form.  Replaces SwitchingFunction::set(int,int,double,double)
 * (src/tools/SwitchingFunction.cpp:1176-1184): MM=0 means 2*NN, D_MAX set
 * where s=1e-5, always stretched. */
int p2a_switch_rational(int nn, int mm, double r0, double d0, p2a_switch *out);

/* One-line description, as SwitchingFunction::description(). */
int p2a_switch_describe(const p2a_switch *sw, char *buf, size_t len);

/* ------------------------------------------------------------------------
 * COORDINATION
 * ------------------------------------------------------------------------ */

enum { P2A_NL_NONE = 0, P2A_NL_CLASSIC = 1, P2A_NL_CELLS = 2 };

/* Parsed action keywords of one COORDINATION line (Coordination.cpp:115-157,
 * CoordinationBase.cpp:30-131).  Atom lists hold 0-based absolute indices.
 * Arrays are owned by the library (p2a_coord_spec_free). */
typedef struct p2a_coord_spec {
  int64_t *groupa;
  int64_t na;
  int64_t *groupb;
  int64_t nb;
  int32_t pbc;       /* !NOPBC                                 */
  int32_t pair;      /* PAIR                                   */
  int32_t serial;    /* SERIAL (accepted, no effect on the GPU) */
  int32_t nl_mode;   /* P2A_NL_*                               */
  double nl_cutoff;  /* NL_CUTOFF                              */
  int32_t nl_stride; /* NL_STRIDE                              */
  int32_t pad_;
  p2a_switch sw;
} p2a_coord_spec;

/* Parse the keywords of a COORDINATION action line (label and action name
 * optional): GROUPA GROUPB (serial numbers, ranges a-b, a-b:s), NN MM D_0 R_0
 * SWITCH NOPBC PAIR SERIAL NLIST NLISTCELLS NL_CUTOFF NL_STRIDE.  Error texts
 * follow the reference ("R_0 should be explicitly specified and positive",
 * "NL_CUTOFF should be explicitly specified and positive", ...).  Host only. */
int p2a_coord_parse(const char *line, p2a_coord_spec *out);
void p2a_coord_spec_free(p2a_coord_spec *spec);

typedef struct p2a_coord p2a_coord;

/* Create a COORDINATION context on HIP device `device` from a parsed spec.
 * Replaces the construction in CoordinationBase::CoordinationBase
 * (src/colvar/CoordinationBase.cpp:42-131) + Coordination::Coordination
 * (src/colvar/Coordination.cpp:128-157). */
int p2a_coord_create(p2a_coord **ctx, int device, const p2a_coord_spec *spec);
int p2a_coord_destroy(p2a_coord *ctx);

/* Run the context's work on the caller's HIP stream (hipStream_t passed as
 * void*); NULL selects the default (null) stream.  Until this is called the
 * context uses a non-blocking stream of its own. */
int p2a_coord_set_stream(p2a_coord *ctx, void *stream);

/* One step of CoordinationBase::prepare + calculate
 * (src/colvar/CoordinationBase.cpp:137-232, NeighborList::prepare
 * NeighborList.cpp:433-456): rebuilds the neighbour structure on the first
 * call and when step % NL_STRIDE == 0, then evaluates value, derivatives and
 * virial.  Host buffers; synchronous.  Any output pointer may be NULL. */
int p2a_coord_calculate(p2a_coord *ctx, int64_t step, const double *box9, const double *pos,
                        double *value, double *deriv, double *virial);

/* Same with device-resident positions and outputs (value: 1 double, deriv:
 * 3*natoms doubles, virial: 9 doubles, all device pointers; any may be NULL).
 * With deriv_dev = virial_dev = NULL only the value is evaluated (a cheaper
 * kernel that visits each pair once).  box9 is a host array.  Asynchronous on
 * the context's stream. */
int p2a_coord_calculate_device(p2a_coord *ctx, int64_t step, const double *box9, const double *pos_dev,
                               double *value_dev, double *deriv_dev, double *virial_dev);

/* Backward pass: derivatives (and the virial if virial_dev is not NULL) of
 * the value of the last calculate_device call, recomputing the pair terms on
 * the neighbour structure that call built (nothing is stored per pair).
 * Positions must be those of that call.  Replaces the derivative part of
 * CoordinationBase::calculate (CoordinationBase.cpp:196-206) when the forward
 * asked for the value only (deriv_dev = virial_dev = NULL).  Asynchronous. */
int p2a_coord_backward_device(p2a_coord *ctx, const double *box9, const double *pos_dev, double *deriv_dev,
                              double *virial_dev);

/* Number of neighbour-structure builds enqueued so far on this context (host
 * counter, no synchronisation).  A caller that keeps the positions of a
 * forward call for a later backward records it after the forward; when it
 * has changed by the time of the backward, another call has replaced the
 * structure, and p2a_coord_rebuild_device restores one for those positions.
 * No reference counterpart (the reference evaluates value and derivatives in
 * one pass, CoordinationBase.cpp:142-232). */
int p2a_coord_list_generation(p2a_coord *ctx, int64_t *generation);

/* Build the neighbour structure now, at these device positions and box, as
 * NeighborList::update does on a rebuild step (NeighborList.cpp:168-308).
 * Asynchronous on the context's stream; no-op on the FP64 paths (NOPBC,
 * PAIR, narrow cells), which keep no structure.  With NLIST/NLISTCELLS the
 * list built here serves the following steps until the next NL_STRIDE step. */
int p2a_coord_rebuild_device(p2a_coord *ctx, const double *box9, const double *pos_dev);

/* Allocation-free COORDINATION: value, derivatives and virial as
 * p2a_coord_calculate_device returns them, without a context, computed by the
 * FP64 all-pairs kernels of that entry point (every pair of GROUPA x GROUPB,
 * or every pair i<j of GROUPA when nb = 0, evaluated at every call).  No
 * counterpart in the reference: it is for callers that cannot allocate
 * device memory (code under HIP graph capture, tight serving loops, harnesses
 * that forbid hipMalloc).  It calls no hipMalloc/hipFree: the per-workgroup
 * partials live in a static device array of 64 KiB per device (8 doubles per
 * 256 atoms), so na + nb is at most 262144 (P2A_ERR_ARG beyond).  Every call
 * on one device shares that array: calls that overlap on one device (other
 * streams, other threads) race and must be ordered by the caller.
 *   sw          resolved switch (p2a_switch_parse / p2a_switch_rational)
 *   box9        host double[9], rows = lattice vectors (ignored if !use_pbc)
 *   na, nb      GROUPA and GROUPB sizes; positions are GROUPA then GROUPB
 *   absidx_dev  device int64[na+nb] absolute atom indices (pairs of one atom
 *               with itself are skipped, CoordinationBase.cpp:183) or NULL
 *               when all requested atoms are distinct
 *   pos_dev     device double[na+nb][3]; value_dev [1], deriv_dev
 *               [na+nb][3], virial_dev [9]: device outputs
 * Runs on the current device, asynchronously on `stream` (NULL: default
 * stream), which must belong to the current device. */
int p2a_coord_calculate_noalloc(void *stream, const p2a_switch *sw, const double *box9, int use_pbc, int64_t na,
                                int64_t nb, const int64_t *absidx_dev, const double *pos_dev, double *value_dev,
                                double *deriv_dev, double *virial_dev);

typedef struct p2a_coord_stats {
  int64_t natoms;
  int64_t contributing_pairs; /* pairs with r^2 <= D_MAX^2 at the last step   */
  int64_t clusters_i, clusters_j;
  int64_t list_entries;       /* cluster-pair list length at the last rebuild */
  int64_t rebuilds;
  int32_t path;               /* 0 cluster-pair kernel, 1 all-pairs FP64, 2 PAIR */
  int32_t pad_;
  double pair_kernel_ms;      /* summed pair-kernel time since profiling was enabled */
  int64_t pair_kernel_launches;
  int64_t total_contributing_pairs; /* summed over the steps since profiling was enabled */
} p2a_coord_stats;

/* Statistics of the last step.  Synchronises the context's stream. */
int p2a_coord_get_stats(p2a_coord *ctx, p2a_coord_stats *stats);
/* Record HIP events around the pair-kernel launches (on the launch stream) of
 * every on-th call (1: every call; 0: off) and accumulate their durations into
 * pair_kernel_ms / pair_kernel_launches; on>0 also resets those sums and
 * total_contributing_pairs, which keeps counting at every call.  Each event
 * record costs a few microseconds of stream time, hence the sampling. */
int p2a_coord_set_profiling(p2a_coord *ctx, int on);

/* Work decomposition of the pair kernel: `chunk` list entries per wave (32 or
 * 64) and a persistent grid of `blocks_per_cu` workgroups per compute unit
 * (0 = one wave per work item).  Results do not depend on it.  No reference
 * counterpart (the reference's OpenMP loop has no such knob). */
int p2a_coord_set_tuning(p2a_coord *ctx, int chunk, int blocks_per_cu);

/* Without a neighbour list (no NLIST/NLISTCELLS): build the pair list at
 * D_MAX + skin and keep it while no atom has moved more than skin/2 from its
 * position at the build; each call then checks the displacements on the
 * device (one host synchronisation) instead of rebuilding.  The pairs
 * evaluated are those of a list rebuilt at every call, as the reference
 * computes every pair at every step (CoordinationBase.cpp:142-232).  skin 0
 * (the default) rebuilds at every call; a cell too narrow for D_MAX + skin
 * also does.  No reference counterpart: the reference has no list without
 * NLIST. */
int p2a_coord_set_skin(p2a_coord *ctx, double skin);

/* ------------------------------------------------------------------------
 * Batched COORDINATION: many independent single-group systems (frames or
 * walkers, each with its own atom count and periodic cell) sharing one
 * switching function and list radius.  Every kernel of a step covers all
 * systems at once, the neighbour structure is rebuilt at every call
 * (NL_STRIDE=1), and the step can be replayed as one captured HIP graph.
 * Replaces running one COORDINATION action per system
 * (CoordinationBase::prepare + calculate, src/colvar/CoordinationBase.cpp:
 * 137-232, with NLISTCELLS, NL_CUTOFF=r_list, NL_STRIDE=1) and summing the
 * per-system outputs.
 *
 * Layouts: positions are the systems' AoS FP64 arrays concatenated in system
 * order ([sum natoms][3]); derivatives likewise; values[nsys]; virials
 * [nsys][9]; boxes [nsys][9] (rows = lattice vectors).
 * ------------------------------------------------------------------------ */
typedef struct p2a_batch p2a_batch;

typedef struct p2a_batch_stats {
  int64_t nsys, natoms;
  int64_t clusters, list_entries, items; /* of the last step                       */
  int64_t contributing_pairs;            /* sum over systems, last step            */
  int32_t chunk, blocks_per_cu;          /* pair-kernel decomposition in use       */
  int32_t graph;                         /* 1 when the last step ran as a graph    */
  int32_t regrowths;                     /* list arrays grown after an overflow    */
  double last_step_ms;                   /* device time of the last step           */
} p2a_batch_stats;

/* natoms[nsys] > 0 atoms per system; sw = the switch (p2a_switch_parse);
 * rlist = neighbour-list radius (>= D_MAX; NL_CUTOFF of the reference). */
int p2a_batch_create(p2a_batch **b, int device, int nsys, const int64_t *natoms, const p2a_switch *sw,
                     double rlist);
int p2a_batch_destroy(p2a_batch *b);
/* Use the caller's stream (NULL = the default stream; graphs then stay off). */
int p2a_batch_set_stream(p2a_batch *b, void *stream);
/* Periodic cells of all systems (host, [nsys][9]).  Each cell must be wider
 * than 2*rlist in every direction (P2A_ERR_UNSUPPORTED names the system). */
int p2a_batch_set_boxes(p2a_batch *b, const double *boxes);
/* One step over every system; device pointers.  deriv and virials may be NULL
 * together (value only is not specialised: they are computed and dropped).
 * Returns when the step has finished. */
int p2a_batch_calculate_device(p2a_batch *b, const double *pos_dev, double *values_dev, double *deriv_dev,
                               double *virials_dev);
/* Contributing pairs per system of the last step (host int64[nsys]). */
int p2a_batch_get_pair_counts(p2a_batch *b, int64_t *out);
/* Capture the step into a HIP graph on the first call and replay it after
 * (on != 0); re-captured when pointers, tuning or array sizes change. */
int p2a_batch_set_graph(p2a_batch *b, int on);
int p2a_batch_set_tuning(p2a_batch *b, int chunk, int blocks_per_cu);
/* Time `reps` steps of every candidate decomposition on these positions,
 * keep the fastest; *chunk / *blocks_per_cu receive the choice (may be NULL). */
int p2a_batch_autotune(p2a_batch *b, const double *pos_dev, int reps, int *chunk, int *blocks_per_cu);
int p2a_batch_get_stats(p2a_batch *b, p2a_batch_stats *stats);

/* ------------------------------------------------------------------------
 * METAD bias on a grid (diagonal Gaussian hills, well-tempered heights)
 * ------------------------------------------------------------------------ */

typedef struct p2a_metad p2a_metad;

typedef struct p2a_metad_spec {
  int32_t ncv;                /* 1..4                                             */
  int32_t periodic[4];        /* per CV (grid wraps, differences use the period)   */
  double gmin[4], gmax[4];    /* GRID_MIN/GRID_MAX (or the CV domain if periodic)  */
  uint32_t nbin[4];           /* GRID_BIN                                          */
  double sigma[4];            /* SIGMA                                             */
  double height;              /* HEIGHT                                            */
  double biasf;               /* BIASFACTOR (<=1: not well-tempered)               */
  double kbt;                 /* k_B T (TEMP * 0.0083144621)                       */
  int32_t pace;               /* PACE                                              */
  int32_t nostretch;          /* 1 mimics PLUMED_DP2CUTOFF_NOSTRETCH               */
} p2a_metad_spec;

/* Grid creation: MetaD::MetaD grid set-up (src/bias/MetaD.cpp:1006-1060,
 * 1466-1474) and GridBase::Init (src/tools/Grid.cpp:239-286): non-periodic
 * directions get nbin+1 nodes; value and gradient per node in FP64. */
int p2a_metad_create(p2a_metad **ctx, int device, const p2a_metad_spec *spec);
int p2a_metad_destroy(p2a_metad *ctx);
int p2a_metad_set_stream(p2a_metad *ctx, void *stream);
/* number of grid nodes */
int64_t p2a_metad_grid_size(const p2a_metad *ctx);

/* Deposit nhills hills in sequence, each with the well-tempered height
 * h = HEIGHT*exp(-V(c)/(kbt*(BIASFACTOR-1))) of the grid left by the previous
 * ones (MetaD::update :2377-2470, getHeight :2280-2305, addGaussian
 * :1886-1937).  centers: host [nhills][ncv]; heights: host output (may be NULL)
 * receiving the deposited Gaussian heights.  One device launch. */
int p2a_metad_deposit(p2a_metad *ctx, int64_t nhills, const double *centers, double *heights);
/* Same with device arrays; asynchronous. */
int p2a_metad_deposit_device(p2a_metad *ctx, int64_t nhills, const double *centers_dev, double *heights_dev);
/* Add hills of given heights (HILLS restart, MetaD::readGaussians :1807-1829). */
int p2a_metad_add_hills(p2a_metad *ctx, int64_t nhills, const double *centers, const double *heights);

/* Bias and its gradient by the grid spline (GridBase::getValueAndDerivatives
 * spline branch, src/tools/Grid.cpp:442-505) at npoints points: host arrays
 * cv [npoints][ncv] -> bias [npoints], der [npoints][ncv] (der may be NULL).
 * Points outside the grid return P2A_ERR_GRID (Grid.cpp:108-118). */
int p2a_metad_bias(p2a_metad *ctx, int64_t npoints, const double *cv, double *bias, double *der);
int p2a_metad_bias_device(p2a_metad *ctx, int64_t npoints, const double *cv_dev, double *bias_dev, double *der_dev);

/* One MD step of METAD on device-resident CVs: MetaD::calculate (:2315-2375:
 * bias and force -dV/ds) followed by MetaD::update (a hill every PACE steps,
 * not on the first step).  cv_dev [ncv]; bias_dev [1]; force_dev [ncv];
 * heights_dev (may be NULL) receives the height of a hill deposited at this
 * step (0 otherwise).  Asynchronous; errors are latched and reported by
 * p2a_metad_check. */
int p2a_metad_step_device(p2a_metad *ctx, int64_t step, const double *cv_dev, double *bias_dev, double *force_dev,
                          double *height_dev);
/* Synchronise and report a latched device error (P2A_ERR_GRID). */
int p2a_metad_check(p2a_metad *ctx);

/* Copy the grid out / in: values [size], derivatives [size][ncv] (Grid.cpp:878-910 layout). */
int p2a_metad_get_grid(p2a_metad *ctx, double *values, double *derivs);
int p2a_metad_set_grid(p2a_metad *ctx, const double *values, const double *derivs);

/* Chain rule of the bias force onto the atoms, ActionWithValue::checkForForces
 * (src/core/ActionWithValue.cpp:337-413) for one CV: forces[3N] += f * deriv[3N],
 * virial[9] += f * virial_cv[9].  Device arrays; f is a device scalar (the
 * METAD force on that CV).  Asynchronous on `stream` (NULL: default stream). */
int p2a_apply_force_device(void *stream, int64_t natoms, const double *f_dev, const double *deriv_dev,
                           const double *virial_cv_dev, double *forces_dev, double *virial_dev);

/* ------------------------------------------------------------------------
 * DISTANCE between the centres of mass of two groups
 *   c1: COM ATOMS=<group1>   c2: COM ATOMS=<group2>   d: DISTANCE ATOMS=c1,c2
 * the second CV of a METAD on COORDINATION and a distance between centres.
 * ------------------------------------------------------------------------ */
typedef struct p2a_cdist p2a_cdist;

/* Construction of the three actions: Center::Center (src/vatom/Center.cpp:
 * 236-317, COM: mass weights) and Distance::Distance (src/colvar/Distance.h:
 * 75-117).  group1/group2: host arrays of 0-based indices into the positions
 * array of natoms atoms (distinct within each group); masses: host [n1+n2]
 * (group 1 then group 2) or NULL for equal masses; use_pbc = !NOPBC (both the
 * groups and the two centres are made whole, ActionAtomistic::makeWhole). */
int p2a_cdist_create(p2a_cdist **ctx, int device, int64_t natoms, int64_t n1, const int64_t *group1, int64_t n2,
                     const int64_t *group2, const double *masses, int use_pbc);
int p2a_cdist_destroy(p2a_cdist *ctx);
/* Value, derivatives and box derivative: Center::calculate (Center.cpp:
 * 319-445) of both groups, then Distance::calculate (Distance.h:141-250).
 * deriv_dev [n1+n2][3]: derivative with respect to each listed atom (group 1
 * then group 2); virial_dev [9]: the box derivative of DISTANCE (the centres
 * are linear in the atoms and add none).  Device arrays, host box; one
 * workgroup, asynchronous on `stream` (NULL: default stream). */
int p2a_cdist_calculate_device(p2a_cdist *ctx, void *stream, const double *box9, const double *pos_dev,
                               double *value_dev, double *deriv_dev, double *virial_dev);
/* Chain rule of a force f on the distance onto the listed atoms
 * (ActionWithVirtualAtom::apply, src/core/ActionWithVirtualAtom.cpp:68-124,
 * through ActionWithValue::checkForForces): forces[3*i+q] += f * deriv, and
 * virial[9] += f * virial_cv.  f_dev is a device scalar (e.g. the METAD
 * force on this CV); forces_dev is the [natoms][3] array of all atoms. */
int p2a_cdist_apply_force_device(p2a_cdist *ctx, void *stream, const double *f_dev, const double *deriv_dev,
                                 const double *virial_cv_dev, double *forces_dev, double *virial_dev);

/* COORDINATIONNUMBER on the cluster pair list of a COORDINATION context
 * (src/symfunc/CoordinationNumbers.cpp:326-372: CONTACT_MATRIX rows times
 * ONES, adjmat/ContactMatrix.h:95-121): the rows are the GROUPA atoms, the
 * environment GROUPA itself (one group, SPECIES=) or GROUPB (SPECIESA /
 * SPECIESB).  rows_dev[na] = sum over the environment of s(r), pairs at one
 * place (|r|^2 < DBL_EPSILON) left out.  With weights_dev[na] also
 * vjp_dev[na+nb][3] = sum_k w_k dcn_k/dx (the derivative layout of
 * p2a_coord_calculate_device) and virial_dev[9] (may be NULL) = the matching
 * box derivative.  The list follows the context's rules (NLIST/NLISTCELLS,
 * NL_STRIDE, skin).  On the FP64 path (cells narrower than twice the list
 * radius, NOPBC, duplicate atoms) the rows are computed over every pair and
 * weights are refused (P2A_ERR_UNSUPPORTED); PAIR is refused.  Device arrays,
 * host box, asynchronous on the context stream.  No counterpart call in the
 * reference's C API: this is what the COORDINATIONNUMBER action computes. */
int p2a_coord_rows_device(p2a_coord *ctx, int64_t step, const double *box9, const double *pos_dev, double *rows_dev,
                          const double *weights_dev, double *vjp_dev, double *virial_dev);

/* COORDINATIONNUMBER (src/symfunc/CoordinationNumbers.cpp:326-372, a shortcut
 * for CONTACT_MATRIX (src/adjmat/ContactMatrix.cpp) times ONES): per-atom
 * coordination numbers cn[k] = sum over env atoms j != rows[k] of
 * s(|x_j - x_rows[k]|), FP64.  rows/env are 0-based indices into pos_dev
 * (natoms x 3 doubles); SPECIES=L is rows = env = L, SPECIESA/SPECIESB is
 * rows = A, env = B.  Device arrays, no derivatives.  A row whose index, or
 * any of whose env indices, lies outside [0, natoms) comes out NaN.
 * Asynchronous on `stream` (NULL: default stream). */
int p2a_coordination_numbers_device(void *stream, const p2a_switch *sw, const double *box9, int use_pbc,
                                    int64_t natoms, int64_t nrows, const int *rows_dev, int64_t nenv,
                                    const int *env_dev, const double *pos_dev, double *cn_dev);

#if defined(__GNUC__)
#pragma GCC visibility pop
#endif

#ifdef __cplusplus
}
#endif
#endif /* PLUMED2_AMD_H */
