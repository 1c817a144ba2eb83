// PLUMED plugin: COORDINATION evaluated by the plumed2_amd HIP library.
//
// Load into a PLUMED kernel with
//     LOAD FILE=libplumed2amd_plugin.so
// after which the input lines
//     c: COORDINATION_HIP GROUPA=... [GROUPB=...] SWITCH={...} | NN= MM= D_0= R_0= ...
// (and, with PLUMED2AMD_OVERRIDE=1 in the environment at load time, plain
// COORDINATION) are served by the GPU.  Keywords, defaults and error texts are
// the reference action's (src/colvar/Coordination.cpp:113-157,
// src/colvar/CoordinationBase.cpp:30-131); the arithmetic goes through the C
// ABI in include/plumed2_amd.h.  This file is host C++ only: it is compiled
// against an installed PLUMED (`plumed info --include-dir`) and linked with
// libplumed2amd.so.
#include "core/ActionRegister.h"
#include "core/Colvar.h"
#include "tools/AtomNumber.h"
#include "tools/Tensor.h"
#include "tools/Vector.h"

#include <cstdlib>
#include <memory>
#include <string>
#include <vector>

#include "plumed2_amd.h"

namespace PLMD {
namespace p2a_plugin {

class CoordinationHIP : public Colvar {
  p2a_coord* ctx_ = nullptr;
  std::vector<double> deriv_;
  size_t natoms_ = 0;

 public:
  explicit CoordinationHIP(const ActionOptions&);
  ~CoordinationHIP() override;
  static void registerKeywords(Keywords& keys);
  void calculate() override;
};

void CoordinationHIP::registerKeywords(Keywords& keys) {
  Colvar::registerKeywords(keys);
  keys.addFlag("SERIAL", false, "Perform the calculation in serial - for debug purpose");
  keys.addFlag("PAIR", false, "Pair only 1st element of the 1st group with 1st element in the second, etc");
  keys.addFlag("NLIST", false, "Use a neighbor list to speed up the calculation");
  keys.addFlag("NLISTCELLS", false, "Use a neighbor list built from link cells");
  keys.add("optional", "NL_CUTOFF", "The cutoff for the neighbor list");
  keys.add("optional", "NL_STRIDE", "The frequency with which we are updating the atoms in the neighbor list");
  keys.add("atoms", "GROUPA", "First list of atoms");
  keys.add("atoms", "GROUPB", "Second list of atoms (if empty, N*(N-1)/2 pairs in GROUPA are counted)");
  keys.add("compulsory", "NN", "6", "The n parameter of the switching function ");
  keys.add("compulsory", "MM", "0", "The m parameter of the switching function; 0 implies 2*NN");
  keys.add("compulsory", "D_0", "0.0", "The d_0 parameter of the switching function");
  keys.add("compulsory", "R_0", "The r_0 parameter of the switching function");
  keys.add("optional", "SWITCH", "Switching function definition (RATIONAL, EXP, GAUSSIAN, SMAP, CUBIC, TANH, COSINUS, Q)");
  keys.add("compulsory", "DEVICE", "0", "HIP device that evaluates this action");
  keys.setValueDescription("scalar", "the value of the coordination");
}

CoordinationHIP::CoordinationHIP(const ActionOptions& ao) : PLUMED_COLVAR_INIT(ao) {
  bool serial = false;
  parseFlag("SERIAL", serial);
  std::vector<AtomNumber> ga, gb;
  parseAtomList("GROUPA", ga);
  parseAtomList("GROUPB", gb);
  bool nopbc = false;
  parseFlag("NOPBC", nopbc);
  bool dopair = false;
  parseFlag("PAIR", dopair);
  bool classic = false, cells = false;
  parseFlag("NLIST", classic);
  parseFlag("NLISTCELLS", cells);
  if (classic && cells) error("Please activate only one of the two version of the NL");
  if (cells && dopair) error("Pair is not compatible with the CELLS implementation of the NL");
  double nl_cut = 0.0;
  int nl_st = 0;
  if (classic || cells) {
    parse("NL_CUTOFF", nl_cut);
    if (nl_cut <= 0.0) error("NL_CUTOFF should be explicitly specified and positive");
    parse("NL_STRIDE", nl_st);
    if (nl_st <= 0) error("NL_STRIDE should be explicitly specified and positive");
  }
  p2a_coord_spec spec{};
  std::string sw;
  parse("SWITCH", sw);
  if (!sw.empty()) {
    if (p2a_switch_parse(sw.c_str(), &spec.sw) != P2A_OK)
      error(std::string("problem reading SWITCH keyword : ") + p2a_last_error());
  } else {
    int nn = 6, mm = 0;
    double d0 = 0.0, r0 = 0.0;
    parse("R_0", r0);
    if (r0 <= 0.0) error("R_0 should be explicitly specified and positive");
    parse("D_0", d0);
    parse("NN", nn);
    parse("MM", mm);
    if (p2a_switch_rational(nn, mm, r0, d0, &spec.sw) != P2A_OK) error(p2a_last_error());
  }
  int device = 0;
  parse("DEVICE", device);
  checkRead();

  std::vector<int64_t> a(ga.size()), b(gb.size());
  for (size_t i = 0; i < ga.size(); i++) a[i] = ga[i].index();
  for (size_t i = 0; i < gb.size(); i++) b[i] = gb[i].index();
  spec.groupa = a.data();
  spec.na = (int64_t)a.size();
  spec.groupb = b.data();
  spec.nb = (int64_t)b.size();
  spec.pbc = !nopbc;
  spec.pair = dopair;
  spec.serial = serial;
  spec.nl_mode = cells ? P2A_NL_CELLS : (classic ? P2A_NL_CLASSIC : P2A_NL_NONE);
  spec.nl_cutoff = nl_cut;
  spec.nl_stride = nl_st;
  if (p2a_coord_create(&ctx_, device, &spec) != P2A_OK) error(p2a_last_error());

  addValueWithDerivatives();
  setNotPeriodic();
  // the GPU keeps its own neighbour structure: always the full atom list
  std::vector<AtomNumber> all(ga);
  all.insert(all.end(), gb.begin(), gb.end());
  requestAtoms(all);
  natoms_ = all.size();
  deriv_.resize(3 * natoms_);

  char desc[256];
  p2a_switch_describe(&spec.sw, desc, sizeof desc);
  log.printf("  between two groups of %u and %u atoms, evaluated by %s\n", (unsigned)ga.size(), (unsigned)gb.size(),
             p2a_version());
  log.printf("  contacts are counted with cutoff %s\n", desc);
  if (classic || cells) log.printf("  using neighbor lists with update every %d steps and cutoff %f\n", nl_st, nl_cut);
}

CoordinationHIP::~CoordinationHIP() { p2a_coord_destroy(ctx_); }

void CoordinationHIP::calculate() {
  const std::vector<Vector>& pos = getPositions();
  const Tensor& box = getBox();
  double b9[9];
  for (int i = 0; i < 3; i++)
    for (int j = 0; j < 3; j++) b9[3 * i + j] = box[i][j];
  double value = 0.0, vir[9];
  // Vector is three contiguous doubles: the positions are already AoS
  if (p2a_coord_calculate(ctx_, getStep(), b9, &pos[0][0], &value, deriv_.data(), vir) != P2A_OK)
    error(p2a_last_error());
  for (size_t i = 0; i < natoms_; i++)
    setAtomsDerivatives((int)i, Vector(deriv_[3 * i], deriv_[3 * i + 1], deriv_[3 * i + 2]));
  setValue(value);
  setBoxDerivatives(Tensor(vir[0], vir[1], vir[2], vir[3], vir[4], vir[5], vir[6], vir[7], vir[8]));
}

PLUMED_REGISTER_ACTION(CoordinationHIP, "COORDINATION_HIP")

// Optional drop-in override of the built-in name (a LOADed registration is
// preferred for the input lines after LOAD).
namespace {
struct OverrideRegistration {
  std::unique_ptr<ActionRegistration<CoordinationHIP>> reg;
  OverrideRegistration() {
    const char* e = std::getenv("PLUMED2AMD_OVERRIDE");
    if (e && std::string(e) == "1") reg = std::make_unique<ActionRegistration<CoordinationHIP>>("COORDINATION");
  }
} override_registration;
}  // namespace

}  // namespace p2a_plugin
}  // namespace PLMD
