// PLUMED plugin: METAD on a grid, deposited and evaluated by the plumed2_amd
// HIP library.
//
// Load into a PLUMED kernel with
//     LOAD FILE=libplumed2amd_plugin.so
// after which
//     m: METAD_HIP ARG=c,d SIGMA=20,0.05 HEIGHT=1.2 PACE=500 BIASFACTOR=10 TEMP=300 \
//        GRID_MIN=0,0 GRID_MAX=5000,6 GRID_BIN=512,512 [FILE=HILLS] [FMT=%f]
// (and, with PLUMED2AMD_OVERRIDE=1 at load time, plain METAD with these keywords)
// keeps the bias on the GPU grid.  Keywords, the first-step rule, the
// BIASFACTOR=1 rule and the HILLS records follow the reference action
// (src/bias/MetaD.cpp: registerKeywords :640-700, calculate :2315-2375,
// update :2377-2470, writeGaussian :1831-1884); diagonal hills on a grid
// only.  Host C++ only, compiled against an installed PLUMED and linked with
// libplumed2amd.so.
#include "bias/Bias.h"
#include "core/ActionRegister.h"
#include "tools/OFile.h"

#include <cstdlib>
#include <memory>
#include <string>
#include <vector>

#include "plumed2_amd.h"

namespace PLMD {
namespace p2a_plugin {

class MetaDHIP : public bias::Bias {
  p2a_metad* ctx_ = nullptr;
  unsigned ncv_ = 0;
  std::vector<double> sigma_;
  double biasf_ = 1.0;
  bool welltemp_ = false;
  int pace_ = 1;
  bool first_step_ = true;
  OFile hills_;

 public:
  explicit MetaDHIP(const ActionOptions&);
  ~MetaDHIP() override;
  static void registerKeywords(Keywords& keys);
  void calculate() override;
  void update() override;
};

void MetaDHIP::registerKeywords(Keywords& keys) {
  Bias::registerKeywords(keys);
  keys.add("compulsory", "SIGMA", "the widths of the Gaussian hills");
  keys.add("compulsory", "PACE", "the frequency for hill addition");
  keys.add("compulsory", "FILE", "HILLS", "a file in which the list of added hills is stored");
  keys.add("compulsory", "HEIGHT", "the heights of the Gaussian hills");
  keys.add("optional", "FMT", "specify format for HILLS files");
  keys.add("optional", "BIASFACTOR", "use well tempered metadynamics and use this bias factor");
  keys.add("optional", "TEMP", "the system temperature - needed for well-tempered metadynamics");
  keys.add("compulsory", "GRID_MIN", "the lower bounds for the grid");
  keys.add("compulsory", "GRID_MAX", "the upper bounds for the grid");
  keys.add("compulsory", "GRID_BIN", "the number of bins for the grid");
  keys.add("compulsory", "DEVICE", "0", "HIP device that holds the grid");
}

MetaDHIP::MetaDHIP(const ActionOptions& ao) : PLUMED_BIAS_INIT(ao) {
  ncv_ = getNumberOfArguments();
  if (ncv_ < 1 || ncv_ > 4) error("METAD_HIP handles 1 to 4 arguments");
  parseVector("SIGMA", sigma_);
  if (sigma_.size() != ncv_) error("number of arguments does not match number of SIGMA parameters");
  double height = 0.0;
  parse("HEIGHT", height);
  parse("PACE", pace_);
  if (pace_ <= 0) error("frequency for hill addition is nonsensical");
  parse("BIASFACTOR", biasf_);
  if (biasf_ < 1.0 && biasf_ != -1.0) error("well tempered bias factor is nonsensical");
  welltemp_ = biasf_ > 1.0;
  const double kbt = getkBT();
  if (welltemp_ && kbt <= 0.0) error("unless the MD engine passes the temperature to plumed, with well-tempered metad you must specify it using TEMP");
  std::vector<std::string> gmin(ncv_), gmax(ncv_);
  parseVector("GRID_MIN", gmin);
  parseVector("GRID_MAX", gmax);
  std::vector<unsigned> nbin(ncv_);
  parseVector("GRID_BIN", nbin);
  if (gmin.size() != ncv_ || gmax.size() != ncv_ || nbin.size() != ncv_)
    error("GRID_MIN, GRID_MAX and GRID_BIN need one entry per argument");
  std::string fname = "HILLS", fmt;
  parse("FILE", fname);
  parse("FMT", fmt);
  int device = 0;
  parse("DEVICE", device);
  checkRead();

  p2a_metad_spec spec{};
  spec.ncv = (int32_t)ncv_;
  for (unsigned i = 0; i < ncv_; i++) {
    Value* v = getPntrToArgument(i);
    spec.periodic[i] = v->isPeriodic();
    if (spec.periodic[i]) {
      // a periodic grid spans the domain of its argument (MetaD.cpp:1006-1060)
      v->getDomain(spec.gmin[i], spec.gmax[i]);
    } else {
      Tools::convert(gmin[i], spec.gmin[i]);
      Tools::convert(gmax[i], spec.gmax[i]);
    }
    spec.nbin[i] = nbin[i];
    spec.sigma[i] = sigma_[i];
  }
  spec.height = height;
  spec.biasf = biasf_;
  spec.kbt = kbt;
  spec.pace = pace_;
  const char* ns = std::getenv("PLUMED_DP2CUTOFF_NOSTRETCH");
  spec.nostretch = (ns && std::string(ns) != "0") ? 1 : 0;
  if (p2a_metad_create(&ctx_, device, &spec) != P2A_OK) error(p2a_last_error());

  hills_.link(*this);
  hills_.open(fname);
  if (!fmt.empty()) hills_.fmtField(fmt);
  hills_.addConstantField("multivariate");
  for (unsigned i = 0; i < ncv_; i++) hills_.setupPrintValue(getPntrToArgument(i));

  log.printf("  Gaussian width");
  for (unsigned i = 0; i < ncv_; i++) log.printf(" %f", sigma_[i]);
  log.printf("\n  Gaussian height %f, deposited every %d steps, grid of %lld nodes on the GPU (%s)\n", height, pace_,
             (long long)p2a_metad_grid_size(ctx_), p2a_version());
  if (welltemp_) log.printf("  Well-Tempered Bias Factor %f\n", biasf_);
  log.printf("  Hills file %s\n", fname.c_str());
}

MetaDHIP::~MetaDHIP() { p2a_metad_destroy(ctx_); }

void MetaDHIP::calculate() {
  double cv[4], der[4] = {0.0, 0.0, 0.0, 0.0}, ene = 0.0;
  for (unsigned i = 0; i < ncv_; i++) cv[i] = getArgument(i);
  // MetaD::calculate: no bias and no force with BIASFACTOR=1
  if (biasf_ != 1.0 && p2a_metad_bias(ctx_, 1, cv, &ene, der) != P2A_OK) error(p2a_last_error());
  setBias(ene);
  for (unsigned i = 0; i < ncv_; i++) setOutputForce((int)i, -der[i]);
}

void MetaDHIP::update() {
  // MetaD::update: a hill every PACE steps, never on the first step
  const bool add = (getStep() % pace_ == 0) && !first_step_;
  if (!add) {
    first_step_ = false;
    return;
  }
  double cv[4], h = 0.0;
  for (unsigned i = 0; i < ncv_; i++) cv[i] = getArgument(i);
  if (p2a_metad_deposit(ctx_, 1, cv, &h) != P2A_OK) error(p2a_last_error());
  // MetaD::writeGaussian: the height of a well-tempered hill is written
  // rescaled by biasf/(biasf-1)
  hills_.printField("time", getTimeStep() * getStep());
  for (unsigned i = 0; i < ncv_; i++) hills_.printField(getPntrToArgument(i), cv[i]);
  hills_.printField("multivariate", "false");
  for (unsigned i = 0; i < ncv_; i++) hills_.printField("sigma_" + getPntrToArgument(i)->getName(), sigma_[i]);
  hills_.printField("height", welltemp_ ? h * biasf_ / (biasf_ - 1.0) : h).printField("biasf", biasf_);
  hills_.printField();
  hills_.flush();
}

PLUMED_REGISTER_ACTION(MetaDHIP, "METAD_HIP")

namespace {
struct OverrideRegistration {
  std::unique_ptr<ActionRegistration<MetaDHIP>> reg;
  OverrideRegistration() {
    const char* e = std::getenv("PLUMED2AMD_OVERRIDE");
    if (e && std::string(e) == "1") reg = std::make_unique<ActionRegistration<MetaDHIP>>("METAD");
  }
} override_registration;
}  // namespace

}  // namespace p2a_plugin
}  // namespace PLMD
