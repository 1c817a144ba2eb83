// PLUMED plugin: METAD on a grid, deposited and evaluated by the plumed2_amd
// HIP library.
//
// Load into a PLUMED kernel with
//     LOAD FILE=libplumed2amd_plugin.so
// after which
//     m: METAD_HIP ARG=c,d SIGMA=20,0.05 HEIGHT=1.2 PACE=500 BIASFACTOR=10 TEMP=300 \
//        GRID_MIN=0,0 GRID_MAX=5000,6 GRID_BIN=512,512 [FILE=HILLS] [FMT=%f]
//        [GRID_WFILE=grid GRID_WSTRIDE=n] [GRID_RFILE=grid]
// (and, with PLUMED2AMD_OVERRIDE=1 at load time, plain METAD with these keywords)
// keeps the bias on the GPU grid.  Keywords, the first-step rule, the
// BIASFACTOR=1 rule and the HILLS records follow the reference action
// (src/bias/MetaD.cpp: registerKeywords :640-700, calculate :2315-2375,
// update :2377-2520, writeGaussian :1831-1884, grid read :1450-1530);
// diagonal hills on a grid only.  Grid files go through PLUMED's own Grid
// class, filled from and into the device grid (same flat node order).  Host C++ only, compiled against an installed PLUMED and linked with
// libplumed2amd.so.
#include "bias/Bias.h"
#include "core/ActionRegister.h"
#include "tools/Grid.h"
#include "tools/IFile.h"
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
  double biasf_ = -1.0;  // BIASFACTOR absent: plain METAD (MetaD.cpp:735)
  bool welltemp_ = false;
  int pace_ = 1;
  bool first_step_ = true;
  OFile hills_;
  int wgridstride_ = 0;
  std::unique_ptr<GridBase> wgrid_;  // host copy of the device grid for GRID_WFILE
  OFile gridfile_;

  void dumpGrid();

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
  keys.add("optional", "GRID_WSTRIDE", "write the grid to a file every N steps");
  keys.add("optional", "GRID_WFILE", "the file on which to write the grid");
  keys.add("optional", "GRID_RFILE", "a grid file from which the bias should be read at the initial step of the simulation");
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
  // well-tempered from BIASFACTOR=1 up (MetaD.cpp:860-865); BIASFACTOR=1 then
  // applies no bias and no force, as in calculate() below
  welltemp_ = biasf_ >= 1.0;
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
  std::string wgridname, rgridname;
  parse("GRID_WSTRIDE", wgridstride_);
  parse("GRID_WFILE", wgridname);
  parse("GRID_RFILE", rgridname);
  if (!wgridname.empty() && wgridstride_ == 0) error("frequency with which to output grid not specified use GRID_WSTRIDE");
  if (wgridstride_ > 0 && wgridname.empty()) error("grid filename not specified use GRID_WFILE");
  checkRead();

  // GRID_RFILE: the grid, its bounds and bins come from the file
  std::unique_ptr<GridBase> rgrid;
  if (!rgridname.empty()) {
    IFile gf;
    gf.link(*this);
    if (!gf.FileExist(rgridname)) error("The GRID file you want to read: " + rgridname + ", cannot be found!");
    gf.open(rgridname);
    rgrid = GridBase::create(getLabel() + ".bias", getArguments(), gf, false, false, true);
    if (rgrid->getDimension() != ncv_) error("mismatch between dimensionality of input grid and number of arguments");
    for (unsigned i = 0; i < ncv_; i++)
      if (getPntrToArgument(i)->isPeriodic() != rgrid->getIsPeriodic()[i])
        error("periodicity mismatch between arguments and input bias");
    gmin = rgrid->getMin();
    gmax = rgrid->getMax();
    const std::vector<unsigned> nodes = rgrid->getNbin();
    for (unsigned i = 0; i < ncv_; i++) nbin[i] = rgrid->getIsPeriodic()[i] ? nodes[i] : nodes[i] - 1;
  }

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

  if (rgrid) {
    const size_t n = (size_t)p2a_metad_grid_size(ctx_);
    if (n != (size_t)rgrid->getSize()) error("grid read from " + rgridname + " does not match the device grid");
    std::vector<double> v(n), d(n * ncv_), der(ncv_);
    for (size_t i = 0; i < n; i++) {
      v[i] = rgrid->getValueAndDerivatives((GridBase::index_t)i, der);
      for (unsigned j = 0; j < ncv_; j++) d[i * ncv_ + j] = der[j];
    }
    if (p2a_metad_set_grid(ctx_, v.data(), d.data()) != P2A_OK) error(p2a_last_error());
    log.printf("  Restarting from %s\n", rgridname.c_str());
  }
  if (wgridstride_ > 0) {
    wgrid_ = std::make_unique<Grid>(getLabel() + ".bias", getArguments(), gmin, gmax, nbin, false, true);
    if ((size_t)wgrid_->getSize() != (size_t)p2a_metad_grid_size(ctx_)) error("grid size mismatch for GRID_WFILE");
    gridfile_.link(*this);
    gridfile_.open(wgridname);
    log.printf("  Grid is written on file %s with stride %d\n", wgridname.c_str(), wgridstride_);
  }

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

// MetaD::update's grid dump (MetaD.cpp:2489-2520): the file is rewound, the
// previous content kept as bck.last.<name>
void MetaDHIP::dumpGrid() {
  const size_t n = (size_t)p2a_metad_grid_size(ctx_);
  std::vector<double> v(n), d(n * ncv_), der(ncv_);
  if (p2a_metad_get_grid(ctx_, v.data(), d.data()) != P2A_OK) error(p2a_last_error());
  for (size_t i = 0; i < n; i++) {
    for (unsigned j = 0; j < ncv_; j++) der[j] = d[i * ncv_ + j];
    wgrid_->setValueAndDerivatives((GridBase::index_t)i, v[i], der);
  }
  gridfile_.rewind();
  wgrid_->writeToFile(gridfile_);
  gridfile_.flush();
}

void MetaDHIP::update() {
  // MetaD::update: a hill every PACE steps, never on the first step; the
  // grid is dumped after the step's hill, first step included
  const bool add = (getStep() % pace_ == 0) && !first_step_;
  first_step_ = false;
  if (wgridstride_ > 0 && getStep() % wgridstride_ == 0 && !add) dumpGrid();
  if (!add) return;
  double cv[4], h = 0.0;
  for (unsigned i = 0; i < ncv_; i++) cv[i] = getArgument(i);
  if (p2a_metad_deposit(ctx_, 1, cv, &h) != P2A_OK) error(p2a_last_error());
  // MetaD::writeGaussian: the height of a well-tempered hill is written
  // rescaled by biasf/(biasf-1)
  hills_.printField("time", getTimeStep() * getStep());
  for (unsigned i = 0; i < ncv_; i++) hills_.printField(getPntrToArgument(i), cv[i]);
  hills_.printField("multivariate", "false");
  for (unsigned i = 0; i < ncv_; i++) hills_.printField("sigma_" + getPntrToArgument(i)->getName(), sigma_[i]);
  hills_.printField("height", (welltemp_ && biasf_ > 1.0) ? h * biasf_ / (biasf_ - 1.0) : h).printField("biasf", biasf_);
  hills_.printField();
  hills_.flush();
  if (wgridstride_ > 0 && getStep() % wgridstride_ == 0) dumpGrid();
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
