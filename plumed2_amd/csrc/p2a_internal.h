// Internal declarations shared by the host and HIP translation units.
#pragma once
#include <cstdint>
#include <string>
#include "plumed2_amd.h"

namespace p2a {

extern thread_local std::string g_last_error;
int fail(int code, const std::string& msg);

int switch_parse(const std::string& definition, p2a_switch& out, std::string& err);
int switch_rational(int nn, int mm, double r0, double d0, p2a_switch& out);
int coord_parse(const std::string& line, p2a_coord_spec& spec, std::string& err);
std::string switch_description(const p2a_switch& s);

// Periodic cell as Pbc::setBox prepares it (Pbc.cpp:165-212): type, inverse,
// reduced lattice and the per-octant shift lists of Pbc::buildShifts.
constexpr int kMaxShift = 8;
struct PbcHost {
  int type = 0;  // 0 unset, 1 orthorhombic, 2 generic
  double box[9] = {0}, inv[9] = {0}, reduced[9] = {0}, invReduced[9] = {0};
  int nshift[8] = {0};
  double shifts[8][kMaxShift][3] = {};
};
void pbc_setup(const double* box9, PbcHost& p);
// perpendicular widths of the cell (1/|reciprocal row|)
void pbc_widths(const double* box9, double w[3]);

}  // namespace p2a
