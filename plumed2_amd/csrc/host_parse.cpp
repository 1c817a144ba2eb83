// Host-side construction of the COORDINATION action: keyword parsing of the
// input line and resolution of the switching function.  No GPU calls here.
//
//   p2a_switch_parse     <- SwitchingFunction::set(std::string)   SwitchingFunction.cpp:1055-1159
//   p2a_switch_rational  <- SwitchingFunction::set(nn,mm,r0,d0)    SwitchingFunction.cpp:1176-1184
//   p2a_coord_parse      <- Coordination::Coordination            Coordination.cpp:128-157
//                           CoordinationBase::CoordinationBase    CoordinationBase.cpp:42-101
#include <algorithm>
#include <cerrno>
#include <cfloat>
#include <cmath>
#include <cstdio>
#include <cstdlib>
#include <cstring>
#include <string>
#include <vector>

#include "p2a_internal.h"
#include "switch_eval.h"

namespace p2a {

thread_local std::string g_last_error;

int fail(int code, const std::string& msg) {
  g_last_error = msg;
  return code;
}

// Tools::getWords with '{' grouping (Tools.cpp:201-270): split on blanks,
// keep blanks inside {...}, drop the outermost braces.
static bool get_words(const std::string& line, std::vector<std::string>& words, std::string& err) {
  words.clear();
  std::string w;
  int level = 0;
  for (char ch : line) {
    if (ch == '}') {
      level--;
      if (level < 0) {
        err = "Extra closed parenthesis in '" + line + "'";
        return false;
      }
      if (level == 0) continue;
    }
    bool sep = (ch == ' ' || ch == '\t' || ch == '\n' || ch == '\r');
    if (level == 0 && sep) {
      if (!w.empty()) words.push_back(w);
      w.clear();
    } else if (ch == '{') {
      if (level > 0) w.push_back(ch);
      level++;
    } else {
      w.push_back(ch);
    }
  }
  if (level != 0) {
    err = "Unmatched parenthesis in '" + line + "'";
    return false;
  }
  if (!w.empty()) words.push_back(w);
  return true;
}

// Tools::findKeyword / Tools::parse / Tools::parseFlag on a word list
static bool find_keyword(const std::vector<std::string>& words, const std::string& key) {
  for (auto& w : words)
    if (w.compare(0, key.size() + 1, key + "=") == 0) return true;
  return false;
}

static bool take_value(std::vector<std::string>& words, const std::string& key, std::string& value) {
  for (size_t i = 0; i < words.size(); i++) {
    if (words[i].compare(0, key.size() + 1, key + "=") == 0) {
      value = words[i].substr(key.size() + 1);
      words.erase(words.begin() + i);
      return true;
    }
  }
  return false;
}

static bool take_flag(std::vector<std::string>& words, const std::string& key) {
  for (size_t i = 0; i < words.size(); i++) {
    if (words[i] == key) {
      words.erase(words.begin() + i);
      return true;
    }
  }
  return false;
}

static bool to_double(const std::string& s, double& v) {
  if (s.empty()) return false;
  const char* p = s.c_str();
  char* end = nullptr;
  errno = 0;
  double x = std::strtod(p, &end);
  if (end == p || *end != '\0' || errno == ERANGE) {
    if (s == "pi") { v = 3.141592653589793238462643383279502884197169399375105820974944592307; return true; }
    return false;
  }
  v = x;
  return true;
}

static bool to_int(const std::string& s, int& v) {
  double d;
  if (!to_double(s, d)) return false;
  if (d != std::floor(d) || std::fabs(d) > 2147483647.0) return false;
  v = (int)d;
  return true;
}

// Data::init (SwitchingFunction.cpp:84-94)
static void data_init(p2a_switch& s, double d0, double dmax, double r0) {
  s.d0 = d0;
  s.dmax = dmax;
  s.dmax_2 = (dmax < std::sqrt(DBL_MAX)) ? dmax * dmax : DBL_MAX;
  s.invr0 = 1.0 / r0;
  s.invr0_2 = s.invr0 * s.invr0;
  s.stretch = 1.0;
  s.shift = 0.0;
}

// SwitchInterface::setupStretch (SwitchingFunction.cpp:63-72)
static void setup_stretch(p2a_switch& s) {
  if (s.dmax < DBL_MAX) {
    double dd;
    s.stretch = 1.0;
    s.shift = 0.0;
    double s0 = switch_calc(s, 0.0, dd);
    double sd = switch_calc(s, s.dmax, dd);
    s.stretch = 1.0 / (s0 - sd);
    s.shift = -sd * s.stretch;
  }
}

// rationalFactory (SwitchingFunction.cpp:307-347) and rational::init (:226-251)
static void rational_init(p2a_switch& s, double d0, double dmax, double r0, int N, int M) {
  std::memset(&s, 0, sizeof(s));
  data_init(s, d0, dmax, r0);
  bool fast = (N % 2 == 0) && (M % 2 == 0) && d0 == 0.0;
  if (((2 * N) == M || M == 0) && fast && N > 0 && N <= 12) {
    s.type = P2A_SW_RATIONAL_FIXED;
    s.nn = N;
    s.mm = 2 * N;
    s.nnf = N / 2;
    s.mmf = N;
    s.fast = 1;
    s.simplified = 1;
    return;
  }
  s.type = P2A_SW_RATIONAL;
  s.nn = N;
  s.mm = (M == 0) ? N * 2 : M;
  s.preRes = (double)s.nn / s.mm;
  s.preDfunc = 0.5 * s.nn * (s.nn - s.mm) / (double)s.mm;
  s.preSecDev = (s.nn * (s.mm * s.mm - 3.0 * s.mm * (-1 + s.nn) + s.nn * (-3 + 2 * s.nn))) / (6.0 * s.mm);
  s.nnf = s.nn / 2;
  s.mmf = s.mm / 2;
  s.preDfuncF = 0.5 * s.nnf * (s.nnf - s.mmf) / (double)s.mmf;
  s.preSecDevF = (s.nnf * (s.mmf * s.mmf - 3.0 * s.mmf * (-1 + s.nnf) + s.nnf * (-3 + 2 * s.nnf))) / (6.0 * s.mmf);
  s.fast = fast;
  s.simplified = (2 * N == M || M == 0);
}

int switch_parse(const std::string& definition, p2a_switch& out, std::string& errormsg) {
  std::vector<std::string> data;
  if (!get_words(definition, data, errormsg)) return P2A_ERR_ARG;
  if (data.empty()) {
    errormsg = "missing all input for switching function";
    return P2A_ERR_ARG;
  }
  std::string name = data[0];
  data.erase(data.begin());
  double d0 = 0.0, dmax = DBL_MAX;
  std::string v;
  auto check_parse_d = [&](const char* key, double& var) {
    if (find_keyword(data, key)) {
      if (!take_value(data, key, v) || !to_double(v, var)) errormsg = std::string("could not parse ") + key;
    }
  };
  auto check_parse_i = [&](const char* key, int& var) {
    if (find_keyword(data, key)) {
      if (!take_value(data, key, v) || !to_int(v, var)) errormsg = std::string("could not parse ") + key;
    }
  };
  auto required_d = [&](const char* key, double& var) {
    if (!take_value(data, key, v) || !to_double(v, var)) errormsg = std::string(key) + " is required for " + name;
  };
  auto required_i = [&](const char* key, int& var) {
    if (!take_value(data, key, v) || !to_int(v, var)) errormsg = std::string(key) + " is required for " + name;
  };
  check_parse_d("D_0", d0);
  check_parse_d("D_MAX", dmax);
  take_flag(data, "STRETCH");  // accepted and ignored (SwitchingFunction.cpp:1078-1079)
  bool dostretch = true;
  if (take_flag(data, "NOSTRETCH")) dostretch = false;

  p2a_switch s;
  std::memset(&s, 0, sizeof(s));
  if (name == "CUBIC") {
    data_init(s, d0, dmax, dmax - d0);  // cubicSwitch::init (:459-462)
    s.type = P2A_SW_CUBIC;
  } else {
    double r0 = 0.0;
    required_d("R_0", r0);
    if (name == "RATIONAL") {
      int nn = 6, mm = 0;
      check_parse_i("NN", nn);
      check_parse_i("MM", mm);
      rational_init(s, d0, dmax, r0, nn, mm);
    } else if (name == "SMAP") {
      int a = 0, b = 0;
      required_i("A", a);
      required_i("B", b);
      data_init(s, d0, dmax, r0);
      s.type = P2A_SW_SMAP;
      s.a = a;
      s.b = b;
      s.c = std::pow(2., (double)a / (double)b) - 1.0;
      s.d = -(double)b / (double)a;
    } else if (name == "Q") {
      double beta = 50.0, lambda = 1.8, ref = 0.0;
      check_parse_d("BETA", beta);
      check_parse_d("LAMBDA", lambda);
      required_d("REF", ref);
      data_init(s, d0, dmax, r0);
      s.type = P2A_SW_Q;
      s.beta = beta;
      s.lambda = lambda;
      s.ref = ref;
    } else if (name == "EXP") {
      data_init(s, d0, dmax, r0);
      s.type = P2A_SW_EXP;
    } else if (name == "GAUSSIAN") {
      if (r0 == 1.0 && d0 == 0.0) {
        data_init(s, 0.0, dmax, 1.0);
        s.type = P2A_SW_FASTGAUSSIAN;
      } else {
        data_init(s, d0, dmax, r0);
        s.type = P2A_SW_GAUSSIAN;
      }
    } else if (name == "TANH") {
      data_init(s, d0, dmax, r0);
      s.type = P2A_SW_TANH;
    } else if (name == "COSINUS") {
      data_init(s, d0, dmax, r0);
      s.type = P2A_SW_COSINUS;
    } else if (name == "MATHEVAL" || name == "CUSTOM") {
      errormsg = "switching function type '" + name + "' (lepton expressions) is not implemented on the GPU";
      return P2A_ERR_UNSUPPORTED;
    } else {
      errormsg = "cannot understand switching function type '" + name + "'";
    }
  }
  if (!data.empty()) {
    errormsg = "found the following rogue keywords in switching function input : ";
    for (auto& w : data) errormsg += w + " ";
  }
  if (!errormsg.empty()) return P2A_ERR_ARG;
  if (dostretch && dmax != DBL_MAX) setup_stretch(s);
  out = s;
  return P2A_OK;
}

int switch_rational(int nn, int mm, double r0, double d0, p2a_switch& out) {
  if (mm == 0) mm = 2 * nn;
  double dmax = d0 + r0 * std::pow(0.00001, 1. / (nn - mm));
  rational_init(out, d0, dmax, r0, nn, mm);
  setup_stretch(out);
  return P2A_OK;
}

// GROUPA/GROUPB atom lists: serial numbers, ranges a-b and a-b:stride,
// comma separated (the numeric forms accepted by Tools::interpretRanges).
static bool parse_atoms(const std::string& value, std::vector<int64_t>& out, std::string& err) {
  size_t pos = 0;
  while (pos <= value.size()) {
    size_t c = value.find(',', pos);
    std::string item = value.substr(pos, c == std::string::npos ? std::string::npos : c - pos);
    pos = (c == std::string::npos) ? value.size() + 1 : c + 1;
    if (item.empty()) continue;
    long long a, b, st = 1;
    char tail;
    if (std::sscanf(item.c_str(), "%lld-%lld:%lld%c", &a, &b, &st, &tail) == 3 ||
        (st = 1, std::sscanf(item.c_str(), "%lld-%lld%c", &a, &b, &tail) == 2)) {
      if (st <= 0 || b < a || a < 1) {
        err = "could not interpret atom range " + item;
        return false;
      }
      for (long long k = a; k <= b; k += st) out.push_back(k - 1);
    } else if (std::sscanf(item.c_str(), "%lld%c", &a, &tail) == 1) {
      if (a < 1) {
        err = "atom serial numbers start from 1: " + item;
        return false;
      }
      out.push_back(a - 1);
    } else {
      err = "cannot interpret atom list item '" + item + "' (only serial numbers and ranges are supported)";
      return false;
    }
  }
  return true;
}

int coord_parse(const std::string& line, p2a_coord_spec& spec, std::string& err) {
  std::memset(&spec, 0, sizeof(spec));
  std::vector<std::string> words;
  if (!get_words(line, words, err)) return P2A_ERR_ARG;
  // optional "label:" and action name
  if (!words.empty() && words[0].size() > 1 && words[0].back() == ':') words.erase(words.begin());
  if (!words.empty() && words[0] == "COORDINATION") words.erase(words.begin());
  for (auto& w : words)
    if (w.compare(0, 6, "LABEL=") == 0) {
      take_value(words, "LABEL", err);
      err.clear();
      break;
    }
  std::string v;
  std::vector<int64_t> ga, gb;
  spec.serial = take_flag(words, "SERIAL");
  if (take_value(words, "GROUPA", v) && !parse_atoms(v, ga, err)) return P2A_ERR_ARG;
  if (take_value(words, "GROUPB", v) && !parse_atoms(v, gb, err)) return P2A_ERR_ARG;
  bool nopbc = take_flag(words, "NOPBC");
  spec.pbc = !nopbc;
  spec.pair = take_flag(words, "PAIR");
  bool classic = take_flag(words, "NLIST");
  bool cells = take_flag(words, "NLISTCELLS");
  if (classic && cells) {
    err = "Please activate only one of the two version of the NL";
    return P2A_ERR_ARG;
  }
  if (cells && spec.pair) {
    err = "Pair is not compatible with the CELLS implementation of the NL";
    return P2A_ERR_ARG;
  }
  spec.nl_mode = cells ? P2A_NL_CELLS : (classic ? P2A_NL_CLASSIC : P2A_NL_NONE);
  if (spec.nl_mode != P2A_NL_NONE) {
    double cut = 0.0;
    if (take_value(words, "NL_CUTOFF", v) && !to_double(v, cut)) cut = 0.0;
    if (cut <= 0.0) {
      err = "NL_CUTOFF should be explicitly specified and positive";
      return P2A_ERR_ARG;
    }
    int st = 0;
    if (take_value(words, "NL_STRIDE", v) && !to_int(v, st)) st = 0;
    if (st <= 0) {
      err = "NL_STRIDE should be explicitly specified and positive";
      return P2A_ERR_ARG;
    }
    spec.nl_cutoff = cut;
    spec.nl_stride = st;
  }
  if (ga.empty()) {
    err = "GROUPA should contain at least one atom";
    return P2A_ERR_ARG;
  }
  if (spec.pair && ga.size() != gb.size()) {
    err = "when using PAIR the two groups should have the same number of atoms";
    return P2A_ERR_ARG;
  }
  // switching function: SWITCH= or NN/MM/D_0/R_0 (Coordination.cpp:128-152)
  std::string sw;
  if (take_value(words, "SWITCH", sw)) {
    std::string errors;
    int rc = switch_parse(sw, spec.sw, errors);
    if (rc != P2A_OK) {
      err = "problem reading SWITCH keyword : " + errors;
      return rc;
    }
  } else {
    int nn = 6, mm = 0;
    double d0 = 0.0, r0 = 0.0;
    if (take_value(words, "R_0", v) && !to_double(v, r0)) r0 = 0.0;
    if (r0 <= 0.0) {
      err = "R_0 should be explicitly specified and positive";
      return P2A_ERR_ARG;
    }
    if (take_value(words, "D_0", v) && !to_double(v, d0)) {
      err = "could not parse D_0";
      return P2A_ERR_ARG;
    }
    if (take_value(words, "NN", v) && !to_int(v, nn)) {
      err = "could not parse NN";
      return P2A_ERR_ARG;
    }
    if (take_value(words, "MM", v) && !to_int(v, mm)) {
      err = "could not parse MM";
      return P2A_ERR_ARG;
    }
    switch_rational(nn, mm, r0, d0, spec.sw);
  }
  // checkRead (Action.cpp:286-294)
  if (!words.empty()) {
    err = "cannot understand the following words from the input line : ";
    for (size_t i = 0; i < words.size(); i++) err += (i ? ", " : "") + words[i];
    return P2A_ERR_ARG;
  }
  spec.na = (int64_t)ga.size();
  spec.nb = (int64_t)gb.size();
  spec.groupa = (int64_t*)std::malloc(sizeof(int64_t) * std::max<size_t>(1, ga.size()));
  spec.groupb = (int64_t*)std::malloc(sizeof(int64_t) * std::max<size_t>(1, gb.size()));
  std::copy(ga.begin(), ga.end(), spec.groupa);
  std::copy(gb.begin(), gb.end(), spec.groupb);
  return P2A_OK;
}

std::string switch_description(const p2a_switch& s) {
  char buf[256];
  const char* names[] = {"rational", "rational", "exponential", "gaussian", "fastgaussian", "smap",
                         "cubic", "tanh", "cosinus", "nativeq"};
  int t = (s.type >= 0 && s.type <= 9) ? s.type : 0;
  std::snprintf(buf, sizeof buf, "%s switching function with parameters d0=%g r0=%g", names[t], s.d0, 1.0 / s.invr0);
  std::string out(buf);
  if (s.type == P2A_SW_RATIONAL || s.type == P2A_SW_RATIONAL_FIXED) {
    std::snprintf(buf, sizeof buf, " nn=%d mm=%d", s.nn, s.mm);
    out += buf;
  }
  if (s.dmax < DBL_MAX) {
    std::snprintf(buf, sizeof buf, " dmax=%g", s.dmax);
    out += buf;
  }
  return out;
}

}  // namespace p2a

using namespace p2a;

extern "C" {

const char* p2a_last_error(void) { return g_last_error.c_str(); }
const char* p2a_version(void) { return "plumed2_amd 0.1.0 (gfx950)"; }

int p2a_switch_parse(const char* definition, p2a_switch* out) {
  if (!definition || !out) return fail(P2A_ERR_ARG, "null argument");
  std::string err;
  int rc = switch_parse(definition, *out, err);
  if (rc != P2A_OK) return fail(rc, err);
  g_last_error.clear();
  return P2A_OK;
}

int p2a_switch_rational(int nn, int mm, double r0, double d0, p2a_switch* out) {
  if (!out) return fail(P2A_ERR_ARG, "null argument");
  if (r0 <= 0.0) return fail(P2A_ERR_ARG, "R_0 should be explicitly specified and positive");
  return switch_rational(nn, mm, r0, d0, *out);
}

int p2a_switch_describe(const p2a_switch* sw, char* buf, size_t len) {
  if (!sw || !buf || len == 0) return fail(P2A_ERR_ARG, "null argument");
  std::string d = switch_description(*sw);
  std::snprintf(buf, len, "%s", d.c_str());
  return P2A_OK;
}

int p2a_coord_parse(const char* line, p2a_coord_spec* out) {
  if (!line || !out) return fail(P2A_ERR_ARG, "null argument");
  std::string err;
  int rc = coord_parse(line, *out, err);
  if (rc != P2A_OK) {
    p2a_coord_spec_free(out);
    return fail(rc, err);
  }
  g_last_error.clear();
  return P2A_OK;
}

void p2a_coord_spec_free(p2a_coord_spec* spec) {
  if (!spec) return;
  std::free(spec->groupa);
  std::free(spec->groupb);
  spec->groupa = spec->groupb = nullptr;
  spec->na = spec->nb = 0;
}

}  // extern "C"
