// Model and data file parsing, restated from crates/lcurve/src/model.rs and
// crates/lcurve/src/types.rs:229-254.
#include "model.h"

#include <cerrno>
#include <cstdlib>
#include <fstream>
#include <map>
#include <sstream>

namespace lc {

namespace {

std::string trim(const std::string& s) {
  size_t b = s.find_first_not_of(" \t\r\n\f\v");
  if (b == std::string::npos) return "";
  size_t e = s.find_last_not_of(" \t\r\n\f\v");
  return s.substr(b, e - b + 1);
}

std::vector<std::string> split_ws(const std::string& s) {
  std::vector<std::string> out;
  std::istringstream is(s);
  std::string tok;
  while (is >> tok) out.push_back(tok);
  return out;
}

bool parse_f64(const std::string& s, double* out) {
  std::string t = trim(s);
  if (t.empty()) return false;
  errno = 0;
  char* end = nullptr;
  double v = std::strtod(t.c_str(), &end);
  if (end != t.c_str() + t.size()) return false;
  *out = v;
  return true;
}

double parse_double(const std::string& s) {
  double v;
  if (!parse_f64(s, &v)) throw LcError("parse error: bad double: '" + s + "'");
  return v;
}

int parse_int(const std::string& s) {
  std::string t = trim(s);
  char* end = nullptr;
  errno = 0;
  long v = std::strtol(t.c_str(), &end, 10);
  if (t.empty() || end != t.c_str() + t.size() || errno != 0 || v > 2147483647L ||
      v < -2147483648L)
    throw LcError("parse error: bad int: '" + s + "'");
  return static_cast<int>(v);
}

// model.rs:48-54
bool parse_bool(const std::string& s) {
  std::string t = trim(s);
  if (t == "1" || t == "true" || t == "yes") return true;
  if (t == "0" || t == "false" || t == "no") return false;
  throw LcError("parse error: bad bool: '" + s + "'");
}

// model.rs:21-38
Pparam pparam_from_str(const std::string& s) {
  auto parts = split_ws(s);
  if (parts.size() < 4) throw LcError("parse error: Pparam: too few fields in '" + s + "'");
  Pparam p;
  if (!parse_f64(parts[0], &p.value)) throw LcError("parse error: bad value: " + parts[0]);
  if (!parse_f64(parts[1], &p.range)) throw LcError("parse error: bad range: " + parts[1]);
  if (!parse_f64(parts[2], &p.dstep)) throw LcError("parse error: bad dstep: " + parts[2]);
  p.vary = parse_bool(parts[3]);
  p.defined = parts.size() > 4 ? parse_bool(parts[4]) : true;
  return p;
}

const std::vector<std::string>& comp_names() {
  static const std::vector<std::string> v = {
      "delta_phase", "nlat1f",   "nlat2f",       "nlat1c",       "nlat2c",   "npole",
      "nlatfill",    "nlngfill", "lfudge",       "llo",          "lhi",      "phase1",
      "phase2",      "wavelength", "roche1",     "roche2",       "eclipse1", "eclipse2",
      "glens1",      "use_radii", "tperiod",     "gdark_bolom1", "gdark_bolom2",
      "mucrit1",     "mucrit2",  "limb1",        "limb2",        "mirror",   "add_disc",
      "nrad",        "opaque",   "add_spot",     "nspot",        "iscale"};
  return v;
}

}  // namespace

const std::vector<std::string>& pparam_names() {
  static const std::vector<std::string> v = {
#define LCH_PNAME(n) #n,
      LCH_PPARAM_LIST(LCH_PNAME)
#undef LCH_PNAME
  };
  return v;
}

Model Model::from_file(const std::string& path) {
  std::ifstream f(path, std::ios::binary);
  if (!f) throw LcError("I/O error: cannot open '" + path + "'");
  std::stringstream ss;
  ss << f.rdbuf();
  return from_text(ss.str());
}

// crates/lcurve/src/model.rs:186-442
Model Model::from_text(const std::string& contents) {
  std::map<std::string, std::string> pv;
  std::map<std::string, bool> found;
  for (const auto& n : pparam_names()) found[n] = false;
  for (const auto& n : comp_names()) found[n] = false;

  std::istringstream in(contents);
  std::string line;
  size_t line_num = 0;
  while (std::getline(in, line)) {
    ++line_num;
    if (!line.empty() && line.back() == '\r') line.pop_back();
    std::string trimmed = trim(line);
    if (trimmed.empty() || trimmed[0] == '#') continue;
    size_t eq = trimmed.find('=');
    if (eq == std::string::npos) continue;
    std::string param = trim(trimmed.substr(0, eq));
    std::string value = trimmed.substr(eq + 1);
    // Strip comments: a '#' not preceded by a backslash.
    size_t nhash = 0;
    for (;;) {
      size_t pos = value.find('#', nhash);
      if (pos == std::string::npos) break;
      if (pos > 0 && value[pos - 1] == '\\') {
        nhash = pos + 1;
      } else {
        value.resize(pos);
        break;
      }
    }
    value = trim(value);
    if (!found.count(param))
      throw LcError("parse error: Parameter '" + param + "' on line " + std::to_string(line_num) +
                    " was not recognised");
    found[param] = true;
    pv[param] = value;
  }

  const char* optional_with_defaults[] = {"pdot", "third", "temp_edge", "absorb_edge"};
  for (const auto& kv : found) {
    if (kv.second) continue;
    bool opt = false;
    for (const char* o : optional_with_defaults) opt = opt || kv.first == o;
    if (opt) {
      pv[kv.first] = "0 1.e-10 1.e-10 0 0";
    } else if (kv.first.rfind("stsp", 0) != 0 && kv.first.rfind("uesp", 0) != 0) {
      throw LcError("parse error: Parameter '" + kv.first + "' was not defined");
    }
  }

  // Star-spot groups: all four or none (model.rs:283-298).
  const char* groups[5] = {"stsp11", "stsp12", "stsp13", "stsp21", "stsp22"};
  const char* fields[4] = {"_long", "_lat", "_fwhm", "_tcen"};
  for (const char* g : groups) {
    bool any = false, all = true;
    for (const char* fl : fields) {
      bool fnd = found[std::string(g) + fl];
      any = any || fnd;
      all = all && fnd;
    }
    if (any && !all)
      throw LcError(std::string("parse error: One or more of ") + g +
                    " parameters were not initialised");
  }
  {
    const char* ue[5] = {"uesp_long1", "uesp_long2", "uesp_lathw", "uesp_taper", "uesp_temp"};
    bool any = false, all = true;
    for (const char* u : ue) {
      any = any || found[u];
      all = all && found[u];
    }
    if (any && !all)
      throw LcError(
          "parse error: One or more uniform equatorial spot parameters were not initialised");
  }

  auto limb = [&](const char* key) {
    auto it = pv.find(key);
    if (it == pv.end()) throw LcError(std::string("parse error: ") + key + " not defined");
    if (it->second == "Poly") return LdcType::Poly;
    if (it->second == "Claret") return LdcType::Claret;
    throw LcError(std::string("parse error: bad ") + key + ": '" + it->second + "'");
  };

  Model m;
  m.limb1 = limb("limb1");
  m.limb2 = limb("limb2");
  m.roche1 = parse_bool(pv.at("roche1"));
  m.glens1 = parse_bool(pv.at("glens1"));
  if (m.glens1 && m.roche1)
    throw LcError("parse error: glens1=1 and roche1=1 are not simultaneously allowed");

#define LCH_READ(n)                                                   \
  {                                                                  \
    auto it = pv.find(#n);                                           \
    m.n = it != pv.end() ? pparam_from_str(it->second) : Pparam();  \
  }
  LCH_PPARAM_LIST(LCH_READ)
#undef LCH_READ

  m.delta_phase = parse_double(pv.at("delta_phase"));
  m.nlat1f = parse_int(pv.at("nlat1f"));
  m.nlat2f = parse_int(pv.at("nlat2f"));
  m.nlat1c = parse_int(pv.at("nlat1c"));
  m.nlat2c = parse_int(pv.at("nlat2c"));
  m.npole = parse_bool(pv.at("npole"));
  m.nlatfill = parse_int(pv.at("nlatfill"));
  m.nlngfill = parse_int(pv.at("nlngfill"));
  m.lfudge = parse_double(pv.at("lfudge"));
  m.llo = parse_double(pv.at("llo"));
  m.lhi = parse_double(pv.at("lhi"));
  m.phase1 = parse_double(pv.at("phase1"));
  m.phase2 = parse_double(pv.at("phase2"));
  m.wavelength = parse_double(pv.at("wavelength"));
  m.roche2 = parse_bool(pv.at("roche2"));
  m.eclipse1 = parse_bool(pv.at("eclipse1"));
  m.eclipse2 = parse_bool(pv.at("eclipse2"));
  m.use_radii = parse_bool(pv.at("use_radii"));
  m.tperiod = parse_double(pv.at("tperiod"));
  m.gdark_bolom1 = parse_bool(pv.at("gdark_bolom1"));
  m.gdark_bolom2 = parse_bool(pv.at("gdark_bolom2"));
  m.mucrit1 = parse_double(pv.at("mucrit1"));
  m.mucrit2 = parse_double(pv.at("mucrit2"));
  m.mirror = parse_bool(pv.at("mirror"));
  m.add_disc = parse_bool(pv.at("add_disc"));
  m.nrad = parse_int(pv.at("nrad"));
  m.opaque = parse_bool(pv.at("opaque"));
  m.add_spot = parse_bool(pv.at("add_spot"));
  m.nspot = parse_int(pv.at("nspot"));
  m.iscale = parse_bool(pv.at("iscale"));
  return m;
}

// crates/lcurve/src/model.rs:445-454
void Model::get_r1r2(double* r1o, double* r2o) const {
  if (use_radii) {
    *r1o = r1.value;
    *r2o = r2.value;
  } else {
    double sini = std::sin(to_radians(iangle.value));
    double c4 = sini * std::cos(2.0 * PI * cphi4.value);
    double c3 = sini * std::cos(2.0 * PI * cphi3.value);
    double r2pr1 = std::sqrt(1.0 - c4 * c4);
    double r2mr1 = std::sqrt(1.0 - c3 * c3);
    *r1o = (r2pr1 - r2mr1) / 2.0;
    *r2o = (r2pr1 + r2mr1) / 2.0;
  }
}

const Pparam* Model::pparam(const std::string& name) const {
#define LCH_FIND(n) \
  if (name == #n) return &n;
  LCH_PPARAM_LIST(LCH_FIND)
#undef LCH_FIND
  return nullptr;
}

Pparam* Model::pparam(const std::string& name) {
  return const_cast<Pparam*>(static_cast<const Model*>(this)->pparam(name));
}

bool Model::set_param_value(const std::string& name, double value) {
  Pparam* p = pparam(name);
  if (!p) return false;
  p->value = value;
  return true;
}

// crates/lcurve/src/model.rs:535-548
LDC Model::get_ldc1() const {
  LDC l;
  l.ldc1 = ldc1_1.value; l.ldc2 = ldc1_2.value; l.ldc3 = ldc1_3.value; l.ldc4 = ldc1_4.value;
  l.mucrit = mucrit1;
  l.ltype = limb1;
  return l;
}
LDC Model::get_ldc2() const {
  LDC l;
  l.ldc1 = ldc2_1.value; l.ldc2 = ldc2_2.value; l.ldc3 = ldc2_3.value; l.ldc4 = ldc2_4.value;
  l.mucrit = mucrit2;
  l.ltype = limb2;
  return l;
}

// crates/lcurve/src/types.rs:229-254
Data read_data(const std::string& path) {
  std::ifstream f(path);
  if (!f) throw LcError("I/O error: cannot open '" + path + "'");
  Data data;
  std::string line;
  while (std::getline(f, line)) {
    std::string t = trim(line);
    if (t.empty() || t[0] == '#') continue;
    auto parts = split_ws(t);
    if (parts.size() < 6) continue;
    Datum d;
    if (!parse_f64(parts[0], &d.time)) throw LcError("parse error: bad time: " + parts[0]);
    if (!parse_f64(parts[1], &d.expose)) throw LcError("parse error: bad expose: " + parts[1]);
    if (!parse_f64(parts[2], &d.flux)) throw LcError("parse error: bad flux: " + parts[2]);
    if (!parse_f64(parts[3], &d.ferr)) throw LcError("parse error: bad ferr: " + parts[3]);
    if (!parse_f64(parts[4], &d.weight)) throw LcError("parse error: bad weight: " + parts[4]);
    d.ndiv = parse_int(parts[5]);
    if (d.ferr < 0.0) {
      d.ferr = -d.ferr;
      d.weight = 0.0;
    }
    data.push_back(d);
  }
  return data;
}

}  // namespace lc
