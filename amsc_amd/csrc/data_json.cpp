// data_json.cpp -- DATA.json reader for amsc_sim_create_from_json.
//
// Schema and defaults of main/mpm_data.h:61-113 (required keys raise an
// error when absent, optional DT_FIXED/NSTEPS/ms_* take the reference's
// defaults).  A small recursive-descent parser: top-level object of numbers,
// arrays of numbers, strings, booleans or null; numbers are converted with
// strtod, which rounds correctly, so the %.17g values the reference's
// generators print round-trip exactly.
#include "amsc_mpm.h"

#include <cctype>
#include <cstdio>
#include <cstdlib>
#include <cstring>
#include <fstream>
#include <map>
#include <sstream>
#include <string>
#include <vector>

namespace amsc {
int fail(int code, const std::string &msg);
}

namespace {

struct Value {
  bool is_array = false;
  double num = 0.0;
  std::vector<double> arr;
  bool valid = false;  // number or array of numbers
};

struct Parser {
  const char *p, *end;
  std::string err;

  void ws() {
    while (p < end && std::isspace((unsigned char)*p)) ++p;
  }
  bool expect(char c) {
    ws();
    if (p < end && *p == c) {
      ++p;
      return true;
    }
    err = std::string("expected '") + c + "'";
    return false;
  }
  bool string(std::string &out) {
    ws();
    if (p >= end || *p != '"') {
      err = "expected string";
      return false;
    }
    ++p;
    out.clear();
    while (p < end && *p != '"') {
      if (*p == '\\' && p + 1 < end) ++p;
      out.push_back(*p++);
    }
    if (p >= end) {
      err = "unterminated string";
      return false;
    }
    ++p;
    return true;
  }
  bool number(double &v) {
    ws();
    char *e = nullptr;
    v = std::strtod(p, &e);
    if (e == p) {
      err = "expected number";
      return false;
    }
    p = e;
    return true;
  }
  bool skip_value() {
    ws();
    if (p >= end) return false;
    if (*p == '"') {
      std::string s;
      return string(s);
    }
    if (*p == '{' || *p == '[') {
      const char open = *p, close = open == '{' ? '}' : ']';
      int depth = 0;
      bool in_str = false;
      for (; p < end; ++p) {
        if (in_str) {
          if (*p == '\\') ++p;
          else if (*p == '"') in_str = false;
        } else if (*p == '"') in_str = true;
        else if (*p == open) ++depth;
        else if (*p == close && --depth == 0) {
          ++p;
          return true;
        }
      }
      err = "unterminated container";
      return false;
    }
    while (p < end && *p != ',' && *p != '}') ++p;
    return true;
  }
  bool value(Value &v) {
    ws();
    if (p >= end) {
      err = "unexpected end";
      return false;
    }
    if (*p == '[') {
      const char *save = p;
      ++p;
      v.is_array = true;
      ws();
      if (p < end && *p == ']') {
        ++p;
        v.valid = true;
        return true;
      }
      while (true) {
        double d;
        if (!number(d)) {
          p = save;
          v = Value();
          err.clear();
          return skip_value();
        }
        v.arr.push_back(d);
        ws();
        if (p < end && *p == ',') {
          ++p;
          continue;
        }
        if (p < end && *p == ']') {
          ++p;
          v.valid = true;
          return true;
        }
        err = "bad array";
        return false;
      }
    }
    if (*p == '-' || *p == '+' || *p == '.' || std::isdigit((unsigned char)*p)) {
      v.valid = number(v.num);
      return v.valid;
    }
    if (!std::strncmp(p, "true", 4)) {
      p += 4;
      v.num = 1.0;
      v.valid = true;
      return true;
    }
    if (!std::strncmp(p, "false", 5)) {
      p += 5;
      v.num = 0.0;
      v.valid = true;
      return true;
    }
    return skip_value();
  }
  bool object(std::map<std::string, Value> &out) {
    if (!expect('{')) return false;
    ws();
    if (p < end && *p == '}') {
      ++p;
      return true;
    }
    while (true) {
      std::string k;
      if (!string(k) || !expect(':')) return false;
      Value v;
      if (!value(v)) return false;
      out[k] = std::move(v);
      ws();
      if (p < end && *p == ',') {
        ++p;
        continue;
      }
      return expect('}');
    }
  }
};

}  // namespace

extern "C" int amsc_sim_create_from_json(const char *path, int max_steps,
                                         int capacity, amsc_sim **out) {
  if (!path || !out) return amsc::fail(AMSC_ERR_INVALID, "NULL argument");
  std::ifstream f(path, std::ios::binary);
  if (!f) return amsc::fail(AMSC_ERR_INVALID, std::string("cannot open ") + path);
  std::stringstream ss;
  ss << f.rdbuf();
  const std::string text = ss.str();
  Parser ps{text.data(), text.data() + text.size(), {}};
  std::map<std::string, Value> j;
  if (!ps.object(j))
    return amsc::fail(AMSC_ERR_INVALID, std::string("DATA.json: ") + ps.err);
  std::string missing;
  auto num = [&](const char *k, double *dst) {
    auto it = j.find(k);
    if (it == j.end() || !it->second.valid || it->second.is_array) {
      if (missing.empty()) missing = k;
      return;
    }
    *dst = it->second.num;
  };
  auto opt = [&](const char *k, double def) {
    auto it = j.find(k);
    return (it == j.end() || !it->second.valid || it->second.is_array)
               ? def
               : it->second.num;
  };
  auto arr = [&](const char *k) -> const std::vector<double> * {
    auto it = j.find(k);
    if (it == j.end() || !it->second.valid || !it->second.is_array) {
      if (missing.empty()) missing = k;
      return nullptr;
    }
    return &it->second.arr;
  };
  // required keys, mpm_data.h:64-97
  const char *pk[] = {"x", "y", "Mp", "Ap", "vpx", "vpy", "hp", "mom_px",
                      "mom_py", "Vp"};
  const std::vector<double> *pa[10];
  for (int i = 0; i < 10; ++i) pa[i] = arr(pk[i]);
  const std::vector<double> *Z = arr("Z");
  (void)arr("dZdx");
  (void)arr("dZdy");
  double Nex = 0, Ney = 0, BC = 0, MS = 0;
  amsc_params p;
  std::memset(&p, 0, sizeof(p));
  num("Nex", &Nex);
  num("Ney", &Ney);
  num("hx", &p.hx);
  num("hy", &p.hy);
  num("rho", &p.rho);
  num("g", &p.g);
  num("T", &p.T);
  num("xi", &p.xi);
  num("mu", &p.mu);
  num("phi", &p.phi);
  num("tauy", &p.tauy);
  num("BINGHAM_ON", &p.bingham_on);
  num("FRICTION_ON", &p.friction_on);
  num("CFL", &p.cfl);
  num("BC_FLAG", &BC);
  num("MERGE_SPLIT_ON", &MS);
  num("eq_level", &p.eq_level);
  if (!missing.empty())
    return amsc::fail(AMSC_ERR_INVALID,
                      "DATA.json: missing or malformed key '" + missing + "'");
  p.ncols = (int)Nex;
  p.nrows = (int)Ney;
  p.bc_flag = (int)BC;
  p.ms_on = (int)MS;
  // optional keys, mpm_data.h:99-112
  p.dt_fixed = opt("DT_FIXED", 0.0);
  p.nsteps = (int)opt("NSTEPS", 0);
  p.max_steps = max_steps > 0 ? max_steps : (int)opt("MAX_STEPS", 0);
  p.ms_alpha = opt("ms_alpha", 0.9);
  p.ms_beta = opt("ms_beta", 0.30);
  p.ms_split_hp_min = opt("ms_split_hp_min", 0.0);
  p.ms_shear_split = opt("ms_shear_split", 1e30);
  p.ms_hp_min = opt("ms_hp_min", 0.05);
  p.ms_max_dv = opt("ms_max_dv", 0.01);
  p.ms_min_level = (int)opt("ms_min_level", -2);
  p.ms_max_level = (int)opt("ms_max_level", 2);
  p.ms_min_ppc = (int)opt("ms_min_particles_per_cell", 2);
  p.ms_interval = (int)opt("ms_call_interval", 10);
  p.ms_max_ops = opt("ms_max_ops", 1e9);
  p.ms_mode = (int)opt("ms_mode", 0);
  p.ms_stretch = opt("ms_stretch_thresh", 1e7);
  const size_t np = pa[0]->size();
  for (int i = 1; i < 10; ++i)
    if (pa[i]->size() != np)
      return amsc::fail(AMSC_ERR_INVALID,
                        std::string("DATA.json: length of '") + pk[i] +
                            "' differs from x");
  if (Z->size() != (size_t)(p.ncols + 1) * (p.nrows + 1))
    return amsc::fail(AMSC_ERR_INVALID, "DATA.json: Z must have (Nex+1)(Ney+1) entries");
  return amsc_sim_create(&p, (int)np, pa[0]->data(), pa[1]->data(),
                         pa[2]->data(), pa[3]->data(), pa[4]->data(),
                         pa[5]->data(), pa[7]->data(), pa[8]->data(),
                         pa[6]->data(), pa[9]->data(), Z->data(), capacity,
                         out);
}
