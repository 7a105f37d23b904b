// params.json reader.  Same keys, defaults and derived values as get_params
// (srcd/set_params.cu:56-298); unknown keys are ignored as there.  The JSON
// reader is a small recursive-descent parser written for this file.
#include "cgfdm_internal.h"
#include <cctype>
#include <cstring>
#include <cmath>
#include <cstdlib>
#include <memory>

namespace cgfdm {

namespace {

struct JVal {
  enum Kind { Null, Bool, Num, Str, Arr, Obj } kind = Null;
  double num = 0;
  bool b = false;
  std::string str;
  std::vector<JVal> arr;
  std::vector<std::pair<std::string, JVal>> obj;
  const JVal* find(const std::string& k) const {
    for (auto& kv : obj) if (kv.first == k) return &kv.second;
    return nullptr;
  }
};

struct JParser {
  const std::string& s;
  size_t i = 0;
  std::string err;
  explicit JParser(const std::string& t) : s(t) {}
  void ws() { while (i < s.size() && std::isspace((unsigned char)s[i])) i++; }
  bool fail(const char* m) { if (err.empty()) err = std::string(m) + " at offset " + std::to_string(i); return false; }
  bool lit(const char* w) {
    size_t n = std::strlen(w);
    if (s.compare(i, n, w) == 0) { i += n; return true; }
    return false;
  }
  bool str(std::string& out) {
    if (i >= s.size() || s[i] != '"') return fail("expected string");
    i++;
    out.clear();
    while (i < s.size() && s[i] != '"') {
      char ch = s[i++];
      if (ch == '\\') {
        if (i >= s.size()) return fail("bad escape");
        char e = s[i++];
        switch (e) {
          case 'n': out += '\n'; break;
          case 't': out += '\t'; break;
          case 'r': out += '\r'; break;
          case 'b': out += '\b'; break;
          case 'f': out += '\f'; break;
          case 'u': i += 4; out += '?'; break;
          default: out += e;
        }
      } else {
        out += ch;
      }
    }
    if (i >= s.size()) return fail("unterminated string");
    i++;
    return true;
  }
  bool val(JVal& v) {
    ws();
    if (i >= s.size()) return fail("unexpected end");
    char ch = s[i];
    if (ch == '{') {
      i++;
      v.kind = JVal::Obj;
      ws();
      if (i < s.size() && s[i] == '}') { i++; return true; }
      while (true) {
        ws();
        std::string k;
        if (!str(k)) return false;
        ws();
        if (i >= s.size() || s[i] != ':') return fail("expected ':'");
        i++;
        JVal x;
        if (!val(x)) return false;
        v.obj.emplace_back(k, std::move(x));
        ws();
        if (i < s.size() && s[i] == ',') { i++; continue; }
        if (i < s.size() && s[i] == '}') { i++; return true; }
        return fail("expected ',' or '}'");
      }
    }
    if (ch == '[') {
      i++;
      v.kind = JVal::Arr;
      ws();
      if (i < s.size() && s[i] == ']') { i++; return true; }
      while (true) {
        JVal x;
        if (!val(x)) return false;
        v.arr.push_back(std::move(x));
        ws();
        if (i < s.size() && s[i] == ',') { i++; continue; }
        if (i < s.size() && s[i] == ']') { i++; return true; }
        return fail("expected ',' or ']'");
      }
    }
    if (ch == '"') { v.kind = JVal::Str; return str(v.str); }
    if (lit("true")) { v.kind = JVal::Bool; v.b = true; v.num = 1; return true; }
    if (lit("false")) { v.kind = JVal::Bool; v.b = false; v.num = 0; return true; }
    if (lit("null")) { v.kind = JVal::Null; return true; }
    const char* b = s.c_str() + i;
    char* e = nullptr;
    double d = std::strtod(b, &e);
    if (e == b) return fail("bad value");
    i += (size_t)(e - b);
    v.kind = JVal::Num;
    v.num = d;
    return true;
  }
};

}  // namespace

void Params::derive() {
  // set_params.cu:286-295; a single rank owns the whole domain
  ni = NX / PX; nj = NY / PY; nk = NZ / PZ;
  nx = ni + 6; ny = nj + 6; nz = nk + 6;
  NT = (int)((float)TMAX / (float)DT);  // real_t division in the reference
}

int Params::set(const std::string& k, double v) {
#define I_(n) if (k == #n) { n = (int)v; return 0; }
#define R_(n) if (k == #n) { n = v; return 0; }
  I_(NX) I_(NY) I_(NZ) I_(PX) I_(PY) I_(PZ) I_(igpu) R_(TMAX) R_(DT) R_(DH)
  I_(INPORT_GRID_TYPE) I_(INPORT_STRESS_TYPE) I_(Friction_type) I_(EXPORT_TIME_SKIP)
  I_(EXPORT_WAVE_SLICE_X) I_(EXPORT_WAVE_SLICE_Y) I_(EXPORT_WAVE_SLICE_Z)
  I_(PML_N) R_(PML_velocity) R_(PML_bmax) R_(PML_fc) I_(DAMP_N)
  R_(mu_s) R_(mu_d) R_(mu_0) R_(Dc) R_(C0) R_(smooth_load_T)
  R_(vs1) R_(vp1) R_(bi_vs1) R_(bi_vp1) R_(bi_rho1) R_(bi_vs2) R_(bi_vp2) R_(bi_rho2)
  R_(viscosity) R_(RupThres) I_(TP_n)
  I_(usePML) I_(useCerjan) I_(FreeSurface) I_(NormalRefine) I_(Rupture) I_(FaultSmooth)
  I_(MetricBF16) I_(MetricPalette) I_(MetricSkipConstant) I_(NewtonMode) R_(NewtonTol) R_(NewtonMLPGate) I_(NewtonFallbackIter) I_(WaveIChunk) I_(WaveBandRound) I_(WaveTallInterior) I_(WaveInteriorKernel) I_(WaveBandKernel) I_(WaveSurfaceKernel) I_(PmlAuxFused) I_(WaveBufferSkew) I_(StageStreams) I_(WaveSurfaceTiles) I_(WaveTopIChunk) I_(WaveBandIChunk)
  if (k == "RS_V0") { V0 = v; return 0; }
  if (k == "RS_Vini") { Vini = v; return 0; }
  if (k == "RS_f0") { f0 = v; return 0; }
  if (k == "RS_fw") { fw = v; return 0; }
  if (k == "RS_L") { L = v; return 0; }
  // "rho1" lands in bi_rho1 (set_params.cu:247-248)
  if (k == "rho1") { bi_rho1 = v; return 0; }
  if (k.size() == 11 && k.compare(0, 10, "Fault_grid") == 0 && k[10] >= '0' && k[10] <= '3') { Fault_grid[k[10] - '0'] = (int)v; return 0; }
  if (k.size() == 9 && k.compare(0, 8, "Asp_grid") == 0 && k[8] >= '0' && k[8] <= '3') { Asp_grid[k[8] - '0'] = (int)v; return 0; }
#undef I_
#undef R_
  return 1;
}

int Params::get(const std::string& k, double* v) const {
#define G_(n) if (k == #n) { *v = (double)n; return 0; }
  G_(NX) G_(NY) G_(NZ) G_(PX) G_(PY) G_(PZ) G_(igpu) G_(TMAX) G_(DT) G_(DH)
  G_(INPORT_GRID_TYPE) G_(INPORT_STRESS_TYPE) G_(Friction_type) G_(EXPORT_TIME_SKIP)
  G_(EXPORT_WAVE_SLICE_X) G_(EXPORT_WAVE_SLICE_Y) G_(EXPORT_WAVE_SLICE_Z)
  G_(PML_N) G_(PML_velocity) G_(PML_bmax) G_(PML_fc) G_(DAMP_N)
  G_(mu_s) G_(mu_d) G_(mu_0) G_(Dc) G_(C0) G_(smooth_load_T) G_(f0) G_(V0) G_(Vini) G_(fw) G_(L)
  G_(vs1) G_(vp1) G_(bi_vs1) G_(bi_vp1) G_(bi_rho1) G_(bi_vs2) G_(bi_vp2) G_(bi_rho2)
  G_(viscosity) G_(RupThres) G_(TP_n) G_(usePML) G_(useCerjan) G_(FreeSurface) G_(NormalRefine) G_(Rupture) G_(FaultSmooth)
  G_(MetricBF16) G_(MetricPalette) G_(MetricSkipConstant) G_(NewtonMode) G_(NewtonTol) G_(NewtonMLPGate) G_(NewtonFallbackIter) G_(WaveIChunk) G_(WaveBandRound) G_(WaveTallInterior) G_(WaveInteriorKernel) G_(WaveBandKernel) G_(WaveSurfaceKernel) G_(PmlAuxFused) G_(WaveBufferSkew) G_(StageStreams) G_(WaveSurfaceTiles) G_(WaveTopIChunk) G_(WaveBandIChunk)
  G_(ni) G_(nj) G_(nk) G_(nx) G_(ny) G_(nz) G_(NT)
#undef G_
  if (k.size() == 11 && k.compare(0, 10, "Fault_grid") == 0 && k[10] >= '0' && k[10] <= '3') { *v = Fault_grid[k[10] - '0']; return 0; }
  if (k.size() == 9 && k.compare(0, 8, "Asp_grid") == 0 && k[8] >= '0' && k[8] <= '3') { *v = Asp_grid[k[8] - '0']; return 0; }
  return 1;
}

int parse_params_json(const std::string& text, Params& p, std::string& err) {
  JParser ps(text);
  JVal root;
  if (!ps.val(root) || root.kind != JVal::Obj) {
    err = "Error at parsing json! " + ps.err;
    return 1;
  }
  for (auto& kv : root.obj) {
    const std::string& k = kv.first;
    const JVal& v = kv.second;
    if (v.kind == JVal::Str) {
      if (k == "OUT") p.OUT = v.str;
      else if (k == "Fault_geometry") p.Fault_geometry = v.str;
      else if (k == "Fault_init_stress") p.Fault_init_stress = v.str;
      else if (k == "Media1D") p.Media1D = v.str;
      else if (k == "Media3D") p.Media3D = v.str;
      else if (k == "NewtonMLP") p.NewtonMLP = v.str;
    } else if (v.kind == JVal::Arr) {
      if (k == "Fault_grid" || k == "Asp_grid") {
        int* dst = k == "Fault_grid" ? p.Fault_grid : p.Asp_grid;
        for (size_t q = 0; q < 4 && q < v.arr.size(); q++) dst[q] = (int)v.arr[q].num;
      }
    } else if (v.kind == JVal::Num || v.kind == JVal::Bool) {
      p.set(k, v.num);  // unknown numeric keys are ignored, as in the reference
    }
  }
  p.derive();
  return 0;
}

}  // namespace cgfdm
