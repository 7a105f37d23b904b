// C ABI (include/cgfdm.h) over the HIP kernels: context lifetime, the
// stage/step driver in the order of the reference time loop
// (srcd/main.cu:420-535), field access and output files.
#include "../../include/cgfdm.h"
#include "cgfdm_internal.h"
#include "mlp.h"
#include "netcdf_classic.h"
#include <cmath>
#include <cstdio>
#include <cstring>
#include <fstream>
#include <sstream>
#include <sys/stat.h>

using namespace cgfdm;

struct cgfdm_ctx {
  Ctx c;
};

namespace {

thread_local std::string g_err = "";

int fail(int code, const std::string& msg) {
  g_err = msg;
  return code;
}

// FD operator direction table, main.cu:86-94; negated after every stage.
const int FD_FLAGS[8][3] = {{-1, -1, -1}, {1, 1, -1}, {1, 1, 1}, {-1, -1, 1}, {-1, 1, -1}, {1, -1, -1}, {1, -1, 1}, {-1, 1, 1}};

void stage_flags(int it, int irk, int& fx, int& fy, int& fz) {
  int s = (irk % 2) ? -1 : 1;
  fx = FD_FLAGS[it % 8][0] * s;
  fy = FD_FLAGS[it % 8][1] * s;
  fz = FD_FLAGS[it % 8][2] * s;
}

// Stage buffers: S = step start (RK mW), A/B = stage values, T = RK
// accumulator (tW).  Stage 0: S->A, 1: A->B, 2: B->A, 3: A->S.
StageBufs bufs_for(Ctx& c, int irk) {
  StageBufs b;
  float* in[4] = {c.S, c.A, c.B, c.A};
  float* out[4] = {c.A, c.B, c.A, c.S};
  b.Win = in[irk];
  b.Wout = out[irk];
  b.S = c.S;
  b.T = c.Tacc;
  return b;
}

int check_hip(const char* what) {
  hipError_t e = hipGetLastError();
  if (e != hipSuccess) return fail(CGFDM_EHIP, std::string(what) + ": " + hipGetErrorString(e));
  return 0;
}

struct FieldRef {
  void* ptr = nullptr;
  size_t count = 0;
  int dtype = 0;   // 0 f32, 1 f64, 2 i32, 3 bf16
  int nf = 0;      // >0: device volume array, nf fields per plane, padded rows (count is the caller's size)
  int planes = 0;  // planes of such an array
  int f0 = 0, fn = 0;  // fn > 0: only the fields f0 .. f0+fn-1 of each plane are exchanged
  bool host = false;   // a host-side value (read only)
};

FieldRef field(Ctx& c, const std::string& n) {
  FieldRef r;
  const long nxyz = c.d.nxyz, nyz = c.d.nyz, n2 = (long)c.d.nj * c.d.nk;
  FaultDev& F = c.f;
  auto f32 = [&](void* p, size_t cnt) { r.ptr = p; r.count = cnt; r.dtype = 0; };
  auto f64 = [&](void* p, size_t cnt) { r.ptr = p; r.count = cnt; r.dtype = 1; };
  auto i32 = [&](void* p, size_t cnt) { r.ptr = p; r.count = cnt; r.dtype = 2; };
  const int nx = c.d.nx;
  if (n == "W") { f32(c.cur, 9 * nxyz); r.nf = 9; r.planes = nx; }
  else if (n == "hW") { f32(c.hW, 9 * nxyz); r.nf = 9; r.planes = nx; }
  else if (n == "M") { f32(c.p.MetricBF16 ? nullptr : c.met, 13 * nxyz); r.nf = 13; r.planes = nx; }
  else if (n == "M_bf16" && c.p.MetricBF16 == 2) {  // 5 words: the 9 metric deviations from M_ref and the palette entry
    r.ptr = c.met; r.count = 5 * nxyz; r.dtype = 2; r.nf = 9; r.planes = nx; r.fn = 5;
  }
  else if (n == "M_jac_bf16") {  // MetricBF16 = 2: the Jacobian's deviation from M_ref[9] (low half)
    r.ptr = c.p.MetricBF16 == 2 ? c.met : nullptr; r.count = nxyz; r.dtype = 2; r.nf = 9; r.planes = nx; r.f0 = 8; r.fn = 1;
  }
  else if (n == "M_palette") {  // MetricBF16 = 2: lambda, mu, rho, 0 of each palette entry (null without a palette)
    f32(c.met_pal, 4 * (size_t)c.met_npal);
  }
  else if (n == "M_bf16") { r.ptr = c.p.MetricBF16 ? c.met : nullptr; r.count = 13 * nxyz; r.dtype = 3; r.nf = 13; r.planes = nx; }
  else if (n == "M_media") {  // MetricBF16 = 2: lambda, mu, rho, the FP32 planes in slots 10-15
    f32(c.p.MetricBF16 == 2 ? c.met : nullptr, 3 * nxyz); r.nf = 9; r.planes = nx; r.f0 = 5; r.fn = 3;
  }
  else if (n == "M_skip") {  // MetricBF16 = 2: bit p set when 32-bit word p of the metric is +-0 everywhere (not loaded)
    r.ptr = c.p.MetricBF16 == 2 ? &c.met_skip : nullptr; r.count = 1; r.dtype = 2; r.host = true;
  }
  else if (n == "M_ref") { r.ptr = c.p.MetricBF16 == 2 ? c.met_ref : nullptr; r.count = 10; r.dtype = 0; r.host = true; }
  else if (n == "matVx2Vz") f32(c.matVx2Vz, (size_t)c.d.nx * c.d.ny * 9);
  else if (n == "matVy2Vz") f32(c.matVy2Vz, (size_t)c.d.nx * c.d.ny * 9);
  else if (n == "Bx") f32(c.pml.B[0], c.d.ni);
  else if (n == "By") f32(c.pml.B[1], c.d.nj);
  else if (n == "Bz") f32(c.pml.B[2], c.d.nk);
  else if (n == "Ax") f32(c.pml.A[0], c.d.ni);
  else if (n == "Ay") f32(c.pml.A[1], c.d.nj);
  else if (n == "Az") f32(c.pml.A[2], c.d.nk);
  else if (n == "Dx") f32(c.pml.D[0], c.d.ni);
  else if (n == "Dy") f32(c.pml.D[1], c.d.nj);
  else if (n == "Dz") f32(c.pml.D[2], c.d.nk);
  else if (n == "cerjan_x") f32(c.cj[0], c.d.nx);
  else if (n == "cerjan_y") f32(c.cj[1], c.d.ny);
  else if (n == "cerjan_z") f32(c.cj[2], c.d.nz);
  else if (n.size() == 5 && n.compare(0, 3, "PML") == 0) {
    int f = (n[3] - 'x') * 2 + (n[4] - '1');
    if (f >= 0 && f < 6 && c.pml.Y[f]) {
      int a = f / 2;
      long cnt = 9L * c.pml.N * (a == 0 ? (long)c.d.nj * c.d.nk : a == 1 ? (long)c.d.nk * c.d.ni : (long)c.d.ni * c.d.nj);
      f32(c.pml.Y[f], cnt);
    }
  } else if (c.p.Rupture && c.setup_done) {
    if (n == "FW") f32(F.W, 18 * nyz);
    else if (n == "FhW") f32(F.hW, 18 * nyz);
    else if (n == "FmW") f32(F.mW, 18 * nyz);
    else if (n == "T11") f32(F.T11, 7 * nyz);
    else if (n == "T12") f32(F.T12, 7 * nyz);
    else if (n == "T13") f32(F.T13, 7 * nyz);
    else if (n == "hT11") f32(F.hT11, nyz);
    else if (n == "hT12") f32(F.hT12, nyz);
    else if (n == "hT13") f32(F.hT13, nyz);
    else if (n == "slabH") { f32(c.slabH, 7 * 9 * nyz); r.nf = 9; r.planes = 7; }
    else if (n == "FWs") f32(F.Ws, F.Ws ? 18 * nyz : 0);
    else if (n == "Tn") f32(F.Tn, n2);
    else if (n == "Ts1") f32(F.Ts1, n2);
    else if (n == "Ts2") f32(F.Ts2, n2);
    else if (n == "tTn") f32(F.tTn, n2);
    else if (n == "tTs1") f32(F.tTs1, n2);
    else if (n == "tTs2") f32(F.tTs2, n2);
    else if (n == "Vs1") f32(F.Vs1, n2);
    // rate-state Newton samples of the last stage: features f0..f3 of the
    // learned guess and the solved log10 slip rate, [5][nk][nj]
    else if (n == "RS_samples") f32(F.rs_carryf ? F.rs_carryf + 4 * n2 : nullptr, 5 * n2);
    // per node: solves, iterations, fallbacks, learned-guess starts, network evaluations
    else if (n == "newton_counts") i32(F.newton_stats, 5 * n2);
    // wavefronts by their largest Newton iteration count (last bin: that many or more)
    else if (n == "newton_wave_hist") i32(F.newton_stats ? F.newton_stats + 5 * n2 : nullptr, kNewtonHistBins);
    else if (n == "Vs2") f32(F.Vs2, n2);
    else if (n == "hslip") f32(F.hslip, n2);
    else if (n == "slip") f64(F.slip, n2);
    else if (n == "State") f64(F.State, n2);
    else if (n == "TP_T") f64(F.TP_T, F.TP_T ? (size_t)F.TP_n * n2 : 0);
    else if (n == "TP_P") f64(F.TP_P, F.TP_P ? (size_t)F.TP_n * n2 : 0);
    else if (n == "TP_hy") f64(F.TP_hy, F.TP_hy ? n2 : 0);
    else if (n == "init_t0") f32(F.init_t0, n2);
    else if (n == "friction") f32(F.friction, n2);
    else if (n == "rup_sensor") f32(F.rup_sensor, n2);
    else if (n == "str_init_x") f32(F.str_init_x, n2);
    else if (n == "str_init_y") f32(F.str_init_y, n2);
    else if (n == "str_init_z") f32(F.str_init_z, n2);
    else if (n == "str_peak") f32(F.str_peak, n2);
    else if (n == "C0") f32(F.C0, n2);
    else if (n == "a") f32(F.a, n2);
    else if (n == "b") f32(F.b, n2);
    else if (n == "Vw") f32(F.Vw, n2);
    else if (n == "vec_n") f32(F.vec_n, 3 * nyz);
    else if (n == "vec_s1") f32(F.vec_s1, 3 * nyz);
    else if (n == "vec_s2") f32(F.vec_s2, 3 * nyz);
    else if (n == "rho_f") f32(F.rho_f, 2 * nyz);
    else if (n == "matT1toVxm") f32(F.matT1toVx, 9 * nyz);
    else if (n == "matT1toVxp") f32(F.matT1toVx + 9 * nyz, 9 * nyz);
    else if (n == "matVytoVxm") f32(F.matVytoVx, 9 * nyz);
    else if (n == "matVytoVxp") f32(F.matVytoVx + 9 * nyz, 9 * nyz);
    else if (n == "matVztoVxm") f32(F.matVztoVx, 9 * nyz);
    else if (n == "matVztoVxp") f32(F.matVztoVx + 9 * nyz, 9 * nyz);
    else if (n == "matVx2Vz1") f32(F.matVx2Vz_f, 9 * c.d.ny);
    else if (n == "matVx2Vz2") f32(F.matVx2Vz_f + 9 * c.d.ny, 9 * c.d.ny);
    else if (n == "matVy2Vz1") f32(F.matVy2Vz_f, 9 * c.d.ny);
    else if (n == "matVy2Vz2") f32(F.matVy2Vz_f + 9 * c.d.ny, 9 * c.d.ny);
    else if (n.size() == 13 && n.compare(0, 11, "matPlus2Min") == 0) f32(F.matP2Mf + (n[11] - '1') * 9 * c.d.ny, 9 * c.d.ny);
    else if (n.size() == 13 && n.compare(0, 11, "matMin2Plus") == 0) f32(F.matM2Pf + (n[11] - '1') * 9 * c.d.ny, 9 * c.d.ny);
    else if (n.size() == 5 && n[0] == 'D' && n[3] == '_') {
      int a = n[1] - '1', b = n[2] - '1', m = n[4] - '1';
      if (a >= 0 && a < 3 && b >= 0 && b < 3 && m >= 0 && m < 2) f32(F.Dmat + (long)(m * 9 + a * 3 + b) * nyz * 9, 9 * nyz);
    }
    else if (n == "united") i32(F.united, n2);
    else if (n == "faultgrid") i32(F.faultgrid, n2);
    else if (n == "rup_index_y") i32(F.rup_index_y, n2);
    else if (n == "rup_index_z") i32(F.rup_index_z, n2);
    else if (n == "init_t0_flag") i32(F.init_t0_flag, n2);
    else if (n == "flag_rup") i32(F.flag_rup, n2);
  }
  return r;
}

size_t dsize(int dt) { return dt == 1 ? 8 : dt == 3 ? 2 : 4; }

// ---------------------------------------------------------------------------
// Output files (io_fault.cu:16-275, io_wave_xy.cu, io_wave_xz.cu, io_wave_yz.cu)
struct OutFiles {
  NcFile fault, xy, xz, yz;
  std::vector<int> fv;  // record var ids of the fault file
  int f_init_t0 = -1;
  int wxy[3], wxz[3], wyz[3];
};

}  // namespace

namespace cgfdm {
struct NcOut : OutFiles {};
}  // namespace cgfdm

extern "C" {

const char* cgfdm_last_error(void) { return g_err.c_str(); }

int cgfdm_create_from_string(const char* json_text, cgfdm_ctx** out) {
  if (!json_text || !out) return fail(CGFDM_EPARAM, "null argument");
  cgfdm_ctx* h = new cgfdm_ctx();
  std::string err;
  if (parse_params_json(json_text, h->c.p, err)) {
    delete h;
    return fail(CGFDM_EJSON, err);
  }
  *out = h;
  return 0;
}

int cgfdm_create(const char* path, cgfdm_ctx** out) {
  if (!path || !out) return fail(CGFDM_EPARAM, "null argument");
  std::ifstream in(path, std::ios::binary);
  if (!in) return fail(CGFDM_EFILE, std::string("Error at opening the configuration file ") + path);
  std::stringstream ss;
  ss << in.rdbuf();
  return cgfdm_create_from_string(ss.str().c_str(), out);
}

int cgfdm_set_param(cgfdm_ctx* h, const char* key, double v) {
  if (!h || !key) return fail(CGFDM_EPARAM, "null argument");
  if (h->c.setup_done) return fail(CGFDM_ESTATE, "parameters are fixed after setup");
  if (h->c.p.set(key, v)) return fail(CGFDM_EPARAM, std::string("unknown parameter ") + key);
  h->c.p.derive();
  return 0;
}

int cgfdm_get_param(cgfdm_ctx* h, const char* key, double* v) {
  if (!h || !key || !v) return fail(CGFDM_EPARAM, "null argument");
  h->c.p.derive();
  if (h->c.p.get(key, v)) return fail(CGFDM_EPARAM, std::string("unknown parameter ") + key);
  return 0;
}

int cgfdm_destroy(cgfdm_ctx* h) {
  if (!h) return 0;
  Ctx& c = h->c;
  if (c.out) { cgfdm_close_outputs(h); }
  // teardown is best effort: each call runs whatever an earlier one returned
  if (c.setup_done) (void)hipSetDevice(c.device);
  if (c.stream) (void)hipStreamSynchronize(c.stream);
  if (c.own_stream) { (void)hipStreamSynchronize(c.own_stream); (void)hipStreamDestroy(c.own_stream); }
  for (hipStream_t st : c.side)
    if (st) { (void)hipStreamSynchronize(st); (void)hipStreamDestroy(st); }
  for (hipEvent_t e : c.fence)
    if (e) (void)hipEventDestroy(e);
  for (void* p : c.allocations) (void)hipFree(p);
  if (c.hW) (void)hipFree(c.hW);
  if (c.wave_work) (void)hipFree(c.wave_work);
  if (c.pml_boxes) (void)hipFree(c.pml_boxes);
  if (c.mlp_dev) (void)hipFree(c.mlp_dev);
  delete h;
  return 0;
}

int cgfdm_set_fault_geometry(cgfdm_ctx* h, const float* x, const float* y, const float* z, size_t n) {
  if (!h || !x || !y || !z) return fail(CGFDM_EPARAM, "null argument");
  h->c.gx.assign(x, x + n); h->c.gy.assign(y, y + n); h->c.gz.assign(z, z + n);
  return 0;
}

int cgfdm_set_coord3d(cgfdm_ctx* h, const double* xyz, size_t n) {
  if (!h || !xyz) return fail(CGFDM_EPARAM, "null argument");
  h->c.coord3d.assign(xyz, xyz + n);
  return 0;
}

int cgfdm_set_topography(cgfdm_ctx* h, const float* t, size_t n) {
  if (!h || !t) return fail(CGFDM_EPARAM, "null argument");
  h->c.topo.assign(t, t + n);
  return 0;
}

int cgfdm_set_media1d(cgfdm_ctx* h, int n, const float* depth, const float* rho, const float* vp, const float* vs) {
  if (!h || n <= 0 || !depth || !rho || !vp || !vs) return fail(CGFDM_EPARAM, "bad media1d arguments");
  Ctx& c = h->c;
  c.m1d_depth.assign(depth, depth + n); c.m1d_rho.assign(rho, rho + n);
  c.m1d_vp.assign(vp, vp + n); c.m1d_vs.assign(vs, vs + n);
  return 0;
}

int cgfdm_set_media3d(cgfdm_ctx* h, const float* vp, const float* vs, const float* rho, size_t n) {
  if (!h || !vp || !vs || !rho) return fail(CGFDM_EPARAM, "null argument");
  Ctx& c = h->c;
  c.p.derive();
  if (n != (size_t)c.p.ni * c.p.nj * c.p.nk) return fail(CGFDM_ESIZE, "Media3D needs NX*NY*NZ values");
  c.m3d_vp.assign(vp, vp + n); c.m3d_vs.assign(vs, vs + n); c.m3d_rho.assign(rho, rho + n);
  return 0;
}

int cgfdm_set_init_stress(cgfdm_ctx* h, const char* name, const float* data, size_t n) {
  if (!h || !name || !data) return fail(CGFDM_EPARAM, "null argument");
  static const char* ok[] = {"Tx", "Ty", "Tz", "dTx", "dTy", "dTz", "mu_s", "mu_d", "Dc", "C0", "a", "b", "L", "Vw", "State",
                             "TP_hy"};
  for (const char* s : ok)
    if (std::strcmp(s, name) == 0) {
      h->c.stress_in[name] = std::vector<float>(data, data + n);
      return 0;
    }
  return fail(CGFDM_EPARAM, std::string("unknown stress variable ") + name);
}

int cgfdm_setup(cgfdm_ctx* h) {
  if (!h) return fail(CGFDM_EPARAM, "null context");
  Ctx& c = h->c;
  if (c.setup_done) return fail(CGFDM_ESTATE, "setup already done");
  int ndev = 0;
  if (hipGetDeviceCount(&ndev) != hipSuccess || ndev == 0) return fail(CGFDM_EHIP, "no HIP device");
  c.device = ndev > 1 ? c.p.igpu % ndev : 0;
  if (hipError_t e = hipSetDevice(c.device)) return fail(CGFDM_EHIP, hipGetErrorString(e));
  std::string err;
  int rc = host_setup(c, err);
  if (rc) {
    int code = rc == 5 ? CGFDM_ECFL : rc == 4 ? CGFDM_ENETCDF : rc == 2 ? CGFDM_EFILE : rc == 6 ? CGFDM_EHIP : CGFDM_EPARAM;
    return fail(code, err);
  }
  if (hipStreamCreateWithFlags(&c.own_stream, hipStreamNonBlocking) != hipSuccess)
    return fail(CGFDM_EHIP, "stream creation failed");
  c.stream = c.own_stream;
  if (c.p.StageStreams) {  // default priority, like every stream here
    for (hipStream_t& st : c.side)
      if (hipStreamCreateWithFlags(&st, hipStreamNonBlocking) != hipSuccess) return fail(CGFDM_EHIP, "stream creation failed");
    for (hipEvent_t& e : c.fence)
      if (hipEventCreateWithFlags(&e, hipEventDisableTiming) != hipSuccess) return fail(CGFDM_EHIP, "event creation failed");
  }
  c.setup_done = true;
  return 0;
}

int cgfdm_stage(cgfdm_ctx* h, int it, int irk) {
  if (!h) return fail(CGFDM_EPARAM, "null context");
  Ctx& c = h->c;
  if (!c.setup_done) return fail(CGFDM_ESTATE, "stage before setup");
  if (irk < 0 || irk > 3 || it < 0) return fail(CGFDM_EPARAM, "bad stage index");
  int fx, fy, fz;
  stage_flags(it, irk, fx, fy, fz);
  StageBufs b = bufs_for(c, irk);
  if (c.p.Rupture) {
    launch_fault_pre(c, b.Win, it, irk, fx, fy, fz);
    if (c.p.Friction_type >= 1) launch_fault_rs(c, b.Win, it, irk, fx, fy, fz);
  }
  launch_wave_stage(c, b, irk, fx, fy, fz, false, nullptr);
  if (c.p.Rupture) {
    launch_fault_deriv(c, b.Win, fx, fy, fz);
    launch_fault_post(c, b, irk);
  }
  c.cur = b.Wout;
  return check_hip("stage launch");
}

int cgfdm_step(cgfdm_ctx* h, int it) {
  if (!h) return fail(CGFDM_EPARAM, "null context");
  Ctx& c = h->c;
  if (!c.setup_done) return fail(CGFDM_ESTATE, "step before setup");
  if (c.p.Rupture) launch_rup_sensor(c);
  if (c.p.Rupture && c.p.Friction_type == 3) launch_thermpress(c);  // main.cu:445-447
  for (int irk = 0; irk < 4; irk++) {
    int rc = cgfdm_stage(h, it, irk);
    if (rc) return rc;
  }
  // smooth_gauss_volume once per step after the stages (main.cu:517-523)
  if (c.p.Rupture && c.p.FaultSmooth) {
    launch_fault_smooth(c);
    return check_hip("fault smoothing launch");
  }
  return 0;
}

int cgfdm_synchronize(cgfdm_ctx* h) {
  if (!h || !h->c.stream) return fail(CGFDM_ESTATE, "no stream");
  hipError_t e = hipStreamSynchronize(h->c.stream);
  if (e != hipSuccess) return fail(CGFDM_EHIP, hipGetErrorString(e));
  return 0;
}

int cgfdm_run(cgfdm_ctx* h, int it0, int nt, int write_outputs) {
  if (!h) return fail(CGFDM_EPARAM, "null context");
  for (int it = it0; it < it0 + nt; it++) {
    if (write_outputs) {
      int rc = cgfdm_write_outputs(h, it);
      if (rc) return rc;
    }
    int rc = cgfdm_step(h, it);
    if (rc) return rc;
  }
  return cgfdm_synchronize(h);
}

int cgfdm_wave_stage(cgfdm_ctx* h, int irk, int fx, int fy, int fz) {
  if (!h) return fail(CGFDM_EPARAM, "null context");
  Ctx& c = h->c;
  if (!c.setup_done) return fail(CGFDM_ESTATE, "before setup");
  if (irk < 0 || irk > 3) return fail(CGFDM_EPARAM, "bad irk");
  StageBufs b = bufs_for(c, irk);
  launch_wave_stage(c, b, irk, fx, fy, fz, false, nullptr);
  c.cur = b.Wout;
  return check_hip("wave_stage launch");
}

int cgfdm_wave_rhs(cgfdm_ctx* h, int fx, int fy, int fz) {
  if (!h) return fail(CGFDM_EPARAM, "null context");
  Ctx& c = h->c;
  if (!c.setup_done) return fail(CGFDM_ESTATE, "before setup");
  if (!c.hW) {
    const long nw = (long)c.d.nx * 9 * c.d.pyz;
    if (hipMalloc(&c.hW, nw * sizeof(float)) != hipSuccess) return fail(CGFDM_EHIP, "hipMalloc hW");
    launch_fill(c, c.hW, nw, 0.0f);
  }
  StageBufs b{c.cur, nullptr, c.S, c.Tacc};
  launch_wave_stage(c, b, 0, fx, fy, fz, true, c.hW);
  return check_hip("wave_rhs launch");
}

int cgfdm_stage_flags(int it, int irk, int* fx, int* fy, int* fz) {
  if (!fx || !fy || !fz) return fail(CGFDM_EPARAM, "null argument");
  if (irk < 0 || irk > 3 || it < 0) return fail(CGFDM_EPARAM, "bad stage index");
  stage_flags(it, irk, *fx, *fy, *fz);
  return 0;
}

int cgfdm_set_stream(cgfdm_ctx* h, void* stream) {
  if (!h) return fail(CGFDM_EPARAM, "null context");
  Ctx& c = h->c;
  if (!c.setup_done) return fail(CGFDM_ESTATE, "set_stream before setup");
  hipStream_t next = stream ? (hipStream_t)stream : c.own_stream;
  if (next == c.stream) return 0;
  // work already queued on the previous stream finishes before any later
  // launch can touch the same buffers
  hipError_t e = hipStreamSynchronize(c.stream);
  if (e != hipSuccess) return fail(CGFDM_EHIP, hipGetErrorString(e));
  c.stream = next;
  return 0;
}

// Fault phases of one RK stage.  Called in the order cgfdm_stage runs them
// (trial, wave stage, derivatives, update) they queue the same launches with
// the same arguments, so the result is the same bit for bit.
static int fault_phase_ready(cgfdm_ctx* h, int it, int irk, int& fx, int& fy, int& fz) {
  if (!h) return fail(CGFDM_EPARAM, "null context");
  Ctx& c = h->c;
  if (!c.setup_done) return fail(CGFDM_ESTATE, "before setup");
  if (!c.p.Rupture) return fail(CGFDM_ESTATE, "no fault (Rupture = 0)");
  if (irk < 0 || irk > 3 || it < 0) return fail(CGFDM_EPARAM, "bad stage index");
  stage_flags(it, irk, fx, fy, fz);
  return 0;
}

int cgfdm_fault_trial(cgfdm_ctx* h, int it, int irk) {
  int fx, fy, fz;
  if (int rc = fault_phase_ready(h, it, irk, fx, fy, fz)) return rc;
  Ctx& c = h->c;
  StageBufs b = bufs_for(c, irk);
  launch_fault_pre(c, b.Win, it, irk, fx, fy, fz);
  if (c.p.Friction_type >= 1) launch_fault_rs(c, b.Win, it, irk, fx, fy, fz);
  return check_hip("fault trial launch");
}

int cgfdm_fault_deriv(cgfdm_ctx* h, int it, int irk) {
  int fx, fy, fz;
  if (int rc = fault_phase_ready(h, it, irk, fx, fy, fz)) return rc;
  Ctx& c = h->c;
  launch_fault_deriv(c, bufs_for(c, irk).Win, fx, fy, fz);
  return check_hip("fault derivative launch");
}

int cgfdm_fault_update(cgfdm_ctx* h, int it, int irk) {
  int fx, fy, fz;
  if (int rc = fault_phase_ready(h, it, irk, fx, fy, fz)) return rc;
  Ctx& c = h->c;
  StageBufs b = bufs_for(c, irk);
  launch_fault_post(c, b, irk);
  c.cur = b.Wout;
  return check_hip("fault update launch");
}

int cgfdm_rup_sensor(cgfdm_ctx* h) {
  int fx, fy, fz;
  if (int rc = fault_phase_ready(h, 0, 0, fx, fy, fz)) return rc;
  launch_rup_sensor(h->c);
  return check_hip("rupture sensor launch");
}

int cgfdm_thermpress(cgfdm_ctx* h) {
  int fx, fy, fz;
  if (int rc = fault_phase_ready(h, 0, 0, fx, fy, fz)) return rc;
  if (h->c.p.Friction_type != 3) return fail(CGFDM_ESTATE, "thermal pressurization needs Friction_type 3");
  launch_thermpress(h->c);
  return check_hip("thermal pressurization launch");
}

int cgfdm_fault_smooth(cgfdm_ctx* h) {
  int fx, fy, fz;
  if (int rc = fault_phase_ready(h, 0, 0, fx, fy, fz)) return rc;
  if (!h->c.p.FaultSmooth) return fail(CGFDM_ESTATE, "FaultSmooth is off");
  launch_fault_smooth(h->c);
  return check_hip("fault smoothing launch");
}

// Between the caller's layout (the reference's: [field][plane][k][j], rows
// of ny) and the device's plane-interleaved [plane][field][k][jpad + j] with
// rows of py; to_dev selects the direction.  The slab buffer is the one
// exception on the caller's side: it is [plane][field][k][j] there too.
static size_t dev_words(const Ctx& c, const FieldRef& r) { return (size_t)r.planes * r.nf * c.d.pyz; }
static void interleave_copy(const Ctx& c, const FieldRef& r, const char* from, char* to, bool to_dev) {
  const size_t es = dsize(r.dtype), nz = (size_t)c.d.nz, ny = (size_t)c.d.ny, py = (size_t)c.d.py;
  const size_t np = (size_t)r.planes, nf = (size_t)r.nf, pyz = (size_t)c.d.pyz, nyz = ny * nz;
  const bool slab = np != (size_t)c.d.nx;
  for (size_t i = 0; i < np; i++)
    for (size_t f = (size_t)r.f0; f < (r.fn ? (size_t)(r.f0 + r.fn) : nf); f++)
      for (size_t k = 0; k < nz; k++) {
        const size_t fo = f - (size_t)r.f0, nfo = r.fn ? (size_t)r.fn : nf;
        const size_t dev = ((i * nf + f) * pyz + k * py + c.d.jpad) * es;
        const size_t ref = ((slab ? i * nfo + fo : fo * np + i) * nyz + k * ny) * es;
        std::memcpy(to + (to_dev ? dev : ref), from + (to_dev ? ref : dev), ny * es);
      }
}

int cgfdm_field_info(cgfdm_ctx* h, const char* name, size_t* count, int* dtype) {
  if (!h || !name) return fail(CGFDM_EPARAM, "null argument");
  FieldRef r = field(h->c, name);
  if (!r.ptr) return fail(CGFDM_EPARAM, std::string("unknown or unavailable field ") + name);
  if (count) *count = r.count;
  if (dtype) *dtype = r.dtype;
  return 0;
}

int cgfdm_get_field(cgfdm_ctx* h, const char* name, void* dst, size_t bytes) {
  if (!h || !name || !dst) return fail(CGFDM_EPARAM, "null argument");
  Ctx& c = h->c;
  if (!c.setup_done) return fail(CGFDM_ESTATE, "before setup");
  FieldRef r = field(c, name);
  if (!r.ptr) return fail(CGFDM_EPARAM, std::string("unknown or unavailable field ") + name);
  if (bytes != r.count * dsize(r.dtype)) return fail(CGFDM_ESIZE, std::string("size mismatch for ") + name);
  if (r.host) {
    std::memcpy(dst, r.ptr, bytes);
    return 0;
  }
  if (hipError_t e = hipStreamSynchronize(c.stream)) return fail(CGFDM_EHIP, hipGetErrorString(e));
  if (!r.nf) {
    if (hipMemcpy(dst, r.ptr, bytes, hipMemcpyDeviceToHost) != hipSuccess) return fail(CGFDM_EHIP, "copy failed");
    return 0;
  }
  std::vector<char> tmp(dev_words(c, r) * dsize(r.dtype));
  if (hipMemcpy(tmp.data(), r.ptr, tmp.size(), hipMemcpyDeviceToHost) != hipSuccess) return fail(CGFDM_EHIP, "copy failed");
  interleave_copy(c, r, tmp.data(), (char*)dst, false);
  return 0;
}

int cgfdm_put_field(cgfdm_ctx* h, const char* name, const void* src, size_t bytes) {
  if (!h || !name || !src) return fail(CGFDM_EPARAM, "null argument");
  Ctx& c = h->c;
  if (!c.setup_done) return fail(CGFDM_ESTATE, "before setup");
  std::string n(name);
  if (n == "W") {  // a new wavefield starts a step: it lands in the step-start buffer
    c.cur = c.S;
  }
  FieldRef r = field(c, name);
  if (!r.ptr) return fail(CGFDM_EPARAM, std::string("unknown or unavailable field ") + name);
  if (bytes != r.count * dsize(r.dtype)) return fail(CGFDM_ESIZE, std::string("size mismatch for ") + name);
  if (r.host) return fail(CGFDM_EPARAM, std::string("read-only field ") + name);
  if (hipError_t e = hipStreamSynchronize(c.stream)) return fail(CGFDM_EHIP, hipGetErrorString(e));
  if (!r.nf) {
    if (hipMemcpy(r.ptr, src, bytes, hipMemcpyHostToDevice) != hipSuccess) return fail(CGFDM_EHIP, "copy failed");
    return 0;
  }
  // read-modify-write keeps the row padding of the device array as it is
  std::vector<char> tmp(dev_words(c, r) * dsize(r.dtype));
  if (hipMemcpy(tmp.data(), r.ptr, tmp.size(), hipMemcpyDeviceToHost) != hipSuccess) return fail(CGFDM_EHIP, "copy failed");
  interleave_copy(c, r, (const char*)src, tmp.data(), true);
  if (hipMemcpy(r.ptr, tmp.data(), tmp.size(), hipMemcpyHostToDevice) != hipSuccess) return fail(CGFDM_EHIP, "copy failed");
  return 0;
}

int cgfdm_get_dims(cgfdm_ctx* h, int dims[8]) {
  if (!h || !dims) return fail(CGFDM_EPARAM, "null argument");
  Params& p = h->c.p;
  p.derive();
  dims[0] = p.ni; dims[1] = p.nj; dims[2] = p.nk; dims[3] = p.nx; dims[4] = p.ny; dims[5] = p.nz;
  dims[6] = p.nx / 2; dims[7] = p.NT;
  return 0;
}

int cgfdm_profile(cgfdm_ctx* h, int enable) {
  if (!h) return fail(CGFDM_EPARAM, "null context");
  Ctx& c = h->c;
  if (c.stream) (void)hipStreamSynchronize(c.stream);  // a failure shows in the event times read next
  c.prof.reset();
  c.prof.on = enable != 0;
  if (c.prof.on) c.prof.grow(8192);
  return check_hip("profile");
}

int cgfdm_profile_classes(cgfdm_ctx* h, const char* names) {
  if (!h) return fail(CGFDM_EPARAM, "null context");
  unsigned mask = 0;
  if (!names || !*names) {
    mask = ~0u;
  } else {
    std::string list(names);
    size_t pos = 0;
    while (pos <= list.size()) {
      const size_t e = std::min(list.find(',', pos), list.size());
      const std::string nm = list.substr(pos, e - pos);
      int k = -1;
      for (int q = 0; q < PK_COUNT; q++)
        if (nm == kProfNames[q]) k = q;
      if (k < 0) return fail(CGFDM_EPARAM, "unknown kernel class " + nm);
      mask |= 1u << k;
      pos = e + 1;
    }
  }
  h->c.prof.mask = mask;
  return 0;
}

int cgfdm_profile_get(cgfdm_ctx* h, const char* name, double* total_ms, long* launches, long* points) {
  if (!h || !name) return fail(CGFDM_EPARAM, "null argument");
  Ctx& c = h->c;
  int k = -1;
  for (int q = 0; q < PK_COUNT; q++)
    if (std::strcmp(name, kProfNames[q]) == 0) k = q;
  if (k < 0) return fail(CGFDM_EPARAM, std::string("unknown kernel class ") + name);
  if (c.stream) (void)hipStreamSynchronize(c.stream);  // a failure shows in the event times read next
  double ms = 0.0;
  long n = 0;
  for (const auto& m : c.prof.marks) {
    if (m.first != k) continue;
    float t = 0.0f;
    if (hipEventElapsedTime(&t, c.prof.pool[m.second], c.prof.pool[m.second + 1]) != hipSuccess)
      return fail(CGFDM_EHIP, "event timing failed");
    ms += t;
    n++;
  }
  if (total_ms) *total_ms = ms;
  if (launches) *launches = n;
  if (points) *points = k <= PK_STAGE_SURFACE ? c.wave_points[k] : 0;
  return 0;
}

int cgfdm_profile_aux(cgfdm_ctx* h, const char* name, long* read_pairs, long* update_pairs) {
  if (!h || !name) return fail(CGFDM_EPARAM, "null argument");
  Ctx& c = h->c;
  if (!c.setup_done) return fail(CGFDM_ESTATE, "cgfdm_setup has not run");
  int k = -1;
  for (int q = 0; q < PK_COUNT; q++)
    if (std::strcmp(name, kProfNames[q]) == 0) k = q;
  if (k < 0) return fail(CGFDM_EPARAM, std::string("unknown kernel class ") + name);
  const int x = k <= PK_STAGE_SURFACE ? k : k == PK_PML_AUX ? 3 : -1;
  if (read_pairs) *read_pairs = x >= 0 ? c.aux_read_pairs[x] : 0;
  if (update_pairs) *update_pairs = x >= 0 ? c.aux_update_pairs[x] : 0;
  return 0;
}

int cgfdm_set_newton_mlp(cgfdm_ctx* h, const float* params, long n) {
  if (!h || !params) return fail(CGFDM_EPARAM, "null argument");
  if (n != kMlpParams) return fail(CGFDM_EPARAM, "the Newton network needs 4555 parameters");
  Ctx& c = h->c;
  c.mlp_blob.assign(params, params + n);
  if (c.setup_done && mlp_upload(c, params, n)) return fail(CGFDM_EHIP, "upload of the Newton network failed");
  return 0;
}

int cgfdm_mlp_forward(const float* params, const float* feat, long n, float* out, int mode) {
  if (!params || n < 0 || (n > 0 && (!feat || !out)) || (mode != 0 && mode != 1)) return fail(CGFDM_EPARAM, "bad argument");
  std::string err;
  if (mlp_forward_host(params, feat, n, out, mode, err)) return fail(CGFDM_EHIP, err);
  return 0;
}

int cgfdm_mlp_forward_noalloc(const float* params, long nparams, const float* feat, long n, float* out, int mode,
                              void* stream) {
  if (nparams != kMlpParams)
    return fail(CGFDM_ESIZE, "cgfdm_mlp_forward_noalloc: the scratch holds a network of exactly " +
                                 std::to_string(kMlpParams) + " parameters, got " + std::to_string(nparams));
  if (n < 0 || n > 0x7fffffffL) return fail(CGFDM_ESIZE, "cgfdm_mlp_forward_noalloc: n must be in 0..2^31-1");
  if (!params || (n > 0 && (!feat || !out)) || (mode != 0 && mode != 1)) return fail(CGFDM_EPARAM, "bad argument");
  std::string err;
  if (mlp_forward_noalloc_host(params, feat, n, out, mode, (hipStream_t)stream, err)) return fail(CGFDM_EHIP, err);
  return 0;
}

int cgfdm_mlp_grad(const float* params, const float* feat, const float* target, long n, float* grad, float* loss) {
  if (!params || !feat || !target || !grad || !loss || n <= 0) return fail(CGFDM_EPARAM, "bad argument");
  std::string err;
  if (mlp_grad_host(params, feat, target, n, grad, loss, err)) return fail(CGFDM_EHIP, err);
  return 0;
}

int cgfdm_mlp_train(float* params, const float* feat, const float* target, long n, int epochs, int batch, float lr,
                    unsigned seed, float* loss_hist) {
  if (!params || !feat || !target || !loss_hist || n <= 0 || n > 0x7fffffffL || epochs <= 0 || batch <= 0 || !(lr > 0))
    return fail(CGFDM_EPARAM, "bad argument");
  std::string err;
  if (mlp_train_host(params, feat, target, n, epochs, batch, lr, seed, loss_hist, err)) return fail(CGFDM_EHIP, err);
  return 0;
}

int cgfdm_check_field(cgfdm_ctx* h, const char* name, long long* nonfinite, long long* ghost_nonzero, double* maxabs) {
  if (!h || !name) return fail(CGFDM_EPARAM, "null argument");
  Ctx& c = h->c;
  if (!c.setup_done) return fail(CGFDM_ESTATE, "before setup");
  FieldRef r = field(c, name);
  if (!r.ptr || r.host) return fail(CGFDM_EPARAM, std::string("unknown or unavailable field ") + name);
  if (r.dtype != 0) return fail(CGFDM_EPARAM, std::string("not an FP32 field: ") + name);
  std::string err;
  const int fn = r.fn ? r.fn : r.nf;
  const int rc = r.nf ? field_check_host(c, (const float*)r.ptr, r.nf, r.f0, fn, 0, nonfinite, ghost_nonzero, maxabs, err)
                      : field_check_host(c, (const float*)r.ptr, 0, 0, 1, (long)r.count, nonfinite, ghost_nonzero, maxabs, err);
  return rc ? fail(CGFDM_EHIP, err) : 0;
}

int cgfdm_print_params(cgfdm_ctx* h) {
  if (!h) return fail(CGFDM_EPARAM, "null context");
  const Params& p = h->c.p;
  std::printf("# Parameters setting:\n"
              "* Tmax           = %g (sec)\n"
              "* DT             = %g (sec)\n"
              "* DH             = %g (m)\n"
              "* NX, NY, NZ, NT = %d %d %d %d\n"
              "* PX, PY, PZ     = %d %d %d\n"
              "* Fault_grid     = %d %d %d %d\n"
              "* Asp_grid       = %d %d %d %d\n"
              "* vp             = %g/%g\n"
              "* vs             = %g/%g\n"
              "* rho            = %g/%g\n"
              "* PML_N          = %d\n"
              "* PML_velocity   = %g\n"
              "* PML_fc         = %g\n"
              "* PML_bmax       = %g\n"
              "* RupThres       = %g\n\n",
              p.TMAX, p.DT, p.DH, p.NX, p.NY, p.NZ, p.NT, p.PX, p.PY, p.PZ, p.Fault_grid[0], p.Fault_grid[1],
              p.Fault_grid[2], p.Fault_grid[3], p.Asp_grid[0], p.Asp_grid[1], p.Asp_grid[2], p.Asp_grid[3], p.bi_vp1,
              p.bi_vp2, p.bi_vs1, p.bi_vs2, p.bi_rho1, p.bi_rho2, p.PML_N, p.PML_velocity, p.PML_fc, p.PML_bmax,
              p.RupThres);
  std::printf("# Friction type: %d\n", p.Friction_type);
  if (p.Friction_type == 0)
    std::printf("* Dc             = %g (m)\n* C0             = %g (Pa)\n* mu_s           = %g\n* mu_d           = %g\n\n",
                p.Dc, p.C0, p.mu_s, p.mu_d);
  else if (p.Friction_type >= 1 && p.Friction_type <= 3)
    std::printf("* RS_V0          = %g (m/s)\n* RS_Vini        = %g (m/s)\n* RS_f0          = %g\n* RS_fw          = %g\n"
                "* RS_L           = %g (m)\n* TP_n           = %d\n\n",
                p.V0, p.Vini, p.f0, p.fw, p.L, p.TP_n);
  std::printf("* smooth_load_T   = %g (s)\n", p.smooth_load_T);
  if (!p.Media1D.empty()) std::printf("Media1D = %s\n", p.Media1D.c_str());
  if (!p.Media3D.empty()) std::printf("Media3D = %s\n", p.Media3D.c_str());
  std::fflush(stdout);
  return 0;
}

int cgfdm_stream_probe(int nread, int nwrite, size_t bytes_per_array, int reps, double* gbps) {
  if (!gbps) return fail(CGFDM_EPARAM, "null argument");
  std::string err;
  if (stream_probe_host(nread, nwrite, bytes_per_array, reps, gbps, err)) return fail(CGFDM_EPARAM, err);
  return 0;
}

int cgfdm_newton_stats(cgfdm_ctx* h, double* solves, double* iterations, double* fallbacks) {
  if (!h) return fail(CGFDM_EPARAM, "null context");
  Ctx& c = h->c;
  unsigned long long s[3] = {0, 0, 0};
  if (c.setup_done && c.p.Rupture && c.f.newton_stats) {
    const long n2 = (long)c.d.nj * c.d.nk;
    std::vector<unsigned int> v(3 * n2);
    if (hipError_t e = hipStreamSynchronize(c.stream)) return fail(CGFDM_EHIP, hipGetErrorString(e));
    if (hipMemcpy(v.data(), c.f.newton_stats, v.size() * sizeof(unsigned int), hipMemcpyDeviceToHost) != hipSuccess)
      return fail(CGFDM_EHIP, "copy failed");
    for (int t = 0; t < 3; t++)
      for (long q = 0; q < n2; q++) s[t] += v[t * n2 + q];
  }
  if (solves) *solves = (double)s[0];
  if (iterations) *iterations = (double)s[1];
  if (fallbacks) *fallbacks = (double)s[2];
  return 0;
}

// ---------------------------------------------------------------------------
int cgfdm_open_outputs(cgfdm_ctx* h) {
  if (!h) return fail(CGFDM_EPARAM, "null context");
  Ctx& c = h->c;
  if (!c.setup_done) return fail(CGFDM_ESTATE, "before setup");
  if (c.out) return 0;
  mkdir(c.p.OUT.c_str(), 0755);
  NcOut* o = new NcOut();
  std::string err;
  const int nj = c.d.nj, nk = c.d.nk, ni = c.d.ni;
  if (c.p.Rupture) {
    NcFile& f = o->fault;
    if (!f.create(c.p.OUT + "/fault_mpi000000.nc", err)) { delete o; return fail(CGFDM_ENETCDF, err); }
    int dt = f.def_dim("nt", 0), dz = f.def_dim("nz", nk), dy = f.def_dim("ny", nj);
    std::vector<int> d2 = {dz, dy}, d3 = {dt, dz, dy};
    int vx = f.def_var("x", NC_FLOAT_, d2), vy = f.def_var("y", NC_FLOAT_, d2), vz = f.def_var("z", NC_FLOAT_, d2);
    o->f_init_t0 = f.def_var("init_t0", NC_FLOAT_, d2);
    int vc0 = f.def_var("C0", NC_FLOAT_, d2), vpk = f.def_var("str_peak", NC_FLOAT_, d2), vun = f.def_var("united", NC_INT_, d2);
    int va = f.def_var("a", NC_FLOAT_, d2), vb = f.def_var("b", NC_FLOAT_, d2), vw = f.def_var("Vw", NC_FLOAT_, d2);
    for (const char* nm : {"Vs1", "Vs2", "tn", "ts1", "ts2", "Us0", "State"}) o->fv.push_back(f.def_var(nm, NC_FLOAT_, d3));
    o->fv.push_back(f.def_var("rup_index_y", NC_INT_, d3));
    o->fv.push_back(f.def_var("rup_index_z", NC_INT_, d3));
    for (const char* nm : {"rup_sensor", "TP_T", "TP_P"}) o->fv.push_back(f.def_var(nm, NC_FLOAT_, d3));
    if (!f.end_def(err)) { delete o; return fail(CGFDM_ENETCDF, err); }
    std::vector<float> buf((size_t)nj * nk);
    std::vector<int> ib((size_t)nj * nk);
    for (int comp = 0; comp < 3; comp++) f.put_var(comp == 0 ? vx : comp == 1 ? vy : vz, &c.out_coord[0][comp * (size_t)nj * nk]);
    auto put2 = [&](const char* fld, int vid) -> int {
      FieldRef r = field(c, fld);
      if (hipMemcpy(buf.data(), r.ptr, buf.size() * 4, hipMemcpyDeviceToHost) != hipSuccess) return fail(CGFDM_EHIP, "copy failed");
      f.put_var(vid, buf.data());
      return 0;
    };
    const std::pair<const char*, int> statics[] = {{"C0", vc0}, {"str_peak", vpk}, {"a", va}, {"b", vb}, {"Vw", vw}};
    for (const auto& fv : statics)
      if (int rc = put2(fv.first, fv.second)) return rc;
    if (hipMemcpy(ib.data(), c.f.united, ib.size() * 4, hipMemcpyDeviceToHost) != hipSuccess) return fail(CGFDM_EHIP, "copy failed");
    f.put_var(vun, ib.data());
  }
  // wave slices: xy at global k, xz at global j, yz at global i
  auto slice_file = [&](NcFile& f, const std::string& name, int d1n, int d2n, const char* n1, const char* n2, int* vids,
                        int axis, int idx) -> bool {
    if (!f.create(c.p.OUT + "/" + name, err)) return false;
    int dt = f.def_dim("nt", 0), da = f.def_dim(n1, d1n), db = f.def_dim(n2, d2n);
    std::vector<int> d2 = {da, db}, d3 = {dt, da, db};
    int vx = f.def_var("x", NC_FLOAT_, d2), vy = f.def_var("y", NC_FLOAT_, d2), vz = f.def_var("z", NC_FLOAT_, d2);
    vids[0] = f.def_var("Vx", NC_FLOAT_, d3); vids[1] = f.def_var("Vy", NC_FLOAT_, d3); vids[2] = f.def_var("Vz", NC_FLOAT_, d3);
    if (!f.end_def(err)) return false;
    // the slice's coordinates, gathered at setup: xy (nx, ny) at k, xz (nx, nz) at j, yz (nz, ny) at i
    const std::vector<float>& C = c.out_coord[axis == 2 ? 1 : axis == 1 ? 2 : 3];
    for (int comp = 0; comp < 3; comp++) f.put_var(comp == 0 ? vx : comp == 1 ? vy : vz, &C[comp * (size_t)d1n * d2n]);
    return true;
  };
  int kz = std::min(std::max(c.p.EXPORT_WAVE_SLICE_Z, 0), nk - 1);
  int jy = std::min(std::max(c.p.EXPORT_WAVE_SLICE_Y, 0), nj - 1);
  int ix = std::min(std::max(c.p.EXPORT_WAVE_SLICE_X, 0), ni - 1);
  if (!slice_file(o->xy, "wave_xy_mpi000000.nc", ni, nj, "nx", "ny", o->wxy, 2, kz) ||
      !slice_file(o->xz, "wave_xz_mpi000000.nc", ni, nk, "nx", "nz", o->wxz, 1, jy) ||
      !slice_file(o->yz, "wave_yz_mpi000000.nc", nk, nj, "nz", "ny", o->wyz, 0, ix)) {
    delete o;
    return fail(CGFDM_ENETCDF, err);
  }
  c.out = o;
  return 0;
}

// Device-to-host copy that keeps the first failure in *e (the output
// writer reports it once the records of the call are written).
static void d2h(void* dst, const void* src, size_t bytes, hipError_t* e) {
  if (*e == hipSuccess) *e = hipMemcpy(dst, src, bytes, hipMemcpyDeviceToHost);
}

int cgfdm_write_outputs(cgfdm_ctx* h, int it) {
  hipError_t herr = hipSuccess;  // first failed copy, reported after the records
  std::string nonfinite;         // first NaN/Inf of the fault record, reported after it
  if (!h) return fail(CGFDM_EPARAM, "null context");
  Ctx& c = h->c;
  if (!c.out) {
    int rc = cgfdm_open_outputs(h);
    if (rc) return rc;
  }
  int skip = std::max(1, c.p.EXPORT_TIME_SKIP);
  if (it % skip) return 0;
  long rec = it / skip;
  NcOut* o = c.out;
  if (hipError_t e = hipStreamSynchronize(c.stream)) return fail(CGFDM_EHIP, hipGetErrorString(e));
  const int nj = c.d.nj, nk = c.d.nk, ni = c.d.ni, ny = c.d.ny;
  const long nxyz = c.d.nxyz, nyz = c.d.nyz, pyz = c.d.pyz;
  const int py = c.d.py, jpad = c.d.jpad;
  (void)nxyz; (void)nyz;
  if (c.p.Rupture) {
    long n2 = (long)nj * nk;
    std::vector<float> buf(n2);
    std::vector<double> db(n2);
    std::vector<int> ib(n2);
    const char* names[7] = {"Vs1", "Vs2", "Tn", "Ts1", "Ts2", "slip", "State"};
    for (int v = 0; v < 7; v++) {
      FieldRef r = field(c, names[v]);
      if (r.dtype == 1) {
        d2h(db.data(), r.ptr, n2 * 8, &herr);
        for (long q = 0; q < n2; q++) buf[q] = (float)db[q];
      } else {
        d2h(buf.data(), r.ptr, n2 * 4, &herr);
      }
      o->fault.put_record(o->fv[v], rec, buf.data());
      // the first non-finite word of the record's fields (reported after the
      // record is written: the file then shows where the run went wrong)
      if (nonfinite.empty() && herr == hipSuccess)
        for (long q = 0; q < n2; q++)
          if (!std::isfinite(buf[q])) {
            nonfinite = std::string(names[v]) + " at node j=" + std::to_string(q % nj) + " k=" + std::to_string(q / nj);
            break;
          }
    }
    d2h(buf.data(), c.f.init_t0, n2 * 4, &herr);
    o->fault.put_var(o->f_init_t0, buf.data());
    d2h(ib.data(), c.f.rup_index_y, n2 * 4, &herr);
    o->fault.put_record(o->fv[7], rec, ib.data());
    d2h(ib.data(), c.f.rup_index_z, n2 * 4, &herr);
    o->fault.put_record(o->fv[8], rec, ib.data());
    d2h(buf.data(), c.f.rup_sensor, n2 * 4, &herr);
    o->fault.put_record(o->fv[9], rec, buf.data());
    // thermal pressurization: the rises on the fault (level 0) as the
    // reference writes them (io_fault.cu:262-266); zeros without it
    for (int v = 0; v < 2; v++) {
      const double* src = v ? c.f.TP_P : c.f.TP_T;
      if (src) {
        d2h(db.data(), src, n2 * 8, &herr);
        for (long q = 0; q < n2; q++) buf[q] = (float)db[q];
      } else {
        std::fill(buf.begin(), buf.end(), 0.0f);
      }
      o->fault.put_record(o->fv[10 + v], rec, buf.data());
    }
    o->fault.sync();
  }
  // slices from the current wavefield
  auto put_slice = [&](NcFile& f, int* vids, int axis, int idx, int d1n, int d2n) {
    std::vector<float> buf((size_t)d1n * d2n);
    std::vector<float> line;
    for (int comp = 0; comp < 3; comp++) {
      if (axis == 2) {  // xy: rows of ny for each i
        line.resize(ny);
        for (int a = 0; a < d1n; a++) {
          d2h(line.data(), c.cur + ((long)(a + 3) * 9 + comp) * pyz + jpad + (long)(idx + 3) * py, ny * 4, &herr);
          std::memcpy(&buf[(long)a * d2n], &line[3], d2n * 4);
        }
      } else if (axis == 1) {  // xz
        std::vector<float> plane(pyz);
        for (int a = 0; a < d1n; a++) {
          d2h(plane.data(), c.cur + ((long)(a + 3) * 9 + comp) * pyz, pyz * 4, &herr);
          for (int b = 0; b < d2n; b++) buf[(long)a * d2n + b] = plane[jpad + (idx + 3) + (long)(b + 3) * py];
        }
      } else {  // yz plane at i
        std::vector<float> plane(pyz);
        d2h(plane.data(), c.cur + ((long)(idx + 3) * 9 + comp) * pyz, pyz * 4, &herr);
        for (int a = 0; a < d1n; a++)
          for (int b = 0; b < d2n; b++) buf[(long)a * d2n + b] = plane[jpad + (b + 3) + (long)(a + 3) * py];
      }
      f.put_record(vids[comp], rec, buf.data());
    }
    f.sync();
  };
  int kz = std::min(std::max(c.p.EXPORT_WAVE_SLICE_Z, 0), nk - 1);
  int jy = std::min(std::max(c.p.EXPORT_WAVE_SLICE_Y, 0), nj - 1);
  int ix = std::min(std::max(c.p.EXPORT_WAVE_SLICE_X, 0), ni - 1);
  put_slice(o->xy, o->wxy, 2, kz, ni, nj);
  put_slice(o->xz, o->wxz, 1, jy, ni, nk);
  put_slice(o->yz, o->wyz, 0, ix, nk, nj);
  if (herr != hipSuccess) return fail(CGFDM_EHIP, std::string("output copy failed: ") + hipGetErrorString(herr));
  if (!nonfinite.empty())
    return fail(CGFDM_ENONFINITE, "non-finite value in fault record " + std::to_string(rec) + " (step " +
                                      std::to_string(it) + "): " + nonfinite + "; the run has become unstable");
  return 0;
}

int cgfdm_close_outputs(cgfdm_ctx* h) {
  if (!h) return fail(CGFDM_EPARAM, "null context");
  Ctx& c = h->c;
  if (!c.out) return 0;
  c.out->fault.close(); c.out->xy.close(); c.out->xz.close(); c.out->yz.close();
  delete c.out;
  c.out = nullptr;
  return 0;
}

}  // extern "C"
