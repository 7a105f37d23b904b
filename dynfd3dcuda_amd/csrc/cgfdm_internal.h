// Internal declarations of the MI355X CGFDM solver (not part of the C ABI).
#pragma once
#include <hip/hip_runtime.h>
#include <cstdint>
#include <string>
#include <vector>
#include <map>

namespace cgfdm {

// bins of the per-wavefront Newton iteration histogram (the last bin holds
// counts of that many iterations or more)
constexpr int kNewtonHistBins = 32;

// ---------------------------------------------------------------------------
// Parameters: the key set and defaults of get_params (set_params.cu:56-298)
// plus runtime forms of the reference's compile-time feature macros.
struct Params {
  int NX = 100, NY = 90, NZ = 80;
  int PX = 1, PY = 1, PZ = 1, igpu = 0;
  double TMAX = 1.0, DT = 0.01, DH = 100.0;
  int INPORT_GRID_TYPE = 0, INPORT_STRESS_TYPE = 0, Friction_type = 0;
  int EXPORT_TIME_SKIP = 1, EXPORT_WAVE_SLICE_X = 0, EXPORT_WAVE_SLICE_Y = 0, EXPORT_WAVE_SLICE_Z = 0;
  int PML_N = 20;  // PML defaults are re-set after the struct defaults (set_params.cu:264-267)
  double PML_velocity = 3000.0, PML_bmax = 3.0, PML_fc = 1.0;
  int DAMP_N = 20;
  double mu_s = 0.677, mu_d = 0.525, mu_0 = 0.0, Dc = 0.4, C0 = 0.0;
  double f0 = 0, V0 = 0, Vini = 0, fw = 0, L = 0, smooth_load_T = 0;
  int Fault_grid[4] = {0, 0, 0, 0};
  int Asp_grid[4] = {0, 0, 0, 0};
  double vp1 = 6000, vs1 = 3464, rho1 = 2670;
  double bi_vp1 = 6000, bi_vs1 = 3464, bi_rho1 = 2670;
  double bi_vp2 = 6000, bi_vs2 = 3464, bi_rho2 = 2670;
  double viscosity = 0, RupThres = 0;
  int TP_n = 1;
  std::string OUT = "output", Fault_geometry = "Fault_geometry.nc", Fault_init_stress = "Fault_init_stress.nc";
  std::string Media1D, Media3D;
  // Makefile feature macros as runtime flags (Makefile:36-121 defaults)
  int usePML = 1, useCerjan = 0, FreeSurface = 1, NormalRefine = 1, Rupture = 1;
  int FaultSmooth = 0;  // Makefile:104: Gaussian smoothing of the fault tractions once per step
  // extensions
  int MetricPalette = 1;  // MetricBF16 = 2: lambda, mu, rho as entries of a palette of the distinct media (a 1-D layered
                          // or two-half model of at most 64 media) instead of FP32 planes; 0: always the planes
  int MetricSkipConstant = 1;  // MetricBF16 = 2: coefficient words equal to their reference on the whole grid are not
                               // loaded by the stage kernels (0: every word loaded)
  int MetricBF16 = 0;   // 1: metric + media stored in bfloat16; 2: metric as bfloat16 deviations from the Cartesian grid, media exact (FP32 compute)
  int NewtonMode = 0;   // 0: reference NRsearch; 1: residual tolerance + plain fallback
  double NewtonTol = 1e-10;
  double NewtonMLPGate = 2.0;  // learned guess consulted where the plain guess's relative residual exceeds this
  int NewtonFallbackIter = 20;
  std::string NewtonMLP; // path of trained MLP weights for the Newton guess
  int WaveIChunk = 0;   // planes per block of the stage kernel (0: automatic)
  int StageStreams = 0;  // 1: interior and free-surface launches of a stage on side streams
  int WaveTallInterior = 0;  // 1: interior tiles 64 x 16 in blocks of 1024 threads
  int WaveInteriorKernel = -1;  // interior tiles: -1 automatic; 0 general kernel; pipelined kernel: 1 64 x 8 with
                                // an LDS plane ring, 2 64 x 6, 3 64 x 4, 4 64 x 8 double-buffered
  int WaveBandKernel = -1;  // absorbing-band tiles: 0 general kernel; pipelined kernel: 1 64 x 8 tiles, 2 64 x 4 tiles;
                            // -1: 2 with 64 x 4 interior tiles, else 1
  int WaveSurfaceKernel = 1;  // free-surface tiles: 0 general kernel; 1 pipelined kernel
  int PmlAuxFused = 1;     // 1: the pipelined band tiles advance the ADE-PML auxiliary fields of their
                           // points on one face; 0: a kernel of its own advances every band point's
  int WaveBufferSkew = 16384;  // words (a multiple of 32) by which buffer q of S, A, B, T starts q times
                               // further into its allocation
  int WaveBandIChunk = 0;  // planes per block of the band tiles (0: automatic)
  int WaveTopIChunk = 0;  // planes per block of the free-surface tiles (0: automatic)
  int WaveSurfaceTiles = 1;  // free-surface tiles = the top TK rows, own line-aligned j tiles (0: earlier layout)
  int WaveBandRound = 0;  // 1: low j band rounded up to whole 64-wide tiles (0: band width exactly)

  // derived (set_params.cu:286-295)
  int ni = 0, nj = 0, nk = 0, nx = 0, ny = 0, nz = 0, NT = 0;
  void derive();
  int set(const std::string& key, double v);   // numeric keys
  int get(const std::string& key, double* v) const;
};

int parse_params_json(const std::string& text, Params& p, std::string& err);

// ---------------------------------------------------------------------------
// Device-side views.  Volume arrays use the reference index
// pos = j + k*ny + i*ny*nz over padded dims (3 ghost layers per side).
// media palette entries (MetricBF16 = 2) at most: 1 KB of LDS per block in
// the stage kernels (4 KB for 256 entries took the 32 x 8 narrow band tiles
// to one block per CU)
constexpr int kMaxMediaPalette = 64;

struct Dims {
  int ni, nj, nk, nx, ny, nz, i0;
  long nxyz, nyz;
  // Device volume arrays (wavefields, metric, fault slab) are plane-
  // interleaved with a padded row: word (field c, padded plane i, padded
  // row k, padded column j) of an array with nf fields per plane sits at
  // (i*nf + c)*pyz + k*py + jpad + j.  py is a multiple of 32 words, not a
  // power of two, and jpad puts the first interior stencil tile on a
  // 128-byte line.
  int py, jpad;
  long pyz;
};

// First column of the interior box of the stage kernel (the absorbing band
// below it is handled by the edge instantiation).
int interior_lo_j(const Params& p);

// Fault state on the device (2-D arrays [nk][nj] unless noted).
struct FaultDev {
  float *W, *mW, *hW, *tW;         // split nodes [9][2][nz][ny]
  float *T11, *T12, *T13;          // [7][nz][ny]
  float *mT11, *mT12, *mT13;       // plane 3 used
  float *hT11, *hT12, *hT13;       // [nz][ny]
  float *rho_f;                    // [2][nz][ny]
  float *Dmat;                     // [2][3][3][nz][ny][9]  D_ab for both sides (x J)
  float *matT1toVx, *matVytoVx, *matVztoVx;   // [2][nz][ny][9]
  float *matVx2Vz_f, *matVy2Vz_f;              // [2][ny][9]
  float *matP2Mf, *matM2Pf;                    // [3][ny][9]
  float *vec_n, *vec_s1, *vec_s2;              // [nz][ny][3]
  float *jacvec, *jac_i0;                      // [nz][ny]
  float *Tn, *Ts1, *Ts2, *tTn, *tTs1, *tTs2;
  float *str_init_x, *str_init_y, *str_init_z;
  float *T0x, *T0y, *T0z, *dT0x, *dT0y, *dT0z;
  float *str_peak, *Dc, *C0, *a, *b, *L, *Vw;
  double *slip, *mslip, *tslip, *State, *mState, *tState;
  float *hslip, *Vs1, *Vs2, *init_t0, *friction, *rsD_y, *rsD_z, *rup_sensor;
  int *init_t0_flag, *rup_index_y, *rup_index_z, *united, *faultgrid, *flag_rup;
  // [5][nk][nj] per-node counts (solves, iterations, fallbacks, solves started
  // from the learned guess, solves whose wavefront evaluated the network),
  // then kNewtonHistBins counts of wavefronts by their largest iteration count
  unsigned int* newton_stats;
  float *Ws;                         // [9][2][nz][ny] smoothed T2*/T3* (FaultSmooth)
  double *TP_T = nullptr, *TP_P = nullptr;  // Friction_type 3: [TP_n][nk][nj] temperature and pore-pressure rise
  double *TP_hy = nullptr;                  // [nk][nj] hydraulic diffusivity
  int TP_n = 1;
  double *rs_carry;                  // [10][nk][nj] rate-state Newton inputs handed from fault_pre to fault_rs
  float *rs_carryf;                  // [9][nk][nj] normal (3), jacvec, guess features (4), solved log10 V
  float *smooth_w;                   // [2][4][7][7] half-volume Gaussian weights (FaultSmooth)
};

struct PmlDev {
  float *A[3], *B[3], *D[3];  // 1-D profiles along x (ni), y (nj), z (nk); D stored as d/beta
  float *Y[6], *S[6], *T[6];  // aux per face x1 x2 y1 y2 z1 z2: [9][pni][pnk][pnj]
  int on[6];
  int N;
};

// ---------------------------------------------------------------------------
// Optional GPU timing of each kernel class with HIP events on the solver's
// stream (off unless enabled; events come from a pool created on enable).
enum ProfKernel {
  PK_STAGE_INTERIOR, PK_STAGE_BAND, PK_STAGE_SURFACE, PK_PML_AUX,
  PK_FAULT_PRE, PK_FAULT_DERIV, PK_FAULT_POST, PK_RUP_SENSOR, PK_FAULT_RS, PK_COUNT
};
extern const char* const kProfNames[PK_COUNT];
struct Profiler {
  bool on = false;
  unsigned mask = ~0u;  // kernel classes timed while on (bit k: class k)
  bool open = false;    // the last begin() recorded
  std::vector<hipEvent_t> pool;
  std::vector<std::pair<int, size_t>> marks;  // (kernel class, index of its start event)
  size_t used = 0;
  void begin(hipStream_t s, int k) {
    open = on && ((mask >> k) & 1u);
    if (!open) return;
    if (used + 2 > pool.size()) grow(used + 2048);
    (void)hipEventRecord(pool[used], s);  // timing is best effort
    marks.push_back({k, used});
    used += 2;
  }
  void end(hipStream_t s) {
    if (open && !marks.empty()) (void)hipEventRecord(pool[marks.back().second + 1], s);
    open = false;
  }
  void grow(size_t n) {
    while (pool.size() < n) {
      hipEvent_t e;
      if (hipEventCreate(&e) != hipSuccess) return;
      pool.push_back(e);
    }
  }
  void reset() { marks.clear(); used = 0; }
  ~Profiler() { for (hipEvent_t e : pool) (void)hipEventDestroy(e); }
};

struct Ctx {
  Params p;
  Dims d;
  bool setup_done = false;
  hipStream_t stream = nullptr;      // where every launch goes
  hipStream_t own_stream = nullptr;  // created at setup, destroyed with the context
  hipStream_t side[2] = {nullptr, nullptr};  // StageStreams: interior, free surface
  hipEvent_t fence[3] = {nullptr, nullptr, nullptr};  // stage input ready, interior done, surface done
  int device = 0;
  // host inputs
  std::vector<float> gx, gy, gz;
  std::vector<double> coord3d;
  std::vector<float> topo;       // [ni][nj] surface height (extension)
  std::vector<float> m1d_depth, m1d_rho, m1d_vp, m1d_vs;
  std::vector<float> m3d_vp, m3d_vs, m3d_rho;  // Media3D [ni][nk][nj]
  std::map<std::string, std::vector<float>> stress_in;
  // device volume buffers
  float *S = nullptr, *A = nullptr, *B = nullptr, *Tacc = nullptr;
  float *cur = nullptr;         // buffer holding W for the last completed stage
  float *hW = nullptr;          // RHS buffer for cgfdm_wave_rhs (tests)
  void* met = nullptr;          // 13 fields, float or bf16
  float *slabH = nullptr;       // [7][9][nz][ny] fault-slab RHS
  float *matVx2Vz = nullptr, *matVy2Vz = nullptr;  // [nx][ny][9]
  float *cj[3] = {nullptr, nullptr, nullptr};      // Cerjan 1-D profiles (nx, ny, nz incl. ghosts)
  void* wave_cold = nullptr;    // device copy of the off-hot-path pointer table
  void* wave_work = nullptr;    // stage-kernel work list (two int4 per block)
  float* trac_top = nullptr;    // free surface: traction vectors of the six top rows [nx][6][9][ny]
  float* img_top = nullptr;     // free surface: image velocity RHS of the three top rows [3][3][ni][nj]
  int wave_nwork = 0;                  // all work items
  bool wave_tall = false;           // interior tiles 64 x 16 (WaveTallInterior)
  bool wave_pipe = false;           // pipelined interior kernel (WaveInteriorKernel)
  bool wave_ring = false;           // its LDS plane ring (WaveInteriorKernel 1)
  bool wave_band_pipe = false;      // pipelined kernel on the absorbing-band tiles (WaveBandKernel)
  int wave_tjn = 32;                // columns of the narrow band tiles (20 or 32)
  int wave_band_tk = 8;             // rows of the 64-wide band tiles (WaveBandKernel 2: 4)
  bool pml_fused = false;           // band tiles advance their ADE-PML auxiliary fields (PmlAuxFused)
  bool wave_surf_pipe = false;      // pipelined kernel on the free-surface tiles (WaveSurfaceKernel)
  int wave_ktop = 0;                // first row of the free-surface tiles (nk without them)
  int* pml_boxes = nullptr;         // boxes of the auxiliary-update kernel (8 ints each), on the device
  int pml_nbox = 0, pml_blocks = 0;
  int wave_tki = 8;                 // rows of an interior tile
  float met_ref[10] = {};           // metric words' reference values (MetricBF16 = 2)
  // MetricBF16 = 2: the media palette on the device (lambda, mu, rho, 0 per
  // entry; null when the media have none) and its entries
  float* met_pal = nullptr;
  int met_npal = 0;
  // MetricBF16 = 2: bit p set when both halves of 32-bit word p are +-0 at
  // every stored point (the stage kernels then skip loading it)
  unsigned met_skip = 0;
  // bfloat16 storage: FP32 metric and media of the planes i0-6..i0+6 that the
  // fault kernels read, and that copy's base moved back by its first plane
  float* met_fault = nullptr;
  const float* met_fault_base = nullptr;
  int wave_nkind[4] = {0, 0, 0, 0};    // items per list: interior, bands, narrow bands, free-surface tiles
  long wave_points[3] = {0, 0, 0};  // grid points per launch of each stage-kernel kind
  // ADE-PML (point, face) pairs per launch of each stage-kernel kind and of
  // pml_aux_kernel (index 3): whose auxiliary value it reads for the -aux/beta
  // term, and whose auxiliary ODE step it takes
  long aux_read_pairs[4] = {0, 0, 0, 0}, aux_update_pairs[4] = {0, 0, 0, 0};
  void* mlp_dev = nullptr;          // MlpDev of the learned Newton guess (NewtonMode 2)
  std::vector<float> mlp_blob;      // its parameters, kept for setup
  Profiler prof;
  PmlDev pml{};
  FaultDev f{};
  unsigned long long* check_counts = nullptr;  // counters of cgfdm_check_field (3 words)
  std::vector<void*> allocations;
  // host copies kept from setup
  std::vector<float> out_coord[4];  // coordinates x, y, z of the output files' grids: fault plane [3][nk][nj],
                                    // wave slices xy [3][ni][nj], xz [3][ni][nk], yz [3][nk][nj]
  double hmin = 0, vpmax = 0;
  // outputs
  struct NcOut* out = nullptr;
};

// setup (host, double precision)
int host_setup(Ctx& c, std::string& err);

// kernel launchers (kernels.hip)
struct StageBufs {
  const float* Win; float* Wout; const float* S; float* T;
};
void launch_wave_stage(Ctx& c, const StageBufs& b, int irk, int fx, int fy, int fz, bool rhs_only, float* rhs_out);
void launch_fault_pre(Ctx& c, const float* Win, int it, int irk, int fx, int fy, int fz);
void launch_fault_deriv(Ctx& c, const float* Win, int fx, int fy, int fz);
void launch_fault_post(Ctx& c, const StageBufs& b, int irk);
void launch_rup_sensor(Ctx& c);
void launch_thermpress(Ctx& c);
void launch_fault_smooth(Ctx& c);
void launch_fault_rs(Ctx& c, const float* Win, int it, int irk, int fx, int fy, int fz);
// learned Newton guess (mlp_train.hip)
int mlp_upload(Ctx& c, const float* blob, long n);  // nonzero on error
int mlp_forward_host(const float* blob, const float* feat, long n, float* out, int mode, std::string& err);
int mlp_grad_host(const float* blob, const float* feat, const float* target, long n, float* grad, float* loss,
                  std::string& err);
int mlp_forward_noalloc_host(const float* P, const float* feat, long n, float* out, int mode, hipStream_t s,
                             std::string& err);
int field_check_host(Ctx& c, const float* ptr, int nf, int f0, int fn, long flat_n, long long* nonfinite,
                     long long* ghost_nonzero, double* maxabs, std::string& err);
int stream_probe_host(int nread, int nwrite, size_t bytes_per_array, int reps, double* gbps, std::string& err);
int mlp_train_host(float* blob, const float* feat, const float* target, long n, int epochs, int batch, float lr,
                   unsigned seed, float* loss_hist, std::string& err);
void launch_fill(Ctx& c, float* p, long n, float v);
int upload_wave_cold(Ctx& c);  // nonzero on a HIP error

}  // namespace cgfdm
