/*
 * cgfdm.h -- C ABI of the MI355X-native CGFDM dynamic-rupture solver.
 *
 * The reference (wenqiang-geophys/dynfd3dcuda) has no library API: its
 * time-step path is a set of void CUDA launchers taking structs of device
 * pointers (srcd/main.cu:49-84), driven by main() (srcd/main.cu:123-580) and
 * configured by params.json (srcd/set_params.cu:56-298).  Each entry point
 * below replaces one of those pieces; the comment names the reference
 * interface it stands in for.  All functions return 0 on success or a
 * positive CGFDM_E* code; the message of the last error is available from
 * cgfdm_last_error().  Pointers are host pointers unless stated otherwise.
 * One context owns one simulation on one GPU; calls on a context are not
 * thread safe.
 */
#ifndef CGFDM_H
#define CGFDM_H

#include <stddef.h>

#ifdef __cplusplus
extern "C" {
#endif

typedef struct cgfdm_ctx cgfdm_ctx;

enum {
  CGFDM_OK = 0,
  CGFDM_EPARAM = 1,    /* bad argument / unknown name                        */
  CGFDM_EFILE = 2,     /* params file missing or unreadable (ref: exit(-1))  */
  CGFDM_EJSON = 3,     /* malformed JSON (ref: exit(-1))                     */
  CGFDM_ENETCDF = 4,   /* netCDF input missing or malformed (ref: exit(2))   */
  CGFDM_ECFL = 5,      /* DT >= 1.3 hmin / vpmax (ref: MPI_Abort 110)        */
  CGFDM_EHIP = 6,      /* HIP runtime / allocation failure                   */
  CGFDM_ESTATE = 7,    /* call out of order (e.g. step before setup)         */
  CGFDM_ESIZE = 8,     /* buffer size does not match the field               */
  CGFDM_ENONFINITE = 9 /* NaN or Inf in an output record (extension: the
                          reference writes such records without notice)      */
};

/* Message of the last error on this thread (never NULL). */
const char* cgfdm_last_error(void);

/* Parse a params.json file: defaults then JSON keys, exactly the key set and
 * defaults of get_params (srcd/set_params.cu:56-298), plus optional runtime
 * forms of the Makefile feature macros ("usePML", "useCerjan", "FreeSurface",
 * "NormalRefine", "Rupture") with the Makefile defaults (Makefile:36-121).
 * Input paths in the JSON are resolved relative to the current directory, as
 * the reference does. */
int cgfdm_create(const char* params_json_path, cgfdm_ctx** out);
/* Same from a JSON text already in memory. */
int cgfdm_create_from_string(const char* json_text, cgfdm_ctx** out);
/* Override one parameter by its JSON key (numbers only; "Fault_grid0".."3"
 * and "Asp_grid0".."3" address array elements).  Must precede setup. */
int cgfdm_set_param(cgfdm_ctx* ctx, const char* key, double value);
int cgfdm_get_param(cgfdm_ctx* ctx, const char* key, double* value);
int cgfdm_destroy(cgfdm_ctx* ctx);

/* Host arrays that replace the netCDF/text inputs (tests and embedding).
 * Fault_geometry x,y,z [nk][nj] (srcd/coord.cu:71-151). */
int cgfdm_set_fault_geometry(cgfdm_ctx* ctx, const float* x, const float* y, const float* z, size_t n);
/* Full 3-D coordinates [3][nx][nz][ny] with ghosts (interior used; extension
 * for topography that varies along the fault normal). */
int cgfdm_set_coord3d(cgfdm_ctx* ctx, const double* xyz, size_t n);
/* Surface topography h [ni][nj] in metres (extension, no counterpart in
 * the reference, whose grid translates the fault surface along x so the
 * surface height cannot vary with x): z(i,j,k) += h(i,j) * k/(NZ-1). */
int cgfdm_set_topography(cgfdm_ctx* ctx, const float* h, size_t n);
/* Media1D layers (srcd/media.cu:63-138). */
int cgfdm_set_media1d(cgfdm_ctx* ctx, int nlayer, const float* depth, const float* rho, const float* vp, const float* vs);
/* Media3D (srcd/media.cu:140-204): vp, vs, rho on the interior grid,
 * [NX][NZ][NY] (the variables of the Media3D netCDF file); overrides the
 * bimaterial and layered media. */
int cgfdm_set_media3d(cgfdm_ctx* ctx, const float* vp, const float* vs, const float* rho, size_t n);
/* Fault_init_stress variable by its netCDF name: Tx Ty Tz dTx dTy dTz mu_s
 * mu_d Dc C0 a b L Vw State, and TP_hy for Friction_type 3, [nk][nj]
 * (srcd/init.cu:904-1022). */
int cgfdm_set_init_stress(cgfdm_ctx* ctx, const char* name, const float* data, size_t n);

/* Build grid, metric, media, PML profiles, Cerjan profiles, fault
 * coefficients, fault initial state and free-surface maps; allocate device
 * memory and upload.  Replaces main.cu:238-396 (alloc_*, construct_coord,
 * cal_metric, init_media*, abs_init, init_cerjan_host, init_fault_coef,
 * init_fault, nc_read_init_stress, init_wave_free, init_wave_3d).  Reads the
 * netCDF/text inputs named in the JSON unless the arrays were set above.
 * Fails with CGFDM_ECFL under the reference's CFL rule (main.cu:323-345). */
int cgfdm_setup(cgfdm_ctx* ctx);

/* One time step it: cal_rup_sensor + 4 RK stages (main.cu:439-516). */
int cgfdm_step(cgfdm_ctx* ctx, int it);
/* One RK stage irk (0..3) of step it with the FD flag triple of that stage
 * (main.cu:448-516): state_rk, trial_sw|trial_rs, fault2wave, wave_deriv,
 * abs_deriv_{x,y,z}, fault_dvelo, fault_dstrs_{f,b}, wave_rk, fault_rk,
 * fault2wave, apply_cerjan. */
int cgfdm_stage(cgfdm_ctx* ctx, int it, int irk);
/* Steps it0 .. it0+nt-1; writes output records when write_outputs != 0
 * (main.cu:420-535). */
int cgfdm_run(cgfdm_ctx* ctx, int it0, int nt, int write_outputs);
/* Wait for all work queued on the context's stream. */
int cgfdm_synchronize(cgfdm_ctx* ctx);

/* Per-kernel entry points on the current stage buffers, for parity tests.
 * wave_deriv + PML + RK combine + Cerjan for the whole volume with the fault
 * slab diverted (wave.cu:8-494, pml{x,y,z}.cu, rk4.cu:154-257,
 * cerjan.cu:166-200). */
int cgfdm_wave_stage(cgfdm_ctx* ctx, int irk, int fx, int fy, int fz);
/* Volume RHS only (no RK): fills the "hW" field (wave_deriv_cu, wave.cu:8-494,
 * plus abs_deriv_{x,y,z} contributions when PML is on). */
int cgfdm_wave_rhs(cgfdm_ctx* ctx, int fx, int fy, int fz);

/* The FD flag triple of stage irk of step it: row it % 8 of the 8-row table
 * (main.cu:86-94, 440-442), negated on odd stages (main.cu:514). */
int cgfdm_stage_flags(int it, int irk, int* fx, int* fy, int* fz);

/* Queue later launches on the caller's HIP stream (a hipStream_t passed as
 * void*); NULL returns to the context's own stream.  Work already queued on
 * the previous stream is waited for first.  The reference launches on the
 * default stream with a device-wide wait per stage (main.cu:505-511). */
int cgfdm_set_stream(cgfdm_ctx* ctx, void* stream);

/* The fault phases of stage irk of step it, for callers that drive a stage a
 * phase at a time (the reference's per-launcher interface, main.cu:49-84).
 * cgfdm_fault_trial, cgfdm_wave_stage (with cgfdm_stage_flags' triple),
 * cgfdm_fault_deriv and cgfdm_fault_update, called in that order, queue what
 * cgfdm_stage queues and give the same result bit for bit.  Each returns
 * CGFDM_ESTATE without a fault (Rupture = 0).
 *   trial:  state_rk + trial_sw | trial_rs (Newton) + fault2wave
 *           (main.cu:455-468; rk4.cu:791-938, trial_slipweakening.cu:23-732,
 *           trial_ratestate.cu:20-893, transform.cu:224-403; the wave2fault
 *           launched before them returns at once, transform.cu:18);
 *   deriv:  fault_dvelo + fault_dstrs_{f,b} (main.cu:486-491;
 *           fault_dvelo.cu:17-389, fault_dstrs_f.cu:17-1254);
 *   update: RK of the planes i0-3..i0+3 diverted from the volume stage +
 *           fault_rk + fault_rk_stage2 + fault2wave + Cerjan on those planes
 *           (main.cu:495-508; rk4.cu:154-266, 466-789). */
int cgfdm_fault_trial(cgfdm_ctx* ctx, int it, int irk);
int cgfdm_fault_deriv(cgfdm_ctx* ctx, int it, int irk);
int cgfdm_fault_update(cgfdm_ctx* ctx, int it, int irk);
/* cal_rup_sensor (rup_sensor.cu:7-97), run before the stages of a step
 * (main.cu:444). */
int cgfdm_rup_sensor(cgfdm_ctx* ctx);
/* thermpress (thermpress.cu:22-233): the temperature and pore-pressure
 * rises normal to the fault, once per step before the stages when
 * Friction_type is 3 (main.cu:445-447); CGFDM_ESTATE for other types. */
int cgfdm_thermpress(cgfdm_ctx* ctx);
/* smooth_gauss_volume (smooth.cu:13-367), after the stages of a step when
 * FaultSmooth is set (main.cu:516-522); CGFDM_ESTATE when it is not. */
int cgfdm_fault_smooth(cgfdm_ctx* ctx);

/* Host copies of named fields (cgfdm_field_info reports each one's element
 * count and type, and rejects an unknown name): "W" (current
 * wavefield, 9 x nx*ny*nz, float), "hW", "M" (13 fields; MetricBF16 = 0),
 * "M_bf16" (MetricBF16 = 1: 13 bfloat16 fields; = 2: 5 int32 fields, each
 * two bfloat16 metric deviations, field 2p in the low half of field p, word =
 * deviation + "M_ref"[field], the high half of field 4 the point's media
 * palette entry), "M_jac_bf16" (MetricBF16 = 2: the Jacobian's deviation
 * from "M_ref"[9] in the low half, int32), "M_media" (MetricBF16 = 2:
 * lambda, mu, rho, float), "M_palette" (MetricBF16 = 2 with a media palette:
 * lambda, mu, rho, 0 of each entry, float), "M_skip" (MetricBF16 = 2: bit p
 * set when both halves of 32-bit metric word p are +-0 at every point, so
 * the stage kernels do not load it; int32, read only), "M_ref" (10 floats,
 * read only),
 * "FW" (split nodes),
 * fault 2-D outputs "Vs1" "Vs2" "Tn" "Ts1" "Ts2" "slip" "State" "init_t0"
 * "rup_sensor" ..., PML aux "PMLx1".."PMLz2".  dtype: 0 float32, 1 float64,
 * 2 int32, 3 bfloat16 (raw 16-bit words). */
int cgfdm_field_info(cgfdm_ctx* ctx, const char* name, size_t* count, int* dtype);
int cgfdm_get_field(cgfdm_ctx* ctx, const char* name, void* dst, size_t bytes);
int cgfdm_put_field(cgfdm_ctx* ctx, const char* name, const void* src, size_t bytes);
/* Dimensions: ni nj nk nx ny nz i0 NT. */
int cgfdm_get_dims(cgfdm_ctx* ctx, int dims[8]);

/* Output files in OUT/ (io_fault.cu:16-275, io_wave_{xy,xz,yz}.cu): define
 * them, write record it/EXPORT_TIME_SKIP, close. */
int cgfdm_open_outputs(cgfdm_ctx* ctx);
int cgfdm_write_outputs(cgfdm_ctx* ctx, int it);
int cgfdm_close_outputs(cgfdm_ctx* ctx);

/* Rate-and-state Newton statistics since setup (extension): number of node
 * solves, total iterations, and solves that fell back to the plain guess. */
int cgfdm_newton_stats(cgfdm_ctx* ctx, double* solves, double* iterations, double* fallbacks);

/* GPU timing per kernel class (extension, no counterpart in the reference):
 * HIP events recorded on the solver's stream around every launch while
 * enabled.  cgfdm_profile(ctx, 1) clears and starts recording, 0 stops.
 * cgfdm_profile_get returns, for kernel class `name` ("stage_interior",
 * "stage_band", "stage_surface", "pml_aux", "fault_pre", "fault_rs", "fault_deriv",
 * "fault_post", "rup_sensor"), the summed GPU time in ms and the launch count
 * since recording started (it synchronizes the stream), and the grid points
 * one launch of a stage-kernel class updates (0 for the others). */
int cgfdm_profile(cgfdm_ctx* ctx, int enable);
/* Restrict the timing to a comma-separated list of kernel classes (NULL or
 * "" for all, the default); the other launches then carry no events. */
int cgfdm_profile_classes(cgfdm_ctx* ctx, const char* names);
int cgfdm_profile_get(cgfdm_ctx* ctx, const char* name, double* total_ms, long* launches, long* points_per_launch);
/* ADE-PML work of one launch of kernel class `name` (extension, no
 * counterpart in the reference; for byte accounting): the (point, face)
 * pairs whose auxiliary value it reads for the -aux/beta term of the stage
 * (a point on two or three absorbing faces counts once per face), and those
 * whose auxiliary ODE step (pmlx.cu:126-210 and the y/z variants) it takes;
 * 0 for classes without such work.  Needs cgfdm_setup. */
int cgfdm_profile_aux(cgfdm_ctx* ctx, const char* name, long* read_pairs, long* update_pairs);

/* Device-side check of an FP32 field (extension, no counterpart in the
 * reference): *nonfinite = NaN/Inf words, *ghost_nonzero = nonzero words in
 * the three ghost layers of a volume field ("W", "hW"; 0 for 2-D and fault
 * fields, which have no ghost layers in their layout here or are checked
 * whole), *maxabs = largest finite magnitude of the interior.  Reads the
 * field once on the device; no host copy of the volume. */
int cgfdm_check_field(cgfdm_ctx* ctx, const char* name, long long* nonfinite, long long* ghost_nonzero, double* maxabs);

/* Print the parameter block of the reference's print_params
 * (set_params.cu:299-388) to stdout, as the reference does after reading
 * params.json (main.cu:156). */
int cgfdm_print_params(cgfdm_ctx* ctx);

/* Streaming-bandwidth probe (extension, no counterpart in the reference;
 * measurement only): a kernel reads `nread` and writes `nwrite` device arrays
 * of `bytes_per_array` once each with non-temporal 16-byte accesses, `reps`
 * times; *gbps is (nread + nwrite) * bytes_per_array * reps / time in GB/s.
 * Forms: 1:0, 1:1, 2:1, 3:1, 2:2.  The 2:1 form has the stage kernels'
 * read:write mix, so it is their practical HBM ceiling. */
int cgfdm_stream_probe(int nread, int nwrite, size_t bytes_per_array, int reps, double* gbps);

/* Learned initial guess of the rate-and-state Newton solve (extension, no
 * counterpart in the reference).  With NewtonMode 2 the solve of
 * trial_ratestate.cu:686-702 starts from the better (smaller residual) of the
 * plain guess x0 = asinh(e^(psi/a) V/(2 V0)) and a 4-64-64-1 network's
 * prediction of log10 V, and stops on the residual rule of NewtonMode 1 with
 * the same fallback.  The network is a float32 parameter blob of
 * CGFDM_MLP_PARAMS values:
 *   xmean[4] xscale[4] ymean yscale W1[64][4] b1[64] W2[64][64] b2[64] w3[64] b3
 * Its inputs per node are the four guess features (field "RS_samples" planes
 * 0-3), its output log10 of the slip rate (plane 4). */
#define CGFDM_MLP_PARAMS 4555
/* Network used by NewtonMode 2 (before or after setup; replaces the file
 * named by the JSON key NewtonMLP). */
int cgfdm_set_newton_mlp(cgfdm_ctx* ctx, const float* params, long n);
/* Evaluate the network on the GPU for n feature rows feat[n][4] (host
 * arrays): mode 0 with the wavefront matrix-instruction kernel the solver
 * uses, mode 1 with the scalar kernel. */
int cgfdm_mlp_forward(const float* params, const float* feat, long n, float* out, int mode);
/* cgfdm_mlp_forward that allocates nothing (no counterpart in the
 * reference): for callers that cannot allocate device memory (code under
 * HIP graph capture, tight serving loops).  params (nparams =
 * CGFDM_MLP_PARAMS floats), feat[n][4] and out[n] are DEVICE arrays of the
 * caller; the packed network goes to a static device scratch of 32 KiB per
 * GPU (held until the process ends).  Queued on `stream` (a hipStream_t, or
 * NULL for the null stream) without synchronising.  Rejects a network of
 * another size and n outside 0..2^31-1 (CGFDM_ESIZE).  Every call on a GPU
 * shares that GPU's scratch, so calls that overlap on one GPU race. */
int cgfdm_mlp_forward_noalloc(const float* params, long nparams, const float* feat, long n, float* out, int mode,
                              void* stream);
/* Gradient (CGFDM_MLP_PARAMS floats, zero for the standardisation entries)
 * and value of the loss mean_s (o_s - (t_s - ymean)/yscale)^2 on the GPU. */
int cgfdm_mlp_grad(const float* params, const float* feat, const float* target, long n, float* grad, float* loss);
/* Train in place on the GPU with Adam (step size decaying from lr to lr/20
 * over the epochs on a cosine, minibatches of `batch` samples in a seeded
 * random order per epoch); loss_hist[e] is the mean loss over epoch e's
 * batches. */
int cgfdm_mlp_train(float* params, const float* feat, const float* target, long n, int epochs, int batch, float lr,
                    unsigned seed, float* loss_hist);

#ifdef __cplusplus
}
#endif
#endif /* CGFDM_H */
