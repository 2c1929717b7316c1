/* C ABI of the MI355X (gfx950) elastic FWI propagation library, libefwi.so.
 *
 * Drop-in boundary for the reference's native path (guoketing/EFWI-3D,
 * Ops/FWI/Src_3d). Plain C types only; every entry point returns 0 on success
 * and a negative code on error, with the message in efwi_last_error(). No
 * entry point exits the process.
 *
 * Volumes use the torch layout [nz][nx][ny] (y fastest) of the PML-padded
 * grid. Seismograms use [3][nrec][nt] with component order x, y, z, the
 * layout of the reference's Shot_x/Shot_y/Shot_z files. "Device" pointers are
 * HIP device memory of the context's GPU; a hipStream_t is passed as void*.
 */
#ifndef EFWI_H_
#define EFWI_H_

#ifdef __cplusplus
extern "C" {
#endif

#define EFWI_OK 0
#define EFWI_ERR_ARG (-1)
#define EFWI_ERR_HIP (-2)
#define EFWI_ERR_IO (-3)
#define EFWI_ERR_STATE (-4)
#define EFWI_ERR_NOMEM (-5)

/* calc_id of the reference (Src/libCUFD.cu:25-28, Src/Parameter.cpp:132-144). */
#define EFWI_CALC_MISFIT 0
#define EFWI_CALC_GRADIENT 1
#define EFWI_CALC_SYNTHETIC 2

typedef struct efwi_ctx efwi_ctx;

/* Grid, time stepping and options. Replaces the para JSON that Parameter
 * reads (Src/Parameter.cpp:41-184): nz/nx/ny are padded sizes, f0 only sets
 * the C-PML alpha profile (Src/utilities.cu:215). */
typedef struct {
  int nz, nx, ny;     /* padded grid */
  int npml;           /* nPoints_pml */
  int nt;             /* nSteps */
  float dz, dx, dy;   /* m */
  float dt;           /* s */
  float f0;           /* Hz */
  int formulation;    /* 0 = reference (default), 1 = consistent (quirks corrected) */
  int if_win;         /* per-trace windows and weights (Src/utilities.cu:524-576) */
  int device;         /* HIP device ordinal */
  int memory_mode;    /* -1 = automatic (0 when its ring fits in 90 % of free device memory, else 1);
                         0 = boundary ring of all nt steps in HBM (the reference's scheme);
                         1 = the same ring rebuilt per segment from checkpoints of the forward
                             state every checkpoint_every steps (bit-identical gradient, memory
                             ~ sqrt(nt) instead of nt, one extra forward pass) */
  int checkpoint_every; /* memory_mode 1 segment length; 0 = ceil(sqrt(nt - 1)) */
} efwi_config;

/* One shot. Replaces one "shot<i>" object of the survey JSON
 * (Src/Src_Rec.cu:74-206); coordinates are on the padded grid (the JSON value
 * plus npml). Host pointers. */
typedef struct {
  int src_z, src_x, src_y;
  const float* stf;       /* nt samples, untapered (the row stf[shot_id]) */
  int nrec;
  const int* rec;         /* [nrec][3] (z, x, y) */
  const float* win_start; /* [nrec] seconds, if_win only */
  const float* win_end;   /* [nrec] seconds, if_win only */
  const float* weights;   /* [nrec] or NULL (= 1), if_win only */
  float src_weight;       /* if_win only (1 when absent from the JSON) */
} efwi_shot;

/* Library / error reporting. */
int efwi_version(void);
const char* efwi_last_error(void);

/* Context lifetime. Allocates the wavefield, C-PML and gradient state on the
 * device; no reference counterpart (cufd allocates and frees per call,
 * Src/libCUFD.cu:154-196,756-795). */
int efwi_create(const efwi_config* cfg, efwi_ctx** out);
void efwi_destroy(efwi_ctx* ctx);

/* Model in the op's units (lambda, mu in the Mu/Lambda units of FWI_3d, rho
 * in g/cc), [nz][nx][ny]. Copies lambda and mu and derives the buoyancy
 * (Src/Model.cu:15-80). on_host = 1 for host pointers. Replaces the model
 * half of cufd's set-up (Src/libCUFD.cu:67-88). Returns EFWI_ERR_ARG, and
 * leaves the context without a model, when the Courant number of the model
 * (efwi_courant_number of the largest sqrt((lambda + 2 mu) / rho)) exceeds 1,
 * where the reference exits (Src/libCUFD.cu:91, Src/utilities.cu:186-202).
 * Synchronises `stream` (the guard reads the maximum back). */
int efwi_set_model(efwi_ctx* ctx, const float* lambda, const float* mu, const float* rho,
                   int on_host, void* stream);

/* Zero the gradient accumulators (Src/Model.cu:62-64). */
int efwi_zero_gradient(efwi_ctx* ctx, void* stream);

/* Run one shot: the forward loop (Src/libCUFD.cu:312-354), and for calc_id
 * 0/1 the misfit and adjoint source (:371-441), and for calc_id 1 the
 * backward loop with reverse reconstruction and imaging (:469-613), which
 * accumulates into the gradients.
 *   obs     device [3][nrec][nt] observed data (calc_id 0/1), else NULL
 *   syn     device [3][nrec][nt] out: raw recorded data, or NULL
 *   res     device [3][nrec][nt] out: windowed residual (adjoint source), or NULL
 *   misfit  host, the shot's sum of squared residuals is ADDED to *misfit (FP64)
 */
int efwi_run_shot(efwi_ctx* ctx, const efwi_shot* shot, int calc_id, const float* obs,
                  float* syn, float* res, double* misfit, void* stream);

/* Copy the accumulated gradients out ([nz][nx][ny]); on_host = 1 for host
 * pointers. Replaces Src/libCUFD.cu:636-652. */
int efwi_get_gradient(efwi_ctx* ctx, float* g_lambda, float* g_mu, float* g_rho, int on_host,
                      void* stream);

/* Source illumination (pseudo-Hessian diagonal, Shin et al. 2001): with
 * enable != 0, every efwi_run_shot adds sum over its forward steps of
 * vz^2 + vx^2 + vy^2 per cell to a device array (padded grid, float) that
 * efwi_zero_gradient clears; the array is zeroed when enabled. Read it with
 * efwi_get_illumination (on_host: out is host memory). Inputs of the learned
 * preconditioner; no reference counterpart. */
int efwi_set_illumination(efwi_ctx* ctx, int enable, void* stream);
int efwi_get_illumination(efwi_ctx* ctx, float* out, int on_host, void* stream);

/* Per-kernel timing with HIP events recorded on the launch stream around
 * every stencil and ring launch of efwi_run_shot (no reference counterpart;
 * the reference only prints wall-clock totals, Src/libCUFD.cu:210-216).
 * efwi_profile_read fills EFWI_NKERNEL totals in milliseconds and launch
 * counts, indexed stress_fwd, velocity_fwd, stress_bwd (reverse stress +
 * adjoint stress), velocity_bwd (reverse velocity + adjoint velocity),
 * velocity_adj, stress_adj (first backward step only), ring, other,
 * fwd_fused, bwd_fused (the fused steps of efwi_set_fusion). */
#define EFWI_NKERNEL 10
int efwi_profile(efwi_ctx* ctx, int enable);
int efwi_profile_read(efwi_ctx* ctx, double* ms, long long* launches, int reset);

/* Launch shape of the four time-loop stencil kernels (kernel index as in
 * efwi_profile_read: 0 stress_fwd, 1 velocity_fwd, 2 stress_bwd,
 * 3 velocity_bwd). zdepth < 16: flat sweep, workgroups of 64 x rows threads
 * (rows 1, 2, 4), each thread updating zdepth consecutive z planes.
 * zdepth >= 16: z-marching sweep, workgroups of 64 x rows (rows 4 or 8; 1 and
 * 2 mean 4) marching zdepth planes with register z queues and LDS x/y tiles.
 * Results do not depend on the shape. efwi_autotune times 12 shapes of each
 * kernel on this grid (flat {1,1} {2,1} {4,1} {1,2} {1,4} {2,2}, marching
 * {4,16} {4,32} {4,64} {8,16} {8,32} {8,64}) with HIP events (reps launches
 * each) and keeps the fastest; report, if not NULL, receives the 4 x 12
 * average milliseconds in that order. No reference counterpart (the
 * reference launches fixed 32 x 1 x 32 blocks, Src/libCUFD.cu:93-97). */
int efwi_set_tile(efwi_ctx* ctx, int kernel, int rows, int zdepth);
int efwi_get_tile(const efwi_ctx* ctx, int kernel, int* rows, int* zdepth);
/* efwi_autotune times the candidates as the time loops run them: on a live
 * wavefield (the state filled with values of a wavefield's magnitude) and
 * with the boundary ring folded in where the loops fold it; it then decides
 * the split below for the chosen backward shapes. */
int efwi_autotune(efwi_ctx* ctx, int reps, double* report, void* stream);

/* Interior/rim split of the two backward kernels (flat shapes): on = 1 runs
 * the wavefronts clear of every C-PML slab (rows inside the x and z
 * interior, 64-cell y segments inside the y interior) as a launch of their
 * own, on a kernel without C-PML terms (fewer registers, twice the
 * wavefronts per SIMD), and the rest as the general kernel's launch. Results
 * are bit-identical either way. Default 0; efwi_autotune sets it. No
 * reference counterpart. */
int efwi_set_split(efwi_ctx* ctx, int on);
int efwi_get_split(const efwi_ctx* ctx, int* on);

/* memory_mode 1: how many of the time loop's last segments keep the boundary
 * ring the forward pass writes (segments >= 1; 0 = automatic, the default: as
 * many as 85 % of the free device memory holds). Only the earlier segments
 * are checkpointed and replayed, so each resident segment saves a replay of
 * checkpoint_every forward steps; the gradient is bit-identical whatever the
 * number. efwi_get_resident_segments reports it and the number of segments.
 * No reference counterpart (the reference keeps the whole ring in device
 * memory: Src/libCUFD.cu:86, Src/Boundary.cu:24). */
int efwi_set_resident_segments(efwi_ctx* ctx, int segments);
int efwi_get_resident_segments(const efwi_ctx* ctx, int* segments, int* total);

/* Fused time steps. pass 0: the forward step's stress and velocity updates in
 * one launch that marches z over workgroups of 64 x (rows + 4) threads, each
 * owning `rows` x rows (4, 8 or 12) of 60 y cells and `zchunk` z planes and
 * recomputing the new stress of a 2-cell ring around them, so that the new
 * velocity reads the new stress from registers and LDS instead of memory.
 * The state then alternates between two sets of wavefield and C-PML arrays
 * (allocated on first use, ~27 fields). rows 0 turns it off (the two-kernel
 * step). Results are bit-identical either way; sources outside the stencil
 * box use the two-kernel step. No reference counterpart (the reference
 * launches each kernel separately, Src/libCUFD.cu:312-354). */
int efwi_set_fusion(efwi_ctx* ctx, int pass, int rows, int zchunk);

/* Batched shots: run `nshots` shots of this grid and model together. Each
 * stencil step of all shots is one persistent launch over a (shot, tile) work
 * queue (per-XCD queues with stealing), the ring, receiver and injection work
 * of all shots one launch each, and with use_graph != 0 each time step's
 * launch sequence is captured once in a hipGraph and replayed (the time step
 * lives in device memory). obs[k] / syn[k] are device arrays [3][nrec_k][nt]
 * (syn may be NULL or hold NULLs); misfits[k] receives shot k's misfit.
 * calc_id 1 keeps a separate gradient per shot (efwi_batch_gradient) and
 * needs the full ring (memory_mode 0 or automatic) for every shot; sources
 * must lie inside the C-PML-free interior; if_win must be 0. Per shot the
 * results equal efwi_run_shot's bit for bit. The call returns when the batch
 * is done. No reference counterpart (the reference runs shots one after
 * another, Src/libCUFD.cu:237-620). */
int efwi_run_batch(efwi_ctx* ctx, const efwi_shot* shots, int nshots, int calc_id, const float* const* obs,
                   float* const* syn, double* misfits, int use_graph, void* stream);
/* Work distribution of the batched stencil launches. Each persistent
 * workgroup first runs a static share of its XCD's queue, static_pct percent
 * (0..100) of an even split, and then claims the rest of the queues, at most
 * `take` tiles per atomic (1..4096; lowered for small batches). Defaults:
 * take 2, static_pct 0. Results do not depend on either. */
int efwi_batch_config(efwi_ctx* ctx, int take, int static_pct);
/* How efwi_run_batch runs its shots. shots_per_launch > 0: groups of that
 * many shots, one group after another, each group one persistent launch per
 * stencil step over its (shot, tile) queue with its own captured graphs
 * (gradient batches keep a boundary ring per shot of a group only); -1: one
 * shot after another through the single-shot kernels (each shot on its own
 * state, as efwi_run_shot runs it); 0 (default): automatic, the single-shot
 * kernels where one shot already gives 16 workgroup tiles per compute unit
 * (the persistent queue's claims and the lost cache reuse between kernels
 * make it slower there), else groups of enough shots to reach that. Results
 * do not depend on it. */
int efwi_batch_group(efwi_ctx* ctx, int shots_per_launch);
/* Gradient (lambda, mu, rho) of shot `shot` of the last efwi_run_batch. */
int efwi_batch_gradient(efwi_ctx* ctx, int shot, float* g_lambda, float* g_mu, float* g_rho, int on_host,
                        void* stream);

/* FP64 dot product of two float device vectors on `stream` (per-block FP64
 * partial sums, then one FP64 reduction); *out is a host double. Used by the
 * L-BFGS two-loop recursion and gradient norms. No reference counterpart (the
 * reference's optimiser runs in scipy on the host). */
int efwi_dot64(const float* a, const float* b, long long n, double* out, void* stream);

/* Bytes of device memory held by the context (state + ring). */
long long efwi_device_bytes(const efwi_ctx* ctx);

/* Host utilities exposed for tests: the 1-D C-PML profile of one axis as six
 * rows K, a, b, K_half, a_half, b_half (Src/utilities.cu:204-320), and the
 * squared time taper w(t)^2 (Src/utilities.cu:578-614). */
int efwi_cpml_profile(int n, int npml, float h, float f0, float dt, float* out6n);
/* The per-trace factors w(t)^2 * weight * src_weight of the if_win misfit
 * (Src/utilities.cu:524-576, window ratio 0.005), [nrec][nt] into out;
 * weights may be NULL (all 1). The plain misfit's factors are
 * efwi_taper_w2(nt, dt, 0.005). */
int efwi_window_factors(int nt, float dt, int nrec, const float* win_start, const float* win_end,
                        const float* weights, float src_weight, float* out);
/* Courant number of the 3-D scheme, vp_max dt sqrt(3) (9/8 + 1/24) /
 * min(dz, dx, dy); efwi_set_model refuses models above 1. The reference's
 * compCourantNumber (Src/utilities.cu:186-202) uses sqrt(2) / min(dz, dx),
 * which is never larger. */
int efwi_courant_number(double vp_max, float dt, float dz, float dx, float dy, double* out);
int efwi_taper_w2(int nt, float dt, float ratio, float* out);

/* The reference's native entry, cufd (Src/libCUFD.h:6-9, Src/libCUFD.cu:37),
 * with a C string for the para file name and a status return instead of
 * exit(). Host pointers throughout, models in [nz][nx][ny]; reads the para
 * and survey JSON and, for calc_id 0/1, data_dir/Shot_{x,y,z}<id>.bin; for
 * calc_id 2 writes them. Gradients and misfit are written (not accumulated)
 * over the shots of this call. gpu_id selects the HIP device. */
int efwi_cufd(float* misfit, float* grad_lambda, float* grad_mu, float* grad_rho,
              const float* lambda, const float* mu, const float* rho, const float* stf,
              int calc_id, int gpu_id, int group_size, const int* shot_ids,
              const char* para_fname);

#ifdef __cplusplus
}
#endif

#endif /* EFWI_H_ */
