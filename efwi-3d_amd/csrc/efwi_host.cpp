// Host orchestration and C ABI of the gfx950 elastic FWI library.
//
// One efwi_ctx owns the device state of one grid: the forward and adjoint
// wavefields, the 18 C-PML memory arrays, the material arrays, the gradient
// accumulators, the 1-D C-PML profiles and (for gradients) the boundary-band
// ring of the reverse reconstruction. efwi_run_shot drives the time loops of
// the reference driver (Src/libCUFD.cu:220-613) on one HIP stream.
#include <hip/hip_runtime.h>

#include <algorithm>
#include <cmath>
#include <cstdio>
#include <cstring>
#include <string>
#include <vector>

#include "../../include/efwi.h"
#include "efwi_common.h"
#include "efwi_host.h"
#include "efwi_kernels.h"

namespace efwi {

static thread_local std::string g_err;

int fail(int code, const std::string& msg) {
  g_err = msg;
  return code;
}


// ---------------------------------------------------------------------------
// 1-D C-PML profile of one axis, Src/utilities.cu:204-320. The float
// variables of the reference (d0_h, depth, normalised depth, damp, K, alpha,
// a, b) are rounded to float where the reference stores them; the libm
// pow/log/exp are evaluated in double except where the reference calls
// expf/powf.
// ---------------------------------------------------------------------------
void cpml_profile_1d(int n, int npml, float h, float f0, float dt, float* K, float* a, float* b,
                     float* Kh, float* ah, float* bh) {
  const double PI = 3.141592653589793238462643383279502884197169;
  const float R = 0.0008f, KMAX = 2.0f, NP = 8.0f;
  const float AMAX = (float)(2.0 * PI * (f0 / 2.0));
  const float thick = npml * h;
  const float vref = 3000.0f;
  const float d0 = (float)(-(NP + 1) * vref * std::log((double)R) / (2.0 * thick));
  std::vector<float> damp(n, 0.0f), damph(n, 0.0f), alpha(n, 0.0f), alphah(n, 0.0f);
  for (int i = 0; i < n; i++) {
    K[i] = 1.0f; Kh[i] = 1.0f; a[i] = 0.0f; ah[i] = 0.0f;
  }
  auto prof = [&](float q, float& dmp, float& kk, float& al, bool kpowf) {
    dmp = (float)(d0 * (0.25f * (double)q + 0.75f * std::pow((double)q, (double)NP) +
                        0.0f * std::pow((double)q, 2.0 * NP)));
    kk = kpowf ? (float)(1.0 + (KMAX - 1.0) * powf(q, NP))
               : (float)(1.0 + (KMAX - 1.0) * std::pow((double)q, (double)NP));
    al = (float)(AMAX * (1.0 - q));
  };
  for (int i = 0; i < n; i++) {
    float dep = (float)(npml - i) * h;
    if (dep >= 0.0f) prof(dep / thick, damp[i], K[i], alpha[i], false);
    dep = (float)((npml - i - 0.5) * h);
    if (dep >= 0.0f) prof(dep / thick, damph[i], Kh[i], alphah[i], false);
    dep = (float)(npml - n + i) * h;
    if (dep >= 0.0f) prof(dep / thick, damp[i], K[i], alpha[i], false);
    dep = (float)((npml - n + i + 0.5) * h);
    if (dep >= 0.0f) prof(dep / thick, damph[i], Kh[i], alphah[i], true);
    if (alpha[i] < 0.0f) alpha[i] = 0.0f;
    if (alphah[i] < 0.0f) alphah[i] = 0.0f;
    b[i] = expf(-(damp[i] / K[i] + alpha[i]) * dt);
    bh[i] = expf(-(damph[i] / Kh[i] + alphah[i]) * dt);
    if (std::fabs(damp[i]) > 1.0e-6)
      a[i] = (float)(damp[i] * (b[i] - 1.0) / (K[i] * (damp[i] + K[i] * alpha[i])));
    if (std::fabs(damph[i]) > 1.0e-6)
      ah[i] = (float)(damph[i] * (bh[i] - 1.0) / (Kh[i] * (damph[i] + Kh[i] * alphah[i])));
  }
}

// w(t) of the time tapers, Src/utilities.cu:578-614 (plain) and :524-576
// (per trace). `valid` is false where the reference returns early and leaves
// the data unchanged.
static float taper_w(float t, float t0, float t1, float t2, float t3) {
  const double PI = 3.141592653589793238462643383279502884197169;
  if (t >= t0 && t < t1) return (float)std::sin(PI / 2.0 * (t - t0) / (t1 - t0));
  if (t >= t1 && t < t2) return 1.0f;
  if (t >= t2 && t < t3) return (float)std::cos(PI / 2.0 * (t - t2) / (t3 - t2));
  return 0.0f;
}

// Courant number of the 3-D 4th-order staggered scheme, vmax dt sqrt(3)
// (9/8 + 1/24) / min(dz, dx, dy): the scheme is unstable above 1. The
// reference's guard (compCourantNumber, Src/utilities.cu:186-202, called at
// Src/libCUFD.cu:91) uses the 2-D form sqrt(2) / min(dz, dx), which never
// exceeds this one, so every model it refuses is refused here too.
double courant_number(double vmax, double dt, double dz, double dx, double dy) {
  const double hmin = std::min(dz, std::min(dx, dy));
  return vmax * dt * std::sqrt(3.0) * (9.0 / 8.0 + 1.0 / 24.0) / hmin;
}
double courant_number_2d(double vmax, double dt, double dz, double dx) {
  return vmax * dt * std::sqrt(2.0) * (1.0 / 24.0 + 9.0 / 8.0) / std::min(dz, dx);
}

float taper_plain_w2(int it, int nt, float dt, float ratio) {
  const float t = it * dt, t0 = 0, t3 = nt * dt, off = nt * dt * ratio;
  if (2.0 * off >= t3 - t0) return 1.0f;
  const float w = taper_w(t, t0, t0 + off, t3 - off, t3);
  return w * w;
}

float taper_window_fac(int it, int nt, float dt, float ratio, float t0, float t3, float weight,
                       float src_weight) {
  const float t = it * dt;
  if (ratio > 0.5) return 1.0f;
  const float tmax = nt * dt;
  if (t0 < 0.0f) t0 = 0.0f;
  if (t0 > tmax) t0 = tmax;
  if (t3 < 0.0f) t3 = 0.0f;
  if (t3 > tmax) t3 = tmax;
  const float off = (t3 - t0) * ratio;
  if (off <= 0.0) return 1.0f;
  const float w = taper_w(t, t0, t0 + off, t3 - off, t3);
  return w * w * weight * src_weight;
}

// ---------------------------------------------------------------------------
// Profiler: an event pair around each timed launch.
// ---------------------------------------------------------------------------
void Profiler::begin(int k, hipStream_t s) {
  if (!on) return;
  if (2 * (used + 1) > pool.size()) {
    for (int j = 0; j < 2; j++) {
      hipEvent_t e;
      hipEventCreate(&e);
      pool.push_back(e);
    }
    kind.push_back(0);
  }
  kind[used] = k;
  hipEventRecord(pool[2 * used], s);
}

void Profiler::end(hipStream_t s) {
  if (!on) return;
  hipEventRecord(pool[2 * used + 1], s);
  used++;
}

int Profiler::collect() {
  if (!used) return EFWI_OK;
  if (hipEventSynchronize(pool[2 * used - 1]) != hipSuccess) return fail(EFWI_ERR_HIP, "event synchronize");
  for (size_t u = 0; u < used; u++) {
    float t = 0.0f;
    if (hipEventElapsedTime(&t, pool[2 * u], pool[2 * u + 1]) != hipSuccess)
      return fail(EFWI_ERR_HIP, "event elapsed time");
    ms[kind[u]] += t;
    launches[kind[u]]++;
  }
  used = 0;
  return EFWI_OK;
}

void Profiler::release() {
  for (hipEvent_t e : pool) hipEventDestroy(e);
  pool.clear();
  kind.clear();
  used = 0;
}

// ---------------------------------------------------------------------------
// Context.
// ---------------------------------------------------------------------------
Ctx::~Ctx() {
  int prev = 0;
  hipGetDevice(&prev);
  hipSetDevice(cfg.device);
  prof.release();
  release_batch();
  for (void* p : allocs) hipFree(p);
  if (ring) hipFree(ring);
  if (ckpt) hipFree(ckpt);
  hipSetDevice(prev);
}

int Ctx::alloc(void** p, size_t bytes) {
  hipError_t e = hipMalloc(p, bytes);
  if (e != hipSuccess) {
    *p = nullptr;
    return fail(EFWI_ERR_NOMEM, "hipMalloc(" + std::to_string(bytes) + " B) failed: " +
                                    hipGetErrorString(e));
  }
  allocs.push_back(*p);
  bytes_held += (long long)bytes;
  return EFWI_OK;
}

int grow(Ctx* c, void** p, size_t* cap, size_t need) {
  if (*cap >= need) return EFWI_OK;
  if (*p) {
    for (auto& q : c->allocs)
      if (q == *p) q = nullptr;
    hipFree(*p);
    c->bytes_held -= (long long)*cap;
  }
  int rc = c->alloc(p, need);
  *cap = rc == EFWI_OK ? need : 0;
  return rc;
}

// Field pointers of a wave group / psi group (fields fsf floats apart).
void wave_of(float* base, size_t fsf, Wave& W) {
  float** f[9] = {&W.vz, &W.vx, &W.vy, &W.szz, &W.sxx, &W.sxz, &W.syy, &W.sxy, &W.syz};
  for (int k = 0; k < F_WAVE_COUNT; k++) *f[k] = base + k * fsf;
}
void psi_of(float* base, size_t fsf, Psi& P) {
  float** f = reinterpret_cast<float**>(&P);
  for (int k = 0; k < P_PSI_COUNT; k++) f[k] = base + k * fsf;
}

int Ctx::init() {
  const efwi_config& c = cfg;
  if (c.nz < 8 || c.nx < 8 || c.ny < 8 || c.npml < 2 || c.nt < 2)
    return fail(EFWI_ERR_ARG, "grid too small (need nz, nx, ny >= 8, npml >= 2, nt >= 2)");
  if (c.nz < 2 * c.npml + 6 || c.nx < 2 * c.npml + 6 || c.ny < 2 * c.npml + 6)
    return fail(EFWI_ERR_ARG, "each padded axis must exceed 2*npml + 6 cells");
  if (!(c.dz > 0 && c.dx > 0 && c.dy > 0 && c.dt > 0))
    return fail(EFWI_ERR_ARG, "grid steps and dt must be positive");
  if (c.formulation != 0 && c.formulation != 1) return fail(EFWI_ERR_ARG, "formulation must be 0 or 1");
  int ndev = 0;
  HIP_TRY(hipGetDeviceCount(&ndev));
  if (c.device < 0 || c.device >= ndev) return fail(EFWI_ERR_ARG, "device ordinal out of range");
  HIP_TRY(hipSetDevice(c.device));
  G.nz = c.nz; G.nx = c.nx; G.ny = c.ny; G.npml = c.npml;
  // rows padded to whole 128-byte lines: a wavefront's 64-cell y segment then
  // covers exactly two lines (the padding cells are never computed or read)
  G.py = (c.ny + 31) / 32 * 32;
  G.pxy = (long)c.nx * G.py;
  G.dz = c.dz; G.dx = c.dx; G.dy = c.dy; G.dt = c.dt;
  N = (long)c.nz * G.pxy;
  // field groups: fields fs bytes apart (256-byte aligned), each group below 4 GiB
  const size_t fs = ((size_t)N * sizeof(float) + 255) / 256 * 256;
  if (fs * P_PSI_COUNT > 0xFFFFFFFFull)
    return fail(EFWI_ERR_ARG, "grid of " + std::to_string(N) + " cells exceeds the 32-bit field groups (at most " +
                                  std::to_string(0xFFFFFFFFull / P_PSI_COUNT / 4) + " cells)");
  G.fs = (unsigned)fs;
  const size_t fsf = fs / sizeof(float);
  int rc;
  if ((rc = alloc((void**)&F.w, fs * F_WAVE_COUNT)) || (rc = alloc((void**)&F.a, fs * F_WAVE_COUNT)) ||
      (rc = alloc((void**)&F.p, fs * P_PSI_COUNT)) || (rc = alloc((void**)&F.m, fs * M_COUNT)))
    return rc;
  wave_of(F.w, fsf, W);
  wave_of(F.a, fsf, A);
  psi_of(F.p, fsf, P);
  lam = F.m + M_LAM * fsf; mu = F.m + M_MU * fsf; byc = F.m + M_BYC * fsf;
  Gr.glam = F.m + M_GLAM * fsf; Gr.gmu = F.m + M_GMU * fsf; Gr.grho = F.m + M_GRHO * fsf;
  M.lam = lam; M.mu = mu; M.byc = byc;
  // C-PML profiles: 3 axes x kProfRows rows (K, a, b, Kh, ah, bh)
  const int n3[3] = {c.nz, c.nx, c.ny};
  const float h3[3] = {c.dz, c.dx, c.dy};
  std::vector<float> host;
  for (int ax = 0; ax < 3; ax++) {
    const int n = n3[ax];
    C.off[ax] = (int)host.size();
    C.n[ax] = n;
    host.resize(host.size() + (size_t)kProfRows * n);
    float* r = host.data() + C.off[ax];
    cpml_profile_1d(n, c.npml, h3[ax], c.f0, c.dt, r + R_K * n, r + R_A * n, r + R_B * n, r + R_KH * n,
                    r + R_AH * n, r + R_BH * n);
    for (int i = 0; i < n; i++) {
      const float K = r[R_K * n + i], Kh = r[R_KH * n + i];
      // the backward kernels drop the C-PML terms of rows interior along x and z
      // (and of wavefronts clear of the y slabs): there a = 0 and K = 1 exactly
      const bool interior = i >= c.npml && i <= n - c.npml - 1;
      if (interior && (r[R_A * n + i] != 0.0f || r[R_AH * n + i] != 0.0f || K != 1.0f || Kh != 1.0f))
        return fail(EFWI_ERR_ARG, "C-PML profile is not trivial inside the PML-free interior");
    }
  }
  float* prof = nullptr;
  if ((rc = alloc((void**)&prof, sizeof(float) * host.size()))) return rc;
  C.p = prof;
  HIP_TRY(hipMemcpy(prof, host.data(), sizeof(float) * host.size(), hipMemcpyHostToDevice));
  // boundary band of the reverse reconstruction (Src/Boundary.cu:14-20)
  for (int ax = 0; ax < 3; ax++) {
    B.lo[ax] = c.npml - 2;
    B.hi[ax] = n3[ax] - c.npml + 1;
  }
  const long nzB = B.hi[0] - B.lo[0] + 1, nxB = B.hi[1] - B.lo[1] + 1, nyB = B.hi[2] - B.lo[2] + 1;
  B.nzs = 2 * 5 * nxB * nyB;
  B.nxs = 2 * 5 * (nzB - 10) * nyB;
  B.nys = 2 * 5 * (nzB - 10) * (nxB - 10);
  B.n = B.nzs + B.nxs + B.nys;
  if ((rc = plan_memory())) return rc;
  if ((rc = alloc((void**)&msum, sizeof(double) * 8))) return rc;
  if ((rc = alloc((void**)&partials, sizeof(double) * kMaxPartials))) return rc;
  if ((rc = alloc((void**)&stf_dev, sizeof(float) * c.nt))) return rc;
  if ((rc = alloc((void**)&w2_dev, sizeof(float) * c.nt))) return rc;
  // plain taper of the data (ratio 0.005, Src/libCUFD.cu:64)
  std::vector<float> w2(c.nt);
  for (int it = 0; it < c.nt; it++) w2[it] = taper_plain_w2(it, c.nt, c.dt, 0.005f);
  HIP_TRY(hipMemcpy(w2_dev, w2.data(), sizeof(float) * c.nt, hipMemcpyHostToDevice));
  HIP_TRY(hipDeviceSynchronize());
  return EFWI_OK;
}

void fields9(const Wave& W, Fields9& F) {
  F.f[0] = W.vz; F.f[1] = W.vx; F.f[2] = W.vy; F.f[3] = W.szz; F.f[4] = W.sxx;
  F.f[5] = W.sxz; F.f[6] = W.syy; F.f[7] = W.sxy; F.f[8] = W.syz;
}

void zero_wave(const Wave& W, long n, hipStream_t s) {
  Fields9 F;
  fields9(W, F);
  for (int k = 0; k < 9; k++) launch_fill(F.f[k], n, 0.0f, s);
}

void zero_psi(const Psi& P, long n, hipStream_t s) {
  float* const* p = reinterpret_cast<float* const*>(&P);
  for (int k = 0; k < 18; k++) launch_fill(p[k], n, 0.0f, s);
}

// Memory plan of the reverse reconstruction.
//   mode 0: the band of every step stays in HBM ([nt-1][9][band]).
//   mode 1: the forward state (v, s and the PML slabs of the 18 psi arrays) is
//           kept every K steps; the backward pass replays each segment of K
//           steps from its checkpoint into separate buffers to rebuild that
//           segment's band ring, then reverses through it. The replayed band
//           values are the values the forward pass produced (the kernels are
//           deterministic), so the gradient is bit-identical to mode 0.
int Ctx::plan_memory() {
  const int steps = cfg.nt - 1;
  if (cfg.memory_mode < 0) {
    // automatic: the full ring when it fits in 90 % of the free device memory
    size_t free_b = 0, total_b = 0;
    HIP_TRY(hipMemGetInfo(&free_b, &total_b));
    const double ring_b = (double)B.n * 9 * steps * sizeof(float);
    cfg.memory_mode = ring_b < 0.9 * (double)free_b ? 0 : 1;
  }
  if (cfg.memory_mode == 0) {
    seg_len = steps;
  } else if (cfg.memory_mode == 1) {
    seg_len = cfg.checkpoint_every > 0 ? cfg.checkpoint_every : (int)std::ceil(std::sqrt((double)steps));
    if (seg_len > steps) seg_len = steps;
  } else {
    return fail(EFWI_ERR_ARG, "memory_mode must be 0 (ring) or 1 (checkpointed ring)");
  }
  nseg = (steps + seg_len - 1) / seg_len;
  return cfg.memory_mode == 1 ? choose_resident() : (res_seg = 1, EFWI_OK);
}

// memory_mode 1: how many of the last segments keep the ring the forward pass
// writes instead of being replayed. Each resident segment costs seg_len ring
// slots and saves one checkpoint and one replay of seg_len forward steps;
// automatic: as many as 85 % of the free device memory holds (at least one).
int Ctx::choose_resident() {
  if (res_req > 0) {
    res_seg = std::min(res_req, nseg);
    return EFWI_OK;
  }
  size_t free_b = 0, total_b = 0;
  HIP_TRY(hipMemGetInfo(&free_b, &total_b));
  // memory already held for the ring and checkpoints counts as available
  const double avail = 0.85 * (double)(free_b + ring_bytes + ckpt_bytes) -
                       (F2.w ? 0.0 : (double)G.fs * (F_WAVE_COUNT + P_PSI_COUNT));
  const double seg_b = (double)B.n * 9 * seg_len * sizeof(float);
  const double ck_b = (double)ckpt_floats_needed() * sizeof(float);
  // R resident segments: R seg_b + (nseg - R) ck_b <= avail
  const double r = seg_b > ck_b ? (avail - nseg * ck_b) / (seg_b - ck_b) : nseg;
  res_seg = (int)std::max(1.0, std::min((double)nseg, std::floor(r)));
  return EFWI_OK;
}

int Ctx::ensure_ring() {
  const size_t need = (size_t)B.n * 9 * (size_t)seg_len * res_seg * sizeof(float);
  if (ring && ring_bytes >= need) return cfg.memory_mode == 1 ? ensure_checkpoints() : EFWI_OK;
  if (ring) {
    hipFree(ring);
    bytes_held -= (long long)ring_bytes;
    ring = nullptr;
  }
  hipError_t e = hipMalloc((void**)&ring, need);
  if (e != hipSuccess) {
    ring = nullptr;
    ring_bytes = 0;
    return fail(EFWI_ERR_NOMEM, "boundary ring of " + std::to_string(need) +
                                    " B does not fit in device memory: " + hipGetErrorString(e) +
                                    " (memory_mode 1 keeps only sqrt(nt) steps of it; "
                                    "efwi_set_resident_segments lowers the resident part)");
  }
  ring_bytes = need;
  bytes_held += (long long)need;
  if (cfg.memory_mode == 1) return ensure_checkpoints();
  return EFWI_OK;
}

// Bytes of one checkpoint: v and s in full, and of each psi array only the
// slabs where its C-PML predicate can be true (z: planes [0,p) and
// [nz-p,nz); x: rows [0,p) and [nx-p,nx) of every plane; y: the first and
// last p cells of every row). Outside those slabs psi is never written.
static const int kPsiAxis[18] = {0, 1, 2, 0, 1, 2, 0, 1, 2,   // vz_z vz_x vz_y vx_z vx_x vx_y vy_z vy_x vy_y
                                 0, 1, 0, 1, 1, 2, 2, 2, 0};  // szz_z sxz_x sxz_z sxx_x sxy_x sxy_y syy_y syz_y syz_z

size_t Ctx::slab_floats(int axis) const {
  const long p = cfg.npml;
  if (axis == 0) return (size_t)(2 * p) * G.pxy;
  if (axis == 1) return (size_t)cfg.nz * 2 * p * G.py;
  return (size_t)cfg.nz * cfg.nx * 2 * p;
}

// Floats of one checkpoint: v and s in full, and of each psi array its
// C-PML slabs (ensure_checkpoints).
size_t Ctx::ckpt_floats_needed() const {
  size_t per = 9 * (size_t)N;
  for (int k = 0; k < 18; k++) per += slab_floats(kPsiAxis[k]);
  return per;
}

int Ctx::ensure_checkpoints() {
  const size_t per = ckpt_floats_needed();
  ckpt_floats = per;
  const size_t need = per * (size_t)std::max(first_res(), 1) * sizeof(float);
  int rc;
  if (!F2.w) {
    const size_t fs = G.fs, fsf = fs / sizeof(float);
    if ((rc = alloc((void**)&F2.w, fs * F_WAVE_COUNT)) || (rc = alloc((void**)&F2.p, fs * P_PSI_COUNT))) return rc;
    F2.a = F.a;
    F2.m = F.m;
    wave_of(F2.w, fsf, W2);
    psi_of(F2.p, fsf, P2);
    // psi outside its slabs is never written: zero it once
    zero_psi(P2, N, nullptr);
    HIP_TRY(hipDeviceSynchronize());
  }
  if (ckpt && ckpt_bytes >= need) return EFWI_OK;
  if (ckpt) {
    hipFree(ckpt);
    bytes_held -= (long long)ckpt_bytes;
    ckpt = nullptr;
  }
  hipError_t e = hipMalloc((void**)&ckpt, need);
  if (e != hipSuccess) {
    ckpt = nullptr;
    ckpt_bytes = 0;
    return fail(EFWI_ERR_NOMEM, "checkpoints of " + std::to_string(need) + " B do not fit: " +
                                    hipGetErrorString(e) + " (raise checkpoint_every)");
  }
  ckpt_bytes = need;
  bytes_held += (long long)need;
  return EFWI_OK;
}

// Copy the forward state (W, P) to checkpoint k (save = true) or checkpoint k
// to (W, P) (save = false).
int Ctx::checkpoint(int k, const Wave& Wf, const Psi& Pf, bool save, hipStream_t s) {
  float* base = ckpt + (size_t)k * ckpt_floats;
  const size_t fb = (size_t)N * sizeof(float);
  Fields9 F;
  fields9(Wf, F);
  for (int f = 0; f < 9; f++) {
    if (save) HIP_TRY(hipMemcpyAsync(base, F.f[f], fb, hipMemcpyDeviceToDevice, s));
    else HIP_TRY(hipMemcpyAsync(F.f[f], base, fb, hipMemcpyDeviceToDevice, s));
    base += N;
  }
  float* const* pf = reinterpret_cast<float* const*>(&Pf);
  const long p = cfg.npml;
  for (int k2 = 0; k2 < 18; k2++) {
    const int ax = kPsiAxis[k2];
    float* a = pf[k2];
    // two slabs per array: (offset, width in floats, rows, pitch in floats)
    long off[2], width, rows, pitch;
    if (ax == 0) {
      off[0] = 0; off[1] = (long)(cfg.nz - p) * G.pxy; width = p * G.pxy; rows = 1; pitch = width;
    } else if (ax == 1) {
      off[0] = 0; off[1] = (long)(cfg.nx - p) * G.py; width = p * G.py; rows = cfg.nz; pitch = G.pxy;
    } else {
      off[0] = 0; off[1] = cfg.ny - p; width = p; rows = (long)cfg.nz * cfg.nx; pitch = G.py;
    }
    for (int h = 0; h < 2; h++) {
      float* dev = a + off[h];
      const size_t wb = width * sizeof(float), pb = pitch * sizeof(float);
      if (save) HIP_TRY(hipMemcpy2DAsync(base, wb, dev, pb, wb, rows, hipMemcpyDeviceToDevice, s));
      else HIP_TRY(hipMemcpy2DAsync(dev, pb, base, wb, wb, rows, hipMemcpyDeviceToDevice, s));
      base += width * rows;
    }
  }
  return EFWI_OK;
}

int Ctx::copy_in(float* field, const float* user, hipMemcpyKind k, hipStream_t s) {
  const size_t row = (size_t)cfg.ny * sizeof(float);
  HIP_TRY(hipMemcpy2DAsync(field, (size_t)G.py * sizeof(float), user, row, row, (size_t)cfg.nz * cfg.nx, k, s));
  return EFWI_OK;
}

int Ctx::copy_out(float* user, const float* field, hipMemcpyKind k, hipStream_t s) {
  const size_t row = (size_t)cfg.ny * sizeof(float);
  HIP_TRY(hipMemcpy2DAsync(user, row, field, (size_t)G.py * sizeof(float), row, (size_t)cfg.nz * cfg.nx, k, s));
  return EFWI_OK;
}

int Ctx::set_model(const float* l, const float* m, const float* rho, int on_host, hipStream_t s) {
  HIP_TRY(hipSetDevice(cfg.device));
  const hipMemcpyKind k = on_host ? hipMemcpyHostToDevice : hipMemcpyDeviceToDevice;
  int rc;
  if ((rc = copy_in(lam, l, k, s)) || (rc = copy_in(mu, m, k, s))) return rc;
  // rho is staged in the row layout through the gradient buffer, then the
  // buffer is cleared again
  if ((rc = copy_in(Gr.grho, rho, k, s))) return rc;
  model_set = false;
  // stability guard (Src/libCUFD.cu:91): refuse a model whose Courant number exceeds 1
  launch_vp2max(G, lam, mu, Gr.grho, partials, kMaxPartials, msum, s);
  HIP_TRY(hipGetLastError());
  double v2 = 0.0;
  HIP_TRY(hipMemcpyAsync(&v2, msum, sizeof(double), hipMemcpyDeviceToHost, s));
  HIP_TRY(hipStreamSynchronize(s));
  const double vmax = std::sqrt(v2);
  const double cn = courant_number(vmax, cfg.dt, cfg.dz, cfg.dx, cfg.dy);
  if (!(cn <= 1.0)) {
    launch_fill(Gr.grho, N, 0.0f, s);
    char msg[256];
    snprintf(msg, sizeof msg,
             "unstable model: Courant number %.4f > 1 (vp max %.2f, dt %g, h min %g; the reference's 2-D "
             "form gives %.4f)",
             cn, vmax, (double)cfg.dt, (double)std::min(cfg.dz, std::min(cfg.dx, cfg.dy)),
             courant_number_2d(vmax, cfg.dt, cfg.dz, cfg.dx));
    return fail(EFWI_ERR_ARG, msg);
  }
  launch_byc(G, Gr.grho, byc, s);
  HIP_TRY(hipGetLastError());
  launch_fill(Gr.grho, N, 0.0f, s);
  model_set = true;
  return EFWI_OK;
}

int Ctx::zero_gradient(hipStream_t s) {
  HIP_TRY(hipSetDevice(cfg.device));
  launch_fill(Gr.glam, N, 0.0f, s);
  launch_fill(Gr.gmu, N, 0.0f, s);
  launch_fill(Gr.grho, N, 0.0f, s);
  if (illum_on) launch_fill(illum, N, 0.0f, s);
  HIP_TRY(hipGetLastError());
  return EFWI_OK;
}

int Ctx::prepare_shot(const efwi_shot* sh, int calc_id) {
  const int nt = cfg.nt;
  if (!sh || !sh->stf) return fail(EFWI_ERR_ARG, "shot or its stf is NULL");
  if (sh->nrec < 0) return fail(EFWI_ERR_ARG, "negative nrec");
  if (sh->nrec > 0 && !sh->rec) return fail(EFWI_ERR_ARG, "receiver array is NULL");
  auto inside = [&](int z, int x, int y) {
    return z >= 0 && z < cfg.nz && x >= 0 && x < cfg.nx && y >= 0 && y < cfg.ny;
  };
  if (!inside(sh->src_z, sh->src_x, sh->src_y)) return fail(EFWI_ERR_ARG, "source outside the padded grid");
  std::vector<long> cells(sh->nrec);
  for (int r = 0; r < sh->nrec; r++) {
    const int z = sh->rec[3 * r], x = sh->rec[3 * r + 1], y = sh->rec[3 * r + 2];
    if (!inside(z, x, y))
      return fail(EFWI_ERR_ARG, "receiver " + std::to_string(r) + " outside the padded grid");
    cells[r] = (long)z * G.pxy + (long)x * G.py + y;
  }
  nrec = sh->nrec;
  src_cell = (long)sh->src_z * G.pxy + (long)sh->src_x * G.py + sh->src_y;
  const int p = cfg.npml;
  src_interior = sh->src_z >= p && sh->src_z <= cfg.nz - 1 - p && sh->src_x >= p &&
                 sh->src_x <= cfg.nx - 1 - p && sh->src_y >= p && sh->src_y <= cfg.ny - 1 - p;
  src_stencil = sh->src_z >= 2 && sh->src_z <= cfg.nz - 3 && sh->src_x >= 2 && sh->src_x <= cfg.nx - 3 &&
                sh->src_y >= 2 && sh->src_y <= cfg.ny - 3;
  // stf tapered with ratio 0.001 (Src/Src_Rec.cu:146)
  std::vector<float> stf(nt);
  for (int it = 0; it < nt; it++) stf[it] = sh->stf[it] * taper_plain_w2(it, nt, cfg.dt, 0.001f);
  HIP_TRY(hipMemcpy(stf_dev, stf.data(), sizeof(float) * nt, hipMemcpyHostToDevice));
  const size_t nrb = (size_t)std::max(nrec, 1);
  int rc;
  if ((rc = grow(this, (void**)&rec_dev, &rec_cap, sizeof(long) * nrb))) return rc;
  if (nrec) HIP_TRY(hipMemcpy(rec_dev, cells.data(), sizeof(long) * nrec, hipMemcpyHostToDevice));
  const size_t seis = sizeof(float) * (size_t)nt * 3 * nrb;
  if ((rc = grow(this, (void**)&syn_t, &syn_cap, seis))) return rc;
  if (calc_id != EFWI_CALC_SYNTHETIC) {
    if ((rc = grow(this, (void**)&obs_t, &obs_cap, seis))) return rc;
    if ((rc = grow(this, (void**)&res_t, &res_cap, seis))) return rc;
    per_trace = cfg.if_win != 0;
    if (per_trace) {
      if (!sh->win_start || !sh->win_end) return fail(EFWI_ERR_ARG, "if_win needs win_start and win_end");
      std::vector<float> fac((size_t)nt * nrb);
      for (int it = 0; it < nt; it++)
        for (int r = 0; r < nrec; r++)
          fac[(size_t)it * nrec + r] =
              taper_window_fac(it, nt, cfg.dt, 0.005f, sh->win_start[r], sh->win_end[r],
                               sh->weights ? sh->weights[r] : 1.0f, sh->src_weight);
      if ((rc = grow(this, (void**)&fac_dev, &fac_cap, sizeof(float) * fac.size()))) return rc;
      HIP_TRY(hipMemcpy(fac_dev, fac.data(), sizeof(float) * fac.size(), hipMemcpyHostToDevice));
    }
  }
  return EFWI_OK;
}

// One forward time step on (Wf, Pf): stress + source, velocity, and (record)
// the receiver gather of step it + 1.
void Ctx::forward_step(const Groups& Ff, const Wave& Wf, const Source& Sf, int it, bool record, hipStream_t s,
                       long save) {
  // save >= 0: ring slot of the state this step produces, written by the
  // stencil kernels themselves where they can (a source added after the stress
  // kernel leaves the stress fields to the separate pass)
  RingFold rf{};
  rf.B = B;
  rf.slot = save >= 0 ? ring + save * 9 * B.n : nullptr;
  prof.begin(K_STRESS_FWD, s);
  const bool fs = launch_stress_fwd(tiles[K_STRESS_FWD], G, Ff, C, Sf, it, s, src_stencil ? &rf : nullptr);
  prof.end(s);
  if (!src_stencil) launch_add_source(Wf.szz, Wf.sxx, Wf.syy, src_cell, stf_dev, it, G.dt, +1, s);
  prof.begin(K_VELOCITY_FWD, s);
  const bool fv = launch_velocity_fwd(tiles[K_VELOCITY_FWD], cfg.formulation, G, Ff, C, s, &rf);
  prof.end(s);
  if (save >= 0 && !(fs && fv)) {
    Fields9 F9;
    fields9(Wf, F9);
    prof.begin(K_RING, s);
    if (!fv) launch_ring(B, G, F9, ring, save, F_VZ, F_VY + 1, 0, s);
    if (!fs) launch_ring(B, G, F9, ring, save, F_SZZ, F_WAVE_COUNT, 0, s);
    prof.end(s);
  }
  if (record) launch_record(Wf, rec_dev, nrec, it + 1, syn_t, s);
}

// The second set of wavefield and C-PML arrays of the fused steps, zeroed
// once (cells a step never writes stay zero in both sets).
int Ctx::ensure_alt() {
  if (Falt.w) return EFWI_OK;
  const size_t fs = G.fs;
  int rc;
  if ((rc = alloc((void**)&Falt.w, fs * F_WAVE_COUNT)) || (rc = alloc((void**)&Falt.p, fs * P_PSI_COUNT))) return rc;
  HIP_TRY(hipMemset(Falt.w, 0, fs * F_WAVE_COUNT));
  HIP_TRY(hipMemset(Falt.p, 0, fs * P_PSI_COUNT));
  HIP_TRY(hipDeviceSynchronize());
  return EFWI_OK;
}

// After a fused step: the new state is in Falt; make it the current one.
void Ctx::swap_state() {
  std::swap(F.w, Falt.w);
  std::swap(F.p, Falt.p);
  const size_t fsf = G.fs / sizeof(float);
  wave_of(F.w, fsf, W);
  psi_of(F.p, fsf, P);
}

void Ctx::fused_forward_step(const Source& Sf, int it, bool record, hipStream_t s, long save) {
  RingFold rf{};
  rf.B = B;
  rf.slot = save >= 0 ? ring + save * 9 * B.n : nullptr;
  Groups Fo = F;
  Fo.w = Falt.w;
  Fo.p = Falt.p;
  const Tile t = tiles[K_FWD_FUSED];
  prof.begin(K_FWD_FUSED, s);
  const bool folded = launch_fwd_fused(t.rows, t.zdepth, cfg.formulation, G, F, Fo, C, Sf, it, s, &rf);
  prof.end(s);
  swap_state();
  if (save >= 0 && !folded) {
    Fields9 F9;
    fields9(W, F9);
    prof.begin(K_RING, s);
    launch_ring(B, G, F9, ring, save, 0, 9, 0, s);
    prof.end(s);
  }
  if (record) launch_record(W, rec_dev, nrec, it + 1, syn_t, s);
}

int Ctx::run_shot(const efwi_shot* sh, int calc_id, const float* obs, float* syn, float* res,
                  double* misfit, hipStream_t s) {
  if (calc_id < 0 || calc_id > 2) return fail(EFWI_ERR_ARG, "invalid calc_id " + std::to_string(calc_id));
  if (!model_set) return fail(EFWI_ERR_STATE, "efwi_set_model was not called");
  if (calc_id != EFWI_CALC_SYNTHETIC && !obs) return fail(EFWI_ERR_ARG, "observed data required");
  HIP_TRY(hipSetDevice(cfg.device));
  // Pageable copies below synchronise with the host; finish earlier work on s first.
  HIP_TRY(hipStreamSynchronize(s));
  int rc = prepare_shot(sh, calc_id);
  if (rc) return rc;
  const bool grad = calc_id == EFWI_CALC_GRADIENT;
  if (grad && (rc = ensure_ring())) return rc;
  const int nt = cfg.nt, form = cfg.formulation;
  const bool fused_fwd = tiles[K_FWD_FUSED].rows > 0 && src_stencil;
  if (fused_fwd && (rc = ensure_alt())) return rc;
  Fields9 F9;

  zero_wave(W, N, s);
  zero_wave(A, N, s);
  zero_psi(P, N, s);
  launch_fill(syn_t, (long)nt * 3 * std::max(nrec, 1), 0.0f, s);

  // forward loop, Src/libCUFD.cu:312-354
  Source Sf;
  Sf.cell = src_stencil ? src_cell : -1;
  Sf.stf = stf_dev;
  const int last_seg = nseg - 1, res0 = first_res();
  for (int it = 0; it <= nt - 2; it++) {
    const int seg = it / seg_len;
    if (grad && cfg.memory_mode == 1 && it % seg_len == 0 && seg < res0)
      if ((rc = checkpoint(seg, W, P, true, s))) return rc;
    // The ring slots of the resident segments are filled by this pass
    // directly: slot it - a holds the state before step it (a: the first
    // resident step), the first by a pass of its own, the others by the step
    // before.
    const int a = res0 * seg_len;
    if (grad && it == a) {
      fields9(W, F9);
      prof.begin(K_RING, s);
      launch_ring(B, G, F9, ring, 0, 0, 9, 0, s);
      prof.end(s);
    }
    const long save = grad && it >= a && it + 1 <= nt - 2 ? it + 1 - a : -1;
    if (fused_fwd) fused_forward_step(Sf, it, true, s, save);
    else forward_step(F, W, Sf, it, true, s, save);
    if (illum_on) launch_illum(W.vz, W.vx, W.vy, illum, N, s);
  }
  fields9(W, F9);
  HIP_TRY(hipGetLastError());
  if (syn && nrec > 0) launch_from_time_major(syn_t, syn, nrec, nt, s);

  if (calc_id != EFWI_CALC_SYNTHETIC && nrec > 0) {
    // misfit and adjoint source, Src/libCUFD.cu:371-441
    launch_to_time_major(obs, obs_t, nrec, nt, s);
    const int nb = launch_misfit(obs_t, syn_t, res_t, per_trace ? fac_dev : w2_dev, per_trace ? 1 : 0,
                                 nt, nrec, partials, kMaxPartials, s);
    launch_sum_partials(partials, nb, msum, s);
    HIP_TRY(hipGetLastError());
    if (res) launch_from_time_major(res_t, res, nrec, nt, s);
    double j = 0.0;
    HIP_TRY(hipMemcpyAsync(&j, msum, sizeof(double), hipMemcpyDeviceToHost, s));
    HIP_TRY(hipStreamSynchronize(s));
    if (misfit) *misfit += j;
  }

  if (grad) {
    // backward loop, Src/libCUFD.cu:469-613. Per step the reference runs
    //   velocity(rev) ; restore v ; source(-) ; stress(rev) ; restore s ;
    //   velocity_adj ; inject ; stress_adj
    // Here velocity_adj joins the reverse velocity and stress_adj the reverse
    // stress: neither reads what the kernels moved past it write (v, s and
    // their rings), and inject only has to precede stress_adj.
    zero_wave(A, N, s);
    zero_psi(P, N, s);
    Source S;
    S.cell = src_interior ? src_cell : -1;
    S.stf = stf_dev;
    launch_velocity_adj(form, G, F, C, s);
    launch_inject(A, rec_dev, nrec, nt - 1, res_t, s);
    launch_stress_adj(G, F, C, s);
    Fields9 F9r;
    fields9(W2, F9r);
    for (int seg = last_seg; seg >= 0; seg--) {
      const int a = seg * seg_len, b = std::min(nt - 2, a + seg_len - 1);
      // slots of a resident segment count from the first resident step; a
      // replayed one fills slots 0 .. seg_len - 1 (those of the first
      // resident segment, consumed by then)
      const int a_slot = seg >= res0 ? res0 * seg_len : a;
      if (seg < res0) {
        // replay the segment from its checkpoint to rebuild its ring
        if ((rc = checkpoint(seg, W2, P2, false, s))) return rc;
        prof.begin(K_RING, s);
        launch_ring(B, G, F9r, ring, 0, 0, 9, 0, s);
        prof.end(s);
        for (int it = a; it <= b; it++) forward_step(F2, W2, Sf, it, false, s, it < b ? it + 1 - a : -1);
      }
      for (int it = b; it >= a; it--) {
        // the restores of slot it - a_slot, folded into the kernels before
        // them where those can take them
        const long slot = it - a_slot;
        RingFold rf{};
        rf.B = B;
        rf.slot = ring + slot * 9 * B.n;
        prof.begin(K_VELOCITY_BWD, s);
        const bool fv = launch_velocity_bwd(tiles[K_VELOCITY_BWD], form, G, F, C, s, &rf, split_bwd == 1);
        prof.end(s);
        if (!fv) launch_ring(B, G, F9, ring, slot, F_VZ, F_VY + 1, 1, s);
        if (!src_interior) launch_add_source(W.szz, W.sxx, W.syy, src_cell, stf_dev, it, G.dt, -1, s);
        launch_inject(A, rec_dev, nrec, it, res_t, s);
        prof.begin(K_STRESS_BWD, s);
        const bool fs = launch_stress_bwd(tiles[K_STRESS_BWD], G, F, C, S, it, s, &rf, split_bwd == 1);
        prof.end(s);
        if (!fs) launch_ring(B, G, F9, ring, slot, F_SZZ, F_WAVE_COUNT, 1, s);
      }
    }
    HIP_TRY(hipGetLastError());
  }
  if (prof.on && (rc = prof.collect())) return rc;
  return EFWI_OK;
}

// Autotuner: time each stencil kernel of the time loop under every tile
// shape (rows per workgroup x planes per thread) on this grid, with HIP events
// on the stream, and keep the fastest; then the interior/rim split of the two
// backward kernels at their chosen shapes. The kernels run as in the time
// loops: on a live wavefield (every field, adjoint and C-PML array filled with
// values of a wavefield's magnitude) and with the boundary ring folded into
// them where the time loops fold it (a one-slot ring of its own). No source.
// The state is zeroed again afterwards.
int Ctx::autotune(int reps, hipStream_t s, double* report) {
  HIP_TRY(hipSetDevice(cfg.device));
  if (!model_set) return fail(EFWI_ERR_STATE, "efwi_set_model was not called");
  if (reps < 1) reps = 1;
  static const Tile cands[] = {{1, 1}, {2, 1}, {4, 1}, {1, 2}, {1, 4}, {2, 2},
                               {4, 16}, {4, 32}, {4, 64}, {8, 16}, {8, 32}, {8, 64}};
  const int ncand = sizeof(cands) / sizeof(cands[0]);
  {
    Fields9 fw, fa;
    fields9(W, fw);
    fields9(A, fa);
    for (int k = 0; k < 9; k++) {
      launch_fill_hash(fw.f[k], N, 17u + k, 1e-3f, s);
      launch_fill_hash(fa.f[k], N, 101u + k, 1e-3f, s);
    }
    float* const* pp = reinterpret_cast<float* const*>(&P);
    for (int k = 0; k < P_PSI_COUNT; k++) launch_fill_hash(pp[k], N, 211u + k, 1e-4f, s);
  }
  if (!tune_slot) {
    const int rc0 = alloc((void**)&tune_slot, sizeof(float) * 9 * (size_t)B.n);
    if (rc0) return rc0;
  }
  launch_fill_hash(tune_slot, 9 * B.n, 307u, 1e-3f, s);
  RingFold rf{};
  rf.B = B;
  rf.slot = tune_slot;
  Source none;
  none.cell = -1;
  none.stf = stf_dev;
  Fields9 f9;
  fields9(W, f9);
  hipEvent_t e0, e1;
  HIP_TRY(hipEventCreate(&e0));
  HIP_TRY(hipEventCreate(&e1));
  const int kinds[4] = {K_STRESS_FWD, K_VELOCITY_FWD, K_VELOCITY_BWD, K_STRESS_BWD};
  for (int q = 0; q < 4; q++) {
    const int k = kinds[q];
    double best = 1e30;
    Tile bt = tiles[k];
    for (int c = 0; c < ncand; c++) {
      const Tile t = cands[c];
      // a shape whose kernel cannot take the ring save or restore is timed
      // with the separate ring pass that the time loop then adds to it
      auto launch = [&]() {
        if (k == K_STRESS_FWD) {
          if (!launch_stress_fwd(t, G, F, C, none, 0, s, &rf))
            launch_ring(B, G, f9, tune_slot, 0, F_SZZ, F_WAVE_COUNT, 0, s);
        } else if (k == K_VELOCITY_FWD) {
          if (!launch_velocity_fwd(t, cfg.formulation, G, F, C, s, &rf))
            launch_ring(B, G, f9, tune_slot, 0, F_VZ, F_VY + 1, 0, s);
        } else if (k == K_VELOCITY_BWD) {
          if (!launch_velocity_bwd(t, cfg.formulation, G, F, C, s, &rf))
            launch_ring(B, G, f9, tune_slot, 0, F_VZ, F_VY + 1, 1, s);
        } else if (!launch_stress_bwd(t, G, F, C, none, 0, s, &rf)) {
          launch_ring(B, G, f9, tune_slot, 0, F_SZZ, F_WAVE_COUNT, 1, s);
        }
      };
      launch();  // warm-up
      HIP_TRY(hipEventRecord(e0, s));
      for (int r = 0; r < reps; r++) launch();
      HIP_TRY(hipEventRecord(e1, s));
      HIP_TRY(hipEventSynchronize(e1));
      float ms = 0.0f;
      HIP_TRY(hipEventElapsedTime(&ms, e0, e1));
      const double avg = ms / reps;
      if (report) report[q * ncand + c] = avg;
      if (avg < best) {
        best = avg;
        bt = t;
      }
    }
    tiles[k] = bt;
  }
  // the interior/rim split of the backward kernels at their chosen shapes
  split_bwd = 0;
  if (has_interior_segments(G) && tiles[K_VELOCITY_BWD].zdepth < 16 && tiles[K_STRESS_BWD].zdepth < 16) {
    double t2[2] = {0.0, 0.0};
    for (int sp = 0; sp < 2; sp++) {
      auto launch = [&]() {
        launch_velocity_bwd(tiles[K_VELOCITY_BWD], cfg.formulation, G, F, C, s, &rf, sp == 1);
        launch_stress_bwd(tiles[K_STRESS_BWD], G, F, C, none, 0, s, &rf, sp == 1);
      };
      launch();
      HIP_TRY(hipEventRecord(e0, s));
      for (int r = 0; r < reps; r++) launch();
      HIP_TRY(hipEventRecord(e1, s));
      HIP_TRY(hipEventSynchronize(e1));
      float ms = 0.0f;
      HIP_TRY(hipEventElapsedTime(&ms, e0, e1));
      t2[sp] = ms / reps;
    }
    split_bwd = t2[1] < t2[0] ? 1 : 0;
  }
  hipEventDestroy(e0);
  hipEventDestroy(e1);
  HIP_TRY(hipGetLastError());
  zero_wave(W, N, s);
  zero_wave(A, N, s);
  zero_psi(P, N, s);
  launch_fill(Gr.glam, N, 0.0f, s);
  launch_fill(Gr.gmu, N, 0.0f, s);
  launch_fill(Gr.grho, N, 0.0f, s);
  HIP_TRY(hipStreamSynchronize(s));
  return EFWI_OK;
}

int Ctx::get_gradient(float* gl, float* gm, float* gr, int on_host, hipStream_t s) {
  HIP_TRY(hipSetDevice(cfg.device));
  const hipMemcpyKind k = on_host ? hipMemcpyDeviceToHost : hipMemcpyDeviceToDevice;
  int rc;
  if (gl && (rc = copy_out(gl, Gr.glam, k, s))) return rc;
  if (gm && (rc = copy_out(gm, Gr.gmu, k, s))) return rc;
  if (gr && (rc = copy_out(gr, Gr.grho, k, s))) return rc;
  if (on_host) HIP_TRY(hipStreamSynchronize(s));
  return EFWI_OK;
}

}  // namespace efwi

// ===========================================================================
// C ABI
// ===========================================================================
using efwi::Ctx;
using efwi::fail;

extern "C" {

int efwi_version(void) { return 100; }

const char* efwi_last_error(void) { return efwi::g_err.c_str(); }

int efwi_create(const efwi_config* cfg, efwi_ctx** out) {
  if (!cfg || !out) return fail(EFWI_ERR_ARG, "NULL argument");
  *out = nullptr;
  Ctx* c = new Ctx();
  c->cfg = *cfg;
  int rc = c->init();
  if (rc) {
    delete c;
    return rc;
  }
  *out = reinterpret_cast<efwi_ctx*>(c);
  return EFWI_OK;
}

void efwi_destroy(efwi_ctx* ctx) { delete reinterpret_cast<Ctx*>(ctx); }

int efwi_set_model(efwi_ctx* ctx, const float* lambda, const float* mu, const float* rho, int on_host,
                   void* stream) {
  if (!ctx || !lambda || !mu || !rho) return fail(EFWI_ERR_ARG, "NULL argument");
  return reinterpret_cast<Ctx*>(ctx)->set_model(lambda, mu, rho, on_host, (hipStream_t)stream);
}

int efwi_zero_gradient(efwi_ctx* ctx, void* stream) {
  if (!ctx) return fail(EFWI_ERR_ARG, "NULL context");
  return reinterpret_cast<Ctx*>(ctx)->zero_gradient((hipStream_t)stream);
}

int efwi_run_shot(efwi_ctx* ctx, const efwi_shot* shot, int calc_id, const float* obs, float* syn,
                  float* res, double* misfit, void* stream) {
  if (!ctx) return fail(EFWI_ERR_ARG, "NULL context");
  return reinterpret_cast<Ctx*>(ctx)->run_shot(shot, calc_id, obs, syn, res, misfit, (hipStream_t)stream);
}

int efwi_get_gradient(efwi_ctx* ctx, float* g_lambda, float* g_mu, float* g_rho, int on_host,
                      void* stream) {
  if (!ctx) return fail(EFWI_ERR_ARG, "NULL context");
  return reinterpret_cast<Ctx*>(ctx)->get_gradient(g_lambda, g_mu, g_rho, on_host, (hipStream_t)stream);
}

int efwi_set_illumination(efwi_ctx* ctx, int enable, void* stream) {
  if (!ctx) return fail(EFWI_ERR_ARG, "NULL context");
  Ctx* c = reinterpret_cast<Ctx*>(ctx);
  HIP_TRY(hipSetDevice(c->cfg.device));
  c->illum_on = enable != 0;
  if (c->illum_on) {
    int rc;
    if (!c->illum && (rc = c->alloc((void**)&c->illum, sizeof(float) * (size_t)c->N))) return rc;
    efwi::launch_fill(c->illum, c->N, 0.0f, (hipStream_t)stream);
    HIP_TRY(hipGetLastError());
  }
  return EFWI_OK;
}

int efwi_get_illumination(efwi_ctx* ctx, float* out, int on_host, void* stream) {
  if (!ctx || !out) return fail(EFWI_ERR_ARG, "NULL argument");
  Ctx* c = reinterpret_cast<Ctx*>(ctx);
  if (!c->illum) return fail(EFWI_ERR_STATE, "illumination is off (efwi_set_illumination)");
  HIP_TRY(hipSetDevice(c->cfg.device));
  hipStream_t s = (hipStream_t)stream;
  int rc = c->copy_out(out, c->illum, on_host ? hipMemcpyDeviceToHost : hipMemcpyDeviceToDevice, s);
  if (rc) return rc;
  if (on_host) HIP_TRY(hipStreamSynchronize(s));
  return EFWI_OK;
}

int efwi_profile(efwi_ctx* ctx, int enable) {
  if (!ctx) return fail(EFWI_ERR_ARG, "NULL context");
  Ctx* c = reinterpret_cast<Ctx*>(ctx);
  c->prof.on = enable != 0;
  return EFWI_OK;
}

int efwi_profile_read(efwi_ctx* ctx, double* ms, long long* launches, int reset) {
  if (!ctx || !ms || !launches) return fail(EFWI_ERR_ARG, "NULL argument");
  Ctx* c = reinterpret_cast<Ctx*>(ctx);
  int rc = c->prof.collect();
  if (rc) return rc;
  for (int k = 0; k < efwi::K_COUNT; k++) {
    ms[k] = c->prof.ms[k];
    launches[k] = c->prof.launches[k];
    if (reset) {
      c->prof.ms[k] = 0.0;
      c->prof.launches[k] = 0;
    }
  }
  return EFWI_OK;
}

int efwi_set_tile(efwi_ctx* ctx, int kernel, int rows, int zdepth) {
  if (!ctx) return fail(EFWI_ERR_ARG, "NULL context");
  if (kernel < 0 || kernel > efwi::K_STRESS_ADJ) return fail(EFWI_ERR_ARG, "kernel index out of range");
  if (zdepth < 1 || zdepth > 256) return fail(EFWI_ERR_ARG, "zdepth must be in [1, 256]");
  if (zdepth < 16 && rows != 1 && rows != 2 && rows != 4)
    return fail(EFWI_ERR_ARG, "rows must be 1, 2 or 4 (flat sweep, zdepth < 16)");
  if (zdepth >= 16 && rows != 1 && rows != 2 && rows != 4 && rows != 8)
    return fail(EFWI_ERR_ARG, "rows must be 1, 2, 4 or 8 (z-marching sweep, zdepth >= 16)");
  reinterpret_cast<Ctx*>(ctx)->tiles[kernel] = efwi::Tile{rows, zdepth};
  return EFWI_OK;
}

int efwi_set_split(efwi_ctx* ctx, int on) {
  if (!ctx) return fail(EFWI_ERR_ARG, "NULL context");
  if (on != 0 && on != 1) return fail(EFWI_ERR_ARG, "on must be 0 or 1");
  reinterpret_cast<Ctx*>(ctx)->split_bwd = on;
  return EFWI_OK;
}

int efwi_get_split(const efwi_ctx* ctx, int* on) {
  if (!ctx || !on) return fail(EFWI_ERR_ARG, "NULL argument");
  *on = reinterpret_cast<const Ctx*>(ctx)->split_bwd;
  return EFWI_OK;
}

int efwi_set_resident_segments(efwi_ctx* ctx, int segments) {
  if (!ctx) return fail(EFWI_ERR_ARG, "NULL context");
  if (segments < 0) return fail(EFWI_ERR_ARG, "segments must be >= 0 (0: automatic)");
  Ctx* c = reinterpret_cast<Ctx*>(ctx);
  HIP_TRY(hipSetDevice(c->cfg.device));
  c->res_req = segments;
  return c->cfg.memory_mode == 1 ? c->choose_resident() : EFWI_OK;
}

int efwi_get_resident_segments(const efwi_ctx* ctx, int* segments, int* total) {
  if (!ctx || !segments || !total) return fail(EFWI_ERR_ARG, "NULL argument");
  const Ctx* c = reinterpret_cast<const Ctx*>(ctx);
  *segments = c->res_seg;
  *total = c->nseg;
  return EFWI_OK;
}

int efwi_set_fusion(efwi_ctx* ctx, int pass, int rows, int zchunk) {
  if (!ctx) return fail(EFWI_ERR_ARG, "NULL context");
  if (pass != 0) return fail(EFWI_ERR_ARG, "pass must be 0 (forward)");
  if (rows != 0 && rows != 4 && rows != 8 && rows != 12) return fail(EFWI_ERR_ARG, "rows must be 0, 4, 8 or 12");
  if (zchunk < 1 || zchunk > 4096) return fail(EFWI_ERR_ARG, "zchunk must be in [1, 4096]");
  reinterpret_cast<Ctx*>(ctx)->tiles[efwi::K_FWD_FUSED] = efwi::Tile{rows, zchunk};
  return EFWI_OK;
}

int efwi_get_tile(const efwi_ctx* ctx, int kernel, int* rows, int* zdepth) {
  if (!ctx || !rows || !zdepth) return fail(EFWI_ERR_ARG, "NULL argument");
  if (kernel < 0 || kernel >= efwi::K_COUNT) return fail(EFWI_ERR_ARG, "kernel index out of range");
  const efwi::Tile t = reinterpret_cast<const Ctx*>(ctx)->tiles[kernel];
  *rows = t.rows;
  *zdepth = t.zdepth;
  return EFWI_OK;
}

int efwi_autotune(efwi_ctx* ctx, int reps, double* report, void* stream) {
  if (!ctx) return fail(EFWI_ERR_ARG, "NULL context");
  return reinterpret_cast<Ctx*>(ctx)->autotune(reps, (hipStream_t)stream, report);
}

int efwi_dot64(const float* a, const float* b, long long n, double* out, void* stream) {
  if (!a || !b || !out || n < 0) return fail(EFWI_ERR_ARG, "bad argument");
  hipStream_t s = (hipStream_t)stream;
  double* buf = nullptr;
  const int maxb = 2048;
  hipError_t e = hipMallocAsync((void**)&buf, sizeof(double) * (maxb + 1), s);
  if (e != hipSuccess) return fail(EFWI_ERR_NOMEM, std::string("dot workspace: ") + hipGetErrorString(e));
  const int nb = efwi::launch_dot64(a, b, (long)n, buf + 1, maxb, s);
  efwi::launch_sum_partials(buf + 1, nb, buf, s);
  e = hipMemcpyAsync(out, buf, sizeof(double), hipMemcpyDeviceToHost, s);
  hipFreeAsync(buf, s);
  if (e != hipSuccess) return fail(EFWI_ERR_HIP, std::string("dot result copy: ") + hipGetErrorString(e));
  e = hipStreamSynchronize(s);
  if (e != hipSuccess) return fail(EFWI_ERR_HIP, std::string("dot synchronize: ") + hipGetErrorString(e));
  return EFWI_OK;
}

long long efwi_device_bytes(const efwi_ctx* ctx) {
  return ctx ? reinterpret_cast<const Ctx*>(ctx)->bytes_held : 0;
}

int efwi_cpml_profile(int n, int npml, float h, float f0, float dt, float* out) {
  if (n <= 0 || !out) return fail(EFWI_ERR_ARG, "bad argument");
  efwi::cpml_profile_1d(n, npml, h, f0, dt, out, out + n, out + 2 * n, out + 3 * n, out + 4 * n, out + 5 * n);
  return EFWI_OK;
}

int efwi_courant_number(double vp_max, float dt, float dz, float dx, float dy, double* out) {
  if (!out || !(dt > 0) || !(dz > 0) || !(dx > 0) || !(dy > 0)) return fail(EFWI_ERR_ARG, "bad argument");
  *out = efwi::courant_number(vp_max, dt, dz, dx, dy);
  return EFWI_OK;
}

int efwi_window_factors(int nt, float dt, int nrec, const float* win_start, const float* win_end,
                        const float* weights, float src_weight, float* out) {
  if (nt <= 0 || nrec < 0 || !out || (nrec > 0 && (!win_start || !win_end))) return fail(EFWI_ERR_ARG, "bad argument");
  for (int r = 0; r < nrec; r++)
    for (int it = 0; it < nt; it++)
      out[(size_t)r * nt + it] = efwi::taper_window_fac(it, nt, dt, 0.005f, win_start[r], win_end[r],
                                                        weights ? weights[r] : 1.0f, src_weight);
  return EFWI_OK;
}

int efwi_taper_w2(int nt, float dt, float ratio, float* out) {
  if (nt <= 0 || !out) return fail(EFWI_ERR_ARG, "bad argument");
  for (int it = 0; it < nt; it++) out[it] = efwi::taper_plain_w2(it, nt, dt, ratio);
  return EFWI_OK;
}

}  // extern "C"
