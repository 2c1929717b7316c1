// Batched shots: several shots of one grid and model advanced together. Every
// stencil step of all shots is one persistent launch over a (shot, tile) work
// queue (efwi_stencil.hip), the receiver, ring and injection work of all
// shots one launch each, and each time step's launch sequence is captured
// once in a hipGraph and replayed nt - 1 times with the time step kept in
// device memory. Per shot the arithmetic is that of efwi_run_shot in the
// reverse-reconstruction mode (memory_mode 0), so each shot's seismograms,
// misfit and gradient equal a single-shot run's bit for bit.
#include <algorithm>
#include <string>
#include <vector>

#include "efwi_host.h"

namespace efwi {

void Ctx::release_batch() {
  if (batch.gf) hipGraphExecDestroy(batch.gf);
  if (batch.gb) hipGraphExecDestroy(batch.gb);
  if (batch.cs) hipStreamDestroy(batch.cs);
  batch.gf = batch.gb = nullptr;
  batch.cs = nullptr;
}

// Shots per persistent launch, or 0: one shot after another through the
// single-shot kernels. The persistent launches pay for their work queue (an
// atomic claim per few tiles) and, over many shots, lose the Infinity Cache
// reuse between consecutive kernels that a single shot has; they win where a
// shot alone leaves the device idle. Automatic (group = 0): where one shot
// already gives kBatchFill workgroup tiles per compute unit, the single-shot
// kernels; otherwise groups of enough shots to reach that many tiles.
int Ctx::batch_group_size(int n) const {
  if (batch.group > 0) return std::min(batch.group, n);
  if (batch.group < 0) return 0;
  const Tile t = tiles[K_STRESS_FWD].zdepth >= 16 ? Tile{4, 1} : tiles[K_STRESS_FWD];
  const long per_shot = make_sweep(2, cfg.nz - 3, 2, cfg.nx - 3, 2, cfg.ny - 3, t).blocks;
  const long want = (long)kBatchFill * std::max(batch.cus, 1);
  if (per_shot >= want) return 0;
  return (int)std::max(1L, std::min((long)n, (want + per_shot - 1) / std::max(per_shot, 1L)));
}

// Point the context's state (wavefields, C-PML arrays, model and gradients)
// at the field groups g.
void Ctx::bind_state(const Groups& g) {
  const size_t fsf = G.fs / sizeof(float);
  F = g;
  wave_of(F.w, fsf, W);
  wave_of(F.a, fsf, A);
  psi_of(F.p, fsf, P);
  lam = F.m + M_LAM * fsf; mu = F.m + M_MU * fsf; byc = F.m + M_BYC * fsf;
  Gr.glam = F.m + M_GLAM * fsf; Gr.gmu = F.m + M_GMU * fsf; Gr.grho = F.m + M_GRHO * fsf;
  M.lam = lam; M.mu = mu; M.byc = byc;
}

int Ctx::ensure_batch(int n, int nring) {
  int rc;
  if (!batch.cs) {
    HIP_TRY(hipStreamCreateWithFlags(&batch.cs, hipStreamNonBlocking));
    if ((rc = alloc((void**)&batch.counters, sizeof(unsigned) * 16 * kQueueStride))) return rc;
    if ((rc = alloc((void**)&batch.it_dev, sizeof(int) * 64))) return rc;
    batch.cus = device_cus(cfg.device);
  }
  const size_t ring_floats = (size_t)B.n * 9 * (size_t)(cfg.nt - 1);
  // rings: one per shot of a group (groups run one after another)
  while ((int)batch.ring.size() < std::min(n, nring)) {
    float* ring_k = nullptr;
    if ((rc = alloc((void**)&ring_k, sizeof(float) * ring_floats))) return rc;
    batch.ring.push_back(ring_k);
  }
  if (n <= batch.n) return EFWI_OK;
  const size_t fs = G.fs;
  for (int k = batch.n; k < n; k++) {
    Groups g{};
    if ((rc = alloc((void**)&g.w, fs * F_WAVE_COUNT)) || (rc = alloc((void**)&g.a, fs * F_WAVE_COUNT)) ||
        (rc = alloc((void**)&g.p, fs * P_PSI_COUNT)) || (rc = alloc((void**)&g.m, fs * M_COUNT)))
      return rc;
    batch.F.push_back(g);
    float* stf_k = nullptr;
    if ((rc = alloc((void**)&stf_k, sizeof(float) * cfg.nt))) return rc;
    batch.stf.push_back(stf_k);
    batch.syn_t.push_back(nullptr);
    batch.obs_t.push_back(nullptr);
    batch.res_t.push_back(nullptr);
    batch.rec.push_back(nullptr);
    batch.seis_cap.push_back(0);
    batch.rec_cap.push_back(0);
    batch.nrec.push_back(0);
  }
  if ((rc = alloc((void**)&batch.shots_dev, sizeof(ShotDev) * n))) return rc;
  batch.n = n;
  return EFWI_OK;
}

int Ctx::run_batch(const efwi_shot* shots, int n, int calc_id, const float* const* obs, float* const* syn,
                   double* misfits, int use_graph, hipStream_t s) {
  if (n < 1 || !shots) return fail(EFWI_ERR_ARG, "a batch needs at least one shot");
  if (calc_id < 0 || calc_id > 2) return fail(EFWI_ERR_ARG, "invalid calc_id " + std::to_string(calc_id));
  if (!model_set) return fail(EFWI_ERR_STATE, "efwi_set_model was not called");
  if (calc_id != EFWI_CALC_SYNTHETIC && !obs) return fail(EFWI_ERR_ARG, "observed data required");
  if (cfg.if_win) return fail(EFWI_ERR_ARG, "batched shots take the plain misfit (if_win = 0)");
  if (cfg.memory_mode == 1) return fail(EFWI_ERR_ARG, "batched shots keep the full ring (memory_mode 0)");
  HIP_TRY(hipSetDevice(cfg.device));
  HIP_TRY(hipStreamSynchronize(s));
  if (!batch.cs) {
    int rc0 = ensure_batch(0, 0);
    if (rc0) return rc0;
  }
  const int gsz0 = batch_group_size(n);
  const int gsz = gsz0 > 0 ? gsz0 : 1;
  int rc = ensure_batch(n, calc_id == EFWI_CALC_GRADIENT && gsz0 > 0 ? gsz : 0);
  if (rc) return rc;
  const int nt = cfg.nt, form = cfg.formulation, p = cfg.npml;
  const bool grad = calc_id == EFWI_CALC_GRADIENT;
  hipStream_t q = batch.cs;
  const size_t fsf = G.fs / sizeof(float);
  // ---- shot table
  std::vector<ShotDev> tab(n);
  for (int k = 0; k < n; k++) {
    const efwi_shot* sh = shots + k;
    if (!sh->stf) return fail(EFWI_ERR_ARG, "shot " + std::to_string(k) + ": stf is NULL");
    if (sh->nrec < 0 || (sh->nrec > 0 && !sh->rec)) return fail(EFWI_ERR_ARG, "shot " + std::to_string(k) + ": bad receivers");
    if (calc_id != EFWI_CALC_SYNTHETIC && !obs[k]) return fail(EFWI_ERR_ARG, "observed data of a shot is NULL");
    auto inside = [&](int z, int x, int y) {
      return z >= 0 && z < cfg.nz && x >= 0 && x < cfg.nx && y >= 0 && y < cfg.ny;
    };
    const bool interior = sh->src_z >= p && sh->src_z <= cfg.nz - 1 - p && sh->src_x >= p &&
                          sh->src_x <= cfg.nx - 1 - p && sh->src_y >= p && sh->src_y <= cfg.ny - 1 - p;
    if (!interior) return fail(EFWI_ERR_ARG, "batched shots need their source inside the C-PML-free interior");
    std::vector<long> cells(sh->nrec);
    for (int r = 0; r < sh->nrec; r++) {
      const int z = sh->rec[3 * r], x = sh->rec[3 * r + 1], y = sh->rec[3 * r + 2];
      if (!inside(z, x, y)) return fail(EFWI_ERR_ARG, "shot " + std::to_string(k) + ": receiver outside the grid");
      cells[r] = (long)z * G.pxy + (long)x * G.py + y;
    }
    const int nrec_k = sh->nrec;
    batch.nrec[k] = nrec_k;
    std::vector<float> stf(nt);
    for (int it = 0; it < nt; it++) stf[it] = sh->stf[it] * taper_plain_w2(it, nt, cfg.dt, 0.001f);
    HIP_TRY(hipMemcpy(batch.stf[k], stf.data(), sizeof(float) * nt, hipMemcpyHostToDevice));
    const size_t nrb = (size_t)std::max(nrec_k, 1);
    if ((rc = grow(this, (void**)&batch.rec[k], &batch.rec_cap[k], sizeof(long) * nrb))) return rc;
    if (nrec_k) HIP_TRY(hipMemcpy(batch.rec[k], cells.data(), sizeof(long) * nrec_k, hipMemcpyHostToDevice));
    const size_t seis = sizeof(float) * (size_t)nt * 3 * nrb;
    if (batch.seis_cap[k] < seis) {
      size_t c1 = batch.seis_cap[k], c2 = batch.seis_cap[k], c3 = batch.seis_cap[k];
      if ((rc = grow(this, (void**)&batch.syn_t[k], &c1, seis)) || (rc = grow(this, (void**)&batch.obs_t[k], &c2, seis)) ||
          (rc = grow(this, (void**)&batch.res_t[k], &c3, seis)))
        return rc;
      batch.seis_cap[k] = seis;
    }
    ShotDev& d = tab[k];
    d.F = batch.F[k];
    d.cell = (long)sh->src_z * G.pxy + (long)sh->src_x * G.py + sh->src_y;
    d.stf = batch.stf[k];
    for (int f = 0; f < 9; f++) {
      d.w9[f] = batch.F[k].w + f * fsf;
      d.a9[f] = batch.F[k].a + f * fsf;
    }
    d.rec = batch.rec[k];
    d.nrec = nrec_k;
    d.syn_t = batch.syn_t[k];
    d.res_t = batch.res_t[k];
    d.ring = calc_id == EFWI_CALC_GRADIENT && gsz0 > 0 ? batch.ring[k % gsz] : nullptr;
  }
  HIP_TRY(hipMemcpy(batch.shots_dev, tab.data(), sizeof(ShotDev) * n, hipMemcpyHostToDevice));
  // ---- state: the model into every shot's model group, zero wavefields and gradients
  for (int k = 0; k < n; k++) {
    const Groups& g = batch.F[k];
    HIP_TRY(hipMemcpyAsync(g.m, F.m, G.fs * 3, hipMemcpyDeviceToDevice, q));  // lambda, mu, buoyancy
    launch_fill(g.m + M_GLAM * fsf, (long)fsf * 3, 0.0f, q);
    launch_fill(g.w, (long)fsf * F_WAVE_COUNT, 0.0f, q);
    launch_fill(g.a, (long)fsf * F_WAVE_COUNT, 0.0f, q);
    launch_fill(g.p, (long)fsf * P_PSI_COUNT, 0.0f, q);
    launch_fill(batch.syn_t[k], (long)nt * 3 * std::max(batch.nrec[k], 1), 0.0f, q);
  }
  HIP_TRY(hipGetLastError());
  if (gsz0 == 0) {
    // one shot after another on the single-shot kernels, each on its own state
    // (its model copy and gradients in batch.F[k]); the context's state, fused
    // step and illumination are left as they were
    const Groups own = F;
    const Tile fused = tiles[K_FWD_FUSED];
    const bool illum0 = illum_on;
    tiles[K_FWD_FUSED].rows = 0;
    illum_on = false;
    for (int k = 0; k < n && !rc; k++) {
      bind_state(batch.F[k]);
      double j = 0.0;
      rc = run_shot(shots + k, calc_id, obs ? obs[k] : nullptr, syn ? syn[k] : nullptr, nullptr, &j, q);
      if (misfits) misfits[k] = calc_id == EFWI_CALC_SYNTHETIC ? 0.0 : j;
    }
    bind_state(own);
    tiles[K_FWD_FUSED] = fused;
    illum_on = illum0;
    if (rc) return rc;
    HIP_TRY(hipStreamSynchronize(q));
    return EFWI_OK;
  }
  for (int g0 = 0; g0 < n; g0 += gsz) {
  const int gn = std::min(gsz, n - g0);
  int gmax_nrec = 0;
  for (int k = g0; k < g0 + gn; k++) gmax_nrec = std::max(gmax_nrec, batch.nrec[k]);
  BatchLaunch L0{batch.shots_dev + g0, gn, batch.counters, batch.it_dev, batch.cus, batch.take, batch.static_pct, 0};
  BatchLaunch L1 = L0;
  L1.counters = batch.counters + 8 * kQueueStride;
  L1.reverse = 1;
  const std::string tiles_key = std::to_string(tiles[K_STRESS_FWD].rows) + "x" + std::to_string(tiles[K_STRESS_FWD].zdepth) +
                                "," + std::to_string(tiles[K_VELOCITY_FWD].rows) + "x" + std::to_string(tiles[K_VELOCITY_FWD].zdepth) +
                                "," + std::to_string(tiles[K_VELOCITY_BWD].rows) + "x" + std::to_string(tiles[K_VELOCITY_BWD].zdepth) +
                                "," + std::to_string(tiles[K_STRESS_BWD].rows) + "x" + std::to_string(tiles[K_STRESS_BWD].zdepth);
  const std::string key = std::to_string(n) + "|" + std::to_string(g0) + "+" + std::to_string(gn) + "|" +
                          std::to_string(gmax_nrec) + "|" + tiles_key + "|" +
                          std::to_string(batch.take) + "|" + std::to_string(batch.static_pct) + "|";
  // ---- forward loop (Src/libCUFD.cu:312-354 per shot): ring save of the
  // state before step it, stress + source, velocity, record of step it + 1.
  // With the ring folded, the stencil kernels of step it save the state they
  // produce (slot it + 1) and slot 0 is saved once before the loop.
  const bool fold = grad && ring_foldable(G, B);
  BatchRing save{};
  save.B = B;
  save.mode = 1;
  save.last = nt - 2;
  BatchRing restore = save;
  restore.mode = 2;
  auto fwd_step = [&](hipStream_t st) {
    if (grad && !fold) launch_ring_batch(B, G, batch.shots_dev + g0, gn, batch.it_dev, 0, 9, 0, st);
    launch_fill(reinterpret_cast<float*>(batch.counters), 8 * kQueueStride, 0.0f, st);
    launch_stress_fwd_batch(tiles[K_STRESS_FWD], G, C, L0, st, fold ? &save : nullptr);
    launch_fill(reinterpret_cast<float*>(batch.counters + 8 * kQueueStride), 8 * kQueueStride, 0.0f, st);
    launch_velocity_fwd_batch(tiles[K_VELOCITY_FWD], form, G, C, L1, st, fold ? &save : nullptr);
    launch_record_batch(batch.shots_dev + g0, gn, gmax_nrec, batch.it_dev, 1, st);
    launch_step_add(batch.it_dev, 1, st);
  };
  auto run_steps = [&](hipGraphExec_t* exec, std::string* have, const std::string& want, int count,
                       auto&& step) -> int {
    if (!use_graph) {
      for (int i = 0; i < count; i++) step(q);
      return EFWI_OK;
    }
    if (*exec == nullptr || *have != want) {
      if (*exec) HIP_TRY(hipGraphExecDestroy(*exec));
      *exec = nullptr;
      hipGraph_t g = nullptr;
      HIP_TRY(hipStreamBeginCapture(q, hipStreamCaptureModeThreadLocal));
      step(q);
      HIP_TRY(hipStreamEndCapture(q, &g));
      hipError_t e = hipGraphInstantiate(exec, g, nullptr, nullptr, 0);
      hipGraphDestroy(g);
      if (e != hipSuccess) return fail(EFWI_ERR_HIP, std::string("hipGraphInstantiate: ") + hipGetErrorString(e));
      *have = want;
    }
    for (int i = 0; i < count; i++) HIP_TRY(hipGraphLaunch(*exec, q));
    return EFWI_OK;
  };
  const int zero = 0;
  HIP_TRY(hipMemcpyAsync(batch.it_dev, &zero, sizeof(int), hipMemcpyHostToDevice, q));
  if (fold) launch_ring_batch(B, G, batch.shots_dev + g0, gn, batch.it_dev, 0, 9, 0, q);
  if ((rc = run_steps(&batch.gf, &batch.key_f, key + (grad ? "g" : "f"), nt - 1, fwd_step))) return rc;
  HIP_TRY(hipGetLastError());
  for (int k = g0; k < g0 + gn; k++)
    if (syn && syn[k] && batch.nrec[k] > 0) launch_from_time_major(batch.syn_t[k], syn[k], batch.nrec[k], nt, q);
  // ---- misfit and adjoint source of every shot (Src/libCUFD.cu:371-441)
  if (calc_id != EFWI_CALC_SYNTHETIC) {
    for (int k = g0; k < g0 + gn; k++) {
      double j = 0.0;
      if (batch.nrec[k] > 0) {
        launch_to_time_major(obs[k], batch.obs_t[k], batch.nrec[k], nt, q);
        const int nb = launch_misfit(batch.obs_t[k], batch.syn_t[k], batch.res_t[k], w2_dev, 0, nt, batch.nrec[k],
                                     partials, kMaxPartials, q);
        launch_sum_partials(partials, nb, msum, q);
        HIP_TRY(hipMemcpyAsync(&j, msum, sizeof(double), hipMemcpyDeviceToHost, q));
        HIP_TRY(hipStreamSynchronize(q));
      }
      if (misfits) misfits[k] = j;
    }
  }
  // ---- backward loop (Src/libCUFD.cu:469-613 per shot), as Ctx::run_shot
  if (grad) {
    for (int k = g0; k < g0 + gn; k++) {
      const Groups& g = batch.F[k];
      launch_fill(g.a, (long)fsf * F_WAVE_COUNT, 0.0f, q);
      launch_fill(g.p, (long)fsf * P_PSI_COUNT, 0.0f, q);
      launch_velocity_adj(form, G, g, C, q);
      Wave Ak;
      wave_of(g.a, fsf, Ak);
      launch_inject(Ak, batch.rec[k], batch.nrec[k], nt - 1, batch.res_t[k], q);
      launch_stress_adj(G, g, C, q);
    }
    auto bwd_step = [&](hipStream_t st) {
      launch_fill(reinterpret_cast<float*>(batch.counters), 8 * kQueueStride, 0.0f, st);
      launch_velocity_bwd_batch(tiles[K_VELOCITY_BWD], form, G, C, L0, st, fold ? &restore : nullptr);
      if (!fold) launch_ring_batch(B, G, batch.shots_dev + g0, gn, batch.it_dev, 0, 3, 1, st);
      launch_inject_batch(batch.shots_dev + g0, gn, gmax_nrec, batch.it_dev, st);
      launch_fill(reinterpret_cast<float*>(batch.counters + 8 * kQueueStride), 8 * kQueueStride, 0.0f, st);
      launch_stress_bwd_batch(tiles[K_STRESS_BWD], G, C, L1, st, fold ? &restore : nullptr);
      if (!fold) launch_ring_batch(B, G, batch.shots_dev + g0, gn, batch.it_dev, 3, 9, 1, st);
      launch_step_add(batch.it_dev, -1, st);
    };
    const int start = nt - 2;
    HIP_TRY(hipMemcpyAsync(batch.it_dev, &start, sizeof(int), hipMemcpyHostToDevice, q));
    if ((rc = run_steps(&batch.gb, &batch.key_b, key + "b", nt - 1, bwd_step))) return rc;
    HIP_TRY(hipGetLastError());
  }
  }  // groups of shots
  HIP_TRY(hipStreamSynchronize(q));
  return EFWI_OK;
}

int Ctx::batch_gradient(int shot, float* gl, float* gm, float* gr, int on_host, hipStream_t s) {
  if (shot < 0 || shot >= batch.n) return fail(EFWI_ERR_ARG, "no such shot in the batch");
  HIP_TRY(hipSetDevice(cfg.device));
  const size_t fsf = G.fs / sizeof(float);
  const hipMemcpyKind k = on_host ? hipMemcpyDeviceToHost : hipMemcpyDeviceToDevice;
  const float* m = batch.F[shot].m;
  int rc;
  if (gl && (rc = copy_out(gl, m + M_GLAM * fsf, k, s))) return rc;
  if (gm && (rc = copy_out(gm, m + M_GMU * fsf, k, s))) return rc;
  if (gr && (rc = copy_out(gr, m + M_GRHO * fsf, k, s))) return rc;
  if (on_host) HIP_TRY(hipStreamSynchronize(s));
  return EFWI_OK;
}

}  // namespace efwi

using efwi::Ctx;
using efwi::fail;

extern "C" {

int efwi_run_batch(efwi_ctx* ctx, const efwi_shot* shots, int nshots, int calc_id, const float* const* obs,
                   float* const* syn, double* misfits, int use_graph, void* stream) {
  if (!ctx) return fail(EFWI_ERR_ARG, "NULL context");
  return reinterpret_cast<Ctx*>(ctx)->run_batch(shots, nshots, calc_id, obs, syn, misfits, use_graph,
                                                (hipStream_t)stream);
}

int efwi_batch_config(efwi_ctx* ctx, int take, int static_pct) {
  if (!ctx) return fail(EFWI_ERR_ARG, "NULL context");
  if (take < 1 || take > 4096) return fail(EFWI_ERR_ARG, "take must be in [1, 4096]");
  if (static_pct < 0 || static_pct > 100) return fail(EFWI_ERR_ARG, "static_pct must be in [0, 100]");
  reinterpret_cast<Ctx*>(ctx)->batch.take = take;
  reinterpret_cast<Ctx*>(ctx)->batch.static_pct = static_pct;
  return EFWI_OK;
}

int efwi_batch_group(efwi_ctx* ctx, int shots_per_launch) {
  if (!ctx) return fail(EFWI_ERR_ARG, "NULL context");
  if (shots_per_launch < -1) return fail(EFWI_ERR_ARG, "shots_per_launch must be >= -1");
  reinterpret_cast<Ctx*>(ctx)->batch.group = shots_per_launch;
  return EFWI_OK;
}

int efwi_batch_gradient(efwi_ctx* ctx, int shot, float* g_lambda, float* g_mu, float* g_rho, int on_host,
                        void* stream) {
  if (!ctx) return fail(EFWI_ERR_ARG, "NULL context");
  return reinterpret_cast<Ctx*>(ctx)->batch_gradient(shot, g_lambda, g_mu, g_rho, on_host, (hipStream_t)stream);
}

}  // extern "C"
