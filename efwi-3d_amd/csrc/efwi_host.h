// Context of the gfx950 elastic FWI library (host side).
#pragma once

#include <hip/hip_runtime.h>

#include <string>
#include <vector>

#include "../../include/efwi.h"
#include "efwi_common.h"
#include "efwi_kernels.h"

namespace efwi {

#define HIP_TRY(call)                                                                   \
  do {                                                                                  \
    hipError_t e_ = (call);                                                             \
    if (e_ != hipSuccess)                                                               \
      return fail(EFWI_ERR_HIP, std::string(#call) + ": " + hipGetErrorString(e_));     \
  } while (0)

int fail(int code, const std::string& msg);
void cpml_profile_1d(int n, int npml, float h, float f0, float dt, float* K, float* a, float* b,
                     float* Kh, float* ah, float* bh);
float taper_plain_w2(int it, int nt, float dt, float ratio);
float taper_window_fac(int it, int nt, float dt, float ratio, float t0, float t3, float weight,
                       float src_weight);

// Optional per-kernel timing with HIP events on the launch stream.
enum KernelId { K_STRESS_FWD, K_VELOCITY_FWD, K_STRESS_BWD, K_VELOCITY_BWD, K_VELOCITY_ADJ, K_STRESS_ADJ,
                K_RING, K_OTHER, K_FWD_FUSED, K_BWD_FUSED, K_COUNT };

struct Profiler {
  bool on = false;
  std::vector<hipEvent_t> pool;
  std::vector<int> kind;  // kernel id of each recorded pair
  size_t used = 0;        // pairs recorded since the last collect
  double ms[K_COUNT] = {0};
  long long launches[K_COUNT] = {0};
  void begin(int k, hipStream_t s);
  void end(hipStream_t s);
  int collect();  // synchronises on the recorded events and accumulates
  void release();
};

// Batched shots (efwi_run_batch): per-shot state, the device shot table, the
// per-XCD queue heads, the device time step, a capture stream and the two
// captured per-step graphs.
struct BatchState {
  int n = 0;
  std::vector<Groups> F;
  std::vector<float*> ring, stf, syn_t, obs_t, res_t;
  std::vector<long*> rec;
  std::vector<size_t> seis_cap, rec_cap;
  std::vector<int> nrec;
  ShotDev* shots_dev = nullptr;
  unsigned* counters = nullptr;  // 2 launches per step x 8 queues
  int* it_dev = nullptr;
  hipStream_t cs = nullptr;
  hipGraphExec_t gf = nullptr, gb = nullptr;
  std::string key_f, key_b;
  int cus = 0;
  int take = 2;        // queue items per atomic (efwi_batch_config)
  int static_pct = 0;  // static share of the queues, percent
  int group = 0;       // shots per persistent launch (efwi_batch_group), 0 = automatic
};

struct Ctx {
  static constexpr int kMaxPartials = 4096;
  Profiler prof;
  Tile tiles[K_COUNT] = {{1, 1}, {1, 1}, {1, 1}, {1, 1}, {1, 1}, {1, 1}, {1, 1}, {1, 1}, {0, 16}, {0, 16}};
  // fused steps (efwi_set_fusion): tiles[K_FWD_FUSED] = {rows, z planes per
  // workgroup}, rows 0 = off. They advance the state from F into Falt (w and p
  // groups), and the two sets are swapped after every step.
  Groups Falt{};
  // interior/rim split of the two backward kernels (efwi_set_split; chosen by
  // efwi_autotune): 0 off, 1 on
  int split_bwd = 0;
  float* tune_slot = nullptr;  // the autotuner's one-slot ring
  efwi_config cfg{};
  Geom G{};
  long N = 0;
  Wave W{}, A{};
  Psi P{};
  Model M{};
  Grad Gr{};
  Prof C{};
  Groups F{};   // field groups of the state; F2: replay copies of w and p (memory_mode 1)
  Groups F2{};
  Band B{};
  float *lam = nullptr, *mu = nullptr, *byc = nullptr;
  bool model_set = false;
  // ring of the reverse reconstruction: [res_seg * seg_len][9][B.n]. The
  // last res_seg segments of the time loop keep the ring the forward pass
  // writes (all of it in memory_mode 0, where there is one segment); in
  // memory_mode 1 the earlier segments are replayed from checkpoints into
  // the first seg_len slots once the resident segments are consumed.
  float* ring = nullptr;
  size_t ring_bytes = 0;
  int seg_len = 0, nseg = 1;
  int res_seg = 1;  // resident segments (1 <= res_seg <= nseg)
  int res_req = 0;  // requested resident segments, 0 = automatic (from the free device memory)
  int first_res() const { return nseg - res_seg; }
  // memory_mode 1: checkpoints [nseg - res_seg][ckpt_floats] and the replay state
  float* ckpt = nullptr;
  size_t ckpt_bytes = 0, ckpt_floats = 0;
  Wave W2{};
  Psi P2{};
  // per-shot buffers
  float* stf_dev = nullptr;
  float* w2_dev = nullptr;
  long* rec_dev = nullptr;
  size_t rec_cap = 0;
  float* syn_t = nullptr;
  size_t syn_cap = 0;
  float* obs_t = nullptr;
  size_t obs_cap = 0;
  float* res_t = nullptr;
  size_t res_cap = 0;
  float* fac_dev = nullptr;
  size_t fac_cap = 0;
  float* illum = nullptr;  // source illumination sum_t |v|^2 over shots (efwi_set_illumination)
  bool illum_on = false;
  double* partials = nullptr;
  double* msum = nullptr;
  int nrec = 0;
  long src_cell = 0;
  bool src_interior = true;   // inside [npml, n-1-npml]^3: subtraction fused into the reverse stress
  bool src_stencil = true;    // inside [2, n-3]^3: injection fused into the forward stress
  bool per_trace = false;
  std::vector<void*> allocs;
  long long bytes_held = 0;

  ~Ctx();
  int alloc(void** p, size_t bytes);
  // user volume [nz][nx][ny] (contiguous) <-> device field (rows G.py apart)
  int copy_in(float* field, const float* user, hipMemcpyKind k, hipStream_t s);
  int copy_out(float* user, const float* field, hipMemcpyKind k, hipStream_t s);
  int init();
  int plan_memory();
  size_t ckpt_floats_needed() const;
  int choose_resident();
  int ensure_ring();
  int ensure_checkpoints();
  size_t slab_floats(int axis) const;
  int checkpoint(int k, const Wave& Wf, const Psi& Pf, bool save, hipStream_t s);
  void forward_step(const Groups& Ff, const Wave& Wf, const Source& Sf, int it, bool record, hipStream_t s,
                    long save = -1);
  int ensure_alt();
  void swap_state();
  void fused_forward_step(const Source& Sf, int it, bool record, hipStream_t s, long save = -1);
  int set_model(const float* l, const float* m, const float* rho, int on_host, hipStream_t s);
  int zero_gradient(hipStream_t s);
  int prepare_shot(const efwi_shot* sh, int calc_id);
  int run_shot(const efwi_shot* sh, int calc_id, const float* obs, float* syn, float* res,
               double* misfit, hipStream_t s);
  int get_gradient(float* gl, float* gm, float* gr, int on_host, hipStream_t s);
  int autotune(int reps, hipStream_t s, double* report);
  // batched shots (efwi_batch.cpp)
  BatchState batch;
  static constexpr int kBatchFill = 16;  // automatic groups: workgroup tiles per CU and launch
  int batch_group_size(int n) const;
  void bind_state(const Groups& g);
  int ensure_batch(int n, int nring);
  int run_batch(const efwi_shot* shots, int n, int calc_id, const float* const* obs, float* const* syn,
                double* misfits, int use_graph, hipStream_t s);
  int batch_gradient(int shot, float* gl, float* gm, float* gr, int on_host, hipStream_t s);
  void release_batch();
};

int grow(Ctx* c, void** p, size_t* cap, size_t need);
void wave_of(float* base, size_t fsf, Wave& W);
void psi_of(float* base, size_t fsf, Psi& P);
void fields9(const Wave& W, Fields9& F);
void zero_wave(const Wave& W, long n, hipStream_t s);
void zero_psi(const Psi& P, long n, hipStream_t s);

}  // namespace efwi
