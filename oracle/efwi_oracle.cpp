// CPU oracle for the 3-D elastic FWI propagation path -- TEST INFRASTRUCTURE ONLY.
//
// This file restates, cell for cell and in the same step order, the arithmetic
// of the reference implementation (guoketing/EFWI-3D, Ops/FWI/Src_3d). It is
// the checker the GPU library is compared against; nothing in the product
// links, loads or calls it. Only tests/, __graft_entry__.smoke() and the
// cpu_baseline leg of bench.py use it.
//
// Layout: torch layout [nz][nx][ny] (y fastest), index (z*nx + x)*ny + y.
// The reference transposes to [ny][nx][nz] on the host (Src/libCUFD.cu:72-80,
// 644-652); the per-cell arithmetic does not depend on the memory layout, so
// the oracle keeps the caller's layout.
//
// Precisions ("precision" argument of the C entry points):
//   0  fp32-pure    all arithmetic in float (what the GPU kernels compute);
//   1  fp32-literal float storage, but the reference's `2.0*` double literals
//                   promote the same sub-expressions to double as in the
//                   CUDA source (Src/stress.cu:57-59,112-114,132-134,
//                   Src/stress_adj.cu:131-140, Src/velocity_adj.cu:218-278);
//   2  fp64         the same stencil evaluated in double (inputs, CPML
//                   coefficients, taper tables and the source amplitude are the
//                   same float values promoted exactly).
// Build with -ffp-contract=off so that no multiply-add is fused.
//
// Formulations ("formulation" argument):
//   0  reference    every quirk of the reference kept (dsxy/dx typo
//                   Src/velocity.cu:79; off-by-one x/y CPML predicate in the
//                   velocity kernel Src/velocity.cu:46-90; swapped adjoint
//                   CPML coefficients Src/velocity_adj.cu:139,142);
//   1  consistent   the three quirks corrected (used by the finite-difference
//                   gradient checks, where the adjoint must match the forward).
//
// Each function cites the reference lines it follows.

#include <cmath>
#include <cstdint>
#include <cstdio>
#include <cstring>
#include <algorithm>
#include <chrono>
#include <memory>
#include <vector>
#include <omp.h>

namespace {

constexpr double kPI = 3.141592653589793238462643383279502884197169;  // Src/utilities.h:15

enum Field {
  VZ, VX, VY, SZZ, SXX, SXZ, SYY, SXY, SYZ,                 // forward wavefield
  AVZ, AVX, AVY, ASZZ, ASXX, ASXZ, ASYY, ASXY, ASYZ,        // adjoint wavefield
  M_DVZ_DZ, M_DVZ_DX, M_DVZ_DY, M_DVX_DZ, M_DVX_DX, M_DVX_DY,
  M_DVY_DZ, M_DVY_DX, M_DVY_DY,                             // psi of the stress kernel
  M_DSZZ_DZ, M_DSXZ_DX, M_DSXZ_DZ, M_DSXX_DX, M_DSXY_DX, M_DSXY_DY,
  M_DSYY_DY, M_DSYZ_DY, M_DSYZ_DZ,                          // psi of the velocity kernel
  LAM, MU, BYC, GLAM, GMU, GRHO,
  NFIELD
};

struct Grid {
  int nz, nx, ny, npml, nt;
  float dz, dx, dy, dt, f0;
};

// ---------------------------------------------------------------------------
// C-PML 1-D profile, Src/utilities.cu:204-320 (called per axis by
// Src/Cpml.cu:33-43 with that axis' grid step). Variables that the reference
// holds in float are rounded to float at the same points; the libm calls are
// evaluated in double (pow/log/exp of the C library) except expf/powf where
// the reference names the float function.
// ---------------------------------------------------------------------------
struct Profile {
  std::vector<float> K, a, b, Kh, ah, bh;
};

void cpml_profile(int N, int npml, float dh, float f0, float dt, Profile& p) {
  p.K.assign(N, 1.0f); p.Kh.assign(N, 1.0f);
  p.a.assign(N, 0.0f); p.ah.assign(N, 0.0f);
  p.b.assign(N, 0.0f); p.bh.assign(N, 0.0f);
  std::vector<float> damp(N, 0.0f), damp_h(N, 0.0f), alpha(N, 0.0f), alpha_h(N, 0.0f);
  const float Rcoef = 0.0008f;
  const float K_MAX = 2.0f;
  const float ALPHA_MAX = (float)(2.0 * kPI * (f0 / 2.0));
  const float NPOWER = 8.0f;
  const float c1 = 0.25f, c2 = 0.75f, c3 = 0.0f;
  const float thick = npml * dh;
  const float VpRef = 3000.0f;  // Src/utilities.cu:221 (model independent)
  const float d0 = (float)(-(NPOWER + 1) * VpRef * std::log((double)Rcoef) / (2.0 * thick));
  auto damp_of = [&](float q) {
    return (float)(d0 * (c1 * (double)q + c2 * std::pow((double)q, (double)NPOWER) +
                         c3 * std::pow((double)q, 2.0 * NPOWER)));
  };
  for (int i = 0; i < N; i++) {
    float depth = (float)(npml - i) * dh;                      // left, integer grid
    if (depth >= 0.0f) {
      float q = depth / thick;
      damp[i] = damp_of(q);
      p.K[i] = (float)(1.0 + (K_MAX - 1.0) * std::pow((double)q, (double)NPOWER));
      alpha[i] = (float)(ALPHA_MAX * (1.0 - q));
    }
    depth = (float)((npml - i - 0.5) * dh);                     // left, half grid
    if (depth >= 0.0f) {
      float q = depth / thick;
      damp_h[i] = damp_of(q);
      p.Kh[i] = (float)(1.0 + (K_MAX - 1.0) * std::pow((double)q, (double)NPOWER));
      alpha_h[i] = (float)(ALPHA_MAX * (1.0 - q));
    }
    depth = (float)(npml - N + i) * dh;                         // right, integer grid
    if (depth >= 0.0f) {
      float q = depth / thick;
      damp[i] = damp_of(q);
      p.K[i] = (float)(1.0 + (K_MAX - 1.0) * std::pow((double)q, (double)NPOWER));
      alpha[i] = (float)(ALPHA_MAX * (1.0 - q));
    }
    depth = (float)((npml - N + i + 0.5) * dh);                 // right, half grid
    if (depth >= 0.0f) {
      float q = depth / thick;
      damp_h[i] = damp_of(q);
      p.Kh[i] = (float)(1.0 + (K_MAX - 1.0) * powf(q, NPOWER));  // powf here, :290
      alpha_h[i] = (float)(ALPHA_MAX * (1.0 - q));
    }
    if (alpha[i] < 0.0f) alpha[i] = 0.0f;
    if (alpha_h[i] < 0.0f) alpha_h[i] = 0.0f;
    p.b[i] = expf(-(damp[i] / p.K[i] + alpha[i]) * dt);
    p.bh[i] = expf(-(damp_h[i] / p.Kh[i] + alpha_h[i]) * dt);
    if (std::fabs(damp[i]) > 1.0e-6)
      p.a[i] = (float)(damp[i] * (p.b[i] - 1.0) / (p.K[i] * (damp[i] + p.K[i] * alpha[i])));
    if (std::fabs(damp_h[i]) > 1.0e-6)
      p.ah[i] = (float)(damp_h[i] * (p.bh[i] - 1.0) /
                        (p.Kh[i] * (damp_h[i] + p.Kh[i] * alpha_h[i])));
  }
}

// Time taper of Src/utilities.cu:578-614 (plain) and :524-576 (per-trace
// window). Returns w(t) for sample it; the caller squares it in float.
float taper_plain(int it, int nt, float dt, float ratio, bool* valid) {
  float t = it * dt;
  float t0 = 0;
  float t3 = nt * dt;
  float offset = nt * dt * ratio;
  if (2.0 * offset >= t3 - t0) { *valid = false; return 1.0f; }
  *valid = true;
  float t1 = t0 + offset;
  float t2 = t3 - offset;
  if (t >= t0 && t < t1) return (float)std::sin(kPI / 2.0 * (t - t0) / (t1 - t0));
  if (t >= t1 && t < t2) return 1.0f;
  if (t >= t2 && t < t3) return (float)std::cos(kPI / 2.0 * (t - t2) / (t3 - t2));
  return 0.0f;
}

float taper_window(int it, int nt, float dt, float ratio, float t0, float t3, bool* valid) {
  float t = it * dt;
  if (ratio > 0.5) { *valid = false; return 1.0f; }
  float t_max = nt * dt;
  if (t0 < 0.0f) t0 = 0.0f;
  if (t0 > t_max) t0 = t_max;
  if (t3 < 0.0f) t3 = 0.0f;
  if (t3 > t_max) t3 = t_max;
  float offset = (t3 - t0) * ratio;
  if (offset <= 0.0) { *valid = false; return 1.0f; }
  *valid = true;
  float t1 = t0 + offset;
  float t2 = t3 - offset;
  if (t >= t0 && t < t1) return (float)std::sin(kPI / 2.0 * (t - t0) / (t1 - t0));
  if (t >= t1 && t < t2) return 1.0f;
  if (t >= t2 && t < t3) return (float)std::cos(kPI / 2.0 * (t - t2) / (t3 - t2));
  return 0.0f;
}

// The reference's misfit reduction, Src/utilities.cu:145-166: one block of
// 512 threads, strided float partial sums, then a shared-memory tree.
float fp32_block_sumsq(const float* e, long n) {
  const int B = 512;
  float s[B];
  for (int t = 0; t < B; t++) s[t] = 0.0f;
  long nb = (n + B - 1) / B;
  for (long k = 0; k < nb; k++)
    for (int t = 0; t < B; t++) {
      long id = k * B + t;
      float a = id < n ? e[id] : 0.0f;
      s[t] += a * a;
    }
  for (int h = B / 2; h >= 1; h /= 2)
    for (int t = 0; t < h; t++) s[t] += s[t + h];
  return s[0];
}

// ---------------------------------------------------------------------------
// The engine. R is the storage/arithmetic type, W the type of the reference's
// `2.0` literals (double in fp32-literal mode, R otherwise).
// ---------------------------------------------------------------------------
struct EngineBase {
  Grid g{};
  int formulation = 0;
  long N = 0, SZ = 0, SX = 0;
  virtual ~EngineBase() {}
  virtual void set(int f, const double* d) = 0;
  virtual void get(int f, double* d) const = 0;
  virtual void zero(int f) = 0;
  virtual void kernel(int k) = 0;
  virtual void add_source(int z, int x, int y, float amp, int sign) = 0;
  virtual void record(const int* rec, long nrec, int it, double* out_x, double* out_y,
                      double* out_z) = 0;
  virtual void set_cpml_off() = 0;
  virtual void get_cpml(int axis, double* out) const = 0;
  virtual void ring_init(const std::vector<long>& band, int nslots) = 0;
  virtual void ring_save(int slot) = 0;
  virtual void ring_restore(int slot, int k0, int k1) = 0;
  virtual void inject(const int* rec, long nrec, int it, const double* rx) = 0;
  // checkpoint mode: the forward state (nine wavefield fields and the 18
  // memory arrays) to and from raw storage, and the band of this engine's
  // state into `dst`'s ring (an engine of the same precision)
  virtual size_t state_bytes() const = 0;
  virtual void save_state(char* out) const = 0;
  virtual void load_state(const char* in) = 0;
  virtual void ring_save_into(EngineBase* dst, int slot) const = 0;
};

// Order of the nine fields in a ring slot (the order of Src/Boundary.cu:59-96).
const int kRingField[9] = {SZZ, SXZ, SXX, SYY, SXY, SYZ, VZ, VX, VY};

template <class R, class W>
struct Engine : EngineBase {
  std::vector<R> F[NFIELD];
  std::vector<R> K[3], A[3], B[3], Kh[3], Ah[3], Bh[3];  // axis 0=z,1=x,2=y
  R c1, c2, dt, hz, hx, hy;
  W two;

  Engine(const Grid& gr, int form) {
    g = gr;
    formulation = form;
    N = (long)g.nz * g.nx * g.ny;
    SZ = (long)g.nx * g.ny;
    SX = g.ny;
    for (int f = 0; f < NFIELD; f++) F[f].assign(N, R(0));
    // Src/stress.cu:33-34: float c1 = 9.0/8.0; float c2 = 1.0/24.0;
    c1 = (R)(9.0 / 8.0);
    c2 = (R)(1.0 / 24.0);
    dt = (R)g.dt; hz = (R)g.dz; hx = (R)g.dx; hy = (R)g.dy;
    two = (W)2.0;
    const int n[3] = {g.nz, g.nx, g.ny};
    const float h[3] = {g.dz, g.dx, g.dy};
    for (int ax = 0; ax < 3; ax++) {
      Profile p;
      cpml_profile(n[ax], g.npml, h[ax], g.f0, g.dt, p);
      K[ax].assign(p.K.begin(), p.K.end());
      A[ax].assign(p.a.begin(), p.a.end());
      B[ax].assign(p.b.begin(), p.b.end());
      Kh[ax].assign(p.Kh.begin(), p.Kh.end());
      Ah[ax].assign(p.ah.begin(), p.ah.end());
      Bh[ax].assign(p.bh.begin(), p.bh.end());
    }
  }

  void set_cpml_off() override {
    for (int ax = 0; ax < 3; ax++) {
      std::fill(K[ax].begin(), K[ax].end(), R(1)); std::fill(Kh[ax].begin(), Kh[ax].end(), R(1));
      std::fill(A[ax].begin(), A[ax].end(), R(0)); std::fill(Ah[ax].begin(), Ah[ax].end(), R(0));
      std::fill(B[ax].begin(), B[ax].end(), R(1)); std::fill(Bh[ax].begin(), Bh[ax].end(), R(1));
    }
  }
  void get_cpml(int ax, double* out) const override {
    const int n = ax == 0 ? g.nz : ax == 1 ? g.nx : g.ny;
    const std::vector<R>* v[6] = {&K[ax], &A[ax], &B[ax], &Kh[ax], &Ah[ax], &Bh[ax]};
    for (int k = 0; k < 6; k++)
      for (int i = 0; i < n; i++) out[(long)k * n + i] = (double)(*v[k])[i];
  }
  void set(int f, const double* d) override {
    for (long i = 0; i < N; i++) F[f][i] = (R)d[i];
  }
  void get(int f, double* d) const override {
    for (long i = 0; i < N; i++) d[i] = (double)F[f][i];
  }
  void zero(int f) override { std::fill(F[f].begin(), F[f].end(), R(0)); }

  // Difference operators as written in the reference (f[i+-k] along stride s).
  inline R Db(const R* f, long i, long s, R h) const {
    return (c1 * (f[i] - f[i - s]) - c2 * (f[i + s] - f[i - 2 * s])) / h;
  }
  inline R Df(const R* f, long i, long s, R h) const {
    return (c1 * (f[i + s] - f[i]) - c2 * (f[i + 2 * s] - f[i - s])) / h;
  }
  // dsxy/dx of the vy update, Src/velocity.cu:79 (the typo uses f[i-s] in
  // place of f[i+s]); the consistent formulation uses Db.
  inline R Dtypo(const R* f, long i, long s, R h) const {
    if (formulation == 1) return Db(f, i, s, h);
    return (c1 * (f[i] - f[i - s]) - c2 * (f[i - s] - f[i - 2 * s])) / h;
  }
  // Adjoint forms, Src/stress_adj.cu:54-59: (-c1*(..) + c2*(..))/h.
  inline R Af(const R* f, long i, long s, R h) const {
    return (-c1 * (f[i + s] - f[i]) + c2 * (f[i + 2 * s] - f[i - s])) / h;
  }
  inline R Ab(const R* f, long i, long s, R h) const {
    return (-c1 * (f[i] - f[i - s]) + c2 * (f[i + s] - f[i - 2 * s])) / h;
  }

  inline bool pml(int i, int n) const { return i < g.npml || i > n - g.npml - 1; }
  // x/y predicate of the forward velocity kernel, Src/velocity.cu:46,50,66,70,86,90.
  inline bool pml_v(int i, int n) const {
    if (formulation == 1) return pml(i, n);
    return i < g.npml || i > n - g.npml;
  }

  // Forward stress, Src/stress.cu:36-101.
  void stress_fwd() {
    R* vz = F[VZ].data(); R* vx = F[VX].data(); R* vy = F[VY].data();
    R* szz = F[SZZ].data(); R* sxx = F[SXX].data(); R* sxz = F[SXZ].data();
    R* syy = F[SYY].data(); R* sxy = F[SXY].data(); R* syz = F[SYZ].data();
    R* m_vz_z = F[M_DVZ_DZ].data(); R* m_vz_x = F[M_DVZ_DX].data(); R* m_vz_y = F[M_DVZ_DY].data();
    R* m_vx_z = F[M_DVX_DZ].data(); R* m_vx_x = F[M_DVX_DX].data(); R* m_vx_y = F[M_DVX_DY].data();
    R* m_vy_z = F[M_DVY_DZ].data(); R* m_vy_x = F[M_DVY_DX].data(); R* m_vy_y = F[M_DVY_DY].data();
    const R* L = F[LAM].data(); const R* M = F[MU].data();
    const int nz = g.nz, nx = g.nx, ny = g.ny;
#pragma omp parallel for schedule(static)
    for (int z = 2; z <= nz - 3; z++)
      for (int x = 2; x <= nx - 3; x++)
        for (int y = 2; y <= ny - 3; y++) {
          const long i = z * SZ + x * SX + y;
          R dvz_dz = Db(vz, i, SZ, hz);
          R dvx_dx = Db(vx, i, SX, hx);
          R dvy_dy = Db(vy, i, 1, hy);
          if (pml(z, nz)) { m_vz_z[i] = B[0][z] * m_vz_z[i] + A[0][z] * dvz_dz; dvz_dz = dvz_dz / K[0][z] + m_vz_z[i]; }
          if (pml(x, nx)) { m_vx_x[i] = B[1][x] * m_vx_x[i] + A[1][x] * dvx_dx; dvx_dx = dvx_dx / K[1][x] + m_vx_x[i]; }
          if (pml(y, ny)) { m_vy_y[i] = B[2][y] * m_vy_y[i] + A[2][y] * dvy_dy; dvy_dy = dvy_dy / K[2][y] + m_vy_y[i]; }
          const R lam = L[i], mu = M[i];
          szz[i] += ((lam + two * mu) * dvz_dz + lam * dvx_dx + lam * dvy_dy) * dt;
          sxx[i] += (lam * dvz_dz + (lam + two * mu) * dvx_dx + lam * dvy_dy) * dt;
          syy[i] += (lam * dvz_dz + lam * dvx_dx + (lam + two * mu) * dvy_dy) * dt;
          R dvx_dz = Df(vx, i, SZ, hz);
          R dvx_dy = Df(vx, i, 1, hy);
          R dvz_dx = Df(vz, i, SX, hx);
          R dvz_dy = Df(vz, i, 1, hy);
          R dvy_dx = Df(vy, i, SX, hx);
          R dvy_dz = Df(vy, i, SZ, hz);
          if (pml(z, nz)) {
            m_vx_z[i] = Bh[0][z] * m_vx_z[i] + Ah[0][z] * dvx_dz; dvx_dz = dvx_dz / Kh[0][z] + m_vx_z[i];
            m_vy_z[i] = Bh[0][z] * m_vy_z[i] + Ah[0][z] * dvy_dz; dvy_dz = dvy_dz / Kh[0][z] + m_vy_z[i];
          }
          if (pml(x, nx)) {
            m_vz_x[i] = Bh[1][x] * m_vz_x[i] + Ah[1][x] * dvz_dx; dvz_dx = dvz_dx / Kh[1][x] + m_vz_x[i];
            m_vy_x[i] = Bh[1][x] * m_vy_x[i] + Ah[1][x] * dvy_dx; dvy_dx = dvy_dx / Kh[1][x] + m_vy_x[i];
          }
          if (pml(y, ny)) {
            m_vz_y[i] = Bh[2][y] * m_vz_y[i] + Ah[2][y] * dvz_dy; dvz_dy = dvz_dy / Kh[2][y] + m_vz_y[i];
            m_vx_y[i] = Bh[2][y] * m_vx_y[i] + Ah[2][y] * dvx_dy; dvx_dy = dvx_dy / Kh[2][y] + m_vx_y[i];
          }
          sxz[i] = sxz[i] + mu * (dvx_dz + dvz_dx) * dt;
          sxy[i] = sxy[i] + mu * (dvx_dy + dvy_dx) * dt;
          syz[i] = syz[i] + mu * (dvy_dz + dvz_dy) * dt;
        }
  }

  // Forward velocity, Src/velocity.cu:175-240.
  void velocity_fwd() {
    R* vz = F[VZ].data(); R* vx = F[VX].data(); R* vy = F[VY].data();
    const R* szz = F[SZZ].data(); const R* sxx = F[SXX].data(); const R* sxz = F[SXZ].data();
    const R* syy = F[SYY].data(); const R* sxy = F[SXY].data(); const R* syz = F[SYZ].data();
    R* m_szz_z = F[M_DSZZ_DZ].data(); R* m_sxz_x = F[M_DSXZ_DX].data(); R* m_sxz_z = F[M_DSXZ_DZ].data();
    R* m_sxx_x = F[M_DSXX_DX].data(); R* m_sxy_x = F[M_DSXY_DX].data(); R* m_sxy_y = F[M_DSXY_DY].data();
    R* m_syy_y = F[M_DSYY_DY].data(); R* m_syz_y = F[M_DSYZ_DY].data(); R* m_syz_z = F[M_DSYZ_DZ].data();
    const R* Byc = F[BYC].data();
    const int nz = g.nz, nx = g.nx, ny = g.ny;
#pragma omp parallel for schedule(static)
    for (int z = 2; z <= nz - 3; z++)
      for (int x = 2; x <= nx - 3; x++)
        for (int y = 2; y <= ny - 3; y++) {
          const long i = z * SZ + x * SX + y;
          const R byc = Byc[i];
          // vz
          R dszz_dz = Df(szz, i, SZ, hz);
          R dsxz_dx = Db(sxz, i, SX, hx);
          R dsyz_dy = Db(syz, i, 1, hy);
          if (pml(z, nz)) { m_szz_z[i] = Bh[0][z] * m_szz_z[i] + Ah[0][z] * dszz_dz; dszz_dz = dszz_dz / Kh[0][z] + m_szz_z[i]; }
          if (pml_v(x, nx)) { m_sxz_x[i] = B[1][x] * m_sxz_x[i] + A[1][x] * dsxz_dx; dsxz_dx = dsxz_dx / K[1][x] + m_sxz_x[i]; }
          if (pml_v(y, ny)) { m_syz_y[i] = B[2][y] * m_syz_y[i] + A[2][y] * dsyz_dy; dsyz_dy = dsyz_dy / K[2][y] + m_syz_y[i]; }
          vz[i] += (dszz_dz + dsxz_dx + dsyz_dy) * byc * dt;
          // vx
          R dsxz_dz = Db(sxz, i, SZ, hz);
          R dsxx_dx = Df(sxx, i, SX, hx);
          R dsxy_dy = Db(sxy, i, 1, hy);
          if (pml(z, nz)) { m_sxz_z[i] = B[0][z] * m_sxz_z[i] + A[0][z] * dsxz_dz; dsxz_dz = dsxz_dz / K[0][z] + m_sxz_z[i]; }
          if (pml_v(x, nx)) { m_sxx_x[i] = Bh[1][x] * m_sxx_x[i] + Ah[1][x] * dsxx_dx; dsxx_dx = dsxx_dx / Kh[1][x] + m_sxx_x[i]; }
          if (pml_v(y, ny)) { m_sxy_y[i] = B[2][y] * m_sxy_y[i] + A[2][y] * dsxy_dy; dsxy_dy = dsxy_dy / K[2][y] + m_sxy_y[i]; }
          vx[i] += (dsxz_dz + dsxx_dx + dsxy_dy) * byc * dt;
          // vy
          R dsyz_dz = Db(syz, i, SZ, hz);
          R dsxy_dx = Dtypo(sxy, i, SX, hx);
          R dsyy_dy = Df(syy, i, 1, hy);
          if (pml(z, nz)) { m_syz_z[i] = B[0][z] * m_syz_z[i] + A[0][z] * dsyz_dz; dsyz_dz = dsyz_dz / K[0][z] + m_syz_z[i]; }
          if (pml_v(x, nx)) { m_sxy_x[i] = B[1][x] * m_sxy_x[i] + A[1][x] * dsxy_dx; dsxy_dx = dsxy_dx / K[1][x] + m_sxy_x[i]; }
          if (pml_v(y, ny)) { m_syy_y[i] = Bh[2][y] * m_syy_y[i] + Ah[2][y] * dsyy_dy; dsyy_dy = dsyy_dy / Kh[2][y] + m_syy_y[i]; }
          vy[i] += (dsyz_dz + dsxy_dx + dsyy_dy) * byc * dt;
        }
  }

  inline bool interior(int z, int x, int y) const {
    const int p = g.npml;
    return z >= p && z <= g.nz - 1 - p && x >= p && x <= g.nx - 1 - p && y >= p && y <= g.ny - 1 - p;
  }

  // Reverse stress with the lambda/mu imaging condition, Src/stress.cu:103-139.
  void stress_rev() {
    const R* vz = F[VZ].data(); const R* vx = F[VX].data(); const R* vy = F[VY].data();
    R* szz = F[SZZ].data(); R* sxx = F[SXX].data(); R* sxz = F[SXZ].data();
    R* syy = F[SYY].data(); R* sxy = F[SXY].data(); R* syz = F[SYZ].data();
    const R* aszz = F[ASZZ].data(); const R* asxx = F[ASXX].data(); const R* asxz = F[ASXZ].data();
    const R* asyy = F[ASYY].data(); const R* asxy = F[ASXY].data(); const R* asyz = F[ASYZ].data();
    R* gl = F[GLAM].data(); R* gm = F[GMU].data();
    const R* L = F[LAM].data(); const R* M = F[MU].data();
    const int p = g.npml;
#pragma omp parallel for schedule(static)
    for (int z = p; z <= g.nz - 1 - p; z++)
      for (int x = p; x <= g.nx - 1 - p; x++)
        for (int y = p; y <= g.ny - 1 - p; y++) {
          const long i = z * SZ + x * SX + y;
          const R lam = L[i], mu = M[i];
          R dvz_dz = Db(vz, i, SZ, hz);
          R dvx_dx = Db(vx, i, SX, hx);
          R dvy_dy = Db(vy, i, 1, hy);
          szz[i] -= ((lam + two * mu) * dvz_dz + lam * dvx_dx + lam * dvy_dy) * dt;
          sxx[i] -= (lam * dvz_dz + (lam + two * mu) * dvx_dx + lam * dvy_dy) * dt;
          syy[i] -= (lam * dvz_dz + lam * dvx_dx + (lam + two * mu) * dvy_dy) * dt;
          R dvx_dz = Df(vx, i, SZ, hz);
          R dvx_dy = Df(vx, i, 1, hy);
          R dvz_dx = Df(vz, i, SX, hx);
          R dvz_dy = Df(vz, i, 1, hy);
          R dvy_dx = Df(vy, i, SX, hx);
          R dvy_dz = Df(vy, i, SZ, hz);
          sxz[i] -= mu * (dvx_dz + dvz_dx) * dt;
          sxy[i] -= mu * (dvx_dy + dvy_dx) * dt;
          syz[i] -= mu * (dvy_dz + dvz_dy) * dt;
          gl[i] += -(aszz[i] + asxx[i] + asyy[i]) * (dvz_dz + dvx_dx + dvy_dy) * dt;
          gm[i] += -(two * aszz[i] * dvz_dz + two * asxx[i] * dvx_dx + two * asyy[i] * dvy_dy +
                     asxz[i] * (dvx_dz + dvz_dx) + asyz[i] * (dvy_dz + dvz_dy) +
                     asxy[i] * (dvy_dx + dvx_dy)) * dt;
        }
  }

  // Reverse velocity with the density imaging condition, Src/velocity.cu:242-275.
  void velocity_rev() {
    R* vz = F[VZ].data(); R* vx = F[VX].data(); R* vy = F[VY].data();
    const R* szz = F[SZZ].data(); const R* sxx = F[SXX].data(); const R* sxz = F[SXZ].data();
    const R* syy = F[SYY].data(); const R* sxy = F[SXY].data(); const R* syz = F[SYZ].data();
    const R* avz = F[AVZ].data(); const R* avx = F[AVX].data(); const R* avy = F[AVY].data();
    R* gr = F[GRHO].data();
    const R* Byc = F[BYC].data();
    const int p = g.npml;
#pragma omp parallel for schedule(static)
    for (int z = p; z <= g.nz - 1 - p; z++)
      for (int x = p; x <= g.nx - 1 - p; x++)
        for (int y = p; y <= g.ny - 1 - p; y++) {
          const long i = z * SZ + x * SX + y;
          const R byc = Byc[i];
          R dszz_dz = Df(szz, i, SZ, hz);
          R dsxz_dx = Db(sxz, i, SX, hx);
          R dsyz_dy = Db(syz, i, 1, hy);
          vz[i] -= (dszz_dz + dsxz_dx + dsyz_dy) * byc * dt;
          R dsxz_dz = Db(sxz, i, SZ, hz);
          R dsxx_dx = Df(sxx, i, SX, hx);
          R dsxy_dy = Db(sxy, i, 1, hy);
          vx[i] -= (dsxz_dz + dsxx_dx + dsxy_dy) * byc * dt;
          R dsyz_dz = Db(syz, i, SZ, hz);
          R dsxy_dx = Dtypo(sxy, i, SX, hx);
          R dsyy_dy = Df(syy, i, 1, hy);
          vy[i] -= (dsyz_dz + dsxy_dx + dsyy_dy) * byc * dt;
          const R b2 = byc * byc;  // pow(d_Byc[id], 2)
          gr[i] += -(avz[i] * (dszz_dz + dsxz_dx + dsyz_dy) * dt * (-b2) +
                     avx[i] * (dsxz_dz + dsxx_dx + dsxy_dy) * dt * (-b2) +
                     avy[i] * (dsyz_dz + dsxy_dx + dsyy_dy) * dt * (-b2));
        }
  }

  // Adjoint velocity, Src/velocity_adj.cu:149-296.
  void velocity_adj() {
    R* vz = F[AVZ].data(); R* vx = F[AVX].data(); R* vy = F[AVY].data();
    const R* szz = F[ASZZ].data(); const R* sxx = F[ASXX].data(); const R* sxz = F[ASXZ].data();
    const R* syy = F[ASYY].data(); const R* sxy = F[ASXY].data(); const R* syz = F[ASYZ].data();
    const R* p_vz_z = F[M_DVZ_DZ].data(); const R* p_vz_x = F[M_DVZ_DX].data(); const R* p_vz_y = F[M_DVZ_DY].data();
    const R* p_vx_z = F[M_DVX_DZ].data(); const R* p_vx_x = F[M_DVX_DX].data(); const R* p_vx_y = F[M_DVX_DY].data();
    const R* p_vy_z = F[M_DVY_DZ].data(); const R* p_vy_x = F[M_DVY_DX].data(); const R* p_vy_y = F[M_DVY_DY].data();
    R* m_szz_z = F[M_DSZZ_DZ].data(); R* m_sxz_x = F[M_DSXZ_DX].data(); R* m_sxz_z = F[M_DSXZ_DZ].data();
    R* m_sxx_x = F[M_DSXX_DX].data(); R* m_sxy_x = F[M_DSXY_DX].data(); R* m_sxy_y = F[M_DSXY_DY].data();
    R* m_syy_y = F[M_DSYY_DY].data(); R* m_syz_y = F[M_DSYZ_DY].data(); R* m_syz_z = F[M_DSYZ_DZ].data();
    const R* L = F[LAM].data(); const R* M = F[MU].data(); const R* Byc = F[BYC].data();
    const int nz = g.nz, nx = g.nx, ny = g.ny;
    const bool cons = formulation == 1;
#pragma omp parallel for schedule(static)
    for (int z = 2; z <= nz - 3; z++)
      for (int x = 2; x <= nx - 3; x++)
        for (int y = 2; y <= ny - 3; y++) {
          const long i = z * SZ + x * SX + y;
          const R lambda = L[i], mu = M[i], byc = Byc[i];
          // vx
          {
            R dpsixx_dx = Af(p_vx_x, i, SX, hx);
            R dszz_dx = Af(szz, i, SX, hx);
            R dsxx_dx = Af(sxx, i, SX, hx);
            R dsyy_dx = Af(syy, i, SX, hx);
            R dpsixz_dz = Ab(p_vx_z, i, SZ, hz);
            R dsxz_dz = Ab(sxz, i, SZ, hz);
            R dpsixy_dy = Ab(p_vx_y, i, 1, hy);
            R dsxy_dy = Ab(sxy, i, 1, hy);
            vx[i] += (A[1][x] * dpsixx_dx + (lambda + two * mu) * dsxx_dx / K[1][x] * dt +
                      lambda * dszz_dx / K[1][x] * dt + lambda * dsyy_dx / K[1][x] * dt +
                      mu / Kh[0][z] * dsxz_dz * dt + Ah[0][z] * dpsixz_dz +
                      mu / Kh[2][y] * dsxy_dy * dt + Ah[2][y] * dpsixy_dy);
            if (pml(x, nx)) m_sxx_x[i] = Bh[1][x] * m_sxx_x[i] + byc * vx[i] * dt;
            if (pml(z, nz)) m_sxz_z[i] = B[0][z] * m_sxz_z[i] + byc * vx[i] * dt;
            if (pml(y, ny)) m_sxy_y[i] = B[2][y] * m_sxy_y[i] + byc * vx[i] * dt;
          }
          // vz
          {
            R dpsizz_dz = Af(p_vz_z, i, SZ, hz);
            R dszz_dz = Af(szz, i, SZ, hz);
            R dsxx_dz = Af(sxx, i, SZ, hz);
            R dsyy_dz = Af(syy, i, SZ, hz);
            R dpsizx_dx = Ab(p_vz_x, i, SX, hx);
            R dsxz_dx = Ab(sxz, i, SX, hx);
            R dpsizy_dy = Ab(p_vz_y, i, 1, hy);
            R dsyz_dy = Ab(syz, i, 1, hy);
            vz[i] += (A[0][z] * dpsizz_dz + (lambda + two * mu) * dszz_dz / K[0][z] * dt +
                      lambda * dsxx_dz / K[0][z] * dt + lambda * dsyy_dz / K[0][z] * dt +
                      mu / Kh[2][y] * dsyz_dy * dt + Ah[2][y] * dpsizy_dy +
                      mu / Kh[1][x] * dsxz_dx * dt + Ah[1][x] * dpsizx_dx);
            if (pml(x, nx)) m_sxz_x[i] = B[1][x] * m_sxz_x[i] + byc * vz[i] * dt;
            if (pml(z, nz)) m_szz_z[i] = Bh[0][z] * m_szz_z[i] + byc * vz[i] * dt;
            if (pml(y, ny)) m_syz_y[i] = B[2][y] * m_syz_y[i] + byc * vz[i] * dt;
          }
          // vy
          {
            R dpsiyy_dy = Af(p_vy_y, i, 1, hy);
            R dsyy_dy = Af(syy, i, 1, hy);
            R dsxx_dy = Af(sxx, i, 1, hy);
            R dszz_dy = Af(szz, i, 1, hy);
            R dpsiyx_dx = Ab(p_vy_x, i, SX, hx);
            R dsxy_dx = Ab(sxy, i, SX, hx);
            R dpsiyz_dz = Ab(p_vy_z, i, SZ, hz);
            R dsyz_dz = Ab(syz, i, SZ, hz);
            vy[i] += (A[2][y] * dpsiyy_dy + (lambda + two * mu) * dsyy_dy / K[2][y] * dt +
                      lambda * dsxx_dy / K[2][y] * dt + lambda * dszz_dy / K[2][y] * dt +
                      mu / Kh[1][x] * dsxy_dx * dt + Ah[1][x] * dpsiyx_dx +
                      mu / Kh[0][z] * dsyz_dz * dt + Ah[0][z] * dpsiyz_dz);
            if (pml(x, nx)) m_sxy_x[i] = B[1][x] * m_sxy_x[i] + byc * vy[i] * dt;
            // Src/velocity_adj.cu:285 uses b_z_half (the forward uses b_z).
            if (pml(z, nz)) m_syz_z[i] = (cons ? B[0][z] : Bh[0][z]) * m_syz_z[i] + byc * vy[i] * dt;
            // Src/velocity_adj.cu:288 uses b_y (the forward uses b_y_half).
            if (pml(y, ny)) m_syy_y[i] = (cons ? Bh[2][y] : B[2][y]) * m_syy_y[i] + byc * vy[i] * dt;
          }
        }
  }

  // Adjoint stress, Src/stress_adj.cu:1-146.
  void stress_adj() {
    const R* vz = F[AVZ].data(); const R* vx = F[AVX].data(); const R* vy = F[AVY].data();
    R* szz = F[ASZZ].data(); R* sxx = F[ASXX].data(); R* sxz = F[ASXZ].data();
    R* syy = F[ASYY].data(); R* sxy = F[ASXY].data(); R* syz = F[ASYZ].data();
    const R* q_szz_z = F[M_DSZZ_DZ].data(); const R* q_sxz_x = F[M_DSXZ_DX].data(); const R* q_sxz_z = F[M_DSXZ_DZ].data();
    const R* q_sxx_x = F[M_DSXX_DX].data(); const R* q_sxy_x = F[M_DSXY_DX].data(); const R* q_sxy_y = F[M_DSXY_DY].data();
    const R* q_syy_y = F[M_DSYY_DY].data(); const R* q_syz_y = F[M_DSYZ_DY].data(); const R* q_syz_z = F[M_DSYZ_DZ].data();
    R* m_vz_z = F[M_DVZ_DZ].data(); R* m_vz_x = F[M_DVZ_DX].data(); R* m_vz_y = F[M_DVZ_DY].data();
    R* m_vx_z = F[M_DVX_DZ].data(); R* m_vx_x = F[M_DVX_DX].data(); R* m_vx_y = F[M_DVX_DY].data();
    R* m_vy_z = F[M_DVY_DZ].data(); R* m_vy_x = F[M_DVY_DX].data(); R* m_vy_y = F[M_DVY_DY].data();
    const R* L = F[LAM].data(); const R* M = F[MU].data(); const R* Byc = F[BYC].data();
    const int nz = g.nz, nx = g.nx, ny = g.ny;
#pragma omp parallel for schedule(static)
    for (int z = 2; z <= nz - 3; z++)
      for (int x = 2; x <= nx - 3; x++)
        for (int y = 2; y <= ny - 3; y++) {
          const long i = z * SZ + x * SX + y;
          const R lambda = L[i], mu = M[i], byc = Byc[i];
          // sxz
          R dphi_xz_x_dx = Af(q_sxz_x, i, SX, hx);
          R dvz_dx = Af(vz, i, SX, hx);
          R dphi_xz_z_dz = Af(q_sxz_z, i, SZ, hz);
          R dvx_dz = Af(vx, i, SZ, hz);
          sxz[i] += A[1][x] * dphi_xz_x_dx + dvz_dx / K[1][x] * byc * dt +
                    A[0][z] * dphi_xz_z_dz + dvx_dz / K[0][z] * byc * dt;
          if (pml(x, nx)) m_vz_x[i] = Bh[1][x] * m_vz_x[i] + sxz[i] * mu * dt;
          if (pml(z, nz)) m_vx_z[i] = Bh[0][z] * m_vx_z[i] + sxz[i] * mu * dt;
          // sxy
          R dphi_xy_x_dx = Af(q_sxy_x, i, SX, hx);
          R dvy_dx = Af(vy, i, SX, hx);
          R dphi_xy_y_dy = Af(q_sxy_y, i, 1, hy);
          R dvx_dy = Af(vx, i, 1, hy);
          sxy[i] += A[1][x] * dphi_xy_x_dx + dvy_dx / K[1][x] * byc * dt +
                    A[2][y] * dphi_xy_y_dy + dvx_dy / K[2][y] * byc * dt;
          if (pml(y, ny)) m_vx_y[i] = Bh[2][y] * m_vx_y[i] + sxy[i] * mu * dt;
          if (pml(x, nx)) m_vy_x[i] = Bh[1][x] * m_vy_x[i] + sxy[i] * mu * dt;
          // syz
          R dphi_yz_y_dy = Af(q_syz_y, i, 1, hy);
          R dvz_dy = Af(vz, i, 1, hy);
          R dphi_yz_z_dz = Af(q_syz_z, i, SZ, hz);
          R dvy_dz = Af(vy, i, SZ, hz);
          syz[i] += A[2][y] * dphi_yz_y_dy + dvz_dy / K[2][y] * byc * dt +
                    A[0][z] * dphi_yz_z_dz + dvy_dz / K[0][z] * byc * dt;
          if (pml(y, ny)) m_vz_y[i] = Bh[2][y] * m_vz_y[i] + syz[i] * mu * dt;
          if (pml(z, nz)) m_vy_z[i] = Bh[0][z] * m_vy_z[i] + syz[i] * mu * dt;
          // normal stresses
          R dphi_xx_x_dx = Ab(q_sxx_x, i, SX, hx);
          R dvx_dx = Ab(vx, i, SX, hx);
          R dphi_zz_z_dz = Ab(q_szz_z, i, SZ, hz);
          R dvz_dz = Ab(vz, i, SZ, hz);
          R dphi_yy_y_dy = Ab(q_syy_y, i, 1, hy);
          R dvy_dy = Ab(vy, i, 1, hy);
          sxx[i] += Ah[1][x] * dphi_xx_x_dx + byc * dvx_dx / Kh[1][x] * dt;
          szz[i] += Ah[0][z] * dphi_zz_z_dz + byc * dvz_dz / Kh[0][z] * dt;
          syy[i] += Ah[2][y] * dphi_yy_y_dy + byc * dvy_dy / Kh[2][y] * dt;
          if (pml(x, nx))
            m_vx_x[i] = B[1][x] * m_vx_x[i] + lambda * szz[i] * dt +
                        (lambda + two * mu) * sxx[i] * dt + lambda * syy[i] * dt;
          if (pml(z, nz))
            m_vz_z[i] = B[0][z] * m_vz_z[i] + (lambda + two * mu) * szz[i] * dt +
                        lambda * sxx[i] * dt + lambda * syy[i] * dt;
          if (pml(y, ny))
            m_vy_y[i] = B[2][y] * m_vy_y[i] + lambda * szz[i] * dt +
                        lambda * sxx[i] * dt + (lambda + two * mu) * syy[i] * dt;
        }
  }

  void kernel(int k) override {
    switch (k) {
      case 0: stress_fwd(); break;
      case 1: velocity_fwd(); break;
      case 2: stress_rev(); break;
      case 3: velocity_rev(); break;
      case 4: velocity_adj(); break;
      case 5: stress_adj(); break;
      default: break;
    }
  }

  // Src/utilities.cu:442-479. The 9^3 amplitude cube is 1 at its centre and
  // expf(-1000 r^2) = 0 elsewhere (Src/utilities.cu:432), so only the centre
  // cell changes.
  void add_source(int z, int x, int y, float amp, int sign) override {
    const float scale = 2250000.0f;  // pow(1500.0, 2)
    const long i = z * SZ + x * SX + y;
    R a = (R)scale * (R)amp * dt * (R)1.0f;
    if (sign > 0) { F[SZZ][i] += a; F[SXX][i] += a; F[SYY][i] += a; }
    else { F[SZZ][i] -= a; F[SXX][i] -= a; F[SYY][i] -= a; }
  }

  std::vector<long> band_;
  std::vector<R> ring_;
  void ring_init(const std::vector<long>& band, int nslots) override {
    band_ = band;
    ring_.assign(band_.size() * 9 * (size_t)nslots, R(0));
  }
  // from_bnd for the nine fields (Src/Boundary.cu:58-95).
  void ring_save(int slot) override {
    const size_t nb = band_.size();
    for (int k = 0; k < 9; k++) {
      const R* f = F[kRingField[k]].data();
      R* dst = ring_.data() + ((size_t)slot * 9 + k) * nb;
      for (size_t b = 0; b < nb; b++) dst[b] = f[band_[b]];
    }
  }
  static bool state_field(int f) { return f <= SYZ || (f >= M_DVZ_DZ && f <= M_DSYZ_DZ); }
  size_t state_bytes() const override { return (size_t)27 * N * sizeof(R); }
  void save_state(char* out) const override {
    for (int f = 0; f < NFIELD; f++)
      if (state_field(f)) {
        std::memcpy(out, F[f].data(), N * sizeof(R));
        out += N * sizeof(R);
      }
  }
  void load_state(const char* in) override {
    for (int f = 0; f < NFIELD; f++)
      if (state_field(f)) {
        std::memcpy(F[f].data(), in, N * sizeof(R));
        in += N * sizeof(R);
      }
  }
  void ring_save_into(EngineBase* dst, int slot) const override {
    auto* d = static_cast<Engine*>(dst);
    const size_t nb = d->band_.size();
    for (int k = 0; k < 9; k++) {
      const R* f = F[kRingField[k]].data();
      R* out = d->ring_.data() + ((size_t)slot * 9 + k) * nb;
      for (size_t b = 0; b < nb; b++) out[b] = f[d->band_[b]];
    }
  }
  // to_bnd (Src/Boundary.cu:97-138): k0..k1 selects the stress (0..6) or
  // velocity (6..9) part.
  void ring_restore(int slot, int k0, int k1) override {
    const size_t nb = band_.size();
    for (int k = k0; k < k1; k++) {
      R* f = F[kRingField[k]].data();
      const R* src = ring_.data() + ((size_t)slot * 9 + k) * nb;
      for (size_t b = 0; b < nb; b++) f[band_[b]] = src[b];
    }
  }
  // res_injection, Src/utilities.cu:497-507 (x, y, z components of rx at
  // [3][nrec][nt]); receivers sharing a cell add sequentially.
  void inject(const int* rec, long nrec, int it, const double* rx) override {
    const int nt = g.nt;
    const int comps[3] = {AVX, AVY, AVZ};
    for (int c = 0; c < 3; c++) {
      R* f = F[comps[c]].data();
      for (long r = 0; r < nrec; r++) {
        const long i = rec[3 * r] * SZ + rec[3 * r + 1] * SX + rec[3 * r + 2];
        f[i] += (R)rx[(long)c * nrec * nt + r * nt + it];
      }
    }
  }

  // Src/utilities.cu:481-495 (written at column `it`).
  void record(const int* rec, long nrec, int it, double* ox, double* oy, double* oz) override {
    const int nt = g.nt;
    for (long r = 0; r < nrec; r++) {
      const long i = rec[3 * r] * SZ + rec[3 * r + 1] * SX + rec[3 * r + 2];
      ox[r * nt + it] = (double)F[VX][i];
      oy[r * nt + it] = (double)F[VY][i];
      oz[r * nt + it] = (double)F[VZ][i];
    }
  }
};

EngineBase* make_engine(const Grid& g, int precision, int formulation) {
  switch (precision) {
    case 0: return new Engine<float, float>(g, formulation);
    case 1: return new Engine<float, double>(g, formulation);
    case 2: return new Engine<double, double>(g, formulation);
    default: return nullptr;
  }
}

// The five-layer boundary band of Src/Boundary.cu:14-41 and
// Src/utilities.cu:323-372, as a cell mask: the union of the six faces that
// from_bnd stores. Faces overlap at edges and corners; the reference restores
// overlapping cells twice with the same value, so the union is what matters.
void band_mask(const Grid& g, std::vector<unsigned char>& mask) {
  const int p = g.npml;
  const int nzB = g.nz - 2 * p + 4, nxB = g.nx - 2 * p + 4, nyB = g.ny - 2 * p + 4;
  const long SZ = (long)g.nx * g.ny, SX = g.ny;
  mask.assign((long)g.nz * SZ, 0);
  auto mark = [&](int z, int x, int y) { mask[z * SZ + x * SX + y] = 1; };
  for (int k = 0; k < 5; k++)
    for (int j = 0; j < nxB; j++)
      for (int i = 0; i < nzB; i++) {
        mark(i + p - 2, j + p - 2, k + p - 2);                 // y low
        mark(i + p - 2, j + p - 2, g.ny - p - k - 1 + 2);      // y high
      }
  for (int i = 0; i < 5; i++)
    for (int k = 0; k < nyB; k++)
      for (int j = 0; j < nxB; j++) {
        mark(i + p - 2, j + p - 2, k + p - 2);                 // z low
        mark(g.nz - p - i - 1 + 2, j + p - 2, k + p - 2);      // z high
      }
  for (int j = 0; j < 5; j++)
    for (int i = 0; i < nzB; i++)
      for (int k = 0; k < nyB; k++) {
        mark(i + p - 2, j + p - 2, k + p - 2);                 // x low
        mark(i + p - 2, g.nx - p - j - 1 + 2, k + p - 2);      // x high
      }
}

}  // namespace

// ===========================================================================
// C entry points (ctypes).
// ===========================================================================
extern "C" {

struct orc_grid {
  int nz, nx, ny, npml, nt;
  float dz, dx, dy, dt, f0;
};

static Grid to_grid(const orc_grid* g) {
  Grid r;
  r.nz = g->nz; r.nx = g->nx; r.ny = g->ny; r.npml = g->npml; r.nt = g->nt;
  r.dz = g->dz; r.dx = g->dx; r.dy = g->dy; r.dt = g->dt; r.f0 = g->f0;
  return r;
}

void* orc_create(const orc_grid* g, int precision, int formulation) {
  return make_engine(to_grid(g), precision, formulation);
}
void orc_destroy(void* e) { delete static_cast<EngineBase*>(e); }
void orc_set(void* e, int f, const double* d) { static_cast<EngineBase*>(e)->set(f, d); }
void orc_get(void* e, int f, double* d) { static_cast<EngineBase*>(e)->get(f, d); }
void orc_kernel(void* e, int k) { static_cast<EngineBase*>(e)->kernel(k); }
void orc_cpml_off(void* e) { static_cast<EngineBase*>(e)->set_cpml_off(); }
void orc_get_cpml(void* e, int axis, double* out) { static_cast<EngineBase*>(e)->get_cpml(axis, out); }
void orc_add_source(void* e, int z, int x, int y, float amp, int sign) {
  static_cast<EngineBase*>(e)->add_source(z, x, y, amp, sign);
}
int orc_nfield() { return NFIELD; }

void orc_band_mask(const orc_grid* g, unsigned char* out) {
  std::vector<unsigned char> m;
  band_mask(to_grid(g), m);
  std::memcpy(out, m.data(), m.size());
}

// w(t)^2 tables of the two taper forms, float as in the reference.
void orc_taper_plain(int nt, float dt, float ratio, float* w2) {
  for (int it = 0; it < nt; it++) {
    bool ok;
    float w = taper_plain(it, nt, dt, ratio, &ok);
    w2[it] = ok ? w * w : 1.0f;
  }
}

float orc_fp32_block_sumsq(const float* e, long n) { return fp32_block_sumsq(e, n); }

void orc_cpml_profile(int N, int npml, float dh, float f0, float dt, float* out6) {
  Profile p;
  cpml_profile(N, npml, dh, f0, dt, p);
  const std::vector<float>* v[6] = {&p.K, &p.a, &p.b, &p.Kh, &p.ah, &p.bh};
  for (int k = 0; k < 6; k++) std::memcpy(out6 + (long)k * N, v[k]->data(), N * sizeof(float));
}

// Survey description: shots are processed in order; receiver coordinates and
// source coordinates are padded-grid (z, x, y) triples (the JSON value plus
// nPoints_pml, Src/Src_Rec.cu:78-118).
struct orc_survey {
  int nshots;
  const int* src;          // [nshots][3]
  const float* stf;        // [nshots][nt] rows (untapered), already selected by shot id
  const long* rec_off;     // [nshots+1]
  const int* rec;          // [rec_off[nshots]][3]
  int if_win;
  const float* win_start;  // [total rec] (if_win)
  const float* win_end;    // [total rec] (if_win)
  const float* weights;    // [total rec] (if_win)
  const float* src_weight; // [nshots]    (if_win)
};

// Full pipeline of cufd (Src/libCUFD.cu:37-799) for calc_id 0, 1, 2.
//   obs      [total rec][3][nt]? -- no: per shot [3][nrec][nt] (x, y, z), shots concatenated
//   syn      out, same layout (raw recorded data, before windowing)
//   res      out, same layout (windowed residual used as adjoint source)
//   misfit64 out, sum over shots and components of r^2 in double
//   misfit32 out, the reference's float accumulation of per-component fp32 block sums
//   glam/gmu/grho out [N] (calc_id 1)
//   seconds  out, wall time of the time loops
// Checkpoint mode of orc_run (0 = the full ring of all nt - 1 steps, the
// reference's scheme): with K > 0 the forward pass keeps the state every K
// steps, and the backward pass replays each segment from its checkpoint in a
// second engine to rebuild that segment's ring. The replayed band values are
// the forward pass's own values, so the results equal the full ring's bit for
// bit with K ring slots instead of nt - 1.
static int g_ckpt_every = 0;
void orc_set_checkpoint(int every) { g_ckpt_every = every > 0 ? every : 0; }

int orc_run(const orc_grid* gg, const orc_survey* sv, int precision, int formulation,
            int calc_id, int cpml_off, const float* lam, const float* mu, const float* rho,
            const float* obs, double* syn, double* res, double* misfit64, float* misfit32,
            double* glam, double* gmu, double* grho, double* seconds) {
  const Grid g = to_grid(gg);
  if (calc_id < 0 || calc_id > 2) return -1;
  const bool if_res = calc_id != 2, with_adj = calc_id == 1;
  std::unique_ptr<EngineBase> E(make_engine(g, precision, formulation));
  if (!E) return -2;
  if (cpml_off) E->set_cpml_off();
  const long N = E->N;
  const int nt = g.nt;
  {
    std::vector<double> tmp(N);
    for (long i = 0; i < N; i++) tmp[i] = lam[i];
    E->set(LAM, tmp.data());
    for (long i = 0; i < N; i++) tmp[i] = mu[i];
    E->set(MU, tmp.data());
    // Byc = 1/rho on [2, n-3]^3, 1/1000 elsewhere (Src/Model.cu:65, Src/utilities.cu:117-128).
    const long SZ = (long)g.nx * g.ny, SX = g.ny;
    for (int z = 0; z < g.nz; z++)
      for (int x = 0; x < g.nx; x++)
        for (int y = 0; y < g.ny; y++) {
          const long i = z * SZ + x * SX + y;
          bool in = z >= 2 && z <= g.nz - 3 && x >= 2 && x <= g.nx - 3 && y >= 2 && y <= g.ny - 3;
          tmp[i] = in ? (double)(float)(1.0 / (double)rho[i]) : (double)(float)(1.0 / 1000.0);
        }
    E->set(BYC, tmp.data());
  }
  // stf taper (Src/Src_Rec.cu:146) and misfit taper (Src/libCUFD.cu:64,400-405).
  std::vector<float> w2_src(nt), w2_dat(nt);
  orc_taper_plain(nt, g.dt, 0.001f, w2_src.data());
  orc_taper_plain(nt, g.dt, 0.005f, w2_dat.data());

  std::vector<long> band;
  if (with_adj) {
    std::vector<unsigned char> mask;
    band_mask(g, mask);
    for (long i = 0; i < N; i++) if (mask[i]) band.push_back(i);
  }
  const int K = with_adj && g_ckpt_every > 0 && g_ckpt_every < nt - 1 ? g_ckpt_every : 0;
  std::unique_ptr<EngineBase> E2;
  std::vector<std::vector<char>> ckpts;
  if (K) {
    E2.reset(make_engine(g, precision, formulation));
    if (cpml_off) E2->set_cpml_off();
    std::vector<double> tmp(N);
    for (int f : {LAM, MU, BYC}) {
      E->get(f, tmp.data());
      E2->set(f, tmp.data());
    }
  }
  float h_l2 = 0.0f;
  double j64 = 0.0;
  auto t_start = std::chrono::steady_clock::now();
  double t_loops = 0.0;

  for (int s = 0; s < sv->nshots; s++) {
    for (int f = VZ; f <= M_DSYZ_DZ; f++) E->zero(f);
    const long r0 = sv->rec_off[s], nrec = sv->rec_off[s + 1] - sv->rec_off[s];
    const int* rec = sv->rec + 3 * r0;
    double* sx = syn + 3 * r0 * nt;
    double* sy = sx + nrec * nt;
    double* szv = sy + nrec * nt;
    for (long k = 0; k < 3 * nrec * nt; k++) sx[k] = 0.0;
    std::vector<float> stf(nt);
    for (int it = 0; it < nt; it++) stf[it] = sv->stf[(long)s * nt + it] * w2_src[it];
    const int zs = sv->src[3 * s], xs = sv->src[3 * s + 1], ys = sv->src[3 * s + 2];

    if (with_adj) E->ring_init(band, K ? K : (nt > 1 ? nt - 1 : 1));
    const int nseg = K ? (nt - 1 + K - 1) / K : 1;
    if (K) ckpts.assign(nseg, std::vector<char>());

    auto t0 = std::chrono::steady_clock::now();
    for (int it = 0; it <= nt - 2; it++) {
      if (K && it % K == 0) {
        ckpts[it / K].resize(E->state_bytes());
        E->save_state(ckpts[it / K].data());
      }
      if (with_adj && !K) E->ring_save(it);
      E->kernel(0);
      E->add_source(zs, xs, ys, stf[it], +1);
      E->kernel(1);
      E->record(rec, nrec, it + 1, sx, sy, szv);
    }
    auto t1 = std::chrono::steady_clock::now();
    t_loops += std::chrono::duration<double>(t1 - t0).count();

    if (!if_res) continue;

    // Residual and misfit (Src/libCUFD.cu:371-441). The precision of the
    // windowing follows the engine precision.
    double* rx = res + 3 * r0 * nt;
    const float* ob = obs + 3 * r0 * nt;
    std::vector<float> w2(nrec * (long)nt);
    for (long r = 0; r < nrec; r++)
      for (int it = 0; it < nt; it++) {
        float f;
        if (sv->if_win) {
          bool ok;
          float w = taper_window(it, nt, g.dt, 0.005f, sv->win_start[r0 + r], sv->win_end[r0 + r], &ok);
          f = ok ? w * w * sv->weights[r0 + r] * sv->src_weight[s] : 1.0f;
        } else {
          f = w2_dat[it];
        }
        w2[r * nt + it] = f;
      }
    float comp32[3];
    for (int c = 0; c < 3; c++) {
      std::vector<float> e32(nrec * (long)nt);
      for (long r = 0; r < nrec; r++)
        for (int it = 0; it < nt; it++) {
          const long k = (long)c * nrec * nt + r * nt + it;
          const float f = w2[r * nt + it];
          double rv;
          if (precision == 2) {
            double o = (double)ob[k] * (double)f;
            double sv2 = sx[k] * (double)f;
            rv = it == 0 ? 0.0 : o - sv2;
            j64 += rv * rv;
            e32[r * nt + it] = (float)rv;
            rx[k] = rv * (double)f;
          } else {
            float o = ob[k] * f;
            float sv2 = (float)sx[k] * f;
            float rf = it == 0 ? 0.0f : o - sv2;
            j64 += (double)rf * (double)rf;
            e32[r * nt + it] = rf;
            rx[k] = (double)(rf * f);
          }
        }
      comp32[c] = fp32_block_sumsq(e32.data(), nrec * (long)nt);
    }
    h_l2 += comp32[0] + comp32[1] + comp32[2];

    if (!with_adj) continue;
    t0 = std::chrono::steady_clock::now();
    // Backward, Src/libCUFD.cu:469-613.
    for (int f = AVZ; f <= M_DSYZ_DZ; f++) E->zero(f);
    E->kernel(4);
    E->inject(rec, nrec, nt - 1, rx);
    E->kernel(5);
    for (int seg = nseg - 1; seg >= 0; seg--) {
      const int a = K ? seg * K : 0, b = K ? std::min(nt - 2, a + K - 1) : nt - 2;
      if (K) {
        // rebuild this segment's ring: slot it - a holds the state before step it
        E2->load_state(ckpts[seg].data());
        for (int it = a; it <= b; it++) {
          E2->ring_save_into(E.get(), it - a);
          E2->kernel(0);
          E2->add_source(zs, xs, ys, stf[it], +1);
          E2->kernel(1);
        }
      }
      for (int it = b; it >= a; it--) {
        E->kernel(3);
        E->ring_restore(it - a, 6, 9);
        E->add_source(zs, xs, ys, stf[it], -1);
        E->kernel(2);
        E->ring_restore(it - a, 0, 6);
        E->kernel(4);
        E->inject(rec, nrec, it, rx);
        E->kernel(5);
      }
    }
    t1 = std::chrono::steady_clock::now();
    t_loops += std::chrono::duration<double>(t1 - t0).count();
  }
  (void)t_start;
  if (misfit64) *misfit64 = j64;
  if (misfit32) *misfit32 = h_l2;
  if (with_adj) {
    E->get(GLAM, glam);
    E->get(GMU, gmu);
    E->get(GRHO, grho);
  }
  if (seconds) *seconds = t_loops;
  return 0;
}

}  // extern "C"
