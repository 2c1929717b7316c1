// Staggered-grid elastic stencil kernels for gfx950: forward, reverse
// reconstruction with imaging condition, and adjoint.
//
// Arithmetic: every cell evaluates the same expression tree as the reference
// kernel it replaces, with IEEE division and no fused multiply-add (this file
// is compiled with -ffp-contract=off), so the float results are the correctly
// rounded results of the reference's float expressions. The only liberty
// taken is to skip a C-PML term a*d(psi) where the profile coefficient a is
// exactly zero (the reference multiplies by zero there), which can change the
// sign of an exact zero and nothing else.
//
// Work distribution ("sweep"): a 1-D grid of one-wavefront workgroups, each a
// 64-cell segment of one y-row. Workgroup b runs on XCD b % 8 (round-robin
// dispatch); the kernels map b so that each XCD owns a contiguous strip of x
// rows and walks it plane by plane in z. The x and z neighbours a cell reads
// were then fetched into that XCD's L2 by the segments just before it, so
// every field is read from HBM about once per launch.
//
// Fusion: the source injection is folded into the forward stress kernel; the
// backward pass runs two kernels per step, {reverse velocity + density
// imaging + adjoint velocity} and {source subtraction + reverse stress +
// lambda/mu imaging + adjoint stress}. Each cell reads its own adjoint field
// before updating it, as the reference's separate launches do, and no kernel
// reads a neighbour value that it writes.
//
// FORM = 0: reference formulation (typo in dsxy/dx, Src/velocity.cu:79;
//           off-by-one x/y C-PML predicate, Src/velocity.cu:46-90; swapped
//           adjoint C-PML coefficients, Src/velocity_adj.cu:139,142).
// FORM = 1: the same three places corrected.
#include "efwi_common.h"
#include "efwi_kernels.h"

namespace efwi {
namespace {

constexpr float C1 = (float)(9.0 / 8.0);   // Src/stress.cu:33
constexpr float C2 = (float)(1.0 / 24.0);  // Src/stress.cu:34

// Backward / forward 4th-order staggered differences (Src/stress.cu:40-68).
__device__ __forceinline__ float Db(const float* __restrict__ f, long i, long s, float h) {
  return (C1 * (f[i] - f[i - s]) - C2 * (f[i + s] - f[i - 2 * s])) / h;
}
__device__ __forceinline__ float Df(const float* __restrict__ f, long i, long s, float h) {
  return (C1 * (f[i + s] - f[i]) - C2 * (f[i + 2 * s] - f[i - s])) / h;
}
// Adjoint forms as written in Src/stress_adj.cu:54-59.
__device__ __forceinline__ float Af(const float* __restrict__ f, long i, long s, float h) {
  return (-C1 * (f[i + s] - f[i]) + C2 * (f[i + 2 * s] - f[i - s])) / h;
}
__device__ __forceinline__ float Ab(const float* __restrict__ f, long i, long s, float h) {
  return (-C1 * (f[i] - f[i - s]) + C2 * (f[i + s] - f[i - 2 * s])) / h;
}
// dsxy/dx of the vy update (Src/velocity.cu:79 and :122).
template <int FORM>
__device__ __forceinline__ float Dsxy_x(const float* __restrict__ f, long i, long s, float h) {
  if (FORM == 1) return Db(f, i, s, h);
  return (C1 * (f[i] - f[i - s]) - C2 * (f[i - s] - f[i - 2 * s])) / h;
}
// a * d(psi), skipped where a == 0 (outside the PML along that axis).
__device__ __forceinline__ float apsi_f(float a, const float* __restrict__ f, long i, long s, float h) {
  return a != 0.0f ? a * Af(f, i, s, h) : 0.0f;
}
__device__ __forceinline__ float apsi_b(float a, const float* __restrict__ f, long i, long s, float h) {
  return a != 0.0f ? a * Ab(f, i, s, h) : 0.0f;
}

__device__ __forceinline__ bool in_pml(int i, int n, int p) { return i < p || i > n - p - 1; }
// x/y predicate of the forward velocity kernel.
template <int FORM>
__device__ __forceinline__ bool in_pml_v(int i, int n, int p) {
  if (FORM == 1) return in_pml(i, n, p);
  return i < p || i > n - p;
}

// psi <- b psi + a d ; d <- d / K + psi
__device__ __forceinline__ float cpml(float* __restrict__ m, long i, float b, float a, float K, float d) {
  float ps = b * m[i] + a * d;
  m[i] = ps;
  return d / K + ps;
}

struct Cell {
  int z, x, y;
  long i;
};

// Workgroup b (64 x R threads: y lanes x rows) covers R rows of the strip of
// XCD b % 8 and Z consecutive planes; the strip is walked plane chunk by
// plane chunk (z slowest), so each XCD's L2 sees a compact, z-ordered region.
template <int Z, typename Body>
__device__ __forceinline__ void sweep(const Sweep& S, const Geom& G, Body&& body) {
  const int b = blockIdx.x;
  const int xcd = b & 7, j = b >> 3;
  const int per_plane = S.rg_max * S.nyb;
  const int zc = j / per_plane;
  const int r = j - zc * per_plane;
  const int rg = r / S.nyb;
  const int yb = r - rg * S.nyb;
  const int rows = S.rows_base + (xcd < S.rows_rem ? 1 : 0);
  const int row = rg * S.R + (int)threadIdx.y;
  if (row >= rows) return;
  Cell c;
  c.x = S.x0 + xcd * S.rows_base + min(xcd, S.rows_rem) + row;
  c.y = yb * 64 + (int)threadIdx.x;
  if (c.y < S.y0 || c.y > S.y1) return;
  const int za = S.z0 + zc * Z;
  const long base = (long)c.x * G.py + c.y;
  if (Z == 1) {
    c.z = za;
    c.i = (long)za * G.pxy + base;
    body(c);
  } else {
    const int zb = min(za + Z, S.z0 + S.nz);
    for (int z = za; z < zb; z++) {
      c.z = z;
      c.i = (long)z * G.pxy + base;
      body(c);
    }
  }
}

__device__ __forceinline__ bool interior(const Geom& G, const Cell& c) {
  const int p = G.npml;
  return c.z >= p && c.z <= G.nz - 1 - p && c.x >= p && c.x <= G.nx - 1 - p && c.y >= p &&
         c.y <= G.ny - 1 - p;
}

// The cell functions below are written in three phases: every global load of
// the cell (neighbours and own values), then the arithmetic, then every store.
// Loads never follow a store in program order, so a wavefront issues all of
// its loads at once and waits once. Each stored value is still the same
// expression tree as the reference's, so the phase order changes no bit.

// ---------------------------------------------------------------------------
// Forward stress of one cell, Src/stress.cu:36-101.
// ---------------------------------------------------------------------------
__device__ __forceinline__ void stress_fwd_cell(const Geom& G, const Wave& W, const Psi& P, const Model& M,
                                                const Cpml& C, const Cell& c) {
  const long i = c.i, SZ = G.pxy, SX = G.py;
  const int z = c.z, x = c.x, y = c.y, p = G.npml;
  float dvz_dz = Db(W.vz, i, SZ, G.dz);
  float dvx_dx = Db(W.vx, i, SX, G.dx);
  float dvy_dy = Db(W.vy, i, 1, G.dy);
  float dvx_dz = Df(W.vx, i, SZ, G.dz);
  float dvx_dy = Df(W.vx, i, 1, G.dy);
  float dvz_dx = Df(W.vz, i, SX, G.dx);
  float dvz_dy = Df(W.vz, i, 1, G.dy);
  float dvy_dx = Df(W.vy, i, SX, G.dx);
  float dvy_dz = Df(W.vy, i, SZ, G.dz);
  const float lam = M.lam[i], mu = M.mu[i], dt = G.dt;
  const float szz = W.szz[i], sxx = W.sxx[i], syy = W.syy[i];
  const float sxz = W.sxz[i], sxy = W.sxy[i], syz = W.syz[i];
  const bool pz = in_pml(z, G.nz, p), px = in_pml(x, G.nx, p), py = in_pml(y, G.ny, p);
  if (pz) {
    dvz_dz = cpml(P.vz_z, i, C.b[0][z], C.a[0][z], C.K[0][z], dvz_dz);
    dvx_dz = cpml(P.vx_z, i, C.bh[0][z], C.ah[0][z], C.Kh[0][z], dvx_dz);
    dvy_dz = cpml(P.vy_z, i, C.bh[0][z], C.ah[0][z], C.Kh[0][z], dvy_dz);
  }
  if (px) {
    dvx_dx = cpml(P.vx_x, i, C.b[1][x], C.a[1][x], C.K[1][x], dvx_dx);
    dvz_dx = cpml(P.vz_x, i, C.bh[1][x], C.ah[1][x], C.Kh[1][x], dvz_dx);
    dvy_dx = cpml(P.vy_x, i, C.bh[1][x], C.ah[1][x], C.Kh[1][x], dvy_dx);
  }
  if (py) {
    dvy_dy = cpml(P.vy_y, i, C.b[2][y], C.a[2][y], C.K[2][y], dvy_dy);
    dvz_dy = cpml(P.vz_y, i, C.bh[2][y], C.ah[2][y], C.Kh[2][y], dvz_dy);
    dvx_dy = cpml(P.vx_y, i, C.bh[2][y], C.ah[2][y], C.Kh[2][y], dvx_dy);
  }
  const float l2m = lam + 2.0f * mu;
  W.szz[i] = szz + (l2m * dvz_dz + lam * dvx_dx + lam * dvy_dy) * dt;
  W.sxx[i] = sxx + (lam * dvz_dz + l2m * dvx_dx + lam * dvy_dy) * dt;
  W.syy[i] = syy + (lam * dvz_dz + lam * dvx_dx + l2m * dvy_dy) * dt;
  W.sxz[i] = sxz + mu * (dvx_dz + dvz_dx) * dt;
  W.sxy[i] = sxy + mu * (dvx_dy + dvy_dx) * dt;
  W.syz[i] = syz + mu * (dvy_dz + dvz_dy) * dt;
}

// ---------------------------------------------------------------------------
// Forward velocity of one cell, Src/velocity.cu:175-240.
// ---------------------------------------------------------------------------
template <int FORM>
__device__ __forceinline__ void velocity_fwd_cell(const Geom& G, const Wave& W, const Psi& P, const Model& M,
                                                  const Cpml& C, const Cell& c) {
  const long i = c.i, SZ = G.pxy, SX = G.py;
  const int z = c.z, x = c.x, y = c.y, p = G.npml;
  float dszz_dz = Df(W.szz, i, SZ, G.dz);
  float dsxz_dx = Db(W.sxz, i, SX, G.dx);
  float dsyz_dy = Db(W.syz, i, 1, G.dy);
  float dsxz_dz = Db(W.sxz, i, SZ, G.dz);
  float dsxx_dx = Df(W.sxx, i, SX, G.dx);
  float dsxy_dy = Db(W.sxy, i, 1, G.dy);
  float dsyz_dz = Db(W.syz, i, SZ, G.dz);
  float dsxy_dx = Dsxy_x<FORM>(W.sxy, i, SX, G.dx);
  float dsyy_dy = Df(W.syy, i, 1, G.dy);
  const float byc = M.byc[i], dt = G.dt;
  const float vz = W.vz[i], vx = W.vx[i], vy = W.vy[i];
  const bool pz = in_pml(z, G.nz, p);
  const bool px = in_pml_v<FORM>(x, G.nx, p), py = in_pml_v<FORM>(y, G.ny, p);
  if (pz) {
    dszz_dz = cpml(P.szz_z, i, C.bh[0][z], C.ah[0][z], C.Kh[0][z], dszz_dz);
    dsxz_dz = cpml(P.sxz_z, i, C.b[0][z], C.a[0][z], C.K[0][z], dsxz_dz);
    dsyz_dz = cpml(P.syz_z, i, C.b[0][z], C.a[0][z], C.K[0][z], dsyz_dz);
  }
  if (px) {
    dsxz_dx = cpml(P.sxz_x, i, C.b[1][x], C.a[1][x], C.K[1][x], dsxz_dx);
    dsxx_dx = cpml(P.sxx_x, i, C.bh[1][x], C.ah[1][x], C.Kh[1][x], dsxx_dx);
    dsxy_dx = cpml(P.sxy_x, i, C.b[1][x], C.a[1][x], C.K[1][x], dsxy_dx);
  }
  if (py) {
    dsyz_dy = cpml(P.syz_y, i, C.b[2][y], C.a[2][y], C.K[2][y], dsyz_dy);
    dsxy_dy = cpml(P.sxy_y, i, C.b[2][y], C.a[2][y], C.K[2][y], dsxy_dy);
    dsyy_dy = cpml(P.syy_y, i, C.bh[2][y], C.ah[2][y], C.Kh[2][y], dsyy_dy);
  }
  W.vz[i] = vz + (dszz_dz + dsxz_dx + dsyz_dy) * byc * dt;
  W.vx[i] = vx + (dsxz_dz + dsxx_dx + dsxy_dy) * byc * dt;
  W.vy[i] = vy + (dsyz_dz + dsxy_dx + dsyy_dy) * byc * dt;
}

// ---------------------------------------------------------------------------
// Backward velocity step of one cell: for interior cells the reverse velocity
// with the density imaging term (Src/velocity.cu:242-275), then for every cell
// of [2, n-3] the adjoint velocity (Src/velocity_adj.cu:149-296). The reverse
// part reads v-dagger before the adjoint part updates it, as the reference's
// launch order does. The adjoint part reads the adjoint stress and the psi
// arrays written by the adjoint stress step, and writes v-dagger and the psi
// arrays the adjoint stress step reads.
// ---------------------------------------------------------------------------
template <int FORM>
__device__ __forceinline__ void velocity_bwd_cell(const Geom& G, const Wave& W, const Wave& A, const Psi& P,
                                                  const Model& M, const Cpml& C, const Grad& Gr, bool rev,
                                                  const Cell& c) {
  const long i = c.i, SZ = G.pxy, SX = G.py;
  const int z = c.z, x = c.x, y = c.y, p = G.npml;
  const float dt = G.dt;
  // ---- loads
  float r_sz = 0.0f, r_sx = 0.0f, r_sy = 0.0f, vz = 0.0f, vx = 0.0f, vy = 0.0f, grho = 0.0f;
  if (rev) {
    r_sz = Df(W.szz, i, SZ, G.dz) + Db(W.sxz, i, SX, G.dx) + Db(W.syz, i, 1, G.dy);
    r_sx = Db(W.sxz, i, SZ, G.dz) + Df(W.sxx, i, SX, G.dx) + Db(W.sxy, i, 1, G.dy);
    r_sy = Db(W.syz, i, SZ, G.dz) + Dsxy_x<FORM>(W.sxy, i, SX, G.dx) + Df(W.syy, i, 1, G.dy);
    vz = W.vz[i]; vx = W.vx[i]; vy = W.vy[i];
    grho = Gr.grho[i];
  }
  const float lambda = M.lam[i], mu = M.mu[i], byc = M.byc[i];
  const float avz = A.vz[i], avx = A.vx[i], avy = A.vy[i];
  const float ax = C.a[1][x], az = C.a[0][z], ay = C.a[2][y];
  const float axh = C.ah[1][x], azh = C.ah[0][z], ayh = C.ah[2][y];
  const float Kx = C.K[1][x], Kz = C.K[0][z], Ky = C.K[2][y];
  const float Kxh = C.Kh[1][x], Kzh = C.Kh[0][z], Kyh = C.Kh[2][y];
  // vx-dagger terms
  const float x1 = apsi_f(ax, P.vx_x, i, SX, G.dx);
  const float x_szz = Af(A.szz, i, SX, G.dx);
  const float x_sxx = Af(A.sxx, i, SX, G.dx);
  const float x_syy = Af(A.syy, i, SX, G.dx);
  const float x6 = apsi_b(azh, P.vx_z, i, SZ, G.dz);
  const float x_sxz = Ab(A.sxz, i, SZ, G.dz);
  const float x8 = apsi_b(ayh, P.vx_y, i, 1, G.dy);
  const float x_sxy = Ab(A.sxy, i, 1, G.dy);
  // vz-dagger terms
  const float z1 = apsi_f(az, P.vz_z, i, SZ, G.dz);
  const float z_szz = Af(A.szz, i, SZ, G.dz);
  const float z_sxx = Af(A.sxx, i, SZ, G.dz);
  const float z_syy = Af(A.syy, i, SZ, G.dz);
  const float z8 = apsi_b(axh, P.vz_x, i, SX, G.dx);
  const float z_sxz = Ab(A.sxz, i, SX, G.dx);
  const float z6 = apsi_b(ayh, P.vz_y, i, 1, G.dy);
  const float z_syz = Ab(A.syz, i, 1, G.dy);
  // vy-dagger terms
  const float y1 = apsi_f(ay, P.vy_y, i, 1, G.dy);
  const float y_syy = Af(A.syy, i, 1, G.dy);
  const float y_sxx = Af(A.sxx, i, 1, G.dy);
  const float y_szz = Af(A.szz, i, 1, G.dy);
  const float y6 = apsi_b(axh, P.vy_x, i, SX, G.dx);
  const float y_sxy = Ab(A.sxy, i, SX, G.dx);
  const float y8 = apsi_b(azh, P.vy_z, i, SZ, G.dz);
  const float y_syz = Ab(A.syz, i, SZ, G.dz);
  const bool pz = in_pml(z, G.nz, p), px = in_pml(x, G.nx, p), py = in_pml(y, G.ny, p);
  // ---- arithmetic and stores
  if (rev) {
    W.vz[i] = vz - r_sz * byc * dt;
    W.vx[i] = vx - r_sx * byc * dt;
    W.vy[i] = vy - r_sy * byc * dt;
    const float nb2 = -(byc * byc);
    Gr.grho[i] = grho + -(avz * r_sz * dt * nb2 + avx * r_sx * dt * nb2 + avy * r_sy * dt * nb2);
  }
  const float l2m = lambda + 2.0f * mu;
  const float nvx = avx + (x1 + l2m * x_sxx / Kx * dt + lambda * x_szz / Kx * dt + lambda * x_syy / Kx * dt +
                           mu / Kzh * x_sxz * dt + x6 + mu / Kyh * x_sxy * dt + x8);
  const float nvz = avz + (z1 + l2m * z_szz / Kz * dt + lambda * z_sxx / Kz * dt + lambda * z_syy / Kz * dt +
                           mu / Kyh * z_syz * dt + z6 + mu / Kxh * z_sxz * dt + z8);
  const float nvy = avy + (y1 + l2m * y_syy / Ky * dt + lambda * y_sxx / Ky * dt + lambda * y_szz / Ky * dt +
                           mu / Kxh * y_sxy * dt + y6 + mu / Kzh * y_syz * dt + y8);
  A.vx[i] = nvx;
  A.vz[i] = nvz;
  A.vy[i] = nvy;
  if (px) {
    P.sxx_x[i] = C.bh[1][x] * P.sxx_x[i] + byc * nvx * dt;
    P.sxz_x[i] = C.b[1][x] * P.sxz_x[i] + byc * nvz * dt;
    P.sxy_x[i] = C.b[1][x] * P.sxy_x[i] + byc * nvy * dt;
  }
  if (pz) {
    P.sxz_z[i] = C.b[0][z] * P.sxz_z[i] + byc * nvx * dt;
    P.szz_z[i] = C.bh[0][z] * P.szz_z[i] + byc * nvz * dt;
    // Src/velocity_adj.cu:285 decays psi_dsyz_dz with b_z_half (the forward uses b_z)
    P.syz_z[i] = (FORM == 1 ? C.b[0][z] : C.bh[0][z]) * P.syz_z[i] + byc * nvy * dt;
  }
  if (py) {
    P.sxy_y[i] = C.b[2][y] * P.sxy_y[i] + byc * nvx * dt;
    P.syz_y[i] = C.b[2][y] * P.syz_y[i] + byc * nvz * dt;
    // Src/velocity_adj.cu:288 decays psi_dsyy_dy with b_y (the forward uses b_y_half)
    P.syy_y[i] = (FORM == 1 ? C.bh[2][y] : C.b[2][y]) * P.syy_y[i] + byc * nvy * dt;
  }
}

// ---------------------------------------------------------------------------
// Backward stress step of one cell: for interior cells the source subtraction
// at the source cell (Src/libCUFD.cu:547-549), the reverse stress and the
// lambda/mu imaging terms (Src/stress.cu:103-139); then for every cell of
// [2, n-3] the adjoint stress (Src/stress_adj.cu:1-146). The imaging terms
// read the adjoint stress before the adjoint part updates it.
// ---------------------------------------------------------------------------
__device__ __forceinline__ void stress_bwd_cell(const Geom& G, const Wave& W, const Wave& A, const Psi& P,
                                                const Model& M, const Cpml& C, const Grad& Gr, const Source& S,
                                                int it, bool rev, const Cell& c) {
  const long i = c.i, SZ = G.pxy, SX = G.py;
  const int z = c.z, x = c.x, y = c.y, p = G.npml;
  const float dt = G.dt;
  const float lambda = M.lam[i], mu = M.mu[i], byc = M.byc[i];
  // ---- loads of the reverse part
  float dvz_dz = 0, dvx_dx = 0, dvy_dy = 0, dvx_dz = 0, dvx_dy = 0, dvz_dx = 0, dvz_dy = 0, dvy_dx = 0, dvy_dz = 0;
  float szz = 0, sxx = 0, syy = 0, sxz = 0, sxy = 0, syz = 0, glam = 0, gmu = 0, amp = 0;
  if (rev) {
    dvz_dz = Db(W.vz, i, SZ, G.dz);
    dvx_dx = Db(W.vx, i, SX, G.dx);
    dvy_dy = Db(W.vy, i, 1, G.dy);
    dvx_dz = Df(W.vx, i, SZ, G.dz);
    dvx_dy = Df(W.vx, i, 1, G.dy);
    dvz_dx = Df(W.vz, i, SX, G.dx);
    dvz_dy = Df(W.vz, i, 1, G.dy);
    dvy_dx = Df(W.vy, i, SX, G.dx);
    dvy_dz = Df(W.vy, i, SZ, G.dz);
    szz = W.szz[i]; sxx = W.sxx[i]; syy = W.syy[i];
    sxz = W.sxz[i]; sxy = W.sxy[i]; syz = W.syz[i];
    glam = Gr.glam[i]; gmu = Gr.gmu[i];
    if (i == S.cell) amp = 2250000.0f * S.stf[it] * dt * 1.0f;
  }
  // ---- loads of the adjoint part
  const float aszz = A.szz[i], asxx = A.sxx[i], asyy = A.syy[i];
  const float asxz = A.sxz[i], asxy = A.sxy[i], asyz = A.syz[i];
  const float ax = C.a[1][x], az = C.a[0][z], ay = C.a[2][y];
  const float Kx = C.K[1][x], Kz = C.K[0][z], Ky = C.K[2][y];
  const float xz1 = apsi_f(ax, P.sxz_x, i, SX, G.dx);
  const float xz_vz = Af(A.vz, i, SX, G.dx);
  const float xz3 = apsi_f(az, P.sxz_z, i, SZ, G.dz);
  const float xz_vx = Af(A.vx, i, SZ, G.dz);
  const float xy1 = apsi_f(ax, P.sxy_x, i, SX, G.dx);
  const float xy_vy = Af(A.vy, i, SX, G.dx);
  const float xy3 = apsi_f(ay, P.sxy_y, i, 1, G.dy);
  const float xy_vx = Af(A.vx, i, 1, G.dy);
  const float yz1 = apsi_f(ay, P.syz_y, i, 1, G.dy);
  const float yz_vz = Af(A.vz, i, 1, G.dy);
  const float yz3 = apsi_f(az, P.syz_z, i, SZ, G.dz);
  const float yz_vy = Af(A.vy, i, SZ, G.dz);
  const float txx = apsi_b(C.ah[1][x], P.sxx_x, i, SX, G.dx);
  const float xx_vx = Ab(A.vx, i, SX, G.dx);
  const float tzz = apsi_b(C.ah[0][z], P.szz_z, i, SZ, G.dz);
  const float zz_vz = Ab(A.vz, i, SZ, G.dz);
  const float tyy = apsi_b(C.ah[2][y], P.syy_y, i, 1, G.dy);
  const float yy_vy = Ab(A.vy, i, 1, G.dy);
  const bool pz = in_pml(z, G.nz, p), px = in_pml(x, G.nx, p), py = in_pml(y, G.ny, p);
  // ---- arithmetic and stores
  const float l2m = lambda + 2.0f * mu;
  if (rev) {
    if (i == S.cell) {
      szz -= amp; sxx -= amp; syy -= amp;
    }
    W.szz[i] = szz - (l2m * dvz_dz + lambda * dvx_dx + lambda * dvy_dy) * dt;
    W.sxx[i] = sxx - (lambda * dvz_dz + l2m * dvx_dx + lambda * dvy_dy) * dt;
    W.syy[i] = syy - (lambda * dvz_dz + lambda * dvx_dx + l2m * dvy_dy) * dt;
    W.sxz[i] = sxz - mu * (dvx_dz + dvz_dx) * dt;
    W.sxy[i] = sxy - mu * (dvx_dy + dvy_dx) * dt;
    W.syz[i] = syz - mu * (dvy_dz + dvz_dy) * dt;
    Gr.glam[i] = glam + -(aszz + asxx + asyy) * (dvz_dz + dvx_dx + dvy_dy) * dt;
    Gr.gmu[i] = gmu + -(2.0f * aszz * dvz_dz + 2.0f * asxx * dvx_dx + 2.0f * asyy * dvy_dy +
                        asxz * (dvx_dz + dvz_dx) + asyz * (dvy_dz + dvz_dy) + asxy * (dvy_dx + dvx_dy)) * dt;
  }
  const float nsxz = asxz + (xz1 + xz_vz / Kx * byc * dt + xz3 + xz_vx / Kz * byc * dt);
  const float nsxy = asxy + (xy1 + xy_vy / Kx * byc * dt + xy3 + xy_vx / Ky * byc * dt);
  const float nsyz = asyz + (yz1 + yz_vz / Ky * byc * dt + yz3 + yz_vy / Kz * byc * dt);
  const float nsxx = asxx + (txx + byc * xx_vx / C.Kh[1][x] * dt);
  const float nszz = aszz + (tzz + byc * zz_vz / C.Kh[0][z] * dt);
  const float nsyy = asyy + (tyy + byc * yy_vy / C.Kh[2][y] * dt);
  A.sxz[i] = nsxz;
  A.sxy[i] = nsxy;
  A.syz[i] = nsyz;
  A.sxx[i] = nsxx;
  A.szz[i] = nszz;
  A.syy[i] = nsyy;
  if (px) {
    P.vz_x[i] = C.bh[1][x] * P.vz_x[i] + nsxz * mu * dt;
    P.vy_x[i] = C.bh[1][x] * P.vy_x[i] + nsxy * mu * dt;
    P.vx_x[i] = C.b[1][x] * P.vx_x[i] + lambda * nszz * dt + l2m * nsxx * dt + lambda * nsyy * dt;
  }
  if (pz) {
    P.vx_z[i] = C.bh[0][z] * P.vx_z[i] + nsxz * mu * dt;
    P.vy_z[i] = C.bh[0][z] * P.vy_z[i] + nsyz * mu * dt;
    P.vz_z[i] = C.b[0][z] * P.vz_z[i] + l2m * nszz * dt + lambda * nsxx * dt + lambda * nsyy * dt;
  }
  if (py) {
    P.vx_y[i] = C.bh[2][y] * P.vx_y[i] + nsxy * mu * dt;
    P.vz_y[i] = C.bh[2][y] * P.vz_y[i] + nsyz * mu * dt;
    P.vy_y[i] = C.b[2][y] * P.vy_y[i] + lambda * nszz * dt + lambda * nsxx * dt + l2m * nsyy * dt;
  }
}

// ---------------------------------------------------------------------------
// Kernels.
// ---------------------------------------------------------------------------
// Forward stress, then the source of this step at its cell (Src/libCUFD.cu:319-330).
template <int Z>
__global__ __launch_bounds__(256) void k_stress_fwd(Sweep Sw, Geom G, Wave W, Psi P, Model M, Cpml C, Source S,
                                                    int it) {
  sweep<Z>(Sw, G, [&](const Cell& c) {
    stress_fwd_cell(G, W, P, M, C, c);
    if (c.i == S.cell) {
      const float a = 2250000.0f * S.stf[it] * G.dt * 1.0f;
      W.szz[c.i] += a; W.sxx[c.i] += a; W.syy[c.i] += a;
    }
  });
}

template <int FORM, int Z>
__global__ __launch_bounds__(256) void k_velocity_fwd(Sweep Sw, Geom G, Wave W, Psi P, Model M, Cpml C) {
  sweep<Z>(Sw, G, [&](const Cell& c) { velocity_fwd_cell<FORM>(G, W, P, M, C, c); });
}

// Backward velocity step: reverse velocity + density imaging (interior), then
// the adjoint velocity ([2, n-3]).
template <int FORM, int Z>
__global__ __launch_bounds__(256) void k_velocity_bwd(Sweep Sw, Geom G, Wave W, Wave A, Psi P, Model M, Cpml C,
                                                      Grad Gr) {
  sweep<Z>(Sw, G, [&](const Cell& c) { velocity_bwd_cell<FORM>(G, W, A, P, M, C, Gr, interior(G, c), c); });
}

// Backward stress step: (source subtraction +) reverse stress + lambda/mu
// imaging (interior), then the adjoint stress ([2, n-3]).
template <int Z>
__global__ __launch_bounds__(256) void k_stress_bwd(Sweep Sw, Geom G, Wave W, Wave A, Psi P, Model M, Cpml C,
                                                    Grad Gr, Source S, int it) {
  sweep<Z>(Sw, G, [&](const Cell& c) { stress_bwd_cell(G, W, A, P, M, C, Gr, S, it, interior(G, c), c); });
}

template <int FORM, int Z>
__global__ __launch_bounds__(256) void k_velocity_adj(Sweep Sw, Geom G, Wave W, Wave A, Psi P, Model M, Cpml C) {
  sweep<Z>(Sw, G, [&](const Cell& c) { velocity_bwd_cell<FORM>(G, W, A, P, M, C, Grad{}, false, c); });
}

template <int Z>
__global__ __launch_bounds__(256) void k_stress_adj(Sweep Sw, Geom G, Wave W, Wave A, Psi P, Model M, Cpml C) {
  sweep<Z>(Sw, G, [&](const Cell& c) {
    Source none;
    none.cell = -1;
    none.stf = nullptr;
    stress_bwd_cell(G, W, A, P, M, C, Grad{}, none, 0, false, c);
  });
}

}  // namespace

// ---------------------------------------------------------------------------
// Launchers.
// ---------------------------------------------------------------------------
Sweep make_sweep(int z0, int z1, int x0, int x1, int y0, int y1, Tile t) {
  Sweep S;
  S.R = t.rows < 1 ? 1 : (t.rows > 4 ? 4 : t.rows);
  S.Z = t.zdepth < 1 ? 1 : t.zdepth;
  S.z0 = z0; S.nz = z1 - z0 + 1;
  S.x0 = x0;
  const int nxr = x1 - x0 + 1;
  S.rows_base = nxr / 8;
  S.rows_rem = nxr % 8;
  const int rows_max = S.rows_base + (S.rows_rem ? 1 : 0);
  S.rg_max = (rows_max + S.R - 1) / S.R;
  S.y0 = y0; S.y1 = y1;
  S.nyb = (y1 + 64) / 64;  // segments of 64 aligned at y = 0
  S.nzc = (S.nz + S.Z - 1) / S.Z;
  S.blocks = 8L * S.rg_max * S.nyb * S.nzc;
  return S;
}

static inline Sweep sweep_full(const Geom& G, Tile t) { return make_sweep(2, G.nz - 3, 2, G.nx - 3, 2, G.ny - 3, t); }

static inline dim3 grid_of(const Sweep& S) { return dim3((unsigned)S.blocks); }
static inline dim3 block_of(const Sweep& S) { return dim3(64, S.R); }

// Dispatch on the compile-time plane depth of a tile (1, 2 or 4).
#define EFWI_DISPATCH_Z(t, LAUNCH) \
  do {                             \
    if ((t).zdepth >= 4) {         \
      constexpr int ZD = 4;        \
      LAUNCH;                      \
    } else if ((t).zdepth == 2) {  \
      constexpr int ZD = 2;        \
      LAUNCH;                      \
    } else {                       \
      constexpr int ZD = 1;        \
      LAUNCH;                      \
    }                              \
  } while (0)

static inline Tile norm_tile(Tile t) {
  t.zdepth = t.zdepth >= 4 ? 4 : (t.zdepth == 2 ? 2 : 1);
  return t;
}

void launch_stress_fwd(Tile t, const Geom& G, const Wave& W, const Psi& P, const Model& M, const Cpml& C,
                       const Source& S, int it, hipStream_t s) {
  t = norm_tile(t);
  const Sweep Sw = sweep_full(G, t);
  EFWI_DISPATCH_Z(t, hipLaunchKernelGGL(k_stress_fwd<ZD>, grid_of(Sw), block_of(Sw), 0, s, Sw, G, W, P, M, C, S, it));
}
void launch_velocity_fwd(Tile t, int form, const Geom& G, const Wave& W, const Psi& P, const Model& M,
                         const Cpml& C, hipStream_t s) {
  t = norm_tile(t);
  const Sweep Sw = sweep_full(G, t);
  if (form == 1)
    EFWI_DISPATCH_Z(t, hipLaunchKernelGGL((k_velocity_fwd<1, ZD>), grid_of(Sw), block_of(Sw), 0, s, Sw, G, W, P, M, C));
  else
    EFWI_DISPATCH_Z(t, hipLaunchKernelGGL((k_velocity_fwd<0, ZD>), grid_of(Sw), block_of(Sw), 0, s, Sw, G, W, P, M, C));
}
void launch_velocity_bwd(Tile t, int form, const Geom& G, const Wave& W, const Wave& A, const Psi& P,
                         const Model& M, const Cpml& C, const Grad& Gr, hipStream_t s) {
  t = norm_tile(t);
  const Sweep Sw = sweep_full(G, t);
  if (form == 1)
    EFWI_DISPATCH_Z(t, hipLaunchKernelGGL((k_velocity_bwd<1, ZD>), grid_of(Sw), block_of(Sw), 0, s, Sw, G, W, A, P,
                                          M, C, Gr));
  else
    EFWI_DISPATCH_Z(t, hipLaunchKernelGGL((k_velocity_bwd<0, ZD>), grid_of(Sw), block_of(Sw), 0, s, Sw, G, W, A, P,
                                          M, C, Gr));
}
void launch_stress_bwd(Tile t, const Geom& G, const Wave& W, const Wave& A, const Psi& P, const Model& M,
                       const Cpml& C, const Grad& Gr, const Source& S, int it, hipStream_t s) {
  t = norm_tile(t);
  const Sweep Sw = sweep_full(G, t);
  EFWI_DISPATCH_Z(t, hipLaunchKernelGGL(k_stress_bwd<ZD>, grid_of(Sw), block_of(Sw), 0, s, Sw, G, W, A, P, M, C, Gr,
                                        S, it));
}
void launch_velocity_adj(int form, const Geom& G, const Wave& A, const Psi& P, const Model& M,
                         const Cpml& C, hipStream_t s) {
  const Sweep Sw = sweep_full(G, Tile{1, 1});
  if (form == 1)
    hipLaunchKernelGGL((k_velocity_adj<1, 1>), grid_of(Sw), block_of(Sw), 0, s, Sw, G, Wave{}, A, P, M, C);
  else
    hipLaunchKernelGGL((k_velocity_adj<0, 1>), grid_of(Sw), block_of(Sw), 0, s, Sw, G, Wave{}, A, P, M, C);
}
void launch_stress_adj(const Geom& G, const Wave& A, const Psi& P, const Model& M, const Cpml& C,
                       hipStream_t s) {
  const Sweep Sw = sweep_full(G, Tile{1, 1});
  hipLaunchKernelGGL(k_stress_adj<1>, grid_of(Sw), block_of(Sw), 0, s, Sw, G, Wave{}, A, P, M, C);
}

}  // namespace efwi
