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
template <typename Body>
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
  const int za = S.z0 + zc * S.Z;
  const int zb = min(za + S.Z, S.z0 + S.nz);
  const long base = (long)c.x * G.py + c.y;
  for (int z = za; z < zb; z++) {
    c.z = z;
    c.i = (long)z * G.pxy + base;
    body(c);
  }
}

__device__ __forceinline__ bool interior(const Geom& G, const Cell& c) {
  const int p = G.npml;
  return c.z >= p && c.z <= G.nz - 1 - p && c.x >= p && c.x <= G.nx - 1 - p && c.y >= p &&
         c.y <= G.ny - 1 - p;
}

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
  const bool pz = in_pml(z, G.nz, p), px = in_pml(x, G.nx, p), py = in_pml(y, G.ny, p);
  if (pz) dvz_dz = cpml(P.vz_z, i, C.b[0][z], C.a[0][z], C.K[0][z], dvz_dz);
  if (px) dvx_dx = cpml(P.vx_x, i, C.b[1][x], C.a[1][x], C.K[1][x], dvx_dx);
  if (py) dvy_dy = cpml(P.vy_y, i, C.b[2][y], C.a[2][y], C.K[2][y], dvy_dy);
  const float lam = M.lam[i], mu = M.mu[i], dt = G.dt;
  const float l2m = lam + 2.0f * mu;
  W.szz[i] += (l2m * dvz_dz + lam * dvx_dx + lam * dvy_dy) * dt;
  W.sxx[i] += (lam * dvz_dz + l2m * dvx_dx + lam * dvy_dy) * dt;
  W.syy[i] += (lam * dvz_dz + lam * dvx_dx + l2m * dvy_dy) * dt;
  float dvx_dz = Df(W.vx, i, SZ, G.dz);
  float dvx_dy = Df(W.vx, i, 1, G.dy);
  float dvz_dx = Df(W.vz, i, SX, G.dx);
  float dvz_dy = Df(W.vz, i, 1, G.dy);
  float dvy_dx = Df(W.vy, i, SX, G.dx);
  float dvy_dz = Df(W.vy, i, SZ, G.dz);
  if (pz) {
    dvx_dz = cpml(P.vx_z, i, C.bh[0][z], C.ah[0][z], C.Kh[0][z], dvx_dz);
    dvy_dz = cpml(P.vy_z, i, C.bh[0][z], C.ah[0][z], C.Kh[0][z], dvy_dz);
  }
  if (px) {
    dvz_dx = cpml(P.vz_x, i, C.bh[1][x], C.ah[1][x], C.Kh[1][x], dvz_dx);
    dvy_dx = cpml(P.vy_x, i, C.bh[1][x], C.ah[1][x], C.Kh[1][x], dvy_dx);
  }
  if (py) {
    dvz_dy = cpml(P.vz_y, i, C.bh[2][y], C.ah[2][y], C.Kh[2][y], dvz_dy);
    dvx_dy = cpml(P.vx_y, i, C.bh[2][y], C.ah[2][y], C.Kh[2][y], dvx_dy);
  }
  W.sxz[i] = W.sxz[i] + mu * (dvx_dz + dvz_dx) * dt;
  W.sxy[i] = W.sxy[i] + mu * (dvx_dy + dvy_dx) * dt;
  W.syz[i] = W.syz[i] + mu * (dvy_dz + dvz_dy) * dt;
}

// ---------------------------------------------------------------------------
// Forward velocity of one cell, Src/velocity.cu:175-240.
// ---------------------------------------------------------------------------
template <int FORM>
__device__ __forceinline__ void velocity_fwd_cell(const Geom& G, const Wave& W, const Psi& P, const Model& M,
                                                  const Cpml& C, const Cell& c) {
  const long i = c.i, SZ = G.pxy, SX = G.py;
  const int z = c.z, x = c.x, y = c.y, p = G.npml;
  const float byc = M.byc[i], dt = G.dt;
  const bool pz = in_pml(z, G.nz, p);
  const bool px = in_pml_v<FORM>(x, G.nx, p), py = in_pml_v<FORM>(y, G.ny, p);
  // vz
  float dszz_dz = Df(W.szz, i, SZ, G.dz);
  float dsxz_dx = Db(W.sxz, i, SX, G.dx);
  float dsyz_dy = Db(W.syz, i, 1, G.dy);
  if (pz) dszz_dz = cpml(P.szz_z, i, C.bh[0][z], C.ah[0][z], C.Kh[0][z], dszz_dz);
  if (px) dsxz_dx = cpml(P.sxz_x, i, C.b[1][x], C.a[1][x], C.K[1][x], dsxz_dx);
  if (py) dsyz_dy = cpml(P.syz_y, i, C.b[2][y], C.a[2][y], C.K[2][y], dsyz_dy);
  W.vz[i] += (dszz_dz + dsxz_dx + dsyz_dy) * byc * dt;
  // vx
  float dsxz_dz = Db(W.sxz, i, SZ, G.dz);
  float dsxx_dx = Df(W.sxx, i, SX, G.dx);
  float dsxy_dy = Db(W.sxy, i, 1, G.dy);
  if (pz) dsxz_dz = cpml(P.sxz_z, i, C.b[0][z], C.a[0][z], C.K[0][z], dsxz_dz);
  if (px) dsxx_dx = cpml(P.sxx_x, i, C.bh[1][x], C.ah[1][x], C.Kh[1][x], dsxx_dx);
  if (py) dsxy_dy = cpml(P.sxy_y, i, C.b[2][y], C.a[2][y], C.K[2][y], dsxy_dy);
  W.vx[i] += (dsxz_dz + dsxx_dx + dsxy_dy) * byc * dt;
  // vy
  float dsyz_dz = Db(W.syz, i, SZ, G.dz);
  float dsxy_dx = Dsxy_x<FORM>(W.sxy, i, SX, G.dx);
  float dsyy_dy = Df(W.syy, i, 1, G.dy);
  if (pz) dsyz_dz = cpml(P.syz_z, i, C.b[0][z], C.a[0][z], C.K[0][z], dsyz_dz);
  if (px) dsxy_dx = cpml(P.sxy_x, i, C.b[1][x], C.a[1][x], C.K[1][x], dsxy_dx);
  if (py) dsyy_dy = cpml(P.syy_y, i, C.bh[2][y], C.ah[2][y], C.Kh[2][y], dsyy_dy);
  W.vy[i] += (dsyz_dz + dsxy_dx + dsyy_dy) * byc * dt;
}

// ---------------------------------------------------------------------------
// Reverse stress + lambda/mu imaging of one interior cell, Src/stress.cu:103-139,
// preceded at the source cell by the source subtraction of the same step
// (Src/libCUFD.cu:547-549). Reads the adjoint stress of the cell as it was
// before this step's adjoint update.
// ---------------------------------------------------------------------------
__device__ __forceinline__ void stress_rev_cell(const Geom& G, const Wave& W, const Wave& A, const Model& M,
                                                const Grad& Gr, const Source& S, int it, const Cell& c) {
  const long i = c.i, SZ = G.pxy, SX = G.py;
  const float lam = M.lam[i], mu = M.mu[i], dt = G.dt;
  float szz = W.szz[i], sxx = W.sxx[i], syy = W.syy[i];
  if (i == S.cell) {
    const float a = 2250000.0f * S.stf[it] * dt * 1.0f;
    szz -= a; sxx -= a; syy -= a;
  }
  const float dvz_dz = Db(W.vz, i, SZ, G.dz);
  const float dvx_dx = Db(W.vx, i, SX, G.dx);
  const float dvy_dy = Db(W.vy, i, 1, G.dy);
  const float l2m = lam + 2.0f * mu;
  W.szz[i] = szz - (l2m * dvz_dz + lam * dvx_dx + lam * dvy_dy) * dt;
  W.sxx[i] = sxx - (lam * dvz_dz + l2m * dvx_dx + lam * dvy_dy) * dt;
  W.syy[i] = syy - (lam * dvz_dz + lam * dvx_dx + l2m * dvy_dy) * dt;
  const float dvx_dz = Df(W.vx, i, SZ, G.dz);
  const float dvx_dy = Df(W.vx, i, 1, G.dy);
  const float dvz_dx = Df(W.vz, i, SX, G.dx);
  const float dvz_dy = Df(W.vz, i, 1, G.dy);
  const float dvy_dx = Df(W.vy, i, SX, G.dx);
  const float dvy_dz = Df(W.vy, i, SZ, G.dz);
  W.sxz[i] -= mu * (dvx_dz + dvz_dx) * dt;
  W.sxy[i] -= mu * (dvx_dy + dvy_dx) * dt;
  W.syz[i] -= mu * (dvy_dz + dvz_dy) * dt;
  const float aszz = A.szz[i], asxx = A.sxx[i], asyy = A.syy[i];
  Gr.glam[i] += -(aszz + asxx + asyy) * (dvz_dz + dvx_dx + dvy_dy) * dt;
  Gr.gmu[i] += -(2.0f * aszz * dvz_dz + 2.0f * asxx * dvx_dx + 2.0f * asyy * dvy_dy +
                 A.sxz[i] * (dvx_dz + dvz_dx) + A.syz[i] * (dvy_dz + dvz_dy) +
                 A.sxy[i] * (dvy_dx + dvx_dy)) * dt;
}

// ---------------------------------------------------------------------------
// Reverse velocity + density imaging of one interior cell, Src/velocity.cu:242-275.
// ---------------------------------------------------------------------------
template <int FORM>
__device__ __forceinline__ void velocity_rev_cell(const Geom& G, const Wave& W, const Wave& A, const Model& M,
                                                  const Grad& Gr, const Cell& c) {
  const long i = c.i, SZ = G.pxy, SX = G.py;
  const float byc = M.byc[i], dt = G.dt;
  const float dszz_dz = Df(W.szz, i, SZ, G.dz);
  const float dsxz_dx = Db(W.sxz, i, SX, G.dx);
  const float dsyz_dy = Db(W.syz, i, 1, G.dy);
  const float sz = dszz_dz + dsxz_dx + dsyz_dy;
  W.vz[i] -= sz * byc * dt;
  const float dsxz_dz = Db(W.sxz, i, SZ, G.dz);
  const float dsxx_dx = Df(W.sxx, i, SX, G.dx);
  const float dsxy_dy = Db(W.sxy, i, 1, G.dy);
  const float sx = dsxz_dz + dsxx_dx + dsxy_dy;
  W.vx[i] -= sx * byc * dt;
  const float dsyz_dz = Db(W.syz, i, SZ, G.dz);
  const float dsxy_dx = Dsxy_x<FORM>(W.sxy, i, SX, G.dx);
  const float dsyy_dy = Df(W.syy, i, 1, G.dy);
  const float sy = dsyz_dz + dsxy_dx + dsyy_dy;
  W.vy[i] -= sy * byc * dt;
  const float nb2 = -(byc * byc);
  Gr.grho[i] += -(A.vz[i] * sz * dt * nb2 + A.vx[i] * sx * dt * nb2 + A.vy[i] * sy * dt * nb2);
}

// ---------------------------------------------------------------------------
// Adjoint velocity of one cell, Src/velocity_adj.cu:149-296: reads the adjoint
// stress and the psi arrays written by the adjoint stress kernel; writes
// v-dagger and the psi arrays the adjoint stress kernel reads.
// ---------------------------------------------------------------------------
template <int FORM>
__device__ __forceinline__ void velocity_adj_cell(const Geom& G, const Wave& A, const Psi& P, const Model& M,
                                                  const Cpml& C, const Cell& c) {
  const long i = c.i, SZ = G.pxy, SX = G.py;
  const int z = c.z, x = c.x, y = c.y, p = G.npml;
  const float lambda = M.lam[i], mu = M.mu[i], byc = M.byc[i], dt = G.dt;
  const float l2m = lambda + 2.0f * mu;
  const bool pz = in_pml(z, G.nz, p), px = in_pml(x, G.nx, p), py = in_pml(y, G.ny, p);
  const float ax = C.a[1][x], az = C.a[0][z], ay = C.a[2][y];
  const float axh = C.ah[1][x], azh = C.ah[0][z], ayh = C.ah[2][y];
  const float Kx = C.K[1][x], Kz = C.K[0][z], Ky = C.K[2][y];
  const float Kxh = C.Kh[1][x], Kzh = C.Kh[0][z], Kyh = C.Kh[2][y];
  // vx
  {
    const float t1 = apsi_f(ax, P.vx_x, i, SX, G.dx);
    const float dszz_dx = Af(A.szz, i, SX, G.dx);
    const float dsxx_dx = Af(A.sxx, i, SX, G.dx);
    const float dsyy_dx = Af(A.syy, i, SX, G.dx);
    const float t6 = apsi_b(azh, P.vx_z, i, SZ, G.dz);
    const float dsxz_dz = Ab(A.sxz, i, SZ, G.dz);
    const float t8 = apsi_b(ayh, P.vx_y, i, 1, G.dy);
    const float dsxy_dy = Ab(A.sxy, i, 1, G.dy);
    const float v = A.vx[i] + (t1 + l2m * dsxx_dx / Kx * dt + lambda * dszz_dx / Kx * dt +
                               lambda * dsyy_dx / Kx * dt + mu / Kzh * dsxz_dz * dt + t6 +
                               mu / Kyh * dsxy_dy * dt + t8);
    A.vx[i] = v;
    if (px) P.sxx_x[i] = C.bh[1][x] * P.sxx_x[i] + byc * v * dt;
    if (pz) P.sxz_z[i] = C.b[0][z] * P.sxz_z[i] + byc * v * dt;
    if (py) P.sxy_y[i] = C.b[2][y] * P.sxy_y[i] + byc * v * dt;
  }
  // vz
  {
    const float t1 = apsi_f(az, P.vz_z, i, SZ, G.dz);
    const float dszz_dz = Af(A.szz, i, SZ, G.dz);
    const float dsxx_dz = Af(A.sxx, i, SZ, G.dz);
    const float dsyy_dz = Af(A.syy, i, SZ, G.dz);
    const float t8 = apsi_b(axh, P.vz_x, i, SX, G.dx);
    const float dsxz_dx = Ab(A.sxz, i, SX, G.dx);
    const float t6 = apsi_b(ayh, P.vz_y, i, 1, G.dy);
    const float dsyz_dy = Ab(A.syz, i, 1, G.dy);
    const float v = A.vz[i] + (t1 + l2m * dszz_dz / Kz * dt + lambda * dsxx_dz / Kz * dt +
                               lambda * dsyy_dz / Kz * dt + mu / Kyh * dsyz_dy * dt + t6 +
                               mu / Kxh * dsxz_dx * dt + t8);
    A.vz[i] = v;
    if (px) P.sxz_x[i] = C.b[1][x] * P.sxz_x[i] + byc * v * dt;
    if (pz) P.szz_z[i] = C.bh[0][z] * P.szz_z[i] + byc * v * dt;
    if (py) P.syz_y[i] = C.b[2][y] * P.syz_y[i] + byc * v * dt;
  }
  // vy
  {
    const float t1 = apsi_f(ay, P.vy_y, i, 1, G.dy);
    const float dsyy_dy = Af(A.syy, i, 1, G.dy);
    const float dsxx_dy = Af(A.sxx, i, 1, G.dy);
    const float dszz_dy = Af(A.szz, i, 1, G.dy);
    const float t6 = apsi_b(axh, P.vy_x, i, SX, G.dx);
    const float dsxy_dx = Ab(A.sxy, i, SX, G.dx);
    const float t8 = apsi_b(azh, P.vy_z, i, SZ, G.dz);
    const float dsyz_dz = Ab(A.syz, i, SZ, G.dz);
    const float v = A.vy[i] + (t1 + l2m * dsyy_dy / Ky * dt + lambda * dsxx_dy / Ky * dt +
                               lambda * dszz_dy / Ky * dt + mu / Kxh * dsxy_dx * dt + t6 +
                               mu / Kzh * dsyz_dz * dt + t8);
    A.vy[i] = v;
    if (px) P.sxy_x[i] = C.b[1][x] * P.sxy_x[i] + byc * v * dt;
    if (pz) P.syz_z[i] = (FORM == 1 ? C.b[0][z] : C.bh[0][z]) * P.syz_z[i] + byc * v * dt;
    if (py) P.syy_y[i] = (FORM == 1 ? C.bh[2][y] : C.b[2][y]) * P.syy_y[i] + byc * v * dt;
  }
}

// ---------------------------------------------------------------------------
// Adjoint stress of one cell, Src/stress_adj.cu:1-146.
// ---------------------------------------------------------------------------
__device__ __forceinline__ void stress_adj_cell(const Geom& G, const Wave& A, const Psi& P, const Model& M,
                                                const Cpml& C, const Cell& c) {
  const long i = c.i, SZ = G.pxy, SX = G.py;
  const int z = c.z, x = c.x, y = c.y, p = G.npml;
  const float lambda = M.lam[i], mu = M.mu[i], byc = M.byc[i], dt = G.dt;
  const bool pz = in_pml(z, G.nz, p), px = in_pml(x, G.nx, p), py = in_pml(y, G.ny, p);
  const float ax = C.a[1][x], az = C.a[0][z], ay = C.a[2][y];
  const float Kx = C.K[1][x], Kz = C.K[0][z], Ky = C.K[2][y];
  // sxz
  {
    const float t1 = apsi_f(ax, P.sxz_x, i, SX, G.dx);
    const float dvz_dx = Af(A.vz, i, SX, G.dx);
    const float t3 = apsi_f(az, P.sxz_z, i, SZ, G.dz);
    const float dvx_dz = Af(A.vx, i, SZ, G.dz);
    const float s = A.sxz[i] + (t1 + dvz_dx / Kx * byc * dt + t3 + dvx_dz / Kz * byc * dt);
    A.sxz[i] = s;
    if (px) P.vz_x[i] = C.bh[1][x] * P.vz_x[i] + s * mu * dt;
    if (pz) P.vx_z[i] = C.bh[0][z] * P.vx_z[i] + s * mu * dt;
  }
  // sxy
  {
    const float t1 = apsi_f(ax, P.sxy_x, i, SX, G.dx);
    const float dvy_dx = Af(A.vy, i, SX, G.dx);
    const float t3 = apsi_f(ay, P.sxy_y, i, 1, G.dy);
    const float dvx_dy = Af(A.vx, i, 1, G.dy);
    const float s = A.sxy[i] + (t1 + dvy_dx / Kx * byc * dt + t3 + dvx_dy / Ky * byc * dt);
    A.sxy[i] = s;
    if (py) P.vx_y[i] = C.bh[2][y] * P.vx_y[i] + s * mu * dt;
    if (px) P.vy_x[i] = C.bh[1][x] * P.vy_x[i] + s * mu * dt;
  }
  // syz
  {
    const float t1 = apsi_f(ay, P.syz_y, i, 1, G.dy);
    const float dvz_dy = Af(A.vz, i, 1, G.dy);
    const float t3 = apsi_f(az, P.syz_z, i, SZ, G.dz);
    const float dvy_dz = Af(A.vy, i, SZ, G.dz);
    const float s = A.syz[i] + (t1 + dvz_dy / Ky * byc * dt + t3 + dvy_dz / Kz * byc * dt);
    A.syz[i] = s;
    if (py) P.vz_y[i] = C.bh[2][y] * P.vz_y[i] + s * mu * dt;
    if (pz) P.vy_z[i] = C.bh[0][z] * P.vy_z[i] + s * mu * dt;
  }
  // normal stresses
  {
    const float txx = apsi_b(C.ah[1][x], P.sxx_x, i, SX, G.dx);
    const float dvx_dx = Ab(A.vx, i, SX, G.dx);
    const float tzz = apsi_b(C.ah[0][z], P.szz_z, i, SZ, G.dz);
    const float dvz_dz = Ab(A.vz, i, SZ, G.dz);
    const float tyy = apsi_b(C.ah[2][y], P.syy_y, i, 1, G.dy);
    const float dvy_dy = Ab(A.vy, i, 1, G.dy);
    const float sxx = A.sxx[i] + (txx + byc * dvx_dx / C.Kh[1][x] * dt);
    const float szz = A.szz[i] + (tzz + byc * dvz_dz / C.Kh[0][z] * dt);
    const float syy = A.syy[i] + (tyy + byc * dvy_dy / C.Kh[2][y] * dt);
    A.sxx[i] = sxx;
    A.szz[i] = szz;
    A.syy[i] = syy;
    const float l2m = lambda + 2.0f * mu;
    if (px) P.vx_x[i] = C.b[1][x] * P.vx_x[i] + lambda * szz * dt + l2m * sxx * dt + lambda * syy * dt;
    if (pz) P.vz_z[i] = C.b[0][z] * P.vz_z[i] + l2m * szz * dt + lambda * sxx * dt + lambda * syy * dt;
    if (py) P.vy_y[i] = C.b[2][y] * P.vy_y[i] + lambda * szz * dt + lambda * sxx * dt + l2m * syy * dt;
  }
}

// ---------------------------------------------------------------------------
// Kernels.
// ---------------------------------------------------------------------------
// Forward stress, then the source of this step at its cell (Src/libCUFD.cu:319-330).
__global__ __launch_bounds__(256) void k_stress_fwd(Sweep Sw, Geom G, Wave W, Psi P, Model M, Cpml C, Source S,
                                                    int it) {
  sweep(Sw, G, [&](const Cell& c) {
    stress_fwd_cell(G, W, P, M, C, c);
    if (c.i == S.cell) {
      const float a = 2250000.0f * S.stf[it] * G.dt * 1.0f;
      W.szz[c.i] += a; W.sxx[c.i] += a; W.syy[c.i] += a;
    }
  });
}

template <int FORM>
__global__ __launch_bounds__(256) void k_velocity_fwd(Sweep Sw, Geom G, Wave W, Psi P, Model M, Cpml C) {
  sweep(Sw, G, [&](const Cell& c) { velocity_fwd_cell<FORM>(G, W, P, M, C, c); });
}

// Backward velocity step: reverse velocity + density imaging (interior), then
// the adjoint velocity ([2, n-3]).
template <int FORM>
__global__ __launch_bounds__(256) void k_velocity_bwd(Sweep Sw, Geom G, Wave W, Wave A, Psi P, Model M, Cpml C,
                                                      Grad Gr) {
  sweep(Sw, G, [&](const Cell& c) {
    if (interior(G, c)) velocity_rev_cell<FORM>(G, W, A, M, Gr, c);
    velocity_adj_cell<FORM>(G, A, P, M, C, c);
  });
}

// Backward stress step: (source subtraction +) reverse stress + lambda/mu
// imaging (interior), then the adjoint stress ([2, n-3]).
__global__ __launch_bounds__(256) void k_stress_bwd(Sweep Sw, Geom G, Wave W, Wave A, Psi P, Model M, Cpml C,
                                                    Grad Gr, Source S, int it) {
  sweep(Sw, G, [&](const Cell& c) {
    if (interior(G, c)) stress_rev_cell(G, W, A, M, Gr, S, it, c);
    stress_adj_cell(G, A, P, M, C, c);
  });
}

template <int FORM>
__global__ __launch_bounds__(256) void k_velocity_adj(Sweep Sw, Geom G, Wave A, Psi P, Model M, Cpml C) {
  sweep(Sw, G, [&](const Cell& c) { velocity_adj_cell<FORM>(G, A, P, M, C, c); });
}

__global__ __launch_bounds__(256) void k_stress_adj(Sweep Sw, Geom G, Wave A, Psi P, Model M, Cpml C) {
  sweep(Sw, G, [&](const Cell& c) { stress_adj_cell(G, A, P, M, C, c); });
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

void launch_stress_fwd(Tile t, const Geom& G, const Wave& W, const Psi& P, const Model& M, const Cpml& C,
                       const Source& S, int it, hipStream_t s) {
  const Sweep Sw = sweep_full(G, t);
  hipLaunchKernelGGL(k_stress_fwd, grid_of(Sw), block_of(Sw), 0, s, Sw, G, W, P, M, C, S, it);
}
void launch_velocity_fwd(Tile t, int form, const Geom& G, const Wave& W, const Psi& P, const Model& M,
                         const Cpml& C, hipStream_t s) {
  const Sweep Sw = sweep_full(G, t);
  if (form == 1)
    hipLaunchKernelGGL(k_velocity_fwd<1>, grid_of(Sw), block_of(Sw), 0, s, Sw, G, W, P, M, C);
  else
    hipLaunchKernelGGL(k_velocity_fwd<0>, grid_of(Sw), block_of(Sw), 0, s, Sw, G, W, P, M, C);
}
void launch_velocity_bwd(Tile t, int form, const Geom& G, const Wave& W, const Wave& A, const Psi& P,
                         const Model& M, const Cpml& C, const Grad& Gr, hipStream_t s) {
  const Sweep Sw = sweep_full(G, t);
  if (form == 1)
    hipLaunchKernelGGL(k_velocity_bwd<1>, grid_of(Sw), block_of(Sw), 0, s, Sw, G, W, A, P, M, C, Gr);
  else
    hipLaunchKernelGGL(k_velocity_bwd<0>, grid_of(Sw), block_of(Sw), 0, s, Sw, G, W, A, P, M, C, Gr);
}
void launch_stress_bwd(Tile t, const Geom& G, const Wave& W, const Wave& A, const Psi& P, const Model& M,
                       const Cpml& C, const Grad& Gr, const Source& S, int it, hipStream_t s) {
  const Sweep Sw = sweep_full(G, t);
  hipLaunchKernelGGL(k_stress_bwd, grid_of(Sw), block_of(Sw), 0, s, Sw, G, W, A, P, M, C, Gr, S, it);
}
void launch_velocity_adj(int form, const Geom& G, const Wave& A, const Psi& P, const Model& M,
                         const Cpml& C, hipStream_t s) {
  const Sweep Sw = sweep_full(G, Tile{1, 1});
  if (form == 1)
    hipLaunchKernelGGL(k_velocity_adj<1>, grid_of(Sw), block_of(Sw), 0, s, Sw, G, A, P, M, C);
  else
    hipLaunchKernelGGL(k_velocity_adj<0>, grid_of(Sw), block_of(Sw), 0, s, Sw, G, A, P, M, C);
}
void launch_stress_adj(const Geom& G, const Wave& A, const Psi& P, const Model& M, const Cpml& C,
                       hipStream_t s) {
  const Sweep Sw = sweep_full(G, Tile{1, 1});
  hipLaunchKernelGGL(k_stress_adj, grid_of(Sw), block_of(Sw), 0, s, Sw, G, A, P, M, C);
}

}  // namespace efwi
