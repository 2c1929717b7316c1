// Staggered-grid elastic stencil kernels for gfx950 (forward, reverse
// reconstruction with imaging condition, adjoint).
//
// Every kernel evaluates, per cell, the same expression tree as the reference
// kernel it replaces, with IEEE division and no fused multiply-add (this file
// is compiled with -ffp-contract=off), so that the float results are the
// correctly rounded results of the reference's float expressions.
//
// Launch shape: 64 lanes along y (the contiguous axis, so every wavefront
// load is one coalesced row segment) x 4 rows along x per workgroup; one
// workgroup row per z plane. Neighbours along x and z come through the L1/L2
// caches.
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

// Map the launch (64 x 4 workgroup, grid (ny/64, nx/4, nz)) to a cell inside
// [lo, n-1-lo] on every axis.
__device__ __forceinline__ bool cell_of(const Geom& G, int lo_z, int hi_z, int lo_x, int hi_x,
                                        int lo_y, int hi_y, Cell& c) {
  c.y = blockIdx.x * 64 + threadIdx.x;
  c.x = blockIdx.y * 4 + threadIdx.y;
  c.z = blockIdx.z;
  if (c.z < lo_z || c.z > hi_z || c.x < lo_x || c.x > hi_x || c.y < lo_y || c.y > hi_y) return false;
  c.i = (long)c.z * G.pxy + (long)c.x * G.py + c.y;
  return true;
}

// ---------------------------------------------------------------------------
// Forward stress, Src/stress.cu:36-101.
// ---------------------------------------------------------------------------
template <int FORM>
__global__ __launch_bounds__(256) void k_stress_fwd(Geom G, Wave W, Psi P, Model M, Cpml C) {
  Cell c;
  if (!cell_of(G, 2, G.nz - 3, 2, G.nx - 3, 2, G.ny - 3, c)) return;
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
// Forward velocity, Src/velocity.cu:175-240.
// ---------------------------------------------------------------------------
template <int FORM>
__global__ __launch_bounds__(256) void k_velocity_fwd(Geom G, Wave W, Psi P, Model M, Cpml C) {
  Cell c;
  if (!cell_of(G, 2, G.nz - 3, 2, G.nx - 3, 2, G.ny - 3, c)) return;
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
// Reverse stress + lambda/mu imaging, Src/stress.cu:103-139. The source
// subtraction of the same step (Src/libCUFD.cu:547-549) is folded in: it
// precedes the update at the source cell, as in the reference's launch order.
// ---------------------------------------------------------------------------
__global__ __launch_bounds__(256) void k_stress_rev(Geom G, Wave W, Wave A, Model M, Grad Gr,
                                                    Source S, int it) {
  Cell c;
  const int p = G.npml;
  if (!cell_of(G, p, G.nz - 1 - p, p, G.nx - 1 - p, p, G.ny - 1 - p, c)) return;
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
// Reverse velocity + density imaging, Src/velocity.cu:242-275.
// ---------------------------------------------------------------------------
template <int FORM>
__global__ __launch_bounds__(256) void k_velocity_rev(Geom G, Wave W, Wave A, Model M, Grad Gr) {
  Cell c;
  const int p = G.npml;
  if (!cell_of(G, p, G.nz - 1 - p, p, G.nx - 1 - p, p, G.ny - 1 - p, c)) return;
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
// Adjoint velocity, Src/velocity_adj.cu:149-296. Reads the adjoint stress and
// the psi arrays written by the adjoint stress kernel; writes v-dagger and the
// psi arrays the adjoint stress kernel reads.
// ---------------------------------------------------------------------------
template <int FORM>
__global__ __launch_bounds__(256) void k_velocity_adj(Geom G, Wave A, Psi P, Model M, Cpml C) {
  Cell c;
  if (!cell_of(G, 2, G.nz - 3, 2, G.nx - 3, 2, G.ny - 3, c)) return;
  const long i = c.i, SZ = G.pxy, SX = G.py;
  const int z = c.z, x = c.x, y = c.y, p = G.npml;
  const float lambda = M.lam[i], mu = M.mu[i], byc = M.byc[i], dt = G.dt;
  const float l2m = lambda + 2.0f * mu;
  const bool pz = in_pml(z, G.nz, p), px = in_pml(x, G.nx, p), py = in_pml(y, G.ny, p);
  // vx
  {
    const float dpsixx_dx = Af(P.vx_x, i, SX, G.dx);
    const float dszz_dx = Af(A.szz, i, SX, G.dx);
    const float dsxx_dx = Af(A.sxx, i, SX, G.dx);
    const float dsyy_dx = Af(A.syy, i, SX, G.dx);
    const float dpsixz_dz = Ab(P.vx_z, i, SZ, G.dz);
    const float dsxz_dz = Ab(A.sxz, i, SZ, G.dz);
    const float dpsixy_dy = Ab(P.vx_y, i, 1, G.dy);
    const float dsxy_dy = Ab(A.sxy, i, 1, G.dy);
    const float Kx = C.K[1][x];
    const float v = A.vx[i] + (C.a[1][x] * dpsixx_dx + l2m * dsxx_dx / Kx * dt +
                               lambda * dszz_dx / Kx * dt + lambda * dsyy_dx / Kx * dt +
                               mu / C.Kh[0][z] * dsxz_dz * dt + C.ah[0][z] * dpsixz_dz +
                               mu / C.Kh[2][y] * dsxy_dy * dt + C.ah[2][y] * dpsixy_dy);
    A.vx[i] = v;
    if (px) P.sxx_x[i] = C.bh[1][x] * P.sxx_x[i] + byc * v * dt;
    if (pz) P.sxz_z[i] = C.b[0][z] * P.sxz_z[i] + byc * v * dt;
    if (py) P.sxy_y[i] = C.b[2][y] * P.sxy_y[i] + byc * v * dt;
  }
  // vz
  {
    const float dpsizz_dz = Af(P.vz_z, i, SZ, G.dz);
    const float dszz_dz = Af(A.szz, i, SZ, G.dz);
    const float dsxx_dz = Af(A.sxx, i, SZ, G.dz);
    const float dsyy_dz = Af(A.syy, i, SZ, G.dz);
    const float dpsizx_dx = Ab(P.vz_x, i, SX, G.dx);
    const float dsxz_dx = Ab(A.sxz, i, SX, G.dx);
    const float dpsizy_dy = Ab(P.vz_y, i, 1, G.dy);
    const float dsyz_dy = Ab(A.syz, i, 1, G.dy);
    const float Kz = C.K[0][z];
    const float v = A.vz[i] + (C.a[0][z] * dpsizz_dz + l2m * dszz_dz / Kz * dt +
                               lambda * dsxx_dz / Kz * dt + lambda * dsyy_dz / Kz * dt +
                               mu / C.Kh[2][y] * dsyz_dy * dt + C.ah[2][y] * dpsizy_dy +
                               mu / C.Kh[1][x] * dsxz_dx * dt + C.ah[1][x] * dpsizx_dx);
    A.vz[i] = v;
    if (px) P.sxz_x[i] = C.b[1][x] * P.sxz_x[i] + byc * v * dt;
    if (pz) P.szz_z[i] = C.bh[0][z] * P.szz_z[i] + byc * v * dt;
    if (py) P.syz_y[i] = C.b[2][y] * P.syz_y[i] + byc * v * dt;
  }
  // vy
  {
    const float dpsiyy_dy = Af(P.vy_y, i, 1, G.dy);
    const float dsyy_dy = Af(A.syy, i, 1, G.dy);
    const float dsxx_dy = Af(A.sxx, i, 1, G.dy);
    const float dszz_dy = Af(A.szz, i, 1, G.dy);
    const float dpsiyx_dx = Ab(P.vy_x, i, SX, G.dx);
    const float dsxy_dx = Ab(A.sxy, i, SX, G.dx);
    const float dpsiyz_dz = Ab(P.vy_z, i, SZ, G.dz);
    const float dsyz_dz = Ab(A.syz, i, SZ, G.dz);
    const float Ky = C.K[2][y];
    const float v = A.vy[i] + (C.a[2][y] * dpsiyy_dy + l2m * dsyy_dy / Ky * dt +
                               lambda * dsxx_dy / Ky * dt + lambda * dszz_dy / Ky * dt +
                               mu / C.Kh[1][x] * dsxy_dx * dt + C.ah[1][x] * dpsiyx_dx +
                               mu / C.Kh[0][z] * dsyz_dz * dt + C.ah[0][z] * dpsiyz_dz);
    A.vy[i] = v;
    if (px) P.sxy_x[i] = C.b[1][x] * P.sxy_x[i] + byc * v * dt;
    if (pz) P.syz_z[i] = (FORM == 1 ? C.b[0][z] : C.bh[0][z]) * P.syz_z[i] + byc * v * dt;
    if (py) P.syy_y[i] = (FORM == 1 ? C.bh[2][y] : C.b[2][y]) * P.syy_y[i] + byc * v * dt;
  }
}

// ---------------------------------------------------------------------------
// Adjoint stress, Src/stress_adj.cu:1-146.
// ---------------------------------------------------------------------------
__global__ __launch_bounds__(256) void k_stress_adj(Geom G, Wave A, Psi P, Model M, Cpml C) {
  Cell c;
  if (!cell_of(G, 2, G.nz - 3, 2, G.nx - 3, 2, G.ny - 3, c)) return;
  const long i = c.i, SZ = G.pxy, SX = G.py;
  const int z = c.z, x = c.x, y = c.y, p = G.npml;
  const float lambda = M.lam[i], mu = M.mu[i], byc = M.byc[i], dt = G.dt;
  const bool pz = in_pml(z, G.nz, p), px = in_pml(x, G.nx, p), py = in_pml(y, G.ny, p);
  // sxz
  {
    const float dphi_x = Af(P.sxz_x, i, SX, G.dx);
    const float dvz_dx = Af(A.vz, i, SX, G.dx);
    const float dphi_z = Af(P.sxz_z, i, SZ, G.dz);
    const float dvx_dz = Af(A.vx, i, SZ, G.dz);
    const float s = A.sxz[i] + (C.a[1][x] * dphi_x + dvz_dx / C.K[1][x] * byc * dt +
                                C.a[0][z] * dphi_z + dvx_dz / C.K[0][z] * byc * dt);
    A.sxz[i] = s;
    if (px) P.vz_x[i] = C.bh[1][x] * P.vz_x[i] + s * mu * dt;
    if (pz) P.vx_z[i] = C.bh[0][z] * P.vx_z[i] + s * mu * dt;
  }
  // sxy
  {
    const float dphi_x = Af(P.sxy_x, i, SX, G.dx);
    const float dvy_dx = Af(A.vy, i, SX, G.dx);
    const float dphi_y = Af(P.sxy_y, i, 1, G.dy);
    const float dvx_dy = Af(A.vx, i, 1, G.dy);
    const float s = A.sxy[i] + (C.a[1][x] * dphi_x + dvy_dx / C.K[1][x] * byc * dt +
                                C.a[2][y] * dphi_y + dvx_dy / C.K[2][y] * byc * dt);
    A.sxy[i] = s;
    if (py) P.vx_y[i] = C.bh[2][y] * P.vx_y[i] + s * mu * dt;
    if (px) P.vy_x[i] = C.bh[1][x] * P.vy_x[i] + s * mu * dt;
  }
  // syz
  {
    const float dphi_y = Af(P.syz_y, i, 1, G.dy);
    const float dvz_dy = Af(A.vz, i, 1, G.dy);
    const float dphi_z = Af(P.syz_z, i, SZ, G.dz);
    const float dvy_dz = Af(A.vy, i, SZ, G.dz);
    const float s = A.syz[i] + (C.a[2][y] * dphi_y + dvz_dy / C.K[2][y] * byc * dt +
                                C.a[0][z] * dphi_z + dvy_dz / C.K[0][z] * byc * dt);
    A.syz[i] = s;
    if (py) P.vz_y[i] = C.bh[2][y] * P.vz_y[i] + s * mu * dt;
    if (pz) P.vy_z[i] = C.bh[0][z] * P.vy_z[i] + s * mu * dt;
  }
  // normal stresses
  {
    const float dphi_xx = Ab(P.sxx_x, i, SX, G.dx);
    const float dvx_dx = Ab(A.vx, i, SX, G.dx);
    const float dphi_zz = Ab(P.szz_z, i, SZ, G.dz);
    const float dvz_dz = Ab(A.vz, i, SZ, G.dz);
    const float dphi_yy = Ab(P.syy_y, i, 1, G.dy);
    const float dvy_dy = Ab(A.vy, i, 1, G.dy);
    const float sxx = A.sxx[i] + (C.ah[1][x] * dphi_xx + byc * dvx_dx / C.Kh[1][x] * dt);
    const float szz = A.szz[i] + (C.ah[0][z] * dphi_zz + byc * dvz_dz / C.Kh[0][z] * dt);
    const float syy = A.syy[i] + (C.ah[2][y] * dphi_yy + byc * dvy_dy / C.Kh[2][y] * dt);
    A.sxx[i] = sxx;
    A.szz[i] = szz;
    A.syy[i] = syy;
    const float l2m = lambda + 2.0f * mu;
    if (px) P.vx_x[i] = C.b[1][x] * P.vx_x[i] + lambda * szz * dt + l2m * sxx * dt + lambda * syy * dt;
    if (pz) P.vz_z[i] = C.b[0][z] * P.vz_z[i] + l2m * szz * dt + lambda * sxx * dt + lambda * syy * dt;
    if (py) P.vy_y[i] = C.b[2][y] * P.vy_y[i] + lambda * szz * dt + lambda * sxx * dt + l2m * syy * dt;
  }
}

}  // namespace

// ---------------------------------------------------------------------------
// Launchers.
// ---------------------------------------------------------------------------
static inline dim3 stencil_grid(const Geom& G) {
  return dim3((G.ny + 63) / 64, (G.nx + 3) / 4, G.nz);
}
static const dim3 kStencilBlock(64, 4, 1);

void launch_stress_fwd(int form, const Geom& G, const Wave& W, const Psi& P, const Model& M,
                       const Cpml& C, hipStream_t s) {
  if (form == 1)
    hipLaunchKernelGGL(k_stress_fwd<1>, stencil_grid(G), kStencilBlock, 0, s, G, W, P, M, C);
  else
    hipLaunchKernelGGL(k_stress_fwd<0>, stencil_grid(G), kStencilBlock, 0, s, G, W, P, M, C);
}
void launch_velocity_fwd(int form, const Geom& G, const Wave& W, const Psi& P, const Model& M,
                         const Cpml& C, hipStream_t s) {
  if (form == 1)
    hipLaunchKernelGGL(k_velocity_fwd<1>, stencil_grid(G), kStencilBlock, 0, s, G, W, P, M, C);
  else
    hipLaunchKernelGGL(k_velocity_fwd<0>, stencil_grid(G), kStencilBlock, 0, s, G, W, P, M, C);
}
void launch_stress_rev(const Geom& G, const Wave& W, const Wave& A, const Model& M, const Grad& Gr,
                       const Source& S, int it, hipStream_t s) {
  hipLaunchKernelGGL(k_stress_rev, stencil_grid(G), kStencilBlock, 0, s, G, W, A, M, Gr, S, it);
}
void launch_velocity_rev(int form, const Geom& G, const Wave& W, const Wave& A, const Model& M,
                         const Grad& Gr, hipStream_t s) {
  if (form == 1)
    hipLaunchKernelGGL(k_velocity_rev<1>, stencil_grid(G), kStencilBlock, 0, s, G, W, A, M, Gr);
  else
    hipLaunchKernelGGL(k_velocity_rev<0>, stencil_grid(G), kStencilBlock, 0, s, G, W, A, M, Gr);
}
void launch_velocity_adj(int form, const Geom& G, const Wave& A, const Psi& P, const Model& M,
                         const Cpml& C, hipStream_t s) {
  if (form == 1)
    hipLaunchKernelGGL(k_velocity_adj<1>, stencil_grid(G), kStencilBlock, 0, s, G, A, P, M, C);
  else
    hipLaunchKernelGGL(k_velocity_adj<0>, stencil_grid(G), kStencilBlock, 0, s, G, A, P, M, C);
}
void launch_stress_adj(const Geom& G, const Wave& A, const Psi& P, const Model& M, const Cpml& C,
                       hipStream_t s) {
  hipLaunchKernelGGL(k_stress_adj, stencil_grid(G), kStencilBlock, 0, s, G, A, P, M, C);
}

}  // namespace efwi
