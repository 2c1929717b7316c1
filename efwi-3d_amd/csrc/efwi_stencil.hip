// Staggered-grid elastic stencil kernels for gfx950: forward, reverse
// reconstruction with imaging condition, and adjoint.
//
// Arithmetic: every cell evaluates the same expression tree as the reference
// kernel it replaces, in FP32 with IEEE division and no fused multiply-add
// (this file is compiled with -ffp-contract=off). Where the reference writes
// a double literal (2.0 * mu in the normal-stress updates and several adjoint
// terms: Src/stress.cu:57-59,112-114,132, Src/stress_adj.cu:131-140,
// Src/velocity_adj.cu:72,100,129), C++ promotes that sub-expression to FP64
// and rounds once at the store; here it stays FP32 (the stencil is FP32
// throughout), so those results can differ from the reference's by the
// rounding of the intermediate terms. Every other float result is the
// correctly rounded result of the reference's float expression. Two more
// liberties: a C-PML term a * d(psi) is skipped where the profile
// coefficient a is exactly zero, and a division by K where K is exactly one
// (the reference multiplies by zero or divides by one there); both can
// change the sign of an exact zero and nothing else.
//
// Addressing: the fields live in four groups (forward wavefield, adjoint
// wavefield, the 18 psi arrays, and lambda/mu/buoyancy/gradients), each one
// allocation with its fields fs bytes apart. Every group is one buffer
// resource: a field access is a buffer_load/store_dword with the field's
// byte offset in the scalar offset, the cell's (or a neighbour's) byte offset
// in a VGPR, and y neighbours folded into the instruction offset (the
// compiler merges a row's y neighbours into wider loads). No 64-bit address
// arithmetic is left per load. The x and z C-PML profile values are uniform
// across a wavefront (a wavefront is one 64-cell segment of one x row) and
// come through scalar loads.
//
// Work distribution ("sweep"): workgroups of 64 x R threads (R rows of 64
// y-consecutive cells), numbered b = 0, 1, ...; workgroup b runs on XCD
// b % 8 (round-robin dispatch), and the kernels map b so that each XCD owns a
// contiguous strip of x rows and walks it plane chunk by plane chunk in z.
// The x and z neighbours a cell reads were then fetched into that XCD's L2 by
// the workgroups just before it.
//
// Per-wavefront specialisation of the backward kernels: a wavefront whose
// row lies in the x or z C-PML runs only the adjoint part (the reverse part
// covers the interior only), and one whose row is interior along x and z
// drops every x/z C-PML term (there a = 0 and K = 1 exactly, on both grids);
// the y C-PML terms are compiled in only for wavefronts that touch a y slab.
//
// Fusion: the source injection is folded into the forward stress kernel; the
// backward pass runs two kernels per step, {reverse velocity + density
// imaging + adjoint velocity} and {source subtraction + reverse stress +
// lambda/mu imaging + adjoint stress}. Each cell reads its own adjoint field
// before updating it, as the reference's separate launches do, and no kernel
// reads a neighbour value that it writes. Each cell function issues all its
// loads before its first store.
//
// FORM = 0: reference formulation (typo in dsxy/dx, Src/velocity.cu:79;
//           off-by-one x/y C-PML predicate, Src/velocity.cu:46-90; swapped
//           adjoint C-PML coefficients, Src/velocity_adj.cu:139,142).
// FORM = 1: the same three places corrected.
#include <algorithm>

#include "efwi_common.h"
#include "efwi_kernels.h"

namespace efwi {
namespace {

// Optional occupancy targets (waves per SIMD) of the forward and backward
// kernels, which bound their VGPR budgets: a build-time tuning knob.
#ifdef EFWI_FWD_WAVES
#define EFWI_FWD_ATTR __attribute__((amdgpu_waves_per_eu(EFWI_FWD_WAVES)))
#else
#define EFWI_FWD_ATTR
#endif
#ifdef EFWI_BWD_WAVES
#define EFWI_BWD_ATTR __attribute__((amdgpu_waves_per_eu(EFWI_BWD_WAVES)))
#else
#define EFWI_BWD_ATTR
#endif

constexpr float C1 = (float)(9.0 / 8.0);   // Src/stress.cu:33
constexpr float C2 = (float)(1.0 / 24.0);  // Src/stress.cu:34

// Word 3 of a gfx9 buffer resource for raw dword access (32-bit data format).
constexpr int kRsrcWord3 = 0x00020000;

struct Grp {
  __amdgpu_buffer_rsrc_t r;
  unsigned fs;
};

__device__ __forceinline__ Grp group(float* base, int nfields, unsigned fs) {
  return Grp{__builtin_amdgcn_make_buffer_rsrc(base, (short)0, (int)((unsigned)nfields * fs), kRsrcWord3), fs};
}
__device__ __forceinline__ float ld(const Grp& g, int k, unsigned o) {
  return __builtin_bit_cast(float, __builtin_amdgcn_raw_buffer_load_b32(g.r, (int)o, (int)(k * g.fs), 0));
}
__device__ __forceinline__ void st(const Grp& g, int k, unsigned o, float v) {
  __builtin_amdgcn_raw_buffer_store_b32(__builtin_bit_cast(unsigned, v), g.r, (int)o, (int)(k * g.fs), 0);
}

// Byte offsets of the neighbours at -2, -1, +1, +2 along one axis.
struct Nb {
  unsigned m2, m1, p1, p2;
};

struct Cell {
  int z, x, y;  // z and x are wavefront-uniform
  unsigned o;   // byte offset of the cell in a field
  Nb zn, xn, yn;
};

// Field groups and axes of an accessor call.
enum GroupId { GW, GA, GP, GM };
enum Axis { AZ, AX, AY };

// Accessor of the flat sweep: every value of the cell or of a neighbour is a
// buffer load; offsets of the four neighbours per axis precomputed.
struct FlatAcc {
  Grp w, a, p, m;
  unsigned o;
  Nb n[3];
  __device__ __forceinline__ const Grp& g(int gi) const { return gi == GW ? w : gi == GA ? a : gi == GP ? p : m; }
  __device__ __forceinline__ float own(int gi, int k) const { return ld(g(gi), k, o); }
  // value at offset d (-2..2) along axis ax
  __device__ __forceinline__ float at(int gi, int k, int ax, int d) const {
    const Nb& b = n[ax];
    return ld(g(gi), k, d == 0 ? o : d == -2 ? b.m2 : d == -1 ? b.m1 : d == 1 ? b.p1 : b.p2);
  }
  __device__ __forceinline__ void st(int gi, int k, float v) const { efwi::st(g(gi), k, o, v); }
};

// Difference numerators of the 4th-order staggered operators along axis ax;
// the caller divides by the grid step. Backward / forward, Src/stress.cu:40-68.
template <class Acc>
__device__ __forceinline__ float n_b(const Acc& f, int g, int k, int ax) {
  return C1 * (f.at(g, k, ax, 0) - f.at(g, k, ax, -1)) - C2 * (f.at(g, k, ax, 1) - f.at(g, k, ax, -2));
}
template <class Acc>
__device__ __forceinline__ float n_f(const Acc& f, int g, int k, int ax) {
  return C1 * (f.at(g, k, ax, 1) - f.at(g, k, ax, 0)) - C2 * (f.at(g, k, ax, 2) - f.at(g, k, ax, -1));
}
// Adjoint forms as written in Src/stress_adj.cu:54-59.
template <class Acc>
__device__ __forceinline__ float n_af(const Acc& f, int g, int k, int ax) {
  return -C1 * (f.at(g, k, ax, 1) - f.at(g, k, ax, 0)) + C2 * (f.at(g, k, ax, 2) - f.at(g, k, ax, -1));
}
template <class Acc>
__device__ __forceinline__ float n_ab(const Acc& f, int g, int k, int ax) {
  return -C1 * (f.at(g, k, ax, 0) - f.at(g, k, ax, -1)) + C2 * (f.at(g, k, ax, 1) - f.at(g, k, ax, -2));
}
// dsxy/dx of the vy update (Src/velocity.cu:79 and :122).
template <int FORM, class Acc>
__device__ __forceinline__ float n_sxy_x(const Acc& f, int g, int k, int ax) {
  if (FORM == 1) return n_b(f, g, k, ax);
  return C1 * (f.at(g, k, ax, 0) - f.at(g, k, ax, -1)) - C2 * (f.at(g, k, ax, -1) - f.at(g, k, ax, -2));
}

// Quotients by the grid steps (IEEE division, as the reference's).
struct Dv {
  const Geom& G;
  __device__ __forceinline__ explicit Dv(const Geom& g) : G(g) {}
  __device__ __forceinline__ float z(float n) const { return n / G.dz; }
  __device__ __forceinline__ float x(float n) const { return n / G.dx; }
  __device__ __forceinline__ float y(float n) const { return n / G.dy; }
};

// C-PML profile value of axis `ax`, row `row`, index i.
__device__ __forceinline__ float pr(const Prof& C, int ax, int row, int i) {
  return C.p[C.off[ax] + row * C.n[ax] + i];
}

// a * d(psi)/dh of the adjoint kernels, skipped where a == 0 (the reference
// multiplies by zero there); ON = false where a is known to be zero.
template <bool ON, class Acc>
__device__ __forceinline__ float apsi_f(const Acc& f, float a, int k, int ax, float h) {
  if (!ON) return 0.0f;
  return a != 0.0f ? a * (n_af(f, GP, k, ax) / h) : 0.0f;
}
template <bool ON, class Acc>
__device__ __forceinline__ float apsi_b(const Acc& f, float a, int k, int ax, float h) {
  if (!ON) return 0.0f;
  return a != 0.0f ? a * (n_ab(f, GP, k, ax) / h) : 0.0f;
}
// v / K; ON = false where K is known to be one.
template <bool ON>
__device__ __forceinline__ float kdiv(float v, float K) {
  return ON ? v / K : v;
}

__device__ __forceinline__ bool in_pml(int i, int n, int p) { return i < p || i > n - p - 1; }
// x/y predicate of the forward velocity kernel.
template <int FORM>
__device__ __forceinline__ bool in_pml_v(int i, int n, int p) {
  if (FORM == 1) return in_pml(i, n, p);
  return i < p || i > n - p;
}

// psi <- b psi + a d (in place) ; returns d / K + psi  (Src/stress.cu:71-99)
__device__ __forceinline__ float cpml(float b, float a, float K, float& psi, float d) {
  psi = b * psi + a * d;
  return d / K + psi;
}

// Axis profile values used by one C-PML correction.
struct Pv {
  float b, a, K;
};
__device__ __forceinline__ Pv pv(const Prof& C, int ax, bool half, int i) {
  return half ? Pv{pr(C, ax, R_BH, i), pr(C, ax, R_AH, i), pr(C, ax, R_KH, i)}
              : Pv{pr(C, ax, R_B, i), pr(C, ax, R_A, i), pr(C, ax, R_K, i)};
}

// Workgroup b (64 x R threads: y lanes x rows) covers R rows of the strip of
// XCD b % 8 and Z consecutive planes; the strip is walked plane chunk by
// plane chunk (z slowest), so each XCD's L2 sees a compact, z-ordered region.
template <typename Body>
__device__ __forceinline__ void sweep_block(const Sweep& S, const Geom& G, int b, Body&& body) {
  const int xcd = b & 7, j = b >> 3;
  const int per_plane = S.rg_max * S.nyb;
  const int zc0 = j / per_plane;
  const int zc = S.zrev ? S.nzc - 1 - zc0 : zc0;
  const int r = j - zc0 * per_plane;
  const int rg = r / S.nyb;
  const int yb = r - rg * S.nyb;
  const int rows = S.rows_base + (xcd < S.rows_rem ? 1 : 0);
  // blockDim.x == 64: threadIdx.y is uniform across a wavefront
  const int row = __builtin_amdgcn_readfirstlane(rg * S.R + (int)threadIdx.y);
  if (row >= rows) return;
  Cell c;
  c.x = S.x0 + xcd * S.rows_base + min(xcd, S.rows_rem) + row;
  c.y = (S.yb0 + yb) * 64 + (int)threadIdx.x;
  if (c.y < S.y0 || c.y > S.y1) return;
  // planes of this wavefront's row segment that another launch covers
  const bool skip_col = c.x >= S.sk_x0 && c.x <= S.sk_x1 && S.yb0 + yb >= S.sk_yb0 && S.yb0 + yb <= S.sk_yb1;
  const int skz0 = skip_col ? S.sk_z0 : 1, skz1 = skip_col ? S.sk_z1 : 0;
  const unsigned sx = 4u * (unsigned)G.py, sz = 4u * (unsigned)G.pxy;
  const int za = S.z0 + zc * S.Z;
  const int zb = min(za + S.Z, S.z0 + S.nz);
  // byte offset of the cell's y-2 neighbour (y >= 2): the cell and its y
  // neighbours are this plus 0..16, which the compiler folds into the
  // instruction offsets (and merges into one dwordx4 + one dword per field)
  unsigned oy = 4u * ((unsigned)za * (unsigned)G.pxy + (unsigned)c.x * (unsigned)G.py + (unsigned)c.y) - 8u;
#pragma unroll 1
  for (int z = za; z < zb; z++) {
    if (z >= skz0 && z <= skz1) {
      oy += sz;
      continue;
    }
    const unsigned o = oy + 8u;
    c.z = z;
    c.o = o;
    c.zn = Nb{o - 2 * sz, o - sz, o + sz, o + 2 * sz};
    c.xn = Nb{o - 2 * sx, o - sx, o + sx, o + 2 * sx};
    c.yn = Nb{oy, oy + 4u, oy + 12u, oy + 16u};
    body(c);
    oy += sz;
  }
}
template <typename Body>
__device__ __forceinline__ void sweep(const Sweep& S, const Geom& G, Body&& body) {
  sweep_block(S, G, (int)blockIdx.x, body);
}

// ---------------------------------------------------------------------------
// Forward stress of one cell, Src/stress.cu:36-101, plus the source of this
// step at its cell (Src/libCUFD.cu:319-330).
// ---------------------------------------------------------------------------
template <class Acc>
__device__ __forceinline__ void stress_fwd_cell(const Geom& G, const Acc& acc, const Prof& C, const Cell& c,
                                                unsigned src_o, float amp) {
  const unsigned o = c.o;
  const int z = c.z, x = c.x, y = c.y, p = G.npml;
  const Dv D(G);
  float dvz_dz = D.z(n_b(acc, GW, F_VZ, AZ));
  float dvx_dx = D.x(n_b(acc, GW, F_VX, AX));
  float dvy_dy = D.y(n_b(acc, GW, F_VY, AY));
  float dvx_dz = D.z(n_f(acc, GW, F_VX, AZ));
  float dvx_dy = D.y(n_f(acc, GW, F_VX, AY));
  float dvz_dx = D.x(n_f(acc, GW, F_VZ, AX));
  float dvz_dy = D.y(n_f(acc, GW, F_VZ, AY));
  float dvy_dx = D.x(n_f(acc, GW, F_VY, AX));
  float dvy_dz = D.z(n_f(acc, GW, F_VY, AZ));
  const float lam = acc.own(GM, M_LAM), mu = acc.own(GM, M_MU), dt = G.dt;
  const float szz = acc.own(GW, F_SZZ), sxx = acc.own(GW, F_SXX), syy = acc.own(GW, F_SYY);
  const float sxz = acc.own(GW, F_SXZ), sxy = acc.own(GW, F_SXY), syz = acc.own(GW, F_SYZ);
  const bool pz = in_pml(z, G.nz, p), px = in_pml(x, G.nx, p), py = in_pml(y, G.ny, p);
  float qz0 = 0, qz1 = 0, qz2 = 0, qx0 = 0, qx1 = 0, qx2 = 0, qy0 = 0, qy1 = 0, qy2 = 0;
  if (pz) { qz0 = acc.own(GP, P_VZ_Z); qz1 = acc.own(GP, P_VX_Z); qz2 = acc.own(GP, P_VY_Z); }
  if (px) { qx0 = acc.own(GP, P_VX_X); qx1 = acc.own(GP, P_VZ_X); qx2 = acc.own(GP, P_VY_X); }
  if (py) { qy0 = acc.own(GP, P_VY_Y); qy1 = acc.own(GP, P_VZ_Y); qy2 = acc.own(GP, P_VX_Y); }
  if (pz) {
    const Pv i = pv(C, 0, false, z), h = pv(C, 0, true, z);
    dvz_dz = cpml(i.b, i.a, i.K, qz0, dvz_dz);
    dvx_dz = cpml(h.b, h.a, h.K, qz1, dvx_dz);
    dvy_dz = cpml(h.b, h.a, h.K, qz2, dvy_dz);
  }
  if (px) {
    const Pv i = pv(C, 1, false, x), h = pv(C, 1, true, x);
    dvx_dx = cpml(i.b, i.a, i.K, qx0, dvx_dx);
    dvz_dx = cpml(h.b, h.a, h.K, qx1, dvz_dx);
    dvy_dx = cpml(h.b, h.a, h.K, qx2, dvy_dx);
  }
  if (py) {
    const Pv i = pv(C, 2, false, y), h = pv(C, 2, true, y);
    dvy_dy = cpml(i.b, i.a, i.K, qy0, dvy_dy);
    dvz_dy = cpml(h.b, h.a, h.K, qy1, dvz_dy);
    dvx_dy = cpml(h.b, h.a, h.K, qy2, dvx_dy);
  }
  const float l2m = lam + 2.0f * mu;
  float nszz = szz + (l2m * dvz_dz + lam * dvx_dx + lam * dvy_dy) * dt;
  float nsxx = sxx + (lam * dvz_dz + l2m * dvx_dx + lam * dvy_dy) * dt;
  float nsyy = syy + (lam * dvz_dz + lam * dvx_dx + l2m * dvy_dy) * dt;
  if (o == src_o) {
    nszz = nszz + amp;
    nsxx = nsxx + amp;
    nsyy = nsyy + amp;
  }
  acc.st(GW, F_SZZ, nszz);
  acc.st(GW, F_SXX, nsxx);
  acc.st(GW, F_SYY, nsyy);
  acc.st(GW, F_SXZ, sxz + mu * (dvx_dz + dvz_dx) * dt);
  acc.st(GW, F_SXY, sxy + mu * (dvx_dy + dvy_dx) * dt);
  acc.st(GW, F_SYZ, syz + mu * (dvy_dz + dvz_dy) * dt);
  if (pz) { acc.st(GP, P_VZ_Z, qz0); acc.st(GP, P_VX_Z, qz1); acc.st(GP, P_VY_Z, qz2); }
  if (px) { acc.st(GP, P_VX_X, qx0); acc.st(GP, P_VZ_X, qx1); acc.st(GP, P_VY_X, qx2); }
  if (py) { acc.st(GP, P_VY_Y, qy0); acc.st(GP, P_VZ_Y, qy1); acc.st(GP, P_VX_Y, qy2); }
}

// ---------------------------------------------------------------------------
// Forward velocity of one cell, Src/velocity.cu:175-240.
// ---------------------------------------------------------------------------
template <int FORM, class Acc>
__device__ __forceinline__ void velocity_fwd_cell(const Geom& G, const Acc& acc, const Prof& C, const Cell& c) {
  const int z = c.z, x = c.x, y = c.y, p = G.npml;
  const Dv D(G);
  float dszz_dz = D.z(n_f(acc, GW, F_SZZ, AZ));
  float dsxz_dx = D.x(n_b(acc, GW, F_SXZ, AX));
  float dsyz_dy = D.y(n_b(acc, GW, F_SYZ, AY));
  float dsxz_dz = D.z(n_b(acc, GW, F_SXZ, AZ));
  float dsxx_dx = D.x(n_f(acc, GW, F_SXX, AX));
  float dsxy_dy = D.y(n_b(acc, GW, F_SXY, AY));
  float dsyz_dz = D.z(n_b(acc, GW, F_SYZ, AZ));
  float dsxy_dx = D.x(n_sxy_x<FORM>(acc, GW, F_SXY, AX));
  float dsyy_dy = D.y(n_f(acc, GW, F_SYY, AY));
  const float byc = acc.own(GM, M_BYC), dt = G.dt;
  const float vz = acc.own(GW, F_VZ), vx = acc.own(GW, F_VX), vy = acc.own(GW, F_VY);
  const bool pz = in_pml(z, G.nz, p);
  const bool px = in_pml_v<FORM>(x, G.nx, p), py = in_pml_v<FORM>(y, G.ny, p);
  float qz0 = 0, qz1 = 0, qz2 = 0, qx0 = 0, qx1 = 0, qx2 = 0, qy0 = 0, qy1 = 0, qy2 = 0;
  if (pz) { qz0 = acc.own(GP, P_SZZ_Z); qz1 = acc.own(GP, P_SXZ_Z); qz2 = acc.own(GP, P_SYZ_Z); }
  if (px) { qx0 = acc.own(GP, P_SXZ_X); qx1 = acc.own(GP, P_SXX_X); qx2 = acc.own(GP, P_SXY_X); }
  if (py) { qy0 = acc.own(GP, P_SYZ_Y); qy1 = acc.own(GP, P_SXY_Y); qy2 = acc.own(GP, P_SYY_Y); }
  if (pz) {
    const Pv i = pv(C, 0, false, z), h = pv(C, 0, true, z);
    dszz_dz = cpml(h.b, h.a, h.K, qz0, dszz_dz);
    dsxz_dz = cpml(i.b, i.a, i.K, qz1, dsxz_dz);
    dsyz_dz = cpml(i.b, i.a, i.K, qz2, dsyz_dz);
  }
  if (px) {
    const Pv i = pv(C, 1, false, x), h = pv(C, 1, true, x);
    dsxz_dx = cpml(i.b, i.a, i.K, qx0, dsxz_dx);
    dsxx_dx = cpml(h.b, h.a, h.K, qx1, dsxx_dx);
    dsxy_dx = cpml(i.b, i.a, i.K, qx2, dsxy_dx);
  }
  if (py) {
    const Pv i = pv(C, 2, false, y), h = pv(C, 2, true, y);
    dsyz_dy = cpml(i.b, i.a, i.K, qy0, dsyz_dy);
    dsxy_dy = cpml(i.b, i.a, i.K, qy1, dsxy_dy);
    dsyy_dy = cpml(h.b, h.a, h.K, qy2, dsyy_dy);
  }
  acc.st(GW, F_VZ, vz + (dszz_dz + dsxz_dx + dsyz_dy) * byc * dt);
  acc.st(GW, F_VX, vx + (dsxz_dz + dsxx_dx + dsxy_dy) * byc * dt);
  acc.st(GW, F_VY, vy + (dsyz_dz + dsxy_dx + dsyy_dy) * byc * dt);
  if (pz) { acc.st(GP, P_SZZ_Z, qz0); acc.st(GP, P_SXZ_Z, qz1); acc.st(GP, P_SYZ_Z, qz2); }
  if (px) { acc.st(GP, P_SXZ_X, qx0); acc.st(GP, P_SXX_X, qx1); acc.st(GP, P_SXY_X, qx2); }
  if (py) { acc.st(GP, P_SYZ_Y, qy0); acc.st(GP, P_SXY_Y, qy1); acc.st(GP, P_SYY_Y, qy2); }
}

// ---------------------------------------------------------------------------
// Backward velocity step of one cell: for interior cells the reverse velocity
// with the density imaging term (Src/velocity.cu:242-275), then for every cell
// of [2, n-3] the adjoint velocity (Src/velocity_adj.cu:149-296). The reverse
// part reads v-dagger before the adjoint part updates it, as the reference's
// launch order does. The adjoint part reads the adjoint stress and the psi
// arrays written by the adjoint stress step, and writes v-dagger and the psi
// arrays the adjoint stress step reads.
// RV: reverse part allowed; XZ: the row lies in the x or z C-PML; Y: the
// wavefront touches a y C-PML slab.
// ---------------------------------------------------------------------------
template <int FORM, bool RV, bool XZ, bool Y, class Acc>
__device__ __forceinline__ void velocity_bwd_cell(const Geom& G, const Acc& acc, const Prof& C, const Cell& c) {
  const int z = c.z, x = c.x, y = c.y, p = G.npml;
  const float dt = G.dt;
  const Dv D(G);
  const bool py = Y && in_pml(y, G.ny, p);
  const bool px = XZ && in_pml(x, G.nx, p), pz = XZ && in_pml(z, G.nz, p);
  const bool rev = RV && !XZ && !py;
  // ---- loads
  float r_sz = 0.0f, r_sx = 0.0f, r_sy = 0.0f, vz = 0.0f, vx = 0.0f, vy = 0.0f, grho = 0.0f;
  if (rev) {
    r_sz = D.z(n_f(acc, GW, F_SZZ, AZ)) + D.x(n_b(acc, GW, F_SXZ, AX)) + D.y(n_b(acc, GW, F_SYZ, AY));
    r_sx = D.z(n_b(acc, GW, F_SXZ, AZ)) + D.x(n_f(acc, GW, F_SXX, AX)) + D.y(n_b(acc, GW, F_SXY, AY));
    r_sy = D.z(n_b(acc, GW, F_SYZ, AZ)) + D.x(n_sxy_x<FORM>(acc, GW, F_SXY, AX)) +
           D.y(n_f(acc, GW, F_SYY, AY));
    vz = acc.own(GW, F_VZ); vx = acc.own(GW, F_VX); vy = acc.own(GW, F_VY);
    grho = acc.own(GM, M_GRHO);
  }
  const float lambda = acc.own(GM, M_LAM), mu = acc.own(GM, M_MU), byc = acc.own(GM, M_BYC);
  const float avz = acc.own(GA, F_VZ), avx = acc.own(GA, F_VX), avy = acc.own(GA, F_VY);
  const float ax = XZ ? pr(C, 1, R_A, x) : 0.0f, az = XZ ? pr(C, 0, R_A, z) : 0.0f;
  const float axh = XZ ? pr(C, 1, R_AH, x) : 0.0f, azh = XZ ? pr(C, 0, R_AH, z) : 0.0f;
  const float ay = Y ? pr(C, 2, R_A, y) : 0.0f, ayh = Y ? pr(C, 2, R_AH, y) : 0.0f;
  const float Kx = XZ ? pr(C, 1, R_K, x) : 1.0f;
  const float Kz = XZ ? pr(C, 0, R_K, z) : 1.0f;
  const float Kxh = XZ ? pr(C, 1, R_KH, x) : 1.0f;
  const float Kzh = XZ ? pr(C, 0, R_KH, z) : 1.0f;
  const float Ky = Y ? pr(C, 2, R_K, y) : 1.0f;
  const float Kyh = Y ? pr(C, 2, R_KH, y) : 1.0f;
  // vx-dagger terms
  const float x1 = apsi_f<XZ>(acc, ax, P_VX_X, AX, G.dx);
  const float x_szz = D.x(n_af(acc, GA, F_SZZ, AX));
  const float x_sxx = D.x(n_af(acc, GA, F_SXX, AX));
  const float x_syy = D.x(n_af(acc, GA, F_SYY, AX));
  const float x6 = apsi_b<XZ>(acc, azh, P_VX_Z, AZ, G.dz);
  const float x_sxz = D.z(n_ab(acc, GA, F_SXZ, AZ));
  const float x8 = apsi_b<Y>(acc, ayh, P_VX_Y, AY, G.dy);
  const float x_sxy = D.y(n_ab(acc, GA, F_SXY, AY));
  // vz-dagger terms
  const float z1 = apsi_f<XZ>(acc, az, P_VZ_Z, AZ, G.dz);
  const float z_szz = D.z(n_af(acc, GA, F_SZZ, AZ));
  const float z_sxx = D.z(n_af(acc, GA, F_SXX, AZ));
  const float z_syy = D.z(n_af(acc, GA, F_SYY, AZ));
  const float z8 = apsi_b<XZ>(acc, axh, P_VZ_X, AX, G.dx);
  const float z_sxz = D.x(n_ab(acc, GA, F_SXZ, AX));
  const float z6 = apsi_b<Y>(acc, ayh, P_VZ_Y, AY, G.dy);
  const float z_syz = D.y(n_ab(acc, GA, F_SYZ, AY));
  // vy-dagger terms
  const float y1 = apsi_f<Y>(acc, ay, P_VY_Y, AY, G.dy);
  const float y_syy = D.y(n_af(acc, GA, F_SYY, AY));
  const float y_sxx = D.y(n_af(acc, GA, F_SXX, AY));
  const float y_szz = D.y(n_af(acc, GA, F_SZZ, AY));
  const float y6 = apsi_b<XZ>(acc, axh, P_VY_X, AX, G.dx);
  const float y_sxy = D.x(n_ab(acc, GA, F_SXY, AX));
  const float y8 = apsi_b<XZ>(acc, azh, P_VY_Z, AZ, G.dz);
  const float y_syz = D.z(n_ab(acc, GA, F_SYZ, AZ));
  // psi values updated by this step
  float qx0 = 0, qx1 = 0, qx2 = 0, qz0 = 0, qz1 = 0, qz2 = 0, qy0 = 0, qy1 = 0, qy2 = 0;
  if (px) { qx0 = acc.own(GP, P_SXX_X); qx1 = acc.own(GP, P_SXZ_X); qx2 = acc.own(GP, P_SXY_X); }
  if (pz) { qz0 = acc.own(GP, P_SXZ_Z); qz1 = acc.own(GP, P_SZZ_Z); qz2 = acc.own(GP, P_SYZ_Z); }
  if (py) { qy0 = acc.own(GP, P_SXY_Y); qy1 = acc.own(GP, P_SYZ_Y); qy2 = acc.own(GP, P_SYY_Y); }
  // ---- arithmetic
  const float l2m = lambda + 2.0f * mu;
  const float nvx = avx + (x1 + kdiv<XZ>(l2m * x_sxx, Kx) * dt + kdiv<XZ>(lambda * x_szz, Kx) * dt +
                           kdiv<XZ>(lambda * x_syy, Kx) * dt + kdiv<XZ>(mu, Kzh) * x_sxz * dt + x6 +
                           kdiv<Y>(mu, Kyh) * x_sxy * dt + x8);
  const float nvz = avz + (z1 + kdiv<XZ>(l2m * z_szz, Kz) * dt + kdiv<XZ>(lambda * z_sxx, Kz) * dt +
                           kdiv<XZ>(lambda * z_syy, Kz) * dt + kdiv<Y>(mu, Kyh) * z_syz * dt + z6 +
                           kdiv<XZ>(mu, Kxh) * z_sxz * dt + z8);
  const float nvy = avy + (y1 + kdiv<Y>(l2m * y_syy, Ky) * dt + kdiv<Y>(lambda * y_sxx, Ky) * dt +
                           kdiv<Y>(lambda * y_szz, Ky) * dt + kdiv<XZ>(mu, Kxh) * y_sxy * dt + y6 +
                           kdiv<XZ>(mu, Kzh) * y_syz * dt + y8);
  // ---- stores
  if (rev) {
    acc.st(GW, F_VZ, vz - r_sz * byc * dt);
    acc.st(GW, F_VX, vx - r_sx * byc * dt);
    acc.st(GW, F_VY, vy - r_sy * byc * dt);
    const float nb2 = -(byc * byc);
    acc.st(GM, M_GRHO, grho + -(avz * r_sz * dt * nb2 + avx * r_sx * dt * nb2 + avy * r_sy * dt * nb2));
  }
  acc.st(GA, F_VX, nvx);
  acc.st(GA, F_VZ, nvz);
  acc.st(GA, F_VY, nvy);
  if (px) {
    const float b = pr(C, 1, R_B, x), bh = pr(C, 1, R_BH, x);
    acc.st(GP, P_SXX_X, bh * qx0 + byc * nvx * dt);
    acc.st(GP, P_SXZ_X, b * qx1 + byc * nvz * dt);
    acc.st(GP, P_SXY_X, b * qx2 + byc * nvy * dt);
  }
  if (pz) {
    const float b = pr(C, 0, R_B, z), bh = pr(C, 0, R_BH, z);
    acc.st(GP, P_SXZ_Z, b * qz0 + byc * nvx * dt);
    acc.st(GP, P_SZZ_Z, bh * qz1 + byc * nvz * dt);
    // Src/velocity_adj.cu:285 decays psi_dsyz_dz with b_z_half (the forward uses b_z)
    acc.st(GP, P_SYZ_Z, (FORM == 1 ? b : bh) * qz2 + byc * nvy * dt);
  }
  if (py) {
    const float b = pr(C, 2, R_B, y), bh = pr(C, 2, R_BH, y);
    acc.st(GP, P_SXY_Y, b * qy0 + byc * nvx * dt);
    acc.st(GP, P_SYZ_Y, b * qy1 + byc * nvz * dt);
    // Src/velocity_adj.cu:288 decays psi_dsyy_dy with b_y (the forward uses b_y_half)
    acc.st(GP, P_SYY_Y, (FORM == 1 ? bh : b) * qy2 + byc * nvy * dt);
  }
}

// ---------------------------------------------------------------------------
// Backward stress step of one cell: for interior cells the source subtraction
// at the source cell (Src/libCUFD.cu:547-549), the reverse stress and the
// lambda/mu imaging terms (Src/stress.cu:103-139); then for every cell of
// [2, n-3] the adjoint stress (Src/stress_adj.cu:1-146). The imaging terms
// read the adjoint stress before the adjoint part updates it.
// ---------------------------------------------------------------------------
template <bool RV, bool XZ, bool Y, class Acc>
__device__ __forceinline__ void stress_bwd_cell(const Geom& G, const Acc& acc, const Prof& C, const Cell& c,
                                                unsigned src_o, float amp) {
  const unsigned o = c.o;
  const int z = c.z, x = c.x, y = c.y, p = G.npml;
  const float dt = G.dt;
  const Dv D(G);
  const bool py = Y && in_pml(y, G.ny, p);
  const bool px = XZ && in_pml(x, G.nx, p), pz = XZ && in_pml(z, G.nz, p);
  const bool rev = RV && !XZ && !py;
  const float lambda = acc.own(GM, M_LAM), mu = acc.own(GM, M_MU), byc = acc.own(GM, M_BYC);
  // ---- loads of the reverse part
  float dvz_dz = 0, dvx_dx = 0, dvy_dy = 0, dvx_dz = 0, dvx_dy = 0, dvz_dx = 0, dvz_dy = 0, dvy_dx = 0, dvy_dz = 0;
  float szz = 0, sxx = 0, syy = 0, sxz = 0, sxy = 0, syz = 0, glam = 0, gmu = 0;
  if (rev) {
    dvz_dz = D.z(n_b(acc, GW, F_VZ, AZ));
    dvx_dx = D.x(n_b(acc, GW, F_VX, AX));
    dvy_dy = D.y(n_b(acc, GW, F_VY, AY));
    dvx_dz = D.z(n_f(acc, GW, F_VX, AZ));
    dvx_dy = D.y(n_f(acc, GW, F_VX, AY));
    dvz_dx = D.x(n_f(acc, GW, F_VZ, AX));
    dvz_dy = D.y(n_f(acc, GW, F_VZ, AY));
    dvy_dx = D.x(n_f(acc, GW, F_VY, AX));
    dvy_dz = D.z(n_f(acc, GW, F_VY, AZ));
    szz = acc.own(GW, F_SZZ); sxx = acc.own(GW, F_SXX); syy = acc.own(GW, F_SYY);
    sxz = acc.own(GW, F_SXZ); sxy = acc.own(GW, F_SXY); syz = acc.own(GW, F_SYZ);
    glam = acc.own(GM, M_GLAM); gmu = acc.own(GM, M_GMU);
  }
  // ---- loads of the adjoint part
  const float aszz = acc.own(GA, F_SZZ), asxx = acc.own(GA, F_SXX), asyy = acc.own(GA, F_SYY);
  const float asxz = acc.own(GA, F_SXZ), asxy = acc.own(GA, F_SXY), asyz = acc.own(GA, F_SYZ);
  const float ax = XZ ? pr(C, 1, R_A, x) : 0.0f, az = XZ ? pr(C, 0, R_A, z) : 0.0f;
  const float axh = XZ ? pr(C, 1, R_AH, x) : 0.0f, azh = XZ ? pr(C, 0, R_AH, z) : 0.0f;
  const float ay = Y ? pr(C, 2, R_A, y) : 0.0f, ayh = Y ? pr(C, 2, R_AH, y) : 0.0f;
  const float Kx = XZ ? pr(C, 1, R_K, x) : 1.0f;
  const float Kz = XZ ? pr(C, 0, R_K, z) : 1.0f;
  const float Ky = Y ? pr(C, 2, R_K, y) : 1.0f;
  const float Kxh = XZ ? pr(C, 1, R_KH, x) : 1.0f;
  const float Kzh = XZ ? pr(C, 0, R_KH, z) : 1.0f;
  const float Kyh = Y ? pr(C, 2, R_KH, y) : 1.0f;
  const float xz1 = apsi_f<XZ>(acc, ax, P_SXZ_X, AX, G.dx);
  const float xz_vz = D.x(n_af(acc, GA, F_VZ, AX));
  const float xz3 = apsi_f<XZ>(acc, az, P_SXZ_Z, AZ, G.dz);
  const float xz_vx = D.z(n_af(acc, GA, F_VX, AZ));
  const float xy1 = apsi_f<XZ>(acc, ax, P_SXY_X, AX, G.dx);
  const float xy_vy = D.x(n_af(acc, GA, F_VY, AX));
  const float xy3 = apsi_f<Y>(acc, ay, P_SXY_Y, AY, G.dy);
  const float xy_vx = D.y(n_af(acc, GA, F_VX, AY));
  const float yz1 = apsi_f<Y>(acc, ay, P_SYZ_Y, AY, G.dy);
  const float yz_vz = D.y(n_af(acc, GA, F_VZ, AY));
  const float yz3 = apsi_f<XZ>(acc, az, P_SYZ_Z, AZ, G.dz);
  const float yz_vy = D.z(n_af(acc, GA, F_VY, AZ));
  const float txx = apsi_b<XZ>(acc, axh, P_SXX_X, AX, G.dx);
  const float xx_vx = D.x(n_ab(acc, GA, F_VX, AX));
  const float tzz = apsi_b<XZ>(acc, azh, P_SZZ_Z, AZ, G.dz);
  const float zz_vz = D.z(n_ab(acc, GA, F_VZ, AZ));
  const float tyy = apsi_b<Y>(acc, ayh, P_SYY_Y, AY, G.dy);
  const float yy_vy = D.y(n_ab(acc, GA, F_VY, AY));
  // psi values updated by this step
  float qx0 = 0, qx1 = 0, qx2 = 0, qz0 = 0, qz1 = 0, qz2 = 0, qy0 = 0, qy1 = 0, qy2 = 0;
  if (px) { qx0 = acc.own(GP, P_VZ_X); qx1 = acc.own(GP, P_VY_X); qx2 = acc.own(GP, P_VX_X); }
  if (pz) { qz0 = acc.own(GP, P_VX_Z); qz1 = acc.own(GP, P_VY_Z); qz2 = acc.own(GP, P_VZ_Z); }
  if (py) { qy0 = acc.own(GP, P_VX_Y); qy1 = acc.own(GP, P_VZ_Y); qy2 = acc.own(GP, P_VY_Y); }
  // ---- arithmetic
  const float l2m = lambda + 2.0f * mu;
  const float nsxz =
      asxz + (xz1 + kdiv<XZ>(xz_vz, Kx) * byc * dt + xz3 + kdiv<XZ>(xz_vx, Kz) * byc * dt);
  const float nsxy =
      asxy + (xy1 + kdiv<XZ>(xy_vy, Kx) * byc * dt + xy3 + kdiv<Y>(xy_vx, Ky) * byc * dt);
  const float nsyz =
      asyz + (yz1 + kdiv<Y>(yz_vz, Ky) * byc * dt + yz3 + kdiv<XZ>(yz_vy, Kz) * byc * dt);
  const float nsxx = asxx + (txx + kdiv<XZ>(byc * xx_vx, Kxh) * dt);
  const float nszz = aszz + (tzz + kdiv<XZ>(byc * zz_vz, Kzh) * dt);
  const float nsyy = asyy + (tyy + kdiv<Y>(byc * yy_vy, Kyh) * dt);
  // ---- stores
  if (rev) {
    if (o == src_o) {
      szz -= amp; sxx -= amp; syy -= amp;
    }
    acc.st(GW, F_SZZ, szz - (l2m * dvz_dz + lambda * dvx_dx + lambda * dvy_dy) * dt);
    acc.st(GW, F_SXX, sxx - (lambda * dvz_dz + l2m * dvx_dx + lambda * dvy_dy) * dt);
    acc.st(GW, F_SYY, syy - (lambda * dvz_dz + lambda * dvx_dx + l2m * dvy_dy) * dt);
    acc.st(GW, F_SXZ, sxz - mu * (dvx_dz + dvz_dx) * dt);
    acc.st(GW, F_SXY, sxy - mu * (dvx_dy + dvy_dx) * dt);
    acc.st(GW, F_SYZ, syz - mu * (dvy_dz + dvz_dy) * dt);
    acc.st(GM, M_GLAM, glam + -(aszz + asxx + asyy) * (dvz_dz + dvx_dx + dvy_dy) * dt);
    acc.st(GM, M_GMU, gmu + -(2.0f * aszz * dvz_dz + 2.0f * asxx * dvx_dx + 2.0f * asyy * dvy_dy +
                            asxz * (dvx_dz + dvz_dx) + asyz * (dvy_dz + dvz_dy) + asxy * (dvy_dx + dvx_dy)) * dt);
  }
  acc.st(GA, F_SXZ, nsxz);
  acc.st(GA, F_SXY, nsxy);
  acc.st(GA, F_SYZ, nsyz);
  acc.st(GA, F_SXX, nsxx);
  acc.st(GA, F_SZZ, nszz);
  acc.st(GA, F_SYY, nsyy);
  if (px) {
    const float b = pr(C, 1, R_B, x), bh = pr(C, 1, R_BH, x);
    acc.st(GP, P_VZ_X, bh * qx0 + nsxz * mu * dt);
    acc.st(GP, P_VY_X, bh * qx1 + nsxy * mu * dt);
    acc.st(GP, P_VX_X, b * qx2 + lambda * nszz * dt + l2m * nsxx * dt + lambda * nsyy * dt);
  }
  if (pz) {
    const float b = pr(C, 0, R_B, z), bh = pr(C, 0, R_BH, z);
    acc.st(GP, P_VX_Z, bh * qz0 + nsxz * mu * dt);
    acc.st(GP, P_VY_Z, bh * qz1 + nsyz * mu * dt);
    acc.st(GP, P_VZ_Z, b * qz2 + l2m * nszz * dt + lambda * nsxx * dt + lambda * nsyy * dt);
  }
  if (py) {
    const float b = pr(C, 2, R_B, y), bh = pr(C, 2, R_BH, y);
    acc.st(GP, P_VX_Y, bh * qy0 + nsxy * mu * dt);
    acc.st(GP, P_VZ_Y, bh * qy1 + nsyz * mu * dt);
    acc.st(GP, P_VY_Y, b * qy2 + lambda * nszz * dt + lambda * nsxx * dt + l2m * nsyy * dt);
  }
}

// Per-wavefront dispatch of the backward cells on (XZ, Y).
struct Flags {
  bool xz, y;
};
__device__ __forceinline__ Flags flags_of(const Geom& G, const Cell& c) {
  const int p = G.npml;
  return Flags{in_pml(c.x, G.nx, p) || in_pml(c.z, G.nz, p), __any(in_pml(c.y, G.ny, p)) != 0};
}

// Source amplitude of step it at the source cell, 1500^2 * stf[it] * dt
// (Src/utilities.cu:442-479; the 9^3 amplitude cube is exactly 1 at its
// centre and 0 elsewhere), and the cell's byte offset (a value no cell has
// when there is no source).
__device__ __forceinline__ float source_amp(const Source& S, const Geom& G, int it) {
  return S.cell >= 0 ? 2250000.0f * S.stf[it] * G.dt * 1.0f : 0.0f;
}
__device__ __forceinline__ unsigned source_off(const Source& S) {
  return S.cell >= 0 ? 4u * (unsigned)S.cell : 0xFFFFFFFFu;
}

// ---------------------------------------------------------------------------
// Kernels.
// ---------------------------------------------------------------------------
__device__ __forceinline__ FlatAcc flat_acc(const Geom& G, const Groups& F, const Cell& c) {
  return FlatAcc{group(F.w, F_WAVE_COUNT, G.fs), group(F.a, F_WAVE_COUNT, G.fs), group(F.p, P_PSI_COUNT, G.fs),
                 group(F.m, M_COUNT, G.fs), c.o, {c.zn, c.xn, c.yn}};
}

// Index of cell (z, x, y) in one field of a ring slot, or -1 outside the band:
// the inverse of the band's enumeration (z slabs, then x slabs, then y slabs,
// each row-major in (z, x, y); Src/utilities.cu:323-420 store the same cells).
__device__ __forceinline__ int band_index(const Band& B, int z, int x, int y) {
  if (z < B.lo[0] || z > B.hi[0] || x < B.lo[1] || x > B.hi[1] || y < B.lo[2] || y > B.hi[2]) return -1;
  const int nxB = B.hi[1] - B.lo[1] + 1, nyB = B.hi[2] - B.lo[2] + 1;
  const int nzM = B.hi[0] - B.lo[0] - 9, nxM = nxB - 10;
  if (z <= B.lo[0] + 4 || z >= B.hi[0] - 4) {
    const int zl = z >= B.hi[0] - 4 ? 5 + z - (B.hi[0] - 4) : z - B.lo[0];
    return (zl * nxB + (x - B.lo[1])) * nyB + (y - B.lo[2]);
  }
  const int zm = z - B.lo[0] - 5;
  if (x <= B.lo[1] + 4 || x >= B.hi[1] - 4) {
    const int sl = x >= B.hi[1] - 4 ? 1 : 0;
    const int xl = sl ? x - (B.hi[1] - 4) : x - B.lo[1];
    return (int)B.nzs + ((sl * nzM + zm) * 5 + xl) * nyB + (y - B.lo[2]);
  }
  if (y <= B.lo[2] + 4 || y >= B.hi[2] - 4) {
    const int sl = y >= B.hi[2] - 4 ? 1 : 0;
    const int yl = sl ? y - (B.hi[2] - 4) : y - B.lo[2];
    return (int)(B.nzs + B.nxs) + ((sl * nzM + zm) * nxM + (x - B.lo[1] - 5)) * 5 + yl;
  }
  return -1;
}
__device__ __forceinline__ int ring_index(const RingFold& R, const Cell& c) {
  return R.slot ? band_index(R.B, c.z, c.x, c.y) : -1;
}

// Accessor of the forward kernels: a band cell's new wavefield values also go
// to the ring slot (rk: the cell's index in a slot's field, or -1), so the
// boundary save needs no pass of its own. The ring's field order is the
// WaveField order.
struct SaveAcc : FlatAcc {
  float* rs;
  long rn;
  int rk;
  __device__ __forceinline__ void st(int gi, int k, float v) const {
    FlatAcc::st(gi, k, v);
    if (gi == GW && rk >= 0) rs[(long)k * rn + rk] = v;
  }
};
__device__ __forceinline__ SaveAcc save_acc(const Geom& G, const Groups& F, const RingFold& R, const Cell& c) {
  return SaveAcc{flat_acc(G, F, c), R.slot, R.B.n, ring_index(R, c)};
}

// Restore folded into a backward kernel: a band cell loads the slot's values
// of wavefield fields K0..K1-1 before its update (so the loads overlap the
// cell's own) and stores them after it, as the reverse reconstruction's
// restore after that kernel would leave them.
template <int K0, int K1>
struct RingRestore {
  int rk;
  float v[K1 - K0];
  __device__ __forceinline__ RingRestore(const RingFold& R, const Cell& c) : rk(ring_index(R, c)) {
    if (rk >= 0) {
#pragma unroll
      for (int k = K0; k < K1; k++) v[k - K0] = R.slot[(long)k * R.B.n + rk];
    }
  }
  __device__ __forceinline__ void store(const FlatAcc& acc) const {
    if (rk >= 0) {
#pragma unroll
      for (int k = K0; k < K1; k++) acc.st(GW, k, v[k - K0]);
    }
  }
};

__global__ __launch_bounds__(256) EFWI_FWD_ATTR void k_stress_fwd(Sweep Sw, Geom G, Groups F, Prof C, Source S, int it,
                                                                  RingFold Rf) {
  const unsigned so = source_off(S);
  const float amp = source_amp(S, G, it);
  sweep(Sw, G, [&](const Cell& c) { stress_fwd_cell(G, save_acc(G, F, Rf, c), C, c, so, amp); });
}

template <int FORM>
__global__ __launch_bounds__(256) EFWI_FWD_ATTR void k_velocity_fwd(Sweep Sw, Geom G, Groups F, Prof C, RingFold Rf) {
  sweep(Sw, G, [&](const Cell& c) { velocity_fwd_cell<FORM>(G, save_acc(G, F, Rf, c), C, c); });
}

// Backward velocity step (RV = true): reverse velocity + density imaging
// (interior), then the adjoint velocity ([2, n-3]); RV = false: adjoint only.
template <int FORM, bool RV>
__global__ __launch_bounds__(256) EFWI_BWD_ATTR void k_velocity_bwd(Sweep Sw, Geom G, Groups F, Prof C, RingFold Rf) {
  sweep(Sw, G, [&](const Cell& c) {
    const FlatAcc acc = flat_acc(G, F, c);
    const Flags f = flags_of(G, c);
    const RingRestore<F_VZ, F_VY + 1> rr(RV ? Rf : RingFold{}, c);
    if (f.xz) {
      if (f.y) velocity_bwd_cell<FORM, RV, true, true>(G, acc, C, c);
      else velocity_bwd_cell<FORM, RV, true, false>(G, acc, C, c);
    } else {
      if (f.y) velocity_bwd_cell<FORM, RV, false, true>(G, acc, C, c);
      else velocity_bwd_cell<FORM, RV, false, false>(G, acc, C, c);
    }
    if (RV) rr.store(acc);
  });
}

// Backward stress step (RV = true): (source subtraction +) reverse stress +
// lambda/mu imaging (interior), then the adjoint stress ([2, n-3]);
// RV = false: adjoint only.
template <bool RV>
__global__ __launch_bounds__(256) EFWI_BWD_ATTR void k_stress_bwd(Sweep Sw, Geom G, Groups F, Prof C, Source S, int it,
                                                                  RingFold Rf) {
  const unsigned so = RV ? source_off(S) : 0xFFFFFFFFu;
  const float amp = RV ? source_amp(S, G, it) : 0.0f;
  sweep(Sw, G, [&](const Cell& c) {
    const FlatAcc acc = flat_acc(G, F, c);
    const Flags f = flags_of(G, c);
    const RingRestore<F_SZZ, F_WAVE_COUNT> rr(RV ? Rf : RingFold{}, c);
    if (f.xz) {
      if (f.y) stress_bwd_cell<RV, true, true>(G, acc, C, c, so, amp);
      else stress_bwd_cell<RV, true, false>(G, acc, C, c, so, amp);
    } else {
      if (f.y) stress_bwd_cell<RV, false, true>(G, acc, C, c, so, amp);
      else stress_bwd_cell<RV, false, false>(G, acc, C, c, so, amp);
    }
    if (RV) rr.store(acc);
  });
}

// The same two backward steps for wavefronts clear of every C-PML slab only
// (the interior launch of the split): one variant, no C-PML terms.
template <int FORM>
__global__ __launch_bounds__(256) void k_velocity_bwd_in(Sweep Sw, Geom G, Groups F, Prof C, RingFold Rf) {
  sweep(Sw, G, [&](const Cell& c) {
    const FlatAcc acc = flat_acc(G, F, c);
    const RingRestore<F_VZ, F_VY + 1> rr(Rf, c);
    velocity_bwd_cell<FORM, true, false, false>(G, acc, C, c);
    rr.store(acc);
  });
}
__global__ __launch_bounds__(256) void k_stress_bwd_in(Sweep Sw, Geom G, Groups F, Prof C, Source S, int it,
                                                       RingFold Rf) {
  const unsigned so = source_off(S);
  const float amp = source_amp(S, G, it);
  sweep(Sw, G, [&](const Cell& c) {
    const FlatAcc acc = flat_acc(G, F, c);
    const RingRestore<F_SZZ, F_WAVE_COUNT> rr(Rf, c);
    stress_bwd_cell<true, false, false>(G, acc, C, c, so, amp);
    rr.store(acc);
  });
}

// ---------------------------------------------------------------------------
// z-marching sweep. A workgroup of 64 x R threads owns R rows x 64 y-cells and
// walks a chunk of planes in z. Fields with z derivatives keep their column's
// planes z-2..z+2 in registers (a queue refilled one plane ahead), fields with
// x or y derivatives are staged per plane in an LDS tile of (R+4) x 68 values
// (the tile plus a 2-cell halo), and every other value (own-cell fields, psi
// in the C-PML) is a buffer load as in the flat sweep. Each field value then
// crosses the memory pipeline about once per cell instead of once per
// neighbour. The cell functions are the flat sweep's, reading through the
// accessor below, so the arithmetic is the same expression for expression.
// ---------------------------------------------------------------------------
constexpr int kTileW = 68;  // 64 + 2 halo columns on each side

// Staged fields of each kernel: slot s = (group G(s), field K(s)) with a z
// queue index Q(s) (or -1) and an LDS tile index L(s) (or -1).
struct StressFwdSlots {
  static constexpr int N = 3, NQ = 3, NL = 3;
  __host__ __device__ static constexpr int G(int) { return GW; }
  __host__ __device__ static constexpr int K(int s) { return s == 0 ? F_VZ : s == 1 ? F_VX : F_VY; }
  __host__ __device__ static constexpr int Q(int s) { return s; }
  __host__ __device__ static constexpr int L(int s) { return s; }
};
struct VelocityFwdSlots {  // szz (z), sxz (x,z), syz (y,z), sxx (x), sxy (x,y), syy (y)
  static constexpr int N = 6, NQ = 3, NL = 5;
  __host__ __device__ static constexpr int G(int) { return GW; }
  __host__ __device__ static constexpr int K(int s) {
    return s == 0 ? F_SZZ : s == 1 ? F_SXZ : s == 2 ? F_SYZ : s == 3 ? F_SXX : s == 4 ? F_SXY : F_SYY;
  }
  __host__ __device__ static constexpr int Q(int s) { return s < 3 ? s : -1; }
  __host__ __device__ static constexpr int L(int s) { return s == 0 ? -1 : s - 1; }
};
struct StressBwdSlots {  // forward and adjoint v, all three axes each
  static constexpr int N = 6, NQ = 6, NL = 6;
  __host__ __device__ static constexpr int G(int s) { return s < 3 ? GW : GA; }
  __host__ __device__ static constexpr int K(int s) { return s % 3 == 0 ? F_VZ : s % 3 == 1 ? F_VX : F_VY; }
  __host__ __device__ static constexpr int Q(int s) { return s; }
  __host__ __device__ static constexpr int L(int s) { return s; }
};
// forward sigma (reverse part): szz (z), sxz (x,z), syz (y,z), sxx (x), sxy (x,y), syy (y);
// adjoint sigma: szz, sxx, syy (x,y,z), sxz (x,z), syz (y,z), sxy (x,y)
struct VelocityBwdSlots {
  static constexpr int N = 12, NQ = 8, NL = 11;
  __host__ __device__ static constexpr int G(int s) { return s < 6 ? GW : GA; }
  __host__ __device__ static constexpr int K(int s) {
    return s == 0 ? F_SZZ : s == 1 ? F_SXZ : s == 2 ? F_SYZ : s == 3 ? F_SXX : s == 4 ? F_SXY : s == 5 ? F_SYY
         : s == 6 ? F_SZZ : s == 7 ? F_SXX : s == 8 ? F_SYY : s == 9 ? F_SXZ : s == 10 ? F_SYZ : F_SXY;
  }
  __host__ __device__ static constexpr int Q(int s) { return s < 3 ? s : s < 6 ? -1 : s < 11 ? s - 3 : -1; }
  __host__ __device__ static constexpr int L(int s) { return s == 0 ? -1 : s - 1; }
};
template <class S>
__host__ __device__ constexpr int slot_of(int g, int k) {
  for (int s = 0; s < S::N; s++)
    if (S::G(s) == g && S::K(s) == k) return s;
  return -1;
}

template <class S, int R>
struct MarchAcc {
  FlatAcc f;
  const float (*q)[5];  // z queues of this thread's column
  const float* t;       // this cell in LDS tile 0
  __device__ __forceinline__ float own(int g, int k) const { return at(g, k, AZ, 0); }
  __device__ __forceinline__ float at(int g, int k, int ax, int d) const {
    const int s = slot_of<S>(g, k);
    if (s < 0) return f.at(g, k, ax, d);
    if ((ax == AZ || d == 0) && S::Q(s) >= 0) return q[S::Q(s)][2 + d];
    if ((ax == AZ && d != 0) || S::L(s) < 0) return f.at(g, k, ax, d);  // not staged along this axis
    const float* p = t + S::L(s) * (R + 4) * kTileW;
    return ax == AX ? p[d * kTileW] : ax == AY ? p[d] : p[0];
  }
  __device__ __forceinline__ void st(int g, int k, float v) const { f.st(g, k, v); }
};

template <class S, int R, typename Body>
__device__ __forceinline__ void march(const Sweep& Sw, const Geom& G, const Groups& F, Body&& body) {
  __shared__ float lds[S::NL * (R + 4) * kTileW];
  const Grp gw = group(F.w, F_WAVE_COUNT, G.fs), ga = group(F.a, F_WAVE_COUNT, G.fs);
  const Grp gp = group(F.p, P_PSI_COUNT, G.fs), gm = group(F.m, M_COUNT, G.fs);
  auto grp = [&](int g) -> const Grp& { return g == GW ? gw : g == GA ? ga : g == GP ? gp : gm; };
  const int b = blockIdx.x;
  const int xcd = b & 7, j = b >> 3;
  const int per_plane = Sw.rg_max * Sw.nyb;
  const int zc0 = j / per_plane;
  const int zc = Sw.zrev ? Sw.nzc - 1 - zc0 : zc0;
  const int r = j - zc0 * per_plane;
  const int rg = r / Sw.nyb;
  const int yb = r - rg * Sw.nyb;
  const int rows = Sw.rows_base + (xcd < Sw.rows_rem ? 1 : 0);
  const int row = __builtin_amdgcn_readfirstlane((int)threadIdx.y);  // blockDim.x == 64
  const int lane = (int)threadIdx.x;
  const int x_base = Sw.x0 + xcd * Sw.rows_base + min(xcd, Sw.rows_rem) + rg * R;
  const int y_base = (Sw.yb0 + yb) * 64;
  Cell c;
  c.x = x_base + row;
  c.y = y_base + lane;
  const bool active = rg * R + row < rows && c.y >= Sw.y0 && c.y <= Sw.y1;
  const unsigned sx = 4u * (unsigned)G.py, sz = 4u * (unsigned)G.pxy;
  // this thread's column, clamped into the grid (inactive threads stage real values)
  const int xc = min(c.x, G.nx - 1), yc = min(c.y, G.ny - 1);
  const unsigned col = 4u * ((unsigned)xc * (unsigned)G.py + (unsigned)yc);
  const int za = Sw.z0 + zc * Sw.Z;
  const int zb = min(za + Sw.Z, Sw.z0 + Sw.nz);
  float q[S::NQ > 0 ? S::NQ : 1][5], nxt[S::NQ > 0 ? S::NQ : 1];
#pragma unroll
  for (int s = 0; s < S::N; s++)
    if (S::Q(s) >= 0)
#pragma unroll
      for (int k = 0; k < 5; k++) q[S::Q(s)][k] = ld(grp(S::G(s)), S::K(s), (unsigned)(za - 2 + k) * sz + col);
  const int tid = row * 64 + lane;
  constexpr int kHalo = 4 * kTileW + 4 * R;
  float* const me = lds + (row + 2) * kTileW + lane + 2;
#pragma unroll 1
  for (int z = za; z < zb; z++) {
    const unsigned plane = (unsigned)z * sz;
    // refill of the queues for the next plane
    if (z + 1 < zb) {
#pragma unroll
      for (int s = 0; s < S::N; s++)
        if (S::Q(s) >= 0) nxt[S::Q(s)] = ld(grp(S::G(s)), S::K(s), plane + 3 * sz + col);
    }
    // stage this plane: own column, then the halo
#pragma unroll
    for (int s = 0; s < S::N; s++)
      if (S::L(s) >= 0)
        me[S::L(s) * (R + 4) * kTileW] =
            S::Q(s) >= 0 ? q[S::Q(s)][2] : ld(grp(S::G(s)), S::K(s), plane + col);
    for (int h = tid; h < kHalo; h += 64 * R) {
      int tr, tc;
      if (h < 4 * kTileW) {
        const int hr = h / kTileW;
        tr = hr < 2 ? hr : R + hr;
        tc = h - hr * kTileW;
      } else {
        // side halo of the R own rows: columns 0, 1 (y - 2, y - 1 of lane 0)
        // and 66, 67 (y + 1, y + 2 of lane 63)
        const int e = h - 4 * kTileW;
        tr = 2 + (e >> 2);
        tc = (e & 3) < 2 ? (e & 3) : 64 + (e & 3);
      }
      const int hx = min(max(x_base - 2 + tr, 0), G.nx - 1);
      const int hy = min(max(y_base - 2 + tc, 0), G.ny - 1);
      const unsigned off = plane + 4u * ((unsigned)hx * (unsigned)G.py + (unsigned)hy);
#pragma unroll
      for (int s = 0; s < S::N; s++)
        if (S::L(s) >= 0) lds[S::L(s) * (R + 4) * kTileW + tr * kTileW + tc] = ld(grp(S::G(s)), S::K(s), off);
    }
    __syncthreads();
    if (active) {
      const unsigned o = plane + 4u * ((unsigned)c.x * (unsigned)G.py + (unsigned)c.y);
      c.z = z;
      c.o = o;
      c.zn = Nb{o - 2 * sz, o - sz, o + sz, o + 2 * sz};
      c.xn = Nb{o - 2 * sx, o - sx, o + sx, o + 2 * sx};
      c.yn = Nb{o - 8u, o - 4u, o + 4u, o + 8u};
      const MarchAcc<S, R> acc{FlatAcc{gw, ga, gp, gm, o, {c.zn, c.xn, c.yn}}, q, me};
      body(c, acc);
    }
    __syncthreads();
#pragma unroll
    for (int s = 0; s < S::N; s++)
      if (S::Q(s) >= 0) {
#pragma unroll
        for (int k = 0; k < 4; k++) q[S::Q(s)][k] = q[S::Q(s)][k + 1];
        q[S::Q(s)][4] = nxt[S::Q(s)];
      }
  }
}

template <int R>
__global__ __launch_bounds__(64 * R) void k_stress_fwd_m(Sweep Sw, Geom G, Groups F, Prof C, Source S, int it) {
  const unsigned so = source_off(S);
  const float amp = source_amp(S, G, it);
  march<StressFwdSlots, R>(Sw, G, F, [&](const Cell& c, const auto& acc) { stress_fwd_cell(G, acc, C, c, so, amp); });
}

template <int FORM, int R>
__global__ __launch_bounds__(64 * R) void k_velocity_fwd_m(Sweep Sw, Geom G, Groups F, Prof C) {
  march<VelocityFwdSlots, R>(Sw, G, F, [&](const Cell& c, const auto& acc) { velocity_fwd_cell<FORM>(G, acc, C, c); });
}

template <int FORM, int R>
__global__ __launch_bounds__(64 * R) void k_velocity_bwd_m(Sweep Sw, Geom G, Groups F, Prof C) {
  march<VelocityBwdSlots, R>(Sw, G, F, [&](const Cell& c, const auto& acc) {
    const Flags f = flags_of(G, c);
    if (f.xz) {
      if (f.y) velocity_bwd_cell<FORM, true, true, true>(G, acc, C, c);
      else velocity_bwd_cell<FORM, true, true, false>(G, acc, C, c);
    } else {
      if (f.y) velocity_bwd_cell<FORM, true, false, true>(G, acc, C, c);
      else velocity_bwd_cell<FORM, true, false, false>(G, acc, C, c);
    }
  });
}

template <int R>
__global__ __launch_bounds__(64 * R) void k_stress_bwd_m(Sweep Sw, Geom G, Groups F, Prof C, Source S, int it) {
  const unsigned so = source_off(S);
  const float amp = source_amp(S, G, it);
  march<StressBwdSlots, R>(Sw, G, F, [&](const Cell& c, const auto& acc) {
    const Flags f = flags_of(G, c);
    if (f.xz) {
      if (f.y) stress_bwd_cell<true, true, true>(G, acc, C, c, so, amp);
      else stress_bwd_cell<true, true, false>(G, acc, C, c, so, amp);
    } else {
      if (f.y) stress_bwd_cell<true, false, true>(G, acc, C, c, so, amp);
      else stress_bwd_cell<true, false, false>(G, acc, C, c, so, amp);
    }
  });
}

// ---------------------------------------------------------------------------
// Fused forward step: the stress and the velocity update of one time step in
// one pass over the grid, reading the state of step it from the "in" field
// groups and writing the state of step it + 1 to the "out" groups (the caller
// alternates the two sets). A workgroup of 64 x (R + 4) threads owns R x rows
// of 60 y cells and a chunk of z planes; it marches z, computing the new
// stress of plane z + 2 for its whole thread grid (the owned cells plus a
// 2-cell ring around them, which neighbouring workgroups own and compute as
// well) and then the new velocity of plane z for the owned cells from the new
// stress: z neighbours from per-thread register queues, x and y neighbours from
// an LDS tile of the plane. Because no workgroup writes what another reads,
// the order of workgroups does not matter; every value is the flat kernels'
// expression, so the results are bit-identical to them.
// ---------------------------------------------------------------------------
constexpr int kFusedY = 60;  // owned y cells per wavefront (lanes 2..61)

// Stress accessor of the fused step: state of step it from the in-groups; new
// stress values go to `sig` (and to the out-group for owned cells), new psi
// values to the out-group for owned cells only.
struct FusedSAcc {
  FlatAcc f;
  Grp wo, po;
  float* sig;  // [6]: new szz, sxx, sxz, syy, sxy, syz (WaveField order from F_SZZ)
  bool owned;
  float* rs;
  long rn;
  int rk;
  __device__ __forceinline__ float own(int gi, int k) const { return f.own(gi, k); }
  __device__ __forceinline__ float at(int gi, int k, int ax, int d) const { return f.at(gi, k, ax, d); }
  __device__ __forceinline__ void st(int gi, int k, float v) const {
    if (gi == GW) {
      sig[k - F_SZZ] = v;
      if (owned) {
        efwi::st(wo, k, f.o, v);
        if (rk >= 0) rs[(long)k * rn + rk] = v;
      }
    } else if (owned) {
      efwi::st(po, k, f.o, v);
    }
  }
};

// Velocity accessor of the fused step: the new stress from the register
// queues (z neighbours) and the LDS tile (x and y neighbours), everything else
// from the in-groups; results to the out-groups.
template <int NR>
struct FusedVAcc {
  FlatAcc f;
  Grp wo, po;
  const float (*q)[5];  // z queues of szz, sxz, syz: planes z-2..z+2
  const float* t;       // this cell in LDS field 0 of the current plane
  float* rs;
  long rn;
  int rk;
  __device__ __forceinline__ float own(int gi, int k) const { return f.own(gi, k); }
  __device__ __forceinline__ float at(int gi, int k, int ax, int d) const {
    if (gi != GW || k < F_SZZ) return f.at(gi, k, ax, d);
    if (ax == AZ) return q[k == F_SZZ ? 0 : k == F_SXZ ? 1 : 2][2 + d];
    // LDS fields: sxz, syz, sxx, sxy, syy
    const int l = k == F_SXZ ? 0 : k == F_SYZ ? 1 : k == F_SXX ? 2 : k == F_SXY ? 3 : 4;
    const float* p = t + l * NR * 64;
    return ax == AX ? p[d * 64] : p[d];
  }
  __device__ __forceinline__ void st(int gi, int k, float v) const {
    if (gi == GW) {
      efwi::st(wo, k, f.o, v);
      if (rk >= 0) rs[(long)k * rn + rk] = v;
    } else {
      efwi::st(po, k, f.o, v);
    }
  }
};

template <int FORM, int R>
__global__ __launch_bounds__(64 * (R + 4)) void k_fwd_fused(FusedSweep S, Geom G, Groups Fi, Groups Fo, Prof C,
                                                            Source Src, int it, RingFold Rf) {
  constexpr int NR = R + 4;
  __shared__ float lds[2][5 * NR * 64];
  const int b = blockIdx.x;
  const int tile = (b & 7) * S.per_xcd + (b >> 3);
  if (tile >= S.tiles) return;  // whole workgroup: uniform
  const int ysg = tile % S.nys;
  const int xt = (tile / S.nys) % S.nxt;
  const int zc0 = tile / (S.nys * S.nxt);
  const int zc = S.zrev ? S.nzc - 1 - zc0 : zc0;
  const int row = __builtin_amdgcn_readfirstlane((int)threadIdx.y);
  const int lane = (int)threadIdx.x;
  const int x = xt * R + row;                 // rows 2.. owned: x = 2 + xt R + (row - 2)
  const int y = ysg * kFusedY + lane;         // lanes 2.. owned
  const bool x_in = x >= 2 && x <= G.nx - 3;  // stress computed here
  const bool y_in = y >= 2 && y <= G.ny - 3;
  const bool own_xy = row >= 2 && row < R + 2 && x_in && lane >= 2 && lane < kFusedY + 2 && y_in;
  const int za = 2 + zc * S.Zc, zb = min(za + S.Zc, G.nz - 2);  // owned planes [za, zb)
  const Grp wi = group(Fi.w, F_WAVE_COUNT, G.fs), pi = group(Fi.p, P_PSI_COUNT, G.fs);
  const Grp wo = group(Fo.w, F_WAVE_COUNT, G.fs), po = group(Fo.p, P_PSI_COUNT, G.fs);
  const Grp gm = group(Fi.m, M_COUNT, G.fs);
  const Grp ga = group(Fi.a, F_WAVE_COUNT, G.fs);
  const unsigned so = source_off(Src);
  const float amp = source_amp(Src, G, it);
  const unsigned sx = 4u * (unsigned)G.py, sz = 4u * (unsigned)G.pxy;
  // this thread's column (clamped into the row pitch for the halo lanes past it)
  const int yc = min(y, G.py - 1);
  const unsigned col = 4u * ((unsigned)min(x, G.nx - 1) * (unsigned)G.py + (unsigned)yc);
  float q[3][5], dq[3][3];
#pragma unroll
  for (int i = 0; i < 3; i++) {
#pragma unroll
    for (int k = 0; k < 5; k++) q[i][k] = 0.0f;
#pragma unroll
    for (int k = 0; k < 3; k++) dq[i][k] = 0.0f;
  }
  float* const me0 = &lds[0][row * 64 + lane];
  float* const me1 = &lds[1][row * 64 + lane];
#pragma unroll 1
  for (int zs = za - 2; zs < zb + 2; zs++) {
    // ---- new stress of plane zs at this thread's column
    float s6[6];
    const unsigned o = (unsigned)zs * sz + col;
    if (x_in && y_in && zs >= 2 && zs <= G.nz - 3) {
      Cell c;
      c.z = zs;
      c.x = x;
      c.y = y;
      c.o = o;
      c.zn = Nb{o - 2 * sz, o - sz, o + sz, o + 2 * sz};
      c.xn = Nb{o - 2 * sx, o - sx, o + sx, o + 2 * sx};
      c.yn = Nb{o - 8u, o - 4u, o + 4u, o + 8u};
      const bool own = own_xy && zs >= za && zs < zb;
      const FusedSAcc acc{FlatAcc{wi, ga, pi, gm, o, {c.zn, c.xn, c.yn}}, wo, po, s6, own, Rf.slot, Rf.B.n,
                          own ? ring_index(Rf, c) : -1};
      stress_fwd_cell(G, acc, C, c, so, amp);
    } else if (zs >= 0 && zs < G.nz && y < G.ny) {
      // outside the stencil box: the state there never changes
#pragma unroll
      for (int k = 0; k < 6; k++) s6[k] = ld(wi, F_SZZ + k, o);
    } else {
#pragma unroll
      for (int k = 0; k < 6; k++) s6[k] = 0.0f;
    }
    // ---- queues: szz, sxz, syz five planes deep, sxx, sxy, syy three
#pragma unroll
    for (int i = 0; i < 3; i++) {
#pragma unroll
      for (int k = 0; k < 4; k++) q[i][k] = q[i][k + 1];
#pragma unroll
      for (int k = 0; k < 2; k++) dq[i][k] = dq[i][k + 1];
    }
    q[0][4] = s6[F_SZZ - F_SZZ];
    q[1][4] = s6[F_SXZ - F_SZZ];
    q[2][4] = s6[F_SYZ - F_SZZ];
    dq[0][2] = s6[F_SXX - F_SZZ];
    dq[1][2] = s6[F_SXY - F_SZZ];
    dq[2][2] = s6[F_SYY - F_SZZ];
    // ---- new velocity of plane zv = zs - 2
    const int zv = zs - 2;
    if (zv >= za) {
      float* const me = (zv & 1) ? me1 : me0;
      me[0 * NR * 64] = q[1][2];
      me[1 * NR * 64] = q[2][2];
      me[2 * NR * 64] = dq[0][0];
      me[3 * NR * 64] = dq[1][0];
      me[4 * NR * 64] = dq[2][0];
      __syncthreads();
      if (own_xy) {
        const unsigned ov = (unsigned)zv * sz + col;
        Cell c;
        c.z = zv;
        c.x = x;
        c.y = y;
        c.o = ov;
        c.zn = Nb{ov - 2 * sz, ov - sz, ov + sz, ov + 2 * sz};
        c.xn = Nb{ov - 2 * sx, ov - sx, ov + sx, ov + 2 * sx};
        c.yn = Nb{ov - 8u, ov - 4u, ov + 4u, ov + 8u};
        const FusedVAcc<NR> acc{FlatAcc{wi, ga, pi, gm, ov, {c.zn, c.xn, c.yn}}, wo, po, q, me, Rf.slot,
                                Rf.B.n, ring_index(Rf, c)};
        velocity_fwd_cell<FORM>(G, acc, C, c);
      }
    }
  }
}

// ---------------------------------------------------------------------------
// Batched shots: one persistent launch per stencil step covers every (shot,
// tile) item of the batch. Items are the flat sweep's workgroup tiles; those of
// XCD strip q (tile index b with b % 8 == q) of every shot form queue q. A
// workgroup drains the queue of the XCD it runs on (hardware register
// XCC_ID), taking `take` consecutive items per atomic, then steals from the
// other queues, so every item is processed once whatever the placement. The time step is
// read from device memory, so a captured step replays unchanged.
// ---------------------------------------------------------------------------
__device__ __forceinline__ int xcc_id() {
  // s_getreg_b32 hwreg(HW_REG_XCC_ID, 0, 4): bits [3:0] hold the XCC of the wave
  return __builtin_amdgcn_s_getreg((3 << 11) | (0 << 6) | 20) & 7;
}

// The value of `v` held by thread 0, in every thread of the workgroup.
__device__ __forceinline__ int wg_broadcast(int v, bool one_wave, int* slot) {
  if (one_wave) return __builtin_amdgcn_readfirstlane(v);
  __syncthreads();
  if (threadIdx.x == 0 && threadIdx.y == 0) *slot = v;
  __syncthreads();
  return __builtin_amdgcn_readfirstlane(*slot);
}

template <typename Item>
__device__ __forceinline__ void persistent_items(const BatchLaunch& L, int per_xcd, Item&& item) {
  __shared__ int slot;
  const int take = L.take;
  const int nq = L.nshots * per_xcd;
  const int me = xcc_id();
  const bool one_wave = blockDim.y == 1;
  const bool leader = threadIdx.x == 0 && threadIdx.y == 0;
  auto run = [&](int q, int i) {
    const int shot = i / per_xcd;
    item(L.reverse ? L.nshots - 1 - shot : shot, 8 * (i - shot * per_xcd) + q);
  };
  // claim the first dynamic items now: the atomic's round trip overlaps the
  // static share
  unsigned* head = L.counters + me * kQueueStride;
  int got = 0;
  if (leader) got = (int)atomicAdd(head, (unsigned)take);
  // static share: workgroup w runs items j, j + wq, ... (L.stat of them) of
  // queue w % 8, j = w / 8, wq = workgroups per queue; the grid is a multiple
  // of 8 and L.stat * wq <= every queue's length
  const int wq = (int)(gridDim.x >> 3);
  const int done = L.stat * wq;
  for (int k = 0, i = (int)(blockIdx.x >> 3); k < L.stat; k++, i += wq) run((int)(blockIdx.x & 7), i);
  // the rest of every queue: `take` items per atomic, own XCD's queue first
  for (int d = 0; d < 8; d++) {
    const int q = (me + d) & 7;
    head = L.counters + q * kQueueStride;
    if (d > 0) {
      // another XCD's queue: skip it without an atomic once it is drained (one
      // load by thread 0, so that every wave of the workgroup takes the same path)
      int h = 0;
      if (leader) h = (int)__hip_atomic_load(head, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
      if (done + wg_broadcast(h, one_wave, &slot) >= nq) continue;
      if (leader) got = (int)atomicAdd(head, (unsigned)take);
    }
    int cur = done + wg_broadcast(got, one_wave, &slot);
    while (cur < nq) {
      // claim the next items before working on these
      int next = 0;
      if (leader) next = (int)atomicAdd(head, (unsigned)take);
      const int end = min(cur + take, nq);
      for (int i = cur; i < end; i++) run(q, i);
      cur = done + wg_broadcast(next, one_wave, &slot);
    }
  }
}

__device__ __forceinline__ int device_step(const int* it) {
  return __builtin_amdgcn_readfirstlane(__hip_atomic_load(it, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT));
}

// The shot table is read with vector loads (the compiler cannot prove that the
// kernel's stores leave it unchanged); the values are wave-uniform, and moving
// them to scalar registers keeps the buffer descriptors built from them scalar
// (descriptors in vector registers turn every buffer access into a loop).
template <class T>
__device__ __forceinline__ T* uniform_ptr(T* p) {
  const unsigned long long v = (unsigned long long)p;
  const unsigned lo = __builtin_amdgcn_readfirstlane((unsigned)v);
  const unsigned hi = __builtin_amdgcn_readfirstlane((unsigned)(v >> 32));
  return (T*)(((unsigned long long)hi << 32) | lo);
}
__device__ __forceinline__ long uniform_long(long x) {
  return (long)(unsigned long long)uniform_ptr((char*)x);
}

// KIND: 0 stress_fwd, 1 velocity_fwd, 2 velocity_bwd, 3 stress_bwd
template <int KIND, int FORM>
__global__ __launch_bounds__(256) void k_batch(Sweep Sw, Geom G, Prof C, BatchLaunch L, int per_xcd, BatchRing R) {
  const int it = device_step(L.it);
  // ring slot of this step (the same index in every shot's ring), or none
  const int slot = R.mode == 1 ? it + 1 : it;
  const bool ring_on = R.mode == 2 || (R.mode == 1 && slot <= R.last);
  int have = -1;
  Groups F{};
  Source S{};
  RingFold Rf{};
  Rf.B = R.B;
  persistent_items(L, per_xcd, [&](int shot, int b) {
    if (shot != have) {  // consecutive items mostly belong to one shot
      const ShotDev& sd = L.shots[shot];
      F.w = uniform_ptr(sd.F.w);
      F.a = uniform_ptr(sd.F.a);
      F.p = uniform_ptr(sd.F.p);
      F.m = uniform_ptr(sd.F.m);
      S.cell = uniform_long(sd.cell);
      S.stf = uniform_ptr(sd.stf);
      Rf.slot = ring_on ? uniform_ptr(sd.ring) + (long)slot * 9 * R.B.n : nullptr;
      have = shot;
    }
    if constexpr (KIND == 0) {
      const unsigned so = source_off(S);
      const float amp = source_amp(S, G, it);
      sweep_block(Sw, G, b, [&](const Cell& c) { stress_fwd_cell(G, save_acc(G, F, Rf, c), C, c, so, amp); });
    } else if constexpr (KIND == 1) {
      sweep_block(Sw, G, b, [&](const Cell& c) { velocity_fwd_cell<FORM>(G, save_acc(G, F, Rf, c), C, c); });
    } else if constexpr (KIND == 2) {
      sweep_block(Sw, G, b, [&](const Cell& c) {
        const FlatAcc acc = flat_acc(G, F, c);
        const Flags f = flags_of(G, c);
        const RingRestore<F_VZ, F_VY + 1> rr(Rf, c);
        if (f.xz) {
          if (f.y) velocity_bwd_cell<FORM, true, true, true>(G, acc, C, c);
          else velocity_bwd_cell<FORM, true, true, false>(G, acc, C, c);
        } else {
          if (f.y) velocity_bwd_cell<FORM, true, false, true>(G, acc, C, c);
          else velocity_bwd_cell<FORM, true, false, false>(G, acc, C, c);
        }
        rr.store(acc);
      });
    } else {
      const unsigned so = source_off(S);
      const float amp = source_amp(S, G, it);
      sweep_block(Sw, G, b, [&](const Cell& c) {
        const FlatAcc acc = flat_acc(G, F, c);
        const Flags f = flags_of(G, c);
        const RingRestore<F_SZZ, F_WAVE_COUNT> rr(Rf, c);
        if (f.xz) {
          if (f.y) stress_bwd_cell<true, true, true>(G, acc, C, c, so, amp);
          else stress_bwd_cell<true, true, false>(G, acc, C, c, so, amp);
        } else {
          if (f.y) stress_bwd_cell<true, false, true>(G, acc, C, c, so, amp);
          else stress_bwd_cell<true, false, false>(G, acc, C, c, so, amp);
        }
        rr.store(acc);
      });
    }
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
  S.yb0 = y0 / 64;  // segments of 64 aligned at y = 0, from the one holding y0
  S.nyb = y1 / 64 - S.yb0 + 1;
  S.nzc = (S.nz + S.Z - 1) / S.Z;
  S.zrev = 0;
  S.sk_z0 = 1; S.sk_z1 = 0; S.sk_x0 = 1; S.sk_x1 = 0; S.sk_yb0 = 1; S.sk_yb1 = 0;
  S.blocks = 8L * S.rg_max * S.nyb * S.nzc;
  return S;
}

static inline Sweep sweep_full(const Geom& G, Tile t) { return make_sweep(2, G.nz - 3, 2, G.nx - 3, 2, G.ny - 3, t); }

static inline dim3 grid_of(const Sweep& S) { return dim3((unsigned)S.blocks); }
static inline dim3 block_of(const Sweep& S) { return dim3(64, S.R); }

// Launch shapes with zdepth >= 16 run the z-marching kernels (rows 4 or 8 per
// workgroup); smaller ones the flat sweep (rows 1, 2 or 4).
static inline bool marching(Tile t) { return t.zdepth >= 16; }
static inline Sweep sweep_march(const Geom& G, Tile t) {
  Sweep S = sweep_full(G, Tile{1, t.zdepth});
  const int R = t.rows >= 8 ? 8 : 4;
  S.R = R;
  const int rows_max = S.rows_base + (S.rows_rem ? 1 : 0);
  S.rg_max = (rows_max + R - 1) / R;
  S.blocks = 8L * S.rg_max * S.nyb * S.nzc;
  return S;
}
#define EFWI_MARCH(KERNEL, ZREV, ...)                                                             \
  do {                                                                                            \
    Sweep Sw = sweep_march(G, t);                                                                 \
    Sw.zrev = (ZREV);                                                                             \
    if (Sw.R == 8) hipLaunchKernelGGL((KERNEL<8>), grid_of(Sw), block_of(Sw), 0, s, __VA_ARGS__); \
    else hipLaunchKernelGGL((KERNEL<4>), grid_of(Sw), block_of(Sw), 0, s, __VA_ARGS__);          \
  } while (0)

// The band inside the flat sweep's box [2, n-3] and at least 10 cells wide on
// every axis (so its slab groups are as band_index enumerates them).
bool ring_foldable(const Geom& G, const Band& B) {
  const int n[3] = {G.nz, G.nx, G.ny};
  for (int ax = 0; ax < 3; ax++)
    if (B.lo[ax] < 2 || B.hi[ax] > n[ax] - 3 || B.hi[ax] - B.lo[ax] + 1 < 10) return false;
  return true;
}
// The fold a flat launch takes: the caller's, or none.
static inline RingFold fold_for(const Geom& G, const RingFold* r) {
  if (r && r->slot && ring_foldable(G, r->B)) return *r;
  RingFold none{};
  none.slot = nullptr;
  return none;
}

bool launch_stress_fwd(Tile t, const Geom& G, const Groups& F, const Prof& C, const Source& S, int it,
                       hipStream_t s, const RingFold* ring) {
  if (marching(t)) {
    EFWI_MARCH(k_stress_fwd_m, 0, Sw, G, F, C, S, it);
    return false;
  }
  const Sweep Sw = sweep_full(G, t);
  const RingFold Rf = fold_for(G, ring);
  hipLaunchKernelGGL(k_stress_fwd, grid_of(Sw), block_of(Sw), 0, s, Sw, G, F, C, S, it, Rf);
  return Rf.slot != nullptr;
}
bool launch_velocity_fwd(Tile t, int form, const Geom& G, const Groups& F, const Prof& C, hipStream_t s,
                         const RingFold* ring) {
  if (marching(t)) {
    Sweep Sw = sweep_march(G, t);
    Sw.zrev = 1;
    if (form == 1) {
      if (Sw.R == 8) hipLaunchKernelGGL((k_velocity_fwd_m<1, 8>), grid_of(Sw), block_of(Sw), 0, s, Sw, G, F, C);
      else hipLaunchKernelGGL((k_velocity_fwd_m<1, 4>), grid_of(Sw), block_of(Sw), 0, s, Sw, G, F, C);
    } else {
      if (Sw.R == 8) hipLaunchKernelGGL((k_velocity_fwd_m<0, 8>), grid_of(Sw), block_of(Sw), 0, s, Sw, G, F, C);
      else hipLaunchKernelGGL((k_velocity_fwd_m<0, 4>), grid_of(Sw), block_of(Sw), 0, s, Sw, G, F, C);
    }
    return false;
  }
  Sweep Sw = sweep_full(G, t);
  Sw.zrev = 1;  // the stress kernel before it ran first to last
  const RingFold Rf = fold_for(G, ring);
  if (form == 1) hipLaunchKernelGGL(k_velocity_fwd<1>, grid_of(Sw), block_of(Sw), 0, s, Sw, G, F, C, Rf);
  else hipLaunchKernelGGL(k_velocity_fwd<0>, grid_of(Sw), block_of(Sw), 0, s, Sw, G, F, C, Rf);
  return Rf.slot != nullptr;
}
// y segments wholly inside the y interior [p, ny - 1 - p]: [lo, hi] (empty if lo > hi)
static inline void interior_segments(const Geom& G, int& lo, int& hi) {
  const int p = G.npml;
  lo = (p + 63) / 64;
  hi = (G.ny - p) / 64 - 1;
}
bool has_interior_segments(const Geom& G) {
  int lo, hi;
  interior_segments(G, lo, hi);
  return lo <= hi && G.nz - 2 * G.npml > 0 && G.nx - 2 * G.npml > 0;
}
// The interior launch of the split and the general sweep that leaves it out.
static inline Sweep split_sweeps(const Geom& G, Tile t, Sweep& rim) {
  int lo, hi;
  interior_segments(G, lo, hi);
  const int p = G.npml;
  Sweep in = make_sweep(p, G.nz - 1 - p, p, G.nx - 1 - p, 64 * lo, 64 * hi + 63, t);
  rim = make_sweep(2, G.nz - 3, 2, G.nx - 3, 2, G.ny - 3, t);
  rim.sk_z0 = p; rim.sk_z1 = G.nz - 1 - p;
  rim.sk_x0 = p; rim.sk_x1 = G.nx - 1 - p;
  rim.sk_yb0 = lo; rim.sk_yb1 = hi;
  return in;
}

bool launch_velocity_bwd(Tile t, int form, const Geom& G, const Groups& F, const Prof& C, hipStream_t s,
                         const RingFold* ring, bool split) {
  if (marching(t)) {
    const Sweep Sw = sweep_march(G, t);
    if (form == 1) {
      if (Sw.R == 8) hipLaunchKernelGGL((k_velocity_bwd_m<1, 8>), grid_of(Sw), block_of(Sw), 0, s, Sw, G, F, C);
      else hipLaunchKernelGGL((k_velocity_bwd_m<1, 4>), grid_of(Sw), block_of(Sw), 0, s, Sw, G, F, C);
    } else {
      if (Sw.R == 8) hipLaunchKernelGGL((k_velocity_bwd_m<0, 8>), grid_of(Sw), block_of(Sw), 0, s, Sw, G, F, C);
      else hipLaunchKernelGGL((k_velocity_bwd_m<0, 4>), grid_of(Sw), block_of(Sw), 0, s, Sw, G, F, C);
    }
    return false;
  }
  const RingFold Rf = fold_for(G, ring);
  if (split && has_interior_segments(G)) {
    Sweep rim;
    const Sweep in = split_sweeps(G, t, rim);
    if (form == 1) {
      hipLaunchKernelGGL(k_velocity_bwd_in<1>, grid_of(in), block_of(in), 0, s, in, G, F, C, Rf);
      hipLaunchKernelGGL((k_velocity_bwd<1, true>), grid_of(rim), block_of(rim), 0, s, rim, G, F, C, Rf);
    } else {
      hipLaunchKernelGGL(k_velocity_bwd_in<0>, grid_of(in), block_of(in), 0, s, in, G, F, C, Rf);
      hipLaunchKernelGGL((k_velocity_bwd<0, true>), grid_of(rim), block_of(rim), 0, s, rim, G, F, C, Rf);
    }
    return Rf.slot != nullptr;
  }
  const Sweep Sw = sweep_full(G, t);
  if (form == 1) hipLaunchKernelGGL((k_velocity_bwd<1, true>), grid_of(Sw), block_of(Sw), 0, s, Sw, G, F, C, Rf);
  else hipLaunchKernelGGL((k_velocity_bwd<0, true>), grid_of(Sw), block_of(Sw), 0, s, Sw, G, F, C, Rf);
  return Rf.slot != nullptr;
}
bool launch_stress_bwd(Tile t, const Geom& G, const Groups& F, const Prof& C, const Source& S, int it,
                       hipStream_t s, const RingFold* ring, bool split) {
  if (marching(t)) {
    EFWI_MARCH(k_stress_bwd_m, 1, Sw, G, F, C, S, it);
    return false;
  }
  const RingFold Rf = fold_for(G, ring);
  if (split && has_interior_segments(G)) {
    Sweep rim;
    Sweep in = split_sweeps(G, t, rim);
    in.zrev = rim.zrev = 1;
    hipLaunchKernelGGL(k_stress_bwd_in, grid_of(in), block_of(in), 0, s, in, G, F, C, S, it, Rf);
    hipLaunchKernelGGL(k_stress_bwd<true>, grid_of(rim), block_of(rim), 0, s, rim, G, F, C, S, it, Rf);
    return Rf.slot != nullptr;
  }
  Sweep Sw = sweep_full(G, t);
  Sw.zrev = 1;  // the velocity kernel before it ran first to last
  hipLaunchKernelGGL(k_stress_bwd<true>, grid_of(Sw), block_of(Sw), 0, s, Sw, G, F, C, S, it, Rf);
  return Rf.slot != nullptr;
}
bool launch_fwd_fused(int R, int zc, int form, const Geom& G, const Groups& Fi, const Groups& Fo, const Prof& C,
                      const Source& S, int it, hipStream_t s, const RingFold* ring) {
  R = R >= 12 ? 12 : R >= 8 ? 8 : 4;
  FusedSweep Sw;
  Sw.Zc = std::max(1, zc);
  Sw.nys = (G.ny - 4 + kFusedY - 1) / kFusedY;
  Sw.nxt = (G.nx - 4 + R - 1) / R;
  Sw.nzc = (G.nz - 4 + Sw.Zc - 1) / Sw.Zc;
  Sw.tiles = Sw.nys * Sw.nxt * Sw.nzc;
  Sw.per_xcd = (Sw.tiles + 7) / 8;
  Sw.zrev = 0;
  const RingFold Rf = fold_for(G, ring);
  const dim3 grid((unsigned)(8 * Sw.per_xcd)), block(64, R + 4);
#define EFWI_FUSED(RR)                                                                                     \
  if (form == 1) hipLaunchKernelGGL((k_fwd_fused<1, RR>), grid, block, 0, s, Sw, G, Fi, Fo, C, S, it, Rf); \
  else hipLaunchKernelGGL((k_fwd_fused<0, RR>), grid, block, 0, s, Sw, G, Fi, Fo, C, S, it, Rf)
  if (R == 12) { EFWI_FUSED(12); }
  else if (R == 8) { EFWI_FUSED(8); }
  else { EFWI_FUSED(4); }
#undef EFWI_FUSED
  return Rf.slot != nullptr;
}

void launch_velocity_adj(int form, const Geom& G, const Groups& F, const Prof& C, hipStream_t s) {
  const Sweep Sw = sweep_full(G, Tile{1, 1});
  const RingFold none = fold_for(G, nullptr);
  if (form == 1) hipLaunchKernelGGL((k_velocity_bwd<1, false>), grid_of(Sw), block_of(Sw), 0, s, Sw, G, F, C, none);
  else hipLaunchKernelGGL((k_velocity_bwd<0, false>), grid_of(Sw), block_of(Sw), 0, s, Sw, G, F, C, none);
}
void launch_stress_adj(const Geom& G, const Groups& F, const Prof& C, hipStream_t s) {
  const Sweep Sw = sweep_full(G, Tile{1, 1});
  Source none;
  none.cell = -1;
  none.stf = nullptr;
  hipLaunchKernelGGL(k_stress_bwd<false>, grid_of(Sw), block_of(Sw), 0, s, Sw, G, F, C, none, 0, fold_for(G, nullptr));
}

// Batched launches (flat tiles only; a zdepth >= 16 shape falls back to its
// flat counterpart rows x 1).
static inline Sweep sweep_batch(const Geom& G, Tile t) {
  if (t.zdepth >= 16) t = Tile{t.rows >= 4 ? 4 : t.rows, 1};
  return sweep_full(G, t);
}
int device_cus(int device) {
  int cus = 256;
  if (hipDeviceGetAttribute(&cus, hipDeviceAttributeMultiprocessorCount, device) != hipSuccess) cus = 256;
  return cus;
}
// Workgroups of `kernel` resident on one CU at `threads` per workgroup (from its
// register and LDS use), cached per block size.
template <class K>
static int resident_per_cu(K kernel, int threads, int* cache) {
  const int slot = threads / 64 - 1;
  if (slot >= 0 && slot < 8 && cache[slot] > 0) return cache[slot];
  int nb = 0;
  if (hipOccupancyMaxActiveBlocksPerMultiprocessor(&nb, kernel, threads, 0) != hipSuccess || nb < 1) nb = 1;
  if (slot >= 0 && slot < 8) cache[slot] = nb;
  return nb;
}
// Persistent workgroups: enough to fill every CU, a multiple of 8 (one set per
// XCD queue), at most one per queue item. Each workgroup first runs its static
// share of L.stat items (stat_pct percent of its even share), then claims the
// rest `take` items per atomic; take is lowered so that every workgroup gets at
// least 2 claims of the dynamic rest.
static inline dim3 batch_plan(const Sweep& Sw, BatchLaunch& L, int per_cu, int stat_pct) {
  const long nq = Sw.blocks / 8 * L.nshots;  // items per queue
  const long wq = std::max(1L, std::min((long)L.cus * per_cu / 8, nq));
  const long even = nq / wq;
  L.stat = (int)(even * std::max(0, std::min(stat_pct, 100)) / 100);
  const long rest = nq - L.stat * wq;
  L.take = (int)std::max(1L, std::min((long)L.take, rest / (wq * 2)));
  return dim3((unsigned)(8 * wq));
}
#define EFWI_BATCH(KERNEL)                                                                     \
  do {                                                                                        \
    static int occ_cache[8] = {0};                                                            \
    Sweep Sw = sweep_batch(G, t);                                                             \
    Sw.zrev = L.reverse; /* shots and planes last to first: start where the last launch ended */ \
    const int per_xcd = (int)(Sw.blocks / 8);                                                 \
    const dim3 blk = block_of(Sw);                                                            \
    const int per_cu = resident_per_cu((KERNEL), (int)(blk.x * blk.y * blk.z), occ_cache);    \
    BatchLaunch Lk = L;                                                                       \
    const dim3 grid = batch_plan(Sw, Lk, per_cu, L.stat);                                             \
    hipLaunchKernelGGL((KERNEL), grid, blk, 0, s, Sw, G, C, Lk, per_xcd, R);                  \
  } while (0)
// The caller's ring fold where it applies, else none.
static inline BatchRing batch_ring(const Geom& G, const BatchRing* ring) {
  if (ring && ring->mode && ring_foldable(G, ring->B)) return *ring;
  BatchRing none{};
  none.mode = 0;
  return none;
}
void launch_stress_fwd_batch(Tile t, const Geom& G, const Prof& C, const BatchLaunch& L, hipStream_t s,
                             const BatchRing* ring) {
  const BatchRing R = batch_ring(G, ring);
  EFWI_BATCH((k_batch<0, 0>));
}
void launch_velocity_fwd_batch(Tile t, int form, const Geom& G, const Prof& C, const BatchLaunch& L, hipStream_t s,
                               const BatchRing* ring) {
  const BatchRing R = batch_ring(G, ring);
  if (form == 1) EFWI_BATCH((k_batch<1, 1>));
  else EFWI_BATCH((k_batch<1, 0>));
}
void launch_velocity_bwd_batch(Tile t, int form, const Geom& G, const Prof& C, const BatchLaunch& L, hipStream_t s,
                               const BatchRing* ring) {
  const BatchRing R = batch_ring(G, ring);
  if (form == 1) EFWI_BATCH((k_batch<2, 1>));
  else EFWI_BATCH((k_batch<2, 0>));
}
void launch_stress_bwd_batch(Tile t, const Geom& G, const Prof& C, const BatchLaunch& L, hipStream_t s,
                             const BatchRing* ring) {
  const BatchRing R = batch_ring(G, ring);
  EFWI_BATCH((k_batch<3, 0>));
}

}  // namespace efwi
