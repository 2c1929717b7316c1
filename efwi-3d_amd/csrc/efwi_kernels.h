// Host-side launchers of the gfx950 kernels (defined in the .hip files).
#pragma once

#include "efwi_common.h"

namespace efwi {

// Boundary band of the reverse reconstruction: the box [lo, hi] (inclusive,
// per axis z, x, y) minus its inner box [lo+5, hi-5], stored as three groups
// of non-overlapping slabs (z slabs, then x slabs, then y slabs), each in
// (z, x, y) row-major order so that consecutive cells are consecutive in y.
struct Band {
  int lo[3], hi[3];
  long nzs, nxs, nys;  // cells in the z-, x- and y-slab groups
  long n;              // total cells per field
};

// Nine field pointers in the ring order vz, vx, vy, szz, sxx, sxz, syy, sxy, syz.
struct Fields9 {
  float* f[9];
};

// Launch shape of a stencil kernel: workgroups of 64 (y lanes) x `rows`
// (x rows) threads, each thread updating `zdepth` consecutive z planes.
struct Tile {
  int rows, zdepth;
};

// XCD-aware sweep of a cell box [z0..z0+nz-1] x [x0..] x [y0..y1]: workgroup
// b belongs to XCD b % 8, which owns a contiguous strip of rows_base (+1 for
// the first rows_rem XCDs) x rows and walks it by chunks of Z planes.
struct Sweep {
  int z0, nz, x0, rows_base, rows_rem, rg_max, y0, y1, yb0, nyb, R, Z, nzc;  // y segments yb0 .. yb0+nyb-1
  int zrev;  // plane chunks last to first (consecutive kernels alternate, so that each starts where the
             // previous one ended, on data the Infinity Cache still holds)
  // cells the sweep leaves out (another launch covers them): planes
  // [sk_z0, sk_z1] of rows [sk_x0, sk_x1] in y segments [sk_yb0, sk_yb1];
  // empty when sk_z0 > sk_z1
  int sk_z0, sk_z1, sk_x0, sk_x1, sk_yb0, sk_yb1;
  long blocks;
};
Sweep make_sweep(int z0, int z1, int x0, int x1, int y0, int y1, Tile t);

// Fused forward step (stress + velocity of one time step in one launch):
// workgroups of 64 x (R + 4) threads own R x rows of 60 y cells and Zc z
// planes; tile t = (zc * nxt + xt) * nys + ysg, t = (b % 8) * per_xcd + b / 8.
struct FusedSweep {
  int nys, nxt, nzc, Zc, tiles, per_xcd, zrev;
};

// stencil kernels (efwi_stencil.hip). F.w is the wavefield a forward kernel
// advances (the forward state or its replay copy); the backward kernels
// reverse F.w and advance the adjoint F.a.
//
// A RingFold with a slot folds the boundary ring into these kernels: the
// forward kernels also write the new F.w values of the band cells to the slot
// (the save of the state they produce), the backward kernels leave the slot's
// values in the band cells of the F.w fields they update (the restore that
// follows them). Only the flat sweep folds, and only when the band lies inside
// its cell box; the launchers return whether they did, and otherwise the
// caller moves the ring with launch_ring.
struct RingFold {
  Band B;
  float* slot;  // ring + slot index * 9 * B.n, or null: no ring traffic
};
bool ring_foldable(const Geom& G, const Band& B);
bool launch_stress_fwd(Tile t, const Geom& G, const Groups& F, const Prof& C, const Source& S, int it,
                       hipStream_t s, const RingFold* ring = nullptr);
bool launch_velocity_fwd(Tile t, int form, const Geom& G, const Groups& F, const Prof& C, hipStream_t s,
                         const RingFold* ring = nullptr);
// split: the wavefronts clear of every C-PML slab (rows inside the x and z
// interior, 64-cell y segments inside the y interior) run as a launch of
// their own on a kernel without the C-PML terms (fewer registers, more
// wavefronts per SIMD), the rest as the general kernel's launch that skips
// them. Flat sweeps only; results do not depend on it.
bool launch_velocity_bwd(Tile t, int form, const Geom& G, const Groups& F, const Prof& C, hipStream_t s,
                         const RingFold* ring = nullptr, bool split = false);
bool launch_stress_bwd(Tile t, const Geom& G, const Groups& F, const Prof& C, const Source& S, int it,
                       hipStream_t s, const RingFold* ring = nullptr, bool split = false);
// whether the grid has wavefronts clear of every C-PML slab
bool has_interior_segments(const Geom& G);
// One fused forward step from the state in Fi (w and p groups) to Fo (Fi.m
// holds the model); rows R in {4, 8, 12}, zc planes per workgroup. The source
// must lie inside the stencil box. Returns whether the ring save was folded.
bool launch_fwd_fused(int R, int zc, int form, const Geom& G, const Groups& Fi, const Groups& Fo, const Prof& C,
                      const Source& S, int it, hipStream_t s, const RingFold* ring = nullptr);
// adjoint-only steps (no reverse reconstruction): the first backward step
void launch_velocity_adj(int form, const Geom& G, const Groups& F, const Prof& C, hipStream_t s);
void launch_stress_adj(const Geom& G, const Groups& F, const Prof& C, hipStream_t s);

// batched persistent stencil kernels: one launch covers the (shot, tile) items
// of all shots of L, with the flat sweep's tiles (t.zdepth < 16).
// A BatchRing folds each shot's boundary ring (ShotDev::ring) into them, the
// slot following the launch's device step it: mode 1 (forward kernels) saves
// the state they produce into slot it + 1 while it + 1 <= last, mode 2
// (backward kernels) restores slot it. Callers pass one only where
// ring_foldable holds.
struct BatchRing {
  Band B;
  int mode;  // 0: no ring traffic
  int last;
};
void launch_stress_fwd_batch(Tile t, const Geom& G, const Prof& C, const BatchLaunch& L, hipStream_t s,
                             const BatchRing* ring = nullptr);
void launch_velocity_fwd_batch(Tile t, int form, const Geom& G, const Prof& C, const BatchLaunch& L, hipStream_t s,
                               const BatchRing* ring = nullptr);
void launch_velocity_bwd_batch(Tile t, int form, const Geom& G, const Prof& C, const BatchLaunch& L, hipStream_t s,
                               const BatchRing* ring = nullptr);
void launch_stress_bwd_batch(Tile t, const Geom& G, const Prof& C, const BatchLaunch& L, hipStream_t s,
                             const BatchRing* ring = nullptr);
// workgroups a persistent launch keeps resident per CU
int device_cus(int device);

// auxiliary kernels (efwi_aux.hip)
void launch_fill(float* p, long n, float v, hipStream_t s);
void launch_fill_hash(float* p, long n, unsigned seed, float scale, hipStream_t s);
void launch_illum(const float* vz, const float* vx, const float* vy, float* illum, long n, hipStream_t s);
void launch_byc(const Geom& G, const float* rho, float* byc, hipStream_t s);
void launch_add_source(float* szz, float* sxx, float* syy, long cell, const float* stf, int it,
                       float dt, int sign, hipStream_t s);
void launch_record(const Wave& W, const long* rec_cell, int nrec, int it, float* syn_t,
                   hipStream_t s);
void launch_inject(const Wave& A, const long* rec_cell, int nrec, int it, const float* res_t,
                   hipStream_t s);
void launch_ring(const Band& B, const Geom& G, const Fields9& F, float* ring, long slot, int k0,
                 int k1, int restore, hipStream_t s);
// obs_t/syn_t/res_t: [nt][3][nrec]; fac: [nt] (per_trace = 0) or [nt][nrec].
// Writes per-block double partial sums of r^2 to partials (nblocks entries) and
// returns nblocks.
int launch_misfit(const float* obs_t, const float* syn_t, float* res_t, const float* fac,
                  int per_trace, int nt, int nrec, double* partials, int max_blocks, hipStream_t s);
void launch_sum_partials(const double* partials, int n, double* out, hipStream_t s);
// max over the grid of (lambda + 2 mu) / rho (row padding skipped), through
// per-block partials (at most max_blocks), into *out (device)
void launch_vp2max(const Geom& G, const float* lam, const float* mu, const float* rho, double* partials,
                   int max_blocks, double* out, hipStream_t s);
// FP64 dot product of float vectors; returns the number of partials written.
int launch_dot64(const float* a, const float* b, long n, double* partials, int max_blocks, hipStream_t s);
// batched per-step aux kernels of the shot table (time step read on the device)
void launch_record_batch(const ShotDev* shots, int nshots, int max_nrec, const int* it, int it_offset,
                         hipStream_t s);
void launch_inject_batch(const ShotDev* shots, int nshots, int max_nrec, const int* it, hipStream_t s);
void launch_ring_batch(const Band& B, const Geom& G, const ShotDev* shots, int nshots, const int* it, int k0,
                       int k1, int restore, hipStream_t s);
void launch_step_add(int* it, int delta, hipStream_t s);
// [3][nrec][nt] <-> [nt][3][nrec]
void launch_to_time_major(const float* src, float* dst, int nrec, int nt, hipStream_t s);
void launch_from_time_major(const float* src, float* dst, int nrec, int nt, hipStream_t s);

}  // namespace efwi
