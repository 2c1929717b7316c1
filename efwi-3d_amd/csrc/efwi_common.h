// Shared device-side descriptors of the elastic FWI library (gfx950).
//
// Memory layout of every volume: torch layout [nz][nx][ny] of the PML-padded
// grid, y fastest, with a row pitch `py` (floats) and a plane pitch `pxy`.
// The caller's tensors (model, gradients) have py = ny; internal state arrays
// use the same pitch so that the model can be read in place.
#pragma once

#include <hip/hip_runtime.h>
#include <cstdint>

namespace efwi {

// Grid geometry passed by value to every stencil kernel.
struct Geom {
  int nz, nx, ny, npml;
  int py;        // row pitch in floats
  long pxy;      // plane pitch in floats
  float dz, dx, dy, dt;
  unsigned fs;  // bytes between consecutive fields of a field group
};

// Field groups: each group is one allocation holding its fields back to back,
// field k at byte k * Geom::fs (fs is a multiple of 256 and every group stays
// below 4 GiB, so a cell's byte offset in a field fits 32 bits). The stencil
// kernels address them through one buffer resource per group.
enum WaveField { F_VZ, F_VX, F_VY, F_SZZ, F_SXX, F_SXZ, F_SYY, F_SXY, F_SYZ, F_WAVE_COUNT };
enum PsiField {
  P_VZ_Z, P_VZ_X, P_VZ_Y, P_VX_Z, P_VX_X, P_VX_Y, P_VY_Z, P_VY_X, P_VY_Y,
  P_SZZ_Z, P_SXZ_X, P_SXZ_Z, P_SXX_X, P_SXY_X, P_SXY_Y, P_SYY_Y, P_SYZ_Y, P_SYZ_Z, P_PSI_COUNT
};
enum ModelField { M_LAM, M_MU, M_BYC, M_GLAM, M_GMU, M_GRHO, M_COUNT };
struct Groups {
  float* w;  // F_WAVE_COUNT fields: forward (or replayed) wavefield
  float* a;  // F_WAVE_COUNT fields: adjoint wavefield
  float* p;  // P_PSI_COUNT fields: C-PML memory variables
  float* m;  // M_COUNT fields: lambda, mu, buoyancy and the three gradients
};

// C-PML profiles of the three axes (0 = z, 1 = x, 2 = y) in one array: axis
// `ax` holds kProfRows rows of n[ax] values starting at off[ax]: K, a, b on
// the integer grid, then Kh, ah, bh on the half grid.
enum ProfRow { R_K, R_A, R_B, R_KH, R_AH, R_BH, kProfRows };
struct Prof {
  const float* p;
  int off[3], n[3];
};

// The nine wavefield components (velocity + stress). Used for the forward
// wavefield and, separately, for the adjoint wavefield.
struct Wave {
  float *vz, *vx, *vy, *szz, *sxx, *sxz, *syy, *sxy, *syz;
};

// C-PML memory variables. The names follow the derivative they correct:
//   sv: psi of the stress kernel (derivatives of v)   -- Src/stress.cu
//   ss: psi of the velocity kernel (derivatives of s) -- Src/velocity.cu
// In the backward pass the same arrays hold the adjoint memory variables, as
// in the reference (Src/libCUFD.cu:481-498).
struct Psi {
  float *vz_z, *vz_x, *vz_y, *vx_z, *vx_x, *vx_y, *vy_z, *vy_x, *vy_y;
  float *szz_z, *sxz_x, *sxz_z, *sxx_x, *sxy_x, *sxy_y, *syy_y, *syz_y, *syz_z;
};

// Material arrays (lambda, mu, buoyancy) and gradient accumulators.
struct Model {
  const float *lam, *mu, *byc;
};
struct Grad {
  float *glam, *gmu, *grho;
};

// Source injected at one cell with amplitude stf_tapered[it] (device array).
struct Source {
  long cell;            // linear index, or -1 for none
  const float* stf;     // tapered source time function, device, nt samples
};

// One shot of a batch as the batched kernels read it (device table).
struct ShotDev {
  Groups F;           // the shot's field groups (its own model copy and gradients in F.m)
  long cell;          // source cell, inside the reverse-reconstruction interior
  const float* stf;   // tapered source time function [nt]
  float* w9[9];       // forward wavefield fields (ring, record), ring order vz vx vy szz sxx sxz syy sxy syz
  float* a9[9];       // adjoint wavefield fields (inject)
  const long* rec;    // receiver cells [nrec]
  int nrec;
  float* syn_t;       // [nt][3][nrec]
  const float* res_t; // [nt][3][nrec]
  float* ring;        // [nt-1][9][band]
};

// A batched launch: the shot table, the eight per-XCD work-queue heads (zeroed
// by a fill kernel before the launch) and the device-resident time step.
// Queue heads of a batched launch sit kQueueStride unsigned apart (one cache
// line each), 8 heads per launch.
constexpr int kQueueStride = 32;
struct BatchLaunch {
  const ShotDev* shots;
  int nshots;
  unsigned* counters;
  const int* it;
  int cus;   // compute units; persistent workgroups = cus x resident workgroups per CU
  int take;  // items taken per atomic
  int stat;  // launcher input: static share in percent; kernel input: static items per workgroup
  int reverse;  // visit the shots last to first (consecutive launches alternate, so that
                // each starts on the shot whose fields the previous one touched last)
};

}  // namespace efwi
