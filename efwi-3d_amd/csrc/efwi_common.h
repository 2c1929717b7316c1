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

// 1-D C-PML profiles per axis (0 = z, 1 = x, 2 = y): K, a, b on the integer
// grid and on the half grid.
struct Cpml {
  const float *K[3], *a[3], *b[3], *Kh[3], *ah[3], *bh[3];
};

// Source injected at one cell with amplitude stf_tapered[it] (device array).
struct Source {
  long cell;            // linear index, or -1 for none
  const float* stf;     // tapered source time function, device, nt samples
};

}  // namespace efwi
