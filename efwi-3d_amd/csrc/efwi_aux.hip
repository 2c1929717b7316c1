// Auxiliary gfx950 kernels of the elastic FWI path: buoyancy, source
// injection, receiver gather and adjoint-source injection, boundary-band
// save/restore for reverse reconstruction, and the misfit with an FP64
// reduction. Compiled with -ffp-contract=off like the stencils.
#include "efwi_common.h"
#include "efwi_kernels.h"

namespace efwi {
namespace {

__global__ void k_fill(float* __restrict__ p, long n, float v) {
  for (long i = (long)blockIdx.x * blockDim.x + threadIdx.x; i < n; i += (long)gridDim.x * blockDim.x)
    p[i] = v;
}

// p[i] = scale * u(i, seed) with u uniform in [-1, 1) from an integer hash
// (values of a live wavefield's magnitude for timing runs)
__global__ void k_fill_hash(float* __restrict__ p, long n, unsigned seed, float scale) {
  for (long i = (long)blockIdx.x * blockDim.x + threadIdx.x; i < n; i += (long)gridDim.x * blockDim.x) {
    unsigned h = (unsigned)i * 0x9E3779B1u ^ seed;
    h ^= h >> 16;
    h *= 0x85EBCA6Bu;
    h ^= h >> 13;
    h *= 0xC2B2AE35u;
    h ^= h >> 16;
    p[i] = scale * ((float)(h >> 8) * (1.0f / 8388608.0f) - 1.0f);
  }
}

// Source illumination of the pseudo-Hessian (diagonal of J^T J for a point
// source, Shin et al. 2001): illum += vz^2 + vx^2 + vy^2 of the forward field
// of this step. No reference counterpart.
__global__ void k_illum(const float* __restrict__ vz, const float* __restrict__ vx, const float* __restrict__ vy,
                        float* __restrict__ illum, long n) {
  for (long i = (long)blockIdx.x * blockDim.x + threadIdx.x; i < n; i += (long)gridDim.x * blockDim.x) {
    const float a = vz[i], b = vx[i], c = vy[i];
    illum[i] += a * a + b * b + c * c;
  }
}

// Byc = 1/rho on [2, n-3]^3 and 1/1000 elsewhere (Src/Model.cu:65,
// Src/utilities.cu:117-128). The reference divides in double and stores
// float; a correctly rounded float division gives the same float because
// double carries more than 2*24+2 bits.
__global__ void k_byc(Geom G, const float* __restrict__ rho, float* __restrict__ byc) {
  const long n = (long)G.nz * G.pxy;
  for (long i = (long)blockIdx.x * blockDim.x + threadIdx.x; i < n; i += (long)gridDim.x * blockDim.x) {
    const int z = (int)(i / G.pxy);
    const long r = i - (long)z * G.pxy;
    const int x = (int)(r / G.py);
    const int y = (int)(r - (long)x * G.py);
    if (y >= G.ny) continue;
    const bool in = z >= 2 && z <= G.nz - 3 && x >= 2 && x <= G.nx - 3 && y >= 2 && y <= G.ny - 3;
    byc[i] = in ? 1.0f / rho[i] : (float)(1.0 / 1000.0);
  }
}

// Largest squared P velocity (lambda + 2 mu) / rho over the grid, per block
// (velInit, Src/utilities.cu:100-113, evaluates it in double on every cell;
// compCourantNumber, :186-202, takes the maximum). Non-finite quotients and
// the row padding are skipped; the per-block maxima go to `partials`.
__global__ __launch_bounds__(256) void k_vp2max(Geom G, const float* __restrict__ lam, const float* __restrict__ mu,
                                                const float* __restrict__ rho, double* __restrict__ partials) {
  const long n = (long)G.nz * G.pxy;
  double m = 0.0;
  for (long i = (long)blockIdx.x * blockDim.x + threadIdx.x; i < n; i += (long)gridDim.x * blockDim.x) {
    const int y = (int)(i % G.py);
    if (y >= G.ny) continue;
    const double v2 = ((double)lam[i] + 2.0 * (double)mu[i]) / (double)rho[i];
    if (v2 > m && v2 < 1e300) m = v2;
  }
  for (int off = 32; off > 0; off >>= 1) m = fmax(m, __shfl_down(m, off, 64));
  __shared__ double part[4];
  const int lane = threadIdx.x & 63, wid = threadIdx.x >> 6;
  if (lane == 0) part[wid] = m;
  __syncthreads();
  if (threadIdx.x == 0) partials[blockIdx.x] = fmax(fmax(part[0], part[1]), fmax(part[2], part[3]));
}
__global__ __launch_bounds__(256) void k_max_partials(const double* __restrict__ p, int n, double* out) {
  double m = 0.0;
  for (int k = threadIdx.x; k < n; k += blockDim.x) m = fmax(m, p[k]);
  for (int off = 32; off > 0; off >>= 1) m = fmax(m, __shfl_down(m, off, 64));
  __shared__ double part[4];
  const int lane = threadIdx.x & 63, wid = threadIdx.x >> 6;
  if (lane == 0) part[wid] = m;
  __syncthreads();
  if (threadIdx.x == 0) *out = fmax(fmax(part[0], part[1]), fmax(part[2], part[3]));
}

// Src/utilities.cu:442-479: +/- 1500^2 * stf[it] * dt at the source cell
// (the 9^3 amplitude cube is exactly 1 at its centre and 0 elsewhere).
__global__ void k_add_source(float* szz, float* sxx, float* syy, long cell, const float* stf, int it,
                             float dt, int sign) {
  if (threadIdx.x != 0) return;
  const float a = 2250000.0f * stf[it] * dt * 1.0f;
  if (sign > 0) {
    szz[cell] += a; sxx[cell] += a; syy[cell] += a;
  } else {
    szz[cell] -= a; sxx[cell] -= a; syy[cell] -= a;
  }
}

// Src/utilities.cu:481-495, time-major: syn_t[it][c][r], c = x, y, z.
__global__ void k_record(Wave W, const long* __restrict__ cell, int nrec, int it, float* __restrict__ syn_t) {
  const int r = blockIdx.x * blockDim.x + threadIdx.x;
  if (r >= nrec) return;
  const long i = cell[r];
  float* o = syn_t + (long)it * 3 * nrec;
  o[r] = W.vx[i];
  o[nrec + r] = W.vy[i];
  o[2 * nrec + r] = W.vz[i];
}

// Src/utilities.cu:497-507. The reference's += is not atomic; receivers that
// share a cell would race there. Here they add atomically.
__global__ void k_inject(Wave A, const long* __restrict__ cell, int nrec, int it, const float* __restrict__ res_t) {
  const int r = blockIdx.x * blockDim.x + threadIdx.x;
  if (r >= nrec) return;
  const long i = cell[r];
  const float* q = res_t + (long)it * 3 * nrec;
  atomicAdd(A.vx + i, q[r]);
  atomicAdd(A.vy + i, q[nrec + r]);
  atomicAdd(A.vz + i, q[2 * nrec + r]);
}

__device__ __forceinline__ long band_cell(const Band& B, const Geom& G, long k) {
  const int nzB = B.hi[0] - B.lo[0] + 1, nxB = B.hi[1] - B.lo[1] + 1, nyB = B.hi[2] - B.lo[2] + 1;
  int z, x, y;
  if (k < B.nzs) {
    const long per = 5L * nxB * nyB;
    const int slab = (int)(k / per);
    const long r = k - slab * per;
    const int zl = (int)(r / ((long)nxB * nyB));
    const long r2 = r - (long)zl * nxB * nyB;
    x = B.lo[1] + (int)(r2 / nyB);
    y = B.lo[2] + (int)(r2 % nyB);
    z = slab == 0 ? B.lo[0] + zl : B.hi[0] - 4 + zl;
  } else if (k < B.nzs + B.nxs) {
    k -= B.nzs;
    const int nzM = nzB - 10;
    const long per = 5L * nzM * nyB;
    const int slab = (int)(k / per);
    const long r = k - slab * per;
    const int zm = (int)(r / (5L * nyB));
    const long r2 = r - (long)zm * 5 * nyB;
    const int xl = (int)(r2 / nyB);
    y = B.lo[2] + (int)(r2 % nyB);
    z = B.lo[0] + 5 + zm;
    x = slab == 0 ? B.lo[1] + xl : B.hi[1] - 4 + xl;
  } else {
    k -= B.nzs + B.nxs;
    const int nzM = nzB - 10, nxM = nxB - 10;
    const long per = 5L * nzM * nxM;
    const int slab = (int)(k / per);
    const long r = k - slab * per;
    const int zm = (int)(r / (5L * nxM));
    const long r2 = r - (long)zm * 5 * nxM;
    const int xm = (int)(r2 / 5);
    const int yl = (int)(r2 % 5);
    z = B.lo[0] + 5 + zm;
    x = B.lo[1] + 5 + xm;
    y = slab == 0 ? B.lo[2] + yl : B.hi[2] - 4 + yl;
  }
  return (long)z * G.pxy + (long)x * G.py + y;
}

// from_bnd / to_bnd (Src/utilities.cu:323-420) for fields k0..k1-1 of one
// time slot: grid.y selects the field.
__global__ void k_ring(Band B, Geom G, Fields9 F, float* __restrict__ ring, long slot, int k0, int restore) {
  const int f = k0 + blockIdx.y;
  float* __restrict__ fld = F.f[f];
  float* __restrict__ dst = ring + (slot * 9 + f) * B.n;
  for (long k = (long)blockIdx.x * blockDim.x + threadIdx.x; k < B.n; k += (long)gridDim.x * blockDim.x) {
    const long i = band_cell(B, G, k);
    if (restore) fld[i] = dst[k];
    else dst[k] = fld[i];
  }
}

// Misfit and adjoint source (Src/libCUFD.cu:371-441 with the windows of
// Src/utilities.cu:524-614): r = obs*w - syn*w (r = 0 at it = 0), J += r^2 in
// double, res = r*w. One element per thread over [nt][3][nrec].
__global__ __launch_bounds__(256) void k_misfit(const float* __restrict__ obs_t, const float* __restrict__ syn_t,
                                                float* __restrict__ res_t, const float* __restrict__ fac,
                                                int per_trace, int nt, int nrec, double* __restrict__ partials) {
  const long n = (long)nt * 3 * nrec;
  double acc = 0.0;
  for (long k = (long)blockIdx.x * blockDim.x + threadIdx.x; k < n; k += (long)gridDim.x * blockDim.x) {
    const int it = (int)(k / (3L * nrec));
    const int r = (int)(k % nrec);
    const float w = per_trace ? fac[(long)it * nrec + r] : fac[it];
    float o = obs_t[k] * w;
    float sy = syn_t[k] * w;
    float e = it == 0 ? 0.0f : o - sy;
    acc += (double)e * (double)e;
    res_t[k] = e * w;
  }
  // wavefront reduction, then across the 4 wavefronts of the workgroup
  for (int off = 32; off > 0; off >>= 1) acc += __shfl_down(acc, off, 64);
  __shared__ double part[4];
  const int lane = threadIdx.x & 63, wid = threadIdx.x >> 6;
  if (lane == 0) part[wid] = acc;
  __syncthreads();
  if (threadIdx.x == 0) partials[blockIdx.x] = (part[0] + part[1]) + (part[2] + part[3]);
}

// FP64 dot product of two float vectors: per-block partial sums in double.
__global__ __launch_bounds__(256) void k_dot64(const float* __restrict__ a, const float* __restrict__ b, long n,
                                               double* __restrict__ partials) {
  double acc = 0.0;
  for (long k = (long)blockIdx.x * blockDim.x + threadIdx.x; k < n; k += (long)gridDim.x * blockDim.x)
    acc += (double)a[k] * (double)b[k];
  for (int off = 32; off > 0; off >>= 1) acc += __shfl_down(acc, off, 64);
  __shared__ double part[4];
  const int lane = threadIdx.x & 63, wid = threadIdx.x >> 6;
  if (lane == 0) part[wid] = acc;
  __syncthreads();
  if (threadIdx.x == 0) partials[blockIdx.x] = (part[0] + part[1]) + (part[2] + part[3]);
}

__global__ __launch_bounds__(256) void k_sum_partials(const double* __restrict__ p, int n, double* out) {
  double acc = 0.0;
  for (int k = threadIdx.x; k < n; k += blockDim.x) acc += p[k];
  for (int off = 32; off > 0; off >>= 1) acc += __shfl_down(acc, off, 64);
  __shared__ double part[4];
  const int lane = threadIdx.x & 63, wid = threadIdx.x >> 6;
  if (lane == 0) part[wid] = acc;
  __syncthreads();
  if (threadIdx.x == 0) *out = (part[0] + part[1]) + (part[2] + part[3]);
}

// [3][nrec][nt] -> [nt][3][nrec] through a 64x64 LDS tile per (component, block).
__global__ __launch_bounds__(256) void k_transpose(const float* __restrict__ src, float* __restrict__ dst,
                                                   int rows, int cols) {
  // src is [comp][rows][cols]; dst is [cols][comp][rows] with 3 components.
  __shared__ float tile[64][65];
  const int c = blockIdx.z;
  const int r0 = blockIdx.y * 64, c0 = blockIdx.x * 64;
  const float* s = src + (long)c * rows * cols;
  for (int j = threadIdx.y; j < 64; j += 4) {
    const int r = r0 + j, cc = c0 + threadIdx.x;
    if (r < rows && cc < cols) tile[j][threadIdx.x] = s[(long)r * cols + cc];
  }
  __syncthreads();
  for (int j = threadIdx.y; j < 64; j += 4) {
    const int cc = c0 + j, r = r0 + threadIdx.x;
    if (r < rows && cc < cols) dst[((long)cc * 3 + c) * rows + r] = tile[threadIdx.x][j];
  }
}

// [nt][3][nrec] -> [3][nrec][nt]
__global__ __launch_bounds__(256) void k_transpose_back(const float* __restrict__ src, float* __restrict__ dst,
                                                        int nrec, int nt) {
  __shared__ float tile[64][65];
  const int c = blockIdx.z;
  const int t0 = blockIdx.y * 64, r0 = blockIdx.x * 64;
  for (int j = threadIdx.y; j < 64; j += 4) {
    const int t = t0 + j, r = r0 + threadIdx.x;
    if (t < nt && r < nrec) tile[j][threadIdx.x] = src[((long)t * 3 + c) * nrec + r];
  }
  __syncthreads();
  for (int j = threadIdx.y; j < 64; j += 4) {
    const int r = r0 + j, t = t0 + threadIdx.x;
    if (t < nt && r < nrec) dst[((long)c * nrec + r) * nt + t] = tile[threadIdx.x][j];
  }
}

inline int blocks_for(long n, int threads, int cap) {
  long b = (n + threads - 1) / threads;
  if (b > cap) b = cap;
  if (b < 1) b = 1;
  return (int)b;
}

// Batched per-step kernels over the shot table (grid.y or grid.z = shot); the
// time step is read from device memory so that a captured step replays.
__device__ __forceinline__ int step_of(const int* it) {
  return __hip_atomic_load(it, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
}
// receiver gather of step *it + off, as k_record
__global__ void k_record_b(const ShotDev* __restrict__ shots, const int* it_dev, int off) {
  const ShotDev& sd = shots[blockIdx.y];
  const int r = blockIdx.x * blockDim.x + threadIdx.x;
  if (r >= sd.nrec) return;
  const int it = step_of(it_dev) + off;
  const long i = sd.rec[r];
  float* o = sd.syn_t + (long)it * 3 * sd.nrec;
  o[r] = sd.w9[1][i];
  o[sd.nrec + r] = sd.w9[2][i];
  o[2 * sd.nrec + r] = sd.w9[0][i];
}
// adjoint-source injection of step *it, as k_inject
__global__ void k_inject_b(const ShotDev* __restrict__ shots, const int* it_dev) {
  const ShotDev& sd = shots[blockIdx.y];
  const int r = blockIdx.x * blockDim.x + threadIdx.x;
  if (r >= sd.nrec) return;
  const int it = step_of(it_dev);
  const long i = sd.rec[r];
  const float* q = sd.res_t + (long)it * 3 * sd.nrec;
  atomicAdd(sd.a9[1] + i, q[r]);
  atomicAdd(sd.a9[2] + i, q[sd.nrec + r]);
  atomicAdd(sd.a9[0] + i, q[2 * sd.nrec + r]);
}
// ring slot *it of fields k0 + blockIdx.y, shot blockIdx.z, as k_ring
__global__ void k_ring_b(Band B, Geom G, const ShotDev* __restrict__ shots, const int* it_dev, int k0, int restore) {
  const ShotDev& sd = shots[blockIdx.z];
  const long slot = step_of(it_dev);
  const int f = k0 + blockIdx.y;
  float* __restrict__ fld = sd.w9[f];
  float* __restrict__ dst = sd.ring + (slot * 9 + f) * B.n;
  for (long k = (long)blockIdx.x * blockDim.x + threadIdx.x; k < B.n; k += (long)gridDim.x * blockDim.x) {
    const long i = band_cell(B, G, k);
    if (restore) fld[i] = dst[k];
    else dst[k] = fld[i];
  }
}
__global__ void k_step_add(int* it, int delta) {
  if (threadIdx.x == 0) *it += delta;
}

}  // namespace

void launch_fill(float* p, long n, float v, hipStream_t s) {
  hipLaunchKernelGGL(k_fill, dim3(blocks_for(n, 256, 8192)), dim3(256), 0, s, p, n, v);
}
void launch_fill_hash(float* p, long n, unsigned seed, float scale, hipStream_t s) {
  hipLaunchKernelGGL(k_fill_hash, dim3(blocks_for(n, 256, 8192)), dim3(256), 0, s, p, n, seed, scale);
}
void launch_illum(const float* vz, const float* vx, const float* vy, float* illum, long n, hipStream_t s) {
  hipLaunchKernelGGL(k_illum, dim3(blocks_for(n, 256, 8192)), dim3(256), 0, s, vz, vx, vy, illum, n);
}
void launch_byc(const Geom& G, const float* rho, float* byc, hipStream_t s) {
  hipLaunchKernelGGL(k_byc, dim3(blocks_for((long)G.nz * G.pxy, 256, 8192)), dim3(256), 0, s, G, rho, byc);
}
void launch_add_source(float* szz, float* sxx, float* syy, long cell, const float* stf, int it,
                       float dt, int sign, hipStream_t s) {
  hipLaunchKernelGGL(k_add_source, dim3(1), dim3(64), 0, s, szz, sxx, syy, cell, stf, it, dt, sign);
}
void launch_record(const Wave& W, const long* rec_cell, int nrec, int it, float* syn_t, hipStream_t s) {
  if (nrec <= 0) return;
  hipLaunchKernelGGL(k_record, dim3((nrec + 255) / 256), dim3(256), 0, s, W, rec_cell, nrec, it, syn_t);
}
void launch_inject(const Wave& A, const long* rec_cell, int nrec, int it, const float* res_t, hipStream_t s) {
  if (nrec <= 0) return;
  hipLaunchKernelGGL(k_inject, dim3((nrec + 255) / 256), dim3(256), 0, s, A, rec_cell, nrec, it, res_t);
}
void launch_ring(const Band& B, const Geom& G, const Fields9& F, float* ring, long slot, int k0, int k1,
                 int restore, hipStream_t s) {
  dim3 grid(blocks_for(B.n, 256, 2048), k1 - k0);
  hipLaunchKernelGGL(k_ring, grid, dim3(256), 0, s, B, G, F, ring, slot, k0, restore);
}
int launch_misfit(const float* obs_t, const float* syn_t, float* res_t, const float* fac, int per_trace,
                  int nt, int nrec, double* partials, int max_blocks, hipStream_t s) {
  const int nb = blocks_for((long)nt * 3 * nrec, 256, max_blocks);
  hipLaunchKernelGGL(k_misfit, dim3(nb), dim3(256), 0, s, obs_t, syn_t, res_t, fac, per_trace, nt, nrec,
                     partials);
  return nb;
}
int launch_dot64(const float* a, const float* b, long n, double* partials, int max_blocks, hipStream_t s) {
  const int nb = blocks_for(n, 256, max_blocks);
  hipLaunchKernelGGL(k_dot64, dim3(nb), dim3(256), 0, s, a, b, n, partials);
  return nb;
}

void launch_vp2max(const Geom& G, const float* lam, const float* mu, const float* rho, double* partials,
                   int max_blocks, double* out, hipStream_t s) {
  const int nb = blocks_for((long)G.nz * G.pxy, 256, max_blocks);
  hipLaunchKernelGGL(k_vp2max, dim3(nb), dim3(256), 0, s, G, lam, mu, rho, partials);
  hipLaunchKernelGGL(k_max_partials, dim3(1), dim3(256), 0, s, partials, nb, out);
}

void launch_sum_partials(const double* partials, int n, double* out, hipStream_t s) {
  hipLaunchKernelGGL(k_sum_partials, dim3(1), dim3(256), 0, s, partials, n, out);
}
void launch_to_time_major(const float* src, float* dst, int nrec, int nt, hipStream_t s) {
  dim3 grid((nt + 63) / 64, (nrec + 63) / 64, 3);
  hipLaunchKernelGGL(k_transpose, grid, dim3(64, 4), 0, s, src, dst, nrec, nt);
}
void launch_from_time_major(const float* src, float* dst, int nrec, int nt, hipStream_t s) {
  dim3 grid((nrec + 63) / 64, (nt + 63) / 64, 3);
  hipLaunchKernelGGL(k_transpose_back, grid, dim3(64, 4), 0, s, src, dst, nrec, nt);
}

void launch_record_batch(const ShotDev* shots, int nshots, int max_nrec, const int* it, int it_offset,
                         hipStream_t s) {
  if (max_nrec <= 0) return;
  hipLaunchKernelGGL(k_record_b, dim3((max_nrec + 255) / 256, nshots), dim3(256), 0, s, shots, it, it_offset);
}
void launch_inject_batch(const ShotDev* shots, int nshots, int max_nrec, const int* it, hipStream_t s) {
  if (max_nrec <= 0) return;
  hipLaunchKernelGGL(k_inject_b, dim3((max_nrec + 255) / 256, nshots), dim3(256), 0, s, shots, it);
}
void launch_ring_batch(const Band& B, const Geom& G, const ShotDev* shots, int nshots, const int* it, int k0,
                       int k1, int restore, hipStream_t s) {
  dim3 grid(blocks_for(B.n, 256, 512), k1 - k0, nshots);
  hipLaunchKernelGGL(k_ring_b, grid, dim3(256), 0, s, B, G, shots, it, k0, restore);
}
void launch_step_add(int* it, int delta, hipStream_t s) {
  hipLaunchKernelGGL(k_step_add, dim3(1), dim3(64), 0, s, it, delta);
}

}  // namespace efwi
