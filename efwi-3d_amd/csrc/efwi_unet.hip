// Kernels of the learned preconditioner: a 3-D U-Net with 3x3x3 convolutions,
// BF16 weights and activations, FP32 accumulation, on gfx950 matrix cores.
//
// Tensors are channels-last volumes: element (voxel v, channel c) at
// p[v * ld + c], voxel v = (z * nx + x) * ny + y (y fastest, as the FWI grid),
// ld (the channel stride) a multiple of 16, so a tensor can be a channel slice
// of a wider one (the U-Net's skip concatenations are written in place).
//
// Convolution as implicit GEMM on v_mfma_f32_16x16x16_bf16 (D = A B + C,
// 16x16x16, one wavefront). Lane l holds A[l % 16][4 (l / 16) + j],
// B[4 (l / 16) + j][l % 16] and D[4 (l / 16) + j][l % 16], j = 0..3. A
// wavefront owns 16 consecutive voxels along y: A rows are those voxels, the
// K dimension runs over the 27 taps x input channels in chunks of 16, and B
// columns are output channels, 16 per accumulator block.
//   forward:  out[v][n] = sum_t sum_k in[v + d_t][k] w[t][n][k]      (+ bias, ReLU)
//   dgrad:    din[v][n] = sum_t sum_k dout[v - d_t][k] w[t][k][n]    (x ReLU mask)
//   wgrad:    dw[t][m][n] = sum_v dpre[v][m] in[v + d_t][n]          (K = voxels)
// with zero padding at the volume faces. No reference counterpart: the
// reference has no learned part.
#include <hip/hip_runtime.h>

#include <algorithm>
#include <cstdint>
#include <mutex>

#include "../../include/efwi_nn.h"

namespace {

typedef short s4 __attribute__((ext_vector_type(4)));
typedef float f4 __attribute__((ext_vector_type(4)));

__device__ __forceinline__ float bf2f(uint16_t h) { return __uint_as_float((unsigned)h << 16); }
// round to nearest even (finite inputs)
__device__ __forceinline__ uint16_t f2bf(float f) {
  unsigned u = __float_as_uint(f);
  u += 0x7FFFu + ((u >> 16) & 1u);
  return (uint16_t)(u >> 16);
}

struct Vol {
  int nz, nx, ny;
  __device__ __forceinline__ long vox(int z, int x, int y) const { return ((long)z * nx + x) * ny + y; }
  __device__ __forceinline__ bool in(int z, int x, int y) const {
    return z >= 0 && z < nz && x >= 0 && x < nx && y >= 0 && y < ny;
  }
};

__device__ __forceinline__ void tap(int t, int& dz, int& dx, int& dy) {
  dz = t / 9 - 1;
  dx = (t / 3) % 3 - 1;
  dy = t % 3 - 1;
}

// One wavefront per row segment of 16 MB voxels along y (MB row blocks of the
// MFMA's A operand), NB output blocks of 16 channels starting at block
// blockIdx.y * NB: every B fragment (weights) is used MB times, every A
// fragment NB times. MODE 0 forward, 1 dgrad.
template <int NB, int MB, int MODE>
__global__ __launch_bounds__(256) void k_conv3(const uint16_t* __restrict__ src, int lds, int kc,
                                               const uint16_t* __restrict__ w, const float* __restrict__ bias,
                                               uint16_t* __restrict__ dst, int ldd, int nc,
                                               const uint16_t* __restrict__ mask, int ldm, Vol V, int relu) {
  const int lane = threadIdx.x & 63;
  const long tile = (long)blockIdx.x * 4 + (threadIdx.x >> 6);
  const int tpr = (V.ny + 16 * MB - 1) / (16 * MB);
  const long rows = (long)V.nz * V.nx;
  if (tile >= rows * tpr) return;  // wavefront-uniform
  const int row = (int)(tile / tpr);
  const int y0 = (int)(tile - (long)row * tpr) * 16 * MB;
  const int z = row / V.nx, x = row - z * V.nx;
  const int i = lane & 15, kq = (lane >> 4) * 4;
  const int nb0 = blockIdx.y * NB, nbt = nc / 16;
  f4 acc[MB][NB];
#pragma unroll
  for (int m = 0; m < MB; m++)
#pragma unroll
    for (int b = 0; b < NB; b++) acc[m][b] = f4{0.f, 0.f, 0.f, 0.f};
  const int sgn = MODE == 0 ? 1 : -1;
  for (int t = 0; t < 27; t++) {
    int dz, dx, dy;
    tap(t, dz, dx, dy);
    const int zs = z + sgn * dz, xs = x + sgn * dx;
    bool ok[MB];
    long vs[MB];
#pragma unroll
    for (int m = 0; m < MB; m++) {
      const int ys = y0 + 16 * m + i + sgn * dy;
      ok[m] = V.in(zs, xs, ys);
      vs[m] = ok[m] ? V.vox(zs, xs, ys) : 0;
    }
    for (int kb = 0; kb < kc; kb += 16) {
      s4 a[MB];
#pragma unroll
      for (int m = 0; m < MB; m++) {
        a[m] = s4{0, 0, 0, 0};
        if (ok[m]) a[m] = *reinterpret_cast<const s4*>(src + vs[m] * lds + kb + kq);
      }
#pragma unroll
      for (int b = 0; b < NB; b++) {
        if (nb0 + b >= nbt) break;
        const int n = (nb0 + b) * 16 + i;
        s4 bw;
        if (MODE == 0) {
          bw = *reinterpret_cast<const s4*>(w + ((long)t * nc + n) * kc + kb + kq);  // w[t][n][k]
        } else {
#pragma unroll
          for (int j = 0; j < 4; j++) bw[j] = (short)w[((long)t * kc + kb + kq + j) * nc + n];  // w[t][k][n]
        }
#pragma unroll
        for (int m = 0; m < MB; m++) acc[m][b] = __builtin_amdgcn_mfma_f32_16x16x16bf16_1k(a[m], bw, acc[m][b], 0, 0, 0);
      }
    }
  }
#pragma unroll
  for (int b = 0; b < NB; b++) {
    if (nb0 + b >= nbt) break;
    const int n = (nb0 + b) * 16 + i;
    const float bn = bias ? bias[n] : 0.f;
#pragma unroll
    for (int m = 0; m < MB; m++)
#pragma unroll
      for (int j = 0; j < 4; j++) {
        const int y = y0 + 16 * m + kq + j;
        if (y >= V.ny) continue;
        const long v = V.vox(z, x, y);
        float r = acc[m][b][j] + bn;
        if (relu) r = fmaxf(r, 0.f);
        if (mask && !(bf2f(mask[v * ldm + n]) > 0.f)) r = 0.f;
        dst[v * ldd + n] = f2bf(r);
      }
  }
}

// Weight gradient partials, operands staged through LDS. A block of three
// wavefronts takes `per` row segments of 16 voxels (a chunk), one 16-channel
// block of m and NB 16-channel blocks of n; wavefront w takes the 9 taps with
// dz = w - 1. Per segment the block loads the dpre tile (16 voxels x 16 m)
// and each wavefront its three rows of src (dx = -1, 0, 1; 18 voxels along y
// for dy = -1, 0, 1) with 16-byte loads, transposes them into LDS (the src
// rows once per dy shift, so every MFMA operand is one aligned 8-byte LDS
// read), and runs 9 x NB MFMAs on one A fragment. Writes
// part[chunk][t][m][n] for its taps, m and n.
template <int NB>
__global__ __launch_bounds__(192) void k_wgrad(const uint16_t* __restrict__ dpre, int ldp, int mc,
                                               const uint16_t* __restrict__ src, int lds, int kc,
                                               float* __restrict__ part, Vol V, int per) {
  __shared__ __align__(16) uint16_t At[16 * 16];             // [m][voxel]
  __shared__ __align__(16) uint16_t Bt[3][3][3][NB * 16][16];  // [dz][dx][dy][n][voxel]
  const int tid = threadIdx.x, wv = tid >> 6, lane = tid & 63;
  const int chunk = blockIdx.x, mb = blockIdx.y, nb0 = blockIdx.z * NB;
  const int nbl = min(NB, kc / 16 - nb0);  // n blocks of this block (uniform)
  const int nvec = nbl * 2;                // 8-channel vectors per voxel
  const int i = lane & 15, kq = (lane >> 4) * 4;
  const int tpr = (V.ny + 15) / 16;
  const long ntiles = (long)V.nz * V.nx * tpr;
  const int dz = wv - 1;
  f4 acc[3][3][NB];
#pragma unroll
  for (int a = 0; a < 3; a++)
#pragma unroll
    for (int b = 0; b < 3; b++)
#pragma unroll
      for (int c = 0; c < NB; c++) acc[a][b][c] = f4{0.f, 0.f, 0.f, 0.f};
  const long t0 = (long)chunk * per, t1 = std::min(ntiles, t0 + per);
  for (long tile = t0; tile < t1; tile++) {
    const int row = (int)(tile / tpr);
    const int y0 = (int)(tile - (long)row * tpr) * 16;
    const int z = row / V.nx, x = row - z * V.nx;
    // A tile: voxel v, channels h..h+7 of the m block (threads 0..31)
    if (tid < 32) {
      const int v = tid >> 1, h = (tid & 1) * 8;
      uint4 q = {0u, 0u, 0u, 0u};
      if (y0 + v < V.ny) q = *reinterpret_cast<const uint4*>(dpre + V.vox(z, x, y0 + v) * ldp + mb * 16 + h);
      const uint16_t* e = reinterpret_cast<const uint16_t*>(&q);
#pragma unroll
      for (int j = 0; j < 8; j++) At[(h + j) * 16 + v] = e[j];
    }
    // this wavefront's src rows: voxel vv (y = y0 - 1 + vv) is operand column
    // k = vv - dyi of shift dyi (dy = dyi - 1)
    const int items = 3 * 18 * nvec;
    for (int it = lane; it < items; it += 64) {
      const int dxi = it / (18 * nvec), r = it - dxi * 18 * nvec;
      const int vv = r / nvec, c8 = (r - vv * nvec) * 8;
      const int zs = z + dz, xs = x + dxi - 1, ys = y0 - 1 + vv;
      uint4 q = {0u, 0u, 0u, 0u};
      if (V.in(zs, xs, ys)) q = *reinterpret_cast<const uint4*>(src + V.vox(zs, xs, ys) * lds + nb0 * 16 + c8);
      const uint16_t* e = reinterpret_cast<const uint16_t*>(&q);
#pragma unroll
      for (int dyi = 0; dyi < 3; dyi++) {
        const int k = vv - dyi;
        if (k >= 0 && k < 16) {
#pragma unroll
          for (int j = 0; j < 8; j++) Bt[wv][dxi][dyi][c8 + j][k] = e[j];
        }
      }
    }
    __syncthreads();
    const s4 a = *reinterpret_cast<const s4*>(&At[i * 16 + kq]);  // A[m][k] = dpre[voxel k][m]
#pragma unroll
    for (int dxi = 0; dxi < 3; dxi++)
#pragma unroll
      for (int dyi = 0; dyi < 3; dyi++)
#pragma unroll
        for (int b = 0; b < NB; b++) {
          if (b >= nbl) break;
          // B[k][n] = src[voxel k + d_t][n]
          const s4 bw = *reinterpret_cast<const s4*>(&Bt[wv][dxi][dyi][b * 16 + i][kq]);
          acc[dxi][dyi][b] = __builtin_amdgcn_mfma_f32_16x16x16bf16_1k(a, bw, acc[dxi][dyi][b], 0, 0, 0);
        }
    __syncthreads();
  }
  // D[m][n]: lane holds m = mb*16 + kq + j, n = (nb0 + b)*16 + i
#pragma unroll
  for (int dxi = 0; dxi < 3; dxi++)
#pragma unroll
    for (int dyi = 0; dyi < 3; dyi++) {
      const int t = 9 * wv + 3 * dxi + dyi;
      float* out = part + ((long)chunk * 27 + t) * mc * kc;
#pragma unroll
      for (int b = 0; b < NB; b++) {
        if (b >= nbl) break;
#pragma unroll
        for (int j = 0; j < 4; j++) out[(long)(mb * 16 + kq + j) * kc + (nb0 + b) * 16 + i] = acc[dxi][dyi][b][j];
      }
    }
}

// out[k] += sum over chunks [blockIdx.y * kChunkGroup, +kChunkGroup) of part[chunk][k]
// (out cleared beforehand; one atomic per element per chunk group)
constexpr int kChunkGroup = 32;
__global__ void k_sum_chunks(const float* __restrict__ part, long width, int nchunks, float* __restrict__ out) {
  const long k = (long)blockIdx.x * blockDim.x + threadIdx.x;
  if (k >= width) return;
  const int c0 = blockIdx.y * kChunkGroup, c1 = min(nchunks, c0 + kChunkGroup);
  float s = 0.f;
  for (int c = c0; c < c1; c++) s += part[(long)c * width + k];
  atomicAdd(out + k, s);
}

// out[c] += sum_v x[v][c] (bias gradients; out cleared beforehand). blockDim
// 256 = (256 / c voxel lanes) x c channels; the lanes are reduced in LDS and
// each block adds one value per channel.
__global__ void k_chan_sum(const uint16_t* __restrict__ x, int ld, int c, long nvox, float* __restrict__ out) {
  __shared__ float red[256];
  const int lanes = 256 / c;
  const int ch = threadIdx.x % c, vl = threadIdx.x / c;
  float s = 0.f;
  if (vl < lanes)
    for (long v = (long)blockIdx.x * lanes + vl; v < nvox; v += (long)gridDim.x * lanes) s += bf2f(x[v * ld + ch]);
  red[threadIdx.x] = s;
  __syncthreads();
  if (threadIdx.x < c) {
    float t = 0.f;
    for (int q = 0; q < lanes; q++) t += red[q * c + threadIdx.x];
    atomicAdd(out + threadIdx.x, t);
  }
}

// 2x2x2 average pooling; (nz, nx, ny) are the fine dimensions (even)
__global__ void k_pool2(const uint16_t* __restrict__ src, int lds, uint16_t* __restrict__ dst, int ldd, int c, Vol F) {
  const Vol Cv{F.nz / 2, F.nx / 2, F.ny / 2};
  const long n = (long)Cv.nz * Cv.nx * Cv.ny * c;
  for (long k = (long)blockIdx.x * blockDim.x + threadIdx.x; k < n; k += (long)gridDim.x * blockDim.x) {
    const int ch = (int)(k % c);
    long v = k / c;
    const int y = (int)(v % Cv.ny); v /= Cv.ny;
    const int x = (int)(v % Cv.nx);
    const int z = (int)(v / Cv.nx);
    float s = 0.f;
    for (int q = 0; q < 8; q++) s += bf2f(src[F.vox(2 * z + (q >> 2), 2 * x + ((q >> 1) & 1), 2 * y + (q & 1)) * lds + ch]);
    dst[Cv.vox(z, x, y) * ldd + ch] = f2bf(0.125f * s);
  }
}
// gradient of the pooling: d_fine[v][c] (=|+=) d_coarse[parent][c] / 8
__global__ void k_pool2_bwd(const uint16_t* __restrict__ dc, int ldc, uint16_t* __restrict__ df, int ldf, int c, Vol F,
                            int accumulate) {
  const long n = (long)F.nz * F.nx * F.ny * c;
  for (long k = (long)blockIdx.x * blockDim.x + threadIdx.x; k < n; k += (long)gridDim.x * blockDim.x) {
    const int ch = (int)(k % c);
    long v = k / c;
    const int y = (int)(v % F.ny); v /= F.ny;
    const int x = (int)(v % F.nx);
    const int z = (int)(v / F.nx);
    const Vol Cv{F.nz / 2, F.nx / 2, F.ny / 2};
    float g = 0.125f * bf2f(dc[Cv.vox(z / 2, x / 2, y / 2) * ldc + ch]);
    const long o = F.vox(z, x, y) * ldf + ch;
    if (accumulate) g += bf2f(df[o]);
    df[o] = f2bf(g);
  }
}
// nearest upsampling x2: fine[v][c] = coarse[parent][c]
__global__ void k_up2(const uint16_t* __restrict__ cs, int ldc, uint16_t* __restrict__ fd, int ldf, int c, Vol F) {
  const long n = (long)F.nz * F.nx * F.ny * c;
  const Vol Cv{F.nz / 2, F.nx / 2, F.ny / 2};
  for (long k = (long)blockIdx.x * blockDim.x + threadIdx.x; k < n; k += (long)gridDim.x * blockDim.x) {
    const int ch = (int)(k % c);
    long v = k / c;
    const int y = (int)(v % F.ny); v /= F.ny;
    const int x = (int)(v % F.nx);
    const int z = (int)(v / F.nx);
    fd[F.vox(z, x, y) * ldf + ch] = cs[Cv.vox(z / 2, x / 2, y / 2) * ldc + ch];
  }
}
// gradient of the upsampling: d_coarse[p][c] = sum of the 8 children, times the ReLU mask of the coarse activation
__global__ void k_up2_bwd(const uint16_t* __restrict__ df, int ldf, uint16_t* __restrict__ dc, int ldc, int c, Vol F,
                          const uint16_t* __restrict__ mask, int ldm) {
  const Vol Cv{F.nz / 2, F.nx / 2, F.ny / 2};
  const long n = (long)Cv.nz * Cv.nx * Cv.ny * c;
  for (long k = (long)blockIdx.x * blockDim.x + threadIdx.x; k < n; k += (long)gridDim.x * blockDim.x) {
    const int ch = (int)(k % c);
    long v = k / c;
    const int y = (int)(v % Cv.ny); v /= Cv.ny;
    const int x = (int)(v % Cv.nx);
    const int z = (int)(v / Cv.nx);
    float s = 0.f;
    for (int q = 0; q < 8; q++) s += bf2f(df[F.vox(2 * z + (q >> 2), 2 * x + ((q >> 1) & 1), 2 * y + (q & 1)) * ldf + ch]);
    const long o = Cv.vox(z, x, y);
    if (mask && !(bf2f(mask[o * ldm + ch]) > 0.f)) s = 0.f;
    dc[o * ldc + ch] = f2bf(s);
  }
}
// a[v][c] = (a[v][c] + b[v][c]) x (mask[v][c] > 0 when given)
__global__ void k_add_masked(uint16_t* __restrict__ a, int lda, const uint16_t* __restrict__ b, int ldb, int c, long nvox,
                             const uint16_t* __restrict__ mask, int ldm) {
  const long n = nvox * c;
  for (long k = (long)blockIdx.x * blockDim.x + threadIdx.x; k < n; k += (long)gridDim.x * blockDim.x) {
    const long v = k / c;
    const int ch = (int)(k - v * c);
    float s = bf2f(a[v * lda + ch]) + (b ? bf2f(b[v * ldb + ch]) : 0.f);
    if (mask && !(bf2f(mask[v * ldm + ch]) > 0.f)) s = 0.f;
    a[v * lda + ch] = f2bf(s);
  }
}

// 1x1x1 output head in FP32: out[v] = sum_c w[c] act[v][c] + b
__global__ void k_head(const uint16_t* __restrict__ act, int ld, int c, const float* __restrict__ w, const float* __restrict__ b,
                       float* __restrict__ out, long nvox) {
  for (long v = (long)blockIdx.x * blockDim.x + threadIdx.x; v < nvox; v += (long)gridDim.x * blockDim.x) {
    float s = b[0];
    for (int ch = 0; ch < c; ch++) s += w[ch] * bf2f(act[v * ld + ch]);
    out[v] = s;
  }
}
// head backward: dact[v][c] = w[c] dout[v] (x ReLU mask of act); dw[c] += sum_v act dout; db += sum_v dout
__global__ void k_head_bwd(const uint16_t* __restrict__ act, int ld, int c, const float* __restrict__ w,
                           const float* __restrict__ dout, uint16_t* __restrict__ dact, int ldd, float* __restrict__ dw,
                           float* __restrict__ db, long nvox) {
  __shared__ float red[17][64];
  float acc[17];
  for (int k = 0; k < 17; k++) acc[k] = 0.f;
  for (long v = (long)blockIdx.x * blockDim.x + threadIdx.x; v < nvox; v += (long)gridDim.x * blockDim.x) {
    const float g = dout[v];
    for (int ch = 0; ch < c; ch++) {
      const float a = bf2f(act[v * ld + ch]);
      dact[v * ldd + ch] = f2bf(a > 0.f ? w[ch] * g : 0.f);
      acc[ch] += a * g;
    }
    acc[16] += g;
  }
  // blockDim.x == 64
  for (int k = 0; k < 17; k++) red[k][threadIdx.x] = acc[k];
  __syncthreads();
  if (threadIdx.x < 17) {
    float s = 0.f;
    for (int q = 0; q < 64; q++) s += red[threadIdx.x][q];
    if (threadIdx.x < c) atomicAdd(dw + threadIdx.x, s);
    if (threadIdx.x == 16) atomicAdd(db, s);
  }
}

// multiplicative output: u[v] = g[v] exp(clamp(o[v], -4, 6)), g = channel 0
// of the input; gradient do[v] = du[v] u[v] inside the clamp, 0 outside
constexpr float kGateLo = -4.f, kGateHi = 6.f;
__global__ void k_gate(const uint16_t* __restrict__ feat, int ld, const float* __restrict__ o, float* __restrict__ u,
                       long nvox) {
  for (long v = (long)blockIdx.x * blockDim.x + threadIdx.x; v < nvox; v += (long)gridDim.x * blockDim.x)
    u[v] = bf2f(feat[v * ld]) * expf(fminf(fmaxf(o[v], kGateLo), kGateHi));
}
__global__ void k_gate_bwd(const float* __restrict__ o, const float* __restrict__ u, const float* __restrict__ du,
                           float* __restrict__ dout, long nvox) {
  for (long v = (long)blockIdx.x * blockDim.x + threadIdx.x; v < nvox; v += (long)gridDim.x * blockDim.x)
    dout[v] = (o[v] > kGateLo && o[v] < kGateHi) ? du[v] * u[v] : 0.f;
}

// masked mean-square loss: dout = 2 (out - target) m / sum(m); loss += sum m (out - target)^2 / sum(m)
__global__ void k_mse(const float* __restrict__ out, const float* __restrict__ target, const float* __restrict__ m,
                      long nvox, float inv_count, float* __restrict__ dout, double* __restrict__ loss) {
  __shared__ double red[256];
  double s = 0.0;
  for (long v = (long)blockIdx.x * blockDim.x + threadIdx.x; v < nvox; v += (long)gridDim.x * blockDim.x) {
    const float d = (out[v] - target[v]) * m[v];
    dout[v] = 2.f * d * inv_count;
    s += (double)d * d;
  }
  red[threadIdx.x] = s;
  __syncthreads();
  for (int k = 128; k > 0; k >>= 1) {
    if (threadIdx.x < k) red[threadIdx.x] += red[threadIdx.x + k];
    __syncthreads();
  }
  if (threadIdx.x == 0) atomicAdd(loss, red[0] * inv_count);
}

// Adam on FP32 master weights; the BF16 copy (when given) is refreshed
__global__ void k_adam(float* __restrict__ w, const float* __restrict__ g, float* __restrict__ m1, float* __restrict__ m2,
                       long n, float lr, float b1, float b2, float eps, float c1, float c2, uint16_t* __restrict__ w16) {
  for (long k = (long)blockIdx.x * blockDim.x + threadIdx.x; k < n; k += (long)gridDim.x * blockDim.x) {
    const float gk = g[k];
    const float a = b1 * m1[k] + (1.f - b1) * gk;
    const float b = b2 * m2[k] + (1.f - b2) * gk * gk;
    m1[k] = a;
    m2[k] = b;
    const float wk = w[k] - lr * (a * c1) / (sqrtf(b * c2) + eps);
    w[k] = wk;
    if (w16) w16[k] = f2bf(wk);
  }
}

__global__ void k_to_bf16(const float* __restrict__ x, uint16_t* __restrict__ y, long n) {
  for (long k = (long)blockIdx.x * blockDim.x + threadIdx.x; k < n; k += (long)gridDim.x * blockDim.x) y[k] = f2bf(x[k]);
}

// max |x| as float bits (non-negative floats order as unsigned integers)
__global__ void k_absmax(const float* __restrict__ x, long n, unsigned* __restrict__ out) {
  __shared__ float red[256];
  float s = 0.f;
  for (long k = (long)blockIdx.x * blockDim.x + threadIdx.x; k < n; k += (long)gridDim.x * blockDim.x) s = fmaxf(s, fabsf(x[k]));
  red[threadIdx.x] = s;
  __syncthreads();
  for (int k = 128; k > 0; k >>= 1) {
    if (threadIdx.x < k) red[threadIdx.x] = fmaxf(red[threadIdx.x], red[threadIdx.x + k]);
    __syncthreads();
  }
  if (threadIdx.x == 0) atomicMax(out, __float_as_uint(red[0]));
}

// Network input from a padded FWI grid (npml cells per face): channel 0 =
// g / gmax, channel 1 = log10(illum / imax + 1e-6) / 6 + 1, channels 2..15 zero.
__global__ void k_pack(const float* __restrict__ g, const float* __restrict__ illum, Vol P, int p, float inv_g,
                       float inv_i, uint16_t* __restrict__ feat) {
  const Vol V{P.nz - 2 * p, P.nx - 2 * p, P.ny - 2 * p};
  const long n = (long)V.nz * V.nx * V.ny;
  for (long v = (long)blockIdx.x * blockDim.x + threadIdx.x; v < n; v += (long)gridDim.x * blockDim.x) {
    const int y = (int)(v % V.ny);
    const long r = v / V.ny;
    const int x = (int)(r % V.nx), z = (int)(r / V.nx);
    const long q = P.vox(z + p, x + p, y + p);
    uint16_t* f = feat + v * 16;
    f[0] = f2bf(g[q] * inv_g);
    f[1] = f2bf(log10f(illum[q] * inv_i + 1e-6f) / 6.f + 1.f);
    for (int c = 2; c < 16; c++) f[c] = 0;
  }
}

__global__ void k_zero(unsigned* __restrict__ p, int n) {
  const int k = blockIdx.x * blockDim.x + threadIdx.x;
  if (k < n) p[k] = 0u;
}

inline unsigned grid_for(long n, int block, long cap = 16384) {
  return (unsigned)std::max(1L, std::min(cap, (n + block - 1) / block));
}

}  // namespace

extern "C" {

static int nn_check(hipStream_t s) {
  (void)s;
  return hipGetLastError() == hipSuccess ? 0 : EFWI_NN_ERR_HIP;
}

int efwi_nn_conv3(const uint16_t* src, int lds, int kc, const uint16_t* w, const float* bias, uint16_t* dst, int ldd,
                  int nc, const uint16_t* mask, int ldm, int nz, int nx, int ny, int mode, int relu, void* stream) {
  if (!src || !w || !dst || kc <= 0 || nc <= 0 || kc % 16 || nc % 16 || lds % 16 || ldd % 16 || lds < kc ||
      ldd < nc || (mask && (ldm % 16 || ldm < nc)) || nz <= 0 || nx <= 0 || ny <= 0 || (mode != 0 && mode != 1))
    return EFWI_NN_ERR_ARG;
  hipStream_t s = (hipStream_t)stream;
  const Vol V{nz, nx, ny};
  // row segments of 64 voxels (4 MFMA row blocks) where the rows are long enough
  const int MBv = ny >= 48 ? 4 : 1;
  const long tiles = (long)nz * nx * ((ny + 16 * MBv - 1) / (16 * MBv));
  const int nbt = nc / 16;
  const int NB = nbt >= 4 ? 4 : nbt;  // output blocks per wavefront
  const dim3 grid((unsigned)((tiles + 3) / 4), (unsigned)((nbt + NB - 1) / NB));
#define EFWI_NN_CONV_M(B, M)                                                                                       \
  if (mode == 0)                                                                                                 \
    hipLaunchKernelGGL((k_conv3<B, M, 0>), grid, dim3(256), 0, s, src, lds, kc, w, bias, dst, ldd, nc, mask, ldm, V, relu); \
  else                                                                                                           \
    hipLaunchKernelGGL((k_conv3<B, M, 1>), grid, dim3(256), 0, s, src, lds, kc, w, bias, dst, ldd, nc, mask, ldm, V, relu)
#define EFWI_NN_CONV(B)                \
  if (MBv == 4) { EFWI_NN_CONV_M(B, 4); } \
  else { EFWI_NN_CONV_M(B, 1); }
  if (NB == 1) { EFWI_NN_CONV(1); }
  else if (NB == 2) { EFWI_NN_CONV(2); }
  else if (NB == 3) { EFWI_NN_CONV(3); }
  else { EFWI_NN_CONV(4); }
#undef EFWI_NN_CONV
#undef EFWI_NN_CONV_M
  return nn_check(s);
}

long efwi_nn_wgrad_scratch(int mc, int kc, int nz, int nx, int ny, int per) {
  const long tiles = (long)nz * nx * ((ny + 15) / 16);
  const long chunks = (tiles + per - 1) / per;
  return chunks * 27L * mc * kc;
}

static bool wgrad_args_ok(const uint16_t* dpre, int ldp, int mc, const uint16_t* src, int lds, int kc, float* dw,
                          int nz, int nx, int ny, int per) {
  return dpre && src && dw && mc % 16 == 0 && kc % 16 == 0 && mc > 0 && kc > 0 && kc <= 96 && per > 0 &&
         ldp % 16 == 0 && lds % 16 == 0 && ldp >= mc && lds >= kc && nz > 0 && nx > 0 && ny > 0 &&
         ((uintptr_t)dpre | (uintptr_t)src) % 16 == 0;
}
static int wgrad_launch(const uint16_t* dpre, int ldp, int mc, const uint16_t* src, int lds, int kc, float* scratch,
                        float* dw, int nz, int nx, int ny, int per, hipStream_t s);

int efwi_nn_wgrad(const uint16_t* dpre, int ldp, int mc, const uint16_t* src, int lds, int kc, float* scratch,
                  long scratch_floats, float* dw, int nz, int nx, int ny, int per, void* stream) {
  if (!scratch || !wgrad_args_ok(dpre, ldp, mc, src, lds, kc, dw, nz, nx, ny, per)) return EFWI_NN_ERR_ARG;
  if (efwi_nn_wgrad_scratch(mc, kc, nz, nx, ny, per) > scratch_floats) return EFWI_NN_ERR_ARG;
  return wgrad_launch(dpre, ldp, mc, src, lds, kc, scratch, dw, nz, nx, ny, per, (hipStream_t)stream);
}

// The partial sums of efwi_nn_wgrad_noalloc: one array per GPU, for the
// process's lifetime. Host code uses the device address that
// hipGetSymbolAddress returns for the current device (kept per device), never
// the array's name.
static __device__ __align__(256) float scratch[EFWI_NN_WGRAD_STATIC_FLOATS];
static constexpr int kMaxDevices = 64;
static float* g_scratch_dev[kMaxDevices];
static std::mutex g_scratch_mu;

static float* static_scratch(int* err) {
  int dev = 0;
  if (hipGetDevice(&dev) != hipSuccess || dev < 0 || dev >= kMaxDevices) {
    *err = EFWI_NN_ERR_HIP;
    return nullptr;
  }
  std::lock_guard<std::mutex> lock(g_scratch_mu);
  if (!g_scratch_dev[dev]) {
    void* ptr = nullptr;
    if (hipGetSymbolAddress(&ptr, HIP_SYMBOL(scratch)) != hipSuccess || !ptr) {
      *err = EFWI_NN_ERR_HIP;
      return nullptr;
    }
    g_scratch_dev[dev] = static_cast<float*>(ptr);
  }
  return g_scratch_dev[dev];
}

int efwi_nn_wgrad_noalloc(const uint16_t* dpre, int ldp, int mc, const uint16_t* src, int lds, int kc, float* dw,
                          int nz, int nx, int ny, int per, void* stream) {
  if (!wgrad_args_ok(dpre, ldp, mc, src, lds, kc, dw, nz, nx, ny, per)) return EFWI_NN_ERR_ARG;
  if (efwi_nn_wgrad_scratch(mc, kc, nz, nx, ny, per) > EFWI_NN_WGRAD_STATIC_FLOATS) return EFWI_NN_ERR_SCRATCH;
  int err = 0;
  float* ptr = static_scratch(&err);
  if (!ptr) return err;
  return wgrad_launch(dpre, ldp, mc, src, lds, kc, ptr, dw, nz, nx, ny, per, (hipStream_t)stream);
}

static int wgrad_launch(const uint16_t* dpre, int ldp, int mc, const uint16_t* src, int lds, int kc, float* scratch,
                        float* dw, int nz, int nx, int ny, int per, hipStream_t s) {
  const Vol V{nz, nx, ny};
  const long tiles = (long)nz * nx * ((ny + 15) / 16);
  const int chunks = (int)((tiles + per - 1) / per);
  // one 16-channel n block per workgroup: two per workgroup (A tiles shared)
  // measured slower, at half the occupancy (185 registers against 112)
  hipLaunchKernelGGL((k_wgrad<1>), dim3((unsigned)chunks, (unsigned)(mc / 16), (unsigned)(kc / 16)), dim3(192), 0, s,
                     dpre, ldp, mc, src, lds, kc, scratch, V, per);
  const long width = 27L * mc * kc;
  hipLaunchKernelGGL(k_zero, dim3((unsigned)((width + 255) / 256)), dim3(256), 0, s, reinterpret_cast<unsigned*>(dw),
                     (int)width);
  hipLaunchKernelGGL(k_sum_chunks, dim3(grid_for(width, 256, 1L << 30), (chunks + kChunkGroup - 1) / kChunkGroup),
                     dim3(256), 0, s, scratch, width, chunks, dw);
  return nn_check(s);
}

int efwi_nn_chan_sum(const uint16_t* x, int ld, int c, long nvox, float* out, void* stream) {
  if (!x || !out || c <= 0 || c > 256 || 256 % c) return EFWI_NN_ERR_ARG;
  hipStream_t s = (hipStream_t)stream;
  hipLaunchKernelGGL(k_zero, dim3((c + 255) / 256), dim3(256), 0, s, reinterpret_cast<unsigned*>(out), c);
  hipLaunchKernelGGL(k_chan_sum, dim3(grid_for(nvox * c, 256, 256)), dim3(256), 0, s, x, ld, c, nvox, out);
  return nn_check(s);
}

int efwi_nn_pool2(const uint16_t* src, int lds, uint16_t* dst, int ldd, int c, int nz, int nx, int ny, void* stream) {
  if (!src || !dst || nz % 2 || nx % 2 || ny % 2 || c <= 0) return EFWI_NN_ERR_ARG;
  hipStream_t s = (hipStream_t)stream;
  hipLaunchKernelGGL(k_pool2, dim3(grid_for((long)nz * nx * ny / 8 * c, 256)), dim3(256), 0, s, src, lds, dst, ldd, c,
                     Vol{nz, nx, ny});
  return nn_check(s);
}

int efwi_nn_pool2_bwd(const uint16_t* dc, int ldc, uint16_t* df, int ldf, int c, int nz, int nx, int ny, int accumulate,
                      void* stream) {
  if (!dc || !df || nz % 2 || nx % 2 || ny % 2 || c <= 0) return EFWI_NN_ERR_ARG;
  hipStream_t s = (hipStream_t)stream;
  hipLaunchKernelGGL(k_pool2_bwd, dim3(grid_for((long)nz * nx * ny * c, 256)), dim3(256), 0, s, dc, ldc, df, ldf, c,
                     Vol{nz, nx, ny}, accumulate);
  return nn_check(s);
}

int efwi_nn_up2(const uint16_t* cs, int ldc, uint16_t* fd, int ldf, int c, int nz, int nx, int ny, void* stream) {
  if (!cs || !fd || nz % 2 || nx % 2 || ny % 2 || c <= 0) return EFWI_NN_ERR_ARG;
  hipStream_t s = (hipStream_t)stream;
  hipLaunchKernelGGL(k_up2, dim3(grid_for((long)nz * nx * ny * c, 256)), dim3(256), 0, s, cs, ldc, fd, ldf, c,
                     Vol{nz, nx, ny});
  return nn_check(s);
}

int efwi_nn_up2_bwd(const uint16_t* df, int ldf, uint16_t* dc, int ldc, int c, int nz, int nx, int ny,
                    const uint16_t* mask, int ldm, void* stream) {
  if (!df || !dc || nz % 2 || nx % 2 || ny % 2 || c <= 0) return EFWI_NN_ERR_ARG;
  hipStream_t s = (hipStream_t)stream;
  hipLaunchKernelGGL(k_up2_bwd, dim3(grid_for((long)nz * nx * ny / 8 * c, 256)), dim3(256), 0, s, df, ldf, dc, ldc, c,
                     Vol{nz, nx, ny}, mask, ldm);
  return nn_check(s);
}

int efwi_nn_add(uint16_t* a, int lda, const uint16_t* b, int ldb, int c, long nvox, const uint16_t* mask, int ldm,
                void* stream) {
  if (!a || c <= 0) return EFWI_NN_ERR_ARG;
  hipStream_t s = (hipStream_t)stream;
  hipLaunchKernelGGL(k_add_masked, dim3(grid_for(nvox * c, 256)), dim3(256), 0, s, a, lda, b, ldb, c, nvox, mask, ldm);
  return nn_check(s);
}

int efwi_nn_head(const uint16_t* act, int ld, int c, const float* w, const float* b, float* out, long nvox, void* stream) {
  if (!act || !w || !b || !out || c <= 0 || c > 16) return EFWI_NN_ERR_ARG;
  hipStream_t s = (hipStream_t)stream;
  hipLaunchKernelGGL(k_head, dim3(grid_for(nvox, 256)), dim3(256), 0, s, act, ld, c, w, b, out, nvox);
  return nn_check(s);
}

int efwi_nn_head_bwd(const uint16_t* act, int ld, int c, const float* w, const float* dout, uint16_t* dact, int ldd,
                     float* dw, float* db, long nvox, void* stream) {
  if (!act || !w || !dout || !dact || !dw || !db || c <= 0 || c > 16) return EFWI_NN_ERR_ARG;
  hipStream_t s = (hipStream_t)stream;
  hipLaunchKernelGGL(k_zero, dim3(1), dim3(64), 0, s, reinterpret_cast<unsigned*>(dw), c);
  hipLaunchKernelGGL(k_zero, dim3(1), dim3(64), 0, s, reinterpret_cast<unsigned*>(db), 1);
  hipLaunchKernelGGL(k_head_bwd, dim3(grid_for(nvox, 64, 2048)), dim3(64), 0, s, act, ld, c, w, dout, dact, ldd, dw, db, nvox);
  return nn_check(s);
}

int efwi_nn_gate(const uint16_t* feat, int ld, const float* o, float* u, long nvox, void* stream) {
  if (!feat || !o || !u || ld < 1) return EFWI_NN_ERR_ARG;
  hipStream_t s = (hipStream_t)stream;
  hipLaunchKernelGGL(k_gate, dim3(grid_for(nvox, 256)), dim3(256), 0, s, feat, ld, o, u, nvox);
  return nn_check(s);
}

int efwi_nn_gate_bwd(const float* o, const float* u, const float* du, float* dout, long nvox, void* stream) {
  if (!o || !u || !du || !dout) return EFWI_NN_ERR_ARG;
  hipStream_t s = (hipStream_t)stream;
  hipLaunchKernelGGL(k_gate_bwd, dim3(grid_for(nvox, 256)), dim3(256), 0, s, o, u, du, dout, nvox);
  return nn_check(s);
}

int efwi_nn_mse(const float* out, const float* target, const float* mask, long nvox, double count, float* dout,
                      double* loss, void* stream) {
  if (!out || !target || !mask || !dout || !loss || !(count > 0)) return EFWI_NN_ERR_ARG;
  hipStream_t s = (hipStream_t)stream;
  hipLaunchKernelGGL(k_zero, dim3(1), dim3(64), 0, s, reinterpret_cast<unsigned*>(loss), 2);
  hipLaunchKernelGGL(k_mse, dim3(grid_for(nvox, 256, 1024)), dim3(256), 0, s, out, target, mask, nvox,
                     (float)(1.0 / count), dout, loss);
  return nn_check(s);
}

int efwi_nn_adam(float* w, const float* g, float* m1, float* m2, long n, float lr, float b1, float b2, float eps, int step,
                 uint16_t* w16, void* stream) {
  if (!w || !g || !m1 || !m2 || n <= 0 || step < 1) return EFWI_NN_ERR_ARG;
  hipStream_t s = (hipStream_t)stream;
  const float c1 = 1.f / (1.f - powf(b1, (float)step)), c2 = 1.f / (1.f - powf(b2, (float)step));
  hipLaunchKernelGGL(k_adam, dim3(grid_for(n, 256)), dim3(256), 0, s, w, g, m1, m2, n, lr, b1, b2, eps, c1, c2, w16);
  return nn_check(s);
}

int efwi_nn_to_bf16(const float* x, uint16_t* y, long n, void* stream) {
  if (!x || !y || n < 0) return EFWI_NN_ERR_ARG;
  hipStream_t s = (hipStream_t)stream;
  hipLaunchKernelGGL(k_to_bf16, dim3(grid_for(n, 256)), dim3(256), 0, s, x, y, n);
  return nn_check(s);
}

int efwi_nn_absmax(const float* x, long n, float* out, void* stream) {
  if (!x || !out || n <= 0) return EFWI_NN_ERR_ARG;
  hipStream_t s = (hipStream_t)stream;
  hipLaunchKernelGGL(k_zero, dim3(1), dim3(64), 0, s, reinterpret_cast<unsigned*>(out), 1);
  hipLaunchKernelGGL(k_absmax, dim3(grid_for(n, 256, 1024)), dim3(256), 0, s, x, n, reinterpret_cast<unsigned*>(out));
  return nn_check(s);
}

int efwi_nn_pack(const float* g, const float* illum, int nz, int nx, int ny, int npml, float inv_g, float inv_i,
                 uint16_t* feat, void* stream) {
  if (!g || !illum || !feat || nz <= 2 * npml || nx <= 2 * npml || ny <= 2 * npml) return EFWI_NN_ERR_ARG;
  hipStream_t s = (hipStream_t)stream;
  const long n = (long)(nz - 2 * npml) * (nx - 2 * npml) * (ny - 2 * npml);
  hipLaunchKernelGGL(k_pack, dim3(grid_for(n, 256)), dim3(256), 0, s, g, illum, Vol{nz, nx, ny}, npml, inv_g, inv_i, feat);
  return nn_check(s);
}

}  // extern "C"
