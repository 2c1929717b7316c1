// Kernels of the learned preconditioner: a 3-D U-Net with 3x3x3 convolutions,
// BF16 weights and activations, FP32 accumulation, on gfx950 matrix cores.
//
// Tensors are channels-last volumes: element (voxel v, channel c) at
// p[v * ld + c], voxel v = (z * nx + x) * ny + y (y fastest, as the FWI grid),
// ld (the channel stride) a multiple of 16, so a tensor can be a channel slice
// of a wider one; a convolution can also take its input from two tensors and
// write its output to two (a skip concatenation kept as its two parts).
//
// Convolutions and weight gradients as implicit GEMMs on
// v_mfma_f32_16x16x32_bf16 (D = A B + C, 16 x 16 x 32, one wavefront). Lane
// l holds 8 K entries of row l % 16 of A and of column l % 16 of B, the K
// entries of lane group l / 16 being the same for A and B (that is all the
// contraction needs), and D[4 (l / 16) + j][l % 16], j = 0..3. In the
// convolution the A rows are 16 consecutive voxels along y, the K dimension
// runs over (tap, input channel) and the B columns are output channels; in
// the weight gradient K runs over voxels.
//   forward:  out[v][n] = sum_t sum_k in[v + d_t][k] w[t][n][k]      (+ bias, ReLU)
//   dgrad:    din[v][n] = sum_t sum_k dout[v - d_t][k] w[t][k][n]    (x ReLU mask)
//   wgrad:    dw[t][m][n] = sum_v dpre[v][m] in[v + d_t][n]          (K = voxels)
// with zero padding at the volume faces. No reference counterpart: the
// reference has no learned part.
#include <hip/hip_runtime.h>

#include <algorithm>
#include <climits>
#include <cstdint>
#include <mutex>

#include "../../include/efwi_nn.h"

namespace {

typedef short s4 __attribute__((ext_vector_type(4)));
typedef float f4 __attribute__((ext_vector_type(4)));
typedef __bf16 bf8 __attribute__((ext_vector_type(8)));

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

// BF16 granules (8 channels, 16 bytes) to and from FP32
__device__ __forceinline__ void unpack8(const uint4& q, float (&f)[8]) {
  const unsigned u[4] = {q.x, q.y, q.z, q.w};
#pragma unroll
  for (int j = 0; j < 4; j++) {
    f[2 * j] = __uint_as_float(u[j] << 16);
    f[2 * j + 1] = __uint_as_float(u[j] & 0xFFFF0000u);
  }
}
__device__ __forceinline__ uint4 pack8(const float (&f)[8]) {
  unsigned u[4];
#pragma unroll
  for (int j = 0; j < 4; j++) u[j] = (unsigned)f2bf(f[2 * j]) | ((unsigned)f2bf(f[2 * j + 1]) << 16);
  return uint4{u[0], u[1], u[2], u[3]};
}
template <typename T>
__device__ __forceinline__ uint4 ld16(const T* p) { return *reinterpret_cast<const uint4*>(p); }
template <typename T>
__device__ __forceinline__ void st16(T* p, const uint4& q) { *reinterpret_cast<uint4*>(p) = q; }

__device__ __forceinline__ void tap(int t, int& dz, int& dx, int& dy) {
  dz = t / 9 - 1;
  dx = (t / 3) % 3 - 1;
  dy = t % 3 - 1;
}

// 3x3x3 convolution as implicit GEMM on v_mfma_f32_16x16x32_bf16, the input
// staged through LDS. A workgroup of 2 TZ wavefronts owns a brick of TZ (z) x
// 2 (x) rows of TY = 16 MB voxels along y and NB blocks of 16 output
// channels; wavefront w takes row (z0 + w / 2, x0 + w % 2), MB row blocks of
// the A operand (16 voxels each) against NB column blocks. Per chunk of CK
// input channels the block copies the brick's halo ((TZ + 2) x 4 rows of TY
// + 2 voxels, zero outside the volume) into LDS with 16-byte loads: an item
// is one (voxel, granule) pair over the 8 rows of two z planes, whose offsets
// are uniform. Every tap then reads its A fragments from there with one 16-byte
// LDS read per lane (the 8-channel granules of a voxel are XOR-swizzled with
// its position so that each 8-lane phase of a read hits distinct banks). The
// K dimension runs over (tap, channel) of the chunk in steps of 32: lane group
// g = l / 16 holds K entries 8 g .. 8 g + 7, i.e. taps t = (32 q + 8 g) / CK,
// so with CK = 16 one MFMA covers two taps (the lower and upper lane halves).
// A lane's LDS offsets are a per-lane base (one per dy) plus a constant per
// tap. B fragments (weights): w[t][n][k] in mode 0, from LDS where the chunk's
// weights are small (CK = 16, NB = 1; the K loop is then unrolled and its
// offsets fold to constants), else one 16-byte global load per lane; mode 1
// (input gradient with the forward layer's weights, taps flipped, B[k][n] =
// w[26 - t][k][n]) gathers them with 2-byte loads, which a caller avoids by
// passing the flipped copy of efwi_nn_flip_weights to mode 0.
template <int CK>
__device__ __forceinline__ int granule_swizzle(int hv) {
  // G = CK / 8 granules of 16 bytes per voxel. A 16-byte LDS read is served
  // 8 lanes at a time across 32 banks: the granules of 8 consecutive voxels
  // must cover distinct 16-byte bank groups (measured on MI355X: no bank
  // conflicts with this swizzle, against 0.66 M conflict cycles without one
  // and 7.5 M with a swizzle for 16-lane, 64-bank service)
  return CK == 32 ? ((hv >> 1) & 3) : ((hv >> 2) & 1);
}

__device__ __forceinline__ uint4 lds_read16(const unsigned char* base, int byte_off) {
  return *reinterpret_cast<const uint4*>(base + byte_off);
}

// rows of a convolution brick along z (x: always 2), two wavefronts per z
// row: 4 for rows of 64 voxels (halo rows per output row 3 instead of 4),
// 2 for the shorter rows of the small volumes (more bricks to fill the GPU)
constexpr int conv_tz(int mb) { return mb == 4 ? 4 : 2; }

template <int CK, int MB, int NB, int MODE>
__global__ __launch_bounds__(512) void k_conv3(const uint16_t* __restrict__ src, int lds, int kc,
                                               const uint16_t* __restrict__ w, const float* __restrict__ bias,
                                               uint16_t* __restrict__ dst, int ldd, int nc,
                                               const uint16_t* __restrict__ mask, int ldm, Vol V, int relu,
                                               int ntz, int ntx, int nty, int per_xcd, int vec_out,
                                               const uint16_t* __restrict__ src2, int lds2, int kc1,
                                               uint16_t* __restrict__ dst2, int ldd2, int nc1, int up1) {
  constexpr int TY = 16 * MB, HW = TY + 2, G = CK / 8;
  constexpr int TZ = conv_tz(MB), NT = 128 * TZ, HR = (TZ + 2) * 4;  // halo rows
  constexpr int NQ = (27 * CK + 31) / 32;
  constexpr int H_PAIRS = HW * G, NRG = HR / 8, H_NPASS = (H_PAIRS * NRG + NT - 1) / NT;
  constexpr bool WL = CK == 16 && NB == 1;
  constexpr int NBC = NB * 16;
  constexpr int VB = CK * 2;  // bytes per voxel in the halo
  constexpr bool TILE_OK = 2 * TZ * TY * NBC * 2 <= HR * HW * VB;  // output tile fits the halo space
  __shared__ __align__(16) unsigned char halo[HR * HW * VB];
  __shared__ __align__(16) unsigned char wl[WL ? 27 * NBC * VB : 16];
  const int tid = threadIdx.x, lane = tid & 63, wv = tid >> 6;
  // blocks b, b + 8, b + 16, ... run on one XCD: give each XCD a contiguous
  // run of bricks (y fastest), whose halos overlap in its L2
  const int b = blockIdx.x;
  const long brick = (long)(b & 7) * per_xcd + (b >> 3);
  if (brick >= (long)ntz * ntx * nty) return;  // whole block
  const int yt = (int)(brick % nty);
  const long r2 = brick / nty;
  const int xt = (int)(r2 % ntx), zt = (int)(r2 / ntx);
  const int z0 = TZ * zt, x0 = 2 * xt, y0 = yt * TY;
  const bool interior = z0 >= 1 && z0 + TZ < V.nz && x0 >= 1 && x0 + 2 < V.nx && y0 >= 1 && y0 + TY < V.ny;
  const int wz = wv >> 1, wx = wv & 1;
  const int i = lane & 15, g = lane >> 4;
  const int hi = CK == 16 ? g >> 1 : 0;    // CK = 16: upper lane half takes the odd tap
  const int grl = CK == 16 ? g & 1 : g;    // granule of the lane's 8 K entries
  const int nb0 = blockIdx.y * NB;
  // LDS byte offsets of the lane's A fragment for tap (dz, dx, dy) = (-1, -1, dy)
  // and row block 0: abase[dy + 1]; other taps and blocks add constants
  int abase[3];
#pragma unroll
  for (int d = 0; d < 3; d++) abase[d] = ((wz * 4 + wx) * HW + i + d) * VB + 16 * (grl ^ granule_swizzle<CK>(i + d));
  const int wbase = (i * CK + 8 * (grl ^ granule_swizzle<CK>(i))) * 2 + hi * NBC * VB;
  // halo corner voxel (z0 - 1, x0 - 1, y0 - 1)
  const long corner = ((long)(z0 - 1) * V.nx + (x0 - 1)) * V.ny + (y0 - 1);
  f4 acc[MB][NB];
#pragma unroll
  for (int m = 0; m < MB; m++)
#pragma unroll
    for (int c = 0; c < NB; c++) acc[m][c] = f4{0.f, 0.f, 0.f, 0.f};
  for (int c0 = 0; c0 < kc; c0 += CK) {
    // input channels [0, kc1) come from src, [kc1, kc) from src2 (a chunk
    // never straddles the two)
    const bool second = c0 >= kc1;
    const uint16_t* const sb = second ? src2 : src;
    const long sl = second ? lds2 : lds;
    const int cs = second ? c0 - kc1 : c0;
    // up1: src is the coarse (nz/2, nx/2, ny/2) tensor whose nearest
    // upsampling is the input; a voxel reads its parent
    const bool up = up1 && !second;
    // halo rows r = 4 hz + hx at (z0 - 1 + hz, x0 - 1 + hx), voxels hv at
    // y0 - 1 + hv. An item is one (hv, granule) pair over 8 rows (two z
    // planes; row offsets uniform), its eight loads in flight before the
    // stores.
#pragma unroll
    for (int v = 0; v < H_NPASS; v++) {
      const int item = tid + NT * v;
      if (item < H_PAIRS * NRG) {
        const int pidx = item % H_PAIRS, rg = item / H_PAIRS;
        const int gr = pidx % G, hv = pidx / G;
        const int ys = y0 - 1 + hv;
        const bool y_ok = interior || (ys >= 0 && ys < V.ny);
        const uint16_t* tp = sb + (corner + (long)(2 * rg) * V.nx * V.ny + hv) * sl + cs + 8 * gr;
        unsigned char* lp = halo + (8 * rg * HW + hv) * VB + 16 * (gr ^ granule_swizzle<CK>(hv));
        uint4 hq[8];
#pragma unroll
        for (int rr = 0; rr < 8; rr++) {
          const int hz = 2 * rg + (rr >> 2), hx = rr & 3;
          const bool ok = interior || (y_ok && (unsigned)(z0 - 1 + hz) < (unsigned)V.nz &&
                                       (unsigned)(x0 - 1 + hx) < (unsigned)V.nx);
          hq[rr] = uint4{0u, 0u, 0u, 0u};
          if (up) {
            if (ok)
              hq[rr] = *reinterpret_cast<const uint4*>(
                  sb + ((((long)(z0 - 1 + hz) >> 1) * (V.nx >> 1) + ((x0 - 1 + hx) >> 1)) * (V.ny >> 1) + (ys >> 1)) * sl +
                  cs + 8 * gr);
          } else if (ok) {
            hq[rr] = *reinterpret_cast<const uint4*>(tp + ((long)(rr >> 2) * V.nx + hx) * V.ny * sl);
          }
        }
#pragma unroll
        for (int rr = 0; rr < 8; rr++) *reinterpret_cast<uint4*>(lp + rr * HW * VB) = hq[rr];
      }
    }
    if (WL) {
      // thread: (n, granule) = (tid % 16, tid / 16 % 2), taps tid / 32 + TS u
      constexpr int TS = NT / 32, WU = (27 + TS - 1) / TS;
      const int nl = tid & 15, gr = (tid >> 4) & 1, t0 = tid >> 5;
      const int n = nb0 * 16 + nl;
      uint4 wq[WU];
#pragma unroll
      for (int u = 0; u < WU; u++) {
        const int t = t0 + TS * u;
        wq[u] = uint4{0u, 0u, 0u, 0u};
        if (t < 27) {
          if (MODE == 0) {
            wq[u] = *reinterpret_cast<const uint4*>(w + ((long)t * nc + n) * kc + c0 + 8 * gr);
          } else {
            const uint16_t* p = w + ((long)(26 - t) * kc + c0 + 8 * gr) * nc + n;
            unsigned e[8];
#pragma unroll
            for (int jj = 0; jj < 8; jj++) e[jj] = p[(long)jj * nc];
            wq[u] = uint4{e[0] | e[1] << 16, e[2] | e[3] << 16, e[4] | e[5] << 16, e[6] | e[7] << 16};
          }
        }
      }
      unsigned char* lp = wl + (t0 * NBC + nl) * VB + 16 * (gr ^ granule_swizzle<CK>(nl));
#pragma unroll
      for (int u = 0; u < WU; u++)
        if (t0 + TS * u < 27) *reinterpret_cast<uint4*>(lp + TS * u * NBC * VB) = wq[u];
    }
    __syncthreads();
    // weights in LDS: unrolled (the lane halves' taps differ; their offsets
    // fold to constants); weights in global memory: not unrolled (the
    // compiler would hoist every step's weight loads)
#pragma unroll(WL ? NQ : 1)
    for (int q = 0; q < NQ; q++) {
      // taps of the lower (tl) and, with CK = 16, upper (th) lane halves
      const int tl = CK == 16 ? 2 * q : q, th = CK == 16 ? 2 * q + 1 : q;
      const int thc = th < 27 ? th : 26;
      auto aoff = [&](int t) {
        const int dz = t / 9, dx = (t / 3) % 3, dy = t % 3;
        return (dy == 0 ? abase[0] : dy == 1 ? abase[1] : abase[2]) + (dz * 4 + dx) * HW * VB;
      };
      const int ao = (CK == 16 && hi) ? aoff(thc) : aoff(tl);
      const bool zero = CK == 16 && hi && th >= 27;
      uint4 a[MB], bw[NB];
#pragma unroll
      for (int m = 0; m < MB; m++) {
        a[m] = lds_read16(halo, ao + 16 * m * VB);
        if (zero) a[m] = uint4{0u, 0u, 0u, 0u};
      }
#pragma unroll
      for (int c = 0; c < NB; c++) {
        const int n = (nb0 + c) * 16 + i;
        if (WL) {
          bw[c] = lds_read16(wl, wbase + tl * NBC * VB);
        } else {
          const int t = (CK == 16 && hi) ? thc : tl;
          const int gr = grl;
          if (MODE == 0) {
            bw[c] = *reinterpret_cast<const uint4*>(w + ((long)t * nc + n) * kc + c0 + 8 * gr);
          } else {
            const uint16_t* p = w + ((long)(26 - t) * kc + c0 + 8 * gr) * nc + n;
            unsigned e[8];
#pragma unroll
            for (int jj = 0; jj < 8; jj++) e[jj] = p[(long)jj * nc];
            bw[c] = uint4{e[0] | e[1] << 16, e[2] | e[3] << 16, e[4] | e[5] << 16, e[6] | e[7] << 16};
          }
        }
        if (zero) bw[c] = uint4{0u, 0u, 0u, 0u};
      }
#pragma unroll
      for (int c = 0; c < NB; c++)
#pragma unroll
        for (int m = 0; m < MB; m++)
          acc[m][c] = __builtin_amdgcn_mfma_f32_16x16x32_bf16(__builtin_bit_cast(bf8, a[m]), __builtin_bit_cast(bf8, bw[c]),
                                                              acc[m][c], 0, 0, 0);
      // keep each step's loads next to its MFMAs (bounded registers)
      __builtin_amdgcn_sched_barrier(0);
    }
    __syncthreads();
  }
  // D[voxel 4 g + j][channel i] of each (row block, column block)
  const int z = z0 + wz, x = x0 + wx;
  const bool row_ok = z < V.nz && x < V.nx;
  const long vrow = V.vox(row_ok ? z : 0, row_ok ? x : 0, y0);
  if (TILE_OK && vec_out) {
    // through LDS (the halo's space): the wavefront's TY x NBC tile as BF16,
    // then 16-byte granules (8 channels of one voxel) to dst, masked
    uint16_t* tile = reinterpret_cast<uint16_t*>(halo) + wv * TY * NBC;
#pragma unroll
    for (int c = 0; c < NB; c++) {
      const float bn = bias ? bias[(nb0 + c) * 16 + i] : 0.f;
#pragma unroll
      for (int m = 0; m < MB; m++)
#pragma unroll
        for (int jj = 0; jj < 4; jj++) {
          float r = acc[m][c][jj] + bn;
          if (relu) r = fmaxf(r, 0.f);
          tile[(16 * m + 4 * g + jj) * NBC + 16 * c + i] = f2bf(r);
        }
    }
    __syncthreads();
    if (!row_ok) return;
    constexpr int GR = NBC / 8;
#pragma unroll
    for (int k0 = 0; k0 < TY * GR; k0 += 64) {
      const int k = k0 + lane;
      if (TY * GR % 64 != 0 && k >= TY * GR) break;
      const int vv = k / GR, gg = k % GR;
      if (!interior && y0 + vv >= V.ny) continue;
      uint4 q = *reinterpret_cast<const uint4*>(tile + vv * NBC + 8 * gg);
      if (mask) {
        float f[8], mk[8];
        unpack8(q, f);
        unpack8(ld16(mask + (vrow + vv) * ldm + nb0 * 16 + 8 * gg), mk);
#pragma unroll
        for (int j = 0; j < 8; j++)
          if (!(mk[j] > 0.f)) f[j] = 0.f;
        q = pack8(f);
      }
      // output channels [0, nc1) go to dst, [nc1, nc) to dst2
      const int ch = nb0 * 16 + 8 * gg;
      st16(ch < nc1 ? dst + (vrow + vv) * ldd + ch : dst2 + (vrow + vv) * ldd2 + (ch - nc1), q);
    }
    return;
  }
  if (!row_ok) return;
#pragma unroll
  for (int c = 0; c < NB; c++) {
    const int n = (nb0 + c) * 16 + i;
    const float bn = bias ? bias[n] : 0.f;
    const long od = n < nc1 ? ldd : ldd2;
    uint16_t* out = n < nc1 ? dst + (vrow + 4 * g) * ldd + n : dst2 + (vrow + 4 * g) * ldd2 + (n - nc1);
    const uint16_t* mk = mask ? mask + (vrow + 4 * g) * ldm + n : nullptr;
#pragma unroll
    for (int m = 0; m < MB; m++)
#pragma unroll
      for (int jj = 0; jj < 4; jj++) {
        const int dyv = 16 * m + jj;
        if (!interior && y0 + 4 * g + dyv >= V.ny) continue;
        float r = acc[m][c][jj] + bn;
        if (relu) r = fmaxf(r, 0.f);
        if (mk && !(bf2f(mk[(long)dyv * ldm]) > 0.f)) r = 0.f;
        out[(long)dyv * od] = f2bf(r);
      }
  }
}

// wt[t][k][n] = w[26 - t][n][k]: the weights of a layer's input gradient as a
// forward convolution (taps mirrored, channels transposed)
__global__ void k_flip_weights(const uint16_t* __restrict__ w, int kc, int nc, uint16_t* __restrict__ wt) {
  const long n_all = 27L * kc * nc;
  for (long e = (long)blockIdx.x * blockDim.x + threadIdx.x; e < n_all; e += (long)gridDim.x * blockDim.x) {
    const int k = (int)(e % kc);
    const long r = e / kc;
    const int n = (int)(r % nc), t = (int)(r / nc);
    wt[((long)(26 - t) * kc + k) * nc + n] = w[e];
  }
}

// The same for up to kFlipMax layers in one launch (blockIdx.y = layer)
constexpr int kFlipMax = 16;
struct FlipSet {
  const uint16_t* w[kFlipMax];
  uint16_t* wt[kFlipMax];
  int kc[kFlipMax], nc[kFlipMax];
};
__global__ void k_flip_many(FlipSet S) {
  const int l = blockIdx.y, kc = S.kc[l], nc = S.nc[l];
  const uint16_t* __restrict__ w = S.w[l];
  uint16_t* __restrict__ wt = S.wt[l];
  const long n_all = 27L * kc * nc;
  for (long e = (long)blockIdx.x * blockDim.x + threadIdx.x; e < n_all; e += (long)gridDim.x * blockDim.x) {
    const int k = (int)(e % kc);
    const long r = e / kc;
    const int n = (int)(r % nc), t = (int)(r / nc);
    wt[((long)(26 - t) * kc + k) * nc + n] = w[e];
  }
}

// Weight gradient partials on v_mfma_f32_16x16x32_bf16: K = voxels. A block
// of four wavefronts takes `per` bricks of 2 (z) x 2 (x) rows of TY voxels
// along y, consecutive along z (bricks are numbered z fastest), one 16-channel
// block of m and one of n. Per brick it copies into LDS, channel-major (y
// contiguous, rows padded to TY + 8 so that each 8-lane phase of a fragment
// read hits distinct banks; 8-byte stores of 4 voxels per channel): dpre of the
// brick (zero outside the volume) and the brick's src halo (4 planes x 4 rows
// of TY + 2 voxels) in a ring of four plane slots (plane z in slot z mod 4):
// a brick that follows its z neighbour in the same block keeps that
// neighbour's last two halo planes and copies only two new ones. Wavefront w takes output row w and every tap: per 32 voxels it
// reads one A fragment (A[m][k] = dpre[voxel k][m]) and, per (dz, dx), two
// 16-byte granules of the halo row from which the three dy shifts of the B
// fragment (B[k][n] = src[voxel k + d_t][n]) follow by dword selection and
// byte alignment. The four wavefronts' sums meet in LDS, and the block writes
// part[chunk][t][m][n] for its m and n blocks. With db, the blocks of the first
// n block also sum the A fragments they read (dpre over their voxels: the bias
// gradient of their m block) into part[chunk][27 mc kc + m]; a chunk's
// partials are 27 mc kc + mc floats.
// Transposed store of 4 voxels x 8 channels (q[k] = channels of voxel k):
// channel j's four values at dst[j * pitch .. + 3], one 8-byte store each.
__device__ __forceinline__ void store_quads(uint16_t* dst, int pitch, const uint4 (&q)[4]) {
  const unsigned* d[4] = {&q[0].x, &q[1].x, &q[2].x, &q[3].x};
#pragma unroll
  for (int j = 0; j < 8; j++) {
    const int h = j >> 1, sh = (j & 1) * 16;
    const unsigned lo = ((d[0][h] >> sh) & 0xFFFFu) | (((d[1][h] >> sh) & 0xFFFFu) << 16);
    const unsigned hi = ((d[2][h] >> sh) & 0xFFFFu) | (((d[3][h] >> sh) & 0xFFFFu) << 16);
    *reinterpret_cast<uint2*>(dst + j * pitch) = uint2{lo, hi};
  }
}

template <int TY>
__global__ __launch_bounds__(256) void k_wgrad(const uint16_t* __restrict__ dpre, int ldp, int mc,
                                               const uint16_t* __restrict__ src, int lds, int kc,
                                               float* __restrict__ part, float* __restrict__ dw,
                                               float* __restrict__ db, Vol V, int ntz, int ntx, int nty, int per,
                                               const uint16_t* __restrict__ src2, int lds2, int kc1, int up1) {
  constexpr int HW = TY + 2, P = TY + 8, NK = TY / 32;
  constexpr int A_BYTES = 4 * 16 * P * 2, B_BYTES = 16 * 16 * P * 2, R_BYTES = 9 * 4 * 256 * 4;
  constexpr int S_BYTES = A_BYTES + (B_BYTES > R_BYTES ? B_BYTES : R_BYTES);
  __shared__ __align__(16) unsigned char smem[S_BYTES];
  uint16_t* Ah = reinterpret_cast<uint16_t*>(smem);            // [row][m][y]
  uint16_t* Bh = reinterpret_cast<uint16_t*>(smem + A_BYTES);  // [halo row][n][hv]
  float* red = reinterpret_cast<float*>(smem + A_BYTES);       // after the last brick
  const int tid = threadIdx.x, wv = tid >> 6, lane = tid & 63;
  const int i = lane & 15, g = lane >> 4;
  const int wz = wv >> 1, wx = wv & 1;
  const int mb = blockIdx.y, nb = blockIdx.z;
  const long nbricks = (long)ntz * ntx * nty;
  const long b0 = (long)blockIdx.x * per, b1 = std::min(nbricks, b0 + per);
  const bool bias = db != nullptr && nb == 0;
  // this block's 16 src channels: from src below kc1, else from src2
  const bool second = nb * 16 >= kc1;
  const uint16_t* const sb = second ? src2 : src;
  const long sl = second ? lds2 : lds;
  const int scol = second ? nb * 16 - kc1 : nb * 16;
  // up1: src is the coarse tensor whose nearest upsampling is the input
  const bool up = up1 && !second;
  float bsum = 0.f;  // lane (i, g): dpre channel mb * 16 + i over its voxels
  f4 acc[27];
#pragma unroll
  for (int t = 0; t < 27; t++) acc[t] = f4{0.f, 0.f, 0.f, 0.f};
  for (long brick = b0; brick < b1; brick++) {
    const int zt = (int)(brick % ntz);
    const long r2 = brick / ntz;
    const int xt = (int)(r2 % ntx), yt = (int)(r2 / ntx);
    const int z0 = 2 * zt, x0 = 2 * xt, y0 = yt * TY;
    // planes z0 - 1 and z0 are in the ring when this block's previous brick
    // was the one below in the same column
    const int hz0 = brick > b0 && zt > 0 ? 2 : 0;
    // channel-major copies, 4 voxels along y per item: each channel's 4 values
    // go out as one 8-byte LDS store (consecutive lanes take consecutive
    // voxel quads of one channel row). Threads keep their (granule, quad,
    // row) for the brick; in a brick clear of the volume faces no voxel needs
    // a bounds check.
    const bool interior = z0 >= 1 && z0 + 2 < V.nz && x0 >= 1 && x0 + 2 < V.nx && y0 >= 1 && y0 + TY + 2 < V.ny;
    if (tid < 4 * (TY / 4) * 2) {
      const int y4 = tid % (TY / 4), rem = tid / (TY / 4);
      const int gr = rem & 1, row = rem >> 1;
      const int z = z0 + (row >> 1), x = x0 + (row & 1), y = y0 + 4 * y4;
      const bool zx_ok = z < V.nz && x < V.nx;
      const uint16_t* p = dpre + V.vox(zx_ok ? z : 0, zx_ok ? x : 0, y) * ldp + mb * 16 + 8 * gr;
      uint4 q[4];
#pragma unroll
      for (int k = 0; k < 4; k++) {
        q[k] = uint4{0u, 0u, 0u, 0u};
        if (interior || (zx_ok && y + k < V.ny)) q[k] = ld16(p + (long)k * ldp);
      }
      store_quads(Ah + (row * 16 + 8 * gr) * P + 4 * y4, P, q);
    }
    constexpr int HW4 = (HW + 3) / 4;
    // halo: thread = (granule, quad, hx) walks hz = 0..3 (uniform plane offsets)
    if (tid < 2 * HW4 * 4) {
      const int y4 = tid % HW4, rem = tid / HW4;
      const int gr = rem & 1, hx = rem >> 1;
      const int xs = x0 - 1 + hx, ys = y0 - 1 + 4 * y4;
      const bool x_ok = xs >= 0 && xs < V.nx;
      const long v0 = V.vox(z0 - 1, x_ok ? xs : 0, ys);
      const long plane = (long)V.nx * V.ny;
#pragma unroll
      for (int hz = 0; hz < 4; hz++) {
        if (hz < hz0) continue;
        const int zs = z0 - 1 + hz;
        const bool zx_ok = x_ok && zs >= 0 && zs < V.nz;
        const uint16_t* p = sb + (v0 + hz * plane) * sl + scol + 8 * gr;
        // coarse row of (zs, xs) when up (read only where zx_ok)
        const uint16_t* pc =
            sb + (((long)(zs >> 1) * (V.nx >> 1) + (xs >> 1)) * (V.ny >> 1)) * sl + scol + 8 * gr;
        uint4 q[4];
#pragma unroll
        for (int k = 0; k < 4; k++) {
          q[k] = uint4{0u, 0u, 0u, 0u};
          if (interior || (zx_ok && ys + k >= 0 && ys + k < V.ny))
            q[k] = up ? ld16(pc + (long)((ys + k) >> 1) * sl) : ld16(p + (long)k * sl);
        }
        store_quads(Bh + (((zs & 3) * 4 + hx) * 16 + 8 * gr) * P + 4 * y4, P, q);
      }
    }
    __syncthreads();
#pragma unroll 1
    for (int c = 0; c < NK; c++) {
      const uint4 a = *reinterpret_cast<const uint4*>(&Ah[(wv * 16 + i) * P + 32 * c + 8 * g]);
      if (bias) {
        const unsigned u[4] = {a.x, a.y, a.z, a.w};
#pragma unroll
        for (int j = 0; j < 4; j++) bsum += __uint_as_float(u[j] << 16) + __uint_as_float(u[j] & 0xFFFF0000u);
      }
#pragma unroll
      for (int dzdx = 0; dzdx < 9; dzdx++) {
        const int hr = ((z0 - 1 + wz + dzdx / 3) & 3) * 4 + (wx + dzdx % 3);
        const uint16_t* bp = &Bh[(hr * 16 + i) * P + 32 * c + 8 * g];
        const uint4 g0 = *reinterpret_cast<const uint4*>(bp);
        const uint4 g1 = *reinterpret_cast<const uint4*>(bp + 8);
        // halo voxels 32c + 8g + (dy + 1) + j, j = 0..7
        const uint4 s1 = {__builtin_amdgcn_alignbyte(g0.y, g0.x, 2), __builtin_amdgcn_alignbyte(g0.z, g0.y, 2),
                          __builtin_amdgcn_alignbyte(g0.w, g0.z, 2), __builtin_amdgcn_alignbyte(g1.x, g0.w, 2)};
        const uint4 s2 = {g0.y, g0.z, g0.w, g1.x};
        const bf8 av = __builtin_bit_cast(bf8, a);
        acc[3 * dzdx] = __builtin_amdgcn_mfma_f32_16x16x32_bf16(av, __builtin_bit_cast(bf8, g0), acc[3 * dzdx], 0, 0, 0);
        acc[3 * dzdx + 1] =
            __builtin_amdgcn_mfma_f32_16x16x32_bf16(av, __builtin_bit_cast(bf8, s1), acc[3 * dzdx + 1], 0, 0, 0);
        acc[3 * dzdx + 2] =
            __builtin_amdgcn_mfma_f32_16x16x32_bf16(av, __builtin_bit_cast(bf8, s2), acc[3 * dzdx + 2], 0, 0, 0);
      }
    }
    __syncthreads();
  }
  // sum of the four wavefronts, 9 taps at a time; lane (i, g) holds
  // D[m = 4 g + j][n = i] of each tap
  const long W = 27L * mc * kc, stride = W + mc;
#pragma unroll
  for (int p = 0; p < 3; p++) {
#pragma unroll
    for (int tt = 0; tt < 9; tt++)
#pragma unroll
      for (int j = 0; j < 4; j++) red[(tt * 4 + wv) * 256 + lane * 4 + j] = acc[9 * p + tt][j];
    __syncthreads();
    const int l = tid >> 2, j = tid & 3;
    const int m = mb * 16 + 4 * (l >> 4) + j, n = nb * 16 + (l & 15);
#pragma unroll
    for (int tt = 0; tt < 9; tt++) {
      const float s = red[(tt * 4 + 0) * 256 + tid] + red[(tt * 4 + 1) * 256 + tid] + red[(tt * 4 + 2) * 256 + tid] +
                      red[(tt * 4 + 3) * 256 + tid];
      part[(long)blockIdx.x * stride + ((long)(9 * p + tt) * mc + m) * kc + n] = s;
      // dw is summed from the partials by atomics after this kernel: the
      // first chunk's blocks clear it
      if (blockIdx.x == 0) dw[((long)(9 * p + tt) * mc + m) * kc + n] = 0.f;
    }
    __syncthreads();
  }
  if (bias) {
    red[tid] = bsum;  // tid = 64 wv + 16 g + i
    __syncthreads();
    if (tid < 16) {
      float s = 0.f;
#pragma unroll
      for (int q = 0; q < 16; q++) s += red[16 * q + tid];
      part[(long)blockIdx.x * stride + W + mb * 16 + tid] = s;
      if (blockIdx.x == 0) db[mb * 16 + tid] = 0.f;
    }
  }
}

// out[k] (k < width) or out2[k - width] (width <= k < total) += sum over the
// chunks [blockIdx.y * kChunkGroup, +kChunkGroup) of part[chunk * stride + k]
// (outputs cleared beforehand; one atomic per element per chunk group)
constexpr int kChunkGroup = 32;
__global__ void k_sum_chunks(const float* __restrict__ part, long stride, long width, long total, int nchunks,
                             float* __restrict__ out, float* __restrict__ out2) {
  const long k = (long)blockIdx.x * blockDim.x + threadIdx.x;
  if (k >= total) return;
  const int c0 = blockIdx.y * kChunkGroup, c1 = min(nchunks, c0 + kChunkGroup);
  float s = 0.f;
  // the loads of 8 chunks in flight at a time; the sum keeps chunk order
#pragma unroll 8
  for (int c = c0; c < c1; c++) s += part[(long)c * stride + k];
  atomicAdd(k < width ? out + k : out2 + (k - width), s);
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
// The same with one 16-byte load (8 channels) per thread and voxel: blockDim
// 1024 = (1024 / (c / 8)) voxel lanes x (c / 8) granules (c and ld multiples
// of 8, x 16-byte aligned). Few large blocks: every block adds its c sums
// with atomics to the same c addresses, which serialise.
__global__ __launch_bounds__(1024) void k_chan_sum8(const uint16_t* __restrict__ x, int ld, int c, long nvox,
                                                    float* __restrict__ out) {
  __shared__ float red[1024 * 8];
  const int G = c / 8, lanes = 1024 / G;
  const int gr = threadIdx.x % G, vl = threadIdx.x / G;
  float s[8] = {0.f, 0.f, 0.f, 0.f, 0.f, 0.f, 0.f, 0.f};
  const long step = (long)gridDim.x * lanes;
  // four voxels in flight per thread
  for (long v = (long)blockIdx.x * lanes + vl; v < nvox; v += 4 * step) {
    uint4 q[4];
#pragma unroll
    for (int k = 0; k < 4; k++) {
      q[k] = uint4{0u, 0u, 0u, 0u};
      if (v + k * step < nvox) q[k] = *reinterpret_cast<const uint4*>(x + (v + k * step) * ld + 8 * gr);
    }
#pragma unroll
    for (int k = 0; k < 4; k++) {
      const unsigned u[4] = {q[k].x, q[k].y, q[k].z, q[k].w};
#pragma unroll
      for (int j = 0; j < 4; j++) {
        s[2 * j] += __uint_as_float(u[j] << 16);
        s[2 * j + 1] += __uint_as_float(u[j] & 0xFFFF0000u);
      }
    }
  }
  // tree over the voxel lanes (lanes is a power of two: G divides 32)
#pragma unroll
  for (int j = 0; j < 8; j++) red[threadIdx.x * 8 + j] = s[j];
  __syncthreads();
  for (int half = lanes / 2; half > 0; half >>= 1) {
    if (vl < half) {
#pragma unroll
      for (int j = 0; j < 8; j++) red[threadIdx.x * 8 + j] += red[(threadIdx.x + half * G) * 8 + j];
    }
    __syncthreads();
  }
  if (threadIdx.x < c) atomicAdd(out + threadIdx.x, red[threadIdx.x]);
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
__global__ void k_zero_tail(float* __restrict__ dw, float* __restrict__ db, double* __restrict__ loss) {
  if (threadIdx.x < 16) dw[threadIdx.x] = 0.f;
  if (threadIdx.x == 16) db[0] = 0.f;
  if (threadIdx.x == 17) loss[0] = 0.0;
}

// gradient of the pooling: d_fine[v][c] (=|+=) d_coarse[parent][c] / 8, then
// zeroed where mask[v][c] <= 0 when a mask is given (rounded to BF16 first)
__global__ void k_pool2_bwd(const uint16_t* __restrict__ dc, int ldc, uint16_t* __restrict__ df, int ldf, int c, Vol F,
                            int accumulate, const uint16_t* __restrict__ mask, int ldm) {
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
    uint16_t r = f2bf(g);
    if (mask && !(bf2f(mask[F.vox(z, x, y) * ldm + ch]) > 0.f)) r = 0;
    df[o] = r;
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

// 8-channel (16-byte) granule forms of the pooling, upsampling and masked-add
// kernels, for channel counts, strides and slices that are multiples of 8
// channels: thread = (voxel, granule) with the granule fastest (a block of
// 256 / G voxels x G granules, G = c / 8), so a wavefront reads and writes
// whole runs of channels. Arithmetic per channel is that of the per-channel
// kernels (same order, same rounding).
// (voxel, channel offset) walk of a granule kernel over n voxels
struct GranuleWalk {
  long v, step;
  int g8;
  __device__ GranuleWalk(int c) {
    const int G = c / 8, per = blockDim.x / G;
    g8 = 8 * (int)(threadIdx.x % G);
    const int vl = (int)(threadIdx.x / G);
    v = vl < per ? (long)blockIdx.x * per + vl : LONG_MAX;
    step = (long)gridDim.x * per;
  }
};

__global__ void k_pool2_v(const uint16_t* __restrict__ src, int lds, uint16_t* __restrict__ dst, int ldd, int c, Vol F) {
  const Vol Cv{F.nz / 2, F.nx / 2, F.ny / 2};
  const long n = (long)Cv.nz * Cv.nx * Cv.ny;
  GranuleWalk gw(c);
  for (long v = gw.v; v < n; v += gw.step) {
    const int y = (int)(v % Cv.ny);
    const long r = v / Cv.ny;
    const int x = (int)(r % Cv.nx), z = (int)(r / Cv.nx);
    float s[8] = {0.f, 0.f, 0.f, 0.f, 0.f, 0.f, 0.f, 0.f};
#pragma unroll
    for (int q = 0; q < 8; q++) {
      float f[8];
      unpack8(ld16(src + F.vox(2 * z + (q >> 2), 2 * x + ((q >> 1) & 1), 2 * y + (q & 1)) * lds + gw.g8), f);
#pragma unroll
      for (int j = 0; j < 8; j++) s[j] += f[j];
    }
#pragma unroll
    for (int j = 0; j < 8; j++) s[j] *= 0.125f;
    st16(dst + v * ldd + gw.g8, pack8(s));
  }
}
__global__ void k_pool2_bwd_v(const uint16_t* __restrict__ dc, int ldc, uint16_t* __restrict__ df, int ldf, int c, Vol F,
                              int accumulate, const uint16_t* __restrict__ mask, int ldm) {
  const Vol Cv{F.nz / 2, F.nx / 2, F.ny / 2};
  const long n = (long)F.nz * F.nx * F.ny;
  GranuleWalk gw(c);
  for (long v = gw.v; v < n; v += gw.step) {
    const int y = (int)(v % F.ny);
    const long r = v / F.ny;
    const int x = (int)(r % F.nx), z = (int)(r / F.nx);
    float gv[8];
    unpack8(ld16(dc + Cv.vox(z / 2, x / 2, y / 2) * ldc + gw.g8), gv);
#pragma unroll
    for (int j = 0; j < 8; j++) gv[j] = 0.125f * gv[j];
    if (accumulate) {
      float a[8];
      unpack8(ld16(df + v * ldf + gw.g8), a);
#pragma unroll
      for (int j = 0; j < 8; j++) gv[j] += a[j];
    }
    if (mask) {
      // the BF16 rounding of the sum, then the mask (exact: a zero or the
      // rounded value), as a separate masking pass would give
      float m[8];
      unpack8(ld16(mask + v * ldm + gw.g8), m);
#pragma unroll
      for (int j = 0; j < 8; j++)
        if (!(m[j] > 0.f)) gv[j] = 0.f;
    }
    st16(df + v * ldf + gw.g8, pack8(gv));
  }
}
__global__ void k_up2_v(const uint16_t* __restrict__ cs, int ldc, uint16_t* __restrict__ fd, int ldf, int c, Vol F) {
  const Vol Cv{F.nz / 2, F.nx / 2, F.ny / 2};
  const long n = (long)F.nz * F.nx * F.ny;
  GranuleWalk gw(c);
  for (long v = gw.v; v < n; v += gw.step) {
    const int y = (int)(v % F.ny);
    const long r = v / F.ny;
    const int x = (int)(r % F.nx), z = (int)(r / F.nx);
    st16(fd + v * ldf + gw.g8, ld16(cs + Cv.vox(z / 2, x / 2, y / 2) * ldc + gw.g8));
  }
}
__global__ void k_up2_bwd_v(const uint16_t* __restrict__ df, int ldf, uint16_t* __restrict__ dc, int ldc, int c, Vol F,
                            const uint16_t* __restrict__ mask, int ldm) {
  const Vol Cv{F.nz / 2, F.nx / 2, F.ny / 2};
  const long n = (long)Cv.nz * Cv.nx * Cv.ny;
  GranuleWalk gw(c);
  for (long v = gw.v; v < n; v += gw.step) {
    const int y = (int)(v % Cv.ny);
    const long r = v / Cv.ny;
    const int x = (int)(r % Cv.nx), z = (int)(r / Cv.nx);
    float s[8] = {0.f, 0.f, 0.f, 0.f, 0.f, 0.f, 0.f, 0.f};
#pragma unroll
    for (int q = 0; q < 8; q++) {
      float f[8];
      unpack8(ld16(df + F.vox(2 * z + (q >> 2), 2 * x + ((q >> 1) & 1), 2 * y + (q & 1)) * ldf + gw.g8), f);
#pragma unroll
      for (int j = 0; j < 8; j++) s[j] += f[j];
    }
    if (mask) {
      float m[8];
      unpack8(ld16(mask + v * ldm + gw.g8), m);
#pragma unroll
      for (int j = 0; j < 8; j++)
        if (!(m[j] > 0.f)) s[j] = 0.f;
    }
    st16(dc + v * ldc + gw.g8, pack8(s));
  }
}
__global__ void k_add_masked_v(uint16_t* __restrict__ a, int lda, const uint16_t* __restrict__ b, int ldb, int c, long nvox,
                               const uint16_t* __restrict__ mask, int ldm) {
  GranuleWalk gw(c);
  for (long v = gw.v; v < nvox; v += gw.step) {
    float s[8];
    unpack8(ld16(a + v * lda + gw.g8), s);
    if (b) {
      float f[8];
      unpack8(ld16(b + v * ldb + gw.g8), f);
#pragma unroll
      for (int j = 0; j < 8; j++) s[j] += f[j];
    }
    if (mask) {
      float m[8];
      unpack8(ld16(mask + v * ldm + gw.g8), m);
#pragma unroll
      for (int j = 0; j < 8; j++)
        if (!(m[j] > 0.f)) s[j] = 0.f;
    }
    st16(a + v * lda + gw.g8, pack8(s));
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

// k_head for 16 channels in two 16-byte granules per voxel (same order of sums)
__global__ void k_head16(const uint16_t* __restrict__ act, int ld, const float* __restrict__ w, const float* __restrict__ b,
                         float* __restrict__ out, long nvox) {
  float wr[16];
#pragma unroll
  for (int ch = 0; ch < 16; ch++) wr[ch] = w[ch];
  const float b0 = b[0];
  for (long v = (long)blockIdx.x * blockDim.x + threadIdx.x; v < nvox; v += (long)gridDim.x * blockDim.x) {
    float a[16];
    unpack8(ld16(act + v * ld), *reinterpret_cast<float(*)[8]>(a));
    unpack8(ld16(act + v * ld + 8), *reinterpret_cast<float(*)[8]>(a + 8));
    float s = b0;
#pragma unroll
    for (int ch = 0; ch < 16; ch++) s += wr[ch] * a[ch];
    out[v] = s;
  }
}

// k_head_bwd for 16 channels in two 16-byte granules per voxel (ld a multiple
// of 8, act and dact 16-byte aligned); the same thread-to-voxel walk and
// arithmetic order
__global__ void k_head_bwd16(const uint16_t* __restrict__ act, int ld, const float* __restrict__ w,
                             const float* __restrict__ dout, uint16_t* __restrict__ dact, int ldd, float* __restrict__ dw,
                             float* __restrict__ db, long nvox) {
  __shared__ float red[17][256];
  float acc[17];
  for (int k = 0; k < 17; k++) acc[k] = 0.f;
  float wr[16];
#pragma unroll
  for (int ch = 0; ch < 16; ch++) wr[ch] = w[ch];
  for (long v = (long)blockIdx.x * blockDim.x + threadIdx.x; v < nvox; v += (long)gridDim.x * blockDim.x) {
    const float g = dout[v];
#pragma unroll
    for (int h = 0; h < 2; h++) {
      float a[8], d[8];
      unpack8(ld16(act + v * ld + 8 * h), a);
#pragma unroll
      for (int j = 0; j < 8; j++) {
        d[j] = a[j] > 0.f ? wr[8 * h + j] * g : 0.f;
        acc[8 * h + j] += a[j] * g;
      }
      st16(dact + v * ldd + 8 * h, pack8(d));
    }
    acc[16] += g;
  }
  // blockDim.x == 256: per sum, 16 partial sums of 16 threads, then their total
  for (int k = 0; k < 17; k++) red[k][threadIdx.x] = acc[k];
  __syncthreads();
  // 17 x 16 partial sums over 256 threads: item e = threadIdx.x + 256 r
  float part[2] = {0.f, 0.f};
#pragma unroll
  for (int r = 0; r < 2; r++) {
    const int e = (int)threadIdx.x + 256 * r;
    if (e < 17 * 16) {
      const int k = e >> 4, q0 = (e & 15) * 16;
      for (int q = q0; q < q0 + 16; q++) part[r] += red[k][q];
    }
  }
  __syncthreads();
#pragma unroll
  for (int r = 0; r < 2; r++) {
    const int e = (int)threadIdx.x + 256 * r;
    if (e < 17 * 16) red[e >> 4][e & 15] = part[r];
  }
  __syncthreads();
  if (threadIdx.x < 17) {
    float t = 0.f;
    for (int q = 0; q < 16; q++) t += red[threadIdx.x][q];
    if (threadIdx.x < 16) atomicAdd(dw + threadIdx.x, t);
    if (threadIdx.x == 16) atomicAdd(db, t);
  }
}

// The loss and the head's gradient in one pass (16 channels): per voxel the
// loss term d = (u - t) m (u given: the gated output) or (o - t) m (u NULL:
// the loss on the head output), its gradient 2 d / count, through the gate
// where u is given (zero outside the clamp), then the head gradient as
// k_head_bwd16. The same FP32 operations per voxel as efwi_nn_mse,
// efwi_nn_gate_bwd and efwi_nn_head_bwd one after another; the sums (loss in
// FP64, dw, db in FP32) add in another order.
__global__ __launch_bounds__(256) void k_loss_head16(const uint16_t* __restrict__ act, int ld, const float* __restrict__ w,
                                                     const float* __restrict__ o, const float* __restrict__ u,
                                                     const float* __restrict__ target, const float* __restrict__ m,
                                                     long nvox, float inv_count, uint16_t* __restrict__ dact, int ldd,
                                                     float* __restrict__ dw, float* __restrict__ db,
                                                     double* __restrict__ loss);
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

__global__ __launch_bounds__(256) void k_loss_head16(const uint16_t* __restrict__ act, int ld, const float* __restrict__ w,
                                                     const float* __restrict__ o, const float* __restrict__ u,
                                                     const float* __restrict__ target, const float* __restrict__ m,
                                                     long nvox, float inv_count, uint16_t* __restrict__ dact, int ldd,
                                                     float* __restrict__ dw, float* __restrict__ db,
                                                     double* __restrict__ loss) {
  __shared__ float red[17][256];
  __shared__ double redl[256];
  float acc[17];
  for (int k = 0; k < 17; k++) acc[k] = 0.f;
  double sl = 0.0;
  float wr[16];
#pragma unroll
  for (int ch = 0; ch < 16; ch++) wr[ch] = w[ch];
  for (long v = (long)blockIdx.x * blockDim.x + threadIdx.x; v < nvox; v += (long)gridDim.x * blockDim.x) {
    float g;
    if (u) {
      const float uv = u[v];
      const float d = (uv - target[v]) * m[v];
      sl += (double)d * d;
      const float dl = 2.f * d * inv_count;
      g = (o[v] > kGateLo && o[v] < kGateHi) ? dl * uv : 0.f;
    } else {
      const float d = (o[v] - target[v]) * m[v];
      sl += (double)d * d;
      g = 2.f * d * inv_count;
    }
#pragma unroll
    for (int h = 0; h < 2; h++) {
      float a[8], dd[8];
      unpack8(ld16(act + v * ld + 8 * h), a);
#pragma unroll
      for (int j = 0; j < 8; j++) {
        dd[j] = a[j] > 0.f ? wr[8 * h + j] * g : 0.f;
        acc[8 * h + j] += a[j] * g;
      }
      st16(dact + v * ldd + 8 * h, pack8(dd));
    }
    acc[16] += g;
  }
  for (int k = 0; k < 17; k++) red[k][threadIdx.x] = acc[k];
  redl[threadIdx.x] = sl;
  __syncthreads();
  // 17 x 16 partial sums over 256 threads: item e = threadIdx.x + 256 r
  float part[2] = {0.f, 0.f};
#pragma unroll
  for (int r = 0; r < 2; r++) {
    const int e = (int)threadIdx.x + 256 * r;
    if (e < 17 * 16) {
      const int k = e >> 4, q0 = (e & 15) * 16;
      for (int q = q0; q < q0 + 16; q++) part[r] += red[k][q];
    }
  }
  for (int k = 128; k > 0; k >>= 1) {
    if ((int)threadIdx.x < k) redl[threadIdx.x] += redl[threadIdx.x + k];
    __syncthreads();
  }
#pragma unroll
  for (int r = 0; r < 2; r++) {
    const int e = (int)threadIdx.x + 256 * r;
    if (e < 17 * 16) red[e >> 4][e & 15] = part[r];
  }
  __syncthreads();
  if (threadIdx.x < 17) {
    float t = 0.f;
    for (int q = 0; q < 16; q++) t += red[threadIdx.x][q];
    if (threadIdx.x < 16) atomicAdd(dw + threadIdx.x, t);
    if (threadIdx.x == 16) atomicAdd(db, t);
  }
  if (threadIdx.x == 0) atomicAdd(loss, redl[0] * inv_count);
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

// resident 256-thread workgroups of a kernel per CU (1 if the query fails)
template <typename K>
int blocks_per_cu(K kernel) {
  int n = 0;
  if (hipOccupancyMaxActiveBlocksPerMultiprocessor(&n, kernel, 256, 0) != hipSuccess || n < 1) n = 1;
  return n;
}

// compute units of the current device (256 if the query fails)
int device_cus() {
  int dev = 0, n = 0;
  if (hipGetDevice(&dev) != hipSuccess ||
      hipDeviceGetAttribute(&n, hipDeviceAttributeMultiprocessorCount, dev) != hipSuccess || n < 1)
    n = 256;
  return n;
}

// a channel slice that the 16-byte granule kernels can take
inline bool granules_ok(int c, const void* p, int ld) {
  return c % 8 == 0 && c <= 2048 && ld % 8 == 0 && (uintptr_t)p % 16 == 0;
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

int efwi_nn_conv3_cat(const uint16_t* src, int lds, int kc1, int up1, const uint16_t* src2, int lds2, int kc2,
                      const uint16_t* w, const float* bias, uint16_t* dst, int ldd, int nc1, uint16_t* dst2, int ldd2,
                      int nc2, const uint16_t* mask, int ldm, int nz, int nx, int ny, int mode, int relu,
                      void* stream) {
  if (up1 && (nz % 2 || nx % 2 || ny % 2)) return EFWI_NN_ERR_ARG;
  if (!src2) kc2 = 0;
  if (!dst2) nc2 = 0;
  const int kc = kc1 + kc2, nc = nc1 + nc2;
  if (!src || !w || !dst || kc1 <= 0 || nc1 <= 0 || kc1 % 16 || nc1 % 16 || lds % 16 || ldd % 16 || lds < kc1 ||
      ldd < nc1 || (mask && (ldm % 16 || ldm < nc)) || nz <= 0 || nx <= 0 || ny <= 0 || (mode != 0 && mode != 1) ||
      ((uintptr_t)src | (uintptr_t)w) % 16)
    return EFWI_NN_ERR_ARG;
  if (src2 && (kc2 <= 0 || kc2 % 16 || lds2 % 16 || lds2 < kc2 || (uintptr_t)src2 % 16)) return EFWI_NN_ERR_ARG;
  if (dst2 && (nc2 <= 0 || nc2 % 16 || ldd2 % 16 || ldd2 < nc2)) return EFWI_NN_ERR_ARG;
  hipStream_t s = (hipStream_t)stream;
  const Vol V{nz, nx, ny};
  // bricks of 2 x 2 rows of 64, 32 or 16 voxels along y; 1 to 4 output blocks
  // per wavefront (a divisor of the block count, so no block is partial)
  const int MB = ny >= 64 ? 4 : ny >= 32 ? 2 : 1;
  const int nbt = nc / 16;
  const int TZ = conv_tz(MB);
  const int ntz = (nz + TZ - 1) / TZ, ntx = (nx + 1) / 2, nty = (ny + 16 * MB - 1) / (16 * MB);
  const long bricks = (long)ntz * ntx * nty;
  // where the bricks times the output-block groups would give fewer than four
  // workgroups per CU (the coarser levels), one block per workgroup: more
  // workgroups, whose wavefronts hide each other's load latency (48x64x64
  // voxels: 32->32 27.7 -> 22.5 us, 96->32 82 -> 75 us)
  int NB = nbt % 4 == 0 ? 4 : nbt % 3 == 0 ? 3 : nbt % 2 == 0 ? 2 : 1;
  if (bricks * (nbt / NB) < 4L * device_cus()) NB = 1;
  const int per_xcd = (int)((bricks + 7) / 8);
  const dim3 grid((unsigned)(8L * per_xcd), (unsigned)(nbt / NB));
  // outputs (and mask) as 16-byte granules where the pointers allow
  const int vec_out = (uintptr_t)dst % 16 == 0 && (uintptr_t)dst2 % 16 == 0 && (!mask || (uintptr_t)mask % 16 == 0);
#define EFWI_NN_CONV_LAUNCH(CK, M, B, MODE)                                                                       \
  hipLaunchKernelGGL((k_conv3<CK, M, B, MODE>), grid, dim3(128 * conv_tz(M)), 0, s, src, lds, kc, w, bias, dst,  \
                     ldd, nc, mask, ldm, V, relu, ntz, ntx, nty, per_xcd, vec_out, src2, lds2, kc1, dst2, ldd2,   \
                     nc1, up1 ? 1 : 0)
#define EFWI_NN_CONV_MODE(CK, M, B)                         \
  if (mode == 0) { EFWI_NN_CONV_LAUNCH(CK, M, B, 0); }    \
  else { EFWI_NN_CONV_LAUNCH(CK, M, B, 1); }
#define EFWI_NN_CONV_NB(CK, M)                              \
  if (NB == 4) { EFWI_NN_CONV_MODE(CK, M, 4) }             \
  else if (NB == 3) { EFWI_NN_CONV_MODE(CK, M, 3) }        \
  else if (NB == 2) { EFWI_NN_CONV_MODE(CK, M, 2) }        \
  else { EFWI_NN_CONV_MODE(CK, M, 1) }
#define EFWI_NN_CONV_MB(CK)                                 \
  if (MB == 4) { EFWI_NN_CONV_NB(CK, 4) }                  \
  else if (MB == 2) { EFWI_NN_CONV_NB(CK, 2) }             \
  else { EFWI_NN_CONV_NB(CK, 1) }
  // 16-channel chunks (the kernel also takes CK = 32): their halos leave room
  // for more workgroups per CU than 32-channel ones (with 4 z-rows 51 against
  // 101 KB of LDS), which measured faster at every level (48x64x64 voxels,
  // 32->32 29 -> 27 us, 96->32 93 -> 77 us; 24x32x32, 64->64 26.5 -> 18.8 us)
  EFWI_NN_CONV_MB(16)
#undef EFWI_NN_CONV_MB
#undef EFWI_NN_CONV_NB
#undef EFWI_NN_CONV_MODE
#undef EFWI_NN_CONV_LAUNCH
  return nn_check(s);
}

int efwi_nn_conv3(const uint16_t* src, int lds, int kc, const uint16_t* w, const float* bias, uint16_t* dst, int ldd,
                  int nc, const uint16_t* mask, int ldm, int nz, int nx, int ny, int mode, int relu, void* stream) {
  return efwi_nn_conv3_cat(src, lds, kc, 0, nullptr, 0, 0, w, bias, dst, ldd, nc, nullptr, 0, 0, mask, ldm, nz, nx,
                           ny, mode, relu, stream);
}

int efwi_nn_flip_weights(const uint16_t* w, int kc, int nc, uint16_t* wt, void* stream) {
  if (!w || !wt || kc <= 0 || nc <= 0 || w == wt) return EFWI_NN_ERR_ARG;
  hipStream_t s = (hipStream_t)stream;
  hipLaunchKernelGGL(k_flip_weights, dim3(grid_for(27L * kc * nc, 256)), dim3(256), 0, s, w, kc, nc, wt);
  return nn_check(s);
}

int efwi_nn_flip_weights_n(int n, const uint16_t* const* w, const int* kc, const int* nc, uint16_t* const* wt,
                           void* stream) {
  if (n < 1 || n > kFlipMax || !w || !kc || !nc || !wt) return EFWI_NN_ERR_ARG;
  FlipSet S{};
  long most = 0;
  for (int l = 0; l < n; l++) {
    if (!w[l] || !wt[l] || kc[l] <= 0 || nc[l] <= 0 || w[l] == wt[l]) return EFWI_NN_ERR_ARG;
    S.w[l] = w[l];
    S.wt[l] = wt[l];
    S.kc[l] = kc[l];
    S.nc[l] = nc[l];
    most = std::max(most, 27L * kc[l] * nc[l]);
  }
  hipStream_t s = (hipStream_t)stream;
  hipLaunchKernelGGL(k_flip_many, dim3(grid_for(most, 256), (unsigned)n), dim3(256), 0, s, S);
  return nn_check(s);
}

// bricks of the weight gradient: 2 x 2 rows of 64 voxels along y, or of 32
// where the rows are that short
static inline int wgrad_ty(int ny) { return ny > 32 ? 64 : 32; }
static inline long wgrad_bricks(int nz, int nx, int ny) {
  const int ty = wgrad_ty(ny);
  return (long)((nz + 1) / 2) * ((nx + 1) / 2) * ((ny + ty - 1) / ty);
}

long efwi_nn_wgrad_scratch(int mc, int kc, int nz, int nx, int ny, int per) {
  if (per <= 0 || nz <= 0 || nx <= 0 || ny <= 0) return 0;
  const long chunks = (wgrad_bricks(nz, nx, ny) + per - 1) / per;
  return chunks * (27L * mc * kc + mc);
}

static bool wgrad_args_ok(const uint16_t* dpre, int ldp, int mc, const uint16_t* src, int lds, int kc, float* dw,
                          int nz, int nx, int ny, int per) {
  return dpre && src && dw && mc % 16 == 0 && kc % 16 == 0 && mc > 0 && kc > 0 && per > 0 &&
         ldp % 16 == 0 && lds % 16 == 0 && ldp >= mc && lds >= kc && nz > 0 && nx > 0 && ny > 0 &&
         ((uintptr_t)dpre | (uintptr_t)src) % 16 == 0;
}
static int wgrad_launch(const uint16_t* dpre, int ldp, int mc, const uint16_t* src, int lds, int kc, float* scratch,
                        float* dw, float* db, int nz, int nx, int ny, int per, hipStream_t s,
                        const uint16_t* src2 = nullptr, int lds2 = 0, int kc1 = 0, int up1 = 0);

int efwi_nn_wgrad(const uint16_t* dpre, int ldp, int mc, const uint16_t* src, int lds, int kc, float* scratch,
                  long scratch_floats, float* dw, float* db, int nz, int nx, int ny, int per, void* stream) {
  if (!scratch || !wgrad_args_ok(dpre, ldp, mc, src, lds, kc, dw, nz, nx, ny, per)) return EFWI_NN_ERR_ARG;
  if (efwi_nn_wgrad_scratch(mc, kc, nz, nx, ny, per) > scratch_floats) return EFWI_NN_ERR_ARG;
  return wgrad_launch(dpre, ldp, mc, src, lds, kc, scratch, dw, db, nz, nx, ny, per, (hipStream_t)stream);
}

int efwi_nn_wgrad_cat(const uint16_t* dpre, int ldp, int mc, const uint16_t* src, int lds, int kc1, int up1,
                      const uint16_t* src2, int lds2, int kc2, float* scratch, long scratch_floats, float* dw,
                      float* db, int nz, int nx, int ny, int per, void* stream) {
  if (up1 && (nz % 2 || nx % 2 || ny % 2)) return EFWI_NN_ERR_ARG;
  const int kc = kc1 + kc2;
  if (!dpre || !src || !src2 || !dw || !scratch || mc <= 0 || mc % 16 || kc1 <= 0 || kc1 % 16 || kc2 <= 0 ||
      kc2 % 16 || ldp % 16 || ldp < mc || lds % 16 || lds < kc1 || lds2 % 16 || lds2 < kc2 || per <= 0 || nz <= 0 ||
      nx <= 0 || ny <= 0 || ((uintptr_t)dpre | (uintptr_t)src | (uintptr_t)src2) % 16)
    return EFWI_NN_ERR_ARG;
  if (efwi_nn_wgrad_scratch(mc, kc, nz, nx, ny, per) > scratch_floats) return EFWI_NN_ERR_ARG;
  return wgrad_launch(dpre, ldp, mc, src, lds, kc, scratch, dw, db, nz, nx, ny, per, (hipStream_t)stream, src2, lds2,
                      kc1, up1 ? 1 : 0);
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
                          float* db, int nz, int nx, int ny, int per, void* stream) {
  if (!wgrad_args_ok(dpre, ldp, mc, src, lds, kc, dw, nz, nx, ny, per)) return EFWI_NN_ERR_ARG;
  if (efwi_nn_wgrad_scratch(mc, kc, nz, nx, ny, per) > EFWI_NN_WGRAD_STATIC_FLOATS) return EFWI_NN_ERR_SCRATCH;
  int err = 0;
  float* ptr = static_scratch(&err);
  if (!ptr) return err;
  return wgrad_launch(dpre, ldp, mc, src, lds, kc, ptr, dw, db, nz, nx, ny, per, (hipStream_t)stream);
}

static int wgrad_launch(const uint16_t* dpre, int ldp, int mc, const uint16_t* src, int lds, int kc, float* scratch,
                        float* dw, float* db, int nz, int nx, int ny, int per, hipStream_t s,
                        const uint16_t* src2, int lds2, int kc1, int up1) {
  const Vol V{nz, nx, ny};
  if (!src2) kc1 = kc;
  const int ty = wgrad_ty(ny);
  const int ntz = (nz + 1) / 2, ntx = (nx + 1) / 2, nty = (ny + ty - 1) / ty;
  const int chunks = (int)((wgrad_bricks(nz, nx, ny) + per - 1) / per);
  const dim3 grid((unsigned)chunks, (unsigned)(mc / 16), (unsigned)(kc / 16));
  if (ty == 64)
    hipLaunchKernelGGL((k_wgrad<64>), grid, dim3(256), 0, s, dpre, ldp, mc, src, lds, kc, scratch, dw, db, V, ntz, ntx,
                       nty, per, src2, lds2, kc1, up1);
  else
    hipLaunchKernelGGL((k_wgrad<32>), grid, dim3(256), 0, s, dpre, ldp, mc, src, lds, kc, scratch, dw, db, V, ntz, ntx,
                       nty, per, src2, lds2, kc1, up1);
  const long width = 27L * mc * kc, total = width + (db ? mc : 0);
  hipLaunchKernelGGL(k_sum_chunks, dim3(grid_for(total, 256, 1L << 30), (chunks + kChunkGroup - 1) / kChunkGroup),
                     dim3(256), 0, s, scratch, width + mc, width, total, chunks, dw, db);
  return nn_check(s);
}

int efwi_nn_chan_sum(const uint16_t* x, int ld, int c, long nvox, float* out, void* stream) {
  if (!x || !out || c <= 0 || c > 256 || 256 % c || ld < c) return EFWI_NN_ERR_ARG;
  hipStream_t s = (hipStream_t)stream;
  hipLaunchKernelGGL(k_zero, dim3((c + 255) / 256), dim3(256), 0, s, reinterpret_cast<unsigned*>(out), c);
  if (c % 8 == 0 && ld % 8 == 0 && (uintptr_t)x % 16 == 0)
    hipLaunchKernelGGL(k_chan_sum8, dim3(grid_for(nvox * c / 8, 4096, 256)), dim3(1024), 0, s, x, ld, c, nvox, out);
  else
    hipLaunchKernelGGL(k_chan_sum, dim3(grid_for(nvox * c, 256, 2048)), dim3(256), 0, s, x, ld, c, nvox, out);
  return nn_check(s);
}

int efwi_nn_pool2(const uint16_t* src, int lds, uint16_t* dst, int ldd, int c, int nz, int nx, int ny, void* stream) {
  if (!src || !dst || nz % 2 || nx % 2 || ny % 2 || c <= 0) return EFWI_NN_ERR_ARG;
  hipStream_t s = (hipStream_t)stream;
  if (granules_ok(c, src, lds) && granules_ok(c, dst, ldd))
    hipLaunchKernelGGL(k_pool2_v, dim3(grid_for((long)nz * nx * ny / 8, 256 / (c / 8))), dim3(256 / (c / 8) * (c / 8)), 0, s, src, lds, dst, ldd, c,
                       Vol{nz, nx, ny});
  else
    hipLaunchKernelGGL(k_pool2, dim3(grid_for((long)nz * nx * ny / 8 * c, 256)), dim3(256), 0, s, src, lds, dst, ldd, c,
                       Vol{nz, nx, ny});
  return nn_check(s);
}

int efwi_nn_pool2_bwd(const uint16_t* dc, int ldc, uint16_t* df, int ldf, int c, int nz, int nx, int ny, int accumulate,
                      const uint16_t* mask, int ldm, void* stream) {
  if (!dc || !df || nz % 2 || nx % 2 || ny % 2 || c <= 0 || (mask && ldm < c)) return EFWI_NN_ERR_ARG;
  hipStream_t s = (hipStream_t)stream;
  if (granules_ok(c, dc, ldc) && granules_ok(c, df, ldf) && (!mask || granules_ok(c, mask, ldm)))
    hipLaunchKernelGGL(k_pool2_bwd_v, dim3(grid_for((long)nz * nx * ny, 256 / (c / 8))), dim3(256 / (c / 8) * (c / 8)), 0, s, dc, ldc, df, ldf, c,
                       Vol{nz, nx, ny}, accumulate, mask, ldm);
  else
    hipLaunchKernelGGL(k_pool2_bwd, dim3(grid_for((long)nz * nx * ny * c, 256)), dim3(256), 0, s, dc, ldc, df, ldf, c,
                       Vol{nz, nx, ny}, accumulate, mask, ldm);
  return nn_check(s);
}

int efwi_nn_up2(const uint16_t* cs, int ldc, uint16_t* fd, int ldf, int c, int nz, int nx, int ny, void* stream) {
  if (!cs || !fd || nz % 2 || nx % 2 || ny % 2 || c <= 0) return EFWI_NN_ERR_ARG;
  hipStream_t s = (hipStream_t)stream;
  if (granules_ok(c, cs, ldc) && granules_ok(c, fd, ldf))
    hipLaunchKernelGGL(k_up2_v, dim3(grid_for((long)nz * nx * ny, 256 / (c / 8))), dim3(256 / (c / 8) * (c / 8)), 0, s, cs, ldc, fd, ldf, c,
                       Vol{nz, nx, ny});
  else
    hipLaunchKernelGGL(k_up2, dim3(grid_for((long)nz * nx * ny * c, 256)), dim3(256), 0, s, cs, ldc, fd, ldf, c,
                       Vol{nz, nx, ny});
  return nn_check(s);
}

int efwi_nn_up2_bwd(const uint16_t* df, int ldf, uint16_t* dc, int ldc, int c, int nz, int nx, int ny,
                    const uint16_t* mask, int ldm, void* stream) {
  if (!df || !dc || nz % 2 || nx % 2 || ny % 2 || c <= 0) return EFWI_NN_ERR_ARG;
  hipStream_t s = (hipStream_t)stream;
  if (granules_ok(c, df, ldf) && granules_ok(c, dc, ldc) && (!mask || granules_ok(c, mask, ldm)))
    hipLaunchKernelGGL(k_up2_bwd_v, dim3(grid_for((long)nz * nx * ny / 8, 256 / (c / 8))), dim3(256 / (c / 8) * (c / 8)), 0, s, df, ldf, dc, ldc, c,
                       Vol{nz, nx, ny}, mask, ldm);
  else
    hipLaunchKernelGGL(k_up2_bwd, dim3(grid_for((long)nz * nx * ny / 8 * c, 256)), dim3(256), 0, s, df, ldf, dc, ldc, c,
                       Vol{nz, nx, ny}, mask, ldm);
  return nn_check(s);
}

int efwi_nn_add(uint16_t* a, int lda, const uint16_t* b, int ldb, int c, long nvox, const uint16_t* mask, int ldm,
                void* stream) {
  if (!a || c <= 0) return EFWI_NN_ERR_ARG;
  hipStream_t s = (hipStream_t)stream;
  if (granules_ok(c, a, lda) && (!b || granules_ok(c, b, ldb)) && (!mask || granules_ok(c, mask, ldm)))
    hipLaunchKernelGGL(k_add_masked_v, dim3(grid_for(nvox, 256 / (c / 8))), dim3(256 / (c / 8) * (c / 8)), 0, s, a, lda, b, ldb, c, nvox, mask, ldm);
  else
    hipLaunchKernelGGL(k_add_masked, dim3(grid_for(nvox * c, 256)), dim3(256), 0, s, a, lda, b, ldb, c, nvox, mask, ldm);
  return nn_check(s);
}

int efwi_nn_head(const uint16_t* act, int ld, int c, const float* w, const float* b, float* out, long nvox, void* stream) {
  if (!act || !w || !b || !out || c <= 0 || c > 16) return EFWI_NN_ERR_ARG;
  hipStream_t s = (hipStream_t)stream;
  if (c == 16 && granules_ok(16, act, ld))
    hipLaunchKernelGGL(k_head16, dim3(grid_for(nvox, 256)), dim3(256), 0, s, act, ld, w, b, out, nvox);
  else
    hipLaunchKernelGGL(k_head, dim3(grid_for(nvox, 256)), dim3(256), 0, s, act, ld, c, w, b, out, nvox);
  return nn_check(s);
}

int efwi_nn_head_bwd(const uint16_t* act, int ld, int c, const float* w, const float* dout, uint16_t* dact, int ldd,
                     float* dw, float* db, long nvox, void* stream) {
  if (!act || !w || !dout || !dact || !dw || !db || c <= 0 || c > 16) return EFWI_NN_ERR_ARG;
  hipStream_t s = (hipStream_t)stream;
  hipLaunchKernelGGL(k_zero, dim3(1), dim3(64), 0, s, reinterpret_cast<unsigned*>(dw), c);
  hipLaunchKernelGGL(k_zero, dim3(1), dim3(64), 0, s, reinterpret_cast<unsigned*>(db), 1);
  if (c == 16 && granules_ok(16, act, ld) && granules_ok(16, dact, ldd))
    hipLaunchKernelGGL(k_head_bwd16, dim3(grid_for(nvox, 256, 2048)), dim3(256), 0, s, act, ld, w, dout, dact, ldd, dw, db,
                       nvox);
  else
    hipLaunchKernelGGL(k_head_bwd, dim3(grid_for(nvox, 64, 2048)), dim3(64), 0, s, act, ld, c, w, dout, dact, ldd, dw, db,
                       nvox);
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

int efwi_nn_loss_head_bwd(const uint16_t* act, int ld, const float* w, const float* o, const float* u,
                          const float* target, const float* mask, long nvox, double count, uint16_t* dact, int ldd,
                          float* dw, float* db, double* loss, void* stream) {
  if (!act || !w || !o || !target || !mask || !dact || !dw || !db || !loss || !(count > 0) || nvox <= 0 ||
      !granules_ok(16, act, ld) || !granules_ok(16, dact, ldd))
    return EFWI_NN_ERR_ARG;
  hipStream_t s = (hipStream_t)stream;
  hipLaunchKernelGGL(k_zero_tail, dim3(1), dim3(64), 0, s, dw, db, loss);
  hipLaunchKernelGGL(k_loss_head16, dim3(grid_for(nvox, 256, 2048)), dim3(256), 0, s, act, ld, w, o, u, target, mask,
                     nvox, (float)(1.0 / count), dact, ldd, dw, db, loss);
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
