// pmx_learned.hip -- kernels of the learned post-smoother (see pmx_learned.hpp).
//
// Convolutions are implicit GEMMs on v_mfma_f32_16x16x32_bf16: a wavefront
// owns 16 consecutive lattice nodes of one x-line (the 16 rows of the MFMA
// A operand) and all 16 output channels (the columns of B); the K dimension
// runs over (tap, input channel) pairs, 32 per instruction.  Activations are
// stored node-major with the channels contiguous ([node][16] BF16, 32 bytes),
// so the 8 consecutive K elements a lane supplies are one 16-byte load.
#include <hip/hip_runtime.h>

#include <algorithm>
#include <cmath>
#include <cstring>
#include <stdexcept>
#include <string>

#include "pmx_kernels.hpp"
#include "pmx_learned.hpp"
#include "pmx_ops.hpp"

namespace pmx {

typedef __bf16 bf16x8 __attribute__((ext_vector_type(8)));
typedef uint16_t u16x8 __attribute__((ext_vector_type(8)));
typedef float floatx4 __attribute__((ext_vector_type(4)));

int net_total_params() {
  int s = 0;
  for (int l = 0; l < kNetLayers; l++) s += net_psize(l);
  return s;
}
int net_param_offset(int l) {
  int s = 0;
  for (int k = 0; k < l; k++) s += net_psize(k);
  return s;
}

// FP32 -> BF16, round to nearest even (NaN stays NaN): the conversion
// instruction of gfx950 (v_cvt_pk_bf16_f32)
__device__ __forceinline__ uint16_t f2bf(float f) { return __builtin_bit_cast(uint16_t, (__bf16)f); }
__device__ __forceinline__ float bf2f(uint16_t h) { return __uint_as_float(((uint32_t)h) << 16); }
// erf by Abramowitz & Stegun 7.1.26 (|error| <= 1.5e-7, far below the BF16
// rounding of the activations): one exponential and one reciprocal (the
// hardware's 1-ulp v_rcp_f32)
__device__ __forceinline__ float erf_as(float x) {
  const float ax = fabsf(x);
  const float t = __builtin_amdgcn_rcpf(1.0f + 0.3275911f * ax);
  const float poly =
      t * (0.254829592f + t * (-0.284496736f + t * (1.421413741f + t * (-1.453152027f + t * 1.061405429f))));
  const float r = 1.0f - poly * __expf(-ax * ax);
  return copysignf(r, x);
}
__device__ __forceinline__ float gelu(float v) { return 0.5f * v * (1.0f + erf_as(v * 0.70710678118654752f)); }
__device__ __forceinline__ float gelu_d(float v) {
  return 0.5f * (1.0f + erff(v * 0.70710678118654752f)) + v * 0.39894228040143268f * expf(-0.5f * v * v);
}
// tap t = (dx+1) + 3 (dy+1) + 9 (dz+1)
__device__ __forceinline__ void tap_off(int t, int &dx, int &dy, int &dz) {
  dx = t % 3 - 1;
  dy = (t / 3) % 3 - 1;
  dz = t / 9 - 1;
}

// ------------------------------------------------------------------ conv
// A workgroup (4 wavefronts) computes a brick of 16 x 4 x 4 output nodes.  The
// input brick with its one-node halo (18 x 6 x 6 nodes, all channels) is
// staged once in LDS; every wavefront then runs 4 x-lines of 16 nodes (4
// independent MFMA accumulation chains) reading its A fragments from LDS.
// The B fragments (weights) stay in registers for the whole brick.
// MODE 0: hidden layer, out_bf = bf16(GELU(conv + b)), pre = conv + b (if set)
// MODE 1: output layer (COUT = 1), out_f = conv + b
// MODE 3: output layer (COUT = 1) rescaled into the correction:
//         out_f = (conv + b) * (*nrm / sq) on interior nodes, 0 on the boundary
// MODE 2: input-gradient convolution, out_bf = bf16(conv), no bias
constexpr int kBX = 16, kBY = 4, kBZ = 4;
constexpr int kHX = kBX + 2, kHY = kBY + 2, kHZ = kBZ + 2, kHalo = kHX * kHY * kHZ;
// LDS offset (elements) of tap t relative to the centre tap in the halo brick
template <int CIN> __host__ __device__ constexpr int tap_lds(int t) {
  return ((t % 3 - 1) + kHX * ((t / 3) % 3 - 1 + kHY * (t / 9 - 1))) * CIN;
}

// IN0 (first layer of a learned step, CIN = kNetIn): the input is built while
// it is staged: `in` holds the operator features and channel 0 becomes
// bf16(uin * sq / ||u||), ||u|| reduced by every workgroup from the npin
// partials pin in a fixed order (a strided sum per thread, then block_sum),
// so every workgroup gets the same value; workgroup 0 stores it in nrmout
// for the output layer's scaling.
template <int CIN, int COUT, int MODE, bool IN0 = false>
__global__ __launch_bounds__(256) void k_conv(int M, const uint16_t *__restrict__ in, const uint16_t *__restrict__ w,
                                              const float *__restrict__ bias, float *__restrict__ pre,
                                              uint16_t *__restrict__ out_bf, float *__restrict__ out_f,
                                              const double *__restrict__ nrm = nullptr, double sq = 1.0,
                                              const float *__restrict__ uin = nullptr,
                                              const double *__restrict__ pin = nullptr, int npin = 0,
                                              double *__restrict__ nrmout = nullptr) {
  // K chunks of 32: CIN 1 -> taps 0..26 in one chunk; CIN 8 -> 4 taps per
  // chunk (7 chunks); CIN 16 -> 2 taps per chunk (14 chunks)
  constexpr int TPC = 32 / CIN;
  constexpr int NCH = CIN == 1 ? 1 : (kTaps + TPC - 1) / TPC;
  __shared__ __attribute__((aligned(16))) uint16_t sh[kHalo * CIN];
  const int nbx = (M + kBX - 1) / kBX, nby = (M + kBY - 1) / kBY;
  const int bx = blockIdx.x % nbx, by = (blockIdx.x / nbx) % nby, bz = blockIdx.x / (nbx * nby);
  const int X0 = bx * kBX, Y0 = by * kBY, Z0 = bz * kBZ;
  float sc_in = 1.f;
  if constexpr (IN0) {
    static_assert(CIN == kNetIn, "IN0: the first layer");
    __shared__ double sh_norm;
    double s = 0;
    for (int i = threadIdx.x; i < npin; i += 256) s += pin[i];
    s = block_sum<double>(s);
    if (threadIdx.x == 0) {
      sh_norm = sqrt(s);
      if (blockIdx.x == 0) nrmout[0] = sh_norm;
    }
    __syncthreads();
    const double nv = sh_norm;
    sc_in = nv > 0 ? (float)(sq / nv) : 0.f;
  }
  // stage the halo brick (zero outside the lattice)
  if (CIN == 1) {
    for (int q = threadIdx.x; q < kHalo; q += 256) {
      const int hx = q % kHX, hy = (q / kHX) % kHY, hz = q / (kHX * kHY);
      const int x = X0 - 1 + hx, y = Y0 - 1 + hy, z = Z0 - 1 + hz;
      sh[q] = (x >= 0 && y >= 0 && z >= 0 && x < M && y < M && z < M) ? in[x + (int64_t)M * (y + (int64_t)M * z)]
                                                                      : (uint16_t)0;
    }
  } else {
    constexpr int PIECES = CIN / 8;  // 16-byte pieces per node
    for (int q = threadIdx.x; q < kHalo * PIECES; q += 256) {
      const int n = q / PIECES, piece = q % PIECES;
      const int hx = n % kHX, hy = (n / kHX) % kHY, hz = n / (kHX * kHY);
      const int x = X0 - 1 + hx, y = Y0 - 1 + hy, z = Z0 - 1 + hz;
      u16x8 v = {0, 0, 0, 0, 0, 0, 0, 0};
      if (x >= 0 && y >= 0 && z >= 0 && x < M && y < M && z < M) {
        const int64_t node = x + (int64_t)M * (y + (int64_t)M * z);
        v = *reinterpret_cast<const u16x8 *>(in + node * CIN + 8 * piece);
        if constexpr (IN0) {
          if (piece == 0) v[0] = f2bf(uin[node] * sc_in);
        }
      }
      *reinterpret_cast<u16x8 *>(sh + n * CIN + 8 * piece) = v;
    }
  }
  const int wave = threadIdx.x >> 6, l = threadIdx.x & 63;
  const int row = l & 15, g = l >> 4;
  // weights -> registers: chunk c holds k = 32 c + 8 g + j
  u16x8 bw[NCH];
#pragma unroll
  for (int c = 0; c < NCH; c++) {
    u16x8 b = {0, 0, 0, 0, 0, 0, 0, 0};
    if (CIN == 1) {
#pragma unroll
      for (int j = 0; j < 8; j++) {
        const int t = 8 * g + j;
        b[j] = t < kTaps ? w[row * kTaps + t] : (uint16_t)0;
      }
    } else {
      const int t = TPC * c + g / (4 / TPC);
      const int cb = 8 * (g % (4 / TPC));
      if (t < kTaps) b = *reinterpret_cast<const u16x8 *>(w + ((int64_t)row * kTaps + t) * CIN + cb);
    }
    bw[c] = b;
  }
  __syncthreads();
  // LDS element of this lane's A fragment for tap 13 (the centre) at q = 0;
  // tap t and line q add tap_lds<CIN>(t) + q kHX kHY CIN, compile-time
  // constants once the chunk loop is unrolled (the lane group only selects
  // among the TPC taps of a chunk)
  const int gsel = CIN == 1 ? 0 : g / (4 / TPC);
  const int abase = ((row + 1) + kHX * ((wave + 1) + kHY)) * CIN + (CIN == 1 ? 0 : 8 * (g % (4 / TPC)));
  // x-lines (y, z) = (wave, 0..3) of the brick
  floatx4 acc[kBZ];
#pragma unroll
  for (int q = 0; q < kBZ; q++) acc[q] = floatx4{0.f, 0.f, 0.f, 0.f};
#pragma unroll
  for (int c = 0; c < NCH; c++) {
#pragma unroll
    for (int q = 0; q < kBZ; q++) {
      u16x8 a = {0, 0, 0, 0, 0, 0, 0, 0};
      if (CIN == 1) {
#pragma unroll
        for (int j = 0; j < 8; j++) {
          const int t = 8 * g + j;
          if (t < kTaps) {
            int dx, dy, dz;
            tap_off(t, dx, dy, dz);
            a[j] = sh[(row + 1 + dx) + kHX * ((wave + 1 + dy) + kHY * (q + 1 + dz))];
          }
        }
      } else {
        int off = tap_lds<CIN>(TPC * c);
#pragma unroll
        for (int k = 1; k < TPC; k++)
          if (gsel == k) off = tap_lds<CIN>(TPC * c + k);
        if (TPC * c + gsel < kTaps)
          a = *reinterpret_cast<const u16x8 *>(sh + abase + q * (kHX * kHY * CIN) + off);
      }
      acc[q] = __builtin_amdgcn_mfma_f32_16x16x32_bf16(__builtin_bit_cast(bf16x8, a), __builtin_bit_cast(bf16x8, bw[c]),
                                                       acc[q], 0, 0, 0);
    }
  }
  // C fragment: row 4 g + i (node X0 + 4 g + i), column l & 15 (output channel)
  const int co = l & 15;
  const float bv = (MODE != 2 && co < COUT) ? bias[co] : 0.f;
  const int y = Y0 + wave;
  if (MODE == 1 || MODE == 3) {
    if (co != 0) return;
    const float sc = MODE == 3 ? (float)(nrm[0] / sq) : 1.f;
#pragma unroll
    for (int q = 0; q < kBZ; q++) {
      const int z = Z0 + q;
      if (y >= M || z >= M) continue;
#pragma unroll
      for (int i = 0; i < 4; i++) {
        const int xn = X0 + 4 * g + i;
        if (xn >= M) continue;
        const float v = acc[q][i] + bv;
        float o = v;
        if (MODE == 3) {
          const bool interior = xn > 0 && y > 0 && z > 0 && xn < M - 1 && y < M - 1 && z < M - 1;
          o = interior ? v * sc : 0.f;
        }
        out_f[xn + (int64_t)M * (y + (int64_t)M * z)] = o;
      }
    }
    return;
  }
  // BF16 outputs (16 channels): transpose through LDS so that every lane
  // stores 16 contiguous bytes (8 channels of one node)
  __shared__ __attribute__((aligned(16))) uint16_t so[4][kBZ * 16 * 16];
  uint16_t *sw = so[wave];
#pragma unroll
  for (int q = 0; q < kBZ; q++) {
    const int z = Z0 + q;
#pragma unroll
    for (int i = 0; i < 4; i++) {
      const int r = 4 * g + i;
      const float v = acc[q][i] + bv;
      if (MODE == 0) {
        const int xn = X0 + r;
        if (pre && xn < M && y < M && z < M) pre[(xn + (int64_t)M * (y + (int64_t)M * z)) * COUT + co] = v;
        sw[(q * 16 + r) * 16 + co] = f2bf(gelu(v));
      } else {
        sw[(q * 16 + r) * 16 + co] = f2bf(v);
      }
    }
  }
  __builtin_amdgcn_wave_barrier();
  __builtin_amdgcn_fence(__ATOMIC_RELEASE, "wavefront");
  __builtin_amdgcn_fence(__ATOMIC_ACQUIRE, "wavefront");
#pragma unroll
  for (int k = 0; k < kBZ * 16 * 16 / 512; k++) {
    const int e = k * 512 + 8 * l;
    const int q = e >> 8, r = (e >> 4) & 15, c0 = e & 15;
    const int xn = X0 + r, z = Z0 + q;
    if (xn < M && y < M && z < M)
      *reinterpret_cast<u16x8 *>(out_bf + (xn + (int64_t)M * (y + (int64_t)M * z)) * 16 + c0) =
          *reinterpret_cast<const u16x8 *>(sw + e);
  }
}

// FP32 -> BF16 input of the first layer
__global__ void k_to_bf16(const float *__restrict__ x, uint16_t *__restrict__ y, int64_t n) {
  const int64_t i = blockIdx.x * (int64_t)blockDim.x + threadIdx.x;
  if (i < n) y[i] = f2bf(x[i]);
}
// master FP32 parameters W[cout][27][cin] -> BF16 weights wb[16][27][cin]
// (outputs padded to 16 with zeros) and the transposed, tap-flipped weights
// wt[16][27][cout] of the input-gradient convolution: wt[c][t][o] = W[o][26-t][c]
// (tap 26-t is the opposite offset of tap t), rows c >= cin zero.
// per layer: offset of W in the parameter vector, cin, cout
struct LayerShapes {
  int off[kNetLayers], cin[kNetLayers], cout[kNetLayers];
};
// BF16 copies of every layer's weights (blockIdx.y = layer): wb = W padded to
// 16 outputs [16][27][cin], wt = transposed, tap-flipped [16][27][cout]
__global__ void k_make_bf16_weights(const float *__restrict__ params, LayerShapes ls, uint16_t *__restrict__ wbf,
                                    uint16_t *__restrict__ wtbf, int wstride) {
  const int l = blockIdx.y;
  const float *__restrict__ W = params + ls.off[l];
  const int cin = ls.cin[l], cout = ls.cout[l];
  uint16_t *__restrict__ wb = wbf + (int64_t)l * wstride;
  uint16_t *__restrict__ wt = wtbf + (int64_t)l * wstride;
  const int i = blockIdx.x * blockDim.x + threadIdx.x;
  if (i < 16 * kTaps * cin) {
    const int c = i % cin, t = (i / cin) % kTaps, o = i / (cin * kTaps);
    wb[i] = (o < cout) ? f2bf(W[(o * kTaps + t) * cin + c]) : (uint16_t)0;
  }
  if (i < 16 * kTaps * cout) {
    const int o = i % cout, t = (i / cout) % kTaps, c = i / (cout * kTaps);
    wt[i] = (c < cin) ? f2bf(W[(o * kTaps + (kTaps - 1 - t)) * cin + c]) : (uint16_t)0;
  }
}

// ------------------------------------------------------------------ backward
// loss = sum over interior nodes of (y - t)^2 / count;  gy = 2 (y - t) / count
__global__ __launch_bounds__(256) void k_loss(int M, const float *__restrict__ y, const float *__restrict__ t,
                                              float inv_count, uint16_t *__restrict__ gy, double *__restrict__ part) {
  const int64_t N = (int64_t)M * M * M;
  double s = 0;
  for (int64_t i = blockIdx.x * (int64_t)blockDim.x + threadIdx.x; i < N; i += (int64_t)gridDim.x * blockDim.x) {
    const int x = (int)(i % M), yy = (int)((i / M) % M), z = (int)(i / ((int64_t)M * M));
    const bool interior = x > 0 && yy > 0 && z > 0 && x < M - 1 && yy < M - 1 && z < M - 1;
    const float d = interior ? y[i] - t[i] : 0.f;
    s += (double)d * d;
    const float g = 2.f * d * inv_count;
    gy[i] = f2bf(g);
  }
  __shared__ double sh[4];
  for (int o = 32; o > 0; o >>= 1) s += __shfl_down(s, o, 64);
  if ((threadIdx.x & 63) == 0) sh[threadIdx.x >> 6] = s;
  __syncthreads();
  if (threadIdx.x == 0) part[blockIdx.x] = (sh[0] + sh[1] + sh[2] + sh[3]) * inv_count;
}
// gradient through GELU: gpre = gpost * gelu'(pre) (BF16 out, [N][16])
__global__ void k_gelu_back(const uint16_t *__restrict__ gpost, const float *__restrict__ pre,
                            uint16_t *__restrict__ gpre, int64_t n) {
  const int64_t i = blockIdx.x * (int64_t)blockDim.x + threadIdx.x;
  if (i >= n) return;
  gpre[i] = f2bf(bf2f(gpost[i]) * gelu_d(pre[i]));
}
// o = sum(p) in a fixed order (one workgroup of kBlock lanes)
__global__ __launch_bounds__(kBlock) void k_sum_parts(const double *__restrict__ p, int n, float *__restrict__ o) {
  double s = 0;
  for (int i = threadIdx.x; i < n; i += kBlock) s += p[i];
  s = block_sum<double>(s);
  if (threadIdx.x == 0) o[0] = (float)s;
}
// Weight and bias gradients on MFMA: for every tap t,
//   dW[o][t][c] = sum over nodes n of gp[n][o] a[n + off_t][c]
// is a 16 x 16 product with the nodes as the K dimension (32 per instruction).
// A persistent workgroup walks over 16 x 4 x 4-node bricks; per brick it
// stages gp channel-major ([o][z][y][x]) and the input a with its halo as
// three copies shifted by dx = -1, 0, 1 ([dx][c][z][y][x]), so that every A
// and B fragment is one aligned 16-byte LDS read.  Wavefront w owns the taps
// t = w (mod 4); the bias gradient is the extra "tap" 27 with B = 1.
// gp: BF16 [N][COUT]; a: BF16 [N][CIN]; part[block][e] as in k_reduce_parts.
template <int CIN, int COUT>
__global__ __launch_bounds__(256) void k_wgrad_mfma(int M, const uint16_t *__restrict__ gp,
                                                    const uint16_t *__restrict__ a, float *__restrict__ part) {
  constexpr int NE = COUT * kTaps * CIN + COUT;
  constexpr int NT = 7;  // taps per wavefront (27 taps + bias over 4 wavefronts)
  __shared__ __attribute__((aligned(16))) uint16_t sa[3 * 16 * kHZ * kHY * kBX];
  __shared__ __attribute__((aligned(16))) uint16_t sg[16 * kBZ * kBY * kBX];
  const int tid = threadIdx.x, wave = tid >> 6, l = tid & 63, row = l & 15, g = l >> 4;
  for (int i = tid; i < 3 * 16 * kHZ * kHY * kBX; i += 256) sa[i] = 0;
  for (int i = tid; i < 16 * kBZ * kBY * kBX; i += 256) sg[i] = 0;
  floatx4 acc[NT];
#pragma unroll
  for (int e = 0; e < NT; e++) acc[e] = floatx4{0.f, 0.f, 0.f, 0.f};
  const int nbx = (M + kBX - 1) / kBX, nby = (M + kBY - 1) / kBY, nbz = (M + kBZ - 1) / kBZ;
  const int nbricks = nbx * nby * nbz;
  for (int b = blockIdx.x; b < nbricks; b += gridDim.x) {
    const int X0 = (b % nbx) * kBX, Y0 = ((b / nbx) % nby) * kBY, Z0 = (b / (nbx * nby)) * kBZ;
    __syncthreads();
    // input brick with halo, three x-shifted copies
    for (int q = tid; q < kHalo * (CIN / 8); q += 256) {
      const int n = q / (CIN / 8), cb = 8 * (q % (CIN / 8));
      const int hx = n % kHX, hy = (n / kHX) % kHY, hz = n / (kHX * kHY);
      const int x = X0 - 1 + hx, y = Y0 - 1 + hy, z = Z0 - 1 + hz;
      u16x8 v = {0, 0, 0, 0, 0, 0, 0, 0};
      if (x >= 0 && y >= 0 && z >= 0 && x < M && y < M && z < M)
        v = *reinterpret_cast<const u16x8 *>(a + (x + (int64_t)M * (y + (int64_t)M * z)) * CIN + cb);
#pragma unroll
      for (int d = 0; d < 3; d++) {
        const int xx = hx - d;  // copy d holds a(X0 + xx + d - 1, ...)
        if (xx < 0 || xx >= kBX) continue;
#pragma unroll
        for (int j = 0; j < 8; j++) sa[(((d * 16 + cb + j) * kHZ + hz) * kHY + hy) * kBX + xx] = v[j];
      }
    }
    // output gradient of the brick, channel-major
    if (COUT == 16) {
      for (int q = tid; q < kBX * kBY * kBZ * 2; q += 256) {
        const int n = q >> 1, ob = 8 * (q & 1);
        const int xx = n % kBX, yy = (n / kBX) % kBY, zz = n / (kBX * kBY);
        const int x = X0 + xx, y = Y0 + yy, z = Z0 + zz;
        u16x8 v = {0, 0, 0, 0, 0, 0, 0, 0};
        if (x < M && y < M && z < M)
          v = *reinterpret_cast<const u16x8 *>(gp + (x + (int64_t)M * (y + (int64_t)M * z)) * 16 + ob);
#pragma unroll
        for (int j = 0; j < 8; j++) sg[(((ob + j) * kBZ + zz) * kBY + yy) * kBX + xx] = v[j];
      }
    } else {
      for (int n = tid; n < kBX * kBY * kBZ; n += 256) {
        const int xx = n % kBX, yy = (n / kBX) % kBY, zz = n / (kBX * kBY);
        const int x = X0 + xx, y = Y0 + yy, z = Z0 + zz;
        sg[(zz * kBY + yy) * kBX + xx] =
            (x < M && y < M && z < M) ? gp[x + (int64_t)M * (y + (int64_t)M * z)] : (uint16_t)0;
      }
    }
    __syncthreads();
    // K chunks of 32 nodes: chunk ch = two x-lines (z = ch / 2, y = 2 (ch % 2) + g / 2), 8 nodes per lane
#pragma unroll 2
    for (int ch = 0; ch < (kBY * kBZ) / 2; ch++) {
      const int zz = ch >> 1, yy = 2 * (ch & 1) + (g >> 1), x0 = 8 * (g & 1);
      const u16x8 af = *reinterpret_cast<const u16x8 *>(sg + ((row * kBZ + zz) * kBY + yy) * kBX + x0);
#pragma unroll
      for (int e = 0; e < NT; e++) {
        const int t = wave + 4 * e;
        if (t > kTaps) continue;
        u16x8 bf;
        if (t == kTaps) {
          bf = u16x8{0x3F80, 0x3F80, 0x3F80, 0x3F80, 0x3F80, 0x3F80, 0x3F80, 0x3F80};  // 1.0
        } else {
          int dx, dy, dz;
          tap_off(t, dx, dy, dz);
          bf = *reinterpret_cast<const u16x8 *>(
              sa + ((((dx + 1) * 16 + row) * kHZ + (zz + 1 + dz)) * kHY + (yy + 1 + dy)) * kBX + x0);
        }
        acc[e] = __builtin_amdgcn_mfma_f32_16x16x32_bf16(__builtin_bit_cast(bf16x8, af), __builtin_bit_cast(bf16x8, bf),
                                                         acc[e], 0, 0, 0);
      }
    }
  }
  // C fragment: row o = 4 g + i, column c = l & 15
  float *pb = part + (int64_t)blockIdx.x * NE;
  const int c = l & 15;
#pragma unroll
  for (int e = 0; e < NT; e++) {
    const int t = wave + 4 * e;
    if (t > kTaps) continue;
#pragma unroll
    for (int i = 0; i < 4; i++) {
      const int o = 4 * g + i;
      if (o >= COUT) continue;
      if (t < kTaps) {
        if (c < CIN) pb[(o * kTaps + t) * CIN + c] = acc[e][i];
      } else if (c == 0) {
        pb[COUT * kTaps * CIN + o] = acc[e][i];
      }
    }
  }
}

// grads[e] += sum_b part[b][e] (fixed order)
constexpr int kRedChunk = 32;  // weight-gradient partials summed per first-pass thread
constexpr int kRedChunks = (512 + kRedChunk - 1) / kRedChunk;  // for the 512 wgrad workgroups
// grads[e] += sum over blocks b of part[b][e], in two fixed-order passes:
// chunk c of kRedChunk blocks into tmp[c][e], then the chunks
__global__ void k_reduce_parts(const float *__restrict__ part, int nb, int ne, float *__restrict__ tmp) {
  const int e = blockIdx.x * blockDim.x + threadIdx.x;
  if (e >= ne) return;
  const int b0 = blockIdx.y * kRedChunk, b1 = min(nb, b0 + kRedChunk);
  float s = 0.f;
#pragma unroll 8
  for (int b = b0; b < b1; b++) s += part[(int64_t)b * ne + e];
  tmp[(int64_t)blockIdx.y * ne + e] = s;
}
__global__ void k_reduce_chunks(const float *__restrict__ tmp, int nc, int ne, float *__restrict__ grads) {
  const int e = blockIdx.x * blockDim.x + threadIdx.x;
  if (e >= ne) return;
  float s = 0.f;
  for (int c = 0; c < nc; c++) s += tmp[(int64_t)c * ne + e];
  grads[e] += s;
}
__global__ void k_adam(float *__restrict__ p, const float *__restrict__ g, float *__restrict__ m,
                       float *__restrict__ v, int n, float lr, float b1, float b2, float eps, float c1, float c2,
                       float gscale) {
  const int i = blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= n) return;
  const float gi = g[i] * gscale;
  const float mi = b1 * m[i] + (1.f - b1) * gi;
  const float vi = b2 * v[i] + (1.f - b2) * gi * gi;
  m[i] = mi;
  v[i] = vi;
  p[i] -= lr * (mi / c1) / (sqrtf(vi / c2) + eps);
}
// deterministic initialisation: uniform(-a, a) with a = sqrt(6 / fan_in) from a counter hash
__global__ void k_init_params(float *__restrict__ p, int n, int off_w, int nw, float a, uint32_t seed) {
  const int i = blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= n) return;
  if (i >= off_w && i < off_w + nw) {
    uint32_t h = (uint32_t)i * 747796405u + seed * 2891336453u + 1442695041u;
    h ^= h >> 16; h *= 0x7feb352dU; h ^= h >> 15; h *= 0x846ca68bU; h ^= h >> 16;
    p[i] = a * (2.0f * ((h >> 8) * (1.0f / 16777216.0f)) - 1.0f);
  } else if (i >= off_w + nw && i < n) {
    p[i] = 0.f;
  }
}

// ------------------------------------------------------------------ host side
void LearnedNet::release() {
  auto fr = [](void *&q) {
    if (q) (void)hipFree(q);
    q = nullptr;
  };
  fr((void *&)params); fr((void *&)grads); fr((void *&)adam_m); fr((void *&)adam_v);
  fr((void *&)wbf); fr((void *&)wtbf); fr((void *&)part); fr((void *&)out);
  for (int l = 0; l < kNetLayers; l++) { fr((void *&)act[l]); fr((void *&)pre[l]); }
  fr((void *&)gbuf[0]); fr((void *&)gbuf[1]);
  fr((void *&)lpart);
  cap = 0;
  part_cap = 0;
  ready = false;
}
void LearnedData::release() {
  if (u) (void)hipFree(u);
  if (t) (void)hipFree(t);
  u = t = nullptr;
  ns = 0;
}

static const int kWStride = 16 * kTaps * 16;  // BF16 weight slots per layer (padded)
static const int kWgradBlocks = 512;
static_assert(kRedChunks * kRedChunk >= kWgradBlocks, "first reduction pass must cover the wgrad partials");
static const int kLossBlocks = 1024;

void net_init(LearnedNet &net, uint32_t seed, hipStream_t st) {
  const int P = net_total_params();
  if (!net.params) {
    check(hipMalloc(&net.params, sizeof(float) * P), "net params");
    check(hipMalloc(&net.grads, sizeof(float) * P), "net grads");
    check(hipMalloc(&net.adam_m, sizeof(float) * P), "adam m");
    check(hipMalloc(&net.adam_v, sizeof(float) * P), "adam v");
    check(hipMalloc(&net.wbf, sizeof(uint16_t) * kWStride * kNetLayers), "net wbf");
    check(hipMalloc(&net.wtbf, sizeof(uint16_t) * kWStride * kNetLayers), "net wtbf");
  }
  for (int l = 0; l < kNetLayers; l++) {
    const int off = net_param_offset(l);
    const float a = std::sqrt(6.0f / (kTaps * net_cin(l))) * (l == kNetLayers - 1 ? 0.1f : 1.0f);
    hipLaunchKernelGGL(k_init_params, dim3((net_psize(l) + 255) / 256), dim3(256), 0, st, net.params + off,
                       net_psize(l), 0, net_wsize(l), a, seed + 977u * (uint32_t)l);
  }
  fill<float>(net.adam_m, 0.f, P, st);
  fill<float>(net.adam_v, 0.f, P, st);
  fill<float>(net.grads, 0.f, P, st);
  net.adam_t = 0;
  net_sync_bf16(net, st);
  net.ready = true;
}

void net_sync_bf16(LearnedNet &net, hipStream_t st) {
  LayerShapes ls;
  for (int l = 0; l < kNetLayers; l++) {
    ls.off[l] = net_param_offset(l);
    ls.cin[l] = net_cin(l);
    ls.cout[l] = net_cout(l);
  }
  hipLaunchKernelGGL(k_make_bf16_weights, dim3((kWStride + 255) / 256, kNetLayers), dim3(256), 0, st, net.params, ls,
                     net.wbf, net.wtbf, kWStride);
}

bool net_reserve(LearnedNet &net, int64_t N, bool training) {
  bool changed = false;
  if (N > net.cap) {
    changed = true;
    for (int l = 0; l < kNetLayers; l++) {
      if (net.act[l]) (void)hipFree(net.act[l]);
      if (net.pre[l]) (void)hipFree(net.pre[l]);
      net.act[l] = nullptr;
      net.pre[l] = nullptr;
      check(hipMalloc(&net.act[l], sizeof(uint16_t) * N * net_cin(l)), "net act");
    }
    if (net.out) (void)hipFree(net.out);
    check(hipMalloc(&net.out, sizeof(float) * N), "net out");
    for (auto &g : net.gbuf) {
      if (g) (void)hipFree(g);
      g = nullptr;
    }
    net.cap = N;
  }
  if (training && !net.pre[0]) {
    changed = true;
    for (int l = 0; l < kNetLayers; l++) check(hipMalloc(&net.pre[l], sizeof(float) * net.cap * kNetCh), "net pre");
    for (auto &g : net.gbuf) check(hipMalloc(&g, sizeof(uint16_t) * net.cap * kNetCh), "net grad buf");
    if (!net.lpart) check(hipMalloc(&net.lpart, sizeof(double) * kLossBlocks), "net loss partials");
  }
  // partials of every workgroup, then the first-pass sums
  const int64_t pc = (int64_t)(kWgradBlocks + kRedChunks) * (16 * kTaps * 16 + 16);
  if (training && net.part_cap < pc) {
    if (net.part) (void)hipFree(net.part);
    check(hipMalloc(&net.part, sizeof(float) * pc), "net part");
    net.part_cap = pc;
  }
  return changed;
}

static unsigned conv_grid(int M) {
  return (unsigned)(((M + kBX - 1) / kBX) * ((M + kBY - 1) / kBY) * ((M + kBZ - 1) / kBZ));
}

void net_forward(LearnedNet &net, int M, const float *in, bool training, hipStream_t st) {
  const int64_t N = (int64_t)M * M * M;
  if (N > net.cap) throw std::runtime_error("net_forward: lattice larger than reserved");
  hipLaunchKernelGGL(k_to_bf16, dim3((unsigned)((N * kNetIn + 255) / 256)), dim3(256), 0, st, in, net.act[0],
                     N * kNetIn);
  net_forward_act(net, M, training, st);
}

void net_forward_act(LearnedNet &net, int M, bool training, hipStream_t st) {
  const unsigned g = conv_grid(M);
  auto P = [&](int l) { return net.params + net_param_offset(l) + net_wsize(l); };  // bias
  hipLaunchKernelGGL((k_conv<kNetIn, 16, 0>), dim3(g), dim3(256), 0, st, M, net.act[0], net.wbf + 0 * kWStride, P(0),
                     training ? net.pre[0] : nullptr, net.act[1], nullptr);
  hipLaunchKernelGGL((k_conv<16, 16, 0>), dim3(g), dim3(256), 0, st, M, net.act[1], net.wbf + 1 * kWStride, P(1),
                     training ? net.pre[1] : nullptr, net.act[2], nullptr);
  hipLaunchKernelGGL((k_conv<16, 16, 0>), dim3(g), dim3(256), 0, st, M, net.act[2], net.wbf + 2 * kWStride, P(2),
                     training ? net.pre[2] : nullptr, net.act[3], nullptr);
  hipLaunchKernelGGL((k_conv<16, 1, 1>), dim3(g), dim3(256), 0, st, M, net.act[3], net.wbf + 3 * kWStride, P(3),
                     nullptr, nullptr, net.out);
}

void net_forward_correction(LearnedNet &net, int M, double *nrm, double sqrt_count, float *c, hipStream_t st,
                            const LearnedInput *fused) {
  const unsigned g = conv_grid(M);
  auto P = [&](int l) { return net.params + net_param_offset(l) + net_wsize(l); };  // bias
  if (fused) {
    // the first layer builds its input (features + scaled defect) itself
    hipLaunchKernelGGL((k_conv<kNetIn, 16, 0, true>), dim3(g), dim3(256), 0, st, M, fused->feat,
                       net.wbf + 0 * kWStride, P(0), nullptr, net.act[1], nullptr, nullptr, sqrt_count, fused->u,
                       fused->part, fused->np, nrm);
  } else {
    hipLaunchKernelGGL((k_conv<kNetIn, 16, 0>), dim3(g), dim3(256), 0, st, M, net.act[0], net.wbf + 0 * kWStride,
                       P(0), nullptr, net.act[1], nullptr, nullptr, 1.0);
  }
  hipLaunchKernelGGL((k_conv<16, 16, 0>), dim3(g), dim3(256), 0, st, M, net.act[1], net.wbf + 1 * kWStride, P(1),
                     nullptr, net.act[2], nullptr, nullptr, 1.0);
  hipLaunchKernelGGL((k_conv<16, 16, 0>), dim3(g), dim3(256), 0, st, M, net.act[2], net.wbf + 2 * kWStride, P(2),
                     nullptr, net.act[3], nullptr, nullptr, 1.0);
  hipLaunchKernelGGL((k_conv<16, 1, 3>), dim3(g), dim3(256), 0, st, M, net.act[3], net.wbf + 3 * kWStride, P(3),
                     nullptr, nullptr, c, nrm, sqrt_count);
}

void net_zero_grads(LearnedNet &net, hipStream_t st) { fill<float>(net.grads, 0.f, net_total_params(), st); }

template <int CIN, int COUT>
static void wgrad(LearnedNet &net, int M, const uint16_t *gp, const uint16_t *a, int layer, hipStream_t st) {
  constexpr int NE = COUT * kTaps * CIN + COUT;
  hipLaunchKernelGGL((k_wgrad_mfma<CIN, COUT>), dim3(kWgradBlocks), dim3(256), 0, st, M, gp, a, net.part);
  float *tmp = net.part + (int64_t)kWgradBlocks * NE;
  hipLaunchKernelGGL(k_reduce_parts, dim3((NE + 255) / 256, kRedChunks), dim3(256), 0, st, net.part, kWgradBlocks,
                     NE, tmp);
  hipLaunchKernelGGL(k_reduce_chunks, dim3((NE + 255) / 256), dim3(256), 0, st, tmp, kRedChunks, NE,
                     net.grads + net_param_offset(layer));
}

static void backward_layers(LearnedNet &net, int M, hipStream_t st);

void net_backward(LearnedNet &net, int M, const float *t, float *loss_dev, hipStream_t st) {
  const int64_t N = (int64_t)M * M * M;
  const double inner = (double)(M - 2) * (M - 2) * (M - 2);
  double *lpart = net.lpart;
  uint16_t *g0 = net.gbuf[0];
  // output layer: gy (BF16 in g0[:N])
  hipLaunchKernelGGL(k_loss, dim3(kLossBlocks), dim3(256), 0, st, M, net.out, t, (float)(1.0 / inner), g0, lpart);
  backward_layers(net, M, st);
  if (loss_dev) hipLaunchKernelGGL(k_sum_parts, dim3(1), dim3(kBlock), 0, st, lpart, kLossBlocks, loss_dev);
}

__global__ void k_split_grad(const float *__restrict__ gy, uint16_t *__restrict__ gb, int64_t n) {
  for (int64_t i = blockIdx.x * (int64_t)blockDim.x + threadIdx.x; i < n; i += (int64_t)gridDim.x * blockDim.x)
    gb[i] = f2bf(gy[i]);
}

void net_backward_from(LearnedNet &net, int M, const float *gy, hipStream_t st) {
  const int64_t N = (int64_t)M * M * M;
  hipLaunchKernelGGL(k_split_grad, dim3((unsigned)std::min<int64_t>((N + 255) / 256, 4096)), dim3(256), 0, st, gy,
                     net.gbuf[0], N);
  backward_layers(net, M, st);
}

// the layers' backward pass, from gy in gbuf[0] (BF16)
static void backward_layers(LearnedNet &net, int M, hipStream_t st) {
  const int64_t N = (int64_t)M * M * M;
  uint16_t *g0 = net.gbuf[0], *g1 = net.gbuf[1];
  wgrad<16, 1>(net, M, g0, net.act[3], 3, st);
  const unsigned g = conv_grid(M);
  // input gradient of layer 3's output: conv of gy with transposed weights (1 -> 16)
  hipLaunchKernelGGL((k_conv<1, 16, 2>), dim3(g), dim3(256), 0, st, M, g0, net.wtbf + 3 * kWStride, nullptr, nullptr,
                     g1, nullptr);
  for (int l = 2; l >= 0; l--) {
    // through GELU of layer l
    hipLaunchKernelGGL(k_gelu_back, dim3((unsigned)((N * kNetCh + 255) / 256)), dim3(256), 0, st, g1, net.pre[l], g0,
                       N * kNetCh);
    if (l == 0) wgrad<kNetIn, 16>(net, M, g0, net.act[0], 0, st);
    else wgrad<16, 16>(net, M, g0, net.act[l], l, st);
    if (l > 0)
      hipLaunchKernelGGL((k_conv<16, 16, 2>), dim3(g), dim3(256), 0, st, M, g0, net.wtbf + l * kWStride, nullptr,
                         nullptr, g1, nullptr);
  }
}

void net_adam(LearnedNet &net, float lr, hipStream_t st, float gscale) {
  net.adam_t++;
  const float b1 = 0.9f, b2 = 0.999f;
  const float c1 = 1.f - std::pow(b1, (float)net.adam_t), c2 = 1.f - std::pow(b2, (float)net.adam_t);
  const int P = net_total_params();
  hipLaunchKernelGGL(k_adam, dim3((P + 255) / 256), dim3(256), 0, st, net.params, net.grads, net.adam_m, net.adam_v, P,
                     lr, b1, b2, 1e-8f, c1, c2, gscale);
  net_sync_bf16(net, st);
}

// ------------------------------------------------------------------ use as a smoother
__global__ void k_diag_inv(const BlockInfo *__restrict__ blocks, const DevSegment *__restrict__ segs, int M,
                           const float *__restrict__ val, float *__restrict__ dinv) {
  const BlockInfo bi = blocks[blockIdx.x];
  const DevSegment s = segs[bi.seg];
  const int idx = bi.base + threadIdx.x;
  if (idx >= s.count) return;
  const int i = idx % s.cx, t = idx / s.cx, j = t % s.cy, k = t / s.cy;
  const int p = (s.x0 + s.stride * i) + M * ((s.y0 + s.stride * j) + M * (s.z0 + s.stride * k));
  const int64_t vb = s.voff + (int64_t)(idx / kSlice) * ((int64_t)s.ns * kSlice) + (idx % kSlice);
  dinv[p] = 1.0f / val[vb + (int64_t)s.dslot * kSlice];
}
__global__ void k_learned_input(const float *__restrict__ dinv, const float *__restrict__ d, float *__restrict__ u,
                                int64_t n) {
  for (int64_t i = blockIdx.x * (int64_t)blockDim.x + threadIdx.x; i < n; i += (int64_t)gridDim.x * blockDim.x)
    u[i] = dinv[i] * d[i];
}
// u = dinv d together with the partial sums of u^2 (the grid and summation
// order of k_dot_partial)
__global__ __launch_bounds__(kBlock) void k_learned_input_sq(const float *__restrict__ dinv,
                                                             const float *__restrict__ d, float *__restrict__ u,
                                                             double *__restrict__ part, int64_t n) {
  double s = 0;
  for (int64_t i = blockIdx.x * (int64_t)kBlock + threadIdx.x; i < n; i += (int64_t)gridDim.x * kBlock) {
    const float v = dinv[i] * d[i];
    u[i] = v;
    s += (double)v * (double)v;
  }
  s = block_sum<float>(s);
  if (threadIdx.x == 0) part[blockIdx.x] = s;
}
// partial sums of (a, b) and (b, b) in one pass (k_dot_partial's order)
__global__ __launch_bounds__(kBlock) void k_dot2_partial(const float *__restrict__ a, const float *__restrict__ b,
                                                         double *__restrict__ pab, double *__restrict__ pbb,
                                                         int64_t n) {
  double s = 0, t = 0;
  for (int64_t i = blockIdx.x * (int64_t)kBlock + threadIdx.x; i < n; i += (int64_t)gridDim.x * kBlock) {
    const double bv = (double)b[i];
    s += (double)a[i] * bv;
    t += bv * bv;
  }
  s = block_sum<float>(s);
  t = block_sum<float>(t);
  if (threadIdx.x == 0) {
    pab[blockIdx.x] = s;
    pbb[blockIdx.x] = t;
  }
}
// act0[p] = feat[p] with channel 0 = bf16(u[p] * sq / *nrm) (no scaling if nrm is null)
__global__ void k_build_input(const float *__restrict__ u, const double *__restrict__ nrm, double sq,
                              const uint16_t *__restrict__ feat, int64_t n, uint16_t *__restrict__ act0) {
  float sc = 1.0f;
  if (nrm) {
    const double nv = nrm[0];
    sc = nv > 0 ? (float)(sq / nv) : 0.f;
  }
  for (int64_t i = blockIdx.x * (int64_t)blockDim.x + threadIdx.x; i < n; i += (int64_t)gridDim.x * blockDim.x) {
#pragma unroll
    for (int piece = 0; piece < kNetIn / 8; piece++) {
      u16x8 v = *reinterpret_cast<const u16x8 *>(feat + i * kNetIn + 8 * piece);
      if (piece == 0) v[0] = f2bf(u[i] * sc);
      *reinterpret_cast<u16x8 *>(act0 + i * kNetIn + 8 * piece) = v;
    }
  }
}
// the six nearest couplings of every interior row, divided by its diagonal
__global__ void k_features(const BlockInfo *__restrict__ blocks, const DevSegment *__restrict__ segs, int M,
                           const float *__restrict__ val, uint16_t *__restrict__ feat) {
  const BlockInfo bi = blocks[blockIdx.x];
  const DevSegment s = segs[bi.seg];
  const int idx = bi.base + threadIdx.x;
  if (idx >= s.count) return;
  const int i = idx % s.cx, t = idx / s.cx, j = t % s.cy, k = t / s.cy;
  const int64_t p = (s.x0 + s.stride * i) + (int64_t)M * ((s.y0 + s.stride * j) + (int64_t)M * (s.z0 + s.stride * k));
  const int64_t vb = s.voff + (int64_t)(idx / kSlice) * ((int64_t)s.ns * kSlice) + (idx % kSlice);
  const int wx = 2 * s.ex + 1, wy = 2 * s.ey + 1;
  auto slot = [&](int dx, int dy, int dz) { return (dx + s.ex) + wx * ((dy + s.ey) + wy * (dz + s.ez)); };
  const float dinv = 1.0f / val[vb + (int64_t)s.dslot * kSlice];
  const int off[6][3] = {{1, 0, 0}, {-1, 0, 0}, {0, 1, 0}, {0, -1, 0}, {0, 0, 1}, {0, 0, -1}};
  const int ext[3] = {s.ex, s.ey, s.ez};
  uint16_t f[16] = {0};
#pragma unroll
  for (int q = 0; q < 6; q++) {
    f[1 + q] = f2bf(val[vb + (int64_t)slot(off[q][0], off[q][1], off[q][2]) * kSlice] * dinv);
    const int axis = q >> 1;
    if (ext[axis] >= 2)
      f[7 + q] = f2bf(val[vb + (int64_t)slot(2 * off[q][0], 2 * off[q][1], 2 * off[q][2]) * kSlice] * dinv);
  }
  u16x8 v0, v1;
#pragma unroll
  for (int j = 0; j < 8; j++) { v0[j] = f[j]; v1[j] = f[8 + j]; }
  *reinterpret_cast<u16x8 *>(feat + p * kNetIn) = v0;
  *reinterpret_cast<u16x8 *>(feat + p * kNetIn + 8) = v1;
}

// c = sign * in * sq / nrm (interior only)
__global__ void k_interior_scale(const float *__restrict__ in, int M, const double *__restrict__ nrm, double sq,
                                 float sign, float *__restrict__ c) {
  const int64_t N = (int64_t)M * M * M;
  const double nv = nrm[0];
  const float sc = nv > 0 ? (float)(sign * sq / nv) : 0.f;
  for (int64_t i = blockIdx.x * (int64_t)blockDim.x + threadIdx.x; i < N; i += (int64_t)gridDim.x * blockDim.x) {
    const int x = (int)(i % M), y = (int)((i / M) % M), z = (int)(i / ((int64_t)M * M));
    const bool interior = x > 0 && y > 0 && z > 0 && x < M - 1 && y < M - 1 && z < M - 1;
    c[i] = interior ? in[i] * sc : 0.f;
  }
}
// x += alpha c, alpha = max(0, (d, Ac) / (Ac, Ac)) reduced by every workgroup
// from the partials of k_dot2_partial in the fixed order of k_reduce_final
__global__ __launch_bounds__(kBlock) void k_learned_step_parts(float *__restrict__ x, const float *__restrict__ c,
                                                               const double *__restrict__ pab,
                                                               const double *__restrict__ pbb, int np, int64_t n) {
  __shared__ float sh_alpha;
  double s = 0, t = 0;
  for (int i = threadIdx.x; i < np; i += kBlock) {
    s += pab[i];
    t += pbb[i];
  }
  s = block_sum<double>(s);
  t = block_sum<double>(t);
  if (threadIdx.x == 0) sh_alpha = (float)(t > 0 ? fmax(0.0, s / t) : 0.0);
  __syncthreads();
  const float af = sh_alpha;
  for (int64_t i = blockIdx.x * (int64_t)kBlock + threadIdx.x; i < n; i += (int64_t)gridDim.x * kBlock)
    x[i] += af * c[i];
}
// the step of k_learned_step_parts, and from the updated defect the next
// step's input: d = d_in - alpha Ac (d_in may be d), u = dinv d (u may be c
// itself: each thread reads c[i] before it writes u[i]) and the partials of
// ||u||^2 in k_learned_input_sq's order (same grid)
__global__ __launch_bounds__(kBlock) void k_learned_step_carry(float *__restrict__ x, const float *c,
                                                               const float *d_in, float *d,
                                                               const float *__restrict__ ac,
                                                               const float *__restrict__ dinv, float *u,
                                                               const double *__restrict__ pab,
                                                               const double *__restrict__ pbb, int np,
                                                               double *__restrict__ pu, int64_t n) {
  __shared__ float sh_alpha;
  double s = 0, t = 0;
  for (int i = threadIdx.x; i < np; i += kBlock) {
    s += pab[i];
    t += pbb[i];
  }
  s = block_sum<double>(s);
  t = block_sum<double>(t);
  if (threadIdx.x == 0) sh_alpha = (float)(t > 0 ? fmax(0.0, s / t) : 0.0);
  __syncthreads();
  const float af = sh_alpha;
  double q = 0;
  for (int64_t i = blockIdx.x * (int64_t)kBlock + threadIdx.x; i < n; i += (int64_t)gridDim.x * kBlock) {
    const float ci = c[i];
    x[i] += af * ci;
    const float dn = d_in[i] - af * ac[i];
    d[i] = dn;
    const float v = dinv[i] * dn;
    u[i] = v;
    q += (double)v * (double)v;
  }
  q = block_sum<float>(q);
  if (threadIdx.x == 0) pu[blockIdx.x] = q;
}
__global__ void k_learned_random(float *__restrict__ xv, int M, uint32_t seed) {
  const int64_t N = (int64_t)M * M * M;
  for (int64_t i = blockIdx.x * (int64_t)blockDim.x + threadIdx.x; i < N; i += (int64_t)gridDim.x * blockDim.x) {
    const int x = (int)(i % M), y = (int)((i / M) % M), z = (int)(i / ((int64_t)M * M));
    const bool interior = x > 0 && y > 0 && z > 0 && x < M - 1 && y < M - 1 && z < M - 1;
    uint32_t h = (uint32_t)i * 2654435761u ^ (seed * 0x9e3779b9u + 0x7f4a7c15u);
    h ^= h >> 16; h *= 0x7feb352dU; h ^= h >> 15; h *= 0x846ca68bU; h ^= h >> 16;
    xv[i] = interior ? 2.0f * ((h >> 8) * (1.0f / 16777216.0f)) - 1.0f : 0.f;
  }
}

static unsigned grid_for(int64_t n) { return (unsigned)std::min<int64_t>((n + 255) / 256, 4096); }

void learned_diag_inv(const DevLayout &D, const float *val, float *dinv, hipStream_t st) {
  fill<float>(dinv, 0.f, D.N, st);
  hipLaunchKernelGGL(k_diag_inv, dim3(D.nblocks_all), dim3(kRowsPerBlock), 0, st, D.blocks_all, D.segs, D.M, val,
                     dinv);
}
void learned_input(const float *dinv, const float *d, float *u, int64_t n, hipStream_t st) {
  hipLaunchKernelGGL(k_learned_input, dim3(grid_for(n)), dim3(256), 0, st, dinv, d, u, n);
}
static unsigned red_grid(int64_t n) {
  return (unsigned)std::max<int64_t>(1, std::min<int64_t>((n + kBlock - 1) / kBlock, kRedBlocks));
}
void learned_input_sq(const float *dinv, const float *d, float *u, int64_t n, double *part, hipStream_t st) {
  hipLaunchKernelGGL(k_learned_input_sq, dim3(red_grid(n)), dim3(kBlock), 0, st, dinv, d, u, part, n);
}
int learned_input_parts(int64_t n) { return (int)red_grid(n); }
void learned_update_carry(const float *d_in, float *d, const float *ac, const float *c, float *x, const float *dinv,
                          float *u, int64_t n, double *part, double *pu, hipStream_t st) {
  const unsigned g = red_grid(n);
  hipLaunchKernelGGL(k_dot2_partial, dim3(g), dim3(kBlock), 0, st, d_in, ac, part, part + kRedBlocks, n);
  hipLaunchKernelGGL(k_learned_step_carry, dim3(g), dim3(kBlock), 0, st, x, c, d_in, d, ac, dinv, u, part,
                     part + kRedBlocks, (int)g, pu, n);
}
void learned_update(const float *d, const float *ac, const float *c, float *x, int64_t n, double *part,
                    hipStream_t st) {
  const unsigned g = red_grid(n);
  hipLaunchKernelGGL(k_dot2_partial, dim3(g), dim3(kBlock), 0, st, d, ac, part, part + kRedBlocks, n);
  hipLaunchKernelGGL(k_learned_step_parts, dim3(grid_for(n)), dim3(kBlock), 0, st, x, c, part, part + kRedBlocks,
                     (int)g, n);
}
void learned_build_input(LearnedNet &net, const float *u, const double *nrm, double sq, const uint16_t *feat,
                         int64_t n, hipStream_t st) {
  hipLaunchKernelGGL(k_build_input, dim3(grid_for(n)), dim3(256), 0, st, u, nrm, sq, feat, n, net.act[0]);
}
void learned_features(const DevLayout &D, const float *val, uint16_t *feat, hipStream_t st) {
  fill<uint16_t>(feat, (uint16_t)0, D.N * kNetIn, st);
  hipLaunchKernelGGL(k_features, dim3(D.nblocks_all), dim3(kRowsPerBlock), 0, st, D.blocks_all, D.segs, D.M, val,
                     feat);
}
void learned_random(float *x, int M, uint32_t seed, hipStream_t st) {
  hipLaunchKernelGGL(k_learned_random, dim3(grid_for((int64_t)M * M * M)), dim3(256), 0, st, x, M, seed);
}
void learned_target(const float *x, int M, const double *nrm, double sq, float sign, float *t, hipStream_t st) {
  hipLaunchKernelGGL(k_interior_scale, dim3(grid_for((int64_t)M * M * M)), dim3(256), 0, st, x, M, nrm, sq, sign,
                     t);
}

// ------------------------------------------------------------------ residual loss
__global__ void k_mask(const float *__restrict__ in, int M, float *__restrict__ out) {
  const int64_t N = (int64_t)M * M * M;
  for (int64_t i = blockIdx.x * (int64_t)blockDim.x + threadIdx.x; i < N; i += (int64_t)gridDim.x * blockDim.x) {
    const int x = (int)(i % M), y = (int)((i / M) % M), z = (int)(i / ((int64_t)M * M));
    const bool interior = x > 0 && y > 0 && z > 0 && x < M - 1 && y < M - 1 && z < M - 1;
    out[i] = interior ? in[i] : 0.f;
  }
}
// r = u - dinv * Ac on interior nodes; g = -2 dinv r / count; part = sum r^2 / count
__global__ __launch_bounds__(256) void k_res_loss(int M, const float *__restrict__ u, const float *__restrict__ dinv,
                                                  const float *__restrict__ ac, float inv_count,
                                                  float *__restrict__ g, double *__restrict__ part) {
  const int64_t N = (int64_t)M * M * M;
  double s = 0;
  for (int64_t i = blockIdx.x * (int64_t)blockDim.x + threadIdx.x; i < N; i += (int64_t)gridDim.x * blockDim.x) {
    const int x = (int)(i % M), y = (int)((i / M) % M), z = (int)(i / ((int64_t)M * M));
    const bool interior = x > 0 && y > 0 && z > 0 && x < M - 1 && y < M - 1 && z < M - 1;
    const float r = interior ? u[i] - dinv[i] * ac[i] : 0.f;
    s += (double)r * r;
    g[i] = -2.f * inv_count * dinv[i] * r;
  }
  __shared__ double sh[4];
  for (int o = 32; o > 0; o >>= 1) s += __shfl_down(s, o, 64);
  if ((threadIdx.x & 63) == 0) sh[threadIdx.x >> 6] = s;
  __syncthreads();
  if (threadIdx.x == 0) part[blockIdx.x] = (sh[0] + sh[1] + sh[2] + sh[3]) * inv_count;
}
// y_j = sum over interior rows i of a_ij g_i, for interior j (0 on the boundary):
// The transpose of the stored stencil operator restricted to interior rows,
// stored in the same CSS layout: slot d of row j holds a_{j+d, j}, the entry
// at offset -d of the interior row j + d (0 where j + d is not an interior
// row or its stencil box lacks -d).  The FE pattern is structurally
// symmetric, so every such entry has its slot in row j's box.  A workgroup
// takes 64 rows j of one residue class, so the class of every neighbour row
// j + d, its stencil shape, the slot of -d and the shift of the
// class-lattice indices are the same across the workgroup (scalar work, S a
// compile-time constant); per lane only the neighbour's row index and the
// bounds remain.
template <int E, int S>
__global__ __launch_bounds__(kRowsPerBlock) void k_css_transpose_values(const BlockInfo *__restrict__ blocks,
                                                                        const DevSegment *__restrict__ segs,
                                                                        const int *__restrict__ segidx,
                                                                        const float *__restrict__ val,
                                                                        float *__restrict__ valT) {
  const BlockInfo bi = blocks[blockIdx.x];
  const DevSegment sj = segs[bi.seg];
  const int idx = bi.base + threadIdx.x;
  if (idx >= sj.count) return;
  const int rx = sj.x0 % S, ry = sj.y0 % S, rz = sj.z0 % S;
  const int i = idx % sj.cx, t = idx / sj.cx, j = t % sj.cy, k = t / sj.cy;
  const int64_t vbj = sj.voff + (int64_t)(idx / kSlice) * ((int64_t)sj.ns * kSlice) + (idx % kSlice);
#pragma unroll
  for (int dz = -E; dz <= E; dz++)
#pragma unroll
    for (int dy = -E; dy <= E; dy++)
#pragma unroll
      for (int dx = -E; dx <= E; dx++) {
        if (dx < -sj.ex || dx > sj.ex || dy < -sj.ey || dy > sj.ey || dz < -sj.ez || dz > sj.ez) continue;
        const int slot_j = (dx + sj.ex) + (2 * sj.ex + 1) * ((dy + sj.ey) + (2 * sj.ey + 1) * (dz + sj.ez));
        float v = 0.f;
        const int si = segidx[((rx + dx + S) % S) + S * (((ry + dy + S) % S) + S * ((rz + dz + S) % S))];
        if (si >= 0) {
          const DevSegment s = segs[si];
          if (dx >= -s.ex && dx <= s.ex && dy >= -s.ey && dy <= s.ey && dz >= -s.ez && dz <= s.ez) {
            const int slot = (-dx + s.ex) + (2 * s.ex + 1) * ((-dy + s.ey) + (2 * s.ey + 1) * (-dz + s.ez));
            // class-lattice index of j + d in the neighbour's class (exact division)
            const int ti = i + (sj.x0 + dx - s.x0) / S, tj = j + (sj.y0 + dy - s.y0) / S,
                      tk = k + (sj.z0 + dz - s.z0) / S;
            if (ti >= 0 && ti < s.cx && tj >= 0 && tj < s.cy && tk >= 0 && tk < s.cz) {
              const int ri = ti + s.cx * (tj + s.cy * tk);
              v = val[s.voff + (int64_t)(ri / kSlice) * ((int64_t)s.ns * kSlice) + (ri % kSlice) +
                      (int64_t)slot * kSlice];
            }
          }
        }
        valT[vbj + (int64_t)slot_j * kSlice] = v;
      }
}

void learned_mask(const float *in, int M, float *out, hipStream_t st) {
  hipLaunchKernelGGL(k_mask, dim3(grid_for((int64_t)M * M * M)), dim3(256), 0, st, in, M, out);
}
void learned_res_loss(LearnedNet &net, int M, const float *u, const float *dinv, const float *ac, float *g,
                      hipStream_t st) {
  const double inner = (double)(M - 2) * (M - 2) * (M - 2);
  hipLaunchKernelGGL(k_res_loss, dim3(kLossBlocks), dim3(256), 0, st, M, u, dinv, ac, (float)(1.0 / inner), g,
                     net.lpart);
}
void learned_loss_sum(LearnedNet &net, float *loss_dev, hipStream_t st) {
  hipLaunchKernelGGL(k_sum_parts, dim3(1), dim3(kBlock), 0, st, net.lpart, kLossBlocks, loss_dev);
}
void css_transpose_values(const DevLayout &D, const float *val, float *valT, hipStream_t st) {
  if (D.S != (D.order == 2 ? 6 : 2)) throw std::logic_error("css_transpose_values: unexpected residue modulus");
  if (D.order == 2)
    hipLaunchKernelGGL((k_css_transpose_values<2, 6>), dim3(D.nblocks_all), dim3(kRowsPerBlock), 0, st,
                       D.blocks_all, D.segs, D.segidx, val, valT);
  else
    hipLaunchKernelGGL((k_css_transpose_values<1, 2>), dim3(D.nblocks_all), dim3(kRowsPerBlock), 0, st,
                       D.blocks_all, D.segs, D.segidx, val, valT);
}
void css_transpose(const DevLayout &D, const float *valT, const float *g, float *y, hipStream_t st) {
  css_apply<float>(D, 0, valT, g, nullptr, y, 0.0, -1, st);
  boundary<float>(D, 3, y, nullptr, nullptr, st);  // y = 0 on the boundary nodes
}

}  // namespace pmx
