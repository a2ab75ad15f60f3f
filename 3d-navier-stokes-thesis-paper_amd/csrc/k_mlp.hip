// Learned subgrid closure: a 10 -> 128 -> 128 -> 128 -> 6 MLP with GELU on
// the hidden layers (no counterpart in the reference).
//
// Inference (k_mlp_infer) runs on the matrix cores in bfloat16 with float32
// accumulation: every layer is H_out^T = W H_in^T + b with the 16 points of a
// tile as the N dimension of v_mfma_f32_16x16x16_bf16 (layer 1, K = 16) and
// v_mfma_f32_16x16x32_bf16 (layers 2-4, K = 32).  In that orientation
// the 16x16 float32 result tile (lane l: rows 4(l/16)..+3, column l%16) is
// already the B operand of the next layer for the matching 16-feature
// K-block (lane l: k = 4(l/16)..+3, column l%16), so activations never leave
// the registers between layers.  The weights, packed per 16x16 block in
// A-operand order (lane l: row l%16, k = 4(l/16)..+3), sit in LDS for the
// whole persistent workgroup.
//
// Training runs in float32 with one tiled GEMM kernel and fused epilogues
// (bias, GELU, GELU derivative), a mean-squared-error gradient and Adam.
#include "chan_internal.h"
#include <cmath>

typedef short bf16x4 __attribute__((ext_vector_type(4)));
typedef __bf16 bf16x8 __attribute__((ext_vector_type(8)));
typedef short s16x8 __attribute__((ext_vector_type(8)));
// two 4-value operands side by side as one 8-value operand
__device__ __forceinline__ bf16x8 cat8(bf16x4 a, bf16x4 b) {
  const s16x8 r = __builtin_shufflevector(a, b, 0, 1, 2, 3, 4, 5, 6, 7);
  return __builtin_bit_cast(bf16x8, r);
}
typedef float f32x4 __attribute__((ext_vector_type(4)));

__device__ __forceinline__ float gelu_f(float x) { return 0.5f * x * (1.0f + erff(x * 0.70710678118654752f)); }
__device__ __forceinline__ float gelu_grad_f(float x) {
  return 0.5f * (1.0f + erff(x * 0.70710678118654752f)) + x * expf(-0.5f * x * x) * 0.39894228040143268f;
}

// Packed bf16 weights: layer 1 (8 out tiles x 1 k-block: 10 inputs padded to
// 16), layers 2 and 3 (8 x 8), layer 4 (1 x 8: 6 outputs padded to 16);
// 512 B per block.
#define MLP_BLK1 0
#define MLP_BLK2 8
#define MLP_BLK3 72
#define MLP_BLK4 136
#define MLP_NBLK 144

// Parameter offsets (floats) in the float32 vector W1,b1,W2,b2,W3,b3,W4,b4.
#define MLP_W1 0
#define MLP_B1 1280
#define MLP_W2 1408
#define MLP_B2 17792
#define MLP_W3 17920
#define MLP_B3 34304
#define MLP_W4 34432
#define MLP_B4 35200
#define MLP_NPARAM 35206

// Four float32 values rounded to bf16 (nearest even) by the hardware
// conversion, as one MFMA operand.
typedef __bf16 bf16x2h __attribute__((ext_vector_type(2)));
typedef unsigned int u32x2 __attribute__((ext_vector_type(2)));
__device__ __forceinline__ bf16x4 to_bf16x4(float a, float b, float c, float d) {
  const bf16x2h lo = {(__bf16)a, (__bf16)b}, hi = {(__bf16)c, (__bf16)d};
  const u32x2 w = {__builtin_bit_cast(unsigned int, lo), __builtin_bit_cast(unsigned int, hi)};
  return __builtin_bit_cast(bf16x4, w);
}

// Layer 1 (K = 16, v_mfma_f32_16x16x16_bf16): block ot holds lane l's A
// operand (row l%16 of output tile ot, k = 4(l/16)..+3).  Layers 2-4
// (K = 32, v_mfma_f32_16x16x32_bf16): the two 512-byte blocks of K-block kb of
// output tile ot hold lane l's 8 values (row l%16, logical k = 8(l/16)..+7) at
// entries 2l, 2l+1.  Logical k = 8q + j stands for input feature
// 32 kb + 16 (j/4) + 4q + j%4, so that a lane's B operand is its two result
// tiles of the previous layer side by side (features 4q..4q+3 of tiles 2kb
// and 2kb+1).
__global__ void k_mlp_pack(const float* __restrict__ p, bf16x4* __restrict__ packed) {
  const int blk = blockIdx.x, l = threadIdx.x;  // 64 threads
  int wofs, nin, nout, row, f0;
  if (blk < MLP_BLK2) {
    wofs = MLP_W1; nin = 10; nout = 128;
    row = blk * 16 + (l & 15);
    f0 = 4 * (l >> 4);
  } else {
    const int b0 = blk < MLP_BLK3 ? MLP_BLK2 : (blk < MLP_BLK4 ? MLP_BLK3 : MLP_BLK4);
    wofs = blk < MLP_BLK3 ? MLP_W2 : (blk < MLP_BLK4 ? MLP_W3 : MLP_W4);
    nin = 128;
    nout = blk < MLP_BLK4 ? 128 : 6;
    const int pi = (blk - b0) / 2, hb = (blk - b0) % 2;
    const int ot = pi / 4, kb = pi % 4;
    const int L = 32 * hb + l / 2, part = l % 2;   // lane of the operand, half of its 8 values
    row = ot * 16 + (L & 15);
    f0 = 32 * kb + 16 * part + 4 * (L >> 4);
  }
  // the hidden layers' weights halved (exactly: a power of two), so that
  // their products are x/2 for the GELU (gelu_half)
  const float h = blk < MLP_BLK4 ? 0.5f : 1.0f;
  float w[4];
  for (int t = 0; t < 4; ++t) {
    const int k = f0 + t;
    w[t] = (row < nout && k < nin) ? h * p[wofs + row * nin + k] : 0.0f;
  }
  packed[blk * 64 + l] = to_bf16x4(w[0], w[1], w[2], w[3]);
}

// GELU for the bf16 inference path, without transcendental instructions or
// branches, from y = x/2: the hidden layers run with halved weights and
// biases, which halves every product, sum and rounding exactly (a power of
// two), so their accumulators hold x/2 bit for bit.  GELU(x) = x/2 + |x| e(w)
// = y + |y| 2e(w), e(w) = erf(4.5 w/sqrt 2)/2 at w = min(|y|/2.25, 1) =
// min(|x|/4.5, 1), with 2e a degree-9 polynomial in w (a least-squares
// Chebyshev fit weighted by t = 4.5 w, made for this kernel, converted to
// monomials in w and doubled).  Doubling the coefficients and halving y are
// exact, so the result is that of the same polynomial on x, rounded once at
// the last fused multiply-add.  In float32 its GELU error is <= 6.3e-5
// absolute for |x| <= 8 and below 1e-5 relative beyond (erf(4.5/sqrt 2)
// differs from 1 by 7e-6), far inside the 2^-9 relative rounding of the bf16
// result the value feeds.  11 VALU operations per value: w is one multiply
// with the clamp-to-[0,1] output modifier (|y| is a source modifier), then
// 9 fused multiply-adds and the last fused multiply-add.
__device__ __forceinline__ float gelu_half(float y) {
  const float w = __builtin_amdgcn_fmed3f(fabsf(y) * (1.0f / 2.25f), 0.0f, 1.0f);
  float e = 2.110172653e+01f;
  e = fmaf(e, w, -1.103641815e+02f);
  e = fmaf(e, w, 2.381811523e+02f);
  e = fmaf(e, w, -2.659908447e+02f);
  e = fmaf(e, w, 1.515909576e+02f);
  e = fmaf(e, w, -2.882330894e+01f);
  e = fmaf(e, w, -7.994132996e+00f);
  e = fmaf(e, w, -2.997743785e-01f);
  e = fmaf(e, w, 3.598392963e+00f);
  e = fmaf(e, w, 1.978126056e-05f);
  return fmaf(fabsf(y), e, y);
}

// GELU on a result tile of a hidden layer (x/2, its halved bias already in
// the accumulator), rounded to bf16 as the next B operand.
__device__ __forceinline__ bf16x4 act_tile(f32x4 acc) {
  return to_bf16x4(gelu_half(acc[0]), gelu_half(acc[1]), gelu_half(acc[2]), gelu_half(acc[3]));
}

// The bias of rows 4(lane/16)..+3 of output tile ot: the accumulator's
// starting value, so the products add onto it.
__device__ __forceinline__ f32x4 bias_tile(const float* __restrict__ bias, int ot, int lane) {
  const float* bb = bias + ot * 16 + 4 * (lane >> 4);
  return f32x4{bb[0], bb[1], bb[2], bb[3]};
}

#define MLP_NT 3  // point tiles of 16 per wavefront (each weight operand read from LDS serves all of them)

// Two workgroups per compute unit (their 2 x 75 KB of LDS fill it): the bound
// keeps the kernel within 256 registers, so it runs at two wavefronts per SIMD
// with its accumulators in VGPRs (175, no AGPRs, no spills), instead of 196
// VGPRs + 72 AGPRs at one wavefront per SIMD with a copy out of the AGPRs
// before every GELU (114 -> 103 us at 589 824 points).  Six-wavefront
// workgroups, three wavefronts per SIMD, are slower (120 us).  What bounds the
// kernel is its vector work: per 48 points 3666 VALU instructions (the GELU's
// 12 per value over 3 x 128 values, at 4 cycles per wavefront instruction
// without packed FP32) against 228 matrix products.
#define MLP_WAVES 4
__global__ __launch_bounds__(64 * MLP_WAVES, 2) void k_mlp_infer(const bf16x4* __restrict__ packed, const float* __restrict__ p, int n,
                                                   const float* __restrict__ x, float* __restrict__ y) {
  __shared__ __align__(16) bf16x4 W[MLP_NBLK * 64];
  const bf16x8* W8 = reinterpret_cast<const bf16x8*>(W);
  __shared__ float bias[128 * 3 + 16];
  for (int i = threadIdx.x; i < MLP_NBLK * 64; i += blockDim.x) W[i] = packed[i];
  for (int i = threadIdx.x; i < 128; i += blockDim.x) {
    bias[i] = 0.5f * p[MLP_B1 + i];  // halved with the hidden layers' weights
    bias[128 + i] = 0.5f * p[MLP_B2 + i];
    bias[256 + i] = 0.5f * p[MLP_B3 + i];
  }
  if (threadIdx.x < 16) bias[384 + threadIdx.x] = threadIdx.x < 6 ? p[MLP_B4 + threadIdx.x] : 0.0f;
  __syncthreads();
  const int lane = threadIdx.x & 63, wv = threadIdx.x >> 6;
  const int nw = gridDim.x * (blockDim.x >> 6);
  const int q = lane >> 4, col = lane & 15;
  // features 4q..4q+3 of the points b + 16 nt + col (zero past n and past
  // feature 10), loaded without branches from clamped in-range addresses (n >= 1)
  float xv[MLP_NT][4];
  auto load_x = [&](long b) {
#pragma unroll
    for (int nt = 0; nt < MLP_NT; ++nt) {
      const long pt = b + 16 * nt + col;
      const float* xp = x + (pt < n ? pt : (long)n - 1) * 10;
#pragma unroll
      for (int t = 0; t < 4; ++t) {
        const int f = 4 * q + t;
        const float v = xp[f < 10 ? f : 9];
        xv[nt][t] = (pt < n && f < 10) ? v : 0.0f;
      }
    }
  };
  for (long base = ((long)blockIdx.x * (blockDim.x >> 6) + wv) * 16 * MLP_NT; base < n; base += (long)nw * 16 * MLP_NT) {
    bf16x4 h[MLP_NT][8];
    // layer 1: B operand = features 4q..4q+3 of point base + 16 nt + col
    load_x(base);
#pragma unroll
    for (int nt = 0; nt < MLP_NT; ++nt) h[nt][0] = to_bf16x4(xv[nt][0], xv[nt][1], xv[nt][2], xv[nt][3]);
    // Each layer is software-pipelined over its 8 output tiles: the matrix
    // products of tile ot+1 are issued before the GELU of tile ot, so that the
    // wavefront's vector instructions run while its products are in flight.
    f32x4 acc[2][MLP_NT];
    bf16x4 g[MLP_NT][8];
    auto layer1 = [&](int ot, f32x4 (&o)[MLP_NT]) {
      const bf16x4 a = W[(MLP_BLK1 + ot) * 64 + lane];
      const f32x4 b0 = bias_tile(bias, ot, lane);
#pragma unroll
      for (int nt = 0; nt < MLP_NT; ++nt) o[nt] = __builtin_amdgcn_mfma_f32_16x16x16bf16_1k(a, h[nt][0], b0, 0, 0, 0);
    };
    layer1(0, acc[0]);
#pragma unroll
    for (int ot = 0; ot < 8; ++ot) {
      if (ot + 1 < 8) layer1(ot + 1, acc[(ot + 1) & 1]);
#pragma unroll
      for (int nt = 0; nt < MLP_NT; ++nt) g[nt][ot] = act_tile(acc[ot & 1][nt]);
      if (ot + 1 < 8) {
        __builtin_amdgcn_sched_group_barrier(0x100, 2, 0);
#pragma unroll
        for (int i = 0; i < MLP_NT; ++i) {
          __builtin_amdgcn_sched_group_barrier(0x008, 1, 0);
          __builtin_amdgcn_sched_group_barrier(0x002, 48, 0);
        }
      }
      __builtin_amdgcn_sched_barrier(0);
    }
#pragma unroll
    for (int nt = 0; nt < MLP_NT; ++nt)
#pragma unroll
      for (int ot = 0; ot < 8; ++ot) h[nt][ot] = g[nt][ot];
    // layers 2 and 3
#pragma unroll
    for (int layer = 0; layer < 2; ++layer) {
      const int blk0 = layer == 0 ? MLP_BLK2 : MLP_BLK3;
      const float* bl = bias + 128 * (layer + 1);
      auto products = [&](int ot, f32x4 (&o)[MLP_NT]) {
        const f32x4 b0 = bias_tile(bl, ot, lane);
#pragma unroll
        for (int nt = 0; nt < MLP_NT; ++nt) o[nt] = b0;
#pragma unroll
        for (int kb = 0; kb < 4; ++kb) {
          const bf16x8 a = W8[(blk0 + (ot * 4 + kb) * 2) * 32 + lane];
#pragma unroll
          for (int nt = 0; nt < MLP_NT; ++nt)
            o[nt] = __builtin_amdgcn_mfma_f32_16x16x32_bf16(a, cat8(h[nt][2 * kb], h[nt][2 * kb + 1]), o[nt], 0, 0, 0);
        }
      };
      products(0, acc[0]);
#pragma unroll
      for (int ot = 0; ot < 8; ++ot) {
        if (ot + 1 < 8) products(ot + 1, acc[(ot + 1) & 1]);
#pragma unroll
        for (int nt = 0; nt < MLP_NT; ++nt) g[nt][ot] = act_tile(acc[ot & 1][nt]);
        if (ot + 1 < 8) {
          // the weight and bias reads, then each product followed by its
          // share of the GELU's vector instructions
          __builtin_amdgcn_sched_group_barrier(0x100, 5, 0);
#pragma unroll
          for (int i = 0; i < 4 * MLP_NT; ++i) {
            __builtin_amdgcn_sched_group_barrier(0x008, 1, 0);
            __builtin_amdgcn_sched_group_barrier(0x002, 12, 0);
          }
        }
        __builtin_amdgcn_sched_barrier(0);
      }
#pragma unroll
      for (int nt = 0; nt < MLP_NT; ++nt)
#pragma unroll
        for (int ot = 0; ot < 8; ++ot) h[nt][ot] = g[nt][ot];
    }
    // layer 4 (linear): rows 4q..4q+3 of the output tile are outputs, column = point
    const f32x4 b4 = bias_tile(bias + 384, 0, lane);
#pragma unroll
    for (int nt = 0; nt < MLP_NT; ++nt) {
      f32x4 o = b4;
#pragma unroll
      for (int kb = 0; kb < 4; ++kb)
        o = __builtin_amdgcn_mfma_f32_16x16x32_bf16(W8[(MLP_BLK4 + kb * 2) * 32 + lane],
                                                    cat8(h[nt][2 * kb], h[nt][2 * kb + 1]), o, 0, 0, 0);
      const long pt = base + 16 * nt + col;
      if (pt < n)
        for (int i = 0; i < 4; ++i) {
          const int f = 4 * q + i;
          if (f < 6) y[pt * 6 + f] = o[i];
        }
    }
  }
}

// ---------------------------------------------------------------------------
// float32 training kernels
// ---------------------------------------------------------------------------
// C(m, n) = sum_k A(m, k) B(k, n), A(m, k) = a[m*as0 + k*as1], B(k, n) = b[k*bs0 + n*bs1],
// C row-major with row stride ldc.  Epilogue: EPI 0 plain; 1 + bias[n], store
// the pre-activation in C and GELU of it in act (or C only when act is null,
// i.e. a linear layer); 2 multiply by gelu'(aux[m*ldc + n]).
template <int EPI>
__global__ __launch_bounds__(256) void k_gemm(int M, int N, int K, const float* __restrict__ a, long as0, long as1,
                                              const float* __restrict__ b, long bs0, long bs1, float* __restrict__ c,
                                              long ldc, const float* __restrict__ bias, float* __restrict__ act,
                                              const float* __restrict__ aux) {
  __shared__ float As[16][65], Bs[16][65];
  const int tx = threadIdx.x & 15, ty = threadIdx.x >> 4;
  const int m0 = blockIdx.y * 64, n0 = blockIdx.x * 64;
  float acc[4][4] = {};
  for (int k0 = 0; k0 < K; k0 += 16) {
    for (int e = threadIdx.x; e < 16 * 64; e += 256) {
      const int kk = e & 15, mm = e >> 4;
      const int gm = m0 + mm, gk = k0 + kk;
      As[kk][mm] = (gm < M && gk < K) ? a[gm * as0 + gk * as1] : 0.0f;
      const int nn = e >> 4;
      const int gn = n0 + nn;
      Bs[kk][nn] = (gn < N && gk < K) ? b[gk * bs0 + gn * bs1] : 0.0f;
    }
    __syncthreads();
#pragma unroll
    for (int kk = 0; kk < 16; ++kk) {
      float av[4], bv[4];
#pragma unroll
      for (int i = 0; i < 4; ++i) { av[i] = As[kk][ty + 16 * i]; bv[i] = Bs[kk][tx + 16 * i]; }
#pragma unroll
      for (int i = 0; i < 4; ++i)
#pragma unroll
        for (int j = 0; j < 4; ++j) acc[i][j] = fmaf(av[i], bv[j], acc[i][j]);
    }
    __syncthreads();
  }
#pragma unroll
  for (int i = 0; i < 4; ++i)
#pragma unroll
    for (int j = 0; j < 4; ++j) {
      const int gm = m0 + ty + 16 * i, gn = n0 + tx + 16 * j;
      if (gm >= M || gn >= N) continue;
      float v = acc[i][j];
      const long o = (long)gm * ldc + gn;
      if (EPI == 1) {
        v += bias[gn];
        c[o] = v;
        if (act) act[o] = gelu_f(v);
      } else if (EPI == 2) {
        c[o] = v * gelu_grad_f(aux[o]);
      } else {
        c[o] = v;
      }
    }
}

// d = 2 (y - t)/(n*m); partial sums of (y - t)^2 per workgroup into part.
__global__ __launch_bounds__(256) void k_mse_grad(long nm, const float* __restrict__ yv, const float* __restrict__ t,
                                                  float* __restrict__ d, double* __restrict__ part) {
  __shared__ double red[4];
  double s = 0.0;
  const float sc = 2.0f / (float)nm;
  for (long i = (long)blockIdx.x * blockDim.x + threadIdx.x; i < nm; i += (long)gridDim.x * blockDim.x) {
    const float e = yv[i] - t[i];
    d[i] = sc * e;
    s += (double)e * (double)e;
  }
  for (int o = 32; o > 0; o >>= 1) s += __shfl_xor(s, o, 64);
  if ((threadIdx.x & 63) == 0) red[threadIdx.x >> 6] = s;
  __syncthreads();
  if (threadIdx.x == 0) part[blockIdx.x] = red[0] + red[1] + red[2] + red[3];
}

// Column sums: g[j] = sum_i d[i*N + j] (bias gradients).
__global__ __launch_bounds__(256) void k_colsum(int M, int N, const float* __restrict__ d, float* __restrict__ g) {
  __shared__ float red[256];
  const int j = blockIdx.x;
  float s = 0.0f;
  for (int i = threadIdx.x; i < M; i += blockDim.x) s += d[(long)i * N + j];
  red[threadIdx.x] = s;
  __syncthreads();
  for (int w = 128; w > 0; w >>= 1) {
    if ((int)threadIdx.x < w) red[threadIdx.x] += red[threadIdx.x + w];
    __syncthreads();
  }
  if (threadIdx.x == 0) g[j] = red[0];
}

// Adam (Kingma & Ba 2015, Algorithm 1), step t >= 1.
__global__ void k_adam(int n, float* __restrict__ p, const float* __restrict__ g, float* __restrict__ m,
                       float* __restrict__ v, float lr, float b1, float b2, float eps, float c1, float c2) {
  const int i = blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= n) return;
  const float gi = g[i];
  const float mi = b1 * m[i] + (1.0f - b1) * gi;
  const float vi = b2 * v[i] + (1.0f - b2) * gi * gi;
  m[i] = mi;
  v[i] = vi;
  p[i] -= lr * (mi / c1) / (sqrtf(vi / c2) + eps);
}

// ---------------------------------------------------------------------------
// host side
// ---------------------------------------------------------------------------
struct chan_mlp {
  hipStream_t stream = nullptr;
  float* d_p = nullptr;      // float32 parameters
  float* d_g = nullptr;      // gradients
  float* d_m = nullptr;      // Adam moments
  float* d_v = nullptr;
  bf16x4* d_packed = nullptr;
  bool packed_ok = false;
  int adam_t = 0;
  // work buffers (capacity in points)
  long cap = 0;
  float *d_x = nullptr, *d_y = nullptr, *d_t = nullptr, *d_d = nullptr;
  float *d_z[4] = {}, *d_a[3] = {}, *d_e[2] = {};
  double* d_part = nullptr;
  std::string last_error;
};

namespace {

const int kIn[4] = {10, 128, 128, 128}, kOut[4] = {128, 128, 128, 6};
const int kW[4] = {MLP_W1, MLP_W2, MLP_W3, MLP_W4}, kB[4] = {MLP_B1, MLP_B2, MLP_B3, MLP_B4};

void mchk(hipError_t e, const char* what) {
  if (e != hipSuccess) throw chan_error(std::string(what) + ": " + hipGetErrorString(e));
}

template <class T>
void mfree(T*& p) {
  if (p) (void)hipFree(p);
  p = nullptr;
}

void ensure_cap(chan_mlp* m, long n) {
  if (n <= m->cap) return;
  mfree(m->d_x); mfree(m->d_y); mfree(m->d_t); mfree(m->d_d); mfree(m->d_part);
  for (auto& z : m->d_z) mfree(z);
  for (auto& a : m->d_a) mfree(a);
  for (auto& e : m->d_e) mfree(e);
  mchk(hipMalloc(&m->d_x, sizeof(float) * 10 * n), "hipMalloc");
  mchk(hipMalloc(&m->d_y, sizeof(float) * 6 * n), "hipMalloc");
  mchk(hipMalloc(&m->d_t, sizeof(float) * 6 * n), "hipMalloc");
  mchk(hipMalloc(&m->d_d, sizeof(float) * 6 * n), "hipMalloc");
  for (int l = 0; l < 4; ++l) mchk(hipMalloc(&m->d_z[l], sizeof(float) * kOut[l] * n), "hipMalloc");
  for (int l = 0; l < 3; ++l) mchk(hipMalloc(&m->d_a[l], sizeof(float) * 128 * n), "hipMalloc");
  for (int l = 0; l < 2; ++l) mchk(hipMalloc(&m->d_e[l], sizeof(float) * 128 * n), "hipMalloc");
  mchk(hipMalloc(&m->d_part, sizeof(double) * 1024), "hipMalloc");
  m->cap = n;
}

template <int EPI>
void gemm(chan_mlp* m, int M, int N, int K, const float* a, long as0, long as1, const float* b, long bs0, long bs1,
          float* c, long ldc, const float* bias = nullptr, float* act = nullptr, const float* aux = nullptr) {
  dim3 grid((N + 63) / 64, (M + 63) / 64);
  hipLaunchKernelGGL(k_gemm<EPI>, grid, dim3(256), 0, m->stream, M, N, K, a, as0, as1, b, bs0, bs1, c, ldc, bias,
                     act, aux);
  LAUNCH_CHECK("k_gemm");
}

// Forward pass of n points at d_x: pre-activations d_z[l], activations d_a[l]
// (layer inputs of layers 2..4), outputs d_z[3].
void forward32(chan_mlp* m, int n) {
  const float* in = m->d_x;
  for (int l = 0; l < 4; ++l) {
    gemm<1>(m, n, kOut[l], kIn[l], in, kIn[l], 1, m->d_p + kW[l], 1, kIn[l], m->d_z[l], kOut[l], m->d_p + kB[l],
            l < 3 ? m->d_a[l] : nullptr);
    if (l < 3) in = m->d_a[l];
  }
}

// Gradient of the mean squared error of n points (targets at d_t) into d_g;
// returns the loss.
double gradient32(chan_mlp* m, int n) {
  forward32(m, n);
  const long nm = 6L * n;
  const int blocks = (int)std::min<long>(1024, (nm + 255) / 256);
  hipLaunchKernelGGL(k_mse_grad, dim3(blocks), dim3(256), 0, m->stream, nm, m->d_z[3], m->d_t, m->d_d, m->d_part);
  LAUNCH_CHECK("k_mse_grad");
  const float* d = m->d_d;  // dL/dz of the current layer, [n][out]
  for (int l = 3; l >= 0; --l) {
    const float* xin = l == 0 ? m->d_x : m->d_a[l - 1];
    // dW[o][i] = sum_b d[b][o] xin[b][i]
    gemm<0>(m, kOut[l], kIn[l], n, d, 1, kOut[l], xin, kIn[l], 1, m->d_g + kW[l], kIn[l]);
    hipLaunchKernelGGL(k_colsum, dim3(kOut[l]), dim3(256), 0, m->stream, n, kOut[l], d, m->d_g + kB[l]);
    LAUNCH_CHECK("k_colsum");
    if (l > 0) {
      float* dn = m->d_e[l & 1];
      // dxin[b][i] = sum_o d[b][o] W[o][i], times gelu'(z[l-1][b][i])
      gemm<2>(m, n, kIn[l], kOut[l], d, kOut[l], 1, m->d_p + kW[l], kIn[l], 1, dn, kIn[l], nullptr, nullptr,
              m->d_z[l - 1]);
      d = dn;
    }
  }
  std::vector<double> part(blocks);
  mchk(hipMemcpyAsync(part.data(), m->d_part, sizeof(double) * blocks, hipMemcpyDeviceToHost, m->stream), "copy");
  mchk(hipStreamSynchronize(m->stream), "sync");
  double s = 0.0;
  for (double v : part) s += v;
  return s / (double)nm;
}

void pack(chan_mlp* m) {
  if (m->packed_ok) return;
  hipLaunchKernelGGL(k_mlp_pack, dim3(MLP_NBLK), dim3(64), 0, m->stream, m->d_p, m->d_packed);
  LAUNCH_CHECK("k_mlp_pack");
  m->packed_ok = true;
}

template <class F>
int mguard(chan_mlp* m, F&& f) {
  if (!m) return 1;
  try {
    f();
    return 0;
  } catch (const chan_error& e) {
    m->last_error = e.msg;
  } catch (const std::exception& e) {
    m->last_error = e.what();
  }
  return 1;
}

}  // namespace

// Workgroups of the persistent inference kernel: as many as fit on the device
// at once (its 75 KB of LDS weights allow two per CU), so that every
// workgroup loads the weights once.
static int mlp_blocks(long n) {
  int dev = 0, cus = 0;
  if (hipGetDevice(&dev) != hipSuccess ||
      hipDeviceGetAttribute(&cus, hipDeviceAttributeMultiprocessorCount, dev) != hipSuccess || cus < 1)
    cus = 256;
  const long tiles = (n + 16 * MLP_NT - 1) / (16 * MLP_NT);
  return (int)std::max<long>(1, std::min<long>(2L * cus, (tiles + MLP_WAVES - 1) / MLP_WAVES));
}

void mlp_prepare_device(chan_mlp* m) {
  pack(m);
  mchk(hipStreamSynchronize(m->stream), "sync");
}

// inference of n points already on the device, on the caller's stream
// (weights changed since the last packing are packed first, synchronised)
void mlp_infer_device(chan_mlp* m, long n, const float* x, float* y, hipStream_t s) {
  if (!m->packed_ok) mlp_prepare_device(m);
  if (n <= 0) return;
  if (n > 0x7fffffffL) throw chan_error("too many points for one inference");
  hipLaunchKernelGGL(k_mlp_infer, dim3(mlp_blocks(n)), dim3(64 * MLP_WAVES), 0, s, m->d_packed, m->d_p, (int)n, x, y);
  LAUNCH_CHECK("k_mlp_infer");
}

extern "C" {

int chan_mlp_create(chan_mlp** out) {
  if (!out) return 1;
  *out = nullptr;
  chan_mlp* m = new chan_mlp();
  if (mguard(m, [&] {
        mchk(hipStreamCreateWithFlags(&m->stream, hipStreamNonBlocking), "stream");
        mchk(hipMalloc(&m->d_p, sizeof(float) * MLP_NPARAM), "hipMalloc");
        mchk(hipMalloc(&m->d_g, sizeof(float) * MLP_NPARAM), "hipMalloc");
        mchk(hipMalloc(&m->d_m, sizeof(float) * MLP_NPARAM), "hipMalloc");
        mchk(hipMalloc(&m->d_v, sizeof(float) * MLP_NPARAM), "hipMalloc");
        mchk(hipMalloc(&m->d_packed, sizeof(bf16x4) * MLP_NBLK * 64), "hipMalloc");
        mchk(hipMemset(m->d_p, 0, sizeof(float) * MLP_NPARAM), "memset");
        mchk(hipMemset(m->d_m, 0, sizeof(float) * MLP_NPARAM), "memset");
        mchk(hipMemset(m->d_v, 0, sizeof(float) * MLP_NPARAM), "memset");
      })) {
    chan_mlp_destroy(m);
    return 1;
  }
  *out = m;
  return 0;
}

int chan_mlp_destroy(chan_mlp* m) {
  if (!m) return 0;
  if (m->stream) (void)hipStreamSynchronize(m->stream);
  mfree(m->d_p); mfree(m->d_g); mfree(m->d_m); mfree(m->d_v); mfree(m->d_packed);
  mfree(m->d_x); mfree(m->d_y); mfree(m->d_t); mfree(m->d_d); mfree(m->d_part);
  for (auto& z : m->d_z) mfree(z);
  for (auto& a : m->d_a) mfree(a);
  for (auto& e : m->d_e) mfree(e);
  if (m->stream) (void)hipStreamDestroy(m->stream);
  delete m;
  return 0;
}

const char* chan_mlp_last_error(const chan_mlp* m) { return m ? m->last_error.c_str() : "null network"; }

int chan_mlp_num_params(void) { return MLP_NPARAM; }

int chan_mlp_set_params(chan_mlp* m, const float* p) {
  return mguard(m, [&] {
    if (!p) throw chan_error("null argument");
    mchk(hipMemcpyAsync(m->d_p, p, sizeof(float) * MLP_NPARAM, hipMemcpyHostToDevice, m->stream), "copy");
    mchk(hipMemsetAsync(m->d_m, 0, sizeof(float) * MLP_NPARAM, m->stream), "memset");
    mchk(hipMemsetAsync(m->d_v, 0, sizeof(float) * MLP_NPARAM, m->stream), "memset");
    mchk(hipStreamSynchronize(m->stream), "sync");
    m->adam_t = 0;
    m->packed_ok = false;
  });
}

int chan_mlp_get_params(chan_mlp* m, float* p) {
  return mguard(m, [&] {
    if (!p) throw chan_error("null argument");
    mchk(hipMemcpyAsync(p, m->d_p, sizeof(float) * MLP_NPARAM, hipMemcpyDeviceToHost, m->stream), "copy");
    mchk(hipStreamSynchronize(m->stream), "sync");
  });
}

int chan_mlp_infer(chan_mlp* m, int n, const float* x, float* y) {
  return mguard(m, [&] {
    if (n < 0 || (n && (!x || !y))) throw chan_error("invalid arguments");
    if (n == 0) return;
    ensure_cap(m, n);
    pack(m);
    mchk(hipMemcpyAsync(m->d_x, x, sizeof(float) * 10 * n, hipMemcpyHostToDevice, m->stream), "copy");
    hipLaunchKernelGGL(k_mlp_infer, dim3(mlp_blocks(n)), dim3(64 * MLP_WAVES), 0, m->stream, m->d_packed, m->d_p, n, m->d_x,
                       m->d_y);
    LAUNCH_CHECK("k_mlp_infer");
    mchk(hipMemcpyAsync(y, m->d_y, sizeof(float) * 6 * n, hipMemcpyDeviceToHost, m->stream), "copy");
    mchk(hipStreamSynchronize(m->stream), "sync");
  });
}

int chan_mlp_forward_fp32(chan_mlp* m, int n, const float* x, float* y) {
  return mguard(m, [&] {
    if (n < 1 || !x || !y) throw chan_error("invalid arguments");
    ensure_cap(m, n);
    mchk(hipMemcpyAsync(m->d_x, x, sizeof(float) * 10 * n, hipMemcpyHostToDevice, m->stream), "copy");
    forward32(m, n);
    mchk(hipMemcpyAsync(y, m->d_z[3], sizeof(float) * 6 * n, hipMemcpyDeviceToHost, m->stream), "copy");
    mchk(hipStreamSynchronize(m->stream), "sync");
  });
}

int chan_mlp_gradient(chan_mlp* m, int n, const float* x, const float* t, float* grad, double* loss) {
  return mguard(m, [&] {
    if (n < 1 || !x || !t) throw chan_error("invalid arguments");
    ensure_cap(m, n);
    mchk(hipMemcpyAsync(m->d_x, x, sizeof(float) * 10 * n, hipMemcpyHostToDevice, m->stream), "copy");
    mchk(hipMemcpyAsync(m->d_t, t, sizeof(float) * 6 * n, hipMemcpyHostToDevice, m->stream), "copy");
    const double l = gradient32(m, n);
    if (grad) {
      mchk(hipMemcpyAsync(grad, m->d_g, sizeof(float) * MLP_NPARAM, hipMemcpyDeviceToHost, m->stream), "copy");
      mchk(hipStreamSynchronize(m->stream), "sync");
    }
    if (loss) *loss = l;
  });
}

// Mini-batch Adam on n points (in the given order, batches of `batch`, the
// incomplete last batch dropped) for `epochs` passes; loss_out receives the
// loss of the last batch.
int chan_mlp_train(chan_mlp* m, int n, const float* x, const float* t, int epochs, int batch, double lr,
                   double* loss_out) {
  return mguard(m, [&] {
    if (n < 1 || !x || !t || epochs < 1 || batch < 1 || batch > n) throw chan_error("invalid arguments");
    ensure_cap(m, n);
    float *dx_all = nullptr, *dt_all = nullptr;
    mchk(hipMalloc(&dx_all, sizeof(float) * 10 * n), "hipMalloc");
    mchk(hipMalloc(&dt_all, sizeof(float) * 6 * n), "hipMalloc");
    double l = 0.0;
    try {
      mchk(hipMemcpyAsync(dx_all, x, sizeof(float) * 10 * n, hipMemcpyHostToDevice, m->stream), "copy");
      mchk(hipMemcpyAsync(dt_all, t, sizeof(float) * 6 * n, hipMemcpyHostToDevice, m->stream), "copy");
      const int nb = n / batch;
      for (int e = 0; e < epochs; ++e)
        for (int bi = 0; bi < nb; ++bi) {
          mchk(hipMemcpyAsync(m->d_x, dx_all + (long)bi * batch * 10, sizeof(float) * 10 * batch,
                              hipMemcpyDeviceToDevice, m->stream), "copy");
          mchk(hipMemcpyAsync(m->d_t, dt_all + (long)bi * batch * 6, sizeof(float) * 6 * batch,
                              hipMemcpyDeviceToDevice, m->stream), "copy");
          l = gradient32(m, batch);
          m->adam_t += 1;
          const float b1 = 0.9f, b2 = 0.999f;
          const float c1 = 1.0f - std::pow(b1, (float)m->adam_t), c2 = 1.0f - std::pow(b2, (float)m->adam_t);
          hipLaunchKernelGGL(k_adam, dim3((MLP_NPARAM + 255) / 256), dim3(256), 0, m->stream, MLP_NPARAM, m->d_p,
                             m->d_g, m->d_m, m->d_v, (float)lr, b1, b2, 1e-8f, c1, c2);
          LAUNCH_CHECK("k_adam");
        }
      mchk(hipStreamSynchronize(m->stream), "sync");
    } catch (...) {
      (void)hipFree(dx_all);
      (void)hipFree(dt_all);
      throw;
    }
    (void)hipFree(dx_all);
    (void)hipFree(dt_all);
    m->packed_ok = false;
    if (loss_out) *loss_out = l;
  });
}

}  // extern "C"
