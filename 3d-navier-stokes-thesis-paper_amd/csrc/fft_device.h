// Device-side FFT building blocks for the channel-flow step on CDNA4 (gfx950).
//
// All transforms are FP64, unnormalised, with FFTW's sign convention
// (forward = exp(-i..), backward = exp(+i..)), which is what the reference's
// plans use (PFC/transform.f90:35,75-158).
//
// A workgroup transforms a batch of B sequences of length N held in LDS
// (sequence s at lds + s*sstride).  Each radix stage is a Stockham autosort
// pass: butterfly i (0 <= i < N/R) reads in[i + r*N/R], multiplies by the
// twiddle w_{pR}^{r*k} (k = i mod p), applies a radix-R DFT and writes
// out[(i/p)*p*R + k + q*p].  The pass is done in place: every thread first
// loads and transforms all of its butterflies into registers, the workgroup
// synchronises, then the results are stored.  The host checks that each
// thread owns at most FFT_MAXV complex values per stage.
#pragma once
#include <hip/hip_runtime.h>

#define FFT_MAXSTAGES 10
#define FFT_MAXV 16

struct FFTPlan {
  int n;                      // transform length
  int nst;                    // number of radix stages
  int rad[FFT_MAXSTAGES];     // radices in application order
  int tw;                     // offset (in double2) of the root table w_n^t = exp(-2 pi i t/n)
};

__device__ __forceinline__ double2 cadd(double2 a, double2 b) { return make_double2(a.x + b.x, a.y + b.y); }
__device__ __forceinline__ double2 csub(double2 a, double2 b) { return make_double2(a.x - b.x, a.y - b.y); }
__device__ __forceinline__ double2 cmul(double2 a, double2 b) {
  return make_double2(a.x * b.x - a.y * b.y, a.x * b.y + a.y * b.x);
}
__device__ __forceinline__ double2 cconj(double2 a) { return make_double2(a.x, -a.y); }
__device__ __forceinline__ double2 cscale(double s, double2 a) { return make_double2(s * a.x, s * a.y); }
// multiply by +i (s=+1) or -i (s=-1)
__device__ __forceinline__ double2 cmuli(double2 a, int s) {
  return s > 0 ? make_double2(-a.y, a.x) : make_double2(a.y, -a.x);
}

// radix-R DFT in registers; SGN = -1 forward, +1 backward
template <int R, int SGN> __device__ __forceinline__ void dft_r(double2* v);

template <> __device__ __forceinline__ void dft_r<2, -1>(double2* v) {
  double2 a = v[0], b = v[1]; v[0] = cadd(a, b); v[1] = csub(a, b);
}
template <> __device__ __forceinline__ void dft_r<2, 1>(double2* v) {
  double2 a = v[0], b = v[1]; v[0] = cadd(a, b); v[1] = csub(a, b);
}

template <int SGN> __device__ __forceinline__ void dft4(double2& a0, double2& a1, double2& a2, double2& a3) {
  double2 s02 = cadd(a0, a2), d02 = csub(a0, a2);
  double2 s13 = cadd(a1, a3), d13 = cmuli(csub(a1, a3), SGN);
  a0 = cadd(s02, s13);
  a2 = csub(s02, s13);
  a1 = cadd(d02, d13);
  a3 = csub(d02, d13);
}
template <> __device__ __forceinline__ void dft_r<4, -1>(double2* v) { dft4<-1>(v[0], v[1], v[2], v[3]); }
template <> __device__ __forceinline__ void dft_r<4, 1>(double2* v) { dft4<1>(v[0], v[1], v[2], v[3]); }

template <int SGN> __device__ __forceinline__ void dft3(double2* v) {
  const double s3 = 0.86602540378443864676372317075294;  // sqrt(3)/2
  double2 a = v[0], b = v[1], c = v[2];
  double2 sbc = cadd(b, c);
  double2 t1 = make_double2(a.x - 0.5 * sbc.x, a.y - 0.5 * sbc.y);
  double2 dbc = csub(b, c);
  // SGN*i*s3*(b-c)
  double2 t2 = cmuli(make_double2(s3 * dbc.x, s3 * dbc.y), SGN);
  v[0] = cadd(a, sbc);
  v[1] = cadd(t1, t2);
  v[2] = csub(t1, t2);
}
template <> __device__ __forceinline__ void dft_r<3, -1>(double2* v) { dft3<-1>(v); }
template <> __device__ __forceinline__ void dft_r<3, 1>(double2* v) { dft3<1>(v); }

template <int SGN> __device__ __forceinline__ void dft8(double2* v) {
  const double h = 0.70710678118654752440084436210485;  // sqrt(2)/2
  double2 e0 = v[0], e1 = v[2], e2 = v[4], e3 = v[6];
  double2 o0 = v[1], o1 = v[3], o2 = v[5], o3 = v[7];
  dft4<SGN>(e0, e1, e2, e3);
  dft4<SGN>(o0, o1, o2, o3);
  // twiddles w8^q, w8 = exp(SGN*2*pi*i/8)
  double2 t1 = make_double2(h * (o1.x - SGN * o1.y), h * (o1.y + SGN * o1.x));       // *(h + SGN*i*h)
  double2 t2 = cmuli(o2, SGN);                                                       // *(SGN*i)
  double2 t3 = make_double2(h * (-o3.x - SGN * o3.y), h * (-o3.y + SGN * o3.x));     // *(-h + SGN*i*h)
  v[0] = cadd(e0, o0); v[4] = csub(e0, o0);
  v[1] = cadd(e1, t1); v[5] = csub(e1, t1);
  v[2] = cadd(e2, t2); v[6] = csub(e2, t2);
  v[3] = cadd(e3, t3); v[7] = csub(e3, t3);
}
template <> __device__ __forceinline__ void dft_r<8, -1>(double2* v) { dft8<-1>(v); }
template <> __device__ __forceinline__ void dft_r<8, 1>(double2* v) { dft8<1>(v); }

// One Stockham stage of radix R over B sequences (in place, register staged).
template <int R, int SGN, int MAXV>
__device__ __forceinline__ void fft_stage(double2* lds, int sstride, int B, int N, int p,
                                          const double2* __restrict__ roots, int tid, int nthr) {
  constexpr int MB = MAXV / R;  // butterflies per thread
  const int nbf = N / R;            // butterflies per sequence
  const int tot = B * nbf;
  const int tstep = N / (p * R);    // root-table stride for this stage
  double2 v[MB][R];
#pragma unroll
  for (int b = 0; b < MB; ++b) {
    int g = tid + b * nthr;
    if (g < tot) {
      int s = g / nbf, i = g - s * nbf;
      int k = i % p;
      const double2* src = lds + s * sstride;
#pragma unroll
      for (int r = 0; r < R; ++r) v[b][r] = src[i + r * nbf];
#pragma unroll
      for (int r = 1; r < R; ++r) {
        if (k) {
          double2 w = roots[r * k * tstep];
          if (SGN > 0) w = cconj(w);
          v[b][r] = cmul(v[b][r], w);
        }
      }
      dft_r<R, SGN>(v[b]);
    }
  }
  __syncthreads();
#pragma unroll
  for (int b = 0; b < MB; ++b) {
    int g = tid + b * nthr;
    if (g < tot) {
      int s = g / nbf, i = g - s * nbf;
      int k = i % p;
      int j = (i / p) * p * R + k;
      double2* dst = lds + s * sstride;
#pragma unroll
      for (int q = 0; q < R; ++q) dst[j + q * p] = v[b][q];
    }
  }
  __syncthreads();
}

// Full in-place FFT of B sequences of length pl.n held in LDS.  Must be called
// by every thread of the workgroup (contains barriers).  MAXV bounds the
// complex values a thread stages per stage (checked on the host).
template <int SGN, int MAXV = FFT_MAXV>
__device__ void fft_batch(double2* lds, int sstride, int B, const FFTPlan& pl,
                          const double2* __restrict__ twbase, int tid, int nthr) {
  const double2* roots = twbase + pl.tw;
  int p = 1;
  for (int st = 0; st < pl.nst; ++st) {
    int R = pl.rad[st];
    switch (R) {
      case 8: fft_stage<8, SGN, MAXV>(lds, sstride, B, pl.n, p, roots, tid, nthr); break;
      case 4: fft_stage<4, SGN, MAXV>(lds, sstride, B, pl.n, p, roots, tid, nthr); break;
      case 3: fft_stage<3, SGN, MAXV>(lds, sstride, B, pl.n, p, roots, tid, nthr); break;
      default: fft_stage<2, SGN, MAXV>(lds, sstride, B, pl.n, p, roots, tid, nthr); break;
    }
    p *= R;
  }
}


// ---------------------------------------------------------------------------
// Two-pass (four-step) FFT with a compile-time length N = N1 * N2.
//
// A sequence is held in LDS as an N1 x N2 array with row stride N2 + 1.
// Input x[n], n = a*N2 + b, sits at row a, column b.  Pass 1 transforms each
// column b over a (length N1, in registers) and multiplies by w_N^{b*k1};
// pass 2 transforms each row k1 over b (length N2).  Both passes read and
// write the same LDS cells, so the transform is in place with one barrier
// between the passes.  The result X[k], k = k1 + N1*k2, ends at row k1,
// column k2 (fft2_out_pos).  With N2 even the odd row stride keeps the
// 16-byte row-wise accesses of pass 2 free of bank conflicts.
//
// The in-register DFTs of length L (4, 6, 8, 9, 12, 16, 18, 24) are built the
// same way from the radix-2/3/4/8 kernels above; their twiddles w_L^e are read
// from the length-N root table at the compile-time index e*N/L (uniform
// scalar loads).
// ---------------------------------------------------------------------------
struct Fft2Split { int n1, n2; };

__host__ __device__ constexpr Fft2Split fft2_split(int n) {
  switch (n) {
    case 16: return {4, 4};
    case 24: return {4, 6};
    case 32: return {4, 8};
    case 36: return {6, 6};
    case 48: return {6, 8};
    case 64: return {8, 8};
    case 72: return {9, 8};
    case 96: return {8, 12};
    case 144: return {12, 12};
    case 192: return {12, 16};
    case 288: return {18, 16};
    default: return {0, 0};
  }
}

// LDS cells one sequence occupies: N1*(N2+1) for a compile-time length, n+1
// for the runtime-planned transform.
__host__ __device__ constexpr int fft_seq_cells(int n) {
  return fft2_split(n).n1 ? fft2_split(n).n1 * (fft2_split(n).n2 + 1) : n + 1;
}

__host__ __device__ constexpr int rdft_first(int L) {
  return (L % 4 == 0 && L > 8) ? 4 : (L % 3 == 0 ? 3 : 2);
}

// x * w_L^e (SGN = -1: w = exp(-2 pi i/L); SGN = +1: conjugate), e a constant
// after unrolling; rt[e*S] = exp(-2 pi i e/L).
template <int L, int SGN, int S>
__device__ __forceinline__ double2 rtwiddle(double2 x, int e, const double2* __restrict__ rt) {
  e %= L;
  if (e == 0) return x;
  if (2 * e == L) return make_double2(-x.x, -x.y);
  if (4 * e == L) return cmuli(x, SGN);
  if (4 * e == 3 * L) return cmuli(x, -SGN);
  double2 w = rt[e * S];
  if (SGN > 0) w = cconj(w);
  return cmul(x, w);
}

template <int L, int SGN, int S>
__device__ __forceinline__ void rdft(double2* v, const double2* __restrict__ rt) {
  if constexpr (L == 1) {
  } else if constexpr (L == 2 || L == 3 || L == 4 || L == 8) {
    dft_r<L, SGN>(v);
  } else {
    constexpr int A = rdft_first(L), B = L / A;
#pragma unroll
    for (int n2 = 0; n2 < B; ++n2) {
      double2 t[A];
#pragma unroll
      for (int a = 0; a < A; ++a) t[a] = v[n2 + B * a];
      dft_r<A, SGN>(t);
#pragma unroll
      for (int k1 = 0; k1 < A; ++k1) v[n2 + B * k1] = rtwiddle<L, SGN, S>(t[k1], n2 * k1, rt);
    }
    double2 o[L];
#pragma unroll
    for (int k1 = 0; k1 < A; ++k1) {
      rdft<B, SGN, S * A>(v + B * k1, rt);
#pragma unroll
      for (int k2 = 0; k2 < B; ++k2) o[k1 + A * k2] = v[B * k1 + k2];
    }
#pragma unroll
    for (int k = 0; k < L; ++k) v[k] = o[k];
  }
}

template <int N>
__device__ __forceinline__ int fft2_in_pos(int n) {
  constexpr int N2 = fft2_split(N).n2;
  return n + n / N2;
}
template <int N>
__device__ __forceinline__ int fft2_out_pos(int k) {
  constexpr int N1 = fft2_split(N).n1, N2 = fft2_split(N).n2;
  return (k % N1) * (N2 + 1) + k / N1;
}

// The two passes of the two-pass FFT for one task each.
template <int N, int SGN>
__device__ __forceinline__ void fft2_pass1(double2* seq, int b, const double2* rl, const double2* __restrict__ rg) {
  constexpr int N1 = fft2_split(N).n1, N2 = fft2_split(N).n2, RS = N2 + 1;
  double2* p = seq + b;
  double2 v[N1];
#pragma unroll
  for (int a = 0; a < N1; ++a) v[a] = p[a * RS];
  rdft<N1, SGN, N2>(v, rg);
#pragma unroll
  for (int k1 = 1; k1 < N1; ++k1) {
    double2 w = rl[k1 * N2 + b];
    if (SGN > 0) w = cconj(w);
    v[k1] = cmul(v[k1], w);
  }
#pragma unroll
  for (int k1 = 0; k1 < N1; ++k1) p[k1 * RS] = v[k1];
}

template <int N, int SGN>
__device__ __forceinline__ void fft2_pass2(double2* seq, int k1, const double2* __restrict__ rg) {
  constexpr int N1 = fft2_split(N).n1, N2 = fft2_split(N).n2, RS = N2 + 1;
  double2* p = seq + k1 * RS;
  double2 v[N2];
#pragma unroll
  for (int b = 0; b < N2; ++b) v[b] = p[b];
  rdft<N2, SGN, N1>(v, rg);
#pragma unroll
  for (int k2 = 0; k2 < N2; ++k2) p[k2] = v[k2];
}

// In-place two-pass FFT of B sequences (sequence s at lds + s*sstride).
// rl: the inter-pass twiddles in LDS, rl[k1*N2 + b] = exp(-2 pi i b k1/N)
// (load_tw2), so that the lanes of a pass-1 sweep (consecutive b) read
// consecutive cells; rg: the length-N root table in global memory
// (scalar-loaded constants of the register DFTs).  Call with
// the input visible to all threads (a barrier before); ends with a barrier.
template <int N, int SGN>
__device__ __forceinline__ void fft2p(double2* lds, int sstride, int B, const double2* rl,
                                      const double2* __restrict__ rg, int tid, int nthr) {
  constexpr int N1 = fft2_split(N).n1, N2 = fft2_split(N).n2;
  static_assert(N1 * N2 == N && N1 > 0, "unsupported two-pass length");
  for (int g = tid; g < B * N2; g += nthr) {
    const int s = g / N2, b = g - s * N2;
    fft2_pass1<N, SGN>(lds + s * sstride, b, rl, rg);
  }
  __syncthreads();
  for (int g = tid; g < B * N1; g += nthr) {
    const int s = g / N1, k1 = g - s * N1;
    fft2_pass2<N, SGN>(lds + s * sstride, k1, rg);
  }
  __syncthreads();
}

// fft2p without the twiddle table in LDS: a pass-1 task b reads w = rg[b]
// from the length-N root table (exp(-2 pi i b/N)) and forms the inter-pass
// twiddles as w^k1 = w^(k1-1) w in registers (w^k1 carries k1 - 1 roundings
// beyond the table's), which leaves the workgroup's LDS to the data alone.
template <int N, int SGN>
__device__ __forceinline__ void fft2_pass1_rec(double2* seq, int b, const double2* __restrict__ rg) {
  constexpr int N1 = fft2_split(N).n1, N2 = fft2_split(N).n2, RS = N2 + 1;
  double2* p = seq + b;
  const double2 wt = rg[b];
  double2 v[N1];
#pragma unroll
  for (int a = 0; a < N1; ++a) v[a] = p[a * RS];
  rdft<N1, SGN, N2>(v, rg);
  const double2 w1 = SGN > 0 ? cconj(wt) : wt;
  double2 w = w1;
#pragma unroll
  for (int k1 = 1; k1 < N1; ++k1) {
    if (k1 > 1) w = cmul(w, w1);
    v[k1] = cmul(v[k1], w);
  }
#pragma unroll
  for (int k1 = 0; k1 < N1; ++k1) p[k1 * RS] = v[k1];
}

template <int N, int SGN>
__device__ __forceinline__ void fft2p_rec(double2* lds, int sstride, int B, const double2* __restrict__ rg,
                                          int tid, int nthr) {
  constexpr int N1 = fft2_split(N).n1, N2 = fft2_split(N).n2;
  static_assert(N1 * N2 == N && N1 > 0, "unsupported two-pass length");
  for (int g = tid; g < B * N2; g += nthr) {
    const int s = g / N2, b = g - s * N2;
    fft2_pass1_rec<N, SGN>(lds + s * sstride, b, rg);
  }
  __syncthreads();
  for (int g = tid; g < B * N1; g += nthr) {
    const int s = g / N1, k1 = g - s * N1;
    fft2_pass2<N, SGN>(lds + s * sstride, k1, rg);
  }
  __syncthreads();
}

// Orders the LDS accesses of one wavefront: everything a lane wrote before
// is visible to every lane of the same wavefront after it.
__device__ __forceinline__ void wave_sync() {
  __builtin_amdgcn_fence(__ATOMIC_RELEASE, "wavefront");
  __builtin_amdgcn_wave_barrier();
  __builtin_amdgcn_fence(__ATOMIC_ACQUIRE, "wavefront");
}

// Two-pass FFT of B sequences owned by one wavefront (lane = 0..63); no
// workgroup barrier, so wavefronts of a workgroup run independently.
template <int N, int SGN>
__device__ __forceinline__ void fft2p_wave(double2* lds, int sstride, int B, const double2* rl,
                                           const double2* __restrict__ rg, int lane) {
  constexpr int N1 = fft2_split(N).n1, N2 = fft2_split(N).n2;
  static_assert(N1 * N2 == N && N1 > 0, "unsupported two-pass length");
  for (int g = lane; g < B * N2; g += 64) {
    const int s = g / N2, b = g - s * N2;
    fft2_pass1<N, SGN>(lds + s * sstride, b, rl, rg);
  }
  wave_sync();
  for (int g = lane; g < B * N1; g += 64) {
    const int s = g / N1, k1 = g - s * N1;
    fft2_pass2<N, SGN>(lds + s * sstride, k1, rg);
  }
  wave_sync();
}

// Position of input element n / output element k of a sequence: the two-pass
// layout for NC > 0, natural order for the runtime-planned transform (NC = 0).
template <int NC>
__device__ __forceinline__ int fft_in_pos(int n) {
  if constexpr (NC > 0) return fft2_in_pos<NC>(n);
  else return n;
}
template <int NC>
__device__ __forceinline__ int fft_out_pos(int k) {
  if constexpr (NC > 0) return fft2_out_pos<NC>(k);
  else return k;
}

// Transform of B sequences: two-pass for NC > 0, runtime plan otherwise.
template <int NC, int SGN, int MAXV>
__device__ __forceinline__ void fft_run(double2* lds, int sstride, int B, const FFTPlan& pl,
                                        const double2* __restrict__ twbase, const double2* rl, int tid,
                                        int nthr) {
  if constexpr (NC > 0) fft2p<NC, SGN>(lds, sstride, B, rl, twbase + pl.tw, tid, nthr);
  else fft_batch<SGN, MAXV>(lds, sstride, B, pl, twbase, tid, nthr);
}

// Copy the length-n root table into LDS (whole workgroup; caller synchronises).
__device__ __forceinline__ void load_roots(double2* dst, const double2* __restrict__ src, int n, int tid,
                                           int nthr) {
  for (int i = tid; i < n; i += nthr) dst[i] = src[i];
}

// The inter-pass twiddles of the two-pass length-NC transform in LDS:
// dst[k1*N2 + b] = roots[b*k1] (b*k1 < N always), NC cells; NC == 0 (the
// runtime-planned transform) copies nothing.  Whole workgroup; caller
// synchronises.
template <int NC>
__device__ __forceinline__ void load_tw2(double2* dst, const double2* __restrict__ roots, int tid, int nthr) {
  if constexpr (NC > 0) {
    // every load first (nthr >= 64), then the LDS writes: one memory round trip
    constexpr int N2 = fft2_split(NC).n2, IT = (NC + 63) / 64;
    double2 v[IT];
#pragma unroll
    for (int i = 0; i < IT; ++i) {
      const int e = tid + i * nthr, ec = e < NC ? e : 0;
      const int k1 = ec / N2, b = ec - k1 * N2;
      v[i] = roots[b * k1];
    }
    // the values are inputs of an (empty) asm statement here, so the loads
    // cannot be sunk into the conditional writes below
#pragma unroll
    for (int i = 0; i < IT; ++i) asm volatile("" ::"v"(v[i].x), "v"(v[i].y));
#pragma unroll
    for (int i = 0; i < IT; ++i) {
      const int e = tid + i * nthr;
      if (e < NC) dst[e] = v[i];
    }
  }
}

// Dispatch a launcher body over the compile-time lengths: FFT_DISPATCH(n, F)
// expands to F(n) for a supported n and F(0) otherwise.
#define FFT_DISPATCH(n, F)        \
  switch (n) {                    \
    case 16: F(16); break;        \
    case 24: F(24); break;        \
    case 32: F(32); break;        \
    case 36: F(36); break;        \
    case 48: F(48); break;        \
    case 64: F(64); break;        \
    case 72: F(72); break;        \
    case 96: F(96); break;        \
    case 144: F(144); break;      \
    case 192: F(192); break;      \
    case 288: F(288); break;      \
    default: F(0); break;         \
  }
