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
