// Batched dealiased x-z transforms on the device (PFC/transform.f90 semantics).
//
//  k_zinv : spectral plane (nz, nx/2+1) -> z-inverse rows [nz2][kxpad]
//           (z padding with the Nyquist split and the x-Nyquist halving of
//           PFC/transform.f90:345-357, backward c2c of length nz2)
//  k_xc2r : rows [nz2][kxpad] -> physical real (nx2, nz2) via a half-length
//           complex FFT (c2r of PFC/transform.f90:396-405)
//  k_xr2c : physical real rows -> x spectra kept for kx < nx/2+1, optionally
//           combined as F(fy) + i kx F(fx) (PFC/transform.f90:265-276,581-617)
//  k_zfwd : x spectra -> forward c2c of length nz2, truncation with the
//           Nyquist fold, 1/(nx2 nz2) scaling, doubling of the last kx
//           column and optional + i kz * (second field)
//           (PFC/transform.f90:316-333, 641-687)
#include "chan_internal.h"

#define ZT_KX 8

__device__ __forceinline__ long map_plane(const PlaneMap& pm, unsigned by) {
  return (long)(by / pm.cnt) * pm.stride + pm.off + by % pm.cnt;
}
__device__ __forceinline__ long map_plane(const PlaneMap& pm) { return map_plane(pm, blockIdx.y); }

// Source of padded row element t (0 <= t < nz2) of a spectral column of nz
// stored kz modes, for a load with a valid address: index, scale and
// conjugation (the Nyquist split: both halves get half of the Nyquist mode;
// zero between the stored halves), and the x-Nyquist halving.
struct PadSrc { int idx; double scale; bool conj; };
__device__ __forceinline__ PadSrc zpad_src(int t, int nz, int nz2, bool pad, bool halve) {
  PadSrc r{0, 1.0, false};
  if (!pad) {
    r.idx = t < nz ? t : nz - 1;
    r.scale = t < nz ? 1.0 : 0.0;
  } else {
    const int h = nz / 2;
    if (t < h) r.idx = t;
    else if (t == h) { r.idx = h; r.scale = 0.5; }
    else if (t == nz2 - h) { r.idx = h; r.scale = 0.5; r.conj = true; }
    else if (t > nz2 - h) r.idx = t - (nz2 - nz);
    else r.scale = 0.0;
    if (halve) r.scale *= 0.5;
  }
  return r;
}
// the loaded value with its scale and conjugation; zero where the element is
// padding or outside the workgroup's columns
__device__ __forceinline__ double2 pad_apply(double2 v, const PadSrc& p, bool valid) {
  if (p.conj) v = cconj(v);
  v = cscale(p.scale, v);
  return (valid && p.scale != 0.0) ? v : make_double2(0.0, 0.0);
}

// Row t and column s of element e of an 8-column tile, with each 8-lane group
// on 4 columns x 2 rows (e in whole groups of 16 covers 2 rows x 8 columns).
// The group's 16-byte LDS writes to s*ss + pos(t) then fall in distinct banks
// for ss = 2 (mod 8) (ss = 306 at nz2 = 288), where the 8 columns of one row
// meet in pairs (measured: LDS bank-conflict cycles of the z-forward 3.6e7 ->
// 1.6e7 per step); every 16 lanes still read two whole 128-byte rows of the
// [z][kx] array.  Needs an even row count.
__device__ __forceinline__ void tile8_ts(int e, int& t, int& s) {
  const int r = e & 15;
  t = 2 * (e >> 4) + ((r >> 2) & 1);
  s = (r & 3) | ((r >> 3) << 2);
}

// (bx, by): the workgroup's kx group and plane row (blockIdx, or a work-list
// entry of the batched form)
template <int NC>
__device__ __forceinline__ void zinv_body(const XZGeom& g, const double2* __restrict__ tw,
                                          const double2* __restrict__ in, long in_ps,
                                          double2* __restrict__ out, long out_ps, const PlaneMap& pm,
                                          int bx, unsigned by) {
  extern __shared__ double2 lds[];
  const int tid = threadIdx.x, nthr = blockDim.x;
  const int kx0 = bx * ZT_KX;
  const long P = map_plane(pm, by);
  const int nz = g.nz, nz2 = NC > 0 ? NC : g.nz2, ss = fft_seq_cells(nz2);
  double2* rl = lds + ZT_KX * ss;
  if constexpr (NC == 0) load_tw2<NC>(rl, tw + g.pz.tw, tid, nthr);
  const int ncol = min(ZT_KX, g.nxh - kx0);
  const double2* src = in + P * in_ps;
  const bool pad = (nz2 > nz) && !g.small;
  const int h = nz / 2;
  if constexpr (NC > 0) {
    // all loads of the thread first (one memory round trip), then the LDS writes
    constexpr int IT = (ZT_KX * NC + 255) / 256;
    double2 v[IT];
#pragma unroll
    for (int i = 0; i < IT; ++i) {
      const int e = tid + i * nthr;
      const int s = e / NC, t = e - s * NC;
      const PadSrc ps = zpad_src(t, nz, NC, pad, g.nx2 > g.nx && kx0 + s == g.nxh - 1);
      const int sc = s < ncol ? s : ncol - 1;
      v[i] = src[(long)(kx0 + sc) * nz + ps.idx];
      v[i] = pad_apply(v[i], ps, e < ZT_KX * NC && s < ncol);
    }
    // the inter-pass twiddles are formed in registers (fft2p_rec): with no
    // twiddle table, the workgroup's 38 KB of LDS let 4 workgroups share a CU
    __builtin_amdgcn_sched_barrier(0);  // keep the loads ahead of the first write
#pragma unroll
    for (int i = 0; i < IT; ++i) {
      const int e = tid + i * nthr;
      const int s = e / NC, t = e - s * NC;
      if (e < ZT_KX * NC) lds[s * ss + fft_in_pos<NC>(t)] = v[i];
    }
  } else {
    for (int e = tid; e < ZT_KX * nz2; e += nthr) {
      int s = e / nz2, t = e - s * nz2;
      double2 v = make_double2(0.0, 0.0);
      if (s < ncol) {
        const double2* col = src + (long)(kx0 + s) * nz;
        if (pad) {
          if (t < h) v = col[t];
          else if (t == h) v = cscale(0.5, col[h]);
          else if (t == nz2 - h) v = cscale(0.5, cconj(col[h]));
          else if (t > nz2 - h) v = col[t - (nz2 - nz)];
          if (g.nx2 > g.nx && kx0 + s == g.nxh - 1) v = cscale(0.5, v);
        } else {
          v = col[t];
        }
      }
      lds[s * ss + fft_in_pos<NC>(t)] = v;
    }
  }
  __syncthreads();
  if constexpr (NC > 0) fft2p_rec<NC, 1>(lds, ss, ncol, tw + g.pz.tw, tid, nthr);  // the valid columns only
  else fft_run<NC, 1, FFT_MAXV>(lds, ss, ZT_KX, g.pz, tw, rl, tid, nthr);
  double2* dst = out + P * out_ps;
  for (int e = tid; e < nz2 * ZT_KX; e += nthr) {
    int t = e / ZT_KX, s = e - t * ZT_KX;
    if (s < ncol) dst[(long)t * g.kxpad + kx0 + s] = lds[s * ss + fft_out_pos<NC>(t)];
  }
}

template <int NC>
__global__ __launch_bounds__(256) void k_zinv(XZGeom g, const double2* __restrict__ tw,
                                              const double2* __restrict__ in, long in_ps,
                                              double2* __restrict__ out, long out_ps, PlaneMap pm) {
  zinv_body<NC>(g, tw, in, in_ps, out, out_ps, pm, blockIdx.x, blockIdx.y);
}

// Batched form: workgroup i runs work-list entry it[i] of member it[i].m.
template <int NC>
__global__ __launch_bounds__(256) void k_zinv_b(const ZBatchArgs* __restrict__ args,
                                                const BatchItem* __restrict__ it) {
  const BatchItem w = it[blockIdx.x];
  const ZBatchArgs& a = args[w.m];
  zinv_body<NC>(a.g, a.tw, a.in, a.in_ps, a.out, a.out_ps, a.pm, w.bx, w.by);
}

// Hermitian half-length packing for an unnormalised c2r of length nx2 = 2M:
// Zc[n] = (Y[n] + conj(Y[M-n])) + i w^n (Y[n] - conj(Y[M-n])), w = exp(2 pi i/nx2),
// with Im Y[0] and Im Y[M] dropped (FFTW c2r semantics).
__device__ __forceinline__ double2 c2r_pack(const double2* __restrict__ row, int nxh, int M, int n,
                                            double2 wn_conjroot) {
  auto Y = [&](int k) -> double2 {
    if (k >= nxh) return make_double2(0.0, 0.0);
    double2 v = row[k];
    if (k == 0 || k == M) v.y = 0.0;
    return v;
  };
  double2 a = Y(n), b = cconj(Y(M - n));
  double2 s = cadd(a, b), d = csub(a, b);
  double2 wd = cmul(wn_conjroot, d);  // w^n * d
  return make_double2(s.x - wd.y, s.y + wd.x);
}

// Post-processing of a half-length forward FFT Z (length M) into the real
// transform X_k = E_k + w^-k O_k, E = (Z_k + conj Z_{M-k})/2, O = -i (Z_k - conj Z_{M-k})/2.
__device__ __forceinline__ double2 r2c_post(const double2* __restrict__ Zs, int M, int k, double2 wk) {
  double2 a = Zs[k % M], b = cconj(Zs[(M - k) % M]);
  double2 E = make_double2(0.5 * (a.x + b.x), 0.5 * (a.y + b.y));
  double2 d = csub(a, b);
  double2 O = make_double2(0.5 * d.y, -0.5 * d.x);
  return cadd(E, cmul(wk, O));
}

template <int NC>
__global__ __launch_bounds__(512) void k_xc2r(XZGeom g, const double2* __restrict__ tw,
                                               const double2* __restrict__ Z, long z_ps,
                                               double* __restrict__ out, long out_ps, int rows) {
  extern __shared__ double2 lds[];
  const int tid = threadIdx.x, nthr = blockDim.x;
  const int M = NC > 0 ? NC : g.nx2 / 2, ss = NC > 0 ? fft_seq_cells(NC) : M + 1;
  const long P = blockIdx.y;
  const int z0 = blockIdx.x * rows;
  const int nr = min(rows, g.nz2 - z0);
  const double2* rootx = tw + g.twx;
  double2* rl = lds + rows * ss;
  load_tw2<NC>(rl, tw + g.px.tw, tid, nthr);
  for (int e = tid; e < rows * M; e += nthr) {
    int r = e / M, n = e - r * M;
    double2 v = make_double2(0.0, 0.0);
    if (r < nr) {
      const double2* row = Z + P * z_ps + (long)(z0 + r) * g.kxpad;
      v = c2r_pack(row, g.nxh, M, n, cconj(rootx[n]));
    }
    lds[r * ss + fft_in_pos<NC>(n)] = v;
  }
  __syncthreads();
  fft_run<NC, 1, FFT_MAXV>(lds, ss, rows, g.px, tw, rl, tid, nthr);
  for (int e = tid; e < nr * M; e += nthr) {
    int r = e / M, n = e - r * M;
    double2 v = lds[r * ss + fft_out_pos<NC>(n)];
    double* o = out + P * out_ps + (long)(z0 + r) * g.nx2 + 2 * n;
    *reinterpret_cast<double2*>(o) = v;
  }
}

template <int NC>
__global__ __launch_bounds__(512) void k_xr2c(XZGeom g, const double2* __restrict__ tw,
                                               const double* __restrict__ fy,
                                               const double* __restrict__ fx,
                                               const double* __restrict__ kxv, long in_ps,
                                               double2* __restrict__ out, long out_ps, int rows) {
  extern __shared__ double2 lds[];
  const int tid = threadIdx.x, nthr = blockDim.x;
  const int M = NC > 0 ? NC : g.nx2 / 2, ss = NC > 0 ? fft_seq_cells(NC) : M + 1;
  const long P = blockIdx.y;
  const int z0 = blockIdx.x * rows;
  const int nr = min(rows, g.nz2 - z0);
  const int nf = fx ? 2 : 1;
  const double2* rootx = tw + g.twx;
  double2* rl = lds + nf * rows * ss;
  load_tw2<NC>(rl, tw + g.px.tw, tid, nthr);
  for (int e = tid; e < nf * rows * M; e += nthr) {
    int q = e / M, n = e - q * M;
    int f = q / rows, r = q - f * rows;
    double2 v = make_double2(0.0, 0.0);
    if (r < nr) {
      const double* src = (f == 0 ? fy : fx) + P * in_ps + (long)(z0 + r) * g.nx2;
      v = *reinterpret_cast<const double2*>(src + 2 * n);
    }
    lds[q * ss + fft_in_pos<NC>(n)] = v;
  }
  __syncthreads();
  fft_run<NC, -1, FFT_MAXV>(lds, ss, nf * rows, g.px, tw, rl, tid, nthr);
  // X_k = E_k + w^-k O_k from the half-length transform (r2c_post), read
  // through the two-pass output order
  auto post = [&](const double2* Zs, int k, double2 wk) {
    double2 a = Zs[fft_out_pos<NC>(k % M)], b = cconj(Zs[fft_out_pos<NC>((M - k) % M)]);
    double2 E = make_double2(0.5 * (a.x + b.x), 0.5 * (a.y + b.y));
    double2 d = csub(a, b);
    double2 O = make_double2(0.5 * d.y, -0.5 * d.x);
    return cadd(E, cmul(wk, O));
  };
  for (int e = tid; e < nr * g.nxh; e += nthr) {
    int r = e / g.nxh, k = e - r * g.nxh;
    double2 wk = rootx[k];
    double2 X = post(lds + r * ss, k, wk);
    if (fx) {
      double2 Xx = post(lds + (rows + r) * ss, k, wk);
      double kk = kxv[k];
      X = make_double2(X.x - kk * Xx.y, X.y + kk * Xx.x);
    }
    out[P * out_ps + (long)(z0 + r) * g.kxpad + k] = X;
  }
}

template <int NC>
__global__ __launch_bounds__(512) void k_zfwd(XZGeom g, const double2* __restrict__ tw,
                                              const double2* __restrict__ A,
                                              const double2* __restrict__ B, long ab_ps,
                                              const double* __restrict__ kzv,
                                              double2* __restrict__ out, long out_ps,
                                              int nl_zero_ny, PlaneMap pm) {
  extern __shared__ double2 lds[];
  const int tid = threadIdx.x, nthr = blockDim.x;
  const int kx0 = blockIdx.x * ZT_KX;
  const long P = map_plane(pm);
  const int nz = g.nz, nz2 = NC > 0 ? NC : g.nz2, ss = fft_seq_cells(nz2);
  const int ncol = min(ZT_KX, g.nxh - kx0);
  double2* dst = out + P * out_ps;
  if (nl_zero_ny > 0) {
    // NL planes that the reference leaves zero: plane 1, plane ny+1, and v at
    // plane ny (PFC/pass1.f90:62,96-99).
    const int ny = nl_zero_ny;
    int c = (int)(P / (ny + 1)), j = (int)(P % (ny + 1));
    if (j == 0 || j == ny || (c == 1 && j == ny - 1)) {
      for (int e = tid; e < ncol * nz; e += nthr) {
        int s = e / nz, q = e - s * nz;
        dst[(long)(kx0 + s) * nz + q] = make_double2(0.0, 0.0);
      }
      return;
    }
  }
  const int nf = B ? 2 : 1;
  double2* rl = lds + nf * ZT_KX * ss;
  load_tw2<NC>(rl, tw + g.pz.tw, tid, nthr);
  for (int e = tid; e < nf * nz2 * ZT_KX; e += nthr) {
    int f = e / (nz2 * ZT_KX);
    int e2 = e - f * nz2 * ZT_KX;
    int t = e2 / ZT_KX, s = e2 - t * ZT_KX;
    double2 v = make_double2(0.0, 0.0);
    if (s < ncol) v = (f == 0 ? A : B)[P * ab_ps + (long)t * g.kxpad + kx0 + s];
    lds[(f * ZT_KX + s) * ss + fft_in_pos<NC>(t)] = v;
  }
  __syncthreads();
  fft_run<NC, -1, FFT_MAXV>(lds, ss, nf * ZT_KX, g.pz, tw, rl, tid, nthr);
  const int h = nz / 2, off = nz2 - nz;
  auto at = [&](const double2* a, int q) { return a[fft_out_pos<NC>(q)]; };
  for (int e = tid; e < ncol * nz; e += nthr) {
    int s = e / nz, q = e - s * nz;
    const double2* a = lds + s * ss;
    double2 val;
    if (g.small) {
      double sc = 1.0 / (double)(g.nx * g.nz);
      val = cscale(sc, at(a, q));
    } else {
      const double scale = 1.0 / (double)(g.nx2 * g.nz2);
      const double f = (kx0 + s == g.nxh - 1) ? 2.0 * scale : scale;
      double2 fa;
      if (q < h) fa = at(a, q);
      else if (q == h) fa = cadd(at(a, h), cconj(at(a, off + h)));
      else fa = at(a, q + off);
      val = cscale(f, fa);
      if (B) {
        const double2* b = lds + (ZT_KX + s) * ss;
        double2 fb;
        if (q < h) fb = at(b, q);
        else if (q == h) fb = cadd(at(b, h), cconj(at(b, off + h)));
        else fb = at(b, q + off);
        fb = cscale(f, fb);
        double kz = kzv[q];
        val = make_double2(val.x - kz * fb.y, val.y + kz * fb.x);
      }
    }
    dst[(long)(kx0 + s) * nz + q] = val;
  }
}

// Same result as k_zfwd with half the LDS: A and B are transformed one after
// the other in one set of 8 sequences.  B's loads are issued into registers
// before A's transform so that their latency overlaps it; A's folded and
// scaled values wait in registers for B's.  Used for the compile-time
// lengths with a second field (the nonlinear pass).
// Compile-time variant of k_zinv with KX columns per workgroup of 32*KX threads.
template <int NC, int KX>
__global__ __launch_bounds__(32 * KX) void k_zinv2(XZGeom g, const double2* __restrict__ tw,
                                                   const double2* __restrict__ in, long in_ps,
                                                   double2* __restrict__ out, long out_ps, PlaneMap pm) {
  extern __shared__ double2 lds[];
  constexpr int NT = 32 * KX, NZ2 = NC, SS = fft_seq_cells(NC);
  const int tid = threadIdx.x;
  const int kx0 = blockIdx.x * KX;
  const long P = map_plane(pm);
  const int nz = g.nz;
  double2* rl = lds + KX * SS;
  const int ncol = min(KX, g.nxh - kx0);
  const double2* src = in + P * in_ps;
  const bool pad = (NZ2 > nz) && !g.small;
  const bool halve = g.nx2 > g.nx;
  constexpr int IT = (KX * NZ2 + NT - 1) / NT;
  double2 v[IT];
#pragma unroll
  for (int i = 0; i < IT; ++i) {
    const int e = tid + i * NT;
    const int s = e / NZ2, t = e - s * NZ2;
    const PadSrc ps = zpad_src(t, nz, NZ2, pad, halve && kx0 + s == g.nxh - 1);
    const int sc = s < ncol ? s : ncol - 1;
    v[i] = src[(long)(kx0 + sc) * nz + ps.idx];
    v[i] = pad_apply(v[i], ps, e < KX * NZ2 && s < ncol);
  }
  __builtin_amdgcn_sched_barrier(0);  // keep the loads ahead of the first write
  load_tw2<NC>(rl, tw + g.pz.tw, tid, NT);  // the twiddles' loads behind the data's
#pragma unroll
  for (int i = 0; i < IT; ++i) {
    const int e = tid + i * NT;
    const int s = e / NZ2, t = e - s * NZ2;
    if (e < KX * NZ2) lds[s * SS + fft2_in_pos<NC>(t)] = v[i];
  }
  __syncthreads();
  fft2p<NC, 1>(lds, SS, ncol, rl, tw + g.pz.tw, tid, NT);  // the valid columns only
  double2* dst = out + P * out_ps;
  for (int e = tid; e < NZ2 * KX; e += NT) {
    const int t = e / KX, s = e - t * KX;
    if (s < ncol) dst[(long)t * g.kxpad + kx0 + s] = lds[s * SS + fft2_out_pos<NC>(t)];
  }
}

template <int NC, int KX>
__device__ __forceinline__ void zfwd2_body(const XZGeom& g, const double2* __restrict__ tw,
                                           const double2* __restrict__ A,
                                           const double2* __restrict__ B, long ab_ps,
                                           const double* __restrict__ kzv,
                                           double2* __restrict__ out, long out_ps,
                                           int nl_zero_ny, const PlaneMap& pm, int bx, unsigned by) {
  extern __shared__ double2 lds[];
  constexpr int NZ2 = NC, SS = fft_seq_cells(NC);
  constexpr int NT = 32 * KX;
  constexpr int NLD = (NZ2 * KX + NT - 1) / NT;       // loads per thread per field
  const int tid = threadIdx.x;
  const int kx0 = bx * KX;
  const long P = map_plane(pm, by);
  const int nz = g.nz;
  const int ncol = min(KX, g.nxh - kx0);
  double2* dst = out + P * out_ps;
  if (nl_zero_ny > 0) {
    // NL planes that the reference leaves zero: plane 1, plane ny+1, and v at
    // plane ny (PFC/pass1.f90:62,96-99).
    const int ny = nl_zero_ny;
    const int c = (int)(P / (ny + 1)), j = (int)(P % (ny + 1));
    if (j == 0 || j == ny || (c == 1 && j == ny - 1)) {
      for (int e = tid; e < ncol * nz; e += NT) {
        const int s = e / nz, q = e - s * nz;
        dst[(long)(kx0 + s) * nz + q] = make_double2(0.0, 0.0);
      }
      return;
    }
  }
  double2* rl = lds + KX * SS;
  const double2* Ap = A + P * ab_ps;
  const double2* Bp = B + P * ab_ps;
  // all loads of the thread first (one memory round trip), then the LDS writes
  double2 ra[NLD], rb[NLD];
#pragma unroll
  for (int k = 0; k < NLD; ++k) {
    const int e = tid + NT * k;
    int t = e / KX, s = e - t * KX;
    if constexpr (KX == 8) tile8_ts(e, t, s);
    const bool ok = e < NZ2 * KX && s < ncol;
    const long o = (long)(t < NZ2 ? t : NZ2 - 1) * g.kxpad + kx0 + (s < ncol ? s : ncol - 1);
    const double2 a = Ap[o], b = Bp[o];
    ra[k] = ok ? a : make_double2(0.0, 0.0);
    rb[k] = ok ? b : make_double2(0.0, 0.0);
  }
  __builtin_amdgcn_sched_barrier(0);  // keep the loads ahead of the first write
  load_tw2<NC>(rl, tw + g.pz.tw, tid, NT);  // the twiddles' loads behind the data's
#pragma unroll
  for (int k = 0; k < NLD; ++k) {
    const int e = tid + NT * k;
    int t = e / KX, s = e - t * KX;
    if constexpr (KX == 8) tile8_ts(e, t, s);
    if (e < NZ2 * KX) lds[s * SS + fft2_in_pos<NC>(t)] = ra[k];
  }
  __syncthreads();
  fft2p<NC, -1>(lds, SS, ncol, rl, tw + g.pz.tw, tid, NT);  // the valid columns only
  const int h = nz / 2, off = NZ2 - nz;
  const double scale = 1.0 / (double)(g.nx2 * NZ2);
  auto fold = [&](const double2* a, int q) {
    if (q < h) return a[fft2_out_pos<NC>(q)];
    if (q == h) return cadd(a[fft2_out_pos<NC>(h)], cconj(a[fft2_out_pos<NC>(off + h)]));
    return a[fft2_out_pos<NC>(q + off)];
  };
  constexpr int NOUT = (KX * 192 + NT - 1) / NT;     // outputs per thread for nz <= 192
  double2 acc[NOUT];
#pragma unroll
  for (int k = 0; k < NOUT; ++k) {
    const int e = tid + NT * k;
    acc[k] = make_double2(0.0, 0.0);
    if (e < ncol * nz) {
      const int s = e / nz, q = e - s * nz;
      const double f = (kx0 + s == g.nxh - 1) ? 2.0 * scale : scale;
      acc[k] = cscale(f, fold(lds + s * SS, q));
    }
  }
  __syncthreads();
#pragma unroll
  for (int k = 0; k < NLD; ++k) {
    const int e = tid + NT * k;
    int t = e / KX, s = e - t * KX;
    if constexpr (KX == 8) tile8_ts(e, t, s);
    if (e < NZ2 * KX) lds[s * SS + fft2_in_pos<NC>(t)] = rb[k];
  }
  __syncthreads();
  fft2p<NC, -1>(lds, SS, ncol, rl, tw + g.pz.tw, tid, NT);  // the valid columns only
#pragma unroll
  for (int k = 0; k < NOUT; ++k) {
    const int e = tid + NT * k;
    if (e < ncol * nz) {
      const int s = e / nz, q = e - s * nz;
      const double f = (kx0 + s == g.nxh - 1) ? 2.0 * scale : scale;
      const double2 fb = cscale(f, fold(lds + s * SS, q));
      const double kz = kzv[q];
      dst[(long)(kx0 + s) * nz + q] = make_double2(acc[k].x - kz * fb.y, acc[k].y + kz * fb.x);
    }
  }
}

template <int NC, int KX>
__global__ __launch_bounds__(32 * KX) void k_zfwd2(XZGeom g, const double2* __restrict__ tw,
                                               const double2* __restrict__ A,
                                               const double2* __restrict__ B, long ab_ps,
                                               const double* __restrict__ kzv,
                                               double2* __restrict__ out, long out_ps,
                                               int nl_zero_ny, PlaneMap pm) {
  zfwd2_body<NC, KX>(g, tw, A, B, ab_ps, kzv, out, out_ps, nl_zero_ny, pm, blockIdx.x, blockIdx.y);
}

template <int NC, int KX>
__global__ __launch_bounds__(32 * KX) void k_zfwd2_b(const ZBatchArgs* __restrict__ args,
                                                 const BatchItem* __restrict__ it) {
  const BatchItem w = it[blockIdx.x];
  const ZBatchArgs& a = args[w.m];
  zfwd2_body<NC, KX>(a.g, a.tw, a.in, a.B, a.in_ps, a.kz, a.out, a.out_ps, a.nl_zero_ny, a.pm, w.bx, w.by);
}

// ---------------------------------------------------------------------------
// host launchers
// ---------------------------------------------------------------------------
static void check_stage_fit(const FFTPlan& p, int nseq, int nthr) {
  for (int i = 0; i < p.nst; ++i) {
    int R = p.rad[i];
    long nbf = (long)nseq * (p.n / R);
    long per = (nbf + nthr - 1) / nthr;
    if (per * R > FFT_MAXV) throw chan_error("FFT batch too large for workgroup");
  }
}

void launch_zinv(const XZGeom& g, const double2* tw, const double2* in, long in_ps, double2* out,
                 long out_ps, int nplanes, hipStream_t s, int kx, int pcnt, int pstride, int poff) {
  const PlaneMap pm = pcnt > 0 ? PlaneMap{pcnt, pstride, poff} : PlaneMap{nplanes, 0, 0};
  const bool ct = fft2_split(g.nz2).n1 > 0;
  if (ct && (kx == 4 || kx == 16)) {
    dim3 gk((g.nxh + kx - 1) / kx, nplanes);
    const size_t lk = sizeof(double2) * (kx * fft_seq_cells(g.nz2) + g.nz2);
#define ZINV2_CASE(N)                                                                                      \
  if constexpr (N > 0) {                                                                                   \
    if (kx == 4) hipLaunchKernelGGL((k_zinv2<N, 4>), gk, dim3(128), lk, s, g, tw, in, in_ps, out, out_ps, pm); \
    else hipLaunchKernelGGL((k_zinv2<N, 16>), gk, dim3(512), lk, s, g, tw, in, in_ps, out, out_ps, pm);         \
  }
    FFT_DISPATCH(g.nz2, ZINV2_CASE)
#undef ZINV2_CASE
    LAUNCH_CHECK("k_zinv2");
    return;
  }
  dim3 grid((g.nxh + ZT_KX - 1) / ZT_KX, nplanes);
  const int nthr = 256;
  if (!ct) check_stage_fit(g.pz, ZT_KX, nthr);
  size_t lds = sizeof(double2) * ZT_KX * fft_seq_cells(g.nz2);  // no twiddle table (k_zinv)
#define ZINV_CASE(N) \
  hipLaunchKernelGGL(k_zinv<N>, grid, dim3(nthr), lds, s, g, tw, in, in_ps, out, out_ps, pm)
  FFT_DISPATCH(g.nz2, ZINV_CASE)
#undef ZINV_CASE
  LAUNCH_CHECK("k_zinv");
}

// Batched z-transforms: one launch runs the work lists of several channels
// of the same transform length (the batched ensemble).  Only the
// compile-time two-pass lengths are batched.
bool zt_batchable(int nz2) { return fft2_split(nz2).n1 > 0; }

void launch_zinv_batched(int nz2, const ZBatchArgs* args, const BatchItem* items, int nitems, hipStream_t s) {
  if (nitems <= 0) return;
  if (!zt_batchable(nz2)) throw chan_error("batched z-inverse: length " + std::to_string(nz2) + " has no compile-time plan");
  const size_t lds = sizeof(double2) * ZT_KX * fft_seq_cells(nz2);
#define ZINVB_CASE(N) \
  if constexpr (N > 0) hipLaunchKernelGGL(k_zinv_b<N>, dim3(nitems), dim3(256), lds, s, args, items)
  FFT_DISPATCH(nz2, ZINVB_CASE)
#undef ZINVB_CASE
  LAUNCH_CHECK("k_zinv_b");
}

void launch_zfwd_batched(int nz2, const ZBatchArgs* args, const BatchItem* items, int nitems, hipStream_t s) {
  if (nitems <= 0) return;
  if (!zt_batchable(nz2)) throw chan_error("batched z-forward: length " + std::to_string(nz2) + " has no compile-time plan");
  const size_t lds = sizeof(double2) * (8 * fft_seq_cells(nz2) + nz2);
#define ZFWDB_CASE(N) \
  if constexpr (N > 0) hipLaunchKernelGGL((k_zfwd2_b<N, 8>), dim3(nitems), dim3(256), lds, s, args, items)
  FFT_DISPATCH(nz2, ZFWDB_CASE)
#undef ZFWDB_CASE
  LAUNCH_CHECK("k_zfwd2_b");
}

// Rows per workgroup of the x c2r / r2c kernels: the largest divisor of nz2
// whose nf * rows * M points fit 3 per thread of a 512-thread workgroup
// (1 per thread for the runtime-planned transform, whose staging is bounded),
// so that the transform passes of short rows still give most threads a task.
static int rows_for(int M, int nz2, int nf, bool ct) {
  const int cap = (ct ? 1536 : 512) / nf;
  int best = 1;
  for (int d = 1; d <= nz2; ++d)
    if (nz2 % d == 0 && d * M <= cap) best = d;
  return best;
}

void launch_xc2r(const XZGeom& g, const double2* tw, const double2* Z, long z_ps, double* out,
                 long out_ps, int nplanes, hipStream_t s) {
  const int M = g.nx2 / 2;
  const bool ct = fft2_split(M).n1 > 0;
  int rows = rows_for(M, g.nz2, 1, ct);
  int nthr = std::min(512, std::max(64, ((rows * M + 63) / 64) * 64));
  if (!ct) check_stage_fit(g.px, rows, nthr);
  dim3 grid(g.nz2 / rows, nplanes);
  size_t lds = sizeof(double2) * (ct ? rows * fft_seq_cells(M) + M : rows * (M + 1));
#define XC2R_CASE(N) hipLaunchKernelGGL(k_xc2r<N>, grid, dim3(nthr), lds, s, g, tw, Z, z_ps, out, out_ps, rows)
  FFT_DISPATCH(M, XC2R_CASE)
#undef XC2R_CASE
  LAUNCH_CHECK("k_xc2r");
}

void launch_xr2c(const XZGeom& g, const double2* tw, const double* fy, const double* fx,
                 const double* kx, long in_ps, double2* out, long out_ps, int nplanes, hipStream_t s) {
  const int M = g.nx2 / 2;
  const int nf = fx ? 2 : 1;
  const bool ct = fft2_split(M).n1 > 0;
  int rows = rows_for(M, g.nz2, nf, ct);
  int nthr = std::min(512, std::max(64, ((nf * rows * M + 63) / 64) * 64));
  if (!ct) check_stage_fit(g.px, nf * rows, nthr);
  dim3 grid(g.nz2 / rows, nplanes);
  size_t lds = sizeof(double2) * (ct ? nf * rows * fft_seq_cells(M) + M : nf * rows * (M + 1));
#define XR2C_CASE(N) \
  hipLaunchKernelGGL(k_xr2c<N>, grid, dim3(nthr), lds, s, g, tw, fy, fx, kx, in_ps, out, out_ps, rows)
  FFT_DISPATCH(M, XR2C_CASE)
#undef XR2C_CASE
  LAUNCH_CHECK("k_xr2c");
}

void launch_zfwd(const XZGeom& g, const double2* tw, const double2* A, const double2* B, long ab_ps,
                 const double* kz, double2* out, long out_ps, int nplanes, int nl_zero_ny,
                 hipStream_t s, int v2, int kx, int pcnt, int pstride, int poff) {
  const PlaneMap pm = pcnt > 0 ? PlaneMap{pcnt, pstride, poff} : PlaneMap{nplanes, 0, 0};
  dim3 grid((g.nxh + ZT_KX - 1) / ZT_KX, nplanes);
  const int nf = B ? 2 : 1;
  const int nthr = 256 * nf;
  const bool ct = fft2_split(g.nz2).n1 > 0;
  if (v2 && ct && B && !g.small && g.nz <= 192) {
    const int k2 = (kx == 4) ? 4 : 8;
    dim3 g2((g.nxh + k2 - 1) / k2, nplanes);
    size_t l2 = sizeof(double2) * (k2 * fft_seq_cells(g.nz2) + g.nz2);
#define ZFWD2_CASE(N)                                                                                       \
  if constexpr (N > 0) {                                                                                    \
    if (k2 == 4) hipLaunchKernelGGL((k_zfwd2<N, 4>), g2, dim3(128), l2, s, g, tw, A, B, ab_ps, kz, out, out_ps, \
                                    nl_zero_ny, pm);                                                        \
    else hipLaunchKernelGGL((k_zfwd2<N, 8>), g2, dim3(256), l2, s, g, tw, A, B, ab_ps, kz, out, out_ps,       \
                            nl_zero_ny, pm);                                                                \
  }
    FFT_DISPATCH(g.nz2, ZFWD2_CASE)
#undef ZFWD2_CASE
    LAUNCH_CHECK("k_zfwd2");
    return;
  }
  if (!ct) check_stage_fit(g.pz, nf * ZT_KX, nthr);
  size_t lds = sizeof(double2) * (nf * ZT_KX * fft_seq_cells(g.nz2) + (ct ? g.nz2 : 0));
#define ZFWD_CASE(N)                                                                              \
  hipLaunchKernelGGL(k_zfwd<N>, grid, dim3(nthr), lds, s, g, tw, A, B, ab_ps, kz, out, out_ps, \
                     nl_zero_ny, pm)
  FFT_DISPATCH(g.nz2, ZFWD_CASE)
#undef ZFWD_CASE
  LAUNCH_CHECK("k_zfwd");
}
