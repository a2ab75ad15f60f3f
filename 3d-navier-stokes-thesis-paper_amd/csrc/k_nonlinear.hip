// Physical-space pass of the nonlinear term (PFC/pass1.f90:26-252) and of
// its linearized variant (PFC/pass1_linearized.f90:135-509), fused per z-row.
//
// A workgroup owns `rz` physical z-rows and a chunk of y-planes.  Thread
// (r, n) owns the two physical points x = 2n, 2n+1 of row r, which are the
// real and imaginary parts of point n of the half-length complex FFT used
// for every x transform.  Walking up the chunk, it keeps in registers the
// velocity at planes k and k+1 and the wall-normal fluxes at y(k-1), yh(k),
// exactly the sliding window of the reference (u = up at :81), so each plane
// is inverse transformed once per chunk plus one halo plane.
//
// Per plane k the workgroup
//   1. x-c2r's the z-inverse rows of u, v, w at plane k+1 (three half-length
//      backward FFTs per row),
//   2. forms the flux products and y-differences (:138-195),
//   3. forward x-transforms the 8 distinct real fields (u^2, d/dy flux_u,
//      w u, d/dy flux_w, w^2, the two interpolated v fluxes, d/dy flux_v),
//   4. combines A = F(fy) + i kx F(fx) and B = F(fz) for each component
//      (PFC/transform.f90:613) and writes them for the z-forward pass.
// At rk_step == 1 it also accumulates the per-plane CFL maxima (:202-217).
#include "chan_internal.h"

__device__ __forceinline__ double2 c2r_pack_nl(const double2* __restrict__ row, int nxh, int M,
                                               int n, double2 wconj) {
  double2 a = make_double2(0.0, 0.0), b = make_double2(0.0, 0.0);
  if (n < nxh) { a = row[n]; if (n == 0 || n == M) a.y = 0.0; }
  int m = M - n;
  if (m < nxh) { b = row[m]; if (m == 0 || m == M) b.y = 0.0; b.y = -b.y; }
  double2 s = cadd(a, b), d = csub(a, b);
  double2 wd = cmul(wconj, d);
  return make_double2(s.x - wd.y, s.y + wd.x);
}

template <int MC>
__device__ __forceinline__ double2 r2c_post_nl(const double2* __restrict__ Zs, int M, int k, double2 wk) {
  double2 a = Zs[fft_out_pos<MC>(k % M)], b = cconj(Zs[fft_out_pos<MC>((M - k) % M)]);
  double2 E = make_double2(0.5 * (a.x + b.x), 0.5 * (a.y + b.y));
  double2 d = csub(a, b);
  double2 O = make_double2(0.5 * d.y, -0.5 * d.x);
  return cadd(E, cmul(wk, O));
}

// elementwise helpers on the pair of points (x = 2n, 2n+1)
__device__ __forceinline__ double2 pmul(double2 a, double2 b) { return make_double2(a.x * b.x, a.y * b.y); }
__device__ __forceinline__ double2 padd(double2 a, double2 b) { return make_double2(a.x + b.x, a.y + b.y); }
__device__ __forceinline__ double2 psub(double2 a, double2 b) { return make_double2(a.x - b.x, a.y - b.y); }
__device__ __forceinline__ double2 psc(double s, double2 a) { return make_double2(s * a.x, s * a.y); }
__device__ __forceinline__ double pabsmax(double2 a) { return fmax(fabs(a.x), fabs(a.y)); }

__device__ __forceinline__ double wave_max(double v) {
#pragma unroll
  for (int o = 32; o > 0; o >>= 1) v = fmax(v, __shfl_xor(v, o, 64));
  return v;
}

template <bool LIN, int MC>
__global__ __launch_bounds__(640) void k_xpass(XZGeom g, const double2* __restrict__ tw, DevOps ops,
                                                XPassArgs a) {
  extern __shared__ double2 lds[];
  const int M = MC > 0 ? MC : g.nx2 / 2, SS = fft_seq_cells(M), RZ = a.rz;
  const int tid = threadIdx.x, nthr = blockDim.x;
  const int r = tid / M, n = tid - r * M;
  const bool act = r < RZ;
  const int zz = blockIdx.x * RZ + (act ? r : 0);
  const int ny = a.ny, nyp = ny + 1;
  const int ka = a.ka + blockIdx.y * a.chunk;
  const int kb = min(a.kb, ka + a.chunk - 1);
  const int nz2 = g.nz2, kxp = g.kxpad, nxh = g.nxh;
  const double2* rootx = tw + g.twx;
  const double2 wconj = cconj(rootx[n < M ? n : 0]);
  double2* rl = lds + 8 * RZ * SS;
  load_tw2<MC>(rl, tw + g.px.tw, tid, nthr);
  const int pin = fft_in_pos<MC>(n < M ? n : 0), pout = fft_out_pos<MC>(n < M ? n : 0);
  const long rowoff = (long)zz * kxp;

  double2 buf = make_double2(0.0, 0.0);
  const bool hasbuf = ops.buffer != nullptr;
  if (hasbuf && act) buf = *reinterpret_cast<const double2*>(ops.buffer + (long)zz * g.nx2 + 2 * n);

  // c2r of plane p (1-based reference plane) for the components in cmask
  auto load_plane = [&](int p, double2& U, double2& V, double2& W, bool needv) {
    const int nc = 3;
    for (int c = 0; c < nc; ++c) {
      double2 v = make_double2(0.0, 0.0);
      if (act && (c != 1 || needv)) {
        const double2* row = a.Z + ((long)(c * nyp + (p - 1)) * nz2) * kxp + rowoff;
        v = c2r_pack_nl(row, nxh, M, n, wconj);
      }
      if (act) lds[(c * RZ + r) * SS + pin] = v;
    }
    __syncthreads();
    fft_run<MC, 1, 12>(lds, SS, 3 * RZ, g.px, tw, rl, tid, nthr);
    if (act) {
      U = lds[(0 * RZ + r) * SS + pout];
      V = lds[(1 * RZ + r) * SS + pout];
      W = lds[(2 * RZ + r) * SS + pout];
    }
    __syncthreads();
  };
  auto load_base = [&](int p, double2& U, double2& V, double2& W) {
    if (!LIN || !act) return;
    const double* b = a.base;
    long o = ((long)(p - 1) * nz2 + zz) * g.nx2 + 2 * n;
    long cs = (long)nyp * nz2 * g.nx2;
    U = *reinterpret_cast<const double2*>(b + o);
    V = *reinterpret_cast<const double2*>(b + cs + o);
    W = *reinterpret_cast<const double2*>(b + 2 * cs + o);
  };

  double2 U0 = {0, 0}, V0 = {0, 0}, W0 = {0, 0}, U1 = {0, 0}, V1 = {0, 0}, W1 = {0, 0};
  double2 BU0 = {0, 0}, BV0 = {0, 0}, BW0 = {0, 0}, BU1 = {0, 0}, BV1 = {0, 0}, BW1 = {0, 0};
  double2 FU, FW, GV;
  const double half = 0.5, two = 2.0;

  // prologue: planes ka-1 and ka, fluxes at y(ka-1) and yh(ka)
  load_plane(ka - 1, U0, V0, W0, true);
  load_base(ka - 1, BU0, BV0, BW0);
  load_plane(ka, U1, V1, W1, true);
  load_base(ka, BU1, BV1, BW1);
  if (!LIN) {
    if (ka - 1 == 1) { FU = pmul(V0, U0); FW = pmul(V0, W0); }                 // :53,56
    else { FU = pmul(V0, psc(half, padd(U0, U1))); FW = pmul(V0, psc(half, padd(W0, W1))); }
    double2 t = psc(half, padd(V0, V1)); GV = pmul(t, t);                        // :54-55
  } else {
    if (ka - 1 == 1) {                                                          // lin :160-170
      FU = padd(pmul(BV0, U0), pmul(V0, BU0));
      FW = padd(pmul(BV0, W0), pmul(V0, BW0));
    } else {                                                                    // lin :265-273
      FU = padd(pmul(psc(half, BV0), padd(U0, U1)), pmul(psc(half, V0), padd(BU0, BU1)));
      FW = padd(pmul(psc(half, BV0), padd(W0, W1)), pmul(psc(half, V0), padd(BW0, BW1)));
    }
    double2 t = psc(half, padd(V0, V1)), tb = psc(half, padd(BV0, BV1));
    GV = pmul(psc(two, t), tb);
  }
  U0 = U1; V0 = V1; W0 = W1; BU0 = BU1; BV0 = BV1; BW0 = BW1;

  for (int k = ka; k <= kb; ++k) {
    const bool top = (k == ny);
    load_plane(k + 1, U1, V1, W1, !top);
    load_base(k + 1, BU1, BV1, BW1);
    const int kk = k - 1;  // 0-based plane index
    const double dh = ops.dh[kk];
    // wall-parallel fluxes at y(k) (:138-146)
    double2 Fu, Fw;
    if (!LIN) {
      if (!top) { Fu = pmul(V0, psc(half, padd(U0, U1))); Fw = pmul(V0, psc(half, padd(W0, W1))); }
      else { Fu = pmul(V0, U1); Fw = pmul(V0, W1); }
    } else {
      if (!top) {
        Fu = padd(pmul(psc(half, BV0), padd(U0, U1)), pmul(psc(half, V0), padd(BU0, BU1)));
        Fw = padd(pmul(psc(half, BV0), padd(W0, W1)), pmul(psc(half, V0), padd(BW0, BW1)));
      } else {
        Fu = padd(pmul(BV0, U1), pmul(V0, BU1));
        Fw = padd(pmul(BV0, W1), pmul(V0, BW1));
      }
    }
    double2 yfu = psc(dh, psub(Fu, FU));
    double2 yfw = psc(dh, psub(Fw, FW));
    if (hasbuf) {
      double tu = ops.target[0 * nyp + kk], tw_ = ops.target[2 * nyp + kk];
      yfu = padd(yfu, pmul(buf, make_double2(U0.x - tu, U0.y - tu)));
      yfw = padd(yfw, pmul(buf, make_double2(W0.x - tw_, W0.y - tw_)));
    }
    const int nf = top ? 5 : 8;
    double2 fld[8];
    if (!LIN) {
      fld[0] = pmul(U0, U0);          // x-flux of u
      fld[2] = pmul(W0, U0);          // z-flux of u == x-flux of w
      fld[4] = pmul(W0, W0);          // z-flux of w
    } else {
      fld[0] = pmul(psc(two, BU0), U0);
      fld[2] = padd(pmul(BW0, U0), pmul(W0, BU0));
      fld[4] = pmul(psc(two, BW0), W0);
    }
    fld[1] = yfu;
    fld[3] = yfw;
    double2 G1 = {0, 0};
    if (!top) {
      const double i1 = ops.iu[2 * kk], i2 = ops.iu[2 * kk + 1];
      if (!LIN) {
        double2 t = psc(half, padd(V0, V1)); G1 = pmul(t, t);
        fld[5] = pmul(V0, padd(psc(i1, U0), psc(i2, U1)));
        fld[7] = pmul(V0, padd(psc(i1, W0), psc(i2, W1)));
      } else {
        double2 t = psc(half, padd(V0, V1)), tb = psc(half, padd(BV0, BV1));
        G1 = pmul(psc(two, t), tb);
        fld[5] = padd(pmul(BV0, padd(psc(i1, U0), psc(i2, U1))), pmul(V0, padd(psc(i1, BU0), psc(i2, BU1))));
        fld[7] = padd(pmul(BV0, padd(psc(i1, W0), psc(i2, W1))), pmul(V0, padd(psc(i1, BW0), psc(i2, BW1))));
      }
      double2 yfv = psc(ops.dn[kk], psub(G1, GV));                                 // :188-189
      if (hasbuf) {
        double tv = ops.target[1 * nyp + kk];
        yfv = padd(yfv, pmul(buf, make_double2(V0.x - tv, V0.y - tv)));
      }
      fld[6] = yfv;
    }
    // CFL maxima for this plane (check_cfl :211-213)
    if (a.cflbits) {
      double au = act ? pabsmax(padd(U0, U1)) : 0.0;
      double av = act ? pabsmax(V0) : 0.0;
      double aw = act ? pabsmax(padd(W0, W1)) : 0.0;
      if (LIN && act) {
        au = fmax(au, pabsmax(padd(BU0, BU1)));
        av = fmax(av, pabsmax(BV0));
        aw = fmax(aw, pabsmax(padd(BW0, BW1)));
      }
      au = wave_max(au); av = wave_max(av); aw = wave_max(aw);
      if ((tid & 63) == 0) {
        unsigned long long* cb = a.cflbits + 3 * kk;
        atomicMax(cb + 0, (unsigned long long)__double_as_longlong(au));
        atomicMax(cb + 1, (unsigned long long)__double_as_longlong(av));
        atomicMax(cb + 2, (unsigned long long)__double_as_longlong(aw));
      }
    }
    // forward x transforms of the nf fields
    if (act) {
#pragma unroll
      for (int f = 0; f < 8; ++f)
        if (f < nf) lds[(f * RZ + r) * SS + pin] = fld[f];
    }
    __syncthreads();
    fft_run<MC, -1, 12>(lds, SS, nf * RZ, g.px, tw, rl, tid, nthr);
    if (act) {
      // outputs kx = n (and kx = M for an undealiased x grid)
      for (int kq = n; kq < nxh; kq += M) {
        const double2 wk = rootx[kq];
        const double kxv = ops.kx[kq];
        auto X = [&](int f) { return r2c_post_nl<MC>(lds + (f * RZ + r) * SS, M, kq, wk); };
        auto ikxm = [&](double2 v) { return make_double2(-kxv * v.y, kxv * v.x); };
        double2 X0 = X(0), X1 = X(1), X2 = X(2), X3 = X(3), X4 = X(4);
        double2 Au = cadd(X1, ikxm(X0));
        double2 Aw = cadd(X3, ikxm(X2));
        const long ab0 = ((long)((0 * nyp + kk) * 2) * nz2 + zz) * kxp + kq;
        const long ab2 = ((long)((2 * nyp + kk) * 2) * nz2 + zz) * kxp + kq;
        const long bst = (long)nz2 * kxp;
        a.AB[ab0] = Au;
        a.AB[ab0 + bst] = X2;
        a.AB[ab2] = Aw;
        a.AB[ab2 + bst] = X4;
        if (!top) {
          double2 X5 = X(5), X6 = X(6), X7 = X(7);
          const long ab1 = ((long)((1 * nyp + kk) * 2) * nz2 + zz) * kxp + kq;
          a.AB[ab1] = cadd(X6, ikxm(X5));
          a.AB[ab1 + bst] = X7;
        }
      }
    }
    __syncthreads();
    // slide the window (:81, :170-171, :198)
    FU = Fu; FW = Fw; GV = G1;
    U0 = U1; V0 = V1; W0 = W1; BU0 = BU1; BV0 = BV1; BW0 = BW1;
  }
}

// ---------------------------------------------------------------------------
// Wave-centric x-pass for the compile-time lengths M <= 192.  A wavefront
// owns RW consecutive z-rows and walks a chunk of planes on its own: the
// RW*M physical point pairs of its rows are spread over the lanes (lane l
// holds the pairs v = l + 64q, q < PPL; pair v is point n = v mod M of row
// v / M), and the x transforms of the rows run in the wavefront's private
// LDS slice, ordered with wavefront-scope fences only.  The eight forward
// fields go through the slice in two batches of four per row: {u^2, d/dy
// flux_u, w u, d/dy flux_w} give Au, Bu, Aw, and {w^2, v-interp u, d/dy
// flux_v, v-interp w} give Bw, Av, Bv.  Sequence (f, r) of a batch, field f
// of row r, sits at f*RW + r.  Short rows take RW > 1: one row of 36 pairs
// leaves 28 of 64 lanes idle in every pointwise step, and its 4-sequence
// transform batches give the lanes 24 tasks.  The four wavefronts of a
// workgroup share only the root table.
// ---------------------------------------------------------------------------
//
// With PF the raw rows of the next plane are streamed into a per-wavefront
// LDS buffer by direct global-to-LDS loads (global_load_lds_dwordx4: lane i
// writes cells base + i) as soon as the current rows have been consumed, so
// their latency overlaps a whole plane of work instead of stalling the
// wavefront at the start of every plane.  A raw row takes nxh cells (M + 1
// with one row per wavefront) rounded up to 64.
// (bx, by): the workgroup's group of 4*RW z-rows and its plane range
// (blockIdx, or a work-list entry of the batched form)
template <bool LIN, int M, bool PF, bool LES = false, int RW = 1>
__device__ __forceinline__ void xpass_w_body(const XZGeom& g, const double2* __restrict__ tw, const DevOps& ops,
                                             const XPassArgs& a, int bx, int by) {
  extern __shared__ double2 lds[];
  constexpr int SEQ = fft_seq_cells(M), NP = RW * M, PPL = (NP + 63) / 64;
  constexpr int FS = RW * SEQ;                      // cells between the fields of a batch
  // the wavefront index is uniform: as a scalar it keeps the row and plane
  // offsets on the scalar unit
  const int tid = threadIdx.x, lane = tid & 63, wv = __builtin_amdgcn_readfirstlane(tid >> 6);
  const int nz2 = g.nz2, kxp = g.kxpad, nxh = g.nxh;
  // one row per wavefront keeps the raw-row stride a constant (M + 1 cells
  // rounded up), so the LDS offsets of its three raw rows fold into the reads
  const int rawn = RW == 1 ? xpass_raw_cells(M + 1) : xpass_raw_cells(nxh);
  double2* rl = lds;
  load_tw2<M>(rl, tw + g.px.tw, tid, 256);
  __syncthreads();
  const int zb = (bx * 4 + wv) * RW;  // first z-row of the wavefront
  if (zb >= nz2) return;  // after the only workgroup barrier
  const long rowoff = (long)zb * kxp, prow = (long)zb * g.nx2;
  double2* W = lds + M + wv * (4 * FS + (PF ? 3 * RW * rawn : 0));
  double2* RAW = W + 4 * FS;
  const double2* rg = tw + g.px.tw;
  const double2* rootx = tw + g.twx;
  const int ny = a.ny, nyp = ny + 1;
  const int ka = a.ka + by * a.chunk;
  const int kb = min(a.kb, ka + a.chunk - 1);
  const bool hasbuf = ops.buffer != nullptr;

  // pair q of the lane: point nq[q] of row rq(q); pin/pout carry the row's
  // sequence offset
  int nq[PPL], pin[PPL], pout[PPL];
  bool vq[PPL];
  double2 wc[PPL];
  auto rq = [&](int q) { return RW == 1 ? 0 : (vq[q] ? (lane + 64 * q) / M : 0); };
#pragma unroll
  for (int q = 0; q < PPL; ++q) {
    const int v = lane + 64 * q;
    vq[q] = v < NP;
    const int vv = vq[q] ? v : 0;
    const int r = RW == 1 ? 0 : vv / M, n = vv - r * M;
    nq[q] = n;
    wc[q] = cconj(rootx[n]);
    pin[q] = r * SEQ + fft2_in_pos<M>(n);
    pout[q] = r * SEQ + fft2_out_pos<M>(n);
  }

  // raw rows of plane p (1-based) into RAW[c*RW + r][0..nxh) (PF only); every
  // global address is clamped to an element of the row
  auto fetch = [&](int p) {
    if constexpr (PF) {
      __builtin_amdgcn_s_waitcnt(0xC07F);  // lgkmcnt(0): earlier reads of RAW are done
#pragma unroll
      for (int c = 0; c < 3; ++c) {
#pragma unroll
        for (int r = 0; r < RW; ++r) {
          const double2* row = a.Z + ((long)(c * nyp + (p - 1)) * nz2 + r) * kxp + rowoff;
          for (int i = 0; 64 * i < nxh; ++i) {
            const int idx = min(64 * i + lane, nxh - 1);
            __builtin_amdgcn_global_load_lds((const void*)(row + idx),
                                             (__attribute__((address_space(3))) void*)(RAW + (c * RW + r) * rawn + 64 * i),
                                             16, 0, 0);
          }
        }
      }
    }
  };
  auto ready = [&]() {
    if constexpr (PF) {
      __builtin_amdgcn_s_waitcnt(0x0F70);  // vmcnt(0): the raw rows have landed
      wave_sync();
    }
  };
  // c2r of plane p (1-based) into U, V, Wv at the owned points
  auto c2r_plane = [&](int p, double2* U, double2* V, double2* Wv, bool needv) {
#pragma unroll
    for (int c = 0; c < 3; ++c) {
#pragma unroll
      for (int q = 0; q < PPL; ++q) {
        if (!vq[q]) continue;
        const double2* row = PF ? RAW + (c * RW + rq(q)) * rawn
                                : a.Z + ((long)(c * nyp + (p - 1)) * nz2 + rq(q)) * kxp + rowoff;
        double2 v = make_double2(0.0, 0.0);
        if (c != 1 || needv) v = c2r_pack_nl(row, nxh, M, nq[q], wc[q]);
        W[c * FS + pin[q]] = v;
      }
    }
    wave_sync();
    fft2p_wave<M, 1>(W, SEQ, 3 * RW, rl, rg, lane);
#pragma unroll
    for (int q = 0; q < PPL; ++q) {
      if (!vq[q]) continue;
      U[q] = W[0 * FS + pout[q]];
      V[q] = W[1 * FS + pout[q]];
      Wv[q] = W[2 * FS + pout[q]];
    }
    wave_sync();
  };
  // offset of pair q in a physical [nz2][nx2] plane
  auto poff = [&](int q) { return prow + (long)rq(q) * g.nx2 + 2 * nq[q]; };
  auto load_base = [&](int p, double2* U, double2* V, double2* Wv) {
    if (!LIN) return;
    const double* b = a.base;
    const long cs = (long)nyp * nz2 * g.nx2;
#pragma unroll
    for (int q = 0; q < PPL; ++q) {
      if (!vq[q]) continue;
      const long o = (long)(p - 1) * nz2 * g.nx2 + poff(q);
      U[q] = *reinterpret_cast<const double2*>(b + o);
      V[q] = *reinterpret_cast<const double2*>(b + cs + o);
      Wv[q] = *reinterpret_cast<const double2*>(b + 2 * cs + o);
    }
  };
  // Everything the plane loop reads from global memory besides the raw rows
  // is loaded ahead of the wait for those rows: with PF a later vector load
  // would make its consumer wait for the in-flight prefetch as well.
  // The output columns of the rows: output u = lane + 64i is column
  // u mod nxh of row u / nxh.  The columns of the first KP iterations (all of
  // them with 3/2 dealiasing, nxh = 2M/3 + 1) have their root and wavenumber
  // in registers; wider rows load the rest in place.
  constexpr int KQ = (RW * (M + 1) + 63) / 64;              // nxh <= M + 1
  constexpr int KPD = (RW * ((2 * M) / 3 + 1) + 63) / 64;
  constexpr int KP = KPD < KQ ? KPD : KQ;
  const int nout = RW * nxh;
  auto orow = [&](int i) {
    const int u = min(lane + 64 * i, nout - 1);
    int r = 0;
#pragma unroll
    for (int j = 1; j < RW; ++j) r += (u >= j * nxh) ? 1 : 0;
    return r;
  };
  auto ocol = [&](int i) { return min(lane + 64 * i, nout - 1) - orow(i) * nxh; };
  double2 rxl[KP];
  double kxl[KP];
#pragma unroll
  for (int i = 0; i < KP; ++i) {
    const int kq = ocol(i);
    rxl[i] = rootx[kq];
    kxl[i] = ops.kx[kq];
  }
  auto bufv = [&](int q) { return *reinterpret_cast<const double2*>(ops.buffer + poff(q)); };
  // Subgrid stress (LES): tau_ij is added to the flux u_i u_j wherever the
  // flux is formed, so the divergence of the fluxes carries -div(tau) with the
  // same discrete operators.  tau lives at the cell centres yh; at the v-points
  // y(k) tau_12 and tau_23 are interpolated with the iu weights, as u and w are.
  constexpr bool les = LES;
  auto tauv = [&](int c, int kk, int q) {
    return *reinterpret_cast<const double2*>(a.tau + ((long)c * nyp + kk) * nz2 * g.nx2 + poff(q));
  };

  const double2 z2 = make_double2(0.0, 0.0);
  double2 U0[PPL], V0[PPL], W0[PPL], U1[PPL], V1[PPL], W1[PPL];
  double2 BU0[PPL], BV0[PPL], BW0[PPL], BU1[PPL], BV1[PPL], BW1[PPL];
  double2 FU[PPL], FW[PPL], GV[PPL];
#pragma unroll
  for (int q = 0; q < PPL; ++q) {
    U0[q] = V0[q] = W0[q] = U1[q] = V1[q] = W1[q] = z2;
    BU0[q] = BV0[q] = BW0[q] = BU1[q] = BV1[q] = BW1[q] = z2;
  }
  const double half = 0.5, two = 2.0;

  // prologue: planes ka-1 and ka, fluxes at y(ka-1) and yh(ka)
  fetch(ka - 1);
  ready();
  c2r_plane(ka - 1, U0, V0, W0, true);
  fetch(ka);
  load_base(ka - 1, BU0, BV0, BW0);
  load_base(ka, BU1, BV1, BW1);
  ready();
  c2r_plane(ka, U1, V1, W1, true);
  fetch(ka + 1);
#pragma unroll
  for (int q = 0; q < PPL; ++q) {
    if (!LIN) {
      if (ka - 1 == 1) { FU[q] = pmul(V0[q], U0[q]); FW[q] = pmul(V0[q], W0[q]); }   // :53,56
      else {
        FU[q] = pmul(V0[q], psc(half, padd(U0[q], U1[q])));
        FW[q] = pmul(V0[q], psc(half, padd(W0[q], W1[q])));
      }
      double2 t = psc(half, padd(V0[q], V1[q]));
      GV[q] = pmul(t, t);                                                                 // :54-55
    } else {
      if (ka - 1 == 1) {                                                                  // lin :160-170
        FU[q] = padd(pmul(BV0[q], U0[q]), pmul(V0[q], BU0[q]));
        FW[q] = padd(pmul(BV0[q], W0[q]), pmul(V0[q], BW0[q]));
      } else {                                                                            // lin :265-273
        FU[q] = padd(pmul(psc(half, BV0[q]), padd(U0[q], U1[q])), pmul(psc(half, V0[q]), padd(BU0[q], BU1[q])));
        FW[q] = padd(pmul(psc(half, BV0[q]), padd(W0[q], W1[q])), pmul(psc(half, V0[q]), padd(BW0[q], BW1[q])));
      }
      double2 t = psc(half, padd(V0[q], V1[q])), tb = psc(half, padd(BV0[q], BV1[q]));
      GV[q] = pmul(psc(two, t), tb);
    }
    if (les && vq[q]) {
      const int k0 = ka - 2;  // 0-based plane ka-1
      if (ka - 1 != 1) {      // the bottom wall carries no subgrid stress
        const double j1 = ldro(ops.iu, 2 * k0), j2 = ldro(ops.iu, 2 * k0 + 1);
        FU[q] = padd(FU[q], padd(psc(j1, tauv(1, k0, q)), psc(j2, tauv(1, k0 + 1, q))));
        FW[q] = padd(FW[q], padd(psc(j1, tauv(4, k0, q)), psc(j2, tauv(4, k0 + 1, q))));
      }
      GV[q] = padd(GV[q], tauv(3, k0 + 1, q));
    }
    U0[q] = U1[q]; V0[q] = V1[q]; W0[q] = W1[q];
    BU0[q] = BU1[q]; BV0[q] = BV1[q]; BW0[q] = BW1[q];
  }

  for (int k = ka; k <= kb; ++k) {
    const bool top = (k == ny);
    const int kk = k - 1;  // 0-based plane index
    const double dh = ldro(ops.dh, kk);
    const double dnk = top ? 0.0 : ldro(ops.dn, kk);
    const double i1 = top ? 0.0 : ldro(ops.iu, 2 * kk), i2 = top ? 0.0 : ldro(ops.iu, 2 * kk + 1);
    double tgu = 0.0, tgv = 0.0, tgw = 0.0;
    if (hasbuf) { tgu = ldro(ops.target, 0 * nyp + kk); tgv = ldro(ops.target, 1 * nyp + kk); tgw = ldro(ops.target, 2 * nyp + kk); }
    load_base(k + 1, BU1, BV1, BW1);
    ready();
    c2r_plane(k + 1, U1, V1, W1, !top);
    if (k < kb) fetch(k + 2);
    double2 Fu[PPL], Fw[PPL], G1[PPL];
    // batch 1: u^2, d/dy flux_u, w u, d/dy flux_w (:138-155)
#pragma unroll
    for (int q = 0; q < PPL; ++q) {
      if (!LIN) {
        if (!top) {
          Fu[q] = pmul(V0[q], psc(half, padd(U0[q], U1[q])));
          Fw[q] = pmul(V0[q], psc(half, padd(W0[q], W1[q])));
        } else {
          Fu[q] = pmul(V0[q], U1[q]);
          Fw[q] = pmul(V0[q], W1[q]);
        }
      } else {
        if (!top) {
          Fu[q] = padd(pmul(psc(half, BV0[q]), padd(U0[q], U1[q])), pmul(psc(half, V0[q]), padd(BU0[q], BU1[q])));
          Fw[q] = padd(pmul(psc(half, BV0[q]), padd(W0[q], W1[q])), pmul(psc(half, V0[q]), padd(BW0[q], BW1[q])));
        } else {
          Fu[q] = padd(pmul(BV0[q], U1[q]), pmul(V0[q], BU1[q]));
          Fw[q] = padd(pmul(BV0[q], W1[q]), pmul(V0[q], BW1[q]));
        }
      }
      if (!vq[q]) continue;
      if (les && !top) {
        Fu[q] = padd(Fu[q], padd(psc(i1, tauv(1, kk, q)), psc(i2, tauv(1, kk + 1, q))));
        Fw[q] = padd(Fw[q], padd(psc(i1, tauv(4, kk, q)), psc(i2, tauv(4, kk + 1, q))));
      }
      double2 yfu = psc(dh, psub(Fu[q], FU[q]));
      double2 yfw = psc(dh, psub(Fw[q], FW[q]));
      if (hasbuf) {
        const double2 bf = bufv(q);
        const double tu = tgu, tw_ = tgw;
        yfu = padd(yfu, pmul(bf, make_double2(U0[q].x - tu, U0[q].y - tu)));
        yfw = padd(yfw, pmul(bf, make_double2(W0[q].x - tw_, W0[q].y - tw_)));
      }
      double2 f0, f2;
      if (!LIN) { f0 = pmul(U0[q], U0[q]); f2 = pmul(W0[q], U0[q]); }
      else { f0 = pmul(psc(two, BU0[q]), U0[q]); f2 = padd(pmul(BW0[q], U0[q]), pmul(W0[q], BU0[q])); }
      if (les) { f0 = padd(f0, tauv(0, kk, q)); f2 = padd(f2, tauv(2, kk, q)); }
      W[0 * FS + pin[q]] = f0;
      W[1 * FS + pin[q]] = yfu;
      W[2 * FS + pin[q]] = f2;
      W[3 * FS + pin[q]] = yfw;
    }
    // CFL maxima for this plane (check_cfl :211-213)
    if (a.cflbits) {
      double au = 0.0, av = 0.0, aw = 0.0;
#pragma unroll
      for (int q = 0; q < PPL; ++q) {
        if (!vq[q]) continue;
        au = fmax(au, pabsmax(padd(U0[q], U1[q])));
        av = fmax(av, pabsmax(V0[q]));
        aw = fmax(aw, pabsmax(padd(W0[q], W1[q])));
        if (LIN) {
          au = fmax(au, pabsmax(padd(BU0[q], BU1[q])));
          av = fmax(av, pabsmax(BV0[q]));
          aw = fmax(aw, pabsmax(padd(BW0[q], BW1[q])));
        }
      }
      au = wave_max(au); av = wave_max(av); aw = wave_max(aw);
      if (lane == 0) {
        unsigned long long* cb = a.cflbits + 3 * kk;
        atomicMax(cb + 0, (unsigned long long)__double_as_longlong(au));
        atomicMax(cb + 1, (unsigned long long)__double_as_longlong(av));
        atomicMax(cb + 2, (unsigned long long)__double_as_longlong(aw));
      }
    }
    wave_sync();
    fft2p_wave<M, -1>(W, SEQ, 4 * RW, rl, rg, lane);
    const long bst = (long)nz2 * kxp;
    const long ab0 = ((long)((0 * nyp + kk) * 2) * nz2) * kxp + rowoff;
    const long ab1 = ((long)((1 * nyp + kk) * 2) * nz2) * kxp + rowoff;
    const long ab2 = ((long)((2 * nyp + kk) * 2) * nz2) * kxp + rowoff;
#pragma unroll
    for (int i = 0; i < KQ; ++i) {
      if (lane + 64 * i >= nout) break;
      const int kq = RW == 1 ? lane + 64 * i : ocol(i);
      const long oo = (long)orow(i) * kxp + kq;
      const double2* S = W + orow(i) * SEQ;
      const double2 wk = i < KP ? rxl[i] : rootx[kq];
      const double kxv = i < KP ? kxl[i] : ops.kx[kq];
      auto ikxm = [&](double2 v) { return make_double2(-kxv * v.y, kxv * v.x); };
      const double2 X0 = r2c_post_nl<M>(S + 0 * FS, M, kq, wk), X1 = r2c_post_nl<M>(S + 1 * FS, M, kq, wk);
      const double2 X2 = r2c_post_nl<M>(S + 2 * FS, M, kq, wk), X3 = r2c_post_nl<M>(S + 3 * FS, M, kq, wk);
      a.AB[ab0 + oo] = cadd(X1, ikxm(X0));
      a.AB[ab0 + bst + oo] = X2;
      a.AB[ab2 + oo] = cadd(X3, ikxm(X2));
    }
    wave_sync();
    // batch 2: w^2 and, below the top plane, the v fluxes (:183-195)
    const int nb2 = top ? 1 : 4;
#pragma unroll
    for (int q = 0; q < PPL; ++q) {
      if (!LIN) {
        if (!top) { double2 t = psc(half, padd(V0[q], V1[q])); G1[q] = pmul(t, t); }
        else G1[q] = z2;
      } else {
        if (!top) {
          double2 t = psc(half, padd(V0[q], V1[q])), tb = psc(half, padd(BV0[q], BV1[q]));
          G1[q] = pmul(psc(two, t), tb);
        } else {
          G1[q] = z2;
        }
      }
      if (!vq[q]) continue;
      if (les && !top) G1[q] = padd(G1[q], tauv(3, kk + 1, q));
      double2 f4 = LIN ? pmul(psc(two, BW0[q]), W0[q]) : pmul(W0[q], W0[q]);
      if (les) f4 = padd(f4, tauv(5, kk, q));
      W[0 * FS + pin[q]] = f4;
      if (!top) {
        double2 f5, f7;
        if (!LIN) {
          f5 = pmul(V0[q], padd(psc(i1, U0[q]), psc(i2, U1[q])));
          f7 = pmul(V0[q], padd(psc(i1, W0[q]), psc(i2, W1[q])));
        } else {
          f5 = padd(pmul(BV0[q], padd(psc(i1, U0[q]), psc(i2, U1[q]))),
                    pmul(V0[q], padd(psc(i1, BU0[q]), psc(i2, BU1[q]))));
          f7 = padd(pmul(BV0[q], padd(psc(i1, W0[q]), psc(i2, W1[q]))),
                    pmul(V0[q], padd(psc(i1, BW0[q]), psc(i2, BW1[q]))));
        }
        if (les) {
          f5 = padd(f5, padd(psc(i1, tauv(1, kk, q)), psc(i2, tauv(1, kk + 1, q))));
          f7 = padd(f7, padd(psc(i1, tauv(4, kk, q)), psc(i2, tauv(4, kk + 1, q))));
        }
        double2 yfv = psc(dnk, psub(G1[q], GV[q]));                                    // :188-189
        if (hasbuf) {
          const double tv = tgv;
          yfv = padd(yfv, pmul(bufv(q), make_double2(V0[q].x - tv, V0[q].y - tv)));
        }
        W[1 * FS + pin[q]] = f5;
        W[2 * FS + pin[q]] = yfv;
        W[3 * FS + pin[q]] = f7;
      }
    }
    wave_sync();
    fft2p_wave<M, -1>(W, SEQ, nb2 * RW, rl, rg, lane);
#pragma unroll
    for (int i = 0; i < KQ; ++i) {
      if (lane + 64 * i >= nout) break;
      const int kq = RW == 1 ? lane + 64 * i : ocol(i);
      const long oo = (long)orow(i) * kxp + kq;
      const double2* S = W + orow(i) * SEQ;
      const double2 wk = i < KP ? rxl[i] : rootx[kq];
      const double kxv = i < KP ? kxl[i] : ops.kx[kq];
      auto ikxm = [&](double2 v) { return make_double2(-kxv * v.y, kxv * v.x); };
      a.AB[ab2 + bst + oo] = r2c_post_nl<M>(S + 0 * FS, M, kq, wk);
      if (!top) {
        const double2 X5 = r2c_post_nl<M>(S + 1 * FS, M, kq, wk), X6 = r2c_post_nl<M>(S + 2 * FS, M, kq, wk);
        const double2 X7 = r2c_post_nl<M>(S + 3 * FS, M, kq, wk);
        a.AB[ab1 + oo] = cadd(X6, ikxm(X5));
        a.AB[ab1 + bst + oo] = X7;
      }
    }
    wave_sync();
    // slide the window (:81, :170-171, :198)
#pragma unroll
    for (int q = 0; q < PPL; ++q) {
      FU[q] = Fu[q]; FW[q] = Fw[q]; GV[q] = G1[q];
      U0[q] = U1[q]; V0[q] = V1[q]; W0[q] = W1[q];
      BU0[q] = BU1[q]; BV0[q] = BV1[q]; BW0[q] = BW1[q];
    }
  }
}

template <bool LIN, int M, bool PF, bool LES = false, int RW = 1>
__global__ __launch_bounds__(256) void k_xpass_w(XZGeom g, const double2* __restrict__ tw, DevOps ops,
                                                  XPassArgs a) {
  xpass_w_body<LIN, M, PF, LES, RW>(g, tw, ops, a, blockIdx.x, blockIdx.y);
}

// Batched form: workgroup i runs work-list entry it[i] of channel it[i].m.
template <bool LIN, int M, bool PF, int RW>
__global__ __launch_bounds__(256) void k_xpass_w_b(const XBatchArgs* __restrict__ args,
                                                    const BatchItem* __restrict__ it) {
  const BatchItem w = it[blockIdx.x];
  const XBatchArgs& x = args[w.m];
  xpass_w_body<LIN, M, PF, false, RW>(x.g, x.tw, x.ops, x.a, w.bx, w.by);
}

using XPassFn = void (*)(XZGeom, const double2*, DevOps, XPassArgs);
using XPassBFn = void (*)(const XBatchArgs*, const BatchItem*);

// Rows per wavefront with a kernel: 1 at every length, 2 for M <= 72 and 4
// for M <= 24, the last two with the row prefetch only.
static bool xpass_rw_built(int M, int rw, bool pf) {
  return rw == 1 || (pf && ((rw == 2 && M <= 72) || (rw == 4 && M <= 24)));
}

// The default: two rows per wavefront where a row's pairs fill at most 36
// lanes; the batched launches also at the lengths where that halves the
// wavefronts per SIMD (their lengths share the GPU, so fewer idle lanes count
// for more than latency hiding: 163 -> 178 steps/s on the 48-member
// ensemble, where a single 64x65x64 channel's x-pass loses 1 %).
int xpass_rows(int M, int nz2, int req, bool pf, bool batched) {
  auto ok = [&](int r) { return xpass_rw_built(M, r, pf) && nz2 % r == 0; };
  if (req > 0 && ok(req)) return req;
  const int d = (M <= 24 || (batched && M <= 72)) ? 2 : 1;
  return ok(d) ? d : 1;
}

// LDS of a wavefront x-pass workgroup: the root table and, per wavefront, 4*rw
// sequences and (with the prefetch) 3*rw raw rows (of M + 1 cells rounded up
// to 64 for one row per wavefront, of nxh cells rounded up otherwise)
size_t xpass_w_lds(int M, int rw, int nxh, bool pf) {
  const int rawn = rw == 1 ? xpass_raw_cells(M + 1) : xpass_raw_cells(nxh);
  return sizeof(double2) * (M + 4 * (4 * rw * fft_seq_cells(M) + (pf ? 3 * rw * rawn : 0)));
}

template <int N, bool PF, int RW>
static XPassFn xpass_w_variant(bool les, bool lin) {
  if (les) return k_xpass_w<false, N, PF, true, RW>;
  return lin ? (XPassFn)k_xpass_w<true, N, PF, false, RW> : (XPassFn)k_xpass_w<false, N, PF, false, RW>;
}

template <int N>
static XPassFn xpass_w_fn(bool les, bool lin, bool pf, int rw) {
  if constexpr (N <= 144) {
    if (pf) {
      if constexpr (N <= 24) {
        if (rw == 4) return xpass_w_variant<N, true, 4>(les, lin);
      }
      if constexpr (N <= 72) {
        if (rw == 2) return xpass_w_variant<N, true, 2>(les, lin);
      }
      return rw == 1 ? xpass_w_variant<N, true, 1>(les, lin) : nullptr;
    }
  }
  return rw == 1 ? xpass_w_variant<N, false, 1>(les, lin) : nullptr;
}

template <int N, int RW>
static XPassBFn xpass_wb_variant(bool lin) {
  constexpr bool pf = N <= 144;
  return lin ? (XPassBFn)k_xpass_w_b<true, N, pf, RW> : (XPassBFn)k_xpass_w_b<false, N, pf, RW>;
}

template <int N>
static XPassBFn xpass_wb_fn(bool lin, int rw) {
  if constexpr (N <= 24) {
    if (rw == 4) return xpass_wb_variant<N, 4>(lin);
  }
  if constexpr (N <= 72) {
    if (rw == 2) return xpass_wb_variant<N, 2>(lin);
  }
  return rw == 1 ? xpass_wb_variant<N, 1>(lin) : nullptr;
}

static XPassBFn xpass_wb_pick(int M, bool lin, int rw) {
  XPassBFn fn = nullptr;
#define XPASSB_CASE(N)                                   \
  if constexpr (N > 0 && N <= 192) fn = xpass_wb_fn<N>(lin, rw);
  FFT_DISPATCH(M, XPASSB_CASE)
#undef XPASSB_CASE
  if (!fn) throw chan_error("batched x-pass: no wavefront kernel for nx2/2 = " + std::to_string(M) + " with " +
                            std::to_string(rw) + " rows per wavefront");
  return fn;
}

// The wavefront x-pass of several channels with the same nx2/2 = M in one
// launch (the batched ensemble); the row prefetch is on wherever the single
// launch uses it (M <= 144).  rw: rows per wavefront (every member's nz2 a
// multiple of it); nxh: the largest nx/2+1 of the members.
bool xpass_batchable(int M) { return fft2_split(M).n1 > 0 && M <= 192; }

void launch_xpass_batched(int M, bool lin, int rw, int nxh, const XBatchArgs* args, const BatchItem* items,
                          int nitems, hipStream_t s) {
  if (nitems <= 0) return;
  if (!xpass_batchable(M)) throw chan_error("batched x-pass: nx2/2 = " + std::to_string(M) + " has no wavefront kernel");
  const XPassBFn fn = xpass_wb_pick(M, lin, rw);
  hipLaunchKernelGGL(fn, dim3(nitems), dim3(256), xpass_w_lds(M, rw, nxh, M <= 144), s, args, items);
  LAUNCH_CHECK("k_xpass_w_b");
}

// Resident workgroups of the batched wavefront x-pass at length M on the
// current device (the plane-range rule of the single launch, over the whole batch).
long xpass_batched_resident(int M, bool lin, int rw, int nxh) {
  const void* fn = (const void*)xpass_wb_pick(M, lin, rw);
  int dev = 0, cus = 0, per_cu = 0;
  if (hipGetDevice(&dev) != hipSuccess ||
      hipDeviceGetAttribute(&cus, hipDeviceAttributeMultiprocessorCount, dev) != hipSuccess || cus < 1)
    cus = 256;
  if (hipOccupancyMaxActiveBlocksPerMultiprocessor(&per_cu, fn, 256, xpass_w_lds(M, rw, nxh, M <= 144)) != hipSuccess ||
      per_cu < 1)
    per_cu = 1;
  return (long)per_cu * cus;
}

// Planes per workgroup of the wavefront x-pass when no chunk is set: the
// plane-rows of the pass spread evenly over the `res` workgroups resident at
// a time, so that one wave of workgroups with full ranges covers nearly all
// of them and the short last ranges fill the remaining slots; at least two
// planes per range.
static int xpass_chunk_auto(long zw, int np, long res) {
  const long c = (zw * np + res - 1) / res;
  return (int)std::max(2L, std::min<long>(c, std::max(2, np)));
}

void launch_xpass(const XZGeom& g, const double2* tw, const DevOps& ops, const XPassArgs& a0,
                  hipStream_t s) {
  XPassArgs a = a0;
  const int M = g.nx2 / 2;
  if (a.kb < a.ka) return;
  // rows per workgroup: largest divisor of nz2 with rz*M <= maxthr
  const int maxthr = (a.maxthr >= 64 && a.maxthr <= 640) ? a.maxthr : 640;
  int rz = 1;
  for (int d = 1; d <= g.nz2; ++d)
    if (g.nz2 % d == 0 && d * M <= std::max(maxthr, M)) rz = d;
  a.rz = rz;
  int nthr = ((rz * M + 63) / 64) * 64;
  if (nthr > 640) throw chan_error("x-pass: nx2/2 above 640 is not supported");
  const bool ct = fft2_split(M).n1 > 0;
  // staging bound of the runtime-planned in-place FFT: 8*rz sequences of length M
  if (!ct)
    for (int i = 0; i < g.px.nst; ++i) {
      int R = g.px.rad[i];
      long nbf = (long)8 * rz * (M / R);
      if (((nbf + nthr - 1) / nthr) * R > 12) throw chan_error("x-pass FFT staging bound");
    }
  const int nplanes = a.kb - a.ka + 1;
  if (a.tau && !(a.wave && ct && M <= 192))
    throw chan_error("the subgrid closure needs the wavefront x-pass (nx2/2 <= 192 of the form 2^a 3^b)");
  if (a.wave && ct && M <= 192) {
    const bool pf = a.prefetch && M <= 144;
    if (a.tau && a.base) throw chan_error("the subgrid closure is not defined for the linearized pass");
    const int rw = xpass_rows(M, g.nz2, a.rows, pf, false);
    const size_t ldw = xpass_w_lds(M, rw, g.nxh, pf);
    XPassFn fn = nullptr;
#define XPASSW_CASE(N)                                                                  \
  if constexpr (N > 0 && N <= 192) fn = xpass_w_fn<N>(a.tau != nullptr, a.base != nullptr, pf, rw);
    FFT_DISPATCH(M, XPASSW_CASE)
#undef XPASSW_CASE
    if (!fn) throw chan_error("x-pass: no wavefront kernel for this length");
    const long zw = (g.nz2 / rw + 3) / 4;
    if (a.chunk <= 0) {
      // resident workgroups of this kernel on the current device
      int dev = 0, cus = 0;
      if (hipGetDevice(&dev) != hipSuccess ||
          hipDeviceGetAttribute(&cus, hipDeviceAttributeMultiprocessorCount, dev) != hipSuccess || cus < 1)
        cus = 256;
      int per_cu = 0;
      if (hipOccupancyMaxActiveBlocksPerMultiprocessor(&per_cu, fn, 256, ldw) != hipSuccess || per_cu < 1) per_cu = 1;
      a.chunk = xpass_chunk_auto(zw, nplanes, (long)per_cu * cus);
    }
    const int nch = (nplanes + a.chunk - 1) / a.chunk;
    hipLaunchKernelGGL(fn, dim3((unsigned)zw, nch), dim3(256), ldw, s, g, tw, ops, a);
    LAUNCH_CHECK("k_xpass_w");
    return;
  }
  if (a.chunk <= 0) a.chunk = 6;
  const int nch = (nplanes + a.chunk - 1) / a.chunk;
  dim3 grid(g.nz2 / rz, nch);
  size_t lds = sizeof(double2) * (8 * rz * fft_seq_cells(M) + (ct ? M : 0));
#define XPASS_CASE(N)                                                                 \
  if (a.base) hipLaunchKernelGGL((k_xpass<true, N>), grid, dim3(nthr), lds, s, g, tw, ops, a); \
  else hipLaunchKernelGGL((k_xpass<false, N>), grid, dim3(nthr), lds, s, g, tw, ops, a)
  FFT_DISPATCH(M, XPASS_CASE)
#undef XPASS_CASE
  LAUNCH_CHECK("k_xpass");
}
