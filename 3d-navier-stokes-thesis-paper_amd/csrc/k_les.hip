// Subgrid-scale closure for large-eddy simulation on the channel grid (no
// counterpart in the reference: it carries only the switches `subgrid`,
// `filter_dim`, `lagr_avg` (PFC/flow.dat:19-22) and commented-out calls of a
// dynamic model, PFC/main_mpi.f90:10-13,74,77-78).
//
// The closure gives the subgrid stress tau_ij at the cell centres yh of the
// dealiased physical grid; the x-pass adds it to the flux u_i u_j wherever the
// flux is formed, so the nonlinear term becomes div(u u + tau) with the
// reference's own discrete operators.
//
// Fields of one update, all at the cell centres of planes 1..ny-1 (zero on the
// wall planes 0 and ny):
//   G[0..8]  = du_i/dx_j  (i = u, v, w; j = x, y, z), G[9..11] = u, v, w
//   d/dx, d/dz spectrally (i kx, i kz); v at yh is the mean of its two faces;
//   du/dy, dw/dy by the central difference over yh(k-1), yh(k+1); dv/dy by the
//   face difference dh(k) (v(k) - v(k-1)).
// Models:
//   1  dynamic Smagorinsky: tau_ij = -2 C(y) D^2 |S| S_ij, |S| = sqrt(2 S:S),
//      D = (dx dz (y(k)-y(k-1)))^(1/3), C(y) from the Germano identity with
//      the trapezoidal test filter (1/4, 1/2, 1/4) along x and z on the
//      dealiased grid (width ratio 2 in the model), least squares over each
//      plane (Lilly), clipped at zero; C(y) is
//      evaluated at the first substep of each time step and held for the
//      other two;
//   2  learned: an MLP maps (G[0..8]/|G|, 1-|yh|) to tau_ij / (D^2 |G|^2);
//      a point whose learned stress removes energy at a rate below -beta times
//      the dynamic Smagorinsky rate, or is not finite, takes the Smagorinsky
//      stress instead (the fallback gate);
//   3  prescribed tau (chan_les_set_tau).
// The stress is evaluated at every substep, or (every_substep = 0) at the
// first substep of each step and held for the other two.
#include "chan_internal.h"
#include <cmath>

namespace {

constexpr int LES_NF = 12;  // fields of G

// symmetric-tensor component order 11, 12, 13, 22, 23, 33 and its (i, j) pairs
__device__ __constant__ int c_ti[6] = {0, 0, 0, 1, 1, 2};
__device__ __constant__ int c_tj[6] = {0, 1, 2, 1, 2, 2};

__device__ __forceinline__ double2 ikm(double k, double2 a) { return make_double2(-k * a.y, k * a.x); }

__global__ void k_les_grad(DevOps ops, int ny, int nxh, int nz, const double2* __restrict__ vel,
                           double2* __restrict__ G) {
  const long nm = (long)nxh * nz;
  const long m = (long)blockIdx.x * blockDim.x + threadIdx.x;
  const int kk = blockIdx.y, nyp = ny + 1;
  if (m >= nm) return;
  auto out = [&](int f, double2 v) { G[((long)f * nyp + kk) * nm + m] = v; };
  if (kk == 0 || kk == ny) {
    for (int f = 0; f < LES_NF; ++f) out(f, make_double2(0.0, 0.0));
    return;
  }
  const int ix = (int)(m / nz), iz = (int)(m - (long)ix * nz);
  const double kx = ops.kx[ix], kz = ops.kz[iz];
  auto V = [&](int c, int k) { return vel[((long)c * nyp + k) * nm + m]; };
  const double2 u = V(0, kk), w = V(2, kk), vm = V(1, kk - 1), v0 = V(1, kk);
  const double2 vc = make_double2(0.5 * (vm.x + v0.x), 0.5 * (vm.y + v0.y));
  const double dyc = 1.0 / (ops.hn[kk] + ops.hn[kk - 1]);
  const double dhk = ops.dh[kk];
  const double2 up = V(0, kk + 1), um = V(0, kk - 1), wp = V(2, kk + 1), wm = V(2, kk - 1);
  out(0, ikm(kx, u));
  out(1, make_double2((up.x - um.x) * dyc, (up.y - um.y) * dyc));
  out(2, ikm(kz, u));
  out(3, ikm(kx, vc));
  out(4, make_double2(dhk * (v0.x - vm.x), dhk * (v0.y - vm.y)));
  out(5, ikm(kz, vc));
  out(6, ikm(kx, w));
  out(7, make_double2((wp.x - wm.x) * dyc, (wp.y - wm.y) * dyc));
  out(8, ikm(kz, w));
  out(9, u);
  out(10, vc);
  out(11, w);
}

struct Strain {
  double s[6];  // S_ij in the symmetric order
  double mag;   // |S| = sqrt(2 S:S)
};
__device__ __forceinline__ Strain strain(const double g[9]) {
  Strain r;
  r.s[0] = g[0];
  r.s[1] = 0.5 * (g[1] + g[3]);
  r.s[2] = 0.5 * (g[2] + g[6]);
  r.s[3] = g[4];
  r.s[4] = 0.5 * (g[5] + g[7]);
  r.s[5] = g[8];
  const double ss = r.s[0] * r.s[0] + r.s[3] * r.s[3] + r.s[5] * r.s[5] +
                    2.0 * (r.s[1] * r.s[1] + r.s[2] * r.s[2] + r.s[4] * r.s[4]);
  r.mag = sqrt(2.0 * ss);
  return r;
}
// sum_ij a_ij b_ij of two symmetric tensors in the symmetric order
__device__ __forceinline__ double contract(const double* a, const double* b) {
  return a[0] * b[0] + a[3] * b[3] + a[5] * b[5] + 2.0 * (a[1] * b[1] + a[2] * b[2] + a[4] * b[4]);
}

// Dynamic procedure with the trapezoidal test filter (1/4, 1/2, 1/4) along x,
// then along z, of the closure fields on the dealiased grid, fused: a
// workgroup takes a 16 x 16 tile of one interior plane with a one-point
// (periodic) halo in LDS, holding the 12 closure fields and the 12 products
// u_i u_j, |S| S_ij at every halo point; every point of the tile filters
// them, forms L_ij = T(u_i u_j) - T(u_i) T(u_j),
// M_ij = 2 D^2 (T(|S| S_ij) - 4 |S^| S^_ij), and the workgroup sums L:M and
// M:M (partials per tile, reduced per plane by k_les_cs2).  With FEAT it also
// writes the network inputs of its tile's points from the gradients it holds
// in LDS, as k_les_feat does (same arithmetic), so that the learned closure
// needs no separate pass over the gradients where C(y) is recomputed.
#define DYN_T 16
#define DYN_H (DYN_T + 2)
template <bool FEAT>
__global__ __launch_bounds__(256) void k_les_dyn(int ny, int nz2, int nx2, const double* __restrict__ Ph,
                                                 const double* __restrict__ delta2, double* __restrict__ part,
                                                 const double* __restrict__ ydist, float* __restrict__ X,
                                                 double* __restrict__ scale) {
  __shared__ double F[24][DYN_H * DYN_H];
  __shared__ double red[2][256];
  const int kk = blockIdx.y, nyp = ny + 1;
  const long ps = (long)nz2 * nx2;
  const int ntx = (nx2 + DYN_T - 1) / DYN_T;
  const int tz0 = (blockIdx.x / ntx) * DYN_T, tx0 = (blockIdx.x % ntx) * DYN_T;
  const bool inner = (kk > 0 && kk < ny);
  double lm = 0.0, mm = 0.0;
  if (inner) {
    for (int p = threadIdx.x; p < DYN_H * DYN_H; p += blockDim.x) {
      const int hz = p / DYN_H, hx = p - hz * DYN_H;
      int z = tz0 + hz - 1, x = tx0 + hx - 1;
      z = (z < 0) ? z + nz2 : (z >= nz2 ? z - nz2 : z);
      x = (x < 0) ? x + nx2 : (x >= nx2 ? x - nx2 : x);
      const double* src = Ph + (long)kk * ps + (long)z * nx2 + x;
      double g[9], u[3];
#pragma unroll
      for (int f = 0; f < 9; ++f) g[f] = src[(long)f * nyp * ps];
#pragma unroll
      for (int c = 0; c < 3; ++c) u[c] = src[(long)(9 + c) * nyp * ps];
      const Strain S = strain(g);
#pragma unroll
      for (int f = 0; f < 9; ++f) F[f][p] = g[f];
#pragma unroll
      for (int c = 0; c < 3; ++c) F[9 + c][p] = u[c];
#pragma unroll
      for (int t = 0; t < 6; ++t) {
        F[12 + t][p] = u[c_ti[t]] * u[c_tj[t]];
        F[18 + t][p] = S.mag * S.s[t];
      }
    }
  }
  __syncthreads();
  if (inner) {
    const int lz = threadIdx.x / DYN_T, lx = threadIdx.x - lz * DYN_T;
    if (tz0 + lz < nz2 && tx0 + lx < nx2) {
      auto filt = [&](int f) {
        const double* a = F[f] + lz * DYN_H + lx;
        const double r0 = 0.25 * a[0] + 0.5 * a[1] + 0.25 * a[2];
        const double r1 = 0.25 * a[DYN_H] + 0.5 * a[DYN_H + 1] + 0.25 * a[DYN_H + 2];
        const double r2 = 0.25 * a[2 * DYN_H] + 0.5 * a[2 * DYN_H + 1] + 0.25 * a[2 * DYN_H + 2];
        return 0.25 * r0 + 0.5 * r1 + 0.25 * r2;
      };
      double tg[9], tu[3], L[6], M[6];
#pragma unroll
      for (int f = 0; f < 9; ++f) tg[f] = filt(f);
#pragma unroll
      for (int c = 0; c < 3; ++c) tu[c] = filt(9 + c);
      const Strain St = strain(tg);
      const double d2 = delta2[kk];
#pragma unroll
      for (int t = 0; t < 6; ++t) {
        L[t] = filt(12 + t) - tu[c_ti[t]] * tu[c_tj[t]];
        M[t] = 2.0 * d2 * (filt(18 + t) - 4.0 * St.mag * St.s[t]);
      }
      lm = contract(L, M);
      mm = contract(M, M);
      if (FEAT) {
        const int hp = (lz + 1) * DYN_H + lx + 1;
        double gp[9], n2 = 0.0;
#pragma unroll
        for (int f = 0; f < 9; ++f) {
          gp[f] = F[f][hp];
          n2 += gp[f] * gp[f];
        }
        const double gn = sqrt(n2);
        const double inv = gn > 0.0 ? 1.0 / gn : 0.0;
        const long p = (long)(kk - 1) * ps + (long)(tz0 + lz) * nx2 + tx0 + lx;
#pragma unroll
        for (int f = 0; f < 9; ++f) X[p * 10 + f] = (float)(gp[f] * inv);
        X[p * 10 + 9] = (float)ydist[kk];
        scale[p] = delta2[kk] * n2;
      }
    }
  }
  red[0][threadIdx.x] = lm;
  red[1][threadIdx.x] = mm;
  __syncthreads();
  for (int o = 128; o > 0; o >>= 1) {
    if ((int)threadIdx.x < o) {
      red[0][threadIdx.x] += red[0][threadIdx.x + o];
      red[1][threadIdx.x] += red[1][threadIdx.x + o];
    }
    __syncthreads();
  }
  if (threadIdx.x == 0) {
    part[((long)kk * gridDim.x + blockIdx.x) * 2 + 0] = red[0][0];
    part[((long)kk * gridDim.x + blockIdx.x) * 2 + 1] = red[1][0];
  }
}

__global__ void k_les_cs2(int ny, int nblk, const double* __restrict__ part, double* __restrict__ cs2) {
  const int kk = blockIdx.x * blockDim.x + threadIdx.x;
  if (kk > ny) return;
  double lm = 0.0, mm = 0.0;
  for (int b = 0; b < nblk; ++b) {
    lm += part[((long)kk * nblk + b) * 2 + 0];
    mm += part[((long)kk * nblk + b) * 2 + 1];
  }
  double c = (mm > 0.0) ? lm / mm : 0.0;
  cs2[kk] = (kk == 0 || kk == ny || !(c > 0.0)) ? 0.0 : c;
}

// Smagorinsky stress at the interior points (zero on the wall planes)
__global__ void k_les_tau_smag(int ny, long ps, const double* __restrict__ Ph, const double* __restrict__ cs2,
                               const double* __restrict__ delta2, double* __restrict__ tau) {
  const long i = (long)blockIdx.x * blockDim.x + threadIdx.x;
  const int kk = blockIdx.y, nyp = ny + 1;
  if (i >= ps) return;
  double* o = tau + (long)kk * ps + i;
  const long cs = (long)nyp * ps;
  if (kk == 0 || kk == ny) {
    for (int t = 0; t < 6; ++t) o[t * cs] = 0.0;
    return;
  }
  double g[9];
  for (int f = 0; f < 9; ++f) g[f] = Ph[((long)f * nyp + kk) * ps + i];
  const Strain S = strain(g);
  const double c = -2.0 * cs2[kk] * delta2[kk] * S.mag;
  for (int t = 0; t < 6; ++t) o[t * cs] = c * S.s[t];
}

// MLP inputs of the interior points p = ((k-1) nz2 + z) nx2 + x and the stress
// scale D^2 |G|^2
__global__ void k_les_feat(int ny, long ps, const double* __restrict__ Ph, const double* __restrict__ ydist,
                           const double* __restrict__ delta2, float* __restrict__ X, double* __restrict__ scale) {
  const long i = (long)blockIdx.x * blockDim.x + threadIdx.x;
  const int kk = blockIdx.y + 1, nyp = ny + 1;
  if (i >= ps) return;
  double g[9], n2 = 0.0;
  for (int f = 0; f < 9; ++f) {
    g[f] = Ph[((long)f * nyp + kk) * ps + i];
    n2 += g[f] * g[f];
  }
  const double gn = sqrt(n2);
  const double inv = gn > 0.0 ? 1.0 / gn : 0.0;
  const long p = (long)(kk - 1) * ps + i;
  for (int f = 0; f < 9; ++f) X[p * 10 + f] = (float)(g[f] * inv);
  X[p * 10 + 9] = (float)ydist[kk];
  scale[p] = delta2[kk] * n2;
}

// learned stress with the fallback gate
__global__ __launch_bounds__(256) void k_les_tau_mlp(int ny, long ps, const double* __restrict__ Ph,
                                                     const float* __restrict__ Y, const double* __restrict__ scale,
                                                     const double* __restrict__ cs2,
                                                     const double* __restrict__ delta2, double beta,
                                                     double* __restrict__ tau, unsigned int* __restrict__ gated) {
  const long i = (long)blockIdx.x * blockDim.x + threadIdx.x;
  const int kk = blockIdx.y + 1, nyp = ny + 1;
  bool fb = false;
  if (i < ps) {
    double g[9];
    for (int f = 0; f < 9; ++f) g[f] = Ph[((long)f * nyp + kk) * ps + i];
    const Strain S = strain(g);
    const long p = (long)(kk - 1) * ps + i;
    double tm[6], ts[6];
    const double c = -2.0 * cs2[kk] * delta2[kk] * S.mag;
    for (int t = 0; t < 6; ++t) {
      tm[t] = scale[p] * (double)Y[p * 6 + t];
      ts[t] = c * S.s[t];
    }
    const double em = -contract(tm, S.s), es = -contract(ts, S.s);
    fb = !isfinite(em) || (beta >= 0.0 && !(em >= -beta * es));
    const long cs = (long)nyp * ps;
    double* o = tau + (long)kk * ps + i;
    for (int t = 0; t < 6; ++t) o[t * cs] = fb ? ts[t] : tm[t];
    if (kk == 1)  // the wall planes carry no stress
      for (int t = 0; t < 6; ++t) {
        tau[t * cs + i] = 0.0;
        tau[t * cs + (long)ny * ps + i] = 0.0;
      }
  }
  // fallback count of the block (one slot per block: no contended atomics)
  __shared__ unsigned int wc[4];
  const unsigned long long b = __ballot(fb);
  if ((threadIdx.x & 63) == 0) wc[threadIdx.x >> 6] = (unsigned int)__popcll(b);
  __syncthreads();
  if (threadIdx.x == 0) gated[(long)blockIdx.y * gridDim.x + blockIdx.x] = wc[0] + wc[1] + wc[2] + wc[3];
}

}  // namespace

// ---------------------------------------------------------------------------
// host side
// ---------------------------------------------------------------------------
struct LesState {
  int mode = 0;
  double beta = 1.0;
  chan_mlp* net = nullptr;
  long ps = 0, nm = 0, npts = 0;
  int nblk = 0;                // Germano partial sums per plane (tiles of k_les_dyn)
  double2* G = nullptr;        // [12][nyp][nm]
  double2* Zl = nullptr;       // [12 nyp][nz2][kxpad] transform intermediate
  double* Ph = nullptr;        // [12][nyp][ps]
  double* tau = nullptr;       // [6][nyp][ps]
  double* cs2 = nullptr;       // [nyp]
  double* delta2 = nullptr;    // [nyp]
  double* ydist = nullptr;     // [nyp]
  double* part = nullptr;      // [nyp][nblk][2]
  float* X = nullptr;          // [npts][10]
  float* Y = nullptr;          // [npts][6]
  double* scale = nullptr;     // [npts]
  unsigned int* gated = nullptr;  // fallback points per block of k_les_tau_mlp
  long ngated = 0;
  bool fresh = false;          // tau is current for the state (mode 3: set by the caller)
  bool cs2_valid = false;      // C(y) computed for the current step
  int every = 1;               // 1: tau at every substep; 0: at the first substep of a step only
  bool tau_valid = false;
};

namespace {

template <class T>
void lalloc(T*& p, size_t n) {
  CHAN_CHECK(hipMalloc((void**)&p, sizeof(T) * std::max<size_t>(n, 1)));
}
template <class T>
void lfree(T*& p) {
  if (p) (void)hipFree(p);
  p = nullptr;
}

void les_c2r(chan_ctx* c, LesState* L, const double2* in, int np, double* out) {
  const XZGeom& g = c->gd;
  const long zps = (long)g.nz2 * g.kxpad;
  launch_zinv(g, c->d_tw, in, L->nm, L->Zl, zps, np, c->stream);
  launch_xc2r(g, c->d_tw, L->Zl, zps, out, L->ps, np, c->stream);
}

// G, its physical fields Ph and (dynamic) the dynamic coefficient
// (with `feat`, the dynamic procedure also writes the network inputs)
void les_fields(chan_ctx* c, LesState* L, bool dynamic, bool feat = false) {
  const int ny = c->prm.ny, nyp = c->nyp;
  dim3 gm((unsigned)((L->nm + 255) / 256), nyp);
  hipLaunchKernelGGL(k_les_grad, gm, dim3(256), 0, c->stream, c->ops, ny, c->nxh, c->prm.nz, c->d_vel, L->G);
  LAUNCH_CHECK("k_les_grad");
  les_c2r(c, L, L->G, LES_NF * nyp, L->Ph);
  if (!dynamic) return;
  const XZGeom& g = c->gd;
  if (feat)
    hipLaunchKernelGGL(k_les_dyn<true>, dim3(L->nblk, nyp), dim3(256), 0, c->stream, ny, g.nz2, g.nx2, L->Ph,
                       L->delta2, L->part, L->ydist, L->X, L->scale);
  else
    hipLaunchKernelGGL(k_les_dyn<false>, dim3(L->nblk, nyp), dim3(256), 0, c->stream, ny, g.nz2, g.nx2, L->Ph,
                       L->delta2, L->part, nullptr, nullptr, nullptr);
  LAUNCH_CHECK("k_les_dyn");
  hipLaunchKernelGGL(k_les_cs2, dim3((nyp + 63) / 64), dim3(64), 0, c->stream, ny, L->nblk, L->part, L->cs2);
  LAUNCH_CHECK("k_les_cs2");
}

// the closure on the current state; C(y) is recomputed when `dyn` (the first
// substep of a time step, an explicit update) or not yet known
void les_compute(chan_ctx* c, LesState* L, bool dyn) {
  const int ny = c->prm.ny, nyp = c->nyp;
  if (L->mode == 3) return;
  const bool d = dyn || !L->cs2_valid;
  les_fields(c, L, d, d && L->mode == 2);
  L->cs2_valid = true;
  dim3 gp((unsigned)((L->ps + 255) / 256), nyp);
  if (L->mode == 1) {
    hipLaunchKernelGGL(k_les_tau_smag, gp, dim3(256), 0, c->stream, ny, L->ps, L->Ph, L->cs2, L->delta2, L->tau);
    LAUNCH_CHECK("k_les_tau_smag");
    return;
  }
  dim3 gi((unsigned)((L->ps + 255) / 256), ny - 1);
  if (!d) {  // otherwise k_les_dyn wrote the network inputs
    hipLaunchKernelGGL(k_les_feat, gi, dim3(256), 0, c->stream, ny, L->ps, L->Ph, L->ydist, L->delta2, L->X,
                       L->scale);
    LAUNCH_CHECK("k_les_feat");
  }
  mlp_infer_device(L->net, L->npts, L->X, L->Y, c->stream);
  hipLaunchKernelGGL(k_les_tau_mlp, gi, dim3(256), 0, c->stream, ny, L->ps, L->Ph, L->Y, L->scale, L->cs2,
                     L->delta2, L->beta, L->tau, L->gated);
  LAUNCH_CHECK("k_les_tau_mlp");
}

LesState* les_need(chan_ctx* c) {
  if (!c->les || c->les->mode == 0) throw chan_error("LES closure not configured (chan_les_configure)");
  return c->les;
}

}  // namespace

void les_destroy(LesState* L) {
  if (!L) return;
  lfree(L->G); lfree(L->Zl); lfree(L->Ph); lfree(L->tau);
  lfree(L->cs2); lfree(L->delta2); lfree(L->ydist); lfree(L->part); lfree(L->X); lfree(L->Y);
  lfree(L->scale); lfree(L->gated);
  delete L;
}

const double* les_tau_for_pass(chan_ctx* c, int rk) {
  LesState* L = c->les;
  if (!L || L->mode == 0) return nullptr;
  if (L->mode == 3) {
    if (!L->fresh) throw chan_error("prescribed subgrid stress not set (chan_les_set_tau)");
    return L->tau;
  }
  if (L->every || rk == 1 || !L->tau_valid) {
    les_compute(c, L, rk == 1);
    L->tau_valid = true;
  }
  return L->tau;
}

void les_prepare_capture(chan_ctx* c) {
  if (c->les && c->les->mode == 2) mlp_prepare_device(c->les->net);
}

extern "C" {

int chan_les_configure(chan_ctx* c, int mode, chan_mlp* net, double beta, int every_substep) {
  return guarded_les(c, [&] {
    if (mode < 0 || mode > 3) throw chan_error("LES mode must be 0 (off), 1, 2 or 3");
    if (mode == 2 && !net) throw chan_error("the learned closure needs a chan_mlp");
    if (beta != beta) throw chan_error("gate beta must be a number (negative: no energy gate)");
    check_ops_les(c);
    if (mode == 0) {
      les_destroy(c->les);
      c->les = nullptr;
      return;
    }
    if ((int)c->h_y.size() != c->prm.ny) throw chan_error("LES closure needs chan_set_coordinates");
    const XZGeom& g = c->gd;
    if (!(c->tune.xpass_wave && fft2_split(g.nx2 / 2).n1 > 0 && g.nx2 / 2 <= 192))
      throw chan_error("the subgrid closure needs the wavefront x-pass (nx2/2 <= 192 of the form 2^a 3^b)");
    les_destroy(c->les);
    c->les = nullptr;
    LesState* L = new LesState();
    c->les = L;
    const int ny = c->prm.ny, nyp = c->nyp;
    L->mode = mode;
    L->beta = beta;
    L->every = every_substep ? 1 : 0;
    L->net = net;
    L->ps = (long)g.nz2 * g.nx2;
    L->nm = (long)c->nxh * c->prm.nz;
    L->npts = (long)(ny - 1) * L->ps;
    L->nblk = (int)(((g.nz2 + DYN_T - 1) / DYN_T) * ((g.nx2 + DYN_T - 1) / DYN_T));
    const long zps = (long)g.nz2 * g.kxpad;
    lalloc(L->G, (size_t)LES_NF * nyp * L->nm);
    lalloc(L->Zl, (size_t)12 * nyp * zps);
    lalloc(L->Ph, (size_t)LES_NF * nyp * L->ps);
    lalloc(L->tau, (size_t)6 * nyp * L->ps);
    lalloc(L->cs2, nyp);
    lalloc(L->delta2, nyp);
    lalloc(L->ydist, nyp);
    lalloc(L->part, (size_t)2 * nyp * L->nblk);
    L->ngated = ((L->ps + 255) / 256) * (long)(ny - 1);
    lalloc(L->gated, (L->ngated + 1) & ~1L);
    if (mode == 2) {
      lalloc(L->X, (size_t)10 * L->npts);
      lalloc(L->Y, (size_t)6 * L->npts);
      lalloc(L->scale, L->npts);
      mlp_prepare_device(net);
    }
    // filter width D^2 and wall distance 1 - |yh| per plane (zero on the walls)
    const std::vector<double>& y = c->h_y;
    const double dx = c->prm.lx / c->prm.nx, dz = c->prm.lz / c->prm.nz;
    std::vector<double> d2(nyp, 0.0), yd(nyp, 0.0);
    for (int k = 1; k < ny; ++k) {
      const double d = std::cbrt(dx * dz * (y[k] - y[k - 1]));
      d2[k] = d * d;
      yd[k] = 1.0 - std::fabs(0.5 * (y[k - 1] + y[k]));
    }
    CHAN_CHECK(hipMemcpy(L->delta2, d2.data(), sizeof(double) * nyp, hipMemcpyHostToDevice));
    CHAN_CHECK(hipMemcpy(L->ydist, yd.data(), sizeof(double) * nyp, hipMemcpyHostToDevice));
    launch_fill_zero(L->cs2, sizeof(double) * nyp, c->stream);
    launch_fill_zero(L->tau, sizeof(double) * 6 * nyp * L->ps, c->stream);
    launch_fill_zero(L->gated, sizeof(unsigned int) * ((L->ngated + 1) & ~1L), c->stream);
    CHAN_CHECK(hipStreamSynchronize(c->stream));
  });
}

int chan_les_update(chan_ctx* c) {
  return guarded_les(c, [&] {
    LesState* L = les_need(c);
    if (L->mode == 3) return;
    les_prepare_state(c);
    les_compute(c, L, true);
    CHAN_CHECK(hipStreamSynchronize(c->stream));
  });
}

int chan_les_set_tau(chan_ctx* c, const double* tau) {
  return guarded_les(c, [&] {
    LesState* L = les_need(c);
    if (L->mode != 3) throw chan_error("chan_les_set_tau needs LES mode 3 (prescribed stress)");
    if (!tau) throw chan_error("null tau");
    CHAN_CHECK(hipMemcpy(L->tau, tau, sizeof(double) * 6 * c->nyp * L->ps, hipMemcpyHostToDevice));
    L->fresh = true;
  });
}

int chan_les_get_tau(chan_ctx* c, double* tau) {
  return guarded_les(c, [&] {
    LesState* L = les_need(c);
    if (!tau) throw chan_error("null tau");
    CHAN_CHECK(hipStreamSynchronize(c->stream));
    CHAN_CHECK(hipMemcpy(tau, L->tau, sizeof(double) * 6 * c->nyp * L->ps, hipMemcpyDeviceToHost));
  });
}

int chan_les_get_coefficient(chan_ctx* c, double* cs2) {
  return guarded_les(c, [&] {
    LesState* L = les_need(c);
    if (!cs2) throw chan_error("null cs2");
    CHAN_CHECK(hipStreamSynchronize(c->stream));
    CHAN_CHECK(hipMemcpy(cs2, L->cs2, sizeof(double) * c->nyp, hipMemcpyDeviceToHost));
  });
}

int chan_les_get_fields(chan_ctx* c, double* fields) {
  return guarded_les(c, [&] {
    LesState* L = les_need(c);
    if (!fields) throw chan_error("null fields");
    les_prepare_state(c);
    les_fields(c, L, false);
    CHAN_CHECK(hipStreamSynchronize(c->stream));
    CHAN_CHECK(hipMemcpy(fields, L->Ph, sizeof(double) * LES_NF * c->nyp * L->ps, hipMemcpyDeviceToHost));
  });
}

int chan_les_features(chan_ctx* c, float* x, double* scale) {
  return guarded_les(c, [&] {
    LesState* L = les_need(c);
    if (L->mode != 2) throw chan_error("chan_les_features needs LES mode 2 (learned closure)");
    if (!x || !scale) throw chan_error("null output");
    les_prepare_state(c);
    les_fields(c, L, false);
    dim3 gi((unsigned)((L->ps + 255) / 256), c->prm.ny - 1);
    hipLaunchKernelGGL(k_les_feat, gi, dim3(256), 0, c->stream, c->prm.ny, L->ps, L->Ph, L->ydist, L->delta2, L->X,
                       L->scale);
    LAUNCH_CHECK("k_les_feat");
    CHAN_CHECK(hipStreamSynchronize(c->stream));
    CHAN_CHECK(hipMemcpy(x, L->X, sizeof(float) * 10 * L->npts, hipMemcpyDeviceToHost));
    CHAN_CHECK(hipMemcpy(scale, L->scale, sizeof(double) * L->npts, hipMemcpyDeviceToHost));
  });
}

int chan_les_gated_points(chan_ctx* c, long* count, long* total) {
  return guarded_les(c, [&] {
    LesState* L = les_need(c);
    std::vector<unsigned int> g(L->ngated, 0u);
    CHAN_CHECK(hipStreamSynchronize(c->stream));
    if (L->mode == 2)
      CHAN_CHECK(hipMemcpy(g.data(), L->gated, sizeof(unsigned int) * L->ngated, hipMemcpyDeviceToHost));
    long sum = 0;
    for (unsigned int v : g) sum += v;
    if (count) *count = sum;
    if (total) *total = L->npts;
  });
}

}  // extern "C"
