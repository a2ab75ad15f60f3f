// Filtered DNS on the device (no counterpart in the reference, which has no
// working LES path): the velocity of a fine channel filtered to a coarse
// channel of the same box, and the exact subgrid stress of that filter, for
// training and judging subgrid closures.
//
// The filter F of a fine spectral field, onto the coarse grid whose faces are
// every other fine face (ny_f = 2 ny_c - 1, same tanh stretching):
//   x, z: the 3-point box average of the fine physical grid (transfer
//         (1 + 2 cos(k dx)) / 3 per direction), then the cut-off to the
//         coarse stored modes without the coarse Nyquist modes;
//   y:    at the cell centres, the average over the two fine cells of a
//         coarse cell weighted by their heights (fine cells 2K-1 and 2K make
//         coarse cell K); at the faces, the shared faces (fine face 2j).
// The restriction keeps the discrete divergence of the filtered field zero.
// The exact stress at the coarse cell centres is
//   tau_ij = F(u_i u_j) - F(u_i) F(u_j)
// with u_i at the fine cell centres (v as the mean of its two faces), the
// products formed on the fine dealiased grid and F(u_i) the filtered
// cell-centre fields; tau is zero on the wall planes.
#include "chan_internal.h"
#include <cmath>

namespace {

// One coarse plane K (blockIdx.y) of one field: in [nyf+1][nxhf][nzf] ->
// out [nyc+1][nxhc][nzc].  kind 0: cell centres, wall planes copied; 1: cell
// centres, wall planes zero; 2: faces (fine face 2K; plane nyc zero).
__global__ void k_filter_spec(const double2* __restrict__ in, int nyf, int nxhf, int nzf, int nxf,
                              double2* __restrict__ out, int nyc, int nxhc, int nzc,
                              const double* __restrict__ hhf, int kind) {
  const long nmc = (long)nxhc * nzc;
  const long m = (long)blockIdx.x * blockDim.x + threadIdx.x;
  const int K = blockIdx.y;
  if (m >= nmc) return;
  const int ix = (int)(m / nzc), iz = (int)(m - (long)ix * nzc);
  const int kzi = (iz <= nzc / 2) ? iz : iz - nzc;
  const bool keep = (ix < nxhc - 1) && (kzi < nzc / 2) && (kzi > -nzc / 2);
  const int src = (kzi >= 0) ? kzi : nzf + kzi;
  const double pi = 3.14159265358979323846;
  const double tx = (1.0 + 2.0 * cos(2.0 * pi * (double)ix / (double)nxf)) / 3.0;
  const double tz = (1.0 + 2.0 * cos(2.0 * pi * (double)kzi / (double)nzf)) / 3.0;
  const double f = keep ? tx * tz : 0.0;
  const long nmf = (long)nxhf * nzf;
  auto F = [&](int k) {
    const double2 v = in[(long)k * nmf + (long)ix * nzf + src];
    return make_double2(f * v.x, f * v.y);
  };
  double2 r = make_double2(0.0, 0.0);
  if (kind == 2) {
    if (K < nyc) r = F(2 * K);
  } else if (K == 0) {
    if (kind == 0) r = F(0);
  } else if (K == nyc) {
    if (kind == 0) r = F(nyf);
  } else {
    const double a = hhf[2 * K - 1], b = hhf[2 * K];
    const double2 p = F(2 * K - 1), q = F(2 * K);
    r = make_double2((a * p.x + b * q.x) / (a + b), (a * p.y + b * q.y) / (a + b));
  }
  out[(long)K * nmc + m] = r;
}

// v at the fine cell centres: planes 1..ny-1 the mean of the two faces, the
// wall planes zero.  [ny+1][nm]
__global__ void k_v_centres(const double2* __restrict__ v, int ny, long nm, double2* __restrict__ vc) {
  const long m = (long)blockIdx.x * blockDim.x + threadIdx.x;
  const int k = blockIdx.y;
  if (m >= nm) return;
  double2 r = make_double2(0.0, 0.0);
  if (k >= 1 && k <= ny - 1) {
    const double2 a = v[(long)(k - 1) * nm + m], b = v[(long)k * nm + m];
    r = make_double2(0.5 * (a.x + b.x), 0.5 * (a.y + b.y));
  }
  vc[(long)k * nm + m] = r;
}

// out = a * b on planes 1..ny-1, zero on the wall planes; [ny+1][ps] reals
__global__ void k_product(const double* __restrict__ a, const double* __restrict__ b, int ny, long ps,
                          double* __restrict__ out) {
  const long i = (long)blockIdx.x * blockDim.x + threadIdx.x;
  const int k = blockIdx.y;
  if (i >= ps) return;
  const long o = (long)k * ps + i;
  out[o] = (k >= 1 && k <= ny - 1) ? a[o] * b[o] : 0.0;
}

// tau = Fp - ua * ub, zero on the wall planes
__global__ void k_tau_exact(const double* __restrict__ Fp, const double* __restrict__ ua,
                            const double* __restrict__ ub, int ny, long ps, double* __restrict__ tau) {
  const long i = (long)blockIdx.x * blockDim.x + threadIdx.x;
  const int k = blockIdx.y;
  if (i >= ps) return;
  const long o = (long)k * ps + i;
  tau[o] = (k >= 1 && k <= ny - 1) ? Fp[o] - ua[o] * ub[o] : 0.0;
}

struct DevBuf {
  void* p = nullptr;
  explicit DevBuf(size_t bytes) { CHAN_CHECK(hipMalloc(&p, bytes ? bytes : 8)); }
  ~DevBuf() { if (p) (void)hipFree(p); }
  template <class T> T* as() const { return static_cast<T*>(p); }
};

}  // namespace

void filter_dns_device(chan_ctx* les, chan_ctx* dns, double* tau_host,
                       void (*c2r)(chan_ctx*, const XZGeom&, const double2*, int, double*),
                       void (*r2c)(chan_ctx*, const XZGeom&, const double*, int, double2*)) {
  const int nyf = dns->prm.ny, nyc = les->prm.ny;
  if (2 * nyc - 1 != nyf) throw chan_error("filter: the fine channel needs ny = 2 ny_coarse - 1");
  if (les->prm.nx > dns->prm.nx || les->prm.nz > dns->prm.nz)
    throw chan_error("filter: the coarse channel has more modes than the fine one");
  if (std::fabs(les->prm.lx - dns->prm.lx) > 1e-12 * dns->prm.lx || std::fabs(les->prm.lz - dns->prm.lz) > 1e-12 * dns->prm.lz)
    throw chan_error("filter: the two channels need the same box");
  if ((int)dns->h_y.size() != nyf || (int)les->h_y.size() != nyc)
    throw chan_error("filter: both channels need their y coordinates (chan_set_coordinates or chan_build_operators)");
  for (int j = 0; j < nyc; ++j)
    if (std::fabs(les->h_y[j] - dns->h_y[2 * j]) > 1e-13)
      throw chan_error("filter: the coarse faces must be every other fine face");
  // everything runs in order on the fine channel's stream; the coarse
  // channel's own work is finished first and its stream restored at the end
  CHAN_CHECK(hipStreamSynchronize(les->stream));
  struct StreamSwap {
    chan_ctx* c;
    hipStream_t saved;
    ~StreamSwap() { c->stream = saved; }
  } swap{les, les->stream};
  les->stream = dns->stream;
  hipStream_t s = dns->stream;
  const long nmf = (long)dns->nxh * dns->prm.nz, nmc = (long)les->nxh * les->prm.nz;
  const int nyfp = nyf + 1, nycp = nyc + 1;
  std::vector<double> hh(nyfp, 0.0);
  for (int j = 1; j < nyf; ++j) hh[j] = dns->h_y[j] - dns->h_y[j - 1];
  DevBuf dhh(sizeof(double) * nyfp);
  CHAN_CHECK(hipMemcpyAsync(dhh.p, hh.data(), sizeof(double) * nyfp, hipMemcpyHostToDevice, s));
  const double2* vel = dns->d_vel;  // the fine velocity (dns must hold it, not the NL term)
  const int nxf = dns->prm.nx;
  auto filt = [&](const double2* in, double2* out, int kind) {
    dim3 g((unsigned)((nmc + 127) / 128), nycp);
    hipLaunchKernelGGL(k_filter_spec, g, dim3(128), 0, s, in, nyf, dns->nxh, dns->prm.nz, nxf, out, nyc, les->nxh,
                       les->prm.nz, dhh.as<double>(), kind);
    LAUNCH_CHECK("k_filter_spec");
  };
  // filtered velocity into the coarse channel
  double2* U = les->d_vel;
  filt(vel + 0L * nyfp * nmf, U + 0L * nycp * nmc, 0);
  filt(vel + 1L * nyfp * nmf, U + 1L * nycp * nmc, 2);
  filt(vel + 2L * nyfp * nmf, U + 2L * nycp * nmc, 0);
  les->data1_is_nl = false;
  if (tau_host) {
    const XZGeom& gf = dns->gd;
    const XZGeom& gc = les->gd;
    const long psf = (long)gf.nz2 * gf.nx2, psc = (long)gc.nz2 * gc.nx2;
    // fine cell-centre velocity: u, vc, w (spectral) and physical
    DevBuf vcb(sizeof(double2) * nyfp * nmf);
    {
      dim3 g((unsigned)((nmf + 127) / 128), nyfp);
      hipLaunchKernelGGL(k_v_centres, g, dim3(128), 0, s, vel + 1L * nyfp * nmf, nyf, nmf, vcb.as<double2>());
      LAUNCH_CHECK("k_v_centres");
    }
    const double2* spec[3] = {vel, vcb.as<double2>(), vel + 2L * nyfp * nmf};
    DevBuf phf(sizeof(double) * 3 * nyfp * psf);
    for (int c = 0; c < 3; ++c) c2r(dns, gf, spec[c], nyfp, phf.as<double>() + (long)c * nyfp * psf);
    // filtered cell-centre velocities on the coarse physical grid (walls zero)
    DevBuf ubs(sizeof(double2) * nycp * nmc);
    DevBuf ub(sizeof(double) * 3 * nycp * psc);
    for (int c = 0; c < 3; ++c) {
      filt(spec[c], ubs.as<double2>(), 1);
      c2r(les, gc, ubs.as<double2>(), nycp, ub.as<double>() + (long)c * nycp * psc);
    }
    // the six products, filtered, and tau
    DevBuf prod(sizeof(double) * nyfp * psf);
    DevBuf ps(sizeof(double2) * nyfp * nmf);
    DevBuf fps(sizeof(double2) * nycp * nmc);
    DevBuf fph(sizeof(double) * nycp * psc);
    DevBuf tau(sizeof(double) * 6 * nycp * psc);
    const int ti[6] = {0, 0, 0, 1, 1, 2}, tj[6] = {0, 1, 2, 1, 2, 2};
    for (int t = 0; t < 6; ++t) {
      dim3 g((unsigned)((psf + 255) / 256), nyfp);
      hipLaunchKernelGGL(k_product, g, dim3(256), 0, s, phf.as<double>() + (long)ti[t] * nyfp * psf,
                         phf.as<double>() + (long)tj[t] * nyfp * psf, nyf, psf, prod.as<double>());
      LAUNCH_CHECK("k_product");
      r2c(dns, gf, prod.as<double>(), nyfp, ps.as<double2>());
      filt(ps.as<double2>(), fps.as<double2>(), 1);
      c2r(les, gc, fps.as<double2>(), nycp, fph.as<double>());
      dim3 gt((unsigned)((psc + 255) / 256), nycp);
      hipLaunchKernelGGL(k_tau_exact, gt, dim3(256), 0, s, fph.as<double>(),
                         ub.as<double>() + (long)ti[t] * nycp * psc, ub.as<double>() + (long)tj[t] * nycp * psc,
                         nyc, psc, tau.as<double>() + (long)t * nycp * psc);
      LAUNCH_CHECK("k_tau_exact");
    }
    CHAN_CHECK(hipMemcpyAsync(tau_host, tau.p, sizeof(double) * 6 * nycp * psc, hipMemcpyDeviceToHost, s));
    CHAN_CHECK(hipStreamSynchronize(s));  // before the work buffers are released
  }
  CHAN_CHECK(hipStreamSynchronize(s));
}
