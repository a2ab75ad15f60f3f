// Small device kernels behind the host-pointer entry points: batched
// SOLVE_HELMHOLTZ / SOLVE_P (PFC/solve.f90:154-338), the wall feedback law of
// boundary (PFC/read_write.f90:1483-1527), and the scale/reduce helpers of
// the Arnoldi matvec (PFC/arnoldi.f90:499-561).
#include "chan_internal.h"

// One thread per system: SOLVE_HELMHOLTZ with Dirichlet rows (the channel
// case, bl_grid == 0), in the reference's elimination order.  f complex(n,
// nsys) in place; scratch holds n doubles per system.
__global__ void k_helmholtz_batch(int n, int nsys, double2* f, const double* alpha, double diff,
                                  const double* d2, const double2* topbc, const double2* botbc,
                                  double* scr) {
  int s = blockIdx.x * blockDim.x + threadIdx.x;
  if (s >= nsys) return;
  double2* F = f + (long)s * n;
  double* C = scr + (long)s * n;
  const double al = alpha[s];
  double cp = 0.0; double2 fp = make_double2(0.0, 0.0);
  for (int i = 0; i < n; ++i) {
    double a, b, c; double2 ff;
    if (i == 0) { a = 0.0; b = 1.0; c = 0.0; ff = botbc[s]; }
    else if (i == n - 1) { a = 0.0; b = 1.0; c = 0.0; ff = topbc[s]; }
    else { a = -diff * d2[3 * i]; b = -diff * d2[3 * i + 1] + al; c = -diff * d2[3 * i + 2]; ff = F[i]; }
    double bp = (i == 0) ? b : b - a * cp;
    double2 g = (i == 0) ? ff : csub(ff, cscale(a, fp));
    double rb = 1.0 / bp;
    cp = c * rb;
    fp = cscale(rb, g);
    F[i] = fp;
    C[i] = cp;
  }
  double2 x = F[n - 1];
  for (int i = n - 2; i >= 0; --i) {
    x = csub(F[i], cscale(C[i], x));
    F[i] = x;
  }
}

// SOLVE_P for nsys systems with zero wall data; p complex(ny+1, nsys).
__global__ void k_solvep_batch(int ny, int nsys, double2* p, const double* alpha, const double* d2h,
                               const double* dh, const double* dn, const double* dm, double* scr) {
  int s = blockIdx.x * blockDim.x + threadIdx.x;
  if (s >= nsys) return;
  const int n = ny + 1;
  double2* F = p + (long)s * n;
  double* C = scr + (long)s * n;
  const double alp = alpha[s];
  double cp = 0.0; double2 fp = make_double2(0.0, 0.0);
  for (int i = 0; i < n; ++i) {
    double a_, b_, c_;
    if (i == 0) { a_ = 0.0; b_ = -dn[0]; c_ = dn[0]; }
    else if (i == 1) { a_ = dh[1] * dn[0]; b_ = -dh[1] * dn[0] - dh[1] * dn[1] + alp; c_ = dh[1] * dn[1]; }
    else if (i == ny - 1) {
      a_ = dh[ny - 1] * dn[ny - 2]; b_ = -dh[ny - 1] * dn[ny - 2] - dh[ny - 1] * dn[ny - 1] + alp;
      c_ = dh[ny - 1] * dn[ny - 1];
    } else if (i == ny) { a_ = -dn[ny - 1]; b_ = dn[ny - 1]; c_ = 0.0; }
    else { a_ = d2h[3 * i]; b_ = d2h[3 * i + 1] + alp; c_ = d2h[3 * i + 2]; }
    double2 rhs = (i == 0 || i == ny) ? make_double2(0.0, 0.0) : F[i];
    double bp = (i == 0) ? b_ : b_ - a_ * cp;
    double2 g = (i == 0) ? rhs : csub(rhs, cscale(a_, fp));
    double rb = 1.0 / bp;
    cp = c_ * rb;
    fp = cscale(rb, g);
    F[i] = fp;
    C[i] = cp;
  }
  double2 x = F[n - 1];
  for (int i = n - 2; i >= 0; --i) {
    x = csub(F[i], cscale(C[i], x));
    F[i] = x;
  }
  double2 p1 = F[1], p2 = F[2], pm2 = F[ny - 2], pm1 = F[ny - 1];
  F[0] = make_double2((0.0 - dm[1] * p1.x - dm[2] * p2.x) / dm[0], (0.0 - dm[1] * p1.y - dm[2] * p2.y) / dm[0]);
  F[ny] = make_double2((0.0 - dm[3] * pm2.x - dm[4] * pm1.x) / dm[5], (0.0 - dm[3] * pm2.y - dm[4] * pm1.y) / dm[5]);
}

// Feedback law: out = (s0*(dmA*p2 + dmB*p3) + add) / den  (pointwise on physical planes)
// the feedback law of both walls (blockIdx.y: law w0 or w1), pointwise on
// the physical wall-adjacent planes
__global__ void k_ctrl_law(int npts, CtrlLaw w0, CtrlLaw w1) {
  const CtrlLaw w = blockIdx.y ? w1 : w0;
  int i = blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= npts) return;
  w.out[i] = w.sgn * (w.ca * w.p2[i] + w.cb * w.p3[i] + w.add) / w.den;
}

// max |x| over n doubles into *out (bits of a nonnegative double)
__global__ void k_absmax(long n, const double* x, unsigned long long* out) {
  double mx = 0.0;
  for (long i = (long)blockIdx.x * blockDim.x + threadIdx.x; i < n; i += (long)gridDim.x * blockDim.x)
    mx = fmax(mx, fabs(x[i]));
  for (int o = 32; o > 0; o >>= 1) mx = fmax(mx, __shfl_xor(mx, o, 64));
  if ((threadIdx.x & 63) == 0) atomicMax(out, (unsigned long long)__double_as_longlong(mx));
}

// Scale each (c, j) plane of a [3][ny+1][plane] complex array by the weight
// factor of the Arnoldi pack (dir 0: scale/sqrt(w), or 0 where w == 0) or
// unpack (dir 1: sqrt(w)/scale); w = dy for v, dyh for u and w.
__global__ void k_arnoldi_scale(int ny, long plane, double2* x, const double* dy, const double* dyh,
                                const unsigned long long* vmax, double eps, int dir) {
  const int nyp = ny + 1;
  const long tot = 3L * nyp * plane;
  const double scale = eps / __longlong_as_double((long long)*vmax);
  for (long e = (long)blockIdx.x * blockDim.x + threadIdx.x; e < tot; e += (long)gridDim.x * blockDim.x) {
    long cj = e / plane;
    int c = (int)(cj / nyp), j = (int)(cj % nyp);
    double w = (c == 1) ? dy[j] : dyh[j];
    double f;
    if (dir == 0) f = (w != 0.0) ? scale / sqrt(w) : 0.0;
    else f = sqrt(w) / scale;
    x[e] = cscale(f, x[e]);
  }
}

void launch_helmholtz_batch(int n, int nsys, double2* f, const double* alpha, double diff,
                            const double* d2, const double2* topbc, const double2* botbc, double* scr,
                            hipStream_t s) {
  hipLaunchKernelGGL(k_helmholtz_batch, dim3((nsys + 63) / 64), dim3(64), 0, s, n, nsys, f, alpha, diff,
                     d2, topbc, botbc, scr);
  LAUNCH_CHECK("k_helmholtz_batch");
}

void launch_solvep_batch(int ny, int nsys, double2* p, const double* alpha, const DevOps& ops,
                         double* scr, hipStream_t s) {
  hipLaunchKernelGGL(k_solvep_batch, dim3((nsys + 63) / 64), dim3(64), 0, s, ny, nsys, p, alpha, ops.d2h,
                     ops.dh, ops.dn, ops.dm, scr);
  LAUNCH_CHECK("k_solvep_batch");
}

// blockIdx.y: component c0 or c1; one thread per (kx, kz) mode m = ix nz + iz
__global__ void k_ctrl_spec(long nm, int nz, CtrlSpec c0, CtrlSpec c1) {
  const CtrlSpec c = blockIdx.y ? c1 : c0;
  const long m = (long)blockIdx.x * blockDim.x + threadIdx.x;
  if (m >= nm) return;
  const int iz = (int)(m % nz);
  const bool kx0 = m < nz;
  auto law = [&](const double2* p2, const double2* p3, double ca, double cb, double add, double sgn, double den) {
    auto lin = [&](long k) {
      const double2 a = p2[k], b = p3[k];
      return make_double2(ca * a.x + cb * b.x, ca * a.y + cb * b.y);
    };
    double2 v = lin(m);
    if (kx0 && iz == 0) {
      v.y = 0.0;
      v.x += add;
    } else if (kx0 && 2 * iz != nz) {
      const double2 w = lin(nz - iz);  // the partner mode -kz of the kx = 0 column
      v = make_double2(0.5 * (v.x + w.x), 0.5 * (v.y - w.y));
    }
    return make_double2(sgn * v.x / den, sgn * v.y / den);
  };
  c.outb[m] = law(c.p2b, c.p3b, c.cab, c.cbb, c.addb, 1.0, c.denb);
  c.outt[m] = law(c.p2t, c.p3t, c.cat, c.cbt, c.addt, -1.0, c.dent);
}

void launch_ctrl_spec(long nm, int nz, const CtrlSpec& c0, const CtrlSpec& c1, int ncomp, hipStream_t s) {
  hipLaunchKernelGGL(k_ctrl_spec, dim3((unsigned)((nm + 255) / 256), ncomp), dim3(256), 0, s, nm, nz, c0, c1);
  LAUNCH_CHECK("k_ctrl_spec");
}

void launch_ctrl_law(int npts, const CtrlLaw& w0, const CtrlLaw& w1, hipStream_t s) {
  hipLaunchKernelGGL(k_ctrl_law, dim3((npts + 255) / 256, 2), dim3(256), 0, s, npts, w0, w1);
  LAUNCH_CHECK("k_ctrl_law");
}

void launch_absmax(long n, const double* x, unsigned long long* out, hipStream_t s) {
  unsigned blocks = (unsigned)std::min<long>((n + 255) / 256, 2048);
  hipLaunchKernelGGL(k_absmax, dim3(blocks), dim3(256), 0, s, n, x, out);
  LAUNCH_CHECK("k_absmax");
}

void launch_arnoldi_scale(int ny, long plane, double2* x, const double* dy, const double* dyh,
                          const unsigned long long* vmax, double eps, int dir, hipStream_t s) {
  long tot = 3L * (ny + 1) * plane;
  unsigned blocks = (unsigned)std::min<long>((tot + 255) / 256, 4096);
  hipLaunchKernelGGL(k_arnoldi_scale, dim3(blocks), dim3(256), 0, s, ny, plane, x, dy, dyh, vmax, eps, dir);
  LAUNCH_CHECK("k_arnoldi_scale");
}
