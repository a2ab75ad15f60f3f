// Diagnostics on the device: the l2 deviation from the laminar profile
// (l2_norm, PFC/read_write.f90:1432-1462), the per-mode kinetic energy
// (PFC/pass2.f90:113-118) and the per-plane velocity moments of
// compute_velocity_moments (PFC/stats.f90:423-440) for one field.
#include "chan_internal.h"

// Sum over the workgroup (256 threads); every thread gets the total.
__device__ __forceinline__ double block_sum(double v, double* red) {
  for (int o = 32; o > 0; o >>= 1) v += __shfl_xor(v, o, 64);
  const int w = threadIdx.x >> 6, l = threadIdx.x & 63;
  __syncthreads();
  if (l == 0) red[w] = v;
  __syncthreads();
  double t = 0.0;
  for (int i = 0; i < (int)(blockDim.x >> 6); ++i) t += red[i];
  return t;
}

// One workgroup per plane j of one component: part[j] = dy(j) * sum (u - lam(j))^2
// over the nx2 x nz2 points of the plane.
__global__ __launch_bounds__(256) void k_l2_planes(const double* __restrict__ phys, long ps, const double* __restrict__ lam,
                                                   const double* __restrict__ dy, double* __restrict__ part) {
  __shared__ double red[4];
  const int j = blockIdx.x;
  const double* p = phys + (long)j * ps;
  const double l = lam ? lam[j] : 0.0;
  double s = 0.0;
  for (long i = threadIdx.x; i < ps; i += blockDim.x) {
    const double d = p[i] - l;
    s += d * d;
  }
  s = block_sum(s, red);
  if (threadIdx.x == 0) part[j] = s * dy[j];
}

// l2 = sqrt(sum of the 3*(ny+1) plane parts / (nx2 nz2) / 2), summed in
// component-then-plane order; drag = d_um_dy(1) - d_um_dy(ny).
__global__ void k_l2_final(const double* __restrict__ part, int n, double npts, const double* __restrict__ dudy, int ny,
                           double* __restrict__ out) {
  if (threadIdx.x != 0 || blockIdx.x != 0) return;
  double t = 0.0;
  for (int i = 0; i < n; ++i) t += part[i];
  out[0] = sqrt(t / npts / 2.0);
  out[1] = dudy[0] - dudy[ny - 1];
}

// KE of every mode: int_y(|u|^2 + |w|^2 + |v at yh|^2, dyh); written as
// ke(x, z) with x fastest.
__global__ __launch_bounds__(256) void k_mode_energy(DevOps ops, int ny, int nxh, int nz, const double2* __restrict__ vel,
                                                     double* __restrict__ ke) {
  const long nm = (long)nxh * nz;
  const long m = (long)blockIdx.x * blockDim.x + threadIdx.x;
  if (m >= nm) return;
  const int nyp = ny + 1;
  auto at = [&](int c, int j) { return vel[((long)c * nyp + j) * nm + m]; };
  auto a2 = [](double2 v) { return v.x * v.x + v.y * v.y; };
  double s = 0.0;
  for (int j = 0; j <= ny; ++j) {
    double t = a2(at(0, j)) + a2(at(2, j));
    if (j == 0) t += a2(at(1, 0));
    else if (j < ny) {
      const double2 a = at(1, j), b = at(1, j - 1);
      t += a2(make_double2(0.5 * (a.x + b.x), 0.5 * (a.y + b.y)));
    } else {
      t += a2(at(1, ny - 1));
    }
    s += t * ops.dyh[j];
  }
  const int ix = (int)(m / nz), iz = (int)(m - (long)ix * nz);
  ke[(long)iz * nxh + ix] = s;
}

// One workgroup per plane: mean, 2nd/3rd/4th central moments of u, v, w
// and <u_i' v'> over the nx x nz points; out(ny+1, 15) with the plane fastest.
__global__ __launch_bounds__(256) void k_plane_moments(const double* __restrict__ phys, long ps, long cs, int nyp,
                                                       double* __restrict__ out) {
  __shared__ double red[4];
  const int j = blockIdx.x;
  const double sc = 1.0 / (double)ps;
  const double* u[3] = {phys + (long)j * ps, phys + cs + (long)j * ps, phys + 2 * cs + (long)j * ps};
  double m[3];
  for (int c = 0; c < 3; ++c) {
    double s = 0.0;
    for (long i = threadIdx.x; i < ps; i += blockDim.x) s += u[c][i];
    m[c] = sc * block_sum(s, red);
  }
  double acc[12];
  for (int k = 0; k < 12; ++k) acc[k] = 0.0;
  for (long i = threadIdx.x; i < ps; i += blockDim.x) {
    const double dv = u[1][i] - m[1];
    for (int c = 0; c < 3; ++c) {
      const double d = u[c][i] - m[c], d2 = d * d;
      acc[c] += d2;
      acc[3 + c] += d2 * d;
      acc[6 + c] += d2 * d2;
      acc[9 + c] += d * dv;
    }
  }
  for (int k = 0; k < 12; ++k) {
    const double t = sc * block_sum(acc[k], red);
    if (threadIdx.x == 0) out[(long)(3 + k) * nyp + j] = t;
  }
  if (threadIdx.x == 0)
    for (int c = 0; c < 3; ++c) out[(long)c * nyp + j] = m[c];
}

void launch_l2_planes(const double* phys, long ps, int nplanes, const double* lam, const double* dy, double* part,
                      hipStream_t s) {
  hipLaunchKernelGGL(k_l2_planes, dim3(nplanes), dim3(256), 0, s, phys, ps, lam, dy, part);
  LAUNCH_CHECK("k_l2_planes");
}

void launch_l2_final(const double* part, int n, double npts, const double* dudy, int ny, double* out, hipStream_t s) {
  hipLaunchKernelGGL(k_l2_final, dim3(1), dim3(64), 0, s, part, n, npts, dudy, ny, out);
  LAUNCH_CHECK("k_l2_final");
}

void launch_mode_energy(const DevOps& ops, int ny, int nxh, int nz, const double2* vel, double* ke, hipStream_t s) {
  const long nm = (long)nxh * nz;
  hipLaunchKernelGGL(k_mode_energy, dim3((unsigned)((nm + 255) / 256)), dim3(256), 0, s, ops, ny, nxh, nz, vel, ke);
  LAUNCH_CHECK("k_mode_energy");
}

void launch_plane_moments(const double* phys, long ps, long cs, int nyp, double* out, hipStream_t s) {
  hipLaunchKernelGGL(k_plane_moments, dim3(nyp), dim3(256), 0, s, phys, ps, cs, nyp, out);
  LAUNCH_CHECK("k_plane_moments");
}

// Spectral field of the VTK output: u and w at plane j, v at yh(j) as
// half*(v(j-1) + v(j)), v(1) at the bottom and v(ny) at the top plane
// (PFC/read_write.f90:418-432, yskip = 1).
__global__ __launch_bounds__(256) void k_vtk_field(int ny, long nm, const double2* __restrict__ vel,
                                                   double2* __restrict__ out) {
  const int nyp = ny + 1;
  const long tot = 3L * nyp * nm;
  for (long e = (long)blockIdx.x * blockDim.x + threadIdx.x; e < tot; e += (long)gridDim.x * blockDim.x) {
    const long m = e % nm, t = e / nm;
    const int j = (int)(t % nyp), c = (int)(t / nyp);
    double2 v;
    if (c != 1) v = vel[e];
    else if (j == 0) v = vel[(1L * nyp + 0) * nm + m];
    else if (j == ny) v = vel[(1L * nyp + ny - 1) * nm + m];
    else {
      const double2 a = vel[(1L * nyp + j - 1) * nm + m], b = vel[(1L * nyp + j) * nm + m];
      v = make_double2(0.5 * (a.x + b.x), 0.5 * (a.y + b.y));
    }
    out[e] = v;
  }
}

void launch_vtk_field(int ny, long nm, const double2* vel, double2* out, hipStream_t s) {
  const long tot = 3L * (ny + 1) * nm;
  const unsigned blocks = (unsigned)std::min<long>((tot + 255) / 256, 8192);
  hipLaunchKernelGGL(k_vtk_field, dim3(blocks), dim3(256), 0, s, ny, nm, vel, out);
  LAUNCH_CHECK("k_vtk_field");
}
