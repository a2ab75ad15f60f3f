// C-ABI of libchanflow: context, state transfer and the drop-in entry points
// for the transform / pass1 / pass2 / solve / pass1_linearized / arnoldi
// interfaces of the reference (each entry cites the interface it replaces in
// include/chanflow.h).  Every compute entry runs on the GPU; there is no host
// fallback.
#include <cmath>
#include <cstring>
#include <map>
#include <mutex>
#include <algorithm>
#include "chan_internal.h"
#include <array>

namespace {

thread_local std::string g_create_error;

inline int fail(chan_ctx* c, const std::string& m) {
  if (c) c->last_error = m;
  else g_create_error = m;
  return 1;
}

template <class F>
int guarded(chan_ctx* c, F&& f) {
  if (!c) return fail(nullptr, "null context");
  try {
    f();
    return 0;
  } catch (const chan_error& e) {
    return fail(c, e.msg);
  } catch (const std::exception& e) {
    return fail(c, e.what());
  }
}

struct KTimer {
  chan_ctx* c;
  int kind;
  hipEvent_t a{}, b{};
  KTimer(chan_ctx* c_, int k) : c(c_), kind(k) {
    if (c->timing) {
      CHAN_CHECK(hipEventCreate(&a));
      CHAN_CHECK(hipEventCreate(&b));
      CHAN_CHECK(hipEventRecord(a, c->stream));
    }
  }
  ~KTimer() {
    if (c->timing) {
      (void)hipEventRecord(b, c->stream);
      c->ev.push_back(a);
      c->ev.push_back(b);
      c->ev_kind.push_back(kind);
    }
  }
};

bool smooth23(int n) {
  if (n < 2) return false;
  while (n % 2 == 0) n /= 2;
  while (n % 3 == 0) n /= 3;
  return n == 1;
}

int roundup(int a, int m) { return ((a + m - 1) / m) * m; }

void* dalloc(size_t bytes) {
  void* p = nullptr;
  if (bytes == 0) bytes = 8;
  CHAN_CHECK(hipMalloc(&p, bytes));
  return p;
}

template <class T>
void dfree(T*& p) {
  if (p) (void)hipFree((void*)p);
  p = nullptr;
}

void* stage(chan_ctx* c, size_t bytes) {
  if (bytes > c->stage_bytes) {
    // queued work may still read the old buffer
    CHAN_CHECK(hipStreamSynchronize(c->stream));
    if (c->d_stage) (void)hipFree(c->d_stage);
    c->d_stage = nullptr;
    c->stage_bytes = 0;
    c->d_stage = dalloc(bytes);
    c->stage_bytes = bytes;
  }
  return c->d_stage;
}

// All host transfers go through a pinned buffer owned by the context
// (hipHostMalloc): user arrays are never handed to the DMA engines directly,
// so the runtime never has to pin pageable memory on the fly.
void* pinned(chan_ctx* c, size_t bytes) {
  if (bytes > c->pin_bytes) {
    if (c->h_pin) (void)hipHostFree(c->h_pin);
    c->h_pin = nullptr;
    c->pin_bytes = 0;
    size_t want = std::max<size_t>(bytes, 1 << 20);
    CHAN_CHECK(hipHostMalloc(&c->h_pin, want, hipHostMallocDefault));
    c->pin_bytes = want;
  }
  return c->h_pin;
}
void h2d(chan_ctx* c, void* d, const void* h, size_t bytes) {
  if (bytes == 0) return;
  void* p = pinned(c, bytes);
  std::memcpy(p, h, bytes);
  CHAN_CHECK(hipMemcpyAsync(d, p, bytes, hipMemcpyHostToDevice, c->stream));
  CHAN_CHECK(hipStreamSynchronize(c->stream));
}
void d2h(chan_ctx* c, void* h, const void* d, size_t bytes) {
  if (bytes == 0) return;
  void* p = pinned(c, bytes);
  CHAN_CHECK(hipMemcpyAsync(p, d, bytes, hipMemcpyDeviceToHost, c->stream));
  CHAN_CHECK(hipStreamSynchronize(c->stream));
  std::memcpy(h, p, bytes);
}
// carve 256-byte aligned sub-buffers out of a staging allocation
inline char* carve(char*& cur, size_t bytes) {
  char* r = cur;
  cur += (bytes + 255) & ~size_t(255);
  return r;
}
void d2d(chan_ctx* c, void* d, const void* s, size_t bytes) {
  CHAN_CHECK(hipMemcpyAsync(d, s, bytes, hipMemcpyDeviceToDevice, c->stream));
}

void check_ops(chan_ctx* c) {
  if (!c->d_ops) throw chan_error("operators not uploaded (chan_upload_operators / chan_build_operators)");
}

long nmodes(chan_ctx* c) { return (long)c->nxh * c->prm.nz; }
size_t vel_bytes(chan_ctx* c) { return sizeof(double2) * 3 * c->nyp * nmodes(c); }
size_t p_bytes(chan_ctx* c) { return sizeof(double2) * c->nyp * nmodes(c); }

// -------- device pipelines --------
void dev_c2r(chan_ctx* c, const XZGeom& g, const double2* in, int np, double* out) {
  const long in_ps = (long)g.nxh * g.nz, zps = (long)g.nz2 * g.kxpad, ops_ = (long)g.nz2 * g.nx2;
  launch_zinv(g, c->d_tw, in, in_ps, c->d_Z, zps, np, c->stream);
  launch_xc2r(g, c->d_tw, c->d_Z, zps, out, ops_, np, c->stream);
}

void dev_r2c(chan_ctx* c, const XZGeom& g, const double* in, int np, double2* out) {
  const long ips = (long)g.nz2 * g.nx2, aps = (long)g.nz2 * g.kxpad, ops_ = (long)g.nxh * g.nz;
  launch_xr2c(g, c->d_tw, in, nullptr, c->ops.kx, ips, c->d_AB, aps, np, c->stream);
  launch_zfwd(g, c->d_tw, c->d_AB, nullptr, aps, c->ops.kz, out, ops_, np, 0, c->stream);
}

void dev_div(chan_ctx* c, const double* fx, const double* fy, const double* fz, int np, double2* out) {
  const XZGeom& g = c->gd;
  const long ips = (long)g.nz2 * g.nx2, aps = (long)g.nz2 * g.kxpad, ops_ = (long)g.nxh * g.nz;
  double2* A = c->d_AB;
  double2* B = c->d_AB + (long)np * aps;
  launch_xr2c(g, c->d_tw, fy, fx, c->ops.kx, ips, A, aps, np, c->stream);
  launch_xr2c(g, c->d_tw, fz, nullptr, c->ops.kx, ips, B, aps, np, c->stream);
  launch_zfwd(g, c->d_tw, A, B, aps, c->ops.kz, out, ops_, np, 0, c->stream);
}

// nonlinear / nonlinear_linearized: d_vel -> d_nl
void nl_pass(chan_ctx* c, int step, int rk, bool lin, bool cfl) {
  const XZGeom& g = c->gd;
  const int nyp = c->nyp, ny = c->prm.ny;
  const long zps = (long)g.nz2 * g.kxpad;
  if (lin && !c->d_base) throw chan_error("base flow not set (chan_set_baseflow)");
  const double* tau = nullptr;
  if (!lin && c->les) {
    KTimer t(c, KT_OTHER);
    tau = les_tau_for_pass(c, rk);
  }
  auto xpass = [&](int ka, int kb) {
    KTimer t(c, KT_XPASS);
    XPassArgs a{};
    a.Z = c->d_Z; a.AB = c->d_AB; a.base = lin ? c->d_base : nullptr;
    a.cflbits = cfl ? c->d_cfl : nullptr;
    a.ny = ny; a.ka = ka; a.kb = kb; a.chunk = c->tune.xpass_chunk; a.rz = 0;
    a.wave = c->tune.xpass_wave; a.maxthr = c->tune.xpass_threads; a.prefetch = c->tune.xpass_prefetch;
    a.tau = tau; a.rows = c->tune.xpass_rows;
    launch_xpass(g, c->d_tw, c->ops, a, c->stream);
  };
  const int nch = std::max(1, std::min(c->tune.nl_chunks, ny - 1));
  if (nch == 1) {
    {
      KTimer t(c, KT_ZINV);
      launch_zinv(g, c->d_tw, c->d_vel, nmodes(c), c->d_Z, zps, 3 * nyp, c->stream, c->tune.zinv_kx);
    }
    xpass(2, ny);
    {
      KTimer t(c, KT_ZFWD);
      launch_zfwd(g, c->d_tw, c->d_AB, c->d_AB + zps, 2 * zps, c->ops.kz, c->d_nl, nmodes(c), 3 * nyp, ny,
                  c->stream, c->tune.zfwd_v2, c->tune.zfwd_kx);
    }
  } else {
    // plane ranges in turn, so that each range's intermediates are re-read
    // soon after they are written; NL planes 1..ny-1 (0-based) are split,
    // planes 0 and ny (zero) go with the first and last range
    int done = -1;  // last z-inverse plane
    for (int ci = 0; ci < nch; ++ci) {
      const int a0 = 1 + (int)((long)(ny - 1) * ci / nch);
      const int b0 = (int)((long)(ny - 1) * (ci + 1) / nch);
      const int zlo = done + 1, zhi = b0 + 1;
      if (zhi >= zlo) {
        KTimer t(c, KT_ZINV);
        launch_zinv(g, c->d_tw, c->d_vel, nmodes(c), c->d_Z, zps, 3 * (zhi - zlo + 1), c->stream,
                    c->tune.zinv_kx, zhi - zlo + 1, nyp, zlo);
      }
      done = zhi;
      xpass(a0 + 1, b0 + 1);
      const int f0 = (ci == 0) ? 0 : a0, f1 = (ci == nch - 1) ? ny : b0;
      {
        KTimer t(c, KT_ZFWD);
        launch_zfwd(g, c->d_tw, c->d_AB, c->d_AB + zps, 2 * zps, c->ops.kz, c->d_nl, nmodes(c), 3 * (f1 - f0 + 1),
                    ny, c->stream, c->tune.zfwd_v2, c->tune.zfwd_kx, f1 - f0 + 1, nyp, f0);
      }
    }
  }
  if (cfl) launch_cfl_finalize(c->ops, c->prm, c->d_cfl, c->d_sc, step, c->stream);
}

void solve_pass(chan_ctx* c, int rk) {
  KTimer t(c, KT_SOLVE);
  SolveArgs a{};
  a.nl = c->d_nl; a.vel = c->d_vel; a.R = c->d_R; a.P = c->d_P; a.cbc = c->d_cbc; a.dudy = c->d_dudy;
  a.sc = c->d_sc; a.scr = c->d_scr; a.rk = rk; a.ny = c->prm.ny; a.nxh = c->nxh; a.nz = c->prm.nz;
  a.re = c->prm.re; a.cosb = std::cos(c->prm.beta); a.sinb = std::sin(c->prm.beta);
  a.fix_mass = c->prm.fix_mass;
  a.lanes = c->tune.solve_lanes; a.bl = c->tune.solve_bl; a.dma = c->tune.solve_dma;
  a.seg = c->tune.solve_seg;
  if (c->d_jet) {  // the jet overwrites the wall-normal velocity at its wall (PFC/pass2.f90:42-43)
    const long nm = nmodes(c);
    launch_jet_bc(c->d_sc, rk, c->jet_forcing, c->jet_freq, c->d_jet,
                  c->d_cbc + ((c->jet_wall == 0 ? 1L : 0L) * 3 + 1) * nm, nm, c->stream);
  }
  launch_solve(c->ops, a, c->stream);
}

__global__ void k_advance_time(DevScalars* sc) {
  if (threadIdx.x == 0 && blockIdx.x == 0) sc->t_total = sc->t_total + sc->dt;
}

// boundary (PFC/read_write.f90:1465-1532)
void boundary_dev(chan_ctx* c) {
  KTimer t(c, KT_OTHER);
  const long nm = nmodes(c);
  launch_fill_zero(c->d_cbc, sizeof(double2) * 6 * nm, c->stream);
  const double gains[2] = {c->prm.ctrlu, c->prm.ctrlw};
  if (gains[0] == 0.0 && gains[1] == 0.0) return;
  const XZGeom& g = c->gd;
  const int ny = c->prm.ny;
  const long np = (long)g.nz2 * g.nx2;
  const double* dm = c->h_dm;
  if (g.nx2 > c->prm.nx && g.nz2 > c->prm.nz) {
    // The law is pointwise and affine in the wall-adjacent velocities, and on
    // a dealiased grid the physical-space round trip (c2r, law, r2c) leaves a
    // spectrum as it is except for the kx = 0 column, of which it keeps the
    // Hermitian part: the law is applied to the Fourier coefficients with that
    // projection, one launch for every controlled component and both walls.
    CtrlSpec cs[2];
    int n = 0;
    for (int q = 0; q < 2; ++q) {
      if (gains[q] == 0.0) continue;
      const int comp = (q == 0) ? 0 : 2;
      const double add = (q == 0) ? 2.0 : 0.0;
      const double2* base = c->d_vel + (long)comp * c->nyp * nm;
      // bottom: (dm(2,1) u2 + dm(3,1) u3 - add)/(gain - dm(1,1));
      // top: -(dm(2,2) u_ny + dm(1,2) u_{ny-1} + add)/(gain + dm(3,2))
      cs[n++] = CtrlSpec{base + 1 * nm, base + 2 * nm, base + (long)(ny - 1) * nm, base + (long)(ny - 2) * nm,
                         dm[1], dm[2], -add, gains[q] - dm[0], dm[4], dm[3], add, gains[q] + dm[5],
                         c->d_cbc + (1L * 3 + comp) * nm, c->d_cbc + (0L * 3 + comp) * nm};
    }
    launch_ctrl_spec(nm, c->prm.nz, cs[0], n == 2 ? cs[1] : cs[0], n, c->stream);
    return;
  }
  double* ph = (double*)stage(c, sizeof(double) * np * 6);
  const long zps = (long)g.nz2 * g.kxpad;
  for (int q = 0; q < 2; ++q) {
    if (gains[q] == 0.0) continue;
    const int comp = (q == 0) ? 0 : 2;
    const double add = (q == 0) ? 2.0 : 0.0;
    const double2* base = c->d_vel + (long)comp * c->nyp * nm;
    // planes 2,3 and ny-1, ny (0-based 1, 2, ny-2, ny-1) in one z-inverse:
    // grid row y is plane 1 + y % 2 + (ny - 3) (y / 2), written to the same
    // plane of the intermediate; then the x c2r of each pair, giving
    // ph[0..3] = u(1), u(2), u(ny-2), u(ny-1)
    if (ny >= 5) {
      launch_zinv(g, c->d_tw, base, nm, c->d_Z, zps, 4, c->stream, 8, 2, ny - 3, 1);
      launch_xc2r(g, c->d_tw, c->d_Z + 1 * zps, zps, ph + 0 * np, np, 2, c->stream);
      launch_xc2r(g, c->d_tw, c->d_Z + (long)(ny - 2) * zps, zps, ph + 2 * np, np, 2, c->stream);
    } else {  // the two pairs overlap
      dev_c2r(c, g, base + 1 * nm, 2, ph + 0 * np);
      dev_c2r(c, g, base + (long)(ny - 2) * nm, 2, ph + 2 * np);
    }
    // top (ph[4]): -(dm(2,2) u_ny + dm(1,2) u_{ny-1} + add)/(gain + dm(3,2));
    // bottom (ph[5]): (dm(2,1) u2 + dm(3,1) u3 - add)/(gain - dm(1,1))
    launch_ctrl_law((int)np, CtrlLaw{ph + 3 * np, ph + 2 * np, dm[4], dm[3], add, -1.0, gains[q] + dm[5], ph + 4 * np},
                    CtrlLaw{ph + 0 * np, ph + 1 * np, dm[1], dm[2], -add, 1.0, gains[q] - dm[0], ph + 5 * np},
                    c->stream);
    // both walls in one x r2c and one z-forward: the top into plane 0 and the
    // bottom into plane 3 of the intermediate and of cbc + comp, the wall
    // data's [wall][component] planes (top = wall 0, bottom = wall 1)
    launch_xr2c(g, c->d_tw, ph + 4 * np, nullptr, c->ops.kx, np, c->d_AB, 3 * zps, 2, c->stream);
    launch_zfwd(g, c->d_tw, c->d_AB, nullptr, zps, c->ops.kz, c->d_cbc + (long)comp * nm, nm, 2, 0, c->stream,
                1, 8, 1, 3, 0);
  }
}

void check_err(chan_ctx* c) {
  DevScalars sc;
  d2h(c, &sc, c->d_sc, sizeof(sc));
  if (sc.err) {
    int e = sc.err;
    sc.err = 0;
    h2d(c, c->d_sc, &sc, sizeof(sc));
    char buf[160];
    snprintf(buf, sizeof(buf), "%s (max CFL = %.6e)", e == 2 ? "DT IS TOO BIG" : "CFL IS TOO BIG", sc.cfl);
    throw chan_error(buf);
  }
}

void collect_times(chan_ctx* c) {
  if (c->ev.empty()) return;
  CHAN_CHECK(hipStreamSynchronize(c->stream));
  for (size_t i = 0; i + 1 < c->ev.size(); i += 2) {
    float ms = 0.f;
    CHAN_CHECK(hipEventElapsedTime(&ms, c->ev[i], c->ev[i + 1]));
    int k = c->ev_kind[i / 2];
    c->ktime_ms[k] += ms;
    c->kcount[k] += 1;
    (void)hipEventDestroy(c->ev[i]);
    (void)hipEventDestroy(c->ev[i + 1]);
  }
  c->ev.clear();
  c->ev_kind.clear();
}

void ensure_nl_in_nl(chan_ctx* c) {
  if (!c->data1_is_nl) d2d(c, c->d_nl, c->d_vel, vel_bytes(c));
}

void ensure_vel_in_vel(chan_ctx* c) {
  if (c->data1_is_nl) {
    d2d(c, c->d_vel, c->d_nl, vel_bytes(c));
    c->data1_is_nl = false;
  }
}

void init_flow(chan_ctx* c, bool lin) {
  check_ops(c);
  ensure_vel_in_vel(c);
  const int ny = c->prm.ny, nxh = c->nxh, nz = c->prm.nz;
  // remove_initial_divergence (PFC/init_vel.f90:518-536)
  launch_project(c->ops, ny, nxh, nz, c->d_vel, c->d_vel, nullptr, c->d_scr, 1.0, 0, nullptr, c->prm.re,
                 c->stream);
  // d_um_dy (PFC/init_vel.f90:333-342)
  launch_dudy(c->ops, ny, nxh, nz, c->d_vel, c->d_dudy, c->stream);
  // data2 = data1 (:344-348)
  d2d(c, c->d_R, c->d_vel, vel_bytes(c));
  // compute_initial_pressure: nonlinear(step=-1, rk_step=3) + solve_continuous_poisson
  nl_pass(c, -1, 3, lin, false);
  launch_project(c->ops, ny, nxh, nz, c->d_nl, nullptr, c->d_P, c->d_scr, 1.0, 1, c->d_vel, c->prm.re,
                 c->stream);
  c->data1_is_nl = false;
}

void roots_for(std::vector<double2>& roots, std::map<int, int>& off, int n) {
  if (off.count(n)) return;
  off[n] = (int)roots.size();
  for (int t = 0; t < n; ++t) {
    long double ang = -2.0L * 3.14159265358979323846264338327950288L * (long double)t / (long double)n;
    roots.push_back(make_double2((double)cosl(ang), (double)sinl(ang)));
  }
}

// One iteration of the timestep loop (PFC/main_mpi.f90:73-90): boundary,
// three (nonlinear, rhs_solve) substeps, t_total += dt.  Kernels only, no
// host synchronisation, so it can be captured into a graph.
void step_body(chan_ctx* c, int step, bool lin) {
  boundary_dev(c);
  for (int rk = 1; rk <= 3; ++rk) {
    nl_pass(c, step, rk, lin, rk == 1);
    solve_pass(c, rk);
  }
  hipLaunchKernelGGL(k_advance_time, dim3(1), dim3(64), 0, c->stream, c->d_sc);
  LAUNCH_CHECK("k_advance_time");
}

// Everything step_body may allocate or synchronise for, done ahead of a capture.
void prepare_capture(chan_ctx* c) {
  check_ops(c);
  ensure_vel_in_vel(c);
  if (c->prm.ctrlu != 0.0 || c->prm.ctrlw != 0.0) (void)stage(c, sizeof(double) * 6 * (size_t)c->gd.nz2 * c->gd.nx2);
  if (c->timing) throw chan_error("disable timing before batching or tuning");
  les_prepare_capture(c);
  CHAN_CHECK(hipStreamSynchronize(c->stream));
}

}  // namespace

void ctx_set_error(chan_ctx* c, const std::string& m) { fail(c, m); }
void check_ops_les(chan_ctx* c) { check_ops(c); }
void les_prepare_state(chan_ctx* c) { ensure_vel_in_vel(c); }

// ---------------------------------------------------------------------------
// Tuning keys: 0 xpass_wave, 1 xpass_threads, 2 zfwd_v2, 3 solve_lanes, 4 solve_bl,
// 5 zinv_kx, 6 zfwd_kx, 7 xpass_prefetch, 8 solve_dma, 9 xpass_chunk, 10 nl_chunks,
// 11 solve_seg, 12 xpass_rows.
Tuning tuning_defaults() {
  Tuning t;
  auto env = [](const char* n, int& v) {
    if (const char* e = getenv(n)) v = atoi(e);
  };
  int v;
  v = t.xpass_wave; env("CHANFLOW_XPASS_WAVE", v); tuning_set(t, 0, v);
  v = t.xpass_threads; env("CHANFLOW_XPASS_THREADS", v); tuning_set(t, 1, v);
  v = t.zfwd_v2; env("CHANFLOW_ZFWD_V2", v); tuning_set(t, 2, v);
  v = t.solve_lanes; env("CHANFLOW_SOLVE_LANES", v); tuning_set(t, 3, v);
  v = t.solve_bl; env("CHANFLOW_SOLVE_BL", v); tuning_set(t, 4, v);
  v = t.zinv_kx; env("CHANFLOW_ZINV_KX", v); tuning_set(t, 5, v);
  v = t.zfwd_kx; env("CHANFLOW_ZFWD_KX", v); tuning_set(t, 6, v);
  v = t.xpass_prefetch; env("CHANFLOW_XPASS_PREFETCH", v); tuning_set(t, 7, v);
  v = t.solve_dma; env("CHANFLOW_SOLVE_DMA", v); tuning_set(t, 8, v);
  v = t.xpass_chunk; env("CHANFLOW_XPASS_CHUNK", v); tuning_set(t, 9, v);
  v = t.nl_chunks; env("CHANFLOW_NL_CHUNKS", v); tuning_set(t, 10, v);
  v = t.solve_seg; env("CHANFLOW_SOLVE_SEG", v); tuning_set(t, 11, v);
  v = t.xpass_rows; env("CHANFLOW_XPASS_ROWS", v); tuning_set(t, 12, v);
  return t;
}

bool tuning_set(Tuning& t, int key, int v) {
  switch (key) {
    case 0: if (v != 0 && v != 1) return false; t.xpass_wave = v; return true;
    case 1: if (v < 64 || v > 640) return false; t.xpass_threads = v; return true;
    case 2: if (v != 0 && v != 1) return false; t.zfwd_v2 = v; return true;
    case 3: if (v != 1 && v != 3) return false; t.solve_lanes = v; return true;
    case 4: if (v != 4 && v != 8 && v != 16) return false; t.solve_bl = v; return true;
    case 5: if (v != 4 && v != 8 && v != 16) return false; t.zinv_kx = v; return true;
    case 6: if (v != 4 && v != 8) return false; t.zfwd_kx = v; return true;
    case 7: if (v != 0 && v != 1) return false; t.xpass_prefetch = v; return true;
    case 8: if (v != 0 && v != 1) return false; t.solve_dma = v; return true;
    case 9: if (v != 0 && (v < 2 || v > 64)) return false; t.xpass_chunk = v; return true;
    case 10: if (v < 1 || v > 64) return false; t.nl_chunks = v; return true;
    case 11: if (v < 0 || v > 3) return false; t.solve_seg = v; return true;
    case 12: if (v != 0 && v != 1 && v != 2 && v != 4) return false; t.xpass_rows = v; return true;
    default: return false;
  }
}

int tuning_get(const Tuning& t, int key) {
  switch (key) {
    case 0: return t.xpass_wave;
    case 1: return t.xpass_threads;
    case 2: return t.zfwd_v2;
    case 3: return t.solve_lanes;
    case 4: return t.solve_bl;
    case 5: return t.zinv_kx;
    case 6: return t.zfwd_kx;
    case 7: return t.xpass_prefetch;
    case 8: return t.solve_dma;
    case 9: return t.xpass_chunk;
    case 10: return t.nl_chunks;
    case 11: return t.solve_seg;
    case 12: return t.xpass_rows;
    default: return -1;
  }
}

FFTPlan make_plan(int n, int tw_offset) {
  FFTPlan p{};
  p.n = n;
  p.tw = tw_offset;
  int a = 0, m = n;
  while (m % 2 == 0) { m /= 2; ++a; }
  int b = 0;
  while (m % 3 == 0) { m /= 3; ++b; }
  if (m != 1) throw chan_error("transform length " + std::to_string(n) + " is not of the form 2^a 3^b");
  // powers of two as radix-8 stages, the remainder as 4, 4x4 or 2
  std::vector<int> r;
  const int n8 = a / 3, rem = a % 3;
  if (rem == 0) { for (int i = 0; i < n8; ++i) r.push_back(8); }
  else if (rem == 2) { for (int i = 0; i < n8; ++i) r.push_back(8); r.push_back(4); }
  else if (n8 >= 1) { for (int i = 0; i < n8 - 1; ++i) r.push_back(8); r.push_back(4); r.push_back(4); }
  else r.push_back(2);
  for (int i = 0; i < b; ++i) r.push_back(3);
  if ((int)r.size() > FFT_MAXSTAGES) throw chan_error("transform length too large");
  p.nst = (int)r.size();
  for (int i = 0; i < p.nst; ++i) p.rad[i] = r[i];
  return p;
}

// One batched launch: the channels of one transform length (or segment
// length) and the work list that covers all their workgroups.
struct BatchLaunch {
  int len = 0;                 // nz2 (z-transforms), nx2/2 (x-pass) or rows per segment (solve)
  int rw = 1, nxh = 0;         // x-pass: z-rows per wavefront, largest nx/2+1 of the members
  void* args = nullptr;        // device array of per-channel arguments
  BatchItem* items = nullptr;  // device work list
  int nitems = 0;
};

struct chan_ensemble {
  std::vector<chan_ctx*> m;
  std::vector<hipStream_t> streams;
  std::vector<hipEvent_t> done;
  hipStream_t master = nullptr;
  hipEvent_t start = nullptr;
  hipGraph_t graph = nullptr;
  hipGraphExec_t exec = nullptr;
  int lin = 0;
  std::string last_error;
  // batched mode (nstreams == 0): every phase of a substage is one launch per
  // transform length over all members
  bool batched = false;
  std::vector<BatchLaunch> zinv, zfwd, xpass[2], solve[3];  // xpass[1]: with the CFL maxima; solve[rk-1]
  CflBatchArgs* cfl_args = nullptr;
  FillBatchArgs* fill_cbc = nullptr;
  DevScalars** scs = nullptr;
  size_t cbc_words = 0;
  std::vector<void*> dmem;     // device tables of the batched launches
  // the launches of one phase touch disjoint members: they run as parallel
  // branches of the graph (fork from the master stream, join back)
  std::vector<hipStream_t> bstreams;
  std::vector<hipEvent_t> bjoin;
  hipEvent_t bfork = nullptr;
  // x-pass plane ranges of the batched launches: each length's rows x planes
  // spread over its resident workgroups times xmul (0: the number of lengths)
  int xmul = 0;
  // launch-plan overrides the autotuner may choose: one z-row per x-pass
  // wavefront at every length; 4 modes per solve wavefront everywhere
  int xrows1 = 0, solve4 = 0;
};

namespace {

void ensemble_release(chan_ensemble* e) {
  if (e->exec) (void)hipGraphExecDestroy(e->exec);
  if (e->graph) (void)hipGraphDestroy(e->graph);
  e->exec = nullptr;
  e->graph = nullptr;
  for (void* p : e->dmem) (void)hipFree(p);
  e->dmem.clear();
  e->zinv.clear();
  e->zfwd.clear();
  for (auto& v : e->xpass) v.clear();
  for (auto& v : e->solve) v.clear();
}

// ---- batched mode ----
// Why a member cannot join the batched launches ("" if it can).
std::string batch_ineligible(const chan_ctx* c, bool lin) {
  const XZGeom& g = c->gd;
  if (c->les) return "an LES closure is configured";
  if (c->d_jet) return "a wall jet is configured";
  if (c->prm.ctrlu != 0.0 || c->prm.ctrlw != 0.0) return "wall feedback gains are set";
  if (lin && !c->d_base) return "no base flow (chan_set_baseflow)";
  if (!zt_batchable(g.nz2)) return "nz2 has no compile-time transform plan";
  if (!xpass_batchable(g.nx2 / 2)) return "nx2/2 has no wavefront x-pass";
  if (g.nz > 192) return "nz above 192";
  if (!solve_seg_rows(c->prm.ny)) return "ny+1 rows do not fit the segmented solve";
  if (3.0 * c->nyp * (double)c->nmodes * 16.0 >= 4.0e9) return "the segmented solve's 32-bit offsets";
  const Tuning& t = c->tune;
  if (!(t.xpass_wave == 1 && t.zfwd_v2 == 1 && t.zinv_kx == 8 && (t.solve_seg == 1 || t.solve_seg == 3) &&
        t.nl_chunks == 1 &&
        t.xpass_prefetch == 1))
    return "its launch shapes are not the defaults";
  if (c->timing) return "kernel timing is on";
  return "";
}

template <class A>
void* upload_table(chan_ensemble* e, const std::vector<A>& v) {
  void* d = dalloc(sizeof(A) * std::max<size_t>(v.size(), 1));
  e->dmem.push_back(d);
  if (!v.empty()) CHAN_CHECK(hipMemcpy(d, v.data(), sizeof(A) * v.size(), hipMemcpyHostToDevice));
  return d;
}

// Members grouped by a length; the group's args and work list.  items(i)
// appends member i's workgroups (bx, by) in launch order; larger members
// first so that the short ones fill the tail.
template <class A, class LenF, class ArgF, class ItemF, class WorkF>
std::vector<BatchLaunch> plan_batches(chan_ensemble* e, LenF len, ArgF arg, ItemF items, WorkF work) {
  std::map<int, std::vector<int>> by_len;
  for (int i = 0; i < (int)e->m.size(); ++i) by_len[len(e->m[i])].push_back(i);
  std::vector<BatchLaunch> out;
  for (auto& kv : by_len) {
    std::vector<int> mem = kv.second;
    std::stable_sort(mem.begin(), mem.end(), [&](int a, int b) { return work(e->m[a]) > work(e->m[b]); });
    std::vector<A> args;
    std::vector<BatchItem> it;
    for (int k = 0; k < (int)mem.size(); ++k) {
      args.push_back(arg(e->m[mem[k]]));
      items(e->m[mem[k]], (unsigned short)k, it);
    }
    BatchLaunch b;
    b.len = kv.first;
    b.args = upload_table(e, args);
    b.items = (BatchItem*)upload_table(e, it);
    b.nitems = (int)it.size();
    out.push_back(b);
  }
  return out;
}

void ensemble_plan_batched(chan_ensemble* e) {
  const bool lin = e->lin != 0;
  for (size_t i = 0; i < e->m.size(); ++i) {
    const std::string why = batch_ineligible(e->m[i], lin);
    if (!why.empty()) throw chan_error("member " + std::to_string(i) + " cannot join the batched launches: " + why);
  }
  if (e->m.size() > 65535) throw chan_error("too many members for one batched launch");
  auto zps = [](chan_ctx* c) { return (long)c->gd.nz2 * c->gd.kxpad; };
  auto push_grid = [](std::vector<BatchItem>& it, unsigned short k, int gx, int gy) {
    for (int y = 0; y < gy; ++y)
      for (int x = 0; x < gx; ++x) it.push_back(BatchItem{k, (unsigned short)x, (unsigned)y});
  };
  auto zwork = [](chan_ctx* c) { return (long)c->gd.nz2 * c->nxh * c->nyp; };
  // z-inverse: d_vel -> d_Z, every plane of every component
  e->zinv = plan_batches<ZBatchArgs>(
      e, [](chan_ctx* c) { return c->gd.nz2; },
      [&](chan_ctx* c) {
        ZBatchArgs a{};
        a.g = c->gd; a.tw = c->d_tw; a.in = c->d_vel; a.in_ps = nmodes(c); a.out = c->d_Z; a.out_ps = zps(c);
        a.pm = PlaneMap{3 * c->nyp, 0, 0};
        return a;
      },
      [&](chan_ctx* c, unsigned short k, std::vector<BatchItem>& it) {
        push_grid(it, k, (c->nxh + 7) / 8, 3 * c->nyp);
      }, zwork);
  // z-forward: A, B -> d_nl with the NL planes the reference leaves zero
  e->zfwd = plan_batches<ZBatchArgs>(
      e, [](chan_ctx* c) { return c->gd.nz2; },
      [&](chan_ctx* c) {
        ZBatchArgs a{};
        a.g = c->gd; a.tw = c->d_tw; a.in = c->d_AB; a.B = c->d_AB + zps(c); a.in_ps = 2 * zps(c);
        a.kz = c->ops.kz; a.out = c->d_nl; a.out_ps = nmodes(c); a.pm = PlaneMap{3 * c->nyp, 0, 0};
        a.nl_zero_ny = c->prm.ny;
        return a;
      },
      [&](chan_ctx* c, unsigned short k, std::vector<BatchItem>& it) {
        push_grid(it, k, (c->nxh + 7) / 8, 3 * c->nyp);
      }, zwork);
  // x-pass: plane ranges from the resident workgroups of the whole batch.
  // The launches of the L lengths run side by side and share the GPU, so
  // together they make L waves of workgroups if each length's z-row groups x
  // planes are spread over its resident workgroups; spreading them over L
  // times as many keeps the workgroups short, so that the last ones to start
  // leave a short tail (at the price of one halo plane per shorter range).
  // Each length takes the fewest z-rows per wavefront of its members' choices
  // (tuning key 12; 1, 2 or 4 each divide the larger), so that every member's
  // nz2 is a multiple of it.
  std::map<int, int> xrw, xnxh;  // per length: rows per wavefront, largest nx/2+1
  for (chan_ctx* c : e->m) {
    const int M = c->gd.nx2 / 2;
    const int r = e->xrows1 ? 1 : xpass_rows(M, c->gd.nz2, c->tune.xpass_rows, M <= 144, true);
    xrw[M] = xrw.count(M) ? std::min(xrw[M], r) : r;
    xnxh[M] = std::max(xnxh[M], c->nxh);
  }
  auto xgroups = [&](chan_ctx* c) { return (c->gd.nz2 / xrw[c->gd.nx2 / 2] + 3) / 4; };
  std::map<int, long> rows;  // per length: z-row groups x planes of all its members
  for (chan_ctx* c : e->m) rows[c->gd.nx2 / 2] += (long)xgroups(c) * (c->prm.ny - 1);
  std::map<int, int> chunk;
  const long nlen = (long)rows.size();
  for (auto& kv : rows) {
    const long res = xpass_batched_resident(kv.first, lin, xrw[kv.first], xnxh[kv.first]) *
                     (e->xmul > 0 ? (long)e->xmul : nlen);
    chunk[kv.first] = (int)std::max(2L, (kv.second + res - 1) / res);
  }
  for (int w = 0; w < 2; ++w) {
    e->xpass[w] = plan_batches<XBatchArgs>(
        e, [](chan_ctx* c) { return c->gd.nx2 / 2; },
        [&](chan_ctx* c) {
          XBatchArgs x{};
          x.g = c->gd; x.tw = c->d_tw; x.ops = c->ops;
          XPassArgs& a = x.a;
          a.Z = c->d_Z; a.AB = c->d_AB; a.base = lin ? c->d_base : nullptr;
          a.cflbits = w ? c->d_cfl : nullptr;
          a.ny = c->prm.ny; a.ka = 2; a.kb = c->prm.ny;
          a.chunk = std::min(chunk[c->gd.nx2 / 2], c->prm.ny - 1);
          a.rz = 0; a.wave = 1; a.maxthr = c->tune.xpass_threads; a.prefetch = 1; a.tau = nullptr;
          a.rows = xrw[c->gd.nx2 / 2];
          return x;
        },
        [&](chan_ctx* c, unsigned short k, std::vector<BatchItem>& it) {
          const int np = c->prm.ny - 1, ch = std::min(chunk[c->gd.nx2 / 2], np);
          push_grid(it, k, xgroups(c), (np + ch - 1) / ch);
        },
        [](chan_ctx* c) { return (long)c->gd.nz2 * c->gd.nx2 * c->nyp; });
    for (BatchLaunch& b : e->xpass[w]) {
      b.rw = xrw[b.len];
      b.nxh = xnxh[b.len];
    }
  }
  // the segmented solve, one table per substep
  auto layout = [e](chan_ctx* c) {
    return e->solve4 ? 64 + solve_seg_rows(c->prm.ny) : solve_seg_layout(c->prm.ny, c->tune.solve_seg);
  };
  for (int rk = 1; rk <= 3; ++rk) {
    e->solve[rk - 1] = plan_batches<SBatchArgs>(
        e, layout,
        [&](chan_ctx* c) {
          SBatchArgs x{};
          x.ops = c->ops;
          SolveArgs& a = x.a;
          a.nl = c->d_nl; a.vel = c->d_vel; a.R = c->d_R; a.P = c->d_P; a.cbc = c->d_cbc; a.dudy = c->d_dudy;
          a.sc = c->d_sc; a.scr = c->d_scr; a.rk = rk; a.ny = c->prm.ny; a.nxh = c->nxh; a.nz = c->prm.nz;
          a.re = c->prm.re; a.cosb = std::cos(c->prm.beta); a.sinb = std::sin(c->prm.beta);
          a.fix_mass = c->prm.fix_mass;
          a.lanes = c->tune.solve_lanes; a.bl = c->tune.solve_bl; a.dma = c->tune.solve_dma; a.seg = 1;
          return x;
        },
        [&](chan_ctx* c, unsigned short k, std::vector<BatchItem>& it) {
          const int modes = layout(c) / 16;
          push_grid(it, k, (int)((nmodes(c) + modes - 1) / modes), 1);
        },
        [](chan_ctx* c) { return (long)c->nmodes * c->nyp; });
  }
  std::vector<CflBatchArgs> cf;
  std::vector<FillBatchArgs> fc;
  std::vector<DevScalars*> sc;
  for (chan_ctx* c : e->m) {
    cf.push_back(CflBatchArgs{c->ops, c->prm, c->d_cfl, c->d_sc});
    const size_t cw = 6 * (size_t)nmodes(c) * sizeof(double2) / 8;
    fc.push_back(FillBatchArgs{(unsigned long long*)c->d_cbc, cw});
    e->cbc_words = std::max(e->cbc_words, cw);
    sc.push_back(c->d_sc);
  }
  e->cfl_args = (CflBatchArgs*)upload_table(e, cf);
  e->fill_cbc = (FillBatchArgs*)upload_table(e, fc);
  e->scs = (DevScalars**)upload_table(e, sc);
}

// One timestep of every member as batched launches (PFC/main_mpi.f90:73-90
// for all of them at once): boundary with zero gains (wall data cleared),
// three substeps of z-inverse, x-pass, z-forward (CFL and dt at the first)
// and solve, then t_total += dt.
void ensemble_step_batched(chan_ensemble* e, hipStream_t s) {
  const int n = (int)e->m.size();
  const bool lin = e->lin != 0;
  // one phase: its launches (one per length, disjoint members) side by side,
  // the longest length first (its workgroups are the longest), so that the
  // shorter ones fill in behind it
  auto phase = [&](const std::vector<BatchLaunch>& v, auto launch) {
    if (v.size() <= 1 || e->bstreams.empty()) {
      for (size_t i = v.size(); i-- > 0;) launch(v[i], s);
      return;
    }
    CHAN_CHECK(hipEventRecord(e->bfork, s));
    for (size_t j = 0; j < v.size(); ++j) {
      const size_t i = v.size() - 1 - j;
      hipStream_t bs = e->bstreams[j % e->bstreams.size()];
      CHAN_CHECK(hipStreamWaitEvent(bs, e->bfork, 0));
      launch(v[i], bs);
      CHAN_CHECK(hipEventRecord(e->bjoin[j % e->bjoin.size()], bs));
      CHAN_CHECK(hipStreamWaitEvent(s, e->bjoin[j % e->bjoin.size()], 0));
    }
  };
  launch_fill_zero_batched(e->fill_cbc, n, e->cbc_words, s);
  for (int rk = 1; rk <= 3; ++rk) {
    phase(e->zinv, [&](const BatchLaunch& b, hipStream_t st) {
      launch_zinv_batched(b.len, (const ZBatchArgs*)b.args, b.items, b.nitems, st);
    });
    phase(e->xpass[rk == 1 ? 1 : 0], [&](const BatchLaunch& b, hipStream_t st) {
      launch_xpass_batched(b.len, lin, b.rw, b.nxh, (const XBatchArgs*)b.args, b.items, b.nitems, st);
    });
    phase(e->zfwd, [&](const BatchLaunch& b, hipStream_t st) {
      launch_zfwd_batched(b.len, (const ZBatchArgs*)b.args, b.items, b.nitems, st);
    });
    if (rk == 1) launch_cfl_finalize_batched(e->cfl_args, n, 1, s);
    phase(e->solve[rk - 1], [&](const BatchLaunch& b, hipStream_t st) {
      launch_solve_batched(b.len, (const SBatchArgs*)b.args, b.items, b.nitems, st);
    });
  }
  launch_advance_time_batched(e->scs, n, s);
}

// Capture one timestep of every member: the members are spread round-robin
// over the ensemble's streams, which fork from and join into the master
// stream, so independent members run concurrently inside one graph.  In
// batched mode the step is the batched launches on the master stream.
void ensemble_capture(chan_ensemble* e) {
  for (chan_ctx* c : e->m) prepare_capture(c);
  if (e->batched) {
    if (e->zinv.empty()) ensemble_plan_batched(e);
    for (chan_ctx* c : e->m) c->data1_is_nl = false;
    CHAN_CHECK(hipStreamBeginCapture(e->master, hipStreamCaptureModeThreadLocal));
    try {
      ensemble_step_batched(e, e->master);
    } catch (...) {
      hipGraph_t g = nullptr;
      (void)hipStreamEndCapture(e->master, &g);
      if (g) (void)hipGraphDestroy(g);
      throw;
    }
    CHAN_CHECK(hipStreamEndCapture(e->master, &e->graph));
    CHAN_CHECK(hipGraphInstantiate(&e->exec, e->graph, nullptr, nullptr, 0));
    return;
  }
  CHAN_CHECK(hipStreamBeginCapture(e->master, hipStreamCaptureModeThreadLocal));
  std::vector<hipStream_t> saved(e->m.size());
  try {
    CHAN_CHECK(hipEventRecord(e->start, e->master));
    for (hipStream_t st : e->streams) CHAN_CHECK(hipStreamWaitEvent(st, e->start, 0));
    for (size_t i = 0; i < e->m.size(); ++i) {
      chan_ctx* c = e->m[i];
      saved[i] = c->stream;
      c->stream = e->streams[i % e->streams.size()];
      step_body(c, 1, e->lin != 0);
      c->stream = saved[i];
    }
    for (size_t k = 0; k < e->streams.size(); ++k) {
      CHAN_CHECK(hipEventRecord(e->done[k], e->streams[k]));
      CHAN_CHECK(hipStreamWaitEvent(e->master, e->done[k], 0));
    }
  } catch (...) {
    for (size_t i = 0; i < e->m.size(); ++i)
      if (saved[i]) e->m[i]->stream = saved[i];
    hipGraph_t g = nullptr;
    (void)hipStreamEndCapture(e->master, &g);
    if (g) (void)hipGraphDestroy(g);
    throw;
  }
  CHAN_CHECK(hipStreamEndCapture(e->master, &e->graph));
  CHAN_CHECK(hipGraphInstantiate(&e->exec, e->graph, nullptr, nullptr, 0));
}

}  // namespace

extern "C" {

const char* chan_version(void) { return "chanflow 1 gfx950 fp64"; }

const char* chan_last_error(const chan_ctx* ctx) {
  if (!ctx) return g_create_error.c_str();
  return ctx->last_error.c_str();
}

int chan_create(const chan_params* prm, chan_ctx** out) {
  if (!prm || !out) return fail(nullptr, "null argument");
  *out = nullptr;
  const chan_params& p = *prm;
  if (p.bl_grid != 0) return fail(nullptr, "bl_grid != 0 (boundary-layer grids) is not supported");
  if (p.nx < 4 || p.nz < 4 || p.nx % 2 || p.nz % 2) return fail(nullptr, "nx, nz must be even and >= 4");
  if (p.ny < 6) return fail(nullptr, "ny must be >= 6");
  if (p.nx2 < p.nx || p.nz2 < p.nz || p.nx2 % 2) return fail(nullptr, "need nx2 >= nx (even), nz2 >= nz");
  if (!smooth23(p.nx2 / 2) || !smooth23(p.nz2) || !smooth23(p.nz) || !smooth23(p.nx / 2))
    return fail(nullptr, "transform lengths must be of the form 2^a 3^b");
  if (p.re <= 0.0 || p.lx <= 0.0 || p.lz <= 0.0) return fail(nullptr, "invalid re/lx/lz");
  chan_ctx* c = new chan_ctx();
  c->prm = p;
  c->tune = tuning_defaults();
  try {
    c->nxh = p.nx / 2 + 1;
    c->nyp = p.ny + 1;
    c->nmodes = c->nxh * p.nz;
    std::vector<double2> roots;
    std::map<int, int> off;
    for (int n : {p.nz2, p.nx2 / 2, p.nx2, p.nz, p.nx / 2, p.nx}) roots_for(roots, off, n);
    auto geo = [&](XZGeom& g, int nx2, int nz2, int small) {
      g.nx = p.nx; g.nz = p.nz; g.nxh = c->nxh; g.nx2 = nx2; g.nz2 = nz2;
      g.kxpad = roundup(c->nxh, 8);
      g.small = small;
      g.pz = make_plan(nz2, off[nz2]);
      g.px = make_plan(nx2 / 2, off[nx2 / 2]);
      g.twx = off[nx2];
    };
    geo(c->gd, p.nx2, p.nz2, 0);
    geo(c->gs, p.nx, p.nz, 1);
    c->kxpad = c->gd.kxpad;
    CHAN_CHECK(hipGetDevice(&c->device));
    CHAN_CHECK(hipStreamCreateWithFlags(&c->stream, hipStreamNonBlocking));
    c->own_stream = true;
    c->d_tw = (double2*)dalloc(sizeof(double2) * roots.size());
    h2d(c, c->d_tw, roots.data(), sizeof(double2) * roots.size());
    const long nm = c->nmodes, nyp = c->nyp;
    c->d_vel = (double2*)dalloc(vel_bytes(c));
    c->d_nl = (double2*)dalloc(vel_bytes(c));
    c->d_R = (double2*)dalloc(vel_bytes(c));
    c->d_P = (double2*)dalloc(p_bytes(c));
    c->d_cbc = (double2*)dalloc(sizeof(double2) * 6 * nm);
    c->d_dudy = (double*)dalloc(sizeof(double) * p.ny);
    c->d_sc = (DevScalars*)dalloc(sizeof(DevScalars));
    c->d_cfl = (unsigned long long*)dalloc(sizeof(unsigned long long) * 3 * nyp);
    const long zsz = 3L * nyp * p.nz2 * c->kxpad;
    c->d_Z = (double2*)dalloc(sizeof(double2) * zsz);
    c->d_AB = (double2*)dalloc(sizeof(double2) * 2 * zsz);
    c->d_scr = (double2*)dalloc((sizeof(double2) * 4 + sizeof(double) * 4) * nyp * nm);
    CHAN_CHECK(hipMemset(c->d_vel, 0, vel_bytes(c)));
    CHAN_CHECK(hipMemset(c->d_nl, 0, vel_bytes(c)));
    CHAN_CHECK(hipMemset(c->d_R, 0, vel_bytes(c)));
    CHAN_CHECK(hipMemset(c->d_P, 0, p_bytes(c)));
    CHAN_CHECK(hipMemset(c->d_cbc, 0, sizeof(double2) * 6 * nm));
    CHAN_CHECK(hipMemset(c->d_dudy, 0, sizeof(double) * p.ny));
    // the CFL maxima start at zero here; the finalize kernel clears them
    // after each read
    CHAN_CHECK(hipMemset(c->d_cfl, 0, sizeof(unsigned long long) * 3 * nyp));
    DevScalars sc{};
    sc.dt = 0.25 * p.lx / (double)p.nx;       // PFC/setup.f90:1377
    sc.p_grad = -2.0 / p.re;                  // PFC/setup.f90:1384
    h2d(c, c->d_sc, &sc, sizeof(sc));
    c->h_dt = sc.dt;
    c->h_dt_known = true;
    CHAN_CHECK(hipDeviceSynchronize());
  } catch (const chan_error& e) {
    g_create_error = e.msg;
    chan_destroy(c);
    return 1;
  }
  *out = c;
  return 0;
}

int chan_destroy(chan_ctx* c) {
  if (!c) return 0;
  if (c->stream) (void)hipStreamSynchronize(c->stream);
  for (auto& e : c->ev) (void)hipEventDestroy(e);
  dfree(c->d_tw); dfree(c->d_ops); dfree(c->d_vel); dfree(c->d_nl); dfree(c->d_R); dfree(c->d_P);
  dfree(c->d_cbc); dfree(c->d_dudy); dfree(c->d_sc); dfree(c->d_cfl); dfree(c->d_Z); dfree(c->d_AB);
  dfree(c->d_scr); dfree(c->d_base);
  les_destroy(c->les);
  dfree(c->d_jet);
  if (c->d_stage) (void)hipFree(c->d_stage);
  if (c->h_pin) (void)hipHostFree(c->h_pin);
  if (c->own_stream && c->stream) (void)hipStreamDestroy(c->stream);
  delete c;
  return 0;
}

int chan_upload_operators(chan_ctx* c, const chan_operators* o) {
  return guarded(c, [&] {
    if (!o || !o->d2h || !o->d2n || !o->dm || !o->dh || !o->dn || !o->hn || !o->iu || !o->dyh || !o->dy ||
        !o->kx || !o->kz)
      throw chan_error("missing operator array");
    const int ny = c->prm.ny, nyp = c->nyp, nxh = c->nxh, nz = c->prm.nz;
    const long nb = (long)c->prm.nx2 * c->prm.nz2;
    bool hasbuf = false;
    if (o->buffer)
      for (long i = 0; i < nb; ++i)
        if (o->buffer[i] != 0.0) { hasbuf = true; break; }
    std::vector<double> h;
    std::vector<size_t> at;
    auto put = [&](const double* src, long n) {
      at.push_back(h.size());
      if (src) h.insert(h.end(), src, src + n);
      else h.insert(h.end(), n, 0.0);
      while (h.size() % 2) h.push_back(0.0);
    };
    put(o->d2h, 3L * nyp); put(o->d2n, 3L * ny); put(o->dm, 12); put(o->dh, nyp); put(o->dn, ny);
    put(o->hn, ny); put(o->iu, 2L * ny); put(o->dyh, nyp); put(o->dy, nyp); put(o->kx, nxh); put(o->kz, nz);
    put(o->target_inflow, 3L * nyp);
    if (hasbuf) put(o->buffer, nb);
    dfree(c->d_ops);
    c->d_ops = (double*)dalloc(sizeof(double) * h.size());
    h2d(c, c->d_ops, h.data(), sizeof(double) * h.size());
    double* b = c->d_ops;
    c->ops.d2h = b + at[0]; c->ops.d2n = b + at[1]; c->ops.dm = b + at[2]; c->ops.dh = b + at[3];
    c->ops.dn = b + at[4]; c->ops.hn = b + at[5]; c->ops.iu = b + at[6]; c->ops.dyh = b + at[7];
    c->ops.dy = b + at[8]; c->ops.kx = b + at[9]; c->ops.kz = b + at[10]; c->ops.target = b + at[11];
    c->ops.buffer = hasbuf ? b + at[12] : nullptr;
    c->has_buffer = hasbuf;
    std::memcpy(c->h_dm, o->dm, sizeof(double) * 12);
  });
}

// define_coordinates (PFC/setup.f90:541-754) and the wavenumbers of
// define_parameters (PFC/setup.f90:1439-1466) for the channel grid.
int chan_build_operators(chan_ctx* c, double stretch, double* out_y, double* out_yh) {
  return guarded(c, [&] {
    const int ny = c->prm.ny, nyp = ny + 1, nx = c->prm.nx, nz = c->prm.nz;
    const double half = 0.5, one = 1.0, two = 2.0;
    std::vector<double> y(ny), yh(nyp), hh(nyp), dh(nyp), hn(ny), dn(ny), iu(2 * ny), dm(12, 0.0),
        d2n(3 * ny, 0.0), d2h(3 * nyp, 0.0), dyh(nyp, 0.0), dy(nyp, 0.0), kx(nx / 2 + 1), kz(nz, 0.0);
    for (int p = 1; p <= ny; ++p) {
      double xi = two * (double)(p - 1) / (double)(ny - 1) - one;
      y[p - 1] = std::tanh(stretch * xi) / std::tanh(stretch);
    }
    yh[0] = y[0];
    for (int j = 1; j < ny; ++j) yh[j] = half * (y[j - 1] + y[j]);
    yh[ny] = y[ny - 1];
    hh[0] = y[1] - y[0];
    for (int j = 1; j < ny; ++j) hh[j] = y[j] - y[j - 1];
    hh[ny] = y[ny - 1] - y[ny - 2];
    for (int j = 0; j < nyp; ++j) dh[j] = one / hh[j];
    for (int j = 0; j < ny; ++j) { hn[j] = yh[j + 1] - yh[j]; dn[j] = one / hn[j]; }
    for (int j = 0; j < ny; ++j) { iu[2 * j] = half * hh[j] / hn[j]; iu[2 * j + 1] = half * hh[j + 1] / hn[j]; }
    auto DM = [&](int i, int k) -> double& { return dm[(i - 1) + 3 * (k - 1)]; };
    DM(1, 1) = -dn[0] * (two * hn[0] + hn[1]) / (hn[0] + hn[1]);
    DM(2, 1) = dn[0] * dn[1] * (hn[0] + hn[1]);
    DM(3, 1) = -dn[1] * hn[0] / (hn[0] + hn[1]);
    DM(1, 2) = dn[ny - 2] * hn[ny - 1] / (hn[ny - 2] + hn[ny - 1]);
    DM(2, 2) = -dn[ny - 1] * dn[ny - 2] * (hn[ny - 1] + hn[ny - 2]);
    DM(3, 2) = dn[ny - 1] * (two * hn[ny - 1] + hn[ny - 2]) / (hn[ny - 1] + hn[ny - 2]);
    DM(1, 3) = (y[0] - y[1] + y[0] - y[2]) / ((y[0] - y[1]) * (y[0] - y[2]));
    DM(2, 3) = (y[0] - y[2]) / ((y[1] - y[0]) * (y[1] - y[2]));
    DM(3, 3) = (y[0] - y[1]) / ((y[2] - y[0]) * (y[2] - y[1]));
    DM(1, 4) = (y[ny - 1] - y[ny - 2]) / ((y[ny - 3] - y[ny - 1]) * (y[ny - 3] - y[ny - 2]));
    DM(2, 4) = (y[ny - 1] - y[ny - 3]) / ((y[ny - 2] - y[ny - 1]) * (y[ny - 2] - y[ny - 3]));
    DM(3, 4) = (y[ny - 1] - y[ny - 2] + y[ny - 1] - y[ny - 3]) / ((y[ny - 1] - y[ny - 2]) * (y[ny - 1] - y[ny - 3]));
    for (int p = 2; p <= ny - 1; ++p) {
      int a = p - 1;
      d2n[3 * a + 0] = dn[a] * dh[a];
      d2n[3 * a + 1] = -dn[a] * dh[a] - dn[a] * dh[a + 1];
      d2n[3 * a + 2] = dn[a] * dh[a + 1];
      d2h[3 * a + 0] = dh[a] * dn[a - 1];
      d2h[3 * a + 1] = -dh[a] * dn[a - 1] - dh[a] * dn[a];
      d2h[3 * a + 2] = dh[a] * dn[a];
    }
    d2h[3 * 1 + 0] = -dh[1] * DM(1, 1);
    d2h[3 * 1 + 1] = -dh[1] * dn[1] - dh[1] * DM(2, 1);
    d2h[3 * 1 + 2] = dh[1] * dn[1] - dh[1] * DM(3, 1);
    d2h[3 * (ny - 1) + 0] = dh[ny - 1] * dn[ny - 2] + dh[ny - 1] * DM(1, 2);
    d2h[3 * (ny - 1) + 1] = -dh[ny - 1] * dn[ny - 2] + dh[ny - 1] * DM(2, 2);
    d2h[3 * (ny - 1) + 2] = dh[ny - 1] * DM(3, 2);
    for (int i = 0; i < 3; ++i) {
      d2h[i] = d2h[3 + i];
      d2n[i] = d2n[3 + i];
      d2n[3 * (ny - 1) + i] = d2n[3 * (ny - 2) + i];
      d2h[3 * ny + i] = d2h[3 * (ny - 1) + i];
    }
    for (int j = 1; j < ny; ++j) dyh[j] = hh[j];
    dy[0] = half * hh[1];
    for (int j = 1; j < ny - 1; ++j) dy[j] = half * (hh[j] + hh[j + 1]);
    dy[ny - 1] = half * hh[ny - 1];
    const double pi = std::atan2(0.0, -1.0);
    for (int x = 0; x <= nx / 2; ++x) kx[x] = (double)x * 2.0 * pi / c->prm.lx;
    for (int z = 0; z <= nz / 2; ++z) kz[z] = (double)z * 2.0 * pi / c->prm.lz;
    for (int z = nz; z >= nz / 2 + 2; --z) kz[z - 1] = -kz[nz + 2 - z - 1];
    chan_operators o{};
    o.d2h = d2h.data(); o.d2n = d2n.data(); o.dm = dm.data(); o.dh = dh.data(); o.dn = dn.data();
    o.hn = hn.data(); o.iu = iu.data(); o.dyh = dyh.data(); o.dy = dy.data(); o.kx = kx.data(); o.kz = kz.data();
    o.buffer = nullptr; o.target_inflow = nullptr;
    if (chan_upload_operators(c, &o)) throw chan_error(c->last_error);
    c->h_y = y;
    if (out_y) std::memcpy(out_y, y.data(), sizeof(double) * ny);
    if (out_yh) std::memcpy(out_yh, yh.data(), sizeof(double) * nyp);
  });
}

int chan_set_state(chan_ctx* c, const double* data1, const double* data2, const double* data4) {
  return guarded(c, [&] {
    const int ny = c->prm.ny;
    if (data1) { h2d(c, c->d_vel, data1, vel_bytes(c)); c->data1_is_nl = false; }
    if (data2) {
      void* s = stage(c, vel_bytes(c));
      h2d(c, s, data2, vel_bytes(c));
      launch_layout(0, ny, c->nxh, c->prm.nz, (const double2*)s, c->d_R, c->stream);
    }
    if (data4) {
      void* s = stage(c, vel_bytes(c));
      h2d(c, s, data4, p_bytes(c));
      launch_layout(2, ny, c->nxh, c->prm.nz, (const double2*)s, c->d_P, c->stream);
    }
    CHAN_CHECK(hipStreamSynchronize(c->stream));
  });
}

int chan_get_state(chan_ctx* c, double* data1, double* data2, double* data4) {
  return guarded(c, [&] {
    const int ny = c->prm.ny;
    if (data1) d2h(c, data1, c->data1_is_nl ? c->d_nl : c->d_vel, vel_bytes(c));
    if (data2) {
      void* s = stage(c, vel_bytes(c));
      launch_layout(1, ny, c->nxh, c->prm.nz, c->d_R, (double2*)s, c->stream);
      d2h(c, data2, s, vel_bytes(c));
    }
    if (data4) {
      void* s = stage(c, vel_bytes(c));
      launch_layout(3, ny, c->nxh, c->prm.nz, c->d_P, (double2*)s, c->stream);
      d2h(c, data4, s, p_bytes(c));
    }
  });
}

int chan_get_planes(chan_ctx* c, int comp, int k0, int k1, double* out) {
  return guarded(c, [&] {
    if (comp < 1 || comp > 3 || k0 < 1 || k1 > c->nyp || k1 < k0) throw chan_error("bad plane range");
    const double2* src = (c->data1_is_nl ? c->d_nl : c->d_vel) + ((long)(comp - 1) * c->nyp + (k0 - 1)) * nmodes(c);
    d2h(c, out, src, sizeof(double2) * nmodes(c) * (k1 - k0 + 1));
  });
}

int chan_set_scalars(chan_ctx* c, const double* dt, const double* t_total, const double* p_grad,
                     const double* d_um_dy) {
  return guarded(c, [&] {
    DevScalars sc;
    d2h(c, &sc, c->d_sc, sizeof(sc));
    if (dt) sc.dt = *dt;
    if (t_total) sc.t_total = *t_total;
    if (p_grad) sc.p_grad = *p_grad;
    h2d(c, c->d_sc, &sc, sizeof(sc));
    c->h_dt = sc.dt;
    c->h_dt_known = true;
    if (d_um_dy) h2d(c, c->d_dudy, d_um_dy, sizeof(double) * c->prm.ny);
    CHAN_CHECK(hipStreamSynchronize(c->stream));
  });
}

int chan_get_scalars(chan_ctx* c, double* dt, double* t_total, double* p_grad, double* d_um_dy) {
  return guarded(c, [&] {
    DevScalars sc;
    d2h(c, &sc, c->d_sc, sizeof(sc));
    if (dt) *dt = sc.dt;
    if (t_total) *t_total = sc.t_total;
    if (p_grad) *p_grad = sc.p_grad;
    if (d_um_dy) d2h(c, d_um_dy, c->d_dudy, sizeof(double) * c->prm.ny);
  });
}

int chan_set_wall_bc(chan_ctx* c, const double* top, const double* bot) {
  return guarded(c, [&] {
    const size_t b = sizeof(double2) * 3 * nmodes(c);
    if (top) h2d(c, c->d_cbc, top, b);
    if (bot) h2d(c, c->d_cbc + 3 * nmodes(c), bot, b);
    CHAN_CHECK(hipStreamSynchronize(c->stream));
  });
}

// The caller's dt goes to the device only when it differs from the value the
// host last knew there.  Only rk_step 1 can change dt (the CFL check and the
// adaptation, PFC/pass1.f90:106-124), so only then does the call wait for
// the kernels and read dt and the CFL number back; at rk_step 2 and 3 it
// returns as soon as the pass is queued, with dt unchanged.
static int nonlinear_common(chan_ctx* c, int step, int rk, double* dt_inout, double* cfl_out, bool lin) {
  return guarded(c, [&] {
    check_ops(c);
    if (dt_inout && !(c->h_dt_known && c->h_dt == *dt_inout)) {
      h2d(c, &c->d_sc->dt, dt_inout, sizeof(double));
      c->h_dt = *dt_inout;
      c->h_dt_known = true;
    }
    ensure_vel_in_vel(c);
    nl_pass(c, step, rk, lin, rk == 1);
    c->data1_is_nl = true;
    if (rk == 1) {
      CHAN_CHECK(hipStreamSynchronize(c->stream));
      check_err(c);
      DevScalars sc;
      d2h(c, &sc, c->d_sc, sizeof(sc));
      c->h_dt = sc.dt;
      c->h_dt_known = true;
      if (dt_inout) *dt_inout = sc.dt;
      if (cfl_out) *cfl_out = sc.cfl;
    } else {
      if (dt_inout && !c->h_dt_known) {
        DevScalars sc;
        d2h(c, &sc, c->d_sc, sizeof(sc));
        c->h_dt = sc.dt;
        c->h_dt_known = true;
      }
      if (dt_inout) *dt_inout = c->h_dt;
      if (cfl_out) *cfl_out = 0.0;
    }
  });
}

int chan_nonlinear(chan_ctx* c, int step, int rk, double* dt_inout, double* cfl_out) {
  return nonlinear_common(c, step, rk, dt_inout, cfl_out, false);
}

int chan_nonlinear_linearized(chan_ctx* c, int step, int rk, double* dt_inout, double* cfl_out) {
  return nonlinear_common(c, step, rk, dt_inout, cfl_out, true);
}

int chan_set_baseflow(chan_ctx* c, const double* base) {
  return guarded(c, [&] {
    const size_t b = sizeof(double) * 3 * c->nyp * (size_t)c->prm.nx2 * c->prm.nz2;
    if (!c->d_base) c->d_base = (double*)dalloc(b);
    h2d(c, c->d_base, base, b);
    CHAN_CHECK(hipStreamSynchronize(c->stream));
  });
}

int chan_jet_configure(chan_ctx* c, int wall, double amplitude, double radius, double xjet, double zjet,
                       double forcing, double frequency) {
  return guarded(c, [&] {
    check_ops(c);
    if (wall != 0 && wall != 1) throw chan_error("jet wall must be 0 (bottom) or 1 (top)");
    if (amplitude == 0.0) {
      dfree(c->d_jet);
      c->d_jet = nullptr;
      return;
    }
    if (!(radius > 0.0)) throw chan_error("jet radius must be > 0");
    const XZGeom& g = c->gs;  // the nx x nz grid of the wall profiles
    const int nx = g.nx2, nz = g.nz2;
    const double lx = c->prm.lx, lz = c->prm.lz, sign = (wall == 0) ? 1.0 : -1.0;
    std::vector<double> f((size_t)nx * nz);
    for (int z = 0; z < nz; ++z) {
      double dz = (double)z * lz / (double)nz - zjet;
      dz -= lz * std::floor(dz / lz + 0.5);
      for (int x = 0; x < nx; ++x) {
        double dx = (double)x * lx / (double)nx - xjet;
        dx -= lx * std::floor(dx / lx + 0.5);
        f[(size_t)z * nx + x] = sign * amplitude * std::exp(-(dx * dx + dz * dz) / (radius * radius));
      }
    }
    const size_t cb = sizeof(double2) * (size_t)g.nxh * g.nz, rb = sizeof(double) * f.size();
    if (!c->d_jet) c->d_jet = (double2*)dalloc(cb);
    char* st = (char*)stage(c, rb);
    h2d(c, st, f.data(), rb);
    dev_r2c(c, g, (const double*)st, 1, c->d_jet);
    // no net mass flux through the wall: the mean mode is removed (uniform suction)
    CHAN_CHECK(hipMemsetAsync(c->d_jet, 0, sizeof(double2), c->stream));
    CHAN_CHECK(hipStreamSynchronize(c->stream));
    c->jet_wall = wall;
    c->jet_forcing = forcing;
    c->jet_freq = frequency;
  });
}

int chan_rhs_solve(chan_ctx* c, int step, int rk, double* p_grad_out) {
  (void)step;
  return guarded(c, [&] {
    check_ops(c);
    if (rk < 1 || rk > 3) throw chan_error("rk_step must be 1..3");
    ensure_nl_in_nl(c);
    solve_pass(c, rk);
    c->data1_is_nl = false;
    if (p_grad_out) {
      DevScalars sc;
      d2h(c, &sc, c->d_sc, sizeof(sc));
      *p_grad_out = sc.p_grad;
    }
  });
}

int chan_step(chan_ctx* c, int step0, int nsteps, int lin) {
  return guarded(c, [&] {
    check_ops(c);
    ensure_vel_in_vel(c);
    c->h_dt_known = false;
    for (int s = 0; s < nsteps; ++s) step_body(c, step0 + s, lin != 0);
    c->data1_is_nl = false;
    CHAN_CHECK(hipStreamSynchronize(c->stream));
    check_err(c);
  });
}

int chan_boundary(chan_ctx* c) {
  return guarded(c, [&] {
    check_ops(c);
    ensure_vel_in_vel(c);
    boundary_dev(c);
    CHAN_CHECK(hipStreamSynchronize(c->stream));
  });
}

int chan_initialize_flow(chan_ctx* c, int lin) {
  return guarded(c, [&] {
    init_flow(c, lin != 0);
    CHAN_CHECK(hipStreamSynchronize(c->stream));
  });
}

int chan_remove_divergence(chan_ctx* c) {
  return guarded(c, [&] {
    check_ops(c);
    ensure_vel_in_vel(c);
    launch_project(c->ops, c->prm.ny, c->nxh, c->prm.nz, c->d_vel, c->d_vel, nullptr, c->d_scr, 1.0, 0,
                   nullptr, c->prm.re, c->stream);
    launch_dudy(c->ops, c->prm.ny, c->nxh, c->prm.nz, c->d_vel, c->d_dudy, c->stream);
    CHAN_CHECK(hipStreamSynchronize(c->stream));
  });
}

int chan_compute_initial_pressure(chan_ctx* c, int lin) {
  return guarded(c, [&] {
    check_ops(c);
    ensure_vel_in_vel(c);
    d2d(c, c->d_R, c->d_vel, vel_bytes(c));
    nl_pass(c, -1, 3, lin != 0, false);
    launch_project(c->ops, c->prm.ny, c->nxh, c->prm.nz, c->d_nl, nullptr, c->d_P, c->d_scr, 1.0, 1,
                   c->d_vel, c->prm.re, c->stream);
    CHAN_CHECK(hipStreamSynchronize(c->stream));
  });
}

static void check_np(chan_ctx* c, int np) {
  if (np < 1 || np > 3 * c->nyp) throw chan_error("nplanes must be in 1..3*(ny+1)");
}

int chan_xz_fft_complex_to_real(chan_ctx* c, const double* cdata, double* data, int np) {
  return guarded(c, [&] {
    check_ops(c); check_np(c, np);
    const XZGeom& g = c->gd;
    const size_t cb = sizeof(double2) * (size_t)np * g.nxh * g.nz, rb = sizeof(double) * (size_t)np * g.nx2 * g.nz2;
    char* s = (char*)stage(c, cb + rb);
    h2d(c, s, cdata, cb);
    dev_c2r(c, g, (const double2*)s, np, (double*)(s + cb));
    d2h(c, data, s + cb, rb);
  });
}

int chan_xz_fft_real_to_complex(chan_ctx* c, const double* data, double* cdata, int np) {
  return guarded(c, [&] {
    check_ops(c); check_np(c, np);
    const XZGeom& g = c->gd;
    const size_t cb = sizeof(double2) * (size_t)np * g.nxh * g.nz, rb = sizeof(double) * (size_t)np * g.nx2 * g.nz2;
    char* s = (char*)stage(c, cb + rb);
    h2d(c, s + cb, data, rb);
    dev_r2c(c, g, (const double*)(s + cb), np, (double2*)s);
    d2h(c, cdata, s, cb);
  });
}

int chan_compute_divergence_of_fluxes(chan_ctx* c, const double* fx, const double* fy, const double* fz,
                                      double* cdata, int np) {
  return guarded(c, [&] {
    check_ops(c);
    if (np < 1 || 2 * np > 3 * c->nyp) throw chan_error("nplanes too large for the staging buffers");
    const XZGeom& g = c->gd;
    const size_t cb = sizeof(double2) * (size_t)np * g.nxh * g.nz, rb = sizeof(double) * (size_t)np * g.nx2 * g.nz2;
    char* s = (char*)stage(c, cb + 3 * rb);
    h2d(c, s + cb, fx, rb);
    h2d(c, s + cb + rb, fy, rb);
    h2d(c, s + cb + 2 * rb, fz, rb);
    dev_div(c, (const double*)(s + cb), (const double*)(s + cb + rb), (const double*)(s + cb + 2 * rb), np,
            (double2*)s);
    d2h(c, cdata, s, cb);
  });
}

// ---------------------------------------------------------------------------
// Allocation-free form of compute_divergence_of_fluxes for callers that cannot
// allocate device memory.  The x-transform intermediates (A = F(fy) + i kx
// F(fx), B = F(fz), [2][nplanes][nz2][kxpad] complex) live in `scratch`, a
// fixed array in the code object's device memory: each GPU holds its own copy
// for the life of the process, and every call on that GPU uses it, so calls
// that overlap on one GPU race.  Every element the z-forward reads is written
// by the x-transforms of the same call, so nothing needs clearing between
// calls.  Host code never passes `scratch` itself to a kernel (its name is a
// host-side stand-in); it uses the address hipGetSymbolAddress returns for
// the current device, looked up once per device.
// ---------------------------------------------------------------------------
static __device__ __align__(256) double2 scratch[CHAN_NOALLOC_SCRATCH_BYTES / sizeof(double2)];

namespace {
constexpr int kMaxDevices = 64;
std::mutex g_scratch_mu;
double2* g_scratch_ptr[kMaxDevices] = {};

double2* scratch_for_device(int dev) {
  if (dev < 0 || dev >= kMaxDevices) throw chan_error("device index out of range for the scratch table");
  std::lock_guard<std::mutex> lk(g_scratch_mu);
  if (!g_scratch_ptr[dev]) {
    void* p = nullptr;
    CHAN_CHECK(hipGetSymbolAddress(&p, HIP_SYMBOL(scratch)));
    g_scratch_ptr[dev] = static_cast<double2*>(p);
  }
  return g_scratch_ptr[dev];
}
}  // namespace

int chan_compute_divergence_of_fluxes_noalloc(chan_ctx* c, const double* fx, const double* fy, const double* fz,
                                              double* cdata, int np, void* stream) {
  return guarded(c, [&] {
    check_ops(c);
    if (np < 1) throw chan_error("nplanes must be >= 1");
    if (!fx || !fy || !fz || !cdata) throw chan_error("device arrays datax, datay, dataz and cdata are required");
    const XZGeom& g = c->gd;
    const long aps = (long)g.nz2 * g.kxpad;
    const size_t need = sizeof(double2) * 2 * (size_t)np * (size_t)aps;
    if (need > CHAN_NOALLOC_SCRATCH_BYTES)
      throw chan_error("compute_divergence_of_fluxes_noalloc: " + std::to_string(np) + " planes need " +
                       std::to_string(need) + " bytes of scratch, more than the " +
                       std::to_string((size_t)CHAN_NOALLOC_SCRATCH_BYTES) + " available (at most " +
                       std::to_string(CHAN_NOALLOC_SCRATCH_BYTES / (sizeof(double2) * 2 * aps)) +
                       " planes per call at this grid)");
    int cur = 0;
    CHAN_CHECK(hipGetDevice(&cur));
    if (cur != c->device) CHAN_CHECK(hipSetDevice(c->device));
    struct Restore {
      int d, was;
      ~Restore() { if (d != was) (void)hipSetDevice(was); }
    } restore{c->device, cur};
    double2* A = scratch_for_device(c->device);
    double2* B = A + (long)np * aps;
    hipStream_t s = stream ? static_cast<hipStream_t>(stream) : c->stream;
    const long ips = (long)g.nz2 * g.nx2, ops_ = (long)g.nxh * g.nz;
    launch_xr2c(g, c->d_tw, fy, fx, c->ops.kx, ips, A, aps, np, s);
    launch_xr2c(g, c->d_tw, fz, nullptr, c->ops.kx, ips, B, aps, np, s);
    launch_zfwd(g, c->d_tw, A, B, aps, c->ops.kz, (double2*)cdata, ops_, np, 0, s);
  });
}

int chan_les_filter_dns(chan_ctx* les, chan_ctx* dns, double* tau) {
  return guarded(les, [&] {
    if (!dns) throw chan_error("filter: null fine channel");
    check_ops(les);
    check_ops(dns);
    ensure_vel_in_vel(dns);
    filter_dns_device(les, dns, tau, dev_c2r, dev_r2c);
  });
}

int chan_small_fft_complex_to_real(chan_ctx* c, const double* cdata, double* data, int np) {
  return guarded(c, [&] {
    check_ops(c); check_np(c, np);
    const XZGeom& g = c->gs;
    const size_t cb = sizeof(double2) * (size_t)np * g.nxh * g.nz, rb = sizeof(double) * (size_t)np * g.nx2 * g.nz2;
    char* s = (char*)stage(c, cb + rb);
    h2d(c, s, cdata, cb);
    dev_c2r(c, g, (const double2*)s, np, (double*)(s + cb));
    d2h(c, data, s + cb, rb);
  });
}

int chan_small_fft_real_to_complex(chan_ctx* c, const double* data, double* cdata, int np) {
  return guarded(c, [&] {
    check_ops(c); check_np(c, np);
    const XZGeom& g = c->gs;
    const size_t cb = sizeof(double2) * (size_t)np * g.nxh * g.nz, rb = sizeof(double) * (size_t)np * g.nx2 * g.nz2;
    char* s = (char*)stage(c, cb + rb);
    h2d(c, s + cb, data, rb);
    dev_r2c(c, g, (const double*)(s + cb), np, (double2*)s);
    d2h(c, cdata, s, cb);
  });
}

int chan_solve_helmholtz(chan_ctx* c, int n, int nsys, double* f, const double* alpha, double diff,
                         const double* d2, const double* topbc, const double* botbc, int top_neumann,
                         int bot_neumann) {
  return guarded(c, [&] {
    if (top_neumann || bot_neumann)
      throw chan_error("Neumann wall rows (bl_grid != 0) are not supported");
    if (n < 3 || nsys < 1) throw chan_error("bad system size");
    const size_t fb = sizeof(double2) * (size_t)n * nsys, ab = sizeof(double) * nsys, db = sizeof(double) * 3 * n,
                 bb = sizeof(double2) * nsys, sb = sizeof(double) * (size_t)n * nsys;
    char* p = (char*)stage(c, fb + ab + db + 2 * bb + sb + 6 * 256);
    double2* df = (double2*)carve(p, fb);
    double* da = (double*)carve(p, ab);
    double* dd = (double*)carve(p, db);
    double2* dt = (double2*)carve(p, bb);
    double2* db_ = (double2*)carve(p, bb);
    double* scr = (double*)carve(p, sb);
    h2d(c, df, f, fb); h2d(c, da, alpha, ab); h2d(c, dd, d2, db); h2d(c, dt, topbc, bb); h2d(c, db_, botbc, bb);
    launch_helmholtz_batch(n, nsys, df, da, diff, dd, dt, db_, scr, c->stream);
    d2h(c, f, df, fb);
  });
}

int chan_solve_p(chan_ctx* c, int nsys, double* pp, const double* alpha) {
  return guarded(c, [&] {
    check_ops(c);
    const int n = c->nyp;
    const size_t fb = sizeof(double2) * (size_t)n * nsys, ab = sizeof(double) * nsys, sb = sizeof(double) * (size_t)n * nsys;
    char* cur = (char*)stage(c, fb + ab + sb + 3 * 256);
    char* s = carve(cur, fb);
    double* da = (double*)carve(cur, ab);
    double* scr = (double*)carve(cur, sb);
    h2d(c, s, pp, fb);
    h2d(c, da, alpha, ab);
    launch_solvep_batch(c->prm.ny, nsys, (double2*)s, da, c->ops, scr, c->stream);
    d2h(c, pp, s, fb);
  });
}

int chan_strip_divergence(chan_ctx* c, double* cu, double* cp, double coef) {
  return guarded(c, [&] {
    check_ops(c);
    const size_t vb = vel_bytes(c), pb = p_bytes(c);
    char* s = (char*)stage(c, vb + 2 * pb);
    h2d(c, s, cu, vb);
    launch_project(c->ops, c->prm.ny, c->nxh, c->prm.nz, (const double2*)s, (double2*)s, (double2*)(s + vb),
                   c->d_scr, coef, 2, nullptr, c->prm.re, c->stream);
    d2h(c, cu, s, vb);
    if (cp) {
      launch_layout(3, c->prm.ny, c->nxh, c->prm.nz, (const double2*)(s + vb), (double2*)(s + vb + pb), c->stream);
      d2h(c, cp, s + vb + pb, pb);
    }
  });
}

// 1-based reference mode indices -> 0-based pairs on the device (after cu in the staging buffer)
static int* stage_cols(chan_ctx* c, char* at, int nsys, const int* p, const int* q, bool reject_mean) {
  if (!p || !q) throw chan_error("mode indices p, q are required");
  std::vector<int> cols(2 * (size_t)nsys);
  for (int s = 0; s < nsys; ++s) {
    const int ix = p[s] - 1, iz = q[s] - 1;
    if (ix < 0 || ix >= c->nxh || iz < 0 || iz >= c->prm.nz)
      throw chan_error("mode index out of range (p = " + std::to_string(p[s]) + ", q = " + std::to_string(q[s]) + ")");
    if (reject_mean && ix == 0 && iz == 0) throw chan_error("strip_divergence: the mean mode has no projection");
    cols[2 * s] = ix;
    cols[2 * s + 1] = iz;
  }
  h2d(c, at, cols.data(), sizeof(int) * cols.size());
  return (int*)at;
}

int chan_strip_divergence_modes(chan_ctx* c, int nsys, double* cu, double* cp, double coef, const int* p,
                                const int* q) {
  return guarded(c, [&] {
    check_ops(c);
    if (nsys < 0) throw chan_error("nsys must be >= 0");
    if (nsys == 0) return;
    const size_t nyp = (size_t)c->nyp;
    const size_t vb = sizeof(double2) * 3 * nyp * nsys, pb = sizeof(double2) * nyp * nsys;
    const size_t sb = sizeof(double2) * nyp * nsys + sizeof(double) * nyp * nsys, ib = sizeof(int) * 2 * nsys;
    char* cur = (char*)stage(c, vb + pb + sb + ib + 4 * 256);
    char* dv = carve(cur, vb);
    char* dp = carve(cur, pb);
    char* ds = carve(cur, sb);
    char* di = carve(cur, ib);
    const int* cols = stage_cols(c, di, nsys, p, q, true);
    h2d(c, dv, cu, vb);
    launch_project(c->ops, c->prm.ny, c->nxh, c->prm.nz, (const double2*)dv, (double2*)dv, (double2*)dp,
                   (double2*)ds, coef, 2, nullptr, c->prm.re, c->stream, cols, nsys);
    d2h(c, cu, dv, vb);
    if (cp) d2h(c, cp, dp, pb);
  });
}

int chan_subtract_pressure_gradient_modes(chan_ctx* c, int nsys, double* crhs, const double* cpress,
                                          double coef, const int* p, const int* q) {
  return guarded(c, [&] {
    check_ops(c);
    if (nsys < 0) throw chan_error("nsys must be >= 0");
    if (nsys == 0) return;
    if (!cpress) throw chan_error("subtract_pressure_gradient: cpress is required");
    const size_t nyp = (size_t)c->nyp;
    const size_t vb = sizeof(double2) * 3 * nyp * nsys, pb = sizeof(double2) * nyp * nsys;
    const size_t ib = sizeof(int) * 2 * nsys;
    char* cur = (char*)stage(c, vb + pb + ib + 3 * 256);
    char* dv = carve(cur, vb);
    char* dp = carve(cur, pb);
    char* di = carve(cur, ib);
    const int* cols = stage_cols(c, di, nsys, p, q, false);
    h2d(c, dv, crhs, vb);
    h2d(c, dp, cpress, pb);
    launch_subtract_grad_cols(c->ops, c->prm.ny, (double2*)dv, (const double2*)dp, coef, cols, nsys, c->stream);
    d2h(c, crhs, dv, vb);
  });
}

int chan_linearized_matvec(chan_ctx* c, const double* v, double* w, int tend, double* t_out) {
  return guarded(c, [&] {
    check_ops(c);
    if (!c->d_base) throw chan_error("base flow not set (chan_set_baseflow)");
    if (c->prm.fix_mass != 0) throw chan_error("the linearized matvec needs fix_mass = 0");
    if (tend < 1) throw chan_error("tend must be >= 1");
    c->h_dt_known = false;
    const XZGeom& g = c->gs;
    const int nyp = c->nyp, ny = c->prm.ny;
    const long plane_r = (long)g.nx2 * g.nz2;          // nx * nz reals
    const long plane_c = (long)g.nxh * g.nz;
    const size_t rb = sizeof(double) * 3 * nyp * plane_r;
    char* s = (char*)stage(c, rb + 16);
    double* dv = (double*)s;
    unsigned long long* vmax = (unsigned long long*)(s + rb);
    h2d(c, dv, v, rb);
    launch_fill_zero(vmax, 8, c->stream);
    launch_absmax(3L * nyp * plane_r, dv, vmax, c->stream);
    dev_r2c(c, g, dv, 3 * nyp, c->d_vel);
    launch_arnoldi_scale(ny, plane_c, c->d_vel, c->ops.dy, c->ops.dyh, vmax, 1.0e-8, 0, c->stream);
    c->data1_is_nl = false;
    d2d(c, c->d_R, c->d_vel, vel_bytes(c));
    launch_fill_zero(c->d_cbc, sizeof(double2) * 6 * plane_c, c->stream);
    DevScalars sc;
    d2h(c, &sc, c->d_sc, sizeof(sc));
    const double dt = sc.dt;
    sc.p_grad = 0.0;   // init_baseflow (PFC/pass1_linearized.f90:126)
    sc.t_total = 0.0;
    h2d(c, c->d_sc, &sc, sizeof(sc));
    // initial pressure (PFC/arnoldi.f90:528-535); the stale rk_step is never 1
    nl_pass(c, -1, 0, true, false);
    launch_project(c->ops, ny, c->nxh, c->prm.nz, c->d_nl, nullptr, c->d_P, c->d_scr, 1.0, 1, c->d_vel, c->prm.re,
                   c->stream);
    double t_total = 0.0;
    for (int st = 1; st <= tend; ++st) {
      for (int rk = 1; rk <= 3; ++rk) {
        nl_pass(c, st, rk, true, rk == 1);
        solve_pass(c, rk);
      }
      t_total = t_total + dt;
    }
    CHAN_CHECK(hipStreamSynchronize(c->stream));
    check_err(c);
    launch_arnoldi_scale(ny, plane_c, c->d_vel, c->ops.dy, c->ops.dyh, vmax, 1.0e-8, 1, c->stream);
    dev_c2r(c, g, c->d_vel, 3 * nyp, dv);
    d2h(c, w, dv, rb);
    if (t_out) *t_out = t_total;
  });
}

int chan_divergence_max(chan_ctx* c, double* out) {
  return guarded(c, [&] {
    check_ops(c);
    unsigned long long* d = (unsigned long long*)stage(c, 8);
    launch_fill_zero(d, 8, c->stream);
    launch_divmax(c->ops, c->prm.ny, c->nxh, c->prm.nz, c->data1_is_nl ? c->d_nl : c->d_vel, d, c->stream);
    unsigned long long h;
    d2h(c, &h, d, 8);
    std::memcpy(out, &h, 8);
  });
}

int chan_set_coordinates(chan_ctx* c, const double* y) {
  return guarded(c, [&] {
    if (!y) throw chan_error("null argument");
    c->h_y.assign(y, y + c->prm.ny);
  });
}

int chan_l2_norm(chan_ctx* c, double* l2, double* drag) {
  return guarded(c, [&] {
    check_ops(c);
    if ((int)c->h_y.size() != c->prm.ny) throw chan_error("coordinates not set (chan_set_coordinates)");
    ensure_vel_in_vel(c);
    const XZGeom& g = c->gd;
    const int ny = c->prm.ny, nyp = c->nyp;
    const long ps = (long)g.nx2 * g.nz2;
    if ((size_t)nyp * ps * sizeof(double) > (size_t)2 * 3 * nyp * g.nz2 * g.kxpad * sizeof(double2))
      throw chan_error("l2_norm: physical planes do not fit the transform buffer");
    // laminar profile at the v-point coordinate y(plane); plane ny+1 uses y(ny)
    // (the reference reads y(ny+1), one past the end of y)
    std::vector<double> lam(nyp);
    for (int j = 0; j < nyp; ++j) {
      const double yy = c->h_y[std::min(j, ny - 1)];
      lam[j] = 1.0 - yy * yy;
    }
    char* s = (char*)stage(c, sizeof(double) * (nyp + 3 * nyp + 2) + 64);
    double* d_lam = (double*)s;
    double* part = d_lam + nyp;
    double* res = part + 3 * nyp;
    h2d(c, d_lam, lam.data(), sizeof(double) * nyp);
    double* phys = (double*)c->d_AB;
    for (int comp = 0; comp < 3; ++comp) {
      dev_c2r(c, g, c->d_vel + (long)comp * nyp * nmodes(c), nyp, phys);
      launch_l2_planes(phys, ps, nyp, comp == 0 ? d_lam : nullptr, c->ops.dy, part + (long)comp * nyp, c->stream);
    }
    launch_l2_final(part, 3 * nyp, (double)ps, c->d_dudy, ny, res, c->stream);
    double h[2];
    d2h(c, h, res, sizeof(h));
    if (l2) *l2 = h[0];
    if (drag) *drag = h[1];
  });
}

int chan_mode_energy(chan_ctx* c, double* ke) {
  return guarded(c, [&] {
    check_ops(c);
    ensure_vel_in_vel(c);
    const size_t b = sizeof(double) * nmodes(c);
    double* d = (double*)stage(c, b);
    launch_mode_energy(c->ops, c->prm.ny, c->nxh, c->prm.nz, c->d_vel, d, c->stream);
    d2h(c, ke, d, b);
  });
}

int chan_plane_moments(chan_ctx* c, double* out) {
  return guarded(c, [&] {
    check_ops(c);
    ensure_vel_in_vel(c);
    const XZGeom& g = c->gs;
    const int nyp = c->nyp;
    const long ps = (long)g.nx2 * g.nz2, cs = (long)nyp * ps;
    if ((size_t)3 * cs * sizeof(double) > (size_t)2 * 3 * nyp * c->gd.nz2 * c->gd.kxpad * sizeof(double2))
      throw chan_error("plane_moments: physical planes do not fit the transform buffer");
    double* phys = (double*)c->d_AB;
    dev_c2r(c, g, c->d_vel, 3 * nyp, phys);
    const size_t b = sizeof(double) * 15 * nyp;
    double* d = (double*)stage(c, b);
    launch_plane_moments(phys, ps, cs, nyp, d, c->stream);
    d2h(c, out, d, b);
  });
}

int chan_write_vtk(chan_ctx* c, const char* path, const double* yh) {
  return guarded(c, [&] {
    check_ops(c);
    if (!path || !yh) throw chan_error("null argument");
    ensure_vel_in_vel(c);
    const XZGeom& g = c->gs;
    const int ny = c->prm.ny, nyp = c->nyp;
    const long nm = nmodes(c), ps = (long)g.nx2 * g.nz2;
    const size_t pb = sizeof(double) * 3 * nyp * ps;
    if (pb > (size_t)2 * 3 * nyp * c->gd.nz2 * c->gd.kxpad * sizeof(double2))
      throw chan_error("write_vtk: physical planes do not fit the transform buffer");
    launch_vtk_field(ny, nm, c->d_vel, c->d_scr, c->stream);
    double* phys = (double*)c->d_AB;
    dev_c2r(c, g, c->d_scr, 3 * nyp, phys);
    std::vector<double> h(3 * nyp * ps);
    d2h(c, h.data(), phys, pb);
    if (chan_vtk_write(path, c->prm.nx, ny, c->prm.nz, c->prm.lx, c->prm.lz, yh, h.data()) != 0)
      throw chan_error(chan_io_last_error());
  });
}

int chan_mass_flux(chan_ctx* c, double* out) {
  return guarded(c, [&] {
    check_ops(c);
    const int nyp = c->nyp;
    std::vector<double2> col(nyp);
    std::vector<double> dyh(nyp);
    const double2* src = c->data1_is_nl ? c->d_nl : c->d_vel;
    double2* pc = (double2*)pinned(c, sizeof(double2) * nyp);
    CHAN_CHECK(hipMemcpy2DAsync(pc, sizeof(double2), src, sizeof(double2) * nmodes(c), sizeof(double2), nyp,
                                hipMemcpyDeviceToHost, c->stream));
    CHAN_CHECK(hipStreamSynchronize(c->stream));
    std::memcpy(col.data(), pc, sizeof(double2) * nyp);
    d2h(c, dyh.data(), c->ops.dyh, sizeof(double) * nyp);
    double q = 0.0;
    for (int j = 0; j < nyp; ++j) q += col[j].x * dyh[j];
    *out = q;
  });
}

int chan_set_timing(chan_ctx* c, int enable) {
  return guarded(c, [&] {
    collect_times(c);
    c->timing = enable;
    for (int k = 0; k < 8; ++k) { c->ktime_ms[k] = 0; c->kcount[k] = 0; }
  });
}

int chan_get_kernel_times(chan_ctx* c, double* ms, long* counts, int nk) {
  return guarded(c, [&] {
    collect_times(c);
    for (int k = 0; k < nk && k < 8; ++k) {
      if (ms) ms[k] = c->ktime_ms[k];
      if (counts) counts[k] = c->kcount[k];
    }
  });
}

int chan_sync(chan_ctx* c) {
  return guarded(c, [&] {
    CHAN_CHECK(hipStreamSynchronize(c->stream));
    check_err(c);
  });
}


// ---------------------------------------------------------------------------
// Launch-shape tuning and batched ensembles
// ---------------------------------------------------------------------------
int chan_set_tuning(chan_ctx* c, int key, int value) {
  return guarded(c, [&] {
    if (!tuning_set(c->tune, key, value)) throw chan_error("invalid tuning key or value");
  });
}

int chan_get_tuning(chan_ctx* c, int key, int* value) {
  return guarded(c, [&] {
    const int v = tuning_get(c->tune, key);
    if (v < 0 || !value) throw chan_error("invalid tuning key");
    *value = v;
  });
}

// Times `steps` timesteps for each candidate launch shape and keeps the
// fastest; the state (velocity, RK register, pressure, scalars, d_um_dy) is
// restored after every trial, so the call leaves the solution unchanged.
int chan_autotune(chan_ctx* c, int steps, double* best_ms) {
  return guarded(c, [&] {
    if (steps < 1) throw chan_error("steps must be >= 1");
    prepare_capture(c);
    c->h_dt_known = false;
    const size_t vb = vel_bytes(c), pb = p_bytes(c), db = sizeof(double) * c->prm.ny;
    char* bk = (char*)dalloc(2 * vb + pb + sizeof(DevScalars) + db + 256);
    auto save = [&](bool back) {
      auto cp = [&](void* a, void* b, size_t n) { d2d(c, back ? b : a, back ? a : b, n); };
      cp(bk, c->d_vel, vb);
      cp(bk + vb, c->d_R, vb);
      cp(bk + 2 * vb, c->d_P, pb);
      cp(bk + 2 * vb + pb, c->d_sc, sizeof(DevScalars));
      cp(bk + 2 * vb + pb + sizeof(DevScalars), c->d_dudy, db);
      CHAN_CHECK(hipStreamSynchronize(c->stream));
    };
    // Candidates: the x-pass kernel (the wavefront kernel only when a
    // subgrid closure is configured: it alone adds the stress to the fluxes)
    // and, for the wavefront kernel, its z-rows per wavefront where the
    // length has them; the z-forward kernel; the segmented solve's layouts
    // (4, 2 or 8 modes per wavefront where ny + 1 rows fit) or the
    // sweep-based solves with their block shapes.
    const int M = c->gd.nx2 / 2;
    const bool pf = c->tune.xpass_prefetch == 1 && M <= 144;
    std::vector<int> rows_c = {0};
    for (int r : {1, 2, 4})
      if (xpass_rows(M, c->gd.nz2, r, pf, false) == r && r != xpass_rows(M, c->gd.nz2, 0, pf, false))
        rows_c.push_back(r);
    std::vector<int> seg_c;
    if (solve_seg_rows(c->prm.ny)) seg_c.push_back(1);
    if (solve_seg_rows(c->prm.ny, 32) == 5) seg_c.push_back(2);
    if (solve_seg_rows(c->prm.ny, 8)) seg_c.push_back(3);
    std::vector<Tuning> cand;
    for (int w : {1, 0})
      for (int z : {1, 0}) {
        if (w == 0 && c->les) continue;
        for (int sg : seg_c)
          for (int r : rows_c) {
            if (r != 0 && w != 1) continue;  // rows per wavefront: the wavefront kernel only
            Tuning t = c->tune;
            t.xpass_wave = w; t.zfwd_v2 = z; t.solve_seg = sg; t.xpass_rows = r;
            cand.push_back(t);
          }
        for (int sl : {3, 1})
          for (int bl : {8, 16})
            for (int dm : {1, 0}) {
              if (dm && (sl != 3 || bl != 8)) continue;  // the ring variant has no block size
              Tuning t = c->tune;
              t.xpass_wave = w; t.zfwd_v2 = z; t.solve_seg = 0;
              t.solve_lanes = sl; t.solve_bl = bl; t.solve_dma = dm;
              cand.push_back(t);
            }
      }
    hipEvent_t a, b;
    CHAN_CHECK(hipEventCreate(&a));
    CHAN_CHECK(hipEventCreate(&b));
    double best = 1e300;
    Tuning bestt = c->tune;
    bool saved = false;
    try {
      save(false);
      saved = true;
      for (const Tuning& t : cand) {
        c->tune = t;
        step_body(c, 1, false);  // first launch of each variant, untimed
        CHAN_CHECK(hipEventRecord(a, c->stream));
        for (int s = 0; s < steps; ++s) step_body(c, 1, false);
        CHAN_CHECK(hipEventRecord(b, c->stream));
        CHAN_CHECK(hipEventSynchronize(b));
        float ms = 0.f;
        CHAN_CHECK(hipEventElapsedTime(&ms, a, b));
        // a candidate whose trial aborted on the CFL check or blew up is not
        // eligible; its state is discarded with the others' by the restore
        DevScalars tsc;
        d2h(c, &tsc, c->d_sc, sizeof(tsc));
        const bool ok = tsc.err == 0 && std::isfinite(tsc.cfl);
        save(true);
        if (ok && ms / steps < best) { best = ms / steps; bestt = t; }
      }
      if (!(best < 1e300)) throw chan_error("autotune: every candidate's trial steps failed the CFL check");
    } catch (...) {
      c->tune = bestt;
      if (saved) {  // put the state back where possible
        try {
          save(true);
        } catch (...) {
        }
      }
      (void)hipEventDestroy(a);
      (void)hipEventDestroy(b);
      dfree(bk);
      throw;
    }
    c->tune = bestt;
    (void)hipEventDestroy(a);
    (void)hipEventDestroy(b);
    dfree(bk);
    if (best_ms) *best_ms = best;
  });
}

int chan_ensemble_create(int n, chan_ctx* const* members, int nstreams, int linearized, chan_ensemble** out) {
  if (!out || !members || n < 1) return fail(nullptr, "invalid ensemble arguments");
  *out = nullptr;
  chan_ensemble* e = new chan_ensemble();
  try {
    for (int i = 0; i < n; ++i) {
      if (!members[i]) throw chan_error("null ensemble member");
      for (int k = 0; k < i; ++k)
        if (members[k] == members[i]) throw chan_error("ensemble member given twice");
      e->m.push_back(members[i]);
    }
    if (nstreams < 0) throw chan_error("nstreams must be >= 0");
    e->batched = (nstreams == 0);
    if (e->batched)
      for (int i = 0; i < n; ++i) {
        const std::string why = batch_ineligible(members[i], linearized != 0);
        if (!why.empty()) throw chan_error("member " + std::to_string(i) + " cannot join the batched launches: " + why);
      }
    const int ns = e->batched ? 0 : std::max(1, std::min(nstreams, n));
    CHAN_CHECK(hipStreamCreateWithFlags(&e->master, hipStreamNonBlocking));
    if (e->batched) {
      // graph branches per phase (CHANFLOW_ENSEMBLE_BRANCHES, default 4, the
      // hardware queues a process gets; 1: the launches of a phase in turn)
      int nbr = 4;
      if (const char* v = getenv("CHANFLOW_ENSEMBLE_BRANCHES")) nbr = std::max(1, std::min(8, atoi(v)));
      const int nb = nbr > 1 ? std::min(n, nbr) : 0;
      e->bstreams.resize(nb, nullptr);
      e->bjoin.resize(nb, nullptr);
      CHAN_CHECK(hipEventCreateWithFlags(&e->bfork, hipEventDisableTiming));
      for (int k = 0; k < nb; ++k) {
        CHAN_CHECK(hipStreamCreateWithFlags(&e->bstreams[k], hipStreamNonBlocking));
        CHAN_CHECK(hipEventCreateWithFlags(&e->bjoin[k], hipEventDisableTiming));
      }
    }
    CHAN_CHECK(hipEventCreateWithFlags(&e->start, hipEventDisableTiming));
    e->streams.resize(ns, nullptr);
    e->done.resize(ns, nullptr);
    for (int k = 0; k < ns; ++k) {
      CHAN_CHECK(hipStreamCreateWithFlags(&e->streams[k], hipStreamNonBlocking));
      CHAN_CHECK(hipEventCreateWithFlags(&e->done[k], hipEventDisableTiming));
    }
    e->lin = linearized;
  } catch (const chan_error& err) {
    g_create_error = err.msg;
    chan_ensemble_destroy(e);
    return 1;
  }
  *out = e;
  return 0;
}

const char* chan_ensemble_last_error(const chan_ensemble* e) {
  return e ? e->last_error.c_str() : g_create_error.c_str();
}

int chan_ensemble_step(chan_ensemble* e, int nsteps) {
  if (!e) return fail(nullptr, "null ensemble");
  try {
    if (nsteps < 0) throw chan_error("nsteps must be >= 0");
    for (chan_ctx* c : e->m) {
      check_ops(c);
      ensure_vel_in_vel(c);
      c->h_dt_known = false;
      CHAN_CHECK(hipStreamSynchronize(c->stream));
    }
    if (!e->exec) ensemble_capture(e);
    for (int s = 0; s < nsteps; ++s) CHAN_CHECK(hipGraphLaunch(e->exec, e->master));
    CHAN_CHECK(hipStreamSynchronize(e->master));
    std::string errs;
    for (size_t i = 0; i < e->m.size(); ++i) {
      try {
        check_err(e->m[i]);
      } catch (const chan_error& err) {
        errs += "member " + std::to_string(i) + ": " + err.msg + "; ";
      }
    }
    if (!errs.empty()) throw chan_error(errs);
    return 0;
  } catch (const chan_error& err) {
    e->last_error = err.msg;
  } catch (const std::exception& err) {
    e->last_error = err.what();
  }
  return 1;
}

// Drops the captured graph, e.g. after a member's tuning or gains changed.
int chan_ensemble_invalidate(chan_ensemble* e) {
  if (!e) return 1;
  ensemble_release(e);
  return 0;
}

// Times `steps` replayed steps of each batched launch plan (the x-pass plane
// ranges spread over 1, 2, 4 or 8 times each length's resident workgroups,
// and the default, as many times as there are lengths), restoring every
// member's state after each trial, and keeps the fastest.  The plans differ
// only in how the x-pass planes are cut into workgroup ranges, so every
// member's arithmetic is the same under each of them.
int chan_ensemble_autotune(chan_ensemble* e, int steps, double* best_ms) {
  if (!e) return fail(nullptr, "null ensemble");
  std::vector<void*> bk;
  hipEvent_t a = nullptr, b = nullptr;
  const int xmul0 = e->xmul, xrows0 = e->xrows1, solve0 = e->solve4;  // kept if the tuning fails
  auto cleanup = [&] {
    for (void* p : bk) dfree(p);
    bk.clear();
    if (a) (void)hipEventDestroy(a);
    if (b) (void)hipEventDestroy(b);
  };
  // every member's state: velocity, RK register, pressure, scalars, d_um_dy
  auto sizes = [](chan_ctx* c) {
    return std::array<size_t, 5>{vel_bytes(c), vel_bytes(c), p_bytes(c), sizeof(DevScalars),
                                 sizeof(double) * (size_t)c->prm.ny};
  };
  auto ptrs = [](chan_ctx* c) {
    return std::array<void*, 5>{c->d_vel, c->d_R, c->d_P, c->d_sc, c->d_dudy};
  };
  auto copy_state = [&](bool back) {
    for (size_t i = 0; i < e->m.size(); ++i) {
      chan_ctx* c = e->m[i];
      const auto sz = sizes(c);
      const auto p = ptrs(c);
      char* q = static_cast<char*>(bk[i]);
      for (int k = 0; k < 5; ++k) {
        if (back) d2d(c, p[k], q, sz[k]);
        else d2d(c, q, p[k], sz[k]);
        q += sz[k];
      }
      CHAN_CHECK(hipStreamSynchronize(c->stream));
    }
  };
  bool saved = false;  // the backups hold every member's state
  auto restore = [&] {  // on failure: put the saved state back where possible
    if (!saved) return;
    try {
      if (e->master) (void)hipStreamSynchronize(e->master);
      copy_state(true);
    } catch (...) {
    }
  };
  try {
    if (steps < 1) throw chan_error("steps must be >= 1");
    if (!e->batched) throw chan_error("autotune applies to the batched launches (nstreams = 0); tune the members with chan_autotune");
    for (chan_ctx* c : e->m) {
      check_ops(c);
      ensure_vel_in_vel(c);
      c->h_dt_known = false;
      CHAN_CHECK(hipStreamSynchronize(c->stream));
    }
    for (chan_ctx* c : e->m) {
      const auto sz = sizes(c);
      bk.push_back(dalloc(sz[0] + sz[1] + sz[2] + sz[3] + sz[4]));
    }
    CHAN_CHECK(hipEventCreate(&a));
    CHAN_CHECK(hipEventCreate(&b));
    // one plan's trial: capture, one untimed and `steps` timed replays, the
    // members' state restored; +inf if a member tripped the CFL check
    auto trial = [&]() {
      ensemble_release(e);
      ensemble_capture(e);
      CHAN_CHECK(hipGraphLaunch(e->exec, e->master));  // first replay, untimed
      CHAN_CHECK(hipEventRecord(a, e->master));
      for (int t = 0; t < steps; ++t) CHAN_CHECK(hipGraphLaunch(e->exec, e->master));
      CHAN_CHECK(hipEventRecord(b, e->master));
      CHAN_CHECK(hipEventSynchronize(b));
      float ms = 0.f;
      CHAN_CHECK(hipEventElapsedTime(&ms, a, b));
      bool ok = true;
      for (chan_ctx* c : e->m) {
        DevScalars tsc;
        d2h(c, &tsc, c->d_sc, sizeof(tsc));
        ok = ok && tsc.err == 0 && std::isfinite(tsc.cfl);
      }
      copy_state(true);
      return ok ? (double)ms / steps : 1e300;
    };
    double best = 1e300;
    int bestm = xmul0, bestr = 0, bests = 0;
    copy_state(false);
    saved = true;
    // the x-pass plane ranges with the default tile shapes, then the tile
    // shapes (one z-row per x-pass wavefront; 4 modes per solve wavefront)
    // with the best ranges
    e->xrows1 = 0;
    e->solve4 = 0;
    for (int m : {0, 1, 2, 4, 8}) {
      e->xmul = m;
      const double ms = trial();
      if (ms < best) { best = ms; bestm = m; }
    }
    e->xmul = bestm;
    for (int v : {1, 2, 3}) {
      e->xrows1 = v & 1;
      e->solve4 = v >> 1;
      const double ms = trial();
      if (ms < best) { best = ms; bestr = e->xrows1; bests = e->solve4; }
    }
    if (!(best < 1e300)) throw chan_error("autotune: every plan's trial steps failed the CFL check");
    e->xmul = bestm;
    e->xrows1 = bestr;
    e->solve4 = bests;
    ensemble_release(e);
    cleanup();
    if (best_ms) *best_ms = best;
    return 0;
  } catch (const chan_error& err) {
    e->last_error = err.msg;
  } catch (const std::exception& err) {
    e->last_error = err.what();
  }
  restore();
  cleanup();
  e->xmul = xmul0;
  e->xrows1 = xrows0;
  e->solve4 = solve0;
  ensemble_release(e);  // the next step plans and captures again
  return 1;
}

int chan_ensemble_destroy(chan_ensemble* e) {
  if (!e) return 0;
  if (e->master) (void)hipStreamSynchronize(e->master);
  ensemble_release(e);
  for (hipEvent_t ev : e->done)
    if (ev) (void)hipEventDestroy(ev);
  for (hipStream_t st : e->streams)
    if (st) (void)hipStreamDestroy(st);
  for (hipEvent_t ev : e->bjoin)
    if (ev) (void)hipEventDestroy(ev);
  for (hipStream_t st : e->bstreams)
    if (st) (void)hipStreamDestroy(st);
  if (e->bfork) (void)hipEventDestroy(e->bfork);
  if (e->start) (void)hipEventDestroy(e->start);
  if (e->master) (void)hipStreamDestroy(e->master);
  delete e;
  return 0;
}

}  // extern "C"
