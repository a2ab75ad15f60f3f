// Fourier-space part of the step: the per-(kx,kz) low-storage RK3 /
// Crank-Nicolson update with fractional-step projection (PFC/pass2.f90:29-352),
// the Helmholtz/Poisson banded solves (PFC/solve.f90:14-338) and the
// mean-mode constant-mass-flux pressure gradient (PFC/pass2.f90:356-375).
//
// One thread owns one (kx,kz) mode; all state arrays are plane-major with
// kz fastest, so a wavefront reads 64 consecutive modes of one plane at a
// time (coalesced).  The y-recurrences stream through three passes:
//   A (ascending)  : assemble the RHS (explicit diffusion and pressure
//                    gradient at rk_step 1) and eliminate the u, w (D2H) and
//                    v (D2N) Helmholtz systems bottom-up,
//   B (descending) : back-substitute, form the divergence of the provisional
//                    velocity and eliminate the Neumann Poisson system
//                    top-down,
//   C (ascending)  : back-substitute the pressure correction, apply the wall
//                    extrapolation, project, update the pressure
//                    (Brown-Cortez-Minion form, :307-309) and assemble the
//                    next RK register.
// The eliminated coefficients live in a plane-major scratch array.
#include "chan_internal.h"

// RK3 coefficients: default-REAL literal quotients widened to double
// (PFC/pass2.f90:14-18).
__constant__ double c_rka[3] = {0.0, (double)(-17.0f / 60.0f), (double)(-5.0f / 12.0f)};
__constant__ double c_rkb[3] = {(double)(8.0f / 15.0f), (double)(5.0f / 12.0f), (double)(3.0f / 4.0f)};
__constant__ double c_rkc[3] = {(double)(29.0f / 96.0f), (double)(-3.0f / 40.0f), (double)(1.0f / 6.0f)};
__constant__ double c_rkd[3] = {(double)(37.0f / 160.0f), (double)(5.0f / 24.0f), (double)(1.0f / 6.0f)};

__device__ __forceinline__ double2 cz() { return make_double2(0.0, 0.0); }
__device__ __forceinline__ double2 rmul(double s, double2 a) { return make_double2(s * a.x, s * a.y); }
// i*s*a
__device__ __forceinline__ double2 imul(double s, double2 a) { return make_double2(-s * a.y, s * a.x); }


// Row-blocked sweeps.  Every pass walks the y rows of its mode with a short
// recurrence; the loads of a block of BL rows do not depend on it, so they
// are all issued before the block's recurrence runs (one memory round trip
// per block instead of one per row).
template <int BL, class LoadF, class CompF>
__device__ __forceinline__ void rows_up(int j0, int j1, LoadF load, CompF comp) {
  using RD = decltype(load(0));
  int j = j0;
  for (; j + BL - 1 <= j1; j += BL) {
    RD d[BL];
#pragma unroll
    for (int t = 0; t < BL; ++t) d[t] = load(j + t);
#pragma unroll
    for (int t = 0; t < BL; ++t) comp(j + t, d[t]);
  }
  for (; j <= j1; ++j) {
    RD d = load(j);
    comp(j, d);
  }
}

template <int BL, class LoadF, class CompF>
__device__ __forceinline__ void rows_down(int j1, int j0, LoadF load, CompF comp) {
  using RD = decltype(load(0));
  int j = j1;
  for (; j - BL + 1 >= j0; j -= BL) {
    RD d[BL];
#pragma unroll
    for (int t = 0; t < BL; ++t) d[t] = load(j - t);
#pragma unroll
    for (int t = 0; t < BL; ++t) comp(j - t, d[t]);
  }
  for (; j >= j0; --j) {
    RD d = load(j);
    comp(j, d);
  }
}

struct RowA { double2 n0, n1, n2, r0, r1, r2, p; };
struct RowB { double2 s0, s1, s2; double c0, c1; };
struct RowC { double2 s3, s0, s1, s2, p, n0, n1, n2; double c0; };

template <int SOLVE_BL>
__global__ __launch_bounds__(64) void k_solve(DevOps ops, SolveArgs a) {
  const long nm = (long)a.nxh * a.nz;
  const long m = (long)blockIdx.x * blockDim.x + threadIdx.x;
  if (m >= nm) return;
  const int ix = (int)(m / a.nz), iz = (int)(m - (long)ix * a.nz);
  const int ny = a.ny, nyp = ny + 1;
  const long MS = nm;
  const double kxv = ops.kx[ix], kzv = ops.kz[iz];
  const double k2 = kxv * kxv + kzv * kzv;
  const bool mean = (k2 == 0.0);
  const int r = a.rk - 1;
  const bool first = (r == 0);
  const double dt = a.sc->dt;
  const double re = a.re;

  const double2* __restrict__ NL = a.nl;
  double2* __restrict__ R = a.R;
  double2* __restrict__ P = a.P;
  double2* __restrict__ U = a.vel;
  double2* __restrict__ S = a.scr;
  double* __restrict__ SR = reinterpret_cast<double*>(a.scr + 4L * nyp * MS);
  auto nl = [&](int c, int j) { return NL[((long)c * nyp + j) * MS + m]; };
  auto rr = [&](int c, int j) { return R[((long)c * nyp + j) * MS + m]; };
  auto pp = [&](int j) { return P[(long)j * MS + m]; };
  auto sidx = [&](int c, int j) { return ((long)c * nyp + j) * MS + m; };
  auto neg = [](double2 v) { return make_double2(-v.x, -v.y); };
  const double* d2h = ops.d2h;
  const double* d2n = ops.d2n;
  const double* dm = ops.dm;

  const double vis_imp = c_rkd[r] * (dt / re);
  const double rk_press = (c_rkc[r] + c_rkd[r]) * dt;
  const double coef = c_rkb[r] * dt;
  const double vis_exp1 = c_rkc[r] * (dt / re);
  const double alpha = 1.0 + vis_imp * k2;

  const double2 top0 = a.cbc[(0L * 3 + 0) * MS + m], top1 = a.cbc[(0L * 3 + 1) * MS + m],
                top2 = a.cbc[(0L * 3 + 2) * MS + m];
  const double2 bot0 = a.cbc[(1L * 3 + 0) * MS + m], bot1 = a.cbc[(1L * 3 + 1) * MS + m],
                bot2 = a.cbc[(1L * 3 + 2) * MS + m];

  // mean-mode pressure gradient (compute_p_grad, :369-373)
  double p_grad = a.sc->p_grad;
  if (mean && a.fix_mass == 1) {
    double tmp = 0.0;
#pragma unroll 8
    for (int j = 1; j < ny; ++j) tmp = tmp + ops.dyh[j] * (-nl(0, j).x);
    p_grad = -tmp + 0.5 * (a.dudy[ny - 1] - a.dudy[0]) / re;
    a.sc->p_grad = p_grad;
  }

  // ---------------- pass A: RHS and forward elimination ----------------
  // Rows 0 and ny of u, w and rows 0 and ny-1 of v are Dirichlet rows whose
  // right-hand side is the wall value (SOLVE_HELMHOLTZ, :174-213).
  double cH = 0.0, cN = 0.0;
  double2 fu = bot0, fw = bot2, fv = bot1;
  S[sidx(0, 0)] = fu;
  S[sidx(1, 0)] = fw;
  SR[sidx(0, 0)] = 0.0;
  if (!mean) {
    S[sidx(2, 0)] = fv;
    SR[sidx(1, 0)] = 0.0;
  }
  // window of the explicit-diffusion stencil: R at rows j-1, j and P at row j
  double2 Rm0 = rr(0, 0), Rm1 = rr(1, 0), Rm2 = rr(2, 0);
  double2 Rc0 = rr(0, 1), Rc1 = rr(1, 1), Rc2 = rr(2, 1);
  double2 Pc = first ? pp(1) : cz();
  auto loadA = [&](int j) {
    RowA d;
    d.n0 = nl(0, j); d.n1 = nl(1, j); d.n2 = nl(2, j);
    d.r0 = rr(0, j + 1); d.r1 = rr(1, j + 1); d.r2 = rr(2, j + 1);
    d.p = first ? pp(j + 1) : cz();
    return d;
  };
  auto compA = [&](int j, const RowA& d) {
    double2 v0 = cadd(Rc0, rmul(coef, neg(d.n0)));
    double2 v1 = cadd(Rc1, rmul(coef, neg(d.n1)));
    double2 v2 = cadd(Rc2, rmul(coef, neg(d.n2)));
    if (first) {
      // add_explicit_diffusive_terms(cv, cru) (:339-351)
      const double a0 = d2h[3 * j], a1 = d2h[3 * j + 1], a2 = d2h[3 * j + 2];
      double2 du = cadd(cadd(rmul(a0, Rm0), rmul(a1, Rc0)), rmul(a2, d.r0));
      v0 = cadd(v0, rmul(vis_exp1, csub(du, rmul(k2, Rc0))));
      double2 dw = cadd(cadd(rmul(a0, Rm2), rmul(a1, Rc2)), rmul(a2, d.r2));
      v2 = cadd(v2, rmul(vis_exp1, csub(dw, rmul(k2, Rc2))));
      if (j <= ny - 2) {
        const double b0 = d2n[3 * j], b1 = d2n[3 * j + 1], b2 = d2n[3 * j + 2];
        double2 dv = cadd(cadd(rmul(b0, Rm1), rmul(b1, Rc1)), rmul(b2, d.r1));
        v1 = cadd(v1, rmul(vis_exp1, csub(dv, rmul(k2, Rc1))));
      }
      if (!mean) {  // subtract_pressure_gradient(cv, cp, rk_press) (:54-77)
        v0 = csub(v0, imul(rk_press * kxv, Pc));
        v2 = csub(v2, imul(rk_press * kzv, Pc));
        if (j <= ny - 2) v1 = csub(v1, rmul(rk_press * ops.dn[j], csub(d.p, Pc)));
      }
    }
    if (mean) {  // mean pressure gradient (:220-221)
      v0 = csub(v0, make_double2(a.cosb * rk_press * p_grad, 0.0));
      v2 = csub(v2, make_double2(a.sinb * rk_press * p_grad, 0.0));
    }
    {
      const double ah = -vis_imp * d2h[3 * j], bh = -vis_imp * d2h[3 * j + 1] + alpha,
                   ch = -vis_imp * d2h[3 * j + 2];
      const double bp = bh - ah * cH;
      const double2 gu = csub(v0, rmul(ah, fu));
      const double2 gw = csub(v2, rmul(ah, fw));
      const double rb = 1.0 / bp;
      cH = ch * rb;
      fu = rmul(rb, gu);
      fw = rmul(rb, gw);
      S[sidx(0, j)] = fu;
      S[sidx(1, j)] = fw;
      SR[sidx(0, j)] = cH;
    }
    if (!mean) {
      if (j <= ny - 2) {
        const double an = -vis_imp * d2n[3 * j], bn = -vis_imp * d2n[3 * j + 1] + alpha,
                     cn = -vis_imp * d2n[3 * j + 2];
        const double bq = bn - an * cN;
        const double2 gv = csub(v1, rmul(an, fv));
        const double rq = 1.0 / bq;
        cN = cn * rq;
        fv = rmul(rq, gv);
      } else {
        cN = 0.0;
        fv = top1;
      }
      S[sidx(2, j)] = fv;
      SR[sidx(1, j)] = cN;
    }
    Rm0 = Rc0; Rm1 = Rc1; Rm2 = Rc2;
    Rc0 = d.r0; Rc1 = d.r1; Rc2 = d.r2;
    Pc = d.p;
  };
  rows_up<SOLVE_BL>(1, ny - 1, loadA, compA);
  S[sidx(0, ny)] = top0;
  S[sidx(1, ny)] = top2;
  SR[sidx(0, ny)] = 0.0;

  // ---------------- pass B: back substitution + Poisson elimination ----------------
  const double tmpc = 1.0 / rk_press;
  const double alp = -k2;
  double ap = 0.0;   // a'_{i+1} of the reversed Poisson elimination
  double2 fp = cz();
  auto poisson_row = [&](int i, double& a_, double& b_, double& c_) {
    const double* dh = ops.dh; const double* dn = ops.dn;
    if (i == 0) { a_ = 0.0; b_ = -dn[0]; c_ = dn[0]; }
    else if (i == 1) { a_ = dh[1] * dn[0]; b_ = -dh[1] * dn[0] - dh[1] * dn[1] + alp; c_ = dh[1] * dn[1]; }
    else if (i == ny - 1) {
      a_ = dh[ny - 1] * dn[ny - 2]; b_ = -dh[ny - 1] * dn[ny - 2] - dh[ny - 1] * dn[ny - 1] + alp;
      c_ = dh[ny - 1] * dn[ny - 1];
    } else if (i == ny) { a_ = -dn[ny - 1]; b_ = dn[ny - 1]; c_ = 0.0; }
    else { a_ = d2h[3 * i]; b_ = d2h[3 * i + 1] + alp; c_ = d2h[3 * i + 2]; }
  };
  auto poisson_elim = [&](int i, double2 rhs) {  // reversed Thomas, row i (descending)
    double a_, b_, c_;
    poisson_row(i, a_, b_, c_);
    const double bp2 = (i == ny) ? b_ : b_ - c_ * ap;
    const double2 g = (i == ny) ? rhs : csub(rhs, rmul(c_, fp));
    const double rb2 = 1.0 / bp2;
    ap = a_ * rb2;
    fp = rmul(rb2, g);
    S[sidx(3, i)] = fp;
    SR[sidx(0, i)] = ap;
  };
  // row ny: Dirichlet rows of u, w; v has no row ny
  double2 xu1 = top0, xw1 = top2, xv1 = cz();
  S[sidx(2, ny)] = cz();
  auto loadB = [&](int j) {
    RowB d;
    d.s0 = S[sidx(0, j)]; d.s1 = S[sidx(1, j)];
    d.s2 = mean ? cz() : S[sidx(2, j)];
    d.c0 = SR[sidx(0, j)]; d.c1 = mean ? 0.0 : SR[sidx(1, j)];
    return d;
  };
  auto compB = [&](int j, const RowB& d) {
    const double2 xu = csub(d.s0, rmul(d.c0, xu1));
    const double2 xw = csub(d.s1, rmul(d.c0, xw1));
    double2 xv;
    if (mean) xv = bot1;  // (:261-263)
    else if (j == ny - 1) xv = d.s2;
    else xv = csub(d.s2, rmul(d.c1, xv1));
    S[sidx(0, j)] = xu;
    S[sidx(1, j)] = xw;
    S[sidx(2, j)] = xv;
    if (!mean) {
      const int i = j + 1;
      double2 rhs = cz();
      if (i <= ny - 1) {  // strip_divergence rhs (:28-31)
        double2 dv = cadd(cadd(imul(kxv, xu1), rmul(ops.dh[i], csub(xv1, xv))), imul(kzv, xw1));
        rhs = rmul(tmpc, dv);
      }
      poisson_elim(i, rhs);
    }
    xu1 = xu; xw1 = xw; xv1 = xv;
  };
  rows_down<SOLVE_BL>(ny - 1, 0, loadB, compB);
  if (!mean) poisson_elim(0, cz());

  // ---------------- pass C: pressure correction, projection, next RK register ----------------
  // The Poisson back substitution (ascending) runs two rows ahead of the
  // projection, which needs phi at rows i-1..i+1; phi at the walls is
  // re-extrapolated from the interior (SOLVE_P :335-336, zero wall data).
  const bool more = (a.rk < 3);
  const double coef2 = more ? dt * c_rka[r + 1] : 0.0;
  const double vis_exp2 = more ? (dt / re) * c_rkc[r + 1] : 0.0;
  const double rk_press2 = more ? (c_rkc[r + 1] + c_rkd[r + 1]) * dt : 0.0;
  const double alpha2 = 1.0 + vis_imp * k2;
  double2 w0[4][3];   // cv at rows i-3..i
  double2 Pw0 = cz(), Pw1 = cz();  // P_new at rows i-1, i
#pragma unroll
  for (int t = 0; t < 4; ++t)
#pragma unroll
    for (int c = 0; c < 3; ++c) w0[t][c] = cz();
  // buildR: next RK register at row j from cv at row j, the D2H stencil rows
  // st (three rows), the D2N stencil rows stv, P_new at rows j, j+1 and NL(j).
  auto buildR = [&](int j, const double2* cvj, const double2 (*st)[3], const double2 (*stv)[3], double2 Pj,
                    double2 Pj1, double2 n0, double2 n1, double2 n2) {
    double2 out[3];
    out[0] = cadd(cvj[0], rmul(coef2, neg(n0)));
    out[1] = cadd(cvj[1], rmul(coef2, neg(n1)));
    out[2] = cadd(cvj[2], rmul(coef2, neg(n2)));
    const double a0 = d2h[3 * j], a1 = d2h[3 * j + 1], a2 = d2h[3 * j + 2];
    double2 du = cadd(cadd(rmul(a0, st[0][0]), rmul(a1, st[1][0])), rmul(a2, st[2][0]));
    double2 dw = cadd(cadd(rmul(a0, st[0][2]), rmul(a1, st[1][2])), rmul(a2, st[2][2]));
    out[0] = cadd(out[0], rmul(vis_exp2, csub(du, rmul(k2, cvj[0]))));
    out[2] = cadd(out[2], rmul(vis_exp2, csub(dw, rmul(k2, cvj[2]))));
    if (j < ny) {
      const double b0 = d2n[3 * j], b1 = d2n[3 * j + 1], b2 = d2n[3 * j + 2];
      double2 dv = cadd(cadd(rmul(b0, stv[0][1]), rmul(b1, stv[1][1])), rmul(b2, stv[2][1]));
      out[1] = cadd(out[1], rmul(vis_exp2, csub(dv, rmul(k2, cvj[1]))));
    }
    if (!mean) {
      if (j >= 1 && j <= ny - 1) {
        out[0] = csub(out[0], imul(rk_press2 * kxv, Pj));
        out[2] = csub(out[2], imul(rk_press2 * kzv, Pj));
      }
      if (j >= 1 && j <= ny - 2) out[1] = csub(out[1], rmul(rk_press2 * ops.dn[j], csub(Pj1, Pj)));
    }
#pragma unroll
    for (int c = 0; c < 3; ++c) R[((long)c * nyp + j) * MS + m] = out[c];
  };
  // projection of row i: s0,s1,s2 = provisional u, w, v; Pold = P(i); f0 = phi(i),
  // fn = phi(i+1), st0..st2 = phi on the D2H stencil rows of i; nb* = NL(i-1).
  auto c2row = [&](int i, double2 s0, double2 s1, double2 s2, double2 Pold, double2 f0, double2 fn,
                   double2 st0, double2 st1, double2 st2, double2 nb0, double2 nb1, double2 nb2) {
    double2 cv[3] = {s0, s2, s1};  // (u, v, w)
    double2 pn = cz();
    if (!mean) {
      if (i >= 1 && i <= ny - 1) {
        cv[0] = csub(cv[0], imul(rk_press * kxv, f0));
        cv[2] = csub(cv[2], imul(rk_press * kzv, f0));
      }
      if (i >= 1 && i <= ny - 2) cv[1] = csub(cv[1], rmul(rk_press * ops.dn[i], csub(fn, f0)));
      double2 d2 = cadd(cadd(rmul(d2h[3 * i], st0), rmul(d2h[3 * i + 1], st1)), rmul(d2h[3 * i + 2], st2));
      pn = csub(cadd(Pold, rmul(alpha2, f0)), rmul(vis_imp, d2));
    }
#pragma unroll
    for (int c = 0; c < 3; ++c) U[((long)c * nyp + i) * MS + m] = cv[c];
    P[(long)i * MS + m] = pn;
#pragma unroll
    for (int t = 0; t < 3; ++t)
#pragma unroll
      for (int c = 0; c < 3; ++c) w0[t][c] = w0[t + 1][c];
#pragma unroll
    for (int c = 0; c < 3; ++c) w0[3][c] = cv[c];
    Pw0 = Pw1; Pw1 = pn;
    if (mean) {  // d_um_dy update from the new mean profile (:84-89)
      if (i >= 2 && i <= ny - 1) a.dudy[i - 1] = (ops.dn[i - 1] * (w0[3][0].x - w0[2][0].x));
      if (i == 2) a.dudy[0] = dm[0] * w0[1][0].x + dm[1] * w0[2][0].x + dm[2] * w0[3][0].x;
      if (i == ny) a.dudy[ny - 1] = dm[3] * w0[1][0].x + dm[4] * w0[2][0].x + dm[5] * w0[3][0].x;
    }
    if (!more) {
#pragma unroll
      for (int c = 0; c < 3; ++c) R[((long)c * nyp + i) * MS + m] = cv[c];
      return;
    }
    if (i >= 2) {
      if (i == 2)  // row 0 (clamped stencils on rows 0..2, no gradient)
        buildR(0, w0[1], &w0[1], &w0[1], cz(), cz(), nl(0, 0), nl(1, 0), nl(2, 0));
      // row j = i-1: centred D2H on rows i-2..i; D2N clamped at the last v row
      const int j = i - 1;
      const double2 (*stv)[3] = (j == ny - 1) ? &w0[0] : &w0[1];
      buildR(j, w0[2], &w0[1], stv, Pw0, Pw1, nb0, nb1, nb2);
    }
  };
  auto s3 = [&](int i) { return mean ? cz() : S[sidx(3, i)]; };
  auto sr0 = [&](int i) { return mean ? 0.0 : SR[sidx(0, i)]; };
  double2 ph0 = s3(0);
  const double2 ph1 = csub(s3(1), rmul(sr0(1), ph0));
  const double2 ph2 = csub(s3(2), rmul(sr0(2), ph1));
  const double2 pb = make_double2((0.0 - dm[1] * ph1.x - dm[2] * ph2.x) / dm[0],
                                  (0.0 - dm[1] * ph1.y - dm[2] * ph2.y) / dm[0]);
  c2row(0, S[sidx(0, 0)], S[sidx(1, 0)], S[sidx(2, 0)], pp(0), pb, cz(), pb, ph1, ph2, cz(), cz(), cz());
  c2row(1, S[sidx(0, 1)], S[sidx(1, 1)], S[sidx(2, 1)], pp(1), ph1, ph2, pb, ph1, ph2, cz(), cz(), cz());
  double2 q0 = pb, q1 = ph1, q2 = ph2, ph = ph2;  // phi at rows i-3..i-1 and the chain value
  auto loadC = [&](int i) {
    RowC d;
    d.s3 = s3(i); d.c0 = sr0(i);
    d.s0 = S[sidx(0, i - 1)]; d.s1 = S[sidx(1, i - 1)]; d.s2 = S[sidx(2, i - 1)];
    d.p = pp(i - 1);
    if (more) { d.n0 = nl(0, i - 2); d.n1 = nl(1, i - 2); d.n2 = nl(2, i - 2); }
    else { d.n0 = d.n1 = d.n2 = cz(); }
    return d;
  };
  auto compC = [&](int i, const RowC& d) {
    ph = csub(d.s3, rmul(d.c0, ph));
    q0 = q1; q1 = q2; q2 = ph;
    c2row(i - 1, d.s0, d.s1, d.s2, d.p, q1, q2, q0, q1, q2, d.n0, d.n1, d.n2);
  };
  rows_up<SOLVE_BL>(3, ny - 1, loadC, compC);
  // top wall: phi(ny) from phi(ny-2), phi(ny-1)
  const double2 pt = make_double2((0.0 - dm[3] * q1.x - dm[4] * q2.x) / dm[5],
                                  (0.0 - dm[3] * q1.y - dm[4] * q2.y) / dm[5]);
  if (more) {
    c2row(ny - 1, S[sidx(0, ny - 1)], S[sidx(1, ny - 1)], S[sidx(2, ny - 1)], pp(ny - 1), q2, pt, q1, q2, pt,
          nl(0, ny - 2), nl(1, ny - 2), nl(2, ny - 2));
    c2row(ny, S[sidx(0, ny)], S[sidx(1, ny)], S[sidx(2, ny)], pp(ny), pt, cz(), q1, q2, pt,
          nl(0, ny - 1), nl(1, ny - 1), nl(2, ny - 1));
    buildR(ny, w0[3], &w0[1], &w0[1], Pw1, cz(), nl(0, ny), nl(1, ny), nl(2, ny));
  } else {
    c2row(ny - 1, S[sidx(0, ny - 1)], S[sidx(1, ny - 1)], S[sidx(2, ny - 1)], pp(ny - 1), q2, pt, q1, q2, pt,
          cz(), cz(), cz());
    c2row(ny, S[sidx(0, ny)], S[sidx(1, ny)], S[sidx(2, ny)], pp(ny), pt, cz(), q1, q2, pt, cz(), cz(), cz());
  }
}

// ---------------------------------------------------------------------------
// Component-parallel variant: three lanes per (kx,kz) mode, lane c owning
// velocity component c (0 = u, 1 = v, 2 = w), five modes per 16-lane row
// (lane 15 of each row idles), 20 modes per wavefront.  Each lane runs the
// RHS assembly, Helmholtz elimination, back substitution, projection and
// next-register assembly of its own component; the Poisson elimination, its
// back substitution and the pressure update are per-mode and run
// redundantly on the three lanes, which exchange their divergence terms
// with lane shuffles.  Arithmetic and its order per value are those of
// k_solve above.
//
// Addressing: every array is indexed as base + (32-bit byte offset), with
// the offset advanced once per row, so a load costs one vector add.
// Scratch (the a.scr buffer): SX[3][ny+1][modes] (f, then x, per
// component), SP[ny+1][modes] (Poisson f), SC[3][ny+1][modes] (Helmholtz
// c'), SA[ny+1][modes] (Poisson a').
// ---------------------------------------------------------------------------
struct RowA3 { double2 n, r, p; };
struct RowB3 { double2 s; double cc; };
struct RowC3 { double2 sp, sx, p, n; double sa; };

__device__ __forceinline__ double2 ldo(const double2* b, uint32_t off) {
  return *reinterpret_cast<const double2*>(reinterpret_cast<const char*>(b) + off);
}
__device__ __forceinline__ double ldo1(const double* b, uint32_t off) {
  return *reinterpret_cast<const double*>(reinterpret_cast<const char*>(b) + off);
}
__device__ __forceinline__ void sto(double2* b, uint32_t off, double2 v) {
  *reinterpret_cast<double2*>(reinterpret_cast<char*>(b) + off) = v;
}
__device__ __forceinline__ void sto1(double* b, uint32_t off, double v) {
  *reinterpret_cast<double*>(reinterpret_cast<char*>(b) + off) = v;
}
__device__ __forceinline__ double2 shfl2(double2 v, int src) {
  return make_double2(__shfl(v.x, src, 64), __shfl(v.y, src, 64));
}

// ---------------------------------------------------------------------------
// Segmented column solve: the whole y column of every mode stays on chip.
//
// A wavefront owns MODES consecutive (kx,kz) modes; lane (s, q) = (lane /
// MODES, lane % MODES) owns rows s*SEG .. s*SEG+SEG-1 of mode q, all three
// velocity components and the pressure correction, in registers.  The last
// segment holds the remaining 3..SEG rows; rows past ny are identity padding.
// Each tridiagonal system (u and w with D2H, v with D2N, the Neumann Poisson
// system) is solved by partition (seg_eliminate, seg_separators, seg_backsub):
//   1. every segment eliminates its interior rows top-down, carrying the
//      coupling to the separator below it (the previous segment's last row)
//      as a second right-hand side; the coupling to the separator above it
//      (its own last row, except in the last segment) stays in c';
//   2. with the interiors eliminated, the separators form a tridiagonal
//      system of (active segments - 1) unknowns per mode, solved by parallel
//      cyclic reduction over lane shuffles;
//   3. every segment back-substitutes with both separators known.
// The state is read once (NL, R and P; the NL and P rows are kept in LDS for
// the pressure update and the next RK register) and written once (u, P, R):
// no scratch array.  The arithmetic per row is that
// of k_solve; only the elimination order of the solves differs, so results
// agree with it to rounding.
// ---------------------------------------------------------------------------

// Partitioned solve of NC right-hand sides X[c][0..SEG) that share one
// matrix, in three calls so that independent systems can be interleaved:
//   seg_eliminate  - step 1 (coef(t, a, b, c) gives row t of this lane's
//                    segment) and the relation of the segment's row 0 to
//                    its two separators;
//   seg_separators - step 2: the separator equations, solved by parallel
//                    cyclic reduction across the segment lanes of each mode
//                    (log2 of the segment count steps of lane shuffles);
//   seg_backsub    - step 3.
// ns segments are active; lanes past them carry identity rows.
template <int SEG, int NC>
struct SegSys {
  double cp[SEG], gp[SEG];  // c' and the coupling to the separator below
  double aT, bT, cT;        // row of this segment's separator (its last row)
  double2 fT[NC];
  double2 Y[NC];            // row 0: x = Y + VL xl + VR xr
  double VL, VR;
};

template <int SEG, int NC, class CoefF>
__device__ __forceinline__ void seg_eliminate(double2 (&X)[NC][SEG], SegSys<SEG, NC>& S, CoefF coef, bool last) {
  S.aT = 0.0; S.bT = 1.0; S.cT = 0.0;
#pragma unroll
  for (int c = 0; c < NC; ++c) S.fT[c] = X[c][SEG - 1];
#pragma unroll
  for (int t = 0; t < SEG; ++t) {
    double a_, b_, c_;
    coef(t, a_, b_, c_);
    if (t == SEG - 1) { S.aT = a_; S.bT = b_; S.cT = c_; }
    const double bp = (t == 0) ? b_ : b_ - a_ * S.cp[t - 1];
    const double inv = 1.0 / bp;
    S.cp[t] = c_ * inv;
    S.gp[t] = (t == 0) ? (-a_) * inv : (-(a_ * S.gp[t - 1])) * inv;
#pragma unroll
    for (int c = 0; c < NC; ++c)
      X[c][t] = (t == 0) ? rmul(inv, X[c][t]) : rmul(inv, csub(X[c][t], rmul(a_, X[c][t - 1])));
  }
  // the separator row itself: x = 0 + 0 * xl - (-1) * xr
  if (!last) {
    S.cp[SEG - 1] = -1.0;
    S.gp[SEG - 1] = 0.0;
#pragma unroll
    for (int c = 0; c < NC; ++c) X[c][SEG - 1] = cz();
  }
  S.VL = S.gp[SEG - 1];
  S.VR = -S.cp[SEG - 1];
#pragma unroll
  for (int c = 0; c < NC; ++c) S.Y[c] = X[c][SEG - 1];
#pragma unroll
  for (int t = SEG - 2; t >= 0; --t) {
#pragma unroll
    for (int c = 0; c < NC; ++c) S.Y[c] = csub(X[c][t], rmul(S.cp[t], S.Y[c]));
    S.VL = S.gp[t] - S.cp[t] * S.VL;
    S.VR = -S.cp[t] * S.VR;
  }
}

template <int MODES, int SEG, int NC>
__device__ __forceinline__ void seg_separators(const double2 (&X)[NC][SEG], const SegSys<SEG, NC>& S, int lane,
                                               int ns, double2 (&xl)[NC], double2 (&xr)[NC]) {
  constexpr int SEGS = 64 / MODES;
  const int s = lane / MODES;
  const bool eq = (s < ns - 1);  // this lane's segment ends in a separator
  // separator s couples row SEG-2 of this segment and row 0 of segment s+1
  const int nb = (s + 1 < SEGS) ? lane + MODES : lane;
  const double VLn = __shfl(S.VL, nb, 64), VRn = __shfl(S.VR, nb, 64);
  double A = eq ? S.aT * S.gp[SEG - 2] : 0.0;
  double B = eq ? S.aT * (-S.cp[SEG - 2]) + S.bT + S.cT * VLn : 1.0;
  double C = eq ? S.cT * VRn : 0.0;
  double2 G[NC];
#pragma unroll
  for (int c = 0; c < NC; ++c) {
    const double2 Yn = shfl2(S.Y[c], nb);
    const double2 g = csub(csub(S.fT[c], rmul(S.aT, X[c][SEG - 2])), rmul(S.cT, Yn));
    G[c] = eq ? g : cz();
  }
  // parallel cyclic reduction; rows outside 0..ns-2 are identity rows
#pragma unroll
  for (int d = 1; d < SEGS - 1; d *= 2) {
    if (d >= ns - 1) continue;
    const bool hasl = (s >= d), hash = (s + d < SEGS);
    const int lo = hasl ? lane - d * MODES : lane, hi = hash ? lane + d * MODES : lane;
    double Al = __shfl(A, lo, 64), Bl = __shfl(B, lo, 64), Cl = __shfl(C, lo, 64);
    double Ah = __shfl(A, hi, 64), Bh = __shfl(B, hi, 64), Ch = __shfl(C, hi, 64);
    double2 Gl[NC], Gh[NC];
#pragma unroll
    for (int c = 0; c < NC; ++c) { Gl[c] = shfl2(G[c], lo); Gh[c] = shfl2(G[c], hi); }
    if (!hasl) { Al = 0.0; Bl = 1.0; Cl = 0.0; }
    if (!hash) { Ah = 0.0; Bh = 1.0; Ch = 0.0; }
    const double k1 = hasl ? A / Bl : 0.0;
    const double k2 = hash ? C / Bh : 0.0;
    const double An = -(Al * k1), Cn = -(Ch * k2);
    const double Bn = B - Cl * k1 - Ah * k2;
#pragma unroll
    for (int c = 0; c < NC; ++c) {
      const double2 gl = hasl ? Gl[c] : cz(), gh = hash ? Gh[c] : cz();
      G[c] = csub(csub(G[c], rmul(k1, gl)), rmul(k2, gh));
    }
    A = An; B = Bn; C = Cn;
  }
  const double rB = 1.0 / B;
  const int pv = (s > 0) ? lane - MODES : lane;
#pragma unroll
  for (int c = 0; c < NC; ++c) {
    const double2 x = eq ? rmul(rB, G[c]) : cz();
    const double2 xp = shfl2(x, pv);
    xr[c] = x;
    xl[c] = (s > 0 && s < ns) ? xp : cz();
  }
}

template <int SEG, int NC>
__device__ __forceinline__ void seg_backsub(double2 (&X)[NC][SEG], const SegSys<SEG, NC>& S, const double2 (&xl)[NC],
                                            const double2 (&xr)[NC]) {
#pragma unroll
  for (int c = 0; c < NC; ++c) {
    double2 nx = xr[c];
#pragma unroll
    for (int t = SEG - 1; t >= 0; --t) {
      X[c][t] = csub(cadd(X[c][t], rmul(S.gp[t], xl[c])), rmul(S.cp[t], nx));
      nx = X[c][t];
    }
  }
}

// bx: the wavefront's mode group; gx: the number of groups of the launch
// (0: no XCD pairing, the batched form)
template <int MODES, int SEG>
__device__ __forceinline__ void solve_seg_body(const DevOps& ops, const SolveArgs& a, long bx, long gx) {
  constexpr int SEGS = 64 / MODES;
  // NL and P rows of this lane's segment, kept from the right-hand-side pass
  // for the pressure update and the next RK register: [row][comp][lane]
  __shared__ double2 stash[SEG][4][64];
  const int lane = threadIdx.x;
  const int s = lane / MODES, q = lane - (lane / MODES) * MODES;
  const long nm = (long)a.nxh * a.nz;
  const int ny = a.ny, nyp = ny + 1;
  const int ns = (nyp + SEG - 1) / SEG;   // active segments
  const int rl = nyp - (ns - 1) * SEG;    // rows of the last one (3..SEG)
  // The G = 8 / MODES mode groups whose rows share 128-byte lines (MODES < 8)
  // go to workgroups b, b+8, .., b+8(G-1), which land on the same XCD one
  // after the other.
  long grp = bx;
  if (MODES < 8 && gx > 0) {
    constexpr long G = 8 / MODES;
    const long full = (gx / (8 * G)) * (8 * G);
    if (grp < full) {
      const long xcd = grp & 7, k = grp >> 3;
      grp = 8 * G * (k / G) + G * xcd + (k % G);
    }
  }
  const long mg = grp * MODES + q;
  const bool valid = (mg < nm) && (s < ns);
  const long m = (mg < nm) ? mg : nm - 1;
  const int ix = (int)(m / a.nz), iz = (int)(m - (long)ix * a.nz);
  const int j0 = s * SEG;
  const bool last = (s == ns - 1);
  const double kxv = ops.kx[ix], kzv = ops.kz[iz];
  const double k2 = kxv * kxv + kzv * kzv;
  const bool mean = (k2 == 0.0);
  const int r = a.rk - 1;
  const double dt = a.sc->dt;
  const double re = a.re;
  const double* d2h = ops.d2h;
  const double* d2n = ops.d2n;
  const double* dn = ops.dn;
  const double* dh = ops.dh;
  const double* dm = ops.dm;
  const double2* __restrict__ NL = a.nl;
  double2* __restrict__ R = a.R;
  double2* __restrict__ P = a.P;
  double2* __restrict__ U = a.vel;
  auto neg = [](double2 v) { return make_double2(-v.x, -v.y); };
  auto sel = [](bool c, double2 x, double2 y) { return make_double2(c ? x.x : y.x, c ? x.y : y.y); };

  const uint32_t rowB = (uint32_t)(nm * 16), cmpB = rowB * (uint32_t)nyp;
  const uint32_t om0 = (uint32_t)(m * 16);
  auto OM = [&](int j) { return om0 + (uint32_t)j * rowB; };
  auto OC = [&](int c, int j) { return (uint32_t)c * cmpB + om0 + (uint32_t)j * rowB; };

  const double vis_imp = c_rkd[r] * (dt / re);
  const double rk_press = (c_rkc[r] + c_rkd[r]) * dt;
  const double coef = c_rkb[r] * dt;
  const double vis_exp1 = c_rkc[r] * (dt / re);
  const double alpha = 1.0 + vis_imp * k2;

  // mean-mode pressure gradient (compute_p_grad, :369-373); the mean mode is mode 0
  double p_grad = a.sc->p_grad;
  if (grp == 0 && a.fix_mass == 1) {
    double part = 0.0;
#pragma unroll
    for (int t = 0; t < SEG; ++t) {
      const int j = j0 + t;
      const int jc = min(j, ny);
      const double term = ops.dyh[jc] * (-ldo(NL, OC(0, jc)).x);
      part = (j >= 1 && j <= ny - 1) ? part + term : part;
    }
    double tmp = 0.0;
#pragma unroll
    for (int k = 0; k < SEGS; ++k) tmp = (k < ns) ? tmp + __shfl(part, k * MODES + q, 64) : tmp;
    if (mean) {
      p_grad = -tmp + 0.5 * (a.dudy[ny - 1] - a.dudy[0]) / re;
      if (lane == 0) a.sc->p_grad = p_grad;
    }
  }
  const double2 msh0 = make_double2(a.cosb * rk_press * p_grad, 0.0);
  const double2 msh2 = make_double2(a.sinb * rk_press * p_grad, 0.0);

  const double2 top0 = a.cbc[(0L * 3 + 0) * nm + m], top1 = a.cbc[(0L * 3 + 1) * nm + m],
                top2 = a.cbc[(0L * 3 + 2) * nm + m];
  const double2 bot0 = a.cbc[(1L * 3 + 0) * nm + m], bot1 = a.cbc[(1L * 3 + 1) * nm + m],
                bot2 = a.cbc[(1L * 3 + 2) * nm + m];

  // ---------------- right-hand sides (rows j0 .. j0+SEG-1) ----------------
  double2 UW[2][SEG], V[1][SEG];
  auto rhs = [&](auto first_c) {
    constexpr bool FIRST = decltype(first_c)::value;
    const int jb = min(j0, ny);  // rows clamped for idle lanes
    double2 Rm0 = cz(), Rm1 = cz(), Rm2 = cz();
    double2 Pc = ldo(P, OM(jb));
    double2 Rc0 = ldo(R, OC(0, jb)), Rc1 = ldo(R, OC(1, jb)), Rc2 = ldo(R, OC(2, jb));
    // The first substep's stencils reach one row into each neighbouring
    // segment: R at j0-1 (the explicit diffusion of row j0) and R, P at
    // j0+SEG (those of the last row).  Every lane reads only its own rows from
    // memory, the rows it overwrites later in this kernel; the neighbours'
    // rows come from the lanes that own them, through register shuffles.
    // Wall rows take the wall values, so what the segments at the ends
    // receive from beyond the walls (their own rows) is never used.
    double2 Rx0 = cz(), Rx1 = cz(), Rx2 = cz(), Px = cz();
    if (FIRST) {
      const int jl = min(j0 + SEG - 1, ny);
      const double2 L0 = ldo(R, OC(0, jl)), L1 = ldo(R, OC(1, jl)), L2 = ldo(R, OC(2, jl));
      const int prv = (s == 0) ? lane : lane - MODES;
      const int nxt = (s + 1 >= SEGS) ? lane : lane + MODES;
      Rm0 = sel(s == 0, Rc0, shfl2(L0, prv));
      Rm1 = sel(s == 0, Rc1, shfl2(L1, prv));
      Rm2 = sel(s == 0, Rc2, shfl2(L2, prv));
      Rx0 = shfl2(Rc0, nxt);
      Rx1 = shfl2(Rc1, nxt);
      Rx2 = shfl2(Rc2, nxt);
      Px = shfl2(Pc, nxt);
    }
#pragma unroll
    for (int t = 0; t < SEG; ++t) {
      const int j = j0 + t;
      const int jc = min(j, ny), jn = min(j + 1, ny), jv = min(j, ny - 1);
      const double2 n0 = ldo(NL, OC(0, jc)), n1 = ldo(NL, OC(1, jc)), n2 = ldo(NL, OC(2, jc));
      stash[t][0][lane] = n0;
      stash[t][1][lane] = n1;
      stash[t][2][lane] = n2;
      stash[t][3][lane] = Pc;
      double2 r0 = cz(), r1 = cz(), r2 = cz(), pn = cz();
      if (t + 1 < SEG) {
        r0 = ldo(R, OC(0, jn)); r1 = ldo(R, OC(1, jn)); r2 = ldo(R, OC(2, jn));
        pn = ldo(P, OM(jn));
      } else if (FIRST) {
        r0 = Rx0; r1 = Rx1; r2 = Rx2;
        pn = Px;
      }
      double2 v0 = cadd(Rc0, rmul(coef, neg(n0)));
      double2 v1 = cadd(Rc1, rmul(coef, neg(n1)));
      double2 v2 = cadd(Rc2, rmul(coef, neg(n2)));
      if (FIRST) {
        // add_explicit_diffusive_terms(cv, cru) (:339-351)
        const double a0 = d2h[3 * jc], a1 = d2h[3 * jc + 1], a2 = d2h[3 * jc + 2];
        const double b0 = d2n[3 * jv], b1 = d2n[3 * jv + 1], b2 = d2n[3 * jv + 2];
        const double dnj = dn[jv];
        const double2 du = cadd(cadd(rmul(a0, Rm0), rmul(a1, Rc0)), rmul(a2, r0));
        v0 = cadd(v0, rmul(vis_exp1, csub(du, rmul(k2, Rc0))));
        const double2 dw = cadd(cadd(rmul(a0, Rm2), rmul(a1, Rc2)), rmul(a2, r2));
        v2 = cadd(v2, rmul(vis_exp1, csub(dw, rmul(k2, Rc2))));
        const double2 dv = cadd(cadd(rmul(b0, Rm1), rmul(b1, Rc1)), rmul(b2, r1));
        v1 = sel(j <= ny - 2, cadd(v1, rmul(vis_exp1, csub(dv, rmul(k2, Rc1)))), v1);
        // subtract_pressure_gradient(cv, cp, rk_press) (:54-77)
        v0 = sel(mean, v0, csub(v0, imul(rk_press * kxv, Pc)));
        v2 = sel(mean, v2, csub(v2, imul(rk_press * kzv, Pc)));
        v1 = sel(!mean && j <= ny - 2, csub(v1, rmul(rk_press * dnj, csub(pn, Pc))), v1);
      }
      // mean pressure gradient (:220-221)
      v0 = sel(mean, csub(v0, msh0), v0);
      v2 = sel(mean, csub(v2, msh2), v2);
      // wall rows carry the wall values (SOLVE_HELMHOLTZ, :174-213); v has no row ny
      v0 = sel(j == 0, bot0, sel(j == ny, top0, sel(j > ny, cz(), v0)));
      v2 = sel(j == 0, bot2, sel(j == ny, top2, sel(j > ny, cz(), v2)));
      v1 = sel(j == 0, bot1, sel(j == ny - 1, top1, sel(j >= ny, cz(), v1)));
      UW[0][t] = v0;
      UW[1][t] = v2;
      V[0][t] = v1;
      Rm0 = Rc0; Rm1 = Rc1; Rm2 = Rc2;
      Rc0 = r0; Rc1 = r1; Rc2 = r2;
      Pc = pn;
    }
  };
  if (r == 0) rhs(std::true_type{});
  else rhs(std::false_type{});

  // ---------------- Helmholtz solves (u and w with D2H, v with D2N) ----------------
  {
    SegSys<SEG, 2> SH;
    SegSys<SEG, 1> SN;
    seg_eliminate(UW, SH, [&](int t, double& a_, double& b_, double& c_) {
      const int j = j0 + t, jc = min(j, ny);
      const double h0 = d2h[3 * jc], h1 = d2h[3 * jc + 1], h2 = d2h[3 * jc + 2];
      const bool wall = (j == 0 || j >= ny);
      a_ = wall ? 0.0 : -vis_imp * h0;
      b_ = wall ? 1.0 : -vis_imp * h1 + alpha;
      c_ = wall ? 0.0 : -vis_imp * h2;
    }, s >= ns - 1);
    seg_eliminate(V, SN, [&](int t, double& a_, double& b_, double& c_) {
      const int j = j0 + t, jv = min(j, ny - 1);
      const double h0 = d2n[3 * jv], h1 = d2n[3 * jv + 1], h2 = d2n[3 * jv + 2];
      const bool wall = (j == 0 || j >= ny - 1);
      a_ = wall ? 0.0 : -vis_imp * h0;
      b_ = wall ? 1.0 : -vis_imp * h1 + alpha;
      c_ = wall ? 0.0 : -vis_imp * h2;
    }, s >= ns - 1);
    double2 xlH[2], xrH[2], xlN[1], xrN[1];
    seg_separators<MODES, SEG, 2>(UW, SH, lane, ns, xlH, xrH);
    seg_separators<MODES, SEG, 1>(V, SN, lane, ns, xlN, xrN);
    seg_backsub(UW, SH, xlH, xrH);
    seg_backsub(V, SN, xlN, xrN);
  }
#pragma unroll
  for (int t = 0; t < SEG; ++t)  // mean mode (:261-263)
    V[0][t] = sel(mean, (j0 + t < ny) ? bot1 : cz(), V[0][t]);

  // ---------------- Neumann Poisson for the pressure correction ----------------
  double2 PH[1][SEG];
  {
    const int pv = (s == 0) ? lane : lane - MODES;
    const double2 vprev = shfl2(V[0][SEG - 1], pv);
    const double tmpc = 1.0 / rk_press;
#pragma unroll
    for (int t = 0; t < SEG; ++t) {
      const int j = j0 + t, jc = min(j, ny);
      const double2 vm = (t == 0) ? vprev : V[0][t - 1];
      // strip_divergence rhs (:28-31)
      const double2 dv = cadd(cadd(imul(kxv, UW[0][t]), rmul(dh[jc], csub(V[0][t], vm))), imul(kzv, UW[1][t]));
      PH[0][t] = sel(j >= 1 && j <= ny - 1 && !mean, rmul(tmpc, dv), cz());
    }
  }
  const double alp = mean ? -1.0 : -k2;  // the mean mode has no pressure correction
  {
    const double dn0 = dn[0], dn1 = dn[1], dh1 = dh[1];
    const double dnt = dn[ny - 1], dnt2 = dn[ny - 2], dht = dh[ny - 1];
    SegSys<SEG, 1> SP;
    seg_eliminate(PH, SP, [&](int t, double& a_, double& b_, double& c_) {
      const int j = j0 + t, jc = min(j, ny);
      const double h0 = d2h[3 * jc], h1 = d2h[3 * jc + 1], h2 = d2h[3 * jc + 2];
      a_ = h0; b_ = h1 + alp; c_ = h2;
      a_ = (j == 1) ? dh1 * dn0 : a_;
      b_ = (j == 1) ? -dh1 * dn0 - dh1 * dn1 + alp : b_;
      c_ = (j == 1) ? dh1 * dn1 : c_;
      a_ = (j == ny - 1) ? dht * dnt2 : a_;
      b_ = (j == ny - 1) ? -dht * dnt2 - dht * dnt + alp : b_;
      c_ = (j == ny - 1) ? dht * dnt : c_;
      a_ = (j == 0) ? 0.0 : (j == ny) ? -dnt : (j > ny) ? 0.0 : a_;
      b_ = (j == 0) ? -dn0 : (j == ny) ? dnt : (j > ny) ? 1.0 : b_;
      c_ = (j == 0) ? dn0 : (j >= ny) ? 0.0 : c_;
    }, s >= ns - 1);
    double2 xl[1], xr[1];
    seg_separators<MODES, SEG, 1>(PH, SP, lane, ns, xl, xr);
    seg_backsub(PH, SP, xl, xr);
  }
  // wall re-extrapolation (SOLVE_P :335-336, zero wall data)
  {
    const double2 pb = make_double2((0.0 - dm[1] * PH[0][1].x - dm[2] * PH[0][2].x) / dm[0],
                                    (0.0 - dm[1] * PH[0][1].y - dm[2] * PH[0][2].y) / dm[0]);
    PH[0][0] = sel(s == 0, pb, PH[0][0]);
    double2 q1 = cz(), q2 = cz();
#pragma unroll
    for (int t = 0; t < SEG; ++t) {
      q1 = sel(t == rl - 3, PH[0][t], q1);
      q2 = sel(t == rl - 2, PH[0][t], q2);
    }
    const double2 pt = make_double2((0.0 - dm[3] * q1.x - dm[4] * q2.x) / dm[5],
                                    (0.0 - dm[3] * q1.y - dm[4] * q2.y) / dm[5]);
#pragma unroll
    for (int t = 0; t < SEG; ++t) PH[0][t] = sel(mean, cz(), sel(last && t == rl - 1, pt, PH[0][t]));
  }

  // ---------------- projection and pressure update ----------------
  const int nxt = (last || s + 1 >= SEGS) ? lane : lane + MODES, prv = (s == 0) ? lane : lane - MODES;
  const double alpha2 = 1.0 + vis_imp * k2;
  double2 PN[SEG];
  {
    const double2 phN = shfl2(PH[0][0], nxt);
    const double2 phP = shfl2(PH[0][SEG - 1], prv);
#pragma unroll
    for (int t = 0; t < SEG; ++t) {
      const int j = j0 + t, jc = min(j, ny), jv = min(j, ny - 1);
      const double2 f0 = PH[0][t];
      const double2 fn = (t + 1 < SEG) ? PH[0][t + 1] : phN;
      double2 s0 = (t > 0) ? PH[0][t - 1] : phP, s1 = f0, s2 = fn;
      if (t == 0) { s0 = sel(j == 0, PH[0][0], s0); s1 = sel(j == 0, PH[0][1], s1); s2 = sel(j == 0, PH[0][2], s2); }
      if (t >= 2) { s0 = sel(j == ny, PH[0][t - 2], s0); s1 = sel(j == ny, PH[0][t - 1], s1); s2 = sel(j == ny, f0, s2); }
      const double2 Pold = stash[t][3][lane];
      const double h0 = d2h[3 * jc], h1 = d2h[3 * jc + 1], h2 = d2h[3 * jc + 2];
      const double dnj = dn[jv];
      const bool inner = !mean && j >= 1 && j <= ny - 1;
      const double2 cu = sel(inner, csub(UW[0][t], imul(rk_press * kxv, f0)), UW[0][t]);
      const double2 cw = sel(inner, csub(UW[1][t], imul(rk_press * kzv, f0)), UW[1][t]);
      const double2 cv = sel(inner && j <= ny - 2, csub(V[0][t], rmul(rk_press * dnj, csub(fn, f0))), V[0][t]);
      const double2 d2 = cadd(cadd(rmul(h0, s0), rmul(h1, s1)), rmul(h2, s2));
      const double2 pn = sel(mean, cz(), csub(cadd(Pold, rmul(alpha2, f0)), rmul(vis_imp, d2)));
      UW[0][t] = cu;
      V[0][t] = cv;
      UW[1][t] = cw;
      PN[t] = pn;
    }
  }
  if (valid) {
#pragma unroll
    for (int t = 0; t < SEG; ++t) {
      const int j = j0 + t;
      if (j <= ny) {
        sto(U, OC(0, j), UW[0][t]);
        sto(U, OC(1, j), V[0][t]);
        sto(U, OC(2, j), UW[1][t]);
        sto(P, OM(j), PN[t]);
      }
    }
  }

  // ---------------- d_um_dy and the next RK register ----------------
  const double2 uN = shfl2(UW[0][0], nxt), vN = shfl2(V[0][0], nxt), wN = shfl2(UW[1][0], nxt);
  const double2 uP = shfl2(UW[0][SEG - 1], prv), vP = shfl2(V[0][SEG - 1], prv), wP = shfl2(UW[1][SEG - 1], prv);
  const double2 pnN = shfl2(PN[0], nxt);
  if (mean && valid) {  // d_um_dy update from the new mean profile (:84-89)
#pragma unroll
    for (int t = 0; t < SEG; ++t) {
      const int j = j0 + t;
      const double un = (t + 1 < SEG) ? UW[0][t + 1].x : uN.x;
      if (j >= 1 && j <= ny - 2) a.dudy[j] = (dn[j] * (un - UW[0][t].x));
      if (t == 0 && j == 0) a.dudy[0] = dm[0] * UW[0][0].x + dm[1] * UW[0][1].x + dm[2] * UW[0][2].x;
      if (t >= 2 && j == ny)
        a.dudy[ny - 1] = dm[3] * UW[0][t - 2].x + dm[4] * UW[0][t - 1].x + dm[5] * UW[0][t].x;
    }
  }
  if (!valid) return;
  if (a.rk == 3) {
#pragma unroll
    for (int t = 0; t < SEG; ++t) {
      const int j = j0 + t;
      if (j <= ny) {
        sto(R, OC(0, j), UW[0][t]);
        sto(R, OC(1, j), V[0][t]);
        sto(R, OC(2, j), UW[1][t]);
      }
    }
    return;
  }
  const double coef2 = dt * c_rka[r + 1];
  const double vis_exp2 = (dt / re) * c_rkc[r + 1];
  const double rk_press2 = (c_rkc[r + 1] + c_rkd[r + 1]) * dt;
#pragma unroll
  for (int t = 0; t < SEG; ++t) {
    const int j = j0 + t, jc = min(j, ny), jv = min(j, ny - 1);
    const double2 n0 = stash[t][0][lane], n1 = stash[t][1][lane], n2 = stash[t][2][lane];
    const double a0 = d2h[3 * jc], a1 = d2h[3 * jc + 1], a2 = d2h[3 * jc + 2];
    const double b0 = d2n[3 * jv], b1 = d2n[3 * jv + 1], b2 = d2n[3 * jv + 2];
    const double dnj = dn[jv];
    // D2H / D2N stencil rows j-1, j, j+1; rows 0..2 at the bottom, the last three at the top
    double2 u0 = (t > 0) ? UW[0][t - 1] : uP, u1 = UW[0][t], u2 = (t + 1 < SEG) ? UW[0][t + 1] : uN;
    double2 w0 = (t > 0) ? UW[1][t - 1] : wP, w1 = UW[1][t], w2 = (t + 1 < SEG) ? UW[1][t + 1] : wN;
    double2 x0 = (t > 0) ? V[0][t - 1] : vP, x1 = V[0][t], x2 = (t + 1 < SEG) ? V[0][t + 1] : vN;
    if (t == 0) {
      const bool b = (j == 0);
      u0 = sel(b, UW[0][0], u0); u1 = sel(b, UW[0][1], u1); u2 = sel(b, UW[0][2], u2);
      w0 = sel(b, UW[1][0], w0); w1 = sel(b, UW[1][1], w1); w2 = sel(b, UW[1][2], w2);
      x0 = sel(b, V[0][0], x0); x1 = sel(b, V[0][1], x1); x2 = sel(b, V[0][2], x2);
    }
    if (t >= 2) {
      const bool b = (j == ny);
      u0 = sel(b, UW[0][t - 2], u0); u1 = sel(b, UW[0][t - 1], u1); u2 = sel(b, UW[0][t], u2);
      w0 = sel(b, UW[1][t - 2], w0); w1 = sel(b, UW[1][t - 1], w1); w2 = sel(b, UW[1][t], w2);
    }
    if (t >= 1) {  // last v row: rows ny-3 .. ny-1
      const bool b = (j == ny - 1);
      x0 = sel(b, (t >= 2) ? V[0][t - 2] : vP, x0);
      x1 = sel(b, V[0][t - 1], x1);
      x2 = sel(b, V[0][t], x2);
    }
    const double2 cu = UW[0][t], cv = V[0][t], cw = UW[1][t];
    double2 o0 = cadd(cu, rmul(coef2, neg(n0)));
    double2 o1 = cadd(cv, rmul(coef2, neg(n1)));
    double2 o2 = cadd(cw, rmul(coef2, neg(n2)));
    const double2 du = cadd(cadd(rmul(a0, u0), rmul(a1, u1)), rmul(a2, u2));
    const double2 dw = cadd(cadd(rmul(a0, w0), rmul(a1, w1)), rmul(a2, w2));
    o0 = cadd(o0, rmul(vis_exp2, csub(du, rmul(k2, cu))));
    o2 = cadd(o2, rmul(vis_exp2, csub(dw, rmul(k2, cw))));
    const double2 dv = cadd(cadd(rmul(b0, x0), rmul(b1, x1)), rmul(b2, x2));
    o1 = sel(j < ny, cadd(o1, rmul(vis_exp2, csub(dv, rmul(k2, cv)))), o1);
    const double2 Pj = PN[t], Pj1 = (t + 1 < SEG) ? PN[t + 1] : pnN;
    const bool inner = !mean && j >= 1 && j <= ny - 1;
    o0 = sel(inner, csub(o0, imul(rk_press2 * kxv, Pj)), o0);
    o2 = sel(inner, csub(o2, imul(rk_press2 * kzv, Pj)), o2);
    o1 = sel(inner && j <= ny - 2, csub(o1, rmul(rk_press2 * dnj, csub(Pj1, Pj))), o1);
    if (j <= ny) {
      sto(R, OC(0, j), o0);
      sto(R, OC(1, j), o1);
      sto(R, OC(2, j), o2);
    }
  }
}

template <int MODES, int SEG>
__global__ __launch_bounds__(64) void k_solve_seg(DevOps ops, SolveArgs a) {
  solve_seg_body<MODES, SEG>(ops, a, blockIdx.x, gridDim.x);
}

// Batched form: workgroup i runs mode group it[i].bx of channel it[i].m.
template <int MODES, int SEG>
__global__ __launch_bounds__(64) void k_solve_seg_b(const SBatchArgs* __restrict__ args,
                                                    const BatchItem* __restrict__ it) {
  const BatchItem w = it[blockIdx.x];
  const SBatchArgs& x = args[w.m];
  solve_seg_body<MODES, SEG>(x.ops, x.a, w.bx, 0);
}

// Rows per segment of the segmented solve for ny+1 rows: the shortest of
// 5..9 with at most maxsegs segments (64 / modes per wavefront) and 3 or more
// rows in the last one (the wall rows 0..2 and ny-2..ny each lie in one
// segment); 0 if none fits.
int solve_seg_rows(int ny, int maxsegs) {
  const int nyp = ny + 1;
  if (ny < 8) return 0;
  for (int seg = 5; seg <= 9; ++seg) {
    const int ns = (nyp + seg - 1) / seg, rl = nyp - (ns - 1) * seg;
    if (ns <= maxsegs && rl >= 3) return seg;
  }
  return 0;
}

// ---- direct global-to-LDS row ring (the DMA variant of k_solve3) ----
//
// Each pass streams its rows through a ring of R slots in the wavefront's
// LDS: global_load_lds writes lane i's 16 (or 4) bytes at slot base + 16i
// (4i), so every lane reads back exactly what it fetched.  A row is consumed
// R rows after its loads were issued; vector memory operations of a
// wavefront complete in issue order, so once at most (R-1)*nld are still
// outstanding (nld = loads per row) the oldest row has landed.  Stores issued
// in between only make that wait stricter.  Rows past the end are replaced by
// loads of each array's first element (one cache line) to keep the count
// uniform.
template <int N>
__device__ __forceinline__ void wait_vm() {
  static_assert(N >= 0 && N < 64, "vmcnt has 6 bits");
  __builtin_amdgcn_s_waitcnt((N & 15) | ((N >> 4) << 14) | 0x0F70);
}
template <int R>
__device__ __forceinline__ void wait_ring(int nld) {
  constexpr int M1 = R - 1;
  switch (nld) {  // a count above 63 is clamped, which only waits for more
    case 2: wait_vm<(2 * M1 < 63 ? 2 * M1 : 63)>(); break;
    case 3: wait_vm<(3 * M1 < 63 ? 3 * M1 : 63)>(); break;
    case 4: wait_vm<(4 * M1 < 63 ? 4 * M1 : 63)>(); break;
    case 5: wait_vm<(5 * M1 < 63 ? 5 * M1 : 63)>(); break;
    case 6: wait_vm<(6 * M1 < 63 ? 6 * M1 : 63)>(); break;
    default: wait_vm<0>(); break;
  }
}
__device__ __forceinline__ void dma16(const void* g, char* lds_uniform) {
  __builtin_amdgcn_global_load_lds(g, (__attribute__((address_space(3))) void*)lds_uniform, 16, 0, 0);
}
__device__ __forceinline__ void dma4(const void* g, char* lds_uniform) {
  __builtin_amdgcn_global_load_lds(g, (__attribute__((address_space(3))) void*)lds_uniform, 4, 0, 0);
}
// n rows; row(t) is the t-th row; fetch(row or -1, slot) issues nld loads;
// read(slot) returns the row data of this lane; comp(row, data) consumes it
template <int R, class RowF, class FetchF, class ReadF, class CompF>
__device__ __forceinline__ void ring_rows(int n, int nld, RowF row, FetchF fetch, ReadF read, CompF comp) {
#pragma unroll
  for (int t = 0; t < R; ++t) fetch(t < n ? row(t) : -1, t);
  int slot = 0;
  for (int t = 0; t < n; ++t) {
    wait_ring<R>(nld);
    const auto d = read(slot);
    comp(row(t), d);
    __builtin_amdgcn_s_waitcnt(0xC07F);  // lgkmcnt(0): this slot has been read
    fetch(t + R < n ? row(t + R) : -1, slot);
    slot = (slot + 1 == R) ? 0 : slot + 1;
  }
  wait_vm<0>();  // nothing lands in the ring after the pass
}
// Ring reads.  The compiler cannot tell ring slots apart, so a plain LDS
// load would make it wait for every DMA in flight (vmcnt(0)) and serialise
// the ring; these reads are issued from inline assembly together with their
// own lgkmcnt(0), after the explicit vmcnt wait above has ordered them.
typedef double v2d __attribute__((ext_vector_type(2)));
__device__ __forceinline__ uint32_t lds_addr(const void* p) {
  return (uint32_t)(uintptr_t)(const __attribute__((address_space(3))) void*)p;
}
__device__ __forceinline__ double2 d2(v2d v) { return make_double2(v.x, v.y); }
// three 16-byte cells at a16 + 0, 1024, 2048
__device__ __forceinline__ void ring_rd_a(uint32_t a16, v2d& x0, v2d& x1, v2d& x2) {
  asm volatile(
      "ds_read_b128 %0, %3\n\t"
      "ds_read_b128 %1, %3 offset:1024\n\t"
      "ds_read_b128 %2, %3 offset:2048\n\t"
      "s_waitcnt lgkmcnt(0)"
      : "=&v"(x0), "=&v"(x1), "=&v"(x2)
      : "v"(a16)
      : "memory");
}
// one 16-byte cell at a16 and two 4-byte cells at a4 + 1024, 1280
__device__ __forceinline__ void ring_rd_b(uint32_t a16, uint32_t a4, v2d& x0, int& lo, int& hi) {
  asm volatile(
      "ds_read_b128 %0, %3\n\t"
      "ds_read_b32 %1, %4 offset:1024\n\t"
      "ds_read_b32 %2, %4 offset:1280\n\t"
      "s_waitcnt lgkmcnt(0)"
      : "=&v"(x0), "=&v"(lo), "=&v"(hi)
      : "v"(a16), "v"(a4)
      : "memory");
}
// four 16-byte cells at a16 + 0 .. 3072 and two 4-byte cells at a4 + 4096, 4352
__device__ __forceinline__ void ring_rd_c(uint32_t a16, uint32_t a4, v2d& x0, v2d& x1, v2d& x2, v2d& x3,
                                          int& lo, int& hi) {
  asm volatile(
      "ds_read_b128 %0, %6\n\t"
      "ds_read_b128 %1, %6 offset:1024\n\t"
      "ds_read_b128 %2, %6 offset:2048\n\t"
      "ds_read_b128 %3, %6 offset:3072\n\t"
      "ds_read_b32 %4, %7 offset:4096\n\t"
      "ds_read_b32 %5, %7 offset:4352\n\t"
      "s_waitcnt lgkmcnt(0)"
      : "=&v"(x0), "=&v"(x1), "=&v"(x2), "=&v"(x3), "=&v"(lo), "=&v"(hi)
      : "v"(a16), "v"(a4)
      : "memory");
}

// ring depths and LDS bytes of the DMA variant (one wavefront per workgroup)
constexpr int SOLVE_RA = 12, SOLVE_RB = 21, SOLVE_RC = 8;
constexpr int SOLVE_SA = 3072, SOLVE_SB = 1536, SOLVE_SC = 4608;   // slot bytes
constexpr int SOLVE_RING_BYTES = 36864;
static_assert(SOLVE_RA * SOLVE_SA <= SOLVE_RING_BYTES && SOLVE_RB * SOLVE_SB <= SOLVE_RING_BYTES &&
              SOLVE_RC * SOLVE_SC <= SOLVE_RING_BYTES, "solve ring exceeds its LDS");

template <int BL, bool DMA>
__global__ __launch_bounds__(64) void k_solve3(DevOps ops, SolveArgs a) {
  extern __shared__ __align__(16) char solve_ring[];
  const uint32_t ring0 = lds_addr(solve_ring);
  const int lane = threadIdx.x;
  const int sub = lane & 15, g5 = sub / 3, c = sub - 3 * g5;
  const long nm = (long)a.nxh * a.nz;
  const long mg = (long)blockIdx.x * 20 + (lane >> 4) * 5 + g5;
  const bool valid = (sub < 15) && (mg < nm);
  const long m = valid ? mg : 0;
  const int ix = (int)(m / a.nz), iz = (int)(m - (long)ix * a.nz);
  const int ny = a.ny, nyp = ny + 1;
  const double kxv = ops.kx[ix], kzv = ops.kz[iz];
  const double k2 = kxv * kxv + kzv * kzv;
  const bool mean = valid && (k2 == 0.0);
  const bool cv1 = (c == 1);
  const bool own = valid && (c == 0);   // stores the per-mode results
  const int gbase = lane - c;           // lane of component 0 of this mode
  const int r = a.rk - 1;
  const bool first = (r == 0);
  const double dt = a.sc->dt;
  const double re = a.re;

  const uint32_t rowB = (uint32_t)(nm * 16), cmpB = rowB * (uint32_t)nyp;
  const uint32_t om0 = (uint32_t)(m * 16);             // mode offset (per-mode arrays)
  const uint32_t oc0 = (uint32_t)c * cmpB + om0;      // component offset
  auto OC = [&](int j) { return oc0 + (uint32_t)j * rowB; };
  auto OM = [&](int j) { return om0 + (uint32_t)j * rowB; };

  const double2* __restrict__ NL = a.nl;
  double2* __restrict__ R = a.R;
  double2* __restrict__ P = a.P;
  double2* __restrict__ U = a.vel;
  double2* __restrict__ SX = a.scr;
  double2* __restrict__ SP = a.scr + 3L * nyp * nm;
  double* __restrict__ SC = reinterpret_cast<double*>(a.scr + 4L * nyp * nm);
  double* __restrict__ SA = SC + 3L * nyp * nm;
  const double* d2h = ops.d2h;
  const double* d2n = ops.d2n;
  const double* dm = ops.dm;
  auto neg = [](double2 v) { return make_double2(-v.x, -v.y); };

  const double vis_imp = c_rkd[r] * (dt / re);
  const double rk_press = (c_rkc[r] + c_rkd[r]) * dt;
  const double coef = c_rkb[r] * dt;
  const double vis_exp1 = c_rkc[r] * (dt / re);
  const double alpha = 1.0 + vis_imp * k2;
  const double kc = (c == 0) ? kxv : kzv;        // wavenumber of the u or w pressure gradient
  const int jt = cv1 ? ny - 1 : ny;              // Dirichlet top row of this component

  const double2 topc = a.cbc[(0L * 3 + c) * nm + m];
  const double2 botc = a.cbc[(1L * 3 + c) * nm + m];

  // mean-mode pressure gradient (compute_p_grad, :369-373)
  double p_grad = a.sc->p_grad;
  if (mean && a.fix_mass == 1) {
    double tmp = 0.0;
#pragma unroll 8
    for (int j = 1; j < ny; ++j) tmp = tmp + ldro(ops.dyh, j) * (-ldo(NL, OM(j)).x);
    p_grad = -tmp + 0.5 * (a.dudy[ny - 1] - a.dudy[0]) / re;
    if (own) a.sc->p_grad = p_grad;
  }
  const double mean_shift = (c == 0) ? a.cosb * rk_press * p_grad : (c == 2 ? a.sinb * rk_press * p_grad : 0.0);

  // ---------------- pass A: RHS and forward elimination (own component) ----------------
  double cp = 0.0;
  double2 f = botc;
  if (valid) { sto(SX, OC(0), f); sto1(SC, OC(0) >> 1, 0.0); }
  double2 Rm = ldo(R, OC(0)), Rc = ldo(R, OC(1));
  double2 Pc = first ? ldo(P, OM(1)) : cz();
  auto loadA = [&](int j) {
    RowA3 d;
    d.n = ldo(NL, OC(j));
    d.r = ldo(R, OC(j + 1));
    d.p = first ? ldo(P, OM(j + 1)) : cz();
    return d;
  };
  auto compA = [&](int j, const RowA3& d) {
    const double h0 = ldro(d2h, 3 * j), h1 = ldro(d2h, 3 * j + 1), h2 = ldro(d2h, 3 * j + 2);
    const double n0 = ldro(d2n, 3 * j), n1 = ldro(d2n, 3 * j + 1), n2 = ldro(d2n, 3 * j + 2);
    const double a0 = cv1 ? n0 : h0, a1 = cv1 ? n1 : h1, a2 = cv1 ? n2 : h2;
    double2 v = cadd(Rc, rmul(coef, neg(d.n)));
    if (first) {
      // add_explicit_diffusive_terms(cv, cru) (:339-351)
      double2 dd = cadd(cadd(rmul(a0, Rm), rmul(a1, Rc)), rmul(a2, d.r));
      double2 ve = cadd(v, rmul(vis_exp1, csub(dd, rmul(k2, Rc))));
      if (j < jt) v = ve;
      if (!mean) {  // subtract_pressure_gradient(cv, cp, rk_press) (:54-77)
        if (cv1) {
          if (j <= ny - 2) v = csub(v, rmul(rk_press * ldro(ops.dn, j), csub(d.p, Pc)));
        } else {
          v = csub(v, imul(rk_press * kc, Pc));
        }
      }
    }
    if (mean && !cv1) v = csub(v, make_double2(mean_shift, 0.0));  // (:220-221)
    if (j < jt) {
      const double ah = -vis_imp * a0, bh = -vis_imp * a1 + alpha, ch = -vis_imp * a2;
      const double bp = bh - ah * cp;
      const double2 gg = csub(v, rmul(ah, f));
      const double rb = 1.0 / bp;
      cp = ch * rb;
      f = rmul(rb, gg);
    } else {
      cp = 0.0;
      f = topc;
    }
    if (valid) { sto(SX, OC(j), f); sto1(SC, OC(j) >> 1, cp); }
    Rm = Rc; Rc = d.r; Pc = d.p;
  };
  if constexpr (DMA) {
    auto fetchA = [&](int j, int slot) {
      char* sb = solve_ring + slot * SOLVE_SA;
      const uint32_t on = j < 0 ? 0u : OC(j), orr = j < 0 ? 0u : OC(j + 1), op = j < 0 ? 0u : OM(j + 1);
      dma16(reinterpret_cast<const char*>(NL) + on, sb);
      dma16(reinterpret_cast<const char*>(R) + orr, sb + 1024);
      if (first) dma16(reinterpret_cast<const char*>(P) + op, sb + 2048);
    };
    auto readA = [&](int slot) {
      v2d x0, x1, x2;
      ring_rd_a(ring0 + slot * SOLVE_SA + 16 * lane, x0, x1, x2);
      RowA3 d;
      d.n = d2(x0);
      d.r = d2(x1);
      d.p = first ? d2(x2) : cz();
      return d;
    };
    ring_rows<SOLVE_RA>(ny - 1, first ? 3 : 2, [](int t) { return 1 + t; }, fetchA, readA, compA);
  } else {
    rows_up<BL>(1, ny - 1, loadA, compA);
  }
  // row ny: Dirichlet row of u and w; v has no row ny (its value there is zero)
  if (valid) { sto(SX, OC(ny), cv1 ? cz() : topc); sto1(SC, OC(ny) >> 1, 0.0); }

  // ---------------- pass B: back substitution + Poisson elimination ----------------
  const double tmpc = 1.0 / rk_press;
  const double alp = -k2;
  double ap = 0.0;
  double2 fp = cz();
  auto poisson_elim = [&](int i, double2 rhs) {  // reversed Thomas, row i (descending)
    auto dh = [&](int k) { return ldro(ops.dh, k); };
    auto dn = [&](int k) { return ldro(ops.dn, k); };
    double a_, b_, c_;
    if (i == 0) { a_ = 0.0; b_ = -dn(0); c_ = dn(0); }
    else if (i == 1) { a_ = dh(1) * dn(0); b_ = -dh(1) * dn(0) - dh(1) * dn(1) + alp; c_ = dh(1) * dn(1); }
    else if (i == ny - 1) {
      a_ = dh(ny - 1) * dn(ny - 2); b_ = -dh(ny - 1) * dn(ny - 2) - dh(ny - 1) * dn(ny - 1) + alp;
      c_ = dh(ny - 1) * dn(ny - 1);
    } else if (i == ny) { a_ = -dn(ny - 1); b_ = dn(ny - 1); c_ = 0.0; }
    else { a_ = ldro(d2h, 3 * i); b_ = ldro(d2h, 3 * i + 1) + alp; c_ = ldro(d2h, 3 * i + 2); }
    const double bp2 = (i == ny) ? b_ : b_ - c_ * ap;
    const double2 gg = (i == ny) ? rhs : csub(rhs, rmul(c_, fp));
    const double rb2 = 1.0 / bp2;
    ap = a_ * rb2;
    fp = rmul(rb2, gg);
    if (own && !mean) { sto(SP, OM(i), fp); sto1(SA, OM(i) >> 1, ap); }
  };
  double2 x1 = cv1 ? cz() : topc;  // x at row j+1
  if (mean && cv1) x1 = cz();
  auto loadB = [&](int j) {
    RowB3 d;
    d.s = ldo(SX, OC(j));
    d.cc = ldo1(SC, OC(j) >> 1);
    return d;
  };
  auto compB = [&](int j, const RowB3& d) {
    double2 x = csub(d.s, rmul(d.cc, x1));
    if (mean && cv1) x = botc;  // (:261-263)
    if (valid) sto(SX, OC(j), x);
    // divergence of the provisional velocity at row i = j+1 (strip_divergence rhs, :28-31)
    const int i = j + 1;
    double2 t;
    if (cv1) t = rmul(ldro(ops.dh, i <= ny - 1 ? i : ny - 1), csub(x1, x));
    else t = imul(kc, x1);
    const double2 t0 = shfl2(t, gbase), t1 = shfl2(t, gbase + 1), t2 = shfl2(t, gbase + 2);
    if (!mean) {
      double2 rhs = cz();
      if (i <= ny - 1) rhs = rmul(tmpc, cadd(cadd(t0, t1), t2));
      poisson_elim(i, rhs);
    }
    x1 = x;
  };
  if constexpr (DMA) {
    auto fetchB = [&](int j, int slot) {
      char* sb = solve_ring + slot * SOLVE_SB;
      const uint32_t o = j < 0 ? 0u : OC(j);
      dma16(reinterpret_cast<const char*>(SX) + o, sb);
      dma4(reinterpret_cast<const char*>(SC) + (o >> 1), sb + 1024);
      dma4(reinterpret_cast<const char*>(SC) + (o >> 1) + 4, sb + 1280);
    };
    auto readB = [&](int slot) {
      v2d x0;
      int lo, hi;
      const uint32_t sb = ring0 + slot * SOLVE_SB;
      ring_rd_b(sb + 16 * lane, sb + 4 * lane, x0, lo, hi);
      RowB3 d;
      d.s = d2(x0);
      d.cc = __hiloint2double(hi, lo);
      return d;
    };
    ring_rows<SOLVE_RB>(ny, 3, [ny](int t) { return ny - 1 - t; }, fetchB, readB, compB);
  } else {
    rows_down<BL>(ny - 1, 0, loadB, compB);
  }
  if (!mean) poisson_elim(0, cz());

  // ---------------- pass C: pressure correction, projection, next RK register ----------------
  const bool more = (a.rk < 3);
  const double coef2 = more ? dt * c_rka[r + 1] : 0.0;
  const double vis_exp2 = more ? (dt / re) * c_rkc[r + 1] : 0.0;
  const double rk_press2 = more ? (c_rkc[r + 1] + c_rkd[r + 1]) * dt : 0.0;
  const double alpha2 = 1.0 + vis_imp * k2;
  double2 w0 = cz(), w1 = cz(), w2 = cz(), w3 = cz();   // own component at rows i-3..i
  double2 Pw0 = cz(), Pw1 = cz();                        // P_new at rows i-1, i
  // next RK register of the own component at row j (from cv at row j, its
  // D2H/D2N stencil rows s0..s2, P_new at j, j+1 and NL at j)
  auto buildR = [&](int j, double2 cvj, double2 s0, double2 s1, double2 s2, double2 Pj, double2 Pj1, double2 nj) {
    double2 out = cadd(cvj, rmul(coef2, neg(nj)));
    const double h0 = ldro(d2h, 3 * j), h1 = ldro(d2h, 3 * j + 1), h2 = ldro(d2h, 3 * j + 2);
    if (!cv1) {
      double2 dd = cadd(cadd(rmul(h0, s0), rmul(h1, s1)), rmul(h2, s2));
      out = cadd(out, rmul(vis_exp2, csub(dd, rmul(k2, cvj))));
    } else if (j < ny) {
      const double b0 = ldro(d2n, 3 * j), b1 = ldro(d2n, 3 * j + 1), b2 = ldro(d2n, 3 * j + 2);
      double2 dd = cadd(cadd(rmul(b0, s0), rmul(b1, s1)), rmul(b2, s2));
      out = cadd(out, rmul(vis_exp2, csub(dd, rmul(k2, cvj))));
    }
    if (!mean) {
      if (cv1) {
        if (j >= 1 && j <= ny - 2) out = csub(out, rmul(rk_press2 * ldro(ops.dn, j), csub(Pj1, Pj)));
      } else if (j >= 1 && j <= ny - 1) {
        out = csub(out, imul(rk_press2 * kc, Pj));
      }
    }
    if (valid) sto(R, OC(j), out);
  };
  // projection of row i: s = provisional own component, Pold = P(i), f0 = phi(i),
  // fn = phi(i+1), st0..st2 = phi on the D2H stencil rows of i, nb = NL(i-1)
  auto c2row = [&](int i, double2 s, double2 Pold, double2 f0, double2 fn, double2 st0, double2 st1,
                   double2 st2, double2 nb) {
    double2 cvv = s;
    double2 pn = cz();
    if (!mean) {
      if (cv1) {
        if (i >= 1 && i <= ny - 2) cvv = csub(cvv, rmul(rk_press * ldro(ops.dn, i), csub(fn, f0)));
      } else if (i >= 1 && i <= ny - 1) {
        cvv = csub(cvv, imul(rk_press * kc, f0));
      }
      double2 d2 = cadd(cadd(rmul(ldro(d2h, 3 * i), st0), rmul(ldro(d2h, 3 * i + 1), st1)), rmul(ldro(d2h, 3 * i + 2), st2));
      pn = csub(cadd(Pold, rmul(alpha2, f0)), rmul(vis_imp, d2));
    }
    if (valid) sto(U, OC(i), cvv);
    if (own) sto(P, OM(i), pn);
    w0 = w1; w1 = w2; w2 = w3; w3 = cvv;
    Pw0 = Pw1; Pw1 = pn;
    if (mean && c == 0) {  // d_um_dy update from the new mean profile (:84-89)
      if (i >= 2 && i <= ny - 1) a.dudy[i - 1] = (ldro(ops.dn, i - 1) * (w3.x - w2.x));
      if (i == 2) a.dudy[0] = ldro(dm, 0) * w1.x + ldro(dm, 1) * w2.x + ldro(dm, 2) * w3.x;
      if (i == ny) a.dudy[ny - 1] = ldro(dm, 3) * w1.x + ldro(dm, 4) * w2.x + ldro(dm, 5) * w3.x;
    }
    if (!more) {
      if (valid) sto(R, OC(i), cvv);
      return;
    }
    if (i >= 2) {
      if (i == 2)  // row 0 (clamped stencils on rows 0..2, no gradient)
        buildR(0, w1, w1, w2, w3, cz(), cz(), ldo(NL, OC(0)));
      // row j = i-1: centred stencil on rows i-2..i; D2N clamped at the last v row
      const int j = i - 1;
      if (j == ny - 1 && cv1) buildR(j, w2, w0, w1, w2, Pw0, Pw1, nb);
      else buildR(j, w2, w1, w2, w3, Pw0, Pw1, nb);
    }
  };
  auto lsp = [&](int i) { return mean ? cz() : ldo(SP, OM(i)); };
  auto lsa = [&](int i) { return mean ? 0.0 : ldo1(SA, OM(i) >> 1); };
  const double2 ph0 = lsp(0);
  const double2 ph1 = csub(lsp(1), rmul(lsa(1), ph0));
  const double2 ph2 = csub(lsp(2), rmul(lsa(2), ph1));
  const double2 pb = make_double2((0.0 - ldro(dm, 1) * ph1.x - ldro(dm, 2) * ph2.x) / ldro(dm, 0),
                                  (0.0 - ldro(dm, 1) * ph1.y - ldro(dm, 2) * ph2.y) / ldro(dm, 0));
  c2row(0, ldo(SX, OC(0)), ldo(P, OM(0)), pb, cz(), pb, ph1, ph2, cz());
  c2row(1, ldo(SX, OC(1)), ldo(P, OM(1)), ph1, ph2, pb, ph1, ph2, cz());
  double2 q0 = pb, q1 = ph1, q2 = ph2, ph = ph2;
  auto loadC = [&](int i) {
    RowC3 d;
    d.sp = lsp(i); d.sa = lsa(i);
    d.sx = ldo(SX, OC(i - 1));
    d.p = ldo(P, OM(i - 1));
    d.n = more ? ldo(NL, OC(i - 2)) : cz();
    return d;
  };
  auto compC = [&](int i, const RowC3& d) {
    ph = csub(d.sp, rmul(d.sa, ph));
    q0 = q1; q1 = q2; q2 = ph;
    c2row(i - 1, d.sx, d.p, q1, q2, q0, q1, q2, d.n);
  };
  if constexpr (DMA) {
    auto fetchC = [&](int i, int slot) {
      char* sb = solve_ring + slot * SOLVE_SC;
      const bool dmy = i < 0;
      const uint32_t om = dmy ? 0u : OM(i), ox = dmy ? 0u : OC(i - 1), opp = dmy ? 0u : OM(i - 1),
                     on = dmy ? 0u : OC(i - 2);
      dma16(reinterpret_cast<const char*>(SP) + om, sb);
      dma16(reinterpret_cast<const char*>(SX) + ox, sb + 1024);
      dma16(reinterpret_cast<const char*>(P) + opp, sb + 2048);
      if (more) dma16(reinterpret_cast<const char*>(NL) + on, sb + 3072);
      dma4(reinterpret_cast<const char*>(SA) + (om >> 1), sb + 4096);
      dma4(reinterpret_cast<const char*>(SA) + (om >> 1) + 4, sb + 4352);
    };
    auto readC = [&](int slot) {
      v2d x0, x1, x2, x3;
      int lo, hi;
      const uint32_t sb = ring0 + slot * SOLVE_SC;
      ring_rd_c(sb + 16 * lane, sb + 4 * lane, x0, x1, x2, x3, lo, hi);
      RowC3 d;
      d.sp = mean ? cz() : d2(x0);
      d.sx = d2(x1);
      d.p = d2(x2);
      d.n = more ? d2(x3) : cz();
      d.sa = mean ? 0.0 : __hiloint2double(hi, lo);
      return d;
    };
    ring_rows<SOLVE_RC>(ny - 3, more ? 6 : 5, [](int t) { return 3 + t; }, fetchC, readC, compC);
  } else {
    rows_up<BL>(3, ny - 1, loadC, compC);
  }
  // top wall: phi(ny) from phi(ny-2), phi(ny-1)
  const double2 pt = make_double2((0.0 - ldro(dm, 3) * q1.x - ldro(dm, 4) * q2.x) / ldro(dm, 5),
                                  (0.0 - ldro(dm, 3) * q1.y - ldro(dm, 4) * q2.y) / ldro(dm, 5));
  c2row(ny - 1, ldo(SX, OC(ny - 1)), ldo(P, OM(ny - 1)), q2, pt, q1, q2, pt, more ? ldo(NL, OC(ny - 2)) : cz());
  c2row(ny, ldo(SX, OC(ny)), ldo(P, OM(ny)), pt, cz(), q1, q2, pt, more ? ldo(NL, OC(ny - 1)) : cz());
  if (more) buildR(ny, w3, w1, w2, w3, Pw1, cz(), ldo(NL, OC(ny)));
}

// Wall-normal velocity of a jet at one wall for substep rk: the spectral
// profile times 1 + forcing sin(2 pi f t) at the substep's end time
// t = t_total + dt sum_{r<=rk} (rkc(r) + rkd(r)) (the time rhs_solve hands to
// jet_profile, PFC/pass2.f90:38-43).
__global__ void k_jet_bc(const DevScalars* __restrict__ sc, int rk, double forcing, double freq,
                         const double2* __restrict__ prof, double2* __restrict__ cbcv, long nm) {
  const long m = (long)blockIdx.x * blockDim.x + threadIdx.x;
  if (m >= nm) return;
  double sum = 0.0;
  for (int r = 0; r < rk; ++r) sum = sum + (c_rkc[r] + c_rkd[r]);
  const double t = sc->t_total + sc->dt * sum;
  const double amp = 1.0 + forcing * sin(6.283185307179586 * freq * t);
  const double2 p = prof[m];
  cbcv[m] = make_double2(amp * p.x, amp * p.y);
}

void launch_jet_bc(const DevScalars* sc, int rk, double forcing, double freq, const double2* prof, double2* cbcv,
                   long nm, hipStream_t s) {
  hipLaunchKernelGGL(k_jet_bc, dim3((unsigned)((nm + 255) / 256)), dim3(256), 0, s, sc, rk, forcing, freq, prof,
                     cbcv, nm);
  LAUNCH_CHECK("k_jet_bc");
}

void launch_solve(const DevOps& ops, const SolveArgs& a, hipStream_t s) {
  long nm = (long)a.nxh * a.nz;
  dim3 grid((unsigned)((nm + 63) / 64));
  const int bl = a.bl, lanes = a.lanes;
  // the segmented and component-parallel kernels need the byte offsets of one
  // array to fit 32 bits
  const bool off32 = 3.0 * (a.ny + 1) * (double)nm * 16.0 < 4.0e9;
  // two modes per wavefront (32 segments): the 5-row segments keep the
  // wavefront within 256 registers, two wavefronts per SIMD
  const int seg2 = (a.seg == 2) ? solve_seg_rows(a.ny, 32) : 0;
  if (seg2 == 5 && off32) {
    dim3 gs((unsigned)((nm + 1) / 2));
    hipLaunchKernelGGL((k_solve_seg<2, 5>), gs, dim3(64), 0, s, ops, a);
    LAUNCH_CHECK("k_solve_seg");
    return;
  }
  // eight modes per wavefront (8 segments) where ny+1 rows fit 8 segments of
  // 5..9 rows: 128-byte row pieces, and fewer idle segment lanes at ny + 1
  // just under 8 SEG (66 rows: 8 segments of 9 rows, against 11 of 16 lanes
  // groups busy with 6-row segments)
  const int seg8 = (a.seg == 3) ? solve_seg_rows(a.ny, 8) : 0;
  if (seg8 && off32) {
    dim3 gs((unsigned)((nm + 7) / 8));
    switch (seg8) {
      case 5: hipLaunchKernelGGL((k_solve_seg<8, 5>), gs, dim3(64), 0, s, ops, a); break;
      case 6: hipLaunchKernelGGL((k_solve_seg<8, 6>), gs, dim3(64), 0, s, ops, a); break;
      case 7: hipLaunchKernelGGL((k_solve_seg<8, 7>), gs, dim3(64), 0, s, ops, a); break;
      case 8: hipLaunchKernelGGL((k_solve_seg<8, 8>), gs, dim3(64), 0, s, ops, a); break;
      default: hipLaunchKernelGGL((k_solve_seg<8, 9>), gs, dim3(64), 0, s, ops, a); break;
    }
    LAUNCH_CHECK("k_solve_seg");
    return;
  }
  const int seg = a.seg ? solve_seg_rows(a.ny) : 0;
  if (seg && off32) {
    dim3 gs((unsigned)((nm + 3) / 4));
    switch (seg) {
      case 5: hipLaunchKernelGGL((k_solve_seg<4, 5>), gs, dim3(64), 0, s, ops, a); break;
      case 6: hipLaunchKernelGGL((k_solve_seg<4, 6>), gs, dim3(64), 0, s, ops, a); break;
      case 7: hipLaunchKernelGGL((k_solve_seg<4, 7>), gs, dim3(64), 0, s, ops, a); break;
      case 8: hipLaunchKernelGGL((k_solve_seg<4, 8>), gs, dim3(64), 0, s, ops, a); break;
      default: hipLaunchKernelGGL((k_solve_seg<4, 9>), gs, dim3(64), 0, s, ops, a); break;
    }
    LAUNCH_CHECK("k_solve_seg");
    return;
  }
  if (lanes == 3 && off32) {
    dim3 g3((unsigned)((nm + 19) / 20));
    if (a.dma) hipLaunchKernelGGL((k_solve3<8, true>), g3, dim3(64), SOLVE_RING_BYTES, s, ops, a);
    else if (bl == 4) hipLaunchKernelGGL((k_solve3<4, false>), g3, dim3(64), 0, s, ops, a);
    else if (bl == 16) hipLaunchKernelGGL((k_solve3<16, false>), g3, dim3(64), 0, s, ops, a);
    else hipLaunchKernelGGL((k_solve3<8, false>), g3, dim3(64), 0, s, ops, a);
    LAUNCH_CHECK("k_solve3");
    return;
  }
  if (bl == 4) hipLaunchKernelGGL(k_solve<4>, grid, dim3(64), 0, s, ops, a);
  else if (bl == 16) hipLaunchKernelGGL(k_solve<16>, grid, dim3(64), 0, s, ops, a);
  else hipLaunchKernelGGL(k_solve<8>, grid, dim3(64), 0, s, ops, a);
  LAUNCH_CHECK("k_solve");
}

// ---------------------------------------------------------------------------
// Projection for the initial field (mode 0: remove_initial_divergence,
// PFC/init_vel.f90:518-536) and the initial pressure (mode 1:
// solve_continuous_poisson, PFC/init_vel.f90:477-513).  Solves the Poisson
// system in the reference's direction (SOLVE_P, PFC/solve.f90:269-338).
// ---------------------------------------------------------------------------
__global__ __launch_bounds__(64) void k_project(DevOps ops, int ny, int nxh, int nz,
                                                const double2* __restrict__ in, double2* out_vel,
                                                double2* out_p, double2* scr, double coef, int mode,
                                                const double2* __restrict__ vel, double re,
                                                const int* __restrict__ cols, int ncols) {
  // Two layouts: the whole field (cols == nullptr; system m = (kx, kz) of
  // [3][ny+1][nxh][nz], pressure [ny+1][nxh][nz]) or ncols independent
  // columns (system m has mode indices cols[2m], cols[2m+1]; velocity
  // complex(ny+1, 3, ncols) and pressure complex(ny+1, ncols), Fortran order).
  const long nm = cols ? (long)ncols : (long)nxh * nz;
  const long m = (long)blockIdx.x * blockDim.x + threadIdx.x;
  if (m >= nm) return;
  const int nyp = ny + 1;
  int ix, iz;
  long vb, vcs, vrs, pb, prs;
  if (cols) {
    ix = cols[2 * m]; iz = cols[2 * m + 1];
    vb = m * 3 * nyp; vcs = nyp; vrs = 1; pb = m * nyp; prs = 1;
  } else {
    ix = (int)(m / nz); iz = (int)(m - (long)ix * nz);
    vb = m; vcs = (long)nyp * nm; vrs = nm; pb = m; prs = nm;
  }
  const double kxv = ops.kx[ix], kzv = ops.kz[iz];
  const double k2 = kxv * kxv + kzv * kzv;
  const bool mean = (k2 == 0.0);
  const double* d2h = ops.d2h; const double* d2n = ops.d2n; const double* dh = ops.dh;
  const double* dn = ops.dn; const double* dm = ops.dm;
  auto VI = [&](int c, int j) -> long { return vb + c * vcs + j * vrs; };
  auto PI = [&](int j) -> long { return pb + j * prs; };
  auto X = [&](int c, int j) -> double2 {
    if (mode != 1) return in[VI(c, j)];
    // mode 1: cru = -NL + (1/Re)(D2 u - k2 u)  (:488-501)
    double2 t = in[VI(c, j)];
    double2 v = make_double2(-t.x, -t.y);
    const double visc = 1.0 / re;
    auto uu = [&](int cc, int jj) { return vel[VI(cc, jj)]; };
    if (c != 1) {
      int j0 = min(max(j - 1, 0), nyp - 3);
      double2 du = cadd(cadd(rmul(d2h[3 * j], uu(c, j0)), rmul(d2h[3 * j + 1], uu(c, j0 + 1))),
                        rmul(d2h[3 * j + 2], uu(c, j0 + 2)));
      v = cadd(v, rmul(visc, csub(du, rmul(k2, uu(c, j)))));
    } else if (j < ny) {
      int j0 = min(max(j - 1, 0), ny - 3);
      double2 du = cadd(cadd(rmul(d2n[3 * j], uu(1, j0)), rmul(d2n[3 * j + 1], uu(1, j0 + 1))),
                        rmul(d2n[3 * j + 2], uu(1, j0 + 2)));
      v = cadd(v, rmul(visc, csub(du, rmul(k2, uu(1, j)))));
    }
    return v;
  };
  double2* F = scr;                                        // f' then phi
  double* C = reinterpret_cast<double*>(scr + (long)nyp * nm);  // c'
  if (mean) {
    if (mode == 0) {
      // set the mean wall-normal velocity to its bottom-wall value (:526-528)
      double2 v0 = in[VI(1, 0)];
      for (int c = 0; c < 3; ++c)
        for (int j = 0; j <= ny; ++j) {
          double2 v = in[VI(c, j)];
          if (c == 1 && j >= 1 && j <= ny - 1) v = v0;
          out_vel[VI(c, j)] = v;
        }
    } else if (mode == 2 && out_vel && out_vel != in) {
      for (int c = 0; c < 3; ++c)
        for (int j = 0; j <= ny; ++j)
          out_vel[VI(c, j)] = in[VI(c, j)];
    }
    if (out_p) for (int j = 0; j <= ny; ++j) out_p[PI(j)] = cz();
    return;
  }
  const double tmp = 1.0 / coef, alp = -k2;
  // forward elimination (ascending), rows as in SOLVE_P
  double cprev = 0.0; double2 fprev = cz();
  for (int i = 0; i <= ny; ++i) {
    double a_, b_, c_;
    if (i == 0) { a_ = 0.0; b_ = -dn[0]; c_ = dn[0]; }
    else if (i == 1) { a_ = dh[1] * dn[0]; b_ = -dh[1] * dn[0] - dh[1] * dn[1] + alp; c_ = dh[1] * dn[1]; }
    else if (i == ny - 1) {
      a_ = dh[ny - 1] * dn[ny - 2]; b_ = -dh[ny - 1] * dn[ny - 2] - dh[ny - 1] * dn[ny - 1] + alp;
      c_ = dh[ny - 1] * dn[ny - 1];
    } else if (i == ny) { a_ = -dn[ny - 1]; b_ = dn[ny - 1]; c_ = 0.0; }
    else { a_ = d2h[3 * i]; b_ = d2h[3 * i + 1] + alp; c_ = d2h[3 * i + 2]; }
    double2 rhs = cz();
    if (i >= 1 && i <= ny - 1) {
      double2 d = cadd(cadd(imul(kxv, X(0, i)), rmul(dh[i], csub(X(1, i), X(1, i - 1)))), imul(kzv, X(2, i)));
      rhs = rmul(tmp, d);
    }
    double bp = (i == 0) ? b_ : b_ - a_ * cprev;
    double2 g = (i == 0) ? rhs : csub(rhs, rmul(a_, fprev));
    double rb = 1.0 / bp;
    cprev = c_ * rb;
    fprev = rmul(rb, g);
    F[PI(i)] = fprev;
    C[PI(i)] = cprev;
  }
  // back substitution (descending)
  double2 x1 = F[PI(ny)];
  double2 p1 = cz(), p2 = cz(), pm1 = cz(), pm2 = cz();
  for (int i = ny - 1; i >= 0; --i) {
    double2 x = csub(F[PI(i)], rmul(C[PI(i)], x1));
    F[PI(i)] = x;
    if (i == 1) p1 = x;
    if (i == 2) p2 = x;
    if (i == ny - 2) pm2 = x;
    if (i == ny - 1) pm1 = x;
    x1 = x;
  }
  double2 pbot = make_double2((0.0 - dm[1] * p1.x - dm[2] * p2.x) / dm[0], (0.0 - dm[1] * p1.y - dm[2] * p2.y) / dm[0]);
  double2 ptop = make_double2((0.0 - dm[3] * pm2.x - dm[4] * pm1.x) / dm[5], (0.0 - dm[3] * pm2.y - dm[4] * pm1.y) / dm[5]);
  F[PI(0)] = pbot;
  F[PI(ny)] = ptop;
  if (out_p) for (int j = 0; j <= ny; ++j) out_p[PI(j)] = F[PI(j)];
  if ((mode == 0 || mode == 2) && out_vel) {
    // subtract_pressure_gradient(cu, cp, coef) (:72-75)
    for (int j = 0; j <= ny; ++j) {
      double2 u = X(0, j), v = X(1, j), w = X(2, j);
      double2 f0 = F[PI(j)];
      if (j >= 1 && j <= ny - 1) {
        u = csub(u, imul(coef * kxv, f0));
        w = csub(w, imul(coef * kzv, f0));
      }
      if (j >= 1 && j <= ny - 2) v = csub(v, rmul(coef * dn[j], csub(F[PI(j + 1)], f0)));
      out_vel[VI(0, j)] = u;
      out_vel[VI(1, j)] = v;
      out_vel[VI(2, j)] = w;
    }
  }
}

// subtract_pressure_gradient (PFC/solve.f90:54-77) for ncols independent
// columns with a given pressure: crhs complex(ny+1, 3, ncols) in place,
// cpress complex(ny+1, ncols); mode indices cols[2m], cols[2m+1].
__global__ __launch_bounds__(64) void k_subtract_grad_cols(DevOps ops, int ny, double2* crhs,
                                                           const double2* __restrict__ cpress, double coef,
                                                           const int* __restrict__ cols, int ncols) {
  const long m = (long)blockIdx.x * blockDim.x + threadIdx.x;
  if (m >= ncols) return;
  const int nyp = ny + 1;
  const double kxv = ops.kx[cols[2 * m]], kzv = ops.kz[cols[2 * m + 1]];
  double2* u = crhs + m * 3 * nyp;
  const double2* p = cpress + m * nyp;
  for (int j = 1; j <= ny - 1; ++j) {
    u[j] = csub(u[j], imul(coef * kxv, p[j]));
    u[2 * nyp + j] = csub(u[2 * nyp + j], imul(coef * kzv, p[j]));
  }
  for (int j = 1; j <= ny - 2; ++j) u[nyp + j] = csub(u[nyp + j], rmul(coef * ops.dn[j], csub(p[j + 1], p[j])));
}

void launch_project(const DevOps& ops, int ny, int nxh, int nz, const double2* in, double2* out_vel,
                    double2* out_p, double2* scr, double coef, int mode, double2* vel_for_visc,
                    double re, hipStream_t s, const int* cols, int ncols) {
  long nm = cols ? (long)ncols : (long)nxh * nz;
  if (mode < 0 || mode > 2) throw chan_error("k_project: bad mode");
  if (mode == 1 && !vel_for_visc) throw chan_error("k_project: mode 1 needs the velocity");
  if (mode != 1 && !out_vel) throw chan_error("k_project: projection needs an output field");
  if (nm <= 0) return;
  dim3 grid((unsigned)((nm + 63) / 64));
  hipLaunchKernelGGL(k_project, grid, dim3(64), 0, s, ops, ny, nxh, nz, in, out_vel, out_p, scr, coef,
                     mode, vel_for_visc, re, cols, ncols);
  LAUNCH_CHECK("k_project");
}

void launch_subtract_grad_cols(const DevOps& ops, int ny, double2* crhs, const double2* cpress, double coef,
                               const int* cols, int ncols, hipStream_t s) {
  if (ncols <= 0) return;
  hipLaunchKernelGGL(k_subtract_grad_cols, dim3((ncols + 63) / 64), dim3(64), 0, s, ops, ny, crhs, cpress, coef,
                     cols, ncols);
  LAUNCH_CHECK("k_subtract_grad_cols");
}

// ---------------------------------------------------------------------------
// CFL number and dt adaptation (check_cfl PFC/pass1.f90:202-250, :106-124)
// ---------------------------------------------------------------------------
__device__ __forceinline__ void cfl_finalize_body(const DevOps& ops, const chan_params& prm,
                                                  unsigned long long* cfl, DevScalars* sc, int step) {
  // one wavefront: lane l takes the planes k = 2 + l, 2 + l + 64, ...; the
  // maximum and the abort flag do not depend on the order of the planes.
  // Each lane clears the maxima it has read, so the next pass starts from zero.
  const int lane = threadIdx.x;
  const int ny = prm.ny;
  const double dt = sc->dt;
  double cmax = 0.0;   // every plane's CFL number is >= 0
  int err = 0;
  for (int k = 2 + lane; k <= ny; k += 64) {
    const int kk = k - 1;
    double amax = __longlong_as_double((long long)cfl[3 * kk + 0]);
    double vmax = __longlong_as_double((long long)cfl[3 * kk + 1]);
    double wmax = __longlong_as_double((long long)cfl[3 * kk + 2]);
    cfl[3 * kk + 0] = 0ull;
    cfl[3 * kk + 1] = 0ull;
    cfl[3 * kk + 2] = 0ull;
    double tempa = 0.5 * amax * (double)prm.nx / (0.55 * prm.lx);
    double tempb = vmax / (1.73 * fmin(ops.hn[k - 1], ops.hn[k - 2]));
    double tempc = 0.5 * wmax * (double)prm.nz / (0.55 * prm.lz);
    double c = (tempa + tempb + tempc) * dt;
    if (prm.dt_flag == 1) { if (c >= 3.0e2) err = 2; }
    else if (c >= 1.15) err = 1;
    if (!(c <= cmax)) cmax = c;   // a NaN plane makes the maximum NaN
  }
  for (int o = 32; o > 0; o >>= 1) {
    const double oc = __shfl_xor(cmax, o, 64);
    if (!(oc <= cmax)) cmax = oc;
    err = max(err, __shfl_xor(err, o, 64));
  }
  if (lane != 0) return;
  sc->cfl = cmax;
  if (err && !sc->err) sc->err = err;
  if (step >= 0 && prm.dt_flag == 1) {
    double dt_temp = dt * prm.cfl_max / cmax;
    sc->dt = fmin(dt_temp, dt + (dt_temp - dt) * dt / (1.0 + dt));
  }
}

__global__ __launch_bounds__(64) void k_cfl_finalize(DevOps ops, chan_params prm,
                                                     unsigned long long* cfl, DevScalars* sc,
                                                     int step) {
  cfl_finalize_body(ops, prm, cfl, sc, step);
}

// One wavefront per channel: the CFL maxima and dt of every channel of a batch.
__global__ __launch_bounds__(64) void k_cfl_finalize_b(const CflBatchArgs* __restrict__ args, int step) {
  const CflBatchArgs& x = args[blockIdx.x];
  cfl_finalize_body(x.ops, x.prm, x.cfl, x.sc, step);
}

void launch_cfl_finalize_batched(const CflBatchArgs* args, int n, int step, hipStream_t s) {
  if (n <= 0) return;
  hipLaunchKernelGGL(k_cfl_finalize_b, dim3(n), dim3(64), 0, s, args, step);
  LAUNCH_CHECK("k_cfl_finalize_b");
}

void launch_cfl_finalize(const DevOps& ops, chan_params prm, unsigned long long* cfl, DevScalars* sc,
                         int step, hipStream_t s) {
  hipLaunchKernelGGL(k_cfl_finalize, dim3(1), dim3(64), 0, s, ops, prm, cfl, sc, step);
  LAUNCH_CHECK("k_cfl_finalize");
}

__global__ void k_fill_zero(unsigned long long* p, size_t n) {
  size_t i = (size_t)blockIdx.x * blockDim.x + threadIdx.x;
  size_t st = (size_t)gridDim.x * blockDim.x;
  for (; i < n; i += st) p[i] = 0ull;
}

void launch_fill_zero(void* p, size_t bytes, hipStream_t s) {
  size_t n = bytes / 8;
  if (n == 0) return;
  unsigned blocks = (unsigned)std::min<size_t>((n + 255) / 256, 4096);
  hipLaunchKernelGGL(k_fill_zero, dim3(blocks), dim3(256), 0, s, (unsigned long long*)p, n);
  LAUNCH_CHECK("k_fill_zero");
}

// Zero fill of several buffers in one launch (the batched ensemble): y-row i
// of the grid clears buffer i (8-byte words).
__global__ void k_fill_zero_b(const FillBatchArgs* __restrict__ args) {
  const FillBatchArgs f = args[blockIdx.y];
  for (size_t i = (size_t)blockIdx.x * blockDim.x + threadIdx.x; i < f.n; i += (size_t)gridDim.x * blockDim.x)
    f.p[i] = 0ull;
}

void launch_fill_zero_batched(const FillBatchArgs* args, int n, size_t max_words, hipStream_t s) {
  if (n <= 0 || max_words == 0) return;
  const unsigned bx = (unsigned)std::min<size_t>((max_words + 255) / 256, 256);
  hipLaunchKernelGGL(k_fill_zero_b, dim3(bx, n), dim3(256), 0, s, args);
  LAUNCH_CHECK("k_fill_zero_b");
}

// t_total += dt of every channel of a batch
__global__ void k_advance_time_b(DevScalars* const* __restrict__ sc, int n) {
  const int i = blockIdx.x * blockDim.x + threadIdx.x;
  if (i < n) sc[i]->t_total = sc[i]->t_total + sc[i]->dt;
}

void launch_advance_time_batched(DevScalars* const* sc, int n, hipStream_t s) {
  if (n <= 0) return;
  hipLaunchKernelGGL(k_advance_time_b, dim3((n + 63) / 64), dim3(64), 0, s, sc, n);
  LAUNCH_CHECK("k_advance_time_b");
}

// The segmented solve of several channels with the same segment length in one launch.
// layout = 16 * modes per wavefront + rows per segment (solve_seg_layout)
void launch_solve_batched(int layout, const SBatchArgs* args, const BatchItem* items, int nitems, hipStream_t s) {
  if (nitems <= 0) return;
  switch (layout) {
    case 64 + 5: hipLaunchKernelGGL((k_solve_seg_b<4, 5>), dim3(nitems), dim3(64), 0, s, args, items); break;
    case 64 + 6: hipLaunchKernelGGL((k_solve_seg_b<4, 6>), dim3(nitems), dim3(64), 0, s, args, items); break;
    case 64 + 7: hipLaunchKernelGGL((k_solve_seg_b<4, 7>), dim3(nitems), dim3(64), 0, s, args, items); break;
    case 64 + 8: hipLaunchKernelGGL((k_solve_seg_b<4, 8>), dim3(nitems), dim3(64), 0, s, args, items); break;
    case 64 + 9: hipLaunchKernelGGL((k_solve_seg_b<4, 9>), dim3(nitems), dim3(64), 0, s, args, items); break;
    case 128 + 5: hipLaunchKernelGGL((k_solve_seg_b<8, 5>), dim3(nitems), dim3(64), 0, s, args, items); break;
    case 128 + 6: hipLaunchKernelGGL((k_solve_seg_b<8, 6>), dim3(nitems), dim3(64), 0, s, args, items); break;
    case 128 + 7: hipLaunchKernelGGL((k_solve_seg_b<8, 7>), dim3(nitems), dim3(64), 0, s, args, items); break;
    case 128 + 8: hipLaunchKernelGGL((k_solve_seg_b<8, 8>), dim3(nitems), dim3(64), 0, s, args, items); break;
    case 128 + 9: hipLaunchKernelGGL((k_solve_seg_b<8, 9>), dim3(nitems), dim3(64), 0, s, args, items); break;
    default: throw chan_error("batched solve: no segmented solve of layout " + std::to_string(layout));
  }
  LAUNCH_CHECK("k_solve_seg_b");
}

int solve_seg_layout(int ny, int solve_seg) {
  // The batched launches take 8 modes per wavefront by default where all 8
  // segments are busy (66 rows: 8 of 9 rows, against 11 of 16 segments of 6
  // rows with 4 modes): 179 -> 182 steps/s on the 48-member ensemble, whose
  // launches keep the GPU full.  Where segments would idle (50 rows: 6 of 8)
  // it was slower, and so is a single channel with fewer than ~4000 modes,
  // whose wavefronts are then too few (64x65x64: 21 -> 26 us per solve).
  int s8 = (solve_seg == 1 || solve_seg == 3) ? solve_seg_rows(ny, 8) : 0;
  if (s8 && solve_seg == 1 && (ny + 1 + s8 - 1) / s8 < 8) s8 = 0;
  if (s8) return 128 + s8;
  const int s4 = solve_seg_rows(ny);
  return s4 ? 64 + s4 : 0;
}

// max |ikx u + dh (v_j - v_{j-1}) + ikz w| over non-mean modes, j = 2..ny
__global__ __launch_bounds__(256) void k_divmax(DevOps ops, int ny, int nxh, int nz,
                                                const double2* __restrict__ vel,
                                                unsigned long long* out) {
  const long nm = (long)nxh * nz;
  const long m = (long)blockIdx.x * blockDim.x + threadIdx.x;
  double mx = 0.0;
  if (m < nm) {
    const int ix = (int)(m / nz), iz = (int)(m - (long)ix * nz);
    const double kxv = ops.kx[ix], kzv = ops.kz[iz];
    const int nyp = ny + 1;
    const long MS = nm;
    if (kxv * kxv + kzv * kzv != 0.0) {
      for (int j = 1; j <= ny - 1; ++j) {
        double2 u = vel[((long)0 * nyp + j) * MS + m];
        double2 v = vel[((long)1 * nyp + j) * MS + m];
        double2 vm = vel[((long)1 * nyp + j - 1) * MS + m];
        double2 w = vel[((long)2 * nyp + j) * MS + m];
        double2 d = cadd(cadd(imul(kxv, u), rmul(ops.dh[j], csub(v, vm))), imul(kzv, w));
        mx = fmax(mx, sqrt(d.x * d.x + d.y * d.y));
      }
    }
  }
  for (int o = 32; o > 0; o >>= 1) mx = fmax(mx, __shfl_xor(mx, o, 64));
  if ((threadIdx.x & 63) == 0) atomicMax(out, (unsigned long long)__double_as_longlong(mx));
}

void launch_divmax(const DevOps& ops, int ny, int nxh, int nz, const double2* vel,
                   unsigned long long* out, hipStream_t s) {
  long nm = (long)nxh * nz;
  hipLaunchKernelGGL(k_divmax, dim3((unsigned)((nm + 255) / 256)), dim3(256), 0, s, ops, ny, nxh, nz,
                     vel, out);
  LAUNCH_CHECK("k_divmax");
}

// d_um_dy from the mean profile (PFC/init_vel.f90:333-342)
__global__ void k_dudy(DevOps ops, int ny, int nz, const double2* __restrict__ vel, double* dudy,
                       long MS) {
  if (threadIdx.x != 0 || blockIdx.x != 0) return;
  const double* dm = ops.dm;
  auto u = [&](int j) { return vel[(long)j * MS].x; };
  dudy[0] = dm[0] * u(0) + dm[1] * u(1) + dm[2] * u(2);
  for (int j = 1; j <= ny - 2; ++j) dudy[j] = ops.dn[j] * (u(j + 1) - u(j));
  dudy[ny - 1] = dm[3] * u(ny - 2) + dm[4] * u(ny - 1) + dm[5] * u(ny);
  (void)nz;
}

void launch_dudy(const DevOps& ops, int ny, int nxh, int nz, const double2* vel, double* dudy,
                 hipStream_t s) {
  // vel points at mode (0,0) of component u; planes are nxh*nz modes apart
  long MS = (long)nxh * nz;
  hipLaunchKernelGGL(k_dudy, dim3(1), dim3(64), 0, s, ops, ny, nz, vel, dudy, MS);
  LAUNCH_CHECK("k_dudy");
}

// Layout conversion between the reference's data2/data4 and the plane-major
// device arrays.  kind 0: data2(ny+1,3,nz,nxh) -> R[3][ny+1][nxh][nz];
// kind 1: inverse; kind 2: data4(ny+1,nz,nxh) -> P[ny+1][nxh][nz]; kind 3: inverse.
__global__ void k_layout(int kind, int ny, int nxh, int nz, const double2* __restrict__ in,
                         double2* __restrict__ out) {
  const int nyp = ny + 1;
  const long nm = (long)nxh * nz;
  const long ncomp = (kind <= 1) ? 3 : 1;
  const long tot = ncomp * nyp * nm;
  for (long e = (long)blockIdx.x * blockDim.x + threadIdx.x; e < tot; e += (long)gridDim.x * blockDim.x) {
    // e indexes the plane-major array: ((c*nyp + j)*nxh + x)*nz + z
    long z = e % nz, t = e / nz;
    long x = t % nxh; t /= nxh;
    long j = t % nyp; long c = t / nyp;
    long f;  // index in the Fortran layout
    if (kind <= 1) f = (((x * nz + z) * 3 + c) * nyp + j);
    else f = ((x * nz + z) * nyp + j);
    if (kind == 0 || kind == 2) out[e] = in[f];
    else out[f] = in[e];
  }
}

void launch_layout(int kind, int ny, int nxh, int nz, const double2* in, double2* out, hipStream_t s) {
  long tot = (long)((kind <= 1) ? 3 : 1) * (ny + 1) * nxh * nz;
  unsigned blocks = (unsigned)std::min<long>((tot + 255) / 256, 8192);
  hipLaunchKernelGGL(k_layout, dim3(blocks), dim3(256), 0, s, kind, ny, nxh, nz, in, out);
  LAUNCH_CHECK("k_layout");
}
