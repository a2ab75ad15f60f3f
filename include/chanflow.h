/*
 * chanflow.h -- C-ABI of the MI355X-native channel-flow time step.
 *
 * Drop-in boundary for the pseudo-spectral RK3/Crank-Nicolson step of the
 * MPI-Fortran channel solver.  The kept Fortran driver (main_mpi.f90, setup,
 * read_write, init_vel, init_jet_buffer) binds these entry points through
 * iso_c_binding in replacement modules that keep the reference's module names
 * and PUBLIC lists (transform, pass1, pass2, solve, pass1_linearized).
 *
 * Conventions
 *  - Every entry point returns int status: 0 = OK, nonzero = error; the
 *    message is available from chan_last_error().  The reference has no
 *    error codes and STOPs (e.g. PFC/pass1.f90:243); the Fortran wrappers
 *    print chan_last_error() and STOP.
 *  - Complex arrays are interleaved double pairs (Fortran COMPLEX(8)); all
 *    host arrays use the reference's column-major layouts.  Calls copy host
 *    arrays and never retain host pointers.
 *  - One HIP stream per context; calls are stream-ordered and synchronise
 *    before returning host-visible outputs.  A context is not thread-safe.
 *  - PFC/ abbreviates ParallelFortran-Code/ of the reference.
 */
#ifndef CHANFLOW_H
#define CHANFLOW_H

#ifdef __cplusplus
extern "C" {
#endif

typedef struct chan_ctx chan_ctx;

/* Run parameters: the subset of runparms/gridparms the step reads
 * (PFC/runparms.f90:22-34, PFC/gridparms.f90:82-85). */
typedef struct chan_params {
  int nx, ny, nz;        /* spectral grid: nx/2+1 stored kx, ny+1 planes, nz kz  */
  int nx2, nz2;          /* dealiased physical grid (3/2 rule)                     */
  double lx, lz;         /* box lengths (already multiplied by pi if scaled)       */
  double re;             /* Reynolds number                                        */
  double beta;           /* mean pressure-gradient angle in radians                */
  int fix_mass;          /* 1: constant mass flux (PFC/pass2.f90:219)              */
  int dt_flag;           /* 1: CFL-adaptive dt (PFC/pass1.f90:120-123)             */
  double cfl_max;        /* target CFL for dt_flag == 1 (PFC/setup.f90:1378)       */
  double ctrlu, ctrlw;   /* wall feedback gains (PFC/setup.f90:61-62)              */
  int bl_grid;           /* must be 0: channel (Dirichlet) walls                   */
} chan_params;

/* Operators produced by define_coordinates/define_parameters
 * (PFC/setup.f90:541-754, 1440-1466), passed verbatim from the Fortran
 * globals so both sides use bit-identical operators.  buffer and
 * target_inflow may be NULL (fringe off: both identically zero). */
typedef struct chan_operators {
  const double* d2h;            /* d2h(3, ny+1)           */
  const double* d2n;            /* d2n(3, ny)             */
  const double* dm;             /* dm(3, 4)               */
  const double* dh;             /* dh(ny+1)               */
  const double* dn;             /* dn(ny)                 */
  const double* hn;             /* hn(ny)                 */
  const double* iu;             /* iu(2, ny)              */
  const double* dyh;            /* dyh(ny+1)              */
  const double* dy;             /* dy(ny+1)               */
  const double* kx;             /* kx(nx/2+1)             */
  const double* kz;             /* kz(nz)                 */
  const double* buffer;         /* buffer(nx2, nz2)       */
  const double* target_inflow;  /* target_inflow(ny+1, 3) */
} chan_operators;

/* Context lifetime.  Replaces the allocation of data1/data2/data4 and the
 * FFTW plans (PFC/setup.f90:441-459, PFC/transform.f90:60-160). */
int chan_create(const chan_params* prm, chan_ctx** out);
int chan_destroy(chan_ctx* ctx);
const char* chan_last_error(const chan_ctx* ctx);
/* Version / capability string, e.g. "chanflow 1 gfx950 fp64". */
const char* chan_version(void);

/* Upload operators (see chan_operators). */
int chan_upload_operators(chan_ctx* ctx, const chan_operators* ops);
/* Build the channel operators on the host exactly as define_coordinates /
 * define_parameters do (tanh grid with the given stretching, which the
 * reference hard-codes as 2.75 at PFC/setup.f90:545) and upload them.  The
 * arrays are also returned when out_* is non-NULL (sizes as in
 * chan_operators; y has ny entries, yh ny+1). */
int chan_build_operators(chan_ctx* ctx, double stretch, double* out_y, double* out_yh);

/* State transfer in the reference's layouts (PFC/data.f90:140-141, shapes
 * PFC/setup.f90:447-449): data1 complex(nz, nx/2+1, ny+1, 3), data2
 * complex(ny+1, 3, nz, nx/2+1), data4 complex(ny+1, nz, nx/2+1).  NULL
 * pointers are skipped.  chan_get_state's data1 is the NL term between
 * chan_nonlinear and chan_rhs_solve, the velocity otherwise (as in the
 * reference, where nonlinear overwrites data1 in place). */
int chan_set_state(chan_ctx* ctx, const double* data1, const double* data2, const double* data4);
int chan_get_state(chan_ctx* ctx, double* data1, double* data2, double* data4);
/* Partial download of planes k0..k1 (1-based, inclusive) of data1 component
 * comp (1..3) into out(nz, nx/2+1, k1-k0+1). */
int chan_get_planes(chan_ctx* ctx, int comp, int k0, int k1, double* out);

/* Module scalars dt, t_total, p_grad and d_um_dy(ny) (PFC/runparms.f90:23,
 * PFC/gridparms.f90:90). NULL pointers are skipped. */
int chan_set_scalars(chan_ctx* ctx, const double* dt, const double* t_total, const double* p_grad,
                     const double* d_um_dy);
int chan_get_scalars(chan_ctx* ctx, double* dt, double* t_total, double* p_grad, double* d_um_dy);

/* Wall boundary data cbc_top/cbc_bot complex(nz, nx/2+1, 3) (PFC/gridparms.f90:113). */
int chan_set_wall_bc(chan_ctx* ctx, const double* cbc_top, const double* cbc_bot);

/* pass1::nonlinear (PFC/pass1.f90:26-252): data1 (velocity) -> data1 (NL
 * term).  At rk_step == 1 the CFL number is checked (abort -> error status)
 * and, with dt_flag == 1 and step >= 0, dt is adapted.  dt_inout and
 * cfl_out may be NULL. */
int chan_nonlinear(chan_ctx* ctx, int step, int rk_step, double* dt_inout, double* cfl_out);

/* pass1_linearized::nonlinear_linearized (PFC/pass1_linearized.f90:135-509)
 * about the base flow set by chan_set_baseflow. */
int chan_nonlinear_linearized(chan_ctx* ctx, int step, int rk_step, double* dt_inout, double* cfl_out);
/* baseflow real(nx2, nz2, ny+1, 3) (PFC/data.f90:149, PFC/pass1_linearized.f90:34). */
int chan_set_baseflow(chan_ctx* ctx, const double* baseflow);

/* pass2::rhs_solve (PFC/pass2.f90:29-180): per-(kx,kz) RK3/CN substep with
 * projection, pressure update and the constant-mass-flux mean mode.  Uses
 * the wall data of the last chan_set_wall_bc / chan_boundary call.
 * p_grad_out may be NULL. */
int chan_rhs_solve(chan_ctx* ctx, int step, int rk_step, double* p_grad_out);

/* One call = nsteps iterations of the timestep loop (PFC/main_mpi.f90:70-87):
 * boundary (zero wall data or feedback control), then three
 * (nonlinear | nonlinear_linearized, rhs_solve) substeps and t_total += dt,
 * all device-resident with no host synchronisation.  step0 is the step
 * number of the first iteration. linearized != 0 uses the base flow. */
int chan_step(chan_ctx* ctx, int step0, int nsteps, int linearized);

/* read_write::boundary (PFC/read_write.f90:1465-1532) on the device. */
int chan_boundary(chan_ctx* ctx);

/* init_vel::remove_initial_divergence (PFC/init_vel.f90:518-536) and the
 * d_um_dy initialisation (:333-342), then data2 = data1 and
 * compute_initial_pressure (:437-514). linearized != 0 uses the base flow. */
int chan_initialize_flow(chan_ctx* ctx, int linearized);
int chan_remove_divergence(chan_ctx* ctx);
int chan_compute_initial_pressure(chan_ctx* ctx, int linearized);

/* Host-pointer variants of the transform module (PFC/transform.f90:6-13)
 * for nplanes consecutive planes.  cdata complex(nz, nx/2+1, nplanes),
 * data real(nx2, nz2, nplanes). */
int chan_xz_fft_complex_to_real(chan_ctx* ctx, const double* cdata, double* data, int nplanes);
int chan_xz_fft_real_to_complex(chan_ctx* ctx, const double* data, double* cdata, int nplanes);
int chan_compute_divergence_of_fluxes(chan_ctx* ctx, const double* datax, const double* datay,
                                      const double* dataz, double* cdata, int nplanes);
/* No counterpart in the reference.  For callers that cannot allocate device
 * memory (code under HIP graph capture, tight serving loops, harnesses that
 * forbid hipMalloc): compute_divergence_of_fluxes on DEVICE arrays, with no
 * allocation, copy or synchronisation.  datax, datay, dataz real(nx2, nz2,
 * nplanes) and cdata complex(nz, nx/2+1, nplanes) are device pointers on the
 * context's GPU; stream is a hipStream_t (NULL: the context's stream).  The
 * x-transform intermediates live in a fixed per-GPU device array of
 * CHAN_NOALLOC_SCRATCH_BYTES, shared by every call on that GPU, so calls that
 * overlap on one GPU race; nplanes * 2 * nz2 * 16 * (nx/2+1 rounded up to 8)
 * bytes must fit (8 planes at nx2 = nz2 = 288), else an error is returned. */
#define CHAN_NOALLOC_SCRATCH_BYTES (8u << 20)
int chan_compute_divergence_of_fluxes_noalloc(chan_ctx* ctx, const double* datax, const double* datay,
                                              const double* dataz, double* cdata, int nplanes, void* stream);
/* small_fft_* (PFC/transform.f90:950-1000, 1071-1142): data real(nx, nz, nplanes). */
int chan_small_fft_complex_to_real(chan_ctx* ctx, const double* cdata, double* data, int nplanes);
int chan_small_fft_real_to_complex(chan_ctx* ctx, const double* data, double* cdata, int nplanes);

/* Host-pointer variants of the solve module (PFC/solve.f90:7-8), run on
 * the device for nsys independent systems sharing alpha-independent data.
 * solve_helmholtz: f complex(n, nsys) in/out, alpha(nsys), d2(3, n),
 * topbc/botbc complex(nsys), n in {ny, ny+1}. */
int chan_solve_helmholtz(chan_ctx* ctx, int n, int nsys, double* f, const double* alpha, double diff,
                         const double* d2, const double* topbc, const double* botbc, int top_neumann,
                         int bot_neumann);
/* solve_p: p complex(ny+1, nsys) in/out, alpha(nsys), zero wall data. */
int chan_solve_p(chan_ctx* ctx, int nsys, double* p, const double* alpha);
/* strip_divergence for every (kx,kz) of a field cu complex(nz, nx/2+1, ny+1, 3)
 * (data1 layout) with scale coef; cp complex(ny+1, nz, nx/2+1) receives the
 * pressure correction (may be NULL). */
int chan_strip_divergence(chan_ctx* ctx, double* cu, double* cp, double coef);
/* The per-mode forms of the same routines, for nsys independent (kx,kz)
 * columns at once.  p(nsys), q(nsys) are the reference's 1-based x- and
 * z-mode indices (x = 1..nx/2+1, z = 1..nz), so that a column is the
 * reference's cu(1:ny+1, 1:3) of mode (q, p).
 * strip_divergence (PFC/solve.f90:14-50): cu complex(ny+1, 3, nsys) in/out,
 *   cp complex(ny+1, nsys) out (may be NULL).  The mean mode (k^2 = 0) is
 *   rejected, as the reference's callers never pass it (its Neumann system
 *   is singular).
 * subtract_pressure_gradient (PFC/solve.f90:54-77): crhs complex(ny+1, 3,
 *   nsys) in/out, cpress complex(ny+1, nsys). */
int chan_strip_divergence_modes(chan_ctx* ctx, int nsys, double* cu, double* cp, double coef, const int* p,
                                const int* q);
int chan_subtract_pressure_gradient_modes(chan_ctx* ctx, int nsys, double* crhs, const double* cpress,
                                          double coef, const int* p, const int* q);

/* arnoldi::compute_time_deriv (PFC/arnoldi.f90:485-567): w = W P_T W^-1 v
 * with v, w real(nx, nz, ny+1, 3), P_T = tend linearized RK3 steps.  Needs
 * chan_set_baseflow.  t_total_out (may be NULL) receives T. */
int chan_linearized_matvec(chan_ctx* ctx, const double* v, double* w, int tend, double* t_total_out);

/* Diagnostics: max discrete divergence over non-mean modes, j = 2..ny
 * (PFC/pass2.f90:380-401); mass flux Q (PFC/pass2.f90:106-108). */
int chan_divergence_max(chan_ctx* ctx, double* out);
int chan_mass_flux(chan_ctx* ctx, double* out);

/* Per-kernel timing with HIP events on the context stream.
 * kinds: 0 z-inverse, 1 x-pass, 2 z-forward, 3 solve, 4 other. */
int chan_set_timing(chan_ctx* ctx, int enable);
int chan_get_kernel_times(chan_ctx* ctx, double* ms_total, long* counts, int nkinds);
/* Synchronise the context stream and report a deferred device error. */
int chan_sync(chan_ctx* ctx);

/* Diagnostics of the device velocity.
 * chan_set_coordinates: y(ny), the v-point coordinates (PFC/setup.f90:601-606),
 *   needed by chan_l2_norm when the operators were uploaded rather than built.
 * chan_l2_norm: l2_norm (PFC/read_write.f90:1432-1462) on the dealiased grid,
 *   l2 = sqrt(sum_planes,comps dy(j) sum_xz (u - lam)^2 / (nx2 nz2) / 2) with
 *   lam = 1 - y(j)^2 for u (the reference reads y(ny+1) one past the end of y;
 *   y(ny) is used for that plane), and drag = d_um_dy(1) - d_um_dy(ny).
 * chan_mode_energy: ke(nx/2+1, nz) = int_y(|u|^2 + |w|^2 + |v at yh|^2, dyh)
 *   per mode (PFC/pass2.f90:113-118).
 * chan_plane_moments: out(ny+1, 15) on the nx x nz grid, per plane: means of
 *   u, v, w, their 2nd, 3rd and 4th central moments and <u_i' v'>
 *   (compute_velocity_moments, PFC/stats.f90:423-440, for one field). */
int chan_set_coordinates(chan_ctx* ctx, const double* y);
int chan_l2_norm(chan_ctx* ctx, double* l2, double* drag);
int chan_mode_energy(chan_ctx* ctx, double* ke);
int chan_plane_moments(chan_ctx* ctx, double* out);

/* Wall jet-in-crossflow (configuration of the linearized/jet-control runs):
 * a Gaussian wall-normal velocity v = +-A exp(-r^2/radius^2) around (xjet,
 * zjet) on the nx x nz wall grid (periodic images; + at the bottom wall
 * (wall 0), - at the top (wall 1), i.e. blowing into the channel), with its
 * mean removed (a uniform suction over the wall, so no net mass flux), and
 * amplitude 1 + forcing sin(2 pi frequency t) at each substep's end time.
 * It overwrites the wall-normal wall velocity at every substep as
 * pass2::rhs_solve's jet_profile call does (PFC/pass2.f90:38-43; the
 * reference's jet is a tanh profile with random-walk forcing and suction in a
 * buffer region, PFC/pass2.f90:423-673).  Combine with the feedback gains
 * (chan_params ctrlu / ctrlw) for controlled runs.  amplitude = 0 removes the
 * jet. */
int chan_jet_configure(chan_ctx* ctx, int wall, double amplitude, double radius, double xjet, double zjet,
                       double forcing, double frequency);

/* Launch-shape tuning of one context (no counterpart in the reference: the
 * reference's only tuning is FFTW_PATIENT planning, PFC/transform.f90:60-160).
 * Keys: 0 x-pass kernel (1: one wavefront per z-row, 0: workgroup per rows),
 * 1 x-pass workgroup thread bound (64..640), 2 z-forward kernel (1: A and B
 * transformed in turn, 0: together), 3 solve lanes per mode (3 or 1),
 * 4 solve rows per prefetch block (4, 8 or 16), 5 z-inverse kx columns per
 * workgroup (4, 8 or 16), 6 z-forward kx columns per workgroup (4 or 8),
 * 7 x-pass prefetch of the next plane's rows into LDS (1 on, 0 off),
 * 8 solve rows streamed through an LDS ring (1 on, 0 off; 3-lane solve only),
 * 9 planes per x-pass workgroup (2..64; 0, the default: the plane-rows of the
 * pass spread evenly over the workgroups resident at a time), 10 plane
 * ranges the nonlinear pass runs in turn (1..64; 1: every kernel over all
 * planes), 11 segmented column solve (1 on with 4 modes per wavefront, 2
 * with 2 modes per wavefront, 3 with 8 modes per wavefront, 0 off) wherever
 * ny+1 rows fit 16 (32, 8) segments of 5 to 9 (5, 5 to 9) rows (keys 3, 4
 * and 8 then do not apply; value 3 falls back to 1 where 8 segments are too
 * few), 12 z-rows per
 * wavefront of the wavefront x-pass (0, the default: by length; 1; 2 for
 * nx2/2 <= 72; 4 for nx2/2 <= 24; with the prefetch, where nz2 is a multiple;
 * otherwise the default applies; a batched ensemble takes the fewest rows its
 * members of one length choose).
 * chan_autotune times `steps` timesteps of every candidate (keys 0 and 2,
 * key 0 fixed at 1 while a subgrid closure is configured; with the wavefront
 * x-pass, key 12's row counts the length has; the
 * segmented solve's layouts of key 11 that ny + 1 rows fit, or the
 * sweep-based solves with keys 3, 4 and 8), keeps the fastest whose trial
 * passed the CFL check and restores the state. */
int chan_set_tuning(chan_ctx* ctx, int key, int value);
int chan_get_tuning(chan_ctx* ctx, int key, int* value);
int chan_autotune(chan_ctx* ctx, int steps, double* best_ms_per_step);

/* Batched ensemble (no counterpart in the reference): independent channel
 * contexts of any sizes and parameters, each set up as usual (operators,
 * state, chan_initialize_flow), advanced together.  One timestep of every
 * member is captured once into a HIP graph; chan_ensemble_step replays it
 * nsteps times and then reports any member's CFL abort.
 * nstreams = 0: batched launches.  Each phase of each RK substage (z-inverse,
 *   x-pass, z-forward, segmented solve; CFL/dt, wall-data and time updates)
 *   is one launch per transform (or segment) length covering the work lists
 *   of every member of that length, whatever their sizes.  Members need the
 *   default launch shapes, no LES closure, no jet and zero feedback gains;
 *   lengths nz2, nx2/2 with compile-time plans and ny+1 rows that fit the
 *   segmented solve.
 * nstreams >= 1: each member's own kernels, the members spread over nstreams
 *   concurrent streams.  Changing a member's tuning or gains
 * needs chan_ensemble_invalidate (the next step captures again).  Members
 * must not be used on their own while an ensemble step runs. */
typedef struct chan_ensemble chan_ensemble;
int chan_ensemble_create(int n, chan_ctx* const* members, int nstreams, int linearized, chan_ensemble** out);
int chan_ensemble_step(chan_ensemble* ens, int nsteps);
int chan_ensemble_invalidate(chan_ensemble* ens);
/* Launch-plan tuning of a batched ensemble (nstreams = 0; no counterpart in
 * the reference): times `steps` replayed steps of each x-pass plane-range
 * plan (each length's planes cut for 1, 2, 4 or 8 times, or as many times as
 * there are lengths, its resident workgroups), then, with the fastest ranges,
 * the tile shapes (one z-row per x-pass wavefront at every length instead of
 * the batched default; 4 modes per solve wavefront everywhere instead of 8
 * where all 8 segments are busy), restores every member's state after each
 * trial, keeps the fastest plan and returns its time per step (on failure
 * the state is restored as well and the previous plan kept).  The ranges
 * give every member the same arithmetic; the tile shapes the same to
 * round-off. */
int chan_ensemble_autotune(chan_ensemble* ens, int steps, double* best_ms_per_step);
int chan_ensemble_destroy(chan_ensemble* ens);
const char* chan_ensemble_last_error(const chan_ensemble* ens);

/* Learned subgrid closure (no counterpart in the reference, whose subgrid
 * model is absent; only its flags remain, PFC/flow.dat:19-22,
 * PFC/main_mpi.f90:10-13,74,77-78): an MLP 10 -> 128 -> 128 -> 128 -> 6 with
 * GELU hidden layers mapping (du_i/dx_j, y+) to (tau_11, tau_12, tau_13,
 * tau_22, tau_23, tau_33).  Parameters are float32, chan_mlp_num_params() of
 * them in the order W1 (128x10), b1, W2 (128x128), b2, W3, b3, W4 (6x128), b4,
 * weights row-major (out, in).  x is float32 (n, 10), y and t (n, 6), row-major.
 * chan_mlp_infer: bfloat16 matrix-core inference with float32 accumulation.
 * chan_mlp_forward_fp32 / chan_mlp_gradient: float32 forward pass and the
 *   gradient of the mean squared error against t (and the loss).
 * chan_mlp_train: mini-batch Adam (lr, beta 0.9/0.999, eps 1e-8), batches
 *   taken in order, the incomplete last batch dropped; loss of the last batch.
 * chan_mlp_set_params resets the Adam moments. */
typedef struct chan_mlp chan_mlp;
int chan_mlp_create(chan_mlp** out);
int chan_mlp_destroy(chan_mlp* net);
const char* chan_mlp_last_error(const chan_mlp* net);
int chan_mlp_num_params(void);
int chan_mlp_set_params(chan_mlp* net, const float* params);
int chan_mlp_get_params(chan_mlp* net, float* params);
int chan_mlp_infer(chan_mlp* net, int n, const float* x, float* y);
int chan_mlp_forward_fp32(chan_mlp* net, int n, const float* x, float* y);
int chan_mlp_gradient(chan_mlp* net, int n, const float* x, const float* t, float* grad, double* loss);
int chan_mlp_train(chan_mlp* net, int n, const float* x, const float* t, int epochs, int batch, double lr,
                   double* loss_out);

/* Subgrid-scale closure for large-eddy simulation (no counterpart in the
 * reference beyond its unused switches subgrid / filter_dim / lagr_avg,
 * PFC/flow.dat:19-22, and commented-out calls of a dynamic model,
 * PFC/main_mpi.f90:10-13,74,77-78).  With a closure configured, every
 * chan_nonlinear (and chan_step) adds the subgrid stress tau_ij to the flux
 * u_i u_j, so the nonlinear term becomes div(u u + tau) with the discrete
 * operators of PFC/pass1.f90:138-195; tau_12, tau_23 are interpolated to the
 * v-points with iu.  tau lives at the cell centres yh of the dealiased
 * physical grid, layout double [6][ny+1][nz2][nx2] (components 11, 12, 13,
 * 22, 23, 33), zero on the wall planes.
 * Modes: 0 off; 1 dynamic Smagorinsky (Germano identity, sharp test filter at
 * half the resolved kx and kz, plane least squares, clipped at 0; C(y) is
 * evaluated at the first substep of each time step); 2 learned
 * closure `net` with a fallback to dynamic Smagorinsky wherever the learned
 * stress's energy transfer -tau:S is below -beta times Smagorinsky's or not
 * finite (beta < 0: only the finiteness check); 3 prescribed tau
 * (chan_les_set_tau).  With every_substep = 1 the
 * stress is evaluated at every substep, with 0 at the first substep of each
 * step and held for the other two.  Needs chan_set_coordinates.
 * chan_les_update evaluates the closure on the current state;
 * chan_les_get_coefficient returns C(y) [ny+1]; chan_les_get_fields the
 * physical closure inputs [12][ny+1][nz2][nx2] (du_i/dx_j row-major, then u,
 * v, w at yh); chan_les_features the network inputs [(ny-1) nz2 nx2][10] and
 * the stress scale D^2 |du/dx|^2 per point (mode 2); chan_les_gated_points the
 * number of points of the last evaluation that fell back, and the total. */
int chan_les_configure(chan_ctx* ctx, int mode, chan_mlp* net, double beta, int every_substep);
int chan_les_update(chan_ctx* ctx);
int chan_les_set_tau(chan_ctx* ctx, const double* tau);
int chan_les_get_tau(chan_ctx* ctx, double* tau);
int chan_les_get_coefficient(chan_ctx* ctx, double* cs2);
int chan_les_get_fields(chan_ctx* ctx, double* fields);
int chan_les_features(chan_ctx* ctx, float* x, double* scale);
int chan_les_gated_points(chan_ctx* ctx, long* count, long* total);
/* Filtered DNS (no counterpart in the reference): the velocity of the fine
 * channel dns filtered onto the coarse channel les, which must have the same
 * box and every other face of the fine grid (ny_dns = 2 ny_les - 1, both with
 * their y coordinates set): a 3-point box average on the fine grid in x and z,
 * the cut-off to the coarse modes (coarse Nyquist modes dropped) and in y
 * the height-weighted average of the two fine cells of a coarse cell (u, w)
 * or the shared faces (v), which keeps the filtered field divergence-free.
 * The result becomes les's velocity.  tau (may be NULL) receives the exact
 * subgrid stress F(u_i u_j) - F(u_i) F(u_j) at the coarse cell centres,
 * [6][ny+1][nz2][nx2] of the coarse grid (components 11, 12, 13, 22, 23, 33,
 * products on the fine dealiased grid, zero on the wall planes). */
int chan_les_filter_dns(chan_ctx* les, chan_ctx* dns, double* tau);

/* Legacy VTK output as output_hdf_dfsd writes it with xskip = yskip = zskip = 1
 * (PFC/read_write.f90:296-500): big-endian RECTILINEAR_GRID, nx x (ny+1) x nz
 * points (x = lx i/nx, y = yh, z = lz k/nz), float32 vectors "flowvec".
 * chan_vtk_write (no context, no GPU) formats phys real(nx, nz, ny+1, 3),
 *   i.e. [3][ny+1][nz][nx], with u, v, w on the yh planes; errors in
 *   chan_io_last_error.
 * chan_write_vtk takes the device velocity: u and w at plane j, v averaged
 *   to yh (v(1) at the bottom plane, v(ny) at the top) in spectral space, then
 *   the small transform to the nx x nz grid; yh(ny+1) is needed for the
 *   coordinates. */
int chan_vtk_write(const char* path, int nx, int ny, int nz, double lx, double lz, const double* yh,
                   const double* phys);
int chan_write_vtk(chan_ctx* ctx, const char* path, const double* yh);

/* Velocity checkpoints in the reference's savevel.out / newvel.out format:
 * read_velocity and SAVE_VELOCITY (PFC/read_write.f90:21-176, 180-292),
 * Fortran sequential unformatted records with 4-byte length markers: one
 * header record NX, NY, NZ (int32), LX, LZ, RE, DT, T_TOTAL (float64), then
 * one record per x-mode x = 1..nx/2+1 holding data1(1:nz, x, 1:ny+1, 1:3).
 * The chan_savevel_* functions need no context and no GPU; their errors are
 * reported by chan_io_last_error. */
const char* chan_io_last_error(void);
int chan_savevel_write(const char* path, int nx, int ny, int nz, double lx, double lz, double re, double dt,
                       double t_total, const double* data1);
int chan_savevel_read_header(const char* path, int* nx, int* ny, int* nz, double* lx, double* lz, double* re,
                             double* dt, double* t_total);
int chan_savevel_read(const char* path, int nx, int ny, int nz, double* data1);
/* SAVE_VELOCITY of the device velocity, dt and t_total; read_velocity into
 * the device: NX, NY, NZ, LX, LZ must match the context (else
 * "INCONSISTENT PARAMETERS"), t_total is taken from the file and dt only
 * when dt_flag == 1 (PFC/read_write.f90:36-63). */
int chan_save_velocity(chan_ctx* ctx, const char* path);
int chan_read_velocity(chan_ctx* ctx, const char* path);

#ifdef __cplusplus
}
#endif
#endif /* CHANFLOW_H */
