// Internal declarations shared by the HIP translation units of libchanflow.
#pragma once
#include <hip/hip_runtime.h>
#include <string>
#include <vector>
#include "fft_device.h"
#include "../../include/chanflow.h"

#define CHAN_CHECK(x)                                                              \
  do {                                                                             \
    hipError_t e_ = (x);                                                           \
    if (e_ != hipSuccess) {                                                        \
      throw chan_error(std::string(#x) + ": " + hipGetErrorString(e_));            \
    }                                                                              \
  } while (0)

#define LAUNCH_CHECK(name)                                                         \
  do {                                                                             \
    hipError_t e_ = hipGetLastError();                                             \
    if (e_ != hipSuccess) throw chan_error(std::string(name) + ": " + hipGetErrorString(e_)); \
  } while (0)

struct chan_error {
  std::string msg;
  explicit chan_error(std::string m) : msg(std::move(m)) {}
};

// Geometry of one dealiased (or small) x-z transform family.
struct XZGeom {
  int nx, nz, nxh;    // spectral sizes: nxh = nx/2+1 stored kx, nz kz
  int nx2, nz2;       // physical sizes
  int kxpad;          // row length (complex) of the [zz][kx] intermediate
  int small;          // 1: "small" transforms (no padding, split, fold or doubling)
  FFTPlan pz;         // complex length nz2
  FFTPlan px;         // complex length nx2/2 (half-length real transform)
  int twx;            // root-table offset for length nx2 (real pre/post twiddles)
};

// Device-resident scalars of one channel instance.
struct DevScalars {
  double dt;
  double p_grad;
  double t_total;
  double cfl;          // last max CFL (rk_step==1)
  int err;             // nonzero: CFL abort (1: CFL IS TOO BIG, 2: DT IS TOO BIG)
  int pad;
};

// Load of data that no kernel writes while it runs, through the constant
// address space: with a wavefront-uniform index it becomes a scalar load,
// which is counted by lgkmcnt and so never waits on vector loads in flight.
template <class T>
__device__ __forceinline__ T ldro(const T* p, long i) {
  return ((const __attribute__((address_space(4))) T*)p)[i];
}

// Operators uploaded from the driver (layouts of the reference's globals).
struct DevOps {
  const double* d2h;   // d2h(3, ny+1)
  const double* d2n;   // d2n(3, ny)
  const double* dm;    // dm(3, 4)
  const double* dh;    // dh(ny+1)
  const double* dn;    // dn(ny)
  const double* hn;    // hn(ny)
  const double* iu;    // iu(2, ny)
  const double* dyh;   // dyh(ny+1)
  const double* dy;    // dy(ny+1)
  const double* kx;    // kx(nx/2+1)
  const double* kz;    // kz(nz)
  const double* buffer;   // buffer(nx2, nz2) or nullptr when identically zero
  const double* target;   // target_inflow(ny+1, 3)
};

// Parameters for one nonlinear x-pass launch.
struct XPassArgs {
  const double2* Z;       // [3][ny+1][nz2][kxpad]
  double2* AB;            // [3][ny+1][2][nz2][kxpad]
  const double* base;     // linearized base flow [3][ny+1][nz2][nx2] or nullptr
  unsigned long long* cflbits;  // [ny+1][3] or nullptr
  int ny;
  int ka, kb;             // reference planes computed: NL(k) for ka <= k <= kb
  int chunk;              // planes per workgroup
  int rz;                 // rows per workgroup
  int wave;               // 1: wave-per-row kernel where the length allows it
  int maxthr;             // workgroup-kernel threads bound (64..640)
  int prefetch;           // 1: direct global-to-LDS prefetch of the next plane's rows
  const double* tau;      // subgrid stress [6][ny+1][nz2][nx2] at the cell centres yh
                          // (11, 12, 13, 22, 23, 33; zero on the wall planes) or nullptr
  int rows;               // z-rows per wavefront of the wave kernel (0: by length; 1, 2, 4)
};

// LDS cells of one raw row of the wave x-pass's prefetch: nxh rounded up to
// the 64 cells one direct global-to-LDS load writes.
__host__ __device__ inline int xpass_raw_cells(int nxh) { return ((nxh + 63) / 64) * 64; }

// Plane of a z-transform workgroup: grid row y covers planes off .. off+cnt-1
// of each field, fields `stride` planes apart (identity: cnt = nplanes, stride = 0).
struct PlaneMap {
  int cnt, stride, off;
};

// Batched launches (the batched ensemble): workgroup i of a launch runs entry
// i of a work list, the workgroup (bx, by) of channel m's own grid.
struct BatchItem {
  unsigned short m, bx;
  unsigned int by;
};
// Per-channel arguments of a batched z-transform: z-inverse in -> out;
// z-forward A = in, B -> out (+ i kz B), in_ps the plane stride of A and B.
struct ZBatchArgs {
  XZGeom g;
  const double2* tw;
  const double2* in;
  const double2* B;
  long in_ps;
  double2* out;
  long out_ps;
  const double* kz;
  PlaneMap pm;
  int nl_zero_ny;
};

// Per-channel arguments of a batched wavefront x-pass.
struct XBatchArgs {
  XZGeom g;
  const double2* tw;
  DevOps ops;
  XPassArgs a;
};

// Launch-shape choices of one context (chan_set_tuning, chan_autotune);
// the defaults can be overridden with CHANFLOW_XPASS_WAVE, CHANFLOW_XPASS_THREADS,
// CHANFLOW_ZFWD_V2, CHANFLOW_SOLVE_LANES and CHANFLOW_SOLVE_BL.
struct Tuning {
  int xpass_wave = 1;     // wave-per-row x-pass (lengths <= 192)
  int xpass_threads = 640;
  int zfwd_v2 = 1;        // sequential-A/B z-forward kernel
  int solve_lanes = 3;    // 3: component-parallel solve, 1: one lane per mode
  int solve_bl = 8;       // rows per prefetch block (4, 8, 16)
  int zinv_kx = 8;        // z-inverse kx columns per workgroup (4, 8, 16)
  int zfwd_kx = 8;        // z-forward kx columns per workgroup (4, 8)
  int xpass_prefetch = 1; // wave x-pass streams the next plane's rows into LDS
  int solve_dma = 1;      // 3-lane solve streams its rows through an LDS ring
  int xpass_chunk = 0;    // planes per x-pass workgroup (2..64; 0: from the resident workgroups)
  int nl_chunks = 1;      // plane ranges the NL pass runs in turn (1: all planes per kernel)
  int solve_seg = 1;      // segmented column solve (columns in registers, no scratch); 2: two modes per
                          // wavefront; 3: eight modes per wavefront where ny+1 rows fit 8 segments
  int xpass_rows = 0;     // z-rows per wavefront of the wave x-pass (0: by length, 1, 2, 4)
};
Tuning tuning_defaults();
bool tuning_set(Tuning& t, int key, int value);
int tuning_get(const Tuning& t, int key);

struct LesState;

struct chan_ctx {
  chan_params prm;
  int nxh, nyp, nmodes, kxpad;
  XZGeom gd;     // dealiased transforms
  XZGeom gs;     // small transforms (nx x nz grid)
  hipStream_t stream = nullptr;
  bool own_stream = false;
  int device = 0;                // the HIP device the context's memory lives on
  std::string last_error;

  // twiddle roots
  double2* d_tw = nullptr;
  // operators
  double* d_ops = nullptr;  // packed
  DevOps ops{};
  double h_dm[12] = {0};         // host copy of dm(3,4) for the feedback law
  std::vector<double> h_y;       // y(ny) (v-point coordinates), for l2_norm
  bool has_buffer = false;
  // state (plane-major, kz fastest): vel and nl are [3][ny+1][nxh][nz]
  double2* d_vel = nullptr;
  double2* d_nl = nullptr;
  double2* d_R = nullptr;
  double2* d_P = nullptr;        // [ny+1][nxh][nz]
  double2* d_cbc = nullptr;      // [2][3][nxh][nz]: top then bottom
  double* d_dudy = nullptr;      // d_um_dy(ny)
  DevScalars* d_sc = nullptr;
  unsigned long long* d_cfl = nullptr;  // [ny+1][3]
  // intermediates
  double2* d_Z = nullptr;
  double2* d_AB = nullptr;
  double2* d_scr = nullptr;      // solve scratch: 4 complex + 2 real per (row, mode)
  double* d_base = nullptr;      // linearized base flow (physical)
  // generic staging buffers for host-pointer entry points
  void* d_stage = nullptr;
  size_t stage_bytes = 0;
  void* h_pin = nullptr;         // pinned host staging (hipHostMalloc)
  size_t pin_bytes = 0;
  bool data1_is_nl = false;      // Fortran's data1 currently holds the NL term
  // The device dt as last set or read by the host; h_dt_known is cleared by
  // every entry point that may change dt on the device without reading it back.
  double h_dt = 0.0;
  bool h_dt_known = false;
  LesState* les = nullptr;       // subgrid closure (chan_les_configure)
  double2* d_jet = nullptr;      // wall jet: spectral v profile [nxh][nz] (chan_jet_configure)
  int jet_wall = 0;              // 0 bottom, 1 top
  double jet_forcing = 0.0, jet_freq = 0.0;
  // options
  Tuning tune;
  int timing = 0;
  std::vector<hipEvent_t> ev;    // timing events (pairs)
  std::vector<int> ev_kind;
  double ktime_ms[8] = {0};
  long kcount[8] = {0};
};

// kernel-kind indices for timing
enum { KT_ZINV = 0, KT_XPASS, KT_ZFWD, KT_SOLVE, KT_OTHER, KT_N };

// plan helpers (host)
FFTPlan make_plan(int n, int tw_offset);
void xz_geom_init(XZGeom& g, int nx, int nz, int nx2, int nz2, int small, std::vector<double2>& roots);

// launchers (k_transform.hip)
// pcnt > 0: grid row y is plane poff + y % pcnt of field y / pcnt, fields
// pstride planes apart (a range of planes of every field); 0: plane y
void launch_zinv(const XZGeom& g, const double2* tw, const double2* in, long in_ps, double2* out,
                 long out_ps, int nplanes, hipStream_t s, int kx = 8, int pcnt = 0, int pstride = 0,
                 int poff = 0);
void launch_xc2r(const XZGeom& g, const double2* tw, const double2* Z, long z_ps, double* out,
                 long out_ps, int nplanes, hipStream_t s);
void launch_xr2c(const XZGeom& g, const double2* tw, const double* fy, const double* fx,
                 const double* kx, long in_ps, double2* out, long out_ps, int nplanes, hipStream_t s);
void launch_zfwd(const XZGeom& g, const double2* tw, const double2* A, const double2* B, long ab_ps,
                 const double* kz, double2* out, long out_ps, int nplanes, int nl_zero_ny,
                 hipStream_t s, int v2 = 1, int kx = 8, int pcnt = 0, int pstride = 0, int poff = 0);
bool zt_batchable(int nz2);
void launch_zinv_batched(int nz2, const ZBatchArgs* args, const BatchItem* items, int nitems, hipStream_t s);
void launch_zfwd_batched(int nz2, const ZBatchArgs* args, const BatchItem* items, int nitems, hipStream_t s);
// (k_nonlinear.hip)
void launch_xpass(const XZGeom& g, const double2* tw, const DevOps& ops, const XPassArgs& a,
                  hipStream_t s);
bool xpass_batchable(int M);
// z-rows per wavefront of the wave x-pass: `req` (1, 2, 4) where the length
// has that kernel and nz2 is a multiple of it, otherwise the default of a
// single or a batched launch
int xpass_rows(int M, int nz2, int req, bool pf, bool batched);
void launch_xpass_batched(int M, bool lin, int rw, int nxh, const XBatchArgs* args, const BatchItem* items,
                          int nitems, hipStream_t s);
long xpass_batched_resident(int M, bool lin, int rw, int nxh);
// (k_solve.hip)
struct SolveArgs {
  const double2* nl;  // [3][ny+1][nxh][nz]
  double2* vel;       // output velocity
  double2* R;
  double2* P;
  const double2* cbc; // [2][3][nxh][nz]
  double* dudy;
  DevScalars* sc;
  double2* scr;
  int rk;             // 1..3
  int ny, nxh, nz;
  double re, cosb, sinb;
  int fix_mass;
  int lanes, bl;      // kernel variant (Tuning::solve_lanes, solve_bl)
  int dma;            // 3-lane variant: rows streamed through an LDS ring (Tuning::solve_dma)
  int seg;            // segmented column solve where the row count allows (Tuning::solve_seg)
};
void launch_solve(const DevOps& ops, const SolveArgs& a, hipStream_t s);
// batched forms (the batched ensemble)
struct SBatchArgs {
  DevOps ops;
  SolveArgs a;
};
struct CflBatchArgs {
  DevOps ops;
  chan_params prm;
  unsigned long long* cfl;
  DevScalars* sc;
};
struct FillBatchArgs {
  unsigned long long* p;
  size_t n;   // 8-byte words
};
void launch_solve_batched(int layout, const SBatchArgs* args, const BatchItem* items, int nitems, hipStream_t s);
// Batched segmented-solve layout of a channel: 16 * modes per wavefront + rows
// per segment (8 modes where ny+1 rows fit 8 segments, all of them busy for
// solve_seg 1, any for solve_seg 3; else 4), 0 if ny+1 rows fit neither
int solve_seg_layout(int ny, int solve_seg);
void launch_cfl_finalize_batched(const CflBatchArgs* args, int n, int step, hipStream_t s);
void launch_fill_zero_batched(const FillBatchArgs* args, int n, size_t max_words, hipStream_t s);
void launch_advance_time_batched(DevScalars* const* sc, int n, hipStream_t s);
int solve_seg_rows(int ny, int maxsegs = 16);  // rows per segment of the segmented solve, 0 if ny+1 rows do not fit
void launch_jet_bc(const DevScalars* sc, int rk, double forcing, double freq, const double2* prof, double2* cbcv,
                   long nm, hipStream_t s);
void launch_l2_planes(const double* phys, long ps, int nplanes, const double* lam, const double* dy, double* part,
                      hipStream_t s);
void launch_l2_final(const double* part, int n, double npts, const double* dudy, int ny, double* out, hipStream_t s);
void launch_mode_energy(const DevOps& ops, int ny, int nxh, int nz, const double2* vel, double* ke, hipStream_t s);
void launch_plane_moments(const double* phys, long ps, long cs, int nyp, double* out, hipStream_t s);
void launch_vtk_field(int ny, long nm, const double2* vel, double2* out, hipStream_t s);
// cols == nullptr: every (kx, kz) of a [3][ny+1][nxh][nz] field; else ncols
// columns complex(ny+1, 3, ncols) with 0-based mode indices cols[2m], cols[2m+1]
void launch_project(const DevOps& ops, int ny, int nxh, int nz, const double2* in, double2* out_vel,
                    double2* out_p, double2* scr, double coef, int mode, double2* vel_for_visc,
                    double re, hipStream_t s, const int* cols = nullptr, int ncols = 0);
void launch_subtract_grad_cols(const DevOps& ops, int ny, double2* crhs, const double2* cpress, double coef,
                               const int* cols, int ncols, hipStream_t s);
void launch_cfl_finalize(const DevOps& ops, chan_params prm, unsigned long long* cfl, DevScalars* sc,
                         int step, hipStream_t s);
void launch_fill_zero(void* p, size_t bytes, hipStream_t s);
void launch_divmax(const DevOps& ops, int ny, int nxh, int nz, const double2* vel,
                   unsigned long long* out, hipStream_t s);
void launch_dudy(const DevOps& ops, int ny, int nxh, int nz, const double2* vel, double* dudy,
                 hipStream_t s);
void launch_layout(int kind, int ny, int nxh, int nz, const double2* in, double2* out, hipStream_t s);
// (k_misc.hip)
void launch_helmholtz_batch(int n, int nsys, double2* f, const double* alpha, double diff,
                            const double* d2, const double2* topbc, const double2* botbc, double* scr,
                            hipStream_t s);
void launch_solvep_batch(int ny, int nsys, double2* p, const double* alpha, const DevOps& ops,
                         double* scr, hipStream_t s);
// out = sgn (ca p2 + cb p3 + add) / den, pointwise over npts
struct CtrlLaw {
  const double* p2;
  const double* p3;
  double ca, cb, add, sgn, den;
  double* out;
};
void launch_ctrl_law(int npts, const CtrlLaw& w0, const CtrlLaw& w1, hipStream_t s);
// The feedback law of one velocity component applied to its Fourier
// coefficients: bottom (sgn +1) from planes 1 and 2, top (sgn -1) from planes
// ny-1 and ny-2, each out = P(sgn (ca p2 + cb p3) / den) + sgn add / den at
// the mean mode, where P keeps the Hermitian part of the kx = 0 column (what
// the physical-space round trip of a dealiased grid leaves of a spectrum)
struct CtrlSpec {
  const double2 *p2b, *p3b, *p2t, *p3t;
  double cab, cbb, addb, denb;
  double cat, cbt, addt, dent;
  double2 *outb, *outt;
};
void launch_ctrl_spec(long nm, int nz, const CtrlSpec& c0, const CtrlSpec& c1, int ncomp, hipStream_t s);
void launch_absmax(long n, const double* x, unsigned long long* out, hipStream_t s);
void launch_arnoldi_scale(int ny, long plane, double2* x, const double* dy, const double* dyh,
                          const unsigned long long* vmax, double eps, int dir, hipStream_t s);

// Error capture shared by the entry points of every translation unit: the
// message goes to the context (chan_last_error) and the status is 1.
void ctx_set_error(chan_ctx* c, const std::string& m);
template <class F>
int guarded_les(chan_ctx* c, F&& f) {
  if (!c) return 1;
  try {
    f();
    return 0;
  } catch (const chan_error& e) {
    ctx_set_error(c, e.msg);
  } catch (const std::exception& e) {
    ctx_set_error(c, e.what());
  }
  return 1;
}
void check_ops_les(chan_ctx* c);        // operators uploaded, else throws
void les_prepare_state(chan_ctx* c);    // the velocity (not the NL term) is in d_vel

// filtered DNS (k_filter.hip): the coarse velocity and, when tau_host is
// given, the exact subgrid stress [6][ny_c+1][nz2_c][nx2_c] on the host
void filter_dns_device(chan_ctx* les, chan_ctx* dns, double* tau_host,
                       void (*c2r)(chan_ctx*, const XZGeom&, const double2*, int, double*),
                       void (*r2c)(chan_ctx*, const XZGeom&, const double*, int, double2*));

// subgrid closure (k_les.hip)
void les_destroy(LesState* s);
const double* les_tau_for_pass(chan_ctx* c, int rk);  // updates and returns tau, nullptr when off
void les_prepare_capture(chan_ctx* c);  // everything the closure may synchronise for, before a capture

// learned closure network (k_mlp.hip)
void mlp_prepare_device(chan_mlp* m);   // packs the bf16 weights, synchronised
void mlp_infer_device(chan_mlp* m, long n, const float* x, float* y, hipStream_t s);
