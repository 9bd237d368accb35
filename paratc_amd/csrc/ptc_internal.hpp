// ptc_internal.hpp -- shared definitions of the gfx950 HIP library.
//
// Device layout of every 3-D field (one context, one case):
//   index(i, j, k) = i + nx * (j + (ny + 2) * k)
//   i = 0..nx-1   (reference ic = i + 1), periodic: no x halo is stored
//   j = 0..ny+1   (reference jc = j), rows 0 and ny+1 are the wall ghost rows
//   k = 0..nz-1   (reference kc = k + 1), periodic: no z halo is stored
// Kernels wrap x and z neighbours with a conditional add, so the reference's
// periodic halo refreshes (myupdate_halo, src/m_Decomp2d.f90:1554-1709) are
// implicit and never read stale data.  Interior-only reference arrays use
// the same layout and ignore their ghost rows.
#pragma once

#include <hip/hip_runtime.h>

#include <cmath>
#include <cstddef>
#include <cstdint>
#include <string>
#include <vector>

#include "../../include/paratc_gpu.h"

namespace ptc {

void set_error(const std::string &msg);

#define PTC_HIP(expr)                                                                  \
  do {                                                                                 \
    hipError_t e_ = (expr);                                                            \
    if (e_ != hipSuccess) {                                                            \
      ::ptc::set_error(std::string(#expr) + ": " + hipGetErrorString(e_));            \
      return PTC_E_HIP;                                                                \
    }                                                                                  \
  } while (0)

#define PTC_CHECK(cond, code, msg)                                                     \
  do {                                                                                 \
    if (!(cond)) {                                                                     \
      ::ptc::set_error(msg);                                                           \
      return (code);                                                                   \
    }                                                                                  \
  } while (0)

#define PTC_TRY(expr)                                                                  \
  do {                                                                                 \
    int r_ = (expr);                                                                   \
    if (r_ != PTC_OK) return r_;                                                       \
  } while (0)

// metric table rows in the device array d_met[(row) * (ny+2) + j]
enum Met {
  M_YC = 0, M_DYP, M_DYC, M_RDYP, M_RDYC, M_YINT,
  M_AM2PH, M_AC2PH, M_AP2PH, M_AM2C, M_AC2C, M_AP2C,
  M_AM2CCN, M_AC2CCN, M_AP2CCN, M_AM2P, M_AC2P, M_AP2P,
  M_AM2CSC, M_AC2CSC, M_AP2CSC, M_AM2CSCCN, M_AC2CSCCN, M_AP2CSCCN,
  M_NMET
};

// Geometry and stencil constants passed by value to every stencil kernel
// (the scalars of m_MeshAndMetries, src/m_MeshAndMetries.f90:10-13,48-52).
struct Geo {
  int nx, ny, nz;
  long long plane;  // nx * (ny + 2)
  double rdx, rdz;
  double dxc1, dxc2, dxc3, dxc4, dzc1, dzc2, dzc3, dzc4;
  double dxxc1, dxxc2, dxxc3, dxxc4, dxxc5, dzzc1, dzzc2, dzzc3, dzzc4, dzzc5;
  double ic1, ic2, ic3, ic4;
  const double *met;  // d_met
  int nyh;            // ny + 2
};

// One complex FFT length: radix plan and twiddle table w^m = exp(-2 pi i m / n).
struct FftLen {
  int n;
  int nstage;
  int radix[16];
  const double2 *tw;
};

constexpr int kMaxRadix = 8;

struct PoissonPlan {
  bool ready = false;
  int prec = PTC_PREC_FP64;
  int tri_req = PTC_TRI_AUTO;      // requested variant
  int tri = PTC_TRI_PARTITIONED;   // variant in use
  int tri_mm_req = 0;              // requested chunk rows (0 = automatic)
  int slab_req = 0;                // requested planes per x/z slab (0 = whole field)
  int slab = 0;                    // planes per slab in use (0 = unblocked)
  int sweeps = 2;
  int h = 0;          // nx / 2
  int pitch = 0;      // complex row pitch of the spectral buffer (>= h + 1)
  int ztile_req = 0;  // requested z tile: 0 = automatic (specialised kernels where built)
  int ztile = 4;      // columns per z-FFT workgroup in use
  int zfast_req = 0;  // column tile of the specialised z kernel (0 = its default)
  bool x_fast = false, z_fast = false;  // compile-time specialised FFT kernels in use
  int z_threads = 256;
  int x_lines = 4;    // lines per x-FFT workgroup
  int x_threads = 256;
  int tri_chunks = 1; // chunks (waves) per line in the partitioned solver
  int tri_mmax = 16;  // register rows per thread (chunk length)
  int tri_lpb = 64;   // lines per workgroup
  int tri_lpb_req = 0; // register y-solve: real lines per workgroup requested (0 = automatic)
  int tri_ll_req = 0; // LDS-staged y-solve: real lines per workgroup requested (0 = register kernel)
  int tri_ll = 0;     // in use
  int tri_cl_req = 0; // complex-line y-solve lines per workgroup (0 = automatic, -1 = real-line kernels)
  int tri_cl = 0;     // in use
  int stream_req = 0; // streaming (nontemporal) x and y passes: -1 off, 0 by size, 1 on
  bool stream = false;  // in use
  int persist_req = 0;  // persistent prefetching x/z passes: 1 where built, else (0, -1) off
  bool persist_x = false, persist_z = false;  // in use
  FftLen fx{}, fz{};          // complex lengths h and nz
  double2 *d_tw = nullptr;    // all twiddle tables
  const double2 *d_wx = nullptr;  // real-FFT post twiddles exp(-2 pi i k / nx), k = 0..h
  void *d_spec = nullptr;     // spectral buffer pitch * ny * nz complex (double2, or float2 for FP32 storage)
  double *d_psum = nullptr;   // FP32-storage form: per-block sums of phi over the j = 1 plane
  double *d_tri_scratch = nullptr;  // Thomas variant: c' per point (real)
  size_t spec_elems = 0;
  size_t spec_bytes = 0;
  bool profile = false;
  std::vector<hipEvent_t> prof_ev;  // 6 events per profiled solve
  int prof_used = 0;                // profiled solves since the last read
};

// Energy spectra of clcStat (ptc_spectra.hip): options, accumulators
// EnergySpecX(nx/2+1, ny, NE1), EnergySpecZ(nz/2+1, ny, NE1) and
// EnergySpec2D(nx, ny2d, nz, NE2) in the reference's column-major order, and
// the work buffers of one sample.
struct SpectraState {
  bool ready = false, on1d = false, on2d = false;
  int jlcs = 1, ny2d = 0, ne1 = 8, ne2 = 5;
  int nspectime = 0;
  std::vector<int> iy2d;  // 1-based rows of the 2-D spectra
  int *d_iy2d = nullptr;
  double *d_x = nullptr, *d_z = nullptr, *d_2d = nullptr;
  int pitch = 0, nch = 0;                      // spectral row pitch and kz chunks of the work buffers
  double2 *d_fa = nullptr, *d_fb = nullptr;    // FP64 2-D spectra pitch * ny * nz
  double *d_cell = nullptr;                    // cell-centre field, field layout
  double2 *d_px = nullptr, *d_pz = nullptr;    // partial sums
  double2 *d_hx = nullptr;                     // symbol of the cell-centre interpolation of ux
};

}  // namespace ptc

struct ptc_ctx {
  int nx = 0, ny = 0, nz = 0;
  double xlx = 0, yly = 0, zlz = 0;
  ptc_params prm{};
  size_t plane = 0;   // nx * (ny + 2)
  size_t n3 = 0;      // nx * (ny + 2) * nz
  ptc::Geo geo{};
  std::vector<double> h_met;  // M_NMET * (ny + 2)
  std::vector<double> h_wnx, h_wnz;
  double normfft = 0;
  double *d_met = nullptr;
  double *d_wnx = nullptr, *d_wnz = nullptr;
  double *d_field[PTC_NFIELDS] = {};
  double *d_red = nullptr;     // reduction partials
  int red_cap = 0;
  double *d_scal = nullptr;    // device scalars
  double *d_cnfac = nullptr;   // Crank-Nicolson factor tables (kCnMaxJobs + kCnStepTables) x kCnFacStride
  // beta each step table (substep ns, kind) holds, NaN when none; the tables
  // also depend on the chunk plan (cn_tab_mm, cn_tab_p) and the metrics
  double cn_tab_beta[9] = {NAN, NAN, NAN, NAN, NAN, NAN, NAN, NAN, NAN};
  int cn_tab_mm = 0, cn_tab_p = 0;
  double *h_pinned = nullptr;  // pinned host scalars
  hipStream_t stream = nullptr;
  int device = 0;
  size_t lds_limit = 65536;
  int n_cu = 256;
  // time scheme (src/m_Parameters.f90:40-61)
  double gconst[3] = {0, 0, 0}, tconst[3] = {0, 0, 0}, aconst[3] = {0, 0, 0};
  int iadvance = 3;
  double pmGamma = 0, pmTheta = 0, pmAlpha = 0, pmAlphaC = 0, pmBeta = 0, pmBetaSc = 0;
  double dt = 0, simtime = 0, prgradaver = 0;
  double divmax_prsrc = 0;
  ptc::PoissonPlan pp;
  double *d_stat = nullptr;  // clcStat accumulator SumStat(NCHASTAT, ny+1)
  void *d_visu = nullptr;    // dump_visu staging (visu_bytes)
  size_t visu_bytes = 0;
  int nfstime = 0;           // samples in it
  ptc::SpectraState spec;
  bool step_graphs = true;               // ptc_channel_iterate replays a captured step when dt is fixed
  hipGraph_t step_graph = nullptr;       // one fixed-dt ChannelIterate step, captured on first use
  hipGraphExec_t step_exec = nullptr;
  // Buffer generation: bumped whenever a buffer that a captured step may
  // hold (reduction partials, Poisson plan buffers) is freed or reallocated.
  // A graph records the generation it was captured at and is re-captured
  // before a replay when the two differ.
  unsigned long long buf_gen = 1;
  unsigned long long step_gen = 0;       // generation of step_graph
  std::string search_log;  // the last Poisson plan search (ptc_poisson_search_log)
  // clcRhsXYZ workgroup shape requests (0 = automatic; ptc_rhs_set)
  int rhs_seg_req = 0, rhs_rows_req = 0, rhs_run_req = 0, rhs_minpl_req = 0;
};

namespace ptc {

// Workgroups go to the 8 XCDs round-robin by linear id (w -> XCD w % 8), and
// each XCD has its own L2.  xcd_group maps w to a work index so that GRP
// consecutive work indices run on one XCD as consecutive workgroups there;
// the last (nwg mod 8 GRP) ids keep their own index.
template <int GRP>
__device__ inline int xcd_group(int w, int nwg) {
  if (GRP == 1) return w;
  constexpr int span = 8 * GRP;
  if (w >= (nwg / span) * span) return w;
  const int x = w % 8, s = w / 8;
  return GRP * (x + 8 * (s / GRP)) + s % GRP;
}

// xcd_group with a run length chosen at launch (1: identity)
__device__ inline int xcd_run(int w, int nwg, int grp) {
  if (grp <= 1) return w;
  const int span = 8 * grp;
  if (w >= (nwg / span) * span) return w;
  const int x = w % 8, s = w / 8;
  return grp * (x + 8 * (s / grp)) + s % grp;
}

// field access (allocating on first use, zero-filled)
int field(ptc_ctx *c, int id, double **out);
// field access without allocating: PTC_E_STATE if the field has no device
// memory yet
int field_existing(ptc_ctx *c, int id, double **out);
int check_field(ptc_ctx *c, int id);
// resolve, once per device, the device address of the file-scope scratch of
// the allocation-free entry points (ptc_noalloc.hip); the device is current
int noalloc_init_device(int device);

// The divergence of clcPrSrc formed in the x forward pass (src/m_TScheme.f90:48-77):
// the pass transforms sudtal * div(ux, uy, uz) and writes each workgroup's
// maximum |div| to part[workgroup].
struct DivSrc {
  const double *ux, *uy, *uz;
  double sudtal;
  double *part;
};

// Poisson
int poisson_build_plan(ptc_ctx *c);
int poisson_free_plan(ptc_ctx *c);
// ghosts: also fill dst's Neumann ghost rows j = 0, ny+1 (SetBC_and_UpdateHalo_pr)
int poisson_solve(ptc_ctx *c, const double *src, double *dst, bool ghosts = false);
// clcPrSrc + clcPPE with the source formed in the x pass (FP64 plans with the
// specialised x kernel and no slabs: poisson_div_fusable)
bool poisson_div_fusable(ptc_ctx *c);
int poisson_solve_div(ptc_ctx *c, const DivSrc &div, double *dst, bool ghosts);
int poisson_prepare(ptc_ctx *c);
int stencil_prepare(ptc_ctx *c);
// one ChannelIterate step (monitors into mon[7] when mon != NULL)
int channel_step(ptc_ctx *c, double *mon);
// device scalar holding the x-z mean of phi's first row before a correction
// (the refinement's x-inverse subtracts it, keeping the (0,0) pin)
constexpr int kSlotPinShift = 4;
// set to 1 by the visualisation kernel where lambda2's cubic has a complex root pair
constexpr int kSlotVisuErr = 32;

// fill n doubles of device memory with v on the context's stream
int fill_device(ptc_ctx *c, double *a, long long n, double v);
// clcStat accumulation and read-back (ptc_stencil.hip)
int launch_stat(ptc_ctx *c, const double *ux, const double *uy, const double *uz, const double *pr,
                const double *sc);
int read_stat(ptc_ctx *c, double *out, double *prgradsum, int reset);
// energy spectra (ptc_spectra.hip)
int spectra_init(ptc_ctx *c, int clc1d, int jlcs, int clc2d, int ny2d, const int *iy2d);
int spectra_sample(ptc_ctx *c, const double *ux, const double *uy, const double *uz, const double *p,
                   const double *sc);
int spectra_read(ptc_ctx *c, double *sx, double *sz, double *s2, int *nspectime, int reset);
void spectra_free(ptc_ctx *c);
// allocate everything a step touches, so the step can be captured in a graph
int prepare_step(ptc_ctx *c);
int capture_step(ptc_ctx *c, hipGraph_t *graph, hipGraphExec_t *exec);
// forget the captured step (its buffers are about to change)
void drop_step_graph(ptc_ctx *c);
// the reduction partials d_red with room for n values; growing them bumps the
// buffer generation and drops the context's captured step
int ensure_red(ptc_ctx *c, int n);
// the clcRhsXYZ workgroup shape in use: threads per row segment (the padded
// row where a row is one segment), rows per workgroup, tiles per XCD run
void rhs_shape_in_use(const ptc_ctx *c, int *seg_threads, int *rows, int *run, int *min_planes);
int poisson_residual(ptc_ctx *c, const double *src, const double *dst, double *rmax);
int poisson_stage_times(ptc_ctx *c, float *ms5, int *nsolves);

// FFT launchers (ptc_fft.hip)
int launch_fft_x_fwd(ptc_ctx *c, const double *src, void *spec, bool f64 = false);
int launch_fft_x_inv(ptc_ctx *c, const void *spec, double *dst, double scale, bool acc, bool ghost);
int launch_fft_z(ptc_ctx *c, void *spec, bool inverse, bool f64 = false);
int fft_make_len(int n, FftLen *L);
bool fft2_has_x(int h);
bool fft2_has_z(int nz, int req, int *tile);
// specialised kernels over the planes j0 .. j0+nj-1 (0-based) of the spectrum;
// the spectrum holds double2 or, in the FP32-storage form, float2 values;
// acc: the x-inverse adds its result to dst
// f64: FP64 spectra whatever the plan's storage precision (the statistics spectra);
// div: form the rows from the velocities instead of reading src
// The FP64 residual of the FP32-storage refinement formed in the x forward
// pass: the pass transforms f - L phi (the operator of k_residual: 4th-order
// staggered x and z, am2ph/ac2ph/ap2ph in y with the Neumann ghosts) and
// writes each workgroup's sum of phi over its lines of row j = 1 to
// psum[workgroup].
struct ResSrc {
  const double *f, *phi;
  double rdx2, rdz2;
  double *psum;
};
int fft2_x_fwd(ptc_ctx *c, const double *src, void *spec, int j0, int nj, bool f64 = false,
               const DivSrc *div = nullptr, const ResSrc *res = nullptr);
// workgroups of one specialised x forward launch over nj planes (0: no such kernel)
long long fft2_x_fwd_blocks(const ptc_ctx *c, int nj);
// persistent prefetching forms built for this x length (nx/2) / z length and tile
bool fft2_has_px(int h);
bool fft2_has_pz(int nz, int tile);
int fft2_x_inv(ptc_ctx *c, const void *spec, double *dst, double scale, int j0, int nj, bool acc, bool ghost);
int fft2_z(ptc_ctx *c, void *spec, bool inverse, int j0, int nj, bool f64 = false);

// tridiagonal launchers (ptc_tridiag.hip)
int launch_tri_poisson(ptc_ctx *c, void *spec);
// Crank-Nicolson y-solves u += (I - beta L_y)^-1 rhs: one job per field;
// kind 0 (u/w), 1 (v, wall row from duy_ym), 2 (scalar); shift: optional
// device scalar added to rhs.  wall (kind 1, may be NULL): afterwards set
// the wall row to the plane wall and row ny+1 to top, as
// SetBC_and_UpdateHaloForPrSrc does (src/m_BC_and_Halo.f90:78-98).
struct CnJob {
  double *u;
  const double *rhs, *duy_ym, *shift, *wall;
  double beta, top;
  int kind, rA, rB, rC;
  // v job with Update_uy_ym folded in: the thread of a line's wall row reads
  // the old uy_ym, uses -old as the wall du, writes uy_ym = 0 and
  // duy_ym = -old, and fills the wall row with the new uy_ym
  double *uym_upd, *duym_upd;
  // factor table of the job (index into d_cnfac) and whether the launch
  // forms it (k_cn_factor) or finds it formed by an earlier launch
  int tab, need;
  // u job of a constant-flow-rate step: the shift added to the right-hand
  // side comes from the flow-rate partials (FlowFold), not from shift
  int fold;
  // scalar job of a step: the solve also writes the wall ghost rows,
  // c(0) = r1m c(1) + r2m and c(ny+1) = r1p c(ny) + r2p
  // (SetBC_and_UpdateHalo_scalar, src/m_BC_and_Halo.f90:137-176)
  int sbc;
  double r1m, r2m, r1p, r2p;
};
// The constant-flow-rate correction (src/m_TScheme.f90:269-300) formed inside
// the Crank-Nicolson launch: the RHS kernel's n partial sums of the bulk flow
// rate and the coefficients k_flowrate would take; dp1ns and prgrad are the
// device scalars it would write (prgrad set rather than added on the first
// substep).
struct FlowFold {
  const double *part;
  int n, scalar_build, first;
  double xnu, ubulk, ucrf, nxnz, yly, alpha, alphaC, dt;
  double *dp1ns, *prgrad;
};
constexpr int kCnMaxJobs = 4;
// tables kept across steps: one per substep (1..3) and kind (0 u/w, 1 v, 2
// scalar), after the kCnMaxJobs per-launch ones
constexpr int kCnStepTables = 9;
// doubles of one job's Crank-Nicolson factor table (chunk rows P * MM <= 256,
// 2P <= 64 reduced rows, at most 6 cyclic-reduction steps: 4 * 256 + 64 +
// 4 * 6 * 64 = 2624)
constexpr int kCnFacStride = 2688;
struct CnJobs {
  CnJob j[kCnMaxJobs];
  int n;
};
CnJob cn_job(int kind, double *u, const double *rhs, const double *duy_ym, double beta, const double *shift);
int launch_tri_cn_jobs(ptc_ctx *c, const CnJobs &jobs);
// the jobs of substep ns of a step: each job's factors in its kept step
// table, formed only where the table holds another beta (or none); with
// solve false, only the factors
int launch_tri_cn_step(ptc_ctx *c, CnJobs jobs, int ns, bool solve = true, const FlowFold *ff = nullptr);
int launch_tri_cn(ptc_ctx *c, int kind, double *u, const double *rhs, const double *duy_ym,
                  double beta, const double *shift);

// stencils (ptc_stencil.hip)
// defer_flowrate: a constant-flow-rate case leaves the flow-rate sum to the
// Crank-Nicolson launch (flow_fold) instead of launching k_flowrate
int flow_fold(ptc_ctx *c, bool first, FlowFold *ff);
// the scalar's wall-row coefficients (r1m, r2m, r1p, r2p)
int scalar_bc_coefs(const ptc_ctx *c, double *r);
int launch_rhs(ptc_ctx *c, const double *ux, const double *uy, const double *uz,
               const double *sc, double *rhsx, double *rhsy, double *rhsz, double *rhsc,
               double *hx, double *hy, double *hz, double *hc, const double *p, bool defer_flowrate = false);
int launch_prsrc(ptc_ctx *c, const double *ux, const double *uy, const double *uz, double *prsrc,
                 double *divmax_host, bool reduce = true);
// the same with the workgroup partials in part (room for prsrc_partials(c)
// values) instead of the context's reduction buffer; allocates nothing
int launch_prsrc_part(ptc_ctx *c, const double *ux, const double *uy, const double *uz, double *prsrc,
                      double *divmax_host, bool reduce, double *part);
int prsrc_partials(const ptc_ctx *c);
// the divergence maximum over n partials into its device slot
int launch_divmax(ptc_ctx *c, int n);
int launch_vel_update(ptc_ctx *c, const double *phi, double *ux, double *uy, double *uz);
int launch_pressure_update(ptc_ctx *c, double *p, const double *phi);
int launch_vel_p_update(ptc_ctx *c, const double *phi, double *ux, double *uy, double *uz, double *p);
int launch_bc_vel(ptc_ctx *c, double *ux, double *uy, double *uz, const double *uy_ym, bool prsrc_only);
int launch_bc_pr(ptc_ctx *c, double *p);
// the velocity/pressure update with the wall and Neumann ghost rows written
// by the same threads (SetBC_and_UpdateHalo + SetBC_and_UpdateHalo_pr)
int launch_vel_p_update_bc(ptc_ctx *c, const double *phi, double *ux, double *uy, double *uz, double *p,
                           const double *uy_ym);
int launch_bc_scalar(ptc_ctx *c, double *s);
int launch_visu(ptc_ctx *c, int which, const double *ux, const double *uy, const double *uz, const double *p,
                const double *sc, int iskip, int jskip, int kskip, int real_bytes, void *dev_out);
int launch_uy_ym(ptc_ctx *c, double *uy_ym, double *duy_ym);
int launch_monitor(ptc_ctx *c, int which, const double *ux, const double *uy, const double *uz,
                   const double *s, double *out);
int reduce_sum_s3(ptc_ctx *c, int nparts, double *host_out);

// The partial results of a launch reduced by one workgroup of
// kReduceThreads threads: every thread has eight loads in flight (indices past
// n read the last partial: a duplicate, masked out of sums, harmless to
// maxima), then a fixed-order tree in LDS.  n >= 1.
constexpr int kReduceThreads = 1024;
template <bool MAX>
__device__ __forceinline__ double reduce_partials(const double *__restrict__ part, int n) {
  __shared__ double sh[kReduceThreads];
  const int t = threadIdx.x;
  double acc = 0.0;
  for (int q0 = 0; q0 < n; q0 += 8 * kReduceThreads) {
    double v[8];
#pragma unroll
    for (int u = 0; u < 8; ++u) v[u] = part[min(q0 + u * kReduceThreads + t, n - 1)];
#pragma unroll
    for (int u = 0; u < 8; ++u) {
      if (MAX) acc = fmax(acc, v[u]);
      else acc += v[u] * ((q0 + u * kReduceThreads + t < n) ? 1.0 : 0.0);
    }
  }
  sh[t] = acc;
  __syncthreads();
  for (int w = kReduceThreads / 2; w > 0; w >>= 1) {
    if (t < w) sh[t] = MAX ? fmax(sh[t], sh[t + w]) : sh[t] + sh[t + w];
    __syncthreads();
  }
  return sh[0];
}

}  // namespace ptc
