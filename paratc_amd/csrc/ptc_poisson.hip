// ptc_poisson.hip -- the pressure Poisson solve (m_Poisson) on one GPU.
//
// The reference solves L phi = f with L = D_x G_x + D_y G_y + D_z G_z
// (4th-order staggered in x and z, 2nd-order stretched in y) by real FFTs in
// x and z, one tridiagonal solve in y per wavenumber pair, and the inverse
// FFTs (clcPPE_x1, src/m_Poisson.f90:352-452).  On one device the six pencil
// transposes of that routine are unnecessary: the whole field is resident and
// every pass reads the dimension it transforms directly.  A solve is five
// streaming passes over a complex half-spectrum buffer S[kz][j][kx]
// (kx = 0..nx/2, row pitch rounded up to 16 complex):
//
//   1. x forward   src (real, device field layout)  -> S     (ptc_fft.hip)
//   2. z forward   S -> S, tiles of T kx columns           (ptc_fft.hip)
//   3. y solve     S -> S, one lane per (kx, kz) line      (ptc_tridiag.hip)
//   4. z inverse   S -> S
//   5. x inverse   S -> dst (real), times normfft = 1/(nx nz)
//
// Each pass reads and writes every point once: 16 bytes per grid point per
// pass in FP64, 80 N bytes per solve.  The modified wavenumber of the complex
// mode kx equals the reference's WaveNumX at the half-complex position of the
// same frequency (the symbol is even in kx), so both solves diagonalise the
// same operator.
#include <algorithm>
#include <cmath>
#include <vector>

#include "ptc_internal.hpp"

namespace ptc {

bool fft_config_ok(const FftLen &L, int ncol, int nthr);
int tri_plan(int ny, size_t lds_limit, bool poisson, int mm_req, int lpb_req, int *P_out, int *mm_out, int *lpb_out);
int tri_plan_lds(int ny, size_t lds_limit, int ll, int mm_req, int *P_out, int *mm_out);
int tri_plan_cplx(int ny, size_t lds_limit, int cl, int mm_req, int *P_out, int *mm_out);

static void make_twiddles(int n, std::vector<double2> &out) {
  const double twopi = 6.283185307179586476925286766559005768;
  out.resize((size_t)n);
  for (int m = 0; m < n; ++m) {
    const double a = twopi * (double)m / (double)n;
    out[(size_t)m] = make_double2(std::cos(a), -std::sin(a));
  }
}

int poisson_free_plan(ptc_ctx *c) {
  PoissonPlan &P = c->pp;
  drop_step_graph(c);  // the captured step holds the plan's buffers
  ++c->buf_gen;        // and so do the ensemble's captured steps
  if (P.d_tw) (void)hipFree(P.d_tw);
  if (P.d_spec) (void)hipFree(P.d_spec);
  if (P.d_psum) (void)hipFree(P.d_psum);
  if (P.d_tri_scratch) (void)hipFree(P.d_tri_scratch);
  for (auto &e : P.prof_ev)
    if (e) (void)hipEventDestroy(e);
  PoissonPlan keep;
  keep.prec = P.prec;
  keep.tri_req = P.tri_req;
  keep.tri_mm_req = P.tri_mm_req;
  keep.slab_req = P.slab_req;
  keep.tri_ll_req = P.tri_ll_req;
  keep.tri_lpb_req = P.tri_lpb_req;
  keep.tri_cl_req = P.tri_cl_req;
  keep.stream_req = P.stream_req;
  keep.persist_req = P.persist_req;
  keep.sweeps = P.sweeps;
  keep.ztile_req = P.ztile_req;
  keep.zfast_req = P.zfast_req;
  keep.profile = P.profile;
  P = keep;
  return PTC_OK;
}

int poisson_build_plan(ptc_ctx *c) {
  PoissonPlan &P = c->pp;
  if (P.ready) return PTC_OK;
  PTC_CHECK(c->nx % 2 == 0 && c->nz % 2 == 0, PTC_E_ARG, "mod(nxc,2)/=0 || mod(nzc,2)/=0 WRONG");
  P.h = c->nx / 2;
  // row pitch: a multiple of 16 complex values, or of a wider requested z tile
  const int pq = std::max(16, P.zfast_req);
  P.pitch = ((P.h + 1 + pq - 1) / pq) * pq;
  PTC_TRY(fft_make_len(P.h, &P.fx));
  PTC_TRY(fft_make_len(c->nz, &P.fz));
  const size_t lds_cap = std::min<size_t>(c->lds_limit, 65536);

  // specialised kernels for the production lengths unless a tile is forced
  P.x_fast = P.ztile_req == 0 && fft2_has_x(P.h);
  if (!P.x_fast) {
    // generic x pass: lines per workgroup so that a tile holds ~2048 complex values
    P.x_lines = std::max(1, 2048 / P.h);
    while (P.x_lines > 1 && sizeof(double2) * (size_t)P.h * P.x_lines > lds_cap) P.x_lines /= 2;
    PTC_CHECK(sizeof(double2) * (size_t)P.h * P.x_lines <= lds_cap, PTC_E_UNSUPPORTED,
              "nx too large for the LDS x-FFT");
    P.x_threads = 0;
    for (int t : {256, 512, 1024})
      if (fft_config_ok(P.fx, P.x_lines, t)) { P.x_threads = t; break; }
    PTC_CHECK(P.x_threads > 0, PTC_E_UNSUPPORTED, "no x-FFT launch shape fits");
  }

  int zt = 0;
  P.z_fast = P.ztile_req == 0 && fft2_has_z(c->nz, P.zfast_req, &zt) && P.pitch % zt == 0;
  if (P.z_fast) {
    P.ztile = zt;
  } else {
    // generic z pass: widest column tile (<= 16, divides pitch) that fits LDS
    int T = P.ztile_req > 0 ? P.ztile_req : 16;
    while (T > 1 && (sizeof(double2) * (size_t)c->nz * T > lds_cap || (size_t)T * c->nz > 4096)) T /= 2;
    PTC_CHECK(P.pitch % T == 0, PTC_E_ARG, "z tile must divide the spectral pitch");
    PTC_CHECK(sizeof(double2) * (size_t)c->nz * T <= lds_cap, PTC_E_UNSUPPORTED, "nz too large for the LDS z-FFT");
    P.ztile = T;
    P.z_threads = 0;
    for (int t : {256, 512, 1024})
      if (fft_config_ok(P.fz, T, t)) { P.z_threads = t; break; }
    PTC_CHECK(P.z_threads > 0, PTC_E_UNSUPPORTED, "no z-FFT launch shape fits");
  }

  // x/z passes over slabs of y planes (specialised kernels only): the z pass
  // then reads a slab the x pass has just written while it is still held in
  // the device's last-level cache, and likewise on the way back
  P.slab = (P.x_fast && P.z_fast && P.slab_req > 0 && P.slab_req < c->ny) ? P.slab_req : 0;
  // persistent prefetching x and z passes where built for these lengths
  // (1 both, 2 z only, 3 x only)
  P.persist_x = (P.persist_req == 1 || P.persist_req == 3) && P.x_fast && fft2_has_px(P.h);
  P.persist_z = (P.persist_req == 1 || P.persist_req == 2) && P.z_fast && fft2_has_pz(c->nz, P.ztile);

  if (P.prec == PTC_PREC_FP32_REFINE) {
    PTC_CHECK(P.x_fast && P.z_fast, PTC_E_UNSUPPORTED,
              "the FP32-storage Poisson form needs specialised FFT kernels for nx/2 and nz");
    PTC_CHECK(P.tri_req != PTC_TRI_THOMAS, PTC_E_UNSUPPORTED, "the Thomas y-solve has no FP32-storage form");
  }

  // y pass: partitioned unless Thomas is asked for, or (automatic choice) no
  // partitioned shape fits this ny.  Without a kernel asked for, the
  // complex-line kernel with 16 lines per workgroup where it fits, else the
  // real-line register kernel.
  P.tri = P.tri_req == PTC_TRI_THOMAS ? PTC_TRI_THOMAS : PTC_TRI_PARTITIONED;
  P.tri_ll = 0;
  P.tri_cl = 0;
  const bool cl_auto = P.tri_cl_req == 0 && P.tri_ll_req == 0 && P.tri_lpb_req == 0;
  if (P.tri == PTC_TRI_PARTITIONED && cl_auto &&
      tri_plan_cplx(c->ny, c->lds_limit, 16, P.tri_mm_req, &P.tri_chunks, &P.tri_mmax) == PTC_OK) {
    P.tri_cl = 16;
    P.tri_lpb = 16;
  } else if (P.tri == PTC_TRI_PARTITIONED && P.tri_cl_req > 0) {
    PTC_TRY(tri_plan_cplx(c->ny, c->lds_limit, P.tri_cl_req, P.tri_mm_req, &P.tri_chunks, &P.tri_mmax));
    P.tri_cl = P.tri_cl_req;
    P.tri_lpb = P.tri_cl;
  } else if (P.tri == PTC_TRI_PARTITIONED && P.tri_ll_req > 0) {
    PTC_TRY(tri_plan_lds(c->ny, c->lds_limit, P.tri_ll_req, P.tri_mm_req, &P.tri_chunks, &P.tri_mmax));
    P.tri_ll = P.tri_ll_req;
    P.tri_lpb = P.tri_ll;
  } else if (P.tri == PTC_TRI_PARTITIONED) {
    const int rc =
        tri_plan(c->ny, c->lds_limit, true, P.tri_mm_req, P.tri_lpb_req, &P.tri_chunks, &P.tri_mmax, &P.tri_lpb);
    if (rc != PTC_OK) {
      if (P.tri_req == PTC_TRI_PARTITIONED || P.prec == PTC_PREC_FP32_REFINE) return rc;
      P.tri = PTC_TRI_THOMAS;
    }
  }

  // twiddle tables: [fx: h][fz: nz][wx: h+1]
  std::vector<double2> tx, tz, all;
  make_twiddles(P.h, tx);
  make_twiddles(c->nz, tz);
  std::vector<double2> tnx;
  make_twiddles(c->nx, tnx);
  all.reserve(tx.size() + tz.size() + (size_t)P.h + 1);
  all.insert(all.end(), tx.begin(), tx.end());
  all.insert(all.end(), tz.begin(), tz.end());
  all.insert(all.end(), tnx.begin(), tnx.begin() + P.h + 1);
  PTC_HIP(hipMalloc(&P.d_tw, sizeof(double2) * all.size()));
  PTC_HIP(hipMemcpy(P.d_tw, all.data(), sizeof(double2) * all.size(), hipMemcpyHostToDevice));
  P.fx.tw = P.d_tw;
  P.fz.tw = P.d_tw + tx.size();
  P.d_wx = P.d_tw + tx.size() + tz.size();

  P.spec_elems = (size_t)P.pitch * c->ny * c->nz;
  const size_t spec_bytes = (P.prec == PTC_PREC_FP32_REFINE ? sizeof(float2) : sizeof(double2)) * P.spec_elems;
  P.spec_bytes = spec_bytes;
  // streaming (nontemporal) x passes and y-solve where the spectrum is larger
  // than the 256 MB last-level cache: such a pass touches every value once,
  // and keeping it cached only evicts other data
  constexpr size_t kStreamBytes = size_t(256) << 20;
  P.stream = P.stream_req > 0 || (P.stream_req == 0 && spec_bytes >= kStreamBytes);
  if (hipMalloc(&P.d_spec, spec_bytes) != hipSuccess) {
    P.d_spec = nullptr;
    set_error("spectral buffer allocation failed");
    return PTC_E_ALLOC;
  }
  PTC_HIP(hipMemsetAsync(P.d_spec, 0, spec_bytes, c->stream));
  if (P.prec == PTC_PREC_FP32_REFINE && P.sweeps > 0) {
    // the correction's x pass forms the residual as it loads its lines and
    // writes per-workgroup sums of phi's first row (a slabbed x pass rounds
    // up once per slab)
    const long long npart = fft2_x_fwd_blocks(c, c->ny) + c->ny;
    if (hipMalloc(&P.d_psum, sizeof(double) * (size_t)npart) != hipSuccess) {
      P.d_psum = nullptr;
      set_error("residual partial-sum allocation failed");
      return PTC_E_ALLOC;
    }
  }
  P.ready = true;
  return PTC_OK;
}

static int ensure_thomas_scratch(ptc_ctx *c) {
  PoissonPlan &P = c->pp;
  if (P.tri != PTC_TRI_THOMAS || P.d_tri_scratch) return PTC_OK;
  if (hipMalloc(&P.d_tri_scratch, sizeof(double) * P.spec_elems) != hipSuccess) {
    P.d_tri_scratch = nullptr;
    set_error("Thomas scratch allocation failed");
    return PTC_E_ALLOC;
  }
  return PTC_OK;
}

// plan and scratch of the solve, built ahead of a captured step
int poisson_prepare(ptc_ctx *c) {
  PTC_TRY(poisson_build_plan(c));
  return ensure_thomas_scratch(c);
}

// One pass sequence src -> spectrum -> dst (acc: dst += result); ev (may be
// NULL) receives six events around the five passes.
int launch_pin_shift(ptc_ctx *c, int nblk);

static int solve_sequence(ptc_ctx *c, const double *src, double *dst, bool acc, bool ghosts, hipEvent_t *ev,
                          const DivSrc *div = nullptr, const ResSrc *res = nullptr) {
  PoissonPlan &P = c->pp;
  if (ev) PTC_HIP(hipEventRecord(ev[0], c->stream));
  if (div) {
    PTC_TRY(fft2_x_fwd(c, nullptr, P.d_spec, 0, c->ny, false, div));
    if (ev) PTC_HIP(hipEventRecord(ev[1], c->stream));
    PTC_TRY(launch_fft_z(c, P.d_spec, false));
  } else if (res && P.slab == 0) {
    PTC_TRY(fft2_x_fwd(c, nullptr, P.d_spec, 0, c->ny, false, nullptr, res));
    PTC_TRY(launch_pin_shift(c, (int)fft2_x_fwd_blocks(c, c->ny)));
    if (ev) PTC_HIP(hipEventRecord(ev[1], c->stream));
    PTC_TRY(launch_fft_z(c, P.d_spec, false));
  } else if (P.slab == 0) {
    PTC_TRY(launch_fft_x_fwd(c, src, P.d_spec));
    if (ev) PTC_HIP(hipEventRecord(ev[1], c->stream));
    PTC_TRY(launch_fft_z(c, P.d_spec, false));
  } else {
    // slabbed: the x and z times are reported together under x_fwd / x_inv;
    // a residual-forming x pass writes each slab's partial sums after the
    // previous slab's
    long long npart = 0;
    for (int j0 = 0; j0 < c->ny; j0 += P.slab) {
      const int nj = std::min(P.slab, c->ny - j0);
      if (res) {
        ResSrc rs = *res;
        rs.psum += npart;
        PTC_TRY(fft2_x_fwd(c, nullptr, P.d_spec, j0, nj, false, nullptr, &rs));
        npart += fft2_x_fwd_blocks(c, nj);
      } else {
        PTC_TRY(fft2_x_fwd(c, src, P.d_spec, j0, nj));
      }
      PTC_TRY(fft2_z(c, P.d_spec, false, j0, nj));
    }
    if (res) PTC_TRY(launch_pin_shift(c, (int)npart));
    if (ev) PTC_HIP(hipEventRecord(ev[1], c->stream));
  }
  if (ev) PTC_HIP(hipEventRecord(ev[2], c->stream));
  PTC_TRY(launch_tri_poisson(c, P.d_spec));
  if (ev) PTC_HIP(hipEventRecord(ev[3], c->stream));
  if (P.slab == 0) {
    PTC_TRY(launch_fft_z(c, P.d_spec, true));
    if (ev) PTC_HIP(hipEventRecord(ev[4], c->stream));
    PTC_TRY(launch_fft_x_inv(c, P.d_spec, dst, c->normfft, acc, ghosts));
  } else {
    if (ev) PTC_HIP(hipEventRecord(ev[4], c->stream));
    for (int j0 = 0; j0 < c->ny; j0 += P.slab) {
      const int nj = std::min(P.slab, c->ny - j0);
      PTC_TRY(fft2_z(c, P.d_spec, true, j0, nj));
      PTC_TRY(fft2_x_inv(c, P.d_spec, dst, c->normfft, j0, nj, acc, ghosts));
    }
  }
  if (ev) PTC_HIP(hipEventRecord(ev[5], c->stream));
  return PTC_OK;
}

int poisson_solve(ptc_ctx *c, const double *src, double *dst, bool ghosts) {
  PTC_TRY(poisson_build_plan(c));
  PoissonPlan &P = c->pp;
  PTC_TRY(ensure_thomas_scratch(c));
  hipEvent_t *ev = nullptr;
  if (P.profile) {
    const size_t need = (size_t)(P.prof_used + 1) * 6;
    while (P.prof_ev.size() < need) {
      hipEvent_t e;
      PTC_HIP(hipEventCreate(&e));
      P.prof_ev.push_back(e);
    }
    ev = P.prof_ev.data() + (size_t)P.prof_used * 6;
    ++P.prof_used;
  }
  PTC_TRY(solve_sequence(c, src, dst, false, ghosts, ev));
  if (P.prec == PTC_PREC_FP32_REFINE) {
    // defect correction: phi += L^-1 (f - L phi), the residual in FP64,
    // formed by the correction's x forward pass as it loads its lines (the
    // FP32 form always has the specialised x kernels), and the correction
    // through the same FP32-storage passes; the correction's x-inverse also
    // removes the mean of phi's first row, so the (0,0) mode keeps the
    // reference's pin phi(j=1) = 0 (src/m_Poisson.f90:399-404)
    const ResSrc res{src, dst, c->geo.rdx * c->geo.rdx, c->geo.rdz * c->geo.rdz, P.d_psum};
    for (int s = 0; s < P.sweeps; ++s) PTC_TRY(solve_sequence(c, nullptr, dst, true, ghosts, nullptr, nullptr, &res));
  }
  return PTC_OK;
}

bool poisson_div_fusable(ptc_ctx *c) {
  if (poisson_build_plan(c) != PTC_OK) return false;
  const PoissonPlan &P = c->pp;
  return P.x_fast && P.slab == 0 && P.prec == PTC_PREC_FP64;
}

int poisson_solve_div(ptc_ctx *c, const DivSrc &div, double *dst, bool ghosts) {
  PTC_TRY(poisson_build_plan(c));
  PoissonPlan &P = c->pp;
  PTC_CHECK(P.x_fast && P.slab == 0 && P.prec == PTC_PREC_FP64, PTC_E_STATE,
            "the divergence-forming solve needs an FP64 plan with the specialised x pass and no slabs");
  PTC_TRY(ensure_thomas_scratch(c));
  hipEvent_t *ev = nullptr;
  if (P.profile) {
    const size_t need = (size_t)(P.prof_used + 1) * 6;
    while (P.prof_ev.size() < need) {
      hipEvent_t e;
      PTC_HIP(hipEventCreate(&e));
      P.prof_ev.push_back(e);
    }
    ev = P.prof_ev.data() + (size_t)P.prof_used * 6;
    ++P.prof_used;
  }
  return solve_sequence(c, nullptr, dst, false, ghosts, ev, &div);
}

int poisson_stage_times(ptc_ctx *c, float *ms5, int *nsolves) {
  PoissonPlan &P = c->pp;
  float acc[5] = {0, 0, 0, 0, 0};
  if (P.prof_used > 0) {
    PTC_HIP(hipStreamSynchronize(c->stream));
    for (int q = 0; q < P.prof_used; ++q)
      for (int s = 0; s < 5; ++s) {
        float ms = 0.f;
        PTC_HIP(hipEventElapsedTime(&ms, P.prof_ev[(size_t)q * 6 + s], P.prof_ev[(size_t)q * 6 + s + 1]));
        acc[s] += ms;
      }
  }
  if (ms5)
    for (int s = 0; s < 5; ++s) ms5[s] = acc[s];
  if (nsolves) *nsolves = P.prof_used;
  return PTC_OK;
}

// ---- FP64 residual r = f - L phi, max |r| over the interior ---------------
// L in x and z is the composed stencil D o G of the 4th-order staggered
// operators: (1, -54, 783, -1460, 783, -54, 1) / 576 / dx^2, whose symbol is
// the modified wavenumber of src/m_Poisson.f90:42,46.
// Order the LDS writes of a wavefront before its lanes read each other's
// values (the lanes of one wavefront run in lockstep; this keeps the compiler
// from moving the accesses across).
__device__ __forceinline__ void wave_sync() {
  __builtin_amdgcn_fence(__ATOMIC_RELEASE, "wavefront");
  __builtin_amdgcn_wave_barrier();
  __builtin_amdgcn_fence(__ATOMIC_ACQUIRE, "wavefront");
}

// r = f - L phi over the interior with the FP64 staggered operator: 4th-order
// x/z (the symbol of the modified wavenumbers, (1, -54, 783, -1460, 783, -54,
// 1) / 576 rdx^2) and am2ph / ac2ph / ap2ph in y with the Neumann ghosts
// phi_0 = phi_1, phi_{ny+1} = phi_ny.
//
// A wavefront owns one y row of 64 x columns and marches along k over planes
// [k0, k1).  Every global load is issued kResAhead planes before its use: the
// column's z stencil sits in a register window that runs kResAhead - 1 planes
// past the stencil's reach, and f, the two y neighbours and the six x-halo
// columns (lanes 0..5) come through register queues of that depth.  The x
// neighbours are the adjacent lanes' centre values, exchanged through a
// per-wavefront LDS row.  part: per-block max |r| (the residual check of
// ptc_poisson_residual; the refinement forms its residual in the x pass).
constexpr int kResAhead = 4;

__global__ void __launch_bounds__(256) k_residual(const double *__restrict__ f, const double *__restrict__ phi,
                                                 Geo g, double rdx2, double rdz2, int kc,
                                                 double *__restrict__ part) {
  constexpr int D = kResAhead, NW = 7 + D - 1;
  __shared__ double xrow[4][2][72];  // per wavefront: one plane row, 3 halo columns each side
  __shared__ double sh[256];
  const int nx = g.nx, ny = g.ny, nz = g.nz;
  const int lane = threadIdx.x, wv = threadIdx.y;
  const int i0 = blockIdx.x * 64, i = i0 + lane;
  const int j = blockIdx.y * 4 + wv + 1;
  const int k0 = blockIdx.z * kc, k1 = min(k0 + kc, nz);
  const bool act = i < nx && j <= ny;
  const int jc = min(j, ny);
  // the column this lane loads: its own (wrapped for lanes past nx), or for
  // lanes 0..5 also an x-halo column (slot 0..2 and 67..69 of the LDS row)
  const int col = (i0 + lane) % nx;
  const int hslot = lane < 3 ? lane : 64 + lane;
  const int hcol = ((i0 - 3 + hslot) % nx + nx) % nx;
  const long long pl = (long long)nx * (ny + 2);
  const double am = g.met[M_AM2PH * g.nyh + jc], ac = g.met[M_AC2PH * g.nyh + jc],
               ap = g.met[M_AP2PH * g.nyh + jc];
  constexpr double st0 = -1460.0 / 576.0, st1 = 783.0 / 576.0, st2 = -54.0 / 576.0, st3 = 1.0 / 576.0;
  auto wz = [&](int kk) { return kk < 0 ? kk + nz : (kk >= nz ? kk - nz : kk); };
  const double *cc = phi + col + (long long)nx * jc;   // this column at row jc
  const double *hc = phi + hcol + (long long)nx * jc;  // halo column at row jc
  const double *fc = f + col + (long long)nx * jc;
  const bool halo = lane < 6;
  double w[NW], fq[D], qm[D], qp[D], hq[D];
#pragma unroll
  for (int s = 0; s < NW; ++s) w[s] = cc[wz(k0 + s - 3) * pl];
#pragma unroll
  for (int s = 0; s < D; ++s) {
    const int kk = wz(k0 + s);
    const long long kp = (long long)kk * pl;
    fq[s] = fc[kp];
    qm[s] = cc[kp - nx];
    qp[s] = cc[kp + nx];
    hq[s] = halo ? hc[kp] : 0.0;
  }
  double mx = 0.0;
  for (int k = k0; k < k1; ++k) {
    double *R = xrow[wv][k & 1];
    const double c0 = w[3];
    R[lane + 3] = c0;
    if (halo) R[hslot] = hq[0];
    wave_sync();
    const double lx = st0 * c0 + st1 * (R[lane + 2] + R[lane + 4]) + st2 * (R[lane + 1] + R[lane + 5]) +
                      st3 * (R[lane] + R[lane + 6]);
    const double lz = st0 * c0 + st1 * (w[2] + w[4]) + st2 * (w[1] + w[5]) + st3 * (w[0] + w[6]);
    const double ym = jc > 1 ? qm[0] : c0;
    const double yp = jc < ny ? qp[0] : c0;
    const double lap = rdx2 * lx + rdz2 * lz + (am * ym + ac * c0 + ap * yp);
    const double res = fq[0] - lap;
    if (act) mx = fmax(mx, fabs(res));
    // advance the window and the queues by one plane; the new entries are
    // the planes k + 1 + 3 + (D - 1) (window) and k + D (queues)
#pragma unroll
    for (int s = 0; s < NW - 1; ++s) w[s] = w[s + 1];
#pragma unroll
    for (int s = 0; s < D - 1; ++s) {
      fq[s] = fq[s + 1];
      qm[s] = qm[s + 1];
      qp[s] = qp[s + 1];
      hq[s] = hq[s + 1];
    }
    const long long kq = (long long)wz(k + D) * pl;
    w[NW - 1] = cc[(long long)wz(k + 3 + D) * pl];
    fq[D - 1] = fc[kq];
    qm[D - 1] = cc[kq - nx];
    qp[D - 1] = cc[kq + nx];
    if (halo) hq[D - 1] = hc[kq];
  }
  const int t = threadIdx.x + 64 * threadIdx.y;
  const long long blk = blockIdx.x + (long long)gridDim.x * (blockIdx.y + (long long)gridDim.y * blockIdx.z);
  sh[t] = mx;
  __syncthreads();
  for (int h = 128; h > 0; h >>= 1) {
    if (t < h) sh[t] = fmax(sh[t], sh[t + h]);
    __syncthreads();
  }
  if (t == 0) part[blk] = sh[0];
}

// launch shape of k_residual: (i, j) tiles of 64 x 4, k split into chunks
// until there are about sixteen workgroups per CU
static dim3 residual_grid(const ptc_ctx *c, int *kc) {
  const int gx = (c->nx + 63) / 64, gy = (c->ny + 3) / 4;
  const int want = std::max(1, (16 * c->n_cu + gx * gy - 1) / (gx * gy));
  const int nch = std::min(c->nz, want);
  *kc = (c->nz + nch - 1) / nch;
  return dim3((unsigned)gx, (unsigned)gy, (unsigned)((c->nz + *kc - 1) / *kc));
}

// out = (sum of the nblk partial sums, in a fixed order) * scale
__global__ void __launch_bounds__(kReduceThreads) k_sum_partials(const double *__restrict__ psum, int nblk,
                                                                  double scale, double *__restrict__ out) {
  const double s = reduce_partials<false>(psum, nblk);
  if (threadIdx.x == 0) out[0] = s * scale;
}

// the mean of phi's first row from nblk per-workgroup sums in d_psum, into the
// device slot the correction's x-inverse subtracts
int launch_pin_shift(ptc_ctx *c, int nblk) {
  hipLaunchKernelGGL(k_sum_partials, dim3(1), dim3(kReduceThreads), 0, c->stream, c->pp.d_psum, nblk,
                     1.0 / ((double)c->nx * (double)c->nz), c->d_scal + kSlotPinShift);
  PTC_HIP(hipGetLastError());
  return PTC_OK;
}

int poisson_residual(ptc_ctx *c, const double *src, const double *dst, double *rmax) {
  int kc = 0;
  const dim3 grid = residual_grid(c, &kc);
  const int nblk = (int)(grid.x * grid.y * grid.z);
  PTC_TRY(ensure_red(c, nblk));
  const double rdx2 = c->geo.rdx * c->geo.rdx, rdz2 = c->geo.rdz * c->geo.rdz;
  hipLaunchKernelGGL(k_residual, grid, dim3(64, 4), 0, c->stream, src, dst, c->geo, rdx2, rdz2, kc, c->d_red);
  PTC_HIP(hipGetLastError());
  std::vector<double> h((size_t)nblk);
  PTC_HIP(hipMemcpyAsync(h.data(), c->d_red, sizeof(double) * nblk, hipMemcpyDeviceToHost, c->stream));
  PTC_HIP(hipStreamSynchronize(c->stream));
  double m = 0.0;
  for (double v : h) m = std::max(m, v);
  *rmax = m;
  return PTC_OK;
}

}  // namespace ptc
