// ptc_stencil.hip -- the explicit stencils of one RK substep on gfx950.
//
// clcRhsXYZ (src/m_TScheme.f90:113-360) builds six 4th-order midpoint
// interpolations into haloed temporaries, refreshes their halos and then
// assembles RhsX/Y/Z (and RhsC).  Here one kernel computes every right-hand
// side of a grid point directly from ux, uy, uz, p (and the scalar): the
// interpolated values a point needs are evaluated in registers from the
// neighbouring velocities, so the six temporaries and their halo refreshes
// never touch HBM.  The operand order of every interpolation and flux follows
// the reference; x and z neighbours wrap periodically (one rank owns the whole
// domain), y uses the stored ghost rows.
//
// clcPrSrc, FluidVelUpdate, PressureUpdate (src/m_TScheme.f90:48-107,512-532),
// the wall boundary fills of m_BC_and_Halo (src/m_BC_and_Halo.f90:23-178),
// Update_uy_ym (src/m_FlowCase.f90:112-122) and the monitors of m_Tools
// (src/m_Tools.f90:24-110,256-301) follow.  Every reduction is a fixed-order
// tree over a fixed grid, so results are run-to-run deterministic.
#include <algorithm>
#include <vector>

#include "ptc_internal.hpp"

namespace ptc {

// device scalar slots in ctx->d_scal
enum { S_DP1NS = 0, S_PRGRAD = 1, S_DIVMAX = 2, S_PRGSUM = 3, S_MON0 = 8 };

__device__ __forceinline__ int wrapi(int i, int n) { return i < 0 ? i + n : (i >= n ? i - n : i); }

struct Ngh {
  int ii[7];
  long long kk[7];
  long long nxl;
  __device__ __forceinline__ Ngh(const Geo &g, int i, int k) {
    nxl = g.nx;
#pragma unroll
    for (int s = 0; s < 7; ++s) {
      ii[s] = wrapi(i + s - 3, g.nx);
      kk[s] = (long long)wrapi(k + s - 3, g.nz) * g.plane;
    }
  }
  // value of a at (i + di, j, k + dk)
  __device__ __forceinline__ double operator()(const double *__restrict__ a, int di, int j, int dk) const {
    return a[ii[di + 3] + nxl * j + kk[dk + 3]];
  }
};

struct RhsCoef {
  double gamma, theta, alpha, beta, beta_sc;
  double xnu, ucrf, gx, gy, gz;
  double sdiff, gev[3], fvv[3];
  int scalar, s3_mode;  // s3_mode: 0 none, 1 viscous (non-scalar build), 2 ux (ScalarFlow build)
};

// Fixed-order tree reductions over all threads of a block (1-D or 2-D); the
// block size must be a power of two no larger than the shared array (256).
__device__ __forceinline__ int lin_tid() { return threadIdx.x + blockDim.x * threadIdx.y; }
__device__ __forceinline__ int lin_size() { return blockDim.x * blockDim.y; }

template <typename T>
__device__ __forceinline__ T block_sum(T v, T *sh) {
  const int t = lin_tid();
  sh[t] = v;
  __syncthreads();
  for (int w = lin_size() / 2; w > 0; w >>= 1) {
    if (t < w) sh[t] += sh[t + w];
    __syncthreads();
  }
  return sh[0];
}

__device__ __forceinline__ double block_max(double v, double *sh) {
  const int t = lin_tid();
  sh[t] = v;
  __syncthreads();
  for (int w = lin_size() / 2; w > 0; w >>= 1) {
    if (t < w) sh[t] = fmax(sh[t], sh[t + w]);
    __syncthreads();
  }
  return sh[0];
}

// ---- clcRhsXYZ (src/m_TScheme.f90:113-360), row-marching form -------------
// The interpolations of Update_VelInterp (:143-167) are formed where they are
// used, from the neighbouring velocities, so the six haloed temporaries never
// reach HBM.  A workgroup owns one x row (one thread per column) at fixed j and
// marches along k over planes [k0, k1).  Each thread keeps its column's values
// in register windows along z (ux, uy, uz, scalar: k-3..k+3; p and uy at j+1:
// k-2..k+1; uz at j-1: k-1..k+2), so a value is loaded once per plane instead
// of once per use, and the next plane's loads are in flight while a plane is
// computed.  The x neighbours are the other threads' values of the same plane,
// exchanged through LDS rows; x wraps inside the row.  Rows longer than a
// workgroup go in segments (seg > 0): the workgroup of lw = seg + 6 threads
// owns seg columns from blockIdx.x * seg and loads three wrapped columns on
// either side; those halo threads take part in the exchange (their products
// two planes ahead included) but store nothing.  The two products whose
// interpolations cross x and z, uxhz * uzhx in h13 (a z difference) and
// uzhx * (uxhz + ucrf) in h31 (an x difference), are formed two planes ahead
// and kept in short queues.  Operand order follows the reference's
// expressions except that those products are formed before their difference
// weight is applied.
enum {
  X_UX = 0,  // ux(., j, k)
  X_UY,      // uy(., j, k)
  X_UZ,      // uz(., j, k)
  X_UZ2,     // uz(., j, k + 2)
  X_P,       // p(., j, k)
  X_UXM,     // ux(., j - 1, k)
  X_UYP,     // uy(., j + 1, k)
  X_R,       // uzhx * (uxhz + ucrf) at (., j, k)
  X_SC,      // scalar(., j, k)
  X_UYF,     // uy(., j, k + 2): the next window entry of row j - 1's uy(j + 1) window
  X_UZF,     // uz(., j, k + 3): the next window entry of row j + 1's uz(j - 1) window
  X_NROWS
};

// Raw buffer access to one plane: the plane's base address and size live in
// scalar registers and a lane adds a 32-bit byte offset, so a load costs one
// VGPR for the address whatever the number of fields in flight.  Offsets past
// the plane read zero and drop stores (the row kernels never form them).
__device__ __forceinline__ __amdgpu_buffer_rsrc_t plane_rsrc(const double *base, int bytes) {
  return __builtin_amdgcn_make_buffer_rsrc(const_cast<double *>(base), (short)0, bytes, 0x00020000);
}
__device__ __forceinline__ double bld(__amdgpu_buffer_rsrc_t r, unsigned off) {
  return __builtin_bit_cast(double, __builtin_amdgcn_raw_buffer_load_b64(r, off, 0, 0));
}
__device__ __forceinline__ void bst(__amdgpu_buffer_rsrc_t r, unsigned off, double v) {
  using V2 = decltype(__builtin_amdgcn_raw_buffer_load_b64(r, 0, 0, 0));
  __builtin_amdgcn_raw_buffer_store_b64(__builtin_bit_cast(V2, v), r, off, 0, 0);
}

template <bool SC, int MAXT>
__global__ void __launch_bounds__(MAXT) k_rhs_rows(Geo g, RhsCoef cf, int kc, const double *__restrict__ ux,
                                                  const double *__restrict__ uy, const double *__restrict__ uz,
                                                  const double *__restrict__ sc, const double *__restrict__ p,
                                                  double *__restrict__ rhsx, double *__restrict__ rhsy,
                                                  double *__restrict__ rhsz, double *__restrict__ rhsc,
                                                  double *__restrict__ hx, double *__restrict__ hy,
                                                  double *__restrict__ hz, double *__restrict__ hc,
                                                  double *__restrict__ s3part, int seg, int lw, int grp) {
  extern __shared__ double xr[];  // X_NROWS x (rows of the workgroup) rows of lw values
  const int nx = g.nx, nz = g.nz;
  const int i = threadIdx.x;
  // tile (segment bx, row group by, z chunk bz): runs of grp consecutive tiles,
  // segments fastest, go to one XCD, so the halo columns and edge rows that
  // neighbouring tiles share are read through the same L2 (with row groups
  // fastest the 4608x140x700 launch ran 17.8 ms against 15.5)
  const int nwg = (int)(gridDim.x * gridDim.y * gridDim.z);
  const int wt = xcd_run((int)(blockIdx.x + gridDim.x * (blockIdx.y + gridDim.y * blockIdx.z)), nwg, grp);
  const int bx = wt % (int)gridDim.x, by = (wt / (int)gridDim.x) % (int)gridDim.y,
            bz = wt / (int)(gridDim.x * gridDim.y);
  // whole row: thread = column; segment: thread t = column bx * seg - 3 + t (wrapped)
  int gcol = seg ? bx * seg - 3 + i : i;
  gcol = gcol < 0 ? gcol + nx : (gcol >= nx ? gcol - nx : gcol);
  // rows j of the workgroup, this thread's (uniform in a wave: blockDim.x is a multiple of 64)
  const int nr = (int)blockDim.y, r = __builtin_amdgcn_readfirstlane((int)threadIdx.y);
  const int jrow = by * nr + r + 1;
  const bool vrow = jrow <= g.ny;  // rows past ny (last workgroup) mirror row ny and store nothing
  const bool act = (seg ? (i >= 3 && i < 3 + seg && bx * seg + i - 3 < nx) : i < nx) && vrow;
  const bool lds_in = seg ? true : i < nx;  // writes its values to the LDS rows
  const int ic = seg ? gcol : (i < nx ? i : 0);
  const int j = vrow ? jrow : g.ny, jp = j + 1;
  const int nyp = g.ny + 1;
  // the j-1 and j+1 rows of a row inside the workgroup are its neighbours'
  // rows: their values come through LDS, and only the workgroup's edge rows
  // load them from HBM
  const bool lo_lds = r > 0;
  const bool hi_lds = r + 1 < nr && jrow + 1 <= g.ny;
  const int k0 = bz * kc, k1 = min(k0 + kc, nz);
  const long long pl = g.plane;
  const double *M = g.met;
  const int W = g.nyh;
  const double rdyp = M[M_RDYP * W + j], rdyc = M[M_RDYC * W + j];
  const double IY1 = M[M_YINT * W + j], IY2 = 1.0 - IY1;
  const double IY3 = M[M_YINT * W + jp], IY4 = 1.0 - IY3;
  // the row's y-Laplacian weights, read once: a load inside the loop could
  // not move across the per-plane barrier and would wait on every load in
  // flight before it
  const double ap2c = M[M_AP2C * W + j], ac2c = M[M_AC2C * W + j], am2c = M[M_AM2C * W + j];
  const double ap2p = M[M_AP2P * W + j], ac2p = M[M_AC2P * W + j], am2p = M[M_AM2P * W + j];
  const double dyp = M[M_DYP * W + j];
  const double ap2s = SC ? M[M_AP2CSC * W + j] : 0.0, ac2s = SC ? M[M_AC2CSC * W + j] : 0.0,
               am2s = SC ? M[M_AM2CSC * W + j] : 0.0;
  // The 4th-order weights (src/m_MeshAndMetries.f90:86-101) are symmetric or
  // antisymmetric in their stored values too (dxc4 = -dxc1, dxc3 = -dxc2,
  // dxxc5 = dxxc1, dxxc4 = dxxc2, the same in z), and the interpolation
  // weights are the exact constants -1/16 and 9/16: only the distinct values
  // are held, as the loop otherwise runs out of scalar registers.
  constexpr double c1 = -0.0625, c2 = 0.5625, c3 = 0.5625, c4 = -0.0625;
  const double ex1 = g.dxc1, ex2 = g.dxc2, ez1 = g.dzc1, ez2 = g.dzc2;
  const double exx1 = g.dxxc1, exx2 = g.dxxc2, exx3 = g.dxxc3, ezz1 = g.dzzc1, ezz2 = g.dzzc2, ezz3 = g.dzzc3;
  const double ucrf = cf.ucrf;
  const bool top = jp == nyp;  // uy interpolations on row ny+1 are zero (src/m_TScheme.f90:157-160)
  // LDS slots of the x neighbours i-3..i+3: wrapped in a whole row; in a
  // segment the halo holds them (clamped for halo threads, whose results at
  // those slots are never read)
  int nb[7];
#pragma unroll
  for (int s = 0; s < 7; ++s) {
    int q = (seg ? i : ic) + s - 3;
    if (seg) q = q < 0 ? 0 : (q >= lw ? lw - 1 : q);
    else q = q < 0 ? q + nx : (q >= nx ? q - nx : q);
    nb[s] = q;
  }
  // LDS row of quantity `row` for workgroup row rr
  auto LR = [&](int row, int rr) { return (row * nr + rr) * lw; };
  auto X = [&](int row, int d) { return xr[LR(row, r) + nb[d + 3]]; };
  // ux(j-1) and uy(j+1) with their x neighbours: the neighbouring rows' ux and
  // uy rows inside the workgroup, else this row's own copies
  const int uxm_row = lo_lds ? LR(X_UX, r - 1) : LR(X_UXM, r), uyp_row = hi_lds ? LR(X_UY, r + 1) : LR(X_UYP, r);
  auto XUXM = [&](int d) { return xr[uxm_row + nb[d + 3]]; };
  auto XUYP = [&](int d) { return xr[uyp_row + nb[d + 3]]; };
  auto wz = [&](int kk) { return kk < 0 ? kk + nz : (kk >= nz ? kk - nz : kk); };
  auto at = [&](int kk) { return (long long)wz(kk) * pl; };
  // byte offsets of this column at rows j, j-1, j+1 inside a plane: loads and
  // stores address a wave-uniform plane pointer plus a 32-bit lane offset
  const unsigned bj = 8u * (unsigned)(ic + nx * j), bjm = bj - 8u * nx, bjp = bj + 8u * nx;
  const int pbytes = (int)(8 * pl);
  auto ldp = [&](const double *a, int kk, unsigned off) { return bld(plane_rsrc(a + at(kk), pbytes), off); };

  double wux[7], wuy[7], wuz[7], wsc[7], wp[4], wuyp[4], wuzm[4];
  double ux_m, ux_p, uy_m, uz_p, p_m, sc_m = 0.0, sc_p = 0.0;
#pragma unroll
  for (int s = 0; s < 7; ++s) {
    wux[s] = ldp(ux, k0 + s - 3, bj);
    wuy[s] = ldp(uy, k0 + s - 3, bj);
    wuz[s] = ldp(uz, k0 + s - 3, bj);
    wsc[s] = SC ? ldp(sc, k0 + s - 3, bj) : 0.0;
  }
#pragma unroll
  for (int s = 0; s < 4; ++s) {
    wp[s] = ldp(p, k0 + s - 2, bj);
    wuyp[s] = ldp(uy, k0 + s - 2, bjp);
    wuzm[s] = ldp(uz, k0 + s - 1, bjm);
  }
  {
    ux_m = ldp(ux, k0, bjm);
    ux_p = ldp(ux, k0, bjp);
    uy_m = ldp(uy, k0, bjm);
    uz_p = ldp(uz, k0, bjp);
    p_m = ldp(p, k0, bjm);
    if (SC) {
      sc_m = ldp(sc, k0, bjm);
      sc_p = ldp(sc, k0, bjp);
    }
  }
  // queues: Q(k') = uxhz * uzhx for k' = k-1..k+2, R(k') = uzhx * (uxhz + ucrf)
  // for k' = k..k+2, uzhx from the uz rows of planes k0-1..k0+1 first
  double Q[4], R[3];
  auto uxhz_at = [&](int s) {  // uxhz at plane k-3+s+2 from the ux window entries s..s+3
    return c1 * wux[s] + c2 * wux[s + 1] + c3 * wux[s + 2] + c4 * wux[s + 3];
  };
  if (lds_in) {
    xr[LR(0, r) + i] = wuz[2];
    xr[LR(1, r) + i] = wuz[3];
    xr[LR(2, r) + i] = wuz[4];
  }
  __syncthreads();
  {
    double zx[3];
#pragma unroll
    for (int q = 0; q < 3; ++q) zx[q] = c1 * X(q, -2) + c2 * X(q, -1) + c3 * X(q, 0) + c4 * X(q, 1);
    const double hz0 = uxhz_at(0), hz1 = uxhz_at(1), hz2 = uxhz_at(2);  // planes k0-1, k0, k0+1
    Q[0] = hz0 * zx[0];
    Q[1] = hz1 * zx[1];
    Q[2] = hz2 * zx[2];
    R[0] = zx[1] * (hz1 + ucrf);
    R[1] = zx[2] * (hz2 + ucrf);
  }
  double s3 = 0.0;
  // the old right-hand sides of a plane are loaded with the previous plane's
  // window entries, a plane ahead of their use
  auto ldh = [&](const double *a, int kk) { return bld(plane_rsrc(a + (long long)kk * pl, pbytes), bj); };
  double hxo = ldh(hx, k0), hyo = ldh(hy, k0), hzo = ldh(hz, k0), hco = SC ? ldh(hc, k0) : 0.0;
  for (int k = k0; k < k1; ++k) {
    const long long kp = (long long)k * pl;
    __syncthreads();  // the previous plane's reads are done
    if (lds_in) {
      xr[LR(X_UX, r) + i] = wux[3];
      xr[LR(X_UY, r) + i] = wuy[3];
      xr[LR(X_UZ, r) + i] = wuz[3];
      xr[LR(X_UZ2, r) + i] = wuz[5];
      xr[LR(X_P, r) + i] = wp[2];
      if (!lo_lds) xr[LR(X_UXM, r) + i] = ux_m;
      if (!hi_lds) xr[LR(X_UYP, r) + i] = wuyp[2];
      xr[LR(X_R, r) + i] = R[0];
      if (SC) xr[LR(X_SC, r) + i] = wsc[3];
      if (nr > 1) {
        xr[LR(X_UYF, r) + i] = wuy[5];
        xr[LR(X_UZF, r) + i] = wuz[6];
      }
    }
    __syncthreads();
    // this plane's j-1 / j+1 values of the rows inside the workgroup, and the
    // next entries of the uy(j+1) and uz(j-1) windows
    double nwuyp = 0.0, nwuzm = 0.0;
    if (lo_lds) {
      ux_m = xr[LR(X_UX, r - 1) + nb[3]];
      uy_m = xr[LR(X_UY, r - 1) + nb[3]];
      p_m = xr[LR(X_P, r - 1) + nb[3]];
      if (SC) sc_m = xr[LR(X_SC, r - 1) + nb[3]];
      nwuzm = xr[LR(X_UZF, r - 1) + nb[3]];
    }
    if (hi_lds) {
      ux_p = xr[LR(X_UX, r + 1) + nb[3]];
      uz_p = xr[LR(X_UZ, r + 1) + nb[3]];
      if (SC) sc_p = xr[LR(X_SC, r + 1) + nb[3]];
      nwuyp = xr[LR(X_UYF, r + 1) + nb[3]];
    }
    // the products two planes ahead
    {
      const double zx2 = c1 * X(X_UZ2, -2) + c2 * X(X_UZ2, -1) + c3 * X(X_UZ2, 0) + c4 * X(X_UZ2, 1);
      const double hz2 = uxhz_at(3);  // plane k+2
      Q[3] = hz2 * zx2;
      R[2] = zx2 * (hz2 + ucrf);
    }
    double xu[7], xy[7];
#pragma unroll
    for (int s = 0; s < 7; ++s) {
      xu[s] = s == 3 ? wux[3] : X(X_UX, s - 3);
      xy[s] = s == 3 ? wuy[3] : X(X_UY, s - 3);
    }
    const double ux0 = wux[3], uy0 = wuy[3], uz0 = wuz[3];

    // ---- RhsX (src/m_TScheme.f90:169-204)
    double rx, cx, d22q1, dcq13x;
    {
      const double uxhyc = IY1 * ux_m + IY2 * ux0;
      const double uxhyp = IY3 * ux0 + IY4 * ux_p;
      const double a0 = c1 * xu[0] + c2 * xu[1] + c3 * xu[2] + c4 * xu[3];
      const double a1 = c1 * xu[1] + c2 * xu[2] + c3 * xu[3] + c4 * xu[4];
      const double a2 = c1 * xu[2] + c2 * xu[3] + c3 * xu[4] + c4 * xu[5];
      const double a3 = c1 * xu[3] + c2 * xu[4] + c3 * xu[5] + c4 * xu[6];
      const double h11 = ex1 * a0 * (a0 + ucrf) + ex2 * a1 * (a1 + ucrf) + (-ex2) * a2 * (a2 + ucrf) +
                         (-ex1) * a3 * (a3 + ucrf);
      const double uyhx_p =
          top ? 0.0 : c1 * XUYP(-2) + c2 * XUYP(-1) + c3 * wuyp[2] + c4 * XUYP(1);
      const double uyhx_c = c1 * xy[1] + c2 * xy[2] + c3 * xy[3] + c4 * xy[4];
      const double h12 = (uyhx_p * uxhyp - uyhx_c * uxhyc) * rdyp;
      const double h13 = ez1 * Q[0] + ez2 * Q[1] + (-ez2) * Q[2] + (-ez1) * Q[3];
      const double d11 =
          exx1 * xu[1] + exx2 * xu[2] + exx3 * xu[3] + exx2 * xu[4] + exx1 * xu[5];
      d22q1 = ap2c * ux_p + ac2c * ux0 + am2c * ux_m;
      const double d33 =
          ezz1 * wux[1] + ezz2 * wux[2] + ezz3 * wux[3] + ezz2 * wux[4] + ezz1 * wux[5];
      dcq13x = d11 + d33;
      cx = -h11 - h12 - h13 + cf.xnu * dcq13x + cf.gx;
      const double gp = ex1 * X(X_P, -2) + ex2 * X(X_P, -1) + (-ex2) * wp[2] + (-ex1) * X(X_P, 1);
      rx = cf.gamma * cx + cf.theta * hxo - cf.alpha * gp + 2.0 * cf.beta * d22q1;
    }
    if (act) {
      if (cf.s3_mode == 1) s3 += (dcq13x + d22q1) * dyp;
      else if (cf.s3_mode == 2) s3 += ux0 * dyp;
    }

    // ---- RhsY (:206-236)
    double ry, cy;
    {
      const double qs = 0.25 * rdyc;
      const double uxhym = IY1 * XUXM(-1) + IY2 * xu[2] + ucrf;
      const double uxhyc = IY1 * ux_m + IY2 * xu[3] + ucrf;
      const double uxhyp = IY1 * XUXM(1) + IY2 * xu[4] + ucrf;
      const double uxhyu = IY1 * XUXM(2) + IY2 * xu[5] + ucrf;
      const double uzhym = IY1 * wuzm[0] + IY2 * wuz[2];
      const double uzhyc = IY1 * wuzm[1] + IY2 * wuz[3];
      const double uzhyp = IY1 * wuzm[2] + IY2 * wuz[4];
      const double uzhyu = IY1 * wuzm[3] + IY2 * wuz[5];
      const double y0 = c1 * xy[0] + c2 * xy[1] + c3 * xy[2] + c4 * xy[3];
      const double y1 = c1 * xy[1] + c2 * xy[2] + c3 * xy[3] + c4 * xy[4];
      const double y2 = c1 * xy[2] + c2 * xy[3] + c3 * xy[4] + c4 * xy[5];
      const double y3 = c1 * xy[3] + c2 * xy[4] + c3 * xy[5] + c4 * xy[6];
      const double h21 = ex1 * y0 * uxhym + ex2 * y1 * uxhyc + (-ex2) * y2 * uxhyp + (-ex1) * y3 * uxhyu;
      const double s1 = wuyp[2] + uy0;
      const double s2 = uy0 + uy_m;
      const double h22 = (s1 * s1 - s2 * s2) * qs;
      const double z0 = c1 * wuy[0] + c2 * wuy[1] + c3 * wuy[2] + c4 * wuy[3];
      const double z1 = c1 * wuy[1] + c2 * wuy[2] + c3 * wuy[3] + c4 * wuy[4];
      const double z2 = c1 * wuy[2] + c2 * wuy[3] + c3 * wuy[4] + c4 * wuy[5];
      const double z3 = c1 * wuy[3] + c2 * wuy[4] + c3 * wuy[5] + c4 * wuy[6];
      const double h23 = ez1 * z0 * uzhym + ez2 * z1 * uzhyc + (-ez2) * z2 * uzhyp + (-ez1) * z3 * uzhyu;
      const double d11 =
          exx1 * xy[1] + exx2 * xy[2] + exx3 * xy[3] + exx2 * xy[4] + exx1 * xy[5];
      const double d22 = ap2p * wuyp[2] + ac2p * uy0 + am2p * uy_m;
      const double d33 =
          ezz1 * wuy[1] + ezz2 * wuy[2] + ezz3 * wuy[3] + ezz2 * wuy[4] + ezz1 * wuy[5];
      cy = -h21 - h22 - h23 + cf.xnu * (d11 + d33) + cf.gy;
      const double gp = (wp[2] - p_m) * rdyc;
      ry = cf.gamma * cy + cf.theta * hyo - cf.alpha * gp + 2.0 * cf.beta * d22;
    }

    // ---- RhsZ (:238-267)
    double rz, cz;
    {
      const double uzhyc = IY1 * wuzm[1] + IY2 * uz0;
      const double uzhyp = IY3 * uz0 + IY4 * uz_p;
      const double h31 = ex1 * X(X_R, -1) + ex2 * R[0] + (-ex2) * X(X_R, 1) + (-ex1) * X(X_R, 2);
      const double uyhz_p = top ? 0.0 : c1 * wuyp[0] + c2 * wuyp[1] + c3 * wuyp[2] + c4 * wuyp[3];
      const double uyhz_c = c1 * wuy[1] + c2 * wuy[2] + c3 * wuy[3] + c4 * wuy[4];
      const double h32 = (uyhz_p * uzhyp - uyhz_c * uzhyc) * rdyp;
      const double b0 = c1 * wuz[0] + c2 * wuz[1] + c3 * wuz[2] + c4 * wuz[3];
      const double b1 = c1 * wuz[1] + c2 * wuz[2] + c3 * wuz[3] + c4 * wuz[4];
      const double b2 = c1 * wuz[2] + c2 * wuz[3] + c3 * wuz[4] + c4 * wuz[5];
      const double b3 = c1 * wuz[3] + c2 * wuz[4] + c3 * wuz[5] + c4 * wuz[6];
      const double h33 = ez1 * b0 * b0 + ez2 * b1 * b1 + (-ez2) * b2 * b2 + (-ez1) * b3 * b3;
      const double d11 = exx1 * X(X_UZ, -2) + exx2 * X(X_UZ, -1) + exx3 * uz0 + exx2 * X(X_UZ, 1) +
                         exx1 * X(X_UZ, 2);
      const double d22 = ap2c * uz_p + ac2c * uz0 + am2c * wuzm[1];
      const double d33 =
          ezz1 * wuz[1] + ezz2 * wuz[2] + ezz3 * wuz[3] + ezz2 * wuz[4] + ezz1 * wuz[5];
      cz = -h31 - h32 - h33 + cf.xnu * (d11 + d33) + cf.gz;
      const double gp = ez1 * wp[0] + ez2 * wp[1] + (-ez2) * wp[2] + (-ez1) * wp[3];
      rz = cf.gamma * cz + cf.theta * hzo - cf.alpha * gp + 2.0 * cf.beta * d22;
    }

    double rc = 0.0, cc = 0.0;
    if (SC) {
      // ---- scalar transport (:302-357)
      double xs[7];
#pragma unroll
      for (int s = 0; s < 7; ++s) xs[s] = s == 3 ? wsc[3] : X(X_SC, s - 3);
      const double sh_m = c1 * xs[0] + c2 * xs[1] + c3 * xs[2] + c4 * xs[3];
      const double x0 = c1 * xs[1] + c2 * xs[2] + c3 * xs[3] + c4 * xs[4];
      const double sh_p = c1 * xs[2] + c2 * xs[3] + c3 * xs[4] + c4 * xs[5];
      const double sh_u = c1 * xs[3] + c2 * xs[4] + c3 * xs[5] + c4 * xs[6];
      const double sz_m = c1 * wsc[0] + c2 * wsc[1] + c3 * wsc[2] + c4 * wsc[3];
      const double z0 = c1 * wsc[1] + c2 * wsc[2] + c3 * wsc[3] + c4 * wsc[4];
      const double sz_p = c1 * wsc[2] + c2 * wsc[3] + c3 * wsc[4] + c4 * wsc[5];
      const double sz_u = c1 * wsc[3] + c2 * wsc[4] + c3 * wsc[5] + c4 * wsc[6];
      const double shyc = IY1 * sc_m + IY2 * wsc[3];
      const double shyp = IY3 * wsc[3] + IY4 * sc_p;
      const double e1 = cf.gev[0] * x0, e2 = cf.gev[1] * shyc, e3 = cf.gev[2] * z0;
      rx = rx + cf.gamma * e1;
      ry = ry + cf.gamma * e2;
      rz = rz + cf.gamma * e3;
      cx = cx + e1;
      cy = cy + e2;
      cz = cz + e3;
      const double f0 = cf.fvv[0], f1 = cf.fvv[1], f2 = cf.fvv[2];
      const double hc1 = ex1 * (xu[2] + f0) * sh_m + ex2 * (xu[3] + f0) * x0 + (-ex2) * (xu[4] + f0) * sh_p +
                         (-ex1) * (xu[5] + f0) * sh_u;
      const double hc2 = ((wuyp[2] + f1) * shyp - (uy0 + f1) * shyc) * rdyp;
      const double hc3 = ez1 * (wuz[2] + f2) * sz_m + ez2 * (wuz[3] + f2) * z0 + (-ez2) * (wuz[4] + f2) * sz_p +
                         (-ez1) * (wuz[5] + f2) * sz_u;
      const double d11 =
          exx1 * xs[1] + exx2 * xs[2] + exx3 * xs[3] + exx2 * xs[4] + exx1 * xs[5];
      const double d22 = ap2s * sc_p + ac2s * wsc[3] + am2s * sc_m;
      const double d33 =
          ezz1 * wsc[1] + ezz2 * wsc[2] + ezz3 * wsc[3] + ezz2 * wsc[4] + ezz1 * wsc[5];
      cc = -hc1 - hc2 - hc3 + cf.sdiff * (d11 + d33);
      rc = cf.gamma * cc + cf.theta * hco + 2.0 * cf.beta_sc * d22;
    }

    if (act) {
      bst(plane_rsrc(rhsx + kp, pbytes), bj, rx);
      bst(plane_rsrc(rhsy + kp, pbytes), bj, ry);
      bst(plane_rsrc(rhsz + kp, pbytes), bj, rz);
      bst(plane_rsrc(hx + kp, pbytes), bj, cx);
      bst(plane_rsrc(hy + kp, pbytes), bj, cy);
      bst(plane_rsrc(hz + kp, pbytes), bj, cz);
      if (SC) {
        bst(plane_rsrc(rhsc + kp, pbytes), bj, rc);
        bst(plane_rsrc(hc + kp, pbytes), bj, cc);
      }
    }

    // advance one plane: shift the windows and queues, load the new entries
#pragma unroll
    for (int s = 0; s < 6; ++s) {
      wux[s] = wux[s + 1];
      wuy[s] = wuy[s + 1];
      wuz[s] = wuz[s + 1];
      if (SC) wsc[s] = wsc[s + 1];
    }
#pragma unroll
    for (int s = 0; s < 3; ++s) {
      wp[s] = wp[s + 1];
      wuyp[s] = wuyp[s + 1];
      wuzm[s] = wuzm[s + 1];
      Q[s] = Q[s + 1];
    }
    R[0] = R[1];
    R[1] = R[2];
    {
      wux[6] = ldp(ux, k + 4, bj);
      wuy[6] = ldp(uy, k + 4, bj);
      wuz[6] = ldp(uz, k + 4, bj);
      if (SC) wsc[6] = ldp(sc, k + 4, bj);
      wp[3] = ldp(p, k + 2, bj);
      // the edge rows load their outside neighbour row; the others take it
      // from LDS in the next plane
      if (lo_lds) {
        wuzm[3] = nwuzm;
      } else {
        wuzm[3] = ldp(uz, k + 3, bjm);
        ux_m = ldp(ux, k + 1, bjm);
        uy_m = ldp(uy, k + 1, bjm);
        p_m = ldp(p, k + 1, bjm);
        if (SC) sc_m = ldp(sc, k + 1, bjm);
      }
      if (hi_lds) {
        wuyp[3] = nwuyp;
      } else {
        wuyp[3] = ldp(uy, k + 2, bjp);
        ux_p = ldp(ux, k + 1, bjp);
        uz_p = ldp(uz, k + 1, bjp);
        if (SC) sc_p = ldp(sc, k + 1, bjp);
      }
      const int kn = k + 1 < k1 ? k + 1 : k;  // the last plane reloads its own
      hxo = ldh(hx, kn);
      hyo = ldh(hy, kn);
      hzo = ldh(hz, kn);
      if (SC) hco = ldh(hc, kn);
    }
  }
  if (cf.s3_mode) {
    // fixed-order pairwise tree over the workgroup's threads
    const int t = r * (int)blockDim.x + i;
    __syncthreads();
    xr[t] = s3;
    __syncthreads();
    for (int n = (int)(blockDim.x * blockDim.y); n > 1;) {
      const int h = (n + 1) / 2;
      if (t < n - h) xr[t] += xr[t + h];
      __syncthreads();
      n = h;
    }
    if (t == 0) s3part[wt] = xr[0];
  }
}

// constant flow rate (:269-300): sums the partials in a fixed order and stores
// the shift added to RhsX and the PrGradAver increment
__global__ void __launch_bounds__(kReduceThreads) k_flowrate(const double *__restrict__ part, int n,
                                                            double *__restrict__ scal, int scalar_build, double xnu,
                                                            double ubulk, double ucrf, double nxnz, double yly,
                                                            double alpha, double alphaC, double dt) {
  const double tot = reduce_partials<false>(part, n);
  if (threadIdx.x == 0) {
    if (scalar_build) {
      const double s3 = tot / nxnz / yly;
      const double dpm = ubulk - ucrf - s3;
      scal[S_DP1NS] = alphaC * dpm;
      scal[S_PRGRAD] = scal[S_PRGRAD] + dpm * alphaC / dt;
    } else {
      const double dpm = xnu * tot / nxnz / yly;
      scal[S_DP1NS] = -(alpha * dpm);
      scal[S_PRGRAD] = scal[S_PRGRAD] + dpm * alphaC;
    }
  }
}

// RhsX += shift over the interior (the standalone clcRhsXYZ result)
__global__ void k_add_shift(double *__restrict__ a, const double *__restrict__ scal, Geo g) {
  const long long total = (long long)g.nx * g.ny * g.nz;
  const double s = scal[S_DP1NS];
  for (long long q = (long long)blockIdx.x * blockDim.x + threadIdx.x; q < total; q += (long long)gridDim.x * blockDim.x) {
    const long long i = q % g.nx, r = q / g.nx;
    const long long j = r % g.ny + 1, k = r / g.ny;
    const long long o = i + (long long)g.nx * (j + (long long)g.nyh * k);
    a[o] = a[o] + s;
  }
}

// ---- clcPrSrc (:48-77) ------------------------------------------------------
__global__ void __launch_bounds__(256) k_prsrc(Geo g, const double *__restrict__ ux, const double *__restrict__ uy,
                                              const double *__restrict__ uz, double *__restrict__ prsrc,
                                              double sudtal, double *__restrict__ part) {
  __shared__ double sh[256];
  const int i = blockIdx.x * 64 + threadIdx.x;
  const int j = blockIdx.y * 4 + threadIdx.y + 1;
  const int k = blockIdx.z;
  double dm = 0.0;
  if (i < g.nx && j <= g.ny) {
    const Ngh N(g, i, k);
    const double rdyp = g.met[M_RDYP * g.nyh + j];
    const double dudx = g.dxc1 * N(ux, -1, j, 0) + g.dxc2 * N(ux, 0, j, 0) + g.dxc3 * N(ux, 1, j, 0) +
                        g.dxc4 * N(ux, 2, j, 0);
    const double dvdy = (N(uy, 0, j + 1, 0) - N(uy, 0, j, 0)) * rdyp;
    const double dwdz = g.dzc1 * N(uz, 0, j, -1) + g.dzc2 * N(uz, 0, j, 0) + g.dzc3 * N(uz, 0, j, 1) +
                        g.dzc4 * N(uz, 0, j, 2);
    const double rdiv = dudx + dvdy + dwdz;
    dm = fabs(rdiv);
    if (prsrc) prsrc[N.ii[3] + N.nxl * j + N.kk[3]] = sudtal * rdiv;
  }
  const double t = block_max(dm, sh);
  if (threadIdx.x == 0 && threadIdx.y == 0) part[blockIdx.x + gridDim.x * (blockIdx.y + gridDim.y * blockIdx.z)] = t;
}

__global__ void __launch_bounds__(kReduceThreads) k_final_max(const double *__restrict__ part, int n,
                                                             double *__restrict__ out) {
  const double t = reduce_partials<true>(part, n);
  if (threadIdx.x == 0) out[0] = t;
}

__global__ void __launch_bounds__(kReduceThreads) k_final_sum(const double *__restrict__ part, int n,
                                                             double *__restrict__ out) {
  const double t = reduce_partials<false>(part, n);
  if (threadIdx.x == 0) out[0] = t;
}

// ---- FluidVelUpdate (:82-107) and PressureUpdate (:512-532) ---------------
// wall values of the velocity boundary conditions (m_BC_and_Halo)
struct BcArgs {
  int mode, flow_type;
  double uxb_m, uxb_p, uyb_p, uzb_m, uzb_p;
  double r1m, r2m, r1p, r2p;
};

// bc: the threads of rows 1 and ny then write the ghost rows of
// SetBC_and_UpdateHalo_pr (p, src/m_BC_and_Halo.f90:115-120) and
// SetBC_and_UpdateHalo (velocities, :32-58, uy's bottom face from uy_ym) from
// the values they have just formed: the boundary pass of the substep folded in.
// The workgroup (64 columns x 4 rows) marches k over a chunk of planes with
// phi's z stencil k-2 .. k+1 in registers, so each phi value is read once from
// HBM by its own column; the x and y neighbours of phi are other lanes' rows
// of the same plane, read through the cache.  Every load of plane k+1 is
// issued before plane k is formed.  Runs of `run` consecutive tiles (row
// groups fastest) go to one XCD, so a tile's edge rows are read through the L2
// that its neighbour tiles fill (4608x140x700: 7.25 ms against 7.64 with x
// tiles fastest).
template <bool VEL, bool PR, bool BC>
__global__ void __launch_bounds__(256) k_vel_p_update(Geo g, const double *__restrict__ phi, double *__restrict__ ux,
                                                     double *__restrict__ uy, double *__restrict__ uz,
                                                     double *__restrict__ p, double alpha, double beta, BcArgs b,
                                                     const double *__restrict__ uy_ym, int kc, int run) {
  const int nwg = (int)(gridDim.x * gridDim.y * gridDim.z);
  const int wt = xcd_run((int)(blockIdx.x + gridDim.x * (blockIdx.y + gridDim.y * blockIdx.z)), nwg, run);
  const int by = wt % (int)gridDim.y, bx = (wt / (int)gridDim.y) % (int)gridDim.x,
            bz = wt / (int)(gridDim.x * gridDim.y);
  const int i = bx * 64 + threadIdx.x;
  const int j = by * 4 + threadIdx.y + 1;
  if (i >= g.nx || j > g.ny) return;
  const int nx = g.nx, nz = g.nz;
  const int k0 = bz * kc, k1 = min(k0 + kc, nz);
  const long long pl = g.plane;
  const int W = g.nyh;
  const int im2 = wrapi(i - 2, nx), im1 = wrapi(i - 1, nx), ip1 = wrapi(i + 1, nx);
  const long long rj = (long long)nx * j;
  const double rdyc = g.met[M_RDYC * W + j];
  const double bp = -beta * g.met[M_AP2PH * W + j];
  const double bcc = -beta * g.met[M_AC2PH * W + j] + 1.0;
  const double bm = -beta * g.met[M_AM2PH * W + j];
  auto ph = [&](int kk) { return phi[i + rj + (long long)wrapi(kk, nz) * pl]; };
  // the values of one plane that the update reads besides the phi window
  struct Plane {
    double x, y, z, p, phm, php, xm2, xm1, xp1;
  };
  auto load = [&](int k) {
    Plane q{};
    const long long pk = (long long)k * pl, o = i + rj + pk;
    q.phm = phi[o - nx];
    if (PR) q.php = phi[o + nx];
    if (VEL) {
      q.xm2 = phi[im2 + rj + pk];
      q.xm1 = phi[im1 + rj + pk];
      q.xp1 = phi[ip1 + rj + pk];
      q.x = ux[o];
      q.y = uy[o];
      q.z = uz[o];
    }
    if (PR) q.p = p[o];
    return q;
  };
  double w0 = ph(k0 - 2), w1 = ph(k0 - 1), w2 = ph(k0), w3 = ph(k0 + 1);
  Plane cur = load(k0);
  for (int k = k0; k < k1; ++k) {
    const Plane nxt = load(k + 1 < k1 ? k + 1 : k);  // the last plane reloads its own
    const double w4 = ph(k + 2);
    const long long o = i + rj + (long long)k * pl;
    double vx = 0.0, vy = 0.0, vz = 0.0, vp = 0.0;
    if (VEL) {
      const double dpdx = g.dxc1 * cur.xm2 + g.dxc2 * cur.xm1 + g.dxc3 * w2 + g.dxc4 * cur.xp1;
      const double dpdy = (w2 - cur.phm) * rdyc;
      const double dpdz = g.dzc1 * w0 + g.dzc2 * w1 + g.dzc3 * w2 + g.dzc4 * w3;
      vx = cur.x - alpha * dpdx;
      vy = cur.y - alpha * dpdy;
      vz = cur.z - alpha * dpdz;
      ux[o] = vx;
      uy[o] = vy;
      uz[o] = vz;
    }
    if (PR) {
      vp = cur.p + bp * cur.php + bcc * w2 + bm * cur.phm;
      p[o] = vp;
    }
    if (BC) {
      const long long dn = (long long)nx;  // one row
      if (j == 1) {
        p[o - dn] = vp;
        ux[o - dn] = b.uxb_m * 2.0 - vx;
        uy[o] = uy_ym[i + (long long)nx * k];
        uz[o - dn] = b.uzb_m * 2.0 - vz;
      }
      if (j == g.ny) {
        p[o + dn] = vp;
        if (b.flow_type == 1) {
          ux[o + dn] = b.uxb_p * 2.0 - vx;
          uy[o + dn] = b.uyb_p;
          uz[o + dn] = b.uzb_p * 2.0 - vz;
        } else {
          ux[o + dn] = vx;
          uy[o + dn] = 0.0;
          uz[o + dn] = vz;
        }
      }
    }
    cur = nxt;
    w0 = w1;
    w1 = w2;
    w2 = w3;
    w3 = w4;
  }
}

// ---- wall boundary conditions (m_BC_and_Halo) on the (i, k) plane ---------
// mode 0: SetBC_and_UpdateHalo, 1: ...ForPrSrc, 2: _pr, 3: _scalar, 4: Update_uy_ym

__global__ void k_bc(Geo g, BcArgs b, double *__restrict__ a0, double *__restrict__ a1, double *__restrict__ a2,
                     double *__restrict__ pl0, double *__restrict__ pl1) {
  const long long n = (long long)g.nx * g.nz;
  for (long long q = (long long)blockIdx.x * blockDim.x + threadIdx.x; q < n; q += (long long)gridDim.x * blockDim.x) {
    const long long i = q % g.nx, k = q / g.nx;
    const long long base = i + g.plane * k;
    const long long r0 = base, r1 = base + g.nx, rn = base + (long long)g.nx * g.ny,
                    rnp = base + (long long)g.nx * (g.ny + 1);
    switch (b.mode) {
      case 5:  // SetBC_and_UpdateHalo_pr on pl1 (:115-120), then case 0 in the same pass
        pl1[r0] = pl1[r1];
        pl1[rnp] = pl1[rn];
        [[fallthrough]];
      case 0:  // src/m_BC_and_Halo.f90:32-58
        if (b.flow_type == 1) {
          a0[rnp] = b.uxb_p * 2.0 - a0[rn];
          a1[rnp] = b.uyb_p;
          a2[rnp] = b.uzb_p * 2.0 - a2[rn];
        } else {
          a0[rnp] = a0[rn];
          a1[rnp] = 0.0;
          a2[rnp] = a2[rn];
        }
        a0[r0] = b.uxb_m * 2.0 - a0[r1];
        a1[r1] = pl0[q];
        a2[r0] = b.uzb_m * 2.0 - a2[r1];
        break;
      case 1:  // :78-98
        a1[rnp] = (b.flow_type == 1) ? b.uyb_p : 0.0;
        a1[r1] = pl0[q];
        break;
      case 2:  // :115-120
        a0[r0] = a0[r1];
        a0[rnp] = a0[rn];
        break;
      case 3:  // :137-176
        a0[r0] = b.r1m * a0[r1] + b.r2m;
        a0[rnp] = b.r1p * a0[rn] + b.r2p;
        break;
      case 4: {  // Update_uy_ym: duy = uy_ym(old) -> uy_ym = 0 -> duy = uy_ym - duy
        const double old = pl0[q];
        pl0[q] = 0.0;
        pl1[q] = pl0[q] - old;
      } break;
      default: break;
    }
  }
}

// ---- dump_visu (src/m_IOAndVisu.f90:182-385) ----------------------------------
// The cell-centre field `which` at the points decomp_2d_write_every keeps
// (every iskip/jskip/kskip-th from the first, src/m_Decomp2d.f90:1232-1389),
// x fastest, as FP64 (R = double) or FP32 (R = float): 0 ux + uCRF, 1 uy,
// 2 uz, 3 p, 4-6 the vorticity components, 7 its magnitude, 8 the scalar.
// One thread per written point: x from the block's x index, the kept y row and
// z plane from the grid's y and z.
// CubicRoot (src/m_Tools.f90:494-570) for x^3 + b x^2 + c x + d, returning its
// Root(2): the middle root of three distinct real roots (the reference's
// ordering branches reduce to that), the repeated root when two coincide
// (0 <= t < 1e-12).  false: a complex pair.
__device__ bool lambda2_middle_root(double b, double c, double d, double &mid) {
  const double SMALL = 1.0e-12, third = 0.3333333333333333333333, pi = 3.141592653589793238462643383279502884;
  const double p = c - third * b * b;
  const double q = (2.0 * b * b * b - 9.0 * b * c + 27.0 * d) / 27.0;
  const double t = q * q / 4.0 + p * p * p / 27.0;
  if (t < 0.0) {
    const double sh = -third * b;
    double rho = pow(-p, 1.5) / sqrt(27.0);
    const double eta = acos(0.5 * q / rho);
    rho = -2.0 * pow(rho, third);
    const double r1 = cos(third * eta) * rho + sh;
    const double r2 = cos(third * (2.0 * pi + eta)) * rho + sh;
    const double r3 = cos(third * (4.0 * pi + eta)) * rho + sh;
    if (r1 < r2 && r1 < r3) mid = r2 < r3 ? r2 : r3;
    else if (r1 > r2 && r1 > r3) mid = r2 < r3 ? r3 : r2;
    else mid = r1;
    return true;
  }
  if (t < SMALL) {
    double rho = 0.5 * q;
    rho = rho < 0.0 ? pow(-rho, third) : -pow(rho, third);
    mid = -rho - third * b;
    return true;
  }
  return false;
}

template <typename R>
__global__ void __launch_bounds__(64) k_visu(Geo g, int which, const double *__restrict__ ux,
                                             const double *__restrict__ uy, const double *__restrict__ uz,
                                             const double *__restrict__ p, const double *__restrict__ sc,
                                             double ucrf, int iskip, int jskip, int kskip, int nxo,
                                             R *__restrict__ out, double *__restrict__ err) {
  const int io = blockIdx.x * 64 + threadIdx.x, jo = blockIdx.y, ko = blockIdx.z;
  if (io >= nxo) return;
  const int nx = g.nx, nz = g.nz;
  const int i = io * iskip, j = 1 + jo * jskip, k = ko * kskip;
  const int ip = i + 1 < nx ? i + 1 : i + 1 - nx, im = i > 0 ? i - 1 : nx - 1;
  const int kp = k + 1 < nz ? k + 1 : k + 1 - nz, km = k > 0 ? k - 1 : nz - 1;
  const int jp = j + 1, jm = j - 1;
  auto at = [&](const double *a, int ii, int jj, int kk) { return a[ii + (long long)nx * jj + g.plane * kk]; };
  const double half = 0.5, quarter = 0.25;
  double v;
  if (which == 0) {
    v = half * (at(ux, ip, j, k) + at(ux, i, j, k)) + ucrf;
  } else if (which == 1) {
    v = half * (at(uy, i, jp, k) + at(uy, i, j, k));
  } else if (which == 2) {
    v = half * (at(uz, i, j, kp) + at(uz, i, j, k));
  } else if (which == 3) {
    v = at(p, i, j, k);
  } else if (which == 8) {
    v = at(sc, i, j, k);
  } else {
    const double *M = g.met;
    const int W = g.nyh;
    const double cac1 = M[M_RDYC * W + j], cac2 = M[M_RDYC * W + jp], cac12 = cac1 - cac2;
    const double rdx = g.rdx, rdz = g.rdz;
    const double dudy = ((at(ux, ip, jp, k) + at(ux, i, jp, k)) * cac2 + (at(ux, ip, j, k) + at(ux, i, j, k)) * cac12 -
                         (at(ux, ip, jm, k) + at(ux, i, jm, k)) * cac1) * quarter;
    const double dudz = (at(ux, ip, j, kp) + at(ux, i, j, kp) - at(ux, ip, j, km) - at(ux, i, j, km)) * rdz * quarter;
    const double dvdx = (at(uy, ip, jp, k) - at(uy, im, jp, k) + at(uy, ip, j, k) - at(uy, im, j, k)) * rdx * quarter;
    const double dvdz = (at(uy, i, jp, kp) + at(uy, i, j, kp) - at(uy, i, jp, km) - at(uy, i, j, km)) * rdz * quarter;
    const double dwdx = (at(uz, ip, j, kp) - at(uz, im, j, kp) + at(uz, ip, j, k) - at(uz, im, j, k)) * rdx * quarter;
    const double dwdy = ((at(uz, i, jp, kp) + at(uz, i, jp, k)) * cac2 + (at(uz, i, j, kp) + at(uz, i, j, k)) * cac12 -
                         (at(uz, i, jm, kp) + at(uz, i, jm, k)) * cac1) * quarter;
    const double vx = dwdy - dvdz, vy = dudz - dwdx, vz = dvdx - dudy;
    if (which <= 7) {
      v = which == 4 ? vx : which == 5 ? vy : which == 6 ? vz : sqrt(vx * vx + vy * vy + vz * vz);
    } else {
      const double dudx = (at(ux, ip, j, k) - at(ux, i, j, k)) * rdx;
      const double dvdy = (at(uy, i, jp, k) - at(uy, i, j, k)) * M[M_RDYP * W + j];
      const double dwdz = (at(uz, i, j, kp) - at(uz, i, j, k)) * rdz;
      if (which == 9) {  // Clc_Q_vor (src/m_Tools.f90:307-342)
        const double udiv = dudx + dvdy + dwdz;
        v = half * (udiv * udiv - dudx * dudx - dvdy * dvdy - dwdz * dwdz) - dudy * dvdx - dudz * dwdx - dvdz * dwdy;
      } else {  // Clc_lamda2 (:347-442): strain and rotation tensors carry a factor 2
        const double o12 = dvdx - dudy, o23 = dwdy - dvdz, o31 = dudz - dwdx;
        const double e11 = dudx + dudx, e12 = dudy + dvdx, e13 = dudz + dwdx;
        const double e22 = dvdy + dvdy, e23 = dvdz + dwdy, e33 = dwdz + dwdz;
        const double l12 = e11 * e12 + e12 * e22 + e13 * e23 + o31 * o23;
        const double l13 = e11 * e13 + e12 * e23 + e13 * e33 + o12 * o23;
        const double l23 = e12 * e13 + e22 * e23 + e23 * e33 + o12 * o31;
        const double l11 = e11 * e11 + e12 * e12 + e13 * e13 - o12 * o12 - o31 * o31;
        const double l22 = e12 * e12 + e22 * e22 + e23 * e23 - o12 * o12 - o23 * o23;
        const double l33 = e13 * e13 + e23 * e23 + e33 * e33 - o31 * o31 - o23 * o23;
        const double rb = -(l11 + l22 + l33);
        const double rc = l11 * l22 + l22 * l33 + l33 * l11 - l12 * l12 - l13 * l13 - l23 * l23;
        const double rd = l12 * l12 * l33 + l13 * l13 * l22 + l23 * l23 * l11 - l11 * l22 * l33 - 2.0 * l12 * l23 * l13;
        double mid;
        if (lambda2_middle_root(rb, rc, rd, mid)) {
          v = quarter * mid;
        } else {
          v = __builtin_nan("");
          err[0] = 1.0;  // the reference aborts ("lamda2 wrong"); the host reports it
        }
      }
    }
  }
  out[io + (long long)nxo * (jo + (long long)gridDim.y * ko)] = (R)v;
}

int launch_visu(ptc_ctx *c, int which, const double *ux, const double *uy, const double *uz, const double *p,
                const double *sc, int iskip, int jskip, int kskip, int real_bytes, void *dev_out) {
  const int nxo = (c->nx + iskip - 1) / iskip, nyo = (c->ny + jskip - 1) / jskip, nzo = (c->nz + kskip - 1) / kskip;
  const dim3 grid((unsigned)((nxo + 63) / 64), (unsigned)nyo, (unsigned)nzo);
  if (real_bytes == 8)
    hipLaunchKernelGGL(k_visu<double>, grid, dim3(64), 0, c->stream, c->geo, which, ux, uy, uz, p, sc, c->prm.ucrf,
                       iskip, jskip, kskip, nxo, (double *)dev_out, c->d_scal + kSlotVisuErr);
  else
    hipLaunchKernelGGL(k_visu<float>, grid, dim3(64), 0, c->stream, c->geo, which, ux, uy, uz, p, sc, c->prm.ucrf,
                       iskip, jskip, kskip, nxo, (float *)dev_out, c->d_scal + kSlotVisuErr);
  PTC_HIP(hipGetLastError());
  return PTC_OK;
}

// ---- monitors (m_Tools) -------------------------------------------------------
// which: 0 CFL max (:24-49), 1 divergence max (:54-81), 2..4 |ux|,|uy|,|uz| max (:86-110),
//        5 sum ux*dyp (:256-275), 6..8 sums dyp*(ux, c, c^2) (:281-301)
__global__ void __launch_bounds__(256) k_monitor(Geo g, int which, const double *__restrict__ ux,
                                                const double *__restrict__ uy, const double *__restrict__ uz,
                                                const double *__restrict__ sc, double *__restrict__ part) {
  __shared__ double sh[256];
  const int i = blockIdx.x * 64 + threadIdx.x;
  const int j = blockIdx.y * 4 + threadIdx.y + 1;
  const int k = blockIdx.z;
  double v = 0.0;
  if (i < g.nx && j <= g.ny) {
    const Ngh N(g, i, k);
    const int W = g.nyh;
    switch (which) {
      case 0:
        v = fabs(N(ux, 0, j, 0) + N(ux, 1, j, 0)) * g.rdx + fabs(N(uy, 0, j, 0) + N(uy, 0, j + 1, 0)) * g.met[M_RDYP * W + j] +
            fabs(N(uz, 0, j, 0) + N(uz, 0, j, 1)) * g.rdz;
        break;
      case 1: {
        const double dudx = g.dxc1 * N(ux, -1, j, 0) + g.dxc2 * N(ux, 0, j, 0) + g.dxc3 * N(ux, 1, j, 0) +
                            g.dxc4 * N(ux, 2, j, 0);
        const double dvdy = (N(uy, 0, j + 1, 0) - N(uy, 0, j, 0)) * g.met[M_RDYP * W + j];
        const double dwdz = g.dzc1 * N(uz, 0, j, -1) + g.dzc2 * N(uz, 0, j, 0) + g.dzc3 * N(uz, 0, j, 1) +
                            g.dzc4 * N(uz, 0, j, 2);
        v = fabs(dudx + dvdy + dwdz);
      } break;
      case 2: v = fabs(N(ux, 0, j, 0)); break;
      case 3: v = fabs(N(uy, 0, j, 0)); break;
      case 4: v = fabs(N(uz, 0, j, 0)); break;
      case 5:
      case 6: v = N(ux, 0, j, 0) * g.met[M_DYP * W + j]; break;
      case 7: v = g.met[M_DYP * W + j] * N(sc, 0, j, 0); break;
      case 8: {
        const double s = N(sc, 0, j, 0);
        v = g.met[M_DYP * W + j] * s * s;
      } break;
      default: break;
    }
  }
  const double t = (which <= 4) ? block_max(v, sh) : block_sum(v, sh);
  if (threadIdx.x == 0 && threadIdx.y == 0) part[blockIdx.x + gridDim.x * (blockIdx.y + gridDim.y * blockIdx.z)] = t;
}

// ---- host launchers ----------------------------------------------------------
static dim3 pt_grid(const ptc_ctx *c) {
  return dim3((unsigned)((c->nx + 63) / 64), (unsigned)((c->ny + 3) / 4), (unsigned)c->nz);
}
static int pt_nblocks(const ptc_ctx *c) {
  const dim3 gr = pt_grid(c);
  return (int)(gr.x * gr.y * gr.z);
}

// row-marching RHS: one workgroup per (row segment, group of rows j, k
// chunk).  Rows of at most 512 points are one segment (a 1024-thread
// workgroup would spill); longer rows go in segments of tx - 6 columns with a
// 3-column halo either side.  A workgroup of ty rows takes the j-1 and j+1
// rows of its inner rows from LDS, so only its edge rows read a neighbour row
// from HBM: at 4608x140x700 (SC) 128-thread segments of two rows ran 16.4-16.7
// ms against 17.2 ms for one row of 256 threads, and four rows of 64 threads
// 16.3 ms (the halo columns cost what the edge rows save); 512-thread
// workgroups of two rows of 256 ran 17.3 ms, of four rows of 128 (one per
// CU, half the edge rows of two) 15.7 against 16.1 ms for two rows of 128.  The k chunks
// are many enough for about 96 wavefronts per CU over the launch (more rounds
// of resident workgroups, so the partial last one costs less: the RHS alone
// 15.3 against 15.4 ms at 4608x140x700 and 3.01 against 3.07 ms at 1536 x 192
// x 256 with 48) and at least 16 planes long, or shorter where that leaves the
// chip short of workgroups (z_chunks).
constexpr int kRhsRowsMaxNx = 512;
constexpr int kRhsSegThreads = 128;
constexpr int kRhsXcdRun = 32;
constexpr int kRhsMinPlanes = 16;
struct RhsShape {
  int tx, ty, seg, run, minpl, nseg, nyb, nch;
};

// Number of z chunks over `tiles` workgroup tiles per plane range: at most
// `want` (the launch's wavefront target), each at least `minpl` planes long.
// Each chunk starts its z window afresh (a prologue of several planes), so
// chunks are long (16 planes) where the grid offers enough tiles; where that
// leaves fewer than 1.5 workgroups per CU (small grids) the automatic length
// (shrink) halves, down to 2 planes, until it does not.  RK substep, 64 x 48
// x 64 channel: 98.5 us with 16-plane chunks, 81.3 with 4; 96 x 64 x 96 with a
// scalar: 133 against 119 us with 8; 128 x 96 x 128 and larger unchanged.
static int z_chunks(int nz, long long tiles, int want, int minpl, bool shrink, int n_cu, int *minpl_used) {
  int nch = std::max(1, std::min(want, nz / minpl));
  while (shrink && minpl > 2 && 2 * tiles * nch < 3LL * n_cu) {
    minpl /= 2;
    nch = std::max(1, std::min(want, nz / minpl));
  }
  *minpl_used = minpl;
  return nch;
}
static RhsShape rhs_shape(const ptc_ctx *c) {
  RhsShape s{};
  if (c->nx <= kRhsRowsMaxNx) {
    s.tx = (c->nx + 63) / 64 * 64;
    s.seg = 0;
  } else {
    s.tx = c->rhs_seg_req ? c->rhs_seg_req : kRhsSegThreads;
    s.seg = s.tx - 6;
  }
  // rows per workgroup: four for segmented rows (the RHS alone at
  // 4608x140x700 15.7 against 16.1 ms with two, at 1536x192x256 2.71 against
  // 2.74), two for whole rows of at most 128 points, else one
  const int ty = c->rhs_rows_req ? c->rhs_rows_req : (s.seg ? 4 : (s.tx <= 128 ? 2 : 1));
  s.ty = std::max(1, std::min({ty, 512 / s.tx, c->ny}));
  s.run = c->rhs_run_req ? c->rhs_run_req : kRhsXcdRun;
  s.nseg = s.seg ? (c->nx + s.seg - 1) / s.seg : 1;
  s.nyb = (c->ny + s.ty - 1) / s.ty;
  const int wpb = s.tx * s.ty / 64;
  const long long tiles = (long long)s.nyb * s.nseg;
  const int want = (int)((96LL * c->n_cu + tiles * wpb - 1) / (tiles * wpb));
  s.nch = z_chunks(c->nz, tiles, want, c->rhs_minpl_req ? c->rhs_minpl_req : kRhsMinPlanes, c->rhs_minpl_req == 0,
                   c->n_cu, &s.minpl);
  return s;
}
void rhs_shape_in_use(const ptc_ctx *c, int *seg_threads, int *rows, int *run, int *min_planes) {
  const RhsShape s = rhs_shape(c);
  *seg_threads = s.tx;
  *rows = s.ty;
  *run = s.run;
  *min_planes = s.minpl;
}
static dim3 rhs_rows_grid(const ptc_ctx *c, int *kc) {
  const RhsShape sh = rhs_shape(c);
  *kc = (c->nz + sh.nch - 1) / sh.nch;
  return dim3((unsigned)sh.nseg, (unsigned)sh.nyb, (unsigned)((c->nz + *kc - 1) / *kc));
}
static int rhs_nblocks(const ptc_ctx *c) {
  int kc = 0;
  const dim3 gr = rhs_rows_grid(c, &kc);
  return (int)(gr.x * gr.y * gr.z);
}

int ensure_red(ptc_ctx *c, int n) {
  if (c->red_cap < n) {
    // a captured step (this context's or an ensemble's) holds the old address
    drop_step_graph(c);
    ++c->buf_gen;
    if (c->d_red) (void)hipFree(c->d_red);
    c->d_red = nullptr;
    c->red_cap = 0;
    PTC_HIP(hipMalloc(&c->d_red, sizeof(double) * (size_t)n));
    c->red_cap = n;
  }
  return PTC_OK;
}

int launch_rhs(ptc_ctx *c, const double *ux, const double *uy, const double *uz, const double *sc, double *rhsx,
               double *rhsy, double *rhsz, double *rhsc, double *hx, double *hy, double *hz, double *hc,
               const double *p, bool defer_flowrate) {
  RhsCoef cf{};
  cf.gamma = c->pmGamma;
  cf.theta = c->pmTheta;
  cf.alpha = c->pmAlpha;
  cf.beta = c->pmBeta;
  cf.beta_sc = c->pmBetaSc;
  cf.xnu = c->prm.xnu;
  cf.ucrf = c->prm.ucrf;
  cf.gx = c->prm.gravity[0];
  cf.gy = c->prm.gravity[1];
  cf.gz = c->prm.gravity[2];
  cf.sdiff = c->prm.sdiff;
  for (int q = 0; q < 3; ++q) {
    cf.gev[q] = c->prm.gev[q];
    cf.fvv[q] = c->prm.fvv[q];
  }
  cf.scalar = c->prm.scalar ? 1 : 0;
  cf.s3_mode = c->prm.is_ux_const ? (c->prm.scalar ? 2 : 1) : 0;
  const int nb = rhs_nblocks(c);
  PTC_TRY(ensure_red(c, nb));
  int kc = 0;
  const dim3 grid = rhs_rows_grid(c, &kc);
  const RhsShape sh = rhs_shape(c);
  const int nt = sh.tx * sh.ty;
  const int lw = sh.seg ? sh.tx : c->nx;  // LDS row width
  const size_t lds = sizeof(double) * (size_t)std::max(X_NROWS * sh.ty * lw, nt);
#define PTC_ROWS(SCAL, MAXT)                                                                                  \
  hipLaunchKernelGGL((k_rhs_rows<SCAL, MAXT>), grid, dim3(sh.tx, sh.ty), lds, c->stream, c->geo, cf, kc, ux, uy, \
                     uz, sc, p, rhsx, rhsy, rhsz, rhsc, hx, hy, hz, hc, c->d_red, sh.seg, lw, sh.run)
  if (nt <= 256) {
    if (cf.scalar) PTC_ROWS(true, 256);
    else PTC_ROWS(false, 256);
  } else {
    if (cf.scalar) PTC_ROWS(true, 512);
    else PTC_ROWS(false, 512);
  }
#undef PTC_ROWS
  PTC_HIP(hipGetLastError());
  if (cf.s3_mode && !defer_flowrate) {
    hipLaunchKernelGGL(k_flowrate, dim3(1), dim3(kReduceThreads), 0, c->stream, c->d_red, nb, c->d_scal, cf.scalar,
                       c->prm.xnu, c->prm.ubulk, c->prm.ucrf, (double)(c->nx * c->nz), c->yly, c->pmAlpha,
                       c->pmAlphaC, c->dt);
    PTC_HIP(hipGetLastError());
  }
  return PTC_OK;
}

int launch_add_shift(ptc_ctx *c, double *rhsx) {
  hipLaunchKernelGGL(k_add_shift, dim3(c->n_cu * 8), dim3(256), 0, c->stream, rhsx, c->d_scal, c->geo);
  PTC_HIP(hipGetLastError());
  return PTC_OK;
}

// what k_flowrate would take, for the Crank-Nicolson launch that forms the
// correction itself (after launch_rhs has written the partials)
int flow_fold(ptc_ctx *c, bool first, FlowFold *ff) {
  PTC_CHECK(c->prm.is_ux_const, PTC_E_STATE, "flow-rate correction without a constant flow rate");
  ff->part = c->d_red;
  ff->n = rhs_nblocks(c);
  PTC_CHECK(ff->n <= c->red_cap, PTC_E_STATE, "flow-rate partials larger than the reduction buffer");
  ff->scalar_build = c->prm.scalar ? 1 : 0;
  ff->first = first ? 1 : 0;
  ff->xnu = c->prm.xnu;
  ff->ubulk = c->prm.ubulk;
  ff->ucrf = c->prm.ucrf;
  ff->nxnz = (double)(c->nx * c->nz);
  ff->yly = c->yly;
  ff->alpha = c->pmAlpha;
  ff->alphaC = c->pmAlphaC;
  ff->dt = c->dt;
  ff->dp1ns = c->d_scal + S_DP1NS;
  ff->prgrad = c->d_scal + S_PRGRAD;
  return PTC_OK;
}

// reduction partials of the RHS and PrSrc kernels, allocated ahead of a
// captured step
int stencil_prepare(ptc_ctx *c) {
  const int nx = (int)fft2_x_fwd_blocks(c, c->ny);
  return ensure_red(c, std::max({pt_nblocks(c), rhs_nblocks(c), nx}));
}

int launch_divmax(ptc_ctx *c, int n) {
  hipLaunchKernelGGL(k_final_max, dim3(1), dim3(kReduceThreads), 0, c->stream, c->d_red, n, c->d_scal + S_DIVMAX);
  PTC_HIP(hipGetLastError());
  return PTC_OK;
}

// reduce: form the divergence maximum in the device slot (always when it is
// read back on the host)
int launch_prsrc(ptc_ctx *c, const double *ux, const double *uy, const double *uz, double *prsrc,
                 double *divmax_host, bool reduce) {
  PTC_TRY(ensure_red(c, pt_nblocks(c)));
  return launch_prsrc_part(c, ux, uy, uz, prsrc, divmax_host, reduce, c->d_red);
}

int prsrc_partials(const ptc_ctx *c) { return pt_nblocks(c); }

int launch_prsrc_part(ptc_ctx *c, const double *ux, const double *uy, const double *uz, double *prsrc,
                      double *divmax_host, bool reduce, double *part) {
  const int nb = pt_nblocks(c);
  hipLaunchKernelGGL(k_prsrc, pt_grid(c), dim3(64, 4), 0, c->stream, c->geo, ux, uy, uz, prsrc, 1.0 / c->pmAlpha,
                     part);
  PTC_HIP(hipGetLastError());
  if (!reduce && !divmax_host) return PTC_OK;
  hipLaunchKernelGGL(k_final_max, dim3(1), dim3(kReduceThreads), 0, c->stream, part, nb, c->d_scal + S_DIVMAX);
  PTC_HIP(hipGetLastError());
  if (divmax_host) {
    PTC_HIP(hipMemcpyAsync(c->h_pinned, c->d_scal + S_DIVMAX, sizeof(double), hipMemcpyDeviceToHost, c->stream));
    PTC_HIP(hipStreamSynchronize(c->stream));
    *divmax_host = c->h_pinned[0];
    c->divmax_prsrc = c->h_pinned[0];
  }
  return PTC_OK;
}

static BcArgs vel_bc_args(const ptc_ctx *c, int mode) {
  BcArgs b{};
  b.mode = mode;
  b.flow_type = c->prm.flow_type;
  b.uxb_m = c->prm.uxbc[2];
  b.uxb_p = c->prm.uxbc[3];
  b.uyb_p = c->prm.uybc[3];
  b.uzb_m = c->prm.uzbc[2];
  b.uzb_p = c->prm.uzbc[3];
  return b;
}

// z chunks of the velocity/pressure update: enough for about 96 wavefronts per
// CU over the launch (the velocity update alone 5.34 against 5.39 ms at
// 4608x140x700 and 0.863 against 0.892 ms at 1536x192x256 with 48), at least
// 16 planes each (each chunk re-reads a 3-plane window prologue), shorter on
// grids with few tiles (z_chunks)
constexpr int kVelXcdRun = 32;
static dim3 vel_grid(const ptc_ctx *c, int *kc) {
  const int gx = (c->nx + 63) / 64, gy = (c->ny + 3) / 4;
  const int want = (96 * c->n_cu + gx * gy * 4 - 1) / (gx * gy * 4);
  int minpl = 0;
  const int nch = z_chunks(c->nz, (long long)gx * gy, want, 16, true, c->n_cu, &minpl);
  *kc = (c->nz + nch - 1) / nch;
  return dim3((unsigned)gx, (unsigned)gy, (unsigned)((c->nz + *kc - 1) / *kc));
}

int launch_vel_p_update(ptc_ctx *c, const double *phi, double *ux, double *uy, double *uz, double *p) {
  const bool dv = ux && uy && uz, dp = p != nullptr;
  PTC_CHECK(dv || dp, PTC_E_ARG, "nothing to update");
  int kc = 0;
  const dim3 grid = vel_grid(c, &kc);
#define PTC_VP(V, P)                                                                                           \
  hipLaunchKernelGGL((k_vel_p_update<V, P, false>), grid, dim3(64, 4), 0, c->stream, c->geo, phi, ux, uy, uz, p, \
                     c->pmAlpha, c->pmBeta, BcArgs{}, nullptr, kc, kVelXcdRun)
  if (dv && dp) PTC_VP(true, true);
  else if (dv) PTC_VP(true, false);
  else PTC_VP(false, true);
#undef PTC_VP
  PTC_HIP(hipGetLastError());
  return PTC_OK;
}

int launch_vel_p_update_bc(ptc_ctx *c, const double *phi, double *ux, double *uy, double *uz, double *p,
                           const double *uy_ym) {
  int kc = 0;
  const dim3 grid = vel_grid(c, &kc);
  hipLaunchKernelGGL((k_vel_p_update<true, true, true>), grid, dim3(64, 4), 0, c->stream, c->geo, phi, ux, uy, uz, p,
                     c->pmAlpha, c->pmBeta, vel_bc_args(c, 5), uy_ym, kc, kVelXcdRun);
  PTC_HIP(hipGetLastError());
  return PTC_OK;
}

int launch_vel_update(ptc_ctx *c, const double *phi, double *ux, double *uy, double *uz) {
  return launch_vel_p_update(c, phi, ux, uy, uz, nullptr);
}

int launch_pressure_update(ptc_ctx *c, double *p, const double *phi) {
  return launch_vel_p_update(c, phi, nullptr, nullptr, nullptr, p);
}

static int launch_bc(ptc_ctx *c, const BcArgs &b, double *a0, double *a1, double *a2, double *p0, double *p1) {
  const long long n = (long long)c->nx * c->nz;
  const int nblk = (int)std::min<long long>((n + 255) / 256, (long long)c->n_cu * 8);
  hipLaunchKernelGGL(k_bc, dim3(std::max(1, nblk)), dim3(256), 0, c->stream, c->geo, b, a0, a1, a2, p0, p1);
  PTC_HIP(hipGetLastError());
  return PTC_OK;
}

int launch_bc_vel(ptc_ctx *c, double *ux, double *uy, double *uz, const double *uy_ym, bool prsrc_only) {
  BcArgs b{};
  b.mode = prsrc_only ? 1 : 0;
  b.flow_type = c->prm.flow_type;
  // uxBcValue(ym_dir = 3), (yp_dir = 4) -> indices 2, 3
  b.uxb_m = c->prm.uxbc[2];
  b.uxb_p = c->prm.uxbc[3];
  b.uyb_p = c->prm.uybc[3];
  b.uzb_m = c->prm.uzbc[2];
  b.uzb_p = c->prm.uzbc[3];
  return launch_bc(c, b, ux, uy, uz, const_cast<double *>(uy_ym), nullptr);
}

// the velocity wall fills and the pressure's Neumann ghosts in one launch
// (the end of a substep, src/m_ChannelSystem.f90:135-136)
int launch_bc_pr(ptc_ctx *c, double *p) {
  BcArgs b{};
  b.mode = 2;
  return launch_bc(c, b, p, nullptr, nullptr, nullptr, nullptr);
}

int scalar_bc_coefs(const ptc_ctx *c, double *r) {
  BcArgs b{};
  const double *v = c->prm.sbc_val;
  const double dyc1 = c->h_met[(size_t)M_DYC * (c->ny + 2) + 1];
  const double dycp = c->h_met[(size_t)M_DYC * (c->ny + 2) + c->ny + 1];
  // src/m_BC_and_Halo.f90:137-171
  switch (c->prm.sbc_opt[0]) {
    case -1: b.r1m = -1.0; b.r2m = 2.0 * v[0]; break;
    case -2: b.r1m = 1.0; b.r2m = -dyc1 * v[2]; break;
    case -3:
      b.r1m = (2.0 + v[0] * dyc1) / (2.0 - v[0] * dyc1);
      b.r2m = -2.0 * v[2] * dyc1 / (2.0 - v[0] * dyc1);
      break;
    default: set_error("ScalarBcOption(1) wrong !!!"); return PTC_E_ARG;
  }
  switch (c->prm.sbc_opt[1]) {
    case -1: b.r1p = -1.0; b.r2p = 2.0 * v[1]; break;
    case -2: b.r1p = 1.0; b.r2p = dycp * v[3]; break;
    case -3:
      b.r1p = (2.0 - v[1] * dycp) / (2.0 + v[1] * dycp);
      b.r2p = 2.0 * v[3] * dycp / (2.0 + v[1] * dycp);
      break;
    default: set_error("ScalarBcOption(2) wrong !!!"); return PTC_E_ARG;
  }
  r[0] = b.r1m;
  r[1] = b.r2m;
  r[2] = b.r1p;
  r[3] = b.r2p;
  return PTC_OK;
}

int launch_bc_scalar(ptc_ctx *c, double *s) {
  BcArgs b{};
  b.mode = 3;
  double r[4];
  PTC_TRY(scalar_bc_coefs(c, r));
  b.r1m = r[0];
  b.r2m = r[1];
  b.r1p = r[2];
  b.r2p = r[3];
  return launch_bc(c, b, s, nullptr, nullptr, nullptr, nullptr);
}

int launch_uy_ym(ptc_ctx *c, double *uy_ym, double *duy_ym) {
  BcArgs b{};
  b.mode = 4;
  return launch_bc(c, b, nullptr, nullptr, nullptr, uy_ym, duy_ym);
}

// result to host (synchronises)
int launch_monitor(ptc_ctx *c, int which, const double *ux, const double *uy, const double *uz, const double *s,
                   double *out) {
  const int nb = pt_nblocks(c);
  PTC_TRY(ensure_red(c, nb));
  hipLaunchKernelGGL(k_monitor, pt_grid(c), dim3(64, 4), 0, c->stream, c->geo, which, ux, uy, uz, s, c->d_red);
  PTC_HIP(hipGetLastError());
  if (which <= 4)
    hipLaunchKernelGGL(k_final_max, dim3(1), dim3(kReduceThreads), 0, c->stream, c->d_red, nb, c->d_scal + S_MON0);
  else
    hipLaunchKernelGGL(k_final_sum, dim3(1), dim3(kReduceThreads), 0, c->stream, c->d_red, nb, c->d_scal + S_MON0);
  PTC_HIP(hipGetLastError());
  PTC_HIP(hipMemcpyAsync(c->h_pinned, c->d_scal + S_MON0, sizeof(double), hipMemcpyDeviceToHost, c->stream));
  PTC_HIP(hipStreamSynchronize(c->stream));
  *out = c->h_pinned[0];
  return PTC_OK;
}

int set_prgrad(ptc_ctx *c, double v) {
  c->h_pinned[1] = v;
  PTC_HIP(hipMemcpyAsync(c->d_scal + S_PRGRAD, c->h_pinned + 1, sizeof(double), hipMemcpyHostToDevice, c->stream));
  PTC_HIP(hipStreamSynchronize(c->stream));
  return PTC_OK;
}

int get_device_scalar(ptc_ctx *c, int slot, double *out) {
  PTC_HIP(hipMemcpyAsync(c->h_pinned + 2, c->d_scal + slot, sizeof(double), hipMemcpyDeviceToHost, c->stream));
  PTC_HIP(hipStreamSynchronize(c->stream));
  *out = c->h_pinned[2];
  return PTC_OK;
}

}  // namespace ptc

namespace ptc {

// ---- clcStat (src/m_FlowCase.f90:283-413) ----------------------------------
// Block (chunk, j): plane j = 1..ny+1, planes k in [chunk*kc, chunk*kc + kc).
// Each thread sums the reference's SumVec terms over its points; the block
// reduces them in a fixed order into part[(j-1)][chunk][q].  Plane ny+1 has
// only the wall terms 27, 33 and 35 (1-based) of the reference.
template <bool SC>
__global__ void __launch_bounds__(256) k_stat_partial(Geo g, double ucrf, const double *__restrict__ ux,
                                                      const double *__restrict__ uy, const double *__restrict__ uz,
                                                      const double *__restrict__ pr, const double *__restrict__ sc,
                                                      int kc, double *__restrict__ part) {
  constexpr int NS = SC ? 40 : 35;
  __shared__ double sh[256];
  const int j = blockIdx.y + 1, ny = g.ny;
  const int k0 = blockIdx.x * kc, k1 = min(k0 + kc, g.nz);
  const double *met = g.met;
  const int nyh = g.nyh;
  const double IY1 = met[M_YINT * nyh + j], IY2 = 1.0 - IY1;
  const double cac = met[M_RDYC * nyh + j];
  const double I1 = g.ic1, I2 = g.ic2, I3 = g.ic3, I4 = g.ic4;
  const double rdxt = g.rdx / 12.0, rdzt = g.rdz / 12.0;
  double sv[NS];
#pragma unroll
  for (int q = 0; q < NS; ++q) sv[q] = 0.0;
  if (j <= ny) {
    const int jm = j - 1, jp = j + 1;
    const double cacU = met[M_RDYC * nyh + jp], caj = met[M_RDYP * nyh + j];
    const double ap2c = met[M_AP2C * nyh + j], ac2c = met[M_AC2C * nyh + j], am2c = met[M_AM2C * nyh + j];
    const double ap2p = met[M_AP2P * nyh + j], ac2p = met[M_AC2P * nyh + j], am2p = met[M_AM2P * nyh + j];
    for (int k = k0; k < k1; ++k)
      for (int i = threadIdx.x; i < g.nx; i += 256) {
        const Ngh n(g, i, k);
        auto U = [&](int di, int jj, int dk) { return n(ux, di, jj, dk); };
        auto V = [&](int di, int jj, int dk) { return n(uy, di, jj, dk); };
        auto W = [&](int di, int jj, int dk) { return n(uz, di, jj, dk); };
        auto Pp = [&](int di, int jj, int dk) { return n(pr, di, jj, dk); };
        const double uxloc = U(0, j, 0) + ucrf, uyloc = V(0, j, 0), uzloc = W(0, j, 0), prloc = Pp(0, j, 0);
        const double uxCell = I1 * U(-1, j, 0) + I2 * U(0, j, 0) + I3 * U(1, j, 0) + I4 * U(2, j, 0) + ucrf;
        const double uyCell = (uyloc + V(0, jp, 0)) * 0.5;
        const double uy_xc = I1 * V(-2, j, 0) + I2 * V(-1, j, 0) + I3 * V(0, j, 0) + I4 * V(1, j, 0);
        const double uy_xp = I1 * V(-2, jp, 0) + I2 * V(-1, jp, 0) + I3 * V(0, jp, 0) + I4 * V(1, jp, 0);
        const double uzCell = I1 * W(0, j, -1) + I2 * W(0, j, 0) + I3 * W(0, j, 1) + I4 * W(0, j, 2);
        const double prloc2 = IY1 * (I1 * Pp(-2, jm, 0) + I2 * Pp(-1, jm, 0) + I3 * Pp(0, jm, 0) + I4 * Pp(1, jm, 0)) +
                              IY2 * (I1 * Pp(-2, j, 0) + I2 * Pp(-1, j, 0) + I3 * Pp(0, j, 0) + I4 * Pp(1, j, 0));
        const double dudx = g.dxc1 * U(-1, j, 0) + g.dxc2 * U(0, j, 0) + g.dxc3 * U(1, j, 0) + g.dxc4 * U(2, j, 0);
        const double dudy = (U(0, j, 0) - U(0, jm, 0)) * cac;
        const double dudz = g.dzc1 * U(0, j, -2) + g.dzc2 * U(0, j, -1) + g.dzc3 * U(0, j, 0) + g.dzc4 * U(0, j, 1);
        const double dvdx = g.dxc1 * V(-2, j, 0) + g.dxc2 * V(-1, j, 0) + g.dxc3 * V(0, j, 0) + g.dxc4 * V(1, j, 0);
        const double dvdy = (V(0, jp, 0) - uyloc) * caj;
        const double dvdz = g.dzc1 * V(0, j, -2) + g.dzc2 * V(0, j, -1) + g.dzc3 * V(0, j, 0) + g.dzc4 * V(0, j, 1);
        const double dwdx = g.dxc1 * W(-2, j, 0) + g.dxc2 * W(-1, j, 0) + g.dxc3 * W(0, j, 0) + g.dxc4 * W(1, j, 0);
        const double dwdy = (uzloc - W(0, jm, 0)) * cac;
        const double dwdz = g.dzc1 * W(0, j, -1) + g.dzc2 * W(0, j, 0) + g.dzc3 * W(0, j, 1) + g.dzc4 * W(0, j, 2);
        auto dxx = [&](auto F) {
          return g.dxxc1 * F(-2, j, 0) + g.dxxc2 * F(-1, j, 0) + g.dxxc3 * F(0, j, 0) + g.dxxc4 * F(1, j, 0) +
                 g.dxxc5 * F(2, j, 0);
        };
        auto dzz = [&](auto F) {
          return g.dzzc1 * F(0, j, -2) + g.dzzc2 * F(0, j, -1) + g.dzzc3 * F(0, j, 0) + g.dzzc4 * F(0, j, 1) +
                 g.dzzc5 * F(0, j, 2);
        };
        const double dudxx = dxx(U), dudzz = dzz(U);
        const double dudyy = ap2c * U(0, jp, 0) + ac2c * U(0, j, 0) + am2c * U(0, jm, 0);
        const double dvdxx = dxx(V), dvdzz = dzz(V);
        const double dvdyy = ap2p * V(0, jp, 0) + ac2p * V(0, j, 0) + am2p * V(0, jm, 0);
        const double dwdxx = dxx(W), dwdzz = dzz(W);
        const double dwdyy = ap2c * W(0, jp, 0) + ac2c * W(0, j, 0) + am2c * W(0, jm, 0);
        const double vor_x = dwdy - dvdz, vor_y = dudz - dwdx, vor_z = dvdx - dudy;
        const double dudxC = (-U(2, j, 0) + 8.0 * U(1, j, 0) - 8.0 * U(-1, j, 0) + U(-2, j, 0)) * rdxt;
        const double dvdxU = g.dxc1 * V(-2, jp, 0) + g.dxc2 * V(-1, jp, 0) + g.dxc3 * V(0, jp, 0) + g.dxc4 * V(1, jp, 0);
        const double dudyU = (U(0, jp, 0) - U(0, j, 0)) * cacU;
        const double dudzC = (IY1 * (-U(0, jm, 2) + 8.0 * U(0, jm, 1) - 8.0 * U(0, jm, -1) + U(0, jm, -2)) +
                              IY2 * (-U(0, j, 2) + 8.0 * U(0, j, 1) - 8.0 * U(0, j, -1) + U(0, j, -2))) *
                             rdzt;
        auto c4z = [&](int di) { return -V(di, j, 2) + 8.0 * V(di, j, 1) - 8.0 * V(di, j, -1) + V(di, j, -2); };
        const double dvdzC = (c4z(-2) * I1 + c4z(-1) * I2 + c4z(0) * I3 + c4z(1) * I4) * rdzt;
        sv[0] += uxloc;
        sv[1] += uyloc;
        sv[2] += uzloc;
        sv[3] += prloc;
        sv[4] += uxloc * uxloc;
        sv[5] += uyloc * uyloc;
        sv[6] += uzloc * uzloc;
        sv[7] += prloc * prloc;
        sv[8] += uxCell * uyCell;
        sv[9] += uyCell * uzCell;
        sv[10] += uxCell * uzCell;
        sv[11] += uxCell * prloc;
        sv[12] += uyCell * prloc;
        sv[13] += uzCell * prloc;
        sv[14] += uxCell * uxCell * uyCell;
        sv[15] += uyloc * uyloc * uyloc;
        sv[16] += uzCell * uzCell * uyCell;
        sv[17] += uxCell * uyCell * uyCell;
        sv[18] += uxloc * uxloc * uxloc;
        sv[19] += uzloc * uzloc * uzloc;
        sv[20] += uxloc * uxloc * uxloc * uxloc;
        sv[21] += uyloc * uyloc * uyloc * uyloc;
        sv[22] += uzloc * uzloc * uzloc * uzloc;
        sv[23] += prloc * dudx;
        sv[24] += prloc * dvdy;
        sv[25] += prloc * dwdz;
        sv[26] += prloc2 * (dudy + dvdx);
        sv[27] += uxloc * (dudxx + dudyy + dudzz);
        sv[28] += uyloc * (dvdxx + dvdyy + dvdzz);
        sv[29] += uzloc * (dwdxx + dwdyy + dwdzz);
        sv[30] += (dudxC * (dvdxU + dvdx) + (dudyU + dudy) * (uy_xp - uy_xc) * caj) * 0.5;
        sv[31] += dudzC * dvdzC;
        sv[32] += vor_x * vor_x;
        sv[33] += vor_y * vor_y;
        sv[34] += vor_z * vor_z;
        if constexpr (SC) {
          const double scloc = n(sc, 0, j, 0);
          sv[35] += scloc;
          sv[36] += scloc * scloc;
          sv[37] += scloc * uxCell;
          sv[38] += scloc * uyCell;
          sv[39] += scloc * uzCell;
        }
      }
  } else {
    const int jm = j - 1;
    for (int k = k0; k < k1; ++k)
      for (int i = threadIdx.x; i < g.nx; i += 256) {
        const Ngh n(g, i, k);
        auto Pp = [&](int di, int jj) { return n(pr, di, jj, 0); };
        const double prloc2 = IY1 * (I1 * Pp(-2, jm) + I2 * Pp(-1, jm) + I3 * Pp(0, jm) + I4 * Pp(1, jm)) +
                              IY2 * (I1 * Pp(-2, j) + I2 * Pp(-1, j) + I3 * Pp(0, j) + I4 * Pp(1, j));
        const double dudy = (n(ux, 0, j, 0) - n(ux, 0, jm, 0)) * cac;
        const double dwdy = (n(uz, 0, j, 0) - n(uz, 0, jm, 0)) * cac;
        sv[26] += prloc2 * (dudy + 0.0);
        sv[32] += dwdy * dwdy;
        sv[34] += dudy * dudy;
      }
  }
  double *out = part + ((long long)blockIdx.y * gridDim.x + blockIdx.x) * NS;
#pragma unroll
  for (int q = 0; q < NS; ++q) {
    const double s = block_sum(sv[q], sh);
    if (threadIdx.x == 0) out[q] = s;
    __syncthreads();
  }
}

// SumStat(q, j) += inxz * (sum over chunks in order); one thread per (q, j)
__global__ void __launch_bounds__(256) k_stat_final(const double *__restrict__ part, int nchunk, int ns, int nyp,
                                                    double inxz, double *__restrict__ sumstat) {
  const int t = blockIdx.x * blockDim.x + threadIdx.x;
  if (t >= ns * nyp) return;
  const int q = t % ns, j = t / ns;
  double s = 0.0;
  for (int c = 0; c < nchunk; ++c) s += part[((long long)j * nchunk + c) * ns + q];
  sumstat[t] += s * inxz;
}

// PrGradsum += PrGradAver (:416), on the device
__global__ void k_stat_prgrad(double *scal) { scal[S_PRGSUM] += scal[S_PRGRAD]; }

int launch_stat(ptc_ctx *c, const double *ux, const double *uy, const double *uz, const double *pr,
                const double *sc) {
  const int ns = c->prm.scalar ? 40 : 35, nyp = c->ny + 1;
  if (!c->d_stat) {
    PTC_HIP(hipMalloc(&c->d_stat, sizeof(double) * (size_t)ns * nyp));
    PTC_TRY(fill_device(c, c->d_stat, (long long)ns * nyp, 0.0));
  }
  // chunks of k planes: ~8 blocks per CU over all planes
  const int want = std::max(1, (c->n_cu * 8) / nyp);
  const int nchunk = std::min(c->nz, want);
  const int kc = (c->nz + nchunk - 1) / nchunk;
  const int nblk = (c->nz + kc - 1) / kc;
  PTC_TRY(ensure_red(c, nblk * nyp * ns));
  if (c->prm.scalar)
    hipLaunchKernelGGL(k_stat_partial<true>, dim3(nblk, nyp), dim3(256), 0, c->stream, c->geo, c->prm.ucrf, ux, uy,
                       uz, pr, sc, kc, c->d_red);
  else
    hipLaunchKernelGGL(k_stat_partial<false>, dim3(nblk, nyp), dim3(256), 0, c->stream, c->geo, c->prm.ucrf, ux, uy,
                       uz, pr, sc, kc, c->d_red);
  PTC_HIP(hipGetLastError());
  const double inxz = 1.0 / ((double)c->nx * (double)c->nz);
  hipLaunchKernelGGL(k_stat_final, dim3((ns * nyp + 255) / 256), dim3(256), 0, c->stream, c->d_red, nblk, ns, nyp,
                     inxz, c->d_stat);
  PTC_HIP(hipGetLastError());
  hipLaunchKernelGGL(k_stat_prgrad, dim3(1), dim3(1), 0, c->stream, c->d_scal);
  PTC_HIP(hipGetLastError());
  ++c->nfstime;
  return PTC_OK;
}

int read_stat(ptc_ctx *c, double *out, double *prgradsum, int reset) {
  const int ns = c->prm.scalar ? 40 : 35, nyp = c->ny + 1;
  if (out) {
    if (c->d_stat) {
      PTC_HIP(hipMemcpyAsync(out, c->d_stat, sizeof(double) * (size_t)ns * nyp, hipMemcpyDeviceToHost, c->stream));
    } else {
      std::fill(out, out + (size_t)ns * nyp, 0.0);
    }
  }
  PTC_HIP(hipMemcpyAsync(c->h_pinned + 3, c->d_scal + S_PRGSUM, sizeof(double), hipMemcpyDeviceToHost, c->stream));
  PTC_HIP(hipStreamSynchronize(c->stream));
  if (prgradsum) *prgradsum = c->h_pinned[3];
  if (reset) {
    // src/m_FlowCase.f90:944: nfstime = 0; SumStat = 0; PrGradsum = 0
    if (c->d_stat) PTC_TRY(fill_device(c, c->d_stat, (long long)ns * nyp, 0.0));
    PTC_TRY(fill_device(c, c->d_scal + S_PRGSUM, 1, 0.0));
    c->nfstime = 0;
  }
  return PTC_OK;
}

}  // namespace ptc
