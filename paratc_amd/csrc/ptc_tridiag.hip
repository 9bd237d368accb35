// ptc_tridiag.hip -- batched wall-normal tridiagonal solves on gfx950.
//
// Two algorithms for the same systems:
//
//  * Partitioned (default).  A line of ny rows is split into P chunks, one
//    wavefront per chunk, one lane per line (64 consecutive lines per
//    workgroup, so every row load is a contiguous 64-lane access).  Each lane
//    keeps its chunk in registers and runs the reference's "modified Thomas"
//    (src/m_Poisson.f90:497-531): afterwards row j reads
//        x_j + a_j x_first + c_j x_last = f_j ,
//    with the chunk's first row coupled to the previous chunk's last row.  The
//    2P first/last rows form a tridiagonal system with unit diagonal
//    (InverseTridiagonal_reduce, src/m_Poisson.f90:1045-1071) solved by wave 0
//    in LDS; every lane then reconstructs its interior rows
//    (src/m_Poisson.f90:665-675).  One HBM read and one write per point.
//  * Thomas.  One thread per line: InverseTridiagonal (src/m_Tools.f90:224-251)
//    with the forward-sweep factor kept in a scratch array.  It moves ~2.5x the
//    bytes and exists as the reference algorithm and autotuner candidate.
//
// Poisson lines (kx, kz) carry complex right-hand sides and the diagonal
// ac2ph(j) + WaveNumX(kx) + WaveNumZ(kz) (src/m_Poisson.f90:383-430); the
// (0,0) line has row j=1 replaced by phi=0 (src/m_Poisson.f90:399-404).
// Crank-Nicolson lines (i, k) of clcU1Hat/U2Hat/U3Hat/clcScalar
// (src/m_TScheme.f90:365-506) are real with j-only coefficients.
#include "ptc_internal.hpp"

namespace ptc {

__device__ __forceinline__ double vscale(double s, double v) { return s * v; }
__device__ __forceinline__ double2 vscale(double s, double2 v) { return make_double2(s * v.x, s * v.y); }
__device__ __forceinline__ double vsub(double a, double b) { return a - b; }
__device__ __forceinline__ double2 vsub(double2 a, double2 b) { return make_double2(a.x - b.x, a.y - b.y); }
__device__ __forceinline__ double vzero(double) { return 0.0; }
__device__ __forceinline__ double2 vzero(double2) { return make_double2(0.0, 0.0); }

// local elimination of one chunk (rows r = 0..m-1, reference row j = s + r + 1)
template <int MMAX, typename V, typename G>
__device__ __forceinline__ void local_eliminate(V *f, double *a, double *c, int m, int s, const G &getc) {
#pragma unroll
  for (int r = 0; r < MMAX; ++r) {
    if (r < m) {
      double aj, bj, cj;
      getc(s + r + 1, aj, bj, cj, f[r]);
      if (r < 2) {
        const double inv = 1.0 / bj;
        a[r] = aj * inv;
        c[r] = cj * inv;
        f[r] = vscale(inv, f[r]);
      } else {
        const double rr = 1.0 / (bj - aj * c[r - 1]);
        f[r] = vscale(rr, vsub(f[r], vscale(aj, f[r - 1])));
        c[r] = rr * cj;
        a[r] = -rr * aj * a[r - 1];
      }
    }
  }
#pragma unroll
  for (int r = MMAX - 3; r >= 1; --r) {
    if (r <= m - 3) {
      f[r] = vsub(f[r], vscale(c[r], f[r + 1]));
      a[r] = a[r] - c[r] * a[r + 1];
      c[r] = -c[r] * c[r + 1];
    }
  }
  const double rr = 1.0 / (1.0 - a[1] * c[0]);
  f[0] = vscale(rr, vsub(f[0], vscale(c[0], f[1])));
  a[0] = rr * a[0];
  c[0] = -rr * c[0] * c[1];
}

// reduced system of 2P rows per line in LDS, solved by wave 0, then every lane
// rebuilds its chunk.  RA/RC/RF are [2P][64].
template <int MMAX, typename V>
__device__ __forceinline__ void reduce_and_rebuild(V *f, const double *a, const double *c, int m, int P, int p,
                                                   int lane, double *RA, double *RC, V *RF) {
  V fl = f[0];
  double al = a[0], cl = c[0];
#pragma unroll
  for (int r = 1; r < MMAX; ++r)
    if (r == m - 1) {
      fl = f[r];
      al = a[r];
      cl = c[r];
    }
  RA[(2 * p) * 64 + lane] = a[0];
  RC[(2 * p) * 64 + lane] = c[0];
  RF[(2 * p) * 64 + lane] = f[0];
  RA[(2 * p + 1) * 64 + lane] = al;
  RC[(2 * p + 1) * 64 + lane] = cl;
  RF[(2 * p + 1) * 64 + lane] = fl;
  __syncthreads();
  if (p == 0) {
    const int n2 = 2 * P;
    // forward (vecm starts at 1: unit diagonal)
    V fp = RF[lane];
    double cp = RC[lane];
    for (int q = 1; q < n2; ++q) {
      const double aq = RA[q * 64 + lane];
      const double den = 1.0 - aq * cp;
      const double inv = 1.0 / den;
      cp = RC[q * 64 + lane] * inv;
      fp = vscale(inv, vsub(RF[q * 64 + lane], vscale(aq, fp)));
      RC[q * 64 + lane] = cp;
      RF[q * 64 + lane] = fp;
    }
    V x = fp;
    for (int q = n2 - 2; q >= 0; --q) {
      x = vsub(RF[q * 64 + lane], vscale(RC[q * 64 + lane], x));
      RF[q * 64 + lane] = x;
    }
  }
  __syncthreads();
  const V xs = RF[(2 * p) * 64 + lane];
  const V xe = RF[(2 * p + 1) * 64 + lane];
#pragma unroll
  for (int r = 0; r < MMAX; ++r) {
    if (r < m) {
      if (r == 0) f[r] = xs;
      else if (r == m - 1) f[r] = xe;
      else f[r] = vsub(vsub(f[r], vscale(a[r], xs)), vscale(c[r], xe));
    }
  }
}

// ---- Poisson y-solve, partitioned ----------------------------------------
template <int MMAX>
__global__ void __launch_bounds__(1024) k_tri_poisson_part(double2 *__restrict__ spec, const double *__restrict__ met,
                                                          int nyh, const double *__restrict__ wnx,
                                                          const double *__restrict__ wnz, int ny, int pitch, int h,
                                                          int P) {
  extern __shared__ double smem[];
  double *RA = smem;
  double *RC = RA + 2 * P * 64;
  double2 *RF = reinterpret_cast<double2 *>(RC + 2 * P * 64);
  const int lane = threadIdx.x & 63;
  const int p = __builtin_amdgcn_readfirstlane(threadIdx.x >> 6);
  const int kx = blockIdx.x * 64 + lane, kz = blockIdx.y;
  const bool valid = kx <= h;
  const int s = (p * ny) / P, m = ((p + 1) * ny) / P - s;
  const double lam = wnx[valid ? kx : 0] + wnz[kz];
  const bool pin = (kx == 0 && kz == 0);
  double2 *col = spec + (long long)kz * ny * pitch + kx;
  double2 f[MMAX];
  double a[MMAX], c[MMAX];
#pragma unroll
  for (int r = 0; r < MMAX; ++r)
    if (r < m) f[r] = valid ? col[(long long)(s + r) * pitch] : make_double2(0.0, 0.0);
  const double *am = met + M_AM2PH * nyh, *ac = met + M_AC2PH * nyh, *ap = met + M_AP2PH * nyh;
  auto getc = [&](int j, double &aj, double &bj, double &cj, double2 &fj) {
    aj = am[j];
    cj = ap[j];
    bj = ac[j] + lam;
    if (pin && j == 1) {
      aj = 0.0;
      bj = 1.0;
      cj = 0.0;
      fj = make_double2(0.0, 0.0);
    }
  };
  local_eliminate<MMAX>(f, a, c, m, s, getc);
  reduce_and_rebuild<MMAX>(f, a, c, m, P, p, lane, RA, RC, RF);
  if (valid) {
#pragma unroll
    for (int r = 0; r < MMAX; ++r)
      if (r < m) col[(long long)(s + r) * pitch] = f[r];
  }
}

// ---- Poisson y-solve, Thomas (reference InverseTridiagonal) --------------
__global__ void __launch_bounds__(256) k_tri_poisson_thomas(double2 *__restrict__ spec, double *__restrict__ cs,
                                                           const double *__restrict__ met, int nyh,
                                                           const double *__restrict__ wnx,
                                                           const double *__restrict__ wnz, int ny, int pitch, int h,
                                                           int nz) {
  const int kx = blockIdx.x * 64 + threadIdx.x;
  const int kz = blockIdx.y * blockDim.y + threadIdx.y;
  if (kx > h || kz >= nz) return;
  const double lam = wnx[kx] + wnz[kz];
  const bool pin = (kx == 0 && kz == 0);
  const double *am = met + M_AM2PH * nyh, *ac = met + M_AC2PH * nyh, *ap = met + M_AP2PH * nyh;
  double2 *col = spec + (long long)kz * ny * pitch + kx;
  double *ccol = cs + (long long)kz * ny * pitch + kx;
  // row 1 (pinned on the (0,0) line)
  double vecm = pin ? 1.0 : ac[1] + lam;
  double cprev = pin ? 0.0 : ap[1];
  double2 fprev = col[0];
  if (pin) fprev = make_double2(0.0, 0.0);
  fprev = make_double2(fprev.x / vecm, fprev.y / vecm);
  col[0] = fprev;
  for (int j = 2; j <= ny; ++j) {
    const long long o = (long long)(j - 1) * pitch;
    const double aj = am[j];
    const double arr = cprev / vecm;
    vecm = (ac[j] + lam) - aj * arr;
    const double2 fj = col[o];
    fprev = make_double2((fj.x - aj * fprev.x) / vecm, (fj.y - aj * fprev.y) / vecm);
    col[o] = fprev;
    ccol[o] = arr;
    cprev = ap[j];
  }
  double2 fnext = fprev;
  for (int j = ny - 1; j >= 1; --j) {
    const long long o = (long long)(j - 1) * pitch;
    const double arr = ccol[o + pitch];
    double2 fj = col[o];
    fj = make_double2(fj.x - arr * fnext.x, fj.y - arr * fnext.y);
    col[o] = fj;
    fnext = fj;
  }
}

// ---- Crank-Nicolson y-solves (u += du), partitioned -----------------------
// kind: 0 = u/w (ForCN tables), 1 = v (row 1 Dirichlet), 2 = scalar (ScCN tables)
template <int MMAX>
__global__ void __launch_bounds__(1024) k_tri_cn_part(double *__restrict__ u, const double *__restrict__ rhs,
                                                     const double *__restrict__ duy_ym,
                                                     const double *__restrict__ shift, const double *__restrict__ met,
                                                     int nyh, int rowA, int rowB, int rowC, double beta, int kind,
                                                     int nx, int ny, int P) {
  extern __shared__ double smem[];
  double *RA = smem;
  double *RC = RA + 2 * P * 64;
  double *RF = RC + 2 * P * 64;
  const int lane = threadIdx.x & 63;
  const int p = __builtin_amdgcn_readfirstlane(threadIdx.x >> 6);
  const int i = blockIdx.x * 64 + lane, k = blockIdx.y;
  const bool valid = i < nx;
  const int s = (p * ny) / P, m = ((p + 1) * ny) / P - s;
  const long long col = (long long)(valid ? i : 0) + (long long)nx * ((long long)(ny + 2) * k);
  const double sh = shift ? shift[0] : 0.0;
  double f[MMAX], a[MMAX], c[MMAX];
#pragma unroll
  for (int r = 0; r < MMAX; ++r)
    if (r < m) f[r] = valid ? rhs[col + (long long)nx * (s + r + 1)] + sh : 0.0;
  const double *am = met + rowA * nyh, *ac = met + rowB * nyh, *ap = met + rowC * nyh;
  const double dym = (kind == 1 && valid) ? duy_ym[i + (long long)nx * k] : 0.0;
  auto getc = [&](int j, double &aj, double &bj, double &cj, double &fj) {
    if (kind == 1 && j == 1) {
      aj = 0.0;
      bj = 1.0;
      cj = 0.0;
      fj = dym;
    } else {
      aj = -beta * am[j];
      bj = -beta * ac[j] + 1.0;
      cj = -beta * ap[j];
    }
  };
  local_eliminate<MMAX>(f, a, c, m, s, getc);
  reduce_and_rebuild<MMAX>(f, a, c, m, P, p, lane, RA, RC, RF);
  if (valid) {
#pragma unroll
    for (int r = 0; r < MMAX; ++r)
      if (r < m) {
        const long long o = col + (long long)nx * (s + r + 1);
        u[o] = u[o] + f[r];
      }
  }
}

// ---- host launchers -------------------------------------------------------

// chunks per line for ny rows with at most mmax rows per chunk, each >= 4 rows
static int choose_chunks(int ny, int mmax) {
  int P = (ny + mmax - 1) / mmax;
  if (P < 1) P = 1;
  while (P > 1 && ny / P < 4) --P;
  return P;
}

int tri_plan(int ny, size_t lds_limit, size_t vbytes, int *P_out, int *mmax_out) {
  for (int mmax : {8, 16, 24, 32}) {
    const int P = choose_chunks(ny, mmax);
    if ((ny + P - 1) / P > mmax) continue;
    if (P > 16) continue;
    const size_t lds = (size_t)2 * P * 64 * (16 + vbytes);
    if (lds > lds_limit) continue;
    *P_out = P;
    *mmax_out = mmax;
    return PTC_OK;
  }
  set_error("ny = " + std::to_string(ny) + " is outside the partitioned tridiagonal kernels");
  return PTC_E_UNSUPPORTED;
}

int launch_tri_poisson(ptc_ctx *c, double2 *spec) {
  PoissonPlan &P = c->pp;
  if (P.tri == PTC_TRI_THOMAS) {
    dim3 blk(64, 4);
    dim3 grid((unsigned)((P.h + 1 + 63) / 64), (unsigned)((c->nz + 3) / 4));
    hipLaunchKernelGGL(k_tri_poisson_thomas, grid, blk, 0, c->stream, spec, P.d_tri_scratch, c->d_met, c->ny + 2,
                       c->d_wnx, c->d_wnz, c->ny, P.pitch, P.h, c->nz);
    PTC_HIP(hipGetLastError());
    return PTC_OK;
  }
  const int nP = P.tri_chunks;
  dim3 grid((unsigned)((P.h + 1 + 63) / 64), (unsigned)c->nz);
  const size_t lds = (size_t)2 * nP * 64 * 32;
  switch (P.tri_mmax) {
    case 8:
      hipLaunchKernelGGL(k_tri_poisson_part<8>, grid, dim3(64 * nP), lds, c->stream, spec, c->d_met, c->ny + 2,
                         c->d_wnx, c->d_wnz, c->ny, P.pitch, P.h, nP);
      break;
    case 16:
      hipLaunchKernelGGL(k_tri_poisson_part<16>, grid, dim3(64 * nP), lds, c->stream, spec, c->d_met, c->ny + 2,
                         c->d_wnx, c->d_wnz, c->ny, P.pitch, P.h, nP);
      break;
    case 24:
      hipLaunchKernelGGL(k_tri_poisson_part<24>, grid, dim3(64 * nP), lds, c->stream, spec, c->d_met, c->ny + 2,
                         c->d_wnx, c->d_wnz, c->ny, P.pitch, P.h, nP);
      break;
    default:
      hipLaunchKernelGGL(k_tri_poisson_part<32>, grid, dim3(64 * nP), lds, c->stream, spec, c->d_met, c->ny + 2,
                         c->d_wnx, c->d_wnz, c->ny, P.pitch, P.h, nP);
      break;
  }
  PTC_HIP(hipGetLastError());
  return PTC_OK;
}

// kind 0 (u/w), 1 (v), 2 (scalar); shift: optional device scalar added to rhs
int launch_tri_cn(ptc_ctx *c, int kind, double *u, const double *rhs, const double *duy_ym, double beta,
                  const double *shift) {
  int nP = 1, mmax = 16;
  PTC_TRY(tri_plan(c->ny, c->lds_limit, 8, &nP, &mmax));
  int rA, rB, rC;
  if (kind == 0) { rA = M_AM2CCN; rB = M_AC2CCN; rC = M_AP2CCN; }
  else if (kind == 1) { rA = M_AM2P; rB = M_AC2P; rC = M_AP2P; }
  else { rA = M_AM2CSCCN; rB = M_AC2CSCCN; rC = M_AP2CSCCN; }
  dim3 grid((unsigned)((c->nx + 63) / 64), (unsigned)c->nz);
  const size_t lds = (size_t)2 * nP * 64 * 24;
#define PTC_CN(MM)                                                                                          \
  hipLaunchKernelGGL(k_tri_cn_part<MM>, grid, dim3(64 * nP), lds, c->stream, u, rhs, duy_ym, shift, c->d_met, \
                     c->ny + 2, rA, rB, rC, beta, kind, c->nx, c->ny, nP)
  switch (mmax) {
    case 8: PTC_CN(8); break;
    case 16: PTC_CN(16); break;
    case 24: PTC_CN(24); break;
    default: PTC_CN(32); break;
  }
#undef PTC_CN
  PTC_HIP(hipGetLastError());
  return PTC_OK;
}

}  // namespace ptc
