// ptc_tridiag.hip -- batched wall-normal tridiagonal solves on gfx950.
//
// Two algorithms for the same systems:
//
//  * Partitioned (default).  A line of ny rows is split into P chunks, one
//    wavefront per chunk, one lane per real line (LPB consecutive lines per
//    workgroup, so every row load is a contiguous LPB-lane access).  Each lane
//    keeps its chunk in registers and runs the reference's "modified Thomas"
//    (src/m_Poisson.f90:497-531): afterwards row j reads
//        x_j + a_j x_first + c_j x_last = f_j ,
//    with the chunk's first row coupled to the previous chunk's last row.  The
//    2P first/last rows form a tridiagonal system with unit diagonal
//    (InverseTridiagonal_reduce, src/m_Poisson.f90:1045-1071) solved by
//    cyclic reduction in LDS; every lane then reconstructs its interior rows
//    (src/m_Poisson.f90:665-675).  One HBM read and one write per point.
//  * Thomas.  One thread per line: InverseTridiagonal (src/m_Tools.f90:224-251)
//    with the forward-sweep factor kept in a scratch array.  It moves ~2.5x the
//    bytes and exists as the reference algorithm and autotuner candidate.
//
// Poisson lines (kx, kz) carry complex right-hand sides with real rows whose
// diagonal is ac2ph(j) + WaveNumX(kx) + WaveNumZ(kz) (src/m_Poisson.f90:383-430);
// the partitioned kernel solves the real and imaginary parts as two real lines.
// The (0,0) line has row j=1 replaced by phi=0 (src/m_Poisson.f90:399-404).
// Crank-Nicolson lines (i, k) of clcU1Hat/U2Hat/U3Hat/clcScalar
// (src/m_TScheme.f90:365-506) are real with j-only coefficients.
#include <algorithm>

#include "ptc_internal.hpp"

namespace ptc {

__device__ __forceinline__ double vscale(double s, double v) { return s * v; }
__device__ __forceinline__ double2 vscale(double s, double2 v) { return make_double2(s * v.x, s * v.y); }
__device__ __forceinline__ double vsub(double a, double b) { return a - b; }
__device__ __forceinline__ double2 vsub(double2 a, double2 b) { return make_double2(a.x - b.x, a.y - b.y); }
__device__ __forceinline__ double vzero(double) { return 0.0; }
__device__ __forceinline__ double2 vzero(double2) { return make_double2(0.0, 0.0); }

// Row coefficients of one line.  A row replaced by an identity row (the
// pinned Poisson row, the v wall row) gets its right-hand side from the kernel.
struct PoissonRowsRe {  // ac2ph(j) + WaveNumX + WaveNumZ, pinned (0,0) line
  const double *am, *ac, *ap;
  double lam;
  int ny;
  bool pin;
  __device__ __forceinline__ void operator()(int j, double &aj, double &bj, double &cj) const {
    if (j > ny) {
      aj = 0.0;
      bj = 1.0;
      cj = 0.0;
    } else {
      aj = am[j];
      cj = j < ny ? ap[j] : 0.0;
      bj = ac[j] + lam;
      if (pin && j == 1) {
        aj = 0.0;
        bj = 1.0;
        cj = 0.0;
      }
    }
  }
};

struct CnRows {  // (I - beta L_y) with the wall rows of the chosen table
  const double *am, *ac, *ap;
  double beta;
  int ny, kind;
  __device__ __forceinline__ void operator()(int j, double &aj, double &bj, double &cj) const {
    if (j > ny) {
      aj = 0.0;
      bj = 1.0;
      cj = 0.0;
    } else if (kind == 1 && j == 1) {
      aj = 0.0;
      bj = 1.0;
      cj = 0.0;
    } else {
      aj = -beta * am[j];
      bj = -beta * ac[j] + 1.0;
      cj = j < ny ? -beta * ap[j] : 0.0;
    }
  }
};

// Every chunk holds exactly MM rows (reference rows j = s+1 .. s+MM).  Rows
// beyond ny are identity rows (a = c = 0, b = 1, f = 0) and the last real row
// drops its super-diagonal, which InverseTridiagonal never reads either
// (src/m_Tools.f90:239-245), so the padded system has the same solution on the
// real rows and all control flow here is compile-time.
//
// Chunk rows r = 0..MM-1 read a_r x_{r-1} + b_r x_r + c_r x_{r+1} = f_r.  The
// downward elimination of rows 1..MM-1 keeps the chunk's first value x_0 as a
// parameter, as the reference's modified Thomas does (src/m_Poisson.f90:497-531):
//     x_r + C_r x_{r+1} = F_r + A_r x_0 ,
// one division per row; F_r overwrites f_r.  Its last row is the chunk's
// reduced last row
//     aE x_0 + x_last + cE x_{next first} = fE ,
// and an upward pass with O(1) state writes x_1 in terms of x_0 and x_last,
// which turns row 0 into the reduced first row
//     aS x_{prev last} + x_0 + cS x_last = fS .
template <int MM, typename V, typename G>
__device__ __forceinline__ void chunk_reduce(V (&f)[MM], double (&A)[MM], double (&C)[MM], int s, G &getc,
                                             double &aS, double &cS, V &fS, double &aE, double &cE, V &fE) {
  static_assert(MM >= 4, "chunks need at least four rows");
  double aj, bj, cj;
  getc(s + 2, aj, bj, cj);
  double rr = 1.0 / bj;
  C[1] = cj * rr;
  A[1] = -aj * rr;
  f[1] = vscale(rr, f[1]);
#pragma unroll
  for (int r = 2; r < MM; ++r) {
    getc(s + r + 1, aj, bj, cj);
    rr = 1.0 / (bj - aj * C[r - 1]);
    C[r] = cj * rr;
    A[r] = -rr * aj * A[r - 1];
    f[r] = vscale(rr, vsub(f[r], vscale(aj, f[r - 1])));
  }
  aE = -A[MM - 1];
  cE = C[MM - 1];
  fE = f[MM - 1];
  // x_r = g + al x_0 + ga x_last, from r = MM-2 down to r = 1
  V g = f[MM - 2];
  double al = A[MM - 2], ga = -C[MM - 2];
#pragma unroll
  for (int r = MM - 3; r >= 1; --r) {
    g = vsub(f[r], vscale(C[r], g));
    al = A[r] - C[r] * al;
    ga = -C[r] * ga;
  }
  getc(s + 1, aj, bj, cj);
  const double r0 = 1.0 / (bj + cj * al);
  aS = r0 * aj;
  cS = r0 * cj * ga;
  fS = vscale(r0, vsub(f[0], vscale(cj, g)));
}

// With x_0 = xs and x_last = xe known: x_r = F_r + A_r x_0 - C_r x_{r+1}
// (no divisions); f receives the chunk.
template <int MM, typename V>
__device__ __forceinline__ void chunk_backsub(V (&f)[MM], const double (&A)[MM], const double (&C)[MM], V xs,
                                              V xe) {
  f[MM - 1] = xe;
#pragma unroll
  for (int r = MM - 2; r >= 1; --r) f[r] = vsub(vsub(f[r], vscale(-A[r], xs)), vscale(C[r], f[r + 1]));
  f[0] = xs;
}

// Reduced system of 2P rows per line (row 2p: first row of chunk p, row 2p+1:
// its last row), unit diagonal, solved by parallel cyclic reduction: at step d
// every row eliminates its neighbours q-d and q+d, so after ceil(log2(2P))
// steps each row stands alone.  Rows outside the system act as decoupled
// identity rows.  The rows live in LDS, double-buffered ([2][2P][LPB] for
// each of a, c, f): a step reads one buffer and writes the other, so it needs
// one barrier and holds nothing in registers across steps.  The reference
// solves the same reduced system by Thomas (InverseTridiagonal_reduce,
// src/m_Poisson.f90:1045-1071); cyclic reduction lets every chunk's lane take
// part instead of one wavefront.
template <int LPB, typename V>
__device__ __forceinline__ void pcr_row(int q, int d, int n2, int lane, const double *RA, const double *RC,
                                        const V *RF, double *WA, double *WC, V *WF) {
  const int o = q * LPB + lane;
  const double a = RA[o], c = RC[o];
  V f = RF[o];
  double am = 0.0, cm = 0.0, ap = 0.0, cp = 0.0;
  V fm = vzero(f), fp = vzero(f);
  if (q - d >= 0) {
    am = RA[o - d * LPB];
    cm = RC[o - d * LPB];
    fm = RF[o - d * LPB];
  }
  if (q + d < n2) {
    ap = RA[o + d * LPB];
    cp = RC[o + d * LPB];
    fp = RF[o + d * LPB];
  }
  // a x_{q-d} + x_q + c x_{q+d} = f, with rows q-d and q+d substituted
  const double inv = 1.0 / (1.0 - a * cm - c * ap);
  WF[o] = vscale(inv, vsub(vsub(f, vscale(a, fm)), vscale(c, fp)));
  WA[o] = -inv * a * am;
  WC[o] = -inv * c * cp;
}

template <int LPB, typename V>
__device__ __forceinline__ void solve_reduced(double aS, double cS, V fS, double aE, double cE, V fE, int P, int p,
                                              int lane, double *RA, double *RC, V *RF, V &xs, V &xe) {
  const int n2 = 2 * P, q0 = 2 * p, q1 = 2 * p + 1;
  const int half = n2 * LPB;  // second buffer
  RA[q0 * LPB + lane] = aS;
  RC[q0 * LPB + lane] = cS;
  RF[q0 * LPB + lane] = fS;
  RA[q1 * LPB + lane] = aE;
  RC[q1 * LPB + lane] = cE;
  RF[q1 * LPB + lane] = fE;
  int in = 0;
  for (int d = 1; d < n2; d *= 2) {
    __syncthreads();
    const int out = half - in;
    pcr_row<LPB>(q0, d, n2, lane, RA + in, RC + in, RF + in, RA + out, RC + out, RF + out);
    pcr_row<LPB>(q1, d, n2, lane, RA + in, RC + in, RF + in, RA + out, RC + out, RF + out);
    in = out;
  }
  // each thread reads back only the rows it wrote itself
  xs = RF[in + q0 * LPB + lane];
  xe = RF[in + q1 * LPB + lane];
}

// chunk index of a thread; with 64 lines per block a chunk is one wavefront
template <int LPB>
__device__ __forceinline__ int chunk_of_thread() {
  if constexpr (LPB == 64) return __builtin_amdgcn_readfirstlane(threadIdx.x >> 6);
  return threadIdx.x / LPB;
}

// ---- Poisson y-solve, partitioned ----------------------------------------
// Largest block of each kernel family: the real-valued Poisson kernels stay
// within 128 VGPRs up to MM = 16; Crank-Nicolson chunks of 16 rows need more,
// so their blocks hold at most 512 threads.
template <int MM>
constexpr int tri_cn_block_cap() {
  return MM >= 16 ? 512 : 1024;
}

// The rows are real, so the real and imaginary parts of a line are two
// independent real systems.  Lane 2m handles the real part of line kx0+m,
// lane 2m+1 its imaginary part: a row is still one contiguous 8-byte-per-lane
// access, the right-hand side takes half the registers, and with LPB = 64 a
// chunk is one wavefront so the row coefficients are wave-uniform.
// S: the spectrum's storage (double2, or float2 in the FP32-storage form);
// the elimination itself is FP64 in both.
template <int MM, int LPB, typename S>
__global__ void __launch_bounds__(1024) k_tri_poisson_split(S *__restrict__ spec, const double *__restrict__ met,
                                                           int nyh, const double *__restrict__ wnx,
                                                           const double *__restrict__ wnz, int ny, int pitch, int h,
                                                           int P) {
  extern __shared__ double smem[];
  double *RA = smem;
  double *RC = RA + 4 * P * LPB;
  double *RF = RC + 4 * P * LPB;
  const int lane = threadIdx.x % LPB;
  const int p = chunk_of_thread<LPB>();
  const int kx = blockIdx.x * (LPB / 2) + lane / 2, kz = blockIdx.y;
  const bool valid = kx <= h;
  const int s = p * MM;
  const bool pin = (kx == 0 && kz == 0);
  // row base addresses are wave-uniform (scalar), the lane adds a 32-bit offset
  using R = decltype(spec->x);
  R *base = reinterpret_cast<R *>(spec + (long long)kz * ny * pitch) + (long long)blockIdx.x * LPB;
  const long long rs = 2LL * pitch;
  double f[MM];
  double lam;
  if constexpr (sizeof(R) == 8) {
    // FP64 storage: predicated row loads (measured fastest for this width)
    lam = valid ? wnx[kx] + wnz[kz] : 1.0;
#pragma unroll
    for (int r = 0; r < MM; ++r) f[r] = (valid && s + r < ny) ? base[(s + r) * rs + lane] : 0.0;
  } else {
    // FP32 storage: a predicated load feeding a conversion is compiled as
    // load-wait-convert per row, so the loads are issued unconditionally
    // (lanes past the last line read column 0, padding rows read row ny-1:
    // finite values, masked by a multiply), the wavenumbers first: the
    // elimination starts on row 0 while the later rows are still arriving.
    const double wx_k = wnx[valid ? kx : 0], wz_k = wnz[kz];
    const int lc = valid ? lane : -(int)(blockIdx.x * LPB);
#pragma unroll
    for (int r = 0; r < MM; ++r) {
      const double v = (double)base[min(s + r, ny - 1) * rs + lc];
      f[r] = v * ((valid && s + r < ny) ? 1.0 : 0.0);
    }
    lam = valid ? wx_k + wz_k : 1.0;
  }
  if (pin && s == 0) f[0] = 0.0;  // phi(1) = 0 on the (0,0) line
  PoissonRowsRe getc{met + M_AM2PH * nyh, met + M_AC2PH * nyh, met + M_AP2PH * nyh, lam, ny, pin};
  double A[MM], C[MM];
  double aS, cS, aE, cE, fS, fE, xs, xe;
  chunk_reduce<MM>(f, A, C, s, getc, aS, cS, fS, aE, cE, fE);
  solve_reduced<LPB>(aS, cS, fS, aE, cE, fE, P, p, lane, RA, RC, RF, xs, xe);
  chunk_backsub<MM>(f, A, C, xs, xe);
  if (valid) {
#pragma unroll
    for (int r = 0; r < MM; ++r)
      if (s + r < ny) base[(s + r) * rs + lane] = (R)f[r];
  }
}

// ---- Poisson y-solve, partitioned, staged through LDS ----------------------
// The same elimination for LL real lines per workgroup: the lines' rows are
// loaded once into an LDS tile with coalesced LL-wide row reads, each thread
// (chunk p, line l) takes its chunk from the tile, and the results go back
// out the same way.  With LL = 16 or 32 a workgroup is small (P * LL
// threads) and several fit on a CU, so one workgroup's loads overlap another's
// elimination.  The row coefficients are copied to LDS as well; the cyclic
// reduction's buffers reuse the tile.
template <int MM, int LL, typename S>
__global__ void __launch_bounds__(512) k_tri_poisson_lds(S *__restrict__ spec, const double *__restrict__ met, int nyh,
                                                         const double *__restrict__ wnx,
                                                         const double *__restrict__ wnz, int ny, int pitch, int h,
                                                         int P) {
  extern __shared__ double smem[];
  constexpr int LP = LL + 1;  // padded tile row
  double *mA = smem, *mC = mA + nyh, *mP = mC + nyh;
  double *tile = mP + nyh;
  const int t = threadIdx.x, nthr = blockDim.x;
  for (int j = t; j < nyh; j += nthr) {
    mA[j] = met[M_AM2PH * nyh + j];
    mC[j] = met[M_AC2PH * nyh + j];
    mP[j] = met[M_AP2PH * nyh + j];
  }
  using R = decltype(spec->x);
  const long long rs = 2LL * pitch;
  const int col0 = blockIdx.x * LL, ncol = 2 * (h + 1), kz = blockIdx.y;
  R *base = reinterpret_cast<R *>(spec + (long long)kz * ny * pitch) + col0;
  for (int e = t; e < ny * LL; e += nthr) {
    const int r = e / LL, q = e - r * LL;
    tile[r * LP + q] = (col0 + q < ncol) ? (double)base[r * rs + q] : 0.0;
  }
  __syncthreads();
  const int line = t % LL, p = t / LL;
  const int s = p * MM;
  double f[MM];
#pragma unroll
  for (int r = 0; r < MM; ++r) f[r] = (s + r < ny) ? tile[(s + r) * LP + line] : 0.0;
  const int kx = (col0 + line) >> 1;
  const bool valid = col0 + line < ncol;
  const double lam = valid ? wnx[kx] + wnz[kz] : 1.0;
  const bool pin = (kx == 0 && kz == 0);
  if (pin && s == 0) f[0] = 0.0;  // phi(1) = 0 on the (0,0) line
  PoissonRowsRe getc{mA, mC, mP, lam, ny, pin};
  double A[MM], C[MM];
  double aS, cS, aE, cE, fS, fE, xs, xe;
  chunk_reduce<MM>(f, A, C, s, getc, aS, cS, fS, aE, cE, fE);
  __syncthreads();  // every chunk has left the tile: it becomes the reduction's buffers
  double *RA = tile, *RC = RA + 4 * P * LL, *RF = RC + 4 * P * LL;
  solve_reduced<LL>(aS, cS, fS, aE, cE, fE, P, p, line, RA, RC, RF, xs, xe);
  chunk_backsub<MM>(f, A, C, xs, xe);
  __syncthreads();  // the reduction's buffers are read: back to the tile
#pragma unroll
  for (int r = 0; r < MM; ++r)
    if (s + r < ny) tile[(s + r) * LP + line] = f[r];
  __syncthreads();
  for (int e = t; e < ny * LL; e += nthr) {
    const int r = e / LL, q = e - r * LL;
    if (col0 + q < ncol) base[r * rs + q] = (R)tile[r * LP + q];
  }
}

// ---- Poisson y-solve, partitioned, one complex line per lane ----------------
// The same elimination with the real and imaginary parts of a line in one
// lane: the row coefficients (and so the divisions of the elimination and of
// the cyclic reduction) are shared by the two parts instead of being formed
// twice.  LC complex lines per workgroup, P chunks: P * LC threads, chunk
// p = thread / LC.  A wavefront holds 64 / LC chunks, so the row
// coefficients are per lane; they are staged in LDS once per workgroup.  A
// row is a contiguous LC * 16-byte access (LC * 8 bytes with FP32 storage).
// spectrum storage <-> FP64 arithmetic; STREAM: nontemporal access (a pass
// over a spectrum larger than the last-level cache)
typedef double tr_d2 __attribute__((ext_vector_type(2)));
typedef float tr_f2 __attribute__((ext_vector_type(2)));
template <bool STREAM>
__device__ __forceinline__ double2 cld(const double2 &v) {
  if constexpr (STREAM) {
    const tr_d2 w = __builtin_nontemporal_load(reinterpret_cast<const tr_d2 *>(&v));
    return make_double2(w.x, w.y);
  } else {
    return v;
  }
}
template <bool STREAM>
__device__ __forceinline__ double2 cld(const float2 &v) {
  if constexpr (STREAM) {
    const tr_f2 w = __builtin_nontemporal_load(reinterpret_cast<const tr_f2 *>(&v));
    return make_double2(w.x, w.y);
  } else {
    return make_double2(v.x, v.y);
  }
}
template <bool STREAM>
__device__ __forceinline__ void cst(double2 &d, double2 v) {
  if constexpr (STREAM) {
    tr_d2 w = {v.x, v.y};
    __builtin_nontemporal_store(w, reinterpret_cast<tr_d2 *>(&d));
  } else {
    d = v;
  }
}
template <bool STREAM>
__device__ __forceinline__ void cst(float2 &d, double2 v) {
  if constexpr (STREAM) {
    tr_f2 w = {(float)v.x, (float)v.y};
    __builtin_nontemporal_store(w, reinterpret_cast<tr_f2 *>(&d));
  } else {
    d = make_float2((float)v.x, (float)v.y);
  }
}

// LDS of the complex-line kernel in doubles: the three coefficient rows
// (padded to 16 bytes) and a, c, f (f complex) of 2P reduced rows x LC lines,
// double-buffered
constexpr size_t tri_cplx_lds_doubles(int nyh, int P, int LC) {
  return (size_t)((3 * nyh + 1) / 2 * 2) + (size_t)16 * P * LC;
}

// chunks of 12 or more complex rows need more than 128 VGPRs: their
// workgroups hold at most 512 threads
constexpr int tri_cplx_cap(int mm) { return mm >= 12 ? 512 : 1024; }

template <int MM, int LC, typename S, bool STREAM>
__global__ void __launch_bounds__(tri_cplx_cap(MM)) k_tri_poisson_cplx(S *__restrict__ spec, const double *__restrict__ met,
                                                          int nyh, const double *__restrict__ wnx,
                                                          const double *__restrict__ wnz, int ny, int pitch, int h,
                                                          int P) {
  extern __shared__ double smem[];
  double *mA = smem, *mC = mA + nyh, *mP = mC + nyh;
  double *RA = smem + (3 * nyh + 1) / 2 * 2;  // 16-byte aligned
  double *RC = RA + 4 * P * LC;
  double2 *RF = reinterpret_cast<double2 *>(RC + 4 * P * LC);
  const int t = threadIdx.x;
  const int lane = t % LC, p = t / LC;
  const int kx = blockIdx.x * LC + lane, kz = blockIdx.y;
  const bool valid = kx <= h;
  const int s = p * MM;
  const bool pin = (kx == 0 && kz == 0);
  // wavenumbers first, then every row of the chunk (lanes past the last line
  // read line 0, padding rows read row ny-1; masked by a multiply), all in
  // flight while the coefficients go to LDS
  const double wx_k = wnx[valid ? kx : 0], wz_k = wnz[kz];
  S *col = spec + (long long)kz * ny * pitch + (valid ? kx : 0);
  double2 f[MM];
#pragma unroll
  for (int r = 0; r < MM; ++r) {
    const double2 v = cld<STREAM>(col[(long long)min(s + r, ny - 1) * pitch]);
    const double m = (valid && s + r < ny) ? 1.0 : 0.0;
    f[r] = make_double2(v.x * m, v.y * m);
  }
  for (int j = t; j < nyh; j += blockDim.x) {
    mA[j] = met[M_AM2PH * nyh + j];
    mC[j] = met[M_AC2PH * nyh + j];
    mP[j] = met[M_AP2PH * nyh + j];
  }
  __syncthreads();
  const double lam = valid ? wx_k + wz_k : 1.0;
  if (pin && s == 0) f[0] = make_double2(0.0, 0.0);  // phi(1) = 0 on the (0,0) line
  PoissonRowsRe getc{mA, mC, mP, lam, ny, pin};
  double A[MM], C[MM];
  double aS, cS, aE, cE;
  double2 fS, fE, xs, xe;
  chunk_reduce<MM>(f, A, C, s, getc, aS, cS, fS, aE, cE, fE);
  solve_reduced<LC>(aS, cS, fS, aE, cE, fE, P, p, lane, RA, RC, RF, xs, xe);
  chunk_backsub<MM>(f, A, C, xs, xe);
  if (valid) {
#pragma unroll
    for (int r = 0; r < MM; ++r)
      if (s + r < ny) cst<STREAM>(col[(long long)(s + r) * pitch], f[r]);
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
//
// The matrix I - beta L_y of a job is the same for every line (i, k).  Its
// chunk elimination factors, the reduced rows and the coefficients of every
// cyclic-reduction step are formed by k_cn_factor (one workgroup per job,
// the same arithmetic chunk_reduce and pcr_row apply to the coefficients;
// once per launch, or once per beta for the tables a step keeps), and the
// solve reads them as wave-uniform scalars: it
// keeps only the right-hand side and the old u in registers and does no
// division.  Table of one job (doubles):
//   [0, PM)        rr_q  reciprocal pivot of chunk row q = s + r (r >= 1)
//   [PM, 2PM)      a_q   sub-diagonal of row q (r >= 2)
//   [2PM, 3PM)     C_q   upper factor (r >= 1)
//   [3PM, 4PM)     A_q   first-row factor (r >= 1)
//   [4PM, 4PM+P)   r0_p  reciprocal pivot of chunk p's reduced first row
//   [4PM+P, +2P)   c_p   super-diagonal of chunk p's first row
//   then (a, c, inv, 0) of reduced row q at cyclic-reduction step l at
//   4PM + 2P + 4 (l 2P + q)
// with PM = P * MM.
struct CnFacIdx {
  int pm, P;
  __device__ __forceinline__ int rr(int q) const { return q; }
  __device__ __forceinline__ int a(int q) const { return pm + q; }
  __device__ __forceinline__ int C(int q) const { return 2 * pm + q; }
  __device__ __forceinline__ int A(int q) const { return 3 * pm + q; }
  __device__ __forceinline__ int r0(int p) const { return 4 * pm + p; }
  __device__ __forceinline__ int c0(int p) const { return 4 * pm + P + p; }
  __device__ __forceinline__ int pcr(int l, int q) const { return 4 * pm + 2 * P + 4 * (l * 2 * P + q); }
};

template <int MM>
__global__ void __launch_bounds__(64) k_cn_factor(CnJobs jobs, const double *__restrict__ met, int nyh, int ny,
                                                  int P, double *__restrict__ fac) {
  __shared__ double sa[2][64], sc[2][64];
  const CnJob &J = jobs.j[blockIdx.x];
  if (!J.need) return;  // formed by an earlier launch, or by another job of this one
  double *T = fac + (long long)J.tab * kCnFacStride;
  const CnFacIdx X{P * MM, P};
  const int t = threadIdx.x, n2 = 2 * P;
  if (t < P) {
    // chunk_reduce on the coefficients alone
    CnRows getc{met + J.rA * nyh, met + J.rB * nyh, met + J.rC * nyh, J.beta, ny, J.kind};
    const int s = t * MM;
    double A[MM], C[MM];
    double aj, bj, cj;
    getc(s + 2, aj, bj, cj);
    double rr = 1.0 / bj;
    C[1] = cj * rr;
    A[1] = -aj * rr;
    T[X.rr(s + 1)] = rr;
    T[X.a(s + 1)] = aj;
#pragma unroll
    for (int r = 2; r < MM; ++r) {
      getc(s + r + 1, aj, bj, cj);
      rr = 1.0 / (bj - aj * C[r - 1]);
      C[r] = cj * rr;
      A[r] = -rr * aj * A[r - 1];
      T[X.rr(s + r)] = rr;
      T[X.a(s + r)] = aj;
    }
    const double aE = -A[MM - 1], cE = C[MM - 1];
    double al = A[MM - 2], ga = -C[MM - 2];
#pragma unroll
    for (int r = MM - 3; r >= 1; --r) {
      al = A[r] - C[r] * al;
      ga = -C[r] * ga;
    }
    getc(s + 1, aj, bj, cj);
    const double r0 = 1.0 / (bj + cj * al);
    T[X.r0(t)] = r0;
    T[X.c0(t)] = cj;
#pragma unroll
    for (int r = 1; r < MM; ++r) {
      T[X.C(s + r)] = C[r];
      T[X.A(s + r)] = A[r];
    }
    sa[0][2 * t] = r0 * aj;
    sc[0][2 * t] = r0 * cj * ga;
    sa[0][2 * t + 1] = aE;
    sc[0][2 * t + 1] = cE;
  }
  // pcr_row on the coefficients: each step's (a, c, inv) of every reduced row
  int in = 0, l = 0;
  for (int d = 1; d < n2; d *= 2, ++l) {
    __syncthreads();
    if (t < n2) {
      const double a = sa[in][t], c = sc[in][t];
      double am = 0.0, cm = 0.0, ap = 0.0, cp = 0.0;
      if (t - d >= 0) {
        am = sa[in][t - d];
        cm = sc[in][t - d];
      }
      if (t + d < n2) {
        ap = sa[in][t + d];
        cp = sc[in][t + d];
      }
      const double inv = 1.0 / (1.0 - a * cm - c * ap);
      double *e = T + X.pcr(l, t);
      e[0] = a;
      e[1] = c;
      e[2] = inv;
      e[3] = 0.0;
      sa[1 - in][t] = -inv * a * am;
      sc[1 - in][t] = -inv * c * cp;
    }
    in = 1 - in;
  }
}

// The constant-flow-rate shift of RhsX (src/m_TScheme.f90:269-300), in the
// arithmetic of k_flowrate: every workgroup of the u job sums the RHS
// kernel's partials in the same fixed order (strided per thread, then a tree
// in red), so all hold the same shift; workgroup (0, 0) stores it and the
// PrGradAver increment, as k_flowrate would.  Called by all threads.
__device__ double flow_fold_shift(const FlowFold &ff, double *red) {
  const int t = threadIdx.x, nt = blockDim.x;
  double acc = 0.0;
  for (int q = t; q < ff.n; q += nt) acc += ff.part[q];
  red[t] = acc;
  __syncthreads();
  for (int n = nt; n > 1;) {
    const int h = (n + 1) / 2;
    if (t < n - h) red[t] += red[t + h];
    __syncthreads();
    n = h;
  }
  const double tot = red[0];
  __syncthreads();  // red is the solve's LDS
  double shift, inc;
  if (ff.scalar_build) {
    const double s3 = tot / ff.nxnz / ff.yly;
    const double dpm = ff.ubulk - ff.ucrf - s3;
    shift = ff.alphaC * dpm;
    inc = dpm * ff.alphaC / ff.dt;
  } else {
    const double dpm = ff.xnu * tot / ff.nxnz / ff.yly;
    shift = -(ff.alpha * dpm);
    inc = dpm * ff.alphaC;
  }
  if (t == 0 && blockIdx.x == 0 && blockIdx.y == 0) {
    *ff.dp1ns = shift;
    *ff.prgrad = (ff.first ? 0.0 : *ff.prgrad) + inc;
  }
  return shift;
}

template <int MM, int LPB>
__global__ void __launch_bounds__(tri_cn_block_cap<MM>()) k_tri_cn_fac(CnJobs jobs, const double *__restrict__ fac,
                                                                     int nx, int ny, int P, FlowFold ff) {
  extern __shared__ double smem[];  // f of the 2P reduced rows x LPB lines, double-buffered
  double *RF = smem;
  const CnJob &J = jobs.j[blockIdx.z];
  const double *__restrict__ T = fac + (long long)J.tab * kCnFacStride;
  const CnFacIdx X{P * MM, P};
  double *__restrict__ u = J.u;
  const double *__restrict__ rhs = J.rhs;
  const int kind = J.kind;
  const int lane = threadIdx.x % LPB;
  const int p = chunk_of_thread<LPB>();
  const int i = blockIdx.x * LPB + lane, k = blockIdx.y;
  const bool valid = i < nx;
  const int s = p * MM;
  const long long col = (long long)(valid ? i : 0) + (long long)nx * ((long long)(ny + 2) * k);
  // the right-hand side and the old u, all rows in flight at once
  // (unconditional loads: padding rows read row ny, masked by a multiply)
  double f[MM], uo[MM];
#pragma unroll
  for (int r = 0; r < MM; ++r) {
    const long long o = col + (long long)nx * (min(s + r, ny - 1) + 1);
    f[r] = rhs[o];
    uo[r] = u[o];
  }
  double sh = J.shift ? J.shift[0] : 0.0;
  if (J.fold) sh = flow_fold_shift(ff, RF);  // uniform over the workgroup (one job)
#pragma unroll
  for (int r = 0; r < MM; ++r) f[r] = (f[r] + sh) * ((valid && s + r < ny) ? 1.0 : 0.0);
  // du at the wall row: duy_ym, or with Update_uy_ym folded in (src/m_FlowCase.f90:112-122)
  // the new uy_ym (zero) minus the old one, both planes written after the solve
  const double wall_new = 0.0;
  double wall_old = 0.0;
  const bool wall_thread = kind == 1 && s == 0 && valid;
  if (wall_thread) {
    const long long q = i + (long long)nx * k;
    if (J.uym_upd) {
      wall_old = J.uym_upd[q];
      f[0] = wall_new - wall_old;
    } else {
      f[0] = J.duy_ym[q];
    }
  }
  if (kind == 1 && s == 0 && !valid) f[0] = 0.0;
  // chunk elimination (chunk_reduce with the factors read from the table)
  f[1] = T[X.rr(s + 1)] * f[1];
#pragma unroll
  for (int r = 2; r < MM; ++r) f[r] = T[X.rr(s + r)] * (f[r] - T[X.a(s + r)] * f[r - 1]);
  const double fE = f[MM - 1];
  double g = f[MM - 2];
#pragma unroll
  for (int r = MM - 3; r >= 1; --r) g = f[r] - T[X.C(s + r)] * g;
  const double fS = T[X.r0(p)] * (f[0] - T[X.c0(p)] * g);
  // reduced rows by cyclic reduction (pcr_row with the step coefficients read)
  const int n2 = 2 * P, q0 = 2 * p, q1 = 2 * p + 1, half = n2 * LPB;
  RF[q0 * LPB + lane] = fS;
  RF[q1 * LPB + lane] = fE;
  int in = 0, l = 0;
  for (int d = 1; d < n2; d *= 2, ++l) {
    __syncthreads();
    const int out = half - in;
#pragma unroll
    for (int h = 0; h < 2; ++h) {
      const int q = h ? q1 : q0;
      const int o = q * LPB + lane;
      const double *e = T + X.pcr(l, q);
      const double fc = RF[in + o];
      const double fm = q - d >= 0 ? RF[in + o - d * LPB] : 0.0;
      const double fp = q + d < n2 ? RF[in + o + d * LPB] : 0.0;
      RF[out + o] = e[2] * ((fc - e[0] * fm) - e[1] * fp);
    }
    in = out;
  }
  const double xs = RF[in + q0 * LPB + lane], xe = RF[in + q1 * LPB + lane];
  // back substitution (chunk_backsub)
  f[MM - 1] = xe;
#pragma unroll
  for (int r = MM - 2; r >= 1; --r) f[r] = (f[r] - (-T[X.A(s + r)]) * xs) - T[X.C(s + r)] * f[r + 1];
  f[0] = xs;
  if (wall_thread && J.uym_upd) {
    const long long q = i + (long long)nx * k;
    J.uym_upd[q] = wall_new;
    J.duym_upd[q] = wall_new - wall_old;
  }
  if (valid) {
    const bool wall = kind == 1 && (J.wall != nullptr || J.uym_upd != nullptr);
#pragma unroll
    for (int r = 0; r < MM; ++r)
      if (s + r < ny) {
        const long long o = col + (long long)nx * (s + r + 1);
        const double v = (wall && s + r == 0) ? (J.uym_upd ? wall_new : J.wall[i + (long long)nx * k]) : uo[r] + f[r];
        u[o] = v;
        // the scalar's wall ghost rows from rows 1 and ny
        if (J.sbc && s + r == 0) u[col] = J.r1m * v + J.r2m;
        if (J.sbc && s + r == ny - 1) u[col + (long long)nx * (ny + 1)] = J.r1p * v + J.r2p;
      }
    if (wall && s <= ny - 1 && ny - 1 < s + MM) u[col + (long long)nx * (ny + 1)] = J.top;
  }
}

// ---- host launchers -------------------------------------------------------

// Launch shape of the partitioned solver: chunk length MM (compile-time), lines
// per block LPB and chunk count P = ceil(ny / MM).  The first candidate with
// at most 1/8 padded rows whose block (LPB * P threads) and LDS fit wins.
// The factored Crank-Nicolson solve of 12-row chunks and 64 lanes holds 79
// VGPRs, so two of its 768-thread workgroups share a CU: one loads while the
// other eliminates (4608x140x700: 8.65 ms for the four solves against 12.8 ms
// for the form that factored every line itself, one workgroup per CU).
struct TriShape {
  int mm, lpb, cap;  // cap: the instantiation's launch bound
};
static const TriShape kPoissonShapes[] = {{4, 64, 1024}, {8, 64, 1024}, {12, 64, 1024}, {16, 64, 1024},
                                          {8, 32, 1024},  {12, 32, 1024}, {16, 32, 512},  {16, 16, 256}};
static const TriShape kCnShapes[] = {{4, 64, 1024}, {8, 64, 1024}, {12, 64, 1024}, {8, 32, 1024}, {16, 32, 512}};

static int tri_choose(int ny, size_t lds_limit, int mm_req, int lpb_req, const TriShape *shapes, int nshapes,
                      int *P_out, int *mm_out, int *lpb_out) {
  for (int pass = 0; pass < 2; ++pass)
    for (int q = 0; q < nshapes; ++q) {
      const int mm = shapes[q].mm, lpb = shapes[q].lpb;
      if (mm_req > 0 && mm != mm_req) continue;
      if (lpb_req > 0 && lpb != lpb_req) continue;
      const int P = (ny + mm - 1) / mm;
      if (lpb * P > shapes[q].cap) continue;
      if (2 * P * lpb > 1 << 16) continue;
      if (pass == 0 && (P * mm - ny) * 8 > ny) continue;  // prefer little padding
      const size_t lds = (size_t)4 * P * lpb * 24;  // a, c, f of the reduced rows, two buffers
      if (lds > lds_limit) continue;
      *P_out = P;
      *mm_out = mm;
      *lpb_out = lpb;
      return PTC_OK;
    }
  set_error("ny = " + std::to_string(ny) + (mm_req > 0 ? " with " + std::to_string(mm_req) + "-row chunks" : "") +
            (lpb_req > 0 ? " and " + std::to_string(lpb_req) + " lines per workgroup" : "") +
            " is outside the partitioned tridiagonal kernels");
  return PTC_E_UNSUPPORTED;
}

// LDS-staged Poisson kernel: LL lines per workgroup, P * LL <= 512 threads,
// LDS = row coefficients + max(tile, reduction buffers)
size_t tri_lds_bytes(int ny, int P, int ll) {
  const size_t tile = (size_t)ny * (ll + 1), red = (size_t)12 * P * ll;
  return sizeof(double) * (3 * (size_t)(ny + 2) + std::max(tile, red));
}

int tri_plan_lds(int ny, size_t lds_limit, int ll, int mm_req, int *P_out, int *mm_out) {
  for (int pass = 0; pass < 2; ++pass)
    for (int mm : {4, 8, 12, 16}) {
      if (mm_req > 0 && mm != mm_req) continue;
      const int P = (ny + mm - 1) / mm;
      if (P * ll > 512) continue;
      if (pass == 0 && (P * mm - ny) * 8 > ny) continue;
      if (tri_lds_bytes(ny, P, ll) > lds_limit) continue;
      *P_out = P;
      *mm_out = mm;
      return PTC_OK;
    }
  set_error("ny = " + std::to_string(ny) + " does not fit the LDS-staged tridiagonal kernel with " +
            std::to_string(ll) + " lines per workgroup");
  return PTC_E_UNSUPPORTED;
}

// complex-line kernel: P * CL threads within the instantiation's launch
// bound, LDS = coefficients + the double-buffered reduction rows
int tri_plan_cplx(int ny, size_t lds_limit, int cl, int mm_req, int *P_out, int *mm_out) {
  // chunk lengths in order of measured speed at ny = 256 (8 rows best; 4-row
  // chunks make the cyclic reduction the larger part)
  for (int pass = 0; pass < 2; ++pass)
    for (int mm : {8, 16, 12, 4}) {
      if (mm_req > 0 && mm != mm_req) continue;
      const int P = (ny + mm - 1) / mm;
      if (P * cl > tri_cplx_cap(mm)) continue;
      if (pass == 0 && (P * mm - ny) * 8 > ny) continue;
      if (sizeof(double) * tri_cplx_lds_doubles(ny + 2, P, cl) > lds_limit) continue;
      *P_out = P;
      *mm_out = mm;
      return PTC_OK;
    }
  set_error("ny = " + std::to_string(ny) + " does not fit the complex-line tridiagonal kernel with " +
            std::to_string(cl) + " lines per workgroup");
  return PTC_E_UNSUPPORTED;
}

int tri_plan(int ny, size_t lds_limit, bool poisson, int mm_req, int lpb_req, int *P_out, int *mm_out,
             int *lpb_out) {
  return poisson ? tri_choose(ny, lds_limit, mm_req, lpb_req, kPoissonShapes, 8, P_out, mm_out, lpb_out)
                 : tri_choose(ny, lds_limit, mm_req, lpb_req, kCnShapes, 5, P_out, mm_out, lpb_out);
}

int launch_tri_poisson(ptc_ctx *c, void *spec) {
  PoissonPlan &P = c->pp;
  const bool f32 = P.prec == PTC_PREC_FP32_REFINE;
  if (P.tri == PTC_TRI_THOMAS) {
    PTC_CHECK(!f32, PTC_E_UNSUPPORTED, "the Thomas y-solve has no FP32-storage form");
    dim3 blk(64, 4);
    dim3 grid((unsigned)((P.h + 1 + 63) / 64), (unsigned)((c->nz + 3) / 4));
    hipLaunchKernelGGL(k_tri_poisson_thomas, grid, blk, 0, c->stream, (double2 *)spec, P.d_tri_scratch, c->d_met,
                       c->ny + 2, c->d_wnx, c->d_wnz, c->ny, P.pitch, P.h, c->nz);
    PTC_HIP(hipGetLastError());
    return PTC_OK;
  }
  if (P.tri_cl > 0) {
    const int nP = P.tri_chunks, cl = P.tri_cl;
    dim3 grid((unsigned)((P.h + 1 + cl - 1) / cl), (unsigned)c->nz);
    const size_t lds = sizeof(double) * tri_cplx_lds_doubles(c->ny + 2, nP, cl);
#define PTC_CPLX(MM, CL)                                                                                       \
  if (P.tri_mmax == MM && cl == CL) {                                                                          \
    if (f32 && P.stream)                                                                                       \
      hipLaunchKernelGGL((k_tri_poisson_cplx<MM, CL, float2, true>), grid, dim3(CL * nP), lds, c->stream,      \
                         (float2 *)spec, c->d_met, c->ny + 2, c->d_wnx, c->d_wnz, c->ny, P.pitch, P.h, nP);    \
    else if (f32)                                                                                              \
      hipLaunchKernelGGL((k_tri_poisson_cplx<MM, CL, float2, false>), grid, dim3(CL * nP), lds, c->stream,     \
                         (float2 *)spec, c->d_met, c->ny + 2, c->d_wnx, c->d_wnz, c->ny, P.pitch, P.h, nP);    \
    else if (P.stream)                                                                                         \
      hipLaunchKernelGGL((k_tri_poisson_cplx<MM, CL, double2, true>), grid, dim3(CL * nP), lds, c->stream,     \
                         (double2 *)spec, c->d_met, c->ny + 2, c->d_wnx, c->d_wnz, c->ny, P.pitch, P.h, nP);   \
    else                                                                                                       \
      hipLaunchKernelGGL((k_tri_poisson_cplx<MM, CL, double2, false>), grid, dim3(CL * nP), lds, c->stream,    \
                         (double2 *)spec, c->d_met, c->ny + 2, c->d_wnx, c->d_wnz, c->ny, P.pitch, P.h, nP);   \
    PTC_HIP(hipGetLastError());                                                                                \
    return PTC_OK;                                                                                             \
  }
    PTC_CPLX(4, 16)
    PTC_CPLX(8, 16)
    PTC_CPLX(12, 16)
    PTC_CPLX(16, 16)
    PTC_CPLX(4, 32)
    PTC_CPLX(8, 32)
    PTC_CPLX(12, 32)
    PTC_CPLX(16, 32)
#undef PTC_CPLX
    set_error("bad complex-line tridiagonal launch shape");
    return PTC_E_STATE;
  }
  if (P.tri_ll > 0) {
    const int nP = P.tri_chunks, ll = P.tri_ll;
    dim3 grid((unsigned)((2 * (P.h + 1) + ll - 1) / ll), (unsigned)c->nz);
    const size_t lds = tri_lds_bytes(c->ny, nP, ll);
#define PTC_LDS(MM, LL)                                                                                        \
  if (P.tri_mmax == MM && ll == LL) {                                                                          \
    if (f32)                                                                                                   \
      hipLaunchKernelGGL((k_tri_poisson_lds<MM, LL, float2>), grid, dim3(LL * nP), lds, c->stream,             \
                         (float2 *)spec, c->d_met, c->ny + 2, c->d_wnx, c->d_wnz, c->ny, P.pitch, P.h, nP);    \
    else                                                                                                       \
      hipLaunchKernelGGL((k_tri_poisson_lds<MM, LL, double2>), grid, dim3(LL * nP), lds, c->stream,            \
                         (double2 *)spec, c->d_met, c->ny + 2, c->d_wnx, c->d_wnz, c->ny, P.pitch, P.h, nP);   \
    PTC_HIP(hipGetLastError());                                                                                \
    return PTC_OK;                                                                                             \
  }
    PTC_LDS(4, 16)
    PTC_LDS(8, 16)
    PTC_LDS(12, 16)
    PTC_LDS(16, 16)
    PTC_LDS(4, 32)
    PTC_LDS(8, 32)
    PTC_LDS(12, 32)
    PTC_LDS(16, 32)
#undef PTC_LDS
    set_error("bad LDS tridiagonal launch shape");
    return PTC_E_STATE;
  }
  const int nP = P.tri_chunks, lpb = P.tri_lpb;
  // lpb lanes per block = lpb/2 complex lines
  dim3 grid((unsigned)((2 * (P.h + 1) + lpb - 1) / lpb), (unsigned)c->nz);
  const size_t lds = (size_t)4 * nP * lpb * 24;
#define PTC_CASE(MM, LPB)                                                                                     \
  if (P.tri_mmax == MM && lpb == LPB) {                                                                       \
    if (f32)                                                                                                  \
      hipLaunchKernelGGL((k_tri_poisson_split<MM, LPB, float2>), grid, dim3(LPB * nP), lds, c->stream,        \
                         (float2 *)spec, c->d_met, c->ny + 2, c->d_wnx, c->d_wnz, c->ny, P.pitch, P.h, nP);   \
    else                                                                                                      \
      hipLaunchKernelGGL((k_tri_poisson_split<MM, LPB, double2>), grid, dim3(LPB * nP), lds, c->stream,       \
                         (double2 *)spec, c->d_met, c->ny + 2, c->d_wnx, c->d_wnz, c->ny, P.pitch, P.h, nP);  \
    PTC_HIP(hipGetLastError());                                                                               \
    return PTC_OK;                                                                                            \
  }
  PTC_CASE(4, 64)
  PTC_CASE(8, 64)
  PTC_CASE(12, 64)
  PTC_CASE(16, 64)
  PTC_CASE(8, 32)
  PTC_CASE(12, 32)
  PTC_CASE(16, 32)
  PTC_CASE(16, 16)
#undef PTC_CASE
  set_error("bad tridiagonal launch shape");
  return PTC_E_STATE;
}

// kind 0 (u/w), 1 (v), 2 (scalar); shift: optional device scalar added to rhs
CnJob cn_job(int kind, double *u, const double *rhs, const double *duy_ym, double beta, const double *shift) {
  CnJob J{};
  J.wall = nullptr;
  J.top = 0.0;
  J.uym_upd = nullptr;
  J.duym_upd = nullptr;
  J.u = u;
  J.rhs = rhs;
  J.duy_ym = duy_ym;
  J.shift = shift;
  J.beta = beta;
  J.kind = kind;
  J.tab = 0;
  J.need = 1;
  J.fold = 0;
  J.sbc = 0;
  J.r1m = J.r2m = J.r1p = J.r2p = 0.0;
  if (kind == 0) { J.rA = M_AM2CCN; J.rB = M_AC2CCN; J.rC = M_AP2CCN; }
  else if (kind == 1) { J.rA = M_AM2P; J.rB = M_AC2P; J.rC = M_AP2P; }
  else { J.rA = M_AM2CSCCN; J.rB = M_AC2CSCCN; J.rC = M_AP2CSCCN; }
  return J;
}

// the factor launch (where a job needs one) and the solve launch (grid z = job)
static int launch_cn(ptc_ctx *c, const CnJobs &jobs, bool factor, bool solve, const FlowFold &ff) {
  int nP = 1, mmax = 16, lpb = 64;
  PTC_TRY(tri_plan(c->ny, c->lds_limit, false, 0, 0, &nP, &mmax, &lpb));
  PTC_CHECK(nP * mmax <= 256 && 2 * nP <= 64, PTC_E_STATE, "Crank-Nicolson chunk plan outside the factor table");
  dim3 grid((unsigned)((c->nx + lpb - 1) / lpb), (unsigned)c->nz, (unsigned)jobs.n);
  const size_t lds = sizeof(double) * 2 * (size_t)(2 * nP) * lpb;
#define PTC_CN(MM, LPB)                                                                                        \
  if (mmax == MM && lpb == LPB) {                                                                              \
    if (factor) {                                                                                              \
      hipLaunchKernelGGL((k_cn_factor<MM>), dim3(jobs.n), dim3(64), 0, c->stream, jobs, c->d_met, c->ny + 2,  \
                         c->ny, nP, c->d_cnfac);                                                               \
      PTC_HIP(hipGetLastError());                                                                              \
    }                                                                                                          \
    if (solve) {                                                                                               \
      hipLaunchKernelGGL((k_tri_cn_fac<MM, LPB>), grid, dim3(LPB * nP), lds, c->stream, jobs, c->d_cnfac,      \
                         c->nx, c->ny, nP, ff);                                                                \
      PTC_HIP(hipGetLastError());                                                                              \
    }                                                                                                          \
    return PTC_OK;                                                                                             \
  }
  PTC_CN(4, 64)
  PTC_CN(8, 64)
  PTC_CN(12, 64)
  PTC_CN(8, 32)
  PTC_CN(16, 32)
#undef PTC_CN
  set_error("bad tridiagonal launch shape");
  return PTC_E_STATE;
}

// one launch for all the jobs, after the factor tables of this launch (the
// per-launch tables 0 .. n-1)
int launch_tri_cn_jobs(ptc_ctx *c, const CnJobs &jobs) {
  PTC_CHECK(jobs.n >= 1 && jobs.n <= kCnMaxJobs, PTC_E_ARG, "bad Crank-Nicolson job count");
  CnJobs J = jobs;
  for (int q = 0; q < J.n; ++q) {
    J.j[q].tab = q;
    J.j[q].need = 1;
    PTC_CHECK(!J.j[q].fold, PTC_E_ARG, "flow-rate folding is for the step's launch");
  }
  return launch_cn(c, J, true, true, FlowFold{});
}

// The matrices of a fixed-dt step repeat every step: per substep, one for u
// and w, one for v, one for the scalar.  Their tables are kept (tables
// kCnMaxJobs + 3 (ns - 1) + kind) with the beta each was formed for, and the
// factor launch runs only for a table whose beta differs (the first step,
// or every substep of a CFL-adaptive one).  At 64 x 48 x 64 that launch is
// about 4.5 us of a 75 us substep.
int launch_tri_cn_step(ptc_ctx *c, CnJobs jobs, int ns, bool solve, const FlowFold *ff) {
  PTC_CHECK(jobs.n >= 1 && jobs.n <= kCnMaxJobs, PTC_E_ARG, "bad Crank-Nicolson job count");
  PTC_CHECK(ns >= 1 && ns <= 3, PTC_E_ARG, "substep index out of range");
  int nP = 1, mmax = 16, lpb = 64;
  PTC_TRY(tri_plan(c->ny, c->lds_limit, false, 0, 0, &nP, &mmax, &lpb));
  if (mmax != c->cn_tab_mm || nP != c->cn_tab_p) {
    for (double &b : c->cn_tab_beta) b = NAN;
    c->cn_tab_mm = mmax;
    c->cn_tab_p = nP;
  }
  bool any = false, formed[kCnStepTables] = {};
  for (int q = 0; q < jobs.n; ++q) {
    CnJob &J = jobs.j[q];
    PTC_CHECK(J.kind >= 0 && J.kind <= 2, PTC_E_ARG, "bad Crank-Nicolson job kind");
    PTC_CHECK(!solve || !J.fold || (ff != nullptr && J.shift == nullptr), PTC_E_ARG,
              "flow-rate folding without its sums");
    const int t = 3 * (ns - 1) + J.kind;
    J.tab = kCnMaxJobs + t;
    for (int r = 0; r < q; ++r)
      PTC_CHECK(jobs.j[r].tab != J.tab || jobs.j[r].beta == J.beta, PTC_E_ARG,
                "two betas for one Crank-Nicolson table");
    // formed by an earlier launch (the cached beta is NaN for none), or by an
    // earlier job of this one (u and w share a table)
    J.need = !(c->cn_tab_beta[t] == J.beta) && !formed[t];
    if (J.need) {
      formed[t] = true;
      any = true;
    }
  }
  PTC_TRY(launch_cn(c, jobs, any, solve, ff ? *ff : FlowFold{}));
  for (int q = 0; q < jobs.n; ++q)
    if (jobs.j[q].need) c->cn_tab_beta[jobs.j[q].tab - kCnMaxJobs] = jobs.j[q].beta;
  return PTC_OK;
}

int launch_tri_cn(ptc_ctx *c, int kind, double *u, const double *rhs, const double *duy_ym, double beta,
                  const double *shift) {
  CnJobs jobs{};
  jobs.j[0] = cn_job(kind, u, rhs, duy_ym, beta, shift);
  jobs.n = 1;
  return launch_tri_cn_jobs(c, jobs);
}

}  // namespace ptc
