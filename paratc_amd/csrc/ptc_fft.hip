// ptc_fft.hip -- batched FFTs of the Poisson solve on gfx950.
//
// Replaces the FFTW r2r plans of the reference (R2HC / HC2R along x with
// stride 1 and along z with stride nx*ny, src/m_Poisson.f90:180-244) by two
// kernels that each make one read and one write pass over HBM:
//
//  * x pass: a real line of nx doubles is read as nx/2 complex pairs
//    z_m = x_{2m} + i x_{2m+1}, transformed with a complex FFT of length
//    h = nx/2 in LDS and unpacked to X_k, k = 0..h (the half spectrum).  The
//    inverse packs X_k back and applies the inverse FFT; HC2R(R2HC(x)) = nx x
//    as in FFTW, and the final 1/(nx nz) of the reference (normfft,
//    src/m_Poisson.f90:48,448) is applied on the store.
//  * z pass: complex FFTs of length nz on tiles of T consecutive kx columns,
//    so every global access moves T*16 contiguous bytes per z plane.
//
// The LDS FFT is a self-sorting (Stockham) mixed-radix transform, radix
// 2/3/4/5/7/8.  Each stage stages its butterfly inputs in registers, then
// writes the outputs back in place after a barrier, so one LDS buffer of the
// tile size suffices.  Twiddles w^m = exp(-2 pi i m / n) come from a table
// computed on the host in FP64.
#include <cmath>
#include <cstring>

#include "ptc_internal.hpp"

namespace ptc {

__device__ __forceinline__ double2 cmul(double2 a, double2 b) {
  return make_double2(a.x * b.x - a.y * b.y, a.x * b.y + a.y * b.x);
}
__device__ __forceinline__ double2 cadd(double2 a, double2 b) { return make_double2(a.x + b.x, a.y + b.y); }
__device__ __forceinline__ double2 csub(double2 a, double2 b) { return make_double2(a.x - b.x, a.y - b.y); }
// multiply by -i (forward) or +i (inverse)
template <bool INV>
__device__ __forceinline__ double2 mul_mi(double2 a) {
  return INV ? make_double2(-a.y, a.x) : make_double2(a.y, -a.x);
}

// ---- small DFTs in registers ---------------------------------------------
template <int R, bool INV>
struct Dft;

template <bool INV>
struct Dft<2, INV> {
  static __device__ __forceinline__ void run(double2 *v) {
    double2 a = v[0], b = v[1];
    v[0] = cadd(a, b);
    v[1] = csub(a, b);
  }
};

template <bool INV>
struct Dft<4, INV> {
  static __device__ __forceinline__ void run(double2 *v) {
    double2 t0 = cadd(v[0], v[2]), t1 = csub(v[0], v[2]);
    double2 t2 = cadd(v[1], v[3]), t3 = mul_mi<INV>(csub(v[1], v[3]));
    v[0] = cadd(t0, t2);
    v[1] = cadd(t1, t3);
    v[2] = csub(t0, t2);
    v[3] = csub(t1, t3);
  }
};

template <bool INV>
struct Dft<8, INV> {
  static __device__ __forceinline__ void run(double2 *v) {
    constexpr double r2 = 0.70710678118654752440;
    double2 e[4] = {v[0], v[2], v[4], v[6]};
    double2 o[4] = {v[1], v[3], v[5], v[7]};
    Dft<4, INV>::run(e);
    Dft<4, INV>::run(o);
    // w8^k, k = 0..3 (forward exp(-i pi k / 4))
    double2 w1 = INV ? make_double2(r2, r2) : make_double2(r2, -r2);
    double2 w3 = INV ? make_double2(-r2, r2) : make_double2(-r2, -r2);
    double2 o1 = cmul(o[1], w1);
    double2 o2 = mul_mi<INV>(o[2]);
    double2 o3 = cmul(o[3], w3);
    v[0] = cadd(e[0], o[0]);
    v[4] = csub(e[0], o[0]);
    v[1] = cadd(e[1], o1);
    v[5] = csub(e[1], o1);
    v[2] = cadd(e[2], o2);
    v[6] = csub(e[2], o2);
    v[3] = cadd(e[3], o3);
    v[7] = csub(e[3], o3);
  }
};

// odd radix by the symmetric (pairwise) form:
//   X_r = v0 + sum_q (v_q + v_{R-q}) cos(2 pi q r / R) -/+ i sum_q (v_q - v_{R-q}) sin(2 pi q r / R)
template <int R>
struct OddTab;
template <>
struct OddTab<3> {
  static __device__ __forceinline__ double c(int m) { return m == 0 ? 1.0 : -0.5; }
  static __device__ __forceinline__ double s(int m) {
    return m == 0 ? 0.0 : (m == 1 ? 0.86602540378443864676 : -0.86602540378443864676);
  }
};
template <>
struct OddTab<5> {
  static __device__ __forceinline__ double c(int m) {
    constexpr double t[5] = {1.0, 0.30901699437494742410, -0.80901699437494742410, -0.80901699437494742410,
                             0.30901699437494742410};
    return t[m];
  }
  static __device__ __forceinline__ double s(int m) {
    constexpr double t[5] = {0.0, 0.95105651629515357212, 0.58778525229247312917, -0.58778525229247312917,
                             -0.95105651629515357212};
    return t[m];
  }
};
template <>
struct OddTab<7> {
  static __device__ __forceinline__ double c(int m) {
    constexpr double t[7] = {1.0, 0.62348980185873353053, -0.22252093395631440429, -0.90096886790241912624,
                             -0.90096886790241912624, -0.22252093395631440429, 0.62348980185873353053};
    return t[m];
  }
  static __device__ __forceinline__ double s(int m) {
    constexpr double t[7] = {0.0, 0.78183148246802980871, 0.97492791218182360702, 0.43388373911755812048,
                             -0.43388373911755812048, -0.97492791218182360702, -0.78183148246802980871};
    return t[m];
  }
};

template <int R, bool INV>
struct DftOdd {
  static __device__ __forceinline__ void run(double2 *v) {
    constexpr int H = (R - 1) / 2;
    double2 a[H + 1], b[H + 1];
#pragma unroll
    for (int q = 1; q <= H; ++q) {
      a[q] = cadd(v[q], v[R - q]);
      b[q] = csub(v[q], v[R - q]);
    }
    double2 x0 = v[0];
#pragma unroll
    for (int q = 1; q <= H; ++q) x0 = cadd(x0, a[q]);
    double2 out[R];
    out[0] = x0;
#pragma unroll
    for (int r = 1; r <= H; ++r) {
      double2 A = v[0];
      double2 B = make_double2(0.0, 0.0);
#pragma unroll
      for (int q = 1; q <= H; ++q) {
        const double cq = OddTab<R>::c((q * r) % R), sq = OddTab<R>::s((q * r) % R);
        A.x += a[q].x * cq;
        A.y += a[q].y * cq;
        B.x += b[q].x * sq;
        B.y += b[q].y * sq;
      }
      // forward: X_r = A - i B, X_{R-r} = A + i B
      if (!INV) {
        out[r] = make_double2(A.x + B.y, A.y - B.x);
        out[R - r] = make_double2(A.x - B.y, A.y + B.x);
      } else {
        out[r] = make_double2(A.x - B.y, A.y + B.x);
        out[R - r] = make_double2(A.x + B.y, A.y - B.x);
      }
    }
#pragma unroll
    for (int r = 0; r < R; ++r) v[r] = out[r];
  }
};
template <bool INV>
struct Dft<3, INV> : DftOdd<3, INV> {};
template <bool INV>
struct Dft<5, INV> : DftOdd<5, INV> {};
template <bool INV>
struct Dft<7, INV> : DftOdd<7, INV> {};

// ---- one Stockham stage on ncol interleaved columns: element (col, m) lives at
//      buf[m * ncol + col].  Every thread stages at most EMAX/R butterflies. ----
template <int R, int EMAX, bool INV>
__device__ __forceinline__ void fft_stage(double2 *buf, int n, int ncol, int Ns, const double2 *__restrict__ tw,
                                          int tid, int nthr) {
  constexpr int BR = EMAX / R;
  const int nb = n / R;
  const int total = nb * ncol;
  double2 v[BR][R];
#pragma unroll
  for (int b = 0; b < BR; ++b) {
    const int t = tid + b * nthr;
    if (t < total) {
      const int col = t % ncol, jb = t / ncol;
#pragma unroll
      for (int r = 0; r < R; ++r) v[b][r] = buf[(jb + r * nb) * ncol + col];
    }
  }
  __syncthreads();
#pragma unroll
  for (int b = 0; b < BR; ++b) {
    const int t = tid + b * nthr;
    if (t < total) {
      const int col = t % ncol, jb = t / ncol;
      const int k = jb % Ns;
      if (k != 0) {
        const int step = k * (n / (Ns * R));  // r * step < n for r < R
#pragma unroll
        for (int r = 1; r < R; ++r) {
          double2 w = tw[r * step];
          if (INV) w.y = -w.y;
          v[b][r] = cmul(v[b][r], w);
        }
      }
      Dft<R, INV>::run(v[b]);
      const int idxD = (jb / Ns) * Ns * R + k;
#pragma unroll
      for (int r = 0; r < R; ++r) buf[(idxD + r * Ns) * ncol + col] = v[b][r];
    }
  }
  __syncthreads();
}

template <int EMAX, bool INV>
__device__ void lds_fft(double2 *buf, const FftLen &L, int ncol, int tid, int nthr) {
  int Ns = 1;
  for (int s = 0; s < L.nstage; ++s) {
    const int R = L.radix[s];
    switch (R) {
      case 8: fft_stage<8, EMAX, INV>(buf, L.n, ncol, Ns, L.tw, tid, nthr); break;
      case 4: fft_stage<4, EMAX, INV>(buf, L.n, ncol, Ns, L.tw, tid, nthr); break;
      case 2: fft_stage<2, EMAX, INV>(buf, L.n, ncol, Ns, L.tw, tid, nthr); break;
      case 3: fft_stage<3, EMAX, INV>(buf, L.n, ncol, Ns, L.tw, tid, nthr); break;
      case 5: fft_stage<5, EMAX, INV>(buf, L.n, ncol, Ns, L.tw, tid, nthr); break;
      case 7: fft_stage<7, EMAX, INV>(buf, L.n, ncol, Ns, L.tw, tid, nthr); break;
      default: break;
    }
    Ns *= R;
  }
}

// ---- x pass, forward: real rows j = 1..ny of src -> half spectra in spec ----
// line index q = (j - 1) + ny * k; spec[q * pitch + kx], kx = 0..h, zero to pitch
template <int EMAX>
__global__ void __launch_bounds__(1024) k_fft_x_fwd(const double *__restrict__ src, double2 *__restrict__ spec,
                                                     FftLen L, const double2 *__restrict__ wx, int ny, int nz,
                                                     int nlpb, int pitch, long long rowlen) {
  extern __shared__ double2 lds[];
  const int h = L.n;
  const int tid = threadIdx.x, nthr = blockDim.x;
  const long long nlines = (long long)ny * nz;
  const long long q0 = (long long)blockIdx.x * nlpb;
  for (int e = tid; e < h * nlpb; e += nthr) {
    const int l = e / h, m = e - l * h;
    const long long q = q0 + l;
    double2 v = make_double2(0.0, 0.0);
    if (q < nlines) {
      const long long j = q % ny + 1, k = q / ny;
      const double2 *row = reinterpret_cast<const double2 *>(src + rowlen * (j + (long long)(ny + 2) * k));
      v = row[m];
    }
    lds[m * nlpb + l] = v;
  }
  __syncthreads();
  lds_fft<EMAX, false>(lds, L, nlpb, tid, nthr);
  for (int e = tid; e < pitch * nlpb; e += nthr) {
    const int l = e / pitch, kx = e - l * pitch;
    const long long q = q0 + l;
    if (q >= nlines) continue;
    double2 X = make_double2(0.0, 0.0);
    if (kx <= h) {
      const double2 zk = lds[(kx == h ? 0 : kx) * nlpb + l];
      double2 zc = lds[(kx == 0 ? 0 : h - kx) * nlpb + l];
      zc.y = -zc.y;
      const double2 s = cadd(zk, zc), d = csub(zk, zc);
      const double2 wd = cmul(wx[kx], d);
      // X = (s - i w d) / 2
      X = make_double2(0.5 * (s.x + wd.y), 0.5 * (s.y - wd.x));
    }
    spec[q * pitch + kx] = X;
  }
}

// ---- x pass, inverse: half spectra -> real rows, scaled ----
template <int EMAX>
__global__ void __launch_bounds__(1024) k_fft_x_inv(const double2 *__restrict__ spec, double *__restrict__ dst,
                                                     FftLen L, const double2 *__restrict__ wx, int ny, int nz,
                                                     int nlpb, int pitch, long long rowlen, double scale) {
  extern __shared__ double2 lds[];
  const int h = L.n;
  const int tid = threadIdx.x, nthr = blockDim.x;
  const long long nlines = (long long)ny * nz;
  const long long q0 = (long long)blockIdx.x * nlpb;
  for (int e = tid; e < h * nlpb; e += nthr) {
    const int l = e / h, k = e - l * h;
    const long long q = q0 + l;
    double2 Z = make_double2(0.0, 0.0);
    if (q < nlines) {
      const double2 xk = spec[q * pitch + k];
      double2 xc = spec[q * pitch + (h - k)];
      xc.y = -xc.y;
      const double2 s = cadd(xk, xc), d = csub(xk, xc);
      double2 w = wx[k];
      w.y = -w.y;
      const double2 wd = cmul(w, d);
      // Z = s + i w^-1 d
      Z = make_double2(s.x - wd.y, s.y + wd.x);
    }
    lds[k * nlpb + l] = Z;
  }
  __syncthreads();
  lds_fft<EMAX, true>(lds, L, nlpb, tid, nthr);
  for (int e = tid; e < h * nlpb; e += nthr) {
    const int l = e / h, m = e - l * h;
    const long long q = q0 + l;
    if (q >= nlines) continue;
    const long long j = q % ny + 1, k = q / ny;
    double2 *row = reinterpret_cast<double2 *>(dst + rowlen * (j + (long long)(ny + 2) * k));
    const double2 v = lds[m * nlpb + l];
    row[m] = make_double2(v.x * scale, v.y * scale);
  }
}

// ---- z pass: complex FFT of length nz on T columns (kx0 .. kx0+T-1) at one row j ----
template <int EMAX, bool INV>
__global__ void __launch_bounds__(1024) k_fft_z(double2 *__restrict__ spec, FftLen L, int ny, int pitch, int T) {
  extern __shared__ double2 lds[];
  const int n = L.n;
  const int tid = threadIdx.x, nthr = blockDim.x;
  const long long zs = (long long)pitch * ny;
  double2 *base = spec + (long long)blockIdx.y * pitch + (long long)blockIdx.x * T;
  for (int e = tid; e < n * T; e += nthr) {
    const int kz = e / T, t = e - kz * T;
    lds[e] = base[kz * zs + t];
  }
  __syncthreads();
  lds_fft<EMAX, INV>(lds, L, T, tid, nthr);
  for (int e = tid; e < n * T; e += nthr) {
    const int kz = e / T, t = e - kz * T;
    base[kz * zs + t] = lds[e];
  }
}

// ---- host side -----------------------------------------------------------

int fft_make_len(int n, FftLen *L) {
  std::memset(L, 0, sizeof(*L));
  L->n = n;
  int m = n;
  int ns = 0;
  while (m % 8 == 0) { L->radix[ns++] = 8; m /= 8; }
  if (m % 4 == 0) { L->radix[ns++] = 4; m /= 4; }
  if (m % 2 == 0) { L->radix[ns++] = 2; m /= 2; }
  for (int f : {3, 5, 7}) {
    while (m % f == 0) {
      if (ns >= 16) return PTC_E_UNSUPPORTED;
      L->radix[ns++] = f;
      m /= f;
    }
  }
  if (m != 1) {
    set_error("FFT length " + std::to_string(n) + " has a prime factor other than 2, 3, 5, 7");
    return PTC_E_UNSUPPORTED;
  }
  L->nstage = ns;
  return PTC_OK;
}

// Largest staged-butterfly demand of a plan: every stage must fit
// ceil(n*ncol/R) <= nthr * (EMAX / R).
static bool stage_fits(const FftLen &L, int ncol, int nthr, int emax) {
  for (int s = 0; s < L.nstage; ++s) {
    const int R = L.radix[s];
    const long long need = (long long)(L.n / R) * ncol;
    if (need > (long long)nthr * (emax / R)) return false;
  }
  return true;
}

template <int EMAX>
static int launch_x_fwd_t(ptc_ctx *c, const double *src, double2 *spec) {
  PoissonPlan &P = c->pp;
  const long long nlines = (long long)c->ny * c->nz;
  const unsigned nblk = (unsigned)((nlines + P.x_lines - 1) / P.x_lines);
  const size_t lds = sizeof(double2) * (size_t)P.fx.n * P.x_lines;
  hipLaunchKernelGGL(k_fft_x_fwd<EMAX>, dim3(nblk), dim3(P.x_threads), lds, c->stream, src, spec, P.fx, P.d_wx,
                     c->ny, c->nz, P.x_lines, P.pitch, (long long)c->nx);
  PTC_HIP(hipGetLastError());
  return PTC_OK;
}

int launch_fft_x_fwd(ptc_ctx *c, const double *src, void *spec_, bool f64) {
  if (c->pp.x_fast) return fft2_x_fwd(c, src, spec_, 0, c->ny, f64);
  double2 *spec = static_cast<double2 *>(spec_);  // generic kernels: FP64 spectra only
  return stage_fits(c->pp.fx, c->pp.x_lines, c->pp.x_threads, 16) ? launch_x_fwd_t<16>(c, src, spec)
                                                                  : launch_x_fwd_t<32>(c, src, spec);
}

template <int EMAX>
static int launch_x_inv_t(ptc_ctx *c, const double2 *spec, double *dst, double scale) {
  PoissonPlan &P = c->pp;
  const long long nlines = (long long)c->ny * c->nz;
  const unsigned nblk = (unsigned)((nlines + P.x_lines - 1) / P.x_lines);
  const size_t lds = sizeof(double2) * (size_t)P.fx.n * P.x_lines;
  hipLaunchKernelGGL(k_fft_x_inv<EMAX>, dim3(nblk), dim3(P.x_threads), lds, c->stream, spec, dst, P.fx, P.d_wx,
                     c->ny, c->nz, P.x_lines, P.pitch, (long long)c->nx, scale);
  PTC_HIP(hipGetLastError());
  return PTC_OK;
}

// ghost: the specialised kernel also fills the Neumann ghost rows; the generic
// one fills them with a separate boundary launch
int launch_fft_x_inv(ptc_ctx *c, const void *spec_, double *dst, double scale, bool acc, bool ghost) {
  if (c->pp.x_fast) return fft2_x_inv(c, spec_, dst, scale, 0, c->ny, acc, ghost);
  PTC_CHECK(!acc, PTC_E_UNSUPPORTED, "accumulating x-inverse needs the specialised kernels");
  if (ghost) {
    PTC_TRY(launch_fft_x_inv(c, spec_, dst, scale, false, false));
    return launch_bc_pr(c, dst);
  }
  const double2 *spec = static_cast<const double2 *>(spec_);
  return stage_fits(c->pp.fx, c->pp.x_lines, c->pp.x_threads, 16) ? launch_x_inv_t<16>(c, spec, dst, scale)
                                                                  : launch_x_inv_t<32>(c, spec, dst, scale);
}

template <int EMAX, bool INV>
static int launch_z_t(ptc_ctx *c, double2 *spec) {
  PoissonPlan &P = c->pp;
  const size_t lds = sizeof(double2) * (size_t)P.fz.n * P.ztile;
  dim3 grid((unsigned)(P.pitch / P.ztile), (unsigned)c->ny);
  hipLaunchKernelGGL((k_fft_z<EMAX, INV>), grid, dim3(P.z_threads), lds, c->stream, spec, P.fz, c->ny, P.pitch,
                     P.ztile);
  PTC_HIP(hipGetLastError());
  return PTC_OK;
}

int launch_fft_z(ptc_ctx *c, void *spec_, bool inverse, bool f64) {
  if (c->pp.z_fast) return fft2_z(c, spec_, inverse, 0, c->ny, f64);
  double2 *spec = static_cast<double2 *>(spec_);
  const bool small = stage_fits(c->pp.fz, c->pp.ztile, c->pp.z_threads, 16);
  if (inverse) return small ? launch_z_t<16, true>(c, spec) : launch_z_t<32, true>(c, spec);
  return small ? launch_z_t<16, false>(c, spec) : launch_z_t<32, false>(c, spec);
}

// exported for the plan builder: does (ncol, nthr) fit with EMAX <= 32?
bool fft_config_ok(const FftLen &L, int ncol, int nthr) { return stage_fits(L, ncol, nthr, 32); }

}  // namespace ptc
