// ptc_spectra.hip -- the energy spectra of clcStat on the device
// (src/m_FlowCase.f90:430-805; the helpers clcEnergySpectraX/Z, clcLTS_*,
// clcCospectraX/Z, clcEnergySpectra2D, clcCosSpectra2D at :954-1375).
//
// The reference transforms each field with R2HC along x (and, separately,
// along z, and for the 2-D spectra along x then z) and sums products of the
// half-complex values over the other direction.  Here each field goes once
// through the Poisson solver's forward x and z passes, in FP64, giving the
// 2-D spectrum F(kx, kz) of every row (kx = 0..nx/2, kz = 0..nz-1), and every
// quantity is read off F:
//   - x spectra: the z sum of Xa(kx, z) conj Xb(kx, z) over z equals
//     (1/nz) sum over kz of Fa conj Fb (Parseval along z);
//   - z spectra: the x sum of Za(x, kz) conj Zb(x, kz) equals (1/nx) times the
//     sum over kx = 0..nx/2 of P(kx, kz) plus the sum over kx = 1..nx/2-1 of
//     conj P(kx, -kz), P = Fa conj Fb (Parseval along x, F(-kx, kz) =
//     conj F(kx, -kz) for real fields);
//   - 2-D spectra: the R2HC along z of the real and imaginary parts of the x
//     transform are (F(kx, kz) + conj F(kx, -kz)) / 2 and
//     (F(kx, kz) - conj F(kx, -kz)) / 2i.
// The linear-coherence spectra are the real and imaginary parts of the same
// sums with b the reference row.  The reference's uCRF shift of ux only moves
// F(0, 0) by nx nz uCRF, which the partial sums add.  The cell-centre ux
// (the 4-point interpolation ic1 ux(i-1) + ic2 ux(i) + ic3 ux(i+1) + ic4 ux(i+2)
// + uCRF) is not transformed separately: its spectrum is H(kx) F_ux(kx, kz)
// with H = ic1 w + ic2 + ic3 conj(w) + ic4 conj(w)^2, w = exp(-2 pi i kx / nx),
// plus the same shift.
#include "ptc_internal.hpp"

#include <algorithm>
#include <cmath>

namespace ptc {

namespace {

constexpr int kSpecThreads = 256;  // partial-sum workgroup
constexpr int kSpecRows = 16;      // kz rows per partial-sum workgroup

// ---- the cell-centre uy (:576-587): the two-face mean into rows 1..ny of a field
__global__ void __launch_bounds__(256) k_spec_cell(Geo g, const double *__restrict__ u, double *__restrict__ out) {
  const int i = blockIdx.x * 256 + threadIdx.x, j = 1 + blockIdx.y, k = blockIdx.z;
  if (i >= g.nx) return;
  const long long o = (long long)g.nx * j + g.plane * k + i;
  out[o] = 0.5 * (u[o] + u[o + g.nx]);
}

__device__ inline double2 cmul(double2 a, double2 b) { return make_double2(a.x * b.x - a.y * b.y, a.x * b.y + a.y * b.x); }

// One pair of spectra to multiply, P = Fa conj Fb.  b's row is the same row,
// or (lcs) the reference row: jr1 below row jsplit, jr2 from it on.  ha: a is
// first multiplied by hx[kx] (the cell-centre interpolation's symbol).
struct SpecPair {
  const double2 *a, *b;
  double sa, sb;  // added to Re F(0, 0) of a and b (after the multiplication)
  int lcs, jr1, jr2, jsplit;
  const double2 *ha;
};

// Partial sums of one row j and kSpecRows kz rows (grid: x = kz chunk, y = j):
//   px[(j * nch + chunk) * (h+1) + kx] = sum over the chunk's kz of P(kx, kz)
//   pz[2 * (j * nz + kz)]     = sum over kx = 0..h of P(kx, kz)
//   pz[2 * (j * nz + kz) + 1] = sum over kx = 1..h-1 of P(kx, kz)
// Thread t owns kx = t + 256 s, s < NS.  Rows go in batches of spec_batch
// (fewer for more kx per thread, to stay within the registers): all their
// loads are issued before the products and the wavefront sums.
template <int NS>
constexpr int spec_batch() { return NS == 1 ? 4 : NS <= 3 ? 2 : 1; }

template <int NS, bool SAME>
__global__ void __launch_bounds__(kSpecThreads) k_spec_part(SpecPair P, int ny, int nz, int h, int pitch,
                                                            double2 *__restrict__ px, double2 *__restrict__ pz) {
  constexpr int NW = kSpecThreads / 64, RB = spec_batch<NS>();
  __shared__ double red[kSpecRows][NW][4];
  const int j = blockIdx.y, ch = blockIdx.x, nch = gridDim.x;
  const int t = threadIdx.x, lane = t & 63, w = t >> 6;
  const int k0 = ch * kSpecRows, k1 = min(nz, k0 + kSpecRows);
  const int jb = P.lcs ? (j < P.jsplit ? P.jr1 : P.jr2) : j;
  double sxr[NS], sxi[NS];
  double2 hs[NS];
#pragma unroll
  for (int s = 0; s < NS; ++s) {
    const int kx = t + kSpecThreads * s;
    sxr[s] = sxi[s] = 0.0;
    hs[s] = (P.ha && kx <= h) ? P.ha[kx] : make_double2(1.0, 0.0);
  }
  for (int kb = k0; kb < k1; kb += RB) {
    double2 va[RB][NS], vb[RB][NS];
#pragma unroll
    for (int r = 0; r < RB; ++r) {
      const bool live = kb + r < k1;
      const long long kz = live ? kb + r : k0;
      const double2 *ra = P.a + (kz * ny + j) * pitch;
      const double2 *rb = P.b + (kz * ny + jb) * pitch;
#pragma unroll
      for (int s = 0; s < NS; ++s) {
        const int kx = t + kSpecThreads * s;
        const bool in = live && kx <= h;
        va[r][s] = in ? ra[kx] : make_double2(0.0, 0.0);
        if (!SAME) vb[r][s] = in ? rb[kx] : make_double2(0.0, 0.0);
      }
    }
    double q[RB][4];
#pragma unroll
    for (int r = 0; r < RB; ++r) {
      q[r][0] = q[r][1] = q[r][2] = q[r][3] = 0.0;
      const bool origin = kb + r == 0;
#pragma unroll
      for (int s = 0; s < NS; ++s) {
        const int kx = t + kSpecThreads * s;
        double2 a = P.ha ? cmul(va[r][s], hs[s]) : va[r][s];
        if (origin && kx == 0) a.x += P.sa;
        double2 b = a;
        if (!SAME) {
          b = vb[r][s];
          if (origin && kx == 0) b.x += P.sb;
        }
        const double pr = a.x * b.x + a.y * b.y, pi = a.y * b.x - a.x * b.y;
        sxr[s] += pr;
        sxi[s] += pi;
        q[r][0] += pr;
        q[r][1] += pi;
        if (kx > 0 && kx < h) {
          q[r][2] += pr;
          q[r][3] += pi;
        }
      }
    }
#pragma unroll
    for (int off = 32; off > 0; off >>= 1)
#pragma unroll
      for (int r = 0; r < RB; ++r)
#pragma unroll
        for (int e = 0; e < 4; ++e) q[r][e] += __shfl_xor(q[r][e], off);
    if (lane == 0) {
#pragma unroll
      for (int r = 0; r < RB; ++r)
        if (kb + r < k1)
#pragma unroll
          for (int e = 0; e < 4; ++e) red[kb + r - k0][w][e] = q[r][e];
    }
  }
  double2 *ox = px + ((long long)j * nch + ch) * (h + 1);
#pragma unroll
  for (int s = 0; s < NS; ++s) {
    const int kx = t + kSpecThreads * s;
    if (kx <= h) ox[kx] = make_double2(sxr[s], sxi[s]);
  }
  __syncthreads();
  if (t < k1 - k0) {
    double v[4] = {0.0, 0.0, 0.0, 0.0};
#pragma unroll
    for (int qq = 0; qq < NW; ++qq)
#pragma unroll
      for (int e = 0; e < 4; ++e) v[e] += red[t][qq][e];
    double2 *oz = pz + 2 * ((long long)j * nz + k0 + t);
    oz[0] = make_double2(v[0], v[1]);
    oz[1] = make_double2(v[2], v[3]);
  }
}

// Adds one sample to the 1-D accumulators of row j (grid: y = j).
//   kind 0: energy or co-spectrum (weight 2 inside, 1 at kx = 0, nx/2),
//           real part into quantity m0;
//   kind 1: linear coherence, real part into m0 and imaginary part (zero at
//           the edge wavenumbers) into m1, no weight.
// ex(kx, j, m) and ez(kz, j, m) are column-major as EnergySpecX/Z.
__global__ void __launch_bounds__(256) k_spec_fin(const double2 *__restrict__ px, const double2 *__restrict__ pz,
                                                  int nch, int ny, int nx, int nz, int kind, int m0, int m1,
                                                  double *__restrict__ ex, double *__restrict__ ez) {
  const int j = blockIdx.y, h = nx / 2, nzh = nz / 2;
  // normEgy of the x and z helpers (:962, :1048)
  const double normx = 1.0 / ((double)nx * (double)nx * (double)nz);
  const double normz = 1.0 / ((double)nx * (double)nz * (double)nz);
  for (int kx = threadIdx.x; kx <= h; kx += blockDim.x) {
    double sr = 0.0, si = 0.0;
    for (int c = 0; c < nch; ++c) {
      const double2 v = px[((long long)j * nch + c) * (h + 1) + kx];
      sr += v.x;
      si += v.y;
    }
    sr = sr / (double)nz;
    si = si / (double)nz;
    double *o0 = ex + kx + (long long)(h + 1) * (j + (long long)ny * m0);
    if (kind == 0) {
      *o0 += sr * ((kx == 0 || kx == h) ? 1.0 : 2.0) * normx;
    } else {
      *o0 += sr * normx;
      if (kx > 0 && kx < h) ex[kx + (long long)(h + 1) * (j + (long long)ny * m1)] += si * normx;
    }
  }
  for (int kz = threadIdx.x; kz <= nzh; kz += blockDim.x) {
    const double2 r = pz[2 * ((long long)j * nz + kz)];
    const int km = kz == 0 ? 0 : nz - kz;
    const double2 q = pz[2 * ((long long)j * nz + km) + 1];
    const double vr = (r.x + q.x) / (double)nx, vi = (r.y - q.y) / (double)nx;
    double *o0 = ez + kz + (long long)(nzh + 1) * (j + (long long)ny * m0);
    if (kind == 0) {
      *o0 += vr * ((kz == 0 || kz == nzh) ? 1.0 : 2.0) * normz;
    } else {
      *o0 += vr * normz;
      if (kz > 0 && kz < nzh) ez[kz + (long long)(nzh + 1) * (j + (long long)ny * m1)] += vi * normz;
    }
  }
}

// The value at half-complex position (p, q) of the R2HC along x then z of
// spectral row `row` (0-based): p <= h is the real part of the x transform at
// kx = p, p > h its imaginary part at kx = nx - p; q likewise along z.
// hx (may be null) multiplies F first and shift is then added to Re F(0, 0).
__device__ inline double hc2d_value(const double2 *__restrict__ F, const double2 *__restrict__ hx, double shift,
                                    int ny, int nx, int nz, int pitch, int row, int p, int q) {
  const int h = nx / 2, nzh = nz / 2;
  const int kx = p <= h ? p : nx - p, kz = q <= nzh ? q : nz - q, km = kz == 0 ? 0 : nz - kz;
  double2 f1 = F[((long long)kz * ny + row) * pitch + kx];
  double2 f2 = F[((long long)km * ny + row) * pitch + kx];
  if (hx) {
    f1 = cmul(f1, hx[kx]);
    f2 = cmul(f2, hx[kx]);
  }
  if (kx == 0 && kz == 0) {
    f1.x += shift;
    f2.x += shift;
  }
  if (p <= h) return q <= nzh ? 0.5 * (f1.x + f2.x) : 0.5 * (f1.y - f2.y);
  return q <= nzh ? 0.5 * (f1.y + f2.y) : 0.5 * (f2.x - f1.x);
}

// Adds one sample to the 2-D accumulator of quantity m (grid: x = p / 64,
// y = chosen row, z = q).  mirror: 0 none (half channel), 1 the row ny+1-jct,
// 2 the face row ny+2-jct (uy); rcoe weighs the mirror's product.
__global__ void __launch_bounds__(64) k_spec_2d(const double2 *__restrict__ a, const double2 *__restrict__ b,
                                                const double2 *__restrict__ ha, double sa, int ny, int nx, int nz,
                                                int pitch, const int *__restrict__ iy, int ny2d, int mirror,
                                                double rcoe, double norm, double *__restrict__ out) {
  const int p = blockIdx.x * 64 + threadIdx.x, jc = blockIdx.y, q = blockIdx.z;
  if (p >= nx) return;
  const int jct = iy[jc];
  double v = hc2d_value(a, ha, sa, ny, nx, nz, pitch, jct - 1, p, q) *
             hc2d_value(b, nullptr, 0.0, ny, nx, nz, pitch, jct - 1, p, q);
  if (mirror) {
    const int jcs = (mirror == 2 ? ny + 2 : ny + 1) - jct;
    v += rcoe * (hc2d_value(a, ha, sa, ny, nx, nz, pitch, jcs - 1, p, q) *
                 hc2d_value(b, nullptr, 0.0, ny, nx, nz, pitch, jcs - 1, p, q));
  }
  out[p + (long long)nx * (jc + (long long)ny2d * q)] += norm * v;
}

template <int NS>
void part_launch(ptc_ctx *c, const SpecPair &P, bool same, int nch, double2 *px, double2 *pz) {
  const dim3 grid((unsigned)nch, (unsigned)c->ny);
  if (same)
    hipLaunchKernelGGL((k_spec_part<NS, true>), grid, dim3(kSpecThreads), 0, c->stream, P, c->ny, c->nz, c->nx / 2,
                       c->spec.pitch, px, pz);
  else
    hipLaunchKernelGGL((k_spec_part<NS, false>), grid, dim3(kSpecThreads), 0, c->stream, P, c->ny, c->nz, c->nx / 2,
                       c->spec.pitch, px, pz);
}

// kx values per thread, rounded up to an instantiated count
int part_slots(int h) {
  const int ns = (h + kSpecThreads) / kSpecThreads;
  for (int n : {1, 2, 3, 4, 6, 10, 19})
    if (ns <= n) return n;
  return 0;
}

int launch_part(ptc_ctx *c, const SpecPair &P, bool same) {
  SpectraState &S = c->spec;
  switch (part_slots(c->nx / 2)) {
    case 1: part_launch<1>(c, P, same, S.nch, S.d_px, S.d_pz); break;
    case 2: part_launch<2>(c, P, same, S.nch, S.d_px, S.d_pz); break;
    case 3: part_launch<3>(c, P, same, S.nch, S.d_px, S.d_pz); break;
    case 4: part_launch<4>(c, P, same, S.nch, S.d_px, S.d_pz); break;
    case 6: part_launch<6>(c, P, same, S.nch, S.d_px, S.d_pz); break;
    case 10: part_launch<10>(c, P, same, S.nch, S.d_px, S.d_pz); break;
    case 19: part_launch<19>(c, P, same, S.nch, S.d_px, S.d_pz); break;
    default: set_error("spectra: nx/2 above 4608"); return PTC_E_UNSUPPORTED;
  }
  PTC_HIP(hipGetLastError());
  return PTC_OK;
}

// one 1-D quantity (kind 0, m0) or the linear-coherence pair (kind 1, m0, m1);
// m are 1-based as the reference's
int spec_1d(ptc_ctx *c, const SpecPair &P, bool same, int kind, int m0, int m1) {
  SpectraState &S = c->spec;
  PTC_TRY(launch_part(c, P, same));
  hipLaunchKernelGGL(k_spec_fin, dim3(1, (unsigned)c->ny), dim3(256), 0, c->stream, S.d_px, S.d_pz, S.nch, c->ny,
                     c->nx, c->nz, kind, m0 - 1, m1 - 1, S.d_x, S.d_z);
  PTC_HIP(hipGetLastError());
  return PTC_OK;
}

int spec_2d(ptc_ctx *c, const double2 *a, const double2 *b, int m, int mirror, double rcoe,
            const double2 *ha = nullptr, double sa = 0.0) {
  SpectraState &S = c->spec;
  const int nx = c->nx, nz = c->nz;
  // normEgy of clcEnergySpectra2D / clcCosSpectra2D (:1253-1255)
  double norm = 1.0 / ((double)nx * (double)nz);
  norm = norm * norm;
  if (c->prm.flow_type == 1) norm = norm * 0.5;
  double *out = S.d_2d + (long long)nx * S.ny2d * nz * (m - 1);
  const dim3 grid((unsigned)((nx + 63) / 64), (unsigned)S.ny2d, (unsigned)nz);
  hipLaunchKernelGGL(k_spec_2d, grid, dim3(64), 0, c->stream, a, b, ha, sa, c->ny, nx, nz, S.pitch, S.d_iy2d, S.ny2d,
                     c->prm.flow_type == 1 ? mirror : 0, rcoe, norm, out);
  PTC_HIP(hipGetLastError());
  return PTC_OK;
}

int cell_uy(ptc_ctx *c, const double *uy) {
  const dim3 grid((unsigned)((c->nx + 255) / 256), (unsigned)c->ny, (unsigned)c->nz);
  hipLaunchKernelGGL(k_spec_cell, grid, dim3(256), 0, c->stream, c->geo, uy, c->spec.d_cell);
  PTC_HIP(hipGetLastError());
  return PTC_OK;
}

// the FP64 2-D spectrum of rows 1..ny of a field
int forward(ptc_ctx *c, const double *src, double2 *spec) {
  PTC_TRY(launch_fft_x_fwd(c, src, spec, true));
  return launch_fft_z(c, spec, false, true);
}

template <typename T>
int alloc(T **p, size_t n) {
  PTC_HIP(hipMalloc((void **)p, sizeof(T) * n));
  return PTC_OK;
}

template <typename T>
void release(T *&p) {
  if (p) (void)hipFree(p);
  p = nullptr;
}

size_t n1d_x(const ptc_ctx *c) { return (size_t)(c->nx / 2 + 1) * c->ny * c->spec.ne1; }
size_t n1d_z(const ptc_ctx *c) { return (size_t)(c->nz / 2 + 1) * c->ny * c->spec.ne1; }
size_t n2d(const ptc_ctx *c) { return (size_t)c->nx * c->spec.ny2d * c->nz * c->spec.ne2; }

// work buffers sized by the current Poisson plan's spectral pitch
int ensure_work(ptc_ctx *c) {
  SpectraState &S = c->spec;
  if (S.d_fa && S.pitch == c->pp.pitch) return PTC_OK;
  release(S.d_fa);
  release(S.d_fb);
  release(S.d_cell);
  release(S.d_px);
  release(S.d_pz);
  release(S.d_hx);
  S.pitch = c->pp.pitch;
  S.nch = (c->nz + kSpecRows - 1) / kSpecRows;
  const size_t nspec = (size_t)S.pitch * c->ny * c->nz;
  PTC_TRY(alloc(&S.d_fa, nspec));
  PTC_TRY(alloc(&S.d_fb, nspec));
  PTC_TRY(alloc(&S.d_cell, (size_t)c->nx * (c->ny + 2) * c->nz));
  PTC_TRY(alloc(&S.d_px, (size_t)c->ny * S.nch * (c->nx / 2 + 1)));
  PTC_TRY(alloc(&S.d_pz, (size_t)2 * c->ny * c->nz));
  // the symbol of the cell-centre interpolation of ux, kx = 0..nx/2
  const int h = c->nx / 2;
  const double ic[4] = {c->geo.ic1, c->geo.ic2, c->geo.ic3, c->geo.ic4};
  std::vector<double2> hx(h + 1);
  for (int kx = 0; kx <= h; ++kx) {
    const double th = 2.0 * M_PI * (double)kx / (double)c->nx;
    const double cs = std::cos(th), sn = std::sin(th), c2 = std::cos(2.0 * th), s2 = std::sin(2.0 * th);
    // ic1 e^{-i th} + ic2 + ic3 e^{i th} + ic4 e^{2 i th}
    hx[kx] = make_double2(ic[0] * cs + ic[1] + ic[2] * cs + ic[3] * c2, -ic[0] * sn + ic[2] * sn + ic[3] * s2);
  }
  PTC_TRY(alloc(&S.d_hx, hx.size()));
  PTC_HIP(hipMemcpy(S.d_hx, hx.data(), sizeof(double2) * hx.size(), hipMemcpyHostToDevice));
  return PTC_OK;
}

}  // namespace

void spectra_free(ptc_ctx *c) {
  SpectraState &S = c->spec;
  release(S.d_x);
  release(S.d_z);
  release(S.d_2d);
  release(S.d_iy2d);
  release(S.d_fa);
  release(S.d_fb);
  release(S.d_cell);
  release(S.d_px);
  release(S.d_pz);
  release(S.d_hx);
  S = SpectraState{};
}

// InitStatVar's spectra part (src/m_FlowCase.f90:140-251)
int spectra_init(ptc_ctx *c, int clc1d, int jlcs, int clc2d, int ny2d, const int *iy2d) {
  PTC_HIP(hipStreamSynchronize(c->stream));
  spectra_free(c);
  SpectraState &S = c->spec;
  const bool ch = c->prm.flow_type == 1;
  if (ny2d < 1) clc2d = 0;  // :209
  PTC_CHECK(!clc1d || (jlcs >= 1 && jlcs <= c->ny), PTC_E_ARG, "jForLCS outside 1..nyc");
  if (clc2d) {
    PTC_CHECK(iy2d != nullptr, PTC_E_ARG, "null iySpec2D");
    for (int q = 0; q < ny2d; ++q) {
      // :220-232
      if (ch) PTC_CHECK(iy2d[q] >= 2 && iy2d[q] <= c->ny / 2, PTC_E_ARG, "iySpec2D WRONG 1");
      else PTC_CHECK(iy2d[q] >= 2 && iy2d[q] <= c->ny, PTC_E_ARG, "iySpec2D WRONG 2");
    }
  }
  S.on1d = clc1d != 0;
  S.on2d = clc2d != 0;
  S.jlcs = jlcs;
  S.ne1 = c->prm.scalar ? 9 : 8;
  S.ne2 = c->prm.scalar ? 6 : 5;
  S.ny2d = S.on2d ? ny2d : 0;
  if (S.on1d) {
    PTC_TRY(alloc(&S.d_x, n1d_x(c)));
    PTC_TRY(alloc(&S.d_z, n1d_z(c)));
    PTC_TRY(fill_device(c, S.d_x, (long long)n1d_x(c), 0.0));
    PTC_TRY(fill_device(c, S.d_z, (long long)n1d_z(c), 0.0));
  }
  if (S.on2d) {
    S.iy2d.assign(iy2d, iy2d + ny2d);
    PTC_TRY(alloc(&S.d_iy2d, (size_t)ny2d));
    PTC_HIP(hipMemcpy(S.d_iy2d, S.iy2d.data(), sizeof(int) * ny2d, hipMemcpyHostToDevice));
    PTC_TRY(alloc(&S.d_2d, n2d(c)));
    PTC_TRY(fill_device(c, S.d_2d, (long long)n2d(c), 0.0));
  }
  S.ready = true;
  return PTC_OK;
}

// One sample (src/m_FlowCase.f90:430-805), in the order that keeps two
// spectra live: the cell-centre uy in fb throughout, each other field in fa.
int spectra_sample(ptc_ctx *c, const double *ux, const double *uy, const double *uz, const double *p,
                   const double *sc) {
  SpectraState &S = c->spec;
  PTC_CHECK(S.ready, PTC_E_STATE, "spectra: call ptc_init_spectra first");
  PTC_CHECK(c->pp.ready, PTC_E_STATE, "spectra: the forward transforms need InitPoissonSolver");
  if (!S.on1d && !S.on2d) {
    ++S.nspectime;
    return PTC_OK;
  }
  PTC_TRY(ensure_work(c));
  const int ny = c->ny;
  const double crf = (double)c->nx * (double)c->nz * c->prm.ucrf;  // ux + uCRF moves F(0, 0) only
  SpecPair P{};
  // cell-centre uy -> fb
  PTC_TRY(cell_uy(c, uy));
  PTC_TRY(forward(c, S.d_cell, S.d_fb));
  // ux -> fa: energy (1), linear coherence (7, 8), co-spectra with the
  // cell-centre uy of ux + uCRF (5) and of the cell-centre ux (6, H(kx) fa);
  // 2-D: ux (1) and the cell-centre co-spectrum (5, IsSymmetry = .false.: rCoe = -1)
  PTC_TRY(forward(c, ux, S.d_fa));
  if (S.on1d) {
    P = SpecPair{S.d_fa, S.d_fa, crf, crf, 0, 0, 0, 0, nullptr};
    PTC_TRY(spec_1d(c, P, true, 0, 1, 1));
    // :473-489: rows 1..nyc/2 against jForLCS, a channel's upper rows against nyc+1-jForLCS
    const bool ch = c->prm.flow_type == 1;
    P = SpecPair{S.d_fa, S.d_fa, crf, crf, 1, S.jlcs - 1, ch ? ny - S.jlcs : S.jlcs - 1, ch ? ny / 2 : ny, nullptr};
    PTC_TRY(spec_1d(c, P, false, 1, 7, 8));
    P = SpecPair{S.d_fa, S.d_fb, crf, 0.0, 0, 0, 0, 0, nullptr};
    PTC_TRY(spec_1d(c, P, false, 0, 5, 5));
    P = SpecPair{S.d_fa, S.d_fb, crf, 0.0, 0, 0, 0, 0, S.d_hx};
    PTC_TRY(spec_1d(c, P, false, 0, 6, 6));
  }
  if (S.on2d) {
    PTC_TRY(spec_2d(c, S.d_fa, S.d_fa, 1, 1, 1.0));
    PTC_TRY(spec_2d(c, S.d_fa, S.d_fb, 5, 1, -1.0, S.d_hx, crf));
  }
  // uy (2-D: the face mirror), uz, p, scalar
  const double *rest[4] = {uy, uz, p, sc};
  const int m1[4] = {2, 3, 4, 9}, m2[4] = {2, 3, 4, 6};
  for (int q = 0; q < 4; ++q) {
    if (q == 3 && !c->prm.scalar) break;
    PTC_TRY(forward(c, rest[q], S.d_fa));
    if (S.on1d) {
      P = SpecPair{S.d_fa, S.d_fa, 0.0, 0.0, 0, 0, 0, 0, nullptr};
      PTC_TRY(spec_1d(c, P, true, 0, m1[q], m1[q]));
    }
    if (S.on2d) PTC_TRY(spec_2d(c, S.d_fa, S.d_fa, m2[q], q == 0 ? 2 : 1, 1.0));
  }
  ++S.nspectime;
  return PTC_OK;
}

// The accumulated sums (not yet divided by nSpectime); reset zeroes them as
// the reference does after writing (src/m_FlowCase.f90:943-947).
int spectra_read(ptc_ctx *c, double *sx, double *sz, double *s2, int *nspectime, int reset) {
  SpectraState &S = c->spec;
  PTC_CHECK(S.ready, PTC_E_STATE, "spectra: call ptc_init_spectra first");
  if (sx && S.on1d) PTC_HIP(hipMemcpyAsync(sx, S.d_x, sizeof(double) * n1d_x(c), hipMemcpyDeviceToHost, c->stream));
  if (sz && S.on1d) PTC_HIP(hipMemcpyAsync(sz, S.d_z, sizeof(double) * n1d_z(c), hipMemcpyDeviceToHost, c->stream));
  if (s2 && S.on2d) PTC_HIP(hipMemcpyAsync(s2, S.d_2d, sizeof(double) * n2d(c), hipMemcpyDeviceToHost, c->stream));
  PTC_HIP(hipStreamSynchronize(c->stream));
  if (nspectime) *nspectime = S.nspectime;
  if (reset) {
    if (S.on1d) {
      PTC_TRY(fill_device(c, S.d_x, (long long)n1d_x(c), 0.0));
      PTC_TRY(fill_device(c, S.d_z, (long long)n1d_z(c), 0.0));
    }
    if (S.on2d) PTC_TRY(fill_device(c, S.d_2d, (long long)n2d(c), 0.0));
    S.nspectime = 0;
  }
  return PTC_OK;
}

}  // namespace ptc
