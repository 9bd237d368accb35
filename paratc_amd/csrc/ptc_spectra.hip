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
// F(0, 0) by nx nz uCRF, which the partial sums add.
#include "ptc_internal.hpp"

#include <algorithm>

namespace ptc {

namespace {

constexpr int kSpecThreads = 256;  // partial-sum workgroup
constexpr int kSpecRows = 16;      // kz rows per partial-sum workgroup

// ---- cell-centre fields: ux (+uCRF) by the 4-point interpolation (:433-444),
// uy by the two-face mean (:576-587), written into rows 1..ny of a field
__global__ void __launch_bounds__(64) k_spec_cell(Geo g, int mode, const double *__restrict__ u, double ucrf,
                                                  double *__restrict__ out) {
  const int i = blockIdx.x * 64 + threadIdx.x, j = 1 + blockIdx.y, k = blockIdx.z;
  if (i >= g.nx) return;
  const long long row = (long long)g.nx * j + g.plane * k;
  if (mode == 0) {
    const int nx = g.nx;
    const int im = i > 0 ? i - 1 : nx - 1, ip = i + 1 < nx ? i + 1 : i + 1 - nx, iu = i + 2 < nx ? i + 2 : i + 2 - nx;
    out[row + i] = g.ic1 * u[row + im] + g.ic2 * u[row + i] + g.ic3 * u[row + ip] + g.ic4 * u[row + iu] + ucrf;
  } else {
    out[row + i] = 0.5 * (u[row + i] + u[row + g.nx + i]);
  }
}

// One pair of spectra to multiply, P = Fa conj Fb.  b's row is the same row,
// or (lcs) the reference row: jr1 below row jsplit, jr2 from it on.
struct SpecPair {
  const double2 *a, *b;
  double sa, sb;  // added to Re F(0, 0) of a and b
  int lcs, jr1, jr2, jsplit;
};

// Partial sums of one row j and kSpecRows kz rows (grid: x = kz chunk, y = j):
//   px[(j * nch + chunk) * (h+1) + kx] = sum over the chunk's kz of P(kx, kz)
//   pz[2 * (j * nz + kz)]     = sum over kx = 0..h of P(kx, kz)
//   pz[2 * (j * nz + kz) + 1] = sum over kx = 1..h-1 of P(kx, kz)
// Thread t owns kx = t + 256 s, s < NS.
template <int NS, bool SAME>
__global__ void __launch_bounds__(kSpecThreads) k_spec_part(SpecPair P, int ny, int nz, int h, int pitch,
                                                            double2 *__restrict__ px, double2 *__restrict__ pz) {
  constexpr int NW = kSpecThreads / 64;
  __shared__ double red[kSpecRows][NW][4];
  const int j = blockIdx.y, ch = blockIdx.x, nch = gridDim.x;
  const int t = threadIdx.x, lane = t & 63, w = t >> 6;
  const int k0 = ch * kSpecRows, k1 = min(nz, k0 + kSpecRows);
  const int jb = P.lcs ? (j < P.jsplit ? P.jr1 : P.jr2) : j;
  double sxr[NS], sxi[NS];
#pragma unroll
  for (int s = 0; s < NS; ++s) sxr[s] = sxi[s] = 0.0;
  for (int kz = k0; kz < k1; ++kz) {
    const double2 *ra = P.a + ((long long)kz * ny + j) * pitch;
    const double2 *rb = P.b + ((long long)kz * ny + jb) * pitch;
    double2 va[NS], vb[NS];
#pragma unroll
    for (int s = 0; s < NS; ++s) {
      const int kx = t + kSpecThreads * s;
      va[s] = kx <= h ? ra[kx] : make_double2(0.0, 0.0);
      if (!SAME) vb[s] = kx <= h ? rb[kx] : make_double2(0.0, 0.0);
    }
    double rr = 0.0, ri = 0.0, ir = 0.0, ii = 0.0;
#pragma unroll
    for (int s = 0; s < NS; ++s) {
      const int kx = t + kSpecThreads * s;
      double2 a = va[s];
      if (kx == 0 && kz == 0) a.x += P.sa;
      double2 b = a;
      if (!SAME) {
        b = vb[s];
        if (kx == 0 && kz == 0) b.x += P.sb;
      }
      const double pr = a.x * b.x + a.y * b.y, pi = a.y * b.x - a.x * b.y;
      sxr[s] += pr;
      sxi[s] += pi;
      rr += pr;
      ri += pi;
      if (kx > 0 && kx < h) {
        ir += pr;
        ii += pi;
      }
    }
#pragma unroll
    for (int off = 32; off > 0; off >>= 1) {
      rr += __shfl_xor(rr, off);
      ri += __shfl_xor(ri, off);
      ir += __shfl_xor(ir, off);
      ii += __shfl_xor(ii, off);
    }
    if (lane == 0) {
      red[kz - k0][w][0] = rr;
      red[kz - k0][w][1] = ri;
      red[kz - k0][w][2] = ir;
      red[kz - k0][w][3] = ii;
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
    for (int q = 0; q < NW; ++q)
#pragma unroll
      for (int e = 0; e < 4; ++e) v[e] += red[t][q][e];
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
__device__ inline double hc2d_value(const double2 *__restrict__ F, int ny, int nx, int nz, int pitch, int row, int p,
                                    int q) {
  const int h = nx / 2, nzh = nz / 2;
  const int kx = p <= h ? p : nx - p, kz = q <= nzh ? q : nz - q, km = kz == 0 ? 0 : nz - kz;
  const double2 f1 = F[((long long)kz * ny + row) * pitch + kx];
  const double2 f2 = F[((long long)km * ny + row) * pitch + kx];
  if (p <= h) return q <= nzh ? 0.5 * (f1.x + f2.x) : 0.5 * (f1.y - f2.y);
  return q <= nzh ? 0.5 * (f1.y + f2.y) : 0.5 * (f2.x - f1.x);
}

// Adds one sample to the 2-D accumulator of quantity m (grid: x = p / 64,
// y = chosen row, z = q).  mirror: 0 none (half channel), 1 the row ny+1-jct,
// 2 the face row ny+2-jct (uy); rcoe weighs the mirror's product.
__global__ void __launch_bounds__(64) k_spec_2d(const double2 *__restrict__ a, const double2 *__restrict__ b, int ny,
                                                int nx, int nz, int pitch, const int *__restrict__ iy, int ny2d,
                                                int mirror, double rcoe, double norm, double *__restrict__ out) {
  const int p = blockIdx.x * 64 + threadIdx.x, jc = blockIdx.y, q = blockIdx.z;
  if (p >= nx) return;
  const int jct = iy[jc];
  double v = hc2d_value(a, ny, nx, nz, pitch, jct - 1, p, q) * hc2d_value(b, ny, nx, nz, pitch, jct - 1, p, q);
  if (mirror) {
    const int jcs = (mirror == 2 ? ny + 2 : ny + 1) - jct;
    v += rcoe * (hc2d_value(a, ny, nx, nz, pitch, jcs - 1, p, q) * hc2d_value(b, ny, nx, nz, pitch, jcs - 1, p, q));
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

int spec_2d(ptc_ctx *c, const double2 *a, const double2 *b, int m, int mirror, double rcoe) {
  SpectraState &S = c->spec;
  const int nx = c->nx, nz = c->nz;
  // normEgy of clcEnergySpectra2D / clcCosSpectra2D (:1253-1255)
  double norm = 1.0 / ((double)nx * (double)nz);
  norm = norm * norm;
  if (c->prm.flow_type == 1) norm = norm * 0.5;
  double *out = S.d_2d + (long long)nx * S.ny2d * nz * (m - 1);
  const dim3 grid((unsigned)((nx + 63) / 64), (unsigned)S.ny2d, (unsigned)nz);
  hipLaunchKernelGGL(k_spec_2d, grid, dim3(64), 0, c->stream, a, b, c->ny, nx, nz, S.pitch, S.d_iy2d, S.ny2d,
                     c->prm.flow_type == 1 ? mirror : 0, rcoe, norm, out);
  PTC_HIP(hipGetLastError());
  return PTC_OK;
}

int cell_field(ptc_ctx *c, int mode, const double *u) {
  const dim3 grid((unsigned)((c->nx + 63) / 64), (unsigned)c->ny, (unsigned)c->nz);
  hipLaunchKernelGGL(k_spec_cell, grid, dim3(64), 0, c->stream, c->geo, mode, u, c->prm.ucrf, c->spec.d_cell);
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
  S.pitch = c->pp.pitch;
  S.nch = (c->nz + kSpecRows - 1) / kSpecRows;
  const size_t nspec = (size_t)S.pitch * c->ny * c->nz;
  PTC_TRY(alloc(&S.d_fa, nspec));
  PTC_TRY(alloc(&S.d_fb, nspec));
  PTC_TRY(alloc(&S.d_cell, (size_t)c->nx * (c->ny + 2) * c->nz));
  PTC_TRY(alloc(&S.d_px, (size_t)c->ny * S.nch * (c->nx / 2 + 1)));
  PTC_TRY(alloc(&S.d_pz, (size_t)2 * c->ny * c->nz));
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
  PTC_TRY(cell_field(c, 1, uy));
  PTC_TRY(forward(c, S.d_cell, S.d_fb));
  // ux -> fa: energy (1), linear coherence (7, 8), co-spectrum with the cell-centre uy (5); 2-D ux (1)
  PTC_TRY(forward(c, ux, S.d_fa));
  if (S.on1d) {
    P = SpecPair{S.d_fa, S.d_fa, crf, crf, 0, 0, 0, 0};
    PTC_TRY(spec_1d(c, P, true, 0, 1, 1));
    // :473-489: rows 1..nyc/2 against jForLCS, a channel's upper rows against nyc+1-jForLCS
    const bool ch = c->prm.flow_type == 1;
    P = SpecPair{S.d_fa, S.d_fa, crf, crf, 1, S.jlcs - 1, ch ? ny - S.jlcs : S.jlcs - 1, ch ? ny / 2 : ny};
    PTC_TRY(spec_1d(c, P, false, 1, 7, 8));
    P = SpecPair{S.d_fa, S.d_fb, crf, 0.0, 0, 0, 0, 0};
    PTC_TRY(spec_1d(c, P, false, 0, 5, 5));
  }
  if (S.on2d) PTC_TRY(spec_2d(c, S.d_fa, S.d_fa, 1, 1, 1.0));
  // cell-centre ux (+uCRF) -> fa: co-spectrum 6; 2-D co-spectrum 5 (IsSymmetry = .false., rCoe = -1)
  PTC_TRY(cell_field(c, 0, ux));
  PTC_TRY(forward(c, S.d_cell, S.d_fa));
  if (S.on1d) {
    P = SpecPair{S.d_fa, S.d_fb, 0.0, 0.0, 0, 0, 0, 0};
    PTC_TRY(spec_1d(c, P, false, 0, 6, 6));
  }
  if (S.on2d) PTC_TRY(spec_2d(c, S.d_fa, S.d_fb, 5, 1, -1.0));
  // uy (2-D: the face mirror), uz, p, scalar
  const double *rest[4] = {uy, uz, p, sc};
  const int m1[4] = {2, 3, 4, 9}, m2[4] = {2, 3, 4, 6};
  for (int q = 0; q < 4; ++q) {
    if (q == 3 && !c->prm.scalar) break;
    PTC_TRY(forward(c, rest[q], S.d_fa));
    if (S.on1d) {
      P = SpecPair{S.d_fa, S.d_fa, 0.0, 0.0, 0, 0, 0, 0};
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
