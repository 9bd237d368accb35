// ptc_seed.hip -- seeded synthetic inputs generated in HBM.
//
// The reference seeds its initial perturbation from the wall clock
// (InitVelocity, src/m_FlowCase.f90:49-107), so a run cannot be reproduced.
// The seeded inputs here are a pure function of (seed, stream, index):
//
//   x_q = mix(key + (q + 1) * 0x9E3779B97F4A7C15),  key = seed + stream * 2^32
//
// with mix the SplitMix64 finaliser (Steele, Lea and Flood, 2014).  Uniform
// doubles take the top 53 bits; a normal value q uses Box-Muller on the pair
// of uniforms (2 floor(q/2), 2 floor(q/2) + 1): r cos t for even q, r sin t
// for odd q.  q counts the interior cells in the reference's column-major
// order, q = i + nx (j - 1 + ny k), so a field made here equals the one the
// host-side generator (paratc_amd/rng.py) makes for the same (seed, stream):
// bit for bit for uniforms, to the last bits of the device log / sqrt / cos /
// sin / tanh for normals and the tanh front.
//
// Generating on the device keeps multi-GB inputs (9216 x 140 x 1400 is 14.5
// GB per field) off the host and off PCIe.
#include <algorithm>
#include <cmath>

#include "ptc_internal.hpp"

namespace ptc {

constexpr int kSlotSeedMean = 40;  // device scalar: the dyp-weighted mean of a seeded source

__device__ __forceinline__ unsigned long long splitmix(unsigned long long key, unsigned long long q) {
  unsigned long long z = key + (q + 1ull) * 0x9E3779B97F4A7C15ull;
  z = (z ^ (z >> 30)) * 0xBF58476D1CE4E5B9ull;
  z = (z ^ (z >> 27)) * 0x94D049BB133111EBull;
  return z ^ (z >> 31);
}

__device__ __forceinline__ double uniform01(unsigned long long key, unsigned long long q) {
  return (double)(splitmix(key, q) >> 11) * (1.0 / 9007199254740992.0);
}

__device__ __forceinline__ double normal01(unsigned long long key, unsigned long long q) {
  const unsigned long long p = q >> 1;
  const double u1 = 1.0 - uniform01(key, 2ull * p);  // (0, 1]
  const double u2 = uniform01(key, 2ull * p + 1ull);
  const double r = sqrt(-2.0 * log(u1));
  const double t = 6.283185307179586476925286766559005768 * u2;
  return (q & 1ull) ? r * sin(t) : r * cos(t);
}

static unsigned long long seed_key(int seed, int stream) {
  return (unsigned long long)(long long)seed + ((unsigned long long)(unsigned)stream << 32);
}

// One workgroup per (y row, z plane): x fastest, interior rows j = 1..ny.
// f = N(0,1); part[row] = the row's sum weighted by dyp(j) (for the
// compatibility mean).
__global__ void __launch_bounds__(256) k_seed_normal(double *__restrict__ f, Geo g, unsigned long long key,
                                                     double *__restrict__ part) {
  const int j = blockIdx.x + 1, k = blockIdx.y;
  const long long base = (long long)g.plane * k + (long long)g.nx * j;
  const unsigned long long q0 = (unsigned long long)g.nx * ((unsigned long long)(j - 1) + (unsigned long long)g.ny * k);
  double s = 0.0;
  for (int i = threadIdx.x; i < g.nx; i += blockDim.x) {
    const double v = normal01(key, q0 + (unsigned long long)i);
    f[base + i] = v;
    s += v;
  }
  __shared__ double sh[256];
  sh[threadIdx.x] = s;
  __syncthreads();
  for (int w = 128; w > 0; w >>= 1) {
    if ((int)threadIdx.x < w) sh[threadIdx.x] += sh[threadIdx.x + w];
    __syncthreads();
  }
  if (threadIdx.x == 0) part[blockIdx.x + gridDim.x * blockIdx.y] = sh[0] * g.met[M_DYP * g.nyh + j];
}

// f -= mean over the interior rows
__global__ void __launch_bounds__(256) k_sub_mean(double *__restrict__ f, Geo g, const double *__restrict__ mean,
                                                  double scale) {
  const int j = blockIdx.x + 1, k = blockIdx.y;
  const long long base = (long long)g.plane * k + (long long)g.nx * j;
  const double m = mean[0] * scale;
  for (int i = threadIdx.x; i < g.nx; i += blockDim.x) f[base + i] -= m;
}

__global__ void __launch_bounds__(kReduceThreads) k_seed_sum(const double *__restrict__ part, int n,
                                                            double *__restrict__ out) {
  const double t = reduce_partials<false>(part, n);
  if (threadIdx.x == 0) out[0] = t;
}

// Initial fields: velocities of uniform noise amp (2 U - 1) on the interior
// (streams 0, 1, 2 for u, v, w), v = 0 on the bottom wall face (row 1);
// kind 0 adds the laminar profile 1.5 (1 - ((yc - h) / h)^2) ubulk to u,
// kind 1 sets the scalar to the front 0.5 (1 - tanh((xc - x0) / delta)) plus
// noise (stream 3).  Ghost rows are zero (the boundary fills set them).
__global__ void __launch_bounds__(256) k_seed_fields(Geo g, int kind, unsigned long long key_base, double amp,
                                                     double ubulk, double half, double x0, double delta, double dx,
                                                     double *__restrict__ ux, double *__restrict__ uy,
                                                     double *__restrict__ uz, double *__restrict__ sc) {
  // no fused multiply-adds: the profile and the noise round as the host
  // generator's separate products and sums do, so the fields match it bit for bit
#pragma clang fp contract(off)
  const int j = blockIdx.x, k = blockIdx.y;  // j = 0..ny+1
  const long long base = (long long)g.plane * k + (long long)g.nx * j;
  const bool interior = j >= 1 && j <= g.ny;
  const unsigned long long q0 =
      interior ? (unsigned long long)g.nx * ((unsigned long long)(j - 1) + (unsigned long long)g.ny * k) : 0ull;
  const unsigned long long s1 = 1ull << 32;
  double prof = 0.0;
  if (interior && kind == 0) {
    const double y = (g.met[M_YC * g.nyh + j] - half) / half;
    prof = 1.5 * (1.0 - y * y) * ubulk;
  }
  for (int i = threadIdx.x; i < g.nx; i += blockDim.x) {
    double u = 0.0, v = 0.0, w = 0.0, c = 0.0;
    if (interior) {
      const unsigned long long q = q0 + (unsigned long long)i;
      u = prof + amp * (2.0 * uniform01(key_base, q) - 1.0);
      v = (j == 1) ? 0.0 : amp * (2.0 * uniform01(key_base + s1, q) - 1.0);
      w = amp * (2.0 * uniform01(key_base + 2ull * s1, q) - 1.0);
      if (kind == 1) {
        const double xc = ((double)i + 0.5) * dx;
        c = 0.5 * (1.0 - tanh((xc - x0) / delta)) + amp * (2.0 * uniform01(key_base + 3ull * s1, q) - 1.0);
      }
    }
    ux[base + i] = u;
    uy[base + i] = v;
    uz[base + i] = w;
    if (sc) sc[base + i] = c;
  }
}

}  // namespace ptc

using namespace ptc;

extern "C" {

int ptc_seed_normal(ptc_ctx *ctx, int fid, int seed, int stream, int compatible) {
  PTC_CHECK(ctx != nullptr, PTC_E_ARG, "null context");
  PTC_CHECK(fid != PTC_F_UY_YM && fid != PTC_F_DUY_YM, PTC_E_ARG, "seeded sources fill 3-D fields");
  double *f = nullptr;
  PTC_TRY(field(ctx, fid, &f));
  const int nrow = ctx->ny * ctx->nz;
  PTC_TRY(ensure_red(ctx, nrow));
  const dim3 grid((unsigned)ctx->ny, (unsigned)ctx->nz);
  hipLaunchKernelGGL(k_seed_normal, grid, dim3(256), 0, ctx->stream, f, ctx->geo, seed_key(seed, stream), ctx->d_red);
  PTC_HIP(hipGetLastError());
  if (compatible) {
    // subtract sum_ijk f dyp_j / (sum_j dyp_j nx nz): Neumann compatibility
    double sdyp = 0.0;
    for (int j = 1; j <= ctx->ny; ++j) sdyp += ctx->h_met[(size_t)M_DYP * (ctx->ny + 2) + j];
    hipLaunchKernelGGL(k_seed_sum, dim3(1), dim3(kReduceThreads), 0, ctx->stream, ctx->d_red, nrow,
                       ctx->d_scal + kSlotSeedMean);
    PTC_HIP(hipGetLastError());
    hipLaunchKernelGGL(k_sub_mean, grid, dim3(256), 0, ctx->stream, f, ctx->geo, ctx->d_scal + kSlotSeedMean,
                       1.0 / (sdyp * (double)ctx->nx * (double)ctx->nz));
    PTC_HIP(hipGetLastError());
  }
  return PTC_OK;
}

int ptc_seed_fields(ptc_ctx *ctx, int kind, int seed, double amp, double x0, double delta) {
  PTC_CHECK(ctx != nullptr, PTC_E_ARG, "null context");
  PTC_CHECK(kind == PTC_SEED_CHANNEL || kind == PTC_SEED_FRONT, PTC_E_ARG, "unknown seeded initial field");
  PTC_CHECK(kind != PTC_SEED_FRONT || ctx->prm.scalar, PTC_E_ARG, "the front initial field needs the scalar");
  PTC_CHECK(kind != PTC_SEED_FRONT || delta > 0.0, PTC_E_ARG, "front width must be positive");
  double *ux, *uy, *uz, *sc = nullptr;
  PTC_TRY(field(ctx, PTC_F_UX, &ux));
  PTC_TRY(field(ctx, PTC_F_UY, &uy));
  PTC_TRY(field(ctx, PTC_F_UZ, &uz));
  if (ctx->prm.scalar) PTC_TRY(field(ctx, PTC_F_SCALAR, &sc));
  const dim3 grid((unsigned)(ctx->ny + 2), (unsigned)ctx->nz);
  hipLaunchKernelGGL(k_seed_fields, grid, dim3(256), 0, ctx->stream, ctx->geo, kind, seed_key(seed, 0), amp,
                     ctx->prm.ubulk, 0.5 * ctx->yly, x0, delta, ctx->xlx / ctx->nx, ux, uy, uz,
                     kind == PTC_SEED_FRONT ? sc : nullptr);
  PTC_HIP(hipGetLastError());
  return PTC_OK;
}

}  // extern "C"
