// ptc_context.hip -- context set-up, device fields and host transfers.
//
// ptc_create does on one device what InitMeshAndMetries
// (src/m_MeshAndMetries.f90:86-335), the RK tables of ReadAndInitParameters
// (src/m_Parameters.f90:159-174), the modified wavenumbers of
// InitPoissonSolver (src/m_Poisson.f90:37-48) and AllocateVariables
// (src/m_Variables.f90:41-105) do on the host: all y tables and constants are
// computed once in FP64 and uploaded; fields are allocated on first use and
// stay in HBM.
#include <algorithm>
#include <cmath>
#include <cstring>
#include <string>
#include <vector>

#include "ptc_internal.hpp"

namespace ptc {

static thread_local std::string g_err;
void set_error(const std::string &msg) { g_err = msg; }
const char *last_error() { return g_err.c_str(); }

__global__ void k_fill(double *__restrict__ a, long long n, double v) {
  for (long long q = (long long)blockIdx.x * blockDim.x + threadIdx.x; q < n; q += (long long)gridDim.x * blockDim.x)
    a[q] = v;
}

int fill_device(ptc_ctx *c, double *a, long long n, double v) {
  const int nblk = (int)std::min<long long>((n + 255) / 256, (long long)c->n_cu * 16);
  hipLaunchKernelGGL(k_fill, dim3(std::max(1, nblk)), dim3(256), 0, c->stream, a, n, v);
  PTC_HIP(hipGetLastError());
  return PTC_OK;
}

// plain copy (fields that fit the last-level cache)
__global__ void __launch_bounds__(256) k_copy(double2 *__restrict__ d, const double2 *__restrict__ s, long long n2) {
  for (long long q = (long long)blockIdx.x * blockDim.x + threadIdx.x; q < n2; q += (long long)gridDim.x * blockDim.x)
    d[q] = s[q];
}

// streaming copy of fields larger than the last-level cache: nontemporal
// 16-byte loads and stores, four per thread in flight per iteration
typedef double nt_d2 __attribute__((ext_vector_type(2)));
__global__ void __launch_bounds__(256) k_copy_nt(double2 *__restrict__ d, const double2 *__restrict__ s, long long n2) {
  const nt_d2 *sv = reinterpret_cast<const nt_d2 *>(s);
  nt_d2 *dv = reinterpret_cast<nt_d2 *>(d);
  const long long st = (long long)gridDim.x * blockDim.x;
  long long q = (long long)blockIdx.x * blockDim.x + threadIdx.x;
  for (; q + 3 * st < n2; q += 4 * st) {
    const nt_d2 a = __builtin_nontemporal_load(sv + q), b = __builtin_nontemporal_load(sv + q + st);
    const nt_d2 e = __builtin_nontemporal_load(sv + q + 2 * st), f = __builtin_nontemporal_load(sv + q + 3 * st);
    __builtin_nontemporal_store(a, dv + q);
    __builtin_nontemporal_store(b, dv + q + st);
    __builtin_nontemporal_store(e, dv + q + 2 * st);
    __builtin_nontemporal_store(f, dv + q + 3 * st);
  }
  for (; q < n2; q += st) __builtin_nontemporal_store(__builtin_nontemporal_load(sv + q), dv + q);
}
__global__ void k_copy_tail(double *__restrict__ d, const double *__restrict__ s, long long n) { d[n - 1] = s[n - 1]; }

constexpr long long kStreamCopyDoubles = 32ll << 20;  // 256 MB

int copy_device(ptc_ctx *c, double *dst, const double *src, long long n) {
  const long long n2 = n / 2;
  const int nblk = (int)std::min<long long>((n2 + 255) / 256, (long long)c->n_cu * 32);
  if (n2 > 0) {
    if (n >= kStreamCopyDoubles)
      hipLaunchKernelGGL(k_copy_nt, dim3(std::max(1, nblk)), dim3(256), 0, c->stream, reinterpret_cast<double2 *>(dst),
                         reinterpret_cast<const double2 *>(src), n2);
    else
      hipLaunchKernelGGL(k_copy, dim3(std::max(1, nblk)), dim3(256), 0, c->stream, reinterpret_cast<double2 *>(dst),
                         reinterpret_cast<const double2 *>(src), n2);
    PTC_HIP(hipGetLastError());
  }
  if (n % 2) {
    hipLaunchKernelGGL(k_copy_tail, dim3(1), dim3(1), 0, c->stream, dst, src, n);
    PTC_HIP(hipGetLastError());
  }
  return PTC_OK;
}

static bool is_plane(int id) { return id == PTC_F_UY_YM || id == PTC_F_DUY_YM; }

size_t field_count(const ptc_ctx *c, int id) { return is_plane(id) ? (size_t)c->nx * c->nz : c->n3; }

int check_field(ptc_ctx *c, int id) {
  PTC_CHECK(c != nullptr, PTC_E_ARG, "null context");
  PTC_CHECK(id >= 0 && id < PTC_NFIELDS, PTC_E_ARG, "field id out of range: " + std::to_string(id));
  return PTC_OK;
}

int field_existing(ptc_ctx *c, int id, double **out) {
  PTC_TRY(check_field(c, id));
  PTC_CHECK(c->d_field[id] != nullptr, PTC_E_STATE,
            "field " + std::to_string(id) + " has no device memory yet (upload or fill it first)");
  *out = c->d_field[id];
  return PTC_OK;
}

int field(ptc_ctx *c, int id, double **out) {
  PTC_TRY(check_field(c, id));
  if (!c->d_field[id]) {
    const size_t n = field_count(c, id);
    double *p = nullptr;
    if (hipMalloc(&p, sizeof(double) * n) != hipSuccess) {
      (void)hipGetLastError();
      set_error("device allocation of field " + std::to_string(id) + " failed (" +
                std::to_string(sizeof(double) * n) + " bytes)");
      return PTC_E_ALLOC;
    }
    c->d_field[id] = p;
    PTC_TRY(fill_device(c, p, (long long)n, 0.0));
  }
  *out = c->d_field[id];
  return PTC_OK;
}

// ---- y metrics: src/m_MeshAndMetries.f90:194-333 --------------------------
static void compute_metrics(ptc_ctx *c, const double *yp_in) {
  const int nyc = c->ny, nyp = c->ny + 1, W = c->ny + 2;
  std::vector<double> &T = c->h_met;
  T.assign((size_t)M_NMET * W, 0.0);
  auto V = [&](int row) { return T.data() + (size_t)row * W; };
  std::vector<double> yp((size_t)W, 0.0);
  for (int j = 1; j <= nyp; ++j) yp[(size_t)j] = yp_in[j - 1];
  double *yc = V(M_YC), *dyp = V(M_DYP), *dyc = V(M_DYC), *rdyp = V(M_RDYP), *rdyc = V(M_RDYC);
  for (int j = 1; j <= nyc; ++j) yc[j] = 0.5 * (yp[j] + yp[j + 1]);
  yc[0] = 2.0 * yp[1] - yc[1];
  yc[nyp] = 2.0 * yp[nyp] - yc[nyc];
  for (int j = 1; j <= nyc; ++j) dyp[j] = yp[j + 1] - yp[j];
  dyp[0] = dyp[1];
  dyp[nyp] = dyp[nyc];
  for (int j = 2; j <= nyc; ++j) dyc[j] = yc[j] - yc[j - 1];
  dyc[1] = dyp[1];
  dyc[nyp] = dyp[nyc];
  for (int j = 0; j <= nyp; ++j) rdyp[j] = 1.0 / dyp[j];
  for (int j = 1; j <= nyp; ++j) rdyc[j] = 1.0 / dyc[j];
  double *yint = V(M_YINT);
  for (int j = 1; j <= nyp; ++j) yint[j] = dyp[j] / (dyp[j] + dyp[j - 1]);
  // pressure Laplacian (Neumann rows 1 and ny)
  double *am = V(M_AM2PH), *ac = V(M_AC2PH), *ap = V(M_AP2PH);
  for (int j = 1; j <= nyc; ++j) {
    am[j] = rdyp[j] * rdyc[j];
    ap[j] = rdyp[j] * rdyc[j + 1];
  }
  am[1] = 0.0;
  ap[1] = rdyp[1] * rdyc[2];
  am[nyc] = rdyp[nyc] * rdyc[nyc];
  ap[nyc] = 0.0;
  for (int j = 1; j <= nyc; ++j) ac[j] = -(am[j] + ap[j]);
  // u/w Laplacian with the one-sided wall rows
  double *um = V(M_AM2C), *uc = V(M_AC2C), *up = V(M_AP2C);
  for (int j = 1; j <= nyc; ++j) {
    um[j] = rdyp[j] * rdyc[j];
    up[j] = rdyp[j] * rdyc[j + 1];
  }
  um[1] = 4.0 * rdyc[1] / (dyc[1] + 2.0 * dyc[2]);
  up[1] = 4.0 * rdyc[2] / (dyc[1] + 2.0 * dyc[2]);
  if (c->prm.flow_type == 1) {
    um[nyc] = 4.0 * rdyc[nyc] / (dyc[nyp] + 2.0 * dyc[nyc]);
    up[nyc] = 4.0 * rdyc[nyp] / (dyc[nyp] + 2.0 * dyc[nyc]);
  } else {
    um[nyc] = rdyp[nyc] * rdyc[nyc];
    up[nyc] = 0.0;
  }
  for (int j = 1; j <= nyc; ++j) uc[j] = -(um[j] + up[j]);
  // Crank-Nicolson form: the ghost reflection u_0 = -u_1 folded into row 1 (and ny)
  double *cm = V(M_AM2CCN), *cc = V(M_AC2CCN), *cp = V(M_AP2CCN);
  for (int j = 1; j <= nyc; ++j) {
    cm[j] = um[j];
    cc[j] = uc[j];
    cp[j] = up[j];
  }
  cp[1] = up[1];
  cc[1] = uc[1] - um[1];
  cm[1] = 2.0 * um[1];
  if (c->prm.flow_type == 1) {
    cm[nyc] = um[nyc];
    cc[nyc] = uc[nyc] - up[nyc];
    cp[nyc] = 2.0 * up[nyc];
  }
  // v Laplacian (cell-centred in y)
  double *vm = V(M_AM2P), *vc = V(M_AC2P), *vp = V(M_AP2P);
  for (int j = 1; j <= nyc; ++j) {
    vm[j] = rdyc[j] * rdyp[j - 1];
    vp[j] = rdyc[j] * rdyp[j];
  }
  for (int j = 1; j <= nyc; ++j) vc[j] = -(vm[j] + vp[j]);
  // scalar Laplacians
  if (c->prm.scalar) {
    double *sm = V(M_AM2CSC), *scc = V(M_AC2CSC), *sp = V(M_AP2CSC);
    for (int j = 1; j <= nyc; ++j) {
      sm[j] = rdyp[j] * rdyc[j];
      sp[j] = rdyp[j] * rdyc[j + 1];
    }
    if (c->prm.sbc_opt[0] == -1) {
      sm[1] = 4.0 * rdyc[1] / (dyc[1] + 2.0 * dyc[2]);
      sp[1] = 4.0 * rdyc[2] / (dyc[1] + 2.0 * dyc[2]);
    }
    if (c->prm.sbc_opt[1] == -1) {
      sm[nyc] = 4.0 * rdyc[nyc] / (dyc[nyp] + 2.0 * dyc[nyc]);
      sp[nyc] = 4.0 * rdyc[nyp] / (dyc[nyp] + 2.0 * dyc[nyc]);
    }
    for (int j = 1; j <= nyc; ++j) scc[j] = -(sm[j] + sp[j]);
    double *tm = V(M_AM2CSCCN), *tc = V(M_AC2CSCCN), *tp = V(M_AP2CSCCN);
    for (int j = 1; j <= nyc; ++j) {
      tm[j] = rdyp[j] * rdyc[j];
      tp[j] = rdyp[j] * rdyc[j + 1];
    }
    for (int j = 1; j <= nyc; ++j) tc[j] = -(tm[j] + tp[j]);
    const double *sv = c->prm.sbc_val;
    switch (c->prm.sbc_opt[0]) {
      case -1:
        tm[1] = 8.0 * rdyc[1] / (dyc[1] + 2.0 * dyc[2]);
        tp[1] = 4.0 * rdyc[2] / (dyc[1] + 2.0 * dyc[2]);
        tc[1] = -(8.0 * rdyc[1] + 4.0 * rdyc[2]) / (dyc[1] + 2.0 * dyc[2]);
        break;
      case -2:
        tp[1] = rdyp[1] * rdyc[2];
        tc[1] = -rdyp[1] * rdyc[2];
        tm[1] = 0.0;
        break;
      case -3:
        tp[1] = rdyp[1] * rdyc[2];
        tc[1] = tc[1] + tm[1] * (2.0 + sv[0] * dyc[1]) / (2.0 - sv[0] * dyc[1]);
        tm[1] = 0.0;
        break;
      default: break;
    }
    switch (c->prm.sbc_opt[1]) {
      case -1:
        tm[nyc] = 4.0 * rdyc[nyc] / (dyc[nyp] + 2.0 * dyc[nyc]);
        tp[nyc] = 8.0 * rdyc[nyp] / (dyc[nyp] + 2.0 * dyc[nyc]);
        tc[nyc] = -(8.0 * rdyc[nyp] + 4.0 * rdyc[nyc]) / (dyc[nyp] + 2.0 * dyc[nyc]);
        break;
      case -2:
        tm[nyc] = rdyp[nyc] * rdyc[nyc];
        tc[nyc] = -rdyp[nyc] * rdyc[nyc];
        tp[nyc] = 0.0;
        break;
      case -3:
        tm[nyc] = rdyp[nyc] * rdyc[nyc];
        tc[nyc] = tc[nyc] + tp[nyc] * (2.0 - sv[1] * dyc[nyp]) / (2.0 + sv[1] * dyc[nyp]);
        tp[nyc] = 0.0;
        break;
      default: break;
    }
  }
}

static void compute_constants(ptc_ctx *c) {
  Geo &g = c->geo;
  g.nx = c->nx;
  g.ny = c->ny;
  g.nz = c->nz;
  g.plane = (long long)c->nx * (c->ny + 2);
  g.nyh = c->ny + 2;
  // src/m_MeshAndMetries.f90:86-101
  const double rdx = (double)c->nx / c->xlx, rdz = (double)c->nz / c->zlz;
  const double rdx2 = rdx * rdx, rdz2 = rdz * rdz;
  g.rdx = rdx;
  g.rdz = rdz;
  g.dxc1 = rdx / 24.0; g.dxc2 = -rdx * 1.125; g.dxc3 = rdx * 1.125; g.dxc4 = -rdx / 24.0;
  g.dzc1 = rdz / 24.0; g.dzc2 = -rdz * 1.125; g.dzc3 = rdz * 1.125; g.dzc4 = -rdz / 24.0;
  g.dxxc1 = -rdx2 / 12.0; g.dxxc2 = 4.0 / 3.0 * rdx2; g.dxxc3 = -2.5 * rdx2; g.dxxc4 = 4.0 / 3.0 * rdx2;
  g.dxxc5 = -rdx2 / 12.0;
  g.dzzc1 = -rdz2 / 12.0; g.dzzc2 = 4.0 / 3.0 * rdz2; g.dzzc3 = -2.5 * rdz2; g.dzzc4 = 4.0 / 3.0 * rdz2;
  g.dzzc5 = -rdz2 / 12.0;
  g.ic1 = -1.0 / 16.0; g.ic2 = 9.0 / 16.0; g.ic3 = 9.0 / 16.0; g.ic4 = -1.0 / 16.0;
  // modified wavenumbers (src/m_Poisson.f90:40-48)
  const double twopi = 6.283185307179586476925286766559005768;
  c->h_wnx.resize((size_t)c->nx);
  c->h_wnz.resize((size_t)c->nz);
  for (int i = 0; i < c->nx; ++i) {
    const double w = twopi * (double)i / (double)c->nx;
    c->h_wnx[(size_t)i] = rdx2 / 288.0 * (std::cos(3.0 * w) - 54.0 * std::cos(2.0 * w) + 783.0 * std::cos(w) - 730.0);
  }
  for (int k = 0; k < c->nz; ++k) {
    const double w = twopi * (double)k / (double)c->nz;
    c->h_wnz[(size_t)k] = rdz2 / 288.0 * (std::cos(3.0 * w) - 54.0 * std::cos(2.0 * w) + 783.0 * std::cos(w) - 730.0);
  }
  c->normfft = 1.0 / ((double)c->nx * (double)c->nz);
  // time-scheme tables (src/m_Parameters.f90:159-174)
  if (c->prm.ischeme == 1) {
    c->iadvance = 1;
    c->gconst[0] = 1.5; c->tconst[0] = -0.5;
  } else if (c->prm.ischeme == 2) {
    c->iadvance = 2;
    c->gconst[0] = 0.5; c->gconst[1] = 1.0;
    c->tconst[0] = 0.0; c->tconst[1] = -0.5;
  } else {
    c->iadvance = 3;
    c->gconst[0] = 8.0 / 15.0; c->gconst[1] = 5.0 / 12.0; c->gconst[2] = 3.0 / 4.0;
    c->tconst[0] = 0.0; c->tconst[1] = -17.0 / 60.0; c->tconst[2] = -5.0 / 12.0;
  }
  for (int s = 0; s < 3; ++s) c->aconst[s] = c->gconst[s] + c->tconst[s];
  c->dt = c->prm.dtmax;
}

static bool fft_smooth(int n) {
  for (int f : {2, 3, 5, 7})
    while (n % f == 0) n /= f;
  return n == 1;
}

int create(const ptc_grid *grid, const ptc_params *prm, const double *yp, ptc_ctx **out) {
  PTC_CHECK(grid && prm && yp && out, PTC_E_ARG, "null argument to ptc_create");
  *out = nullptr;
  PTC_CHECK(grid->nx >= 8 && grid->ny >= 4 && grid->nz >= 8, PTC_E_ARG, "grid too small (nx, nz >= 8, ny >= 4)");
  PTC_CHECK(grid->nx % 2 == 0 && grid->nz % 2 == 0, PTC_E_ARG, "mod(nxc,2)/=0 || mod(nzc,2)/=0 WRONG");
  PTC_CHECK(!(prm->flow_type == 1 && grid->ny % 2), PTC_E_ARG, "mod(nyc,2)/=0 For channel geomotry. WRONG");
  PTC_CHECK(prm->flow_type == 1 || prm->flow_type == 2, PTC_E_ARG, "FlowType must be 1 or 2");
  PTC_CHECK(prm->ischeme >= 1 && prm->ischeme <= 3, PTC_E_ARG, "Time scheme WRONG!!!");
  PTC_CHECK(fft_smooth(grid->nx / 2) && fft_smooth(grid->nz), PTC_E_UNSUPPORTED,
            "nx/2 and nz must factor into 2, 3, 5 and 7");
  if (prm->scalar)
    for (int s = 0; s < 2; ++s)
      PTC_CHECK(prm->sbc_opt[s] >= -3 && prm->sbc_opt[s] <= -1, PTC_E_ARG, "ScalarBcOption wrong !!!");
  for (int j = 0; j < grid->ny; ++j)
    PTC_CHECK(yp[j + 1] > yp[j], PTC_E_ARG, "yp must increase strictly");
  ptc_ctx *c = new ptc_ctx();
  c->nx = grid->nx;
  c->ny = grid->ny;
  c->nz = grid->nz;
  c->xlx = grid->xlx;
  c->yly = grid->yly;
  c->zlz = grid->zlz;
  c->prm = *prm;
  c->plane = (size_t)c->nx * (c->ny + 2);
  c->n3 = c->plane * c->nz;
  compute_metrics(c, yp);
  compute_constants(c);
  auto fail = [&](int code) {
    ptc_destroy(c);
    return code;
  };
  if (hipGetDevice(&c->device) != hipSuccess) {
    set_error("no HIP device");
    return fail(PTC_E_HIP);
  }
  if (noalloc_init_device(c->device) != PTC_OK) return fail(PTC_E_HIP);
  hipDeviceProp_t prop;
  if (hipGetDeviceProperties(&prop, c->device) == hipSuccess) {
    c->lds_limit = prop.sharedMemPerBlock;
    c->n_cu = prop.multiProcessorCount;
  }
  if (hipStreamCreate(&c->stream) != hipSuccess) {
    set_error("hipStreamCreate failed");
    return fail(PTC_E_HIP);
  }
  const size_t W = (size_t)c->ny + 2;
  if (hipMalloc(&c->d_met, sizeof(double) * M_NMET * W) != hipSuccess ||
      hipMalloc(&c->d_wnx, sizeof(double) * c->nx) != hipSuccess ||
      hipMalloc(&c->d_wnz, sizeof(double) * c->nz) != hipSuccess ||
      hipMalloc(&c->d_scal, sizeof(double) * 64) != hipSuccess ||
      hipMalloc(&c->d_cnfac, sizeof(double) * (kCnMaxJobs + kCnStepTables) * kCnFacStride) != hipSuccess ||
      hipHostMalloc(&c->h_pinned, sizeof(double) * 64) != hipSuccess) {
    set_error("device allocation failed in ptc_create");
    return fail(PTC_E_ALLOC);
  }
  if (hipMemcpy(c->d_met, c->h_met.data(), sizeof(double) * M_NMET * W, hipMemcpyHostToDevice) != hipSuccess ||
      hipMemcpy(c->d_wnx, c->h_wnx.data(), sizeof(double) * c->nx, hipMemcpyHostToDevice) != hipSuccess ||
      hipMemcpy(c->d_wnz, c->h_wnz.data(), sizeof(double) * c->nz, hipMemcpyHostToDevice) != hipSuccess) {
    set_error("upload of metric tables failed");
    return fail(PTC_E_HIP);
  }
  c->geo.met = c->d_met;
  if (fill_device(c, c->d_scal, 64, 0.0) != PTC_OK) return fail(PTC_E_HIP);
  *out = c;
  return PTC_OK;
}

void destroy(ptc_ctx *c) {
  if (!c) return;
  if (c->stream) (void)hipStreamSynchronize(c->stream);
  poisson_free_plan(c);
  spectra_free(c);
  for (auto &p : c->d_field)
    if (p) (void)hipFree(p);
  if (c->d_met) (void)hipFree(c->d_met);
  if (c->d_wnx) (void)hipFree(c->d_wnx);
  if (c->d_wnz) (void)hipFree(c->d_wnz);
  if (c->d_red) (void)hipFree(c->d_red);
  if (c->d_stat) (void)hipFree(c->d_stat);
  if (c->d_visu) (void)hipFree(c->d_visu);
  if (c->d_scal) (void)hipFree(c->d_scal);
  if (c->d_cnfac) (void)hipFree(c->d_cnfac);
  if (c->h_pinned) (void)hipHostFree(c->h_pinned);
  if (c->stream) (void)hipStreamDestroy(c->stream);
  delete c;
}

// ---- host <-> device transfers in the reference layouts -------------------
int upload(ptc_ctx *c, int id, const double *host, int layout) {
  PTC_TRY(check_field(c, id));
  PTC_CHECK(host != nullptr, PTC_E_ARG, "null host pointer");
  double *d = nullptr;
  PTC_TRY(field(c, id, &d));
  const int nx = c->nx, ny = c->ny, nz = c->nz;
  if (is_plane(id)) {
    PTC_CHECK(layout == PTC_L_PLANE, PTC_E_ARG, "wall-plane fields take PTC_L_PLANE");
    PTC_HIP(hipMemcpyAsync(d, host, sizeof(double) * nx * nz, hipMemcpyHostToDevice, c->stream));
    PTC_HIP(hipStreamSynchronize(c->stream));
    return PTC_OK;
  }
  PTC_CHECK(layout == PTC_L_HALO || layout == PTC_L_INTERIOR, PTC_E_ARG, "3-D fields take PTC_L_HALO or PTC_L_INTERIOR");
  std::vector<double> st(c->n3, 0.0);
  const size_t hx = (size_t)nx + 4, hy = (size_t)ny + 2;
  for (int k = 0; k < nz; ++k)
    for (int j = 0; j <= ny + 1; ++j) {
      double *row = st.data() + (size_t)nx * ((size_t)j + hy * k);
      if (layout == PTC_L_HALO) {
        const double *src = host + 2 + hx * ((size_t)j + hy * (size_t)(k + 2));
        std::memcpy(row, src, sizeof(double) * nx);
      } else if (j >= 1 && j <= ny) {
        const double *src = host + (size_t)nx * ((size_t)(j - 1) + (size_t)ny * k);
        std::memcpy(row, src, sizeof(double) * nx);
      }
    }
  PTC_HIP(hipMemcpyAsync(d, st.data(), sizeof(double) * c->n3, hipMemcpyHostToDevice, c->stream));
  PTC_HIP(hipStreamSynchronize(c->stream));
  return PTC_OK;
}

int download(ptc_ctx *c, int id, double *host, int layout) {
  PTC_TRY(check_field(c, id));
  PTC_CHECK(host != nullptr, PTC_E_ARG, "null host pointer");
  double *d = nullptr;
  PTC_TRY(field(c, id, &d));
  const int nx = c->nx, ny = c->ny, nz = c->nz;
  if (is_plane(id)) {
    PTC_CHECK(layout == PTC_L_PLANE, PTC_E_ARG, "wall-plane fields take PTC_L_PLANE");
    PTC_HIP(hipMemcpyAsync(host, d, sizeof(double) * nx * nz, hipMemcpyDeviceToHost, c->stream));
    PTC_HIP(hipStreamSynchronize(c->stream));
    return PTC_OK;
  }
  PTC_CHECK(layout == PTC_L_HALO || layout == PTC_L_INTERIOR, PTC_E_ARG, "3-D fields take PTC_L_HALO or PTC_L_INTERIOR");
  std::vector<double> st(c->n3);
  PTC_HIP(hipMemcpyAsync(st.data(), d, sizeof(double) * c->n3, hipMemcpyDeviceToHost, c->stream));
  PTC_HIP(hipStreamSynchronize(c->stream));
  const size_t hx = (size_t)nx + 4, hy = (size_t)ny + 2;
  if (layout == PTC_L_INTERIOR) {
    for (int k = 0; k < nz; ++k)
      for (int j = 1; j <= ny; ++j)
        std::memcpy(host + (size_t)nx * ((size_t)(j - 1) + (size_t)ny * k),
                    st.data() + (size_t)nx * ((size_t)j + hy * k), sizeof(double) * nx);
    return PTC_OK;
  }
  // haloed: every halo cell holds its periodic image (x then z, corners included)
  for (int kk = -1; kk <= nz + 2; ++kk) {
    const int k = ((kk - 1) % nz + nz) % nz;  // reference kc = kk -> device k
    for (int j = 0; j <= ny + 1; ++j) {
      const double *row = st.data() + (size_t)nx * ((size_t)j + hy * k);
      double *dst = host + hx * ((size_t)j + hy * (size_t)(kk + 1));
      for (int ii = -1; ii <= nx + 2; ++ii) {
        const int i = ((ii - 1) % nx + nx) % nx;
        dst[ii + 1] = row[i];
      }
    }
  }
  return PTC_OK;
}

}  // namespace ptc
