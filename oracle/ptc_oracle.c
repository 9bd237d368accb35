/*
 * ptc_oracle.c -- CPU restatement (FP64) of ParaTC's per-time-step hot path.
 *
 * TEST INFRASTRUCTURE ONLY: the checker for the HIP library.  Only tests/,
 * __graft_entry__.smoke() and bench.py's cpu_baseline leg may load it; the
 * product path (paratc_amd/) never links, loads or calls anything here.
 *
 * Each routine restates one reference routine, cited as file:line into the
 * ParaTC tree (src/...).  Loop order and operand order follow the reference so
 * that rounding matches it as closely as a C compiler allows (built with
 * -ffp-contract=off).  Outer loops over z planes are OpenMP-parallel; every
 * reduction is done serially in the reference's (k, j, i) order so results do
 * not depend on the thread count.
 *
 * Third-party arithmetic: the reference calls FFTW 3.3.9 (R2HC / HC2R r2r
 * kinds, src/m_Poisson.f90:208-209,230-231).  FFTW is not linked here; the
 * transforms are restated from their published definition
 *     R2HC: r_k = Re X_k (k = 0..n/2),  r_{n-k} = Im X_k (0 < k < n/2),
 *           X_k = sum_j x_j exp(-2 pi i j k / n)
 *     HC2R: the unnormalised inverse, HC2R(R2HC(x)) = n x
 * with an own mixed-radix complex FFT.  Tests pin it against the O(n^2) DFT
 * and numpy.fft.
 */
#include "ptc_oracle.h"

#include <math.h>
#include <stdint.h>
#include <stdlib.h>
#include <string.h>
#ifdef _OPENMP
#include <omp.h>
#endif

/* ------------------------------------------------------------------------ */
/* indexing                                                                  */
/* ------------------------------------------------------------------------ */
#define NXH(c) ((size_t)(c)->nx + 4)
#define NYH(c) ((size_t)(c)->ny + 2)
/* haloed: i in [-1,nx+2], j in [0,ny+1], k in [-1,nz+2] (reference indices) */
#define H(c, a, i, j, k) (a)[(size_t)((i) + 1) + NXH(c) * ((size_t)(j) + NYH(c) * (size_t)((k) + 1))]
/* interior: i in [1,nx], j in [1,ny], k in [1,nz] */
#define I3(c, a, i, j, k) (a)[(size_t)((i)-1) + (size_t)(c)->nx * ((size_t)((j)-1) + (size_t)(c)->ny * (size_t)((k)-1))]
/* wall plane: i in [1,nx], k in [1,nz] */
#define P2(c, a, i, k) (a)[(size_t)((i)-1) + (size_t)(c)->nx * (size_t)((k)-1)]

typedef struct { double re, im; } cpx;

typedef struct fftplan {
  int n;
  int nfac;
  int fac[40];
  cpx *tw; /* tw[m] = exp(-2 pi i m / n) */
} fftplan;

struct orc_ctx {
  orc_cfg cfg;
  int nx, ny, nz;
  /* m_MeshAndMetries scalars (src/m_MeshAndMetries.f90:10-13,48-52) */
  double dx, dz, rdx, rdz, rdx2, rdz2;
  double dxc1, dxc2, dxc3, dxc4, dzc1, dzc2, dzc3, dzc4;
  double dxxc1, dxxc2, dxxc3, dxxc4, dxxc5, dzzc1, dzzc2, dzzc3, dzzc4, dzzc5;
  double ic1, ic2, ic3, ic4;
  /* y vectors, index j = 0..ny+1 */
  double *yp, *yc, *dyp, *dyc, *rdyp, *rdyc, *yint;
  double *am2ph, *ac2ph, *ap2ph, *am2c, *ac2c, *ap2c;
  double *am2cCN, *ac2cCN, *ap2cCN, *am2p, *ac2p, *ap2p;
  double *am2cSc, *ac2cSc, *ap2cSc, *am2cScCN, *ac2cScCN, *ap2cScCN;
  double *wnx, *wnz; /* WaveNumX(1..nx) at [0..nx-1] */
  double normfft;
  /* m_Parameters time-scheme state (src/m_Parameters.f90:40-61) */
  double gconst[3], tconst[3], aconst[3];
  int iadvance;
  double pmGamma, pmTheta, pmAlpha, pmAlphaC, pmBeta, pmBetaSc, dt, simtime, prgradaver;
  /* FFT plans */
  fftplan px, pz;
  /* haloed temporaries of clcRhsXYZ (src/m_TScheme.f90:130,137) */
  double *uxhx, *uyhx, *uzhx, *uxhz, *uyhz, *uzhz;
  double *s3buf; /* per-point s3tot contributions, summed serially */
};

static double *dalloc(size_t n) {
  double *p = (double *)calloc(n ? n : 1, sizeof(double));
  return p;
}

int orc_set_threads(int n) {
#ifdef _OPENMP
  if (n > 0) omp_set_num_threads(n);
  return omp_get_max_threads();
#else
  (void)n;
  return 1;
#endif
}

/* ------------------------------------------------------------------------ */
/* FFT (restated from the DFT definition; see header)                        */
/* ------------------------------------------------------------------------ */
static void plan_init(fftplan *p, int n) {
  p->n = n;
  p->nfac = 0;
  int m = n;
  const int pref[] = {4, 2, 3, 5, 7};
  for (int t = 0; t < 5; t++) {
    while (m % pref[t] == 0) {
      p->fac[p->nfac++] = pref[t];
      m /= pref[t];
    }
  }
  for (int f = 11; m > 1; f += 2) {
    while (m % f == 0) {
      p->fac[p->nfac++] = f;
      m /= f;
    }
  }
  p->tw = (cpx *)malloc(sizeof(cpx) * (size_t)(n > 0 ? n : 1));
  const double twopi = 6.283185307179586476925286766559005768;
  for (int k = 0; k < n; k++) {
    double a = twopi * (double)k / (double)n;
    p->tw[k].re = cos(a);
    p->tw[k].im = -sin(a);
  }
}

static void plan_free(fftplan *p) { free(p->tw); p->tw = NULL; }

/* out[k] = sum_j in[j*is] * w^(sign*j*k), w = exp(-2 pi i / n0 * step);
 * sign = -1 forward (exp(-...)), +1 inverse. */
static void fft_rec(cpx *out, const cpx *in, int n, int is, const int *fac,
                    const fftplan *pl, int step, int inv) {
  if (n == 1) {
    out[0] = in[0];
    return;
  }
  const int p = fac[0], m = n / p, N = pl->n;
  for (int q = 0; q < p; q++) fft_rec(out + (size_t)q * m, in + (size_t)q * is, m, is * p, fac + 1, pl, step * p, inv);
  cpx t[64];
  for (int k = 0; k < m; k++) {
    for (int q = 0; q < p; q++) {
      cpx y = out[(size_t)q * m + k];
      if (q == 0 || k == 0) {
        t[q] = y;
      } else {
        size_t e = ((size_t)q * (size_t)k * (size_t)step) % (size_t)N;
        cpx w = pl->tw[e];
        if (inv) w.im = -w.im;
        t[q].re = y.re * w.re - y.im * w.im;
        t[q].im = y.re * w.im + y.im * w.re;
      }
    }
    if (p == 2) {
      out[k].re = t[0].re + t[1].re;
      out[k].im = t[0].im + t[1].im;
      out[(size_t)m + k].re = t[0].re - t[1].re;
      out[(size_t)m + k].im = t[0].im - t[1].im;
    } else if (p == 4) {
      /* X_r = sum_q t_q (-i)^(q r) forward, (+i)^(q r) inverse */
      cpx a0 = {t[0].re + t[2].re, t[0].im + t[2].im};
      cpx a1 = {t[0].re - t[2].re, t[0].im - t[2].im};
      cpx b0 = {t[1].re + t[3].re, t[1].im + t[3].im};
      cpx b1 = {t[1].re - t[3].re, t[1].im - t[3].im};
      /* multiply b1 by -i (forward) or +i (inverse) */
      cpx jb1 = inv ? (cpx){-b1.im, b1.re} : (cpx){b1.im, -b1.re};
      out[k].re = a0.re + b0.re;
      out[k].im = a0.im + b0.im;
      out[(size_t)m + k].re = a1.re + jb1.re;
      out[(size_t)m + k].im = a1.im + jb1.im;
      out[(size_t)2 * m + k].re = a0.re - b0.re;
      out[(size_t)2 * m + k].im = a0.im - b0.im;
      out[(size_t)3 * m + k].re = a1.re - jb1.re;
      out[(size_t)3 * m + k].im = a1.im - jb1.im;
    } else {
      for (int r = 0; r < p; r++) {
        double sr = 0.0, si = 0.0;
        for (int q = 0; q < p; q++) {
          size_t e = (((size_t)q * (size_t)r) % (size_t)p) * (size_t)(N / p);
          cpx w = pl->tw[e];
          if (inv) w.im = -w.im;
          sr += t[q].re * w.re - t[q].im * w.im;
          si += t[q].re * w.im + t[q].im * w.re;
        }
        out[(size_t)r * m + k].re = sr;
        out[(size_t)r * m + k].im = si;
      }
    }
  }
}

static void cfft(const fftplan *pl, cpx *out, const cpx *in, int inv) {
  fft_rec(out, in, pl->n, 1, pl->fac, pl, 1, inv);
}

/* one R2HC line: x (stride is) -> hc in place */
static void r2hc_line(const fftplan *pl, double *x, size_t is, cpx *b1, cpx *b2) {
  const int n = pl->n;
  for (int j = 0; j < n; j++) { b1[j].re = x[(size_t)j * is]; b1[j].im = 0.0; }
  cfft(pl, b2, b1, 0);
  x[0] = b2[0].re;
  for (int k = 1; k < (n + 1) / 2; k++) {
    x[(size_t)k * is] = b2[k].re;
    x[(size_t)(n - k) * is] = b2[k].im;
  }
  if (n % 2 == 0) x[(size_t)(n / 2) * is] = b2[n / 2].re;
}

static void hc2r_line(const fftplan *pl, double *x, size_t is, cpx *b1, cpx *b2) {
  const int n = pl->n;
  b1[0].re = x[0];
  b1[0].im = 0.0;
  for (int k = 1; k < (n + 1) / 2; k++) {
    double re = x[(size_t)k * is], im = x[(size_t)(n - k) * is];
    b1[k].re = re; b1[k].im = im;
    b1[n - k].re = re; b1[n - k].im = -im;
  }
  if (n % 2 == 0) { b1[n / 2].re = x[(size_t)(n / 2) * is]; b1[n / 2].im = 0.0; }
  cfft(pl, b2, b1, 1);
  for (int j = 0; j < n; j++) x[(size_t)j * is] = b2[j].re;
}

static void r2r_batch(const fftplan *pl, int howmany, size_t istride, size_t idist,
                      double *data, int inverse) {
  const int n = pl->n;
#pragma omp parallel
  {
    cpx *b1 = (cpx *)malloc(sizeof(cpx) * (size_t)n);
    cpx *b2 = (cpx *)malloc(sizeof(cpx) * (size_t)n);
#pragma omp for schedule(static)
    for (int h = 0; h < howmany; h++) {
      double *x = data + (size_t)h * idist;
      if (inverse) hc2r_line(pl, x, istride, b1, b2);
      else r2hc_line(pl, x, istride, b1, b2);
    }
    free(b1);
    free(b2);
  }
}

void orc_r2hc(int n, int howmany, int istride, int idist, double *data) {
  fftplan p;
  plan_init(&p, n);
  r2r_batch(&p, howmany, (size_t)istride, (size_t)idist, data, 0);
  plan_free(&p);
}

void orc_hc2r(int n, int howmany, int istride, int idist, double *data) {
  fftplan p;
  plan_init(&p, n);
  r2r_batch(&p, howmany, (size_t)istride, (size_t)idist, data, 1);
  plan_free(&p);
}

/* ------------------------------------------------------------------------ */
/* Mesh and metrics: src/m_MeshAndMetries.f90:86-333                          */
/* ------------------------------------------------------------------------ */
static void init_metrics(orc_ctx *c) {
  const orc_cfg *g = &c->cfg;
  const int nyc = c->ny, nyp = c->ny + 1;
  const double nxr = (double)c->nx, nzr = (double)c->nz;
  /* :86-101 */
  c->dx = g->xlx / nxr;
  c->dz = g->zlz / nzr;
  c->rdx = nxr / g->xlx;
  c->rdz = nzr / g->zlz;
  c->rdx2 = c->rdx * c->rdx;
  c->rdz2 = c->rdz * c->rdz;
  c->dxc1 = c->rdx / 24.0; c->dxc2 = -c->rdx * 1.125; c->dxc3 = c->rdx * 1.125; c->dxc4 = -c->rdx / 24.0;
  c->dzc1 = c->rdz / 24.0; c->dzc2 = -c->rdz * 1.125; c->dzc3 = c->rdz * 1.125; c->dzc4 = -c->rdz / 24.0;
  c->dxxc1 = -c->rdx2 / 12.0; c->dxxc2 = 4.0 / 3.0 * c->rdx2; c->dxxc3 = -2.5 * c->rdx2;
  c->dxxc4 = 4.0 / 3.0 * c->rdx2; c->dxxc5 = -c->rdx2 / 12.0;
  c->dzzc1 = -c->rdz2 / 12.0; c->dzzc2 = 4.0 / 3.0 * c->rdz2; c->dzzc3 = -2.5 * c->rdz2;
  c->dzzc4 = 4.0 / 3.0 * c->rdz2; c->dzzc5 = -c->rdz2 / 12.0;
  c->ic1 = -1.0 / 16.0; c->ic2 = 9.0 / 16.0; c->ic3 = 9.0 / 16.0; c->ic4 = -1.0 / 16.0;

  double *yp = c->yp, *yc = c->yc, *dyp = c->dyp, *dyc = c->dyc, *rdyp = c->rdyp, *rdyc = c->rdyc;
  /* :195-199 */
  for (int j = 1; j <= nyc; j++) yc[j] = 0.5 * (yp[j] + yp[j + 1]);
  yc[0] = 2.0 * yp[1] - yc[1];
  yc[nyp] = 2.0 * yp[nyp] - yc[nyc];
  /* :202-206 */
  for (int j = 1; j <= nyc; j++) dyp[j] = yp[j + 1] - yp[j];
  dyp[0] = dyp[1];
  dyp[nyp] = dyp[nyc];
  /* :219-223 */
  for (int j = 2; j <= nyc; j++) dyc[j] = yc[j] - yc[j - 1];
  dyc[1] = dyp[1];
  dyc[nyp] = dyp[nyc];
  /* :226-231 */
  for (int j = 0; j <= nyp; j++) rdyp[j] = 1.0 / dyp[j];
  for (int j = 1; j <= nyp; j++) rdyc[j] = 1.0 / dyc[j];
  /* :234-236 */
  for (int j = 1; j <= nyp; j++) c->yint[j] = dyp[j] / (dyp[j] + dyp[j - 1]);
  /* :239-247 pressure Laplacian */
  for (int j = 1; j <= nyc; j++) {
    c->am2ph[j] = rdyp[j] * rdyc[j];
    c->ap2ph[j] = rdyp[j] * rdyc[j + 1];
  }
  c->am2ph[1] = 0.0;
  c->ap2ph[1] = rdyp[1] * rdyc[2];
  c->am2ph[nyc] = rdyp[nyc] * rdyc[nyc];
  c->ap2ph[nyc] = 0.0;
  for (int j = 1; j <= nyc; j++) c->ac2ph[j] = -(c->am2ph[j] + c->ap2ph[j]);
  /* :250-263 u/w Laplacian */
  for (int j = 1; j <= nyc; j++) {
    c->am2c[j] = rdyp[j] * rdyc[j];
    c->ap2c[j] = rdyp[j] * rdyc[j + 1];
  }
  c->am2c[1] = 4.0 * rdyc[1] / (dyc[1] + 2.0 * dyc[2]);
  c->ap2c[1] = 4.0 * rdyc[2] / (dyc[1] + 2.0 * dyc[2]);
  if (g->flow_type == 1) {
    c->am2c[nyc] = 4.0 * rdyc[nyc] / (dyc[nyp] + 2.0 * dyc[nyc]);
    c->ap2c[nyc] = 4.0 * rdyc[nyp] / (dyc[nyp] + 2.0 * dyc[nyc]);
  } else if (g->flow_type == 2) {
    c->am2c[nyc] = rdyp[nyc] * rdyc[nyc];
    c->ap2c[nyc] = 0.0;
  }
  for (int j = 1; j <= nyc; j++) c->ac2c[j] = -(c->am2c[j] + c->ap2c[j]);
  /* :266-274 Crank-Nicolson variant */
  for (int j = 1; j <= nyc; j++) {
    c->ap2cCN[j] = c->ap2c[j];
    c->ac2cCN[j] = c->ac2c[j];
    c->am2cCN[j] = c->am2c[j];
  }
  c->ap2cCN[1] = c->ap2c[1];
  c->ac2cCN[1] = c->ac2c[1] - c->am2c[1];
  c->am2cCN[1] = 2.0 * c->am2c[1];
  if (g->flow_type == 1) {
    c->am2cCN[nyc] = c->am2c[nyc];
    c->ac2cCN[nyc] = c->ac2c[nyc] - c->ap2c[nyc];
    c->ap2cCN[nyc] = 2.0 * c->ap2c[nyc];
  }
  /* :277-281 v Laplacian */
  for (int j = 1; j <= nyc; j++) {
    c->am2p[j] = rdyc[j] * rdyp[j - 1];
    c->ap2p[j] = rdyc[j] * rdyp[j];
  }
  for (int j = 1; j <= nyc; j++) c->ac2p[j] = -(c->am2p[j] + c->ap2p[j]);
  /* :284-333 scalar Laplacians */
  if (g->scalar) {
    for (int j = 1; j <= nyc; j++) {
      c->am2cSc[j] = rdyp[j] * rdyc[j];
      c->ap2cSc[j] = rdyp[j] * rdyc[j + 1];
    }
    if (g->sbc_opt[0] == -1) {
      c->am2cSc[1] = 4.0 * rdyc[1] / (dyc[1] + 2.0 * dyc[2]);
      c->ap2cSc[1] = 4.0 * rdyc[2] / (dyc[1] + 2.0 * dyc[2]);
    }
    if (g->sbc_opt[1] == -1) {
      c->am2cSc[nyc] = 4.0 * rdyc[nyc] / (dyc[nyp] + 2.0 * dyc[nyc]);
      c->ap2cSc[nyc] = 4.0 * rdyc[nyp] / (dyc[nyp] + 2.0 * dyc[nyc]);
    }
    for (int j = 1; j <= nyc; j++) c->ac2cSc[j] = -(c->am2cSc[j] + c->ap2cSc[j]);
    for (int j = 1; j <= nyc; j++) {
      c->am2cScCN[j] = rdyp[j] * rdyc[j];
      c->ap2cScCN[j] = rdyp[j] * rdyc[j + 1];
    }
    for (int j = 1; j <= nyc; j++) c->ac2cScCN[j] = -(c->am2cScCN[j] + c->ap2cScCN[j]);
    const double *sv = g->sbc_val;
    switch (g->sbc_opt[0]) {
      case -1:
        c->am2cScCN[1] = 8.0 * rdyc[1] / (dyc[1] + 2.0 * dyc[2]);
        c->ap2cScCN[1] = 4.0 * rdyc[2] / (dyc[1] + 2.0 * dyc[2]);
        c->ac2cScCN[1] = -(8.0 * rdyc[1] + 4.0 * rdyc[2]) / (dyc[1] + 2.0 * dyc[2]);
        break;
      case -2:
        c->ap2cScCN[1] = rdyp[1] * rdyc[2];
        c->ac2cScCN[1] = -rdyp[1] * rdyc[2];
        c->am2cScCN[1] = 0.0;
        break;
      case -3:
        c->ap2cScCN[1] = rdyp[1] * rdyc[2];
        c->ac2cScCN[1] = c->ac2cScCN[1] + c->am2cScCN[1] * (2.0 + sv[0] * dyc[1]) / (2.0 - sv[0] * dyc[1]);
        c->am2cScCN[1] = 0.0;
        break;
      default: break;
    }
    switch (g->sbc_opt[1]) {
      case -1:
        c->am2cScCN[nyc] = 4.0 * rdyc[nyc] / (dyc[nyp] + 2.0 * dyc[nyc]);
        c->ap2cScCN[nyc] = 8.0 * rdyc[nyp] / (dyc[nyp] + 2.0 * dyc[nyc]);
        c->ac2cScCN[nyc] = -(8.0 * rdyc[nyp] + 4.0 * rdyc[nyc]) / (dyc[nyp] + 2.0 * dyc[nyc]);
        break;
      case -2:
        c->am2cScCN[nyc] = rdyp[nyc] * rdyc[nyc];
        c->ac2cScCN[nyc] = -rdyp[nyc] * rdyc[nyc];
        c->ap2cScCN[nyc] = 0.0;
        break;
      case -3:
        c->am2cScCN[nyc] = rdyp[nyc] * rdyc[nyc];
        c->ac2cScCN[nyc] = c->ac2cScCN[nyc] + c->ap2cScCN[nyc] * (2.0 - sv[1] * dyc[nyp]) / (2.0 + sv[1] * dyc[nyp]);
        c->ap2cScCN[nyc] = 0.0;
        break;
      default: break;
    }
  }
  /* Poisson wavenumbers: src/m_Poisson.f90:37-48 */
  const double twopi = 6.283185307179586476925286766559005768;
  for (int i = 1; i <= c->nx; i++) {
    double wa1 = twopi * (double)(i - 1) / nxr;
    c->wnx[i - 1] = c->rdx2 / 288.0 * (cos(3.0 * wa1) - 54.0 * cos(2.0 * wa1) + 783.0 * cos(wa1) - 730.0);
  }
  for (int k = 1; k <= c->nz; k++) {
    double wa3 = twopi * (double)(k - 1) / nzr;
    c->wnz[k - 1] = c->rdz2 / 288.0 * (cos(3.0 * wa3) - 54.0 * cos(2.0 * wa3) + 783.0 * cos(wa3) - 730.0);
  }
  c->normfft = 1.0 / (nxr * nzr);
}

orc_ctx *orc_create(const orc_cfg *cfg, const double *yp) {
  orc_ctx *c = (orc_ctx *)calloc(1, sizeof(orc_ctx));
  c->cfg = *cfg;
  c->nx = cfg->nx;
  c->ny = cfg->ny;
  c->nz = cfg->nz;
  const size_t ny2 = (size_t)c->ny + 2;
  double **vecs[] = {&c->yp, &c->yc, &c->dyp, &c->dyc, &c->rdyp, &c->rdyc, &c->yint,
                     &c->am2ph, &c->ac2ph, &c->ap2ph, &c->am2c, &c->ac2c, &c->ap2c,
                     &c->am2cCN, &c->ac2cCN, &c->ap2cCN, &c->am2p, &c->ac2p, &c->ap2p,
                     &c->am2cSc, &c->ac2cSc, &c->ap2cSc, &c->am2cScCN, &c->ac2cScCN, &c->ap2cScCN};
  for (size_t v = 0; v < sizeof(vecs) / sizeof(vecs[0]); v++) *vecs[v] = dalloc(ny2);
  for (int j = 1; j <= c->ny + 1; j++) c->yp[j] = yp[j - 1];
  c->wnx = dalloc((size_t)c->nx);
  c->wnz = dalloc((size_t)c->nz);
  init_metrics(c);
  /* RK tables: src/m_Parameters.f90:159-174 */
  if (cfg->ischeme == 1) {
    c->iadvance = 1;
    c->gconst[0] = 3.0 / 2.0; c->gconst[1] = 0.0; c->gconst[2] = 0.0;
    c->tconst[0] = -0.5; c->tconst[1] = 0.0; c->tconst[2] = 0.0;
  } else if (cfg->ischeme == 2) {
    c->iadvance = 2;
    c->gconst[0] = 0.5; c->gconst[1] = 1.0; c->gconst[2] = 0.0;
    c->tconst[0] = 0.0; c->tconst[1] = -0.5; c->tconst[2] = 0.0;
  } else {
    c->iadvance = 3;
    c->gconst[0] = 8.0 / 15.0; c->gconst[1] = 5.0 / 12.0; c->gconst[2] = 3.0 / 4.0;
    c->tconst[0] = 0.0; c->tconst[1] = -17.0 / 60.0; c->tconst[2] = -5.0 / 12.0;
  }
  for (int s = 0; s < 3; s++) c->aconst[s] = c->gconst[s] + c->tconst[s];
  c->dt = cfg->dtmax; /* ChannelInitialize: dt = dtMax (m_ChannelSystem.f90:64) */
  c->simtime = 0.0;
  plan_init(&c->px, c->nx);
  plan_init(&c->pz, c->nz);
  const size_t nh = NXH(c) * NYH(c) * ((size_t)c->nz + 4);
  c->uxhx = dalloc(nh); c->uyhx = dalloc(nh); c->uzhx = dalloc(nh);
  c->uxhz = dalloc(nh); c->uyhz = dalloc(nh); c->uzhz = dalloc(nh);
  c->s3buf = dalloc((size_t)c->nx * c->ny * c->nz);
  return c;
}

void orc_destroy(orc_ctx *c) {
  if (!c) return;
  double *vecs[] = {c->yp, c->yc, c->dyp, c->dyc, c->rdyp, c->rdyc, c->yint,
                    c->am2ph, c->ac2ph, c->ap2ph, c->am2c, c->ac2c, c->ap2c,
                    c->am2cCN, c->ac2cCN, c->ap2cCN, c->am2p, c->ac2p, c->ap2p,
                    c->am2cSc, c->ac2cSc, c->ap2cSc, c->am2cScCN, c->ac2cScCN, c->ap2cScCN,
                    c->wnx, c->wnz, c->uxhx, c->uyhx, c->uzhx, c->uxhz, c->uyhz, c->uzhz, c->s3buf};
  for (size_t v = 0; v < sizeof(vecs) / sizeof(vecs[0]); v++) free(vecs[v]);
  plan_free(&c->px);
  plan_free(&c->pz);
  free(c);
}

int orc_get_metric(const orc_ctx *c, int which, double *out) {
  const double *tab[] = {c->yc, c->dyp, c->dyc, c->rdyp, c->rdyc, c->yint,
                         c->am2ph, c->ac2ph, c->ap2ph, c->am2c, c->ac2c, c->ap2c,
                         c->am2cCN, c->ac2cCN, c->ap2cCN, c->am2p, c->ac2p, c->ap2p,
                         c->am2cSc, c->ac2cSc, c->ap2cSc, c->am2cScCN, c->ac2cScCN, c->ap2cScCN};
  if (which < 0 || which >= (int)(sizeof(tab) / sizeof(tab[0]))) return -1;
  memcpy(out, tab[which], sizeof(double) * ((size_t)c->ny + 2));
  return 0;
}

void orc_wavenumbers(const orc_ctx *c, double *wnx, double *wnz) {
  memcpy(wnx, c->wnx, sizeof(double) * (size_t)c->nx);
  memcpy(wnz, c->wnz, sizeof(double) * (size_t)c->nz);
}

/* ------------------------------------------------------------------------ */
/* InverseTridiagonal: src/m_Tools.f90:224-251                               */
/* aj sub-diagonal, bj diagonal, cj super-diagonal, (m lines) x (n rows)     */
/* ------------------------------------------------------------------------ */
static void thomas(const double *aj, const double *bj, const double *cj, double *fj,
                   int m, int n, double *vecm, double *arrmn) {
#define A2(a, i, j) (a)[(size_t)((i)-1) + (size_t)m * (size_t)((j)-1)]
  for (int i = 1; i <= m; i++) {
    vecm[i - 1] = A2(bj, i, 1);
    A2(fj, i, 1) = A2(fj, i, 1) / vecm[i - 1];
  }
  for (int j = 2; j <= n; j++) {
    for (int i = 1; i <= m; i++) {
      A2(arrmn, i, j) = A2(cj, i, j - 1) / vecm[i - 1];
      vecm[i - 1] = A2(bj, i, j) - A2(aj, i, j) * A2(arrmn, i, j);
      A2(fj, i, j) = (A2(fj, i, j) - A2(aj, i, j) * A2(fj, i, j - 1)) / vecm[i - 1];
    }
  }
  for (int j = n - 1; j >= 1; j--) {
    for (int i = 1; i <= m; i++) A2(fj, i, j) = A2(fj, i, j) - A2(arrmn, i, j + 1) * A2(fj, i, j + 1);
  }
#undef A2
}

void orc_inverse_tridiagonal(const double *aj, const double *bj, const double *cj,
                             double *fj, int m, int n) {
  double *vecm = dalloc((size_t)m);
  double *arrmn = dalloc((size_t)m * n);
  thomas(aj, bj, cj, fj, m, n, vecm, arrmn);
  free(vecm);
  free(arrmn);
}

/* ------------------------------------------------------------------------ */
/* Poisson: src/m_Poisson.f90:352-452 (choice 1), 700-798 (choice 2),        */
/* 457-695 (choice 3), 803-1040 (choice 4), on one rank.                     */
/* ------------------------------------------------------------------------ */

/* y-solve of one kc plane, plain Thomas (:383-430).  f is (nx, ny) at plane kc. */
static void poisson_plane_thomas(const orc_ctx *c, int kc, double *f, double *m_, double *cc,
                                 double *p_, double *vecm, double *arrmn) {
  const int nx = c->nx, ny = c->ny;
  for (int jc = 1; jc <= ny; jc++)
    for (int ic = 1; ic <= nx; ic++) {
      m_[(size_t)(ic - 1) + (size_t)nx * (jc - 1)] = c->am2ph[jc];
      p_[(size_t)(ic - 1) + (size_t)nx * (jc - 1)] = c->ap2ph[jc];
      cc[(size_t)(ic - 1) + (size_t)nx * (jc - 1)] = c->ac2ph[jc] + c->wnx[ic - 1] + c->wnz[kc - 1];
    }
  if (kc == 1) {
    /* (:400-404) the (kx,kz)=(0,0) line: row j=1 replaced by phi=0 */
    p_[0] = 0.0;
    cc[0] = 1.0;
    m_[0] = 0.0;
    f[0] = 0.0;
  }
  thomas(m_, cc, p_, f, nx, ny, vecm, arrmn);
}

/* y-solve of one kc plane with the partition reduction on one rank (p_row = 1):
 * local modified Thomas (:497-539), reduced 2-row system
 * (InverseTridiagonal_reduce :1045-1071 with Initialize_ReduceMatrix :249-334),
 * reconstruction (:665-675). */
static void poisson_plane_reduce(const orc_ctx *c, int kc, double *f, double *m_, double *cc,
                                 double *p_) {
  const int nx = c->nx, ny = c->ny;
#define A2(a, i, j) (a)[(size_t)((i)-1) + (size_t)nx * (size_t)((j)-1)]
  for (int jc = 1; jc <= ny; jc++)
    for (int ic = 1; ic <= nx; ic++) {
      A2(m_, ic, jc) = c->am2ph[jc];
      A2(cc, ic, jc) = c->ac2ph[jc] + c->wnx[ic - 1] + c->wnz[kc - 1];
      A2(p_, ic, jc) = c->ap2ph[jc];
    }
  if (kc == 1) {
    A2(m_, 1, 1) = 0.0;
    A2(cc, 1, 1) = 1.0;
    A2(p_, 1, 1) = 0.0;
    A2(f, 1, 1) = 0.0;
  }
  for (int ic = 1; ic <= nx; ic++) {
    for (int jc = 1; jc <= 2; jc++) {
      A2(m_, ic, jc) = A2(m_, ic, jc) / A2(cc, ic, jc);
      A2(p_, ic, jc) = A2(p_, ic, jc) / A2(cc, ic, jc);
      A2(f, ic, jc) = A2(f, ic, jc) / A2(cc, ic, jc);
    }
  }
  for (int jc = 3; jc <= ny; jc++) {
    const int jm = jc - 1;
    for (int ic = 1; ic <= nx; ic++) {
      double r = 1.0 / (A2(cc, ic, jc) - A2(m_, ic, jc) * A2(p_, ic, jm));
      A2(f, ic, jc) = r * (A2(f, ic, jc) - A2(m_, ic, jc) * A2(f, ic, jm));
      A2(p_, ic, jc) = r * A2(p_, ic, jc);
      A2(m_, ic, jc) = -r * A2(m_, ic, jc) * A2(m_, ic, jm);
    }
  }
  for (int jc = ny - 2; jc >= 2; jc--) {
    const int jp = jc + 1;
    for (int ic = 1; ic <= nx; ic++) {
      A2(f, ic, jc) = A2(f, ic, jc) - A2(p_, ic, jc) * A2(f, ic, jp);
      A2(m_, ic, jc) = A2(m_, ic, jc) - A2(p_, ic, jc) * A2(m_, ic, jp);
      A2(p_, ic, jc) = -A2(p_, ic, jc) * A2(p_, ic, jp);
    }
  }
  for (int ic = 1; ic <= nx; ic++) {
    const int jc = 1, jp = 2;
    double r = 1.0 / (1.0 - A2(m_, ic, jp) * A2(p_, ic, jc));
    A2(f, ic, jc) = r * (A2(f, ic, jc) - A2(p_, ic, jc) * A2(f, ic, jp));
    A2(m_, ic, jc) = r * A2(m_, ic, jc);
    A2(p_, ic, jc) = -r * A2(p_, ic, jc) * A2(p_, ic, jp);
  }
  /* reduced system of 2 rows per line: unit diagonal, a = (m_1, m_ny), c = (p_1, p_ny) */
  for (int ic = 1; ic <= nx; ic++) {
    double f1 = A2(f, ic, 1), f2 = A2(f, ic, ny);
    double a2 = A2(m_, ic, ny), c1 = A2(p_, ic, 1);
    double vecm = 1.0;
    double arr = c1 / vecm;
    vecm = 1.0 - a2 * arr;
    f2 = (f2 - a2 * f1) / vecm;
    f1 = f1 - arr * f2;
    A2(f, ic, 1) = f1;
    A2(f, ic, ny) = f2;
    for (int jc = 2; jc <= ny - 1; jc++)
      A2(f, ic, jc) = A2(f, ic, jc) - A2(m_, ic, jc) * f1 - A2(p_, ic, jc) * f2;
  }
#undef A2
}

void orc_poisson(orc_ctx *c, int choice, double *prsrc, double *prphi) {
  const int nx = c->nx, ny = c->ny, nz = c->nz;
  const size_t plane = (size_t)nx * ny;
  const int xfirst = (choice == 1 || choice == 3);
  const int reduce = (choice == 3 || choice == 4);
  /* on one rank every transpose_* is an identity copy of the (nx,ny,nz) array
   * (src/m_Decomp2d.f90:606-611); the transforms act on prsrc in place
   * (-DOverWriteFFT, src/makefile:20). */
  if (xfirst) {
    r2r_batch(&c->px, ny * nz, 1, (size_t)nx, prsrc, 0);
    r2r_batch(&c->pz, nx * ny, plane, 1, prsrc, 0);
  } else {
    r2r_batch(&c->pz, nx * ny, plane, 1, prsrc, 0);
    r2r_batch(&c->px, ny * nz, 1, (size_t)nx, prsrc, 0);
  }
  if (ny >= 4 || !reduce) {
#pragma omp parallel
    {
      double *m_ = dalloc(plane), *cc = dalloc(plane), *p_ = dalloc(plane);
      double *vecm = dalloc((size_t)nx), *arrmn = dalloc(plane);
#pragma omp for schedule(static)
      for (int kc = 1; kc <= nz; kc++) {
        double *f = prsrc + plane * (size_t)(kc - 1);
        if (reduce) poisson_plane_reduce(c, kc, f, m_, cc, p_);
        else poisson_plane_thomas(c, kc, f, m_, cc, p_, vecm, arrmn);
      }
      free(m_); free(cc); free(p_); free(vecm); free(arrmn);
    }
  }
  if (xfirst) {
    r2r_batch(&c->pz, nx * ny, plane, 1, prsrc, 1);
    r2r_batch(&c->px, ny * nz, 1, (size_t)nx, prsrc, 1);
  } else {
    r2r_batch(&c->px, ny * nz, 1, (size_t)nx, prsrc, 1);
    r2r_batch(&c->pz, nx * ny, plane, 1, prsrc, 1);
  }
  /* :445-451 */
#pragma omp parallel for schedule(static)
  for (int kc = 1; kc <= nz; kc++)
    for (int jc = 1; jc <= ny; jc++)
      for (int ic = 1; ic <= nx; ic++) H(c, prphi, ic, jc, kc) = I3(c, prsrc, ic, jc, kc) * c->normfft;
}

/* ------------------------------------------------------------------------ */
/* PMcoeUpdate: src/m_Parameters.f90:215-230                                 */
/* ------------------------------------------------------------------------ */
void orc_pm_coe_update(orc_ctx *c, int ns, double dt) {
  const double a = c->aconst[ns - 1];
  c->dt = dt;
  c->pmGamma = c->gconst[ns - 1] * dt;
  c->pmTheta = c->tconst[ns - 1] * dt;
  c->pmAlpha = a * dt;
  c->pmBeta = 0.5 * a * c->cfg.xnu * dt;
  c->pmBetaSc = 0.5 * a * c->cfg.sdiff * dt;
  c->simtime = c->simtime + dt * a;
  c->pmAlphaC = a;
  if (ns == 1) c->prgradaver = 0.0;
}

void orc_get_coe(const orc_ctx *c, double *o) {
  o[0] = c->pmGamma; o[1] = c->pmTheta; o[2] = c->pmAlpha; o[3] = c->pmAlphaC;
  o[4] = c->pmBeta; o[5] = c->pmBetaSc; o[6] = c->dt; o[7] = c->simtime; o[8] = c->prgradaver;
}

void orc_set_coe(orc_ctx *c, const double *o) {
  c->pmGamma = o[0]; c->pmTheta = o[1]; c->pmAlpha = o[2]; c->pmAlphaC = o[3];
  c->pmBeta = o[4]; c->pmBetaSc = o[5]; c->dt = o[6]; c->simtime = o[7]; c->prgradaver = o[8];
}

/* ------------------------------------------------------------------------ */
/* myupdate_halo, y-pencil, one rank: src/m_Decomp2d.f90:1554-1709            */
/* (the y-periodic steps 1-2 are never requested: every HaloInfo has         */
/*  ymh = yph = 0, src/m_Variables.f90:52-70, src/m_TScheme.f90:31-41)        */
/* ------------------------------------------------------------------------ */
void orc_update_halo(const orc_ctx *c, double *mat, int xmh, int xph, int zmh, int zph) {
  const int nx = c->nx, ny = c->ny, nz = c->nz;
  /* step 3: x- halo, full y and z memory range */
  if (xmh > 0)
    for (int k = -1; k <= nz + 2; k++)
      for (int j = 0; j <= ny + 1; j++)
        for (int h = 1; h <= xmh; h++) H(c, mat, 1 - h, j, k) = H(c, mat, nx + 1 - h, j, k);
  /* step 4: x+ halo */
  if (xph > 0)
    for (int k = -1; k <= nz + 2; k++)
      for (int j = 0; j <= ny + 1; j++)
        for (int h = 1; h <= xph; h++) H(c, mat, nx + h, j, k) = H(c, mat, h, j, k);
  /* step 5: z- halo, full x and y memory range */
  if (zmh > 0)
    for (int h = 1; h <= zmh; h++)
      for (int j = 0; j <= ny + 1; j++)
        for (int i = -1; i <= nx + 2; i++) H(c, mat, i, j, 1 - h) = H(c, mat, i, j, nz + 1 - h);
  /* step 6: z+ halo */
  if (zph > 0)
    for (int h = 1; h <= zph; h++)
      for (int j = 0; j <= ny + 1; j++)
        for (int i = -1; i <= nx + 2; i++) H(c, mat, i, j, nz + h) = H(c, mat, i, j, h);
}

/* ------------------------------------------------------------------------ */
/* clcRhsXYZ: src/m_TScheme.f90:113-360                                      */
/* ------------------------------------------------------------------------ */
void orc_clc_rhs_xyz(orc_ctx *c, const double *ux, const double *uy, const double *uz,
                     const double *scalar, double *rhsx, double *rhsy, double *rhsz,
                     double *rhsc, double *histx, double *histy, double *histz,
                     double *histc, const double *pressure) {
  const int nx = c->nx, ny = c->ny, nz = c->nz, nyp = ny + 1;
  const int sc = c->cfg.scalar;
  const double IC1 = c->ic1, IC2 = c->ic2, IC3 = c->ic3, IC4 = c->ic4;
  double *uxhx = c->uxhx, *uyhx = c->uyhx, *uxhz = c->uxhz, *uyhz = c->uyhz, *uzhz = c->uzhz;
  /* ScalarFlow build: uzhx is associated with RhsC (:141) */
  double *uzhx = sc ? rhsc : c->uzhx;
  const double uCRF = c->cfg.ucrf, xnu = c->cfg.xnu;
  const double *grav = c->cfg.gravity;
  const double pmGamma = c->pmGamma, pmTheta = c->pmTheta, pmAlpha = c->pmAlpha, pmBeta = c->pmBeta;

  /* Update_VelInterp_uvw (:143-161) */
#pragma omp parallel for schedule(static)
  for (int kc = 1; kc <= nz; kc++) {
    const int ks = kc - 2, km = kc - 1, kp = kc + 1, ku = kc + 2;
    for (int jc = 1; jc <= ny; jc++)
      for (int ic = 1; ic <= nx; ic++) {
        const int is = ic - 2, im = ic - 1, ip = ic + 1, iu = ic + 2;
        H(c, uxhx, ic, jc, kc) = IC1 * H(c, ux, im, jc, kc) + IC2 * H(c, ux, ic, jc, kc) + IC3 * H(c, ux, ip, jc, kc) + IC4 * H(c, ux, iu, jc, kc);
        H(c, uyhx, ic, jc, kc) = IC1 * H(c, uy, is, jc, kc) + IC2 * H(c, uy, im, jc, kc) + IC3 * H(c, uy, ic, jc, kc) + IC4 * H(c, uy, ip, jc, kc);
        H(c, uzhx, ic, jc, kc) = IC1 * H(c, uz, is, jc, kc) + IC2 * H(c, uz, im, jc, kc) + IC3 * H(c, uz, ic, jc, kc) + IC4 * H(c, uz, ip, jc, kc);
        H(c, uxhz, ic, jc, kc) = IC1 * H(c, ux, ic, jc, ks) + IC2 * H(c, ux, ic, jc, km) + IC3 * H(c, ux, ic, jc, kc) + IC4 * H(c, ux, ic, jc, kp);
        H(c, uyhz, ic, jc, kc) = IC1 * H(c, uy, ic, jc, ks) + IC2 * H(c, uy, ic, jc, km) + IC3 * H(c, uy, ic, jc, kc) + IC4 * H(c, uy, ic, jc, kp);
        H(c, uzhz, ic, jc, kc) = IC1 * H(c, uz, ic, jc, km) + IC2 * H(c, uz, ic, jc, kc) + IC3 * H(c, uz, ic, jc, kp) + IC4 * H(c, uz, ic, jc, ku);
      }
    for (int ic = 1; ic <= nx; ic++) {
      H(c, uyhx, ic, nyp, kc) = 0.0;
      H(c, uyhz, ic, nyp, kc) = 0.0;
    }
  }
  orc_update_halo(c, uxhx, 2, 1, 0, 0);
  orc_update_halo(c, uyhx, 1, 2, 0, 0);
  orc_update_halo(c, uzhx, 1, 2, 1, 2);
  orc_update_halo(c, uxhz, 1, 2, 1, 2);
  orc_update_halo(c, uyhz, 0, 0, 1, 2);
  orc_update_halo(c, uzhz, 0, 0, 2, 1);

  /* RhsX (:169-204) */
#pragma omp parallel for schedule(static)
  for (int kc = 1; kc <= nz; kc++) {
    const int ks = kc - 2, km = kc - 1, kp = kc + 1, ku = kc + 2;
    for (int jc = 1; jc <= ny; jc++) {
      const int jm = jc - 1, jp = jc + 1;
      const double sucaj = c->rdyp[jc];
      const double IY1 = c->yint[jc], IY2 = 1.0 - IY1, IY3 = c->yint[jp], IY4 = 1.0 - IY3;
      for (int ic = 1; ic <= nx; ic++) {
        const int is = ic - 2, im = ic - 1, ip = ic + 1, iu = ic + 2;
        double uxhyc = IY1 * H(c, ux, ic, jm, kc) + IY2 * H(c, ux, ic, jc, kc);
        double uxhyp = IY3 * H(c, ux, ic, jc, kc) + IY4 * H(c, ux, ic, jp, kc);
        double h11 = c->dxc1 * H(c, uxhx, is, jc, kc) * (H(c, uxhx, is, jc, kc) + uCRF) + c->dxc2 * H(c, uxhx, im, jc, kc) * (H(c, uxhx, im, jc, kc) + uCRF)
                   + c->dxc3 * H(c, uxhx, ic, jc, kc) * (H(c, uxhx, ic, jc, kc) + uCRF) + c->dxc4 * H(c, uxhx, ip, jc, kc) * (H(c, uxhx, ip, jc, kc) + uCRF);
        double h12 = (H(c, uyhx, ic, jp, kc) * uxhyp - H(c, uyhx, ic, jc, kc) * uxhyc) * sucaj;
        double h13 = c->dzc1 * H(c, uxhz, ic, jc, km) * H(c, uzhx, ic, jc, km) + c->dzc2 * H(c, uxhz, ic, jc, kc) * H(c, uzhx, ic, jc, kc)
                   + c->dzc3 * H(c, uxhz, ic, jc, kp) * H(c, uzhx, ic, jc, kp) + c->dzc4 * H(c, uxhz, ic, jc, ku) * H(c, uzhx, ic, jc, ku);
        double d11q1 = c->dxxc1 * H(c, ux, is, jc, kc) + c->dxxc2 * H(c, ux, im, jc, kc) + c->dxxc3 * H(c, ux, ic, jc, kc) + c->dxxc4 * H(c, ux, ip, jc, kc) + c->dxxc5 * H(c, ux, iu, jc, kc);
        double d22q1 = c->ap2c[jc] * H(c, ux, ic, jp, kc) + c->ac2c[jc] * H(c, ux, ic, jc, kc) + c->am2c[jc] * H(c, ux, ic, jm, kc);
        double d33q1 = c->dzzc1 * H(c, ux, ic, jc, ks) + c->dzzc2 * H(c, ux, ic, jc, km) + c->dzzc3 * H(c, ux, ic, jc, kc) + c->dzzc4 * H(c, ux, ic, jc, kp) + c->dzzc5 * H(c, ux, ic, jc, ku);
        double dcq13 = d11q1 + d33q1;
        if (sc) I3(c, c->s3buf, ic, jc, kc) = H(c, ux, ic, jc, kc) * c->dyp[jc];
        else I3(c, c->s3buf, ic, jc, kc) = (dcq13 + d22q1) * c->dyp[jc];
        double convEd1 = -h11 - h12 - h13 + xnu * dcq13 + grav[0];
        double gradp1 = c->dxc1 * H(c, pressure, is, jc, kc) + c->dxc2 * H(c, pressure, im, jc, kc) + c->dxc3 * H(c, pressure, ic, jc, kc) + c->dxc4 * H(c, pressure, ip, jc, kc);
        I3(c, rhsx, ic, jc, kc) = pmGamma * convEd1 + pmTheta * I3(c, histx, ic, jc, kc) - pmAlpha * gradp1 + 2.0 * pmBeta * d22q1;
        I3(c, histx, ic, jc, kc) = convEd1;
      }
    }
  }

  /* RhsY (:206-236) */
#pragma omp parallel for schedule(static)
  for (int kc = 1; kc <= nz; kc++) {
    const int ks = kc - 2, km = kc - 1, kp = kc + 1, ku = kc + 2;
    for (int jc = 1; jc <= ny; jc++) {
      const int jm = jc - 1, jp = jc + 1;
      const double sucac = c->rdyc[jc], qsucac = 0.25 * sucac;
      const double IY1 = c->yint[jc], IY2 = 1.0 - IY1;
      for (int ic = 1; ic <= nx; ic++) {
        const int is = ic - 2, im = ic - 1, ip = ic + 1, iu = ic + 2;
        double uxhym = IY1 * H(c, ux, im, jm, kc) + IY2 * H(c, ux, im, jc, kc) + uCRF;
        double uzhym = IY1 * H(c, uz, ic, jm, km) + IY2 * H(c, uz, ic, jc, km);
        double uxhyc = IY1 * H(c, ux, ic, jm, kc) + IY2 * H(c, ux, ic, jc, kc) + uCRF;
        double uzhyc = IY1 * H(c, uz, ic, jm, kc) + IY2 * H(c, uz, ic, jc, kc);
        double uxhyp = IY1 * H(c, ux, ip, jm, kc) + IY2 * H(c, ux, ip, jc, kc) + uCRF;
        double uzhyp = IY1 * H(c, uz, ic, jm, kp) + IY2 * H(c, uz, ic, jc, kp);
        double uxhyu = IY1 * H(c, ux, iu, jm, kc) + IY2 * H(c, ux, iu, jc, kc) + uCRF;
        double uzhyu = IY1 * H(c, uz, ic, jm, ku) + IY2 * H(c, uz, ic, jc, ku);
        double h21 = c->dxc1 * H(c, uyhx, im, jc, kc) * uxhym + c->dxc2 * H(c, uyhx, ic, jc, kc) * uxhyc + c->dxc3 * H(c, uyhx, ip, jc, kc) * uxhyp + c->dxc4 * H(c, uyhx, iu, jc, kc) * uxhyu;
        double s1 = H(c, uy, ic, jp, kc) + H(c, uy, ic, jc, kc);
        double s2 = H(c, uy, ic, jc, kc) + H(c, uy, ic, jm, kc);
        double h22 = (s1 * s1 - s2 * s2) * qsucac;
        double h23 = c->dzc1 * H(c, uyhz, ic, jc, km) * uzhym + c->dzc2 * H(c, uyhz, ic, jc, kc) * uzhyc + c->dzc3 * H(c, uyhz, ic, jc, kp) * uzhyp + c->dzc4 * H(c, uyhz, ic, jc, ku) * uzhyu;
        double d11q2 = c->dxxc1 * H(c, uy, is, jc, kc) + c->dxxc2 * H(c, uy, im, jc, kc) + c->dxxc3 * H(c, uy, ic, jc, kc) + c->dxxc4 * H(c, uy, ip, jc, kc) + c->dxxc5 * H(c, uy, iu, jc, kc);
        double d22q2 = c->ap2p[jc] * H(c, uy, ic, jp, kc) + c->ac2p[jc] * H(c, uy, ic, jc, kc) + c->am2p[jc] * H(c, uy, ic, jm, kc);
        double d33q2 = c->dzzc1 * H(c, uy, ic, jc, ks) + c->dzzc2 * H(c, uy, ic, jc, km) + c->dzzc3 * H(c, uy, ic, jc, kc) + c->dzzc4 * H(c, uy, ic, jc, kp) + c->dzzc5 * H(c, uy, ic, jc, ku);
        double dcq13 = d11q2 + d33q2;
        double convEd2 = -h21 - h22 - h23 + xnu * dcq13 + grav[1];
        double gradp2 = (H(c, pressure, ic, jc, kc) - H(c, pressure, ic, jm, kc)) * sucac;
        I3(c, rhsy, ic, jc, kc) = pmGamma * convEd2 + pmTheta * I3(c, histy, ic, jc, kc) - pmAlpha * gradp2 + 2.0 * pmBeta * d22q2;
        I3(c, histy, ic, jc, kc) = convEd2;
      }
    }
  }

  /* RhsZ (:238-267) */
#pragma omp parallel for schedule(static)
  for (int kc = 1; kc <= nz; kc++) {
    const int ks = kc - 2, km = kc - 1, kp = kc + 1, ku = kc + 2;
    for (int jc = 1; jc <= ny; jc++) {
      const int jm = jc - 1, jp = jc + 1;
      const double sucaj = c->rdyp[jc];
      const double IY1 = c->yint[jc], IY2 = 1.0 - IY1, IY3 = c->yint[jp], IY4 = 1.0 - IY3;
      for (int ic = 1; ic <= nx; ic++) {
        const int is = ic - 2, im = ic - 1, ip = ic + 1, iu = ic + 2;
        double uzhyc = IY1 * H(c, uz, ic, jm, kc) + IY2 * H(c, uz, ic, jc, kc);
        double uzhyp = IY3 * H(c, uz, ic, jc, kc) + IY4 * H(c, uz, ic, jp, kc);
        double h31 = c->dxc1 * H(c, uzhx, im, jc, kc) * (H(c, uxhz, im, jc, kc) + uCRF) + c->dxc2 * H(c, uzhx, ic, jc, kc) * (H(c, uxhz, ic, jc, kc) + uCRF)
                   + c->dxc3 * H(c, uzhx, ip, jc, kc) * (H(c, uxhz, ip, jc, kc) + uCRF) + c->dxc4 * H(c, uzhx, iu, jc, kc) * (H(c, uxhz, iu, jc, kc) + uCRF);
        double h32 = (H(c, uyhz, ic, jp, kc) * uzhyp - H(c, uyhz, ic, jc, kc) * uzhyc) * sucaj;
        double h33 = c->dzc1 * H(c, uzhz, ic, jc, ks) * H(c, uzhz, ic, jc, ks) + c->dzc2 * H(c, uzhz, ic, jc, km) * H(c, uzhz, ic, jc, km)
                   + c->dzc3 * H(c, uzhz, ic, jc, kc) * H(c, uzhz, ic, jc, kc) + c->dzc4 * H(c, uzhz, ic, jc, kp) * H(c, uzhz, ic, jc, kp);
        double d11q3 = c->dxxc1 * H(c, uz, is, jc, kc) + c->dxxc2 * H(c, uz, im, jc, kc) + c->dxxc3 * H(c, uz, ic, jc, kc) + c->dxxc4 * H(c, uz, ip, jc, kc) + c->dxxc5 * H(c, uz, iu, jc, kc);
        double d22q3 = c->ap2c[jc] * H(c, uz, ic, jp, kc) + c->ac2c[jc] * H(c, uz, ic, jc, kc) + c->am2c[jc] * H(c, uz, ic, jm, kc);
        double d33q3 = c->dzzc1 * H(c, uz, ic, jc, ks) + c->dzzc2 * H(c, uz, ic, jc, km) + c->dzzc3 * H(c, uz, ic, jc, kc) + c->dzzc4 * H(c, uz, ic, jc, kp) + c->dzzc5 * H(c, uz, ic, jc, ku);
        double dcq13 = d11q3 + d33q3;
        double convEd3 = -h31 - h32 - h33 + xnu * dcq13 + grav[2];
        double gradp3 = c->dzc1 * H(c, pressure, ic, jc, ks) + c->dzc2 * H(c, pressure, ic, jc, km) + c->dzc3 * H(c, pressure, ic, jc, kc) + c->dzc4 * H(c, pressure, ic, jc, kp);
        H(c, rhsz, ic, jc, kc) = pmGamma * convEd3 + pmTheta * I3(c, histz, ic, jc, kc) - pmAlpha * gradp3 + 2.0 * pmBeta * d22q3;
        I3(c, histz, ic, jc, kc) = convEd3;
      }
    }
  }

  /* constant flow rate (:269-300); the sum runs in the reference's order */
  if (c->cfg.is_ux_const) {
    double s3tot = 0.0;
    const size_t n = (size_t)nx * ny * nz;
    for (size_t q = 0; q < n; q++) s3tot += c->s3buf[q];
    const double nxnz = (double)(nx * nz);
    if (sc) {
      double s3tot1 = s3tot / nxnz / c->cfg.yly;
      double dpmdxns = c->cfg.ubulk - uCRF - s3tot1;
      double dp1ns = c->pmAlphaC * dpmdxns;
      for (size_t q = 0; q < n; q++) rhsx[q] = rhsx[q] + dp1ns;
      c->prgradaver = c->prgradaver + dpmdxns * c->pmAlphaC / c->dt;
    } else {
      double dpmdxns = xnu * s3tot / nxnz / c->cfg.yly;
      double dp1ns = pmAlpha * dpmdxns;
      for (size_t q = 0; q < n; q++) rhsx[q] = rhsx[q] - dp1ns;
      c->prgradaver = c->prgradaver + dpmdxns * c->pmAlphaC;
    }
  }

  if (!sc) return;
  /* scalar transport (:302-357); shx => uyhx, shz => uyhz */
  double *shx = uyhx, *shz = uyhz;
  const double *s = scalar;
#pragma omp parallel for schedule(static)
  for (int kc = 1; kc <= nz; kc++) {
    const int ks = kc - 2, km = kc - 1, kp = kc + 1;
    for (int jc = 1; jc <= ny; jc++)
      for (int ic = 1; ic <= nx; ic++) {
        const int is = ic - 2, im = ic - 1, ip = ic + 1;
        H(c, shx, ic, jc, kc) = IC1 * H(c, s, is, jc, kc) + IC2 * H(c, s, im, jc, kc) + IC3 * H(c, s, ic, jc, kc) + IC4 * H(c, s, ip, jc, kc);
        H(c, shz, ic, jc, kc) = IC1 * H(c, s, ic, jc, ks) + IC2 * H(c, s, ic, jc, km) + IC3 * H(c, s, ic, jc, kc) + IC4 * H(c, s, ic, jc, kp);
      }
  }
  orc_update_halo(c, shx, 1, 2, 0, 0);
  orc_update_halo(c, shz, 0, 0, 1, 2);
  const double *gev = c->cfg.gev, *fvv = c->cfg.fvv;
  const double sdiff = c->cfg.sdiff, pmBetaSc = c->pmBetaSc;
#pragma omp parallel for schedule(static)
  for (int kc = 1; kc <= nz; kc++) {
    const int ks = kc - 2, km = kc - 1, kp = kc + 1, ku = kc + 2;
    for (int jc = 1; jc <= ny; jc++) {
      const int jm = jc - 1, jp = jc + 1;
      const double sucaj = c->rdyp[jc];
      const double IY1 = c->yint[jc], IY2 = 1.0 - IY1, IY3 = c->yint[jp], IY4 = 1.0 - IY3;
      for (int ic = 1; ic <= nx; ic++) {
        const int is = ic - 2, im = ic - 1, ip = ic + 1, iu = ic + 2;
        double shyc = IY1 * H(c, s, ic, jm, kc) + IY2 * H(c, s, ic, jc, kc);
        double shyp = IY3 * H(c, s, ic, jc, kc) + IY4 * H(c, s, ic, jp, kc);
        double convEd1 = gev[0] * H(c, shx, ic, jc, kc);
        double convEd2 = gev[1] * shyc;
        double convEd3 = gev[2] * H(c, shz, ic, jc, kc);
        I3(c, rhsx, ic, jc, kc) = I3(c, rhsx, ic, jc, kc) + pmGamma * convEd1;
        I3(c, rhsy, ic, jc, kc) = I3(c, rhsy, ic, jc, kc) + pmGamma * convEd2;
        H(c, rhsz, ic, jc, kc) = H(c, rhsz, ic, jc, kc) + pmGamma * convEd3;
        I3(c, histx, ic, jc, kc) = I3(c, histx, ic, jc, kc) + convEd1;
        I3(c, histy, ic, jc, kc) = I3(c, histy, ic, jc, kc) + convEd2;
        I3(c, histz, ic, jc, kc) = I3(c, histz, ic, jc, kc) + convEd3;
        double hc1 = c->dxc1 * (H(c, ux, im, jc, kc) + fvv[0]) * H(c, shx, im, jc, kc) + c->dxc2 * (H(c, ux, ic, jc, kc) + fvv[0]) * H(c, shx, ic, jc, kc)
                   + c->dxc3 * (H(c, ux, ip, jc, kc) + fvv[0]) * H(c, shx, ip, jc, kc) + c->dxc4 * (H(c, ux, iu, jc, kc) + fvv[0]) * H(c, shx, iu, jc, kc);
        double hc2 = ((H(c, uy, ic, jp, kc) + fvv[1]) * shyp - (H(c, uy, ic, jc, kc) + fvv[1]) * shyc) * sucaj;
        double hc3 = c->dzc1 * (H(c, uz, ic, jc, km) + fvv[2]) * H(c, shz, ic, jc, km) + c->dzc2 * (H(c, uz, ic, jc, kc) + fvv[2]) * H(c, shz, ic, jc, kc)
                   + c->dzc3 * (H(c, uz, ic, jc, kp) + fvv[2]) * H(c, shz, ic, jc, kp) + c->dzc4 * (H(c, uz, ic, jc, ku) + fvv[2]) * H(c, shz, ic, jc, ku);
        double d11sc = c->dxxc1 * H(c, s, is, jc, kc) + c->dxxc2 * H(c, s, im, jc, kc) + c->dxxc3 * H(c, s, ic, jc, kc) + c->dxxc4 * H(c, s, ip, jc, kc) + c->dxxc5 * H(c, s, iu, jc, kc);
        double d22sc = c->ap2cSc[jc] * H(c, s, ic, jp, kc) + c->ac2cSc[jc] * H(c, s, ic, jc, kc) + c->am2cSc[jc] * H(c, s, ic, jm, kc);
        double d33sc = c->dzzc1 * H(c, s, ic, jc, ks) + c->dzzc2 * H(c, s, ic, jc, km) + c->dzzc3 * H(c, s, ic, jc, kc) + c->dzzc4 * H(c, s, ic, jc, kp) + c->dzzc5 * H(c, s, ic, jc, ku);
        double convEdc = -hc1 - hc2 - hc3 + sdiff * (d11sc + d33sc);
        H(c, rhsc, ic, jc, kc) = pmGamma * convEdc + pmTheta * I3(c, histc, ic, jc, kc) + 2.0 * pmBetaSc * d22sc;
        I3(c, histc, ic, jc, kc) = convEdc;
      }
    }
  }
}

/* ------------------------------------------------------------------------ */
/* Crank-Nicolson y-solves: src/m_TScheme.f90:365-506                         */
/* ------------------------------------------------------------------------ */
typedef enum { CN_U, CN_V, CN_C } cn_kind;

static void cn_solve(orc_ctx *c, cn_kind kind, double *u, const double *rhs, int rhs_haloed,
                     const double *duy_ym) {
  const int nx = c->nx, ny = c->ny, nz = c->nz;
  const size_t plane = (size_t)nx * ny;
  const double beta = (kind == CN_C) ? c->pmBetaSc : c->pmBeta;
  const double *am = (kind == CN_U) ? c->am2cCN : (kind == CN_V) ? c->am2p : c->am2cScCN;
  const double *ac = (kind == CN_U) ? c->ac2cCN : (kind == CN_V) ? c->ac2p : c->ac2cScCN;
  const double *ap = (kind == CN_U) ? c->ap2cCN : (kind == CN_V) ? c->ap2p : c->ap2cScCN;
#pragma omp parallel
  {
    double *tm = dalloc(plane), *tc = dalloc(plane), *tp = dalloc(plane), *tf = dalloc(plane);
    double *vecm = dalloc((size_t)nx), *arrmn = dalloc(plane);
#define A2(a, i, j) (a)[(size_t)((i)-1) + (size_t)nx * (size_t)((j)-1)]
    for (int jc = 1; jc <= ny; jc++)
      for (int ic = 1; ic <= nx; ic++) {
        if (kind == CN_V && jc == 1) {
          A2(tp, ic, 1) = 0.0;
          A2(tc, ic, 1) = 1.0;
          A2(tm, ic, 1) = 0.0;
        } else {
          A2(tp, ic, jc) = -beta * ap[jc];
          A2(tc, ic, jc) = -beta * ac[jc] + 1.0;
          A2(tm, ic, jc) = -beta * am[jc];
        }
      }
#pragma omp for schedule(static)
    for (int kc = 1; kc <= nz; kc++) {
      for (int jc = 1; jc <= ny; jc++)
        for (int ic = 1; ic <= nx; ic++) {
          if (kind == CN_V && jc == 1) A2(tf, ic, 1) = P2(c, duy_ym, ic, kc);
          else A2(tf, ic, jc) = rhs_haloed ? H(c, rhs, ic, jc, kc) : I3(c, rhs, ic, jc, kc);
        }
      thomas(tm, tc, tp, tf, nx, ny, vecm, arrmn);
      for (int jc = 1; jc <= ny; jc++)
        for (int ic = 1; ic <= nx; ic++) H(c, u, ic, jc, kc) = H(c, u, ic, jc, kc) + A2(tf, ic, jc);
    }
#undef A2
    free(tm); free(tc); free(tp); free(tf); free(vecm); free(arrmn);
  }
}

void orc_clc_u1hat(orc_ctx *c, double *ux, const double *rhsx) { cn_solve(c, CN_U, ux, rhsx, 0, NULL); }
void orc_clc_u2hat(orc_ctx *c, double *uy, const double *rhsy, const double *duy_ym) { cn_solve(c, CN_V, uy, rhsy, 0, duy_ym); }
void orc_clc_u3hat(orc_ctx *c, double *uz, const double *rhsz) { cn_solve(c, CN_U, uz, rhsz, 1, NULL); }
void orc_clc_scalar(orc_ctx *c, double *s, const double *rhsc) { cn_solve(c, CN_C, s, rhsc, 1, NULL); }

/* ------------------------------------------------------------------------ */
/* clcPrSrc :48-77, FluidVelUpdate :82-107, PressureUpdate :512-532          */
/* ------------------------------------------------------------------------ */
double orc_clc_prsrc(orc_ctx *c, const double *ux, const double *uy, const double *uz, double *prsrc) {
  const int nx = c->nx, ny = c->ny, nz = c->nz;
  const double sudtal = 1.0 / c->pmAlpha;
  double *pmax = dalloc((size_t)nz + 1);
#pragma omp parallel for schedule(static)
  for (int kc = 1; kc <= nz; kc++) {
    const int km = kc - 1, kp = kc + 1, ku = kc + 2;
    double dm = 0.0;
    for (int jc = 1; jc <= ny; jc++) {
      const int jp = jc + 1;
      const double sucaj = c->rdyp[jc];
      for (int ic = 1; ic <= nx; ic++) {
        const int im = ic - 1, ip = ic + 1, iu = ic + 2;
        double dudx = c->dxc1 * H(c, ux, im, jc, kc) + c->dxc2 * H(c, ux, ic, jc, kc) + c->dxc3 * H(c, ux, ip, jc, kc) + c->dxc4 * H(c, ux, iu, jc, kc);
        double dvdy = (H(c, uy, ic, jp, kc) - H(c, uy, ic, jc, kc)) * sucaj;
        double dwdz = c->dzc1 * H(c, uz, ic, jc, km) + c->dzc2 * H(c, uz, ic, jc, kc) + c->dzc3 * H(c, uz, ic, jc, kp) + c->dzc4 * H(c, uz, ic, jc, ku);
        double rdiv = dudx + dvdy + dwdz;
        double a = fabs(rdiv);
        dm = (a > dm) ? a : dm;
        I3(c, prsrc, ic, jc, kc) = sudtal * rdiv;
      }
    }
    pmax[kc] = dm;
  }
  double divmax = 0.0;
  for (int kc = 1; kc <= nz; kc++) divmax = pmax[kc] > divmax ? pmax[kc] : divmax;
  free(pmax);
  return divmax;
}

void orc_fluid_vel_update(orc_ctx *c, const double *phi, double *ux, double *uy, double *uz) {
  const int nx = c->nx, ny = c->ny, nz = c->nz;
  const double pmAlpha = c->pmAlpha;
#pragma omp parallel for schedule(static)
  for (int kc = 1; kc <= nz; kc++) {
    const int ks = kc - 2, km = kc - 1, kp = kc + 1;
    for (int jc = 1; jc <= ny; jc++) {
      const int jm = jc - 1;
      const double sucac = c->rdyc[jc];
      for (int ic = 1; ic <= nx; ic++) {
        const int is = ic - 2, im = ic - 1, ip = ic + 1;
        double dpdx = c->dxc1 * H(c, phi, is, jc, kc) + c->dxc2 * H(c, phi, im, jc, kc) + c->dxc3 * H(c, phi, ic, jc, kc) + c->dxc4 * H(c, phi, ip, jc, kc);
        double dpdy = (H(c, phi, ic, jc, kc) - H(c, phi, ic, jm, kc)) * sucac;
        double dpdz = c->dzc1 * H(c, phi, ic, jc, ks) + c->dzc2 * H(c, phi, ic, jc, km) + c->dzc3 * H(c, phi, ic, jc, kc) + c->dzc4 * H(c, phi, ic, jc, kp);
        H(c, ux, ic, jc, kc) = H(c, ux, ic, jc, kc) - pmAlpha * dpdx;
        H(c, uy, ic, jc, kc) = H(c, uy, ic, jc, kc) - pmAlpha * dpdy;
        H(c, uz, ic, jc, kc) = H(c, uz, ic, jc, kc) - pmAlpha * dpdz;
      }
    }
  }
}

void orc_pressure_update(orc_ctx *c, double *p, const double *phi) {
  const int nx = c->nx, ny = c->ny, nz = c->nz;
  const double beta = c->pmBeta;
#pragma omp parallel for schedule(static)
  for (int kc = 1; kc <= nz; kc++)
    for (int jc = 1; jc <= ny; jc++) {
      const int jp = jc + 1, jm = jc - 1;
      const double bp = -beta * c->ap2ph[jc], bc = -beta * c->ac2ph[jc] + 1.0, bm = -beta * c->am2ph[jc];
      for (int ic = 1; ic <= nx; ic++)
        H(c, p, ic, jc, kc) = H(c, p, ic, jc, kc) + bp * H(c, phi, ic, jp, kc) + bc * H(c, phi, ic, jc, kc) + bm * H(c, phi, ic, jm, kc);
    }
}

/* ------------------------------------------------------------------------ */
/* Boundary conditions: src/m_BC_and_Halo.f90:23-178; Update_uy_ym:          */
/* src/m_FlowCase.f90:112-122                                                 */
/* ------------------------------------------------------------------------ */
void orc_set_bc_and_update_halo(orc_ctx *c, double *ux, double *uy, double *uz, const double *uy_ym) {
  const int nx = c->nx, ny = c->ny, nz = c->nz, nyp = ny + 1;
  const double *xb = c->cfg.uxbc, *yb = c->cfg.uybc, *zb = c->cfg.uzbc;
  for (int kc = 1; kc <= nz; kc++)
    for (int ic = 1; ic <= nx; ic++) {
      if (c->cfg.flow_type == 1) {
        H(c, ux, ic, nyp, kc) = xb[3] * 2.0 - H(c, ux, ic, ny, kc);
        H(c, uy, ic, nyp, kc) = yb[3];
        H(c, uz, ic, nyp, kc) = zb[3] * 2.0 - H(c, uz, ic, ny, kc);
      } else if (c->cfg.flow_type == 2) {
        H(c, ux, ic, nyp, kc) = H(c, ux, ic, ny, kc);
        H(c, uy, ic, nyp, kc) = 0.0;
        H(c, uz, ic, nyp, kc) = H(c, uz, ic, ny, kc);
      }
    }
  for (int kc = 1; kc <= nz; kc++)
    for (int ic = 1; ic <= nx; ic++) {
      H(c, ux, ic, 0, kc) = xb[2] * 2.0 - H(c, ux, ic, 1, kc);
      H(c, uy, ic, 1, kc) = P2(c, uy_ym, ic, kc);
      H(c, uz, ic, 0, kc) = zb[2] * 2.0 - H(c, uz, ic, 1, kc);
    }
  orc_update_halo(c, ux, 2, 2, 2, 2);
  orc_update_halo(c, uy, 2, 2, 2, 2);
  orc_update_halo(c, uz, 2, 2, 2, 2);
}

void orc_set_bc_for_prsrc(orc_ctx *c, double *ux, double *uy, double *uz, const double *uy_ym) {
  const int nx = c->nx, ny = c->ny, nz = c->nz, nyp = ny + 1;
  for (int kc = 1; kc <= nz; kc++)
    for (int ic = 1; ic <= nx; ic++) H(c, uy, ic, nyp, kc) = (c->cfg.flow_type == 1) ? c->cfg.uybc[3] : 0.0;
  for (int kc = 1; kc <= nz; kc++)
    for (int ic = 1; ic <= nx; ic++) H(c, uy, ic, 1, kc) = P2(c, uy_ym, ic, kc);
  orc_update_halo(c, ux, 1, 2, 0, 0);
  orc_update_halo(c, uz, 0, 0, 1, 2);
}

void orc_set_bc_pr(orc_ctx *c, double *p) {
  const int nx = c->nx, ny = c->ny, nz = c->nz, nyp = ny + 1;
  for (int kc = 1; kc <= nz; kc++)
    for (int ic = 1; ic <= nx; ic++) {
      H(c, p, ic, 0, kc) = H(c, p, ic, 1, kc);
      H(c, p, ic, nyp, kc) = H(c, p, ic, ny, kc);
    }
  orc_update_halo(c, p, 2, 1, 2, 1);
}

void orc_set_bc_scalar(orc_ctx *c, double *s) {
  const int nx = c->nx, ny = c->ny, nz = c->nz, nyp = ny + 1;
  const double *v = c->cfg.sbc_val;
  double r1 = 0.0, r2 = 0.0;
  switch (c->cfg.sbc_opt[0]) {
    case -1: r1 = -1.0; r2 = 2.0 * v[0]; break;
    case -2: r1 = 1.0; r2 = -c->dyc[1] * v[2]; break;
    case -3:
      r1 = (2.0 + v[0] * c->dyc[1]) / (2.0 - v[0] * c->dyc[1]);
      r2 = -2.0 * v[2] * c->dyc[1] / (2.0 - v[0] * c->dyc[1]);
      break;
    default: break;
  }
  for (int kc = 1; kc <= nz; kc++)
    for (int ic = 1; ic <= nx; ic++) H(c, s, ic, 0, kc) = r1 * H(c, s, ic, 1, kc) + r2;
  switch (c->cfg.sbc_opt[1]) {
    case -1: r1 = -1.0; r2 = 2.0 * v[1]; break;
    case -2: r1 = 1.0; r2 = c->dyc[nyp] * v[3]; break;
    case -3:
      r1 = (2.0 - v[1] * c->dyc[nyp]) / (2.0 + v[1] * c->dyc[nyp]);
      r2 = 2.0 * v[3] * c->dyc[nyp] / (2.0 + v[1] * c->dyc[nyp]);
      break;
    default: r1 = 0.0; r2 = 0.0; break;
  }
  for (int kc = 1; kc <= nz; kc++)
    for (int ic = 1; ic <= nx; ic++) H(c, s, ic, nyp, kc) = r1 * H(c, s, ic, ny, kc) + r2;
  orc_update_halo(c, s, 2, 2, 2, 2);
}

void orc_update_uy_ym(orc_ctx *c, double *uy_ym, double *duy_ym) {
  const size_t n = (size_t)c->nx * c->nz;
  for (size_t q = 0; q < n; q++) duy_ym[q] = uy_ym[q];
  for (size_t q = 0; q < n; q++) uy_ym[q] = 0.0;
  for (size_t q = 0; q < n; q++) duy_ym[q] = uy_ym[q] - duy_ym[q];
}

/* ------------------------------------------------------------------------ */
/* Monitors: src/m_Tools.f90:24-110, 256-301                                 */
/* ------------------------------------------------------------------------ */
double orc_calc_max_cfl(const orc_ctx *c, const double *ux, const double *uy, const double *uz) {
  double m = 0.0;
  for (int kc = 1; kc <= c->nz; kc++)
    for (int jc = 1; jc <= c->ny; jc++)
      for (int ic = 1; ic <= c->nx; ic++) {
        double v = fabs(H(c, ux, ic, jc, kc) + H(c, ux, ic + 1, jc, kc)) * c->rdx
                 + fabs(H(c, uy, ic, jc, kc) + H(c, uy, ic, jc + 1, kc)) * c->rdyp[jc]
                 + fabs(H(c, uz, ic, jc, kc) + H(c, uz, ic, jc, kc + 1)) * c->rdz;
        if (v > m) m = v;
      }
  return m * 0.5;
}

double orc_check_divergence(const orc_ctx *c, const double *ux, const double *uy, const double *uz) {
  double m = 0.0;
  for (int kc = 1; kc <= c->nz; kc++)
    for (int jc = 1; jc <= c->ny; jc++)
      for (int ic = 1; ic <= c->nx; ic++) {
        double dudx = c->dxc1 * H(c, ux, ic - 1, jc, kc) + c->dxc2 * H(c, ux, ic, jc, kc) + c->dxc3 * H(c, ux, ic + 1, jc, kc) + c->dxc4 * H(c, ux, ic + 2, jc, kc);
        double dvdy = (H(c, uy, ic, jc + 1, kc) - H(c, uy, ic, jc, kc)) * c->rdyp[jc];
        double dwdz = c->dzc1 * H(c, uz, ic, jc, kc - 1) + c->dzc2 * H(c, uz, ic, jc, kc) + c->dzc3 * H(c, uz, ic, jc, kc + 1) + c->dzc4 * H(c, uz, ic, jc, kc + 2);
        double d = fabs(dudx + dvdy + dwdz);
        if (d > m) m = d;
      }
  return m;
}

void orc_calc_vmax(const orc_ctx *c, const double *ux, const double *uy, const double *uz, double *v) {
  v[0] = v[1] = v[2] = 0.0;
  for (int kc = 1; kc <= c->nz; kc++)
    for (int jc = 1; jc <= c->ny; jc++)
      for (int ic = 1; ic <= c->nx; ic++) {
        double a = fabs(H(c, ux, ic, jc, kc)); if (a > v[0]) v[0] = a;
        a = fabs(H(c, uy, ic, jc, kc)); if (a > v[1]) v[1] = a;
        a = fabs(H(c, uz, ic, jc, kc)); if (a > v[2]) v[2] = a;
      }
}

double orc_calc_ux_aver(const orc_ctx *c, const double *ux) {
  double r = 0.0;
  for (int kc = 1; kc <= c->nz; kc++)
    for (int jc = 1; jc <= c->ny; jc++)
      for (int ic = 1; ic <= c->nx; ic++) r = r + H(c, ux, ic, jc, kc) * c->dyp[jc];
  return r / (double)(c->nx * c->nz) / c->cfg.yly;
}

void orc_clc_mean_value(const orc_ctx *c, const double *ux, const double *s, double *res) {
  res[0] = res[1] = res[2] = 0.0;
  for (int kc = 1; kc <= c->nz; kc++)
    for (int jc = 1; jc <= c->ny; jc++) {
      const double d = c->dyp[jc];
      for (int ic = 1; ic <= c->nx; ic++) {
        const double v = H(c, s, ic, jc, kc);
        res[0] = res[0] + d * H(c, ux, ic, jc, kc);
        res[1] = res[1] + d * v;
        res[2] = res[2] + d * v * v;
      }
    }
  for (int q = 0; q < 3; q++) res[q] = res[q] / (double)(c->nx * c->nz) / c->cfg.yly;
}

/* ------------------------------------------------------------------------ */
/* clcStat: src/m_FlowCase.f90:283-413                                        */
/* ------------------------------------------------------------------------ */
void orc_clc_stat(const orc_ctx *c, const double *ux, const double *uy, const double *uz, const double *pr,
                  const double *sc, double *sumstat) {
  const int nx = c->nx, ny = c->ny, nz = c->nz, nyp = ny + 1;
  const int ns = c->cfg.scalar ? 40 : 35;
  const double uCRF = c->cfg.ucrf, half = 0.5;
  const double IC1 = c->ic1, IC2 = c->ic2, IC3 = c->ic3, IC4 = c->ic4;
  const double rdxt = c->rdx / 12.0, rdzt = c->rdz / 12.0;
  const double inxz = 1.0 / ((double)nx * (double)nz);
  double sv[40];
  for (int jc = 1; jc <= ny; jc++) {
    const int jp = jc + 1, jm = jc - 1;
    const double InterpY1 = c->yint[jc], InterpY2 = 1.0 - InterpY1;
    const double cac = c->rdyc[jc], cacU = c->rdyc[jp], caj = c->rdyp[jc];
    for (int q = 0; q < 40; q++) sv[q] = 0.0;
    for (int kc = 1; kc <= nz; kc++) {
      const int ks = kc - 2, km = kc - 1, kp = kc + 1, ku = kc + 2;
      for (int ic = 1; ic <= nx; ic++) {
        const int is = ic - 2, im = ic - 1, ip = ic + 1, iu = ic + 2;
        const double uxloc = H(c, ux, ic, jc, kc) + uCRF;
        const double uyloc = H(c, uy, ic, jc, kc);
        const double uzloc = H(c, uz, ic, jc, kc);
        const double prloc = H(c, pr, ic, jc, kc);
        const double uxCell = IC1 * H(c, ux, im, jc, kc) + IC2 * H(c, ux, ic, jc, kc) + IC3 * H(c, ux, ip, jc, kc) + IC4 * H(c, ux, iu, jc, kc) + uCRF;
        const double uyCell = (uyloc + H(c, uy, ic, jp, kc)) * half;
        const double uy_xc = IC1 * H(c, uy, is, jc, kc) + IC2 * H(c, uy, im, jc, kc) + IC3 * H(c, uy, ic, jc, kc) + IC4 * H(c, uy, ip, jc, kc);
        const double uy_xp = IC1 * H(c, uy, is, jp, kc) + IC2 * H(c, uy, im, jp, kc) + IC3 * H(c, uy, ic, jp, kc) + IC4 * H(c, uy, ip, jp, kc);
        const double uzCell = IC1 * H(c, uz, ic, jc, km) + IC2 * H(c, uz, ic, jc, kc) + IC3 * H(c, uz, ic, jc, kp) + IC4 * H(c, uz, ic, jc, ku);
        const double prloc2 = InterpY1 * (IC1 * H(c, pr, is, jm, kc) + IC2 * H(c, pr, im, jm, kc) + IC3 * H(c, pr, ic, jm, kc) + IC4 * H(c, pr, ip, jm, kc))
                            + InterpY2 * (IC1 * H(c, pr, is, jc, kc) + IC2 * H(c, pr, im, jc, kc) + IC3 * H(c, pr, ic, jc, kc) + IC4 * H(c, pr, ip, jc, kc));
        const double dudx = c->dxc1 * H(c, ux, im, jc, kc) + c->dxc2 * H(c, ux, ic, jc, kc) + c->dxc3 * H(c, ux, ip, jc, kc) + c->dxc4 * H(c, ux, iu, jc, kc);
        const double dudy = (H(c, ux, ic, jc, kc) - H(c, ux, ic, jm, kc)) * cac;
        const double dudz = c->dzc1 * H(c, ux, ic, jc, ks) + c->dzc2 * H(c, ux, ic, jc, km) + c->dzc3 * H(c, ux, ic, jc, kc) + c->dzc4 * H(c, ux, ic, jc, kp);
        const double dvdx = c->dxc1 * H(c, uy, is, jc, kc) + c->dxc2 * H(c, uy, im, jc, kc) + c->dxc3 * H(c, uy, ic, jc, kc) + c->dxc4 * H(c, uy, ip, jc, kc);
        const double dvdy = (H(c, uy, ic, jp, kc) - uyloc) * caj;
        const double dvdz = c->dzc1 * H(c, uy, ic, jc, ks) + c->dzc2 * H(c, uy, ic, jc, km) + c->dzc3 * H(c, uy, ic, jc, kc) + c->dzc4 * H(c, uy, ic, jc, kp);
        const double dwdx = c->dxc1 * H(c, uz, is, jc, kc) + c->dxc2 * H(c, uz, im, jc, kc) + c->dxc3 * H(c, uz, ic, jc, kc) + c->dxc4 * H(c, uz, ip, jc, kc);
        const double dwdy = (uzloc - H(c, uz, ic, jm, kc)) * cac;
        const double dwdz = c->dzc1 * H(c, uz, ic, jc, km) + c->dzc2 * H(c, uz, ic, jc, kc) + c->dzc3 * H(c, uz, ic, jc, kp) + c->dzc4 * H(c, uz, ic, jc, ku);
        const double dudxx = c->dxxc1 * H(c, ux, is, jc, kc) + c->dxxc2 * H(c, ux, im, jc, kc) + c->dxxc3 * H(c, ux, ic, jc, kc) + c->dxxc4 * H(c, ux, ip, jc, kc) + c->dxxc5 * H(c, ux, iu, jc, kc);
        const double dudyy = c->ap2c[jc] * H(c, ux, ic, jp, kc) + c->ac2c[jc] * H(c, ux, ic, jc, kc) + c->am2c[jc] * H(c, ux, ic, jm, kc);
        const double dudzz = c->dzzc1 * H(c, ux, ic, jc, ks) + c->dzzc2 * H(c, ux, ic, jc, km) + c->dzzc3 * H(c, ux, ic, jc, kc) + c->dzzc4 * H(c, ux, ic, jc, kp) + c->dzzc5 * H(c, ux, ic, jc, ku);
        const double dvdxx = c->dxxc1 * H(c, uy, is, jc, kc) + c->dxxc2 * H(c, uy, im, jc, kc) + c->dxxc3 * H(c, uy, ic, jc, kc) + c->dxxc4 * H(c, uy, ip, jc, kc) + c->dxxc5 * H(c, uy, iu, jc, kc);
        const double dvdyy = c->ap2p[jc] * H(c, uy, ic, jp, kc) + c->ac2p[jc] * H(c, uy, ic, jc, kc) + c->am2p[jc] * H(c, uy, ic, jm, kc);
        const double dvdzz = c->dzzc1 * H(c, uy, ic, jc, ks) + c->dzzc2 * H(c, uy, ic, jc, km) + c->dzzc3 * H(c, uy, ic, jc, kc) + c->dzzc4 * H(c, uy, ic, jc, kp) + c->dzzc5 * H(c, uy, ic, jc, ku);
        const double dwdxx = c->dxxc1 * H(c, uz, is, jc, kc) + c->dxxc2 * H(c, uz, im, jc, kc) + c->dxxc3 * H(c, uz, ic, jc, kc) + c->dxxc4 * H(c, uz, ip, jc, kc) + c->dxxc5 * H(c, uz, iu, jc, kc);
        const double dwdyy = c->ap2c[jc] * H(c, uz, ic, jp, kc) + c->ac2c[jc] * H(c, uz, ic, jc, kc) + c->am2c[jc] * H(c, uz, ic, jm, kc);
        const double dwdzz = c->dzzc1 * H(c, uz, ic, jc, ks) + c->dzzc2 * H(c, uz, ic, jc, km) + c->dzzc3 * H(c, uz, ic, jc, kc) + c->dzzc4 * H(c, uz, ic, jc, kp) + c->dzzc5 * H(c, uz, ic, jc, ku);
        const double vor_x = dwdy - dvdz, vor_y = dudz - dwdx, vor_z = dvdx - dudy;
        const double dudxC = (-H(c, ux, iu, jc, kc) + 8.0 * H(c, ux, ip, jc, kc) - 8.0 * H(c, ux, im, jc, kc) + H(c, ux, is, jc, kc)) * rdxt;
        const double dvdxU = c->dxc1 * H(c, uy, is, jp, kc) + c->dxc2 * H(c, uy, im, jp, kc) + c->dxc3 * H(c, uy, ic, jp, kc) + c->dxc4 * H(c, uy, ip, jp, kc);
        const double dudyU = (H(c, ux, ic, jp, kc) - H(c, ux, ic, jc, kc)) * cacU;
        const double dudzC = (InterpY1 * (-H(c, ux, ic, jm, ku) + 8.0 * H(c, ux, ic, jm, kp) - 8.0 * H(c, ux, ic, jm, km) + H(c, ux, ic, jm, ks))
                            + InterpY2 * (-H(c, ux, ic, jc, ku) + 8.0 * H(c, ux, ic, jc, kp) - 8.0 * H(c, ux, ic, jc, km) + H(c, ux, ic, jc, ks))) * rdzt;
        const double dvdzC = ((-H(c, uy, is, jc, ku) + 8.0 * H(c, uy, is, jc, kp) - 8.0 * H(c, uy, is, jc, km) + H(c, uy, is, jc, ks)) * IC1
                            + (-H(c, uy, im, jc, ku) + 8.0 * H(c, uy, im, jc, kp) - 8.0 * H(c, uy, im, jc, km) + H(c, uy, im, jc, ks)) * IC2
                            + (-H(c, uy, ic, jc, ku) + 8.0 * H(c, uy, ic, jc, kp) - 8.0 * H(c, uy, ic, jc, km) + H(c, uy, ic, jc, ks)) * IC3
                            + (-H(c, uy, ip, jc, ku) + 8.0 * H(c, uy, ip, jc, kp) - 8.0 * H(c, uy, ip, jc, km) + H(c, uy, ip, jc, ks)) * IC4) * rdzt;
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
        sv[30] += (dudxC * (dvdxU + dvdx) + (dudyU + dudy) * (uy_xp - uy_xc) * caj) * half;
        sv[31] += dudzC * dvdzC;
        sv[32] += vor_x * vor_x;
        sv[33] += vor_y * vor_y;
        sv[34] += vor_z * vor_z;
        if (ns == 40) {
          const double scloc = H(c, sc, ic, jc, kc);
          sv[35] += scloc;
          sv[36] += scloc * scloc;
          sv[37] += scloc * uxCell;
          sv[38] += scloc * uyCell;
          sv[39] += scloc * uzCell;
        }
      }
    }
    for (int q = 0; q < ns; q++) sumstat[q + (size_t)ns * (jc - 1)] += sv[q] * inxz;
  }
  /* nyp only (:393-413) */
  {
    const int jc = nyp, jm = jc - 1;
    const double cac = c->rdyc[jc];
    const double InterpY1 = c->yint[jc], InterpY2 = 1.0 - InterpY1;
    for (int q = 0; q < 40; q++) sv[q] = 0.0;
    for (int kc = 1; kc <= nz; kc++)
      for (int ic = 1; ic <= nx; ic++) {
        const int is = ic - 2, im = ic - 1, ip = ic + 1;
        const double prloc2 = InterpY1 * (IC1 * H(c, pr, is, jm, kc) + IC2 * H(c, pr, im, jm, kc) + IC3 * H(c, pr, ic, jm, kc) + IC4 * H(c, pr, ip, jm, kc))
                            + InterpY2 * (IC1 * H(c, pr, is, jc, kc) + IC2 * H(c, pr, im, jc, kc) + IC3 * H(c, pr, ic, jc, kc) + IC4 * H(c, pr, ip, jc, kc));
        const double dudy = (H(c, ux, ic, jc, kc) - H(c, ux, ic, jm, kc)) * cac;
        const double dwdy = (H(c, uz, ic, jc, kc) - H(c, uz, ic, jm, kc)) * cac;
        const double vor_x = dwdy, vor_z = -dudy;
        sv[26] += prloc2 * (dudy + 0.0);
        sv[32] += vor_x * vor_x;
        sv[34] += vor_z * vor_z;
      }
    for (int q = 0; q < ns; q++) sumstat[q + (size_t)ns * (jc - 1)] += sv[q] * inxz;
  }
}

/* ------------------------------------------------------------------------ */
/* ChannelIterate: src/m_ChannelSystem.f90:80-145                            */
/* ------------------------------------------------------------------------ */
void orc_channel_iterate(orc_ctx *c, int poisson_choice,
                         double *ux, double *uy, double *uz, double *p, double *scalar,
                         double *histx, double *histy, double *histz, double *histc,
                         double *arr1, double *arr2, double *halo, double *arrc,
                         double *uy_ym, double *duy_ym, double *mon) {
  double uddxmax = orc_calc_max_cfl(c, ux, uy, uz);
  double cflmp = uddxmax * c->dt;
  double dt;
  if (c->cfg.icfl == 1) {
    dt = c->cfg.cflc / uddxmax;
    dt = dt < c->cfg.dtmax ? dt : c->cfg.dtmax;
  } else {
    dt = c->cfg.dtmax;
  }
  double divmax1 = 0.0;
  for (int ns = 1; ns <= c->iadvance; ns++) {
    orc_pm_coe_update(c, ns, dt);
    orc_update_uy_ym(c, uy_ym, duy_ym);
    orc_clc_rhs_xyz(c, ux, uy, uz, scalar, arr1, arr2, halo, arrc, histx, histy, histz, histc, p);
    if (c->cfg.scalar) orc_clc_scalar(c, scalar, arrc);
    orc_clc_u1hat(c, ux, arr1);
    orc_clc_u2hat(c, uy, arr2, duy_ym);
    orc_clc_u3hat(c, uz, halo);
    orc_set_bc_for_prsrc(c, ux, uy, uz, uy_ym);
    divmax1 = orc_clc_prsrc(c, ux, uy, uz, arr1);
    orc_poisson(c, poisson_choice, arr1, halo);
    orc_set_bc_pr(c, halo);
    orc_fluid_vel_update(c, halo, ux, uy, uz);
    orc_pressure_update(c, p, halo);
    orc_set_bc_and_update_halo(c, ux, uy, uz, uy_ym);
    orc_set_bc_pr(c, p);
    if (c->cfg.scalar) orc_set_bc_scalar(c, scalar);
  }
  mon[0] = divmax1;
  mon[1] = orc_check_divergence(c, ux, uy, uz);
  orc_calc_vmax(c, ux, uy, uz, mon + 2);
  mon[5] = cflmp;
  mon[6] = dt;
}

/* ---- dump_visu (src/m_IOAndVisu.f90:182-385), one rank, cell-centre form --
 * The field `which` (0 ux, 1 uy, 2 uz, 3 pr, 4 wx, 5 wy, 6 wz, 7 wMag,
 * 8 scalar) at the cell centres that decomp_2d_write_every keeps with
 * from1 = .true. (src/m_Decomp2d.f90:1232-1389): i = 1, 1+iskip, ... (and the
 * same in j and k), written x fastest.  The velocity, pressure and scalar
 * arrays are haloed (x/z halos current, y ghost rows set). */
/* CubicRoot (src/m_Tools.f90:494-570): the roots of x^3 + b x^2 + c x + d.
 * Returns 1 with root[] in the reference's order when they are all real (for
 * three distinct roots root[1] is the middle one), 0 for a complex pair. */
static int cubic_root(double b, double c, double d, double root[3]) {
  const double SMALL = 1.0e-12, third = 0.3333333333333333333333, pi = 3.141592653589793238462643383279502884;
  const double p = c - third * b * b;
  const double q = (2.0 * b * b * b - 9.0 * b * c + 27.0 * d) / 27.0;
  double t = q * q / 4.0 + p * p * p / 27.0;
  if (t < 0.0) {
    t = -third * b;
    double rho = pow(-p, 1.5) / sqrt(27.0);
    const double eta = acos(0.5 * q / rho);
    rho = -2.0 * pow(rho, third);
    const double r1 = cos(third * eta) * rho + t;
    const double r2 = cos(third * (2.0 * pi + eta)) * rho + t;
    const double r3 = cos(third * (4.0 * pi + eta)) * rho + t;
    double o[3];
    if (r1 < r2 && r1 < r3) {
      o[0] = r1; o[1] = r2 < r3 ? r2 : r3; o[2] = r2 < r3 ? r3 : r2;
    } else if (r1 > r2 && r1 > r3) {
      o[0] = r2 < r3 ? r2 : r3; o[1] = r2 < r3 ? r3 : r2; o[2] = r1;
    } else {
      o[0] = r2 < r3 ? r2 : r3; o[1] = r1; o[2] = r2 < r3 ? r3 : r2;
    }
    root[0] = o[0]; root[1] = o[1]; root[2] = o[2];
    return 1;
  }
  if (t < SMALL) {
    double rho = 0.5 * q;
    rho = rho < 0.0 ? pow(-rho, third) : -pow(rho, third);
    root[0] = 2.0 * rho - third * b;
    root[1] = -rho - third * b;
    root[2] = root[1];
    return 1;
  }
  return 0;
}

static double visu_point(const orc_ctx *c, int which, const double *ux, const double *uy, const double *uz,
                         const double *pr, const double *sc, int ic, int jc, int kc) {
  const double half = 0.5, quarter = 0.25;
  const int ip = ic + 1, im = ic - 1, jp = jc + 1, jm = jc - 1, kp = kc + 1, km = kc - 1;
  const double rdx = c->rdx, rdz = c->rdz;
  const double cac1 = c->rdyc[jc], cac2 = c->rdyc[jp], cac12 = cac1 - cac2;
  switch (which) {
    case 0: return half * (H(c, ux, ip, jc, kc) + H(c, ux, ic, jc, kc)) + c->cfg.ucrf; /* :199-214 */
    case 1: return half * (H(c, uy, ic, jp, kc) + H(c, uy, ic, jc, kc));               /* :216-231 */
    case 2: return half * (H(c, uz, ic, jc, kp) + H(c, uz, ic, jc, kc));               /* :233-248 */
    case 3: return H(c, pr, ic, jc, kc);                                               /* :250-254 */
    case 8: return H(c, sc, ic, jc, kc);                                               /* :257-260 */
    default: break;
  }
  /* :263-369 */
  const double dudy = ((H(c, ux, ip, jp, kc) + H(c, ux, ic, jp, kc)) * cac2 +
                       (H(c, ux, ip, jc, kc) + H(c, ux, ic, jc, kc)) * cac12 -
                       (H(c, ux, ip, jm, kc) + H(c, ux, ic, jm, kc)) * cac1) * quarter;
  const double dudz = (H(c, ux, ip, jc, kp) + H(c, ux, ic, jc, kp) - H(c, ux, ip, jc, km) - H(c, ux, ic, jc, km)) * rdz * quarter;
  const double dvdx = (H(c, uy, ip, jp, kc) - H(c, uy, im, jp, kc) + H(c, uy, ip, jc, kc) - H(c, uy, im, jc, kc)) * rdx * quarter;
  const double dvdz = (H(c, uy, ic, jp, kp) + H(c, uy, ic, jc, kp) - H(c, uy, ic, jp, km) - H(c, uy, ic, jc, km)) * rdz * quarter;
  const double dwdx = (H(c, uz, ip, jc, kp) - H(c, uz, im, jc, kp) + H(c, uz, ip, jc, kc) - H(c, uz, im, jc, kc)) * rdx * quarter;
  const double dwdy = ((H(c, uz, ic, jp, kp) + H(c, uz, ic, jp, kc)) * cac2 +
                       (H(c, uz, ic, jc, kp) + H(c, uz, ic, jc, kc)) * cac12 -
                       (H(c, uz, ic, jm, kp) + H(c, uz, ic, jm, kc)) * cac1) * quarter;
  const double vor_x = dwdy - dvdz, vor_y = dudz - dwdx, vor_z = dvdx - dudy;
  if (which == 9 || which == 10) {
    const double dudx = (H(c, ux, ip, jc, kc) - H(c, ux, ic, jc, kc)) * rdx;
    const double dvdy = (H(c, uy, ic, jp, kc) - H(c, uy, ic, jc, kc)) * c->rdyp[jc];
    const double dwdz = (H(c, uz, ic, jc, kp) - H(c, uz, ic, jc, kc)) * rdz;
    if (which == 9) { /* Clc_Q_vor, src/m_Tools.f90:307-342 */
      const double udiv = dudx + dvdy + dwdz;
      return half * (udiv * udiv - dudx * dudx - dvdy * dvdy - dwdz * dwdz) - dudy * dvdx - dudz * dwdx - dvdz * dwdy;
    }
    /* Clc_lamda2, src/m_Tools.f90:347-442: the tensors carry a factor 2, so
     * the root is scaled by a quarter; a complex pair (the reference aborts
     * with "lamda2 wrong") gives NaN here */
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
    double root[3];
    return cubic_root(rb, rc, rd, root) ? quarter * root[1] : NAN;
  }
  switch (which) {
    case 4: return vor_x;
    case 5: return vor_y;
    case 6: return vor_z;
    default: return sqrt(vor_x * vor_x + vor_y * vor_y + vor_z * vor_z);
  }
}

long long orc_visu_count(const orc_ctx *c, int iskip, int jskip, int kskip) {
  const long long nxo = (c->nx + iskip - 1) / iskip, nyo = (c->ny + jskip - 1) / jskip, nzo = (c->nz + kskip - 1) / kskip;
  return nxo * nyo * nzo;
}

void orc_dump_visu(const orc_ctx *c, int which, const double *ux, const double *uy, const double *uz, const double *pr,
                   const double *sc, int iskip, int jskip, int kskip, double *out) {
  const int nxo = (c->nx + iskip - 1) / iskip, nyo = (c->ny + jskip - 1) / jskip, nzo = (c->nz + kskip - 1) / kskip;
  size_t q = 0;
  for (int k = 0; k < nzo; k++)
    for (int j = 0; j < nyo; j++)
      for (int i = 0; i < nxo; i++)
        out[q++] = visu_point(c, which, ux, uy, uz, pr, sc, 1 + i * iskip, 1 + j * jskip, 1 + k * kskip);
}

/* ---- energy spectra (clcStat, src/m_FlowCase.f90:430-805; helpers :954-1375) --
 * One rank holds the whole field, so every pencil is the full (nx, ny, nz)
 * box and the transposes are copies.  Work arrays are interior boxes indexed
 * B(i, j, k) = i + nx*(j + ny*k), 0-based.  R2HC lines go through the
 * oracle's own FFTW-semantics transform. */
#define B(i, j, k) ((size_t)(i) + (size_t)nx * ((size_t)(j) + (size_t)ny * (size_t)(k)))

static double *spec_box(const orc_ctx *c) { return (double *)calloc((size_t)c->nx * c->ny * c->nz, sizeof(double)); }

/* interior copy of a haloed array plus a constant (transpose_y1_to_x1 + arrFFT + uCRF) */
static void spec_copy(const orc_ctx *c, const double *a, double add, double *o) {
  const int nx = c->nx, ny = c->ny, nz = c->nz;
  for (int k = 0; k < nz; k++)
    for (int j = 0; j < ny; j++)
      for (int i = 0; i < nx; i++) o[B(i, j, k)] = H(c, a, i + 1, j + 1, k + 1) + add;
}

/* ux at the cell centre + uCRF (:433-444) */
static void spec_ux_cell(const orc_ctx *c, const double *ux, double *o) {
  const int nx = c->nx, ny = c->ny, nz = c->nz;
  for (int k = 0; k < nz; k++)
    for (int j = 0; j < ny; j++)
      for (int i = 0; i < nx; i++) {
        const int ic = i + 1, jc = j + 1, kc = k + 1;
        o[B(i, j, k)] = c->ic1 * H(c, ux, ic - 1, jc, kc) + c->ic2 * H(c, ux, ic, jc, kc) +
                        c->ic3 * H(c, ux, ic + 1, jc, kc) + c->ic4 * H(c, ux, ic + 2, jc, kc) + c->cfg.ucrf;
      }
}

/* uy at the cell centre (:576-587) */
static void spec_uy_cell(const orc_ctx *c, const double *uy, double *o) {
  const int nx = c->nx, ny = c->ny, nz = c->nz;
  for (int k = 0; k < nz; k++)
    for (int j = 0; j < ny; j++)
      for (int i = 0; i < nx; i++)
        o[B(i, j, k)] = 0.5 * (H(c, uy, i + 1, j + 1, k + 1) + H(c, uy, i + 1, j + 2, k + 1));
}

static void spec_fft_x(const orc_ctx *c, double *a) { orc_r2hc(c->nx, c->ny * c->nz, 1, c->nx, a); }
static void spec_fft_z(const orc_ctx *c, double *a) { orc_r2hc(c->nz, c->nx * c->ny, c->nx * c->ny, 1, a); }

/* The 1-D accumulators of one quantity m (1-based as the reference):
 * kind 0 = energy / cospectrum (clcEnergySpectraX, clcCospectraX and their z
 * forms, :954-980, :1040-1066, :1071-1097, :1157-1183), 1 = LCS real part
 * (:985-1010, :1102-1127), 2 = LCS imaginary part (:1015-1035, :1132-1152).
 * a and b are the transformed boxes (b = a for an energy spectrum). */
static void spec_acc_x(const orc_ctx *c, const double *a, const double *b, int kind, int m, double *specx) {
  const int nx = c->nx, ny = c->ny, nz = c->nz, nxh = nx / 2, nxhp = nxh + 1;
  const double norm = 1.0 / ((double)nx * (double)nx * (double)nz);
  double *egy = (double *)malloc(sizeof(double) * (size_t)nxhp);
  for (int j = 0; j < ny; j++) {
    for (int q = 0; q < nxhp; q++) egy[q] = 0.0;
    for (int k = 0; k < nz; k++) {
      if (kind == 0) {
        egy[0] += a[B(0, j, k)] * b[B(0, j, k)];
        egy[nxhp - 1] += a[B(nxhp - 1, j, k)] * b[B(nxhp - 1, j, k)];
      } else if (kind == 1) {
        egy[0] += a[B(0, j, k)];
        egy[nxhp - 1] += a[B(nxhp - 1, j, k)];
      } else {
        egy[0] = 0.0;
        egy[nxhp - 1] = 0.0;
      }
      for (int ic = 2; ic <= nxh; ic++) {
        const int cn = nx + 2 - ic;  /* 1-based counterpart */
        const size_t p = B(ic - 1, j, k), q = B(cn - 1, j, k);
        if (kind == 0) egy[ic - 1] += (a[p] * b[p] + a[q] * b[q]) * 2.0;
        else if (kind == 1) egy[ic - 1] += a[p] + a[q];
        else egy[ic - 1] += a[p] - a[q];
      }
    }
    for (int q = 0; q < nxhp; q++) specx[(size_t)q + (size_t)nxhp * ((size_t)j + (size_t)ny * (size_t)(m - 1))] += egy[q] * norm;
  }
  free(egy);
}

static void spec_acc_z(const orc_ctx *c, const double *a, const double *b, int kind, int m, double *specz) {
  const int nx = c->nx, ny = c->ny, nz = c->nz, nzh = nz / 2, nzhp = nzh + 1;
  const double norm = 1.0 / ((double)nx * (double)nz * (double)nz);
  double *egy = (double *)malloc(sizeof(double) * (size_t)nzhp);
  for (int j = 0; j < ny; j++) {
    for (int q = 0; q < nzhp; q++) egy[q] = 0.0;
    for (int i = 0; i < nx; i++) {
      if (kind == 0) {
        egy[0] += a[B(i, j, 0)] * b[B(i, j, 0)];
        egy[nzhp - 1] += a[B(i, j, nzhp - 1)] * b[B(i, j, nzhp - 1)];
      } else if (kind == 1) {
        egy[0] += a[B(i, j, 0)];
        egy[nzhp - 1] += a[B(i, j, nzhp - 1)];
      }
      for (int kc = 2; kc <= nzh; kc++) {
        const int cn = nz + 2 - kc;
        const size_t p = B(i, j, kc - 1), q = B(i, j, cn - 1);
        if (kind == 0) egy[kc - 1] += (a[p] * b[p] + a[q] * b[q]) * 2.0;
        else if (kind == 1) egy[kc - 1] += a[p] + a[q];
        else egy[kc - 1] += a[p] - a[q];
      }
    }
    for (int q = 0; q < nzhp; q++) specz[(size_t)q + (size_t)nzhp * ((size_t)j + (size_t)ny * (size_t)(m - 1))] += egy[q] * norm;
  }
  free(egy);
}

/* LCS products (:470-549 for x, :632-711 for z): the transformed ux box a (or
 * its real/imaginary swap when imag) times the reference row's values at the
 * same position; a channel's upper half uses the mirrored row. */
static void spec_lcs(const orc_ctx *c, const double *a, int jlcs, int dir, int imag, double *o) {
  const int nx = c->nx, ny = c->ny, nz = c->nz;
  const int jr1 = jlcs - 1, jr2 = c->cfg.flow_type == 1 ? ny - jlcs : jlcs - 1;
  for (int k = 0; k < nz; k++)
    for (int j = 0; j < ny; j++) {
      const int jr = (c->cfg.flow_type == 1 && j + 1 > ny / 2) ? jr2 : jr1;
      for (int i = 0; i < nx; i++) {
        int is = i, ks = k;
        if (imag && dir == 0 && i > 0) is = nx - i;  /* arrx2(ic) = arrx1(nxc+2-ic) */
        if (imag && dir == 1 && k > 0) ks = nz - k;  /* arrz2(kc) = arrz1(nzc+2-kc) */
        o[B(i, j, k)] = a[B(is, j, ks)] * a[B(i, jr, k)];
      }
    }
}

int orc_spectra_count(const orc_ctx *c, int which) {
  const int ne1 = c->cfg.scalar ? 9 : 8;
  if (which == 0) return (c->nx / 2 + 1) * c->ny * ne1;
  if (which == 1) return (c->nz / 2 + 1) * c->ny * ne1;
  return c->cfg.scalar ? 6 : 5;
}

void orc_spectra_1d(const orc_ctx *c, const double *ux, const double *uy, const double *uz, const double *pr,
                    const double *sc, int jlcs, double *specx, double *specz) {
  double *a1 = spec_box(c), *a2 = spec_box(c), *t = spec_box(c);
  for (int dir = 0; dir < 2; dir++) {
    void (*fft)(const orc_ctx *, double *) = dir == 0 ? spec_fft_x : spec_fft_z;
    void (*acc)(const orc_ctx *, const double *, const double *, int, int, double *) = dir == 0 ? spec_acc_x : spec_acc_z;
    double *spec = dir == 0 ? specx : specz;
    spec_copy(c, ux, c->cfg.ucrf, a1);
    fft(c, a1);
    acc(c, a1, a1, 0, 1, spec);
    spec_lcs(c, a1, jlcs, dir, 0, t);
    acc(c, t, t, 1, 7, spec);
    spec_lcs(c, a1, jlcs, dir, 1, t);
    acc(c, t, t, 2, 8, spec);
    const double *f3[3] = {uy, uz, pr};
    for (int q = 0; q < 3; q++) {
      spec_copy(c, f3[q], 0.0, a2);
      fft(c, a2);
      acc(c, a2, a2, 0, 2 + q, spec);
    }
    spec_uy_cell(c, uy, a2);
    fft(c, a2);
    acc(c, a1, a2, 0, 5, spec);
    spec_ux_cell(c, ux, a1);
    fft(c, a1);
    acc(c, a1, a2, 0, 6, spec);
    if (c->cfg.scalar) {
      spec_copy(c, sc, 0.0, a2);
      fft(c, a2);
      acc(c, a2, a2, 0, 9, spec);
    }
  }
  free(a1);
  free(a2);
  free(t);
}

/* clcEnergySpectra2D / clcCosSpectra2D (:1231-1375): R2HC along x, then along
 * z, of the two boxes; the products at the chosen rows (and, in a channel,
 * their mirrors) go to spec2d(nx, ny2d, nz, NEnergySpec2D). */
static void spec_acc_2d(const orc_ctx *c, const double *a, const double *b, int m, int ny2d, const int *iy2d,
                        int is_uy, double rcoe, double *spec2d) {
  const int nx = c->nx, ny = c->ny, nz = c->nz;
  double norm = 1.0 / ((double)nx * (double)nz);
  norm = norm * norm;
  if (c->cfg.flow_type == 1) norm = norm * 0.5;
  for (int jc = 0; jc < ny2d; jc++) {
    const int jct = iy2d[jc], jcs = is_uy ? ny + 2 - jct : ny + 1 - jct;
    for (int k = 0; k < nz; k++)
      for (int i = 0; i < nx; i++) {
        const double t1 = a[B(i, jct - 1, k)] * b[B(i, jct - 1, k)];
        double v = t1;
        if (c->cfg.flow_type == 1) v = t1 + rcoe * (a[B(i, jcs - 1, k)] * b[B(i, jcs - 1, k)]);
        spec2d[(size_t)i + (size_t)nx * ((size_t)jc + (size_t)ny2d * ((size_t)k + (size_t)nz * (size_t)(m - 1)))] += norm * v;
      }
  }
}

void orc_spectra_2d(const orc_ctx *c, const double *ux, const double *uy, const double *uz, const double *pr,
                    const double *sc, int ny2d, const int *iy2d, double *spec2d) {
  double *a = spec_box(c), *b = spec_box(c);
  const double *f[4] = {ux, uy, uz, pr};
  for (int q = 0; q < 4; q++) {
    spec_copy(c, f[q], 0.0, a);
    spec_fft_x(c, a);
    spec_fft_z(c, a);
    /* clcEnergySpectra2D: RealTemp1^2 + RealTemp2^2 (:1303-1306) */
    spec_acc_2d(c, a, a, 1 + q, ny2d, iy2d, q == 1, 1.0, spec2d);
  }
  spec_ux_cell(c, ux, a);
  spec_fft_x(c, a);
  spec_fft_z(c, a);
  spec_uy_cell(c, uy, b);
  spec_fft_x(c, b);
  spec_fft_z(c, b);
  spec_acc_2d(c, a, b, 5, ny2d, iy2d, 0, -1.0, spec2d); /* IsSymmetry = .false.: rCoe = -1 (:1359-1372) */
  if (c->cfg.scalar) {
    spec_copy(c, sc, 0.0, a);
    spec_fft_x(c, a);
    spec_fft_z(c, a);
    spec_acc_2d(c, a, a, 6, ny2d, iy2d, 0, 1.0, spec2d);
  }
  free(a);
  free(b);
}
#undef B
