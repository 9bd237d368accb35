/*
 * ptc_oracle.h -- CPU restatement of ParaTC's per-time-step hot path.
 *
 * TEST INFRASTRUCTURE ONLY.  This oracle is the checker for the HIP library
 * in paratc_amd/csrc: only tests/, __graft_entry__.smoke() and bench.py's
 * cpu_baseline leg may load it.  The product path never links or calls it.
 *
 * Layout conventions follow the reference exactly (column-major FP64):
 *   haloed arrays  (mb1 bounds, src/m_Variables.f90:47-50):
 *       x -1..nx+2, y 0..ny+1, z -1..nz+2   -> (nx+4)*(ny+2)*(nz+4) doubles
 *   interior arrays (y1start:y1end on one rank):
 *       x 1..nx, y 1..ny, z 1..nz            -> nx*ny*nz doubles
 *   wall planes (uy_ym, duy_ym): x 1..nx, z 1..nz -> nx*nz doubles
 *   y vectors: index j stored at [j] for j = 0..ny+1 (ny+2 doubles).
 *
 * Every function names the reference routine it restates (file:line).
 */
#ifndef PTC_ORACLE_H
#define PTC_ORACLE_H

#ifdef __cplusplus
extern "C" {
#endif

/* Case description: the subset of the reference namelists the hot path reads
 * (src/m_Parameters.f90:118-129, src/m_MeshAndMetries.f90:77-80). */
typedef struct orc_cfg {
  int nx, ny, nz;          /* nxc, nyc, nzc (cells) */
  int flow_type;           /* 1 = channel (FT_CH), 2 = half channel (FT_HC) */
  int scalar;              /* 1 = the -DScalarFlow build */
  int is_ux_const;         /* IsUxConst */
  int ischeme;             /* 1 AB2, 2 RK2, 3 RK3 */
  int icfl;                /* 1 adaptive dt, 2 fixed dt */
  int sbc_opt[2];          /* ScalarBcOption (y-, y+): -1/-2/-3 */
  int pad_;
  double xlx, yly, zlz;
  double ubulk, xnu, ucrf, dtmax, cflc;
  double gravity[3];
  double uxbc[6], uybc[6], uzbc[6];  /* already shifted by -uCRF if IsUseCRF */
  double sdiff;            /* SDiffCoe = xnu / Sc */
  double gev[3];           /* GravityEffVec */
  double fvv[3];           /* FallingVelVec (x component already includes uCRF) */
  double sbc_val[4];       /* ScalarBcValues (F_b, F_t, S_b, S_t) */
} orc_cfg;

typedef struct orc_ctx orc_ctx;

orc_ctx *orc_create(const orc_cfg *cfg, const double *yp /* yp(1..ny+1) */);
void orc_destroy(orc_ctx *c);
int orc_set_threads(int n);

/* Metric tables, written into caller arrays of ny+2 doubles (index j).
 * which: 0 yc 1 dyp 2 dyc 3 rdyp 4 rdyc 5 yinterp 6 am2ph 7 ac2ph 8 ap2ph
 *        9 am2c 10 ac2c 11 ap2c 12 am2cForCN 13 ac2cForCN 14 ap2cForCN
 *        15 am2p 16 ac2p 17 ap2p 18 am2cSc 19 ac2cSc 20 ap2cSc
 *        21 am2cScCN 22 ac2cScCN 23 ap2cScCN                               */
int orc_get_metric(const orc_ctx *c, int which, double *out);
/* Modified wavenumbers WaveNumX(1..nx) / WaveNumZ(1..nz) (m_Poisson.f90:37-47) */
void orc_wavenumbers(const orc_ctx *c, double *wnx, double *wnz);

/* Batched building blocks */
void orc_inverse_tridiagonal(const double *aj, const double *bj, const double *cj,
                             double *fj, int m, int n);
void orc_r2hc(int n, int howmany, int istride, int idist, double *data);
void orc_hc2r(int n, int howmany, int istride, int idist, double *data);

/* Poisson: choice 1..4 (clcPPE_x1, clcPPE_z1, clcPPE_x1_reduce, clcPPE_z1_reduce).
 * prsrc is interior and is destroyed; prphi is haloed, interior written only. */
void orc_poisson(orc_ctx *c, int choice, double *prsrc, double *prphi);

/* Substep coefficients (PMcoeUpdate, m_Parameters.f90:215-230). */
void orc_pm_coe_update(orc_ctx *c, int ns, double dt);
void orc_get_coe(const orc_ctx *c, double *out /* 9: gamma theta alpha alphaC beta betaSc dt simtime prgradaver */);
void orc_set_coe(orc_ctx *c, const double *in /* same 9 */);

/* m_TScheme */
void orc_clc_rhs_xyz(orc_ctx *c, const double *ux, const double *uy, const double *uz,
                     const double *scalar, double *rhsx, double *rhsy, double *rhsz,
                     double *rhsc, double *histx, double *histy, double *histz,
                     double *histc, const double *pressure);
void orc_clc_u1hat(orc_ctx *c, double *ux, const double *rhsx);
void orc_clc_u2hat(orc_ctx *c, double *uy, const double *rhsy, const double *duy_ym);
void orc_clc_u3hat(orc_ctx *c, double *uz, const double *rhsz);
void orc_clc_scalar(orc_ctx *c, double *scalar, const double *rhsc);
double orc_clc_prsrc(orc_ctx *c, const double *ux, const double *uy, const double *uz, double *prsrc);
void orc_fluid_vel_update(orc_ctx *c, const double *prphi, double *ux, double *uy, double *uz);
void orc_pressure_update(orc_ctx *c, double *pressure, const double *prphi);

/* m_BC_and_Halo + myupdate_halo */
void orc_update_halo(const orc_ctx *c, double *mat, int xmh, int xph, int zmh, int zph);
void orc_set_bc_and_update_halo(orc_ctx *c, double *ux, double *uy, double *uz, const double *uy_ym);
void orc_set_bc_for_prsrc(orc_ctx *c, double *ux, double *uy, double *uz, const double *uy_ym);
void orc_set_bc_pr(orc_ctx *c, double *p);
void orc_set_bc_scalar(orc_ctx *c, double *s);
void orc_update_uy_ym(orc_ctx *c, double *uy_ym, double *duy_ym);

/* m_Tools monitors */
double orc_calc_max_cfl(const orc_ctx *c, const double *ux, const double *uy, const double *uz);
double orc_check_divergence(const orc_ctx *c, const double *ux, const double *uy, const double *uz);
void orc_calc_vmax(const orc_ctx *c, const double *ux, const double *uy, const double *uz, double *vmax3);
double orc_calc_ux_aver(const orc_ctx *c, const double *ux);
void orc_clc_mean_value(const orc_ctx *c, const double *ux, const double *scalar, double *res3);

/* clcStat accumulation (m_FlowCase.f90:283-413): adds the x-z means of the
 * NCHASTAT = 40 (scalar build) or 35 statistics of planes j = 1..ny+1 to
 * sumstat(NCHASTAT, ny+1), column-major as the reference's SumStat. */
void orc_clc_stat(const orc_ctx *c, const double *ux, const double *uy, const double *uz, const double *pressure,
                  const double *scalar, double *sumstat);

/* dump_visu's cell-centre fields (src/m_IOAndVisu.f90:182-385) on the points
 * decomp_2d_write_every keeps (every iskip/jskip/kskip-th, from the first),
 * x fastest: which = 0 ux (+uCRF), 1 uy, 2 uz, 3 pr, 4 wx, 5 wy, 6 wz,
 * 7 |w|, 8 scalar, 9 Q (Clc_Q_vor, src/m_Tools.f90:307-342), 10 lambda2
 * (Clc_lamda2 with CubicRoot, :347-442, :494-570; NaN where the reference
 * aborts on a complex root pair).  orc_visu_count: the number of points. */
long long orc_visu_count(const orc_ctx *c, int iskip, int jskip, int kskip);
void orc_dump_visu(const orc_ctx *c, int which, const double *ux, const double *uy, const double *uz, const double *pr,
                   const double *sc, int iskip, int jskip, int kskip, double *out);

/* Energy spectra of clcStat (src/m_FlowCase.f90:430-805, helpers :954-1375),
 * added to the accumulators as the reference does per sample.
 * orc_spectra_1d: specx(nx/2+1, ny, NE1) and specz(nz/2+1, ny, NE1), column-
 * major, NE1 = 9 (scalar build) or 8: 1 ux+uCRF, 2 uy, 3 uz, 4 p, 5 co-spectrum
 * of ux+uCRF and the cell-centre uy, 6 co-spectrum of the cell-centre ux and
 * uy, 7/8 real/imaginary linear coherence with row jlcs, 9 scalar.
 * orc_spectra_2d: spec2d(nx, ny2d, nz, NE2) on the rows iy2d (1-based), NE2 =
 * 6 or 5: 1 ux, 2 uy, 3 uz, 4 p, 5 cell-centre ux x uy, 6 scalar.
 * orc_spectra_count: 0 -> specx length, 1 -> specz length, 2 -> NE2. */
int orc_spectra_count(const orc_ctx *c, int which);
void orc_spectra_1d(const orc_ctx *c, const double *ux, const double *uy, const double *uz, const double *pr,
                    const double *sc, int jlcs, double *specx, double *specz);
void orc_spectra_2d(const orc_ctx *c, const double *ux, const double *uy, const double *uz, const double *pr,
                    const double *sc, int ny2d, const int *iy2d, double *spec2d);

/* One ChannelIterate time step (m_ChannelSystem.f90:80-145).  Buffers follow
 * the reference's aliasing: arr1 = RealArr1, arr2 = RealArr2, halo = RealHalo,
 * arrc = RealArrC.  mon receives [divmax1 (last substep), divmax2, vmax x3,
 * cfl*dt, dt].  poisson_choice selects the clcPPE variant.
 * Each procedure it calls is pinned on its own; the sequence of many steps is
 * parity unpinned beyond its invariants (post-projection divergence at
 * round-off, the constant flow rate held). */
void orc_channel_iterate(orc_ctx *c, int poisson_choice,
                         double *ux, double *uy, double *uz, double *p, double *scalar,
                         double *histx, double *histy, double *histz, double *histc,
                         double *arr1, double *arr2, double *halo, double *arrc,
                         double *uy_ym, double *duy_ym, double *mon);

#ifdef __cplusplus
}
#endif
#endif
