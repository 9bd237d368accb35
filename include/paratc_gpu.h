/*
 * paratc_gpu.h -- C-ABI of the MI355X (gfx950) implementation of ParaTC's
 * per-time-step hot path: the m_Poisson pressure solve, the m_TScheme stencils
 * and wall-normal Crank-Nicolson solves, and the m_BC_and_Halo boundary fills.
 *
 * Design of the boundary
 *  - One context per case (ptc_create) holds what the reference keeps in module
 *    globals (m_Parameters, m_MeshAndMetries, m_Variables): grid, physics,
 *    y-metrics and the device-resident field arrays.
 *  - Fields live in HBM and are named by PTC_F_* ids, so the caller can alias
 *    buffers exactly as the reference does (RhsX => RealArr1, prphiHalo =>
 *    RealHalo, src/m_ChannelSystem.f90:108,124).  ptc_upload / ptc_download move
 *    a field between the host and the device in the reference's own layouts.
 *  - Every routine returns 0 on success and a negative PTC_E_* code on failure;
 *    ptc_last_error() returns the message.  The reference aborts instead
 *    (MainLog%CheckForError, src/m_LogInfo.f90:164-170); a host maps a nonzero
 *    code to that abort.
 *  - All work is ordered on the context's stream.  Only the calls that return a
 *    host scalar (monitors, divmax) synchronise.
 *  - Plain C types only: ints, doubles, pointers, sizes.
 */
#ifndef PARATC_GPU_H
#define PARATC_GPU_H

#include <stddef.h>

#ifdef __cplusplus
extern "C" {
#endif

#define PTC_ABI_VERSION 1

/* error codes */
#define PTC_OK 0
#define PTC_E_ARG (-1)         /* bad argument (size, id, layout, option) */
#define PTC_E_HIP (-2)         /* a HIP runtime call failed */
#define PTC_E_ALLOC (-3)       /* device allocation failed */
#define PTC_E_STATE (-4)       /* call out of order (e.g. no Poisson plan) */
#define PTC_E_UNSUPPORTED (-5) /* grid or option outside the built kernels */
#define PTC_E_LIMIT (-6)       /* div_limit / vel_limit guard tripped */

/* field ids: the arrays of src/m_Variables.f90:10-34 */
enum {
  PTC_F_UX = 0,   /* ux        haloed */
  PTC_F_UY,       /* uy        haloed */
  PTC_F_UZ,       /* uz        haloed */
  PTC_F_PR,       /* pressure  haloed */
  PTC_F_SCALAR,   /* scalar    haloed (ScalarFlow) */
  PTC_F_HISTX,    /* HistxOld  interior */
  PTC_F_HISTY,    /* HistyOld  interior */
  PTC_F_HISTZ,    /* HistzOld  interior */
  PTC_F_HISTC,    /* HistCOld  interior (ScalarFlow) */
  PTC_F_ARR1,     /* RealArr1  interior (RhsX, prsrc) */
  PTC_F_ARR2,     /* RealArr2  interior (RhsY) */
  PTC_F_HALO,     /* RealHalo  haloed   (RhsZ, prphiHalo) */
  PTC_F_ARRC,     /* RealArrC  haloed   (RhsC, ScalarFlow) */
  PTC_F_WORK1,    /* spare work arrays for callers */
  PTC_F_WORK2,
  PTC_F_UY_YM,    /* uy_ym   wall plane (nx, nz) */
  PTC_F_DUY_YM,   /* duy_ym  wall plane (nx, nz) */
  PTC_NFIELDS
};

/* host layouts for upload/download (Fortran column-major FP64) */
enum {
  PTC_L_HALO = 0,     /* (nx+4)*(ny+2)*(nz+4): x -1..nx+2, y 0..ny+1, z -1..nz+2 (mb1) */
  PTC_L_INTERIOR = 1, /* nx*ny*nz: x 1..nx, y 1..ny, z 1..nz (y1start:y1end) */
  PTC_L_PLANE = 2     /* nx*nz (uy_ym, duy_ym) */
};

/* Poisson solve precision / variant flags (ptc_poisson_config) */
enum {
  PTC_PREC_FP64 = 0,        /* plain FP64 */
  PTC_PREC_FP32_REFINE = 1  /* FP32 bulk + FP64 defect correction */
};
enum {
  PTC_TRI_AUTO = 0,
  PTC_TRI_THOMAS = 1,       /* one thread per line, the reference's InverseTridiagonal */
  PTC_TRI_PARTITIONED = 2   /* register-resident chunks + reduced system (Choice-3 idea) */
};

/* scalar ids for ptc_get_scalar */
enum {
  PTC_S_DIVMAX_PRSRC = 0, /* last clcPrSrc max|div| */
  PTC_S_PRGRADAVER = 1,   /* PrGradAver (src/m_TScheme.f90:284,298) */
  PTC_S_SIMTIME = 2,      /* SimTime */
  PTC_S_DT = 3            /* current dt */
};

typedef struct ptc_ctx ptc_ctx;

/* Grid: nxc, nyc, nzc and domain lengths (src/m_Parameters.f90:17-19). */
typedef struct ptc_grid {
  int nx, ny, nz, pad_;
  double xlx, yly, zlz;
} ptc_grid;

/* Physics and options: the BasicParam / ScalarFlowOptions keys the hot path
 * reads (src/m_Parameters.f90:118-129), after the derivations of
 * src/m_Parameters.f90:153-187 (uCRF shift of the wall values, SDiffCoe,
 * GravityEffVec, FallingVelVec). */
typedef struct ptc_params {
  int flow_type;    /* 1 channel, 2 half channel */
  int scalar;       /* 1 = the -DScalarFlow variant */
  int is_ux_const;  /* IsUxConst */
  int ischeme;      /* 1 AB2, 2 RK2, 3 RK3 */
  int icfl;         /* 1 adaptive dt, 2 fixed */
  int sbc_opt[2];   /* ScalarBcOption (y-, y+) */
  int pad_;
  double ubulk, xnu, ucrf, dtmax, cflc;
  double gravity[3];
  double uxbc[6], uybc[6], uzbc[6];
  double sdiff;
  double gev[3];
  double fvv[3];
  double sbc_val[4];
  double vel_limit, div_limit;
} ptc_params;

/* ---- context: replaces InitMeshAndMetries + AllocateVariables
 *      (src/m_MeshAndMetries.f90:71-335, src/m_Variables.f90:41-105).
 *      yp holds the nodes yp(1..ny+1) computed by the host's clcYcoord. */
int ptc_create(const ptc_grid *grid, const ptc_params *prm, const double *yp, ptc_ctx **out);
int ptc_destroy(ptc_ctx *ctx);
const char *ptc_last_error(void);
int ptc_abi_version(void);

/* ---- field transfer (device-resident between calls) */
int ptc_upload(ptc_ctx *ctx, int field, const double *host, int layout);
int ptc_download(ptc_ctx *ctx, int field, double *host, int layout);
int ptc_fill(ptc_ctx *ctx, int field, double value);
int ptc_copy(ptc_ctx *ctx, int dst, int src);
/* raw device address of a field (nx*(ny+2)*nz doubles, x fastest, y rows 0..ny+1) */
int ptc_field_device_ptr(ptc_ctx *ctx, int field, void **dptr, long long *count);
int ptc_synchronize(ptc_ctx *ctx);
/* the context's HIP stream (a hipStream_t), so a caller can order its own work
 * or time the library's kernels with events on the same stream */
int ptc_get_stream(ptc_ctx *ctx, void **stream);
/* y-metric table, ny+2 values (same numbering as the oracle header) */
int ptc_get_metric(ptc_ctx *ctx, int which, double *out);

/* ---- m_Parameters: PMcoeUpdate (src/m_Parameters.f90:215-230) */
int ptc_pm_coe_update(ptc_ctx *ctx, int ns, double dt);
/* coefficient vector [gamma theta alpha alphaC beta betaSc dt simtime prgradaver] */
int ptc_get_coe(ptc_ctx *ctx, double *out9);
int ptc_set_coe(ptc_ctx *ctx, const double *in9);

/* ---- m_Poisson (src/m_Poisson.f90:22) */
/* InitPoissonSolver (:28-175): builds the FFT plans and autotunes the variant
 * (tile shapes, Thomas vs partitioned) by timing, like the reference's 4-way
 * Choice search; leaves src and dst zeroed as the reference does (:174). */
int ptc_init_poisson_solver(ptc_ctx *ctx, int src, int dst);
/* clcPPE (:16, :352-452): dst interior = L^-1 src with the (kx,kz)=(0,0) line
 * pinned at j=1.  src is preserved (the reference destroys it; callers may not
 * rely on either). */
int ptc_clc_ppe(ptc_ctx *ctx, int src, int dst);
/* Destory_Poisson_FFT_Plan (:339-347) */
int ptc_destroy_poisson_plan(ptc_ctx *ctx);
/* explicit variant selection (overrides the autotuner); prec = PTC_PREC_*,
 * tri = PTC_TRI_*, tile = 0 selects the compile-time specialised FFT kernels
 * where they exist for nx/2 and nz (generic kernels otherwise), tile = 1..16
 * forces the generic kernels with that z column tile; sweeps = refinement
 * sweeps for PTC_PREC_FP32_REFINE */
int ptc_poisson_config(ptc_ctx *ctx, int prec, int tri, int tile, int sweeps);
/* Further plan knobs (each rebuilds the plan; InitPoissonSolver's search sets
 * them too):
 *   PTC_POISSON_CHUNK_ROWS  rows per chunk of the partitioned y-solve (4, 8,
 *                           12 or 16; 0 = the shortest chunk whose workgroup
 *                           fits).  With tri = PTC_TRI_AUTO a grid no
 *                           partitioned shape fits falls back to Thomas.
 *   PTC_POISSON_SLAB_PLANES run the x and z passes slab by slab of this many
 *                           y planes (0 = whole field; specialised FFT
 *                           kernels only), so each z pass finds its input in
 *                           the last-level cache.
 *   PTC_POISSON_TRI_LINES   0: the partitioned y-solve loads its rows straight
 *                           into registers (64 real lines per workgroup);
 *                           16 or 32: it stages that many lines through LDS
 *                           in smaller workgroups.
 *   PTC_POISSON_FAST_ZTILE  column tile of the specialised z kernel (0 = its
 *                           default; 16 for nz = 512 and 32 for nz = 256 are
 *                           the wider compiled alternatives).  A tile with no
 *                           compiled kernel falls back to the generic z pass.
 *   PTC_POISSON_TRI_LANES   real lines per workgroup of the register y-solve
 *                           (0 = automatic, 64, 32 or 16): narrower
 *                           workgroups are smaller, so several share a CU.
 *   PTC_POISSON_TRI_COMPLEX 16 or 32: the partitioned y-solve runs one lane
 *                           per complex line (shared row coefficients and
 *                           divisions), that many lines per workgroup; -1:
 *                           the real and imaginary parts as separate lanes;
 *                           0 (automatic): complex lines, 16 per workgroup,
 *                           where a shape fits and no other y-solve knob is
 *                           set.
 *   PTC_POISSON_STREAMING   nontemporal (cache-bypassing) loads and stores in
 *                           the x passes and the complex-line y-solve: 1 on,
 *                           -1 off, 0 (automatic) on where the spectrum is
 *                           larger than the last-level cache.
 *   PTC_POISSON_PERSIST     1: the x and z passes of the lengths that have
 *                           one (x: nx = 1024, 4608, 9216; z: nz = 1400)
 *                           run as persistent workgroups that load
 *                           their next line or column tile while they
 *                           transform the current one; 2: the z passes
 *                           only; 3: the x passes only; 0 or -1: one
 *                           workgroup per line or tile.  The search of
 *                           InitPoissonSolver tries them. */
enum {
  PTC_POISSON_CHUNK_ROWS = 1,
  PTC_POISSON_SLAB_PLANES = 2,
  PTC_POISSON_TRI_LINES = 3,
  PTC_POISSON_FAST_ZTILE = 4,
  PTC_POISSON_TRI_LANES = 5,
  PTC_POISSON_TRI_COMPLEX = 6,
  PTC_POISSON_STREAMING = 7,
  PTC_POISSON_PERSIST = 8
};
int ptc_poisson_set(ptc_ctx *ctx, int knob, int value);
/* the value a plan knob has in the plan in use (chunk rows, slab planes,
 * LDS-staged lines, z tile, lanes per workgroup, complex lines per
 * workgroup, streaming on (1) or off (0), persistent passes (1 x, 2 z,
 * 3 both); 0 where the knob's variant is not in use) */
int ptc_poisson_get_knob(ptc_ctx *ctx, int knob, int *value);
/* Workgroup shape of the row-marching clcRhsXYZ kernel (0 = automatic):
 *   PTC_RHS_SEGMENT_THREADS  threads per row segment where rows are longer
 *                            than 512 points: 64, 128 or 256 (automatic 128)
 *   PTC_RHS_ROWS             rows j per workgroup, at most 512 threads in
 *                            all: the inner rows' j-1 / j+1 values come
 *                            through LDS (automatic: 4 for rows in
 *                            segments, 2 where a whole row is at most 128
 *                            threads, else 1)
 *   PTC_RHS_XCD_RUN          consecutive tiles run on one XCD (1 = plain
 *                            order; automatic 32)
 *   PTC_RHS_MIN_PLANES       z planes per chunk at least (each chunk starts
 *                            its z window afresh; automatic 16, halved down
 *                            to 2 on grids that would otherwise give fewer
 *                            than 1.5 workgroups per CU); the chunks are
 *                            otherwise many enough for about 96 wavefronts
 *                            per CU
 * A change drops the captured step graph.  Results do not depend on the
 * shape except through the order of the IsUxConst flow-rate sum. */
enum {
  PTC_RHS_SEGMENT_THREADS = 1,
  PTC_RHS_ROWS = 2,
  PTC_RHS_XCD_RUN = 3,
  PTC_RHS_MIN_PLANES = 4
};
int ptc_rhs_set(ptc_ctx *ctx, int knob, int value);
/* the shape in use for this grid: threads per row segment (the padded row
 * where a row is one segment), rows per workgroup, tiles per XCD run */
int ptc_rhs_get(ptc_ctx *ctx, int knob, int *value);
/* the plan in use: out10 = [prec, tri, z tile, sweeps, specialised-kernel
 * mask (1 = x pass, 2 = z pass), chunk rows, chunks per line, lanes per
 * workgroup, slab planes, LDS-staged lines (0 = register kernel)] (after
 * autotuning; chunk fields are 0 for Thomas) */
int ptc_poisson_get_config(ptc_ctx *ctx, int *out10);
/* The last ptc_init_poisson_solver search as text, one line per timing: each
 * candidate plan's knobs and ms per solve, the re-timings of the closest
 * ones, the pick (the reference prints its four choices' times,
 * src/m_Poisson.f90:50-171).  Copies at most cap - 1 bytes and a NUL into
 * buf (buf may be NULL with cap 0); *len (may be NULL) receives the full
 * length. */
int ptc_poisson_search_log(ptc_ctx *ctx, char *buf, int cap, int *len);
/* FP64 residual max|src - L dst| over the interior (test and refinement aid) */
int ptc_poisson_residual(ptc_ctx *ctx, int src, int dst, double *rmax);
/* Per-stage kernel times: with enable = 1 every later solve records HIP events
 * around its five passes; ms5 (may be NULL) receives the summed times in ms and
 * *nsolves (may be NULL) the number of solves since the last reset.  enable =
 * 0 stops recording, -1 only reads.  enable >= 0 resets the sums after reading. */
int ptc_poisson_stage_times(ptc_ctx *ctx, int enable, float *ms5, int *nsolves);

/* ---- m_TScheme (src/m_TScheme.f90:14-19) */
int ptc_init_time_scheme(ptc_ctx *ctx);
/* clcRhsXYZ (:113-360): pass -1 for scalar/rhsc/histc when not ScalarFlow */
int ptc_clc_rhs_xyz(ptc_ctx *ctx, int ux, int uy, int uz, int scalar, int rhsx, int rhsy,
                    int rhsz, int rhsc, int histx, int histy, int histz, int histc,
                    int pressure);
int ptc_clc_u1hat(ptc_ctx *ctx, int ux, int rhsx);                /* :365-394 */
int ptc_clc_u2hat(ptc_ctx *ctx, int uy, int rhsy, int duy_ym);    /* :399-437 */
int ptc_clc_u3hat(ptc_ctx *ctx, int uz, int rhsz);                /* :442-471 */
int ptc_clc_scalar(ptc_ctx *ctx, int scalar, int rhsc);           /* :477-506 */
/* clcPrSrc (:48-77); divmax may be NULL (then no host synchronisation) */
int ptc_clc_prsrc(ptc_ctx *ctx, int ux, int uy, int uz, int prsrc, double *divmax);
/* ptc_clc_prsrc without any device allocation.  No counterpart in the
 * reference: it is for callers that cannot allocate (code under HIP graph
 * capture, tight serving loops, harnesses that forbid hipMalloc).  Same
 * kernels and results as ptc_clc_prsrc, with the per-workgroup partial
 * maxima in a 16 MiB device array of the library instead of a buffer
 * allocated on first use.  Limits: ceil(nx/64)*ceil(ny/4)*nz <= 2^21 (else
 * PTC_E_ARG); the four fields must already have device memory (uploaded,
 * filled or written before; else PTC_E_STATE).  Every call on one GPU shares
 * that array, so calls that overlap on one GPU (contexts on different
 * streams) race.  divmax = NULL under graph capture (a non-NULL divmax
 * synchronises the stream). */
int ptc_clc_prsrc_noalloc(ptc_ctx *ctx, int ux, int uy, int uz, int prsrc, double *divmax);
int ptc_fluid_vel_update(ptc_ctx *ctx, int phi, int ux, int uy, int uz); /* :82-107 */
int ptc_pressure_update(ptc_ctx *ctx, int pressure, int phi);           /* :512-532 */

/* ---- m_BC_and_Halo (src/m_BC_and_Halo.f90:13-17) and Update_uy_ym
 *      (src/m_FlowCase.f90:112-122).  x and z are periodic in the device
 *      layout, so halo refreshes (myupdate_halo) are implicit; these set the
 *      wall ghost rows. */
int ptc_set_bc_and_update_halo(ptc_ctx *ctx, int ux, int uy, int uz, int uy_ym);
int ptc_set_bc_and_update_halo_for_prsrc(ptc_ctx *ctx, int ux, int uy, int uz, int uy_ym);
int ptc_set_bc_and_update_halo_pr(ptc_ctx *ctx, int p);
int ptc_set_bc_and_update_halo_scalar(ptc_ctx *ctx, int scalar);
int ptc_update_uy_ym(ptc_ctx *ctx, int uy_ym, int duy_ym);

/* ---- m_Tools monitors (src/m_Tools.f90:24-110, 256-301) */
int ptc_calc_max_cfl(ptc_ctx *ctx, int ux, int uy, int uz, double *uddxmax);
int ptc_check_divergence(ptc_ctx *ctx, int ux, int uy, int uz, double *divmax);
int ptc_calc_vmax(ptc_ctx *ctx, int ux, int uy, int uz, double *vmax3);
int ptc_calc_ux_aver(ptc_ctx *ctx, int ux, double *mean);
int ptc_clc_mean_value(ptc_ctx *ctx, int ux, int scalar, double *res3);

/* ---- m_ChannelSystem: nsteps calls of ChannelIterate (src/m_ChannelSystem.f90:80-145)
 *      on the standard buffers.  mon (may be NULL) receives, for the last step,
 *      [divmax1, divmax2, vmax x3, cfl*dt, dt]; returns PTC_E_LIMIT when the
 *      reference's div_limit / vel_limit guard (:142-145) would abort. */
int ptc_channel_iterate(ptc_ctx *ctx, int nsteps, double *mon7);
/* With a fixed dt (iCFL != 1) ptc_channel_iterate captures one step into a
 * HIP graph on first use and replays it (enable = 1, the default); enable = 0
 * launches every kernel of every step instead.  No reference counterpart. */
int ptc_step_graphs(ptc_ctx *ctx, int enable);
int ptc_get_scalar(ptc_ctx *ctx, int which, double *value);

/* ---- clcStat (src/m_FlowCase.f90:257-949), the statistics side.
 *  ptc_clc_stat adds the x-z means of the NCHASTAT statistics (40 with a
 *  scalar, else 35) of planes j = 1..ny+1 to a device SumStat (:283-413),
 *  PrGradAver to PrGradsum (:416) and counts the sample (nfstime, :428);
 *  scalar = -1 without a scalar.  ptc_get_stat copies SumStat(NCHASTAT, ny+1)
 *  (column-major, as the reference's array; may be NULL), the statistic count,
 *  nfstime and PrGradsum; reset = 1 zeroes them as after the reference writes
 *  its stats file (:944).  The spectra are the three calls below. */
int ptc_clc_stat(ptc_ctx *ctx, int ux, int uy, int uz, int pressure, int scalar);
int ptc_get_stat(ptc_ctx *ctx, double *sumstat, int *nstat, int *nfstime, double *prgradsum, int reset);

/* ---- clcStat's energy spectra (src/m_FlowCase.f90:430-805 and the helpers
 *      clcEnergySpectraX/Z, clcLTS_*, clcCospectraX/Z, clcEnergySpectra2D,
 *      clcCosSpectra2D at :954-1375).
 *  ptc_init_spectra takes InitStatVar's SpectraOptions / Spectra2DOpt
 *  (:140-251): clc_1d = clcSpectra1D, jfor_lcs = jForLCS (1..ny), clc_2d =
 *  clcSpectra2D, iy_spec2d[ny_spec2d] = iySpec2D (1-based rows; 2..ny/2 in a
 *  channel, 2..ny in a half channel, else PTC_E_ARG as the reference aborts;
 *  ny_spec2d < 1 turns the 2-D spectra off).  It zeroes the accumulators.
 *  ptc_clc_spectra adds one sample of the current fields, as clcStat does when
 *  mod(itime, ivSpec) == 0 (the host keeps that schedule); it needs the
 *  Poisson plan (InitPoissonSolver), whose forward transforms it uses in FP64
 *  whatever the plan's storage precision.  scalar = -1 without a scalar.
 *  ptc_get_spectra copies the sums, not yet divided by nSpectime:
 *    specx  EnergySpecX(nx/2+1, ny, NE1), specz EnergySpecZ(nz/2+1, ny, NE1),
 *    NE1 = 9 with a scalar, else 8 (1 ux+uCRF, 2 uy, 3 uz, 4 p, 5 co-spectrum
 *    of ux+uCRF and the cell-centre uy, 6 of the cell-centre ux and uy,
 *    7/8 real/imaginary linear coherence with row jForLCS, 9 scalar);
 *    spec2d EnergySpec2D(nx, ny_spec2d, nz, NE2), NE2 = 6 or 5 (1 ux, 2 uy,
 *    3 uz, 4 p, 5 cell-centre ux x uy, 6 scalar), values in FFTW's
 *    half-complex order along x and z;
 *  all column-major, any pointer may be NULL; reset = 1 zeroes the sums and
 *  nSpectime as after the reference writes SpecX/SpecZ/Spec2D (:943-947). */
int ptc_init_spectra(ptc_ctx *ctx, int clc_1d, int jfor_lcs, int clc_2d, int ny_spec2d, const int *iy_spec2d);
int ptc_clc_spectra(ptc_ctx *ctx, int ux, int uy, int uz, int pressure, int scalar);
int ptc_get_spectra(ptc_ctx *ctx, double *specx, double *specz, double *spec2d, int *nspectime, int reset);

/* dump_visu (src/m_IOAndVisu.f90:182-385): the cell-centre field `which` of
 * the current ux, uy, uz, pressure (and scalar) at every iskip/jskip/kskip-th
 * point from the first, as decomp_2d_write_every keeps them with from1
 * (src/m_Decomp2d.f90:1232-1389), x fastest, as real_bytes = 4 or 8 reals
 * (mytype_save) into host_out of count = ceil(nx/iskip) ceil(ny/jskip)
 * ceil(nz/kskip) values.  ux carries uCRF; PTC_VISU_WMAG is |vorticity|.
 * PTC_VISU_Q is Clc_Q_vor (src/m_Tools.f90:307-342), PTC_VISU_LAMBDA2
 * Clc_lamda2 with CubicRoot (:347-442, :494-570); where the reference aborts
 * with "lamda2 wrong" (a complex root pair) this returns PTC_E_LIMIT. */
enum {
  PTC_VISU_UX = 0,
  PTC_VISU_UY = 1,
  PTC_VISU_UZ = 2,
  PTC_VISU_PR = 3,
  PTC_VISU_WX = 4,
  PTC_VISU_WY = 5,
  PTC_VISU_WZ = 6,
  PTC_VISU_WMAG = 7,
  PTC_VISU_SCALAR = 8,
  PTC_VISU_Q = 9,
  PTC_VISU_LAMBDA2 = 10
};
int ptc_visu_field(ptc_ctx *ctx, int which, int iskip, int jskip, int kskip, int real_bytes, void *host_out,
                   long long count);

/* ---- Seeded synthetic inputs, generated in HBM.  No counterpart in the
 *      reference, whose initial perturbation is seeded from the wall clock
 *      (InitVelocity, src/m_FlowCase.f90:49-107) and cannot be reproduced.
 *  Values are a pure function of (seed, stream, cell): SplitMix64 on the
 *  counter q = i + nx (j - 1 + ny k) of the interior cell in column-major
 *  order (the host generator paratc_amd/rng.py gives the same values).
 *  ptc_seed_normal: field interior = N(0,1) per cell (Box-Muller on uniform
 *  pairs); compatible = 1 then subtracts the dyp-weighted mean, so the field
 *  is a compatible right-hand side of the Neumann Poisson problem (the
 *  Poisson workloads' sources).
 *  ptc_seed_fields: the initial ux, uy, uz (and scalar) of a case: uniform
 *  noise amp (2 U - 1) on u, v, w (streams 0, 1, 2; v = 0 on the bottom wall
 *  face); PTC_SEED_CHANNEL adds the laminar profile 1.5 (1 - ((y - h)/h)^2)
 *  ubulk to u (h = yly / 2); PTC_SEED_FRONT sets the scalar to the front
 *  0.5 (1 - tanh((x - x0) / delta)) plus noise (stream 3), velocities noise
 *  only.  Ghost rows are zeroed: call the boundary fills afterwards. */
enum {
  PTC_SEED_CHANNEL = 0,
  PTC_SEED_FRONT = 1
};
int ptc_seed_normal(ptc_ctx *ctx, int field, int seed, int stream, int compatible);
int ptc_seed_fields(ptc_ctx *ctx, int kind, int seed, double amp, double x0, double delta);

/* ---- Ensembles: many independent cases on one GPU.  No counterpart in the
 *      reference, which runs one case per job; this is the serving layer
 *      around ChannelIterate (src/m_ChannelSystem.f90:80-145).
 *
 *  A case is a context the caller has created, filled (ptc_upload, boundary
 *  fills) and given a Poisson plan (ptc_init_poisson_solver or
 *  ptc_poisson_config).  ptc_ensemble_admit takes ownership when the case's
 *  device footprint fits the ensemble's HBM budget, else returns PTC_E_LIMIT
 *  and the caller keeps the context (admission control: queue it and retry).
 *  With a fixed time step (iCFL != 1) the case's step is captured once in a
 *  HIP graph and each later step is one graph launch on the case's stream;
 *  CFL-adaptive cases are stepped with ordinary launches.
 *  ptc_ensemble_step enqueues nsteps steps of every admitted case (one step
 *  per case in turn, each case on its own stream) and returns without a host
 *  barrier: a case runs ahead by at most four steps, beyond which the call
 *  waits on that case's own oldest step only.  Work on a case's context
 *  (downloads, monitors) is ordered after its enqueued steps.
 *  ptc_ensemble_sync waits for every enqueued step.  A case-step's latency
 *  is the time from the completion of the case's previous step (or its
 *  admission) to its own; ptc_ensemble_latency waits for the enqueued steps
 *  and gives the 50th and 99th percentiles (ms).  ptc_ensemble_case's steps
 *  count the enqueued steps.  ptc_ensemble_release waits for the case's
 *  steps, removes it and hands its context back (ctx != NULL) or destroys
 *  it.  One host thread per ensemble. */
typedef struct ptc_ensemble ptc_ensemble;
int ptc_ensemble_create(size_t hbm_budget_bytes, ptc_ensemble **out);
int ptc_ensemble_destroy(ptc_ensemble *e);
/* device bytes a case holds once everything its step touches is allocated */
int ptc_case_footprint(ptc_ctx *ctx, size_t *bytes);
int ptc_ensemble_admit(ptc_ensemble *e, ptc_ctx *ctx, int *case_id);
int ptc_ensemble_step(ptc_ensemble *e, int nsteps);
int ptc_ensemble_sync(ptc_ensemble *e);
int ptc_ensemble_case(ptc_ensemble *e, int case_id, ptc_ctx **ctx, long long *steps);
int ptc_ensemble_release(ptc_ensemble *e, int case_id, ptc_ctx **ctx);
int ptc_ensemble_usage(ptc_ensemble *e, size_t *used_bytes, size_t *budget_bytes, int *live_cases);
int ptc_ensemble_latency(ptc_ensemble *e, int reset, double *p50_ms, double *p99_ms, long long *nsamples);

#ifdef __cplusplus
}
#endif
#endif /* PARATC_GPU_H */
