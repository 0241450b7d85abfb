/*
 * fvm_capi.h -- C ABI of the MI355X fractional 4-step incompressible Navier-Stokes solver.
 *
 * The reference (Johann426/FVM, LES/v2.3_LES-Reynolds stresses/, cited as RLES/) keeps all
 * state in the Fortran module COMDAT_SHARED and advances it with implicit-interface external
 * subroutines called from MAIN (RLES/main.f90:35-161).  This header is the drop-in boundary a
 * Fortran driver binds with ISO_C_BINDING (fvm_amd/fortran/fvm_capi_mod.f90): the module state
 * becomes an opaque device context, each subroutine becomes one entry point that acts on the
 * device-resident state.  Host arrays cross the boundary only at set-up and restart I/O.
 *
 * Conventions: plain pointers and sizes, no torch or HIP types.  Cell arrays are FP64,
 * (NI+1)(NJ+1)(NK+1) values, I fastest, indices 0..NI x 0..NJ x 0..NK incl. ghosts (the
 * reference's (IS:IE,JS:JE,KS:KE) arrays with one rank).  Node arrays are NI*NJ*NK values,
 * I fastest (the records of the reference's grid file).  Every function returns an int status:
 * 0 ok, 1 bad argument, 2 HIP runtime error, 3 non-finite value, 4 Poisson iteration hit its
 * sweep cap, 5 grid too large for a line kernel.  The reference has no status returns
 * (failures surface as NaN); the codes are an addition.  All calls on one context are ordered
 * on one HIP stream; only scalars are returned synchronously.
 */
#ifndef FVM_CAPI_H
#define FVM_CAPI_H

#ifdef __cplusplus
extern "C" {
#endif

typedef struct FvmCtx fvm_ctx;

/* Run constants.  RLES/comdat_param.f90:6-48 holds them as compile-time PARAMETERs. */
typedef struct fvm_params {
  int ni, nj, nk;       /* node counts NI (radial), NJ (circumferential), NK (span)      */
  double re;            /* Reynolds number RE                                            */
  double dtime;         /* time step DTIME                                               */
  double ref_a;         /* reference area REF_A of the force coefficients               */
  int use_les;          /* USE_LES: dynamic Smagorinsky model on (1) / off (0)           */
  int nswpp;            /* NSWPP: Poisson sweep cap                                       */
  double resimax;       /* RESIMAX: scaled-L1 Poisson residual tolerance                 */
  int poisson_mode;     /* 0: FP64 zebra radial-line Gauss-Seidel relaxation;
                           1: FP32 zebra sweeps inside FP64 defect correction;
                           2: FP64 defect correction with the FP64 separable inverse
                              (J/K Fourier modes x radial line solves);
                           3: the same with the separable inverse in FP32 (mixed form)  */
  int check_every;      /* sweeps between residual evaluations (>=1)                     */
  int af_mode;          /* AF line sweeps in the reference's lexicographic order, computed
                           0: by chunked block iteration to convergence (default);
                           1: by anti-diagonal wavefront launches;
                           2: as 0, with the I and J stages of the momentum step (fvm_step,
                              fvm_iter_u) solved for their change in FP32 against the FP64
                              residual of the stage input, the change added to the FP64
                              field (mixed form; the K stage, ITER_T and the one-component
                              entries stay FP64)                                            */
  double af_tol;        /* chunked iteration: stop when max change <= af_tol * max |u|
                           (<= 0: 1e-14)                                                 */
} fvm_params;

/* Per-step diagnostics (RLES/calccoef.f90:11-201 scalars and the ITER_P counters). */
typedef struct fvm_step_info {
  double resip;         /* final scaled-L1 residual of the pressure equation (FP64)     */
  int niter;            /* Poisson line sweeps executed                                  */
  int nouter;           /* defect-correction passes (mixed mode), else residual checks  */
  double pref;          /* span-mean far-field pressure subtracted from p                */
  double c_avg;         /* dynamic Smagorinsky coefficient C = <LM>/<2MM>                */
  double cfl, cpb, cd, cl, cd_p, cd_v, cl_p, cl_v;
  int af_iters[3];      /* passes of the three AF stage sweeps                          */
  double nu;            /* Nusselt number NU (calccoef.f90:117-184); 0 without the
                           passive temperature                                           */
  double c_t;           /* dynamic coefficient of the subgrid heat flux
                           (stress_tensor_dy.f90:686-711); 0 without the temperature     */
} fvm_step_info;

/* Field selectors for fvm_set_field / fvm_get_field. */
enum {
  FVM_U = 0, FVM_V = 1, FVM_W = 2, FVM_P = 3, FVM_NLX = 4, FVM_NLY = 5, FVM_NLZ = 6,
  FVM_DIFX = 7, FVM_DIFY = 8, FVM_DIFZ = 9, FVM_SU = 10, FVM_SV = 11, FVM_SW = 12,
  FVM_DJAC = 13, FVM_XIX = 14, FVM_ETX = 15, FVM_ZTX = 16, FVM_XIY = 17, FVM_ETY = 18,
  FVM_ZTY = 19, FVM_XIZ = 20, FVM_ETZ = 21, FVM_ZTZ = 22, FVM_X = 23, FVM_Y = 24, FVM_Z = 25,
  FVM_THETA = 26, FVM_NUT = 27, FVM_SGSX = 28, FVM_SGSY = 29, FVM_SGSZ = 30, FVM_UN = 31,
  FVM_VN = 32, FVM_WN = 33,
  /* passive temperature (after fvm_set_thermal): T, its convective term ADV (setting it sets
     ADV_NM1 too), DIFT, SGS_T, the eddy conductivity, the test-filtered T^, dT/dx_i, T_N */
  FVM_T = 34, FVM_ADV = 35, FVM_DIFT = 36, FVM_SGST = 37, FVM_KAPPA = 38, FVM_THAT = 39,
  FVM_DTDX = 40, FVM_DTDY = 41, FVM_DTDZ = 42, FVM_TN = 43,
  /* field output (after fvm_writedata_fields): |omega| and lambda_2 on the cells, and the
     nodal averages PG, UG, VG, WG, TG, |omega| and lambda_2 held at the node's cell index
     (nodes I = 1..NI, J = 1..NJ, K = 1..NK of the cell-array layout); the wall y+ of CALCCOEF
     on the I = 1 cells (zero elsewhere) */
  FVM_VORM = 44, FVM_LAMDA2 = 45, FVM_PG = 46, FVM_UG = 47, FVM_VG = 48, FVM_WG = 49,
  FVM_TG = 50, FVM_VORG = 51, FVM_LAMDAG = 52, FVM_YPLUS = 53,
  /* time averages (after fvm_post_begin): the running sums of u, v, w, p, T, of uu, uv, uw, vv,
     vw, ww and of the SGS stresses -2 nu_t S_ij (11, 12, 13, 22, 23, 33) on every cell, which
     fvm_post_finalize leaves as the reference leaves U_AVG..WW_AVG and SGS11..SGS33: the means,
     the Reynolds stresses and the SGS means */
  FVM_SUM_U = 54, FVM_SUM_V = 55, FVM_SUM_W = 56, FVM_SUM_P = 57, FVM_SUM_T = 58,
  FVM_SUM_UU = 59, FVM_SUM_UV = 60, FVM_SUM_UW = 61, FVM_SUM_VV = 62, FVM_SUM_VW = 63,
  FVM_SUM_WW = 64, FVM_SUM_SGS11 = 65, FVM_SUM_SGS12 = 66, FVM_SUM_SGS13 = 67,
  FVM_SUM_SGS22 = 68, FVM_SUM_SGS23 = 69, FVM_SUM_SGS33 = 70,
  /* after fvm_post_finalize: node averages held at the node's cell index (nodes I = 1..NI,
     J = 1..NJ, K = 1..NK) of the means of p, u, v, w, T (Avrg000.PLT), of the rms
     sqrt(<u_i u_i>) (Urms000.PLT) and of the Reynolds stresses with the SGS part (Strs000.PLT),
     and the wall friction, pressure and Nusselt coefficients on the I = 1 cells (Surf000.PLT,
     zero elsewhere) */
  FVM_AVG_PG = 71, FVM_AVG_UG = 72, FVM_AVG_VG = 73, FVM_AVG_WG = 74, FVM_AVG_TG = 75,
  FVM_RMS_UG = 76, FVM_RMS_VG = 77, FVM_RMS_WG = 78, FVM_RS_UU = 79, FVM_RS_UV = 80,
  FVM_RS_UW = 81, FVM_RS_VV = 82, FVM_RS_VW = 83, FVM_RS_WW = 84, FVM_SURF_CF = 85,
  FVM_SURF_CP = 86, FVM_SURF_NU = 87,
  /* the convective history NLX_NM1, NLY_NM1, NLZ_NM1, ADV_NM1 (the previous step's NL) */
  FVM_NLX_NM1 = 88, FVM_NLY_NM1 = 89, FVM_NLZ_NM1 = 90, FVM_ADV_NM1 = 91, FVM_NFIELDS = 92
};

/* GRID + METRIC + the geometric part of INIT (RLES/grid.f90:1-367, RLES/metric.f90:1-538,
 * RLES/init.f90:210-259), on the device.  xg/yg/zg: node coordinates (NI*NJ*NK, I fastest). */
int fvm_create(const fvm_params* prm, const double* xg, const double* yg, const double* zg,
               fvm_ctx** out);
int fvm_destroy(fvm_ctx* ctx);

/* The 2D solver (cylinder/cylinder_v3.0.3_AF_original 4step_face_nl/: GRID grid.f90, METRIC
 * metric.f90, INIT init.f90, MAIN Main.f90:48-110) on a one-cell span slab of the same kernels.
 * xg/yg: NI*NJ node coordinates of the 2D grid file (I fastest); prm->nk and prm->ref_a are
 * ignored (forces per unit span), prm->use_les must be 0 and the wall does not rotate
 * (VEL_ROT = 0).  Field arrays of a 2D context are one (NI+1)(NJ+1) plane; fvm_step skips the
 * pressure reference and fvm_calccoef returns the 2D force coefficients
 * (calccoef.f90:11-144). */
int fvm_create2d(const fvm_params* prm, const double* xg, const double* yg, fvm_ctx** out);

/* Restart read / write of the module arrays (RLES/init.f90:183-208, RLES/writedata.f90:193-216):
 * copies one full cell array host <-> device.  Setting NLX..NLZ sets both the current and the
 * previous-step convective term (the restart holds only the current one). */
int fvm_set_field(fvm_ctx* ctx, int field, const double* host);
int fvm_get_field(fvm_ctx* ctx, int field, double* host);

/* One time step of MAIN (RLES/main.f90:39-150): U_N copy, CONTRA_VEL, NL, STRESS_TENSOR_DY,
 * DIF, P_SOURCE, ITER_U x3, predictor correction, BCU x3, CONTRA_VEL, ITER_P, BCP, P_SOURCE,
 * projection, pressure reference, BCU x3, [ITER_T, BCT,] CALCCOEF.  The passive temperature is
 * advanced only after fvm_set_thermal.  info may be NULL. */
int fvm_step(fvm_ctx* ctx, fvm_step_info* info);

/* Switch on the passive temperature T of the 3D solver: T_N copy, NL(T_N, ADV), the subgrid
 * heat flux of STRESS_TENSOR_DY (stress_tensor_dy.f90:589-738), DIF(T_N, DIFT), ITER_T + BCT
 * (RLES/iter_t.f90:1-247, bct.f90:1-49; main.f90:41-147) and the Nusselt number of CALCCOEF,
 * with T_WALL = 1 and TIN = 0 (comdat_param.f90:39-40) and Peclet number pe (the reference's
 * PE, comdat_param.f90:48).  T and ADV start at zero; set them with fvm_set_field.  Needs the
 * exact K-line stage and the line-scan form of the I and J stages (extruded grids);
 * FVM_ERR_ARG otherwise and for 2D contexts. */
int fvm_set_thermal(fvm_ctx* ctx, double pe);

/* Field output of WRITEDATA (RLES/writedata.f90:8-133, calc_vor.f90:1-141, find_root.f90:1-27):
 * the vorticity magnitude of CALC_VOR, lambda_2 (middle eigenvalue of S^2 + Omega^2, kept where
 * no root lies strictly between the other two, extrapolated to the wall and far-field ghost
 * planes, periodic ghosts), the nodal averages of p, u, v, w, T, |omega| and lambda_2 and the
 * wall y+ that CALCCOEF computes and WRITEDATA writes (calccoef.f90:121, 159; writedata.f90:
 * 14-32), for the current state; read them with fvm_get_field.  3D contexts with NI >= 4. */
int fvm_writedata_fields(fvm_ctx* ctx);

/* Time averages of POST_PROCESS (RLES/post_process.f90:9-276) with their accumulation at the
 * end of CALCCOEF (RLES/calccoef.f90:286-307).  fvm_post_begin is POST_PROCESS's first call: it
 * zeroes the running sums (FVM_SUM_*) and starts the accumulation, after which every fvm_step
 * and fvm_calccoef of a 3D context adds the state it leaves: u, v, w, p, T, the products u_i u_j
 * and, with the LES model on, the SGS stresses -2 nu_t S_ij of the step's STRESS_TENSOR_DY.
 * t_const is the temperature averaged when the passive temperature is off (the field the
 * reference carries unchanged; NULL: zero); it is ignored after fvm_set_thermal.  The caller
 * switches DTIME to DTIME_POST itself with fvm_set_dtime, as POST_PROCESS does
 * (post_process.f90:59).  fvm_post_finalize is its second call: the means, the rms, the
 * Reynolds stresses <u_i u_j> - <u_i><u_j> + <tau_ij>, their node averages and the wall
 * coefficients Cf = 2/sqrt(Re) |du_t/dn|, Cp = 2 p, Nu = -dT/dn from the means
 * (post_process.f90:62-276), read with fvm_get_field; it ends the accumulation and returns
 * the number of accumulated steps in niter_pos (may be NULL).  FVM_ERR_ARG for 2D contexts and
 * for a finalization with no accumulated step. */
int fvm_post_begin(fvm_ctx* ctx, const double* t_const);
int fvm_post_finalize(fvm_ctx* ctx, int* niter_pos);

/* One time step of each of n independent cases (contexts of any size, 2D or 3D), stepped
 * together: the ensemble of 2D cylinder cases of uneven grid size and Reynolds number.  Each
 * case is the time step of fvm_step (2D: MAIN of cylinder/cylinder_v3.0.3_AF_original
 * 4step_face_nl/Main.f90:48-110; the reference has no ensemble driver).  Every context runs on
 * its own stream; nthreads host threads take the next context from a shared counter, so the
 * kernels of different cases overlap on the device while a thread waits for one case's
 * convergence checks.  infos (n entries) may be NULL.  Returns the first error status, else
 * FVM_ERR_NOCONV if any case hit its Poisson sweep cap, else FVM_OK.  The contexts must be
 * distinct and not used by other threads during the call. */
int fvm_ensemble_step(fvm_ctx* const* ctxs, int n, int nthreads, fvm_step_info* infos);

/* ---------------------------------------------------------------------------------------------
 * Batched native 2D solver: many independent 2D cylinder cases of uneven grid size and Reynolds
 * number (the time step of cylinder/cylinder_v3.0.3_AF_original 4step_face_nl/, Main.f90:48-118;
 * the reference has no ensemble driver: it runs one case per process) advanced together.  Every
 * phase of the step is one launch over all cases; the Poisson defect correction runs a fixed
 * number of passes (converged cases skip them), so the whole step is one captured hipGraph,
 * replayed per step.  AF sweeps keep the reference's line order (lines one after the other).
 * Fields are (NI+1)(NJ+1) planes, I fastest, as the 2D reference's RESTART.DAT holds them.
 * ------------------------------------------------------------------------------------------- */
typedef struct fvm_batch2d fvm_batch2d;

typedef struct fvm_case2d {
  int ni, nj;            /* node counts NI (radial), NJ (circumferential)                    */
  double re, dtime;      /* RE, DTIME                                                        */
  const double* xg;      /* NI*NJ node coordinates, I fastest (the 2D grid file's records)  */
  const double* yg;
} fvm_case2d;

/* GRID, METRIC, INIT geometry and the AF / Poisson factors of every case (host set-up in the
 * reference's operations), device planes, block tables and the captured step.  mixed, bit set:
 * 1: the Poisson correction's transforms and mode-line storage in FP32 (residual, metrics, p in
 * FP64); 2: the AF line sweeps of both stages solve for the change of the stage in FP32 against
 * the stage input's FP64 residual (the field stays FP64; fvm_batch2d_adjoint_init refuses such a
 * batch).  npass: Poisson correction passes per step (each case stops once RESIP <= resimax).
 * NJ-1 must be even with prime factors 2, 3, 5, 7, 11, 13; NI-1, NJ-1 <= 512. */
int fvm_batch2d_create(const fvm_case2d* cases, int n, int mixed, double resimax, int npass,
                       fvm_batch2d** out);
int fvm_batch2d_destroy(fvm_batch2d* b);
/* Restart I/O of case c: FVM_U, FVM_V, FVM_P, FVM_NLX, FVM_NLY (set and get); get also
 * FVM_DIFX, FVM_DIFY, FVM_SU, FVM_SV (P_SOURCE of the last pressure), FVM_UN, FVM_VN, the
 * geometry FVM_X, FVM_Y, FVM_DJAC, FVM_XIX, FVM_XIY, FVM_ETX, FVM_ETY, FVM_THETA. */
int fvm_batch2d_set_field(fvm_batch2d* b, int c, int field, const double* host);
int fvm_batch2d_get_field(fvm_batch2d* b, int c, int field, double* host);
int fvm_batch2d_set_dtime(fvm_batch2d* b, int c, double dtime);
/* Wall velocity U_F, V_F of case c per wall cell J (NJ+1 values, J = 1..NJ-1 used): the hook
 * of BCU/BCV (bcu.f90:10-20, bcv.f90:10-20; zero: the reference's fixed wall). */
int fvm_batch2d_set_wall_velocity(fvm_batch2d* b, int c, const double* uf, const double* vf);
/* nsteps time steps of every case; infos (n entries, may be NULL): the last step's RESIP,
 * Poisson passes, CFL, C_pb, C_D, C_L and parts.  FVM_ERR_NAN / FVM_ERR_NOCONV as fvm_step. */
int fvm_batch2d_step(fvm_batch2d* b, int nsteps, fvm_step_info* infos);
/* The last nsteps steps' C_D, C_L, C_D pressure / friction, C_L pressure / friction, RESIP, CFL
 * per case, oldest first: out[(s n + c) 8 + q]. */
int fvm_batch2d_history(fvm_batch2d* b, int nsteps, double* out);
/* Tile choices {stencil tile height 1|2|4|8, FFT columns per workgroup 1..16}; the autotuner times
 * every admissible pair with reps steps each (state saved and restored) and keeps the fastest;
 * ms_out (20 values, tile height major, may be NULL) receives the timings. */
int fvm_batch2d_set_tuning(fvm_batch2d* b, const int* choice);
int fvm_batch2d_get_tuning(fvm_batch2d* b, int* choice);
int fvm_batch2d_autotune(fvm_batch2d* b, int reps, int* choice, double* ms_out);
/* nsteps steps launched eagerly (the captured step's launches) with HIP events around each
 * launch: per kernel the total ms and launch count, up to cap entries, names 48 chars each.
 * Returns the number of kernels (negative: an error status). */
int fvm_batch2d_profile(fvm_batch2d* b, int nsteps, char* names, double* ms, long* calls,
                        int cap);

/* Discrete adjoint of the batched step (the reference has no adjoint: these serve gradient-based
 * flow control on top of its time step).  adjoint_init allocates the adjoint planes (once) and
 * zeroes the adjoint state and the wall-velocity gradients.  adjoint_step maps the adjoint of
 * the state after a step to the adjoint before it; the batch's state must be the state BEFORE
 * that step (recomputed from a checkpoint).  weights (n values): the weight of each case's C_D
 * after the step in the objective.  The adjoint Poisson equation is solved to the scaled
 * residual tol (at most maxpass corrections; ares, may be NULL, gets the worst case's residual).
 * adjoint_wall: the accumulated d(objective)/d(U_F, V_F) of case c per wall cell J (NJ+1
 * values).  adjoint_get: the adjoint state of case c (0 u, 1 v, 2 p, 3 NLX, 4 NLY).
 * save / load: device snapshots of every case's u, v, p, NLX, NLY in numbered slots (the
 * checkpoints of the recompute). */
int fvm_batch2d_adjoint_init(fvm_batch2d* b);
int fvm_batch2d_adjoint_step(fvm_batch2d* b, const double* weights, double tol, int maxpass,
                             double* ares);
int fvm_batch2d_adjoint_wall(fvm_batch2d* b, int c, double* gu, double* gv);
int fvm_batch2d_adjoint_get(fvm_batch2d* b, int c, int which, double* host);
int fvm_batch2d_save(fvm_batch2d* b, int slot);
int fvm_batch2d_load(fvm_batch2d* b, int slot);

/* Host-only (no device call): the 2D set-up arrays X, Y, DJAC, XIX, XIY, ETX, ETY, THETA,
 * DJAC Q12 of one case (9 planes of (NI+1)(NJ+1)) (grid.f90:47-115, metric.f90:18-145,
 * init.f90:71-99). */
int fvm_geometry2d(int ni, int nj, const double* xg, const double* yg, double* out);

/* Change DTIME between steps.  DTIME is a run-time module variable of the reference, reassigned
 * by POST_PROCESS (RLES/post_process.f90:59) and by the CFL control of CALCCOEF
 * (RLES/calccoef.f90:27-33, commented out there); the dt-dependent coefficient 2 Re J^-1/DTIME
 * is recomputed.  The convective history NL is kept, as on a restart with a new DTIME. */
int fvm_set_dtime(fvm_ctx* ctx, double dtime);

/* The subroutines of the time step, for drivers that keep MAIN's call sequence. */
int fvm_begin_step(fvm_ctx* ctx);                 /* U_N = U, T_N = T (main.f90:41-45)      */
int fvm_contra_vel(fvm_ctx* ctx);                 /* CONTRA_VEL (RLES/contra_vel.f90:1-58)   */
int fvm_nl(fvm_ctx* ctx);                         /* NL_NM1 = NL; NL(U_N..W_N) (nl.f90:1-33) */
int fvm_stress_tensor_dy(fvm_ctx* ctx, double* c_avg); /* SGS + NL += J SGS (main.f90:59-69)
                                                     (RLES/stress_tensor_dy.f90:1-586)      */
int fvm_dif(fvm_ctx* ctx);                        /* DIF(U_N..W_N) (RLES/dif.f90:1-25)       */
int fvm_p_source(fvm_ctx* ctx);                   /* P_SOURCE(P) (RLES/p_source.f90:1-44)    */
int fvm_iter_u(fvm_ctx* ctx);                     /* ITER_U for U, V, W (iter_u.f90:1-256)   */
int fvm_iter_t(fvm_ctx* ctx);                     /* ITER_T + BCT(T,T_N) (iter_t.f90:1-247,
                                                     bct.f90:1-49, main.f90:145-147)        */
int fvm_correct(fvm_ctx* ctx, double sign);       /* U += sign*DTIME*S/J, then BCU x3
                                                     (RLES/main.f90:90-98, 108-111, 137-142) */
int fvm_iter_p(fvm_ctx* ctx, double* resip, int* niter); /* ITER_P + BCP (iter_p.f90:1-66,
                                                     lisolv.f90:1-57, bcp.f90:1-38)         */
int fvm_pressure_reference(fvm_ctx* ctx, double* pref); /* RLES/main.f90:125-135          */
int fvm_calccoef(fvm_ctx* ctx, fvm_step_info* info);    /* CALCCOEF (calccoef.f90:1-201)  */
/* The discrete divergence of the face volume fluxes of the current velocity (CONTRA_VEL's UE..WB,
 * RLES/contra_vel.f90:13-56, summed as ITER_P's source, iter_p.f90:31-34), reported beside RESIP:
 * out[0] max |div|, out[1] mean |div| over the interior cells, out[2] sum |div| / sum (|UE| + |VN|
 * + |WT|).  No counterpart in the reference (it prints RESIP only); recomputes the context's face
 * fluxes, which the next step recomputes anyway. */
int fvm_divergence(fvm_ctx* ctx, double* out);

/* fvm_calccoef without any device allocation: the same kernels and results, with the CFL and
 * force partial sums in a static device array of 1 MiB (131072 doubles) per GPU instead of the
 * context's reduction buffer.  No counterpart in the reference; it is for callers that cannot
 * allocate device memory at call time (graph capture, tight serving loops, harnesses that forbid
 * hipMalloc).  Grids needing more than the array holds (about 3.3 million wall cells) return
 * FVM_ERR_SIZE.  Every call on one GPU shares the array: calls that overlap on one GPU race. */
int fvm_calccoef_noalloc(fvm_ctx* ctx, fvm_step_info* info);

/* The reference's subroutines one array at a time, with their own argument lists: each array
 * argument is a field selector (the device array standing for the reference's module array).
 * For a driver that keeps the reference's MAIN statement for statement (the Fortran shims of
 * fvm_amd/fortran/fvm_shims.f90 bind the reference's names to these):
 *   fvm_copy_field   MAIN's whole-array assignments: U_N = U (RLES/main.f90:41-44),
 *                    NLX_NM1 = NLX (main.f90:51-54)
 *   fvm_nl_c         NL(IN, OUT) (nl.f90:1-49): OUT from IN and CONTRA_VEL's face fluxes
 *   fvm_dif_c        DIF(IN, OUT) (dif.f90:1-40, dif_diagonal.f90)
 *   fvm_stress_tensor_dy_c  STRESS_TENSOR_DY (stress_tensor_dy.f90): SGS_X..Z and NU_T, NL
 *                    unchanged; returns the dynamic coefficient
 *   fvm_add_sgs_c    NLX = NLX + SGS_X * DJAC (main.f90:61-66)
 *   fvm_iter_u_c     ITER_U(IN, INM, INA, INB, INC, IND) (iter_u.f90:1-262): IN = U, V or W
 *                    (its BCU far-field value is the UIN MAIN sets for it: 1 for U, 0 else);
 *                    the 2D context's ITER_U / ITER_V (2D iter_u.f90, iter_v.f90) likewise
 *   fvm_correct_c    U = U + sign DTIME S_U / DJAC (main.f90:90-92, 108-111)
 *   fvm_bcu          BCU(A, B) (bcu.f90:1-50): A = U, V or W, B its step-start copy
 *   fvm_bcp          BCP (bcp.f90:1-38)
 * P_SOURCE(P), CONTRA_VEL, ITER_P, CALCCOEF are fvm_p_source, fvm_contra_vel, fvm_iter_p,
 * fvm_calccoef below. */
int fvm_copy_field(fvm_ctx* ctx, int dst, int src);
int fvm_nl_c(fvm_ctx* ctx, int in, int out);
int fvm_dif_c(fvm_ctx* ctx, int in, int out);
int fvm_stress_tensor_dy_c(fvm_ctx* ctx, double* c_avg);
int fvm_add_sgs_c(fvm_ctx* ctx, int nl, int sgs);
int fvm_iter_u_c(fvm_ctx* ctx, int in, int inm, int ina, int inb, int inc, int ind);
int fvm_correct_c(fvm_ctx* ctx, int f, int s, double sign);
int fvm_bcu(fvm_ctx* ctx, int a, int b);
int fvm_bcp(fvm_ctx* ctx);

/* Kernel timing (HIP events on the context's stream).  fvm_profile_get fills up to cap
 * entries of names (each <= 47 chars) / total ms / call counts and returns the count. */
int fvm_profile_enable(fvm_ctx* ctx, int on);
int fvm_profile_get(fvm_ctx* ctx, char* names, double* ms, long* calls, int cap);
int fvm_profile_reset(fvm_ctx* ctx);
int fvm_sync(fvm_ctx* ctx);

#ifdef __cplusplus
}
#endif
#endif
