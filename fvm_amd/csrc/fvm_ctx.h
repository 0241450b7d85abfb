// Device context of the solver: every field, metric and scratch array lives in HBM for the
// whole run (allocated once at fvm_create, freed at fvm_destroy), all work is ordered on one
// HIP stream owned by the context.
#pragma once
#include <map>
#include <string>
#include <vector>

#include "fvm_common.h"
#include "../../include/fvm_capi.h"

struct ProfAcc {
  double ms = 0.0;
  long calls = 0;
};

struct FvmCtx {
  fvm_params prm;
  Dims d;
  int dev = 0;
  hipStream_t stream = nullptr;
  double pi = 0.0;

  // geometry (node arrays use the cell layout and hold nodes 1..NI, 1..NJ, 1..NK)
  double *xg = nullptr, *yg = nullptr, *zg = nullptr, *x = nullptr, *y = nullptr, *z = nullptr;
  double* djac = nullptr;
  double* m[9] = {};  // metric grouped per Cartesian component s: (xi_s, eta_s, zeta_s), s=x,y,z
  double** dm = nullptr;    // the same nine pointers as a device-resident table
  double *jq11 = nullptr, *jq22 = nullptr, *jq33 = nullptr;  // J^-1 Q_ii
  double *jq12 = nullptr, *jq13 = nullptr, *jq23 = nullptr;  // J^-1 Q_ij (cross terms)
  double* theta = nullptr;
  double* filt2 = nullptr;  // test-filter width squared, (J^-1)^(2/3)
  double* ap0 = nullptr;    // 2 Re J^-1 / dt
  double* fa[3][3] = {};    // AF face coefficients per stage s and direction: I, J, K faces
  double** dfa[3] = {};     // device tables of fa[s]
  double *pdI = nullptr, *pdJ = nullptr, *pdK = nullptr;  // Poisson face coefficients D_E, D_N, D_T
  double* wall = nullptr;   // per wall triangle: A, B, C, |n|, d_n, area (6 x 2 x (NJ-1)(NK-1))

  // fields
  double *u = nullptr, *v = nullptr, *w = nullptr, *p = nullptr;
  double *un = nullptr, *vn = nullptr, *wn = nullptr;
  // u, v, w still equal U_N, V_N, W_N bit for bit (set by the step-start copy, cleared by every
  // entry point that may write them): the first AF stage then takes U_N from its staged input
  bool u_is_un = false;
  bool ufar_stale = false;  // U_N changed since the far-field values were formed
  double* nl[3] = {};
  double* nl1[3] = {};
  double* dif[3] = {};
  double* s[3] = {};
  double* su[3] = {};
  double* ju[3] = {};   // J^-1 x contravariant velocity (volume-flux base)
  double* ufar = nullptr;  // far-field boundary value per component, (NJ+1)(NK+1) each
  double* rhs = nullptr;   // Poisson right-hand side div/dt
  double* psu = nullptr;   // Poisson explicit source rhs - cross terms
  float* p32 = nullptr;    // FP32 correction of the mixed-precision Poisson solve
  float* su32 = nullptr;   // FP32 defect
  // LES
  double* lss[6] = {};
  double** dlss = nullptr;  // device table of lss
  double* mag = nullptr;
  double* uh[3] = {};
  double* nut = nullptr;
  double* sgs[3] = {};
  bool sgs_stale = false;   // sgs[] not yet formed from the last SGS pass's nu_t and S_ij
  bool nut_stale = false;   // nut not yet formed from the last SGS pass's C, Delta^2 and |S|
  // AF chunked iteration: two iterate buffers per component, device table of prev/next
  double* afb[3][2] = {};
  double** afptr = nullptr;
  int af_iters[3] = {};
  // stage couplings that vanish identically on this grid (a: along the sweep, b: across it)
  bool af_zero_a[3] = {}, af_zero_b[3] = {};
  double af_b_ratio[3] = {};
  bool xz_zero = false;       // xi_z = eta_z = 0 on every cell (metric kernels skip them)
  bool dd2_zero = false;      // xi_z = eta_z = 0 around every cell: stage-3 cross terms vanish  // max over cells of (|b couplings|) / (2 Re J^-1/dt)
  // Thomas factors A, D, B of the K-line stage when its lines are independent
  double *tk_a = nullptr, *tk_d = nullptr, *tk_b = nullptr;
  // line-scan form of the I and J stages when the K couplings vanish: per stage the line
  // factors (beta = -A/D, 1/D, B, coupling to the previous line) and the chunk boundary lines
  bool af_scan[2] = {};
  double* lf[2][4] = {};
  float* lf32[2][4] = {};   // FP32 copies of lf (af_mode 2: FP32 change form)
  double* bnd[2] = {};
  long nbnd = 0;
  // reductions
  double* red = nullptr;     // partial sums
  double** scantab[2] = {nullptr, nullptr};  // device pointer tables of the I/J scan stages
  double* scantab_h[2][9] = {};               // their last uploaded contents
  double* afp[2][6] = {};                     // metric products of the stage 0/1 cross terms
  double** dafp[2] = {nullptr, nullptr};      // their device pointer tables
  double* cflpart = nullptr; // CFL maxima of the projection kernel (one per workgroup)
  int ncflpart = 0;          // capacity of cflpart
  int ncflused = 0;          // entries the last projection wrote
  double* scal = nullptr;    // device scalars (c_avg, residual sums, pref, ...)
  double* sink = nullptr;    // 64 doubles that lanes past a line's end store to (never read)
  double* hscal = nullptr;   // pinned host mirror of scal
  // host mailbox: fine-grained pinned memory that k_publish writes scalars and then a sequence
  // number into, so the host reads a convergence test's inputs without a copy kernel or a
  // stream synchronisation (fvm_fetch)
  double* hmail = nullptr;
  unsigned long long* hseq = nullptr;
  unsigned long long seq = 0;

  int nred = 0;
  // separable approximate inverse of the pressure operator (fvm_poisson_spec.hip)
  void* sep = nullptr;
  double sep_dev = 0.0;     // largest relative deviation of a face coefficient from its J,K mean
  long sep_fallbacks = 0;   // FP32 passes handed over to the FP64 inverse
  // 2D solver on a one-cell span slab
  bool two_d = false;
  double* wall2d = nullptr;  // per wall cell J: length, DN1, sin TH2, cos TH2, cos th, sin th
  // passive temperature (ITER_T, BCT, SGS_T; RLES/iter_t.f90, bct.f90,
  // stress_tensor_dy.f90:589-738), allocated by fvm_set_thermal
  bool thermal = false;
  double pe = 0.0;
  double *t = nullptr, *tn = nullptr, *adv = nullptr, *adv1 = nullptr, *dift = nullptr;
  double *sgst = nullptr, *that = nullptr, *dtd[3] = {}, *kappa = nullptr, *magh = nullptr;
  double *tfar = nullptr;   // far-field boundary value of T, (NJ+1)(NK+1)
  double *ap0t = nullptr;   // 2 Pe J^-1 / dt
  double *tk_at = nullptr, *tk_dt = nullptr, *tk_bt = nullptr;  // K-line factors with ap0t
  double* lft[2][4] = {};   // I/J line-scan factors with ap0t
  double** scantab_t[2] = {nullptr, nullptr};
  double* scantab_th[2][9] = {};
  double sum_area = 0.0;    // SUM_AREA of the wall triangles (init.f90:240-261)
  // field output of WRITEDATA (allocated on first use): |omega|, lambda_2 on the cells and the
  // nodal averages of p, u, v, w, T, |omega|, lambda_2 at the node's cell index
  double *vorm = nullptr, *lam2 = nullptr;
  double* nodal[7] = {};
  double* yplus = nullptr;  // wall y+ of CALCCOEF on the I = 1 cells (zero elsewhere)
  // time averages: 17 running sums per cell (u v w p T, uu uv uw vv vw ww, SGS 11 12 13 22 23
  // 33), the cell rms of the finalization, 14 node averages and the wall Cf, Cp, Nu
  double* post_sum[17] = {};
  double* post_rms[3] = {};
  double* post_node[14] = {};
  double* post_surf[3] = {};
  double* post_t = nullptr;  // averaged temperature without the passive temperature model
  bool post_on = false;
  long niter_pos = 0;
  // last-step diagnostics
  fvm_step_info last{};
  // profiling
  bool prof_on = false;
  std::map<std::string, ProfAcc> prof;
  std::vector<hipEvent_t> ev_pool;
  std::vector<std::pair<std::string, std::pair<hipEvent_t, hipEvent_t>>> ev_pending;
  size_t ev_next = 0;
  std::vector<void*> allocs;
};

// Kernel timing: wraps one launch in a pair of events when profiling is on.
// arguments of the AF stage kernels
struct AFArgs {
  int nc = 3;             // components: 3 (u, v, w) or 1 (T)
  double* f[3];           // stage input / output field per component
  const double* fn[3];    // step-start field U_N per component
  const double* nl[3];
  const double* nl1[3];
  const double* dif[3];
  const double* s[3];
  double* su[3];
  const double* ufar;     // far-field boundary values, (NJ+1)(NK+1) per component
};

struct ProfScope {
  FvmCtx* c;
  const char* name;
  hipEvent_t a = nullptr, b = nullptr;
  ProfScope(FvmCtx* ctx, const char* nm);
  ~ProfScope();
};

// stage launchers (return FVM_OK or an error code)
int fvm_fetch(FvmCtx* c, const double* src, int n, double* dst);
Mailbox fvm_mail(FvmCtx* c);  // a fresh sequence number for a kernel that publishes itself
int fvm_mail_wait(FvmCtx* c, const Mailbox& mb, int n, double* dst);
int fvm_geom_setup(FvmCtx* c);
int fvm_k_set_dtime(FvmCtx* c, double dtime);
int fvm_af_setup(FvmCtx* c);
int fvm_af_scan_setup(FvmCtx* c);
int fvm_af_scan_stage(FvmCtx* c, int s, const AFArgs& A, int* iters, bool temp = false,
                      bool delta = false);
int fvm_ghost_fill(FvmCtx* c, double* a);
int fvm_ghost_fill_n(FvmCtx* c, double** a, int cnt);
int fvm_k_contra_vel(FvmCtx* c);
int fvm_k_nl(FvmCtx* c);
int fvm_k_nl_dif(FvmCtx* c);
int fvm_k_dif(FvmCtx* c);
int fvm_k_p_source(FvmCtx* c);
int fvm_k_ufar(FvmCtx* c);
int fvm_k_bcu_ghosts(FvmCtx* c, double** f, int cnt);
int fvm_k_bct_ghosts(FvmCtx* c, double* f);
int fvm_k_tfar(FvmCtx* c);
int fvm_k_step_copy(FvmCtx* c);
int fvm_k_iter_t(FvmCtx* c);
int fvm_thermal_setup(FvmCtx* c);
int fvm_k_writedata_fields(FvmCtx* c);
int fvm_k_post_begin(FvmCtx* c, const double* t_const);
int fvm_k_post_accum(FvmCtx* c);
int fvm_k_post_finalize(FvmCtx* c);
int fvm_k_correct(FvmCtx* c, double sign);
// one component at a time (the per-subroutine entry points)
int fvm_k_nl_one(FvmCtx* c, const double* in, double* out);
int fvm_k_dif_one(FvmCtx* c, const double* in, double* out);
int fvm_k_correct_one(FvmCtx* c, double* f, const double* s, double sign);
int fvm_k_add_sgs(FvmCtx* c, double* nl, const double* sgs);
int fvm_k_bcu_ghosts_one(FvmCtx* c, double* f, int comp);
int fvm_k_iter_u_one(FvmCtx* c, int comp, double* f, const double* fn, const double* nl,
                     const double* nl1, const double* dif, const double* s);
double* fvm_field(FvmCtx* c, int f);
int fvm_k_psource_correct(FvmCtx* c, double sign, bool with_cfl);
int fvm_k_correct_bcu_contra(FvmCtx* c);
int fvm_k_pref(FvmCtx* c, double* pref_host);
// CALCCOEF; part (cap doubles) holds the CFL and force partial sums, nullptr: the context's own
// reduction buffer.  fvm_calccoef_scratch: the doubles of part one call needs.
int fvm_k_calccoef(FvmCtx* c, double* out8, bool cfl_ready = false,
                   const Mailbox* mb = nullptr, double* part = nullptr, long cap = 0);
int fvm_k_calccoef2d(FvmCtx* c, double* out8, bool cfl_ready = false,
                     const Mailbox* mb = nullptr, double* part = nullptr, long cap = 0);
long fvm_calccoef_scratch(const FvmCtx* c, bool cfl_ready);
int fvm_calccoef_info(FvmCtx* c, const double* o8, fvm_step_info* info);
int fvm_wall2d_setup(FvmCtx* c, const double* xg, const double* yg);
int fvm_k_sgs(FvmCtx* c, bool sgs_out = true, bool nl_update = true);
int fvm_k_sgs_out(FvmCtx* c);
int fvm_k_nut_out(FvmCtx* c);
int fvm_k_iter_u(FvmCtx* c);
int fvm_k_iter_p(FvmCtx* c, double* resip, int* niter);
int fvm_k_div_norm(FvmCtx* c, double* out);
int fvm_k_bcp(FvmCtx* c);
int fvm_sep_setup(FvmCtx* c);
void fvm_sep_destroy(FvmCtx* c);
int fvm_k_iter_p_sep(FvmCtx* c, bool mixed, double* resip, int* niter);
