// C ABI of the solver: context life cycle, restart I/O, the time-step sequence and kernel
// timing.  See include/fvm_capi.h for the contract.
#include <math.h>
#include <string.h>

#include <algorithm>
#include <atomic>
#include <thread>
#include <vector>

#include "fvm_ctx.h"

// Copies n <= 48 device scalars into the host mailbox, then (after a system-scope fence) the
// sequence number s: a host that sees s also sees the scalars.  One thread; vector stores.
__global__ void k_publish(const double* __restrict__ src, int n, volatile double* dst,
                          volatile unsigned long long* seq, unsigned long long s) {
  if (threadIdx.x != 0) return;
  for (int q = 0; q < n; ++q) dst[q] = src[q];
  __threadfence_system();
  *seq = s;
}

Mailbox fvm_mail(FvmCtx* c) {
  return Mailbox{c->hmail, c->hseq, ++c->seq};
}

// Spins on the mailbox's sequence number until mb's kernel has published, then copies its n
// scalars to dst.  The stream keeps running (no synchronisation), so the device idles only for
// the host's reaction and the next launch.  A stream error or a drained stream without the
// expected number ends the wait with FVM_ERR_HIP.
int fvm_mail_wait(FvmCtx* c, const Mailbox& mb, int n, double* dst) {
  volatile unsigned long long* f = c->hseq;
  for (long it = 1; *f < mb.s; ++it) {
    if ((it & 1023) == 0) {
      const hipError_t e = hipStreamQuery(c->stream);
      if (e == hipSuccess) {
        if (*f >= mb.s) break;
        fprintf(stderr, "fvm_amd: stream drained without publishing scalar set %llu\n", mb.s);
        return FVM_ERR_HIP;
      }
      if (e != hipErrorNotReady) {
        fprintf(stderr, "fvm_amd: HIP error %s while waiting for scalars\n", hipGetErrorString(e));
        return FVM_ERR_HIP;
      }
    }
  }
  std::atomic_thread_fence(std::memory_order_acquire);
  const volatile double* m = c->hmail;
  for (int q = 0; q < n; ++q) dst[q] = m[q];
  return FVM_OK;
}

// The n device scalars at src, in stream order, into dst on the host (k_publish, then the wait).
int fvm_fetch(FvmCtx* c, const double* src, int n, double* dst) {
  if (n < 1 || n > 48) return FVM_ERR_ARG;
  const Mailbox mb = fvm_mail(c);
  hipLaunchKernelGGL(k_publish, dim3(1), dim3(64), 0, c->stream, src, n, c->hmail, c->hseq, mb.s);
  HIP_CHECK(hipGetLastError());
  return fvm_mail_wait(c, mb, n, dst);
}

ProfScope::ProfScope(FvmCtx* ctx, const char* nm) : c(ctx), name(nm) {
  if (!c->prof_on) return;
  if (hipEventCreate(&a) != hipSuccess || hipEventCreate(&b) != hipSuccess) { a = b = nullptr; return; }
  (void)hipEventRecord(a, c->stream);
}

ProfScope::~ProfScope() {
  if (!c->prof_on || !a) return;
  (void)hipEventRecord(b, c->stream);
  c->ev_pending.push_back({name, {a, b}});
}

static void prof_flush(FvmCtx* c) {
  if (c->ev_pending.empty()) return;
  (void)hipStreamSynchronize(c->stream);
  for (auto& e : c->ev_pending) {
    float ms = 0.f;
    (void)hipEventElapsedTime(&ms, e.second.first, e.second.second);
    ProfAcc& acc = c->prof[e.first];
    acc.ms += ms;
    acc.calls += 1;
    (void)hipEventDestroy(e.second.first);
    (void)hipEventDestroy(e.second.second);
  }
  c->ev_pending.clear();
}

static int dalloc(FvmCtx* c, double** p, long n) {
  if (hipMalloc((void**)p, n * sizeof(double)) != hipSuccess) return FVM_ERR_HIP;
  c->allocs.push_back(*p);
  if (hipMemsetAsync(*p, 0, n * sizeof(double), c->stream) != hipSuccess) return FVM_ERR_HIP;
  return FVM_OK;
}

// Passive temperature: arrays (once), 2 Pe J^-1/dt, SUM_AREA of the wall triangles in the
// reference's order (init.f90:240-259) and the temperature's line factors.
int fvm_thermal_setup(FvmCtx* c) {
  const Dims d = c->d;
  const long n = d.n;
  if (!c->t) {
    int r = 0;
    double** arr[] = {&c->t, &c->tn, &c->adv, &c->adv1, &c->dift, &c->sgst, &c->that,
                      &c->dtd[0], &c->dtd[1], &c->dtd[2], &c->kappa, &c->magh, &c->ap0t};
    // T, like u, v, w, may pass through the AF work arrays (padded J-line-major size)
    const long nt = std::max((long)n, scan_jline_elems(d.ni, d.nj, d.nk));
    for (double** q : arr) r |= dalloc(c, q, q == &c->t ? nt : n);
    r |= dalloc(c, &c->tfar, (long)(d.nj + 1) * (d.nk + 1));
    for (int s = 0; s < 2; ++s) {
      if (hipMalloc((void**)&c->scantab_t[s], 9 * sizeof(double*)) != hipSuccess) r = FVM_ERR_HIP;
      else c->allocs.push_back(c->scantab_t[s]);
    }
    if (r) return FVM_ERR_HIP;
    const long nw = (long)(d.nj - 1) * (d.nk - 1);
    std::vector<double> w(12 * nw);
    HIP_CHECK(hipMemcpy(w.data(), c->wall, w.size() * sizeof(double), hipMemcpyDeviceToHost));
    double sa = 0.0;
    for (long t = 0; t < nw; ++t) sa = sa + w[6 * t + 5] + w[6 * (nw + t) + 5];
    c->sum_area = sa;
  }
  c->thermal = true;
  int r = fvm_k_set_dtime(c, c->prm.dtime);  // ap0t and every factor set
  if (r) return r;
  // the temperature uses the exact K-line stage and the line-scan form of the I and J stages
  const double ratio = c->prm.re / c->pe;
  const bool ok = (c->prm.af_mode == 0 || c->prm.af_mode == 2) && c->tk_at && c->af_zero_a[2] && c->af_zero_b[2] &&
                  c->af_scan[0] && c->af_scan[1] && c->af_b_ratio[0] * ratio <= 1e-15 &&
                  c->af_b_ratio[1] * ratio <= 1e-15 && c->lft[0][0] && c->lft[1][0];
  if (!ok) {
    fprintf(stderr, "fvm_amd: the passive temperature needs an extruded grid (independent K "
                    "lines, negligible K couplings of the I and J stages) and af_mode 0 or 2\n");
    c->thermal = false;
    return FVM_ERR_ARG;
  }
  return FVM_OK;
}

static double* field_ptr(FvmCtx* c, int f) {
  switch (f) {
    case FVM_U: return c->u;
    case FVM_V: return c->v;
    case FVM_W: return c->w;
    case FVM_P: return c->p;
    case FVM_NLX: return c->nl[0];
    case FVM_NLY: return c->nl[1];
    case FVM_NLZ: return c->nl[2];
    case FVM_DIFX: return c->dif[0];
    case FVM_DIFY: return c->dif[1];
    case FVM_DIFZ: return c->dif[2];
    case FVM_SU: return c->s[0];
    case FVM_SV: return c->s[1];
    case FVM_SW: return c->s[2];
    case FVM_DJAC: return c->djac;
    case FVM_XIX: return c->m[0];
    case FVM_ETX: return c->m[1];
    case FVM_ZTX: return c->m[2];
    case FVM_XIY: return c->m[3];
    case FVM_ETY: return c->m[4];
    case FVM_ZTY: return c->m[5];
    case FVM_XIZ: return c->m[6];
    case FVM_ETZ: return c->m[7];
    case FVM_ZTZ: return c->m[8];
    case FVM_X: return c->x;
    case FVM_Y: return c->y;
    case FVM_Z: return c->z;
    case FVM_THETA: return c->theta;
    case FVM_NUT: return c->nut;
    case FVM_SGSX: return c->sgs[0];
    case FVM_SGSY: return c->sgs[1];
    case FVM_SGSZ: return c->sgs[2];
    case FVM_UN: return c->un;
    case FVM_VN: return c->vn;
    case FVM_WN: return c->wn;
    case FVM_T: return c->t;
    case FVM_ADV: return c->adv;
    case FVM_DIFT: return c->dift;
    case FVM_SGST: return c->sgst;
    case FVM_KAPPA: return c->kappa;
    case FVM_THAT: return c->that;
    case FVM_DTDX: return c->dtd[0];
    case FVM_DTDY: return c->dtd[1];
    case FVM_DTDZ: return c->dtd[2];
    case FVM_TN: return c->tn;
    case FVM_VORM: return c->vorm;
    case FVM_LAMDA2: return c->lam2;
    case FVM_PG: case FVM_UG: case FVM_VG: case FVM_WG: case FVM_TG: case FVM_VORG:
    case FVM_LAMDAG: return c->nodal[f - FVM_PG];
    case FVM_YPLUS: return c->yplus;
    case FVM_NLX_NM1: return c->nl1[0];
    case FVM_NLY_NM1: return c->nl1[1];
    case FVM_NLZ_NM1: return c->nl1[2];
    case FVM_ADV_NM1: return c->adv1;
    default:
      if (f >= FVM_SUM_U && f <= FVM_SUM_SGS33) return c->post_sum[f - FVM_SUM_U];
      if (f >= FVM_AVG_PG && f <= FVM_RS_WW) return c->post_node[f - FVM_AVG_PG];
      if (f >= FVM_SURF_CF && f <= FVM_SURF_NU) return c->post_surf[f - FVM_SURF_CF];
      return nullptr;
  }
}

double* fvm_field(FvmCtx* c, int f) { return field_ptr(c, f); }

// The far-field values of BCU are formed from U_N; after U_N changed through the
// per-subroutine entry points they are formed again before the next BCU.
static int ufar_fresh(FvmCtx* c) {
  if (!c->ufar_stale) return FVM_OK;
  c->ufar_stale = false;
  return fvm_k_ufar(c);
}

// velocity component (0 u, 1 v, 2 w) of a field selector, -1 for other fields
static int vel_comp(int f) {
  return f == FVM_U ? 0 : (f == FVM_V ? 1 : (f == FVM_W ? 2 : -1));
}

// The rest of CALCCOEF after the force kernels: the time-average accumulation and the scalars
// of info from the eight coefficients o8 and the step's dynamic coefficients.
int fvm_calccoef_info(FvmCtx* c, const double* o, fvm_step_info* info) {
  int r = fvm_k_post_accum(c);  // calccoef.f90:286-307 while the time averages run
  if (r) return r;
  if (info) {
    info->cfl = o[0]; info->cpb = o[1]; info->cd = o[2]; info->cl = o[3];
    info->cd_p = o[4]; info->cd_v = o[5]; info->cl_p = o[6]; info->cl_v = o[7];
    // the dynamic coefficients of this step's STRESS_TENSOR_DY and the Nusselt number
    double sc[2] = {0.0, 0.0};
    HIP_CHECK(hipMemcpy(sc, c->scal, 2 * sizeof(double), hipMemcpyDeviceToHost));
    info->c_avg = c->prm.use_les ? sc[0] : 0.0;
    info->c_t = (c->thermal && c->prm.use_les) ? sc[1] : 0.0;
    info->nu = 0.0;
    if (c->thermal && !c->two_d) {
      HIP_CHECK(hipMemcpy(&info->nu, c->scal + 16, sizeof(double), hipMemcpyDeviceToHost));
    }
  }
  return FVM_OK;
}

extern "C" {

static int create_impl(const fvm_params* prm, const double* xg, const double* yg,
                       const double* zg, int min_nk, fvm_ctx** out);

int fvm_create(const fvm_params* prm, const double* xg, const double* yg, const double* zg,
               fvm_ctx** out) {
  return create_impl(prm, xg, yg, zg, 3, out);
}

// The 2D solver (cylinder/cylinder_v3.0.3_AF_original 4step_face_nl/) on a one-cell span slab:
// nodes at z = 0 and z = LZ2D with periodic span ghosts.  Every span term then vanishes (equal
// neighbours, zeta_x = zeta_y = xi_z = eta_z = 0) and the span face coefficients are ~1/LZ2D^2
// of the others, so the step reduces to the 2D scheme; the span length scales J^-1, the
// fluxes and the operators alike.  No pressure reference is subtracted (the 2D MAIN has none)
// and CALCCOEF follows the 2D definitions.
static const double LZ2D = 1.0e6;

int fvm_create2d(const fvm_params* prm, const double* xg, const double* yg, fvm_ctx** out) {
  if (!prm || !xg || !yg || !out) return FVM_ERR_ARG;
  if (prm->use_les) return FVM_ERR_ARG;
  const int ni = prm->ni, nj = prm->nj;
  if (ni < 4 || nj < 4) return FVM_ERR_ARG;
  const long nn = (long)ni * nj;
  std::vector<double> x3(2 * nn), y3(2 * nn), z3(2 * nn);
  for (int k = 0; k < 2; ++k)
    for (long q = 0; q < nn; ++q) {
      x3[q + k * nn] = xg[q];
      y3[q + k * nn] = yg[q];
      z3[q + k * nn] = k == 0 ? 0.0 : LZ2D;
    }
  fvm_params p3 = *prm;
  p3.nk = 2;
  p3.ref_a = LZ2D;  // forces per unit span
  fvm_ctx* c = nullptr;
  int r = create_impl(&p3, x3.data(), y3.data(), z3.data(), 2, &c);
  if (r) return r;
  c->two_d = true;
  r = fvm_wall2d_setup(c, xg, yg);
  if (r) { fvm_destroy(c); return r; }
  *out = c;
  return FVM_OK;
}

static int create_impl(const fvm_params* prm, const double* xg, const double* yg,
                       const double* zg, int min_nk, fvm_ctx** out) {
  if (!prm || !xg || !yg || !zg || !out) return FVM_ERR_ARG;
  if (prm->ni < 4 || prm->nj < 4 || prm->nk < min_nk) return FVM_ERR_ARG;
  if (prm->ni - 1 > FVM_MAX_LINE || prm->nj - 1 > FVM_MAX_LINE || prm->nk - 1 > FVM_MAX_LINE)
    return FVM_ERR_SIZE;
  if (!(prm->re > 0) || !(prm->dtime > 0) || prm->nswpp < 1) return FVM_ERR_ARG;
  FvmCtx* c = new FvmCtx();
  c->prm = *prm;
  if (c->prm.check_every < 1) c->prm.check_every = 1;
  c->d.ni = prm->ni; c->d.nj = prm->nj; c->d.nk = prm->nk;
  c->d.sj = prm->ni + 1;
  c->d.sk = (long)(prm->ni + 1) * (prm->nj + 1);
  c->d.n = c->d.sk * (prm->nk + 1);
  c->pi = 4.0 * atan(1.0);
  int dev_count = 0;
  if (hipGetDeviceCount(&dev_count) != hipSuccess || dev_count < 1) {
    fprintf(stderr, "fvm_amd: no HIP device\n");
    delete c;
    return FVM_ERR_HIP;
  }
  (void)hipGetDevice(&c->dev);
  if (hipStreamCreate(&c->stream) != hipSuccess) { delete c; return FVM_ERR_HIP; }
  const long n = c->d.n;
  int r = 0;
  double** cells[] = {&c->xg, &c->yg, &c->zg, &c->x, &c->y, &c->z, &c->djac, &c->jq11, &c->jq22,
                      &c->jq33, &c->jq12, &c->jq13, &c->jq23, &c->theta, &c->filt2, &c->ap0,
                      &c->pdI, &c->pdJ, &c->pdK, &c->u, &c->v, &c->w, &c->p, &c->un, &c->vn,
                      &c->wn, &c->rhs, &c->psu, &c->mag, &c->nut};
  // the general chunked AF stages rotate u, v, w with the AF work arrays, which also hold the
  // J-line scans' padded line-major copies: all of them get the larger size
  const long na = std::max((long)n, scan_jline_elems(c->d.ni, c->d.nj, c->d.nk));
  for (double** q : cells)
    r |= dalloc(c, q, (q == &c->u || q == &c->v || q == &c->w) ? na : n);
  for (int a = 0; a < 9; ++a) r |= dalloc(c, &c->m[a], n);
  for (int s = 0; s < 3; ++s)
    for (int q = 0; q < 3; ++q) r |= dalloc(c, &c->fa[s][q], n);
  for (int q = 0; q < 3; ++q) {
    r |= dalloc(c, &c->nl[q], n);
    r |= dalloc(c, &c->nl1[q], n);
    r |= dalloc(c, &c->dif[q], n);
    r |= dalloc(c, &c->s[q], n);
    r |= dalloc(c, &c->su[q], n);
    r |= dalloc(c, &c->ju[q], n);
    r |= dalloc(c, &c->uh[q], n);
    r |= dalloc(c, &c->sgs[q], n);
    r |= dalloc(c, &c->afb[q][0], na);
    r |= dalloc(c, &c->afb[q][1], na);
  }
  for (int q = 0; q < 6; ++q) r |= dalloc(c, &c->lss[q], n);
  r |= dalloc(c, &c->ufar, 3L * (prm->nj + 1) * (prm->nk + 1));
  r |= dalloc(c, &c->wall, 12L * (prm->nj - 1) * (prm->nk - 1));
  long nblk = (long)((prm->ni - 1 + 63) / 64) * ((prm->nj - 1 + 3) / 4) * (prm->nk - 1);
  // AF chunk iteration: one partial pair per workgroup, chunks of 8 lines
  long nafw = 3L * ((std::max(prm->ni, prm->nj) + 7) / 8 + 1) * std::max(prm->nj, prm->nk) + 8;
  c->nred = (int)std::max(std::max(nblk * 2, 2 * nafw), 4096L);
  r |= dalloc(c, &c->red, c->nred);
  c->ncflpart = (int)nblk;
  r |= dalloc(c, &c->cflpart, nblk);
  r |= dalloc(c, &c->scal, 128);
  c->sink = c->scal + 64;
  double *p32 = nullptr, *su32 = nullptr;
  r |= dalloc(c, &p32, (n + 1) / 2);
  r |= dalloc(c, &su32, (n + 1) / 2);
  c->p32 = (float*)p32;
  c->su32 = (float*)su32;
  if (r) { fvm_destroy(c); return FVM_ERR_HIP; }
  if (hipHostMalloc((void**)&c->hscal, 64 * sizeof(double)) != hipSuccess) {
    fvm_destroy(c);
    return FVM_ERR_HIP;
  }
  if (hipHostMalloc((void**)&c->hmail, 64 * sizeof(double), hipHostMallocCoherent) !=
      hipSuccess) {
    fvm_destroy(c);
    return FVM_ERR_HIP;
  }
  c->hseq = reinterpret_cast<unsigned long long*>(c->hmail + 56);
  *c->hseq = 0;
  // device pointer tables
  double** tabs[8];
  for (int t = 0; t < 8; ++t) {
    if (hipMalloc((void**)&tabs[t], 9 * sizeof(double*)) != hipSuccess) { fvm_destroy(c); return FVM_ERR_HIP; }
    c->allocs.push_back(tabs[t]);
  }
  c->dm = tabs[0];
  c->dlss = tabs[1];
  for (int s = 0; s < 3; ++s) c->dfa[s] = tabs[2 + s];
  c->afptr = tabs[5];
  c->scantab[0] = tabs[6];
  c->scantab[1] = tabs[7];
  (void)hipMemcpy(c->dm, c->m, 9 * sizeof(double*), hipMemcpyHostToDevice);
  (void)hipMemcpy(c->dlss, c->lss, 6 * sizeof(double*), hipMemcpyHostToDevice);
  for (int s = 0; s < 3; ++s) (void)hipMemcpy(c->dfa[s], c->fa[s], 3 * sizeof(double*), hipMemcpyHostToDevice);
  // node coordinates: NI*NJ*NK host records -> cell layout, nodes 1..N
  const int ni = prm->ni, nj = prm->nj, nk = prm->nk;
  std::vector<double> h(n, 0.0);
  const double* src[3] = {xg, yg, zg};
  double* dst[3] = {c->xg, c->yg, c->zg};
  for (int a = 0; a < 3; ++a) {
    for (int k = 0; k < nk; ++k)
      for (int j = 0; j < nj; ++j)
        for (int i = 0; i < ni; ++i)
          h[cidx(c->d, i + 1, j + 1, k + 1)] = src[a][i + (long)ni * (j + (long)nj * k)];
    if (hipMemcpy(dst[a], h.data(), n * sizeof(double), hipMemcpyHostToDevice) != hipSuccess) {
      fvm_destroy(c);
      return FVM_ERR_HIP;
    }
  }
  r = fvm_geom_setup(c);
  if (r) { fvm_destroy(c); return r; }
  r = fvm_af_setup(c);
  if (r) { fvm_destroy(c); return r; }
  r = fvm_sep_setup(c);
  if (r) { fvm_destroy(c); return r; }
  *out = c;
  return FVM_OK;
}

int fvm_destroy(fvm_ctx* c) {
  if (!c) return FVM_ERR_ARG;
  if (c->stream) (void)hipStreamSynchronize(c->stream);
  fvm_sep_destroy(c);
  for (auto& e : c->ev_pending) {
    (void)hipEventDestroy(e.second.first);
    (void)hipEventDestroy(e.second.second);
  }
  for (void* p : c->allocs) (void)hipFree(p);
  if (c->hscal) (void)hipHostFree(c->hscal);
  if (c->hmail) (void)hipHostFree(c->hmail);
  if (c->stream) (void)hipStreamDestroy(c->stream);
  delete c;
  return FVM_OK;
}

// Fields that carry the cell volume J^-1 (fluxes summed over the faces, NL, DIF, S_U, the
// volume itself): on the 2D slab they are LZ2D times the 2D solver's per-unit-span values
// (cylinder/.../init.f90:46-56 reads and writedata.f90 writes NLX, NLY, ADV per unit span), so
// a 2D context exchanges them with the host per unit span.
static bool span_scaled(int f) {
  switch (f) {
    case FVM_NLX: case FVM_NLY: case FVM_NLZ:
    case FVM_NLX_NM1: case FVM_NLY_NM1: case FVM_NLZ_NM1:
    case FVM_DIFX: case FVM_DIFY: case FVM_DIFZ:
    case FVM_SU: case FVM_SV: case FVM_SW:
    case FVM_DJAC: case FVM_ADV: case FVM_ADV_NM1: case FVM_DIFT:
      return true;
    default:
      return false;
  }
}

int fvm_set_field(fvm_ctx* c, int field, const double* host) {
  if (!c || !host) return FVM_ERR_ARG;
  c->u_is_un = false;
  double* dptr = field_ptr(c, field);
  if (!dptr) return FVM_ERR_ARG;
  int rs = fvm_k_sgs_out(c);  // form SGS_x,y,z and nu_t before any of their inputs can change
  if (rs) return rs;
  rs = fvm_k_nut_out(c);
  if (rs) return rs;
  if (field >= FVM_XIX && field <= FVM_ZTZ) c->xz_zero = false;  // re-derived only at set-up
  if (field == FVM_UN || field == FVM_VN || field == FVM_WN) c->ufar_stale = true;
  std::vector<double> slab;
  if (c->two_d) {  // one (NI+1)(NJ+1) plane, repeated on the slab's three K planes
    slab.resize(c->d.n);
    const double sc = span_scaled(field) ? LZ2D : 1.0;
    for (int k = 0; k < 3; ++k)
      for (long q = 0; q < c->d.sk; ++q) slab[k * c->d.sk + q] = host[q] * sc;
    host = slab.data();
  }
  HIP_CHECK(hipMemcpyAsync(dptr, host, c->d.n * sizeof(double), hipMemcpyHostToDevice, c->stream));
  if (field >= FVM_NLX && field <= FVM_NLZ)
    HIP_CHECK(hipMemcpyAsync(c->nl1[field - FVM_NLX], host, c->d.n * sizeof(double),
                             hipMemcpyHostToDevice, c->stream));
  if (field == FVM_ADV)
    HIP_CHECK(hipMemcpyAsync(c->adv1, host, c->d.n * sizeof(double), hipMemcpyHostToDevice,
                             c->stream));
  HIP_CHECK(hipStreamSynchronize(c->stream));
  return FVM_OK;
}

int fvm_get_field(fvm_ctx* c, int field, double* host) {
  if (!c || !host) return FVM_ERR_ARG;
  double* dptr = field_ptr(c, field);
  if (!dptr) return FVM_ERR_ARG;
  if (field >= FVM_SGSX && field <= FVM_SGSZ) {
    const int rs = fvm_k_sgs_out(c);
    if (rs) return rs;
  }
  if (field == FVM_NUT) {
    const int rs = fvm_k_nut_out(c);
    if (rs) return rs;
  }
  if (c->two_d) dptr += c->d.sk;  // the slab's interior plane K = 1
  const long cnt = c->two_d ? c->d.sk : c->d.n;
  HIP_CHECK(hipMemcpyAsync(host, dptr, cnt * sizeof(double), hipMemcpyDeviceToHost, c->stream));
  HIP_CHECK(hipStreamSynchronize(c->stream));
  if (c->two_d && span_scaled(field))
    for (long q = 0; q < cnt; ++q) host[q] /= LZ2D;
  return FVM_OK;
}

int fvm_begin_step(fvm_ctx* c) {
  if (!c) return FVM_ERR_ARG;
  int r = fvm_k_step_copy(c);  // U_N = U ... (T_N = T)
  if (r) return r;
  if (c->thermal) {  // BCT's far-field values depend only on T_N
    r = fvm_k_tfar(c);
    if (r) return r;
  }
  // far-field boundary values depend only on U_N: one evaluation serves every BCU of the step
  c->ufar_stale = false;
  return fvm_k_ufar(c);
}

int fvm_set_thermal(fvm_ctx* c, double pe) {
  if (!c || c->two_d || !(pe > 0)) return FVM_ERR_ARG;
  c->pe = pe;
  return fvm_thermal_setup(c);
}

int fvm_iter_t(fvm_ctx* c) { return c ? fvm_k_iter_t(c) : FVM_ERR_ARG; }

int fvm_writedata_fields(fvm_ctx* c) { return c ? fvm_k_writedata_fields(c) : FVM_ERR_ARG; }

int fvm_set_dtime(fvm_ctx* c, double dtime) {
  if (!c || !(dtime > 0)) return FVM_ERR_ARG;
  return fvm_k_set_dtime(c, dtime);
}

int fvm_ensemble_step(fvm_ctx* const* ctxs, int n, int nthreads, fvm_step_info* infos) {
  if (!ctxs || n < 0 || nthreads < 1) return FVM_ERR_ARG;
  for (int q = 0; q < n; ++q)
    if (!ctxs[q]) return FVM_ERR_ARG;
  {  // one context twice would be stepped by two threads at once on one stream and mailbox
    std::vector<fvm_ctx*> sorted(ctxs, ctxs + n);
    std::sort(sorted.begin(), sorted.end());
    if (std::adjacent_find(sorted.begin(), sorted.end()) != sorted.end()) return FVM_ERR_ARG;
  }
  std::vector<int> status(n, FVM_OK);
  const int nt = std::min(nthreads, std::max(n, 1));
  // each thread takes the next case from a shared counter, so uneven cases balance and any
  // number of started threads (the calling one included) covers every case
  std::atomic<int> next{0};
  auto work = [&]() {
    for (int q; (q = next.fetch_add(1)) < n;)
      status[q] = fvm_step(ctxs[q], infos ? infos + q : nullptr);
  };
  std::vector<std::thread> pool;
  try {
    for (int t = 1; t < nt; ++t) pool.emplace_back(work);
  } catch (...) {  // fewer threads: the cases are still all taken
  }
  work();
  for (std::thread& th : pool) th.join();
  int r = FVM_OK;
  for (int q = 0; q < n; ++q) {
    if (status[q] != FVM_OK && status[q] != FVM_ERR_NOCONV) return status[q];
    if (status[q] == FVM_ERR_NOCONV) r = FVM_ERR_NOCONV;
  }
  return r;
}

int fvm_post_begin(fvm_ctx* c, const double* t_const) {
  return c ? fvm_k_post_begin(c, t_const) : FVM_ERR_ARG;
}

int fvm_post_finalize(fvm_ctx* c, int* niter_pos) {
  if (!c) return FVM_ERR_ARG;
  const long np = c->niter_pos;
  int r = fvm_k_post_finalize(c);
  if (r) return r;
  HIP_CHECK(hipStreamSynchronize(c->stream));
  if (niter_pos) *niter_pos = (int)np;
  return FVM_OK;
}

// ---- the reference's subroutines with their own arguments (one array per call) ----

int fvm_copy_field(fvm_ctx* c, int dst, int src) {
  if (!c) return FVM_ERR_ARG;
  double* a = field_ptr(c, dst);
  const double* b = field_ptr(c, src);
  if (!a || !b) return FVM_ERR_ARG;
  if (a == b) return FVM_OK;
  int rs = fvm_k_sgs_out(c);  // the lazily formed SGS_x,y,z and nu_t, before inputs change
  if (rs) return rs;
  rs = fvm_k_nut_out(c);
  if (rs) return rs;
  HIP_CHECK(hipMemcpyAsync(a, b, c->d.n * sizeof(double), hipMemcpyDeviceToDevice, c->stream));
  if (dst == FVM_UN || dst == FVM_VN || dst == FVM_WN) c->ufar_stale = true;
  if (dst == FVM_U || dst == FVM_V || dst == FVM_W || dst == FVM_UN || dst == FVM_VN ||
      dst == FVM_WN)
    c->u_is_un = false;
  return FVM_OK;
}

int fvm_nl_c(fvm_ctx* c, int in, int out) {
  if (!c) return FVM_ERR_ARG;
  const double* a = field_ptr(c, in);
  double* b = field_ptr(c, out);
  if (!a || !b || a == b) return FVM_ERR_ARG;
  return fvm_k_nl_one(c, a, b);
}

int fvm_dif_c(fvm_ctx* c, int in, int out) {
  if (!c) return FVM_ERR_ARG;
  const double* a = field_ptr(c, in);
  double* b = field_ptr(c, out);
  if (!a || !b || a == b) return FVM_ERR_ARG;
  return fvm_k_dif_one(c, a, b);
}

int fvm_stress_tensor_dy_c(fvm_ctx* c, double* c_avg) {
  if (!c || c->two_d) return FVM_ERR_ARG;
  int r = ufar_fresh(c);  // BCU of the test-filtered velocities takes the far-field values
  if (r) return r;
  r = fvm_k_sgs(c, true, false);
  if (r) return r;
  if (c_avg) {
    HIP_CHECK(hipMemcpyAsync(c_avg, c->scal, sizeof(double), hipMemcpyDeviceToHost, c->stream));
    HIP_CHECK(hipStreamSynchronize(c->stream));
  }
  return FVM_OK;
}

int fvm_add_sgs_c(fvm_ctx* c, int nl, int sgs) {
  if (!c) return FVM_ERR_ARG;
  double* a = field_ptr(c, nl);
  const double* b = field_ptr(c, sgs);
  if (!a || !b || a == b) return FVM_ERR_ARG;
  return fvm_k_add_sgs(c, a, b);
}

int fvm_iter_u_c(fvm_ctx* c, int in, int inm, int ina, int inb, int inc, int ind) {
  if (!c) return FVM_ERR_ARG;
  const int comp = vel_comp(in);
  double* f = field_ptr(c, in);
  const double* fn = field_ptr(c, inm);
  const double* nl = field_ptr(c, ina);
  const double* nl1 = field_ptr(c, inb);
  const double* dif = field_ptr(c, inc);
  const double* s = field_ptr(c, ind);
  if (comp < 0 || !f || !fn || !nl || !nl1 || !dif || !s || fn == f) return FVM_ERR_ARG;
  int r = ufar_fresh(c);
  if (r) return r;
  return fvm_k_iter_u_one(c, comp, f, fn, nl, nl1, dif, s);
}

int fvm_correct_c(fvm_ctx* c, int f, int s, double sign) {
  if (!c) return FVM_ERR_ARG;
  double* a = field_ptr(c, f);
  const double* b = field_ptr(c, s);
  if (!a || !b || a == b) return FVM_ERR_ARG;
  c->u_is_un = false;
  return fvm_k_correct_one(c, a, b, sign);
}

int fvm_bcu(fvm_ctx* c, int a, int b) {
  if (!c) return FVM_ERR_ARG;
  const int comp = vel_comp(a);
  if (comp < 0 || b != FVM_UN + comp) return FVM_ERR_ARG;
  int r = ufar_fresh(c);
  if (r) return r;
  return fvm_k_bcu_ghosts_one(c, field_ptr(c, a), comp);
}

int fvm_bcp(fvm_ctx* c) { return c ? fvm_k_bcp(c) : FVM_ERR_ARG; }

int fvm_contra_vel(fvm_ctx* c) { return c ? fvm_k_contra_vel(c) : FVM_ERR_ARG; }
int fvm_nl(fvm_ctx* c) { return c ? fvm_k_nl(c) : FVM_ERR_ARG; }
int fvm_dif(fvm_ctx* c) { return c ? fvm_k_dif(c) : FVM_ERR_ARG; }
int fvm_p_source(fvm_ctx* c) { return c ? fvm_k_p_source(c) : FVM_ERR_ARG; }
int fvm_iter_u(fvm_ctx* c) { return c ? fvm_k_iter_u(c) : FVM_ERR_ARG; }

int fvm_stress_tensor_dy(fvm_ctx* c, double* c_avg) {
  if (!c) return FVM_ERR_ARG;
  int r = fvm_k_sgs(c);
  if (r) return r;
  if (c_avg) {
    HIP_CHECK(hipMemcpyAsync(c_avg, c->scal, sizeof(double), hipMemcpyDeviceToHost, c->stream));
    HIP_CHECK(hipStreamSynchronize(c->stream));
  }
  return FVM_OK;
}

int fvm_correct(fvm_ctx* c, double sign) {
  if (!c) return FVM_ERR_ARG;
  c->u_is_un = false;
  int r = fvm_k_correct(c, sign);
  if (r) return r;
  double* f[3] = {c->u, c->v, c->w};
  return fvm_k_bcu_ghosts(c, f, 3);
}

int fvm_iter_p(fvm_ctx* c, double* resip, int* niter) {
  if (!c) return FVM_ERR_ARG;
  double res = 0.0;
  int it = 0;
  int r = fvm_k_iter_p(c, &res, &it);
  if (resip) *resip = res;
  if (niter) *niter = it;
  return r;
}

int fvm_pressure_reference(fvm_ctx* c, double* pref) {
  if (!c) return FVM_ERR_ARG;
  if (c->two_d) {  // the 2D MAIN keeps p unshifted
    if (pref) *pref = 0.0;
    return FVM_OK;
  }
  return fvm_k_pref(c, pref);
}

int fvm_calccoef(fvm_ctx* c, fvm_step_info* info) {
  if (!c) return FVM_ERR_ARG;
  double o[8];
  int r = c->two_d ? fvm_k_calccoef2d(c, o) : fvm_k_calccoef(c, o);
  if (r) return r;
  return fvm_calccoef_info(c, o, info);
}

int fvm_step(fvm_ctx* c, fvm_step_info* info) {
  if (!c) return FVM_ERR_ARG;
  fvm_step_info st{};
  int r;
#define CK(x) do { r = (x); if (r && r != FVM_ERR_NOCONV) return r; } while (0)
  CK(fvm_begin_step(c));                          // main.f90:42-45
  CK(fvm_k_contra_vel(c));                        // :48
  CK(fvm_k_nl_dif(c));                            // :51-57 and DIF :72-75 in one pass
  if (c->prm.use_les) {                           // :59-69
    CK(fvm_k_sgs(c, false));  // SGS_x,y,z formed on first read (fvm_get_field)
  }
  CK(fvm_k_p_source(c));                          // :78
  CK(fvm_k_iter_u(c));                            // :82-88
  CK(fvm_k_correct_bcu_contra(c));                // :90-101 (correction, BCU, CONTRA_VEL)
  double res = 0.0;
  int it = 0;
  int rp = fvm_k_iter_p(c, &res, &it);            // :103-105 (BCP included)
  if (rp && rp != FVM_ERR_NOCONV) return rp;
  CK(fvm_k_psource_correct(c, +1.0, true));       // :108-111 (P_SOURCE, projection; CFL maxima)
  if (!c->two_d) CK(fvm_k_pref(c, nullptr));     // :125-135 (none in 2D)
  double* f[3] = {c->u, c->v, c->w};
  CK(fvm_k_bcu_ghosts(c, f, 3));                  // :138-142
  if (c->thermal) CK(fvm_k_iter_t(c));            // :145-147 (ITER_T, BCT)
  if (c->two_d) {
    const Mailbox mb = fvm_mail(c);               // the 2D CALCCOEF publishes scal[0..16]
    CK(fvm_k_calccoef2d(c, nullptr, true, &mb));  // 2D CALCCOEF
    r = fvm_mail_wait(c, mb, 17, c->hscal);
  } else {
    const Mailbox mb = fvm_mail(c);               // the force kernel publishes scal[0..16]
    CK(fvm_k_calccoef(c, nullptr, true, &mb));    // :150
    CK(fvm_k_post_accum(c));                      // calccoef.f90:286-307 (time averages)
    r = fvm_mail_wait(c, mb, 17, c->hscal);
  }
#undef CK
  if (r) return r;
  st.c_avg = c->prm.use_les ? c->hscal[0] : 0.0;
  st.c_t = (c->thermal && c->prm.use_les) ? c->hscal[1] : 0.0;
  st.nu = c->thermal ? c->hscal[16] : 0.0;
  st.pref = c->hscal[2];
  st.cfl = c->hscal[8]; st.cpb = c->hscal[9]; st.cd = c->hscal[10]; st.cl = c->hscal[11];
  st.cd_p = c->hscal[12]; st.cd_v = c->hscal[13]; st.cl_p = c->hscal[14]; st.cl_v = c->hscal[15];
  st.resip = res;
  st.niter = it;
  st.nouter = c->last.nouter;
  for (int q = 0; q < 3; ++q) st.af_iters[q] = c->af_iters[q];
  c->last = st;
  if (info) *info = st;
  prof_flush(c);
  if (!(st.cd == st.cd) || !(st.cfl == st.cfl)) {
    fprintf(stderr, "fvm_amd: non-finite force coefficients / CFL after the step\n");
    return FVM_ERR_NAN;
  }
  return rp;
}

int fvm_profile_enable(fvm_ctx* c, int on) {
  if (!c) return FVM_ERR_ARG;
  prof_flush(c);
  c->prof_on = on != 0;
  return FVM_OK;
}

int fvm_profile_reset(fvm_ctx* c) {
  if (!c) return FVM_ERR_ARG;
  prof_flush(c);
  c->prof.clear();
  return FVM_OK;
}

int fvm_profile_get(fvm_ctx* c, char* names, double* ms, long* calls, int cap) {
  if (!c) return -1;
  prof_flush(c);
  int q = 0;
  for (auto& kv : c->prof) {
    if (q >= cap) break;
    if (names) {
      strncpy(names + 48 * q, kv.first.c_str(), 47);
      names[48 * q + 47] = 0;
    }
    if (ms) ms[q] = kv.second.ms;
    if (calls) calls[q] = kv.second.calls;
    ++q;
  }
  return q;
}

int fvm_divergence(fvm_ctx* c, double* out) {
  if (!c || !out) return FVM_ERR_ARG;
  return fvm_k_div_norm(c, out);
}

int fvm_sync(fvm_ctx* c) {
  if (!c) return FVM_ERR_ARG;
  prof_flush(c);
  HIP_CHECK(hipStreamSynchronize(c->stream));
  return FVM_OK;
}

}  // extern "C"
