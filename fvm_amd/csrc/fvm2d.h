// Batched native 2D solver: many independent 2D cylinder cases (the time step of
// cylinder/cylinder_v3.0.3_AF_original 4step_face_nl/, Main.f90:48-118) advanced together.
//
// Every case owns planes of (NI+1)(NJ+1) FP64 cells, I fastest, indices 0..NI x 0..NJ incl.
// ghosts, exactly the reference's (0:NI,0:NJ) arrays at its single plane K = 2.  A device table
// holds one Case2D per case (sizes, constants, field pointers); every phase of the step is ONE
// launch over all cases, its workgroups mapped to (case, tile) through a per-launch block table,
// so cases of uneven size and Reynolds number share each launch.  The step has a fixed shape:
// the Poisson defect correction runs a fixed number of passes, and each case's kernels skip the
// work once the case has converged (a per-case flag on the device), so the whole step is
// captured once as a hipGraph and replayed.
#pragma once
#include "fvm_common.h"
#include "fvm_fft.h"

// per-case plane fields
enum F2 {
  // geometry and constants (set-up, per DTIME)
  G_X, G_Y, G_DJAC, G_XIX, G_XIY, G_ETX, G_ETY, G_THETA,
  G_JQ12,              // DJAC Q12 (= DJAC Q21), the cross-term weight of DIF_DIAGONAL
  G_DE, G_DN,          // DIF face coefficients (DW(i) = DE(i-1), DS(j) = DN(j-1) bit for bit)
  G_JXE1, G_JXE2,      // (DJAC XIX) ETX and (DJAC XIY) ETY, the AF cross-term weights
  G_AP0,               // 2 Re DJAC / DTIME
  G_AN1, G_AE2, G_AW2, // AF line couplings lagged in the sweeps: stage 1 AN (AS(j) = AN(j-1)),
                       // stage 2 AE, AW (with the wall / far-field folds)
  G_INVD1, G_NA1, G_B1,  // stage-1 (I-line) Thomas factors 1/D, -A/D, B
  G_INVD2, G_NA2, G_B2,  // stage-2 (J-line) Thomas factors
  G_NAT1, G_BT1, G_NAT2, G_BT2,  // transposed line factors of the two stages (adjoint)
  G_GA2, G_GC2,        // stage-2 line couplings AS, AN (the line ends' ghost terms)
  G_TINVD2, G_TNA2, G_TB2, G_TAW2,  // stage-2 factors and AW in the J-line-major layout
                                    // [J + (NJ+1) I] (J-line sweeps read rows contiguously)
  G_AE1,               // stage-1 AE (folded AE(NI-1) = 0; AW(I) = AE(I-1)): the FP64 residual
                       // of the FP32 change form
  // state
  F_U, F_V, F_P, F_NLX, F_NLY,
  // step work arrays
  F_UN, F_VN, F_DIFX, F_DIFY, F_SU, F_SV, F_SU1, F_SV1, F_RHS,
  F_TQU, F_TQV, F_TXU, F_TXV,  // J-line sweep: right-hand sides and iterates, J-line-major
  NF2
};

// per-J boundary arrays (NJ+1 each): far-field U_F, V_F of the step; pre-fold AW(1), AE(NI-1)
// of the two AF stages; the wall velocity U_F, V_F (blowing / suction; zero: the reference's
// fixed wall); the wall-cell geometry of CALCCOEF
// (B_WALL: 6 values per J, see Geo2D::wall)
enum B2 { B_UFAR, B_VFAR, B_AW1W, B_AE1F, B_AW2W, B_AE2F, B_UWALL, B_VWALL, B_WALL, NB2 };

// adjoint planes (allocated by fvm_batch2d_adjoint_init): the state adjoints of u, v (full
// planes), p (full plane), NL (interior); work planes; the Poisson adjoint's solution, source and
// operator image; per-J accumulators of the wall and far-field velocity adjoints
enum A2 { A_LSU, A_LSV, A_LP, A_LNX, A_LNY, A_W1U, A_W1V, A_W2U, A_W2V, A_LQU, A_LQV, A_LSXU,
          A_LSXV, A_MU, A_B, A_WP, NA2 };
enum AB2 { AB_UFW, AB_VFW, AB_UFF, AB_VFF, NAB2 };

// FP32 planes of the change form of the AF sweeps (a stage solves T (x - x0) = r, r the FP64
// residual of the stage input x0, in FP32): the residual (natural layout), the change iterate
// (natural layout in the I-line stage, J-line-major in the J-line stage), the J-line-major
// residual, and the line factors and lagged couplings of the two stages in FP32
enum H2 { H_RU, H_RV, H_DU, H_DV, H_TRU, H_TRV, H_INVD1, H_NA1, H_B1, H_AN1, H_TINVD2, H_TNA2,
          H_TB2, H_TAW2, NH2 };

// The FP32 J-line-major planes (H_TRU, H_TRV, H_T*2 and H_DU / H_DV in the J-line stage) hold
// cell (I, J) at (J + J32_OFF) + tls32(NJ) I: a line stride that is a multiple of four floats and
// row J = 1 on a 16-byte boundary, so a lane's R consecutive rows (R = 2, 4, 6, 8) are read as
// float2 / float4 vectors.
constexpr int J32_OFF = 3;
__host__ __device__ inline long tls32(int nj) { return ((long)nj + 1 + J32_OFF + 3) & ~3L; }
__host__ __device__ inline long tidx32(int nj, int i, int j) { return j + J32_OFF + tls32(nj) * i; }

struct Case2D {
  int ni, nj;          // node counts NI (radial), NJ (circumferential)
  int sj;              // stride of J = NI + 1
  int n;               // cells per plane incl. ghosts
  double re, dt;
  double resimax;
  double* f[NF2];
  double* b[NB2];
  // Poisson separable inverse (FP64 and FP32 forms share the factors)
  FftPlan pj;          // plan of the J transform (length NJ - 1)
  int mj;              // half-spectrum length (NJ - 1) / 2 + 1
  double* pdw;         // <D_W> per radial row (n1)
  double* invd;        // mode-line factors [r + n1 m]
  double* bf;
  void* twj;           // twiddles exp(-2 pi i t / (NJ - 1)): Cx<double> or Cx<float>
  void* twj32;
  void* R;             // residual, compact (n1 (NJ-1)), double or float
  void* S;             // half spectrum (n1 mj complex)
  void* R32;
  void* S32;
  double* part;        // per-block partial sums of this case (residual pairs, CFL maxima)
  double* stat;        // per-case scalars (see ST_*)
  int ntile;           // stencil tiles of this case (partials per pass)
  double* swb;         // AF sweep chunk boundary lines [pass slot][component][chunk][row]
  float* h[NH2];       // FP32 change form planes
  double* a[NA2];      // adjoint planes (null until the adjoint is initialised)
  double* ab[NAB2];
};

// lines per chunk of the batched AF sweeps, and chunks of a sweep over nl lines
constexpr int SW_C = 16;
__host__ __device__ inline int sw_nchunk(int nl) { return (nl + SW_C - 1) / SW_C; }

// per-case device scalars
enum ST {
  ST_RESIP = 0, ST_S0, ST_S1, ST_DONE, ST_NPASS, ST_CFL, ST_CPB, ST_CD, ST_CL, ST_CDP, ST_CDV,
  ST_CLP, ST_CLV, ST_RES0, ST_NAN, ST_ARES,
  ST_AFCH,  // largest relative change of a chunk's boundary line in the last AF sweep pass
  NST = 17
};
