// Kernels of the batched 2D step (cylinder/cylinder_v3.0.3_AF_original 4step_face_nl/,
// Main.f90:48-118).  Each launch covers every case of the batch: its workgroups are mapped to
// (case, tile) through block tables built on the host.  Expressions follow the reference's
// association order (the library is built with -ffp-contract=off).
#pragma once
#include "fvm2d.h"

#define C2_DONE(cs) ((cs).stat[ST_DONE] != 0.0)

// A 64 x TJ (I x J) tile of interior cells; blockIdx.x indexes the block table (case -1: a
// padding entry, no tile).
struct TileIdx {
  int c, i, j, tile;
};

__device__ inline TileIdx tile_index(const int* __restrict__ bcase, const int* __restrict__ btile,
                                     const Case2D* __restrict__ tab) {
  TileIdx t;
  t.c = bcase[blockIdx.x];
  t.tile = btile[blockIdx.x];
  if (t.c < 0) {  // padding entry of an XCD-ordered table: the caller returns
    t.i = t.j = 0;
    return t;
  }
  const int tx = (tab[t.c].ni - 1 + 63) >> 6;
  t.i = 1 + (t.tile % tx) * 64 + (int)threadIdx.x;
  t.j = 1 + (t.tile / tx) * (int)blockDim.y + (int)threadIdx.y;
  return t;
}

// workgroup sum / max over blockDim.x * blockDim.y threads (a fixed-order tree)
__device__ inline double wg_sum(double v, double* sh) {
  const int t = threadIdx.x + blockDim.x * threadIdx.y, nt = blockDim.x * blockDim.y;
  sh[t] = v;
  __syncthreads();
  for (int s = nt / 2; s > 0; s >>= 1) {
    if (t < s) sh[t] = sh[t] + sh[t + s];
    __syncthreads();
  }
  const double r = sh[0];
  __syncthreads();
  return r;
}

__device__ inline double wg_max(double v, double* sh) {
  const int t = threadIdx.x + blockDim.x * threadIdx.y, nt = blockDim.x * blockDim.y;
  sh[t] = v;
  __syncthreads();
  for (int s = nt / 2; s > 0; s >>= 1) {
    if (t < s) sh[t] = fmax(sh[t], sh[t + s]);
    __syncthreads();
  }
  const double r = sh[0];
  __syncthreads();
  return r;
}

// DJAC x contravariant velocity (contra_vel.f90:9-10: U1 = XIX U + XIY V, U2 = ETX U + ETY V;
// the face fluxes carry DJAC U1)
// (the helpers below take generic or global-address-space pointers alike)
template <class P>
struct GeomT {
  P dj, xix, xiy, etx, ety;
};
using Geom = GeomT<const double*>;
using GeomG = GeomT<gptr<const double>>;
__device__ inline GeomG geom2(const Case2D& cs) {
  return GeomG{glb((const double*)cs.f[G_DJAC]), glb((const double*)cs.f[G_XIX]),
               glb((const double*)cs.f[G_XIY]), glb((const double*)cs.f[G_ETX]),
               glb((const double*)cs.f[G_ETY])};
}
template <class G, class P>
__device__ inline double ju1(const GeomT<G>& g, P U, P V, long q) {
  return g.dj[q] * (g.xix[q] * U[q] + g.xiy[q] * V[q]);
}
template <class G, class P>
__device__ inline double ju2(const GeomT<G>& g, P U, P V, long q) {
  return g.dj[q] * (g.etx[q] * U[q] + g.ety[q] * V[q]);
}

// NL(IN, OUT) (nl.f90:21-27) with the face fluxes UE, UW, VN, VS
template <class P>
__device__ inline double nl2(P F, long c, int sj, double ue, double uw, double vn, double vs) {
  const double p = F[c];
  double t = ue * 0.5 * (p + F[c + 1]);
  t = t - uw * 0.5 * (p + F[c - 1]);
  t = t + vn * 0.5 * (p + F[c + sj]);
  t = t - vs * 0.5 * (p + F[c - sj]);
  return 1.0 * t;
}

// DIF_DIAGONAL-type cross term 0.25 (w_E d_J(F)_E - w_W d_J(F)_W + w_N d_I(F)_N - w_S d_I(F)_S)
// (dif_diagonal.f90:13-23 with w = DJAC Q12; iter_u.f90:30-41 with w = DJAC XI_s ET_s)
template <class W, class P>
__device__ inline double cross2(W __restrict__ w, P F, long c, int sj) {
  double t = w[c + 1] * (F[c + 1 + sj] - F[c + 1 - sj]);
  t = t - w[c - 1] * (F[c - 1 + sj] - F[c - 1 - sj]);
  t = t + w[c + sj] * (F[c + 1 + sj] - F[c - 1 + sj]);
  t = t - w[c - sj] * (F[c + 1 - sj] - F[c - 1 - sj]);
  return 0.25 * t;
}

// DIF(IN, OUT) (dif.f90:13-23): orthogonal part with DW(i) = DE(i-1), DS(j) = DN(j-1)
template <class W, class P>
__device__ inline double dif2(W __restrict__ de, W __restrict__ dn, W __restrict__ jq12, P F,
                              long c, int sj) {
  const double p = F[c];
  double t = de[c] * (F[c + 1] - p) + de[c - 1] * (F[c - 1] - p);
  t = t + dn[c] * (F[c + sj] - p);
  t = t + dn[c - sj] * (F[c - sj] - p);
  t = t + cross2(jq12, F, c, sj);
  return 1.0 * t;
}

// P_SOURCE (p_source.f90:9-27): S = -0.5 ((DJ XI)_E p_E - (DJ XI)_W p_W + (DJ ET)_N p_N -
// (DJ ET)_S p_S)
template <class W, class Q>
__device__ inline double psrc(W __restrict__ dj, W __restrict__ mxi, W __restrict__ met, Q P,
                              long c, int sj) {
  double t = dj[c + 1] * mxi[c + 1] * P[c + 1];
  t = t - dj[c - 1] * mxi[c - 1] * P[c - 1];
  t = t + dj[c + sj] * met[c + sj] * P[c + sj];
  t = t - dj[c - sj] * met[c - sj] * P[c - sj];
  return -0.5 * t;
}
