// Host-side set-up data of one 2D case (fvm2d_setup.hip).
#pragma once
#include <vector>

struct Geo2D {
  int ni = 0, nj = 0;
  std::vector<double> x, y, djac, xix, xiy, etx, ety, theta, jq12, de, dn;
  // per wall cell J (index 6 J .. 6 J + 5, J = 1..NJ-1): DN1, sin TH2, cos TH2, AREA,
  // cos THETA(1,J), sin THETA(1,J)
  std::vector<double> wall;
};

struct AF2D {
  std::vector<double> ap0, an1, ae2, aw2;
  std::vector<double> invd[2], na[2], b[2], jxe[2];
  // transposed line systems (the discrete adjoint): T^T has the pivots of T, so with
  // T = tridiag(a, d, c): y(r) = q(r)/D(r) + naT(r) y(r-1), naT(r) = -c(r-1)/D(r), and
  // x(r) = y(r) - bT(r) x(r+1), bT(r) = a(r+1)/D(r); stage 1 (J lines) also needs its line-end
  // couplings to the ghosts: ga = -a (AS at the first row), gc = -c (AN at the last row)
  std::vector<double> nat[2], bt[2], ga2, gc2;
  std::vector<double> awwall[2], aefar[2];  // pre-fold AW(1, J), AE(NI-1, J) per stage
  std::vector<double> ae1;                  // stage-1 AE after the far-field fold
};

int geo2d_build(int ni, int nj, const double* xg, const double* yg, Geo2D& g);
void geo2d_af(const Geo2D& g, double re, double dt, AF2D& a);
