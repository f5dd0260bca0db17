// CPU oracle, part 1: 1D polynomial basis, Vandermonde matrices, DG operators and index mappings.
// TEST INFRASTRUCTURE ONLY (see fxg_oracle.h). References are relative to the FLEXI root.
#include "fxg_oracle.h"
#include "oracle_internal.h"

#include <cmath>
#include <cstdlib>
#include <cstring>
#include <vector>

namespace orc {

const double PP_Pi = std::acos(-1.0);            // globals/preprocessing.f90:28
const double PP_RealTolerance = 2.220446049250313e-16;  // EPSILON(1.0D0), preprocessing.f90:27

// LegendrePolynomialAndDerivative, basis.f90:142-183 (normalised by sqrt(N+1/2)).
void legendre_and_derivative(int N, double x, double* L, double* Lder) {
  if (N == 0) {
    *L = 1.0; *Lder = 0.0;
  } else if (N == 1) {
    *L = x; *Lder = 1.0;
  } else {
    double Lm2 = 1.0, Lm1 = x, Dm1 = 1.0, l = 0.0, ld = 0.0;
    for (int i = 2; i <= N; ++i) {
      l = ((double)(2 * i - 1) * x * Lm1 - (double)(i - 1) * Lm2) / (double)i;
      ld = i * Lm1 + x * Dm1;
      Lm2 = Lm1; Lm1 = l; Dm1 = ld;
    }
    *L = l; *Lder = ld;
  }
  *L = *L * std::sqrt((double)N + 0.5);
  *Lder = *Lder * std::sqrt((double)N + 0.5);
}

// qAndLEvaluation, basis.f90:487-515.
static void q_and_l(int N, double x, double* q, double* qder, double* L) {
  double Lm2 = 1.0, Lm1 = x, l = 0.0;
  for (int i = 2; i <= N; ++i) {
    l = ((double)(2 * i - 1) * x * Lm1 - (double)(i - 1) * Lm2) / (double)i;
    Lm2 = Lm1; Lm1 = l;
  }
  *q = (double)(2 * N + 1) / (double)(N + 1) * (x * l - Lm2);
  *qder = (double)(2 * N + 1) * l;
  *L = l;
}

// LegendreGaussNodesAndWeights, basis.f90:370-439.
static void gauss_nodes(int N, double* x, double* w) {
  if (N == 0) { x[0] = 0.0; if (w) w[0] = 2.0; return; }
  if (N == 1) {
    x[0] = -std::sqrt(1.0 / 3.0); x[1] = -x[0];
    if (w) { w[0] = 1.0; w[1] = 1.0; }
    return;
  }
  const double cheb_tmp = 2.0 * std::atan(1.0) / (double)(N + 1);
  double L, Ld;
  for (int i = 0; i <= (N + 1) / 2 - 1; ++i) {
    x[i] = -std::cos(cheb_tmp * (double)(2 * i + 1));
    for (int it = 0; it <= 10; ++it) {
      legendre_and_derivative(N + 1, x[i], &L, &Ld);
      const double dx = -L / Ld;
      x[i] += dx;
      if (std::fabs(dx) < 1.0e-15 * std::fabs(x[i])) break;
    }
    legendre_and_derivative(N + 1, x[i], &L, &Ld);
    x[N - i] = -x[i];
    if (w) {
      w[i] = (2.0 * N + 3) / ((1.0 - x[i] * x[i]) * Ld * Ld);
      w[N - i] = w[i];
    }
  }
  if (N % 2 == 0) {
    x[N / 2] = 0.0;
    legendre_and_derivative(N + 1, x[N / 2], &L, &Ld);
    if (w) w[N / 2] = (2.0 * N + 3) / (Ld * Ld);
  }
}

// LegGaussLobNodesAndWeights, basis.f90:521-586.
static void gauss_lobatto_nodes(int N, double* x, double* w) {
  x[0] = -1.0; x[N] = 1.0;
  if (w) { w[0] = 2.0 / (double)(N * (N + 1)); w[N] = w[0]; }
  double q, qd, L;
  if (N > 1) {
    const double cont1 = PP_Pi / (double)N;
    const double cont2 = 3.0 / ((double)(8 * N) * PP_Pi);
    for (int i = 1; i <= (N + 1) / 2 - 1; ++i) {
      x[i] = -std::cos(cont1 * ((double)i + 0.25) - cont2 / ((double)i + 0.25));
      for (int it = 0; it <= 10; ++it) {
        q_and_l(N, x[i], &q, &qd, &L);
        const double dx = -q / qd;
        x[i] += dx;
        if (std::fabs(dx) < 1.0e-15 * std::fabs(x[i])) break;
      }
      q_and_l(N, x[i], &q, &qd, &L);
      x[N - i] = -x[i];
      if (w) { w[i] = w[0] / (L * L); w[N - i] = w[i]; }
    }
  }
  if (N % 2 == 0) {
    x[N / 2] = 0.0;
    q_and_l(N, x[N / 2], &q, &qd, &L);
    if (w) w[N / 2] = w[0] / (L * L);
  }
}

// ChebyGaussLobNodesAndWeights, basis.f90:217-241.
static void cheb_lobatto_nodes(int N, double* x, double* w) {
  for (int i = 0; i <= N; ++i) x[i] = -std::cos(i / (double)N * PP_Pi);
  if (w) {
    for (int i = 0; i <= N; ++i) w[i] = PP_Pi / (double)N;
    w[0] *= 0.5; w[N] *= 0.5;
  }
}

// BarycentricWeights, basis.f90:592-613.
void barycentric_weights(int N, const double* x, double* wb) {
  for (int i = 0; i <= N; ++i) wb[i] = 1.0;
  for (int i = 1; i <= N; ++i)
    for (int j = 0; j <= i - 1; ++j) {
      wb[j] *= (x[j] - x[i]);
      wb[i] *= (x[i] - x[j]);
    }
  for (int i = 0; i <= N; ++i) wb[i] = 1.0 / wb[i];
}

// ALMOSTEQUAL, basis.f90:651-666.
static bool almost_equal(double x, double y) {
  if (x == 0.0 || y == 0.0) return std::fabs(x - y) <= 2.0 * PP_RealTolerance;
  return std::fabs(x - y) <= PP_RealTolerance * std::fabs(x) &&
         std::fabs(x - y) <= PP_RealTolerance * std::fabs(y);
}

// LagrangeInterpolationPolys, basis.f90:679-716.
void lagrange_polys(double x, int N, const double* xGP, const double* wb, double* L) {
  bool eq = false;
  for (int i = 0; i <= N; ++i) {
    L[i] = 0.0;
    if (almost_equal(x, xGP[i])) { L[i] = 1.0; eq = true; }
  }
  if (eq) return;
  double s = 0.0;
  for (int i = 0; i <= N; ++i) { L[i] = wb[i] / (x - xGP[i]); s += L[i]; }
  for (int i = 0; i <= N; ++i) L[i] /= s;
}

// PolynomialDerivativeMatrix, basis.f90:619-643. D(iGP,iLagrange), column-major.
void derivative_matrix(int N, const double* x, double* D) {
  std::vector<double> wb(N + 1);
  barycentric_weights(N, x, wb.data());
  const int n = N + 1;
  for (int i = 0; i < n * n; ++i) D[i] = 0.0;
  for (int l = 0; l <= N; ++l)
    for (int i = 0; i <= N; ++i)
      if (l != i) {
        D[i + n * l] = wb[l] / (wb[i] * (x[i] - x[l]));
        D[i + n * i] -= D[i + n * l];
      }
}

// Dense inverse by LU with partial pivoting; stands in for the LAPACK DGETRF/DGETRI pair
// used by INVERSE (globals/mathtools.f90:41-97). Results agree to round-off.
void invert(int n, const double* A, double* Ainv) {
  std::vector<double> a(A, A + n * n);
  std::vector<double> inv(n * n, 0.0);
  for (int i = 0; i < n; ++i) inv[i + n * i] = 1.0;
  for (int c = 0; c < n; ++c) {
    int p = c;
    for (int r = c + 1; r < n; ++r)
      if (std::fabs(a[r + n * c]) > std::fabs(a[p + n * c])) p = r;
    if (p != c)
      for (int k = 0; k < n; ++k) {
        std::swap(a[c + n * k], a[p + n * k]);
        std::swap(inv[c + n * k], inv[p + n * k]);
      }
    const double piv = a[c + n * c];
    for (int k = 0; k < n; ++k) { a[c + n * k] /= piv; inv[c + n * k] /= piv; }
    for (int r = 0; r < n; ++r) {
      if (r == c) continue;
      const double f = a[r + n * c];
      if (f == 0.0) continue;
      for (int k = 0; k < n; ++k) {
        a[r + n * k] -= f * a[c + n * k];
        inv[r + n * k] -= f * inv[c + n * k];
      }
    }
  }
  std::memcpy(Ainv, inv.data(), sizeof(double) * n * n);
}

// buildLegendreVdm, basis.f90:47-109 (LAPACK branch).
static void legendre_vdm(int N, const double* x, double* V, double* sV) {
  const int n = N + 1;
  double d;
  for (int i = 0; i <= N; ++i)
    for (int j = 0; j <= N; ++j) legendre_and_derivative(j, x[i], &V[i + n * j], &d);
  invert(n, V, sV);
}

void nodes(int N, int type, double* x, double* w) {
  switch (type) {
    case OR_NODE_G: gauss_nodes(N, x, w); break;
    case OR_NODE_GL: gauss_lobatto_nodes(N, x, w); break;
    case OR_NODE_CL: cheb_lobatto_nodes(N, x, w); break;
    case OR_NODE_VISU:
      for (int i = 0; i <= N; ++i) x[i] = 2.0 * (double)i / (double)N - 1.0;
      if (w) {
        for (int i = 0; i <= N; ++i) w[i] = 2.0 / (double)N;
        w[0] *= 0.5; w[N] *= 0.5;
      }
      break;
    default: std::abort();
  }
}

// GetVandermonde, interpolation.f90:291-358. Vdm(0:Nout,0:Nin), column-major.
void vandermonde(int Nin, int tin, int Nout, int tout, int modal, double* Vdm) {
  const int ni = Nin + 1, no = Nout + 1;
  if (tin == tout && Nin == Nout) {
    for (int i = 0; i < no * ni; ++i) Vdm[i] = 0.0;
    for (int i = 0; i <= Nin; ++i) Vdm[i + no * i] = 1.0;
    return;
  }
  std::vector<double> xi(ni), wb(ni), xo(no);
  nodes(Nin, tin, xi.data(), nullptr);
  barycentric_weights(Nin, xi.data(), wb.data());
  nodes(Nout, tout, xo.data(), nullptr);
  if (modal && Nout < Nin) {
    std::vector<double> Vi(ni * ni), sVi(ni * ni), Vo(no * no), sVo(no * no);
    legendre_vdm(Nin, xi.data(), Vi.data(), sVi.data());
    legendre_vdm(Nout, xo.data(), Vo.data(), sVo.data());
    // Vdm = MATMUL(Vdm_Leg_Out(0:Nout,0:Nout), sVdm_Leg_In(0:Nout,0:Nin))
    for (int i = 0; i <= Nout; ++i)
      for (int j = 0; j <= Nin; ++j) {
        double s = 0.0;
        for (int k = 0; k <= Nout; ++k) s += Vo[i + no * k] * sVi[k + ni * j];
        Vdm[i + no * j] = s;
      }
  } else {
    std::vector<double> L(ni);
    for (int i = 0; i <= Nout; ++i) {
      lagrange_polys(xo[i], Nin, xi.data(), wb.data(), L.data());
      for (int j = 0; j <= Nin; ++j) Vdm[i + no * j] = L[j];
    }
  }
}

// ---- mappings.f90 ----
// Flip_S2M, mappings.f90:188-221 (3D).
void flip_s2m(int N, int p, int q, int flip, int* o) {
  switch (flip) {
    case 0: o[0] = p; o[1] = q; break;
    case 1: o[0] = q; o[1] = p; break;
    case 2: o[0] = N - p; o[1] = q; break;
    case 3: o[0] = N - q; o[1] = N - p; break;
    case 4: o[0] = p; o[1] = N - q; break;
  }
}
// CGNS_VolToSide, mappings.f90:233-268 (3D). Side ids: 1 zeta-, 2 eta-, 3 xi+, 4 eta+, 5 xi-, 6 zeta+.
static void cgns_vol_to_side(int N, int i, int j, int k, int s, int* o) {
  o[0] = o[1] = o[2] = 0;
  switch (s) {
    case 5: o[0] = k; o[1] = j; o[2] = i; break;
    case 3: o[0] = j; o[1] = k; o[2] = N - i; break;
    case 2: o[0] = i; o[1] = k; o[2] = j; break;
    case 4: o[0] = N - i; o[1] = k; o[2] = N - j; break;
    case 1: o[0] = j; o[1] = i; o[2] = k; break;
    case 6: o[0] = i; o[1] = j; o[2] = N - k; break;
  }
}
// CGNS_SideToVol, mappings.f90:274-309 (3D).
static void cgns_side_to_vol(int N, int l, int p, int q, int s, int* o) {
  switch (s) {
    case 5: o[0] = l; o[1] = q; o[2] = p; break;
    case 3: o[0] = N - l; o[1] = p; o[2] = q; break;
    case 2: o[0] = p; o[1] = l; o[2] = q; break;
    case 4: o[0] = N - p; o[1] = N - l; o[2] = q; break;
    case 1: o[0] = q; o[1] = p; o[2] = l; break;
    case 6: o[0] = p; o[1] = q; o[2] = N - l; break;
  }
}
// CGNS_SideToVol2, mappings.f90:315-346 (3D).
static void cgns_side_to_vol2(int N, int p, int q, int s, int* o) {
  switch (s) {
    case 5: o[0] = q; o[1] = p; break;
    case 3: o[0] = p; o[1] = q; break;
    case 2: o[0] = p; o[1] = q; break;
    case 4: o[0] = N - p; o[1] = q; break;
    case 1: o[0] = q; o[1] = p; break;
    case 6: o[0] = p; o[1] = q; break;
  }
}
// SideToVol2, mappings.f90:436-446.
void side_to_vol2(int N, int p, int q, int flip, int s, int* o) {
  int pq[2];
  flip_s2m(N, p, q, flip, pq);  // Flip_M2S == Flip_S2M
  cgns_side_to_vol2(N, pq[0], pq[1], s, o);
}
// SideToVol, mappings.f90:421-431.
void side_to_vol(int N, int l, int p, int q, int flip, int s, int* o) {
  int pq[2];
  flip_s2m(N, p, q, flip, pq);
  cgns_side_to_vol(N, l, pq[0], pq[1], s, o);
}
// VolToSide, mappings.f90:404-415.
void vol_to_side(int N, int i, int j, int k, int flip, int s, int* o) {
  int pq[3] = {0, 0, 0};
  cgns_vol_to_side(N, i, j, k, s, pq);
  flip_s2m(N, pq[0], pq[1], flip, o);
  o[2] = pq[2];
}

}  // namespace orc

using namespace orc;

extern "C" {

void or_nodes_weights(int N, int type, double* x, double* w, double* wBary) {
  nodes(N, type, x, w);
  if (wBary) barycentric_weights(N, x, wBary);
}

// GetDerivativeMatrix, interpolation.f90:364-377.
void or_derivative_matrix(int N, int type, double* D) {
  std::vector<double> x(N + 1);
  nodes(N, type, x.data(), nullptr);
  derivative_matrix(N, x.data(), D);
}

void or_lagrange_polys(double x, int N, const double* xGP, const double* wb, double* L) {
  lagrange_polys(x, N, xGP, wb, L);
}

void or_vandermonde(int Nin, int tin, int Nout, int tout, int modal, double* Vdm) {
  vandermonde(Nin, tin, Nout, tout, modal, Vdm);
}

void or_dg_basis(int N, int type, double* xGP, double* wGP, double* wBary, double* L_Minus,
                 double* L_Plus, double* L_HatMinus, double* L_HatPlus, double* D, double* D_T,
                 double* D_Hat, double* D_Hat_T, double* DVolSurf) {
  const int n = N + 1;
  // InitInterpolationBasis, interpolation.f90:127-160
  nodes(N, type, xGP, wGP);
  barycentric_weights(N, xGP, wBary);
  lagrange_polys(1.0, N, xGP, wBary, L_Plus);
  lagrange_polys(-1.0, N, xGP, wBary, L_Minus);
  // InitDGBasis, dg.f90:132-209 (diagonal mass matrix: EXACT_MM is off)
  derivative_matrix(N, xGP, D);
  for (int i = 0; i < n; ++i)
    for (int j = 0; j < n; ++j) D_T[i + n * j] = D[j + n * i];
  // D_Hat = -Minv * D^T * M  ->  D_Hat(i,j) = -D(j,i) * w(j) / w(i)
  for (int i = 0; i < n; ++i)
    for (int j = 0; j < n; ++j) D_Hat[i + n * j] = -(1.0 / wGP[i]) * (D[j + n * i] * wGP[j]);
  for (int i = 0; i < n; ++i)
    for (int j = 0; j < n; ++j) D_Hat_T[i + n * j] = D_Hat[j + n * i];
  if (type == OR_NODE_GL) {
    // dg.f90:194-201
    for (int i = 0; i < n * n; ++i) DVolSurf[i] = D_T[i];
    DVolSurf[0] = DVolSurf[0] + 1.0 / (2.0 * wGP[0]);
    DVolSurf[N + n * N] = DVolSurf[N + n * N] - 1.0 / (2.0 * wGP[N]);
  } else {
    // Chan's S = D - 0.5 Minv Vf^T B Vf, transposed (dg.f90:184-194)
    for (int i = 0; i < n; ++i)
      for (int j = 0; j < n; ++j) {
        const double vbv = -L_Minus[i] * L_Minus[j] + L_Plus[i] * L_Plus[j];
        DVolSurf[j + n * i] = D[i + n * j] - 0.5 * (1.0 / wGP[i]) * vbv;
      }
  }
  for (int i = 0; i < n; ++i) {
    L_HatPlus[i] = (1.0 / wGP[i]) * L_Plus[i];
    L_HatMinus[i] = (1.0 / wGP[i]) * L_Minus[i];
  }
}

// buildMappings, mappings.f90:41-182. Layouts:
// S2V2(2,0:N,0:N,0:4,1:6), S2V(3,0:N,0:N,0:N,0:4,1:6), V2S(3,0:N,0:N,0:N,0:4,1:6).
void or_mappings(int N, int* S2V2, int* S2V, int* V2S) {
  const int n = N + 1;
  for (int s = 1; s <= 6; ++s)
    for (int f = 0; f <= 4; ++f) {
      for (int q = 0; q <= N; ++q)
        for (int p = 0; p <= N; ++p) {
          int o[2];
          side_to_vol2(N, p, q, f, s, o);
          const int b = 2 * (p + n * (q + n * (f + 5 * (s - 1))));
          if (S2V2) { S2V2[b] = o[0]; S2V2[b + 1] = o[1]; }
        }
      for (int k = 0; k <= N; ++k)
        for (int j = 0; j <= N; ++j)
          for (int i = 0; i <= N; ++i) {
            const int b = 3 * (i + n * (j + n * (k + n * (f + 5 * (s - 1)))));
            int o[3];
            if (S2V) {
              side_to_vol(N, i, j, k, f, s, o);  // (l,p,q) = (i,j,k)
              S2V[b] = o[0]; S2V[b + 1] = o[1]; S2V[b + 2] = o[2];
            }
            if (V2S) {
              vol_to_side(N, i, j, k, f, s, o);
              V2S[b] = o[0]; V2S[b + 1] = o[1]; V2S[b + 2] = o[2];
            }
          }
    }
}

}  // extern "C"
