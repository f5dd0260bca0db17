// Host-side 1D polynomial tools: Legendre/Lobatto/Chebyshev/equidistant node sets, barycentric
// Lagrange interpolation, differentiation and modal Vandermonde matrices. Textbook algorithms
// (Kopriva, "Implementing Spectral Methods for Partial Differential Equations", ch. 3), used to
// build the operators FLEXI sets up in src/interpolation/{basis,interpolation}.f90.
#pragma once
#include <cmath>
#include <vector>

namespace fxg {
namespace basis {

enum NodeSet { GAUSS = 1, GAUSS_LOBATTO = 2, CHEBYSHEV_LOBATTO = 3, EQUIDISTANT = 4 };

// Legendre polynomial L_N(x) and L_N'(x) by the three-term recurrence (not normalised).
inline void legendre(int N, double x, double& L, double& dL) {
  if (N == 0) { L = 1.0; dL = 0.0; return; }
  double Lm = 1.0, Lc = x, dLm = 0.0, dLc = 1.0;
  for (int k = 2; k <= N; ++k) {
    const double Ln = ((2.0 * k - 1.0) * x * Lc - (k - 1.0) * Lm) / k;
    const double dLn = dLm + (2.0 * k - 1.0) * Lc;
    Lm = Lc; Lc = Ln; dLm = dLc; dLc = dLn;
  }
  L = Lc;
  dL = dLc;
}

// Gauss-Legendre nodes/weights: Newton on L_{N+1} from Chebyshev guesses.
inline void gauss(int N, double* x, double* w) {
  const int n = N + 1;
  for (int i = 0; i < (n + 1) / 2; ++i) {
    double z = -std::cos(M_PI * (2.0 * i + 1.0) / (2.0 * N + 2.0));
    double L, dL;
    for (int it = 0; it < 100; ++it) {
      legendre(n, z, L, dL);
      const double dz = -L / dL;
      z += dz;
      if (std::fabs(dz) <= 1e-16 * std::fabs(z)) break;
    }
    legendre(n, z, L, dL);
    x[i] = z;
    x[N - i] = -z;
    if (w) w[i] = w[N - i] = 2.0 / ((1.0 - z * z) * dL * dL);
  }
  if (N % 2 == 0) {
    double L, dL;
    legendre(n, 0.0, L, dL);
    x[N / 2] = 0.0;
    if (w) w[N / 2] = 2.0 / (dL * dL);
  }
}

// Gauss-Lobatto nodes/weights: Newton on L_{N+1} - L_{N-1}.
inline void gauss_lobatto(int N, double* x, double* w) {
  x[0] = -1.0; x[N] = 1.0;
  if (w) w[0] = w[N] = 2.0 / (N * (N + 1.0));
  for (int i = 1; i < (N + 1) / 2; ++i) {
    double z = -std::cos(M_PI * (i + 0.25) / N - 3.0 / (8.0 * N * M_PI * (i + 0.25)));
    for (int it = 0; it < 100; ++it) {
      double Lp, dLp, Lm, dLm;
      legendre(N + 1, z, Lp, dLp);
      legendre(N - 1, z, Lm, dLm);
      const double dz = -(Lp - Lm) / (dLp - dLm);
      z += dz;
      if (std::fabs(dz) <= 1e-16 * std::fabs(z)) break;
    }
    double L, dL;
    legendre(N, z, L, dL);
    x[i] = z;
    x[N - i] = -z;
    if (w) w[i] = w[N - i] = 2.0 / (N * (N + 1.0) * L * L);
  }
  if (N % 2 == 0) {
    double L, dL;
    legendre(N, 0.0, L, dL);
    x[N / 2] = 0.0;
    if (w) w[N / 2] = 2.0 / (N * (N + 1.0) * L * L);
  }
}

inline void nodes(int N, int set, double* x, double* w = nullptr) {
  switch (set) {
    case GAUSS: gauss(N, x, w); break;
    case GAUSS_LOBATTO: gauss_lobatto(N, x, w); break;
    case CHEBYSHEV_LOBATTO:
      for (int i = 0; i <= N; ++i) x[i] = -std::cos(M_PI * i / N);
      break;
    default:
      for (int i = 0; i <= N; ++i) x[i] = 2.0 * i / N - 1.0;
      break;
  }
}

inline void bary_weights(int N, const double* x, double* wb) {
  for (int j = 0; j <= N; ++j) {
    double p = 1.0;
    for (int k = 0; k <= N; ++k)
      if (k != j) p *= (x[j] - x[k]);
    wb[j] = 1.0 / p;
  }
}

// Lagrange basis values at z (barycentric formula of the second kind).
inline void lagrange_at(double z, int N, const double* x, const double* wb, double* L) {
  for (int j = 0; j <= N; ++j)
    if (std::fabs(z - x[j]) <= 2.2e-16 * std::fmax(1.0, std::fabs(z))) {
      for (int k = 0; k <= N; ++k) L[k] = (k == j) ? 1.0 : 0.0;
      return;
    }
  double s = 0.0;
  for (int j = 0; j <= N; ++j) { L[j] = wb[j] / (z - x[j]); s += L[j]; }
  for (int j = 0; j <= N; ++j) L[j] /= s;
}

// D(i,l) = l_l'(x_i), column-major, diagonal from the negative sum.
inline void diff_matrix(int N, const double* x, double* D) {
  std::vector<double> wb(N + 1);
  bary_weights(N, x, wb.data());
  const int n = N + 1;
  for (int i = 0; i < n; ++i) {
    double diag = 0.0;
    for (int l = 0; l < n; ++l) {
      if (l == i) continue;
      D[i + n * l] = wb[l] / (wb[i] * (x[i] - x[l]));
      diag -= D[i + n * l];
    }
    D[i + n * i] = diag;
  }
}

inline bool invert(int n, std::vector<double> a, std::vector<double>& inv) {
  inv.assign((size_t)n * n, 0.0);
  for (int i = 0; i < n; ++i) inv[i + n * i] = 1.0;
  for (int c = 0; c < n; ++c) {
    int p = c;
    for (int r = c + 1; r < n; ++r)
      if (std::fabs(a[r + n * c]) > std::fabs(a[p + n * c])) p = r;
    if (a[p + n * c] == 0.0) return false;
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
  return true;
}

// Interpolation matrix V(0:Nout,0:Nin) from node set `sin` to `sout`; with `modal` and
// Nout < Nin the high Legendre modes are truncated instead (GetVandermonde semantics,
// src/interpolation/interpolation.f90:291-358).
inline std::vector<double> vandermonde(int Nin, int sin, int Nout, int sout, bool modal) {
  const int ni = Nin + 1, no = Nout + 1;
  std::vector<double> V((size_t)no * ni, 0.0);
  if (sin == sout && Nin == Nout) {
    for (int i = 0; i < ni; ++i) V[i + no * i] = 1.0;
    return V;
  }
  std::vector<double> xi(ni), xo(no), wb(ni), L(ni);
  nodes(Nin, sin, xi.data());
  nodes(Nout, sout, xo.data());
  bary_weights(Nin, xi.data(), wb.data());
  if (modal && Nout < Nin) {
    // nodal -> Legendre coefficients on the input nodes, truncate, evaluate on the output nodes
    std::vector<double> Vi((size_t)ni * ni), sVi;
    for (int i = 0; i < ni; ++i)
      for (int j = 0; j < ni; ++j) {
        double l, dl;
        legendre(j, xi[i], l, dl);
        Vi[i + ni * j] = l;
      }
    invert(ni, Vi, sVi);
    for (int i = 0; i < no; ++i)
      for (int j = 0; j < ni; ++j) {
        double s = 0.0;
        for (int k = 0; k < no; ++k) {
          double l, dl;
          legendre(k, xo[i], l, dl);
          s += l * sVi[k + ni * j];
        }
        V[i + no * j] = s;
      }
    return V;
  }
  for (int i = 0; i < no; ++i) {
    lagrange_at(xo[i], Nin, xi.data(), wb.data(), L.data());
    for (int j = 0; j < ni; ++j) V[i + no * j] = L[j];
  }
  return V;
}

}  // namespace basis
}  // namespace fxg
