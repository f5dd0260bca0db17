// Internal helpers shared by the oracle translation units. TEST INFRASTRUCTURE ONLY.
#pragma once

namespace orc {
extern const double PP_Pi;
extern const double PP_RealTolerance;
void legendre_and_derivative(int N, double x, double* L, double* Lder);
void barycentric_weights(int N, const double* x, double* wb);
void lagrange_polys(double x, int N, const double* xGP, const double* wb, double* L);
void derivative_matrix(int N, const double* x, double* D);
void invert(int n, const double* A, double* Ainv);
void nodes(int N, int type, double* x, double* w);
void vandermonde(int Nin, int tin, int Nout, int tout, int modal, double* Vdm);
void flip_s2m(int N, int p, int q, int flip, int* o);
void side_to_vol2(int N, int p, int q, int flip, int s, int* o);
void side_to_vol(int N, int l, int p, int q, int flip, int s, int* o);
void vol_to_side(int N, int i, int j, int k, int flip, int s, int* o);
}  // namespace orc
