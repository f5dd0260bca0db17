// Boundary conditions of the Navier-Stokes equations at boundary side points, FP64
// (src/equations/navierstokes/idealgas/getboundaryflux.f90). Supported BC types: 2 (Dirichlet
// state RefStatePrim(:,BCState), Riemann solver against it), 3 (adiabatic no-slip wall),
// 4 (isothermal no-slip wall at RefStatePrim(TEMP,BCState)), 9 (slip wall).
#pragma once
#include "fxg_physics.h"

namespace fxg {

enum { BC_DIRICHLET = 2, BC_WALL_ADIABATIC = 3, BC_WALL_ISOTHERMAL = 4, BC_WALL_SLIP = 9 };

// PRESSURE_RIEMANN (idealgas/eos.f90:561-584): pressure of the Riemann problem of the inner state
// against a wall; vn is the normal velocity (positive: towards the wall, shock).
__device__ __forceinline__ double wall_pressure(const Gas& g, double rho, double vn, double p) {
  const double kappaFac = 2.0 * g.kappa * g.sKappaM1;
  if (vn <= 0.0) return p * pow(fmax(0.0001, 1.0 + 0.5 * g.kappaM1 * vn / sqrt(g.kappa * p / rho)), kappaFac);
  const double sKappaP1 = 1.0 / (g.kappa + 1.0);
  const double ar = 2.0 * sKappaP1 / rho;
  const double br = g.kappaM1 * sKappaP1 * p;
  return p + vn / ar * 0.5 * (vn + sqrt(vn * vn + 4.0 * ar * (p + br)));
}

// GetBoundaryState (getboundaryflux.f90:224-472): the outer primitive state Pb[6] =
// (rho,u,v,w,p,T) of a boundary point with inner state Pm and side frame (nv, t1, t2).
__device__ __forceinline__ void boundary_state(const Gas& g, int type, const double* ref, const double* Pm,
                                               const double* nv, const double* t1, const double* t2, double* Pb) {
  if (type == BC_DIRICHLET) {
#pragma unroll
    for (int v = 0; v < 6; ++v) Pb[v] = ref[v];
    return;
  }
  double rho = Pm[0];
  double a = Pm[1] * nv[0] + Pm[2] * nv[1] + Pm[3] * nv[2];
  double b = Pm[1] * t1[0] + Pm[2] * t1[1] + Pm[3] * t1[2];
  double c = Pm[1] * t2[0] + Pm[2] * t2[1] + Pm[3] * t2[2];
  const double p = wall_pressure(g, rho, a, Pm[4]);
  double T;
  if (type == BC_WALL_SLIP) {
    a = 0.0;
    T = p / (rho * g.R);
  } else {
    a = b = c = 0.0;
    T = type == BC_WALL_ADIABATIC ? Pm[5] : ref[5];
    rho = p / (T * g.R);
  }
  Pb[0] = rho;
#pragma unroll
  for (int d = 0; d < 3; ++d) Pb[1 + d] = a * nv[d] + b * t1[d] + c * t2[d];
  Pb[4] = p;
  Pb[5] = T;
}

// Lifting_GetBoundaryFlux (getboundaryflux.f90:855-939), strong form, before SurfElem: boundary
// value minus inner value of the lifting variables (u,v,w,T).
__device__ __forceinline__ void boundary_lifting_flux(int type, const double* Pm, const double* Pb, double* F) {
#pragma unroll
  for (int v = 0; v < 4; ++v) {
    const int k = v < 3 ? 1 + v : 5;
    double f;
    if (type == BC_DIRICHLET) f = 0.5 * (Pm[k] + Pb[k]);
    else if (type == BC_WALL_SLIP) f = v < 3 ? Pb[k] : Pm[5];
    else f = v < 3 ? 0.0 : Pb[5];
    F[v] = f - Pm[k];
  }
}

}  // namespace fxg
