// Boundary conditions of the Navier-Stokes equations at boundary side points, FP64
// (src/equations/navierstokes/idealgas/getboundaryflux.f90). Supported BC types:
//  Riemann type (Riemann solver and BR1 average against a boundary state, viscous flux with the
//  inner gradients): 2 (RefStatePrim(:,BCState)), 12 and 121 (a per-point state given by the host,
//  BCDataPrim: a State file, or an exact function evaluated once), 23 (outflow at the Mach number
//  RefStatePrim(2,BCState)), 24 (pressure outflow at RefStatePrim(5,BCState)), 25 (subsonic outflow
//  at that pressure), 27 (subsonic inflow at total temperature RefStatePrim(1,BCState), direction
//  RefStatePrim(2:4,BCState) and total pressure RefStatePrim(5,BCState), as InitBC leaves them);
//  walls (inviscid flux (0, p_wall n, 0)): 3 (adiabatic no-slip), 4 (isothermal no-slip at
//  RefStatePrim(TEMP,BCState)), 9 (slip), 91 (slip, velocity gradients with the wall-normal
//  derivative of the tangential velocity and the tangential derivatives of the normal velocity
//  removed).
#pragma once
#include "fxg_physics.h"

namespace fxg {

enum {
  BC_DIRICHLET = 2,
  BC_WALL_ADIABATIC = 3,
  BC_WALL_ISOTHERMAL = 4,
  BC_WALL_SLIP = 9,
  BC_STATE_DATA = 12,
  BC_EXACT_DATA = 121,
  BC_OUTFLOW_MACH = 23,
  BC_OUTFLOW_PRESSURE = 24,
  BC_OUTFLOW_SUBSONIC = 25,
  BC_INFLOW_SUBSONIC = 27,
  BC_WALL_SLIP_GRAD = 91
};

// Riemann-type boundaries: the boundary flux is the Riemann flux against the boundary state.
__host__ __device__ __forceinline__ bool bc_riemann_type(int type) {
  return type == BC_DIRICHLET || type == BC_STATE_DATA || type == BC_EXACT_DATA || type == BC_OUTFLOW_MACH ||
         type == BC_OUTFLOW_PRESSURE || type == BC_OUTFLOW_SUBSONIC || type == BC_INFLOW_SUBSONIC;
}
__host__ __device__ __forceinline__ bool bc_slip_wall(int type) {
  return type == BC_WALL_SLIP || type == BC_WALL_SLIP_GRAD;
}

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
// (rho,u,v,w,p,T) of a boundary point with inner state Pm and side frame (nv, t1, t2). ref is
// RefStatePrim(:,BCState) (types 2, 4, 23, 24, 25, 27), data the point's BCDataPrim (12, 121).
__device__ __forceinline__ void boundary_state(const Gas& g, int type, const double* ref, const double* data,
                                               const double* Pm, const double* nv, const double* t1,
                                               const double* t2, double* Pb) {
  if (type == BC_DIRICHLET || type == BC_STATE_DATA || type == BC_EXACT_DATA) {
    const double* s = type == BC_DIRICHLET ? ref : data;
#pragma unroll
    for (int v = 0; v < 6; ++v) Pb[v] = s[v];
    return;
  }
  // the inner state in the side frame (a: normal, b, c: tangential velocity)
  double rho = Pm[0], p = Pm[4], T = Pm[5];
  double a = Pm[1] * nv[0] + Pm[2] * nv[1] + Pm[3] * nv[2];
  double b = Pm[1] * t1[0] + Pm[2] * t1[1] + Pm[3] * t1[2];
  double c = Pm[1] * t2[0] + Pm[2] * t2[1] + Pm[3] * t2[2];
  if (type == BC_WALL_ADIABATIC || type == BC_WALL_ISOTHERMAL) {
    p = wall_pressure(g, rho, a, p);
    a = b = c = 0.0;
    T = type == BC_WALL_ADIABATIC ? Pm[5] : ref[5];
    rho = p / (T * g.R);
  } else if (bc_slip_wall(type)) {
    p = wall_pressure(g, rho, a, p);
    a = 0.0;
    T = p / (rho * g.R);
  } else if (type == BC_OUTFLOW_MACH) {
    // NASA/TM-2011-217181 eqs. (19)-(24): the total pressure at the prescribed Mach number
    const double MaOut = ref[1];
    const double cs = sqrt(g.kappa * p / rho);
    const double Ma = a / cs;
    double pb;
    if (Ma < 1.0) {
      const double pt = p * pow(1.0 + 0.5 * g.kappaM1 * Ma * Ma, g.kappa * g.sKappaM1);
      pb = pt * pow(1.0 + 0.5 * g.kappaM1 * MaOut * MaOut, -g.kappa * g.sKappaM1);
    } else {
      pb = p + 0.5 * rho * (a * a + b * b + c * c);
    }
    rho = g.kappa * pb / (cs * cs);
    p = pb;
    T = p / (g.R * rho);
  } else if (type == BC_OUTFLOW_PRESSURE) {
    const double cs = sqrt(g.kappa * p / rho);
    if (a / cs < 1.0) {  // subsonic: the prescribed pressure; supersonic: the inner state
      p = ref[4];
      rho = g.kappa * p / (cs * cs);
      T = p / (g.R * rho);
    }
  } else if (type == BC_OUTFLOW_SUBSONIC) {
    const double cs = sqrt(g.kappa * p / rho);
    const double pb = (a / cs < 1.0) ? ref[4] : p + 0.5 * rho * (a * a + b * b + c * c);
    if (a < 0.0) {  // inflow through an outflow boundary: the normal velocity turned outwards
      a = fabs(a);
      b = c = 0.0;
    }
    rho = g.kappa * pb / (cs * cs);
    p = ref[4];
    T = p / (g.R * rho);
  } else {  // BC_INFLOW_SUBSONIC, NASA/TM-2011-217181 with the corrections of Slater (AIAA 2001-3882)
    const double Tt = ref[0], pt = ref[4];
    const double A = -(ref[1] * nv[0] + ref[2] * nv[1] + ref[3] * nv[2]);
    const double cs = sqrt(g.kappa * p / rho);
    const double Rplus = -a - 2.0 * cs / g.kappaM1;
    const double q1 = A * A + 2.0 / g.kappaM1;
    const double q2 = 2.0 * Rplus;
    const double q3 = g.kappaM1 / 2.0 * Rplus * Rplus - g.kappa * g.R * Tt * A * A;
    const double disc = sqrt(q2 * q2 - 4.0 * q1 * q3);
    const double cb = fmax((-q2 + disc) / (2.0 * q1), (-q2 - disc) / (2.0 * q1));
    const double Tb = cb * cb / (g.kappa * g.R);
    const double Ma = sqrt(2.0 / g.kappaM1 * (Tt / Tb - 1.0));
    const double pb = pt * pow(1.0 + 0.5 * g.kappaM1 * Ma * Ma, -g.kappa / g.kappaM1);
    const double Ub = Ma * sqrt(g.kappa * g.R * Tb);
    rho = pb / (g.R * Tb);
    a = Ub * (ref[1] * nv[0] + ref[2] * nv[1] + ref[3] * nv[2]);
    b = Ub * (ref[1] * t1[0] + ref[2] * t1[1] + ref[3] * t1[2]);
    c = Ub * (ref[1] * t2[0] + ref[2] * t2[1] + ref[3] * t2[2]);
    p = pb;
    T = Tb;
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
    if (bc_riemann_type(type)) f = 0.5 * (Pm[k] + Pb[k]);
    else if (bc_slip_wall(type)) f = v < 3 ? Pb[k] : Pm[5];
    else f = v < 3 ? 0.0 : Pb[5];
    F[v] = f - Pm[k];
  }
}

// The gradients a slip wall's viscous flux sees (getboundaryflux.f90:622-776), in place on
// q = (d/dx, d/dy, d/dz) x (u, v, w, T) (index 4*direction + variable): type 9 removes the
// wall-normal part of every gradient, (I - n n^T) grad; type 91 does so for the temperature and,
// for the velocity gradient G (G_ab = d u_a / d x_b), zeroes in Z = E^T G E, E = (n, t1, t2), the
// wall-normal derivatives of the tangential velocity and the tangential derivatives of the normal
// velocity, G' = E Z E^T.
template <class T>
__device__ __forceinline__ void slip_wall_gradients(int type, const T* nv, const T* t1, const T* t2, T* q) {
  T M[3][3];
#pragma unroll
  for (int a = 0; a < 3; ++a)
#pragma unroll
    for (int b = 0; b < 3; ++b) M[a][b] = a == b ? T(1) - nv[a] * nv[a] : -nv[a] * nv[b];
  const int nProj = type == BC_WALL_SLIP ? 4 : 1;
#pragma unroll
  for (int i = 0; i < 4; ++i) {
    if (i >= nProj) break;
    const int v = type == BC_WALL_SLIP ? i : 3;
    const T gx = q[v], gy = q[4 + v], gz = q[8 + v];
#pragma unroll
    for (int a = 0; a < 3; ++a) q[4 * a + v] = M[a][0] * gx + M[a][1] * gy + M[a][2] * gz;
  }
  if (type != BC_WALL_SLIP_GRAD) return;
  const T* E[3] = {nv, t1, t2};
  // W[c][b] = d(u.e_c)/dx_b
  T W[3][3];
#pragma unroll
  for (int cc = 0; cc < 3; ++cc)
#pragma unroll
    for (int b = 0; b < 3; ++b) W[cc][b] = E[cc][0] * q[4 * b] + E[cc][1] * q[4 * b + 1] + E[cc][2] * q[4 * b + 2];
  // Z[c][d] = derivative of component c along e_d
  T Z[3][3];
#pragma unroll
  for (int cc = 0; cc < 3; ++cc)
#pragma unroll
    for (int d = 0; d < 3; ++d) Z[cc][d] = E[d][0] * W[cc][0] + E[d][1] * W[cc][1] + E[d][2] * W[cc][2];
  Z[1][0] = Z[2][0] = Z[0][1] = Z[0][2] = T(0);
  // back: W'[c][b] = sum_d e_d[b] Z[c][d]; G'_ab = sum_c e_c[a] W'[c][b]
#pragma unroll
  for (int cc = 0; cc < 3; ++cc)
#pragma unroll
    for (int b = 0; b < 3; ++b) W[cc][b] = nv[b] * Z[cc][0] + t1[b] * Z[cc][1] + t2[b] * Z[cc][2];
#pragma unroll
  for (int b = 0; b < 3; ++b)
#pragma unroll
    for (int a = 0; a < 3; ++a) q[4 * b + a] = nv[a] * W[0][b] + t1[a] * W[1][b] + t2[a] * W[2][b];
}

}  // namespace fxg
