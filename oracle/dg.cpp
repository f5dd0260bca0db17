// CPU oracle, part 3: point physics, the DG right-hand side with BR1 lifting, the CFL time step,
// the LSERK stepper and the Taylor-Green vortex test case.
// TEST INFRASTRUCTURE ONLY. Data layout is FLEXI's (column-major, variable fastest).
#include "fxg_oracle.h"
#include "oracle_internal.h"

#include <omp.h>

#include <cmath>
#include <cstring>
#include <vector>

using namespace orc;

namespace {

enum { DENS = 0, MOM1 = 1, MOM2 = 2, MOM3 = 3, ENER = 4 };                    // eos.h
enum { VEL1 = 1, VEL2 = 2, VEL3 = 3, PRES = 4, TEMP = 5 };                     // eos.h
// extended state U(1:11): cons 0..4, then SRHO 5, VEL 6..8, PRES 9, TEMP 10 (eos.h EXT_*)
enum { E_DENS = 0, E_MOM1 = 1, E_ENER = 4, E_SRHO = 5, E_VEL1 = 6, E_VEL2 = 7, E_VEL3 = 8, E_PRES = 9 };

struct EosC {
  double kappa, kappaM1, sKappaM1, R, cp, cv, Pr, mu0, KappasPr, s43, s23;
};

EosC make_eos(const or_params* p) {
  // InitEos, idealgas/eos.f90:106-196; s43/s23 from equation.f90
  EosC e;
  e.kappa = p->kappa;
  e.kappaM1 = p->kappa - 1.0;
  e.sKappaM1 = 1.0 / e.kappaM1;
  e.R = p->R;
  e.cp = e.R * e.kappa / (e.kappa - 1.0);
  e.cv = e.R / (e.kappa - 1.0);
  e.Pr = p->Pr;
  e.KappasPr = e.kappa / e.Pr;
  e.mu0 = p->mu0;
  e.s43 = 4.0 / 3.0;
  e.s23 = 2.0 / 3.0;
  return e;
}

// ConsToPrim, idealgas/eos.f90:203-230
inline void cons_to_prim(const EosC& e, const double* c, double* p) {
  const double sRho = 1.0 / c[DENS];
  p[DENS] = c[DENS];
  p[VEL1] = c[MOM1] * sRho;
  p[VEL2] = c[MOM2] * sRho;
  p[VEL3] = c[MOM3] * sRho;
  p[PRES] = e.kappaM1 * (c[ENER] - 0.5 * (c[MOM1] * p[VEL1] + c[MOM2] * p[VEL2] + c[MOM3] * p[VEL3]));
  p[TEMP] = p[PRES] * sRho / e.R;
}

// PrimToCons, idealgas/eos.f90:305-328
inline void prim_to_cons(const EosC& e, const double* p, double* c) {
  c[DENS] = p[DENS];
  c[MOM1] = p[VEL1] * p[DENS];
  c[MOM2] = p[VEL2] * p[DENS];
  c[MOM3] = p[VEL3] * p[DENS];
  c[ENER] = e.sKappaM1 * p[PRES] + 0.5 * (c[MOM1] * p[VEL1] + c[MOM2] * p[VEL2] + c[MOM3] * p[VEL3]);
}

// EvalFlux3D_Point, navierstokes/flux.f90:66-119
inline void eval_flux(const double* U, const double* P, double* f, double* g, double* h) {
  const double Ep = U[ENER] + P[PRES];
  f[DENS] = U[MOM1];
  f[MOM1] = U[MOM1] * P[VEL1] + P[PRES];
  f[MOM2] = U[MOM1] * P[VEL2];
  f[MOM3] = U[MOM1] * P[VEL3];
  f[ENER] = Ep * P[VEL1];
  g[DENS] = U[MOM2];
  g[MOM1] = f[MOM2];
  g[MOM2] = U[MOM2] * P[VEL2] + P[PRES];
  g[MOM3] = U[MOM2] * P[VEL3];
  g[ENER] = Ep * P[VEL2];
  h[DENS] = U[MOM3];
  h[MOM1] = f[MOM3];
  h[MOM2] = g[MOM3];
  h[MOM3] = U[MOM3] * P[VEL3] + P[PRES];
  h[ENER] = Ep * P[VEL3];
}

// EvalDiffFlux3D_Point, navierstokes/flux.f90:149-241 (constant viscosity, PP_VISC=0).
// Gradients hold the lifting variables (u,v,w,T).
inline void eval_diff_flux(const EosC& e, const double* P, const double* gx, const double* gy,
                           const double* gz, double* f, double* g, double* h) {
  const double muS = e.mu0;
  const double lambda = muS * e.cp / e.Pr;  // THERMAL_CONDUCTIVITY_H, eos.h
  const double v1 = P[VEL1], v2 = P[VEL2], v3 = P[VEL3];
  const double tau_xx = muS * (e.s43 * gx[0] - e.s23 * gy[1] - e.s23 * gz[2]);
  const double tau_yy = muS * (-e.s23 * gx[0] + e.s43 * gy[1] - e.s23 * gz[2]);
  const double tau_zz = muS * (-e.s23 * gx[0] - e.s23 * gy[1] + e.s43 * gz[2]);
  const double tau_xy = muS * (gy[0] + gx[1]);
  const double tau_xz = muS * (gz[0] + gx[2]);
  const double tau_yz = muS * (gz[1] + gy[2]);
  f[DENS] = 0.0;
  f[MOM1] = -tau_xx;
  f[MOM2] = -tau_xy;
  f[MOM3] = -tau_xz;
  f[ENER] = -tau_xx * v1 - tau_xy * v2 - tau_xz * v3 - lambda * gx[3];
  g[DENS] = 0.0;
  g[MOM1] = -tau_xy;
  g[MOM2] = -tau_yy;
  g[MOM3] = -tau_yz;
  g[ENER] = -tau_xy * v1 - tau_yy * v2 - tau_yz * v3 - lambda * gy[3];
  h[DENS] = 0.0;
  h[MOM1] = -tau_xz;
  h[MOM2] = -tau_yz;
  h[MOM3] = -tau_zz;
  h[ENER] = -tau_xz * v1 - tau_yz * v2 - tau_zz * v3 - lambda * gz[3];
}

// EvalEulerFlux1D_fast, navierstokes/flux.f90:396-417
inline void euler_flux_1d_fast(const double* UE, double* F) {
  F[DENS] = UE[E_MOM1];
  F[MOM1] = UE[E_MOM1] * UE[E_VEL1] + UE[E_PRES];
  F[MOM2] = UE[E_MOM1] * UE[E_VEL2];
  F[MOM3] = UE[E_MOM1] * UE[E_VEL3];
  F[ENER] = (UE[E_ENER] + UE[E_PRES]) * UE[E_VEL1];
}

// SplitSurfaceFluxPI, navierstokes/splitflux.f90:672-700
inline void split_surface_flux_pi(const double* L, const double* R, double* F) {
  const double H_LL = (L[E_ENER] + L[E_PRES]) / L[E_DENS];
  const double H_RR = (R[E_ENER] + R[E_PRES]) / R[E_DENS];
  const double rs = L[E_DENS] + R[E_DENS];
  const double us = L[E_VEL1] + R[E_VEL1];
  F[DENS] = 0.25 * rs * us;
  F[MOM1] = 0.125 * rs * (us * us) + 0.5 * (L[E_PRES] + R[E_PRES]);
  F[MOM2] = 0.125 * rs * us * (L[E_VEL2] + R[E_VEL2]);
  F[MOM3] = 0.125 * rs * us * (L[E_VEL3] + R[E_VEL3]);
  F[ENER] = 0.125 * rs * (H_LL + H_RR) * us;
}

// SplitVolumeFluxPI, navierstokes/splitflux.f90:603-666
void split_volume_flux_pi(const double* URef, const double* PRef, const double* U, const double* P,
                          const double* MRef, const double* M, double* Flux) {
  const double HRef = (URef[ENER] + PRef[PRES]) / URef[DENS];
  const double H = (U[ENER] + P[PRES]) / U[DENS];
  const double rs = URef[DENS] + U[DENS];
  const double u = PRef[VEL1] + P[VEL1], v = PRef[VEL2] + P[VEL2], w = PRef[VEL3] + P[VEL3];
  const double ps = PRef[PRES] + P[PRES];
  double f[5], g[5], h[5];
  f[DENS] = 0.5 * rs * u;
  f[MOM1] = 0.25 * rs * (u * u) + ps;
  f[MOM2] = 0.25 * rs * u * v;
  f[MOM3] = 0.25 * rs * u * w;
  f[ENER] = 0.25 * rs * u * (HRef + H);
  g[DENS] = 0.5 * rs * v;
  g[MOM1] = f[MOM2];
  g[MOM2] = 0.25 * rs * (v * v) + ps;
  g[MOM3] = 0.25 * rs * v * w;
  g[ENER] = 0.25 * rs * v * (HRef + H);
  h[DENS] = 0.5 * rs * w;
  h[MOM1] = f[MOM3];
  h[MOM2] = g[MOM3];
  h[MOM3] = 0.25 * rs * (w * w) + ps;
  h[ENER] = 0.25 * rs * w * (HRef + H);
  const double m1 = 0.5 * (MRef[0] + M[0]), m2 = 0.5 * (MRef[1] + M[1]), m3 = 0.5 * (MRef[2] + M[2]);
  for (int v5 = 0; v5 < 5; ++v5) Flux[v5] = m1 * f[v5] + m3 * h[v5] + m2 * g[v5];
}

// Riemann_LF (riemann.f90:558-589), Riemann_Roe (:668-740), Riemann_RoeEntropyFix (:748-855),
// weak (average of F_L,F_R) or split (SplitSurfaceFluxPI) central part.
void riemann_solver(const EosC& e, int which, int split, const double* F_L, const double* F_R,
                    const double* L, const double* R, double* F) {
  if (which == OR_RIEMANN_LF) {
    const double cL = std::sqrt(e.kappa * L[E_PRES] * L[E_SRHO]);
    const double cR = std::sqrt(e.kappa * R[E_PRES] * R[E_SRHO]);
    const double lmax = std::fmax(std::fabs(R[E_VEL1]), std::fabs(L[E_VEL1])) + std::fmax(cL, cR);
    if (!split) {
      for (int v = 0; v < 5; ++v) F[v] = 0.5 * ((F_L[v] + F_R[v]) - lmax * (R[v] - L[v]));
    } else {
      split_surface_flux_pi(L, R, F);
      for (int v = 0; v < 5; ++v) F[v] = F[v] - 0.5 * lmax * (R[v] - L[v]);
    }
    return;
  }
  const double H_L = (L[E_ENER] + L[E_PRES]) * L[E_SRHO];  // TOTALENTHALPY_HE
  const double H_R = (R[E_ENER] + R[E_PRES]) * R[E_SRHO];
  const double sqL = std::sqrt(L[E_DENS]), sqR = std::sqrt(R[E_DENS]);
  const double ss = 1.0 / (sqL + sqR);
  double RoeVel[3];
  for (int c = 0; c < 3; ++c) RoeVel[c] = (sqR * R[E_VEL1 + c] + sqL * L[E_VEL1 + c]) * ss;
  const double absVel = RoeVel[0] * RoeVel[0] + RoeVel[1] * RoeVel[1] + RoeVel[2] * RoeVel[2];
  const double RoeH = (sqR * H_R + sqL * H_L) * ss;
  const double Roec = std::sqrt(e.kappaM1 * (RoeH - 0.5 * absVel));  // ROEC_RIEMANN_H
  double a[5] = {RoeVel[0] - Roec, RoeVel[0], RoeVel[0], RoeVel[0], RoeVel[0] + Roec};
  const double r1[5] = {1., a[0], RoeVel[1], RoeVel[2], RoeH - RoeVel[0] * Roec};
  const double r2[5] = {1., RoeVel[0], RoeVel[1], RoeVel[2], 0.5 * absVel};
  const double r3[5] = {0., 0., 1., 0., RoeVel[1]};
  const double r4[5] = {0., 0., 0., 1., RoeVel[2]};
  const double r5[5] = {1., a[4], RoeVel[1], RoeVel[2], RoeH + RoeVel[0] * Roec};
  double al[5];
  if (which == OR_RIEMANN_ROE) {
    double dU[6];
    for (int v = 0; v < 5; ++v) dU[v] = R[v] - L[v];
    dU[5] = dU[4] - (dU[2] - RoeVel[1] * dU[0]) * RoeVel[1] - (dU[3] - RoeVel[2] * dU[0]) * RoeVel[2];
    const double A3 = dU[2] - RoeVel[1] * dU[0];
    const double A4 = dU[3] - RoeVel[2] * dU[0];
    // ALPHA2_RIEMANN_H, eos.h
    const double A2 = e.kappaM1 / (Roec * Roec) * (dU[0] * (RoeH - RoeVel[0] * RoeVel[0]) - dU[5] + RoeVel[0] * dU[1]);
    const double A1 = 0.5 / Roec * (dU[0] * (RoeVel[0] + Roec) - dU[1] - Roec * A2);
    const double A5 = dU[0] - A1 - A2;
    al[0] = A1; al[1] = A2; al[2] = A3; al[3] = A4; al[4] = A5;
    for (int k = 0; k < 5; ++k) a[k] = std::fabs(a[k]);
  } else {
    const double cL = std::sqrt(e.kappa * L[E_PRES] * L[E_SRHO]);
    const double cR = std::sqrt(e.kappa * R[E_PRES] * R[E_SRHO]);
    const double RoeDens = std::sqrt(L[E_DENS] * R[E_DENS]);
    const double d1 = R[E_DENS] - L[E_DENS];
    const double dv[3] = {R[E_VEL1] - L[E_VEL1], R[E_VEL2] - L[E_VEL2], R[E_VEL3] - L[E_VEL3]};
    const double d5 = R[E_PRES] - L[E_PRES];
    const double tmp = 0.5 / (Roec * Roec);
    al[0] = tmp * (d5 - RoeDens * Roec * dv[0]);
    al[1] = d1 - d5 * 2. * tmp;
    al[2] = RoeDens * dv[1];
    al[3] = RoeDens * dv[2];
    al[4] = tmp * (d5 + RoeDens * Roec * dv[0]);
    const double aL[5] = {L[E_VEL1] - cL, L[E_VEL1], L[E_VEL1], L[E_VEL1], L[E_VEL1] + cL};
    const double aR[5] = {R[E_VEL1] - cR, R[E_VEL1], R[E_VEL1], R[E_VEL1], R[E_VEL1] + cR};
    for (int k = 0; k < 5; ++k) {
      const double da = std::fmax(std::fmax(0.0, a[k] - aL[k]), aR[k] - a[k]);
      if (std::fabs(a[k]) < da)
        a[k] = 0.5 * (a[k] * a[k] / da + da);
      else
        a[k] = std::fabs(a[k]);
    }
  }
  if (!split) {
    for (int v = 0; v < 5; ++v)
      F[v] = 0.5 * ((F_L[v] + F_R[v]) - al[0] * a[0] * r1[v] - al[1] * a[1] * r2[v] - al[2] * a[2] * r3[v] -
                    al[3] * a[3] * r4[v] - al[4] * a[4] * r5[v]);
  } else {
    split_surface_flux_pi(L, R, F);
    for (int v = 0; v < 5; ++v)
      F[v] = F[v] - 0.5 * (al[0] * a[0] * r1[v] + al[1] * a[1] * r2[v] + al[2] * a[2] * r3[v] +
                           al[3] * a[3] * r4[v] + al[4] * a[4] * r5[v]);
  }
}

// Riemann_Point, riemann.f90:362-440: rotate into (n,t1,t2), solve, rotate back.
void riemann_point(const EosC& e, int which, int split, const double* U_L, const double* U_R,
                   const double* P_L, const double* P_R, const double* nv, const double* t1,
                   const double* t2, double* Fout) {
  double L[11], R[11], F_L[5], F_R[5], F[5];
  auto fill = [&](double* X, const double* U, const double* P) {
    X[E_DENS] = U[DENS];
    X[E_SRHO] = 1.0 / X[E_DENS];
    X[E_ENER] = U[ENER];
    X[E_PRES] = P[PRES];
    X[E_VEL1] = P[VEL1] * nv[0] + P[VEL2] * nv[1] + P[VEL3] * nv[2];
    X[E_VEL2] = P[VEL1] * t1[0] + P[VEL2] * t1[1] + P[VEL3] * t1[2];
    X[E_MOM1] = X[E_DENS] * X[E_VEL1];
    X[E_MOM1 + 1] = X[E_DENS] * X[E_VEL2];
    X[E_VEL3] = P[VEL1] * t2[0] + P[VEL2] * t2[1] + P[VEL3] * t2[2];
    X[E_MOM1 + 2] = X[E_DENS] * X[E_VEL3];
    X[10] = 0.0;
  };
  fill(L, U_L, P_L);
  fill(R, U_R, P_R);
  if (!split) {
    euler_flux_1d_fast(L, F_L);
    euler_flux_1d_fast(R, F_R);
  }
  riemann_solver(e, which, split, F_L, F_R, L, R, F);
  Fout[DENS] = F[DENS];
  for (int c = 0; c < 3; ++c) Fout[MOM1 + c] = nv[c] * F[MOM1] + t1[c] * F[MOM2] + t2[c] * F[MOM3];
  Fout[ENER] = F[ENER];
}

}  // namespace

struct or_ctx {
  or_params p;
  EosC e;
  int N, n, nf, np, nElems, nSides;
  std::vector<int> S2E, E2S, S2V2;
  std::vector<double> Mf, Mg, Mh, sJ, NormVec, TangVec1, TangVec2, SurfElem;
  std::vector<double> xGP, wGP, wBary, L_Minus, L_Plus, L_HatMinus, L_HatPlus, D, D_T, D_Hat, D_Hat_T, DVolSurf;
  // work arrays
  std::vector<double> UPrim, U_master, U_slave, UPrim_master, UPrim_slave, Flux_master, Flux_slave;
  std::vector<double> gx, gy, gz, gx_m, gy_m, gz_m, gx_s, gy_s, gz_s, Ut_tmp_rk;
  std::vector<double> MetricsAdv, MetricsVisc;
  // boundary sides 1..nBCSides (mesh_vars.f90: BC, BoundaryType; equation_vars.f90: RefStatePrim)
  int nBCSides = 0, nBCs = 0, nRefState = 0;
  std::vector<int> BC, BoundaryType;  // BC(nBCSides), BoundaryType(nBCs,3): type, state, alpha
  std::vector<double> RefStatePrim;   // (6,nRefState)
  std::vector<double> BCDataPrim;     // (6,0:N,0:N,nBCSides) for types 12/121, or empty
};

namespace {

inline int vidx(const or_ctx* c, int i, int j, int k) { return i + c->n * (j + c->n * k); }

// PRESSURE_RIEMANN, idealgas/eos.f90:561-584: pressure of the local Riemann problem against a
// wall, P in the side frame (VEL1 = normal velocity).
double pressure_riemann(const EosC& e, const double* P) {
  const double kappaFac = 2. * e.kappa * e.sKappaM1;
  if (P[VEL1] <= 0.) {  // rarefaction
    return P[PRES] * std::pow(std::fmax(0.0001, 1. + 0.5 * e.kappaM1 * P[VEL1] / std::sqrt(e.kappa * P[PRES] / P[DENS])),
                              kappaFac);
  }
  const double sKappaP1 = 1. / (e.kappa + 1.);  // shock
  const double ar = 2. * sKappaP1 / P[DENS];
  const double br = e.kappaM1 * sKappaP1 * P[PRES];
  return P[PRES] + P[VEL1] / ar * 0.5 * (P[VEL1] + std::sqrt(P[VEL1] * P[VEL1] + 4. * ar * (P[PRES] + br)));
}

inline bool riemann_type_bc(int type) {
  return type == 2 || type == 12 || type == 121 || type == 23 || type == 24 || type == 25 || type == 27;
}

// GetBoundaryState, idealgas/getboundaryflux.f90:224-472: types 2 (RefState, BCState >= 1), 12/121
// (BCDataPrim of the point, data), 3 (adiabatic wall), 4 (isothermal wall), 9/91 (slip wall), 23
// (outflow Mach number), 24 (pressure outflow), 25 (subsonic outflow), 27 (subsonic inflow); ref =
// RefStatePrim(:,BCState), as InitBC leaves it (type 27: Tt, direction a(1:3), pt).
void boundary_state(const EosC& e, int type, const double* ref, const double* data, const double* Pm, const double* nv,
                    const double* t1, const double* t2, double* Pb) {
  if (type == 2) {
    for (int v = 0; v < 6; ++v) Pb[v] = ref[v];
    return;
  }
  if (type == 12 || type == 121) {
    for (int v = 0; v < 6; ++v) Pb[v] = data[v];
    return;
  }
  // transform the inner state into the normal system (:302-312)
  Pb[DENS] = Pm[DENS];
  Pb[VEL1] = Pm[VEL1] * nv[0] + Pm[VEL2] * nv[1] + Pm[VEL3] * nv[2];
  Pb[VEL2] = Pm[VEL1] * t1[0] + Pm[VEL2] * t1[1] + Pm[VEL3] * t1[2];
  Pb[VEL3] = Pm[VEL1] * t2[0] + Pm[VEL2] * t2[1] + Pm[VEL3] * t2[2];
  Pb[PRES] = Pm[PRES];
  Pb[TEMP] = Pm[TEMP];
  const double kappa = e.kappa, kappaM1 = e.kappaM1, sKappaM1 = e.sKappaM1, R = e.R;
  if (type == 3 || type == 4) {  // :314-336
    Pb[PRES] = pressure_riemann(e, Pb);
    Pb[VEL1] = Pb[VEL2] = Pb[VEL3] = 0.;
    Pb[TEMP] = type == 3 ? Pm[TEMP] : ref[TEMP];
    Pb[DENS] = Pb[PRES] / (Pb[TEMP] * R);
  } else if (type == 9 || type == 91) {  // :338-350
    Pb[PRES] = pressure_riemann(e, Pb);
    Pb[VEL1] = 0.;
    Pb[DENS] = Pm[DENS];
    Pb[TEMP] = Pb[PRES] / (Pb[DENS] * R);
  } else if (type == 23) {  // :358-383
    const double MaOut = ref[1];
    const double c = std::sqrt(kappa * Pb[PRES] / Pb[DENS]);
    const double Ma = Pb[VEL1] / c;
    double pb;
    if (Ma < 1) {
      const double pt = Pb[PRES] * std::pow(1 + 0.5 * (kappa - 1) * Ma * Ma, kappa * sKappaM1);
      pb = pt * std::pow(1 + 0.5 * (kappa - 1) * MaOut * MaOut, -kappa * sKappaM1);
    } else {
      pb = Pb[PRES] + 0.5 * Pb[DENS] * (Pb[VEL1] * Pb[VEL1] + Pb[VEL2] * Pb[VEL2] + Pb[VEL3] * Pb[VEL3]);
    }
    Pb[DENS] = kappa * pb / (c * c);
    Pb[PRES] = pb;
    Pb[TEMP] = Pb[PRES] / (R * Pb[DENS]);
  } else if (type == 24) {  // :385-403
    const double c = std::sqrt(kappa * Pb[PRES] / Pb[DENS]);
    const double Ma = Pb[VEL1] / c;
    if (Ma < 1) {
      const double pb = ref[PRES];
      Pb[DENS] = kappa * pb / (c * c);
      Pb[PRES] = pb;
      Pb[TEMP] = Pb[PRES] / (R * Pb[DENS]);
    }
  } else if (type == 25) {  // :405-431
    const double c = std::sqrt(kappa * Pb[PRES] / Pb[DENS]);
    const double Ma = Pb[VEL1] / c;
    double pb;
    if (Ma < 1) pb = ref[PRES];
    else pb = Pb[PRES] + 0.5 * Pb[DENS] * (Pb[VEL1] * Pb[VEL1] + Pb[VEL2] * Pb[VEL2] + Pb[VEL3] * Pb[VEL3]);
    if (Pb[VEL1] < 0.) {
      Pb[VEL1] = std::fabs(Pb[VEL1]);
      Pb[VEL2] = 0.;
      Pb[VEL3] = 0.;
    }
    Pb[DENS] = kappa * pb / (c * c);
    Pb[PRES] = ref[PRES];
    Pb[TEMP] = Pb[PRES] / (R * Pb[DENS]);
  } else {  // 27, :433-472
    const double Tt = ref[0], pt = ref[4];
    const double* a = ref + 1;
    const double A = -1. * (a[0] * nv[0] + a[1] * nv[1] + a[2] * nv[2]);
    const double c = std::sqrt(kappa * Pb[PRES] / Pb[DENS]);
    const double Rplus = -Pb[VEL1] - 2. * c / kappaM1;
    const double tmp1 = A * A + 2. / kappaM1;
    const double tmp2 = 2 * Rplus;
    const double tmp3 = kappaM1 / 2. * Rplus * Rplus - kappa * R * Tt * A * A;
    const double cb = std::fmax((-tmp2 + std::sqrt(tmp2 * tmp2 - 4 * tmp1 * tmp3)) / (2 * tmp1),
                                (-tmp2 - std::sqrt(tmp2 * tmp2 - 4 * tmp1 * tmp3)) / (2 * tmp1));
    const double Tb = cb * cb / (kappa * R);
    const double Ma = std::sqrt(2. / kappaM1 * (Tt / Tb - 1.));
    const double pb = pt * std::pow(1. + 0.5 * kappaM1 * Ma * Ma, -kappa / kappaM1);
    const double U = Ma * std::sqrt(kappa * R * Tb);
    Pb[DENS] = pb / (R * Tb);
    Pb[VEL1] = U * (a[0] * nv[0] + a[1] * nv[1] + a[2] * nv[2]);
    Pb[VEL2] = U * (a[0] * t1[0] + a[1] * t1[1] + a[2] * t1[2]);
    Pb[VEL3] = U * (a[0] * t2[0] + a[1] * t2[1] + a[2] * t2[2]);
    Pb[PRES] = pb;
    Pb[TEMP] = Tb;
  }
  // rotate the state back to the physical system (:476-481)
  const double va = Pb[VEL1], vb = Pb[VEL2], vc = Pb[VEL3];
  for (int d = 0; d < 3; ++d) Pb[VEL1 + d] = va * nv[d] + vb * t1[d] + vc * t2[d];
}

// Lifting_GetBoundaryFlux, idealgas/getboundaryflux.f90:855-939 (strong form, lifting variables
// u,v,w,T), before the SurfElem factor: the boundary value minus the inner one.
void lifting_boundary_flux(int type, const double* Pm, const double* Pb, double* F) {
  const int PL[4] = {VEL1, VEL2, VEL3, TEMP};
  for (int v = 0; v < 4; ++v) {
    double f;
    if (riemann_type_bc(type)) f = 0.5 * (Pm[PL[v]] + Pb[PL[v]]);
    else if (type == 3 || type == 4) f = v < 3 ? 0. : Pb[TEMP];
    else f = v < 3 ? Pb[PL[v]] : Pm[TEMP];  // 9, 91
    F[v] = f - Pm[PL[v]];
  }
}

// The gradients the viscous flux of a slip wall sees, getboundaryflux.f90:622-776 (lifting
// variables u,v,w,T in g*[0..3]): 9 projects every gradient on the wall, (I - n n^T) grad; 91 does
// so for T, and for the velocity removes in the normal/tangential frame the normal derivatives of
// the tangential components and the tangential derivatives of the normal component.
void slip_gradients(int type, const double* nv, const double* tv1, const double* tv2, const double* gx,
                    const double* gy, const double* gz, double* qx, double* qy, double* qz) {
  double M[3][3];
  M[0][0] = 1. - nv[0] * nv[0];
  M[1][1] = 1. - nv[1] * nv[1];
  M[2][2] = 1. - nv[2] * nv[2];
  M[0][1] = -nv[0] * nv[1];
  M[0][2] = -nv[0] * nv[2];
  M[2][1] = -nv[2] * nv[1];
  M[1][0] = M[0][1];
  M[2][0] = M[0][2];
  M[1][2] = M[2][1];
  if (type == 9) {
    for (int v = 0; v < 4; ++v) {
      qx[v] = M[0][0] * gx[v] + M[0][1] * gy[v] + M[0][2] * gz[v];
      qy[v] = M[1][0] * gx[v] + M[1][1] * gy[v] + M[1][2] * gz[v];
      qz[v] = M[2][0] * gx[v] + M[2][1] * gy[v] + M[2][2] * gz[v];
    }
    return;
  }
  const int T = 3;
  qx[T] = M[0][0] * gx[T] + M[0][1] * gy[T] + M[0][2] * gz[T];
  qy[T] = M[1][0] * gx[T] + M[1][1] * gy[T] + M[1][2] * gz[T];
  qz[T] = M[2][0] * gx[T] + M[2][1] * gy[T] + M[2][2] * gz[T];
  auto dot = [](const double* a, const double* g) { return a[0] * g[0] + a[1] * g[1] + a[2] * g[2]; };
  double gx_vN = dot(nv, gx), gx_vT1 = dot(tv1, gx), gx_vT2 = dot(tv2, gx);
  double gy_vN = dot(nv, gy), gy_vT1 = dot(tv1, gy), gy_vT2 = dot(tv2, gy);
  double gz_vN = dot(nv, gz), gz_vT1 = dot(tv1, gz), gz_vT2 = dot(tv2, gz);
  const double gn_vN = nv[0] * gx_vN + nv[1] * gy_vN + nv[2] * gz_vN;
  const double gn_vT1 = 0., gn_vT2 = 0., gt1_vN = 0., gt2_vN = 0.;
  const double gt1_vT1 = tv1[0] * gx_vT1 + tv1[1] * gy_vT1 + tv1[2] * gz_vT1;
  const double gt1_vT2 = tv1[0] * gx_vT2 + tv1[1] * gy_vT2 + tv1[2] * gz_vT2;
  const double gt2_vT1 = tv2[0] * gx_vT1 + tv2[1] * gy_vT1 + tv2[2] * gz_vT1;
  const double gt2_vT2 = tv2[0] * gx_vT2 + tv2[1] * gy_vT2 + tv2[2] * gz_vT2;
  gx_vN = nv[0] * gn_vN + tv1[0] * gt1_vN + tv2[0] * gt2_vN;
  gx_vT1 = nv[0] * gn_vT1 + tv1[0] * gt1_vT1 + tv2[0] * gt2_vT1;
  gx_vT2 = nv[0] * gn_vT2 + tv1[0] * gt1_vT2 + tv2[0] * gt2_vT2;
  gy_vN = nv[1] * gn_vN + tv1[1] * gt1_vN + tv2[1] * gt2_vN;
  gy_vT1 = nv[1] * gn_vT1 + tv1[1] * gt1_vT1 + tv2[1] * gt2_vT1;
  gy_vT2 = nv[1] * gn_vT2 + tv1[1] * gt1_vT2 + tv2[1] * gt2_vT2;
  gz_vN = nv[2] * gn_vN + tv1[2] * gt1_vN + tv2[2] * gt2_vN;
  gz_vT1 = nv[2] * gn_vT1 + tv1[2] * gt1_vT1 + tv2[2] * gt2_vT1;
  gz_vT2 = nv[2] * gn_vT2 + tv1[2] * gt1_vT2 + tv2[2] * gt2_vT2;
  for (int a = 0; a < 3; ++a) {
    qx[a] = nv[a] * gx_vN + tv1[a] * gx_vT1 + tv2[a] * gx_vT2;
    qy[a] = nv[a] * gy_vN + tv1[a] * gy_vT1 + tv2[a] * gy_vT2;
    qz[a] = nv[a] * gz_vN + tv1[a] * gz_vT1 + tv2[a] * gz_vT2;
  }
}

// GetBoundaryFlux, idealgas/getboundaryflux.f90:480-789, before the SurfElem factor.
void boundary_flux(const or_ctx* c, int type, const double* Pm, const double* Pb, const double* gx,
                   const double* gy, const double* gz, const double* nv, const double* t1, const double* t2,
                   double* F) {
  const EosC& e = c->e;
  if (riemann_type_bc(type)) {  // Riemann-type: Riemann solver against the boundary state + ViscousFlux
    double Um[5], Ub[5];
    prim_to_cons(e, Pm, Um);
    prim_to_cons(e, Pb, Ub);
    riemann_point(e, c->p.riemann, c->p.splitDG, Um, Ub, Pm, Pb, nv, t1, t2, F);
    double fL[5], gL[5], hL[5], fR[5], gR[5], hR[5];
    eval_diff_flux(e, Pm, gx, gy, gz, fL, gL, hL);
    eval_diff_flux(e, Pb, gx, gy, gz, fR, gR, hR);
    for (int v = 0; v < 5; ++v)
      F[v] = F[v] + 0.5 * (nv[0] * (fL[v] + fR[v]) + nv[1] * (gL[v] + gR[v]) + nv[2] * (hL[v] + hR[v]));
    return;
  }
  // walls: Euler flux with zero normal velocity, F = (0, p n, 0)
  F[DENS] = 0.;
  for (int d = 0; d < 3; ++d) F[MOM1 + d] = Pb[PRES] * nv[d];
  F[ENER] = 0.;
  double fd[5], gd[5], hd[5];
  if (type == 3 || type == 4) {
    eval_diff_flux(e, Pb, gx, gy, gz, fd, gd, hd);
    if (type == 3) fd[ENER] = gd[ENER] = hd[ENER] = 0.;
  } else {  // 9, 91
    double qx[4], qy[4], qz[4];
    slip_gradients(type, nv, t1, t2, gx, gy, gz, qx, qy, qz);
    eval_diff_flux(e, Pb, qx, qy, qz, fd, gd, hd);
  }
  for (int v = 1; v < 5; ++v) F[v] = F[v] + nv[0] * fd[v] + nv[1] * gd[v] + nv[2] * hd[v];
}

const double* bc_ref(const or_ctx* c, int s) {
  const int iBC = c->BC[s] - 1;
  const int type = c->BoundaryType[iBC], state = c->BoundaryType[iBC + c->nBCs];
  if (type == 12 || type == 121 || type == 3 || type == 9 || type == 91) return nullptr;
  return state >= 1 ? &c->RefStatePrim[6 * (state - 1)] : nullptr;
}
// BCDataPrim(:,p,q,SideID) of boundary side point fi (types 12/121), or null
const double* bc_data(const or_ctx* c, size_t fi) {
  return c->BCDataPrim.empty() ? nullptr : &c->BCDataPrim[6 * fi];
}
inline int bc_type(const or_ctx* c, int s) { return c->BoundaryType[c->BC[s] - 1]; }

// ProlongToFace, interpolation/prolongtoface.t90:21-144 (EvalElemFaceG / EvalElemFaceGL).
void prolong_to_face(const or_ctx* c, int nVar, const double* Uvol, double* Um, double* Us) {
  const int N = c->N, n = c->n, nf = c->nf, np = c->np;
#pragma omp parallel for schedule(static)
  for (int s = 0; s < c->nSides; ++s) {
    std::vector<double> Uface(nVar * nf);
    for (int side = 0; side < 2; ++side) {
      const int elem = c->S2E[(side == 0 ? 0 : 1) + 5 * s];
      if (elem <= 0) continue;
      const int ls = c->S2E[(side == 0 ? 2 : 3) + 5 * s];
      const int flip = side == 0 ? 0 : c->S2E[4 + 5 * s];
      const double* U = Uvol + (size_t)nVar * np * (elem - 1);
      for (int b = 0; b < n; ++b)
        for (int a = 0; a < n; ++a) {
          // face-local (a,b) -> volume index of the face plane and the normal line
          for (int v = 0; v < nVar; ++v) {
            double val;
            auto at = [&](int l) {
              switch (ls) {
                case 5: case 3: return U[v + nVar * vidx(c, l, a, b)];
                case 2: case 4: return U[v + nVar * vidx(c, a, l, b)];
                default: return U[v + nVar * vidx(c, a, b, l)];
              }
            };
            if (c->p.nodeType == OR_NODE_G) {
              const double* Lv = (ls == 5 || ls == 2 || ls == 1) ? c->L_Minus.data() : c->L_Plus.data();
              val = at(0) * Lv[0];
              for (int l = 1; l <= N; ++l) val = val + at(l) * Lv[l];
            } else {
              val = (ls == 5 || ls == 2 || ls == 1) ? at(0) : at(N);
            }
            Uface[v + nVar * (a + n * b)] = val;
          }
        }
      double* out = (side == 0 ? Um : Us) + (size_t)nVar * nf * s;
      for (int q = 0; q < n; ++q)
        for (int p = 0; p < n; ++p) {
          const int* ij = &c->S2V2[2 * (p + n * (q + n * (flip + 5 * (ls - 1))))];
          for (int v = 0; v < nVar; ++v) out[v + nVar * (p + n * q)] = Uface[v + nVar * (ij[0] + n * ij[1])];
        }
    }
  }
}

// DoSurfInt, dg/surfint.t90:683-766. F is the side flux already in volume (a,b) frame.
void do_surf_int(const or_ctx* c, int nVar, const double* F, int ls, double* Ut) {
  const int N = c->N, n = c->n;
  const bool minus = (ls == 5 || ls == 2 || ls == 1);
  if (c->p.nodeType == OR_NODE_G) {
    const double* Lh = minus ? c->L_HatMinus.data() : c->L_HatPlus.data();
    for (int b = 0; b < n; ++b)
      for (int a = 0; a < n; ++a)
        for (int l = 0; l < n; ++l) {
          int id;
          switch (ls) {
            case 5: case 3: id = vidx(c, l, a, b); break;
            case 2: case 4: id = vidx(c, a, l, b); break;
            default: id = vidx(c, a, b, l);
          }
          for (int v = 0; v < nVar; ++v) Ut[v + nVar * id] = Ut[v + nVar * id] + F[v + nVar * (a + n * b)] * Lh[l];
        }
  } else {
    const double Lh = minus ? c->L_HatMinus[0] : c->L_HatPlus[N];
    const int l = minus ? 0 : N;
    for (int b = 0; b < n; ++b)
      for (int a = 0; a < n; ++a) {
        int id;
        switch (ls) {
          case 5: case 3: id = vidx(c, l, a, b); break;
          case 2: case 4: id = vidx(c, a, l, b); break;
          default: id = vidx(c, a, b, l);
        }
        for (int v = 0; v < nVar; ++v) Ut[v + nVar * id] = Ut[v + nVar * id] + F[v + nVar * (a + n * b)] * Lh;
      }
  }
}

// SurfInt (Gauss: surfint.t90:21-239 weak branch; GL: :471-616), also used by the lifting with
// slave_sign = +1 (Lifting_SurfInt_BR1, lifting_br1.t90:195-292, strong form).
void surf_int(const or_ctx* c, int nVar, const double* Flux_m, const double* Flux_s, double slave_sign, double* Ut) {
  const int n = c->n, nf = c->nf, np = c->np;
  std::vector<double> tmp(nVar * nf);
  for (int s = 0; s < c->nSides; ++s) {
    const int e = c->S2E[0 + 5 * s], nb = c->S2E[1 + 5 * s];
    if (e > 0) {
      const int ls = c->S2E[2 + 5 * s];
      for (int q = 0; q < n; ++q)
        for (int p = 0; p < n; ++p) {
          const int* ij = &c->S2V2[2 * (p + n * (q + n * (0 + 5 * (ls - 1))))];
          for (int v = 0; v < nVar; ++v) tmp[v + nVar * (ij[0] + n * ij[1])] = Flux_m[v + nVar * (p + n * (q + n * s))];
        }
      do_surf_int(c, nVar, tmp.data(), ls, Ut + (size_t)nVar * np * (e - 1));
    }
    if (nb > 0) {
      const int ls = c->S2E[3 + 5 * s], flip = c->S2E[4 + 5 * s];
      for (int q = 0; q < n; ++q)
        for (int p = 0; p < n; ++p) {
          const int* ij = &c->S2V2[2 * (p + n * (q + n * (flip + 5 * (ls - 1))))];
          for (int v = 0; v < nVar; ++v)
            tmp[v + nVar * (ij[0] + n * ij[1])] =
                slave_sign < 0 ? -Flux_s[v + nVar * (p + n * (q + n * s))] : Flux_s[v + nVar * (p + n * (q + n * s))];
        }
      do_surf_int(c, nVar, tmp.data(), ls, Ut + (size_t)nVar * np * (nb - 1));
    }
  }
}

// Lifting_BR1, dg/lifting/lifting_br1.t90:48-181, strong form, non-conservative volume integral.
void lifting_br1(or_ctx* c) {
  const int N = c->N, n = c->n, nf = c->nf, np = c->np, K = c->nElems;
  const int PL[4] = {VEL1, VEL2, VEL3, TEMP};  // PRIM_LIFT with the rho gradient dropped
  // Lifting_VolInt_Nonconservative, lifting_volint.t90:26-118
#pragma omp parallel for schedule(static)
  for (int e = 0; e < K; ++e) {
    const double* P = &c->UPrim[(size_t)6 * np * e];
    for (int k = 0; k < n; ++k)
      for (int j = 0; j < n; ++j)
        for (int i = 0; i < n; ++i) {
          double gxi[4], geta[4], gzeta[4];
          for (int v = 0; v < 4; ++v) {
            gxi[v] = c->D_T[0 + n * i] * P[PL[v] + 6 * vidx(c, 0, j, k)];
            geta[v] = c->D_T[0 + n * j] * P[PL[v] + 6 * vidx(c, i, 0, k)];
            gzeta[v] = c->D_T[0 + n * k] * P[PL[v] + 6 * vidx(c, i, j, 0)];
          }
          for (int l = 1; l <= N; ++l)
            for (int v = 0; v < 4; ++v) {
              gxi[v] = gxi[v] + c->D_T[l + n * i] * P[PL[v] + 6 * vidx(c, l, j, k)];
              geta[v] = geta[v] + c->D_T[l + n * j] * P[PL[v] + 6 * vidx(c, i, l, k)];
              gzeta[v] = gzeta[v] + c->D_T[l + n * k] * P[PL[v] + 6 * vidx(c, i, j, l)];
            }
          const size_t id = (size_t)np * e + vidx(c, i, j, k);
          const double* mf = &c->Mf[3 * id];
          const double* mg = &c->Mg[3 * id];
          const double* mh = &c->Mh[3 * id];
          for (int v = 0; v < 4; ++v) {
            c->gx[v + 4 * id] = mf[0] * gxi[v] + mg[0] * geta[v] + mh[0] * gzeta[v];
            c->gy[v + 4 * id] = mf[1] * gxi[v] + mg[1] * geta[v] + mh[1] * gzeta[v];
            c->gz[v + 4 * id] = mf[2] * gxi[v] + mg[2] * geta[v] + mh[2] * gzeta[v];
          }
        }
  }
  // Lifting_FillFlux (dir=0) and Lifting_FillFlux_NormVec, lifting_fillflux.t90:40-117,189-239
  // Lifting_FillFlux_BC (lifting_fillflux.t90:133-180) for sides 1..nBCSides
  for (int s = 0; s < c->nBCSides; ++s) {
    const int type = bc_type(c, s);
    const double* ref = bc_ref(c, s);
    for (int pq = 0; pq < nf; ++pq) {
      const size_t fi = (size_t)nf * s + pq;
      double Pb[6], F[4];
      boundary_state(c->e, type, ref, bc_data(c, fi), &c->UPrim_master[6 * fi], &c->NormVec[3 * fi], &c->TangVec1[3 * fi],
                     &c->TangVec2[3 * fi], Pb);
      lifting_boundary_flux(type, &c->UPrim_master[6 * fi], Pb, F);
      for (int v = 0; v < 4; ++v) {
        const double f = F[v] * c->SurfElem[fi];
        c->gx_m[v + 4 * fi] = f * c->NormVec[0 + 3 * fi];
        c->gy_m[v + 4 * fi] = f * c->NormVec[1 + 3 * fi];
        c->gz_m[v + 4 * fi] = f * c->NormVec[2 + 3 * fi];
      }
    }
  }
  for (int s = c->nBCSides; s < c->nSides; ++s) {
    for (int pq = 0; pq < nf; ++pq) {
      const size_t fi = (size_t)nf * s + pq;
      for (int v = 0; v < 4; ++v) {
        double f = -1 * c->UPrim_master[PL[v] + 6 * fi] + c->UPrim_slave[PL[v] + 6 * fi];
        f = 0.5 * f * c->SurfElem[fi];
        c->gx_m[v + 4 * fi] = f * c->NormVec[0 + 3 * fi];
        c->gy_m[v + 4 * fi] = f * c->NormVec[1 + 3 * fi];
        c->gz_m[v + 4 * fi] = f * c->NormVec[2 + 3 * fi];
      }
    }
  }
  // Lifting_SurfInt_BR1 with the same flux on master and slave (strong form)
  surf_int(c, 4, c->gx_m.data(), c->gx_m.data(), +1.0, c->gx.data());
  surf_int(c, 4, c->gy_m.data(), c->gy_m.data(), +1.0, c->gy.data());
  surf_int(c, 4, c->gz_m.data(), c->gz_m.data(), +1.0, c->gz.data());
  // ApplyJacobianLifting
#pragma omp parallel for schedule(static)
  for (int e = 0; e < K; ++e)
    for (int id = 0; id < np; ++id) {
      const size_t g = (size_t)np * e + id;
      for (int v = 0; v < 4; ++v) {
        c->gx[v + 4 * g] = c->gx[v + 4 * g] * c->sJ[g];
        c->gy[v + 4 * g] = c->gy[v + 4 * g] * c->sJ[g];
        c->gz[v + 4 * g] = c->gz[v + 4 * g] * c->sJ[g];
      }
    }
  // ProlongToFaceLifting
  prolong_to_face(c, 4, c->gx.data(), c->gx_m.data(), c->gx_s.data());
  prolong_to_face(c, 4, c->gy.data(), c->gy_m.data(), c->gy_s.data());
  prolong_to_face(c, 4, c->gz.data(), c->gz_m.data(), c->gz_s.data());
}

// VolInt_weakForm (volint.f90:117-187) and VolInt_splitForm (:211-371) with VOLINT_VISC=1.
void vol_int(or_ctx* c, const double* Uall, double* Utall) {
  const int N = c->N, n = c->n, np = c->np, K = c->nElems;
  const double* DH = c->D_Hat_T.data();
  const double* DV = c->DVolSurf.data();
#pragma omp parallel for schedule(static)
  for (int e = 0; e < K; ++e) {
    const double* U = Uall + (size_t)5 * np * e;
    const double* P = &c->UPrim[(size_t)6 * np * e];
    const double* mf = &c->Mf[(size_t)3 * np * e];
    const double* mg = &c->Mg[(size_t)3 * np * e];
    const double* mh = &c->Mh[(size_t)3 * np * e];
    double* Ut = Utall + (size_t)5 * np * e;
    std::vector<double> f(5 * np), g(5 * np), h(5 * np), fv(5 * np), gv(5 * np), hv(5 * np);
    for (int id = 0; id < np; ++id) {
      const size_t gi = (size_t)np * e + id;
      eval_diff_flux(c->e, &P[6 * id], &c->gx[4 * gi], &c->gy[4 * gi], &c->gz[4 * gi], &fv[5 * id], &gv[5 * id], &hv[5 * id]);
    }
    auto metrics = [&](double* F, double* G, double* H) {
      // VolInt_Metrics, volint.f90:378-422
      for (int id = 0; id < np; ++id)
        for (int v = 0; v < 5; ++v) {
          const double ft = F[v + 5 * id], gt = G[v + 5 * id], ht = H[v + 5 * id];
          F[v + 5 * id] = ft * mf[0 + 3 * id] + gt * mf[1 + 3 * id] + ht * mf[2 + 3 * id];
          G[v + 5 * id] = ft * mg[0 + 3 * id] + gt * mg[1 + 3 * id] + ht * mg[2 + 3 * id];
          H[v + 5 * id] = ft * mh[0 + 3 * id] + gt * mh[1 + 3 * id] + ht * mh[2 + 3 * id];
        }
    };
    if (!c->p.splitDG) {
      for (int id = 0; id < np; ++id) {
        eval_flux(&U[5 * id], &P[6 * id], &f[5 * id], &g[5 * id], &h[5 * id]);
        for (int v = 0; v < 5; ++v) {
          f[v + 5 * id] = f[v + 5 * id] + fv[v + 5 * id];
          g[v + 5 * id] = g[v + 5 * id] + gv[v + 5 * id];
          h[v + 5 * id] = h[v + 5 * id] + hv[v + 5 * id];
        }
      }
      metrics(f.data(), g.data(), h.data());
      for (int k = 0; k < n; ++k)
        for (int j = 0; j < n; ++j)
          for (int i = 0; i < n; ++i) {
            double* o = &Ut[5 * vidx(c, i, j, k)];
            for (int v = 0; v < 5; ++v)
              o[v] = DH[0 + n * i] * f[v + 5 * vidx(c, 0, j, k)] + DH[0 + n * k] * h[v + 5 * vidx(c, i, j, 0)] +
                     DH[0 + n * j] * g[v + 5 * vidx(c, i, 0, k)];
            for (int l = 1; l <= N; ++l)
              for (int v = 0; v < 5; ++v)
                o[v] = o[v] + DH[l + n * i] * f[v + 5 * vidx(c, l, j, k)] + DH[l + n * k] * h[v + 5 * vidx(c, i, j, l)] +
                       DH[l + n * j] * g[v + 5 * vidx(c, i, l, k)];
          }
    } else {
      metrics(fv.data(), gv.data(), hv.data());
      for (int k = 0; k < n; ++k)
        for (int j = 0; j < n; ++j)
          for (int i = 0; i < n; ++i) {
            const int id = vidx(c, i, j, k);
            for (int v = 0; v < 5; ++v)
              Ut[v + 5 * id] = DH[i + n * i] * fv[v + 5 * id] + DH[j + n * j] * gv[v + 5 * id] + DH[k + n * k] * hv[v + 5 * id];
          }
      double F[5];
      for (int k = 0; k < n; ++k)
        for (int j = 0; j < n; ++j)
          for (int i = 0; i < n; ++i) {
            const int a = vidx(c, i, j, k);
            for (int l = i + 1; l <= N; ++l) {
              const int b = vidx(c, l, j, k);
              split_volume_flux_pi(&U[5 * a], &P[6 * a], &U[5 * b], &P[6 * b], &mf[3 * a], &mf[3 * b], F);
              for (int v = 0; v < 5; ++v) {
                Ut[v + 5 * a] = Ut[v + 5 * a] + DV[l + n * i] * F[v];
                Ut[v + 5 * b] = Ut[v + 5 * b] + DV[i + n * l] * F[v];
              }
              for (int v = 0; v < 5; ++v) {
                Ut[v + 5 * a] = Ut[v + 5 * a] + DH[l + n * i] * fv[v + 5 * b];
                Ut[v + 5 * b] = Ut[v + 5 * b] + DH[i + n * l] * fv[v + 5 * a];
              }
            }
            for (int l = j + 1; l <= N; ++l) {
              const int b = vidx(c, i, l, k);
              split_volume_flux_pi(&U[5 * a], &P[6 * a], &U[5 * b], &P[6 * b], &mg[3 * a], &mg[3 * b], F);
              for (int v = 0; v < 5; ++v) {
                Ut[v + 5 * a] = Ut[v + 5 * a] + DV[l + n * j] * F[v];
                Ut[v + 5 * b] = Ut[v + 5 * b] + DV[j + n * l] * F[v];
              }
              for (int v = 0; v < 5; ++v) {
                Ut[v + 5 * a] = Ut[v + 5 * a] + DH[l + n * j] * gv[v + 5 * b];
                Ut[v + 5 * b] = Ut[v + 5 * b] + DH[j + n * l] * gv[v + 5 * a];
              }
            }
            for (int l = k + 1; l <= N; ++l) {
              const int b = vidx(c, i, j, l);
              split_volume_flux_pi(&U[5 * a], &P[6 * a], &U[5 * b], &P[6 * b], &mh[3 * a], &mh[3 * b], F);
              for (int v = 0; v < 5; ++v) {
                Ut[v + 5 * a] = Ut[v + 5 * a] + DV[l + n * k] * F[v];
                Ut[v + 5 * b] = Ut[v + 5 * b] + DV[k + n * l] * F[v];
              }
              for (int v = 0; v < 5; ++v) {
                Ut[v + 5 * a] = Ut[v + 5 * a] + DH[l + n * k] * hv[v + 5 * b];
                Ut[v + 5 * b] = Ut[v + 5 * b] + DH[k + n * l] * hv[v + 5 * a];
              }
            }
          }
    }
  }
}

// FillFlux, dg/fillflux.f90:40-215 (inner sides: Riemann_Side + ViscousFlux_Side, times SurfElem).
void fill_flux(or_ctx* c) {
  const int nf = c->nf;
  const int split = c->p.splitDG;
  // 2. boundary sides (fillflux.f90:172-190), then 3. times SurfElem
  for (int s = 0; s < c->nBCSides; ++s) {
    const int type = bc_type(c, s);
    const double* ref = bc_ref(c, s);
    for (int pq = 0; pq < nf; ++pq) {
      const size_t fi = (size_t)nf * s + pq;
      const double *Pm = &c->UPrim_master[6 * fi], *nv = &c->NormVec[3 * fi], *t1 = &c->TangVec1[3 * fi],
                   *t2 = &c->TangVec2[3 * fi];
      double Pb[6], F[5];
      boundary_state(c->e, type, ref, bc_data(c, fi), Pm, nv, t1, t2, Pb);
      boundary_flux(c, type, Pm, Pb, &c->gx_m[4 * fi], &c->gy_m[4 * fi], &c->gz_m[4 * fi], nv, t1, t2, F);
      for (int v = 0; v < 5; ++v) c->Flux_master[v + 5 * fi] = F[v] * c->SurfElem[fi];
    }
  }
#pragma omp parallel for schedule(static)
  for (int s = c->nBCSides; s < c->nSides; ++s) {
    for (int pq = 0; pq < nf; ++pq) {
      const size_t fi = (size_t)nf * s + pq;
      double F[5];
      riemann_point(c->e, c->p.riemann, split, &c->U_master[5 * fi], &c->U_slave[5 * fi], &c->UPrim_master[6 * fi],
                    &c->UPrim_slave[6 * fi], &c->NormVec[3 * fi], &c->TangVec1[3 * fi], &c->TangVec2[3 * fi], F);
      // ViscousFlux_Side, riemann.f90:448-498
      double fL[5], gL[5], hL[5], fR[5], gR[5], hR[5];
      eval_diff_flux(c->e, &c->UPrim_master[6 * fi], &c->gx_m[4 * fi], &c->gy_m[4 * fi], &c->gz_m[4 * fi], fL, gL, hL);
      eval_diff_flux(c->e, &c->UPrim_slave[6 * fi], &c->gx_s[4 * fi], &c->gy_s[4 * fi], &c->gz_s[4 * fi], fR, gR, hR);
      const double* nv = &c->NormVec[3 * fi];
      for (int v = 0; v < 5; ++v) {
        const double Fv = 0.5 * (nv[0] * (fL[v] + fR[v]) + nv[1] * (gL[v] + gR[v]) + nv[2] * (hL[v] + hR[v]));
        F[v] = F[v] + Fv;
      }
      for (int v = 0; v < 5; ++v) {
        c->Flux_master[v + 5 * fi] = F[v] * c->SurfElem[fi];
        c->Flux_slave[v + 5 * fi] = c->Flux_master[v + 5 * fi];
      }
    }
  }
}

}  // namespace

extern "C" {

void or_set_threads(int n) { omp_set_num_threads(n); }

void or_cons_to_prim(const or_params* p, const double* cons, double* prim) {
  cons_to_prim(make_eos(p), cons, prim);
}
void or_prim_to_cons(const or_params* p, const double* prim, double* cons) {
  prim_to_cons(make_eos(p), prim, cons);
}
void or_riemann_point(const or_params* p, const double* U_L, const double* U_R, const double* P_L,
                      const double* P_R, const double* nv, const double* t1, const double* t2, double* F) {
  riemann_point(make_eos(p), p->riemann, p->splitDG, U_L, U_R, P_L, P_R, nv, t1, t2, F);
}
void or_split_volume_flux_pi(const double* URef, const double* UPrimRef, const double* U, const double* UPrim,
                             const double* MRef, const double* M, double* F) {
  split_volume_flux_pi(URef, UPrimRef, U, UPrim, MRef, M, F);
}
void or_eval_diff_flux_point(const or_params* p, const double* UPrim, const double* gx, const double* gy,
                             const double* gz, double* f, double* g, double* h) {
  eval_diff_flux(make_eos(p), UPrim, gx, gy, gz, f, g, h);
}
void or_boundary_point(const or_params* p, int type, const double* ref, const double* data, const double* Pm,
                       const double* nv, const double* t1, const double* t2, const double* gx, const double* gy,
                       const double* gz, double* Pb, double* Flift, double* Flux) {
  or_ctx c;
  c.p = *p;
  c.e = make_eos(p);
  boundary_state(c.e, type, ref, data, Pm, nv, t1, t2, Pb);
  lifting_boundary_flux(type, Pm, Pb, Flift);
  boundary_flux(&c, type, Pm, Pb, gx, gy, gz, nv, t1, t2, Flux);
}
void or_slip_gradients(int type, const double* nv, const double* t1, const double* t2, const double* gx,
                       const double* gy, const double* gz, double* qx, double* qy, double* qz) {
  slip_gradients(type, nv, t1, t2, gx, gy, gz, qx, qy, qz);
}

or_ctx* or_init(const or_params* p, int nElems, int nSides, const int* SideToElem, const int* ElemToSide,
                const double* Mf, const double* Mg, const double* Mh, const double* sJ, const double* NormVec,
                const double* TangVec1, const double* TangVec2, const double* SurfElem) {
  or_ctx* c = new or_ctx;
  c->p = *p;
  c->e = make_eos(p);
  c->N = p->N;
  c->n = p->N + 1;
  c->nf = c->n * c->n;
  c->np = c->n * c->n * c->n;
  c->nElems = nElems;
  c->nSides = nSides;
  const size_t np = c->np, nf = c->nf;
  c->S2E.assign(SideToElem, SideToElem + 5 * nSides);
  c->E2S.assign(ElemToSide, ElemToSide + 12 * nElems);
  c->Mf.assign(Mf, Mf + 3 * np * nElems);
  c->Mg.assign(Mg, Mg + 3 * np * nElems);
  c->Mh.assign(Mh, Mh + 3 * np * nElems);
  c->sJ.assign(sJ, sJ + np * nElems);
  c->NormVec.assign(NormVec, NormVec + 3 * nf * nSides);
  c->TangVec1.assign(TangVec1, TangVec1 + 3 * nf * nSides);
  c->TangVec2.assign(TangVec2, TangVec2 + 3 * nf * nSides);
  c->SurfElem.assign(SurfElem, SurfElem + nf * nSides);
  const int n = c->n;
  c->xGP.resize(n); c->wGP.resize(n); c->wBary.resize(n); c->L_Minus.resize(n); c->L_Plus.resize(n);
  c->L_HatMinus.resize(n); c->L_HatPlus.resize(n);
  c->D.resize(n * n); c->D_T.resize(n * n); c->D_Hat.resize(n * n); c->D_Hat_T.resize(n * n); c->DVolSurf.resize(n * n);
  or_dg_basis(c->N, p->nodeType, c->xGP.data(), c->wGP.data(), c->wBary.data(), c->L_Minus.data(), c->L_Plus.data(),
              c->L_HatMinus.data(), c->L_HatPlus.data(), c->D.data(), c->D_T.data(), c->D_Hat.data(),
              c->D_Hat_T.data(), c->DVolSurf.data());
  c->S2V2.resize(2 * nf * 5 * 6);
  or_mappings(c->N, c->S2V2.data(), nullptr, nullptr);
  c->UPrim.resize(6 * np * nElems);
  c->U_master.resize(5 * nf * nSides); c->U_slave.resize(5 * nf * nSides);
  c->UPrim_master.resize(6 * nf * nSides); c->UPrim_slave.resize(6 * nf * nSides);
  c->Flux_master.resize(5 * nf * nSides); c->Flux_slave.resize(5 * nf * nSides);
  for (auto* v : {&c->gx, &c->gy, &c->gz}) v->assign(4 * np * nElems, 0.0);
  for (auto* v : {&c->gx_m, &c->gy_m, &c->gz_m, &c->gx_s, &c->gy_s, &c->gz_s}) v->assign(4 * nf * nSides, 0.0);
  // InitCalctimestep, calctimestep.f90:42-83
  c->MetricsAdv.resize(3 * np * nElems);
  c->MetricsVisc.resize(3 * np * nElems);
  const double KappasPr_max = std::fmax(4. / 3., c->e.KappasPr);
  for (size_t g = 0; g < np * nElems; ++g) {
    const double* M[3] = {&c->Mf[3 * g], &c->Mg[3 * g], &c->Mh[3 * g]};
    for (int d = 0; d < 3; ++d) {
      c->MetricsAdv[d + 3 * g] = c->sJ[g] * std::sqrt(M[d][0] * M[d][0] + M[d][1] * M[d][1] + M[d][2] * M[d][2]);
      double s = 0.0;
      for (int q = 0; q < 3; ++q) { const double a = M[d][q] * c->sJ[g]; s += a * a; }
      c->MetricsVisc[d + 3 * g] = KappasPr_max * s;
    }
  }
  return c;
}

void or_free(or_ctx* c) { delete c; }

int or_set_boundaries(or_ctx* c, int nBCSides, const int* BC, int nBCs, const int* BoundaryType, int nRefState,
                      const double* RefStatePrim) {
  if (nBCSides < 0 || nBCSides > c->nSides) return 1;
  for (int s = 0; s < nBCSides; ++s) {
    if (c->S2E[1 + 5 * s] > 0) return 2;  // a boundary side has no neighbour element
    if (BC[s] < 1 || BC[s] > nBCs) return 3;
    const int type = BoundaryType[BC[s] - 1], state = BoundaryType[BC[s] - 1 + nBCs];
    const bool usesRef = type == 2 || type == 4 || type == 23 || type == 24 || type == 25 || type == 27;
    if (!usesRef && type != 3 && type != 9 && type != 91 && type != 12 && type != 121) return 4;
    if (usesRef && (state < 1 || state > nRefState)) return 5;
  }
  c->nBCSides = nBCSides;
  c->nBCs = nBCs;
  c->nRefState = nRefState;
  c->BC.assign(BC, BC + nBCSides);
  c->BoundaryType.assign(BoundaryType, BoundaryType + 3 * nBCs);
  c->RefStatePrim.assign(RefStatePrim, RefStatePrim + 6 * nRefState);
  return 0;
}

void or_set_boundary_data(or_ctx* c, const double* BCDataPrim) {
  c->BCDataPrim.assign(BCDataPrim, BCDataPrim + 6 * (size_t)c->nf * c->nBCSides);
}

void or_dg_time_derivative(or_ctx* c, double t, const double* U, double* Ut) {
  (void)t;
  const size_t np = c->np, K = c->nElems, nf = c->nf;
  // 2. ConsToPrim on the volume (dg.f90:334)
#pragma omp parallel for schedule(static)
  for (long g = 0; g < (long)(np * K); ++g) cons_to_prim(c->e, &U[5 * g], &c->UPrim[6 * g]);
  // 3. ProlongToFaceCons (dg.f90:374,412)
  prolong_to_face(c, 5, U, c->U_master.data(), c->U_slave.data());
  // 4. GetPrimitiveStateSurface (equation.f90:189-235)
#pragma omp parallel for schedule(static)
  for (long f = 0; f < (long)(nf * c->nSides); ++f) {
    cons_to_prim(c->e, &c->U_master[5 * f], &c->UPrim_master[6 * f]);
    if (f >= (long)nf * c->nBCSides) cons_to_prim(c->e, &c->U_slave[5 * f], &c->UPrim_slave[6 * f]);
  }
  // 6. Lifting
  lifting_br1(c);
  // 8. VolInt
  vol_int(c, U, Ut);
  // 11. FillFlux + SurfIntCons
  fill_flux(c);
  surf_int(c, 5, c->Flux_master.data(), c->Flux_slave.data(), -1.0, Ut);
  // 12. Ut=-Ut ; 14. ApplyJacobianCons
#pragma omp parallel for schedule(static)
  for (long g = 0; g < (long)(np * K); ++g)
    for (int v = 0; v < 5; ++v) Ut[v + 5 * g] = (-Ut[v + 5 * g]) * c->sJ[g];
}

void or_get_gradients(or_ctx* c, double* gx, double* gy, double* gz) {
  const size_t n = 4 * (size_t)c->np * c->nElems;
  std::memcpy(gx, c->gx.data(), n * sizeof(double));
  std::memcpy(gy, c->gy.data(), n * sizeof(double));
  std::memcpy(gz, c->gz.data(), n * sizeof(double));
}

// The lifted density gradient of FLEXI's default build (PP_OPTLIFT=0, LIFT_DENS first of the
// PRIM_LIFT variables, idealgas/eos.h:108-117) for the state of the last right-hand side: the
// steps of Lifting_BR1 above (lifting_br1.t90:48-181; volume lifting_volint.t90:26-118, fluxes
// lifting_fillflux.t90:40-239) for the one variable rho, with Lifting_GetBoundaryFlux's density
// (getboundaryflux.f90:890-921: Riemann-type 0.5*(rho_in + rho_b), no-slip walls the boundary
// state's rho, slip walls the inner rho). gx/gy/gz: (0:N,0:N,0:N,nElems).
void or_get_density_gradient(or_ctx* c, double* gx, double* gy, double* gz) {
  const int N = c->N, n = c->n, nf = c->nf, np = c->np, K = c->nElems;
  const size_t nn = (size_t)np * K;
  std::fill(gx, gx + nn, 0.0);
  std::fill(gy, gy + nn, 0.0);
  std::fill(gz, gz + nn, 0.0);
#pragma omp parallel for schedule(static)
  for (int e = 0; e < K; ++e) {
    const double* P = &c->UPrim[(size_t)6 * np * e];
    for (int k = 0; k < n; ++k)
      for (int j = 0; j < n; ++j)
        for (int i = 0; i < n; ++i) {
          double gxi = c->D_T[0 + n * i] * P[DENS + 6 * vidx(c, 0, j, k)];
          double geta = c->D_T[0 + n * j] * P[DENS + 6 * vidx(c, i, 0, k)];
          double gzeta = c->D_T[0 + n * k] * P[DENS + 6 * vidx(c, i, j, 0)];
          for (int l = 1; l <= N; ++l) {
            gxi = gxi + c->D_T[l + n * i] * P[DENS + 6 * vidx(c, l, j, k)];
            geta = geta + c->D_T[l + n * j] * P[DENS + 6 * vidx(c, i, l, k)];
            gzeta = gzeta + c->D_T[l + n * k] * P[DENS + 6 * vidx(c, i, j, l)];
          }
          const size_t id = (size_t)np * e + vidx(c, i, j, k);
          const double* mf = &c->Mf[3 * id];
          const double* mg = &c->Mg[3 * id];
          const double* mh = &c->Mh[3 * id];
          gx[id] = mf[0] * gxi + mg[0] * geta + mh[0] * gzeta;
          gy[id] = mf[1] * gxi + mg[1] * geta + mh[1] * gzeta;
          gz[id] = mf[2] * gxi + mg[2] * geta + mh[2] * gzeta;
        }
  }
  std::vector<double> fx((size_t)nf * c->nSides), fy(fx.size()), fz(fx.size());
  for (int s = 0; s < c->nSides; ++s)
    for (int pq = 0; pq < nf; ++pq) {
      const size_t fi = (size_t)nf * s + pq;
      double f;
      if (s < c->nBCSides) {
        const int type = bc_type(c, s);
        double Pb[6];
        const double* Pm = &c->UPrim_master[6 * fi];
        boundary_state(c->e, type, bc_ref(c, s), bc_data(c, fi), Pm, &c->NormVec[3 * fi], &c->TangVec1[3 * fi],
                       &c->TangVec2[3 * fi],
                       Pb);
        const double rb = riemann_type_bc(type) ? 0.5 * (Pm[DENS] + Pb[DENS])
                                                : ((type == 9 || type == 91) ? Pm[DENS] : Pb[DENS]);
        f = (rb - Pm[DENS]) * c->SurfElem[fi];
      } else {
        f = -1 * c->UPrim_master[DENS + 6 * fi] + c->UPrim_slave[DENS + 6 * fi];
        f = 0.5 * f * c->SurfElem[fi];
      }
      fx[fi] = f * c->NormVec[0 + 3 * fi];
      fy[fi] = f * c->NormVec[1 + 3 * fi];
      fz[fi] = f * c->NormVec[2 + 3 * fi];
    }
  surf_int(c, 1, fx.data(), fx.data(), +1.0, gx);
  surf_int(c, 1, fy.data(), fy.data(), +1.0, gy);
  surf_int(c, 1, fz.data(), fz.data(), +1.0, gz);
  for (size_t g = 0; g < nn; ++g) {
    gx[g] = gx[g] * c->sJ[g];
    gy[g] = gy[g] * c->sJ[g];
    gz[g] = gz[g] * c->sJ[g];
  }
}

void or_get_faces(or_ctx* c, double* Um, double* Us, double* Fm) {
  const size_t n = 5 * (size_t)c->nf * c->nSides;
  if (Um) std::memcpy(Um, c->U_master.data(), n * sizeof(double));
  if (Us) std::memcpy(Us, c->U_slave.data(), n * sizeof(double));
  if (Fm) std::memcpy(Fm, c->Flux_master.data(), n * sizeof(double));
}

// CALCTIMESTEP, calctimestep.f90:89-227 (FV off, constant viscosity, single rank).
double or_calc_timestep(or_ctx* c, const double* U, double CFLScale, double DFLScale, int* errType, double* dtElem) {
  const int np = c->np;
  double TimeStepConv = 1.79769313486231571e308, TimeStepVisc = 1.79769313486231571e308;  // HUGE(1.)
  int err = 0;
  for (int e = 0; e < c->nElems; ++e) {
    double ML[3] = {0, 0, 0}, MLv[3] = {0, 0, 0};
    for (int id = 0; id < np; ++id) {
      const size_t g = (size_t)np * e + id;
      const double* Uc = &U[5 * g];
      double UE[11];
      for (int v = 0; v < 5; ++v) UE[v] = Uc[v];
      UE[E_SRHO] = 1.0 / UE[E_DENS];
      for (int q = 0; q < 3; ++q) UE[E_VEL1 + q] = UE[E_MOM1 + q] * UE[E_SRHO];
      UE[E_PRES] = c->e.kappaM1 * (UE[E_ENER] - 0.5 * (UE[E_VEL1] * UE[1] + UE[E_VEL2] * UE[2] + UE[E_VEL3] * UE[3]));
      UE[10] = UE[E_PRES] * UE[E_SRHO] / c->e.R;
      if (std::isnan(UE[E_DENS])) err = 1;
      const double cs = std::sqrt(c->e.kappa * UE[E_PRES] * UE[E_SRHO]);
      double vsJ[3];
      for (int q = 0; q < 3; ++q) vsJ[q] = UE[E_VEL1 + q] * c->sJ[g];
      const double* M[3] = {&c->Mf[3 * g], &c->Mg[3 * g], &c->Mh[3 * g]};
      for (int d = 0; d < 3; ++d)
        ML[d] = std::fmax(ML[d], std::fabs(M[d][0] * vsJ[0] + M[d][1] * vsJ[1] + M[d][2] * vsJ[2]) +
                                     cs * c->MetricsAdv[d + 3 * g]);
      const double mu = c->e.mu0;
      for (int d = 0; d < 3; ++d) MLv[d] = std::fmax(MLv[d], mu * UE[E_SRHO] * c->MetricsVisc[d + 3 * g]);
    }
    double dte = CFLScale * 2. / (ML[0] + ML[1] + ML[2]);
    TimeStepConv = std::fmin(TimeStepConv, dte);
    if (std::isnan(TimeStepConv)) err = 2;
    const double sv = MLv[0] + MLv[1] + MLv[2];
    if (sv > 0.0) {
      dte = std::fmin(dte, DFLScale * 4. / sv);
      TimeStepVisc = std::fmin(TimeStepVisc, DFLScale * 4. / sv);
    }
    if (std::isnan(TimeStepVisc)) err = 3;
    if (dtElem) dtElem[e] = dte;
  }
  *errType = err;
  return std::fmin(TimeStepConv, TimeStepVisc);
}

void or_fill_cfl_dfl(int N, int nodeType, double CFL_in, double DFL_in, double* CFL_out, double* DFL_out) {
  // CFLScaleAlpha (Carpenter RK4-5), timedisc_vars.f90:175-190; DFLScaleAlpha :93-98; RelativeDFL :197
  static const double cflG[15] = {2.0351, 1.7595, 1.5401, 1.3702, 1.2375, 1.1318, 1.0440, 0.9709,
                                  0.9079, 0.8539, 0.8066, 0.7650, 0.7290, 0.6952, 0.6660};
  static const double cflGL[15] = {4.7497, 3.4144, 2.8451, 2.4739, 2.2027, 1.9912, 1.8225, 1.6830,
                                   1.5682, 1.4692, 1.3849, 1.3106, 1.2454, 1.1880, 1.1362};
  static const double dflG[10] = {1.12, 0.76, 0.55, 0.41, 0.32, 0.25, 0.20, 0.17, 0.14, 0.12};
  static const double dflGL[10] = {4.50, 1.95, 1.18, 0.79, 0.56, 0.42, 0.32, 0.25, 0.20, 0.17};
  const double RelativeDFL = 1.853;
  double alpha = (nodeType == OR_NODE_G ? cflG : cflGL)[(N < 15 ? N : 15) - 1];
  double cfl = CFL_in * alpha;
  cfl = cfl / (2. * N + 1.);
  alpha = (nodeType == OR_NODE_G ? dflG : dflGL)[(N < 10 ? N : 10) - 1] * RelativeDFL;
  double dfl = DFL_in * alpha;
  dfl = dfl / ((2. * N + 1.) * (2. * N + 1.));
  *CFL_out = cfl;
  *DFL_out = dfl;
}

void or_rk_coefs(double* RKA, double* RKb, double* RKc) {
  // CarpenterRK4-5, timedisc_vars.f90:200-214. RKA/RKc are indexed 2..5 (stored at [1..4]).
  RKA[0] = 0.0;
  RKA[1] = 567301805773.0 / 1357537059087.0;
  RKA[2] = 2404267990393.0 / 2016746695238.0;
  RKA[3] = 3550918686646.0 / 2091501179385.0;
  RKA[4] = 1275806237668.0 / 842570457699.0;
  RKb[0] = 1432997174477.0 / 9575080441755.0;
  RKb[1] = 5161836677717.0 / 13612068292357.0;
  RKb[2] = 1720146321549.0 / 2090206949498.0;
  RKb[3] = 3134564353537.0 / 4481467310338.0;
  RKb[4] = 2277821191437.0 / 14882151754819.0;
  RKc[0] = 0.0;
  RKc[1] = 1432997174477.0 / 9575080441755.0;
  RKc[2] = 2526269341429.0 / 6820363962896.0;
  RKc[3] = 2006345519317.0 / 3224310063776.0;
  RKc[4] = 2802321613138.0 / 2924317926251.0;
}

void or_timestep_lserkw2(or_ctx* c, double t, double dt, double* U, double* Ut_tmp) {
  // TimeStepByLSERKW2, timestep.f90:82-152; VCopy/VAXPBY, globals/vector.f90:74-125
  double RKA[5], RKb[5], RKc[5];
  or_rk_coefs(RKA, RKb, RKc);
  const long n = 5L * c->np * c->nElems;
  std::vector<double> Ut(n);
  for (int s = 0; s < 5; ++s) {
    const double b_dt = RKb[s] * dt;
    const double tStage = (s == 0) ? t : t + RKc[s] * dt;
    or_dg_time_derivative(c, tStage, U, Ut.data());
    if (s == 0) {
#pragma omp parallel for schedule(static)
      for (long i = 0; i < n; ++i) Ut_tmp[i] = Ut[i];
    } else {
      const double a = -RKA[s];
#pragma omp parallel for schedule(static)
      for (long i = 0; i < n; ++i) Ut_tmp[i] = Ut_tmp[i] * a + Ut[i];
    }
#pragma omp parallel for schedule(static)
    for (long i = 0; i < n; ++i) U[i] = U[i] + Ut_tmp[i] * b_dt;
  }
}

// ExactFuncTestcase (TGV), testcase/taylorgreenvortex/testcase.f90:176-219 with rho0=U0=1,
// IniConstDens=T, p0=(U0/Ma)^2/kappa*rho0 (InitTestcase :106).
void or_tgv_ic(const or_params* p, double Ma, int npts, const double* x, double* U) {
  const EosC e = make_eos(p);
  const double rho0 = 1.0, U0 = 1.0;
  const double p0 = (U0 / Ma) * (U0 / Ma) / e.kappa * rho0;
  for (int i = 0; i < npts; ++i) {
    const double* X = &x[3 * i];
    double prim[6];
    prim[VEL1] = U0 * std::sin(X[0]) * std::cos(X[1]) * std::cos(X[2]);
    prim[VEL2] = -U0 * std::cos(X[0]) * std::sin(X[1]) * std::cos(X[2]);
    prim[VEL3] = 0.0;
    prim[PRES] = p0 + (rho0 * U0 * U0) / 16. *
                          (std::cos(2 * X[0]) * std::cos(2. * X[2]) + 2. * std::cos(2. * X[1]) +
                           2. * std::cos(2. * X[0]) + std::cos(2 * X[1]) * std::cos(2. * X[2]));
    prim[DENS] = rho0;
    prim[TEMP] = prim[PRES] / (prim[DENS] * e.R);
    prim_to_cons(e, prim, &U[5 * i]);
  }
}

// AnalyzeTestcase, testcase.f90:241-479, with the analysis basis of analyze.f90:228-270
// (NAnalyze Gauss-Lobatto points, Vdm from the solution nodes) and Vol from analyze.f90:152-160.
void or_tgv_analyze(or_ctx* c, const double* U, const double* gx, const double* gy, const double* gz,
                    int NAnalyze, double* out) {
  const int N = c->N, n = c->n, np = c->np, na = NAnalyze + 1, nap = na * na * na;
  std::vector<double> xA(na), wA(na), Vdm(na * n), L(n);
  nodes(NAnalyze, OR_NODE_GL, xA.data(), wA.data());
  for (int i = 0; i < na; ++i) {
    lagrange_polys(xA[i], N, c->xGP.data(), c->wBary.data(), L.data());
    for (int j = 0; j < n; ++j) Vdm[i + na * j] = L[j];
  }
  double Vol = 0.0;
  for (int e = 0; e < c->nElems; ++e) {
    double ev = 0.0;
    for (int k = 0; k < n; ++k)
      for (int j = 0; j < n; ++j)
        for (int i = 0; i < n; ++i)
          ev = ev + c->wGP[i] * c->wGP[j] * c->wGP[k] / c->sJ[(size_t)np * e + vidx(c, i, j, k)];
    Vol += ev;
  }
  // 3D tensor interpolation in ChangeBasis3D order
  auto cb3 = [&](int dim, const double* X, double* Y) {
    std::vector<double> b1(dim * na * n * n, 0.0), b2(dim * na * na * n, 0.0);
    for (int kI = 0; kI < n; ++kI) for (int jI = 0; jI < n; ++jI) for (int iI = 0; iI < n; ++iI)
      for (int iO = 0; iO < na; ++iO) for (int d = 0; d < dim; ++d)
        b1[d + dim * (iO + na * (jI + n * kI))] += Vdm[iO + na * iI] * X[d + dim * (iI + n * (jI + n * kI))];
    for (int kI = 0; kI < n; ++kI) for (int jI = 0; jI < n; ++jI) for (int jO = 0; jO < na; ++jO)
      for (int iO = 0; iO < na; ++iO) for (int d = 0; d < dim; ++d)
        b2[d + dim * (iO + na * (jO + na * kI))] += Vdm[jO + na * jI] * b1[d + dim * (iO + na * (jI + n * kI))];
    for (int i = 0; i < dim * nap; ++i) Y[i] = 0.0;
    for (int kI = 0; kI < n; ++kI) for (int kO = 0; kO < na; ++kO) for (int jO = 0; jO < na; ++jO)
      for (int iO = 0; iO < na; ++iO) for (int d = 0; d < dim; ++d)
        Y[d + dim * (iO + na * (jO + na * kO))] += Vdm[kO + na * kI] * b2[d + dim * (iO + na * (jO + na * kI))];
  };
  double Ekin = 0, Ekin_comp = 0, T_mean = 0, Entropy = 0, DR_u = 0, DR_S = 0, DR_Sd = 0, DR_p = 0, ED_S = 0,
         ED_D = 0, Enstr = 0, Vmax = 0;
  std::vector<double> sJa(nap), Ua(5 * nap), gxa(3 * nap), gya(3 * nap), gza(3 * nap), tmp(3 * np);
  for (int e = 0; e < c->nElems; ++e) {
    cb3(1, &c->sJ[(size_t)np * e], sJa.data());
    cb3(5, &U[(size_t)5 * np * e], Ua.data());
    for (int id = 0; id < np; ++id) for (int q = 0; q < 3; ++q) tmp[q + 3 * id] = gx[q + 4 * ((size_t)np * e + id)];
    cb3(3, tmp.data(), gxa.data());
    for (int id = 0; id < np; ++id) for (int q = 0; q < 3; ++q) tmp[q + 3 * id] = gy[q + 4 * ((size_t)np * e + id)];
    cb3(3, tmp.data(), gya.data());
    for (int id = 0; id < np; ++id) for (int q = 0; q < 3; ++q) tmp[q + 3 * id] = gz[q + 4 * ((size_t)np * e + id)];
    cb3(3, tmp.data(), gza.data());
    for (int k = 0; k < na; ++k) for (int j = 0; j < na; ++j) for (int i = 0; i < na; ++i) {
      const int a = i + na * (j + na * k);
      double P[6];
      cons_to_prim(c->e, &Ua[5 * a], P);
      double G[3][3];  // GradVel(:,1)=GradVelx -> G[row][col], col = derivative direction
      for (int q = 0; q < 3; ++q) { G[q][0] = gxa[q + 3 * a]; G[q][1] = gya[q + 3 * a]; G[q][2] = gza[q + 3 * a]; }
      const double divU = G[0][0] + G[1][1] + G[2][2];
      double S[3][3], Sd[3][3];
      for (int r = 0; r < 3; ++r) for (int q = 0; q < 3; ++q) S[r][q] = 0.5 * (G[r][q] + G[q][r]);
      for (int r = 0; r < 3; ++r) for (int q = 0; q < 3; ++q) Sd[r][q] = S[r][q];
      for (int r = 0; r < 3; ++r) Sd[r][r] = Sd[r][r] - 1. / 3. * divU;
      const double vort[3] = {G[2][1] - G[1][2], G[0][2] - G[2][0], G[1][0] - G[0][1]};
      Vmax = std::fmax(Vmax, std::sqrt(vort[0] * vort[0] + vort[1] * vort[1] + vort[2] * vort[2]));
      const double IF = wA[i] * wA[j] * wA[k] / sJa[a];
      const double vv = P[1] * P[1] + P[2] * P[2] + P[3] * P[3];
      T_mean += IF * P[TEMP];
      Entropy += IF * (-c->e.sKappaM1) * P[DENS] * (std::log(P[PRES]) - c->e.kappa * std::log(P[DENS]));
      Ekin += IF * 0.5 * vv;
      Ekin_comp += IF * 0.5 * P[DENS] * vv;
      const double w2 = vort[0] * vort[0] + vort[1] * vort[1] + vort[2] * vort[2];
      Enstr += IF * 0.5 * P[DENS] * w2;
      double sG = 0, sS = 0, sSd = 0;
      for (int q = 0; q < 3; ++q) for (int r = 0; r < 3; ++r) {  // SUM over a column-major 3x3
        sG += G[r][q] * G[r][q]; sS += S[r][q] * S[r][q]; sSd += Sd[r][q] * Sd[r][q];
      }
      DR_u += IF * sG; DR_S += IF * sS; DR_Sd += IF * sSd; DR_p += IF * P[PRES] * divU;
      ED_S += IF * c->e.mu0 * w2; ED_D += IF * c->e.mu0 * divU * divU;
    }
  }
  const double rho0 = 1.0, mu0 = c->e.mu0;
  Ekin = Ekin / Vol; Ekin_comp = Ekin_comp / Vol / rho0; T_mean = T_mean / Vol; Entropy = Entropy / Vol;
  Enstr = Enstr / (rho0 * Vol); DR_u = DR_u * mu0 / Vol; DR_S = DR_S * 2. * mu0 / (rho0 * Vol);
  DR_Sd = DR_Sd * 2. * mu0 / (rho0 * Vol); DR_p = -DR_p / (rho0 * Vol); ED_S = ED_S / Vol; ED_D = ED_D * 4. / 3. / Vol;
  const double uPrime = std::sqrt(2. / 3. * Ekin);
  const double o[15] = {DR_S, DR_Sd + DR_p, Ekin, Ekin_comp, Enstr, DR_u, DR_S, DR_Sd, DR_p, Vmax, T_mean, uPrime, Entropy, ED_S, ED_D};
  for (int i = 0; i < 15; ++i) out[i] = o[i];
}

}  // extern "C"

// Stand-alone entry points for FLEXI's unit tests on a single element (unitTests/ProlongToFace.f90,
// unitTests/SurfInt.f90): the caller supplies SideToElem and the 1D prolongation vectors.
extern "C" void or_prolong_to_face_raw(int N, int nodeType, int nSides, const int* S2E, const double* L_Minus,
                                       const double* L_Plus, int nVar, const double* Uvol, double* Um, double* Us) {
  or_ctx c;
  c.p.nodeType = nodeType;
  c.N = N; c.n = N + 1; c.nf = c.n * c.n; c.np = c.nf * c.n; c.nSides = nSides; c.nElems = 0;
  c.S2E.assign(S2E, S2E + 5 * nSides);
  c.L_Minus.assign(L_Minus, L_Minus + c.n);
  c.L_Plus.assign(L_Plus, L_Plus + c.n);
  c.S2V2.resize(2 * c.nf * 5 * 6);
  or_mappings(N, c.S2V2.data(), nullptr, nullptr);
  prolong_to_face(&c, nVar, Uvol, Um, Us);
}

extern "C" void or_surf_int_raw(int N, int nodeType, int nSides, const int* S2E, const double* L_HatMinus,
                                const double* L_HatPlus, int nVar, const double* Fm, const double* Fs,
                                double slave_sign, double* Ut) {
  or_ctx c;
  c.p.nodeType = nodeType;
  c.N = N; c.n = N + 1; c.nf = c.n * c.n; c.np = c.nf * c.n; c.nSides = nSides; c.nElems = 0;
  c.S2E.assign(S2E, S2E + 5 * nSides);
  c.L_HatMinus.assign(L_HatMinus, L_HatMinus + c.n);
  c.L_HatPlus.assign(L_HatPlus, L_HatPlus + c.n);
  c.S2V2.resize(2 * c.nf * 5 * 6);
  or_mappings(N, c.S2V2.data(), nullptr, nullptr);
  surf_int(&c, nVar, Fm, Fs, slave_sign, Ut);
}
