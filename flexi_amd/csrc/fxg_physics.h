// Point-wise physics of the compressible Navier-Stokes equations for the gfx950 kernels.
// Every function is templated on the arithmetic type T of the flux work (double for the FP64
// path, float for the FP32-bulk path); callers convert state to T after forming it in FP64.
#pragma once
#include <hip/hip_runtime.h>

namespace fxg {

// Gas and transport constants, precomputed on the host in FP64 (src/equations/navierstokes/
// idealgas/eos.f90:106-196): KappaM1, sKappaM1, cp = R*kappa/(kappa-1), lambda = mu0*cp/Pr.
struct Gas {
  double kappa, kappaM1, sKappaM1, R, sR, mu0, lambda, s43, s23;
  double kpmax;  // KAPPASPR_MAX_TIMESTEP_H (eos.h): max(4/3, kappa/Pr)
};

template <class T>
struct GasT {
  T kappa, kappaM1, sR, mu0, lambda, s43, s23;
  __device__ __forceinline__ explicit GasT(const Gas& g)
      : kappa((T)g.kappa), kappaM1((T)g.kappaM1), sR((T)(1.0 / g.R)), mu0((T)g.mu0),
        lambda((T)g.lambda), s43((T)g.s43), s23((T)g.s23) {}
};

// Primitive state (rho, u, v, w, p, T) from conservative variables: ConsToPrim
// (idealgas/eos.f90:203-230). Always evaluated in FP64: the pressure is a difference of two
// large numbers at low Mach number.
struct Prim {
  double rho, u, v, w, p, T, srho;  // srho = 1/rho
};

__device__ __forceinline__ Prim cons_to_prim(const double* c, const Gas& g) {
  Prim q;
  const double sRho = 1.0 / c[0];
  q.rho = c[0];
  q.u = c[1] * sRho;
  q.v = c[2] * sRho;
  q.w = c[3] * sRho;
  q.p = g.kappaM1 * (c[4] - 0.5 * (c[1] * q.u + c[2] * q.v + c[3] * q.w));
  q.T = q.p * sRho * g.sR;
  q.srho = sRho;
  return q;
}

// Viscous (diffusive) flux f_v . n for constant viscosity, following EvalDiffFlux3D_Point
// (src/equations/navierstokes/flux.f90:149-241); g* hold the lifting variables (u,v,w,T).
// fx/fy/fz receive the three Cartesian flux vectors' momentum and energy components
// (the density component is zero).
template <class T>
__device__ __forceinline__ void diff_flux(const GasT<T>& g, T u, T v, T w, const T* gx, const T* gy,
                                          const T* gz, T* fx, T* fy, T* fz) {
  const T mu = g.mu0;
  const T txx = mu * (g.s43 * gx[0] - g.s23 * gy[1] - g.s23 * gz[2]);
  const T tyy = mu * (-g.s23 * gx[0] + g.s43 * gy[1] - g.s23 * gz[2]);
  const T tzz = mu * (-g.s23 * gx[0] - g.s23 * gy[1] + g.s43 * gz[2]);
  const T txy = mu * (gy[0] + gx[1]);
  const T txz = mu * (gz[0] + gx[2]);
  const T tyz = mu * (gz[1] + gy[2]);
  fx[0] = -txx; fx[1] = -txy; fx[2] = -txz;
  fx[3] = -txx * u - txy * v - txz * w - g.lambda * gx[3];
  fy[0] = -txy; fy[1] = -tyy; fy[2] = -tyz;
  fy[3] = -txy * u - tyy * v - tyz * w - g.lambda * gy[3];
  fz[0] = -txz; fz[1] = -tyz; fz[2] = -tzz;
  fz[3] = -txz * u - tyz * v - tzz * w - g.lambda * gz[3];
}

// State rotated into the side frame (n, t1, t2), the extended vector of Riemann_Side
// (src/equations/navierstokes/riemann.f90:295-350).
template <class T>
struct RotState {
  T rho, srho, E, p, vn, vt1, vt2;
};

template <class T>
__device__ __forceinline__ RotState<T> rotate(T rho, T E, T p, T u, T v, T w, const T* n, const T* t1,
                                              const T* t2) {
  RotState<T> s;
  s.rho = rho;
  s.srho = T(1) / rho;
  s.E = E;
  s.p = p;
  s.vn = u * n[0] + v * n[1] + w * n[2];
  s.vt1 = u * t1[0] + v * t1[1] + w * t1[2];
  s.vt2 = u * t2[0] + v * t2[1] + w * t2[2];
  return s;
}

// Central part of the split-form surface flux with the Pirozzoli average
// (SplitSurfaceFluxPI, src/equations/navierstokes/splitflux.f90:672-700), in the side frame.
template <class T>
__device__ __forceinline__ void pi_surface_flux(const RotState<T>& L, const RotState<T>& R, T* F) {
  const T HL = (L.E + L.p) * L.srho;
  const T HR = (R.E + R.p) * R.srho;
  const T rs = L.rho + R.rho;
  const T us = L.vn + R.vn;
  F[0] = T(0.25) * rs * us;
  F[1] = T(0.125) * rs * (us * us) + T(0.5) * (L.p + R.p);
  F[2] = T(0.125) * rs * us * (L.vt1 + R.vt1);
  F[3] = T(0.125) * rs * us * (L.vt2 + R.vt2);
  F[4] = T(0.125) * rs * (HL + HR) * us;
}

// One-dimensional Euler flux in the side frame (EvalEulerFlux1D_fast, flux.f90:396-417).
template <class T>
__device__ __forceinline__ void euler_flux_1d(const RotState<T>& s, T* F) {
  const T m1 = s.rho * s.vn;
  F[0] = m1;
  F[1] = m1 * s.vn + s.p;
  F[2] = m1 * s.vt1;
  F[3] = m1 * s.vt2;
  F[4] = (s.E + s.p) * s.vn;
}

// Numerical flux in the side frame: Lax-Friedrichs (Riemann_LF, riemann.f90:558-589), Roe
// (Riemann_Roe, :668-740) or Roe with Harten-Hyman entropy fix (Riemann_RoeEntropyFix,
// :748-855). The central part is the average of the 1D Euler fluxes (weak form) or the
// Pirozzoli split surface flux (split form).
template <class T>
__device__ __forceinline__ void riemann_rot(const GasT<T>& g, int solver, bool split, const RotState<T>& L,
                                           const RotState<T>& R, T* F) {
  T UL[5] = {L.rho, L.rho * L.vn, L.rho * L.vt1, L.rho * L.vt2, L.E};
  T UR[5] = {R.rho, R.rho * R.vn, R.rho * R.vt1, R.rho * R.vt2, R.E};
  T Fc[5];
  if (split) {
    pi_surface_flux(L, R, Fc);
  } else {
    T FL[5], FR[5];
    euler_flux_1d(L, FL);
    euler_flux_1d(R, FR);
    for (int k = 0; k < 5; ++k) Fc[k] = T(0.5) * (FL[k] + FR[k]);
  }
  if (solver == 1) {  // LF
    const T cL = sqrt(g.kappa * L.p * L.srho);
    const T cR = sqrt(g.kappa * R.p * R.srho);
    const T lmax = fmax(fabs(R.vn), fabs(L.vn)) + fmax(cL, cR);
    for (int k = 0; k < 5; ++k) F[k] = Fc[k] - T(0.5) * lmax * (UR[k] - UL[k]);
    return;
  }
  const T HL = (L.E + L.p) * L.srho;
  const T HR = (R.E + R.p) * R.srho;
  const T sqL = sqrt(L.rho), sqR = sqrt(R.rho);
  const T ss = T(1) / (sqL + sqR);
  const T ru = (sqR * R.vn + sqL * L.vn) * ss;
  const T rv = (sqR * R.vt1 + sqL * L.vt1) * ss;
  const T rw = (sqR * R.vt2 + sqL * L.vt2) * ss;
  const T absVel = ru * ru + rv * rv + rw * rw;
  const T H = (sqR * HR + sqL * HL) * ss;
  const T c = sqrt(g.kappaM1 * (H - T(0.5) * absVel));
  T a[5] = {ru - c, ru, ru, ru, ru + c};
  T al[5];
  if (solver == 3) {  // Roe
    T d[6];
    for (int k = 0; k < 5; ++k) d[k] = UR[k] - UL[k];
    d[5] = d[4] - (d[2] - rv * d[0]) * rv - (d[3] - rw * d[0]) * rw;
    al[2] = d[2] - rv * d[0];
    al[3] = d[3] - rw * d[0];
    const T sc = T(1) / c;
    al[1] = g.kappaM1 * (sc * sc) * (d[0] * (H - ru * ru) - d[5] + ru * d[1]);
    al[0] = T(0.5) * sc * (d[0] * (ru + c) - d[1] - c * al[1]);
    al[4] = d[0] - al[0] - al[1];
    for (int k = 0; k < 5; ++k) a[k] = fabs(a[k]);
  } else {  // RoeEntropyFix
    const T cL = sqrt(g.kappa * L.p * L.srho);
    const T cR = sqrt(g.kappa * R.p * R.srho);
    const T rd = sqrt(L.rho * R.rho);
    const T d1 = R.rho - L.rho, du = R.vn - L.vn, dv = R.vt1 - L.vt1, dw = R.vt2 - L.vt2, dp = R.p - L.p;
    const T sc = T(1) / c;
    const T tmp = T(0.5) * (sc * sc);
    al[0] = tmp * (dp - rd * c * du);
    al[1] = d1 - dp * T(2) * tmp;
    al[2] = rd * dv;
    al[3] = rd * dw;
    al[4] = tmp * (dp + rd * c * du);
    const T aL[5] = {L.vn - cL, L.vn, L.vn, L.vn, L.vn + cL};
    const T aR[5] = {R.vn - cR, R.vn, R.vn, R.vn, R.vn + cR};
    for (int k = 0; k < 5; ++k) {
      const T da = fmax(fmax(T(0), a[k] - aL[k]), aR[k] - a[k]);
      a[k] = (fabs(a[k]) < da) ? T(0.5) * (a[k] * a[k] / da + da) : fabs(a[k]);
    }
  }
  // Dissipation sum_k alpha_k |a_k| r_k with the right eigenvectors of the Roe matrix
  // (riemann.f90:697-701); r1 and r5 carry the unmodified eigenvalues ru -/+ c.
  const T w1 = al[0] * a[0], w2 = al[1] * a[1], w3 = al[2] * a[2], w4 = al[3] * a[3], w5 = al[4] * a[4];
  T D[5];
  D[0] = w1 + w2 + w5;
  D[1] = w1 * (ru - c) + w2 * ru + w5 * (ru + c);
  D[2] = w1 * rv + w2 * rv + w3 + w5 * rv;
  D[3] = w1 * rw + w2 * rw + w4 + w5 * rw;
  D[4] = w1 * (H - ru * c) + w2 * (T(0.5) * absVel) + w3 * rv + w4 * rw + w5 * (H + ru * c);
  for (int k = 0; k < 5; ++k) F[k] = Fc[k] - T(0.5) * D[k];
}

// Per-node spectral radii of CALCTIMESTEP (src/equations/navierstokes/calctimestep.f90:121-190):
// lam[d] convective = |Ja^d.v sJ| + c*MetricsAdv_d, lam[3+d] viscous = mu/rho*MetricsVisc_d. The
// time-step kernel and the fused last LSERK stage both use this, with contraction into FMAs off,
// so their time steps agree bitwise.
__device__ __forceinline__ void dt_node_lambda(const double* c, const Gas& gas, double sj, const double* M0,
                                               const double* M1, const double* M2, double* lam, int& isnan_rho) {
#pragma clang fp contract(off)
  const double srho = 1.0 / c[0];
  const double vel[3] = {c[1] * srho, c[2] * srho, c[3] * srho};
  const double p = gas.kappaM1 * (c[4] - 0.5 * (vel[0] * c[1] + vel[1] * c[2] + vel[2] * c[3]));
  isnan_rho = (c[0] != c[0]) ? 1 : 0;
  const double cs = sqrt(gas.kappa * p * srho);
  const double vsJ[3] = {vel[0] * sj, vel[1] * sj, vel[2] * sj};
  const double* M[3] = {M0, M1, M2};
#pragma unroll
  for (int d = 0; d < 3; ++d) {
    // MetricsAdv and MetricsVisc of InitCalctimestep (calctimestep.f90:42-83), formed here from the
    // metric terms the spectral radius reads anyway instead of being read as two more arrays
    const double adv = sj * sqrt(M[d][0] * M[d][0] + M[d][1] * M[d][1] + M[d][2] * M[d][2]);
    const double a0 = M[d][0] * sj, a1 = M[d][1] * sj, a2 = M[d][2] * sj;
    const double visc = gas.kpmax * (a0 * a0 + a1 * a1 + a2 * a2);
    lam[d] = fabs(M[d][0] * vsJ[0] + M[d][1] * vsJ[1] + M[d][2] * vsJ[2]) + cs * adv;
    lam[3 + d] = gas.mu0 * srho * visc;
  }
}

// Block-wide maxima of lam[6] (NaN propagates, as MAX(..) in the reference) and OR of the NaN-density
// flags; red needs 6*(NT/64) doubles and nanflag NT/64 ints of LDS. Every thread must call it; the
// result is valid in thread 0.
template <int NT>
__device__ __forceinline__ void block_lambda_max(double* lam, int isnan_rho, double* red, int* nanflag, int t,
                                                 double* mx, int& nanr) {
  const int lane = t & 63, wv = t >> 6;
#pragma unroll
  for (int c = 0; c < 6; ++c) {
    double x = lam[c];
    for (int off = 32; off > 0; off >>= 1) {
      const double y = __shfl_down(x, off, 64);
      x = (y > x || y != y) ? y : x;
    }
    if (lane == 0) red[c * (NT / 64) + wv] = x;
  }
  const unsigned long long any = __ballot(isnan_rho);
  if (lane == 0) nanflag[wv] = any ? 1 : 0;
  __syncthreads();
  if (t == 0) {
    for (int c = 0; c < 6; ++c) mx[c] = red[c * (NT / 64)];
    nanr = nanflag[0];
    for (int w = 1; w < NT / 64; ++w) {
      for (int c = 0; c < 6; ++c) {
        const double r = red[c * (NT / 64) + w];
        mx[c] = (r > mx[c] || r != r) ? r : mx[c];
      }
      nanr |= nanflag[w];
    }
  }
}

// The same maxima through LDS rows: every node writes its six values, then the waves of the block
// reduce one value each (64 lanes over the nodes, then a cross-lane tree; a wave takes several
// values when the block has fewer than six). The values leave the registers at once instead of
// being carried through six interleaved cross-lane trees. red needs 6*NP doubles, part 6 doubles
// and nanflag NT/64 ints of LDS (nanflag is written before the first barrier: it must not overlap
// anything another wave may still read); every thread must call it with lam[] set for t < NP; the
// result is valid in thread 0. Two barriers.
template <int NT, int NP>
__device__ __forceinline__ void block_lambda_max_rows(const double* lam, int isnan_rho, double* red, double* part,
                                                      int* nanflag, int t, double* mx, int& nanr) {
  if (t < NP) {
#pragma unroll
    for (int c = 0; c < 6; ++c) red[c * NP + t] = lam[c];
  }
  const unsigned long long any = __ballot(isnan_rho);
  if ((t & 63) == 0) nanflag[t >> 6] = any ? 1 : 0;
  __syncthreads();
  const int l = t & 63;
  for (int c = t >> 6; c < 6; c += NT / 64) {
    // the lanes t >= NP of block_lambda_max contribute zeros; so does the start value here
    double x = NP < NT ? 0.0 : -__builtin_inf();
    for (int p = l; p < NP; p += 64) {
      const double y = red[c * NP + p];
      x = (y > x || y != y) ? y : x;
    }
    for (int off = 32; off > 0; off >>= 1) {
      const double y = __shfl_down(x, off, 64);
      x = (y > x || y != y) ? y : x;
    }
    if (l == 0) part[c] = x;
  }
  __syncthreads();
  if (t == 0) {
#pragma unroll
    for (int c = 0; c < 6; ++c) mx[c] = part[c];
    nanr = 0;
    for (int w = 0; w < NT / 64; ++w) nanr |= nanflag[w];
  }
}

}  // namespace fxg
