// Discrete adjoint of the Gauss-Lobatto split-form right-hand side (the fused stage's arithmetic):
// out = (dUt/dU)^T lam at a state U, written as three element kernels and the face-state kernel.
//
// Forward, per element and node i (mu_i = -sJ_i lam_i is the adjoint of the accumulated terms):
//   Ut_i = -sJ_i [ sum_d sum_l DVolSurf(l,i) F#(P_i,P_l; Ja^d_i,Ja^d_l)            split Euler part
//                + sum_d sum_l D_Hat_T(l,i) Ja^d_l . f_v(l)                           viscous volume
//                + sum_{faces} L_hat [ ms F*(U_m,U_s) SurfElem + (g_own - g_nb)/2 ] ] surface
//   with P = (rho,u,v,w,p,H), f_v(q, G(q)) the viscous flux of the BR1 gradients G of the
//   lifting variables q = (u,v,w,T), g = f_v . n SurfElem at a face node.
// Reverse:
//   k_adj_euler  split-flux adjoint (each pair's flux depends on sums of the two nodes' primitive
//                variables, so one pair Jacobian serves both nodes), Riemann-flux adjoint per own
//                face point (directional derivatives with dual numbers, 10 per point), the
//                primitive-to-conservative chain; writes out and, per own face point, the parts
//                that belong to the neighbour: XU (its face state) and Xg (its viscous face flux).
//   k_adj_visc   BR1 gradients recomputed; viscous-flux adjoint (volume, own face half, the
//                neighbour's half from Xg); lifting adjoint (transposed line derivatives and face
//                corrections); adds to out and writes Xq (the neighbour's face-state part).
//   k_adj_gather adds XU and Xq of the neighbours at the face nodes.
// References for the forward terms: src/dg/volint.f90:211-371, src/equations/navierstokes/
// splitflux.f90:603-700, riemann.f90:267-855, fillflux.f90:40-215, surfint.t90:471-616,
// lifting_br1.t90:48-292, flux.f90:149-241.
#include "fxg_dual.h"
#include "fxg_elem.h"

namespace fxg {
namespace {

// Adjoint of the primitive variables (rho,u,v,w,p,T,H) of one node mapped to the conservative
// state: ConsToPrim (idealgas/eos.f90:203-230) and H = (E+p)/rho, differentiated.
__device__ __forceinline__ void prim_adjoint(const Prim& q, double E, const Gas& g, double lrho, double lu, double lv,
                                             double lw, double lp, double lT, double lH, double* lU) {
  const double sr = 1.0 / q.rho;
  const double H = (E + q.p) * sr;
  double lE = lH * sr;
  lp += lH * sr;
  lrho -= H * lH * sr;
  lp += lT * sr * g.sR;
  lrho -= q.T * lT * sr;
  lE += g.kappaM1 * lp;
  lrho += g.kappaM1 * 0.5 * (q.u * q.u + q.v * q.v + q.w * q.w) * lp;
  lU[1] = lu * sr - g.kappaM1 * q.u * lp;
  lU[2] = lv * sr - g.kappaM1 * q.v * lp;
  lU[3] = lw * sr - g.kappaM1 * q.w * lp;
  lrho -= (q.u * lu + q.v * lv + q.w * lw) * sr;
  lU[0] = lrho;
  lU[4] = lE;
}

// Rotated state of conservative variables in dual arithmetic (cons_to_prim + rotate).
__device__ __forceinline__ RotState<Dual> rot_state(const Dual* c, double kappaM1, const Dual* nv, const Dual* t1,
                                                     const Dual* t2) {
  const Dual sr = Dual(1.0) / c[0];
  const Dual u = c[1] * sr, v = c[2] * sr, w = c[3] * sr;
  const Dual p = Dual(kappaM1) * (c[4] - Dual(0.5) * (c[1] * u + c[2] * v + c[3] * w));
  return rotate<Dual>(c[0], c[4], p, u, v, w, nv, t1, t2);
}

// ---------------------------------------------------------------------------------------------
template <int N, int SOLVER>
__global__ __launch_bounds__(D<N>::NT) void k_adj_euler(const double* __restrict__ U, DevMesh m,
                                                      const double* __restrict__ FUa, const double* __restrict__ lam,
                                                      Gas gas, double* __restrict__ out, double* __restrict__ XU,
                                                      double* __restrict__ Xg) {
  constexpr int n = D<N>::n, nf = D<N>::nf, np = D<N>::np, NT = D<N>::NT, NF = D<N>::NF;
  __shared__ double sA[6 * np];   // rho, p, H, u, v, w
  __shared__ double sMu[5 * np];  // -sJ*lam
  __shared__ double sM[3 * np];   // Ja^d of the direction loop
  __shared__ double sLf[5 * NF];  // Riemann adjoint w.r.t. the own face state, per own face point
  __shared__ unsigned sIdx[NF];
  __shared__ double sB[3 * n * n + 4 * n];
  const BasisLayout B{n};
  const int e = blockIdx.x, t = threadIdx.x;
  const bool act = t < np;
  const size_t gid = (size_t)np * e + (act ? t : 0);
  for (int x = t; x < 3 * n * n + 4 * n; x += NT) sB[x] = m.basis[x];
  if (act) {
    const Prim q = cons_to_prim(U + 5 * gid, gas);
    sA[t] = q.rho; sA[np + t] = q.p; sA[2 * np + t] = (U[5 * gid + 4] + q.p) / q.rho;
    sA[3 * np + t] = q.u; sA[4 * np + t] = q.v; sA[5 * np + t] = q.w;
    const double sj = m.sJ[gid];
#pragma unroll
    for (int v = 0; v < 5; ++v) sMu[v * np + t] = -sj * lam[5 * gid + v];
  }
  for (int x = t; x < NF; x += NT) sIdx[x] = face_slot<N>(m, e, x);
  __syncthreads();

  // own face points: the neighbour's share of the viscous face term, and the Riemann adjoint
  for (int x = t; x < NF; x += NT) {
    const int ls = x / nf + 1, ab = x - (ls - 1) * nf;
    const int id = vidx<N>(ls, ab % n, ab / n);
    const unsigned own = sIdx[x];
    const bool master = slot_is_master<N>(own);
    const unsigned nb = slot_nbr<N>(own);
    const double lh = minus_side(ls) ? sB[B.L_HatMinus()] : sB[B.L_HatPlus() + N];
#pragma unroll
    for (int c = 0; c < 4; ++c) Xg[4 * (size_t)own + c] = -0.5 * lh * sMu[(1 + c) * np + id];
    double nv[3], t1[3], t2[3];
    const double se = face_frame<N, double>(m, e, ls, id, own, master, nv, t1, t2);
    const double ms = master ? 1.0 : -1.0;
    double lF[5];
#pragma unroll
    for (int v = 0; v < 5; ++v) lF[v] = ms * se * lh * sMu[v * np + id];
    const double* Uo = U + 5 * ((size_t)np * e + id);
    const double* Un = FUa + 5 * (size_t)nb;
    const Dual nd[3] = {nv[0], nv[1], nv[2]}, t1d[3] = {t1[0], t1[1], t1[2]}, t2d[3] = {t2[0], t2[1], t2[2]};
    const GasT<Dual> gd(gas);
#pragma unroll 1
    for (int k = 0; k < 10; ++k) {
      Dual a[5], b[5];
#pragma unroll
      for (int v = 0; v < 5; ++v) {
        a[v] = Dual(Uo[v], k == v ? 1.0 : 0.0);
        b[v] = Dual(Un[v], k == 5 + v ? 1.0 : 0.0);
      }
      const RotState<Dual> So = rot_state(a, gas.kappaM1, nd, t1d, t2d);
      const RotState<Dual> Sn = rot_state(b, gas.kappaM1, nd, t1d, t2d);
      Dual F[5];
      if (master) riemann_rot<Dual>(gd, SOLVER, true, So, Sn, F);
      else riemann_rot<Dual>(gd, SOLVER, true, Sn, So, F);
      double s = lF[0] * F[0].d + lF[4] * F[4].d;
#pragma unroll
      for (int c = 0; c < 3; ++c) s += lF[1 + c] * (nv[c] * F[1].d + t1[c] * F[2].d + t2[c] * F[3].d);
      if (k < 5) sLf[k * NF + x] = s;
      else XU[5 * (size_t)own + (k - 5)] = s;
    }
  }

  // split-flux adjoint: each pair (i,l) contributes J_s^T (DV(l,i) mu_i + DV(i,l) mu_l) to both
  // nodes' primitive adjoints, J_s the flux Jacobian w.r.t. the pair sums (rho,u,v,w,p,H)
  const int i = t % n, j = (t / n) % n, k = t / (n * n);
  double lrho = 0, lu = 0, lv = 0, lw = 0, lp = 0, lH = 0;
  double rho = 0, uu = 0, vv = 0, ww = 0, pp = 0, HH = 0, mu[5] = {0, 0, 0, 0, 0};
  if (act) {
    rho = sA[t]; pp = sA[np + t]; HH = sA[2 * np + t]; uu = sA[3 * np + t]; vv = sA[4 * np + t]; ww = sA[5 * np + t];
#pragma unroll
    for (int v = 0; v < 5; ++v) mu[v] = sMu[v * np + t];
  }
  const double* DV = sB + B.DVolSurf();
#pragma unroll 1
  for (int d = 0; d < 3; ++d) {
    double m0 = 0, m1 = 0, m2 = 0;
    if (act) {
      const double* Jg = (d == 0 ? m.Mf : (d == 1 ? m.Mg : m.Mh)) + 3 * gid;
      m0 = Jg[0]; m1 = Jg[1]; m2 = Jg[2];
      sM[t] = m0; sM[np + t] = m1; sM[2 * np + t] = m2;
    }
    __syncthreads();
    if (act) {
      const int me = d == 0 ? i : (d == 1 ? j : k);
      const int stride = d == 0 ? 1 : (d == 1 ? n : n * n);
      const int base = t - me * stride;
#pragma unroll 1
      for (int l = 0; l <= N; ++l) {
        if (l == me) continue;
        const int id = base + l * stride;
        const double rs = rho + sA[id];
        const double us = uu + sA[3 * np + id], vs = vv + sA[4 * np + id], ws = ww + sA[5 * np + id];
        const double Hs = HH + sA[2 * np + id];
        const double a0 = 0.5 * (m0 + sM[id]), a1 = 0.5 * (m1 + sM[np + id]), a2 = 0.5 * (m2 + sM[2 * np + id]);
        const double mv = a0 * us + a1 * vs + a2 * ws;
        const double qq = 0.25 * rs * mv;
        const double wi = DV[l + n * me], wl = DV[me + n * l];
        double vec[5];
#pragma unroll
        for (int c = 0; c < 5; ++c) vec[c] = wi * mu[c] + wl * sMu[c * np + id];
        const double beta = 0.5 * vec[0] + 0.25 * (vec[1] * us + vec[2] * vs + vec[3] * ws) + 0.25 * Hs * vec[4];
        lrho += mv * beta;
        lu += a0 * rs * beta + qq * vec[1];
        lv += a1 * rs * beta + qq * vec[2];
        lw += a2 * rs * beta + qq * vec[3];
        lp += a0 * vec[1] + a1 * vec[2] + a2 * vec[3];
        lH += qq * vec[4];
      }
    }
    __syncthreads();
  }
  if (!act) return;
  const Prim q = cons_to_prim(U + 5 * gid, gas);
  double lU[5];
  prim_adjoint(q, U[5 * gid + 4], gas, lrho, lu, lv, lw, lp, 0.0, lH, lU);
  const int sides = node_sides<N>(i, j, k);
#pragma unroll 1
  for (int ls = 1; ls <= 6; ++ls) {
    if (!(sides & (1 << (ls - 1)))) continue;
    const int x = face_point<N>(ls, i, j, k);
#pragma unroll
    for (int v = 0; v < 5; ++v) lU[v] += sLf[v * NF + x];
  }
#pragma unroll
  for (int v = 0; v < 5; ++v) out[5 * gid + v] = lU[v];
}

// ---------------------------------------------------------------------------------------------
template <int N>
__global__ __launch_bounds__(D<N>::NT) void k_adj_visc(const double* __restrict__ U, DevMesh m,
                                                     const double* __restrict__ FUa, const double* __restrict__ lam,
                                                     const double* __restrict__ Xg, Gas gas, double* __restrict__ out,
                                                     double* __restrict__ Xq) {
  constexpr int n = D<N>::n, np = D<N>::np, NT = D<N>::NT, NF = D<N>::NF;
  __shared__ double sP[4 * np];    // u, v, w, T
  __shared__ double sNb[4 * NF];   // neighbour u, v, w, T per own face point
  __shared__ double sMu[4 * np];   // -sJ*lam, momentum and energy
  __shared__ double sK[12 * np];   // lifting adjoint per reference derivative: rows 4*d + var
  __shared__ unsigned sIdx[NF];
  __shared__ double sB[3 * n * n + 4 * n];
  const BasisLayout B{n};
  const int e = blockIdx.x, t = threadIdx.x;
  const bool act = t < np;
  const size_t gid = (size_t)np * e + (act ? t : 0);
  for (int x = t; x < 3 * n * n + 4 * n; x += NT) sB[x] = m.basis[x];
  Prim q{};
  if (act) {
    q = cons_to_prim(U + 5 * gid, gas);
    sP[t] = q.u; sP[np + t] = q.v; sP[2 * np + t] = q.w; sP[3 * np + t] = q.T;
    const double sj = m.sJ[gid];
#pragma unroll
    for (int c = 0; c < 4; ++c) sMu[c * np + t] = -sj * lam[5 * gid + 1 + c];
  }
  for (int x = t; x < NF; x += NT) {
    const unsigned own = face_slot<N>(m, e, x);
    sIdx[x] = own;
    const Prim qn = cons_to_prim(FUa + 5 * (size_t)slot_nbr<N>(own), gas);
    sNb[x] = qn.u; sNb[NF + x] = qn.v; sNb[2 * NF + x] = qn.w; sNb[3 * NF + x] = qn.T;
  }
  __syncthreads();
  const int i = t % n, j = (t / n) % n, k = t / (n * n);
  const int sides = act ? node_sides<N>(i, j, k) : 0;
  double Ja[9], sj = 0;
  double lq[4] = {0, 0, 0, 0};  // adjoint of u, v, w, T at this node
  if (act) {
    load_metrics<double>(m, gid, Ja, sj);
    double g[12];
    lift_node<N, double>(t, sP, sB, sNb, Ja, sj, g);
    // adjoint of the viscous flux vectors fx, fy, fz (momentum and energy rows)
    double lf[3][4] = {{0, 0, 0, 0}, {0, 0, 0, 0}, {0, 0, 0, 0}};
    const double* DH = sB + B.D_Hat_T();
#pragma unroll
    for (int d = 0; d < 3; ++d) {
      const int me = d == 0 ? i : (d == 1 ? j : k);
      const int stride = d == 0 ? 1 : (d == 1 ? n : n * n);
      const int base = t - me * stride;
      double lc[4] = {0, 0, 0, 0};
      for (int l = 0; l <= N; ++l) {
        const double w = DH[me + n * l];
#pragma unroll
        for (int c = 0; c < 4; ++c) lc[c] += w * sMu[c * np + base + l * stride];
      }
#pragma unroll
      for (int x = 0; x < 3; ++x)
#pragma unroll
        for (int c = 0; c < 4; ++c) lf[x][c] += Ja[3 * d + x] * lc[c];
    }
#pragma unroll 1
    for (int ls = 1; ls <= 6; ++ls) {
      if (!(sides & (1 << (ls - 1)))) continue;
      const int dir = dir_of(ls);
      const double sg = normal_sign(ls);
      const double lh = minus_side(ls) ? sB[B.L_HatMinus()] : sB[B.L_HatPlus() + N];
      const unsigned own = sIdx[face_point<N>(ls, i, j, k)];
      const double* xg = Xg + 4 * (size_t)slot_nbr<N>(own);
#pragma unroll
      for (int c = 0; c < 4; ++c) {
        const double a = 0.5 * lh * sMu[c * np + t] + xg[c];
#pragma unroll
        for (int x = 0; x < 3; ++x) lf[x][c] += a * sg * Ja[3 * dir + x];
      }
    }
    // diff_flux adjoint (flux.f90:149-241): stresses, velocities, temperature gradient
    const double mu0 = gas.mu0;
    const double txx = mu0 * (gas.s43 * g[0] - gas.s23 * g[5] - gas.s23 * g[10]);
    const double tyy = mu0 * (-gas.s23 * g[0] + gas.s43 * g[5] - gas.s23 * g[10]);
    const double tzz = mu0 * (-gas.s23 * g[0] - gas.s23 * g[5] + gas.s43 * g[10]);
    const double txy = mu0 * (g[4] + g[1]), txz = mu0 * (g[8] + g[2]), tyz = mu0 * (g[9] + g[6]);
    const double u = q.u, v = q.v, w = q.w;
    const double ltxx = -lf[0][0] - u * lf[0][3];
    const double ltxy = -lf[0][1] - v * lf[0][3] - lf[1][0] - u * lf[1][3];
    const double ltxz = -lf[0][2] - w * lf[0][3] - lf[2][0] - u * lf[2][3];
    const double ltyy = -lf[1][1] - v * lf[1][3];
    const double ltyz = -lf[1][2] - w * lf[1][3] - lf[2][1] - v * lf[2][3];
    const double ltzz = -lf[2][2] - w * lf[2][3];
    lq[0] = -(txx * lf[0][3] + txy * lf[1][3] + txz * lf[2][3]);
    lq[1] = -(txy * lf[0][3] + tyy * lf[1][3] + tyz * lf[2][3]);
    lq[2] = -(txz * lf[0][3] + tyz * lf[1][3] + tzz * lf[2][3]);
    double lg[12];
    lg[0] = mu0 * (gas.s43 * ltxx - gas.s23 * ltyy - gas.s23 * ltzz);   // du/dx
    lg[5] = mu0 * (-gas.s23 * ltxx + gas.s43 * ltyy - gas.s23 * ltzz);  // dv/dy
    lg[10] = mu0 * (-gas.s23 * ltxx - gas.s23 * ltyy + gas.s43 * ltzz); // dw/dz
    lg[4] = mu0 * ltxy; lg[1] = mu0 * ltxy;                              // du/dy, dv/dx
    lg[8] = mu0 * ltxz; lg[2] = mu0 * ltxz;                              // du/dz, dw/dx
    lg[9] = mu0 * ltyz; lg[6] = mu0 * ltyz;                              // dv/dz, dw/dy
    lg[3] = -gas.lambda * lf[0][3]; lg[7] = -gas.lambda * lf[1][3]; lg[11] = -gas.lambda * lf[2][3];
    // lifting adjoint: nu = sJ*lg; reference-derivative adjoint kappa^d = Ja^d . nu
    double nu[12];
#pragma unroll
    for (int c = 0; c < 12; ++c) nu[c] = sj * lg[c];
#pragma unroll
    for (int d = 0; d < 3; ++d)
#pragma unroll
      for (int v2 = 0; v2 < 4; ++v2)
        sK[(4 * d + v2) * np + t] = Ja[3 * d] * nu[v2] + Ja[3 * d + 1] * nu[4 + v2] + Ja[3 * d + 2] * nu[8 + v2];
    // face corrections 0.5*L_hat*(q_nb - q_own).(n SurfElem): own part here, neighbour's via Xq
    const double lhm = sB[B.L_HatMinus()], lhp = sB[B.L_HatPlus() + N];
#pragma unroll 1
    for (int ls = 1; ls <= 6; ++ls) {
      if (!(sides & (1 << (ls - 1)))) continue;
      const int dir = dir_of(ls);
      const double s = normal_sign(ls) * 0.5 * (minus_side(ls) ? lhm : lhp);
      const unsigned own = sIdx[face_point<N>(ls, i, j, k)];
#pragma unroll
      for (int v2 = 0; v2 < 4; ++v2) {
        const double z = s * (Ja[3 * dir] * nu[v2] + Ja[3 * dir + 1] * nu[4 + v2] + Ja[3 * dir + 2] * nu[8 + v2]);
        lq[v2] -= z;
        Xq[4 * (size_t)own + v2] = z;
      }
    }
  }
  __syncthreads();
  if (!act) return;
  // transposed line derivatives: dq/dxi_d at node i = sum_l D_T(l,i) q_l
  const double* DT = sB + B.D_T();
#pragma unroll
  for (int d = 0; d < 3; ++d) {
    const int me = d == 0 ? i : (d == 1 ? j : k);
    const int stride = d == 0 ? 1 : (d == 1 ? n : n * n);
    const int base = t - me * stride;
    for (int l = 0; l <= N; ++l) {
      const double w = DT[me + n * l];
#pragma unroll
      for (int v2 = 0; v2 < 4; ++v2) lq[v2] += w * sK[(4 * d + v2) * np + base + l * stride];
    }
  }
  double lU[5];
  prim_adjoint(q, U[5 * gid + 4], gas, 0.0, lq[0], lq[1], lq[2], 0.0, lq[3], 0.0, lU);
#pragma unroll
  for (int v = 0; v < 5; ++v) out[5 * gid + v] += lU[v];
}

// ---------------------------------------------------------------------------------------------
template <int N>
__global__ __launch_bounds__(D<N>::NT) void k_adj_gather(const double* __restrict__ U, DevMesh m,
                                                       const double* __restrict__ XU, const double* __restrict__ Xq,
                                                       Gas gas, double* __restrict__ out) {
  constexpr int n = D<N>::n, np = D<N>::np, NT = D<N>::NT, NF = D<N>::NF;
  __shared__ unsigned sIdx[NF];
  const int e = blockIdx.x, t = threadIdx.x;
  for (int x = t; x < NF; x += NT) sIdx[x] = face_slot<N>(m, e, x);
  __syncthreads();
  if (t >= np) return;
  const int i = t % n, j = (t / n) % n, k = t / (n * n);
  const int sides = node_sides<N>(i, j, k);
  if (!sides) return;
  const size_t gid = (size_t)np * e + t;
  const Prim q = cons_to_prim(U + 5 * gid, gas);
  double lU[5] = {0, 0, 0, 0, 0}, lq[4] = {0, 0, 0, 0};
#pragma unroll 1
  for (int ls = 1; ls <= 6; ++ls) {
    if (!(sides & (1 << (ls - 1)))) continue;
    const unsigned nb = slot_nbr<N>(sIdx[face_point<N>(ls, i, j, k)]);
#pragma unroll
    for (int v = 0; v < 5; ++v) lU[v] += XU[5 * (size_t)nb + v];
#pragma unroll
    for (int v = 0; v < 4; ++v) lq[v] += Xq[4 * (size_t)nb + v];
  }
  double lP[5];
  prim_adjoint(q, U[5 * gid + 4], gas, 0.0, lq[0], lq[1], lq[2], 0.0, lq[3], 0.0, lP);
#pragma unroll
  for (int v = 0; v < 5; ++v) out[5 * gid + v] += lU[v] + lP[v];
}

// y += a*x and y = a*y over n doubles (the LSERK adjoint's register updates)
__global__ __launch_bounds__(256) void k_axpy(size_t n, double a, const double* __restrict__ x, double* __restrict__ y) {
  const size_t i = (size_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (i < n) y[i] += a * x[i];
}
__global__ __launch_bounds__(256) void k_scale(size_t n, double a, double* __restrict__ y) {
  const size_t i = (size_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (i < n) y[i] *= a;
}

template <int N>
hipError_t vjp_n(const DevMesh& m, const AdjointWork& w, const Gas& gas, int riemann, const double* U,
                 const double* lam, double* out, hipStream_t s) {
  const int K = m.nElems;
  constexpr int NT = D<N>::NT;
  hipError_t err = launch_faces_of(m, U, w.FUa, s);
  if (err != hipSuccess) return err;
  if (riemann == 1)
    hipLaunchKernelGGL((k_adj_euler<N, 1>), dim3(K), dim3(NT), 0, s, U, m, w.FUa, lam, gas, out, w.XU, w.Xg);
  else if (riemann == 3)
    hipLaunchKernelGGL((k_adj_euler<N, 3>), dim3(K), dim3(NT), 0, s, U, m, w.FUa, lam, gas, out, w.XU, w.Xg);
  else
    hipLaunchKernelGGL((k_adj_euler<N, 33>), dim3(K), dim3(NT), 0, s, U, m, w.FUa, lam, gas, out, w.XU, w.Xg);
  hipLaunchKernelGGL((k_adj_visc<N>), dim3(K), dim3(NT), 0, s, U, m, w.FUa, lam, w.Xg, gas, out, w.Xq);
  hipLaunchKernelGGL((k_adj_gather<N>), dim3(K), dim3(NT), 0, s, U, m, w.XU, w.Xq, gas, out);
  return hipGetLastError();
}

}  // namespace

hipError_t launch_rhs_vjp(const DevMesh& m, const AdjointWork& w, const Gas& gas, int riemann, const double* U,
                          const double* lam, double* out, hipStream_t s) {
  switch (m.N) {
    case 2: return vjp_n<2>(m, w, gas, riemann, U, lam, out, s);
    case 3: return vjp_n<3>(m, w, gas, riemann, U, lam, out, s);
    case 4: return vjp_n<4>(m, w, gas, riemann, U, lam, out, s);
    case 5: return vjp_n<5>(m, w, gas, riemann, U, lam, out, s);
    case 6: return vjp_n<6>(m, w, gas, riemann, U, lam, out, s);
    case 7: return vjp_n<7>(m, w, gas, riemann, U, lam, out, s);
    default: return hipErrorInvalidValue;
  }
}

hipError_t launch_axpy(size_t n, double a, const double* x, double* y, hipStream_t s) {
  hipLaunchKernelGGL(k_axpy, dim3((unsigned)((n + 255) / 256)), dim3(256), 0, s, n, a, x, y);
  return hipGetLastError();
}

hipError_t launch_scale(size_t n, double a, double* y, hipStream_t s) {
  hipLaunchKernelGGL(k_scale, dim3((unsigned)((n + 255) / 256)), dim3(256), 0, s, n, a, y);
  return hipGetLastError();
}

}  // namespace fxg
