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
//                face point (the hand-written reverse of riemann_rot; with -DFXG_ADJ_DUAL=1
//                directional derivatives in dual numbers instead), the primitive-to-conservative
//                chain; writes out and, per own face point, the parts that belong to the
//                neighbour: XU (its face state) and Xg (its viscous face flux). In the reverse
//                stepper it first applies the previous reverse stage's register update.
//   k_adj_visc   BR1 gradients recomputed; viscous-flux adjoint (volume, own face half, the
//                neighbour's half from Xg); lifting adjoint (transposed line derivatives and face
//                corrections); adds to out, and adds the neighbour's lifting part, taken to its
//                conservative variables, to XU.
//   k_adj_gather adds the neighbours' XU at the face nodes (k_adj_stage_update: and the register
//                update of the last reverse stage).
// References for the forward terms: src/dg/volint.f90:211-371, src/equations/navierstokes/
// splitflux.f90:603-700, riemann.f90:267-855, fillflux.f90:40-215, surfint.t90:471-616,
// lifting_br1.t90:48-292, flux.f90:149-241.
#include "fxg_dual.h"
#include "fxg_elem.h"

// FXG_ADJ_DUAL=1 takes the Riemann-flux derivatives from dual-number evaluations instead of the
// hand-written reverse (a cross-check of the reverse; slower).
#ifndef FXG_ADJ_DUAL
#define FXG_ADJ_DUAL 0
#endif

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

// Reverse of riemann_rot (split form) at one side point: lF is the adjoint of F in the side
// frame; the adjoints of the two rotated states' (rho, E, p, vn, vt1, vt2) are returned in aL, aR.
// Line by line the transpose of the forward arithmetic in fxg_physics.h (riemann.f90:558-855,
// splitflux.f90:672-700); srho = 1/rho is folded into rho. Divisions are taken as products with
// reciprocals formed once (1/rho, 1/c, 1/sqrt(rho)): the adjoint's rounding needs no match.
struct RotAdj {
  double rho, E, p, vn, vt1, vt2;
};

template <int SOLVER>
__device__ __forceinline__ void riemann_rot_vjp(const Gas& g, const RotState<double>& L, const RotState<double>& R,
                                                const double* lF, RotAdj& aL, RotAdj& aR) {
  aL = RotAdj{0, 0, 0, 0, 0, 0};
  aR = RotAdj{0, 0, 0, 0, 0, 0};
  // central part, Pirozzoli surface flux: F += Fc
  {
    const double HL = (L.E + L.p) * L.srho, HR = (R.E + R.p) * R.srho;
    const double rs = L.rho + R.rho, us = L.vn + R.vn, sv1 = L.vt1 + R.vt1, sv2 = L.vt2 + R.vt2, Hs = HL + HR;
    const double lrs = 0.25 * us * lF[0] + 0.125 * us * us * lF[1] + 0.125 * us * sv1 * lF[2] +
                       0.125 * us * sv2 * lF[3] + 0.125 * Hs * us * lF[4];
    const double lus = 0.25 * rs * lF[0] + 0.25 * rs * us * lF[1] + 0.125 * rs * sv1 * lF[2] +
                       0.125 * rs * sv2 * lF[3] + 0.125 * rs * Hs * lF[4];
    const double lsv1 = 0.125 * rs * us * lF[2], lsv2 = 0.125 * rs * us * lF[3], lHs = 0.125 * rs * us * lF[4];
    aL.p += 0.5 * lF[1]; aR.p += 0.5 * lF[1];
    aL.rho += lrs - HL * lHs * L.srho; aR.rho += lrs - HR * lHs * R.srho;
    aL.vn += lus; aR.vn += lus;
    aL.vt1 += lsv1; aR.vt1 += lsv1;
    aL.vt2 += lsv2; aR.vt2 += lsv2;
    aL.E += lHs * L.srho; aR.E += lHs * R.srho;
    aL.p += lHs * L.srho; aR.p += lHs * R.srho;
  }
  const double UL[5] = {L.rho, L.rho * L.vn, L.rho * L.vt1, L.rho * L.vt2, L.E};
  const double UR[5] = {R.rho, R.rho * R.vn, R.rho * R.vt1, R.rho * R.vt2, R.E};
  double ld[5] = {0, 0, 0, 0, 0};  // adjoint of UR - UL
  // d(sqrt(kappa p / rho)): c = sqrt(kappa p srho)
  auto sound_adj = [&](const RotState<double>& S, double c, double lc, RotAdj& a) {
    const double f = 0.5 * g.kappa / c;
    a.p += f * S.srho * lc;
    a.rho -= f * S.p * S.srho * S.srho * lc;
  };
  if (SOLVER == 1) {  // Lax-Friedrichs: F -= 0.5*lmax*(UR - UL)
    const double cL = sqrt(g.kappa * L.p * L.srho), cR = sqrt(g.kappa * R.p * R.srho);
    const double vmax = fmax(fabs(R.vn), fabs(L.vn)), cmax = fmax(cL, cR);
    const double lmax = vmax + cmax;
    double llmax = 0.0;
    for (int k = 0; k < 5; ++k) {
      ld[k] = -0.5 * lmax * lF[k];
      llmax -= 0.5 * (UR[k] - UL[k]) * lF[k];
    }
    if (fabs(R.vn) >= fabs(L.vn)) aR.vn += (R.vn < 0.0 ? -llmax : llmax);
    else aL.vn += (L.vn < 0.0 ? -llmax : llmax);
    if (cL >= cR) sound_adj(L, cL, llmax, aL);
    else sound_adj(R, cR, llmax, aR);
  } else {
    // Roe averages
    const double HLr = (L.E + L.p) * L.srho, HRr = (R.E + R.p) * R.srho;
    const double sqL = sqrt(L.rho), sqR = sqrt(R.rho);
    const double ss = 1.0 / (sqL + sqR), isqL = 1.0 / sqL, isqR = 1.0 / sqR;
    const double ru = (sqR * R.vn + sqL * L.vn) * ss;
    const double rv = (sqR * R.vt1 + sqL * L.vt1) * ss;
    const double rw = (sqR * R.vt2 + sqL * L.vt2) * ss;
    const double absVel = ru * ru + rv * rv + rw * rw;
    const double H = (sqR * HRr + sqL * HLr) * ss;
    const double c = sqrt(g.kappaM1 * (H - 0.5 * absVel));
    const double ic = 1.0 / c;
    double al[5], a[5];
    double d[5];
    for (int k = 0; k < 5; ++k) d[k] = UR[k] - UL[k];
    // entropy-fix quantities
    double cL = 0, cR = 0, rd = 0, tmp = 0, du = 0, dp = 0;
    if (SOLVER == 3) {
      const double d5 = d[4] - (d[2] - rv * d[0]) * rv - (d[3] - rw * d[0]) * rw;
      al[2] = d[2] - rv * d[0];
      al[3] = d[3] - rw * d[0];
      al[1] = g.kappaM1 * ic * ic * (d[0] * (H - ru * ru) - d5 + ru * d[1]);
      al[0] = 0.5 * ic * (d[0] * (ru + c) - d[1] - c * al[1]);
      al[4] = d[0] - al[0] - al[1];
      a[0] = fabs(ru - c); a[1] = a[2] = a[3] = fabs(ru); a[4] = fabs(ru + c);
    } else {
      cL = sqrt(g.kappa * L.p * L.srho); cR = sqrt(g.kappa * R.p * R.srho);
      rd = sqrt(L.rho * R.rho);
      const double d1 = R.rho - L.rho, dv = R.vt1 - L.vt1, dw = R.vt2 - L.vt2;
      du = R.vn - L.vn; dp = R.p - L.p;
      tmp = 0.5 * ic * ic;
      al[0] = tmp * (dp - rd * c * du);
      al[1] = d1 - dp * 2.0 * tmp;
      al[2] = rd * dv;
      al[3] = rd * dw;
      al[4] = tmp * (dp + rd * c * du);
      const double ar[5] = {ru - c, ru, ru, ru, ru + c};
      const double aLk[5] = {L.vn - cL, L.vn, L.vn, L.vn, L.vn + cL};
      const double aRk[5] = {R.vn - cR, R.vn, R.vn, R.vn, R.vn + cR};
      for (int k = 0; k < 5; ++k) {
        const double da = fmax(fmax(0.0, ar[k] - aLk[k]), aRk[k] - ar[k]);
        a[k] = (fabs(ar[k]) < da) ? 0.5 * (ar[k] * ar[k] / da + da) : fabs(ar[k]);
      }
    }
    const double w[5] = {al[0] * a[0], al[1] * a[1], al[2] * a[2], al[3] * a[3], al[4] * a[4]};
    // F -= 0.5*D with D the dissipation sum_k w_k r_k
    double lD[5];
    for (int k = 0; k < 5; ++k) lD[k] = -0.5 * lF[k];
    double lw[5];
    lw[0] = lD[0] + (ru - c) * lD[1] + rv * lD[2] + rw * lD[3] + (H - ru * c) * lD[4];
    lw[1] = lD[0] + ru * lD[1] + rv * lD[2] + rw * lD[3] + 0.5 * absVel * lD[4];
    lw[2] = lD[2] + rv * lD[4];
    lw[3] = lD[3] + rw * lD[4];
    lw[4] = lD[0] + (ru + c) * lD[1] + rv * lD[2] + rw * lD[3] + (H + ru * c) * lD[4];
    const double wsum = w[0] + w[1] + w[4];
    double lru = lD[1] * wsum + lD[4] * (w[4] - w[0]) * c;
    double lc = lD[1] * (w[4] - w[0]) + lD[4] * (w[4] - w[0]) * ru;
    double lrv = lD[2] * wsum + lD[4] * w[2];
    double lrw = lD[3] * wsum + lD[4] * w[3];
    double lH = lD[4] * (w[0] + w[4]);
    double labs = 0.5 * w[1] * lD[4];
    double lal[5], la[5];
    for (int k = 0; k < 5; ++k) {
      lal[k] = a[k] * lw[k];
      la[k] = al[k] * lw[k];
    }
    if (SOLVER == 3) {
      // al4 = d0 - al0 - al1
      ld[0] += lal[4]; lal[0] -= lal[4]; lal[1] -= lal[4];
      // al0 = 0.5/c*Y, Y = d0(ru+c) - d1 - c al1
      const double lY = 0.5 * lal[0] * ic;
      lc -= al[0] * ic * lal[0];
      ld[0] += (ru + c) * lY; lru += d[0] * lY; lc += (d[0] - al[1]) * lY; ld[1] -= lY; lal[1] -= c * lY;
      // al1 = kappaM1/c^2*X, X = d0(H - ru^2) - d5 + ru d1
      const double lX = g.kappaM1 * lal[1] * ic * ic;
      lc -= 2.0 * al[1] * ic * lal[1];
      double ld5 = -lX;
      ld[0] += (H - ru * ru) * lX; lH += d[0] * lX; lru += (d[1] - 2.0 * ru * d[0]) * lX; ld[1] += ru * lX;
      // al2 = d2 - rv d0, al3 = d3 - rw d0
      ld[2] += lal[2]; lrv -= d[0] * lal[2]; ld[0] -= rv * lal[2];
      ld[3] += lal[3]; lrw -= d[0] * lal[3]; ld[0] -= rw * lal[3];
      // d5 = d4 - e2 rv - e3 rw, e2 = d2 - rv d0, e3 = d3 - rw d0
      const double e2 = d[2] - rv * d[0], e3 = d[3] - rw * d[0];
      ld[4] += ld5;
      const double le2 = -rv * ld5, le3 = -rw * ld5;
      lrv -= e2 * ld5; lrw -= e3 * ld5;
      ld[2] += le2; lrv -= d[0] * le2; ld[0] -= rv * le2;
      ld[3] += le3; lrw -= d[0] * le3; ld[0] -= rw * le3;
      // a = |ru -/+ c|, |ru|
      const double s0 = (ru - c) < 0.0 ? -1.0 : 1.0, s1 = ru < 0.0 ? -1.0 : 1.0, s4 = (ru + c) < 0.0 ? -1.0 : 1.0;
      lru += s0 * la[0] + s1 * (la[1] + la[2] + la[3]) + s4 * la[4];
      lc += -s0 * la[0] + s4 * la[4];
    } else {
      // entropy-fixed |a_k|
      const double ar[5] = {ru - c, ru, ru, ru, ru + c};
      const double aLk[5] = {L.vn - cL, L.vn, L.vn, L.vn, L.vn + cL};
      const double aRk[5] = {R.vn - cR, R.vn, R.vn, R.vn, R.vn + cR};
      double lar[5] = {0, 0, 0, 0, 0}, laL[5] = {0, 0, 0, 0, 0}, laR[5] = {0, 0, 0, 0, 0};
      for (int k = 0; k < 5; ++k) {
        const double x = ar[k] - aLk[k], y = aRk[k] - ar[k];
        const double m1 = fmax(0.0, x);
        const double da = fmax(m1, y);
        double lda = 0.0;
        if (fabs(ar[k]) < da) {
          lar[k] += ar[k] / da * la[k];
          lda = 0.5 * (1.0 - ar[k] * ar[k] / (da * da)) * la[k];
        } else {
          lar[k] += (ar[k] < 0.0 ? -la[k] : la[k]);
        }
        if (m1 >= y) {
          if (x > 0.0) { lar[k] += lda; laL[k] -= lda; }
        } else {
          laR[k] += lda; lar[k] -= lda;
        }
      }
      lru += lar[0] + lar[1] + lar[2] + lar[3] + lar[4];
      lc += -lar[0] + lar[4];
      aL.vn += laL[0] + laL[1] + laL[2] + laL[3] + laL[4];
      aR.vn += laR[0] + laR[1] + laR[2] + laR[3] + laR[4];
      double lcL = -laL[0] + laL[4], lcR = -laR[0] + laR[4];
      // wave strengths from primitive jumps
      const double dv = R.vt1 - L.vt1, dw = R.vt2 - L.vt2;
      double ltmp = (dp - rd * c * du) * lal[0] + (dp + rd * c * du) * lal[4] - 2.0 * dp * lal[1];
      const double ldp = tmp * (lal[0] + lal[4]) - 2.0 * tmp * lal[1];
      const double lrd = -tmp * c * du * lal[0] + tmp * c * du * lal[4] + dv * lal[2] + dw * lal[3];
      lc += -tmp * rd * du * lal[0] + tmp * rd * du * lal[4];
      const double ldu = -tmp * rd * c * lal[0] + tmp * rd * c * lal[4];
      const double ld1 = lal[1];
      const double ldv = rd * lal[2], ldw = rd * lal[3];
      lc -= ltmp * ic * ic * ic;  // tmp = 0.5/c^2
      const double hrd = 0.5 / rd;
      aL.rho += R.rho * hrd * lrd; aR.rho += L.rho * hrd * lrd;
      aR.rho += ld1; aL.rho -= ld1;
      aR.vn += ldu; aL.vn -= ldu;
      aR.vt1 += ldv; aL.vt1 -= ldv;
      aR.vt2 += ldw; aL.vt2 -= ldw;
      aR.p += ldp; aL.p -= ldp;
      sound_adj(L, cL, lcL, aL);
      sound_adj(R, cR, lcR, aR);
    }
    // c = sqrt(kappaM1 (H - absVel/2))
    lH += 0.5 * g.kappaM1 * ic * lc;
    labs -= 0.25 * g.kappaM1 * ic * lc;
    lru += 2.0 * ru * labs; lrv += 2.0 * rv * labs; lrw += 2.0 * rw * labs;
    // H, ru, rv, rw: Roe averages with weights sqL*ss, sqR*ss
    double lsqL = HLr * ss * lH + L.vn * ss * lru + L.vt1 * ss * lrv + L.vt2 * ss * lrw;
    double lsqR = HRr * ss * lH + R.vn * ss * lru + R.vt1 * ss * lrv + R.vt2 * ss * lrw;
    const double lss = (sqR * HRr + sqL * HLr) * lH + (sqR * R.vn + sqL * L.vn) * lru +
                       (sqR * R.vt1 + sqL * L.vt1) * lrv + (sqR * R.vt2 + sqL * L.vt2) * lrw;
    lsqL -= ss * ss * lss; lsqR -= ss * ss * lss;
    aL.vn += sqL * ss * lru; aR.vn += sqR * ss * lru;
    aL.vt1 += sqL * ss * lrv; aR.vt1 += sqR * ss * lrv;
    aL.vt2 += sqL * ss * lrw; aR.vt2 += sqR * ss * lrw;
    const double lHL = sqL * ss * lH, lHR = sqR * ss * lH;
    aL.E += L.srho * lHL; aL.p += L.srho * lHL; aL.rho -= (L.E + L.p) * L.srho * L.srho * lHL;
    aR.E += R.srho * lHR; aR.p += R.srho * lHR; aR.rho -= (R.E + R.p) * R.srho * R.srho * lHR;
    aL.rho += 0.5 * lsqL * isqL; aR.rho += 0.5 * lsqR * isqR;
  }
  // U = (rho, rho vn, rho vt1, rho vt2, E) of both sides through d = UR - UL
  aR.rho += ld[0] + R.vn * ld[1] + R.vt1 * ld[2] + R.vt2 * ld[3];
  aR.vn += R.rho * ld[1]; aR.vt1 += R.rho * ld[2]; aR.vt2 += R.rho * ld[3]; aR.E += ld[4];
  aL.rho -= ld[0] + L.vn * ld[1] + L.vt1 * ld[2] + L.vt2 * ld[3];
  aL.vn -= L.rho * ld[1]; aL.vt1 -= L.rho * ld[2]; aL.vt2 -= L.rho * ld[3]; aL.E -= ld[4];
}

// Adjoint of a rotated state (rho, E, p, vn, vt1, vt2) to the conservative state it was built
// from (cons_to_prim + rotation by n, t1, t2).
__device__ __forceinline__ void rot_state_adjoint(const double* c, const Gas& g, const double* nv, const double* t1,
                                                  const double* t2, const RotAdj& a, double* lU) {
  const double sr = 1.0 / c[0];
  const double u = c[1] * sr, v = c[2] * sr, w = c[3] * sr;
  const double lu = a.vn * nv[0] + a.vt1 * t1[0] + a.vt2 * t2[0];
  const double lv = a.vn * nv[1] + a.vt1 * t1[1] + a.vt2 * t2[1];
  const double lw = a.vn * nv[2] + a.vt1 * t1[2] + a.vt2 * t2[2];
  // p = kappaM1 (E - (m.u)/2)
  lU[4] = a.E + g.kappaM1 * a.p;
  lU[1] = lu * sr - g.kappaM1 * u * a.p;
  lU[2] = lv * sr - g.kappaM1 * v * a.p;
  lU[3] = lw * sr - g.kappaM1 * w * a.p;
  lU[0] = a.rho + g.kappaM1 * 0.5 * (u * u + v * v + w * w) * a.p - (u * lu + v * lv + w * lw) * sr;
}

// ---------------------------------------------------------------------------------------------
// Roe without entropy fix at N=5,6 (256/384-lane workgroups of 216/343 active lanes): three waves
// per SIMD at 168 VGPRs with a 20-28 byte spill instead of two at 174-176. (At N=7 a 512-lane
// workgroup cannot gain a wave that way; at N=3,4 the spill is larger.)
template <int N, int SOLVER>
__global__ __launch_bounds__(D<N>::NT)
__attribute__((amdgpu_waves_per_eu(SOLVER == 3 && (N == 5 || N == 6) ? 3 : 1))) void k_adj_euler(
    const double* __restrict__ U, DevMesh m, const double* __restrict__ FUa, double* __restrict__ lam,
    double* __restrict__ ub, double cub, Gas gas, double* __restrict__ out, double* __restrict__ XU,
    double* __restrict__ Xg, Pending pd) {
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
  const int i = t % n, j = (t / n) % n, k = t / (n * n);
  // Staging. The independent loads (ElemToSide of the lane's first own face point, the basis, the
  // node's U, sJ, lam, ub and out) are issued together, before the first wait; all stores come
  // after the loads (a store between two loads makes the second wait for the first's write-back).
  constexpr int NBAS = 3 * n * n + 4 * n;
  static_assert(NBAS <= NT, "one basis entry per lane");
  const int2 sfe = face_side_entry<N>(m, e, t < NF ? t : 0);  // (lanes t >= NF do not use it)
  const double bas = t < NBAS ? m.basis[t] : 0.0;
  double c5[5] = {1.0, 1.0, 1.0, 1.0, 1.0}, l0[5] = {0, 0, 0, 0, 0}, u0[5] = {0, 0, 0, 0, 0},
         o5[5] = {0, 0, 0, 0, 0}, sj = 0.0;
  if (act) {
#pragma unroll
    for (int v = 0; v < 5; ++v) c5[v] = U[5 * gid + v];
    sj = m.sJ[gid];
#pragma unroll
    for (int v = 0; v < 5; ++v) l0[v] = lam[5 * gid + v];
    if (ub) {
#pragma unroll
      for (int v = 0; v < 5; ++v) u0[v] = ub[5 * gid + v];
    }
    if (pd.XU) {
#pragma unroll
      for (int v = 0; v < 5; ++v) o5[v] = out[5 * gid + v];
    }
  }
  if (t < NBAS) sB[t] = bas;
  for (int x = t; x < NF; x += NT) {
    const unsigned own = face_slot_from<N>(m, x == t ? sfe : face_side_entry<N>(m, e, x), x);
    sIdx[x] = own;
    if (pd.XU) {  // the neighbours' face parts, staged in sLf (free until the surface phase)
      const size_t nb = slot_nbr<N>(own);
#pragma unroll
      for (int v = 0; v < 5; ++v) sLf[v * NF + x] = pd.XU[5 * nb + v];
    }
  }
  if (pd.XU) __syncthreads();  // uniform over the launch
  if (act) {
    const Prim q = cons_to_prim(c5, gas);
    sA[t] = q.rho; sA[np + t] = q.p; sA[2 * np + t] = (c5[4] + q.p) / q.rho;
    sA[3 * np + t] = q.u; sA[4 * np + t] = q.v; sA[5 * np + t] = q.w;
    double l5[5];
    if (pd.XU) {
      // the register update of the previous reverse stage: ub += its out + the neighbours' face
      // parts, rbar = a*(rbar + cub*ub) (rbar lives in lam)
      double g5[5];
#pragma unroll
      for (int v = 0; v < 5; ++v) g5[v] = o5[v];
      const int sides = node_sides<N>(i, j, k);
#pragma unroll 1
      for (int ls = 1; ls <= 6; ++ls) {
        if (!(sides & (1 << (ls - 1)))) continue;
        const int x = face_point<N>(ls, i, j, k);
#pragma unroll
        for (int v = 0; v < 5; ++v) g5[v] += sLf[v * NF + x];
      }
#pragma unroll
      for (int v = 0; v < 5; ++v) {
        const double r = pd.a * (l0[v] + pd.cub * u0[v]);
        const double u1 = u0[v] + g5[v];
        lam[5 * gid + v] = r;
        ub[5 * gid + v] = u1;
        l5[v] = r + cub * u1;
      }
    } else {
#pragma unroll
      for (int v = 0; v < 5; ++v) l5[v] = l0[v] + (ub ? cub * u0[v] : 0.0);
    }
#pragma unroll
    for (int v = 0; v < 5; ++v) sMu[v * np + t] = -sj * l5[v];
  }
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
#if FXG_ADJ_DUAL
    // directional derivatives in dual arithmetic (10 evaluations): the reference for the reverse
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
#else
    // flux adjoint in the side frame, then the reverse of the Riemann solver and of the rotation
    double lFr[5];
    lFr[0] = lF[0];
    lFr[1] = nv[0] * lF[1] + nv[1] * lF[2] + nv[2] * lF[3];
    lFr[2] = t1[0] * lF[1] + t1[1] * lF[2] + t1[2] * lF[3];
    lFr[3] = t2[0] * lF[1] + t2[1] * lF[2] + t2[2] * lF[3];
    lFr[4] = lF[4];
    const Prim po = cons_to_prim(Uo, gas), pn = cons_to_prim(Un, gas);
    const RotState<double> So = rotate<double>(po.rho, Uo[4], po.p, po.u, po.v, po.w, nv, t1, t2);
    const RotState<double> Sn = rotate<double>(pn.rho, Un[4], pn.p, pn.u, pn.v, pn.w, nv, t1, t2);
    // the master's state on the left, operands selected rather than branched on (one copy of the
    // reverse; a wave may hold master and slave points)
    const RotState<double> SL = master ? So : Sn, SR = master ? Sn : So;
    RotAdj aL, aR;
    riemann_rot_vjp<SOLVER>(gas, SL, SR, lFr, aL, aR);
    const RotAdj ao = master ? aL : aR, an = master ? aR : aL;
    double gOwn[5], gNb[5];
    rot_state_adjoint(Uo, gas, nv, t1, t2, ao, gOwn);
    rot_state_adjoint(Un, gas, nv, t1, t2, an, gNb);
#pragma unroll
    for (int v = 0; v < 5; ++v) {
      sLf[v * NF + x] = gOwn[v];
      XU[5 * (size_t)own + v] = gNb[v];
    }
#endif
  }

  // split-flux adjoint: each pair (i,l) contributes J_s^T (DV(l,i) mu_i + DV(i,l) mu_l) to both
  // nodes' primitive adjoints, J_s the flux Jacobian w.r.t. the pair sums (rho,u,v,w,p,H)
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
#pragma unroll
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
// XU: the Riemann part of the neighbour's face adjoint (k_adj_euler); the lifting part w.r.t. the
// neighbour's (u,v,w,T), taken to its conservative variables with its face state, is added here.
template <int N>
__global__ __launch_bounds__(D<N>::NT) __attribute__((amdgpu_waves_per_eu(4))) void k_adj_visc(const double* __restrict__ U, DevMesh m,
                                                     const double* __restrict__ FUa, const double* __restrict__ lam,
                                                     const double* __restrict__ ub, double cub,
                                                     const double* __restrict__ Xg, Gas gas, double* __restrict__ out,
                                                     double* __restrict__ XU) {
  constexpr int n = D<N>::n, np = D<N>::np, NT = D<N>::NT, NF = D<N>::NF;
  // sW: u, v, w, T (rows 0-3) and -sJ*lam of momentum and energy (rows 4-7) until the node
  // values are formed, then the lifting adjoint per reference derivative (rows 4*d + var)
  __shared__ double sW[12 * np];
  double* sP = sW;
  double* sMu = sW + 4 * np;
  double* sK = sW;
  __shared__ double sNb[5 * NF];   // neighbour u, v, w, T, rho per own face point
  __shared__ unsigned sIdx[NF];
  __shared__ double sB[3 * n * n + 4 * n];
  const BasisLayout B{n};
  const int e = blockIdx.x, t = threadIdx.x;
  const bool act = t < np;
  const size_t gid = (size_t)np * e + (act ? t : 0);
  // Staging: the independent loads (ElemToSide of the lane's first own face point, the basis, the
  // node's U, sJ, lam and ub) are issued together, before the first wait.
  constexpr int NBAS = 3 * n * n + 4 * n;
  static_assert(NBAS <= NT, "one basis entry per lane");
  const int2 sfe = face_side_entry<N>(m, e, t < NF ? t : 0);  // (lanes t >= NF do not use it)
  const double bas = t < NBAS ? m.basis[t] : 0.0;
  double c5[5] = {1.0, 1.0, 1.0, 1.0, 1.0}, l4[4] = {0, 0, 0, 0}, u4[4] = {0, 0, 0, 0}, sjn = 0.0;
  if (act) {
#pragma unroll
    for (int v = 0; v < 5; ++v) c5[v] = U[5 * gid + v];
    sjn = m.sJ[gid];
#pragma unroll
    for (int c = 0; c < 4; ++c) l4[c] = lam[5 * gid + 1 + c];
    if (ub) {
#pragma unroll
      for (int c = 0; c < 4; ++c) u4[c] = ub[5 * gid + 1 + c];
    }
  }
  if (t < NBAS) sB[t] = bas;
  Prim q{};
  if (act) {
    q = cons_to_prim(c5, gas);
    sP[t] = q.u; sP[np + t] = q.v; sP[2 * np + t] = q.w; sP[3 * np + t] = q.T;
#pragma unroll
    for (int c = 0; c < 4; ++c) sMu[c * np + t] = -sjn * (l4[c] + (ub ? cub * u4[c] : 0.0));
  }
  // where rows 8..11 of sW (free until the lifting adjoint is formed) hold NF values, the
  // neighbours' viscous shares Xg are staged there with the face states, so that the side loop
  // below reads them from LDS instead of waiting for one load per side
  constexpr bool XG_LDS = np >= NF;  // N >= 5
  double* sXg = sW + 8 * np;
  for (int x = t; x < NF; x += NT) {
    const unsigned own = face_slot_from<N>(m, x == t ? sfe : face_side_entry<N>(m, e, x), x);
    sIdx[x] = own;
    const size_t nb = slot_nbr<N>(own);
    if constexpr (XG_LDS) {
#pragma unroll
      for (int c = 0; c < 4; ++c) sXg[c * NF + x] = Xg[4 * nb + c];
    }
    const Prim qn = cons_to_prim(FUa + 5 * nb, gas);
    sNb[x] = qn.u; sNb[NF + x] = qn.v; sNb[2 * NF + x] = qn.w; sNb[3 * NF + x] = qn.T; sNb[4 * NF + x] = qn.rho;
  }
  __syncthreads();
  const int i = t % n, j = (t / n) % n, k = t / (n * n);
  const int sides = act ? node_sides<N>(i, j, k) : 0;
  double Ja[9], sj = 0;
  double lq[4] = {0, 0, 0, 0};  // adjoint of u, v, w, T at this node
  double nu[12];
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
      const int x = face_point<N>(ls, i, j, k);
      const double* xg = Xg + 4 * (size_t)slot_nbr<N>(sIdx[x]);
#pragma unroll
      for (int c = 0; c < 4; ++c) {
        const double a = 0.5 * lh * sMu[c * np + t] + (XG_LDS ? sXg[c * NF + x] : xg[c]);
#pragma unroll
        for (int r = 0; r < 3; ++r) lf[r][c] += a * sg * Ja[3 * dir + r];
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
    // lifting adjoint: nu = sJ*lg
#pragma unroll
    for (int c = 0; c < 12; ++c) nu[c] = sj * lg[c];
  }
  __syncthreads();  // sP and sMu are no longer read: sK takes their place
  if (act) {
    // reference-derivative adjoint kappa^d = Ja^d . nu
#pragma unroll
    for (int d = 0; d < 3; ++d)
#pragma unroll
      for (int v2 = 0; v2 < 4; ++v2)
        sK[(4 * d + v2) * np + t] = Ja[3 * d] * nu[v2] + Ja[3 * d + 1] * nu[4 + v2] + Ja[3 * d + 2] * nu[8 + v2];
    // face corrections 0.5*L_hat*(q_nb - q_own).(n SurfElem): own part here, the neighbour's
    // part to its conservative variables (its state is the face state FUa of the neighbour slot)
    // and into its exchange slot
    const double lhm = sB[B.L_HatMinus()], lhp = sB[B.L_HatPlus() + N];
#pragma unroll 1
    for (int ls = 1; ls <= 6; ++ls) {
      if (!(sides & (1 << (ls - 1)))) continue;
      const int dir = dir_of(ls);
      const double s = normal_sign(ls) * 0.5 * (minus_side(ls) ? lhm : lhp);
      const int x = face_point<N>(ls, i, j, k);
      double z[4];
#pragma unroll
      for (int v2 = 0; v2 < 4; ++v2) {
        z[v2] = s * (Ja[3 * dir] * nu[v2] + Ja[3 * dir + 1] * nu[4 + v2] + Ja[3 * dir + 2] * nu[8 + v2]);
        lq[v2] -= z[v2];
      }
      const Prim qn{sNb[4 * NF + x], sNb[x], sNb[NF + x], sNb[2 * NF + x], 0.0, sNb[3 * NF + x]};
      double lP[5];
      prim_adjoint(qn, 0.0, gas, 0.0, z[0], z[1], z[2], 0.0, z[3], 0.0, lP);
      // in place of the neighbour values of this face point (read above; no other lane has x)
#pragma unroll
      for (int v = 0; v < 5; ++v) sNb[v * NF + x] = lP[v];
    }
  }
  __syncthreads();
  // into the neighbour's exchange slots, one lane per own face point: the read-modify-writes of
  // a node's sides then do not wait for each other's stores
  for (int x = t; x < NF; x += NT) {
    const unsigned own = sIdx[x];
#pragma unroll
    for (int v = 0; v < 5; ++v) XU[5 * (size_t)own + v] += sNb[v * NF + x];
  }
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
// out += the neighbours' face parts (XU of the neighbour slot of every own face point)
template <int N>
__global__ __launch_bounds__(D<N>::NT) void k_adj_gather(DevMesh m, const double* __restrict__ XU,
                                                       double* __restrict__ out) {
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
  double lU[5] = {0, 0, 0, 0, 0};
#pragma unroll 1
  for (int ls = 1; ls <= 6; ++ls) {
    if (!(sides & (1 << (ls - 1)))) continue;
    const unsigned nb = slot_nbr<N>(sIdx[face_point<N>(ls, i, j, k)]);
#pragma unroll
    for (int v = 0; v < 5; ++v) lU[v] += XU[5 * (size_t)nb + v];
  }
#pragma unroll
  for (int v = 0; v < 5; ++v) out[5 * gid + v] += lU[v];
}

// The register update of a reverse LSERK stage (the last one of a chain; the others are applied
// by the next stage's k_adj_euler): ub += out + the neighbours' face parts, rbar = a*(rbar + cub*ub).
template <int N>
__global__ __launch_bounds__(D<N>::NT) void k_adj_stage_update(DevMesh m, const double* __restrict__ XU,
                                                             const double* __restrict__ out, double* __restrict__ ub,
                                                             double* __restrict__ rbar, double cub, double a) {
  constexpr int n = D<N>::n, np = D<N>::np, NT = D<N>::NT, NF = D<N>::NF;
  __shared__ unsigned sIdx[NF];
  const int e = blockIdx.x, t = threadIdx.x;
  for (int x = t; x < NF; x += NT) sIdx[x] = face_slot<N>(m, e, x);
  __syncthreads();
  if (t >= np) return;
  const int i = t % n, j = (t / n) % n, k = t / (n * n);
  const int sides = node_sides<N>(i, j, k);
  const size_t gid = (size_t)np * e + t;
  double g5[5];
#pragma unroll
  for (int v = 0; v < 5; ++v) g5[v] = out[5 * gid + v];
#pragma unroll 1
  for (int ls = 1; ls <= 6; ++ls) {
    if (!(sides & (1 << (ls - 1)))) continue;
    const unsigned nb = slot_nbr<N>(sIdx[face_point<N>(ls, i, j, k)]);
#pragma unroll
    for (int v = 0; v < 5; ++v) g5[v] += XU[5 * (size_t)nb + v];
  }
#pragma unroll
  for (int v = 0; v < 5; ++v) {
    const double u0 = ub[5 * gid + v];
    rbar[5 * gid + v] = a * (rbar[5 * gid + v] + cub * u0);
    ub[5 * gid + v] = u0 + g5[v];
  }
}

// y += a*x over n doubles
__global__ __launch_bounds__(256) void k_axpy(size_t n, double a, const double* __restrict__ x, double* __restrict__ y) {
  const size_t i = (size_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (i < n) y[i] += a * x[i];
}

// The transposed-Jacobian parts of one right-hand side at U (face states FU): out = the element's
// own parts, XU = the parts of the neighbours' face nodes. lam (+ cub*ub when ub is given) is the
// adjoint of Ut; with pd.XU a previous stage's register update is applied first (k_adj_euler).
template <int N>
hipError_t adj_parts(const DevMesh& m, const AdjointWork& w, const Gas& gas, int riemann, const double* U,
                     const double* FU, double* lam, double* ub, double cub, double* out, double* XU, Pending pd,
                     hipStream_t s) {
  const int K = m.nElems;
  constexpr int NT = D<N>::NT;
  if (riemann == 1)
    hipLaunchKernelGGL((k_adj_euler<N, 1>), dim3(K), dim3(NT), 0, s, U, m, FU, lam, ub, cub, gas, out, XU, w.Xg, pd);
  else if (riemann == 3)
    hipLaunchKernelGGL((k_adj_euler<N, 3>), dim3(K), dim3(NT), 0, s, U, m, FU, lam, ub, cub, gas, out, XU, w.Xg, pd);
  else
    hipLaunchKernelGGL((k_adj_euler<N, 33>), dim3(K), dim3(NT), 0, s, U, m, FU, lam, ub, cub, gas, out, XU, w.Xg,
                       pd);
  hipLaunchKernelGGL((k_adj_visc<N>), dim3(K), dim3(NT), 0, s, U, m, FU, (const double*)lam, (const double*)ub, cub,
                     (const double*)w.Xg, gas, out, XU);
  return hipGetLastError();
}

template <int N>
hipError_t adj_gather(const DevMesh& m, const double* XU, double* out, hipStream_t s) {
  hipLaunchKernelGGL((k_adj_gather<N>), dim3(m.nElems), dim3(D<N>::NT), 0, s, m, XU, out);
  return hipGetLastError();
}

template <int N>
hipError_t adj_update(const DevMesh& m, const double* XU, const double* out, double* ub, double* rbar, double cub,
                      double a, hipStream_t s) {
  hipLaunchKernelGGL((k_adj_stage_update<N>), dim3(m.nElems), dim3(D<N>::NT), 0, s, m, XU, out, ub, rbar, cub, a);
  return hipGetLastError();
}

}  // namespace

#define FXG_ADJ_SWITCH(CALL)                                                                     \
  switch (m.N) {                                                                                 \
    case 2: return CALL(2);                                                                      \
    case 3: return CALL(3);                                                                      \
    case 4: return CALL(4);                                                                      \
    case 5: return CALL(5);                                                                      \
    case 6: return CALL(6);                                                                      \
    case 7: return CALL(7);                                                                      \
    default: return hipErrorInvalidValue;                                                        \
  }

hipError_t launch_rhs_vjp(const DevMesh& m, const AdjointWork& w, const Gas& gas, int riemann, const double* U,
                          const double* lam, double* out, hipStream_t s) {
  hipError_t err = launch_faces_of(m, U, w.FUa, s);
  if (err != hipSuccess) return err;
  double* lw = const_cast<double*>(lam);  // read only without a pending update
#define FXG_PARTS(NN) adj_parts<NN>(m, w, gas, riemann, U, w.FUa, lw, nullptr, 0.0, out, w.XU[0], Pending{}, s)
  err = [&]() -> hipError_t { FXG_ADJ_SWITCH(FXG_PARTS) }();
#undef FXG_PARTS
  if (err != hipSuccess) return err;
#define FXG_GATHER(NN) adj_gather<NN>(m, w.XU[0], out, s)
  FXG_ADJ_SWITCH(FXG_GATHER)
#undef FXG_GATHER
}

hipError_t launch_adjoint_stage(const DevMesh& m, const AdjointWork& w, const Gas& gas, int riemann, const double* U,
                                const double* FU, double* rbar, double* ubar, double cub, double a, double* out,
                                AdjointChain& ch, hipStream_t s) {
  if (!FU) {
    const hipError_t err = launch_faces_of(m, U, w.FUa, s);
    if (err != hipSuccess) return err;
    FU = w.FUa;
  }
  const Pending pd{ch.pending ? w.XU[ch.cur] : nullptr, ch.cub, ch.a};
  const int nxt = ch.pending ? 1 - ch.cur : ch.cur;
#define FXG_PARTS(NN) adj_parts<NN>(m, w, gas, riemann, U, FU, rbar, ubar, cub, out, w.XU[nxt], pd, s)
  const hipError_t err = [&]() -> hipError_t { FXG_ADJ_SWITCH(FXG_PARTS) }();
#undef FXG_PARTS
  ch.cur = nxt;
  ch.pending = true;
  ch.cub = cub;
  ch.a = a;
  return err;
}

hipError_t launch_adjoint_flush(const DevMesh& m, const AdjointWork& w, double* rbar, double* ubar,
                                const double* out, AdjointChain& ch, hipStream_t s) {
  if (!ch.pending) return hipSuccess;
  ch.pending = false;
#define FXG_UPDATE(NN) adj_update<NN>(m, w.XU[ch.cur], out, ubar, rbar, ch.cub, ch.a, s)
  FXG_ADJ_SWITCH(FXG_UPDATE)
#undef FXG_UPDATE
}
#undef FXG_ADJ_SWITCH

hipError_t launch_axpy(size_t n, double a, const double* x, double* y, hipStream_t s) {
  hipLaunchKernelGGL(k_axpy, dim3((unsigned)((n + 255) / 256)), dim3(256), 0, s, n, a, x, y);
  return hipGetLastError();
}

}  // namespace fxg
