// Fused two-kernel stage for Gauss-Lobatto nodes and the split form (the benchmark path).
//
// Per stage, instead of prolongation / lifting / side fluxes / volume integral:
//   A  k_visc_faces (fxg_visc.hip): per element, the BR1 gradients at its boundary nodes and, for each of its
//      six faces, the viscous normal flux f_v . (n SurfElem) seen from that element (4 values per
//      face point, momentum and energy) into the side array FG.
//   B  k_dg_fused:   per element, the face fluxes of its six sides (Riemann solver with the
//      neighbour's face state, plus 0.5*(own - neighbour) viscous normal fluxes), the BR1
//      gradients at all nodes (recomputed, never stored), the split-form volume integral,
//      Ut=-Ut*sJ, the LSERK register update and the new state's face values for the next stage.
// Face states FU are double-buffered: stage s reads FU[s%2] and writes FU[(s+1)%2], so an element
// never reads a face value that a neighbour is rewriting. Nothing is accumulated atomically.
//
// Both kernels first stage, per own face point, the side-array index and the neighbour's face
// state in LDS (one lane per face point); every later phase reads them from there, so the only
// dependent global loads (ElemToSide -> S2V2 -> face array) happen once, at the start.
//
// On Gauss-Lobatto nodes the face value of an element is its boundary node, and the face metric
// n*SurfElem equals NormalSign*Ja^{NormalDir} at that node (src/mesh/metrics.f90:751-787). With
// host-supplied metrics the Riemann flux uses the side's NormVec/SurfElem on both sides, as FLEXI
// does; with metrics computed by the library each element uses its own boundary metric (the
// curl-form face metric depends only on the face, so the two agree to round-off, and the volume and
// surface terms of an element then cancel exactly for a uniform state). The viscous normal fluxes
// and the lifting corrections always take n*SurfElem from the element's own metric terms (equal to
// round-off, and multiplied by small quantities there). Each element gets FLEXI's contributions:
// master +Flux, slave -Flux (src/dg/surfint.t90:471-616) with
// Flux = [F*(U_m,U_s;n_m) + (f_m+f_s).n_m/2]*SurfElem (src/dg/fillflux.f90:40-215).
#include "fxg_bc.h"
#include "fxg_elem.h"

namespace fxg {
namespace {

// ---------------------------------------------------------------------------------------------
// Kernel B: the rest of the stage.
// One LDS array, carved by phase (values in the working precision RF):
//  staging to the gather of the surface terms: rows of np 0 rho, 1 p, 2 H, 3 u, 4 v, 5 w, 6 T, 7 E,
//    then the neighbour face states (7 rows of NF: u, v, w, T, rho, p, E) and the surface
//    contributions (5 rows of NF);
//  direction loop: rows 0..5 as before, the metric terms Ja^1, Ja^2, Ja^3 of every node in rows
//    6..14 (written once, from the registers the lifting loaded them into) and the contravariant
//    viscous flux of the current direction in rows 15..18.
// At N=7 that is 19 rows of 512 doubles (76 KB): with the side slots and the basis two workgroups
// still share a CU's 160 KB, and the metric terms are read from memory once per stage. In FP32-bulk
// (STAGE false) only the current direction's Ja^d is staged (rows 6..8, read from memory in each
// direction, viscous flux in rows 9..12), so that four workgroups still share a CU.
template <int N, bool STAGE>
struct LdsB {
  static constexpr int np = D<N>::np, NF = D<N>::NF;
  static constexpr int FACE = 8 * np;                 // neighbour face states, then surface contributions
  static constexpr int JA = 6 * np;                   // Ja^1..3 (9 rows), or Ja^d (3 rows)
  static constexpr int FV = (STAGE ? 15 : 9) * np;    // contravariant viscous flux, 4 rows
  static constexpr int A = 8 * np + 12 * NF, Bm = FV + 4 * np;
  static constexpr int SIZE = A > Bm ? A : Bm;
};

// The Riemann solver is a template parameter: each instance then carries only its own solver's
// code, which keeps the kernel within 128 VGPRs (two workgroups per CU at N=7) without spills.
// In FP32-bulk the LDS arrays are FP32 (38 KB at N=7) and the kernel is held to 64 VGPRs, so
// four workgroups share a CU (six waves per SIMD below N=7, where 8 would spill, and in the last
// stage of a step).
// LAM: the last stage of a step, which also forms CALCTIMESTEP's per-element spectral radii of the
// new state; in FP64 from the metric terms staged in LDS (the same values as the time-step kernel
// reads), so they are not read from memory a second time.
template <int N, class RF, int SOLVER, bool BC, bool LAM>
__global__ __launch_bounds__(D<N>::NT)
__attribute__((amdgpu_waves_per_eu(sizeof(RF) == 4 ? (N == 7 && !LAM ? 8 : 6) : 1))) void k_dg_fused(const double* U, DevMesh m, const double* __restrict__ FU,
                                                     double* __restrict__ FUnext, const RF* __restrict__ FG,
                                                     Gas gas, double* Ut, RKStage st, RF* __restrict__ Gout,
                                                     const double* __restrict__ bcP) {
  constexpr int n = D<N>::n, nf = D<N>::nf, np = D<N>::np, NT = D<N>::NT, NF = D<N>::NF;
  constexpr bool STAGE = sizeof(RF) == 8;  // all metric terms staged once (FP64)
  using L = LdsB<N, STAGE>;
  // LDS holds the working precision RF: the values are rounded to RF where they are stored, as
  // they were rounded where they are read. sA rows: 0 rho, 1 p, 2 H, 3 u, 4 v, 5 w, 6 T, 7 E
  __shared__ __align__(16) RF sL[L::SIZE];
  __shared__ unsigned sIdx[NF];
  __shared__ double sB[3 * n * n + 4 * n];
  RF* sA = sL;
  // neighbour rows: 0 u, 1 v, 2 w, 3 T, 4 rho, 5 p, 6 E; surface rows follow
  RF* sNb = sL + L::FACE;
  RF* sSurf = sNb + 7 * NF;
  const BasisLayout B{n};
  const GasT<RF> gt(gas);
  const int e = blockIdx.x, t = threadIdx.x;
  const bool act = t < np;
  const size_t gid = (size_t)np * e + (act ? t : 0);
  // the basis, with the DVolSurf block halved (exact) for the split-flux loop
  for (int x = t; x < 3 * n * n + 4 * n; x += NT) {
    const double b = m.basis[x];
    sB[x] = (x >= B.DVolSurf() && x < B.DVolSurf() + n * n) ? 0.5 * b : b;
  }
  if (act) {
    const double* c = U + 5 * gid;
    const Prim q = cons_to_prim(c, gas);
    const double E = c[4];
    sA[t] = q.rho; sA[np + t] = q.p; sA[2 * np + t] = (E + q.p) * q.srho;
    sA[3 * np + t] = q.u; sA[4 * np + t] = q.v; sA[5 * np + t] = q.w; sA[6 * np + t] = q.T; sA[7 * np + t] = E;
  }
  for (int x = t; x < NF; x += NT) {
    const unsigned own = face_slot<N>(m, e, x);
    sIdx[x] = own;
    const double* c = FU + 5 * (size_t)slot_nbr<N>(own);
    const Prim q = cons_to_prim(c, gas);
    sNb[x] = q.u; sNb[NF + x] = q.v; sNb[2 * NF + x] = q.w; sNb[3 * NF + x] = q.T;
    sNb[4 * NF + x] = q.rho; sNb[5 * NF + x] = q.p; sNb[6 * NF + x] = c[4];
  }
  __syncthreads();

  // Surface phase, one lane per own face point: Riemann flux with the neighbour's state and the
  // neighbour's half of the viscous normal flux; the own half is added by the node itself below.
  for (int x = t; x < NF; x += NT) {
    const int ls = x / nf + 1, ab = x - (ls - 1) * nf;
    const int id = vidx<N>(ls, ab % n, ab / n);
    const unsigned own = sIdx[x];
    const bool master = slot_is_master<N>(own);
    const unsigned nb = slot_nbr<N>(own);
    const double ms = master ? 1.0 : -1.0;
    RF gn[4];
#pragma unroll
    for (int c = 0; c < 4; ++c) gn[c] = FG[4 * (size_t)nb + c];
    // (the same arithmetic as face_frame, written out: the register allocation of this kernel
    // stays within 128 VGPRs this way)
    double se;
    RF nv[3];  // master normal
    if (m.ownFaceMetric) {
      // own outward normal times SurfElem from the own metric terms at the boundary node
      const int dir = dir_of(ls);
      const double sgn = normal_sign(ls);
      const double* Jd = (dir == 0 ? m.Mf : (dir == 1 ? m.Mg : m.Mh)) + 3 * ((size_t)np * e + id);
      const double ns0 = sgn * Jd[0], ns1 = sgn * Jd[1], ns2 = sgn * Jd[2];
      se = sqrt(ns0 * ns0 + ns1 * ns1 + ns2 * ns2);
      const double mse = ms / se;
      nv[0] = (RF)(ns0 * mse); nv[1] = (RF)(ns1 * mse); nv[2] = (RF)(ns2 * mse);
    } else {
      // the side's normal and surface element, shared by master and slave as in FLEXI
      const size_t sp = slot_side_point<N>(own);
      se = m.SurfElem[sp];
      nv[0] = (RF)m.NormVec[3 * sp]; nv[1] = (RF)m.NormVec[3 * sp + 1]; nv[2] = (RF)m.NormVec[3 * sp + 2];
    }
    // tangential basis: Gram-Schmidt of the coordinate axis least aligned with n
    RF t1[3], t2[3];
    {
      const RF ax = fabs(nv[0]), ay = fabs(nv[1]), az = fabs(nv[2]);
      RF h[3] = {RF(0), RF(0), RF(0)};
      if (ax <= ay && ax <= az) h[0] = RF(1); else if (ay <= az) h[1] = RF(1); else h[2] = RF(1);
      const RF dt_ = h[0] * nv[0] + h[1] * nv[1] + h[2] * nv[2];
      RF tt[3] = {h[0] - dt_ * nv[0], h[1] - dt_ * nv[1], h[2] - dt_ * nv[2]};
      const RF stn = RF(1) / sqrt(tt[0] * tt[0] + tt[1] * tt[1] + tt[2] * tt[2]);
      t1[0] = tt[0] * stn; t1[1] = tt[1] * stn; t1[2] = tt[2] * stn;
      t2[0] = nv[1] * t1[2] - nv[2] * t1[1];
      t2[1] = nv[2] * t1[0] - nv[0] * t1[2];
      t2[2] = nv[0] * t1[1] - nv[1] * t1[0];
    }
    const RotState<RF> So = rotate<RF>((RF)sA[id], (RF)sA[7 * np + id], (RF)sA[np + id], (RF)sA[3 * np + id],
                                       (RF)sA[4 * np + id], (RF)sA[5 * np + id], nv, t1, t2);
    const RotState<RF> Sn = rotate<RF>((RF)sNb[4 * NF + x], (RF)sNb[6 * NF + x], (RF)sNb[5 * NF + x], (RF)sNb[x],
                                       (RF)sNb[NF + x], (RF)sNb[2 * NF + x], nv, t1, t2);
    RF F[5];
    bool wall = false;
    if constexpr (BC) {
      // a wall side: the inviscid flux is (0, p_wall n, 0) (getboundaryflux.f90:580-592)
      const int side = (int)(own / (2u * (unsigned)nf));
      if (side < m.nBC && !bc_riemann_type(m.bcInfo[2 * side])) {
        wall = true;
        F[0] = RF(0); F[1] = (RF)bcP[4 * slot_side_point<N>(own)]; F[2] = F[3] = F[4] = RF(0);
      }
    }
    if (!wall) {
      // the master's state on the left, by selecting the operands rather than branching: one copy
      // of the solver serves both sides, and a wave that holds master and slave points (N < 7)
      // runs it once
      const RotState<RF> SL = master ? So : Sn, SR = master ? Sn : So;
      riemann_rot<RF>(gt, SOLVER, true, SL, SR, F);
    }
    const double lh = minus_side(ls) ? sB[B.L_HatMinus()] : sB[B.L_HatPlus() + N];
    sSurf[x] = ms * (double)F[0] * se * lh;
#pragma unroll
    for (int c = 0; c < 3; ++c) {
      const RF Fc = nv[c] * F[1] + t1[c] * F[2] + t2[c] * F[3];
      sSurf[(1 + c) * NF + x] = (ms * (double)Fc * se - 0.5 * (double)gn[c]) * lh;
    }
    sSurf[4 * NF + x] = (ms * (double)F[4] * se - 0.5 * (double)gn[3]) * lh;
  }

  // Lifting and viscous fluxes at every node; the own half of the viscous normal face fluxes.
  const int i = t % n, j = (t / n) % n, k = t / (n * n);
  const int sides = act ? node_sides<N>(i, j, k) : 0;
  RF acc[5] = {0, 0, 0, 0, 0};
  RF fv[12];  // contravariant viscous flux Ja^d . (f,g,h)_v, d-major
  // FP64: the node's metric terms, read once, held until the surface terms are gathered and then
  // staged in LDS for the direction loop
  RF JaK[STAGE ? 9 : 1] = {};
  double sjd = 0.0;
  if (act) {
    RF Ja[9], sj;
    load_metrics<RF>(m, gid, Ja, sj);
    if constexpr (STAGE) {
#pragma unroll
      for (int r = 0; r < 9; ++r) JaK[r] = Ja[r];
    }
    RF g[12];
    lift_node<N, RF>(t, sA + 3 * np, sB, sNb, Ja, sj, g);  // u,v,w,T rows
    if (Gout) {  // (np, 12, nElems) like the four-kernel path
#pragma unroll
      for (int c = 0; c < 12; ++c) Gout[t + (size_t)np * (c + 12 * (size_t)e)] = g[c];
    }
    RF fx[4], fy[4], fz[4];
    diff_flux<RF>(gt, (RF)sA[3 * np + t], (RF)sA[4 * np + t], (RF)sA[5 * np + t], g, g + 4, g + 8, fx, fy, fz);
#pragma unroll 1
    for (int ls = 1; ls <= 6; ++ls) {
      if (!(sides & (1 << (ls - 1)))) continue;
      const int dir = dir_of(ls);
      const RF s = (RF)normal_sign(ls);
      const RF n0 = Ja[3 * dir] * s, n1 = Ja[3 * dir + 1] * s, n2 = Ja[3 * dir + 2] * s;
      const RF lh = (RF)(minus_side(ls) ? sB[B.L_HatMinus()] : sB[B.L_HatPlus() + N]);
#pragma unroll
      for (int c = 0; c < 4; ++c) acc[1 + c] += RF(0.5) * (fx[c] * n0 + fy[c] * n1 + fz[c] * n2) * lh;
    }
#pragma unroll
    for (int d = 0; d < 3; ++d)
#pragma unroll
      for (int c = 0; c < 4; ++c) fv[4 * d + c] = Ja[3 * d] * fx[c] + Ja[3 * d + 1] * fy[c] + Ja[3 * d + 2] * fz[c];
  }
  __syncthreads();  // surface contributions complete
  if (sides) {
#pragma unroll 1
    for (int ls = 1; ls <= 6; ++ls) {
      if (!(sides & (1 << (ls - 1)))) continue;
      const int x = face_point<N>(ls, i, j, k);
#pragma unroll
      for (int v = 0; v < 5; ++v) acc[v] += (RF)sSurf[v * NF + x];
    }
  }
  __syncthreads();  // the face rows are free: the metric terms take their place

  // split-form volume integral with the viscous weak-form part (volint.f90:211-371)
  const double* DVh = sB + B.DVolSurf();  // DVolSurf/2
  const double* DH = sB + B.D_Hat_T();
  RF rho = 0, uu = 0, vv = 0, ww = 0, pp = 0, HH = 0;
  RF* sJa = sL + L::JA;
  RF* sF = sL + L::FV;
  if (act) {
    rho = (RF)sA[t]; pp = (RF)sA[np + t]; HH = (RF)sA[2 * np + t];
    uu = (RF)sA[3 * np + t]; vv = (RF)sA[4 * np + t]; ww = (RF)sA[5 * np + t];
    if constexpr (STAGE) {
#pragma unroll
      for (int r = 0; r < 9; ++r) sJa[r * np + t] = JaK[r];
    }
  }
  // the register R and the state U of the update are loaded before the last direction (whose
  // arithmetic then hides their latency; fv of the first two directions no longer holds registers)
  double rpre[5] = {0, 0, 0, 0, 0}, upre[5] = {0, 0, 0, 0, 0};
#pragma unroll
  for (int d = 0; d < 3; ++d) {
    RF* sM = sJa + (STAGE ? 3 * d * np : 0);
    if (act) {
      if constexpr (!STAGE) {
        const double* Jg = (d == 0 ? m.Mf : (d == 1 ? m.Mg : m.Mh)) + 3 * gid;
        sM[t] = (RF)Jg[0]; sM[np + t] = (RF)Jg[1]; sM[2 * np + t] = (RF)Jg[2];
      }
#pragma unroll
      for (int c = 0; c < 4; ++c) sF[c * np + t] = fv[4 * d + c];
      if (d == 2 && st.mode != 0) {
#pragma unroll
        for (int v = 0; v < 5; ++v) upre[v] = U[5 * gid + v];
        if (st.mode == 2)
#pragma unroll
          for (int v = 0; v < 5; ++v) rpre[v] = st.R[5 * gid + v];
      }
    }
    __syncthreads();
    if (act) {
      const RF m0 = sM[t], m1 = sM[np + t], m2 = sM[2 * np + t];
      const int me = d == 0 ? i : (d == 1 ? j : k);
      const int stride = d == 0 ? 1 : (d == 1 ? n : n * n);
      const int base = t - me * stride;
#pragma unroll
      for (int l = 0; l <= N; ++l) {
        const int id = base + l * stride;
        const RF dh = (RF)DH[l + n * me];
#pragma unroll
        for (int c = 0; c < 4; ++c) acc[1 + c] += dh * (RF)sF[c * np + id];
        if (l == me) continue;
        const RF rs = rho + (RF)sA[id];
        const RF us = uu + (RF)sA[3 * np + id], vs = vv + (RF)sA[4 * np + id], ws = ww + (RF)sA[5 * np + id];
        const RF ps = pp + (RF)sA[np + id], Hs = HH + (RF)sA[2 * np + id];
        // Pirozzoli two-point flux with the metric average a = b/2 (SplitVolumeFluxPI,
        // splitflux.f90:603-666), weighted by DV: with b = Ja_i + Ja_l, h = DV/2, Qh = h rs (b.u_sum)/4
        // and P = h p_sum, DV*F = (2 Qh, Qh u_sum + b P, Qh H_sum)
        const RF b0 = m0 + (RF)sM[id], b1 = m1 + (RF)sM[np + id], b2 = m2 + (RF)sM[2 * np + id];
        const RF mb = b0 * us + b1 * vs + b2 * ws;  // 2 (a . u_sum)
        const RF h = (RF)DVh[l + n * me];
        const RF Qh = h * (RF(0.25) * rs * mb);
        const RF P = h * ps;
        acc[0] += RF(2) * Qh;
        acc[1] += Qh * us;
        acc[1] += b0 * P;
        acc[2] += Qh * vs;
        acc[2] += b1 * P;
        acc[3] += Qh * ws;
        acc[3] += b2 * P;
        acc[4] += Qh * Hs;
      }
    }
    __syncthreads();
  }
  // Ut = -Ut*sJ, the register update and the new state's face values for the next stage; in the
  // last stage of a step (LAM, always an LSERK stage) also CALCTIMESTEP's per-element spectral
  // radii of the new state (the time-step kernel's arithmetic, dt_node_lambda)
  double un[5] = {0, 0, 0, 0, 0};
  if constexpr (!LAM) {
    if (act) {
      sjd = m.sJ[gid];
      double ut[5];
#pragma unroll
      for (int v = 0; v < 5; ++v) ut[v] = (-(double)acc[v]) * sjd;
      const size_t o = 5 * gid;
      if (st.mode == 0) {
#pragma unroll
        for (int v = 0; v < 5; ++v) Ut[o + v] = ut[v];
      } else {
#pragma unroll
        for (int v = 0; v < 5; ++v) {
          const double r = (st.mode == 1) ? ut[v] : rpre[v] * st.minusA + ut[v];
          st.R[o + v] = r;
          un[v] = upre[v] + r * st.b_dt;
        }
#pragma unroll
        for (int v = 0; v < 5; ++v) st.U[o + v] = un[v];
#pragma unroll 1
        for (int ls = 1; ls <= 6; ++ls) {
          if (!(sides & (1 << (ls - 1)))) continue;
          double* dst = FUnext + 5 * (size_t)sIdx[face_point<N>(ls, i, j, k)];
#pragma unroll
          for (int v = 0; v < 5; ++v) dst[v] = un[v];
        }
      }
    }
  } else {
    // The spectral radii come before the global stores. The compiler puts a wait for every
    // outstanding vector-memory operation of the wave before the reduction's cross-lane moves (and
    // before the metric loads' use in FP32-bulk), which after the stores took the time of their
    // write-back (about 300 us per launch at 32^3 elements, N=7). The reduction goes through LDS
    // rows, so that the new state and register stay in registers beside it: within 128 VGPRs in
    // FP64; in FP32-bulk this instance is held to six waves per SIMD (76 VGPRs) instead of eight,
    // which spilled.
    double rn[5] = {0, 0, 0, 0, 0};
    if (act) {
      sjd = m.sJ[gid];
#pragma unroll
      for (int v = 0; v < 5; ++v) {
        const double ut = (-(double)acc[v]) * sjd;
        rn[v] = (st.mode == 1) ? ut : rpre[v] * st.minusA + ut;
        un[v] = upre[v] + rn[v] * st.b_dt;
      }
    }
    double lam[6] = {0, 0, 0, 0, 0, 0};
    int isnan_rho = 0;
    if (act) {
      if constexpr (STAGE) {
        double Jk[9];  // Ja^1..3 of the node as staged (the values the time-step kernel reads)
#pragma unroll
        for (int r = 0; r < 9; ++r) Jk[r] = (double)sJa[r * np + t];
        dt_node_lambda(un, gas, sjd, Jk, Jk + 3, Jk + 6, lam, isnan_rho);
      } else {  // FP32-bulk: the metric terms from memory (not staged in full)
        dt_node_lambda(un, gas, sjd, m.Mf + 3 * gid, m.Mg + 3 * gid, m.Mh + 3 * gid, lam, isnan_rho);
      }
    }
    double mx[6];
    int nanr = 0;
    // LDS freed by the direction loop: the rows of 6 NP doubles from the start, the partial maxima
    // and flags in the viscous-flux rows (FP64; not in the metric rows, which other waves may still
    // be reading above) or right after them (FP32-bulk, whose metric terms come from memory)
    double* sRd = reinterpret_cast<double*>(sL);
    double* sPart = STAGE ? reinterpret_cast<double*>(sF) : sRd + 6 * np;
    block_lambda_max_rows<NT, np>(lam, isnan_rho, sRd, sPart, reinterpret_cast<int*>(sPart + 6), t, mx, nanr);
    if (t == 0) {
      st.lamSum[2 * e] = mx[0] + mx[1] + mx[2];
      st.lamSum[2 * e + 1] = mx[3] + mx[4] + mx[5];
      st.lamNan[e] = nanr;
    }
    if (act) {
      const size_t o = 5 * gid;
#pragma unroll
      for (int v = 0; v < 5; ++v) st.R[o + v] = rn[v];
#pragma unroll
      for (int v = 0; v < 5; ++v) st.U[o + v] = un[v];
#pragma unroll 1
      for (int ls = 1; ls <= 6; ++ls) {
        if (!(sides & (1 << (ls - 1)))) continue;
        double* dst = FUnext + 5 * (size_t)sIdx[face_point<N>(ls, i, j, k)];
#pragma unroll
        for (int v = 0; v < 5; ++v) dst[v] = un[v];
      }
    }
  }
}

// Face states of U into FU (all six faces of every element).
template <int N>
__global__ __launch_bounds__(D<N>::NT) void k_faces_of(const double* __restrict__ U, DevMesh m, double* __restrict__ FU) {
  constexpr int n = D<N>::n, np = D<N>::np, NT = D<N>::NT, NF = D<N>::NF;
  __shared__ unsigned sIdx[NF];
  const int e = blockIdx.x, t = threadIdx.x;
  for (int x = t; x < NF; x += NT) sIdx[x] = face_slot_tab<N>(m, e, x);
  __syncthreads();
  if (t >= np) return;
  const int i = t % n, j = (t / n) % n, k = t / (n * n);
  const int sides = node_sides<N>(i, j, k);
  if (!sides) return;
  const double* src = U + 5 * ((size_t)np * e + t);
  double u[5];
#pragma unroll
  for (int v = 0; v < 5; ++v) u[v] = src[v];
#pragma unroll 1
  for (int ls = 1; ls <= 6; ++ls) {
    if (!(sides & (1 << (ls - 1)))) continue;
    double* dst = FU + 5 * (size_t)sIdx[face_point<N>(ls, i, j, k)];
#pragma unroll
    for (int v = 0; v < 5; ++v) dst[v] = u[v];
  }
}

struct Scope {
  KernelTimer* t;
  int id;
  hipStream_t s;
  Scope(KernelTimer* t_, int id_, hipStream_t s_) : t(t_), id(id_), s(s_) { if (t) t->mark(id, true, s); }
  ~Scope() { if (t) t->mark(id, false, s); }
};

template <int N, class RF>
hipError_t fused_n(const DevMesh& m, const FusedWork& w, const Gas& gas, int riemann, const double* U, double* Ut,
                   const RKStage& st, bool faces_fresh, int cur, void* G, hipStream_t s, KernelTimer* tm) {
  const int K = m.nElems;
  constexpr int NT = D<N>::NT;
  if (!faces_fresh) {
    Scope sc(tm, K_FACES, s);
    hipLaunchKernelGGL((k_faces_of<N>), dim3(K), dim3(NT), 0, s, U, m, w.FU[cur]);
  }
  const bool bc = m.nBC > 0;
  {
    Scope sc(tm, K_VISC_FACES, s);
    if (bc) {
      const hipError_t err = launch_bc_ghost(m, w.FU[cur], w.bcP, gas, s);
      if (err != hipSuccess) return err;
    }
    const hipError_t err = launch_visc_faces(m, sizeof(RF) == 4, U, w.Jsoa, w.FU[cur], w.FG, gas, w.gBC, w.bcP, s);
    if (err != hipSuccess) return err;
  }
  {
    Scope sc(tm, K_DG_FUSED, s);
    const double* FUc = w.FU[cur];
    double* FUn = w.FU[1 - cur];
    const RF* FG = (const RF*)w.FG;
#define FXG_LAUNCH(SOLV, BCV)                                                                                       \
  if (st.lamSum)                                                                                                   \
    hipLaunchKernelGGL((k_dg_fused<N, RF, SOLV, BCV, true>), dim3(K), dim3(NT), 0, s, U, m, FUc, FUn, FG, gas, Ut,  \
                       st, (RF*)G, w.bcP);                                                                         \
  else                                                                                                             \
    hipLaunchKernelGGL((k_dg_fused<N, RF, SOLV, BCV, false>), dim3(K), dim3(NT), 0, s, U, m, FUc, FUn, FG, gas, Ut, \
                       st, (RF*)G, w.bcP)
    if (riemann == 1) {
      if (bc) FXG_LAUNCH(1, true); else FXG_LAUNCH(1, false);
    } else if (riemann == 3) {
      if (bc) FXG_LAUNCH(3, true); else FXG_LAUNCH(3, false);
    } else {
      if (bc) FXG_LAUNCH(33, true); else FXG_LAUNCH(33, false);
    }
#undef FXG_LAUNCH
  }
  return hipGetLastError();
}

// Packs the FLEXI-layout metric terms (3,np,K) x 3 and sJ (np,K) into rows of np per element.
// Packs the metric terms of the boundary nodes, (3,np,K) x 3 and sJ (np,K), into rows of NB
// per element in boundary-node order (bnd_rank).
template <int N>
__global__ __launch_bounds__(256) void k_pack_metrics(DevMesh m, size_t nNodes, double* __restrict__ J) {
  constexpr int n = N + 1, np = D<N>::np, NB = Bnd<N>::NB;
  const size_t x = (size_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (x >= nNodes) return;
  const size_t e = x / np;
  const int t = (int)(x - e * np);
  const int i = t % n, j = (t / n) % n, k = t / (n * n);
  if (!node_sides<N>(i, j, k)) return;
  double* Je = J + 10 * NB * e + bnd_rank<N>(i, j, k);
#pragma unroll
  for (int c = 0; c < 3; ++c) {
    Je[(0 + c) * NB] = m.Mf[3 * x + c];
    Je[(3 + c) * NB] = m.Mg[3 * x + c];
    Je[(6 + c) * NB] = m.Mh[3 * x + c];
  }
  Je[9 * NB] = m.sJ[x];
}

template <int N>
__global__ __launch_bounds__(256) void k_slot_table(DevMesh m, size_t total, unsigned* __restrict__ slot) {
  constexpr int NF = 6 * (N + 1) * (N + 1);
  const size_t g = (size_t)blockIdx.x * 256 + threadIdx.x;
  if (g >= total) return;
  const int e = (int)(g / NF), x = (int)(g - (size_t)e * NF);
  slot[g] = face_slot<N>(m, e, x);
}

}  // namespace

hipError_t launch_slot_table(const DevMesh& m, unsigned* slot, hipStream_t s) {
  if (m.slotTab) return hipErrorInvalidValue;
  const size_t total = (size_t)6 * (m.N + 1) * (m.N + 1) * m.nElems;
  const dim3 g((unsigned)((total + 255) / 256)), b(256);
  switch (m.N) {
#define FXG_CASE(NN) \
  case NN: hipLaunchKernelGGL((k_slot_table<NN>), g, b, 0, s, m, total, slot); break;
    FXG_CASE(2) FXG_CASE(3) FXG_CASE(4) FXG_CASE(5) FXG_CASE(6) FXG_CASE(7)
#undef FXG_CASE
    default: return hipErrorInvalidValue;
  }
  return hipGetLastError();
}

hipError_t launch_faces_of(const DevMesh& m, const double* U, double* FU, hipStream_t s) {
  switch (m.N) {
#define FXG_CASE(NN) \
  case NN: hipLaunchKernelGGL((k_faces_of<NN>), dim3(m.nElems), dim3(D<NN>::NT), 0, s, U, m, FU); break;
    FXG_CASE(2) FXG_CASE(3) FXG_CASE(4) FXG_CASE(5) FXG_CASE(6) FXG_CASE(7)
#undef FXG_CASE
    default: return hipErrorInvalidValue;
  }
  return hipGetLastError();
}

size_t packed_metrics_size(int N, int nElems) {
  const int n = N + 1;
  return (size_t)10 * (2 * n * n + (N - 1) * 4 * N) * nElems;
}

hipError_t launch_pack_metrics(const DevMesh& m, double* Jsoa, hipStream_t s) {
  const int np = (m.N + 1) * (m.N + 1) * (m.N + 1);
  const size_t nNodes = (size_t)np * m.nElems;
  const dim3 g((unsigned)((nNodes + 255) / 256)), b(256);
  switch (m.N) {
#define FXG_CASE(NN) \
  case NN: hipLaunchKernelGGL((k_pack_metrics<NN>), g, b, 0, s, m, nNodes, Jsoa); break;
    FXG_CASE(2) FXG_CASE(3) FXG_CASE(4) FXG_CASE(5) FXG_CASE(6) FXG_CASE(7)
#undef FXG_CASE
    default: return hipErrorInvalidValue;
  }
  return hipGetLastError();
}

hipError_t launch_rhs_fused(const DevMesh& m, const FusedWork& w, const Gas& gas, int riemann, int fp32,
                            const double* U, double* Ut, const RKStage& st, bool faces_fresh, int cur, void* G,
                            hipStream_t s, KernelTimer* tm) {
  switch (m.N) {
#define FXG_CASE(NN)                                                                                   \
  case NN:                                                                                             \
    return fp32 ? fused_n<NN, float>(m, w, gas, riemann, U, Ut, st, faces_fresh, cur, G, s, tm)           \
                : fused_n<NN, double>(m, w, gas, riemann, U, Ut, st, faces_fresh, cur, G, s, tm);
    FXG_CASE(2) FXG_CASE(3) FXG_CASE(4) FXG_CASE(5) FXG_CASE(6) FXG_CASE(7)
#undef FXG_CASE
    default: return hipErrorInvalidValue;
  }
}

// Face slots are 32-bit: nSides * 2 * (N+1)^2 must stay below 2^32.
bool fused_supported(int N, int nodeType, int split, long long nSides) {
  return nodeType == 2 && split && N >= 2 && N <= 7 && nSides * 2LL * (N + 1) * (N + 1) < (1LL << 32);
}

}  // namespace fxg
