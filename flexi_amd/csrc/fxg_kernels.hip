// gfx950 kernels of the DGSEM right-hand side (FLEXI's DGTimeDerivative_weakForm) with BR1
// lifting, the fused LSERK register update and the CFL time step.
//
// Element kernels run one workgroup per element with one lane per volume node (N+1)^3, rounded
// up to whole 64-lane wavefronts; element data are staged through LDS with coalesced loads.
// Side kernels run one lane per side point. No atomics: every element gathers the fluxes of its
// own six sides, so results do not depend on scheduling.
#include "fxg_kernels.h"
#include "fxg_bc.h"

namespace fxg {

template <int N>
struct Dim {
  static constexpr int n = N + 1;
  static constexpr int nf = n * n;
  static constexpr int np = n * n * n;
  static constexpr int NT = ((np + 63) / 64) * 64;
};

// Local side ids of flexi.h: 1 zeta-, 2 eta-, 3 xi+, 4 eta+, 5 xi-, 6 zeta+.
__device__ __forceinline__ bool is_minus(int ls) { return ls == 5 || ls == 2 || ls == 1; }
__device__ __forceinline__ int side_dir(int ls) { return (ls == 5 || ls == 3) ? 0 : (ls == 2 || ls == 4) ? 1 : 2; }

// Volume node of local side ls at face coordinates (a,b) and normal coordinate l; the face
// coordinates follow EvalElemFaceG/DoSurfInt (prolongtoface.t90:211-264, surfint.t90:683-766).
template <int N>
__device__ __forceinline__ int vol_index(int ls, int a, int b, int l) {
  constexpr int n = N + 1;
  switch (side_dir(ls)) {
    case 0: return l + n * (a + n * b);
    case 1: return a + n * (l + n * b);
    default: return a + n * (b + n * l);
  }
}

// Face coordinates (a,b) and normal coordinate l of node (i,j,k) with respect to local side ls.
__device__ __forceinline__ void face_coords(int ls, int i, int j, int k, int& a, int& b, int& l) {
  switch (side_dir(ls)) {
    case 0: a = j; b = k; l = i; break;
    case 1: a = i; b = k; l = j; break;
    default: a = i; b = j; l = k; break;
  }
}

template <int N>
__device__ __forceinline__ void load_basis(const double* __restrict__ g, double* s, int t, int nt) {
  constexpr int n = N + 1;
  for (int x = t; x < 3 * n * n + 4 * n; x += nt) s[x] = g[x];
}

// The element's U (5*np doubles, contiguous lanes) and the basis into LDS, every load issued
// before the first wait: a loop with the wait inside would spend one memory latency per
// iteration (five at N=7).
template <int N, int NT>
__device__ __forceinline__ void stage_u_basis(const double* __restrict__ Ue, double* sU, const double* __restrict__ g,
                                              double* sB, int t) {
  constexpr int n = N + 1, np = n * n * n, NBAS = 3 * n * n + 4 * n, RU = (5 * np + NT - 1) / NT;
  static_assert(NBAS <= NT, "one basis entry per lane");
  double ul[RU];
#pragma unroll
  for (int r = 0; r < RU; ++r) ul[r] = t + r * NT < 5 * np ? Ue[t + r * NT] : 0.0;
  const double b = t < NBAS ? g[t] : 0.0;
#pragma unroll
  for (int r = 0; r < RU; ++r)
    if (t + r * NT < 5 * np) sU[t + r * NT] = ul[r];
  if (t < NBAS) sB[t] = b;
}

// ---------------------------------------------------------------------------------------------
// ProlongToFace of the conservative state (prolongtoface.t90:21-144): every element writes the
// face values of its six sides into the master or slave face array of the side, reordered into
// the side frame with S2V2.
template <int N, bool GL>
__global__ __launch_bounds__(Dim<N>::NT) void k_prolong_u(const double* __restrict__ U, DevMesh m,
                                                        double* __restrict__ Um, double* __restrict__ Us) {
  constexpr int n = Dim<N>::n, nf = Dim<N>::nf, np = Dim<N>::np, NT = Dim<N>::NT;
  __shared__ double sU[5 * np];
  __shared__ double sB[3 * n * n + 4 * n];
  const int e = blockIdx.x;
  const int t = threadIdx.x;
  const double* Ue = U + (size_t)5 * np * e;
  stage_u_basis<N, NT>(Ue, sU, m.basis, sB, t);
  __syncthreads();
  const BasisLayout B{n};
  for (int x = t; x < 6 * nf; x += NT) {
    const int ls = x / nf + 1, pq = x - (ls - 1) * nf;
    const int side = m.E2S[2 * (ls - 1) + 12 * e] - 1;
    const int flip = m.E2S[2 * (ls - 1) + 1 + 12 * e];
    const int* ab = m.s2v2 + 2 * (pq + nf * (flip + 5 * (ls - 1)));
    double out[5];
    if (GL) {
      const int id = vol_index<N>(ls, ab[0], ab[1], is_minus(ls) ? 0 : N);
#pragma unroll
      for (int v = 0; v < 5; ++v) out[v] = sU[5 * id + v];
    } else {
      const double* L = sB + (is_minus(ls) ? B.L_Minus() : B.L_Plus());
      const int id0 = vol_index<N>(ls, ab[0], ab[1], 0);
#pragma unroll
      for (int v = 0; v < 5; ++v) out[v] = sU[5 * id0 + v] * L[0];
      for (int l = 1; l <= N; ++l) {
        const int id = vol_index<N>(ls, ab[0], ab[1], l);
#pragma unroll
        for (int v = 0; v < 5; ++v) out[v] = out[v] + sU[5 * id + v] * L[l];
      }
    }
    double* dst = (flip == 0 ? Um : Us) + 5 * ((size_t)pq + (size_t)nf * side);
#pragma unroll
    for (int v = 0; v < 5; ++v) dst[v] = out[v];
  }
}

// ---------------------------------------------------------------------------------------------
// BR1 lifting (Lifting_BR1, lifting_br1.t90:48-181; strong form, non-conservative volume
// integral lifting_volint.t90:26-118; surface flux lifting_fillflux.t90:40-239), Jacobian,
// and prolongation of the lifted gradients to the element's six sides.
template <int N, bool GL, class RF>
__global__ __launch_bounds__(Dim<N>::NT) void k_lift(const double* __restrict__ U, DevMesh m,
                                                   const double* __restrict__ Um, const double* __restrict__ Us,
                                                   const double* __restrict__ bcLift, Gas gas, RF* __restrict__ G,
                                                   RF* __restrict__ Gm, RF* __restrict__ Gs) {
  constexpr int n = Dim<N>::n, nf = Dim<N>::nf, np = Dim<N>::np, NT = Dim<N>::NT;
  __shared__ double sA[12 * np];
  __shared__ double sB[3 * n * n + 4 * n];
  const BasisLayout B{n};
  const int e = blockIdx.x;
  const int t = threadIdx.x;
  const double* Ue = U + (size_t)5 * np * e;
  stage_u_basis<N, NT>(Ue, sA, m.basis, sB, t);
  __syncthreads();
  double* sP = sA + 5 * np;  // lifting variables (u,v,w,T), SoA
  if (t < np) {
    const Prim q = cons_to_prim(sA + 5 * t, gas);
    sP[t] = q.u;
    sP[np + t] = q.v;
    sP[2 * np + t] = q.w;
    sP[3 * np + t] = q.T;
  }
  __syncthreads();
  RF g[12];
  if (t < np) {
    const int i = t % n, j = (t / n) % n, k = t / (n * n);
    const double* DT = sB + B.D_T();
    RF gxi[4], geta[4], gzeta[4];
#pragma unroll
    for (int v = 0; v < 4; ++v) {
      gxi[v] = (RF)DT[0 + n * i] * (RF)sP[v * np + (0 + n * (j + n * k))];
      geta[v] = (RF)DT[0 + n * j] * (RF)sP[v * np + (i + n * (0 + n * k))];
      gzeta[v] = (RF)DT[0 + n * k] * (RF)sP[v * np + (i + n * (j + n * 0))];
    }
    for (int l = 1; l <= N; ++l) {
      const RF dxi = (RF)DT[l + n * i], deta = (RF)DT[l + n * j], dzeta = (RF)DT[l + n * k];
#pragma unroll
      for (int v = 0; v < 4; ++v) {
        gxi[v] = gxi[v] + dxi * (RF)sP[v * np + (l + n * (j + n * k))];
        geta[v] = geta[v] + deta * (RF)sP[v * np + (i + n * (l + n * k))];
        gzeta[v] = gzeta[v] + dzeta * (RF)sP[v * np + (i + n * (j + n * l))];
      }
    }
    const size_t gid = (size_t)np * e + t;
    RF mf[3], mg[3], mh[3];
#pragma unroll
    for (int c = 0; c < 3; ++c) {
      mf[c] = (RF)m.Mf[3 * gid + c];
      mg[c] = (RF)m.Mg[3 * gid + c];
      mh[c] = (RF)m.Mh[3 * gid + c];
    }
#pragma unroll
    for (int d = 0; d < 3; ++d)
#pragma unroll
      for (int v = 0; v < 4; ++v) g[4 * d + v] = mf[d] * gxi[v] + mg[d] * geta[v] + mh[d] * gzeta[v];
    // surface contributions: 0.5*(U_slave - U_master)*SurfElem*n, same sign on both sides
#pragma unroll 1
    for (int ls = 1; ls <= 6; ++ls) {
      int a, b, l;
      face_coords(ls, i, j, k, a, b, l);
      if (GL && l != (is_minus(ls) ? 0 : N)) continue;
      const int side = m.E2S[2 * (ls - 1) + 12 * e] - 1;
      const int flip = m.E2S[2 * (ls - 1) + 1 + 12 * e];
      const int* pq2 = m.v2s2 + 2 * ((a + n * b) + nf * (flip + 5 * (ls - 1)));
      const size_t fp = (size_t)(pq2[0] + n * pq2[1]) + (size_t)nf * side;
      const RF se = (RF)m.SurfElem[fp];
      const RF nv[3] = {(RF)m.NormVec[3 * fp], (RF)m.NormVec[3 * fp + 1], (RF)m.NormVec[3 * fp + 2]};
      const RF lh = GL ? (RF)(is_minus(ls) ? sB[B.L_HatMinus()] : sB[B.L_HatPlus() + N])
                       : (RF)(is_minus(ls) ? sB[B.L_HatMinus() + l] : sB[B.L_HatPlus() + l]);
      RF f[4];
      if (side < m.nBC) {  // boundary side: Lifting_GetBoundaryFlux (k_bc_lift) times SurfElem
#pragma unroll
        for (int v = 0; v < 4; ++v) f[v] = (RF)bcLift[4 * fp + v] * se;
      } else {
        const Prim pm = cons_to_prim(Um + 5 * fp, gas);
        const Prim ps = cons_to_prim(Us + 5 * fp, gas);
        const RF dq[4] = {(RF)ps.u - (RF)pm.u, (RF)ps.v - (RF)pm.v, (RF)ps.w - (RF)pm.w, (RF)ps.T - (RF)pm.T};
#pragma unroll
        for (int v = 0; v < 4; ++v) f[v] = RF(0.5) * dq[v] * se;
      }
#pragma unroll
      for (int v = 0; v < 4; ++v)
#pragma unroll
        for (int d = 0; d < 3; ++d) g[4 * d + v] = g[4 * d + v] + f[v] * nv[d] * lh;
    }
    const RF sj = (RF)m.sJ[gid];
#pragma unroll
    for (int c = 0; c < 12; ++c) g[c] = g[c] * sj;
  }
  __syncthreads();  // all reads of sP are done before sA is overwritten
  if (t < np) {
    RF* Ge = G + (size_t)12 * np * e;
#pragma unroll
    for (int c = 0; c < 12; ++c) {
      Ge[c * np + t] = g[c];
      sA[c * np + t] = (double)g[c];
    }
  }
  __syncthreads();
  // ProlongToFaceLifting of the gradients into the side arrays (12 values per side point)
  for (int x = t; x < 6 * nf; x += NT) {
    const int ls = x / nf + 1, pq = x - (ls - 1) * nf;
    const int side = m.E2S[2 * (ls - 1) + 12 * e] - 1;
    const int flip = m.E2S[2 * (ls - 1) + 1 + 12 * e];
    const int* ab = m.s2v2 + 2 * (pq + nf * (flip + 5 * (ls - 1)));
    RF* dst = (flip == 0 ? Gm : Gs) + 12 * ((size_t)pq + (size_t)nf * side);
    if (GL) {
      const int id = vol_index<N>(ls, ab[0], ab[1], is_minus(ls) ? 0 : N);
#pragma unroll
      for (int c = 0; c < 12; ++c) dst[c] = (RF)sA[c * np + id];
    } else {
      const double* L = sB + (is_minus(ls) ? B.L_Minus() : B.L_Plus());
      const int id0 = vol_index<N>(ls, ab[0], ab[1], 0);
      RF out[12];
#pragma unroll
      for (int c = 0; c < 12; ++c) out[c] = (RF)sA[c * np + id0] * (RF)L[0];
      for (int l = 1; l <= N; ++l) {
        const int id = vol_index<N>(ls, ab[0], ab[1], l);
#pragma unroll
        for (int c = 0; c < 12; ++c) out[c] = out[c] + (RF)sA[c * np + id] * (RF)L[l];
      }
#pragma unroll
      for (int c = 0; c < 12; ++c) dst[c] = out[c];
    }
  }
}

// ---------------------------------------------------------------------------------------------
// FillFlux for inner sides (fillflux.f90:40-215): Riemann_Side + ViscousFlux_Side, times
// SurfElem. One lane per side point.
template <class RF, bool SPLIT>
__global__ __launch_bounds__(256) void k_flux(DevMesh m, const double* __restrict__ Um, const double* __restrict__ Us,
                                            const RF* __restrict__ Gm, const RF* __restrict__ Gs, Gas gas,
                                            int solver, double* __restrict__ Flux, double* __restrict__ FluxS,
                                            int first, int nPts) {
  const int fp = first + blockIdx.x * blockDim.x + threadIdx.x;
  if (fp >= nPts) return;
  const GasT<RF> g(gas);
  const Prim pl = cons_to_prim(Um + 5 * (size_t)fp, gas);
  const Prim pr = cons_to_prim(Us + 5 * (size_t)fp, gas);
  RF gl[12], gr[12];
#pragma unroll
  for (int c = 0; c < 12; ++c) {
    gl[c] = Gm[12 * (size_t)fp + c];
    gr[c] = Gs[12 * (size_t)fp + c];
  }
  RF fxl[4], fyl[4], fzl[4], fxr[4], fyr[4], fzr[4];
  diff_flux<RF>(g, (RF)pl.u, (RF)pl.v, (RF)pl.w, gl, gl + 4, gl + 8, fxl, fyl, fzl);
  diff_flux<RF>(g, (RF)pr.u, (RF)pr.v, (RF)pr.w, gr, gr + 4, gr + 8, fxr, fyr, fzr);
  // the side's flux with the master's face metric (FLEXI's Flux_master = Flux_slave) and, with
  // metrics computed by the library, once more with the slave element's own face metric
  const int nFlux = m.slaveNS ? 2 : 1;
  for (int side = 0; side < nFlux; ++side) {
    RF nv[3], t1[3], t2[3];
    double se;
    if (side == 0) {
#pragma unroll
      for (int c = 0; c < 3; ++c) {
        nv[c] = (RF)m.NormVec[3 * (size_t)fp + c];
        t1[c] = (RF)m.TangVec1[3 * (size_t)fp + c];
        t2[c] = (RF)m.TangVec2[3 * (size_t)fp + c];
      }
      se = m.SurfElem[fp];
    } else {
      const double* ns = m.slaveNS + 3 * (size_t)fp;
      se = sqrt(ns[0] * ns[0] + ns[1] * ns[1] + ns[2] * ns[2]);
      nv[0] = (RF)(ns[0] / se); nv[1] = (RF)(ns[1] / se); nv[2] = (RF)(ns[2] / se);
      // tangential basis: Gram-Schmidt of the coordinate axis least aligned with n
      const RF ax = fabs(nv[0]), ay = fabs(nv[1]), az = fabs(nv[2]);
      RF h[3] = {RF(0), RF(0), RF(0)};
      if (ax <= ay && ax <= az) h[0] = RF(1); else if (ay <= az) h[1] = RF(1); else h[2] = RF(1);
      const RF dt_ = h[0] * nv[0] + h[1] * nv[1] + h[2] * nv[2];
      RF tt[3] = {h[0] - dt_ * nv[0], h[1] - dt_ * nv[1], h[2] - dt_ * nv[2]};
      const RF tn = sqrt(tt[0] * tt[0] + tt[1] * tt[1] + tt[2] * tt[2]);
      t1[0] = tt[0] / tn; t1[1] = tt[1] / tn; t1[2] = tt[2] / tn;
      t2[0] = nv[1] * t1[2] - nv[2] * t1[1];
      t2[1] = nv[2] * t1[0] - nv[0] * t1[2];
      t2[2] = nv[0] * t1[1] - nv[1] * t1[0];
    }
    const RotState<RF> L = rotate<RF>((RF)pl.rho, (RF)Um[5 * (size_t)fp + 4], (RF)pl.p, (RF)pl.u, (RF)pl.v, (RF)pl.w,
                                      nv, t1, t2);
    const RotState<RF> R = rotate<RF>((RF)pr.rho, (RF)Us[5 * (size_t)fp + 4], (RF)pr.p, (RF)pr.u, (RF)pr.v, (RF)pr.w,
                                      nv, t1, t2);
    RF F[5];
    riemann_rot<RF>(g, solver, SPLIT, L, R, F);
    RF Fo[5];
    Fo[0] = F[0];
#pragma unroll
    for (int c = 0; c < 3; ++c) Fo[1 + c] = nv[c] * F[1] + t1[c] * F[2] + t2[c] * F[3];
    Fo[4] = F[4];
    // ViscousFlux_Side (riemann.f90:448-498): 0.5*(f_L + f_R).n
#pragma unroll
    for (int k = 0; k < 4; ++k)
      Fo[1 + k] = Fo[1 + k] + RF(0.5) * (nv[0] * (fxl[k] + fxr[k]) + nv[1] * (fyl[k] + fyr[k]) + nv[2] * (fzl[k] + fzr[k]));
    double* dst = side == 0 ? Flux : FluxS;
#pragma unroll
    for (int v = 0; v < 5; ++v) dst[5 * (size_t)fp + v] = (double)Fo[v] * se;
  }
}

// ---------------------------------------------------------------------------------------------
// Boundary sides 0..nBC-1 (FLEXI numbers them first). One lane per boundary side point; the inner
// state comes from the prolonged conservative variables as GetPrimitiveStateSurface does.
__device__ __forceinline__ int bc_point(const DevMesh& m, const double* __restrict__ Um, const Gas& gas, int fp,
                                        double* Pm, double* Pb, double* nv, double* t1, double* t2) {
  const Prim q = cons_to_prim(Um + 5 * (size_t)fp, gas);
  Pm[0] = q.rho; Pm[1] = q.u; Pm[2] = q.v; Pm[3] = q.w; Pm[4] = q.p; Pm[5] = q.T;
#pragma unroll
  for (int c = 0; c < 3; ++c) {
    nv[c] = m.NormVec[3 * (size_t)fp + c];
    t1[c] = m.TangVec1[3 * (size_t)fp + c];
    t2[c] = m.TangVec2[3 * (size_t)fp + c];
  }
  const int side = fp / ((m.N + 1) * (m.N + 1));
  const int type = m.bcInfo[2 * side], st = m.bcInfo[2 * side + 1];
  boundary_state(gas, type, st >= 0 ? m.refPrim + 6 * st : nullptr, m.bcData ? m.bcData + 6 * (size_t)fp : nullptr,
                 Pm, nv, t1, t2, Pb);
  return type;
}

// Lifting flux of the boundary sides (Lifting_FillFlux_BC, lifting_fillflux.t90:133-180) before
// SurfElem, for k_lift.
__global__ __launch_bounds__(256) void k_bc_lift(DevMesh m, const double* __restrict__ Um, Gas gas,
                                               double* __restrict__ bcLift, int nPts) {
  const int fp = blockIdx.x * blockDim.x + threadIdx.x;
  if (fp >= nPts) return;
  double Pm[6], Pb[6], nv[3], t1[3], t2[3], F[4];
  const int type = bc_point(m, Um, gas, fp, Pm, Pb, nv, t1, t2);
  boundary_lifting_flux(type, Pm, Pb, F);
#pragma unroll
  for (int v = 0; v < 4; ++v) bcLift[4 * (size_t)fp + v] = F[v];
}

// GetBoundaryFlux (getboundaryflux.f90:480-789) times SurfElem (fillflux.f90:192-199): Riemann
// solver against the boundary state plus ViscousFlux for Riemann-type sides; for walls the Euler
// flux (0, p n, 0) plus the viscous flux of the wall state with the inner gradients (slip walls:
// as slip_wall_gradients leaves them; no heat flux at an adiabatic wall).
template <class RF>
__global__ __launch_bounds__(256) void k_bc_flux(DevMesh m, const double* __restrict__ Um, const RF* __restrict__ Gm,
                                               Gas gas, int solver, int split, double* __restrict__ Flux, int nPts) {
  const int fp = blockIdx.x * blockDim.x + threadIdx.x;
  if (fp >= nPts) return;
  double Pm[6], Pb[6], nv[3], t1[3], t2[3], F[5], q[12];
  const int type = bc_point(m, Um, gas, fp, Pm, Pb, nv, t1, t2);
#pragma unroll
  for (int c = 0; c < 12; ++c) q[c] = (double)Gm[12 * (size_t)fp + c];
  const GasT<double> g(gas);
  if (bc_riemann_type(type)) {
    // both states' conservative variables from the primitive ones (PrimToCons), as GetBoundaryFlux
    const double Em = gas.sKappaM1 * Pm[4] + 0.5 * ((Pm[1] * Pm[0]) * Pm[1] + (Pm[2] * Pm[0]) * Pm[2] + (Pm[3] * Pm[0]) * Pm[3]);
    const double Eb = gas.sKappaM1 * Pb[4] + 0.5 * ((Pb[1] * Pb[0]) * Pb[1] + (Pb[2] * Pb[0]) * Pb[2] + (Pb[3] * Pb[0]) * Pb[3]);
    const RotState<double> L = rotate<double>(Pm[0], Em, Pm[4], Pm[1], Pm[2], Pm[3], nv, t1, t2);
    const RotState<double> R = rotate<double>(Pb[0], Eb, Pb[4], Pb[1], Pb[2], Pb[3], nv, t1, t2);
    double Fr[5];
    riemann_rot<double>(g, solver, split != 0, L, R, Fr);
    F[0] = Fr[0];
#pragma unroll
    for (int c = 0; c < 3; ++c) F[1 + c] = nv[c] * Fr[1] + t1[c] * Fr[2] + t2[c] * Fr[3];
    F[4] = Fr[4];
    double fxl[4], fyl[4], fzl[4], fxr[4], fyr[4], fzr[4];
    diff_flux<double>(g, Pm[1], Pm[2], Pm[3], q, q + 4, q + 8, fxl, fyl, fzl);
    diff_flux<double>(g, Pb[1], Pb[2], Pb[3], q, q + 4, q + 8, fxr, fyr, fzr);
#pragma unroll
    for (int k = 0; k < 4; ++k)
      F[1 + k] = F[1 + k] + 0.5 * (nv[0] * (fxl[k] + fxr[k]) + nv[1] * (fyl[k] + fyr[k]) + nv[2] * (fzl[k] + fzr[k]));
  } else {
    F[0] = 0.0;
#pragma unroll
    for (int d = 0; d < 3; ++d) F[1 + d] = Pb[4] * nv[d];
    F[4] = 0.0;
    if (bc_slip_wall(type)) slip_wall_gradients<double>(type, nv, t1, t2, q);
    double fx[4], fy[4], fz[4];
    diff_flux<double>(g, Pb[1], Pb[2], Pb[3], q, q + 4, q + 8, fx, fy, fz);
    if (type == BC_WALL_ADIABATIC) fx[3] = fy[3] = fz[3] = 0.0;
#pragma unroll
    for (int k = 0; k < 4; ++k) F[1 + k] = F[1 + k] + nv[0] * fx[k] + nv[1] * fy[k] + nv[2] * fz[k];
  }
  const double se = m.SurfElem[fp];
#pragma unroll
  for (int v = 0; v < 5; ++v) Flux[5 * (size_t)fp + v] = F[v] * se;
}

// ---------------------------------------------------------------------------------------------
// Volume integral (VolInt_weakForm volint.f90:117-187, or VolInt_splitForm :211-371 with the
// Pirozzoli flux, splitflux.f90:603-666), surface integral (surfint.t90), Ut=-Ut, Jacobian
// (dg.f90:656-676), and optionally the LSERK update of TimeStepByLSERKW2 (timestep.f90:124-131).
//
// The split flux is evaluated in contravariant form: with m = (Ja_i + Ja_l)/2 and the sums
// rho_s, v_s, p_s, H_s of the two states, F = (rho_s m.v_s/2, rho_s m.v_s v_s/4 + p_s m,
// rho_s m.v_s H_s/4), algebraically identical to the reference's m.(f,g,h).
template <int N, bool GL, bool SPLIT, class RF>
__global__ __launch_bounds__(Dim<N>::NT) void k_volint(const double* __restrict__ U, DevMesh m,
                                                     const RF* __restrict__ G, const double* __restrict__ Flux,
                                                     const double* __restrict__ FluxS, Gas gas,
                                                     double* __restrict__ Ut, RKStage st) {
  constexpr int n = Dim<N>::n, nf = Dim<N>::nf, np = Dim<N>::np, NT = Dim<N>::NT;
  constexpr int NPRIM = SPLIT ? 6 : 5;   // split: rho,u,v,w,p,H ; weak: staging of f~ (5 comps)
  constexpr int NDIR = SPLIT ? 7 : 0;    // split: Ja^d (3) + viscous contravariant flux (4)
  __shared__ double sA[(NPRIM + NDIR) * np > 5 * np ? (NPRIM + NDIR) * np : 5 * np];
  __shared__ double sB[3 * n * n + 4 * n];
  const BasisLayout B{n};
  const GasT<RF> g(gas);
  const int e = blockIdx.x;
  const int t = threadIdx.x;
  const double* Ue = U + (size_t)5 * np * e;
  stage_u_basis<N, NT>(Ue, sA, m.basis, sB, t);
  __syncthreads();
  const bool act = t < np;
  const int i = t % n, j = (t / n) % n, k = t / (n * n);
  double Uown[5] = {0, 0, 0, 0, 0};
  Prim q{};
  if (act) {
#pragma unroll
    for (int v = 0; v < 5; ++v) Uown[v] = sA[5 * t + v];
    q = cons_to_prim(Uown, gas);
  }
  // viscous Cartesian fluxes at the own node (EvalDiffFlux3D_Point)
  RF fvx[4] = {0, 0, 0, 0}, fvy[4] = {0, 0, 0, 0}, fvz[4] = {0, 0, 0, 0};
  RF Ja[9];
  const size_t gid = (size_t)np * e + (act ? t : 0);
  if (act) {
    RF gg[12];
    const RF* Ge = G + (size_t)12 * np * e;
#pragma unroll
    for (int c = 0; c < 12; ++c) gg[c] = Ge[c * np + t];
    diff_flux<RF>(g, (RF)q.u, (RF)q.v, (RF)q.w, gg, gg + 4, gg + 8, fvx, fvy, fvz);
#pragma unroll
    for (int c = 0; c < 3; ++c) {
      Ja[c] = (RF)m.Mf[3 * gid + c];
      Ja[3 + c] = (RF)m.Mg[3 * gid + c];
      Ja[6 + c] = (RF)m.Mh[3 * gid + c];
    }
  }
  __syncthreads();  // sA (U staging) is reused below
  RF acc[5] = {0, 0, 0, 0, 0};
  if (SPLIT) {
    double* sR = sA;           // rho
    double* sUu = sA + np;     // u
    double* sV = sA + 2 * np;  // v
    double* sW = sA + 3 * np;  // w
    double* sPp = sA + 4 * np; // p
    double* sH = sA + 5 * np;  // H = (E+p)/rho
    double* sM = sA + 6 * np;  // Ja^d (3 x np)
    double* sF = sA + 9 * np;  // viscous contravariant flux (4 x np)
    if (act) {
      sR[t] = q.rho; sUu[t] = q.u; sV[t] = q.v; sW[t] = q.w; sPp[t] = q.p;
      sH[t] = (Uown[4] + q.p) / Uown[0];
    }
    const RF rho = (RF)q.rho, uu = (RF)q.u, vv = (RF)q.v, ww = (RF)q.w, pp = (RF)q.p;
    const RF HH = act ? (RF)((Uown[4] + q.p) / Uown[0]) : RF(0);
    const double* DV = sB + B.DVolSurf();
    const double* DH = sB + B.D_Hat_T();
#pragma unroll 1
    for (int d = 0; d < 3; ++d) {
      if (act) {
        const RF* Jd = Ja + 3 * d;
        sM[t] = (double)Jd[0]; sM[np + t] = (double)Jd[1]; sM[2 * np + t] = (double)Jd[2];
#pragma unroll
        for (int c = 0; c < 4; ++c) sF[c * np + t] = (double)(Jd[0] * fvx[c] + Jd[1] * fvy[c] + Jd[2] * fvz[c]);
      }
      __syncthreads();
      if (act) {
        const int me = d == 0 ? i : (d == 1 ? j : k);
        const int stride = d == 0 ? 1 : (d == 1 ? n : n * n);
        const int base = t - me * stride;
        const RF m0 = Ja[3 * d], m1 = Ja[3 * d + 1], m2 = Ja[3 * d + 2];
#pragma unroll
        for (int l = 0; l <= N; ++l) {
          const int id = base + l * stride;
          const RF dh = (RF)DH[l + n * me];
#pragma unroll
          for (int c = 0; c < 4; ++c) acc[1 + c] = acc[1 + c] + dh * (RF)sF[c * np + id];
          if (l == me) continue;
          const RF rs = rho + (RF)sR[id];
          const RF us = uu + (RF)sUu[id], vs = vv + (RF)sV[id], ws = ww + (RF)sW[id];
          const RF ps = pp + (RF)sPp[id], Hs = HH + (RF)sH[id];
          const RF a0 = RF(0.5) * (m0 + (RF)sM[id]), a1 = RF(0.5) * (m1 + (RF)sM[np + id]),
                   a2 = RF(0.5) * (m2 + (RF)sM[2 * np + id]);
          const RF mv = a0 * us + a1 * vs + a2 * ws;
          const RF qq = RF(0.25) * rs * mv;
          const RF dv = (RF)DV[l + n * me];
          acc[0] = acc[0] + dv * (RF(0.5) * rs * mv);
          acc[1] = acc[1] + dv * (qq * us + a0 * ps);
          acc[2] = acc[2] + dv * (qq * vs + a1 * ps);
          acc[3] = acc[3] + dv * (qq * ws + a2 * ps);
          acc[4] = acc[4] + dv * (qq * Hs);
        }
      }
      __syncthreads();
    }
  } else {
    // weak form: per direction stage the contravariant flux Ja^d.(f_euler + f_visc) (5 comps)
    const double* DH = sB + B.D_Hat_T();
    RF fe[5], ge[5], he[5];
    if (act) {
      const RF Ep = (RF)(Uown[4] + q.p);
      const RF m1 = (RF)Uown[1], m2 = (RF)Uown[2], m3 = (RF)Uown[3];
      const RF uu = (RF)q.u, vv = (RF)q.v, ww = (RF)q.w, pp = (RF)q.p;
      fe[0] = m1; fe[1] = m1 * uu + pp; fe[2] = m1 * vv; fe[3] = m1 * ww; fe[4] = Ep * uu;
      ge[0] = m2; ge[1] = fe[2]; ge[2] = m2 * vv + pp; ge[3] = m2 * ww; ge[4] = Ep * vv;
      he[0] = m3; he[1] = fe[3]; he[2] = ge[3]; he[3] = m3 * ww + pp; he[4] = Ep * ww;
#pragma unroll
      for (int c = 0; c < 4; ++c) {
        fe[1 + c] = fe[1 + c] + fvx[c];
        ge[1 + c] = ge[1 + c] + fvy[c];
        he[1 + c] = he[1 + c] + fvz[c];
      }
    }
#pragma unroll 1
    for (int d = 0; d < 3; ++d) {
      if (act) {
        const RF* Jd = Ja + 3 * d;
#pragma unroll
        for (int v = 0; v < 5; ++v) sA[v * np + t] = (double)(fe[v] * Jd[0] + ge[v] * Jd[1] + he[v] * Jd[2]);
      }
      __syncthreads();
      if (act) {
        const int me = d == 0 ? i : (d == 1 ? j : k);
        const int stride = d == 0 ? 1 : (d == 1 ? n : n * n);
        const int base = t - me * stride;
#pragma unroll
        for (int l = 0; l <= N; ++l) {
          const RF dh = (RF)DH[l + n * me];
          const int id = base + l * stride;
#pragma unroll
          for (int v = 0; v < 5; ++v) acc[v] = acc[v] + dh * (RF)sA[v * np + id];
        }
      }
      __syncthreads();
    }
  }
  if (!act) return;  // no barrier follows
  // surface integral: master sides +Flux, slave sides -Flux (surfint.t90:21-239, 471-616)
#pragma unroll 1
  for (int ls = 1; ls <= 6; ++ls) {
    int a, b, l;
    face_coords(ls, i, j, k, a, b, l);
    if (GL && l != (is_minus(ls) ? 0 : N)) continue;
    const int side = m.E2S[2 * (ls - 1) + 12 * e] - 1;
    const int flip = m.E2S[2 * (ls - 1) + 1 + 12 * e];
    const int* pq2 = m.v2s2 + 2 * ((a + n * b) + nf * (flip + 5 * (ls - 1)));
    const size_t fp = (size_t)(pq2[0] + n * pq2[1]) + (size_t)nf * side;
    const double lh = GL ? (is_minus(ls) ? sB[B.L_HatMinus()] : sB[B.L_HatPlus() + N])
                         : (is_minus(ls) ? sB[B.L_HatMinus() + l] : sB[B.L_HatPlus() + l]);
    const RF sl = (RF)(flip == 0 ? lh : -lh);
    const double* Fs = (flip != 0 && FluxS) ? FluxS : Flux;
#pragma unroll
    for (int v = 0; v < 5; ++v) acc[v] = acc[v] + (RF)Fs[5 * fp + v] * sl;
  }
  const double sj = m.sJ[gid];
  double ut[5];
#pragma unroll
  for (int v = 0; v < 5; ++v) ut[v] = (-(double)acc[v]) * sj;
  const size_t o = (size_t)5 * gid;
  if (st.mode == 0) {
#pragma unroll
    for (int v = 0; v < 5; ++v) Ut[o + v] = ut[v];
  } else {
    // VCopy / VAXPBY of TimeStepByLSERKW2 (vector.f90:74-125)
#pragma unroll
    for (int v = 0; v < 5; ++v) {
      const double r = (st.mode == 1) ? ut[v] : st.R[o + v] * st.minusA + ut[v];
      st.R[o + v] = r;
      st.U[o + v] = Uown[v] + r * st.b_dt;
    }
  }
}

// ---------------------------------------------------------------------------------------------
// CALCTIMESTEP (calctimestep.f90:89-227): per element the maxima of the convective and viscous
// eigenvalue sums, then dtElem; a second kernel reduces over elements.
template <int N>
__global__ __launch_bounds__(Dim<N>::NT) void k_dt_elem(const double* __restrict__ U, DevMesh m, Gas gas, double CFL,
                                                      double DFL, double* __restrict__ dtConv,
                                                      double* __restrict__ dtVisc, int* __restrict__ err) {
  constexpr int np = Dim<N>::np, NT = Dim<N>::NT;
  __shared__ double red[6 * (NT / 64)];
  __shared__ int nanflag[NT / 64];
  const int e = blockIdx.x, t = threadIdx.x;
  double lam[6] = {0, 0, 0, 0, 0, 0};
  int isnan_rho = 0;
  if (t < np) {
    const size_t gid = (size_t)np * e + t;
    dt_node_lambda(U + 5 * gid, gas, m.sJ[gid], m.Mf + 3 * gid, m.Mg + 3 * gid, m.Mh + 3 * gid, lam, isnan_rho);
  }
  double mx[6];
  int nanr = 0;
  block_lambda_max<NT>(lam, isnan_rho, red, nanflag, t, mx, nanr);
  if (t == 0) {
    const double dc = CFL * 2. / (mx[0] + mx[1] + mx[2]);
    const double sv = mx[3] + mx[4] + mx[5];
    const double dv = (sv > 0.0) ? DFL * 4. / sv : 1.79769313486231571e308;
    dtConv[e] = dc;
    dtVisc[e] = dv;
    err[e] = nanr ? 1 : (dc != dc ? 2 : (dv != dv ? 3 : 0));
  }
}

// Per-element steps from the spectral-radius sums the fused last LSERK stage stored
// (lamSum[2e]: convective, lamSum[2e+1]: viscous), as k_dt_elem finishes them.
__global__ __launch_bounds__(256) void k_dt_finish(const double* __restrict__ lamSum, const int* __restrict__ lamNan,
                                                 int K, double CFL, double DFL, double* __restrict__ dtConv,
                                                 double* __restrict__ dtVisc, int* __restrict__ err) {
  const int e = blockIdx.x * blockDim.x + threadIdx.x;
  if (e >= K) return;
  const double dc = CFL * 2. / lamSum[2 * e];
  const double sv = lamSum[2 * e + 1];
  const double dv = (sv > 0.0) ? DFL * 4. / sv : 1.79769313486231571e308;
  dtConv[e] = dc;
  dtVisc[e] = dv;
  err[e] = lamNan[e] ? 1 : (dc != dc ? 2 : (dv != dv ? 3 : 0));
}

// Minimum over elements of the convective and viscous steps and the error code.
__global__ __launch_bounds__(1024) void k_dt_reduce(const double* __restrict__ dtConv, const double* __restrict__ dtVisc,
                                                  const int* __restrict__ err, int K, double* __restrict__ out,
                                                  int* __restrict__ errOut) {
  __shared__ double sc[16], sv[16];
  __shared__ int se[16];
  const int t = threadIdx.x;
  double c = 1.79769313486231571e308, v = 1.79769313486231571e308;
  int ec = 0;
  for (int x = t; x < K; x += 1024) {
    c = fmin(c, dtConv[x]);
    v = fmin(v, dtVisc[x]);
    const int ex = err[x];
    // priority to the root cause: 1 (NaN density) over 2 (convective) over 3 (viscous)
    if (ex != 0 && (ec == 0 || ex < ec)) ec = ex;
  }
  for (int off = 32; off > 0; off >>= 1) {
    c = fmin(c, __shfl_down(c, off, 64));
    v = fmin(v, __shfl_down(v, off, 64));
    const int eo = __shfl_down(ec, off, 64);
    if (eo != 0 && (ec == 0 || eo < ec)) ec = eo;
  }
  if ((t & 63) == 0) { sc[t >> 6] = c; sv[t >> 6] = v; se[t >> 6] = ec; }
  __syncthreads();
  if (t == 0) {
    for (int w = 1; w < 16; ++w) {
      c = fmin(c, sc[w]);
      v = fmin(v, sv[w]);
      if (se[w] != 0 && (ec == 0 || se[w] < ec)) ec = se[w];
    }
    out[0] = c;
    out[1] = v;
    errOut[0] = ec;
  }
}

template <class RF>
__global__ __launch_bounds__(256) void k_unpack_grad(const RF* __restrict__ G, int np, int nNodes, double* __restrict__ gx,
                                                   double* __restrict__ gy, double* __restrict__ gz) {
  const int x = blockIdx.x * blockDim.x + threadIdx.x;
  if (x >= nNodes) return;
  const int e = x / np, t = x - e * np;
  const RF* Ge = G + (size_t)12 * np * e;
  for (int v = 0; v < 4; ++v) {
    gx[4 * (size_t)x + v] = (double)Ge[v * np + t];
    gy[4 * (size_t)x + v] = (double)Ge[(4 + v) * np + t];
    gz[4 * (size_t)x + v] = (double)Ge[(8 + v) * np + t];
  }
}

// ---------------------------------------------------------------------------------------------
// Lifted density gradient, the first lifting variable of FLEXI's default build (PP_OPTLIFT=0,
// LIFT_DENS, idealgas/eos.h:108-117): the BR1 lifting of k_lift applied to rho (strong form,
// lifting_br1.t90:48-181) with the boundary values of Lifting_GetBoundaryFlux
// (getboundaryflux.f90:855-939): Dirichlet 0.5*(rho_in + rho_b), no-slip walls the wall state's
// density, slip walls the inner density. The solver itself does not need it (its viscous fluxes
// use only the velocity and temperature gradients); FP64 in both precisions.
// rhoG(np, 3, nElems): d/dx, d/dy, d/dz of rho at each node.
template <int N, bool GL>
__global__ __launch_bounds__(Dim<N>::NT) void k_lift_density(const double* __restrict__ U, DevMesh m,
                                                           const double* __restrict__ Um,
                                                           const double* __restrict__ Us, Gas gas,
                                                           double* __restrict__ rhoG) {
  constexpr int n = Dim<N>::n, nf = Dim<N>::nf, np = Dim<N>::np, NT = Dim<N>::NT;
  __shared__ double sR[np];
  __shared__ double sB[3 * n * n + 4 * n];
  const int e = blockIdx.x;
  const int t = threadIdx.x;
  if (t < np) sR[t] = U[(size_t)5 * ((size_t)np * e + t)];
  load_basis<N>(m.basis, sB, t, NT);
  __syncthreads();
  if (t >= np) return;
  const BasisLayout B{n};
  const int i = t % n, j = (t / n) % n, k = t / (n * n);
  const double* DT = sB + B.D_T();
  double gxi = DT[0 + n * i] * sR[0 + n * (j + n * k)];
  double geta = DT[0 + n * j] * sR[i + n * (0 + n * k)];
  double gzeta = DT[0 + n * k] * sR[i + n * (j + n * 0)];
  for (int l = 1; l <= N; ++l) {
    gxi = gxi + DT[l + n * i] * sR[l + n * (j + n * k)];
    geta = geta + DT[l + n * j] * sR[i + n * (l + n * k)];
    gzeta = gzeta + DT[l + n * k] * sR[i + n * (j + n * l)];
  }
  const size_t gid = (size_t)np * e + t;
  double g[3];
#pragma unroll
  for (int d = 0; d < 3; ++d) g[d] = m.Mf[3 * gid + d] * gxi + m.Mg[3 * gid + d] * geta + m.Mh[3 * gid + d] * gzeta;
#pragma unroll 1
  for (int ls = 1; ls <= 6; ++ls) {
    int a, b, l;
    face_coords(ls, i, j, k, a, b, l);
    if (GL && l != (is_minus(ls) ? 0 : N)) continue;
    const int side = m.E2S[2 * (ls - 1) + 12 * e] - 1;
    const int flip = m.E2S[2 * (ls - 1) + 1 + 12 * e];
    const int* pq2 = m.v2s2 + 2 * ((a + n * b) + nf * (flip + 5 * (ls - 1)));
    const int fp = pq2[0] + n * pq2[1] + nf * side;
    const double se = m.SurfElem[fp];
    const double lh = GL ? (is_minus(ls) ? sB[B.L_HatMinus()] : sB[B.L_HatPlus() + N])
                         : (is_minus(ls) ? sB[B.L_HatMinus() + l] : sB[B.L_HatPlus() + l]);
    double f;
    if (side < m.nBC) {
      double Pm[6], Pb[6], nv[3], t1[3], t2[3];
      const int type = bc_point(m, Um, gas, fp, Pm, Pb, nv, t1, t2);
      const double rb = bc_riemann_type(type) ? 0.5 * (Pm[0] + Pb[0]) : (bc_slip_wall(type) ? Pm[0] : Pb[0]);
      f = (rb - Pm[0]) * se;
    } else {
      f = 0.5 * (Us[5 * (size_t)fp] - Um[5 * (size_t)fp]) * se;
    }
#pragma unroll
    for (int d = 0; d < 3; ++d) g[d] = g[d] + f * m.NormVec[3 * (size_t)fp + d] * lh;
  }
  const double sj = m.sJ[gid];
#pragma unroll
  for (int d = 0; d < 3; ++d) rhoG[t + (size_t)np * (d + 3 * (size_t)e)] = g[d] * sj;
}

// One direction of the lifted gradients in FLEXI's layout gradU{x,y,z}(nVar,0:N,0:N,0:N,nElems):
// nVar = 4 (u,v,w,T; PP_OPTLIFT=1) or 5 (rho first, from rhoG; PP_OPTLIFT=0).
template <class RF>
__global__ __launch_bounds__(256) void k_unpack_grad_dir(const RF* __restrict__ G, const double* __restrict__ rhoG,
                                                        int np, int nNodes, int nVar, int dir,
                                                        double* __restrict__ out) {
  const int x = blockIdx.x * blockDim.x + threadIdx.x;
  if (x >= nNodes) return;
  const int e = x / np, t = x - e * np;
  const RF* Ge = G + (size_t)12 * np * e;
  double* o = out + (size_t)nVar * x;
  if (nVar == 5) *o++ = rhoG[t + (size_t)np * (dir + 3 * (size_t)e)];
#pragma unroll
  for (int v = 0; v < 4; ++v) o[v] = (double)Ge[(4 * dir + v) * np + t];
}

// ---------------------------------------------------------------------------------------------
// Dispatch on the compile-time degree.

struct TimedScope {
  KernelTimer* t;
  int id;
  hipStream_t s;
  TimedScope(KernelTimer* t_, int id_, hipStream_t s_) : t(t_), id(id_), s(s_) {
    if (t) t->mark(id, true, s);
  }
  ~TimedScope() {
    if (t) t->mark(id, false, s);
  }
};

template <int N, bool GL, class RF>
static hipError_t rhs_n(const DevMesh& m, const DevWork& w, const Gas& gas, int split, int riemann,
                        const double* U, double* Ut, const RKStage& st, hipStream_t s, KernelTimer* tm) {
  const int K = m.nElems;
  constexpr int NT = Dim<N>::NT;
  {
    TimedScope ts(tm, K_PROLONG, s);
    hipLaunchKernelGGL((k_prolong_u<N, GL>), dim3(K), dim3(NT), 0, s, U, m, w.Uf_m, w.Uf_s);
  }
  const int nBCPts = m.nBC * Dim<N>::nf;
  {
    TimedScope ts(tm, K_LIFT, s);
    if (nBCPts > 0)
      hipLaunchKernelGGL(k_bc_lift, dim3((nBCPts + 255) / 256), dim3(256), 0, s, m, w.Uf_m, gas, w.bcLift, nBCPts);
    hipLaunchKernelGGL((k_lift<N, GL, RF>), dim3(K), dim3(NT), 0, s, U, m, w.Uf_m, w.Uf_s, w.bcLift, gas, (RF*)w.G,
                       (RF*)w.Gf_m, (RF*)w.Gf_s);
  }
  const int nPts = m.nSides * Dim<N>::nf, nIn = nPts - nBCPts;
  {
    TimedScope ts(tm, K_FLUX, s);
    if (nIn > 0) {
      if (split)
        hipLaunchKernelGGL((k_flux<RF, true>), dim3((nIn + 255) / 256), dim3(256), 0, s, m, w.Uf_m, w.Uf_s,
                           (const RF*)w.Gf_m, (const RF*)w.Gf_s, gas, riemann, w.Flux, w.FluxS, nBCPts, nPts);
      else
        hipLaunchKernelGGL((k_flux<RF, false>), dim3((nIn + 255) / 256), dim3(256), 0, s, m, w.Uf_m, w.Uf_s,
                           (const RF*)w.Gf_m, (const RF*)w.Gf_s, gas, riemann, w.Flux, w.FluxS, nBCPts, nPts);
    }
    if (nBCPts > 0)
      hipLaunchKernelGGL((k_bc_flux<RF>), dim3((nBCPts + 255) / 256), dim3(256), 0, s, m, w.Uf_m, (const RF*)w.Gf_m,
                         gas, riemann, split, w.Flux, nBCPts);
  }
  {
    TimedScope ts(tm, K_VOLINT, s);
    if (split)
      hipLaunchKernelGGL((k_volint<N, GL, true, RF>), dim3(K), dim3(NT), 0, s, U, m, (const RF*)w.G, w.Flux, w.FluxS,
                         gas, Ut, st);
    else
      hipLaunchKernelGGL((k_volint<N, GL, false, RF>), dim3(K), dim3(NT), 0, s, U, m, (const RF*)w.G, w.Flux,
                         w.FluxS, gas, Ut, st);
  }
  return hipGetLastError();
}

template <int N>
static hipError_t rhs_dispatch(const DevMesh& m, const DevWork& w, const Gas& gas, int nodeType, int split,
                               int riemann, int fp32, const double* U, double* Ut, const RKStage& st, hipStream_t s,
                               KernelTimer* tm) {
  if (nodeType == 2) {
    if (fp32) return rhs_n<N, true, float>(m, w, gas, split, riemann, U, Ut, st, s, tm);
    return rhs_n<N, true, double>(m, w, gas, split, riemann, U, Ut, st, s, tm);
  }
  if (split) return hipErrorInvalidValue;  // Gauss split form (entropy variables) is not built
  if (fp32) return rhs_n<N, false, float>(m, w, gas, split, riemann, U, Ut, st, s, tm);
  return rhs_n<N, false, double>(m, w, gas, split, riemann, U, Ut, st, s, tm);
}

bool supported_degree(int N) { return N >= 1 && N <= 7; }

hipError_t launch_rhs(const DevMesh& m, const DevWork& w, const Gas& gas, int nodeType, int split, int riemann,
                      int fp32, const double* U, double* Ut, const RKStage& st, hipStream_t s, KernelTimer* tm) {
  switch (m.N) {
    case 1: return rhs_dispatch<1>(m, w, gas, nodeType, split, riemann, fp32, U, Ut, st, s, tm);
    case 2: return rhs_dispatch<2>(m, w, gas, nodeType, split, riemann, fp32, U, Ut, st, s, tm);
    case 3: return rhs_dispatch<3>(m, w, gas, nodeType, split, riemann, fp32, U, Ut, st, s, tm);
    case 4: return rhs_dispatch<4>(m, w, gas, nodeType, split, riemann, fp32, U, Ut, st, s, tm);
    case 5: return rhs_dispatch<5>(m, w, gas, nodeType, split, riemann, fp32, U, Ut, st, s, tm);
    case 6: return rhs_dispatch<6>(m, w, gas, nodeType, split, riemann, fp32, U, Ut, st, s, tm);
    case 7: return rhs_dispatch<7>(m, w, gas, nodeType, split, riemann, fp32, U, Ut, st, s, tm);
    default: return hipErrorInvalidValue;
  }
}

template <int N>
static hipError_t dt_n(const DevMesh& m, const DevWork& w, const Gas& gas, const double* U, double CFL, double DFL,
                       hipStream_t s, KernelTimer* tm) {
  const int K = m.nElems;
  double* dtConv = w.dtElem;
  double* dtVisc = w.dtElem + K;
  {
    TimedScope ts(tm, K_DT_ELEM, s);
    hipLaunchKernelGGL((k_dt_elem<N>), dim3(K), dim3(Dim<N>::NT), 0, s, U, m, gas, CFL, DFL, dtConv, dtVisc,
                       w.errElem);
  }
  {
    TimedScope ts(tm, K_DT_REDUCE, s);
    hipLaunchKernelGGL(k_dt_reduce, dim3(1), dim3(1024), 0, s, dtConv, dtVisc, w.errElem, K, w.dtOut, w.errOut);
  }
  return hipGetLastError();
}

hipError_t launch_calc_timestep_from_sums(const DevMesh& m, const DevWork& w, const double* lamSum, const int* lamNan,
                                          double CFL, double DFL, hipStream_t s, KernelTimer* tm) {
  const int K = m.nElems;
  double* dtConv = w.dtElem;
  double* dtVisc = w.dtElem + K;
  {
    TimedScope ts(tm, K_DT_ELEM, s);
    hipLaunchKernelGGL(k_dt_finish, dim3((K + 255) / 256), dim3(256), 0, s, lamSum, lamNan, K, CFL, DFL, dtConv, dtVisc,
                       w.errElem);
  }
  {
    TimedScope ts(tm, K_DT_REDUCE, s);
    hipLaunchKernelGGL(k_dt_reduce, dim3(1), dim3(1024), 0, s, dtConv, dtVisc, w.errElem, K, w.dtOut, w.errOut);
  }
  return hipGetLastError();
}

hipError_t launch_calc_timestep(const DevMesh& m, const DevWork& w, const Gas& gas, const double* U, double CFL,
                                double DFL, hipStream_t s, KernelTimer* tm) {
  switch (m.N) {
    case 1: return dt_n<1>(m, w, gas, U, CFL, DFL, s, tm);
    case 2: return dt_n<2>(m, w, gas, U, CFL, DFL, s, tm);
    case 3: return dt_n<3>(m, w, gas, U, CFL, DFL, s, tm);
    case 4: return dt_n<4>(m, w, gas, U, CFL, DFL, s, tm);
    case 5: return dt_n<5>(m, w, gas, U, CFL, DFL, s, tm);
    case 6: return dt_n<6>(m, w, gas, U, CFL, DFL, s, tm);
    case 7: return dt_n<7>(m, w, gas, U, CFL, DFL, s, tm);
    default: return hipErrorInvalidValue;
  }
}


hipError_t launch_lift_density(const DevMesh& m, const DevWork& w, const Gas& gas, int nodeType, const double* U,
                               double* rhoG, hipStream_t s) {
  const bool gl = nodeType == 2;
  switch (m.N) {
#define FXG_CASE(NN)                                                                                               \
  case NN:                                                                                                         \
    if (gl) {                                                                                                      \
      hipLaunchKernelGGL((k_prolong_u<NN, true>), dim3(m.nElems), dim3(Dim<NN>::NT), 0, s, U, m, w.Uf_m, w.Uf_s);   \
      hipLaunchKernelGGL((k_lift_density<NN, true>), dim3(m.nElems), dim3(Dim<NN>::NT), 0, s, U, m, w.Uf_m, w.Uf_s, \
                         gas, rhoG);                                                                               \
    } else {                                                                                                       \
      hipLaunchKernelGGL((k_prolong_u<NN, false>), dim3(m.nElems), dim3(Dim<NN>::NT), 0, s, U, m, w.Uf_m, w.Uf_s);  \
      hipLaunchKernelGGL((k_lift_density<NN, false>), dim3(m.nElems), dim3(Dim<NN>::NT), 0, s, U, m, w.Uf_m,        \
                         w.Uf_s, gas, rhoG);                                                                       \
    }                                                                                                              \
    break;
    FXG_CASE(1) FXG_CASE(2) FXG_CASE(3) FXG_CASE(4) FXG_CASE(5) FXG_CASE(6) FXG_CASE(7)
#undef FXG_CASE
    default: return hipErrorInvalidValue;
  }
  return hipGetLastError();
}

hipError_t launch_unpack_gradient_dir(const DevMesh& m, const void* G, int fp32, const double* rhoG, int nVar, int dir,
                                      double* out, hipStream_t s) {
  const int n = m.N + 1, np = n * n * n;
  const int nNodes = np * m.nElems;
  if (fp32)
    hipLaunchKernelGGL(k_unpack_grad_dir<float>, dim3((nNodes + 255) / 256), dim3(256), 0, s, (const float*)G, rhoG,
                       np, nNodes, nVar, dir, out);
  else
    hipLaunchKernelGGL(k_unpack_grad_dir<double>, dim3((nNodes + 255) / 256), dim3(256), 0, s, (const double*)G, rhoG,
                       np, nNodes, nVar, dir, out);
  return hipGetLastError();
}

hipError_t launch_unpack_gradients(const DevMesh& m, const void* G, int fp32, double* gx, double* gy, double* gz,
                                   hipStream_t s) {
  const int n = m.N + 1, np = n * n * n;
  const int nNodes = np * m.nElems;
  if (fp32)
    hipLaunchKernelGGL(k_unpack_grad<float>, dim3((nNodes + 255) / 256), dim3(256), 0, s, (const float*)G, np, nNodes,
                       gx, gy, gz);
  else
    hipLaunchKernelGGL(k_unpack_grad<double>, dim3((nNodes + 255) / 256), dim3(256), 0, s, (const double*)G, np, nNodes,
                       gx, gy, gz);
  return hipGetLastError();
}

}  // namespace fxg
