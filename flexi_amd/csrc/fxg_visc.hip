// Kernel A of the fused Gauss-Lobatto split-form stage (the rest is in fxg_fused.hip): for each
// element, the BR1 gradients at its boundary nodes and, for each of its six faces, the viscous
// normal flux f_v . (n SurfElem) seen from that element (4 values per face point, momentum and
// energy) into the side array FG (lifting_br1.t90:48-181, flux.f90:149-241). A translation unit
// of its own, so that its code does not change the register allocation of k_dg_fused.
#include "fxg_bc.h"
#include "fxg_elem.h"

namespace fxg {
namespace {

// FP64 stages the element's U in LDS with contiguous lanes (held to 80 VGPRs at N=7, 52 KB:
// three workgroups per CU); FP32-bulk converts U from global memory directly and is held to 64
// VGPRs (18 KB: four workgroups per CU, 0.42 ms instead of 0.53 ms at 32^3, N=7). On meshes with
// boundary sides (BC) it also keeps the gradients of the boundary face points for k_bc_visc.
template <int N, class RF, bool BC>
__global__ __launch_bounds__(D<N>::NT)
__attribute__((amdgpu_waves_per_eu(sizeof(RF) == 4 ? (N == 7 ? 8 : 6) : (N == 7 ? 6 : 1)))) void k_visc_faces(
    const double* __restrict__ U, DevMesh m, const double* __restrict__ Jsoa, const double* __restrict__ FU,
    RF* __restrict__ FG, Gas gas, RF* __restrict__ gBC) {
  constexpr int n = D<N>::n, np = D<N>::np, NT = D<N>::NT, NF = D<N>::NF;
  constexpr bool raw = sizeof(RF) == 8;
  __shared__ double sRaw[raw ? 5 * np : 1];  // the element's U as stored, loaded with contiguous lanes
  __shared__ RF sP[4 * np];        // u, v, w, T
  __shared__ RF sNb[4 * NF];       // neighbour u, v, w, T per own face point
  __shared__ unsigned sIdx[NF];
  __shared__ double sB[3 * n * n + 4 * n];
  const int e = blockIdx.x, t = threadIdx.x;
  const double* Ue = U + (size_t)5 * np * e;
  constexpr int NB = Bnd<N>::NB;         // boundary nodes
  constexpr int NBAS = 3 * n * n + 4 * n;
  static_assert(NBAS <= NT, "one basis entry per lane");
  constexpr int RS = (NF + NT - 1) / NT;  // own face points per lane
  constexpr int RU = raw ? (5 * np + NT - 1) / NT : 5;
  // Every global load of the staging is issued before the first wait, instead of one wait per
  // loop of loads: the side slots first (the neighbour states depend on them), then U (FP64: the
  // element's values with contiguous lanes; FP32-bulk: the lane's node) and the basis.
  unsigned own[RS];
#pragma unroll
  for (int r = 0; r < RS; ++r) {
    const int x = t + r * NT;
    own[r] = x < NF ? face_slot_tab<N>(m, e, x) : 0u;
  }
  double ul[RU];
#pragma unroll
  for (int r = 0; r < RU; ++r) {
    if constexpr (raw) {
      const int x = t + r * NT;
      ul[r] = x < 5 * np ? Ue[x] : 0.0;
    } else {
      ul[r] = t < np ? Ue[5 * t + r] : 1.0;
    }
  }
  const double bas = t < NBAS ? m.basis[t] : 0.0;
  double cn[RS][5];
#pragma unroll
  for (int r = 0; r < RS; ++r) {
    const double* c = FU + 5 * (size_t)slot_nbr<N>(own[r]);
#pragma unroll
    for (int v = 0; v < 5; ++v) cn[r][v] = t + r * NT < NF ? c[v] : 1.0;
  }
  if constexpr (raw) {
#pragma unroll
    for (int r = 0; r < RU; ++r)
      if (t + r * NT < 5 * np) sRaw[t + r * NT] = ul[r];
  } else if (t < np) {
    const Prim q = cons_to_prim(ul, gas);
    sP[t] = q.u; sP[np + t] = q.v; sP[2 * np + t] = q.w; sP[3 * np + t] = q.T;
  }
  if (t < NBAS) sB[t] = bas;
#pragma unroll
  for (int r = 0; r < RS; ++r) {
    const int x = t + r * NT;
    if (x < NF) {
      sIdx[x] = own[r];
      const Prim q = cons_to_prim(cn[r], gas);
      sNb[x] = q.u; sNb[NF + x] = q.v; sNb[2 * NF + x] = q.w; sNb[3 * NF + x] = q.T;
    }
  }
  __syncthreads();
  if (raw) {
    if (t < np) {
      const Prim q = cons_to_prim(sRaw + 5 * t, gas);
      sP[t] = q.u; sP[np + t] = q.v; sP[2 * np + t] = q.w; sP[3 * np + t] = q.T;
    }
    __syncthreads();
  }
  // one lane per boundary node, in boundary-rank order (the Jsoa rows): lanes 0..NB-1, so the
  // waves beyond NB/64 end here instead of idling through the lifting
  if (t >= NB) return;
  int i, j, k;
  bnd_node<N>(t, i, j, k);
  const int nd = i + n * (j + n * k);
  const int sides = node_sides<N>(i, j, k);
  const GasT<RF> gt(gas);
  // metric terms of the boundary nodes only, rows of NB per element (Jsoa layout)
  const double* Je = Jsoa + (size_t)10 * NB * e;
  RF Ja[9], sj;
#pragma unroll
  for (int r = 0; r < 9; ++r) Ja[r] = (RF)Je[r * NB + t];
  sj = (RF)Je[9 * NB + t];
  RF g[12];
  lift_node<N, RF>(nd, sP, sB, sNb, Ja, sj, g);
  RF fx[4], fy[4], fz[4];
  diff_flux<RF>(gt, (RF)sP[nd], (RF)sP[np + nd], (RF)sP[2 * np + nd], g, g + 4, g + 8, fx, fy, fz);
#pragma unroll 1
  for (int ls = 1; ls <= 6; ++ls) {
    if (!(sides & (1 << (ls - 1)))) continue;
    const int dir = dir_of(ls);
    const RF s = (RF)normal_sign(ls);
    const RF n0 = Ja[3 * dir] * s, n1 = Ja[3 * dir + 1] * s, n2 = Ja[3 * dir + 2] * s;
    const unsigned own = sIdx[face_point<N>(ls, i, j, k)];
    RF* dst = FG + 4 * (size_t)own;
#pragma unroll
    for (int c = 0; c < 4; ++c) dst[c] = fx[c] * n0 + fy[c] * n1 + fz[c] * n2;
    if constexpr (BC) {
      // a boundary side: keep the gradients at this face point for k_bc_visc
      const int side = (int)(own / (2u * (unsigned)(n * n)));
      if (side < m.nBC) {
        RF* gb = gBC + 12 * slot_side_point<N>(own);
#pragma unroll
        for (int c = 0; c < 12; ++c) gb[c] = g[c];
      }
    }
  }
}

// One lane per boundary side point, after k_visc_faces: the outer slot's viscous flux, so that with
// the inner element's half it makes the boundary's viscous flux (getboundaryflux.f90:561-776):
// Riemann type 0.5 (f_in + f_b).n with the boundary state's velocity, walls f_wall.n with the
// wall's velocity (none, or the tangential part at a slip wall) and the inner gradients (slip walls:
// as slip_wall_gradients leaves them; no heat flux at an adiabatic wall). n SurfElem is the side's.
template <class RF>
__global__ __launch_bounds__(256) void k_bc_visc(DevMesh m, const double* __restrict__ FU, const RF* __restrict__ gBC,
                                               RF* __restrict__ FG, const double* __restrict__ bcP, Gas gas, int nf,
                                               int nPts) {
  const int fp = blockIdx.x * blockDim.x + threadIdx.x;
  if (fp >= nPts) return;
  const int side = fp / nf, pq = fp - side * nf;
  const size_t in = ((size_t)side * 2) * nf + pq, out = in + nf;
  const int type = m.bcInfo[2 * side];
  const GasT<RF> gt(gas);
  const RF se = (RF)m.SurfElem[fp];
  const RF nh[3] = {(RF)m.NormVec[3 * (size_t)fp], (RF)m.NormVec[3 * (size_t)fp + 1], (RF)m.NormVec[3 * (size_t)fp + 2]};
  RF q[12];
#pragma unroll
  for (int c = 0; c < 12; ++c) q[c] = gBC[12 * (size_t)fp + c];
  RF ub[3] = {RF(0), RF(0), RF(0)};
  const bool riemann = bc_riemann_type(type);
  if (type == BC_DIRICHLET) {
    const double* ref = m.refPrim + 6 * m.bcInfo[2 * side + 1];
    ub[0] = (RF)ref[1]; ub[1] = (RF)ref[2]; ub[2] = (RF)ref[3];
  } else if (riemann) {  // the boundary state's velocity, as k_bc_ghost formed it
    ub[0] = (RF)bcP[4 * (size_t)fp + 1]; ub[1] = (RF)bcP[4 * (size_t)fp + 2]; ub[2] = (RF)bcP[4 * (size_t)fp + 3];
  } else if (bc_slip_wall(type)) {
    const Prim p = cons_to_prim(FU + 5 * in, gas);
    const RF vn = (RF)p.u * nh[0] + (RF)p.v * nh[1] + (RF)p.w * nh[2];
    ub[0] = (RF)p.u - vn * nh[0]; ub[1] = (RF)p.v - vn * nh[1]; ub[2] = (RF)p.w - vn * nh[2];
    const RF t1[3] = {(RF)m.TangVec1[3 * (size_t)fp], (RF)m.TangVec1[3 * (size_t)fp + 1],
                      (RF)m.TangVec1[3 * (size_t)fp + 2]};
    const RF t2[3] = {(RF)m.TangVec2[3 * (size_t)fp], (RF)m.TangVec2[3 * (size_t)fp + 1],
                      (RF)m.TangVec2[3 * (size_t)fp + 2]};
    slip_wall_gradients<RF>(type, nh, t1, t2, q);
  }
  RF bx[4], by[4], bz[4];
  diff_flux<RF>(gt, ub[0], ub[1], ub[2], q, q + 4, q + 8, bx, by, bz);
  if (type == BC_WALL_ADIABATIC) bx[3] = by[3] = bz[3] = RF(0);
  const RF* fi = FG + 4 * in;
  RF* fo = FG + 4 * out;
#pragma unroll
  for (int c = 0; c < 4; ++c) {
    const RF fb = (bx[c] * nh[0] + by[c] * nh[1] + bz[c] * nh[2]) * se;
    fo[c] = riemann ? -fb : fi[c] - RF(2) * fb;
  }
}

// One lane per boundary side point: the outer slot of the boundary side, from the inner face
// state and GetBoundaryState (getboundaryflux.f90:224-472). Riemann type: the boundary state itself
// (the Riemann solver and the lifting then see it as FLEXI's boundary flux does); its velocity also
// goes to bcP for k_bc_visc. Walls: a state whose lifting variables are 2 f - q_in, f the boundary
// lifting value (Lifting_GetBoundaryFlux, :855-939), so that the BR1 difference (q_out - q_in)/2 is
// FLEXI's f - q_in; their inviscid flux (0, p_wall n, 0) is formed by the stage kernel from bcP.
__global__ __launch_bounds__(256) void k_bc_ghost(DevMesh m, double* __restrict__ FU, double* __restrict__ bcP,
                                                Gas gas, int nf, int nPts) {
  const int fp = blockIdx.x * blockDim.x + threadIdx.x;
  if (fp >= nPts) return;
  const int side = fp / nf, pq = fp - side * nf;
  const size_t in = ((size_t)side * 2) * nf + pq, out = in + nf;
  const Prim q = cons_to_prim(FU + 5 * in, gas);
  const double Pm[6] = {q.rho, q.u, q.v, q.w, q.p, q.T};
  double nv[3], t1[3], t2[3];
#pragma unroll
  for (int c = 0; c < 3; ++c) {
    nv[c] = m.NormVec[3 * (size_t)fp + c];
    t1[c] = m.TangVec1[3 * (size_t)fp + c];
    t2[c] = m.TangVec2[3 * (size_t)fp + c];
  }
  const int type = m.bcInfo[2 * side], st = m.bcInfo[2 * side + 1];
  double Pb[6];
  boundary_state(gas, type, st >= 0 ? m.refPrim + 6 * st : nullptr, m.bcData ? m.bcData + 6 * (size_t)fp : nullptr,
                 Pm, nv, t1, t2, Pb);
  double G[6];
  double* bp = bcP + 4 * (size_t)fp;
  if (bc_riemann_type(type)) {
#pragma unroll
    for (int v = 0; v < 6; ++v) G[v] = Pb[v];
    bp[0] = Pb[4]; bp[1] = Pb[1]; bp[2] = Pb[2]; bp[3] = Pb[3];
  } else {
    double F[4];
    boundary_lifting_flux(type, Pm, Pb, F);
    G[0] = Pm[0];
    G[1] = Pm[1] + 2.0 * F[0]; G[2] = Pm[2] + 2.0 * F[1]; G[3] = Pm[3] + 2.0 * F[2];
    G[5] = Pm[5] + 2.0 * F[3];
    G[4] = G[0] * gas.R * G[5];
    bp[0] = Pb[4];
  }
  double* o = FU + 5 * out;
  o[0] = G[0];
  o[1] = G[1] * G[0]; o[2] = G[2] * G[0]; o[3] = G[3] * G[0];
  o[4] = gas.sKappaM1 * G[4] + 0.5 * (o[1] * G[1] + o[2] * G[2] + o[3] * G[3]);
}

}  // namespace

hipError_t launch_bc_ghost(const DevMesh& m, double* FU, double* bcP, const Gas& gas, hipStream_t s) {
  const int nf = (m.N + 1) * (m.N + 1), nPts = m.nBC * nf;
  if (nPts == 0) return hipSuccess;
  hipLaunchKernelGGL(k_bc_ghost, dim3((nPts + 255) / 256), dim3(256), 0, s, m, FU, bcP, gas, nf, nPts);
  return hipGetLastError();
}

hipError_t launch_visc_faces(const DevMesh& m, bool fp32, const double* U, const double* Jsoa, const double* FU,
                             void* FG, const Gas& gas, void* gBC, const double* bcP, hipStream_t s) {
  const bool bc = m.nBC > 0;
  switch (m.N) {
#define FXG_LAUNCH(NN, RF, BC)                                                                                    \
  hipLaunchKernelGGL((k_visc_faces<NN, RF, BC>), dim3(m.nElems), dim3(D<NN>::NT), 0, s, U, m, Jsoa, FU, (RF*)FG, \
                     gas, (RF*)gBC)
#define FXG_CASE(NN)                                                        \
  case NN:                                                                  \
    if (fp32) {                                                             \
      if (bc) FXG_LAUNCH(NN, float, true); else FXG_LAUNCH(NN, float, false); \
    } else {                                                                \
      if (bc) FXG_LAUNCH(NN, double, true); else FXG_LAUNCH(NN, double, false); \
    }                                                                       \
    break;
    FXG_CASE(2) FXG_CASE(3) FXG_CASE(4) FXG_CASE(5) FXG_CASE(6) FXG_CASE(7)
#undef FXG_CASE
#undef FXG_LAUNCH
    default: return hipErrorInvalidValue;
  }
  if (bc) {
    const int nf = (m.N + 1) * (m.N + 1), nPts = m.nBC * nf;
    if (fp32)
      hipLaunchKernelGGL((k_bc_visc<float>), dim3((nPts + 255) / 256), dim3(256), 0, s, m, FU, (const float*)gBC,
                         (float*)FG, bcP, gas, nf, nPts);
    else
      hipLaunchKernelGGL((k_bc_visc<double>), dim3((nPts + 255) / 256), dim3(256), 0, s, m, FU, (const double*)gBC,
                         (double*)FG, bcP, gas, nf, nPts);
  }
  return hipGetLastError();
}

}  // namespace fxg
