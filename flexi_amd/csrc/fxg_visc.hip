// Kernel A of the fused Gauss-Lobatto split-form stage (the rest is in fxg_fused.hip): for each
// element, the BR1 gradients at its boundary nodes and, for each of its six faces, the viscous
// normal flux f_v . (n SurfElem) seen from that element (4 values per face point, momentum and
// energy) into the side array FG (lifting_br1.t90:48-181, flux.f90:149-241). A translation unit
// of its own, so that its code does not change the register allocation of k_dg_fused.
#include "fxg_elem.h"

namespace fxg {
namespace {

// FP64 stages the element's U in LDS with contiguous lanes (held to 80 VGPRs at N=7, 52 KB:
// three workgroups per CU); FP32-bulk converts U from global memory directly and is held to 64
// VGPRs (18 KB: four workgroups per CU, 0.42 ms instead of 0.53 ms at 32^3, N=7).
template <int N, class RF>
__global__ __launch_bounds__(D<N>::NT)
__attribute__((amdgpu_waves_per_eu(sizeof(RF) == 4 ? (N == 7 ? 8 : 6) : (N == 7 ? 6 : 1)))) void k_visc_faces(
    const double* __restrict__ U, DevMesh m, const double* __restrict__ Jsoa, const double* __restrict__ FU,
    RF* __restrict__ FG, Gas gas) {
  constexpr int n = D<N>::n, np = D<N>::np, NT = D<N>::NT, NF = D<N>::NF;
  constexpr bool raw = sizeof(RF) == 8;
  __shared__ double sRaw[raw ? 5 * np : 1];  // the element's U as stored, loaded with contiguous lanes
  __shared__ RF sP[4 * np];        // u, v, w, T
  __shared__ RF sNb[4 * NF];       // neighbour u, v, w, T per own face point
  __shared__ unsigned sIdx[NF];
  __shared__ double sB[3 * n * n + 4 * n];
  const int e = blockIdx.x, t = threadIdx.x;
  const double* Ue = U + (size_t)5 * np * e;
  if (raw)
    for (int x = t; x < 5 * np; x += NT) sRaw[x] = Ue[x];
  else if (t < np) {
    const Prim q = cons_to_prim(Ue + 5 * t, gas);
    sP[t] = q.u; sP[np + t] = q.v; sP[2 * np + t] = q.w; sP[3 * np + t] = q.T;
  }
  for (int x = t; x < 3 * n * n + 4 * n; x += NT) sB[x] = m.basis[x];
  for (int x = t; x < NF; x += NT) {
    const unsigned own = face_slot<N>(m, e, x);
    sIdx[x] = own;
    const Prim q = cons_to_prim(FU + 5 * (size_t)slot_nbr<N>(own), gas);
    sNb[x] = q.u; sNb[NF + x] = q.v; sNb[2 * NF + x] = q.w; sNb[3 * NF + x] = q.T;
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
  constexpr int NB = Bnd<N>::NB;
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
    RF* dst = FG + 4 * (size_t)sIdx[face_point<N>(ls, i, j, k)];
#pragma unroll
    for (int c = 0; c < 4; ++c) dst[c] = fx[c] * n0 + fy[c] * n1 + fz[c] * n2;
  }
}

}  // namespace

hipError_t launch_visc_faces(const DevMesh& m, bool fp32, const double* U, const double* Jsoa, const double* FU,
                             void* FG, const Gas& gas, hipStream_t s) {
  switch (m.N) {
#define FXG_CASE(NN)                                                                                     \
  case NN:                                                                                               \
    if (fp32)                                                                                            \
      hipLaunchKernelGGL((k_visc_faces<NN, float>), dim3(m.nElems), dim3(D<NN>::NT), 0, s, U, m, Jsoa, FU,  \
                         (float*)FG, gas);                                                               \
    else                                                                                                 \
      hipLaunchKernelGGL((k_visc_faces<NN, double>), dim3(m.nElems), dim3(D<NN>::NT), 0, s, U, m, Jsoa, FU, \
                         (double*)FG, gas);                                                              \
    break;
    FXG_CASE(2) FXG_CASE(3) FXG_CASE(4) FXG_CASE(5) FXG_CASE(6) FXG_CASE(7)
#undef FXG_CASE
    default: return hipErrorInvalidValue;
  }
  return hipGetLastError();
}

}  // namespace fxg
