// Element-local helpers shared by the fused stage kernels and the adjoint kernels (Gauss-Lobatto
// nodes, one workgroup per element, one lane per node): local side geometry of flexi.h (1 zeta-,
// 2 eta-, 3 xi+, 4 eta+, 5 xi-, 6 zeta+), side-array slots, BR1 lifting at a node.
#pragma once
#include "fxg_kernels.h"

namespace fxg {

template <int N>
struct D {
  static constexpr int n = N + 1, nf = n * n, np = n * n * n, NT = ((np + 63) / 64) * 64;
  static constexpr int NF = 6 * nf;  // own face points of an element
};

__device__ __forceinline__ bool minus_side(int ls) { return ls == 5 || ls == 2 || ls == 1; }
__device__ __forceinline__ int dir_of(int ls) { return (ls == 5 || ls == 3) ? 0 : (ls == 2 || ls == 4) ? 1 : 2; }
__device__ __forceinline__ double normal_sign(int ls) { return (ls == 3 || ls == 4 || ls == 6) ? 1.0 : -1.0; }

// Volume node of own face point (ls, a, b).
template <int N>
__device__ __forceinline__ int vidx(int ls, int a, int b) {
  constexpr int n = N + 1;
  const int l = minus_side(ls) ? 0 : N;
  switch (dir_of(ls)) {
    case 0: return l + n * (a + n * b);
    case 1: return a + n * (l + n * b);
    default: return a + n * (b + n * l);
  }
}

// Local sides a node lies on (bit ls-1 set).
template <int N>
__device__ __forceinline__ int node_sides(int i, int j, int k) {
  int m = 0;
  if (k == 0) m |= 1 << 0;   // zeta-
  if (j == 0) m |= 1 << 1;   // eta-
  if (i == N) m |= 1 << 2;   // xi+
  if (j == N) m |= 1 << 3;   // eta+
  if (i == 0) m |= 1 << 4;   // xi-
  if (k == N) m |= 1 << 5;   // zeta+
  return m;
}

// Own face-point index x = (ls-1)*nf + a + n*b of node (i,j,k) on local side ls.
template <int N>
__device__ __forceinline__ int face_point(int ls, int i, int j, int k) {
  constexpr int n = N + 1, nf = n * n;
  switch (dir_of(ls)) {
    case 0: return (ls - 1) * nf + j + n * k;
    case 1: return (ls - 1) * nf + i + n * k;
    default: return (ls - 1) * nf + i + n * j;
  }
}

// Side-array slot of own face point x of element e: ((side*2 + slot)*nf + pq), slot 0 master.
// In two steps for callers that issue other loads between them: the ElemToSide entry (side, flip)
// of the point's local side, then the slot through the inverse S2V2 map.
template <int N>
__device__ __forceinline__ int2 face_side_entry(const DevMesh& m, int e, int x) {
  constexpr int nf = (N + 1) * (N + 1);
  const int ls = x / nf + 1;
  return make_int2(m.E2S[2 * (ls - 1) + 12 * e] - 1, m.E2S[2 * (ls - 1) + 1 + 12 * e]);
}
template <int N>
__device__ __forceinline__ unsigned face_slot_from(const DevMesh& m, int2 sf, int x) {
  constexpr int n = N + 1, nf = n * n;
  const int ls = x / nf + 1, ab = x - (ls - 1) * nf;
  const int* pq = m.v2s2 + 2 * (ab + nf * (sf.y + 5 * (ls - 1)));
  return (unsigned)(((size_t)sf.x * 2 + (sf.y ? 1 : 0)) * nf + pq[0] + n * pq[1]);
}
template <int N>
__device__ __forceinline__ unsigned face_slot(const DevMesh& m, int e, int x) {
  return face_slot_from<N>(m, face_side_entry<N>(m, e, x), x);
}
// The same from the table of the fused path (m.slotTab, filled by launch_slot_table): one load
// instead of ElemToSide followed by the inverse S2V2 map.
template <int N>
__device__ __forceinline__ unsigned face_slot_tab(const DevMesh& m, int e, int x) {
  return m.slotTab[(size_t)6 * (N + 1) * (N + 1) * e + x];
}
template <int N>
__device__ __forceinline__ bool slot_is_master(unsigned own) { return ((own / (unsigned)((N + 1) * (N + 1))) & 1u) == 0; }
template <int N>
__device__ __forceinline__ unsigned slot_nbr(unsigned own) {
  constexpr unsigned nf = (N + 1) * (N + 1);
  return slot_is_master<N>(own) ? own + nf : own - nf;
}
template <int N>
__device__ __forceinline__ size_t slot_side_point(unsigned own) {
  constexpr unsigned nf = (N + 1) * (N + 1);
  return (size_t)(own / (2 * nf)) * nf + own % nf;
}

// BR1 lifting at node t (strong form, non-conservative): D_T contractions of the lifting
// variables held in LDS (sP: u,v,w,T as SoA rows of np), metric transform, surface corrections
// 0.5*(U_nb - U_own).(n SurfElem)*L_hat at face nodes with the neighbour values staged per own
// face point (sNb: u,v,w,T as SoA rows of NF), times sJ (lifting_br1.t90:48-181).
template <int N, class RF, class LT>
__device__ __forceinline__ void lift_node(int t, const LT* sP, const double* sB, const LT* sNb,
                                          const RF* Ja, RF sj, RF* g) {
  constexpr int n = N + 1, np = n * n * n, NF = D<N>::NF;
  const BasisLayout B{n};
  const int i = t % n, j = (t / n) % n, k = t / (n * n);
  const double* DT = sB + B.D_T();
  RF gxi[4], geta[4], gzeta[4];
#pragma unroll
  for (int v = 0; v < 4; ++v) gxi[v] = geta[v] = gzeta[v] = RF(0);
#pragma unroll
  for (int l = 0; l <= N; ++l) {
    const RF dxi = (RF)DT[l + n * i], deta = (RF)DT[l + n * j], dzeta = (RF)DT[l + n * k];
#pragma unroll
    for (int v = 0; v < 4; ++v) {
      gxi[v] += dxi * (RF)sP[v * np + (l + n * (j + n * k))];
      geta[v] += deta * (RF)sP[v * np + (i + n * (l + n * k))];
      gzeta[v] += dzeta * (RF)sP[v * np + (i + n * (j + n * l))];
    }
  }
#pragma unroll
  for (int d = 0; d < 3; ++d)
#pragma unroll
    for (int v = 0; v < 4; ++v) g[4 * d + v] = Ja[d] * gxi[v] + Ja[3 + d] * geta[v] + Ja[6 + d] * gzeta[v];
  const int sides = node_sides<N>(i, j, k);
  if (sides) {
    const RF lhm = (RF)sB[B.L_HatMinus()], lhp = (RF)sB[B.L_HatPlus() + N];
    const RF own[4] = {(RF)sP[t], (RF)sP[np + t], (RF)sP[2 * np + t], (RF)sP[3 * np + t]};
#pragma unroll 1
    for (int ls = 1; ls <= 6; ++ls) {
      if (!(sides & (1 << (ls - 1)))) continue;
      const int x = face_point<N>(ls, i, j, k);
      const int dir = dir_of(ls);
      const RF s = (RF)normal_sign(ls) * RF(0.5) * (minus_side(ls) ? lhm : lhp);
      const RF nse[3] = {Ja[3 * dir] * s, Ja[3 * dir + 1] * s, Ja[3 * dir + 2] * s};
      RF dq[4];
#pragma unroll
      for (int v = 0; v < 4; ++v) dq[v] = (RF)sNb[v * NF + x] - own[v];
#pragma unroll
      for (int d = 0; d < 3; ++d)
#pragma unroll
        for (int v = 0; v < 4; ++v) g[4 * d + v] += dq[v] * nse[d];
    }
  }
#pragma unroll
  for (int c = 0; c < 12; ++c) g[c] *= sj;
}

template <class RF>
__device__ __forceinline__ void load_metrics(const DevMesh& m, size_t gid, RF* Ja, RF& sj) {
#pragma unroll
  for (int c = 0; c < 3; ++c) {
    Ja[c] = (RF)m.Mf[3 * gid + c];
    Ja[3 + c] = (RF)m.Mg[3 * gid + c];
    Ja[6 + c] = (RF)m.Mh[3 * gid + c];
  }
  sj = (RF)m.sJ[gid];
}

// Metric terms of node t from the element's block of the packed array (rows of np: Ja^1 x,y,z,
// Ja^2 x,y,z, Ja^3 x,y,z, sJ), so that consecutive lanes load consecutive doubles.
template <int N, class RF>
__device__ __forceinline__ void load_metrics(const double* Je, int t, RF* Ja, RF& sj) {
  constexpr int np = D<N>::np;
#pragma unroll
  for (int r = 0; r < 9; ++r) Ja[r] = (RF)Je[r * np + t];
  sj = (RF)Je[9 * np + t];
}

// Geometry of own face point (local side ls, boundary node id), the same arithmetic as the surface
// phase of the fused stage kernel: the master normal nv, SurfElem (returned) and a tangential
// basis (Gram-Schmidt of the coordinate axis least aligned with n, t2 = n x t1). With
// ownFaceMetric the own boundary metric NormalSign*Ja^{NormalDir} gives n*SurfElem (times -1 on a
// slave side for the master normal); otherwise the side's NormVec/SurfElem are used for both
// sides, as FLEXI does.
template <int N, class RF>
__device__ __forceinline__ double face_frame(const DevMesh& m, int e, int ls, int id, unsigned own, bool master, RF* nv,
                                             RF* t1, RF* t2) {
  constexpr int np = D<N>::np;
  double se;
  if (m.ownFaceMetric) {
    const int dir = dir_of(ls);
    const double sgn = normal_sign(ls), ms = master ? 1.0 : -1.0;
    const double* Jd = (dir == 0 ? m.Mf : (dir == 1 ? m.Mg : m.Mh)) + 3 * ((size_t)np * e + id);
    const double ns0 = sgn * Jd[0], ns1 = sgn * Jd[1], ns2 = sgn * Jd[2];
    se = sqrt(ns0 * ns0 + ns1 * ns1 + ns2 * ns2);
    const double mse = ms / se;
    nv[0] = (RF)(ns0 * mse); nv[1] = (RF)(ns1 * mse); nv[2] = (RF)(ns2 * mse);
  } else {
    const size_t sp = slot_side_point<N>(own);
    se = m.SurfElem[sp];
    nv[0] = (RF)m.NormVec[3 * sp]; nv[1] = (RF)m.NormVec[3 * sp + 1]; nv[2] = (RF)m.NormVec[3 * sp + 2];
  }
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
  return se;
}

// Boundary nodes of an element (on at least one face): NB of them, numbered plane by plane in
// volume order (k = 0 plane: all; 0 < k < N: the ring j = 0, i = 0 | N, j = N; k = N: all), so
// that consecutive boundary lanes address consecutive entries.
template <int N>
struct Bnd {
  static constexpr int n = N + 1, nf = n * n, NB = 2 * nf + (N - 1) * 4 * N;
};
template <int N>
__device__ __forceinline__ int bnd_rank(int i, int j, int k) {
  constexpr int n = N + 1, nf = n * n;
  if (k == 0) return i + n * j;
  if (k == N) return nf + (N - 1) * 4 * N + i + n * j;
  const int ring = (j == 0) ? i : (j == N ? n + 2 * (N - 1) + i : n + 2 * (j - 1) + (i == 0 ? 0 : 1));
  return nf + (k - 1) * 4 * N + ring;
}
// Inverse of bnd_rank: the node (i,j,k) of boundary rank b, 0 <= b < Bnd<N>::NB.
template <int N>
__device__ __forceinline__ void bnd_node(int b, int& i, int& j, int& k) {
  constexpr int n = N + 1, nf = n * n, NR = (N - 1) * 4 * N;
  if (b < nf) {
    k = 0; i = b % n; j = b / n;
  } else if (b >= nf + NR) {
    const int r = b - nf - NR;
    k = N; i = r % n; j = r / n;
  } else {
    const int r = b - nf, ring = r % (4 * N);
    k = 1 + r / (4 * N);
    if (ring < n) {
      j = 0; i = ring;
    } else if (ring >= n + 2 * (N - 1)) {
      j = N; i = ring - n - 2 * (N - 1);
    } else {
      const int q = ring - n;
      j = 1 + q / 2; i = (q % 2) ? N : 0;
    }
  }
}

}  // namespace fxg
