// Mesh metrics on the GPU (FLEXI src/mesh/metrics.f90:161-787, curl form).
//
// Pipeline per element: equidistant NodeCoords -> Chebyshev-Lobatto (CL) geometry of degree NGeo
// -> CL geometry and its derivatives on degree N -> Jacobian determinant sampled exactly at
// NGeoRef = 3*NGeo nodes and projected (modal truncation) to N -> sJ; curl-form metric terms
// Ja = -curl(X_l grad X_m)/2-type combinations on the CL-N grid, interpolated to the solution
// nodes; surface metrics of each master side from the interpolated face metrics.
#include "fxg_metrics.h"

#include <vector>

#include "fxg_basis.h"

namespace fxg {
namespace {

// Offsets of the small matrices in the device buffer (all column-major, rows = output nodes).
struct MetricMats {
  int ng, n, nr;
  size_t V_EQ_CLG, V_CLG_CLN, D_CLG, V_CLG_REF, V_REF_N, D_CLN, V_CLN_N, V_EQ_N, total;
  MetricMats(int N, int NGeo) {
    n = N + 1; ng = NGeo + 1; nr = 3 * NGeo + 1;
    size_t o = 0;
    V_EQ_CLG = o; o += (size_t)ng * ng;
    V_CLG_CLN = o; o += (size_t)n * ng;
    D_CLG = o; o += (size_t)ng * ng;
    V_CLG_REF = o; o += (size_t)nr * ng;
    V_REF_N = o; o += (size_t)n * nr;
    D_CLN = o; o += (size_t)n * n;
    V_CLN_N = o; o += (size_t)n * n;
    V_EQ_N = o; o += (size_t)n * ng;
    total = o;
  }
};

__device__ __forceinline__ int vol_id(int n, int i, int j, int k) { return i + n * (j + n * k); }

// Geometry on CL nodes: XCL_N(3,0:N^3), dXCL_NGeo(9,0:NGeo^3) [index d + 3*c: derivative
// direction d of coordinate c], Elem_xGP(3,0:N^3). One workgroup per element.
__global__ __launch_bounds__(256) void k_geo(const double* __restrict__ NC, const double* __restrict__ mats, MetricMats M,
                                           double* __restrict__ XCLN, double* __restrict__ dXCLG,
                                           double* __restrict__ xGP) {
  extern __shared__ double sh[];
  const int ng = M.ng, n = M.n, ngp = ng * ng * ng, np = n * n * n;
  double* sNC = sh;            // 3*ngp
  double* sX = sh + 3 * ngp;   // XCL_NGeo, 3*ngp
  const int e = blockIdx.x, t = threadIdx.x, NT = blockDim.x;
  const double* nce = NC + (size_t)3 * ngp * e;
  for (int x = t; x < 3 * ngp; x += NT) sNC[x] = nce[x];
  __syncthreads();
  const double* Veq = mats + M.V_EQ_CLG;
  for (int o = t; o < ngp; o += NT) {
    const int io = o % ng, jo = (o / ng) % ng, ko = o / (ng * ng);
    double acc[3] = {0, 0, 0};
    for (int c3 = 0; c3 < ng; ++c3)
      for (int b = 0; b < ng; ++b)
        for (int a = 0; a < ng; ++a) {
          const double w = Veq[io + ng * a] * Veq[jo + ng * b] * Veq[ko + ng * c3];
          const int id = vol_id(ng, a, b, c3);
          for (int c = 0; c < 3; ++c) acc[c] += w * sNC[3 * id + c];
        }
    for (int c = 0; c < 3; ++c) sX[3 * o + c] = acc[c];
  }
  // Elem_xGP = Vdm_EQNGeo_N applied to NodeCoords (BuildCoords, metrics.f90:75-156)
  const double* Vxn = mats + M.V_EQ_N;
  for (int o = t; o < np; o += NT) {
    const int io = o % n, jo = (o / n) % n, ko = o / (n * n);
    double acc[3] = {0, 0, 0};
    for (int c3 = 0; c3 < ng; ++c3)
      for (int b = 0; b < ng; ++b)
        for (int a = 0; a < ng; ++a) {
          const double w = Vxn[io + n * a] * Vxn[jo + n * b] * Vxn[ko + n * c3];
          const int id = vol_id(ng, a, b, c3);
          for (int c = 0; c < 3; ++c) acc[c] += w * sNC[3 * id + c];
        }
    for (int c = 0; c < 3; ++c) xGP[(size_t)3 * ((size_t)np * e + o) + c] = acc[c];
  }
  __syncthreads();
  const double* Vcn = mats + M.V_CLG_CLN;
  for (int o = t; o < np; o += NT) {
    const int io = o % n, jo = (o / n) % n, ko = o / (n * n);
    double acc[3] = {0, 0, 0};
    for (int c3 = 0; c3 < ng; ++c3)
      for (int b = 0; b < ng; ++b)
        for (int a = 0; a < ng; ++a) {
          const double w = Vcn[io + n * a] * Vcn[jo + n * b] * Vcn[ko + n * c3];
          const int id = vol_id(ng, a, b, c3);
          for (int c = 0; c < 3; ++c) acc[c] += w * sX[3 * id + c];
        }
    for (int c = 0; c < 3; ++c) XCLN[(size_t)3 * ((size_t)np * e + o) + c] = acc[c];
  }
  const double* D = mats + M.D_CLG;
  for (int o = t; o < ngp; o += NT) {
    const int i = o % ng, j = (o / ng) % ng, k = o / (ng * ng);
    double d[9] = {0, 0, 0, 0, 0, 0, 0, 0, 0};
    for (int l = 0; l < ng; ++l)
      for (int c = 0; c < 3; ++c) {
        d[0 + 3 * c] += D[i + ng * l] * sX[3 * vol_id(ng, l, j, k) + c];
        d[1 + 3 * c] += D[j + ng * l] * sX[3 * vol_id(ng, i, l, k) + c];
        d[2 + 3 * c] += D[k + ng * l] * sX[3 * vol_id(ng, i, j, l) + c];
      }
    for (int q = 0; q < 9; ++q) dXCLG[(size_t)9 * ((size_t)ngp * e + o) + q] = d[q];
  }
}

// Jacobian determinant at the NGeoRef^3 nodes of the solution node type (exact for the
// polynomial mapping, metrics.f90:253-275). One lane per (element, reference node).
__global__ __launch_bounds__(256) void k_jac_ref(const double* __restrict__ dXCLG, const double* __restrict__ mats,
                                               MetricMats M, int nElems, double* __restrict__ detRef) {
  const int ng = M.ng, nr = M.nr, ngp = ng * ng * ng, nrp = nr * nr * nr;
  const long x = (long)blockIdx.x * blockDim.x + threadIdx.x;
  if (x >= (long)nrp * nElems) return;
  const int e = (int)(x / nrp), r = (int)(x - (long)e * nrp);
  const int ir = r % nr, jr = (r / nr) % nr, kr = r / (nr * nr);
  const double* V = mats + M.V_CLG_REF;
  const double* dX = dXCLG + (size_t)9 * ngp * e;
  double X[9] = {0, 0, 0, 0, 0, 0, 0, 0, 0};
  for (int c3 = 0; c3 < ng; ++c3)
    for (int b = 0; b < ng; ++b)
      for (int a = 0; a < ng; ++a) {
        const double w = V[ir + nr * a] * V[jr + nr * b] * V[kr + nr * c3];
        const double* src = dX + 9 * vol_id(ng, a, b, c3);
        for (int q = 0; q < 9; ++q) X[q] += w * src[q];
      }
  // X[d + 3*c] = dX_c/dxi_d ; det of the 3x3 Jacobian (metrics.f90:266-268)
  auto A = [&](int d, int c) { return X[(d - 1) + 3 * (c - 1)]; };
  detRef[x] = A(1, 1) * (A(2, 2) * A(3, 3) - A(3, 2) * A(2, 3)) + A(2, 1) * (A(3, 2) * A(1, 3) - A(1, 2) * A(3, 3)) +
              A(3, 1) * (A(1, 2) * A(2, 3) - A(2, 2) * A(1, 3));
}

// Everything on the degree-N grid: sJ, dXCL_N, curl-form JaCL_N, interpolation of the metric
// terms to the solution nodes, surface metrics of the element's master sides.
__global__ __launch_bounds__(512) void k_metrics_n(const double* __restrict__ XCLN, const double* __restrict__ dXCLG,
                                                 const double* __restrict__ detRef, const double* __restrict__ mats,
                                                 MetricMats M, int NGeo, DevMesh m, double* __restrict__ Mf,
                                                 double* __restrict__ Mg, double* __restrict__ Mh,
                                                 double* __restrict__ sJ, double* __restrict__ NormVec,
                                                 double* __restrict__ TangVec1, double* __restrict__ TangVec2,
                                                 double* __restrict__ SurfElem, double* __restrict__ SlaveNS) {
  extern __shared__ double sh[];
  const int n = M.n, ng = M.ng, nr = M.nr, np = n * n * n, ngp = ng * ng * ng, nrp = nr * nr * nr, nf = n * n;
  const int N = n - 1;
  double* sR = sh;  // R_CL_N, then JaCL_N: [q*np + node], q = d + 3*c (9 values)
  const int e = blockIdx.x, t = threadIdx.x, NT = blockDim.x;
  const double* Vcn = mats + M.V_CLG_CLN;
  const double* DN = mats + M.D_CLN;
  const double* xe = XCLN + (size_t)3 * np * e;
  for (int o = t; o < np; o += NT) {
    const int i = o % n, j = (o / n) % n, k = o / (n * n);
    // sJ = 1/detJ, detJ projected from the reference nodes (modal truncation when N < NGeoRef)
    const double* W = mats + M.V_REF_N;
    const double* dr = detRef + (size_t)nrp * e;
    double det = 0.0;
    for (int c3 = 0; c3 < nr; ++c3)
      for (int b = 0; b < nr; ++b) {
        const double wbc = W[j + n * b] * W[k + n * c3];
        for (int a = 0; a < nr; ++a) det += W[i + n * a] * wbc * dr[vol_id(nr, a, b, c3)];
      }
    sJ[(size_t)np * e + o] = 1.0 / det;
    // dXCL_N at this node
    double d[9] = {0, 0, 0, 0, 0, 0, 0, 0, 0};
    if (N >= NGeo) {
      const double* src = dXCLG + (size_t)9 * ngp * e;
      for (int c3 = 0; c3 < ng; ++c3)
        for (int b = 0; b < ng; ++b)
          for (int a = 0; a < ng; ++a) {
            const double w = Vcn[i + n * a] * Vcn[j + n * b] * Vcn[k + n * c3];
            for (int q = 0; q < 9; ++q) d[q] += w * src[9 * vol_id(ng, a, b, c3) + q];
          }
    } else {
      for (int l = 0; l < n; ++l)
        for (int c = 0; c < 3; ++c) {
          d[0 + 3 * c] += DN[i + n * l] * xe[3 * vol_id(n, l, j, k) + c];
          d[1 + 3 * c] += DN[j + n * l] * xe[3 * vol_id(n, i, l, k) + c];
          d[2 + 3 * c] += DN[k + n * l] * xe[3 * vol_id(n, i, j, l) + c];
        }
    }
    // R_CL_N(:,1) = (z dY - y dZ)/2, R(:,2) = (x dZ - z dX)/2, R(:,3) = (y dX - x dY)/2  (metrics.f90:443-447).
    // Coordinates are taken relative to the element's first CL node: the discrete curl of
    // c*grad(X) vanishes for constant c (the tensor-product derivatives commute), so the metric
    // terms are unchanged, but the products stay O(h^2) instead of O(|x| h). Translated copies
    // of an element (periodic neighbours) then get the same face metrics to round-off of h^2.
    const double X = xe[3 * o] - xe[0], Y = xe[3 * o + 1] - xe[1], Z = xe[3 * o + 2] - xe[2];
    for (int dd = 0; dd < 3; ++dd) {
      sR[(dd + 3 * 0) * np + o] = 0.5 * (Z * d[dd + 3 * 1] - Y * d[dd + 3 * 2]);
      sR[(dd + 3 * 1) * np + o] = 0.5 * (X * d[dd + 3 * 2] - Z * d[dd + 3 * 0]);
      sR[(dd + 3 * 2) * np + o] = 0.5 * (Y * d[dd + 3 * 0] - X * d[dd + 3 * 1]);
    }
  }
  __syncthreads();
  // JaCL_N(dd,c): discrete curl of R (metrics.f90:449-464)
  double ja[4][9];
  int cnt = 0;
  for (int o = t; o < np; o += NT, ++cnt) {
    const int i = o % n, j = (o / n) % n, k = o / (n * n);
    double J[9] = {0, 0, 0, 0, 0, 0, 0, 0, 0};  // J[dd + 3*c]
    auto R = [&](int d, int c, int a, int b, int cc) { return sR[(d + 3 * c) * np + vol_id(n, a, b, cc)]; };
    for (int q = 0; q < n; ++q)
      for (int c = 0; c < 3; ++c) {
        J[0 + 3 * c] -= DN[j + n * q] * R(2, c, i, q, k);
        J[1 + 3 * c] -= DN[k + n * q] * R(0, c, i, j, q);
        J[2 + 3 * c] -= DN[i + n * q] * R(1, c, q, j, k);
      }
    for (int q = 0; q < n; ++q)
      for (int c = 0; c < 3; ++c) {
        J[0 + 3 * c] += DN[k + n * q] * R(1, c, i, j, q);
        J[1 + 3 * c] += DN[i + n * q] * R(2, c, q, j, k);
        J[2 + 3 * c] += DN[j + n * q] * R(0, c, i, q, k);
      }
    for (int q = 0; q < 9; ++q) ja[cnt][q] = J[q];
  }
  __syncthreads();
  cnt = 0;
  for (int o = t; o < np; o += NT, ++cnt)
    for (int q = 0; q < 9; ++q) sR[q * np + o] = ja[cnt][q];
  __syncthreads();
  // Metrics_{f,g,h}Tilde = Vdm_CLN_N applied to JaCL_N(dd,:,...)
  const double* V = mats + M.V_CLN_N;
  double* Mout[3] = {Mf, Mg, Mh};
  for (int o = t; o < np; o += NT) {
    const int io = o % n, jo = (o / n) % n, ko = o / (n * n);
    double acc[9] = {0, 0, 0, 0, 0, 0, 0, 0, 0};
    for (int c3 = 0; c3 < n; ++c3)
      for (int b = 0; b < n; ++b) {
        const double wbc = V[jo + n * b] * V[ko + n * c3];
        for (int a = 0; a < n; ++a) {
          const double w = V[io + n * a] * wbc;
          const int id = vol_id(n, a, b, c3);
          for (int q = 0; q < 9; ++q) acc[q] += w * sR[q * np + id];
        }
      }
    for (int dd = 0; dd < 3; ++dd)
      for (int c = 0; c < 3; ++c) Mout[dd][(size_t)3 * ((size_t)np * e + o) + c] = acc[dd + 3 * c];
  }
  // Surface metrics of master sides (CalcSurfMetrics + SurfMetricsFromJa, metrics.f90:601-787)
  const int NormalDirs[7] = {0, 3, 2, 1, 2, 1, 3}, TangDirs[7] = {0, 1, 3, 2, 3, 2, 1};
  const double NormalSigns[7] = {0, -1., -1., 1., 1., -1., 1.};
  for (int x = t; x < 6 * nf; x += NT) {
    const int ls = x / nf + 1, pq = x - (ls - 1) * nf;
    const int side = m.E2S[2 * (ls - 1) + 12 * e] - 1;
    const int flip = m.E2S[2 * (ls - 1) + 1 + 12 * e];
    if (flip != 0 && !SlaveNS) continue;
    const int* ab = m.s2v2 + 2 * (pq + nf * (flip + 5 * (ls - 1)));
    const int a0 = ab[0], b0 = ab[1];
    const int l = (ls == 5 || ls == 2 || ls == 1) ? 0 : N;
    const int dir = (ls == 5 || ls == 3) ? 0 : (ls == 2 || ls == 4) ? 1 : 2;
    double jf[9] = {0, 0, 0, 0, 0, 0, 0, 0, 0};
    for (int b = 0; b < n; ++b)
      for (int a = 0; a < n; ++a) {
        const double w = V[a0 + n * a] * V[b0 + n * b];
        const int id = dir == 0 ? vol_id(n, l, a, b) : (dir == 1 ? vol_id(n, a, l, b) : vol_id(n, a, b, l));
        for (int q = 0; q < 9; ++q) jf[q] += w * sR[q * np + id];
      }
    const int nd = NormalDirs[ls] - 1, td = TangDirs[ls] - 1;
    const double ns = NormalSigns[ls];
    if (flip != 0) {
      // the slave element's own n*SurfElem of the side point, oriented as the master's normal
      const size_t fs = (size_t)pq + (size_t)nf * side;
      for (int c = 0; c < 3; ++c) SlaveNS[3 * fs + c] = -ns * jf[nd + 3 * c];
      continue;
    }
    const double se = sqrt(jf[nd] * jf[nd] + jf[nd + 3] * jf[nd + 3] + jf[nd + 6] * jf[nd + 6]);
    double nv[3], t1[3];
    for (int c = 0; c < 3; ++c) nv[c] = ns * jf[nd + 3 * c] / se;
    const double dot = jf[td] * nv[0] + jf[td + 3] * nv[1] + jf[td + 6] * nv[2];
    for (int c = 0; c < 3; ++c) t1[c] = jf[td + 3 * c] - dot * nv[c];
    const double tn = sqrt(t1[0] * t1[0] + t1[1] * t1[1] + t1[2] * t1[2]);
    for (int c = 0; c < 3; ++c) t1[c] /= tn;
    const size_t fp = (size_t)pq + (size_t)nf * side;
    SurfElem[fp] = se;
    for (int c = 0; c < 3; ++c) {
      NormVec[3 * fp + c] = nv[c];
      TangVec1[3 * fp + c] = t1[c];
    }
    TangVec2[3 * fp + 0] = nv[1] * t1[2] - nv[2] * t1[1];
    TangVec2[3 * fp + 1] = nv[2] * t1[0] - nv[0] * t1[2];
    TangVec2[3 * fp + 2] = nv[0] * t1[1] - nv[1] * t1[0];
  }
}

// ---------------------------------------------------------------------------------------------
// The reference's operation order (FXG_METRICS_REFERENCE). The same quantities as k_geo,
// k_jac_ref and k_metrics_n, evaluated as metrics.f90:75-787 evaluates them: absolute coordinates
// in the curl form, every change of basis as ChangeBasis3D/2D's one-dimensional passes
// (changeBasis.t90:103-143,231-279), every sum accumulated from zero in the order of its loop
// index, no FMA contraction. One lane forms each output value and recomputes the partial sums of
// the passes it needs, which are the same values a pass-by-pass evaluation stores.

// ChangeBasis3D at the output node (iO,jO,kO): component `comp` of X (node stride `dim`),
// V(0:no-1,0:ni-1).
__device__ double cb3_at(const double* __restrict__ V, int no, int ni, const double* X, int dim, int comp, int iO,
                         int jO, int kO) {
#pragma clang fp contract(off)
  double y = 0.0;
  for (int kI = 0; kI < ni; ++kI) {
    double b2 = 0.0;
    for (int jI = 0; jI < ni; ++jI) {
      double b1 = 0.0;
      for (int iI = 0; iI < ni; ++iI) b1 += V[iO + no * iI] * X[comp + dim * (iI + ni * (jI + ni * kI))];
      b2 += V[jO + no * jI] * b1;
    }
    y += V[kO + no * kI] * b2;
  }
  return y;
}

// Dynamic LDS of k_metrics_ref, in doubles: XCL_N (3 np), dXCL_NGeo (9 ngp), then a region used
// first for NodeCoords, XCL_NGeo and detJ at the NGeoRef nodes, then for R_CL_N, which JaCL_N
// replaces in place.
struct RefLds {
  int xn, dxg, nc, xg, det, r, total;
  __host__ __device__ RefLds(int np, int ngp, int nrp) {
    xn = 0;
    dxg = xn + 3 * np;
    nc = dxg + 9 * ngp;
    xg = nc + 3 * ngp;
    det = xg + 3 * ngp;
    r = nc;
    const int end1 = det + nrp, end2 = r + 9 * np;
    total = end1 > end2 ? end1 : end2;
  }
};

// One workgroup of 512 lanes per element, up to 4 nodes per lane (np <= 2048).
__global__ __launch_bounds__(512) void k_metrics_ref(const double* __restrict__ NC, const double* __restrict__ mats,
                                                   MetricMats M, int NGeo, DevMesh m, double* __restrict__ Mf,
                                                   double* __restrict__ Mg, double* __restrict__ Mh,
                                                   double* __restrict__ sJ, double* __restrict__ xGP,
                                                   double* __restrict__ NormVec, double* __restrict__ TangVec1,
                                                   double* __restrict__ TangVec2, double* __restrict__ SurfElem) {
#pragma clang fp contract(off)
  extern __shared__ double sh[];
  const int n = M.n, ng = M.ng, nr = M.nr, np = n * n * n, ngp = ng * ng * ng, nrp = nr * nr * nr, nf = n * n;
  const int N = n - 1;
  const RefLds L(np, ngp, nrp);
  double* sXN = sh + L.xn;
  double* sdXG = sh + L.dxg;  // [d1 + 3*(c + 3*id)]: derivative direction d1 of coordinate c
  double* sNC = sh + L.nc;
  double* sXG = sh + L.xg;
  double* sDet = sh + L.det;
  double* sR = sh + L.r;   // R_CL_N [c + 3*(d + 3*id)]
  double* sJa = sh + L.r;  // JaCL_N [dd + 3*(c + 3*id)], in place of R_CL_N
  const int e = blockIdx.x, t = threadIdx.x, NT = blockDim.x;
  const double* nce = NC + (size_t)3 * ngp * e;
  for (int x = t; x < 3 * ngp; x += NT) sNC[x] = nce[x];
  __syncthreads();
  // BuildCoords (metrics.f90:75-156): XCL_NGeo and Elem_xGP from the equidistant NodeCoords
  for (int x = t; x < 3 * ngp; x += NT) {
    const int c = x % 3, o = x / 3;
    sXG[x] = cb3_at(mats + M.V_EQ_CLG, ng, ng, sNC, 3, c, o % ng, (o / ng) % ng, o / (ng * ng));
  }
  for (int x = t; x < 3 * np; x += NT) {
    const int c = x % 3, o = x / 3;
    xGP[(size_t)3 * np * e + x] = cb3_at(mats + M.V_EQ_N, n, ng, sNC, 3, c, o % n, (o / n) % n, o / (n * n));
  }
  __syncthreads();
  // XCL_N, and dXCL_NGeo by the CL derivative matrix (metrics.f90:228-245)
  for (int x = t; x < 3 * np; x += NT) {
    const int c = x % 3, o = x / 3;
    sXN[x] = cb3_at(mats + M.V_CLG_CLN, n, ng, sXG, 3, c, o % n, (o / n) % n, o / (n * n));
  }
  const double* DG = mats + M.D_CLG;
  for (int x = t; x < 9 * ngp; x += NT) {
    const int d1 = x % 3, c = (x / 3) % 3, o = x / 9;
    const int i = o % ng, j = (o / ng) % ng, k = o / (ng * ng);
    double s = 0.0;
    for (int l = 0; l < ng; ++l) {
      const int id = d1 == 0 ? vol_id(ng, l, j, k) : (d1 == 1 ? vol_id(ng, i, l, k) : vol_id(ng, i, j, l));
      const double Dv = DG[(d1 == 0 ? i : (d1 == 1 ? j : k)) + ng * l];
      s += Dv * sXG[c + 3 * id];
    }
    sdXG[x] = s;
  }
  __syncthreads();
  // detJ at the NGeoRef nodes (metrics.f90:253-275)
  for (int r = t; r < nrp; r += NT) {
    const int ir = r % nr, jr = (r / nr) % nr, kr = r / (nr * nr);
    double X[9];
    for (int q = 0; q < 9; ++q) X[q] = cb3_at(mats + M.V_CLG_REF, nr, ng, sdXG, 9, q, ir, jr, kr);
    auto A = [&](int a, int b) { return X[(a - 1) + 3 * (b - 1)]; };
    double det = 0.0;
    det = det + A(1, 1) * (A(2, 2) * A(3, 3) - A(3, 2) * A(2, 3)) + A(2, 1) * (A(3, 2) * A(1, 3) - A(1, 2) * A(3, 3)) +
          A(3, 1) * (A(1, 2) * A(2, 3) - A(2, 2) * A(1, 3));
    sDet[r] = det;
  }
  __syncthreads();
  // sJ = 1/detJ, detJ projected to the solution nodes (modal truncation when N < NGeoRef)
  for (int o = t; o < np; o += NT)
    sJ[(size_t)np * e + o] = 1.0 / cb3_at(mats + M.V_REF_N, n, nr, sDet, 1, 0, o % n, (o / n) % n, o / (n * n));
  __syncthreads();  // the first region is free: R_CL_N and JaCL_N take it
  // R_CL_N from dXCL_N and the absolute CL coordinates (metrics.f90:441-447)
  const double* DN = mats + M.D_CLN;
  for (int o = t; o < np; o += NT) {
    const int i = o % n, j = (o / n) % n, k = o / (n * n);
    double d[9];  // [d1 + 3*c]
    if (N >= NGeo) {
      for (int q = 0; q < 9; ++q) d[q] = cb3_at(mats + M.V_CLG_CLN, n, ng, sdXG, 9, q, i, j, k);
    } else {
      for (int q = 0; q < 9; ++q) {
        const int d1 = q % 3, c = q / 3;
        double s = 0.0;
        for (int l = 0; l < n; ++l) {
          const int id = d1 == 0 ? vol_id(n, l, j, k) : (d1 == 1 ? vol_id(n, i, l, k) : vol_id(n, i, j, l));
          s += DN[(d1 == 0 ? i : (d1 == 1 ? j : k)) + n * l] * sXN[c + 3 * id];
        }
        d[q] = s;
      }
    }
    const double* xc = sXN + 3 * o;
    for (int c = 0; c < 3; ++c) {
      sR[c + 3 * (0 + 3 * o)] = 0.5 * (xc[2] * d[c + 3 * 1] - xc[1] * d[c + 3 * 2]);
      sR[c + 3 * (1 + 3 * o)] = 0.5 * (xc[0] * d[c + 3 * 2] - xc[2] * d[c + 3 * 0]);
      sR[c + 3 * (2 + 3 * o)] = 0.5 * (xc[1] * d[c + 3 * 0] - xc[0] * d[c + 3 * 1]);
    }
  }
  __syncthreads();
  // JaCL_N: discrete curl of R (metrics.f90:449-464), the subtracted terms first; held in
  // registers until every lane has read R
  double ja[4][9];
  int cnt = 0;
  for (int o = t; o < np; o += NT, ++cnt) {
    const int i = o % n, j = (o / n) % n, k = o / (n * n);
    auto R = [&](int a, int b, int idx) { return sR[a + 3 * (b + 3 * idx)]; };
    double J[9] = {0, 0, 0, 0, 0, 0, 0, 0, 0};  // [dd + 3*c]
    for (int q = 0; q < n; ++q)
      for (int c = 0; c < 3; ++c) {
        J[0 + 3 * c] = J[0 + 3 * c] - DN[j + n * q] * R(2, c, vol_id(n, i, q, k));
        J[1 + 3 * c] = J[1 + 3 * c] - DN[k + n * q] * R(0, c, vol_id(n, i, j, q));
        J[2 + 3 * c] = J[2 + 3 * c] - DN[i + n * q] * R(1, c, vol_id(n, q, j, k));
      }
    for (int q = 0; q < n; ++q)
      for (int c = 0; c < 3; ++c) {
        J[0 + 3 * c] = J[0 + 3 * c] + DN[k + n * q] * R(1, c, vol_id(n, i, j, q));
        J[1 + 3 * c] = J[1 + 3 * c] + DN[i + n * q] * R(2, c, vol_id(n, q, j, k));
        J[2 + 3 * c] = J[2 + 3 * c] + DN[j + n * q] * R(0, c, vol_id(n, i, q, k));
      }
    for (int q = 0; q < 9; ++q) ja[cnt][q] = J[q];
  }
  __syncthreads();
  cnt = 0;
  for (int o = t; o < np; o += NT, ++cnt)
    for (int q = 0; q < 9; ++q) sJa[q + 9 * o] = ja[cnt][q];
  __syncthreads();
  // Metrics_{f,g,h}Tilde: ChangeBasis3D(Vdm_CLN_N) of JaCL_N(dd,:,...) (metrics.f90:536-541)
  double* Mout[3] = {Mf, Mg, Mh};
  for (int x = t; x < 9 * np; x += NT) {
    const int c = x % 3, dd = (x / 3) % 3, o = x / 9;
    Mout[dd][(size_t)3 * ((size_t)np * e + o) + c] =
        cb3_at(mats + M.V_CLN_N, n, n, sJa, 9, dd + 3 * c, o % n, (o / n) % n, o / (n * n));
  }
  // Surface metrics of the master sides: ChangeBasis2D of the face slice of JaCL_N, then
  // SurfMetricsFromJa (metrics.f90:601-787)
  const double* V = mats + M.V_CLN_N;
  const int NormalDirs[7] = {0, 3, 2, 1, 2, 1, 3}, TangDirs[7] = {0, 1, 3, 2, 3, 2, 1};
  const double NormalSigns[7] = {0, -1., -1., 1., 1., -1., 1.};
  for (int x = t; x < 6 * nf; x += NT) {
    const int ls = x / nf + 1, pq = x - (ls - 1) * nf;
    const int side = m.E2S[2 * (ls - 1) + 12 * e] - 1;
    const int flip = m.E2S[2 * (ls - 1) + 1 + 12 * e];
    if (flip != 0) continue;
    const int* ab = m.s2v2 + 2 * (pq + nf * (5 * (ls - 1)));
    const int a0 = ab[0], b0 = ab[1];
    const int l = (ls == 5 || ls == 2 || ls == 1) ? 0 : N;
    const int dir = (ls == 5 || ls == 3) ? 0 : (ls == 2 || ls == 4) ? 1 : 2;
    double jf[9];  // [dd + 3*c]
    for (int q = 0; q < 9; ++q) {
      double y = 0.0;
      for (int bI = 0; bI < n; ++bI) {
        double b1 = 0.0;
        for (int aI = 0; aI < n; ++aI) {
          const int id = dir == 0 ? vol_id(n, l, aI, bI) : (dir == 1 ? vol_id(n, aI, l, bI) : vol_id(n, aI, bI, l));
          b1 += V[a0 + n * aI] * sJa[q + 9 * id];
        }
        y += V[b0 + n * bI] * b1;
      }
      jf[q] = y;
    }
    const int nd = NormalDirs[ls] - 1, td = TangDirs[ls] - 1;
    const double ns = NormalSigns[ls];
    double s2 = 0.0;
    for (int c = 0; c < 3; ++c) s2 += jf[nd + 3 * c] * jf[nd + 3 * c];
    const double se = sqrt(s2);
    double nv[3], t1[3];
    for (int c = 0; c < 3; ++c) nv[c] = ns * jf[nd + 3 * c] / se;
    double dot = 0.0;
    for (int c = 0; c < 3; ++c) dot += jf[td + 3 * c] * nv[c];
    for (int c = 0; c < 3; ++c) t1[c] = jf[td + 3 * c] - dot * nv[c];
    double tn = 0.0;
    for (int c = 0; c < 3; ++c) tn += t1[c] * t1[c];
    tn = sqrt(tn);
    for (int c = 0; c < 3; ++c) t1[c] = t1[c] / tn;
    const size_t fp = (size_t)pq + (size_t)nf * side;
    SurfElem[fp] = se;
    for (int c = 0; c < 3; ++c) {
      NormVec[3 * fp + c] = nv[c];
      TangVec1[3 * fp + c] = t1[c];
    }
    TangVec2[3 * fp + 0] = nv[1] * t1[2] - nv[2] * t1[1];
    TangVec2[3 * fp + 1] = nv[2] * t1[0] - nv[0] * t1[2];
    TangVec2[3 * fp + 2] = nv[0] * t1[1] - nv[1] * t1[0];
  }
}

}  // namespace

int compute_metrics_gpu(int N, int nodeType, int NGeo, int nElems, const double* NodeCoords, const DevMesh& m,
                        double* Mf, double* Mg, double* Mh, double* sJ, double* xGP, double* NormVec,
                        double* TangVec1, double* TangVec2, double* SurfElem, double* SlaveNS, bool reference,
                        hipStream_t s, std::string& err) {
  using namespace basis;
  MetricMats M(N, NGeo);
  const int n = N + 1, ng = NGeo + 1, NGeoRef = 3 * NGeo;
  const int sol = nodeType == 1 ? GAUSS : GAUSS_LOBATTO;
  std::vector<double> h(M.total, 0.0);
  auto put = [&](size_t off, const std::vector<double>& v) { std::copy(v.begin(), v.end(), h.begin() + off); };
  put(M.V_EQ_CLG, vandermonde(NGeo, EQUIDISTANT, NGeo, CHEBYSHEV_LOBATTO, false));
  put(M.V_CLG_CLN, vandermonde(NGeo, CHEBYSHEV_LOBATTO, N, CHEBYSHEV_LOBATTO, false));
  {
    std::vector<double> x(ng), D((size_t)ng * ng);
    nodes(NGeo, CHEBYSHEV_LOBATTO, x.data());
    diff_matrix(NGeo, x.data(), D.data());
    put(M.D_CLG, D);
  }
  put(M.V_CLG_REF, vandermonde(NGeo, CHEBYSHEV_LOBATTO, NGeoRef, sol, false));
  put(M.V_REF_N, vandermonde(NGeoRef, sol, N, sol, true));
  {
    std::vector<double> x(n), D((size_t)n * n);
    nodes(N, CHEBYSHEV_LOBATTO, x.data());
    diff_matrix(N, x.data(), D.data());
    put(M.D_CLN, D);
  }
  const std::vector<double> Vcln = vandermonde(N, CHEBYSHEV_LOBATTO, N, sol, false);
  put(M.V_CLN_N, Vcln);
  {
    const std::vector<double> V1 = vandermonde(NGeo, EQUIDISTANT, N, CHEBYSHEV_LOBATTO, false);
    std::vector<double> V((size_t)n * ng, 0.0);
    for (int i = 0; i < n; ++i)
      for (int j = 0; j < ng; ++j) {
        double acc = 0.0;
        for (int k = 0; k < n; ++k) acc += Vcln[i + n * k] * V1[k + n * j];
        V[i + n * j] = acc;
      }
    put(M.V_EQ_N, V);
  }
  const int np = n * n * n, ngp = ng * ng * ng, nr = NGeoRef + 1, nrp = nr * nr * nr;
  if (reference) {
    const RefLds L(np, ngp, nrp);
    const size_t sh = (size_t)L.total * sizeof(double);
    if (SlaveNS) {
      err = "compute_metrics_gpu: the reference order shares the master's face metric (no SlaveNS)";
      return 1;
    }
    if (np > 4 * 512 || sh > 160 * 1024) {
      err = "compute_metrics_gpu: N/NGeo too large for the LDS of the reference-order metrics";
      return 1;
    }
    double* dm = nullptr;
    hipError_t e = hipMalloc((void**)&dm, M.total * sizeof(double));
    if (e == hipSuccess) e = hipMemcpyAsync(dm, h.data(), M.total * sizeof(double), hipMemcpyHostToDevice, s);
    if (e == hipSuccess && sh > 64 * 1024)
      e = hipFuncSetAttribute(reinterpret_cast<const void*>(&k_metrics_ref), hipFuncAttributeMaxDynamicSharedMemorySize,
                              (int)sh);
    if (e == hipSuccess) {
      hipLaunchKernelGGL(k_metrics_ref, dim3(nElems), dim3(512), sh, s, NodeCoords, dm, M, NGeo, m, Mf, Mg, Mh, sJ, xGP,
                         NormVec, TangVec1, TangVec2, SurfElem);
      e = hipGetLastError();
    }
    if (e == hipSuccess) e = hipStreamSynchronize(s);
    (void)hipFree(dm);
    if (e != hipSuccess) {
      err = std::string("compute_metrics_gpu: ") + hipGetErrorString(e);
      return 1;
    }
    return 0;
  }
  double *dm = nullptr, *XCLN = nullptr, *dXCLG = nullptr, *detRef = nullptr;
  hipError_t e = hipMalloc((void**)&dm, M.total * sizeof(double));
  if (e == hipSuccess) e = hipMalloc((void**)&XCLN, (size_t)3 * np * nElems * sizeof(double));
  if (e == hipSuccess) e = hipMalloc((void**)&dXCLG, (size_t)9 * ngp * nElems * sizeof(double));
  if (e == hipSuccess) e = hipMalloc((void**)&detRef, (size_t)nrp * nElems * sizeof(double));
  if (e == hipSuccess) e = hipMemcpyAsync(dm, h.data(), M.total * sizeof(double), hipMemcpyHostToDevice, s);
  const size_t shGeo = (size_t)6 * ngp * sizeof(double);
  const size_t shN = (size_t)9 * np * sizeof(double);
  if (e == hipSuccess && (shGeo > 65536 || shN > 65536)) {
    err = "compute_metrics_gpu: N/NGeo too large for the LDS staging";
    e = hipErrorInvalidValue;
  }
  if (e == hipSuccess) {
    hipLaunchKernelGGL(k_geo, dim3(nElems), dim3(256), shGeo, s, NodeCoords, dm, M, XCLN, dXCLG, xGP);
    e = hipGetLastError();
  }
  if (e == hipSuccess) {
    const long tot = (long)nrp * nElems;
    hipLaunchKernelGGL(k_jac_ref, dim3((unsigned)((tot + 255) / 256)), dim3(256), 0, s, dXCLG, dm, M, nElems, detRef);
    e = hipGetLastError();
  }
  if (e == hipSuccess) {
    // k_metrics_n keeps up to 4 nodes per lane in registers: np <= 4*512
    hipLaunchKernelGGL(k_metrics_n, dim3(nElems), dim3(512), shN, s, XCLN, dXCLG, detRef, dm, M, NGeo, m, Mf, Mg, Mh,
                       sJ, NormVec, TangVec1, TangVec2, SurfElem, SlaveNS);
    e = hipGetLastError();
  }
  if (e == hipSuccess) e = hipStreamSynchronize(s);
  (void)hipFree(dm);
  (void)hipFree(XCLN);
  (void)hipFree(dXCLG);
  (void)hipFree(detRef);
  if (e != hipSuccess) {
    if (err.empty()) err = std::string("compute_metrics_gpu: ") + hipGetErrorString(e);
    return 1;
  }
  return 0;
}

}  // namespace fxg
