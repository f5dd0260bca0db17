// CPU oracle, part 2: mesh metrics (curl form), Jacobian and surface metrics.
// TEST INFRASTRUCTURE ONLY. Follows mesh/metrics.f90:75-787 (non-tree, curl-form branch).
#include "fxg_oracle.h"
#include "oracle_internal.h"

#include <cmath>
#include <vector>

using namespace orc;

namespace {

// ChangeBasis3D, interpolation/changeBasis.t90:231-279. X(dim,0:Nin^3) -> Y(dim,0:Nout^3).
void change_basis_3d(int dim, int Nin, int Nout, const double* Vdm, const double* X, double* Y) {
  const int ni = Nin + 1, no = Nout + 1;
  std::vector<double> b1(dim * no * ni * ni, 0.0), b2(dim * no * no * ni, 0.0);
  for (int kI = 0; kI < ni; ++kI)
    for (int jI = 0; jI < ni; ++jI)
      for (int iI = 0; iI < ni; ++iI)
        for (int iO = 0; iO < no; ++iO)
          for (int d = 0; d < dim; ++d)
            b1[d + dim * (iO + no * (jI + ni * kI))] +=
                Vdm[iO + no * iI] * X[d + dim * (iI + ni * (jI + ni * kI))];
  for (int kI = 0; kI < ni; ++kI)
    for (int jI = 0; jI < ni; ++jI)
      for (int jO = 0; jO < no; ++jO)
        for (int iO = 0; iO < no; ++iO)
          for (int d = 0; d < dim; ++d)
            b2[d + dim * (iO + no * (jO + no * kI))] +=
                Vdm[jO + no * jI] * b1[d + dim * (iO + no * (jI + ni * kI))];
  for (int i = 0; i < dim * no * no * no; ++i) Y[i] = 0.0;
  for (int kI = 0; kI < ni; ++kI)
    for (int kO = 0; kO < no; ++kO)
      for (int jO = 0; jO < no; ++jO)
        for (int iO = 0; iO < no; ++iO)
          for (int d = 0; d < dim; ++d)
            Y[d + dim * (iO + no * (jO + no * kO))] +=
                Vdm[kO + no * kI] * b2[d + dim * (iO + no * (jO + no * kI))];
}

// ChangeBasis2D, interpolation/changeBasis.t90:103-143. X(dim,0:Nin^2) -> Y(dim,0:Nout^2).
void change_basis_2d(int dim, int Nin, int Nout, const double* Vdm, const double* X, double* Y) {
  const int ni = Nin + 1, no = Nout + 1;
  std::vector<double> b1(dim * no * ni, 0.0);
  for (int jI = 0; jI < ni; ++jI)
    for (int iI = 0; iI < ni; ++iI)
      for (int iO = 0; iO < no; ++iO)
        for (int d = 0; d < dim; ++d)
          b1[d + dim * (iO + no * jI)] += Vdm[iO + no * iI] * X[d + dim * (iI + ni * jI)];
  for (int i = 0; i < dim * no * no; ++i) Y[i] = 0.0;
  for (int jI = 0; jI < ni; ++jI)
    for (int jO = 0; jO < no; ++jO)
      for (int iO = 0; iO < no; ++iO)
        for (int d = 0; d < dim; ++d)
          Y[d + dim * (iO + no * jO)] += Vdm[jO + no * jI] * b1[d + dim * (iO + no * jI)];
}

const int NormalDirs[7] = {0, 3, 2, 1, 2, 1, 3};          // mesh_vars.f90:105
const int TangDirs[7] = {0, 1, 3, 2, 3, 2, 1};            // mesh_vars.f90:107
const double NormalSigns[7] = {0, -1., -1., 1., 1., -1., 1.};  // mesh_vars.f90:111

}  // namespace

extern "C" void or_calc_metrics(int N, int nodeType, int NGeo, int nElems, const double* NodeCoords,
                                int nSides, const int* ElemToSide, double* Mf, double* Mg,
                                double* Mh, double* sJ, double* Elem_xGP, double* NormVec,
                                double* TangVec1, double* TangVec2, double* SurfElem,
                                double* Face_xGP) {
  (void)nSides;
  const int n = N + 1, ng = NGeo + 1, NGeoRef = 3 * NGeo, nr = NGeoRef + 1;  // mesh.f90:336
  const int np = n * n * n, ngp = ng * ng * ng, nrp = nr * nr * nr;
  std::vector<double> Vdm_EQNGeo_CLNGeo(ng * ng), Vdm_CLNGeo_CLN(n * ng), DCL_NGeo(ng * ng),
      Vdm_CLNGeo_NGeoRef(nr * ng), Vdm_NGeoRef_N(n * nr), DCL_N(n * n), Vdm_CLN_N(n * n),
      xCL(n), Vdm_EQ_N(n * ng);
  vandermonde(NGeo, OR_NODE_VISU, NGeo, OR_NODE_CL, 0, Vdm_EQNGeo_CLNGeo.data());
  vandermonde(NGeo, OR_NODE_CL, N, OR_NODE_CL, 0, Vdm_CLNGeo_CLN.data());
  {
    std::vector<double> x(ng);
    nodes(NGeo, OR_NODE_CL, x.data(), nullptr);
    derivative_matrix(NGeo, x.data(), DCL_NGeo.data());
  }
  vandermonde(NGeo, OR_NODE_CL, NGeoRef, nodeType, 0, Vdm_CLNGeo_NGeoRef.data());
  vandermonde(NGeoRef, nodeType, N, nodeType, 1, Vdm_NGeoRef_N.data());
  nodes(N, OR_NODE_CL, xCL.data(), nullptr);
  derivative_matrix(N, xCL.data(), DCL_N.data());
  vandermonde(N, OR_NODE_CL, N, nodeType, 0, Vdm_CLN_N.data());
  // BuildCoords (metrics.f90:75-156): Vdm_EQNGeo_N = Vdm_CLN_N * Vdm_EQNGeo_CLN
  {
    std::vector<double> V1(n * ng);
    vandermonde(NGeo, OR_NODE_VISU, N, OR_NODE_CL, 0, V1.data());
    for (int i = 0; i < n; ++i)
      for (int j = 0; j < ng; ++j) {
        double s = 0.0;
        for (int k = 0; k < n; ++k) s += Vdm_CLN_N[i + n * k] * V1[k + n * j];
        Vdm_EQ_N[i + n * j] = s;
      }
  }

  std::vector<double> XCL_NGeo(3 * ngp), XCL_N(3 * np), dXCL_NGeo(9 * ngp), dX_Ref(9 * nrp),
      tmpo(3 * nrp), detRef(nrp), detN(np), dXCL_N(9 * np), R_CL(9 * np),
      JaCL(9 * np), bufI(3 * ngp), bufO(3 * np);
  // dXCL_NGeo(d1,d2,i,j,k): d1 = derivative direction, d2 = coordinate
  auto dx = [&](std::vector<double>& a, int d1, int d2, int idx) -> double& { return a[d1 + 3 * (d2 + 3 * idx)]; };

  for (int e = 0; e < nElems; ++e) {
    const double* NC = NodeCoords + (size_t)3 * ngp * e;
    change_basis_3d(3, NGeo, N, Vdm_EQ_N.data(), NC, Elem_xGP + (size_t)3 * np * e);
    change_basis_3d(3, NGeo, NGeo, Vdm_EQNGeo_CLNGeo.data(), NC, XCL_NGeo.data());
    change_basis_3d(3, NGeo, N, Vdm_CLNGeo_CLN.data(), XCL_NGeo.data(), XCL_N.data());
    for (auto& v : dXCL_NGeo) v = 0.0;
    for (int k = 0; k < ng; ++k)
      for (int j = 0; j < ng; ++j)
        for (int i = 0; i < ng; ++i) {
          const int id = i + ng * (j + ng * k);
          for (int ll = 0; ll < ng; ++ll) {
            for (int c = 0; c < 3; ++c) {
              dx(dXCL_NGeo, 0, c, id) += DCL_NGeo[i + ng * ll] * XCL_NGeo[c + 3 * (ll + ng * (j + ng * k))];
              dx(dXCL_NGeo, 1, c, id) += DCL_NGeo[j + ng * ll] * XCL_NGeo[c + 3 * (i + ng * (ll + ng * k))];
            }
            for (int c = 0; c < 3; ++c)
              dx(dXCL_NGeo, 2, c, id) += DCL_NGeo[k + ng * ll] * XCL_NGeo[c + 3 * (i + ng * (j + ng * ll))];
          }
        }
    // dX_NGeoRef(:,d,...) = ChangeBasisVolume(Vdm_CLNGeo_NGeoRef, dXCL_NGeo(:,d,...))
    for (int d = 0; d < 3; ++d) {
      for (int id = 0; id < ngp; ++id)
        for (int c = 0; c < 3; ++c) bufI[c + 3 * id] = dx(dXCL_NGeo, c, d, id);
      change_basis_3d(3, NGeo, NGeoRef, Vdm_CLNGeo_NGeoRef.data(), bufI.data(), tmpo.data());
      for (int id = 0; id < nrp; ++id)
        for (int c = 0; c < 3; ++c) dx(dX_Ref, c, d, id) = tmpo[c + 3 * id];
    }
    for (int id = 0; id < nrp; ++id) {
      auto X = [&](int a, int b) { return dx(dX_Ref, a - 1, b - 1, id); };
      double det = 0.0;
      det = det + X(1, 1) * (X(2, 2) * X(3, 3) - X(3, 2) * X(2, 3)) +
            X(2, 1) * (X(3, 2) * X(1, 3) - X(1, 2) * X(3, 3)) +
            X(3, 1) * (X(1, 2) * X(2, 3) - X(2, 2) * X(1, 3));
      detRef[id] = det;
    }
    change_basis_3d(1, NGeoRef, N, Vdm_NGeoRef_N.data(), detRef.data(), detN.data());
    for (int id = 0; id < np; ++id) sJ[(size_t)np * e + id] = 1.0 / detN[id];

    if (N >= NGeo) {
      for (int d = 0; d < 3; ++d) {
        for (int id = 0; id < ngp; ++id)
          for (int c = 0; c < 3; ++c) bufI[c + 3 * id] = dx(dXCL_NGeo, c, d, id);
        change_basis_3d(3, NGeo, N, Vdm_CLNGeo_CLN.data(), bufI.data(), bufO.data());
        for (int id = 0; id < np; ++id)
          for (int c = 0; c < 3; ++c) dx(dXCL_N, c, d, id) = bufO[c + 3 * id];
      }
    } else {
      for (auto& v : dXCL_N) v = 0.0;
      for (int k = 0; k < n; ++k)
        for (int j = 0; j < n; ++j)
          for (int i = 0; i < n; ++i) {
            const int id = i + n * (j + n * k);
            for (int ll = 0; ll < n; ++ll)
              for (int c = 0; c < 3; ++c) {
                dx(dXCL_N, 0, c, id) += DCL_N[i + n * ll] * XCL_N[c + 3 * (ll + n * (j + n * k))];
                dx(dXCL_N, 1, c, id) += DCL_N[j + n * ll] * XCL_N[c + 3 * (i + n * (ll + n * k))];
                dx(dXCL_N, 2, c, id) += DCL_N[k + n * ll] * XCL_N[c + 3 * (i + n * (j + n * ll))];
              }
          }
    }
    // Curl-form metrics, metrics.f90:441-474. R_CL_N(:,d,...) and JaCL_N(dd,:,...).
    for (int id = 0; id < np; ++id) {
      const double* xc = &XCL_N[3 * id];
      for (int c = 0; c < 3; ++c) {
        R_CL[c + 3 * (0 + 3 * id)] = 0.5 * (xc[2] * dx(dXCL_N, c, 1, id) - xc[1] * dx(dXCL_N, c, 2, id));
        R_CL[c + 3 * (1 + 3 * id)] = 0.5 * (xc[0] * dx(dXCL_N, c, 2, id) - xc[2] * dx(dXCL_N, c, 0, id));
        R_CL[c + 3 * (2 + 3 * id)] = 0.5 * (xc[1] * dx(dXCL_N, c, 0, id) - xc[0] * dx(dXCL_N, c, 1, id));
      }
    }
    // JaCL(dd,c,id) layout: [dd + 3*(c + 3*id)]
    for (auto& v : JaCL) v = 0.0;
    auto R = [&](int c, int d, int i, int j, int k) { return R_CL[c + 3 * (d + 3 * (i + n * (j + n * k)))]; };
    for (int k = 0; k < n; ++k)
      for (int j = 0; j < n; ++j)
        for (int i = 0; i < n; ++i) {
          double* J = &JaCL[9 * (i + n * (j + n * k))];
          for (int q = 0; q < n; ++q)
            for (int c = 0; c < 3; ++c) {
              J[0 + 3 * c] = J[0 + 3 * c] - DCL_N[j + n * q] * R(2, c, i, q, k);
              J[1 + 3 * c] = J[1 + 3 * c] - DCL_N[k + n * q] * R(0, c, i, j, q);
              J[2 + 3 * c] = J[2 + 3 * c] - DCL_N[i + n * q] * R(1, c, q, j, k);
            }
          for (int q = 0; q < n; ++q)
            for (int c = 0; c < 3; ++c) {
              J[0 + 3 * c] = J[0 + 3 * c] + DCL_N[k + n * q] * R(1, c, i, j, q);
              J[1 + 3 * c] = J[1 + 3 * c] + DCL_N[i + n * q] * R(2, c, q, j, k);
              J[2 + 3 * c] = J[2 + 3 * c] + DCL_N[j + n * q] * R(0, c, i, q, k);
            }
        }
    // Metrics_*Tilde = ChangeBasisVolume(Vdm_CLN_N, JaCL_N(dd,:,...)), metrics.f90:536-541
    double* Mout[3] = {Mf + (size_t)3 * np * e, Mg + (size_t)3 * np * e, Mh + (size_t)3 * np * e};
    for (int dd = 0; dd < 3; ++dd) {
      for (int id = 0; id < np; ++id)
        for (int c = 0; c < 3; ++c) bufO[c + 3 * id] = JaCL[dd + 3 * (c + 3 * id)];
      change_basis_3d(3, N, N, Vdm_CLN_N.data(), bufO.data(), Mout[dd]);
    }
    // CalcSurfMetrics, metrics.f90:601-745 (conforming sides only)
    std::vector<double> tmp(3 * n * n), tmp2(3 * n * n), JaFace(9 * n * n);
    for (int ls = 1; ls <= 6; ++ls) {
      const int SideID = ElemToSide[0 + 2 * ((ls - 1) + 6 * e)] - 1;
      const int flip = ElemToSide[1 + 2 * ((ls - 1) + 6 * e)];
      auto slice = [&](auto get) {
        for (int b = 0; b < n; ++b)
          for (int a = 0; a < n; ++a)
            for (int c = 0; c < 3; ++c) tmp[c + 3 * (a + n * b)] = get(c, a, b);
      };
      auto vol_index = [&](int a, int b) {
        switch (ls) {
          case 5: return 0 + n * (a + n * b);
          case 3: return N + n * (a + n * b);
          case 2: return a + n * (0 + n * b);
          case 4: return a + n * (N + n * b);
          case 1: return a + n * (b + n * 0);
          default: return a + n * (b + n * N);
        }
      };
      if (flip == 0) {
        slice([&](int c, int a, int b) { return XCL_N[c + 3 * vol_index(a, b)]; });
        change_basis_2d(3, N, N, Vdm_CLN_N.data(), tmp.data(), tmp2.data());
        for (int q = 0; q < n; ++q)
          for (int p = 0; p < n; ++p) {
            int pq[2];
            side_to_vol2(N, p, q, 0, ls, pq);
            for (int c = 0; c < 3; ++c)
              Face_xGP[c + 3 * (p + n * (q + n * SideID))] = tmp2[c + 3 * (pq[0] + n * pq[1])];
          }
      }
      for (auto& v : JaFace) v = 0.0;
      for (int dd = 0; dd < 3; ++dd) {
        slice([&](int c, int a, int b) { return JaCL[dd + 3 * (c + 3 * vol_index(a, b))]; });
        change_basis_2d(3, N, N, Vdm_CLN_N.data(), tmp.data(), tmp2.data());
        for (int q = 0; q < n; ++q)
          for (int p = 0; p < n; ++p) {
            int pq[2];
            side_to_vol2(N, p, q, flip, ls, pq);
            for (int c = 0; c < 3; ++c) JaFace[dd + 3 * (c + 3 * (p + n * q))] = tmp2[c + 3 * (pq[0] + n * pq[1])];
          }
      }
      if (flip != 0) continue;
      // SurfMetricsFromJa, metrics.f90:751-787
      const int nd = NormalDirs[ls] - 1, td = TangDirs[ls] - 1;
      const double ns = NormalSigns[ls];
      for (int q = 0; q < n; ++q)
        for (int p = 0; p < n; ++p) {
          const int fi = p + n * (q + n * SideID);
          const double* J = &JaFace[9 * (p + n * q)];
          double s2 = 0.0;
          for (int c = 0; c < 3; ++c) s2 += J[nd + 3 * c] * J[nd + 3 * c];
          const double se = std::sqrt(s2);
          SurfElem[fi] = se;
          double nv[3], t1[3];
          for (int c = 0; c < 3; ++c) nv[c] = ns * J[nd + 3 * c] / se;
          double dot = 0.0;
          for (int c = 0; c < 3; ++c) dot += J[td + 3 * c] * nv[c];
          for (int c = 0; c < 3; ++c) t1[c] = J[td + 3 * c] - dot * nv[c];
          double tn = 0.0;
          for (int c = 0; c < 3; ++c) tn += t1[c] * t1[c];
          tn = std::sqrt(tn);
          for (int c = 0; c < 3; ++c) t1[c] = t1[c] / tn;
          for (int c = 0; c < 3; ++c) {
            NormVec[c + 3 * fi] = nv[c];
            TangVec1[c + 3 * fi] = t1[c];
          }
          // CROSS, globals/mathtools.f90
          TangVec2[0 + 3 * fi] = nv[1] * t1[2] - nv[2] * t1[1];
          TangVec2[1 + 3 * fi] = nv[2] * t1[0] - nv[0] * t1[2];
          TangVec2[2 + 3 * fi] = nv[0] * t1[1] - nv[1] * t1[0];
        }
    }
  }
}
