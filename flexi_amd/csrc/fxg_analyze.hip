// Taylor-Green vortex testcase analysis on the device: AnalyzeTestcase
// (src/testcase/taylorgreenvortex/testcase.f90:241-479) and the domain volume of InitAnalyze
// (src/analyze/analyze.f90:153-160).
//
// One workgroup per element stages the element's state U(5), the velocity gradients
// (GradVel(q,d) = d u_q / d x_d, 9 values) and sJ in LDS, interpolates them to the
// (NAnalyze+1)^3 Gauss-Lobatto points one z-plane at a time by sum factorisation (z, then y,
// then x contraction; FLEXI's ChangeBasis3D contracts x first, which changes the result only by
// rounding), evaluates the integrands at each point and reduces them over the workgroup into 13
// per-element partial sums (11 integrals, the maximum vorticity and the element volume). A
// single-workgroup kernel then sums the partials over the elements in a fixed order, so the
// result does not depend on scheduling. Only 13 doubles leave the device.
#include <hip/hip_runtime.h>

#include "fxg_kernels.h"
#include "fxg_physics.h"

namespace fxg {

namespace {

constexpr int NF = 15;  // interpolated fields: U(5), GradVel(3,3) column by column, sJ
constexpr int NAMAX = TGV_NANALYZE_MAX + 1;

// Partial sums per element, in this order.
enum { A_T, A_ENTROPY, A_EKIN, A_EKINC, A_ENSTR, A_DRU, A_DRS, A_DRSD, A_DRP, A_EDS, A_EDD, A_VMAX, A_VOL, A_N };

__device__ __forceinline__ double wave_sum(double v) {
  for (int off = 32; off > 0; off >>= 1) v += __shfl_down(v, off, 64);
  return v;
}
__device__ __forceinline__ double wave_max(double v) {
  for (int off = 32; off > 0; off >>= 1) v = fmax(v, __shfl_down(v, off, 64));
  return v;
}

// tab: V(na, n) row-major (V[a*n+i] = l_i(xi_a), the Lagrange polynomials of the solution nodes
// at the analysis points), then wA(na) and wGP(n).
template <int N, class RF>
__global__ __launch_bounds__(256) void k_tgv_analyze(const double* __restrict__ U, const RF* __restrict__ G,
                                                    const double* __restrict__ sJ, const double* __restrict__ tab,
                                                    int na, Gas gas, double* __restrict__ part) {
  constexpr int n = N + 1, nn = n * n, np = n * n * n;
  __shared__ double F[NF * np];
  __shared__ double P[NF * nn];
  __shared__ double Q[NF * n * NAMAX];
  __shared__ double V[NAMAX * n], wA[NAMAX], red[4 * A_N];
  const int t = threadIdx.x, e = blockIdx.x;
  for (int x = t; x < na * n; x += 256) V[x] = tab[x];
  for (int x = t; x < na; x += 256) wA[x] = tab[na * n + x];
  const double* wGP = tab + na * n + na;
  double acc[A_N];
#pragma unroll
  for (int q = 0; q < A_N; ++q) acc[q] = 0.0;
  const double* Ue = U + (size_t)5 * np * e;
  const RF* Ge = G + (size_t)12 * np * e;
  for (int x = t; x < np; x += 256) {
#pragma unroll
    for (int v = 0; v < 5; ++v) F[v * np + x] = Ue[5 * x + v];
    // GradVel(:,d) = gradient in direction d of (u,v,w): G holds (dir*4 + var) blocks
#pragma unroll
    for (int d = 0; d < 3; ++d)
#pragma unroll
      for (int q = 0; q < 3; ++q) F[(5 + 3 * d + q) * np + x] = (double)Ge[(4 * d + q) * np + x];
    const double s = sJ[(size_t)np * e + x];
    F[14 * np + x] = s;
    const int i = x % n, j = (x / n) % n, k = x / nn;
    acc[A_VOL] += wGP[i] * wGP[j] * wGP[k] / s;
  }
  __syncthreads();
  for (int c = 0; c < na; ++c) {
    // z contraction: P(f, i, j) = sum_k V(c, k) F(f, i, j, k)
    for (int x = t; x < NF * nn; x += 256) {
      const int f = x / nn, ij = x - f * nn;
      double s = 0.0;
#pragma unroll
      for (int k = 0; k < n; ++k) s += V[c * n + k] * F[f * np + ij + nn * k];
      P[x] = s;
    }
    __syncthreads();
    // y contraction: Q(f, i, b) = sum_j V(b, j) P(f, i, j)
    for (int x = t; x < NF * n * na; x += 256) {
      const int f = x / (n * na), r = x - f * n * na, b = r / n, i = r - b * n;
      double s = 0.0;
#pragma unroll
      for (int j = 0; j < n; ++j) s += V[b * n + j] * P[f * nn + i + n * j];
      Q[x] = s;
    }
    __syncthreads();
    // x contraction and the integrands at the points (a, b, c)
    for (int x = t; x < na * na; x += 256) {
      const int b = x / na, a = x - b * na;
      double y[NF];
#pragma unroll
      for (int f = 0; f < NF; ++f) {
        double s = 0.0;
#pragma unroll
        for (int i = 0; i < n; ++i) s += V[a * n + i] * Q[f * n * na + i + n * b];
        y[f] = s;
      }
      const Prim pr = cons_to_prim(y, gas);
      // G(q, d) = y[5 + 3d + q]
      const double g00 = y[5], g10 = y[6], g20 = y[7], g01 = y[8], g11 = y[9], g21 = y[10], g02 = y[11],
                   g12 = y[12], g22 = y[13];
      const double divU = g00 + g11 + g22;
      const double s01 = 0.5 * (g01 + g10), s02 = 0.5 * (g02 + g20), s12 = 0.5 * (g12 + g21);
      const double d3 = 1.0 / 3.0 * divU;
      const double sd00 = g00 - d3, sd11 = g11 - d3, sd22 = g22 - d3;
      const double w0 = g21 - g12, w1 = g02 - g20, w2 = g10 - g01;
      const double ww = w0 * w0 + w1 * w1 + w2 * w2;
      acc[A_VMAX] = fmax(acc[A_VMAX], sqrt(ww));
      const double IF = wA[a] * wA[b] * wA[c] / y[14];
      const double vv = pr.u * pr.u + pr.v * pr.v + pr.w * pr.w;
      acc[A_T] += IF * pr.T;
      acc[A_ENTROPY] += IF * (-gas.sKappaM1) * pr.rho * (log(pr.p) - gas.kappa * log(pr.rho));
      acc[A_EKIN] += IF * 0.5 * vv;
      acc[A_EKINC] += IF * 0.5 * pr.rho * vv;
      acc[A_ENSTR] += IF * 0.5 * pr.rho * ww;
      acc[A_DRU] += IF * (g00 * g00 + g10 * g10 + g20 * g20 + g01 * g01 + g11 * g11 + g21 * g21 + g02 * g02 +
                          g12 * g12 + g22 * g22);
      const double SS_off = 2.0 * (s01 * s01 + s02 * s02 + s12 * s12);
      acc[A_DRS] += IF * (g00 * g00 + g11 * g11 + g22 * g22 + SS_off);
      acc[A_DRSD] += IF * (sd00 * sd00 + sd11 * sd11 + sd22 * sd22 + SS_off);
      acc[A_DRP] += IF * pr.p * divU;
      acc[A_EDS] += IF * gas.mu0 * ww;
      acc[A_EDD] += IF * gas.mu0 * divU * divU;
    }
    __syncthreads();
  }
  const int wv = t >> 6;
#pragma unroll
  for (int q = 0; q < A_N; ++q) {
    const double r = q == A_VMAX ? wave_max(acc[q]) : wave_sum(acc[q]);
    if ((t & 63) == 0) red[wv * A_N + q] = r;
  }
  __syncthreads();
  if (t < A_N) {
    double r = red[t];
    for (int w = 1; w < 4; ++w) r = t == A_VMAX ? fmax(r, red[w * A_N + t]) : r + red[w * A_N + t];
    part[(size_t)A_N * e + t] = r;
  }
}

// Sum (maximum for the vorticity) of the per-element partials over the elements, in an order
// fixed by the launch shape.
__global__ __launch_bounds__(1024) void k_tgv_reduce(const double* __restrict__ part, int K, double* __restrict__ out) {
  __shared__ double red[16 * A_N];
  const int t = threadIdx.x;
  double acc[A_N];
#pragma unroll
  for (int q = 0; q < A_N; ++q) acc[q] = 0.0;
  for (int e = t; e < K; e += 1024)
#pragma unroll
    for (int q = 0; q < A_N; ++q)
      acc[q] = q == A_VMAX ? fmax(acc[q], part[(size_t)A_N * e + q]) : acc[q] + part[(size_t)A_N * e + q];
#pragma unroll
  for (int q = 0; q < A_N; ++q) {
    const double r = q == A_VMAX ? wave_max(acc[q]) : wave_sum(acc[q]);
    if ((t & 63) == 0) red[(t >> 6) * A_N + q] = r;
  }
  __syncthreads();
  if (t < A_N) {
    double r = red[t];
    for (int w = 1; w < 16; ++w) r = t == A_VMAX ? fmax(r, red[w * A_N + t]) : r + red[w * A_N + t];
    out[t] = r;
  }
}

template <int N>
hipError_t tgv_n(const DevMesh& m, const double* U, const void* G, int fp32, const double* tab, int na, const Gas& gas,
                 double* part, double* out, hipStream_t s) {
  if (fp32)
    hipLaunchKernelGGL((k_tgv_analyze<N, float>), dim3(m.nElems), dim3(256), 0, s, U, (const float*)G, m.sJ, tab, na,
                       gas, part);
  else
    hipLaunchKernelGGL((k_tgv_analyze<N, double>), dim3(m.nElems), dim3(256), 0, s, U, (const double*)G, m.sJ, tab,
                       na, gas, part);
  hipLaunchKernelGGL(k_tgv_reduce, dim3(1), dim3(1024), 0, s, part, m.nElems, out);
  return hipGetLastError();
}

}  // namespace

hipError_t launch_tgv_analyze(const DevMesh& m, const double* U, const void* G, int fp32, const double* tab, int na,
                              const Gas& gas, double* part, double* out, hipStream_t s) {
  if (na < 2 || na > NAMAX) return hipErrorInvalidValue;
  switch (m.N) {
    case 1: return tgv_n<1>(m, U, G, fp32, tab, na, gas, part, out, s);
    case 2: return tgv_n<2>(m, U, G, fp32, tab, na, gas, part, out, s);
    case 3: return tgv_n<3>(m, U, G, fp32, tab, na, gas, part, out, s);
    case 4: return tgv_n<4>(m, U, G, fp32, tab, na, gas, part, out, s);
    case 5: return tgv_n<5>(m, U, G, fp32, tab, na, gas, part, out, s);
    case 6: return tgv_n<6>(m, U, G, fp32, tab, na, gas, part, out, s);
    case 7: return tgv_n<7>(m, U, G, fp32, tab, na, gas, part, out, s);
    default: return hipErrorInvalidValue;
  }
}

}  // namespace fxg
