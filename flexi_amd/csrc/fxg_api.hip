// Host side of libfxg: context set-up, the DG basis, device memory and the C ABI of include/fxg.h.
#include <hip/hip_runtime.h>

#include <algorithm>
#include <cmath>
#include <cstdio>
#include <cstdlib>
#include <cstring>
#include <mutex>
#include <string>
#include <vector>

#include "../../include/fxg.h"
#include "fxg_basis.h"
#include "fxg_kernels.h"
#include "fxg_metrics.h"

namespace {

thread_local std::string g_err;

int fail(const std::string& msg) {
  g_err = msg;
  return 1;
}

#define HIPCHECK(call)                                                                  \
  do {                                                                                  \
    hipError_t e_ = (call);                                                             \
    if (e_ != hipSuccess) return fail(std::string(#call) + ": " + hipGetErrorString(e_)); \
  } while (0)

// Operators of InitDGBasis (src/dg/dg.f90:132-209) in the BasisLayout order.
std::vector<double> build_basis(int N, int nodeType) {
  const int n = N + 1;
  std::vector<double> x(n), w(n), wb(n), D(n * n), Lm(n), Lp(n);
  fxg::basis::nodes(N, nodeType == FXG_NODE_GAUSS ? fxg::basis::GAUSS : fxg::basis::GAUSS_LOBATTO, x.data(), w.data());
  fxg::basis::bary_weights(N, x.data(), wb.data());
  fxg::basis::diff_matrix(N, x.data(), D.data());
  fxg::basis::lagrange_at(-1.0, N, x.data(), wb.data(), Lm.data());
  fxg::basis::lagrange_at(1.0, N, x.data(), wb.data(), Lp.data());
  fxg::BasisLayout B{n};
  std::vector<double> out(B.size(), 0.0);
  for (int a = 0; a < n; ++a)
    for (int b = 0; b < n; ++b) {
      out[B.D_T() + a + n * b] = D[b + n * a];
      // D_Hat = -M^{-1} D^T M  ->  D_Hat(b,a) = -D(a,b) w(a)/w(b);  D_Hat_T(a,b) = D_Hat(b,a)
      out[B.D_Hat_T() + a + n * b] = -(1.0 / w[b]) * (D[a + n * b] * w[a]);
      out[B.DVolSurf() + a + n * b] = D[b + n * a];
    }
  if (nodeType == FXG_NODE_GAUSS_LOBATTO) {
    // corner terms of the split-form operator cancel for Gauss-Lobatto (dg.f90:196-201)
    out[B.DVolSurf() + 0] += 1.0 / (2.0 * w[0]);
    out[B.DVolSurf() + N + n * N] -= 1.0 / (2.0 * w[N]);
  }
  for (int a = 0; a < n; ++a) {
    out[B.L_Minus() + a] = Lm[a];
    out[B.L_Plus() + a] = Lp[a];
    out[B.L_HatMinus() + a] = (1.0 / w[a]) * Lm[a];
    out[B.L_HatPlus() + a] = (1.0 / w[a]) * Lp[a];
  }
  return out;
}

// Side-to-volume index map S2V2 (src/mesh/mappings.f90:188-446) and its inverse.
void flip_s2m(int N, int p, int q, int flip, int& a, int& b) {
  switch (flip) {
    case 0: a = p; b = q; break;
    case 1: a = q; b = p; break;
    case 2: a = N - p; b = q; break;
    case 3: a = N - q; b = N - p; break;
    default: a = p; b = N - q; break;
  }
}
void cgns_s2v2(int N, int p, int q, int ls, int& a, int& b) {
  switch (ls) {
    case 5: a = q; b = p; break;
    case 3: a = p; b = q; break;
    case 2: a = p; b = q; break;
    case 4: a = N - p; b = q; break;
    case 1: a = q; b = p; break;
    default: a = p; b = q; break;
  }
}
void build_maps(int N, std::vector<int>& s2v2, std::vector<int>& v2s2) {
  const int n = N + 1, nf = n * n;
  s2v2.assign(2 * nf * 30, 0);
  v2s2.assign(2 * nf * 30, 0);
  for (int ls = 1; ls <= 6; ++ls)
    for (int f = 0; f <= 4; ++f)
      for (int q = 0; q <= N; ++q)
        for (int p = 0; p <= N; ++p) {
          int pp, qq, a, b;
          flip_s2m(N, p, q, f, pp, qq);
          cgns_s2v2(N, pp, qq, ls, a, b);
          const int o = nf * (f + 5 * (ls - 1));
          s2v2[2 * (p + n * q + o)] = a;
          s2v2[2 * (p + n * q + o) + 1] = b;
          v2s2[2 * (a + n * b + o)] = p;
          v2s2[2 * (a + n * b + o) + 1] = q;
        }
}

// Event pairs around every kernel launch, read back after the timed region.
struct EventTimer : fxg::KernelTimer {
  std::vector<hipEvent_t> pool;
  std::vector<std::pair<int, size_t>> marks;  // (kernel id, event index)
  hipError_t err = hipSuccess;
  void mark(int id, bool start, hipStream_t s) override {
    (void)start;
    hipEvent_t ev;
    if (marks.size() < pool.size()) {
      ev = pool[marks.size()];
    } else {
      hipError_t e = hipEventCreate(&ev);
      if (e != hipSuccess) { err = e; return; }
      pool.push_back(ev);
    }
    hipError_t e = hipEventRecord(ev, s);
    if (e != hipSuccess) err = e;
    marks.emplace_back(id, marks.size());
  }
  ~EventTimer() override {
    for (auto ev : pool) (void)hipEventDestroy(ev);
  }
};

template <class T>
int dev_alloc(T** p, size_t count) {
  if (count == 0) count = 1;
  hipError_t e = hipMalloc((void**)p, count * sizeof(T));
  if (e != hipSuccess) return fail(std::string("hipMalloc: ") + hipGetErrorString(e));
  return 0;
}

}  // namespace

struct fxg_ctx {
  fxg_params p{};
  int device = 0;
  int N = 0, n = 0, nf = 0, np = 0, K = 0, S = 0;
  hipStream_t stream = nullptr;
  fxg::Gas gas{};
  fxg::DevMesh m{};
  fxg::DevWork w{};
  double *U = nullptr, *Ut = nullptr, *R = nullptr;
  // owned device arrays
  int *E2S = nullptr, *s2v2 = nullptr, *v2s2 = nullptr;
  double *Mf = nullptr, *Mg = nullptr, *Mh = nullptr, *sJ = nullptr, *NV = nullptr, *T1 = nullptr, *T2 = nullptr,
         *SE = nullptr, *basis = nullptr, *xGP = nullptr;
  double *dtHost = nullptr;
  int* errHost = nullptr;
  std::vector<void*> owned;
  EventTimer* timer = nullptr;  // non-null while kernel timing is enabled
  // fused Gauss-Lobatto split-form path: face states double-buffered across LSERK stages
  bool fused = false;
  fxg::FusedWork fw{};
  int cur = 0;
  // the state whose face values FU[cur] holds (only trusted for the context's own state, which
  // changes only through fxg_set_state and the library's own updates)
  const double* faces_of = nullptr;
  // Ut and the lifted gradients (G) of the context's last right-hand side. After an LSERK step on
  // the context's state that is the last stage's, at the state before the last update,
  // U - RKb_last*dt*R: the fused stages do not store Ut or the gradients, the four-kernel stages
  // not Ut, so they are recomputed there on request (rhs_stale). The device-pointer calls
  // overwrite G with the gradients of the caller's state and also mark them stale.
  bool rhs_stale = false;
  int last_rhs = 0;  // 0: none since creation or fxg_set_state; 1: a right-hand side of U; 2: an LSERK step of U
  double last_b_dt = 0.0;
  double* Uprev = nullptr;
  // the state whose lifted gradients G holds (c->U or Uprev), null when unknown
  const double* grads_state = nullptr;
  // fxg_get_gradients: one direction in FLEXI's layout (5 per node) and the density gradient
  double* gradOut = nullptr;
  double* rhoG = nullptr;
  // testcase analysis: per-element partial sums, the interpolation tables and the 13 sums
  // (allocated at the first fxg_tgv_analyze)
  double* anaPart = nullptr;
  double* anaTab = nullptr;
  double* anaOut = nullptr;
  // CALCTIMESTEP spectral-radius sums per element, computed by a fused last stage for lam_of
  double* lamSum = nullptr;
  int* lamNan = nullptr;
  const double* lam_of = nullptr;
  // discrete adjoint scratch, allocated on first use
  fxg::AdjointWork aw{};
  double* adjTmp = nullptr;
  // boundary sides 1..nBC: their conditions must be set (fxg_set_boundaries) before a right-hand side
  int nBC = 0;
  bool bcSet = false;
  // types 12/121 take BCDataPrim(6,0:N,0:N,nBC) from the host (fxg_set_boundary_data)
  bool bcNeedsData = false;
  double* bcData = nullptr;
  int* bcInfo = nullptr;
  double* refPrim = nullptr;
};

namespace {

int check_params(const fxg_params* p) {
  if (!p) return fail("params is NULL");
  if (!fxg::supported_degree(p->N)) return fail("N must be between 1 and 7");
  if (p->nodeType != FXG_NODE_GAUSS && p->nodeType != FXG_NODE_GAUSS_LOBATTO) return fail("unknown nodeType");
  if (p->splitDG != FXG_SPLITDG_NONE && p->splitDG != FXG_SPLITDG_PI) return fail("SplitDG must be 0 (weak) or PI (4)");
  if (p->splitDG == FXG_SPLITDG_PI && p->nodeType != FXG_NODE_GAUSS_LOBATTO)
    return fail("the split form is built for Gauss-Lobatto nodes only");
  if (p->riemann != FXG_RIEMANN_LF && p->riemann != FXG_RIEMANN_ROE && p->riemann != FXG_RIEMANN_ROEENTROPYFIX)
    return fail("Riemann must be LF (1), Roe (3) or RoeEntropyFix (33)");
  if (p->doWeakLifting || p->doConservativeLifting)
    return fail("only the strong, non-conservative BR1 lifting is built (doWeakLifting=F, doConservativeLifting=F)");
  if (p->precision != FXG_PREC_FP64 && p->precision != FXG_PREC_FP32_BULK) return fail("unknown precision");
  if (p->metrics != FXG_METRICS_OWN && p->metrics != FXG_METRICS_REFERENCE)
    return fail("metrics must be FXG_METRICS_OWN (0) or FXG_METRICS_REFERENCE (1)");
  if (!(p->kappa > 1.0) || !(p->R > 0.0) || !(p->Pr > 0.0) || !(p->mu0 >= 0.0)) return fail("invalid gas constants");
  return 0;
}

// Connectivity must be a conforming mesh: every side has one master (flip 0) and one slave,
// consistent between SideToElem and ElemToSide, except the boundary sides, which have a master
// only (SideToElem(2,:) = -1) and come first, 1..nBC, as setLocalSideIDs numbers them
// (prepare_mesh.f90:196-222). Checked on the host so that no kernel can index out of bounds.
int check_connectivity(int K, int S, const int* S2E, const int* E2S, int* nBCOut = nullptr) {
  if (K <= 0 || S <= 0) return fail("nElems and nSides must be positive");
  if (!S2E || !E2S) return fail("SideToElem/ElemToSide is NULL");
  int nBC = 0;
  while (nBC < S && S2E[5 * nBC + 1] == -1) ++nBC;
  std::vector<int> seen(2 * (size_t)S, 0);
  for (int e = 0; e < K; ++e)
    for (int ls = 1; ls <= 6; ++ls) {
      const int sid = E2S[2 * (ls - 1) + 12 * e], flip = E2S[2 * (ls - 1) + 1 + 12 * e];
      if (sid < 1 || sid > S) return fail("ElemToSide side id out of range");
      if (flip < 0 || flip > 4) return fail("ElemToSide flip out of range");
      const int* se = S2E + 5 * (sid - 1);
      if (flip == 0) {
        if (se[0] != e + 1 || se[2] != ls) return fail("ElemToSide/SideToElem master mismatch");
        seen[2 * (sid - 1)]++;
      } else {
        if (se[1] != e + 1 || se[3] != ls || se[4] != flip) return fail("ElemToSide/SideToElem slave mismatch");
        seen[2 * (sid - 1) + 1]++;
      }
    }
  for (int s = 0; s < S; ++s)
    if (seen[2 * s] != 1 || seen[2 * s + 1] != (s < nBC ? 0 : 1))
      return fail("every inner side needs exactly one master and one slave, every boundary side (the first "
                  "sides, SideToElem(2,:) = -1) one master");
  if (nBCOut) *nBCOut = nBC;
  return 0;
}

template <class T>
int upload(fxg_ctx* c, T** dst, const T* src, size_t count) {
  if (dev_alloc(dst, count)) return 1;
  c->owned.push_back(*dst);
  if (src) {
    hipError_t e = hipMemcpy(*dst, src, count * sizeof(T), hipMemcpyHostToDevice);
    if (e != hipSuccess) return fail(std::string("hipMemcpy: ") + hipGetErrorString(e));
  }
  return 0;
}

template <class T>
int alloc_owned(fxg_ctx* c, T** dst, size_t count) { return upload<T>(c, dst, nullptr, count); }

// ---------------------------------------------------------------------------------------------
// Scratch of fxg_calc_timestep_dev_noalloc: a file-scope device array, so that entry point
// allocates nothing. Each GPU holds its own copy until the process ends; its address on a device
// is looked up once, with that device current, when the first context on it is set up.
// Layout (doubles): a region of 2*kNoAllocMaxElems for dtConv[K] followed by dtVisc[K] (K the
// context's elements), dtOut[2], then the ints errElem[kNoAllocMaxElems], errOut[1]. Every call
// writes each entry it reads (k_dt_elem writes all
// nElems per-element entries, k_dt_reduce the three outputs), so nothing needs clearing; calls that
// overlap on one GPU share the array and race.
constexpr int kNoAllocMaxElems = 1 << 19;
constexpr size_t kNoAllocInts = (size_t)kNoAllocMaxElems + 1;
constexpr size_t kNoAllocDoubles = 2 * (size_t)kNoAllocMaxElems + 2 + (kNoAllocInts + 1) / 2;
static __device__ __align__(256) double scratch[kNoAllocDoubles];
constexpr int kMaxDevices = 64;
std::mutex g_scratch_mu;
double* g_scratch_ptr[kMaxDevices] = {};

// The device address of `scratch` on `device`, which must be current.
int scratch_init(int device) {
  if (device < 0 || device >= kMaxDevices) return fail("device ordinal out of range for the scratch table");
  std::lock_guard<std::mutex> lk(g_scratch_mu);
  if (g_scratch_ptr[device]) return 0;
  double* ptr = nullptr;
  HIPCHECK(hipGetSymbolAddress((void**)&ptr, HIP_SYMBOL(scratch)));
  if (!ptr) return fail("hipGetSymbolAddress(scratch) returned NULL");
  g_scratch_ptr[device] = ptr;
  return 0;
}

__global__ void k_dt_store(const double* __restrict__ dtOut, const int* __restrict__ errOut, double* __restrict__ dt,
                           int* __restrict__ errType) {
  dt[0] = fmin(dtOut[0], dtOut[1]);
  errType[0] = errOut[0];
}

// Allocations and tables shared by both creation paths. Geometry arrays are allocated but filled
// by the caller.
int ctx_setup(fxg_ctx* c, const fxg_params* p, int K, int S, const int* S2E, const int* E2S, int nBC) {
  c->p = *p;
  c->nBC = nBC;
  HIPCHECK(hipGetDevice(&c->device));
  if (scratch_init(c->device)) return 1;
  c->N = p->N; c->n = p->N + 1; c->nf = c->n * c->n; c->np = c->nf * c->n; c->K = K; c->S = S;
  HIPCHECK(hipStreamCreateWithFlags(&c->stream, hipStreamNonBlocking));
  const double kappaM1 = p->kappa - 1.0;
  const double cp = p->R * p->kappa / kappaM1;
  c->gas = fxg::Gas{p->kappa, kappaM1, 1.0 / kappaM1, p->R, 1.0 / p->R, p->mu0, p->mu0 * cp / p->Pr, 4.0 / 3.0, 2.0 / 3.0,
                    std::fmax(4.0 / 3.0, p->kappa / p->Pr)};
  const size_t vol = (size_t)c->np * K, face = (size_t)c->nf * S;
  std::vector<int> s2v2, v2s2;
  build_maps(c->N, s2v2, v2s2);
  std::vector<double> basis = build_basis(c->N, p->nodeType);
  if (upload(c, &c->E2S, E2S, (size_t)12 * K) || upload(c, &c->s2v2, s2v2.data(), s2v2.size()) ||
      upload(c, &c->v2s2, v2s2.data(), v2s2.size()) || upload(c, &c->basis, basis.data(), basis.size()))
    return 1;
  if (alloc_owned(c, &c->Mf, 3 * vol) || alloc_owned(c, &c->Mg, 3 * vol) || alloc_owned(c, &c->Mh, 3 * vol) ||
      alloc_owned(c, &c->sJ, vol) || alloc_owned(c, &c->NV, 3 * face) || alloc_owned(c, &c->T1, 3 * face) ||
      alloc_owned(c, &c->T2, 3 * face) || alloc_owned(c, &c->SE, face) || alloc_owned(c, &c->xGP, 3 * vol))
    return 1;
  // Uprev: the last stage's input state, formed when Ut or the gradients of a step are requested
  if (alloc_owned(c, &c->U, 5 * vol) || alloc_owned(c, &c->Ut, 5 * vol) || alloc_owned(c, &c->R, 5 * vol) ||
      alloc_owned(c, &c->Uprev, 5 * vol))
    return 1;
  HIPCHECK(hipMemsetAsync(c->U, 0, 5 * vol * sizeof(double), c->stream));
  HIPCHECK(hipMemsetAsync(c->Ut, 0, 5 * vol * sizeof(double), c->stream));
  HIPCHECK(hipMemsetAsync(c->R, 0, 5 * vol * sizeof(double), c->stream));
  const size_t gsz = p->precision == FXG_PREC_FP32_BULK ? sizeof(float) : sizeof(double);
  char *G, *Gm, *Gs;
  if (alloc_owned(c, &c->w.Uf_m, 5 * face) || alloc_owned(c, &c->w.Uf_s, 5 * face) ||
      alloc_owned(c, &G, 12 * vol * gsz) || alloc_owned(c, &Gm, 12 * face * gsz) || alloc_owned(c, &Gs, 12 * face * gsz) ||
      alloc_owned(c, &c->w.Flux, 5 * face) || alloc_owned(c, &c->w.dtElem, 2 * (size_t)K) ||
      alloc_owned(c, &c->w.errElem, (size_t)K) || alloc_owned(c, &c->w.dtOut, 2) || alloc_owned(c, &c->w.errOut, 1) ||
      alloc_owned(c, &c->w.bcLift, 4 * (size_t)c->nf * nBC))
    return 1;
  c->w.G = G; c->w.Gf_m = Gm; c->w.Gf_s = Gs;
  HIPCHECK(hipMemsetAsync(G, 0, 12 * vol * gsz, c->stream));
  // FXG_PIPELINE=classic keeps the four-kernel path for cases the fused one covers
  const char* pipe = std::getenv("FXG_PIPELINE");
  c->fused = fxg::fused_supported(c->N, p->nodeType, p->splitDG == FXG_SPLITDG_PI, S) &&
             !(pipe && std::strcmp(pipe, "classic") == 0);
  if (c->fused) {
    char* FG;
    if (alloc_owned(c, &c->fw.FU[0], 10 * face) || alloc_owned(c, &c->fw.FU[1], 10 * face) ||
        alloc_owned(c, &FG, 8 * face * gsz) || alloc_owned(c, &c->fw.Jsoa, fxg::packed_metrics_size(c->N, K)) ||
        alloc_owned(c, &c->lamSum, 2 * (size_t)K) || alloc_owned(c, &c->lamNan, (size_t)K) ||
        alloc_owned(c, &c->fw.bcP, 4 * (size_t)c->nf * nBC))
      return 1;
    char* gBC;
    if (alloc_owned(c, &gBC, 12 * (size_t)c->nf * nBC * gsz)) return 1;
    c->fw.gBC = gBC;
    c->fw.FG = FG;
  }
  HIPCHECK(hipHostMalloc((void**)&c->dtHost, 2 * sizeof(double), hipHostMallocDefault));
  HIPCHECK(hipHostMalloc((void**)&c->errHost, sizeof(int), hipHostMallocDefault));
  c->m = fxg::DevMesh{c->N, K, S, c->E2S, c->s2v2, c->v2s2, c->Mf, c->Mg, c->Mh, c->sJ,
                      c->NV, c->T1, c->T2, c->SE, c->basis, 0};
  return 0;
}

int finish_setup(fxg_ctx* c) {
  if (c->fused) HIPCHECK(fxg::launch_pack_metrics(c->m, c->fw.Jsoa, c->stream));
  if (c->fused) {  // the fused kernels read the side slots of their face points from this table
    unsigned* slot = nullptr;
    if (alloc_owned(c, &slot, (size_t)6 * c->nf * c->K)) return 1;
    HIPCHECK(fxg::launch_slot_table(c->m, slot, c->stream));
    c->m.slotTab = slot;
  }
  HIPCHECK(hipStreamSynchronize(c->stream));
  return 0;
}

void ctx_free(fxg_ctx* c) {
  if (!c) return;
  delete c->timer;
  for (void* ptr : c->owned) (void)hipFree(ptr);
  if (c->dtHost) (void)hipHostFree(c->dtHost);
  if (c->errHost) (void)hipHostFree(c->errHost);
  if (c->stream) (void)hipStreamDestroy(c->stream);
  delete c;
}

// One right-hand side (st.mode 0) or LSERK stage. On the fused path `faces_fresh` says that
// FU[cur] already holds the face states of U (written by the previous stage).
int rhs(fxg_ctx* c, const double* U, double* Ut, const fxg::RKStage& st, hipStream_t s, bool faces_fresh = false) {
  if (c->nBC > 0 && !c->bcSet) return fail("the mesh has boundary sides: set their conditions with fxg_set_boundaries");
  if (c->bcNeedsData && !c->bcData)
    return fail("boundary types 12/121 need their states: set BCDataPrim with fxg_set_boundary_data");
  if (c->fused) {
    HIPCHECK(fxg::launch_rhs_fused(c->m, c->fw, c->gas, c->p.riemann, c->p.precision == FXG_PREC_FP32_BULK, U, Ut, st,
                                   faces_fresh, c->cur, st.mode == 0 ? c->w.G : nullptr, s, c->timer));
    if (st.mode != 0) c->cur ^= 1;
    c->faces_of = U;  // FU[cur] now holds the face values of U (updated in place in RK modes)
    return 0;
  }
  HIPCHECK(fxg::launch_rhs(c->m, c->w, c->gas, c->p.nodeType, c->p.splitDG == FXG_SPLITDG_PI, c->p.riemann,
                           c->p.precision == FXG_PREC_FP32_BULK, U, Ut, st, s, c->timer));
  return 0;
}

// One LSERK stage (st = 0 .. nStages-1) of TimeStepByLSERKW2 on (U, R).
int lserk_stage(fxg_ctx* c, double dt, int st, const double* RKA, const double* RKb, double* U, double* R,
                hipStream_t s) {
  fxg::RKStage stage{st == 0 ? 1 : 2, st == 0 ? 0.0 : -RKA[st - 1], RKb[st] * dt, U, R};
  return rhs(c, U, nullptr, stage, s, st > 0);
}

// One TimeStepByLSERKW2 step. `faces_known`: FU[cur] already holds the face values of U (the
// previous step of the same call, or of the context's own state), so the first stage reuses them.
// `want_lambda`: the fused last stage also stores CALCTIMESTEP's per-element spectral radii of the
// new state (used by the next fxg_calc_timestep on the context's state).
int lserk(fxg_ctx* c, double dt, int nStages, const double* RKA, const double* RKb, double* U, double* R,
          hipStream_t s, bool faces_known = false, bool want_lambda = false) {
  if (nStages < 1) return fail("nStages must be positive");
  if (!RKb || (nStages > 1 && !RKA)) return fail("RK coefficient arrays are NULL");
  for (int st = 0; st < nStages; ++st) {
    fxg::RKStage stage{st == 0 ? 1 : 2, st == 0 ? 0.0 : -RKA[st - 1], RKb[st] * dt, U, R};
    if (want_lambda && c->fused && st == nStages - 1) {
      stage.lamSum = c->lamSum;
      stage.lamNan = c->lamNan;
    }
    if (rhs(c, U, nullptr, stage, s, st > 0 || faces_known)) return 1;
  }
  return 0;
}

int check_rk(int nStages, const double* RKA, const double* RKb) {
  if (nStages < 1) return fail("nStages must be positive");
  if (!RKb || (nStages > 1 && !RKA)) return fail("RK coefficient arrays are NULL");
  return 0;
}

int ensure_adjoint(fxg_ctx* c) {
  if (!c->fused || c->p.precision != FXG_PREC_FP64 || c->nBC > 0)
    return fail("the discrete adjoint is built for the FP64 Gauss-Lobatto split form (N 2..7) on meshes without "
                "boundary sides");
  if (c->aw.FUa) return 0;
  const size_t face = (size_t)c->nf * c->S, vol = (size_t)c->np * c->K;
  if (alloc_owned(c, &c->aw.FUa, 10 * face) || alloc_owned(c, &c->aw.XU[0], 10 * face) ||
      alloc_owned(c, &c->aw.XU[1], 10 * face) || alloc_owned(c, &c->aw.Xg, 8 * face) ||
      alloc_owned(c, &c->adjTmp, 5 * vol))
    return 1;
  return 0;
}

// Device buffers of one adjoint run, released on every exit path.
struct DevBuffers {
  std::vector<void*> p;
  ~DevBuffers() {
    for (void* x : p) (void)hipFree(x);
  }
  double* get(size_t count) {
    double* d = nullptr;
    if (hipMalloc((void**)&d, count * sizeof(double)) != hipSuccess) return nullptr;
    p.push_back(d);
    return d;
  }
};

// Ut and gradients of the context's last right-hand side when they are not in Ut and G
// (rhs_stale): after an LSERK step the last stage's input was U_{s-1} = U - RKb_s*dt*R (to
// round-off), and one right-hand side there restores them; after a right-hand side of U, or with
// none yet, the right-hand side of U.
int refresh_last_rhs(fxg_ctx* c) {
  if (!c->rhs_stale) return 0;
  const double* X = c->U;
  if (c->last_rhs == 2) {
    const size_t vol5 = (size_t)5 * c->np * c->K;
    HIPCHECK(hipMemcpyAsync(c->Uprev, c->U, vol5 * sizeof(double), hipMemcpyDeviceToDevice, c->stream));
    HIPCHECK(fxg::launch_axpy(vol5, -c->last_b_dt, c->R, c->Uprev, c->stream));
    X = c->Uprev;
  }
  fxg::RKStage st{0, 0.0, 0.0, nullptr, nullptr};
  if (rhs(c, X, c->Ut, st, c->stream)) return 1;
  c->rhs_stale = false;
  c->grads_state = X;
  return 0;
}

}  // namespace

extern "C" {

const char* fxg_last_error(void) { return g_err.c_str(); }
const char* fxg_version(void) { return "fxg 0.1 (gfx950)"; }

int fxg_create(const fxg_params* p, const fxg_mesh* mesh, fxg_ctx** out) {
  if (!out) return fail("ctx output pointer is NULL");
  *out = nullptr;
  if (check_params(p)) return 1;
  if (!mesh) return fail("mesh is NULL");
  int nBC = 0;
  if (check_connectivity(mesh->nElems, mesh->nSides, mesh->SideToElem, mesh->ElemToSide, &nBC)) return 1;
  if (mesh->firstInnerSide != nBC + 1 || mesh->lastInnerSide != mesh->nSides)
    return fail("side ranges: boundary sides 1..nBCSides, then inner sides up to nSides (no mortar or MPI sides); "
                "firstInnerSide must be nBCSides+1 and lastInnerSide nSides");
  if (!mesh->Metrics_fTilde || !mesh->Metrics_gTilde || !mesh->Metrics_hTilde || !mesh->sJ || !mesh->NormVec ||
      !mesh->TangVec1 || !mesh->TangVec2 || !mesh->SurfElem)
    return fail("a metric array is NULL");
  fxg_ctx* c = new fxg_ctx;
  if (ctx_setup(c, p, mesh->nElems, mesh->nSides, mesh->SideToElem, mesh->ElemToSide, nBC)) { ctx_free(c); return 1; }
  const size_t vol = (size_t)c->np * c->K, face = (size_t)c->nf * c->S;
  struct { double* d; const double* h; size_t n; } cp[] = {
      {c->Mf, mesh->Metrics_fTilde, 3 * vol}, {c->Mg, mesh->Metrics_gTilde, 3 * vol}, {c->Mh, mesh->Metrics_hTilde, 3 * vol},
      {c->sJ, mesh->sJ, vol}, {c->NV, mesh->NormVec, 3 * face}, {c->T1, mesh->TangVec1, 3 * face},
      {c->T2, mesh->TangVec2, 3 * face}, {c->SE, mesh->SurfElem, face}};
  for (auto& x : cp) {
    hipError_t e = hipMemcpy(x.d, x.h, x.n * sizeof(double), hipMemcpyHostToDevice);
    if (e != hipSuccess) { fail(std::string("hipMemcpy: ") + hipGetErrorString(e)); ctx_free(c); return 1; }
  }
  if (finish_setup(c)) { ctx_free(c); return 1; }
  *out = c;
  return 0;
}

int fxg_create_from_nodes(const fxg_params* p, int nElems, int nSides, const int* S2E, const int* E2S, int NGeo,
                          const double* NodeCoords, fxg_ctx** out) {
  if (!out) return fail("ctx output pointer is NULL");
  *out = nullptr;
  if (check_params(p)) return 1;
  int nBC = 0;
  if (check_connectivity(nElems, nSides, S2E, E2S, &nBC)) return 1;
  if (NGeo < 1 || NGeo > 8) return fail("NGeo must be between 1 and 8");
  if (!NodeCoords) return fail("NodeCoords is NULL");
  fxg_ctx* c = new fxg_ctx;
  if (ctx_setup(c, p, nElems, nSides, S2E, E2S, nBC)) { ctx_free(c); return 1; }
  const size_t ng = (size_t)(NGeo + 1) * (NGeo + 1) * (NGeo + 1);
  double* dNC = nullptr;
  if (upload(c, &dNC, NodeCoords, 3 * ng * nElems)) { ctx_free(c); return 1; }
  std::string err;
  const bool own = p->metrics == FXG_METRICS_OWN;
  // four-kernel path: each side of a face takes its own face metric too (the slave's in slaveNS,
  // its side fluxes in FluxS), as the fused path does with its boundary nodes
  double* slaveNS = nullptr;
  if (own && !c->fused) {
    const size_t face = (size_t)c->nf * nSides;
    if (alloc_owned(c, &slaveNS, 3 * face) || alloc_owned(c, &c->w.FluxS, 5 * face)) { ctx_free(c); return 1; }
  }
  if (fxg::compute_metrics_gpu(c->N, p->nodeType, NGeo, nElems, dNC, c->m, c->Mf, c->Mg, c->Mh, c->sJ, c->xGP, c->NV,
                               c->T1, c->T2, c->SE, slaveNS, !own, c->stream, err)) {
    fail(err); ctx_free(c); return 1;
  }
  c->m.slaveNS = slaveNS;
  // FXG_METRICS_OWN: the curl-form face metric depends only on the face, so each element's own
  // boundary metric is the face metric; using it on both sides keeps the freestream exact to
  // flux round-off. FXG_METRICS_REFERENCE: the master's NormVec/SurfElem on both sides, as FLEXI.
  c->m.ownFaceMetric = own ? 1 : 0;
  if (finish_setup(c)) { ctx_free(c); return 1; }
  *out = c;
  return 0;
}

int fxg_destroy(fxg_ctx* c) {
  if (!c) return 0;
  (void)hipSetDevice(c->device);
  (void)hipStreamSynchronize(c->stream);
  ctx_free(c);
  return 0;
}

int fxg_set_boundaries(fxg_ctx* c, int nBCs, const int* BC, const int* BoundaryType, int nRefState,
                       const double* RefStatePrim) {
  if (!c) return fail("ctx is NULL");
  if (c->nBC == 0) return 0;
  if (!BC || !BoundaryType || nBCs < 1) return fail("BC and BoundaryType are required for a mesh with boundary sides");
  if (nRefState < 0 || (nRefState > 0 && !RefStatePrim)) return fail("RefStatePrim is NULL");
  std::vector<int> info(2 * (size_t)c->nBC);
  bool needsData = false;
  for (int s = 0; s < c->nBC; ++s) {
    const int iBC = BC[s];
    if (iBC < 1 || iBC > nBCs) return fail("BC(SideID) out of range 1..nBCs");
    const int type = BoundaryType[iBC - 1], state = BoundaryType[iBC - 1 + nBCs];
    const bool known = type == 2 || type == 3 || type == 4 || type == 9 || type == 91 || type == 12 || type == 121 ||
                       type == 23 || type == 24 || type == 25 || type == 27;
    if (!known)
      return fail("boundary type " + std::to_string(type) + " is not supported (2, 3, 4, 9, 12, 23, 24, 25, 27, 91 "
                  "and 121 are; 22, an exact function of time, is evaluated by the host as 121 at each call time)");
    // the types that read RefStatePrim(:,BCState) (InitBC, getboundaryflux.f90:64-218)
    const bool usesRef = type == 2 || type == 4 || type == 23 || type == 24 || type == 25 || type == 27;
    if (usesRef && (state < 1 || state > nRefState))
      return fail("boundary type " + std::to_string(type) + " needs a RefState index 1..nRefState as its state "
                  "(state 0, an exact function, is not supported)");
    const double* ref = usesRef ? RefStatePrim + 6 * (state - 1) : nullptr;
    if (type == 4 && !(ref[5] > 0.0)) return fail("an isothermal wall needs a positive wall temperature");
    if ((type == 24 || type == 25) && !(ref[4] > 0.0)) return fail("an outflow boundary needs a positive pressure");
    if (type == 23 && !(ref[1] >= 0.0)) return fail("an outflow Mach number must not be negative");
    if (type == 27) {
      const double a2 = ref[1] * ref[1] + ref[2] * ref[2] + ref[3] * ref[3];
      if (!(ref[0] > 0.0) || !(ref[4] > 0.0) || !(std::fabs(a2 - 1.0) < 1e-12))
        return fail("a subsonic inflow (27) needs RefStatePrim = (Tt>0, unit direction a(1:3), pt>0), as InitBC "
                    "leaves it");
    }
    needsData = needsData || type == 12 || type == 121;
    info[2 * s] = type;
    info[2 * s + 1] = usesRef ? state - 1 : -1;
  }
  HIPCHECK(hipSetDevice(c->device));
  HIPCHECK(hipStreamSynchronize(c->stream));
  if (!c->bcInfo && alloc_owned(c, &c->bcInfo, info.size())) return 1;
  HIPCHECK(hipMemcpy(c->bcInfo, info.data(), info.size() * sizeof(int), hipMemcpyHostToDevice));
  if (c->refPrim) {
    (void)hipFree(c->refPrim);
    c->owned.erase(std::remove(c->owned.begin(), c->owned.end(), (void*)c->refPrim), c->owned.end());
    c->refPrim = nullptr;
  }
  if (upload(c, &c->refPrim, nRefState > 0 ? RefStatePrim : nullptr, 6 * (size_t)std::max(1, nRefState))) return 1;
  c->m.nBC = c->nBC;
  c->m.bcInfo = c->bcInfo;
  c->m.refPrim = c->refPrim;
  c->bcNeedsData = needsData;
  c->bcSet = true;
  return 0;
}

int fxg_set_boundary_data(fxg_ctx* c, const double* BCDataPrim) {
  if (!c) return fail("ctx is NULL");
  if (c->nBC == 0) return 0;
  if (!BCDataPrim) return fail("BCDataPrim is NULL");
  const size_t n = 6 * (size_t)c->nf * c->nBC;
  for (size_t x = 0; x < n; x += 6)
    if (!(BCDataPrim[x] > 0.0) || !(BCDataPrim[x + 4] > 0.0) || !(BCDataPrim[x + 5] > 0.0))
      return fail("BCDataPrim needs positive density, pressure and temperature at every boundary point");
  HIPCHECK(hipSetDevice(c->device));
  HIPCHECK(hipStreamSynchronize(c->stream));
  if (!c->bcData && alloc_owned(c, &c->bcData, n)) return 1;
  HIPCHECK(hipMemcpy(c->bcData, BCDataPrim, n * sizeof(double), hipMemcpyHostToDevice));
  c->m.bcData = c->bcData;
  return 0;
}

int fxg_sizes(fxg_ctx* c, int* N, int* K, int* S) {
  if (!c) return fail("ctx is NULL");
  if (N) *N = c->N;
  if (K) *K = c->K;
  if (S) *S = c->S;
  return 0;
}

int fxg_set_state(fxg_ctx* c, const double* U) {
  if (!c || !U) return fail("NULL argument");
  c->faces_of = nullptr;
  c->rhs_stale = false;
  c->last_rhs = 0;
  c->grads_state = nullptr;
  c->lam_of = nullptr;
  HIPCHECK(hipMemcpyAsync(c->U, U, 5 * (size_t)c->np * c->K * sizeof(double), hipMemcpyHostToDevice, c->stream));
  HIPCHECK(hipStreamSynchronize(c->stream));
  return 0;
}

int fxg_get_state(fxg_ctx* c, double* U) {
  if (!c || !U) return fail("NULL argument");
  HIPCHECK(hipMemcpyAsync(U, c->U, 5 * (size_t)c->np * c->K * sizeof(double), hipMemcpyDeviceToHost, c->stream));
  HIPCHECK(hipStreamSynchronize(c->stream));
  return 0;
}

int fxg_get_ut(fxg_ctx* c, double* Ut) {
  if (!c || !Ut) return fail("NULL argument");
  if (refresh_last_rhs(c)) return 1;
  HIPCHECK(hipMemcpyAsync(Ut, c->Ut, 5 * (size_t)c->np * c->K * sizeof(double), hipMemcpyDeviceToHost, c->stream));
  HIPCHECK(hipStreamSynchronize(c->stream));
  return 0;
}

int fxg_get_gradients(fxg_ctx* c, int nVarLifting, double* gx, double* gy, double* gz) {
  if (!c || !gx || !gy || !gz) return fail("NULL argument");
  if (nVarLifting != 4 && nVarLifting != 5)
    return fail("fxg_get_gradients: nVarLifting must be 4 (PP_OPTLIFT=1) or 5 (PP_OPTLIFT=0)");
  if (refresh_last_rhs(c)) return 1;
  const size_t nodes = (size_t)c->np * c->K;
  if (!c->gradOut && alloc_owned(c, &c->gradOut, 5 * nodes)) return 1;
  if (nVarLifting == 5) {
    if (!c->grads_state)
      return fail("fxg_get_gradients: no right-hand side has been evaluated for the current state");
    if (!c->rhoG && alloc_owned(c, &c->rhoG, 3 * nodes)) return 1;
    HIPCHECK(fxg::launch_lift_density(c->m, c->w, c->gas, c->p.nodeType, c->grads_state, c->rhoG, c->stream));
  }
  double* host[3] = {gx, gy, gz};
  for (int d = 0; d < 3; ++d) {
    HIPCHECK(fxg::launch_unpack_gradient_dir(c->m, c->w.G, c->p.precision == FXG_PREC_FP32_BULK, c->rhoG, nVarLifting,
                                             d, c->gradOut, c->stream));
    HIPCHECK(hipMemcpyAsync(host[d], c->gradOut, nVarLifting * nodes * sizeof(double), hipMemcpyDeviceToHost,
                            c->stream));
    // the next direction reuses gradOut
    HIPCHECK(hipStreamSynchronize(c->stream));
  }
  return 0;
}

}  // extern "C"

namespace {

// AnalyzeTestcase on the device with the given scratch: part(TGV_NSUMS, K) partial sums,
// tab(TGV_TAB_MAX) the interpolation table, sums(TGV_NSUMS) the results. Allocates nothing itself
// (the refresh of a step's gradients uses the context's Uprev, allocated at creation).
int tgv_analyze_impl(fxg_ctx* c, int NAnalyze, double* out, double* part, double* tabDev, double* sumsDev) {
  if (NAnalyze < 1 || NAnalyze > fxg::TGV_NANALYZE_MAX)
    return fail("fxg_tgv_analyze: NAnalyze must be in 1.." + std::to_string(fxg::TGV_NANALYZE_MAX));
  if (refresh_last_rhs(c)) return 1;
  const int n = c->n, na = NAnalyze + 1;
  // V(na, n): Lagrange polynomials of the solution nodes at the NAnalyze Gauss-Lobatto points
  // (Vdm_GaussN_NAnalyze, analyze.f90:258-264), then their weights and the solution weights
  double x[8], w[8], wb[8], xa[fxg::TGV_NANALYZE_MAX + 1], wa[fxg::TGV_NANALYZE_MAX + 1], L[8];
  double tab[fxg::TGV_TAB_MAX];
  fxg::basis::nodes(c->N, c->p.nodeType == FXG_NODE_GAUSS ? fxg::basis::GAUSS : fxg::basis::GAUSS_LOBATTO, x, w);
  fxg::basis::bary_weights(c->N, x, wb);
  fxg::basis::nodes(NAnalyze, fxg::basis::GAUSS_LOBATTO, xa, wa);
  for (int a = 0; a < na; ++a) {
    fxg::basis::lagrange_at(xa[a], c->N, x, wb, L);
    for (int i = 0; i < n; ++i) tab[(size_t)a * n + i] = L[i];
    tab[(size_t)na * n + a] = wa[a];
  }
  for (int i = 0; i < n; ++i) tab[(size_t)na * n + na + i] = w[i];
  const size_t ntab = (size_t)na * n + na + n;
  double sums[fxg::TGV_NSUMS];
  HIPCHECK(hipMemcpyAsync(tabDev, tab, ntab * sizeof(double), hipMemcpyHostToDevice, c->stream));
  HIPCHECK(fxg::launch_tgv_analyze(c->m, c->U, c->w.G, c->p.precision == FXG_PREC_FP32_BULK, tabDev, na, c->gas,
                                   part, sumsDev, c->stream));
  HIPCHECK(hipMemcpyAsync(sums, sumsDev, sizeof(sums), hipMemcpyDeviceToHost, c->stream));
  HIPCHECK(hipStreamSynchronize(c->stream));
  // normalisation (testcase.f90:443-459), rho0 = 1 (testcase_vars.f90:35)
  const double rho0 = 1.0, Vol = sums[12], mu0 = c->gas.mu0;
  const double Ekin = sums[2] / Vol, DR_S = sums[6] * 2.0 * mu0 / (rho0 * Vol),
               DR_Sd = sums[7] * 2.0 * mu0 / (rho0 * Vol), DR_p = -sums[8] / (rho0 * Vol);
  const double v[15] = {DR_S, DR_Sd + DR_p, Ekin, sums[3] / Vol / rho0, sums[4] / (rho0 * Vol), sums[5] * mu0 / Vol,
                        DR_S, DR_Sd, DR_p, sums[11], sums[0] / Vol, std::sqrt(2.0 / 3.0 * Ekin), sums[1] / Vol,
                        sums[9] / Vol, sums[10] * 4.0 / 3.0 / Vol};
  for (int q = 0; q < 15; ++q) out[q] = v[q];
  return 0;
}

}  // namespace

extern "C" {

int fxg_tgv_analyze(fxg_ctx* c, int NAnalyze, double* out) {
  if (!c || !out) return fail("NULL argument");
  if (!c->anaPart && (alloc_owned(c, &c->anaPart, (size_t)fxg::TGV_NSUMS * c->K) ||
                      alloc_owned(c, &c->anaTab, (size_t)fxg::TGV_TAB_MAX) ||
                      alloc_owned(c, &c->anaOut, (size_t)fxg::TGV_NSUMS)))
    return 1;
  return tgv_analyze_impl(c, NAnalyze, out, c->anaPart, c->anaTab, c->anaOut);
}

int fxg_get_metrics(fxg_ctx* c, double* Mf, double* Mg, double* Mh, double* sJ, double* NV, double* T1, double* T2,
                    double* SE, double* xGP) {
  if (!c) return fail("ctx is NULL");
  const size_t vol = (size_t)c->np * c->K, face = (size_t)c->nf * c->S;
  struct { double* h; const double* d; size_t n; } cp[] = {
      {Mf, c->Mf, 3 * vol}, {Mg, c->Mg, 3 * vol}, {Mh, c->Mh, 3 * vol}, {sJ, c->sJ, vol}, {NV, c->NV, 3 * face},
      {T1, c->T1, 3 * face}, {T2, c->T2, 3 * face}, {SE, c->SE, face}, {xGP, c->xGP, 3 * vol}};
  for (auto& x : cp)
    if (x.h) HIPCHECK(hipMemcpyAsync(x.h, x.d, x.n * sizeof(double), hipMemcpyDeviceToHost, c->stream));
  HIPCHECK(hipStreamSynchronize(c->stream));
  return 0;
}

int fxg_dg_time_derivative(fxg_ctx* c, double t) {
  (void)t;  // no source terms, and the boundary states (RefStates, walls) do not depend on time
  if (!c) return fail("ctx is NULL");
  fxg::RKStage st{0, 0.0, 0.0, nullptr, nullptr};
  c->rhs_stale = false;
  c->last_rhs = 1;
  c->grads_state = nullptr;
  if (rhs(c, c->U, c->Ut, st, c->stream)) return 1;
  c->grads_state = c->U;
  HIPCHECK(hipStreamSynchronize(c->stream));
  return 0;
}

int fxg_timestep_lserkw2(fxg_ctx* c, double t, double dt, int nStages, const double* RKA, const double* RKb,
                         const double* RKc) {
  (void)t;
  (void)RKc;  // stage times only enter time-dependent sources or boundary states, which this path has none of
  if (!c) return fail("ctx is NULL");
  c->lam_of = nullptr;
  c->grads_state = nullptr;
  if (lserk(c, dt, nStages, RKA, RKb, c->U, c->R, c->stream, c->fused && c->faces_of == c->U, true)) return 1;
  // Ut of the last stage is not kept by either pipeline, the gradients not by the fused one
  c->rhs_stale = true;
  c->last_rhs = 2;
  if (c->fused) c->lam_of = c->U;
  c->last_b_dt = RKb[nStages - 1] * dt;
  HIPCHECK(hipStreamSynchronize(c->stream));
  return 0;
}

int fxg_dg_time_derivative_dev(fxg_ctx* c, double t, const double* U, double* Ut, void* stream) {
  (void)t;
  if (!c || !U || !Ut) return fail("NULL argument");
  fxg::RKStage st{0, 0.0, 0.0, nullptr, nullptr};
  const int rc = rhs(c, U, Ut, st, stream ? (hipStream_t)stream : c->stream);
  c->faces_of = nullptr;  // the caller's buffer may change before the next call
  // G now holds the gradients of the caller's state: the context's are recomputed on request
  c->rhs_stale = true;
  c->grads_state = nullptr;
  return rc;
}

int fxg_timestep_lserkw2_dev(fxg_ctx* c, double t, double dt, int nStages, const double* RKA, const double* RKb,
                             const double* RKc, double* U, double* R, void* stream) {
  (void)t;
  (void)RKc;
  if (!c || !U || !R) return fail("NULL argument");
  const int rc = lserk(c, dt, nStages, RKA, RKb, U, R, stream ? (hipStream_t)stream : c->stream);
  c->faces_of = nullptr;
  if (!c->fused) {  // the four-kernel stages wrote the caller's gradients into G
    c->rhs_stale = true;
    c->grads_state = nullptr;
  }
  return rc;
}

const char* fxg_pipeline(fxg_ctx* c) { return !c ? "" : (c->fused ? "fused" : "classic"); }

int fxg_sync(fxg_ctx* c) {
  if (!c) return fail("ctx is NULL");
  HIPCHECK(hipStreamSynchronize(c->stream));
  return 0;
}

int fxg_profile_enable(fxg_ctx* c, int enable) {
  if (!c) return fail("ctx is NULL");
  HIPCHECK(hipStreamSynchronize(c->stream));
  delete c->timer;
  c->timer = enable ? new EventTimer : nullptr;
  return 0;
}

int fxg_profile_read(fxg_ctx* c, int kernel, double* total_ms, int* launches) {
  if (!c || !total_ms || !launches) return fail("NULL argument");
  if (!c->timer) return fail("kernel timing is not enabled");
  if (c->timer->err != hipSuccess) return fail(std::string("event timer: ") + hipGetErrorString(c->timer->err));
  HIPCHECK(hipStreamSynchronize(c->stream));
  double tot = 0.0;
  int cnt = 0;
  const auto& mk = c->timer->marks;
  for (size_t i = 0; i + 1 < mk.size(); i += 2) {
    if (mk[i].first != kernel) continue;
    float ms = 0.f;
    HIPCHECK(hipEventElapsedTime(&ms, c->timer->pool[mk[i].second], c->timer->pool[mk[i + 1].second]));
    tot += ms;
    ++cnt;
  }
  *total_ms = tot;
  *launches = cnt;
  return 0;
}

int fxg_rhs_vjp_dev(fxg_ctx* c, double t, const double* U, const double* lam, double* out, void* stream) {
  (void)t;
  if (!c || !U || !lam || !out) return fail("NULL argument");
  if (ensure_adjoint(c)) return 1;
  HIPCHECK(fxg::launch_rhs_vjp(c->m, c->aw, c->gas, c->p.riemann, U, lam, out,
                               stream ? (hipStream_t)stream : c->stream));
  return 0;
}

int fxg_adjoint_storage(fxg_ctx* c, int nSteps, int nStages, int checkpointEvery, size_t* doubles) {
  if (!c || !doubles) return fail("NULL argument");
  if (nSteps < 0 || nStages < 1 || checkpointEvery < 0) return fail("invalid nSteps, nStages or checkpointEvery");
  const size_t vol5 = (size_t)5 * c->np * c->K, face10 = (size_t)10 * c->nf * c->S;
  if (checkpointEvery == 0)
    *doubles = (size_t)nSteps * nStages * (vol5 + face10);
  else
    *doubles = (size_t)((nSteps + checkpointEvery - 1) / checkpointEvery) * vol5;
  return 0;
}

int fxg_forward_n_steps(fxg_ctx* c, int nSteps, const double* dt, int nStages, const double* RKA, const double* RKb,
                        const double* RKc, double* U, double* ckpt, int checkpointEvery, void* stream) {
  (void)RKc;
  if (!c || !dt || !U) return fail("NULL argument");
  if (nSteps < 0) return fail("nSteps must not be negative");
  if (ckpt && checkpointEvery < 0) return fail("checkpointEvery must not be negative");
  if (check_rk(nStages, RKA, RKb)) return 1;
  const hipStream_t s = stream ? (hipStream_t)stream : c->stream;
  const size_t vol5 = (size_t)5 * c->np * c->K;
  if (ckpt && checkpointEvery == 0) {
    // trajectory: every stage's input state and its face states, written by the stages themselves
    // (each stage reads state x and writes state x+1 out of place, so storing costs no extra bytes)
    if (ensure_adjoint(c)) return 1;
    c->faces_of = nullptr;
    if (nSteps == 0) return 0;
    const size_t face10 = (size_t)10 * c->nf * c->S, nTot = (size_t)nSteps * nStages;
    double* S = ckpt;
    double* F = ckpt + nTot * vol5;
    HIPCHECK(hipMemcpyAsync(S, U, vol5 * sizeof(double), hipMemcpyDeviceToDevice, s));
    HIPCHECK(fxg::launch_faces_of(c->m, S, F, s));
    for (size_t x = 0; x < nTot; ++x) {
      const int it = (int)(x / nStages), st = (int)(x % nStages);
      const bool last = x + 1 == nTot;
      fxg::FusedWork fw = c->fw;
      fw.FU[0] = F + x * face10;
      fw.FU[1] = last ? c->fw.FU[0] : F + (x + 1) * face10;
      const fxg::RKStage stage{st == 0 ? 1 : 2, st == 0 ? 0.0 : -RKA[st - 1], RKb[st] * dt[it],
                               last ? U : S + (x + 1) * vol5, c->R};
      HIPCHECK(fxg::launch_rhs_fused(c->m, fw, c->gas, c->p.riemann, 0, S + x * vol5, nullptr, stage, true, 0, nullptr,
                                     s, c->timer));
    }
    return 0;
  }
  for (int it = 0; it < nSteps; ++it) {
    if (ckpt && it % checkpointEvery == 0)
      HIPCHECK(hipMemcpyAsync(ckpt + (size_t)(it / checkpointEvery) * vol5, U, vol5 * sizeof(double),
                              hipMemcpyDeviceToDevice, s));
    if (lserk(c, dt[it], nStages, RKA, RKb, U, c->R, s, it > 0 && c->fused)) {
      c->faces_of = nullptr;
      return 1;
    }
  }
  c->faces_of = nullptr;
  return 0;
}

int fxg_adjoint_n_steps(fxg_ctx* c, int nSteps, const double* dt, int nStages, const double* RKA, const double* RKb,
                        const double* RKc, const double* U0, const double* ckptIn, double* ubar, int checkpointEvery,
                        void* stream) {
  (void)RKc;
  if (!c || !dt || !U0 || !ubar) return fail("NULL argument");
  if (nSteps < 0) return fail("nSteps must not be negative");
  if (checkpointEvery < 0) return fail("checkpointEvery must not be negative");
  if (checkpointEvery == 0 && !ckptIn)
    return fail("checkpointEvery 0 reads the trajectory stored by fxg_forward_n_steps: ckpt_dev is required");
  if (check_rk(nStages, RKA, RKb)) return 1;
  if (ensure_adjoint(c)) return 1;
  c->faces_of = nullptr;  // the recomputation below runs on scratch states
  if (nSteps == 0) return 0;
  const hipStream_t s = stream ? (hipStream_t)stream : c->stream;
  const size_t vol5 = (size_t)5 * c->np * c->K, bytes = vol5 * sizeof(double);
  const size_t face10 = (size_t)10 * c->nf * c->S;  // both slots of every side, 5 per point
  if (checkpointEvery == 0) {
    // the stored trajectory: the reverse stages only, no recomputation
    const size_t nTot = (size_t)nSteps * nStages;
    const double* S = ckptIn;
    const double* F = ckptIn + nTot * vol5;
    DevBuffers buf;
    double* rbar = buf.get(vol5);
    if (!rbar) return fail("fxg_adjoint_n_steps: out of device memory");
    HIPCHECK(hipMemsetAsync(rbar, 0, bytes, s));
    fxg::AdjointChain chain;
    for (size_t x = nTot; x-- > 0;) {
      const int it = (int)(x / nStages), st = (int)(x % nStages);
      HIPCHECK(fxg::launch_adjoint_stage(c->m, c->aw, c->gas, c->p.riemann, S + x * vol5, F + x * face10, rbar, ubar,
                                         RKb[st] * dt[it], st == 0 ? 0.0 : -RKA[st - 1], c->adjTmp, chain, s));
    }
    HIPCHECK(fxg::launch_adjoint_flush(c->m, c->aw, rbar, ubar, c->adjTmp, chain, s));
    HIPCHECK(hipStreamSynchronize(s));
    return 0;
  }
  const int every = checkpointEvery < nSteps ? checkpointEvery : nSteps;
  const int nck = (nSteps + every - 1) / every;
  DevBuffers buf;
  const double* ck = ckptIn;
  // the stage states of one segment and their face states, written by the recomputation itself
  double* seg = buf.get((size_t)every * nStages * vol5);
  double* segF = buf.get((size_t)every * nStages * face10);
  double* Uw = buf.get(vol5);
  double* R = buf.get(vol5);
  double* rbar = buf.get(vol5);
  if (!seg || !segF || !Uw || !R || !rbar)
    return fail("fxg_adjoint_n_steps: out of device memory for the stage states");
  if (!ck) {
    // forward sweep with a checkpoint of U every `every` steps
    double* ckw = buf.get((size_t)nck * vol5);
    if (!ckw) return fail("fxg_adjoint_n_steps: out of device memory for checkpoints");
    HIPCHECK(hipMemcpyAsync(Uw, U0, bytes, hipMemcpyDeviceToDevice, s));
    for (int it = 0; it < nSteps; ++it) {
      if (it % every == 0)
        HIPCHECK(hipMemcpyAsync(ckw + (size_t)(it / every) * vol5, Uw, bytes, hipMemcpyDeviceToDevice, s));
      if (lserk(c, dt[it], nStages, RKA, RKb, Uw, R, s, it > 0)) return 1;
    }
    ck = ckw;
  }
  // reverse sweep, one segment at a time: recompute the stage states of the segment from its
  // checkpoint, then run the adjoint of each stage backwards (R_s = Ut_s - RKA_s R_{s-1},
  // U_s = U_{s-1} + RKb_s dt R_s):  rbar += RKb_s dt ubar;  ubar += (dUt/dU)^T rbar;
  // rbar *= -RKA_s (0 for the first stage)
  HIPCHECK(hipMemsetAsync(rbar, 0, bytes, s));
  // each reverse stage's register update is applied by the next stage's first kernel (the
  // pending update only needs the exchange arrays, so it also spans the recomputation of a segment)
  fxg::AdjointChain chain;
  for (int sg = nck - 1; sg >= 0; --sg) {
    const int s0 = sg * every, s1 = (s0 + every < nSteps) ? s0 + every : nSteps;
    const int nst = (s1 - s0) * nStages;
    HIPCHECK(hipMemcpyAsync(seg, ck + (size_t)sg * vol5, bytes, hipMemcpyDeviceToDevice, s));
    HIPCHECK(fxg::launch_faces_of(c->m, seg, segF, s));
    // stage x reads seg[x] and its faces segF[x] and writes the next stage state and its faces
    // to seg[x+1], segF[x+1] (the end state of the segment to scratch)
    for (int x = 0; x < nst; ++x) {
      const int it = s0 + x / nStages, st = x % nStages;
      const bool last = x + 1 == nst;
      fxg::FusedWork fw = c->fw;
      fw.FU[0] = segF + (size_t)x * face10;
      fw.FU[1] = last ? c->fw.FU[0] : segF + (size_t)(x + 1) * face10;
      const fxg::RKStage stage{st == 0 ? 1 : 2, st == 0 ? 0.0 : -RKA[st - 1], RKb[st] * dt[it],
                               last ? Uw : seg + (size_t)(x + 1) * vol5, R};
      HIPCHECK(fxg::launch_rhs_fused(c->m, fw, c->gas, c->p.riemann, 0, seg + (size_t)x * vol5, nullptr, stage, true,
                                     0, nullptr, s, c->timer));
    }
    for (int x = nst - 1; x >= 0; --x) {
      const int it = s0 + x / nStages, st = x % nStages;
      HIPCHECK(fxg::launch_adjoint_stage(c->m, c->aw, c->gas, c->p.riemann, seg + (size_t)x * vol5,
                                         segF + (size_t)x * face10, rbar, ubar, RKb[st] * dt[it],
                                         st == 0 ? 0.0 : -RKA[st - 1], c->adjTmp, chain, s));
    }
  }
  HIPCHECK(fxg::launch_adjoint_flush(c->m, c->aw, rbar, ubar, c->adjTmp, chain, s));
  HIPCHECK(hipStreamSynchronize(s));
  c->faces_of = nullptr;
  return 0;
}

int fxg_calc_timestep(fxg_ctx* c, double CFL, double DFL, double* dt, int* errType, double* dtElem) {
  if (!c || !dt || !errType) return fail("NULL argument");
  if (c->fused && c->lam_of == c->U)
    HIPCHECK(fxg::launch_calc_timestep_from_sums(c->m, c->w, c->lamSum, c->lamNan, CFL, DFL, c->stream, c->timer));
  else
    HIPCHECK(fxg::launch_calc_timestep(c->m, c->w, c->gas, c->U, CFL, DFL, c->stream, c->timer));
  HIPCHECK(hipMemcpyAsync(c->dtHost, c->w.dtOut, 2 * sizeof(double), hipMemcpyDeviceToHost, c->stream));
  HIPCHECK(hipMemcpyAsync(c->errHost, c->w.errOut, sizeof(int), hipMemcpyDeviceToHost, c->stream));
  if (dtElem) {
    // dtElem = MIN(convective, viscous) per element, as CALCTIMESTEP stores it
    std::vector<double> tmp(2 * (size_t)c->K);
    HIPCHECK(hipMemcpyAsync(tmp.data(), c->w.dtElem, 2 * (size_t)c->K * sizeof(double), hipMemcpyDeviceToHost, c->stream));
    HIPCHECK(hipStreamSynchronize(c->stream));
    for (int e = 0; e < c->K; ++e) dtElem[e] = std::fmin(tmp[e], tmp[c->K + e]);
  }
  HIPCHECK(hipStreamSynchronize(c->stream));
  *dt = std::fmin(c->dtHost[0], c->dtHost[1]);
  *errType = c->errHost[0];
  if (*errType != 0) return fail("CALCTIMESTEP: NaN in density (1), convective (2) or viscous (3) time step: " +
                                 std::to_string(*errType));
  return 0;
}

int fxg_calc_timestep_dev_noalloc(fxg_ctx* c, const double* U, double CFL, double DFL, double* dt, int* errType,
                                  void* stream) {
  if (!c || !U || !dt || !errType) return fail("NULL argument");
  if (c->K > kNoAllocMaxElems)
    return fail("fxg_calc_timestep_dev_noalloc: " + std::to_string(c->K) + " elements exceed its scratch maximum of " +
                std::to_string(kNoAllocMaxElems));
  double* p = (c->device >= 0 && c->device < kMaxDevices) ? g_scratch_ptr[c->device] : nullptr;
  if (!p) return fail("fxg_calc_timestep_dev_noalloc: no scratch address for the context's device");
  // the per-element and reduced values in the scratch array, in the layout described at `scratch`
  fxg::DevWork w{};
  w.dtElem = p;
  w.dtOut = p + 2 * (size_t)kNoAllocMaxElems;
  w.errElem = reinterpret_cast<int*>(w.dtOut + 2);
  w.errOut = w.errElem + kNoAllocMaxElems;
  const hipStream_t s = stream ? (hipStream_t)stream : c->stream;
  HIPCHECK(fxg::launch_calc_timestep(c->m, w, c->gas, U, CFL, DFL, s, nullptr));
  hipLaunchKernelGGL(k_dt_store, dim3(1), dim3(1), 0, s, w.dtOut, w.errOut, dt, errType);
  HIPCHECK(hipGetLastError());
  return 0;
}

}  // extern "C"
