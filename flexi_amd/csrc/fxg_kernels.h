// Launch interface between the host context (fxg_api.hip) and the gfx950 kernels.
#pragma once
#include <hip/hip_runtime.h>

#include "fxg_physics.h"

namespace fxg {

// Offsets (in doubles) of the 1D operators inside the device basis buffer; every block is
// (N+1)^2 or (N+1) long. Column-major like FLEXI: M(a,b) at a + (N+1)*b.
struct BasisLayout {
  int n;
  __host__ __device__ int D_T() const { return 0; }
  __host__ __device__ int D_Hat_T() const { return n * n; }
  __host__ __device__ int DVolSurf() const { return 2 * n * n; }
  __host__ __device__ int L_Minus() const { return 3 * n * n; }
  __host__ __device__ int L_Plus() const { return 3 * n * n + n; }
  __host__ __device__ int L_HatMinus() const { return 3 * n * n + 2 * n; }
  __host__ __device__ int L_HatPlus() const { return 3 * n * n + 3 * n; }
  __host__ __device__ int size() const { return 3 * n * n + 4 * n; }
};

// Device views of one context's arrays (FLEXI layouts unless noted).
struct DevMesh {
  int N, nElems, nSides;
  const int* E2S;    // ElemToSide(2,6,nElems), 1-based side ids
  const int* s2v2;   // S2V2(2,0:N,0:N,0:4,1:6)
  const int* v2s2;   // inverse of S2V2: volume-face (a,b) -> side point (p,q), same shape
  const double* Mf;  // Metrics_fTilde(3,0:N,0:N,0:N,nElems)
  const double* Mg;
  const double* Mh;
  const double* sJ;
  const double* NormVec;  // (3,0:N,0:N,nSides)
  const double* TangVec1;
  const double* TangVec2;
  const double* SurfElem;  // (0:N,0:N,nSides)
  const double* basis;        // BasisLayout
  // Gauss-Lobatto fused stage: 1 = each element takes n*SurfElem of a face from its own metric
  // terms at the face node (metrics computed by the library), 0 = the side's NormVec/SurfElem as
  // supplied by the host, shared by master and slave as FLEXI does
  int ownFaceMetric;
  // four-kernel path with metrics computed by the library: the slave element's own n*SurfElem of
  // each side point, (3,0:N,0:N,nSides), oriented as the master's normal; null otherwise
  const double* slaveNS = nullptr;
  // boundary sides 0..nBC-1 (FLEXI numbers them first): per side (BC type, RefState index or -1),
  // and RefStatePrim(6,nRefState)
  int nBC = 0;
  const int* bcInfo = nullptr;
  const double* refPrim = nullptr;
  // BCDataPrim(6,0:N,0:N,nBC) of the boundaries of types 12 and 121 (host-supplied), or null
  const double* bcData = nullptr;
  // face_slot of every own face point of every element, (6*(N+1)^2, nElems), or null (then
  // formed from ElemToSide and the inverse S2V2 map)
  const unsigned* slotTab = nullptr;
};

// Work arrays of one right-hand side evaluation. Face arrays of U are FP64 (U(5,p,q,side));
// gradient arrays hold T = double or float (FP32-bulk).
struct DevWork {
  double* Uf_m;    // (5,0:N,0:N,nSides)
  double* Uf_s;
  void* G;         // volume gradients: (0:N^3, 12, nElems), components gx(u,v,w,T),gy(..),gz(..)
  void* Gf_m;      // face gradients (12,0:N,0:N,nSides)
  void* Gf_s;
  double* Flux;    // Flux_master(5,0:N,0:N,nSides); Flux_slave equals it (fillflux.f90:199)
  double* FluxS = nullptr;  // with DevMesh::slaveNS: Flux_slave with the slave's own face metric
  double* bcLift;  // lifting flux of the boundary sides before SurfElem, (4,0:N,0:N,nBC)
  double* dtElem;  // (nElems)
  int* errElem;    // (nElems)
  double* dtOut;   // [2]: convective and viscous minima
  int* errOut;     // [1]
};

struct RKStage {
  int mode;        // 0: write Ut; 1: LSERK register update (first stage); 2: later stages
  double minusA;   // -RKA(iStage)
  double b_dt;     // RKb(iStage)*dt
  double* U;       // state, updated in place
  double* R;       // LSERK register Ut_tmp
  // fused path, optional: per-element CALCTIMESTEP spectral-radius sums of the new state
  // (2 per element: convective, viscous) and NaN-density flags
  double* lamSum = nullptr;
  int* lamNan = nullptr;
};

// Kernel ids for the optional event timer (fxg_profile_read).
// 0-3: four-kernel path; 6-8: fused Gauss-Lobatto split-form path.
enum KernelId {
  K_PROLONG = 0, K_LIFT = 1, K_FLUX = 2, K_VOLINT = 3, K_DT_ELEM = 4, K_DT_REDUCE = 5,
  K_FACES = 6, K_VISC_FACES = 7, K_DG_FUSED = 8, K_COUNT = 9
};

// Records a HIP event before (start = true) and after each kernel launch when timing is on.
struct KernelTimer {
  virtual void mark(int id, bool start, hipStream_t s) = 0;
  virtual ~KernelTimer() = default;
};

// Runs prolongation, lifting, fluxes and the volume/surface integrals for one RHS.
// Ut may be null when stage.mode != 0. `timer` may be null.
hipError_t launch_rhs(const DevMesh& m, const DevWork& w, const Gas& gas, int nodeType, int split,
                      int riemann, int fp32, const double* U, double* Ut, const RKStage& stage,
                      hipStream_t s, KernelTimer* timer);
// Work arrays of the fused Gauss-Lobatto split-form stage (fxg_fused.hip). Face arrays are
// indexed ((side*2 + slot)*(N+1)^2 + pq) with slot 0 = master, 1 = slave.
struct FusedWork {
  double* FU[2];  // face states, 5 per side point and slot, double-buffered across stages
  void* FG;       // viscous normal face fluxes f_v.(n SurfElem), 4 per side point and slot
  double* Jsoa;   // metric terms of the boundary nodes, rows of NB per element: Ja^1..3 x,y,z, sJ
  // per boundary side point, (4,0:N,0:N,nBC): walls the pressure (PRESSURE_RIEMANN) first, Riemann-type
  // boundaries the boundary state's velocity in 1..3
  double* bcP = nullptr;
  void* gBC = nullptr;    // lifted gradients at the boundary side points, 12 per point (gradient type)
};
// Doubles of FusedWork::Jsoa.
size_t packed_metrics_size(int N, int nElems);
// Face states of U into a face array laid out like FusedWork::FU (Gauss-Lobatto).
hipError_t launch_faces_of(const DevMesh& m, const double* U, double* FU, hipStream_t s);
// Fills FusedWork::Jsoa (boundary-node metric terms) from the FLEXI-layout metric arrays of `m`.
hipError_t launch_pack_metrics(const DevMesh& m, double* Jsoa, hipStream_t s);
// Fills slot(6*(N+1)^2, nElems) with face_slot of every own face point (m.slotTab must be null).
hipError_t launch_slot_table(const DevMesh& m, unsigned* slot, hipStream_t s);
bool fused_supported(int N, int nodeType, int split, long long nSides);
// One stage of the fused path. FU[cur] must hold the face states of U unless faces_fresh is
// false, in which case they are rebuilt first. In RK modes the new state's faces go to FU[1-cur].
// G (optional, (np,12,nElems) of the gradient type) receives the lifted gradients.
// Kernel A of the fused stage (fxg_visc.hip): viscous normal face fluxes into FG (float if fp32).
hipError_t launch_visc_faces(const DevMesh& m, bool fp32, const double* U, const double* Jsoa, const double* FU,
                             void* FG, const Gas& gas, void* gBC, const double* bcP, hipStream_t s);
// Boundary sides on the fused path (fxg_visc.hip): the outer slot of each boundary side point gets a
// face state whose BR1 lifting difference with the inner state is the boundary lifting flux (and,
// for Riemann-type sides, whose Riemann flux is the boundary flux); bcP the walls' pressure and the
// Riemann-type boundary states' velocity.
hipError_t launch_bc_ghost(const DevMesh& m, double* FU, double* bcP, const Gas& gas, hipStream_t s);
hipError_t launch_rhs_fused(const DevMesh& m, const FusedWork& w, const Gas& gas, int riemann, int fp32,
                            const double* U, double* Ut, const RKStage& st, bool faces_fresh, int cur,
                            void* G, hipStream_t s, KernelTimer* timer);

// Scratch of the discrete adjoint of the Gauss-Lobatto split-form right-hand side
// (fxg_adjoint.hip); face arrays use the FusedWork::FU slot layout.
struct AdjointWork {
  double* FUa;    // face states of the linearisation state, 5 per side slot
  double* XU[2];  // adjoint parts of the neighbour's face node (Riemann flux and lifting), 5 per
                  // side slot; two, so that a stage can read the previous stage's while writing
  double* Xg;     // adjoint of the neighbour's viscous normal face flux, 4 per side slot
};
// A previous reverse LSERK stage whose register update is still to be applied (XU == nullptr: none).
struct Pending {
  const double* XU;
  double cub, a;
};
// Reverse LSERK stages launched back to back: the register update of a stage is applied by the
// first kernel of the next one (or by launch_adjoint_flush).
struct AdjointChain {
  int cur = 0;
  bool pending = false;
  double cub = 0.0, a = 0.0;
};
// out = (dUt/dU)^T lam at U (FLEXI layout, FP64, Gauss-Lobatto split form, N 2..7).
hipError_t launch_rhs_vjp(const DevMesh& m, const AdjointWork& w, const Gas& gas, int riemann, const double* U,
                          const double* lam, double* out, hipStream_t s);
// One reverse LSERK stage at the stage state U: lam = rbar + cub*ubar; ubar += (dUt/dU)^T lam;
// rbar = a*lam. `FU`: the face states of U in the fused path's face layout (nullptr: computed
// into w.FUa). `out` is a state-sized scratch array that must stay untouched until the chain's
// next stage or flush; ubar and rbar are complete only after launch_adjoint_flush.
hipError_t launch_adjoint_stage(const DevMesh& m, const AdjointWork& w, const Gas& gas, int riemann, const double* U,
                                const double* FU, double* rbar, double* ubar, double cub, double a, double* out,
                                AdjointChain& chain, hipStream_t s);
hipError_t launch_adjoint_flush(const DevMesh& m, const AdjointWork& w, double* rbar, double* ubar,
                                const double* out, AdjointChain& chain, hipStream_t s);
// y += a*x over n doubles
hipError_t launch_axpy(size_t n, double a, const double* x, double* y, hipStream_t s);

hipError_t launch_calc_timestep(const DevMesh& m, const DevWork& w, const Gas& gas, const double* U,
                                double CFL, double DFL, hipStream_t s, KernelTimer* timer);
// CALCTIMESTEP from the spectral-radius sums a fused last stage stored for the current state.
hipError_t launch_calc_timestep_from_sums(const DevMesh& m, const DevWork& w, const double* lamSum, const int* lamNan,
                                          double CFL, double DFL, hipStream_t s, KernelTimer* timer);
// Copies a (12, np, nElems) gradient array of type float/double into three FLEXI-layout
// FP64 arrays (4,0:N,0:N,0:N,nElems).
hipError_t launch_unpack_gradients(const DevMesh& m, const void* G, int fp32, double* gx, double* gy,
                                   double* gz, hipStream_t s);
// Lifted density gradient rhoG(np, 3, nElems) of U (PP_OPTLIFT=0's LIFT_DENS), through the face
// states written into w.Uf_m/Uf_s (overwritten).
hipError_t launch_lift_density(const DevMesh& m, const DevWork& w, const Gas& gas, int nodeType, const double* U,
                               double* rhoG, hipStream_t s);
// Direction `dir` of the lifted gradients in FLEXI's layout (nVar,0:N,0:N,0:N,nElems): nVar 4
// (u,v,w,T) from G, or 5 with rho first from rhoG.
hipError_t launch_unpack_gradient_dir(const DevMesh& m, const void* G, int fp32, const double* rhoG, int nVar, int dir,
                                      double* out, hipStream_t s);

// Taylor-Green vortex analysis (AnalyzeTestcase): per-element partial sums part(13, nElems) of
// the integrals at na^3 Gauss-Lobatto points, then their sum (maximum for the vorticity) over
// the elements into out[13]: T, entropy, Ekin, Ekin_comp, enstrophy, DR_u, DR_S, DR_Sd, DR_p,
// ED_S, ED_D (unnormalised), max |vorticity|, domain volume. tab holds V(na, n) row-major
// (Lagrange polynomials of the solution nodes at the analysis points), wA(na), wGP(n).
constexpr int TGV_NANALYZE_MAX = 31;
constexpr int TGV_NSUMS = 13;
// interpolation table of the analysis: V(na, n), wA(na), wGP(n), at most this many doubles (N <= 7)
constexpr int TGV_TAB_MAX = (TGV_NANALYZE_MAX + 1) * (8 + 1) + 8;

hipError_t launch_tgv_analyze(const DevMesh& m, const double* U, const void* G, int fp32, const double* tab, int na,
                              const Gas& gas, double* part, double* out, hipStream_t s);
bool supported_degree(int N);

}  // namespace fxg
