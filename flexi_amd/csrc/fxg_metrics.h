// Mesh metrics on the GPU: CalcMetrics/CalcSurfMetrics of FLEXI (src/mesh/metrics.f90:161-787),
// curl form, non-conforming trees excluded.
#pragma once
#include <hip/hip_runtime.h>

#include <string>

#include "fxg_kernels.h"

namespace fxg {

// NodeCoords(3,0:NGeo,0:NGeo,0:NGeo,nElems) on the device -> Metrics_{f,g,h}Tilde, sJ, Elem_xGP
// and, for master sides, NormVec/TangVec1/TangVec2/SurfElem. With SlaveNS (may be null), also the
// slave element's own n*SurfElem of every inner side point, (3,0:N,0:N,nSides), oriented as the
// master's normal. `m` supplies ElemToSide and S2V2. With `reference`, the metrics.f90 operation
// order (absolute coordinates, one-dimensional basis changes in sequence, no FMA contraction;
// SlaveNS must be null); otherwise coordinates relative to each element's first node. Returns
// nonzero and sets `err` on failure.
int compute_metrics_gpu(int N, int nodeType, int NGeo, int nElems, const double* NodeCoords, const DevMesh& m,
                        double* Mf, double* Mg, double* Mh, double* sJ, double* xGP, double* NormVec,
                        double* TangVec1, double* TangVec2, double* SurfElem, double* SlaveNS, bool reference,
                        hipStream_t s, std::string& err);

}  // namespace fxg
