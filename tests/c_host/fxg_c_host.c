/* A C host of the fxg library, as a FLEXI-like driver would use it: HOPR mesh -> context with the
 * metrics computed on the GPU, restart from a FLEXI State file, nSteps TimeStepByLSERKW2 steps
 * (CarpenterRK4-5, fixed dt), State file out. Only the C ABI (include/fxg.h, include/fxg_io.h).
 *
 *   fxg_c_host <mesh.h5> <N> <nSteps> <dt> <state_in.h5> <state_out.h5> [<CFLScale> <DFLScale>]
 *
 * With CFLScale and DFLScale, each step's dt comes from CALCTIMESTEP first, as in FLEXI's time
 * loop (src/timedisc/timedisc.f90:160-212), and <dt> is not used.
 *
 * Scheme: Gauss-Lobatto, split form PI, RoeEntropyFix, BR1, kappa 1.4, R 71.42857, Pr 0.72,
 * mu0 6.25e-4 (FLEXI's tgv/split regression parameters). Prints one line "ok t=<t> rho_sum=<sum>"
 * and returns 0, or prints the library's message and returns 1. */
#include <stdio.h>
#include <stdlib.h>

#include "fxg.h"
#include "fxg_io.h"

static int fail_fxg(const char* what) {
  fprintf(stderr, "%s: %s\n", what, fxg_last_error());
  return 1;
}
static int fail_io(const char* what) {
  fprintf(stderr, "%s: %s\n", what, fxg_io_last_error());
  return 1;
}

int main(int argc, char** argv) {
  if (argc != 7 && argc != 9) {
    fprintf(stderr, "usage: %s mesh.h5 N nSteps dt state_in.h5 state_out.h5 [CFLScale DFLScale]\n", argv[0]);
    return 2;
  }
  const char* meshFile = argv[1];
  const int N = atoi(argv[2]), nSteps = atoi(argv[3]);
  const int cfl = argc == 9;
  const double cflScale = cfl ? atof(argv[7]) : 0.0, dflScale = cfl ? atof(argv[8]) : 0.0;
  double dt = atof(argv[4]);
  int K = 0, S = 0, NGeo = 0;
  if (fxg_io_mesh_size(meshFile, &K, &S, &NGeo)) return fail_io("mesh size");
  const int ng = NGeo + 1, n = N + 1;
  int* s2e = malloc(sizeof(int) * 5 * (size_t)S);
  int* e2s = malloc(sizeof(int) * 12 * (size_t)K);
  double* nodes = malloc(sizeof(double) * 3 * (size_t)ng * ng * ng * K);
  double* U = malloc(sizeof(double) * 5 * (size_t)n * n * n * K);
  if (!s2e || !e2s || !nodes || !U) return 1;
  if (fxg_io_read_mesh(meshFile, K, S, NGeo, s2e, e2s, nodes)) return fail_io("read mesh");

  fxg_params p = {N, FXG_NODE_GAUSS_LOBATTO, FXG_SPLITDG_PI, FXG_RIEMANN_ROEENTROPYFIX,
                  1.4, 71.42857, 0.72, 6.25e-4, 0, 0, FXG_PREC_FP64, FXG_METRICS_OWN};
  fxg_ctx* ctx = NULL;
  if (fxg_create_from_nodes(&p, K, S, s2e, e2s, NGeo, nodes, &ctx)) return fail_fxg("create");

  double t = 0.0;
  char nodeType[64], stateMesh[256];
  if (fxg_io_read_state(argv[5], 5, N, K, U, &t, nodeType, sizeof nodeType, stateMesh, sizeof stateMesh))
    return fail_io("read state");
  if (fxg_set_state(ctx, U)) return fail_fxg("set state");

  /* CarpenterRK4-5 (src/timedisc/timedisc_vars.f90:200-214): RKA(2:5), RKb(1:5), RKc(2:5) */
  const double RKA[4] = {567301805773.0 / 1357537059087.0, 2404267990393.0 / 2016746695238.0,
                         3550918686646.0 / 2091501179385.0, 1275806237668.0 / 842570457699.0};
  const double RKb[5] = {1432997174477.0 / 9575080441755.0, 5161836677717.0 / 13612068292357.0,
                         1720146321549.0 / 2090206949498.0, 3134564353537.0 / 4481467310338.0,
                         2277821191437.0 / 14882151754819.0};
  const double RKc[4] = {1432997174477.0 / 9575080441755.0, 2526269341429.0 / 6820363962896.0,
                         2006345519317.0 / 3224310063776.0, 2802321613138.0 / 2924317926251.0};
  for (int it = 0; it < nSteps; ++it) {
    if (cfl) {
      int errType = 0;
      if (fxg_calc_timestep(ctx, cflScale, dflScale, &dt, &errType, NULL)) return fail_fxg("CALCTIMESTEP");
    }
    if (fxg_timestep_lserkw2(ctx, t, dt, 5, RKA, RKb, RKc)) return fail_fxg("time step");
    t += dt;
    if (cfl) printf("dt=%.17e\n", dt);
  }
  if (fxg_get_state(ctx, U)) return fail_fxg("get state");
  if (fxg_io_write_state(argv[6], "c_host", meshFile, nodeType, N, K, t, U)) return fail_io("write state");
  double rs = 0.0;
  for (size_t i = 0; i < (size_t)n * n * n * K; ++i) rs += U[5 * i];
  printf("ok t=%.17g rho_sum=%.17g\n", t, rs);
  fxg_destroy(ctx);
  free(s2e); free(e2s); free(nodes); free(U);
  return 0;
}
