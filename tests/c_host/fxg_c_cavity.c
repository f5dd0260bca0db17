/* FLEXI's parabolic/cavity_3D regression case driven from C through the C ABI only
 * (include/fxg.h, include/fxg_io.h), as a FLEXI-like host would run it: HOPR mesh with boundary
 * sides and the parameter file's BoundaryName/BoundaryType overrides, RefStates, boundary
 * conditions, initial state RefStateCons(IniRefState=2), then TimeDisc's loop (CALCTIMESTEP and
 * UpdateTimeStep's dt rules, timedisc_func.f90:215-268) with CarpenterRK4-5 to TEnd = 1, State
 * file out.
 *
 *   fxg_c_cavity <cavity4x4x4_mesh.h5> <state_out.h5>
 *
 * Parameters of the case's parameter.ini: N=2 on Gauss nodes, weak form, RoeEntropyFix, BR1,
 * kappa 1.4, R 1, Pr 0.72, mu0 0.01, CFLScale 0.99, DFLScale 0.4. Prints
 * "ok t=<t> steps=<n>" and returns 0, or prints the library's message and returns 1. */
#include <math.h>
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

/* UpdateTimeStep: the smallest of the stable step and the times to the next analysis and to TEnd;
 * a step leaving less than dt/100 before either is stretched to reach it. */
static double update_dt(double dt_min, double t, double t_an, double tend, int* fin) {
  const double dt_an = t_an - t, dt_end = tend - t;
  double dt = dt_min;
  if (dt_an < dt) dt = dt_an;
  if (dt_end < dt) dt = dt_end;
  *fin = dt == dt_end;
  dt = INFINITY;
  if (dt_min > 0 && dt_min < dt) dt = dt_min;
  if (dt_an > 0 && dt_an < dt) dt = dt_an;
  if (dt_end > 0 && dt_end < dt) dt = dt_end;
  if (dt_an - dt < dt / 100.0 && dt_an > 0) dt = dt_an;
  if (dt_end - dt < dt / 100.0 && dt_end > 0) {
    dt = dt_end;
    *fin = 1;
  }
  return dt;
}

int main(int argc, char** argv) {
  if (argc != 3) {
    fprintf(stderr, "usage: %s cavity4x4x4_mesh.h5 state_out.h5\n", argv[0]);
    return 2;
  }
  const char* meshFile = argv[1];
  const int N = 2, n = N + 1;
  const double kappa = 1.4, R = 1.0;
  /* BoundaryName / BoundaryType of the parameter file */
  const char* userNames[3] = {"BC_wall_left", "BC_wall_right", "BC_free"};
  const int userTypes[6] = {4, 1, 4, 1, 2, 1};
  int K = 0, S = 0, NGeo = 0, nBCSides = 0, nBCs = 0;
  if (fxg_io_mesh_size_bc(meshFile, 3, userNames, userTypes, &K, &S, &NGeo, &nBCSides, &nBCs))
    return fail_io("mesh size");
  const int ng = NGeo + 1;
  int* s2e = malloc(sizeof(int) * 5 * (size_t)S);
  int* e2s = malloc(sizeof(int) * 12 * (size_t)K);
  double* nodes = malloc(sizeof(double) * 3 * (size_t)ng * ng * ng * K);
  int* BC = malloc(sizeof(int) * (size_t)(nBCSides > 0 ? nBCSides : 1));
  int* btype = malloc(sizeof(int) * 3 * (size_t)(nBCs > 0 ? nBCs : 1));
  double* U = malloc(sizeof(double) * 5 * (size_t)n * n * n * K);
  if (!s2e || !e2s || !nodes || !BC || !btype || !U) return 1;
  if (fxg_io_read_mesh_bc(meshFile, 3, userNames, userTypes, K, S, NGeo, nBCSides, nBCs, s2e, e2s, nodes, BC, btype,
                          NULL, 0))
    return fail_io("read mesh");

  fxg_params p = {N, FXG_NODE_GAUSS, FXG_SPLITDG_NONE, FXG_RIEMANN_ROEENTROPYFIX, kappa, R, 0.72, 0.01, 0, 0,
                  FXG_PREC_FP64, FXG_METRICS_REFERENCE};
  fxg_ctx* ctx = NULL;
  if (fxg_create_from_nodes(&p, K, S, s2e, e2s, NGeo, nodes, &ctx)) return fail_fxg("create");

  /* RefStatePrim(6,2) from RefState (rho,u,v,w,p): T = p*(1/rho)/R (equation.f90:126-141) */
  const double ref[2][5] = {{1.0, 1.0, 0.0, 0.0, 71.4285714286}, {1.0, 0.0, 0.0, 0.0, 71.4285714286}};
  double refPrim[12];
  for (int r = 0; r < 2; ++r) {
    for (int v = 0; v < 5; ++v) refPrim[6 * r + v] = ref[r][v];
    refPrim[6 * r + 5] = ref[r][4] * (1.0 / ref[r][0]) / R;
  }
  if (fxg_set_boundaries(ctx, nBCs, BC, btype, 2, refPrim)) return fail_fxg("boundaries");

  /* IniExactFunc = 1: RefStateCons(IniRefState = 2) everywhere (PrimToCons, eos.f90:305-328) */
  const double* q = refPrim + 6;
  double c[5] = {q[0], q[1] * q[0], q[2] * q[0], q[3] * q[0], 0.0};
  c[4] = (1.0 / (kappa - 1.0)) * q[4] + 0.5 * (c[1] * q[1] + c[2] * q[2] + c[3] * q[3]);
  for (size_t i = 0; i < (size_t)n * n * n * K; ++i)
    for (int v = 0; v < 5; ++v) U[5 * i + v] = c[v];
  if (fxg_set_state(ctx, U)) return fail_fxg("set state");

  /* CarpenterRK4-5 (timedisc_vars.f90:175-214) and FillCFL_DFL (timedisc_func.f90:463-527) */
  const double RKA[4] = {567301805773.0 / 1357537059087.0, 2404267990393.0 / 2016746695238.0,
                         3550918686646.0 / 2091501179385.0, 1275806237668.0 / 842570457699.0};
  const double RKb[5] = {1432997174477.0 / 9575080441755.0, 5161836677717.0 / 13612068292357.0,
                         1720146321549.0 / 2090206949498.0, 3134564353537.0 / 4481467310338.0,
                         2277821191437.0 / 14882151754819.0};
  const double RKc[4] = {1432997174477.0 / 9575080441755.0, 2526269341429.0 / 6820363962896.0,
                         2006345519317.0 / 3224310063776.0, 2802321613138.0 / 2924317926251.0};
  double CFL = 0.99 * 1.7595;  /* CFLScaleAlpha(2), Gauss nodes */
  CFL = CFL / (2.0 * N + 1.0);
  double DFL = 0.4 * (0.76 * 1.853);  /* DFLScaleAlpha(2) * RelativeDFL */
  DFL = DFL / ((2.0 * N + 1.0) * (2.0 * N + 1.0));

  const double tend = 1.0;
  double t = 0.0;
  int steps = 0, fin = 0;
  while (!fin) {
    double dt_min = 0.0;
    int err = 0;
    if (fxg_calc_timestep(ctx, CFL, DFL, &dt_min, &err, NULL)) return fail_fxg("time step size");
    const double dt = update_dt(dt_min, t, tend, tend, &fin);
    if (fxg_timestep_lserkw2(ctx, t, dt, 5, RKA, RKb, RKc)) return fail_fxg("time step");
    t += dt;
    ++steps;
  }
  if (fxg_get_state(ctx, U)) return fail_fxg("get state");
  if (fxg_io_write_state(argv[2], "cavity_Re100", "cavity4x4x4_mesh.h5", "GAUSS", N, K, t, U))
    return fail_io("write state");
  printf("ok t=%.17g steps=%d\n", t, steps);
  fxg_destroy(ctx);
  free(s2e); free(e2s); free(nodes); free(BC); free(btype); free(U);
  return 0;
}
