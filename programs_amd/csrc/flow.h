// Device state of a SIMPLE flow on an fvk context.
#pragma once
#include <vector>

#include "ctx.h"

namespace fvk {

// Scalars of the outer iteration kept in HBM (module variables of the
// reference: cfd_data.f90:15-16, ieq%res_n, the saved pp_res_n of
// cfd_setup_pp).
struct FlowSc {
  double res_n_u, pp_res_n, mx_res_n, c_in, c_out, c_rat;
  double phi_mx_cell, phi_mx_all;
  double h_reduc, h_redpp, h_redpp_sum, h_red_p, h_red_p_sum;
  double res0_n_u[3], res0_n_pp, res0_n_p;
  int it_u[3], it_pp, it_p, diverged;
  int hist_k;     // next slot of the device history (graph replays)
  int fallbacks;  // pressure solves that fell back to plain PCG
};

struct Flow {
  Ctx *c = nullptr;
  fvk_flow_cfg cfg{};
  int nvc = 0;
  // fields (device rows, component-major)
  double *u = nullptr, *ub = nullptr, *p = nullptr, *dpdx = nullptr;
  double *spd = nullptr, *frc_dns = nullptr, *cs = nullptr, *frc_vol = nullptr;
  double *ddx = nullptr, *g0 = nullptr, *rhs = nullptr, *rhs_p = nullptr;
  double *cor = nullptr, *cfl = nullptr, *cpp = nullptr, *fc1 = nullptr, *cinout = nullptr;
  double *pl_r = nullptr, *pc = nullptr, *ddxp = nullptr, *red_scratch = nullptr;
  FlowSc *sc = nullptr;
  // inlet (class 10) and outlet (class 20) boundary faces, ascending
  int *io_faces = nullptr;
  int n_in = 0, n_out = 0;
  // cells with outlet faces: per cell up to 4 outlet faces (ascending) and
  // the velocity after each of them (outlet back-flow projection,
  // cfd_setup.f90:349-356, modifies the cell value in face order)
  int n_oc = 0;
  int *oc_cell = nullptr, *oc_face = nullptr, *cell_oc = nullptr;
  double *oc_u = nullptr;
  int row_of_cell1 = 0;
  // launches are being captured into a HIP graph: no host round trips
  bool capture = false;
  // learned pressure preconditioner (p_solver 2) and optional training on
  // every pressure-correction matrix the flow assembles
  fvk_model *model = nullptr;
  int train_steps = 0;
  double train_lr = 1e-3;
  // graphs of one outer iteration (segments between host-driven solves)
  std::vector<hipGraphExec_t> graphs;
  FlowSc *hist = nullptr;
  int hist_cap = 0;
};

int mom_setup(Flow &F, int it, bool &skipped);
int pp_setup(Flow &F, int it);
int outer_begin(Flow &F);
// n outer iterations from it0: it0 eagerly, the rest as replays of HIP
// graphs of one outer iteration; history into h (n entries, may be null)
int flow_run_graph(Flow &F, int it0, int n, fvk_outer_hist *h, int &replayed);

}  // namespace fvk
