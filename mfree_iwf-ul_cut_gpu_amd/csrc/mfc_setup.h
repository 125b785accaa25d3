// mfc_setup.h — host-side scenario generation of the orthogonal cut (setup_ref_cut).
#pragma once

#include <string>
#include <vector>

#include "../../include/mfc.h"
#include "mfc_consts.h"

namespace mfc {

// constants of one simulation (phys/corr/joco/trml/geom bundles, constants_structs.h:27-85)
struct sim_constants {
  int N, n_tool;
  // grid (grid.cu:139-158)
  int gnx, gny, num_cell;
  double bbmin_x, bbmin_y, cell;
  // time step (benchmarks.cpp:530-531)
  double dt, t_final;
  // material
  double mass, rho0, K, G, c0;  // c0 = sqrtf(K/rho0) as float (actions_gpu.cu:47)
  // correctors
  double wdeltap, av_alpha, av_beta, av_eta, xsph_eps, stress_eps;
  // thermal
  double alpha_wp, alpha_tool, cp, tq, eta_heat, T_init;
  // Johnson-Cook
  double jc_A, jc_B, jc_C, jc_n, jc_m, jc_Tmelt, jc_Tref, jc_eps_dot_ref;
  // contact
  double contact_alpha, mu, tool_vx, tool_vy;
  // switches derived from the constants (actions_gpu.cu:448-468, interactions_gpu.cu:443)
  int thermal_bc, fric_heat, thermal_wp, pse, pse_quirk;
  int promote;        // float truncations of the reference evaluated in FP64
  double secant_tol;  // radial-return secant tolerance
  int jc_tanh;        // the secant solves with the tanh-modified flow stress
  double tanh_a, tanh_b, tanh_c, tanh_d;
  int plastic;        // Johnson-Cook return on (the reference's m_plastic, actions_gpu.cu:25)
  int adaptive;       // the grid follows the particles every step (grid.cu:97-136)
  int ncell_table;    // cells of the binning table: num_cell (locked) or max_cell (adaptive)
  int rothlin;        // grid_gpu_rothlin semantics (MFC_GRID_ROTHLIN)
};

struct scenario {
  sim_constants k;
  std::vector<double> px, py, rho, h, T;
  std::vector<double> vx, vy;  // initial velocities; empty: at rest
  std::vector<int> flags;
  mfc_tool_state tool;
};

// the geometry of an adaptive grid over positions and smoothing lengths (all N particles,
// blanked ones included): box of the positions +- 1e-6, cell size max h (kernel width 2),
// nx = ceil(lx / cell), ny = ceil(ly / cell) (grid_base::update_geometry, grid.cu:97-136)
void adaptive_geometry(const double *px, const double *py, const double *h, int N, double &bbmin_x, double &bbmin_y,
                       double &cell, int64_t &gnx, int64_t &gny);

// fills a scenario from a configuration; returns MFC_OK or an MFC_E_* code with err set
int build_scenario(const mfc_config &cfg, scenario &out, std::string &err);

// the constants as passed to the kernels
void make_dev_consts(const sim_constants &c, dev_consts &k);

// one incremental tool move by dt (tool.cpp:211-221, geometric_primitives.cpp:70-85)
void tool_advance(mfc_tool_state &t, double dt);

// the literal constants of setup_ref_cut (benchmarks.cpp:255-332,384-421)
void config_ref_cut(mfc_config &c, int ny, double rake, double clear, double chamfer, double speed, double feed);
void config_rings(mfc_config &c, int nbox);
void config_impact(mfc_config &c, int nbox);

}  // namespace mfc
