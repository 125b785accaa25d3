// mfc_consts.h — constants of one simulation as seen by the kernels (plain struct, usable from
// host and device code).
#pragma once

namespace mfc {

constexpr int TILE = 256;                      // particles per work-group of the particle kernels
constexpr double PI_FAC = 0.454728408833987;   // 10/(7 pi), kernels.cuh:28
constexpr double FLT_MAX_D = 3.4028234663852886e38;

// physical, corrector, Johnson-Cook, thermal, contact and grid constants
// (phys/corr/joco/trml/geom bundles of constants_structs.h:27-85), passed by value to every kernel
struct dev_consts {
  int N, gnx, gny, num_cell;
  double bbmin_x, bbmin_y, cell;
  double dt, mass, K, G, rho0, c0;
  double wdeltap, av_alpha, av_beta, av_eta, xsph_eps, stress_eps;
  double alpha_wp, alpha_tool, cp, tq, eta_heat, T_init;
  double jc_A, jc_B, jc_C, jc_n, jc_m, jc_Tmelt, jc_Tref, jc_eps_dot_ref;
  double contact_alpha, mu, tool_vx, tool_vy;
  int thermal_bc, fric_heat, thermal_wp, pse, pse_quirk;
  int promote;        // reference float truncations evaluated in FP64
  double secant_tol;  // absolute stop tolerance of the radial-return secant
  int jc_tanh;        // the secant's flow stress: tanh-modified JC (the yield test keeps plain JC)
  double tanh_a, tanh_b, tanh_c, tanh_d;
  int plastic;        // Johnson-Cook return on
  int stage_limit;    // most window candidates a tile stages in LDS (0: all it has room for)
  int sort_fast;      // most particles of a chunk the sort holds in registers (0: all it can)
  int sort_list;      // most list entries a chunk of the sort reads (0: all the list holds)
  // adaptive grid: the geometry fields above (bbmin, cell, gnx, gny, num_cell) are rewritten on
  // the device at the start of every step; ncell_table is the size of the binning table (the
  // counting sort runs over it), num_cell for a locked grid
  int adaptive;
  int ncell_table;
  // grid_gpu_rothlin: cell ranges rebuilt from the sorted ids; ft not carried through the sort
  int rothlin;
};

// the geometry of an adaptive grid for one step, kept on the device for the host to read
struct grid_geo {
  double bbmin_x, bbmin_y, cell;
  int gnx, gny, num_cell;
  int overflow;  // steps whose geometry did not fit the cell table
};

}  // namespace mfc
