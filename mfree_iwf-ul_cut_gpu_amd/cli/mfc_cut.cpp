// mfc_cut — command-line driver of the orthogonal cut, the C++ host side over the C ABI.
//
// Same options and force log as the reference's main() (mfree_gpu.cpp:146-323):
//   -r rake angle [deg]   -c clearance angle [deg]   -f edge radius (fillet) [cm]
//   -v cutting speed [cm/us]   -n rows of the reference lattice   -q feed [cm]
// plus -s steps (default: the reference's t_final/dt), -p precision (0 FP64, 1 FP32C) and
// -o output directory (default "results"), -b rings|impact for the reference's verification
// setups instead of the cut (main()'s commented lines, mfree_gpu.cpp:209-210; -n is then nbox,
// default 80 / 20; no tool: no tool VTK files). Every last_step/135 steps it appends the tool force
// "Fx Fy" (%.9g) to <out>/force_<fillet>_<feed> and the Usui wear "min max mean accumulated"
// (%e) to <out>/wear_<speed>_<feed>, writes <out>/vtk_out_NNNNNN.vtk and vtk_tool_NNNNNN.vtk
// (-x: no VTK files), as mfree_gpu.cpp:238-304 does, and prints the progress line.
#include <sys/stat.h>
#include <unistd.h>

#include <chrono>
#include <cstdio>
#include <cstdlib>
#include <string>

#include "mfc.h"

int main(int argc, char **argv) {
  double rake = 10., clear = 8., fillet = 0.002, speed = 83.333328e-5 / 500. * 70., feed = 0.01;
  int ny = -1, precision = MFC_PREC_FP64;  // ny: 30 for the cut, nbox 80 / 20 for -b
  bool vtk = true;
  long steps = -1;
  std::string out = "results", bench;
  int c;
  while ((c = getopt(argc, argv, "r:c:f:v:n:q:s:p:o:b:xh")) != -1) {
    switch (c) {
      case 'r': rake = std::atof(optarg); break;
      case 'c': clear = std::atof(optarg); break;
      case 'f': fillet = std::atof(optarg); break;
      case 'v': speed = std::atof(optarg); break;
      case 'n': ny = std::atoi(optarg); break;
      case 'q': feed = std::atof(optarg); break;
      case 's': steps = std::atol(optarg); break;
      case 'p': precision = std::atoi(optarg); break;
      case 'o': out = optarg; break;
      case 'x': vtk = false; break;
      case 'b': bench = optarg; break;
      default:
        std::fprintf(stderr, "usage: %s [-r rake] [-c clearance] [-f fillet] [-v speed] [-n ny] [-q feed] "
                             "[-s steps] [-p 0|1] [-o dir] [-x] [-b rings|impact]\n", argv[0]);
        return c == 'h' ? 0 : 2;
    }
  }
  std::printf("arguments received %.9g %.9g %.9g %.9g %.9g\n", rake, clear, fillet, speed, feed);

  mfc_config cfg;
  if (bench == "rings") {
    mfc_config_rings(&cfg, ny > 0 ? ny : 80);
  } else if (bench == "impact") {
    mfc_config_impact(&cfg, ny > 0 ? ny : 20);
  } else if (bench.empty()) {
    mfc_config_ref_cut(&cfg, ny > 0 ? ny : 30, rake, clear, fillet, speed, feed);
  } else {
    std::fprintf(stderr, "unknown setup %s (rings or impact)\n", bench.c_str());
    return 2;
  }
  cfg.precision = precision;
  mfc_sim *sim = nullptr;
  if (mfc_create(&cfg, &sim) != MFC_OK) {
    std::fprintf(stderr, "mfc_create: %s\n", mfc_last_error());
    return 1;
  }
  mfc_info info;
  mfc_get_info(sim, &info);
  mfc_tool_state tool_state;
  const bool has_tool = mfc_get_tool(sim, &tool_state) == MFC_OK && tool_state.num_seg > 0;
  const long last_step = steps >= 0 ? steps : (long)(info.t_final / info.dt);
  long freq = (long)(info.t_final / info.dt) / 135;
  if (freq < 1) freq = 1;
  std::printf("%d particles (%d tool), %d cells, dt %.9g us, %ld steps\n", info.n_particles, info.n_tool,
              info.num_cell, info.dt, last_step);

  mkdir(out.c_str(), 0777);
  char name[512];
  std::snprintf(name, sizeof(name), "%s/force_%f_%f", out.c_str(), fillet, feed);
  FILE *fp = std::fopen(name, "w+");
  if (!fp) {
    std::fprintf(stderr, "cannot open %s\n", name);
    mfc_destroy(sim);
    return 1;
  }
  // the wear log with the reference's Usui constants and tool velocity (benchmarks.cpp:384,508-512)
  std::snprintf(name, sizeof(name), "%s/wear_%f_%f", out.c_str(), speed, feed);
  FILE *fw = std::fopen(name, "w+");
  if (!fw) {
    std::fprintf(stderr, "cannot open %s\n", name);
    std::fclose(fp);
    mfc_destroy(sim);
    return 1;
  }
  const double usui_K = 1. / 100 * 7.8, usui_alpha = 2500.;
  const double vel_tool = speed == 0. ? 83.333328 * 1e-5 : speed;
  const double wear_tool_vel[2] = {0., vel_tool};
  const double time_btwn_dumps = freq * info.dt;
  const auto t0 = std::chrono::steady_clock::now();
  int rc = MFC_OK;
  unsigned print_iter = 0;
  for (long step = 0; step < last_step + 1 && rc == MFC_OK;) {
    if (step % freq == 0) {
      double F[2];
      if ((rc = mfc_tool_force(sim, F)) != MFC_OK) break;
      std::fprintf(fp, "%.9g %.9g\n", F[0], F[1]);
      std::fflush(fp);
      double W[4];
      if ((rc = mfc_tool_wear(sim, usui_K, usui_alpha, wear_tool_vel, time_btwn_dumps, W)) != MFC_OK) break;
      std::fprintf(fw, "%e %e %e %e\n", W[0], W[1], W[2], W[3]);
      std::fflush(fw);
      if (vtk) {
        char vname[600];
        std::snprintf(vname, sizeof(vname), "%s/vtk_out_%06u.vtk", out.c_str(), print_iter);
        if ((rc = mfc_write_vtk(sim, vname)) != MFC_OK) break;
        std::snprintf(vname, sizeof(vname), "%s/vtk_tool_%06u.vtk", out.c_str(), print_iter);
        if (has_tool && (rc = mfc_write_tool_vtk(sim, vname)) != MFC_OK) break;
      }
      print_iter++;
      const double s = std::chrono::duration<double>(std::chrono::steady_clock::now() - t0).count();
      std::printf("%06ld of %06ld: %04.1f percent done, %f s\n", step, last_step, 100.0 * step / last_step, s);
    }
    // run up to the next report (hipGraph replay inside mfc_step)
    long next = (step / freq + 1) * freq;
    if (next > last_step + 1) next = last_step + 1;
    rc = mfc_step(sim, (int)(next - step));
    step = next;
  }
  if (rc == MFC_OK) rc = mfc_sync(sim);
  std::fclose(fp);
  std::fclose(fw);
  const double s = std::chrono::duration<double>(std::chrono::steady_clock::now() - t0).count();
  if (rc != MFC_OK) std::fprintf(stderr, "step failed: %s\n", mfc_last_error());
  else std::printf("gpu time: (seconds) %f\n", s);
  mfc_destroy(sim);
  return rc == MFC_OK ? 0 : 1;
}
