// C++ host loop driving the GPU engine through its C ABI.  It mirrors the
// reference driver's time loop (src/meshImportTest.cpp:194-313) without MPI:
// one process owns the whole mesh on one GPU.  Per step it calls the engine's
// stages in the reference order (emit, move, regroup, collide, sample), writes
// the cell fields at the reference's output cadence (statistic_Result,
// :377-406) and prints a stage-time line every NDSIPLAY = 50 steps (:26-74).
//
// usage: dsmc_run [apollo|box] [steps] [ppc] [--mixed] [--surface] [--out DIR]
#include <chrono>
#include <cstdio>
#include <cstdlib>
#include <cstring>
#include <string>
#include <vector>

#include "dsmc_engine.h"

static void die(int rc, const char* what) {
  if (rc != DSMC_OK) {
    std::fprintf(stderr, "%s failed (%d): %s\n", what, rc, dsmc_last_error());
    std::exit(1);
  }
}

static bool output_step(int istep) {  // src/meshImportTest.cpp:383-386
  if (istep <= 2000 && istep % 100 == 0) return true;
  if (istep > 2000 && istep <= 10000 && istep % 1000 == 0) return true;
  if (istep > 10000 && istep <= 50000 && istep % 2000 == 0) return true;
  return istep % 50000 == 0;
}

int main(int argc, char** argv) {
  std::string cas = argc > 1 ? argv[1] : "apollo";
  const int nsteps = argc > 2 ? std::atoi(argv[2]) : 500;
  const double ppc = argc > 3 ? std::atof(argv[3]) : 20.0;
  bool mixed = false, surface = false;
  std::string outdir;
  for (int i = 4; i < argc; ++i) {
    if (!std::strcmp(argv[i], "--mixed")) mixed = true;
    else if (!std::strcmp(argv[i], "--surface")) surface = true;
    else if (!std::strcmp(argv[i], "--out") && i + 1 < argc) outdir = argv[++i];
  }
  dsmc_mesh* mesh = nullptr;
  dsmc_params p;
  if (cas == "box") {
    die(dsmc_mesh_box_tets(64, DSMC_TAG_WALL, &mesh), "box mesh");
    die(dsmc_params_box(&p, 0.05), "params");
  } else {
    die(dsmc_mesh_capsule(25, 62, 50, 3.912, 0, 0.0, 0, &mesh), "capsule mesh");
    die(dsmc_params_apollo(&p, 1), "params");
    p.NSS = nsteps / 2;
  }
  p.Npinitial = (int)ppc;
  p.precision = mixed ? DSMC_PREC_MIXED : DSMC_PREC_FP64;
  p.surface_stats = surface ? 1 : 0;
  dsmc_mesh_view view;
  die(dsmc_mesh_get_view(mesh, &view), "view");
  die(dsmc_params_derive(&view, &p), "derive");
  int64_t n = 0;
  dsmc_particles_soa soa{};
  die(dsmc_fill_freestream(&view, &p, ppc, cas == "box" ? 1.0 : 0.0, 20261203ull, &soa, &n), "fill count");
  std::vector<double> x(n), y(n), z(n), vx(n), vy(n), vz(n), ir(n);
  std::vector<int32_t> cell(n);
  std::vector<uint32_t> id(n);
  soa = {n, x.data(), y.data(), z.data(), vx.data(), vy.data(), vz.data(), ir.data(), cell.data(), id.data()};
  die(dsmc_fill_freestream(&view, &p, ppc, cas == "box" ? 1.0 : 0.0, 20261203ull, &soa, &n), "fill");
  dsmc_engine* eng = nullptr;
  die(dsmc_engine_create(&view, &p, 20261203ull, 0, &eng), "engine");
  die(dsmc_particles_upload(eng, &soa), "upload");
  die(dsmc_statistic_pre(eng), "statistic_pre");
  std::printf("case=%s cells=%d particles=%lld dtime=%.6g Neff=%.6g\n", cas.c_str(), view.ncell, (long long)n,
              p.dtime, p.Neff);
  std::vector<double> macro((size_t)view.ncell * DSMC_MACRO_WIDTH);
  dsmc_profile_enable(eng, 1);
  auto t0 = std::chrono::steady_clock::now();
  for (int istep = 1; istep <= nsteps; ++istep) {
    die(dsmc_step(eng, istep, DSMC_STAGE_ALL), "step");
    if (output_step(istep)) {
      die(dsmc_macro_get(eng, DSMC_MACRO_OUTPUT, macro.data()), "macro");
      if (!outdir.empty()) {
        // out2dat's file (src/ProcessDSMC.cpp:3421-3424): Fluent cell fields
        char fn[256];
        std::snprintf(fn, sizeof(fn), "%s/DSMCKn%.1f_iter%d_CELL%d.dat", outdir.c_str(), p.Kn, istep, view.ncell);
        die(dsmc_write_fluent_cells(fn, &p, view.ncell, macro.data(), 2, 1), "cell fields file");
        if (surface && istep > p.NSS) {
          // outBoundaryStressHeat's file (src/ProcessDSMC.cpp:3618-3621)
          std::vector<double> wall((size_t)view.nface * DSMC_SURF_WIDTH);
          double cd = 0.0, qpeak = 0.0;
          die(dsmc_surface_fields(eng, istep, p.L0, wall.data(), &cd, &qpeak), "surface fields");
          std::snprintf(fn, sizeof(fn), "%s/BoundaryStressHeatFluentKn%.1f_iter%d.dat", outdir.c_str(), p.Kn, istep);
          die(dsmc_write_fluent_walls(fn, &view, wall.data(), 7, 1), "wall fields file");
          std::printf("step %d: C_D = %.6g, peak q = %.6g\n", istep, cd, qpeak);
        }
      }
    }
    if (istep % 50 == 0) {
      dsmc_kernel_times kt;
      dsmc_profile_get(eng, &kt);
      dsmc_stats st;
      dsmc_stats_get(eng, &st);
      std::printf("Finished %d of %d steps. Emit: %.4f  Move: %.4f  Sort: %.4f  Collide+Sample: %.4f ms  N=%lld\n",
                  istep, nsteps, kt.ms[0], kt.ms[1], kt.ms[2] + kt.ms[3] + kt.ms[4] + kt.ms[5], kt.ms[6],
                  (long long)st.particles);
      dsmc_profile_enable(eng, 1);
    }
  }
  const double sec = std::chrono::duration<double>(std::chrono::steady_clock::now() - t0).count();
  dsmc_stats st;
  dsmc_stats_get(eng, &st);
  std::printf("Total running time = %f s, particle-steps/s = %.4g\n", sec, (double)st.particle_steps / sec);
  dsmc_engine_destroy(eng);
  dsmc_mesh_destroy(mesh);
  return 0;
}
