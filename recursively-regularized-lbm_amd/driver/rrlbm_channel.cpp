// rrlbm_channel -- channel-flow driver over the C ABI, the counterpart of
// RR27V33_code_DNS_LES/main.cu:7-152.
//
// Time loop: [forcing_GP below --gp-steps | forcing] -> fused steps (or the
// drop-in kernels with --reference) with diagnostics every --ndiag steps
// (diag_flow: "Step ... umean = U_b/u_tau ..." on stdout and
// <dir>/statistics.dat, output.cu:3-66), plane profiles every --nout steps
// (<dir>/Profiles.txt, output.cu:142-225, with the tau_t column printed as a
// floating-point number), flow fields every --nflowout steps and the
// populations at the end (data/%09d.dat, data/f%09d.dat, output.cu:68-140,
// 227-258), restart from --continue N (initial.cu:184-265).
//
// Options (defaults = parameters.h / globals.cu):
//   --lx/--ly/--lz N   grid (64^3)          --steps N      time steps (1000)
//   --re-tau R         180                  --visc V       0.000622 (globals.cu:58)
//   --les              WALE (USE_LES 1)     --storage f64|f32|f16 (f32)
//   --init arrays|channel  init_arrays (initial.cu:268) or the seeded
//                      log-law + streak + noise field (channel)
//   --seed S           channel seed         --gp-steps N   forcing_GP phase (0: off;
//                      the reference uses 50000 with --reference)
//   --reference        drop-in kernels in the main.cu order instead of fused steps
//   --ndiag/--nout/--nflowout N   output periods (1000 / 1000 / 100000)
//   --dir D            output directory (data)   --continue N  restart step
//   --wall fdbc|rrbc|bb  wall treatment: RRBC (main.cu:74-75, default), FDBC in its
//                      place (main.cu:76-77, BC.cu:129-309) or plain bounce-back
//   --rr-order 3|6     RR expansion order (6: the fourth- to sixth-order terms the
//                      reference keeps commented out in collision.cu / BC.cu)
//   --bgk              the BGK operator of BGK/collision.cu (plain bounce-back walls)
//   --phase-seed S     seed of the forcing_GP phases (12345, main.cu:30)
//   --les-exact        LES in the reference order (a velocity pass, then the step)
//                      instead of the one-launch lagged-gradient step
#include <math.h>
#include <stdio.h>
#include <stdlib.h>
#include <string.h>
#include <errno.h>
#include <sys/stat.h>
#include <sys/time.h>

#include <string>
#include <vector>

#include "rrlbm.h"

static void die(rrlbm_ctx* ctx, int st, const char* what) {
  fprintf(stderr, "%s failed (status %d): %s\n", what, st, rrlbm_last_error(ctx));
  exit(1);
}
#define CALL(ctx, expr)                 \
  do {                                  \
    int st_ = (expr);                   \
    if (st_) die((ctx), st_, #expr);    \
  } while (0)

static double now() {
  struct timeval tv;
  gettimeofday(&tv, nullptr);
  return tv.tv_sec + 1e-6 * tv.tv_usec;
}

// The forcing_GP phase: the C library's rand() scaled to [0, 1], seeded once
// with srand (main.cu:3-5,30), drawn every 2000 steps (main.cu:38-41).
static double random_phase() { return (double)rand() / RAND_MAX; }

int main(int argc, char** argv) {
  rrlbm_params p;
  rrlbm_params_default(&p);
  int steps = 1000, ndiag = 1000, nout = 1000, nflowout = 100000, gp_steps = 0, cont = 0;
  double re_tau = 180.0, visc = 0.000622;
  bool reference = false, init_channel = false;
  unsigned long long seed = 2024;
  unsigned phase_seed = 12345;
  std::string dir = "data";
  for (int i = 1; i < argc; ++i) {
    std::string a = argv[i];
    auto next = [&]() -> const char* {
      if (i + 1 >= argc) {
        fprintf(stderr, "missing value for %s\n", a.c_str());
        exit(2);
      }
      return argv[++i];
    };
    if (a == "--lx") p.lx = atoi(next());
    else if (a == "--ly") p.ly = atoi(next());
    else if (a == "--lz") p.lz = atoi(next());
    else if (a == "--steps") steps = atoi(next());
    else if (a == "--re-tau") re_tau = atof(next());
    else if (a == "--visc") visc = atof(next());
    else if (a == "--les") p.use_les = 1;
    else if (a == "--storage") {
      std::string s = next();
      p.storage = s == "f64" ? RRLBM_STORE_F64 : (s == "f16" ? RRLBM_STORE_F16 : RRLBM_STORE_F32);
    } else if (a == "--init") init_channel = std::string(next()) == "channel";
    else if (a == "--seed") seed = strtoull(next(), nullptr, 10);
    else if (a == "--gp-steps") gp_steps = atoi(next());
    else if (a == "--reference") reference = true;
    else if (a == "--ndiag") ndiag = atoi(next());
    else if (a == "--nout") nout = atoi(next());
    else if (a == "--nflowout") nflowout = atoi(next());
    else if (a == "--dir") dir = next();
    else if (a == "--continue") cont = atoi(next());
    else if (a == "--wall") {
      const std::string w = next();
      p.wall_rr = w == "fdbc" ? 2 : (w == "bb" ? 0 : 1);
    } else if (a == "--rr-order") p.rr_order = atoi(next());
    else if (a == "--phase-seed") phase_seed = (unsigned)strtoul(next(), nullptr, 10);
    else if (a == "--les-exact") p.les_lagged = 0;
    else if (a == "--bgk") {
      p.collision_model = RRLBM_COLLISION_BGK;
      p.wall_rr = 0;
    } else {
      fprintf(stderr, "unknown option %s\n", a.c_str());
      return 2;
    }
  }
  // globals.cu:57-63 with LY the channel width 2h
  const double u0 = 2.0 * re_tau * visc / p.ly;
  const double ystar = visc / u0;
  p.tau = 3.0 * visc + 0.5;
  p.force[0] = 2.0 * u0 * u0 / p.ly;
  printf("Non-dimensional time, %.6e\n", p.ly / u0 / 2);
  printf("Driving force, %.6e\n", p.force[0]);

  // the output directory (statistics.dat, Profiles.txt, the field files)
  if (mkdir(dir.c_str(), 0755) != 0 && errno != EEXIST) {
    fprintf(stderr, "cannot create %s: %s\n", dir.c_str(), strerror(errno));
    return 1;
  }
  rrlbm_ctx* ctx = nullptr;
  CALL(nullptr, rrlbm_create(&p, &ctx));
  if (cont > 0) {
    CALL(ctx, rrlbm_read_restart(ctx, dir.c_str(), cont));
  } else if (init_channel) {
    CALL(ctx, rrlbm_init_channel(ctx, u0, visc, seed, 0.1, 0.01));
  } else {
    CALL(ctx, rrlbm_init_arrays(ctx, u0, ystar));
  }
  const long n = (long)p.lx * p.ly * p.lz;
  std::vector<double> prof((size_t)p.ly * 10);
  srand(phase_seed);
  double phase = 0.0;  // d_rand_phase before the first draw (globals.cu:16)
  const double t0 = now();
  double t_out = 0.0;
  bool packed = false;
  for (int istep = cont + 1; istep <= steps; ++istep) {
    // main.cu:33-46: forcing_GP (new phase every 2000 steps) below gp_steps,
    // the uniform force after
    const bool gp = istep < gp_steps;
    if (gp && istep % 2000 == 0) phase = random_phase();
    if (reference || !packed) {
      if (gp) CALL(ctx, rrlbm_forcing_gp(ctx, istep, phase));
      else CALL(ctx, rrlbm_forcing(ctx));
    }
    if (reference) {
      CALL(ctx, rrlbm_step_reference(ctx, 1));
    } else if (!packed) {
      CALL(ctx, rrlbm_fused_pack(ctx));   // this step's collision
      packed = true;
    } else if (gp) {
      CALL(ctx, rrlbm_fused_step_gp(ctx, istep, phase));   // force computed in the kernel
    } else {
      CALL(ctx, rrlbm_fused_step(ctx, 1));
    }
    const bool want_diag = istep % ndiag == 0, want_prof = istep % nout == 0;
    const bool want_flow = istep % nflowout == 0 || istep == steps, want_f = istep == steps;
    if (want_diag || want_prof || want_flow || want_f) {
      // finish the queued steps first, so that they count as computation
      CALL(ctx, rrlbm_synchronize(ctx));
      const double ts = now();
      if (packed) {
        // the fused state holds this step's post-collision populations: finish
        // the step (streaming, BB, macrovar, RRBC, macrovar_BN) into the fields
        CALL(ctx, rrlbm_fused_unpack(ctx));
        packed = false;
        // the next step restarts from the finished fields: collide them again
      }
      if (want_diag) {
        rrlbm_diag d;
        CALL(ctx, rrlbm_diag_flow(ctx, &d));
        const double um = d.u_sum[0] / n, vm = d.u_sum[1] / n, wm = d.u_sum[2] / n, rm = 1.0 + d.mass_excess / n;
        printf("Step %-7d umean = %.17e, vmean = %.17e, wmean = %.17e, rhomean = %14.6e\n", istep, um / u0, vm, wm, rm);
        fflush(stdout);
        std::string path = dir + "/statistics.dat";
        FILE* fs = fopen(path.c_str(), "a");
        if (!fs) {
          fprintf(stderr, "cannot open %s\n", path.c_str());
          return 1;
        }
        fprintf(fs, "%-7d %14.6e %14.6e %14.6e %14.6e\n", istep, um, vm, wm, rm);
        fclose(fs);
      }
      if (want_prof) {
        CALL(ctx, rrlbm_plane_profiles(ctx, prof.data(), prof.size()));
        std::string path = dir + "/Profiles.txt";
        FILE* fp = fopen(path.c_str(), "a");
        if (!fp) {
          fprintf(stderr, "cannot open %s\n", path.c_str());
          return 1;
        }
        fprintf(fp, "%d\n", istep);
        for (int iy = 0; iy < p.ly; ++iy) {
          const double* o = &prof[(size_t)iy * 10];
          fprintf(fp, "%.1f, %14.6e %14.6e %14.6e %14.6e %14.6e %14.6e %14.6e %14.6e %14.6e %14.6e %14.6e\n",
                  iy + 0.5, (iy + 0.5) / ystar, o[0], o[1], o[2], o[3], o[4], o[5], o[6], o[7], o[8], o[9]);
        }
        fclose(fp);
      }
      if (want_flow) CALL(ctx, rrlbm_write_flow(ctx, dir.c_str(), istep));
      if (want_f) CALL(ctx, rrlbm_write_f(ctx, dir.c_str(), istep));
      t_out += now() - ts;
    }
  }
  CALL(ctx, rrlbm_synchronize(ctx));
  const double total = now() - t0;
  printf("Simulation completed.\n");
  printf("Total iterations: %d\n", steps);
  printf("Output time: %.2f seconds\n", t_out);
  printf("Total computation time: %.2f seconds (%.1f MLUPS)\n", total,
         (double)n * (steps - cont) / (total - t_out) / 1e6);
  rrlbm_destroy(ctx);
  return 0;
}
