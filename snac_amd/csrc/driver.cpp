// The SNaC main program (main.f90) on the library: run it in a directory
// holding dns.nml and blocks.nml, as the reference's binary is run; it
// writes the same data/ files (grid and geometry files, visualisation
// fields and their log, checkpoints, time.out).  The argument is the FFT
// axis (0..3), the reference's compile-time FFT_AXIS; the device ordinal
// comes second (default 0); a third argument "impdiff" selects implicit
// diffusion (the reference's IMPDIFF=1 build).
//
// Order of main.f90: set-up and sanity checks (219-524, in the context's
// creation), grid files (412-437), initial condition or checkpoint
// (526-561), ghost fills (562-639), the initial out3d (654-656), the first
// dt (660-662), then the time loop (893-1136): one RK3 step, the stopping
// criteria (1037-1047), the stability and divergence checks every icheck
// steps (1048-1069), time.out every iout0d, out3d every iout3d and the
// checkpoint every isave steps or at the end (1070-1134).  The 0D/1D/2D
// diagnostics beyond time.out and the wall-force integrals are not written.
#include <algorithm>
#include <chrono>
#include <cmath>
#include <cstdio>
#include <cstdlib>
#include <filesystem>
#include <limits>
#include <string>
#include <vector>

#include "../../include/snac_amd.h"
#include "io.hpp"

namespace {

snac_ctx *g_ctx = nullptr;

void check(int rc, const char *what) {
  if (rc == 0) return;
  std::fprintf(stderr, "ERROR: %s: %s\n", what, snac_last_error());
  if (g_ctx) snac_ctx_destroy(g_ctx);
  std::exit(1);
}

std::string i7(int n) {
  char b[16];
  std::snprintf(b, sizeof b, "%07d", n);
  return b;
}

// out3d.h90: u, v, w, p and every scalar of every block
void out3d(snac_ctx *c, const std::string &dir, int nblocks, int nscal, double time, int istep) {
  const std::string fld = i7(istep);
  const int one[3] = {1, 1, 1};
  for (int b = 0; b < nblocks; b++) {
    int n[3], lo[3], nb[6], per[3];
    check(snac_ctx_block_info(c, b, n, lo, nb, per), "block_info");
    const std::string cb = snac::block_num(b), log = "log_visu_3d_b_" + cb + ".out";
    const char *pre[4] = {"vex", "vey", "vez", "pre"};
    const char *var[4] = {"Velocity_X", "Velocity_Y", "Velocity_Z", "Pressure_P"};
    for (int f = 0; f < 4; f++)
      check(snac_write_visu_3d(c, dir.c_str(), (std::string(pre[f]) + "_fld_" + fld + "_b_" + cb + ".bin").c_str(),
                               log.c_str(), var[f], b, f, one, n, one, time, istep),
            "write_visu_3d");
    for (int is = 0; is < nscal; is++) {
      const std::string sn = snac::block_num(is);
      check(snac_write_visu_3d(c, dir.c_str(), ("sca_" + sn + "_fld_" + fld + "_b_" + cb + ".bin").c_str(),
                               log.c_str(), ("Scalar_S" + sn).c_str(), b, SNAC_SCAL0 + 2 * is, one, n, one, time,
                               istep),
            "write_visu_3d");
    }
  }
}

// out0d: appends the values with '(nE15.7)'
void out0d(const std::string &path, const std::vector<double> &v) {
  std::FILE *f = std::fopen(path.c_str(), "a");
  if (!f) return;
  for (double x : v) std::fputs(snac::fortran_e(x, 15, 7).c_str(), f);
  std::fputc('\n', f);
  std::fclose(f);
}

}  // namespace

int main(int argc, char **argv) {
  const int fft_axis = argc > 1 ? std::atoi(argv[1]) : 0;
  const int device = argc > 2 ? std::atoi(argv[2]) : 0;
  // the reference's build options IMPDIFF=1 and BOUSSINESQ_BUOYANCY=1 as words after the device
  bool impdiff = false, buoyancy = false;
  for (int q = 3; q < argc; q++) {
    const std::string o = argv[q];
    if (o == "impdiff") impdiff = true;
    else if (o == "buoyancy") buoyancy = true;
    else {
      std::fprintf(stderr, "unknown option %s (impdiff, buoyancy)\n", argv[q]);
      return 2;
    }
  }
  const std::string datadir = "data/";
  std::filesystem::create_directories(datadir);
  check(snac_ctx_create_from_nml(&g_ctx, "dns.nml", "blocks.nml", fft_axis, device), "set-up");
  snac_ctx *c = g_ctx;
  if (impdiff) check(snac_set_impdiff(c, 1), "implicit diffusion");
  if (buoyancy) check(snac_set_buoyancy(c, 1), "Boussinesq buoyancy");
  double dp[11];
  int ip[8], rp[8];
  check(snac_ctx_get_params(c, dp, ip), "params");
  check(snac_ctx_get_run_params(c, rp), "params");
  const double cfl = dp[1], dtmax = dp[2], dt_f = dp[3], time_max = dp[8], tw_max = dp[9];
  const int nstep = ip[0], icheck = ip[1], nscal = ip[2];
  const bool stop_type[3] = {ip[5] != 0, ip[6] != 0, ip[7] != 0};
  const int iout0d = rp[0], iout3d = rp[3], isave = rp[4], nsaves_max = rp[5];
  const bool is_overwrite_save = rp[6] != 0, restart = rp[7] != 0;
  const int nblocks = snac_ctx_num_blocks(c);
  const double small = std::numeric_limits<double>::epsilon() * std::pow(10.0, 15 / 2.0);

  check(snac_save_grid(c, datadir.c_str()), "save_grid");
  double time = 0.0;
  int istep = 0;
  if (!restart) {
    check(snac_initflow(c), "initflow");
    std::printf(" *** Initial condition succesfully set ***\n");
  } else {
    check(snac_restart_io(c, 'r', datadir.c_str(), &time, &istep), "checkpoint");
    std::printf(" *** Checkpoint loaded at time = %s time step = %d. ***\n", snac::fortran_e(time, 15, 7).c_str(), istep);
  }
  check(snac_init_bounds(c), "init_bounds");
  out3d(c, datadir, nblocks, nscal, time, istep);
  double dt_cfl = 0.0;
  check(snac_chkdt(c, &dt_cfl), "chkdt");
  double dt = dt_f > 0.0 ? dt_f : std::min(cfl * dt_cfl, dtmax);
  std::printf(" dt_cfl = %.16E dt = %.16E\n", dt_cfl, dt);
  std::printf(" *** Calculation loop starts now ***\n");
  const auto twi = std::chrono::steady_clock::now();
  bool is_done = false, kill = false;
  int savecounter = 0;
  while (!is_done) {
    istep = istep + 1;
    time = time + dt;
    std::printf(" Timestep # %d Time = %.16E\n", istep, time);
    check(snac_step(c, dt), "step");
    if (stop_type[0] && istep >= nstep) is_done = true;
    if (stop_type[1] && time >= time_max) is_done = true;
    if (stop_type[2]) {
      const double tw = std::chrono::duration<double>(std::chrono::steady_clock::now() - twi).count() / 3600.0;
      if (tw >= tw_max) is_done = true;
    }
    if (icheck > 0 && istep % icheck == 0) {
      std::printf(" Checking stability and divergence...\n");
      check(snac_chkdt(c, &dt_cfl), "chkdt");
      dt = dt_f > 0.0 ? dt_f : std::min(cfl * dt_cfl, dtmax);
      std::printf(" dt_cfl = %.16E dt = %.16E\n", dt_cfl, dt);
      if (dt_cfl < small) {
        std::fprintf(stderr, " ERROR: timestep is too small.\n Aborting...\n");
        is_done = true;
        kill = true;
      }
      double divtot = 0.0, divmax = 0.0;
      check(snac_chkdiv(c, &divtot, &divmax), "chkdiv");
      std::printf(" Total divergence = %.16E | Maximum divergence = %.16E\n", divtot, divmax);
      if (divtot != divtot) {
        std::fprintf(stderr, " ERROR: maximum divergence is too large.\n Aborting...\n");
        is_done = true;
        kill = true;
      }
    }
    if (iout0d > 0 && istep % iout0d == 0) out0d(datadir + "time.out", {1.0 * istep, dt, time});
    if (iout3d > 0 && istep % iout3d == 0) out3d(c, datadir, nblocks, nscal, time, istep);
    if (isave > 0 && (istep % isave == 0 || (is_done && !kill))) {
      for (int b = 0; b < nblocks; b++) {
        const std::string cb = snac::block_num(b);
        std::string base = is_overwrite_save ? "fld_b_" + cb : "fld_" + i7(istep) + "_b_" + cb;
        if (!is_overwrite_save && nsaves_max > 0) {
          if (b == 0) {
            savecounter = savecounter + 1;
            const int ichkptnum = savecounter % nsaves_max + 1;
            out0d(datadir + "log_saves.out", {1.0 * istep, time, 1.0 * ichkptnum});
          }
          char num[16];
          std::snprintf(num, sizeof num, "%04d", savecounter % nsaves_max + 1);
          base = std::string("fld_") + num + "_b_" + cb;
        }
        std::vector<std::string> names = {"_u", "_v", "_w", "_p"};
        for (int is = 0; is < nscal; is++) names.push_back("_s_" + snac::block_num(is));
        for (size_t f = 0; f < names.size(); f++) {
          const int fid = f < 4 ? (int)f : SNAC_SCAL0 + 2 * ((int)f - 4);
          int is_ = istep;
          double t_ = time;
          check(snac_load_one(c, 'w', (datadir + base + names[f] + ".bin").c_str(), b, fid, &t_, &is_), "checkpoint");
          if (!is_overwrite_save) {  // fld_b_NNN_* -> the last checkpoint (symbolic links)
            const std::filesystem::path link = datadir + "fld_b_" + cb + names[f] + ".bin";
            std::error_code ec;
            std::filesystem::remove(link, ec);
            std::filesystem::create_symlink(base + names[f] + ".bin", link, ec);
          }
        }
      }
      std::printf(" *** Checkpoint saved at time = %s time step = %d. ***\n", snac::fortran_e(time, 15, 7).c_str(),
                  istep);
    }
  }
  if (!kill) std::printf(" *** Fim ***\n");
  snac_ctx_destroy(c);
  return 0;
}
