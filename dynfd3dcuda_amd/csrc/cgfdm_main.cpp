// cgfdm: command-line driver with the reference's argv, exit codes and
// progress output (srcd/main.cu:123-580, srcd/set_params.cu:108-132):
//   cgfdm [params.json]
// exit -1 (255) for params file / JSON errors, 2 for netCDF errors, 110 for
// the CFL check; prints the parameter block (set_params.cu:299-388, after
// reading params.json as main.cu:156 does) and "> %6d, total %d, %8.4lf s"
// every 10 steps.  A run whose fault record turns non-finite stops with
// exit 1 after writing that record (the reference writes on silently).
#include "../../include/cgfdm.h"
#include <chrono>
#include <cstdio>
#include <cstring>

int main(int argc, char** argv) {
  const char* path = argc > 1 ? argv[1] : "params.json";
  std::printf("PARAMS_FILE = %s\n", path);
  cgfdm_ctx* ctx = nullptr;
  int rc = cgfdm_create(path, &ctx);
  if (rc) {
    std::printf("%s\n", cgfdm_last_error());
    return -1;
  }
  cgfdm_print_params(ctx);
  rc = cgfdm_setup(ctx);
  if (rc) {
    std::printf("%s\n", cgfdm_last_error());
    cgfdm_destroy(ctx);
    if (rc == CGFDM_ECFL) return 110;
    if (rc == CGFDM_ENETCDF) return 2;
    if (rc == CGFDM_EFILE || rc == CGFDM_EJSON) return -1;
    return 1;
  }
  int dims[8];
  cgfdm_get_dims(ctx, dims);
  int nt = dims[7];
  rc = cgfdm_open_outputs(ctx);
  if (rc) {
    std::printf("%s\n", cgfdm_last_error());
    cgfdm_destroy(ctx);
    return 2;
  }
  double total = 0;
  for (int it = 0; it < nt; it++) {
    auto t0 = std::chrono::steady_clock::now();
    if ((rc = cgfdm_write_outputs(ctx, it)) || (rc = cgfdm_step(ctx, it)) || (rc = cgfdm_synchronize(ctx))) {
      std::printf("%s\n", cgfdm_last_error());
      cgfdm_destroy(ctx);
      return rc == CGFDM_ENETCDF ? 2 : 1;
    }
    double dt = std::chrono::duration<double>(std::chrono::steady_clock::now() - t0).count();
    total += dt;
    if (it % 10 == 0) {
      std::printf("> %6d, total %d, %8.4lf s\n", it, nt, dt);
      std::fflush(stdout);
    }
  }
  cgfdm_close_outputs(ctx);
  std::printf("----------------------------------\n");
  std::printf("> Total %6d: wall %e s\n", nt, total);
  cgfdm_destroy(ctx);
  return 0;
}
