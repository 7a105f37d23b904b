// Test helper: exercises the solver's netCDF classic reader/writer
// (dynfd3dcuda_amd/csrc/netcdf_classic.cpp) from the command line.
//   nc_tool write <path>        writes the fixed test file (see test_netcdf.py)
//   nc_tool read <path> <var>   prints the variable as doubles, one per line
#include "netcdf_classic.h"
#include <cstdio>
#include <cstring>
#include <string>
#include <vector>

using cgfdm::NcFile;

int main(int argc, char** argv) {
  if (argc < 3) return 2;
  std::string err;
  if (!std::strcmp(argv[1], "write")) {
    NcFile f;
    if (!f.create(argv[2], err)) { std::fprintf(stderr, "%s\n", err.c_str()); return 1; }
    int dt = f.def_dim("nt", 0), dz = f.def_dim("nz", 3), dy = f.def_dim("ny", 4);
    int vx = f.def_var("x", cgfdm::NC_FLOAT_, {dz, dy});
    int vu = f.def_var("united", cgfdm::NC_INT_, {dz, dy});
    int vv = f.def_var("Vs1", cgfdm::NC_FLOAT_, {dt, dz, dy});
    int vd = f.def_var("t", cgfdm::NC_DOUBLE_, {dt});
    if (!f.end_def(err)) { std::fprintf(stderr, "%s\n", err.c_str()); return 1; }
    std::vector<float> x(12);
    std::vector<int> u(12);
    for (int q = 0; q < 12; q++) { x[q] = 0.5f * q - 1.0f; u[q] = q % 3; }
    f.put_var(vx, x.data());
    f.put_var(vu, u.data());
    for (int r = 0; r < 3; r++) {
      std::vector<float> v(12);
      for (int q = 0; q < 12; q++) v[q] = 100.0f * r + q;
      double t = 0.25 * r;
      f.put_record(vv, r, v.data());
      f.put_record(vd, r, &t);
    }
    f.close();
    return 0;
  }
  if (!std::strcmp(argv[1], "read") && argc == 4) {
    NcFile f;
    if (!f.open_read(argv[2], err)) { std::fprintf(stderr, "%s\n", err.c_str()); return 1; }
    std::vector<double> v;
    if (!f.read_double(argv[3], v, err)) { std::fprintf(stderr, "%s\n", err.c_str()); return 1; }
    for (double d : v) std::printf("%.17g\n", d);
    return 0;
  }
  return 2;
}
