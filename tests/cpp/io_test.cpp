// Output and restart round trips through the reference's interface: linear
// fields set from host callbacks (so every interpolation the writers make is
// exact), written with writeVTK / writeDat / writeVTKFullMesh, and dumped and
// reloaded in the text and binary formats.
// Usage: io_test <output directory> <N> <periodic z: 0|1>
// Prints one line per check; exit status 0 when every round trip holds.
#include <cmath>
#include <cstdio>
#include <cstdlib>
#include <iostream>
#include <stdexcept>
#include <string>

#include "PressureTensor.h"
#include "VTKDatExport.h"
#include "VelocityTensor.h"

double Reynolds = 1.0;

namespace {

Real fu(Real, Real x, Real y, Real z) { return 1 + 2 * x + 3 * y + 5 * z; }
Real fv(Real, Real x, Real y, Real z) { return 2 - x + 4 * y + z; }
Real fw(Real, Real x, Real y, Real z) { return -1 + x + y - 2 * z; }
Real fp(Real x, Real y, Real z) { return 3 + x - y + 2 * z; }

// largest |a - b| over the field, relative to the largest |b|
double rel_diff(const mif::StaggeredTensor &a, const mif::StaggeredTensor &b) {
  double num = 0, den = 0;
  const auto &s = a.sizes();
  for (size_t k = 0; k < s[2]; k++)
    for (size_t j = 0; j < s[1]; j++)
      for (size_t i = 0; i < s[0]; i++) {
        num = std::fmax(num, std::fabs(a(i, j, k) - b(i, j, k)));
        den = std::fmax(den, std::fabs(b(i, j, k)));
      }
  return num / den;
}

int run(int argc, char *argv[]) {
  using namespace mif;
  if (argc != 4) {
    std::fprintf(stderr, "usage: %s <output directory> <N> <periodic z: 0|1>\n", argv[0]);
    return 2;
  }
  const std::string dir = argv[1];
  const size_t N = std::atol(argv[2]);
  const bool pz = std::atoi(argv[3]) != 0;
  const std::array<bool, 3> periodic{false, false, pz};
  // the driver's two geometries: [0,1]x[0,1]x[-1,1] or [-0.5,0.5]^3
  const Constants c(N, N, N, 1.0, 1.0, pz ? 1.0 : 2.0, pz ? -0.5 : 0.0, pz ? -0.5 : 0.0, pz ? -0.5 : -1.0, 1e3, 1e-3,
                    1, 1, 1, 0, periodic);
  PressureSolverStructures structures(c);
  VelocityTensor vel(c);
  StaggeredTensor p(c, StaggeringDirection::none);
  TimeVectorFunction f(fu, fv, fw);
  vel.set(f.set_time(0.0), true);
  p.set(fp, true);

  writeVTK(dir + "/solution.vtk", vel, p);
  writeVTKFullMesh(dir + "/full.vtk", vel, p);
  // lines through the domain's centre (a node when N - 1 is even) and lines
  // halfway between node planes on one or both other axes
  const Real cx = c.min_x_global + c.x_size / 2, cy = c.min_y_global + c.y_size_global / 2,
             cz = c.min_z_global + c.z_size_global / 2;
  writeDat(dir + "/line_x.dat", vel, p, 0, cx, cy, cz);
  writeDat(dir + "/line_y.dat", vel, p, 1, cx, cy, cz);
  writeDat(dir + "/line_z.dat", vel, p, 2, cx, cy, cz);
  writeDat(dir + "/line_x_half.dat", vel, p, 0, cx, cy + c.dy / 2, cz + c.dz / 2);
  writeDat(dir + "/line_y_half.dat", vel, p, 1, cx + c.dx / 2, cy, cz + c.dz / 2);
  writeDat(dir + "/line_z_half.dat", vel, p, 2, cx + c.dx / 2, cy + c.dy / 2, cz);

  int bad = 0;
  // text dump / load: 6 significant digits, as the reference prints
  vel.u.dump(dir + "/u.txt");
  VelocityTensor back(c);
  back.u.load(dir + "/u.txt");
  const double e_text = rel_diff(back.u, vel.u);
  std::printf("text_round_trip %.3e\n", e_text);
  bad += !(e_text < 1e-5);
  // binary dump / load: exact
  for (int q = 0; q < 3; q++) vel.components[q]->dump_binary(dir + "/vel" + std::to_string(q) + ".bin");
  p.dump_binary(dir + "/p.bin");
  StaggeredTensor p2(c, StaggeringDirection::none);
  for (int q = 0; q < 3; q++) bad += !back.components[q]->load_binary(dir + "/vel" + std::to_string(q) + ".bin");
  bad += !p2.load_binary(dir + "/p.bin");
  double e_bin = rel_diff(p2, p);
  for (int q = 0; q < 3; q++) e_bin = std::fmax(e_bin, rel_diff(*back.components[q], *vel.components[q]));
  std::printf("binary_round_trip %.3e\n", e_bin);
  bad += !(e_bin == 0.0);
  // a restart file of another staggering is refused
  const int wrong = back.v.load_binary(dir + "/vel0.bin");
  std::printf("binary_wrong_staggering_refused %d\n", wrong == 0);
  bad += wrong != 0;
  // a plain value list loads into the first values, the rest unchanged
  {
    std::FILE *fl = std::fopen((dir + "/list.txt").c_str(), "w");
    std::fprintf(fl, "1.5 2.5\n-3.25\n");
    std::fclose(fl);
    StaggeredTensor q(c, StaggeringDirection::none);
    q.set(fp, true);
    const Real keep = q(3, 0, 0);
    q.load(dir + "/list.txt");
    const bool ok = q(0, 0, 0) == 1.5 && q(1, 0, 0) == 2.5 && q(2, 0, 0) == -3.25 && q(3, 0, 0) == keep;
    std::printf("list_load %d\n", ok);
    bad += !ok;
  }
  bad += p2.load(dir + "/does_not_exist.txt") != 0;
  return bad == 0 ? 0 : 1;
}

}  // namespace

int main(int argc, char *argv[]) {
  try {
    return run(argc, argv);
  } catch (const std::exception &e) {
    std::fprintf(stderr, "io_test: %s\n", e.what());
    return 1;
  }
}
