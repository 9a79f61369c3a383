// Output writers of the `mif` driver for one rank holding the whole domain
// (src/VTKDatExport.cpp of the reference).  The reference reads its host
// tensors directly; here every writer runs twice over the same sequence of
// points: the first pass records which value of which field it reads, one
// gather per field fetches exactly those values from the device, and the
// second pass computes the printed numbers from them in the same order.
#include "VTKDatExport.h"

#include <algorithm>
#include <cmath>
#include <cstdint>
#include <cstdio>
#include <cstring>
#include <fstream>
#include <numeric>
#include <stdexcept>
#include <tuple>
#include <vector>

namespace mif {

namespace {

// Reads of u, v, w (fields 0..2) and p (field 3), recorded then replayed.
class Fetcher {
 public:
  Fetcher(const VelocityTensor &velocity, const StaggeredTensor &pressure)
      : fields_{&velocity.u, &velocity.v, &velocity.w, &pressure}, ctx_(pressure.constants.context()) {}

  Real operator()(int f, long i, long j, long k) {
    if (recording_) {
      idx_[f].push_back(static_cast<int32_t>(i));
      idx_[f].push_back(static_cast<int32_t>(j));
      idx_[f].push_back(static_cast<int32_t>(k));
      return 0.0;
    }
    return val_[f].at(pos_[f]++);
  }
  bool recording() const { return recording_; }
  void fetch() {
    for (int f = 0; f < 4; f++) {
      const int64_t n = static_cast<int64_t>(idx_[f].size() / 3);
      val_[f].assign(static_cast<size_t>(n), 0.0);
      check_status(mif_field_gather(fields_[f]->device(), idx_[f].data(), n, val_[f].data()), ctx_,
                   "mif_field_gather");
    }
    recording_ = false;
  }

 private:
  const StaggeredTensor *fields_[4];
  const mif_ctx *ctx_;
  std::vector<int32_t> idx_[4];
  std::vector<double> val_[4];
  size_t pos_[4] = {0, 0, 0, 0};
  bool recording_ = true;
};

// Runs body twice: recording the reads, then with the fetched values.
template <typename Body>
void two_passes(Fetcher &fx, Body body) {
  body();
  fx.fetch();
  body();
}

// Index of the pressure plane at or before pos and the weight of that
// plane in a linear interpolation (src/VTKDatExport.cpp:45-51; the weight
// passes through a float there, and here).
std::tuple<size_t, float> pos_to_index(Real pos, Real min_pos_global, Real delta) {
  const Real offset = pos - min_pos_global;
  const Real float_index = offset / delta;
  const Real int_index_1 = std::floor(float_index);
  const Real index_1_importance = 1.0 - (float_index - int_index_1);
  return {static_cast<size_t>(int_index_1), static_cast<float>(index_1_importance)};
}

// The written index ranges at one rank (COMPUTE_INDEXING of the reference
// with Py = Pz = 1): x skips the periodic ghost at 0, y and z start at 0 and
// span the global node counts.
struct Ranges {
  int start_local[3], end_local[3], start_global[3], end_global[3];
  size_t N[3];
};

Ranges ranges_of(const Constants &c) {
  Ranges r;
  r.N[0] = c.Nx_global;
  r.N[1] = c.periodic_bc[1] ? c.Ny_owner + 1 : c.Ny_owner;
  r.N[2] = c.periodic_bc[2] ? c.Nz_owner + 1 : c.Nz_owner;
  r.start_local[0] = c.periodic_bc[0] ? 1 : 0;
  r.start_local[1] = c.prev_proc_y == -1 ? 0 : 1;
  r.start_local[2] = c.prev_proc_z == -1 ? 0 : 1;
  const int base[3] = {c.base_i, c.base_j, c.base_k};
  for (int a = 0; a < 3; a++) {
    r.end_local[a] = r.start_local[a] + static_cast<int>(r.N[a]);
    r.start_global[a] = base[a] + r.start_local[a];
    r.end_global[a] = r.start_global[a] + static_cast<int>(r.N[a]);
  }
  return r;
}

double to_big_endian(double x) {
  uint64_t b;
  std::memcpy(&b, &x, sizeof(b));
  b = __builtin_bswap64(b);
  std::memcpy(&x, &b, sizeof(b));
  return x;
}

void write_values(std::FILE *file, std::vector<Real> values) {
  for (Real &v : values) v = to_big_endian(v);
  std::fwrite(values.data(), sizeof(Real), values.size(), file);
}

std::FILE *open_or_throw(const std::string &filename, const char *mode) {
  std::FILE *file = std::fopen(filename.c_str(), mode);
  if (!file) throw std::runtime_error("cannot open " + filename + " for writing");
  return file;
}

}  // namespace

void writeVTK(const std::string &filename, const VelocityTensor &velocity, const StaggeredTensor &pressure) {
  const Constants &c = velocity.constants;
  const Ranges r = ranges_of(c);
  Fetcher fx(velocity, pressure);
  std::vector<Real> xyz, pu, pv, pw, pp;

  two_passes(fx, [&] {
    xyz.clear();
    pu.clear();
    pv.clear();
    pw.clear();
    pp.clear();
    auto write_point = [&](int i, int j, int k) {
      xyz.push_back(c.min_x_global + (c.base_i + i) * c.dx);
      xyz.push_back(c.min_y_global + (c.base_j + j) * c.dy);
      xyz.push_back(c.min_z_global + (c.base_k + k) * c.dz);
      pu.push_back((fx(0, i, j, k) + fx(0, i + 1, j, k)) / 2);
      pv.push_back((fx(1, i, j, k) + fx(1, i, j + 1, k)) / 2);
      pw.push_back((fx(2, i, j, k) + fx(2, i, j, k + 1)) / 2);
      pp.push_back(fx(3, i, j, k));
    };
    // z = 0 plane first, then x = 0, then y = 0 (src/VTKDatExport.cpp:160-203)
    {
      const int k_global = static_cast<int>(std::get<0>(pos_to_index(0.0, c.min_z_global, c.dz)));
      if (k_global >= r.start_global[2] && k_global < r.end_global[2]) {
        const int k = k_global - r.start_global[2] + r.start_local[2];
        for (int i = r.start_local[0]; i < r.end_local[0]; i++)
          for (int j = r.start_local[1]; j < r.end_local[1]; j++) write_point(i, j, k);
      }
    }
    {
      const int i_global = static_cast<int>(std::get<0>(pos_to_index(0.0, c.min_x_global, c.dx)));
      const int i = i_global - r.start_global[0] + r.start_local[0];
      for (int j = r.start_local[1]; j < r.end_local[1]; j++)
        for (int k = r.start_local[2]; k < r.end_local[2]; k++) write_point(i, j, k);
    }
    {
      const int j_global = static_cast<int>(std::get<0>(pos_to_index(0.0, c.min_y_global, c.dy)));
      if (j_global >= r.start_global[1] && j_global < r.end_global[1]) {
        const int j = j_global - r.start_global[1] + r.start_local[1];
        for (int i = r.start_local[0]; i < r.end_local[0]; i++)
          for (int k = r.start_local[2]; k < r.end_local[2]; k++) write_point(i, j, k);
      }
    }
  });

  const int n = static_cast<int>(pu.size());
  const char *typestr = sizeof(Real) == 8 ? "double" : "float";
  std::FILE *file = open_or_throw(filename, "wb");
  std::fprintf(file, "# vtk DataFile Version 2.0\nvtk output\nBINARY\nDATASET UNSTRUCTURED_GRID \nPOINTS %d %s\n", n,
               typestr);
  write_values(file, xyz);
  std::fprintf(file, "\nPOINT_DATA %d\nSCALARS u %s 1\nLOOKUP_TABLE default\n", n, typestr);
  write_values(file, pu);
  std::fprintf(file, "\nSCALARS v %s 1\nLOOKUP_TABLE default\n", typestr);
  write_values(file, pv);
  std::fprintf(file, "\nSCALARS w %s 1\nLOOKUP_TABLE default\n", typestr);
  write_values(file, pw);
  std::fprintf(file, "\nSCALARS p %s 1\nLOOKUP_TABLE default\n", typestr);
  write_values(file, pp);
  std::fclose(file);
}

void writeDat(const std::string &filename, const VelocityTensor &velocity, const StaggeredTensor &pressure,
              const int direction, const Real x, const Real y, const Real z) {
  if (direction < 0 || direction > 2) throw std::invalid_argument("writeDat: direction must be 0, 1 or 2");
  const Constants &c = velocity.constants;
  const Ranges r = ranges_of(c);
  const auto index_1 = pos_to_index(x, c.min_x_global, c.dx);
  const auto index_2 = pos_to_index(y, c.min_y_global, c.dy);
  const auto index_3 = pos_to_index(z, c.min_z_global, c.dz);
  const int i_global = static_cast<int>(std::get<0>(index_1));
  const int j_global = static_cast<int>(std::get<0>(index_2));
  const int k_global = static_cast<int>(std::get<0>(index_3));
  // weights of the plane at or before the point: 1 (on a plane) or 0.5
  // (halfway between two)
  const Real weight_i = std::get<1>(index_1);
  const Real weight_j = std::get<1>(index_2);
  const Real weight_k = std::get<1>(index_3);
  constexpr Real precision = 1e-6;
  const bool aligned_x = std::abs(weight_i - 1.0) < precision;
  const bool aligned_y = std::abs(weight_j - 1.0) < precision;
  const bool aligned_z = std::abs(weight_k - 1.0) < precision;
  for (const Real wgt : {weight_i, weight_j, weight_k})
    if (!(std::abs(wgt - 1.0) < precision || std::abs(wgt - 0.5) < precision))
      throw std::invalid_argument("writeDat: the line must lie on a pressure plane or halfway between two");

  Fetcher fx(velocity, pressure);
  auto u = [&](int i, int j, int k) { return fx(0, i, j, k); };
  auto v = [&](int i, int j, int k) { return fx(1, i, j, k); };
  auto w = [&](int i, int j, int k) { return fx(2, i, j, k); };
  auto p = [&](int i, int j, int k) { return fx(3, i, j, k); };
  std::vector<Real> coord, pu, pv, pw, pp;

  // src/VTKDatExport.cpp:384-514, including its mixed-up component in the
  // unaligned-x branches of the y and z lines (v read where u is meant)
  two_passes(fx, [&] {
    coord.clear();
    pu.clear();
    pv.clear();
    pw.clear();
    pp.clear();
    if (direction == 0) {
      if (j_global >= r.start_global[1] && j_global < r.end_global[1] && k_global >= r.start_global[2] &&
          k_global < r.end_global[2]) {
        const int j = j_global - r.start_global[1] + r.start_local[1];
        const int k = k_global - r.start_global[2] + r.start_local[2];
        for (int i = r.start_local[0]; i < r.end_local[0]; i++) {
          coord.push_back(c.min_x_global + (c.base_i + i) * c.dx);
          pu.push_back(weight_j * weight_k * (u(i, j, k) + u(i + 1, j, k)) / 2 +
                       (1 - weight_j) * weight_k * (u(i, j + 1, k) + u(i + 1, j + 1, k)) / 2 +
                       weight_j * (1 - weight_k) * (u(i, j, k + 1) + u(i + 1, j, k + 1)) / 2 +
                       (1 - weight_j) * (1 - weight_k) * (u(i, j + 1, k + 1) + u(i + 1, j + 1, k + 1)) / 2);
          if (aligned_y)
            pv.push_back(weight_k * (v(i, j, k) + v(i, j + 1, k)) / 2 +
                         (1 - weight_k) * (v(i, j, k + 1) + v(i, j + 1, k + 1)) / 2);
          else
            pv.push_back(weight_k * v(i, j + 1, k) + (1 - weight_k) * v(i, j + 1, k + 1));
          if (aligned_z)
            pw.push_back(weight_j * (w(i, j, k) + w(i, j, k + 1)) / 2 +
                         (1 - weight_j) * (w(i, j + 1, k) + w(i, j + 1, k + 1)) / 2);
          else
            pw.push_back(weight_j * w(i, j, k + 1) + (1 - weight_j) * w(i, j + 1, k + 1));
          pp.push_back(weight_j * weight_k * p(i, j, k) + (1 - weight_j) * weight_k * p(i, j + 1, k) +
                       weight_j * (1 - weight_k) * p(i, j, k + 1) + (1 - weight_j) * (1 - weight_k) * p(i, j + 1, k + 1));
        }
      }
    } else if (direction == 1) {
      if (i_global >= r.start_global[0] && i_global < r.end_global[0] && k_global >= r.start_global[2] &&
          k_global < r.end_global[2]) {
        const int i = i_global - r.start_global[0] + r.start_local[0];
        const int k = k_global - r.start_global[2] + r.start_local[2];
        for (int j = r.start_local[1]; j < r.end_local[1]; j++) {
          coord.push_back(c.min_y_global + (c.base_j + j) * c.dy);
          if (aligned_x)
            pu.push_back(weight_k * (u(i, j, k) + u(i + 1, j, k)) / 2 +
                         (1 - weight_k) * (u(i, j, k + 1) + u(i + 1, j, k + 1)) / 2);
          else
            pu.push_back(weight_k * u(i + 1, j, k) + (1 - weight_k) * v(i + 1, j, k + 1));
          pv.push_back(weight_i * weight_k * (v(i, j, k) + v(i, j + 1, k)) / 2 +
                       (1 - weight_i) * weight_k * (v(i + 1, j, k) + v(i + 1, j + 1, k)) / 2 +
                       weight_i * (1 - weight_k) * (v(i, j, k + 1) + v(i, j + 1, k + 1)) / 2 +
                       (1 - weight_i) * (1 - weight_k) * (v(i + 1, j, k + 1) + v(i + 1, j + 1, k + 1)) / 2);
          if (aligned_z)
            pw.push_back(weight_i * (w(i, j, k) + w(i, j, k + 1)) / 2 +
                         (1 - weight_i) * (w(i + 1, j, k) + w(i + 1, j, k + 1)) / 2);
          else
            pw.push_back(weight_i * w(i, j, k + 1) + (1 - weight_i) * w(i + 1, j, k + 1));
          pp.push_back(weight_i * weight_k * p(i, j, k) + (1 - weight_i) * weight_k * p(i + 1, j, k) +
                       weight_i * (1 - weight_k) * p(i, j, k + 1) + (1 - weight_i) * (1 - weight_k) * p(i + 1, j, k + 1));
        }
      }
    } else {
      if (i_global >= r.start_global[0] && i_global < r.end_global[0] && j_global >= r.start_global[1] &&
          j_global < r.end_global[1]) {
        const int i = i_global - r.start_global[0] + r.start_local[0];
        const int j = j_global - r.start_global[1] + r.start_local[1];
        for (int k = r.start_local[2]; k < r.end_local[2]; k++) {
          coord.push_back(c.min_z_global + (c.base_k + k) * c.dz);
          if (aligned_x)
            pu.push_back(weight_j * (u(i, j, k) + u(i + 1, j, k)) / 2 +
                         (1 - weight_j) * (u(i, j + 1, k) + u(i + 1, j + 1, k)) / 2);
          else
            pu.push_back(weight_j * u(i + 1, j, k) + (1 - weight_j) * v(i + 1, j + 1, k));
          if (aligned_y)
            pv.push_back(weight_i * (v(i, j, k) + v(i, j + 1, k)) / 2 +
                         (1 - weight_i) * (v(i + 1, j, k) + v(i + 1, j + 1, k)) / 2);
          else
            pv.push_back(weight_i * v(i, j + 1, k) + (1 - weight_i) * v(i + 1, j + 1, k));
          pw.push_back(weight_i * weight_j * (w(i, j, k) + w(i, j, k + 1)) / 2 +
                       (1 - weight_i) * weight_j * (w(i + 1, j, k) + w(i + 1, j, k + 1)) / 2 +
                       weight_i * (1 - weight_j) * (w(i, j + 1, k) + w(i, j + 1, k + 1)) / 2 +
                       (1 - weight_i) * (1 - weight_j) * (w(i + 1, j + 1, k) + w(i + 1, j + 1, k + 1)) / 2);
          pp.push_back(weight_i * weight_j * p(i, j, k) + (1 - weight_i) * weight_j * p(i + 1, j, k) +
                       weight_i * (1 - weight_j) * p(i, j + 1, k) + (1 - weight_i) * (1 - weight_j) * p(i + 1, j + 1, k));
        }
      }
    }
  });

  // sorted by the running coordinate (stable, as the reference's insertion sort)
  std::vector<size_t> order(coord.size());
  std::iota(order.begin(), order.end(), size_t(0));
  std::stable_sort(order.begin(), order.end(), [&](size_t a, size_t b) { return coord[a] < coord[b]; });
  std::FILE *file = open_or_throw(filename, "w");
  for (size_t q : order) {
    const Real X = direction == 0 ? coord[q] : x, Y = direction == 1 ? coord[q] : y, Z = direction == 2 ? coord[q] : z;
    std::fprintf(file, "%.8f %.8f %.8f %.8e %.8e %.8e %.8e\n", X, Y, Z, pu[q], pv[q], pw[q], pp[q]);
  }
  std::fclose(file);
}

void writeVTKFullMesh(const std::string &filename, const VelocityTensor &velocity, const StaggeredTensor &pressure) {
  const Constants &c = velocity.constants;
  const Ranges r = ranges_of(c);
  Fetcher fx(velocity, pressure);
  std::vector<Real> pu, pv, pw, pp;
  two_passes(fx, [&] {
    pu.clear();
    pv.clear();
    pw.clear();
    pp.clear();
    for (int k = r.start_local[2]; k < r.end_local[2]; ++k)
      for (int j = r.start_local[1]; j < r.end_local[1]; ++j)
        for (int i = r.start_local[0]; i < r.end_local[0]; ++i) {
          pu.push_back((fx(0, i, j, k) + fx(0, i + 1, j, k)) / 2);
          pv.push_back((fx(1, i, j, k) + fx(1, i, j + 1, k)) / 2);
          pw.push_back((fx(2, i, j, k) + fx(2, i, j, k + 1)) / 2);
          pp.push_back(fx(3, i, j, k));
        }
  });
  std::ofstream out(filename);
  if (!out) throw std::runtime_error("cannot open " + filename + " for writing");
  out << "# vtk DataFile Version 3.0\n"
      << "pressure mesh solution\n"
      << "ASCII\n"
      << "DATASET STRUCTURED_POINTS\n"
      << "DIMENSIONS " << r.N[0] << ' ' << r.N[1] << ' ' << r.N[2] << '\n'
      << "ORIGIN " << c.min_x_global << " " << c.min_y_global << " " << c.min_z_global << "\n"
      << "SPACING " << c.dx << ' ' << c.dy << ' ' << c.dz << '\n'
      << "POINT_DATA " << r.N[0] * r.N[1] * r.N[2] << '\n';
  auto block = [&](const char *name, const std::vector<Real> &vals) {
    out << "SCALARS " << name << " double 1\n"
        << "LOOKUP_TABLE default\n";
    for (const Real v : vals) out << v << ' ';
  };
  block("u", pu);
  block("v", pv);
  block("w", pw);
  std::vector<Real> mag(pu.size());
  for (size_t q = 0; q < pu.size(); q++) mag[q] = std::sqrt(pu[q] * pu[q] + pv[q] * pv[q] + pw[q] * pw[q]);
  block("|u|", mag);
  block("p", pp);
}

}  // namespace mif
