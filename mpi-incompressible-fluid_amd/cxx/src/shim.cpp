// The reference's host interface for the timestep path over the C ABI of
// the HIP library.  Every numerical routine runs on the device; the host
// side keeps the reference's objects, converts callbacks to descriptors and
// tabulates callbacks that have none.
#include <cstring>
#include <fstream>
#include <iostream>
#include <memory>
#include <sstream>
#include <stdexcept>
#include <string>

#include "Constants.h"
#include "Norms.h"
#include "PressureEquation.h"
#include "PressureGradient.h"
#include "PressureSolverStructures.h"
#include "PressureTensor.h"
#include "StaggeredTensor.h"
#include "Timestep.h"
#include "TimestepVelocity.h"
#include "VelocityTensor.h"

namespace mif {

void check_status(int rc, const mif_ctx *ctx, const char *what) {
  if (rc == MIF_OK) return;
  const char *msg = mif_last_error(ctx);
  throw std::runtime_error(std::string(what) + ": " + (msg ? msg : "error") + " (status " + std::to_string(rc) + ")");
}

namespace {

mif_grid grid_of(size_t nx, size_t ny, size_t nz, Real lx, Real ly, Real lz, Real x0, Real y0, Real z0, Real Re,
                 Real T, unsigned steps, const std::array<bool, 3> &periodic) {
  mif_grid g{};
  g.Nx_global = nx;
  g.Ny_global = ny;
  g.Nz_global = nz;
  g.x_size = lx;
  g.y_size = ly;
  g.z_size = lz;
  g.min_x = x0;
  g.min_y = y0;
  g.min_z = z0;
  g.Re = Re;
  g.final_time = T;
  g.num_time_steps = steps;
  for (int a = 0; a < 3; a++) g.periodic[a] = periodic[a] ? 1 : 0;
  return g;
}

size_t periodic_extra(bool p) { return p ? 1 : 0; }

}  // namespace

// ---- Constants -----------------------------------------------------------------
// Derived values as src/Constants.cpp:63-101 computes them, at one rank.
Constants::Constants(size_t Nx_global, size_t Ny_global, size_t Nz_global, Real x_size, Real y_size_global,
                     Real z_size_global, Real min_x_global, Real min_y_global, Real min_z_global, Real Re,
                     Real final_time, unsigned int num_time_steps, int Py, int Pz, int rank,
                     const std::array<bool, 3> &periodic_bc, int precision, int device)
    : Nx_global(Nx_global),
      Ny_global(Ny_global),
      Nz_global(Nz_global),
      x_size(x_size),
      y_size_global(y_size_global),
      z_size_global(z_size_global),
      min_x_global(min_x_global),
      min_y_global(min_y_global),
      min_z_global(min_z_global),
      Re(Re),
      final_time(final_time),
      num_time_steps(num_time_steps),
      periodic_bc(periodic_bc),
      Py(Py),
      Pz(Pz),
      rank(rank),
      y_rank(0),
      z_rank(0),
      dt(final_time / num_time_steps),
      Nx_domains(Nx_global - 1),
      Ny_domains_global(Ny_global - 1),
      Nz_domains_global(Nz_global - 1),
      dx(x_size / Nx_domains),
      dy(y_size_global / Ny_domains_global),
      dz(z_size_global / Nz_domains_global),
      one_over_2_dx(1 / (2 * dx)),
      one_over_2_dy(1 / (2 * dy)),
      one_over_2_dz(1 / (2 * dz)),
      one_over_8_dx(1 / (8 * dx)),
      one_over_8_dy(1 / (8 * dy)),
      one_over_8_dz(1 / (8 * dz)),
      one_over_dx2_Re(1 / (Re * dx * dx)),
      one_over_dy2_Re(1 / (Re * dy * dy)),
      one_over_dz2_Re(1 / (Re * dz * dz)),
      dx_over_2(dx / 2),
      dy_over_2(dy / 2),
      dz_over_2(dz / 2),
      one_over_dx(1 / dx),
      one_over_dy(1 / dy),
      one_over_dz(1 / dz),
      P(1),
      Ny_owner(Ny_global - periodic_extra(periodic_bc[1])),
      Nz_owner(Nz_global - periodic_extra(periodic_bc[2])),
      Nx(Nx_global + periodic_extra(periodic_bc[0])),
      Ny(Ny_global + periodic_extra(periodic_bc[1])),
      Nz(Nz_global + periodic_extra(periodic_bc[2])),
      Nx_staggered(periodic_bc[0] ? Nx : Nx + 1),
      Ny_staggered(Ny + 1),
      Nz_staggered(Nz + 1),
      base_i(periodic_bc[0] ? -1 : 0),
      base_j(periodic_bc[1] ? -1 : 0),
      base_k(periodic_bc[2] ? -1 : 0),
      prev_proc_y(-1),
      next_proc_y(-1),
      prev_proc_z(-1),
      next_proc_z(-1) {
  if (Py != 1 || Pz != 1 || rank != 0)
    throw std::invalid_argument("the solver holds the whole domain on one GPU: Py = Pz = 1 and rank 0 are required");
  const mif_grid g = grid_of(Nx_global, Ny_global, Nz_global, x_size, y_size_global, z_size_global, min_x_global,
                             min_y_global, min_z_global, Re, final_time, num_time_steps, periodic_bc);
  check_status(mif_ctx_create(&g, precision, device, &ctx_), nullptr, "mif_ctx_create");
}

Constants::~Constants() { mif_ctx_destroy(ctx_); }

// ---- StaggeredTensor -----------------------------------------------------------
StaggeredTensor::StaggeredTensor(const Constants &constants, const StaggeringDirection &staggering)
    : constants(constants), staggering(staggering) {
  check_status(mif_field_create(constants.context(), static_cast<int>(staggering), &field_), constants.context(),
               "mif_field_create");
  uint64_t e[3];
  check_status(mif_field_extents(field_, e), constants.context(), "mif_field_extents");
  for (int a = 0; a < 3; a++) sizes_[a] = static_cast<size_t>(e[a]);
  // the device field starts at zero, as the reference's tensors do
  host_.assign(sizes_[0] * sizes_[1] * sizes_[2], 0.0);
  state_ = SYNCED;
}

StaggeredTensor::~StaggeredTensor() { mif_field_destroy(field_); }

void StaggeredTensor::pull() const {
  if (state_ != DEVICE_NEWER) return;
  check_status(mif_field_download(field_, host_.data()), constants.context(), "mif_field_download");
  state_ = SYNCED;
}

mif_field *StaggeredTensor::device() const {
  if (state_ == HOST_NEWER) {
    check_status(mif_field_upload(field_, host_.data()), constants.context(), "mif_field_upload");
    state_ = SYNCED;
  }
  return field_;
}

void StaggeredTensor::apply_periodic_bc() {
  check_status(mif_field_apply_periodic_bc(device()), constants.context(), "apply_periodic_bc");
  mark_device_written();
}

void StaggeredTensor::set(const std::function<Real(Real, Real, Real)> &f, bool include_border) {
  pull();
  const size_t lo = include_border ? 0 : 1;
  for (size_t k = lo; k + lo < sizes_[2]; k++)
    for (size_t j = lo; j + lo < sizes_[1]; j++)
      for (size_t i = lo; i + lo < sizes_[0]; i++)
        host_[i + sizes_[0] * (j + sizes_[1] * k)] =
            evaluate_function_at_index(static_cast<int>(i), static_cast<int>(j), static_cast<int>(k), f);
  state_ = HOST_NEWER;
}

void StaggeredTensor::swap_data(StaggeredTensor &other) {
  if (other.staggering != staggering) throw std::invalid_argument("swap_data: different staggering");
  check_status(mif_field_swap(field_, other.field_), constants.context(), "mif_field_swap");
  host_.swap(other.host_);
  std::swap(state_, other.state_);
}

// ---- dump / load -----------------------------------------------------------------
uint8_t StaggeredTensor::dump(const std::string &file_name) const {
  std::ofstream out(file_name);
  if (!out) {
    std::cerr << "Error opening " << file_name << " file." << std::endl;
    return 0;
  }
  pull();
  out << "Tensor (row major). Dimensions: [ ";
  for (size_t v : sizes_) out << v << " ";
  out << "]" << std::endl << "[ ";
  for (const Real v : host_) out << v << " ";
  out << "]" << std::endl;
  return 1;
}

uint8_t StaggeredTensor::load(const std::string &file_name) {
  std::ifstream in(file_name);
  if (!in) {
    std::cerr << "Error opening " << file_name << " file." << std::endl;
    return 0;
  }
  std::stringstream whole;
  whole << in.rdbuf();
  std::string text = whole.str();
  // a dump's header ends with the dimensions' closing bracket; its values
  // follow the next opening bracket
  if (text.compare(0, 6, "Tensor") == 0) {
    const size_t close = text.find(']');
    const size_t open = close == std::string::npos ? close : text.find('[', close);
    text = open == std::string::npos ? std::string() : text.substr(open + 1);
  }
  pull();
  std::istringstream values(text);
  for (size_t q = 0; q < host_.size(); q++) {
    Real v;
    if (!(values >> v)) break;
    host_[q] = v;
  }
  state_ = HOST_NEWER;
  return 1;
}

namespace {
constexpr char kBinaryMagic[8] = {'M', 'I', 'F', 'T', 'N', 'S', 'R', '1'};
}

uint8_t StaggeredTensor::dump_binary(const std::string &file_name) const {
  std::ofstream out(file_name, std::ios::binary);
  if (!out) {
    std::cerr << "Error opening " << file_name << " file." << std::endl;
    return 0;
  }
  pull();
  const int64_t head[5] = {static_cast<int64_t>(sizes_[0]), static_cast<int64_t>(sizes_[1]),
                           static_cast<int64_t>(sizes_[2]), static_cast<int64_t>(staggering),
                           static_cast<int64_t>(sizeof(Real))};
  out.write(kBinaryMagic, sizeof(kBinaryMagic));
  out.write(reinterpret_cast<const char *>(head), sizeof(head));
  out.write(reinterpret_cast<const char *>(host_.data()), static_cast<std::streamsize>(host_.size() * sizeof(Real)));
  return out ? 1 : 0;
}

uint8_t StaggeredTensor::load_binary(const std::string &file_name) {
  std::ifstream in(file_name, std::ios::binary);
  if (!in) {
    std::cerr << "Error opening " << file_name << " file." << std::endl;
    return 0;
  }
  char magic[8];
  int64_t head[5];
  in.read(magic, sizeof(magic));
  in.read(reinterpret_cast<char *>(head), sizeof(head));
  if (!in || std::memcmp(magic, kBinaryMagic, sizeof(magic)) != 0 || head[0] != static_cast<int64_t>(sizes_[0]) ||
      head[1] != static_cast<int64_t>(sizes_[1]) || head[2] != static_cast<int64_t>(sizes_[2]) ||
      head[3] != static_cast<int64_t>(staggering) || head[4] != static_cast<int64_t>(sizeof(Real))) {
    std::cerr << file_name << ": not a restart file of a tensor with these extents and staggering" << std::endl;
    return 0;
  }
  std::vector<Real> values(host_.size());
  in.read(reinterpret_cast<char *>(values.data()), static_cast<std::streamsize>(values.size() * sizeof(Real)));
  if (!in) {
    std::cerr << file_name << ": truncated restart file" << std::endl;
    return 0;
  }
  host_.swap(values);
  state_ = HOST_NEWER;
  return 1;
}

// ---- VelocityTensor ------------------------------------------------------------
VelocityTensor::VelocityTensor(const Constants &constants)
    : u(constants), v(constants), w(constants), components{&u, &v, &w}, constants(constants) {}

void VelocityTensor::swap_data(VelocityTensor &other) {
  u.swap_data(other.u);
  v.swap_data(other.v);
  w.swap_data(other.w);
}

std::array<mif_field *, 3> VelocityTensor::device() const { return {u.device(), v.device(), w.device()}; }

void VelocityTensor::mark_device_written() const {
  u.mark_device_written();
  v.mark_device_written();
  w.mark_device_written();
}

void VelocityTensor::set(const VectorFunction &f, bool include_border) {
  if (f.has_descriptor()) {
    const mif_function d = f.descriptor();
    for (int c = 0; c < 3; c++) {
      mif_field *fld = components[c]->device();
      check_status(mif_field_set(fld, &d, c, f.time(), include_border ? 1 : 0), constants.context(), "mif_field_set");
      components[c]->mark_device_written();
    }
    return;
  }
  for (int c = 0; c < 3; c++) components[c]->set(*f.components[c], include_border);
}

namespace {

// ---- boundary data of callbacks without a descriptor ----------------------
// The device evaluates the named exact solutions itself.  Any other callback
// is evaluated here, on the host, into the face tables of the C ABI's
// *_tabulated entry points (layout in mif_gpu.h): O(N^2) values per stage.

// The two axes other than `axis`, lower first.
void other_axes(int axis, int *a, int *b) {
  *a = axis == 0 ? 1 : 0;
  *b = axis == 2 ? 1 : 2;
}

// Appends one velocity face table: the values VelocityTensor::apply_bc
// writes (src/VelocityTensor.cpp:36-233).  On a face of axis `axis` the
// tangential components take f at their own points of that face, at index 0
// or extent - 1; the normal component takes f at the wall node (index 0, or
// the last node N - 1) plus or minus h/2 times the tangential divergence of
// f taken at the components' staggered points of that node.
void append_velocity_faces(const VelocityTensor &vel, const VectorFunction &f, std::vector<double> &out) {
  const Constants &c = vel.constants;
  const size_t nodes[3] = {c.Nx, c.Ny, c.Nz};
  const Real half[3] = {c.dx_over_2, c.dy_over_2, c.dz_over_2};
  const Real inv_h[3] = {c.one_over_dx, c.one_over_dy, c.one_over_dz};
  for (int q = 0; q < 3; q++) {
    const StaggeredTensor &t = *vel.components[q];
    const auto &e = t.sizes();
    const SpaceFunction &fq = *f.components[q];
    for (int face = 0; face < 6; face++) {
      const int axis = 2 - face / 2;
      const bool hi = face & 1;
      if (c.periodic_bc[axis]) continue;
      int o1, o2;
      other_axes(axis, &o1, &o2);
      for (size_t b = 0; b < e[o2]; b++)
        for (size_t a = 0; a < e[o1]; a++) {
          int idx[3];
          idx[o1] = static_cast<int>(a);
          idx[o2] = static_cast<int>(b);
          if (q != axis) {
            idx[axis] = hi ? static_cast<int>(e[axis]) - 1 : 0;
            out.push_back(t.evaluate_function_at_index(idx[0], idx[1], idx[2], fq));
            continue;
          }
          idx[axis] = hi ? static_cast<int>(nodes[axis]) - 1 : 0;
          const Real at_boundary = t.evaluate_function_at_index_unstaggered(idx[0], idx[1], idx[2], fq);
          Real div = 0;
          for (const int o : {o1, o2}) {
            const StaggeredTensor &s = *vel.components[o];
            int nxt[3] = {idx[0], idx[1], idx[2]};
            nxt[o] += 1;
            const Real d = (s.evaluate_function_at_index(nxt[0], nxt[1], nxt[2], *f.components[o]) -
                            s.evaluate_function_at_index(idx[0], idx[1], idx[2], *f.components[o])) *
                           inv_h[o];
            div = o == o1 ? d : div + d;
          }
          out.push_back(hi ? at_boundary - half[axis] * div : at_boundary + half[axis] * div);
        }
    }
  }
}

// Appends one Neumann table: the normal component of g at every node of the
// six wall faces of the node extents (apply_non_homogeneous_neumann,
// src/PressureEquation.cpp:10-57, reads g there).
void append_neumann_faces(const StaggeredTensor &pressure, const VectorFunction &g, std::vector<double> &out) {
  const auto &e = pressure.sizes();
  for (int face = 0; face < 6; face++) {
    const int axis = 2 - face / 2;
    const bool hi = face & 1;
    int o1, o2;
    other_axes(axis, &o1, &o2);
    for (size_t b = 0; b < e[o2]; b++)
      for (size_t a = 0; a < e[o1]; a++) {
        int idx[3];
        idx[o1] = static_cast<int>(a);
        idx[o2] = static_cast<int>(b);
        idx[axis] = hi ? static_cast<int>(e[axis]) - 1 : 0;
        out.push_back(pressure.evaluate_function_at_index(idx[0], idx[1], idx[2], *g.components[axis]));
      }
  }
}

// The three stage times of FULL_TIMESTEP (src/Timestep.cpp:99-104): t_n +
// 64/120 dt, then + 16/120 dt, then t_n + dt; index 0 is t_n.
std::array<Real, 4> full_step_times(Real dt, Real t_n) {
  const Real time_1 = t_n + 64.0 / 120.0 * dt;
  const Real time_2 = time_1 + 16.0 / 120.0 * dt;
  return {t_n, time_1, time_2, t_n + dt};
}

}  // namespace

void VelocityTensor::apply_bc(const VectorFunction &exact_velocity) {
  auto f = device();
  if (exact_velocity.has_descriptor()) {
    const mif_function d = exact_velocity.descriptor();
    check_status(mif_velocity_apply_bc(f.data(), &d, exact_velocity.time()), constants.context(), "apply_bc");
  } else {
    std::vector<double> faces;
    append_velocity_faces(*this, exact_velocity, faces);
    check_status(mif_velocity_apply_bc_tabulated(f.data(), faces.data()), constants.context(), "apply_bc");
  }
  mark_device_written();
}

// ---- pressure solver objects ---------------------------------------------------
PressureSolverStructures::PressureSolverStructures(const Constants &constants)
    : constants(constants),
      periodic_bc{constants.periodic_bc[0], constants.periodic_bc[1], constants.periodic_bc[2]},
      Nx_points(static_cast<int>(constants.Nx_global - (constants.periodic_bc[0] ? 1 : 0))),
      Ny_points(static_cast<int>(constants.Ny_global - (constants.periodic_bc[1] ? 1 : 0))),
      Nz_points(static_cast<int>(constants.Nz_global - (constants.periodic_bc[2] ? 1 : 0))) {}

PressureTensor::PressureTensor(PressureSolverStructures &structures)
    : Tensor<Real, 1U, int>({structures.Nx_points * structures.Ny_points * structures.Nz_points}),
      structures(structures),
      max_size(structures.Nx_points * structures.Ny_points * structures.Nz_points) {}

void PressureTensor::copy_from_staggered(const StaggeredTensor &other) {
  const auto &s = other.sizes();
  const auto &pbc = other.constants.periodic_bc;
  int q = 0;
  for (size_t k = pbc[2]; k < s[2] - pbc[2]; k++)
    for (size_t j = pbc[1]; j < s[1] - pbc[1]; j++)
      for (size_t i = pbc[0]; i < s[0] - pbc[0]; i++) (*this)(q++) = other(i, j, k);
}

void PressureTensor::copy_to_staggered(StaggeredTensor &other, int) const {
  const auto s = other.sizes();
  const auto &pbc = other.constants.periodic_bc;
  int q = 0;
  for (size_t k = pbc[2]; k < s[2] - pbc[2]; k++)
    for (size_t j = pbc[1]; j < s[1] - pbc[1]; j++)
      for (size_t i = pbc[0]; i < s[0] - pbc[0]; i++) other(i, j, k) = (*this)(q++);
  other.apply_periodic_bc();
}

// ---- timestep ------------------------------------------------------------------
void timestep(VelocityTensor &velocity, VelocityTensor &velocity_buffer, VelocityTensor &velocity_buffer_2,
              const TimeVectorFunction &exact_velocity, Real t_n, StaggeredTensor &pressure,
              StaggeredTensor &pressure_buffer, PressureTensor &) {
  auto a = velocity.device(), b = velocity_buffer.device(), c = velocity_buffer_2.device();
  if (exact_velocity.has_descriptor()) {
    const mif_function ev = exact_velocity.descriptor();
    check_status(mif_timestep(a.data(), b.data(), c.data(), &ev, t_n, pressure.device(), pressure_buffer.device()),
                 velocity.constants.context(), "timestep");
  } else {
    const auto times = full_step_times(velocity.constants.dt, t_n);
    std::vector<double> faces;
    for (int s = 1; s <= 3; s++) append_velocity_faces(velocity, exact_velocity.set_time(times[s]), faces);
    check_status(
        mif_timestep_tabulated(a.data(), b.data(), c.data(), faces.data(), t_n, pressure.device(), pressure_buffer.device()),
        velocity.constants.context(), "timestep");
  }
  velocity.mark_device_written();
  velocity_buffer.mark_device_written();
  velocity_buffer_2.mark_device_written();
  pressure.mark_device_written();
  pressure_buffer.mark_device_written();
}

void timestep_nhn(VelocityTensor &velocity, VelocityTensor &velocity_buffer, VelocityTensor &velocity_buffer_2,
                  const TimeVectorFunction &exact_velocity, const TimeVectorFunction &exact_pressure_gradient,
                  Real t_n, StaggeredTensor &pressure, StaggeredTensor &pressure_buffer, PressureTensor &) {
  auto a = velocity.device(), b = velocity_buffer.device(), c = velocity_buffer_2.device();
  if (exact_velocity.has_descriptor() && exact_pressure_gradient.has_descriptor()) {
    const mif_function ev = exact_velocity.descriptor();
    const mif_function gp = exact_pressure_gradient.descriptor();
    check_status(
        mif_timestep_nhn(a.data(), b.data(), c.data(), &ev, &gp, t_n, pressure.device(), pressure_buffer.device()),
        velocity.constants.context(), "timestep_nhn");
  } else {
    // PRESSURE_EQUATION_true hands stage s's solve
    // exact_pressure_gradient.get_difference_over_time(new_time, prev_time)
    // (src/Timestep.cpp:89-95)
    const auto times = full_step_times(velocity.constants.dt, t_n);
    std::vector<double> faces, neumann;
    for (int s = 1; s <= 3; s++) {
      append_velocity_faces(velocity, exact_velocity.set_time(times[s]), faces);
      append_neumann_faces(pressure, exact_pressure_gradient.get_difference_over_time(times[s], times[s - 1]),
                           neumann);
    }
    check_status(mif_timestep_nhn_tabulated(a.data(), b.data(), c.data(), faces.data(), neumann.data(), t_n,
                                            pressure.device(), pressure_buffer.device()),
                 velocity.constants.context(), "timestep_nhn");
  }
  velocity.mark_device_written();
  velocity_buffer.mark_device_written();
  velocity_buffer_2.mark_device_written();
  pressure.mark_device_written();
  pressure_buffer.mark_device_written();
}

void timestep_velocity(VelocityTensor &velocity, VelocityTensor &velocity_buffer, VelocityTensor &rhs_buffer,
                       const TimeVectorFunction &exact_velocity, Real t_n) {
  mif_function forcing{};
  forcing.kind = MIF_FN_VT_FORCING;
  forcing.Re = current_reynolds();
  auto a = velocity.device(), b = velocity_buffer.device(), c = rhs_buffer.device();
  if (exact_velocity.has_descriptor()) {
    const mif_function ev = exact_velocity.descriptor();
    check_status(mif_timestep_velocity(a.data(), b.data(), c.data(), &ev, &forcing, t_n),
                 velocity.constants.context(), "timestep_velocity");
  } else {
    // boundary data at t_n + 8/15 dt, t_n + 2/3 dt, t_n + dt
    // (src/TimestepVelocity.cpp:60-90)
    const Real dt = velocity.constants.dt;
    const Real times[3] = {t_n + (8.0 / 15.0) * dt, t_n + (2.0 / 3.0) * dt, t_n + dt};
    std::vector<double> faces;
    for (const Real t : times) append_velocity_faces(velocity, exact_velocity.set_time(t), faces);
    check_status(mif_timestep_velocity_tabulated(a.data(), b.data(), c.data(), faces.data(), &forcing, t_n),
                 velocity.constants.context(), "timestep_velocity");
  }
  // the library exchanged the storage of velocity and velocity_buffer
  // behind the same field handles; all three now hold device results
  velocity.mark_device_written();
  velocity_buffer.mark_device_written();
  rhs_buffer.mark_device_written();
}

// ---- pressure equation ---------------------------------------------------------
void solve_pressure_equation_homogeneous_periodic(StaggeredTensor &pressure, PressureTensor &,
                                                  const VelocityTensor &velocity, Real dt) {
  auto v = velocity.device();
  check_status(mif_solve_pressure_homogeneous_periodic(pressure.device(), v.data(), dt), pressure.constants.context(),
               "solve_pressure_equation_homogeneous_periodic");
  pressure.mark_device_written();
}

void solve_pressure_equation_non_homogeneous_neumann(StaggeredTensor &pressure, PressureTensor &,
                                                     const VelocityTensor &velocity,
                                                     const VectorFunction &exact_pressure_gradient, Real dt) {
  auto v = velocity.device();
  if (exact_pressure_gradient.has_descriptor()) {
    const mif_function gp = exact_pressure_gradient.descriptor();
    check_status(mif_solve_pressure_non_homogeneous_neumann(pressure.device(), v.data(), &gp,
                                                           exact_pressure_gradient.time(), dt),
                 pressure.constants.context(), "solve_pressure_equation_non_homogeneous_neumann");
  } else {
    std::vector<double> neumann;
    append_neumann_faces(pressure, exact_pressure_gradient, neumann);
    check_status(mif_solve_pressure_non_homogeneous_neumann_tabulated(pressure.device(), v.data(), neumann.data(), dt),
                 pressure.constants.context(), "solve_pressure_equation_non_homogeneous_neumann");
  }
  pressure.mark_device_written();
}

namespace {

// A node field of the pressure's context holding f at every node.
std::unique_ptr<StaggeredTensor> tabulate_nodes(const Constants &c, const std::function<Real(Real, Real, Real)> &f) {
  auto t = std::make_unique<StaggeredTensor>(c, StaggeringDirection::none);
  t->set(f, true);
  return t;
}

}  // namespace

void adjust_pressure(StaggeredTensor &pressure, const std::function<Real(Real, Real, Real)> &exact_pressure) {
  auto table = tabulate_nodes(pressure.constants, exact_pressure);
  check_status(mif_adjust_pressure_tabulated(pressure.device(), table->device()), pressure.constants.context(),
               "adjust_pressure");
  pressure.mark_device_written();
}

// ---- norms ---------------------------------------------------------------------
namespace {

std::array<Real, 3> velocity_norms(const VelocityTensor &velocity, const TimeVectorFunction &exact, Real time) {
  std::array<Real, 3> out{};
  auto v = velocity.device();
  if (exact.has_descriptor()) {
    const mif_function d = exact.descriptor();
    check_status(mif_error_norms_velocity(v.data(), &d, time, out.data()), velocity.constants.context(),
                 "error norms");
    return out;
  }
  std::unique_ptr<StaggeredTensor> tab[3];
  mif_field *f[3];
  for (int c = 0; c < 3; c++) {
    const TimeFunction &fc = *exact.components[c];
    tab[c] = tabulate_nodes(velocity.constants, [&fc, time](Real x, Real y, Real z) { return fc(time, x, y, z); });
    f[c] = tab[c]->device();
  }
  check_status(mif_error_norms_velocity_tabulated(v.data(), f, out.data()), velocity.constants.context(),
               "error norms");
  return out;
}

std::array<Real, 3> pressure_norms(const StaggeredTensor &pressure, const TimeFunction &exact, Real time) {
  std::array<Real, 3> out{};
  mif_function d;
  if (describe_scalar(exact, &d)) {
    check_status(mif_error_norms_pressure(pressure.device(), &d, time, out.data()), pressure.constants.context(),
                 "error norms");
    return out;
  }
  auto tab = tabulate_nodes(pressure.constants, [&exact, time](Real x, Real y, Real z) { return exact(time, x, y, z); });
  check_status(mif_error_norms_pressure_tabulated(pressure.device(), tab->device(), out.data()),
               pressure.constants.context(), "error norms");
  return out;
}

}  // namespace

Real ErrorL1Norm(const VelocityTensor &v, const TimeVectorFunction &e, Real t) { return velocity_norms(v, e, t)[0]; }
Real ErrorL2Norm(const VelocityTensor &v, const TimeVectorFunction &e, Real t) { return velocity_norms(v, e, t)[1]; }
Real ErrorLInfNorm(const VelocityTensor &v, const TimeVectorFunction &e, Real t) { return velocity_norms(v, e, t)[2]; }

Real ErrorL1Norm(const StaggeredTensor &p, const TimeFunction &e, Real t) { return pressure_norms(p, e, t)[0]; }
Real ErrorL2Norm(const StaggeredTensor &p, const TimeFunction &e, Real t) { return pressure_norms(p, e, t)[1]; }
Real ErrorLInfNorm(const StaggeredTensor &p, const TimeFunction &e, Real t) { return pressure_norms(p, e, t)[2]; }

Real accumulate_error_mpi_l1(Real local_error, const Constants &) { return local_error; }
Real accumulate_error_mpi_l2(Real local_error, const Constants &) { return local_error; }
Real accumulate_error_mpi_linf(Real local_error, const Constants &) { return local_error; }

// ---- pressure gradient -----------------------------------------------------------
void pressure_gradient(const StaggeredTensor &pressure, VelocityTensor &out) {
  auto o = out.device();
  check_status(mif_pressure_gradient(pressure.device(), o.data()), pressure.constants.context(), "pressure_gradient");
  out.mark_device_written();
}

}  // namespace mif
