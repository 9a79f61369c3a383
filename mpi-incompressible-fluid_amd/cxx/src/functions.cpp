// Host side of the analytic functions: the closed forms behind the named
// exact solutions, the lid data of the driver cases, and the callback
// triples that carry device descriptors.
#include <cmath>
#include <stdexcept>

#include "Manufactured.h"
#include "ManufacturedPressure.h"
#include "ManufacturedVelocity.h"
#include "TestCaseBoundaries.h"
#include "VectorFunction.h"

// The generated functions read a global Reynolds number that the calling
// program defines (include/Manufactured.h:4).  Programs that never use them
// need not define it, so the reference is weak.
#pragma weak Reynolds

namespace {

constexpr double kPi = 3.14159265358979323846;
constexpr double kA = kPi / 4;  // Ethier-Steinman a
constexpr double kD = kPi / 2;  // Ethier-Steinman d

// Ethier-Steinman velocity component with the arguments rotated so that
// (x, y, z) -> (y, z, x) gives v and (z, x, y) gives w.
double es_component(double t, double x, double y, double z) {
  return -kA * (std::exp(kA * x) * std::sin(kA * y + kD * z) + std::exp(kA * z) * std::cos(kA * x + kD * y)) *
         std::exp(-kD * kD * t / mif::current_reynolds());
}

double es_pressure(double t, double x, double y, double z) {
  const double s = std::exp(2 * kA * x) + std::exp(2 * kA * y) + std::exp(2 * kA * z) +
                   2 * std::sin(kA * x + kD * y) * std::cos(kA * z + kD * x) * std::exp(kA * (y + z)) +
                   2 * std::sin(kA * y + kD * z) * std::cos(kA * x + kD * y) * std::exp(kA * (z + x)) +
                   2 * std::sin(kA * z + kD * x) * std::cos(kA * y + kD * z) * std::exp(kA * (x + y));
  return -kA * kA / 2 * s * std::exp(-2 * kD * kD * t / mif::current_reynolds());
}

// d/dx of es_pressure; the y and z derivatives follow by rotating arguments.
double es_pressure_dx(double t, double x, double y, double z) {
  const double s1 = std::sin(kA * x + kD * y), c1 = std::cos(kA * z + kD * x);
  const double s2 = std::sin(kA * y + kD * z), c2 = std::cos(kA * x + kD * y);
  const double s3 = std::sin(kA * z + kD * x), c3 = std::cos(kA * y + kD * z);
  const double s = 2 * kA * std::exp(2 * kA * x) + 2 * std::exp(kA * (y + z)) * (kA * c2 * c1 - kD * s1 * s3) +
                   2 * kA * std::exp(kA * (z + x)) * s2 * (c2 - s1) +
                   2 * std::exp(kA * (x + y)) * c3 * (kD * c1 + kA * s3);
  return -kA * kA / 2 * s * std::exp(-2 * kD * kD * t / mif::current_reynolds());
}

bool on_plane(double x, double plane) {
  return x < plane + mif::exact_solution_precision && x > plane - mif::exact_solution_precision;
}

}  // namespace

// ---- Ethier-Steinman (include/Manufactured.h) -------------------------------
double u_exact(double t, double x, double y, double z) { return es_component(t, x, y, z); }
double v_exact(double t, double x, double y, double z) { return es_component(t, y, z, x); }
double w_exact(double t, double x, double y, double z) { return es_component(t, z, x, y); }
double p_exact(double t, double x, double y, double z) { return es_pressure(t, x, y, z); }
double dp_dx_exact(double t, double x, double y, double z) { return es_pressure_dx(t, x, y, z); }
double dp_dy_exact(double t, double x, double y, double z) { return es_pressure_dx(t, y, z, x); }
double dp_dz_exact(double t, double x, double y, double z) { return es_pressure_dx(t, z, x, y); }

// ---- velocity-only field and forcing (include/ManufacturedVelocity.h) ------
double u_exact_v_test(double t, double x, double y, double z) {
  return std::sin(x) * std::cos(y) * std::sin(z) * std::sin(t);
}
double v_exact_v_test(double t, double x, double y, double z) {
  return std::cos(x) * std::sin(y) * std::sin(z) * std::sin(t);
}
double w_exact_v_test(double t, double x, double y, double z) {
  return 2 * std::cos(x) * std::cos(y) * std::cos(z) * std::sin(t);
}
// f = du/dt + (u . grad) u - lap(u) / Re of the field above.
double forcing_x(double t, double x, double y, double z) {
  const double st = std::sin(t), ct = std::cos(t), sx = std::sin(x), cx = std::cos(x), sy = std::sin(y),
               cy = std::cos(y), sz = std::sin(z);
  return sx * cx * st * st * (2 - 2 * sy * sy - sz * sz) + sx * sz * cy * (ct + 3 * st / mif::current_reynolds());
}
double forcing_y(double t, double x, double y, double z) {
  const double st = std::sin(t), ct = std::cos(t), sx = std::sin(x), cx = std::cos(x), sy = std::sin(y),
               cy = std::cos(y), sz = std::sin(z);
  return sy * cy * st * st * (2 - 2 * sx * sx - sz * sz) + sy * sz * cx * (ct + 3 * st / mif::current_reynolds());
}
double forcing_z(double t, double x, double y, double z) {
  const double st = std::sin(t), ct = std::cos(t), sx = std::sin(x), cx = std::cos(x), sy = std::sin(y),
               cy = std::cos(y), sz = std::sin(z), cz = std::cos(z);
  return sz * cz * st * st * (2 * sx * sx + 2 * sy * sy - 4) +
         cx * cy * cz * (2 * ct + 6 * st / mif::current_reynolds());
}

// ---- Poisson test field (include/ManufacturedPressure.h) --------------------
// grad p equals the velocity field.
double u_exact_p_test(double t, double x, double y, double z) { return -std::sin(x) * std::cos(y) * std::cos(z) * t; }
double v_exact_p_test(double t, double x, double y, double z) { return -std::cos(x) * std::sin(y) * std::cos(z) * t; }
double w_exact_p_test(double t, double x, double y, double z) { return -std::cos(x) * std::cos(y) * std::sin(z) * t; }
double p_exact_p_test(double t, double x, double y, double z) { return std::cos(x) * std::cos(y) * std::cos(z) * t; }
double dp_dx_exact_p_test(double t, double x, double y, double z) { return u_exact_p_test(t, x, y, z); }
double dp_dy_exact_p_test(double t, double x, double y, double z) { return v_exact_p_test(t, x, y, z); }
double dp_dz_exact_p_test(double t, double x, double y, double z) { return w_exact_p_test(t, x, y, z); }

namespace mif {

// ---- lid-driven cavity data (include/TestCaseBoundaries.h) ------------------
Real exact_u_t1(Real, Real, Real, Real) { return 0.0; }
Real exact_v_t1(Real, Real x, Real, Real) { return on_plane(x, 1.0) ? 1.0 : 0.0; }
Real exact_w_t1(Real, Real, Real, Real) { return 0.0; }
Real exact_p_initial_t1(Real, Real, Real) { return 0.0; }
Real exact_u_t2(Real, Real, Real, Real) { return 0.0; }
Real exact_v_t2(Real, Real x, Real, Real) { return on_plane(x, -0.5) ? 1.0 : 0.0; }
Real exact_w_t2(Real, Real, Real, Real) { return 0.0; }
Real exact_p_initial_t2(Real, Real, Real) { return 0.0; }

double current_reynolds() { return &::Reynolds != nullptr ? ::Reynolds : 1.0; }

namespace {

using TimeFn = Real (*)(Real, Real, Real, Real);

struct NamedTriple {
  TimeFn u, v, w;
  int kind;
};

const NamedTriple kNamedTriples[] = {
    {u_exact, v_exact, w_exact, MIF_FN_ES_VELOCITY},
    {dp_dx_exact, dp_dy_exact, dp_dz_exact, MIF_FN_ES_GRAD_P},
    {u_exact_v_test, v_exact_v_test, w_exact_v_test, MIF_FN_VT_VELOCITY},
    {forcing_x, forcing_y, forcing_z, MIF_FN_VT_FORCING},
    {u_exact_p_test, v_exact_p_test, w_exact_p_test, MIF_FN_PT_VELOCITY},
    {dp_dx_exact_p_test, dp_dy_exact_p_test, dp_dz_exact_p_test, MIF_FN_PT_GRAD_P},
    {exact_u_t1, exact_v_t1, exact_w_t1, MIF_FN_LID_1},
    {exact_u_t2, exact_v_t2, exact_w_t2, MIF_FN_LID_2},
};

TimeFn pointer_of(const TimeFunction &f) {
  const TimeFn *p = f.target<TimeFn>();
  return p ? *p : nullptr;
}

mif_function make_descriptor(int kind) {
  mif_function d{};
  d.kind = kind;
  return d;
}

// Evaluate a descriptor on the host (used to give descriptor-built functions
// host callbacks as well).
Real eval_descriptor(const mif_function &d, int comp, Real t, Real x, Real y, Real z);

Real eval_kind(int kind, const mif_function &d, int comp, Real t, Real x, Real y, Real z) {
  switch (kind) {
    case MIF_FN_ZERO:
      return 0.0;
    case MIF_FN_ES_VELOCITY:
      return comp == 0 ? u_exact(t, x, y, z) : (comp == 1 ? v_exact(t, x, y, z) : w_exact(t, x, y, z));
    case MIF_FN_ES_PRESSURE:
      return p_exact(t, x, y, z);
    case MIF_FN_ES_GRAD_P:
      return comp == 0 ? dp_dx_exact(t, x, y, z) : (comp == 1 ? dp_dy_exact(t, x, y, z) : dp_dz_exact(t, x, y, z));
    case MIF_FN_VT_VELOCITY:
      return comp == 0 ? u_exact_v_test(t, x, y, z)
                       : (comp == 1 ? v_exact_v_test(t, x, y, z) : w_exact_v_test(t, x, y, z));
    case MIF_FN_VT_FORCING:
      return comp == 0 ? forcing_x(t, x, y, z) : (comp == 1 ? forcing_y(t, x, y, z) : forcing_z(t, x, y, z));
    case MIF_FN_PT_VELOCITY:
    case MIF_FN_PT_GRAD_P:
      return comp == 0 ? u_exact_p_test(t, x, y, z)
                       : (comp == 1 ? v_exact_p_test(t, x, y, z) : w_exact_p_test(t, x, y, z));
    case MIF_FN_PT_PRESSURE:
      return p_exact_p_test(t, x, y, z);
    case MIF_FN_LID_1:
      return comp == 1 ? exact_v_t1(t, x, y, z) : 0.0;
    case MIF_FN_LID_2:
      return comp == 1 ? exact_v_t2(t, x, y, z) : 0.0;
    case MIF_FN_CONSTANT:
      return d.c[comp];
    default:
      throw std::invalid_argument("unknown function descriptor kind");
  }
}

Real eval_descriptor(const mif_function &d, int comp, Real t, Real x, Real y, Real z) {
  if (d.kind == MIF_FN_TIME_DIFFERENCE)
    return eval_kind(d.base, d, comp, d.t1, x, y, z) - eval_kind(d.base, d, comp, d.t0, x, y, z);
  return eval_kind(d.kind, d, comp, t, x, y, z);
}

TimeFunction time_callback(const mif_function &d, int comp) {
  return [d, comp](Real t, Real x, Real y, Real z) { return eval_descriptor(d, comp, t, x, y, z); };
}

SpaceFunction space_callback(const mif_function &d, int comp, Real time) {
  return [d, comp, time](Real x, Real y, Real z) { return eval_descriptor(d, comp, time, x, y, z); };
}

SpaceFunction at_time(const TimeFunction &f, Real time) {
  return [f, time](Real x, Real y, Real z) { return f(time, x, y, z); };
}

}  // namespace

bool describe_scalar(const TimeFunction &f, mif_function *out) {
  const TimeFn p = pointer_of(f);
  if (p == static_cast<TimeFn>(p_exact)) {
    *out = make_descriptor(MIF_FN_ES_PRESSURE);
  } else if (p == static_cast<TimeFn>(p_exact_p_test)) {
    *out = make_descriptor(MIF_FN_PT_PRESSURE);
  } else {
    return false;
  }
  out->Re = current_reynolds();
  return true;
}

// ---- VectorFunction ----------------------------------------------------------
VectorFunction::VectorFunction(const SpaceFunction f_u, const SpaceFunction f_v, const SpaceFunction f_w)
    : f_u(f_u), f_v(f_v), f_w(f_w), components{&this->f_u, &this->f_v, &this->f_w} {}

VectorFunction::VectorFunction(const mif_function &descriptor, Real time)
    : f_u(space_callback(descriptor, 0, time)),
      f_v(space_callback(descriptor, 1, time)),
      f_w(space_callback(descriptor, 2, time)),
      components{&this->f_u, &this->f_v, &this->f_w},
      has_desc_(true),
      desc_(descriptor),
      time_(time) {}

VectorFunction::VectorFunction(const VectorFunction &o)
    : f_u(o.f_u),
      f_v(o.f_v),
      f_w(o.f_w),
      components{&this->f_u, &this->f_v, &this->f_w},
      has_desc_(o.has_desc_),
      desc_(o.desc_),
      time_(o.time_) {}

VectorFunction VectorFunction::operator+(const VectorFunction &other) const {
  const SpaceFunction a_u = f_u, a_v = f_v, a_w = f_w, b_u = other.f_u, b_v = other.f_v, b_w = other.f_w;
  return VectorFunction([a_u, b_u](Real x, Real y, Real z) { return a_u(x, y, z) + b_u(x, y, z); },
                        [a_v, b_v](Real x, Real y, Real z) { return a_v(x, y, z) + b_v(x, y, z); },
                        [a_w, b_w](Real x, Real y, Real z) { return a_w(x, y, z) + b_w(x, y, z); });
}

VectorFunction VectorFunction::operator*(const Real scalar) const {
  const SpaceFunction a_u = f_u, a_v = f_v, a_w = f_w;
  return VectorFunction([a_u, scalar](Real x, Real y, Real z) { return a_u(x, y, z) * scalar; },
                        [a_v, scalar](Real x, Real y, Real z) { return a_v(x, y, z) * scalar; },
                        [a_w, scalar](Real x, Real y, Real z) { return a_w(x, y, z) * scalar; });
}

mif_function VectorFunction::descriptor() const {
  mif_function d = desc_;
  d.Re = current_reynolds();
  return d;
}

// ---- TimeVectorFunction ------------------------------------------------------
TimeVectorFunction::TimeVectorFunction(const TimeFunction f_u, const TimeFunction f_v, const TimeFunction f_w)
    : f_u(f_u), f_v(f_v), f_w(f_w), components{&this->f_u, &this->f_v, &this->f_w} {
  const TimeFn pu = pointer_of(f_u), pv = pointer_of(f_v), pw = pointer_of(f_w);
  if (!pu || !pv || !pw) return;
  for (const NamedTriple &n : kNamedTriples) {
    if (pu == n.u && pv == n.v && pw == n.w) {
      has_desc_ = true;
      desc_ = make_descriptor(n.kind);
      return;
    }
  }
}

TimeVectorFunction::TimeVectorFunction(const mif_function &descriptor)
    : f_u(time_callback(descriptor, 0)),
      f_v(time_callback(descriptor, 1)),
      f_w(time_callback(descriptor, 2)),
      components{&this->f_u, &this->f_v, &this->f_w},
      has_desc_(true),
      desc_(descriptor) {}

TimeVectorFunction::TimeVectorFunction(const TimeVectorFunction &o)
    : f_u(o.f_u),
      f_v(o.f_v),
      f_w(o.f_w),
      components{&this->f_u, &this->f_v, &this->f_w},
      has_desc_(o.has_desc_),
      desc_(o.desc_) {}

mif_function TimeVectorFunction::descriptor() const {
  mif_function d = desc_;
  d.Re = current_reynolds();
  return d;
}

VectorFunction TimeVectorFunction::set_time(Real time) const {
  if (has_desc_) return VectorFunction(desc_, time);
  return VectorFunction(at_time(f_u, time), at_time(f_v, time), at_time(f_w, time));
}

VectorFunction TimeVectorFunction::get_difference_over_time(Real time_1, Real time_2) const {
  if (has_desc_ && desc_.kind != MIF_FN_TIME_DIFFERENCE) {
    mif_function d{};
    d.kind = MIF_FN_TIME_DIFFERENCE;
    d.base = desc_.kind;
    for (int c = 0; c < 3; c++) d.c[c] = desc_.c[c];
    d.t0 = time_1;
    d.t1 = time_2;
    return VectorFunction(d, 0.0);
  }
  const TimeFunction a_u = f_u, a_v = f_v, a_w = f_w;
  return VectorFunction(
      [a_u, time_1, time_2](Real x, Real y, Real z) { return a_u(time_2, x, y, z) - a_u(time_1, x, y, z); },
      [a_v, time_1, time_2](Real x, Real y, Real z) { return a_v(time_2, x, y, z) - a_v(time_1, x, y, z); },
      [a_w, time_1, time_2](Real x, Real y, Real z) { return a_w(time_2, x, y, z) - a_w(time_1, x, y, z); });
}

}  // namespace mif
