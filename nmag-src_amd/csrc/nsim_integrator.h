// Device-resident time integrator (BDF with Newton-GMRES); filled in below the
// rhs path.
#pragma once

#include <hip/hip_runtime.h>

#include <string>

struct nsim_handle;

namespace nsim {

struct Integrator {
  virtual ~Integrator() = default;
};

// Entry points of the rhs path used by the integrator.
int device_rhs(nsim_handle* h, const double* m, double* dmdt);
int device_jtimes(nsim_handle* h, const double* m, const double* v, double* Jv);
hipStream_t handle_stream(nsim_handle* h);
int handle_n(nsim_handle* h);
const double* handle_volumes_inv(nsim_handle* h);
double* handle_m(nsim_handle* h);
int handle_error_flags(nsim_handle* h, unsigned* out);
void set_error(const std::string& msg);

}  // namespace nsim
