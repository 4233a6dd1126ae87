// dropin_internal.h -- recognising the user's fvhyper plugin as device kinds.
#pragma once

#include <fvhyper/mesh.h>

#include <vector>

#include "../../../include/fvh.h"

namespace fvhyper {
namespace dropin {

// Builds the device configuration for run(): the solver:: flags, and for
// each plugin function (calc_flux, limiter_func, calc_dt, every
// boundaries::bounds entry the mesh uses) the device kind and parameters
// that reproduce it bit for bit on probe inputs.  Throws std::runtime_error
// naming the function when no device kind matches.
fvh_config device_config(mesh& m, const std::vector<double>& q_initial);

// The same without the time step (fixed, unused): the configuration of the
// per-phase functions, which take the time step from the caller.
fvh_config phase_config(mesh& m);

}  // namespace dropin
}  // namespace fvhyper
