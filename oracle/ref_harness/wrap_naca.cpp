// Reference-side plugin unit: compiles examples/naca/main.cpp from the
// reference tree unchanged, with every plugin symbol it defines renamed under
// the "nc_" prefix, so that several example plugins can be linked into one
// harness binary (driver.cpp dispatches between them at run time).
// The fvhyper headers are included first so that their declarations keep the
// library's own names; only what the example defines is renamed.
#include <fvhyper/mesh.h>
#include <fvhyper/explicit.h>
#include <fvhyper/parallel.h>
#include <fvhyper/post.h>

#define main nc_main
#define vars nc_vars
#define var_names nc_var_names
#define do_calc_gradients nc_do_calc_gradients
#define do_calc_limiters nc_do_calc_limiters
#define linear_interpolate nc_linear_interpolate
#define diffusive_gradients nc_diffusive_gradients
#define global_dt nc_global_dt
#define smooth_residuals nc_smooth_residuals
#define consts nc_consts
#define generate_initial_solution nc_generate_initial_solution
#define limiter_func nc_limiter_func
#define calc_flux nc_calc_flux
#define calc_dt nc_calc_dt
#define calc_p nc_calc_p
#define boundaries nc_boundaries
#define post nc_post

#include "examples/naca/main.cpp"
