// Reference-side plugin unit: compiles examples/burgers/main.cpp from the
// reference tree unchanged, with every plugin symbol it defines renamed under
// the "bg_" prefix, so that several example plugins can be linked into one
// harness binary (driver.cpp dispatches between them at run time).
// The fvhyper headers are included first so that their declarations keep the
// library's own names; only what the example defines is renamed.
#include <fvhyper/mesh.h>
#include <fvhyper/explicit.h>
#include <fvhyper/parallel.h>
#include <fvhyper/post.h>

#define main bg_main
#define vars bg_vars
#define var_names bg_var_names
#define do_calc_gradients bg_do_calc_gradients
#define do_calc_limiters bg_do_calc_limiters
#define linear_interpolate bg_linear_interpolate
#define diffusive_gradients bg_diffusive_gradients
#define global_dt bg_global_dt
#define smooth_residuals bg_smooth_residuals
#define consts bg_consts
#define generate_initial_solution bg_generate_initial_solution
#define limiter_func bg_limiter_func
#define calc_flux bg_calc_flux
#define calc_dt bg_calc_dt
#define calc_p bg_calc_p
#define boundaries bg_boundaries
#define post bg_post

#include "examples/burgers/main.cpp"
