// Reference-side plugin unit: compiles examples/shockTube/main.cpp from the
// reference tree unchanged, with every plugin symbol it defines renamed under
// the "st_" prefix, so that several example plugins can be linked into one
// harness binary (driver.cpp dispatches between them at run time).
// The fvhyper headers are included first so that their declarations keep the
// library's own names; only what the example defines is renamed.
#include <fvhyper/mesh.h>
#include <fvhyper/explicit.h>
#include <fvhyper/parallel.h>
#include <fvhyper/post.h>

#define main st_main
#define vars st_vars
#define var_names st_var_names
#define do_calc_gradients st_do_calc_gradients
#define do_calc_limiters st_do_calc_limiters
#define linear_interpolate st_linear_interpolate
#define diffusive_gradients st_diffusive_gradients
#define global_dt st_global_dt
#define smooth_residuals st_smooth_residuals
#define consts st_consts
#define generate_initial_solution st_generate_initial_solution
#define limiter_func st_limiter_func
#define calc_flux st_calc_flux
#define calc_dt st_calc_dt
#define calc_p st_calc_p
#define boundaries st_boundaries
#define post st_post

#include "examples/shockTube/main.cpp"
