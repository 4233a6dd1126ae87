// Reference-side plugin unit: compiles examples/forwardStep/main.cpp from the
// reference tree unchanged, with every plugin symbol it defines renamed under
// the "fs_" prefix, so that several example plugins can be linked into one
// harness binary (driver.cpp dispatches between them at run time).
// The fvhyper headers are included first so that their declarations keep the
// library's own names; only what the example defines is renamed.
#include <fvhyper/mesh.h>
#include <fvhyper/explicit.h>
#include <fvhyper/parallel.h>
#include <fvhyper/post.h>

#define main fs_main
#define vars fs_vars
#define var_names fs_var_names
#define do_calc_gradients fs_do_calc_gradients
#define do_calc_limiters fs_do_calc_limiters
#define linear_interpolate fs_linear_interpolate
#define diffusive_gradients fs_diffusive_gradients
#define global_dt fs_global_dt
#define smooth_residuals fs_smooth_residuals
#define consts fs_consts
#define generate_initial_solution fs_generate_initial_solution
#define limiter_func fs_limiter_func
#define calc_flux fs_calc_flux
#define calc_dt fs_calc_dt
#define calc_p fs_calc_p
#define boundaries fs_boundaries
#define post fs_post

#include "examples/forwardStep/main.cpp"
