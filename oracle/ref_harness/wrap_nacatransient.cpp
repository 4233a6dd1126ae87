// Reference-side plugin unit: compiles examples/nacaTransient/main.cpp from the
// reference tree unchanged, with every plugin symbol it defines renamed under
// the "nt_" prefix, so that several example plugins can be linked into one
// harness binary (driver.cpp dispatches between them at run time).
// The fvhyper headers are included first so that their declarations keep the
// library's own names; only what the example defines is renamed.
#include <fvhyper/mesh.h>
#include <fvhyper/explicit.h>
#include <fvhyper/parallel.h>
#include <fvhyper/post.h>

#define main nt_main
#define vars nt_vars
#define var_names nt_var_names
#define do_calc_gradients nt_do_calc_gradients
#define do_calc_limiters nt_do_calc_limiters
#define linear_interpolate nt_linear_interpolate
#define diffusive_gradients nt_diffusive_gradients
#define global_dt nt_global_dt
#define smooth_residuals nt_smooth_residuals
#define consts nt_consts
#define generate_initial_solution nt_generate_initial_solution
#define limiter_func nt_limiter_func
#define calc_flux nt_calc_flux
#define calc_dt nt_calc_dt
#define calc_p nt_calc_p
#define calc_grad_p nt_calc_grad_p
#define calc_grad_t nt_calc_grad_t
#define calc_grad_u nt_calc_grad_u
#define calc_grad_v nt_calc_grad_v
#define boundaries nt_boundaries
#define post nt_post

#include "examples/nacaTransient/main.cpp"
