// Reference-side plugin unit: compiles examples/flatPlate/main.cpp from the
// reference tree unchanged, with every plugin symbol it defines renamed under
// the "fp_" prefix, so that several example plugins can be linked into one
// harness binary (driver.cpp dispatches between them at run time).
// The fvhyper headers are included first so that their declarations keep the
// library's own names; only what the example defines is renamed.
#include <fvhyper/mesh.h>
#include <fvhyper/explicit.h>
#include <fvhyper/parallel.h>
#include <fvhyper/post.h>

#define main fp_main
#define vars fp_vars
#define var_names fp_var_names
#define do_calc_gradients fp_do_calc_gradients
#define do_calc_limiters fp_do_calc_limiters
#define linear_interpolate fp_linear_interpolate
#define diffusive_gradients fp_diffusive_gradients
#define global_dt fp_global_dt
#define smooth_residuals fp_smooth_residuals
#define consts fp_consts
#define generate_initial_solution fp_generate_initial_solution
#define limiter_func fp_limiter_func
#define calc_flux fp_calc_flux
#define calc_dt fp_calc_dt
#define calc_p fp_calc_p
#define calc_grad_p fp_calc_grad_p
#define calc_grad_t fp_calc_grad_t
#define calc_grad_u fp_calc_grad_u
#define calc_grad_v fp_calc_grad_v
#define boundaries fp_boundaries
#define post fp_post

#include "examples/flatPlate/main.cpp"
