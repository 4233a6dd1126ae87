// fvhyper/explicit.h -- drop-in for fvhyper's explicit solver API
// (reference: fvhyper/explicit.h).  The plugin symbols are declared exactly
// as the reference declares them; run() and the phase functions execute on
// the GPU through libfvh.  The user's host plugin functions are probed once
// and mapped onto the library's device kinds (Roe flux spelling and gamma,
// limiter function, time-step form and CFL, boundary kinds and freestream);
// a plugin with no device counterpart makes run() throw std::runtime_error
// naming it, never a silently different result.
#pragma once

#include <math.h>

#include <map>
#include <string>
#include <unordered_map>
#include <vector>

#include <fvhyper/mesh.h>
#include <fvhyper/parallel.h>

namespace fvhyper {

extern const int vars;
extern const std::vector<std::string> var_names;

namespace solver {
    extern const bool do_calc_gradients;
    extern const bool do_calc_limiters;
    extern const bool linear_interpolate;
    extern const bool diffusive_gradients;
    extern const bool global_dt;
    extern const bool smooth_residuals;
    extern const double limiter_k_value;
}

double limiter_func(const double& r);

void generate_initial_solution(std::vector<double>& v, const mesh& m);

void calc_flux(double* f, const double* qi, const double* qj, const double* gx, const double* gy,
               const double* n);

void calc_dt(std::vector<double>& dt, const std::vector<double>& q, mesh& m);

// The phases of one stage, each on the GPU through libfvh (reference:
// explicit.h:88-149, bodies explicit.cpp:112-417, 481-564, 625-653).  Every
// call copies its vectors to the device, runs the phase there and copies the
// results back: for programs that compose their own loop; run() keeps the
// state on the device instead.  The plugin kinds come from probing the
// user's plugin functions once per mesh, as run() does.  Output vectors are
// resized to vars values per cell when they are shorter.
//
// The reference's header and its source differ for two of them: explicit.h
// declares calc_gradients with a const mesh& and calc_limiters without qmin
// and qmax, while explicit.cpp defines them with mesh& and with qmin, qmax.
// Both forms are defined here.
void calc_gradients(std::vector<double>& gx, std::vector<double>& gy, const std::vector<double>& q,
                    const mesh& m);
void calc_gradients(std::vector<double>& gx, std::vector<double>& gy, const std::vector<double>& q, mesh& m);
void calc_limiters(std::vector<double>& limiters, const std::vector<double>& q, const std::vector<double>& gx,
                   const std::vector<double>& gy, const mesh& m);
void calc_limiters(std::vector<double>& limiters, std::vector<double>& qmin, std::vector<double>& qmax,
                   const std::vector<double>& q, const std::vector<double>& gx, const std::vector<double>& gy,
                   mesh& m);
void calc_time_derivatives(std::vector<double>& qt, const std::vector<double>& q, const std::vector<double>& gx,
                           const std::vector<double>& gy, const std::vector<double>& limiters, mesh& m);
void update_cells(std::vector<double>& q, std::vector<double>& ql, const std::vector<double>& qt,
                  const std::vector<double>& dt, const double v);
void update_bounds(std::vector<double>& q, std::vector<double>& gx, std::vector<double>& gy,
                   std::vector<double>& limiters, mesh& m);
// one device holds the whole mesh: no halo to exchange (explicit.cpp:420-477)
void update_comms(std::vector<double>& q, mesh& m);
void calc_residuals(double* R, std::vector<double>& qt, mesh& m, mpi_wrapper& pool);
void min_dt(std::vector<double>& dt, mesh& m);
// one rank: the local minimum is the global one (explicit.cpp:567-621)
void validate_dt(std::vector<double>& dt, mpi_wrapper& pool);
void complete_calc_qt(std::vector<double>& qt, std::vector<double>& q, std::vector<double>& gx,
                      std::vector<double>& gy, std::vector<double>& qmin, std::vector<double>& qmax,
                      std::vector<double>& limiters, mesh& m, mpi_wrapper& pool);

struct solverOptions {
    double max_time = 1e10;
    uint max_step = 1e8;
    uint print_interval = 1;
    bool verbose = true;
    double tolerance = 1e-16;
    bool save_time_series = false;
    double time_series_interval = 0.2;
};

// reference: explicit.cpp:657-817
void run(const std::string name, std::vector<double>& q, mpi_wrapper& pool, mesh& m, solverOptions& opt);

}  // namespace fvhyper
