// probe.cpp -- map the user's fvhyper plugin functions onto device kinds.
//
// fvhyper plugins are host function pointers (examples/*/main.cpp), which a
// GPU cannot call.  The library implements every plugin kind the reference's
// Euler examples use (fvh_physics.h).  Here each user function is evaluated
// on the host on a deterministic battery of inputs and compared bit for bit
// with the candidate device kinds (the same fvh_physics.h code, compiled for
// the host); the time step is compared cell by cell on the real mesh and
// initial state through the library itself.  The first exact match is used;
// no match throws.
#include <fvhyper/explicit.h>
#include <fvhyper/mesh.h>

#include <cmath>
#include <cstdio>
#include <cstdlib>
#include <cstring>
#include <set>
#include <stdexcept>
#include <string>

#include "../../csrc/fvh_physics.h"
#include "dropin_internal.h"

namespace fvhyper {
namespace dropin {
namespace {

struct Rng {
    uint64_t s;
    explicit Rng(uint64_t seed) : s(seed) {}
    double u01() {
        uint64_t z = (s += 0x9E3779B97F4A7C15ull);
        z = (z ^ (z >> 30)) * 0xBF58476D1CE4E5B9ull;
        z = (z ^ (z >> 27)) * 0x94D049BB133111EBull;
        z ^= z >> 31;
        return (double)(z >> 11) * 0x1.0p-53;
    }
    double in(double a, double b) { return a + (b - a) * u01(); }
};

bool same(const double* a, const double* b, int n) { return std::memcmp(a, b, sizeof(double) * n) == 0; }

// x and its decimal roundings to 1..17 significant digits
std::vector<double> nice(double x) {
    std::set<double> s;
    s.insert(x);
    char buf[64];
    for (int d = 1; d <= 17; ++d) {
        std::snprintf(buf, sizeof buf, "%.*g", d, x);
        s.insert(std::strtod(buf, nullptr));
    }
    return std::vector<double>(s.begin(), s.end());
}

void conserved(double* q, double rho, double u, double v, double p) {
    q[0] = rho;
    q[1] = rho * u;
    q[2] = rho * v;
    q[3] = p / 0.4 + 0.5 * rho * (u * u + v * v);
}

// ---- flux: Roe spelling and gamma ------------------------------------------
struct Visc { double pr, cp, r; };
struct FluxKind {
    int kind;                 // FVH_FLUX_*
    double gamma;
    double mu;                // viscous kind only
    std::vector<Visc> visc;   // (pr, cp, r) that reproduce the viscous terms bit for bit
};

// Laminar terms: the user's flux with a gradient, against the device kind.
// A shear layer gives mu; heat conduction gives cp/(pr r), and every
// (pr, cp, r) from a list of usual values and decimal roundings that
// reproduces the flux on the battery is kept (pr is settled by calc_dt).
void probe_viscous(FluxKind& fk, Rng& r, int NS, const std::vector<double>& qi, const std::vector<double>& qj,
                   const std::vector<double>& nn) {
    const double g = fk.gamma, zero[4] = {0, 0, 0, 0};
    std::vector<double> gx(4 * NS), gy(4 * NS), fu(4 * NS);
    for (int k = 0; k < 4 * NS; ++k) { gx[k] = r.in(-2, 2); gy[k] = r.in(-2, 2); }
    for (int k = 0; k < NS; ++k) calc_flux(&fu[4 * k], &qi[4 * k], &qj[4 * k], &gx[4 * k], &gy[4 * k], &nn[2 * k]);
    // shear u = a y at constant rho, p: f1 changes by -mu a through n = (0, 1)
    double q[4], f0[4], f1[4];
    conserved(q, 1.3, 0.4, 0.0, 0.9);
    q[3] = 0.9 / (g - 1) + 0.5 * (q[1] * q[1]) / q[0];
    const double a = 1.0, n[2] = {0.0, 1.0};
    const double gys[4] = {0.0, q[0] * a, 0.0, q[1] * a};
    calc_flux(f0, q, q, zero, zero, n);
    calc_flux(f1, q, q, zero, gys, n);
    const double mu_est = -(f1[1] - f0[1]) / a;
    // heat conduction at rest: f3 changes by -(cp mu / pr) (dp/dx) / (rho r)
    conserved(q, 1.3, 0.0, 0.0, 0.9);
    q[3] = 0.9 / (g - 1);
    const double b = 0.7, nx[2] = {1.0, 0.0};
    const double gxs[4] = {0.0, 0.0, 0.0, b / (g - 1)};
    calc_flux(f0, q, q, zero, zero, nx);
    calc_flux(f1, q, q, gxs, zero, nx);
    const double heat = -(f1[3] - f0[3]) * q[0] / b;   // cp mu / (pr r)
    auto matches = [&](double mu, double pr, double cp, double rg) {
        for (int k = 0; k < NS; ++k) {
            double fd[4];
            fvh::roe_flux<double, false>(g, fd, &qi[4 * k], &qj[4 * k], nn[2 * k], nn[2 * k + 1]);
            fvh::viscous_terms<double>(g, mu, pr, cp, rg, fd, &qi[4 * k], &qj[4 * k], &gx[4 * k], &gy[4 * k],
                                       nn[2 * k], nn[2 * k + 1]);
            if (!same(fd, &fu[4 * k], 4)) return false;
        }
        return true;
    };
    for (double mu : nice(mu_est)) {
        if (!(mu > 0)) continue;
        for (double pr : {0.72, 0.71, 0.7, 0.73, 0.75, 0.9, 1.0, 0.5})
            for (double cp : {1.0, 1004.5, 1005.0, 1004.0, 1006.0, 1004.68506, 718.0, 717.5})
                for (double rg : nice(cp * mu / pr / heat))
                    if (fk.visc.size() < 64 && matches(mu, pr, cp, rg)) {
                        fk.mu = mu;
                        fk.visc.push_back({pr, cp, rg});
                    }
        if (!fk.visc.empty()) return;
    }
    throw std::runtime_error("calc_flux: depends on the gradients, but no laminar viscous kind reproduces it (the "
                             "library implements the viscous terms of the flatPlate/nacaTransient examples)");
}

FluxKind probe_flux() {
    Rng r(11);
    const int NS = 256;
    std::vector<double> qi(4 * NS), qj(4 * NS), nn(2 * NS), fu(4 * NS);
    const double zero[4] = {0, 0, 0, 0};
    for (int k = 0; k < NS; ++k) {
        conserved(&qi[4 * k], r.in(0.2, 3), r.in(-3, 3), r.in(-3, 3), r.in(0.2, 3));
        if (k % 8 == 0) std::memcpy(&qj[4 * k], &qi[4 * k], 32);
        else conserved(&qj[4 * k], r.in(0.2, 3), r.in(-3, 3), r.in(-3, 3), r.in(0.2, 3));
        const double t = r.in(0, 6.283185307179586);
        nn[2 * k] = std::cos(t);
        nn[2 * k + 1] = std::sin(t);
        calc_flux(&fu[4 * k], &qi[4 * k], &qj[4 * k], zero, zero, &nn[2 * k]);
    }
    // gamma from the consistent flux: f1 = rho u^2 + p with qi = qj, n = (1, 0)
    double q[4] = {1.0, 0.3, 0.1, 3.0}, f[4], n[2] = {1.0, 0.0};
    calc_flux(f, q, q, zero, zero, n);
    const double p = f[1] - q[1] * (q[1] / q[0]);
    const double g_est = 1.0 + p / (q[3] - 0.5 / q[0] * (q[1] * q[1] + q[2] * q[2]));
    std::vector<double> gs = nice(g_est);
    for (double g : {1.4, 5.0 / 3.0, 1.3, 1.2, 1.1, 1.67}) gs.push_back(g);
    for (double g : gs) {
        for (int joined = 0; joined < 2; ++joined) {
            bool ok = true;
            for (int k = 0; k < NS && ok; ++k) {
                double fd[4];
                if (joined) fvh::roe_flux<double, true>(g, fd, &qi[4 * k], &qj[4 * k], nn[2 * k], nn[2 * k + 1]);
                else fvh::roe_flux<double, false>(g, fd, &qi[4 * k], &qj[4 * k], nn[2 * k], nn[2 * k + 1]);
                ok = same(fd, &fu[4 * k], 4);
            }
            if (!ok) continue;
            FluxKind fk{joined ? FVH_FLUX_ROE_JOINED : FVH_FLUX_ROE_SPLIT, g, 0.0, {}};
            // gradient dependence: the viscous kind (split Roe + laminar terms)
            bool grad_free = true;
            for (int k = 0; k < NS && grad_free; ++k) {
                double gxk[4], gyk[4], fg[4];
                for (int j = 0; j < 4; ++j) { gxk[j] = r.in(-2, 2); gyk[j] = r.in(-2, 2); }
                calc_flux(fg, &qi[4 * k], &qj[4 * k], gxk, gyk, &nn[2 * k]);
                grad_free = same(fg, &fu[4 * k], 4);
            }
            if (grad_free) return fk;
            if (joined) continue;
            fk.kind = FVH_FLUX_ROE_VISCOUS;
            probe_viscous(fk, r, NS, qi, qj, nn);
            return fk;
        }
    }
    throw std::runtime_error("calc_flux: no device flux kind reproduces it (the library implements the Roe flux "
                             "with the entropy cure in the shockTube/naca and forwardStep spellings, and with the "
                             "laminar terms of the flatPlate/nacaTransient examples)");
}

// ---- limiter function --------------------------------------------------------
struct PhiKind { int kind; double yt; };

PhiKind probe_phi() {
    std::vector<double> ys;
    for (int k = 0; k <= 300; ++k) ys.push_back(-3.0 + 15.0 * k / 300.0);
    for (double y : {0.0, 1.0, 2.0, 1e-3, 1e3, 1e6, -1e3, 0.5, 1.999, 2.001}) ys.push_back(y);
    std::vector<double> fu(ys.size());
    for (size_t k = 0; k < ys.size(); ++k) fu[k] = limiter_func(ys[k]);
    auto matches = [&](int kind, double yt) {
        for (size_t k = 0; k < ys.size(); ++k) {
            double v;
            switch (kind) {
                case FVH_PHI_ZERO: v = fvh::phi<double, fvh::PHI_ZERO>(ys[k], yt); break;
                case FVH_PHI_MICHALAK: v = fvh::phi<double, fvh::PHI_MICHALAK>(ys[k], yt); break;
                case FVH_PHI_VENKATAKRISHNAN: v = fvh::phi<double, fvh::PHI_VENKAT>(ys[k], yt); break;
                default: v = fvh::phi<double, fvh::PHI_MINMOD>(ys[k], yt); break;
            }
            if (!same(&v, &fu[k], 1)) return false;
        }
        return true;
    };
    if (matches(FVH_PHI_ZERO, 2.0)) return {FVH_PHI_ZERO, 2.0};
    if (matches(FVH_PHI_MINMOD, 2.0)) return {FVH_PHI_MINMOD, 2.0};
    if (matches(FVH_PHI_VENKATAKRISHNAN, 2.0)) return {FVH_PHI_VENKATAKRISHNAN, 2.0};
    // Michalak: phi == 1 from y_t on; locate the threshold by bisection
    double lo = 0.0, hi = 1e3;
    if (limiter_func(hi) == 1.0 && limiter_func(lo) != 1.0) {
        for (int it = 0; it < 200; ++it) {
            const double mid = 0.5 * (lo + hi);
            if (limiter_func(mid) == 1.0) hi = mid; else lo = mid;
        }
        for (double yt : nice(hi))
            if (matches(FVH_PHI_MICHALAK, yt)) return {FVH_PHI_MICHALAK, yt};
    }
    throw std::runtime_error("limiter_func: no device limiter kind reproduces it (zero, Michalak, "
                             "Venkatakrishnan, minmod)");
}

// ---- boundary functions --------------------------------------------------------
fvh_bc_spec probe_bc(const std::string& name, void (*fn)(double*, double*, double*), double gamma) {
    Rng r(29);
    const int NS = 192;
    std::vector<double> qs(4 * NS), ns(2 * NS), bu(4 * NS);
    for (int k = 0; k < NS; ++k) {
        const double t = r.in(0, 6.283185307179586);
        ns[2 * k] = std::cos(t);
        ns[2 * k + 1] = std::sin(t);
        const double rho = r.in(0.3, 2.0), p = r.in(0.3, 2.0);
        const double c = std::sqrt(gamma * p / rho);
        const double mach = (k % 2) ? r.in(0.1, 0.95) : r.in(1.1, 3.0);
        const double a = r.in(0, 6.283185307179586);
        double q[4];
        q[0] = rho;
        q[1] = rho * mach * c * std::cos(a);
        q[2] = rho * mach * c * std::sin(a);
        q[3] = p / (gamma - 1) + 0.5 * (q[1] * q[1] + q[2] * q[2]) / rho;
        std::memcpy(&qs[4 * k], q, 32);
        double b[4];
        fn(b, q, &ns[2 * k]);
        std::memcpy(&bu[4 * k], b, 32);
    }
    auto matches = [&](const fvh::BcParam<double>& P) {
        for (int k = 0; k < NS; ++k) {
            double b[4];
            fvh::boundary(P, gamma, b, &qs[4 * k], ns[2 * k], ns[2 * k + 1]);
            if (!same(b, &bu[4 * k], 4)) return false;
        }
        return true;
    };
    fvh_bc_spec spec;
    std::memset(&spec, 0, sizeof spec);
    std::snprintf(spec.name, sizeof spec.name, "%s", name.c_str());
    fvh::BcParam<double> P{fvh::BC_COPY, 0, 0, 0, 0};
    if (matches(P)) { spec.kind = FVH_BC_COPY; return spec; }
    P.kind = fvh::BC_MIRROR;
    if (matches(P)) { spec.kind = FVH_BC_MIRROR; return spec; }
    P.kind = fvh::BC_NOSLIP;
    if (matches(P)) { spec.kind = FVH_BC_NOSLIP; return spec; }
    // characteristic kinds: a supersonic state entering through n returns the freestream
    double q[4], n[2] = {1.0, 0.0}, bv[4];
    conserved(q, 1.0, -2.5 * std::sqrt(gamma), 0.0, 1.0);
    q[3] = 1.0 / (gamma - 1) + 0.5 * (q[1] * q[1] + q[2] * q[2]) / q[0];
    fn(bv, q, n);
    const double rho = bv[0];
    std::vector<double> us, vs, ps;
    for (double u : nice(bv[1] / rho)) if (rho * u == bv[1]) us.push_back(u);
    for (double v : nice(bv[2] / rho)) if (rho * v == bv[2]) vs.push_back(v);
    const double p_est = (gamma - 1) * (bv[3] - 0.5 / rho * (bv[1] * bv[1] + bv[2] * bv[2]));
    for (double p : nice(p_est))
        if (p / (gamma - 1) + 0.5 / rho * (bv[1] * bv[1] + bv[2] * bv[2]) == bv[3]) ps.push_back(p);
    for (int kind : {fvh::BC_INLET_OUTLET, fvh::BC_FARFIELD, fvh::BC_FARFIELD_V})
        for (double u : us)
            for (double v : vs)
                for (double p : ps) {
                    P = fvh::BcParam<double>{kind, rho, u, v, p};
                    if (matches(P)) {
                        spec.kind = kind;
                        spec.rho = rho; spec.u = u; spec.v = v; spec.p = p;
                        return spec;
                    }
                }
    throw std::runtime_error("boundaries::bounds[\"" + name + "\"]: no device boundary kind reproduces it "
                             "(copy, mirror wall, no-slip wall, characteristic inlet/outlet, farfield in the naca "
                             "and flatPlate spellings)");
}

// ---- time step: compared on the mesh itself through the library --------------
std::vector<double> device_dt(mesh& m, fvh_config c, const std::vector<double>& q) {
    c.flags &= ~FVH_FLAG_GLOBAL_DT;
    fvh_solver* s = nullptr;
    if (fvh_solver_create(m.device_mesh(), &c, &s) != FVH_OK) throw std::runtime_error(fvh_last_error());
    std::vector<double> dt(m.cellsAreas.size());
    int32_t rc = fvh_solver_set_state(s, q.data());
    if (rc == FVH_OK) rc = fvh_solver_phases(s, 0.0, dt.data(), nullptr, nullptr, nullptr, nullptr, nullptr);
    fvh_solver_free(s);
    if (rc != FVH_OK) throw std::runtime_error(fvh_last_error());
    return dt;
}

void probe_dt(mesh& m, fvh_config& c, const std::vector<double>& q_initial, const std::vector<Visc>& visc) {
    // the state the first calc_dt sees: the initial solution after the
    // initial boundary update (explicit.cpp:712, 729)
    std::vector<double> q(q_initial);
    {
        fvh_solver* s = nullptr;
        if (fvh_solver_create(m.device_mesh(), &c, &s) != FVH_OK) throw std::runtime_error(fvh_last_error());
        int32_t rc = fvh_solver_set_state(s, q.data());
        if (rc == FVH_OK) rc = fvh_solver_run_begin(s);
        if (rc == FVH_OK) rc = fvh_solver_get_state(s, q.data());
        fvh_solver_free(s);
        if (rc != FVH_OK) throw std::runtime_error(fvh_last_error());
    }
    const size_t N = m.cellsAreas.size();
    std::vector<double> dtv(vars * N, 0.0);
    calc_dt(dtv, q, m);
    std::vector<double> du(N);
    bool all_equal = true;
    for (size_t i = 0; i < N; ++i) {
        du[i] = dtv[vars * i];
        for (int k = 1; k < vars; ++k)
            if (!same(&dtv[vars * i + k], &du[i], 1))
                throw std::runtime_error("calc_dt: the variables of a cell get different time steps");
        if (!same(&du[i], &du[0], 1)) all_equal = false;
    }
    if (all_equal) {
        c.dt_kind = FVH_DT_FIXED;
        c.dt_fixed = du[0];
        if (same(device_dt(m, c, q).data(), du.data(), (int)N)) return;
    }
    // the viscous kinds once per (pr, cp, r) the flux left open: pr enters
    // the time step through mu/pr
    std::vector<std::pair<int, Visc>> tries = {{FVH_DT_SUM_CFL, Visc{c.pr, c.cp, c.r_gas}},
                                               {FVH_DT_EDGE_MIN, Visc{c.pr, c.cp, c.r_gas}}};
    for (const Visc& v : visc) {
        tries.push_back({FVH_DT_EDGE_MIN_VISCOUS, v});
        tries.push_back({FVH_DT_SUM_VISCOUS, v});
    }
    for (const auto& t : tries) {
        c.dt_kind = t.first;
        c.pr = t.second.pr;
        c.cp = t.second.cp;
        c.r_gas = t.second.r;
        c.cfl = 1.0;
        const std::vector<double> d1 = device_dt(m, c, q);
        // the time step is linear in the CFL number away from the edge-min cap of 1
        size_t i = 0;
        while (i + 1 < N && !(d1[i] < 0.5 && d1[i] > 0)) ++i;
        for (double cfl : nice(du[i] / d1[i])) {
            c.cfl = cfl;
            if (same(device_dt(m, c, q).data(), du.data(), (int)N)) return;
        }
    }
    throw std::runtime_error("calc_dt: no device time-step kind reproduces it on this mesh (fixed, summed CFL, "
                             "edge-minimum local, and their viscous forms)");
}

}  // namespace

namespace {

fvh_config config_without_dt(mesh& m, FluxKind& fk) {
    if (vars != 4)
        throw std::runtime_error("vars = " + std::to_string(vars) +
                                 ": the device path implements the 4-variable Euler plugins");
    fvh_config c;
    fvh_config_default(&c);
    c.flags = 0;
    if (solver::do_calc_gradients) c.flags |= FVH_FLAG_GRADIENTS;
    if (solver::do_calc_limiters) c.flags |= FVH_FLAG_LIMITERS;
    if (solver::linear_interpolate) c.flags |= FVH_FLAG_LINEAR;
    if (solver::global_dt) c.flags |= FVH_FLAG_GLOBAL_DT;
    if (solver::smooth_residuals) c.flags |= FVH_FLAG_SMOOTH;
    if (solver::diffusive_gradients) c.flags |= FVH_FLAG_DIFFUSIVE;
    c.K = solver::limiter_k_value;
    fk = probe_flux();
    c.flux = fk.kind;
    c.gamma = fk.gamma;
    if (fk.kind == FVH_FLUX_ROE_VISCOUS) {
        c.mu = fk.mu;
        c.pr = fk.visc[0].pr;
        c.cp = fk.visc[0].cp;
        c.r_gas = fk.visc[0].r;
    }
    if (solver::do_calc_limiters) {
        const PhiKind pk = probe_phi();
        c.phi = pk.kind;
        c.phi_yt = pk.yt;
    }
    // one BC slot per physical name the mesh uses
    std::vector<std::string> seen;
    c.n_bc = 0;
    for (size_t b = 0; b < m.boundaryNames.size(); ++b) {
        const std::string& nm = m.boundaryNames[b];
        if (std::find(seen.begin(), seen.end(), nm) != seen.end()) continue;
        if (c.n_bc == FVH_MAX_BC) throw std::runtime_error("more than 16 boundary names");
        seen.push_back(nm);
        c.bc[c.n_bc++] = probe_bc(nm, m.boundaryFuncs[b], c.gamma);
    }
    return c;
}

}  // namespace

fvh_config phase_config(mesh& m) {
    FluxKind fk;
    return config_without_dt(m, fk);
}

fvh_config device_config(mesh& m, const std::vector<double>& q_initial) {
    FluxKind fk;
    fvh_config c = config_without_dt(m, fk);
    probe_dt(m, c, q_initial, fk.visc);
    return c;
}

}  // namespace dropin
}  // namespace fvhyper
