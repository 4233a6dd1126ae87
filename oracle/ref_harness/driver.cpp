// Reference harness: links the unmodified fvhyper library sources
// (fvhyper/src/*.cpp) with the renamed example plugins (wrap_*.cpp) and
// defines the plugin symbols the library expects (fvhyper/explicit.h:31-85,
// fvhyper/mesh.h:36-39, fvhyper/post.h:24-30) as run-time dispatchers between
// those examples.  It is test infrastructure: it produces the golden vectors
// and the CPU timing baseline, and is never part of the product.
//
// Usage:  fvh_ref <mode> <mesh name without .msh> <out file> [key=value ...]
//   modes: mesh    dump the mesh arrays (mesh.cpp:803-855 result)
//          run     fvhyper::run (explicit.cpp:657-817) then dump q
//          loop    the run() loop with a caller-chosen alpha list (alpha=a,b,..)
//          phases  one stage of every phase from the state in ic=file:<path>
//          vtk     fvhyper::run then fvhyper::writeVtk (post.cpp:21-200)
//   keys:  flux=st|fs|nc|fp|nt  dt=st|fs|nc|fp|nt  phi=zero|fs|nc|venk|minmod
//          ic=st|fs|nc|file:<path>  gamma= cfl_fs= cfl_nc= cfl_fp= cfl_nt= mach=
//          (fp, nt: the viscous flatPlate and nacaTransient examples)
//          max_step= max_time= tol= print= alpha=
//   solver flags come from FVH_REF_FLAGS (comma list of grad,lim,lin,gdt,smooth,diff)
//   because the library reads them as link-time constants.
#include <fvhyper/mesh.h>
#include <fvhyper/explicit.h>
#include <fvhyper/parallel.h>
#include <fvhyper/post.h>

#include <cstdio>
#include <cstdlib>
#include <cstring>
#include <cstdint>
#include <fstream>
#include <string>
#include <vector>
#include <map>
#include <chrono>

namespace fvhyper {

// ---- renamed example plugins (wrap_*.cpp) ---------------------------------
#define FVH_DECLARE_PLUGIN(P)                                                  \
    void P##_generate_initial_solution(std::vector<double>&, const mesh&);     \
    double P##_limiter_func(const double&);                                    \
    void P##_calc_flux(double*, const double*, const double*, const double*,   \
                       const double*, const double*);                          \
    void P##_calc_dt(std::vector<double>&, const std::vector<double>&, mesh&); \
    namespace P##_consts { extern double gamma; }
FVH_DECLARE_PLUGIN(st)
FVH_DECLARE_PLUGIN(fs)
FVH_DECLARE_PLUGIN(nc)
FVH_DECLARE_PLUGIN(fp)
FVH_DECLARE_PLUGIN(nt)
namespace fp_consts { extern double cfl; }
namespace nt_consts { extern double cfl; extern double mach; }
namespace fp_boundaries { void wall(double*, double*, double*);
                          void farfield(double*, double*, double*); }
namespace nt_boundaries { void farfield(double*, double*, double*); }
double bg_limiter_func(const double&);
namespace fs_consts { extern double cfl; extern double mach; }
namespace nc_consts { extern double cfl; }
namespace st_boundaries { void zero_flux(double*, double*, double*); }
namespace fs_boundaries { void wall(double*, double*, double*);
                          void inlet_outlet(double*, double*, double*); }
namespace nc_boundaries { void wall(double*, double*, double*);
                          void farfield(double*, double*, double*); }
namespace st_post { void calc_output_u(double*, double*);
                    void calc_output_p(double*, double*); }

// Library functions defined in explicit.cpp with signatures that explicit.h
// does not declare (explicit.cpp:112-117, 174-182).
void calc_gradients(std::vector<double>&, std::vector<double>&,
                    const std::vector<double>&, mesh&);
void calc_limiters(std::vector<double>&, std::vector<double>&,
                   std::vector<double>&, const std::vector<double>&,
                   const std::vector<double>&, const std::vector<double>&,
                   mesh&);

// ---- run-time selection ---------------------------------------------------
static bool env_flag(const char* name) {
    const char* f = std::getenv("FVH_REF_FLAGS");
    if (!f) return false;
    std::string s = std::string(",") + f + ",";
    return s.find(std::string(",") + name + ",") != std::string::npos;
}

static std::string g_flux = "st", g_dt = "st", g_phi = "zero", g_ic = "st";

extern const int vars = 4;
extern const std::vector<std::string> var_names = {"rho", "rhou", "rhov", "rhoe"};
namespace solver {
    extern const bool do_calc_gradients = env_flag("grad");
    extern const bool do_calc_limiters = env_flag("lim");
    extern const bool linear_interpolate = env_flag("lin");
    extern const bool diffusive_gradients = env_flag("diff");
    extern const bool global_dt = env_flag("gdt");
    extern const bool smooth_residuals = env_flag("smooth");
}

// Venkatakrishnan's limiter function phi(r) = (r^2 + 2r)/(r^2 + r + 2).
// No reference example defines it; this is the form the product uses.
static double venkatakrishnan(const double& r) {
    return (r*r + 2.0*r)/(r*r + r + 2.0);
}

double limiter_func(const double& r) {
    if (g_phi == "zero") return st_limiter_func(r);
    if (g_phi == "fs") return fs_limiter_func(r);
    if (g_phi == "nc") return nc_limiter_func(r);
    if (g_phi == "minmod") return bg_limiter_func(r);
    if (g_phi == "venk") return venkatakrishnan(r);
    std::fprintf(stderr, "unknown phi %s\n", g_phi.c_str()); std::exit(2);
}

void calc_flux(double* f, const double* qi, const double* qj,
               const double* gx, const double* gy, const double* n) {
    if (g_flux == "st") st_calc_flux(f, qi, qj, gx, gy, n);
    else if (g_flux == "fs") fs_calc_flux(f, qi, qj, gx, gy, n);
    else if (g_flux == "fp") fp_calc_flux(f, qi, qj, gx, gy, n);
    else if (g_flux == "nt") nt_calc_flux(f, qi, qj, gx, gy, n);
    else nc_calc_flux(f, qi, qj, gx, gy, n);
}

void calc_dt(std::vector<double>& dt, const std::vector<double>& q, mesh& m) {
    if (g_dt == "st") st_calc_dt(dt, q, m);
    else if (g_dt == "fs") fs_calc_dt(dt, q, m);
    else if (g_dt == "fp") fp_calc_dt(dt, q, m);
    else if (g_dt == "nt") nt_calc_dt(dt, q, m);
    else nc_calc_dt(dt, q, m);
}

static std::vector<double> g_ic_data;

void generate_initial_solution(std::vector<double>& v, const mesh& m) {
    if (g_ic == "st") st_generate_initial_solution(v, m);
    else if (g_ic == "fs") fs_generate_initial_solution(v, m);
    else if (g_ic == "nc") nc_generate_initial_solution(v, m);
    else {
        if (g_ic_data.size() != v.size()) {
            std::fprintf(stderr, "ic file holds %zu values, mesh needs %zu\n",
                         g_ic_data.size(), v.size());
            std::exit(2);
        }
        v = g_ic_data;
    }
}

namespace boundaries {
    std::map<std::string, void (*)(double*, double*, double*)> bounds = {
        {"transmissive", st_boundaries::zero_flux},
        {"wall", fs_boundaries::wall},
        {"inlet", fs_boundaries::inlet_outlet},
        {"outlet", fs_boundaries::inlet_outlet},
        {"airfoil", nc_boundaries::wall},
        {"farfield", nc_boundaries::farfield},
        {"noslip", fp_boundaries::wall},
        {"farfield_fp", fp_boundaries::farfield},
        {"farfield_nt", nt_boundaries::farfield},
    };
}

namespace post {
    std::map<std::string, void (*)(double*, double*)> extra_scalars = {
        {"p", st_post::calc_output_p}};
    std::map<std::string, void (*)(double*, double*)> extra_vectors = {
        {"U", st_post::calc_output_u}};
}

}  // namespace fvhyper

// ---- dump container: records of (name, type, count, raw little-endian data)
struct Dump {
    std::FILE* f;
    explicit Dump(const std::string& path) {
        f = std::fopen(path.c_str(), "wb");
        if (!f) { std::perror(path.c_str()); std::exit(2); }
    }
    ~Dump() { std::fclose(f); }
    void head(const std::string& name, char type, uint64_t n) {
        uint32_t len = (uint32_t)name.size();
        std::fwrite(&len, 4, 1, f);
        std::fwrite(name.data(), 1, len, f);
        std::fwrite(&type, 1, 1, f);
        std::fwrite(&n, 8, 1, f);
    }
    void d(const std::string& name, const std::vector<double>& v) {
        head(name, 'd', v.size());
        std::fwrite(v.data(), 8, v.size(), f);
    }
    void u(const std::string& name, const std::vector<uint32_t>& v) {
        head(name, 'u', v.size());
        std::fwrite(v.data(), 4, v.size(), f);
    }
};

static std::vector<double> read_doubles(const std::string& path) {
    std::FILE* f = std::fopen(path.c_str(), "rb");
    if (!f) { std::perror(path.c_str()); std::exit(2); }
    std::fseek(f, 0, SEEK_END);
    long n = std::ftell(f) / 8;
    std::fseek(f, 0, SEEK_SET);
    std::vector<double> v(n);
    if (std::fread(v.data(), 8, n, f) != (size_t)n) { std::exit(2); }
    std::fclose(f);
    return v;
}

static void dump_mesh(Dump& out, fvhyper::mesh& m) {
    using std::vector;
    const uint32_t E = m.edgesNodes.cols();
    const uint32_t N = m.cellsAreas.size();
    vector<uint32_t> en(2 * E), ec(2 * E), cn(4 * N), tri(N), be(m.boundaryEdges.begin(), m.boundaryEdges.end());
    for (uint32_t e = 0; e < E; ++e) {
        en[2*e] = m.edgesNodes(e, 0); en[2*e+1] = m.edgesNodes(e, 1);
        ec[2*e] = m.edgesCells(e, 0); ec[2*e+1] = m.edgesCells(e, 1);
    }
    for (uint32_t c = 0; c < N; ++c) {
        for (int k = 0; k < 4; ++k) cn[4*c+k] = m.cellsNodes(c, k);
        tri[c] = m.cellsIsTriangle[c] ? 1 : 0;
    }
    out.u("nReal", {m.nRealCells});
    out.u("edgesNodes", en);
    out.u("edgesCells", ec);
    out.u("cellsNodes", cn);
    out.u("cellsIsTriangle", tri);
    out.u("boundaryEdges", be);
    out.d("nodesX", m.nodesX);
    out.d("nodesY", m.nodesY);
    out.d("edgesLengths", m.edgesLengths);
    out.d("edgesNormalsX", m.edgesNormalsX);
    out.d("edgesNormalsY", m.edgesNormalsY);
    out.d("edgesCentersX", m.edgesCentersX);
    out.d("edgesCentersY", m.edgesCentersY);
    out.d("cellsAreas", m.cellsAreas);
    out.d("cellsCentersX", m.cellsCentersX);
    out.d("cellsCentersY", m.cellsCentersY);
}

int main(int argc, char** argv) {
    using namespace fvhyper;
    if (argc < 4) {
        std::fprintf(stderr, "usage: fvh_ref <mode> <mesh> <out> [key=value ...]\n");
        return 2;
    }
    const std::string mode = argv[1], name = argv[2], outpath = argv[3];
    std::map<std::string, std::string> kv;
    for (int a = 4; a < argc; ++a) {
        std::string s = argv[a];
        size_t p = s.find('=');
        if (p == std::string::npos) { std::fprintf(stderr, "bad arg %s\n", argv[a]); return 2; }
        kv[s.substr(0, p)] = s.substr(p + 1);
    }
    auto get = [&](const char* k, const char* d) {
        return kv.count(k) ? kv[k] : std::string(d);
    };
    g_flux = get("flux", "st");
    g_dt = get("dt", "st");
    g_phi = get("phi", "zero");
    g_ic = get("ic", "st");
    if (g_ic.rfind("file:", 0) == 0) g_ic_data = read_doubles(g_ic.substr(5));
    if (kv.count("gamma")) {
        double g = std::stod(kv["gamma"]);
        st_consts::gamma = g; fs_consts::gamma = g; nc_consts::gamma = g;
        fp_consts::gamma = g; nt_consts::gamma = g;
    }
    if (kv.count("cfl_fs")) fs_consts::cfl = std::stod(kv["cfl_fs"]);
    if (kv.count("cfl_nc")) nc_consts::cfl = std::stod(kv["cfl_nc"]);
    if (kv.count("cfl_fp")) fp_consts::cfl = std::stod(kv["cfl_fp"]);
    if (kv.count("cfl_nt")) nt_consts::cfl = std::stod(kv["cfl_nt"]);
    if (kv.count("mach")) fs_consts::mach = std::stod(kv["mach"]);

    mpi_wrapper pool;
    mesh m;
    m.read_file(name, pool);

    solverOptions opt;
    opt.max_step = (uint)std::stoul(get("max_step", "10"));
    opt.max_time = std::stod(get("max_time", "1e10"));
    opt.tolerance = std::stod(get("tol", "1e-16"));
    opt.print_interval = (uint)std::stoul(get("print", "1"));
    opt.verbose = get("verbose", "1") == "1";

    // one dump per rank when several ranks run (<out>.<rank>)
    Dump out(pool.size > 1 ? outpath + "." + std::to_string(pool.rank) : outpath);
    if (pool.size > 1) {
        // element tag - 1 of each of this rank's cells (its numbering of the
        // partition file), and whether it is another rank's cell
        std::vector<uint32_t> orig(m.nRealCells), ghost(m.nRealCells);
        for (uint c = 0; c < m.nRealCells; ++c) {
            orig[c] = m.currentToOriginalCells.at(c);
            ghost[c] = m.cellsIsGhost[c] ? 1 : 0;
        }
        out.u("cellsOriginal", orig);
        out.u("cellsIsGhost", ghost);
    }
    if (mode == "mesh") {
        dump_mesh(out, m);
    } else if (mode == "run" || mode == "vtk") {
        std::vector<double> q;
        run(name, q, pool, m, opt);
        out.d("q", q);
        if (mode == "vtk") writeVtk(get("vtk_name", "out"), q, m, pool.rank, pool.size);
    } else if (mode == "loop") {
        // The run() loop of explicit.cpp:657-817 with a caller-chosen alpha
        // list, used for integrators the library hard-codes away (:681-687).
        std::vector<double> alpha;
        {
            std::string s = get("alpha", "1"), tok;
            size_t p0 = 0;
            while (p0 <= s.size()) {
                size_t p1 = s.find(',', p0);
                if (p1 == std::string::npos) p1 = s.size();
                alpha.push_back(std::stod(s.substr(p0, p1 - p0)));
                p0 = p1 + 1;
            }
        }
        std::vector<double> q(vars * m.cellsAreas.size());
        generate_initial_solution(q, m);
        const size_t n = q.size();
        std::vector<double> qk(n), qt(n), gx(n), gy(n), lim(n), qmin(n), qmax(n), dt(n), s0(n), s1(n);
        update_bounds(q, gx, gy, lim, m);
        uint step = 0;
        double time = 0.0;
        auto t0 = std::chrono::steady_clock::now();
        while (!((step >= opt.max_step) | (time >= opt.max_time))) {
            calc_dt(dt, q, m);
            if (solver::global_dt) min_dt(dt, m);
            qk = q;
            for (const double& a : alpha) {
                complete_calc_qt(qt, qk, gx, gy, qmin, qmax, lim, m, pool);
                if (solver::smooth_residuals) smooth_residuals(qt, s0, s1, m);
                update_cells(qk, q, qt, dt, a);
                update_bounds(qk, gx, gy, lim, m);
            }
            q = qk;
            step += 1;
            time += dt[0];
        }
        auto t1 = std::chrono::steady_clock::now();
        std::printf("%u, %.17g, %.6f\n", step, time,
                    std::chrono::duration<double>(t1 - t0).count());
        out.d("q", q);
        out.d("qt", qt);
    } else if (mode == "phases") {
        // One stage of every phase (explicit.cpp:729-747) from a given state.
        std::vector<double> q(vars * m.cellsAreas.size());
        generate_initial_solution(q, m);
        const size_t n = q.size();
        std::vector<double> dt(n), gx(n), gy(n), lim(n), qmin(n), qmax(n), qt(n), qk(n);
        const double a = std::stod(get("alpha", "0.5"));
        calc_dt(dt, q, m);
        out.d("dt_local", dt);
        if (solver::global_dt) min_dt(dt, m);
        out.d("dt", dt);
        if (solver::do_calc_gradients) calc_gradients(gx, gy, q, m);
        out.d("gx", gx);
        out.d("gy", gy);
        if (solver::do_calc_limiters) calc_limiters(lim, qmin, qmax, q, gx, gy, m);
        out.d("lim", lim);
        out.d("qmin", qmin);
        out.d("qmax", qmax);
        calc_time_derivatives(qt, q, gx, gy, lim, m);
        if (solver::smooth_residuals) {
            // as run() (explicit.cpp:743): update and residuals see the smoothed qt
            out.d("qt_raw", qt);
            std::vector<double> s0(n), s1(n);
            smooth_residuals(qt, s0, s1, m);
        }
        out.d("qt", qt);
        update_cells(qk, q, qt, dt, a);
        out.d("qk_cells", qk);
        update_bounds(qk, gx, gy, lim, m);
        out.d("qk", qk);
        std::vector<double> R(vars);
        calc_residuals(R.data(), qt, m, pool);
        out.d("R", R);
    } else {
        std::fprintf(stderr, "unknown mode %s\n", mode.c_str());
        return 2;
    }
    return pool.exit();
}
