// fvhyper_dropin.cpp -- fvhyper's mesh / run / writeVtk entry points on top
// of libfvh (the C ABI in include/fvh.h).
//
// A user translation unit written against fvhyper (examples/*/main.cpp)
// compiles unchanged against dropin/include and links this library instead
// of fvhyper/src/*.cpp.  run() executes the whole RK loop on the GPU; the
// user's host plugin functions are only called on the host to recognise
// which device kind they are (see probe.cpp) and to build the initial state.
#include <fvhyper/explicit.h>
#include <fvhyper/mesh.h>
#include <fvhyper/parallel.h>
#include <fvhyper/post.h>

#include <chrono>
#include <cstdio>
#include <cstdlib>
#include <fstream>
#include <iostream>
#include <stdexcept>

#include "../../../include/fvh.h"
#include "dropin_internal.h"

namespace fvhyper {

namespace solver {
// library constant of the reference (explicit.cpp:26-28)
extern const double limiter_k_value = 7.5;
}

// ---------------------------------------------------------------------------
// parallel
// ---------------------------------------------------------------------------
mpi_wrapper::mpi_wrapper() : rank(0), size(1) {}
int mpi_wrapper::exit() { return 0; }

// ---------------------------------------------------------------------------
// mesh
// ---------------------------------------------------------------------------
mesh::mesh() : nRealCells(0), handle_(nullptr), phase_(nullptr) {}
mesh::~mesh() {
    if (phase_) fvh_solver_free(phase_);
    if (handle_) fvh_mesh_free(handle_);
}

fvh_solver* mesh::phase_solver() const {
    if (!phase_) {
        if (!handle_) throw std::runtime_error("mesh not read");
        fvh_config c = dropin::phase_config(const_cast<mesh&>(*this));
        if (fvh_solver_create(handle_, &c, &phase_) != FVH_OK) throw std::runtime_error(fvh_last_error());
    }
    return phase_;
}

namespace {

template <class V>
void export_field(::fvh_mesh* h, const char* name, V& out) {
    const int64_t n = fvh_mesh_export(h, name, nullptr, 0);
    if (n < 0) throw std::runtime_error(fvh_last_error());
    out.resize((size_t)n);
    if (n && fvh_mesh_export(h, name, out.data(), n) < 0) throw std::runtime_error(fvh_last_error());
}

std::vector<uint> export_uint(::fvh_mesh* h, const char* name) {
    std::vector<int32_t> v;
    export_field(h, name, v);
    return std::vector<uint>(v.begin(), v.end());
}

}  // namespace

void mesh::read_file(std::string name, mpi_wrapper& pool) {
    if (pool.size > 1)
        throw std::runtime_error("one device reads whole meshes; partitioned name_<rank>.msh files need MPI ranks");
    filename = name + ".msh";
    ::fvh_mesh* h = nullptr;
    const int32_t rc = fvh_mesh_read_msh(filename.c_str(), &h);
    if (rc == FVH_E_MESH) throw std::invalid_argument(fvh_last_error());
    if (rc != FVH_OK) throw std::runtime_error(fvh_last_error());
    if (phase_) { fvh_solver_free(phase_); phase_ = nullptr; }
    if (handle_) fvh_mesh_free(handle_);
    handle_ = h;
    export_field(h, "nodesX", nodesX);
    export_field(h, "nodesY", nodesY);
    edgesNodes.assign(export_uint(h, "edgesNodes"));
    edgesCells.assign(export_uint(h, "edgesCells"));
    export_field(h, "edgesLengths", edgesLengths);
    export_field(h, "edgesNormalsX", edgesNormalsX);
    export_field(h, "edgesNormalsY", edgesNormalsY);
    export_field(h, "edgesCentersX", edgesCentersX);
    export_field(h, "edgesCentersY", edgesCentersY);
    cellsNodes.assign(export_uint(h, "cellsNodes"));
    {
        std::vector<int32_t> t;
        export_field(h, "cellsIsTriangle", t);
        cellsIsTriangle.assign(t.begin(), t.end());
    }
    export_field(h, "cellsAreas", cellsAreas);
    export_field(h, "cellsCentersX", cellsCentersX);
    export_field(h, "cellsCentersY", cellsCentersY);
    boundaryEdges = export_uint(h, "boundaryEdges");
    int64_t counts[6];
    fvh_mesh_counts(h, counts);
    nRealCells = (uint)counts[1];
    cellsIsGhost.assign(cellsAreas.size(), false);
    // physical names and boundary functions (mesh.cpp:660-667)
    physicalNames.clear();
    std::vector<std::string> names((size_t)counts[5]);
    char buf[256];
    for (int32_t k = 0; k < (int32_t)counts[5]; ++k) {
        fvh_mesh_name(h, k, buf, sizeof buf);
        names[k] = buf;
        physicalNames[(uint)k] = buf;
    }
    const std::vector<uint> tags = export_uint(h, "boundaryTags");
    boundaryFuncs.clear();
    boundaryNames.clear();
    for (size_t b = 0; b < tags.size(); ++b) {
        const std::string& nm = names[tags[b]];
        auto it = boundaries::bounds.find(nm);
        if (it == boundaries::bounds.end())
            throw std::invalid_argument("Physical name " + nm + " of tag " + std::to_string(tags[b]) + " not in mesh");
        boundaryFuncs.push_back(it->second);
        boundaryNames.push_back(nm);
    }
}

// ---------------------------------------------------------------------------
// run (explicit.cpp:657-817)
// ---------------------------------------------------------------------------
namespace {

void print_row(uint step, double time, double seconds, const double* R, bool first) {
    std::cout << step << ", " << time << ", " << seconds << ", ";
    for (int i = 0; i < vars; ++i) {
        if (first) std::cout << "1.0";
        else std::cout << R[i];
        if (i < vars - 1) std::cout << ", ";
    }
    std::cout << std::endl;
}

}  // namespace

void run(const std::string name, std::vector<double>& q, mpi_wrapper& pool, mesh& m, solverOptions& opt) {
    if (!m.device_mesh()) throw std::runtime_error("run(): mesh not read");
    q.resize(vars * m.cellsAreas.size());
    generate_initial_solution(q, m);

    fvh_config cfg = dropin::device_config(m, q);
    fvh_solver* s = nullptr;
    if (fvh_solver_create(m.device_mesh(), &cfg, &s) != FVH_OK) throw std::runtime_error(fvh_last_error());
    struct Guard { fvh_solver* s; ~Guard() { fvh_solver_free(s); } } guard{s};
    auto check = [](int32_t rc) { if (rc != FVH_OK) throw std::runtime_error(fvh_last_error()); };
    check(fvh_solver_set_state(s, q.data()));

    if (opt.verbose & (pool.rank == 0)) {
        std::cout << "Step, Time, RealTime, ";
        for (int i = 0; i < vars; ++i) {
            std::cout << "R(q[" << i << "])";
            if (i < vars - 1) std::cout << ", ";
        }
        std::cout << std::endl;
    }
    double save_time = opt.time_series_interval;
    uint time_step = 0;
    check(fvh_solver_run_begin(s));
    fvh_run_options o;
    o.max_time = opt.max_time;
    o.max_step = opt.max_step;
    o.print_interval = opt.print_interval ? opt.print_interval : 1;
    o.tolerance = opt.tolerance;
    o.stop_var = -1;
    o.reserved = 0;
    const auto begin = std::chrono::steady_clock::now();
    auto seconds = [&]() {
        const auto end = std::chrono::steady_clock::now();
        const int us = (int)std::chrono::duration_cast<std::chrono::microseconds>(end - begin).count();
        return ((double)us) / 1e6;
    };
    fvh_status st{};
    uint done = 0;
    for (;;) {
        // advance to the next step that prints or saves, one step at a time
        // when a time series is written
        uint n = 1;
        if (!opt.save_time_series && done > 0) {
            // stop right after the next step whose index is a multiple of print_interval
            const uint next = ((done - 1) / o.print_interval + 1) * o.print_interval;
            n = next + 1 - done;
        }
        check(fvh_solver_run_continue(s, &o, n, &st));
        const bool moved = st.step > done;
        done = st.step;
        if (!moved) break;
        const uint step = done - 1;   // the step just taken, as the reference numbers it
        if (opt.verbose & (pool.rank == 0)) {
            if (step == 0) print_row(step, st.time_prev, seconds(), st.R, true);
            else if (step % o.print_interval == 0) print_row(step, st.time_prev, seconds(), st.R, false);
        }
        if (opt.save_time_series && st.time_prev > save_time) {
            check(fvh_solver_get_state(s, q.data()));
            writeVtk(name, q, m, pool.rank, pool.size, std::to_string(time_step));
            save_time += opt.time_series_interval;
            time_step += 1;
        }
        if (st.stopped) break;
    }
    check(fvh_solver_get_state(s, q.data()));
    if (opt.verbose & (pool.rank == 0)) print_row(st.step, st.time, seconds(), st.R, false);
}

// ---------------------------------------------------------------------------
// the phases one at a time (explicit.cpp:112-417, 481-564, 625-653)
// ---------------------------------------------------------------------------
namespace {

void check_rc(int32_t rc) {
    if (rc != FVH_OK) throw std::runtime_error(fvh_last_error());
}

// vars values per cell over all cells
size_t state_size(const mesh& m) { return (size_t)vars * m.cellsAreas.size(); }

void need(const std::vector<double>& v, const mesh& m, const char* what) {
    if (v.size() < state_size(m))
        throw std::invalid_argument(std::string(what) + ": fewer than vars values per cell");
}

void fit(std::vector<double>& v, const mesh& m) {
    if (v.size() < state_size(m)) v.resize(state_size(m));
}

}  // namespace

void calc_gradients(std::vector<double>& gx, std::vector<double>& gy, const std::vector<double>& q, mesh& m) {
    need(q, m, "calc_gradients");
    fit(gx, m);
    fit(gy, m);
    check_rc(fvh_phase_gradients(m.phase_solver(), q.data(), gx.data(), gy.data()));
}

void calc_gradients(std::vector<double>& gx, std::vector<double>& gy, const std::vector<double>& q,
                    const mesh& m) {
    calc_gradients(gx, gy, q, const_cast<mesh&>(m));
}

void calc_limiters(std::vector<double>& limiters, std::vector<double>& qmin, std::vector<double>& qmax,
                   const std::vector<double>& q, const std::vector<double>& gx, const std::vector<double>& gy,
                   mesh& m) {
    need(q, m, "calc_limiters");
    need(gx, m, "calc_limiters");
    need(gy, m, "calc_limiters");
    fit(limiters, m);
    fit(qmin, m);
    fit(qmax, m);
    check_rc(fvh_phase_limiters(m.phase_solver(), q.data(), gx.data(), gy.data(), limiters.data(), qmin.data(),
                                qmax.data()));
}

void calc_limiters(std::vector<double>& limiters, const std::vector<double>& q, const std::vector<double>& gx,
                   const std::vector<double>& gy, const mesh& m) {
    need(q, m, "calc_limiters");
    need(gx, m, "calc_limiters");
    need(gy, m, "calc_limiters");
    fit(limiters, m);
    check_rc(fvh_phase_limiters(const_cast<mesh&>(m).phase_solver(), q.data(), gx.data(), gy.data(),
                                limiters.data(), nullptr, nullptr));
}

void calc_time_derivatives(std::vector<double>& qt, const std::vector<double>& q, const std::vector<double>& gx,
                           const std::vector<double>& gy, const std::vector<double>& limiters, mesh& m) {
    need(q, m, "calc_time_derivatives");
    need(gx, m, "calc_time_derivatives");
    need(gy, m, "calc_time_derivatives");
    need(limiters, m, "calc_time_derivatives");
    fit(qt, m);
    check_rc(fvh_phase_time_derivatives(m.phase_solver(), q.data(), gx.data(), gy.data(), limiters.data(),
                                        qt.data()));
}

void update_cells(std::vector<double>& q, std::vector<double>& ql, const std::vector<double>& qt,
                  const std::vector<double>& dt, const double v) {
    if (ql.size() < q.size() || qt.size() < q.size() || dt.size() < q.size())
        throw std::invalid_argument("update_cells: ql, qt and dt need as many values as q");
    check_rc(fvh_phase_update_cells(nullptr, q.data(), ql.data(), qt.data(), dt.data(), v, (int64_t)q.size()));
}

void update_bounds(std::vector<double>& q, std::vector<double>& gx, std::vector<double>& gy,
                   std::vector<double>& limiters, mesh& m) {
    need(q, m, "update_bounds");
    need(gx, m, "update_bounds");
    need(gy, m, "update_bounds");
    need(limiters, m, "update_bounds");
    check_rc(fvh_phase_update_bounds(m.phase_solver(), q.data(), gx.data(), gy.data(), limiters.data()));
}

void update_comms(std::vector<double>&, mesh&) {}

void calc_residuals(double* R, std::vector<double>& qt, mesh& m, mpi_wrapper&) {
    need(qt, m, "calc_residuals");
    check_rc(fvh_phase_residuals(m.phase_solver(), qt.data(), R));
}

void min_dt(std::vector<double>& dt, mesh& m) {
    if (dt.empty()) return;
    check_rc(fvh_phase_min_dt(m.phase_solver(), dt.data(), (int64_t)dt.size()));
}

void validate_dt(std::vector<double>&, mpi_wrapper&) {}

void complete_calc_qt(std::vector<double>& qt, std::vector<double>& q, std::vector<double>& gx,
                      std::vector<double>& gy, std::vector<double>& qmin, std::vector<double>& qmax,
                      std::vector<double>& limiters, mesh& m, mpi_wrapper& pool) {
    if (solver::do_calc_gradients) calc_gradients(gx, gy, q, m);
    if (solver::do_calc_limiters) calc_limiters(limiters, qmin, qmax, q, gx, gy, m);
    (void)pool;
    calc_time_derivatives(qt, q, gx, gy, limiters, m);
}

// ---------------------------------------------------------------------------
// writeVtk (post.cpp:21-200): same names, layout and std::to_string numbers
// ---------------------------------------------------------------------------
bool vtk_full_precision = false;

void writeVtk(const std::string name, std::vector<double>& q, mesh& m, int rank, int world_size,
              const std::string time) {
    // full precision (an extension, off by default): %.17g numbers and
    // Float64 arrays, for post-processing that needs the exact state
    const char* fp = std::getenv("FVH_VTK_FULL_PRECISION");
    const bool full = vtk_full_precision || (fp && fp[0] == '1');
    auto num = [full](double x) {
        if (!full) return std::to_string(x);
        char buf[40];
        std::snprintf(buf, sizeof buf, "%.17g", x);
        return std::string(buf);
    };
    const std::string ftype = full ? "Float64" : "Float32";
    const std::string dash_time = time.empty() ? "" : "_" + time;
    std::string filename = (world_size > 1) ? name + "_" + std::to_string(rank) + dash_time + ".vtu"
                                            : name + dash_time + ".vtu";
    if (!time.empty()) filename = "times/" + filename;
    const size_t V = var_names.size();
    const uint N = m.nRealCells;
    std::string s;
    s.reserve(64 * (size_t)N + 64 * m.nodesX.size());
    s += "<VTKFile type=\"UnstructuredGrid\" version=\"0.1\" byte_order=\"BigEndian\">\n";
    s += "  <UnstructuredGrid>\n";
    s += "    <Piece NumberOfPoints=\"" + std::to_string(m.nodesX.size()) + "\" NumberOfCells=\"" +
         std::to_string(N) + "\">\n";
    s += "      <Points>\n";
    s += "        <DataArray type=\"" + ftype + "\" NumberOfComponents=\"3\" Format=\"ascii\">\n";
    for (size_t i = 0; i < m.nodesX.size(); ++i)
        s += "          " + num(m.nodesX[i]) + " " + num(m.nodesY[i]) + " 0.0\n";
    s += "        </DataArray>\n";
    s += "      </Points>\n";
    s += "      <Cells>\n";
    s += "        <DataArray type=\"Int32\" Name=\"connectivity\" Format=\"ascii\">\n";
    s += "          ";
    for (uint c = 0; c < N; ++c) {
        const uint nn = m.cellsIsTriangle[c] ? 3 : 4;
        for (uint k = 0; k < nn; ++k) s += std::to_string(m.cellsNodes(c, k)) + "  ";
        s += "\n          ";
    }
    s += "        </DataArray>\n";
    s += "        <DataArray type=\"Int32\" Name=\"offsets\" Format=\"ascii\">\n";
    s += "          ";
    uint offset = 0;
    for (uint c = 0; c < N; ++c) {
        offset += m.cellsIsTriangle[c] ? 3 : 4;
        s += std::to_string(offset) + "  ";
    }
    s += "\n";
    s += "        </DataArray>\n";
    s += "        <DataArray type=\"Int32\" Name=\"types\" Format=\"ascii\">\n";
    s += "          ";
    for (uint c = 0; c < N; ++c) s += m.cellsIsTriangle[c] ? "5  " : "9  ";
    s += "\n";
    s += "        </DataArray>\n";
    s += "      </Cells>\n";
    s += "      <CellData Scalars=\"scalars\">\n";
    s += "        <DataArray type=\"Float32\" Name=\"rank\" Format=\"ascii\">\n";
    for (uint c = 0; c < N; ++c) s += "          " + std::to_string(rank) + "\n";
    s += "        </DataArray>\n";
    for (size_t k = 0; k < V; ++k) {
        s += "        <DataArray type=\"" + ftype + "\" Name=\"" + var_names[k] + "\" Format=\"ascii\">\n";
        for (uint c = 0; c < N; ++c) s += "          " + num(q[V * c + k]) + "\n";
        s += "        </DataArray>\n";
    }
    for (auto& kv : post::extra_scalars) {
        s += "        <DataArray type=\"" + ftype + "\" Name=\"" + kv.first + "\" Format=\"ascii\">\n";
        for (uint c = 0; c < N; ++c) {
            double out[1];
            kv.second(out, &q[V * c]);
            s += "          " + num(out[0]) + "\n";
        }
        s += "        </DataArray>\n";
    }
    for (auto& kv : post::extra_vectors) {
        s += "        <DataArray type=\"" + ftype + "\" Name=\"" + kv.first +
             "\" NumberOfComponents=\"3\" Format=\"ascii\">\n";
        for (uint c = 0; c < N; ++c) {
            double out[2];
            kv.second(out, &q[V * c]);
            s += "          " + num(out[0]) + " " + num(out[1]) + " 0.0\n";
        }
        s += "        </DataArray>\n";
    }
    s += "      </CellData>\n";
    s += "    </Piece>\n";
    s += "  </UnstructuredGrid>\n";
    s += "</VTKFile>";
    std::ofstream out(filename);
    out << s;
    if (rank == 0 && world_size > 1) {
        // the parallel master file is only written for several ranks (post.cpp:40-81)
    }
}

}  // namespace fvhyper
