// fvh_solver.hip -- device layout, RK driver and the C ABI of the solver.
//
// Replaces fvhyper::run (fvhyper/src/explicit.cpp:657-817) and the per-phase
// functions it calls, on one GPU.  The host code here only builds the device
// layout and queues kernels; every per-cell, per-edge and per-step operation
// runs on the GPU.
//
// A solver holds one or more independent instances (meshes with their own
// boundary tables, step counters, time, residuals and stop flags).  A single
// mesh is a batch of one.  Device layout of a batch: the real cells of every
// instance (instance by instance, each in its own locality order), then the
// ghost cells of every instance; tiles of at most 256 cells never straddle
// two instances, so every workgroup works for exactly one instance and the
// per-instance reductions are segmented sums over tile ranges.  Batching
// changes no arithmetic: each instance's results are bit-identical to a run
// of that mesh alone.
#include <hip/hip_runtime.h>

#include <algorithm>
#include <chrono>
#include <cmath>
#include <cstdio>
#include <cstring>
#include <memory>
#include <mutex>
#include <string>
#include <vector>

#include "fvh_internal.h"
#include "fvh_kernels.h"

namespace fvh {
namespace {

const char* kKernelNames[] = {"dt", "dt_min", "gradlim", "flux", "ctrl_begin", "ctrl_end", "bounds_init", "stage",
                              "smooth"};
enum { K_DT = 0, K_DTMIN, K_GRADLIM, K_FLUX, K_BEGIN, K_END, K_BINIT, K_STAGE, K_SMOOTH, K_COUNT };

#define FVH_HIP(call)                                                                       \
    do {                                                                                    \
        hipError_t e_ = (call);                                                             \
        if (e_ != hipSuccess) {                                                             \
            set_error(std::string(#call) + ": " + hipGetErrorString(e_));                   \
            return FVH_E_HIP;                                                               \
        }                                                                                   \
    } while (0)

// Calls F<T, S>(args...) for the solver's arithmetic type T and state
// storage type S.
#define FVH_DISPATCH(s, F, ...)                                                     \
    ((s)->mixed ? F<double, float>(__VA_ARGS__)                                     \
                : (s)->f32 ? F<float, float>(__VA_ARGS__) : F<double, double>(__VA_ARGS__))

struct DevBuf {
    void* p = nullptr;
    size_t bytes = 0;
    DevBuf() = default;
    DevBuf(const DevBuf&) = delete;
    DevBuf& operator=(const DevBuf&) = delete;
    ~DevBuf() { if (p) (void)hipFree(p); }
    int alloc(size_t n) {
        if (p) { (void)hipFree(p); p = nullptr; }
        bytes = n;
        if (n == 0) n = 8;
        return hipMalloc(&p, n) == hipSuccess ? 0 : -1;
    }
};

constexpr int kAdjMaxStages = FVH_MAX_ALPHA;

// One captured RK step, replayed with hipGraphLaunch.
struct StepGraph {
    int cur = 0, need = 0, run_mode = 0;
    fvh_run_options o{};
    int cur_after = 0;
    hipGraph_t graph = nullptr;
    hipGraphExec_t exec = nullptr;
};

}  // namespace
}  // namespace fvh

struct fvh_solver {
    fvh_config cfg{};                     // physics, the same for every instance
    std::vector<fvh_config> icfg;         // per instance (boundary tables)
    int device = 0;
    hipStream_t stream = nullptr;
    bool f32 = false;                     // FP32 arithmetic, geometry and states
    bool mixed = false;                   // FP32 states, FP64 arithmetic and geometry
    bool second = false;                  // gradient/limiter/reconstruction pass present
    int32_t n_inst = 0;
    int32_t n_real = 0, n_total = 0, n_edges = 0, n_he = 0, n_bnd = 0;
    std::vector<int32_t> inv;             // device cell -> cell of the concatenated reference-order state
    // per instance
    std::vector<int32_t> i_ncells, i_real0, i_nreal, i_ghost0, i_nghost, i_dev0;
    std::vector<int64_t> i_state0;        // first cell of the instance in the concatenated state
    std::vector<int4> i_tiles;            // (rt0, rt1, gt0, gt1)
    // tiles
    int32_t n_rt = 0, n_all = 0;
    fvh::DevBuf tile_c0, tile_inst, inst_tiles, inst_dev0;
    // device buffers
    fvh::DevBuf he_off, he_nbr, he_edge, he_twin, he_bc, he_slot;
    fvh::DevBuf e_geo, e_c, c_area, c_inv_area, c_sqrt_area, c_c;
    fvh::DevBuf t_off, t_edge, t_he, t_cell;
    bool use_tiles = false;
    // fused stage pass: tile halos
    bool fused = false;
    fvh::DevBuf x_ext, x_cells, x_r1_start, x_r1_loc, x_r1_slot, x_he_loc;
    int32_t x_max_ext = 0, x_max_g = 0, x_max_edges = 0, x_mean_g = 0, x_lds = 0;
    fvh::DevBuf q[3], dt, dt_global, face, bincr, blk_min, res_blk, ctrl, bcs, sums_blk, sums;
    fvh::DevBuf soft_part;   // smoothed kinks: fixed-point partial sums of the p-norm time step, per tile
    fvh::DevBuf stage_queue; // persistent stage form: kXcd ticket counters and a done counter (zero between launches)
    fvh::DevBuf unphysical;                  // per instance: a stage left a non-physical cell state
    fvh::DevBuf dbg[4];
    fvh::DevBuf ph[6], ph_part;              // per-phase entry points: inputs and outputs, min_dt partials
    fvh::DevBuf qt_s[2];                     // residual smoothing: raw qt, first sweep's qt
    fvh::DevBuf gxy[2];                      // viscous flux: cell gradients gx, gy
    int cur = 0;                          // index of the current state buffer
    // discrete adjoint: per device edge (half-edge side 0, side 1, cell side 0,
    // cell side 1); boundary element geometry (instance 0); work buffers
    fvh::DevBuf epair;
    std::vector<double> b_nx, b_ny, b_len;
    fvh::DevBuf a_qs[fvh::kAdjMaxStages + 1], a_gx[fvh::kAdjMaxStages], a_gy[fvh::kAdjMaxStages],
        a_lm[fvh::kAdjMaxStages], a_qt[fvh::kAdjMaxStages];
    fvh::DevBuf a_lam, a_lam_q0, a_lam_qt, a_lam_bq, a_lam_bincr, a_lam_face, a_lam_direct, a_nb, a_ld,
        a_tmp;
    fvh::DevBuf a_obj_ghost, a_obj_gw, a_obj_terms;
    fvh::DevBuf a_ref, a_amin_v, a_amin_r, a_amin_d, a_amin;   // device -> reference index; argmin scratch
    std::vector<std::unique_ptr<fvh::DevBuf>> a_ck;
    std::vector<std::unique_ptr<fvh::DevBuf>> a_tape;   // stage records of the last steps (one buffer per step)
    hipGraph_t a_graph = nullptr;            // one step's replay + reverse, captured once
    hipGraphExec_t a_graph_exec = nullptr;
    bool a_graph_geo = false;                // the captured reverse forms the geometry adjoint
    // geometry adjoint (node-coordinate gradients): per half-edge, half-edge
    // centre, edge, cell; host maps to the reference numbering (one instance)
    fvh::DevBuf a_g_he, a_g_hc, a_g_e, a_g_c, a_obj_p;
    std::vector<uint32_t> h_he_edge;         // device half-edge -> device edge | flags
    std::vector<int32_t> h_edge_ref;         // device edge -> reference edge of its instance
    std::unique_ptr<fvh_mesh> mesh0;         // instance 0's mesh (single-mesh solvers)
    int32_t max_real_deg = 0;                // most half-edges of any real cell
    // graphs
    bool use_graphs = false;
    std::vector<fvh::StepGraph> graphs;
    // profiling
    bool profiling = false;
    std::vector<std::pair<int, std::pair<hipEvent_t, hipEvent_t>>> pending;
    std::vector<hipEvent_t> event_pool;
    int64_t k_count[fvh::K_COUNT] = {};
    double k_ms[fvh::K_COUNT] = {};
    double stage_bytes = 0.0;

    void drop_graphs() {
        for (auto& g : graphs) {
            if (g.exec) (void)hipGraphExecDestroy(g.exec);
            if (g.graph) (void)hipGraphDestroy(g.graph);
        }
        graphs.clear();
    }
    ~fvh_solver() {
        if (stream) (void)hipStreamSynchronize(stream);
        drop_graphs();
        if (a_graph_exec) (void)hipGraphExecDestroy(a_graph_exec);
        if (a_graph) (void)hipGraphDestroy(a_graph);
        for (auto& pe : pending) { event_pool.push_back(pe.second.first); event_pool.push_back(pe.second.second); }
        for (hipEvent_t e : event_pool) (void)hipEventDestroy(e);
        if (stream) (void)hipStreamDestroy(stream);
    }
};

namespace fvh {
namespace {

template <class T>
MeshView<T> view(fvh_solver* s) {
    MeshView<T> m;
    m.n_real = s->n_real;
    m.n_total = s->n_total;
    m.he_off = (const int32_t*)s->he_off.p;
    m.he_nbr = (const int32_t*)s->he_nbr.p;
    m.he_edge = (const uint32_t*)s->he_edge.p;
    m.he_twin = (const int32_t*)s->he_twin.p;
    m.he_bc = (const int32_t*)s->he_bc.p;
    m.he_slot = (const uint16_t*)s->he_slot.p;
    m.e_geo = (const Q4<T>*)s->e_geo.p;
    m.e_c = (const G2<T>*)s->e_c.p;
    m.c_area = (const T*)s->c_area.p;
    m.c_inv_area = (const T*)s->c_inv_area.p;
    m.c_sqrt_area = (const T*)s->c_sqrt_area.p;
    m.c_c = (const G2<T>*)s->c_c.p;
    return m;
}

TileMap tile_map(fvh_solver* s) {
    TileMap t;
    t.n_rt = s->n_rt;
    t.n_all = s->n_all;
    t.c0 = (const int32_t*)s->tile_c0.p;
    t.inst = (const int32_t*)s->tile_inst.p;
    return t;
}

InstView inst_view(fvh_solver* s) {
    InstView v;
    v.n = s->n_inst;
    v.tiles = (const int4*)s->inst_tiles.p;
    v.dev0 = (const int32_t*)s->inst_dev0.p;
    return v;
}

HaloView halo_view(fvh_solver* s) {
    HaloView h;
    h.ext = (const int4*)s->x_ext.p;
    h.cells = (const int32_t*)s->x_cells.p;
    h.r1_start = (const int32_t*)s->x_r1_start.p;
    h.r1_loc = (const uint16_t*)s->x_r1_loc.p;
    h.he_loc = (const uint16_t*)s->x_he_loc.p;
    h.r1_slot = (const uint16_t*)s->x_r1_slot.p;
    h.max_ext = s->x_max_ext;
    h.max_g = s->x_max_g;
    h.mean_g = s->x_mean_g;
    h.max_edges = s->x_max_edges;
    h.lds_bytes = s->x_lds;
    return h;
}

TileView tile_view(fvh_solver* s) {
    TileView t;
    t.n_tiles = s->n_rt;
    t.off = (const int32_t*)s->t_off.p;
    t.edge = (const int32_t*)s->t_edge.p;
    t.he = (const int2*)s->t_he.p;
    t.cell = (const int2*)s->t_cell.p;
    return t;
}

// Kernel kinds of a configuration: with smoothed kinks, the limiter, flux and
// summed-CFL time-step kinds of the smoothed kernels.
inline bool kinks(const fvh_config& c) { return (c.flags & FVH_FLAG_SMOOTH_KINKS) != 0; }
inline int kphi(const fvh_config& c) { return kinks(c) ? (int)PHI_VENKAT_S : c.phi; }
inline int kflux(const fvh_config& c) { return kinks(c) ? c.flux + 3 : c.flux; }
inline int kdt(const fvh_config& c) { return kinks(c) && c.dt_kind == FVH_DT_SUM_CFL ? 5 : c.dt_kind; }

PhysParams phys(const fvh_config& c) {
    PhysParams p;
    p.gamma = c.gamma;
    p.cfl = c.cfl;
    p.dt_fixed = c.dt_fixed;
    p.K = c.K;
    p.phi_yt = c.phi_yt;
    p.positivity = (c.flags & FVH_FLAG_POSITIVITY) ? 1 : 0;
    p.mu = c.mu;
    p.pr = c.pr;
    p.cp = c.cp;
    p.r = c.r_gas;
    p.diffusive = (c.flags & FVH_FLAG_DIFFUSIVE) ? 1 : 0;
    p.unphysical = nullptr;
    // widths of the smoothed kinks: kink_width times kink_scale, a zero scale
    // from the first characteristic boundary's freestream (or 1)
    double sc[5] = {1.0, 1.0, 1.0, 1.0, 1.0};
    for (int k = 0; k < c.n_bc; ++k) {
        const fvh_bc_spec& b = c.bc[k];
        if (b.kind != FVH_BC_INLET_OUTLET && b.kind != FVH_BC_FARFIELD && b.kind != FVH_BC_FARFIELD_V) continue;
        const double speed = std::sqrt(b.u * b.u + b.v * b.v);
        sc[0] = b.rho;
        sc[1] = sc[2] = b.rho * (speed + std::sqrt(c.gamma * b.p / b.rho));
        sc[3] = b.p / (c.gamma - 1.0) + 0.5 * b.rho * speed * speed;
        sc[4] = speed + std::sqrt(c.gamma * b.p / b.rho);
        break;
    }
    const double w = c.kink_width > 0.0 ? c.kink_width : 1e-2;
    for (int k = 0; k < 4; ++k) p.kw[k] = w * (c.kink_scale[k] > 0.0 ? c.kink_scale[k] : sc[k]);
    p.kwv = w * (c.kink_scale[4] > 0.0 ? c.kink_scale[4] : sc[4]);
    p.kwl = w;
    return p;
}

// kMaxBc slots per instance, instance-major
template <class T>
std::vector<BcParam<T>> bc_table(const std::vector<fvh_config>& cs) {
    std::vector<BcParam<T>> t(kMaxBc * cs.size());
    for (size_t i = 0; i < cs.size(); ++i) {
        const fvh_config& c = cs[i];
        for (int k = 0; k < kMaxBc; ++k) {
            BcParam<T>& b = t[i * kMaxBc + k];
            b.kind = k < c.n_bc ? c.bc[k].kind : 0;
            b.rho = T(k < c.n_bc ? c.bc[k].rho : 0);
            b.u = T(k < c.n_bc ? c.bc[k].u : 0);
            b.v = T(k < c.n_bc ? c.bc[k].v : 0);
            b.p = T(k < c.n_bc ? c.bc[k].p : 0);
        }
    }
    return t;
}

// Event bracket around one launch when profiling is on.
struct Bracket {
    fvh_solver* s;
    int k;
    hipEvent_t a = nullptr, b = nullptr;
    Bracket(fvh_solver* s_, int k_) : s(s_), k(k_) {
        if (!s->profiling) return;
        a = take();
        b = take();
        (void)hipEventRecord(a, s->stream);
    }
    ~Bracket() {
        if (!s->profiling) return;
        (void)hipEventRecord(b, s->stream);
        s->pending.push_back({k, {a, b}});
    }
    hipEvent_t take() {
        if (!s->event_pool.empty()) {
            hipEvent_t e = s->event_pool.back();
            s->event_pool.pop_back();
            return e;
        }
        hipEvent_t e;
        (void)hipEventCreate(&e);
        return e;
    }
};

int32_t drain_events(fvh_solver* s) {
    if (s->pending.empty()) return FVH_OK;
    FVH_HIP(hipStreamSynchronize(s->stream));
    for (auto& pe : s->pending) {
        float ms = 0.f;
        (void)hipEventElapsedTime(&ms, pe.second.first, pe.second.second);
        s->k_count[pe.first] += 1;
        s->k_ms[pe.first] += ms;
        s->event_pool.push_back(pe.second.first);
        s->event_pool.push_back(pe.second.second);
    }
    s->pending.clear();
    return FVH_OK;
}

int32_t upload(DevBuf& b, const void* src, size_t bytes) {
    if (b.alloc(bytes)) { set_error("hipMalloc failed"); return FVH_E_HIP; }
    if (bytes) FVH_HIP(hipMemcpy(b.p, src, bytes, hipMemcpyHostToDevice));
    return FVH_OK;
}

// Host geometry in device order (doubles; converted to the state type on upload).
struct HostGeo {
    std::vector<double> e_nx, e_ny, e_len, e_gg, e_cx, e_cy;   // per device edge
    std::vector<double> c_area, c_cx, c_cy;                    // per device cell
};

template <class T, class S>
int32_t upload_geometry(fvh_solver* s, const HostGeo& h) {
    const int32_t E = s->n_edges, N = s->n_total;
    std::vector<Q4<T>> geo(E);
    std::vector<G2<T>> ec(E);
    for (int32_t d = 0; d < E; ++d) {
        geo[d].v[0] = T(h.e_nx[d]);
        geo[d].v[1] = T(h.e_ny[d]);
        geo[d].v[2] = T(h.e_len[d]);
        geo[d].v[3] = T(h.e_gg[d]);
        ec[d].v[0] = T(h.e_cx[d]);
        ec[d].v[1] = T(h.e_cy[d]);
    }
    std::vector<T> area(N), ia(N), sa(N);
    std::vector<G2<T>> cc(N);
    for (int32_t d = 0; d < N; ++d) {
        const double A = h.c_area[d];
        area[d] = T(A);
        ia[d] = T(1. / A);          // invA of explicit.cpp:158
        sa[d] = T(std::sqrt(A));    // sqrt_area of explicit.cpp:218
        cc[d].v[0] = T(h.c_cx[d]);
        cc[d].v[1] = T(h.c_cy[d]);
    }
    int32_t rc;
    const std::vector<BcParam<T>> bt = bc_table<T>(s->icfg);
    if ((rc = upload(s->bcs, bt.data(), bt.size() * sizeof(BcParam<T>)))) return rc;
    if ((rc = upload(s->e_geo, geo.data(), geo.size() * sizeof(Q4<T>)))) return rc;
    if ((rc = upload(s->e_c, ec.data(), ec.size() * sizeof(G2<T>)))) return rc;
    if ((rc = upload(s->c_area, area.data(), N * sizeof(T)))) return rc;
    if ((rc = upload(s->c_inv_area, ia.data(), N * sizeof(T)))) return rc;
    if ((rc = upload(s->c_sqrt_area, sa.data(), N * sizeof(T)))) return rc;
    if ((rc = upload(s->c_c, cc.data(), N * sizeof(G2<T>)))) return rc;
    for (int k = 0; k < 3; ++k)
        if (s->q[k].alloc((size_t)N * sizeof(Q4<S>))) { set_error("hipMalloc failed"); return FVH_E_HIP; }
    if (s->dt.alloc((size_t)N * sizeof(T)) || s->dt_global.alloc((size_t)s->n_inst * 8) ||
        s->face.alloc((size_t)(s->second ? s->n_he : 1) * sizeof(Q4<T>)) ||
        s->bincr.alloc((size_t)(s->n_bnd + 1) * sizeof(Q4<T>)) ||
        s->blk_min.alloc((size_t)(s->n_all + 1) * 8) || s->res_blk.alloc((size_t)(s->n_rt + 1) * 32) ||
        s->sums_blk.alloc((size_t)(s->n_rt + 1) * 32) || s->sums.alloc((size_t)s->n_inst * 32) ||
        s->ctrl.alloc((size_t)s->n_inst * sizeof(Ctrl)) || s->unphysical.alloc((size_t)s->n_inst * 4)) {
        set_error("hipMalloc failed");
        return FVH_E_HIP;
    }
    if ((s->cfg.flags & FVH_FLAG_SMOOTH_KINKS) && s->soft_part.alloc((size_t)(s->n_all + 1) * 16)) {
        set_error("hipMalloc failed");
        return FVH_E_HIP;
    }
    if (s->stage_queue.alloc(64) || hipMemset(s->stage_queue.p, 0, 64) != hipSuccess) {
        set_error("hipMalloc failed");
        return FVH_E_HIP;
    }
    if (s->cfg.flags & FVH_FLAG_SMOOTH)
        for (auto& b : s->qt_s)
            if (b.alloc((size_t)(s->n_real + 1) * sizeof(Q4<T>))) {
                set_error("hipMalloc failed");
                return FVH_E_HIP;
            }
    if (s->cfg.flux == FVH_FLUX_ROE_VISCOUS && (s->cfg.flags & FVH_FLAG_DIFFUSIVE))
        for (auto& b : s->gxy) {
            // zero until a gradient pass writes them (and for good without one)
            if (b.alloc((size_t)N * sizeof(Q4<T>))) {
                set_error("hipMalloc failed");
                return FVH_E_HIP;
            }
            FVH_HIP(hipMemset(b.p, 0, b.bytes));
        }
    FVH_HIP(hipMemset(s->ctrl.p, 0, s->ctrl.bytes));
    FVH_HIP(hipMemset(s->unphysical.p, 0, s->unphysical.bytes));
    FVH_HIP(hipMemset(s->face.p, 0, s->face.bytes));
    FVH_HIP(hipMemset(s->bincr.p, 0, s->bincr.bytes));
    return FVH_OK;
}

// Queue the kernels of one RK step (explicit.cpp:729-747) reading state
// buffer s->cur.  run_mode adds the stop test and bookkeeping kernels.
// dbg (tests): 4 device buffers for gx, gy, limiter and qt of one stage run
// with alpha_override, leaving the current state buffer unchanged.
// The work-queue counters of the persistent stage form, or null for one
// workgroup per tile.
inline int32_t* stage_queue(fvh_solver* s) {
    return s->cfg.stage_launch == 1 ? (int32_t*)s->stage_queue.p : nullptr;
}

// The global time step of each instance from the per-cell steps: min_dt,
// or with smoothed kinks and a CFL step its p-norm form.
template <class T>
void global_dt(fvh_solver* s, hipStream_t st, const TileMap& tm, const InstView& iv, const double* blk, T* dt,
               T* dtg, const Ctrl* cc) {
    launch_dt_min(st, iv, blk, dt, s->f32, dtg, cc);
    if (kinks(s->cfg) && s->cfg.dt_kind != FVH_DT_FIXED)
        launch_dt_soft(st, tm, iv, (const double*)dt, (double*)dtg, (uint64_t*)s->soft_part.p, cc);
}

template <class T, class S>
int32_t enqueue_step(fvh_solver* s, bool run_mode, const fvh_run_options* o, int need_res,
                     Q4<T>* const* dbg = nullptr, const double* alpha_override = nullptr) {
    const fvh_config& c = s->cfg;
    const MeshView<T> m = view<T>(s);
    const TileMap tm = tile_map(s);
    const InstView iv = inst_view(s);
    const TileView tv = tile_view(s);
    PhysParams pp = phys(c);
    pp.unphysical = (uint32_t*)s->unphysical.p;
    const BcParam<T>* bcs = (const BcParam<T>*)s->bcs.p;
    const bool grad = c.flags & FVH_FLAG_GRADIENTS, lim = c.flags & FVH_FLAG_LIMITERS,
               lin = c.flags & FVH_FLAG_LINEAR, gdt = c.flags & FVH_FLAG_GLOBAL_DT,
               smooth = c.flags & FVH_FLAG_SMOOTH;
    Ctrl* ctrl = (Ctrl*)s->ctrl.p;
    const Ctrl* cc = run_mode ? ctrl : nullptr;
    hipStream_t st = s->stream;
    if (run_mode) {
        Bracket b(s, K_BEGIN);
        launch_ctrl_begin(st, s->n_inst, ctrl, o->max_step, o->max_time, o->tolerance, o->stop_var);
    }
    Q4<S>* q0 = (Q4<S>*)s->q[s->cur].p;
    Q4<S>* qa = (Q4<S>*)s->q[(s->cur + 1) % 3].p;
    Q4<S>* qb = (Q4<S>*)s->q[(s->cur + 2) % 3].p;
    T* dt = (T*)s->dt.p;
    T* dtg = (T*)s->dt_global.p;
    double* blk = (double*)s->blk_min.p;
    Q4<T>* face = (Q4<T>*)s->face.p;
    Q4<T>* bincr = (Q4<T>*)s->bincr.p;
    Q4<T>* gx_c = (Q4<T>*)s->gxy[0].p;   // null unless the viscous flux takes face gradients
    Q4<T>* gy_c = (Q4<T>*)s->gxy[1].p;
    const bool fused = s->fused && !dbg;
    const bool kk = kinks(c);
    // the time-step kernel: first order, a fixed step (the second-order
    // passes compute only the CFL kinds), the fused form's global step and
    // every step with smoothed kinks
    if (!s->second || c.dt_kind == FVH_DT_FIXED || (fused && gdt) || kk) {
        Bracket b(s, K_DT);
        launch_dt<T>(st, m, tm, kdt(c), pp, q0, dt, blk, cc);
    }
    const Q4<S>* qin = q0;
    Q4<S>* out = qa;
    const int n_stage = alpha_override ? 1 : c.n_alpha;
    if (fused) {
        const HaloView hv = halo_view(s);
        if (gdt) {
            Bracket b(s, K_DTMIN);
            global_dt(s, st, tm, iv, blk, dt, dtg, cc);
        }
        for (int k = 0; k < n_stage; ++k) {
            Bracket b(s, K_STAGE);
            const bool last = k == n_stage - 1;
            launch_stage<T>(st, m, tm, tv, hv, pp, kphi(c), c.flux == FVH_FLUX_ROE_JOINED, grad, lim,
                            (k == 0 && !gdt && !kk) ? c.dt_kind : -1, bcs, qin, q0, out, (Q4<T>*)s->bincr.p, dt,
                            gdt ? dtg : nullptr,
                            T(c.alpha[k]), (last && need_res) ? (double*)s->res_blk.p : nullptr, cc,
                            c.stage_threads, StageRecord<T>(), stage_queue(s));
            qin = out;
            out = (out == qa) ? qb : qa;
        }
    }
    for (int k = 0; k < (fused ? 0 : n_stage); ++k) {
        if (s->second) {
            // the first stage's pass also computes the time step from q0
            Bracket b(s, K_GRADLIM);
            launch_gradlim<T>(st, m, tm, pp, kphi(c), grad, lim, lin, qin, face, bincr,
                              (k == 0 && !kk) ? c.dt_kind : -1, dt, blk, dbg ? dbg[0] : gx_c, dbg ? dbg[1] : gy_c,
                              dbg ? dbg[2] : nullptr, cc);
        }
        if (k == 0 && gdt) {
            Bracket b(s, K_DTMIN);
            global_dt(s, st, tm, iv, blk, dt, dtg, cc);
        }
        const bool last = k == n_stage - 1;
        const T a = T(alpha_override ? *alpha_override : c.alpha[k]);
        double* res = (last && need_res) ? (double*)s->res_blk.p : nullptr;
        const Q4<T>* gxv = dbg ? dbg[0] : gx_c;
        const Q4<T>* gyv = dbg ? dbg[1] : gy_c;
        if (smooth) {
            // flux pass stores qt; two smoothing sweeps, the second finishes the stage
            Q4<T>* qs0 = (Q4<T>*)s->qt_s[0].p;
            Q4<T>* qs1 = (Q4<T>*)s->qt_s[1].p;
            {
                Bracket b(s, K_FLUX);
                launch_flux<T>(st, m, tm, s->use_tiles ? &tv : nullptr, pp, kflux(c), lin, bcs, qin, q0, out, face,
                               bincr, dt, gdt ? dtg : nullptr, a, nullptr, nullptr, cc, qs0, gxv, gyv);
            }
            {
                Bracket b(s, K_SMOOTH);
                launch_smooth<T>(st, m, tm, pp, lin, bcs, qs0, qs0, qs1, q0, out, bincr, dt, gdt ? dtg : nullptr, a,
                                 nullptr, nullptr, cc);
            }
            Bracket b(s, K_SMOOTH);
            launch_smooth<T>(st, m, tm, pp, lin, bcs, qs1, qs0, nullptr, q0, out, bincr, dt, gdt ? dtg : nullptr, a,
                             res, dbg ? dbg[3] : nullptr, cc);
        } else {
            Bracket b(s, K_FLUX);
            launch_flux<T>(st, m, tm, s->use_tiles ? &tv : nullptr, pp, kflux(c), lin, bcs, qin, q0, out, face, bincr,
                           dt, gdt ? dtg : nullptr, a, res, dbg ? dbg[3] : nullptr, cc, nullptr, gxv, gyv);
        }
        qin = out;
        out = (out == qa) ? qb : qa;
    }
    if (!dbg) {
        Bracket b(s, K_END);
        launch_ctrl_end(st, iv, ctrl, (const double*)s->res_blk.p, need_res, gdt ? s->dt_global.p : s->dt.p,
                        s->f32, gdt ? 1 : 0);
    }
    FVH_HIP(hipGetLastError());
    // the last stage wrote qin; it becomes the current state
    if (!dbg) s->cur = (qin == qa) ? (s->cur + 1) % 3 : (s->cur + 2) % 3;
    return FVH_OK;
}

bool same_options(const fvh_run_options& a, const fvh_run_options& b) {
    return a.max_time == b.max_time && a.max_step == b.max_step && a.tolerance == b.tolerance &&
           a.stop_var == b.stop_var;
}

// One step, replayed from a captured graph when graphs are on (a graph per
// state buffer, residual flag and run options; kernels take their buffers
// and options by value, so a replay is the same launch sequence).
template <class T, class S>
int32_t step(fvh_solver* s, bool run_mode, const fvh_run_options* o, int need_res) {
    if (!s->use_graphs || s->profiling) return enqueue_step<T, S>(s, run_mode, o, need_res);
    fvh_run_options key{};
    if (run_mode) key = *o;
    for (auto& g : s->graphs) {
        if (g.cur == s->cur && g.need == need_res && g.run_mode == (int)run_mode && same_options(g.o, key)) {
            FVH_HIP(hipGraphLaunch(g.exec, s->stream));
            s->cur = g.cur_after;
            return FVH_OK;
        }
    }
    if (s->graphs.size() >= 24) s->drop_graphs();
    StepGraph g;
    g.cur = s->cur;
    g.need = need_res;
    g.run_mode = run_mode;
    g.o = key;
    FVH_HIP(hipStreamBeginCapture(s->stream, hipStreamCaptureModeThreadLocal));
    int32_t rc = enqueue_step<T, S>(s, run_mode, o, need_res);
    hipGraph_t graph = nullptr;
    const hipError_t e = hipStreamEndCapture(s->stream, &graph);
    if (rc) {
        if (graph) (void)hipGraphDestroy(graph);
        s->cur = g.cur;
        return rc;
    }
    if (e != hipSuccess) {
        s->cur = g.cur;
        set_error(std::string("hipStreamEndCapture: ") + hipGetErrorString(e));
        return FVH_E_HIP;
    }
    g.cur_after = s->cur;
    g.graph = graph;
    FVH_HIP(hipGraphInstantiate(&g.exec, graph, nullptr, nullptr, 0));
    s->graphs.push_back(g);
    FVH_HIP(hipGraphLaunch(g.exec, s->stream));
    return FVH_OK;
}

template <class T, class S>
int32_t set_state_t(fvh_solver* s, const double* q) {
    std::vector<Q4<S>> h(s->n_total);
    for (int32_t d = 0; d < s->n_total; ++d) {
        const int64_t r = s->inv[d];
        for (int k = 0; k < 4; ++k) h[d].v[k] = S(q[4 * r + k]);
    }
    FVH_HIP(hipStreamSynchronize(s->stream));
    for (int b = 0; b < 3; ++b)
        FVH_HIP(hipMemcpy(s->q[b].p, h.data(), h.size() * sizeof(Q4<S>), hipMemcpyHostToDevice));
    return FVH_OK;
}

template <class T>
int32_t get_q4(fvh_solver* s, const void* dev, double* out, int32_t count) {
    std::vector<Q4<T>> h(count);
    FVH_HIP(hipStreamSynchronize(s->stream));
    FVH_HIP(hipMemcpy(h.data(), dev, h.size() * sizeof(Q4<T>), hipMemcpyDeviceToHost));
    for (int32_t d = 0; d < count; ++d) {
        const int64_t r = s->inv[d];
        for (int k = 0; k < 4; ++k) out[4 * r + k] = (double)h[d].v[k];
    }
    return FVH_OK;
}

template <class T, class S>
int32_t phases_t(fvh_solver* s, double alpha, double* dt_out, double* gx, double* gy, double* lim,
                 double* qt, double* qk) {
    const bool gdt = s->cfg.flags & FVH_FLAG_GLOBAL_DT;
    const size_t qbytes = (size_t)s->n_total * sizeof(Q4<T>);
    for (int k = 0; k < 4; ++k) {
        if (!s->dbg[k].p && s->dbg[k].alloc(qbytes)) { set_error("hipMalloc failed"); return FVH_E_HIP; }
        FVH_HIP(hipMemsetAsync(s->dbg[k].p, 0, qbytes, s->stream));
    }
    Q4<S>* q0 = (Q4<S>*)s->q[s->cur].p;
    Q4<S>* out = (Q4<S>*)s->q[(s->cur + 1) % 3].p;
    // update_cells runs over every cell; ghost cells keep q (their qt is 0)
    FVH_HIP(hipMemcpyAsync(out, q0, (size_t)s->n_total * sizeof(Q4<S>), hipMemcpyDeviceToDevice, s->stream));
    Q4<T>* dbg[4] = {(Q4<T>*)s->dbg[0].p, (Q4<T>*)s->dbg[1].p, (Q4<T>*)s->dbg[2].p, (Q4<T>*)s->dbg[3].p};
    int32_t rc = enqueue_step<T, S>(s, false, nullptr, 0, dbg, &alpha);
    if (rc) return rc;
    FVH_HIP(hipStreamSynchronize(s->stream));
    if (dt_out) {
        std::vector<T> h(s->n_total);
        T g = T(0);
        FVH_HIP(hipMemcpy(h.data(), s->dt.p, h.size() * sizeof(T), hipMemcpyDeviceToHost));
        if (gdt) FVH_HIP(hipMemcpy(&g, s->dt_global.p, sizeof(T), hipMemcpyDeviceToHost));
        for (int32_t d = 0; d < s->n_total; ++d) dt_out[s->inv[d]] = (double)(gdt ? g : h[d]);
    }
    if (gx && (rc = get_q4<T>(s, dbg[0], gx, s->n_total))) return rc;
    if (gy && (rc = get_q4<T>(s, dbg[1], gy, s->n_total))) return rc;
    if (lim && (rc = get_q4<T>(s, dbg[2], lim, s->n_total))) return rc;
    if (qt && (rc = get_q4<T>(s, dbg[3], qt, s->n_total))) return rc;
    if (qk && (rc = get_q4<S>(s, out, qk, s->n_total))) return rc;
    return FVH_OK;
}

template <class T, class S>
int32_t apply_bounds_t(fvh_solver* s) {
    const fvh_config& c = s->cfg;
    {
        Bracket b(s, K_BINIT);
        launch_bounds_init<T>(s->stream, view<T>(s), phys(c), (c.flags & FVH_FLAG_LINEAR) != 0,
                              (const BcParam<T>*)s->bcs.p, (Q4<S>*)s->q[s->cur].p);
    }
    FVH_HIP(hipGetLastError());
    std::vector<Ctrl> z(s->n_inst);
    std::memset(z.data(), 0, z.size() * sizeof(Ctrl));
    for (auto& c0 : z)
        for (int k = 0; k < 4; ++k) c0.R[k] = 1.0;
    FVH_HIP(hipMemcpyAsync(s->ctrl.p, z.data(), z.size() * sizeof(Ctrl), hipMemcpyHostToDevice, s->stream));
    FVH_HIP(hipMemsetAsync(s->unphysical.p, 0, s->unphysical.bytes, s->stream));
    FVH_HIP(hipStreamSynchronize(s->stream));
    return FVH_OK;
}

int32_t read_ctrl(fvh_solver* s, std::vector<Ctrl>& h) {
    h.resize(s->n_inst);
    FVH_HIP(hipStreamSynchronize(s->stream));
    FVH_HIP(hipMemcpy(h.data(), s->ctrl.p, h.size() * sizeof(Ctrl), hipMemcpyDeviceToHost));
    return FVH_OK;
}

void to_status(const Ctrl& h, fvh_status* st) {
    st->step = h.step;
    st->stopped = h.stop;
    st->time = h.time;
    st->time_prev = h.time_prev;
    for (int k = 0; k < 4; ++k) { st->R[k] = h.R[k]; st->R0[k] = h.R0[k]; }
}

// Status of the whole solver: instance 0's counters and residuals, with
// `stopped` set once every instance has stopped and `step` the largest count.
int32_t read_status(fvh_solver* s, fvh_status* st) {
    std::vector<Ctrl> h;
    int32_t rc = read_ctrl(s, h);
    if (rc) return rc;
    to_status(h[0], st);
    uint32_t all = 1, mx = 0;
    for (const Ctrl& c : h) { all &= c.stop ? 1u : 0u; mx = std::max(mx, c.step); }
    st->stopped = all;
    st->step = mx;
    return FVH_OK;
}

// Failure detection: instances whose stage updates left a real cell with
// rho <= 0, p <= 0 or a non-finite value, or whose residuals are not finite.
// Fills msg (empty when all is well).
int32_t check_state(fvh_solver* s, std::string& msg) {
    std::vector<uint32_t> f(s->n_inst);
    std::vector<Ctrl> h;
    int32_t rc = read_ctrl(s, h);
    if (rc) return rc;
    FVH_HIP(hipMemcpy(f.data(), s->unphysical.p, f.size() * 4, hipMemcpyDeviceToHost));
    msg.clear();
    int shown = 0, n_bad = 0;
    for (int i = 0; i < s->n_inst; ++i) {
        if (!f[i] && !h[i].nonfinite) continue;
        ++n_bad;
        if (shown++ < 4)
            msg += (msg.empty() ? "" : "; ") + std::string("instance ") + std::to_string(i) + " by step " +
                   std::to_string(h[i].step) + (f[i] ? ": a cell state with rho <= 0, p <= 0 or not finite"
                                                     : ": non-finite residual");
    }
    if (n_bad > shown) msg += "; and " + std::to_string(n_bad - shown) + " more instances";
    if (n_bad) msg = "non-physical state: " + msg;
    return FVH_OK;
}

// Executes up to n more steps of run() from the current status (steps are
// queued in growing chunks; kernels queued after an instance's stop are
// no-ops for it).  Afterwards every instance's state is in buffer s->cur.
template <class T, class S>
int32_t run_continue_t(fvh_solver* s, const fvh_run_options* o, uint64_t n, fvh_status* st) {
    int32_t rc;
    std::vector<Ctrl> before, after;
    if ((rc = read_ctrl(s, before))) return rc;
    fvh_status now;
    if ((rc = read_status(s, &now))) return rc;
    const uint32_t print = o->print_interval ? o->print_interval : 1;
    const bool every = o->tolerance > 1.01e-16;
    uint64_t step = now.step;
    const int cur0 = s->cur;
    // each executed step moves the state forward by one or two buffers
    const int delta = (s->cfg.n_alpha % 2) ? 1 : 2;
    uint64_t chunk = 8, left = n;
    std::string bad;
    while (!now.stopped && left > 0) {
        const uint64_t k_n = std::min<uint64_t>(chunk, left);
        for (uint64_t k = 0; k < k_n; ++k, ++step) {
            const int need = (step == 0) || (step % print == 0) || every;
            if ((rc = fvh::step<T, S>(s, true, o, need))) return rc;
        }
        left -= k_n;
        if ((rc = read_status(s, &now))) return rc;
        if ((rc = check_state(s, bad))) return rc;
        if (!bad.empty()) break;
        if (chunk < 256) chunk *= 2;
    }
    if ((rc = read_ctrl(s, after))) return rc;
    uint64_t most = 0;
    for (int i = 0; i < s->n_inst; ++i) most = std::max<uint64_t>(most, after[i].step - before[i].step);
    s->cur = (int)((cur0 + most * delta) % 3);
    // an instance that stopped early holds its state in another buffer
    for (int i = 0; i < s->n_inst; ++i) {
        const uint64_t done = after[i].step - before[i].step;
        const int b = (int)((cur0 + done * delta) % 3);
        if (b == s->cur) continue;
        const size_t qs = sizeof(Q4<S>);
        char* dst = (char*)s->q[s->cur].p;
        const char* src = (const char*)s->q[b].p;
        FVH_HIP(hipMemcpyAsync(dst + (size_t)s->i_real0[i] * qs, src + (size_t)s->i_real0[i] * qs,
                               (size_t)s->i_nreal[i] * qs, hipMemcpyDeviceToDevice, s->stream));
        if (s->i_nghost[i])
            FVH_HIP(hipMemcpyAsync(dst + (size_t)s->i_ghost0[i] * qs, src + (size_t)s->i_ghost0[i] * qs,
                                   (size_t)s->i_nghost[i] * qs, hipMemcpyDeviceToDevice, s->stream));
    }
    FVH_HIP(hipStreamSynchronize(s->stream));
    *st = now;
    if ((rc = drain_events(s))) return rc;
    if (!bad.empty()) {
        set_error(bad);
        return FVH_E_STATE;
    }
    return FVH_OK;
}

template <class T, class S>
int32_t run_t(fvh_solver* s, const fvh_run_options* o, fvh_status* st) {
    int32_t rc = apply_bounds_t<T, S>(s);
    if (rc) return rc;
    return run_continue_t<T, S>(s, o, ~0ull, st);
}

// Tile edge lists for the tiled flux pass.  Returns false when a tile touches
// more edges than its LDS holds.
bool build_tiles(fvh_solver* s, const std::vector<int32_t>& tile_c0, const std::vector<int32_t>& off,
                 const std::vector<uint32_t>& he_edge, std::vector<uint16_t>& he_slot, std::vector<int32_t>& t_off,
                 std::vector<int32_t>& t_edge, std::vector<int32_t>& t_he, std::vector<int32_t>& t_cell) {
    const int32_t E = s->n_edges;
    std::vector<int32_t> eh(2 * (size_t)E, -1), ecell(2 * (size_t)E, -1);
    for (int32_t d = 0; d < s->n_total; ++d)
        for (int32_t h = off[d]; h < off[d + 1]; ++h) {
            const uint32_t ue = he_edge[h];
            const int32_t e = (int32_t)(ue & kEdgeMask);
            const int sd = (ue & kSideBit) ? 1 : 0;
            eh[2 * (size_t)e + sd] = h;
            ecell[2 * (size_t)e + sd] = d;
        }
    t_off.assign(s->n_rt + 1, 0);
    std::vector<int32_t> local;
    std::vector<int32_t> slot_of(E, -1);
    for (int32_t t = 0; t < s->n_rt; ++t) {
        const int32_t c0 = tile_c0[t], c1 = tile_c0[t + 1];
        local.clear();
        for (int32_t d = c0; d < c1; ++d)
            for (int32_t h = off[d]; h < off[d + 1]; ++h) local.push_back((int32_t)(he_edge[h] & kEdgeMask));
        std::sort(local.begin(), local.end());
        local.erase(std::unique(local.begin(), local.end()), local.end());
        if ((int)local.size() > kTileEdges) return false;
        for (size_t k = 0; k < local.size(); ++k) {
            const int32_t e = local[k];
            slot_of[e] = (int32_t)k;
            t_edge.push_back(e);
            t_he.push_back(eh[2 * (size_t)e]);
            t_he.push_back(eh[2 * (size_t)e + 1]);
            t_cell.push_back(ecell[2 * (size_t)e]);
            t_cell.push_back(ecell[2 * (size_t)e + 1]);
        }
        for (int32_t d = c0; d < c1; ++d)
            for (int32_t h = off[d]; h < off[d + 1]; ++h) he_slot[h] = (uint16_t)slot_of[he_edge[h] & kEdgeMask];
        t_off[t + 1] = t_off[t] + (int32_t)local.size();
    }
    return true;
}

// LDS bytes of the fused stage pass for a tile with n_ext local cells and
// n_edges edges, in the solver's precision.
size_t halo_lds_bytes(const fvh_solver* s, int32_t n_ext, int32_t n_edges) {
    // with the positivity fallback, the face states' pressures follow the
    // edge slots (two per tile edge, in the arithmetic type)
    const bool pos = (s->cfg.flags & FVH_FLAG_POSITIVITY) != 0;
    if (s->mixed) return stage_lds_tile<double, float>(n_ext, n_edges, pos);
    return s->f32 ? stage_lds_tile<float, float>(n_ext, n_edges, pos)
                  : stage_lds_tile<double, double>(n_ext, n_edges, pos);
}

// Halos of the real tiles for the fused stage pass.  Returns false when a
// tile's local cell list does not fit 16-bit indices.
struct HaloHost {
    std::vector<int4> ext;
    std::vector<int32_t> cells, r1_start;
    std::vector<uint16_t> r1_loc, r1_slot, he_loc;
    int32_t max_ext = 0, max_g = 0, max_edges = 0;
    size_t max_lds = 0;  // the largest tile's LDS bytes
    int64_t sum_g = 0;   // tile cells plus real ring-1 cells, over all tiles
};

bool build_halo(fvh_solver* s, const std::vector<int32_t>& tile_c0, const std::vector<int32_t>& off,
                const std::vector<int32_t>& he_nbr, const std::vector<int32_t>& t_off,
                const std::vector<int32_t>& t_he, HaloHost& hh) {
    const int32_t N = s->n_total;
    std::vector<int32_t> mark(N, -1);
    std::vector<int32_t> r1r, r1g, r2;
    std::vector<int32_t> slot_of_he(s->n_he, -1);
    hh.ext.resize(s->n_rt);
    hh.he_loc.assign(s->n_he, 0);
    for (int32_t t = 0; t < s->n_rt; ++t) {
        const int32_t c0 = tile_c0[t], c1 = tile_c0[t + 1], nt = c1 - c0;
        for (int32_t c = c0; c < c1; ++c) mark[c] = c - c0;
        r1r.clear(); r1g.clear(); r2.clear();
        for (int32_t c = c0; c < c1; ++c)
            for (int32_t h = off[c]; h < off[c + 1]; ++h) {
                const int32_t nb = he_nbr[h];
                if (mark[nb] != -1) continue;
                mark[nb] = -2;
                (nb < s->n_real ? r1r : r1g).push_back(nb);
            }
        std::sort(r1r.begin(), r1r.end());
        std::sort(r1g.begin(), r1g.end());
        const int32_t n1r = (int32_t)r1r.size(), n1 = n1r + (int32_t)r1g.size();
        for (int32_t k = 0; k < n1r; ++k) mark[r1r[k]] = nt + k;
        for (size_t k = 0; k < r1g.size(); ++k) mark[r1g[k]] = nt + n1r + (int32_t)k;
        for (int32_t c : r1r)
            for (int32_t h = off[c]; h < off[c + 1]; ++h) {
                const int32_t nb = he_nbr[h];
                if (mark[nb] != -1) continue;
                mark[nb] = -2;
                r2.push_back(nb);
            }
        std::sort(r2.begin(), r2.end());
        for (size_t k = 0; k < r2.size(); ++k) mark[r2[k]] = nt + n1 + (int32_t)k;
        const int32_t n_ext = nt + n1 + (int32_t)r2.size();
        if (n_ext > 65535) return false;
        for (int32_t c = c0; c < c1; ++c)
            for (int32_t h = off[c]; h < off[c + 1]; ++h) hh.he_loc[h] = (uint16_t)mark[he_nbr[h]];
        const int32_t base = (int32_t)hh.cells.size();
        hh.ext[t] = make_int4(base, n1r, n1, (int32_t)r2.size());
        // the tile's edge slot of each half-edge on one of its edges
        for (int32_t k = t_off[t]; k < t_off[t + 1]; ++k) {
            slot_of_he[t_he[2 * (size_t)k]] = k - t_off[t];
            slot_of_he[t_he[2 * (size_t)k + 1]] = k - t_off[t];
        }
        for (int32_t c : r1r) {
            hh.cells.push_back(c);
            hh.r1_start.push_back((int32_t)hh.r1_loc.size());
            for (int32_t h = off[c]; h < off[c + 1]; ++h) {
                hh.r1_loc.push_back((uint16_t)mark[he_nbr[h]]);
                hh.r1_slot.push_back(slot_of_he[h] < 0 ? (uint16_t)0xffff : (uint16_t)slot_of_he[h]);
            }
        }
        for (int32_t c : r1g) { hh.cells.push_back(c); hh.r1_start.push_back(0); }
        for (int32_t c : r2) { hh.cells.push_back(c); hh.r1_start.push_back(0); }
        for (int32_t k = t_off[t]; k < t_off[t + 1]; ++k) {
            slot_of_he[t_he[2 * (size_t)k]] = -1;
            slot_of_he[t_he[2 * (size_t)k + 1]] = -1;
        }
        if (t_off[t + 1] - t_off[t] > 65535) return false;
        hh.max_ext = std::max(hh.max_ext, n_ext);
        hh.max_g = std::max(hh.max_g, nt + n1r);
        hh.sum_g += nt + n1r;
        hh.max_edges = std::max(hh.max_edges, t_off[t + 1] - t_off[t]);
        hh.max_lds = std::max(hh.max_lds, halo_lds_bytes(s, n_ext, t_off[t + 1] - t_off[t]));
        for (int32_t c = c0; c < c1; ++c) mark[c] = -1;
        for (int32_t c : r1r) mark[c] = -1;
        for (int32_t c : r1g) mark[c] = -1;
        for (int32_t c : r2) mark[c] = -1;
    }
    if (hh.r1_loc.empty()) { hh.r1_loc.push_back(0); hh.r1_slot.push_back(0xffff); }
    if (hh.cells.empty()) { hh.cells.push_back(0); hh.r1_start.push_back(0); }
    return true;
}

// LDS bytes of the fused stage pass: its largest tile's.
size_t halo_lds_bytes(const HaloHost& hh) { return hh.max_lds; }

int32_t check_config(const fvh_config* c) {
    if (c->n_alpha < 1 || c->n_alpha > FVH_MAX_ALPHA || c->n_bc < 0 || c->n_bc > FVH_MAX_BC) {
        set_error("config out of range (alpha or bc count)");
        return FVH_E_ARG;
    }
    if (c->precision != FVH_F64 && c->precision != FVH_F32 && c->precision != FVH_F32_F64) {
        set_error("unknown precision");
        return FVH_E_ARG;
    }
    if (c->flux < 0 || c->flux > 2 || c->phi < 0 || c->phi > 3 || c->dt_kind < 0 || c->dt_kind > 4 ||
        c->renumber < 0 || c->renumber > 2 || c->flux_mode < 0 || c->flux_mode > 3) {
        set_error("unknown flux, limiter, time-step, order or flux-mode kind");
        return FVH_E_ARG;
    }
    if (c->stage_threads != 0 && c->stage_threads != kStageBlock / 2 && c->stage_threads != kStageBlock) {
        set_error("stage_threads must be 0, 256 or 512");
        return FVH_E_ARG;
    }
    if (c->stage_launch != 0 && c->stage_launch != 1) {
        set_error("stage_launch must be 0 (a workgroup per tile) or 1 (persistent work queues)");
        return FVH_E_ARG;
    }
    if (c->tile_cells != 0 && (c->tile_cells < 32 || c->tile_cells > kBlock)) {
        set_error("tile_cells must be 0 or 32..256");
        return FVH_E_ARG;
    }
    for (int k = 0; k < c->n_bc; ++k)
        if (c->bc[k].kind < 0 || c->bc[k].kind > 5) { set_error("unknown boundary kind"); return FVH_E_ARG; }
    const bool visc = c->flux == FVH_FLUX_ROE_VISCOUS || c->dt_kind == FVH_DT_EDGE_MIN_VISCOUS ||
                      c->dt_kind == FVH_DT_SUM_VISCOUS;
    if (visc && (!(c->pr != 0.0) || !(c->r_gas != 0.0))) {
        set_error("viscous flux or time step: pr and r_gas must be non-zero");
        return FVH_E_ARG;
    }
    if (c->flags & FVH_FLAG_SMOOTH_KINKS) {
        const uint32_t second = FVH_FLAG_GRADIENTS | FVH_FLAG_LIMITERS | FVH_FLAG_LINEAR;
        if (c->precision != FVH_F64 || c->phi != FVH_PHI_VENKATAKRISHNAN || (c->flags & second) != second ||
            c->flux == FVH_FLUX_ROE_VISCOUS || (c->dt_kind != FVH_DT_SUM_CFL && c->dt_kind != FVH_DT_FIXED)) {
            set_error("smoothed kinks take FP64, the Venkatakrishnan limiter with gradients, limiters and linear "
                      "reconstruction, an inviscid Roe flux and the summed-CFL or fixed time step");
            return FVH_E_UNSUPPORTED;
        }
        if (!(c->kink_width >= 0.0) || !(c->kink_width < 1.0)) {
            set_error("kink_width must lie in [0, 1)");
            return FVH_E_ARG;
        }
        for (int k = 0; k < 5; ++k)
            if (!(c->kink_scale[k] >= 0.0) || !std::isfinite(c->kink_scale[k])) {
                set_error("kink_scale must be finite and non-negative");
                return FVH_E_ARG;
            }
    }
    return FVH_OK;
}


// Everything but the boundary table must agree across a batch.
bool same_physics(const fvh_config& a, const fvh_config& b) {
    if (a.precision != b.precision || a.flux != b.flux || a.gamma != b.gamma || a.phi != b.phi ||
        a.phi_yt != b.phi_yt || a.K != b.K || a.dt_kind != b.dt_kind || a.dt_fixed != b.dt_fixed ||
        a.cfl != b.cfl || a.flags != b.flags || a.n_alpha != b.n_alpha || a.renumber != b.renumber ||
        a.flux_mode != b.flux_mode || a.tile_cells != b.tile_cells || a.mu != b.mu || a.pr != b.pr ||
        a.cp != b.cp || a.r_gas != b.r_gas || a.stage_threads != b.stage_threads || a.kink_width != b.kink_width ||
        a.stage_launch != b.stage_launch)
        return false;
    for (int k = 0; k < 5; ++k)
        if (a.kink_scale[k] != b.kink_scale[k]) return false;
    for (int k = 0; k < a.n_alpha; ++k)
        if (a.alpha[k] != b.alpha[k]) return false;
    return true;
}

// Build the device layout of n instances and upload it.  splits marks, per
// instance, the RCB tiles to bisect once more (null: none).
int32_t create(int32_t n, const fvh_mesh* const* ms, const fvh_config* cs, fvh_solver** out,
               const std::vector<std::vector<uint8_t>>* splits = nullptr, bool auto_shrunk = false) {
    if (n < 1 || !ms || !cs || !out) { set_error("null argument or empty batch"); return FVH_E_ARG; }
    int32_t rc;
    for (int32_t i = 0; i < n; ++i) {
        if (!ms[i]) { set_error("null mesh"); return FVH_E_ARG; }
        if ((rc = check_config(&cs[i]))) return rc;
        if (!same_physics(cs[0], cs[i])) {
            set_error("batch instance " + std::to_string(i) + ": configs may differ only in their boundary tables");
            return FVH_E_ARG;
        }
    }
    std::unique_ptr<fvh_solver> s(new fvh_solver());
    const fvh_config& c = cs[0];
    s->cfg = c;
    s->icfg.assign(cs, cs + n);
    s->n_inst = n;
    s->f32 = c.precision == FVH_F32;
    s->mixed = c.precision == FVH_F32_F64;
    s->second = (c.flags & (FVH_FLAG_GRADIENTS | FVH_FLAG_LIMITERS | FVH_FLAG_LINEAR)) != 0;
    // physical name -> BC slot of each instance (mesh.cpp:662-667)
    std::vector<std::vector<int32_t>> slot_of_name(n);
    int64_t tot_real = 0, tot_cells = 0, tot_edges = 0;
    for (int32_t i = 0; i < n; ++i) {
        const fvh_mesh& m = *ms[i];
        slot_of_name[i].assign(m.names.size(), -1);
        for (size_t t = 0; t < m.names.size(); ++t)
            for (int k = 0; k < cs[i].n_bc; ++k)
                if (m.names[t] == std::string(cs[i].bc[k].name)) slot_of_name[i][t] = k;
        for (size_t b = 0; b < m.bnd_tag.size(); ++b) {
            const int32_t t = m.bnd_tag[b];
            if (slot_of_name[i][t] < 0) {
                set_error("Physical name " + m.names[t] + " of tag " + std::to_string(t) + " not in mesh");
                return FVH_E_BC;
            }
        }
        tot_real += m.n_real;
        tot_cells += m.n_cells;
        tot_edges += m.n_edges();
    }
    if (tot_edges >= (int64_t)kEdgeMask || tot_cells >= (int64_t)INT32_MAX / 4) {
        set_error("too many edges or cells");
        return FVH_E_UNSUPPORTED;
    }
    if (hipGetDevice(&s->device) != hipSuccess) { set_error("no HIP device"); return FVH_E_HIP; }
    s->n_real = (int32_t)tot_real;
    s->n_total = (int32_t)tot_cells;
    s->n_edges = (int32_t)tot_edges;
    const int32_t N = s->n_total, E = s->n_edges;
    const int32_t tile = c.tile_cells ? c.tile_cells : kBlock;
    s->inv.assign(N, 0);
    s->i_ncells.resize(n); s->i_real0.resize(n); s->i_nreal.resize(n); s->i_ghost0.resize(n);
    s->i_nghost.resize(n); s->i_dev0.resize(n); s->i_state0.resize(n); s->i_tiles.resize(n);
    // global ids: edge_base[i] + local edge; device cell of (instance, local reference cell)
    std::vector<int32_t> edge_base(n), cell_dev;     // cell_dev over the concatenated state
    cell_dev.assign(N, 0);
    std::vector<int32_t> tile_c0, tile_inst;
    // tile sizes of the real cells per instance (empty: tiles of `tile` cells)
    std::vector<std::vector<int32_t>> leaves(n);
    {
        int32_t r0 = 0, g0 = s->n_real, eb = 0;
        int64_t st0 = 0;
        for (int32_t i = 0; i < n; ++i) {
            const fvh_mesh& m = *ms[i];
            std::vector<int32_t> order;
            if (c.renumber == FVH_ORDER_RCM) order = rcm_order(m);
            else if (c.renumber == FVH_ORDER_TILES)
                order = rcb_order(m, tile, splits ? &(*splits)[i] : nullptr, &leaves[i]);
            else { order.resize(m.n_real); for (int32_t k = 0; k < m.n_real; ++k) order[k] = k; }
            s->i_ncells[i] = m.n_cells;
            s->i_real0[i] = r0;
            s->i_nreal[i] = m.n_real;
            s->i_ghost0[i] = g0;
            s->i_nghost[i] = m.n_cells - m.n_real;
            s->i_state0[i] = st0;
            for (int32_t k = 0; k < m.n_real; ++k) {
                cell_dev[st0 + order[k]] = r0 + k;
                s->inv[r0 + k] = (int32_t)(st0 + order[k]);
            }
            for (int32_t g = m.n_real; g < m.n_cells; ++g) {
                const int32_t d = g0 + (g - m.n_real);
                cell_dev[st0 + g] = d;
                s->inv[d] = (int32_t)(st0 + g);
            }
            s->i_dev0[i] = m.n_cells > 0 ? cell_dev[st0] : 0;
            edge_base[i] = eb;
            r0 += m.n_real;
            g0 += m.n_cells - m.n_real;
            st0 += m.n_cells;
            eb += m.n_edges();
        }
        // real tiles, instance by instance, then ghost tiles
        std::vector<int32_t> sizes;
        for (int32_t i = 0; i < n; ++i) {
            s->i_tiles[i].x = (int32_t)tile_inst.size();
            if (leaves[i].empty())
                for (int32_t k = 0; k < s->i_nreal[i]; k += tile) leaves[i].push_back(std::min(tile, s->i_nreal[i] - k));
            int32_t k = 0;
            for (const int32_t sz : leaves[i]) {
                tile_c0.push_back(s->i_real0[i] + k);
                tile_inst.push_back(i);
                sizes.push_back(sz);
                k += sz;
            }
            if (k != s->i_nreal[i]) { set_error("internal: tile sizes"); return FVH_E_UNSUPPORTED; }
            s->i_tiles[i].y = (int32_t)tile_inst.size();
        }
        s->n_rt = (int32_t)tile_inst.size();
        for (int32_t i = 0; i < n; ++i) {
            s->i_tiles[i].z = (int32_t)tile_inst.size();
            for (int32_t k = 0; k < s->i_nghost[i]; k += kBlock) {
                tile_c0.push_back(s->i_ghost0[i] + k);
                tile_inst.push_back(i);
                sizes.push_back(std::min(kBlock, s->i_nghost[i] - k));
            }
            s->i_tiles[i].w = (int32_t)tile_inst.size();
        }
        s->n_all = (int32_t)tile_inst.size();
        // tile t spans [c0[t], c0[t+1]): the next tile starts where this one
        // ends because instances are packed back to back within the real and
        // the ghost ranges
        std::vector<int32_t> ends(s->n_all);
        for (int32_t t = 0; t < s->n_all; ++t) ends[t] = tile_c0[t] + sizes[t];
        for (int32_t t = 0; t + 1 < s->n_all; ++t)
            if (ends[t] != tile_c0[t + 1]) { set_error("internal: tile layout"); return FVH_E_UNSUPPORTED; }
        tile_c0.push_back(s->n_all ? ends[s->n_all - 1] : 0);
    }
    // half-edges in CSR by device cell, each list sorted by (reference edge, side)
    std::vector<int32_t> off(N + 1, 0);
    for (int32_t i = 0; i < n; ++i) {
        const fvh_mesh& m = *ms[i];
        const int64_t st0 = s->i_state0[i];
        for (int32_t e = 0; e < m.n_edges(); ++e)
            for (int sd = 0; sd < 2; ++sd) off[cell_dev[st0 + m.edge_cells[2 * (int64_t)e + sd]] + 1] += 1;
    }
    for (int32_t d = 0; d < N; ++d) off[d + 1] += off[d];
    s->n_he = off[N];
    for (int32_t d = 0; d < s->n_real; ++d) s->max_real_deg = std::max(s->max_real_deg, off[d + 1] - off[d]);
    std::vector<int32_t> cursor(off.begin(), off.end() - 1);
    std::vector<int32_t> he_nbr(s->n_he), he_twin(s->n_he), he_bc(s->n_he, -1), he_ref(s->n_he);
    std::vector<uint32_t> he_side(s->n_he);
    s->n_bnd = 0;
    for (int32_t i = 0; i < n; ++i) {
        const fvh_mesh& m = *ms[i];
        const int64_t st0 = s->i_state0[i];
        std::vector<int32_t> bc_of_edge(m.n_edges(), -1);
        for (size_t b = 0; b < m.bnd_edge.size(); ++b)
            bc_of_edge[m.bnd_edge[b]] = i * kMaxBc + slot_of_name[i][m.bnd_tag[b]];
        s->n_bnd += (int32_t)m.bnd_edge.size();
        for (int32_t e = 0; e < m.n_edges(); ++e) {
            const int32_t d0 = cell_dev[st0 + m.edge_cells[2 * (int64_t)e]];
            const int32_t d1 = cell_dev[st0 + m.edge_cells[2 * (int64_t)e + 1]];
            const int32_t h0 = cursor[d0]++;
            const int32_t h1 = cursor[d1]++;
            const int32_t ge = edge_base[i] + e;
            he_nbr[h0] = d1; he_nbr[h1] = d0;
            he_twin[h0] = h1; he_twin[h1] = h0;
            he_ref[h0] = ge; he_ref[h1] = ge;
            he_side[h0] = 0; he_side[h1] = 1;
            if (bc_of_edge[e] >= 0) he_bc[h0] = bc_of_edge[e];
        }
    }
    // device edge numbering: first use in half-edge order; host geometry in device order
    HostGeo hg;
    hg.e_nx.resize(E); hg.e_ny.resize(E); hg.e_len.resize(E); hg.e_gg.resize(E); hg.e_cx.resize(E); hg.e_cy.resize(E);
    hg.c_area.resize(N); hg.c_cx.resize(N); hg.c_cy.resize(N);
    std::vector<int32_t> edge_inst(E);
    for (int32_t i = 0; i < n; ++i)
        for (int32_t e = 0; e < ms[i]->n_edges(); ++e) edge_inst[edge_base[i] + e] = i;
    std::vector<int32_t> dev_of_edge(E, -1);
    int32_t n_dev_edges = 0;
    std::vector<uint32_t> he_edge(s->n_he);
    for (int32_t h = 0; h < s->n_he; ++h) {
        const int32_t ge = he_ref[h];
        if (dev_of_edge[ge] < 0) {
            const int32_t d = n_dev_edges++;
            dev_of_edge[ge] = d;
            const int32_t i = edge_inst[ge];
            const fvh_mesh& m = *ms[i];
            const int32_t e = ge - edge_base[i];
            hg.e_nx[d] = m.e_nx[e];
            hg.e_ny[d] = m.e_ny[e];
            hg.e_len[d] = m.e_len[e];
            hg.e_cx[d] = m.e_cx[e];
            hg.e_cy[d] = m.e_cy[e];
            // Green-Gauss weight, as calc_gradients computes it (explicit.cpp:132-140)
            const int32_t ci = m.edge_cells[2 * (int64_t)e], cj = m.edge_cells[2 * (int64_t)e + 1];
            const double dxif = m.e_cx[e] - m.c_cx[ci];
            const double dyif = m.e_cy[e] - m.c_cy[ci];
            const double dif = std::sqrt(dxif * dxif + dyif * dyif);
            const double dxij = m.c_cx[ci] - m.c_cx[cj];
            const double dyij = m.c_cy[ci] - m.c_cy[cj];
            const double dij = std::sqrt(dxij * dxij + dyij * dyij);
            hg.e_gg[d] = dif / dij;
            s->h_edge_ref.push_back(e);
        }
        uint32_t w = (uint32_t)dev_of_edge[ge];
        if (he_side[h]) w |= kSideBit;
        if (he_bc[h] >= 0) w |= kBcBit;
        he_edge[h] = w;
    }
    std::vector<int4> ep(E, make_int4(-1, -1, -1, -1));
    {
        for (int32_t d = 0; d < N; ++d)
            for (int32_t h = off[d]; h < off[d + 1]; ++h) {
                const uint32_t ue = he_edge[h];
                int4& x = ep[ue & kEdgeMask];
                if (ue & kSideBit) { x.y = h; x.w = d; } else { x.x = h; x.z = d; }
            }
        const fvh_mesh& m0 = *ms[0];
        for (size_t b = 0; b < m0.bnd_edge.size(); ++b) {
            const int32_t e = m0.bnd_edge[b];
            s->b_nx.push_back(m0.e_nx[e]);
            s->b_ny.push_back(m0.e_ny[e]);
            s->b_len.push_back(m0.e_len[e]);
        }
    }
    for (int32_t i = 0; i < n; ++i) {
        const fvh_mesh& m = *ms[i];
        const int64_t st0 = s->i_state0[i];
        for (int32_t r = 0; r < m.n_cells; ++r) {
            const int32_t d = cell_dev[st0 + r];
            hg.c_area[d] = m.c_area[r];
            hg.c_cx[d] = m.c_cx[r];
            hg.c_cy[d] = m.c_cy[r];
        }
    }
    // tiles for the tiled flux pass
    std::vector<uint16_t> he_slot(s->n_he, 0);
    std::vector<int32_t> t_off, t_edge, t_he, t_cell;
    s->use_tiles = false;
    if (c.flux_mode != FVH_FLUX_MODE_CELL && s->n_real > 0)
        s->use_tiles = build_tiles(s.get(), tile_c0, off, he_edge, he_slot, t_off, t_edge, t_he, t_cell);
    if ((c.flux_mode == FVH_FLUX_MODE_TILE || c.flux_mode == FVH_FLUX_MODE_FUSED) && !s->use_tiles &&
        s->n_real > 0) {
        set_error("a tile touches more than 1024 edges; use the per-cell flux mode");
        return FVH_E_UNSUPPORTED;
    }
    // fused stage pass: linear reconstruction only (otherwise there is
    // nothing to fuse), and only when the largest tile halo fits in LDS
    HaloHost hh;
    const bool lin_on = (c.flags & FVH_FLAG_LINEAR) != 0;
    s->fused = false;
    // residual smoothing needs every cell's qt before any update: not fusable;
    // the viscous flux and time steps exist in the two-pass forms only
    const bool fusable = !(c.flags & FVH_FLAG_SMOOTH) && c.flux != FVH_FLUX_ROE_VISCOUS && c.dt_kind <= FVH_DT_EDGE_MIN;
    if ((c.flux_mode == FVH_FLUX_MODE_FUSED || c.flux_mode == FVH_FLUX_MODE_AUTO) && s->use_tiles && s->second &&
        lin_on && fusable && build_halo(s.get(), tile_c0, off, he_nbr, t_off, t_he, hh)) {
        s->fused = halo_lds_bytes(hh) <= 160 * 1024;
    }
    // Automatic tile size (tile_cells 0): tiles of 256 cells first.  When an
    // average tile with its real ring-1 cells outnumbers 256 lanes, the
    // gradient phase of a 256-lane workgroup takes two rounds; tiles shrunk
    // until it fits one round do the same work in fewer instructions per
    // lane.  They are rebuilt when the fused pass runs 256-lane workgroups,
    // or when it runs 512 and the smaller halos would let three 256-lane
    // workgroups share a CU (the 975,744-cell step mesh: 208 cells per tile,
    // 12% faster; the 2 M-cell airfoil: 208 cells, 256 lanes, 5% faster).
    if (c.tile_cells == 0 && s->fused && c.renumber == FVH_ORDER_TILES && s->n_real > 0 && s->n_rt > 0 &&
        c.stage_threads != kStageBlock) {
        const size_t lds = halo_lds_bytes(hh);
        const int32_t mean_g = (int32_t)((hh.sum_g + s->n_rt - 1) / s->n_rt);
        const int half = kStageBlock / 2;
        const int lanes = stage_lanes(lds, c.stage_threads, mean_g);
        const double per_cell = (double)hh.sum_g / (double)s->n_real;
        const int32_t t = std::max<int32_t>(128, (int32_t)(half / per_cell) / 16 * 16);
        // halo bytes scale slightly slower than the cell count: 5% margin
        const bool three_fit = (double)lds * t / kBlock * 1.05 + 1024 <= 160.0 * 1024 / 3;
        if (t < kBlock && mean_g > half && (lanes == half || (c.stage_threads == 0 && three_fit))) {
            // the layout so far is host data only (the stream and the
            // device arrays come after these decisions)
            std::vector<fvh_config> shrunk(cs, cs + n);
            for (fvh_config& x : shrunk) x.tile_cells = t;
            return create(n, ms, shrunk.data(), out, nullptr, true);
        }
    }
    // A shrink is kept only if it gave what it was made for: 256-lane
    // workgroups whose gradient phase takes one round; otherwise the tiles
    // of 256 cells are rebuilt.
    if (auto_shrunk && s->fused && s->n_rt > 0) {
        const int32_t mean_g = (int32_t)((hh.sum_g + s->n_rt - 1) / s->n_rt);
        if (mean_g > kStageBlock / 2 || stage_lanes(halo_lds_bytes(hh), c.stage_threads, mean_g) != kStageBlock / 2) {
            std::vector<fvh_config> full(cs, cs + n);
            for (fvh_config& x : full) x.tile_cells = kBlock;
            return create(n, ms, full.data(), out);
        }
    }
    // Oversized tiles: when a few tiles' halos set an LDS size that lets one
    // 256-lane workgroup fewer share a CU than the others would need, those
    // tiles are bisected once more (the 2 M-cell airfoil: 74 of 9,616 tiles
    // above 40 KB, up to 51 KB, so three workgroups per CU instead of four).
    if (s->fused && c.renumber == FVH_ORDER_TILES && !splits && s->n_rt > 0) {
        const size_t lds = halo_lds_bytes(hh);
        const int32_t mean_g = (int32_t)((hh.sum_g + s->n_rt - 1) / s->n_rt);
        const int per_cu = (int)((160 * 1024) / (lds + 1024));
        if (stage_lanes(lds, c.stage_threads, mean_g) == kStageBlock / 2 && per_cu < 4) {
            const size_t target = (160 * 1024) / (per_cu + 1) - 1024;
            std::vector<std::vector<uint8_t>> mark(n);
            int32_t over = 0;
            for (int32_t t = 0; t < s->n_rt; ++t) {
                const int32_t i = tile_inst[t];
                const int4 x = hh.ext[t];
                const int32_t n_ext = (tile_c0[t + 1] - tile_c0[t]) + x.z + x.w;
                if (halo_lds_bytes(s.get(), n_ext, t_off[t + 1] - t_off[t]) <= target) continue;
                if (mark[i].empty()) mark[i].assign(s->i_tiles[i].y - s->i_tiles[i].x, 0);
                mark[i][t - s->i_tiles[i].x] = 1;
                ++over;
            }
            if (over > 0 && over * 100 <= 3 * s->n_rt) return create(n, ms, cs, out, &mark);
        }
    }
    if (c.flux_mode == FVH_FLUX_MODE_FUSED && !s->fused && s->n_real > 0) {
        set_error("fused stage pass unavailable: needs linear reconstruction, an inviscid flux and time step, no "
                  "residual smoothing and a tile halo that fits in LDS");
        return FVH_E_UNSUPPORTED;
    }
    // the layout is final: device work starts here
    FVH_HIP(hipStreamCreateWithFlags(&s->stream, hipStreamNonBlocking));
    if ((rc = upload(s->epair, ep.data(), ep.size() * sizeof(int4)))) return rc;
    if (s->fused) {
        s->x_max_ext = hh.max_ext;
        s->x_max_g = hh.max_g;
        s->x_mean_g = s->n_rt ? (int32_t)((hh.sum_g + s->n_rt - 1) / s->n_rt) : 0;
        s->x_max_edges = hh.max_edges;
        s->x_lds = (int32_t)hh.max_lds;
        if ((rc = upload(s->x_ext, hh.ext.data(), hh.ext.size() * sizeof(int4)))) return rc;
        if ((rc = upload(s->x_cells, hh.cells.data(), hh.cells.size() * 4))) return rc;
        if ((rc = upload(s->x_r1_start, hh.r1_start.data(), hh.r1_start.size() * 4))) return rc;
        if ((rc = upload(s->x_r1_loc, hh.r1_loc.data(), hh.r1_loc.size() * 2))) return rc;
        if ((rc = upload(s->x_he_loc, hh.he_loc.data(), hh.he_loc.size() * 2))) return rc;
        if ((rc = upload(s->x_r1_slot, hh.r1_slot.data(), hh.r1_slot.size() * 2))) return rc;
    }
    if ((rc = upload(s->he_off, off.data(), off.size() * 4))) return rc;
    if ((rc = upload(s->he_nbr, he_nbr.data(), he_nbr.size() * 4))) return rc;
    if ((rc = upload(s->he_edge, he_edge.data(), he_edge.size() * 4))) return rc;
    if (n == 1) {
        s->h_he_edge = he_edge;
        s->mesh0.reset(new fvh_mesh(*ms[0]));
    }
    if ((rc = upload(s->he_twin, he_twin.data(), he_twin.size() * 4))) return rc;
    if ((rc = upload(s->he_bc, he_bc.data(), he_bc.size() * 4))) return rc;
    if ((rc = upload(s->he_slot, he_slot.data(), he_slot.size() * 2))) return rc;
    if ((rc = upload(s->tile_c0, tile_c0.data(), tile_c0.size() * 4))) return rc;
    if ((rc = upload(s->tile_inst, tile_inst.data(), tile_inst.size() * 4))) return rc;
    if ((rc = upload(s->inst_tiles, s->i_tiles.data(), s->i_tiles.size() * sizeof(int4)))) return rc;
    if ((rc = upload(s->inst_dev0, s->i_dev0.data(), s->i_dev0.size() * 4))) return rc;
    if (s->use_tiles) {
        if ((rc = upload(s->t_off, t_off.data(), t_off.size() * 4))) return rc;
        if ((rc = upload(s->t_edge, t_edge.data(), t_edge.size() * 4))) return rc;
        if ((rc = upload(s->t_he, t_he.data(), t_he.size() * 4))) return rc;
        if ((rc = upload(s->t_cell, t_cell.data(), t_cell.size() * 4))) return rc;
    }
    rc = FVH_DISPATCH(s, upload_geometry, s.get(), hg);
    if (rc) return rc;
    // algorithmic bytes per stage (state s, geometry g bytes)
    const double sb = (s->f32 || s->mixed) ? 4.0 : 8.0, gb = s->f32 ? 4.0 : 8.0, V = 4.0;
    const double C = s->n_real, B = s->n_bnd;
    double bytes = C * (3 * V * sb + sb + gb) + E * (3 * gb + 8) + B * (2 * V * sb + 2 * gb + 4);
    if (s->second) bytes += C * (7 * V * sb + 2 * gb) + E * (2 * gb);
    s->stage_bytes = bytes;
    *out = s.release();
    return FVH_OK;
}

template <class T, class S>
int32_t sums_t(fvh_solver* s, double* out) {
    launch_sums<T>(s->stream, view<T>(s), tile_map(s), inst_view(s), (const Q4<S>*)s->q[s->cur].p,
                   (double*)s->sums_blk.p, (double*)s->sums.p);
    FVH_HIP(hipGetLastError());
    FVH_HIP(hipStreamSynchronize(s->stream));
    FVH_HIP(hipMemcpy(out, s->sums.p, (size_t)s->n_inst * 32, hipMemcpyDeviceToHost));
    return FVH_OK;
}

// ---------------------------------------------------------------------------
// discrete adjoint driver (FP64, single instance)
// ---------------------------------------------------------------------------
// What the reverse of one step reads: the state entering every stage and
// the step's output, each stage's gradient, limiter and qt, the cell time
// steps and the global one.  A kept record (`light`) holds only the states,
// qt and the time steps; its gradients and limiters point to the work
// buffers and are recomputed from the stage inputs by the reverse (the
// gradient pass writes the bits the fused stage pass computes).
struct StepRecord {
    void* qs[kAdjMaxStages + 1];
    void* gx[kAdjMaxStages];
    void* gy[kAdjMaxStages];
    void* lm[kAdjMaxStages];
    void* qt[kAdjMaxStages];
    void* dt;
    void* dtg;
    bool light;
};

struct Adjoint {
    fvh_solver* s;
    int n_stage = 0;
    size_t qbytes = 0;
    int64_t fwd_steps = 0;
    GeoAdj<double> ga;   // geometry adjoint (he == nullptr: not formed)
    StepRecord work{};   // the replay buffers
    std::vector<StepRecord> tapes;   // records kept by the forward sweep

    // bytes of one kept step record: the stage states, qt, the time steps
    size_t record_bytes() const {
        return (size_t)(n_stage + 1 + n_stage) * qbytes + (size_t)s->n_total * 8 + 256;
    }
    StepRecord work_record() const {
        StepRecord r{};
        for (int k = 0; k <= n_stage; ++k) r.qs[k] = s->a_qs[k].p;
        for (int k = 0; k < n_stage; ++k) {
            r.gx[k] = s->a_gx[k].p; r.gy[k] = s->a_gy[k].p; r.lm[k] = s->a_lm[k].p; r.qt[k] = s->a_qt[k].p;
        }
        r.dt = s->dt.p;
        r.dtg = s->dt_global.p;
        return r;
    }
    // n records carved out of one allocation each
    int32_t alloc_tapes(int n) {
        tapes.clear();
        const size_t q = (qbytes + 255) / 256 * 256, d = ((size_t)s->n_total * 8 + 255) / 256 * 256;
        for (int i = 0; i < n; ++i) {
            const size_t bytes = record_bytes() + (size_t)(2 * n_stage + 1) * 256 + d;
            if ((int)s->a_tape.size() <= i || s->a_tape[i]->bytes < bytes) {
                std::unique_ptr<DevBuf> b(new DevBuf());
                if (b->alloc(bytes)) return oom();
                if ((int)s->a_tape.size() <= i) s->a_tape.push_back(std::move(b));
                else s->a_tape[i] = std::move(b);
            }
            char* p = (char*)s->a_tape[i]->p;
            StepRecord r = work_record();
            r.light = true;
            for (int k = 0; k <= n_stage; ++k) { r.qs[k] = p; p += q; }
            for (int k = 0; k < n_stage; ++k) { r.qt[k] = p; p += q; }
            r.dt = p; p += d;
            r.dtg = p;
            tapes.push_back(r);
        }
        return FVH_OK;
    }

    // zeroed geometry-adjoint accumulators
    int32_t alloc_geo() {
        const size_t nh = (size_t)std::max(s->n_he, 1);
        if ((!s->a_g_he.p && s->a_g_he.alloc(nh * sizeof(Q4<double>))) ||
            (!s->a_g_hc.p && s->a_g_hc.alloc((nh + 1) / 2 * sizeof(Q4<double>))) ||
            (!s->a_g_e.p && s->a_g_e.alloc((size_t)std::max(s->n_edges, 1) * sizeof(Q4<double>))) ||
            (!s->a_g_c.p && s->a_g_c.alloc((size_t)std::max(s->n_total, 1) * sizeof(Q4<double>))) ||
            (!s->a_obj_p.p && s->a_obj_p.alloc((size_t)(s->n_bnd + 1) * 8)))
            return oom();
        ga.he = (Q4<double>*)s->a_g_he.p;
        ga.hc = (G2<double>*)s->a_g_hc.p;
        ga.e = (Q4<double>*)s->a_g_e.p;
        ga.c = (Q4<double>*)s->a_g_c.p;
        launch_fill<double>(s->stream, ga.he, s->n_he);
        launch_fill<double>(s->stream, (Q4<double>*)ga.hc, (s->n_he + 1) / 2);   // G2 pairs
        launch_fill<double>(s->stream, ga.e, s->n_edges);
        launch_fill<double>(s->stream, ga.c, s->n_total);
        FVH_HIP(hipGetLastError());
        return FVH_OK;
    }

    int32_t alloc(int snaps) {
        const size_t N = (size_t)s->n_total;
        qbytes = N * sizeof(Q4<double>);
        n_stage = s->cfg.n_alpha;
        for (int k = 0; k <= n_stage; ++k)
            if (!s->a_qs[k].p && s->a_qs[k].alloc(qbytes)) return oom();
        for (int k = 0; k < n_stage; ++k)
            if ((!s->a_gx[k].p && s->a_gx[k].alloc(qbytes)) || (!s->a_gy[k].p && s->a_gy[k].alloc(qbytes)) ||
                (!s->a_lm[k].p && s->a_lm[k].alloc(qbytes)) || (!s->a_qt[k].p && s->a_qt[k].alloc(qbytes)))
                return oom();
        const size_t he = (size_t)std::max(s->n_he, 1) * sizeof(Q4<double>);
        if ((!s->a_lam.p && s->a_lam.alloc(qbytes)) ||
            (!s->a_lam_q0.p && s->a_lam_q0.alloc(qbytes)) || (!s->a_lam_qt.p && s->a_lam_qt.alloc(qbytes)) ||
            (!s->a_lam_bincr.p && s->a_lam_bincr.alloc((size_t)(s->n_bnd + 1) * sizeof(Q4<double>))) ||
            (!s->a_lam_bq.p && s->a_lam_bq.alloc((size_t)(s->n_total - s->n_real + 1) * sizeof(Q4<double>))) ||
            (!s->a_lam_face.p && s->a_lam_face.alloc(he)) || (!s->a_nb.p && s->a_nb.alloc(he)) ||
            (!s->a_lam_direct.p && s->a_lam_direct.alloc(qbytes)) ||
            (!s->a_ld.p && s->a_ld.alloc((size_t)(kAdjMaxStages * (s->n_rt + 1)) * 8)) ||
            (!s->a_tmp.p && s->a_tmp.alloc(qbytes)) ||
            (!s->a_amin_v.p && s->a_amin_v.alloc(kArgminBlocks * 8)) ||
            (!s->a_amin_r.p && s->a_amin_r.alloc(kArgminBlocks * 4)) ||
            (!s->a_amin_d.p && s->a_amin_d.alloc(kArgminBlocks * 4)) || (!s->a_amin.p && s->a_amin.alloc(8)))
            return oom();
        if (!s->a_ref.p && upload(s->a_ref, s->inv.data(), s->inv.size() * 4)) return oom();
        while ((int)s->a_ck.size() < snaps) {
            std::unique_ptr<DevBuf> b(new DevBuf());
            if (b->alloc(qbytes)) return oom();
            s->a_ck.push_back(std::move(b));
        }
        return FVH_OK;
    }
    int32_t oom() {
        set_error("hipMalloc failed (adjoint buffers)");
        return FVH_E_HIP;
    }
    double work_bytes() const {
        return (double)qbytes * (n_stage + 1 + 4 * n_stage + 6) +
               2.0 * (double)std::max(s->n_he, 1) * sizeof(Q4<double>);
    }

    // k forward steps from state src into state dst (the fast path)
    int32_t advance(const void* src, void* dst, int k) {
        launch_copy<double>(s->stream, src, s->q[s->cur].p, s->n_total);
        fvh_run_options o{1e300, 0xffffffffu, 1, 0.0, -1, 0};
        for (int i = 0; i < k; ++i) {
            int32_t rc = fvh::step<double, double>(s, false, &o, 0);
            if (rc) return rc;
        }
        fwd_steps += k;
        launch_copy<double>(s->stream, s->q[s->cur].p, dst, s->n_total);
        return FVH_OK;
    }

    // one step from R.qs[0], keeping every stage's input, gradient, limiter and qt
    int32_t replay_body(const StepRecord& R) {
        const fvh_config& c = s->cfg;
        const MeshView<double> m = view<double>(s);
        const TileMap tm = tile_map(s);
        const InstView iv = inst_view(s);
        const TileView tv = tile_view(s);
        const PhysParams pp = phys(c);
        const BcParam<double>* bcs = (const BcParam<double>*)s->bcs.p;
        const bool grad = c.flags & FVH_FLAG_GRADIENTS, lim = c.flags & FVH_FLAG_LIMITERS,
                   lin = c.flags & FVH_FLAG_LINEAR, gdt = c.flags & FVH_FLAG_GLOBAL_DT;
        hipStream_t st = s->stream;
        double* dt = (double*)R.dt;
        double* blk = (double*)s->blk_min.p;
        const bool kk = kinks(c);
        // the time-step kernel fills the record's own dt (a fixed step too:
        // the gradient pass computes only the CFL kinds)
        if (!s->second || s->fused || c.dt_kind == FVH_DT_FIXED || kk)
            launch_dt<double>(st, m, tm, kdt(c), pp, (const Q4<double>*)R.qs[0], dt, blk, nullptr);
        if (s->fused) {
            // the fused stage pass, storing what the two-pass form would (every
            // cell's time step comes from the time-step kernel above)
            const HaloView hv = halo_view(s);
            if (gdt) global_dt(s, st, tm, iv, blk, dt, (double*)R.dtg, nullptr);
            for (int k = 0; k < n_stage; ++k) {
                StageRecord<double> rec;
                if (!R.light) {
                    rec.gx = (Q4<double>*)R.gx[k];
                    rec.gy = (Q4<double>*)R.gy[k];
                    rec.lm = (Q4<double>*)R.lm[k];
                }
                rec.qt = (Q4<double>*)R.qt[k];
                if (!launch_stage<double>(st, m, tm, tv, hv, pp, kphi(c), c.flux == FVH_FLUX_ROE_JOINED, grad, lim, -1,
                                          bcs, (const Q4<double>*)R.qs[k], (const Q4<double>*)R.qs[0],
                                          (Q4<double>*)R.qs[k + 1], (Q4<double>*)s->bincr.p, dt,
                                          gdt ? (const double*)R.dtg : nullptr, c.alpha[k], nullptr, nullptr,
                                          c.stage_threads, rec, stage_queue(s))) {
                    set_error("fused stage pass: tile halo does not fit LDS");
                    return FVH_E_UNSUPPORTED;
                }
            }
            FVH_HIP(hipGetLastError());
            return FVH_OK;
        }
        for (int k = 0; k < n_stage; ++k) {
            const Q4<double>* qin = (const Q4<double>*)R.qs[k];
            if (s->second)
                launch_gradlim<double>(st, m, tm, pp, kphi(c), grad, lim, lin, qin, (Q4<double>*)s->face.p,
                                       (Q4<double>*)s->bincr.p, (k == 0 && !kk) ? c.dt_kind : -1, dt, blk,
                                       R.light ? nullptr : (Q4<double>*)R.gx[k],
                                       R.light ? nullptr : (Q4<double>*)R.gy[k],
                                       R.light ? nullptr : (Q4<double>*)R.lm[k], nullptr);
            if (k == 0 && gdt) global_dt(s, st, tm, iv, blk, dt, (double*)R.dtg, nullptr);
            launch_flux<double>(st, m, tm, s->use_tiles ? &tv : nullptr, pp, kflux(c), lin, bcs,
                                qin, (const Q4<double>*)R.qs[0], (Q4<double>*)R.qs[k + 1],
                                (const Q4<double>*)s->face.p, (const Q4<double>*)s->bincr.p, dt,
                                gdt ? (const double*)R.dtg : nullptr, c.alpha[k], nullptr, (Q4<double>*)R.qt[k],
                                nullptr);
        }
        FVH_HIP(hipGetLastError());
        return FVH_OK;
    }

    // reverse of a recorded step: a_lam (adjoint of the step's output) ->
    // a_lam (adjoint of its input)
    int32_t reverse(const StepRecord& R) {
        const fvh_config& c = s->cfg;
        const MeshView<double> m = view<double>(s);
        const TileMap tm = tile_map(s);
        const PhysParams pp = phys(c);
        const bool grad = c.flags & FVH_FLAG_GRADIENTS, lim = c.flags & FVH_FLAG_LIMITERS,
                   lin = c.flags & FVH_FLAG_LINEAR, gdt = c.flags & FVH_FLAG_GLOBAL_DT;
        hipStream_t st = s->stream;
        Q4<double>* lam = (Q4<double>*)s->a_lam.p;
        Q4<double>* lam_q0 = (Q4<double>*)s->a_lam_q0.p;
        launch_fill<double>(st, lam_q0, s->n_total);
        for (int k = n_stage - 1; k >= 0; --k) {
            if (R.light && s->second)
                // the stage's gradients and limiters from its input, as the
                // forward computed them (cells and ghosts; no face states)
                launch_gradlim<double>(st, m, tm, pp, kphi(c), grad, lim, false, (const Q4<double>*)R.qs[k], nullptr,
                                       nullptr, -1, nullptr, nullptr, (Q4<double>*)R.gx[k], (Q4<double>*)R.gy[k],
                                       (Q4<double>*)R.lm[k], nullptr);
            launch_adj_stage<double>(st, m, tm, pp, kphi(c), grad, lim, lin, (const BcParam<double>*)s->bcs.p,
                                     s->n_edges, (const int4*)s->epair.p, (const Q4<double>*)R.qs[k],
                                     (const Q4<double>*)R.qs[k + 1], (const Q4<double>*)R.gx[k],
                                     (const Q4<double>*)R.gy[k], (const Q4<double>*)R.lm[k],
                                     (const Q4<double>*)R.qt[k], (const double*)R.dt,
                                     gdt ? (const double*)R.dtg : nullptr, c.alpha[k],
                                     k == n_stage - 1 ? lam : nullptr, lam_q0, (Q4<double>*)s->a_lam_qt.p,
                                     (Q4<double>*)s->a_lam_bq.p, (Q4<double>*)s->a_lam_bincr.p,
                                     (Q4<double>*)s->a_lam_face.p, (Q4<double>*)s->a_lam_direct.p,
                                     (Q4<double>*)s->a_nb.p, k == 0 ? lam_q0 : nullptr, k == 0 ? 1 : 0, nullptr,
                                     (double*)s->a_ld.p + (size_t)k * s->n_rt, ga);
        }
        if (gdt && c.dt_kind != FVH_DT_FIXED && kinks(c)) {
            // the p-norm step: every cell's step reversed (a_nb is free once
            // the stage reverses above have been gathered)
            launch_adj_dt_soft<double>(st, m, pp, (const Q4<double>*)R.qs[0], (const double*)R.dt,
                                       (const double*)R.dtg, (const double*)s->a_ld.p, n_stage * s->n_rt,
                                       (double*)s->a_amin_v.p, lam_q0, (Q4<double>*)s->a_nb.p, ga);
        } else if (gdt && c.dt_kind != FVH_DT_FIXED) {
            // the cell whose step is the minimum, first in reference order (min_dt, explicit.cpp:555-564)
            launch_dt_argmin<double>(st, (const double*)R.dt, (const int32_t*)s->a_ref.p, s->n_total,
                                     (double*)s->a_amin_v.p, (int32_t*)s->a_amin_r.p, (int32_t*)s->a_amin_d.p,
                                     (int32_t*)s->a_amin.p);
            launch_adj_dt<double>(st, m, pp, c.dt_kind, (const Q4<double>*)R.qs[0], (const int32_t*)s->a_amin.p,
                                  (const double*)s->a_ld.p, n_stage * s->n_rt, lam_q0, ga);
        }
        FVH_HIP(hipGetLastError());
        launch_copy<double>(st, s->a_lam_q0.p, s->a_lam.p, s->n_total);
        return FVH_OK;
    }

    // replay of the step starting from state src, then its reverse; the
    // launch sequence never changes, so it is captured once as a HIP graph
    int32_t step_back(const void* src) {
        launch_copy<double>(s->stream, src, s->a_qs[0].p, s->n_total);
        if (s->a_graph_exec && s->a_graph_geo != (ga.he != nullptr)) {
            // captured with the other geometry choice
            FVH_HIP(hipStreamSynchronize(s->stream));
            (void)hipGraphExecDestroy(s->a_graph_exec);
            (void)hipGraphDestroy(s->a_graph);
            s->a_graph_exec = nullptr;
            s->a_graph = nullptr;
        }
        if (!s->a_graph_exec) {
            FVH_HIP(hipStreamBeginCapture(s->stream, hipStreamCaptureModeThreadLocal));
            work = work_record();
            int32_t rc = replay_body(work);
            if (!rc) rc = reverse(work);
            hipGraph_t g = nullptr;
            const hipError_t e = hipStreamEndCapture(s->stream, &g);
            if (rc) { if (g) (void)hipGraphDestroy(g); return rc; }
            if (e != hipSuccess) {
                set_error(std::string("hipStreamEndCapture: ") + hipGetErrorString(e));
                return FVH_E_HIP;
            }
            s->a_graph = g;
            s->a_graph_geo = ga.he != nullptr;
            FVH_HIP(hipGraphInstantiate(&s->a_graph_exec, g, nullptr, nullptr, 0));
        }
        FVH_HIP(hipGraphLaunch(s->a_graph_exec, s->stream));
        fwd_steps += 1;
        return FVH_OK;
    }

    // binomial checkpointing: the state of step a is in slot sa; a_lam holds
    // the adjoint of state b; on return it holds the adjoint of state a
    static double beta(int sn, int r) {   // C(sn + r, sn)
        double v = 1.0;
        for (int i = 1; i <= sn; ++i) v = v * (double)(r + i) / (double)i;
        return v;
    }
    int32_t rev(int a, int b, int sa, std::vector<int>& free_slots) {
        const int n = b - a;
        int32_t rc;
        if (n <= 0) return FVH_OK;
        if (n == 1) return step_back(s->a_ck[sa]->p);
        if (free_slots.empty()) {
            for (int k = b - 1; k >= a; --k) {
                if ((rc = advance(s->a_ck[sa]->p, s->a_tmp.p, k - a))) return rc;
                if ((rc = step_back(s->a_tmp.p))) return rc;
            }
            return FVH_OK;
        }
        const int sn = (int)free_slots.size();
        int r = 0;
        while (beta(sn, r) < (double)n) ++r;
        int mstep = n - (int)beta(sn - 1, r);
        mstep = std::max(1, std::min(n - 1, mstep));
        const int t = free_slots.back();
        free_slots.pop_back();
        if ((rc = advance(s->a_ck[sa]->p, s->a_ck[t]->p, mstep))) return rc;
        if ((rc = rev(a + mstep, b, t, free_slots))) return rc;
        free_slots.push_back(t);
        return rev(a, a + mstep, sa, free_slots);
    }
};

// Boundary-element data of the objective, uploaded per call.
int32_t objective_setup(fvh_solver* s, const double* wx, const double* wy) {
    const int32_t nb = s->i_nghost.empty() ? 0 : s->i_nghost[0];
    std::vector<int32_t> ghost(nb + 1, 0);
    std::vector<Q4<double>> gw(nb + 1);
    for (int32_t b = 0; b < nb; ++b) {
        ghost[b] = s->i_ghost0[0] + b;
        gw[b].v[0] = (wx ? wx[b] : 0.0) * s->b_len[b];
        gw[b].v[1] = (wy ? wy[b] : 0.0) * s->b_len[b];
        gw[b].v[2] = s->b_nx[b];
        gw[b].v[3] = s->b_ny[b];
    }
    int32_t rc;
    if ((rc = upload(s->a_obj_ghost, ghost.data(), ghost.size() * 4))) return rc;
    if ((rc = upload(s->a_obj_gw, gw.data(), gw.size() * sizeof(Q4<double>)))) return rc;
    if (!s->a_obj_terms.p && s->a_obj_terms.alloc((size_t)(nb + 1) * 8)) { set_error("hipMalloc failed"); return FVH_E_HIP; }
    return FVH_OK;
}

int32_t objective_eval(fvh_solver* s, const void* q, bool seed, double* J, double* p_out = nullptr) {
    const int32_t nb = s->i_nghost[0];
    if (seed) launch_fill<double>(s->stream, (Q4<double>*)s->a_lam.p, s->n_total);
    launch_objective<double>(s->stream, nb, phys(s->cfg), (const int32_t*)s->a_obj_ghost.p,
                             (const Q4<double>*)s->a_obj_gw.p, (const Q4<double>*)q, (double*)s->a_obj_terms.p,
                             seed ? (Q4<double>*)s->a_lam.p : nullptr, p_out);
    FVH_HIP(hipGetLastError());
    std::vector<double> t(nb);
    FVH_HIP(hipStreamSynchronize(s->stream));
    if (nb) FVH_HIP(hipMemcpy(t.data(), s->a_obj_terms.p, (size_t)nb * 8, hipMemcpyDeviceToHost));
    double sum = 0.0;
    for (int32_t b = 0; b < nb; ++b) sum += t[b];
    *J = sum;
    return FVH_OK;
}

// Geometry adjoint accumulated on the device (per half-edge, edge, cell) and
// the objective's own dependence on the boundary normals and lengths, summed
// per reference edge and cell, then chained to the node coordinates.
int32_t node_gradient(fvh_solver* s, const double* wx, const double* wy, double* gxy) {
    const fvh_mesh& m = *s->mesh0;
    const int32_t nh = s->n_he, ne = s->n_edges, nc = s->n_total, nb = (int32_t)m.bnd_edge.size();
    std::vector<Q4<double>> he(nh), e(ne), c(nc);
    std::vector<G2<double>> hc(nh);
    std::vector<double> p(nb);
    FVH_HIP(hipStreamSynchronize(s->stream));
    FVH_HIP(hipMemcpy(he.data(), s->a_g_he.p, (size_t)nh * sizeof(Q4<double>), hipMemcpyDeviceToHost));
    FVH_HIP(hipMemcpy(hc.data(), s->a_g_hc.p, (size_t)nh * sizeof(G2<double>), hipMemcpyDeviceToHost));
    FVH_HIP(hipMemcpy(e.data(), s->a_g_e.p, (size_t)ne * sizeof(Q4<double>), hipMemcpyDeviceToHost));
    FVH_HIP(hipMemcpy(c.data(), s->a_g_c.p, (size_t)nc * sizeof(Q4<double>), hipMemcpyDeviceToHost));
    if (nb) FVH_HIP(hipMemcpy(p.data(), s->a_obj_p.p, (size_t)nb * 8, hipMemcpyDeviceToHost));
    GeoBar b;
    b.resize(m.n_edges(), m.n_cells);
    for (int32_t h = 0; h < nh; ++h) {
        const int32_t r = s->h_edge_ref[s->h_he_edge[h] & kEdgeMask];
        b.e_nx[r] += he[h].v[0];
        b.e_ny[r] += he[h].v[1];
        b.e_len[r] += he[h].v[2];
        b.e_w[r] += he[h].v[3];
        b.e_cx[r] += hc[h].v[0];
        b.e_cy[r] += hc[h].v[1];
    }
    for (int32_t d = 0; d < ne; ++d) {
        const int32_t r = s->h_edge_ref[d];
        b.e_nx[r] += e[d].v[0];
        b.e_ny[r] += e[d].v[1];
        b.e_len[r] += e[d].v[2];
    }
    for (int32_t d = 0; d < nc; ++d) {
        const int32_t r = s->inv[d];
        b.c_area[r] += c[d].v[0];
        b.c_cx[r] += c[d].v[1];
        b.c_cy[r] += c[d].v[2];
    }
    // J = sum_b p_b (wx_b n_x + wy_b n_y) len_b
    for (int32_t k = 0; k < nb; ++k) {
        const int32_t r = m.bnd_edge[k];
        const double ax = wx ? wx[k] : 0.0, ay = wy ? wy[k] : 0.0;
        b.e_nx[r] += p[k] * ax * m.e_len[r];
        b.e_ny[r] += p[k] * ay * m.e_len[r];
        b.e_len[r] += p[k] * (ax * m.e_nx[r] + ay * m.e_ny[r]);
    }
    std::fill(gxy, gxy + 2 * m.x.size(), 0.0);
    mesh_geometry_vjp(m, b, gxy);
    return FVH_OK;
}

int32_t adjoint_check(fvh_solver* s) {
    if (s->f32 || s->mixed) { set_error("the adjoint is FP64 only"); return FVH_E_UNSUPPORTED; }
    if (s->n_inst != 1) { set_error("the adjoint takes a single-mesh solver"); return FVH_E_UNSUPPORTED; }
    if (s->max_real_deg > kMaxHe) {
        set_error("the adjoint takes triangles and quadrilaterals (cells of at most " + std::to_string(kMaxHe) +
                  " edges)");
        return FVH_E_UNSUPPORTED;
    }
    const fvh_config& c = s->cfg;
    if (!(c.flags & FVH_FLAG_GLOBAL_DT) && c.dt_kind != FVH_DT_FIXED) {
        set_error("the adjoint needs a global or fixed time step");
        return FVH_E_UNSUPPORTED;
    }
    if (c.flags & FVH_FLAG_SMOOTH) {
        set_error("the adjoint does not take residual smoothing");
        return FVH_E_UNSUPPORTED;
    }
    if (c.flux == FVH_FLUX_ROE_VISCOUS || c.dt_kind > FVH_DT_EDGE_MIN) {
        set_error("the adjoint takes the inviscid fluxes and time steps");
        return FVH_E_UNSUPPORTED;
    }
    for (int k = 0; k < c.n_bc; ++k)
        if (c.bc[k].kind > FVH_BC_FARFIELD) {
            set_error("the adjoint takes the copy, mirror, inlet/outlet and farfield boundaries");
            return FVH_E_UNSUPPORTED;
        }
    return FVH_OK;
}

// ---------------------------------------------------------------------------
// Scratch memory of the entry point that allocates nothing
// (fvh_min_dt_device): a file-scope device array, one copy per GPU, whose
// address each device resolves once.  Every call on a GPU shares its copy.
// ---------------------------------------------------------------------------
static __device__ __align__(256) double scratch[kArrayMinBlocks];

constexpr int kMaxDevices = 64;
std::mutex g_scratch_mu;
double* g_scratch_ptr[kMaxDevices] = {};

int32_t scratch_address(double** out) {
    int dev = 0;
    FVH_HIP(hipGetDevice(&dev));
    if (dev < 0 || dev >= kMaxDevices) { set_error("device index out of range"); return FVH_E_HIP; }
    std::lock_guard<std::mutex> lk(g_scratch_mu);
    if (!g_scratch_ptr[dev]) {
        void* p = nullptr;
        const hipError_t e = hipGetSymbolAddress(&p, HIP_SYMBOL(scratch));
        if (e != hipSuccess || !p) {
            set_error(std::string("hipGetSymbolAddress(scratch): ") + hipGetErrorString(e));
            return FVH_E_HIP;
        }
        g_scratch_ptr[dev] = (double*)p;
    }
    *out = g_scratch_ptr[dev];
    return FVH_OK;
}

// ---------------------------------------------------------------------------
// The phases one at a time on caller arrays (FP64, single instance)
// ---------------------------------------------------------------------------
int32_t phase_setup(fvh_solver* s) {
    if (s->f32 || s->mixed) { set_error("the per-phase functions take FP64 solvers"); return FVH_E_UNSUPPORTED; }
    if (s->n_inst != 1) { set_error("the per-phase functions take single-instance solvers"); return FVH_E_UNSUPPORTED; }
    if (kinks(s->cfg)) {
        set_error("the per-phase functions are the reference's phases; smoothed kinks run through run/advance");
        return FVH_E_UNSUPPORTED;
    }
    const size_t qb = (size_t)s->n_total * sizeof(Q4<double>);
    for (auto& b : s->ph)
        if (!b.p && b.alloc(qb)) { set_error("hipMalloc failed"); return FVH_E_HIP; }
    if (!s->ph_part.p && s->ph_part.alloc((size_t)kArrayMinBlocks * 8)) {
        set_error("hipMalloc failed");
        return FVH_E_HIP;
    }
    return FVH_OK;
}

// host array in the reference layout (4 per cell, all cells) -> device order
int32_t put_q4(fvh_solver* s, const double* host, void* dev) {
    std::vector<Q4<double>> h(s->n_total);
    for (int32_t d = 0; d < s->n_total; ++d) {
        const int64_t r = s->inv[d];
        for (int k = 0; k < 4; ++k) h[d].v[k] = host[4 * r + k];
    }
    FVH_HIP(hipMemcpyAsync(dev, h.data(), h.size() * sizeof(Q4<double>), hipMemcpyHostToDevice, s->stream));
    FVH_HIP(hipStreamSynchronize(s->stream));
    return FVH_OK;
}

int32_t put_flat(fvh_solver* s, const double* host, void* dev, int64_t n) {
    FVH_HIP(hipMemcpyAsync(dev, host, (size_t)n * 8, hipMemcpyHostToDevice, s->stream));
    return FVH_OK;
}

int32_t get_flat(fvh_solver* s, const void* dev, double* host, int64_t n) {
    FVH_HIP(hipMemcpyAsync(host, dev, (size_t)n * 8, hipMemcpyDeviceToHost, s->stream));
    FVH_HIP(hipStreamSynchronize(s->stream));
    return FVH_OK;
}

#define FVH_TRY(x) do { int32_t rc_ = (x); if (rc_) return rc_; } while (0)

Q4<double>* ph_q4(fvh_solver* s, int k) { return (Q4<double>*)s->ph[k].p; }

}  // namespace
}  // namespace fvh

// ---------------------------------------------------------------------------
// C ABI
// ---------------------------------------------------------------------------
using namespace fvh;

namespace fvh {
namespace {
typedef float v4f __attribute__((ext_vector_type(4)));
// grid-stride copy, 16 B per lane per iteration; NT: non-temporal loads and
// stores (the data is touched once)
template <bool NT>
__global__ void k_stream_copy(const v4f* __restrict__ a, v4f* __restrict__ b, size_t n) {
    for (size_t i = (size_t)blockIdx.x * blockDim.x + threadIdx.x; i < n; i += (size_t)gridDim.x * blockDim.x) {
        if (NT) {
            const v4f v = __builtin_nontemporal_load(a + i);
            __builtin_nontemporal_store(v, b + i);
        } else {
            b[i] = a[i];
        }
    }
}
}  // namespace
}  // namespace fvh

extern "C" {

void fvh_config_default(fvh_config* c) {
    if (!c) return;
    std::memset(c, 0, sizeof *c);
    c->precision = FVH_F64;
    c->flux = FVH_FLUX_ROE_SPLIT;
    c->gamma = 1.4;
    c->phi = FVH_PHI_ZERO;
    c->phi_yt = 2.0;
    c->K = 7.5;
    c->dt_kind = FVH_DT_FIXED;
    c->dt_fixed = 2e-5;
    c->cfl = 1.0;
    c->flags = FVH_FLAG_GLOBAL_DT;
    const double rk5[5] = {0.0695, 0.1602, 0.2898, 0.506, 1.};
    c->n_alpha = 5;
    for (int k = 0; k < 5; ++k) c->alpha[k] = rk5[k];
    c->renumber = FVH_ORDER_TILES;
    c->flux_mode = FVH_FLUX_MODE_AUTO;
    c->tile_cells = 0;
    // consts:: of examples/flatPlate/main.cpp:28-35
    c->mu = 2e-5;
    c->pr = 0.72;
    c->cp = 1.0;
    c->r_gas = 8.3145;
}

int32_t fvh_solver_create(const fvh_mesh* m, const fvh_config* c, fvh_solver** out) {
    if (!m || !c || !out) { set_error("null argument"); return FVH_E_ARG; }
    return create(1, &m, c, out);
}

int32_t fvh_batch_create(int32_t n, const fvh_mesh* const* meshes, const fvh_config* configs, fvh_solver** out) {
    return create(n, meshes, configs, out);
}

int32_t fvh_batch_size(const fvh_solver* s) { return s ? s->n_inst : 0; }

int32_t fvh_solver_stage_info(const fvh_solver* s, int64_t* info) {
    if (!s || !info) { set_error("null argument"); return FVH_E_ARG; }
    for (int k = 0; k < 8; ++k) info[k] = 0;
    info[0] = s->fused ? 1 : 0;
    info[1] = s->n_rt;
    info[6] = s->cfg.tile_cells ? s->cfg.tile_cells : kBlock;
    if (!s->fused) return FVH_OK;
    HaloView hv = halo_view(const_cast<fvh_solver*>(s));
    const size_t lds = s->mixed ? stage_lds_bytes<double, float>(hv)
                                : s->f32 ? stage_lds_bytes<float, float>(hv) : stage_lds_bytes<double, double>(hv);
    info[2] = (int64_t)lds;
    info[3] = stage_lanes(lds, s->cfg.stage_threads, s->x_mean_g);
    info[4] = s->x_max_ext;
    info[5] = s->x_max_edges;
    info[7] = s->x_mean_g;
    return FVH_OK;
}

int32_t fvh_solver_tile_cells(const fvh_solver* s) {
    if (!s) { set_error("null solver"); return FVH_E_ARG; }
    return s->cfg.tile_cells ? s->cfg.tile_cells : kBlock;
}

int32_t fvh_batch_instance(const fvh_solver* s, int32_t k, int64_t* info) {
    if (!s || !info || k < 0 || k >= s->n_inst) { set_error("bad argument"); return FVH_E_ARG; }
    info[0] = s->i_ncells[k];
    info[1] = s->i_nreal[k];
    info[2] = s->i_state0[k];
    return FVH_OK;
}

int32_t fvh_batch_status(fvh_solver* s, int32_t k, fvh_status* st) {
    if (!s || !st || k < 0 || k >= s->n_inst) { set_error("bad argument"); return FVH_E_ARG; }
    std::vector<Ctrl> h;
    int32_t rc = read_ctrl(s, h);
    if (rc) return rc;
    to_status(h[k], st);
    return FVH_OK;
}

int32_t fvh_batch_sums(fvh_solver* s, double* out) {
    if (!s || !out) { set_error("null argument"); return FVH_E_ARG; }
    return FVH_DISPATCH(s, sums_t, s, out);
}

int32_t fvh_autotune(int32_t n, const fvh_mesh* const* meshes, const fvh_config* configs, const double* q,
                     int32_t steps, int32_t* best_mode, int32_t* best_tile, double* report, int32_t max_report) {
    if (n < 1 || !meshes || !configs || !q || steps < 1 || !best_mode || !best_tile) {
        set_error("bad argument");
        return FVH_E_ARG;
    }
    const int32_t modes[3] = {FVH_FLUX_MODE_TILE, FVH_FLUX_MODE_FUSED, FVH_FLUX_MODE_CELL};
    const int32_t tiles[3] = {128, 192, 256};
    double best = 1e300;
    int32_t rows = 0;
    *best_mode = FVH_FLUX_MODE_TILE;
    *best_tile = kBlock;
    std::vector<fvh_config> cs(configs, configs + n);
    hipEvent_t e0 = nullptr, e1 = nullptr;
    FVH_HIP(hipEventCreate(&e0));
    FVH_HIP(hipEventCreate(&e1));
    int32_t rc = FVH_OK;
    for (int32_t mi = 0; mi < 3 && rc == FVH_OK; ++mi)
        for (int32_t ti = 0; ti < 3 && rc == FVH_OK; ++ti) {
            for (auto& c : cs) { c.flux_mode = modes[mi]; c.tile_cells = tiles[ti]; }
            fvh_solver* s = nullptr;
            double ms = -1.0;
            if (create(n, meshes, cs.data(), &s) == FVH_OK) {
                fvh_run_options o{1e300, 0xffffffffu, 1, 0.0, -1, 0};
                rc = FVH_DISPATCH(s, set_state_t, s, q);
                if (!rc) rc = FVH_DISPATCH(s, apply_bounds_t, s);
                for (int k = 0; k < 2 && !rc; ++k) rc = FVH_DISPATCH(s, fvh::step, s, false, &o, 0);
                if (!rc && hipEventRecord(e0, s->stream) != hipSuccess) rc = FVH_E_HIP;
                for (int k = 0; k < steps && !rc; ++k) rc = FVH_DISPATCH(s, fvh::step, s, false, &o, 0);
                if (!rc && hipEventRecord(e1, s->stream) != hipSuccess) rc = FVH_E_HIP;
                if (!rc && hipEventSynchronize(e1) != hipSuccess) rc = FVH_E_HIP;
                float t = 0.f;
                if (!rc && hipEventElapsedTime(&t, e0, e1) == hipSuccess) ms = (double)t / steps;
                delete s;
            }
            if (report && rows < max_report) {
                report[3 * rows] = modes[mi];
                report[3 * rows + 1] = tiles[ti];
                report[3 * rows + 2] = ms;
            }
            ++rows;
            if (ms > 0.0 && ms < best) { best = ms; *best_mode = modes[mi]; *best_tile = tiles[ti]; }
        }
    (void)hipEventDestroy(e0);
    (void)hipEventDestroy(e1);
    if (rc) return rc;
    return rows;
}

int32_t fvh_objective(fvh_solver* s, const double* q, int32_t n_steps, const double* wx, const double* wy,
                      double* J) {
    if (!s || !q || !J || n_steps < 0) { set_error("bad argument"); return FVH_E_ARG; }
    int32_t rc = adjoint_check(s);
    if (rc) return rc;
    if ((rc = objective_setup(s, wx, wy))) return rc;
    s->cur = 0;
    if ((rc = set_state_t<double, double>(s, q)) || (rc = apply_bounds_t<double, double>(s))) return rc;
    fvh_run_options o{1e300, 0xffffffffu, 1, 0.0, -1, 0};
    for (int32_t k = 0; k < n_steps; ++k)
        if ((rc = fvh::step<double, double>(s, false, &o, 0))) return rc;
    return objective_eval(s, s->q[s->cur].p, false, J);
}

int32_t fvh_adjoint(fvh_solver* s, const double* q, int32_t n_steps, const double* wx, const double* wy,
                    int32_t snaps, double mem_cap_bytes, double* J, double* grad, double* grad_nodes,
                    fvh_adjoint_stats* st) {
    if (!s || !q || !J || !grad || n_steps < 0 || snaps < 0) { set_error("bad argument"); return FVH_E_ARG; }
    int32_t rc = adjoint_check(s);
    if (rc) return rc;
    if (grad_nodes && !s->mesh0) { set_error("node gradients: the solver keeps no mesh"); return FVH_E_UNSUPPORTED; }
    const auto t0 = std::chrono::steady_clock::now();
    Adjoint A{s};
    A.qbytes = (size_t)s->n_total * sizeof(Q4<double>);
    A.n_stage = s->cfg.n_alpha;
    const double cap = mem_cap_bytes > 0 ? mem_cap_bytes : 8e9;
    const double avail = cap - A.work_bytes() - (grad_nodes ? 2.0 * A.qbytes + 3.0 * s->n_he * 32.0 : 0.0);
    int32_t n_rec = 0;   // last steps kept whole by the forward sweep (no replay in the reverse)
    if (snaps == 0) {
        snaps = (int32_t)std::max(1.0, std::min((double)std::max(n_steps, 1), std::floor(avail / A.qbytes)));
        if (snaps >= n_steps && n_steps > 0) {
            // the rest of the budget: step records instead of checkpoints, latest steps first
            const double extra = (double)A.record_bytes() + 2048.0 * (A.n_stage + 2) - (double)A.qbytes;
            n_rec = (int32_t)std::max(0.0, std::min((double)n_steps, std::floor((avail - (double)n_steps * A.qbytes) / extra)));
            snaps = std::max(1, n_steps - n_rec);
        }
    }
    snaps = std::max(1, std::min(snaps, std::max(n_steps, 1)));
    if ((rc = A.alloc(snaps)) || (rc = objective_setup(s, wx, wy))) return rc;
    if (n_rec && (rc = A.alloc_tapes(n_rec))) return rc;
    if (grad_nodes && (rc = A.alloc_geo())) return rc;
    // initial state: bounds applied (explicit.cpp:712), kept in slot 0
    s->cur = 0;
    if ((rc = set_state_t<double, double>(s, q)) || (rc = apply_bounds_t<double, double>(s))) return rc;
    launch_copy<double>(s->stream, s->q[s->cur].p, s->a_ck[0]->p, s->n_total);
    // forward to the objective, then seed the adjoint of the final state.
    // With a slot per step the sweep keeps every step's state and each step
    // is replayed once; otherwise the binomial schedule recomputes.
    const bool every = n_rec > 0 || snaps >= n_steps;
    const int32_t i0 = n_steps - n_rec;   // first recorded step
    const void* final_state = s->a_tmp.p;
    if (every) {
        // step in the solver's own buffers, keeping a copy of each step's start
        fvh_run_options o{1e300, 0xffffffffu, 1, 0.0, -1, 0};
        for (int k = 0; k < i0; ++k) {
            if ((rc = fvh::step<double, double>(s, false, &o, 0))) return rc;
            void* dst = k + 1 < i0 ? s->a_ck[k + 1]->p : (n_rec ? A.tapes[0].qs[0] : s->a_tmp.p);
            launch_copy<double>(s->stream, s->q[s->cur].p, dst, s->n_total);
        }
        if (n_rec) {
            // the last steps recorded whole; each record's output is the next one's input
            if (i0 == 0)
                launch_copy<double>(s->stream, s->a_ck[0]->p, A.tapes[0].qs[0], s->n_total);
            for (int k = 0; k < n_rec; ++k) {
                if (k > 0) A.tapes[k].qs[0] = A.tapes[k - 1].qs[A.n_stage];
                if ((rc = A.replay_body(A.tapes[k]))) return rc;
            }
            final_state = A.tapes[n_rec - 1].qs[A.n_stage];
        }
        A.fwd_steps += n_steps;
        if (n_steps == 0)
            launch_copy<double>(s->stream, s->a_ck[0]->p, s->a_tmp.p, s->n_total);
    } else if ((rc = A.advance(s->a_ck[0]->p, s->a_tmp.p, n_steps))) {
        return rc;
    }
    FVH_HIP(hipStreamSynchronize(s->stream));
    const auto t1 = std::chrono::steady_clock::now();
    if ((rc = objective_eval(s, final_state, true, J, grad_nodes ? (double*)s->a_obj_p.p : nullptr))) return rc;
    if (every) {
        for (int k = n_rec - 1; k >= 0; --k)
            if ((rc = A.reverse(A.tapes[k]))) return rc;
        for (int k = i0 - 1; k >= 0; --k)
            if ((rc = A.step_back(s->a_ck[k]->p))) return rc;
    } else {
        std::vector<int> free_slots;
        for (int k = snaps - 1; k >= 1; --k) free_slots.push_back(k);
        if ((rc = A.rev(0, n_steps, 0, free_slots))) return rc;
    }
    launch_adj_bounds_init<double>(s->stream, view<double>(s), phys(s->cfg), (const BcParam<double>*)s->bcs.p,
                                   (const Q4<double>*)s->a_ck[0]->p, (Q4<double>*)s->a_lam.p, A.ga);
    FVH_HIP(hipGetLastError());
    if ((rc = get_q4<double>(s, s->a_lam.p, grad, s->n_total))) return rc;
    if (grad_nodes && (rc = node_gradient(s, wx, wy, grad_nodes))) return rc;
    // ghost entries: run() overwrites them, so J does not depend on them
    for (int32_t g = s->n_real; g < s->n_total; ++g)
        for (int k = 0; k < 4; ++k) grad[4 * (int64_t)s->inv[g] + k] = 0.0;
    const auto t2 = std::chrono::steady_clock::now();
    if (st) {
        st->steps = n_steps;
        st->snaps = snaps;
        st->forward_steps = A.fwd_steps;
        st->checkpoint_bytes = (double)snaps * A.qbytes + (double)n_rec * A.record_bytes();
        st->recorded_steps = n_rec;
        st->work_bytes = A.work_bytes();
        st->forward_ms = std::chrono::duration<double, std::milli>(t1 - t0).count();
        st->total_ms = std::chrono::duration<double, std::milli>(t2 - t0).count();
    }
    return FVH_OK;
}

int32_t fvh_solver_use_graphs(fvh_solver* s, int32_t enable) {
    if (!s) { set_error("null argument"); return FVH_E_ARG; }
    FVH_HIP(hipStreamSynchronize(s->stream));
    if (!enable) s->drop_graphs();
    s->use_graphs = enable != 0;
    return FVH_OK;
}

void fvh_solver_free(fvh_solver* s) { delete s; }

int32_t fvh_solver_set_state(fvh_solver* s, const double* q) {
    if (!s || !q) { set_error("null argument"); return FVH_E_ARG; }
    s->cur = 0;
    return FVH_DISPATCH(s, set_state_t, s, q);
}

int32_t fvh_solver_get_state(fvh_solver* s, double* q) {
    if (!s || !q) { set_error("null argument"); return FVH_E_ARG; }
    return (s->f32 || s->mixed) ? get_q4<float>(s, s->q[s->cur].p, q, s->n_total)
                                : get_q4<double>(s, s->q[s->cur].p, q, s->n_total);
}

// Restart file, little-endian: "FVHSTAT1", int32 {version 1, precision,
// instances, 0}, int64 value count (4 per cell, all cells), one Ctrl record
// per instance (step, time, residuals, R0, stop flag), then the state in the
// reference layout as doubles (exact for either precision).
namespace {
const char kRestartMagic[8] = {'F', 'V', 'H', 'S', 'T', 'A', 'T', '1'};
static_assert(sizeof(fvh::Ctrl) == 104, "restart record layout");
}  // namespace

int32_t fvh_solver_save(fvh_solver* s, const char* path) {
    if (!s || !path) { set_error("null argument"); return FVH_E_ARG; }
    const int64_t nv = 4 * (int64_t)s->n_total;
    std::vector<double> q((size_t)nv);
    int32_t rc = fvh_solver_get_state(s, q.data());
    if (rc) return rc;
    std::vector<Ctrl> h;
    if ((rc = read_ctrl(s, h))) return rc;
    std::FILE* f = std::fopen(path, "wb");
    if (!f) { set_error(std::string("cannot open ") + path + " for writing"); return FVH_E_IO; }
    const int32_t head[4] = {1, s->cfg.precision, s->n_inst, 0};
    bool ok = std::fwrite(kRestartMagic, 1, 8, f) == 8 && std::fwrite(head, 4, 4, f) == 4 &&
              std::fwrite(&nv, 8, 1, f) == 1 && std::fwrite(h.data(), sizeof(Ctrl), h.size(), f) == h.size() &&
              std::fwrite(q.data(), 8, q.size(), f) == q.size();
    ok = (std::fclose(f) == 0) && ok;
    if (!ok) { set_error(std::string("write failed: ") + path); return FVH_E_IO; }
    return FVH_OK;
}

int32_t fvh_solver_load(fvh_solver* s, const char* path) {
    if (!s || !path) { set_error("null argument"); return FVH_E_ARG; }
    std::FILE* f = std::fopen(path, "rb");
    if (!f) { set_error(std::string("cannot open ") + path); return FVH_E_IO; }
    char magic[8];
    int32_t head[4];
    int64_t nv = 0;
    bool ok = std::fread(magic, 1, 8, f) == 8 && std::memcmp(magic, kRestartMagic, 8) == 0 &&
              std::fread(head, 4, 4, f) == 4 && std::fread(&nv, 8, 1, f) == 1;
    if (!ok || head[0] != 1) {
        std::fclose(f);
        set_error(std::string("not a restart file: ") + path);
        return FVH_E_FORMAT;
    }
    if (head[2] != s->n_inst || nv != 4 * (int64_t)s->n_total) {
        std::fclose(f);
        set_error("restart file holds " + std::to_string(head[2]) + " instances and " + std::to_string(nv / 4) +
                  " cells; the solver has " + std::to_string(s->n_inst) + " and " + std::to_string(s->n_total));
        return FVH_E_ARG;
    }
    std::vector<Ctrl> h((size_t)s->n_inst);
    std::vector<double> q((size_t)nv);
    ok = std::fread(h.data(), sizeof(Ctrl), h.size(), f) == h.size() &&
         std::fread(q.data(), 8, q.size(), f) == q.size();
    std::fclose(f);
    if (!ok) { set_error(std::string("truncated restart file: ") + path); return FVH_E_FORMAT; }
    int32_t rc = fvh_solver_set_state(s, q.data());
    if (rc) return rc;
    FVH_HIP(hipMemcpyAsync(s->ctrl.p, h.data(), h.size() * sizeof(Ctrl), hipMemcpyHostToDevice, s->stream));
    FVH_HIP(hipMemsetAsync(s->unphysical.p, 0, s->unphysical.bytes, s->stream));
    FVH_HIP(hipStreamSynchronize(s->stream));
    return FVH_OK;
}

int32_t fvh_solver_apply_bounds(fvh_solver* s) {
    if (!s) { set_error("null argument"); return FVH_E_ARG; }
    return FVH_DISPATCH(s, apply_bounds_t, s);
}

int32_t fvh_solver_phases(fvh_solver* s, double alpha, double* dt, double* gx, double* gy, double* lim,
                          double* qt, double* qk) {
    if (!s) { set_error("null argument"); return FVH_E_ARG; }
    if (s->n_inst != 1) { set_error("phases: single-instance solvers only"); return FVH_E_UNSUPPORTED; }
    return FVH_DISPATCH(s, phases_t, s, alpha, dt, gx, gy, lim, qt, qk);
}

int32_t fvh_phase_gradients(fvh_solver* s, const double* q, double* gx, double* gy) {
    if (!s || !q || !gx || !gy) { set_error("null argument"); return FVH_E_ARG; }
    FVH_TRY(phase_setup(s));
    FVH_TRY(put_q4(s, q, s->ph[0].p));
    launch_gradlim<double>(s->stream, view<double>(s), tile_map(s), phys(s->cfg), s->cfg.phi, true, false, false,
                           ph_q4(s, 0), (Q4<double>*)s->face.p, (Q4<double>*)s->bincr.p, -1, (double*)s->dt.p,
                           (double*)s->blk_min.p, ph_q4(s, 1), ph_q4(s, 2), nullptr, nullptr);
    FVH_HIP(hipGetLastError());
    FVH_TRY(get_q4<double>(s, s->ph[1].p, gx, s->n_total));
    return get_q4<double>(s, s->ph[2].p, gy, s->n_total);
}

int32_t fvh_phase_limiters(fvh_solver* s, const double* q, const double* gx, const double* gy, double* lim,
                           double* qmin, double* qmax) {
    if (!s || !q || !gx || !gy || !lim) { set_error("null argument"); return FVH_E_ARG; }
    FVH_TRY(phase_setup(s));
    FVH_TRY(put_q4(s, q, s->ph[0].p));
    FVH_TRY(put_q4(s, gx, s->ph[1].p));
    FVH_TRY(put_q4(s, gy, s->ph[2].p));
    launch_phase_limiters<double>(s->stream, view<double>(s), phys(s->cfg), s->cfg.phi, ph_q4(s, 0), ph_q4(s, 1),
                                  ph_q4(s, 2), ph_q4(s, 3), qmin ? ph_q4(s, 4) : nullptr,
                                  qmax ? ph_q4(s, 5) : nullptr);
    FVH_HIP(hipGetLastError());
    FVH_TRY(get_q4<double>(s, s->ph[3].p, lim, s->n_total));
    if (qmin) FVH_TRY(get_q4<double>(s, s->ph[4].p, qmin, s->n_total));
    if (qmax) FVH_TRY(get_q4<double>(s, s->ph[5].p, qmax, s->n_total));
    return FVH_OK;
}

int32_t fvh_phase_time_derivatives(fvh_solver* s, const double* q, const double* gx, const double* gy,
                                   const double* lim, double* qt) {
    if (!s || !q || !gx || !gy || !lim || !qt) { set_error("null argument"); return FVH_E_ARG; }
    FVH_TRY(phase_setup(s));
    const fvh_config& c = s->cfg;
    const bool lin = c.flags & FVH_FLAG_LINEAR;
    if (lin && !s->second) { set_error("internal: no face buffer"); return FVH_E_UNSUPPORTED; }
    FVH_TRY(put_q4(s, q, s->ph[0].p));
    FVH_TRY(put_q4(s, gx, s->ph[1].p));
    FVH_TRY(put_q4(s, gy, s->ph[2].p));
    FVH_TRY(put_q4(s, lim, s->ph[3].p));
    const MeshView<double> m = view<double>(s);
    if (lin) launch_phase_faces<double>(s->stream, m, ph_q4(s, 0), ph_q4(s, 1), ph_q4(s, 2), ph_q4(s, 3),
                                        (Q4<double>*)s->face.p);
    // ghost cells keep qt = 0 (explicit.cpp:352-362)
    FVH_HIP(hipMemsetAsync(s->ph[4].p, 0, s->ph[4].bytes, s->stream));
    const TileView tv = tile_view(s);
    launch_flux<double>(s->stream, m, tile_map(s), s->use_tiles ? &tv : nullptr, phys(c), c.flux, lin,
                        (const BcParam<double>*)s->bcs.p, ph_q4(s, 0), ph_q4(s, 0), ph_q4(s, 5),
                        (const Q4<double>*)s->face.p, (const Q4<double>*)s->bincr.p, (const double*)s->dt.p, nullptr,
                        0.0, nullptr, nullptr, nullptr, ph_q4(s, 4), ph_q4(s, 1), ph_q4(s, 2));
    FVH_HIP(hipGetLastError());
    return get_q4<double>(s, s->ph[4].p, qt, s->n_total);
}

int32_t fvh_phase_update_cells(fvh_solver* s, double* q, const double* ql, const double* qt, const double* dt,
                               double v, int64_t n) {
    if (!q || !ql || !qt || !dt || n < 0) { set_error("bad argument"); return FVH_E_ARG; }
    if (!s) {
        // no mesh: temporary device buffers on the current device's null stream
        DevBuf b[4];
        for (auto& x : b)
            if (x.alloc((size_t)n * 8)) { set_error("hipMalloc failed"); return FVH_E_HIP; }
        FVH_HIP(hipMemcpy(b[1].p, ql, (size_t)n * 8, hipMemcpyHostToDevice));
        FVH_HIP(hipMemcpy(b[2].p, qt, (size_t)n * 8, hipMemcpyHostToDevice));
        FVH_HIP(hipMemcpy(b[3].p, dt, (size_t)n * 8, hipMemcpyHostToDevice));
        launch_update_cells(nullptr, (double*)b[0].p, (const double*)b[1].p, (const double*)b[2].p,
                            (const double*)b[3].p, v, n);
        FVH_HIP(hipGetLastError());
        FVH_HIP(hipMemcpy(q, b[0].p, (size_t)n * 8, hipMemcpyDeviceToHost));
        return FVH_OK;
    }
    FVH_TRY(phase_setup(s));
    if (n > 4 * (int64_t)s->n_total) { set_error("update_cells: more values than 4 per cell"); return FVH_E_ARG; }
    FVH_TRY(put_flat(s, ql, s->ph[1].p, n));
    FVH_TRY(put_flat(s, qt, s->ph[2].p, n));
    FVH_TRY(put_flat(s, dt, s->ph[3].p, n));
    launch_update_cells(s->stream, (double*)s->ph[0].p, (const double*)s->ph[1].p, (const double*)s->ph[2].p,
                        (const double*)s->ph[3].p, v, n);
    FVH_HIP(hipGetLastError());
    return get_flat(s, s->ph[0].p, q, n);
}

int32_t fvh_phase_update_bounds(fvh_solver* s, double* q, const double* gx, const double* gy, const double* lim) {
    if (!s || !q || !gx || !gy || !lim) { set_error("null argument"); return FVH_E_ARG; }
    FVH_TRY(phase_setup(s));
    FVH_TRY(put_q4(s, q, s->ph[0].p));
    FVH_TRY(put_q4(s, gx, s->ph[1].p));
    FVH_TRY(put_q4(s, gy, s->ph[2].p));
    FVH_TRY(put_q4(s, lim, s->ph[3].p));
    launch_bounds_init<double>(s->stream, view<double>(s), phys(s->cfg), (s->cfg.flags & FVH_FLAG_LINEAR) != 0,
                               (const BcParam<double>*)s->bcs.p, ph_q4(s, 0), ph_q4(s, 1), ph_q4(s, 2),
                               ph_q4(s, 3));
    FVH_HIP(hipGetLastError());
    return get_q4<double>(s, s->ph[0].p, q, s->n_total);
}

int32_t fvh_phase_min_dt(fvh_solver* s, double* dt, int64_t n) {
    if (!s || !dt || n < 1) { set_error("bad argument"); return FVH_E_ARG; }
    FVH_TRY(phase_setup(s));
    if (n > 4 * (int64_t)s->n_total) { set_error("min_dt: more values than 4 per cell"); return FVH_E_ARG; }
    FVH_TRY(put_flat(s, dt, s->ph[0].p, n));
    launch_array_min(s->stream, (double*)s->ph[0].p, n, (double*)s->ph_part.p);
    FVH_HIP(hipGetLastError());
    return get_flat(s, s->ph[0].p, dt, n);
}

int32_t fvh_min_dt_device(double* dt, int64_t n, void* stream) {
    if (!dt || n < 1) { set_error("bad argument"); return FVH_E_ARG; }
    if (n > kMinDtDeviceMax) {
        set_error("min_dt on a device array: at most " + std::to_string(kMinDtDeviceMax) + " values, got " +
                  std::to_string(n));
        return FVH_E_ARG;
    }
    double* part = nullptr;
    int32_t rc = scratch_address(&part);
    if (rc) return rc;
    // every block overwrites its own slot: nothing to clear beforehand
    launch_array_min((hipStream_t)stream, dt, n, part);
    FVH_HIP(hipGetLastError());
    return FVH_OK;
}

int32_t fvh_phase_residuals(fvh_solver* s, const double* qt, double* R) {
    if (!s || !qt || !R) { set_error("null argument"); return FVH_E_ARG; }
    FVH_TRY(phase_setup(s));
    FVH_TRY(put_q4(s, qt, s->ph[0].p));
    launch_residuals<double>(s->stream, view<double>(s), tile_map(s), inst_view(s), ph_q4(s, 0),
                             (double*)s->sums_blk.p, (double*)s->sums.p);
    FVH_HIP(hipGetLastError());
    return get_flat(s, s->sums.p, R, 4);
}

static bool bad_options(const fvh_run_options* o) {
    if (o->stop_var < -1 || o->stop_var > 3) { set_error("stop_var must be -1 or a variable index 0..3"); return true; }
    return false;
}

int32_t fvh_solver_run(fvh_solver* s, const fvh_run_options* o, fvh_status* st) {
    if (!s || !o || !st) { set_error("null argument"); return FVH_E_ARG; }
    if (bad_options(o)) return FVH_E_ARG;
    return FVH_DISPATCH(s, run_t, s, o, st);
}

int32_t fvh_solver_run_begin(fvh_solver* s) {
    if (!s) { set_error("null argument"); return FVH_E_ARG; }
    return FVH_DISPATCH(s, apply_bounds_t, s);
}

int32_t fvh_solver_run_continue(fvh_solver* s, const fvh_run_options* o, uint32_t n, fvh_status* st) {
    if (!s || !o || !st) { set_error("null argument"); return FVH_E_ARG; }
    if (bad_options(o)) return FVH_E_ARG;
    return FVH_DISPATCH(s, run_continue_t, s, o, n, st);
}

int32_t fvh_solver_run_log(fvh_solver* s, const fvh_run_options* o, fvh_row_cb cb, void* user, fvh_status* st) {
    if (!s || !o || !st) { set_error("null argument"); return FVH_E_ARG; }
    if (bad_options(o)) return FVH_E_ARG;
    int32_t rc = fvh_solver_run_begin(s);
    if (rc) return rc;
    const uint32_t print = o->print_interval ? o->print_interval : 1;
    uint32_t done = 0;
    for (;;) {
        // to the next step the reference prints: step 0, then every print-th
        const uint32_t n = done == 0 ? 1 : ((done - 1) / print + 1) * print + 1 - done;
        if ((rc = FVH_DISPATCH(s, run_continue_t, s, o, n, st))) return rc;
        if (st->step <= done) break;
        done = st->step;
        const uint32_t step = done - 1;
        if (cb && (step == 0 || step % print == 0)) cb(step, st, user);
        if (st->stopped) break;
    }
    return FVH_OK;
}

int32_t fvh_stream_copy(double mbytes, int32_t reps, double* gbs) {
    if (!(mbytes >= 1.0) || reps < 1 || !gbs) { set_error("bad argument"); return FVH_E_ARG; }
    using fvh::v4f;
    const size_t n = (size_t)(mbytes * 1e6) / sizeof(v4f);
    DevBuf a, b;
    if (a.alloc(n * sizeof(v4f)) || b.alloc(n * sizeof(v4f))) { set_error("hipMalloc failed"); return FVH_E_HIP; }
    FVH_HIP(hipMemset(a.p, 0, a.bytes));
    hipStream_t st = nullptr;
    FVH_HIP(hipStreamCreateWithFlags(&st, hipStreamNonBlocking));
    hipEvent_t e0 = nullptr, e1 = nullptr;
    int32_t rc = FVH_OK;
    if (hipEventCreate(&e0) != hipSuccess || hipEventCreate(&e1) != hipSuccess) rc = FVH_E_HIP;
    // the best of a few launch shapes: 8, 32 and 128 workgroups of 256 per CU
    // (256 CUs) or one lane per element, plain and non-temporal
    double best = 0.0;
    const size_t shapes[4] = {256 * 8, 256 * 32, 256 * 128, (n + 255) / 256};
    for (int nt = 0; nt < 2 && !rc; ++nt)
        for (size_t g : shapes) {
            const int blocks = (int)std::min<size_t>(std::max<size_t>(g, 1), (n + 255) / 256);
            auto launch = [&]() {
                if (nt) hipLaunchKernelGGL(k_stream_copy<true>, dim3(blocks), dim3(256), 0, st, (const v4f*)a.p,
                                           (v4f*)b.p, n);
                else hipLaunchKernelGGL(k_stream_copy<false>, dim3(blocks), dim3(256), 0, st, (const v4f*)a.p,
                                        (v4f*)b.p, n);
            };
            float ms = 0.f;
            launch();
            if (hipEventRecord(e0, st) != hipSuccess) { rc = FVH_E_HIP; break; }
            for (int r = 0; r < reps; ++r) launch();
            if (hipEventRecord(e1, st) != hipSuccess || hipEventSynchronize(e1) != hipSuccess ||
                hipEventElapsedTime(&ms, e0, e1) != hipSuccess || hipGetLastError() != hipSuccess) {
                rc = FVH_E_HIP;
                break;
            }
            best = std::max(best, 2.0 * (double)(n * sizeof(v4f)) * reps / (ms * 1e-3) / 1e9);
        }
    if (e0) (void)hipEventDestroy(e0);
    if (e1) (void)hipEventDestroy(e1);
    (void)hipStreamDestroy(st);
    if (rc) { set_error("stream copy failed"); return rc; }
    *gbs = best;
    return FVH_OK;
}

int32_t fvh_solver_advance(fvh_solver* s, int32_t n) {
    if (!s || n < 0) { set_error("bad argument"); return FVH_E_ARG; }
    fvh_run_options o{1e300, 0xffffffffu, 1, 0.0, -1, 0};
    for (int32_t k = 0; k < n; ++k) {
        int32_t rc = FVH_DISPATCH(s, fvh::step, s, false, &o, 0);
        if (rc) return rc;
    }
    return FVH_OK;
}

int32_t fvh_solver_synchronize(fvh_solver* s) {
    if (!s) { set_error("null argument"); return FVH_E_ARG; }
    FVH_HIP(hipStreamSynchronize(s->stream));
    return FVH_OK;
}

int32_t fvh_solver_status(fvh_solver* s, fvh_status* st) {
    if (!s || !st) { set_error("null argument"); return FVH_E_ARG; }
    return read_status(s, st);
}

int32_t fvh_solver_profile(fvh_solver* s, int32_t enable) {
    if (!s) { set_error("null argument"); return FVH_E_ARG; }
    int32_t rc = drain_events(s);
    s->profiling = enable != 0;
    if (enable) {
        for (int k = 0; k < K_COUNT; ++k) { s->k_count[k] = 0; s->k_ms[k] = 0.0; }
    }
    return rc;
}

int32_t fvh_solver_kernel_times(fvh_solver* s, int32_t max, int64_t* counts, double* total_ms) {
    if (!s) { set_error("null argument"); return FVH_E_ARG; }
    int32_t rc = drain_events(s);
    if (rc) return rc;
    for (int k = 0; k < K_COUNT && k < max; ++k) {
        if (counts) counts[k] = s->k_count[k];
        if (total_ms) total_ms[k] = s->k_ms[k];
    }
    return K_COUNT;
}

const char* fvh_solver_kernel_name(int32_t k) {
    return (k >= 0 && k < K_COUNT) ? kKernelNames[k] : "";
}

double fvh_solver_stage_bytes(const fvh_solver* s) { return s ? s->stage_bytes : 0.0; }

}  // extern "C"
