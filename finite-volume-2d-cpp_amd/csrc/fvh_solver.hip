// fvh_solver.hip -- device layout, RK driver and the C ABI of the solver.
//
// Replaces fvhyper::run (fvhyper/src/explicit.cpp:657-817) and the per-phase
// functions it calls, on one GPU.  The host code here only builds the device
// layout and queues kernels; every per-cell, per-edge and per-step operation
// runs on the GPU.
#include <hip/hip_runtime.h>

#include <algorithm>
#include <cmath>
#include <cstring>
#include <memory>
#include <string>
#include <vector>

#include "fvh_internal.h"
#include "fvh_kernels.h"

namespace fvh {
namespace {

const char* kKernelNames[] = {"dt", "dt_min", "gradlim", "flux", "ctrl_begin", "ctrl_end", "bounds_init"};
enum { K_DT = 0, K_DTMIN, K_GRADLIM, K_FLUX, K_BEGIN, K_END, K_BINIT, K_COUNT };

#define FVH_HIP(call)                                                                       \
    do {                                                                                    \
        hipError_t e_ = (call);                                                             \
        if (e_ != hipSuccess) {                                                             \
            set_error(std::string(#call) + ": " + hipGetErrorString(e_));                   \
            return FVH_E_HIP;                                                               \
        }                                                                                   \
    } while (0)

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

}  // namespace
}  // namespace fvh

struct fvh_solver {
    fvh_config cfg{};
    int device = 0;
    hipStream_t stream = nullptr;
    bool f32 = false;
    bool second = false;                  // gradient/limiter/reconstruction pass present
    int32_t n_real = 0, n_total = 0, n_edges = 0, n_he = 0, n_bnd = 0;
    int32_t dev_ref0 = 0;                 // device position of reference cell 0
    std::vector<int32_t> pos;             // reference cell -> device position
    std::vector<int32_t> inv;             // device position -> reference cell
    // device buffers
    fvh::DevBuf he_off, he_nbr, he_edge, he_twin, he_bc, he_slot;
    fvh::DevBuf e_geo, e_c, c_area, c_inv_area, c_sqrt_area, c_c;
    fvh::DevBuf t_off, t_edge, t_he, t_cell;
    int32_t n_tiles = 0;
    bool use_tiles = false;
    fvh::DevBuf q[3], dt, dt_global, face, bincr, blk_min, res_blk, ctrl, bcs;
    fvh::DevBuf dbg[4];
    int cur = 0;                          // index of the current state buffer
    // profiling
    bool profiling = false;
    std::vector<std::pair<int, std::pair<hipEvent_t, hipEvent_t>>> pending;
    std::vector<hipEvent_t> event_pool;
    int64_t k_count[fvh::K_COUNT] = {};
    double k_ms[fvh::K_COUNT] = {};
    double stage_bytes = 0.0;

    ~fvh_solver() {
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

TileView tile_view(fvh_solver* s) {
    TileView t;
    t.n_tiles = s->n_tiles;
    t.off = (const int32_t*)s->t_off.p;
    t.edge = (const int32_t*)s->t_edge.p;
    t.he = (const int2*)s->t_he.p;
    t.cell = (const int2*)s->t_cell.p;
    return t;
}

PhysParams phys(const fvh_config& c) {
    PhysParams p;
    p.gamma = c.gamma;
    p.cfl = c.cfl;
    p.dt_fixed = c.dt_fixed;
    p.K = c.K;
    p.phi_yt = c.phi_yt;
    p.positivity = (c.flags & FVH_FLAG_POSITIVITY) ? 1 : 0;
    return p;
}

template <class T>
std::vector<BcParam<T>> bc_table(const fvh_config& c) {
    std::vector<BcParam<T>> t(kMaxBc);
    for (int k = 0; k < kMaxBc; ++k) {
        t[k].kind = k < c.n_bc ? c.bc[k].kind : 0;
        t[k].rho = T(k < c.n_bc ? c.bc[k].rho : 0);
        t[k].u = T(k < c.n_bc ? c.bc[k].u : 0);
        t[k].v = T(k < c.n_bc ? c.bc[k].v : 0);
        t[k].p = T(k < c.n_bc ? c.bc[k].p : 0);
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

template <class T>
int32_t upload_geometry(fvh_solver* s, const fvh_mesh& m, const std::vector<int32_t>& dev_edge_ref,
                        const std::vector<double>& gg) {
    const int32_t E = s->n_edges, N = s->n_total;
    std::vector<Q4<T>> geo(E);
    std::vector<G2<T>> ec(E);
    for (int32_t d = 0; d < E; ++d) {
        const int32_t e = dev_edge_ref[d];
        geo[d].v[0] = T(m.e_nx[e]);
        geo[d].v[1] = T(m.e_ny[e]);
        geo[d].v[2] = T(m.e_len[e]);
        geo[d].v[3] = T(gg[e]);
        ec[d].v[0] = T(m.e_cx[e]);
        ec[d].v[1] = T(m.e_cy[e]);
    }
    std::vector<T> area(N), ia(N), sa(N);
    std::vector<G2<T>> cc(N);
    for (int32_t d = 0; d < N; ++d) {
        const int32_t r = s->inv[d];
        const double A = m.c_area[r];
        area[d] = T(A);
        ia[d] = T(1. / A);          // invA of explicit.cpp:158
        sa[d] = T(std::sqrt(A));    // sqrt_area of explicit.cpp:218
        cc[d].v[0] = T(m.c_cx[r]);
        cc[d].v[1] = T(m.c_cy[r]);
    }
    int32_t rc;
    const std::vector<BcParam<T>> bt = bc_table<T>(s->cfg);
    if ((rc = upload(s->bcs, bt.data(), bt.size() * sizeof(BcParam<T>)))) return rc;
    if ((rc = upload(s->e_geo, geo.data(), geo.size() * sizeof(Q4<T>)))) return rc;
    if ((rc = upload(s->e_c, ec.data(), ec.size() * sizeof(G2<T>)))) return rc;
    if ((rc = upload(s->c_area, area.data(), N * sizeof(T)))) return rc;
    if ((rc = upload(s->c_inv_area, ia.data(), N * sizeof(T)))) return rc;
    if ((rc = upload(s->c_sqrt_area, sa.data(), N * sizeof(T)))) return rc;
    if ((rc = upload(s->c_c, cc.data(), N * sizeof(G2<T>)))) return rc;
    for (int k = 0; k < 3; ++k)
        if (s->q[k].alloc((size_t)N * sizeof(Q4<T>))) { set_error("hipMalloc failed"); return FVH_E_HIP; }
    if (s->dt.alloc((size_t)N * sizeof(T)) || s->dt_global.alloc(8) ||
        s->face.alloc((size_t)(s->second ? s->n_he : 1) * sizeof(Q4<T>)) ||
        s->bincr.alloc((size_t)(s->n_bnd + 1) * sizeof(Q4<T>)) ||
        s->blk_min.alloc((size_t)((N + kBlock - 1) / kBlock + 1) * 8) ||
        s->res_blk.alloc((size_t)((s->n_real + kBlock - 1) / kBlock + 1) * 32) || s->ctrl.alloc(sizeof(Ctrl))) {
        set_error("hipMalloc failed");
        return FVH_E_HIP;
    }
    FVH_HIP(hipMemset(s->ctrl.p, 0, sizeof(Ctrl)));
    FVH_HIP(hipMemset(s->face.p, 0, s->face.bytes));
    FVH_HIP(hipMemset(s->bincr.p, 0, s->bincr.bytes));
    return FVH_OK;
}

// Queue the kernels of one RK step (explicit.cpp:729-747) reading state
// buffer s->cur.  run_mode adds the stop test and bookkeeping kernels.
// dbg (tests): 4 device buffers for gx, gy, limiter and qt of one stage run
// with alpha_override, leaving the current state buffer unchanged.
template <class T>
int32_t enqueue_step(fvh_solver* s, bool run_mode, const fvh_run_options* o, int need_res,
                     Q4<T>* const* dbg = nullptr, const double* alpha_override = nullptr) {
    const fvh_config& c = s->cfg;
    const MeshView<T> m = view<T>(s);
    const TileView tv = tile_view(s);
    const PhysParams pp = phys(c);
    const BcParam<T>* bcs = (const BcParam<T>*)s->bcs.p;
    const bool grad = c.flags & FVH_FLAG_GRADIENTS, lim = c.flags & FVH_FLAG_LIMITERS,
               lin = c.flags & FVH_FLAG_LINEAR, gdt = c.flags & FVH_FLAG_GLOBAL_DT;
    Ctrl* ctrl = (Ctrl*)s->ctrl.p;
    const Ctrl* cc = run_mode ? ctrl : nullptr;
    hipStream_t st = s->stream;
    if (run_mode) {
        Bracket b(s, K_BEGIN);
        launch_ctrl_begin(st, ctrl, o->max_step, o->max_time, o->tolerance, o->stop_var);
    }
    Q4<T>* q0 = (Q4<T>*)s->q[s->cur].p;
    Q4<T>* qa = (Q4<T>*)s->q[(s->cur + 1) % 3].p;
    Q4<T>* qb = (Q4<T>*)s->q[(s->cur + 2) % 3].p;
    T* dt = (T*)s->dt.p;
    double* blk = (double*)s->blk_min.p;
    Q4<T>* face = (Q4<T>*)s->face.p;
    Q4<T>* bincr = (Q4<T>*)s->bincr.p;
    if (!s->second) {
        Bracket b(s, K_DT);
        launch_dt<T>(st, m, c.dt_kind, pp, q0, dt, blk, cc);
    }
    const T* dtp = gdt ? (const T*)s->dt_global.p : dt;
    const int dt_stride = gdt ? 0 : 1;
    const Q4<T>* qin = q0;
    Q4<T>* out = qa;
    const int n_stage = alpha_override ? 1 : c.n_alpha;
    for (int k = 0; k < n_stage; ++k) {
        if (s->second) {
            // the first stage's pass also computes the time step from q0
            Bracket b(s, K_GRADLIM);
            launch_gradlim<T>(st, m, pp, c.phi, grad, lim, lin, qin, face, bincr, k == 0 ? c.dt_kind : -1, dt,
                              blk, dbg ? dbg[0] : nullptr, dbg ? dbg[1] : nullptr, dbg ? dbg[2] : nullptr, cc);
        }
        if (k == 0 && gdt) {
            Bracket b(s, K_DTMIN);
            launch_dt_min(st, blk, (s->n_total + kBlock - 1) / kBlock, dt, s->f32, s->dev_ref0,
                          (double*)s->dt_global.p, cc);
        }
        const bool last = k == n_stage - 1;
        {
            Bracket b(s, K_FLUX);
            launch_flux<T>(st, m, s->use_tiles ? &tv : nullptr, pp, c.flux == FVH_FLUX_ROE_JOINED, lin, bcs, qin,
                           q0, out, face, bincr, dtp, dt_stride,
                           T(alpha_override ? *alpha_override : c.alpha[k]),
                           (last && need_res) ? (double*)s->res_blk.p : nullptr, dbg ? dbg[3] : nullptr, cc);
        }
        qin = out;
        out = (out == qa) ? qb : qa;
    }
    if (!dbg) {
        Bracket b(s, K_END);
        launch_ctrl_end(st, ctrl, (const double*)s->res_blk.p, (s->n_real + kBlock - 1) / kBlock, need_res,
                        gdt ? s->dt_global.p : s->dt.p, s->f32, gdt ? 0 : s->dev_ref0);
    }
    FVH_HIP(hipGetLastError());
    // the last stage wrote qin; it becomes the current state
    if (!dbg) s->cur = (qin == qa) ? (s->cur + 1) % 3 : (s->cur + 2) % 3;
    return FVH_OK;
}

template <class T>
int32_t set_state_t(fvh_solver* s, const double* q) {
    std::vector<Q4<T>> h(s->n_total);
    for (int32_t d = 0; d < s->n_total; ++d) {
        const int32_t r = s->inv[d];
        for (int k = 0; k < 4; ++k) h[d].v[k] = T(q[4 * (int64_t)r + k]);
    }
    for (int b = 0; b < 3; ++b)
        FVH_HIP(hipMemcpy(s->q[b].p, h.data(), h.size() * sizeof(Q4<T>), hipMemcpyHostToDevice));
    return FVH_OK;
}

template <class T>
int32_t get_q4(fvh_solver* s, const void* dev, double* out, int32_t count) {
    std::vector<Q4<T>> h(count);
    FVH_HIP(hipStreamSynchronize(s->stream));
    FVH_HIP(hipMemcpy(h.data(), dev, h.size() * sizeof(Q4<T>), hipMemcpyDeviceToHost));
    for (int32_t d = 0; d < count; ++d) {
        const int32_t r = s->inv[d];
        for (int k = 0; k < 4; ++k) out[4 * (int64_t)r + k] = (double)h[d].v[k];
    }
    return FVH_OK;
}

template <class T>
int32_t phases_t(fvh_solver* s, double alpha, double* dt_out, double* gx, double* gy, double* lim,
                 double* qt, double* qk) {
    const bool gdt = s->cfg.flags & FVH_FLAG_GLOBAL_DT;
    const size_t qbytes = (size_t)s->n_total * sizeof(Q4<T>);
    for (int k = 0; k < 4; ++k) {
        if (!s->dbg[k].p && s->dbg[k].alloc(qbytes)) { set_error("hipMalloc failed"); return FVH_E_HIP; }
        FVH_HIP(hipMemsetAsync(s->dbg[k].p, 0, qbytes, s->stream));
    }
    Q4<T>* q0 = (Q4<T>*)s->q[s->cur].p;
    Q4<T>* out = (Q4<T>*)s->q[(s->cur + 1) % 3].p;
    // update_cells runs over every cell; ghost cells keep q (their qt is 0)
    FVH_HIP(hipMemcpyAsync(out, q0, qbytes, hipMemcpyDeviceToDevice, s->stream));
    Q4<T>* dbg[4] = {(Q4<T>*)s->dbg[0].p, (Q4<T>*)s->dbg[1].p, (Q4<T>*)s->dbg[2].p, (Q4<T>*)s->dbg[3].p};
    int32_t rc = enqueue_step<T>(s, false, nullptr, 0, dbg, &alpha);
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
    if (qk && (rc = get_q4<T>(s, out, qk, s->n_total))) return rc;
    return FVH_OK;
}

template <class T>
int32_t apply_bounds_t(fvh_solver* s) {
    const fvh_config& c = s->cfg;
    {
        Bracket b(s, K_BINIT);
        launch_bounds_init<T>(s->stream, view<T>(s), phys(c), (c.flags & FVH_FLAG_LINEAR) != 0,
                              (const BcParam<T>*)s->bcs.p, (Q4<T>*)s->q[s->cur].p);
    }
    FVH_HIP(hipGetLastError());
    Ctrl z;
    std::memset(&z, 0, sizeof z);
    for (int k = 0; k < 4; ++k) z.R[k] = 1.0;
    FVH_HIP(hipMemcpyAsync(s->ctrl.p, &z, sizeof z, hipMemcpyHostToDevice, s->stream));
    FVH_HIP(hipStreamSynchronize(s->stream));
    return FVH_OK;
}

int32_t read_status(fvh_solver* s, fvh_status* st) {
    Ctrl h;
    FVH_HIP(hipStreamSynchronize(s->stream));
    FVH_HIP(hipMemcpy(&h, s->ctrl.p, sizeof h, hipMemcpyDeviceToHost));
    st->step = h.step;
    st->stopped = h.stop;
    st->time = h.time;
    st->time_prev = h.time_prev;
    for (int k = 0; k < 4; ++k) { st->R[k] = h.R[k]; st->R0[k] = h.R0[k]; }
    return FVH_OK;
}

// Executes up to n more steps of run() from the current status (steps are
// queued in growing chunks; kernels queued after the stop are no-ops).
template <class T>
int32_t run_continue_t(fvh_solver* s, const fvh_run_options* o, uint64_t n, fvh_status* st) {
    int32_t rc;
    fvh_status now;
    if ((rc = read_status(s, &now))) return rc;
    const uint32_t print = o->print_interval ? o->print_interval : 1;
    const bool every = o->tolerance > 1.01e-16;
    uint64_t step = now.step;
    const uint64_t first = step;
    const int cur0 = s->cur;
    // each executed step moves the state forward by one or two buffers
    const int delta = (s->cfg.n_alpha % 2) ? 1 : 2;
    uint64_t chunk = 8, left = n;
    while (!now.stopped && left > 0) {
        const uint64_t k_n = std::min<uint64_t>(chunk, left);
        for (uint64_t k = 0; k < k_n; ++k, ++step) {
            const int need = (step == 0) || (step % print == 0) || every;
            if ((rc = enqueue_step<T>(s, true, o, need))) return rc;
        }
        left -= k_n;
        if ((rc = read_status(s, &now))) return rc;
        if (chunk < 256) chunk *= 2;
    }
    s->cur = (int)((cur0 + (uint64_t)(now.step - first) * delta) % 3);
    *st = now;
    return drain_events(s);
}

template <class T>
int32_t run_t(fvh_solver* s, const fvh_run_options* o, fvh_status* st) {
    int32_t rc = apply_bounds_t<T>(s);
    if (rc) return rc;
    return run_continue_t<T>(s, o, ~0ull, st);
}

// Tile edge lists for the tiled flux pass.  Returns false when a tile touches
// more edges than its LDS holds.
bool build_tiles(fvh_solver* s, const std::vector<int32_t>& off, const std::vector<uint32_t>& he_edge,
                 std::vector<uint16_t>& he_slot, std::vector<int32_t>& t_off, std::vector<int32_t>& t_edge,
                 std::vector<int32_t>& t_he, std::vector<int32_t>& t_cell) {
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
    s->n_tiles = (s->n_real + kBlock - 1) / kBlock;
    t_off.assign(s->n_tiles + 1, 0);
    std::vector<int32_t> local;
    std::vector<int32_t> slot_of(E, -1);
    for (int32_t t = 0; t < s->n_tiles; ++t) {
        const int32_t c0 = t * kBlock, c1 = std::min(s->n_real, c0 + kBlock);
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

}  // namespace
}  // namespace fvh

// ---------------------------------------------------------------------------
// C ABI
// ---------------------------------------------------------------------------
using namespace fvh;

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
}

int32_t fvh_solver_create(const fvh_mesh* m, const fvh_config* c, fvh_solver** out) {
    if (!m || !c || !out) { set_error("null argument"); return FVH_E_ARG; }
    if (c->n_alpha < 1 || c->n_alpha > FVH_MAX_ALPHA || c->n_bc < 0 || c->n_bc > FVH_MAX_BC) {
        set_error("config out of range (alpha or bc count)");
        return FVH_E_ARG;
    }
    if (c->precision != FVH_F64 && c->precision != FVH_F32) { set_error("unknown precision"); return FVH_E_ARG; }
    if (c->flux < 0 || c->flux > 1 || c->phi < 0 || c->phi > 3 || c->dt_kind < 0 || c->dt_kind > 2 ||
        c->renumber < 0 || c->renumber > 2 || c->flux_mode < 0 || c->flux_mode > 2) {
        set_error("unknown flux, limiter, time-step, order or flux-mode kind");
        return FVH_E_ARG;
    }
    for (int k = 0; k < c->n_bc; ++k)
        if (c->bc[k].kind < 0 || c->bc[k].kind > 3) { set_error("unknown boundary kind"); return FVH_E_ARG; }
    // physical name -> BC slot (mesh.cpp:662-667)
    std::vector<int32_t> slot_of_name(m->names.size(), -1);
    for (size_t t = 0; t < m->names.size(); ++t) {
        for (int k = 0; k < c->n_bc; ++k)
            if (m->names[t] == std::string(c->bc[k].name)) slot_of_name[t] = k;
    }
    for (size_t b = 0; b < m->bnd_tag.size(); ++b) {
        const int32_t t = m->bnd_tag[b];
        if (slot_of_name[t] < 0) {
            set_error("Physical name " + m->names[t] + " of tag " + std::to_string(t) + " not in mesh");
            return FVH_E_BC;
        }
    }
    if (m->n_edges() >= (int32_t)kEdgeMask) { set_error("too many edges"); return FVH_E_UNSUPPORTED; }
    std::unique_ptr<fvh_solver> s(new fvh_solver());
    s->cfg = *c;
    s->f32 = c->precision == FVH_F32;
    s->second = (c->flags & (FVH_FLAG_GRADIENTS | FVH_FLAG_LIMITERS | FVH_FLAG_LINEAR)) != 0;
    if (hipGetDevice(&s->device) != hipSuccess) { set_error("no HIP device"); return FVH_E_HIP; }
    FVH_HIP(hipStreamCreateWithFlags(&s->stream, hipStreamNonBlocking));
    s->n_real = m->n_real;
    s->n_total = m->n_cells;
    s->n_edges = m->n_edges();
    s->n_bnd = (int32_t)m->bnd_edge.size();
    // device order of the real cells; ghosts keep theirs
    s->pos.assign(s->n_total, 0);
    s->inv.assign(s->n_total, 0);
    std::vector<int32_t> order;
    if (c->renumber == FVH_ORDER_RCM) order = rcm_order(*m);
    else if (c->renumber == FVH_ORDER_TILES) order = rcb_order(*m, kBlock);
    else { order.resize(m->n_real); for (int32_t i = 0; i < m->n_real; ++i) order[i] = i; }
    for (int32_t k = 0; k < m->n_real; ++k) { s->pos[order[k]] = k; s->inv[k] = order[k]; }
    for (int32_t g = m->n_real; g < m->n_cells; ++g) { s->pos[g] = g; s->inv[g] = g; }
    s->dev_ref0 = s->n_total > 0 ? s->pos[0] : 0;
    // half-edges in CSR by device cell, each list sorted by (reference edge, side)
    const int32_t E = s->n_edges, N = s->n_total;
    std::vector<int32_t> off(N + 1, 0);
    for (int32_t e = 0; e < E; ++e)
        for (int sd = 0; sd < 2; ++sd) off[s->pos[m->edge_cells[2 * (int64_t)e + sd]] + 1] += 1;
    for (int32_t d = 0; d < N; ++d) off[d + 1] += off[d];
    s->n_he = off[N];
    std::vector<int32_t> cursor(off.begin(), off.end() - 1);
    std::vector<int32_t> he_nbr(s->n_he), he_twin(s->n_he), he_bc(s->n_he, -1), he_ref(s->n_he);
    std::vector<uint32_t> he_side(s->n_he);
    std::vector<int32_t> bc_of_edge(E, -1);
    for (size_t b = 0; b < m->bnd_edge.size(); ++b) bc_of_edge[m->bnd_edge[b]] = slot_of_name[m->bnd_tag[b]];
    for (int32_t e = 0; e < E; ++e) {
        const int32_t d0 = s->pos[m->edge_cells[2 * (int64_t)e]], d1 = s->pos[m->edge_cells[2 * (int64_t)e + 1]];
        const int32_t h0 = cursor[d0]++;
        const int32_t h1 = cursor[d1]++;
        he_nbr[h0] = d1; he_nbr[h1] = d0;
        he_twin[h0] = h1; he_twin[h1] = h0;
        he_ref[h0] = e; he_ref[h1] = e;
        he_side[h0] = 0; he_side[h1] = 1;
        if (bc_of_edge[e] >= 0) he_bc[h0] = bc_of_edge[e];
    }
    // device edge numbering: first use in half-edge order
    std::vector<int32_t> dev_of_edge(E, -1), dev_edge_ref;
    dev_edge_ref.reserve(E);
    std::vector<uint32_t> he_edge(s->n_he);
    for (int32_t h = 0; h < s->n_he; ++h) {
        const int32_t e = he_ref[h];
        if (dev_of_edge[e] < 0) { dev_of_edge[e] = (int32_t)dev_edge_ref.size(); dev_edge_ref.push_back(e); }
        uint32_t w = (uint32_t)dev_of_edge[e];
        if (he_side[h]) w |= kSideBit;
        if (he_bc[h] >= 0) w |= kBcBit;
        he_edge[h] = w;
    }
    // Green-Gauss weight per edge, as calc_gradients computes it (explicit.cpp:132-140)
    std::vector<double> gg(E);
    for (int32_t e = 0; e < E; ++e) {
        const int32_t i = m->edge_cells[2 * (int64_t)e], j = m->edge_cells[2 * (int64_t)e + 1];
        const double dxif = m->e_cx[e] - m->c_cx[i];
        const double dyif = m->e_cy[e] - m->c_cy[i];
        const double dif = std::sqrt(dxif * dxif + dyif * dyif);
        const double dxij = m->c_cx[i] - m->c_cx[j];
        const double dyij = m->c_cy[i] - m->c_cy[j];
        const double dij = std::sqrt(dxij * dxij + dyij * dyij);
        gg[e] = dif / dij;
    }
    // tiles for the tiled flux pass
    std::vector<uint16_t> he_slot(s->n_he, 0);
    std::vector<int32_t> t_off, t_edge, t_he, t_cell;
    s->use_tiles = false;
    if (c->flux_mode != FVH_FLUX_MODE_CELL && s->n_real > 0)
        s->use_tiles = build_tiles(s.get(), off, he_edge, he_slot, t_off, t_edge, t_he, t_cell);
    if (c->flux_mode == FVH_FLUX_MODE_TILE && !s->use_tiles && s->n_real > 0) {
        set_error("a tile touches more than 1024 edges; use the per-cell flux mode");
        return FVH_E_UNSUPPORTED;
    }
    int32_t rc;
    if ((rc = upload(s->he_off, off.data(), off.size() * 4))) return rc;
    if ((rc = upload(s->he_nbr, he_nbr.data(), he_nbr.size() * 4))) return rc;
    if ((rc = upload(s->he_edge, he_edge.data(), he_edge.size() * 4))) return rc;
    if ((rc = upload(s->he_twin, he_twin.data(), he_twin.size() * 4))) return rc;
    if ((rc = upload(s->he_bc, he_bc.data(), he_bc.size() * 4))) return rc;
    if ((rc = upload(s->he_slot, he_slot.data(), he_slot.size() * 2))) return rc;
    if (s->use_tiles) {
        if ((rc = upload(s->t_off, t_off.data(), t_off.size() * 4))) return rc;
        if ((rc = upload(s->t_edge, t_edge.data(), t_edge.size() * 4))) return rc;
        if ((rc = upload(s->t_he, t_he.data(), t_he.size() * 4))) return rc;
        if ((rc = upload(s->t_cell, t_cell.data(), t_cell.size() * 4))) return rc;
    }
    rc = s->f32 ? upload_geometry<float>(s.get(), *m, dev_edge_ref, gg)
                : upload_geometry<double>(s.get(), *m, dev_edge_ref, gg);
    if (rc) return rc;
    // algorithmic bytes per stage (state s, geometry g bytes)
    const double sb = s->f32 ? 4.0 : 8.0, gb = sb, V = 4.0;
    const double C = s->n_real, B = s->n_bnd;
    double bytes = C * (3 * V * sb + sb + gb) + E * (3 * gb + 8) + B * (2 * V * sb + 2 * gb + 4);
    if (s->second) bytes += C * (7 * V * sb + 2 * gb) + E * (2 * gb);
    s->stage_bytes = bytes;
    *out = s.release();
    return FVH_OK;
}

void fvh_solver_free(fvh_solver* s) { delete s; }

int32_t fvh_solver_set_state(fvh_solver* s, const double* q) {
    if (!s || !q) { set_error("null argument"); return FVH_E_ARG; }
    s->cur = 0;
    return s->f32 ? set_state_t<float>(s, q) : set_state_t<double>(s, q);
}

int32_t fvh_solver_get_state(fvh_solver* s, double* q) {
    if (!s || !q) { set_error("null argument"); return FVH_E_ARG; }
    return s->f32 ? get_q4<float>(s, s->q[s->cur].p, q, s->n_total) : get_q4<double>(s, s->q[s->cur].p, q, s->n_total);
}

int32_t fvh_solver_apply_bounds(fvh_solver* s) {
    if (!s) { set_error("null argument"); return FVH_E_ARG; }
    return s->f32 ? apply_bounds_t<float>(s) : apply_bounds_t<double>(s);
}

int32_t fvh_solver_phases(fvh_solver* s, double alpha, double* dt, double* gx, double* gy, double* lim,
                          double* qt, double* qk) {
    if (!s) { set_error("null argument"); return FVH_E_ARG; }
    return s->f32 ? phases_t<float>(s, alpha, dt, gx, gy, lim, qt, qk)
                  : phases_t<double>(s, alpha, dt, gx, gy, lim, qt, qk);
}

static bool bad_options(const fvh_run_options* o) {
    if (o->stop_var < -1 || o->stop_var > 3) { set_error("stop_var must be -1 or a variable index 0..3"); return true; }
    return false;
}

int32_t fvh_solver_run(fvh_solver* s, const fvh_run_options* o, fvh_status* st) {
    if (!s || !o || !st) { set_error("null argument"); return FVH_E_ARG; }
    if (bad_options(o)) return FVH_E_ARG;
    return s->f32 ? run_t<float>(s, o, st) : run_t<double>(s, o, st);
}

int32_t fvh_solver_run_begin(fvh_solver* s) {
    if (!s) { set_error("null argument"); return FVH_E_ARG; }
    return s->f32 ? apply_bounds_t<float>(s) : apply_bounds_t<double>(s);
}

int32_t fvh_solver_run_continue(fvh_solver* s, const fvh_run_options* o, uint32_t n, fvh_status* st) {
    if (!s || !o || !st) { set_error("null argument"); return FVH_E_ARG; }
    if (bad_options(o)) return FVH_E_ARG;
    return s->f32 ? run_continue_t<float>(s, o, n, st) : run_continue_t<double>(s, o, n, st);
}

int32_t fvh_solver_advance(fvh_solver* s, int32_t n) {
    if (!s || n < 0) { set_error("bad argument"); return FVH_E_ARG; }
    fvh_run_options o{1e300, 0xffffffffu, 1, 0.0, -1, 0};
    for (int32_t k = 0; k < n; ++k) {
        int32_t rc = s->f32 ? enqueue_step<float>(s, false, &o, 0) : enqueue_step<double>(s, false, &o, 0);
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
