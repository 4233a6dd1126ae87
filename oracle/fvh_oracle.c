/*
 * fvh_oracle.c -- plain-C restatement of fvhyper's explicit path.
 * TEST INFRASTRUCTURE ONLY (see fvh_oracle.h).  Build: oracle/Makefile
 * (-ffp-contract=off, no fast-math), matching the reference's x86-64
 * baseline build in which no multiply-add is ever fused.
 *
 * Arithmetic is written in the reference's association on purpose: a
 * different but algebraically equal order changes the last bits, and the
 * second-order path amplifies such differences over steps.
 */
#include "fvh_oracle.h"

#include <math.h>
#include <stdlib.h>
#include <string.h>

#define V 4

/* std::min / std::max as the reference's libstdc++ defines them:
 * min(a,b) = (b < a) ? b : a ; max(a,b) = (a < b) ? b : a. */
static inline double smin(double a, double b) { return (b < a) ? b : a; }
static inline double smax(double a, double b) { return (a < b) ? b : a; }
/* smoothed kinks (extension, no reference counterpart): |x|_w, min_w, max_w */
static inline double kabs(double x, double w) { return x * x / sqrt(x * x + w * w); }
static inline double kmin(double a, double b, double w) { return 0.5 * (a + b - kabs(a - b, w)); }
static inline double kmax(double a, double b, double w) { return 0.5 * (a + b + kabs(a - b, w)); }

/* ---------------------------------------------------------------------- */
/* edge lookup: (min node, max node) -> edge id  (mesh.cpp:103-126 edgesRef) */
typedef struct { uint64_t* keys; int32_t* vals; uint64_t mask; } emap;

static uint64_t mix64(uint64_t z) {
    z = (z ^ (z >> 30)) * 0xBF58476D1CE4E5B9ULL;
    z = (z ^ (z >> 27)) * 0x94D049BB133111EBULL;
    return z ^ (z >> 31);
}
static int emap_init(emap* h, int64_t n) {
    uint64_t cap = 16;
    while (cap < (uint64_t)(2 * n + 16)) cap <<= 1;
    h->keys = (uint64_t*)malloc(cap * sizeof(uint64_t));
    h->vals = (int32_t*)malloc(cap * sizeof(int32_t));
    if (!h->keys || !h->vals) return -1;
    for (uint64_t i = 0; i < cap; ++i) h->keys[i] = ~0ULL;
    h->mask = cap - 1;
    return 0;
}
static void emap_free(emap* h) { free(h->keys); free(h->vals); }
static uint64_t ekey(int32_t a, int32_t b) {
    uint32_t lo = (uint32_t)(a < b ? a : b), hi = (uint32_t)(a < b ? b : a);
    return ((uint64_t)lo << 32) | hi;
}
static int32_t emap_find(const emap* h, uint64_t k) {
    uint64_t i = mix64(k) & h->mask;
    while (h->keys[i] != ~0ULL) {
        if (h->keys[i] == k) return h->vals[i];
        i = (i + 1) & h->mask;
    }
    return -1;
}
static void emap_put(emap* h, uint64_t k, int32_t v) {
    uint64_t i = mix64(k) & h->mask;
    while (h->keys[i] != ~0ULL) i = (i + 1) & h->mask;
    h->keys[i] = k;
    h->vals[i] = v;
}

/* ---------------------------------------------------------------------- */
orc_mesh* orc_mesh_build(int32_t n_nodes, const double* x, const double* y,
                         int32_t n_cells, const int32_t* cells,
                         int32_t n_bnd, const int32_t* bnd, const int32_t* bnd_bc,
                         int* err) {
    *err = 0;
    orc_mesh* m = (orc_mesh*)calloc(1, sizeof(orc_mesh));
    const int64_t NC = (int64_t)n_cells + n_bnd;
    const int64_t EMAX = 4 * (int64_t)n_cells + 4;
    m->n_nodes = n_nodes;
    m->x = (double*)malloc(sizeof(double) * (n_nodes ? n_nodes : 1));
    m->y = (double*)malloc(sizeof(double) * (n_nodes ? n_nodes : 1));
    memcpy(m->x, x, sizeof(double) * n_nodes);
    memcpy(m->y, y, sizeof(double) * n_nodes);
    m->cell_nodes = (int32_t*)calloc(4 * NC + 4, sizeof(int32_t));
    m->cell_tri = (uint8_t*)calloc(NC + 1, 1);
    m->cell_area = (double*)calloc(NC + 1, sizeof(double));
    m->cell_cx = (double*)calloc(NC + 1, sizeof(double));
    m->cell_cy = (double*)calloc(NC + 1, sizeof(double));
    m->edge_nodes = (int32_t*)malloc(sizeof(int32_t) * 2 * EMAX);
    m->edge_cells = (int32_t*)malloc(sizeof(int32_t) * 2 * EMAX);
    m->bnd_edge = (int32_t*)malloc(sizeof(int32_t) * (n_bnd + 1));
    m->bnd_bc = (int32_t*)malloc(sizeof(int32_t) * (n_bnd + 1));

    /* cells in file order (mesh.cpp:528-560); 4th node 0 for triangles */
    for (int32_t c = 0; c < n_cells; ++c) {
        int tri = cells[4 * c + 3] < 0;
        m->cell_tri[c] = (uint8_t)tri;
        for (int k = 0; k < 3; ++k) m->cell_nodes[4 * c + k] = cells[4 * c + k];
        m->cell_nodes[4 * c + 3] = tri ? 0 : cells[4 * c + 3];
    }

    /* edges in order of first appearance (mesh.cpp:130-154, 821-823) */
    emap h;
    emap_init(&h, EMAX);
    int32_t E = 0;
    for (int32_t c = 0; c < n_cells; ++c) {
        int sz = m->cell_tri[c] ? 3 : 4;
        for (int i = 0; i < sz; ++i) {
            int j = (i < sz - 1) ? i + 1 : 0;
            int32_t a = m->cell_nodes[4 * c + i], b = m->cell_nodes[4 * c + j];
            uint64_t k = ekey(a, b);
            if (emap_find(&h, k) < 0) {
                m->edge_nodes[2 * E] = a;
                m->edge_nodes[2 * E + 1] = b;
                m->edge_cells[2 * E] = c;
                m->edge_cells[2 * E + 1] = 0;
                emap_put(&h, k, E);
                ++E;
            }
        }
    }
    m->n_edges = E;
    m->n_real = n_cells;

    /* second owner (mesh.cpp:157-187) */
    for (int32_t c = 0; c < n_cells; ++c) {
        int32_t ce[4];
        const int32_t* cn = &m->cell_nodes[4 * c];
        ce[0] = emap_find(&h, ekey(cn[0], cn[1]));
        ce[1] = emap_find(&h, ekey(cn[1], cn[2]));
        if (m->cell_tri[c]) {
            ce[2] = emap_find(&h, ekey(cn[2], cn[0]));
            ce[3] = -1;
        } else {
            ce[2] = emap_find(&h, ekey(cn[2], cn[3]));
            ce[3] = emap_find(&h, ekey(cn[3], cn[0]));
        }
        for (int k = 0; k < 4; ++k) {
            int32_t e = ce[k];
            if (e != -1 && c != m->edge_cells[2 * e]) m->edge_cells[2 * e + 1] = c;
        }
    }

    /* geometry (mesh.cpp:190-260) */
    m->edge_len = (double*)malloc(sizeof(double) * (E + 1));
    m->edge_nx = (double*)malloc(sizeof(double) * (E + 1));
    m->edge_ny = (double*)malloc(sizeof(double) * (E + 1));
    m->edge_cx = (double*)malloc(sizeof(double) * (E + 1));
    m->edge_cy = (double*)malloc(sizeof(double) * (E + 1));
    for (int32_t c = 0; c < n_cells; ++c) {
        double cx = 0.0, cy = 0.0;
        int n = m->cell_tri[c] ? 3 : 4;
        for (int j = 0; j < n; ++j) {
            cx += x[m->cell_nodes[4 * c + j]] / ((double)n);
            cy += y[m->cell_nodes[4 * c + j]] / ((double)n);
        }
        m->cell_cx[c] = cx;
        m->cell_cy[c] = cy;
    }
    for (int32_t e = 0; e < E; ++e) {
        int32_t n0 = m->edge_nodes[2 * e], n1 = m->edge_nodes[2 * e + 1];
        m->edge_cx[e] = (x[n1] + x[n0]) * 0.5;
        m->edge_cy[e] = (y[n1] + y[n0]) * 0.5;
        double dex = x[n1] - x[n0];
        double dey = y[n1] - y[n0];
        double l = sqrt(dex * dex + dey * dey);
        m->edge_len[e] = l;
        m->edge_nx[e] = -dey / l;
        m->edge_ny[e] = dex / l;
        int32_t c0 = m->edge_cells[2 * e];
        double dot = m->edge_nx[e] * (m->edge_cx[e] - m->cell_cx[c0])
                   + m->edge_ny[e] * (m->edge_cy[e] - m->cell_cy[c0]);
        if (dot < 0) {
            m->edge_nx[e] *= -1.;
            m->edge_ny[e] *= -1.;
        }
    }
    for (int32_t c = 0; c < n_cells; ++c) {
        const int32_t* cn = &m->cell_nodes[4 * c];
        double x1 = x[cn[0]], x2 = x[cn[1]], x3 = x[cn[2]];
        double y1 = y[cn[0]], y2 = y[cn[1]], y3 = y[cn[2]];
        if (m->cell_tri[c]) {
            m->cell_area[c] = 0.5 * fabs(x1 * (y2 - y3) + x2 * (y3 - y1) + x3 * (y1 - y2));
        } else {
            double x4 = x[cn[3]], y4 = y[cn[3]];
            m->cell_area[c] = 0.5 * fabs(x1 * (y2 - y3) + x2 * (y3 - y1) + x3 * (y1 - y2))
                            + 0.5 * fabs(x1 * (y3 - y4) + x3 * (y4 - y1) + x4 * (y1 - y3));
        }
    }

    /* boundary ghost cells (mesh.cpp:623-669) */
    int32_t nc = n_cells;
    for (int32_t b = 0; b < n_bnd; ++b) {
        int32_t a = bnd[2 * b], bb = bnd[2 * b + 1];
        int32_t e = emap_find(&h, ekey(a, bb));
        if (e < 0) { *err = 1; emap_free(&h); orc_mesh_free(m); return NULL; }
        int32_t cell = m->edge_cells[2 * e];
        double area = m->cell_area[cell];
        double dx = m->edge_cx[e] - m->cell_cx[cell];
        double dy = m->edge_cy[e] - m->cell_cy[cell];
        double dist = sqrt(dx * dx + dy * dy);
        m->cell_area[nc] = area;
        m->cell_cx[nc] = m->edge_cx[e] + dist * m->edge_nx[e];
        m->cell_cy[nc] = m->edge_cy[e] + dist * m->edge_ny[e];
        m->cell_tri[nc] = 1;
        m->cell_nodes[4 * nc] = a < bb ? a : bb;
        m->cell_nodes[4 * nc + 1] = a < bb ? bb : a;
        m->edge_cells[2 * e + 1] = nc;
        m->bnd_edge[b] = e;
        m->bnd_bc[b] = bnd_bc[b];
        ++nc;
    }
    m->n_cells = nc;
    m->n_bnd = n_bnd;
    emap_free(&h);
    return m;
}

void orc_mesh_free(orc_mesh* m) {
    if (!m) return;
    free(m->x); free(m->y); free(m->edge_nodes); free(m->edge_cells);
    free(m->edge_len); free(m->edge_nx); free(m->edge_ny); free(m->edge_cx); free(m->edge_cy);
    free(m->cell_nodes); free(m->cell_tri); free(m->cell_area); free(m->cell_cx); free(m->cell_cy);
    free(m->bnd_edge); free(m->bnd_bc);
    free(m);
}

/* ---------------------------------------------------------------------- */
/* plugin kinds                                                            */

static inline double pressure(double g, const double* q) {
    /* calc_p: forwardStep/main.cpp:37-39, naca/main.cpp:36-38 */
    return (g - 1) * (q[3] - 0.5 / q[0] * (q[1] * q[1] + q[2] * q[2]));
}

/* Laminar stresses and heat flux of the viscous examples, subtracted from
 * the Roe flux: flatPlate/main.cpp:193-224 with its helpers calc_grad_p
 * (:43-57), calc_grad_t (:65-78), calc_grad_u (:80-85), calc_grad_v
 * (:87-92), all on the mean face state; nacaTransient/main.cpp holds the
 * same code. */
static void viscous_terms(const orc_config* cf, double* f, const double* qi, const double* qj,
                          const double* gx, const double* gy, const double* n) {
    const double g = cf->gamma, mu = cf->mu;
    double q[4];
    for (int k = 0; k < 4; ++k) q[k] = 0.5 * (qi[k] + qj[k]);
    double gp[2];
    gp[0] = 2. * gx[3];
    gp[0] -= gx[1] * q[1] / q[0] + q[1] * (gx[1] * q[0] - gx[0] * q[1]) / (q[0] * q[0]);
    gp[0] -= gx[2] * q[2] / q[0] + q[2] * (gx[2] * q[0] - gx[0] * q[2]) / (q[0] * q[0]);
    gp[0] *= 0.5 * (g - 1);
    gp[1] = 2. * gy[3];
    gp[1] -= gy[1] * q[1] / q[0] + q[1] * (gy[1] * q[0] - gy[0] * q[1]) / (q[0] * q[0]);
    gp[1] -= gy[2] * q[2] / q[0] + q[2] * (gy[2] * q[0] - gy[0] * q[2]) / (q[0] * q[0]);
    gp[1] *= 0.5 * (g - 1);
    const double p = pressure(g, q);
    double gt[2];
    gt[0] = (1. / cf->r) * ((gp[0] * q[0] - gx[0] * p) / (q[0] * q[0]));
    gt[1] = (1. / cf->r) * ((gp[1] * q[0] - gy[0] * p) / (q[0] * q[0]));
    double gu[2], gv[2];
    gu[0] = (q[0] * gx[1] - q[1] * gx[0]) / (q[0] * q[0]);
    gu[1] = (q[0] * gy[1] - q[1] * gy[0]) / (q[0] * q[0]);
    gv[0] = (q[0] * gx[2] - q[2] * gx[0]) / (q[0] * q[0]);
    gv[1] = (q[0] * gy[2] - q[2] * gy[0]) / (q[0] * q[0]);
    const double div_v = gu[0] + gv[1];
    const double txx = 2. * mu * (gu[0] - div_v / 3.);
    const double tyy = 2. * mu * (gv[1] - div_v / 3.);
    const double txy = mu * (gu[1] + gv[0]);
    const double k = cf->cp * mu / cf->pr;
    double ph[2];
    ph[0] = q[1] / q[0] * txx + q[2] / q[0] * txy + k * gt[0];
    ph[1] = q[1] / q[0] * txy + q[2] / q[0] * tyy + k * gt[1];
    f[1] -= n[0] * txx + n[1] * txy;
    f[2] -= n[0] * txy + n[1] * tyy;
    f[3] -= n[0] * ph[0] + n[1] * ph[1];
}

static void roe_flux(const orc_config* cf, double* f, const double* qi, const double* qj, const double* n);

void orc_flux(const orc_config* cf, double* f, const double* qi, const double* qj,
              const double* n) {
    const double z[4] = {0., 0., 0., 0.};
    orc_flux_grad(cf, f, qi, qj, z, z, n);
}

void orc_flux_grad(const orc_config* cf, double* f, const double* qi, const double* qj,
                   const double* gx, const double* gy, const double* n) {
    roe_flux(cf, f, qi, qj, n);
    if (cf->flux == ORC_FLUX_ROE_VISCOUS) viscous_terms(cf, f, qi, qj, gx, gy, n);
}

/* gradient_for_diffusion (explicit.cpp:70-109) */
void orc_face_gradient(double* gx, double* gy, const double* gxi, const double* gyi,
                       const double* gxj, const double* gyj, const double* qi, const double* qj,
                       const double* ci, const double* cj) {
    double t[2];
    t[0] = ci[0] - cj[0];
    t[1] = ci[1] - cj[1];
    const double l = sqrt(t[0] * t[0] + t[1] * t[1]);
    t[0] = t[0] / l;
    t[1] = t[1] / l;
    for (int k = 0; k < 4; ++k) {
        const double dir = (qi[k] - qj[k]) / l;
        const double xb = (gxi[k] + gxj[k]) * 0.5;
        const double yb = (gyi[k] + gyj[k]) * 0.5;
        const double dot = xb * t[0] + yb * t[1];
        gx[k] = xb - (dot - dir) * t[0];
        gy[k] = yb - (dot - dir) * t[1];
    }
}

/* Roe flux with the entropy cure: shockTube/main.cpp:68-135 (SPLIT: central
 * part accumulated per side, as also naca/main.cpp:71-137 and the viscous
 * examples' flatPlate/main.cpp:126-191) and forwardStep/main.cpp:72-135
 * (JOINED: central part in one sum). */
static void roe_flux(const orc_config* cf, double* f, const double* qi, const double* qj, const double* n) {
    const double g = cf->gamma;
    double pi = pressure(g, qi);
    double pj = pressure(g, qj);
    double Vi = (qi[1] * n[0] + qi[2] * n[1]) / qi[0];
    double Vj = (qj[1] * n[0] + qj[2] * n[1]) / qj[0];
    if (cf->flux == ORC_FLUX_ROE_JOINED) {
        f[0] = (qi[0] * Vi + qj[0] * Vj) * 0.5;
        f[1] = (qi[1] * Vi + pi * n[0] + qj[1] * Vj + pj * n[0]) * 0.5;
        f[2] = (qi[2] * Vi + pi * n[1] + qj[2] * Vj + pj * n[1]) * 0.5;
        f[3] = ((qi[3] + pi) * Vi + (qj[3] + pj) * Vj) * 0.5;
    } else {
        f[0] = qi[0] * Vi;
        f[1] = qi[1] * Vi + pi * n[0];
        f[2] = qi[2] * Vi + pi * n[1];
        f[3] = (qi[3] + pi) * Vi;
        f[0] += qj[0] * Vj;
        f[1] += qj[1] * Vj + pj * n[0];
        f[2] += qj[2] * Vj + pj * n[1];
        f[3] += (qj[3] + pj) * Vj;
        for (int i = 0; i < 4; ++i) f[i] *= 0.5;
    }
    const double pL = pi, pR = pj;
    const double uL = qi[1] / qi[0];
    const double uR = qj[1] / qj[0];
    const double vL = qi[2] / qi[0];
    const double vR = qj[2] / qj[0];
    const double srhoL = sqrt(qi[0]);
    const double srhoR = sqrt(qj[0]);
    const double rho = srhoR * srhoL;
    const double u = (uL * srhoL + uR * srhoR) / (srhoL + srhoR);
    const double v = (vL * srhoL + vR * srhoR) / (srhoL + srhoR);
    const double h = ((qi[3] + pL) / qi[0] * srhoL + (qj[3] + pR) / qj[0] * srhoR) / (srhoL + srhoR);
    const double q2 = u * u + v * v;
    const double c = sqrt((g - 1.) * (h - 0.5 * q2));
    const double Vn = u * n[0] + v * n[1];
    const double VR = uR * n[0] + vR * n[1];
    const double VL = uL * n[0] + vL * n[1];
    const double w = cf->kwv;
    const double lcm = cf->kinks ? kabs(kmin(Vn - c, VL - c, w), w) : fabs(smin(Vn - c, VL - c));
    const double lc = cf->kinks ? kabs(Vn, w) : fabs(Vn);
    const double lcp = cf->kinks ? kabs(kmax(Vn + c, VR + c, w), w) : fabs(smax(Vn + c, VR + c));
    const double kF1 = lcm * ((pR - pL) - rho * c * (VR - VL)) / (2. * c * c);
    const double kF234_0 = lc * ((qj[0] - qi[0]) - (pR - pL) / (c * c));
    const double kF234_1 = lc * rho;
    const double kF5 = lcp * ((pR - pL) + rho * c * (VR - VL)) / (2. * c * c);
    f[0] -= 0.5 * (kF1 + kF234_0 + kF5);
    f[1] -= 0.5 * (kF1 * (u - c * n[0]) + kF234_0 * u + kF234_1 * (uR - uL - (VR - VL) * n[0]) + kF5 * (u + c * n[0]));
    f[2] -= 0.5 * (kF1 * (v - c * n[1]) + kF234_0 * v + kF234_1 * (vR - vL - (VR - VL) * n[1]) + kF5 * (v + c * n[1]));
    f[3] -= 0.5 * (kF1 * (h - c * Vn) + kF234_0 * q2 * 0.5 + kF234_1 * (u * (uR - uL) + v * (vR - vL) - Vn * (VR - VL)) + kF5 * (h + c * Vn));
}

/* limiter_func kinds: zero (shockTube/main.cpp:60-62), Michalak
 * (forwardStep/main.cpp:57-66), minmod (burgers/main.cpp:51-53),
 * Venkatakrishnan (r^2+2r)/(r^2+r+2) (no reference example). */
double orc_phi(const orc_config* cf, double y) {
    switch (cf->phi) {
    case ORC_PHI_ZERO:
        return 0.;
    case ORC_PHI_MICHALAK: {
        const double yt = cf->phi_yt;
        if (y >= yt) return 1.0;
        const double a = 1.0 / (yt * yt) - 2.0 / (yt * yt * yt);
        const double b = -3.0 / 2.0 * a * yt - 0.5 / yt;
        return a * y * y * y + b * y * y + y;
    }
    case ORC_PHI_VENKAT:
        return (y * y + 2.0 * y) / (y * y + y + 2.0);
    default:
        return smax(0., smin(1., y));
    }
}

/* boundary kinds: copy (shockTube/main.cpp:158-163), mirror wall
 * (forwardStep/main.cpp:262-269), characteristic inlet/outlet
 * (forwardStep/main.cpp:194-261), farfield (naca/main.cpp:192-259), no-slip
 * wall (flatPlate/main.cpp:361-366) and the viscous examples' farfield
 * (flatPlate/main.cpp:286-353: b[0] where naca has b[1]). */
void orc_bc(const orc_config* cf, int slot, double* b, const double* q, const double* n) {
    const int kind = cf->bc_kind[slot];
    const double g = cf->gamma;
    if (kind == ORC_BC_COPY) {
        b[0] = q[0]; b[1] = q[1]; b[2] = q[2]; b[3] = q[3];
        return;
    }
    if (kind == ORC_BC_MIRROR) {
        b[0] = q[0];
        b[1] = q[1] - 2.0 * n[0] * (n[0] * q[1] + n[1] * q[2]);
        b[2] = q[2] - 2.0 * n[1] * (n[0] * q[1] + n[1] * q[2]);
        b[3] = q[3];
        return;
    }
    if (kind == ORC_BC_NOSLIP) {
        b[0] = q[0];
        b[1] = -q[1];
        b[2] = -q[2];
        b[3] = q[3];
        return;
    }
    const double* P = cf->bc_param[slot];
    double bv[4];
    bv[0] = P[0];
    bv[1] = bv[0] * P[1];
    bv[2] = bv[0] * P[2];
    bv[3] = P[3] / (g - 1) + 0.5 / bv[0] * (bv[1] * bv[1] + bv[2] * bv[2]);
    double U0 = q[1] / q[0];
    double U1 = q[2] / q[0];
    double u_norm = sqrt(U0 * U0 + U1 * U1);
    double u_dot_n = U0 * n[0] + U1 * n[1];
    double p = pressure(g, q);
    double c = sqrt(g * p / q[0]);
    double mach = u_norm / c;
    if (mach > 1) {
        if (u_dot_n < 0) {
            b[0] = bv[0]; b[1] = bv[1]; b[2] = bv[2]; b[3] = bv[3];
        } else {
            b[0] = q[0]; b[1] = q[1]; b[2] = q[2]; b[3] = q[3];
        }
        return;
    }
    double pa = pressure(g, bv);
    double rhoa = bv[0];
    double ua = bv[1] / bv[0];
    double va = bv[2] / bv[0];
    double pd = p;
    double rhod = q[0];
    double ud = q[1] / q[0];
    double vd = q[2] / q[0];
    double rho0 = q[0];
    double c0 = c;
    if (u_dot_n < 0) {
        if (kind == ORC_BC_INLET_OUTLET) {
            double pb = pd;
            b[0] = bv[0]; b[1] = bv[1]; b[2] = bv[2];
            b[3] = pb / (g - 1.) + 0.5 / b[0] * (b[1] * b[1] + b[2] * b[2]);
        } else {
            double pb = 0.5 * (pa + pd - rho0 * c0 * (n[0] * (ua - ud) + n[1] * (va - vd)));
            b[0] = rhoa + (pb - pa) / (c0 * c0);
            b[1] = b[0] * (ua - n[0] * (pa - pb) / (rho0 * c0));
            /* b[1] as naca/main.cpp:246, b[0] as flatPlate/main.cpp:340 */
            b[2] = (kind == ORC_BC_FARFIELD_V ? b[0] : b[1]) * (va - n[1] * (pa - pb) / (rho0 * c0));
            b[3] = pb / (g - 1) + 0.5 / b[0] * (b[1] * b[1] + b[2] * b[2]);
        }
    } else {
        double pb = pa;
        b[0] = rhod + (pb - pd) / (c0 * c0);
        b[1] = b[0] * (ud + n[0] * (pd - pb) / (rho0 * c0));
        /* va in both spellings; b[1] as naca/main.cpp:253, b[0] as flatPlate/main.cpp:347 */
        b[2] = (kind == ORC_BC_FARFIELD_V ? b[0] : b[1]) * (va + n[1] * (pd - pb) / (rho0 * c0));
        b[3] = pb / (g - 1.) + 0.5 / b[0] * (b[1] * b[1] + b[2] * b[2]);
    }
}

/* ---------------------------------------------------------------------- */
/* phases                                                                  */

/* calc_dt kinds: fixed (shockTube/main.cpp:140-147), summed CFL
 * (forwardStep/main.cpp:140-185), edge-minimum local (naca/main.cpp:142-183).
 * dt holds one value per cell (the reference stores 4 identical copies). */
void orc_calc_dt(const orc_mesh* m, const orc_config* cf, const double* q, double* dt) {
    const int32_t N = m->n_cells;
    const double g = cf->gamma;
    if (cf->dt_kind == ORC_DT_FIXED) {
        for (int32_t i = 0; i < N; ++i) dt[i] = cf->dt_fixed;
        return;
    }
    if (cf->dt_kind == ORC_DT_SUM_CFL) {
        for (int32_t i = 0; i < N; ++i) dt[i] = 0.;
        for (int32_t e = 0; e < m->n_edges; ++e) {
            int32_t i = m->edge_cells[2 * e], j = m->edge_cells[2 * e + 1];
            double le = m->edge_len[e], n0 = m->edge_nx[e], n1 = m->edge_ny[e];
            const double* qi = &q[V * i];
            const double* qj = &q[V * j];
            double ci = sqrt(pressure(g, qi) * g / qi[0]);
            double cj = sqrt(pressure(g, qj) * g / qj[0]);
            double ei, ej, center;
            if (cf->kinks) {
                ei = ci + kabs(qi[1] / qi[0] * n0 + qi[2] / qi[0] * n1, cf->kwv);
                ej = cj + kabs(qj[1] / qj[0] * n0 + qj[2] / qj[0] * n1, cf->kwv);
                center = kmax(ei, ej, cf->kwv);
            } else {
                ei = ci + fabs(qi[1] / qi[0] * n0 + qi[2] / qi[0] * n1);
                ej = cj + fabs(qj[1] / qj[0] * n0 + qj[2] / qj[0] * n1);
                center = smax(ei, ej);
            }
            dt[i] += center * le;
            dt[j] += center * le;
        }
        for (int32_t i = 0; i < N; ++i) dt[i] = cf->cfl * m->cell_area[i] / dt[i];
        return;
    }
    if (cf->dt_kind == ORC_DT_EDGE_MIN_VISCOUS || cf->dt_kind == ORC_DT_SUM_VISCOUS) {
        /* flatPlate/main.cpp:230-278 (local minimum over edges) and
         * nacaTransient/main.cpp:234-286 (summed): convective plus C = 2
         * times the viscous spectral radius */
        const int sum = cf->dt_kind == ORC_DT_SUM_VISCOUS;
        const double C = 2;
        for (int32_t i = 0; i < N; ++i) dt[i] = sum ? 0. : 1.0;
        for (int32_t e = 0; e < m->n_edges; ++e) {
            int32_t i = m->edge_cells[2 * e], j = m->edge_cells[2 * e + 1];
            double le = m->edge_len[e], n0 = m->edge_nx[e], n1 = m->edge_ny[e];
            const double* qi = &q[V * i];
            const double* qj = &q[V * j];
            double ci = sqrt(pressure(g, qi) * g / qi[0]);
            double cj = sqrt(pressure(g, qj) * g / qj[0]);
            double evi = smax(4. / (3. * qi[0]), g / qi[0]) * (cf->mu / cf->pr) * le * le / m->cell_area[i];
            double evj = smax(4. / (3. * qj[0]), g / qj[0]) * (cf->mu / cf->pr) * le * le / m->cell_area[j];
            if (sum) {
                double eci = ci + fabs(qi[1] / qi[0] * n0 + qi[2] / qi[0] * n1);
                double ecj = cj + fabs(qj[1] / qj[0] * n0 + qj[2] / qj[0] * n1);
                double center = smax(eci + C * evi, ecj + C * evj);
                dt[i] += center * le;
                dt[j] += center * le;
            } else {
                double eci = (ci + fabs(n0 * qi[1] / qi[0]) + fabs(n1 * qi[2] / qi[0])) * le;
                double ecj = (cj + fabs(n0 * qj[1] / qj[0]) + fabs(n1 * qj[2] / qj[0])) * le;
                dt[i] = smin(dt[i], cf->cfl * m->cell_area[i] / (eci + C * evi));
                dt[j] = smin(dt[j], cf->cfl * m->cell_area[j] / (ecj + C * evj));
            }
        }
        if (sum)
            for (int32_t i = 0; i < N; ++i) dt[i] = cf->cfl * m->cell_area[i] / dt[i];
        return;
    }
    for (int32_t i = 0; i < N; ++i) dt[i] = 1.0;
    for (int32_t e = 0; e < m->n_edges; ++e) {
        int32_t i = m->edge_cells[2 * e], j = m->edge_cells[2 * e + 1];
        double le = m->edge_len[e], n0 = m->edge_nx[e], n1 = m->edge_ny[e];
        const double* qi = &q[V * i];
        const double* qj = &q[V * j];
        double ci = sqrt(pressure(g, qi) * g / qi[0]);
        double cj = sqrt(pressure(g, qj) * g / qj[0]);
        double ei = (ci + fabs(qi[1] / qi[0])) * n0 + (ci + fabs(qi[2] / qi[0])) * n1;
        double ej = (cj + fabs(qj[1] / qj[0])) * n0 + (cj + fabs(qj[2] / qj[0])) * n1;
        double dti = cf->cfl * (m->cell_area[i] / le) / fabs(ei);
        double dtj = cf->cfl * (m->cell_area[j] / le) / fabs(ej);
        dt[i] = smin(dt[i], dti);
        dt[j] = smin(dt[j], dtj);
    }
}

/* min_dt (explicit.cpp:555-564) */
void orc_min_dt(const orc_mesh* m, double* dt) {
    double mn = dt[0];
    for (int32_t i = 0; i < m->n_cells; ++i) mn = smin(mn, dt[i]);
    for (int32_t i = 0; i < m->n_cells; ++i) dt[i] = mn;
}

/* Extension (smoothed kinks): the p-norm (sum_i dt_i^-128)^(-1/128) as
 * m / S^(1/128), m = min_dt's value, S = sum_i (m/dt_i)^128; each term in
 * [0, 1] added in 64.64 fixed point (order-free), S^(1/128) by seven square
 * roots; a NaN term makes the step NaN. */
void orc_min_dt_soft(const orc_mesh* m, double* dt) {
    const int32_t N = m->n_cells;
    double mn = dt[0];
    for (int32_t i = 0; i < N; ++i) mn = smin(mn, dt[i]);
    unsigned __int128 S = 0;
    for (int32_t i = 0; i < N; ++i) {
        double y = mn / dt[i];
        if (!(y >= 0.0 && y <= 1.0)) { S += (unsigned __int128)1 << 100; continue; }
        for (int k = 0; k < 7; ++k) y = y * y;
        if (y >= 1.0) S += (unsigned __int128)1 << 64;
        else S += (unsigned __int128)(uint64_t)(y * 18446744073709551616.0);
    }
    double root = (double)(uint64_t)(S >> 64) + (double)(uint64_t)S * 0x1p-64;
    for (int k = 0; k < 7; ++k) root = sqrt(root);
    const double v = (S >> 100) != 0 ? NAN : mn / root;
    for (int32_t i = 0; i < N; ++i) dt[i] = v;
}

/* calc_gradients (explicit.cpp:112-171) */
void orc_calc_gradients(const orc_mesh* m, const double* q, double* gx, double* gy) {
    const int64_t n = (int64_t)V * m->n_cells;
    for (int64_t i = 0; i < n; ++i) { gx[i] = 0.; gy[i] = 0.; }
    for (int32_t e = 0; e < m->n_edges; ++e) {
        int32_t i = m->edge_cells[2 * e], j = m->edge_cells[2 * e + 1];
        double nx = m->edge_nx[e], ny = m->edge_ny[e], le = m->edge_len[e];
        double dxif = m->edge_cx[e] - m->cell_cx[i];
        double dyif = m->edge_cy[e] - m->cell_cy[i];
        double dif = sqrt(dxif * dxif + dyif * dyif);
        double dxij = m->cell_cx[i] - m->cell_cx[j];
        double dyij = m->cell_cy[i] - m->cell_cy[j];
        double dij = sqrt(dxij * dxij + dyij * dyij);
        double gf = dif / dij;
        if (i != j) {
            double f[V];
            for (int k = 0; k < V; ++k) f[k] = (q[V * i + k] * (1.0 - gf) + q[V * j + k] * gf) * le;
            for (int k = 0; k < V; ++k) {
                gx[V * i + k] += f[k] * nx;
                gy[V * i + k] += f[k] * ny;
                gx[V * j + k] -= f[k] * nx;
                gy[V * j + k] -= f[k] * ny;
            }
        }
    }
    for (int32_t i = 0; i < m->n_real; ++i) {
        const double invA = 1. / m->cell_area[i];
        for (int k = 0; k < V; ++k) { gx[V * i + k] *= invA; gy[V * i + k] *= invA; }
    }
    for (int32_t i = m->n_real; i < m->n_cells; ++i)
        for (int k = 0; k < V; ++k) { gx[V * i + k] = 0.; gy[V * i + k] = 0.; }
}

/* the sigma blend of calc_limiters over y = d2 / K3a - 1 in (0, 1]
 * (explicit.cpp:230-233); with smoothed kinks (extension) 1 - the
 * quartic-order smoothstep, four times differentiable at both ends */
double orc_sigma(const orc_config* cf, double yy) {
    if (cf->kinks) {
        double y2 = yy * yy;
        return 1.0 - y2 * y2 * yy * (126.0 + yy * (-420.0 + yy * (540.0 + yy * (-315.0 + 70.0 * yy))));
    }
    return 2.0 * yy * yy * yy - 3.0 * yy * yy + 1.0;
}

/* calc_limiters (explicit.cpp:174-258) */
void orc_calc_limiters(const orc_mesh* m, const orc_config* cf, const double* q,
                       const double* gx, const double* gy,
                       double* lim, double* qmin, double* qmax) {
    const int64_t n = (int64_t)V * m->n_cells;
    for (int64_t i = 0; i < n; ++i) { lim[i] = 1.; qmin[i] = q[i]; qmax[i] = q[i]; }
    for (int32_t e = 0; e < m->n_edges; ++e) {
        int32_t i = m->edge_cells[2 * e], j = m->edge_cells[2 * e + 1];
        for (int k = 0; k < V; ++k) {
            if (cf->kinks) {
                const double w = cf->kw[k];
                qmin[V * i + k] = kmin(qmin[V * i + k], q[V * j + k], w);
                qmin[V * j + k] = kmin(qmin[V * j + k], q[V * i + k], w);
                qmax[V * i + k] = kmax(qmax[V * i + k], q[V * j + k], w);
                qmax[V * j + k] = kmax(qmax[V * j + k], q[V * i + k], w);
                continue;
            }
            qmin[V * i + k] = smin(qmin[V * i + k], q[V * j + k]);
            qmin[V * j + k] = smin(qmin[V * j + k], q[V * i + k]);
            qmax[V * i + k] = smax(qmax[V * i + k], q[V * j + k]);
            qmax[V * j + k] = smax(qmax[V * j + k], q[V * i + k]);
        }
    }
    const double tol = 1e-15;
    for (int32_t e = 0; e < m->n_edges; ++e) {
        int32_t ids[2] = {m->edge_cells[2 * e], m->edge_cells[2 * e + 1]};
        for (int s = 0; s < 2; ++s) {
            int32_t id = ids[s];
            if (id >= m->n_real) continue;
            double dx = m->edge_cx[e] - m->cell_cx[id];
            double dy = m->edge_cy[e] - m->cell_cy[id];
            double sqrt_area = sqrt(m->cell_area[id]);
            for (int k = 0; k < V; ++k) {
                double dqg = gx[V * id + k] * dx + gy[V * id + k] * dy;
                double dmax = qmax[V * id + k] - q[V * id + k];
                double dmin = qmin[V * id + k] - q[V * id + k];
                const double Ka = cf->K * sqrt_area;
                const double K3a = Ka * Ka * Ka;
                const double d2 = (dmax - dmin) * (dmax - dmin);
                double sig;
                if (d2 <= K3a) {
                    sig = 1.;
                } else if (d2 <= 2 * K3a) {
                    sig = orc_sigma(cf, d2 / K3a - 1.0);
                } else {
                    sig = 0.;
                }
                double l = 1.0;
                if (sig < 1.0 && cf->kinks) {
                    /* smoothed kinks: the extremum by a blend on the sign of
                     * dqg; Venkatakrishnan's phi(d/dqg) without the quotient,
                     * regularised with w^2 as in his original form */
                    const double w = cf->kw[k];
                    const double t = dqg / sqrt(dqg * dqg + w * w);
                    const double d = 0.5 * (dmax + dmin) + 0.5 * (dmax - dmin) * t;
                    l = (d * d + 2.0 * d * dqg + w * w) / (d * d + d * dqg + 2.0 * dqg * dqg + w * w);
                } else if (sig < 1.0) {
                    if (dqg > tol) l = orc_phi(cf, dmax / dqg);
                    else if (dqg < -tol) l = orc_phi(cf, dmin / dqg);
                    else l = 1.0;
                }
                l = sig + (1.0 - sig) * l;
                lim[V * id + k] = cf->kinks ? kmin(lim[V * id + k], l, cf->kwl) : smin(lim[V * id + k], l);
            }
        }
    }
}

/* calc_time_derivatives (explicit.cpp:261-364), without the viscous term */
void orc_calc_time_derivatives(const orc_mesh* m, const orc_config* cf, const double* q,
                               const double* gx, const double* gy, const double* lim,
                               double* qt) {
    const int64_t n = (int64_t)V * m->n_cells;
    for (int64_t i = 0; i < n; ++i) qt[i] = 0.;
    for (int32_t e = 0; e < m->n_edges; ++e) {
        int32_t i = m->edge_cells[2 * e], j = m->edge_cells[2 * e + 1];
        double le = m->edge_len[e];
        double nn[2] = {m->edge_nx[e], m->edge_ny[e]};
        double cx = m->edge_cx[e], cy = m->edge_cy[e];
        double di0 = cx - m->cell_cx[i], di1 = cy - m->cell_cy[i];
        double dj0 = cx - m->cell_cx[j], dj1 = cy - m->cell_cy[j];
        double qi[V], qj[V], f[V];
        if (cf->lin) {
            for (int k = 0; k < V; ++k) {
                int64_t ki = V * (int64_t)i + k, kj = V * (int64_t)j + k;
                qi[k] = q[ki] + (gx[ki] * di0 + gy[ki] * di1) * lim[ki];
                qj[k] = q[kj] + (gx[kj] * dj0 + gy[kj] * dj1) * lim[kj];
            }
        } else {
            for (int k = 0; k < V; ++k) { qi[k] = q[V * i + k]; qj[k] = q[V * j + k]; }
        }
        if (cf->lin && cf->pos) {
            /* extension (no reference counterpart): fall back to the cell
             * value where the reconstructed face state is not physical */
            if (!(qi[0] > 0) || !(pressure(cf->gamma, qi) > 0))
                for (int k = 0; k < V; ++k) qi[k] = q[V * i + k];
            if (!(qj[0] > 0) || !(pressure(cf->gamma, qj) > 0))
                for (int k = 0; k < V; ++k) qj[k] = q[V * j + k];
        }
        /* viscous face gradient from the cell gradients and cell states
         * (explicit.cpp:324-338); zero without diffusive_gradients */
        double gxv[V] = {0., 0., 0., 0.}, gyv[V] = {0., 0., 0., 0.};
        if (cf->diff) {
            const double ci[2] = {m->cell_cx[i], m->cell_cy[i]};
            const double cj[2] = {m->cell_cx[j], m->cell_cy[j]};
            orc_face_gradient(gxv, gyv, &gx[V * i], &gy[V * i], &gx[V * j], &gy[V * j], &q[V * i], &q[V * j], ci,
                              cj);
        }
        orc_flux_grad(cf, f, qi, qj, gxv, gyv, nn);
        for (int k = 0; k < V; ++k) {
            qt[V * i + k] -= f[k] * le / m->cell_area[i];
            qt[V * j + k] += f[k] * le / m->cell_area[j];
        }
    }
    for (int32_t i = m->n_real; i < m->n_cells; ++i)
        for (int k = 0; k < V; ++k) qt[V * i + k] = 0.;
}

/* update_cells (explicit.cpp:368-378); dt is one value per cell here */
void orc_update_cells(int64_t ncell, double* q, const double* ql, const double* qt,
                      const double* dt, double a) {
    for (int64_t c = 0; c < ncell; ++c)
        for (int k = 0; k < V; ++k) q[V * c + k] = ql[V * c + k] + qt[V * c + k] * dt[c] * a;
}

/* update_bounds (explicit.cpp:380-417) */
void orc_update_bounds(const orc_mesh* m, const orc_config* cf, double* q,
                       const double* gx, const double* gy, const double* lim) {
    for (int32_t b = 0; b < m->n_bnd; ++b) {
        int32_t e = m->bnd_edge[b];
        double nn[2] = {m->edge_nx[e], m->edge_ny[e]};
        int32_t ii = m->edge_cells[2 * e], ib = m->edge_cells[2 * e + 1];
        double qint[V];
        if (cf->lin) {
            double d0 = m->edge_cx[e] - m->cell_cx[ii];
            double d1 = m->edge_cy[e] - m->cell_cy[ii];
            for (int k = 0; k < V; ++k) {
                int64_t kk = V * (int64_t)ii + k;
                qint[k] = q[kk] + (gx[kk] * d0 + gy[kk] * d1) * lim[kk];
            }
            if (cf->pos && (!(qint[0] > 0) || !(pressure(cf->gamma, qint) > 0)))
                for (int k = 0; k < V; ++k) qint[k] = q[V * ii + k];
        } else {
            for (int k = 0; k < V; ++k) qint[k] = q[V * ii + k];
        }
        orc_bc(cf, m->bnd_bc[b], &q[V * ib], qint, nn);
    }
}

/* calc_residuals (explicit.cpp:481-552), single rank */
void orc_calc_residuals(const orc_mesh* m, const double* qt, double* R) {
    for (int k = 0; k < V; ++k) R[k] = 0.;
    for (int32_t i = 0; i < m->n_real; ++i)
        for (int k = 0; k < V; ++k) R[k] += qt[V * i + k] * qt[V * i + k] * m->cell_area[i];
    for (int k = 0; k < V; ++k) R[k] = sqrt(R[k]);
}

/* smooth_residuals (explicit.cpp:32-66): two Jacobi sweeps with eps = 0.6.
 * Each sweep clears the smoother over every cell (:49-51), subtracts
 * qt_j * eps from both sides of every edge in edge order (:52-59), then
 * rewrites the real cells only (:60-65), so a boundary ghost keeps the
 * qt = 0 that calc_time_derivatives gave it (:352-362) in both sweeps.
 * n_e is 3 for a triangle and 4 otherwise (:61). */
void orc_smooth_residuals(const orc_mesh* m, double* qt) {
    const double epsilon = 0.6;
    const int64_t n = (int64_t)V * m->n_cells;
    double* qt0 = (double*)malloc(sizeof(double) * (size_t)n);
    double* sm = (double*)malloc(sizeof(double) * (size_t)n);
    if (!qt0 || !sm) { free(qt0); free(sm); return; }
    memcpy(qt0, qt, sizeof(double) * (size_t)n);
    for (int it = 0; it < 2; ++it) {
        for (int64_t i = 0; i < n; ++i) sm[i] = 0.;
        for (int32_t e = 0; e < m->n_edges; ++e) {
            const int64_t i = m->edge_cells[2 * e], j = m->edge_cells[2 * e + 1];
            for (int k = 0; k < V; ++k) {
                sm[V * i + k] -= qt[V * j + k] * epsilon;
                sm[V * j + k] -= qt[V * i + k] * epsilon;
            }
        }
        for (int32_t i = 0; i < m->n_real; ++i) {
            const double ne = m->cell_tri[i] ? 3 : 4;
            for (int k = 0; k < V; ++k)
                qt[V * i + k] = (qt0[V * i + k] - sm[V * i + k]) / (1. + ne * epsilon);
        }
    }
    free(qt0);
    free(sm);
}

/* run (explicit.cpp:657-817), single rank, no output files */
int orc_run(const orc_mesh* m, const orc_config* cf, double* q,
            uint32_t max_step, double max_time, double tol, int32_t stop_var, uint32_t print_interval,
            uint32_t* steps_out, double* time_out, double* R_out, double* R0_out,
            double* qt_out) {
    const int64_t N = m->n_cells, n = V * N;
    double* qk = (double*)calloc(n, sizeof(double));
    double* qt = (double*)calloc(n, sizeof(double));
    double* gx = (double*)calloc(n, sizeof(double));
    double* gy = (double*)calloc(n, sizeof(double));
    double* lim = (double*)calloc(n, sizeof(double));
    double* qmin = (double*)calloc(n, sizeof(double));
    double* qmax = (double*)calloc(n, sizeof(double));
    double* dt = (double*)calloc(N, sizeof(double));
    if (!qk || !qt || !gx || !gy || !lim || !qmin || !qmax || !dt) return -1;
    double R0[V], R[V];
    for (int k = 0; k < V; ++k) { R[k] = 1.0; R0[k] = 0.0; }
    orc_update_bounds(m, cf, q, gx, gy, lim);
    uint32_t step = 0;
    double time = 0;
    for (;;) {
        double Rmax = 0.0;
        if (step == 0) Rmax = 1.0;
        else if (stop_var >= 0 && stop_var < V) Rmax = R[stop_var];
        else { for (int k = 0; k < V; ++k) Rmax = smax(R[k], Rmax); }
        if ((step >= max_step) | (time >= max_time) | (Rmax < tol)) break;
        orc_calc_dt(m, cf, q, dt);
        if (cf->gdt) {
            if (cf->kinks && cf->dt_kind != ORC_DT_FIXED) orc_min_dt_soft(m, dt);
            else orc_min_dt(m, dt);
        }
        memcpy(qk, q, sizeof(double) * n);
        for (int s = 0; s < cf->n_alpha; ++s) {
            if (cf->grad) orc_calc_gradients(m, qk, gx, gy);
            if (cf->lim) orc_calc_limiters(m, cf, qk, gx, gy, lim, qmin, qmax);
            orc_calc_time_derivatives(m, cf, qk, gx, gy, lim, qt);
            if (cf->smooth) orc_smooth_residuals(m, qt);   /* explicit.cpp:743 */
            orc_update_cells(N, qk, q, qt, dt, cf->alpha[s]);
            orc_update_bounds(m, cf, qk, gx, gy, lim);
        }
        memcpy(q, qk, sizeof(double) * n);
        if (step == 0) {
            orc_calc_residuals(m, qt, R0);
            for (int k = 0; k < V; ++k) R[k] = R0[k];
        } else if ((step % print_interval == 0) | (tol > 1.01e-16)) {
            orc_calc_residuals(m, qt, R);
            for (int k = 0; k < V; ++k) R[k] = R[k] / R0[k];
        }
        step += 1;
        time += dt[0];
    }
    *steps_out = step;
    *time_out = time;
    for (int k = 0; k < V; ++k) { R_out[k] = R[k]; R0_out[k] = R0[k]; }
    if (qt_out) memcpy(qt_out, qt, sizeof(double) * n);
    free(qk); free(qt); free(gx); free(gy); free(lim); free(qmin); free(qmax); free(dt);
    return 0;
}
