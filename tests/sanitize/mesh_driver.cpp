// Host mesh code under AddressSanitizer and UndefinedBehaviorSanitizer: reads
// every mesh file named on the command line (MSH 4.1 or SU2 by extension)
// through the C ABI, exports every member, builds both device orders and
// the geometry reverse, and reads the malformed files expecting an error
// code.  Exit status 0 when all went as expected.
//
//   mesh_asan good1.msh good2.su2 ... -- bad1.msh bad2.su2 ...
#include <cstdio>
#include <cstring>
#include <string>
#include <vector>

#include "fvh_internal.h"

static int fail(const char* what, const char* path) {
    std::fprintf(stderr, "%s: %s (%s)\n", path, what, fvh_last_error());
    return 1;
}

static int32_t read_any(const char* path, fvh_mesh** m) {
    const size_t n = std::strlen(path);
    if (n > 4 && std::strcmp(path + n - 4, ".su2") == 0) return fvh_mesh_read_su2(path, m);
    return fvh_mesh_read_msh(path, m);
}

int main(int argc, char** argv) {
    bool bad = false;
    int errors = 0;
    for (int a = 1; a < argc; ++a) {
        if (std::strcmp(argv[a], "--") == 0) { bad = true; continue; }
        fvh_mesh* m = nullptr;
        const int32_t rc = read_any(argv[a], &m);
        if (bad) {
            if (rc == FVH_OK) { errors += fail("malformed file accepted", argv[a]); fvh_mesh_free(m); }
            continue;
        }
        if (rc != FVH_OK) { errors += fail("read failed", argv[a]); continue; }
        int64_t c[6];
        fvh_mesh_counts(m, c);
        const char* fields[] = {"nodesX", "nodesY", "edgesLengths", "edgesNormalsX", "edgesNormalsY",
                                "edgesCentersX", "edgesCentersY", "cellsAreas", "cellsCentersX", "cellsCentersY",
                                "edgesNodes", "edgesCells", "cellsNodes", "cellsIsTriangle", "boundaryEdges",
                                "boundaryTags", "rcmPermutation", "tilePermutation"};
        for (const char* f : fields) {
            const int64_t n = fvh_mesh_export(m, f, nullptr, 0);
            if (n < 0) { errors += fail(f, argv[a]); continue; }
            std::vector<double> buf((size_t)n + 1);
            if (fvh_mesh_export(m, f, buf.data(), n) != n) errors += fail(f, argv[a]);
            // a short destination is refused, not overrun
            if (n > 1 && fvh_mesh_export(m, f, buf.data(), n - 1) >= 0) errors += fail("short export", argv[a]);
        }
        char name[64];
        for (int32_t k = 0; k < (int32_t)c[5]; ++k) fvh_mesh_name(m, k, name, sizeof name);
        if (fvh_mesh_name(m, (int32_t)c[5], name, sizeof name) >= 0) errors += fail("name out of range", argv[a]);
        const std::vector<int32_t> r1 = fvh::rcm_order(*m);
        const std::vector<int32_t> r2 = fvh::rcb_order(*m, 208);
        if ((int64_t)r1.size() != c[1] || (int64_t)r2.size() != c[1]) errors += fail("order size", argv[a]);
        fvh::GeoBar b;
        b.resize(m->n_edges(), m->n_cells);
        for (size_t e = 0; e < b.e_len.size(); ++e) {
            b.e_nx[e] = 0.1; b.e_ny[e] = -0.2; b.e_len[e] = 0.3; b.e_w[e] = 0.05; b.e_cx[e] = 0.01; b.e_cy[e] = 0.02;
        }
        for (size_t k = 0; k < b.c_area.size(); ++k) { b.c_area[k] = 0.5; b.c_cx[k] = 0.1; b.c_cy[k] = -0.1; }
        std::vector<double> gxy(2 * m->x.size(), 0.0);
        fvh::mesh_geometry_vjp(*m, b, gxy.data());
        fvh_mesh_free(m);
        std::printf("%s: %lld cells, %lld edges ok\n", argv[a], (long long)c[1], (long long)c[3]);
    }
    return errors ? 1 : 0;
}
