// fvh_internal.h -- host-side types shared by the mesh builder and the solver.
#pragma once

#include <cstdint>
#include <string>
#include <vector>

#include "../../include/fvh.h"

// Mesh in the reference's numbering and geometry (fvhyper/mesh.h:136-206),
// single device: no MPI ghost cells, no halo maps.
struct fvh_mesh {
    std::vector<double> x, y;                 // nodesX, nodesY (file order)
    int32_t n_real = 0;                       // nRealCells
    int32_t n_cells = 0;                      // real + boundary ghost cells
    std::vector<int32_t> edge_nodes;          // edgesNodes, 2 per edge
    std::vector<int32_t> edge_cells;          // edgesCells, 2 per edge
    std::vector<double> e_len, e_nx, e_ny, e_cx, e_cy;
    std::vector<int32_t> cell_nodes;          // cellsNodes, 4 per cell
    std::vector<uint8_t> cell_tri;            // cellsIsTriangle
    std::vector<double> c_area, c_cx, c_cy;   // cellsAreas, cellsCenters
    std::vector<int32_t> bnd_edge;            // boundaryEdges
    std::vector<int32_t> bnd_tag;             // index into names per boundary element
    std::vector<std::string> names;           // physical names of boundary groups
    std::vector<int32_t> perm;                // device position of each reference cell (RCM), lazily
    int32_t n_edges() const { return (int32_t)e_len.size(); }
};

namespace fvh {

void set_error(const std::string& msg);

// Builds reference numbering + geometry from file-order arrays.  Returns an
// FVH_* code; on error the message is set.
int32_t build_mesh(fvh_mesh& m, int32_t n_nodes, const double* x, const double* y,
                   int32_t n_cells, const int32_t* cells, int32_t n_bnd,
                   const int32_t* bnd, const int32_t* bnd_tag,
                   const std::vector<std::string>& names);

int32_t read_msh(const char* path, fvh_mesh& m);
int32_t read_su2(const char* path, fvh_mesh& m);

// Adjoints of the quantities derived from the node coordinates: per edge
// (normal, length, Green-Gauss weight, centre) and per cell, ghosts included
// (area, centre).
struct GeoBar {
    std::vector<double> e_nx, e_ny, e_len, e_w, e_cx, e_cy;
    std::vector<double> c_area, c_cx, c_cy;
    void resize(int32_t n_edges, int32_t n_cells);
};

// Reverse of the geometry build_mesh derives from the node coordinates
// (mesh.cpp:190-260, ghost cells mesh.cpp:642-655) and of the Green-Gauss
// weight |x_f - x_i| / |x_i - x_j| (explicit.cpp:132-140): adds dJ/dx_k and
// dJ/dy_k to gxy[2k], gxy[2k+1].  b is consumed.  The orientation flip of a
// normal and the sign of an area are piecewise constant and kept.
void mesh_geometry_vjp(const fvh_mesh& m, GeoBar& b, double* gxy);

// Reverse Cuthill-McKee order of the real cells: order[k] = reference cell
// placed at device position k.  Deterministic (ties by cell index).
std::vector<int32_t> rcm_order(const fvh_mesh& m);

// Recursive coordinate bisection of the real cells' centres into tiles of
// `tile` cells (all full but the last; cuts at multiples of tile along the
// longer side of the centroids' bounding box), tiles in recursion order,
// cells in ascending reference index within a tile.  order[k] as for
// rcm_order.  A tile whose index (in the order without splits) is marked in
// `split` is bisected once more, into two tiles; leaf_sizes receives the
// size of every tile, in order.
std::vector<int32_t> rcb_order(const fvh_mesh& m, int32_t tile, const std::vector<uint8_t>* split = nullptr,
                               std::vector<int32_t>* leaf_sizes = nullptr);

}  // namespace fvh
