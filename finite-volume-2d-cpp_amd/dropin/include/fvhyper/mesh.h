// fvhyper/mesh.h -- drop-in for fvhyper's mesh (reference: fvhyper/mesh.h).
//
// The public members keep the reference's names, types and numbering
// (cells in file order, then one boundary ghost cell per boundary element;
// edges by first appearance), so plugin code that reads them (calc_dt,
// generate_initial_solution, ...) works unchanged.  read_file() parses the
// gmsh MSH 4.1 file through libfvh and keeps the library handle for run().
#pragma once

#include <sys/types.h>
#include <math.h>

#include <algorithm>
#include <iostream>
#include <map>
#include <stdexcept>
#include <string>
#include <tuple>
#include <unordered_map>
#include <vector>

#include <fvhyper/parallel.h>

struct fvh_mesh;
struct fvh_solver;

namespace fvhyper {

extern const int vars;

namespace boundaries {
    extern std::map<std::string, void (*)(double*, double*, double*)> bounds;
}

// Fixed-width rows of node or cell indices (reference: mesh.h:64-131).
template <uint N>
class meshArray {
public:
    meshArray() : n(0) {}
    uint& operator()(const uint& i, const uint& j) { return nodes[i * N + j]; }
    const uint& operator()(const uint& i, const uint& j) const { return nodes[i * N + j]; }
    void push_back(const std::vector<uint> v) {
        if (v.size() != N)
            throw std::invalid_argument("size of input array " + std::to_string(v.size()) + " not equal to " +
                                        std::to_string(N));
        nodes.insert(nodes.end(), v.begin(), v.end());
        n += 1;
    }
    uint cols() const { return n; }
    uint rows() const { return N; }
    void dump() {
        for (uint i = 0; i < n; ++i) {
            for (uint j = 0; j < N; ++j) std::cout << (*this)(i, j) << " ";
            std::cout << "\n";
        }
    }
    const std::vector<uint>& get_vector() { return nodes; }
    void assign(std::vector<uint>&& v) { nodes = std::move(v); n = (uint)(nodes.size() / N); }

private:
    std::vector<uint> nodes;
    uint n;
};

class mesh {
public:
    mesh();
    ~mesh();
    mesh(const mesh&) = delete;
    mesh& operator=(const mesh&) = delete;

    std::string filename;
    std::map<uint, std::string> physicalNames;

    std::vector<double> nodesX;
    std::vector<double> nodesY;

    meshArray<2> edgesNodes;
    meshArray<2> edgesCells;
    std::vector<double> edgesLengths;
    std::vector<double> edgesNormalsX;
    std::vector<double> edgesNormalsY;
    std::vector<double> edgesCentersX;
    std::vector<double> edgesCentersY;

    std::vector<uint> boundaryEdges;
    std::vector<void (*)(double*, double*, double*)> boundaryFuncs;
    std::vector<std::string> boundaryNames;   // physical name of each boundary element

    meshArray<4> cellsNodes;
    std::vector<bool> cellsIsTriangle;
    std::vector<double> cellsAreas;
    std::vector<double> cellsCentersX;
    std::vector<double> cellsCentersY;
    std::vector<bool> cellsIsGhost;

    uint nRealCells;

    // Reads name.msh (reference: mesh.cpp:803-855).  Throws
    // std::invalid_argument as the reference does for a boundary element
    // that is not a cell edge or a physical name missing from
    // boundaries::bounds.
    void read_file(std::string name, mpi_wrapper& pool);

    // libfvh mesh handle (owned)
    ::fvh_mesh* device_mesh() const { return handle_; }
    // libfvh solver of the per-phase functions (calc_gradients, ...), built on
    // first use from the probed plugin kinds (owned)
    ::fvh_solver* phase_solver() const;

private:
    ::fvh_mesh* handle_;
    mutable ::fvh_solver* phase_;
};

}  // namespace fvhyper
