// fvhyper/post.h -- drop-in for fvhyper's VTU/PVTU writer
// (reference: fvhyper/post.h:24-42, fvhyper/src/post.cpp:21-200): same file
// names, layout and number formatting.
#pragma once

#include <map>
#include <string>
#include <vector>

#include <fvhyper/mesh.h>

namespace fvhyper {

namespace post {
    extern std::map<std::string, void (*)(double*, double*)> extra_scalars;
    extern std::map<std::string, void (*)(double*, double*)> extra_vectors;
}

extern const std::vector<std::string> var_names;

void writeVtk(const std::string name, std::vector<double>& q, mesh& m, int rank, int world_size,
              const std::string time = "");

// Extension (no reference counterpart): write every number with %.17g and
// the arrays as Float64 instead of the reference's std::to_string text.
// Also switched on by the environment variable FVH_VTK_FULL_PRECISION=1.
extern bool vtk_full_precision;

}  // namespace fvhyper
