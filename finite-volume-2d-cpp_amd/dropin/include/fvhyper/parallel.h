// fvhyper/parallel.h -- drop-in for fvhyper's process wrapper
// (reference: fvhyper/parallel.h:24-33).  This build runs one partition on
// one GPU, so the wrapper reports rank 0 of 1 and needs no MPI.
#pragma once

namespace fvhyper {

class mpi_wrapper {
public:
    int rank;
    int size;
    mpi_wrapper();
    int exit();
};

}  // namespace fvhyper
