// Communicator size and rank (reference: include/mpi_dispatcher/misc.hpp).  With <mpi.h> they ask
// MPI (the caller links MPI); without it the single-process stand-in of pomerol/Misc.hpp has one
// rank.
#pragma once

#include "../pomerol/Misc.hpp"

namespace pMPI {

#if __has_include(<mpi.h>)
inline int size(MPI_Comm const& Comm) {
    int s = 1;
    MPI_Comm_size(Comm, &s);
    return s;
}
inline int rank(MPI_Comm const& Comm) {
    int r = 0;
    MPI_Comm_rank(Comm, &r);
    return r;
}
#else
inline int size(MPI_Comm const&) { return 1; }
inline int rank(MPI_Comm const&) { return 0; }
#endif

} // namespace pMPI
