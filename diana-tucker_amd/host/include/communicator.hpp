// Single-rank communicator (include/communicator.hpp:14-97 collapsed to one
// process): every collective of the reference is a copy or a no-op at one rank,
// so this build keeps only the rank/size queries and the init/finalize pair.
#ifndef DIANA_AMD_COMMUNICATOR_HPP
#define DIANA_AMD_COMMUNICATOR_HPP
inline void mpi_init() {}
inline int mpi_rank() { return 0; }
inline int mpi_size() { return 1; }
inline void mpi_finalize() {}

template <typename Ty>
class Communicator {
public:
    int rank() const { return 0; }
    int size() const { return 1; }
    static void bcast(Ty *, int, int) {}
    static void barrier() {}
};
#endif
