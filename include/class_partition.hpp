// class_partition.hpp -- drop-in for c++/source/class_partition.hpp:29-48.
// One entry per rank; on one GPU there is one rank holding the whole grid.
#ifndef PS_CLASS_PARTITION_HPP
#define PS_CLASS_PARTITION_HPP

#include <vector>

class class_partition
{
public:
    int ncell[3];
    int icell[3];
    double dx[3];
};

// Same arguments as the reference's partition_setup (partition_setup.cpp:9-15):
// pencil direction (0: split y, 1: split x), domain cells, boundary
// conditions, cell size and per-direction extension flags (a direction with
// bc 0 and its flag set is doubled).  Single rank: one partition, icell = 0.
std::vector<class_partition> partition_setup(int pencil_dir, int domain_ncell[3], int bound_cond[3],
                                             double dx[3], bool extendX, bool extendY, bool extendZ);

#endif
