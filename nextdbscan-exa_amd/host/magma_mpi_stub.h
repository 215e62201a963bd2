// Single-rank stand-in for the reference's magmaMPI handle (magma_mpi.h:28-137).
// The reference's own build without MPI_ON makes every collective a no-op on
// one rank (magma_mpi.h:61,80,100,132-133); this keeps only that case, so the
// host entry points stay source-compatible.
#pragma once

struct magmaMPI {
    int rank = 0;
    int n_nodes = 1;
    static magmaMPI build() { return magmaMPI(); }
};
