// trianglecounting -- MPI-free drop-in for the reference's
// trianglecounting.bin (Without-graph-partition/mpi_main.cpp:75-119 and
// With-graph-partition/mpi_main.cpp:75-127).
//
//   trianglecounting <folder> <ranks> <threads> <blocks> <chunk> [<n>]
//
// Same positional arguments.  The reference starts one MPI process per
// rank; here one process runs the logical ranks 0..ranks-1 in turn on one
// GPU, which is what the reference does when launched with `mpirun -n 1`
// (mpi_main.cpp:95-104): counts are summed and times reduced to max/min.
// Output lines keep the reference's fields and order:
//   unpartitioned: folder,V,E,count,max_time,GTEPS
//   partitioned:   \nfolder,triangle:count,min_time,max_time
// (integers printed with 64-bit formats so E >= 2^31 is not truncated).
// Errors print the library's message and exit with status 1.
#include <cstdio>
#include <cstdlib>

#include "trust.h"

static int usage()
{
    fprintf(stderr,
            "usage: trianglecounting <folder> <ranks> <threads> <blocks> <chunk> [<n>]\n"
            "  folder   graph folder (begin.bin, adjacent.bin; or partition{i}_{j}/ tiles)\n"
            "  ranks    logical ranks (for <n>: a multiple of n^3)\n"
            "  threads  work-items per workgroup (multiple of 64, <= 1024; 0 = auto)\n"
            "  blocks   workgroups (0 = fill the device)\n"
            "  chunk    vertices per queue grab (0 = auto)\n"
            "  n        partition count: count the n x n tile partition\n");
    return 2;
}

int main(int argc, char *argv[])
{
    if (argc != 6 && argc != 7)
        return usage();
    const char *folder = argv[1];
    const int total_rank = atoi(argv[2]);
    const int threads = atoi(argv[3]);
    const int blocks = atoi(argv[4]);
    const int chunk = atoi(argv[5]);
    const int n = argc == 7 ? atoi(argv[6]) : 0;
    if (total_rank < 1 || (argc == 7 && n < 1))
        return usage();
    unsigned long long sum = 0;
    double tmin = 9999, tmax = 0;
    trust_arguments args = {0, 0, 0, 0, 0};
    for (int rank = 0; rank < total_rank; rank++) {
        int rc = n > 0 ? trust_triangle_count_partition(rank, folder, total_rank, threads, blocks,
                                                        chunk, n, &args)
                       : trust_triangle_count(rank, folder, total_rank, threads, blocks, chunk,
                                              &args);
        if (rc != TRUST_OK) {
            fprintf(stderr, "trianglecounting: rank %d: error %d: %s\n", rank, rc,
                    trust_last_error());
            return 1;
        }
        sum += args.count;
        if (args.time < tmin)
            tmin = args.time;
        if (args.time > tmax)
            tmax = args.time;
    }
    if (n > 0)
        printf("\n%s,triangle:%llu,%f,%f\n", folder, sum, tmin, tmax);
    else
        printf("%s,%lld,%lld,%llu,%f,%f \n", folder, (long long)args.vertices,
               (long long)args.edge_count, sum, tmax,
               tmax > 0 ? (double)args.edge_count / tmax / 1e9 : 0.0);
    return 0;
}
