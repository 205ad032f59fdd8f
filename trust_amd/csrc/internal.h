// Internal declarations shared by the host-side translation units of
// libtrust_amd.so.  Not part of the public ABI (include/trust.h is).
#pragma once

#include <cstddef>
#include <cstdint>
#include <string>
#include <vector>

#include "trust.h"

namespace trust {

// Record a message for trust_last_error() and return `code`.
int fail(int code, const char *fmt, ...) __attribute__((format(printf, 2, 3)));
void clear_error();

// Owned host CSR.
struct HostCsr {
    int64_t V = 0;
    int64_t E = 0;
    std::vector<int64_t> beg;
    std::vector<int32_t> adj;
};

// Hand an owned CSR to a caller-visible trust_csr (malloc'd buffers).
int export_csr(HostCsr &&src, trust_csr *out);

// File helpers.
int64_t file_size(const std::string &path);
int read_file(const std::string &path, void *dst, size_t bytes);
int write_file(const std::string &path, const void *src, size_t bytes);
int load_folder(const std::string &folder, HostCsr &out);

// Largest beg[u+1]-beg[u] (OpenMP, host).
int64_t host_max_degree(const int64_t *beg, int64_t V);

// memcpy with all OpenMP threads, in 4 MB pieces.
void host_parallel_copy(void *dst, const void *src, size_t bytes);

// Row sums of tile class c of an n x n partition (tiles[c*n + b], b < n):
// the global out-degree of each of its `rows` vertices, saturated to int32.
void host_tile_row_degrees(const trust_csr *tiles, int n, int c, int64_t rows, int32_t *out);

// Parallel LSD radix sort of 64-bit keys (stable); tmp must hold n keys.
void radix_sort_u64(uint64_t *keys, uint64_t *tmp, size_t n, int key_bits);

// Input contract (reference preprocessing), see preprocess.cpp.
struct PreprocessOpts {
    // >= 0: ids are already 0..n_fixed-1 and all are kept (a 1.mmio header
    // or prop.txt fixes the vertex count); < 0: ids are compacted.
    int64_t n_fixed = -1;
    // what a self-loop adds to its vertex's degree before orientation
    int self_loop_weight = 2;
    int order = TRUST_ORDER_OUTDEGREE;
};
int preprocess_edges(const int32_t *src, const int32_t *dst, int64_t m, HostCsr &out,
                     const PreprocessOpts &o = PreprocessOpts());

// A batch CSR: vbase[0] = 0, non-decreasing, vbase[ng] = V, and every edge
// of graph g stays inside [vbase[g], vbase[g+1]) (batch.cpp, OpenMP).
int batch_validate(const int64_t *beg, const int32_t *adj, int64_t V, const int64_t *vbase,
                   int64_t ngraphs);

// Pairs [e0, e0 + count) of the seeded R-MAT stream of trust_gen_rmat.
void gen_rmat_range(int scale, double a, double b, double c, uint64_t seed, int64_t e0,
                    int64_t count, int32_t *src, int32_t *dst);

// 2D hash partition tiles, see partition.cpp.
int partition_tiles(const int64_t *beg, const int32_t *adj, int64_t V, int n,
                    std::vector<HostCsr> &tiles);

}  // namespace trust
