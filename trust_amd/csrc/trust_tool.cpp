// trust_tool -- input-contract tools (drop-ins for the reference's
// Preprocess/ executables) and synthetic graph generation.
//
//   trust_tool undirect <edgelist.txt> [out=1.mmio]
//       Preprocess/fromDirectToUndirect.cpp: text "u v" lines (lines not
//       starting with a digit are skipped) -> compact 1-based ids, u<=v,
//       deduplicated, written as "V V M" + "u v" lines.
//   trust_tool preprocess [in=1.mmio] [outdir=.]
//       Preprocess/preprocess_outdegree.cpp: 1.mmio -> begin.bin, adjacent.bin
//       (byte-identical to the reference tool).
//   trust_tool partition <n> [dir=.]
//       Preprocess/partition.cpp: begin.bin/adjacent.bin in dir ->
//       dir/partition{i}_{j}/{begin.bin,adjacent.bin,map,division}.
//   trust_tool gen rmat <scale> <edge_factor> <seed> <outdir>
//   trust_tool gen chunglu <n> <m> <gamma> <max_degree> <seed> <outdir>
//   trust_tool gen er <n> <m> <seed> <outdir>
//   trust_tool gen ba <n> <k> <seed> <outdir>
//       seeded edge list -> preprocessed CSR folder (a=.57 b=.19 c=.19 for R-MAT).
#include <sys/stat.h>

#include <algorithm>
#include <cstdio>
#include <cstdlib>
#include <cstring>
#include <fstream>
#include <sstream>
#include <string>
#include <vector>

#include "internal.h"

using namespace trust;

static int die(int rc)
{
    fprintf(stderr, "trust_tool: error %d: %s\n", rc, trust_last_error());
    return 1;
}

static bool read_pairs(const std::string &path, std::vector<int32_t> &a, std::vector<int32_t> &b,
                       bool header, long long *hv)
{
    std::ifstream in(path);
    if (!in)
        return false;
    std::string line;
    bool first = true;
    while (std::getline(in, line)) {
        if (line.empty() || line[0] < '0' || line[0] > '9')
            continue;
        std::istringstream ss(line);
        long long u, v;
        if (header && first) {
            long long v2, m;
            ss >> u >> v2 >> m;
            *hv = u;
            first = false;
            continue;
        }
        if (!(ss >> u >> v))
            continue;
        a.push_back((int32_t)u);
        b.push_back((int32_t)v);
    }
    return true;
}

static int cmd_undirect(int argc, char **argv)
{
    if (argc < 1)
        return 2;
    std::string out = argc > 1 ? argv[1] : "1.mmio";
    std::vector<int32_t> a, b;
    if (!read_pairs(argv[0], a, b, false, nullptr)) {
        fprintf(stderr, "trust_tool: cannot read %s\n", argv[0]);
        return 1;
    }
    int32_t maxid = -1;
    for (size_t e = 0; e < a.size(); e++)
        maxid = std::max(maxid, std::max(a[e], b[e]));
    std::vector<int32_t> id((size_t)maxid + 1, 0);
    for (size_t e = 0; e < a.size(); e++)
        id[a[e]] = id[b[e]] = 1;
    int64_t k = 0;
    for (auto &x : id)
        x = x ? (int32_t)++k : 0;  // 1-based
    std::vector<uint64_t> keys(a.size()), tmp(a.size());
    for (size_t e = 0; e < a.size(); e++) {
        uint64_t u = (uint64_t)id[a[e]], v = (uint64_t)id[b[e]];
        if (u > v)
            std::swap(u, v);
        keys[e] = (u << 32) | v;
    }
    radix_sort_u64(keys.data(), tmp.data(), keys.size(), 64);
    keys.erase(std::unique(keys.begin(), keys.end()), keys.end());
    FILE *f = fopen(out.c_str(), "w");
    if (!f) {
        fprintf(stderr, "trust_tool: cannot write %s\n", out.c_str());
        return 1;
    }
    fprintf(f, "%lld %lld %zu\n", (long long)k, (long long)k, keys.size());
    for (uint64_t x : keys)
        fprintf(f, "%u %u\n", (unsigned)(x >> 32), (unsigned)(x & 0xffffffffu));
    fclose(f);
    return 0;
}

static int cmd_preprocess(int argc, char **argv)
{
    std::string in = argc > 0 ? argv[0] : "1.mmio";
    std::string outdir = argc > 1 ? argv[1] : ".";
    std::vector<int32_t> a, b;
    long long hv = -1;
    if (!read_pairs(in, a, b, true, &hv) || hv < 0) {
        fprintf(stderr, "trust_tool: cannot read %s\n", in.c_str());
        return 1;
    }
    for (size_t e = 0; e < a.size(); e++) {  // 1-based ids in the file
        a[e]--;
        b[e]--;
    }
    HostCsr g;
    int rc = preprocess_edges(a.data(), b.data(), (int64_t)a.size(), g, hv);
    if (rc)
        return die(rc);
    trust_csr c{g.V, g.E, g.beg.data(), g.adj.data()};
    rc = trust_csr_save(outdir.c_str(), &c);
    return rc ? die(rc) : 0;
}

static int cmd_partition(int argc, char **argv)
{
    if (argc < 1)
        return 2;
    int n = atoi(argv[0]);
    std::string dir = argc > 1 ? argv[1] : ".";
    trust_csr c;
    int rc = trust_csr_load(dir.c_str(), &c);
    if (rc)
        return die(rc);
    rc = trust_partition_write(dir.c_str(), &c, n);
    trust_csr_free(&c);
    return rc ? die(rc) : 0;
}

static int cmd_gen(int argc, char **argv)
{
    if (argc < 1)
        return 2;
    std::string kind = argv[0];
    std::vector<int32_t> a, b;
    std::string outdir;
    int rc = 0;
    if (kind == "rmat" && argc == 5) {
        int scale = atoi(argv[1]);
        int64_t m = (int64_t)atoll(argv[2]) << scale;
        a.resize((size_t)m);
        b.resize((size_t)m);
        rc = trust_gen_rmat(scale, m, 0.57, 0.19, 0.19, strtoull(argv[3], 0, 10), a.data(),
                            b.data());
        outdir = argv[4];
    } else if (kind == "chunglu" && argc == 7) {
        int64_t n = atoll(argv[1]), m = atoll(argv[2]);
        a.resize((size_t)m);
        b.resize((size_t)m);
        rc = trust_gen_chung_lu(n, m, atof(argv[3]), atof(argv[4]), strtoull(argv[5], 0, 10),
                                a.data(), b.data());
        outdir = argv[6];
    } else if (kind == "er" && argc == 5) {
        int64_t n = atoll(argv[1]), m = atoll(argv[2]);
        a.resize((size_t)m);
        b.resize((size_t)m);
        rc = trust_gen_erdos_renyi(n, m, strtoull(argv[3], 0, 10), a.data(), b.data());
        outdir = argv[4];
    } else if (kind == "ba" && argc == 5) {
        int64_t n = atoll(argv[1]);
        int k = atoi(argv[2]);
        a.resize((size_t)(n * k));
        b.resize((size_t)(n * k));
        int64_t m = 0;
        rc = trust_gen_barabasi_albert(n, k, strtoull(argv[3], 0, 10), a.data(), b.data(), &m);
        a.resize((size_t)m);
        b.resize((size_t)m);
        outdir = argv[4];
    } else {
        return 2;
    }
    if (rc)
        return die(rc);
    HostCsr g;
    rc = preprocess_edges(a.data(), b.data(), (int64_t)a.size(), g);
    if (rc)
        return die(rc);
    trust_csr c{g.V, g.E, g.beg.data(), g.adj.data()};
    rc = trust_csr_save(outdir.c_str(), &c);
    if (rc)
        return die(rc);
    printf("%s: V=%lld E=%lld\n", outdir.c_str(), (long long)g.V, (long long)g.E);
    return 0;
}

int main(int argc, char **argv)
{
    int rc = 2;
    if (argc >= 2) {
        std::string cmd = argv[1];
        if (cmd == "undirect")
            rc = cmd_undirect(argc - 2, argv + 2);
        else if (cmd == "preprocess")
            rc = cmd_preprocess(argc - 2, argv + 2);
        else if (cmd == "partition")
            rc = cmd_partition(argc - 2, argv + 2);
        else if (cmd == "gen")
            rc = cmd_gen(argc - 2, argv + 2);
    }
    if (rc == 2)
        fprintf(stderr,
                "usage: trust_tool undirect <edges.txt> [out]\n"
                "       trust_tool preprocess [1.mmio] [outdir]\n"
                "       trust_tool partition <n> [dir]\n"
                "       trust_tool gen rmat <scale> <ef> <seed> <outdir>\n"
                "       trust_tool gen chunglu <n> <m> <gamma> <maxdeg> <seed> <outdir>\n"
                "       trust_tool gen er <n> <m> <seed> <outdir>\n"
                "       trust_tool gen ba <n> <k> <seed> <outdir>\n");
    return rc;
}
