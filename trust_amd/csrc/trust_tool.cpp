// trust_tool -- input-contract tools (drop-ins for the reference's
// Preprocess/ executables) and synthetic graph generation.
//
//   trust_tool undirect <edgelist.txt> [out=1.mmio]
//       Preprocess/fromDirectToUndirect.cpp: text "u v" lines (lines not
//       starting with a digit are skipped) -> compact 1-based ids, u<=v,
//       deduplicated, written as "V V M" + "u v" lines.
//   trust_tool preprocess [in=1.mmio] [outdir=.] [--order outdegree|plain|indegree]
//       Preprocess/preprocess_outdegree.cpp: 1.mmio -> begin.bin, adjacent.bin
//       (byte-identical to the reference tool).  --order plain is
//       Preprocess/otherPreprocess/preprocess.cpp (also writes `edge`, the
//       oriented (u,v) int32 pairs), --order indegree is
//       otherPreprocess/preprocess_indegree.cpp (also writes `edge`, one int32
//       per row: entries <= the row index found by the tool's binary search).
//   trust_tool txt2bin <edgelist.txt> [edge_count_hint]
//       Preprocess/speedupIO/TransformTxt2Binary.cpp: a whitespace-separated
//       stream of id pairs -> prop.txt ("pairs vertices"), degree (int32 per
//       compact id: endpoint occurrences), BinEdgelist (int32 pairs of compact
//       ids, input order), written to the current directory.  A character
//       that cannot start a number comments out the rest of its line (the
//       reference tool stops making progress there).
//   trust_tool preprocess-bin [dir=.]
//       Preprocess/speedupIO/preprocess_out.cpp: dir/{prop.txt,BinEdgelist}
//       -> dir/begin.bin, dir/adjacent.bin (byte-identical to the reference).
//   trust_tool partition <n> [dir=.]
//       Preprocess/partition.cpp: begin.bin/adjacent.bin in dir ->
//       dir/partition{i}_{j}/{begin.bin,adjacent.bin,map,division}; the
//       input is mapped from disk and one tile row is held at a time, any
//       n up to 1024 (the reference: whole graph in memory, n <= 10).
//   trust_tool gen rmat <scale> <edge_factor> <seed> <outdir>
//   trust_tool gen chunglu <n> <m> <gamma> <max_degree> <seed> <outdir>
//   trust_tool gen er <n> <m> <seed> <outdir>
//   trust_tool gen ba <n> <k> <seed> <outdir>
//       seeded edge list -> preprocessed CSR folder (a=.57 b=.19 c=.19 for R-MAT).
#include <sys/stat.h>

#include <algorithm>
#include <climits>
#include <cstdio>
#include <memory>
#include <cstdlib>
#include <cstring>
#include <fstream>
#include <string>
#include <vector>

#include "internal.h"

using namespace trust;

static int die(int rc)
{
    fprintf(stderr, "trust_tool: error %d: %s\n", rc, trust_last_error());
    return 1;
}

// A file read in 64 MiB blocks, one byte at a time.
struct ByteReader {
    FILE *f;
    std::vector<char> buf;
    size_t pos = 0, len = 0;
    explicit ByteReader(const std::string &path)
        : f(fopen(path.c_str(), "rb")), buf(f ? (size_t)64 << 20 : 0) {}
    ~ByteReader()
    {
        if (f)
            fclose(f);
    }
    bool ok() const { return f != nullptr; }
    int next()
    {
        if (pos == len) {
            len = fread(buf.data(), 1, buf.size(), f);
            pos = 0;
            if (len == 0)
                return EOF;
        }
        return (unsigned char)buf[pos++];
    }
};

static bool is_blank(int c)
{
    return c == ' ' || c == '\t' || c == '\r' || c == '\v' || c == '\f';
}

// Parses an optionally signed decimal starting at c; false if there are no
// digits or the value leaves [-2^31, 2^31).  c ends on the first byte after.
static bool parse_int(ByteReader &r, int &c, long long &x)
{
    const bool neg = c == '-';
    if (c == '-' || c == '+')
        c = r.next();
    if (c < '0' || c > '9')
        return false;
    x = 0;
    while (c >= '0' && c <= '9') {
        x = x * 10 + (c - '0');
        if (x > (long long)INT32_MAX + 1)
            return false;
        c = r.next();
    }
    if (neg)
        x = -x;
    return x <= INT32_MAX;
}

// Id pairs, interleaved (u, v), in fixed-size blocks: reading a large list
// needs no reallocation copy (a growing vector holds its old and new buffers
// at once), and the pairs are rewritten in place and written block by block.
struct PairBlocks {
    static constexpr size_t kBlock = (size_t)1 << 22;  // pairs per block (32 MiB)
    std::vector<std::unique_ptr<int32_t[]>> blocks;
    size_t n = 0;
    void push(int32_t u, int32_t v)
    {
        if (n % kBlock == 0)
            blocks.emplace_back(new int32_t[2 * kBlock]);
        int32_t *p = blocks.back().get() + 2 * (n % kBlock);
        p[0] = u;
        p[1] = v;
        n++;
    }
    size_t size_of(size_t k) const { return std::min(kBlock, n - k * kBlock); }  // pairs in block k
};

// Line-oriented edge text as the reference's readers take it
// (fromDirectToUndirect.cpp:36-52, preprocess_outdegree.cpp:49-78): lines
// that do not start with a digit are skipped, a line gives the first two
// integers on it (a third, the 1.mmio weight column, is ignored), and with
// `header` the first such line is "rows cols entries" (hv[0], hv[1]).  Each
// pair goes to push(u, v).  An id below min_id (the reference's readers index
// arrays with it: undefined behaviour) is an error: *bad_id holds it and the
// call returns false (*bad_id stays LLONG_MIN on other failures).
template <typename Push>
static bool read_pairs(const std::string &path, Push &&push, bool header, long long *hv,
                       long long min_id, long long *bad_id)
{
    ByteReader r(path);
    *bad_id = LLONG_MIN;
    if (!r.ok())
        return false;
    bool first = true;
    int c = r.next();
    while (c != EOF) {
        if (c < '0' || c > '9') {  // not a data line
            while (c != EOF && c != '\n')
                c = r.next();
            if (c == '\n')
                c = r.next();
            continue;
        }
        long long v[3] = {-1, -1, -1};
        int nv = 0;
        while (c != EOF && c != '\n') {
            if (is_blank(c)) {
                c = r.next();
                continue;
            }
            long long x;
            if (!((c >= '0' && c <= '9') || c == '-' || c == '+') || !parse_int(r, c, x)) {
                while (c != EOF && c != '\n')  // the stream reader stops at this token
                    c = r.next();
                break;
            }
            if (nv < 3)
                v[nv++] = x;
        }
        if (c == '\n')
            c = r.next();
        if (header && first) {
            hv[0] = v[0];
            hv[1] = v[1];
            first = false;
        } else if (nv >= 2) {
            if (v[0] < min_id || v[1] < min_id) {
                *bad_id = v[0] < min_id ? v[0] : v[1];
                return false;
            }
            push((int32_t)v[0], (int32_t)v[1]);
        }
    }
    return true;
}

// Whitespace-separated integer pairs (the fscanf "%d%d" loop of
// TransformTxt2Binary.cpp:85-95, which pairs tokens across line ends).  Ids
// must lie in [0, 2^31).
static bool read_int_pairs(const std::string &path, PairBlocks &pb, std::string &err)
{
    ByteReader r(path);
    if (!r.ok()) {
        err = "cannot open";
        return false;
    }
    bool have_u = false;
    int32_t u = 0;
    int c = r.next();
    while (c != EOF) {
        if (is_blank(c) || c == '\n') {
            c = r.next();
            continue;
        }
        if (c == '-' || c == '+' || (c >= '0' && c <= '9')) {
            long long x;
            if (!parse_int(r, c, x)) {
                err = "malformed or out-of-range id";
                return false;
            }
            if (x < 0) {
                err = "negative id";
                return false;
            }
            if (!have_u) {
                u = (int32_t)x;
                have_u = true;
            } else {
                pb.push(u, (int32_t)x);
                have_u = false;
            }
            continue;
        }
        while (c != EOF && c != '\n')  // comment to the end of the line
            c = r.next();
    }
    if (have_u) {
        err = "odd number of ids (the last pair has no second id)";
        return false;
    }
    return true;
}

static int write_bytes(const std::string &path, const void *p, size_t n)
{
    FILE *f = fopen(path.c_str(), "wb");
    if (!f || (n && fwrite(p, 1, n, f) != n)) {
        if (f)
            fclose(f);
        fprintf(stderr, "trust_tool: cannot write %s\n", path.c_str());
        return 1;
    }
    return fclose(f) == 0 ? 0 : 1;
}

// The edge_count_hint argument is accepted for the reference's argv and not
// needed: the pairs are read into blocks.
static int cmd_txt2bin(int argc, char **argv)
{
    if (argc < 1)
        return 2;
    PairBlocks pb;
    std::string err;
    if (!read_int_pairs(argv[0], pb, err)) {
        fprintf(stderr, "trust_tool: %s: %s\n", argv[0], err.c_str());
        return 1;
    }
    // ids that occur, compacted in ascending order (TransformTxt2Binary.cpp:97-121)
    int64_t maxid = -1;
    for (size_t k = 0; k < pb.blocks.size(); k++) {
        const int32_t *p = pb.blocks[k].get();
        for (size_t i = 0; i < 2 * pb.size_of(k); i++)
            maxid = std::max<int64_t>(maxid, p[i]);
    }
    std::vector<int64_t> occ((size_t)(maxid + 1), 0);
    for (size_t k = 0; k < pb.blocks.size(); k++) {
        const int32_t *p = pb.blocks[k].get();
        for (size_t i = 0; i < 2 * pb.size_of(k); i++)
            occ[p[i]]++;
    }
    std::vector<int32_t> degree;
    std::vector<int32_t> cid((size_t)(maxid + 1), -1);
    for (int64_t x = 0; x <= maxid; x++)
        if (occ[x]) {
            if (occ[x] > INT32_MAX) {
                fprintf(stderr, "trust_tool: id %lld occurs more than 2^31-1 times\n",
                        (long long)x);
                return 1;
            }
            cid[x] = (int32_t)degree.size();
            degree.push_back((int32_t)occ[x]);
        }
    std::vector<int64_t>().swap(occ);
    char prop[64];
    const int np = snprintf(prop, sizeof prop, "%zu %zu", pb.n, degree.size());
    if (write_bytes("prop.txt", prop, (size_t)np) ||
        write_bytes("degree", degree.data(), degree.size() * sizeof(int32_t)))
        return 1;
    // BinEdgelist: the compacted pairs, rewritten in place and written block by block
    FILE *f = fopen("BinEdgelist", "wb");
    bool ok = f != nullptr;
    for (size_t k = 0; ok && k < pb.blocks.size(); k++) {
        int32_t *p = pb.blocks[k].get();
        const size_t m = 2 * pb.size_of(k);
        for (size_t i = 0; i < m; i++)
            p[i] = cid[p[i]];
        ok = fwrite(p, sizeof(int32_t), m, f) == m;
        pb.blocks[k].reset();
    }
    if (f)
        ok = fclose(f) == 0 && ok;
    if (!ok) {
        fprintf(stderr, "trust_tool: cannot write BinEdgelist\n");
        return 1;
    }
    printf("pairs=%zu vertices=%zu\n", pb.n, degree.size());
    return 0;
}

static int cmd_preprocess_bin(int argc, char **argv)
{
    const std::string dir = argc > 0 ? argv[0] : ".";
    long long m = -1, k = -1;
    {
        std::ifstream prop(dir + "/prop.txt");
        if (!(prop >> m >> k) || m < 0 || k < 0 || k > INT32_MAX) {
            fprintf(stderr, "trust_tool: %s/prop.txt: expected \"pairs vertices\"\n", dir.c_str());
            return 1;
        }
    }
    std::vector<int32_t> pairs((size_t)(2 * m));
    FILE *f = fopen((dir + "/BinEdgelist").c_str(), "rb");
    if (!f || fread(pairs.data(), sizeof(int32_t), pairs.size(), f) != pairs.size()) {
        if (f)
            fclose(f);
        fprintf(stderr, "trust_tool: %s/BinEdgelist holds fewer than %lld pairs\n", dir.c_str(), m);
        return 1;
    }
    fclose(f);
    std::vector<int32_t> a((size_t)m), b((size_t)m);
    for (long long e = 0; e < m; e++) {
        a[e] = pairs[2 * e];
        b[e] = pairs[2 * e + 1];
    }
    std::vector<int32_t>().swap(pairs);
    PreprocessOpts o;
    o.n_fixed = k;  // ids are already compact (preprocess_out.cpp:65-72)
    o.self_loop_weight = 1;  // adjacency lists are deduplicated (:101-126)
    HostCsr g;
    int rc = preprocess_edges(a.data(), b.data(), (int64_t)m, g, o);
    if (rc)
        return die(rc);
    trust_csr c{g.V, g.E, g.beg.data(), g.adj.data()};
    rc = trust_csr_save(dir.c_str(), &c);
    return rc ? die(rc) : 0;
}

static int cmd_undirect(int argc, char **argv)
{
    if (argc < 1)
        return 2;
    std::string out = argc > 1 ? argv[1] : "1.mmio";
    PairBlocks pb;
    long long bad = LLONG_MIN;
    if (!read_pairs(argv[0], [&](int32_t u, int32_t v) { pb.push(u, v); }, false, nullptr, 0,
                    &bad)) {
        if (bad != LLONG_MIN)
            fprintf(stderr, "trust_tool: %s: negative vertex id %lld\n", argv[0], bad);
        else
            fprintf(stderr, "trust_tool: cannot read %s\n", argv[0]);
        return 1;
    }
    int32_t maxid = -1;
    for (size_t q = 0; q < pb.blocks.size(); q++) {
        const int32_t *p = pb.blocks[q].get();
        for (size_t i = 0; i < 2 * pb.size_of(q); i++)
            maxid = std::max(maxid, p[i]);
    }
    std::vector<int32_t> id((size_t)maxid + 1, 0);
    for (size_t q = 0; q < pb.blocks.size(); q++) {
        const int32_t *p = pb.blocks[q].get();
        for (size_t i = 0; i < 2 * pb.size_of(q); i++)
            id[p[i]] = 1;
    }
    int64_t k = 0;
    for (auto &x : id)
        x = x ? (int32_t)++k : 0;  // 1-based
    // keys (u < v) block by block, each block freed once converted; the sort's
    // scratch is allocated after: 16 bytes per pair at the peak
    std::vector<uint64_t> keys(pb.n);
    size_t e = 0;
    for (size_t q = 0; q < pb.blocks.size(); q++) {
        const int32_t *p = pb.blocks[q].get();
        for (size_t i = 0; i < pb.size_of(q); i++, e++) {
            uint64_t u = (uint64_t)id[p[2 * i]], v = (uint64_t)id[p[2 * i + 1]];
            if (u > v)
                std::swap(u, v);
            keys[e] = (u << 32) | v;
        }
        pb.blocks[q].reset();
    }
    std::vector<int32_t>().swap(id);
    std::vector<uint64_t> tmp(keys.size());
    radix_sort_u64(keys.data(), tmp.data(), keys.size(), 64);
    std::vector<uint64_t>().swap(tmp);
    keys.erase(std::unique(keys.begin(), keys.end()), keys.end());
    FILE *f = fopen(out.c_str(), "w");
    if (!f) {
        fprintf(stderr, "trust_tool: cannot write %s\n", out.c_str());
        return 1;
    }
    fprintf(f, "%lld %lld %zu\n", (long long)k, (long long)k, keys.size());
    std::vector<char> obuf((size_t)1 << 24);
    size_t n = 0;
    auto put = [&](uint32_t x, char end) {
        char tmp[12];
        int t = 0;
        do {
            tmp[t++] = (char)('0' + x % 10);
            x /= 10;
        } while (x);
        while (t)
            obuf[n++] = tmp[--t];
        obuf[n++] = end;
    };
    for (uint64_t x : keys) {
        put((uint32_t)(x >> 32), ' ');
        put((uint32_t)(x & 0xffffffffu), '\n');
        if (n > obuf.size() - 32) {
            fwrite(obuf.data(), 1, n, f);
            n = 0;
        }
    }
    fwrite(obuf.data(), 1, n, f);
    return fclose(f) == 0 ? 0 : 1;
}

// The `edge` side file of the otherPreprocess tools (preprocess.cpp:115-131,
// preprocess_indegree.cpp:115-125), written from the final CSR.
static int write_edge_file(const std::string &outdir, const HostCsr &g, int order)
{
    std::vector<int32_t> out;
    if (order == TRUST_ORDER_PLAIN) {  // (row, id) for every entry
        out.resize(2 * (size_t)g.E);
        for (int64_t i = 0; i < g.V; i++)
            for (int64_t e = g.beg[i]; e < g.beg[i + 1]; e++) {
                out[2 * e] = (int32_t)i;
                out[2 * e + 1] = g.adj[e];
            }
    } else {
        // k = the tools' binary_search(32) over row lengths (:28-39), then per
        // row the std::upper_bound of k in the row as stored (:122-123)
        int64_t l = 0, r = g.V - 1;
        while (l < r - 1) {
            const int64_t mid = (l + r) >> 1;
            if (g.beg[mid + 1] - g.beg[mid] >= 32)
                l = mid;
            else
                r = mid;
        }
        const int32_t k = (int32_t)l;
        out.resize((size_t)g.V);
        for (int64_t i = 0; i < g.V; i++) {
            const int32_t *row = g.adj.data() + g.beg[i], *end = g.adj.data() + g.beg[i + 1];
            out[i] = (int32_t)(std::upper_bound(row, end, k) - row);
        }
    }
    return write_bytes(outdir + "/edge", out.data(), out.size() * sizeof(int32_t));
}

static int cmd_preprocess(int argc, char **argv)
{
    std::vector<std::string> pos;
    int order = TRUST_ORDER_OUTDEGREE;
    for (int i = 0; i < argc; i++) {
        if (!strcmp(argv[i], "--order") && i + 1 < argc) {
            const std::string o = argv[++i];
            if (o == "outdegree")
                order = TRUST_ORDER_OUTDEGREE;
            else if (o == "plain")
                order = TRUST_ORDER_PLAIN;
            else if (o == "indegree")
                order = TRUST_ORDER_INDEGREE;
            else
                return 2;
        } else {
            pos.push_back(argv[i]);
        }
    }
    std::string in = pos.size() > 0 ? pos[0] : "1.mmio";
    std::string outdir = pos.size() > 1 ? pos[1] : ".";
    std::vector<int32_t> a, b;
    long long hv[2] = {-1, -1}, bad = LLONG_MIN;
    if (!read_pairs(
            in,
            [&](int32_t u, int32_t v) {
                a.push_back(u);
                b.push_back(v);
            },
            true, hv, 1, &bad) ||
        hv[0] < 0 || hv[1] < 0) {
        if (bad != LLONG_MIN)
            fprintf(stderr, "trust_tool: %s: vertex id %lld (1.mmio ids start at 1)\n",
                    in.c_str(), bad);
        else
            fprintf(stderr, "trust_tool: cannot read %s\n", in.c_str());
        return 1;
    }
    for (size_t e = 0; e < a.size(); e++) {  // 1-based ids in the file
        a[e]--;
        b[e]--;
    }
    PreprocessOpts o;
    // the vertex count is the header's first field in preprocess_outdegree.cpp:62
    // and its second in otherPreprocess/preprocess*.cpp:60
    o.n_fixed = order == TRUST_ORDER_OUTDEGREE ? hv[0] : hv[1];
    o.order = order;
    HostCsr g;
    int rc = preprocess_edges(a.data(), b.data(), (int64_t)a.size(), g, o);
    if (rc)
        return die(rc);
    trust_csr c{g.V, g.E, g.beg.data(), g.adj.data()};
    rc = trust_csr_save(outdir.c_str(), &c);
    if (rc)
        return die(rc);
    return order == TRUST_ORDER_OUTDEGREE ? 0 : write_edge_file(outdir, g, order);
}

static int cmd_partition(int argc, char **argv)
{
    if (argc < 1)
        return 2;
    int n = atoi(argv[0]);
    std::string dir = argc > 1 ? argv[1] : ".";
    int rc = trust_partition_folder(dir.c_str(), n);  // one tile row in memory at a time
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

// gen-ooc rmat <scale> <ef> <seed> <outdir> [workdir] [bucket_pairs]: the
// R-MAT graph preprocessed out of core (trust_rmat_csr_ooc); follow with
// "partition <n> <outdir>" for its tiles.
static int cmd_gen_ooc(int argc, char **argv)
{
    if (argc < 5 || std::string(argv[0]) != "rmat")
        return 2;
    const int scale = atoi(argv[1]), ef = atoi(argv[2]);
    const uint64_t seed = strtoull(argv[3], nullptr, 10);
    const std::string outdir = argv[4];
    const std::string work = argc > 5 ? argv[5] : outdir;
    const long long bucket = argc > 6 ? atoll(argv[6]) : (1ll << 27);
    if (scale < 1 || scale > 30 || ef < 1)
        return 2;
    const int rc = trust_rmat_csr_ooc(scale, (int64_t)ef << scale, 0.57, 0.19, 0.19, seed,
                                      outdir.c_str(), work.c_str(), bucket);
    if (rc)
        return die(rc);
    printf("%s: R-MAT scale %d, edge factor %d, seed %llu, preprocessed out of core\n",
           outdir.c_str(), scale, ef, (unsigned long long)seed);
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
        else if (cmd == "txt2bin")
            rc = cmd_txt2bin(argc - 2, argv + 2);
        else if (cmd == "preprocess-bin")
            rc = cmd_preprocess_bin(argc - 2, argv + 2);
        else if (cmd == "partition")
            rc = cmd_partition(argc - 2, argv + 2);
        else if (cmd == "gen")
            rc = cmd_gen(argc - 2, argv + 2);
        else if (cmd == "gen-ooc")
            rc = cmd_gen_ooc(argc - 2, argv + 2);
    }
    if (rc == 2)
        fprintf(stderr,
                "usage: trust_tool undirect <edges.txt> [out]\n"
                "       trust_tool preprocess [1.mmio] [outdir] [--order outdegree|plain|indegree]\n"
                "       trust_tool txt2bin <edges.txt> [edge_count_hint]\n"
                "       trust_tool preprocess-bin [dir]\n"
                "       trust_tool partition <n> [dir]\n"
                "       trust_tool gen rmat <scale> <ef> <seed> <outdir>\n"
                "       trust_tool gen chunglu <n> <m> <gamma> <maxdeg> <seed> <outdir>\n"
                "       trust_tool gen er <n> <m> <seed> <outdir>\n"
                "       trust_tool gen ba <n> <k> <seed> <outdir>\n"
                "       trust_tool gen-ooc rmat <scale> <ef> <seed> <outdir> [workdir] [bucket_pairs]\n");
    return rc;
}
