// Host-side graph substrate: generators, CSR construction, `.graph` I/O.
//
// Interfaces follow include/belewitte_graph.h.  Everything here is plain
// C++17 + OpenMP so it loads and runs without a GPU.  Results never depend
// on the thread count: edges are drawn from a counter-based hash indexed by
// edge number, and the CSR is canonicalised by a per-row sort + unique after
// an (order-scrambling) parallel scatter.

#include "belewitte_graph.h"

#include <omp.h>

#include <algorithm>
#include <atomic>
#include <cmath>
#include <cstdio>
#include <cstring>
#include <memory>
#include <numeric>
#include <stdexcept>
#include <string>
#include <unordered_set>
#include <vector>

namespace {

thread_local std::string g_error;

int fail(int code, const std::string &msg)
{
    g_error = msg;
    return code;
}

}  // namespace

struct bw_graph {
    uint64_t V = 0, E = 0;
    bool undirected = false;
    std::vector<uint64_t> off, roff;  // roff/rcol empty when undirected
    std::vector<uint32_t> col, rcol;

    const std::vector<uint64_t> &offsets(bool rev) const
    { return (rev && !undirected) ? roff : off; }
    const std::vector<uint32_t> &cols(bool rev) const
    { return (rev && !undirected) ? rcol : col; }
};

extern "C" uint64_t bw_hash64(uint64_t seed, uint64_t stream, uint64_t ctr)
{
    // splitmix64 finaliser (Steele, Lea, Flood 2014) over a Weyl sequence.
    uint64_t z = seed + 0x9E3779B97F4A7C15ULL * (ctr + 1)
                      + 0xD1B54A32D192ED03ULL * (stream + 1);
    z = (z ^ (z >> 30)) * 0xBF58476D1CE4E5B9ULL;
    z = (z ^ (z >> 27)) * 0x94D049BB133111EBULL;
    return z ^ (z >> 31);
}

namespace {

inline uint64_t uniform_below(uint64_t h, uint64_t n)
{ return static_cast<uint64_t>((static_cast<unsigned __int128>(h) * n) >> 64); }

// Sorts the pairs into rows of `key` without atomics: a stable parallel
// partition by the key's high bits into at most 1024 contiguous row ranges,
// then one thread per range places its pairs row by row.  Each row is then
// sorted and deduplicated.  Returns offsets/cols with rows sorted ascending
// and duplicates merged.
void build_csr(uint64_t V, const std::vector<uint32_t> &key,
               const std::vector<uint32_t> &val, bool drop_self_loops,
               std::vector<uint64_t> &off, std::vector<uint32_t> &col)
{
    const uint64_t n = key.size();
    uint32_t vbits = 0;
    while ((1ULL << vbits) < V) vbits++;
    const uint32_t bbits = std::min<uint32_t>(10, vbits);
    const uint32_t shift = vbits - bbits;
    const uint64_t nb = 1ULL << bbits;
    const int T = omp_get_max_threads();
    const uint64_t chunk = (n + T - 1) / T;

    // Pass 1: per-thread bucket histograms.
    std::vector<uint64_t> hist(static_cast<size_t>(T) * nb, 0);
#pragma omp parallel num_threads(T)
    {
        const int t = omp_get_thread_num();
        uint64_t *h = &hist[static_cast<size_t>(t) * nb];
        const uint64_t b0 = std::min(n, t * chunk), b1 = std::min(n, b0 + chunk);
        for (uint64_t i = b0; i < b1; i++) {
            if (drop_self_loops && key[i] == val[i]) continue;
            h[key[i] >> shift]++;
        }
    }
    // Bucket-major, thread-minor positions keep the partition stable.
    std::vector<uint64_t> bstart(nb + 1, 0);
    uint64_t acc = 0;
    for (uint64_t b = 0; b < nb; b++) {
        bstart[b] = acc;
        for (int t = 0; t < T; t++) {
            uint64_t c = hist[static_cast<size_t>(t) * nb + b];
            hist[static_cast<size_t>(t) * nb + b] = acc;
            acc += c;
        }
    }
    bstart[nb] = acc;
    const uint64_t m = acc;
    std::vector<uint32_t> pk(m), pv(m);
#pragma omp parallel num_threads(T)
    {
        const int t = omp_get_thread_num();
        uint64_t *h = &hist[static_cast<size_t>(t) * nb];
        const uint64_t b0 = std::min(n, t * chunk), b1 = std::min(n, b0 + chunk);
        for (uint64_t i = b0; i < b1; i++) {
            if (drop_self_loops && key[i] == val[i]) continue;
            uint64_t pos = h[key[i] >> shift]++;
            pk[pos] = key[i];
            pv[pos] = val[i];
        }
    }

    // Pass 2: row counts per bucket (rows of a bucket are disjoint).
    std::vector<uint64_t> cnt(V + 1, 0);
#pragma omp parallel for schedule(dynamic, 1)
    for (int64_t b = 0; b < static_cast<int64_t>(nb); b++)
        for (uint64_t i = bstart[b]; i < bstart[b + 1]; i++) cnt[pk[i]]++;
    std::vector<uint64_t> start(V + 1);
    acc = 0;
    for (uint64_t v = 0; v < V; v++) { start[v] = acc; acc += cnt[v]; }
    start[V] = acc;

    std::vector<uint32_t> tmp(m);
#pragma omp parallel for schedule(dynamic, 1)
    for (int64_t b = 0; b < static_cast<int64_t>(nb); b++) {
        const uint64_t r0 = static_cast<uint64_t>(b) << shift;
        const uint64_t r1 = std::min<uint64_t>(V, static_cast<uint64_t>(b + 1) << shift);
        if (r0 >= r1) continue;
        std::vector<uint64_t> cur(start.begin() + static_cast<int64_t>(r0),
                                  start.begin() + static_cast<int64_t>(r1));
        for (uint64_t i = bstart[b]; i < bstart[b + 1]; i++) tmp[cur[pk[i] - r0]++] = pv[i];
    }
    std::vector<uint32_t>().swap(pk);
    std::vector<uint32_t>().swap(pv);

    // Per-row canonicalisation; cnt[v] becomes the deduplicated degree.
#pragma omp parallel for schedule(dynamic, 4096)
    for (int64_t v = 0; v < static_cast<int64_t>(V); v++) {
        auto b = tmp.begin() + static_cast<int64_t>(start[v]);
        auto e = tmp.begin() + static_cast<int64_t>(start[v + 1]);
        std::sort(b, e);
        cnt[v] = static_cast<uint64_t>(std::unique(b, e) - b);
    }

    off.assign(V + 1, 0);
    acc = 0;
    for (uint64_t v = 0; v < V; v++) { off[v] = acc; acc += cnt[v]; }
    off[V] = acc;
    col.resize(acc);
#pragma omp parallel for schedule(dynamic, 4096)
    for (int64_t v = 0; v < static_cast<int64_t>(V); v++) {
        std::copy_n(tmp.begin() + static_cast<int64_t>(start[v]), cnt[v],
                    col.begin() + static_cast<int64_t>(off[v]));
    }
}

// Reverse CSR from a canonical forward CSR.  Rows come out sorted by source
// because the scatter is followed by the same per-row sort.
void build_reverse(uint64_t V, const std::vector<uint64_t> &off,
                   const std::vector<uint32_t> &col,
                   std::vector<uint64_t> &roff, std::vector<uint32_t> &rcol)
{
    const uint64_t E = col.size();
    std::vector<uint32_t> src(E);
#pragma omp parallel for schedule(dynamic, 4096)
    for (int64_t v = 0; v < static_cast<int64_t>(V); v++) {
        for (uint64_t j = off[v]; j < off[v + 1]; j++) src[j] = static_cast<uint32_t>(v);
    }
    build_csr(V, col, src, false, roff, rcol);
}

std::unique_ptr<bw_graph> finish(uint64_t V, std::vector<uint32_t> &src,
                                 std::vector<uint32_t> &dst, bool undirected,
                                 bool keep_self_loops)
{
    auto g = std::make_unique<bw_graph>();
    g->V = V;
    g->undirected = undirected;
    if (undirected) {
        const size_t n = src.size();
        src.resize(2 * n);
        dst.resize(2 * n);
        std::copy_n(dst.begin(), n, src.begin() + static_cast<int64_t>(n));
        std::copy_n(src.begin(), n, dst.begin() + static_cast<int64_t>(n));
    }
    build_csr(V, src, dst, !keep_self_loops, g->off, g->col);
    std::vector<uint32_t>().swap(src);
    std::vector<uint32_t>().swap(dst);
    g->E = g->col.size();
    if (!undirected) build_reverse(V, g->off, g->col, g->roff, g->rcol);
    return g;
}

struct RmatThresholds {
    uint64_t ta, tab, tabc;  // 32-bit fixed-point cumulative probabilities
    RmatThresholds(double a, double b, double c)
      : ta(static_cast<uint64_t>(std::llround(a * 4294967296.0)))
      , tab(static_cast<uint64_t>(std::llround((a + b) * 4294967296.0)))
      , tabc(static_cast<uint64_t>(std::llround((a + b + c) * 4294967296.0)))
    {}

    // One edge of an R-MAT graph of the given scale; draws use counters
    // ctr*32 .. ctr*32 + levels/2 of stream 0.
    void edge(uint64_t seed, uint64_t ctr, uint32_t scale, uint32_t &s,
              uint32_t &d) const
    {
        uint32_t su = 0, du = 0;
        uint64_t h = 0;
        for (uint32_t l = 0; l < scale; l++) {
            uint64_t r;
            if ((l & 1) == 0) {
                h = bw_hash64(seed, 0, ctr * 32 + l / 2);
                r = h >> 32;
            } else {
                r = h & 0xFFFFFFFFULL;
            }
            uint32_t sb, db;
            if (r < ta) { sb = 0; db = 0; }
            else if (r < tab) { sb = 0; db = 1; }
            else if (r < tabc) { sb = 1; db = 0; }
            else { sb = 1; db = 1; }
            su = (su << 1) | sb;
            du = (du << 1) | db;
        }
        s = su;
        d = du;
    }
};

bool valid_probs(double a, double b, double c)
{ return a >= 0 && b >= 0 && c >= 0 && a + b + c <= 1.0; }

template<typename F>
int guarded(F &&f)
{
    try {
        return f();
    } catch (const std::bad_alloc &) {
        return fail(BW_ERR_NOMEM, "out of host memory");
    } catch (const std::exception &e) {
        return fail(BW_ERR_ARG, e.what());
    }
}

}  // namespace

extern "C" const char *bw_graph_last_error(void) { return g_error.c_str(); }

extern "C" int bw_graph_rmat(bw_graph **out, uint32_t scale, uint32_t ef,
                             double a, double b, double c, uint64_t seed,
                             int undirected)
{
    if (!out || scale == 0 || scale > 31 || ef == 0 || !valid_probs(a, b, c))
        return fail(BW_ERR_ARG, "bw_graph_rmat: invalid argument");
    return guarded([&] {
        const uint64_t V = 1ULL << scale;
        const uint64_t n = static_cast<uint64_t>(ef) * V;
        RmatThresholds th(a, b, c);
        std::vector<uint32_t> src(n), dst(n);
#pragma omp parallel for schedule(static)
        for (int64_t i = 0; i < static_cast<int64_t>(n); i++) {
            th.edge(seed, static_cast<uint64_t>(i), scale, src[i], dst[i]);
        }
        *out = finish(V, src, dst, undirected != 0, false).release();
        return BW_OK;
    });
}

extern "C" int bw_graph_rmat_n(bw_graph **out, uint64_t V, uint64_t n,
                               double a, double b, double c, uint64_t seed,
                               int undirected)
{
    if (!out || V < 2 || V > (1ULL << 31) || !valid_probs(a, b, c))
        return fail(BW_ERR_ARG, "bw_graph_rmat_n: invalid argument");
    return guarded([&] {
        uint32_t scale = 0;
        while ((1ULL << scale) < V) scale++;
        RmatThresholds th(a, b, c);
        std::vector<uint32_t> src(n), dst(n);
#pragma omp parallel for schedule(static)
        for (int64_t i = 0; i < static_cast<int64_t>(n); i++) {
            // Redraw from counters i + k*n until both endpoints are < V.
            uint64_t ctr = static_cast<uint64_t>(i);
            uint32_t s, d;
            for (;;) {
                th.edge(seed, ctr, scale, s, d);
                if (s < V && d < V) break;
                ctr += n;
            }
            src[i] = s;
            dst[i] = d;
        }
        *out = finish(V, src, dst, undirected != 0, false).release();
        return BW_OK;
    });
}

extern "C" int bw_graph_erdos_renyi(bw_graph **out, uint64_t V, uint64_t pairs,
                                    uint64_t seed, int undirected)
{
    if (!out || V == 0 || V > (1ULL << 32))
        return fail(BW_ERR_ARG, "bw_graph_erdos_renyi: invalid argument");
    return guarded([&] {
        std::vector<uint32_t> src(pairs), dst(pairs);
#pragma omp parallel for schedule(static)
        for (int64_t i = 0; i < static_cast<int64_t>(pairs); i++) {
            src[i] = static_cast<uint32_t>(uniform_below(bw_hash64(seed, 1, 2 * i), V));
            dst[i] = static_cast<uint32_t>(uniform_below(bw_hash64(seed, 1, 2 * i + 1), V));
        }
        *out = finish(V, src, dst, undirected != 0, false).release();
        return BW_OK;
    });
}

extern "C" int bw_graph_grid2d(bw_graph **out, uint32_t W, uint32_t H,
                               double rewire, uint64_t seed)
{
    if (!out || W == 0 || H == 0 || rewire < 0 || rewire > 1 ||
        static_cast<uint64_t>(W) * H > (1ULL << 32))
        return fail(BW_ERR_ARG, "bw_graph_grid2d: invalid argument");
    return guarded([&] {
        const uint64_t V = static_cast<uint64_t>(W) * H;
        // Canonical undirected edge order of evolve/gen-graph.cpp:99-123:
        // row-major over vertices, right edge then down edge.
        const uint64_t M = static_cast<uint64_t>(W - 1) * H
                         + static_cast<uint64_t>(H - 1) * W;
        std::vector<uint32_t> src(M), dst(M);
        {
            uint64_t k = 0;
            for (uint64_t y = 0; y < H; y++) {
                for (uint64_t x = 0; x < W; x++) {
                    uint32_t v = static_cast<uint32_t>(y * W + x);
                    if (x + 1 < W) { src[k] = v; dst[k] = v + 1; k++; }
                    if (y + 1 < H) { src[k] = v; dst[k] = static_cast<uint32_t>(v + W); k++; }
                }
            }
        }
        const uint64_t R = static_cast<uint64_t>(std::llround(rewire * static_cast<double>(M)));
        if (R > 0) {
            // The R edges with the smallest hash are rewired.
            std::vector<std::pair<uint64_t, uint64_t>> rank(M);
#pragma omp parallel for schedule(static)
            for (int64_t k = 0; k < static_cast<int64_t>(M); k++)
                rank[k] = {bw_hash64(seed, 2, k), static_cast<uint64_t>(k)};
            std::nth_element(rank.begin(), rank.begin() + static_cast<int64_t>(R - 1), rank.end());
            std::vector<uint64_t> chosen(R);
            for (uint64_t i = 0; i < R; i++) chosen[i] = rank[i].second;
            std::vector<std::pair<uint64_t, uint64_t>>().swap(rank);
            std::sort(chosen.begin(), chosen.end());

            // Presence test against the current (symmetric) edge set; the
            // rewiring is applied sequentially in edge order so it is
            // deterministic.
            std::vector<uint64_t> keys;
            keys.reserve(2 * M);
            for (uint64_t k = 0; k < M; k++) {
                keys.push_back((static_cast<uint64_t>(src[k]) << 32) | dst[k]);
                keys.push_back((static_cast<uint64_t>(dst[k]) << 32) | src[k]);
            }
            std::sort(keys.begin(), keys.end());
            std::unordered_set<uint64_t> added;
            auto present = [&](uint32_t u, uint32_t w) {
                uint64_t key = (static_cast<uint64_t>(u) << 32) | w;
                if (std::binary_search(keys.begin(), keys.end(), key)) return true;
                return added.count(key) != 0;
            };
            for (uint64_t i = 0; i < R; i++) {
                uint64_t k = chosen[i];
                uint32_t u = src[k];
                for (uint64_t attempt = 0; attempt < 64; attempt++) {
                    uint32_t w = static_cast<uint32_t>(
                        uniform_below(bw_hash64(seed, 3, k * 64 + attempt), V));
                    if (w == u || present(u, w)) continue;
                    dst[k] = w;
                    added.insert((static_cast<uint64_t>(u) << 32) | w);
                    added.insert((static_cast<uint64_t>(w) << 32) | u);
                    break;
                }
            }
        }
        *out = finish(V, src, dst, true, false).release();
        return BW_OK;
    });
}

extern "C" int bw_graph_chain(bw_graph **out, uint64_t n, int undirected)
{
    if (!out || n >= (1ULL << 32)) return fail(BW_ERR_ARG, "bw_graph_chain: invalid argument");
    return guarded([&] {
        std::vector<uint32_t> s(n), d(n);
        for (uint64_t i = 0; i < n; i++) { s[i] = static_cast<uint32_t>(i); d[i] = static_cast<uint32_t>(i + 1); }
        *out = finish(n + 1, s, d, undirected != 0, false).release();
        return BW_OK;
    });
}

extern "C" int bw_graph_star(bw_graph **out, uint64_t n, int undirected)
{
    if (!out || n >= (1ULL << 32)) return fail(BW_ERR_ARG, "bw_graph_star: invalid argument");
    return guarded([&] {
        std::vector<uint32_t> s(n, 0), d(n);
        for (uint64_t i = 0; i < n; i++) d[i] = static_cast<uint32_t>(i + 1);
        *out = finish(n + 1, s, d, undirected != 0, false).release();
        return BW_OK;
    });
}

extern "C" int bw_graph_mesh(bw_graph **out, uint64_t H, uint64_t W, int undirected)
{
    if (!out || H == 0 || W == 0 || H * W > (1ULL << 32))
        return fail(BW_ERR_ARG, "bw_graph_mesh: invalid argument");
    return guarded([&] {
        std::vector<uint32_t> s, d;
        for (uint64_t j = 0; j < H; j++) {
            for (uint64_t i = 0; i < W; i++) {
                if (i < W - 1) { s.push_back(static_cast<uint32_t>(j * W + i)); d.push_back(static_cast<uint32_t>(j * W + i + 1)); }
                if (j < H - 1) { s.push_back(static_cast<uint32_t>(j * W + i)); d.push_back(static_cast<uint32_t>((j + 1) * W + i)); }
            }
        }
        *out = finish(H * W, s, d, undirected != 0, false).release();
        return BW_OK;
    });
}

extern "C" int bw_graph_from_edges(bw_graph **out, uint64_t V, uint64_t n,
                                   const uint32_t *src, const uint32_t *dst,
                                   int undirected, int keep_self_loops)
{
    if (!out || (n && (!src || !dst)) || V > (1ULL << 32))
        return fail(BW_ERR_ARG, "bw_graph_from_edges: invalid argument");
    for (uint64_t i = 0; i < n; i++) {
        if (src[i] >= V || dst[i] >= V)
            return fail(BW_ERR_ARG, "bw_graph_from_edges: vertex id out of range at edge "
                                    + std::to_string(i));
    }
    return guarded([&] {
        std::vector<uint32_t> s(src, src + n), d(dst, dst + n);
        *out = finish(V, s, d, undirected != 0, keep_self_loops != 0).release();
        return BW_OK;
    });
}

// ---------------------------------------------------------------- file I/O

namespace {

struct FileCloser { void operator()(FILE *f) const { if (f) fclose(f); } };
using File = std::unique_ptr<FILE, FileCloser>;

bool read_exact(FILE *f, void *p, size_t n) { return fread(p, 1, n, f) == n; }
bool write_exact(FILE *f, const void *p, size_t n) { return fwrite(p, 1, n, f) == n; }

template<typename T>
bool read_array(FILE *f, uint32_t width, uint64_t count, std::vector<T> &dst)
{
    dst.resize(count);
    const size_t chunk = 1 << 20;
    if (width == sizeof(T)) return read_exact(f, dst.data(), count * sizeof(T));
    std::vector<uint8_t> buf;
    for (uint64_t i = 0; i < count; i += chunk) {
        uint64_t m = std::min<uint64_t>(chunk, count - i);
        buf.resize(m * width);
        if (!read_exact(f, buf.data(), buf.size())) return false;
        for (uint64_t j = 0; j < m; j++) {
            uint64_t v;
            if (width == 4) { uint32_t x; std::memcpy(&x, &buf[j * 4], 4); v = x; }
            else { std::memcpy(&v, &buf[j * 8], 8); }
            dst[i + j] = static_cast<T>(v);
        }
    }
    return true;
}

template<typename T>
bool write_array(FILE *f, uint32_t width, const std::vector<T> &src)
{
    if (width == sizeof(T)) return write_exact(f, src.data(), src.size() * sizeof(T));
    const size_t chunk = 1 << 20;
    std::vector<uint8_t> buf;
    for (size_t i = 0; i < src.size(); i += chunk) {
        size_t m = std::min(chunk, src.size() - i);
        buf.resize(m * width);
        for (size_t j = 0; j < m; j++) {
            uint64_t v = src[i + j];
            if (width == 4) { uint32_t x = static_cast<uint32_t>(v); std::memcpy(&buf[j * 4], &x, 4); }
            else std::memcpy(&buf[j * 8], &v, 8);
        }
        if (!write_exact(f, buf.data(), buf.size())) return false;
    }
    return true;
}

}  // namespace

extern "C" int bw_graph_load(bw_graph **out, const char *path)
{
    if (!out || !path) return fail(BW_ERR_ARG, "bw_graph_load: invalid argument");
    return guarded([&]() -> int {
        File f(fopen(path, "rb"));
        if (!f) return fail(BW_ERR_IO, std::string("Failed to open graph: ") + path);
        fseeko(f.get(), 0, SEEK_END);
        const uint64_t size = static_cast<uint64_t>(ftello(f.get()));
        fseeko(f.get(), 0, SEEK_SET);
        if (size < 12) return fail(BW_ERR_FORMAT, "Invalid file size! Wrong format?");

        uint32_t w[10] = {0};
        if (!read_exact(f.get(), w, 12)) return fail(BW_ERR_IO, "read failed");
        // Version detection of utils/Graph.hpp:788-792.
        const uint32_t version = (w[1] == 0 && w[2] == 0) ? w[0] : 0;
        auto g = std::make_unique<bw_graph>();
        uint32_t offw = 4, idw = 4;
        uint64_t expect;
        if (version == 0) {
            g->undirected = w[0] != 0;
            g->V = static_cast<int32_t>(w[1]);
            g->E = static_cast<int32_t>(w[2]);
            expect = 3 + g->V + 1 + g->E;
            if (!g->undirected) expect += g->V + 1 + g->E;
            expect *= 4;
        } else if (version == 1) {
            if (size < 40 || !read_exact(f.get(), &w[3], 28)) return fail(BW_ERR_FORMAT, "Invalid file size! Wrong format?");
            g->undirected = w[3] != 0;
            offw = w[4];
            idw = w[5];
            std::memcpy(&g->V, &w[6], 8);
            std::memcpy(&g->E, &w[8], 8);
            if ((offw != 4 && offw != 8) || (idw != 4 && idw != 8))
                return fail(BW_ERR_FORMAT, "Invalid graph file version or vertex/edge size!");
            expect = 40 + (g->V + 1) * offw + g->E * idw;
            if (!g->undirected) expect += (g->V + 1) * offw + g->E * idw;
        } else {
            return fail(BW_ERR_FORMAT, "Invalid graph file version or vertex/edge size!");
        }
        if (size != expect)
            return fail(BW_ERR_FORMAT, "Invalid file size! Wrong format? Expected: "
                        + std::to_string(expect) + " Found: " + std::to_string(size));
        if (g->V > (1ULL << 32)) return fail(BW_ERR_FORMAT, "vertex count exceeds 32-bit ids");
        if (!read_array(f.get(), offw, g->V + 1, g->off) || !read_array(f.get(), idw, g->E, g->col))
            return fail(BW_ERR_IO, "short read");
        if (!g->undirected) {
            if (!read_array(f.get(), offw, g->V + 1, g->roff) || !read_array(f.get(), idw, g->E, g->rcol))
                return fail(BW_ERR_IO, "short read");
        }
        // Structural validation so the kernels can trust the arrays.
        for (int rev = 0; rev < (g->undirected ? 1 : 2); rev++) {
            const auto &o = rev ? g->roff : g->off;
            const auto &c = rev ? g->rcol : g->col;
            if (o[0] != 0 || o[g->V] != g->E) return fail(BW_ERR_FORMAT, "corrupt offsets");
            for (uint64_t v = 0; v < g->V; v++)
                if (o[v] > o[v + 1]) return fail(BW_ERR_FORMAT, "corrupt offsets");
            for (uint32_t x : c)
                if (x >= g->V) return fail(BW_ERR_FORMAT, "edge endpoint out of range");
        }
        *out = g.release();
        return BW_OK;
    });
}

extern "C" int bw_graph_save(const bw_graph *g, const char *path)
{
    if (!g || !path) return fail(BW_ERR_ARG, "bw_graph_save: invalid argument");
    return guarded([&]() -> int {
        File f(fopen(path, "wb"));
        if (!f) return fail(BW_ERR_IO, std::string("cannot create ") + path);
        // Field widths as in utils/Graph.hpp:1064-1080 (smallest_size).
        const uint32_t offw = g->E > 0xFFFFFFFFULL ? 8 : 4;
        const uint32_t idw = g->V > 0xFFFFFFFFULL ? 8 : 4;
        uint32_t w[10] = {1, 0, 0, g->undirected ? 1u : 0u, offw, idw, 0, 0, 0, 0};
        std::memcpy(&w[6], &g->V, 8);
        std::memcpy(&w[8], &g->E, 8);
        bool ok = write_exact(f.get(), w, sizeof w)
               && write_array(f.get(), offw, g->off)
               && write_array(f.get(), idw, g->col);
        if (ok && !g->undirected)
            ok = write_array(f.get(), offw, g->roff) && write_array(f.get(), idw, g->rcol);
        if (!ok) return fail(BW_ERR_IO, "write failed");
        return BW_OK;
    });
}

extern "C" void bw_graph_free(bw_graph *g) { delete g; }
extern "C" uint64_t bw_graph_vertex_count(const bw_graph *g) { return g ? g->V : 0; }
extern "C" uint64_t bw_graph_edge_count(const bw_graph *g) { return g ? g->E : 0; }
extern "C" int bw_graph_is_undirected(const bw_graph *g) { return g && g->undirected; }

extern "C" const uint64_t *bw_graph_offsets(const bw_graph *g, int reverse)
{ return g ? g->offsets(reverse != 0).data() : nullptr; }

extern "C" const uint32_t *bw_graph_cols(const bw_graph *g, int reverse)
{ return g ? g->cols(reverse != 0).data() : nullptr; }

extern "C" int bw_graph_offsets32(const bw_graph *g, int reverse, uint32_t *dst)
{
    if (!g || !dst) return fail(BW_ERR_ARG, "bw_graph_offsets32: invalid argument");
    if (g->E > 0xFFFFFFFFULL) return fail(BW_ERR_ARG, "edge count does not fit uint32 offsets");
    const auto &o = g->offsets(reverse != 0);
#pragma omp parallel for schedule(static)
    for (int64_t i = 0; i < static_cast<int64_t>(o.size()); i++) dst[i] = static_cast<uint32_t>(o[i]);
    return BW_OK;
}

extern "C" uint32_t bw_graph_max_out_degree_vertex(const bw_graph *g)
{
    uint32_t best = 0;
    uint64_t bestDeg = 0;
    for (uint64_t v = 0; v < g->V; v++) {
        uint64_t d = g->off[v + 1] - g->off[v];
        if (d > bestDeg) { bestDeg = d; best = static_cast<uint32_t>(v); }
    }
    return best;
}

extern "C" int bw_graph_degree_stats(const bw_graph *g, int kind, double out[7])
{
    if (!g || !out || kind < 0 || kind > 2 || g->V == 0)
        return fail(BW_ERR_ARG, "bw_graph_degree_stats: invalid argument");
    return guarded([&] {
        std::vector<double> d(g->V);
        const auto &ro = g->offsets(true);
        for (uint64_t v = 0; v < g->V; v++) {
            uint64_t x = 0;
            if (kind == 1 || kind == 2) x += g->off[v + 1] - g->off[v];
            if (kind == 0 || (kind == 2 && !g->undirected)) x += ro[v + 1] - ro[v];
            d[v] = static_cast<double>(x);
        }
        // Index pairs of utils/StatisticalSummary.hpp:11-46.
        auto mid = [](size_t n) -> std::pair<size_t, size_t> {
            if (n % 2 == 0) return {n / 2 - 1, n / 2};
            return {n / 2, n / 2};
        };
        const size_t n = d.size(), half = n / 2;
        std::pair<size_t, size_t> med{0, 0}, lo{0, 0}, hi{0, 0};
        if (n > 1) {
            med = mid(n);
            lo = mid(half);
            hi = lo;
            hi.first += half + n % 2;
            hi.second += half + n % 2;
        }
        std::sort(d.begin(), d.end());
        double total = 0, M = 0, S = 0;
        int64_t k = 1;
        for (double v : d) {
            total += v;
            double tM = M;
            M += (v - tM) / static_cast<double>(k);
            S += (v - tM) * (v - M);
            k++;
        }
        out[0] = d.front();
        out[1] = (d[lo.first] + d[lo.second]) / 2.0;
        out[2] = total / static_cast<double>(n);
        out[3] = (d[med.first] + d[med.second]) / 2.0;
        out[4] = (d[hi.first] + d[hi.second]) / 2.0;
        out[5] = d.back();
        out[6] = (k - 2 == 0) ? 0.0 : std::sqrt(S / static_cast<double>(k - 2));
        return BW_OK;
    });
}

// Relabels one stored CSR: new row i is old row order[i], its columns
// mapped through perm and sorted.  A pure permutation: every stored edge,
// duplicates and self-loops included, keeps its multiplicity.
static void permute_csr(uint64_t V, const std::vector<uint64_t> &off, const std::vector<uint32_t> &col,
                        const std::vector<uint32_t> &order, const std::vector<uint32_t> &perm,
                        std::vector<uint64_t> &noff, std::vector<uint32_t> &ncol)
{
    noff.assign(V + 1, 0);
    for (uint64_t i = 0; i < V; i++) noff[i + 1] = noff[i] + (off[order[i] + 1] - off[order[i]]);
    ncol.resize(noff[V]);
#pragma omp parallel for schedule(dynamic, 4096)
    for (int64_t i = 0; i < static_cast<int64_t>(V); i++) {
        const uint64_t v = order[static_cast<uint64_t>(i)];
        uint64_t k = noff[static_cast<uint64_t>(i)];
        for (uint64_t j = off[v]; j < off[v + 1]; j++) ncol[k++] = perm[col[j]];
        std::sort(ncol.begin() + static_cast<int64_t>(noff[static_cast<uint64_t>(i)]),
                  ncol.begin() + static_cast<int64_t>(k));
    }
}

extern "C" int bw_graph_reorder_by_degree(bw_graph **out, const bw_graph *g,
                                          uint32_t *perm_out)
{
    if (!out || !g) return fail(BW_ERR_ARG, "bw_graph_reorder_by_degree: invalid argument");
    return guarded([&] {
        const uint64_t V = g->V;
        std::vector<uint32_t> order(V);
        std::iota(order.begin(), order.end(), 0u);
        std::stable_sort(order.begin(), order.end(), [&](uint32_t a, uint32_t b) {
            uint64_t da = g->off[a + 1] - g->off[a], db = g->off[b + 1] - g->off[b];
            return da > db || (da == db && a < b);
        });
        std::vector<uint32_t> perm(V);
        for (uint64_t i = 0; i < V; i++) perm[order[i]] = static_cast<uint32_t>(i);
        // Both stored directions are permuted as they are (an undirected
        // graph stores one CSR), so the edge count never changes.
        auto r = std::make_unique<bw_graph>();
        r->V = V;
        r->E = g->E;
        r->undirected = g->undirected;
        permute_csr(V, g->off, g->col, order, perm, r->off, r->col);
        if (!g->undirected) permute_csr(V, g->roff, g->rcol, order, perm, r->roff, r->rcol);
        if (perm_out) std::copy(perm.begin(), perm.end(), perm_out);
        *out = r.release();
        return BW_OK;
    });
}
