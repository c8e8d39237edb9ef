// CPU check of the binned PageRank plan (belewitte_amd/csrc/hip/
// pagerank_bins_plan.hpp): builds the plan from a reverse CSR, checks that
// every index the two kernels derive from it stays in bounds and that every
// vertex, pair, task and slot is covered exactly once, then replays the
// kernels' arithmetic sequentially (the accumulate kernel's per-lane flag
// prefix included) and prints the largest difference from an FP64 PageRank
// of the same graph, and that of the sequential FP32 pull iteration (the
// reference's per-row CSR-order sums).  Usage: bins_plan_check OFF COLS
// ROWS_LOG2 CHUNK_TASKS ITERATIONS UNIFORM.  Exit status 3 on a failed
// check, 4 when the plan is refused.  Rows accumulate in FP64 and each
// chunk's partial is rounded to FP32, as in the accumulate kernel.
#include "pagerank_bins_plan.hpp"

#include <cmath>
#include <cstdio>
#include <cstdlib>
#include <vector>

#define CHECK(c)                                                                 \
    do {                                                                         \
        if (!(c)) {                                                              \
            std::fprintf(stderr, "check failed: %s (line %d)\n", #c, __LINE__);  \
            std::exit(3);                                                        \
        }                                                                        \
    } while (0)

using namespace bw::bins;

static std::vector<uint32_t> read_u32(const char *path)
{
    FILE *f = std::fopen(path, "rb");
    CHECK(f != nullptr);
    std::fseek(f, 0, SEEK_END);
    const long n = std::ftell(f) / 4;
    std::fseek(f, 0, SEEK_SET);
    std::vector<uint32_t> v(n);
    CHECK(n == 0 || std::fread(v.data(), 4, n, f) == static_cast<size_t>(n));
    std::fclose(f);
    return v;
}

int main(int argc, char **argv)
{
    if (argc != 7) return 2;
    const auto off = read_u32(argv[1]);
    const auto cols = read_u32(argv[2]);
    const uint32_t L = std::atoi(argv[3]), chunkTasks = std::atoi(argv[4]);
    const int iters = std::atoi(argv[5]);
    const bool uniform = std::atoi(argv[6]) != 0;
    const uint32_t V = static_cast<uint32_t>(off.size() - 1), E = static_cast<uint32_t>(cols.size());
    Plan P;
    std::string err;
    if (!build_plan(off.data(), cols.data(), V, E, L, chunkTasks, P, err)) {
        std::printf("plan error: %s\n", err.c_str());
        return 4;
    }
    const uint32_t D = 1u << L;
    const uint32_t V_ = V;
    (void)V_;

    // Coverage and bounds.
    std::vector<int> vseen(V, 0), pseen(P.pairs, 0), tseen(P.tasks.size(), 0), sseen(P.slots, 0);
    uint32_t flatNext = 0;
    for (const Quad &b : P.blocks) {
        CHECK(b.x < b.y && b.y <= V && b.y - b.x <= kBlockVerts);
        CHECK((b.x >> L) == ((b.y - 1) >> L));
        CHECK(b.w <= kMaxRuns && static_cast<size_t>(b.z) + b.w < P.run_flat.size());
        for (uint32_t v = b.x; v < b.y; v++) vseen[v]++;
        CHECK(P.run_flat[b.z] == flatNext);  // blocks' flat ranges follow each other
        flatNext = P.run_flat[b.z + b.w];
        CHECK(flatNext <= P.pairs);
        for (uint32_t i = 0; i < b.w; i++) {
            const uint32_t n = P.run_flat[b.z + i + 1] - P.run_flat[b.z + i];
            CHECK(n > 0);
            for (uint32_t j = 0; j < n; j++) {
                const uint32_t p = P.run_bin[b.z + i] + j;
                CHECK(p < P.pairs);
                pseen[p]++;
                CHECK(P.psrc[P.run_flat[b.z + i] + j] < b.y - b.x);
            }
        }
    }
    CHECK(flatNext == P.pairs);
    for (uint32_t v = 0; v < V; v++) CHECK(vseen[v] == 1);
    for (uint32_t p = 0; p < P.pairs; p++) CHECK(pseen[p] == 1);
    for (const Quad &c : P.chunks) {
        CHECK(c.x < c.y && c.y <= P.tasks.size() && c.z < P.K && c.w < P.slots);
        CHECK(c.w >= P.chunk_base[c.z] && c.w < P.chunk_base[c.z + 1]);
        sseen[c.w]++;
        for (uint32_t t = c.x; t < c.y; t++) {
            const Quad &tk = P.tasks[t];
            tseen[t]++;
            CHECK(tk.w == c.z && tk.x % 8 == 0 && tk.x < tk.y);
            CHECK(tk.x >= P.edge_start[c.z] && tk.y <= P.edge_start[c.z + 1]);
            CHECK(P.edge_start[c.z + 1] % 8 == 0 && P.edge_start[c.z + 1] <= P.codes.size());
        }
    }
    for (int x : tseen) CHECK(x == 1);
    for (int x : sseen) CHECK(x == 1);
    // The tasks of each partition cover its edges exactly once, and a chunk
    // has at most one task per wave of the accumulate workgroup.
    {
        std::vector<uint8_t> eseen(P.codes.size(), 0);
        for (const Quad &tk : P.tasks)
            for (uint32_t e = tk.x; e < tk.y; e++) eseen[e]++;
        for (uint32_t k = 0; k < P.K; k++) {
            const uint32_t r0 = k << L, r1 = std::min<uint64_t>(static_cast<uint64_t>(r0) + D, V);
            const uint32_t n = off[r1] - off[r0];
            for (uint32_t e = P.edge_start[k]; e < P.edge_start[k + 1]; e++)
                CHECK(eseen[e] == (e < P.edge_start[k] + n ? 1 : 0));
        }
        for (const Quad &c : P.chunks) CHECK(c.y - c.x <= kChunkWaves);
    }

    // The kernels' arithmetic, sequentially.
    std::vector<uint32_t> deg(V, 0);
    for (uint32_t e = 0; e < E; e++) deg[cols[e]]++;
    std::vector<float> bins(std::max<uint32_t>(P.pairs, 1)), rank(V);
    std::vector<float> partials(static_cast<size_t>(std::max<uint32_t>(P.slots, 1)) * D);
    const float kDamp = 0.85f;
    auto finalise = [&](bool phase, bool scatter, double Dm, double &dangl) {
        dangl = 0.0;
        std::vector<float> contrib(kBlockVerts);
        for (const Quad &b : P.blocks) {
            const uint32_t k = b.x >> L, rbase = k << L;
            for (uint32_t v = b.x; v < b.y; v++) {
                float nr, cv;
                const uint32_t dv = deg[v];
                if (!phase) {
                    nr = 1.0f / static_cast<float>(V);
                    cv = dv ? nr / static_cast<float>(dv) : nr;
                } else {
                    double s = 0.0;
                    for (uint32_t q = P.chunk_base[k]; q < P.chunk_base[k + 1]; q++)
                        s += partials[static_cast<size_t>(q) * D + (v - rbase)];
                    const float sum = static_cast<float>(s);
                    nr = uniform ? static_cast<float>(0.15 / V + 0.85 * (static_cast<double>(sum) + Dm / V))
                                 : std::fmaf(kDamp, sum, (1.0f - kDamp) / static_cast<float>(V));
                    cv = dv ? nr / static_cast<float>(dv) : 0.0f;
                }
                if (dv == 0) dangl += nr;
                rank[v] = nr;
                contrib[v - b.x] = cv;
            }
            if (!scatter) continue;
            // The finalise kernel's walk: 8 waves of 64 lanes, wave w over
            // the flat range [w * per, (w + 1) * per), each lane's run found
            // by binary search for its first flat and followed forward.
            const uint32_t *rflat = &P.run_flat[b.z];
            const uint32_t f0 = rflat[0], total = rflat[b.w] - f0, waves = 8;
            const uint32_t per = ((total + waves - 1) / waves + 63) & ~63u;
            for (uint32_t w = 0; w < waves; w++) {
                const uint32_t fbeg = f0 + std::min(w * per, total), fend = std::min(fbeg + per, f0 + total);
                for (uint32_t lane = 0; lane < 64; lane++) {
                    uint32_t lo = 0;
                    if (fbeg + lane < fend) {
                        uint32_t hi = b.w;
                        while (hi - lo > 1) {
                            const uint32_t mid = (lo + hi) >> 1;
                            if (rflat[mid] <= fbeg + lane) lo = mid;
                            else hi = mid;
                        }
                    }
                    for (uint32_t f = fbeg + lane; f < fend; f += 64) {
                        for (;;) {
                            CHECK(lo + 1 <= b.w);
                            if (rflat[lo + 1] > f) break;
                            lo++;
                        }
                        const uint32_t p = P.run_bin[b.z + lo] + (f - rflat[lo]);
                        CHECK(p < P.pairs && f < P.pairs && P.psrc[f] < b.y - b.x);
                        bins[p] = contrib[P.psrc[f]];
                    }
                }
            }
        }
    };
    auto accumulate = [&] {
        std::vector<double> acc(D);
        for (const Quad &ch : P.chunks) {
            std::fill(acc.begin(), acc.end(), 0.0);
            for (uint32_t t = ch.x; t < ch.y; t++) {
                const Quad &tk = P.tasks[t];
                uint32_t base = tk.z;
                for (uint32_t e0 = tk.x; e0 < tk.y; e0 += 512) {
                    uint32_t pre[64], n[64], tot = 0;
                    for (int l = 0; l < 64; l++) {
                        const uint32_t e = e0 + 8 * l;
                        n[l] = e < tk.y ? std::min(8u, tk.y - e) : 0;
                        pre[l] = tot;
                        for (uint32_t j = 0; j < n[l]; j++) tot += (P.codes[e + j] & kNewPair) ? 1 : 0;
                    }
                    for (int l = 0; l < 64; l++) {
                        uint32_t p = base + pre[l];
                        for (uint32_t j = 0; j < n[l]; j++) {
                            const uint16_t c = P.codes[e0 + 8 * l + j];
                            p += (c & kNewPair) ? 1 : 0;
                            CHECK(p < P.pairs && (c & (D - 1)) == (c & 0x7FFFu));
                            acc[c & (D - 1)] += bins[p];
                        }
                    }
                    base += tot;
                }
            }
            for (uint32_t i = 0; i < D; i++)
                partials[static_cast<size_t>(ch.w) * D + i] = static_cast<float>(acc[i]);
        }
    };
    double dangl = 0.0;
    finalise(false, true, 0.0, dangl);
    for (int it = 1; it <= iters; it++) {
        const double Dm = dangl;
        accumulate();
        finalise(true, it < iters, Dm, dangl);
    }

    // FP64 PageRank of the same graph.
    std::vector<double> x(V, 1.0 / V), y(V);
    for (int it = 1; it <= iters; it++) {
        double dm = 0.0;
        for (uint32_t v = 0; v < V; v++)
            if (!deg[v]) dm += x[v];
        for (uint32_t r = 0; r < V; r++) {
            double s = 0.0;
            for (uint32_t e = off[r]; e < off[r + 1]; e++) s += x[cols[e]] / deg[cols[e]];
            y[r] = 0.15 / V + 0.85 * (s + (uniform ? dm / V : 0.0));
        }
        x.swap(y);
    }
    // The sequential FP32 pull iteration.
    std::vector<float> xr(V, 1.0f / static_cast<float>(V)), yr(V), cr(V);
    for (int it = 1; it <= iters; it++) {
        double dm = 0.0;
        for (uint32_t v = 0; v < V; v++) {
            if (!deg[v]) dm += xr[v];
            cr[v] = deg[v] ? xr[v] / static_cast<float>(deg[v]) : 0.0f;
        }
        for (uint32_t r = 0; r < V; r++) {
            float s = 0.0f;
            for (uint32_t e = off[r]; e < off[r + 1]; e++) s += cr[cols[e]];
            yr[r] = uniform ? static_cast<float>(0.15 / V + 0.85 * (static_cast<double>(s) + dm / V))
                            : std::fmaf(kDamp, s, (1.0f - kDamp) / static_cast<float>(V));
        }
        xr.swap(yr);
    }
    double maxErr = 0.0, refErr = 0.0, maxRank = 0.0;
    for (uint32_t v = 0; v < V; v++) {
        maxErr = std::max(maxErr, std::fabs(static_cast<double>(rank[v]) - x[v]));
        refErr = std::max(refErr, std::fabs(static_cast<double>(xr[v]) - x[v]));
        maxRank = std::max(maxRank, x[v]);
    }
    std::printf("pairs %u slots %u blocks %zu max_abs_err %.3e ref_err %.3e max_rank %.3e\n", P.pairs, P.slots,
                P.blocks.size(), maxErr, refErr, maxRank);
    return 0;
}
