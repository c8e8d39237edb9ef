// Host-side plan of the binned PageRank iteration (propagation blocking:
// S. Beamer, K. Asanovic, D. Patterson, "Reducing PageRank Communication via
// Propagation Blocking", IPDPS 2017), restated for LDS accumulation on
// gfx950.  Plain C++, no HIP types, so that a CPU program can build and
// check a plan.
//
// The vertices (rows) are cut into partitions of 2^rows_log2 rows.  A pair
// is a distinct (source s, partition k) such that s has at least one
// out-edge into k; each pair holds one value per iteration, s's
// contribution rank[s]/outdeg(s), in a "bin" array ordered by partition,
// then by source.  One iteration is:
//   1. accumulate: per partition, its in-edges ordered by (source, row)
//      stream their 16-bit local row ids; each edge adds the value of its
//      pair to the row's LDS accumulator.  Long partitions are cut into
//      chunks, one workgroup each, which write one partial vector per chunk.
//   2. finalise + scatter: per block of consecutive vertices (inside one
//      partition), the partials of the block's rows are summed in chunk
//      order and finalised into new ranks; the new contributions of the
//      block's vertices are staged in LDS and written to every bin slot of
//      the block's pairs -- a set of runs, one per partition the block's
//      vertices point into, each contiguous in the bin array.  The pairs'
//      sources are stored in this scatter order ("flat" order), so the
//      scatter reads them sequentially.
// Pairs deduplicate the edges a source has into one partition, so the bin
// traffic is 4 bytes per pair written and read instead of one gathered
// cache line per few edges.
#pragma once

#include <algorithm>
#include <cstdint>
#include <string>
#include <vector>

namespace bw {
namespace bins {

// Scatter block limits (measured on R-MAT scale 22: 2048 / 32768 beat
// 512 / 8192, 1024 / 16384, 4096 / 65536 and 8192 / 131072).
#ifndef BW_BINS_BLOCK_VERTS
#define BW_BINS_BLOCK_VERTS 2048
#endif
#ifndef BW_BINS_BLOCK_WORK
#define BW_BINS_BLOCK_WORK 32768
#endif
constexpr uint32_t kTaskEdges = 2048;    // unit of the chunk size (chunk_tasks x kTaskEdges edges at most)
constexpr uint32_t kChunkWaves = 16;     // tasks per chunk: one per wave of the accumulate workgroup
constexpr uint32_t kMaxRuns = 2048;      // partitions one scatter block may point into
constexpr uint32_t kBlockVerts = BW_BINS_BLOCK_VERTS;  // vertices of one scatter block at most
constexpr uint32_t kBlockWork = BW_BINS_BLOCK_WORK;    // vertices x chunks of their partition + pairs, per block
constexpr uint16_t kNewPair = 0x8000;    // edge code flag: the edge opens a new pair

struct Quad {
    uint32_t x, y, z, w;
};

struct Plan {
    uint32_t V = 0, E = 0, rows_log2 = 0, K = 0;
    uint32_t pairs = 0;                  // U, the number of (source, partition) pairs
    uint32_t slots = 0;                  // partial vectors (chunks)
    std::vector<uint16_t> codes;         // per edge, partition-major, each partition padded to 64:
                                         // (row - partition base) | kNewPair on a pair's first edge
    std::vector<uint32_t> edge_start;    // K + 1: padded start of each partition's edges
    std::vector<uint32_t> pair_start;    // K + 1: first pair of each partition
    std::vector<Quad> tasks;             // (edge begin, edge end, pair base, partition), 8-aligned begin:
                                         // pair(e) = base + #flags in [begin, e]
    std::vector<Quad> chunks;            // (task begin, task end, partition, slot), longest first;
                                         // a partition's edges split evenly into chunks, a chunk's
                                         // evenly into kChunkWaves tasks
    std::vector<uint32_t> chunk_base;    // K + 1: slots of partition k are [chunk_base[k], chunk_base[k+1])
    std::vector<Quad> blocks;            // (first vertex, end vertex, run begin, run count)
    std::vector<uint32_t> run_flat;      // per block: flat index of each run's first pair, then the block's end
    std::vector<uint32_t> run_bin;       // per block: bin index of each run's first pair, then 0
    std::vector<uint16_t> psrc;          // per pair in flat order (blocks in vertex order, each block's
                                         // runs in partition order): source - first vertex of its block
};

// Builds the plan from a reverse CSR (row r's in-edge sources are
// cols[off[r] .. off[r+1])).  Returns false with `err` set on bad input or
// on a graph the binned layout cannot hold.
inline bool build_plan(const uint32_t *off, const uint32_t *cols, uint32_t V, uint32_t E, uint32_t rows_log2,
                       uint32_t chunk_tasks, Plan &P, std::string &err)
{
    if (rows_log2 < 8 || rows_log2 > 15) {
        err = "rows_log2 must be in [8, 15]";
        return false;
    }
    if (chunk_tasks == 0) {
        err = "chunk_tasks must be positive";
        return false;
    }
    if (V == 0 || off[V] != E) {
        err = "reverse offsets do not end at the edge count";
        return false;
    }
    const uint32_t L = rows_log2, D = 1u << L;
    const uint32_t K = static_cast<uint32_t>((static_cast<uint64_t>(V) + D - 1) >> L);
    if (K > kMaxRuns) {
        err = "too many partitions for the binned layout: raise rows_log2";
        return false;
    }
    P = Plan();
    P.V = V;
    P.E = E;
    P.rows_log2 = L;
    P.K = K;

    // Forward lists: the rows of each source's out-edges, in increasing
    // row order (so their partitions come in increasing order too).
    std::vector<uint32_t> foff(static_cast<size_t>(V) + 1, 0);
    for (uint32_t e = 0; e < E; e++) {
        if (cols[e] >= V) {
            err = "column id out of range";
            return false;
        }
        foff[cols[e] + 1]++;
    }
    for (uint32_t v = 0; v < V; v++) foff[v + 1] += foff[v];
    std::vector<uint32_t> frow(std::max<uint32_t>(E, 1));
    {
        std::vector<uint32_t> cur(foff.begin(), foff.end() - 1);
        for (uint32_t r = 0; r < V; r++)
            for (uint32_t e = off[r]; e < off[r + 1]; e++) frow[cur[cols[e]]++] = r;
    }

    // Pairs per partition and per source.
    std::vector<uint64_t> pairsK(K, 0);
    std::vector<uint32_t> pairsS(V, 0);
    for (uint32_t s = 0; s < V; s++) {
        uint32_t prev = UINT32_MAX;
        for (uint32_t e = foff[s]; e < foff[s + 1]; e++) {
            const uint32_t k = frow[e] >> L;
            if (k != prev) {
                pairsK[k]++;
                pairsS[s]++;
                prev = k;
            }
        }
    }
    P.pair_start.assign(K + 1, 0);
    P.edge_start.assign(K + 1, 0);
    uint64_t u = 0, ep = 0;
    std::vector<uint32_t> edgesK(K);
    for (uint32_t k = 0; k < K; k++) {
        P.pair_start[k] = static_cast<uint32_t>(u);
        P.edge_start[k] = static_cast<uint32_t>(ep);
        const uint32_t r0 = k << L, r1 = std::min<uint64_t>(static_cast<uint64_t>(r0) + D, V);
        edgesK[k] = off[r1] - off[r0];
        u += pairsK[k];
        ep += (static_cast<uint64_t>(edgesK[k]) + 63) & ~63ull;  // 64 codes = one 128-byte line
        // pairs and codes are addressed with 32-bit byte offsets
        if (u >= (1ull << 30) || ep >= (1ull << 31) - 8) {
            err = "graph too large for the binned layout";
            return false;
        }
    }
    P.pair_start[K] = static_cast<uint32_t>(u);
    P.edge_start[K] = static_cast<uint32_t>(ep);
    P.pairs = static_cast<uint32_t>(u);

    // Chunks per partition: ceil(edges / (chunk_tasks x kTaskEdges)).
    const uint64_t chunkEdges = static_cast<uint64_t>(chunk_tasks) * kTaskEdges;
    P.chunk_base.assign(K + 1, 0);
    for (uint32_t k = 0; k < K; k++)
        P.chunk_base[k + 1] = P.chunk_base[k] + static_cast<uint32_t>((edgesK[k] + chunkEdges - 1) / chunkEdges);
    P.slots = P.chunk_base[K];

    // Scatter blocks: consecutive vertices of one partition, cut so that
    // vertices x chunks of the partition + pairs stays within kBlockWork,
    // at most kBlockVerts vertices and kMaxRuns runs.
    std::vector<uint32_t> stamp(K, UINT32_MAX), blen(K, 0), done(K, 0), touched;
    std::vector<uint32_t> blockOf(V), runPart;
    uint32_t flat = 0;
    for (uint32_t k = 0; k < K; k++) {
        const uint32_t vEnd = std::min<uint64_t>((static_cast<uint64_t>(k) << L) + D, V);
        const uint64_t perVertex = std::max<uint32_t>(1, P.chunk_base[k + 1] - P.chunk_base[k]);
        uint32_t s = k << L;
        while (s < vEnd) {
            const uint32_t id = static_cast<uint32_t>(P.blocks.size()), v0 = s;
            uint64_t work = 0;
            touched.clear();
            while (s < vEnd && s - v0 < kBlockVerts) {
                uint32_t fresh = 0, prev = UINT32_MAX;
                for (uint32_t e = foff[s]; e < foff[s + 1]; e++) {
                    const uint32_t kk = frow[e] >> L;
                    if (kk != prev && stamp[kk] != id) fresh++;
                    prev = kk;
                }
                const uint64_t w = perVertex + pairsS[s];
                if (s > v0 && (work + w > kBlockWork || touched.size() + fresh > kMaxRuns)) break;
                prev = UINT32_MAX;
                for (uint32_t e = foff[s]; e < foff[s + 1]; e++) {
                    const uint32_t kk = frow[e] >> L;
                    if (kk == prev) continue;
                    prev = kk;
                    if (stamp[kk] != id) {
                        stamp[kk] = id;
                        blen[kk] = 0;
                        touched.push_back(kk);
                    }
                    blen[kk]++;
                }
                blockOf[s] = id;
                work += w;
                s++;
            }
            std::sort(touched.begin(), touched.end());
            const uint32_t runBegin = static_cast<uint32_t>(P.run_flat.size());
            for (uint32_t kk : touched) {
                P.run_flat.push_back(flat);
                P.run_bin.push_back(P.pair_start[kk] + done[kk]);
                runPart.push_back(kk);
                flat += blen[kk];
                done[kk] += blen[kk];
            }
            P.run_flat.push_back(flat);
            P.run_bin.push_back(0);
            runPart.push_back(UINT32_MAX);
            P.blocks.push_back({v0, s, runBegin, static_cast<uint32_t>(touched.size())});
        }
    }

    // Edge codes in (partition, source, row) order; pair sources in flat
    // order (cursor[k]: next flat slot of the current block's run into k).
    if (flat != P.pairs) {
        err = "internal error: pair count";
        return false;
    }
    P.codes.assign(std::max<uint64_t>(ep, 8), 0);
    P.psrc.assign(std::max<uint32_t>(P.pairs, 1), 0);
    {
        std::vector<uint32_t> ec(K, 0), cursor(K, 0);
        uint32_t cur = UINT32_MAX;
        for (uint32_t s = 0; s < V; s++) {
            const Quad &b = P.blocks[blockOf[s]];
            if (blockOf[s] != cur) {
                cur = blockOf[s];
                for (uint32_t i = b.z; i < b.z + b.w; i++) cursor[runPart[i]] = P.run_flat[i];
            }
            uint32_t prev = UINT32_MAX;
            for (uint32_t e = foff[s]; e < foff[s + 1]; e++) {
                const uint32_t r = frow[e], k = r >> L;
                const bool fresh = k != prev;
                prev = k;
                if (fresh) P.psrc[cursor[k]++] = static_cast<uint16_t>(s - b.x);
                P.codes[P.edge_start[k] + ec[k]++] = static_cast<uint16_t>((r & (D - 1)) | (fresh ? kNewPair : 0));
            }
        }
    }

    // Chunks: each partition's edges split evenly (boundaries rounded down
    // to 64 edges, so a wave's 512 codes are 8 whole lines), each chunk's
    // evenly into kChunkWaves tasks (empty ones dropped).
    for (uint32_t k = 0; k < K; k++) {
        const uint64_t n = edgesK[k], nch = P.chunk_base[k + 1] - P.chunk_base[k];
        const uint32_t es = P.edge_start[k];
        uint32_t flags = 0, counted = 0;  // new-pair flags in [es, es + counted)
        auto flags_before = [&](uint32_t off) {
            for (; counted < off; counted++) flags += (P.codes[es + counted] & kNewPair) ? 1u : 0u;
            return flags;
        };
        auto round64 = [](uint64_t x) { return static_cast<uint32_t>(x & ~63ull); };
        for (uint64_t c = 0; c < nch; c++) {
            const uint32_t cb = round64(n * c / nch), ce = c + 1 == nch ? static_cast<uint32_t>(n) : round64(n * (c + 1) / nch);
            const uint32_t t0 = static_cast<uint32_t>(P.tasks.size());
            for (uint32_t w = 0; w < kChunkWaves; w++) {
                const uint32_t tb = cb + round64(static_cast<uint64_t>(ce - cb) * w / kChunkWaves);
                const uint32_t te = w + 1 == kChunkWaves ? ce : cb + round64(static_cast<uint64_t>(ce - cb) * (w + 1) / kChunkWaves);
                if (te > tb) P.tasks.push_back({es + tb, es + te, P.pair_start[k] - 1u + flags_before(tb), k});
            }
            P.chunks.push_back({t0, static_cast<uint32_t>(P.tasks.size()), k, P.chunk_base[k] + static_cast<uint32_t>(c)});
        }
    }
    std::stable_sort(P.chunks.begin(), P.chunks.end(),
                     [](const Quad &a, const Quad &b) { return a.y - a.x > b.y - b.x; });
    return true;
}

}  // namespace bins
}  // namespace bw
