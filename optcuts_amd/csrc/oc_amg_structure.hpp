// Host-side structure of the smoothed-aggregation hierarchy: aggregates,
// prolongator pattern and the contribution maps of its numeric phase. Pure
// C++ (no device calls); oc_amg.hip uploads the result.
#pragma once

#include <algorithm>
#include <thread>
#include <numeric>
#include <string>
#include <utility>
#include <vector>

namespace oc {

// Row-block parallel loop for the symbolic set-up: fn(chunk, r0, r1) on
// nChunks contiguous row ranges, each on its own host thread. Callers write
// chunk-local outputs and concatenate them in chunk order, so the result does
// not depend on the thread count.
template <typename Fn>
inline void parallel_row_chunks(int n, int nChunks, Fn&& fn)
{
    std::vector<std::thread> th;
    for (int c = 0; c < nChunks; ++c) {
        const int r0 = (int)((long)n * c / nChunks), r1 = (int)((long)n * (c + 1) / nChunks);
        th.emplace_back([&fn, c, r0, r1] { fn(c, r0, r1); });
    }
    for (auto& t : th) t.join();
}

inline int setup_threads(long work)
{
    const unsigned hw = std::thread::hardware_concurrency();
    const int t = (int)std::min<unsigned>(hw ? hw : 1, 16);
    return work < 20000 ? 1 : t;
}

// Symbolic product rows: for each output row, the distinct columns (sorted)
// and the contributing (left, right) pairs grouped by column in generation
// order. gen(row, emit) calls emit(col, left, right) for every pair of the row
// in a fixed order; rows [r0, r1).
struct SymRows {
    std::vector<int> nCols;   // per row
    std::vector<int> cols;    // concatenated
    std::vector<int> grpEnd;  // per column entry: end of its pair group (row-local offsets)
    std::vector<int> a, b;    // pairs
};

template <typename Gen>
inline void symbolic_rows(int r0, int r1, int ncols, Gen&& gen, SymRows& out)
{
    struct Item {
        int c, x, y;
    };
    std::vector<Item> buf;
    std::vector<int> slot(ncols, -1), cols, cnt, fill;
    for (int r = r0; r < r1; ++r) {
        buf.clear();
        gen(r, [&](int c, int x, int y) { buf.push_back(Item{c, x, y}); });
        cols.clear();
        for (const Item& it : buf)
            if (slot[it.c] < 0) {
                slot[it.c] = 0;
                cols.push_back(it.c);
            }
        std::sort(cols.begin(), cols.end());
        const int k = (int)cols.size();
        cnt.assign(k + 1, 0);
        for (int q = 0; q < k; ++q) slot[cols[q]] = q;
        for (const Item& it : buf) cnt[slot[it.c] + 1]++;
        for (int q = 0; q < k; ++q) cnt[q + 1] += cnt[q];
        const size_t base = out.a.size();
        out.a.resize(base + buf.size());
        out.b.resize(base + buf.size());
        fill.assign(cnt.begin(), cnt.end() - 1);
        for (const Item& it : buf) {
            const int q = fill[slot[it.c]]++;
            out.a[base + q] = it.x;
            out.b[base + q] = it.y;
        }
        out.nCols.push_back(k);
        for (int q = 0; q < k; ++q) {
            out.cols.push_back(cols[q]);
            out.grpEnd.push_back(cnt[q + 1]);
            slot[cols[q]] = -1;
        }
    }
}

// Concatenates chunk results into CSR row pointers, columns, pair-group
// pointers and pairs.
inline void concat_rows(int n, std::vector<SymRows>& parts, std::vector<int>& rowPtr, std::vector<int>& col,
                        std::vector<int>& grpPtr, std::vector<int>& pa, std::vector<int>& pb)
{
    rowPtr.assign(n + 1, 0);
    col.clear();
    grpPtr.assign(1, 0);
    pa.clear();
    pb.clear();
    size_t nc = 0, np = 0;
    for (auto& P : parts) {
        nc += P.cols.size();
        np += P.a.size();
    }
    col.reserve(nc);
    grpPtr.reserve(nc + 1);
    pa.reserve(np);
    pb.reserve(np);
    int r = 0;
    for (auto& P : parts) {
        size_t ci = 0, pbase = 0;
        for (int k : P.nCols) {
            // pairs of this row start at pbase in P.a; group ends are row-local
            int rowPairs = 0;
            for (int q = 0; q < k; ++q) {
                col.push_back(P.cols[ci + q]);
                grpPtr.push_back((int)(pa.size() + P.grpEnd[ci + q]));
                rowPairs = P.grpEnd[ci + q];
            }
            pa.insert(pa.end(), P.a.begin() + pbase, P.a.begin() + pbase + rowPairs);
            pb.insert(pb.end(), P.b.begin() + pbase, P.b.begin() + pbase + rowPairs);
            pbase += rowPairs;
            ci += k;
            rowPtr[r + 1] = rowPtr[r] + k;
            ++r;
        }
    }
}

constexpr int kAmgNullDim = 3;  // two translations and a rotation of the UV plane

struct AmgHostLevel {
    int n = 0, bs = 0;                       // nodes and block size of A at this level
    std::vector<int> rowPtr, colIdx, diagPos;
    std::vector<unsigned char> excluded;     // nodes outside every aggregate (fixed vertices)
    // transfer to the next level
    int nc = 0;                              // aggregates (= nodes of the next level)
    std::vector<int> aggPtr, aggMem;         // members of each aggregate
    std::vector<int> pPtr, pCol, pRow;       // prolongator pattern (bs x 3 blocks)
    std::vector<unsigned char> pIsTent;      // block (u, agg(u)): carries the tentative block
    std::vector<int> psPtr, psA;             // per P block: A blocks of its row whose column is in its aggregate
    std::vector<int> rPtr, rP;               // per aggregate: P blocks of its column (restriction)
    std::vector<int> tPtr, tCol;             // T = A P pattern
    std::vector<int> tcPtr, tcA, tcP;        // per T block: (A block, P block) pairs
    std::vector<int> acPtr, acP, acT;        // per next-level block: (P block, T block) pairs
};

// Greedy aggregation (phase 1: roots whose neighbours are all free take
// them; phase 2: leftovers join a neighbouring aggregate; phase 3: isolated
// nodes become singletons). Excluded nodes get -1.
inline int aggregate_nodes(int n, const std::vector<int>& rowPtr, const std::vector<int>& colIdx,
                           const std::vector<unsigned char>& excluded, std::vector<int>& agg)
{
    agg.assign(n, -1);
    int na = 0;
    for (int i = 0; i < n; ++i) {
        if (excluded[i] || agg[i] >= 0) continue;
        bool free = true;
        for (int k = rowPtr[i]; k < rowPtr[i + 1] && free; ++k) {
            const int w = colIdx[k];
            if (w != i && !excluded[w] && agg[w] >= 0) free = false;
        }
        if (!free) continue;
        agg[i] = na;
        for (int k = rowPtr[i]; k < rowPtr[i + 1]; ++k) {
            const int w = colIdx[k];
            if (!excluded[w]) agg[w] = na;
        }
        ++na;
    }
    std::vector<int> agg1 = agg;
    for (int i = 0; i < n; ++i) {
        if (excluded[i] || agg1[i] >= 0) continue;
        for (int k = rowPtr[i]; k < rowPtr[i + 1]; ++k) {
            const int w = colIdx[k];
            if (w != i && !excluded[w] && agg1[w] >= 0) {
                agg[i] = agg1[w];
                break;
            }
        }
    }
    for (int i = 0; i < n; ++i)
        if (!excluded[i] && agg[i] < 0) agg[i] = na++;
    return na;
}


// Builds every level down to at most coarseMaxDofs unknowns (or maxLevels
// levels, or until coarsening stalls). Returns 0, or -1 with a message.
inline int amg_build_structure(int nV, const std::vector<int>& rowPtr, const std::vector<int>& colIdx,
                               const std::vector<unsigned char>& fixedMask, long coarseMaxDofs, int maxLevels,
                               std::vector<AmgHostLevel>& L, std::string& err)
{
    L.clear();
    L.emplace_back();
    L[0].n = nV;
    L[0].bs = 2;
    L[0].rowPtr = rowPtr;
    L[0].colIdx = colIdx;
    L[0].excluded = fixedMask;
    for (;;) {
        AmgHostLevel& A = L.back();
        A.diagPos.assign(A.n, -1);
        for (int u = 0; u < A.n; ++u)
            for (int k = A.rowPtr[u]; k < A.rowPtr[u + 1]; ++k)
                if (A.colIdx[k] == u) A.diagPos[u] = k;
        for (int u = 0; u < A.n; ++u)
            if (A.diagPos[u] < 0) {
                err = "AMG: a row has no diagonal block";
                return -1;
            }
        if ((long)A.n * A.bs <= coarseMaxDofs || (int)L.size() >= maxLevels) break;
        // Isolated nodes (only a diagonal block: a disconnected component
        // already reduced to one node, e.g. one of many independent charts)
        // stay singleton aggregates; once every node is isolated the level is
        // block diagonal and its diagonal-block inverse is the exact coarse
        // solve.
        int nIso = 0, nIn = 0;
        for (int u = 0; u < A.n; ++u) {
            if (A.excluded[u]) continue;
            ++nIn;
            nIso += A.rowPtr[u + 1] - A.rowPtr[u] == 1;
        }
        if (nIso == nIn) break;
        std::vector<int> agg;
        const int nc = aggregate_nodes(A.n, A.rowPtr, A.colIdx, A.excluded, agg);
        if (nc == 0 || (long)(nc - nIso) * 10 > (long)(nIn - nIso) * 9) break;  // coupled part stopped coarsening
        A.nc = nc;
        // aggregate members
        A.aggPtr.assign(nc + 1, 0);
        for (int u = 0; u < A.n; ++u)
            if (agg[u] >= 0) A.aggPtr[agg[u] + 1]++;
        for (int c = 0; c < nc; ++c) A.aggPtr[c + 1] += A.aggPtr[c];
        A.aggMem.assign(A.aggPtr[nc], 0);
        {
            std::vector<int> f(A.aggPtr.begin(), A.aggPtr.end() - 1);
            for (int u = 0; u < A.n; ++u)
                if (agg[u] >= 0) A.aggMem[f[agg[u]]++] = u;
        }
        // P pattern: row u -> aggregates of its (non-excluded) closed
        // neighbourhood; with it the smoothing contributions (the A blocks of
        // row u grouped by the aggregate of their column, k ascending).
        {
            const int nt = setup_threads((long)A.colIdx.size());
            std::vector<SymRows> parts(nt);
            parallel_row_chunks(A.n, nt, [&](int ch, int r0, int r1) {
                symbolic_rows(r0, r1, nc,
                              [&](int u, auto&& emit) {
                                  if (A.excluded[u]) return;
                                  for (int k = A.rowPtr[u]; k < A.rowPtr[u + 1]; ++k) {
                                      const int w = A.colIdx[k];
                                      if (!A.excluded[w]) emit(agg[w], k, 0);
                                  }
                              },
                              parts[ch]);
            });
            std::vector<int> unused;
            concat_rows(A.n, parts, A.pPtr, A.pCol, A.psPtr, A.psA, unused);
        }
        const int nP = (int)A.pCol.size();
        A.pRow.assign(nP, 0);
        A.pIsTent.assign(nP, 0);
        for (int u = 0; u < A.n; ++u)
            for (int j = A.pPtr[u]; j < A.pPtr[u + 1]; ++j) {
                A.pRow[j] = u;
                A.pIsTent[j] = A.pCol[j] == agg[u] ? 1 : 0;
            }
        // restriction: P blocks by column
        A.rPtr.assign(nc + 1, 0);
        for (int j = 0; j < nP; ++j) A.rPtr[A.pCol[j] + 1]++;
        for (int c = 0; c < nc; ++c) A.rPtr[c + 1] += A.rPtr[c];
        A.rP.assign(nP, 0);
        {
            std::vector<int> f(A.rPtr.begin(), A.rPtr.end() - 1);
            for (int j = 0; j < nP; ++j) A.rP[f[A.pCol[j]]++] = j;
        }
        // T = A P: pattern and (A block, P block) pairs, per row grouped by
        // coarse column in generation order (k ascending, then j).
        {
            const int nt = setup_threads((long)A.pCol.size() * 8);
            std::vector<SymRows> parts(nt);
            parallel_row_chunks(A.n, nt, [&](int ch, int r0, int r1) {
                symbolic_rows(r0, r1, nc,
                              [&](int u, auto&& emit) {
                                  for (int k = A.rowPtr[u]; k < A.rowPtr[u + 1]; ++k) {
                                      const int w = A.colIdx[k];
                                      for (int j = A.pPtr[w]; j < A.pPtr[w + 1]; ++j) emit(A.pCol[j], k, j);
                                  }
                              },
                              parts[ch]);
            });
            concat_rows(A.n, parts, A.tPtr, A.tCol, A.tcPtr, A.tcA, A.tcP);
        }
        // coarse A = P^T T: pattern and (P block, T block) pairs, rows by coarse
        // node: over the P blocks (u, c1) in increasing block order (the
        // restriction lists), the T entries of row u.
        AmgHostLevel nxl;
        AmgHostLevel* nx = &nxl;
        nx->n = nc;
        nx->bs = kAmgNullDim;
        nx->excluded.assign(nc, 0);
        {
            const int nt = setup_threads((long)A.tCol.size() * 4);
            std::vector<SymRows> parts(nt);
            parallel_row_chunks(nc, nt, [&](int ch, int r0, int r1) {
                symbolic_rows(r0, r1, nc,
                              [&](int c1, auto&& emit) {
                                  for (int q = A.rPtr[c1]; q < A.rPtr[c1 + 1]; ++q) {
                                      const int j = A.rP[q];
                                      const int u = A.pRow[j];
                                      for (int t = A.tPtr[u]; t < A.tPtr[u + 1]; ++t) emit(A.tCol[t], j, t);
                                  }
                              },
                              parts[ch]);
            });
            concat_rows(nc, parts, nx->rowPtr, nx->colIdx, A.acPtr, A.acP, A.acT);
        }
        L.push_back(std::move(nxl));
    }
    return 0;
}

}  // namespace oc
