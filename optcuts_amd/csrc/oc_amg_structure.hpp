// Host-side structure of the smoothed-aggregation hierarchy: aggregates,
// prolongator pattern and the contribution maps of its numeric phase. Pure
// C++ (no device calls); oc_amg.hip uploads the result.
#pragma once

#include <algorithm>
#include <condition_variable>
#include <functional>
#include <mutex>
#include <atomic>
#include <thread>
#include <numeric>
#include <string>
#include <utility>
#include <vector>

namespace oc {

// Persistent host workers for the row-chunk loops of the symbolic set-up
// (pattern, slot maps, multigrid structure): a set-up runs a dozen parallel
// regions, and starting threads per region would cost more than some of them.
// One region at a time; a caller that finds the pool busy (another context's
// set-up on another host thread) starts its own threads instead.
class RowPool {
public:
    static RowPool& instance()
    {
        static RowPool pool;
        return pool;
    }
    int workers() const { return (int)workers_.size(); }
    // Runs task(c) for c in [0, n) on the workers and the calling thread;
    // false (nothing run) when the pool is busy.
    bool run(int n, const std::function<void(int)>& task)
    {
        std::unique_lock<std::mutex> busy(busy_, std::try_to_lock);
        if (!busy.owns_lock()) return false;
        {
            std::lock_guard<std::mutex> g(m_);
            task_ = &task;
            n_ = n;
            next_ = 0;
            pending_ = n;
            ++gen_;
        }
        cv_.notify_all();
        work(&task, n);
        std::unique_lock<std::mutex> g(m_);
        done_.wait(g, [&] { return pending_ == 0 && active_ == 0; });
        task_ = nullptr;
        return true;
    }

private:
    RowPool()
    {
        const unsigned hw = std::thread::hardware_concurrency();
        const int nw = (int)std::min<unsigned>(hw ? hw : 1, 16) - 1;
        for (int i = 0; i < nw; ++i) workers_.emplace_back([this] { loop(); });
    }
    ~RowPool()
    {
        {
            std::lock_guard<std::mutex> g(m_);
            stop_ = true;
        }
        cv_.notify_all();
        for (auto& t : workers_) t.join();
    }
    // Claims chunks until none is left. The region's task and size are read
    // under the lock by the worker before it enters, and run() returns only
    // after every worker that entered has left, so a late worker never sees
    // the next region's counter with this region's task.
    void work(const std::function<void(int)>* task, int n)
    {
        for (;;) {
            int c;
            {
                std::lock_guard<std::mutex> g(m_);
                c = next_ < n ? next_++ : -1;
            }
            if (c < 0) return;
            (*task)(c);
            std::lock_guard<std::mutex> g(m_);
            if (--pending_ == 0) done_.notify_all();
        }
    }
    void loop()
    {
        unsigned long seen = 0;
        for (;;) {
            const std::function<void(int)>* task;
            int n;
            {
                std::unique_lock<std::mutex> g(m_);
                cv_.wait(g, [&] { return stop_ || (gen_ != seen && task_); });
                if (stop_) return;
                seen = gen_;
                task = task_;
                n = n_;
                ++active_;
            }
            work(task, n);
            std::lock_guard<std::mutex> g(m_);
            if (--active_ == 0) done_.notify_all();
        }
    }

    std::mutex busy_, m_;
    std::condition_variable cv_, done_;
    std::vector<std::thread> workers_;
    const std::function<void(int)>* task_ = nullptr;
    int n_ = 0, next_ = 0, pending_ = 0, active_ = 0;
    unsigned long gen_ = 0;
    bool stop_ = false;
};

// Row-block parallel loop for the symbolic set-up: fn(chunk, r0, r1) on
// nChunks contiguous row ranges, in parallel. Callers write chunk-local
// outputs or disjoint row ranges, so the result does not depend on the number
// or order of threads.
template <typename Fn>
inline void parallel_row_chunks(int n, int nChunks, Fn&& fn)
{
    const auto range = [&](int c) {
        const int r0 = (int)((long)n * c / nChunks), r1 = (int)((long)n * (c + 1) / nChunks);
        fn(c, r0, r1);
    };
    if (nChunks <= 1) {
        if (nChunks == 1) range(0);
        return;
    }
    const std::function<void(int)> task(range);
    if (RowPool::instance().run(nChunks, task)) return;
    // pool busy (another context is building): own threads, at most 16, claiming chunks in turn
    const int nT = std::min(nChunks, 16);
    std::atomic<int> next{0};
    std::vector<std::thread> th;
    for (int t = 0; t < nT; ++t)
        th.emplace_back([&task, &next, nChunks] {
            for (int c = next++; c < nChunks; c = next++) task(c);
        });
    for (auto& t : th) t.join();
}

inline int setup_threads(long work)
{
    const unsigned hw = std::thread::hardware_concurrency();
    const int t = (int)std::min<unsigned>(hw ? hw : 1, 16);
    return work < 20000 ? 1 : t;
}

// Row chunks for a parallel region: 8 per host thread, claimed dynamically, so that rows of uneven cost (the
// rows an air mesh touches are numbered, and aggregated, last) spread over the threads.
inline int setup_chunks(long work)
{
    const int t = setup_threads(work);
    return t > 1 ? 8 * t : 1;
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

template <typename Gen, typename Skip>
inline void symbolic_rows(int r0, int r1, int ncols, Gen&& gen, Skip&& skip, SymRows& out)
{
    struct Item {
        int c, x, y;
    };
    std::vector<Item> buf;
    std::vector<int> slot(ncols, -1), cols, cnt, fill;
    for (int r = r0; r < r1; ++r) {
        if (skip(r)) continue;
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

// A symbolic product of an earlier set-up (CSR rows, pair-group ends, pairs),
// whose rows are reused where they come out the same.
struct ProductRef {
    const std::vector<int>* rowPtr = nullptr;
    const std::vector<int>* col = nullptr;
    const std::vector<int>* grp = nullptr;
    const std::vector<int>* pa = nullptr;
    const std::vector<int>* pb = nullptr;
};

// Rows [0, n) of a symbolic product: CSR row pointers and columns, per column
// entry the end of its pair group (grpPtr[0] = 0) and the pairs pa/pb (pb may
// be null). Row r comes from gen unless keep(r): then it is row r of the
// earlier product `old` with its pairs renumbered by mapA(r, a) and
// mapB(r, a, b) (the blocks they name have moved). Generated rows run in row
// chunks on host threads and every row is written at its final offset, so
// the result does not depend on the thread count.
template <typename Gen, typename Keep, typename MapA, typename MapB>
inline void product_rows(int n, int ncols, int nt, Gen&& gen, Keep&& keep, const ProductRef& old, MapA&& mapA,
                         MapB&& mapB, std::vector<int>& rowPtr, std::vector<int>& col, std::vector<int>& grpPtr,
                         std::vector<int>& pa, std::vector<int>* pb)
{
    std::vector<SymRows> parts(nt);
    std::vector<int> nCol(n), nPair(n);
    parallel_row_chunks(n, nt, [&](int ch, int r0, int r1) {
        SymRows& S = parts[ch];
        symbolic_rows(r0, r1, ncols, gen, keep, S);
        size_t ci = 0, ri = 0;
        for (int r = r0; r < r1; ++r) {
            if (keep(r)) {
                const int e0 = (*old.rowPtr)[r], e1 = (*old.rowPtr)[r + 1];
                nCol[r] = e1 - e0;
                nPair[r] = (*old.grp)[e1] - (*old.grp)[e0];
            } else {
                const int k = S.nCols[ri++];
                nCol[r] = k;
                nPair[r] = k ? S.grpEnd[ci + k - 1] : 0;
                ci += k;
            }
        }
    });
    std::vector<int> pairStart(n + 1, 0);
    rowPtr.assign(n + 1, 0);
    for (int r = 0; r < n; ++r) {
        rowPtr[r + 1] = rowPtr[r] + nCol[r];
        pairStart[r + 1] = pairStart[r] + nPair[r];
    }
    col.resize(rowPtr[n]);
    grpPtr.resize((size_t)rowPtr[n] + 1);
    grpPtr[0] = 0;
    pa.resize(pairStart[n]);
    if (pb) pb->resize(pairStart[n]);
    parallel_row_chunks(n, nt, [&](int ch, int r0, int r1) {
        const SymRows& S = parts[ch];
        size_t ci = 0, ri = 0, pbase = 0;
        for (int r = r0; r < r1; ++r) {
            const int e = rowPtr[r], p = pairStart[r];
            if (keep(r)) {
                const int e0 = (*old.rowPtr)[r], p0 = (*old.grp)[e0];
                for (int q = 0; q < nCol[r]; ++q) {
                    col[e + q] = (*old.col)[e0 + q];
                    grpPtr[e + q + 1] = (*old.grp)[e0 + q + 1] - p0 + p;
                }
                for (int q = 0; q < nPair[r]; ++q) {
                    const int a = (*old.pa)[p0 + q];
                    pa[p + q] = mapA(r, a);
                    if (pb) (*pb)[p + q] = mapB(r, a, (*old.pb)[p0 + q]);
                }
            } else {
                const int k = S.nCols[ri++];
                for (int q = 0; q < k; ++q) {
                    col[e + q] = S.cols[ci + q];
                    grpPtr[e + q + 1] = p + S.grpEnd[ci + q];
                }
                std::copy(S.a.begin() + pbase, S.a.begin() + pbase + nPair[r], pa.begin() + p);
                if (pb) std::copy(S.b.begin() + pbase, S.b.begin() + pbase + nPair[r], pb->begin() + p);
                ci += k;
                pbase += nPair[r];
            }
        }
    });
}

constexpr int kAmgNullDim = 3;  // two translations and a rotation of the UV plane

struct AmgHostLevel {
    int n = 0, bs = 0;                       // nodes and block size of A at this level
    std::vector<int> rowPtr, colIdx, diagPos;
    std::vector<unsigned char> excluded;     // nodes outside every aggregate (fixed vertices)
    std::vector<unsigned char> late;         // nodes that become aggregate roots only after all others (may be empty)
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
    std::vector<int> agg;                    // aggregate of each node (-1: excluded)
};

// Greedy aggregation (phase 1: roots whose neighbours are all free take
// them; phase 2: leftovers join a neighbouring aggregate; phase 3: isolated
// nodes become singletons). Excluded nodes get -1. Phase 1 visits nodes in
// order, the `late` ones (if given) after all others: when only the rows of
// late nodes change between two builds (the air mesh's rows of a scaffolded
// system), every root among the other nodes takes the same neighbours and its
// aggregate the same number, so the later levels' rows away from the change
// come out the same and are copied.
inline int aggregate_nodes(int n, const std::vector<int>& rowPtr, const std::vector<int>& colIdx,
                           const std::vector<unsigned char>& excluded, std::vector<int>& agg,
                           const std::vector<unsigned char>* late = nullptr)
{
    agg.assign(n, -1);
    int na = 0;
    const bool twoPass = late && !late->empty();
    for (int pass = 0; pass < (twoPass ? 2 : 1); ++pass)
        for (int i = 0; i < n; ++i) {
            if (excluded[i] || agg[i] >= 0) continue;
            if (twoPass && (*late)[i] != pass) continue;
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


// The levels of an earlier structure build: rows of a new level that come out
// the same as that build's level of the same depth (same matrix row, same
// aggregates around it) are copied from it instead of regenerated; the result
// is identical to a build without it.
struct AmgPrevious {
    std::vector<const AmgHostLevel*> levels;
};

// Builds every level down to at most coarseMaxDofs unknowns (or maxLevels
// levels, or until coarsening stalls). Returns 0, or -1 with a message.
inline int amg_build_structure(int nV, const std::vector<int>& rowPtr, const std::vector<int>& colIdx,
                               const std::vector<unsigned char>& fixedMask, long coarseMaxDofs, int maxLevels,
                               std::vector<AmgHostLevel>& L, std::string& err, const AmgPrevious* prev = nullptr,
                               const std::vector<unsigned char>* late0 = nullptr)
{
    // Levels reuse the storage of L's elements on entry (a hierarchy two
    // set-ups old); every array is rewritten, and nc marks the levels that
    // get a transfer.
    std::vector<AmgHostLevel> store = std::move(L);
    L.clear();
    const auto take = [&](size_t l) {
        AmgHostLevel x = l < store.size() ? std::move(store[l]) : AmgHostLevel();
        x.nc = 0;
        return x;
    };
    L.push_back(take(0));
    L[0].n = nV;
    L[0].bs = 2;
    L[0].rowPtr = rowPtr;
    L[0].colIdx = colIdx;
    L[0].excluded = fixedMask;
    if (late0) L[0].late = *late0;
    else L[0].late.clear();
    for (;;) {
        AmgHostLevel& A = L.back();
        A.diagPos.assign(A.n, -1);
        parallel_row_chunks(A.n, setup_chunks((long)A.colIdx.size()), [&](int, int r0, int r1) {
            for (int u = r0; u < r1; ++u)
                for (int k = A.rowPtr[u]; k < A.rowPtr[u + 1]; ++k)
                    if (A.colIdx[k] == u) A.diagPos[u] = k;
        });
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
        std::vector<int>& agg = A.agg;
        const int nc = aggregate_nodes(A.n, A.rowPtr, A.colIdx, A.excluded, agg, &A.late);
        if (nc == 0 || (long)(nc - nIso) * 10 > (long)(nIn - nIso) * 9) break;  // coupled part stopped coarsening
        A.nc = nc;
        // Rows of the earlier build's level that come out the same: matrix row
        // (cleanA), prolongator row (cleanP: same aggregates in its closed
        // neighbourhood), A P row (cleanT: cleanA and every neighbour cleanP);
        // dA/dP/dT: how far each such row moved in its CSR arrays.
        const size_t lv = L.size() - 1;
        const AmgHostLevel* O = (prev && lv + 1 < prev->levels.size() && prev->levels[lv]->nc > 0)
                                    ? prev->levels[lv]
                                    : nullptr;
        const AmgHostLevel* Onext = O ? prev->levels[lv + 1] : nullptr;
        std::vector<unsigned char> cleanA, cleanP, cleanT;
        std::vector<int> dA, dP, dT;
        const int nt0 = setup_chunks((long)A.colIdx.size());
        if (O) {
            cleanA.assign(A.n, 0);
            cleanP.assign(A.n, 0);
            cleanT.assign(A.n, 0);
            dA.assign(A.n, 0);
            dP.assign(A.n, 0);
            dT.assign(A.n, 0);
            parallel_row_chunks(A.n, nt0, [&](int, int r0, int r1) {
                for (int u = r0; u < r1; ++u) {
                    if (u >= O->n || A.excluded[u] != O->excluded[u]) continue;
                    const int k0 = A.rowPtr[u], k1 = A.rowPtr[u + 1], o0 = O->rowPtr[u];
                    if (k1 - k0 != O->rowPtr[u + 1] - o0 ||
                        !std::equal(A.colIdx.begin() + k0, A.colIdx.begin() + k1, O->colIdx.begin() + o0))
                        continue;
                    cleanA[u] = 1;
                    dA[u] = k0 - o0;
                }
            });
            parallel_row_chunks(A.n, nt0, [&](int, int r0, int r1) {
                for (int u = r0; u < r1; ++u) {
                    if (!cleanA[u]) continue;
                    bool same = true;
                    for (int k = A.rowPtr[u]; k < A.rowPtr[u + 1] && same; ++k) same = agg[A.colIdx[k]] == O->agg[A.colIdx[k]];
                    cleanP[u] = same;
                }
            });
            parallel_row_chunks(A.n, nt0, [&](int, int r0, int r1) {
                for (int u = r0; u < r1; ++u) {
                    if (!cleanA[u]) continue;
                    bool same = true;
                    for (int k = A.rowPtr[u]; k < A.rowPtr[u + 1] && same; ++k) same = cleanP[A.colIdx[k]];
                    cleanT[u] = same;
                }
            });
        }
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
        const auto none = [](int) { return false; };
        // P pattern: row u -> aggregates of its (non-excluded) closed
        // neighbourhood; with it the smoothing contributions (the A blocks of
        // row u grouped by the aggregate of their column, k ascending).
        {
            const auto gen = [&](int u, auto&& emit) {
                if (A.excluded[u]) return;
                for (int k = A.rowPtr[u]; k < A.rowPtr[u + 1]; ++k) {
                    const int w = A.colIdx[k];
                    if (!A.excluded[w]) emit(agg[w], k, 0);
                }
            };
            const auto mapA = [&](int u, int k) { return k + dA[u]; };
            const auto mapB = [](int, int, int b) { return b; };
            if (O)
                product_rows(A.n, nc, nt0, gen, [&](int u) { return cleanP[u] != 0; },
                             ProductRef{&O->pPtr, &O->pCol, &O->psPtr, &O->psA, nullptr}, mapA, mapB, A.pPtr,
                             A.pCol, A.psPtr, A.psA, nullptr);
            else
                product_rows(A.n, nc, nt0, gen, none, ProductRef{}, mapA, mapB, A.pPtr, A.pCol, A.psPtr, A.psA,
                             nullptr);
        }
        const int nP = (int)A.pCol.size();
        A.pRow.resize(nP);
        A.pIsTent.resize(nP);
        parallel_row_chunks(A.n, nt0, [&](int, int r0, int r1) {
            for (int u = r0; u < r1; ++u) {
                for (int j = A.pPtr[u]; j < A.pPtr[u + 1]; ++j) {
                    A.pRow[j] = u;
                    A.pIsTent[j] = A.pCol[j] == agg[u] ? 1 : 0;
                }
                if (O && cleanP[u]) dP[u] = A.pPtr[u] - O->pPtr[u];
            }
        });
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
            const int nt = setup_chunks((long)A.pCol.size() * 8);
            const auto gen = [&](int u, auto&& emit) {
                for (int k = A.rowPtr[u]; k < A.rowPtr[u + 1]; ++k) {
                    const int w = A.colIdx[k];
                    for (int j = A.pPtr[w]; j < A.pPtr[w + 1]; ++j) emit(A.pCol[j], k, j);
                }
            };
            const auto mapA = [&](int u, int k) { return k + dA[u]; };
            const auto mapB = [&](int, int, int j) { return j + dP[O->pRow[j]]; };
            if (O)
                product_rows(A.n, nc, nt, gen, [&](int u) { return cleanT[u] != 0; },
                             ProductRef{&O->tPtr, &O->tCol, &O->tcPtr, &O->tcA, &O->tcP}, mapA, mapB, A.tPtr, A.tCol,
                             A.tcPtr, A.tcA, &A.tcP);
            else
                product_rows(A.n, nc, nt, gen, none, ProductRef{}, mapA, mapB, A.tPtr, A.tCol, A.tcPtr, A.tcA,
                             &A.tcP);
        }
        if (O)
            for (int u = 0; u < A.n; ++u)
                if (cleanT[u]) dT[u] = A.tPtr[u] - O->tPtr[u];
        // coarse A = P^T T: pattern and (P block, T block) pairs, rows by coarse
        // node: over the P blocks (u, c1) in increasing block order (the
        // restriction lists), the T entries of row u. A coarse row is reused
        // when its restriction list has the same rows, each cleanP and cleanT.
        AmgHostLevel nxl = take(L.size());
        AmgHostLevel* nx = &nxl;
        nx->n = nc;
        nx->bs = kAmgNullDim;
        nx->excluded.assign(nc, 0);
        // an aggregate holding a late node is late on the next level
        nx->late.clear();
        if (!A.late.empty()) {
            nx->late.assign(nc, 0);
            for (int u = 0; u < A.n; ++u)
                if (agg[u] >= 0 && A.late[u]) nx->late[agg[u]] = 1;
        }
        {
            const int nt = setup_chunks((long)A.tCol.size() * 4);
            std::vector<unsigned char> cleanC;
            if (O) {
                cleanC.assign(nc, 0);
                parallel_row_chunks(nc, nt, [&](int, int r0, int r1) {
                    for (int c1 = r0; c1 < r1 && c1 < O->nc; ++c1) {
                        const int q0 = A.rPtr[c1], q1 = A.rPtr[c1 + 1], o0 = O->rPtr[c1];
                        if (q1 - q0 != O->rPtr[c1 + 1] - o0) continue;
                        bool same = true;
                        for (int q = 0; q < q1 - q0 && same; ++q) {
                            const int u = A.pRow[A.rP[q0 + q]];
                            same = u == O->pRow[O->rP[o0 + q]] && cleanP[u] && cleanT[u];
                        }
                        cleanC[c1] = same;
                    }
                });
            }
            const auto gen = [&](int c1, auto&& emit) {
                for (int q = A.rPtr[c1]; q < A.rPtr[c1 + 1]; ++q) {
                    const int j = A.rP[q];
                    const int u = A.pRow[j];
                    for (int t = A.tPtr[u]; t < A.tPtr[u + 1]; ++t) emit(A.tCol[t], j, t);
                }
            };
            const auto mapA = [&](int, int j) { return j + dP[O->pRow[j]]; };
            const auto mapB = [&](int, int j, int t) { return t + dT[O->pRow[j]]; };
            if (O)
                product_rows(nc, nc, nt, gen, [&](int c1) { return cleanC[c1] != 0; },
                             ProductRef{&Onext->rowPtr, &Onext->colIdx, &O->acPtr, &O->acP, &O->acT},
                             mapA, mapB, nx->rowPtr, nx->colIdx, A.acPtr, A.acP, &A.acT);
            else
                product_rows(nc, nc, nt, gen, none, ProductRef{}, mapA, mapB, nx->rowPtr, nx->colIdx, A.acPtr,
                             A.acP, &A.acT);
        }
        L.push_back(std::move(nxl));
    }
    // the coarsest level has no transfer: drop what its storage held
    AmgHostLevel& C = L.back();
    for (auto* v : {&C.agg, &C.aggPtr, &C.aggMem, &C.pPtr, &C.pCol, &C.pRow, &C.psPtr, &C.psA, &C.rPtr, &C.rP,
                    &C.tPtr, &C.tCol, &C.tcPtr, &C.tcA, &C.tcP, &C.acPtr, &C.acP, &C.acT})
        v->clear();
    C.pIsTent.clear();
    return 0;
}

}  // namespace oc
