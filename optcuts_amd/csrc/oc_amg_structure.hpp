// Host-side structure of the smoothed-aggregation hierarchy: aggregates,
// prolongator pattern and the contribution maps of its numeric phase. Pure
// C++ (no device calls); oc_amg.hip uploads the result.
#pragma once

#include <algorithm>
#include <numeric>
#include <string>
#include <utility>
#include <vector>

namespace oc {

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
        // P pattern: row u -> aggregates of its (non-excluded) closed neighbourhood
        A.pPtr.assign(A.n + 1, 0);
        A.pCol.clear();
        A.pRow.clear();
        A.pIsTent.clear();
        for (int u = 0; u < A.n; ++u) {
            if (!A.excluded[u]) {
                std::vector<int> cs;
                for (int k = A.rowPtr[u]; k < A.rowPtr[u + 1]; ++k) {
                    const int w = A.colIdx[k];
                    if (!A.excluded[w]) cs.push_back(agg[w]);
                }
                std::sort(cs.begin(), cs.end());
                cs.erase(std::unique(cs.begin(), cs.end()), cs.end());
                for (int c : cs) {
                    A.pCol.push_back(c);
                    A.pRow.push_back(u);
                    A.pIsTent.push_back(c == agg[u] ? 1 : 0);
                }
            }
            A.pPtr[u + 1] = (int)A.pCol.size();
        }
        const int nP = (int)A.pCol.size();
        auto pfind = [&](int u, int c) {
            const int* b = A.pCol.data() + A.pPtr[u];
            const int* e = A.pCol.data() + A.pPtr[u + 1];
            const int* p = std::lower_bound(b, e, c);
            return (p != e && *p == c) ? int(p - A.pCol.data()) : -1;
        };
        // smoothing contributions: A blocks of row u bucketed by aggregate of their column
        A.psPtr.assign(nP + 1, 0);
        for (int u = 0; u < A.n; ++u) {
            if (A.excluded[u]) continue;
            for (int k = A.rowPtr[u]; k < A.rowPtr[u + 1]; ++k) {
                const int w = A.colIdx[k];
                if (!A.excluded[w]) A.psPtr[pfind(u, agg[w]) + 1]++;
            }
        }
        for (int j = 0; j < nP; ++j) A.psPtr[j + 1] += A.psPtr[j];
        A.psA.assign(A.psPtr[nP], 0);
        {
            std::vector<int> f(A.psPtr.begin(), A.psPtr.end() - 1);
            for (int u = 0; u < A.n; ++u) {
                if (A.excluded[u]) continue;
                for (int k = A.rowPtr[u]; k < A.rowPtr[u + 1]; ++k) {
                    const int w = A.colIdx[k];
                    if (!A.excluded[w]) A.psA[f[pfind(u, agg[w])]++] = k;
                }
            }
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
        // T = A P: pattern and (A block, P block) pairs
        A.tPtr.assign(A.n + 1, 0);
        A.tCol.clear();
        A.tcPtr.assign(1, 0);
        A.tcA.clear();
        A.tcP.clear();
        for (int u = 0; u < A.n; ++u) {
            std::vector<std::pair<int, std::pair<int, int>>> items;  // (coarse col, (k, j))
            for (int k = A.rowPtr[u]; k < A.rowPtr[u + 1]; ++k) {
                const int w = A.colIdx[k];
                for (int j = A.pPtr[w]; j < A.pPtr[w + 1]; ++j) items.push_back({A.pCol[j], {k, j}});
            }
            std::stable_sort(items.begin(), items.end(),
                             [](const auto& a, const auto& b) { return a.first < b.first; });
            for (size_t q = 0; q < items.size(); ++q) {
                if (q == 0 || items[q].first != items[q - 1].first) {
                    if (q) A.tcPtr.push_back((int)A.tcA.size());
                    A.tCol.push_back(items[q].first);
                }
                A.tcA.push_back(items[q].second.first);
                A.tcP.push_back(items[q].second.second);
            }
            if (!items.empty()) A.tcPtr.push_back((int)A.tcA.size());
            A.tPtr[u + 1] = (int)A.tCol.size();
        }
        // coarse A = P^T T: pattern and (P block, T block) pairs, rows by coarse node
        AmgHostLevel nxl;
        AmgHostLevel* nx = &nxl;
        nx->n = nc;
        nx->bs = kAmgNullDim;
        nx->excluded.assign(nc, 0);
        {
            std::vector<int> cnt(nc + 1, 0);
            for (int u = 0; u < A.n; ++u)
                for (int j = A.pPtr[u]; j < A.pPtr[u + 1]; ++j) cnt[A.pCol[j] + 1] += A.tPtr[u + 1] - A.tPtr[u];
            for (int c = 0; c < nc; ++c) cnt[c + 1] += cnt[c];
            std::vector<int> c2(cnt[nc]), jp(cnt[nc]), tp(cnt[nc]);
            std::vector<int> f(cnt.begin(), cnt.end() - 1);
            for (int u = 0; u < A.n; ++u)
                for (int j = A.pPtr[u]; j < A.pPtr[u + 1]; ++j) {
                    const int c1 = A.pCol[j];
                    for (int t = A.tPtr[u]; t < A.tPtr[u + 1]; ++t) {
                        c2[f[c1]] = A.tCol[t];
                        jp[f[c1]] = j;
                        tp[f[c1]] = t;
                        ++f[c1];
                    }
                }
            nx->rowPtr.assign(nc + 1, 0);
            nx->colIdx.clear();
            A.acPtr.assign(1, 0);
            A.acP.clear();
            A.acT.clear();
            std::vector<int> idx;
            for (int c1 = 0; c1 < nc; ++c1) {
                idx.resize(cnt[c1 + 1] - cnt[c1]);
                std::iota(idx.begin(), idx.end(), cnt[c1]);
                std::stable_sort(idx.begin(), idx.end(), [&](int a, int b) { return c2[a] < c2[b]; });
                for (size_t q = 0; q < idx.size(); ++q) {
                    if (q == 0 || c2[idx[q]] != c2[idx[q - 1]]) {
                        if (q) A.acPtr.push_back((int)A.acP.size());
                        nx->colIdx.push_back(c2[idx[q]]);
                    }
                    A.acP.push_back(jp[idx[q]]);
                    A.acT.push_back(tp[idx[q]]);
                }
                if (!idx.empty()) A.acPtr.push_back((int)A.acP.size());
                nx->rowPtr[c1 + 1] = (int)nx->colIdx.size();
            }
        }
        L.push_back(std::move(nxl));
    }
    return 0;
}

}  // namespace oc
