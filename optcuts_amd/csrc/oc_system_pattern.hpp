// Host-side pattern of the Newton system over a triangle set (mesh triangles,
// then air-mesh triangles): per-vertex corner lists, the 2x2-block CSR pattern
// and the assembly map of its upper blocks (LinSysSolver::set_pattern,
// LinSysSolver.hpp:37-135, and the slot map that update_a walks,
// LinSysSolver.hpp:138-159). Pure C++ (no device calls); oc_capi.hip uploads
// the result.
#pragma once

#include "oc_amg_structure.hpp"  // parallel_row_chunks, setup_threads

#include <algorithm>
#include <vector>

namespace oc {

// Block CSR pattern: sorted rows, each with its diagonal.
struct HostPattern {
    int nV = 0;
    std::vector<int> rowPtr, colIdx, diagPos;
    long nnzb() const { return (long)colIdx.size(); }
    int find(int u, int w) const
    {
        const int* b = colIdx.data() + rowPtr[u];
        const int* e = colIdx.data() + rowPtr[u + 1];
        const int* p = std::lower_bound(b, e, w);
        return (p != e && *p == w) ? int(p - colIdx.data()) : -1;
    }
};

// vertex -> (triangle*3 + corner) lists, triangles ascending
inline void build_vertex_lists(int nV, int nF, const int* F, std::vector<int>& ptr, std::vector<int>& src)
{
    ptr.assign(nV + 1, 0);
    for (int t = 0; t < nF; ++t)
        for (int k = 0; k < 3; ++k) ptr[F[3 * t + k] + 1]++;
    for (int v = 0; v < nV; ++v) ptr[v + 1] += ptr[v];
    src.resize(ptr[nV]);
    static thread_local std::vector<int> fill;  // kept per host thread for its memory
    fill.assign(ptr.begin(), ptr.end() - 1);
    for (int t = 0; t < nF; ++t)
        for (int k = 0; k < 3; ++k) src[fill[F[3 * t + k]]++] = 3 * t + k;
}

// Upper blocks of the pattern and their assembly map: upper block b = (ubRow,
// ubCol) with ubCol >= ubRow sits at ubPosU in the CSR and its mirror at
// ubPosL (-1 on the diagonal); ubCSrc[ubCPtr[b] .. ubCPtr[b+1]) lists the
// element blocks that assemble it, in triangle order, as 6 t + id (id: 0..5
// for the corner pairs (0,0) (0,1) (0,2) (1,1) (1,2) (2,2)), with the top bit
// set when the block is the transpose of the element's (a > b).
struct SystemMaps {
    std::vector<int> ubRow, ubCol, ubPosU, ubPosL, ubCPtr, ubCSrc;
};

// Pattern of the Newton matrix of triangles Ft over nVt vertices (one row per
// vertex from its triangles: fixed rows keep only the diagonal, free rows drop
// fixed columns) and its upper-block maps; vcPtr/vcSrc are Ft's vertex lists.
//
// prevP/prevM/keep (optional): an earlier build and, per vertex u < nKeep
// (nKeep <= prevP->nV), keep[u] = 1 when u's triangles are the same triangles
// (same indices, same vertices, same fixed flags) in both. Such rows are
// copied from the earlier build instead of regenerated; the result is
// identical to a build without them. Rows run in chunks on host threads and are written at their final
// offsets, so the output does not depend on the thread count.
// Work arrays of build_system_maps, kept per host thread so that repeated builds (one per air re-mesh) reuse their
// memory instead of allocating and first-touching megabytes each time.
struct PatternScratch {
    std::vector<std::vector<int>> part;
    std::vector<int> len, ubStart, ubStartO, posToUB, cnt;
};
inline PatternScratch& pattern_scratch()
{
    static thread_local PatternScratch s;
    return s;
}

// P and M may hold an earlier build's arrays (their memory is reused; every entry is rewritten).
inline void build_system_maps(int nVt, const std::vector<int>& Ft, const unsigned char* fixed,
                              const std::vector<int>& vcPtr, const std::vector<int>& vcSrc, HostPattern& P,
                              SystemMaps& M, const HostPattern* prevP = nullptr, const SystemMaps* prevM = nullptr,
                              const unsigned char* keep = nullptr, int nKeep = 0)
{
    const long nFt = (long)Ft.size() / 3;
    // rows in 8 chunks per host thread, claimed dynamically: the rows an air mesh changes (which are regenerated,
    // not copied) sit together at the end of the numbering, so one chunk per thread would leave one thread with
    // most of the work
    const int nth = setup_threads(3L * nFt);
    const int nch = nth > 1 ? 8 * nth : 1;
    const bool reuse = prevP && prevM && keep;
    const auto kept = [&](int u) { return reuse && u < nKeep && u < prevP->nV && keep[u]; };
    PatternScratch& S = pattern_scratch();

    // ---- block pattern
    P.nV = nVt;
    {
        std::vector<std::vector<int>>& part = S.part;
        part.resize(nch);
        for (auto& q : part) q.clear();
        std::vector<int>& len = S.len;
        len.resize(nVt);
        parallel_row_chunks(nVt, nch, [&](int ch, int r0, int r1) {
            std::vector<int> loc;
            auto& out = part[ch];
            for (int v = r0; v < r1; ++v) {
                if (kept(v)) {
                    len[v] = prevP->rowPtr[v + 1] - prevP->rowPtr[v];
                    continue;
                }
                const size_t n0 = out.size();
                if (fixed[v]) {
                    out.push_back(v);
                } else {
                    loc.assign(1, v);
                    for (int q = vcPtr[v]; q < vcPtr[v + 1]; ++q) {
                        const int t = vcSrc[q] / 3;
                        for (int k = 0; k < 3; ++k) loc.push_back(Ft[3 * (size_t)t + k]);
                    }
                    std::sort(loc.begin(), loc.end());
                    loc.erase(std::unique(loc.begin(), loc.end()), loc.end());
                    for (int w : loc)
                        if (w == v || !fixed[w]) out.push_back(w);
                }
                len[v] = (int)(out.size() - n0);
            }
        });
        P.rowPtr.assign(nVt + 1, 0);
        for (int v = 0; v < nVt; ++v) P.rowPtr[v + 1] = P.rowPtr[v] + len[v];
        P.colIdx.resize(P.rowPtr[nVt]);
        P.diagPos.resize(nVt);
        parallel_row_chunks(nVt, nch, [&](int ch, int r0, int r1) {
            size_t o = 0;
            const auto& in = part[ch];
            for (int v = r0; v < r1; ++v) {
                const int k0 = P.rowPtr[v];
                if (kept(v)) {
                    const int p0 = prevP->rowPtr[v];
                    std::copy(prevP->colIdx.begin() + p0, prevP->colIdx.begin() + p0 + len[v],
                              P.colIdx.begin() + k0);
                    P.diagPos[v] = prevP->diagPos[v] - p0 + k0;
                } else {
                    std::copy(in.begin() + o, in.begin() + o + len[v], P.colIdx.begin() + k0);
                    o += len[v];
                    P.diagPos[v] = -1;
                    for (int k = k0; k < k0 + len[v]; ++k)
                        if (P.colIdx[k] == v) P.diagPos[v] = k;
                }
            }
        });
    }

    // ---- upper blocks (w >= u) in row order: rows are sorted and hold their
    // diagonal, so row u's upper blocks are its entries from the diagonal on
    std::vector<int>& ubStart = S.ubStart;
    ubStart.assign(nVt + 1, 0);
    for (int u = 0; u < nVt; ++u) ubStart[u + 1] = ubStart[u] + (P.rowPtr[u + 1] - P.diagPos[u]);
    const int nUB = ubStart[nVt];
    std::vector<int>& posToUB = S.posToUB;
    posToUB.assign(P.nnzb(), -1);
    M.ubRow.resize(nUB);
    M.ubCol.resize(nUB);
    M.ubPosU.resize(nUB);
    M.ubPosL.resize(nUB);
    parallel_row_chunks(nVt, nch, [&](int, int r0, int r1) {
        for (int u = r0; u < r1; ++u) {
            int b = ubStart[u];
            for (int k = P.diagPos[u]; k < P.rowPtr[u + 1]; ++k, ++b) {
                const int w = P.colIdx[k];
                posToUB[k] = b;
                M.ubRow[b] = u;
                M.ubCol[b] = w;
                M.ubPosU[b] = k;
                M.ubPosL[b] = w == u ? -1 : P.find(w, u);
            }
        }
    });

    // ---- contributions per upper block in triangle order. Block (u, w),
    // u <= w, receives the element block (a, b) of every triangle t with
    // f[a] = u and f[b] = w (for u == w only a == b); walking u's triangles in
    // order gives each of its blocks its contributions in triangle order.
    std::vector<int>& ubStartO = S.ubStartO;
    if (reuse) {
        ubStartO.assign(prevP->nV + 1, 0);
        for (int u = 0; u < prevP->nV; ++u)
            ubStartO[u + 1] = ubStartO[u] + (prevP->rowPtr[u + 1] - prevP->diagPos[u]);
    }
    auto row_contribs = [&](int u, auto&& fn) {
        if (fixed[u]) return;
        for (int q = vcPtr[u]; q < vcPtr[u + 1]; ++q) {
            const int t = vcSrc[q] / 3, a = vcSrc[q] % 3;
            const int* f = &Ft[3 * (size_t)t];
            for (int b = 0; b < 3; ++b) {
                const int w = f[b];
                if (fixed[w] || w < u || (w == u && a != b)) continue;
                const int lo = a < b ? a : b, hi = a < b ? b : a;
                const int id = lo == 0 ? hi : (lo == 1 ? 2 + hi : 5);
                const unsigned int code = (unsigned int)(6 * t + id) | (a > b ? 0x80000000u : 0u);
                fn(posToUB[P.find(u, w)], code);
            }
        }
    };
    std::vector<int>& cnt = S.cnt;
    cnt.assign(nUB, 0);
    parallel_row_chunks(nVt, nch, [&](int, int r0, int r1) {
        for (int u = r0; u < r1; ++u) {
            if (kept(u)) {
                const int bo = ubStartO[u];
                for (int b = ubStart[u]; b < ubStart[u + 1]; ++b)
                    cnt[b] = prevM->ubCPtr[bo + b - ubStart[u] + 1] - prevM->ubCPtr[bo + b - ubStart[u]];
            } else {
                row_contribs(u, [&](int ub, unsigned int) { cnt[ub]++; });
            }
        }
    });
    M.ubCPtr.resize((size_t)nUB + 1);
    M.ubCPtr[0] = 0;
    for (int b = 0; b < nUB; ++b) M.ubCPtr[b + 1] = M.ubCPtr[b] + cnt[b];
    M.ubCSrc.resize(M.ubCPtr[nUB]);
    parallel_row_chunks(nVt, nch, [&](int, int r0, int r1) {
        std::vector<int> fill;
        for (int u = r0; u < r1; ++u) {
            const int b0 = ubStart[u], b1 = ubStart[u + 1];
            if (kept(u)) {
                const int s0 = prevM->ubCPtr[ubStartO[u]], s1 = prevM->ubCPtr[ubStartO[u + 1]];
                std::copy(prevM->ubCSrc.begin() + s0, prevM->ubCSrc.begin() + s1, M.ubCSrc.begin() + M.ubCPtr[b0]);
            } else {
                fill.assign(M.ubCPtr.begin() + b0, M.ubCPtr.begin() + b1);
                row_contribs(u, [&](int ub, unsigned int code) { M.ubCSrc[fill[ub - b0]++] = (int)code; });
            }
        }
    });
}

}  // namespace oc
