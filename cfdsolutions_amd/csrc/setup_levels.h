// Host-side set-up of the ILU(0) factorisation and sweep schedules, on up to
// 16 threads (cfd_parallel_for must be visible where this is included).  The
// results equal those of one sequential pass over the rows.
#pragma once

#include <algorithm>
#include <climits>
#include <cstdint>
#include <memory>
#include <vector>

// ILU(0) merge map of a BSR pattern (ia, ja sorted per row, iau the diagonal
// slots): for the kk-th lower block (i, k) of row i, moff[kk] is the first of
// the map entries of the upper blocks (k, j) of row k, and mqq[moff[kk] + t]
// is the slot of block (i, j) in row i for the t-th of them, or -1 when (i, j)
// is outside the pattern (ILU(0) drops it).  Entries are numbered row by row,
// lower block by lower block, as a sequential pass would number them; moff is
// 0 at the other slots.  Every entry of both arrays is written once, by the
// thread that owns its row (no separate initialisation pass over them).
// Returns false when there are more than INT32_MAX entries.
inline bool build_merge_map_par(const int *ia, const int *ja, const int *iau, int N, std::unique_ptr<int[]> &moff,
                                std::unique_ptr<int[]> &mqq, int64_t &total) {
    auto count = [&](int kk) {
        const int k = ja[kk];
        return (int64_t)(ia[k + 1] - (iau[k] + 1));
    };
    std::vector<int64_t> start((size_t)N + 1, 0);
    cfd_parallel_for(N, [&](int64_t i0, int64_t i1) {
        for (int i = (int)i0; i < (int)i1; i++) {
            int64_t s = 0;
            for (int kk = ia[i]; kk < iau[i]; kk++) s += count(kk);
            start[i + 1] = s;
        }
    });
    for (int i = 0; i < N; i++) start[i + 1] += start[i];
    const int64_t tot = start[N];
    if (tot > INT32_MAX) return false;
    moff.reset(new int[(size_t)ia[N] + 1]);
    mqq.reset(new int[(size_t)std::max<int64_t>(tot, 1)]);
    mqq[0] = -1;  // keeps the array non-empty when there are no entries
    cfd_parallel_for(N, [&](int64_t i0, int64_t i1) {
        for (int i = (int)i0; i < (int)i1; i++) {
            int64_t t = start[i];
            for (int kk = ia[i]; kk < iau[i]; kk++) {
                moff[kk] = (int)t;
                const int k = ja[kk];
                int qq = kk + 1;
                const int qe = ia[i + 1];
                for (int pp = iau[k] + 1; pp < ia[k + 1]; pp++, t++) {
                    while (qq < qe && ja[qq] < ja[pp]) qq++;
                    mqq[(size_t)t] = qq < qe && ja[qq] == ja[pp] ? qq : -1;
                }
            }
            for (int kk = iau[i]; kk < ia[i + 1]; kk++) moff[kk] = 0;
        }
    });
    total = tot;
    return true;
}

// Level schedule of the lower (or upper) triangular sweep of a multicolour
// ordered pattern: rows [cstart[c], cstart[c+1]) have colour c and no block
// couples two rows of one colour, so a row's lower neighbours all have
// smaller colours (upper: larger).  The colours are taken in order (upper:
// in reverse), the rows of each in parallel.  level(i) = 1 + the largest
// level of the rows it depends on (0 for none); ptr[l]..ptr[l+1] index the
// rows of level l in rows, ascending within a level; maxw is the widest level.
inline void build_levels_by_colour(const int *ia, const int *ja, const int *iau, int N, bool upper,
                                   const std::vector<int> &cstart, std::vector<int> &ptr, std::vector<int> &rows,
                                   int &maxw) {
    std::vector<int> lev((size_t)N, 0);
    const int nc = (int)cstart.size() - 1;
    for (int q = 0; q < nc; q++) {
        const int c = upper ? nc - 1 - q : q;
        cfd_parallel_for(cstart[c + 1] - cstart[c], [&](int64_t r0, int64_t r1) {
            for (int i = cstart[c] + (int)r0; i < cstart[c] + (int)r1; i++) {
                int l = 0;
                const int k0 = upper ? iau[i] + 1 : ia[i], k1 = upper ? ia[i + 1] : iau[i];
                for (int k = k0; k < k1; k++) l = std::max(l, lev[ja[k]] + 1);
                lev[i] = l;
            }
        });
    }
    int nl = 0;
    for (int i = 0; i < N; i++) nl = std::max(nl, lev[i] + 1);
    ptr.assign((size_t)nl + 1, 0);
    for (int i = 0; i < N; i++) ptr[lev[i] + 1]++;
    maxw = 0;
    for (int l = 0; l < nl; l++) {
        maxw = std::max(maxw, ptr[l + 1]);
        ptr[l + 1] += ptr[l];
    }
    rows.assign((size_t)N, 0);
    std::vector<int> pos(ptr.begin(), ptr.end() - 1);
    for (int i = 0; i < N; i++) rows[pos[lev[i]]++] = i;
}
