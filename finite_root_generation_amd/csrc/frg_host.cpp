// frg_host.cpp -- host post-processing of a round (frgsupport.cu), behind the C ABI.
//
// frg_sort_dedupe_merge: the reference sorts the new roots on Re(z_0) on the device
// (frg.cu:432-435), merges them into the sorted collection through a std::list of indices,
// marks duplicates by scanning left and right neighbours while |dRe z_0| <= tol, and folds the
// result into a std::deque (frgsupport.cu:10-215).  Here the same result is produced with one
// stable sort of the new roots and a linear merge; the duplicate rule is unchanged:
//   a new root with code 0 becomes code 6 if a neighbour in the merged order, reached before
//   |dRe z_0| exceeds tol, is an old root, or a new root of smaller index that still has code 0,
//   and all 42 reals of the two roots differ by at most tol.
// New roots are examined in index order, so "still has code 0" sees earlier decisions, as in
// the reference's loop over tloc (frgsupport.cu:69).
#include <algorithm>
#include <cmath>
#include <cstdint>
#include <cstring>
#include <new>
#include <numeric>
#include <vector>

#include "../../include/frg.h"

namespace {

inline bool same_root(const frg_c64 *a, const frg_c64 *b, double tol) {
    for (int j = 0; j < FRG_NV; ++j) {
        if (std::fabs(a[j].re - b[j].re) > tol || std::fabs(a[j].im - b[j].im) > tol) return false;
    }
    return true;
}

// Principal branch W0 on [-1/e, inf) by Halley's iteration on w e^w - z.
double lambertw0(double z) {
    const double em1 = 0.36787944117144233;  // 1/e
    if (std::isnan(z) || z < -em1) return NAN;
    if (z == 0.0) return 0.0;
    double w;
    if (z < 1.0) {
        // expansion about the branch point in p = sqrt(2(e z + 1))
        const double p = std::sqrt(std::max(0.0, 2.0 * (2.718281828459045 * z + 1.0)));
        w = -1.0 + p - p * p / 3.0 + 11.0 / 72.0 * p * p * p;
        if (z > -0.25) w = z * (1.0 - z + 1.5 * z * z);  // series about 0
    } else {
        const double l = std::log(z);
        w = l - std::log(std::max(l, 1e-300));
    }
    for (int it = 0; it < 60; ++it) {
        const double ew = std::exp(w);
        const double f = w * ew - z;
        const double wp1 = w + 1.0;
        if (wp1 == 0.0) break;
        const double dw = f / (ew * wp1 - (w + 2.0) * f / (2.0 * wp1));
        w -= dw;
        if (std::fabs(dw) <= 1e-16 * (1.0 + std::fabs(w))) break;
    }
    return w;
}

}  // namespace

extern "C" {

double frg_lambertw(double z) { return lambertw0(z); }

double frg_completeness(int64_t successes, int64_t duplicates) {
    if (successes + duplicates <= 0 || successes <= 0) return 0.0;
    const double r = (double)successes / ((double)successes + (double)duplicates);
    return r * lambertw0((-1.0 / r) * std::exp(-1.0 / r)) + 1.0;
}

int64_t frg_sort_dedupe_merge(const frg_c64 *coll, int64_t m, const frg_c64 *fpt, int32_t *err, int64_t n, double tol,
                              frg_c64 *coll_out) try {
    if (m < 0 || n < 0 || m > FRG_MAX_COUNT || n > FRG_MAX_COUNT || (m > 0 && !coll) || (n > 0 && (!fpt || !err)) ||
        !coll_out || !(tol >= 0.0))
        return FRG_EINVAL;
    auto key = [&](int64_t g) -> double { return g < m ? coll[g * FRG_NV].re : fpt[(g - m) * FRG_NV].re; };
    // new roots in ascending Re z_0 (stable; NaN keys last)
    std::vector<int64_t> ord((size_t)n);
    std::iota(ord.begin(), ord.end(), 0);
    std::stable_sort(ord.begin(), ord.end(), [&](int64_t a, int64_t b) {
        const double ka = fpt[a * FRG_NV].re, kb = fpt[b * FRG_NV].re;
        const bool na = std::isnan(ka), nb = std::isnan(kb);
        if (na != nb) return nb;
        return !na && ka < kb;
    });
    // merged order of global ids (old g < m, new m + i); old first on equal keys
    std::vector<int64_t> merged((size_t)(m + n));
    std::vector<int64_t> rank((size_t)(m + n));
    {
        int64_t io = 0, in = 0, r = 0;
        while (io < m || in < n) {
            const bool take_old = io < m && (in >= n || !(fpt[ord[in] * FRG_NV].re < coll[io * FRG_NV].re));
            merged[r++] = take_old ? io++ : m + ord[in++];
        }
        for (int64_t i = 0; i < m + n; ++i) rank[merged[i]] = i;
    }
    for (int64_t t = 0; t < n; ++t) {
        if (err[t] != 0) continue;
        const frg_c64 *rt = fpt + t * FRG_NV;
        const double kt = rt[0].re;
        const int64_t r0 = rank[m + t];
        for (int dir = -1; dir <= 1 && err[t] == 0; dir += 2) {
            for (int64_t i = r0 + dir; i >= 0 && i < m + n; i += dir) {
                const int64_t g = merged[i];
                if (std::fabs(kt - key(g)) > tol) break;
                if (g < m) {
                    if (same_root(rt, coll + g * FRG_NV, tol)) { err[t] = 6; break; }
                } else {
                    const int64_t u = g - m;
                    if (u < t && err[u] == 0 && same_root(rt, fpt + u * FRG_NV, tol)) { err[t] = 6; break; }
                }
            }
        }
    }
    int64_t k = 0;
    for (int64_t i = 0; i < m + n; ++i) {
        const int64_t g = merged[i];
        const frg_c64 *src = g < m ? coll + g * FRG_NV : (err[g - m] == 0 ? fpt + (g - m) * FRG_NV : nullptr);
        if (src) std::memcpy(coll_out + (k++) * FRG_NV, src, sizeof(frg_c64) * FRG_NV);
    }
    return k;
} catch (const std::bad_alloc &) {
    return FRG_ENOMEM;
}

}  // extern "C"
