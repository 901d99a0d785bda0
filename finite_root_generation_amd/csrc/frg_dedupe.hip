// frg_dedupe.hip -- device-side sort, duplicate marking and merge of a round's roots.
//
// Replaces the device sort of the new roots on Re z_0 (frg.cu:432-435, thrust::sort_by_key) and
// the host SortingDeletingMerging (frgsupport.cu:10-215), with the same result as the host
// frg_sort_dedupe_merge (bit-identical collection and codes):
//   1. the new roots are sorted stably on Re z_0 (NaN last) by a radix sort of an order-preserving
//      64-bit image of the key (-0 folded onto +0);
//   2. merge ranks by binary search: an old root goes before a new root of equal key, and old
//      roots with a NaN key (they sit at the end of the sorted collection) follow the last finite
//      old root;
//   3. the merged order is cut into runs wherever two consecutive keys differ by more than tol.  A
//      duplicate scan from a new root with a finite key stops at the first key farther than tol,
//      so it never leaves the root's run; the run's new roots are put in index order by a radix
//      sort of (run, index) keys, and one thread per run then applies the sequential rule in that
//      order: a code-0 root becomes code 6 if a neighbour reached before |d Re z_0| > tol is an old
//      root, or a new root of smaller index still at code 0, and all 42 reals differ by at most
//      tol.  A run of L roots costs O(L) plus its duplicate scans (the non-finite keys of failed
//      paths, which never break a run, form one long run at the end).  If a code-0 new root has a
//      NaN key its scan is not bounded, and one thread runs the whole rule serially instead;
//   4. old roots and the new roots still at code 0 are compacted into the merged collection.
#include <hip/hip_runtime.h>

#include <cstring>

#include <rocprim/device/device_radix_sort.hpp>
#include <rocprim/device/device_scan.hpp>

#include "frg_dedupe_kernels.h"

namespace frg {
namespace {

constexpr int kNV = 21;

__device__ __forceinline__ uint64_t order_key(double x) {
    if (x != x) return ~0ull;
    if (x == 0.0) x = 0.0;
    const uint64_t b = (uint64_t)__double_as_longlong(x);
    return (b >> 63) ? ~b : (b | 0x8000000000000000ull);
}

__device__ __forceinline__ double key_of(const double2 *r) { return r[0].x; }

__device__ __forceinline__ bool same_root(const double2 *a, const double2 *b, double tol) {
    for (int j = 0; j < kNV; ++j) {
        if (fabs(a[j].x - b[j].x) > tol || fabs(a[j].y - b[j].y) > tol) return false;
    }
    return true;
}

__global__ void keys_kernel(const double2 *fpt, int64_t n, uint64_t *keys, int64_t *idx) {
    const int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= n) return;
    keys[i] = order_key(key_of(fpt + i * kNV));
    idx[i] = i;
}

// number of old roots with a finite key (they precede the NaN keys)
__global__ void finite_old_kernel(const double2 *coll, int64_t m, int64_t *m_f) {
    if (blockIdx.x != 0 || threadIdx.x != 0) return;
    int64_t lo = 0, hi = m;
    while (lo < hi) {
        const int64_t mid = (lo + hi) >> 1;
        const double k = key_of(coll + mid * kNV);
        if (k == k) lo = mid + 1;
        else hi = mid;
    }
    *m_f = lo;
}

__global__ void rank_new_kernel(const double2 *coll, int64_t m, const int64_t *m_f_p, const double2 *fpt,
                                const int64_t *sidx, int64_t n, int64_t *mid, double *mkey) {
    const int64_t r = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (r >= n) return;
    const int64_t m_f = *m_f_p;
    const int64_t i = sidx[r];
    const double kd = key_of(fpt + i * kNV);
    // first finite old root with kd < key; none: the new root comes after every old root
    int64_t lo = 0, hi = m_f;
    while (lo < hi) {
        const int64_t md = (lo + hi) >> 1;
        if (kd < key_of(coll + md * kNV)) hi = md;
        else lo = md + 1;
    }
    const int64_t before = lo < m_f ? lo : m;
    mid[r + before] = m + i;
    mkey[r + before] = kd;
}

__global__ void rank_old_kernel(const double2 *coll, int64_t m, const int64_t *m_f_p, const double2 *fpt,
                                const int64_t *sidx, int64_t n, int64_t *mid, double *mkey) {
    const int64_t g = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (g >= m) return;
    const int64_t m_f = *m_f_p;
    // a NaN old root follows the last finite old root: the new roots ahead of it are those that
    // precede that finite root
    const double kg = key_of(coll + g * kNV);
    const double kref = g < m_f ? kg : (m_f > 0 ? key_of(coll + (m_f - 1) * kNV) : 0.0);
    int64_t cnt = 0;
    if (g < m_f || m_f > 0) {
        int64_t lo = 0, hi = n;  // first sorted new root that is not < kref
        while (lo < hi) {
            const int64_t md = (lo + hi) >> 1;
            if (key_of(fpt + sidx[md] * kNV) < kref) lo = md + 1;
            else hi = md;
        }
        cnt = lo;
    }
    mid[g + cnt] = g;
    mkey[g + cnt] = kg;
}

__global__ void breaks_kernel(const double *mkey, int64_t total, double tol, int32_t *brk) {
    const int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= total) return;
    brk[i] = (i == 0 || mkey[i] - mkey[i - 1] > tol) ? 1 : 0;
}

__global__ void run_starts_kernel(const int32_t *brk, const int32_t *rid, int64_t total, int64_t *start) {
    const int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= total) return;
    if (brk[i]) start[rid[i]] = i;
}

__global__ void zero_flag_kernel(int32_t *flag) {
    if (blockIdx.x == 0 && threadIdx.x == 0) *flag = 0;
}

// flags code-0 new roots whose key is NaN
__global__ void nan_check_kernel(const double2 *fpt, const int32_t *err, int64_t n, int32_t *flag) {
    const int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= n) return;
    const double k = key_of(fpt + i * kNV);
    if (err[i] == 0 && k != k) atomicOr(flag, 1);
}

// The duplicate scan of new root t from merged position p over [lo, hi).
__device__ bool is_duplicate(const double2 *coll, int64_t m, const double2 *fpt, const int32_t *err,
                             const int64_t *mid, const double *mkey, int64_t lo, int64_t hi, int64_t t, int64_t p,
                             double tol) {
    const double2 *rt = fpt + t * kNV;
    const double kt = key_of(rt);
    for (int dir = -1; dir <= 1; dir += 2) {
        for (int64_t i = p + dir; i >= lo && i < hi; i += dir) {
            if (fabs(kt - mkey[i]) > tol) break;
            const int64_t g = mid[i];
            if (g < m) {
                if (same_root(rt, coll + g * kNV, tol)) return true;
            } else {
                const int64_t u = g - m;
                if (u < t && err[u] == 0 && same_root(rt, fpt + u * kNV, tol)) return true;
            }
        }
    }
    return false;
}

// Key (run, new-root index) and merged position of every new root; the run of merged position i
// is rid[i] + brk[i] - 1 (rid: exclusive sum of the run starts brk).
__global__ void run_order_kernel(const int64_t *mid, int64_t m, const int32_t *brk, const int32_t *rid,
                                 int64_t total, uint64_t *okey, int64_t *opos) {
    const int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= total) return;
    const int64_t g = mid[i];
    if (g < m) return;
    const uint64_t run = (uint64_t)(rid[i] + brk[i] - 1);
    okey[g - m] = (run << 32) | (uint64_t)(g - m);
    opos[g - m] = i;
}

// One thread per run with new roots (the first of its run in the (run, index) order): the run's
// new roots in increasing index order.
__global__ void dedupe_runs_kernel(const double2 *coll, int64_t m, const double2 *fpt, int32_t *err,
                                   const int64_t *mid, const double *mkey, const int64_t *start, int64_t nruns,
                                   int64_t total, const uint64_t *okey, const int64_t *opos, int64_t n, double tol) {
    const int64_t s = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (s >= n) return;
    const uint64_t run = okey[s] >> 32;
    if (s > 0 && (okey[s - 1] >> 32) == run) return;
    const int64_t lo = start[run], hi = (int64_t)run + 1 < nruns ? start[run + 1] : total;
    for (int64_t k = s; k < n && (okey[k] >> 32) == run; ++k) {
        const int64_t t = (int64_t)(okey[k] & 0xffffffffull);
        if (err[t] == 0 && is_duplicate(coll, m, fpt, err, mid, mkey, lo, hi, t, opos[k], tol)) err[t] = 6;
    }
}

// The whole rule serially (a code-0 new root has a NaN key, so its scan is unbounded).
__global__ void dedupe_serial_kernel(const double2 *coll, int64_t m, const double2 *fpt, int32_t *err,
                                     const int64_t *mid, const double *mkey, const int64_t *rank, int64_t n,
                                     int64_t total, double tol) {
    if (blockIdx.x != 0 || threadIdx.x != 0) return;
    for (int64_t t = 0; t < n; ++t)
        if (err[t] == 0 && is_duplicate(coll, m, fpt, err, mid, mkey, 0, total, t, rank[t], tol)) err[t] = 6;
}

__global__ void new_rank_kernel(const int64_t *mid, int64_t m, int64_t total, int64_t *rank) {
    const int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= total) return;
    if (mid[i] >= m) rank[mid[i] - m] = i;
}

__global__ void keep_kernel(const int64_t *mid, int64_t m, const int32_t *err, int64_t total, int32_t *keep) {
    const int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= total) return;
    const int64_t g = mid[i];
    keep[i] = (g < m || err[g - m] == 0) ? 1 : 0;
}

// 64 threads per root row copy (21 complex values)
__global__ void scatter_kernel(const double2 *coll, int64_t m, const double2 *fpt, const int64_t *mid,
                               const int32_t *keep, const int32_t *pos, int64_t total, double2 *out) {
    const int64_t i = (int64_t)blockIdx.x * (blockDim.x / 32) + threadIdx.x / 32;
    const int j = threadIdx.x % 32;
    if (i >= total || !keep[i] || j >= kNV) return;
    const int64_t g = mid[i];
    const double2 *src = g < m ? coll + g * kNV : fpt + (g - m) * kNV;
    out[(int64_t)pos[i] * kNV + j] = src[j];
}

inline unsigned blocks(int64_t n, int t) { return (unsigned)((n + t - 1) / t); }

}  // namespace

size_t dedupe_temp_bytes(int64_t m, int64_t n) {
    const int64_t total = m + n;
    size_t sort_bytes = 0, scan_bytes = 0;
    (void)rocprim::radix_sort_pairs(nullptr, sort_bytes, (const uint64_t *)nullptr, (uint64_t *)nullptr,
                                    (const int64_t *)nullptr, (int64_t *)nullptr, (size_t)(n > 0 ? n : 1));
    (void)rocprim::exclusive_scan(nullptr, scan_bytes, (const int32_t *)nullptr, (int32_t *)nullptr, 0,
                                  (size_t)(total > 0 ? total : 1), rocprim::plus<int32_t>());
    const size_t a = 256;
    auto up = [&](size_t x) { return (x + a - 1) / a * a; };
    return up(sizeof(uint64_t) * n) * 2 + up(sizeof(int64_t) * n) * 3 + up(sizeof(int64_t) * total) * 2 +
           up(sizeof(double) * total) + up(sizeof(int32_t) * total) * 3 + up(sizeof(int64_t)) +
           up(sizeof(int32_t)) + up(std::max(sort_bytes, scan_bytes)) + 4 * a;
}

hipError_t sort_dedupe_merge_device(const double2 *coll, int64_t m, const double2 *fpt, int32_t *err, int64_t n,
                                    double tol, double2 *out, int64_t *m_out, void *temp, size_t temp_bytes,
                                    hipStream_t st) {
    const int64_t total = m + n;
    *m_out = 0;
    if (total == 0) return hipSuccess;
    char *p = static_cast<char *>(temp);
    const size_t a = 256;
    auto take = [&](size_t bytes) {
        char *r = p;
        p += (bytes + a - 1) / a * a;
        return r;
    };
    uint64_t *keys = (uint64_t *)take(sizeof(uint64_t) * n), *skeys = (uint64_t *)take(sizeof(uint64_t) * n);
    int64_t *idx = (int64_t *)take(sizeof(int64_t) * n), *sidx = (int64_t *)take(sizeof(int64_t) * n);
    int64_t *rank = (int64_t *)take(sizeof(int64_t) * n);
    int64_t *mid = (int64_t *)take(sizeof(int64_t) * total), *start = (int64_t *)take(sizeof(int64_t) * total);
    double *mkey = (double *)take(sizeof(double) * total);
    int32_t *brk = (int32_t *)take(sizeof(int32_t) * total), *rid = (int32_t *)take(sizeof(int32_t) * total);
    int32_t *keep = (int32_t *)take(sizeof(int32_t) * total);
    int64_t *m_f = (int64_t *)take(sizeof(int64_t));
    int32_t *flag = (int32_t *)take(sizeof(int32_t));
    void *cub_tmp = p;
    const size_t used = (size_t)(p - static_cast<char *>(temp));
    if (used > temp_bytes) return hipErrorInvalidValue;
    size_t cub_bytes = temp_bytes - used;
    const int T = 256;
    hipError_t e;

    hipLaunchKernelGGL(finite_old_kernel, dim3(1), dim3(1), 0, st, coll, m, m_f);
    if (n > 0) {
        hipLaunchKernelGGL(keys_kernel, dim3(blocks(n, T)), dim3(T), 0, st, fpt, n, keys, idx);
        size_t sb = cub_bytes;
        e = rocprim::radix_sort_pairs(cub_tmp, sb, keys, skeys, idx, sidx, (size_t)n, 0, 64, st);
        if (e != hipSuccess) return e;
        hipLaunchKernelGGL(rank_new_kernel, dim3(blocks(n, T)), dim3(T), 0, st, coll, m, m_f, fpt, sidx, n, mid,
                           mkey);
    }
    if (m > 0)
        hipLaunchKernelGGL(rank_old_kernel, dim3(blocks(m, T)), dim3(T), 0, st, coll, m, m_f, fpt, sidx, n, mid,
                           mkey);
    int32_t nan_flag = 0;
    if (n > 0) {
        hipLaunchKernelGGL(zero_flag_kernel, dim3(1), dim3(1), 0, st, flag);
        hipLaunchKernelGGL(nan_check_kernel, dim3(blocks(n, T)), dim3(T), 0, st, fpt, err, n, flag);
        e = hipMemcpyAsync(&nan_flag, flag, sizeof(int32_t), hipMemcpyDeviceToHost, st);
        if (e != hipSuccess) return e;
        e = hipStreamSynchronize(st);
        if (e != hipSuccess) return e;
    }
    if (n > 0 && nan_flag) {
        hipLaunchKernelGGL(new_rank_kernel, dim3(blocks(total, T)), dim3(T), 0, st, mid, m, total, rank);
        hipLaunchKernelGGL(dedupe_serial_kernel, dim3(1), dim3(1), 0, st, coll, m, fpt, err, mid, mkey, rank, n,
                           total, tol);
    } else if (n > 0) {
        hipLaunchKernelGGL(breaks_kernel, dim3(blocks(total, T)), dim3(T), 0, st, mkey, total, tol, brk);
        size_t sb = cub_bytes;
        e = rocprim::exclusive_scan(cub_tmp, sb, brk, rid, 0, (size_t)total, rocprim::plus<int32_t>(), st);
        if (e != hipSuccess) return e;
        hipLaunchKernelGGL(run_starts_kernel, dim3(blocks(total, T)), dim3(T), 0, st, brk, rid, total, start);
        // the new roots of each run in index order: (run, index) keys sorted, reusing the buffers
        // of the first sort (keys, idx -> skeys, rank)
        hipLaunchKernelGGL(run_order_kernel, dim3(blocks(total, T)), dim3(T), 0, st, mid, m, brk, rid, total, keys,
                           idx);
        sb = cub_bytes;
        e = rocprim::radix_sort_pairs(cub_tmp, sb, keys, skeys, idx, rank, (size_t)n, 0, 64, st);
        if (e != hipSuccess) return e;
        int32_t last_rid = 0, last_brk = 0;
        e = hipMemcpyAsync(&last_rid, rid + (total - 1), sizeof(int32_t), hipMemcpyDeviceToHost, st);
        if (e == hipSuccess) e = hipMemcpyAsync(&last_brk, brk + (total - 1), sizeof(int32_t), hipMemcpyDeviceToHost, st);
        if (e == hipSuccess) e = hipStreamSynchronize(st);
        if (e != hipSuccess) return e;
        const int64_t nruns = (int64_t)last_rid + last_brk;
        hipLaunchKernelGGL(dedupe_runs_kernel, dim3(blocks(n, 64)), dim3(64), 0, st, coll, m, fpt, err, mid, mkey,
                           start, nruns, total, skeys, rank, n, tol);
    }
    hipLaunchKernelGGL(keep_kernel, dim3(blocks(total, T)), dim3(T), 0, st, mid, m, err, total, keep);
    size_t sb = cub_bytes;
    e = rocprim::exclusive_scan(cub_tmp, sb, keep, rid, 0, (size_t)total, rocprim::plus<int32_t>(), st);
    if (e != hipSuccess) return e;
    hipLaunchKernelGGL(scatter_kernel, dim3(blocks(total, 8)), dim3(256), 0, st, coll, m, fpt, mid, keep, rid, total,
                       out);
    int32_t last_pos = 0, last_keep = 0;
    e = hipMemcpyAsync(&last_pos, rid + (total - 1), sizeof(int32_t), hipMemcpyDeviceToHost, st);
    if (e == hipSuccess) e = hipMemcpyAsync(&last_keep, keep + (total - 1), sizeof(int32_t), hipMemcpyDeviceToHost, st);
    if (e == hipSuccess) e = hipStreamSynchronize(st);
    if (e != hipSuccess) return e;
    *m_out = (int64_t)last_pos + last_keep;
    return hipGetLastError();
}

}  // namespace frg
