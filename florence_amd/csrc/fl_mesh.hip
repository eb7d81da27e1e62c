// High-order meshes from linear ones on the device, numbered as the reference numbers
// them: GetHighOrderMesh (Florence/MeshGeneration/Mesh.py:1339-1409) through
// HighOrderMeshHex / HighOrderMeshTet_SEMISTABLE (HigherOrderMeshing/
// HigherOrderMeshingHex.py:19-133, HigherOrderMeshingTet.py:19-133) with Decimals=10 and
// check_duplicates=True:
//
//  1. candidates: the linear nodes, then every element's non-vertex nodes at
//     Neval^T x_vertices, element by element (GetInteriorCoordinates.py:3-13);
//  2. the new candidates are merged on their coordinates rounded to `decimals`
//     (makezero at 1e-14, np.round = rint(x 10^d) / 10^d); numpy's unique of the rows as
//     raw bytes orders them by memcmp of the little-endian doubles, i.e. by the
//     byte-swapped 64-bit words (x, y, z) as unsigned integers, keeping the first
//     occurrence (unique2d, Tensors.py:15-74): the merged list is the linear nodes
//     followed by the merged new nodes in that order;
//  3. remove_duplicates_2D (Tensors.py:428-441) repeats the merge over that whole list
//     at decimals - 2 (after makezero in place), which gives every node its final number:
//     the rank of its rounded coordinates in the same byte order.
//
// Both merges are a sort of (key words, position) -- a bitonic sort over a power-of-two
// padding, so ties keep the first position -- a flag per run start and a scan.
#include <cstdint>

#include "fl_common.hpp"

namespace fl {
namespace {

constexpr int MT = 256;

struct Key4 {
    uint64_t w0, w1, w2;   // byte-swapped bits of the rounded x, y, z
    uint64_t pos;          // position in the list being merged (tie-break: first kept)
};

__device__ inline bool key_less(const Key4 &a, const Key4 &b) {
    if (a.w0 != b.w0) return a.w0 < b.w0;
    if (a.w1 != b.w1) return a.w1 < b.w1;
    if (a.w2 != b.w2) return a.w2 < b.w2;
    return a.pos < b.pos;
}

__device__ inline bool key_same(const Key4 &a, const Key4 &b) {
    return a.w0 == b.w0 && a.w1 == b.w1 && a.w2 == b.w2;
}

// np.round(makezero(x), d) as its bytes, byte-swapped: memcmp order as an integer
__device__ inline uint64_t round_key(double x, double scale) {
    if (fabs(x) < 1e-14) x = 0.0;
    const double r = rint(x * scale) / scale;
    return __builtin_bswap64((uint64_t)__double_as_longlong(r));
}

__device__ inline double zero_small(double x) { return fabs(x) < 1e-14 ? 0.0 : x; }

// coordinates of the candidates: linear nodes, then element-major non-vertex nodes
__global__ __launch_bounds__(MT) void candidates_kernel(const double *__restrict__ X, int64_t nlin,
                                                        const int32_t *__restrict__ el, int64_t nelem, int nv,
                                                        int npe, const double *__restrict__ Neval,
                                                        double *__restrict__ cand) {
    const int64_t nnew = npe - nv;
    const int64_t i = (int64_t)blockIdx.x * MT + threadIdx.x;
    if (i >= nlin + nelem * nnew) return;
    if (i < nlin) {
#pragma unroll
        for (int d = 0; d < 3; ++d) cand[3 * i + d] = X[3 * i + d];
        return;
    }
    const int64_t e = (i - nlin) / nnew;
    const int l = nv + (int)((i - nlin) - e * nnew);
    double s[3] = {0.0, 0.0, 0.0};
    for (int v = 0; v < nv; ++v) {
        const double w = Neval[v * npe + l];
        const int64_t nd = el[e * nv + v];
#pragma unroll
        for (int d = 0; d < 3; ++d) s[d] += w * X[3 * nd + d];
    }
#pragma unroll
    for (int d = 0; d < 3; ++d) cand[3 * i + d] = s[d];
}

// keys of n points (coordinates src[3 * (first + t)]) at scale 10^d; padding keys sort last
__global__ __launch_bounds__(MT) void keys_kernel(const double *__restrict__ src, int64_t first, int64_t n,
                                                  int64_t npad, double scale, Key4 *__restrict__ keys) {
    const int64_t t = (int64_t)blockIdx.x * MT + threadIdx.x;
    if (t >= npad) return;
    Key4 k;
    if (t < n) {
        const double *c = src + 3 * (first + t);
        k.w0 = round_key(c[0], scale);
        k.w1 = round_key(c[1], scale);
        k.w2 = round_key(c[2], scale);
        k.pos = (uint64_t)t;
    } else {
        k.w0 = k.w1 = k.w2 = ~0ull;
        k.pos = ~0ull;
    }
    keys[t] = k;
}

// one compare-exchange pass of the bitonic network (stage k, distance j)
__global__ __launch_bounds__(MT) void bitonic_kernel(Key4 *__restrict__ keys, int64_t npad, int64_t j, int64_t k) {
    const int64_t i = (int64_t)blockIdx.x * MT + threadIdx.x;
    if (i >= npad) return;
    const int64_t l = i ^ j;
    if (l <= i) return;
    const Key4 a = keys[i], b = keys[l];
    const bool up = (i & k) == 0;
    if (up ? key_less(b, a) : key_less(a, b)) {
        keys[i] = b;
        keys[l] = a;
    }
}

// run starts of the sorted keys (n real ones) as 0/1
__global__ __launch_bounds__(MT) void run_flags_kernel(const Key4 *__restrict__ keys, int64_t n,
                                                       int64_t *__restrict__ flag) {
    const int64_t t = (int64_t)blockIdx.x * MT + threadIdx.x;
    if (t >= n) return;
    flag[t] = (t == 0 || !key_same(keys[t], keys[t - 1])) ? 1 : 0;
}

// after the exclusive scan of the flags: rank of every sorted entry, the first position
// of every run (its representative) and the rank of every position
__global__ __launch_bounds__(MT) void run_ranks_kernel(const Key4 *__restrict__ keys, int64_t n,
                                                       const int64_t *__restrict__ flag_scan,
                                                       const int64_t *__restrict__ flag_raw,
                                                       int64_t *__restrict__ rep, int64_t *__restrict__ rank_of_pos) {
    const int64_t t = (int64_t)blockIdx.x * MT + threadIdx.x;
    if (t >= n) return;
    const int64_t r = flag_scan[t] + flag_raw[t] - 1;
    const int64_t pos = (int64_t)keys[t].pos;
    rank_of_pos[pos] = r;
    if (flag_raw[t]) rep[r] = pos;
}

// the merged list of step 2: linear nodes, then the merged new nodes (coordinates of
// their first occurrence), made zero below 1e-14 as remove_duplicates_2D does in place
__global__ __launch_bounds__(MT) void merged_list_kernel(const double *__restrict__ cand, int64_t nlin,
                                                         const int64_t *__restrict__ rep_new, int64_t nuniq,
                                                         double *__restrict__ list) {
    const int64_t p = (int64_t)blockIdx.x * MT + threadIdx.x;
    if (p >= nlin + nuniq) return;
    const int64_t src = p < nlin ? p : nlin + rep_new[p - nlin];
#pragma unroll
    for (int d = 0; d < 3; ++d) list[3 * p + d] = zero_small(cand[3 * src + d]);
}

__global__ __launch_bounds__(MT) void final_points_kernel(const double *__restrict__ list,
                                                          const int64_t *__restrict__ rep, int64_t nfinal,
                                                          double *__restrict__ points) {
    const int64_t r = (int64_t)blockIdx.x * MT + threadIdx.x;
    if (r >= nfinal) return;
#pragma unroll
    for (int d = 0; d < 3; ++d) points[3 * r + d] = list[3 * rep[r] + d];
}

__global__ __launch_bounds__(MT) void final_elements_kernel(const int32_t *__restrict__ el, int64_t nelem, int nv,
                                                            int npe, int64_t nlin,
                                                            const int64_t *__restrict__ rank_new,
                                                            const int64_t *__restrict__ rank_final,
                                                            int32_t *__restrict__ out) {
    const int64_t t = (int64_t)blockIdx.x * MT + threadIdx.x;
    if (t >= nelem * npe) return;
    const int64_t e = t / npe;
    const int l = (int)(t - e * npe);
    const int64_t p = l < nv ? (int64_t)el[e * nv + l] : nlin + rank_new[e * (npe - nv) + (l - nv)];
    out[t] = (int32_t)rank_final[p];
}

int64_t pow2_at_least(int64_t n) {
    int64_t p = 1;
    while (p < n) p <<= 1;
    return p;
}

int bitonic_sort(Key4 *keys, int64_t npad, hipStream_t st) {
    const unsigned nb = (unsigned)((npad + MT - 1) / MT);
    for (int64_t k = 2; k <= npad; k <<= 1)
        for (int64_t j = k >> 1; j > 0; j >>= 1) {
            bitonic_kernel<<<nb, MT, 0, st>>>(keys, npad, j, k);
            FL_LAUNCH_CHECK();
        }
    return 0;
}

struct RefineWork {
    double *cand, *list;
    Key4 *keys;
    int64_t *flag, *scan, *rep_new, *rank_new, *rep, *rank_final, *scan_scratch;
    size_t scan_count;
};

RefineWork carve_refine(void *ws, size_t bytes, int64_t nlin, int64_t nelem, int nv, int npe, size_t *need,
                        bool *ok) {
    const int64_t nnew = nelem * (int64_t)(npe - nv), M = nlin + nnew;
    const int64_t npad = pow2_at_least(M > 1 ? M : 2);
    Carve cv(ws, bytes);
    RefineWork w;
    w.cand = cv.take<double>(3 * (size_t)M);
    w.list = cv.take<double>(3 * (size_t)M);
    w.keys = cv.take<Key4>((size_t)npad);
    w.flag = cv.take<int64_t>((size_t)M + 1);
    w.scan = cv.take<int64_t>((size_t)M + 1);
    w.rep_new = cv.take<int64_t>((size_t)M);
    w.rank_new = cv.take<int64_t>((size_t)M);
    w.rep = cv.take<int64_t>((size_t)M);
    w.rank_final = cv.take<int64_t>((size_t)M);
    w.scan_count = scan_scratch_count(M + 1);
    w.scan_scratch = cv.take<int64_t>(w.scan_count);
    *need = cv.off + 256;
    *ok = cv.ok;
    return w;
}

// sort n keys of src[first ..] at 10^d; ranks of the positions, representatives, count
int merge_points(RefineWork &w, const double *src, int64_t first, int64_t n, int decimals, int64_t *rank_of_pos,
                 int64_t *rep, int64_t *count, hipStream_t st) {
    const int64_t npad = pow2_at_least(n > 1 ? n : 2);
    double scale = 1.0;
    for (int i = 0; i < decimals; ++i) scale *= 10.0;
    keys_kernel<<<(unsigned)((npad + MT - 1) / MT), MT, 0, st>>>(src, first, n, npad, scale, w.keys);
    FL_LAUNCH_CHECK();
    int rc = bitonic_sort(w.keys, npad, st);
    if (rc) return rc;
    const unsigned nb = (unsigned)((n + MT - 1) / MT);
    run_flags_kernel<<<nb, MT, 0, st>>>(w.keys, n, w.flag);
    FL_LAUNCH_CHECK();
    FL_HIP_CHECK(hipMemcpyAsync(w.scan, w.flag, sizeof(int64_t) * n, hipMemcpyDeviceToDevice, st));
    rc = exclusive_scan_i64(w.scan, n, w.scan_scratch, w.scan_count, st);
    if (rc) return rc;
    run_ranks_kernel<<<nb, MT, 0, st>>>(w.keys, n, w.scan, w.flag, rep, rank_of_pos);
    FL_LAUNCH_CHECK();
    int64_t last[2] = {0, 0};
    FL_HIP_CHECK(hipMemcpyAsync(&last[0], w.scan + n - 1, sizeof(int64_t), hipMemcpyDeviceToHost, st));
    FL_HIP_CHECK(hipMemcpyAsync(&last[1], w.flag + n - 1, sizeof(int64_t), hipMemcpyDeviceToHost, st));
    FL_HIP_CHECK(hipStreamSynchronize(st));
    *count = last[0] + last[1];
    return 0;
}

}  // namespace
}  // namespace fl

using namespace fl;

extern "C" {

size_t fl_mesh_refine_workspace_size(int64_t nnode_lin, int64_t nelem, int32_t nv, int32_t npe) {
    if (nnode_lin <= 0 || nelem <= 0 || nv <= 0 || npe <= nv) return 0;
    size_t need = 0;
    bool ok = false;
    carve_refine(nullptr, 0, nnode_lin, nelem, nv, npe, &need, &ok);
    return need;
}

int fl_mesh_refine(const double *points_lin, int64_t nnode_lin, const int32_t *elements_lin, int64_t nelem,
                   int32_t nv, const double *Neval, int32_t npe, int32_t decimals, double *points,
                   int32_t *elements, int64_t *nnode, void *workspace, size_t workspace_bytes, void *stream) {
    hipStream_t st = as_stream(stream);
    if (!points_lin || !elements_lin || !Neval || !points || !elements || !nnode || nnode_lin <= 0 || nelem <= 0 ||
        nv <= 0 || npe <= nv || decimals < 1 || decimals > 15) {
        set_error("fl_mesh_refine: invalid argument");
        return FL_E_ARG;
    }
    const int64_t nnew = nelem * (int64_t)(npe - nv), M = nnode_lin + nnew;
    if (M >= INT32_MAX || nelem * (int64_t)npe >= INT32_MAX) {
        set_error("fl_mesh_refine: %lld candidate nodes exceed int32 ids", (long long)M);
        return FL_E_LIMIT;
    }
    size_t need = 0;
    bool ok = false;
    RefineWork w = carve_refine(workspace, workspace_bytes, nnode_lin, nelem, nv, npe, &need, &ok);
    if (!ok) {
        set_error("fl_mesh_refine: workspace too small (%zu < %zu)", workspace_bytes, need);
        return FL_E_SPACE;
    }
    candidates_kernel<<<(unsigned)((M + MT - 1) / MT), MT, 0, st>>>(points_lin, nnode_lin, elements_lin, nelem, nv,
                                                                     npe, Neval, w.cand);
    FL_LAUNCH_CHECK();
    // step 2: merge the new nodes at `decimals`
    int64_t nuniq = 0;
    int rc = merge_points(w, w.cand, nnode_lin, nnew, decimals, w.rank_new, w.rep_new, &nuniq, st);
    if (rc) return rc;
    // step 3: merge the whole list at decimals - 2 (decimals when below 6)
    const int64_t nlist = nnode_lin + nuniq;
    merged_list_kernel<<<(unsigned)((nlist + MT - 1) / MT), MT, 0, st>>>(w.cand, nnode_lin, w.rep_new, nuniq,
                                                                         w.list);
    FL_LAUNCH_CHECK();
    int64_t nfinal = 0;
    rc = merge_points(w, w.list, 0, nlist, decimals >= 6 ? decimals - 2 : decimals, w.rank_final, w.rep, &nfinal,
                      st);
    if (rc) return rc;
    final_points_kernel<<<(unsigned)((nfinal + MT - 1) / MT), MT, 0, st>>>(w.list, w.rep, nfinal, points);
    final_elements_kernel<<<(unsigned)((nelem * npe + MT - 1) / MT), MT, 0, st>>>(
        elements_lin, nelem, nv, npe, nnode_lin, w.rank_new, w.rank_final, elements);
    FL_LAUNCH_CHECK();
    FL_HIP_CHECK(hipStreamSynchronize(st));
    *nnode = nfinal;
    return 0;
}

}  // extern "C"
