// Sparsity pattern of the displacement-formulation tangent, built on the device.
//
// Same pattern as ComputeSparsityPattern (Florence/FiniteElements/Assembly/_Assembly_/
// ComputeSparsityPattern.pyx:44-90, ComputeSparsityPattern.h:22-62): for node a the
// sorted unique set of nodes sharing an element with it; its three rows 3a+i list the
// columns 3b+k of every neighbour b in ascending order.  The reference builds this per
// node with std::sort/std::unique on the host; here one workgroup per node gathers the
// candidate nodes of its incident elements into LDS, bitonic-sorts them and compacts.
#include <cstdarg>
#include <climits>
#include <algorithm>
#include <vector>

#include "fl_common.hpp"

namespace fl {

static thread_local std::string g_last_error;

void set_error(const char *fmt, ...) {
    char buf[1024];
    va_list ap;
    va_start(ap, fmt);
    vsnprintf(buf, sizeof(buf), fmt, ap);
    va_end(ap);
    g_last_error = buf;
}

// ------------------------------------------------------------------ prefix sum
constexpr int SCAN_T = 256;
constexpr int SCAN_I = 8;
constexpr int SCAN_B = SCAN_T * SCAN_I;

__global__ __launch_bounds__(SCAN_T) void scan_blocks_kernel(int64_t *d, int64_t n, int64_t *sums) {
    __shared__ int64_t sh[SCAN_T];
    const int t = threadIdx.x;
    const int64_t base = (int64_t)blockIdx.x * SCAN_B + (int64_t)t * SCAN_I;
    int64_t v[SCAN_I];
    int64_t s = 0;
#pragma unroll
    for (int i = 0; i < SCAN_I; ++i) {
        const int64_t idx = base + i;
        v[i] = idx < n ? d[idx] : 0;
        s += v[i];
    }
    sh[t] = s;
    __syncthreads();
    for (int off = 1; off < SCAN_T; off <<= 1) {
        int64_t x = t >= off ? sh[t - off] : 0;
        __syncthreads();
        sh[t] += x;
        __syncthreads();
    }
    int64_t run = sh[t] - s;
#pragma unroll
    for (int i = 0; i < SCAN_I; ++i) {
        const int64_t idx = base + i;
        if (idx < n) d[idx] = run;
        run += v[i];
    }
    if (t == SCAN_T - 1 && sums != nullptr) sums[blockIdx.x] = sh[t];
}

__global__ void scan_add_kernel(int64_t *d, int64_t n, const int64_t *sums) {
    const int64_t idx = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (idx < n) d[idx] += sums[idx / SCAN_B];
}

size_t scan_scratch_count(int64_t n) {
    size_t total = 0;
    int64_t m = n;
    while (m > SCAN_B) {
        m = (m + SCAN_B - 1) / SCAN_B;
        total += (size_t)m + 1;
    }
    return total + 1;
}

int exclusive_scan_i64(int64_t *data, int64_t n, int64_t *scratch, size_t scratch_count, hipStream_t st) {
    if (n <= 0) return 0;
    const int64_t nb = (n + SCAN_B - 1) / SCAN_B;
    if (nb == 1) {
        scan_blocks_kernel<<<1, SCAN_T, 0, st>>>(data, n, nullptr);
        FL_LAUNCH_CHECK();
        return 0;
    }
    if ((size_t)nb + 1 > scratch_count) {
        set_error("scan scratch too small");
        return FL_E_SPACE;
    }
    scan_blocks_kernel<<<(unsigned)nb, SCAN_T, 0, st>>>(data, n, scratch);
    FL_LAUNCH_CHECK();
    int rc = exclusive_scan_i64(scratch, nb, scratch + nb + 1, scratch_count - (size_t)nb - 1, st);
    if (rc) return rc;
    scan_add_kernel<<<(unsigned)((n + 255) / 256), 256, 0, st>>>(data, n, scratch);
    FL_LAUNCH_CHECK();
    return 0;
}

// ------------------------------------------------------- node -> element map
__global__ void count_incidence_kernel(const int32_t *__restrict__ elements, int64_t nentries, int64_t nnode,
                                       int64_t *__restrict__ counts, int32_t *__restrict__ bad) {
    const int64_t t = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (t >= nentries) return;
    const int32_t nd = elements[t];
    if (nd < 0 || nd >= nnode) {
        atomicOr(bad, 1);
        return;
    }
    atomicAdd(reinterpret_cast<unsigned long long *>(counts + nd), 1ull);
}

__global__ void fill_incidence_kernel(const int32_t *__restrict__ elements, int64_t nentries,
                                      const int64_t *__restrict__ offsets, int32_t *__restrict__ cursor,
                                      int32_t *__restrict__ node_elem) {
    const int64_t t = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (t >= nentries) return;
    const int32_t nd = elements[t];
    const int32_t slot = atomicAdd(cursor + nd, 1);
    node_elem[offsets[nd] + slot] = (int32_t)t;   // t = e*npe + local
}

// Sort each node's incidence list (ascending element, hence ascending e*npe+local)
// and emit the int32 pointer array.  Lists are short (<= a few dozen entries).
__global__ void sort_incidence_kernel(const int64_t *__restrict__ offsets, int64_t nnode,
                                      int32_t *__restrict__ node_elem, int32_t *__restrict__ node_elem_ptr) {
    const int64_t a = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (a > nnode) return;
    node_elem_ptr[a] = (int32_t)offsets[a];
    if (a == nnode) return;
    const int64_t b = offsets[a], e = offsets[a + 1];
    for (int64_t i = b + 1; i < e; ++i) {
        const int32_t v = node_elem[i];
        int64_t j = i - 1;
        while (j >= b && node_elem[j] > v) {
            node_elem[j + 1] = node_elem[j];
            --j;
        }
        node_elem[j + 1] = v;
    }
}

// ----------------------------------------------------------- neighbour lists
constexpr int NB_T = 256;
constexpr int NB_MAX = 8192;   // candidate nodes per node (incident elements x npe)

// Gather candidates of node a into LDS, sort, mark uniques.  Returns unique count.
__device__ int gather_sort_unique(const int32_t *__restrict__ elements, int32_t npe,
                                  const int32_t *__restrict__ node_elem_ptr,
                                  const int32_t *__restrict__ node_elem, int64_t a, int32_t *cand,
                                  int32_t *flag_sum, int32_t *overflow) {
    const int t = threadIdx.x;
    const int32_t b0 = node_elem_ptr[a], b1 = node_elem_ptr[a + 1];
    const int ncand = (b1 - b0) * npe;
    if (ncand > NB_MAX) {
        if (t == 0) atomicOr(overflow, 1);
        return -1;
    }
    int S = 64;
    while (S < ncand) S <<= 1;
    for (int i = t; i < S; i += NB_T) {
        int32_t v = INT_MAX;
        if (i < ncand) {
            const int32_t ent = node_elem[b0 + i / npe];
            const int32_t e = ent / npe;
            v = elements[(int64_t)e * npe + i % npe];
        }
        cand[i] = v;
    }
    __syncthreads();
    // bitonic sort, ascending
    for (int k = 2; k <= S; k <<= 1) {
        for (int j = k >> 1; j > 0; j >>= 1) {
            for (int i = t; i < S; i += NB_T) {
                const int p = i ^ j;
                if (p > i) {
                    const int32_t x = cand[i], y = cand[p];
                    const bool up = (i & k) == 0;
                    if ((x > y) == up) {
                        cand[i] = y;
                        cand[p] = x;
                    }
                }
            }
            __syncthreads();
        }
    }
    // per-thread contiguous chunk unique counts, then exclusive scan in flag_sum
    const int per = (S + NB_T - 1) / NB_T;
    const int lo = t * per, hi = min(lo + per, ncand);
    int c = 0;
    for (int i = lo; i < hi; ++i) c += (i == 0 || cand[i] != cand[i - 1]) ? 1 : 0;
    flag_sum[t] = c;
    __syncthreads();
    for (int off = 1; off < NB_T; off <<= 1) {
        int x = t >= off ? flag_sum[t - off] : 0;
        __syncthreads();
        flag_sum[t] += x;
        __syncthreads();
    }
    return flag_sum[NB_T - 1];
}

__global__ __launch_bounds__(NB_T) void count_neighbours_kernel(
    const int32_t *__restrict__ elements, int32_t npe, const int32_t *__restrict__ node_elem_ptr,
    const int32_t *__restrict__ node_elem, int64_t nnode, int64_t *__restrict__ nbr_count,
    int32_t *__restrict__ max_degree, int32_t *__restrict__ overflow) {
    __shared__ int32_t cand[NB_MAX];
    __shared__ int32_t flag_sum[NB_T];
    const int64_t a = blockIdx.x;
    const int cnt = gather_sort_unique(elements, npe, node_elem_ptr, node_elem, a, cand, flag_sum, overflow);
    if (threadIdx.x == 0) {
        nbr_count[a] = cnt < 0 ? 0 : cnt;
        if (cnt > 0) atomicMax(max_degree, cnt);
    }
}

__global__ __launch_bounds__(NB_T) void fill_neighbours_kernel(
    const int32_t *__restrict__ elements, int32_t npe, const int32_t *__restrict__ node_elem_ptr,
    const int32_t *__restrict__ node_elem, int64_t nnode, const int32_t *__restrict__ node_nbr_ptr,
    int32_t *__restrict__ node_nbrs, int32_t *__restrict__ indptr, int32_t *__restrict__ indices,
    int32_t *__restrict__ overflow) {
    __shared__ int32_t cand[NB_MAX];
    __shared__ int32_t flag_sum[NB_T];
    const int64_t a = blockIdx.x;
    const int t = threadIdx.x;
    const int cnt = gather_sort_unique(elements, npe, node_elem_ptr, node_elem, a, cand, flag_sum, overflow);
    if (cnt < 0) return;
    const int32_t p0 = node_nbr_ptr[a];
    if (cnt != node_nbr_ptr[a + 1] - p0) {
        if (t == 0) atomicOr(overflow, 2);
        return;
    }
    const int32_t b0 = node_elem_ptr[a], b1 = node_elem_ptr[a + 1];
    const int ncand = (b1 - b0) * npe;
    int S = 64;
    while (S < ncand) S <<= 1;
    const int per = (S + NB_T - 1) / NB_T;
    const int lo = t * per, hi = min(lo + per, ncand);
    int r = t == 0 ? 0 : flag_sum[t - 1];
    for (int i = lo; i < hi; ++i) {
        if (i == 0 || cand[i] != cand[i - 1]) {
            node_nbrs[p0 + r] = cand[i];
            ++r;
        }
    }
    __syncthreads();
    // CSR rows 3a+i: indptr = 9*p0 + i*3*cnt; columns 3b+k
    const int64_t rowlen = 3 * (int64_t)cnt;
    if (t < 3) indptr[3 * a + t] = (int32_t)(9 * (int64_t)p0 + t * rowlen);
    if (a == nnode - 1 && t == 0) indptr[3 * nnode] = (int32_t)(9 * (int64_t)node_nbr_ptr[nnode]);
    if (indices != nullptr) {
        for (int64_t q = t; q < 3 * rowlen; q += NB_T) {
            const int64_t i = q / rowlen, c = q % rowlen;
            const int32_t nb = node_nbrs[p0 + c / 3];
            indices[9 * (int64_t)p0 + i * rowlen + c] = 3 * nb + (int32_t)(c % 3);
        }
    }
}

__global__ void i64_to_i32_kernel(const int64_t *__restrict__ src, int32_t *__restrict__ dst, int64_t n) {
    const int64_t t = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (t < n) dst[t] = (int32_t)src[t];
}

// ranks[(e npe + la) npe + lb] = position of node elements[e][lb] in the sorted
// neighbour list of node elements[e][la] (the CSR column block of that entry); status
// bit FL_S_PATTERN if a node is missing or a list exceeds the int16 range.
__global__ void rank_table_kernel(const int32_t *__restrict__ elements, int64_t nelem, int npe,
                                  const int32_t *__restrict__ nbr_ptr, const int32_t *__restrict__ nbrs,
                                  int16_t *__restrict__ ranks, int32_t *__restrict__ status) {
    const int64_t t = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    const int64_t npe2 = (int64_t)npe * npe;
    if (t >= nelem * npe2) return;
    const int64_t e = t / npe2;
    const int la = (int)((t - e * npe2) / npe), lb = (int)(t - e * npe2 - (int64_t)la * npe);
    const int32_t a = elements[e * npe + la], b = elements[e * npe + lb];
    const int32_t p0 = nbr_ptr[a], deg = nbr_ptr[a + 1] - p0;
    int lo = 0, hi = deg;
    while (hi - lo > 1) {
        const int mid = (lo + hi) >> 1;
        if (b < nbrs[p0 + mid]) hi = mid; else lo = mid;
    }
    const bool ok = deg > 0 && deg <= 32767 && nbrs[p0 + lo] == b;
    if (!ok) atomicOr(status, FL_S_PATTERN);
    ranks[t] = ok ? (int16_t)lo : (int16_t)0;
}

// The direct scatter's table, one entry per slot of sf_scatter_index's order for hexes
// of p1^3 nodes: the position of the column node b in the sorted neighbour list of the
// row node a, with the sign bit set when an element of a lower colour holds both a and b
// (that element has already written the entry); unused slots hold 0.  Tensor nodes map
// to the element's local nodes through local_of_t of the tensor basis.  Status bit
// FL_S_PATTERN on a missing node, an oversized list, a basis of another p1, or two
// elements of one colour that share a node.
__global__ void scatter_table_kernel(const int32_t *__restrict__ elements, int64_t nelem, int p1,
                                     const int32_t *__restrict__ node_elem_ptr, const int32_t *__restrict__ node_elem,
                                     const int32_t *__restrict__ nbr_ptr, const int32_t *__restrict__ nbrs,
                                     const int32_t *__restrict__ colour, const int32_t *__restrict__ basis,
                                     int16_t *__restrict__ table, int32_t *__restrict__ status) {
    const int64_t t = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    const int64_t per = sf_scatter_per_element(p1);
    if (t >= nelem * per) return;
    const int64_t e = t / per;
    int64_t u = t - e * per;
    const int mt = sf_row_tiles(p1), nt = sf_col_tiles(p1);
    const int r = (int)(u % 4);
    u /= 4;
    const int tm = (int)(u % mt);
    u /= mt;
    const int lane = (int)(u % 64);
    u /= 64;
    const int s = (int)(u % nt);
    const int az = (int)(u / nt);
    const int pp = p1 * p1, npe = pp * p1;
    const int row = 16 * tm + (lane >> 4) + 4 * r, col = 16 * s + (lane & 15);
    if (row >= pp || col >= npe) {
        table[t] = 0;
        return;
    }
    const int ab = row / p1, bb = row % p1, ae = col / pp, be = (col / p1) % p1, bz = col % p1;
    const int ta = (ab * p1 + ae) * p1 + az, tb = (bb * p1 + be) * p1 + bz;
    const int32_t *loc = basis + 4 + basis[1] * basis[1] * basis[1];
    const int la = min(max(loc[ta], 0), npe - 1), lb = min(max(loc[tb], 0), npe - 1);
    const int32_t a = elements[e * npe + la], b = elements[e * npe + lb];
    const int32_t p0 = nbr_ptr[a], deg = nbr_ptr[a + 1] - p0;
    int lo = 0, hi = deg;
    while (hi - lo > 1) {
        const int mid = (lo + hi) >> 1;
        if (b < nbrs[p0 + mid]) hi = mid; else lo = mid;
    }
    bool ok = basis[0] == p1 && deg > 0 && deg <= 32767 && nbrs[p0 + lo] == b;
    const int32_t ce = colour[e];
    bool add = false;
    for (int32_t q = node_elem_ptr[a]; q < node_elem_ptr[a + 1]; ++q) {
        const int32_t e2 = node_elem[q] / npe;
        if (e2 != e && colour[e2] == ce) ok = false;
        if (colour[e2] >= ce || add) continue;
        for (int32_t q2 = node_elem_ptr[b]; q2 < node_elem_ptr[b + 1]; ++q2)
            if (node_elem[q2] / npe == e2) add = true;
    }
    if (!ok) atomicOr(status, FL_S_PATTERN);
    table[t] = ok ? (int16_t)(add ? lo - 32768 : lo) : (int16_t)0;
}

}  // namespace fl

using namespace fl;

extern "C" {

int fl_version(void) { return 100; }

const char *fl_last_error(void) { return g_last_error.c_str(); }

size_t fl_pattern_workspace_size(int64_t nnode) {
    // counts/offsets (nnode+1) x2 int64, cursor int32, scan scratch, flags
    return align_up((size_t)(nnode + 1) * 8, 256) * 2 + align_up((size_t)(nnode + 1) * 4, 256) +
           align_up(scan_scratch_count(nnode + 1) * 8, 256) + 1024;
}

int fl_pattern_graph(const int32_t *elements, int64_t nelem, int32_t npe, int64_t nnode, int32_t *node_elem_ptr,
                     int32_t *node_elem, int32_t *node_nbr_ptr, int64_t *node_nnz, int32_t *max_degree,
                     void *workspace, size_t workspace_bytes, void *stream) {
    hipStream_t st = as_stream(stream);
    if (nelem <= 0 || npe <= 0 || nnode <= 0 || !elements || !node_elem_ptr || !node_elem || !node_nbr_ptr ||
        !node_nnz || !max_degree) {
        set_error("fl_pattern_graph: invalid argument");
        return FL_E_ARG;
    }
    const int64_t nent = nelem * (int64_t)npe;
    if (nent >= INT32_MAX) {
        set_error("fl_pattern_graph: nelem*npe exceeds int32");
        return FL_E_LIMIT;
    }
    Carve cv(workspace, workspace_bytes);
    int64_t *counts = cv.take<int64_t>(nnode + 1);
    int64_t *nbrc = cv.take<int64_t>(nnode + 1);
    int32_t *cursor = cv.take<int32_t>(nnode + 1);
    const size_t sc = scan_scratch_count(nnode + 1);
    int64_t *scratch = cv.take<int64_t>(sc);
    int32_t *flags = cv.take<int32_t>(4);   // [0] bad node id, [1] overflow, [2] max degree
    if (!cv.ok) {
        set_error("fl_pattern_graph: workspace too small (%zu < %zu)", workspace_bytes, cv.off);
        return FL_E_SPACE;
    }
    FL_HIP_CHECK(hipMemsetAsync(counts, 0, sizeof(int64_t) * (nnode + 1), st));
    FL_HIP_CHECK(hipMemsetAsync(cursor, 0, sizeof(int32_t) * (nnode + 1), st));
    FL_HIP_CHECK(hipMemsetAsync(flags, 0, sizeof(int32_t) * 4, st));
    const unsigned gb = (unsigned)((nent + 255) / 256);
    count_incidence_kernel<<<gb, 256, 0, st>>>(elements, nent, nnode, counts, flags);
    FL_LAUNCH_CHECK();
    int rc = exclusive_scan_i64(counts, nnode + 1, scratch, sc, st);
    if (rc) return rc;
    fill_incidence_kernel<<<gb, 256, 0, st>>>(elements, nent, counts, cursor, node_elem);
    FL_LAUNCH_CHECK();
    sort_incidence_kernel<<<(unsigned)((nnode + 1 + 255) / 256), 256, 0, st>>>(counts, nnode, node_elem,
                                                                               node_elem_ptr);
    FL_LAUNCH_CHECK();
    FL_HIP_CHECK(hipMemsetAsync(nbrc, 0, sizeof(int64_t) * (nnode + 1), st));
    count_neighbours_kernel<<<(unsigned)nnode, NB_T, 0, st>>>(elements, npe, node_elem_ptr, node_elem, nnode, nbrc,
                                                              flags + 2, flags + 1);
    FL_LAUNCH_CHECK();
    rc = exclusive_scan_i64(nbrc, nnode + 1, scratch, sc, st);
    if (rc) return rc;
    int64_t total = 0;
    int32_t hflags[4] = {0, 0, 0, 0};
    FL_HIP_CHECK(hipMemcpyAsync(&total, nbrc + nnode, sizeof(int64_t), hipMemcpyDeviceToHost, st));
    FL_HIP_CHECK(hipMemcpyAsync(hflags, flags, sizeof(hflags), hipMemcpyDeviceToHost, st));
    FL_HIP_CHECK(hipStreamSynchronize(st));
    if (hflags[0]) {
        set_error("fl_pattern_graph: element connectivity references a node outside [0, nnode)");
        return FL_E_ARG;
    }
    if (hflags[1]) {
        set_error("fl_pattern_graph: a node touches more than %d candidate nodes", NB_MAX);
        return FL_E_LIMIT;
    }
    if (9 * total >= (int64_t)INT32_MAX) {
        set_error("fl_pattern_graph: nnz = %lld exceeds int32 CSR indices", (long long)(9 * total));
        return FL_E_LIMIT;
    }
    i64_to_i32_kernel<<<(unsigned)((nnode + 1 + 255) / 256), 256, 0, st>>>(nbrc, node_nbr_ptr, nnode + 1);
    FL_LAUNCH_CHECK();
    FL_HIP_CHECK(hipStreamSynchronize(st));
    *node_nnz = total;
    *max_degree = hflags[2];
    return 0;
}

int fl_pattern_fill(const int32_t *elements, int64_t nelem, int32_t npe, int64_t nnode, const int32_t *node_elem_ptr,
                    const int32_t *node_elem, const int32_t *node_nbr_ptr, int32_t *node_nbrs, int32_t *indptr,
                    int32_t *indices, void *stream) {
    hipStream_t st = as_stream(stream);
    if (nelem <= 0 || npe <= 0 || nnode <= 0 || !elements || !node_elem_ptr || !node_elem || !node_nbr_ptr ||
        !node_nbrs || !indptr) {
        set_error("fl_pattern_fill: invalid argument");
        return FL_E_ARG;
    }
    int32_t *flag = nullptr;
    FL_HIP_CHECK(hipMallocAsync(reinterpret_cast<void **>(&flag), sizeof(int32_t), st));
    FL_HIP_CHECK(hipMemsetAsync(flag, 0, sizeof(int32_t), st));
    fill_neighbours_kernel<<<(unsigned)nnode, NB_T, 0, st>>>(elements, npe, node_elem_ptr, node_elem, nnode,
                                                             node_nbr_ptr, node_nbrs, indptr, indices, flag);
    FL_LAUNCH_CHECK();
    int32_t h = 0;
    FL_HIP_CHECK(hipMemcpyAsync(&h, flag, sizeof(int32_t), hipMemcpyDeviceToHost, st));
    FL_HIP_CHECK(hipFreeAsync(flag, st));
    FL_HIP_CHECK(hipStreamSynchronize(st));
    if (h) {
        set_error("fl_pattern_fill: neighbour lists inconsistent with fl_pattern_graph (flag %d)", h);
        return FL_E_ARG;
    }
    return 0;
}

size_t fl_pattern_rank_table_bytes(int64_t nelem, int32_t npe) {
    return nelem > 0 && npe > 0 ? sizeof(int16_t) * (size_t)nelem * npe * npe : 0;
}

int fl_pattern_rank_table(const int32_t *elements, int64_t nelem, int32_t npe, const int32_t *node_nbr_ptr,
                          const int32_t *node_nbrs, int16_t *ranks, int32_t *status, void *stream) {
    if (!elements || !node_nbr_ptr || !node_nbrs || !ranks || !status || nelem <= 0 || npe <= 0) {
        set_error("fl_pattern_rank_table: invalid argument");
        return FL_E_ARG;
    }
    const int64_t n = nelem * npe * (int64_t)npe;
    rank_table_kernel<<<(unsigned)((n + 255) / 256), 256, 0, as_stream(stream)>>>(elements, nelem, npe, node_nbr_ptr,
                                                                                 node_nbrs, ranks, status);
    FL_LAUNCH_CHECK();
    return 0;
}

int fl_pattern_colour(const int32_t *elements, int64_t nelem, int32_t npe, int64_t nnode, int32_t max_colours,
                      int32_t *colour_of_elem, int32_t *colour_elems, int32_t *colour_offsets, int32_t *ncolour) {
    if (!elements || !colour_of_elem || !colour_elems || !colour_offsets || !ncolour || nelem <= 0 || npe <= 0 ||
        nnode <= 0 || max_colours < 1 || max_colours > 64 || nelem >= INT32_MAX) {
        set_error("fl_pattern_colour: invalid argument");
        return FL_E_ARG;
    }
    // greedy first fit in element order; used[a] = colours already holding node a
    std::vector<uint64_t> used((size_t)nnode, 0);
    std::vector<int64_t> count(64, 0);
    int nc = 0;
    for (int64_t e = 0; e < nelem; ++e) {
        const int32_t *conn = elements + e * npe;
        uint64_t taken = 0;
        for (int a = 0; a < npe; ++a) {
            if (conn[a] < 0 || conn[a] >= nnode) {
                set_error("fl_pattern_colour: element %lld has node %d outside [0, %lld)", (long long)e, conn[a],
                          (long long)nnode);
                return FL_E_ARG;
            }
            taken |= used[conn[a]];
        }
        int c = 0;
        while (c < 64 && ((taken >> c) & 1)) ++c;
        if (c >= max_colours) {
            set_error("fl_pattern_colour: element %lld needs more than %d colours", (long long)e, max_colours);
            return FL_E_LIMIT;
        }
        colour_of_elem[e] = c;
        ++count[c];
        nc = std::max(nc, c + 1);
        for (int a = 0; a < npe; ++a) used[conn[a]] |= 1ull << c;
    }
    colour_offsets[0] = 0;
    for (int c = 0; c < nc; ++c) colour_offsets[c + 1] = colour_offsets[c] + (int32_t)count[c];
    std::vector<int32_t> cur(colour_offsets, colour_offsets + nc);
    for (int64_t e = 0; e < nelem; ++e) colour_elems[cur[colour_of_elem[e]]++] = (int32_t)e;
    *ncolour = nc;
    return 0;
}

// p1 of a hex of npe = p1^3 nodes (0 otherwise)
static int cube_root(int32_t npe) {
    for (int p1 = 1; p1 * p1 * p1 <= npe; ++p1)
        if (p1 * p1 * p1 == npe) return p1;
    return 0;
}

size_t fl_pattern_scatter_table_bytes(int64_t nelem, int32_t npe) {
    const int p1 = npe > 0 ? cube_root(npe) : 0;
    return nelem > 0 && p1 > 0 ? sizeof(int16_t) * (size_t)(nelem * sf_scatter_per_element(p1)) : 0;
}

int fl_pattern_scatter_table(const int32_t *elements, int64_t nelem, int32_t npe, const int32_t *node_elem_ptr,
                             const int32_t *node_elem, const int32_t *node_nbr_ptr, const int32_t *node_nbrs,
                             const int32_t *colour_of_elem, const void *tensor_basis, int16_t *table, int32_t *status,
                             void *stream) {
    const int p1 = npe > 0 ? cube_root(npe) : 0;
    if (!elements || !node_elem_ptr || !node_elem || !node_nbr_ptr || !node_nbrs || !colour_of_elem ||
        !tensor_basis || !table || !status || nelem <= 0 || p1 <= 0) {
        set_error("fl_pattern_scatter_table: invalid argument (hexes of p1^3 nodes and their tensor basis)");
        return FL_E_ARG;
    }
    const int64_t n = nelem * sf_scatter_per_element(p1);
    if ((n + 255) / 256 > (int64_t)UINT32_MAX) {
        set_error("fl_pattern_scatter_table: too many entries for one launch");
        return FL_E_LIMIT;
    }
    scatter_table_kernel<<<(unsigned)((n + 255) / 256), 256, 0, as_stream(stream)>>>(
        elements, nelem, p1, node_elem_ptr, node_elem, node_nbr_ptr, node_nbrs, colour_of_elem,
        static_cast<const int32_t *>(tensor_basis), table, status);
    FL_LAUNCH_CHECK();
    return 0;
}

int fl_pattern_gather_plan(const int32_t *node_nbr_ptr, int64_t nnode, int32_t *node_order, int32_t *bucket_offsets,
                           int32_t *bucket_degree, void *stream) {
    hipStream_t st = as_stream(stream);
    if (!node_nbr_ptr || !node_order || !bucket_offsets || !bucket_degree || nnode <= 0) {
        set_error("fl_pattern_gather_plan: invalid argument");
        return FL_E_ARG;
    }
    // class 0: the nodes of the smallest neighbour count (in meshes with element-interior
    // nodes, exactly those: a node of two distinct elements has more neighbours than an
    // element has nodes); classes 1.. by the limits
    static const int32_t limits[FL_GATHER_BUCKETS] = {0, 128, 256, 512, 1024, 2048};
    std::vector<int32_t> ptr((size_t)nnode + 1);
    FL_HIP_CHECK(hipMemcpyAsync(ptr.data(), node_nbr_ptr, sizeof(int32_t) * (nnode + 1), hipMemcpyDeviceToHost, st));
    FL_HIP_CHECK(hipStreamSynchronize(st));
    int32_t dmin = INT32_MAX;
    for (int64_t a = 0; a < nnode; ++a) dmin = std::min(dmin, ptr[a + 1] - ptr[a]);
    std::vector<int32_t> cls((size_t)nnode);
    int64_t count[FL_GATHER_BUCKETS] = {};
    for (int b = 0; b < FL_GATHER_BUCKETS; ++b) bucket_degree[b] = 0;
    for (int64_t a = 0; a < nnode; ++a) {
        const int32_t d = ptr[a + 1] - ptr[a];
        int b = 0;
        if (d != dmin) {
            b = 1;
            while (b < FL_GATHER_BUCKETS && d > limits[b]) ++b;
        }
        if (b == FL_GATHER_BUCKETS) {
            set_error("fl_pattern_gather_plan: node %lld has %d neighbours (limit %d)", (long long)a, d,
                      limits[FL_GATHER_BUCKETS - 1]);
            return FL_E_LIMIT;
        }
        cls[a] = b;
        ++count[b];
        if (d > bucket_degree[b]) bucket_degree[b] = d;
    }
    bucket_offsets[0] = 0;
    for (int b = 0; b < FL_GATHER_BUCKETS; ++b) bucket_offsets[b + 1] = bucket_offsets[b] + (int32_t)count[b];
    std::vector<int32_t> order((size_t)nnode);
    int64_t cur[FL_GATHER_BUCKETS];
    for (int b = 0; b < FL_GATHER_BUCKETS; ++b) cur[b] = bucket_offsets[b];
    for (int64_t a = 0; a < nnode; ++a) order[cur[cls[a]]++] = (int32_t)a;
    FL_HIP_CHECK(hipMemcpyAsync(node_order, order.data(), sizeof(int32_t) * nnode, hipMemcpyHostToDevice, st));
    FL_HIP_CHECK(hipStreamSynchronize(st));
    return 0;
}

}  // extern "C"
