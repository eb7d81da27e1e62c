// Shared helpers of the florence_amd HIP library (gfx950).
#pragma once
#include <hip/hip_runtime.h>

#include <cstdint>
#include <cstdio>
#include <string>

#include "../../include/florence_amd.h"

namespace fl {

void set_error(const char *fmt, ...);

#define FL_HIP_CHECK(expr)                                                                 \
    do {                                                                                   \
        hipError_t _e = (expr);                                                            \
        if (_e != hipSuccess) {                                                            \
            ::fl::set_error("%s failed at %s:%d: %s", #expr, __FILE__, __LINE__,           \
                            hipGetErrorString(_e));                                        \
            return FL_E_HIP;                                                               \
        }                                                                                  \
    } while (0)

#define FL_LAUNCH_CHECK() FL_HIP_CHECK(hipGetLastError())

inline hipStream_t as_stream(void *s) { return reinterpret_cast<hipStream_t>(s); }

inline size_t align_up(size_t v, size_t a) { return (v + a - 1) / a * a; }

// Bump allocator over a caller-provided workspace.
struct Carve {
    char *base;
    size_t cap;
    size_t off = 0;
    bool ok = true;
    Carve(void *b, size_t c) : base(static_cast<char *>(b)), cap(c) {}
    template <class T>
    T *take(size_t count) {
        size_t start = align_up(off, 256);
        size_t end = start + count * sizeof(T);
        if (end > cap || base == nullptr) {
            ok = false;
            off = end;
            return nullptr;
        }
        off = end;
        return reinterpret_cast<T *>(base + start);
    }
};

// Kernel-variant knobs (set through fl_set_tuning; 0 = library default).
enum TuneKey { TUNE_STIFF_MINW = 1, TUNE_STIFF_VARIANT_BASE = 2, TUNE_STIFF_SF = 7, TUNE_GATHER = 8, TUNE_KEYS = 9 };
int tuning_value(int key, int dflt);

// Order of the direct scatter table of hexes of p1^3 nodes: the order in which the
// factorised stiffness kernel stores an element's entries.  Per element and slice az,
// per column tile s of 16 columns, per lane of the wave that owns the tile, the lane's
// sf_row_tiles x 4 rows.  Row 16 tm + lane / 16 + 4 r is (ab, bb) = divmod(row, p1),
// column 16 s + lane % 16 is ((ae p1 + be) p1 + bz); the entry couples row node
// (ab p1 + ae) p1 + az with column node (bb p1 + be) p1 + bz.  Slots outside the p1^2
// rows or p1^3 columns are unused.
__host__ __device__ inline int sf_row_tiles(int p1) { return (p1 * p1 + 15) / 16; }
__host__ __device__ inline int sf_col_tiles(int p1) { return (p1 * p1 * p1 + 15) / 16; }
__host__ __device__ inline int64_t sf_scatter_per_element(int p1) {
    return (int64_t)p1 * sf_col_tiles(p1) * 64 * sf_row_tiles(p1) * 4;
}
__host__ __device__ inline int64_t sf_scatter_index(int p1, int64_t e, int az, int s, int lane, int tm, int r) {
    return ((((e * p1 + az) * sf_col_tiles(p1) + s) * 64 + lane) * sf_row_tiles(p1) + tm) * 4 + r;
}

// The assembled tangent in matrix-free form (fl_mf_apply, the multigrid's fine-level
// smoother): the element products run where the point-record layout lives
// (fl_assemble.hip), the per-node sums with the solver (fl_mg.hip).
struct MfOp {
    const double *qp;          // point records of the assembly
    const int32_t *basis;      // tensor basis table
    const int32_t *elements;
    int64_t nelem, nnode;
    int npe, G, p1, q;
    const int32_t *node_elem_ptr, *node_elem;
    double *ye;                // element vectors [nelem][npe][3], local node order
    const float *c32;          // optional FP32 copy of the C tensors [nelem][54][G]
    const uint32_t *c16;       // optional BF16 copy, two per word [nelem][27][G] (fields 2j, 2j+1)
};
// Checks a FlMgMatrixFree and resolves its records (FL_E_* with the error set).
int mf_prepare(const FlMgMatrixFree *in, MfOp *op);
// ye_e = K_e x_e for every element (x read as 0 on fixed dofs), from op.c32 when set.
int mf_element_apply(const MfOp &op, const uint8_t *free_mask, const double *x, hipStream_t st);
// c32[e][f][g] = (float) C field f of point g of element e (the records' 54 C values).
int mf_round_records(const MfOp &op, float *c32, hipStream_t st);
int mf_round_records16(const MfOp &op, uint32_t *c16, hipStream_t st);

// Exclusive prefix sum of int64 values in place (n entries, last entry = total), device.
int exclusive_scan_i64(int64_t *data, int64_t n, int64_t *scratch, size_t scratch_count, hipStream_t st);
size_t scan_scratch_count(int64_t n);

}  // namespace fl
