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

// Exclusive prefix sum of int64 values in place (n entries, last entry = total), device.
int exclusive_scan_i64(int64_t *data, int64_t n, int64_t *scratch, size_t scratch_count, hipStream_t st);
size_t scan_scratch_count(int64_t n);

}  // namespace fl
