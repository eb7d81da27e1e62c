// Optional per-phase timing of the library's launches with HIP events recorded on the
// caller's stream (off by default; no cost when disabled).
#include <mutex>
#include <vector>

#include "fl_common.hpp"
#include "fl_profile.hpp"

namespace fl {

namespace {
struct CallRecord {
    int kind;
    std::vector<hipEvent_t> ev;
};
std::mutex g_mu;
bool g_on = false;
std::vector<CallRecord> g_calls;
// events are reused: creating one per mark stalls the host for milliseconds once many
// exist, which would show up as idle GPU time between short launches
std::vector<hipEvent_t> g_pool;

bool take_event(hipEvent_t *e) {
    if (!g_pool.empty()) {
        *e = g_pool.back();
        g_pool.pop_back();
        return true;
    }
    return hipEventCreate(e) == hipSuccess;
}
}  // namespace

void prof_begin(hipStream_t st, int kind) {
    std::lock_guard<std::mutex> lk(g_mu);
    if (!g_on) return;
    CallRecord r;
    r.kind = kind;
    hipEvent_t e;
    if (!take_event(&e)) return;
    hipEventRecord(e, st);
    r.ev.push_back(e);
    g_calls.push_back(r);
}

void prof_mark(hipStream_t st) {
    std::lock_guard<std::mutex> lk(g_mu);
    if (!g_on || g_calls.empty()) return;
    hipEvent_t e;
    if (!take_event(&e)) return;
    hipEventRecord(e, st);
    g_calls.back().ev.push_back(e);
}

namespace {
int g_tune[TUNE_KEYS] = {};
}

int tuning_value(int key, int dflt) {
    if (key < 0 || key >= TUNE_KEYS) return dflt;
    const int v = g_tune[key];
    return v == 0 ? dflt : v;
}

static void clear_locked() {
    for (auto &c : g_calls)
        for (auto e : c.ev) g_pool.push_back(e);
    g_calls.clear();
}

}  // namespace fl

using namespace fl;

extern "C" {

int fl_set_tuning(int32_t key, int32_t value) {
    if (key <= 0 || key >= TUNE_KEYS) {
        set_error("fl_set_tuning: unknown key %d", key);
        return FL_E_ARG;
    }
    g_tune[key] = value;
    return 0;
}

int fl_profile_enable(int on) {
    std::lock_guard<std::mutex> lk(g_mu);
    clear_locked();
    g_on = on != 0;
    return 0;
}

int fl_profile_read(int32_t *kinds, double *phase_ms, int32_t max_calls, int32_t max_phases) {
    std::lock_guard<std::mutex> lk(g_mu);
    int n = 0;
    for (auto &c : g_calls) {
        if (n >= max_calls) break;
        if (!c.ev.empty()) FL_HIP_CHECK(hipEventSynchronize(c.ev.back()));
        kinds[n] = c.kind;
        for (int p = 0; p < max_phases; ++p) {
            float ms = -1.0f;
            if (p + 1 < (int)c.ev.size()) FL_HIP_CHECK(hipEventElapsedTime(&ms, c.ev[p], c.ev[p + 1]));
            phase_ms[(size_t)n * max_phases + p] = ms;
        }
        ++n;
    }
    clear_locked();
    return n;
}

}  // extern "C"
