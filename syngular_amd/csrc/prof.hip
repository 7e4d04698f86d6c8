// Live per-kernel timing with HIP events, for the benchmark's roofline figure.
//
// While profiling is on, every launch of an instrumented kernel whose tag is
// selected (syn_prof_select; all tags by default) is bracketed by two events
// recorded on the launch stream, and tagged with the kernel and its
// algorithmic flop count.  syn_prof_end synchronises the events and sums
// elapsed time and flops per tag.  Off by default; costs nothing when off.
// The two event packets per launch cost the stream a few microseconds each,
// so a timing that should stay close to the uninstrumented run selects only
// the kernels it reports.
#include <mutex>
#include <vector>

#include "common.hpp"
#include "prof.hpp"

namespace syn {

namespace {
struct Rec {
    hipEvent_t a, b;
    int tag;
    double flops;
};
std::mutex g_mu;
bool g_on = false;
uint32_t g_mask = 0xffffffffu;   // tags bracketed by events
std::vector<Rec> g_recs;
std::vector<hipEvent_t> g_pool;
// the work figure the next grouped launch of the calling thread carries
// (lanes of the server note and launch from threads of their own)
thread_local double g_pending_flops = 0.0;
double g_extra[SYN_PROF_NTAGS] = {};   // work added without events (totals per call)

hipEvent_t take_event() {
    if (!g_pool.empty()) {
        hipEvent_t e = g_pool.back();
        g_pool.pop_back();
        return e;
    }
    hipEvent_t e = nullptr;
    if (hipEventCreate(&e) != hipSuccess) return nullptr;
    return e;
}
}  // namespace

bool prof_on() { return g_on; }

bool prof_wants(int tag) { return g_on && tag >= 0 && tag < 32 && ((g_mask >> tag) & 1u); }

void prof_add_work(int tag, double work) {
    if (!g_on || tag < 0 || tag >= SYN_PROF_NTAGS) return;
    std::lock_guard<std::mutex> lk(g_mu);
    g_extra[tag] += work;
}

ProfScope::ProfScope(int tag, double flops, hipStream_t s) : tag_(tag), flops_(flops), s_(s), a_(nullptr) {
    if (!prof_wants(tag)) return;
    std::lock_guard<std::mutex> lk(g_mu);
    a_ = take_event();
    if (a_ && hipEventRecord(a_, s_) != hipSuccess) a_ = nullptr;
}

ProfScope::~ProfScope() {
    if (!a_) return;
    std::lock_guard<std::mutex> lk(g_mu);
    hipEvent_t b = take_event();
    if (!b || hipEventRecord(b, s_) != hipSuccess) return;
    if (flops_ < 0.0) {
        flops_ = g_pending_flops;
        g_pending_flops = 0.0;
    }
    g_recs.push_back({a_, b, tag_, flops_});
}

}  // namespace syn

extern "C" int syn_prof_begin(void) {
    std::lock_guard<std::mutex> lk(syn::g_mu);
    for (auto &r : syn::g_recs) {
        syn::g_pool.push_back(r.a);
        syn::g_pool.push_back(r.b);
    }
    syn::g_recs.clear();
    syn::g_pending_flops = 0.0;
    for (double &e : syn::g_extra) e = 0.0;
    syn::g_on = true;
    return SYN_OK;
}

extern "C" int syn_prof_select(uint32_t mask) {
    std::lock_guard<std::mutex> lk(syn::g_mu);
    syn::g_mask = mask;
    return SYN_OK;
}

extern "C" int syn_prof_note_flops(double flops) {
    std::lock_guard<std::mutex> lk(syn::g_mu);
    syn::g_pending_flops = flops;
    return SYN_OK;
}

extern "C" int syn_prof_end(int64_t *launches, double *ms, double *flops, int ntags) {
    using namespace syn;
    std::lock_guard<std::mutex> lk(g_mu);
    g_on = false;
    for (int t = 0; t < ntags; ++t) {
        launches[t] = 0;
        ms[t] = 0.0;
        flops[t] = 0.0;
    }
    for (int t = 0; t < ntags && t < SYN_PROF_NTAGS; ++t) flops[t] += g_extra[t];
    for (auto &r : g_recs) {
        if (int rc = hip_check(hipEventSynchronize(r.b), "prof event sync")) return rc;
        float el = 0.0f;
        if (int rc = hip_check(hipEventElapsedTime(&el, r.a, r.b), "prof elapsed")) return rc;
        if (r.tag >= 0 && r.tag < ntags) {
            launches[r.tag] += 1;
            ms[r.tag] += el;
            flops[r.tag] += r.flops;
        }
    }
    return SYN_PROF_NTAGS;
}
