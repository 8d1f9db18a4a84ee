// plx_events.cpp -- per-handle HIP-event timers around the library's kernel launches.
//
// The device counterpart of PyLith's EventLogger (petsc/EventLogger.icc:23-47): named events
// with a count and an accumulated time, here measured with hipEvents recorded on the handle's
// stream around each launch, so the times are those of the kernels themselves.
#include <map>

#include "plx_internal.hpp"

namespace plx {

ScopedEvent::ScopedEvent(plx_dm dm, const char *name) : dm_(dm), name_(name) {
    if (!dm_->events.on || !name_) return;
    if (hipEventCreate(&a_) == hipSuccess && hipEventRecord(a_, dm_->stream) == hipSuccess) active_ = true;
}

ScopedEvent::~ScopedEvent() {
    if (!active_) return;
    hipEvent_t b = nullptr;
    if (hipEventCreate(&b) == hipSuccess && hipEventRecord(b, dm_->stream) == hipSuccess)
        dm_->events.pending.push_back({name_, a_, b});
}

static void fold_pending(plx_dm dm) {
    for (auto &p : dm->events.pending) {
        float ms = 0.0f;
        (void)hipEventSynchronize(p.end);
        if (hipEventElapsedTime(&ms, p.begin, p.end) == hipSuccess) {
            auto &s = dm->events.stats[p.name];
            s.first += 1;
            s.second += ms;
        }
        (void)hipEventDestroy(p.begin);
        (void)hipEventDestroy(p.end);
    }
    dm->events.pending.clear();
}

void events_release(plx_dm dm) { fold_pending(dm); }

}  // namespace plx

using namespace plx;

extern "C" {

int plx_dm_set_event_logging(plx_dm dm, int on) {
    if (dm) dm->epoch++;
    if (!dm) return 1;
    fold_pending(dm);
    dm->events.on = on != 0;
    return 0;
}

int plx_dm_get_event_stats(plx_dm dm, const char *name, int *count, double *total_ms) {
    if (!dm || !name) return 1;
    fold_pending(dm);
    auto it = dm->events.stats.find(name);
    if (count) *count = it == dm->events.stats.end() ? 0 : it->second.first;
    if (total_ms) *total_ms = it == dm->events.stats.end() ? 0.0 : it->second.second;
    return 0;
}

int plx_dm_reset_event_stats(plx_dm dm) {
    if (!dm) return 1;
    fold_pending(dm);
    dm->events.stats.clear();
    return 0;
}

}  // extern "C"
