// mpm_events.h -- kernel-group timing on a stream (profiling mode): events from a pool are
// recorded between the kernel groups of a step and folded into per-group totals after the stream
// has been waited for. The event calls come from the Api parameter (HipEvents in the library), so
// the bookkeeping is also built on the host alone with a stand-in Api for its tests.
//
// Api provides: typedef Event, Stream; static std::string create(Event *), record(Event, Stream),
// elapsed(float *ms, Event start, Event end) returning "" on success or the failure's message;
// static void destroy(Event).
#pragma once

#include <string>
#include <utility>
#include <vector>

template <class Api, int NG>
struct EventMarks {
    typedef typename Api::Event Event;
    typedef typename Api::Stream Stream;

    std::vector<Event> pool;                                  // created events, reused after a fold
    std::vector<std::pair<int, std::pair<int, int>>> marks;   // (group, (start, end)) pool indices
    size_t used = 0;                                          // pool entries recorded since the fold
    int open = -1;                                            // event that starts the next group
    double totalMs[NG] = {};                                  // per group, folded so far
    long long launches[NG] = {};
    std::string err;                                          // first failure since clear()

    // records the next pool event on the stream; -1 (and err) when it cannot be created or recorded
    int mark(Stream s) {
        if (used == pool.size()) {
            Event e{};
            std::string m = Api::create(&e);
            if (!m.empty()) { fail("event create: " + m); return -1; }
            pool.push_back(e);
        }
        std::string m = Api::record(pool[used], s);
        if (!m.empty()) { fail("event record: " + m); return -1; }
        return (int)used++;
    }
    // the event before the first group of a step
    void begin(Stream s) { open = mark(s); }
    // the end of a group: timed from the previous mark when both events were recorded
    void end(int group, Stream s) {
        const int e = mark(s);
        if (open >= 0 && e >= 0 && group >= 0 && group < NG) marks.push_back({group, {open, e}});
        open = e;
    }
    // events recorded since the last fold (the caller folds, after a wait, when this grows large)
    size_t live() const { return used; }
    // fold the recorded groups into the totals; call only after the stream has been waited for
    void fold() {
        for (const auto &m : marks) {
            const int a = m.second.first, b = m.second.second;
            if (a < 0 || b < 0 || a >= (int)used || b >= (int)used) { fail("event index out of range"); continue; }
            float ms = 0.f;
            std::string e = Api::elapsed(&ms, pool[a], pool[b]);
            if (!e.empty()) { fail("event elapsed time: " + e); continue; }
            totalMs[m.first] += ms;
            launches[m.first] += 1;
        }
        marks.clear();
        used = 0;
        open = -1;
    }
    void reset_totals() {
        for (int g = 0; g < NG; g++) { totalMs[g] = 0.; launches[g] = 0; }
    }
    void clear_error() { err.clear(); }
    void destroy() {
        for (Event e : pool) Api::destroy(e);
        pool.clear();
        marks.clear();
        used = 0;
        open = -1;
    }

private:
    void fail(const std::string &m) {
        if (err.empty()) err = m;
    }
};
