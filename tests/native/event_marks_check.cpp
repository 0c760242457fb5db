// Host-only check of the kernel-group timing bookkeeping (csrc/mpm_events.h) with a stand-in for
// the HIP event calls whose failures the test chooses: a failed create or record never leaves an
// index that is read later, every fold reads only recorded events, and failures are reported.
#include <cassert>
#include <cstdio>
#include <set>
#include <string>

#include "../../nairn-mpm-fea_amd/csrc/mpm_events.h"

struct Fake {
    typedef int Event;
    typedef int Stream;
    static int created, failCreateAt, failRecordAt, records;
    static std::set<int> recorded;           // events recorded since the last fold
    static std::string create(Event *e) {
        if (created == failCreateAt) { ++created; return "out of events"; }
        *e = 1000 + created++;
        return "";
    }
    static std::string record(Event e, Stream) {
        if (records++ == failRecordAt) return "record refused";
        recorded.insert(e);
        return "";
    }
    static std::string elapsed(float *ms, Event a, Event b) {
        // only events recorded since the last fold may be read
        assert(recorded.count(a) && recorded.count(b));
        *ms = (float)(b - a);
        return "";
    }
    static void destroy(Event) {}
};
int Fake::created = 0, Fake::failCreateAt = -1, Fake::failRecordAt = -1, Fake::records = 0;
std::set<int> Fake::recorded;

static void reset(int failCreate, int failRecord) {
    Fake::created = 0; Fake::records = 0;
    Fake::failCreateAt = failCreate; Fake::failRecordAt = failRecord;
    Fake::recorded.clear();
}

// three steps of four groups each, folding after every step
static EventMarks<Fake, 4> run(int failCreate, int failRecord) {
    reset(failCreate, failRecord);
    EventMarks<Fake, 4> ev;
    for (int step = 0; step < 3; step++) {
        ev.begin(0);
        for (int g = 0; g < 4; g++) ev.end(g, 0);
        ev.fold();
        Fake::recorded.clear();
    }
    return ev;
}

int main() {
    // no failures: every group timed once per step
    {
        EventMarks<Fake, 4> ev = run(-1, -1);
        assert(ev.err.empty());
        for (int g = 0; g < 4; g++) assert(ev.launches[g] == 3);
        assert(ev.pool.size() == 5);          // the pool is reused after each fold
    }
    // event creation fails at each possible call: the groups next to the missing event are not
    // timed, the others are, and the failure is reported
    for (int k = 0; k < 5; k++) {
        EventMarks<Fake, 4> ev = run(k, -1);
        assert(ev.err.find("event create") == 0);
        long long n = 0;
        for (int g = 0; g < 4; g++) n += ev.launches[g];
        assert(n >= 12 - 2 && n < 12);
    }
    // a record fails: the same
    for (int k = 0; k < 15; k++) {
        EventMarks<Fake, 4> ev = run(-1, k);
        assert(ev.err.find("event record") == 0);
        long long n = 0;
        for (int g = 0; g < 4; g++) n += ev.launches[g];
        assert(n >= 12 - 2 && n < 12);
    }
    // a group index outside the table is not kept
    {
        reset(-1, -1);
        EventMarks<Fake, 2> ev;
        ev.begin(0);
        ev.end(5, 0);
        ev.end(-1, 0);
        ev.end(1, 0);
        ev.fold();
        assert(ev.launches[0] == 0 && ev.launches[1] == 1 && ev.err.empty());
    }
    // many steps without a fold never read past the pool
    {
        reset(-1, -1);
        EventMarks<Fake, 1> ev;
        for (int i = 0; i < 2000; i++) { ev.begin(0); ev.end(0, 0); }
        assert(ev.live() == 4000);
        ev.fold();
        assert(ev.launches[0] == 2000 && ev.live() == 0 && ev.err.empty());
    }
    std::puts("event marks ok");
    return 0;
}
