// Event-bracketed kernel timing registry (see ozl_profile.h).
#include "ozl_profile.h"

#include <mutex>
#include <stdio.h>
#include <string.h>
#include <string>
#include <vector>

namespace {
struct Rec {
    std::string name;
    hipEvent_t a, b;
};
std::mutex g_mu;
bool g_on = false;
std::string g_tag;
std::vector<Rec> g_recs;
std::vector<hipEvent_t> g_pool;

hipEvent_t take()
{
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

ozl_prof_token ozl_prof_start(const char* name, hipStream_t st)
{
    std::lock_guard<std::mutex> g(g_mu);
    if (!g_on) return { -1 };
    Rec r{ g_tag + name, take(), take() };
    if (!r.a || !r.b) return { -1 };
    (void)hipEventRecord(r.a, st);
    g_recs.push_back(r);
    return { (int)g_recs.size() - 1 };
}

void ozl_prof_stop(ozl_prof_token t, hipStream_t st)
{
    std::lock_guard<std::mutex> g(g_mu);
    if (t.slot < 0 || t.slot >= (int)g_recs.size()) return;
    (void)hipEventRecord(g_recs[t.slot].b, st);
}

extern "C" {

// Starts recording an event pair around every kernel the library launches.
void ozl_profile_begin(void)
{
    std::lock_guard<std::mutex> g(g_mu);
    g_on = true;
}

// Prefixes the names of the launches recorded from now on (e.g. "fse/"), so
// that launches of one kernel in different calls are told apart.
void ozl_profile_tag(const char* tag)
{
    std::lock_guard<std::mutex> g(g_mu);
    g_tag = tag ? tag : "";
}

// Stops recording, waits for the events, and writes one line per kernel name:
// "<name> <launches> <total_ms>\n". Returns the bytes written (or needed).
size_t ozl_profile_end(char* out, size_t cap)
{
    std::lock_guard<std::mutex> g(g_mu);
    g_on = false;
    std::vector<std::string> names;
    std::vector<int> counts;
    std::vector<double> ms;
    for (Rec& r : g_recs) {
        float t = 0;
        (void)hipEventSynchronize(r.b);
        (void)hipEventElapsedTime(&t, r.a, r.b);
        size_t k = 0;
        for (; k < names.size(); k++)
            if (names[k] == r.name) break;
        if (k == names.size()) { names.push_back(r.name); counts.push_back(0); ms.push_back(0); }
        counts[k]++;
        ms[k] += t;
        g_pool.push_back(r.a);
        g_pool.push_back(r.b);
    }
    g_recs.clear();
    size_t used = 0;
    for (size_t k = 0; k < names.size(); k++) {
        char line[256];
        int len = snprintf(line, sizeof(line), "%s %d %.6f\n", names[k].c_str(), counts[k], ms[k]);
        if (out && used + (size_t)len < cap) memcpy(out + used, line, (size_t)len);
        used += (size_t)len;
    }
    if (out && cap) out[used < cap ? used : cap - 1] = 0;
    return used;
}

}  // extern "C"
