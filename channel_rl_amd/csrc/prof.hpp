// prof.hpp -- optional per-kernel HIP-event timing of the non-captured step.
// A context turns it on around a step; every launcher brackets its kernels
// with PROF_BEGIN/PROF_END, which do nothing while it is off.
#pragma once
#include <hip/hip_runtime.h>

#include <map>
#include <string>
#include <utility>
#include <vector>

namespace chan {

struct Prof {
    std::vector<std::pair<hipEvent_t, hipEvent_t>> pool;
    std::vector<std::string> names;  // per used pair
    size_t used = 0;
    std::map<std::string, std::pair<double, long>> acc;  // name -> (ms, launches)
    hipEvent_t open = nullptr;
    std::string open_name;

    void begin(const char* name, hipStream_t s)
    {
        if (used == pool.size()) {
            hipEvent_t a, b;
            (void)hipEventCreate(&a);
            (void)hipEventCreate(&b);
            pool.emplace_back(a, b);
        }
        (void)hipEventRecord(pool[used].first, s);
        if (names.size() <= used) names.resize(used + 1);
        names[used] = name;
    }
    void end(hipStream_t s)
    {
        (void)hipEventRecord(pool[used].second, s);
        used++;
    }
    // after a stream sync: fold the recorded pairs into the totals
    void collect()
    {
        for (size_t q = 0; q < used; q++) {
            float ms = 0.f;
            (void)hipEventElapsedTime(&ms, pool[q].first, pool[q].second);
            auto& a = acc[names[q]];
            a.first += ms;
            a.second += 1;
        }
        used = 0;
    }
    ~Prof()
    {
        for (auto& p : pool) {
            (void)hipEventDestroy(p.first);
            (void)hipEventDestroy(p.second);
        }
    }
};

extern thread_local Prof* g_prof;

}  // namespace chan

#define PROF_BEGIN(name, s) \
    do { if (::chan::g_prof) ::chan::g_prof->begin(name, s); } while (0)
#define PROF_END(s) \
    do { if (::chan::g_prof) ::chan::g_prof->end(s); } while (0)
