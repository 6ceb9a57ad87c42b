// Per-kernel device timing with HIP events on the context's stream.
#pragma once
#include <map>
#include <string>
#include <vector>

#include "erp_internal.h"

namespace erp {

struct KernelTimer {
    explicit KernelTimer(hipStream_t stream) : s(stream) {}
    ~KernelTimer() {
        drop_pending();
        for (hipEvent_t e : pool) (void)hipEventDestroy(e);
    }
    void begin(const char *name) {
        cur = name;
        cur_a = get();
        (void)hipEventRecord(cur_a, s);
    }
    void end() {
        hipEvent_t b = get();
        (void)hipEventRecord(b, s);
        pending.push_back({cur, cur_a, b});
    }
    // Wait for the recorded kernels and add their durations to the totals.
    int collect() {
        ERP_HIP_TRY(hipStreamSynchronize(s));
        for (const Rec &r : pending) {
            float ms = 0.f;
            ERP_HIP_TRY(hipEventElapsedTime(&ms, r.a, r.b));
            auto it = totals.find(r.name);
            if (it == totals.end()) {
                order.push_back(r.name);
                it = totals.emplace(r.name, Total{}).first;
            }
            it->second.ms += ms;
            it->second.count += 1;
            pool.push_back(r.a);
            pool.push_back(r.b);
        }
        pending.clear();
        return ERP_OK;
    }
    void clear() { drop_pending(); }
    void reset_totals() {
        drop_pending();
        totals.clear();
        order.clear();
    }
    // Average ms per launch for each kernel name, in first-seen order.
    int report(double *ms, int32_t cap, int32_t *count, const char **names) const {
        int32_t k = 0;
        for (const std::string &nm : order) {
            if (k >= cap) break;
            const Total &t = totals.at(nm);
            if (ms) ms[k] = t.count ? t.ms / (double)t.count : 0.0;
            if (names) names[k] = nm.c_str();
            ++k;
        }
        *count = k;
        return ERP_OK;
    }

  private:
    struct Rec {
        const char *name;
        hipEvent_t a, b;
    };
    struct Total {
        double ms = 0.0;
        int64_t count = 0;
    };
    hipEvent_t get() {
        if (!pool.empty()) {
            hipEvent_t e = pool.back();
            pool.pop_back();
            return e;
        }
        hipEvent_t e = nullptr;
        (void)hipEventCreate(&e);
        return e;
    }
    void drop_pending() {
        if (pending.empty()) return;
        (void)hipStreamSynchronize(s);
        for (const Rec &r : pending) {
            pool.push_back(r.a);
            pool.push_back(r.b);
        }
        pending.clear();
    }
    hipStream_t s;
    const char *cur = nullptr;
    hipEvent_t cur_a = nullptr;
    std::vector<Rec> pending;
    std::vector<hipEvent_t> pool;
    std::map<std::string, Total> totals;
    std::vector<std::string> order;
};

}  // namespace erp
