#include "common.h"

#include <sstream>

namespace lra {

Profiler &profiler() {
    static Profiler p;
    return p;
}

void Profiler::begin(const std::string &name, hipStream_t s) {
    if (!on_) return;
    hipEvent_t e;
    HIP_CHECK(hipEventCreate(&e));
    HIP_CHECK(hipEventRecord(e, s));
    open_[name] = e;
}

void Profiler::end(const std::string &name, hipStream_t s, double flops, double bytes) {
    if (!on_) return;
    auto it = open_.find(name);
    if (it == open_.end()) LRA_THROW("profiler: end without begin: " + name);
    hipEvent_t e;
    HIP_CHECK(hipEventCreate(&e));
    HIP_CHECK(hipEventRecord(e, s));
    pending_.push_back({name, it->second, e, flops, bytes});
    open_.erase(it);
    if (pending_.size() > 4096) flush();
}

void Profiler::flush() {
    if (pending_.empty()) return;
    HIP_CHECK(hipEventSynchronize(pending_.back().b));
    for (auto &p : pending_) {
        HIP_CHECK(hipEventSynchronize(p.b));
        float ms = 0;
        HIP_CHECK(hipEventElapsedTime(&ms, p.a, p.b));
        auto &r = recs_[p.name];
        r.ms += ms;
        r.calls += 1;
        r.flops += p.flops;
        r.bytes += p.bytes;
        (void)hipEventDestroy(p.a);
        (void)hipEventDestroy(p.b);
    }
    pending_.clear();
}

std::string Profiler::report() {
    flush();
    std::ostringstream os;
    char line[256];
    snprintf(line, sizeof line, "%-32s %8s %12s %12s %12s\n", "region", "calls", "ms", "TFLOP/s", "GB/s");
    os << line;
    for (auto &kv : recs_) {
        const auto &r = kv.second;
        const double s = r.ms * 1e-3;
        snprintf(line, sizeof line, "%-32s %8ld %12.3f %12.3f %12.1f\n", kv.first.c_str(), r.calls, r.ms,
                 s > 0 ? r.flops / s * 1e-12 : 0.0, s > 0 ? r.bytes / s * 1e-9 : 0.0);
        os << line;
    }
    return os.str();
}

}  // namespace lra
