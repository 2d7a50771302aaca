// profiler.hpp -- optional per-kernel timing with HIP events on the stream the
// kernels are launched on, with the algorithmic HBM bytes of each launch.
// Off by default; enabled through rmetis_hip_profile_enable().
#pragma once
#include <map>
#include <mutex>
#include <utility>
#include <vector>

#include "common.hpp"

namespace rk {

enum ProfTag {
    kProfGain = 0,      // cut-pass gain launches over every tile (boundary, gain tables, candidates)
    kProfCompact = 1,   // tile-count scan + ordered compaction
    kProfAdmit = 2,     // ordered admission (one workgroup)
    kProfMatch = 3,     // matching (suitor init, chains, pairing)
    kProfContract = 4,  // coarse numbering + contraction
    kProfOther = 5,     // part weights, projection, edge cut
    kProfGainInc = 6,   // gain launches over a data-dependent part of the level: incremental
                        // cut passes (tiles near the last pass's moves) and balance-repair
                        // passes (tiles with a vertex in an overweight part)
    kProfTags = 7
};

class Profiler {
public:
    static Profiler& get() {
        static Profiler p;
        return p;
    }
    bool on() const { return on_; }
    // Records only launches of tag (and, with size >= 0, on levels of that
    // many vertices); tag < 0 records every launch.
    void filter(int tag, long long size) {
        std::lock_guard<std::mutex> lk(mu_);
        ftag_ = tag;
        fsize_ = size;
    }
    bool wants(int tag, long long size) const {
        return on_ && (ftag_ < 0 || (tag == ftag_ && (fsize_ < 0 || size == fsize_)));
    }
    void enable(bool v) {
        std::lock_guard<std::mutex> lk(mu_);
        harvest_locked();
        on_ = v;
        acc_.clear();
    }
    hipEvent_t start(hipStream_t s) {
        hipEvent_t e = take();
        (void)hipEventRecord(e, s);
        return e;
    }
    void stop(int tag, hipEvent_t e0, double bytes, long long size, hipStream_t s) {
        hipEvent_t e1 = take();
        (void)hipEventRecord(e1, s);
        std::lock_guard<std::mutex> lk(mu_);
        pending_.push_back({tag, e0, e1, bytes, size});
    }
    // Totals of the launches of one tag; size >= 0 restricts them to launches
    // on a level of that many vertices.
    void read(int tag, long long size, long long* launches, double* ms, double* bytes) {
        std::lock_guard<std::mutex> lk(mu_);
        harvest_locked();
        *launches = 0;
        *ms = *bytes = 0;
        for (auto& kv : acc_) {
            if (kv.first.first != tag || (size >= 0 && kv.first.second != size)) continue;
            *launches += kv.second.launches;
            *ms += kv.second.ms;
            *bytes += kv.second.bytes;
        }
    }

private:
    struct Rec {
        int tag;
        hipEvent_t e0, e1;
        double bytes;
        long long size;
    };
    struct Acc {
        long long launches = 0;
        double ms = 0, bytes = 0;
    };
    hipEvent_t take() {
        std::lock_guard<std::mutex> lk(mu_);
        if (!free_.empty()) {
            hipEvent_t e = free_.back();
            free_.pop_back();
            return e;
        }
        hipEvent_t e = nullptr;
        (void)hipEventCreate(&e);
        return e;
    }
    void harvest_locked() {
        for (auto& r : pending_) {
            (void)hipEventSynchronize(r.e1);
            float ms = 0.f;
            if (hipEventElapsedTime(&ms, r.e0, r.e1) == hipSuccess) {
                Acc& a = acc_[{r.tag, r.size}];
                a.launches += 1;
                a.ms += ms;
                a.bytes += r.bytes;
            }
            free_.push_back(r.e0);
            free_.push_back(r.e1);
        }
        pending_.clear();
    }
    std::mutex mu_;
    bool on_ = false;
    int ftag_ = -1;
    long long fsize_ = -1;
    std::vector<Rec> pending_;
    std::vector<hipEvent_t> free_;
    std::map<std::pair<int, long long>, Acc> acc_;
};

// Algorithmic HBM bytes of one level's matching: xadj, adjncy, adjwgt and
// the neighbour-state gather at 4 B per entry, the match write.
inline double match_bytes(long long n, long long m, bool weighted) {
    return 4.0 * (n + 1) + (weighted ? 12.0 : 8.0) * m + 4.0 * n;
}
// ... and of its contraction into n' vertices and m' entries: the coarse
// numbering (leader flags and cmap, read and written), the fine CSR with the
// cmap[adjncy] gather, cmap and vertex weights, the coarse CSR and weights.
inline double contract_bytes(long long n, long long m, int ncon, bool weighted, long long cn, long long cm) {
    return 8.0 * n + 4.0 * (n + 1) + (weighted ? 12.0 : 8.0) * m + 4.0 * n + 4.0 * n * ncon + 4.0 * (cn + 1) +
           8.0 * cm + 4.0 * cn * ncon;
}

// Times the launches issued during its lifetime when profiling is on.
struct ProfScope {
    int tag;
    double bytes;
    long long size;
    hipStream_t s;
    hipEvent_t e0 = nullptr;
    ProfScope(int t, double b, long long sz, hipStream_t st) : tag(t), bytes(b), size(sz), s(st) {
        if (Profiler::get().wants(t, sz)) e0 = Profiler::get().start(s);
    }
    ~ProfScope() {
        if (e0) Profiler::get().stop(tag, e0, bytes, size, s);
    }
};

}  // namespace rk
