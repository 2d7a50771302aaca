// serve.hpp -- request loop with continuous batching (serve.cpp).
#pragma once
#include <chrono>
#include <condition_variable>
#include <atomic>
#include <deque>
#include <map>
#include <memory>
#include <mutex>
#include <string>
#include <thread>
#include <unordered_map>
#include <vector>

#include "driver.hpp"

namespace rk {

constexpr int kMaxServeSlots = 8;   // batched calls a server runs at once, at most

// One k-way request: a copy of the caller's graph, its part count, shared
// options and its own seed; the result is filled in by the worker that ran it.
struct ServeRequest {
    int64_t id = 0;
    int32_t n = 0, ncon = 1, k = 2;
    std::vector<int32_t> xadj, adjncy, vwgt, adjwgt;   // graph-local ids; empty weights mean ones
    // a graph already in device memory (the caller keeps it valid until the
    // result is collected): read in place, and the part written to d_part
    bool on_device = false;
    int64_t m = 0;
    const int32_t *d_xadj = nullptr, *d_adjncy = nullptr, *d_vwgt = nullptr, *d_adjwgt = nullptr;
    int32_t* d_part = nullptr;
    std::vector<float> ubvec;                           // ncon entries, or empty for 1 + ufactor/1000
    Options opt;
    uint64_t seed = 0;
    int64_t footprint = 0;                              // estimated device bytes while it runs
    std::chrono::steady_clock::time_point t_submit, t_start, t_done;
    bool done = false;
    int32_t status = 0;                                 // METIS return code once done
    std::string error;
    std::vector<int32_t> part;
    int32_t objval = 0;
    int32_t batch_graphs = 0;                           // graphs of the batched call that served it

    bool has_vwgt() const { return on_device ? d_vwgt != nullptr : !vwgt.empty(); }
    bool has_adjwgt() const { return on_device ? d_adjwgt != nullptr : !adjwgt.empty(); }
};

struct ServerStats {
    int64_t batches = 0, served = 0;
    int64_t inflight_batches = 0, max_inflight_batches = 0, max_batch_graphs = 0;
};

// Device memory of a server: segments taken with hipMalloc while there is
// room under the budget and kept until the server ends, carved into one
// contiguous range per batch (first fit).  No segment is freed while batches
// run: a hipFree waits for every stream of the device.  All calls are made
// with the server's mutex held.
class DevicePool {
public:
    explicit DevicePool(int64_t budget) : budget_(budget) {}
    ~DevicePool();
    // A free range of `bytes`, from a new segment when none is free and the
    // budget allows one (extra: device bytes held elsewhere that count
    // against the budget); null when neither is possible.
    char* take(int64_t bytes, int64_t extra);
    bool can_take(int64_t bytes, int64_t extra) const;
    void give(char* p, int64_t bytes);
    // Frees every segment; only while no range is taken (no batch runs).
    void release_all();
    int64_t held() const { return held_; }
    int64_t taken() const { return taken_; }

private:
    struct Seg {
        char* base;
        int64_t size;
        std::map<int64_t, int64_t> free;   // offset -> length, coalesced
    };
    static int64_t round(int64_t b) { return (b + 255) & ~(int64_t)255; }
    bool fits_free(int64_t bytes) const;
    int64_t budget_, held_ = 0, taken_ = 0;
    std::vector<Seg> segs_;
};

class Server {
public:
    // slots: batched calls that may run at once (one worker thread and one
    // HIP stream each); budget: device bytes the batches in flight may take;
    // max_batch_vertices: vertices per batched call at most.
    Server(int device, int slots, int64_t budget, int64_t max_batch_vertices);
    ~Server();
    int64_t submit(std::shared_ptr<ServeRequest> r);
    // The finished request (removed from the server), or null while pending.
    // wait_ms: 0 returns at once, < 0 waits until it is done, > 0 waits at
    // most that long and then fails with a report of every slot's phase.
    std::shared_ptr<ServeRequest> result(int64_t id, int64_t wait_ms);
    ServerStats stats();
    // One line per slot: what it runs and for how long (for diagnostics).
    std::string status();

private:
    void worker(int slot);
    std::string status_locked();
    bool fits_locked(int64_t bytes) const;
    int device_;
    int64_t budget_, max_batch_v_;
    int64_t in_use_ = 0;
    int64_t next_id_ = 1;
    DevicePool pool_;
    struct Slot {
        std::atomic<int32_t> phase{0};
        int64_t graphs = 0, vertices = 0;
        int64_t overflow = 0;                       // device bytes its arena allocated past its ranges
        std::chrono::steady_clock::time_point since;
    };
    std::vector<std::unique_ptr<Slot>> slots_;
    bool stop_ = false;
    ServerStats st_;
    std::mutex mu_;
    std::condition_variable cv_work_, cv_done_;
    std::deque<std::shared_ptr<ServeRequest>> pending_;
    std::unordered_map<int64_t, std::shared_ptr<ServeRequest>> reqs_;
    std::vector<std::thread> threads_;
};

}  // namespace rk
