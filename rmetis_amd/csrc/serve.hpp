// serve.hpp -- request loop with continuous batching (serve.cpp).
#pragma once
#include <chrono>
#include <condition_variable>
#include <deque>
#include <memory>
#include <mutex>
#include <string>
#include <thread>
#include <unordered_map>
#include <vector>

#include "driver.hpp"

namespace rk {

constexpr int kMaxServeSlots = 3;   // batched calls a server runs at once, at most

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

class Server {
public:
    // slots: batched calls that may run at once (one worker thread and one
    // HIP stream each); budget: device bytes the batches in flight may take;
    // max_batch_vertices: vertices per batched call at most.
    Server(int device, int slots, int64_t budget, int64_t max_batch_vertices);
    ~Server();
    int64_t submit(std::shared_ptr<ServeRequest> r);
    // The finished request (removed from the server), or null while pending
    // (waits for it when wait is true).
    std::shared_ptr<ServeRequest> result(int64_t id, bool wait);
    ServerStats stats();

private:
    void worker();
    int device_;
    int64_t budget_, max_batch_v_;
    int64_t in_use_ = 0;
    int64_t next_id_ = 1;
    bool stop_ = false;
    ServerStats st_;
    std::mutex mu_;
    std::condition_variable cv_work_, cv_done_;
    std::deque<std::shared_ptr<ServeRequest>> pending_;
    std::unordered_map<int64_t, std::shared_ptr<ServeRequest>> reqs_;
    std::vector<std::thread> threads_;
};

}  // namespace rk
