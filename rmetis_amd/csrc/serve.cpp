// serve.cpp -- k-way requests served with continuous batching.
//
// Requests (one graph each, with its own nparts and seed) are queued on
// submit.  Every slot is a worker thread with its own pooled context (HIP
// stream and arena): when idle it takes the oldest pending request and every
// later pending request that can share a batched call with it (same ncon,
// vertex-weight presence, ubvec and options other than the seed), as far as
// the memory budget and
// the per-batch vertex cap allow, and runs them through kway_batch_device.
// Requests that arrive while batches are on the GPU are picked up by the next
// free slot, so the work of several batches overlaps on the device.  Every
// result equals that of a separate partition_kway call with the request's
// seed (kway.rs:20-141), whatever the batch it ran in.
#include "serve.hpp"

#include <algorithm>

namespace rk {

namespace {

using Clock = std::chrono::steady_clock;

// Device memory a request's hierarchy and work buffers may take: the input
// CSR, the coarse levels (at most as large again), the refinement buffers
// (tile records, movers, candidate pool) and the contraction tables.
int64_t footprint_bytes(int64_t n, int64_t m, int64_t ncon) {
    const int64_t input = 4 * (n + 1) + 8 * m + 4 * n * ncon;
    return 3 * input + 36 * m + 160 * n + (1 << 20);
}

bool same_setup(const ServeRequest& a, const ServeRequest& b) {
    const Options &x = a.opt, &y = b.opt;
    // graphs without vertex weights keep unit coarse weights, so they batch
    // only with each other
    return a.ncon == b.ncon && a.has_vwgt() == b.has_vwgt() && a.ubvec == b.ubvec &&
           x.ctype_shem == y.ctype_shem && x.iptype == y.iptype &&
           x.rtype_greedy == y.rtype_greedy && x.ncuts == y.ncuts && x.niter == y.niter && x.ufactor == y.ufactor;
}

constexpr int64_t kSegmentMin = 1ll << 30;   // device memory is taken in segments of 1 GiB or more

const char* phase_name(int32_t p) {
    switch (p) {
        case kPhaseIdle: return "idle";
        case kPhasePack: return "packing the batch";
        case kPhaseCoarsen: return "coarsening level";
        case kPhaseInitial: return "growing the coarsest graphs";
        case kPhaseRefine: return "refining level";
        case kPhaseFinish: return "finishing";
        default: return "?";
    }
}

}  // namespace

DevicePool::~DevicePool() { release_all(); }

bool DevicePool::fits_free(int64_t bytes) const {
    for (const Seg& sg : segs_)
        for (const auto& f : sg.free)
            if (f.second >= bytes) return true;
    return false;
}

bool DevicePool::can_take(int64_t bytes, int64_t extra) const {
    bytes = round(bytes);
    return fits_free(bytes) || held_ + extra + bytes <= budget_;
}

char* DevicePool::take(int64_t bytes, int64_t extra) {
    bytes = round(bytes);
    for (Seg& sg : segs_)
        for (auto it = sg.free.begin(); it != sg.free.end(); ++it)
            if (it->second >= bytes) {
                const int64_t off = it->first, len = it->second;
                sg.free.erase(it);
                if (len > bytes) sg.free[off + bytes] = len - bytes;
                taken_ += bytes;
                return sg.base + off;
            }
    if (held_ + extra + bytes > budget_) return nullptr;
    const int64_t room = budget_ - held_ - extra;
    const int64_t size = std::min<int64_t>(room, std::max<int64_t>(bytes, kSegmentMin));
    void* p = nullptr;
    RK_HIP(hipMalloc(&p, (size_t)size));
    segs_.push_back({(char*)p, size, {}});
    held_ += size;
    if (size > bytes) segs_.back().free[bytes] = size - bytes;
    taken_ += bytes;
    return (char*)p;
}

void DevicePool::give(char* p, int64_t bytes) {
    bytes = round(bytes);
    for (Seg& sg : segs_) {
        if (p < sg.base || p >= sg.base + sg.size) continue;
        int64_t off = p - sg.base, len = bytes;
        auto next = sg.free.lower_bound(off);
        if (next != sg.free.end() && next->first == off + len) {   // merge with the following range
            len += next->second;
            next = sg.free.erase(next);
        }
        if (next != sg.free.begin()) {                              // and with the preceding one
            auto prev = std::prev(next);
            if (prev->first + prev->second == off) {
                off = prev->first;
                len += prev->second;
                sg.free.erase(prev);
            }
        }
        sg.free[off] = len;
        taken_ -= bytes;
        return;
    }
}

void DevicePool::release_all() {
    for (Seg& sg : segs_) (void)hipFree(sg.base);
    segs_.clear();
    held_ = 0;
    taken_ = 0;
}

Server::Server(int device, int slots, int64_t budget, int64_t max_batch_vertices)
    : device_(device), budget_(budget), max_batch_v_(max_batch_vertices), pool_(budget) {
    if (slots < 1) throw Error(-2, "a server needs at least one slot");
    for (int i = 0; i < slots; ++i) {
        slots_.push_back(std::make_unique<Slot>());
        slots_.back()->since = Clock::now();
    }
    for (int i = 0; i < slots; ++i) threads_.emplace_back([this, i]() { worker(i); });
}

Server::~Server() {
    {
        std::lock_guard<std::mutex> lk(mu_);
        stop_ = true;
    }
    cv_work_.notify_all();
    for (auto& t : threads_) t.join();
    std::lock_guard<std::mutex> lk(mu_);
    for (auto& r : pending_) {
        r->status = -4;
        r->error = "the server was destroyed before the request ran";
        r->done = true;
    }
    pending_.clear();
    cv_done_.notify_all();
}

int64_t Server::submit(std::shared_ptr<ServeRequest> r) {
    if (!r->on_device) r->m = (int64_t)r->adjncy.size();
    r->footprint = footprint_bytes(r->n, r->m, r->ncon);
    if (r->footprint > budget_)
        throw Error(-3, "request needs about " + std::to_string(r->footprint >> 20) +
                            " MiB of device memory, above the server's budget");
    std::lock_guard<std::mutex> lk(mu_);
    r->id = next_id_++;
    r->t_submit = Clock::now();
    reqs_[r->id] = r;
    pending_.push_back(r);
    cv_work_.notify_one();
    return r->id;
}

std::shared_ptr<ServeRequest> Server::result(int64_t id, int64_t wait_ms) {
    std::unique_lock<std::mutex> lk(mu_);
    auto it = reqs_.find(id);
    if (it == reqs_.end()) throw Error(-2, "unknown request id " + std::to_string(id));
    std::shared_ptr<ServeRequest> r = it->second;
    if (wait_ms < 0) {
        cv_done_.wait(lk, [&]() { return r->done; });
    } else if (wait_ms > 0) {
        if (!cv_done_.wait_for(lk, std::chrono::milliseconds(wait_ms), [&]() { return r->done; }))
            throw Error(-4, "request " + std::to_string(id) + " not finished after " + std::to_string(wait_ms) +
                                " ms; " + status_locked());
    }
    if (!r->done) return nullptr;
    reqs_.erase(it);
    return r;
}

ServerStats Server::stats() {
    std::lock_guard<std::mutex> lk(mu_);
    return st_;
}

std::string Server::status() {
    std::lock_guard<std::mutex> lk(mu_);
    return status_locked();
}

std::string Server::status_locked() {
    std::string out = std::to_string(pending_.size()) + " pending, " + std::to_string(st_.inflight_batches) +
                      " batches in flight, device memory " + std::to_string(pool_.taken() >> 20) + " of " +
                      std::to_string(pool_.held() >> 20) + " MiB held in use (budget " +
                      std::to_string(budget_ >> 20) + " MiB)";
    const auto now = Clock::now();
    for (size_t i = 0; i < slots_.size(); ++i) {
        const Slot& sl = *slots_[i];
        const int32_t ph = sl.phase.load(std::memory_order_relaxed);
        const double ms = std::chrono::duration<double, std::milli>(now - sl.since).count();
        out += "; slot " + std::to_string(i) + ": " + phase_name(ph / 1000);
        if (ph / 1000 == kPhaseCoarsen || ph / 1000 == kPhaseRefine) out += " " + std::to_string(ph % 1000);
        if (ph / 1000 != kPhaseIdle)
            out += " (" + std::to_string(sl.graphs) + " graphs, " + std::to_string(sl.vertices) + " vertices)";
        out += " for " + std::to_string((int64_t)ms) + " ms";
    }
    return out;
}

bool Server::fits_locked(int64_t bytes) const {
    int64_t extra = 0;
    for (const auto& sl : slots_) extra += sl->overflow;
    return pool_.can_take(bytes, extra);
}

void Server::worker(int slot) {
    (void)hipSetDevice(device_);
    Context* cx = nullptr;
    try {
        cx = acquire_context();
    } catch (const Error&) {
        return;   // no usable device: the queue is failed by the destructor
    }
    Slot& me = *slots_[slot];
    cx->phase = &me.phase;
    for (;;) {
        std::vector<std::shared_ptr<ServeRequest>> batch;
        int64_t fp = 0;
        char* range = nullptr;
        {
            std::unique_lock<std::mutex> lk(mu_);
            // the oldest request runs as soon as its device memory is free
            // next to the batches in flight; with none in flight it always
            // runs (submit checked it against the budget), after the pool
            // drops segments it cannot use
            cv_work_.wait(lk, [&]() {
                return stop_ ||
                       (!pending_.empty() && (st_.inflight_batches == 0 || fits_locked(pending_.front()->footprint)));
            });
            if (stop_) break;
            if (st_.inflight_batches == 0 && !fits_locked(pending_.front()->footprint)) pool_.release_all();
            const std::shared_ptr<ServeRequest> head = pending_.front();
            int64_t nv = 0;
            for (auto it = pending_.begin(); it != pending_.end();) {
                const auto& r = *it;
                const bool first = batch.empty();
                if (!first && (!same_setup(*head, *r) || !fits_locked(fp + r->footprint) ||
                               nv + r->n > max_batch_v_)) {
                    ++it;
                    continue;
                }
                fp += r->footprint;
                nv += r->n;
                batch.push_back(r);
                it = pending_.erase(it);
            }
            int64_t extra = 0;
            for (const auto& sl : slots_) extra += sl->overflow;
            // null only when the slots' own overflow chunks leave no room
            // under the budget: the batch then runs in its arena's chunks
            range = pool_.take(fp, extra);
            in_use_ += fp;
            st_.inflight_batches += 1;
            st_.max_inflight_batches = std::max(st_.max_inflight_batches, st_.inflight_batches);
            me.graphs = (int64_t)batch.size();
            me.vertices = nv;
            me.since = Clock::now();
            me.phase.store(kPhasePack * 1000, std::memory_order_relaxed);
        }
        const auto t_start = Clock::now();
        int32_t status = 1;
        std::string error;
        const int32_t G = (int32_t)batch.size();
        std::vector<int32_t> objval(G, 0);
        std::vector<int32_t> host_part;
        std::vector<int32_t> gvtx(G + 1, 0);
        try {
            // the union CSR: graphs one after another, global vertex ids
            const int32_t ncon = batch[0]->ncon;
            int64_t n = 0, m = 0;
            bool any_w = false;
            for (int32_t i = 0; i < G; ++i) {
                gvtx[i] = (int32_t)n;
                n += batch[i]->n;
                m += batch[i]->m;
                any_w = any_w || batch[i]->has_adjwgt();
            }
            if (n > INT32_MAX || m > INT32_MAX) throw Error(-4, "batch too large for int32 CSR");
            gvtx[G] = (int32_t)n;
            Arena& ar = cx->arena;
            if (range) ar.attach(range, (size_t)fp, false);
            ar.reset();
            cx->upload_reset();
            // pinned staging only for what comes from the host
            bool any_host = false, fill_w = false;
            for (int32_t i = 0; i < G; ++i) {
                any_host = any_host || !batch[i]->on_device;
                fill_w = fill_w || !batch[i]->on_device || !batch[i]->has_adjwgt();
            }
            const bool unit = !batch[0]->has_vwgt();
            auto stage = [&](int64_t cnt) { return (int32_t*)cx->upload_stage(sizeof(int32_t) * (size_t)std::max<int64_t>(cnt, 1)); };
            int32_t* hx = any_host ? stage(n) : nullptr;
            int32_t* ha = any_host ? stage(m) : nullptr;
            int32_t* hv = !unit && any_host ? stage(n * ncon) : nullptr;
            int32_t* hw = any_w && fill_w ? stage(m) : nullptr;
            int32_t* hend = stage(1);
            std::vector<int32_t> nparts(G);
            std::vector<uint64_t> seeds(G);
            // the union in HBM: graphs from the host through the pinned
            // arrays (one copy each per array), graphs already on the device
            // by the packing kernel and device copies
            DevGraph g;
            g.n = (int32_t)n;
            g.m = (int32_t)m;
            g.ncon = ncon;
            g.xadj = ar.get<int32_t>(n + 1);
            g.adjncy = ar.get<int32_t>(m);
            g.vwgt = unit ? nullptr : ar.get<int32_t>(n * ncon);
            g.adjwgt = any_w ? ar.get<int32_t>(m) : nullptr;
            hipStream_t s = cx->stream;
            auto h2d = [&](int32_t* dst, const int32_t* src, int64_t cnt) {
                if (cnt > 0) RK_HIP(hipMemcpyAsync(dst, src, sizeof(int32_t) * (size_t)cnt, hipMemcpyHostToDevice, s));
            };
            auto d2d = [&](int32_t* dst, const int32_t* src, int64_t cnt) {
                if (cnt > 0)
                    RK_HIP(hipMemcpyAsync(dst, src, sizeof(int32_t) * (size_t)cnt, hipMemcpyDeviceToDevice, s));
            };
            int64_t eo = 0;
            for (int32_t i = 0; i < G; ++i) {
                const ServeRequest& r = *batch[i];
                const int32_t v0 = gvtx[i];
                if (r.on_device) {
                    pack_csr(r.n, r.m, r.d_xadj, r.d_adjncy, v0, (int32_t)eo, g.xadj, g.adjncy, s);
                    if (!unit) d2d(g.vwgt + (int64_t)v0 * ncon, r.d_vwgt, (int64_t)r.n * ncon);
                    if (any_w) {
                        if (r.d_adjwgt) {
                            d2d(g.adjwgt + eo, r.d_adjwgt, r.m);
                        } else {
                            std::fill(hw + eo, hw + eo + r.m, 1);
                            h2d(g.adjwgt + eo, hw + eo, r.m);
                        }
                    }
                } else {
                    for (int32_t v = 0; v < r.n; ++v) hx[v0 + v] = (int32_t)(eo + r.xadj[v]);
                    for (int64_t j = 0; j < r.m; ++j) ha[eo + j] = r.adjncy[j] + v0;
                    h2d(g.xadj + v0, hx + v0, r.n);
                    h2d(g.adjncy + eo, ha + eo, r.m);
                    if (!unit) {
                        std::copy(r.vwgt.begin(), r.vwgt.end(), hv + (int64_t)v0 * ncon);
                        h2d(g.vwgt + (int64_t)v0 * ncon, hv + (int64_t)v0 * ncon, (int64_t)r.n * ncon);
                    }
                    if (any_w) {
                        if (r.adjwgt.empty()) std::fill(hw + eo, hw + eo + r.m, 1);
                        else std::copy(r.adjwgt.begin(), r.adjwgt.end(), hw + eo);
                        h2d(g.adjwgt + eo, hw + eo, r.m);
                    }
                }
                eo += r.m;
                nparts[i] = r.k;
                seeds[i] = r.seed;
            }
            *hend = (int32_t)m;
            h2d(g.xadj + n, hend, 1);
            int32_t* d_part = ar.get<int32_t>(n);
            const float* ub = batch[0]->ubvec.empty() ? nullptr : batch[0]->ubvec.data();
            kway_batch_device(*cx, g, G, gvtx.data(), nparts.data(), ub, batch[0]->opt, seeds.data(), d_part,
                              objval.data(), nullptr, 1);   // one attempt at a time: the footprint is per copy
            host_part.resize(n);
            int32_t* hp = (int32_t*)cx->host_stage2(sizeof(int32_t) * (size_t)n);
            for (int32_t i = 0; i < G; ++i) {
                const ServeRequest& r = *batch[i];
                if (r.on_device)
                    d2d(r.d_part, d_part + gvtx[i], r.n);
                else
                    RK_HIP(hipMemcpyAsync(hp + gvtx[i], d_part + gvtx[i], sizeof(int32_t) * (size_t)r.n,
                                          hipMemcpyDeviceToHost, s));
            }
            RK_HIP(hipStreamSynchronize(s));
            std::copy(hp, hp + n, host_part.begin());
        } catch (const Error& e) {
            status = e.code;
            error = e.what();
        } catch (const std::bad_alloc&) {
            status = -3;
            error = "host out of memory";
        } catch (const std::exception& e) {
            status = -4;
            error = e.what();
        }
        cx->arena.detach();     // the range goes back to the pool below
        cx->arena.reset();
        cx->upload_reset();
        const auto t_done = Clock::now();
        {
            std::lock_guard<std::mutex> lk(mu_);
            if (range) pool_.give(range, fp);
            me.overflow = (int64_t)cx->arena.owned_bytes();
            me.phase.store(kPhaseIdle * 1000, std::memory_order_relaxed);
            me.since = t_done;
            for (int32_t i = 0; i < G; ++i) {
                ServeRequest& r = *batch[i];
                r.t_start = t_start;
                r.t_done = t_done;
                r.batch_graphs = G;
                r.status = status;
                r.error = error;
                if (status == 1) {
                    if (!r.on_device) r.part.assign(host_part.begin() + gvtx[i], host_part.begin() + gvtx[i + 1]);
                    r.objval = objval[i];
                }
                r.done = true;
            }
            in_use_ -= fp;
            st_.inflight_batches -= 1;
            st_.batches += 1;
            st_.served += G;
            st_.max_batch_graphs = std::max<int64_t>(st_.max_batch_graphs, G);
        }
        cv_done_.notify_all();
        cv_work_.notify_all();
    }
    cx->phase = nullptr;   // the slot's counter dies with the server; the context lives on in the pool
    release_context(cx);
}

}  // namespace rk
