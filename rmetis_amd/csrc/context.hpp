// context.hpp -- per-call device context: one HIP stream (default priority),
// a bump arena of device memory and pinned staging.  Contexts are pooled per
// device so that concurrent C-ABI calls each own a stream and arena (the
// reference's entry points are reentrant, c_api.rs:55-81) and repeated calls
// reuse their memory instead of calling hipMalloc again.
#pragma once
#include <algorithm>
#include <atomic>
#include <memory>
#include <mutex>
#include <string>
#include <vector>

#include "common.hpp"

namespace rk {

class Arena {
public:
    struct Mark {
        size_t chunk, off;
    };
    ~Arena() { release(); }
    // Puts [base, base + bytes), memory the arena does not own, first in its
    // chain (replacing the previous borrowed range; owned chunks stay after
    // it).  fixed: running past the borrowed range is an error instead of a
    // hipMalloc of one more owned chunk.
    void attach(void* base, size_t bytes, bool fixed) {
        if (!chunks_.empty() && !chunks_[0].owned) chunks_.erase(chunks_.begin());
        chunks_.insert(chunks_.begin(), Chunk{(char*)base, bytes, false});
        fixed_ = fixed;
        reset();
    }
    // Drops the borrowed range (owned chunks stay).
    void detach() {
        if (!chunks_.empty() && !chunks_[0].owned) chunks_.erase(chunks_.begin());
        fixed_ = false;
        reset();
    }
    void* alloc(size_t bytes) {
        bytes = (bytes + 255) & ~(size_t)255;
        if (bytes == 0) bytes = 256;
        while (cur_ < chunks_.size()) {
            if (off_ + bytes <= chunks_[cur_].size) {
                char* p = chunks_[cur_].base + off_;
                off_ += bytes;
                return p;
            }
            ++cur_;
            off_ = 0;
        }
        if (fixed_)
            throw Error(-3, "the fixed device workspace (" + std::to_string(chunks_.empty() ? 0 : chunks_[0].size >> 20) +
                                " MiB) is too small for this graph's hierarchy");
        size_t sz = bytes > kChunk ? bytes : kChunk;
        void* p = nullptr;
        RK_HIP(hipMalloc(&p, sz));
        chunks_.push_back({(char*)p, sz, true});
        cur_ = chunks_.size() - 1;
        off_ = bytes;
        return p;
    }
    template <typename T>
    T* get(int64_t count) {
        return (T*)alloc((size_t)(count > 0 ? count : 1) * sizeof(T));
    }
    Mark mark() const { return {cur_, off_}; }
    void rewind(Mark m) {
        cur_ = m.chunk;
        off_ = m.off;
    }
    void reset() {
        cur_ = 0;
        off_ = 0;
    }
    void release() {
        for (auto& c : chunks_)
            if (c.owned) (void)hipFree(c.base);
        chunks_.clear();
        fixed_ = false;
        reset();
    }
    // Bytes of the chunks the arena allocated itself (not borrowed ranges).
    size_t owned_bytes() const {
        size_t t = 0;
        for (auto& c : chunks_)
            if (c.owned) t += c.size;
        return t;
    }

private:
    static constexpr size_t kChunk = 64ull << 20;
    struct Chunk {
        char* base;
        size_t size;
        bool owned;
    };
    std::vector<Chunk> chunks_;
    size_t cur_ = 0, off_ = 0;
    bool fixed_ = false;
};

struct Context {
    int device = 0;
    hipStream_t stream = nullptr;
    Arena arena;
    // Progress of the call for observers (a server slot's status): phase *
    // 1000 + level, see kPhase*.  Null when nobody watches.
    std::atomic<int32_t>* phase = nullptr;
    void set_phase(int32_t p, int32_t level = 0) {
        if (phase) phase->store(p * 1000 + level, std::memory_order_relaxed);
    }
    // pageable: host staging is ordinary heap memory instead of pinned
    // hipHostMalloc buffers (for callers that allow no HIP allocation)
    bool pageable = false;
    void* pinned = nullptr;
    size_t pinned_size = 0;

    void* pinned2 = nullptr;
    size_t pinned2_size = 0;
    int32_t* small = nullptr;   // pinned scratch for device-to-host reads of a few counters

    int32_t* host_small() {
        if (!small) {
            if (pageable) small = (int32_t*)heap_block(kSmallBytes);
            else RK_HIP(hipHostMalloc((void**)&small, kSmallBytes, hipHostMallocDefault));
        }
        return small;
    }
    static constexpr size_t kSmallBytes = 4096;

    // Two independent pinned staging buffers; a buffer may be reused only
    // after the stream has been synchronised past its last copy.
    void* host_stage(size_t bytes) { return grow(pinned, pinned_size, bytes); }
    void* host_stage2(size_t bytes) { return grow(pinned2, pinned2_size, bytes); }
    // Heap blocks of a pageable context; kept until the context ends, so a
    // pointer handed out stays valid while copies from it may be in flight.
    std::vector<std::unique_ptr<char[]>> heap;
    void* heap_block(size_t bytes) {
        heap.emplace_back(new char[bytes > 0 ? bytes : 1]);
        return heap.back().get();
    }
    // Pinned bump area for many small uploads within one call: nothing is
    // reused before upload_reset(), which the caller issues only after the
    // stream has been synchronised past the last copy.  Chunks are kept for
    // later calls.
    void* upload_stage(size_t bytes) {
        bytes = (bytes + 255) & ~(size_t)255;
        while (up_cur < up_chunks.size()) {
            if (up_off + bytes <= up_chunks[up_cur].second) {
                char* p = up_chunks[up_cur].first + up_off;
                up_off += bytes;
                return p;
            }
            ++up_cur;
            up_off = 0;
        }
        const size_t sz = bytes > kUpChunk ? bytes : kUpChunk;
        void* p = nullptr;
        if (pageable) p = heap_block(sz);
        else RK_HIP(hipHostMalloc(&p, sz, hipHostMallocDefault));
        up_chunks.push_back({(char*)p, sz});
        up_cur = up_chunks.size() - 1;
        up_off = bytes;
        return p;
    }
    void upload_reset() {
        up_cur = 0;
        up_off = 0;
    }
    static constexpr size_t kUpChunk = 4u << 20;
    std::vector<std::pair<char*, size_t>> up_chunks;
    size_t up_cur = 0, up_off = 0;

    // Grows a staging buffer (at least doubling it).  The old buffer is not
    // freed until the context ends: a hipHostFree waits for every stream of
    // the device, and other calls may be running on theirs.
    void* grow(void*& p, size_t& cap, size_t bytes) {
        if (bytes > cap) {
            size_t sz = std::max<size_t>(bytes, std::max<size_t>(2 * cap, 1u << 20));
            if (pageable) {
                p = heap_block(sz);     // the old block stays owned by `heap`
            } else {
                if (p) retired.push_back(p);
                p = nullptr;
                RK_HIP(hipHostMalloc(&p, sz, hipHostMallocDefault));
            }
            cap = sz;
        }
        return p;
    }
    std::vector<void*> retired;   // outgrown pinned buffers
    ~Context() {
        if (!pageable) {
            if (pinned) (void)hipHostFree(pinned);
            if (pinned2) (void)hipHostFree(pinned2);
            if (small) (void)hipHostFree(small);
            for (auto& c : up_chunks) (void)hipHostFree(c.first);
            for (void* r : retired) (void)hipHostFree(r);
        }
        arena.release();
        if (stream) (void)hipStreamDestroy(stream);
    }
};

// Phases reported through Context::phase.
enum : int32_t {
    kPhaseIdle = 0,
    kPhasePack = 1,       // building the batch's union CSR
    kPhaseCoarsen = 2,    // level = the level being matched and contracted
    kPhaseInitial = 3,    // initial partitions of the coarsest graphs
    kPhaseRefine = 4,     // level = the level being refined
    kPhaseFinish = 5,     // cuts, best attempts, results to the caller
};

// Acquire a context for the current device (creates one when the pool is
// empty).  Throws rk::Error(-4) when no HIP device is usable.
Context* acquire_context();
void release_context(Context* c);

struct ContextGuard {
    Context* c;
    ContextGuard() : c(acquire_context()) {}
    ~ContextGuard() {
        c->arena.reset();
        c->upload_reset();
        release_context(c);
    }
    Context* operator->() { return c; }
};

}  // namespace rk
