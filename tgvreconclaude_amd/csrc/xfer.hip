// Host <-> device copies of large caller buffers (no reference counterpart).
//
// The callers hand over pageable host memory (the reference's std::vectors,
// NumPy arrays). A pageable copy is staged by the runtime through its own
// buffers by one host thread; here the bytes go through two pinned buffers of
// kChunk bytes that alternate, filled and drained by a small pool of host
// threads, so that the host copy of one chunk overlaps the DMA of the other
// and the host copy itself runs at several threads' memory bandwidth. Every
// DMA is on the context's stream, so later work on the stream sees the data.
#include <algorithm>
#include <condition_variable>
#include <cstring>
#include <functional>
#include <mutex>
#include <thread>

#include "tgvx_internal.h"

namespace tgvx {

namespace {
constexpr size_t kChunk = 32ull << 20;  // bytes per pinned buffer
constexpr size_t kSmall = 2ull << 20;   // below this the runtime's own path
constexpr size_t kSlice = 1ull << 20;   // smallest share of one copy thread
}  // namespace

// A fixed pool of host threads. run(parts, f) calls f(0) .. f(parts - 1),
// part 0 on the calling thread, the others on the pool's threads, and
// returns when all have returned; copy() splits one byte range that way.
class CopyPool {
  public:
    explicit CopyPool(int n) : n_(std::max(1, n)) {
        for (int i = 1; i < n_; ++i) th_.emplace_back([this, i] { loop(i); });
    }
    ~CopyPool() {
        {
            std::lock_guard<std::mutex> lk(mu_);
            stop_ = true;
        }
        cv_.notify_all();
        for (auto& t : th_) t.join();
    }
    int threads() const { return n_; }
    void run(int parts, const std::function<void(int)>& f) {
        parts = std::max(1, std::min(parts, n_));
        if (parts == 1) {
            f(0);
            return;
        }
        {
            std::lock_guard<std::mutex> lk(mu_);
            fn_ = &f;
            parts_ = parts;
            left_ = parts - 1;
            ++gen_;
        }
        cv_.notify_all();
        f(0);
        std::unique_lock<std::mutex> lk(mu_);
        done_.wait(lk, [this] { return left_ == 0; });
        fn_ = nullptr;
    }
    void copy(void* dst, const void* src, size_t bytes) {
        const int parts = (int)std::min<size_t>((size_t)n_, std::max<size_t>(1, bytes / kSlice));
        char* d = static_cast<char*>(dst);
        const char* s = static_cast<const char*>(src);
        const size_t per = (bytes + parts - 1) / parts;
        run(parts, [&](int i) {
            const size_t lo = std::min(bytes, per * (size_t)i), hi = std::min(bytes, lo + per);
            if (hi > lo) std::memcpy(d + lo, s + lo, hi - lo);
        });
    }

  private:
    void loop(int i) {
        uint64_t seen = 0;
        for (;;) {
            const std::function<void(int)>* f;
            {
                std::unique_lock<std::mutex> lk(mu_);
                cv_.wait(lk, [&] { return stop_ || gen_ != seen; });
                if (stop_) return;
                seen = gen_;
                if (i >= parts_) continue;
                f = fn_;
            }
            (*f)(i);
            {
                std::lock_guard<std::mutex> lk(mu_);
                if (--left_ == 0) done_.notify_one();
            }
        }
    }
    const int n_;
    std::vector<std::thread> th_;
    std::mutex mu_;
    std::condition_variable cv_, done_;
    bool stop_ = false;
    uint64_t gen_ = 0;
    const std::function<void(int)>* fn_ = nullptr;
    int parts_ = 0, left_ = 0;
};

namespace {
thread_local BlockCache* t_cache = nullptr;
}  // namespace

BlockCache* current_cache() { return t_cache; }
CacheScope::CacheScope(BlockCache* c) : prev(t_cache) { t_cache = c; }
CacheScope::~CacheScope() { t_cache = prev; }

BlockCache::~BlockCache() { trim(0); }

void* BlockCache::get(size_t b, size_t* got) {
    {
        std::lock_guard<std::mutex> lk(mu_);
        auto it = free_.lower_bound(b);
        if (it != free_.end() && it->first <= b + b / 8 + (2ull << 20)) {
            void* p = it->second;
            *got = it->first;
            held_ -= it->first;
            free_.erase(it);
            return p;
        }
    }
    void* p = nullptr;
    if (hipMalloc(&p, b) != hipSuccess) {
        (void)hipGetLastError();
        trim(0);
        TGVX_HIP(hipMalloc(&p, b));
    }
    *got = b;
    return p;
}

void BlockCache::put(void* p, size_t bytes) {
    std::lock_guard<std::mutex> lk(mu_);
    free_.emplace(bytes, p);
    held_ += bytes;
    while (held_ > kMaxHeld && !free_.empty()) {
        auto last = std::prev(free_.end());
        (void)hipFree(last->second);
        held_ -= last->first;
        free_.erase(last);
    }
}

void BlockCache::trim(size_t keep) {
    std::lock_guard<std::mutex> lk(mu_);
    while (held_ > keep && !free_.empty()) {
        auto last = std::prev(free_.end());
        (void)hipFree(last->second);
        held_ -= last->first;
        free_.erase(last);
    }
}

Staging::~Staging() {
    for (int i = 0; i < 2; ++i) {
        if (done[i]) (void)hipEventSynchronize(done[i]);
        if (done[i]) (void)hipEventDestroy(done[i]);
        if (buf[i]) (void)hipHostFree(buf[i]);
    }
}

namespace {
Staging& staging_of(tgvx_ctx* c) {
    if (!c->stage) c->stage = std::make_unique<Staging>();
    Staging& s = *c->stage;
    if (!s.pool) {
        const unsigned hw = std::thread::hardware_concurrency();
        s.pool = std::make_shared<CopyPool>((int)std::min(8u, std::max(1u, hw / 2)));
    }
    return s;
}
Staging& staging(tgvx_ctx* c) {
    Staging& s = staging_of(c);
    for (int i = 0; i < 2; ++i) {
        if (!s.buf[i]) TGVX_HIP(hipHostMalloc(&s.buf[i], kChunk, hipHostMallocDefault));
        if (!s.done[i]) TGVX_HIP(hipEventCreateWithFlags(&s.done[i], hipEventDisableTiming));
    }
    return s;
}
}  // namespace

int host_threads(tgvx_ctx* c) { return staging_of(c).pool->threads(); }

void parallel_for(tgvx_ctx* c, int parts, const std::function<void(int)>& f) { staging_of(c).pool->run(parts, f); }

int parallel_split(tgvx_ctx* c, size_t count, size_t grain, const std::function<void(int, size_t, size_t)>& f) {
    if (count == 0) return 0;
    CopyPool& pool = *staging_of(c).pool;
    const int parts = (int)std::min<size_t>((size_t)pool.threads(), std::max<size_t>(1, count / std::max<size_t>(grain, 1)));
    const size_t per = (count + parts - 1) / parts;
    pool.run(parts, [&](int k) {
        const size_t lo = std::min(count, per * (size_t)k), hi = std::min(count, lo + per);
        if (hi > lo) f(k, lo, hi);
    });
    return parts;
}

void parallel_range(tgvx_ctx* c, size_t count, size_t grain, const std::function<void(size_t, size_t)>& f) {
    parallel_split(c, count, grain, [&](int, size_t lo, size_t hi) { f(lo, hi); });
}

void h2dv(tgvx_ctx* c, void* dst, const HostPiece* pieces, int count, hipStream_t st) {
    if (!st) st = c->stream;
    // the pieces' start offsets in the concatenated stream
    std::vector<size_t> at((size_t)std::max(count, 0) + 1, 0);
    for (int i = 0; i < count; ++i) at[i + 1] = at[i] + pieces[i].bytes;
    const size_t bytes = at[count > 0 ? count : 0];
    if (bytes == 0) return;
    char* out = static_cast<char*>(dst);
    if (bytes < kSmall) {
        for (int i = 0; i < count; ++i)
            if (pieces[i].bytes)
                TGVX_HIP(hipMemcpyAsync(out + at[i], pieces[i].src, pieces[i].bytes, hipMemcpyHostToDevice, st));
        return;
    }
    Staging& s = staging(c);
    // bytes [lo, hi) of the concatenated stream to d
    auto gather = [&](char* d, size_t lo, size_t hi) {
        size_t i = (size_t)(std::upper_bound(at.begin(), at.end(), lo) - at.begin()) - 1;
        while (lo < hi) {
            while (at[i + 1] <= lo) ++i;
            const size_t len = std::min(hi, at[i + 1]) - lo;
            std::memcpy(d, static_cast<const char*>(pieces[i].src) + (lo - at[i]), len);
            d += len;
            lo += len;
        }
    };
    for (size_t off = 0; off < bytes; off += kChunk) {
        const size_t len = std::min(kChunk, bytes - off);
        const int b = s.next;
        s.next ^= 1;
        if (s.used[b]) TGVX_HIP(hipEventSynchronize(s.done[b]));  // its previous DMA has read it
        char* pin = static_cast<char*>(s.buf[b]);
        const int parts = (int)std::min<size_t>((size_t)s.pool->threads(), std::max<size_t>(1, len / kSlice));
        const size_t per = (len + parts - 1) / parts;
        s.pool->run(parts, [&](int k) {
            const size_t lo = std::min(len, per * (size_t)k), hi = std::min(len, lo + per);
            if (hi > lo) gather(pin + lo, off + lo, off + hi);
        });
        TGVX_HIP(hipMemcpyAsync(out + off, pin, len, hipMemcpyHostToDevice, st));
        TGVX_HIP(hipEventRecord(s.done[b], st));
        s.used[b] = true;
    }
}

void h2d(tgvx_ctx* c, void* dst, const void* src, size_t bytes, hipStream_t st) {
    const HostPiece one{src, bytes};
    h2dv(c, dst, &one, 1, st);
}

void d2h(tgvx_ctx* c, void* dst, const void* src, size_t bytes, hipStream_t st) {
    if (bytes == 0) return;
    if (!st) st = c->stream;
    if (bytes < kSmall) {
        TGVX_HIP(hipMemcpyAsync(dst, src, bytes, hipMemcpyDeviceToHost, st));
        TGVX_HIP(hipStreamSynchronize(st));
        return;
    }
    Staging& s = staging(c);
    const char* in = static_cast<const char*>(src);
    char* out = static_cast<char*>(dst);
    const size_t nch = (bytes + kChunk - 1) / kChunk;
    auto issue = [&](size_t k) {
        const size_t off = k * kChunk, len = std::min(kChunk, bytes - off);
        const int b = (int)(k & 1);
        // an earlier copy's DMA out of this buffer (possibly on another
        // stream) has read it
        if (s.used[b]) TGVX_HIP(hipEventSynchronize(s.done[b]));
        TGVX_HIP(hipMemcpyAsync(s.buf[b], in + off, len, hipMemcpyDeviceToHost, st));
        TGVX_HIP(hipEventRecord(s.done[b], st));
        s.used[b] = true;
    };
    issue(0);
    for (size_t k = 0; k < nch; ++k) {
        // the next chunk's DMA into the other buffer, whose host copy of chunk
        // k - 1 has finished, runs while this chunk is copied out
        if (k + 1 < nch) issue(k + 1);
        const int b = (int)(k & 1);
        TGVX_HIP(hipEventSynchronize(s.done[b]));
        const size_t off = k * kChunk, len = std::min(kChunk, bytes - off);
        s.pool->copy(out + off, s.buf[b], len);
    }
    s.next = (int)(nch & 1);
}

}  // namespace tgvx
