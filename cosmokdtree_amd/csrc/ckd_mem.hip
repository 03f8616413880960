// ckd_mem.hip -- device memory of the library.
//
// Every device array the library owns or uses as scratch (tree arrays,
// build ping-pong buffers, per-level work arrays, query workspaces, CSR
// scan temporaries, sort buffers) comes from a caching allocator with one
// free list per device.  A freed block stays mapped, tagged with an event
// recorded on the stream that freed it; the next allocation of a fitting
// size takes it back (best fit, at most twice the rounded request) and,
// when it runs on another stream, makes that stream wait for the event.
// So repeated builds and queries take their memory without hipMalloc's page
// mapping and without hipFree's device-wide synchronisation and unmapping
// (both cost tens to hundreds of microseconds of host time per call, and
// the GPU idles once the queue drains).  Blocks go back to the driver in
// ckd_release_memory(), and when hipMalloc runs out of memory (then the
// cached blocks are freed and the allocation is tried once more).
#include <cstdint>
#include <map>
#include <mutex>
#include <unordered_map>
#include <vector>

#include "ckd_internal.h"

namespace ckd {
namespace {
constexpr int MAX_DEVICES = 64;

struct Block {
    void *p;
    size_t size;
    hipEvent_t ready;  // recorded on `stream` at the free; null when no work uses the block
    hipStream_t stream;
};

struct DeviceCache {
    std::multimap<size_t, Block> free_blocks;  // by size
    std::vector<hipEvent_t> spare_events;
    size_t cached_bytes = 0;
};

struct Live {
    int dev;
    size_t size;
};

std::mutex g_mu;
DeviceCache g_cache[MAX_DEVICES];
std::unordered_map<void *, Live> g_live;

// Requests are rounded so that blocks of nearby sizes serve each other:
// powers of two up to 1 MiB, multiples of 2 MiB above.
size_t round_size(size_t b) {
    if (b <= 256) return 256;
    if (b <= ((size_t)1 << 20)) {
        size_t r = 256;
        while (r < b) r <<= 1;
        return r;
    }
    const size_t g = (size_t)2 << 20;
    return (b + g - 1) / g * g;
}

// Frees every cached block of device `dev` (the current device) after the
// work on the device has finished.  Called with g_mu held.
hipError_t flush_device(int dev) {
    DeviceCache &c = g_cache[dev];
    if (c.free_blocks.empty() && c.spare_events.empty()) return hipSuccess;
    hipError_t e = hipDeviceSynchronize();
    for (auto &kv : c.free_blocks) {
        hipError_t f = hipFree(kv.second.p);
        if (kv.second.ready) (void)hipEventDestroy(kv.second.ready);
        if (e == hipSuccess) e = f;
    }
    for (hipEvent_t ev : c.spare_events) (void)hipEventDestroy(ev);
    c.free_blocks.clear();
    c.spare_events.clear();
    c.cached_bytes = 0;
    return e;
}

// Device switch for the scope of a call that touches another device's
// memory or events.
struct OnDevice {
    int prev = -1;
    explicit OnDevice(int dev) {
        int cur = 0;
        if (hipGetDevice(&cur) == hipSuccess && cur != dev && hipSetDevice(dev) == hipSuccess) prev = cur;
    }
    ~OnDevice() {
        if (prev >= 0) (void)hipSetDevice(prev);
    }
};
}  // namespace

int dev_alloc(void **p, size_t bytes, hipStream_t s) {
    *p = nullptr;
    int dev = 0;
    CKD_HIP(hipGetDevice(&dev));
    if (dev < 0 || dev >= MAX_DEVICES) {
        set_error("device %d is beyond the %d devices the library supports", dev, MAX_DEVICES);
        return CKD_ERR_ARG;
    }
    const size_t want = round_size(bytes);
    std::lock_guard<std::mutex> lock(g_mu);
    DeviceCache &c = g_cache[dev];
    auto it = c.free_blocks.lower_bound(want);
    if (it != c.free_blocks.end() && it->first <= 2 * want) {
        Block b = it->second;
        c.free_blocks.erase(it);
        c.cached_bytes -= b.size;
        if (b.ready) {
            hipError_t e = b.stream == s ? hipSuccess : hipStreamWaitEvent(s, b.ready, 0);
            c.spare_events.push_back(b.ready);
            if (e != hipSuccess) {
                c.free_blocks.emplace(b.size, Block{b.p, b.size, nullptr, nullptr});
                c.cached_bytes += b.size;
                (void)hipDeviceSynchronize();  // the block is idle after this
                return hip_fail(e, "stream wait for a reused block", __FILE__, __LINE__);
            }
        }
        g_live[b.p] = Live{dev, b.size};
        *p = b.p;
        return CKD_OK;
    }
    hipError_t e = hipMalloc(p, want);
    if (e == hipErrorOutOfMemory) {
        (void)hipGetLastError();
        e = flush_device(dev);
        if (e == hipSuccess) e = hipMalloc(p, want);
    }
    if (e != hipSuccess) {
        *p = nullptr;
        (void)hipGetLastError();
        set_error("device allocation of %zu bytes failed: %s", bytes, hipGetErrorString(e));
        return e == hipErrorOutOfMemory ? CKD_ERR_MEMORY : CKD_ERR_HIP;
    }
    g_live[*p] = Live{dev, want};
    return CKD_OK;
}

static void cache_block(void *p, hipStream_t s, bool in_use) {
    if (!p) return;
    std::lock_guard<std::mutex> lock(g_mu);
    auto it = g_live.find(p);
    if (it == g_live.end()) return;  // not ours (never happens for library buffers)
    const Live l = it->second;
    g_live.erase(it);
    DeviceCache &c = g_cache[l.dev];
    hipEvent_t ev = nullptr;
    if (in_use) {
        OnDevice on(l.dev);
        if (!c.spare_events.empty()) {
            ev = c.spare_events.back();
            c.spare_events.pop_back();
        } else if (hipEventCreateWithFlags(&ev, hipEventDisableTiming) != hipSuccess) {
            ev = nullptr;
        }
        if (!ev || hipEventRecord(ev, s) != hipSuccess) {
            // no event to order reuse by: wait here instead
            (void)hipGetLastError();
            if (ev) c.spare_events.push_back(ev);
            ev = nullptr;
            (void)hipStreamSynchronize(s);
        }
    }
    c.free_blocks.emplace(l.size, Block{p, l.size, ev, s});
    c.cached_bytes += l.size;
}

void dev_free(void *p, hipStream_t s) { cache_block(p, s, true); }

void dev_free_idle(void *p) { cache_block(p, nullptr, false); }

}  // namespace ckd

extern "C" int ckd_release_memory(void) {
    using ckd::hip_fail;
    std::lock_guard<std::mutex> lock(ckd::g_mu);
    int rc = CKD_OK;
    for (int d = 0; d < ckd::MAX_DEVICES; d++) {
        ckd::DeviceCache &c = ckd::g_cache[d];
        if (c.free_blocks.empty() && c.spare_events.empty()) continue;
        ckd::OnDevice on(d);
        hipError_t e = ckd::flush_device(d);
        if (e != hipSuccess && rc == CKD_OK) rc = hip_fail(e, "releasing cached device memory", __FILE__, __LINE__);
    }
    return rc;
}
