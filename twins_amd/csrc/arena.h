// A device-memory arena: one allocation made at start-up, handed out first-fit in 256-byte
// units, so that a long-running server never calls hipMalloc or hipFree (hipFree may wait for the
// device) while its kernels run. Shared by the TCP and RDMA serving loops.
#pragma once
#include <cstddef>
#include <cstdint>
#include <map>
#include <mutex>
#include <vector>

#include <hip/hip_runtime.h>

// A stream for a serving loop's long-running kernels, on a hardware queue of its own. HIP maps
// streams onto at most GPU_MAX_HW_QUEUES hardware queues (4 by default) and, beyond that, shares
// one queue between streams; a queue runs its kernels in order, so work of a stream that shares
// the persistent kernel's queue waits until that kernel leaves, and a host that waits on such work
// stalls the loop (it publishes no request, so the kernel leaves only when its teams go idle). A
// stream made with a CU mask gets a hardware queue that no other stream shares; the mask here
// names every CU, so kernels on it run where they would on any stream.
inline hipError_t serving_stream_create(hipStream_t *s)
{
    int dev = 0, cus = 0;
    hipError_t e = hipGetDevice(&dev);
    if (e == hipSuccess)
        e = hipDeviceGetAttribute(&cus, hipDeviceAttributeMultiprocessorCount, dev);
    if (e != hipSuccess)
        return e;
    std::vector<uint32_t> mask((size_t)(cus + 31) / 32, 0xFFFFFFFFu);
    return hipExtStreamCreateWithCUMask(s, (uint32_t)mask.size(), mask.data()); // (its size in 32-bit words)
}

struct Arena
{
    char *base = nullptr;
    size_t size = 0;
    std::map<size_t, size_t> free_; // offset -> bytes
    size_t used = 0;
    std::mutex mu;
    void *alloc(size_t n)
    {
        n = (n + 255) & ~(size_t)255;
        std::lock_guard<std::mutex> g(mu);
        for (auto it = free_.begin(); it != free_.end(); ++it)
            if (it->second >= n)
            {
                size_t off = it->first, left = it->second - n;
                free_.erase(it);
                if (left)
                    free_[off + n] = left;
                used += n;
                sizes[off] = n;
                return base + off;
            }
        return nullptr;
    }
    void release(void *p)
    {
        std::lock_guard<std::mutex> g(mu);
        size_t off = (char *)p - base;
        auto sz = sizes.find(off);
        if (sz == sizes.end())
            return;
        size_t n = sz->second;
        sizes.erase(sz);
        used -= n;
        auto nx = free_.lower_bound(off);
        if (nx != free_.end() && nx->first == off + n)
        {
            n += nx->second;
            free_.erase(nx);
        }
        auto pv = free_.lower_bound(off);
        if (pv != free_.begin())
        {
            --pv;
            if (pv->first + pv->second == off)
            {
                pv->second += n;
                return;
            }
        }
        free_[off] = n;
    }
    std::map<size_t, size_t> sizes;
};
