// A device-memory arena: one allocation made at start-up, handed out first-fit in 256-byte
// units, so that a long-running server never calls hipMalloc or hipFree (hipFree may wait for the
// device) while its kernels run. Shared by the TCP and RDMA serving loops.
#pragma once
#include <cstddef>
#include <cstdint>
#include <map>
#include <mutex>

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
