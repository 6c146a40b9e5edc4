// Device-coherent accesses, shared by the TCP and RDMA models: relaxed atomics at agent scope,
// which every XCD sees in one order. The data a team's workgroups hand each other from slot to
// slot (link frames and their count words, a world's ticket and pending-slot flags, event-chain
// tail nodes, a sender's end state) goes only through these, so a slot's team barrier needs no
// acquire/release fence: on a GPU of several XCDs an agent-scope fence writes back and
// invalidates the L2 (buffer_wbl2 / buffer_inv), and every slot would start with cold caches for
// the entities' own records. CPU builds: plain accesses.
#pragma once
#include <cstdint>

namespace twc
{

template <typename T>
#if defined(__HIPCC__)
__host__ __device__
#endif
inline T coh_load(const T *p)
{
    static_assert(sizeof(T) == 4 || sizeof(T) == 8, "a 4- or 8-byte word");
#if defined(__HIP_DEVICE_COMPILE__)
    return __hip_atomic_load(p, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
#else
    return *p;
#endif
}
template <typename T>
#if defined(__HIPCC__)
__host__ __device__
#endif
inline void coh_store(T *p, T v)
{
    static_assert(sizeof(T) == 4 || sizeof(T) == 8, "a 4- or 8-byte word");
#if defined(__HIP_DEVICE_COMPILE__)
    __hip_atomic_store(p, v, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
#else
    *p = v;
#endif
}
#if defined(__HIP_DEVICE_COMPILE__)
// (the same through pointers that carry the global address space: global, not flat, instructions)
template <typename T>
inline __device__ T coh_load(const __attribute__((address_space(1))) T *p)
{
    static_assert(sizeof(T) == 4 || sizeof(T) == 8, "a 4- or 8-byte word");
    return __hip_atomic_load(p, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
}
template <typename T>
inline __device__ void coh_store(__attribute__((address_space(1))) T *p, T v)
{
    static_assert(sizeof(T) == 4 || sizeof(T) == 8, "a 4- or 8-byte word");
    __hip_atomic_store(p, v, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
}
#endif
// a record of whole 8-byte words, word by word (the words' loads or stores issued together)
template <typename R>
#if defined(__HIPCC__)
__host__ __device__
#endif
inline R coh_get(const R *p)
{
    static_assert(sizeof(R) % 8 == 0, "whole 8-byte words");
    R r;
    const uint64_t *q = reinterpret_cast<const uint64_t *>(p);
#pragma unroll
    for (int i = 0; i < (int)(sizeof(R) / 8); i++)
    {
        const uint64_t w = coh_load(q + i);
        __builtin_memcpy(reinterpret_cast<char *>(&r) + 8 * i, &w, 8);
    }
    return r;
}
template <typename R>
#if defined(__HIPCC__)
__host__ __device__
#endif
inline void coh_put(R *p, const R &r)
{
    static_assert(sizeof(R) % 8 == 0, "whole 8-byte words");
    uint64_t *d = reinterpret_cast<uint64_t *>(p);
#pragma unroll
    for (int i = 0; i < (int)(sizeof(R) / 8); i++)
    {
        uint64_t w;
        __builtin_memcpy(&w, reinterpret_cast<const char *>(&r) + 8 * i, 8);
        coh_store(d + i, w);
    }
}

#if defined(__HIP_DEVICE_COMPILE__)
// 16 bytes at byte offset `off` of a buffer, with the cache policy of an agent-scope atomic store
// (sc1: written through, device-coherent): half the store instructions of 8-byte atomic stores
inline __device__ void coh_store16(__amdgpu_buffer_rsrc_t r, uint32_t off, uint64_t a, uint64_t b)
{
    typedef unsigned v4u __attribute__((ext_vector_type(4)));
    const v4u v = {(unsigned)a, (unsigned)(a >> 32), (unsigned)b, (unsigned)(b >> 32)};
    __builtin_amdgcn_raw_buffer_store_b128(v, r, off, 0, 16 /* sc1 */);
}
// a raw buffer over `bytes` bytes from p (at most 2^31), made wave-uniform: p and bytes must be
// the same on every lane
inline __device__ __amdgpu_buffer_rsrc_t coh_buffer(const void *p, uint32_t bytes)
{
    const uint64_t a = (uint64_t)p;
    // (readfirstlane returns an int: each half is widened through uint32_t, not sign-extended)
    const uint64_t lo = (uint32_t)__builtin_amdgcn_readfirstlane((unsigned)a), hi = (uint32_t)__builtin_amdgcn_readfirstlane((unsigned)(a >> 32));
    const uint64_t u = lo | hi << 32;
    return __builtin_amdgcn_make_buffer_rsrc((void *)u, 0, (int)__builtin_amdgcn_readfirstlane(bytes), 0x00020000);
}
#endif
#if defined(__HIP_DEVICE_COMPILE__)
// (records through pointers that carry the global address space)
template <typename R>
inline __device__ R coh_get(const __attribute__((address_space(1))) R *p)
{
    static_assert(sizeof(R) % 8 == 0, "whole 8-byte words");
    R r;
    const __attribute__((address_space(1))) uint64_t *q = reinterpret_cast<const __attribute__((address_space(1))) uint64_t *>(p);
#pragma unroll
    for (int i = 0; i < (int)(sizeof(R) / 8); i++)
    {
        const uint64_t w = coh_load(q + i);
        __builtin_memcpy(reinterpret_cast<char *>(&r) + 8 * i, &w, 8);
    }
    return r;
}
template <typename R>
inline __device__ void coh_put(__attribute__((address_space(1))) R *p, const R &r)
{
    static_assert(sizeof(R) % 8 == 0, "whole 8-byte words");
    __attribute__((address_space(1))) uint64_t *d = reinterpret_cast<__attribute__((address_space(1))) uint64_t *>(p);
#pragma unroll
    for (int i = 0; i < (int)(sizeof(R) / 8); i++)
    {
        uint64_t w;
        __builtin_memcpy(&w, reinterpret_cast<const char *>(&r) + 8 * i, 8);
        coh_store(d + i, w);
    }
}
#endif

} // namespace twc
