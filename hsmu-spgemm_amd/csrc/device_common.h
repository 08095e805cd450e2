// device_common.h -- wave64 / workgroup helpers shared by the HIP kernel files.
#pragma once

#include <hip/hip_runtime.h>

namespace hsmu {

typedef unsigned long long u64;

#define DEVI __device__ __forceinline__

// ---- wave64 helpers ---------------------------------------------------------------

// number of set bits of m below this lane (v_mbcnt_lo/hi)
DEVI int mbcnt(u64 m)
{
    return __builtin_amdgcn_mbcnt_hi((unsigned)(m >> 32), __builtin_amdgcn_mbcnt_lo((unsigned)m, 0u));
}

DEVI u64 below(int b) { return (1ull << b) - 1ull; }

// The first active lane's value, in a scalar register: for values that are the same in
// every lane (loaded from one address), so that no per-lane copy is kept.
DEVI int uniform(int v) { return __builtin_amdgcn_readfirstlane(v); }
DEVI unsigned uniform(unsigned v) { return (unsigned)__builtin_amdgcn_readfirstlane((int)v); }
DEVI long long uniform64(long long v)
{
    const unsigned lo = (unsigned)__builtin_amdgcn_readfirstlane((int)(unsigned)(v & 0xffffffffLL));
    const unsigned hi = (unsigned)__builtin_amdgcn_readfirstlane((int)(v >> 32));
    return (long long)(((unsigned long long)hi << 32) | lo);
}
DEVI u64 uniform64(u64 v) { return (u64)uniform64((long long)v); }

// Reductions and scans over the 64 lanes of a wave (every lane must take part).
template <typename T> DEVI T wave_sum(T v)
{
#pragma unroll
    for (int o = 32; o > 0; o >>= 1)
        v += __shfl_xor(v, o, 64);
    return v;
}

template <typename T> DEVI T wave_max(T v)
{
#pragma unroll
    for (int o = 32; o > 0; o >>= 1) {
        const T u = __shfl_xor(v, o, 64);
        v = u > v ? u : v;
    }
    return v;
}

template <typename T> DEVI T wave_incl_scan(T v)
{
    const int lane = threadIdx.x & 63;
#pragma unroll
    for (int o = 1; o < 64; o <<= 1) {
        const T u = __shfl_up(v, o, 64);
        if (lane >= o)
            v += u;
    }
    return v;
}

// Map the hardware block index so that each XCD (blocks are dealt round-robin over the
// nx XCDs of the device: 8 on an MI355X) works on a contiguous range of logical blocks:
// neighbouring rows, which share rows of B, then share one XCD's L2.
DEVI int xcd_block(int b, int n, int nx)
{
    const int x = b % nx, q = b / nx, per = n / nx, rem = n % nx;
    return x * per + (x < rem ? x : rem) + q;
}

// Device work queue of the persistent row kernels: the n row groups of a launch are cut
// into nx contiguous ranges, one per XCD, each with a counter q[x] (zero at launch) that
// hands out chunks of `chunk` consecutive groups.  A workgroup takes the next chunk of its
// own XCD's range (xq starts at blockIdx.x % nx) and, once that range is used up, of the
// following ranges in turn, so the groups one XCD works on stay neighbours (shared rows
// of B in its L2) until the queues drain.  Thread 0 pulls; the chunk [s[0], s[1]) goes to
// every thread through LDS, s[0] = n when nothing is left.  The caller's loop body must
// end with a workgroup barrier before the next pull.
DEVI void pull_chunk(unsigned *q, int n, int chunk, int nx, int &xq, int *s)
{
    if (threadIdx.x == 0) {
        int got = n, end = n;
        for (int k = 0; k < nx; ++k) {
            const int x = (xq + k) % nx;
            const int lo = (int)((long long)n * x / nx), hi = (int)((long long)n * (x + 1) / nx);
            const unsigned nch = (unsigned)((hi - lo + chunk - 1) / chunk);
            if (__hip_atomic_load(&q[x], __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT) >= nch)
                continue;
            const unsigned v = atomicAdd(&q[x], 1u);
            if (v < nch) {
                got = lo + (int)v * chunk;
                end = got + chunk < hi ? got + chunk : hi;
                xq = x;
                break;
            }
        }
        s[0] = got;
        s[1] = end;
    }
    __syncthreads();
}

// The row groups a workgroup of a row kernel works on: with a queue, chunks pulled from
// it; without, groups xcd_block(blockIdx) + k * gridDim (one group at a time).
// nx: XCDs of the device (1..MAX_XCDS).
struct GroupWalk {
    unsigned *q;
    int n, chunk, nx, xq, grp, end;
    int *s; // two ints of LDS

    DEVI GroupWalk(unsigned *queue, int ngroups, int *lds2, int nxcd)
        : q(queue), n(ngroups), nx(nxcd), xq(blockIdx.x % nxcd), s(lds2)
    {
        // chunks of about an eighth of a workgroup's share of its XCD's range
        const long long per = (long long)ngroups / ((long long)gridDim.x * 8);
        chunk = per > 1 ? (int)per : 1;
        if (q) {
            pull_chunk(q, n, chunk, nx, xq, s);
            grp = s[0];
            end = s[1];
        } else {
            grp = xcd_block(blockIdx.x, gridDim.x, nx);
            end = grp + 1;
        }
    }
    DEVI bool more() const { return grp < n; }
    DEVI void next()
    {
        if (++grp < end)
            return;
        if (q) {
            pull_chunk(q, n, chunk, nx, xq, s);
            grp = s[0];
            end = s[1];
        } else {
            grp += (int)gridDim.x - 1;
            end = grp + 1;
        }
    }
};

// Sum over a group of G consecutive lanes.  G <= 64: shuffles; G > 64: through LDS
// (every thread of the block must call it).
template <int G, typename T> DEVI T group_sum(T v, T *red)
{
    if constexpr (G <= 64) {
#pragma unroll
        for (int o = G / 2; o > 0; o >>= 1)
            v += __shfl_xor(v, o, G);
        return v;
    } else {
#pragma unroll
        for (int o = 32; o > 0; o >>= 1)
            v += __shfl_xor(v, o, 64);
        const int w = threadIdx.x >> 6;
        __syncthreads();
        if ((threadIdx.x & 63) == 0)
            red[w] = v;
        __syncthreads();
        const int w0 = (threadIdx.x / G) * (G / 64);
        T s = 0;
#pragma unroll
        for (int k = 0; k < G / 64; ++k)
            s += red[w0 + k];
        return s;
    }
}

template <int G, typename T> DEVI T group_min(T v)
{
    static_assert(G <= 64, "group_min is a wave operation");
#pragma unroll
    for (int o = G / 2; o > 0; o >>= 1) {
        T y = __shfl_xor(v, o, G);
        v = y < v ? y : v;
    }
    return v;
}

template <int G, typename T> DEVI T group_max(T v)
{
    static_assert(G <= 64, "group_max is a wave operation");
#pragma unroll
    for (int o = G / 2; o > 0; o >>= 1) {
        T y = __shfl_xor(v, o, G);
        v = y > v ? y : v;
    }
    return v;
}

// Exclusive scan over a group of G lanes; total = group sum.  G > 64 goes through LDS
// (every thread of the block must call it).
template <int G, typename T> DEVI T group_excl_scan(T v, T *red, T &total)
{
    constexpr int W = G < 64 ? G : 64;
    const int l = threadIdx.x & (W - 1);
    T x = v;
#pragma unroll
    for (int o = 1; o < W; o <<= 1) {
        T y = __shfl_up(x, o, W);
        if (l >= o)
            x += y;
    }
    if constexpr (G <= 64) {
        total = __shfl(x, W - 1, W);
        return x - v;
    } else {
        const int w = threadIdx.x >> 6;
        __syncthreads();
        if (l == 63)
            red[w] = x;
        __syncthreads();
        const int w0 = (threadIdx.x / G) * (G / 64);
        T before = 0, tot = 0;
#pragma unroll
        for (int k = 0; k < G / 64; ++k) {
            const T r = red[w0 + k];
            if (w0 + k < w)
                before += r;
            tot += r;
        }
        total = tot;
        return before + x - v;
    }
}

// Workgroup barrier that also orders global memory for the whole workgroup at device
// scope: prior stores are complete in L2 and this CU's L1 is invalidated afterwards,
// so plain loads see values other waves wrote with stores or L2 atomics.
DEVI void block_sync_global()
{
    __builtin_amdgcn_fence(__ATOMIC_RELEASE, "agent");
    __syncthreads();
    __builtin_amdgcn_fence(__ATOMIC_ACQUIRE, "agent");
}

} // namespace hsmu
