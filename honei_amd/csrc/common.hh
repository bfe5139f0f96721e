// common.hh -- error plumbing, stream helpers and launch-shape constants
// shared by the gfx950 kernels and the C-ABI of libhsolve.
#pragma once

#include <hip/hip_runtime.h>

#include <cstdint>
#include <cstdio>
#include <stdexcept>
#include <string>

#include "hsolve.h"

namespace hsolve
{
    // Error raised inside the library; converted to a status code at the C-ABI.
    struct Error : public std::runtime_error
    {
        int code;
        Error(int c, const std::string & msg) : std::runtime_error(msg), code(c) {}
    };

    void set_last_error(const std::string & msg);

    inline void hip_check(hipError_t e, const char * what, const char * file, int line)
    {
        if (e != hipSuccess)
        {
            throw Error(HSOLVE_ERR_HIP, std::string(what) + " failed: " + hipGetErrorString(e) + " at " +
                    file + ":" + std::to_string(line));
        }
    }

    inline void require(bool cond, int code, const std::string & msg)
    {
        if (! cond)
            throw Error(code, msg);
    }

    inline hipStream_t as_stream(void * s) { return reinterpret_cast<hipStream_t>(s); }

    // Launch-shape constants for gfx950: 64-wide waves, 256 CUs in 8 XCDs.
    constexpr int kWave = 64;
    constexpr int kBlock = 256;
    constexpr int kReduceBlocks = 2048;   // partial sums per one-pass reduction

    // SplitMix64 finaliser: output number i of the stream keyed by (seed, s)
    // is mix64(key + G (i + 1)) with key = mix64(seed + G (s + 1)),
    // G = 0x9E3779B97F4A7C15.
    __host__ __device__ inline uint64_t mix64(uint64_t z)
    {
        z = (z ^ (z >> 30)) * 0xBF58476D1CE4E5B9ull;
        z = (z ^ (z >> 27)) * 0x94D049BB133111EBull;
        return z ^ (z >> 31);
    }
    constexpr uint64_t kGolden = 0x9E3779B97F4A7C15ull;

    // Workgroup barrier for LDS traffic only: waits for this wave's LDS
    // operations, not for its outstanding global loads (which
    // __syncthreads() would drain), so loads issued before it stay in flight.
    __device__ inline void lds_barrier()
    {
        asm volatile("s_waitcnt lgkmcnt(0)\n\ts_barrier" ::: "memory");
    }

    // The same pointer in the global address space.  Device pointers that a
    // kernel reads from memory (slot tables, pointer arrays) are generic, and
    // the compiler then uses flat instructions, which also count against the
    // LDS counter: every wait on them drains LDS traffic as well and the
    // compiler falls back to waiting for all outstanding loads.
    template <typename T> using gptr = __attribute__((address_space(1))) T *;
    template <typename T> __device__ __forceinline__ gptr<T> as_global(T * p) { return (gptr<T>)p; }
    // base[i] and base[i] = v through a 32-bit byte offset (i * sizeof(T) <
    // 2^32): the address is a wave-uniform base plus a 32-bit per-lane
    // offset, which the load and store instructions take as they are, with
    // no 64-bit address arithmetic per lane
    template <typename T> __device__ __forceinline__ T ld32(gptr<const T> base, uint32_t i)
    {
        return *(gptr<const T>)((gptr<const char>)base + (uint32_t)(i * (uint32_t)sizeof(T)));
    }
    template <typename T> __device__ __forceinline__ void st32(gptr<T> base, uint32_t i, T v)
    {
        *(gptr<T>)((gptr<char>)base + (uint32_t)(i * (uint32_t)sizeof(T))) = v;
    }

    inline unsigned grid_for(uint64_t work, unsigned per_block)
    {
        uint64_t g = (work + per_block - 1) / per_block;
        if (g == 0)
            g = 1;
        return (unsigned)g;
    }

    // Per-stream scratch of the dot / norm reductions: the block partials,
    // summed by a following one-block launch (no memset, so a reduction can
    // be replayed inside a captured graph).
    struct ReduceScratch
    {
        double * partials;      // kReduceBlocks doubles
    };
    ReduceScratch & reduce_scratch(hipStream_t stream);
}

#define HIP_CHECK(expr) ::hsolve::hip_check((expr), #expr, __FILE__, __LINE__)

// Wrap a C-ABI body: exceptions become status codes and a stored message.
#define HSOLVE_API_BEGIN try {
#define HSOLVE_API_END                                                         \
    }                                                                          \
    catch (const ::hsolve::Error & e)                                          \
    {                                                                          \
        ::hsolve::set_last_error(e.what());                                    \
        return e.code;                                                         \
    }                                                                          \
    catch (const std::bad_alloc &)                                             \
    {                                                                          \
        ::hsolve::set_last_error("host allocation failed");                    \
        return HSOLVE_ERR_ALLOC;                                               \
    }                                                                          \
    catch (const std::exception & e)                                           \
    {                                                                          \
        ::hsolve::set_last_error(e.what());                                    \
        return HSOLVE_ERR_INVALID;                                             \
    }                                                                          \
    return HSOLVE_OK;
