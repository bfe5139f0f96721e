// sim.hh -- a host-only stand-in for the device under the request server's
// host code (honei_amd/csrc/serve.hip), for AddressSanitizer runs on a CPU.
//
// sim_hip.cc implements the HIP runtime calls that serve.hip makes.  A
// stream is a queue of deferred operations that run only when the host
// waits on the stream, queries an event recorded on it, or at random
// moments of other calls (so work of different streams interleaves in
// varying orders).  Device and pinned allocations are ordinary heap blocks:
// an operation that touches freed slot memory is a heap-use-after-free.
// While an asynchronous copy from or to pinned memory is queued, its host
// range is poisoned, so any host access to it before the copy ran is an
// ASan error, as it would be a race on the device.  Calls that HIP forbids
// during a thread-local stream capture (allocation, synchronous copies,
// synchronizing a capturing stream) abort with a message.
//
// sim_mg.cc stands in for the engine behind the server (mg_engine.hh): a
// slot's "solve" takes a number of ticks fixed by its right-hand side, and
// every tick reads the slot's right-hand side and writes its whole result
// array through the device table, so stale table entries and freed slots
// are touched.
#pragma once

#include <hip/hip_runtime_api.h>

#include <cstdint>
#include <functional>

namespace sim
{
    // Queues an operation on a stream (or on the graph the stream captures).
    void enqueue(hipStream_t st, std::function<void()> op);
    // Seeds the random device progress.
    void seed(uint64_t s);
    // Reports a broken contract and aborts.
    [[noreturn]] void fail(const char * fmt, ...);
    // Bytes currently held by device allocations.
    uint64_t device_bytes();

    // The simulated engine's right-hand side entry i for a seeded request,
    // the ticks a solve of right-hand side b takes, and the result entry j.
    double rhs_value(uint64_t seed, uint64_t i);
    uint64_t ticks_for(const double * b);
    inline double result_value(double bj, uint64_t ticks) { return 2.0 * bj + (double)ticks; }
}

extern "C" const char * sim_last_error();
