// honei/util/memory_arbiter.hh -- host/device coherence of container memory.
//
// Role of HONEI's MemoryArbiter + MemoryBackend<Tag> pair
// (honei/util/memory_arbiter-impl.hh:404-547, memory_backend.hh:55-274):
// a container's array lives on the host and is mirrored into HBM on first
// device use.  lock(mode, memory_value) returns the pointer valid in that
// memory, copying only when the other side holds newer data; a write lock
// marks the other side stale.  Unlike HONEI there is no worker-thread funnel
// and no global map: each block carries its own state and mutex, and device
// copies stay resident until the host block dies.
#pragma once

#include <honei/backends/mi355x/hsolve.hh>
#include <honei/util/tags.hh>

#include <cstdlib>
#include <cstring>
#include <memory>
#include <mutex>

namespace honei
{
    class MemoryBlock
    {
        private:
            size_t _bytes;
            void * _host;
            void * _device;
            bool _host_valid, _device_valid;
            std::mutex _mutex;

        public:
            explicit MemoryBlock(size_t bytes) :
                _bytes(bytes), _host(nullptr), _device(nullptr), _host_valid(true), _device_valid(false)
            {
                _host = std::calloc(bytes ? bytes : 1, 1);
                if (! _host)
                    throw InternalError("host allocation of " + std::to_string(bytes) + " bytes failed");
            }

            MemoryBlock(const MemoryBlock &) = delete;
            MemoryBlock & operator=(const MemoryBlock &) = delete;

            ~MemoryBlock()
            {
                if (_device)
                    hsolve_free(_device);
                std::free(_host);
            }

            size_t bytes() const { return _bytes; }

            void * lock(LockMode mode, TagValue memory)
            {
                std::lock_guard<std::mutex> guard(_mutex);
                if (memory == tv_gpu_hip)
                {
                    if (! _device)
                        mi355x::check(hsolve_malloc(&_device, _bytes ? _bytes : 1), "MemoryBlock device alloc");
                    if (mode != lm_write_only && ! _device_valid)
                        mi355x::check(hsolve_upload(_host, _device, _bytes, mi355x::stream()), "upload");
                    _device_valid = true;
                    if (mode != lm_read_only)
                        _host_valid = false;
                    return _device;
                }
                if (mode != lm_write_only && ! _host_valid)
                    mi355x::check(hsolve_download(_device, _host, _bytes, mi355x::stream()), "download");
                else if (! _host_valid)
                    mi355x::check(hsolve_synchronize(mi355x::stream()), "synchronize");
                _host_valid = true;
                if (mode != lm_read_only)
                    _device_valid = false;
                return _host;
            }

            void unlock(LockMode)
            {
            }

            bool device_valid() const { return _device_valid; }
            bool host_valid() const { return _host_valid; }
    };
}
