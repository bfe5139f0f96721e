// honei/backends/mi355x/hsolve.hh -- C++ view of the C-ABI (include/hsolve.h):
// status codes become HONEI exceptions, and one process-wide stream is used
// by the template layer.
#pragma once

#include <hsolve.h>

#include <honei/util/exception.hh>

#include <string>

namespace honei
{
    namespace mi355x
    {
        inline void check(int status, const char * what)
        {
            if (status == HSOLVE_OK)
                return;
            const std::string msg = std::string(what) + ": " + hsolve_last_error();
            if (status == HSOLVE_ERR_SIZE)
                throw VectorSizeDoesNotMatch(0, 0);
            if (status == HSOLVE_ERR_NOT_CONVERGED)
                throw NotConverged(msg);
            throw InternalError(msg);
        }

        // The stream every template-layer operation is issued on (the null
        // stream: operations are ordered and host reads synchronise it).
        inline void * stream() { return nullptr; }

        template <typename DT> struct Suffix;
        template <> struct Suffix<float> { static const int dtype = 0; };
        template <> struct Suffix<double> { static const int dtype = 1; };
    }
}
