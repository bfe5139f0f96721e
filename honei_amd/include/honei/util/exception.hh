// honei/util/exception.hh -- exceptions of the template layer, named as in
// HONEI (honei/util/exception.hh, honei/la/vector_error.hh): size mismatches
// throw VectorSizeDoesNotMatch, backend failures InternalError.
#pragma once

#include <stdexcept>
#include <string>

namespace honei
{
    class Exception : public std::runtime_error
    {
        public:
            explicit Exception(const std::string & message) : std::runtime_error(message) {}
            const std::string message() const { return what(); }
    };

    class InternalError : public Exception
    {
        public:
            explicit InternalError(const std::string & message) : Exception("Internal error: " + message) {}
    };

    class VectorSizeDoesNotMatch : public Exception
    {
        public:
            VectorSizeDoesNotMatch(unsigned long size, unsigned long expected_size) :
                Exception("Vector size '" + std::to_string(size) + "' does not match expected size '" +
                          std::to_string(expected_size) + "'")
            {
            }
    };

    class NotConverged : public Exception
    {
        public:
            explicit NotConverged(const std::string & message) : Exception(message) {}
    };
}
