// honei/math/vectorpool.hh -- create_vectorpool (HONEI's
// honei/math/vectorpool.hh): `count` fresh vectors of one size.
#pragma once

#include <vector>

namespace honei
{
    template <typename VectorType_>
    std::vector<VectorType_> create_vectorpool(unsigned long count, unsigned long size)
    {
        std::vector<VectorType_> pool;
        for (unsigned long i(0) ; i < count ; ++i)
            pool.push_back(VectorType_(size));
        return pool;
    }
}
