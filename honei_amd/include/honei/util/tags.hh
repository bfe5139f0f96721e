// honei/util/tags.hh -- backend tags of the MI355X build.
//
// Interface kept alike to HONEI (honei/util/tags.hh:36-283): a tag is a type
// with `tag_value`, `memory_value` and `name`, and every operation is a
// class template specialised per tag.  This build provides one backend tag,
// tags::GPU::MI355X, whose memory lives in HBM of one gfx950 device and
// whose operations run hand-written HIP kernels through include/hsolve.h.
#pragma once

namespace honei
{
    namespace tags
    {
        enum TagValue
        {
            tv_cpu = 1,           // host memory (containers' home)
            tv_gpu_hip = 100,     // MI355X device memory
            tv_none
        };

        struct CPU
        {
            static const TagValue tag_value = tv_cpu;
            static const TagValue memory_value = tv_cpu;
            static constexpr const char * name = "cpu";
        };

        struct GPU
        {
            // The MI355X backend: HBM-resident data, gfx950 kernels.
            struct MI355X
            {
                static const TagValue tag_value = tv_gpu_hip;
                static const TagValue memory_value = tv_gpu_hip;
                static constexpr const char * name = "mi355x";
            };
        };
    }

    using tags::TagValue;
    using tags::tv_cpu;
    using tags::tv_gpu_hip;

    // Lock modes of container memory (honei/util/memory_arbiter.hh:41-47).
    enum LockMode
    {
        lm_read_only = 0,
        lm_write_only,
        lm_read_and_write
    };
}
