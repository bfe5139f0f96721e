// honei/math/methods.hh -- method, cycle, precision and problem tags.
// The names form the solver templates' interface and are kept alike to
// HONEI's (honei/math/methods.hh:25-110) so callers compile unchanged.
#pragma once

namespace honei
{
    namespace methods
    {
        struct CG {};
        struct JAC {};
        struct NONE {};
        struct VAR {};
        struct PCG
        {
            struct JAC;
        };
        // Multigrid<..., CYCLE::V, ...> takes the shape itself; the operator-list
        // MGCycleCreation takes the ::STATIC member
        struct CYCLE
        {
            struct F
            {
                struct V
                {
                    struct STATIC;
                };
                struct W
                {
                    struct STATIC;
                };
            };
            struct W
            {
                struct STATIC;
            };
            struct V
            {
                struct STATIC;
            };
        };
        struct FIXED;
        struct MIXED;
        struct PROLMAT;
        struct NATURAL;
    }

    namespace applications
    {
        struct POISSON;
    }

    namespace boundary_types
    {
        struct DIRICHLET
        {
            struct DIRICHLET_0;
        };
        struct DIRICHLET_NEUMANN;
    }

    namespace io_formats
    {
        struct ELL;
        struct EXP;
        struct DV;
    }
}
