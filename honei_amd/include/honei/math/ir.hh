// honei/math/ir.hh -- iterative refinement around the operator-list
// multigrid (HONEI's honei/math/ir.hh:30-79): the defect, its norm and the
// solution update in the outer type, each correction solved by MGSolver on
// inner-type data.  As in the reference, the inner finest iterate is not
// reset between outer steps: each MGSolver call starts from the previous
// correction.
#pragma once

#include <honei/la/operations.hh>
#include <honei/math/mg.hh>

namespace honei
{
    template <typename OuterTag_, typename InnerTag_, typename NormType_> struct IRSolver
    {
        template <typename OuterMatrixType_, typename OuterVectorType_, typename InnerMatrixType_,
                  typename InnerVectorType_, typename InnerTransferContType_, typename InnerPreconContType_,
                  typename OuterType_, typename InnerType_>
        static void value(OuterMatrixType_ A, OuterVectorType_ b, OuterVectorType_ x,
                          MGData<InnerMatrixType_, InnerVectorType_, InnerTransferContType_, InnerPreconContType_,
                                 InnerType_> & data,
                          OperatorList & cycle, OuterType_ eps, unsigned long max_iters, unsigned long & used_iters)
        {
            if (cycle.size() == 0)
                throw InternalError("IRSolver: the operator list is empty");
            OuterVectorType_ r(x.size());
            OuterVectorType_ c(x.size());
            OuterType_ rnorm_initial(1e16), rnorm_current(1e16);
            for (unsigned long i(0) ; i < max_iters ; ++i)
            {
                Defect<OuterTag_>::value(r, b, A, x);
                if (i == 0)
                    rnorm_initial = NormType_::value(r);
                else
                {
                    rnorm_current = NormType_::value(r);
                    used_iters = i;
                    if (rnorm_current < eps * rnorm_initial)
                        break;
                }
                convert<OuterTag_>(data.b.back(), r);
                MGSolver<InnerTag_, NormType_>::value(data, cycle);
                convert<OuterTag_>(c, data.x.back());
                Sum<OuterTag_>::value(x, c);
            }
        }
    };
}
