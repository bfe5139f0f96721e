// honei/math/transposition.hh -- Transposition<Tag_>::value(SparseMatrixELL)
// (HONEI's honei/math/transposition.hh:30-43): the transpose is assembled on
// the host through a SparseMatrix, so each of its rows holds its entries in
// ascending column order, and converted back to ELL with the ell::threads
// layout.  Setup only; the result uploads on first device use.
#pragma once

#include <honei/la/sparse_matrix.hh>
#include <honei/la/sparse_matrix_ell.hh>

namespace honei
{
    template <typename Tag_> struct Transposition
    {
        template <typename DT_> static SparseMatrixELL<DT_> value(const SparseMatrixELL<DT_> & src)
        {
            SparseMatrix<DT_> target(src.columns(), src.rows());
            const unsigned long * aj = src.Aj().read_elements();
            const DT_ * ax = src.Ax().read_elements();
            const unsigned long * arl = src.Arl().read_elements();
            const unsigned long T = src.threads(), S = src.stride();
            for (unsigned long row(0) ; row < src.rows() ; ++row)
                for (unsigned long k(0) ; k < arl[row] ; ++k)
                    for (unsigned long t(0) ; t < T ; ++t)
                    {
                        const unsigned long pos = row * T + t + k * S;
                        if (ax[pos] != DT_(0))
                            target(aj[pos], row, ax[pos]);
                    }
            return SparseMatrixELL<DT_>(target);
        }
    };
}
