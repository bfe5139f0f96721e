// honei/la/sparse_matrix_csr.hh -- SparseMatrixCSR<DT> of the MI355X build.
//
// Host layout and constructors as HONEI's blocked CSR
// (honei/la/sparse_matrix_csr.hh:41-160, sparse_matrix_csr-impl.hh:42-240):
// Ar[r] .. Ar[r + 1] are the blocks of row r, Aj[i] the first column of
// block i, Ax[i * blocksize + k] the value at column Aj[i] + k (0 where the
// row stores nothing).  csr::blocksize (Configuration, default 1) is read by
// the converting constructors.  On first device use the matrix is uploaded
// once into an hsolve_csr handle, which is rebuilt only after a write access
// to the host arrays.
#pragma once

#include <honei/la/dense_vector.hh>
#include <honei/la/sparse_matrix.hh>
#include <honei/la/sparse_matrix_ell.hh>
#include <honei/util/configuration.hh>

#include <memory>
#include <vector>

namespace honei
{
    template <typename DataType_> class SparseMatrixCSR
    {
        private:
            struct Handle
            {
                hsolve_csr h;
                explicit Handle(hsolve_csr hh) : h(hh) {}
                ~Handle() { hsolve_csr_destroy(h); }
            };

            struct Impl
            {
                unsigned long blocksize, rows, columns, used_elements;
                DenseVector<unsigned long> Aj;
                DenseVector<DataType_> Ax;
                DenseVector<unsigned long> Ar;
                std::shared_ptr<Handle> handle;

                Impl(unsigned long bs, unsigned long r, unsigned long c, const DenseVector<unsigned long> & aj,
                     const DenseVector<DataType_> & ax, const DenseVector<unsigned long> & ar, unsigned long ue) :
                    blocksize(bs), rows(r), columns(c), used_elements(ue), Aj(aj), Ax(ax), Ar(ar)
                {
                }
            };

            std::shared_ptr<Impl> _imp;

            // sparse_matrix_csr-impl.hh:98-136: a block starts at the first
            // unplaced entry's column rounded down to a multiple of bs and
            // takes every following entry of the row that lands on one of its
            // bs positions; stored zeros stay stored.
            void _create(const SparseMatrix<DataType_> & src)
            {
                const unsigned long bs = (unsigned long)Configuration::instance()->get_value("csr::blocksize", 1);
                if (bs == 0)
                    throw InternalError("SparseMatrixCSR: csr::blocksize must be at least 1");
                std::vector<unsigned long> ar(src.rows() + 1), aj;
                std::vector<DataType_> ax;
                unsigned long ue(0);
                for (unsigned long row(0) ; row < src.rows() ; ++row)
                {
                    ar[row] = aj.size();
                    const auto & entries = src.row(row);
                    auto it = entries.begin();
                    while (it != entries.end())
                    {
                        const unsigned long start = it->first - it->first % bs;
                        aj.push_back(start);
                        for (unsigned long k(0) ; k < bs ; ++k)
                        {
                            if (it != entries.end() && it->first == start + k)
                            {
                                ax.push_back(it->second);
                                ++ue;
                                ++it;
                            }
                            else
                                ax.push_back(DataType_(0));
                        }
                    }
                }
                ar[src.rows()] = aj.size();
                DenseVector<unsigned long> Ar(src.rows() + 1), Aj(aj.size() ? aj.size() : 1, 0ul);
                DenseVector<DataType_> Ax(ax.size() ? ax.size() : 1, DataType_(0));
                for (unsigned long i(0) ; i < ar.size() ; ++i)
                    Ar[i] = ar[i];
                for (unsigned long i(0) ; i < aj.size() ; ++i)
                    Aj[i] = aj[i];
                for (unsigned long i(0) ; i < ax.size() ; ++i)
                    Ax[i] = ax[i];
                _imp = std::make_shared<Impl>(bs, src.rows(), src.columns(), Aj, Ax, Ar, ue);
            }

            // SparseMatrix(SparseMatrixELL) (sparse_matrix.hh:157-186): every
            // slot holding a non-zero value
            static SparseMatrix<DataType_> rows_of(const SparseMatrixELL<DataType_> & src)
            {
                SparseMatrix<DataType_> sm(src.rows(), src.columns());
                const unsigned long * aj = src.Aj().read_elements();
                const DataType_ * ax = src.Ax().read_elements();
                const unsigned long size = src.Aj().size(), t = src.threads(), stride = src.stride();
                for (unsigned long row(0) ; row < src.rows() ; ++row)
                    for (unsigned long i(row * t) ; i < size ; i += stride)
                        for (unsigned long th(0) ; th < t ; ++th)
                            if (ax[i + th] != DataType_(0))
                                sm(row, aj[i + th], ax[i + th]);
                return sm;
            }

        public:
            typedef DataType_ iDT_;

            /// From CSR vectors (sparse_matrix_csr-impl.hh:139-145).
            SparseMatrixCSR(unsigned long blocksize, unsigned long rows, unsigned long columns,
                            const DenseVector<unsigned long> & Aj, const DenseVector<DataType_> & Ax,
                            const DenseVector<unsigned long> & Ar, unsigned long used_elements) :
                _imp(std::make_shared<Impl>(blocksize, rows, columns, Aj, Ax, Ar, used_elements))
            {
                if (blocksize == 0 || Ar.size() != rows + 1)
                    throw InternalError("SparseMatrixCSR: blocksize 0 or Ar.size() != rows + 1");
            }

            explicit SparseMatrixCSR(const SparseMatrix<DataType_> & src) : _imp() { _create(src); }

            explicit SparseMatrixCSR(const SparseMatrixELL<DataType_> & src) : _imp() { _create(rows_of(src)); }

            SparseMatrixCSR(const SparseMatrixCSR &) = default;
            SparseMatrixCSR & operator=(const SparseMatrixCSR &) = default;

            unsigned long blocksize() const { return _imp->blocksize; }
            unsigned long rows() const { return _imp->rows; }
            unsigned long columns() const { return _imp->columns; }
            unsigned long size() const { return _imp->rows * _imp->columns; }
            unsigned long used_elements() const { return _imp->used_elements; }

            /// Host arrays; a non-const access drops the device handle.
            DenseVector<unsigned long> & Aj() { _imp->handle.reset(); return _imp->Aj; }
            DenseVector<DataType_> & Ax() { _imp->handle.reset(); return _imp->Ax; }
            DenseVector<unsigned long> & Ar() { _imp->handle.reset(); return _imp->Ar; }
            const DenseVector<unsigned long> & Aj() const { return _imp->Aj; }
            const DenseVector<DataType_> & Ax() const { return _imp->Ax; }
            const DenseVector<unsigned long> & Ar() const { return _imp->Ar; }

            /// Element (row, column) (sparse_matrix_csr-impl.hh:228-240).
            DataType_ operator()(unsigned long row, unsigned long column) const
            {
                const unsigned long * ar = _imp->Ar.read_elements();
                const unsigned long * aj = _imp->Aj.read_elements();
                const DataType_ * ax = _imp->Ax.read_elements();
                const unsigned long bs = blocksize();
                for (unsigned long i(ar[row]) ; i < ar[row + 1] ; ++i)
                    if (column >= aj[i] && column < aj[i] + bs)
                        return ax[i * bs + (column - aj[i])];
                return DataType_(0);
            }

            /// The device-resident handle (uploaded on first use).
            hsolve_csr device_handle() const
            {
                if (! _imp->handle)
                {
                    hsolve_csr h = nullptr;
                    mi355x::check(hsolve_csr_create(mi355x::Suffix<DataType_>::dtype, rows(), columns(), blocksize(),
                                                    (const uint64_t *)_imp->Ar.read_elements(),
                                                    (const uint64_t *)_imp->Aj.read_elements(),
                                                    _imp->Ax.read_elements(), used_elements(), &h),
                                  "SparseMatrixCSR upload");
                    _imp->handle = std::make_shared<Handle>(h);
                }
                return _imp->handle->h;
            }

            void lock(LockMode mode) const
            {
                _imp->Aj.lock(mode);
                _imp->Ax.lock(mode);
                _imp->Ar.lock(mode);
            }

            void unlock(LockMode mode) const
            {
                _imp->Aj.unlock(mode);
                _imp->Ax.unlock(mode);
                _imp->Ar.unlock(mode);
            }

            SparseMatrixCSR copy() const
            {
                SparseMatrixCSR r(*this);
                r._imp = std::make_shared<Impl>(blocksize(), rows(), columns(), _imp->Aj.copy(), _imp->Ax.copy(),
                                                _imp->Ar.copy(), used_elements());
                return r;
            }
    };
}
