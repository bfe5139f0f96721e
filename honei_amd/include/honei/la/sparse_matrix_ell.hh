// honei/la/sparse_matrix_ell.hh -- SparseMatrixELL<DT> of the MI355X build.
//
// Host layout and constructors as HONEI's (honei/la/sparse_matrix_ell.hh:44-129,
// sparse_matrix_ell-impl.hh:42-354): element k of lane t of row r at
// r*threads + t + k*stride, stride = 32*ceil(rows*threads/32), uint64
// column indices, Arl[r] = row step count, padding (column 0, value 0).
// On first device use the matrix is uploaded once into an hsolve_ell handle
// (backend layout with compressed indices); the handle stays resident and
// is rebuilt only after a write access to the host arrays.
#pragma once

#include <honei/la/dense_vector.hh>
#include <honei/la/sparse_matrix.hh>
#include <honei/util/configuration.hh>

#include <cmath>
#include <memory>

namespace honei
{
    template <typename DataType_> class SparseMatrixELL
    {
        private:
            struct Handle
            {
                hsolve_ell h;
                explicit Handle(hsolve_ell hh) : h(hh) {}
                ~Handle() { hsolve_ell_destroy(h); }
            };

            struct Impl
            {
                unsigned long rows, columns, stride, num_cols_per_row, threads;
                DenseVector<unsigned long> Aj;
                DenseVector<DataType_> Ax;
                DenseVector<unsigned long> Arl;
                std::shared_ptr<Handle> handle;

                Impl(unsigned long r, unsigned long c, unsigned long s, unsigned long n, unsigned long t,
                     const DenseVector<unsigned long> & aj, const DenseVector<DataType_> & ax,
                     const DenseVector<unsigned long> & arl) :
                    rows(r), columns(c), stride(s), num_cols_per_row(n), threads(t), Aj(aj), Ax(ax), Arl(arl)
                {
                }
            };

            std::shared_ptr<Impl> _imp;

            // sparse_matrix_ell-impl.hh:186-207: count slots until the first zero value
            static DenseVector<unsigned long> row_length(unsigned long rows, unsigned long stride, unsigned long threads,
                                                         const DenseVector<DataType_> & Ax)
            {
                DenseVector<unsigned long> rl(rows, 0ul);
                const DataType_ * ax = Ax.read_elements();
                unsigned long * r = rl.elements();
                const unsigned long size = Ax.size();
                for (unsigned long row(0) ; row < rows ; ++row)
                {
                    unsigned long count(0);
                    bool stop(false);
                    for (unsigned long i(row * threads) ; i < size && ! stop ; i += stride)
                        for (unsigned long t(0) ; t < threads ; ++t)
                        {
                            if (ax[i + t] == DataType_(0))
                            {
                                stop = true;
                                break;
                            }
                            if (t == 0)
                                ++count;
                        }
                    r[row] = count;
                }
                return rl;
            }

        public:
            typedef DataType_ iDT_;

            /// Raw constructor (as used by MatrixIO<io_formats::ELL>::read_matrix).
            SparseMatrixELL(unsigned long rows, unsigned long columns, unsigned long stride,
                            unsigned long num_cols_per_row, const DenseVector<unsigned long> & Aj,
                            const DenseVector<DataType_> & Ax, unsigned long threads) :
                _imp(std::make_shared<Impl>(rows, columns, stride, num_cols_per_row, threads, Aj, Ax,
                                            row_length(rows, stride, threads, Ax)))
            {
            }

            /// From a SparseMatrix with ell::threads lanes per row
            /// (sparse_matrix_ell-impl.hh:65-117): zeros are skipped, Arl counts
            /// stored entries.
            explicit SparseMatrixELL(const SparseMatrix<DataType_> & src) : _imp()
            {
                const unsigned long threads = (unsigned long)Configuration::instance()->get_value("ell::threads", 1);
                const unsigned long rows = src.rows();
                DenseVector<unsigned long> arl(rows, 0ul);
                unsigned long ncpr = 1;
                unsigned long * a = arl.elements();
                for (unsigned long i(0) ; i < rows ; ++i)
                {
                    a[i] = (unsigned long)std::ceil(double(src.used_elements(i)) / double(threads));
                    if (src.used_elements(i) > ncpr)
                        ncpr = src.used_elements(i);
                }
                ncpr = (unsigned long)std::ceil(double(ncpr) / double(threads));
                const unsigned long alignment(32);
                const unsigned long stride = alignment * (((rows * threads) + alignment - 1) / alignment);
                DenseVector<unsigned long> Aj(ncpr * stride, 0ul);
                DenseVector<DataType_> Ax(ncpr * stride, DataType_(0));
                unsigned long * aj = Aj.elements();
                DataType_ * ax = Ax.elements();
                for (unsigned long row(0) ; row < rows ; ++row)
                {
                    unsigned long target(0);
                    for (const auto & e : src.row(row))
                    {
                        if (e.second != DataType_(0))
                        {
                            const unsigned long pos = (target % threads) + row * threads + (target / threads) * stride;
                            aj[pos] = e.first;
                            ax[pos] = e.second;
                            ++target;
                        }
                    }
                }
                _imp = std::make_shared<Impl>(rows, src.columns(), stride, ncpr, threads, Aj, Ax, arl);
            }

            SparseMatrixELL(const SparseMatrixELL &) = default;
            SparseMatrixELL & operator=(const SparseMatrixELL &) = default;

            unsigned long rows() const { return _imp->rows; }
            unsigned long columns() const { return _imp->columns; }
            unsigned long size() const { return _imp->rows * _imp->columns; }
            unsigned long stride() const { return _imp->stride; }
            unsigned long num_cols_per_row() const { return _imp->num_cols_per_row; }
            unsigned long threads() const { return _imp->threads; }

            unsigned long used_elements() const
            {
                unsigned long ue(0);
                const DataType_ * ax = _imp->Ax.read_elements();
                for (unsigned long i(0) ; i < _imp->Ax.size() ; ++i)
                    if (ax[i] != DataType_(0))
                        ++ue;
                return ue;
            }

            /// Host arrays; a non-const access drops the device handle.
            DenseVector<unsigned long> & Aj() { _imp->handle.reset(); return _imp->Aj; }
            DenseVector<DataType_> & Ax() { _imp->handle.reset(); return _imp->Ax; }
            DenseVector<unsigned long> & Arl() { _imp->handle.reset(); return _imp->Arl; }
            const DenseVector<unsigned long> & Aj() const { return _imp->Aj; }
            const DenseVector<DataType_> & Ax() const { return _imp->Ax; }
            const DenseVector<unsigned long> & Arl() const { return _imp->Arl; }

            /// Element (row, column) (sparse_matrix_ell-impl.hh:321-333).
            DataType_ operator()(unsigned long row, unsigned long column) const
            {
                const unsigned long * aj = _imp->Aj.read_elements();
                const DataType_ * ax = _imp->Ax.read_elements();
                const unsigned long max = _imp->Arl.read_elements()[row];
                for (unsigned long i(row * threads()), j(0) ; j < max && aj[i] <= column ; i += stride(), ++j)
                    for (unsigned long t(0) ; t < threads() ; ++t)
                        if (aj[i + t] == column)
                            return ax[i + t];
                return DataType_(0);
            }

            /// The device-resident handle (uploaded on first use).
            hsolve_ell device_handle() const
            {
                if (! _imp->handle)
                {
                    hsolve_ell h = nullptr;
                    mi355x::check(hsolve_ell_create(mi355x::Suffix<DataType_>::dtype, rows(), columns(), stride(),
                                                    num_cols_per_row(), threads(),
                                                    (const uint64_t *)_imp->Aj.read_elements(),
                                                    _imp->Ax.read_elements(),
                                                    (const uint64_t *)_imp->Arl.read_elements(), &h),
                                  "SparseMatrixELL upload");
                    _imp->handle = std::make_shared<Handle>(h);
                }
                return _imp->handle->h;
            }

            void lock(LockMode mode) const
            {
                _imp->Aj.lock(mode);
                _imp->Ax.lock(mode);
                _imp->Arl.lock(mode);
            }

            void unlock(LockMode mode) const
            {
                _imp->Aj.unlock(mode);
                _imp->Ax.unlock(mode);
                _imp->Arl.unlock(mode);
            }

            SparseMatrixELL copy() const
            {
                SparseMatrixELL r(*this);
                r._imp = std::make_shared<Impl>(rows(), columns(), stride(), num_cols_per_row(), threads(),
                                                _imp->Aj.copy(), _imp->Ax.copy(), _imp->Arl.copy());
                return r;
            }
    };
}
