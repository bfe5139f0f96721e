// honei/la/banded_matrix_qx.hh -- BandedMatrixQx<Q1Type, DT>: nine full-length
// bands LL LD LU DL DD DU UL UD UU at offsets -m-1, -m, -m+1, -1, 0, 1,
// m-1, m, m+1 with m = root() = sqrt(size) (honei/la/band_type.hh,
// banded_matrix_qx-impl.hh:45-210).  Each band is a DenseVector, so it is
// mirrored to HBM on first device use and stays resident.
#pragma once

#include <honei/la/dense_vector.hh>

#include <cmath>
#include <vector>

namespace honei
{
    enum BandType
    {
        Q1Type = 0
    };

    enum Q1BandIndex
    {
        LL = 0, LD, LU, DL, DD, DU, UL, UD, UU
    };

    template <BandType BandType_, typename DataType_> class BandedMatrixQx;

    template <typename DataType_> class BandedMatrixQx<Q1Type, DataType_>
    {
        private:
            std::vector<DenseVector<DataType_>> _bands;
            unsigned long _size;
            long _root;

        public:
            BandedMatrixQx(unsigned long size, const DenseVector<DataType_> & ll, const DenseVector<DataType_> & ld,
                           const DenseVector<DataType_> & lu, const DenseVector<DataType_> & dl,
                           const DenseVector<DataType_> & dd, const DenseVector<DataType_> & du,
                           const DenseVector<DataType_> & ul, const DenseVector<DataType_> & ud,
                           const DenseVector<DataType_> & uu) :
                _bands{ll, ld, lu, dl, dd, du, ul, ud, uu}, _size(size), _root((long)std::sqrt((double)size))
            {
                for (const auto & b : _bands)
                    if (b.size() != size)
                        throw VectorSizeDoesNotMatch(b.size(), size);
                if ((unsigned long)(_root * _root) != size)
                    throw InternalError("BandedMatrixQx: size " + std::to_string(size) + " is not a square");
            }

            /// size x size matrix with zero bands
            explicit BandedMatrixQx(unsigned long size) :
                BandedMatrixQx(size, DenseVector<DataType_>(size, DataType_(0)), DenseVector<DataType_>(size, DataType_(0)),
                               DenseVector<DataType_>(size, DataType_(0)), DenseVector<DataType_>(size, DataType_(0)),
                               DenseVector<DataType_>(size, DataType_(0)), DenseVector<DataType_>(size, DataType_(0)),
                               DenseVector<DataType_>(size, DataType_(0)), DenseVector<DataType_>(size, DataType_(0)),
                               DenseVector<DataType_>(size, DataType_(0)))
            {
            }

            unsigned long size() const { return _size; }
            unsigned long rows() const { return _size; }
            unsigned long columns() const { return _size; }
            signed long root() const { return _root; }

            DenseVector<DataType_> & band(unsigned long index) { return _bands.at(index); }
            const DenseVector<DataType_> & band(unsigned long index) const { return _bands.at(index); }

            /// Device pointers of the nine bands (read access), LL..UU.
            void device_bands(const DataType_ * out[9]) const
            {
                for (int i = 0; i < 9; ++i)
                    out[i] = static_cast<const DataType_ *>(_bands[i].lock(lm_read_only, tv_gpu_hip));
            }

            BandedMatrixQx copy() const
            {
                return BandedMatrixQx(_size, _bands[0].copy(), _bands[1].copy(), _bands[2].copy(), _bands[3].copy(),
                                      _bands[4].copy(), _bands[5].copy(), _bands[6].copy(), _bands[7].copy(),
                                      _bands[8].copy());
            }
    };
}
