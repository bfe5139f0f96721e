// honei/la/dense_vector.hh -- DenseVector<DT> of the MI355X build.
//
// Same handle semantics as HONEI (honei/la/dense_vector.hh:48-210): copying
// a DenseVector shares its storage, copy() makes a deep copy, elements() and
// operator[] give host access, lock()/unlock() give the pointer of a memory
// space (tv_cpu or tv_gpu_hip) with the coherence rules of MemoryBlock.
#pragma once

#include <honei/util/memory_arbiter.hh>

#include <memory>
#include <ostream>

namespace honei
{
    template <typename DataType_> class DenseVector
    {
        private:
            std::shared_ptr<MemoryBlock> _block;
            unsigned long _size;

        public:
            typedef DataType_ DataType;

            explicit DenseVector(unsigned long size) :
                _block(std::make_shared<MemoryBlock>(sizeof(DataType_) * size)), _size(size)
            {
            }

            DenseVector(unsigned long size, DataType_ value) : DenseVector(size)
            {
                DataType_ * e = elements();
                for (unsigned long i(0) ; i < size ; ++i)
                    e[i] = value;
            }

            DenseVector(const DenseVector &) = default;
            DenseVector & operator=(const DenseVector &) = default;

            unsigned long size() const { return _size; }

            /// Host elements, coherent; marks the device copy stale.
            DataType_ * elements() const
            {
                return static_cast<DataType_ *>(_block->lock(lm_read_and_write, tv_cpu));
            }

            /// Host elements for reading only.
            const DataType_ * read_elements() const
            {
                return static_cast<const DataType_ *>(_block->lock(lm_read_only, tv_cpu));
            }

            DataType_ & operator[](unsigned long i) { return elements()[i]; }
            const DataType_ & operator[](unsigned long i) const { return read_elements()[i]; }

            void * lock(LockMode mode, TagValue memory = tv_cpu) const { return _block->lock(mode, memory); }
            void unlock(LockMode mode) const { _block->unlock(mode); }

            /// Identity of the storage (HONEI's memid()).
            const void * memid() const { return _block.get(); }

            /// Deep copy.
            DenseVector copy() const
            {
                DenseVector result(_size);
                if (_block->device_valid() && ! _block->host_valid())
                {
                    void * dst = result.lock(lm_write_only, tv_gpu_hip);
                    const void * src = lock(lm_read_only, tv_gpu_hip);
                    mi355x::check(hsolve_copy(src, dst, sizeof(DataType_) * _size, mi355x::stream()), "copy");
                    result.unlock(lm_write_only);
                    unlock(lm_read_only);
                }
                else
                {
                    const DataType_ * s = read_elements();
                    DataType_ * d = result.elements();
                    for (unsigned long i(0) ; i < _size ; ++i)
                        d[i] = s[i];
                }
                return result;
            }
    };

    template <typename DT_> std::ostream & operator<<(std::ostream & lhs, const DenseVector<DT_> & v)
    {
        lhs << "[";
        for (unsigned long i(0) ; i < v.size() ; ++i)
            lhs << (i ? ", " : "") << v[i];
        return lhs << "]";
    }
}
