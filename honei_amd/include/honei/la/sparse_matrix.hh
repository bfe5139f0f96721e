// honei/la/sparse_matrix.hh -- row-wise sparse matrix used to build ELL
// matrices (HONEI's SparseMatrix, honei/la/sparse_matrix.hh:221): rows keep
// their entries sorted by column, like SparseVector.
#pragma once

#include <honei/util/exception.hh>

#include <map>
#include <vector>

namespace honei
{
    template <typename DataType_> class SparseMatrix
    {
        private:
            unsigned long _rows, _columns;
            std::vector<std::map<unsigned long, DataType_>> _row;

        public:
            SparseMatrix(unsigned long rows, unsigned long columns) : _rows(rows), _columns(columns), _row(rows) {}

            unsigned long rows() const { return _rows; }
            unsigned long columns() const { return _columns; }

            /// Stores value at (row, column); a stored zero stays stored.
            void operator()(unsigned long row, unsigned long column, DataType_ value)
            {
                if (row >= _rows || column >= _columns)
                    throw InternalError("SparseMatrix index out of range");
                _row[row][column] = value;
            }

            DataType_ operator()(unsigned long row, unsigned long column) const
            {
                auto it = _row[row].find(column);
                return it == _row[row].end() ? DataType_(0) : it->second;
            }

            const std::map<unsigned long, DataType_> & row(unsigned long r) const { return _row[r]; }

            unsigned long used_elements(unsigned long r) const { return _row[r].size(); }
    };
}
