// honei/math/io.hh -- fixture formats and Poisson problem generators, host
// code independent of the backend (HONEI keeps them tag-independent too):
//   MatrixIO<io_formats::ELL>   honei/math/matrix_io.hh:433-583
//       5 x uint64 header (size, rows, columns, stride, num_cols_per_row),
//       then uint64 Aj[size], float64 Ax[size]
//   VectorIO<io_formats::EXP>   honei/math/vector_io.hh:52-116 (text, '#' lines skipped)
//   VectorIO<io_formats::DV>    honei/math/vector_io.hh:205-250 (uint64 size + float64[])
//   FillMatrix<Tag, POISSON, DIRICHLET_0 / DIRICHLET_NEUMANN>   honei/math/fill_matrix.hh:36-133
//   FillVector<Tag, POISSON, DIRICHLET_NEUMANN>                 honei/math/fill_vector.hh:38-160
#pragma once

#include <honei/la/banded_matrix_qx.hh>
#include <honei/la/sparse_matrix_ell.hh>
#include <honei/math/methods.hh>

#include <cmath>
#include <cstdint>
#include <cstdio>
#include <fstream>
#include <string>
#include <vector>

namespace honei
{
    template <typename Format_> struct MatrixIO;
    template <typename Format_> struct VectorIO;

    template <> struct MatrixIO<io_formats::ELL>
    {
        template <typename DT_> static SparseMatrixELL<DT_> read_matrix(const std::string & input, DT_)
        {
            FILE * file = std::fopen(input.c_str(), "rb");
            if (! file)
                throw InternalError("File " + input + " not found!");
            uint64_t hdr[5];
            if (std::fread(hdr, sizeof(uint64_t), 5, file) != 5)
            {
                std::fclose(file);
                throw InternalError("fread error!");
            }
            const uint64_t size = hdr[0];
            DenseVector<unsigned long> aj(size);
            std::vector<double> ax(size);
            const bool ok = std::fread(aj.elements(), sizeof(uint64_t), size, file) == size &&
                            std::fread(ax.data(), sizeof(double), size, file) == size;
            std::fclose(file);
            if (! ok)
                throw InternalError("fread error!");
            DenseVector<DT_> axc(size);
            DT_ * e = axc.elements();
            for (uint64_t i = 0; i < size; ++i)
                e[i] = (DT_)ax[i];
            SparseMatrixELL<DT_> m(hdr[1], hdr[2], hdr[3], hdr[4], aj, axc, 1);
            // matrix_io.hh:572-579: with ell::threads != 1 the file's matrix is
            // laid out again through a SparseMatrix
            if (Configuration::instance()->get_value("ell::threads", 1) != 1)
                return SparseMatrixELL<DT_>(sparse_of(m));
            return m;
        }

        /// SparseMatrix(SparseMatrixELL) (honei/la/sparse_matrix.hh:157-185):
        /// every slot of every row (not only the first Arl[row]), zeros skipped.
        template <typename DT_> static SparseMatrix<DT_> sparse_of(const SparseMatrixELL<DT_> & m)
        {
            SparseMatrix<DT_> sm(m.rows(), m.columns());
            const unsigned long * aj = m.Aj().read_elements();
            const DT_ * ax = m.Ax().read_elements();
            const unsigned long size = m.Aj().size();
            for (unsigned long row(0) ; row < m.rows() ; ++row)
                for (unsigned long i(row * m.threads()) ; i < size ; i += m.stride())
                    for (unsigned long t(0) ; t < m.threads() ; ++t)
                        if (ax[i + t] != DT_(0))
                            sm(row, aj[i + t], ax[i + t]);
            return sm;
        }

        static void write_matrix(const std::string & output, const SparseMatrixELL<double> & m)
        {
            if (m.threads() != 1)
                throw InternalError("Only Matrices with 1 threads data layout are supported for export");
            FILE * file = std::fopen(output.c_str(), "wb");
            if (! file)
                throw InternalError("cannot open " + output);
            const uint64_t hdr[5] = {m.Aj().size(), m.rows(), m.columns(), m.stride(), m.num_cols_per_row()};
            std::fwrite(hdr, sizeof(uint64_t), 5, file);
            std::fwrite(m.Aj().read_elements(), sizeof(uint64_t), hdr[0], file);
            std::fwrite(m.Ax().read_elements(), sizeof(double), hdr[0], file);
            std::fclose(file);
        }
    };

    template <> struct VectorIO<io_formats::EXP>
    {
        template <typename DT_> static DenseVector<DT_> read_vector(const std::string & filename, DT_)
        {
            std::ifstream file(filename.c_str());
            if (! file.is_open())
                throw InternalError("Unable to open Vector file " + filename);
            std::vector<DT_> data;
            while (! file.eof())
            {
                std::string line;
                std::getline(file, line);
                if (line.find('#') != std::string::npos)
                    continue;
                if (file.eof())
                    break;
                const std::string::size_type first = line.find_first_not_of(' ');
                line.erase(0, first);
                data.push_back((DT_)std::atof(line.c_str()));
            }
            DenseVector<DT_> result(data.size());
            DT_ * e = result.elements();
            for (size_t i = 0; i < data.size(); ++i)
                e[i] = data[i];
            return result;
        }
    };

    template <> struct VectorIO<io_formats::DV>
    {
        template <typename DT_> static DenseVector<DT_> read_vector(const std::string & input, DT_)
        {
            FILE * file = std::fopen(input.c_str(), "rb");
            if (! file)
                throw InternalError("File " + input + " not found!");
            uint64_t size = 0;
            bool ok = std::fread(&size, sizeof(uint64_t), 1, file) == 1;
            std::vector<double> v(ok ? size : 0);
            ok = ok && std::fread(v.data(), sizeof(double), size, file) == size;
            std::fclose(file);
            if (! ok)
                throw InternalError("fread error!");
            DenseVector<DT_> result(size);
            DT_ * e = result.elements();
            for (uint64_t i = 0; i < size; ++i)
                e[i] = (DT_)v[i];
            return result;
        }

        template <typename DT_> static void write_vector(const std::string & output, const DenseVector<DT_> & dv)
        {
            FILE * file = std::fopen(output.c_str(), "wb");
            if (! file)
                throw InternalError("cannot open " + output);
            const uint64_t size = dv.size();
            std::vector<double> v(size);
            for (uint64_t i = 0; i < size; ++i)
                v[i] = (double)dv[i];
            std::fwrite(&size, sizeof(uint64_t), 1, file);
            std::fwrite(v.data(), sizeof(double), size, file);
            std::fclose(file);
        }
    };

    template <typename Tag_, typename Application_, typename BoundaryType_> struct FillMatrix;
    template <typename Tag_, typename Application_, typename BoundaryType_> struct FillVector;

    template <typename Tag_> struct FillMatrix<Tag_, applications::POISSON, boundary_types::DIRICHLET::DIRICHLET_0>
    {
        /// dd = 8/3, off-diagonals -1/3, identity rows on the boundary
        template <typename DT_> static void value(BandedMatrixQx<Q1Type, DT_> & target)
        {
            const unsigned long N = target.size();
            const unsigned long M = (unsigned long)std::sqrt((double)N);
            DT_ * b[9];
            for (int k = 0; k < 9; ++k)
                b[k] = target.band(k).elements();
            for (unsigned long i(1) ; i <= N ; ++i)
            {
                const bool bnd = i <= M || i > N - M || (i % M) == 1 || (i % M) == 0;
                for (int k = 0; k < 9; ++k)
                    b[k][i - 1] = bnd ? DT_(0.0) : DT_(-1.0 / 3.0);
                b[DD][i - 1] = bnd ? DT_(1.0) : DT_(8.0 / 3.0);
            }
        }
    };

    template <typename Tag_> struct FillMatrix<Tag_, applications::POISSON, boundary_types::DIRICHLET_NEUMANN>
    {
        /// Dirichlet on the first/last row and first column, Neumann on the right edge
        template <typename DT_> static void value(BandedMatrixQx<Q1Type, DT_> & target)
        {
            const unsigned long N = target.size();
            const unsigned long M = (unsigned long)std::sqrt((double)N);
            DT_ * b[9];
            for (int k = 0; k < 9; ++k)
                b[k] = target.band(k).elements();
            for (unsigned long i(1) ; i <= N ; ++i)
            {
                const unsigned long r = i - 1;
                if (i <= M || i > N - M || (i % M) == 1)
                {
                    for (int k = 0; k < 9; ++k)
                        b[k][r] = DT_(0.0);
                    b[DD][r] = DT_(1.0);
                }
                else if ((i % M) == 0 && i > M && i < N)
                {
                    b[DD][r] = DT_(4.0 / 3.0);
                    b[LL][r] = DT_(-1.0 / 3.0);
                    b[LD][r] = DT_(-1.0 / 6.0);
                    b[DL][r] = DT_(-1.0 / 3.0);
                    b[UL][r] = DT_(-1.0 / 3.0);
                    b[UD][r] = DT_(-1.0 / 6.0);
                    b[LU][r] = b[DU][r] = b[UU][r] = DT_(0.0);
                }
                else
                {
                    for (int k = 0; k < 9; ++k)
                        b[k][r] = DT_(-1.0 / 3.0);
                    b[DD][r] = DT_(8.0 / 3.0);
                }
            }
        }
    };

    template <typename Tag_> struct FillVector<Tag_, applications::POISSON, boundary_types::DIRICHLET_NEUMANN>
    {
        template <typename DT_> static void value(DenseVector<DT_> & target)
        {
            const unsigned long N = target.size();
            const unsigned long M = (unsigned long)std::sqrt((double)N);
            int level = 0;
            while (level <= 30 && (unsigned long)(((1ul << level) + 1) * ((1ul << level) + 1)) != N)
                ++level;
            if (level > 30)
                throw InternalError("Uknown size!");
            DT_ * t = target.elements();
            for (unsigned long i(1) ; i <= N ; ++i)
            {
                if (i <= M || i > N - M || (i % M) == 1)
                    t[i - 1] = DT_(2.0);
                else if ((i % M) == 0 && i > M && i < N)
                    t[i - 1] = DT_(0.5 * (1.0 / std::pow(4.0, level)));
                else
                    t[i - 1] = DT_(1.0 / std::pow(4.0, level));
            }
        }
    };
}
