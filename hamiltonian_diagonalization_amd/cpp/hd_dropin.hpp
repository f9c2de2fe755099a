// hd_dropin.hpp -- C++ host classes with the reference's names and calling
// conventions, implemented over the C ABI of libhdiag_hip.so (include/hdiag.h).
//
// A caller written against the reference's FermionFockBasis,
// SpinlessFermionHamiltonian<T>, HamiltonianBase<T,B>, CrsSparseMatrix<T>,
// MatrixVectorFunction<T,L1,L2> and ArpackWrapper keeps its call sequence:
//
//   FermionFockBasis basis;
//   basis.GenerateFockSpace(N, Norb, AngularMomentumSector(lzTot, maxLz), mpi);
//   ham.SetFockBasis(basis);
//   ham.Initialize(N, Norb, utilities::_SPARSE_MAPPED_, mpi);
//   ham.Add_CdC_Terms(&quadraticTables, 0, ham.m_data.m_fockSpaceDim, mpi);   // optional
//   ham.Add_CdCdCC_Terms(&quarticTables, 0, ham.m_data.m_fockSpaceDim, mpi);
//   MatrixVectorFunction<double, Quad, Quart> mv(ham, mpi);
//   ham.LanczosDiagonalize(&quadraticTables, &quarticTables, mv, nev, mpi);
//
// Differences a caller sees (each forced by running on the device):
//  - GenerateFockSpace takes the sector as a descriptor functor
//    (AngularMomentumSector, LinearMomentumSector2D, LinearMomentumSector1D,
//    NoConstraint).  They are callable like the reference's predicates
//    (std::bind(&TestAngularMomentumSector, _1, lzTot, maxLz)), so they convert
//    to std::function<bool(fock_t)>; an opaque std::function is rejected
//    because a device basis cannot call a host predicate.
//  - Add_CdC_Terms only records the quadratic table; Add_CdCdCC_Terms runs the
//    whole build (count, scan, fill, sort) on the device over all rows, which
//    is the only range the model callers use (pseudopotential_model.cpp:246-250).
//  - The matrix lives in device memory in full (both triangles) storage.
//    m_crsSparseMatrix presents the reference's upper-triangle layout
//    (int32 row starts and columns, ascending columns, M[i][j] = <j|H|i>) and
//    copies it from the device on first access.
//  - LanczosDiagonalize runs the device thick-restart Lanczos instead of
//    calling back into MatrixVectorFunction::Simple for every product; Simple
//    stays available for callers that multiply host vectors themselves.
//
// Errors follow the reference: messages on stderr, mpi.m_exitFlag = true, and
// MpiWrapper::ExitFlagTest() ends the process where the reference calls it.
#ifndef HD_DROPIN_HPP
#define HD_DROPIN_HPP

#include <hdiag.h>

#include <algorithm>
#include <cmath>
#include <complex>
#include <cstdint>
#include <cstdio>
#include <cstdlib>
#include <cstring>
#include <fstream>
#include <functional>
#include <iomanip>
#include <iostream>
#include <map>
#include <memory>
#include <sstream>
#include <string>
#include <type_traits>
#include <vector>

// src/utilities/general/orbital_and_state_defs.hpp:36-58
typedef uint64_t fock_t;
typedef uint16_t kState_t;
typedef int16_t mState_t;
typedef uint32_t iSize_t;
typedef int32_t crsIndex_t;
typedef std::complex<double> dcmplx;
#ifndef _SEARCH_ERROR_CODE_
#define _SEARCH_ERROR_CODE_ UINT64_MAX
#endif
#ifndef _SPARSE_ZERO_TOL_
#define _SPARSE_ZERO_TOL_ 1e-14
#endif

#ifndef HD_DROPIN_REFERENCE_MPI
// src/utilities/wrappers/io_wrapper.hpp:21-27
namespace io
{
    enum fileFormat_t { _TEXT_ = 0, _BINARY_ = 1 };
    enum io_t { _IN_, _OUT_ };
}
#endif

namespace utilities
{
#ifndef HD_DROPIN_REFERENCE_MPI
    // Single-process stand-in for the reference's MpiWrapper
    // (src/utilities/wrappers/mpi_wrapper.hpp): one host thread drives one GPU,
    // so there is one rank and every stage owns all rows.  Inside the
    // reference's build, define HD_DROPIN_REFERENCE_MPI and include its
    // mpi_wrapper.hpp and sparse_matrix.hpp first; only m_id, m_nbrProcs,
    // m_exitFlag and ExitFlagTest() are used.
    struct MpiWrapper
    {
        int m_id = 0;
        int m_nbrProcs = 1;
        fock_t m_firstTask = 0;
        fock_t m_lastTask = 0;
        bool m_exitFlag = false;

        // mpi_wrapper.cpp:327-338: stop the run once an error was flagged
        void ExitFlagTest() const
        {
            if (m_exitFlag) {
                std::cerr << "\n\tEXIT FLAG SET: TERMINATING" << std::endl;
                std::exit(EXIT_FAILURE);
            }
        }
    };

    // src/utilities/data_structures/sparse_matrix.hpp:55
    enum storageMethod_t { _SPARSE_MAPPED_, _SPARSE_CRS_, _DENSE_ };

    // io_wrapper.hpp:35-70: open a text or binary stream; on failure the
    // message goes to stderr and mpi.m_exitFlag is set.
    template <typename F>
    void GenFileStream(F& stream, const std::string fileName, const io::fileFormat_t format, MpiWrapper& mpi)
    {
        constexpr bool in = std::is_same<F, std::ifstream>::value;
        std::ios::openmode mode = in ? std::ios::in : std::ios::out;
        if (format == io::_BINARY_) mode |= std::ios::binary;
        else if (format != io::_TEXT_) {
            std::cerr << "Unknown file format " << format << std::endl;
            mpi.m_exitFlag = true;
            return;
        }
        stream.open(fileName.c_str(), mode);
        if (!stream.is_open()) {
            std::cerr << "Could not open file " << fileName << std::endl;
            mpi.m_exitFlag = true;
        }
    }

    // io_wrapper.hpp:77-119: text fields are printed with 15 significant
    // digits, a complex number as "re<TAB>im".
    template <typename T>
    std::string ToStream(const T variable)
    {
        std::stringstream ss;
        ss << std::setprecision(15) << variable;
        return ss.str();
    }
    inline std::string ToStream(const dcmplx variable)
    {
        std::stringstream ss;
        ss << std::setprecision(15) << variable.real() << "\t" << std::setprecision(15) << variable.imag();
        return ss.str();
    }
    inline void FromStream(std::ifstream& stream, double& value) { stream >> value; }
    inline void FromStream(std::ifstream& stream, unsigned short& value) { stream >> value; }
    inline void FromStream(std::ifstream& stream, dcmplx& value)
    {
        double re = 0.0, im = 0.0;
        stream >> re >> im;
        value = dcmplx(re, im);
    }
#endif

    namespace detail
    {
        inline bool Ok(int status, const char* where)
        {
            if (status == HD_OK) return true;
            std::cerr << "\n\tERROR in " << where << ": " << hd_last_error() << std::endl;
            return false;
        }
        struct BasisDeleter { void operator()(hd_basis* b) const { hd_basis_destroy(b); } };
        struct TablesDeleter { void operator()(hd_tables* t) const { hd_tables_destroy(t); } };
        struct HamDeleter { void operator()(hd_hamiltonian* h) const { hd_ham_destroy(h); } };
        template <typename T> struct IsComplex : std::false_type {};
        template <> struct IsComplex<dcmplx> : std::true_type {};
        // array term tables expose their k vector (term_tables_base.hpp:105-115)
        template <class L, class = void> struct IsArrayQuartic : std::false_type {};
        template <class L>
        struct IsArrayQuartic<L, typename std::enable_if<std::is_same<
                                     decltype(std::declval<L&>().GetKTable()), std::vector<kState_t>*>::value>::type>
            : std::true_type {};
        template <class L, class = void> struct HasDimension : std::false_type {};
        template <class L>
        struct HasDimension<L, decltype((void)std::declval<const L&>().GetDimension())> : std::true_type {};
        // entries of a quadratic table: its dimension for array tables, else
        // the orbital count
        template <class L>
        iSize_t TableKMax(const L* t, const iSize_t nbrOrbitals)
        {
            if constexpr (HasDimension<L>::value) return t->GetDimension();
            else return nbrOrbitals;
        }
    }

    // CrsSparseMatrix<T> (src/utilities/data_structures/crs_sparse_matrix.hpp:123-816):
    // the reference's upper-triangle CRS of H^T.  The arrays are a host copy of
    // a device build in that layout, made on first access.
    template <typename T>
    class CrsSparseMatrix
    {
        public:
        CrsSparseMatrix() = default;

        void Initialize(const int nbrRows, const int nbrCols, const long int nbrNonZerosEstimate)
        {
            Clear();
            m_nbrRows = nbrRows;
            m_nbrCols = nbrCols;
            m_values.reserve((size_t)std::max<long int>(nbrNonZerosEstimate, 0));
            m_cols.reserve((size_t)std::max<long int>(nbrNonZerosEstimate, 0));
            m_rowStarts.assign((size_t)nbrRows + 1, 0);
            m_materialized = true;
        }

        void Clear()
        {
            m_values.clear();
            m_cols.clear();
            m_rowStarts.clear();
            m_nbrRows = m_nbrCols = 0;
            m_materialized = false;
            m_source.reset();
            m_basis.reset();
        }

        // Records where the host copy comes from: the device basis and tables,
        // built again in upper-triangle storage when an accessor is first used.
        void AdoptDevice(std::shared_ptr<hd_basis> basis, std::shared_ptr<hd_tables> tables,
                         const double zeroTol, const int dim)
        {
            Clear();
            m_basis = std::move(basis);
            m_source = std::move(tables);
            m_zeroTol = zeroTol;
            m_nbrRows = m_nbrCols = dim;
        }

        long int GetNbrNonZeros() const { Materialize(); return (long int)m_values.size(); }
        int GetNbrRows() const { return m_nbrRows; }
        const int* GetRowStarts() const { Materialize(); return m_rowStarts.data(); }
        int* GetRowStarts() { Materialize(); return m_rowStarts.data(); }
        const int* GetCols() const { Materialize(); return m_cols.data(); }
        int* GetCols() { Materialize(); return m_cols.data(); }
        const T* GetData() const { Materialize(); return m_values.data(); }
        T* GetData() { Materialize(); return m_values.data(); }

        // crs_sparse_matrix.hpp Value(i, j): the stored element or zero
        T Value(const int i, const int j) const
        {
            Materialize();
            const int* b = m_cols.data() + m_rowStarts[i];
            const int* e = m_cols.data() + m_rowStarts[i + 1];
            const int* p = std::lower_bound(b, e, j);
            return (p != e && *p == j) ? m_values[(size_t)(p - m_cols.data())] : T(0.0);
        }

        // crs_sparse_matrix.hpp:643-660: row-major dense copy of the stored entries
        void ConvertToArray(T* denseMatrix, const int nbrCols) const
        {
            Materialize();
            std::fill(denseMatrix, denseMatrix + (size_t)m_nbrRows * nbrCols, T(0.0));
            for (int i = 0; i < m_nbrRows; ++i)
                for (int k = m_rowStarts[i]; k < m_rowStarts[i + 1]; ++k)
                    denseMatrix[(size_t)i * nbrCols + m_cols[k]] = m_values[k];
        }

        void Print() const
        {
            Materialize();
            for (int i = 0; i < m_nbrRows; ++i)
                for (int k = m_rowStarts[i]; k < m_rowStarts[i + 1]; ++k)
                    std::cout << "\t(" << i << "," << m_cols[k] << ")\t" << m_values[k] << "\n";
        }

        // crs_sparse_matrix.hpp:662-700.  Binary: uint32 nnz, uint32 rows, int32
        // row starts (rows + 1), int32 columns (nnz), T values (nnz).  Text: nnz,
        // rows, the first `rows` row starts one per line, then "row col value"
        // per entry.
        void ToFile(const std::string fileName, const io::fileFormat_t format, MpiWrapper& mpi) const
        {
            Materialize();
            std::ofstream f_out;
            GenFileStream(f_out, fileName, format, mpi);
            mpi.ExitFlagTest();
            const unsigned int nnz = (unsigned int)m_values.size();
            const unsigned int rows = m_rowStarts.empty() ? 0u : (unsigned int)(m_rowStarts.size() - 1);
            if (format == io::_TEXT_) {
                f_out << nnz << "\n" << rows << "\n";
                for (unsigned int i = 0; i < rows; ++i) f_out << m_rowStarts[i] << "\n";
                for (unsigned int i = 0; i < rows; ++i)
                    for (int j = m_rowStarts[i]; j < m_rowStarts[i + 1]; ++j)
                        f_out << i << " " << m_cols[j] << " " << ToStream(m_values[j]) << "\n";
            } else {
                f_out.write(reinterpret_cast<const char*>(&nnz), sizeof(unsigned int));
                f_out.write(reinterpret_cast<const char*>(&rows), sizeof(unsigned int));
                const int zero = 0;
                const int* rs = m_rowStarts.empty() ? &zero : m_rowStarts.data();
                f_out.write(reinterpret_cast<const char*>(rs), ((std::streamsize)rows + 1) * sizeof(int));
                f_out.write(reinterpret_cast<const char*>(m_cols.data()), (std::streamsize)nnz * sizeof(int));
                f_out.write(reinterpret_cast<const char*>(m_values.data()), (std::streamsize)nnz * sizeof(T));
            }
        }

        // crs_sparse_matrix.hpp:705-751.  The text layout omits the final row
        // start; it is nnz, which the header line holds (the reference leaves it
        // zero and so drops the last row's entries on reading).  A short or
        // inconsistent file is an error here.
        void FromFile(const std::string fileName, const io::fileFormat_t format, MpiWrapper& mpi)
        {
            std::ifstream f_in;
            GenFileStream(f_in, fileName, format, mpi);
            mpi.ExitFlagTest();
            Clear();
            unsigned int nnz = 0, rows = 0;
            bool ok = true;
            if (format == io::_TEXT_) {
                ok = (bool)(f_in >> nnz >> rows);
                m_values.resize(nnz);
                m_cols.resize(nnz);
                m_rowStarts.assign((size_t)rows + 1, 0);
                for (unsigned int i = 0; ok && i < rows; ++i) ok = (bool)(f_in >> m_rowStarts[i]);
                m_rowStarts[rows] = (int)nnz;
                for (unsigned int i = 0; ok && i < rows; ++i) {
                    if (m_rowStarts[i] > m_rowStarts[i + 1] || m_rowStarts[i] < 0) { ok = false; break; }
                    for (int j = m_rowStarts[i]; ok && j < m_rowStarts[i + 1]; ++j) {
                        int row = 0;
                        ok = (bool)(f_in >> row >> m_cols[j]);
                        FromStream(f_in, m_values[j]);
                        ok = ok && (bool)f_in && row == (int)i;
                    }
                }
            } else {
                f_in.read(reinterpret_cast<char*>(&nnz), sizeof(unsigned int));
                f_in.read(reinterpret_cast<char*>(&rows), sizeof(unsigned int));
                m_values.resize(nnz);
                m_cols.resize(nnz);
                m_rowStarts.assign((size_t)rows + 1, 0);
                f_in.read(reinterpret_cast<char*>(m_rowStarts.data()), ((std::streamsize)rows + 1) * sizeof(int));
                f_in.read(reinterpret_cast<char*>(m_cols.data()), (std::streamsize)nnz * sizeof(int));
                f_in.read(reinterpret_cast<char*>(m_values.data()), (std::streamsize)nnz * sizeof(T));
                ok = (bool)f_in && m_rowStarts[rows] == (int)nnz;
            }
            if (!ok) {
                std::cerr << "\n\tERROR in CrsSparseMatrix::FromFile: " << fileName << " is truncated or inconsistent"
                          << std::endl;
                mpi.m_exitFlag = true;
                Clear();
                return;
            }
            m_nbrRows = m_nbrCols = (int)rows;
            m_materialized = true;
        }

        private:
        void Materialize() const
        {
            if (m_materialized) return;
            m_materialized = true;
            if (!m_source) return;
            hd_hamiltonian* raw = nullptr;
            if (!detail::Ok(hd_ham_build(m_basis.get(), m_source.get(), HD_STORAGE_UPPER_REF, m_zeroTol, 1, &raw),
                            "CrsSparseMatrix (upper-triangle build)"))
                return;
            std::unique_ptr<hd_hamiltonian, detail::HamDeleter> h(raw);
            hd_ham_info_t info;
            if (!detail::Ok(hd_ham_info(h.get(), &info), "CrsSparseMatrix")) return;
            // crsIndex_t is int32 in the reference: nnz < 2^31 (crs_sparse_matrix.hpp:127-133)
            if (info.nnz >= ((int64_t)1 << 31)) {
                std::cerr << "\n\tERROR in CrsSparseMatrix: " << info.nnz
                          << " non-zeros exceed the int32 row starts of the reference layout" << std::endl;
                return;
            }
            std::vector<int64_t> rp((size_t)info.dim + 1);
            m_cols.resize((size_t)info.nnz);
            m_values.resize((size_t)info.nnz);
            if (!detail::Ok(hd_ham_copy_crs(h.get(), 0, info.dim, HD_F64, rp.data(), m_cols.data(), m_values.data()),
                            "CrsSparseMatrix"))
                return;
            m_rowStarts.assign(rp.begin(), rp.end());
        }

        mutable std::vector<T> m_values;
        mutable std::vector<int> m_cols;
        mutable std::vector<int> m_rowStarts;
        mutable bool m_materialized = false;
        int m_nbrRows = 0;
        int m_nbrCols = 0;
        double m_zeroTol = _SPARSE_ZERO_TOL_;
        std::shared_ptr<hd_tables> m_source;
        std::shared_ptr<hd_basis> m_basis;
    };

    namespace linearAlgebra
    {
        // Options of ArpackWrapper (src/utilities/wrappers/arpack_wrapper.hpp,
        // defaults arpack_wrapper.cpp:38-49).  The iteration itself runs on the
        // device (hd_lanczos); the initial/final vector files keep the
        // reference's raw T[nodeDim] binary format (arpack_wrapper.hpp:753-774).
        class ArpackWrapper
        {
            public:
            void SetEigenvalueMagnitude(const char which) { m_which = which; }
            void UseInitialVector() { m_provideInitial = true; }
            void StoreFinalVector() { m_storeFinal = true; }
            void SetInitialVectorFile(const std::string& name) { m_initialVectorFile = name; }
            void SetFinalVectorFile(const std::string& name) { m_finalVectorFile = name; }
            char GetUpperOrLower() const { return 'U'; }
            // The reference leaves ARPACK's TOL uninitialised (arpack_wrapper.hpp:624);
            // here it is explicit: ||Hx - lx|| <= tol * max(1, |l|).
            void SetTolerance(const double tol) { m_tol = tol; }
            void SetNbrKrylovVectors(const int ncv) { m_ncv = ncv; }
            void SetMaxRestarts(const int n) { m_maxRestarts = n; }
            void SetSeed(const uint64_t seed) { m_seed = seed; }
            // HD_F64 (default), HD_CODED, or HD_F32 with FP64 refinement.
            void SetPrecision(const int precision) { m_precision = precision; }
            // A device start vector (e.g. the learned start of hd_mlp_start_vector);
            // takes precedence over the initial-vector file.
            void SetDeviceStartVector(const void* d_start) { m_deviceStart = d_start; }
            // Partial reorthogonalisation (default) or a full pass at every step.
            void SetFullReorthogonalisation(const bool full) { m_fullReorth = full; }

            char m_which = 'S';
            bool m_provideInitial = false;
            bool m_storeFinal = false;
            std::string m_initialVectorFile;
            std::string m_finalVectorFile;
            double m_tol = 1e-12;
            int m_ncv = 0;
            int m_maxRestarts = 0;
            uint64_t m_seed = 0;
            int m_precision = HD_F64;
            const void* m_deviceStart = nullptr;
            bool m_fullReorth = false;
            hd_lanczos_stats m_lastStats{};
        };
    }
}

namespace diagonalization
{
    // Sector descriptors.  Each evaluates the same predicate as the reference
    // function it names, so it can stand wherever that predicate was bound.

    // TestAngularMomentumSector (angular_momentum_constraint.hpp:41-62)
    struct AngularMomentumSector
    {
        int m_lzTot, m_maxLz;
        AngularMomentumSector(const int lzTot, const int maxLz) : m_lzTot(lzTot), m_maxLz(maxLz) {}
        bool operator()(const fock_t state) const
        {
            long s = 0;
            for (fock_t b = state; b; b &= b - 1) s += 2 * __builtin_ctzll(b) - m_maxLz;
            return s == m_lzTot;
        }
        hd_sector Descriptor() const { return hd_sector{HD_SECTOR_LZ, m_lzTot, m_maxLz, 0, 0}; }
    };

    // TestLinearMomentumSector2 (linear_momentum_constraints_2D.cpp:38-79), k = kx*dimY + ky
    struct LinearMomentumSector2D
    {
        int m_kx, m_ky, m_dimX, m_dimY;
        LinearMomentumSector2D(const int kx, const int ky, const int dimX, const int dimY)
            : m_kx(kx), m_ky(ky), m_dimX(dimX), m_dimY(dimY) {}
        bool operator()(const fock_t state) const
        {
            long sx = 0, sy = 0;
            for (fock_t b = state; b; b &= b - 1) {
                const int k = __builtin_ctzll(b);
                sx += k / m_dimY;
                sy += k % m_dimY;
            }
            return sx % m_dimX == m_kx && sy % m_dimY == m_ky;
        }
        hd_sector Descriptor() const { return hd_sector{HD_SECTOR_K2D, m_kx, m_ky, m_dimX, m_dimY}; }
    };

    // TestLinearMomentumSector1 (linear_momentum_constraints_2D.cpp:87-110)
    struct LinearMomentumSector1D
    {
        int m_ky, m_dimY;
        LinearMomentumSector1D(const int ky, const int dimY) : m_ky(ky), m_dimY(dimY) {}
        bool operator()(const fock_t state) const
        {
            long s = 0;
            for (fock_t b = state; b; b &= b - 1) s += __builtin_ctzll(b);
            return s % m_dimY == m_ky;
        }
        hd_sector Descriptor() const { return hd_sector{HD_SECTOR_K1D, m_ky, m_dimY, 0, 0}; }
    };

    // [](const fock_t state){return true;} (pseudopotential_model.cpp:197)
    struct NoConstraint
    {
        bool operator()(const fock_t) const { return true; }
        hd_sector Descriptor() const { return hd_sector{HD_SECTOR_NONE, 0, 0, 0, 0}; }
    };

    // FermionFockBasis (src/basis/fermion_fock_basis.hpp, .cpp): m_basis is the
    // host copy of the device basis, in the reference's ascending order.
    class FermionFockBasis
    {
        public:
        std::vector<fock_t> m_basis;
        bool m_basisStored = false;
        fock_t m_firstState = 0;

        void Clear()
        {
            m_basis.clear();
            m_basisStored = false;
            m_firstState = 0;
            m_device.reset();
        }

        // fermion_fock_basis.cpp:86-92: C(nbrOrbitals, nbrParticles), particles first
        fock_t GetFullFockSpaceDimension(const iSize_t nbrParticles, const iSize_t nbrOrbitals) const
        {
            if (nbrParticles > nbrOrbitals) return 0;
            unsigned __int128 c = 1;
            for (iSize_t i = 0; i < nbrParticles; ++i) c = c * (nbrOrbitals - i) / (i + 1);
            return (fock_t)c;
        }

        fock_t GetReducedFockSpaceDimension(const iSize_t, const iSize_t, const utilities::MpiWrapper&) const
        {
            return m_basis.size();
        }

        // fermion_fock_basis.cpp:134-161: offset + position of the state, or
        // _SEARCH_ERROR_CODE_; the search is over the host copy, whose order is
        // the device ranking's.
        fock_t FindFockStateIndex(const fock_t offset, const fock_t state) const
        {
            if (!m_basisStored) {
                std::cerr << "ERROR in FindFockStateIndex: BASIS WAS NOT STORED " << std::endl;
                return _SEARCH_ERROR_CODE_;
            }
            auto it = std::lower_bound(m_basis.begin(), m_basis.end(), state);
            if (it == m_basis.end() || *it != state) return _SEARCH_ERROR_CODE_;
            return offset + (fock_t)(it - m_basis.begin());
        }

        fock_t GetFirstState(const fock_t firstNodeIndex, const iSize_t,
                             std::vector<fock_t>::const_iterator& it_state) const
        {
            if (firstNodeIndex != m_firstState || m_basis.empty()) {
                std::cerr << "\n\tERROR in GetFirstState: non-matching first state index." << std::endl;
                std::exit(EXIT_FAILURE);
            }
            it_state = m_basis.begin();
            return m_basis[0];
        }

        fock_t GetNextState(const fock_t, std::vector<fock_t>::const_iterator& it_state) const
        {
            ++it_state;
            return *it_state;
        }

        // fermion_fock_basis.cpp:241-354: returns true when the sector is empty.
        bool GenerateFockSpace(const iSize_t nbrParticles, const iSize_t nbrOrbitals,
                               const std::function<bool(const fock_t state)>& Constraint,
                               utilities::MpiWrapper& mpi)
        {
            hd_sector s;
            if (auto* p = Constraint.target<AngularMomentumSector>()) s = p->Descriptor();
            else if (auto* p2 = Constraint.target<LinearMomentumSector2D>()) s = p2->Descriptor();
            else if (auto* p1 = Constraint.target<LinearMomentumSector1D>()) s = p1->Descriptor();
            else if (Constraint.target<NoConstraint>()) s = NoConstraint().Descriptor();
            else return GenerateFockSpaceOnHost(nbrParticles, nbrOrbitals, Constraint, mpi);
            return GenerateFockSpace(nbrParticles, nbrOrbitals, s, mpi);
        }

        // An opaque predicate (the reference's callers pass std::bind of
        // TestAngularMomentumSector / TestLinearMomentumSector2, or lambdas):
        // every N-particle state in ascending order (Gosper's successor, as
        // fermion_fock_basis.cpp:266-307) is tested on the host.  When the kept
        // states are exactly one of the sectors the device generates (no
        // constraint; L_z; k_y on a dimX x dimY grid; (k_x, k_y)), that device
        // basis is used; otherwise the list is uploaded as it is
        // (hd_basis_create_from_states) and the Hamiltonian takes the general
        // build.  A hint names the path taken.
        bool GenerateFockSpaceOnHost(const iSize_t nbrParticles, const iSize_t nbrOrbitals,
                                     const std::function<bool(const fock_t state)>& Constraint,
                                     utilities::MpiWrapper& mpi)
        {
            Clear();
            if (nbrParticles == 0 || nbrParticles > nbrOrbitals || nbrOrbitals > 64 || nbrParticles >= 64) {
                std::cerr << "\n\tERROR in GenerateFockSpace: invalid particle or orbital count" << std::endl;
                mpi.m_exitFlag = true;
                mpi.ExitFlagTest();
                return true;
            }
            std::vector<fock_t> kept;
            const fock_t last = (nbrParticles == 64 ? ~(fock_t)0 : (((fock_t)1 << nbrParticles) - 1))
                                << (nbrOrbitals - nbrParticles);
            for (fock_t x = ((fock_t)1 << nbrParticles) - 1;;) {
                if (Constraint(x)) kept.push_back(x);
                if (x == last) break;
                const fock_t c = x & (0 - x), r = x + c;
                x = (((r ^ x) >> 2) / c) | r;
            }
            if (kept.empty()) return true;
            // the sectors a device basis can have for these states
            struct Candidate {
                hd_sector sector;
                std::function<bool(fock_t)> test;
                std::string name;
            };
            std::vector<Candidate> cand;
            if ((fock_t)kept.size() == GetFullFockSpaceDimension(nbrParticles, nbrOrbitals))
                cand.push_back({NoConstraint().Descriptor(), NoConstraint(), "no constraint"});
            int lz = 0;
            for (iSize_t k = 0; k < nbrOrbitals; ++k)
                if (kept[0] >> k & 1) lz += 2 * (int)k - (int)(nbrOrbitals - 1);
            const AngularMomentumSector am(lz, (int)nbrOrbitals - 1);
            cand.push_back({am.Descriptor(), am, "AngularMomentumSector"});
            for (iSize_t dimY = 1; dimY <= nbrOrbitals; ++dimY) {
                if (nbrOrbitals % dimY) continue;
                const int dimX = (int)(nbrOrbitals / dimY);
                int kx = 0, ky = 0;
                for (iSize_t k = 0; k < nbrOrbitals; ++k)
                    if (kept[0] >> k & 1) {
                        kx += (int)(k / dimY);
                        ky += (int)(k % dimY);
                    }
                const LinearMomentumSector1D s1(ky % (int)dimY, (int)dimY);
                const LinearMomentumSector2D s2(kx % dimX, ky % (int)dimY, dimX, (int)dimY);
                cand.push_back({s1.Descriptor(), s1, "LinearMomentumSector1D"});
                cand.push_back({s2.Descriptor(), s2, "LinearMomentumSector2D"});
            }
            for (const auto& c : cand) {
                // every kept state must lie in the sector before its device
                // basis is generated and compared
                bool inside = true;
                for (fock_t x : kept)
                    if (!c.test(x)) {
                        inside = false;
                        break;
                    }
                if (!inside) continue;
                hd_basis* raw = nullptr;
                if (hd_basis_create(&c.sector, nbrParticles, nbrOrbitals, &raw) != HD_OK) continue;
                std::shared_ptr<hd_basis> dev(raw, utilities::detail::BasisDeleter());
                int64_t dim = 0;
                hd_basis_dim(raw, &dim);
                if ((size_t)dim != kept.size()) continue;
                std::vector<fock_t> states((size_t)dim);
                if (hd_basis_copy_states(raw, 0, dim, states.data()) != HD_OK || states != kept) continue;
                std::cerr << "\n\tHINT GenerateFockSpace: the constraint selects the states of a device sector ("
                          << c.name << "); pass the descriptor to skip the host enumeration." << std::endl;
                m_device = dev;
                m_basis.swap(states);
                m_basisStored = true;
                return false;
            }
            std::cerr << "\n\tHINT GenerateFockSpace: the constraint was evaluated on the host for "
                      << kept.size() << " states; the Hamiltonian takes the general build." << std::endl;
            hd_basis* raw = nullptr;
            if (!utilities::detail::Ok(hd_basis_create_from_states(nbrParticles, nbrOrbitals, kept.data(),
                                                                   (int64_t)kept.size(), &raw),
                                       "GenerateFockSpace")) {
                mpi.m_exitFlag = true;
                mpi.ExitFlagTest();
                return true;
            }
            m_device.reset(raw, utilities::detail::BasisDeleter());
            m_basis.swap(kept);
            m_basisStored = true;
            return false;
        }

        bool GenerateFockSpace(const iSize_t nbrParticles, const iSize_t nbrOrbitals, const hd_sector& sector,
                               utilities::MpiWrapper& mpi)
        {
            Clear();
            hd_basis* raw = nullptr;
            const int st = hd_basis_create(&sector, nbrParticles, nbrOrbitals, &raw);
            if (st == HD_ERR_EMPTY) return true;
            if (!utilities::detail::Ok(st, "GenerateFockSpace")) {
                mpi.m_exitFlag = true;
                mpi.ExitFlagTest();
                return true;
            }
            m_device.reset(raw, utilities::detail::BasisDeleter());
            int64_t dim = 0;
            hd_basis_dim(raw, &dim);
            m_basis.resize((size_t)dim);
            if (!utilities::detail::Ok(hd_basis_copy_states(raw, 0, dim, m_basis.data()), "GenerateFockSpace")) {
                mpi.m_exitFlag = true;
                mpi.ExitFlagTest();
            }
            m_basisStored = true;
            m_firstState = 0;
            return dim == 0;
        }

        fock_t GetStoredDimension() const { return m_basis.size(); }

        // fermion_fock_basis.cpp:360-377
        void GetFockBasis(fock_t* buffer, const fock_t dim, utilities::MpiWrapper& mpi) const
        {
            if (m_basis.size() >= dim) {
                std::memcpy(buffer, m_basis.data(), dim * sizeof(fock_t));
            } else {
                std::cerr << "\n\tERROR in GetFockBasis: Given copy buffer is longer than the allocated "
                             "Fock space buffer dimension" << std::endl;
                mpi.m_exitFlag = true;
            }
            mpi.ExitFlagTest();
        }

        // Device basis handle (shared by copies, as the reference copies the basis
        // into the Hamiltonian: hamiltonian_base.hpp:597-600).
        std::shared_ptr<hd_basis> Device() const { return m_device; }

        private:
        std::shared_ptr<hd_basis> m_device;
    };

    // TwoLevelFockState (two_level_fermion_fock_basis.hpp:49-59, .cpp:36-70):
    // occupations of the first (LLL) and second Landau level, ordered by
    // (m_first, m_second).  On the device the pair is the combined state
    // (m_first << (n1+2)) | m_second, whose ascending order is the same.
    struct TwoLevelFockState
    {
        fock_t m_first = 0;
        fock_t m_second = 0;
        TwoLevelFockState() = default;
        TwoLevelFockState(const fock_t first, const fock_t second) : m_first(first), m_second(second) {}
        friend bool operator<(const TwoLevelFockState& a, const TwoLevelFockState& b)
        {
            return a.m_first < b.m_first || (a.m_first == b.m_first && a.m_second < b.m_second);
        }
        friend bool operator>(const TwoLevelFockState& a, const TwoLevelFockState& b) { return b < a; }
        friend bool operator==(const TwoLevelFockState& a, const TwoLevelFockState& b)
        {
            return a.m_first == b.m_first && a.m_second == b.m_second;
        }
        friend std::ostream& operator<<(std::ostream& os, const TwoLevelFockState& s)
        {
            return os << s.m_first << "\t" << s.m_second;
        }
    };

    // TestAngularMomentumSector2LL (two_level_angular_momentum_constraint.hpp:41-73)
    // with max = m_maxLz2: first-level orbital k has m = 2k - (max - 2),
    // second-level orbital k has m = 2k - max.
    struct AngularMomentumSector2LL
    {
        int m_lzTot, m_max;
        AngularMomentumSector2LL(const int lzTot, const int max) : m_lzTot(lzTot), m_max(max) {}
        bool operator()(const TwoLevelFockState& state) const
        {
            long s = 0;
            for (fock_t b = state.m_first; b; b &= b - 1) s += 2 * __builtin_ctzll(b) - (m_max - 2);
            for (fock_t b = state.m_second; b; b &= b - 1) s += 2 * __builtin_ctzll(b) - m_max;
            return s == m_lzTot;
        }
        hd_sector Descriptor(const iSize_t nbrOrbitals1) const
        {
            return hd_sector{HD_SECTOR_LZ2, m_lzTot, m_max, (int32_t)nbrOrbitals1, 0};
        }
    };

    // [](const TwoLevelFockState& state){return true;} (two_level_pseudopotential_model.cpp:151)
    struct TwoLevelNoConstraint
    {
        bool operator()(const TwoLevelFockState&) const { return true; }
        hd_sector Descriptor(const iSize_t) const { return hd_sector{HD_SECTOR_NONE, 0, 0, 0, 0}; }
    };

    // TwoLevelFermionFockBasis (src/basis/two_level_fermion_fock_basis.hpp:64-89):
    // the device basis over the 2 n1 + 2 combined orbitals; m_basis is its host
    // copy split into pairs, in the reference's (m_first, m_second) order.  The
    // second level holds n1 + 2 orbitals (the model's nbrOrbitalsLLL + 2).
    class TwoLevelFermionFockBasis
    {
        public:
        std::vector<TwoLevelFockState> m_basis;
        bool m_basisStored = false;
        fock_t m_firstState = 0;

        void Clear()
        {
            m_basis.clear();
            m_basisStored = false;
            m_firstState = 0;
            m_n1 = m_n2 = 0;
            m_device.reset();
        }

        // two_level_fermion_fock_basis.cpp:140-155
        fock_t GetFullFockSpaceDimension(const iSize_t nbrParticles, const iSize_t nbrOrbitals1,
                                         const iSize_t nbrOrbitals2) const
        {
            auto binom = [](iSize_t n, iSize_t k) -> unsigned __int128 {
                if (k > n) return 0;
                unsigned __int128 c = 1;
                for (iSize_t i = 0; i < k; ++i) c = c * (n - i) / (i + 1);
                return c;
            };
            unsigned __int128 total = 0;
            for (iSize_t n0 = 0; n0 <= nbrParticles; ++n0)
                total += binom(nbrOrbitals1, n0) * binom(nbrOrbitals2, nbrParticles - n0);
            return (fock_t)total;
        }

        fock_t GetReducedFockSpaceDimension(const iSize_t, const iSize_t, const iSize_t,
                                            const utilities::MpiWrapper&) const
        {
            return m_basis.size();
        }

        void AddState(const fock_t state1, const fock_t state2)
        {
            m_basis.push_back(TwoLevelFockState(state1, state2));
            m_basisStored = true;
        }

        // two_level_fermion_fock_basis.cpp:211-241: offset + position, or
        // _SEARCH_ERROR_CODE_ when the state is not in the basis
        void FindFockStateIndex(const fock_t offset, const TwoLevelFockState& state, fock_t& index) const
        {
            if (!m_basisStored) {
                std::cerr << "ERROR in FindFockStateIndex: BASIS WAS NOT STORED " << std::endl;
                return;
            }
            auto it = std::lower_bound(m_basis.begin(), m_basis.end(), state);
            if (it == m_basis.end() || !(*it == state)) {
                index = _SEARCH_ERROR_CODE_;
                return;
            }
            index = offset + (fock_t)(it - m_basis.begin());
        }

        void GetFirstState(const fock_t firstNodeIndex, std::vector<TwoLevelFockState>::const_iterator& it_state) const
        {
            if (firstNodeIndex != m_firstState || m_basis.empty()) {
                std::cerr << "\n\tERROR in GetFirstState: non-matching first state index." << std::endl;
                std::exit(EXIT_FAILURE);
            }
            it_state = m_basis.begin();
        }

        void GetNextState(std::vector<TwoLevelFockState>::const_iterator& it_state) const { ++it_state; }

        // two_level_fermion_fock_basis.cpp:214-337: returns true when the
        // sector is empty.  The constraint must be one of the descriptors above.
        bool GenerateFockSpace(const iSize_t nbrParticles, const iSize_t nbrOrbitals1, const iSize_t nbrOrbitals2,
                               const std::function<bool(const TwoLevelFockState& state)>& Constraint,
                               utilities::MpiWrapper& mpi)
        {
            Clear();
            hd_sector s;
            if (auto* p = Constraint.target<AngularMomentumSector2LL>()) s = p->Descriptor(nbrOrbitals1);
            else if (Constraint.target<TwoLevelNoConstraint>()) s = TwoLevelNoConstraint().Descriptor(nbrOrbitals1);
            else {
                std::cerr << "\n\tERROR in GenerateFockSpace: the constraint must be AngularMomentumSector2LL or "
                             "TwoLevelNoConstraint; a device basis cannot evaluate a host predicate." << std::endl;
                mpi.m_exitFlag = true;
                mpi.ExitFlagTest();
                return true;
            }
            if (nbrOrbitals2 != nbrOrbitals1 + 2) {
                std::cerr << "\n\tERROR in GenerateFockSpace: the second level holds nbrOrbitals1 + 2 orbitals"
                          << std::endl;
                mpi.m_exitFlag = true;
                mpi.ExitFlagTest();
                return true;
            }
            hd_basis* raw = nullptr;
            const int st = hd_basis_create(&s, nbrParticles, nbrOrbitals1 + nbrOrbitals2, &raw);
            if (st == HD_ERR_EMPTY) return true;
            if (!utilities::detail::Ok(st, "GenerateFockSpace")) {
                mpi.m_exitFlag = true;
                mpi.ExitFlagTest();
                return true;
            }
            m_device.reset(raw, utilities::detail::BasisDeleter());
            m_n1 = nbrOrbitals1;
            m_n2 = nbrOrbitals2;
            int64_t dim = 0;
            hd_basis_dim(raw, &dim);
            std::vector<fock_t> comb((size_t)dim);
            if (!utilities::detail::Ok(hd_basis_copy_states(raw, 0, dim, comb.data()), "GenerateFockSpace")) {
                mpi.m_exitFlag = true;
                mpi.ExitFlagTest();
            }
            m_basis.resize((size_t)dim);
            const fock_t lo = ((fock_t)1 << m_n2) - 1;
            for (size_t i = 0; i < comb.size(); ++i) m_basis[i] = TwoLevelFockState(comb[i] >> m_n2, comb[i] & lo);
            m_basisStored = true;
            m_firstState = 0;
            return dim == 0;
        }

        fock_t GetStoredDimension() const { return m_basis.size(); }

        // two_level_fermion_fock_basis.cpp:494-522 (each pair to its own slot;
        // the reference's loop writes every pair to buffer[0])
        void GetFockBasis(fock_t* buffer1, fock_t* buffer2, const fock_t dim, utilities::MpiWrapper& mpi) const
        {
            if (m_basis.size() >= dim) {
                for (fock_t i = 0; i < dim; ++i) {
                    buffer1[i] = m_basis[i].m_first;
                    buffer2[i] = m_basis[i].m_second;
                }
            } else {
                std::cerr << "\n\tERROR in GetFockBasis: Given copy buffer is longer than the allocated "
                             "Fock space buffer dimension" << std::endl;
                mpi.m_exitFlag = true;
            }
            mpi.ExitFlagTest();
        }

        std::shared_ptr<hd_basis> Device() const { return m_device; }
        iSize_t NbrOrbitals1() const { return m_n1; }

        private:
        std::shared_ptr<hd_basis> m_device;
        iSize_t m_n1 = 0, m_n2 = 0;
    };

    // HamiltonianData (src/hamiltonians/hamiltonian_data.hpp:58-85)
    struct HamiltonianData
    {
        iSize_t m_nbrParticles = 0;
        iSize_t m_nbrOrbitals = 0;
        fock_t m_highestOrbital = 0;
        utilities::storageMethod_t m_storageMethod = utilities::_SPARSE_CRS_;
        bool m_matrixAllocated = false;
        bool m_eigenvaluesCalculated = false;
        bool m_eigenvectorsCalculated = false;
        fock_t m_fockSpaceDim = 0;
        fock_t m_nodeDim = 0;
        iSize_t m_nbrEigenvalues = 0;

        HamiltonianData() = default;
        HamiltonianData(const iSize_t nbrParticles, const iSize_t nbrOrbitals,
                        const utilities::storageMethod_t storageMethod)
            : m_nbrParticles(nbrParticles), m_nbrOrbitals(nbrOrbitals),
              m_highestOrbital(nbrOrbitals ? (fock_t)1 << (nbrOrbitals - 1) : 0),  // hamiltonian_data.cpp:67
              m_storageMethod(storageMethod) {}
        void Clear() { *this = HamiltonianData(); }
        void SetDataDimensions(const fock_t fockSpaceDimension, const iSize_t, utilities::MpiWrapper&)
        {
            m_fockSpaceDim = fockSpaceDimension;
            m_nodeDim = fockSpaceDimension;
        }
    };

    template <typename T, class L1, class L2> class MatrixVectorFunction;
    template <typename T> class TwoLevelSpinlessFermionHamiltonian;

    // HamiltonianBase<T,B> (src/hamiltonians/hamiltonian_base.hpp)
    template <typename T, class B>
    class HamiltonianBase
    {
        static_assert(std::is_same<T, double>::value || std::is_same<T, dcmplx>::value,
                      "matrix elements are double or std::complex<double>");

        public:
        using ValueType = T;
        HamiltonianData m_data;
        B m_fockBasis;
        std::vector<double> m_eigenvalues;
        std::vector<T> m_eigenvectors;        // nev x nodeDim, eigenvector i at [i*nodeDim]
        utilities::CrsSparseMatrix<T> m_crsSparseMatrix;
        utilities::linearAlgebra::ArpackWrapper m_arpackWrapper;

        void SetFockBasis(const B& fockBasis) { m_fockBasis = fockBasis; }

        // hamiltonian_base.hpp:605-623
        void SetNbrEigenvalues(const iSize_t nbrEigenvalues, const utilities::MpiWrapper&)
        {
            if (nbrEigenvalues > m_data.m_fockSpaceDim) {
                std::cerr << "\n\tWARNING in SpinlessFermionHamiltonian: nbrEigenvalues larger than Fock space dimension:"
                          << std::endl << "\tSet number of eigenvalues to " << m_data.m_fockSpaceDim << std::endl;
                m_data.m_nbrEigenvalues = (iSize_t)m_data.m_fockSpaceDim;
            } else {
                m_data.m_nbrEigenvalues = nbrEigenvalues;
            }
        }

        // hamiltonian_base.hpp:378-449: true when every stored entry is diagonal
        bool MatrixDiagonalCheck(const int, const utilities::MpiWrapper&)
        {
            if (!m_device) return false;
            hd_ham_info_t info;
            if (!utilities::detail::Ok(hd_ham_info(m_device.get(), &info), "MatrixDiagonalCheck")) return false;
            if (info.nnz > info.dim) return false;
            std::vector<int64_t> rp((size_t)info.dim + 1);
            std::vector<int32_t> col((size_t)info.nnz);
            std::vector<T> val((size_t)info.nnz);
            if (!utilities::detail::Ok(hd_ham_copy_crs(m_device.get(), 0, info.dim, HD_F64, rp.data(), col.data(),
                                                       val.data()), "MatrixDiagonalCheck"))
                return false;
            for (int64_t i = 0; i < info.dim; ++i)
                for (int64_t k = rp[i]; k < rp[i + 1]; ++k)
                    if (col[k] != i) return false;
            m_diagonal.assign((size_t)info.dim, 0.0);
            for (int64_t i = 0; i < info.dim; ++i)
                for (int64_t k = rp[i]; k < rp[i + 1]; ++k) m_diagonal[i] = std::real(val[k]);
            return true;
        }

        // hamiltonian_base.hpp:634-758 (dense diagonalisation of small sectors)
        bool FullDiagonalize(iSize_t& nbrEigenvalues, utilities::MpiWrapper& mpi)
        {
            SetNbrEigenvalues(nbrEigenvalues, mpi);
            nbrEigenvalues = m_data.m_nbrEigenvalues;
            if (!m_device) {
                std::cerr << "\n\tERROR in FullDiagonalize: the Hamiltonian was not built" << std::endl;
                return false;
            }
            m_eigenvalues.assign(m_data.m_nbrEigenvalues, 0.0);
            const bool ok = utilities::detail::Ok(
                hd_full_diagonalize(m_device.get(), (int32_t)m_data.m_nbrEigenvalues, m_eigenvalues.data()),
                "FullDiagonalize");
            m_data.m_eigenvaluesCalculated = ok;
            m_data.m_eigenvectorsCalculated = false;  // eigenvalues only on this path
            DeallocateMatrixMemory();
            return ok;
        }

        // hamiltonian_base.hpp:769-931
        template <class L1, class L2>
        bool LanczosDiagonalize(L1*, L2*, MatrixVectorFunction<T, L1, L2>&, iSize_t& nbrEigenvalues,
                                utilities::MpiWrapper& mpi)
        {
            SetNbrEigenvalues(nbrEigenvalues, mpi);
            nbrEigenvalues = m_data.m_nbrEigenvalues;
            m_data.m_eigenvaluesCalculated = m_data.m_eigenvectorsCalculated = false;
            if (!m_data.m_matrixAllocated || !m_device) {
                std::cerr << "\n\tERROR in LanczosDiagonalize: the Hamiltonian was not built" << std::endl;
                return false;
            }
            const iSize_t nev = m_data.m_nbrEigenvalues;
            const fock_t dim = m_data.m_nodeDim;
            m_eigenvalues.assign(nev, 0.0);
            m_eigenvectors.assign((size_t)nev * dim, T(0.0));
            if (MatrixDiagonalCheck(0, mpi)) {
                // hamiltonian_base.hpp:803-872: sorted diagonal, unit vectors on each
                // degenerate set, normalised by the set's size as in the reference
                std::vector<double> sorted(m_diagonal);
                std::sort(sorted.begin(), sorted.end());
                for (iSize_t i = 0; i < nev; ++i) {
                    m_eigenvalues[i] = sorted[i];
                    double norm = 0.0;
                    for (fock_t j = 0; j < dim; ++j)
                        if (std::fabs(m_eigenvalues[i] - m_diagonal[j]) < _SPARSE_ZERO_TOL_) {
                            m_eigenvectors[(size_t)i * dim + j] = T(1.0);
                            norm += 1.0;
                        }
                    for (fock_t j = 0; j < dim; ++j) m_eigenvectors[(size_t)i * dim + j] /= norm;
                }
            } else {
                hd_lanczos_opts o{};
                o.nev = (int32_t)nev;
                o.ncv = m_arpackWrapper.m_ncv;
                o.max_restarts = m_arpackWrapper.m_maxRestarts;
                o.tol = m_arpackWrapper.m_tol;
                o.seed = m_arpackWrapper.m_seed;
                o.precision = m_arpackWrapper.m_precision;
                o.full_reorth = m_arpackWrapper.m_fullReorth ? 1 : 0;
                o.want_vectors = 1;
                o.start = HD_START_RANDOM;
                void* d_file = nullptr;
                if (m_arpackWrapper.m_deviceStart) {
                    o.start = HD_START_USER;
                    o.d_start = m_arpackWrapper.m_deviceStart;
                } else if (m_arpackWrapper.m_provideInitial) {
                    // arpack_wrapper.hpp:753-774: raw T[nodeDim]; a missing file falls
                    // back to the random start with a message
                    std::vector<T> init(dim);
                    std::ifstream f(m_arpackWrapper.m_initialVectorFile, std::ios::binary);
                    if (!f.is_open()) {
                        std::cerr << "\n\tERROR in DiagonalizeSymmetricMatrix: Cannot open initial vector file "
                                  << m_arpackWrapper.m_initialVectorFile << std::endl
                                  << "\n\tUsing random starting vector instead." << std::endl;
                    } else {
                        f.read(reinterpret_cast<char*>(init.data()), (std::streamsize)(dim * sizeof(T)));
                        if (hd_malloc(&d_file, dim * sizeof(T)) == HD_OK &&
                            hd_memcpy_h2d(d_file, init.data(), dim * sizeof(T)) == HD_OK) {
                            o.start = HD_START_USER;
                            o.d_start = d_file;
                        }
                    }
                }
                if (o.precision == HD_TILED) {
                    // the tiled layout is built from the coded values on first use
                    hd_tile_info_t ti;
                    if (hd_ham_tile_info(m_device.get(), &ti) != HD_OK &&
                        !utilities::detail::Ok(hd_ham_tile(m_device.get(), 0, 0), "LanczosDiagonalize")) {
                        if (d_file) hd_free(d_file);
                        return false;
                    }
                }
                const int st = hd_lanczos(m_device.get(), &o, m_eigenvalues.data(), &m_arpackWrapper.m_lastStats);
                if (d_file) hd_free(d_file);
                if (st != HD_OK && st != HD_ERR_NOT_CONVERGED) {
                    utilities::detail::Ok(st, "LanczosDiagonalize");
                    return false;
                }
                if (st == HD_ERR_NOT_CONVERGED)
                    std::cerr << "\n\tWARNING in LanczosDiagonalize: " << hd_last_error() << std::endl;
                for (iSize_t i = 0; i < nev; ++i)
                    if (!utilities::detail::Ok(hd_lanczos_vectors(m_device.get(), (int32_t)i, nullptr,
                                                                  m_eigenvectors.data() + (size_t)i * dim),
                                               "LanczosDiagonalize"))
                        return false;
                if (m_arpackWrapper.m_storeFinal && !m_arpackWrapper.m_finalVectorFile.empty()) {
                    // the ground state in the same raw format, for a later --use-initial run
                    std::ofstream f(m_arpackWrapper.m_finalVectorFile, std::ios::binary);
                    f.write(reinterpret_cast<const char*>(m_eigenvectors.data()), (std::streamsize)(dim * sizeof(T)));
                }
                if (st == HD_ERR_NOT_CONVERGED) return false;
            }
            DeallocateMatrixMemory();
            m_data.m_eigenvaluesCalculated = m_data.m_eigenvectorsCalculated = true;
            return true;
        }

        // hamiltonian_base.hpp:1188-1200
        void GetEigenvalues(double* buffer, const iSize_t nbrEigenvalues) const
        {
            if (nbrEigenvalues > m_eigenvalues.size())
                std::cerr << "\n\t- ERROR in GetEigenvalues: incorrect number specified" << std::endl;
            std::memcpy(buffer, m_eigenvalues.data(),
                        sizeof(double) * std::min<size_t>(nbrEigenvalues, m_eigenvalues.size()));
        }

        fock_t GetEigenvectorNodeDim() const { return m_data.m_nodeDim; }

        // hamiltonian_base.hpp:1123-1153: the upper-triangle CRS in the
        // CrsSparseMatrix file format
        void HamiltonianToFile(const std::string fileName, const io::fileFormat_t format, utilities::MpiWrapper& mpi)
        {
            if (!m_data.m_matrixAllocated) {
                std::cerr << "\n\t- Hamiltonain not allocated - failed to srite to file" << std::endl;
                return;
            }
            m_crsSparseMatrix.ToFile(fileName, format, mpi);
        }

        // hamiltonian_base.hpp:1204-1298.  Binary: uint32 nev, double
        // eigenvalues[nev], then (with vectors) the dimension in 4 bytes and
        // T vectors[nev][dim].  Text: "##  NBR EIGENVALUES", nev,
        // "##  EIGENVALUES", one value per line (15 digits), then
        // "##  FOCK SPACE DIMENSION", dim, "##  EIGENVECTORS", one entry per line.
        // Writing vectors implies writing values.
        void EigensystemToFile(const std::string fileName, bool writeValues, const bool writeVectors,
                               const io::fileFormat_t format, utilities::MpiWrapper& mpi)
        {
            if (writeVectors) writeValues = true;
            if (!m_data.m_eigenvaluesCalculated && writeValues) {
                std::cerr << "ERROR eigenvalues cannot be written to file as they are not calculated" << std::endl;
                mpi.m_exitFlag = true;
                return;
            }
            if (!m_data.m_eigenvectorsCalculated && writeVectors) {
                std::cerr << "ERROR eigenvectors cannot be written to file as they are not calculated" << std::endl;
                mpi.m_exitFlag = true;
                return;
            }
            std::ofstream f_out;
            utilities::GenFileStream(f_out, fileName, format, mpi);
            if (mpi.m_exitFlag) return;
            const iSize_t nev = m_data.m_nbrEigenvalues;
            const fock_t dim = m_data.m_fockSpaceDim;
            if (writeValues) {
                if (format == io::_TEXT_) {
                    f_out << "##  NBR EIGENVALUES\n" << nev << "\n" << "##  EIGENVALUES\n";
                    for (iSize_t i = 0; i < nev; ++i) f_out << std::setprecision(15) << m_eigenvalues[i] << "\n";
                } else {
                    f_out.write(reinterpret_cast<const char*>(&nev), sizeof(unsigned int));
                    f_out.write(reinterpret_cast<const char*>(m_eigenvalues.data()), (std::streamsize)nev * sizeof(double));
                }
            }
            if (writeVectors) {
                if (format == io::_TEXT_) {
                    f_out << "##  FOCK SPACE DIMENSION\n" << dim << "\n" << "##  EIGENVECTORS\n";
                    for (size_t i = 0; i < (size_t)nev * dim; ++i) f_out << utilities::ToStream(m_eigenvectors[i]) << "\n";
                } else {
                    const unsigned int dim32 = (unsigned int)dim;
                    f_out.write(reinterpret_cast<const char*>(&dim32), sizeof(unsigned int));
                    f_out.write(reinterpret_cast<const char*>(m_eigenvectors.data()),
                                (std::streamsize)((size_t)nev * dim * sizeof(T)));
                }
            }
        }

        // hamiltonian_base.hpp:1304-1432.  Reading vectors implies reading
        // values.  The vectors are flagged as present only when they were read
        // (the reference flags them whenever the file opened), and a truncated
        // file sets the exit flag.
        void EigensystemFromFile(const std::string fileName, bool readValues, const bool readVectors,
                                 const io::fileFormat_t format, utilities::MpiWrapper& mpi)
        {
            if (readVectors) readValues = true;
            std::ifstream f_in;
            utilities::GenFileStream(f_in, fileName, format, mpi);
            if (mpi.m_exitFlag) return;
            bool ok = true;
            if (readValues) {
                iSize_t nev = 0;
                if (format == io::_TEXT_) {
                    std::string line;
                    while (std::getline(f_in, line) && line != "##  NBR EIGENVALUES") {}
                    f_in >> nev;
                    while (std::getline(f_in, line) && line != "##  EIGENVALUES") {}
                    m_eigenvalues.assign(nev, 0.0);
                    for (iSize_t i = 0; i < nev; ++i) f_in >> m_eigenvalues[i];
                } else {
                    f_in.read(reinterpret_cast<char*>(&nev), sizeof(unsigned int));
                    m_eigenvalues.assign(f_in ? nev : 0, 0.0);
                    f_in.read(reinterpret_cast<char*>(m_eigenvalues.data()), (std::streamsize)m_eigenvalues.size() * sizeof(double));
                }
                ok = (bool)f_in;
                m_data.m_nbrEigenvalues = nev;
                m_data.m_eigenvaluesCalculated = ok;
            }
            if (readVectors && ok) {
                fock_t dim = 0;
                if (format == io::_TEXT_) {
                    std::string line;
                    while (std::getline(f_in, line) && line != "##  FOCK SPACE DIMENSION") {}
                    f_in >> dim;
                    while (std::getline(f_in, line) && line != "##  EIGENVECTORS") {}
                    m_eigenvectors.assign((size_t)m_data.m_nbrEigenvalues * dim, T(0.0));
                    for (auto& v : m_eigenvectors) utilities::FromStream(f_in, v);
                } else {
                    unsigned int dim32 = 0;
                    f_in.read(reinterpret_cast<char*>(&dim32), sizeof(unsigned int));
                    dim = dim32;
                    m_eigenvectors.assign((size_t)m_data.m_nbrEigenvalues * dim, T(0.0));
                    f_in.read(reinterpret_cast<char*>(m_eigenvectors.data()),
                              (std::streamsize)(m_eigenvectors.size() * sizeof(T)));
                }
                ok = (bool)f_in;
                m_data.m_fockSpaceDim = m_data.m_nodeDim = dim;
                m_data.m_eigenvectorsCalculated = ok;
            }
            if (!ok) {
                std::cerr << "\n\tERROR in EigensystemFromFile: " << fileName << " is truncated" << std::endl;
                mpi.m_exitFlag = true;
            }
        }

        void PrintEigensystem() const
        {
            std::cout << "\n\tEIGENVALUES:\n";
            for (double e : m_eigenvalues) std::printf("\t%.15g\n", e);
        }

        // Device handle of the full-storage matrix (null after DeallocateMatrixMemory).
        hd_hamiltonian* Device() const { return m_device.get(); }

        protected:
        // hamiltonian_base.hpp:149-180: the matrix is released after diagonalisation
        void DeallocateMatrixMemory()
        {
            m_device.reset();
            m_data.m_matrixAllocated = false;
        }

        std::shared_ptr<hd_hamiltonian> m_device;
        std::vector<double> m_diagonal;
    };

    // SpinlessFermionHamiltonian<T> (src/hamiltonians/spinless_fermion_hamiltonian.hpp)
    template <typename T>
    class SpinlessFermionHamiltonian : public HamiltonianBase<T, FermionFockBasis>
    {
        public:
        // spinless_fermion_hamiltonian.hpp:59-105
        void Initialize(const iSize_t nbrParticles, const iSize_t nbrOrbitals,
                        const utilities::storageMethod_t storageMethod, utilities::MpiWrapper& mpi)
        {
            this->m_data = HamiltonianData(nbrParticles, nbrOrbitals, storageMethod);
            if (0 >= (int)this->m_data.m_nbrParticles) {
                std::cerr << "\n\tERROR: MUST SPECIFY nbr particles > 0" << std::endl;
                mpi.m_exitFlag = true;
            }
            if (0 >= (int)this->m_data.m_nbrOrbitals) {
                std::cerr << "\n\tERROR: MUST SPECIFY nbr orbitals > 0" << std::endl;
                mpi.m_exitFlag = true;
            }
            if (this->m_data.m_nbrParticles >= 64) {
                std::cerr << "\n\n\tERROR: CANNOT ADDRESS 64 PARTICLES OR MORE.\n" << std::endl;
                mpi.m_exitFlag = true;
            }
            if (this->m_data.m_nbrParticles > this->m_data.m_nbrOrbitals) {
                std::cerr << "\n\n\tERROR: NBR PARTICLES > NBR ORBITALS - NOT ALLOWED\n" << std::endl;
                mpi.m_exitFlag = true;
            }
            mpi.ExitFlagTest();
            this->m_data.SetDataDimensions(this->m_fockBasis.GetStoredDimension(), mpi.m_nbrProcs, mpi);
            m_quadratic.clear();
            m_haveQuadratic = false;
            m_quadDiagonal = true;
        }

        void Clear()
        {
            this->DeallocateMatrixMemory();
            this->m_fockBasis.Clear();
            this->m_data.Clear();
            this->m_crsSparseMatrix.Clear();
            this->m_eigenvalues.clear();
            this->m_eigenvectors.clear();
            m_quadratic.clear();
            m_haveQuadratic = false;
            m_quadDiagonal = true;
        }

        // spinless_fermion_hamiltonian.hpp:143-257.  The k1 lists of every k2
        // (GetK1) and their E (GetEkk) are recorded and enter the device build
        // together with the quartic terms, quadratic terms first as in the
        // reference.  Diagonal-only tables (k1 = k2, both reference array
        // models) keep the pair-structure build; lists with hops (the Wannier
        // basis, any hash table) take the general build.
        template <class L>
        void Add_CdC_Terms(L* termTables, const fock_t minRow, const fock_t maxRow, utilities::MpiWrapper& mpi)
        {
            if (!CheckRange(minRow, maxRow, "Add_CdC_Terms", mpi)) return;
            const iSize_t kMax = utilities::detail::TableKMax(termTables, this->m_data.m_nbrOrbitals);
            const iSize_t w = utilities::detail::IsComplex<T>::value ? 2 : 1;
            const iSize_t maxK = std::max<iSize_t>(termTables->GetMaxKCount(), 1);
            std::vector<kState_t> buf(maxK);
            m_quadratic.assign(kMax * w, 0.0);
            m_nq1.assign(kMax, 0);
            m_q1.assign((size_t)kMax * maxK, 0);
            m_e.assign((size_t)kMax * maxK * w, 0.0);
            m_maxQ1 = maxK;
            m_quadDiagonal = true;
            for (iSize_t k = 0; k < kMax; ++k) {
                iSize_t n = 0;
                termTables->GetK1(buf.data(), n, (kState_t)k);
                m_nq1[k] = (uint8_t)n;
                for (iSize_t j = 0; j < n; ++j) {
                    const T e = termTables->GetEkk(buf[j], (kState_t)k);
                    m_q1[(size_t)k * maxK + j] = buf[j];
                    std::memcpy(&m_e[((size_t)k * maxK + j) * w], &e, sizeof(T));
                    if (buf[j] != k) m_quadDiagonal = false;
                }
                if (n > 1) m_quadDiagonal = false;
                const T e = n ? termTables->GetEkk((kState_t)k, (kState_t)k) : T(0.0);
                std::memcpy(&m_quadratic[k * w], &e, sizeof(T));
            }
            m_haveQuadratic = true;
        }

        // spinless_fermion_hamiltonian.hpp:281-436 plus the map -> CRS conversion
        // (mapped_sparse_matrix.hpp:423-462, crs_sparse_matrix.hpp:193-254).
        // Array tables (GetKTable() gives the k vector) with diagonal quadratic
        // terms use the pair-structure build; any other table class is read
        // through GetK1 / GetVkkkk into k1 lists for the general build
        // (QuarticTermHashTables, term_hash_tables.cpp:96-130).
        template <class L>
        void Add_CdCdCC_Terms(L* termTables, const fock_t minRow, const fock_t maxRow, utilities::MpiWrapper& mpi)
        {
            if (!CheckRange(minRow, maxRow, "Add_CdCdCC_Terms", mpi)) return;
            std::shared_ptr<hd_tables> tables;
            if constexpr (utilities::detail::IsArrayQuartic<L>::value) {
                if (!m_haveQuadratic || m_quadDiagonal) {
                    tables = ArrayTables(termTables, mpi);
                    if (!tables) return;
                }
            }
            if (!tables) {
                tables = ListTables(termTables, mpi);
                if (!tables) return;
            }
            hd_hamiltonian* rawH = nullptr;
            const int mask = tables_is_hash_ ? (m_valueMask & 3) | 1 : m_valueMask;
            if (!utilities::detail::Ok(hd_ham_build(this->m_fockBasis.Device().get(), tables.get(), HD_STORAGE_FULL,
                                                    _SPARSE_ZERO_TOL_, mask, &rawH),
                                       "Add_CdCdCC_Terms")) {
                mpi.m_exitFlag = true;
                return;
            }
            this->m_device.reset(rawH, utilities::detail::HamDeleter());
            this->m_data.m_matrixAllocated = true;
            this->m_data.m_storageMethod = utilities::_SPARSE_CRS_;
            this->m_crsSparseMatrix.AdoptDevice(this->m_fockBasis.Device(), tables, _SPARSE_ZERO_TOL_,
                                                (int)this->m_data.m_fockSpaceDim);
        }

        // Resident value copies of the device build (hdiag.h: bit 0 FP64, bit 1
        // FP32, bit 2 coded FP64); FP64 is always kept for Simple() on host vectors.
        void SetValueFormats(const int mask) { m_valueMask = mask | 1; }

        private:
        bool CheckRange(const fock_t minRow, const fock_t maxRow, const char* where, utilities::MpiWrapper& mpi)
        {
            if (!this->m_fockBasis.Device()) {
                std::cerr << "\n\tERROR in " << where << ": no device Fock basis (GenerateFockSpace first)" << std::endl;
                mpi.m_exitFlag = true;
                return false;
            }
            if (minRow != 0 || maxRow < this->m_data.m_fockSpaceDim) {
                std::cerr << "\n\tERROR in " << where << ": the device build covers all rows [0, "
                          << this->m_data.m_fockSpaceDim << "); got [" << minRow << ", " << maxRow << ")" << std::endl;
                mpi.m_exitFlag = true;
                return false;
            }
            return true;
        }

        template <class L>
        std::shared_ptr<hd_tables> ArrayTables(L* termTables, utilities::MpiWrapper& mpi)
        {
            tables_is_hash_ = false;
            std::vector<kState_t>* kt = termTables->GetKTable();
            auto* vt = termTables->GetVTable();
            const size_t n = vt->size();
            // the table's layout follows from its size: kMax^3 (sphere,
            // term_tables.cpp:36-44) or (kMax+1) kMax^2 / 2 (flux lattice fold)
            int layout = -1;
            uint32_t kMax = 0;
            for (uint32_t K = 1; K <= 64 && layout < 0; ++K) {
                if ((size_t)K * K * K == n && kt->size() == n) { layout = HD_LAYOUT_SPHERE; kMax = K; }
                else if ((size_t)(K + 1) * K * K / 2 == n && kt->size() == n) { layout = HD_LAYOUT_OFL; kMax = K; }
            }
            if (layout < 0) {
                std::cerr << "\n\tERROR in Add_CdCdCC_Terms: table size " << n
                          << " matches neither the sphere nor the flux-lattice layout" << std::endl;
                mpi.m_exitFlag = true;
                return nullptr;
            }
            if (m_haveQuadratic && m_quadratic.size() != kMax * (utilities::detail::IsComplex<T>::value ? 2u : 1u)) {
                std::cerr << "\n\tERROR in Add_CdCdCC_Terms: quadratic and quartic tables disagree on kMax" << std::endl;
                mpi.m_exitFlag = true;
                return nullptr;
            }
            hd_tables* rawT = nullptr;
            if (!utilities::detail::Ok(hd_tables_create(layout, kMax, kt->data(), reinterpret_cast<const double*>(vt->data()),
                                                        utilities::detail::IsComplex<T>::value ? 1 : 0,
                                                        m_haveQuadratic ? m_quadratic.data() : nullptr, &rawT),
                                       "Add_CdCdCC_Terms")) {
                mpi.m_exitFlag = true;
                return nullptr;
            }
            return std::shared_ptr<hd_tables>(rawT, utilities::detail::TablesDeleter());
        }

        // every (k2,k3,k4) of the table through GetK1 / GetVkkkk
        template <class L>
        std::shared_ptr<hd_tables> ListTables(L* termTables, utilities::MpiWrapper& mpi)
        {
            tables_is_hash_ = true;
            const size_t K = this->m_data.m_nbrOrbitals;
            const size_t w = utilities::detail::IsComplex<T>::value ? 2 : 1;
            const iSize_t maxK = std::max<iSize_t>(termTables->GetMaxKCount(), 1);
            if (maxK > 255 || (m_haveQuadratic && m_maxQ1 > 255)) {
                std::cerr << "\n\tERROR in Add_CdCdCC_Terms: more than 255 k1 per key" << std::endl;
                mpi.m_exitFlag = true;
                return nullptr;
            }
            std::vector<kState_t> buf(maxK);
            std::vector<uint8_t> nk1(K * K * K, 0);
            std::vector<uint16_t> k1(K * K * K * maxK, 0);
            std::vector<double> v(K * K * K * maxK * w, 0.0);
            for (size_t k4 = 0; k4 < K; ++k4)
                for (size_t k3 = 0; k3 < K; ++k3)
                    for (size_t k2 = 0; k2 < K; ++k2) {
                        const size_t e = (k4 * K + k3) * K + k2;
                        iSize_t n = 0;
                        termTables->GetK1(buf.data(), n, (kState_t)k2, (kState_t)k3, (kState_t)k4);
                        nk1[e] = (uint8_t)n;
                        for (iSize_t j = 0; j < n; ++j) {
                            k1[e * maxK + j] = buf[j];
                            const T val = termTables->GetVkkkk(buf[j], (kState_t)k2, (kState_t)k3, (kState_t)k4);
                            std::memcpy(&v[(e * maxK + j) * w], &val, sizeof(T));
                        }
                    }
            hd_tables* rawT = nullptr;
            if (!utilities::detail::Ok(
                    hd_tables_create_hash((uint32_t)K, utilities::detail::IsComplex<T>::value ? 1 : 0, (int32_t)maxK,
                                          nk1.data(), k1.data(), v.data(), (int32_t)m_maxQ1,
                                          m_haveQuadratic ? m_nq1.data() : nullptr, m_q1.data(), m_e.data(), &rawT),
                    "Add_CdCdCC_Terms")) {
                mpi.m_exitFlag = true;
                return nullptr;
            }
            return std::shared_ptr<hd_tables>(rawT, utilities::detail::TablesDeleter());
        }

        std::vector<double> m_quadratic;
        std::vector<uint8_t> m_nq1;
        std::vector<uint16_t> m_q1;
        std::vector<double> m_e;
        iSize_t m_maxQ1 = 1;
        bool m_quadDiagonal = true;
        bool m_haveQuadratic = false;
        bool tables_is_hash_ = false;
        int m_valueMask = 1 | 2;
    };

    // TwoLevelSpinlessFermionHamiltonian (src/hamiltonians/two_level_spinless_fermion_hamiltonian.hpp):
    // same-level terms only.  The term tables of each call are recorded and the
    // device matrix is rebuilt after every Add_CdCdCC_Terms call from all
    // terms added so far, so that after the level-0 call the matrix holds the
    // level-0 terms as the reference's does.  The model's order is required:
    // Add_CdC_Terms (level 0, then 1) before Add_CdCdCC_Terms level 0, then 1
    // (two_level_pseudopotential_model.cpp:188-200); the device build sums each
    // diagonal entry in that order with the Trim between the levels.
    template <typename T>
    class TwoLevelSpinlessFermionHamiltonian : public HamiltonianBase<T, TwoLevelFermionFockBasis>
    {
        static_assert(std::is_same<T, double>::value, "the two-level Hamiltonian is real");

        public:
        // two_level_spinless_fermion_hamiltonian.hpp:66-114
        void Initialize(const iSize_t nbrParticles, const iSize_t nbrOrbitals,
                        const utilities::storageMethod_t storageMethod, utilities::MpiWrapper& mpi)
        {
            this->m_data = HamiltonianData(nbrParticles, nbrOrbitals, storageMethod);
            m_nbrOrbitals1 = (nbrOrbitals - 2) / 2;
            m_nbrOrbitals2 = m_nbrOrbitals1 + 2;
            if (0 >= (int)nbrParticles || nbrParticles >= 64 || nbrParticles > nbrOrbitals) {
                std::cerr << "\n\tERROR: invalid number of particles " << nbrParticles << std::endl;
                mpi.m_exitFlag = true;
            }
            if (nbrOrbitals <= 2 || (nbrOrbitals & 1)) {
                std::cerr << "\n\tERROR: MUST SPECIFY nbr orbitals > 2 and an even number of them." << std::endl;
                mpi.m_exitFlag = true;
            }
            mpi.ExitFlagTest();
            this->m_data.SetDataDimensions(this->m_fockBasis.GetStoredDimension(), mpi.m_nbrProcs, mpi);
            for (int l = 0; l < 2; ++l) {
                m_eps[l].clear();
                m_kt[l].clear();
                m_vt[l].clear();
            }
            m_quarticLevels = 0;
        }

        void Clear()
        {
            this->DeallocateMatrixMemory();
            this->m_fockBasis.Clear();
            this->m_data.Clear();
            this->m_crsSparseMatrix.Clear();
            this->m_eigenvalues.clear();
            this->m_eigenvectors.clear();
            for (int l = 0; l < 2; ++l) {
                m_eps[l].clear();
                m_kt[l].clear();
                m_vt[l].clear();
            }
            m_quarticLevels = 0;
        }

        // two_level_spinless_fermion_hamiltonian.hpp:143-262 (diagonal E_k only,
        // GetK1(k2) = k2 on the sphere tables)
        template <class Tab>
        void Add_CdC_Terms(Tab* termTables, const iSize_t level1, const iSize_t level2, const fock_t minRow,
                           const fock_t maxRow, utilities::MpiWrapper& mpi)
        {
            if (!CheckLevels(level1 == level2, level1, "Add_CdC_Terms", mpi) || !CheckRange(minRow, maxRow, mpi)) return;
            if (m_quarticLevels) {
                std::cerr << "\n\tERROR in Add_CdC_Terms: quadratic terms must come before the quartic terms" << std::endl;
                mpi.m_exitFlag = true;
                return;
            }
            const iSize_t L = level1 ? m_nbrOrbitals2 : m_nbrOrbitals1;
            if (termTables->GetDimension() != L) {
                std::cerr << "\n\tERROR in Add_CdC_Terms: level " << level1 << " table has "
                          << termTables->GetDimension() << " entries, expected " << L << std::endl;
                mpi.m_exitFlag = true;
                return;
            }
            m_eps[level1].assign(L, 0.0);
            for (iSize_t k = 0; k < L; ++k) {
                kState_t buf[8];
                iSize_t n = 0;
                termTables->GetK1(buf, n, (kState_t)k);
                if (n > 1 || (n == 1 && buf[0] != k)) {
                    std::cerr << "\n\tERROR in Add_CdC_Terms: only diagonal quadratic terms are supported" << std::endl;
                    mpi.m_exitFlag = true;
                    return;
                }
                m_eps[level1][k] = n ? termTables->GetEkk((kState_t)k, (kState_t)k) : 0.0;
            }
        }

        // two_level_spinless_fermion_hamiltonian.hpp:280-480
        template <class Tab>
        void Add_CdCdCC_Terms(Tab* termTables, const iSize_t level1, const iSize_t level2, const iSize_t level3,
                              const iSize_t level4, const fock_t minRow, const fock_t maxRow, utilities::MpiWrapper& mpi)
        {
            const bool same = level1 == level2 && level2 == level3 && level3 == level4;
            if (!CheckLevels(same, level1, "Add_CdCdCC_Terms", mpi) || !CheckRange(minRow, maxRow, mpi)) return;
            if ((int)level1 != m_quarticLevels) {
                std::cerr << "\n\tERROR in Add_CdCdCC_Terms: the levels are added in the order 0, 1" << std::endl;
                mpi.m_exitFlag = true;
                return;
            }
            const iSize_t L = level1 ? m_nbrOrbitals2 : m_nbrOrbitals1;
            if (termTables->GetKTable()->size() != (size_t)L * L * L || termTables->GetVTable()->size() != (size_t)L * L * L) {
                std::cerr << "\n\tERROR in Add_CdCdCC_Terms: level " << level1 << " needs a sphere table with kMax "
                          << L << std::endl;
                mpi.m_exitFlag = true;
                return;
            }
            m_kt[level1] = *termTables->GetKTable();
            m_vt[level1] = *termTables->GetVTable();
            ++m_quarticLevels;
            Rebuild(mpi);
        }

        void SetValueFormats(const int mask) { m_valueMask = mask | 1; }

        private:
        bool CheckLevels(const bool same, const iSize_t level, const char* where, utilities::MpiWrapper& mpi)
        {
            if (level > 1) {
                std::cerr << "\n\tERROR in " << where << ": level specification incorrect (set to 0,1 only)." << std::endl;
                mpi.m_exitFlag = true;
                return false;
            }
            if (!same) {
                std::cerr << "\n\tERROR in " << where << ": only coded for terms in the same level so far." << std::endl;
                mpi.m_exitFlag = true;
                return false;
            }
            return true;
        }

        bool CheckRange(const fock_t minRow, const fock_t maxRow, utilities::MpiWrapper& mpi)
        {
            if (!this->m_fockBasis.Device() || this->m_fockBasis.NbrOrbitals1() != m_nbrOrbitals1) {
                std::cerr << "\n\tERROR: no matching device two-level basis (GenerateFockSpace first)" << std::endl;
                mpi.m_exitFlag = true;
                return false;
            }
            if (minRow != 0 || maxRow < this->m_data.m_fockSpaceDim) {
                std::cerr << "\n\tERROR: the device build covers all rows" << std::endl;
                mpi.m_exitFlag = true;
                return false;
            }
            return true;
        }

        void Rebuild(utilities::MpiWrapper& mpi)
        {
            const iSize_t L[2] = {m_nbrOrbitals1, m_nbrOrbitals2};
            std::vector<kState_t> kt[2];
            std::vector<double> vt[2];
            for (int l = 0; l < 2; ++l) {
                if (m_kt[l].empty()) {  // not added yet: no terms in this level
                    kt[l].assign((size_t)L[l] * L[l] * L[l], (kState_t)0xFFFF);
                    vt[l].assign((size_t)L[l] * L[l] * L[l], 0.0);
                } else {
                    kt[l] = m_kt[l];
                    vt[l] = m_vt[l];
                }
            }
            hd_tables* rawT = nullptr;
            if (!utilities::detail::Ok(hd_tables_create_two_level(
                                           m_nbrOrbitals1, kt[0].data(), vt[0].data(),
                                           m_eps[0].empty() ? nullptr : m_eps[0].data(), kt[1].data(), vt[1].data(),
                                           m_eps[1].empty() ? nullptr : m_eps[1].data(), &rawT),
                                       "Add_CdCdCC_Terms")) {
                mpi.m_exitFlag = true;
                return;
            }
            std::shared_ptr<hd_tables> tables(rawT, utilities::detail::TablesDeleter());
            hd_hamiltonian* rawH = nullptr;
            this->m_device.reset();
            if (!utilities::detail::Ok(hd_ham_build(this->m_fockBasis.Device().get(), tables.get(), HD_STORAGE_FULL,
                                                    _SPARSE_ZERO_TOL_, m_valueMask, &rawH),
                                       "Add_CdCdCC_Terms")) {
                mpi.m_exitFlag = true;
                return;
            }
            this->m_device.reset(rawH, utilities::detail::HamDeleter());
            this->m_data.m_matrixAllocated = true;
            this->m_data.m_storageMethod = utilities::_SPARSE_CRS_;
            this->m_crsSparseMatrix.AdoptDevice(this->m_fockBasis.Device(), tables, _SPARSE_ZERO_TOL_,
                                                (int)this->m_data.m_fockSpaceDim);
        }

        iSize_t m_nbrOrbitals1 = 0, m_nbrOrbitals2 = 0;
        std::vector<double> m_eps[2];
        std::vector<kState_t> m_kt[2];
        std::vector<double> m_vt[2];
        int m_quarticLevels = 0;
        int m_valueMask = 1 | 2;
    };

    // MatrixVectorFunction<T,L1,L2> (src/hamiltonians/matrix_vector_routines.hpp:40-135)
    template <typename T, class L1, class L2>
    class MatrixVectorFunction
    {
        public:
        // the one-level and the two-level Hamiltonian constructors (:56-91)
        MatrixVectorFunction(SpinlessFermionHamiltonian<T>& hamiltonian, const utilities::MpiWrapper&)
            : m_device([&hamiltonian]() { return hamiltonian.Device(); }),
              UPLO(hamiltonian.m_arpackWrapper.GetUpperOrLower()) {}
        MatrixVectorFunction(TwoLevelSpinlessFermionHamiltonian<T>& hamiltonian, const utilities::MpiWrapper&)
            : m_device([&hamiltonian]() { return hamiltonian.Device(); }),
              UPLO(hamiltonian.m_arpackWrapper.GetUpperOrLower()) {}

        // matrix_vector_routines.hpp:94-104: y = H x on host vectors of length nodeDim
        void Simple(T* x, T* y, const utilities::MpiWrapper& mpi)
        {
            if (!utilities::detail::Ok(hd_spmv_host(m_device(), x, y), "MatrixVectorFunction::Simple"))
                const_cast<utilities::MpiWrapper&>(mpi).m_exitFlag = true;
        }

        // The same product on device vectors (asynchronous on the matrix's stream).
        int SimpleDevice(const T* d_x, T* d_y, const int precision = HD_F64)
        {
            return hd_spmv(m_device(), precision, d_x, d_y);
        }

        private:
        std::function<hd_hamiltonian*()> m_device;
        char UPLO;
    };

    // Observables (src/observables/fermion_observables.hpp:38-626) of the
    // Hamiltonian's eigenvectors, computed on the device (hd_obs_*): the
    // eigenvector is uploaded, the basis is the Hamiltonian's device basis.
    // Differences a caller sees: an eigenvector index equal to the number of
    // eigenvalues is rejected (the reference's test lets it through and reads
    // past the vectors), and GetMostProbable returns real eigenvectors'
    // amplitudes as (value, 0) where the reference reinterprets the doubles.
    struct Observables
    {
        template <class H>
        void DensityFunction(const H& hamiltonian, const iSize_t eigenvectorIndex, double* probabilityList,
                             utilities::MpiWrapper& mpi)
        {
            Device<H> d;
            if (!d.Load(hamiltonian, eigenvectorIndex, "DensityFunction", mpi)) return;
            Check(hd_obs_density(d.basis, d.cplx, d.vec, probabilityList), mpi);
        }

        template <class H>
        void DensityDensityFunction(const H& hamiltonian, const iSize_t eigenvectorIndex, double* probabilityList,
                                    utilities::MpiWrapper& mpi)
        {
            Device<H> d;
            if (!d.Load(hamiltonian, eigenvectorIndex, "DensityDensityFunction", mpi)) return;
            Check(hd_obs_density_density(d.basis, d.cplx, d.vec, probabilityList), mpi);
        }

        template <class H>
        void TranslationalDensityDensityFunction(const H& hamiltonian, const iSize_t eigenvectorIndex,
                                                 std::function<kState_t(const kState_t, const kState_t)> StateAddition,
                                                 std::function<kState_t(const kState_t, const kState_t)> StateSubtraction,
                                                 std::vector<double>& functionList, utilities::MpiWrapper& mpi)
        {
            Device<H> d;
            if (!d.Load(hamiltonian, eigenvectorIndex, "TranslationalDensityDensityFunction", mpi)) return;
            const iSize_t norb = hamiltonian.m_data.m_nbrOrbitals;
            const iSize_t nq = (iSize_t)functionList.size();
            std::vector<kState_t> add((size_t)norb * nq), sub((size_t)norb * nq);
            for (iSize_t k = 0; k < norb; ++k)
                for (iSize_t q = 0; q < nq; ++q) {
                    add[(size_t)k * nq + q] = StateAddition((kState_t)k, (kState_t)q);
                    sub[(size_t)k * nq + q] = StateSubtraction((kState_t)k, (kState_t)q);
                }
            Check(hd_obs_translational_dd(d.basis, d.cplx, d.vec, (int32_t)nq, add.data(), sub.data(),
                                          functionList.data()), mpi);
        }

        template <class H>
        void RotationalDensityDensityFunction(const H& hamiltonian, const iSize_t eigenvectorIndex,
                                              std::function<kState_t(const kState_t, const int)> R60,
                                              std::function<kState_t(const kState_t, const int)> InverseR60,
                                              dcmplx* functionList, utilities::MpiWrapper& mpi)
        {
            (void)R60;  // the reference evaluates only the inverse rotation
            Device<H> d;
            if (!d.Load(hamiltonian, eigenvectorIndex, "RotationalDensityDensityFunction", mpi)) return;
            const iSize_t norb = hamiltonian.m_data.m_nbrOrbitals;
            std::vector<kState_t> inv((size_t)norb * 6);
            for (iSize_t k = 0; k < norb; ++k)
                for (int a = 0; a < 6; ++a) inv[(size_t)k * 6 + a] = InverseR60((kState_t)k, a);
            double out[12];
            if (Check(hd_obs_rotational_dd(d.basis, d.cplx, d.vec, inv.data(), out), mpi))
                for (int m = 0; m < 6; ++m) functionList[m] = dcmplx(out[2 * m], out[2 * m + 1]);
        }

        template <class H>
        double GetParticipationRatio(const H& hamiltonian, const iSize_t eigenvectorIndex, utilities::MpiWrapper& mpi)
        {
            Device<H> d;
            if (!d.Load(hamiltonian, eigenvectorIndex, "GetParticipationRatio", mpi)) return 0.0;
            double r = 0.0;
            Check(hd_obs_participation_ratio(d.basis, d.cplx, d.vec, &r), mpi);
            return r;
        }

        // states and amplitudes hold nbrStates entries per eigenvector
        template <class H>
        void GetMostProbable(const H& hamiltonian, const iSize_t nbrStates, fock_t* states, dcmplx* amplitudes,
                             utilities::MpiWrapper& mpi)
        {
            using T = typename std::remove_const<
                typename std::remove_reference<decltype(hamiltonian.m_eigenvectors[0])>::type>::type;
            if (!hamiltonian.m_data.m_eigenvectorsCalculated) return;
            const iSize_t n = (iSize_t)std::min<fock_t>(hamiltonian.m_data.m_fockSpaceDim, nbrStates);
            std::vector<T> amp(n);
            for (iSize_t e = 0; e < hamiltonian.m_data.m_nbrEigenvalues; ++e) {
                Device<H> d;
                if (!d.Load(hamiltonian, e, "GetMostProbable", mpi)) return;
                if (!Check(hd_obs_most_probable(d.basis, d.cplx, d.vec, (int32_t)n, states + (size_t)e * n, amp.data()),
                           mpi))
                    return;
                for (iSize_t i = 0; i < n; ++i) amplitudes[(size_t)e * n + i] = dcmplx(amp[i]);
            }
        }

        private:
        static bool Check(const int status, utilities::MpiWrapper& mpi)
        {
            if (utilities::detail::Ok(status, "Observables")) return true;
            mpi.m_exitFlag = true;
            return false;
        }

        // one eigenvector uploaded to the device, released with the helper
        template <class H>
        struct Device
        {
            const hd_basis* basis = nullptr;
            void* vec = nullptr;
            int32_t cplx = 0;
            ~Device()
            {
                if (vec) hd_free(vec);
            }
            bool Load(const H& hamiltonian, const iSize_t index, const char* where, utilities::MpiWrapper& mpi)
            {
                if (!hamiltonian.m_data.m_eigenvectorsCalculated) {
                    std::cerr << "\n\tERROR in " << where << ": eigenvectors were not calculated" << std::endl;
                    std::exit(EXIT_FAILURE);  // fermion_observables.hpp:101-104
                }
                if (index >= hamiltonian.m_data.m_nbrEigenvalues) {
                    std::cerr << "\n\tERROR in " << where << ": eigenvectorIndex out of range." << std::endl;
                    return false;
                }
                basis = hamiltonian.m_fockBasis.Device().get();
                if (!basis) {
                    std::cerr << "\n\tERROR in " << where << ": no device Fock basis" << std::endl;
                    mpi.m_exitFlag = true;
                    return false;
                }
                using T = typename std::remove_const<
                    typename std::remove_reference<decltype(hamiltonian.m_eigenvectors[0])>::type>::type;
                cplx = utilities::detail::IsComplex<T>::value ? 1 : 0;
                const fock_t dim = hamiltonian.GetEigenvectorNodeDim();
                const size_t bytes = (size_t)dim * sizeof(T);
                if (!Check(hd_malloc(&vec, std::max<size_t>(bytes, 8)), mpi)) return false;
                return Check(hd_memcpy_h2d(vec, hamiltonian.m_eigenvectors.data() + (size_t)index * dim, bytes), mpi);
            }
        };
    };

    // Term tables with the reference's accessors and file formats
    // (term_tables_base.hpp:44-345; sphere term_tables.cpp:36-169; flux-lattice
    // term_tables.cpp:36-221), for callers that have no table classes of their own.
    template <typename T>
    class TermTables
    {
        public:
        TermTables() : m_fileHeader("# File contains term table data in array format") {}
        virtual ~TermTables() = default;

        // term_tables_base.hpp:77-93: both arrays get CalculateDim(kMax) entries
        virtual void Initialize(const kState_t kMax, const utilities::MpiWrapper&)
        {
            m_kMax = kMax;
            const iSize_t dim = CalculateDim(kMax);
            m_kTable.assign(dim, 0);
            m_vTable.assign(dim, T(0.0));
        }
        void Clear()
        {
            m_kTable.clear();
            m_vTable.clear();
        }

        virtual void GetK1(kState_t*, iSize_t&, const kState_t, const kState_t, const kState_t) const {}
        virtual void SetK1(const kState_t, const kState_t, const kState_t, const kState_t) {}
        virtual T GetVkkkk(const kState_t, const kState_t, const kState_t, const kState_t) const { return T(0.0); }
        virtual void SetVkkkk(const T, const kState_t, const kState_t, const kState_t, const kState_t) {}
        virtual void GetK1(kState_t*, iSize_t&, const kState_t) const {}
        virtual T GetEkk(const kState_t, const kState_t) const { return T(0.0); }
        virtual void ToFile(const std::string fileName, const io::fileFormat_t format,
                            utilities::MpiWrapper& mpi) const = 0;
        virtual void FromFile(const std::string fileName, const io::fileFormat_t format,
                              utilities::MpiWrapper& mpi) = 0;

        iSize_t GetMaxKCount() const { return 1; }
        iSize_t GetDimension() const { return (iSize_t)m_vTable.size(); }
        std::vector<kState_t>* GetKTable() { return &m_kTable; }
        std::vector<T>* GetVTable() { return &m_vTable; }
        kState_t GetKMax() const { return m_kMax; }

        // term_tables_base.hpp:161-169
        void SetCoefficient(const T coefficient)
        {
            for (auto& v : m_vTable) v *= coefficient;
        }

        protected:
        virtual iSize_t CalculateDim(const kState_t kMax) const = 0;

        // term_tables_base.hpp:179-247.  Binary: uint32 header length, the
        // header, uint32 label count (2 or 4), kMax in a 4-byte field, then the
        // k table (4 labels only, uint16 x dim) and the value table (T x dim).
        // The reference writes that field from its 2-byte m_kMax, so its upper
        // two bytes are whatever follows m_kMax in memory; they are written as
        // zero here and ignored on reading.  Text: the header line, the label
        // count, kMax, then "k1 k1 value" per k (2 labels) or "k1 k2 k3 k4
        // value" over k4, k3, k2 with k1 from GetK1 (4 labels), tab separated.
        void ToFileBase(const std::string fileName, const io::fileFormat_t format, const iSize_t nbrLabels,
                        utilities::MpiWrapper& mpi) const
        {
            std::ofstream f_out;
            utilities::GenFileStream(f_out, fileName, format, mpi);
            if (!mpi.m_exitFlag) {
                if (format == io::_BINARY_) {
                    const iSize_t strSize = (iSize_t)m_fileHeader.size();
                    f_out.write(reinterpret_cast<const char*>(&strSize), sizeof(iSize_t));
                    f_out.write(m_fileHeader.data(), strSize);
                    f_out.write(reinterpret_cast<const char*>(&nbrLabels), sizeof(iSize_t));
                    const iSize_t kMaxField = m_kMax;
                    f_out.write(reinterpret_cast<const char*>(&kMaxField), sizeof(iSize_t));
                    if (nbrLabels == 4)
                        f_out.write(reinterpret_cast<const char*>(m_kTable.data()),
                                    (std::streamsize)(m_kTable.size() * sizeof(kState_t)));
                    if (nbrLabels == 2 || nbrLabels == 4)
                        f_out.write(reinterpret_cast<const char*>(m_vTable.data()),
                                    (std::streamsize)(m_vTable.size() * sizeof(T)));
                } else {
                    f_out << m_fileHeader << "\n" << nbrLabels << "\n" << m_kMax << "\n";
                    if (nbrLabels == 2) {
                        for (kState_t k1 = 0; k1 < m_kMax; ++k1)
                            f_out << k1 << "\t" << k1 << "\t" << utilities::ToStream(m_vTable[k1]) << "\n";
                    } else if (nbrLabels == 4) {
                        for (kState_t k4 = 0; k4 < m_kMax; ++k4)
                            for (kState_t k3 = 0; k3 < m_kMax; ++k3)
                                for (kState_t k2 = 0; k2 < m_kMax; ++k2) {
                                    kState_t k1 = 0;
                                    iSize_t n = 0;
                                    this->GetK1(&k1, n, k2, k3, k4);
                                    f_out << k1 << "\t" << k2 << "\t" << k3 << "\t" << k4 << "\t"
                                          << utilities::ToStream(this->GetVkkkk(k1, k2, k3, k4)) << "\n";
                                }
                    }
                }
            }
            mpi.ExitFlagTest();
        }

        // term_tables_base.hpp:252-345: the header must match; a mismatch is
        // reported and sets the exit flag, as in the reference.  A truncated
        // file is reported the same way.
        void FromFileBase(const std::string fileName, const io::fileFormat_t format, utilities::MpiWrapper& mpi)
        {
            std::ifstream f_in;
            utilities::GenFileStream(f_in, fileName, format, mpi);
            if (!mpi.m_exitFlag) {
                iSize_t nbrLabels = 0;
                std::string fileHeader;
                bool ok = true;
                if (format == io::_BINARY_) {
                    iSize_t strSize = 0;
                    f_in.read(reinterpret_cast<char*>(&strSize), sizeof(iSize_t));
                    if (!f_in || strSize > 4096) strSize = 0;
                    fileHeader.resize(strSize);
                    if (strSize) f_in.read(&fileHeader[0], strSize);
                    if (fileHeader != m_fileHeader) return BadHeader(fileHeader, mpi);
                    iSize_t kMaxField = 0;
                    f_in.read(reinterpret_cast<char*>(&nbrLabels), sizeof(iSize_t));
                    f_in.read(reinterpret_cast<char*>(&kMaxField), sizeof(iSize_t));
                    this->Initialize((kState_t)(kMaxField & 0xFFFFu), mpi);
                    if (nbrLabels == 4)
                        f_in.read(reinterpret_cast<char*>(m_kTable.data()),
                                  (std::streamsize)(m_kTable.size() * sizeof(kState_t)));
                    if (nbrLabels == 2 || nbrLabels == 4)
                        f_in.read(reinterpret_cast<char*>(m_vTable.data()),
                                  (std::streamsize)(m_vTable.size() * sizeof(T)));
                    ok = (bool)f_in;
                } else {
                    std::getline(f_in, fileHeader);
                    if (fileHeader != m_fileHeader) return BadHeader(fileHeader, mpi);
                    kState_t kMax = 0;
                    f_in >> nbrLabels >> kMax;
                    this->Initialize(kMax, mpi);
                    if (nbrLabels == 2) {
                        for (iSize_t i = 0; i < m_vTable.size(); ++i) {
                            kState_t k1, k2;
                            f_in >> k1 >> k2;
                            utilities::FromStream(f_in, m_vTable[i]);
                        }
                    } else if (nbrLabels == 4) {
                        for (kState_t k4 = 0; k4 < m_kMax; ++k4)
                            for (kState_t k3 = 0; k3 < m_kMax; ++k3)
                                for (kState_t k2 = 0; k2 < m_kMax; ++k2) {
                                    kState_t k1, dummy;
                                    T v;
                                    f_in >> k1 >> dummy >> dummy >> dummy;
                                    utilities::FromStream(f_in, v);
                                    this->SetK1(k1, k2, k3, k4);
                                    this->SetVkkkk(v, k1, k2, k3, k4);
                                }
                    }
                    ok = (bool)f_in;
                }
                if (!ok) {
                    std::cerr << "\n\tFile read in error - " << fileName << " is truncated" << std::endl;
                    mpi.m_exitFlag = true;
                }
            }
            mpi.ExitFlagTest();
        }

        std::vector<T> m_vTable;
        std::vector<kState_t> m_kTable;
        kState_t m_kMax = 0;
        std::string m_fileHeader;

        private:
        void BadHeader(const std::string& fileHeader, utilities::MpiWrapper& mpi)
        {
            std::cerr << "\n\tFile read in error - wrong file header " << fileHeader << std::endl;
            mpi.m_exitFlag = true;
            mpi.ExitFlagTest();
        }
    };

    // Sphere QuarticTermTables (fqhe_sphere term_tables.cpp:96-187): index
    // (k4*kMax+k3)*kMax+k2, kMax^3 entries
    class SphereQuarticTermTables : public TermTables<double>
    {
        public:
        // BaseGenerate4KTable + BaseTermsFromPseudopotentials
        // (pseudopotential_model_base.hpp:97-229) on the C ABI.  FindM1 stores
        // the label of m_maxLz2 + 2 where m1 is out of range; m_maxLz2 = maxLz
        // for one level (label 0xFFFF), n1 + 1 for the two-level model's first
        // level (maxLz = n1 - 1: label 0xFFFE).
        bool FromPseudopotentials(const int maxLz, const std::vector<double>& pseudopotentials,
                                  const utilities::MpiWrapper& mpi, const int maxLz2 = -1)
        {
            Initialize((kState_t)(maxLz + 1), mpi);
            const bool ok = utilities::detail::Ok(hd_sphere_ktable(maxLz, m_kTable.data()), "QuarticTermTables") &&
                            utilities::detail::Ok(hd_sphere_vtable(maxLz, pseudopotentials.data(),
                                                                   (int32_t)pseudopotentials.size(), m_vTable.data()),
                                                  "QuarticTermTables");
            if (ok && maxLz2 >= 0) {
                const kState_t err = (kState_t)((-(maxLz2 + 2) + maxLz) / 2);
                for (auto& k : m_kTable)
                    if (k >= m_kMax) k = err;
            }
            return ok;
        }
        void GetK1(kState_t* buf, iSize_t& n, const kState_t k2, const kState_t k3, const kState_t k4) const override
        {
            buf[0] = m_kTable[((size_t)k4 * m_kMax + k3) * m_kMax + k2];
            n = buf[0] >= m_kMax ? 0 : 1;
        }
        void SetK1(const kState_t k1, const kState_t k2, const kState_t k3, const kState_t k4) override
        {
            m_kTable[((size_t)k4 * m_kMax + k3) * m_kMax + k2] = k1;
        }
        double GetVkkkk(const kState_t, const kState_t k2, const kState_t k3, const kState_t k4) const override
        {
            if (k2 >= m_kMax || k3 >= m_kMax || k4 >= m_kMax)
                std::cerr << "\n\tERROR in GetVkkkk: k index out of range" << std::endl;
            return m_vTable[((size_t)k4 * m_kMax + k3) * m_kMax + k2];
        }
        void SetVkkkk(const double v, const kState_t, const kState_t k2, const kState_t k3, const kState_t k4) override
        {
            m_vTable[((size_t)k4 * m_kMax + k3) * m_kMax + k2] = v;
        }
        void ToFile(const std::string fileName, const io::fileFormat_t format, utilities::MpiWrapper& mpi) const override
        {
            ToFileBase(fileName, format, 4, mpi);
        }
        void FromFile(const std::string fileName, const io::fileFormat_t format, utilities::MpiWrapper& mpi) override
        {
            FromFileBase(fileName, format, mpi);
        }

        protected:
        iSize_t CalculateDim(const kState_t kMax) const override { return (iSize_t)kMax * kMax * kMax; }
    };

    // Flux-lattice QuarticTermTables (optical_flux_lattice term_tables.cpp:95-221):
    // folded on k4 <= k3, (kMax+1) kMax^2 / 2 entries; with k4 > k3 the entry of
    // (k3, k4) is used and k1 takes the place of k2 in the value index.
    class OflQuarticTermTables : public TermTables<dcmplx>
    {
        public:
        // Generate4KTable + GenerateQuarticTerms (interacting_ofl_model.cpp:258-296,
        // 457-590) on the device from one Bloch table: bloch holds, per k, the
        // band-0 block of NonInteractingOflModel::m_blochCoefficients
        // (2 (2c+1)^2 values, spin fastest).  BuildTermTables then scales the
        // table with SetCoefficient(0.5 * interactionStrength).
        bool FromBlochCoefficients(const int dimX, const int dimY, const int cutoff, const dcmplx* bloch,
                                   const utilities::MpiWrapper& mpi)
        {
            Initialize((kState_t)(dimX * dimY), mpi);
            return utilities::detail::Ok(hd_ofl_ktable(dimX, dimY, m_kTable.data(), nullptr, nullptr),
                                         "OflQuarticTermTables") &&
                   utilities::detail::Ok(hd_ofl_vtable(dimX, dimY, cutoff, reinterpret_cast<const double*>(bloch),
                                                        reinterpret_cast<double*>(m_vTable.data()), nullptr),
                                         "OflQuarticTermTables");
        }
        void GetK1(kState_t* buf, iSize_t& n, const kState_t k2, const kState_t k3, const kState_t k4) const override
        {
            n = 1;
            buf[0] = m_kTable[Fold(k3, k4) * m_kMax + k2];
        }
        void SetK1(const kState_t k1, const kState_t k2, const kState_t k3, const kState_t k4) override
        {
            m_kTable[Fold(k3, k4) * m_kMax + k2] = k1;
        }
        dcmplx GetVkkkk(const kState_t k1, const kState_t k2, const kState_t k3, const kState_t k4) const override
        {
            return m_vTable[Fold(k3, k4) * m_kMax + (k4 > k3 ? k1 : k2)];
        }
        void SetVkkkk(const dcmplx v, const kState_t k1, const kState_t k2, const kState_t k3, const kState_t k4) override
        {
            m_vTable[Fold(k3, k4) * m_kMax + (k4 > k3 ? k1 : k2)] = v;
        }
        void ToFile(const std::string fileName, const io::fileFormat_t format, utilities::MpiWrapper& mpi) const override
        {
            ToFileBase(fileName, format, 4, mpi);
        }
        void FromFile(const std::string fileName, const io::fileFormat_t format, utilities::MpiWrapper& mpi) override
        {
            FromFileBase(fileName, format, mpi);
        }

        protected:
        iSize_t CalculateDim(const kState_t kMax) const override { return (iSize_t)(kMax + 1) * kMax * kMax / 2; }

        private:
        // row of the folded (min, max) pair: t1*kMax + t2 - t1(t1+1)/2
        size_t Fold(const kState_t k3, const kState_t k4) const
        {
            const size_t t1 = std::min(k3, k4), t2 = std::max(k3, k4);
            return t1 * m_kMax + t2 - (t1 * (t1 + 1)) / 2;
        }
    };

    // QuadraticTermTables of both models (term_tables.cpp:36-88): E_k on the
    // diagonal, GetK1(k2) gives k1 = k2, kMax entries, written with 2 labels
    template <typename T>
    class DiagonalQuadraticTermTables : public TermTables<T>
    {
        public:
        void GetK1(kState_t* buf, iSize_t& n, const kState_t k2) const override { buf[0] = k2; n = 1; }
        T GetEkk(const kState_t k1, const kState_t) const override { return this->m_vTable[k1]; }
        void SetEkk(const T e, const kState_t k1) { this->m_vTable[k1] = e; }
        void ToFile(const std::string fileName, const io::fileFormat_t format, utilities::MpiWrapper& mpi) const override
        {
            this->ToFileBase(fileName, format, 2, mpi);
        }
        void FromFile(const std::string fileName, const io::fileFormat_t format, utilities::MpiWrapper& mpi) override
        {
            this->FromFileBase(fileName, format, mpi);
        }

        protected:
        iSize_t CalculateDim(const kState_t kMax) const override { return kMax; }
    };

    // TermHashTables (src/hamiltonians/term_hash_tables_base.hpp:44-276): a
    // multimap of k1 lists keyed by (k2,k3,k4) or k2 and a map of values keyed
    // by the labels, with the keys of utilities::Key (Pack4x16: k1<<48 | k2<<32
    // | k3<<16 | k4; Pack2x32: k1<<32 | k2; serialize.hpp:96-130).  The order
    // of a k1 list does not change the Hamiltonian (each k1 of one list gives
    // a different out state), so ordered maps stand in for the reference's
    // hash maps and files are written in ascending key order.
    class TermHashTablesBase
    {
        public:
        TermHashTablesBase() : m_fileHeader("# File contains term table data in hash table format") {}
        virtual ~TermHashTablesBase() = default;

        void Initialize(const kState_t kMax) { m_kMax = kMax; }
        void Clear()
        {
            m_k1.clear();
            m_v.clear();
        }
        kState_t GetKMax() const { return m_kMax; }
        // MultiHashMultiMap::GetMaxCount (multi_key_hash.hpp:942-952)
        iSize_t GetMaxKCount() const
        {
            iSize_t m = 1;
            for (const auto& kv : m_k1) m = std::max<iSize_t>(m, (iSize_t)kv.second.size());
            return m;
        }
        // GetKTable()->Insert(k1, key) and GetVTable()->Insert(key) = value
        void InsertK1(const kState_t k1, const uint64_t key) { m_k1[key].push_back(k1); }
        dcmplx& InsertV(const uint64_t key) { return m_v[key]; }

        static uint64_t Key4(const kState_t a, const kState_t b, const kState_t c, const kState_t d)
        {
            return ((uint64_t)a << 48) | ((uint64_t)b << 32) | ((uint64_t)c << 16) | (uint64_t)d;
        }
        static uint64_t Key2(const uint32_t a, const uint32_t b) { return ((uint64_t)a << 32) | (uint64_t)b; }

        protected:
        // MultiHashMultiMap::Value (:958-983); a key the table does not hold
        // gives an empty list here (the reference leaves the count unset)
        void List(kState_t* buf, iSize_t& n, const uint64_t key) const
        {
            n = 0;
            auto it = m_k1.find(key);
            if (it == m_k1.end()) return;
            for (kState_t k : it->second) buf[n++] = k;
        }
        // MultiHashMap::Value (:783-797): 0 when absent (m_emptyReturn)
        dcmplx Value(const uint64_t key) const
        {
            auto it = m_v.find(key);
            return it == m_v.end() ? dcmplx(0.0) : it->second;
        }

        // TermHashTables::ToFileBase (term_hash_tables_base.hpp:139-175) with
        // MultiHashBase::ToFileBase (multi_key_hash.hpp:480-566): the header,
        // then the k table with nbrLabels - 1 key labels and k1 values, then the
        // value table with nbrLabels labels.  Binary: uint32 size, uint64 keys,
        // values.  Text: size, label count, then the unpacked keys and the value
        // per line, tab separated.
        void ToFileBase(const std::string& fileName, const io::fileFormat_t format, const iSize_t nbrLabels,
                        utilities::MpiWrapper& mpi) const
        {
            std::ofstream f_out;
            utilities::GenFileStream(f_out, fileName, format, mpi);
            if (mpi.m_exitFlag) return;
            if (format == io::_BINARY_) {
                const iSize_t strSize = (iSize_t)m_fileHeader.size();
                f_out.write(reinterpret_cast<const char*>(&strSize), sizeof(iSize_t));
                f_out.write(m_fileHeader.data(), strSize);
            } else {
                f_out << m_fileHeader << "\n";
            }
            std::vector<uint64_t> kkeys;
            std::vector<kState_t> kvals;
            for (const auto& kv : m_k1)
                for (kState_t k : kv.second) {
                    kkeys.push_back(kv.first);
                    kvals.push_back(k);
                }
            WriteMap(f_out, format, nbrLabels - 1, kkeys, kvals);
            std::vector<uint64_t> vkeys;
            std::vector<dcmplx> vvals;
            for (const auto& kv : m_v) {
                vkeys.push_back(kv.first);
                vvals.push_back(kv.second);
            }
            WriteMap(f_out, format, nbrLabels, vkeys, vvals);
        }

        // TermHashTables::FromFileBase (:180-219) with MultiHashBase::FromFileBase
        // (:572-669): a header mismatch or a short file sets the exit flag.
        void FromFileBase(const std::string& fileName, const io::fileFormat_t format, utilities::MpiWrapper& mpi)
        {
            std::ifstream f_in;
            utilities::GenFileStream(f_in, fileName, format, mpi);
            if (mpi.m_exitFlag) return;
            std::string header;
            if (format == io::_BINARY_) {
                iSize_t strSize = 0;
                f_in.read(reinterpret_cast<char*>(&strSize), sizeof(iSize_t));
                if (!f_in || strSize > 4096) strSize = 0;
                header.resize(strSize);
                if (strSize) f_in.read(&header[0], strSize);
            } else {
                std::getline(f_in, header);
            }
            if (header != m_fileHeader) {
                std::cerr << "\n\tFile read in error - wrong file header " << header << std::endl;
                mpi.m_exitFlag = true;
                return;
            }
            Clear();
            std::vector<uint64_t> keys;
            std::vector<kState_t> kvals;
            std::vector<dcmplx> vvals;
            bool ok = ReadMap(f_in, format, keys, kvals);
            for (size_t i = 0; ok && i < keys.size(); ++i) m_k1[keys[i]].push_back(kvals[i]);
            ok = ok && ReadMap(f_in, format, keys, vvals);
            for (size_t i = 0; ok && i < keys.size(); ++i) m_v[keys[i]] = vvals[i];
            if (!ok) {
                std::cerr << "\n\tFile read in error - " << fileName << " is truncated" << std::endl;
                mpi.m_exitFlag = true;
            }
        }

        std::map<uint64_t, std::vector<kState_t>> m_k1;
        std::map<uint64_t, dcmplx> m_v;
        kState_t m_kMax = 0;
        std::string m_fileHeader;

        private:
        static std::string Labels(const uint64_t key, const iSize_t n)
        {
            std::stringstream ss;
            if (n == 1) ss << key;
            else if (n == 2) ss << (key >> 32) << "\t" << (key & 0xFFFFFFFFull);
            else if (n == 3) ss << (key >> 48) << "\t" << ((key >> 32) & 0xFFFF) << "\t" << ((key >> 16) & 0xFFFF);
            else ss << (key >> 48) << "\t" << ((key >> 32) & 0xFFFF) << "\t" << ((key >> 16) & 0xFFFF) << "\t"
                    << (key & 0xFFFF);
            return ss.str();
        }
        template <typename V>
        static void WriteMap(std::ofstream& f, const io::fileFormat_t format, const iSize_t n,
                             const std::vector<uint64_t>& keys, const std::vector<V>& vals)
        {
            const unsigned int size = (unsigned int)keys.size();
            if (format == io::_BINARY_) {
                f.write(reinterpret_cast<const char*>(&size), sizeof(unsigned int));
                f.write(reinterpret_cast<const char*>(keys.data()), (std::streamsize)size * sizeof(uint64_t));
                f.write(reinterpret_cast<const char*>(vals.data()), (std::streamsize)size * sizeof(V));
            } else {
                f << size << "\n" << n << "\n";
                for (unsigned int i = 0; i < size; ++i) f << Labels(keys[i], n) << "\t" << utilities::ToStream(vals[i]) << "\n";
            }
        }
        template <typename V>
        static bool ReadMap(std::ifstream& f, const io::fileFormat_t format, std::vector<uint64_t>& keys,
                            std::vector<V>& vals)
        {
            unsigned int size = 0;
            if (format == io::_BINARY_) {
                f.read(reinterpret_cast<char*>(&size), sizeof(unsigned int));
                if (!f) return false;
                keys.resize(size);
                vals.resize(size);
                f.read(reinterpret_cast<char*>(keys.data()), (std::streamsize)size * sizeof(uint64_t));
                f.read(reinterpret_cast<char*>(vals.data()), (std::streamsize)size * sizeof(V));
                return (bool)f;
            }
            uint64_t n = 0;
            if (!(f >> size >> n)) return false;
            keys.resize(size);
            vals.resize(size);
            for (unsigned int i = 0; i < size; ++i) {
                uint64_t key = 0;
                if (n == 1) {
                    f >> key;
                } else if (n == 2) {
                    uint64_t a = 0, b = 0;
                    f >> a >> b;
                    key = Key2((uint32_t)a, (uint32_t)b);
                } else {
                    uint64_t a = 0, b = 0, c = 0, d = 0;
                    f >> a >> b >> c;
                    if (n == 4) f >> d;
                    key = Key4((kState_t)a, (kState_t)b, (kState_t)c, (kState_t)d);
                }
                keys[i] = key;
                utilities::FromStream(f, vals[i]);
                if (!f) return false;
            }
            return true;
        }
    };

    // QuarticTermHashTables (optical_flux_lattice term_hash_tables.cpp:96-172)
    class QuarticTermHashTables : public TermHashTablesBase
    {
        public:
        void GetK1(kState_t* buf, iSize_t& n, const kState_t k2, const kState_t k3, const kState_t k4) const
        {
            List(buf, n, Key4(k2, k3, k4, 0));
        }
        dcmplx GetVkkkk(const kState_t k1, const kState_t k2, const kState_t k3, const kState_t k4) const
        {
            return Value(Key4(k1, k2, k3, k4));
        }
        // SetQuarticFromArray (term_hash_tables_base.hpp:231-252): every
        // (k2,k3,k4) with the array's k1 and GetVkkkk
        template <class C>
        void SetQuarticFromArray(C* quarticArray)
        {
            for (kState_t k2 = 0; k2 < m_kMax; ++k2)
                for (kState_t k3 = 0; k3 < m_kMax; ++k3)
                    for (kState_t k4 = 0; k4 < m_kMax; ++k4) {
                        kState_t k1 = 0;
                        iSize_t n = 0;
                        quarticArray->GetK1(&k1, n, k2, k3, k4);
                        InsertK1(k1, Key4(k2, k3, k4, 0));
                        InsertV(Key4(k1, k2, k3, k4)) = quarticArray->GetVkkkk(k1, k2, k3, k4);
                    }
        }
        void ToFile(const std::string fileName, const io::fileFormat_t format, utilities::MpiWrapper& mpi) const
        {
            ToFileBase(fileName, format, 4, mpi);
        }
        void FromFile(const std::string fileName, const io::fileFormat_t format, utilities::MpiWrapper& mpi)
        {
            FromFileBase(fileName, format, mpi);
            kState_t km = 0;
            for (const auto& kv : m_k1) km = std::max<kState_t>(km, (kState_t)((kv.first >> 48) + 1));
            if (m_kMax == 0) m_kMax = km;
        }
    };

    // QuadraticTermHashTables (term_hash_tables.cpp:36-94)
    class QuadraticTermHashTables : public TermHashTablesBase
    {
        public:
        void GetK1(kState_t* buf, iSize_t& n, const kState_t k2) const { List(buf, n, (uint64_t)k2); }
        dcmplx GetEkk(const kState_t k1, const kState_t k2) const { return Value(Key2(k1, k2)); }
        // SetQuadraticFromArray (term_hash_tables_base.hpp:264-275)
        template <class C>
        void SetQuadraticFromArray(C* quadraticArray)
        {
            for (kState_t k2 = 0; k2 < m_kMax; ++k2) {
                kState_t k1 = 0;
                iSize_t n = 0;
                quadraticArray->GetK1(&k1, n, k2);
                InsertK1(k1, (uint64_t)k2);
                InsertV(Key2(k1, k2)) = quadraticArray->GetEkk(k1, k2);
            }
        }
        void ToFile(const std::string fileName, const io::fileFormat_t format, utilities::MpiWrapper& mpi) const
        {
            ToFileBase(fileName, format, 2, mpi);
        }
        void FromFile(const std::string fileName, const io::fileFormat_t format, utilities::MpiWrapper& mpi)
        {
            FromFileBase(fileName, format, mpi);
            kState_t km = 0;
            for (const auto& kv : m_k1) km = std::max<kState_t>(km, (kState_t)(kv.first + 1));
            if (m_kMax == 0) m_kMax = km;
        }
    };

    // NonInteractingOflModelGrid::GenerateWannierCoefficients
    // (noninteracting_ofl_model_grid.cpp:782-932): the change-of-basis matrix M
    // (K x K, row x*dimY+ky, column kx*dimY+ky') from the band-0 Bloch vectors
    // (blochDim values per k): A_x(kx,ky) = <u(kx,ky)|u(kx+1,ky)> / |.|,
    // lambda(ky) = (prod_kx A_x)^(1/dimX), aProd(kx,ky) = prod_{kappa<kx} A_x,
    // M = exp(-2 pi i (kx + offsetX) x / dimX) lambda^kx / (sqrt(dimX) aProd).
    inline void GenerateWannierCoefficients(const iSize_t dimX, const iSize_t dimY, const double offsetX,
                                            const dcmplx* bloch, const iSize_t blochDim, dcmplx* M)
    {
        const double pi = 3.1415926535897932384626433832795028841972;
        const dcmplx I(0.0, 1.0);
        std::vector<dcmplx> Ax((size_t)dimX * dimY), lambdaX(dimY, dcmplx(1.0)), aProd((size_t)dimX * dimY);
        for (iSize_t kx = 0; kx < dimX; ++kx)
            for (iSize_t ky = 0; ky < dimY; ++ky) {
                const iSize_t kn = kx == dimX - 1 ? 0 : kx + 1;
                dcmplx a(0.0);
                for (iSize_t i = 0; i < blochDim; ++i)
                    a += std::conj(bloch[((size_t)kx * dimY + ky) * blochDim + i]) * bloch[((size_t)kn * dimY + ky) * blochDim + i];
                Ax[(size_t)kx * dimY + ky] = a / std::abs(a);
            }
        for (iSize_t kx = 0; kx < dimX; ++kx)
            for (iSize_t ky = 0; ky < dimY; ++ky) lambdaX[ky] *= Ax[(size_t)kx * dimY + ky];
        for (iSize_t ky = 0; ky < dimY; ++ky) lambdaX[ky] = std::pow(lambdaX[ky], 1.0 / dcmplx(dimX, 0));
        for (iSize_t ky = 0; ky < dimY; ++ky) aProd[ky] = 1.0;
        for (iSize_t kx = 1; kx < dimX; ++kx)
            for (iSize_t ky = 0; ky < dimY; ++ky)
                aProd[(size_t)kx * dimY + ky] = Ax[(size_t)(kx - 1) * dimY + ky] * aProd[(size_t)(kx - 1) * dimY + ky];
        const iSize_t dim = dimX * dimY;
        for (iSize_t x = 0; x < dimX; ++x)
            for (iSize_t ky = 0; ky < dimY; ++ky)
                for (iSize_t kx = 0; kx < dimX; ++kx)
                    for (iSize_t ky1 = 0; ky1 < dimY; ++ky1) {
                        const size_t i1 = (size_t)x * dimY + ky, i2 = (size_t)kx * dimY + ky1;
                        M[i1 * dim + i2] = ky == ky1 ? std::exp(-2.0 * I * pi * ((double)kx + offsetX) * (double)x / (double)dimX) *
                                                           std::pow(lambdaX[ky], dcmplx(kx, 0)) /
                                                           (std::sqrt((double)dimX) * aProd[i2])
                                                     : dcmplx(0.0);
                    }
    }

    // InteractingOflModel::ChangeToWannierBasis (interacting_ofl_model.cpp:728-923)
    // through hd_ofl_wannier: the hash tables of the hybrid Wannier basis from
    // the Bloch-basis array tables (quad may be null: no quadratic terms), the
    // change-of-basis matrix and matrixElementTol.
    inline bool ChangeToWannierBasis(const int dimX, const int dimY, const double offsetX, OflQuarticTermTables& quart,
                                     DiagonalQuadraticTermTables<dcmplx>* quad, const dcmplx* basisChange,
                                     const double tol, QuarticTermHashTables& hashQuart,
                                     QuadraticTermHashTables& hashQuad)
    {
        const size_t K = (size_t)dimX * dimY, slots = K * K * K * dimX;
        std::vector<uint8_t> nk1(K * K * K), nq1(K);
        std::vector<uint16_t> k1(slots), q1(K * dimX);
        std::vector<dcmplx> v(slots), e(K * dimX);
        if (!utilities::detail::Ok(
                hd_ofl_wannier(dimX, dimY, offsetX, quart.GetKTable()->data(),
                               reinterpret_cast<const double*>(quart.GetVTable()->data()),
                               quad ? reinterpret_cast<const double*>(quad->GetVTable()->data()) : nullptr,
                               reinterpret_cast<const double*>(basisChange), tol, nk1.data(), k1.data(),
                               reinterpret_cast<double*>(v.data()), nq1.data(), q1.data(),
                               reinterpret_cast<double*>(e.data()), nullptr),
                "ChangeToWannierBasis"))
            return false;
        hashQuart.Clear();
        hashQuad.Clear();
        hashQuart.Initialize((kState_t)K);
        hashQuad.Initialize((kState_t)K);
        for (size_t ent = 0; ent < K * K * K; ++ent) {
            const kState_t k2 = (kState_t)(ent % K), k3 = (kState_t)((ent / K) % K), k4 = (kState_t)(ent / (K * K));
            for (int j = 0; j < nk1[ent]; ++j) {
                const size_t s = ent * dimX + j;
                hashQuart.InsertK1(k1[s], TermHashTablesBase::Key4(k2, k3, k4, 0));
                // the reference inserts a value only where |sum| > tol
                if (v[s] != dcmplx(0.0)) hashQuart.InsertV(TermHashTablesBase::Key4(k1[s], k2, k3, k4)) = v[s];
            }
        }
        if (quad)
            for (size_t k2 = 0; k2 < K; ++k2)
                for (int j = 0; j < nq1[k2]; ++j) {
                    const size_t s = k2 * dimX + j;
                    hashQuad.InsertK1(q1[s], (uint64_t)k2);
                    hashQuad.InsertV(TermHashTablesBase::Key2(q1[s], (uint32_t)k2)) = e[s];
                }
        return true;
    }
}

#endif
