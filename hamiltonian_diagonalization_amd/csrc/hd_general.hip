// hd_general.hip -- the Hamiltonian build for general (hash) term tables.
//
// Reference: SpinlessFermionHamiltonian<T>::Add_CdC_Terms / Add_CdCdCC_Terms
// (src/hamiltonians/spinless_fermion_hamiltonian.hpp:143-257, 281-436) reading
// QuarticTermHashTables / QuadraticTermHashTables through GetK1 (a list of k1
// per key), GetVkkkk and GetEkk
// (src/optical_flux_lattice/interacting_ofl_model/term_hash_tables.cpp:51-130),
// then MappedSparseMatrix::Trim and the map -> CRS conversion
// (mapped_sparse_matrix.hpp:355-462, crs_sparse_matrix.hpp:193-254).
//
// With general tables a (k2,k3,k4) carries a list of k1, quadratic terms hop
// (k1 != k2), and the terms that land on one column are no longer the four
// terms of one pair hop, so the pair tables of hd_build.hip do not apply.
// Every term is evaluated per row, and duplicates are summed afterwards:
//   1. k_gcount: one thread per row walks the terms in the reference's order
//      and counts the ones the reference inserts into its map (|el| > tol,
//      out >= in).
//   2. exclusive scan -> raw row offsets.
//   3. k_gfill: the same walk writes (column, value) in term order; an out
//      state outside the sector (FindFockStateIndex's miss, skipped by the
//      reference) gets a sentinel column and is dropped in 4.
//   4. k_gmerge: one workgroup per row sorts (column << 32 | position) keys in
//      LDS (bitonic; the position makes equal columns keep term order), sums
//      each run of equal columns from +0.0 in term order -- the reference's
//      map += sequence -- drops |sum| < tol (Trim, counted) and writes the
//      kept entries packed at the row's raw offset.
//   5. scan of the kept counts -> CRS row pointers; k_gcompact copies each
//      row into the final FP64 / FP32 arrays.
// Full storage is built from the upper triangle: row r holds
// conj(U[i][r]) for the i < r with U[i][r] stored, in ascending i, then the
// upper entries of row r.  That is the operator U + (U - D)^H the reference's
// product applies to its upper CRS (sparse_linear_algebra.hpp:142-165), also
// where the tables are not exactly Hermitian (the hybrid Wannier quadratic
// terms of ChangeToWannierBasis take |x2 - x1| on both sides of the diagonal).
// The lower entries are scattered per column with atomics (k_glower_scatter)
// and put back in column order per row (k_gassemble), so the result does not
// depend on the scatter order.
// Values are bit-identical to the reference's: the same terms, the same
// products (sign * value is an exact negation or copy), the same additions in
// the same order.
#include <hip/hip_runtime.h>

#include <algorithm>
#include <cmath>
#include <string>

#include "hd_internal.hpp"

namespace hd {

namespace {

template <bool CPLX> struct GT;
template <> struct GT<false> { using T = double; using F = float; };
template <> struct GT<true> { using T = double2; using F = float2; };

HD_DEV double gabs(double x) { return fabs(x); }
HD_DEV double gabs(double2 x) { return hypot(x.x, x.y); }
HD_DEV double gadd(double a, double b) { return a + b; }
HD_DEV double2 gadd(double2 a, double2 b) { return make_double2(a.x + b.x, a.y + b.y); }
HD_DEV double gneg(double a) { return -a; }
HD_DEV double2 gneg(double2 a) { return make_double2(-a.x, -a.y); }
template <typename T> HD_DEV T gzero();
template <> HD_DEV double gzero<double>() { return 0.0; }
template <> HD_DEV double2 gzero<double2>() { return make_double2(0.0, 0.0); }
HD_DEV float gf32(double v) { return (float)v; }
HD_DEV float2 gf32(double2 v) { return make_float2((float)v.x, (float)v.y); }

constexpr int32_t kNoColumn = 0x7FFFFFFF;  // out state outside the sector
constexpr int kMergeThreads = 256;

struct GenArgs {
    int K, max_k1, max_q1, has_quad;
    double tol;
    uint64_t mask;         // (1 << Norb) - 1: the reference's k2 candidate mask
    const uint8_t* nk1;    // [K^3] list lengths, entry (k4*K + k3)*K + k2
    const uint8_t* k1;     // [K^3 * max_k1]
    const void* v;         // [K^3 * max_k1] T
    const uint8_t* nq1;    // [K]
    const uint8_t* q1;     // [K * max_q1]
    const void* e;         // [K * max_q1] T
};

// The terms of one row in the reference's order, f(out, el) for each term the
// reference inserts into its map.  Quadratic terms (Add_CdC_Terms :172-228):
// k2 over occupied orbitals ascending, k1 over GetK1(k2); k1 == k2 is the
// diagonal E, otherwise k1 must be empty and c+_k1 c_k2 gives the sign
// (Evaluate_CdC, fermions.hpp:155-227).  Quartic terms (Add_CdCdCC_Terms
// :316-398): k4 over occupied orbitals, k3 over the other occupied ones, k2
// over mask ^ (occupied orbitals above k3 other than k4), k1 over
// GetK1(k2,k3,k4), the occupation filter of :351-355, the sign of
// c+_k1 c+_k2 c_k3 c_k4 (Evaluate_CdCdCC, fermions.hpp:52-144).
template <bool CPLX, bool UPPER, typename Fn>
HD_DEV void for_each_term(uint64_t s, const GenArgs& A, Fn&& f)
{
    using T = typename GT<CPLX>::T;
    if (A.has_quad) {
        const T* E = static_cast<const T*>(A.e);
        for (uint64_t t2 = s; t2; t2 &= t2 - 1) {
            const int k2 = __builtin_ctzll(t2);
            const uint64_t b2 = 1ull << k2;
            const int n = A.nq1[k2];
            for (int j = 0; j < n; ++j) {
                const int idx = k2 * A.max_q1 + j;
                const int k1 = A.q1[idx];
                if (k1 == k2) {
                    const T el = E[idx];
                    if (gabs(el) > A.tol) f(s, el);
                    continue;
                }
                const uint64_t b1 = 1ull << k1;
                if (b1 & s) continue;  // c+_k1 on an occupied orbital
                int c = __popcll(s & (b2 - 1));
                uint64_t out = s ^ b2;
                c += __popcll(out & (b1 - 1));
                out |= b1;
                if (UPPER && out < s) continue;
                T el = E[idx];
                if (c & 1) el = gneg(el);
                if (gabs(el) > A.tol) f(out, el);
            }
        }
    }
    const T* V = static_cast<const T*>(A.v);
    const int K = A.K;
    for (uint64_t t4 = s; t4; t4 &= t4 - 1) {
        const int k4 = __builtin_ctzll(t4);
        const uint64_t b4 = 1ull << k4;
        uint64_t t3 = s ^ b4;
        while (t3) {
            const uint64_t b3 = t3 & (0 - t3);
            const int k3 = __builtin_ctzll(t3);
            const uint64_t b34 = b3 | b4;
            const int ebase = (k4 * K + k3) * K;
            for (uint64_t t2 = A.mask ^ (t3 ^ b3); t2; t2 &= t2 - 1) {
                const int k2 = __builtin_ctzll(t2);
                const uint64_t b2 = 1ull << k2;
                if ((b2 & s) && !(b2 & b34)) continue;
                const int ent = ebase + k2;
                const int n = A.nk1[ent];
                for (int j = 0; j < n; ++j) {
                    const int idx = ent * A.max_k1 + j;
                    const uint64_t b1 = 1ull << A.k1[idx];
                    if ((b1 & s) && !(b1 & b34)) continue;
                    if (b1 & b2) continue;
                    int c = __popcll(s & (b4 - 1));
                    uint64_t out = s ^ b4;
                    c += __popcll(out & (b3 - 1));
                    out ^= b3;
                    c += __popcll(out & (b2 - 1));
                    out |= b2;
                    c += __popcll(out & (b1 - 1));
                    out |= b1;
                    if (UPPER && out < s) continue;
                    T el = V[idx];
                    if (c & 1) el = gneg(el);
                    if (gabs(el) > A.tol) f(out, el);
                }
            }
            t3 ^= b3;
        }
    }
}

template <bool CPLX, bool UPPER>
__global__ void __launch_bounds__(256)
k_gcount(const uint64_t* __restrict__ states, int64_t dim, GenArgs A, int32_t* __restrict__ counts)
{
    const int64_t stride = (int64_t)gridDim.x * blockDim.x;
    for (int64_t row = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; row < dim; row += stride) {
        int cnt = 0;
        for_each_term<CPLX, UPPER>(states[row], A, [&](uint64_t, const typename GT<CPLX>::T&) { ++cnt; });
        counts[row] = cnt;
    }
}

template <bool CPLX, bool UPPER, bool LZ>
__global__ void __launch_bounds__(256)
k_gfill(const uint64_t* __restrict__ states, int64_t dim, GenArgs A, const int32_t* __restrict__ rank_blob,
        int rank_words, const int64_t* __restrict__ rowptr, int32_t* __restrict__ col_out,
        typename GT<CPLX>::T* __restrict__ val_out, unsigned int* __restrict__ err)
{
    using T = typename GT<CPLX>::T;
    extern __shared__ int32_t lds[];
    block_copy_words(lds, rank_blob, rank_words);
    __syncthreads();
    const RankView rv = rank_view(lds);
    const int64_t stride = (int64_t)gridDim.x * blockDim.x;
    for (int64_t row = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; row < dim; row += stride) {
        int64_t pos = rowptr[row];
        const int64_t end = rowptr[row + 1];
        for_each_term<CPLX, UPPER>(states[row], A, [&](uint64_t out, const T& el) {
            if (pos >= end) {  // more terms than the count pass saw: never write past the row
                atomicOr(err, 4u);
                return;
            }
            uint64_t r;
            col_out[pos] = rank_state<LZ>(rv, out, r) ? (int32_t)r : kNoColumn;
            val_out[pos] = el;
            ++pos;
        });
        if (pos != end) atomicOr(err, 2u);
    }
}

// One workgroup per row: stable sort by column, run sums in term order, Trim,
// packed write of the kept entries at the row's raw offset.
template <bool CPLX, int CAP>
__global__ void __launch_bounds__(kMergeThreads)
k_gmerge(int64_t dim, const int64_t* __restrict__ rowptr, const int32_t* __restrict__ col_in,
         const typename GT<CPLX>::T* __restrict__ val_in, double tol, int32_t* __restrict__ col_out,
         typename GT<CPLX>::T* __restrict__ val_out, int32_t* __restrict__ kept,
         unsigned long long* __restrict__ ntrim, int min_len)
{
    using T = typename GT<CPLX>::T;
    __shared__ unsigned long long key[CAP];
    __shared__ int part[kMergeThreads];
    const int tid = threadIdx.x;
    unsigned long long trimmed = 0;
    for (int64_t row = blockIdx.x; row < dim; row += gridDim.x) {
        const int64_t start = rowptr[row];
        const int L = (int)(rowptr[row + 1] - start);
        if (L <= min_len) continue;  // merged by k_gmerge_wave (uniform over the block)
        int n = 1;
        while (n < L) n <<= 1;
        for (int q = tid; q < n; q += kMergeThreads)
            key[q] = q < L ? (((unsigned long long)(uint32_t)col_in[start + q] << 32) | (unsigned)q) : ~0ull;
        __syncthreads();
        for (int k = 2; k <= n; k <<= 1) {
            for (int j = k >> 1; j > 0; j >>= 1) {
                for (int i = tid; i < n; i += kMergeThreads) {
                    const int l = i ^ j;
                    if (l > i) {
                        const unsigned long long ki = key[i], kl = key[l];
                        if ((ki > kl) == ((i & k) == 0)) {
                            key[i] = kl;
                            key[l] = ki;
                        }
                    }
                }
                __syncthreads();
            }
        }
        // contiguous chunk of sorted positions per thread; the head of a run
        // of equal columns sums the whole run (which may cross chunks)
        const int per = (L + kMergeThreads - 1) / kMergeThreads;
        const int i0 = min(L, tid * per), i1 = min(L, i0 + per);
        auto run_sum = [&](int i, int32_t col) {
            T acc = gzero<T>();
            for (int q = i; q < L && (int32_t)(key[q] >> 32) == col; ++q)
                acc = gadd(acc, val_in[start + (int64_t)(key[q] & 0xFFFFFFFFull)]);
            return acc;
        };
        auto is_head = [&](int i, int32_t col) { return i == 0 || (int32_t)(key[i - 1] >> 32) != col; };
        int c = 0;
        for (int i = i0; i < i1; ++i) {
            const int32_t col = (int32_t)(key[i] >> 32);
            if (col == kNoColumn || !is_head(i, col)) continue;
            if (gabs(run_sum(i, col)) < tol) ++trimmed;
            else ++c;
        }
        part[tid] = c;
        __syncthreads();
        if (tid == 0) {
            int acc = 0;
            for (int t = 0; t < kMergeThreads; ++t) {
                const int v = part[t];
                part[t] = acc;
                acc += v;
            }
            kept[row] = acc;
        }
        __syncthreads();
        int64_t dst = start + part[tid];
        for (int i = i0; i < i1; ++i) {
            const int32_t col = (int32_t)(key[i] >> 32);
            if (col == kNoColumn || !is_head(i, col)) continue;
            const T sum = run_sum(i, col);
            if (gabs(sum) < tol) continue;
            col_out[dst] = col;
            val_out[dst] = sum;
            ++dst;
        }
        __syncthreads();
    }
    if (trimmed) atomicAdd(ntrim, trimmed);
}

// Lower part of full storage: per-column counts of the strictly upper
// entries (row r, column c > r) of the merged rows.
__global__ void __launch_bounds__(256)
k_glower_count(int64_t dim, const int64_t* __restrict__ raw_ptr, const int32_t* __restrict__ kept,
               const int32_t* __restrict__ mcol, int32_t* __restrict__ lower)
{
    for (int64_t row = blockIdx.x; row < dim; row += gridDim.x) {
        const int64_t src = raw_ptr[row];
        for (int q = threadIdx.x; q < kept[row]; q += blockDim.x) {
            const int32_t c = mcol[src + q];
            if (c > row) atomicAdd(&lower[c], 1);
        }
    }
}

template <bool CPLX>
__global__ void __launch_bounds__(256)
k_glower_scatter(int64_t dim, const int64_t* __restrict__ raw_ptr, const int32_t* __restrict__ kept,
                 const int32_t* __restrict__ mcol, const typename GT<CPLX>::T* __restrict__ mval,
                 const int64_t* __restrict__ lptr, int32_t* __restrict__ cursor, int32_t* __restrict__ tcol,
                 typename GT<CPLX>::T* __restrict__ tval)
{
    using T = typename GT<CPLX>::T;
    for (int64_t row = blockIdx.x; row < dim; row += gridDim.x) {
        const int64_t src = raw_ptr[row];
        for (int q = threadIdx.x; q < kept[row]; q += blockDim.x) {
            const int32_t c = mcol[src + q];
            if (c <= row) continue;
            const int64_t p = lptr[c] + atomicAdd(&cursor[c], 1);
            tcol[p] = (int32_t)row;
            T v = mval[src + q];
            if (CPLX) reinterpret_cast<double*>(&v)[1] = -reinterpret_cast<double*>(&v)[1];
            tval[p] = v;
        }
    }
}

// One workgroup per row of full storage: the row's lower entries sorted by
// column (keys column << 32 | slot, columns distinct), then its upper entries.
template <bool CPLX, int CAP>
__global__ void __launch_bounds__(kMergeThreads)
k_gassemble(int64_t dim, const int64_t* __restrict__ raw_ptr, const int32_t* __restrict__ kept,
            const int32_t* __restrict__ mcol, const typename GT<CPLX>::T* __restrict__ mval,
            const int64_t* __restrict__ lptr, const int32_t* __restrict__ tcol,
            const typename GT<CPLX>::T* __restrict__ tval, const int64_t* __restrict__ rowptr,
            int32_t* __restrict__ col, typename GT<CPLX>::T* __restrict__ val64, typename GT<CPLX>::F* __restrict__ val32)
{
    using T = typename GT<CPLX>::T;
    __shared__ unsigned long long key[CAP];
    const int tid = threadIdx.x;
    for (int64_t row = blockIdx.x; row < dim; row += gridDim.x) {
        const int64_t lp = lptr[row], dst = rowptr[row];
        const int L = (int)(lptr[row + 1] - lp);
        int n = 1;
        while (n < L) n <<= 1;
        for (int q = tid; q < n; q += kMergeThreads)
            key[q] = q < L ? (((unsigned long long)(uint32_t)tcol[lp + q] << 32) | (unsigned)q) : ~0ull;
        __syncthreads();
        for (int k = 2; k <= n; k <<= 1) {
            for (int j = k >> 1; j > 0; j >>= 1) {
                for (int i = tid; i < n; i += kMergeThreads) {
                    const int l = i ^ j;
                    if (l > i) {
                        const unsigned long long ki = key[i], kl = key[l];
                        if ((ki > kl) == ((i & k) == 0)) {
                            key[i] = kl;
                            key[l] = ki;
                        }
                    }
                }
                __syncthreads();
            }
        }
        for (int q = tid; q < L; q += kMergeThreads) {
            const T v = tval[lp + (int64_t)(key[q] & 0xFFFFFFFFull)];
            col[dst + q] = (int32_t)(key[q] >> 32);
            if (val64) val64[dst + q] = v;
            if (val32) val32[dst + q] = gf32(v);
        }
        const int64_t src = raw_ptr[row];
        for (int q = tid; q < kept[row]; q += kMergeThreads) {
            const T v = mval[src + q];
            col[dst + L + q] = mcol[src + q];
            if (val64) val64[dst + L + q] = v;
            if (val32) val32[dst + L + q] = gf32(v);
        }
        __syncthreads();
    }
}

__global__ void __launch_bounds__(256) k_gadd_counts(int64_t dim, const int32_t* __restrict__ a,
                                                     const int32_t* __restrict__ b, int32_t* __restrict__ out)
{
    const int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (i < dim) out[i] = a[i] + b[i];
}

template <bool CPLX, int CAP>
void launch_assemble(int grid, hipStream_t st, int64_t dim, const int64_t* raw_ptr, const int32_t* kept,
                     const int32_t* mcol, const typename GT<CPLX>::T* mval, const int64_t* lptr, const int32_t* tcol,
                     const typename GT<CPLX>::T* tval, const int64_t* rowptr, int32_t* col,
                     typename GT<CPLX>::T* val64, typename GT<CPLX>::F* val32)
{
    hipLaunchKernelGGL((k_gassemble<CPLX, CAP>), dim3(grid), dim3(kMergeThreads), 0, st, dim, raw_ptr, kept, mcol,
                       mval, lptr, tcol, tval, rowptr, col, val64, val32);
}

template <bool CPLX>
__global__ void __launch_bounds__(256)
k_gcompact(int64_t dim, const int64_t* __restrict__ raw_ptr, const int64_t* __restrict__ rowptr,
           const int32_t* __restrict__ mcol, const typename GT<CPLX>::T* __restrict__ mval,
           int32_t* __restrict__ col, typename GT<CPLX>::T* __restrict__ val64, typename GT<CPLX>::F* __restrict__ val32)
{
    for (int64_t row = blockIdx.x; row < dim; row += gridDim.x) {
        const int64_t src = raw_ptr[row], dst = rowptr[row];
        const int len = (int)(rowptr[row + 1] - dst);
        for (int q = threadIdx.x; q < len; q += blockDim.x) {
            col[dst + q] = mcol[src + q];
            const typename GT<CPLX>::T v = mval[src + q];
            if (val64) val64[dst + q] = v;
            if (val32) val32[dst + q] = gf32(v);
        }
    }
}

// k_gmerge for the rows of at most 512 raw terms, one wavefront per row (longer
// rows are skipped, for k_gmerge with min_len = 512): the keys
// (column << 9 | position) sorted in registers (sort512_regs), then each lane
// sums the runs that start among its 8 sorted keys, in term order from +0.0 as
// k_gmerge does, and the kept sums are packed at the row's raw offset through a
// wavefront prefix sum of the lanes' counts.
template <bool CPLX>
__global__ void __launch_bounds__(256)
k_gmerge_wave(int64_t dim, const int64_t* __restrict__ rowptr, const int32_t* __restrict__ col_in,
              const typename GT<CPLX>::T* __restrict__ val_in, double tol, int32_t* __restrict__ col_out,
              typename GT<CPLX>::T* __restrict__ val_out, int32_t* __restrict__ kept,
              unsigned long long* __restrict__ ntrim)
{
    using T = typename GT<CPLX>::T;
    __shared__ uint64_t s_key[4 * 512];  // one row's sorted keys per wavefront
    const int lane = threadIdx.x & 63;
    uint64_t* sk = s_key + (threadIdx.x >> 6) * 512;
    unsigned long long trimmed = 0;
    const int64_t waves = (int64_t)gridDim.x * (blockDim.x >> 6);
    for (int64_t row = (int64_t)blockIdx.x * (blockDim.x >> 6) + (threadIdx.x >> 6); row < dim; row += waves) {
        const int64_t start = rowptr[row];
        const int L = (int)(rowptr[row + 1] - start);
        if (L > 512) continue;  // left to k_gmerge (uniform over the wavefront)
        uint64_t key[8];
#pragma unroll
        for (int t = 0; t < 8; ++t) {
            const int q = lane * 8 + t;
            key[t] = q < L ? (((uint64_t)(uint32_t)col_in[start + q] << 9) | (uint64_t)q) : ~0ull;
        }
        sort512_regs(key, lane);
#pragma unroll
        for (int t = 0; t < 8; ++t) sk[lane * 8 + t] = key[t];
        __builtin_amdgcn_wave_barrier();
        T sum[8];
        unsigned keep = 0;
        int c = 0;
#pragma unroll
        for (int t = 0; t < 8; ++t) {
            sum[t] = gzero<T>();
            const int i = lane * 8 + t;
            if (i >= L) continue;
            const int32_t col = (int32_t)(uint32_t)(key[t] >> 9);
            if (col == kNoColumn) continue;
            if (i > 0 && (int32_t)(uint32_t)(sk[i - 1] >> 9) == col) continue;  // not the head of its run
            T acc = gzero<T>();
            for (int q = i; q < L && (int32_t)(uint32_t)(sk[q] >> 9) == col; ++q)
                acc = gadd(acc, val_in[start + (int64_t)(sk[q] & 511ull)]);
            if (gabs(acc) < tol) {
                ++trimmed;
            } else {
                sum[t] = acc;
                keep |= 1u << t;
                ++c;
            }
        }
        int incl = c;
#pragma unroll
        for (int d = 1; d < 64; d <<= 1) {
            const int v = __shfl_up(incl, d);
            if (lane >= d) incl += v;
        }
        if (lane == 63) kept[row] = incl;
        int64_t dst = start + (incl - c);
#pragma unroll
        for (int t = 0; t < 8; ++t) {
            if (!(keep & (1u << t))) continue;
            col_out[dst] = (int32_t)(uint32_t)(key[t] >> 9);
            val_out[dst] = sum[t];
            ++dst;
        }
        __builtin_amdgcn_wave_barrier();
    }
    if (trimmed) atomicAdd(ntrim, trimmed);
}

// Rows of at most 512 raw terms by k_gmerge_wave, the longer ones by
// k_gmerge<CAP> (CAP >= the longest row).  (A 512-thread block sorting rows of
// up to 4,096 terms in registers was no faster than k_gmerge on them: 11.3
// against 10.4 ms on the 6x5, N=8 hash tables.)
template <bool CPLX, int CAP>
void launch_merge(int grid, hipStream_t st, int64_t dim, const int64_t* raw_ptr, const int32_t* rcol,
                  const typename GT<CPLX>::T* rval, double tol, int32_t* mcol, typename GT<CPLX>::T* mval,
                  int32_t* kept, unsigned long long* ntrim)
{
    hipLaunchKernelGGL((k_gmerge_wave<CPLX>), dim3((unsigned)std::max<int64_t>(1, std::min<int64_t>((dim + 3) / 4, 65536))),
                       dim3(256), 0, st, dim, raw_ptr, rcol, rval, tol, mcol, mval, kept, ntrim);
    if (CAP > 512)
        hipLaunchKernelGGL((k_gmerge<CPLX, CAP>), dim3(grid), dim3(kMergeThreads), 0, st, dim, raw_ptr, rcol, rval,
                           tol, mcol, mval, kept, ntrim, 512);
}

template <bool CPLX>
int run_general(hd_hamiltonian* h, const hd_basis* b, const hd_tables* t, double tol, int value_mask)
{
    using T = typename GT<CPLX>::T;
    using F = typename GT<CPLX>::F;
    const hipStream_t st = h->stream;
    const int64_t dim = b->dim;
    const int K = t->K;
    int n_cu = 256;
    HD_CHECK(device_props(&n_cu, nullptr));
    uint8_t *d_nk1 = nullptr, *d_k1 = nullptr, *d_nq1 = nullptr, *d_q1 = nullptr;
    double *d_v = nullptr, *d_e = nullptr;
    int32_t *d_counts = nullptr, *d_rcol = nullptr, *d_mcol = nullptr;
    int64_t* d_raw = nullptr;
    T *d_rval = nullptr, *d_mval = nullptr;
    unsigned long long* d_misc = nullptr;  // [0] trimmed, [1] error flags
    auto cleanup = [&]() {
        dfree(d_nk1); dfree(d_k1); dfree(d_nq1); dfree(d_q1); dfree(d_v); dfree(d_e);
        dfree(d_counts); dfree(d_rcol); dfree(d_mcol); dfree(d_raw); dfree(d_rval); dfree(d_mval); dfree(d_misc);
    };
    int status = HD_OK;
#define HD_GTRY(x)             \
    do {                       \
        status = (x);          \
        if (status != HD_OK) { \
            cleanup();         \
            return status;     \
        }                      \
    } while (0)
    HipErrors q;
    HD_GTRY(dalloc_t(&d_nk1, t->nk1.size(), "k1 list lengths"));
    HD_GTRY(dalloc_t(&d_k1, t->k1.size(), "k1 lists"));
    HD_GTRY(dalloc_t(&d_v, t->hv.size(), "quartic values"));
    HD_GTRY(dalloc_t(&d_misc, 2, "counters"));
    HD_ENQ(q, hipMemcpyAsync(d_nk1, t->nk1.data(), t->nk1.size(), hipMemcpyHostToDevice, st));
    HD_ENQ(q, hipMemcpyAsync(d_k1, t->k1.data(), t->k1.size(), hipMemcpyHostToDevice, st));
    HD_ENQ(q, hipMemcpyAsync(d_v, t->hv.data(), t->hv.size() * 8, hipMemcpyHostToDevice, st));
    HD_ENQ(q, hipMemsetAsync(d_misc, 0, 2 * sizeof(unsigned long long), st));
    if (t->has_quad) {
        HD_GTRY(dalloc_t(&d_nq1, t->nq1.size(), "quadratic list lengths"));
        HD_GTRY(dalloc_t(&d_q1, t->q1.size(), "quadratic lists"));
        HD_GTRY(dalloc_t(&d_e, t->he.size(), "quadratic values"));
        HD_ENQ(q, hipMemcpyAsync(d_nq1, t->nq1.data(), t->nq1.size(), hipMemcpyHostToDevice, st));
        HD_ENQ(q, hipMemcpyAsync(d_q1, t->q1.data(), t->q1.size(), hipMemcpyHostToDevice, st));
        HD_ENQ(q, hipMemcpyAsync(d_e, t->he.data(), t->he.size() * 8, hipMemcpyHostToDevice, st));
    }
    GenArgs A;
    A.K = K;
    A.max_k1 = t->max_k1;
    A.max_q1 = t->max_q1;
    A.has_quad = t->has_quad;
    A.tol = tol;
    A.mask = b->Norb >= 64 ? ~0ull : ((1ull << b->Norb) - 1ull);
    A.nk1 = d_nk1;
    A.k1 = d_k1;
    A.v = d_v;
    A.nq1 = d_nq1;
    A.q1 = d_q1;
    A.e = d_e;
    const bool upper = h->storage == HD_STORAGE_UPPER_REF;
    const int grid_rows = (int)std::max<int64_t>(1, std::min<int64_t>((dim + 255) / 256, (int64_t)n_cu * 8));
    EventTimer ev;

    // ---- count
    HD_GTRY(dalloc_t(&d_counts, (size_t)dim, "row counts"));
    ev.start(st);
    // the terms of the upper triangle (out >= in) in both storages; full
    // storage completes it below
    hipLaunchKernelGGL((k_gcount<CPLX, true>), dim3(grid_rows), dim3(256), 0, st, b->d_states, dim, A, d_counts);
    HD_ENQ(q, hipGetLastError());
    h->t_count = ev.stop(st);
    HD_GTRY(q.status());

    // ---- scan of the raw counts
    HD_GTRY(dalloc_t(&d_raw, (size_t)dim + 1, "raw row offsets"));
    int64_t raw_nnz = 0;
    int raw_max = 0;
    ev.start(st);
    HD_GTRY(scan_counts(d_counts, dim, d_raw, &raw_nnz, &raw_max, st));
    h->t_scan = ev.stop(st);
    if (raw_max > 16384) {
        cleanup();
        return fail(HD_ERR_UNSUPPORTED, "a row has " + std::to_string(raw_max) +
                                            " terms before merging; the per-row merge takes at most 16384");
    }

    // ---- fill in term order
    HD_GTRY(dalloc_t(&d_rcol, (size_t)std::max<int64_t>(raw_nnz, 1), "raw columns"));
    HD_GTRY(dalloc_t(&d_rval, (size_t)std::max<int64_t>(raw_nnz, 1), "raw values"));
    {
        const int rank_words = (int)b->rt.blob.size();
        const size_t lds = (size_t)rank_words * 4;
        if (lds > 160 * 1024) {
            cleanup();
            return fail(HD_ERR_UNSUPPORTED, "sector rank table exceeds LDS");
        }
        const bool lz = b->rt.is_lz != 0;
        unsigned int* d_err = reinterpret_cast<unsigned int*>(d_misc + 1);
#define HD_GFILL(U, LZV)                                                                                      \
    do {                                                                                                      \
        if (lds > 64 * 1024)                                                                                  \
            HD_ENQ(q, hipFuncSetAttribute((const void*)k_gfill<CPLX, U, LZV>,                                 \
                                          hipFuncAttributeMaxDynamicSharedMemorySize, (int)lds));             \
        hipLaunchKernelGGL((k_gfill<CPLX, U, LZV>), dim3(grid_rows), dim3(256), lds, st, b->d_states, dim, A, \
                           b->d_rank, rank_words, d_raw, d_rcol, d_rval, d_err);                              \
    } while (0)
        ev.start(st);
        if (lz) HD_GFILL(true, true);
        else HD_GFILL(true, false);
#undef HD_GFILL
        HD_ENQ(q, hipGetLastError());
        h->t_fill = ev.stop(st);
        HD_GTRY(q.status());
        unsigned long long host_misc[2] = {0, 0};
        HD_ENQ(q, hipMemcpyAsync(host_misc, d_misc, sizeof(host_misc), hipMemcpyDeviceToHost, st));
        HD_ENQ(q, hipStreamSynchronize(st));
        HD_GTRY(q.status());
        if (host_misc[1] != 0) {
            cleanup();
            return fail(HD_ERR_UNSUPPORTED, "general fill pass inconsistency (flags " + std::to_string(host_misc[1]) + ")");
        }
    }
    dfree(d_nk1); dfree(d_k1); dfree(d_v); dfree(d_nq1); dfree(d_q1); dfree(d_e);
    d_nk1 = d_k1 = d_nq1 = d_q1 = nullptr;
    d_v = d_e = nullptr;

    // ---- merge duplicates per row, Trim
    HD_GTRY(dalloc_t(&d_mcol, (size_t)std::max<int64_t>(raw_nnz, 1), "merged columns"));
    HD_GTRY(dalloc_t(&d_mval, (size_t)std::max<int64_t>(raw_nnz, 1), "merged values"));
    int64_t nnz = 0;
    int max_row = 0;
    {
        const int grid = (int)std::max<int64_t>(1, std::min<int64_t>(dim, (int64_t)n_cu * 8));
        ev.start(st);
        if (raw_max <= 512) launch_merge<CPLX, 512>(grid, st, dim, d_raw, d_rcol, d_rval, tol, d_mcol, d_mval, d_counts, d_misc);
        else if (raw_max <= 1024) launch_merge<CPLX, 1024>(grid, st, dim, d_raw, d_rcol, d_rval, tol, d_mcol, d_mval, d_counts, d_misc);
        else if (raw_max <= 4096) launch_merge<CPLX, 4096>(grid, st, dim, d_raw, d_rcol, d_rval, tol, d_mcol, d_mval, d_counts, d_misc);
        else if (raw_max <= 8192) launch_merge<CPLX, 8192>(grid, st, dim, d_raw, d_rcol, d_rval, tol, d_mcol, d_mval, d_counts, d_misc);
        else launch_merge<CPLX, 16384>(grid, st, dim, d_raw, d_rcol, d_rval, tol, d_mcol, d_mval, d_counts, d_misc);
        HD_ENQ(q, hipGetLastError());
        HD_GTRY(q.status());
        dfree(d_rcol);
        dfree(d_rval);
        d_rcol = nullptr;
        d_rval = nullptr;
        HD_GTRY(dalloc_t(&h->d_rowptr, (size_t)dim + 1, "row pointers"));
        h->bytes += ((size_t)dim + 1) * 8;
        // upper storage: the kept counts are the row lengths; full storage
        // first needs the lower part of every row
        int64_t* d_uptr = nullptr;
        int32_t *d_lower = nullptr, *d_cursor = nullptr, *d_tcol = nullptr;
        int64_t* d_lptr = nullptr;
        T* d_tval = nullptr;
        auto free_full = [&]() { dfree(d_uptr); dfree(d_lower); dfree(d_cursor); dfree(d_tcol); dfree(d_lptr); dfree(d_tval); };
#define HD_FTRY(x)                 \
    do {                           \
        const int _s = (x);        \
        if (_s != HD_OK) {         \
            free_full();           \
            cleanup();             \
            return _s;             \
        }                          \
    } while (0)
        int64_t total_lower = 0;
        int max_lower = 0;
        if (upper) {
            HD_GTRY(scan_counts(d_counts, dim, h->d_rowptr, &nnz, &max_row, st));
        } else {
            int64_t nnz_u = 0;
            int max_u = 0;
            HD_FTRY(dalloc_t(&d_uptr, (size_t)dim + 1, "upper row pointers"));
            HD_FTRY(scan_counts(d_counts, dim, d_uptr, &nnz_u, &max_u, st));
            HD_FTRY(dalloc_t(&d_lower, (size_t)dim, "lower counts"));
            HD_ENQ(q, hipMemsetAsync(d_lower, 0, (size_t)dim * sizeof(int32_t), st));
            hipLaunchKernelGGL(k_glower_count, dim3(grid), dim3(256), 0, st, dim, d_raw, d_counts, d_mcol, d_lower);
            HD_ENQ(q, hipGetLastError());
            HD_FTRY(dalloc_t(&d_lptr, (size_t)dim + 1, "lower row offsets"));
            HD_FTRY(scan_counts(d_lower, dim, d_lptr, &total_lower, &max_lower, st));
            if (max_lower > 16384) {
                free_full();
                cleanup();
                return fail(HD_ERR_UNSUPPORTED, "a column holds " + std::to_string(max_lower) +
                                                    " upper entries; full storage takes at most 16384");
            }
            HD_FTRY(dalloc_t(&d_cursor, (size_t)dim, "lower cursors"));
            HD_ENQ(q, hipMemsetAsync(d_cursor, 0, (size_t)dim * sizeof(int32_t), st));
            HD_FTRY(dalloc_t(&d_tcol, (size_t)std::max<int64_t>(total_lower, 1), "lower columns"));
            HD_FTRY(dalloc_t(&d_tval, (size_t)std::max<int64_t>(total_lower, 1), "lower values"));
            hipLaunchKernelGGL((k_glower_scatter<CPLX>), dim3(grid), dim3(256), 0, st, dim, d_raw, d_counts, d_mcol,
                               d_mval, d_lptr, d_cursor, d_tcol, d_tval);
            HD_ENQ(q, hipGetLastError());
            hipLaunchKernelGGL(k_gadd_counts, dim3((unsigned)((dim + 255) / 256)), dim3(256), 0, st, dim, d_counts,
                               d_lower, d_cursor);  // the cursors now hold the full row lengths
            HD_ENQ(q, hipGetLastError());
            HD_FTRY(scan_counts(d_cursor, dim, h->d_rowptr, &nnz, &max_row, st));
            HD_FTRY(q.status());
        }
        h->nnz = nnz;
        h->max_row = max_row;
        unsigned long long trimmed = 0;
        HD_ENQ(q, hipMemcpyAsync(&trimmed, d_misc, sizeof(trimmed), hipMemcpyDeviceToHost, st));
        HD_ENQ(q, hipStreamSynchronize(st));
        HD_FTRY(q.status());
        h->n_trimmed = (int64_t)trimmed;
        // ---- final arrays, padded by 8 zeroed entries (the SpMV reads aligned
        // groups of 4 that may run past the last row)
        const size_t n = (size_t)nnz;
        HD_FTRY(dalloc_t(&h->d_col, n + 8, "columns"));
        HD_ENQ(q, hipMemsetAsync(h->d_col + n, 0, 8 * sizeof(int32_t), st));
        h->bytes += n * 4;
        if (value_mask & 1) {
            HD_FTRY(dalloc(&h->d_val64, (n + 8) * sizeof(T), "FP64 values"));
            HD_ENQ(q, hipMemsetAsync(static_cast<T*>(h->d_val64) + n, 0, 8 * sizeof(T), st));
            h->bytes += n * sizeof(T);
        }
        if (value_mask & 2) {
            HD_FTRY(dalloc(&h->d_val32, (n + 8) * sizeof(F), "FP32 values"));
            HD_ENQ(q, hipMemsetAsync(static_cast<F*>(h->d_val32) + n, 0, 8 * sizeof(F), st));
            h->bytes += n * sizeof(F);
        }
        T* v64 = static_cast<T*>(h->d_val64);
        F* v32 = static_cast<F*>(h->d_val32);
        if (upper) {
            hipLaunchKernelGGL((k_gcompact<CPLX>), dim3(grid), dim3(256), 0, st, dim, d_raw, h->d_rowptr, d_mcol,
                               d_mval, h->d_col, v64, v32);
        } else {
#define HD_ASM(C) launch_assemble<CPLX, C>(grid, st, dim, d_raw, d_counts, d_mcol, d_mval, d_lptr, d_tcol, d_tval, \
                                            h->d_rowptr, h->d_col, v64, v32)
            if (max_lower <= 256) HD_ASM(256);
            else if (max_lower <= 1024) HD_ASM(1024);
            else if (max_lower <= 4096) HD_ASM(4096);
            else if (max_lower <= 8192) HD_ASM(8192);
            else HD_ASM(16384);
#undef HD_ASM
        }
        HD_ENQ(q, hipGetLastError());
        h->t_sort = ev.stop(st);
        HD_ENQ(q, hipStreamSynchronize(st));
        HD_FTRY(q.status());
        free_full();
#undef HD_FTRY
    }
    HD_GTRY(ev.status("general build timing"));
    cleanup();
#undef HD_GTRY
    return HD_OK;
}

}  // namespace

int general_build(hd_hamiltonian* h, const hd_basis* b, const hd_tables* t, double tol, int value_mask)
{
    if (t->layout != HD_LAYOUT_HASH) return fail(HD_ERR_ARG, "general_build: not a hash table");
    if (b->dim > (int64_t)0x7FFFFFFE) return fail(HD_ERR_UNSUPPORTED, "sector too large for int32 columns");
    return t->is_complex ? run_general<true>(h, b, t, tol, value_mask) : run_general<false>(h, b, t, tol, value_mask);
}

}  // namespace hd
