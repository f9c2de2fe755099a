// hd_build.hip -- term tables and the Hamiltonian build into CRS.
//
// Reference: SpinlessFermionHamiltonian<T>::Add_CdC_Terms / Add_CdCdCC_Terms
// (src/hamiltonians/spinless_fermion_hamiltonian.hpp:143-436).  For each row
// (in-state) the reference walks k4 over occupied orbitals, k3 over the other
// occupied orbitals, k2 over the remaining candidates, takes k1 from the table,
// evaluates the Jordan-Wigner sign (fermions.hpp:52-144), binary-searches the
// out-state and accumulates sign*V into a hash map keyed (row, col), in that
// order, then trims |v| < 1e-14 and sorts into CRS.
//
// Pair structure used here.  With a table that conserves the pair (sphere
// L_z and lattice momentum both do), every term is either diagonal
// ({k1,k2} = {k3,k4}) or a two-particle hop {a,b} -> {c,d}.  A hop is reached by
// exactly four terms, and in the reference's loop order they are
//   (k4,k3,k2) = (a,b,c), (a,b,d), (b,a,c), (b,a,d)    (a<b, c<d)
// with signs s, -s, -s, s where s is the sign of the first.  So the stored
// value is s * P(a,b,c) with P = (((+V1) - V2) - V3) + V4 (terms with
// |V| <= tol skipped), bit-identical to the reference's += sequence because
// negation is exact in IEEE arithmetic.  P depends only on the orbitals and is
// tabulated once on the host; the kernels only add the sign (masked popcounts)
// and the column (combinatorial rank of the out-state).  Diagonal entries are
// state dependent and are summed per row in the reference's exact order.
// Tables that break the pair structure are rejected (HD_ERR_UNSUPPORTED).
#include <hip/hip_runtime.h>

#include <algorithm>
#include <cmath>
#include <complex>
#include <cstdlib>
#include <cstring>
#include <map>

#include "hd_internal.hpp"

namespace hd {

using cd = std::complex<double>;

// ---------------------------------------------------------- table helpers

static int get_k1(const hd_tables* t, int k2, int k3, int k4)
{
    const int K = t->K;
    if (t->layout == HD_LAYOUT_SPHERE) {
        const int v = t->ktab[((size_t)k4 * K + k3) * K + k2];
        return v < K ? v : -1;  // GetK1 returns no k1 when k1 >= kMax
    }
    int t1 = k4, t2 = k3;
    if (k4 > k3) { t1 = k3; t2 = k4; }
    const int v = t->ktab[((size_t)t1 * K + t2 - (t1 * (t1 + 1)) / 2) * K + k2];
    return v < K ? v : -1;  // a folded-table k1 >= kMax cannot be a valid orbital
}

static cd get_v(const hd_tables* t, int k1, int k2, int k3, int k4)
{
    const int K = t->K;
    size_t i;
    if (t->layout == HD_LAYOUT_SPHERE) {
        i = ((size_t)k4 * K + k3) * K + k2;
    } else {
        int t1 = k4, t2 = k3, t3 = k2;
        if (k4 > k3) { t1 = k3; t2 = k4; t3 = k1; }
        i = ((size_t)t1 * K + t2 - (t1 * (t1 + 1)) / 2) * K + t3;
    }
    if (t->is_complex) return cd(t->vtab[2 * i], t->vtab[2 * i + 1]);
    return cd(t->vtab[i], 0.0);
}

static double absv(const cd& v, bool cplx) { return cplx ? std::abs(v) : std::fabs(v.real()); }

// The array tables as one-k1 lists (ConvertTableFormat's SetQuarticFromArray /
// SetQuadraticFromArray, term_hash_tables_base.hpp:231-275): k1 = GetK1 and
// V = GetVkkkk for every (k2,k3,k4), none where the sphere table has no k1;
// k1 = k2 with E_k for the quadratic terms.
static void array_to_lists(const hd_tables* t, hd_tables& out)
{
    const int K = t->K;
    const size_t w = t->is_complex ? 2 : 1, ne = (size_t)K * K * K;
    out = hd_tables();
    out.layout = HD_LAYOUT_HASH;
    out.K = K;
    out.is_complex = t->is_complex;
    out.max_k1 = 1;
    out.nk1.assign(ne, 0);
    out.k1.assign(ne, 0);
    out.hv.assign(ne * w, 0.0);
    for (int k4 = 0; k4 < K; ++k4)
        for (int k3 = 0; k3 < K; ++k3)
            for (int k2 = 0; k2 < K; ++k2) {
                const size_t e = ((size_t)k4 * K + k3) * K + k2;
                const int k1 = get_k1(t, k2, k3, k4);
                if (k1 < 0) continue;
                const cd v = get_v(t, k1, k2, k3, k4);
                out.nk1[e] = 1;
                out.k1[e] = (uint8_t)k1;
                out.hv[e * w] = v.real();
                if (w == 2) out.hv[e * w + 1] = v.imag();
            }
    if (!t->eps.empty()) {
        out.has_quad = 1;
        out.max_q1 = 1;
        out.nq1.assign((size_t)K, 1);
        out.q1.resize((size_t)K);
        for (int k = 0; k < K; ++k) out.q1[(size_t)k] = (uint8_t)k;
        out.he = t->eps;
    }
}

// Device-side derived tables for one build.
struct Derived {
    int K = 0, npairs = 0, cplx = 0, has_eps = 0, split = 0;
    std::vector<uint8_t> kpair;  // npairs*K: d | 0x80 if trimmed; 0xFF none
    std::vector<double> pval;    // npairs*K (x2 complex)
    std::vector<double> dterm;   // K*K*2 (x2 complex): diagonal terms of ordered (k4,k3)
    std::vector<uint8_t> dflag;  // K*K: bit0 first term present, bit1 second
    std::vector<double> eps;     // K (x2 complex)
    std::vector<uint8_t> epsflag;
    // coded format: remap[pair*K + c] -> index into dict (distinct values)
    std::vector<int16_t> remap;
    std::vector<double> dict;  // (x2 complex)
    int ndict = 0;
    bool codable = true;
};

// Dictionary of the distinct kept hop values (bitwise); the code of an entry is
// its dictionary index, the sign of the hop goes into bit 15.
static void build_dictionary(Derived& D)
{
    const int w = D.cplx ? 2 : 1;
    D.remap.assign(D.kpair.size(), (int16_t)-1);
    D.dict.clear();
    std::map<std::pair<uint64_t, uint64_t>, int> seen;
    for (size_t i = 0; i < D.kpair.size(); ++i) {
        const uint8_t e = D.kpair[i];
        if (e == 0xFF || (e & 0x80)) continue;
        uint64_t re, im = 0;
        std::memcpy(&re, &D.pval[i * w], 8);
        if (D.cplx) std::memcpy(&im, &D.pval[i * w + 1], 8);
        auto it = seen.find({re, im});
        int code;
        if (it == seen.end()) {
            code = (int)seen.size();
            seen[{re, im}] = code;
            for (int q = 0; q < w; ++q) D.dict.push_back(D.pval[i * w + q]);
        } else {
            code = it->second;
        }
        if (code >= 0x7FFF) {
            D.codable = false;
            return;
        }
        D.remap[i] = (int16_t)code;
    }
    D.ndict = (int)seen.size();
}

static int derive_tables(const hd_tables* t, const hd_basis* b, double tol, Derived& D)
{
    if (t->layout == HD_LAYOUT_HASH)
        return fail(HD_ERR_UNSUPPORTED, "hash term tables are built by the general build only (not batched)");
    const int K = t->K;
    const bool cplx = t->is_complex != 0;
    D.K = K;
    D.cplx = cplx;
    D.split = t->split;
    D.npairs = K * (K - 1) / 2;
    const int w = cplx ? 2 : 1;
    D.kpair.assign((size_t)D.npairs * K, 0xFF);
    D.pval.assign((size_t)D.npairs * K * w, 0.0);
    D.dterm.assign((size_t)K * K * 2 * w, 0.0);
    D.dflag.assign((size_t)K * K, 0);
    auto put = [&](std::vector<double>& v, size_t i, const cd& x) {
        if (cplx) { v[2 * i] = x.real(); v[2 * i + 1] = x.imag(); }
        else v[i] = x.real();
    };
    auto unsupported = [&](int k1, int k2, int k3, int k4) {
        return fail(HD_ERR_UNSUPPORTED,
                    "term table does not conserve the pair structure at (k1,k2,k3,k4) = (" +
                        std::to_string(k1) + "," + std::to_string(k2) + "," +
                        std::to_string(k3) + "," + std::to_string(k4) + ")");
    };
    // which hops keep the sector: group sums of orbital pairs, computed once
    std::vector<int> sums;
    const bool by_sums = b->rt.Norb == K && sector_pair_sums(b->rt, b->sector, sums) == HD_OK;
    auto keeps_sector = [&](int a, int bb, int c, int d) {
        if (by_sums) return sums[(size_t)a * K + bb] == sums[(size_t)c * K + d];
        return sector_weight_delta_ok(b->rt, b->sector, a, bb, c, d) != 0;
    };
    for (int a = 0; a < K; ++a)
        for (int bb = a + 1; bb < K; ++bb) {
            const int pair = a * (2 * K - a - 1) / 2 + (bb - a - 1);
            for (int c = 0; c < K; ++c) {
                if (c == a || c == bb) continue;
                const int d = get_k1(t, c, bb, a);
                if (d < 0) {
                    if (get_k1(t, c, a, bb) >= 0) return unsupported(get_k1(t, c, a, bb), c, a, bb);
                    continue;
                }
                if (d == a || d == bb) return unsupported(d, c, bb, a);
                if (d == c) continue;  // c+_k c+_k vanishes (fermions.hpp:70)
                if (get_k1(t, d, bb, a) != c || get_k1(t, c, a, bb) != d || get_k1(t, d, a, bb) != c)
                    return unsupported(d, c, bb, a);
                if (c > d) continue;  // the hop {a,b}->{c,d} is tabulated at c < d
                if (!keeps_sector(a, bb, c, d)) continue;
                const cd terms[4] = {get_v(t, d, c, bb, a), -get_v(t, c, d, bb, a),
                                     -get_v(t, d, c, a, bb), get_v(t, c, d, a, bb)};
                cd acc(0.0, 0.0);
                bool any = false;
                for (const cd& x : terms)
                    if (absv(x, cplx) > tol) {
                        acc += x;
                        any = true;
                    }
                if (!any) continue;
                const size_t i = (size_t)pair * K + c;
                D.kpair[i] = (uint8_t)(d | (absv(acc, cplx) < tol ? 0x80 : 0));
                put(D.pval, i, acc);
            }
        }
    // diagonal terms for each ordered (k4, k3): k2 runs over {k3, k4} ascending
    for (int k4 = 0; k4 < K; ++k4)
        for (int k3 = 0; k3 < K; ++k3) {
            if (k3 == k4) continue;
            const int order[2] = {std::min(k3, k4), std::max(k3, k4)};
            for (int q = 0; q < 2; ++q) {
                const int k2 = order[q];
                const int k1 = get_k1(t, k2, k3, k4);
                if (k1 < 0 || k1 == k2) continue;
                if (k1 != k3 && k1 != k4) return unsupported(k1, k2, k3, k4);
                // c+_{k4} c+_{k3} c_{k3} c_{k4} = +n n ; c+_{k3} c+_{k4} c_{k3} c_{k4} = -n n
                const double sign = (k1 == k4) ? 1.0 : -1.0;
                const cd el = sign * get_v(t, k1, k2, k3, k4);
                if (absv(el, cplx) > tol) {
                    put(D.dterm, ((size_t)k4 * K + k3) * 2 + q, el);
                    D.dflag[(size_t)k4 * K + k3] |= (uint8_t)(1 << q);
                }
            }
        }
    D.has_eps = t->eps.empty() ? 0 : 1;
    D.eps.assign((size_t)K * w, 0.0);
    D.epsflag.assign(K, 0);
    if (D.has_eps)
        for (int k = 0; k < K; ++k) {
            const cd e = cplx ? cd(t->eps[2 * k], t->eps[2 * k + 1]) : cd(t->eps[k], 0.0);
            if (absv(e, cplx) > tol) {
                put(D.eps, k, e);
                D.epsflag[k] = 1;
            }
        }
    build_dictionary(D);
    return HD_OK;
}

// ---------------------------------------------------------- device kernels

template <bool CPLX> struct VT;
template <> struct VT<false> { using T = double; using F = float; };
template <> struct VT<true> { using T = double2; using F = float2; };

HD_DEV double vabs(double x) { return fabs(x); }
HD_DEV double vabs(double2 x) { return hypot(x.x, x.y); }
HD_DEV double vadd(double a, double b) { return a + b; }
HD_DEV double2 vadd(double2 a, double2 b) { return make_double2(a.x + b.x, a.y + b.y); }
HD_DEV double vneg(double a) { return -a; }
HD_DEV double2 vneg(double2 a) { return make_double2(-a.x, -a.y); }
template <typename T> HD_DEV T vzero();
template <> HD_DEV double vzero<double>() { return 0.0; }
template <> HD_DEV double2 vzero<double2>() { return make_double2(0.0, 0.0); }
HD_DEV float to_f32(double v) { return (float)v; }
HD_DEV float2 to_f32(double2 v) { return make_float2((float)v.x, (float)v.y); }

struct BuildArgs {
    int N, Norb, K, upper, has_eps, split;
    double tol;
    uint64_t norb_mask;
    const uint8_t* kpair;   // global copy (kernels stage it into LDS)
    const void* pval;
    const void* dterm;
    const uint8_t* dflag;
    const void* eps;
    const uint8_t* epsflag;
    int kpair_bytes;
    const int16_t* remap;  // coded format only (pair*K + c -> dictionary index)
};

// Parity of the sign of c+_d c+_c c_b c_a |s> (k4 = a acts first), the first of
// the four terms of the hop; fermions.hpp:87-143 restated with popcounts.
HD_DEV int hop_parity(uint64_t s, int a, int b, int c, int d)
{
    const uint64_t ba = 1ull << a, bb = 1ull << b, bc = 1ull << c, bd = 1ull << d;
    int p = __popcll(s & (ba - 1));
    const uint64_t s1 = s ^ ba;
    p += __popcll(s1 & (bb - 1));
    const uint64_t s2 = s1 ^ bb;
    p += __popcll(s2 & (bc - 1));
    const uint64_t s3 = s2 | bc;
    p += __popcll(s3 & (bd - 1));
    return p & 1;
}

// Walk the two-particle hops of state s: occupied a<b, unoccupied c (< d),
// d from the pair table.  f(a,b,c,d,table_index,out,trimmed).
template <bool UPPER, typename F>
HD_DEV void for_each_hop(uint64_t s, int K, uint64_t norb_mask, const uint8_t* kp, F&& f)
{
    const uint64_t unocc = ~s & norb_mask;
    for (uint64_t ta = s; ta; ta &= ta - 1) {
        const int a = __builtin_ctzll(ta);
        const int pbase = a * (2 * K - a - 1) / 2 - a - 1;
        for (uint64_t tb = ta & (ta - 1); tb; tb &= tb - 1) {
            const int b = __builtin_ctzll(tb);
            const int rowi = (pbase + b) * K;
            for (uint64_t tc = unocc; tc; tc &= tc - 1) {
                const int c = __builtin_ctzll(tc);
                const int e = kp[rowi + c];
                if (e == 0xFF) continue;
                const int d = e & 0x3F;
                if ((s >> d) & 1ull) continue;  // Pauli: d already occupied
                const uint64_t out = (s ^ (1ull << a) ^ (1ull << b)) | (1ull << c) | (1ull << d);
                if (UPPER && out < s) continue;
                f(a, b, c, d, rowi + c, out, (e & 0x80) != 0);
            }
        }
    }
}

template <bool CPLX>
__global__ void __launch_bounds__(256)
k_count(const uint64_t* __restrict__ states, int64_t dim, BuildArgs A, int32_t* __restrict__ counts,
        typename VT<CPLX>::T* __restrict__ diag, uint8_t* __restrict__ dkeep,
        unsigned long long* __restrict__ ntrim)
{
    using T = typename VT<CPLX>::T;
    extern __shared__ int32_t lds[];
    const int K = A.K;
    // stage kpair, dterm, dflag, eps, epsflag in LDS
    uint8_t* s_kp = reinterpret_cast<uint8_t*>(lds);
    const int kpw = (A.kpair_bytes + 3) / 4;
    T* s_dt = reinterpret_cast<T*>(lds + ((kpw + 3) & ~3));
    uint8_t* s_df = reinterpret_cast<uint8_t*>(s_dt + 2 * K * K);
    T* s_eps = reinterpret_cast<T*>(s_df + ((K * K + 15) & ~15));
    uint8_t* s_ef = reinterpret_cast<uint8_t*>(s_eps + K);
    block_copy_words(lds, reinterpret_cast<const int32_t*>(A.kpair), kpw);
    for (int i = threadIdx.x; i < 2 * K * K; i += blockDim.x)
        s_dt[i] = reinterpret_cast<const T*>(A.dterm)[i];
    for (int i = threadIdx.x; i < K * K; i += blockDim.x) s_df[i] = A.dflag[i];
    for (int i = threadIdx.x; i < K; i += blockDim.x) {
        s_eps[i] = reinterpret_cast<const T*>(A.eps)[i];
        s_ef[i] = A.epsflag[i];
    }
    __syncthreads();
    unsigned long long trimmed = 0;
    const int64_t stride = (int64_t)gridDim.x * blockDim.x;
    for (int64_t row = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; row < dim; row += stride) {
        const uint64_t s = states[row];
        // diagonal: quadratic terms first (Add_CdC_Terms runs before
        // Add_CdCdCC_Terms), then the quartic terms in (k4, k3, k2) order.
        // Two levels (two_level_pseudopotential_model.cpp:188-200): E terms of
        // the first level (orbitals >= split), of the second, then the first
        // level's quartic terms, the Trim that ends that call (|v| < tol
        // restarts the sum at 0), then the second level's quartic terms.
        T dsum = vzero<T>();
        bool any = false;
        const uint64_t lo_mask = A.split ? (1ull << A.split) - 1ull : ~0ull;
        const uint64_t part[2] = {A.split ? (s & ~lo_mask) : s, A.split ? (s & lo_mask) : 0ull};
        if (A.has_eps)
            for (int q = 0; q < 2; ++q)
                for (uint64_t t = part[q]; t; t &= t - 1) {
                    const int k = __builtin_ctzll(t);
                    if (s_ef[k]) {
                        dsum = vadd(dsum, s_eps[k]);
                        any = true;
                    }
                }
        for (int q = 0; q < 2; ++q) {
            const uint64_t sq = part[q];
            for (uint64_t t4 = sq; t4; t4 &= t4 - 1) {
                const int k4 = __builtin_ctzll(t4);
                for (uint64_t t3 = sq ^ (1ull << k4); t3; t3 &= t3 - 1) {
                    const int k3 = __builtin_ctzll(t3);
                    const int idx = k4 * K + k3;
                    const int f = s_df[idx];
                    if (f & 1) dsum = vadd(dsum, s_dt[2 * idx]);
                    if (f & 2) dsum = vadd(dsum, s_dt[2 * idx + 1]);
                    any |= (f != 0);
                }
            }
            if (q == 0 && A.split && any && vabs(dsum) < A.tol) dsum = vzero<T>();
        }
        const bool keep = any && !(vabs(dsum) < A.tol);
        if (any && !keep) ++trimmed;
        int cnt = keep ? 1 : 0;
        auto body = [&](int, int, int, int, int, uint64_t, bool tr) {
            if (tr) ++trimmed;
            else ++cnt;
        };
        if (A.upper) for_each_hop<true>(s, K, A.norb_mask, s_kp, body);
        else for_each_hop<false>(s, K, A.norb_mask, s_kp, body);
        counts[row] = cnt;
        diag[row] = dsum;
        dkeep[row] = keep ? 1 : 0;
    }
    if (trimmed) atomicAdd(ntrim, trimmed);
}

template <bool CPLX, bool UPPER, bool LZ>
__global__ void __launch_bounds__(1024)
k_fill(const uint64_t* __restrict__ states, int64_t dim, BuildArgs A,
       const int32_t* __restrict__ rank_blob, int rank_words, const int64_t* __restrict__ rowptr,
       const typename VT<CPLX>::T* __restrict__ diag, const uint8_t* __restrict__ dkeep,
       int32_t* __restrict__ col_out, typename VT<CPLX>::T* __restrict__ val_out,
       uint16_t* __restrict__ code_out, unsigned int* __restrict__ err)
{
    using T = typename VT<CPLX>::T;
    extern __shared__ int32_t lds[];
    block_copy_words(lds, rank_blob, rank_words);
    uint8_t* s_kp = reinterpret_cast<uint8_t*>(lds + rank_words);
    block_copy_words(lds + rank_words, reinterpret_cast<const int32_t*>(A.kpair),
                     (A.kpair_bytes + 3) / 4);
    __syncthreads();
    const RankView rv = rank_view(lds);
    const T* P = reinterpret_cast<const T*>(A.pval);
    const int64_t stride = (int64_t)gridDim.x * blockDim.x;
    for (int64_t row = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; row < dim; row += stride) {
        const uint64_t s = states[row];
        int64_t pos = rowptr[row];
        const int64_t end = rowptr[row + 1];
        if (dkeep[row] && pos < end) {
            col_out[pos] = (int32_t)row;
            val_out[pos] = diag[row];
            if (code_out) code_out[pos] = 0xFFFF;
            ++pos;
        }
        auto body = [&](int a, int b, int c, int d, int ti, uint64_t out, bool tr) {
            if (tr) return;
            if (pos >= end) {  // more entries than the count pass saw: never write past the row
                atomicOr(err, 4u);
                return;
            }
            uint64_t r;
            if (!rank_state<LZ>(rv, out, r)) {
                atomicOr(err, 1u);
                r = 0;
            }
            T v = P[ti];
            const int par = hop_parity(s, a, b, c, d);
            if (par) v = vneg(v);
            col_out[pos] = (int32_t)r;
            val_out[pos] = v;
            if (code_out) code_out[pos] = (uint16_t)((uint16_t)A.remap[ti] | (par ? 0x8000u : 0u));
            ++pos;
        };
        for_each_hop<UPPER>(s, A.K, A.norb_mask, s_kp, body);
        if (pos != end) atomicOr(err, 2u);
    }
}

// The same rows as k_fill, one wavefront per row.  A row's candidates (occupied
// a < b, unoccupied c) are walked 64 at a time; those that are hops (pair table
// entry present, d unoccupied, not trimmed, out >= s for upper storage) are
// packed as (a, b, c, d) bytes into a per-wave queue in LDS, and every 64
// queued hops are ranked by the 64 lanes together and stored to 64 consecutive
// entries of the row, so no lane idles through the ranking and the stores are
// contiguous.  Entries come out in the order k_fill writes them.
// LDS: rank table, pair table, then per block the list of (ia, ib) occupied-index
// pairs (2 bytes each) and per wave a 128-hop queue and the row's occupied and
// unoccupied orbitals (64 bytes).
// Split ranking.  The rank of a sector state is a sum over its occupied
// orbitals p, taken in descending order, of W(p, c, g): the number of ways to
// place the c particles left (p's included) below p with the group weight g left
// (the table of make_rank_table).  For a state of the sector, c and g at p are
// fixed by the orbitals above p, and also by those at or below p (the count and
// weight that remain).  So the orbitals below m1 contribute L(low bits), those
// from m2 up contribute H(high bits), both tabulated over all bit patterns, and
// only the occupied orbitals of the middle [m1, m2) are stepped, from the c and g
// that H leaves.  Valid for states of the sector only: the fill ranks hops of
// the pair table, which keep the sector (derive_tables drops the others).
template <bool LZ>
HD_DEV uint32_t split_rank(const int32_t* sb, const SplitArgs& S, uint64_t s)
{
    const uint32_t* L = reinterpret_cast<const uint32_t*>(sb);
    const uint2* H = reinterpret_cast<const uint2*>(sb + S.oH);
    const uint2* pk = reinterpret_cast<const uint2*>(sb + S.oPk);
    const int32_t* sub = sb + S.oSub;
    const uint32_t* tab = reinterpret_cast<const uint32_t*>(sb + S.oTab);
    const int nmid = S.m2 - S.m1;
    uint32_t r = L[s & ((1ull << S.m1) - 1ull)];
    const uint2 he = H[S.m2 < 64 ? (s >> S.m2) : 0ull];
    r += he.x;
    int c = (int)(he.y & 0xFFu), g = (int)(he.y >> 8);
    uint64_t t = nmid > 0 ? (s >> S.m1) & ((1ull << nmid) - 1ull) : 0ull;
    while (t) {
        const int q = 63 - __clzll(t);
        t ^= 1ull << q;
        const uint2 e = pk[q * (S.N + 1) + c];
        const int lo = (int)(int16_t)(e.y & 0xFFFFu), hi = (int)(int16_t)(e.y >> 16);
        if (g >= lo && g <= hi) r += tab[(int)e.x + g];
        g = LZ ? g - sub[q] : sub[g * nmid + q];
        c = c > 0 ? c - 1 : 0;
    }
    return r;
}

constexpr int kFillWaves = 16;
static size_t fill_wave_lds(int rank_words, int kp_bytes, int N)
{
    const size_t kpw = (size_t)(kp_bytes + 3) / 4;
    const size_t npair = (size_t)N * (N - 1) / 2;
    return 4 * (size_t)rank_words + 4 * kpw + 4 * ((npair * 2 + 3) / 4) +
           (size_t)kFillWaves * (128 * 4 + 64);
}

// SPLIT: rank_blob holds the split tables (S) instead of the rank table.
template <bool CPLX, bool UPPER, bool LZ, bool SPLIT>
__global__ void __launch_bounds__(64 * kFillWaves)
k_fill_wave(const uint64_t* __restrict__ states, int64_t dim, BuildArgs A, SplitArgs S,
            const int32_t* __restrict__ rank_blob, int rank_words, const int64_t* __restrict__ rowptr,
            const typename VT<CPLX>::T* __restrict__ diag, const uint8_t* __restrict__ dkeep,
            int32_t* __restrict__ col_out, typename VT<CPLX>::T* __restrict__ val_out,
            uint16_t* __restrict__ code_out, unsigned int* __restrict__ err)
{
    using T = typename VT<CPLX>::T;
    extern __shared__ int32_t lds[];
    const int K = A.K, N = A.N;
    const int kpw = (A.kpair_bytes + 3) / 4;
    const int npair = N * (N - 1) / 2;
    block_copy_words(lds, rank_blob, rank_words);
    const uint8_t* s_kp = reinterpret_cast<const uint8_t*>(lds + rank_words);
    block_copy_words(lds + rank_words, reinterpret_cast<const int32_t*>(A.kpair), kpw);
    uint8_t* s_pair = reinterpret_cast<uint8_t*>(lds + rank_words + kpw);
    for (int P = threadIdx.x; P < npair; P += blockDim.x) {
        int ia = 0, rem = P;
        while (rem >= N - 1 - ia) {
            rem -= N - 1 - ia;
            ++ia;
        }
        s_pair[2 * P] = (uint8_t)ia;
        s_pair[2 * P + 1] = (uint8_t)(ia + 1 + rem);
    }
    const int wave = threadIdx.x >> 6, lane = threadIdx.x & 63;
    uint32_t* queue = reinterpret_cast<uint32_t*>(lds + rank_words + kpw + (npair * 2 + 3) / 4) + wave * (128 + 16);
    uint8_t* s_orb = reinterpret_cast<uint8_t*>(queue + 128);  // [0, N) occupied, [N, K) unoccupied
    __syncthreads();
    const RankView rv = rank_view(lds);
    const T* P = reinterpret_cast<const T*>(A.pval);
    const int nun = K - N;
    const int ncand = npair * nun;
    const uint32_t magic = nun > 0 ? (uint32_t)((0x100000000ull + nun - 1) / (uint64_t)nun) : 0u;
    const uint64_t below = (1ull << lane) - 1ull;
    const int64_t nw = (int64_t)gridDim.x * (blockDim.x >> 6);
    for (int64_t row = (int64_t)blockIdx.x * (blockDim.x >> 6) + wave; row < dim; row += nw) {
        const uint64_t s = states[row];
        int64_t pos = rowptr[row];
        const int64_t end = rowptr[row + 1];
        if (dkeep[row] && pos < end) {
            if (lane == 0) {
                col_out[pos] = (int32_t)row;
                val_out[pos] = diag[row];
                if (code_out) code_out[pos] = 0xFFFF;
            }
            ++pos;
        }
        if (lane < K) {
            const int nb = __popcll(s & below);
            if ((s >> lane) & 1ull) s_orb[nb] = (uint8_t)lane;
            else s_orb[N + lane - nb] = (uint8_t)lane;
        }
        __builtin_amdgcn_wave_barrier();
        // rank and store the first n queued hops (n <= 64), one per lane
        auto drain = [&](int n) {
            if (lane < n) {
                const uint32_t h = queue[lane];
                const int a = h & 0xFF, b = (h >> 8) & 0xFF, c = (h >> 16) & 0xFF, d = h >> 24;
                const uint64_t out = (s ^ (1ull << a) ^ (1ull << b)) | (1ull << c) | (1ull << d);
                const int ti = ((a * (2 * K - a - 1) / 2 - a - 1) + b) * K + c;
                const int64_t q = pos + lane;
                if (q < end) {
                    uint64_t r;
                    if (SPLIT) {
                        r = split_rank<LZ>(lds, S, out);
                        if (r >= (uint64_t)dim) {  // never a column outside the sector
                            atomicOr(err, 1u);
                            r = 0;
                        }
                    } else if (!rank_state<LZ>(rv, out, r)) {
                        atomicOr(err, 1u);
                        r = 0;
                    }
                    T v = P[ti];
                    const int par = hop_parity(s, a, b, c, d);
                    if (par) v = vneg(v);
                    col_out[q] = (int32_t)r;
                    val_out[q] = v;
                    if (code_out) code_out[q] = (uint16_t)((uint16_t)A.remap[ti] | (par ? 0x8000u : 0u));
                } else {  // more entries than the count pass saw: never write past the row
                    atomicOr(err, 4u);
                }
            }
            pos += n;
        };
        int nq = 0;
        for (int base = 0; base < ncand; base += 64) {
            const int qi = base + lane;
            bool ok = false;
            uint32_t h = 0;
            if (qi < ncand) {
                const int pi = (int)__umulhi((uint32_t)qi, magic);
                const int ci = qi - pi * nun;
                const int a = s_orb[s_pair[2 * pi]], b = s_orb[s_pair[2 * pi + 1]], c = s_orb[N + ci];
                const int e = s_kp[((a * (2 * K - a - 1) / 2 - a - 1) + b) * K + c];
                if (e != 0xFF && !(e & 0x80)) {
                    const int d = e & 0x3F;
                    if (!((s >> d) & 1ull)) {
                        const uint64_t out = (s ^ (1ull << a) ^ (1ull << b)) | (1ull << c) | (1ull << d);
                        ok = !(UPPER && out < s);
                        h = (uint32_t)a | ((uint32_t)b << 8) | ((uint32_t)c << 16) | ((uint32_t)d << 24);
                    }
                }
            }
            const uint64_t m = __ballot(ok);
            if (ok) queue[nq + __popcll(m & below)] = h;
            nq += __popcll(m);
            __builtin_amdgcn_wave_barrier();
            if (nq >= 64) {
                drain(64);
                nq -= 64;
                const uint32_t carry = lane < nq ? queue[64 + lane] : 0u;
                __builtin_amdgcn_wave_barrier();
                if (lane < nq) queue[lane] = carry;
                __builtin_amdgcn_wave_barrier();
            }
        }
        if (nq > 0) drain(nq);
        __builtin_amdgcn_wave_barrier();
        if (lane == 0 && pos != end) atomicOr(err, 2u);
    }
}

HD_DEV double vmad(double acc, double v, double x) { return acc + v * x; }
HD_DEV double2 vmad(double2 acc, double2 v, double2 x)
{
    return make_double2(acc.x + (v.x * x.x - v.y * x.y), acc.y + (v.x * x.y + v.y * x.x));
}

// Matrix-free product: y[row] = diag[row] x[row] + sum over the row's hops (the
// full-storage rows k_fill writes, trimmed hops skipped) of sign * P * x[rank(out)].
// One thread per row; the rank table and the pair table sit in LDS.
template <bool CPLX, bool LZ>
__global__ void __launch_bounds__(256)
k_spmv_free(const uint64_t* __restrict__ states, int64_t dim, int K, uint64_t norb_mask,
            const uint8_t* __restrict__ kpair, int kp_bytes, const int32_t* __restrict__ rank_blob, int rank_words,
            const typename VT<CPLX>::T* __restrict__ P, const typename VT<CPLX>::T* __restrict__ diag,
            const typename VT<CPLX>::T* __restrict__ x, typename VT<CPLX>::T* __restrict__ y)
{
    using T = typename VT<CPLX>::T;
    extern __shared__ int32_t lds[];
    block_copy_words(lds, rank_blob, rank_words);
    uint8_t* s_kp = reinterpret_cast<uint8_t*>(lds + rank_words);
    block_copy_words(lds + rank_words, reinterpret_cast<const int32_t*>(kpair), (kp_bytes + 3) / 4);
    __syncthreads();
    const RankView rv = rank_view(lds);
    const int64_t stride = (int64_t)gridDim.x * blockDim.x;
    for (int64_t row = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; row < dim; row += stride) {
        const uint64_t s = states[row];
        T acc = vmad(vzero<T>(), diag[row], x[row]);
        for_each_hop<false>(s, K, norb_mask, s_kp, [&](int a, int b, int c, int d, int ti, uint64_t out, bool tr) {
            if (tr) return;
            uint64_t r;
            if (!rank_state<LZ>(rv, out, r)) return;  // hops stay in the sector (sector_weight_delta_ok)
            T v = P[ti];
            if (hop_parity(s, a, b, c, d)) v = vneg(v);
            acc = vmad(acc, v, x[r]);
        });
        y[row] = acc;
    }
}

HD_DEV double shfl_xor_v(double v, int d) { return __shfl_xor(v, d); }
HD_DEV double2 shfl_xor_v(double2 v, int d) { return make_double2(__shfl_xor(v.x, d), __shfl_xor(v.y, d)); }

// The matrix-free product with k_fill_wave's row walk: one wavefront per row,
// candidates 64 at a time through a per-wave queue of hops, every 64 hops ranked
// by the split tables (rank_blob) and gathered by the 64 lanes together, then a
// wavefront sum.  y[row] = diag[row] x[row] + the row's full-storage hops; the
// order of the sum differs from k_spmv_free's (and from the column-sorted CRS)
// by rounding only.  LDS as in k_fill_wave.
template <bool CPLX, bool LZ>
__global__ void __launch_bounds__(64 * kFillWaves)
k_spmv_free_wave(const uint64_t* __restrict__ states, int64_t dim, int N, int K, SplitArgs S,
                 const int32_t* __restrict__ split_blob, const uint8_t* __restrict__ kpair, int kp_bytes,
                 const typename VT<CPLX>::T* __restrict__ pval, const typename VT<CPLX>::T* __restrict__ diag,
                 const typename VT<CPLX>::T* __restrict__ x, typename VT<CPLX>::T* __restrict__ y)
{
    using T = typename VT<CPLX>::T;
    extern __shared__ int32_t lds[];
    const int kpw = (kp_bytes + 3) / 4;
    const int npair = N * (N - 1) / 2;
    block_copy_words(lds, split_blob, S.words);
    const uint8_t* s_kp = reinterpret_cast<const uint8_t*>(lds + S.words);
    block_copy_words(lds + S.words, reinterpret_cast<const int32_t*>(kpair), kpw);
    uint8_t* s_pair = reinterpret_cast<uint8_t*>(lds + S.words + kpw);
    for (int P = threadIdx.x; P < npair; P += blockDim.x) {
        int ia = 0, rem = P;
        while (rem >= N - 1 - ia) {
            rem -= N - 1 - ia;
            ++ia;
        }
        s_pair[2 * P] = (uint8_t)ia;
        s_pair[2 * P + 1] = (uint8_t)(ia + 1 + rem);
    }
    const int wave = threadIdx.x >> 6, lane = threadIdx.x & 63;
    uint32_t* queue = reinterpret_cast<uint32_t*>(lds + S.words + kpw + (npair * 2 + 3) / 4) + wave * (128 + 16);
    uint8_t* s_orb = reinterpret_cast<uint8_t*>(queue + 128);
    __syncthreads();
    const int nun = K - N;
    const int ncand = npair * nun;
    const uint32_t magic = nun > 0 ? (uint32_t)((0x100000000ull + nun - 1) / (uint64_t)nun) : 0u;
    const uint64_t below = (1ull << lane) - 1ull;
    const int64_t nw = (int64_t)gridDim.x * (blockDim.x >> 6);
    for (int64_t row = (int64_t)blockIdx.x * (blockDim.x >> 6) + wave; row < dim; row += nw) {
        const uint64_t s = states[row];
        T acc = lane == 0 ? vmad(vzero<T>(), diag[row], x[row]) : vzero<T>();
        if (lane < K) {
            const int nb = __popcll(s & below);
            if ((s >> lane) & 1ull) s_orb[nb] = (uint8_t)lane;
            else s_orb[N + lane - nb] = (uint8_t)lane;
        }
        __builtin_amdgcn_wave_barrier();
        auto drain = [&](int n) {
            if (lane < n) {
                const uint32_t h = queue[lane];
                const int a = h & 0xFF, b = (h >> 8) & 0xFF, c = (h >> 16) & 0xFF, d = h >> 24;
                const uint64_t out = (s ^ (1ull << a) ^ (1ull << b)) | (1ull << c) | (1ull << d);
                const int ti = ((a * (2 * K - a - 1) / 2 - a - 1) + b) * K + c;
                const uint32_t r = split_rank<LZ>(lds, S, out);
                if ((int64_t)r < dim) {  // hops keep the sector
                    T v = pval[ti];
                    if (hop_parity(s, a, b, c, d)) v = vneg(v);
                    acc = vmad(acc, v, x[r]);
                }
            }
        };
        int nq = 0;
        for (int base = 0; base < ncand; base += 64) {
            const int qi = base + lane;
            bool ok = false;
            uint32_t h = 0;
            if (qi < ncand) {
                const int pi = (int)__umulhi((uint32_t)qi, magic);
                const int ci = qi - pi * nun;
                const int a = s_orb[s_pair[2 * pi]], b = s_orb[s_pair[2 * pi + 1]], c = s_orb[N + ci];
                const int e = s_kp[((a * (2 * K - a - 1) / 2 - a - 1) + b) * K + c];
                if (e != 0xFF && !(e & 0x80)) {
                    const int d = e & 0x3F;
                    ok = !((s >> d) & 1ull);
                    h = (uint32_t)a | ((uint32_t)b << 8) | ((uint32_t)c << 16) | ((uint32_t)d << 24);
                }
            }
            const uint64_t m = __ballot(ok);
            if (ok) queue[nq + __popcll(m & below)] = h;
            nq += __popcll(m);
            __builtin_amdgcn_wave_barrier();
            if (nq >= 64) {
                drain(64);
                nq -= 64;
                const uint32_t carry = lane < nq ? queue[64 + lane] : 0u;
                __builtin_amdgcn_wave_barrier();
                if (lane < nq) queue[lane] = carry;
                __builtin_amdgcn_wave_barrier();
            }
        }
        if (nq > 0) drain(nq);
        __builtin_amdgcn_wave_barrier();
#pragma unroll
        for (int d = 32; d >= 1; d >>= 1) acc = vadd(acc, shfl_xor_v(acc, d));
        if (lane == 0) y[row] = acc;
    }
}

template <typename T>
__global__ void __launch_bounds__(256) k_fold_diag(int64_t dim, T* __restrict__ diag, const uint8_t* __restrict__ keep)
{
    const int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (i < dim && !keep[i]) diag[i] = vzero<T>();
}

// Per-row sort by column: one wavefront per row, bitonic network in LDS over
// CAP slots, then the permuted values are gathered from the unsorted copy and
// written in the final precisions.
template <bool CPLX, int CAP>
__global__ void __launch_bounds__(64)
k_sort(int64_t dim, const int64_t* __restrict__ rowptr, const int32_t* __restrict__ col_in,
       const typename VT<CPLX>::T* __restrict__ val_in, const uint16_t* __restrict__ code_in,
       int32_t* __restrict__ col_out, typename VT<CPLX>::T* __restrict__ val64,
       typename VT<CPLX>::F* __restrict__ val32, uint16_t* __restrict__ code_out)
{
    __shared__ int32_t key[CAP];
    __shared__ int16_t idx[CAP];
    const int lane = threadIdx.x;
    for (int64_t row = blockIdx.x; row < dim; row += gridDim.x) {
        const int64_t start = rowptr[row];
        const int L = (int)(rowptr[row + 1] - start);
        int n = 1;
        while (n < L) n <<= 1;
        for (int q = lane; q < n; q += 64) {
            key[q] = q < L ? col_in[start + q] : 0x7FFFFFFF;
            idx[q] = (int16_t)q;
        }
        __syncthreads();
        for (int k = 2; k <= n; k <<= 1) {
            for (int j = k >> 1; j > 0; j >>= 1) {
                for (int i = lane; i < n; i += 64) {
                    const int l = i ^ j;
                    if (l > i) {
                        const int ki = key[i], kl = key[l];
                        const bool up = (i & k) == 0;
                        if ((ki > kl) == up) {
                            key[i] = kl;
                            key[l] = ki;
                            const int16_t t = idx[i];
                            idx[i] = idx[l];
                            idx[l] = t;
                        }
                    }
                }
                __syncthreads();
            }
        }
        for (int q = lane; q < L; q += 64) {
            col_out[start + q] = key[q];
            const typename VT<CPLX>::T v = val_in[start + idx[q]];
            if (val64) val64[start + q] = v;
            if (val32) val32[start + q] = to_f32(v);
            if (code_out) code_out[start + q] = code_in[start + idx[q]];
        }
        __syncthreads();
    }
}

// The same per-row sort for rows of at most 512 entries with the keys in
// registers: one wavefront per row, lane l holding elements 8l..8l+7 of a
// 512-slot bitonic network; distances of 8 and more swap between lanes by
// shuffles, shorter ones inside a lane.  A key is (column << S | position): 32
// bits (S = 9) when every column is below 2^23, else 64 bits (S = 16).  The
// columns of a row are distinct, so the result equals k_sort's.

template <bool CPLX, typename Key>
__global__ void __launch_bounds__(256)
k_sort512(int64_t dim, const int64_t* __restrict__ rowptr, const int32_t* __restrict__ col_in,
          const typename VT<CPLX>::T* __restrict__ val_in, const uint16_t* __restrict__ code_in,
          int32_t* __restrict__ col_out, typename VT<CPLX>::T* __restrict__ val64,
          typename VT<CPLX>::F* __restrict__ val32, uint16_t* __restrict__ code_out)
{
    constexpr int S = sizeof(Key) == 4 ? 9 : 16;
    constexpr Key kPosMask = (Key(1) << S) - 1;
    __shared__ Key s_key[4 * 512];  // one 512-key row per wavefront of the block
    const int lane = threadIdx.x & 63;
    const int64_t waves = (int64_t)gridDim.x * (blockDim.x >> 6);
    // the next row's keys are loaded while the current row is sorted
    auto load = [&](int64_t r, int64_t& start, int& L, Key (&key)[8]) {
        start = rowptr[r];
        L = (int)(rowptr[r + 1] - start);
#pragma unroll
        for (int t = 0; t < 8; ++t) {
            const int q = lane * 8 + t;
            key[t] = q < L ? (Key)(((Key)(uint32_t)col_in[start + q] << S) | (Key)q) : ~Key(0);
        }
    };
    int64_t row = (int64_t)blockIdx.x * (blockDim.x >> 6) + (threadIdx.x >> 6);
    int64_t nstart = 0;
    int nL = 0;
    Key nkey[8];
    if (row < dim) load(row, nstart, nL, nkey);
    for (; row < dim; row += waves) {
        const int64_t start = nstart;
        const int L = nL;
        Key key[8];
#pragma unroll
        for (int t = 0; t < 8; ++t) key[t] = nkey[t];
        if (row + waves < dim) load(row + waves, nstart, nL, nkey);
        sort512_regs(key, lane);
        // through LDS to the striped order q = 64 t + lane, so that the stores of
        // a wavefront are contiguous
        Key* sk = s_key + (threadIdx.x >> 6) * 512;
#pragma unroll
        for (int t = 0; t < 8; ++t) sk[lane * 8 + t] = key[t];
        __builtin_amdgcn_wave_barrier();
#pragma unroll
        for (int t = 0; t < 8; ++t) {
            const int q = t * 64 + lane;
            if (q < L) {
                const Key kq = sk[q];
                const int idx = (int)(kq & kPosMask);
                col_out[start + q] = (int32_t)(kq >> S);
                const typename VT<CPLX>::T v = val_in[start + idx];
                if (val64) val64[start + q] = v;
                if (val32) val32[start + q] = to_f32(v);
                if (code_out) code_out[start + q] = code_in[start + idx];
            }
        }
        __builtin_amdgcn_wave_barrier();
    }
}

// The fill and the row sort in one pass, for rows of at most 512 entries and
// columns below 2^23: one wavefront per row walks the candidates as k_fill_wave
// does, but keeps all of the row's hops in LDS, ranks them 64 at a time into
// keys (column << 9 | hop slot; the diagonal is slot n), sorts the 512 keys in
// registers (sort512_regs) and writes the final column, FP64, FP32 and code
// arrays in striped order.  Nothing is staged in HBM.  The values and codes
// are those k_fill_wave stages, so the arrays equal the fill-then-sort result.
// LDS: split tables, pair table, pair list, then per wave 512 hop words, 512
// key words and the row's orbitals (16 words).
constexpr int kSortedWaveWords = 512 + 512 + 16;
static size_t fill_sorted_lds(int split_words, int kp_bytes, int N)
{
    const size_t kpw = (size_t)(kp_bytes + 3) / 4;
    const size_t npair = (size_t)N * (N - 1) / 2;
    return 4 * ((size_t)split_words + kpw + (npair * 2 + 3) / 4) + 4 * (size_t)kFillWaves * kSortedWaveWords;
}

template <bool CPLX, bool UPPER, bool LZ>
__global__ void __launch_bounds__(64 * kFillWaves)
k_fill_sorted(const uint64_t* __restrict__ states, int64_t dim, BuildArgs A, SplitArgs S,
              const int32_t* __restrict__ split_blob, const int64_t* __restrict__ rowptr,
              const typename VT<CPLX>::T* __restrict__ diag, const uint8_t* __restrict__ dkeep,
              int32_t* __restrict__ col_out, typename VT<CPLX>::T* __restrict__ val64,
              typename VT<CPLX>::F* __restrict__ val32, uint16_t* __restrict__ code_out,
              unsigned int* __restrict__ err)
{
    using T = typename VT<CPLX>::T;
    extern __shared__ int32_t lds[];
    const int K = A.K, N = A.N;
    const int kpw = (A.kpair_bytes + 3) / 4;
    const int npair = N * (N - 1) / 2;
    block_copy_words(lds, split_blob, S.words);
    const uint8_t* s_kp = reinterpret_cast<const uint8_t*>(lds + S.words);
    block_copy_words(lds + S.words, reinterpret_cast<const int32_t*>(A.kpair), kpw);
    uint8_t* s_pair = reinterpret_cast<uint8_t*>(lds + S.words + kpw);
    for (int P = threadIdx.x; P < npair; P += blockDim.x) {
        int ia = 0, rem = P;
        while (rem >= N - 1 - ia) {
            rem -= N - 1 - ia;
            ++ia;
        }
        s_pair[2 * P] = (uint8_t)ia;
        s_pair[2 * P + 1] = (uint8_t)(ia + 1 + rem);
    }
    const int wave = threadIdx.x >> 6, lane = threadIdx.x & 63;
    uint32_t* hops = reinterpret_cast<uint32_t*>(lds + S.words + kpw + (npair * 2 + 3) / 4) + wave * kSortedWaveWords;
    uint32_t* keys = hops + 512;
    uint8_t* s_orb = reinterpret_cast<uint8_t*>(keys + 512);
    __syncthreads();
    const T* P = reinterpret_cast<const T*>(A.pval);
    const int nun = K - N;
    const int ncand = npair * nun;
    const uint32_t magic = nun > 0 ? (uint32_t)((0x100000000ull + nun - 1) / (uint64_t)nun) : 0u;
    const uint64_t below = (1ull << lane) - 1ull;
    const int64_t nw = (int64_t)gridDim.x * (blockDim.x >> 6);
    for (int64_t row = (int64_t)blockIdx.x * (blockDim.x >> 6) + wave; row < dim; row += nw) {
        const uint64_t s = states[row];
        const int64_t start = rowptr[row];
        const int L = (int)(rowptr[row + 1] - start);
        if (lane < K) {
            const int nb = __popcll(s & below);
            if ((s >> lane) & 1ull) s_orb[nb] = (uint8_t)lane;
            else s_orb[N + lane - nb] = (uint8_t)lane;
        }
        __builtin_amdgcn_wave_barrier();
        // the row's hops, in the reference's candidate order
        int n = 0;
        for (int base = 0; base < ncand; base += 64) {
            const int qi = base + lane;
            bool ok = false;
            uint32_t h = 0;
            if (qi < ncand) {
                const int pi = (int)__umulhi((uint32_t)qi, magic);
                const int ci = qi - pi * nun;
                const int a = s_orb[s_pair[2 * pi]], b = s_orb[s_pair[2 * pi + 1]], c = s_orb[N + ci];
                const int e = s_kp[((a * (2 * K - a - 1) / 2 - a - 1) + b) * K + c];
                if (e != 0xFF && !(e & 0x80)) {
                    const int d = e & 0x3F;
                    if (!((s >> d) & 1ull)) {
                        const uint64_t out = (s ^ (1ull << a) ^ (1ull << b)) | (1ull << c) | (1ull << d);
                        ok = !(UPPER && out < s);
                        h = (uint32_t)a | ((uint32_t)b << 8) | ((uint32_t)c << 16) | ((uint32_t)d << 24);
                    }
                }
            }
            const uint64_t m = __ballot(ok);
            const int slot = n + __popcll(m & below);
            if (ok && slot < 512) hops[slot] = h;
            n += __popcll(m);
        }
        const bool has_diag = dkeep[row] != 0;
        const int total = n + (has_diag ? 1 : 0);
        if (total != L || total > 512) {  // the count pass saw another row: write nothing
            if (lane == 0) atomicOr(err, total > 512 ? 4u : 2u);
            __builtin_amdgcn_wave_barrier();
            continue;
        }
        __builtin_amdgcn_wave_barrier();
        // keys: rank << 9 | slot; the diagonal is slot n; padding sorts last
        for (int j = lane; j < 512; j += 64) {
            uint32_t kj = 0xFFFFFFFFu;
            if (j < n) {
                const uint32_t h = hops[j];
                const int a = h & 0xFF, b = (h >> 8) & 0xFF, c = (h >> 16) & 0xFF, d = h >> 24;
                const uint64_t out = (s ^ (1ull << a) ^ (1ull << b)) | (1ull << c) | (1ull << d);
                uint32_t r = split_rank<LZ>(lds, S, out);
                if ((int64_t)r >= dim) {  // never a column outside the sector
                    atomicOr(err, 1u);
                    r = 0;
                }
                kj = (r << 9) | (uint32_t)j;
            } else if (j == n && has_diag) {
                kj = ((uint32_t)row << 9) | (uint32_t)j;
            }
            keys[j] = kj;
        }
        __builtin_amdgcn_wave_barrier();
        uint32_t key[8];
#pragma unroll
        for (int t = 0; t < 8; ++t) key[t] = keys[lane * 8 + t];
        sort512_regs(key, lane);
        __builtin_amdgcn_wave_barrier();
#pragma unroll
        for (int t = 0; t < 8; ++t) keys[lane * 8 + t] = key[t];
        __builtin_amdgcn_wave_barrier();
#pragma unroll
        for (int t = 0; t < 8; ++t) {
            const int q = t * 64 + lane;
            if (q < total) {
                const uint32_t kq = keys[q];
                const int idx = (int)(kq & 511u);
                T v;
                uint16_t code;
                if (idx == n) {
                    v = diag[row];
                    code = 0xFFFF;
                } else {
                    const uint32_t h = hops[idx];
                    const int a = h & 0xFF, b = (h >> 8) & 0xFF, c = (h >> 16) & 0xFF, d = h >> 24;
                    const int ti = ((a * (2 * K - a - 1) / 2 - a - 1) + b) * K + c;
                    v = P[ti];
                    const int par = hop_parity(s, a, b, c, d);
                    if (par) v = vneg(v);
                    code = code_out ? (uint16_t)((uint16_t)A.remap[ti] | (par ? 0x8000u : 0u)) : (uint16_t)0;
                }
                col_out[start + q] = (int32_t)(kq >> 9);
                if (val64) val64[start + q] = v;
                if (val32) val32[start + q] = to_f32(v);
                if (code_out) code_out[start + q] = code;
            }
        }
        __builtin_amdgcn_wave_barrier();
    }
}

// ------------------------------------------------------------ scan kernels

constexpr int kScanBlock = 256;
constexpr int kScanItems = 16;

__global__ void __launch_bounds__(kScanBlock)
k_scan_partial(const int32_t* __restrict__ counts, int64_t n, long long* __restrict__ bsum,
               int* __restrict__ bmax)
{
    __shared__ long long ss[kScanBlock];
    __shared__ int sm[kScanBlock];
    const int64_t base = (int64_t)blockIdx.x * kScanBlock * kScanItems;
    long long s = 0;
    int m = 0;
    for (int i = 0; i < kScanItems; ++i) {
        const int64_t j = base + (int64_t)i * kScanBlock + threadIdx.x;
        if (j < n) {
            s += counts[j];
            m = max(m, counts[j]);
        }
    }
    ss[threadIdx.x] = s;
    sm[threadIdx.x] = m;
    __syncthreads();
    for (int o = kScanBlock / 2; o > 0; o >>= 1) {
        if ((int)threadIdx.x < o) {
            ss[threadIdx.x] += ss[threadIdx.x + o];
            sm[threadIdx.x] = max(sm[threadIdx.x], sm[threadIdx.x + o]);
        }
        __syncthreads();
    }
    if (threadIdx.x == 0) {
        bsum[blockIdx.x] = ss[0];
        bmax[blockIdx.x] = sm[0];
    }
}

__global__ void __launch_bounds__(1024)
k_scan_top(long long* __restrict__ bsum, const int* __restrict__ bmax, int nb,
           long long* __restrict__ total, int* __restrict__ gmax)
{
    // serial exclusive scan of block sums in chunks of 1024
    __shared__ long long buf[1024];
    __shared__ int mb[1024];
    __shared__ long long carry;
    if (threadIdx.x == 0) carry = 0;
    int m = 0;
    __syncthreads();
    for (int base = 0; base < nb; base += 1024) {
        const int i = base + threadIdx.x;
        const long long v = i < nb ? bsum[i] : 0;
        if (i < nb) m = max(m, bmax[i]);
        buf[threadIdx.x] = v;
        __syncthreads();
        for (int o = 1; o < 1024; o <<= 1) {
            const long long t = (int)threadIdx.x >= o ? buf[threadIdx.x - o] : 0;
            __syncthreads();
            buf[threadIdx.x] += t;
            __syncthreads();
        }
        if (i < nb) bsum[i] = carry + buf[threadIdx.x] - v;  // exclusive
        __syncthreads();
        if (threadIdx.x == 1023) carry += buf[1023];
        __syncthreads();
    }
    mb[threadIdx.x] = m;
    __syncthreads();
    for (int o = 512; o > 0; o >>= 1) {
        if ((int)threadIdx.x < o) mb[threadIdx.x] = max(mb[threadIdx.x], mb[threadIdx.x + o]);
        __syncthreads();
    }
    if (threadIdx.x == 0) {
        *total = carry;
        *gmax = mb[0];
    }
}

__global__ void __launch_bounds__(kScanBlock)
k_scan_final(const int32_t* __restrict__ counts, int64_t n, const long long* __restrict__ bsum,
             int64_t* __restrict__ rowptr)
{
    __shared__ long long ss[kScanBlock];
    const int64_t base = (int64_t)blockIdx.x * kScanBlock * kScanItems;
    // each thread owns kScanItems consecutive items
    long long loc[kScanItems];
    long long s = 0;
    for (int i = 0; i < kScanItems; ++i) {
        const int64_t j = base + (int64_t)threadIdx.x * kScanItems + i;
        loc[i] = j < n ? counts[j] : 0;
        s += loc[i];
    }
    ss[threadIdx.x] = s;
    __syncthreads();
    for (int o = 1; o < kScanBlock; o <<= 1) {
        const long long t = (int)threadIdx.x >= o ? ss[threadIdx.x - o] : 0;
        __syncthreads();
        ss[threadIdx.x] += t;
        __syncthreads();
    }
    long long run = bsum[blockIdx.x] + ss[threadIdx.x] - s;
    for (int i = 0; i < kScanItems; ++i) {
        const int64_t j = base + (int64_t)threadIdx.x * kScanItems + i;
        if (j < n) rowptr[j] = run;
        run += loc[i];
        if (j == n - 1) rowptr[n] = run;
    }
}

int scan_counts(const int32_t* d_counts, int64_t n, int64_t* d_rowptr, int64_t* total, int* max_count,
                hipStream_t st)
{
    const int64_t per = (int64_t)kScanBlock * kScanItems;
    const int nb = (int)std::max<int64_t>(1, (n + per - 1) / per);
    long long* d_bsum = nullptr;
    int* d_bmax = nullptr;
    long long* d_res = nullptr;  // [0] total, [1] max (as int)
    auto cleanup = [&]() { dfree(d_bsum); dfree(d_bmax); dfree(d_res); };
    int status = dalloc_t(&d_bsum, (size_t)nb, "scan partials");
    if (status == HD_OK) status = dalloc_t(&d_bmax, (size_t)nb, "scan maxima");
    if (status == HD_OK) status = dalloc_t(&d_res, 2, "scan results");
    if (status != HD_OK) {
        cleanup();
        return status;
    }
    long long host[2] = {0, 0};
    HipErrors q;
    HD_ENQ(q, hipMemsetAsync(d_res, 0, 2 * sizeof(long long), st));
    HD_ENQ(q, hipMemsetAsync(d_rowptr, 0, sizeof(int64_t), st));
    if (n > 0) {
        hipLaunchKernelGGL(k_scan_partial, dim3(nb), dim3(kScanBlock), 0, st, d_counts, n, d_bsum, d_bmax);
        hipLaunchKernelGGL(k_scan_top, dim3(1), dim3(1024), 0, st, d_bsum, d_bmax, nb, d_res,
                           reinterpret_cast<int*>(d_res + 1));
        hipLaunchKernelGGL(k_scan_final, dim3(nb), dim3(kScanBlock), 0, st, d_counts, n, d_bsum, d_rowptr);
        HD_ENQ(q, hipGetLastError());
    }
    HD_ENQ(q, hipMemcpyAsync(host, d_res, sizeof(host), hipMemcpyDeviceToHost, st));
    HD_ENQ(q, hipStreamSynchronize(st));
    cleanup();
    HD_CHECK(q.status());
    *total = (int64_t)host[0];
    *max_count = (int)host[1];
    return HD_OK;
}

// --------------------------------------------------------------- launchers

static int set_lds(const void* fn, size_t lds)
{
    if (lds > 160 * 1024)
        return fail(HD_ERR_UNSUPPORTED, "build tables need " + std::to_string(lds) +
                                            " bytes of LDS (limit 160 KiB)");
    if (lds > 64 * 1024) HD_HIP(hipFuncSetAttribute(fn, hipFuncAttributeMaxDynamicSharedMemorySize, (int)lds));
    return HD_OK;
}

// Row-walk kernel of the fill and of the matrix-free product: by default the
// fill and sort in one pass (k_fill_sorted) where rows and tables allow, else
// the wavefront walk with split ranking where its tables fit, else the
// wavefront walk with the rank table, else one thread per row; the last three
// stage their entries for k_sort512 / k_sort.  HD_BUILD_FILL = "split", "rank"
// or "thread" forces the later choices (when they apply), and with them the
// 32-bit-key k_sort512, the 64-bit-key k_sort512 and the LDS k_sort<1024>, so
// every kernel can be checked against the others on the same matrices.
enum class FillPath { kAuto, kSplit, kRank, kThread };
static FillPath fill_path()
{
    const char* e = std::getenv("HD_BUILD_FILL");
    if (!e) return FillPath::kAuto;
    if (!std::strcmp(e, "split")) return FillPath::kSplit;
    if (!std::strcmp(e, "rank")) return FillPath::kRank;
    if (!std::strcmp(e, "thread")) return FillPath::kThread;
    return FillPath::kAuto;
}

// Host side of split_rank: L, H and the middle's tables from the rank table of
// make_rank_table, with m1 and m2 chosen for the fewest middle orbitals whose
// tables fit in `budget` bytes.  Returns false (no split) for a list basis or
// when nothing fits.
static bool make_split_rank(const RankTable& rt, size_t budget, SplitArgs& S, std::vector<int32_t>& out)
{
    if (rt.empty || (rt.is_lz != 0 && rt.is_lz != 1)) return false;
    const int32_t* bl = rt.blob.data();
    const int N = bl[0], K = bl[1], G = bl[2], target = bl[3];
    const bool lz = bl[4] == 1;
    if (K > 64 || N < 1 || N > 255) return false;
    const int npc = (K + 1) * (N + 1);
    const int32_t* off = bl + kRankHeader;
    const int32_t* gmm = off + npc;
    const int32_t* sub = gmm + 2 * npc;
    const uint32_t* tab = reinterpret_cast<const uint32_t*>(sub + (lz ? K : G * K));
    auto W = [&](int p, int c, int g) -> uint32_t {
        if (c < 0 || c > N || p < 0 || p > K) return 0u;
        const int i = p * (N + 1) + c;
        if (g < gmm[2 * i] || g > gmm[2 * i + 1]) return 0u;
        return tab[off[i] + g];
    };
    auto gsub = [&](int g, int k) { return lz ? g - sub[k] : sub[g * K + k]; };
    // entries of the middle's counting table for orbitals [a, b)
    auto mid_entries = [&](int a, int b) {
        size_t n = 0;
        for (int p = a; p < b; ++p)
            for (int c = 0; c <= N; ++c) {
                const int i = p * (N + 1) + c;
                if (gmm[2 * i + 1] >= gmm[2 * i]) n += (size_t)(gmm[2 * i + 1] - gmm[2 * i] + 1);
            }
        return n;
    };
    int best_m1 = -1, best_hb = -1;
    size_t best_bytes = 0;
    for (int m1 = std::min(K, 16); m1 >= std::min(K, 4); --m1)
        for (int hb = std::min(K - m1, 14); hb >= 0; --hb) {
            const int nmid = K - m1 - hb;
            const size_t bytes = 4 * ((size_t)1 << m1) + 8 * ((size_t)1 << hb) + 8 * (size_t)nmid * (N + 1) +
                                 4 * (size_t)nmid * (lz ? 1 : G) + 4 * mid_entries(m1, m1 + nmid);
            if (bytes > budget) continue;
            const int bnmid = best_m1 < 0 ? 1 << 30 : K - best_m1 - best_hb;
            if (nmid < bnmid || (nmid == bnmid && bytes < best_bytes)) {
                best_m1 = m1;
                best_hb = hb;
                best_bytes = bytes;
            }
        }
    if (best_m1 < 0) return false;
    const int m1 = best_m1, m2 = K - best_hb, nmid = m2 - m1;
    // group sum (the inverse of gsub) for the low table
    std::vector<int> gadd;
    if (!lz) {
        gadd.assign((size_t)G * K, 0);
        for (int g2 = 0; g2 < G; ++g2)
            for (int k = 0; k < K; ++k) gadd[(size_t)sub[g2 * K + k] * K + k] = g2;
    }
    out.clear();
    // L: ascending over the bits of l, c and g counted up to and including p
    for (uint32_t l = 0; l < (1u << m1); ++l) {
        uint32_t r = 0;
        int c = 0, g = 0;
        for (int p = 0; p < m1; ++p) {
            if (!((l >> p) & 1u)) continue;
            ++c;
            g = lz ? g + sub[p] : gadd[(size_t)g * K + p];
            r += W(p, c, g);
        }
        out.push_back((int32_t)r);
    }
    S.oH = (int)out.size();
    // H: descending from (N, target); impossible patterns get zeros
    for (uint32_t hpat = 0; hpat < (1u << (K - m2)); ++hpat) {
        uint32_t r = 0;
        int c = N, g = target;
        bool ok = true;
        for (int p = K - 1; p >= m2 && ok; --p) {
            if (!((hpat >> (p - m2)) & 1u)) continue;
            if (c <= 0) {
                ok = false;
                break;
            }
            r += W(p, c, g);
            g = gsub(g, p);
            --c;
            if (lz && g < 0) ok = false;
        }
        if (!ok || g >= (1 << 24)) r = 0, c = 0, g = 0;
        out.push_back((int32_t)r);
        out.push_back((int32_t)((uint32_t)c | ((uint32_t)g << 8)));
    }
    // the middle: (p, c) ranges, weights, counts
    S.oPk = (int)out.size();
    const int pk_words = 2 * nmid * (N + 1);
    out.resize(out.size() + (size_t)pk_words);
    S.oSub = (int)out.size();
    if (lz) {
        for (int q = 0; q < nmid; ++q) out.push_back(sub[m1 + q]);
    } else {
        for (int g = 0; g < G; ++g)
            for (int q = 0; q < nmid; ++q) out.push_back(sub[g * K + m1 + q]);
    }
    S.oTab = (int)out.size();
    for (int q = 0; q < nmid; ++q)
        for (int c = 0; c <= N; ++c) {
            const int i = (m1 + q) * (N + 1) + c;
            const int lo = gmm[2 * i], hi = gmm[2 * i + 1];
            int32_t* e = out.data() + S.oPk + 2 * (q * (N + 1) + c);
            if (hi < lo) {
                e[0] = 0;
                e[1] = (int32_t)(1u | (0u << 16));  // lo 1 > hi 0: empty
                continue;
            }
            if (lo < -32768 || hi > 32767) return false;
            const int base = (int)out.size() - S.oTab;
            for (int g = lo; g <= hi; ++g) out.push_back((int32_t)tab[off[i] + g]);
            e = out.data() + S.oPk + 2 * (q * (N + 1) + c);  // push_back may have moved the storage
            e[0] = base - lo;
            e[1] = (int32_t)(((uint32_t)(uint16_t)(int16_t)lo) | ((uint32_t)(uint16_t)(int16_t)hi << 16));
        }
    S.words = (int)out.size();
    S.m1 = m1;
    S.m2 = m2;
    S.N = N;
    return 4 * out.size() <= budget;
}

template <bool CPLX, bool UPPER, bool LZ>
static int launch_fill(int grid, size_t lds, hipStream_t st, const hd_basis* b, const BuildArgs& A,
                       int rank_words, const int64_t* rowptr, const typename VT<CPLX>::T* diag,
                       const uint8_t* dkeep, int32_t* ctmp, typename VT<CPLX>::T* vtmp,
                       uint16_t* codetmp, unsigned int* err, const int32_t* d_split, const SplitArgs& S)
{
    int n_cu = 256;
    HD_CHECK(device_props(&n_cu, nullptr));
    const FillPath path = fill_path();
    if (d_split && A.N >= 2 && A.K > A.N && (path == FillPath::kAuto || path == FillPath::kSplit)) {
        const size_t wlds = fill_wave_lds(S.words, A.kpair_bytes, A.N);
        HD_CHECK(set_lds((const void*)k_fill_wave<CPLX, UPPER, LZ, true>, wlds));
        const int wgrid = (int)std::max<int64_t>(
            1, std::min<int64_t>((b->dim + kFillWaves - 1) / kFillWaves, (int64_t)n_cu * (wlds > 80 * 1024 ? 1 : 2)));
        hipLaunchKernelGGL((k_fill_wave<CPLX, UPPER, LZ, true>), dim3(wgrid), dim3(64 * kFillWaves), wlds, st,
                           b->d_states, b->dim, A, S, d_split, S.words, rowptr, diag, dkeep, ctmp, vtmp,
                           codetmp, err);
        return HD_OK;
    }
    const size_t wlds = fill_wave_lds(rank_words, A.kpair_bytes, A.N);
    if (A.N >= 2 && A.K > A.N && wlds <= 160 * 1024 && path != FillPath::kThread) {
        HD_CHECK(set_lds((const void*)k_fill_wave<CPLX, UPPER, LZ, false>, wlds));
        const int wgrid = (int)std::max<int64_t>(
            1, std::min<int64_t>((b->dim + kFillWaves - 1) / kFillWaves, (int64_t)n_cu * (wlds > 80 * 1024 ? 1 : 2)));
        hipLaunchKernelGGL((k_fill_wave<CPLX, UPPER, LZ, false>), dim3(wgrid), dim3(64 * kFillWaves), wlds, st,
                           b->d_states, b->dim, A, S, b->d_rank, rank_words, rowptr, diag, dkeep, ctmp, vtmp,
                           codetmp, err);
        return HD_OK;
    }
    HD_CHECK(set_lds((const void*)k_fill<CPLX, UPPER, LZ>, lds));
    hipLaunchKernelGGL((k_fill<CPLX, UPPER, LZ>), dim3(grid), dim3(1024), lds, st, b->d_states, b->dim, A,
                       b->d_rank, rank_words, rowptr, diag, dkeep, ctmp, vtmp, codetmp, err);
    return HD_OK;
}

template <bool CPLX>
static int run_build(hd_hamiltonian* h, const hd_basis* b, const Derived& D, double tol,
                     int value_mask)
{
    using T = typename VT<CPLX>::T;
    using F = typename VT<CPLX>::F;
    const hipStream_t st = h->stream;
    const int64_t dim = b->dim;
    const int K = D.K;
    int n_cu = 256;
    HD_CHECK(device_props(&n_cu, nullptr));
    // upload derived tables
    uint8_t *d_kpair = nullptr, *d_dflag = nullptr, *d_epsflag = nullptr, *d_dkeep = nullptr;
    T *d_pval = nullptr, *d_dterm = nullptr, *d_eps = nullptr, *d_diag = nullptr, *d_vtmp = nullptr;
    int32_t *d_counts = nullptr, *d_ctmp = nullptr;
    uint16_t* d_codetmp = nullptr;
    int16_t* d_remap = nullptr;
    const bool coded = (value_mask & 4) != 0;
    long long* d_bsum = nullptr;
    int* d_bmax = nullptr;
    unsigned long long* d_misc = nullptr;  // [0] trimmed, [1] total, [2] max, [3] error flags
    int32_t* d_split = nullptr;
    int status = HD_OK;
    const int kp_bytes = (int)D.kpair.size();
    const size_t kp_alloc = ((size_t)kp_bytes + 3) & ~(size_t)3;
    auto cleanup = [&]() {
        dfree(d_kpair); dfree(d_dflag); dfree(d_epsflag); dfree(d_dkeep); dfree(d_pval);
        dfree(d_dterm); dfree(d_eps); dfree(d_diag); dfree(d_vtmp); dfree(d_counts);
        dfree(d_ctmp); dfree(d_bsum); dfree(d_bmax); dfree(d_misc); dfree(d_codetmp); dfree(d_remap);
        dfree(d_split);
    };
#define HD_TRY(x)                  \
    do {                           \
        status = (x);              \
        if (status != HD_OK) {     \
            cleanup();             \
            return status;         \
        }                          \
    } while (0)
#define HD_TRYHIP(x)                                                                      \
    do {                                                                                  \
        hipError_t _e = (x);                                                              \
        if (_e != hipSuccess) {                                                           \
            cleanup();                                                                    \
            return fail(HD_ERR_HIP, std::string(#x) + ": " + hipGetErrorString(_e));     \
        }                                                                                 \
    } while (0)
    HD_TRY(dalloc(reinterpret_cast<void**>(&d_kpair), kp_alloc, "pair table"));
    HD_TRY(dalloc_t(&d_pval, D.pval.size() / (CPLX ? 2 : 1), "pair values"));
    HD_TRY(dalloc_t(&d_dterm, (size_t)2 * K * K, "diagonal terms"));
    HD_TRY(dalloc_t(&d_dflag, (size_t)K * K, "diagonal flags"));
    HD_TRY(dalloc_t(&d_eps, (size_t)K, "quadratic terms"));
    HD_TRY(dalloc_t(&d_epsflag, (size_t)K, "quadratic flags"));
    HD_TRY(dalloc_t(&d_misc, 4, "counters"));
    HD_TRYHIP(hipMemsetAsync(d_kpair, 0xFF, kp_alloc, st));
    HD_TRYHIP(hipMemcpyAsync(d_kpair, D.kpair.data(), kp_bytes, hipMemcpyHostToDevice, st));
    HD_TRYHIP(hipMemcpyAsync(d_pval, D.pval.data(), D.pval.size() * 8, hipMemcpyHostToDevice, st));
    HD_TRYHIP(hipMemcpyAsync(d_dterm, D.dterm.data(), D.dterm.size() * 8, hipMemcpyHostToDevice, st));
    HD_TRYHIP(hipMemcpyAsync(d_dflag, D.dflag.data(), D.dflag.size(), hipMemcpyHostToDevice, st));
    HD_TRYHIP(hipMemcpyAsync(d_eps, D.eps.data(), D.eps.size() * 8, hipMemcpyHostToDevice, st));
    HD_TRYHIP(hipMemcpyAsync(d_epsflag, D.epsflag.data(), D.epsflag.size(), hipMemcpyHostToDevice, st));
    HD_TRYHIP(hipMemsetAsync(d_misc, 0, 4 * sizeof(unsigned long long), st));

    BuildArgs A;
    A.N = b->N;
    A.Norb = b->Norb;
    A.K = K;
    A.upper = h->storage == HD_STORAGE_UPPER_REF;
    A.has_eps = D.has_eps;
    A.split = D.split;
    A.tol = tol;
    A.norb_mask = b->Norb >= 64 ? ~0ull : ((1ull << b->Norb) - 1ull);
    A.kpair = d_kpair;
    A.pval = d_pval;
    A.dterm = d_dterm;
    A.dflag = d_dflag;
    A.eps = d_eps;
    A.epsflag = d_epsflag;
    A.kpair_bytes = kp_bytes;
    A.remap = nullptr;
    if (coded) {
        HD_TRY(dalloc_t(&d_remap, D.remap.size(), "code remap"));
        HD_TRYHIP(hipMemcpyAsync(d_remap, D.remap.data(), D.remap.size() * 2, hipMemcpyHostToDevice, st));
        A.remap = d_remap;
    }

    // ---- count
    HD_TRY(dalloc_t(&d_counts, (size_t)dim, "row counts"));
    HD_TRY(dalloc_t(&d_diag, (size_t)dim, "diagonal"));
    HD_TRY(dalloc_t(&d_dkeep, (size_t)dim, "diagonal flags"));
    HD_TRY(dalloc_t(&h->d_rowptr, (size_t)dim + 1, "row pointers"));
    h->bytes += ((size_t)dim + 1) * 8;
    EventTimer ev;
    {
        const size_t lds = (((kp_alloc / 4) + 3) & ~(size_t)3) * 4 + sizeof(T) * 2 * K * K +
                           (((size_t)K * K + 15) & ~(size_t)15) + sizeof(T) * K + K + 16;
        HD_TRY(set_lds((const void*)k_count<CPLX>, lds));
        const int grid = (int)std::max<int64_t>(1, std::min<int64_t>((dim + 255) / 256, (int64_t)n_cu * 8));
        ev.start(st);
        hipLaunchKernelGGL(k_count<CPLX>, dim3(grid), dim3(256), lds, st, b->d_states, dim, A,
                           d_counts, d_diag, d_dkeep, d_misc);
        HD_TRYHIP(hipGetLastError());
        h->t_count = ev.stop(st);
    }
    // ---- scan
    unsigned long long host_misc[4] = {0, 0, 0, 0};
    {
        const int64_t per = (int64_t)kScanBlock * kScanItems;
        const int nb = (int)((dim + per - 1) / per);
        HD_TRY(dalloc_t(&d_bsum, (size_t)nb, "scan partials"));
        HD_TRY(dalloc_t(&d_bmax, (size_t)nb, "scan maxima"));
        ev.start(st);
        hipLaunchKernelGGL(k_scan_partial, dim3(nb), dim3(kScanBlock), 0, st, d_counts, dim, d_bsum, d_bmax);
        hipLaunchKernelGGL(k_scan_top, dim3(1), dim3(1024), 0, st, d_bsum, d_bmax, nb,
                           reinterpret_cast<long long*>(d_misc + 1), reinterpret_cast<int*>(d_misc + 2));
        hipLaunchKernelGGL(k_scan_final, dim3(nb), dim3(kScanBlock), 0, st, d_counts, dim, d_bsum,
                           h->d_rowptr);
        HD_TRYHIP(hipGetLastError());
        h->t_scan = ev.stop(st);
        HD_TRYHIP(hipMemcpyAsync(host_misc, d_misc, sizeof(host_misc), hipMemcpyDeviceToHost, st));
        HD_TRYHIP(hipStreamSynchronize(st));
    }
    h->nnz = (int64_t)host_misc[1];
    h->n_trimmed = (int64_t)host_misc[0];
    h->max_row = (int64_t)(int)host_misc[2];
    dfree(d_counts);
    d_counts = nullptr;
    if (h->storage == HD_STORAGE_MATRIX_FREE) {
        // keep the pair tables and the diagonal (0 where the CRS holds none)
        hipLaunchKernelGGL(k_fold_diag<T>, dim3((unsigned)((dim + 255) / 256)), dim3(256), 0, st, dim, d_diag, d_dkeep);
        HD_TRYHIP(hipGetLastError());
        HD_TRYHIP(hipStreamSynchronize(st));
        h->d_diag = d_diag;
        h->d_mf_kpair = d_kpair;
        h->d_mf_pval = d_pval;
        h->mf_K = K;
        h->mf_kp_bytes = kp_bytes;
        h->mf_norb_mask = A.norb_mask;
        d_diag = nullptr;
        d_kpair = nullptr;
        d_pval = nullptr;
        {
            // the product's ranking tables (k_spmv_free_wave), when they fit
            SplitArgs S{};
            std::vector<int32_t> split;
            const size_t wfix = fill_wave_lds(0, kp_bytes, b->N);
            const FillPath mf_path = fill_path();
            if (b->N >= 2 && K > b->N && wfix < 160 * 1024 &&
                (mf_path == FillPath::kAuto || mf_path == FillPath::kSplit) &&
                make_split_rank(b->rt, 160 * 1024 - wfix, S, split)) {
                HD_TRY(dalloc_t(&h->d_mf_split, split.size(), "split rank tables"));
                HD_TRYHIP(hipMemcpyAsync(h->d_mf_split, split.data(), split.size() * 4, hipMemcpyHostToDevice, st));
                HD_TRYHIP(hipStreamSynchronize(st));
                h->mf_split = S;
            }
        }
        dfree(h->d_rowptr);
        h->d_rowptr = nullptr;
        h->bytes = (size_t)dim * sizeof(T) + kp_alloc + D.pval.size() * 8;
        cleanup();
        return HD_OK;
    }
    if (h->max_row > 4096) {
        cleanup();
        return fail(HD_ERR_UNSUPPORTED, "row of " + std::to_string(h->max_row) +
                                            " entries exceeds the 4096-entry row sort");
    }
    // ---- fill
    const size_t nnz = (size_t)h->nnz;
    // every entry array is padded by 8 zeroed entries: the SpMV reads aligned
    // groups of 4 that may run past the last row
    auto alloc_final = [&]() -> int {
        int rc = dalloc_t(&h->d_col, nnz + 8, "columns");
        if (rc != HD_OK) return rc;
        HipErrors q;
        HD_ENQ(q, hipMemsetAsync(h->d_col + nnz, 0, 8 * sizeof(int32_t), st));
        h->bytes += nnz * 4;
        if (value_mask & 1) {
            if ((rc = dalloc(&h->d_val64, (nnz + 8) * sizeof(T), "FP64 values")) != HD_OK) return rc;
            HD_ENQ(q, hipMemsetAsync(static_cast<T*>(h->d_val64) + nnz, 0, 8 * sizeof(T), st));
            h->bytes += nnz * sizeof(T);
        }
        if (value_mask & 2) {
            if ((rc = dalloc(&h->d_val32, (nnz + 8) * sizeof(F), "FP32 values")) != HD_OK) return rc;
            HD_ENQ(q, hipMemsetAsync(static_cast<F*>(h->d_val32) + nnz, 0, 8 * sizeof(F), st));
            h->bytes += nnz * sizeof(F);
        }
        if (coded) {
            // padded by 8 entries so the SpMV's aligned 4-wide loads stay in bounds
            if ((rc = dalloc_t(&h->d_code, nnz + 8, "value codes")) != HD_OK) return rc;
            if ((rc = dalloc_t(reinterpret_cast<T**>(&h->d_dict), (size_t)std::max(D.ndict, 1), "value dictionary")) !=
                HD_OK)
                return rc;
            HD_ENQ(q, hipMemcpyAsync(h->d_dict, D.dict.data(), D.dict.size() * 8, hipMemcpyHostToDevice, st));
            HD_ENQ(q, hipMemsetAsync(h->d_code + nnz, 0xFF, 16, st));
            h->ndict = D.ndict;
            h->bytes += (nnz + 8) * 2 + (size_t)D.ndict * sizeof(T);
        }
        return q.status();
    };
    // fill and sort in one pass where the rows and the split tables allow
    {
        SplitArgs FS{};
        std::vector<int32_t> fsplit;
        const size_t ffix = fill_sorted_lds(0, A.kpair_bytes, A.N);
        if (fill_path() == FillPath::kAuto && A.N >= 2 && K > A.N && h->max_row <= 512 &&
            dim < (int64_t(1) << 23) && ffix < 160 * 1024 &&
            make_split_rank(b->rt, 160 * 1024 - ffix, FS, fsplit)) {
            HD_TRY(alloc_final());
            HD_TRY(dalloc_t(&d_split, fsplit.size(), "split rank tables"));
            HD_TRYHIP(hipMemcpyAsync(d_split, fsplit.data(), fsplit.size() * 4, hipMemcpyHostToDevice, st));
            const size_t flds = fill_sorted_lds(FS.words, A.kpair_bytes, A.N);
            const bool lz = b->rt.is_lz != 0;
            const int fgrid = (int)std::max<int64_t>(
                1, std::min<int64_t>((dim + kFillWaves - 1) / kFillWaves, (int64_t)n_cu * (flds > 80 * 1024 ? 1 : 2)));
            unsigned int* d_err = reinterpret_cast<unsigned int*>(d_misc + 3);
            T* v64 = reinterpret_cast<T*>(h->d_val64);
            F* v32 = reinterpret_cast<F*>(h->d_val32);
#define HD_FS(U, L)                                                                                            \
    do {                                                                                                       \
        HD_TRY(set_lds((const void*)k_fill_sorted<CPLX, U, L>, flds));                                         \
        hipLaunchKernelGGL((k_fill_sorted<CPLX, U, L>), dim3(fgrid), dim3(64 * kFillWaves), flds, st,          \
                           b->d_states, dim, A, FS, d_split, h->d_rowptr, d_diag, d_dkeep, h->d_col, v64, v32, \
                           h->d_code, d_err);                                                                  \
    } while (0)
            ev.start(st);
            if (A.upper && lz) HD_FS(true, true);
            else if (A.upper) HD_FS(true, false);
            else if (lz) HD_FS(false, true);
            else HD_FS(false, false);
#undef HD_FS
            HD_TRYHIP(hipGetLastError());
            h->t_fill = ev.stop(st);
            h->t_sort = 0.0;
            HD_TRYHIP(hipMemcpyAsync(host_misc, d_misc, sizeof(host_misc), hipMemcpyDeviceToHost, st));
            HD_TRYHIP(hipStreamSynchronize(st));
            HD_TRY(ev.status("build stage timing"));
            if (host_misc[3] != 0) {
                cleanup();
                return fail(HD_ERR_UNSUPPORTED, "fill pass inconsistency (flags " +
                                                    std::to_string(host_misc[3]) + ")");
            }
            if (coded) {
                h->d_diag = d_diag;  // the coded format reads the diagonal from here
                h->bytes += (size_t)dim * sizeof(T);
                d_diag = nullptr;
            }
            cleanup();
            return HD_OK;
        }
    }
    HD_TRY(dalloc_t(&d_ctmp, nnz, "unsorted columns"));
    HD_TRY(dalloc_t(&d_vtmp, nnz, "unsorted values"));
    if (coded) HD_TRY(dalloc_t(&d_codetmp, nnz, "unsorted codes"));
    {
        const int rank_words = (int)b->rt.blob.size();
        const size_t lds = (size_t)rank_words * 4 + kp_alloc;
        const bool lz = b->rt.is_lz != 0;
        const int per_cu = lds > 80 * 1024 ? 1 : (lds > 40 * 1024 ? 2 : 4);
        const int grid = (int)std::max<int64_t>(1, std::min<int64_t>((dim + 1023) / 1024, (int64_t)n_cu * per_cu));
        unsigned int* d_err = reinterpret_cast<unsigned int*>(d_misc + 3);
        // split rank tables in the LDS the wave fill leaves (its other tables at
        // 0 words of rank table)
        SplitArgs S{};
        std::vector<int32_t> split;
        const size_t wfix = fill_wave_lds(0, A.kpair_bytes, A.N);
        if (wfix < 160 * 1024 && make_split_rank(b->rt, 160 * 1024 - wfix, S, split)) {
            HD_TRY(dalloc_t(&d_split, split.size(), "split rank tables"));
            HD_TRYHIP(hipMemcpyAsync(d_split, split.data(), split.size() * 4, hipMemcpyHostToDevice, st));
        }
        ev.start(st);
        if (A.upper && lz)
            HD_TRY((launch_fill<CPLX, true, true>(grid, lds, st, b, A, rank_words, h->d_rowptr, d_diag, d_dkeep, d_ctmp, d_vtmp, d_codetmp, d_err, d_split, S)));
        else if (A.upper)
            HD_TRY((launch_fill<CPLX, true, false>(grid, lds, st, b, A, rank_words, h->d_rowptr, d_diag, d_dkeep, d_ctmp, d_vtmp, d_codetmp, d_err, d_split, S)));
        else if (lz)
            HD_TRY((launch_fill<CPLX, false, true>(grid, lds, st, b, A, rank_words, h->d_rowptr, d_diag, d_dkeep, d_ctmp, d_vtmp, d_codetmp, d_err, d_split, S)));
        else
            HD_TRY((launch_fill<CPLX, false, false>(grid, lds, st, b, A, rank_words, h->d_rowptr, d_diag, d_dkeep, d_ctmp, d_vtmp, d_codetmp, d_err, d_split, S)));
        HD_TRYHIP(hipGetLastError());
        h->t_fill = ev.stop(st);
        HD_TRYHIP(hipMemcpyAsync(host_misc, d_misc, sizeof(host_misc), hipMemcpyDeviceToHost, st));
        HD_TRYHIP(hipStreamSynchronize(st));
        if (host_misc[3] != 0) {
            cleanup();
            return fail(HD_ERR_UNSUPPORTED, "fill pass inconsistency (flags " +
                                                std::to_string(host_misc[3]) + ")");
        }
    }
    if (coded) {
        h->d_diag = d_diag;  // the coded format reads the diagonal from here
        h->bytes += (size_t)dim * sizeof(T);
    } else {
        dfree(d_diag);
    }
    d_diag = nullptr;
    dfree(d_dkeep); d_dkeep = nullptr;
    // ---- sort into the final arrays
    HD_TRY(alloc_final());
    {
        const int grid = (int)std::max<int64_t>(1, std::min<int64_t>(dim, (int64_t)n_cu * 32));
        ev.start(st);
        T* v64 = reinterpret_cast<T*>(h->d_val64);
        F* v32 = reinterpret_cast<F*>(h->d_val32);
        const dim3 g512((unsigned)std::max<int64_t>(1, std::min<int64_t>((dim + 3) / 4, (int64_t)n_cu * 32)));
        // HD_BUILD_FILL=rank takes the 64-bit keys and =thread the LDS sort, so
        // that every sort kernel is checked on small matrices too
        const FillPath sp = fill_path();
        if (sp == FillPath::kThread && h->max_row <= 1024)
            hipLaunchKernelGGL((k_sort<CPLX, 1024>), dim3(grid), dim3(64), 0, st, dim, h->d_rowptr, d_ctmp, d_vtmp, d_codetmp, h->d_col, v64, v32, h->d_code);
        else if (h->max_row <= 512 && dim < (int64_t(1) << 23) && sp != FillPath::kRank)
            hipLaunchKernelGGL((k_sort512<CPLX, uint32_t>), g512, dim3(256), 0, st, dim, h->d_rowptr, d_ctmp, d_vtmp,
                               d_codetmp, h->d_col, v64, v32, h->d_code);
        else if (h->max_row <= 512)
            hipLaunchKernelGGL((k_sort512<CPLX, uint64_t>), g512, dim3(256), 0, st, dim, h->d_rowptr, d_ctmp, d_vtmp,
                               d_codetmp, h->d_col, v64, v32, h->d_code);
        else if (h->max_row <= 1024)
            hipLaunchKernelGGL((k_sort<CPLX, 1024>), dim3(grid), dim3(64), 0, st, dim, h->d_rowptr, d_ctmp, d_vtmp, d_codetmp, h->d_col, v64, v32, h->d_code);
        else if (h->max_row <= 2048)
            hipLaunchKernelGGL((k_sort<CPLX, 2048>), dim3(grid), dim3(64), 0, st, dim, h->d_rowptr, d_ctmp, d_vtmp, d_codetmp, h->d_col, v64, v32, h->d_code);
        else
            hipLaunchKernelGGL((k_sort<CPLX, 4096>), dim3(grid), dim3(64), 0, st, dim, h->d_rowptr, d_ctmp, d_vtmp, d_codetmp, h->d_col, v64, v32, h->d_code);
        HD_TRYHIP(hipGetLastError());
        h->t_sort = ev.stop(st);
    }
    HD_TRYHIP(hipStreamSynchronize(st));
    HD_TRY(ev.status("build stage timing"));
    cleanup();
#undef HD_TRY
#undef HD_TRYHIP
    return HD_OK;
}

// One sector of a batch: tables[0..n) share their k table and differ in the
// values.  The CRS pattern keeps every hop that survives the trim under at
// least one table; the code of an entry indexes the tuple of its hop value
// under all tables (0.0 where the hop is trimmed or absent), numbered in order
// of first appearance across the sectors of the batch.
int matrix_free_launch(const hd_hamiltonian* h, const void* x, void* y, hipStream_t stream)
{
    const hd_basis* b = h->basis;
    if (h->dim == 0) return HD_OK;
    if (h->d_mf_split) {
        const SplitArgs& S = h->mf_split;
        const size_t wlds = fill_wave_lds(S.words, h->mf_kp_bytes, S.N);
        int n_cu = 256;
        HD_CHECK(device_props(&n_cu, nullptr));
        const int wgrid = (int)std::max<int64_t>(
            1, std::min<int64_t>((h->dim + kFillWaves - 1) / kFillWaves, (int64_t)n_cu * (wlds > 80 * 1024 ? 1 : 2)));
        const bool lzw = b->rt.is_lz != 0;
#define HD_MFW(C, L)                                                                                            \
    do {                                                                                                        \
        using T = typename VT<C>::T;                                                                            \
        HD_CHECK(set_lds((const void*)k_spmv_free_wave<C, L>, wlds));                                           \
        hipLaunchKernelGGL((k_spmv_free_wave<C, L>), dim3(wgrid), dim3(64 * kFillWaves), wlds, stream,          \
                           b->d_states, h->dim, S.N, h->mf_K, S, h->d_mf_split, h->d_mf_kpair, h->mf_kp_bytes, \
                           static_cast<const T*>(h->d_mf_pval), static_cast<const T*>(h->d_diag),              \
                           static_cast<const T*>(x), static_cast<T*>(y));                                      \
    } while (0)
        if (h->is_complex) {
            if (lzw) HD_MFW(true, true);
            else HD_MFW(true, false);
        } else {
            if (lzw) HD_MFW(false, true);
            else HD_MFW(false, false);
        }
#undef HD_MFW
        HD_HIP(hipGetLastError());
        return HD_OK;
    }
    const int rank_words = (int)b->rt.blob.size();
    const size_t kp_alloc = ((size_t)h->mf_kp_bytes + 3) & ~(size_t)3;
    const size_t lds = (size_t)rank_words * 4 + kp_alloc;
    int n_cu = 256;
    HD_CHECK(device_props(&n_cu, nullptr));
    const int per_cu = lds > 80 * 1024 ? 1 : (lds > 40 * 1024 ? 2 : 4);
    const int grid = (int)std::max<int64_t>(1, std::min<int64_t>((h->dim + 255) / 256, (int64_t)n_cu * per_cu * 4));
    const bool lz = b->rt.is_lz != 0;
#define HD_MF(C, L)                                                                                             \
    do {                                                                                                        \
        using T = typename VT<C>::T;                                                                            \
        HD_CHECK(set_lds((const void*)k_spmv_free<C, L>, lds));                                                 \
        hipLaunchKernelGGL((k_spmv_free<C, L>), dim3(grid), dim3(256), lds, stream, b->d_states, h->dim,        \
                           h->mf_K, h->mf_norb_mask, h->d_mf_kpair, h->mf_kp_bytes, b->d_rank, rank_words,      \
                           static_cast<const T*>(h->d_mf_pval), static_cast<const T*>(h->d_diag),              \
                           static_cast<const T*>(x), static_cast<T*>(y));                                      \
    } while (0)
    if (h->is_complex) {
        if (lz) HD_MF(true, true);
        else HD_MF(true, false);
    } else {
        if (lz) HD_MF(false, true);
        else HD_MF(false, false);
    }
#undef HD_MF
    HD_HIP(hipGetLastError());
    return HD_OK;
}

// The coded full-storage build of one batch sector from its merged tables.
static int build_batch_merged(const hd_basis* b, const Derived& M, double tol, hd_hamiltonian** out)
{
    hd_hamiltonian* h = new hd_hamiltonian();
    h->basis = b;
    h->dim = b->dim;
    h->storage = HD_STORAGE_FULL;
    if (hipGetDevice(&h->device) != hipSuccess ||
        hipStreamCreateWithFlags(&h->stream, hipStreamNonBlocking) != hipSuccess) {
        delete h;
        return fail(HD_ERR_HIP, "stream creation failed");
    }
    const int st = run_build<false>(h, b, M, tol, 4);
    if (st != HD_OK) {
        const std::string msg = hd_last_error();
        hd_ham_destroy(h);
        set_error(msg);
        return st;
    }
    *out = h;
    return HD_OK;
}

int build_batch_sector(const hd_basis* b, const hd_tables* const* tables, int n, double tol, BatchTables& bt,
                       hd_hamiltonian** out)
{
    *out = nullptr;
    const hd_tables* t0 = tables[0];
    for (int p = 0; p < n; ++p)
        if (tables[p]->split) return fail(HD_ERR_UNSUPPORTED, "two-level tables are not batched");
    const int K = t0->K;
    // the merged tables depend on the sector only through the hops it keeps
    std::vector<int> sums;
    const bool cacheable = b->rt.Norb == K && sector_pair_sums(b->rt, b->sector, sums) == HD_OK;
    if (cacheable)
        for (size_t i = 0; i < bt.merged_keys.size(); ++i)
            if (bt.merged_keys[i] == sums)
                return build_batch_merged(b, *static_cast<const Derived*>(bt.merged[i].get()), tol, out);
    std::vector<Derived> Ds((size_t)n);
    for (int p = 0; p < n; ++p) HD_CHECK(derive_tables(tables[p], b, tol, Ds[(size_t)p]));
    const size_t nk = Ds[0].kpair.size();
    Derived M = Ds[0];  // diagonal flags of table 0 only mark where the pattern holds a diagonal entry
    M.remap.assign(nk, (int16_t)-1);
    std::vector<double> tuple((size_t)n);
    std::vector<uint64_t> key((size_t)n);
    for (size_t i = 0; i < nk; ++i) {
        int d = -1;
        bool kept = false;
        for (int p = 0; p < n; ++p) {
            const uint8_t e = Ds[(size_t)p].kpair[i];
            if (e == 0xFF) {
                tuple[(size_t)p] = 0.0;
                continue;
            }
            if (d >= 0 && (e & 0x3F) != d)
                return fail(HD_ERR_UNSUPPORTED, "batch tables disagree on the hop structure");
            d = e & 0x3F;
            const bool k = !(e & 0x80);
            kept |= k;
            tuple[(size_t)p] = k ? Ds[(size_t)p].pval[i] : 0.0;
        }
        if (d < 0) {
            M.kpair[i] = 0xFF;
            continue;
        }
        M.kpair[i] = (uint8_t)(kept ? d : (d | 0x80));
        if (!kept) continue;
        for (int p = 0; p < n; ++p) std::memcpy(&key[(size_t)p], &tuple[(size_t)p], 8);
        auto it = bt.index.find(key);
        int code;
        if (it == bt.index.end()) {
            code = (int)bt.index.size();
            if (code >= 0x7FFF)
                return fail(HD_ERR_UNSUPPORTED, "more than 32766 distinct hop-value tuples in the batch");
            bt.index.emplace(key, code);
            bt.dict.insert(bt.dict.end(), tuple.begin(), tuple.end());
        } else {
            code = it->second;
        }
        M.remap[i] = (int16_t)code;
    }
    M.codable = true;
    if (bt.dterm.empty()) {
        // diagonal terms and E_k of every table, 0.0 where the reference skips
        // the term (|el| <= tol), point index fastest
        bt.K = K;
        bt.n = n;
        bt.has_eps = Ds[0].has_eps;
        bt.dterm.assign((size_t)K * K * 2 * n, 0.0);
        bt.eps.assign((size_t)K * n, 0.0);
        for (int p = 0; p < n; ++p) {
            const Derived& D = Ds[(size_t)p];
            for (size_t idx = 0; idx < (size_t)K * K; ++idx)
                for (int q = 0; q < 2; ++q)
                    if (D.dflag[idx] & (1 << q)) bt.dterm[(idx * 2 + q) * n + p] = D.dterm[idx * 2 + q];
            for (int k = 0; k < K; ++k)
                if (D.has_eps && D.epsflag[(size_t)k]) bt.eps[(size_t)k * n + p] = D.eps[(size_t)k];
        }
    }
    if (cacheable) {
        bt.merged_keys.push_back(sums);
        bt.merged.push_back(std::make_shared<Derived>(M));
    }
    return build_batch_merged(b, M, tol, out);
}

}  // namespace hd

using namespace hd;

extern "C" {

int hd_tables_create(int32_t layout, uint32_t kmax, const uint16_t* h_ktable, const double* h_vtable,
                     int32_t is_complex, const double* h_eps, hd_tables** out)
{
    if (!out || !h_ktable || !h_vtable) return fail(HD_ERR_ARG, "hd_tables_create: null argument");
    *out = nullptr;
    if (kmax == 0 || kmax > 64) return fail(HD_ERR_ARG, "kMax must be in 1..64");
    if (layout != HD_LAYOUT_SPHERE && layout != HD_LAYOUT_OFL)
        return fail(HD_ERR_ARG, "unknown table layout");
    const size_t K = kmax;
    const size_t n = layout == HD_LAYOUT_SPHERE ? K * K * K : (K + 1) * K * K / 2;
    hd_tables* t = new hd_tables();
    t->layout = layout;
    t->K = (int)kmax;
    t->is_complex = is_complex ? 1 : 0;
    t->ktab.assign(h_ktable, h_ktable + n);
    const size_t w = t->is_complex ? 2 : 1;
    t->vtab.assign(h_vtable, h_vtable + n * w);
    if (h_eps) t->eps.assign(h_eps, h_eps + K * w);
    *out = t;
    return HD_OK;
}

int hd_tables_create_two_level(uint32_t n1, const uint16_t* h_ktable1, const double* h_vtable1,
                               const double* h_eps1, const uint16_t* h_ktable2, const double* h_vtable2,
                               const double* h_eps2, hd_tables** out)
{
    if (!out || !h_ktable1 || !h_vtable1 || !h_ktable2 || !h_vtable2)
        return fail(HD_ERR_ARG, "hd_tables_create_two_level: null argument");
    *out = nullptr;
    if (n1 == 0 || 2 * n1 + 2 > 64) return fail(HD_ERR_ARG, "n1 must be in 1..31");
    const int K1 = (int)n1, K2 = (int)n1 + 2, K = K1 + K2;
    // combined orbital of level-local label k: level 0 (first) at K2 + k,
    // level 1 (second) at k
    auto comb = [&](int level, int k) { return level == 0 ? K2 + k : k; };
    hd_tables* t = new hd_tables();
    t->layout = HD_LAYOUT_SPHERE;
    t->K = K;
    t->is_complex = 0;
    t->split = K2;
    t->ktab.assign((size_t)K * K * K, (uint16_t)0xFFFF);
    t->vtab.assign((size_t)K * K * K, 0.0);
    for (int level = 0; level < 2; ++level) {
        const int L = level == 0 ? K1 : K2;
        const uint16_t* kt = level == 0 ? h_ktable1 : h_ktable2;
        const double* vt = level == 0 ? h_vtable1 : h_vtable2;
        for (int k4 = 0; k4 < L; ++k4)
            for (int k3 = 0; k3 < L; ++k3)
                for (int k2 = 0; k2 < L; ++k2) {
                    const size_t i = ((size_t)k4 * L + k3) * L + k2;
                    const size_t j = ((size_t)comb(level, k4) * K + comb(level, k3)) * K + comb(level, k2);
                    if (kt[i] < L) t->ktab[j] = (uint16_t)comb(level, kt[i]);  // GetK1: k1 >= kMax gives none
                    t->vtab[j] = vt[i];
                }
    }
    if (h_eps1 || h_eps2) {
        t->eps.assign((size_t)K, 0.0);
        for (int k = 0; k < K1 && h_eps1; ++k) t->eps[(size_t)comb(0, k)] = h_eps1[k];
        for (int k = 0; k < K2 && h_eps2; ++k) t->eps[(size_t)comb(1, k)] = h_eps2[k];
    }
    *out = t;
    return HD_OK;
}

int hd_tables_create_hash(uint32_t kmax, int32_t is_complex, int32_t max_k1, const uint8_t* h_nk1,
                          const uint16_t* h_k1, const double* h_v, int32_t max_q1, const uint8_t* h_nq1,
                          const uint16_t* h_q1, const double* h_e, hd_tables** out)
{
    if (!out || !h_nk1 || (max_k1 > 0 && (!h_k1 || !h_v)))
        return fail(HD_ERR_ARG, "hd_tables_create_hash: null argument");
    *out = nullptr;
    if (kmax == 0 || kmax > 64) return fail(HD_ERR_ARG, "kMax must be in 1..64");
    if (max_k1 < 0 || max_k1 > 255 || max_q1 < 0 || max_q1 > 255)
        return fail(HD_ERR_ARG, "list lengths must be in 0..255");
    if (h_nq1 && max_q1 > 0 && (!h_q1 || !h_e)) return fail(HD_ERR_ARG, "hd_tables_create_hash: null quadratic list");
    const size_t K = kmax, ne = K * K * K, w = is_complex ? 2 : 1;
    for (size_t e = 0; e < ne; ++e) {
        if (h_nk1[e] > max_k1) return fail(HD_ERR_ARG, "k1 list longer than max_k1 at entry " + std::to_string(e));
        for (int j = 0; j < h_nk1[e]; ++j)
            if (h_k1[e * (size_t)max_k1 + j] >= kmax)
                return fail(HD_ERR_ARG, "k1 label >= kMax at entry " + std::to_string(e));
    }
    if (h_nq1)
        for (size_t k = 0; k < K; ++k) {
            if (h_nq1[k] > max_q1) return fail(HD_ERR_ARG, "quadratic list longer than max_q1");
            for (int j = 0; j < h_nq1[k]; ++j)
                if (h_q1[k * (size_t)max_q1 + j] >= kmax) return fail(HD_ERR_ARG, "quadratic k1 label >= kMax");
        }
    hd_tables* t = new hd_tables();
    t->layout = HD_LAYOUT_HASH;
    t->K = (int)kmax;
    t->is_complex = is_complex ? 1 : 0;
    t->max_k1 = std::max(max_k1, 1);
    t->nk1.assign(h_nk1, h_nk1 + ne);
    t->k1.assign(ne * (size_t)t->max_k1, 0);
    t->hv.assign(ne * (size_t)t->max_k1 * w, 0.0);
    for (size_t e = 0; e < ne; ++e)
        for (int j = 0; j < h_nk1[e]; ++j) {
            const size_t i = e * (size_t)max_k1 + j, o = e * (size_t)t->max_k1 + j;
            t->k1[o] = (uint8_t)h_k1[i];
            for (size_t q = 0; q < w; ++q) t->hv[o * w + q] = h_v[i * w + q];
        }
    if (h_nq1) {
        t->has_quad = 1;
        t->max_q1 = std::max(max_q1, 1);
        t->nq1.assign(h_nq1, h_nq1 + K);
        t->q1.assign(K * (size_t)t->max_q1, 0);
        t->he.assign(K * (size_t)t->max_q1 * w, 0.0);
        for (size_t k = 0; k < K; ++k)
            for (int j = 0; j < h_nq1[k]; ++j) {
                const size_t i = k * (size_t)max_q1 + j, o = k * (size_t)t->max_q1 + j;
                t->q1[o] = (uint8_t)h_q1[i];
                for (size_t q = 0; q < w; ++q) t->he[o * w + q] = h_e[i * w + q];
            }
    }
    *out = t;
    return HD_OK;
}

int hd_tables_destroy(hd_tables* t)
{
    delete t;
    return HD_OK;
}

int hd_ham_build(const hd_basis* b, const hd_tables* t, int32_t storage, double zero_tol,
                 int32_t value_mask, hd_hamiltonian** out)
{
    if (!b || !t || !out) return fail(HD_ERR_ARG, "hd_ham_build: null argument");
    *out = nullptr;
    if (storage != HD_STORAGE_FULL && storage != HD_STORAGE_UPPER_REF && storage != HD_STORAGE_MATRIX_FREE)
        return fail(HD_ERR_ARG, "unknown storage");
    if (storage == HD_STORAGE_MATRIX_FREE && value_mask != 1)
        return fail(HD_ERR_ARG, "matrix-free storage computes in FP64 only (value_mask 1)");
    if (value_mask < 1 || value_mask > 7) return fail(HD_ERR_ARG, "value_mask must be in 1..7");
    if ((value_mask & 4) && storage != HD_STORAGE_FULL)
        return fail(HD_ERR_ARG, "the coded value format needs full storage");
    if (!(zero_tol >= 0.0)) return fail(HD_ERR_ARG, "zero_tol must be >= 0");
    if (t->K != b->Norb)
        return fail(HD_ERR_UNSUPPORTED, "kMax (" + std::to_string(t->K) + ") must equal the number of orbitals (" +
                                            std::to_string(b->Norb) + ")");
    hd_tables lists;  // array tables read as one-k1 lists, for a basis given as a state list
    if (b->sector.kind == HD_SECTOR_LIST && t->layout != HD_LAYOUT_HASH) {
        if (t->split) return fail(HD_ERR_UNSUPPORTED, "two-level tables need a two-level sector basis");
        array_to_lists(t, lists);
        t = &lists;
    }
    if (t->layout == HD_LAYOUT_HASH) {
        if (storage == HD_STORAGE_MATRIX_FREE)
            return fail(HD_ERR_UNSUPPORTED, "hash term tables: full or upper storage only");
        if (value_mask & 4)
            return fail(HD_ERR_UNSUPPORTED, "hash term tables: FP64/FP32 values only (no coded copy)");
        hd_hamiltonian* h = new hd_hamiltonian();
        h->basis = b;
        h->dim = b->dim;
        h->storage = storage;
        h->is_complex = t->is_complex;
        if (hipGetDevice(&h->device) != hipSuccess ||
            hipStreamCreateWithFlags(&h->stream, hipStreamNonBlocking) != hipSuccess) {
            delete h;
            return fail(HD_ERR_HIP, "stream creation failed");
        }
        const int st = general_build(h, b, t, zero_tol, value_mask);
        if (st != HD_OK) {
            const std::string msg = hd_last_error();
            hd_ham_destroy(h);
            set_error(msg);
            return st;
        }
        *out = h;
        return HD_OK;
    }
    Derived D;
    HD_CHECK(derive_tables(t, b, zero_tol, D));
    if ((value_mask & 4) && !D.codable)
        return fail(HD_ERR_UNSUPPORTED, "more than 32766 distinct values: the coded format does not apply");
    hd_hamiltonian* h = new hd_hamiltonian();
    h->basis = b;
    h->dim = b->dim;
    h->storage = storage;
    h->is_complex = t->is_complex;
    if (hipGetDevice(&h->device) != hipSuccess ||
        hipStreamCreateWithFlags(&h->stream, hipStreamNonBlocking) != hipSuccess) {
        delete h;
        return fail(HD_ERR_HIP, "stream creation failed");
    }
    const int st = t->is_complex ? run_build<true>(h, b, D, zero_tol, value_mask)
                                 : run_build<false>(h, b, D, zero_tol, value_mask);
    if (st != HD_OK) {
        const std::string msg = hd_last_error();
        hd_ham_destroy(h);
        set_error(msg);
        return st;
    }
    *out = h;
    return HD_OK;
}

int hd_ham_destroy(hd_hamiltonian* h)
{
    if (!h) return HD_OK;
    if (h->lz) lanczos_state_free(h->lz);
    dfree(h->d_rowptr);
    dfree(h->d_col);
    dfree(h->d_val64);
    dfree(h->d_val32);
    dfree(h->d_code);
    dfree(h->d_diag);
    dfree(h->d_dict);
    dfree(h->d_mf_kpair);
    dfree(h->d_mf_pval);
    dfree(h->d_mf_split);
    tiled_free(h);
    if (h->stream) (void)hipStreamDestroy(h->stream);
    delete h;
    return HD_OK;
}

int hd_ham_info(const hd_hamiltonian* h, hd_ham_info_t* info)
{
    if (!h || !info) return fail(HD_ERR_ARG, "hd_ham_info: null argument");
    info->dim = h->dim;
    info->nnz = h->nnz;
    info->storage = h->storage;
    info->is_complex = h->is_complex;
    info->has_f32 = h->d_val32 != nullptr;
    info->has_f64 = h->d_val64 != nullptr;
    info->n_trimmed = h->n_trimmed;
    info->max_row = h->max_row;
    info->t_count_ms = h->t_count;
    info->t_scan_ms = h->t_scan;
    info->t_fill_ms = h->t_fill;
    info->t_sort_ms = h->t_sort;
    info->device_bytes = h->bytes;
    return HD_OK;
}

int hd_ham_device_crs(const hd_hamiltonian* h, const int64_t** d_rowptr, const int32_t** d_col,
                      const void** d_val64, const void** d_val32)
{
    if (!h) return fail(HD_ERR_ARG, "hd_ham_device_crs: null handle");
    if (d_rowptr) *d_rowptr = h->d_rowptr;
    if (d_col) *d_col = h->d_col;
    if (d_val64) *d_val64 = h->d_val64;
    if (d_val32) *d_val32 = h->d_val32;
    return HD_OK;
}

int hd_ham_device_coded(const hd_hamiltonian* h, const uint16_t** d_code, const void** d_diag,
                        const void** d_dict, int32_t* ndict)
{
    if (!h) return fail(HD_ERR_ARG, "hd_ham_device_coded: null handle");
    if (d_code) *d_code = h->d_code;
    if (d_diag) *d_diag = h->d_diag;
    if (d_dict) *d_dict = h->d_dict;
    if (ndict) *ndict = h->ndict;
    return HD_OK;
}

int hd_ham_copy_crs(const hd_hamiltonian* h, int64_t row0, int64_t row1, int32_t precision,
                    int64_t* h_rowptr, int32_t* h_col, void* h_val)
{
    if (!h || row0 < 0 || row1 > h->dim || row0 > row1 || !h_rowptr)
        return fail(HD_ERR_ARG, "hd_ham_copy_crs: bad row range");
    if (!h->d_rowptr) return fail(HD_ERR_ARG, "hd_ham_copy_crs: matrix-free storage holds no CRS");
    HD_HIP(hipMemcpy(h_rowptr, h->d_rowptr + row0, (size_t)(row1 - row0 + 1) * 8, hipMemcpyDeviceToHost));
    const int64_t base = h_rowptr[0];
    const int64_t n = h_rowptr[row1 - row0] - base;
    for (int64_t i = 0; i <= row1 - row0; ++i) h_rowptr[i] -= base;
    if (h_col && n > 0)
        HD_HIP(hipMemcpy(h_col, h->d_col + base, (size_t)n * 4, hipMemcpyDeviceToHost));
    if (h_val && n > 0) {
        const void* src = precision == HD_F32 ? h->d_val32 : h->d_val64;
        if (!src) return fail(HD_ERR_ARG, "requested value precision is not resident");
        const size_t es = (precision == HD_F32 ? 4 : 8) * (h->is_complex ? 2 : 1);
        HD_HIP(hipMemcpy(h_val, static_cast<const char*>(src) + base * es, (size_t)n * es,
                         hipMemcpyDeviceToHost));
    }
    return HD_OK;
}

}  // extern "C"
