// hd_obs.hip -- observables of eigenvectors on the device.
//
// Reference: struct Observables (src/observables/fermion_observables.hpp:38-626).
// Each observable is a weighted sum over the sector's states with weight
// p_i = |psi_i|^2:
//   density            <n_k>          = sum_i p_i n_k(s_i)                (:47-106)
//   density-density    <n_k n_q>      = sum_i p_i n_k(s_i) n_q(s_i)       (:115-198)
//   translational      S(q)           = 1/N sum_i p_i c_q(s_i)             (:216-317)
//   rotational         R(m)           = 1/(6N) sum_alpha C_alpha e^{i pi m alpha/3} (:334-428)
//   participation      1 / sum_i p_i^2                                      (:528-563)
//   most probable      the n largest |psi_i| with their states              (:437-517)
// c_q(s) and C_alpha(s) are integers per state (the reference's operator
// loops evaluated on one Fock state).  Every reduction is deterministic: a
// fixed grid walks 256-state chunks; the chunk's states and weights sit in
// LDS and each thread owns a set of bins, adding the chunk's contributions in
// state order; per-block partials are then summed per bin in block order.
#include <hip/hip_runtime.h>

#include <algorithm>
#include <cmath>
#include <cstring>
#include <vector>

#include "hd_internal.hpp"

namespace hd {

namespace {

constexpr int kOB = 256;        // states per chunk = threads per block
constexpr int kMaxBins = 4096;  // Norb^2 for Norb <= 64
constexpr int kBinsPerThread = kMaxBins / kOB;
constexpr int kMaxQ = 64;       // translations per call
constexpr int kRot = 6;

enum ObsKind { kDensity = 0, kDensityDensity = 1, kParticipation = 2, kTranslational = 3, kRotational = 4 };

struct ObsArgs {
    int64_t dim;
    const uint64_t* states;
    const double* vec;  // real, or interleaved complex
    int cplx;
    int norb, nbins, nq;
    const uint16_t* add;  // [norb * nq]: StateAddition(k, q)
    const uint16_t* sub;  // [norb * nq]: StateSubtraction(k, q)
    const uint16_t* inv;  // [norb * 6]: InverseR60(k, alpha)
    double* partial;      // [gridDim.x * nbins]
};

// Evaluate_CdC (src/utilities/mathematics/fermions.hpp:155-225): c+_{k1} c_{k2}
// on `state`; returns the sign (0 when the result vanishes) and the out-state.
__device__ __forceinline__ int cdc(uint64_t k1bit, uint64_t k2bit, uint64_t state, uint64_t& out)
{
    if ((state & k2bit) == 0) return 0;
    if (k1bit & k2bit) {
        out = state;
        return 1;
    }
    if (state & k1bit) return 0;
    int n = __popcll(state & (k2bit - 1));
    const uint64_t mid = state ^ k2bit;
    n += __popcll(mid & (k1bit - 1));
    out = mid | k1bit;
    return (n & 1) ? -1 : 1;
}

template <int KIND>
__global__ void __launch_bounds__(kOB) k_obs(ObsArgs a)
{
    __shared__ uint64_t ss[kOB];
    __shared__ double sp[kOB];
    __shared__ int16_t cnt[KIND == kTranslational ? kOB * kMaxQ : KIND == kRotational ? kOB * kRot : 1];
    double acc[kBinsPerThread];
#pragma unroll
    for (int k = 0; k < kBinsPerThread; ++k) acc[k] = 0.0;
    const int64_t nchunks = (a.dim + kOB - 1) / kOB;
    for (int64_t ch = blockIdx.x; ch < nchunks; ch += gridDim.x) {
        const int64_t i = ch * kOB + threadIdx.x;
        uint64_t s = 0;
        double p = 0.0;
        if (i < a.dim) {
            s = a.states[i];
            if (a.cplx) {
                const double re = a.vec[2 * i], im = a.vec[2 * i + 1];
                p = re * re + im * im;
            } else {
                p = a.vec[i] * a.vec[i];
            }
        }
        ss[threadIdx.x] = s;
        sp[threadIdx.x] = p;
        if constexpr (KIND == kTranslational) {
            // c_q(s) = sum over k1 occupied, k2 occupied after c+_{k1+q} c_{k1},
            // of the two signs when c+_{k2-q} c_{k2} restores s
            for (int q = 0; q < a.nq; ++q) {
                int c = 0;
                for (uint64_t t1 = s; t1; t1 &= t1 - 1) {
                    const int k1 = __ffsll((unsigned long long)t1) - 1;
                    uint64_t out1 = 0;
                    const int s1 = cdc(1ull << a.add[k1 * a.nq + q], 1ull << k1, s, out1);
                    if (!s1) continue;
                    for (uint64_t t2 = out1; t2; t2 &= t2 - 1) {
                        const int k2 = __ffsll((unsigned long long)t2) - 1;
                        uint64_t out2 = 0;
                        const int s2 = cdc(1ull << a.sub[k2 * a.nq + q], 1ull << k2, out1, out2);
                        if (s2 && out2 == s) c += s1 * s2;
                    }
                }
                cnt[threadIdx.x * kMaxQ + q] = (int16_t)c;
            }
        }
        if constexpr (KIND == kRotational) {
            // C_alpha(s) = #{(k1, k2) both occupied : InverseR60(k2, alpha) == k1},
            // counted only where p > 1e-11 as in the reference
            for (int al = 0; al < kRot; ++al) {
                int c = 0;
                if (p > 1e-11)
                    for (uint64_t t2 = s; t2; t2 &= t2 - 1) {
                        const int k2 = __ffsll((unsigned long long)t2) - 1;
                        c += (int)((s >> a.inv[k2 * kRot + al]) & 1ull);
                    }
                cnt[threadIdx.x * kRot + al] = (int16_t)c;
            }
        }
        __syncthreads();
        const int n_here = (int)min<int64_t>(kOB, a.dim - ch * kOB);
#pragma unroll
        for (int k = 0; k < kBinsPerThread; ++k) {
            const int bin = threadIdx.x + k * kOB;
            if (bin >= a.nbins) break;
            double v = acc[k];
            if constexpr (KIND == kDensity) {
                for (int j = 0; j < n_here; ++j)
                    if ((ss[j] >> bin) & 1ull) v += sp[j];
            } else if constexpr (KIND == kDensityDensity) {
                const int kk = bin / a.norb, qq = bin % a.norb;
                for (int j = 0; j < n_here; ++j)
                    if ((ss[j] >> kk) & (ss[j] >> qq) & 1ull) v += sp[j];
            } else if constexpr (KIND == kParticipation) {
                for (int j = 0; j < n_here; ++j) v += sp[j] * sp[j];
            } else if constexpr (KIND == kTranslational) {
                for (int j = 0; j < n_here; ++j) v += sp[j] * (double)cnt[j * kMaxQ + bin];
            } else {
                for (int j = 0; j < n_here; ++j) v += sp[j] * (double)cnt[j * kRot + bin];
            }
            acc[k] = v;
        }
        __syncthreads();
    }
#pragma unroll
    for (int k = 0; k < kBinsPerThread; ++k) {
        const int bin = threadIdx.x + k * kOB;
        if (bin < a.nbins) a.partial[(int64_t)blockIdx.x * a.nbins + bin] = acc[k];
    }
}

__global__ void __launch_bounds__(kOB) k_obs_reduce(const double* __restrict__ partial, int nblocks, int nbins,
                                                    double* __restrict__ out)
{
    const int bin = blockIdx.x * kOB + threadIdx.x;
    if (bin >= nbins) return;
    double s = 0.0;
    for (int b = 0; b < nblocks; ++b) s += partial[(int64_t)b * nbins + bin];
    out[bin] = s;
}

int run_obs(int kind, const hd_basis* b, int is_complex, const void* d_vec, int nq, const uint16_t* h_add,
            const uint16_t* h_sub, const uint16_t* h_inv, int nbins, double* h_out)
{
    if (!b || !d_vec || !h_out) return fail(HD_ERR_ARG, "observable: null argument");
    if (b->dim == 0) {
        std::fill(h_out, h_out + nbins, 0.0);
        return HD_OK;
    }
    if (nbins > kMaxBins) return fail(HD_ERR_UNSUPPORTED, "observable: too many bins");
    const hipStream_t st = b->stream;
    const int nblocks = (int)std::min<int64_t>((b->dim + kOB - 1) / kOB, 1024);
    double *d_part = nullptr, *d_out = nullptr;
    uint16_t* d_tab = nullptr;
    const size_t tab_n = kind == kTranslational ? 2 * (size_t)b->Norb * nq : kind == kRotational ? (size_t)b->Norb * kRot : 0;
    int rc = dalloc_t(&d_part, (size_t)nblocks * nbins, "observable partials");
    if (rc == HD_OK) rc = dalloc_t(&d_out, (size_t)nbins, "observable output");
    if (rc == HD_OK && tab_n) rc = dalloc_t(&d_tab, tab_n, "observable tables");
    HipErrors q;
    if (rc == HD_OK) {
        ObsArgs a{};
        a.dim = b->dim;
        a.states = b->d_states;
        a.vec = static_cast<const double*>(d_vec);
        a.cplx = is_complex ? 1 : 0;
        a.norb = b->Norb;
        a.nbins = nbins;
        a.nq = nq;
        a.partial = d_part;
        if (kind == kTranslational) {
            HD_ENQ(q, hipMemcpyAsync(d_tab, h_add, sizeof(uint16_t) * b->Norb * nq, hipMemcpyHostToDevice, st));
            HD_ENQ(q, hipMemcpyAsync(d_tab + (size_t)b->Norb * nq, h_sub, sizeof(uint16_t) * b->Norb * nq,
                                     hipMemcpyHostToDevice, st));
            a.add = d_tab;
            a.sub = d_tab + (size_t)b->Norb * nq;
        } else if (kind == kRotational) {
            HD_ENQ(q, hipMemcpyAsync(d_tab, h_inv, sizeof(uint16_t) * b->Norb * kRot, hipMemcpyHostToDevice, st));
            a.inv = d_tab;
        }
        switch (kind) {
        case kDensity: hipLaunchKernelGGL(k_obs<kDensity>, dim3(nblocks), dim3(kOB), 0, st, a); break;
        case kDensityDensity: hipLaunchKernelGGL(k_obs<kDensityDensity>, dim3(nblocks), dim3(kOB), 0, st, a); break;
        case kParticipation: hipLaunchKernelGGL(k_obs<kParticipation>, dim3(nblocks), dim3(kOB), 0, st, a); break;
        case kTranslational: hipLaunchKernelGGL(k_obs<kTranslational>, dim3(nblocks), dim3(kOB), 0, st, a); break;
        default: hipLaunchKernelGGL(k_obs<kRotational>, dim3(nblocks), dim3(kOB), 0, st, a); break;
        }
        HD_ENQ(q, hipGetLastError());
        hipLaunchKernelGGL(k_obs_reduce, dim3((nbins + kOB - 1) / kOB), dim3(kOB), 0, st, d_part, nblocks, nbins, d_out);
        HD_ENQ(q, hipGetLastError());
        HD_ENQ(q, hipMemcpyAsync(h_out, d_out, sizeof(double) * nbins, hipMemcpyDeviceToHost, st));
        HD_ENQ(q, hipStreamSynchronize(st));
        rc = q.status();
    }
    dfree(d_part);
    dfree(d_out);
    dfree(d_tab);
    return rc;
}

// Top-n selection: keys = |psi_i|^2 as ordered 64-bit patterns (non-negative
// doubles order like their bits).  A bitwise search finds the n-th largest key
// t with counting passes; the keys above t and then the first ties at t (in
// ascending index order) are compacted, exactly n of them, and sorted on one
// block.
__global__ void __launch_bounds__(kOB) k_key_count(int64_t dim, const double* __restrict__ vec, int cplx,
                                                   uint64_t thresh, unsigned long long* __restrict__ count)
{
    __shared__ unsigned long long red[kOB];
    unsigned long long c = 0;
    for (int64_t i = (int64_t)blockIdx.x * kOB + threadIdx.x; i < dim; i += (int64_t)gridDim.x * kOB) {
        const double p = cplx ? vec[2 * i] * vec[2 * i] + vec[2 * i + 1] * vec[2 * i + 1] : vec[i] * vec[i];
        c += (uint64_t)__double_as_longlong(p) >= thresh ? 1ull : 0ull;
    }
    red[threadIdx.x] = c;
    __syncthreads();
    for (int o = kOB / 2; o > 0; o >>= 1) {
        if ((int)threadIdx.x < o) red[threadIdx.x] += red[threadIdx.x + o];
        __syncthreads();
    }
    if (threadIdx.x == 0) atomicAdd(count, red[0]);
}

__device__ __forceinline__ uint64_t key_of(const double* vec, int cplx, int64_t i)
{
    const double p = cplx ? vec[2 * i] * vec[2 * i] + vec[2 * i + 1] * vec[2 * i + 1] : vec[i] * vec[i];
    return (uint64_t)__double_as_longlong(p);
}

// Entries with key > thresh (fewer than n of them) are appended in any order.
__global__ void __launch_bounds__(kOB) k_key_above(int64_t dim, const double* __restrict__ vec, int cplx,
                                                   uint64_t thresh, int64_t* __restrict__ out, int cap,
                                                   unsigned int* __restrict__ n_out)
{
    for (int64_t i = (int64_t)blockIdx.x * kOB + threadIdx.x; i < dim; i += (int64_t)gridDim.x * kOB)
        if (key_of(vec, cplx, i) > thresh) {
            const unsigned int s = atomicAdd(n_out, 1u);
            if ((int)s < cap) out[s] = i;
        }
}

// Ties (key == thresh), block b over the contiguous range [b*len, (b+1)*len).
__global__ void __launch_bounds__(kOB) k_tie_count(int64_t dim, const double* __restrict__ vec, int cplx,
                                                   uint64_t thresh, int64_t len, unsigned int* __restrict__ counts)
{
    __shared__ unsigned int red[kOB];
    const int64_t lo = (int64_t)blockIdx.x * len, hi = min(lo + len, dim);
    unsigned int c = 0;
    for (int64_t i = lo + threadIdx.x; i < hi; i += kOB) c += key_of(vec, cplx, i) == thresh ? 1u : 0u;
    red[threadIdx.x] = c;
    __syncthreads();
    for (int o = kOB / 2; o > 0; o >>= 1) {
        if ((int)threadIdx.x < o) red[threadIdx.x] += red[threadIdx.x + o];
        __syncthreads();
    }
    if (threadIdx.x == 0) counts[blockIdx.x] = red[0];
}

// Writes the ties of global rank < want (ranks in ascending index order) to
// out[base + rank]; first[b] = ties in blocks before b.
__global__ void __launch_bounds__(kOB) k_tie_write(int64_t dim, const double* __restrict__ vec, int cplx,
                                                   uint64_t thresh, int64_t len,
                                                   const unsigned int* __restrict__ first, unsigned int want,
                                                   int64_t* __restrict__ out, unsigned int base)
{
    __shared__ unsigned int scan[kOB];
    const int64_t lo = (int64_t)blockIdx.x * len, hi = min(lo + len, dim);
    unsigned int run = first[blockIdx.x];
    for (int64_t i0 = lo; i0 < hi && run < want; i0 += kOB) {
        const int64_t i = i0 + threadIdx.x;
        const unsigned int f = (i < hi && key_of(vec, cplx, i) == thresh) ? 1u : 0u;
        scan[threadIdx.x] = f;
        __syncthreads();
        for (int o = 1; o < kOB; o <<= 1) {  // inclusive Hillis-Steele scan
            const unsigned int v = (int)threadIdx.x >= o ? scan[threadIdx.x - o] : 0u;
            __syncthreads();
            scan[threadIdx.x] += v;
            __syncthreads();
        }
        const unsigned int rank = run + scan[threadIdx.x] - f;
        if (f && rank < want) out[base + rank] = i;
        const unsigned int total = scan[kOB - 1];
        __syncthreads();
        run += total;
    }
}

// One block: sort the collected indices by (key descending, index ascending)
// with a bitonic sort in LDS, then write the first n states and amplitudes.
__global__ void __launch_bounds__(1024) k_top_sort(const int64_t* __restrict__ idx, int m, int n,
                                                   const double* __restrict__ vec, int cplx,
                                                   const uint64_t* __restrict__ states, uint64_t* __restrict__ o_states,
                                                   double* __restrict__ o_amp)
{
    constexpr int kCap = 4096;
    __shared__ uint64_t key[kCap];
    __shared__ int64_t ix[kCap];
    int P = 1;
    while (P < m) P <<= 1;
    for (int i = threadIdx.x; i < P; i += 1024) {
        if (i < m) {
            const int64_t r = idx[i];
            const double p = cplx ? vec[2 * r] * vec[2 * r] + vec[2 * r + 1] * vec[2 * r + 1] : vec[r] * vec[r];
            key[i] = (uint64_t)__double_as_longlong(p);
            ix[i] = r;
        } else {
            key[i] = 0;
            ix[i] = INT64_MAX;
        }
    }
    __syncthreads();
    // "before" = larger key, then smaller index
    for (int k = 2; k <= P; k <<= 1)
        for (int j = k >> 1; j > 0; j >>= 1) {
            for (int i = threadIdx.x; i < P; i += 1024) {
                const int l = i ^ j;
                if (l > i) {
                    const bool i_first = key[i] > key[l] || (key[i] == key[l] && ix[i] < ix[l]);
                    const bool up = (i & k) == 0;
                    if (i_first != up) {
                        const uint64_t tk = key[i];
                        key[i] = key[l];
                        key[l] = tk;
                        const int64_t ti = ix[i];
                        ix[i] = ix[l];
                        ix[l] = ti;
                    }
                }
            }
            __syncthreads();
        }
    for (int i = threadIdx.x; i < n; i += 1024) {
        const int64_t r = ix[i];
        o_states[i] = states[r];
        if (cplx) {
            o_amp[2 * i] = vec[2 * r];
            o_amp[2 * i + 1] = vec[2 * r + 1];
        } else {
            o_amp[i] = vec[r];
        }
    }
}

}  // namespace

}  // namespace hd

using namespace hd;

extern "C" {

int hd_obs_density(const hd_basis* b, int32_t is_complex, const void* d_vec, double* h_out)
{
    if (!b) return fail(HD_ERR_ARG, "hd_obs_density: null basis");
    return run_obs(kDensity, b, is_complex, d_vec, 0, nullptr, nullptr, nullptr, b->Norb, h_out);
}

int hd_obs_density_density(const hd_basis* b, int32_t is_complex, const void* d_vec, double* h_out)
{
    if (!b) return fail(HD_ERR_ARG, "hd_obs_density_density: null basis");
    return run_obs(kDensityDensity, b, is_complex, d_vec, 0, nullptr, nullptr, nullptr, b->Norb * b->Norb, h_out);
}

int hd_obs_participation_ratio(const hd_basis* b, int32_t is_complex, const void* d_vec, double* h_out)
{
    if (!b || !h_out) return fail(HD_ERR_ARG, "hd_obs_participation_ratio: null argument");
    double s = 0.0;
    HD_CHECK(run_obs(kParticipation, b, is_complex, d_vec, 0, nullptr, nullptr, nullptr, 1, &s));
    *h_out = 1.0 / s;
    return HD_OK;
}

int hd_obs_translational_dd(const hd_basis* b, int32_t is_complex, const void* d_vec, int32_t nq,
                            const uint16_t* h_add, const uint16_t* h_sub, double* h_out)
{
    if (!b || !h_add || !h_sub) return fail(HD_ERR_ARG, "hd_obs_translational_dd: null argument");
    if (nq <= 0 || nq > kMaxQ) return fail(HD_ERR_ARG, "hd_obs_translational_dd: nq must be in 1..64");
    for (int64_t i = 0; i < (int64_t)b->Norb * nq; ++i)
        if (h_add[i] >= b->Norb || h_sub[i] >= b->Norb)
            return fail(HD_ERR_ARG, "hd_obs_translational_dd: table entry outside the orbitals");
    if ((int64_t)b->N * b->N > 32767) return fail(HD_ERR_UNSUPPORTED, "too many particles for the 16-bit counts");
    HD_CHECK(run_obs(kTranslational, b, is_complex, d_vec, nq, h_add, h_sub, nullptr, nq, h_out));
    for (int q = 0; q < nq; ++q) h_out[q] /= b->N;
    return HD_OK;
}

int hd_obs_rotational_dd(const hd_basis* b, int32_t is_complex, const void* d_vec, const uint16_t* h_inv_r60,
                         double* h_out)
{
    if (!b || !h_inv_r60 || !h_out) return fail(HD_ERR_ARG, "hd_obs_rotational_dd: null argument");
    for (int64_t i = 0; i < (int64_t)b->Norb * kRot; ++i)
        if (h_inv_r60[i] >= b->Norb) return fail(HD_ERR_ARG, "hd_obs_rotational_dd: table entry outside the orbitals");
    double c[kRot];
    HD_CHECK(run_obs(kRotational, b, is_complex, d_vec, 0, nullptr, nullptr, h_inv_r60, kRot, c));
    // R(m) = 1/(6N) sum_alpha C_alpha exp(i pi m alpha / 3)
    const double pi = 3.14159265358979323846;
    for (int m = 0; m < kRot; ++m) {
        double re = 0.0, im = 0.0;
        for (int al = 0; al < kRot; ++al) {
            const double ph = pi / 3.0 * (double)(m * al);
            re += c[al] * std::cos(ph);
            im += c[al] * std::sin(ph);
        }
        h_out[2 * m] = re / (kRot * b->N);
        h_out[2 * m + 1] = im / (kRot * b->N);
    }
    return HD_OK;
}

int hd_obs_most_probable(const hd_basis* b, int32_t is_complex, const void* d_vec, int32_t n, uint64_t* h_states,
                         void* h_amplitudes)
{
    if (!b || !d_vec || !h_states || !h_amplitudes) return fail(HD_ERR_ARG, "hd_obs_most_probable: null argument");
    if (n <= 0 || n > 1024) return fail(HD_ERR_ARG, "hd_obs_most_probable: n must be in 1..1024");
    n = (int)std::min<int64_t>(n, b->dim);
    if (n == 0) return HD_OK;
    const hipStream_t st = b->stream;
    const double* v = static_cast<const double*>(d_vec);
    const int grid = (int)std::min<int64_t>((b->dim + kOB - 1) / kOB, 1024);
    unsigned long long* d_cnt = nullptr;
    int64_t* d_idx = nullptr;
    uint64_t* d_ost = nullptr;
    double* d_oamp = nullptr;
    constexpr int kCap = 4096;
    int rc = dalloc_t(&d_cnt, 2, "top-n counter");
    if (rc == HD_OK) rc = dalloc_t(&d_idx, kCap, "top-n indices");
    if (rc == HD_OK) rc = dalloc_t(&d_ost, (size_t)n, "top-n states");
    if (rc == HD_OK) rc = dalloc_t(&d_oamp, (size_t)n * (is_complex ? 2 : 1), "top-n amplitudes");
    HipErrors q;
    // the n-th largest key t: the largest t with count(key >= t) >= n, bit by bit
    uint64_t t = 0;
    for (int bit = 62; bit >= 0 && rc == HD_OK; --bit) {
        const uint64_t cand = t | (1ull << bit);
        unsigned long long c = 0;
        HD_ENQ(q, hipMemsetAsync(d_cnt, 0, sizeof(unsigned long long), st));
        hipLaunchKernelGGL(k_key_count, dim3(grid), dim3(kOB), 0, st, b->dim, v, is_complex ? 1 : 0, cand, d_cnt);
        HD_ENQ(q, hipGetLastError());
        HD_ENQ(q, hipMemcpyAsync(&c, d_cnt, sizeof(c), hipMemcpyDeviceToHost, st));
        HD_ENQ(q, hipStreamSynchronize(st));
        rc = q.status();
        if (rc == HD_OK && c >= (unsigned long long)n) t = cand;
    }
    // keys > t: fewer than n; then the first (n - that many) ties at t
    unsigned int m = 0;
    const int64_t len = (b->dim + grid - 1) / grid;
    std::vector<unsigned int> tie_counts(grid), first(grid);
    unsigned int* d_ties = nullptr;
    if (rc == HD_OK) rc = dalloc_t(&d_ties, 2 * (size_t)grid, "top-n tie counts");
    if (rc == HD_OK) {
        HD_ENQ(q, hipMemsetAsync(d_cnt, 0, sizeof(unsigned long long), st));
        hipLaunchKernelGGL(k_key_above, dim3(grid), dim3(kOB), 0, st, b->dim, v, is_complex ? 1 : 0, t, d_idx, kCap,
                           reinterpret_cast<unsigned int*>(d_cnt));
        hipLaunchKernelGGL(k_tie_count, dim3(grid), dim3(kOB), 0, st, b->dim, v, is_complex ? 1 : 0, t, len, d_ties);
        HD_ENQ(q, hipGetLastError());
        HD_ENQ(q, hipMemcpyAsync(&m, d_cnt, sizeof(m), hipMemcpyDeviceToHost, st));
        HD_ENQ(q, hipMemcpyAsync(tie_counts.data(), d_ties, sizeof(unsigned int) * grid, hipMemcpyDeviceToHost, st));
        HD_ENQ(q, hipStreamSynchronize(st));
        rc = q.status();
        if (rc == HD_OK && m >= (unsigned int)n) rc = fail(HD_ERR_HIP, "hd_obs_most_probable: inconsistent counts");
    }
    if (rc == HD_OK) {
        unsigned int run = 0;
        for (int g = 0; g < grid; ++g) {
            first[g] = run;
            run += tie_counts[g];
        }
        if (m + run < (unsigned int)n) rc = fail(HD_ERR_HIP, "hd_obs_most_probable: too few ties");
    }
    if (rc == HD_OK) {
        HD_ENQ(q, hipMemcpyAsync(d_ties + grid, first.data(), sizeof(unsigned int) * grid, hipMemcpyHostToDevice, st));
        hipLaunchKernelGGL(k_tie_write, dim3(grid), dim3(kOB), 0, st, b->dim, v, is_complex ? 1 : 0, t, len,
                           d_ties + grid, (unsigned int)n - m, d_idx, m);
        HD_ENQ(q, hipGetLastError());
        m = (unsigned int)n;
    }
    dfree(d_ties);
    if (rc == HD_OK) {
        hipLaunchKernelGGL(k_top_sort, dim3(1), dim3(1024), 0, st, d_idx, (int)m, n, v, is_complex ? 1 : 0,
                           b->d_states, d_ost, d_oamp);
        HD_ENQ(q, hipGetLastError());
        HD_ENQ(q, hipMemcpyAsync(h_states, d_ost, sizeof(uint64_t) * n, hipMemcpyDeviceToHost, st));
        HD_ENQ(q, hipMemcpyAsync(h_amplitudes, d_oamp, sizeof(double) * n * (is_complex ? 2 : 1),
                                 hipMemcpyDeviceToHost, st));
        HD_ENQ(q, hipStreamSynchronize(st));
        rc = q.status();
    }
    dfree(d_cnt);
    dfree(d_idx);
    dfree(d_ost);
    dfree(d_oamp);
    return rc;
}

}  // extern "C"
