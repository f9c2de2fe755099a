// hd_block.hip -- block thick-restart Lanczos for one real Hamiltonian.
//
// Reference: ArpackWrapper::DiagonalizeSymmetricMatrix (src/utilities/wrappers/
// arpack_wrapper.hpp:595-937) iterates a single Krylov vector.  On MI355X the
// single-vector SpMV is bound by its x gathers: every stored entry pulls an
// x line into L2 for 8 useful bytes (measured L2-miss traffic 4x the
// algorithmic bytes).  A block of B vectors stored row-major, X[row][B],
// turns each gather into B contiguous doubles, so one line serves B matvecs.
//
// Method: block Lanczos with full reorthogonalisation (classical Gram-Schmidt
// applied twice against the whole basis), each new block orthonormalised by
// Cholesky QR applied twice, and thick restart by Rayleigh-Ritz on the
// projected matrix (a block form of the Krylov-Schur restart, Stewart, SIAM
// J. Matrix Anal. Appl. 23 (2001) 601): keep q*B Ritz vectors, append the
// residual block, continue.  The projected matrix is assembled on the host
// from the Gram-Schmidt coefficients and the QR factors and diagonalised by
// cyclic Jacobi.  Residuals are ||R_m z_i||, checked at the end against the
// FP64 operator with the single-vector SpMV.
#include <hip/hip_runtime.h>

#include <algorithm>
#include <chrono>
#include <cmath>
#include <vector>

#include "hd_internal.hpp"

namespace hd {

namespace {

constexpr int kBT = 256;           // threads of the block vector kernels
constexpr int64_t kBChunk = 8192;  // rows per reduction chunk

HD_DEV uint64_t bmix(uint64_t z)
{
    z += 0x9E3779B97F4A7C15ull;
    z = (z ^ (z >> 30)) * 0xBF58476D1CE4E5B9ull;
    z = (z ^ (z >> 27)) * 0x94D049BB133111EBull;
    return z ^ (z >> 31);
}

// Y = H X for X, Y [row][B].  LPR lanes share a row: each walks every LPR-th
// entry, gathers the B doubles of x at the entry's column (16-byte loads) and
// keeps B partial sums, reduced across the row's lanes at the end.  Coded
// values: sign * dict[code & 0x7FFF] from LDS, code 0xFFFF = diag[row].
template <int B, int LPR, bool CODED>
__global__ void __launch_bounds__(256)
k_spmm_blk(int64_t dim, const int64_t* __restrict__ rowptr, const int32_t* __restrict__ col,
           const double* __restrict__ val, const uint16_t* __restrict__ code, const double* __restrict__ dict,
           int ndict, const double* __restrict__ diag, const double* __restrict__ x, double* __restrict__ y)
{
    extern __shared__ double s_dict[];
    if (CODED) {
        for (int i = threadIdx.x; i < ndict; i += blockDim.x) s_dict[i] = dict[i];
        __syncthreads();
    }
    const int lane = threadIdx.x & (LPR - 1);
    const int64_t rows_per_block = blockDim.x / LPR;
    for (int64_t r = (int64_t)blockIdx.x * rows_per_block + threadIdx.x / LPR; r < dim;
         r += (int64_t)gridDim.x * rows_per_block) {
        double acc[B];
#pragma unroll
        for (int q = 0; q < B; ++q) acc[q] = 0.0;
        const int64_t e1 = rowptr[r + 1];
        for (int64_t e = rowptr[r] + lane; e < e1; e += LPR) {
            const int c = __builtin_nontemporal_load(col + e);
            double v;
            if (CODED) {
                const unsigned cd = __builtin_nontemporal_load(code + e);
                const bool dg = cd == 0xFFFFu;
                v = dg ? diag[r] : s_dict[cd & 0x7FFFu];
                v = (!dg && (cd & 0x8000u)) ? -v : v;
            } else {
                v = __builtin_nontemporal_load(val + e);
            }
            const double2* xp = reinterpret_cast<const double2*>(x + (size_t)c * B);
#pragma unroll
            for (int q = 0; q < B / 2; ++q) {
                const double2 xv = xp[q];
                acc[2 * q] = fma(v, xv.x, acc[2 * q]);
                acc[2 * q + 1] = fma(v, xv.y, acc[2 * q + 1]);
            }
        }
#pragma unroll
        for (int o = LPR / 2; o > 0; o >>= 1)
#pragma unroll
            for (int q = 0; q < B; ++q) acc[q] += __shfl_xor(acc[q], o, LPR);
#pragma unroll
        for (int q = 0; q < B; ++q)
            if (q % LPR == lane) y[(size_t)r * B + q] = acc[q];
    }
}

// partial[(c*nb + i)*B*B + a*B + d] = sum over rows of chunk c of V_i[r][a] W[r][d]
template <int B>
__global__ void __launch_bounds__(kBT)
k_blk_dots(int nb, const double* __restrict__ V0, int64_t bstride, const double* __restrict__ W, int64_t n,
           double* __restrict__ partial)
{
    __shared__ double red[kBT * B];
    const int c = blockIdx.x, i = blockIdx.y;
    const int d = threadIdx.x % B, rl = threadIdx.x / B;
    constexpr int RL = kBT / B;
    const int64_t r0 = (int64_t)c * kBChunk, r1 = min(n, r0 + kBChunk);
    const double* Vi = V0 + (size_t)i * bstride;
    double acc[B];
#pragma unroll
    for (int a = 0; a < B; ++a) acc[a] = 0.0;
    for (int64_t r = r0 + rl; r < r1; r += RL) {
        const double w = W[(size_t)r * B + d];
        const double* vr = Vi + (size_t)r * B;
#pragma unroll
        for (int a = 0; a < B; ++a) acc[a] = fma(vr[a], w, acc[a]);
    }
#pragma unroll
    for (int a = 0; a < B; ++a) red[threadIdx.x * B + a] = acc[a];
    __syncthreads();
    if (threadIdx.x < B * B) {
        const int a = threadIdx.x / B, dd = threadIdx.x % B;
        double s = 0.0;
        for (int q = 0; q < RL; ++q) s += red[(q * B + dd) * B + a];
        partial[((size_t)c * nb + i) * B * B + a * B + dd] = s;
    }
}

// out[i*B*B + e] = sum over chunks, in chunk order
__global__ void k_blk_reduce(int nchunks, int nb, int bb, const double* __restrict__ partial, double* __restrict__ out)
{
    const int i = blockIdx.x;
    for (int e = threadIdx.x; e < bb; e += blockDim.x) {
        double s = 0.0;
        for (int c = 0; c < nchunks; ++c) s += partial[((size_t)c * nb + i) * bb + e];
        out[(size_t)i * bb + e] = s;
    }
}

// W[r][d] -= sum_{i<nb} sum_a V_i[r][a] C[i][a][d]
template <int B>
__global__ void __launch_bounds__(kBT)
k_blk_update(int nb, const double* __restrict__ V0, int64_t bstride, const double* __restrict__ C, int64_t n,
             double* __restrict__ W)
{
    extern __shared__ double sC[];
    for (int t = threadIdx.x; t < nb * B * B; t += blockDim.x) sC[t] = C[t];
    __syncthreads();
    const int d = threadIdx.x % B, rl = threadIdx.x / B;
    constexpr int RL = kBT / B;
    for (int64_t r = (int64_t)blockIdx.x * RL + rl; r < n; r += (int64_t)gridDim.x * RL) {
        double w = W[(size_t)r * B + d];
        for (int i = 0; i < nb; ++i) {
            const double* vr = V0 + (size_t)i * bstride + (size_t)r * B;
#pragma unroll
            for (int a = 0; a < B; ++a) w = fma(-vr[a], sC[(i * B + a) * B + d], w);
        }
        W[(size_t)r * B + d] = w;
    }
}

// Y[r][d] = sum_a X[r][a] M[a][d]  (X != Y)
template <int B>
__global__ void __launch_bounds__(kBT)
k_blk_mul(const double* __restrict__ X, const double* __restrict__ M, int64_t n, double* __restrict__ Y)
{
    __shared__ double sM[B * B];
    for (int t = threadIdx.x; t < B * B; t += blockDim.x) sM[t] = M[t];
    __syncthreads();
    const int d = threadIdx.x % B, rl = threadIdx.x / B;
    constexpr int RL = kBT / B;
    for (int64_t r = (int64_t)blockIdx.x * RL + rl; r < n; r += (int64_t)gridDim.x * RL) {
        const double* xr = X + (size_t)r * B;
        double s = 0.0;
#pragma unroll
        for (int a = 0; a < B; ++a) s = fma(xr[a], sM[a * B + d], s);
        Y[(size_t)r * B + d] = s;
    }
}

// Column l of the basis is V[(l/B)*bstride + r*B + l%B].  out block q' column d
// = sum_{l<L} V_l Z[l*ldz + q'*B + d] for q' < qn.
template <int B>
__global__ void __launch_bounds__(kBT)
k_blk_combine(int L, int qn, const double* __restrict__ V0, int64_t bstride, const double* __restrict__ Z, int ldz,
              int64_t n, double* __restrict__ out)
{
    extern __shared__ double sZ[];
    for (int t = threadIdx.x; t < L * ldz; t += blockDim.x) sZ[t] = Z[t];
    __syncthreads();
    const int d = threadIdx.x % B, rl = threadIdx.x / B;
    constexpr int RL = kBT / B;
    for (int64_t r = (int64_t)blockIdx.x * RL + rl; r < n; r += (int64_t)gridDim.x * RL) {
        for (int qq = 0; qq < qn; ++qq) {
            double s = 0.0;
            for (int l = 0; l < L; ++l)
                s = fma(V0[(size_t)(l / B) * bstride + (size_t)r * B + (l % B)], sZ[l * ldz + qq * B + d], s);
            out[(size_t)qq * bstride + (size_t)r * B + d] = s;
        }
    }
}

// Ritz vectors as contiguous columns: out[i*n + r] = sum_l V_l[r] Z[l*ldz + i], i < nv
template <int B>
__global__ void __launch_bounds__(kBT)
k_blk_extract(int L, int nv, const double* __restrict__ V0, int64_t bstride, const double* __restrict__ Z, int ldz,
              int64_t n, double* __restrict__ out)
{
    extern __shared__ double sZ[];
    for (int t = threadIdx.x; t < L * ldz; t += blockDim.x) sZ[t] = Z[t];
    __syncthreads();
    for (int64_t r = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; r < n; r += (int64_t)gridDim.x * blockDim.x)
        for (int i = 0; i < nv; ++i) {
            double s = 0.0;
            for (int l = 0; l < L; ++l)
                s = fma(V0[(size_t)(l / B) * bstride + (size_t)r * B + (l % B)], sZ[l * ldz + i], s);
            out[(size_t)i * n + r] = s;
        }
}

// Random-normal block (splitmix64 + Box-Muller); column `skip` keeps its value
// when skip >= 0 (a caller's start vector).
template <int B>
__global__ void __launch_bounds__(kBT)
k_blk_randn(uint64_t seed, int64_t n, int skip, double* __restrict__ out)
{
    for (int64_t t = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; t < n * B; t += (int64_t)gridDim.x * blockDim.x) {
        if ((int)(t % B) == skip) continue;
        const uint64_t h = bmix(seed ^ bmix((uint64_t)t));
        const double u1 = ((h >> 11) + 1) * (1.0 / 9007199254740993.0);
        const double u2 = (bmix(h) >> 11) * (1.0 / 9007199254740992.0);
        out[t] = sqrt(-2.0 * log(u1)) * cos(6.283185307179586 * u2);
    }
}

// column 0 of a block from a contiguous vector
template <int B>
__global__ void k_blk_set_col(const double* __restrict__ v, int64_t n, double* __restrict__ out)
{
    for (int64_t r = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; r < n; r += (int64_t)gridDim.x * blockDim.x)
        out[(size_t)r * B] = v[r];
}

__global__ void k_resid_sq(const double* __restrict__ hx, const double* __restrict__ x, double theta, int64_t n,
                           double* __restrict__ partial)
{
    __shared__ double red[256];
    double s = 0.0;
    for (int64_t r = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; r < n; r += (int64_t)gridDim.x * blockDim.x) {
        const double d = fma(-theta, x[r], hx[r]);
        s = fma(d, d, s);
    }
    red[threadIdx.x] = s;
    __syncthreads();
    for (int o = 128; o > 0; o >>= 1) {
        if ((int)threadIdx.x < o) red[threadIdx.x] += red[threadIdx.x + o];
        __syncthreads();
    }
    if (threadIdx.x == 0) partial[blockIdx.x] = red[0];
}

// ---------------------------------------------------------------- host side

// Upper Cholesky G = R^T R of a B x B Gram matrix; false when not positive.
bool cholesky_upper(int B, const std::vector<double>& G, std::vector<double>& R)
{
    R.assign((size_t)B * B, 0.0);
    double scale = 0.0;
    for (int i = 0; i < B; ++i) scale = std::max(scale, G[(size_t)i * B + i]);
    for (int i = 0; i < B; ++i) {
        double s = G[(size_t)i * B + i];
        for (int k = 0; k < i; ++k) s -= R[(size_t)k * B + i] * R[(size_t)k * B + i];
        if (!(s > 1e-28 * scale) || !(scale > 0)) return false;
        const double rii = std::sqrt(s);
        R[(size_t)i * B + i] = rii;
        for (int j = i + 1; j < B; ++j) {
            double t = G[(size_t)i * B + j];
            for (int k = 0; k < i; ++k) t -= R[(size_t)k * B + i] * R[(size_t)k * B + j];
            R[(size_t)i * B + j] = t / rii;
        }
    }
    return true;
}

// inverse of an upper triangular B x B matrix
void upper_inverse(int B, const std::vector<double>& R, std::vector<double>& M)
{
    M.assign((size_t)B * B, 0.0);
    for (int j = 0; j < B; ++j) {
        M[(size_t)j * B + j] = 1.0 / R[(size_t)j * B + j];
        for (int i = j - 1; i >= 0; --i) {
            double s = 0.0;
            for (int k = i + 1; k <= j; ++k) s += R[(size_t)i * B + k] * M[(size_t)k * B + j];
            M[(size_t)i * B + j] = -s / R[(size_t)i * B + i];
        }
    }
}

template <int B>
struct BlockLanczos {
    hd_hamiltonian* h = nullptr;
    hipStream_t st = nullptr;
    int64_t n = 0;
    int64_t bstride = 0;  // doubles per block (n * B)
    int m = 0;            // blocks in the basis before a restart
    int n_cu = 256;
    int lpr = 8;
    double* V = nullptr;      // (m+1) blocks
    double* W = nullptr;      // one block
    double* tmp = nullptr;    // restart blocks
    double* partial = nullptr;
    double* coef = nullptr;   // (m+1)*B*B reduced dots
    double* small = nullptr;  // B*B matrices and Z uploads
    int nchunks = 0;
    int qmax = 0;
    double t_spmm = 0.0;
    int spmms = 0;
    EventTimer timer;

    ~BlockLanczos()
    {
        dfree(V);
        dfree(W);
        dfree(tmp);
        dfree(partial);
        dfree(coef);
        dfree(small);
    }

    double* blk(int j) { return V + (size_t)j * bstride; }

    int alloc(int q)
    {
        qmax = q;
        bstride = n * B;
        nchunks = (int)((n + kBChunk - 1) / kBChunk);
        HD_CHECK(dalloc_t(&V, (size_t)(m + 1) * bstride, "block Lanczos basis"));
        HD_CHECK(dalloc_t(&W, (size_t)bstride, "block work"));
        HD_CHECK(dalloc_t(&tmp, (size_t)q * bstride, "restart blocks"));
        HD_CHECK(dalloc_t(&partial, std::max<size_t>((size_t)nchunks * (m + 1) * B * B, 256), "block dot partials"));
        HD_CHECK(dalloc_t(&coef, (size_t)(m + 1) * B * B, "block coefficients"));
        HD_CHECK(dalloc_t(&small, (size_t)(m + 1) * B * (m + 1) * B + B * B, "small matrices"));
        return HD_OK;
    }

    int grid_rows(int64_t rows_per_block) const
    {
        return (int)std::max<int64_t>(1, std::min<int64_t>((n + rows_per_block - 1) / rows_per_block, (int64_t)n_cu * 8));
    }

    template <int LPR>
    int spmm_lpr(const double* x, double* y)
    {
        const bool coded = h->d_code && (!h->d_val64 || prefer_coded);
        const int grid = grid_rows(256 / LPR);
        if (coded) {
            const size_t lds = (size_t)std::max(h->ndict, 1) * sizeof(double);
            if (lds > 160 * 1024) return fail(HD_ERR_UNSUPPORTED, "value dictionary larger than LDS");
            if (lds > 64 * 1024)
                HD_HIP(hipFuncSetAttribute((const void*)k_spmm_blk<B, LPR, true>,
                                           hipFuncAttributeMaxDynamicSharedMemorySize, (int)lds));
            hipLaunchKernelGGL((k_spmm_blk<B, LPR, true>), dim3(grid), dim3(256), lds, st, n, h->d_rowptr, h->d_col,
                               nullptr, h->d_code, static_cast<const double*>(h->d_dict), h->ndict,
                               static_cast<const double*>(h->d_diag), x, y);
        } else {
            hipLaunchKernelGGL((k_spmm_blk<B, LPR, false>), dim3(grid), dim3(256), 0, st, n, h->d_rowptr, h->d_col,
                               static_cast<const double*>(h->d_val64), nullptr, nullptr, 0, nullptr, x, y);
        }
        HD_HIP(hipGetLastError());
        return HD_OK;
    }
    bool prefer_coded = true;

    int spmm_raw(const double* x, double* y)
    {
        switch (lpr) {
        case 4: return spmm_lpr<4>(x, y);
        case 8: return spmm_lpr<8>(x, y);
        case 16: return spmm_lpr<16>(x, y);
        default: return spmm_lpr<32>(x, y);
        }
    }

    int spmm(const double* x, double* y)
    {
        timer.start(st);
        HD_CHECK(spmm_raw(x, y));
        t_spmm += timer.stop(st);
        ++spmms;
        return timer.status("block SpMM timing");
    }

    // lanes per row by timing (x, y: block vectors)
    int autotune(const double* x, double* y)
    {
        float best = 1e30f;
        int best_l = 8;
        for (int l : {4, 8, 16, 32}) {
            lpr = l;
            HD_CHECK(spmm_raw(x, y));
            timer.start(st);
            for (int r = 0; r < 2; ++r) HD_CHECK(spmm_raw(x, y));
            const float ms = timer.stop(st);
            HD_CHECK(timer.status("block SpMM autotune"));
            if (ms < best) {
                best = ms;
                best_l = l;
            }
        }
        lpr = best_l;
        return HD_OK;
    }

    // coef[i] = V_i^T w for i < nb (B x B each), copied to host
    int dots(int nb, const double* base, const double* w, std::vector<double>& host)
    {
        hipLaunchKernelGGL(k_blk_dots<B>, dim3((unsigned)nchunks, (unsigned)nb), dim3(kBT), 0, st, nb, base, bstride,
                           w, n, partial);
        hipLaunchKernelGGL(k_blk_reduce, dim3((unsigned)nb), dim3(B * B), 0, st, nchunks, nb, B * B, partial, coef);
        HD_HIP(hipGetLastError());
        host.resize((size_t)nb * B * B);
        HD_HIP(hipMemcpyAsync(host.data(), coef, host.size() * 8, hipMemcpyDeviceToHost, st));
        HD_HIP(hipStreamSynchronize(st));
        return HD_OK;
    }

    // w -= sum_i V_i coef[i] with the coefficients currently in `coef`
    int update(int nb, double* w)
    {
        const size_t lds = (size_t)nb * B * B * sizeof(double);
        hipLaunchKernelGGL(k_blk_update<B>, dim3(grid_rows(kBT / B)), dim3(kBT), lds, st, nb, V, bstride, coef, n, w);
        HD_HIP(hipGetLastError());
        return HD_OK;
    }

    // y = x M (B x B host matrix)
    int mul(const double* x, const std::vector<double>& M, double* y)
    {
        HD_HIP(hipMemcpyAsync(small, M.data(), (size_t)B * B * 8, hipMemcpyHostToDevice, st));
        hipLaunchKernelGGL(k_blk_mul<B>, dim3(grid_rows(kBT / B)), dim3(kBT), 0, st, x, small, n, y);
        HD_HIP(hipGetLastError());
        return HD_OK;
    }

    // Orthogonalise w against blocks 0..nb-1 (two CGS passes); the sum of the
    // coefficients goes to ctot ((nb*B) x B, row-major [i*B + a][d]).
    int orthogonalize(int nb, double* w, std::vector<double>& ctot)
    {
        std::vector<double> c1, c2;
        HD_CHECK(dots(nb, V, w, c1));
        HD_CHECK(update(nb, w));
        HD_CHECK(dots(nb, V, w, c2));
        HD_CHECK(update(nb, w));
        ctot.resize(c1.size());
        for (size_t i = 0; i < c1.size(); ++i) ctot[i] = c1[i] + c2[i];
        return HD_OK;
    }

    // w = V_{next} R with V_{next} orthonormal (Cholesky QR twice); false in R
    // signals rank deficiency
    int qr(double* w, double* vnext, std::vector<double>& R, bool& ok)
    {
        std::vector<double> G, R1, R2, M;
        HD_CHECK(dots(1, w, w, G));
        ok = cholesky_upper(B, G, R1);
        if (!ok) return HD_OK;
        upper_inverse(B, R1, M);
        HD_CHECK(mul(w, M, vnext));
        HD_CHECK(dots(1, vnext, vnext, G));
        ok = cholesky_upper(B, G, R2);
        if (!ok) return HD_OK;
        upper_inverse(B, R2, M);
        HD_CHECK(mul(vnext, M, w));
        HD_HIP(hipMemcpyAsync(vnext, w, (size_t)bstride * 8, hipMemcpyDeviceToDevice, st));
        R.assign((size_t)B * B, 0.0);
        for (int i = 0; i < B; ++i)
            for (int j = i; j < B; ++j) {
                double s = 0.0;
                for (int k = i; k <= j; ++k) s += R2[(size_t)i * B + k] * R1[(size_t)k * B + j];
                R[(size_t)i * B + j] = s;
            }
        return HD_OK;
    }
};

}  // namespace

template <int B>
static int run_block(hd_hamiltonian* h, const hd_lanczos_opts& o, double* evals, hd_lanczos_stats* stats)
{
    const auto t_start = std::chrono::steady_clock::now();
    BlockLanczos<B> L;
    L.h = h;
    L.st = h->stream;
    L.n = h->dim;
    L.prefer_coded = o.precision == HD_CODED || !h->d_val64;
    if (!h->d_val64 && !h->d_code) return fail(HD_ERR_ARG, "block Lanczos needs FP64 or coded values");
    HD_CHECK(device_props(&L.n_cu, nullptr));
    const int nev = (int)std::min<int64_t>(o.nev, h->dim);
    const int qneed = (nev + B - 1) / B;
    int m = o.ncv > 0 ? (o.ncv + B - 1) / B : std::max(2 * qneed + 4, 6);
    const int q = std::max(qneed + 1, m / 2);
    m = std::max(m, q + 2);
    if ((int64_t)(m + 1) * B > h->dim) return fail(HD_ERR_ARG, "matrix too small for the block basis; use block = 1");
    if ((m + 1) * B > 160) return fail(HD_ERR_ARG, "block basis above 160 columns");
    L.m = m;
    HD_CHECK(L.alloc(q));
    const double tol = o.tol > 0 ? o.tol : 1e-10;
    const int max_restarts = o.max_restarts > 0 ? o.max_restarts : 2000;
    const int grid = L.grid_rows(kBT);
    // start block: caller's vector in column 0 (if any), random normal elsewhere
    if (o.start == HD_START_USER) {
        if (!o.d_start) return fail(HD_ERR_ARG, "HD_START_USER needs d_start");
        hipLaunchKernelGGL(k_blk_set_col<B>, dim3(grid), dim3(kBT), 0, L.st, static_cast<const double*>(o.d_start),
                           L.n, L.W);
    }
    hipLaunchKernelGGL(k_blk_randn<B>, dim3(grid), dim3(kBT), 0, L.st, o.seed, L.n,
                       o.start == HD_START_USER ? 0 : -1, L.W);
    HD_HIP(hipGetLastError());
    std::vector<double> R, ctot;
    bool ok = false;
    HD_CHECK(L.qr(L.W, L.blk(0), R, ok));
    if (!ok) return fail(HD_ERR_ARG, "start block is rank deficient");
    HD_CHECK(L.autotune(L.blk(0), L.W));

    const int LT = (m + 1) * B;  // projected matrix (with the residual block)
    std::vector<double> T((size_t)LT * LT, 0.0), Tm, w, Z, Rm;
    int k = 0;  // kept blocks
    int restarts = 0;
    bool converged = false;
    uint64_t rng = 1;
    std::vector<double> res(nev, 0.0);
    for (;;) {
        for (int j = k; j < m; ++j) {
            HD_CHECK(L.spmm(L.blk(j), L.W));
            HD_CHECK(L.orthogonalize(j + 1, L.W, ctot));
            // column block j of the projected matrix; on the first step after a
            // restart the rows of the kept Ritz vectors keep the restart coupling
            const int a0 = (j == k && k > 0) ? k * B : 0;
            for (int a = a0; a < (j + 1) * B; ++a)
                for (int d = 0; d < B; ++d) T[(size_t)a * LT + j * B + d] = ctot[(size_t)a * B + d];
            HD_CHECK(L.qr(L.W, L.blk(j + 1), R, ok));
            if (!ok) {
                // invariant subspace reached: continue with a fresh random block
                hipLaunchKernelGGL(k_blk_randn<B>, dim3(grid), dim3(kBT), 0, L.st, o.seed + 7919 * (rng++), L.n, -1,
                                   L.W);
                HD_HIP(hipGetLastError());
                std::vector<double> cc;
                HD_CHECK(L.orthogonalize(j + 1, L.W, cc));
                HD_CHECK(L.qr(L.W, L.blk(j + 1), R, ok));
                if (!ok) return fail(HD_ERR_NOT_CONVERGED, "block Lanczos: repeated breakdown");
                R.assign((size_t)B * B, 0.0);
            }
            for (int a = 0; a < B; ++a)
                for (int d = 0; d < B; ++d) T[(size_t)((j + 1) * B + a) * LT + j * B + d] = R[(size_t)a * B + d];
            Rm = R;
        }
        // Rayleigh-Ritz on the symmetrised m*B x m*B projected matrix
        const int nm = m * B;
        Tm.assign((size_t)nm * nm, 0.0);
        for (int a = 0; a < nm; ++a)
            for (int c = 0; c < nm; ++c) Tm[(size_t)a * nm + c] = 0.5 * (T[(size_t)a * LT + c] + T[(size_t)c * LT + a]);
        if (!sym_eig(nm, Tm, w, Z)) return fail(HD_ERR_NOT_CONVERGED, "projected eigenproblem did not converge");
        // residual of Ritz pair i: ||R_m z_i||, z_i = rows of the last block
        bool all = true;
        for (int i = 0; i < nev; ++i) {
            double s = 0.0;
            for (int a = 0; a < B; ++a) {
                double t = 0.0;
                for (int d = 0; d < B; ++d) t += Rm[(size_t)a * B + d] * Z[(size_t)((m - 1) * B + d) * nm + i];
                s += t * t;
            }
            res[i] = std::sqrt(s) / std::max(1.0, std::fabs(w[i]));
            if (!(res[i] <= tol)) all = false;
        }
        ++restarts;
        if (all) {
            converged = true;
            break;
        }
        if (restarts >= max_restarts) break;
        // thick restart: blocks 0..q-1 <- Ritz vectors 0..qB-1, block q <- residual block
        const int qb = q * B;
        std::vector<double> Zq((size_t)nm * qb);
        for (int l = 0; l < nm; ++l)
            for (int c = 0; c < qb; ++c) Zq[(size_t)l * qb + c] = Z[(size_t)l * nm + c];
        HD_HIP(hipMemcpyAsync(L.small, Zq.data(), Zq.size() * 8, hipMemcpyHostToDevice, L.st));
        const size_t zl = Zq.size() * 8;
        if (zl > 64 * 1024)
            HD_HIP(hipFuncSetAttribute((const void*)k_blk_combine<B>, hipFuncAttributeMaxDynamicSharedMemorySize,
                                       (int)zl));
        hipLaunchKernelGGL(k_blk_combine<B>, dim3(L.grid_rows(kBT / B)), dim3(kBT), zl, L.st, nm, q, L.V, L.bstride,
                           L.small, qb, L.n, L.tmp);
        HD_HIP(hipGetLastError());
        HD_HIP(hipMemcpyAsync(L.blk(q), L.blk(m), (size_t)L.bstride * 8, hipMemcpyDeviceToDevice, L.st));
        HD_HIP(hipMemcpyAsync(L.blk(0), L.tmp, (size_t)q * L.bstride * 8, hipMemcpyDeviceToDevice, L.st));
        std::fill(T.begin(), T.end(), 0.0);
        for (int i = 0; i < qb; ++i) T[(size_t)i * LT + i] = w[i];
        // coupling of the residual block to the kept Ritz vectors: R_m z_i
        for (int a = 0; a < B; ++a)
            for (int i = 0; i < qb; ++i) {
                double t = 0.0;
                for (int d = 0; d < B; ++d) t += Rm[(size_t)a * B + d] * Z[(size_t)((m - 1) * B + d) * nm + i];
                T[(size_t)(qb + a) * LT + i] = t;
                T[(size_t)i * LT + qb + a] = t;
            }
        k = q;
    }
    // eigenvectors (contiguous columns) and FP64 residuals with the single-vector operator
    const int nm = m * B;
    hd_lanczos_state* S = h->lz;
    if (!S) S = h->lz = new hd_lanczos_state();
    dfree(S->d_vec);
    S->d_vec = nullptr;
    S->dim = L.n;
    S->cplx = 0;
    S->nev = nev;
    HD_CHECK(dalloc(&S->d_vec, sizeof(double) * (size_t)L.n * nev, "eigenvectors"));
    {
        std::vector<double> Zn((size_t)nm * nev);
        for (int l = 0; l < nm; ++l)
            for (int i = 0; i < nev; ++i) Zn[(size_t)l * nev + i] = Z[(size_t)l * nm + i];
        HD_HIP(hipMemcpyAsync(L.small, Zn.data(), Zn.size() * 8, hipMemcpyHostToDevice, L.st));
        hipLaunchKernelGGL(k_blk_extract<B>, dim3(grid), dim3(kBT), Zn.size() * 8, L.st, nm, nev, L.V, L.bstride,
                           L.small, nev, L.n, static_cast<double*>(S->d_vec));
        HD_HIP(hipGetLastError());
    }
    S->evals.assign(w.begin(), w.begin() + nev);
    for (int i = 0; i < nev; ++i) evals[i] = w[i];
    if (stats)
        for (int i = 0; i < 64; ++i) stats->residual[i] = 0.0;
    int nconv = 0;
    double* hx = L.W;  // reuse the work block as one vector
    const int rgrid = 256;
    double* part = L.partial;
    const int prec64 = L.prefer_coded && h->d_code ? HD_CODED : HD_F64;
    for (int i = 0; i < nev; ++i) {
        const double* xi = static_cast<const double*>(S->d_vec) + (size_t)i * L.n;
        HD_CHECK(spmv_launch(h, prec64, xi, hx, L.st));
        hipLaunchKernelGGL(k_resid_sq, dim3(rgrid), dim3(256), 0, L.st, hx, xi, w[i], L.n, part);
        HD_HIP(hipGetLastError());
        std::vector<double> ps(rgrid);
        HD_HIP(hipMemcpyAsync(ps.data(), part, rgrid * 8, hipMemcpyDeviceToHost, L.st));
        HD_HIP(hipStreamSynchronize(L.st));
        double s = 0.0;
        for (double v : ps) s += v;
        const double r = std::sqrt(s) / std::max(1.0, std::fabs(w[i]));
        if (stats && i < 64) stats->residual[i] = r;
        if (r <= tol) ++nconv;
    }
    if (stats) {
        stats->converged = converged && nconv == nev;
        stats->nconv = nconv;
        stats->restarts = restarts;
        stats->matvecs = L.spmms * B;
        stats->matvecs_f32 = 0;
        stats->matvecs_f64 = L.spmms * B;
        stats->time_spmv_ms = L.t_spmm;
        stats->time_ortho_ms = 0.0;
        stats->time_total_ms =
            std::chrono::duration<double, std::milli>(std::chrono::steady_clock::now() - t_start).count();
    }
    if (!converged) return fail(HD_ERR_NOT_CONVERGED, "block Lanczos reached the restart limit before tol");
    return HD_OK;
}

int run_block_lanczos(hd_hamiltonian* h, const hd_lanczos_opts& o, double* evals, hd_lanczos_stats* stats)
{
    if (h->is_complex) return fail(HD_ERR_UNSUPPORTED, "block Lanczos handles real matrices");
    switch (o.block) {
    case 4: return run_block<4>(h, o, evals, stats);
    case 8: return run_block<8>(h, o, evals, stats);
    case 16: return run_block<16>(h, o, evals, stats);
    default: return fail(HD_ERR_ARG, "block size must be 1, 4, 8 or 16");
    }
}

}  // namespace hd
