// hd_lanczos.hip -- device-resident thick-restart Lanczos (FP64 recurrence).
//
// Reference: ArpackWrapper::DiagonalizeSymmetricMatrix (src/utilities/wrappers/arpack_wrapper.hpp:595-937)
// drives PARPACK p?naupd/p?neupd by reverse communication, WHICH = 'SR',
// NCV = min(N, 2*NEV+21), exact shifts, maxiter 2000 (arpack_wrapper.cpp:38-49),
// with every mat-vec a host round trip through MatrixVectorFunction::Simple.
// Here the Krylov basis, the operator and all vector work stay in HBM; the host
// only sees the small projected matrix.  Method: thick-restart Lanczos (Wu and
// Simon, SIAM J. Matrix Anal. Appl. 22 (2000) 602) with full
// reorthogonalisation: each step first removes the three-term part locally
// (w -= beta v_{j-1}; alpha = v_j.w; w -= alpha v_j), then one classical
// Gram-Schmidt pass against the whole basis, with a second pass when the norm
// drops by more than 100x (the DGKS criterion); the first step after a restart
// (whose projection is the arrow, not tridiagonal) always takes two passes
// (Giraud et al., Numer. Math. 101 (2005) 87).  The projected
// tridiagonal-plus-arrow matrix is diagonalised by cyclic Jacobi rotations on
// the host.
//
// Mixed precision (precision = HD_F32) chooses the operator per restart
// cycle.  An operator error F_j in the product of basis vector v_j enters the
// residual of a Ritz vector y = sum_j s_j v_j as sum_j s_j F_j, which the
// recurrence never sees (Simoncini, SIAM J. Numer. Anal. 43 (2005) 1155:
// products may be inexact where the Ritz coefficients are small).  The kept
// Ritz vectors carry O(1) coefficients on every vector of the cycles that made
// them, so the early cycles use the exact FP64 operator; once the wanted Ritz
// vectors' coefficients on a cycle's new vectors are small, the FP32-input
// operator (HD_TILED_X32, exact for x rounded to FP32, or the FP32-value CRS)
// takes the cycle.  The rounding such a cycle adds to the residual of wanted
// pair i is estimated by eps32 |T| |s_i,new| (eps32 the operator's relative
// rounding, |T| the largest projected row, s_i,new the pair's coefficients on
// the cycle's new vectors); a cycle runs in FP32 only while the sum of these
// estimates, with the next cycle's predicted coefficients, stays within
// f32_budget * tol * max(1, |theta_i|).  The prediction cannot see an
// eigenvalue that enters the wanted window during the cycle, so the estimate
// is checked again with the cycle's actual coefficients, and a cycle beyond
// the budget is discarded (its starting state was kept) and run in FP64.
// Convergence is the usual Ritz test;
// the eigenpairs are then checked with the FP64 operator, and hd_lanczos
// repeats the solve with the FP64 operator should a residual miss tol.  All
// reductions are FP64 and deterministic (per-block partials reduced in a fixed
// order).
#include <hip/hip_runtime.h>

#include <algorithm>
#include <chrono>
#include <cmath>
#include <cstdio>
#include <cstdlib>
#include <cstring>
#include <limits>
#include <numeric>

#include "hd_internal.hpp"


namespace hd {

void lanczos_state_free(hd_lanczos_state* s)
{
    if (!s) return;
    dfree(s->d_vec);
    delete s;
}

// ------------------------------------------------------------ device side

template <bool C> struct E;
template <> struct E<false> { using T = double; using F = float; };
template <> struct E<true> { using T = double2; using F = float2; };

HD_DEV double conj_mul(double a, double b) { return a * b; }          // conj(a) b
HD_DEV double2 conj_mul(double2 a, double2 b)
{
    return make_double2(a.x * b.x + a.y * b.y, a.x * b.y - a.y * b.x);
}
HD_DEV double mul(double a, double b) { return a * b; }
HD_DEV double2 mul(double2 a, double2 b) { return make_double2(a.x * b.x - a.y * b.y, a.x * b.y + a.y * b.x); }
HD_DEV double sub(double a, double b) { return a - b; }
HD_DEV double2 sub(double2 a, double2 b) { return make_double2(a.x - b.x, a.y - b.y); }
HD_DEV double add2(double a, double b) { return a + b; }
HD_DEV double2 add2(double2 a, double2 b) { return make_double2(a.x + b.x, a.y + b.y); }
HD_DEV double norm2(double a) { return a * a; }
HD_DEV double norm2(double2 a) { return a.x * a.x + a.y * a.y; }
HD_DEV double scale(double a, double s) { return a * s; }
HD_DEV double2 scale(double2 a, double s) { return make_double2(a.x * s, a.y * s); }
template <typename T> HD_DEV T zero();
template <> HD_DEV double zero<double>() { return 0.0; }
template <> HD_DEV double2 zero<double2>() { return make_double2(0.0, 0.0); }
template <typename T> HD_DEV T real_part(double x);  // x + 0i
template <> HD_DEV double real_part<double>(double x) { return x; }
template <> HD_DEV double2 real_part<double2>(double x) { return make_double2(x, 0.0); }
HD_DEV float narrow(double a) { return (float)a; }
HD_DEV float2 narrow(double2 a) { return make_float2((float)a.x, (float)a.y); }
HD_DEV double widen(float a) { return (double)a; }
HD_DEV double2 widen(float2 a) { return make_double2((double)a.x, (double)a.y); }

constexpr int kVB = 256;  // block size of the vector kernels
constexpr int kNV = 16;   // vectors per dot-product pass
constexpr int kMaxBasis = 256;               // Krylov basis limit (ncv)
constexpr int kNormSlot = kMaxBasis + 8;     // coefficient slot of the staged norm

// partial[j * nb + block] = sum_r conj(V_j[r]) w[r]  for j in [j0, j0+nv);
// with want_norm, partial[(kNV) * nb + block] = sum |w|^2 (as .x).
template <bool C>
__global__ void __launch_bounds__(kVB)
k_dots(int64_t dim, const typename E<C>::T* __restrict__ V, int64_t ld, int j0, int nv,
       const typename E<C>::T* __restrict__ w, typename E<C>::T* __restrict__ partial, int want_norm)
{
    using T = typename E<C>::T;
    __shared__ T red[kVB];
    T acc[kNV];
#pragma unroll
    for (int q = 0; q < kNV; ++q) acc[q] = zero<T>();
    double nrm = 0.0;
    const int64_t stride = (int64_t)gridDim.x * kVB;
    for (int64_t r = (int64_t)blockIdx.x * kVB + threadIdx.x; r < dim; r += stride) {
        const T wr = w[r];
        if (want_norm) nrm += norm2(wr);
#pragma unroll
        for (int q = 0; q < kNV; ++q)
            if (q < nv) acc[q] = add2(acc[q], conj_mul(V[(int64_t)(j0 + q) * ld + r], wr));
    }
    const int nout = nv + (want_norm ? 1 : 0);
    for (int q = 0; q < nout; ++q) {
        // select acc[q] with static indices, so that acc stays in registers (a
        // runtime index would put it in scratch memory)
        T v = real_part<T>(nrm);
#pragma unroll
        for (int u = 0; u < kNV; ++u)
            if (u == q && q < nv) v = acc[u];
        red[threadIdx.x] = v;
        __syncthreads();
        for (int o = kVB / 2; o > 0; o >>= 1) {
            if ((int)threadIdx.x < o) red[threadIdx.x] = add2(red[threadIdx.x], red[threadIdx.x + o]);
            __syncthreads();
        }
        if (threadIdx.x == 0) partial[(int64_t)q * gridDim.x + blockIdx.x] = red[0];
        __syncthreads();
    }
}

// out[q] = sum_b partial[q * nb + b]: one block per output, strided partial
// sums then a fixed tree, so the result does not depend on scheduling.
template <bool C>
__global__ void __launch_bounds__(kVB)
k_reduce(const typename E<C>::T* __restrict__ partial, int nb, int nout, typename E<C>::T* __restrict__ out)
{
    using T = typename E<C>::T;
    __shared__ T red[kVB];
    const int q = blockIdx.x;
    if (q >= nout) return;
    T s = zero<T>();
    for (int b = threadIdx.x; b < nb; b += kVB) s = add2(s, partial[(int64_t)q * nb + b]);
    red[threadIdx.x] = s;
    __syncthreads();
    for (int o = kVB / 2; o > 0; o >>= 1) {
        if ((int)threadIdx.x < o) red[threadIdx.x] = add2(red[threadIdx.x], red[threadIdx.x + o]);
        __syncthreads();
    }
    if (threadIdx.x == 0) out[q] = red[0];
}

// w -= sum_{j<nv} V_j coef[j]; partial norm^2 of the result.
template <bool C>
__global__ void __launch_bounds__(kVB)
k_project_out(int64_t dim, const typename E<C>::T* __restrict__ V, int64_t ld, int nv,
              const typename E<C>::T* __restrict__ coef, typename E<C>::T* __restrict__ w,
              double* __restrict__ partial)
{
    using T = typename E<C>::T;
    __shared__ T sc[kMaxBasis];
    __shared__ double red[kVB];
    for (int i = threadIdx.x; i < nv; i += kVB) sc[i] = coef[i];
    __syncthreads();
    double nrm = 0.0;
    const int64_t stride = (int64_t)gridDim.x * kVB;
    for (int64_t r = (int64_t)blockIdx.x * kVB + threadIdx.x; r < dim; r += stride) {
        T acc = w[r];
        for (int j = 0; j < nv; ++j) acc = sub(acc, mul(V[(int64_t)j * ld + r], sc[j]));
        w[r] = acc;
        nrm += norm2(acc);
    }
    red[threadIdx.x] = nrm;
    __syncthreads();
    for (int o = kVB / 2; o > 0; o >>= 1) {
        if ((int)threadIdx.x < o) red[threadIdx.x] += red[threadIdx.x + o];
        __syncthreads();
    }
    if (threadIdx.x == 0) partial[blockIdx.x] = red[0];
}

__global__ void __launch_bounds__(kVB) k_reduce_real(const double* __restrict__ partial, int nb, double* __restrict__ out)
{
    __shared__ double red[kVB];
    double s = 0.0;
    for (int b = threadIdx.x; b < nb; b += kVB) s += partial[b];
    red[threadIdx.x] = s;
    __syncthreads();
    for (int o = kVB / 2; o > 0; o >>= 1) {
        if ((int)threadIdx.x < o) red[threadIdx.x] += red[threadIdx.x + o];
        __syncthreads();
    }
    if (threadIdx.x == 0) *out = red[0];
}

template <bool C>
__global__ void k_scale(int64_t dim, const typename E<C>::T* __restrict__ src, double s,
                        typename E<C>::T* __restrict__ dst)
{
    const int64_t stride = (int64_t)gridDim.x * blockDim.x;
    for (int64_t r = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; r < dim; r += stride)
        dst[r] = scale(src[r], s);
}

// out_l = sum_i V_i Y[i][l] for l in [l0, l0+nl), nl <= NL, with the m x nl
// block of Y staged in LDS (m * nl doubles of dynamic LDS).
template <bool C, int NL>
__global__ void __launch_bounds__(kVB)
k_combine(int64_t dim, const typename E<C>::T* __restrict__ V, int64_t ld, int m,
          const double* __restrict__ Y, int ldy, int l0, int nl, typename E<C>::T* __restrict__ out)
{
    using T = typename E<C>::T;
    extern __shared__ double sy[];
    for (int i = threadIdx.x; i < m * nl; i += kVB) sy[i] = Y[(i / nl) * ldy + l0 + (i % nl)];
    __syncthreads();
    const int64_t stride = (int64_t)gridDim.x * kVB;
    for (int64_t r = (int64_t)blockIdx.x * kVB + threadIdx.x; r < dim; r += stride) {
        T acc[NL];
#pragma unroll
        for (int l = 0; l < NL; ++l) acc[l] = zero<T>();
        // four basis vectors' loads in flight per thread; the sum over i keeps
        // its order
        for (int i0 = 0; i0 < m; i0 += 4) {
            T v[4];
#pragma unroll
            for (int u = 0; u < 4; ++u) v[u] = i0 + u < m ? V[(int64_t)(i0 + u) * ld + r] : zero<T>();
#pragma unroll
            for (int u = 0; u < 4; ++u) {
                if (i0 + u >= m) break;
#pragma unroll
                for (int l = 0; l < NL; ++l)
                    if (l < nl) acc[l] = add2(acc[l], scale(v[u], sy[(i0 + u) * nl + l]));
            }
        }
#pragma unroll
        for (int l = 0; l < NL; ++l)
            if (l < nl) out[(int64_t)l * ld + r] = acc[l];
    }
}

// w -= a * p; partial[block] = sum conj(q) w (the updated w).
template <bool C>
__global__ void __launch_bounds__(kVB)
k_axpy_dot(int64_t dim, const typename E<C>::T* __restrict__ p, double a, const typename E<C>::T* __restrict__ q,
           typename E<C>::T* __restrict__ w, typename E<C>::T* __restrict__ partial)
{
    using T = typename E<C>::T;
    __shared__ T red[kVB];
    T s = zero<T>();
    const int64_t stride = (int64_t)gridDim.x * kVB;
    for (int64_t r = (int64_t)blockIdx.x * kVB + threadIdx.x; r < dim; r += stride) {
        const T v = sub(w[r], scale(p[r], a));
        w[r] = v;
        s = add2(s, conj_mul(q[r], v));
    }
    red[threadIdx.x] = s;
    __syncthreads();
    for (int o = kVB / 2; o > 0; o >>= 1) {
        if ((int)threadIdx.x < o) red[threadIdx.x] = add2(red[threadIdx.x], red[threadIdx.x + o]);
        __syncthreads();
    }
    if (threadIdx.x == 0) partial[blockIdx.x] = red[0];
}

// w -= a * p; partial[block] = sum |w|^2 (the updated w, as .x).
template <bool C>
__global__ void __launch_bounds__(kVB)
k_axpy_norm(int64_t dim, const typename E<C>::T* __restrict__ p, double a, typename E<C>::T* __restrict__ w,
            typename E<C>::T* __restrict__ partial)
{
    using T = typename E<C>::T;
    __shared__ T red[kVB];
    T s = zero<T>();
    const int64_t stride = (int64_t)gridDim.x * kVB;
    for (int64_t r = (int64_t)blockIdx.x * kVB + threadIdx.x; r < dim; r += stride) {
        const T v = sub(w[r], scale(p[r], a));
        w[r] = v;
        s = add2(s, conj_mul(v, v));
    }
    red[threadIdx.x] = s;
    __syncthreads();
    for (int o = kVB / 2; o > 0; o >>= 1) {
        if ((int)threadIdx.x < o) red[threadIdx.x] = add2(red[threadIdx.x], red[threadIdx.x + o]);
        __syncthreads();
    }
    if (threadIdx.x == 0) partial[blockIdx.x] = red[0];
}

template <bool C>
__global__ void k_accumulate(int64_t dim, const typename E<C>::T* __restrict__ src, typename E<C>::T* __restrict__ dst)
{
    const int64_t stride = (int64_t)gridDim.x * blockDim.x;
    for (int64_t r = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; r < dim; r += stride)
        dst[r] = add2(dst[r], src[r]);
}

template <bool C>
__global__ void k_to_f32(int64_t dim, const typename E<C>::T* __restrict__ src, typename E<C>::F* __restrict__ dst)
{
    const int64_t stride = (int64_t)gridDim.x * blockDim.x;
    for (int64_t r = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; r < dim; r += stride)
        dst[r] = narrow(src[r]);
}

template <bool C>
__global__ void k_to_f64(int64_t dim, const typename E<C>::F* __restrict__ src, typename E<C>::T* __restrict__ dst)
{
    const int64_t stride = (int64_t)gridDim.x * blockDim.x;
    for (int64_t r = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; r < dim; r += stride)
        dst[r] = widen(src[r]);
}

// Philox4x32-10 (Salmon, Moraes, Dror, Shaw, SC'11) counter-based generator.
HD_DEV void philox(uint32_t c[4], uint32_t k0, uint32_t k1)
{
    for (int i = 0; i < 10; ++i) {
        const uint64_t p0 = (uint64_t)0xD2511F53u * c[0];
        const uint64_t p1 = (uint64_t)0xCD9E8D57u * c[2];
        const uint32_t n0 = (uint32_t)(p1 >> 32) ^ c[1] ^ k0;
        const uint32_t n2 = (uint32_t)(p0 >> 32) ^ c[3] ^ k1;
        c[1] = (uint32_t)p1;
        c[3] = (uint32_t)p0;
        c[0] = n0;
        c[2] = n2;
        k0 += 0x9E3779B9u;
        k1 += 0xBB67AE85u;
    }
}

// Standard normal entries (Box-Muller on Philox uniforms); element r uses
// counter (r, stream).
template <bool C>
__global__ void k_randn(int64_t dim, uint64_t seed, uint32_t stream, typename E<C>::T* __restrict__ out)
{
    const int64_t stride = (int64_t)gridDim.x * blockDim.x;
    for (int64_t r = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; r < dim; r += stride) {
        uint32_t c[4] = {(uint32_t)r, (uint32_t)((uint64_t)r >> 32), stream, 0x5EEDu};
        philox(c, (uint32_t)seed, (uint32_t)(seed >> 32));
        const double u1 = ((double)c[0] + 1.0) * 2.3283064365386963e-10;  // (0,1]
        const double u2 = (double)c[1] * 2.3283064365386963e-10;
        const double rad = sqrt(-2.0 * log(u1));
        const double z0 = rad * cos(6.283185307179586 * u2);
        if constexpr (C) {
            const double z1 = rad * sin(6.283185307179586 * u2);
            out[r] = make_double2(z0, z1);
        } else {
            out[r] = z0;
        }
    }
}

// partial[b] = sum |y - theta x|^2
template <bool C>
__global__ void __launch_bounds__(kVB)
k_resid(int64_t dim, const typename E<C>::T* __restrict__ y, const typename E<C>::T* __restrict__ x,
        double theta, double* __restrict__ partial)
{
    __shared__ double red[kVB];
    double s = 0.0;
    const int64_t stride = (int64_t)gridDim.x * kVB;
    for (int64_t r = (int64_t)blockIdx.x * kVB + threadIdx.x; r < dim; r += stride)
        s += norm2(sub(y[r], scale(x[r], theta)));
    red[threadIdx.x] = s;
    __syncthreads();
    for (int o = kVB / 2; o > 0; o >>= 1) {
        if ((int)threadIdx.x < o) red[threadIdx.x] += red[threadIdx.x + o];
        __syncthreads();
    }
    if (threadIdx.x == 0) partial[blockIdx.x] = red[0];
}

// -------------------------------------------------------- host: Jacobi

// Symmetric eigen-decomposition of the n x n row-major matrix A (destroyed)
// by cyclic Jacobi rotations; eigenvalues ascending in w, eigenvectors as the
// columns of Z (row-major n x n).
void jacobi_eigh(int n, std::vector<double>& A, std::vector<double>& w, std::vector<double>& Z)
{
    Z.assign((size_t)n * n, 0.0);
    for (int i = 0; i < n; ++i) Z[(size_t)i * n + i] = 1.0;
    auto a = [&](int i, int j) -> double& { return A[(size_t)i * n + j]; };
    for (int sweep = 0; sweep < 100; ++sweep) {
        double off = 0.0, tot = 0.0;
        for (int i = 0; i < n; ++i)
            for (int j = 0; j < n; ++j) {
                tot += a(i, j) * a(i, j);
                if (i != j) off += a(i, j) * a(i, j);
            }
        if (off <= 1e-34 * tot || off == 0.0) break;
        for (int p = 0; p < n - 1; ++p)
            for (int q = p + 1; q < n; ++q) {
                const double apq = a(p, q);
                if (std::fabs(apq) < 1e-300) continue;
                const double theta = (a(q, q) - a(p, p)) / (2.0 * apq);
                const double t = (theta >= 0 ? 1.0 : -1.0) / (std::fabs(theta) + std::sqrt(theta * theta + 1.0));
                const double c = 1.0 / std::sqrt(t * t + 1.0), s = t * c;
                for (int k = 0; k < n; ++k) {
                    const double akp = a(k, p), akq = a(k, q);
                    a(k, p) = c * akp - s * akq;
                    a(k, q) = s * akp + c * akq;
                }
                for (int k = 0; k < n; ++k) {
                    const double apk = a(p, k), aqk = a(q, k);
                    a(p, k) = c * apk - s * aqk;
                    a(q, k) = s * apk + c * aqk;
                }
                for (int k = 0; k < n; ++k) {
                    const double zkp = Z[(size_t)k * n + p], zkq = Z[(size_t)k * n + q];
                    Z[(size_t)k * n + p] = c * zkp - s * zkq;
                    Z[(size_t)k * n + q] = s * zkp + c * zkq;
                }
            }
    }
    std::vector<int> idx(n);
    std::iota(idx.begin(), idx.end(), 0);
    std::sort(idx.begin(), idx.end(), [&](int x, int y) { return a(x, x) < a(y, y); });
    w.resize(n);
    std::vector<double> Zs((size_t)n * n);
    for (int c = 0; c < n; ++c) {
        w[c] = a(idx[c], idx[c]);
        for (int r = 0; r < n; ++r) Zs[(size_t)r * n + c] = Z[(size_t)r * n + idx[c]];
    }
    Z.swap(Zs);
}

// ------------------------------------------------------------ host driver

template <bool C>
struct Lanczos {
    using T = typename E<C>::T;
    using F = typename E<C>::F;
    hd_hamiltonian* h;
    hipStream_t st;
    int64_t n;
    int m;
    int nb;  // blocks for reductions
    T* V = nullptr;     // (m+1) x n
    T* W = nullptr;     // n (operator output / work)
    T* tmp = nullptr;   // up to m x n (restart combination)
    T* bak = nullptr;   // mixed precision: the restart state before an FP32 cycle
    F* x32 = nullptr;
    F* y32 = nullptr;
    T* partial = nullptr;
    T* coef = nullptr;  // device coefficients (kMaxBasis) + staged norm at kNormSlot
    double* dpart = nullptr;
    double* dscal = nullptr;
    double* dY = nullptr;
    double t_spmv = 0, t_ortho = 0;
    int mv32 = 0, mv64 = 0, second_passes = 0, skipped_projections = 0;
    int prec64 = HD_F64;  // the FP64 operator: FP64 values or the coded format
    EventTimer timer;     // brackets every operator application on st

    ~Lanczos()
    {
        dfree(V); dfree(W); dfree(tmp); dfree(bak); dfree(x32); dfree(y32); dfree(partial); dfree(coef);
        dfree(dpart); dfree(dscal); dfree(dY);
    }

    int grid() const { return nb; }

    int alloc(bool need_x32, bool need_y32)
    {
        HD_CHECK(dalloc_t(&V, (size_t)(m + 1) * n, "Lanczos basis"));
        HD_CHECK(dalloc_t(&W, (size_t)n, "Lanczos work vector"));
        HD_CHECK(dalloc_t(&tmp, (size_t)m * n, "restart buffer"));
        if (need_x32) HD_CHECK(dalloc_t(&x32, (size_t)n, "FP32 x"));
        if (need_y32) HD_CHECK(dalloc_t(&y32, (size_t)n, "FP32 y"));
        HD_CHECK(dalloc_t(&partial, (size_t)(kNV + 1) * nb, "dot partials"));
        HD_CHECK(dalloc_t(&coef, kNormSlot + 8, "coefficients"));
        HD_CHECK(dalloc_t(&dpart, (size_t)nb, "norm partials"));
        HD_CHECK(dalloc_t(&dscal, 4, "scalars"));
        HD_CHECK(dalloc_t(&dY, (size_t)kMaxBasis * kMaxBasis, "Ritz coefficients"));
        return HD_OK;
    }

    T* vec(int j) { return V + (size_t)j * n; }

    // W = H V_j in the requested precision
    int apply(const T* x, T* y, int prec)
    {
        timer.start(st);
        int s;
        if (prec == HD_F32 || prec == HD_TILED_X32) {
            hipLaunchKernelGGL(k_to_f32<C>, dim3(nb), dim3(kVB), 0, st, n, x, x32);
            if (prec == HD_F32) {
                s = spmv_launch(h, HD_F32, x32, y32, st);
                hipLaunchKernelGGL(k_to_f64<C>, dim3(nb), dim3(kVB), 0, st, n, y32, y);
            } else {
                s = spmv_launch(h, HD_TILED_X32, x32, y, st);
            }
            ++mv32;
        } else {
            s = spmv_launch(h, prec, x, y, st);
            ++mv64;
        }
        t_spmv += timer.stop(st);
        if (s == HD_OK) s = timer.status("SpMV timing");
        return s;
    }

    // coefficients c[0..nv) = V[0..nv)^H w (device, also copied to h_out),
    // optionally |w|^2 into *nrm
    int dots(int nv, const T* w, std::vector<T>& h_out, double* wnorm2 = nullptr)
    {
        h_out.resize(nv + 1);
        for (int j0 = 0; j0 < nv; j0 += kNV) {
            const int cnt = std::min(kNV, nv - j0);
            const int want = (wnorm2 && j0 == 0) ? 1 : 0;
            hipLaunchKernelGGL(k_dots<C>, dim3(nb), dim3(kVB), 0, st, n, V, n, j0, cnt, w, partial, want);
            hipLaunchKernelGGL(k_reduce<C>, dim3(cnt + want), dim3(kVB), 0, st, partial, nb, cnt + want,
                               coef + (want ? kNormSlot - cnt : j0));
            if (want) {
                // the norm lands in coef[kNormSlot]; move the dots to their slots
                HD_HIP(hipMemcpyAsync(coef + j0, coef + kNormSlot - cnt, sizeof(T) * cnt, hipMemcpyDeviceToDevice,
                                      st));
            }
        }
        HD_HIP(hipGetLastError());
        HD_HIP(hipMemcpyAsync(h_out.data(), coef, sizeof(T) * nv, hipMemcpyDeviceToHost, st));
        if (wnorm2) HD_HIP(hipMemcpyAsync(&h_out[nv], coef + kNormSlot, sizeof(T), hipMemcpyDeviceToHost, st));
        HD_HIP(hipStreamSynchronize(st));
        if (wnorm2) *wnorm2 = *reinterpret_cast<const double*>(&h_out[nv]);
        h_out.resize(nv);
        return HD_OK;
    }

    // w -= V[0..nv) coef (coef on device); returns |w|^2
    int project(int nv, T* w, double* nrm2)
    {
        hipLaunchKernelGGL(k_project_out<C>, dim3(nb), dim3(kVB), 0, st, n, V, n, nv, coef, w, dpart);
        hipLaunchKernelGGL(k_reduce_real, dim3(1), dim3(kVB), 0, st, dpart, nb, dscal);
        HD_HIP(hipGetLastError());
        HD_HIP(hipMemcpyAsync(nrm2, dscal, sizeof(double), hipMemcpyDeviceToHost, st));
        HD_HIP(hipStreamSynchronize(st));
        return HD_OK;
    }

    int norm2_of(const T* w, double* out)
    {
        hipLaunchKernelGGL(k_dots<C>, dim3(nb), dim3(kVB), 0, st, n, V, n, 0, 0, w, partial, 1);
        hipLaunchKernelGGL(k_reduce<C>, dim3(1), dim3(kVB), 0, st, partial, nb, 1, coef);
        T v;
        HD_HIP(hipMemcpyAsync(&v, coef, sizeof(T), hipMemcpyDeviceToHost, st));
        HD_HIP(hipStreamSynchronize(st));
        *out = *reinterpret_cast<double*>(&v);
        return HD_OK;
    }

    // Full reorthogonalisation of w against V[0..nv) by classical Gram-Schmidt:
    // one pass, and a second one only when the pass cancelled more than a
    // factor `kappa` of the norm (the "twice is enough" test of Daniel, Gragg,
    // Kaufman and Stewart, Math. Comp. 30 (1976) 772; one CGS pass leaves an
    // orthogonality error of about eps * |w_before| / |w_after|), or always
    // when `two_pass`.  Returns the summed coefficients and the final |w|.
    //
    // After the local step, w's components along the basis are at rounding
    // level.  When every one of them is below kSkipTol * |w| the projection
    // is skipped and the coefficients are returned as zero (the recurrence then
    // holds for the w actually kept): each new vector is measured against the
    // whole basis, so the basis stays orthogonal to kSkipTol throughout.
    static constexpr double kSkipTol = 1e-12;
    int orthogonalize(int nv, T* w, std::vector<T>& c, double* beta, bool two_pass = false)
    {
        const auto t0 = std::chrono::steady_clock::now();
        std::vector<T> c2;
        double before2 = 0.0, nrm2 = 0.0;
        HD_CHECK(dots(nv, w, c, &before2));
        if (!two_pass) {
            double cmax2 = 0.0;
            for (int j = 0; j < nv; ++j) cmax2 = std::max(cmax2, host_abs2(c[j]));
            if (cmax2 <= kSkipTol * kSkipTol * before2) {
                c.assign(nv, host_zero());
                *beta = std::sqrt(std::max(before2, 0.0));
                ++skipped_projections;
                t_ortho += std::chrono::duration<double, std::milli>(std::chrono::steady_clock::now() - t0).count();
                return HD_OK;
            }
        }
        HD_CHECK(project(nv, w, &nrm2));
        const double kappa = 100.0;
        if (two_pass || nrm2 * kappa * kappa < before2) {
            HD_CHECK(dots(nv, w, c2));
            HD_CHECK(project(nv, w, &nrm2));
            for (int j = 0; j < nv; ++j) c[j] = host_add(c[j], c2[j]);
            ++second_passes;
        }
        *beta = std::sqrt(std::max(nrm2, 0.0));
        t_ortho += std::chrono::duration<double, std::milli>(std::chrono::steady_clock::now() - t0).count();
        return HD_OK;
    }

    // The local three-term part of a Lanczos step, done before the full
    // reorthogonalisation so that its large coefficients never multiply the
    // basis' own orthogonality error (which a single classical Gram-Schmidt
    // pass would amplify by alpha/beta per step):
    //   w -= beta_prev * v_prev;  a = <v_cur, w>;  w -= a * v_cur.
    // With wnorm2, the second pass also returns |w|^2 of the updated w.
    int local_step(const T* v_prev, double beta_prev, const T* v_cur, T* w, double* a, double* wnorm2 = nullptr)
    {
        hipLaunchKernelGGL(k_axpy_dot<C>, dim3(nb), dim3(kVB), 0, st, n, v_prev, beta_prev, v_cur, w, partial);
        hipLaunchKernelGGL(k_reduce<C>, dim3(1), dim3(kVB), 0, st, partial, nb, 1, coef);
        T v;
        HD_HIP(hipGetLastError());
        HD_HIP(hipMemcpyAsync(&v, coef, sizeof(T), hipMemcpyDeviceToHost, st));
        HD_HIP(hipStreamSynchronize(st));
        *a = host_re(v);
        if (wnorm2) {
            hipLaunchKernelGGL(k_axpy_norm<C>, dim3(nb), dim3(kVB), 0, st, n, v_cur, *a, w, partial);
            hipLaunchKernelGGL(k_reduce<C>, dim3(1), dim3(kVB), 0, st, partial, nb, 1, coef);
            HD_HIP(hipGetLastError());
            HD_HIP(hipMemcpyAsync(&v, coef, sizeof(T), hipMemcpyDeviceToHost, st));
            HD_HIP(hipStreamSynchronize(st));
            *wnorm2 = host_re(v);
        } else {
            hipLaunchKernelGGL(k_axpy_dot<C>, dim3(nb), dim3(kVB), 0, st, n, v_cur, *a, v_cur, w, partial);
        }
        HD_HIP(hipGetLastError());
        return HD_OK;
    }

    static double host_abs2(double a) { return a * a; }
    static double host_abs2(double2 a) { return a.x * a.x + a.y * a.y; }
    static T host_zero() { T z; std::memset(&z, 0, sizeof(T)); return z; }
    static double host_add(double a, double b) { return a + b; }
    static double2 host_add(double2 a, double2 b) { return make_double2(a.x + b.x, a.y + b.y); }
    static double host_re(double a) { return a; }
    static double host_re(double2 a) { return a.x; }

    int scale_into(const T* src, double s, T* dst)
    {
        hipLaunchKernelGGL(k_scale<C>, dim3(nb), dim3(kVB), 0, st, n, src, s, dst);
        HD_HIP(hipGetLastError());
        return HD_OK;
    }

    // out[0..kk) = V[0..mm) Y[:, 0..kk)  (Y row-major mm x ldy on the host)
    int combine(int mm, const std::vector<double>& Y, int ldy, int kk, T* out)
    {
        HD_HIP(hipMemcpyAsync(dY, Y.data(), sizeof(double) * mm * ldy, hipMemcpyHostToDevice, st));
        // real vectors: up to 32 outputs per pass over the basis; complex: 16
        constexpr int NL = C ? 16 : 32;
        for (int l0 = 0; l0 < kk; l0 += NL) {
            const int nl = std::min(NL, kk - l0);
            hipLaunchKernelGGL((k_combine<C, NL>), dim3(nb), dim3(kVB), sizeof(double) * mm * nl, st, n, V, n, mm, dY,
                               ldy, l0, nl, out + (size_t)l0 * n);
        }
        HD_HIP(hipGetLastError());
        HD_HIP(hipStreamSynchronize(st));
        return HD_OK;
    }

    int random_start(uint64_t seed, uint32_t stream_id, T* out)
    {
        hipLaunchKernelGGL(k_randn<C>, dim3(nb), dim3(kVB), 0, st, n, seed, stream_id, out);
        HD_HIP(hipGetLastError());
        return HD_OK;
    }

    int residual_norm(const T* y, const T* x, double theta, double* out)
    {
        hipLaunchKernelGGL(k_resid<C>, dim3(nb), dim3(kVB), 0, st, n, y, x, theta, dpart);
        hipLaunchKernelGGL(k_reduce_real, dim3(1), dim3(kVB), 0, st, dpart, nb, dscal);
        HD_HIP(hipGetLastError());
        HD_HIP(hipMemcpyAsync(out, dscal, sizeof(double), hipMemcpyDeviceToHost, st));
        HD_HIP(hipStreamSynchronize(st));
        *out = std::sqrt(std::max(*out, 0.0));
        return HD_OK;
    }
};

template <bool C>
static int run_lanczos(hd_hamiltonian* h, const hd_lanczos_opts& o, double* evals, hd_lanczos_stats* stats)
{
    using T = typename E<C>::T;
    const auto t_start = std::chrono::steady_clock::now();
    Lanczos<C> L;
    L.h = h;
    L.st = h->stream;
    L.n = h->dim;
    const int nev = (int)std::min<int64_t>(o.nev, h->dim);
    int m = o.ncv > 0 ? o.ncv : 2 * nev + 21;
    m = (int)std::min<int64_t>(m, h->dim);
    m = std::min(m, kMaxBasis);
    if (m < nev + 2) return fail(HD_ERR_ARG, "ncv must be at least nev + 2 (and dim > nev + 1)");
    L.m = m;
    int n_cu = 256;
    HD_CHECK(device_props(&n_cu, nullptr));
    L.nb = (int)std::max<int64_t>(1, std::min<int64_t>((h->dim + kVB - 1) / kVB, (int64_t)n_cu * 4));
    const bool mixed = o.precision == HD_F32;
    // the FP32-input operator of the mixed solve: the tiled layout with FP32 x
    // (exact for FP32 inputs), else the FP32-value CRS (FP32 accumulation)
    const int prec32 = (h->d_tg_base && !h->is_complex) ? HD_TILED_X32 : HD_F32;
    if (mixed && prec32 == HD_F32 && !h->d_val32)
        return fail(HD_ERR_ARG, "FP32 Lanczos needs FP32 values (value_mask & 2) or a real tiled layout");
    const bool free_ = h->storage == HD_STORAGE_MATRIX_FREE;
    if (!free_ && !h->d_val64 && !h->d_code && !h->d_tg_base)
        return fail(HD_ERR_ARG, "Lanczos needs FP64, coded or tiled values for the FP64 operator");
    if (o.precision == HD_TILED && !h->d_tg_base) return fail(HD_ERR_ARG, "no tiled layout (hd_ham_tile)");
    if (o.precision < HD_F64 || o.precision > HD_TILED) return fail(HD_ERR_ARG, "unknown Lanczos precision");
    // the exact FP64 operator: the tiled layout when asked for (or, in a mixed
    // solve, when built: it is the fastest exact one), else FP64 values, else coded
    L.prec64 = free_ ? HD_F64
               : (o.precision == HD_TILED || (mixed && h->d_tg_base)) ? HD_TILED
               : (o.precision == HD_CODED || !h->d_val64) ? (h->d_code ? HD_CODED : HD_TILED)
                                                          : HD_F64;
    HD_CHECK(L.alloc(mixed, mixed && prec32 == HD_F32));
    const double tol = o.tol > 0 ? o.tol : 1e-10;
    const double budget = o.f32_budget > 0 ? std::min(o.f32_budget, 1.0) : 1.0;
    // relative rounding of one FP32-input product: x rounded to FP32, plus
    // FP32 values and accumulation over a row for the CRS form
    const double u32 = 5.9604644775390625e-8;  // 2^-24
    const double eps32 = prec32 == HD_TILED_X32 ? u32 : u32 * (1.0 + std::sqrt((double)std::max<int64_t>(h->max_row, 1)));
    const int max_restarts = o.max_restarts > 0 ? o.max_restarts : 2000;
    int prec = L.prec64;
    // HD_LANCZOS_TRACE=1: one line per restart cycle on stderr
    const char* trace_env = std::getenv("HD_LANCZOS_TRACE");
    const bool trace = trace_env && trace_env[0] == '1';
    int kk_keep = nev + (m - nev) / 2;
    kk_keep = std::min(std::max(kk_keep, nev + 1), m - 2 >= nev ? m - 2 : nev);
    kk_keep = std::max(kk_keep, nev);
    if (mixed) HD_CHECK(dalloc_t(&L.bak, (size_t)(kk_keep + 1) * L.n, "mixed-precision restart backup"));

    // start vector
    if (o.start == HD_START_USER) {
        if (!o.d_start) return fail(HD_ERR_ARG, "HD_START_USER needs d_start");
        HD_HIP(hipMemcpyAsync(L.vec(0), o.d_start, sizeof(T) * L.n, hipMemcpyDeviceToDevice, L.st));
    } else {
        HD_CHECK(L.random_start(o.seed, 0u, L.vec(0)));
    }
    uint32_t rng_stream = 1;
    double nrm;
    HD_CHECK(L.norm2_of(L.vec(0), &nrm));
    nrm = std::sqrt(nrm);
    if (!(nrm > 0)) return fail(HD_ERR_ARG, "start vector is zero");
    HD_CHECK(L.scale_into(L.vec(0), 1.0 / nrm, L.vec(0)));

    std::vector<double> alpha(m, 0.0), beta(m, 0.0), theta, arrow;
    std::vector<double> Tm, w, Y;
    std::vector<T> c;
    int k = 0;  // kept Ritz vectors at the head of V
    int restarts = 0;
    bool converged = false;
    std::vector<double> res_est(m, 0.0);
    double anorm = 0.0;
    // Partial reorthogonalisation (H. D. Simon, Math. Comp. 42 (1984) 115):
    // om_cur[i] ~ |<v_j, v_i>| and om_prev[i] ~ |<v_{j-1}, v_i>| are propagated
    // by the recurrence the basis obeys, A V = V T + rounding, with T the arrow
    // of the kept Ritz pairs followed by the tridiagonal of the cycle; a new
    // vector is orthogonalised against the whole basis only when an estimate
    // exceeds sqrt(eps), and then the next one too (its recurrence still holds
    // the old error).  The first vector of a cycle (the arrow) and the last
    // (which becomes the next cycle's first) are always orthogonalised, so every
    // cycle starts from an orthogonal pair.
    const bool pro = o.full_reorth == 0;
    const double eps = std::numeric_limits<double>::epsilon();
    // a pass removes coefficients of the size of the estimates, which the
    // recurrence (and so the Ritz vectors' residuals) then misses: keeping them
    // below tol keeps the residuals at tol (semi-orthogonality, sqrt(eps), at most)
    const double om_max = std::min(std::sqrt(eps), o.tol > 0 ? o.tol : 1e-10);
    std::vector<double> om_prev(m + 1, 0.0), om_cur(m + 1, 0.0), om_next(m + 1, 0.0);
    bool force_next = false;
    int reorth_steps = 0;
    // mixed precision: bound of the rounding each wanted pair has taken from
    // FP32 cycles, and its coefficient norm on the previous cycle's new vectors
    std::vector<double> err32(nev, 0.0), nb_prev(nev, 1.0);
    // An FP32 cycle is speculative: the restart state it starts from is kept,
    // and when the cycle's own coefficients put a wanted pair's rounding
    // estimate beyond the budget (a pair that entered the wanted window during
    // the cycle carries O(1) coefficients on its vectors), the cycle is
    // discarded and run again with the FP64 operator.
    std::vector<double> theta_b, arrow_b, err32_b, nb_prev_b;
    int rollbacks = 0;
    // Projected matrix of the first n basis vectors: diag(theta) + arrow at row
    // k, tridiagonal below.
    auto projected = [&](int n) {
        Tm.assign((size_t)n * n, 0.0);
        for (int i = 0; i < k; ++i) {
            Tm[(size_t)i * n + i] = theta[i];
            Tm[(size_t)i * n + k] = Tm[(size_t)k * n + i] = arrow[i];
        }
        for (int j = k; j < n; ++j) {
            Tm[(size_t)j * n + j] = alpha[j];
            if (j + 1 < n) Tm[(size_t)j * n + j + 1] = Tm[(size_t)(j + 1) * n + j] = beta[j];
        }
    };
    // Convergence inside a cycle: once the previous cycle ended with every
    // wanted Ritz estimate within 1e4 tol, the Ritz pairs of the basis built so
    // far are tested after each step (a (j+1)-square eigenproblem on the host),
    // and the cycle ends at the first step where all wanted pairs meet tol.  An
    // FP32 cycle ends there only if its rounding estimate stays in the budget.
    double r_prev = std::numeric_limits<double>::infinity();
    int mc = m;  // basis size at which the current cycle ends
    for (;;) {
        const bool cyc32 = prec == prec32 && mixed;
        const bool check_in_cycle = r_prev <= 1e4 * tol;
        mc = m;
        if (cyc32) {
            HD_HIP(hipMemcpyAsync(L.bak, L.vec(0), sizeof(T) * L.n * (size_t)(k + 1), hipMemcpyDeviceToDevice, L.st));
            theta_b = theta;
            arrow_b = arrow;
            err32_b = err32;
            nb_prev_b = nb_prev;
        }
        // A rounding error of the FP32-input product is a vector A*delta with
        // delta ~ eps32 |x| spread over the sector: its overlap with a basis
        // vector is ~ eps32 |A| / sqrt(dim).  Such a cycle tolerates overlaps up
        // to eps32 (counted in err32 below).
        const double eps_op = cyc32 ? eps + eps32 / std::sqrt((double)std::max<int64_t>(L.n, 1)) : eps;
        const double om_cycle = cyc32 ? std::max(om_max, eps32) : om_max;
        for (int j = k; j < m; ++j) {
            HD_CHECK(L.apply(L.vec(j), L.W, prec));
            double b;
            if (j == k) {
                // first vector of a cycle: its coupling to the kept Ritz
                // vectors is the arrow, removed by two full CGS passes
                HD_CHECK(L.orthogonalize(j + 1, L.W, c, &b, true));
                alpha[j] = Lanczos<C>::host_re(c[j]);
                ++reorth_steps;
                std::fill(om_cur.begin(), om_cur.end(), eps);  // v_k: orthogonal to the kept vectors
                om_cur[j] = 1.0;
                std::fill(om_next.begin(), om_next.end(), eps);
                force_next = false;
            } else {
                double a, wn2 = 0.0;
                HD_CHECK(L.local_step(L.vec(j - 1), beta[j - 1], L.vec(j), L.W, &a, pro ? &wn2 : nullptr));
                bool full = !pro || force_next || j == m - 1;
                force_next = false;
                if (!full) {
                    const double bj = std::sqrt(std::max(wn2, 0.0));
                    anorm = std::max(anorm, std::fabs(a) + bj);
                    if (!(bj > 1e-12 * std::max(anorm, 1e-300))) {
                        full = true;  // breakdown: handled below with the orthogonalised w
                    } else {
                        // om_next[i] ~ <v_{j+1}, v_i>, i < j, from
                        // b_j v_{j+1} = A v_j - a_j v_j - b_{j-1} v_{j-1}
                        double worst = 0.0;
                        const double noise = 2.0 * eps_op * anorm / bj;
                        // Real: Simon's signed recurrence.  Complex: the overlaps are
                        // complex, so their magnitudes are propagated by the triangle
                        // inequality (an upper bound).
                        auto f = [](double x) { return C ? std::fabs(x) : x; };
                        for (int i = 0; i < j; ++i) {
                            double tv;  // (A v_i)^H v_j through the projected matrix, diagonal term apart
                            double dii;
                            if (i < k) {
                                dii = theta[i];
                                tv = f(arrow[i]) * om_cur[k];
                            } else if (i == k) {
                                dii = alpha[k];
                                tv = f(beta[k]) * om_cur[k + 1];
                                for (int l = 0; l < k; ++l) tv += f(arrow[l]) * om_cur[l];
                            } else {
                                dii = alpha[i];
                                tv = f(beta[i - 1]) * om_cur[i - 1] + f(beta[i]) * om_cur[i + 1];
                            }
                            const double v = (tv + f(dii - a) * om_cur[i] + (C ? 1.0 : -1.0) * beta[j - 1] * om_prev[i]) / bj;
                            om_next[i] = v + (v >= 0 ? noise : -noise);
                            worst = std::max(worst, std::fabs(om_next[i]));
                        }
                        om_next[j] = eps_op * anorm / bj;
                        if (worst > om_cycle) {
                            full = true;
                            force_next = true;
                        } else {
                            alpha[j] = a;
                            b = bj;
                        }
                    }
                }
                if (full) {
                    HD_CHECK(L.orthogonalize(j + 1, L.W, c, &b, false));
                    alpha[j] = a + Lanczos<C>::host_re(c[j]);
                    ++reorth_steps;
                    std::fill(om_next.begin(), om_next.end(), eps);
                }
            }
            anorm = std::max(anorm, std::fabs(alpha[j]) + b);
            if (b <= 1e-12 * std::max(anorm, 1e-300)) {
                // invariant subspace: continue with a fresh random direction
                HD_CHECK(L.random_start(o.seed, rng_stream++, L.W));
                double bb;
                std::vector<T> cc;
                HD_CHECK(L.orthogonalize(j + 1, L.W, cc, &bb));
                HD_CHECK(L.scale_into(L.W, 1.0 / bb, L.vec(j + 1)));
                beta[j] = 0.0;
                std::fill(om_next.begin(), om_next.end(), eps);
                force_next = true;
            } else {
                beta[j] = b;
                HD_CHECK(L.scale_into(L.W, 1.0 / b, L.vec(j + 1)));
            }
            om_next[j + 1] = 1.0;
            std::swap(om_prev, om_cur);
            std::swap(om_cur, om_next);
            const int n1 = j + 1;
            if (check_in_cycle && n1 < m && n1 > k + 1 && n1 >= nev && beta[j] > 0.0) {
                projected(n1);
                if (sym_eig(n1, Tm, w, Y)) {
                    bool done = true;
                    for (int i = 0; i < nev && done; ++i) {
                        const double scale_i = std::max(1.0, std::fabs(w[i]));
                        if (!(std::fabs(beta[j] * Y[(size_t)(n1 - 1) * n1 + i]) / scale_i <= tol)) done = false;
                        if (done && cyc32) {
                            double nb2 = 0.0;
                            for (int jj = k; jj < n1; ++jj) nb2 += Y[(size_t)jj * n1 + i] * Y[(size_t)jj * n1 + i];
                            if (err32[i] + eps32 * anorm * std::sqrt(nb2) > budget * tol * scale_i) done = false;
                        }
                    }
                    if (done) {
                        mc = n1;
                        break;
                    }
                }
            }
        }
        projected(mc);
        if (!sym_eig(mc, Tm, w, Y)) return fail(HD_ERR_NOT_CONVERGED, "projected eigenproblem did not converge");
        const double bm = beta[mc - 1];
        bool all = true;
        bool allow32 = mixed;
        bool overshoot = false;
        for (int i = 0; i < nev; ++i) {
            const double scale_i = std::max(1.0, std::fabs(w[i]));
            res_est[i] = std::fabs(bm * Y[(size_t)(mc - 1) * mc + i]) / scale_i;
            if (mixed) {
                // the wanted Ritz vector's coefficients on this cycle's new vectors
                double nb2 = 0.0;
                for (int j = k; j < mc; ++j) nb2 += Y[(size_t)j * mc + i] * Y[(size_t)j * mc + i];
                const double nb = std::sqrt(nb2);
                if (cyc32) {
                    err32[i] += eps32 * anorm * nb;
                    if (err32[i] > budget * tol * scale_i) overshoot = true;
                }
                // the next cycle's coefficients, predicted from this cycle's decrease
                const double pred = nb * std::min(1.0, nb / std::max(nb_prev[i], 1e-300));
                nb_prev[i] = nb;
                if (err32[i] + eps32 * anorm * pred > budget * tol * scale_i) allow32 = false;
            }
            if (!(res_est[i] <= tol)) all = false;
        }
        r_prev = 0.0;
        for (int i = 0; i < nev; ++i) r_prev = std::max(r_prev, res_est[i]);
        ++restarts;
        if (trace) {
            double rmax = 0.0, emax = 0.0;
            for (int i = 0; i < nev; ++i) {
                rmax = std::max(rmax, res_est[i]);
                emax = std::max(emax, err32[i] / std::max(1.0, std::fabs(w[i])));
            }
            std::fprintf(stderr, "[hd_lanczos] cycle %d op %d k %d ritz-resid %.3e fp32-err %.3e next-fp32 %d anorm %.3e full-passes %d%s\n",
                         restarts, prec, k, rmax, emax, (int)(mixed && allow32 && !overshoot), anorm, reorth_steps,
                         overshoot ? " (discarded: run again in FP64)" : "");
        }
        if (overshoot) {
            HD_HIP(hipMemcpyAsync(L.vec(0), L.bak, sizeof(T) * L.n * (size_t)(k + 1), hipMemcpyDeviceToDevice, L.st));
            theta = theta_b;
            arrow = arrow_b;
            err32 = err32_b;
            nb_prev = nb_prev_b;
            prec = L.prec64;
            ++rollbacks;
            if (restarts >= max_restarts) break;
            continue;
        }
        if (all) {
            converged = true;
            break;
        }
        if (restarts >= max_restarts) break;
        // thick restart: keep kk Ritz vectors, continue from the residual direction
        const int kk = kk_keep;
        HD_CHECK(L.combine(mc, Y, mc, kk, L.tmp));
        HD_HIP(hipMemcpyAsync(L.vec(kk), L.vec(mc), sizeof(T) * L.n, hipMemcpyDeviceToDevice, L.st));
        HD_HIP(hipMemcpyAsync(L.vec(0), L.tmp, sizeof(T) * L.n * kk, hipMemcpyDeviceToDevice, L.st));
        theta.assign(w.begin(), w.begin() + kk);
        arrow.resize(kk);
        for (int i = 0; i < kk; ++i) arrow[i] = bm * Y[(size_t)(mc - 1) * mc + i];
        k = kk;
        if (mixed) prec = allow32 ? prec32 : L.prec64;
    }
    // Ritz vectors and FP64 residuals
    hd_lanczos_state* S = h->lz;
    if (!S) S = h->lz = new hd_lanczos_state();
    dfree(S->d_vec);
    S->d_vec = nullptr;
    S->dim = L.n;
    S->cplx = C;
    S->nev = nev;
    HD_CHECK(dalloc(&S->d_vec, sizeof(T) * (size_t)L.n * nev, "eigenvectors"));
    HD_CHECK(L.combine(mc, Y, mc, nev, static_cast<T*>(S->d_vec)));
    S->evals.assign(w.begin(), w.begin() + nev);
    for (int i = 0; i < nev; ++i) evals[i] = w[i];
    int nconv = 0;
    if (stats) {
        for (int i = 0; i < 64; ++i) stats->residual[i] = 0;
    }
    for (int i = 0; i < nev; ++i) {
        T* xi = static_cast<T*>(S->d_vec) + (size_t)i * L.n;
        HD_CHECK(L.apply(xi, L.W, L.prec64));
        double r;
        HD_CHECK(L.residual_norm(L.W, xi, w[i], &r));
        r /= std::max(1.0, std::fabs(w[i]));
        if (trace) std::fprintf(stderr, "[hd_lanczos] pair %d: theta %.15g FP64 residual %.3e (Ritz estimate %.3e)\n", i, w[i], r, res_est[i]);
        if (stats && i < 64) stats->residual[i] = r;
        if (r <= tol) ++nconv;
    }
    if (stats) {
        stats->converged = converged && nconv == nev;
        stats->nconv = nconv;
        stats->restarts = restarts;
        stats->matvecs = L.mv32 + L.mv64;
        stats->matvecs_f32 = L.mv32;
        stats->matvecs_f64 = L.mv64;
        stats->time_spmv_ms = L.t_spmv;
        stats->time_ortho_ms = L.t_ortho;
        stats->reorth_steps = reorth_steps;
        stats->f32_cycles_discarded = rollbacks;
        stats->time_total_ms =
            std::chrono::duration<double, std::milli>(std::chrono::steady_clock::now() - t_start).count();
    }
    if (!converged) return fail(HD_ERR_NOT_CONVERGED, "Lanczos reached the restart limit before tol");
    return HD_OK;
}

}  // namespace hd

using namespace hd;

extern "C" {

int hd_lanczos(hd_hamiltonian* h, const hd_lanczos_opts* opts, double* h_evals, hd_lanczos_stats* stats)
{
    if (!h || !opts || !h_evals) return fail(HD_ERR_ARG, "hd_lanczos: null argument");
    if (opts->nev <= 0) return fail(HD_ERR_ARG, "nev must be positive");
    if (h->storage != HD_STORAGE_FULL && h->storage != HD_STORAGE_MATRIX_FREE)
        return fail(HD_ERR_ARG, "Lanczos needs full or matrix-free storage");
    if (h->storage == HD_STORAGE_MATRIX_FREE && (opts->precision != HD_F64 || opts->block > 1))
        return fail(HD_ERR_ARG, "matrix-free storage runs the single-vector Lanczos in FP64 (precision HD_F64)");
    if (!h->lanes_per_row) HD_CHECK(spmv_autotune(h));
    if (opts->block > 1) return run_block_lanczos(h, *opts, h_evals, stats);
    auto run = [&](const hd_lanczos_opts& o, hd_lanczos_stats* s) {
        return h->is_complex ? run_lanczos<true>(h, o, h_evals, s) : run_lanczos<false>(h, o, h_evals, s);
    };
    hd_lanczos_stats s1{};
    const int rc = run(*opts, &s1);
    const bool mixed = opts->precision == HD_F32;
    if ((rc != HD_OK && rc != HD_ERR_NOT_CONVERGED) || s1.converged || (opts->full_reorth && !mixed)) {
        if (stats) *stats = s1;
        return rc;
    }
    // the Ritz estimates met tol but an explicit FP64 residual did not: solve
    // again with the FP64 operator (mixed precision) or with full
    // reorthogonalisation (partial reorthogonalisation)
    hd_lanczos_opts o2 = *opts;
    if (mixed) o2.precision = h->d_tg_base ? HD_TILED : (h->d_val64 ? HD_F64 : HD_CODED);
    else o2.full_reorth = 1;
    hd_lanczos_stats s2{};
    const int rc2 = run(o2, &s2);
    s2.matvecs += s1.matvecs;
    s2.matvecs_f32 += s1.matvecs_f32;
    s2.matvecs_f64 += s1.matvecs_f64;
    s2.restarts += s1.restarts;
    s2.reorth_steps += s1.reorth_steps;
    s2.f32_cycles_discarded += s1.f32_cycles_discarded;
    s2.time_total_ms += s1.time_total_ms;
    s2.time_spmv_ms += s1.time_spmv_ms;
    s2.time_ortho_ms += s1.time_ortho_ms;
    if (stats) *stats = s2;
    return rc2;
}

int hd_lanczos_vectors(const hd_hamiltonian* h, int32_t i, void* d_out, void* h_out)
{
    if (!h || !h->lz || !h->lz->d_vec) return fail(HD_ERR_ARG, "no Lanczos result on this handle");
    if (i < 0 || i >= h->lz->nev) return fail(HD_ERR_ARG, "eigenvector index out of range");
    const size_t es = 8 * (h->lz->cplx ? 2 : 1);
    const char* src = static_cast<const char*>(h->lz->d_vec) + (size_t)i * h->lz->dim * es;
    if (d_out) HD_HIP(hipMemcpy(d_out, src, h->lz->dim * es, hipMemcpyDeviceToDevice));
    else if (h_out) HD_HIP(hipMemcpy(h_out, src, h->lz->dim * es, hipMemcpyDeviceToHost));
    else return fail(HD_ERR_ARG, "no output buffer");
    return HD_OK;
}

int hd_full_diagonalize(hd_hamiltonian* h, int32_t nev, double* h_evals)
{
    if (!h || !h_evals || nev <= 0) return fail(HD_ERR_ARG, "hd_full_diagonalize: bad argument");
    if (h->is_complex) return fail(HD_ERR_UNSUPPORTED, "dense fallback handles real matrices; use hd_lanczos");
    if (h->dim > 4096) return fail(HD_ERR_ARG, "dense fallback limited to dim <= 4096");
    const int64_t n = h->dim;
    std::vector<int64_t> rp(n + 1);
    std::vector<int32_t> col((size_t)h->nnz);
    std::vector<double> val((size_t)h->nnz);
    HD_CHECK(hd_ham_copy_crs(h, 0, n, HD_F64, rp.data(), col.data(), val.data()));
    std::vector<double> A((size_t)n * n, 0.0), w, Z;
    for (int64_t i = 0; i < n; ++i)
        for (int64_t j = rp[i]; j < rp[i + 1]; ++j) {
            A[(size_t)i * n + col[j]] = val[j];
            if (h->storage == HD_STORAGE_UPPER_REF && col[j] != i) A[(size_t)col[j] * n + i] = val[j];
        }
    if (!sym_eig((int)n, A, w, Z)) return fail(HD_ERR_NOT_CONVERGED, "dense eigenproblem did not converge");
    for (int i = 0; i < std::min<int64_t>(nev, n); ++i) h_evals[i] = w[i];
    return HD_OK;
}

}  // extern "C"
