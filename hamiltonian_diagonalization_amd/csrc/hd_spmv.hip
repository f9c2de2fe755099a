// hd_spmv.hip -- y = H x over the full (both triangles) CRS.
//
// Reference: MatrixVectorFunction::Simple -> ParallelSymmetricMatrixVectorMultiply
// -> SymmetricMatrixVectorMultiply, the naive loop over the upper triangle U
// applying y = U x + (U - D)^H x (src/utilities/mathematics/sparse_linear_algebra.hpp:142-165,
// :218-312).  The scattered y[col] += ... of the transposed half cannot run
// race-free in parallel, so the device keeps both triangles and every row is a
// pure gather: LPR lanes of a wavefront stream the row's column indices and
// values (non-temporal, so the resident x vector keeps the caches), gather x,
// and reduce with cross-lane shuffles.  Consecutive rows have near-identical
// column sets, so each XCD is given a contiguous range of row blocks and its
// private L2 sees the x entries those rows share.
#include <hip/hip_runtime.h>

#include <algorithm>

#include "hd_internal.hpp"

namespace hd {

template <typename V> struct Acc;
template <> struct Acc<double> { using T = double; };
template <> struct Acc<float> { using T = float; };
template <> struct Acc<double2> { using T = double2; };
template <> struct Acc<float2> { using T = float2; };

HD_DEV double fmadd(double v, double x, double a) { return fma(v, x, a); }
HD_DEV float fmadd(float v, float x, float a) { return fmaf(v, x, a); }
HD_DEV double2 fmadd(double2 v, double2 x, double2 a)
{
    a.x = fma(v.x, x.x, a.x);
    a.x = fma(-v.y, x.y, a.x);
    a.y = fma(v.x, x.y, a.y);
    a.y = fma(v.y, x.x, a.y);
    return a;
}
HD_DEV float2 fmadd(float2 v, float2 x, float2 a)
{
    a.x = fmaf(v.x, x.x, a.x);
    a.x = fmaf(-v.y, x.y, a.x);
    a.y = fmaf(v.x, x.y, a.y);
    a.y = fmaf(v.y, x.x, a.y);
    return a;
}
template <typename T> HD_DEV T zero_of();
template <> HD_DEV double zero_of<double>() { return 0.0; }
template <> HD_DEV float zero_of<float>() { return 0.0f; }
template <> HD_DEV double2 zero_of<double2>() { return make_double2(0.0, 0.0); }
template <> HD_DEV float2 zero_of<float2>() { return make_float2(0.f, 0.f); }

HD_DEV double shfl_xor(double v, int m) { return __shfl_xor(v, m, 64); }
HD_DEV float shfl_xor(float v, int m) { return __shfl_xor(v, m, 64); }
HD_DEV double2 shfl_xor(double2 v, int m) { return make_double2(__shfl_xor(v.x, m, 64), __shfl_xor(v.y, m, 64)); }
HD_DEV float2 shfl_xor(float2 v, int m) { return make_float2(__shfl_xor(v.x, m, 64), __shfl_xor(v.y, m, 64)); }
HD_DEV double add(double a, double b) { return a + b; }
HD_DEV float add(float a, float b) { return a + b; }
HD_DEV double2 add(double2 a, double2 b) { return make_double2(a.x + b.x, a.y + b.y); }
HD_DEV float2 add(float2 a, float2 b) { return make_float2(a.x + b.x, a.y + b.y); }

template <typename T> HD_DEV T ntload(const T* p) { return __builtin_nontemporal_load(p); }
template <> HD_DEV double2 ntload<double2>(const double2* p)
{
    const double* q = reinterpret_cast<const double*>(p);
    return make_double2(__builtin_nontemporal_load(q), __builtin_nontemporal_load(q + 1));
}
template <> HD_DEV float2 ntload<float2>(const float2* p)
{
    const float* q = reinterpret_cast<const float*>(p);
    return make_float2(__builtin_nontemporal_load(q), __builtin_nontemporal_load(q + 1));
}

constexpr int kSpmvBlock = 256;

// Logical block index with the blocks of each XCD (dispatched round-robin,
// block b on XCD b % 8) mapped to one contiguous range.
HD_DEV int64_t xcd_block(int64_t b, int64_t nb)
{
    const int64_t x = b & 7, i = b >> 3, q = nb >> 3, r = nb & 7;
    return x * q + (x < r ? x : r) + i;
}

template <typename V, int LPR>
__global__ void __launch_bounds__(kSpmvBlock)
k_spmv(int64_t dim, const int64_t* __restrict__ rowptr, const int32_t* __restrict__ col,
       const V* __restrict__ val, const V* __restrict__ x, V* __restrict__ y)
{
    using A = typename Acc<V>::T;
    constexpr int RPB = kSpmvBlock / LPR;
    const int64_t lb = xcd_block(blockIdx.x, gridDim.x);
    const int64_t row = lb * RPB + threadIdx.x / LPR;
    const int sub = threadIdx.x & (LPR - 1);
    A acc0 = zero_of<A>(), acc1 = zero_of<A>();
    if (row < dim) {
        const int64_t start = rowptr[row], end = rowptr[row + 1];
        int64_t j = start + sub;
        for (; j + LPR < end; j += 2 * LPR) {
            const int c0 = ntload(col + j), c1 = ntload(col + j + LPR);
            const V v0 = ntload(val + j), v1 = ntload(val + j + LPR);
            acc0 = fmadd(v0, x[c0], acc0);
            acc1 = fmadd(v1, x[c1], acc1);
        }
        if (j < end) acc0 = fmadd(ntload(val + j), x[ntload(col + j)], acc0);
    }
    A acc = add(acc0, acc1);
#pragma unroll
    for (int m = LPR / 2; m > 0; m >>= 1) acc = add(acc, shfl_xor(acc, m));
    if (row < dim && sub == 0) y[row] = acc;
}

template <typename V, int LPR>
static void launch_one(int64_t dim, const hd_hamiltonian* h, const void* val, const void* x, void* y,
                       hipStream_t s)
{
    constexpr int RPB = kSpmvBlock / LPR;
    const int64_t nb = (dim + RPB - 1) / RPB;
    hipLaunchKernelGGL((k_spmv<V, LPR>), dim3((unsigned)nb), dim3(kSpmvBlock), 0, s, dim, h->d_rowptr,
                       h->d_col, static_cast<const V*>(val), static_cast<const V*>(x), static_cast<V*>(y));
}

template <typename V>
static void launch_lpr(int lpr, int64_t dim, const hd_hamiltonian* h, const void* val, const void* x,
                       void* y, hipStream_t s)
{
    switch (lpr) {
    case 4: launch_one<V, 4>(dim, h, val, x, y, s); break;
    case 8: launch_one<V, 8>(dim, h, val, x, y, s); break;
    case 16: launch_one<V, 16>(dim, h, val, x, y, s); break;
    case 32: launch_one<V, 32>(dim, h, val, x, y, s); break;
    default: launch_one<V, 64>(dim, h, val, x, y, s); break;
    }
}

static int pick_default_lanes(const hd_hamiltonian* h)
{
    const double avg = h->dim ? (double)h->nnz / (double)h->dim : 0.0;
    if (avg > 96) return 64;
    if (avg > 48) return 32;
    if (avg > 24) return 16;
    if (avg > 12) return 8;
    return 4;
}

int spmv_launch(const hd_hamiltonian* h, int precision, const void* x, void* y, hipStream_t stream)
{
    if (h->storage != HD_STORAGE_FULL)
        return fail(HD_ERR_ARG, "SpMV needs full (both-triangle) storage");
    const void* val = precision == HD_F32 ? h->d_val32 : h->d_val64;
    if (!val) return fail(HD_ERR_ARG, "requested SpMV precision has no resident values");
    const int lpr = h->lanes_per_row ? h->lanes_per_row : pick_default_lanes(h);
    if (h->dim == 0) return HD_OK;
    if (precision == HD_F32) {
        if (h->is_complex) launch_lpr<float2>(lpr, h->dim, h, val, x, y, stream);
        else launch_lpr<float>(lpr, h->dim, h, val, x, y, stream);
    } else {
        if (h->is_complex) launch_lpr<double2>(lpr, h->dim, h, val, x, y, stream);
        else launch_lpr<double>(lpr, h->dim, h, val, x, y, stream);
    }
    HD_HIP(hipGetLastError());
    return HD_OK;
}

// Times every lanes-per-row variant on the matrix and keeps the fastest
// (FP32 values when resident, else FP64).
int spmv_autotune(hd_hamiltonian* h)
{
    const int prec = h->d_val32 ? HD_F32 : HD_F64;
    const size_t es = (prec == HD_F32 ? 4 : 8) * (h->is_complex ? 2 : 1);
    void *x = nullptr, *y = nullptr;
    HD_CHECK(dalloc(&x, (size_t)h->dim * es, "autotune x"));
    int st = dalloc(&y, (size_t)h->dim * es, "autotune y");
    if (st != HD_OK) {
        dfree(x);
        return st;
    }
    hipMemsetAsync(x, 0, (size_t)h->dim * es, h->stream);
    hipEvent_t e0, e1;
    hipEventCreate(&e0);
    hipEventCreate(&e1);
    const int cand[5] = {4, 8, 16, 32, 64};
    float best = 1e30f;
    int best_l = 64;
    for (int l : cand) {
        h->lanes_per_row ? (void)0 : (void)0;
        const_cast<hd_hamiltonian*>(h)->lanes_per_row = l;
        spmv_launch(h, prec, x, y, h->stream);  // warm
        hipEventRecord(e0, h->stream);
        for (int r = 0; r < 3; ++r) spmv_launch(h, prec, x, y, h->stream);
        hipEventRecord(e1, h->stream);
        hipEventSynchronize(e1);
        float ms = 0;
        hipEventElapsedTime(&ms, e0, e1);
        if (ms < best) {
            best = ms;
            best_l = l;
        }
    }
    h->lanes_per_row = best_l;
    hipEventDestroy(e0);
    hipEventDestroy(e1);
    dfree(x);
    dfree(y);
    HD_HIP(hipGetLastError());
    return HD_OK;
}

}  // namespace hd

using namespace hd;

extern "C" {

int hd_spmv(const hd_hamiltonian* h, int32_t precision, const void* d_x, void* d_y)
{
    if (!h || !d_x || !d_y) return fail(HD_ERR_ARG, "hd_spmv: null argument");
    return spmv_launch(h, precision, d_x, d_y, h->stream);
}

int hd_spmv_host(const hd_hamiltonian* h, const void* h_x, void* h_y)
{
    if (!h || !h_x || !h_y) return fail(HD_ERR_ARG, "hd_spmv_host: null argument");
    const size_t bytes = (size_t)h->dim * 8 * (h->is_complex ? 2 : 1);
    void *x = nullptr, *y = nullptr;
    HD_CHECK(dalloc(&x, bytes, "Simple() x"));
    int st = dalloc(&y, bytes, "Simple() y");
    if (st == HD_OK) {
        hipMemcpyAsync(x, h_x, bytes, hipMemcpyHostToDevice, h->stream);
        st = spmv_launch(h, HD_F64, x, y, h->stream);
        if (st == HD_OK) {
            hipMemcpyAsync(h_y, y, bytes, hipMemcpyDeviceToHost, h->stream);
            if (hipStreamSynchronize(h->stream) != hipSuccess) st = fail(HD_ERR_HIP, "Simple() copy failed");
        }
    }
    dfree(x);
    dfree(y);
    return st;
}

int hd_spmv_set_lanes(hd_hamiltonian* h, int32_t lanes)
{
    if (!h) return fail(HD_ERR_ARG, "null handle");
    if (lanes == 0) return spmv_autotune(h);
    if (lanes != 4 && lanes != 8 && lanes != 16 && lanes != 32 && lanes != 64)
        return fail(HD_ERR_ARG, "lanes per row must be 4, 8, 16, 32 or 64");
    h->lanes_per_row = lanes;
    return HD_OK;
}

int hd_spmv_get_lanes(const hd_hamiltonian* h, int32_t* lanes)
{
    if (!h || !lanes) return fail(HD_ERR_ARG, "null argument");
    *lanes = h->lanes_per_row ? h->lanes_per_row : pick_default_lanes(h);
    return HD_OK;
}

}  // extern "C"
