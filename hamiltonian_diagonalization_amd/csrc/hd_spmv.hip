// hd_spmv.hip -- y = H x over the full (both triangles) CRS.
//
// Reference: MatrixVectorFunction::Simple -> ParallelSymmetricMatrixVectorMultiply
// -> SymmetricMatrixVectorMultiply, the naive loop over the upper triangle U
// applying y = U x + (U - D)^H x (src/utilities/mathematics/sparse_linear_algebra.hpp:142-165,
// :218-312).  The scattered y[col] += ... of the transposed half cannot run
// race-free in parallel, so the device keeps both triangles and every row is a
// pure gather: LPR lanes of a wavefront stream the row's column indices and
// values with scalar non-temporal loads (the stream is read once and must not
// evict x from L2/MALL), gather x, and reduce with cross-lane shuffles.  One
// block handles one chunk of NT/LPR rows; (LPR, NT) is autotuned per matrix.
// Measured on MI355X the kernel is bound by the L2-miss traffic of the x
// gathers, not by the stream, so the block -> chunk order is chosen for L2
// reuse (window_chunk); 16-byte vector loads, persistent grids and four
// entries per lane in flight (FP64 8.9 -> 11.1 ms, complex 1.76 -> 1.86 ms:
// more gathers in flight evict more lines) were slower.
//
// Coded format (HD_CODED): every off-diagonal value is sign * P[hop] from the
// build's pair table, and the distinct P values are few (1,496 for the
// N=12, 2S=33 sphere), so an entry is stored as int32 column + uint16 code
// (bit 15 = sign, low 15 bits = dictionary index, 0xFFFF = the row's diagonal
// from a separate FP64 array).  The FP64 dictionary lives in LDS.  That is the
// exact FP64 operator in 6 bytes per entry instead of 12.
#include <hip/hip_runtime.h>

#include <mutex>

#include <algorithm>

#include "hd_internal.hpp"

namespace hd {

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
HD_DEV double negv(double a) { return -a; }
HD_DEV double2 negv(double2 a) { return make_double2(-a.x, -a.y); }

// scalar non-temporal loads (complex values as two component loads)
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

// Block -> chunk order.  Blocks are dispatched in index order and block b runs
// on XCD b % 8.  Within each window of 8 * kWinG blocks, XCD x takes kWinG
// consecutive chunks, so every XCD's L2 sees a contiguous band of rows whose
// column sets overlap; the window keeps the rows in flight across the chip
// narrow.  (Measured on the N=12, 2S=33 sphere: 9.1 ms against 10.6 ms for the
// plain order and 12.6 ms for a persistent XCD-partitioned walk.)
constexpr int64_t kWinG = 1024;

HD_DEV int64_t window_chunk(int64_t b, int64_t nb)
{
    const int64_t W = 8 * kWinG;
    const int64_t base = (b / W) * W;
    if (base + W > nb) return b;  // the last partial window keeps the plain order
    const int64_t o = b - base;
    return base + (o & 7) * kWinG + (o >> 3);
}

// y = H x, one block per chunk of NT / LPR rows, LPR lanes per row.
template <typename V, int LPR, int NT>
__global__ void __launch_bounds__(NT)
k_spmv(int64_t dim, const int64_t* __restrict__ rowptr, const int32_t* __restrict__ col,
       const V* __restrict__ val, const V* __restrict__ x, V* __restrict__ y)
{
    constexpr int RPB = NT / LPR;
    const int sub = threadIdx.x & (LPR - 1);
    const int64_t row = window_chunk(blockIdx.x, gridDim.x) * RPB + threadIdx.x / LPR;
    V acc = zero_of<V>();
    if (row < dim) {
        const int64_t start = rowptr[row], end = rowptr[row + 1];
        for (int64_t j = start + sub; j < end; j += LPR)
            acc = fmadd(ntload(val + j), x[ntload(col + j)], acc);
    }
#pragma unroll
    for (int m = LPR / 2; m > 0; m >>= 1) acc = add(acc, shfl_xor(acc, m));
    if (row < dim && sub == 0) y[row] = acc;
}

// Coded values: the block copies the dictionary into LDS, then as k_spmv.  A
// dictionary above kLdsDictBytes (random complex couplings give tens of
// thousands of distinct values) is read in place instead: at most 32767
// entries, 512 KiB complex, it stays resident in each XCD's 4 MiB L2.
constexpr size_t kLdsDictBytes = 64 * 1024;

template <typename V, int LPR, int NT, bool LDS_DICT>
__global__ void __launch_bounds__(NT)
k_spmv_coded(int64_t dim, const int64_t* __restrict__ rowptr, const int32_t* __restrict__ col,
             const uint16_t* __restrict__ code, const V* __restrict__ diag, const V* __restrict__ dict,
             int ndict, const V* __restrict__ x, V* __restrict__ y)
{
    extern __shared__ int32_t lds_raw[];
    const V* sdict = dict;
    if constexpr (LDS_DICT) {
        V* t = reinterpret_cast<V*>(lds_raw);
        for (int i = threadIdx.x; i < ndict; i += NT) t[i] = dict[i];
        __syncthreads();
        sdict = t;
    }
    constexpr int RPB = NT / LPR;
    const int sub = threadIdx.x & (LPR - 1);
    const int64_t row = window_chunk(blockIdx.x, gridDim.x) * RPB + threadIdx.x / LPR;
    V acc = zero_of<V>();
    if (row < dim) {
        const int64_t start = rowptr[row], end = rowptr[row + 1];
        for (int64_t j = start + sub; j < end; j += LPR) {
            const uint32_t q = ntload(code + j);
            V v;
            if (q == 0xFFFFu) v = diag[row];
            else {
                v = sdict[q & 0x7FFFu];
                if (q & 0x8000u) v = negv(v);
            }
            acc = fmadd(v, x[ntload(col + j)], acc);
        }
    }
#pragma unroll
    for (int m = LPR / 2; m > 0; m >>= 1) acc = add(acc, shfl_xor(acc, m));
    if (row < dim && sub == 0) y[row] = acc;
}

// Schedules (lanes per row, threads per block) the autotune chooses from.
struct Sched {
    int lpr, nt;
};
constexpr Sched kScheds[] = {{4, 256},  {8, 256},  {16, 256}, {16, 512}, {32, 256},
                             {32, 512}, {32, 1024}, {64, 512}, {64, 1024}};
constexpr int kNumScheds = (int)(sizeof(kScheds) / sizeof(kScheds[0]));

template <typename V, int LPR, int NT>
static void launch_one(const hd_hamiltonian* h, const void* val, const void* x, void* y, hipStream_t s)
{
    constexpr int RPB = NT / LPR;
    const unsigned grid = (unsigned)((h->dim + RPB - 1) / RPB);
    hipLaunchKernelGGL((k_spmv<V, LPR, NT>), dim3(grid), dim3(NT), 0, s, h->dim, h->d_rowptr, h->d_col,
                       static_cast<const V*>(val), static_cast<const V*>(x), static_cast<V*>(y));
}

template <typename V, int LPR, int NT>
static void launch_coded(const hd_hamiltonian* h, const void* x, void* y, hipStream_t s)
{
    constexpr int RPB = NT / LPR;
    const unsigned grid = (unsigned)((h->dim + RPB - 1) / RPB);
    const size_t lds = sizeof(V) * (size_t)std::max(h->ndict, 1);
    if (lds <= kLdsDictBytes)
        hipLaunchKernelGGL((k_spmv_coded<V, LPR, NT, true>), dim3(grid), dim3(NT), lds, s, h->dim, h->d_rowptr,
                           h->d_col, h->d_code, static_cast<const V*>(h->d_diag),
                           static_cast<const V*>(h->d_dict), h->ndict, static_cast<const V*>(x),
                           static_cast<V*>(y));
    else
        hipLaunchKernelGGL((k_spmv_coded<V, LPR, NT, false>), dim3(grid), dim3(NT), 0, s, h->dim, h->d_rowptr,
                           h->d_col, h->d_code, static_cast<const V*>(h->d_diag),
                           static_cast<const V*>(h->d_dict), h->ndict, static_cast<const V*>(x),
                           static_cast<V*>(y));
}

template <typename V, bool CODED>
static void launch_sched(int si, const hd_hamiltonian* h, const void* val, const void* x, void* y, hipStream_t s)
{
#define HD_SCHED(I, L, T)                                   \
    case I:                                                 \
        if constexpr (CODED) launch_coded<V, L, T>(h, x, y, s); \
        else launch_one<V, L, T>(h, val, x, y, s);          \
        break;
    switch (si) {
        HD_SCHED(0, 4, 256)
        HD_SCHED(1, 8, 256)
        HD_SCHED(2, 16, 256)
        HD_SCHED(3, 16, 512)
        HD_SCHED(4, 32, 256)
        HD_SCHED(6, 32, 1024)
        HD_SCHED(7, 64, 512)
        HD_SCHED(8, 64, 1024)
    default:  // 5: 32 lanes, 512 threads
        if constexpr (CODED) launch_coded<V, 32, 512>(h, x, y, s);
        else launch_one<V, 32, 512>(h, val, x, y, s);
        break;
    }
#undef HD_SCHED
}

// schedule index for a lanes-per-row request: 512 threads from 16 lanes up
static int sched_for_lanes(int lpr)
{
    const int nt = lpr >= 16 ? 512 : 256;
    for (int i = 0; i < kNumScheds; ++i)
        if (kScheds[i].lpr == lpr && kScheds[i].nt == nt) return i;
    return 5;
}

static int pick_default_sched(const hd_hamiltonian* h)
{
    const double avg = h->dim ? (double)h->nnz / (double)h->dim : 0.0;
    if (avg > 96) return 5;  // 32 lanes, 512 threads
    if (avg > 48) return 3;
    if (avg > 24) return 1;
    return 0;
}

int spmv_launch(const hd_hamiltonian* h, int precision, const void* x, void* y, hipStream_t stream)
{
    if (h->storage == HD_STORAGE_MATRIX_FREE) {
        if (precision != HD_F64) return fail(HD_ERR_ARG, "matrix-free storage computes the product in FP64 only");
        return matrix_free_launch(h, x, y, stream);
    }
    if (h->storage != HD_STORAGE_FULL)
        return fail(HD_ERR_ARG, "SpMV needs full (both-triangle) storage");
    if (precision < 0 || precision > HD_TILED_X32) return fail(HD_ERR_ARG, "unknown SpMV precision");
    if (precision == HD_TILED || precision == HD_TILED_X32) return tiled_launch(h, x, y, stream, precision == HD_TILED_X32);
    const int si = h->spmv_sched[precision] >= 0 ? h->spmv_sched[precision] : pick_default_sched(h);
    if (h->dim == 0) return HD_OK;
    if (precision == HD_CODED) {
        if (!h->d_code) return fail(HD_ERR_ARG, "no coded values resident (value_mask & 4)");
        if (h->ndict > 0x7FFF) return fail(HD_ERR_UNSUPPORTED, "value dictionary above 32767 entries");
        if (h->is_complex) launch_sched<double2, true>(si, h, nullptr, x, y, stream);
        else launch_sched<double, true>(si, h, nullptr, x, y, stream);
        HD_HIP(hipGetLastError());
        return HD_OK;
    }
    const void* val = precision == HD_F32 ? h->d_val32 : h->d_val64;
    if (!val) return fail(HD_ERR_ARG, "requested SpMV precision has no resident values");
    if (precision == HD_F32) {
        if (h->is_complex) launch_sched<float2, false>(si, h, val, x, y, stream);
        else launch_sched<float, false>(si, h, val, x, y, stream);
    } else {
        if (h->is_complex) launch_sched<double2, false>(si, h, val, x, y, stream);
        else launch_sched<double, false>(si, h, val, x, y, stream);
    }
    HD_HIP(hipGetLastError());
    return HD_OK;
}

static int default_precision(const hd_hamiltonian* h)
{
    if (h->d_code) return HD_CODED;
    if (h->d_val32) return HD_F32;
    return HD_F64;
}

// Times every lanes-per-row variant on the matrix (in its fastest resident
// format) and keeps the fastest.
// Plain streaming copy, 16 bytes per lane per iteration (the byte-counter
// calibration kernel).
__global__ void __launch_bounds__(256) k_stream_copy(const int4* __restrict__ src, int4* __restrict__ dst, int64_t n)
{
    const int64_t stride = (int64_t)gridDim.x * blockDim.x;
    for (int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; i < n; i += stride) dst[i] = src[i];
}

int spmv_autotune(hd_hamiltonian* h)
{
    if (h->storage == HD_STORAGE_MATRIX_FREE) {  // one kernel, nothing to choose
        h->lanes_per_row = 1;
        return HD_OK;
    }
    // one schedule per resident value format, each timed on zero vectors
    const size_t es = 8 * (h->is_complex ? 2 : 1);
    void *x = nullptr, *y = nullptr;
    HD_CHECK(dalloc(&x, (size_t)h->dim * es, "autotune x"));
    int st = dalloc(&y, (size_t)h->dim * es, "autotune y");
    if (st != HD_OK) {
        dfree(x);
        return st;
    }
    HipErrors q;
    HD_ENQ(q, hipMemsetAsync(x, 0, (size_t)h->dim * es, h->stream));
    st = q.status();
    EventTimer timer;
    const bool have[3] = {h->d_val64 != nullptr, h->d_val32 != nullptr, h->d_code != nullptr};
    for (int prec = 0; prec < 3 && st == HD_OK; ++prec) {
        if (!have[prec]) continue;
        float best = 1e30f;
        int best_s = pick_default_sched(h);
        for (int si = 0; si < kNumScheds && st == HD_OK; ++si) {
            h->spmv_sched[prec] = si;
            if ((st = spmv_launch(h, prec, x, y, h->stream)) != HD_OK) break;  // warm
            timer.start(h->stream);
            for (int r = 0; r < 3 && st == HD_OK; ++r) st = spmv_launch(h, prec, x, y, h->stream);
            const float ms = timer.stop(h->stream);
            if (st == HD_OK) st = timer.status("SpMV autotune timing");
            if (st == HD_OK && ms < best) {
                best = ms;
                best_s = si;
            }
        }
        h->spmv_sched[prec] = best_s;
    }
    h->lanes_per_row = kScheds[h->spmv_sched[default_precision(h)] >= 0 ? h->spmv_sched[default_precision(h)]
                                                                        : pick_default_sched(h)].lpr;
    dfree(x);
    dfree(y);
    return st;
}

// Staging of hd_spmv_host_noalloc: one file-scope device array per GPU (each
// device holds its own copy for the life of the process), x in the first
// half and y in the second.  Host code never passes the array's name to a
// kernel or a copy: the name is a host-side stand-in, and its device address
// is looked up once per device with hipGetSymbolAddress.
constexpr size_t kNoAllocHalf = (size_t)32 << 20;  // bytes per staged vector
constexpr int kNoAllocMaxDevices = 64;
static __device__ __align__(256) unsigned char scratch[2 * kNoAllocHalf];
static std::mutex g_scratch_mutex;
static void* g_scratch_ptr[kNoAllocMaxDevices] = {};

// Device address of `scratch` on the current device `dev` (looked up on the
// first call for that device).
static int scratch_address(int dev, unsigned char** out)
{
    if (dev < 0 || dev >= kNoAllocMaxDevices) return fail(HD_ERR_ARG, "device index beyond the no-alloc table");
    std::lock_guard<std::mutex> lock(g_scratch_mutex);
    if (!g_scratch_ptr[dev]) {
        void* p = nullptr;
        const hipError_t e = hipGetSymbolAddress(&p, HIP_SYMBOL(scratch));
        if (e != hipSuccess || !p)
            return fail(HD_ERR_HIP, std::string("hipGetSymbolAddress(scratch): ") + hipGetErrorString(e));
        g_scratch_ptr[dev] = p;
    }
    *out = static_cast<unsigned char*>(g_scratch_ptr[dev]);
    return HD_OK;
}

// The FP64 operator Simple() applies: FP64 values, else coded, else tiled.
static int simple_precision(const hd_hamiltonian* h)
{
    return h->d_val64 ? HD_F64 : h->d_code ? HD_CODED : h->d_tg_base ? HD_TILED : -1;
}

}  // namespace hd

using namespace hd;

extern "C" {

int hd_spmv_host_noalloc(const hd_hamiltonian* h, const void* h_x, void* h_y)
{
    if (!h || !h_x || !h_y) return fail(HD_ERR_ARG, "hd_spmv_host_noalloc: null argument");
    const size_t bytes = (size_t)h->dim * 8 * (h->is_complex ? 2 : 1);
    if (bytes > kNoAllocHalf)
        return fail(HD_ERR_ARG, "hd_spmv_host_noalloc: a vector of " + std::to_string(bytes) +
                                    " bytes exceeds the 32 MiB staging buffer (dim <= 4194304 real, "
                                    "2097152 complex); use hd_spmv_host or hd_spmv");
    const int prec = simple_precision(h);
    if (prec < 0) return fail(HD_ERR_ARG, "Simple() needs FP64, coded or tiled values");
    int dev = -1;
    HD_HIP(hipGetDevice(&dev));
    if (dev != h->device) return fail(HD_ERR_ARG, "hd_spmv_host_noalloc: the handle's device must be current");
    unsigned char* base = nullptr;
    HD_CHECK(scratch_address(dev, &base));
    if (bytes == 0) return HD_OK;
    void* x = base;
    void* y = base + kNoAllocHalf;
    HipErrors q;
    HD_ENQ(q, hipMemcpyAsync(x, h_x, bytes, hipMemcpyHostToDevice, h->stream));
    HD_CHECK(q.status());
    HD_CHECK(spmv_launch(h, prec, x, y, h->stream));
    HD_ENQ(q, hipMemcpyAsync(h_y, y, bytes, hipMemcpyDeviceToHost, h->stream));
    HD_ENQ(q, hipStreamSynchronize(h->stream));
    return q.status();
}

int hd_spmv(const hd_hamiltonian* h, int32_t precision, const void* d_x, void* d_y)
{
    if (!h || !d_x || !d_y) return fail(HD_ERR_ARG, "hd_spmv: null argument");
    return spmv_launch(h, precision, d_x, d_y, h->stream);
}

int hd_spmv_host(const hd_hamiltonian* h, const void* h_x, void* h_y)
{
    if (!h || !h_x || !h_y) return fail(HD_ERR_ARG, "hd_spmv_host: null argument");
    const size_t bytes = (size_t)h->dim * 8 * (h->is_complex ? 2 : 1);
    const int prec = simple_precision(h);
    if (prec < 0) return fail(HD_ERR_ARG, "Simple() needs FP64, coded or tiled values");
    void *x = nullptr, *y = nullptr;
    HD_CHECK(dalloc(&x, bytes, "Simple() x"));
    int st = dalloc(&y, bytes, "Simple() y");
    if (st == HD_OK) {
        // copies and product in the order of the handle's stream, then one wait
        HipErrors q;
        HD_ENQ(q, hipMemcpyAsync(x, h_x, bytes, hipMemcpyHostToDevice, h->stream));
        st = q.status();
        if (st == HD_OK) st = spmv_launch(h, prec, x, y, h->stream);
        if (st == HD_OK) {
            HD_ENQ(q, hipMemcpyAsync(h_y, y, bytes, hipMemcpyDeviceToHost, h->stream));
            HD_ENQ(q, hipStreamSynchronize(h->stream));
            st = q.status();
        }
    }
    dfree(x);
    dfree(y);
    return st;
}

int hd_spmv_time(hd_hamiltonian* h, int32_t precision, int32_t iters, double* avg_ms)
{
    if (!h || !avg_ms || iters <= 0) return fail(HD_ERR_ARG, "hd_spmv_time: bad argument");
    if (!h->lanes_per_row) HD_CHECK(spmv_autotune(h));
    const size_t z = h->is_complex ? 2 : 1;
    const size_t ex = (precision == HD_F32 || precision == HD_TILED_X32 ? 4 : 8) * z;
    const size_t ey = (precision == HD_F32 ? 4 : 8) * z;
    void *x = nullptr, *y = nullptr;
    HD_CHECK(dalloc(&x, (size_t)h->dim * ex, "timing x"));
    int st = dalloc(&y, (size_t)h->dim * ey, "timing y");
    if (st != HD_OK) {
        dfree(x);
        return st;
    }
    if (hipMemsetAsync(x, 0, (size_t)h->dim * ex, h->stream) != hipSuccess) st = fail(HD_ERR_HIP, "timing fill failed");
    if (st == HD_OK) st = spmv_launch(h, precision, x, y, h->stream);  // warm
    EventTimer timer;
    timer.start(h->stream);
    for (int i = 0; i < iters && st == HD_OK; ++i) st = spmv_launch(h, precision, x, y, h->stream);
    const float ms = timer.stop(h->stream);
    if (st == HD_OK) st = timer.status("SpMV timing");
    *avg_ms = ms / iters;
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
    for (int& si : h->spmv_sched) si = sched_for_lanes(lanes);
    return HD_OK;
}

int hd_stream_copy(const void* d_src, void* d_dst, size_t bytes, int32_t iters, double* avg_ms)
{
    if (!d_src || !d_dst || !avg_ms || iters <= 0 || bytes == 0 || bytes % 16)
        return fail(HD_ERR_ARG, "hd_stream_copy: bad argument");
    const int64_t n = (int64_t)(bytes / 16);
    int dev = 0;
    HD_HIP(hipGetDevice(&dev));
    hipDeviceProp_t prop;
    HD_HIP(hipGetDeviceProperties(&prop, dev));
    const int grid = (int)std::min<int64_t>((n + 255) / 256, (int64_t)prop.multiProcessorCount * 16);
    EventTimer timer;
    timer.start(nullptr);
    for (int i = 0; i < iters; ++i)
        hipLaunchKernelGGL(k_stream_copy, dim3(grid), dim3(256), 0, nullptr, (const int4*)d_src, (int4*)d_dst, n);
    HD_HIP(hipGetLastError());
    const float ms = timer.stop(nullptr);
    HD_CHECK(timer.status("copy timing"));
    *avg_ms = ms / iters;
    return HD_OK;
}

int hd_spmv_get_lanes(const hd_hamiltonian* h, int32_t* lanes)
{
    if (!h || !lanes) return fail(HD_ERR_ARG, "null argument");
    const int si = h->spmv_sched[default_precision(h)];
    *lanes = kScheds[si >= 0 ? si : pick_default_sched(h)].lpr;
    return HD_OK;
}

int hd_spmv_get_schedule(const hd_hamiltonian* h, int32_t precision, int32_t* lanes, int32_t* threads)
{
    if (!h || !lanes || !threads || precision < 0 || precision > HD_TILED_X32) return fail(HD_ERR_ARG, "bad argument");
    if (precision == HD_TILED || precision == HD_TILED_X32) {  // one schedule: 1024 threads per row group
        *lanes = h->tg_rows ? 1024 / h->tg_rows : 0;
        *threads = 1024;
        return HD_OK;
    }
    const int si = h->spmv_sched[precision] >= 0 ? h->spmv_sched[precision] : pick_default_sched(h);
    *lanes = kScheds[si].lpr;
    *threads = kScheds[si].nt;
    return HD_OK;
}

}  // extern "C"
