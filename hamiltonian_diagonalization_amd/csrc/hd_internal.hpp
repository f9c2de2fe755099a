// hd_internal.hpp -- shared definitions of libhdiag_hip (not part of the ABI).
#pragma once

#include <hip/hip_runtime.h>

#include <cstdint>
#include <cstdio>
#include <map>
#include <memory>
#include <string>
#include <vector>

#include "../../include/hdiag.h"

namespace hd {

// ------------------------------------------------------------------ errors

void set_error(const std::string& msg);
int fail(int code, const std::string& msg);

#define HD_HIP(expr)                                                                   \
    do {                                                                               \
        hipError_t _e = (expr);                                                        \
        if (_e != hipSuccess)                                                          \
            return ::hd::fail(HD_ERR_HIP, std::string(#expr) + ": " +                 \
                                              hipGetErrorString(_e));                  \
    } while (0)

#define HD_CHECK(expr)                      \
    do {                                    \
        int _s = (expr);                    \
        if (_s != HD_OK) return _s;         \
    } while (0)

// A pair of HIP events timing work on one stream.  The first HIP error of the
// sequence is kept (later calls are skipped) and reported by status(); the
// events are released by the destructor.
struct EventTimer {
    hipEvent_t e0 = nullptr, e1 = nullptr;
    hipError_t err = hipSuccess;
    EventTimer()
    {
        note(hipEventCreate(&e0));
        note(hipEventCreate(&e1));
    }
    ~EventTimer()
    {
        if (e0) (void)hipEventDestroy(e0);
        if (e1) (void)hipEventDestroy(e1);
    }
    EventTimer(const EventTimer&) = delete;
    EventTimer& operator=(const EventTimer&) = delete;
    void note(hipError_t e)
    {
        if (err == hipSuccess) err = e;
    }
    void start(hipStream_t s)
    {
        if (err == hipSuccess) note(hipEventRecord(e0, s));
    }
    // Waits for the work recorded since start(); 0 ms after an error.
    float stop(hipStream_t s)
    {
        float ms = 0.f;
        if (err == hipSuccess) note(hipEventRecord(e1, s));
        if (err == hipSuccess) note(hipEventSynchronize(e1));
        if (err == hipSuccess) note(hipEventElapsedTime(&ms, e0, e1));
        return err == hipSuccess ? ms : 0.f;
    }
    int status(const char* what) const
    {
        return err == hipSuccess ? HD_OK : fail(HD_ERR_HIP, std::string(what) + ": " + hipGetErrorString(err));
    }
};

// The first failure of a run of enqueued HIP calls (copies, fills, launches)
// whose status is checked once, after the run: HD_ENQ(q, hipMemcpyAsync(...)).
struct HipErrors {
    hipError_t err = hipSuccess;
    const char* what = "";
    void operator()(hipError_t e, const char* w)
    {
        if (err == hipSuccess && e != hipSuccess) {
            err = e;
            what = w;
        }
    }
    int status() const
    {
        return err == hipSuccess ? HD_OK : fail(HD_ERR_HIP, std::string(what) + ": " + hipGetErrorString(err));
    }
};
#define HD_ENQ(q, expr) (q)((expr), #expr)

// Device allocation that reports HD_ERR_NOMEM with the size.
int dalloc(void** p, size_t bytes, const char* what);
template <typename T>
int dalloc_t(T** p, size_t count, const char* what)
{
    return dalloc(reinterpret_cast<void**>(p), count * sizeof(T), what);
}
void dfree(void* p);

// -------------------------------------------------------- sector rank table

// Counting table of the sector: W[p][c][g] = number of subsets of orbitals
// {0..p-1} with c elements whose weight is g.  Two families:
//   LZ:     weight(k) = k (integer sum, the target is sum k = (lzTot+N*maxLz)/2),
//           stored compactly per (p,c) over g in [gmin, gmax].
//   finite: weight in a finite abelian group (Z_X x Z_Y for K2D, Z_Y for K1D,
//           trivial for NONE), dense over g, subtraction through a table.
// Layout of the device blob (int32 words), copied into LDS by the kernels:
//   [0] N  [1] Norb  [2] G  [3] target  [4] is_lz  [5] tab_len
//   off [(Norb+1)*(N+1)]    offset of (p,c) into tab, minus gmin
//   gmm [(Norb+1)*(N+1)*2]  (gmin, gmax); gmax < gmin marks an empty row
//   sub [G*Norb]            finite groups: g - weight(k); LZ kinds: weight(k) [Norb]
//   tab [tab_len]           uint32 counts
struct RankTable {
    int N = 0, Norb = 0, G = 1, target = 0, is_lz = 0;
    std::vector<int32_t> blob;
    uint64_t dim = 0;
    bool empty = true;
    size_t bytes() const { return blob.size() * sizeof(int32_t); }
};

// Builds the table; returns HD_OK, HD_ERR_ARG or HD_ERR_UNSUPPORTED (counts
// beyond 32 bits).  An empty sector gives rt.empty = true.
int make_rank_table(const hd_sector& sec, int N, int Norb, RankTable& rt);
// Group weight helpers on the host (used to test hop sector preservation).
int sector_weight_delta_ok(const RankTable& rt, const hd_sector& sec, int a, int b, int c,
                           int d);
// The same test for all pairs at once: sums[a*Norb + b] is the group sum of the
// weights of orbitals a and b, so a hop {a,b} -> {c,d} keeps the sector exactly
// when sums[a*Norb+b] == sums[c*Norb+d] (all zero without a sector constraint).
int sector_pair_sums(const RankTable& rt, const hd_sector& sec, std::vector<int>& sums);

constexpr int kRankHeader = 6;
// is_lz value of a basis given as a list of states (HD_SECTOR_LIST): the
// blob is the header followed by the device address of the ascending states
// (2 words) and their count (2 words); ranking is a binary search.
constexpr int kRankList = 2;

// ------------------------------------------------------------ device side

#define HD_DEV __device__ __forceinline__

// The hardware hands consecutive workgroup ids to the 8 XCDs in turn, and each
// XCD has its own L2.  For a grid of n blocks this maps block b to the work
// item it should take so that XCD k runs one contiguous range of items, in
// order: neighbouring row groups, which read mostly the same x lines, then
// share one L2.  The map is a bijection of [0, n).
constexpr int kXcds = 8;
HD_DEV int64_t xcd_block(int64_t b, int64_t n)
{
    const int64_t q = n / kXcds, rem = n % kXcds, k = b % kXcds;
    return k * q + (k < rem ? k : rem) + b / kXcds;
}
// The same within windows of kXcds * win blocks (the last window may be
// partial): every XCD takes `win` consecutive items of each window, which keeps
// the items in flight across the chip close together.  win <= 0 keeps the
// launch order; win >= ceil(n / kXcds) is one window over the whole grid.
HD_DEV int64_t xcd_window(int64_t b, int64_t n, int64_t win)
{
    if (win <= 0) return b;
    const int64_t W = kXcds * win, base = (b / W) * W;
    return base + xcd_block(b - base, n - base < W ? n - base : W);
}

// Layout of the split ranking tables (hd_build.hip: make_split_rank, split_rank).
struct SplitArgs {
    int words;   // LDS words of the tables
    int m1, m2;  // middle orbitals [m1, m2)
    int N;
    int oH, oPk, oSub, oTab;  // word offsets of H, the middle's (p, c) ranges, weights, counts
};

struct RankView {
    const int32_t* off;
    const int32_t* gmm;
    const int32_t* sub;
    const uint32_t* tab;
    int N, Norb, G, target, is_lz;
    const uint64_t* list;  // kRankList: the states, ascending
    int64_t list_dim;
};

HD_DEV RankView rank_view(const int32_t* blob)
{
    RankView v;
    v.N = blob[0];
    v.Norb = blob[1];
    v.G = blob[2];
    v.target = blob[3];
    v.is_lz = blob[4];
    const int npc = (v.Norb + 1) * (v.N + 1);
    v.off = blob + kRankHeader;
    v.gmm = v.off + npc;
    v.sub = v.gmm + 2 * npc;
    v.tab = reinterpret_cast<const uint32_t*>(v.sub + (v.is_lz ? v.Norb : v.G * v.Norb));
    v.list = nullptr;
    v.list_dim = 0;
    if (v.is_lz == kRankList) {
        const uint32_t* w = reinterpret_cast<const uint32_t*>(blob + kRankHeader);
        v.list = reinterpret_cast<const uint64_t*>(((uint64_t)w[1] << 32) | w[0]);
        v.list_dim = (int64_t)(((uint64_t)w[3] << 32) | w[2]);
    }
    return v;
}

// Binary search of s in the list of a kRankList basis.
HD_DEV bool list_rank(const RankView& v, uint64_t s, uint64_t& r_out)
{
    int64_t lo = 0, hi = v.list_dim;
    while (lo < hi) {
        const int64_t mid = (lo + hi) >> 1;
        if (v.list[mid] < s) lo = mid + 1;
        else hi = mid;
    }
    if (lo < v.list_dim && v.list[lo] == s) {
        r_out = (uint64_t)lo;
        return true;
    }
    return false;
}

template <bool LZ>
HD_DEV uint32_t rank_lookup(const RankView& v, int p, int c, int g)
{
    const int i = p * (v.N + 1) + c;
    const int lo = v.gmm[2 * i], hi = v.gmm[2 * i + 1];
    if (g < lo || g > hi) return 0u;
    return v.tab[v.off[i] + g];
}

template <bool LZ>
HD_DEV int group_sub(const RankView& v, int g, int k)
{
    if (LZ) return g - v.sub[k];  // integer weights: sub holds w(k)
    return v.sub[g * v.Norb + k];
}

// Index of `s` within the sector in ascending order; returns false when s is
// not a sector state (wrong particle number, bits beyond Norb or wrong weight).
template <bool LZ>
HD_DEV bool rank_state(const RankView& v, uint64_t s, uint64_t& r_out)
{
    if (v.is_lz == kRankList) return list_rank(v, s, r_out);
    if (__popcll(s) != v.N) return false;
    if (v.Norb < 64 && (s >> v.Norb) != 0) return false;
    uint64_t r = 0;
    int c = v.N, g = v.target;
    uint64_t t = s;
    while (t) {
        const int p = 63 - __clzll(t);
        r += rank_lookup<LZ>(v, p, c, g);
        g = group_sub<LZ>(v, g, p);
        --c;
        t ^= 1ull << p;
    }
    r_out = r;
    return g == 0;
}

// Sector state of index r (ascending order); r < dim.
template <bool LZ>
HD_DEV uint64_t unrank_state(const RankView& v, uint64_t r)
{
    uint64_t s = 0;
    int c = v.N, g = v.target;
    for (int p = v.Norb - 1; p >= 0 && c > 0; --p) {
        const uint64_t z = rank_lookup<LZ>(v, p, c, g);
        if (r >= z) {
            r -= z;
            s |= 1ull << p;
            g = group_sub<LZ>(v, g, p);
            --c;
        }
    }
    return s;
}

HD_DEV uint32_t shfl_xor_key(uint32_t v, int m) { return (uint32_t)__shfl_xor((int)v, m); }
HD_DEV uint64_t shfl_xor_key(uint64_t v, int m)
{
    const int lo = __shfl_xor((int)(uint32_t)v, m);
    const int hi = __shfl_xor((int)(uint32_t)(v >> 32), m);
    return ((uint64_t)(uint32_t)hi << 32) | (uint32_t)lo;
}

// Ascending bitonic sort of a wavefront's 512 keys, lane l holding elements
// 8l .. 8l+7 in registers: distances below 8 within a lane, the others
// between lanes (shfl_xor).
template <typename Key>
HD_DEV void sort512_regs(Key (&key)[8], int lane)
{
#pragma unroll
    for (int k = 2; k <= 512; k <<= 1) {
#pragma unroll
        for (int j = k >> 1; j > 0; j >>= 1) {
            if (j >= 8) {
                const int lj = j >> 3;
                const bool lower = (lane & lj) == 0;
#pragma unroll
                for (int t = 0; t < 8; ++t) {
                    const Key other = shfl_xor_key(key[t], lj);
                    const bool up = ((lane * 8 + t) & k) == 0;
                    key[t] = (lower == up) ? (key[t] < other ? key[t] : other) : (key[t] > other ? key[t] : other);
                }
            } else {
#pragma unroll
                for (int t = 0; t < 8; ++t) {
                    if (t & j) continue;
                    const int t2 = t | j;
                    const bool up = ((lane * 8 + t) & k) == 0;
                    const Key a = key[t], b = key[t2];
                    const Key lo = a < b ? a : b, hi = a < b ? b : a;
                    key[t] = up ? lo : hi;
                    key[t2] = up ? hi : lo;
                }
            }
        }
    }
}

// Ascending bitonic sort of v[0, P) (P a power of two, P <= 8 * THREADS) by a
// block of THREADS threads, in registers: thread t holds elements 8t .. 8t+7.
// Stages of distance below 8 run within a thread, those below 512 between the
// lanes of a wavefront (shfl_xor), and only the longer ones exchange through
// v, with a barrier.  Elements from P to 8 * THREADS take part (as a separate
// network) and must hold values that may move among themselves (all equal).
template <typename Key, int THREADS>
HD_DEV void bitonic_blocked(Key* v, int P)
{
    const int t = threadIdx.x;
    Key a[8];
#pragma unroll
    for (int e = 0; e < 8; ++e) a[e] = v[8 * t + e];
    for (int k = 2; k <= P; k <<= 1) {
        for (int j = k >> 1; j >= 8; j >>= 1) {
            if (j >= 512) {
                __syncthreads();
#pragma unroll
                for (int e = 0; e < 8; ++e) v[8 * t + e] = a[e];
                __syncthreads();
#pragma unroll
                for (int e = 0; e < 8; ++e) {
                    const int i = 8 * t + e;
                    const Key b = v[i ^ j];
                    const bool keep_min = ((i & j) == 0) == ((i & k) == 0);
                    a[e] = keep_min ? (a[e] < b ? a[e] : b) : (a[e] < b ? b : a[e]);
                }
            } else {
                const int d = j >> 3;
#pragma unroll
                for (int e = 0; e < 8; ++e) {
                    const int i = 8 * t + e;
                    const Key b = shfl_xor_key(a[e], d);
                    const bool keep_min = ((i & j) == 0) == ((i & k) == 0);
                    a[e] = keep_min ? (a[e] < b ? a[e] : b) : (a[e] < b ? b : a[e]);
                }
            }
        }
#pragma unroll
        for (int j = 4; j >= 1; j >>= 1) {
            if (2 * j > k) continue;
#pragma unroll
            for (int e = 0; e < 8; ++e) {
                if (e & j) continue;
                const bool asc = ((8 * t + e) & k) == 0;
                const Key lo = a[e], hi = a[e | j];
                if ((lo > hi) == asc) {
                    a[e] = hi;
                    a[e | j] = lo;
                }
            }
        }
    }
    __syncthreads();
#pragma unroll
    for (int e = 0; e < 8; ++e) v[8 * t + e] = a[e];
    __syncthreads();
}

// Cooperative copy of n 32-bit words into LDS by the whole block.
HD_DEV void block_copy_words(int32_t* dst, const int32_t* src, int n)
{
    for (int i = threadIdx.x; i < n; i += blockDim.x) dst[i] = src[i];
}

// complex helpers (interleaved re, im)
HD_DEV double2 cadd(double2 a, double2 b) { return make_double2(a.x + b.x, a.y + b.y); }
HD_DEV float2 cadd(float2 a, float2 b) { return make_float2(a.x + b.x, a.y + b.y); }

int device_props(int* n_cu, size_t* max_lds);

}  // namespace hd

// ------------------------------------------------------------- handle types

struct hd_basis {
    int device = 0;
    hipStream_t stream = nullptr;
    hd_sector sector{};
    int N = 0, Norb = 0;
    int64_t dim = 0;
    hd::RankTable rt;
    int32_t* d_rank = nullptr;  // blob
    uint64_t* d_states = nullptr;
};

struct hd_tables {
    int layout = 0, K = 0, is_complex = 0;
    std::vector<uint16_t> ktab;
    std::vector<double> vtab;  // interleaved when complex
    std::vector<double> eps;   // empty when no quadratic table
    // Two-level tables: orbitals >= split form the first level, those below
    // the second; the diagonal sums the first level's terms before the
    // second's.  0 for one level.
    int split = 0;
    // HD_LAYOUT_HASH: k1 lists per (k2,k3,k4) at entry (k4*K + k3)*K + k2 and
    // per k2, with one value per list slot (interleaved when complex)
    int max_k1 = 0, max_q1 = 0, has_quad = 0;
    std::vector<uint8_t> nk1, k1, nq1, q1;
    std::vector<double> hv, he;
};

// Eigenpairs of the last Lanczos call on a Hamiltonian handle.
struct hd_lanczos_state {
    int64_t dim = 0;
    int cplx = 0;
    int nev = 0;
    void* d_vec = nullptr;  // nev vectors, contiguous
    std::vector<double> evals;
};

struct hd_hamiltonian {
    int device = 0;
    hipStream_t stream = nullptr;
    const hd_basis* basis = nullptr;
    int64_t dim = 0, nnz = 0, n_trimmed = 0, max_row = 0;
    int storage = HD_STORAGE_FULL;
    int is_complex = 0;
    int64_t* d_rowptr = nullptr;
    int32_t* d_col = nullptr;
    void* d_val64 = nullptr;  // double or double2
    void* d_val32 = nullptr;  // float or float2
    // coded values: entry = sign * dict[code & 0x7FFF], code 0xFFFF = diag[row]
    uint16_t* d_code = nullptr;
    void* d_diag = nullptr;   // double or double2, dim entries
    void* d_dict = nullptr;   // double or double2, ndict entries
    int ndict = 0;
    int lanes_per_row = 0;    // SpMV lanes per row (0 = pick on first use)
    int spmv_sched[4] = {-1, -1, -1, -1};  // SpMV schedule per precision (hd_precision); -1 = default
    // tiled coded layout (hd_tiled.hip): row groups, x lines staged in LDS
    int64_t tg_groups = 0, tg_entries = 0, tg_lines = 0, tg_offs = 0;
    int tg_rows = 0;  // rows per group (128, or 64 for denser matrices)
    int tg_entry_bytes = 4;  // 4, or 3 for packed entries (dictionary <= 2047 values)
    int64_t tg_win = 0;      // block order of the tiled product (groups per XCD window)
    size_t tg_bytes = 0;     // HBM of the tiled layout, included in `bytes`
    int64_t* d_tg_base = nullptr;  // entry, line and offset-table bases, 3 x (groups + 1)
    uint32_t* d_tg_entries = nullptr;
    uint32_t* d_tg_lines = nullptr;
    uint32_t* d_tg_offs = nullptr;
    double* d_tg_diag = nullptr;
    double t_tile = 0;
    // matrix-free storage: pair table, pair values and the full-storage
    // diagonal (d_diag); the basis handle supplies states and ranking
    uint8_t* d_mf_kpair = nullptr;
    void* d_mf_pval = nullptr;  // double or double2
    int mf_K = 0, mf_kp_bytes = 0;
    uint64_t mf_norb_mask = 0;
    int32_t* d_mf_split = nullptr;  // split ranking tables (null: rank table walk)
    hd::SplitArgs mf_split{};
    double t_count = 0, t_scan = 0, t_fill = 0, t_sort = 0;
    size_t bytes = 0;
    hd_lanczos_state* lz = nullptr;  // eigenvectors of the last Lanczos call
};

namespace hd {
// Hop-value tuples of a batch of tables (one value per table), shared by all
// sectors of the batch, and the per-table diagonal terms.
struct BatchTables {
    int K = 0, n = 0, has_eps = 0;
    std::map<std::vector<uint64_t>, int> index;  // tuple bits -> dictionary index
    std::vector<double> dict;                    // tuple i at [i*n, (i+1)*n)
    std::vector<double> dterm;                   // [((k4*K + k3)*2 + q)*n + p], 0.0 where skipped
    std::vector<double> eps;                     // [k*n + p]
    // merged pair tables already derived for this batch, by the sectors' pair
    // sums (sector_pair_sums): sectors with equal sums keep the same hops
    std::vector<std::vector<int>> merged_keys;
    std::vector<std::shared_ptr<void>> merged;
};
// Coded full-storage pattern of one sector for a batch of real tables
// (hd_build.hip); the codes index bt's tuples.
int build_batch_sector(const hd_basis* b, const hd_tables* const* tables, int n, double tol, BatchTables& bt,
                       hd_hamiltonian** out);
int spmv_launch(const hd_hamiltonian* h, int precision, const void* x, void* y,
                hipStream_t stream);
// Exclusive scan of n int32 counts into n+1 int64 row pointers on `stream`,
// with the total and the largest count returned to the host (hd_build.hip).
int scan_counts(const int32_t* d_counts, int64_t n, int64_t* d_rowptr, int64_t* total, int* max_count,
                hipStream_t stream);
// Build of HD_LAYOUT_HASH tables: full or upper storage, FP64/FP32 values
// (hd_general.hip).
int general_build(hd_hamiltonian* h, const hd_basis* b, const hd_tables* t, double tol, int value_mask);
// Matrix-free FP64 product of a HD_STORAGE_MATRIX_FREE handle (hd_build.hip).
int matrix_free_launch(const hd_hamiltonian* h, const void* x, void* y, hipStream_t stream);
// Tiled coded layout (hd_tiled.hip).
int tiled_build(hd_hamiltonian* h, int rows_per_group, int entry_bytes);
void tiled_free(hd_hamiltonian* h);
// x32: x in FP32 (HD_TILED_X32), y in FP64; real matrices.
int tiled_launch(const hd_hamiltonian* h, const void* x, void* y, hipStream_t s, bool x32 = false);
size_t tiled_bytes(const hd_hamiltonian* h);
// Cyclic Jacobi on a symmetric row-major matrix (destroyed): ascending
// eigenvalues w, eigenvectors as the columns of Z (hd_lanczos.hip).
void jacobi_eigh(int n, std::vector<double>& A, std::vector<double>& w, std::vector<double>& Z);
// Householder + implicit QL, same contract (hd_eigsym.cpp); false if QL fails.
bool sym_eig(int n, std::vector<double>& A, std::vector<double>& w, std::vector<double>& Z);
// Block thick-restart Lanczos for opts.block > 1 (hd_block.hip).
int run_block_lanczos(hd_hamiltonian* h, const hd_lanczos_opts& o, double* evals, hd_lanczos_stats* stats);
int spmv_autotune(hd_hamiltonian* h);
void lanczos_state_free(hd_lanczos_state* s);
}  // namespace hd
