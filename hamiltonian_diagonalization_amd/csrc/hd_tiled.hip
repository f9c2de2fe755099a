// hd_tiled.hip -- the coded operator re-laid out for x reuse through LDS.
//
// The CRS SpMV is bound by its x gathers: every 8-byte gather pulls a whole
// cache line from L2 into L1, and a row's ~330 columns lie in ~265 distinct
// 128-byte lines.  Consecutive rows are numerically adjacent states that differ
// in their low orbitals, and one hop applied to each of them lands on nearby
// ranks, so 64 consecutive rows together touch only 0.06-0.08 distinct lines
// per stored entry (N=12, 2S=33 sphere).  The tiled layout makes that reuse
// explicit:
//
//   group    = 64 consecutive rows (one 1024-thread block, 16 lanes per row)
//   lines    = the sorted distinct x lines (16 doubles) the group's off-diagonal
//              entries read, split into chunks of 256 lines (32 KiB of LDS)
//   entry    = uint32: bits 24-31 line slot within the chunk, 20-23 word
//              within the line, 0-19 value code (bit 19 sign, bits 0-18
//              dictionary index into the coded format's FP64 dictionary);
//              with at most 2047 dictionary values the entries are packed
//              into 3 bytes (slot 16-23, word 12-15, sign 11, index 0-10)
//   chunks   = 256 consecutive lines; per chunk the 64 row starts and the
//              chunk end (the entries of (chunk c, row r) are
//              contiguous and in the row's column order; each chunk's block
//              is padded to a multiple of 4 entries), then the chunk line
//              starts: 66 words per chunk plus one
//   diagonal = FP64 per row, applied after the off-diagonal sum
//
// The SpMV block stages each chunk's lines into LDS with coalesced 16-byte
// loads (the next chunk's lines are in flight while the current one is
// processed), then every lane walks its row's entries of that chunk reading x
// and the dictionary from LDS.  HBM carries 3 or 4 bytes per entry plus the line lists
// and offsets; the operator is the coded format's exact FP64 matrix.  Complex
// (Hermitian) matrices use the same layout with 8 values per line (a 3-bit word
// field) and read their larger dictionary from L2 (k_spmv_tiled_c).
//
// The layout is built on the device from the coded CRS, one block per group:
// an LDS hash set collects the group's lines, a bitonic sort orders them,
// per-(chunk, row) counts give the row starts, and each entry is written to
// its slot.  Every write is
// bounds-checked against the group's counts.
#include <hip/hip_runtime.h>

#include <algorithm>
#include <type_traits>
#include <vector>

#include "hd_internal.hpp"

namespace hd {

namespace {

constexpr int kTC = 256;           // lines per chunk
constexpr int kLineWords = 16;     // doubles per 128-byte line
constexpr int kHashSlots = 8192;   // LDS hash set of a group's lines
constexpr int kMaxLines = 5120;    // distinct lines per group (load factor 0.625)
constexpr int kMaxChunks = kMaxLines / kTC;
constexpr int kU = 8;              // entries each SpMV lane has in flight
constexpr uint32_t kEmpty = 0xFFFFFFFFu;
constexpr int kBuildThreads = 512;
constexpr int kLU = 8;             // entries each build thread loads ahead
constexpr int kSpmvThreads = 1024;  // one block per group: kSpmvThreads / TG lanes per row

__device__ __forceinline__ uint32_t line_hash(uint32_t line)
{
    return (line * 2654435761u) >> 19;  // 13 bits
}

// Inserts `line` (if absent) and returns its slot, or -1 when the probe
// sequence found no room (bounded by the table size).
__device__ __forceinline__ int hash_insert(uint32_t* keys, uint32_t line)
{
    uint32_t h = line_hash(line);
    for (int probe = 0; probe < kHashSlots; ++probe) {
        uint32_t old = keys[h];  // most lines are already in: no atomic for them
        if (old == kEmpty) old = atomicCAS(&keys[h], kEmpty, line);
        if (old == kEmpty || old == line) return (int)h;
        h = (h + 1) & (kHashSlots - 1);
    }
    return -1;
}

__device__ __forceinline__ int hash_find(const uint32_t* keys, uint32_t line)
{
    uint32_t h = line_hash(line);
    for (int probe = 0; probe < kHashSlots; ++probe) {
        const uint32_t k = keys[h];
        if (k == line) return (int)h;
        if (k == kEmpty) return -1;
        h = (h + 1) & (kHashSlots - 1);
    }
    return -1;
}

// The plan of one row group, shared by both build passes: the sorted distinct
// lines, their chunks (kTC consecutive lines each), entry counts per (chunk,
// row), the per-chunk row starts T (chunk blocks padded to 4 entries) and the
// row's entries in earlier chunks (cum).
template <int TG>
struct TilePlanLds {
    uint32_t keys[kHashSlots];
    uint32_t sorted[kHashSlots];
    uint16_t rank_of[kHashSlots];
    uint8_t chunk_of[kHashSlots];
    uint32_t counts[kMaxChunks * TG];
    uint32_t cum[kMaxChunks * TG];
    uint32_t T[kMaxChunks * (TG + 1)];
    uint32_t ls[kMaxChunks + 1];
    uint32_t cbase[kMaxChunks + 1];
    int64_t rp[TG + 1];
    int64_t dpos[TG];
    int n_lines, n_chunks, flags;
};

// Row (within the group) of entry j: the last r with rp[r] <= j.
template <int TG>
__device__ __forceinline__ int row_of(const TilePlanLds<TG>& S, int nrows, int64_t j)
{
    int lo = 0, hi = nrows - 1;
    while (lo < hi) {
        const int mid = (lo + hi + 1) >> 1;
        if (S.rp[mid] <= j) lo = mid;
        else hi = mid - 1;
    }
    return lo;
}

// Builds the plan in LDS; S.flags != 0 when the group does not fit.
// LS: log2 of the x values per 128-byte line (4 real, 3 complex).  QA: every
// (chunk, row) block is padded to a multiple of 4 entries (the 3-byte layout,
// read as 12-byte quads).
template <int TG, int LS, bool QA>
__device__ __forceinline__ void tile_plan(TilePlanLds<TG>& S, int64_t dim, const int64_t* __restrict__ rowptr,
                          const int32_t* __restrict__ col, const uint16_t* __restrict__ code, int64_t g)
{
    const int t = threadIdx.x;
    const int64_t r0 = g * TG, r1 = min(r0 + TG, dim);
    const int nrows = (int)(r1 - r0);
    for (int i = t; i < kHashSlots; i += kBuildThreads) {
        S.keys[i] = kEmpty;
        S.sorted[i] = kEmpty;
    }
    for (int i = t; i <= nrows; i += kBuildThreads) S.rp[i] = rowptr[r0 + i];
    for (int i = t; i < TG; i += kBuildThreads) S.dpos[i] = -1;
    if (t == 0) S.n_lines = S.n_chunks = S.flags = 0;
    __syncthreads();
    const int64_t e0 = S.rp[0], e1 = S.rp[nrows];
    // kLU entries per thread loaded before any is inserted: one block per CU
    // leaves few waves to hide the load latency
    for (int64_t jb = e0; jb < e1; jb += (int64_t)kBuildThreads * kLU) {
        uint32_t qv[kLU], cv[kLU];
#pragma unroll
        for (int u = 0; u < kLU; ++u) {  // clamped, not guarded: the loads issue together
            const int64_t j = min(jb + u * kBuildThreads + t, e1 - 1);
            qv[u] = code[j];
            cv[u] = (uint32_t)col[j];
        }
#pragma unroll
        for (int u = 0; u < kLU; ++u) {
            const int64_t j = jb + u * kBuildThreads + t;
            if (j >= e1) continue;
            if (qv[u] == 0xFFFFu) {
                S.dpos[row_of(S, nrows, j)] = j;
                continue;
            }
            if (hash_insert(S.keys, cv[u] >> LS) < 0) atomicOr(&S.flags, 1);
        }
    }
    __syncthreads();
    {
        const int lane = t & 63;
        for (int i = t; i < kHashSlots; i += kBuildThreads) {
            const uint32_t k = S.keys[i];
            const uint64_t m = __ballot(k != kEmpty);
            int basei = 0;
            if (lane == 0 && m) basei = atomicAdd(&S.n_lines, __popcll(m));
            basei = __shfl(basei, 0);
            if (k != kEmpty) S.sorted[basei + __popcll(m & ((1ull << lane) - 1ull))] = k;
        }
    }
    __syncthreads();
    const int L = S.n_lines;
    if (L > kMaxLines) {
        if (t == 0) S.flags |= 2;
        __syncthreads();
        return;
    }
    int P = 1;
    while (P < L) P <<= 1;
    if (P <= 8 * kBuildThreads) {
        bitonic_blocked<uint32_t, kBuildThreads>(S.sorted, P);
    } else {
        for (int k = 2; k <= P; k <<= 1)
            for (int jj = k >> 1; jj > 0; jj >>= 1) {
                for (int i = t; i < P; i += kBuildThreads) {
                    const int ixj = i ^ jj;
                    if (ixj > i) {
                        const uint32_t a = S.sorted[i], b = S.sorted[ixj];
                        if ((a > b) == ((i & k) == 0)) {
                            S.sorted[i] = b;
                            S.sorted[ixj] = a;
                        }
                    }
                }
                __syncthreads();
            }
    }
    // ranks, and the chunks: 256 consecutive lines each
    for (int i = t; i < L; i += kBuildThreads) {
        const int h = hash_find(S.keys, S.sorted[i]);
        S.rank_of[h] = (uint16_t)i;
        S.chunk_of[i] = (uint8_t)(i / kTC);
    }
    const int nch = (L + kTC - 1) / kTC;
    for (int c = t; c <= nch; c += kBuildThreads) S.ls[c] = (uint32_t)min(c * kTC, L);
    if (t == 0) S.n_chunks = nch;
    __syncthreads();
    // entries per (chunk, row).  A row's columns are sorted and a chunk is a run
    // of consecutive sorted lines, so the row's entries in chunks below c are
    // those with columns below chunk c's first column: cum is a binary search
    // of the row's columns (less its diagonal), and counts are differences.
    for (int i = t; i < nch * TG; i += kBuildThreads) {
        const int c = i / TG, r = i % TG;
        uint32_t below = 0;
        if (r < nrows && c > 0) {
            const uint32_t bound = S.sorted[S.ls[c]] << LS;
            int64_t lo = S.rp[r], hi = S.rp[r + 1];
            while (lo < hi) {
                const int64_t mid = (lo + hi) >> 1;
                if ((uint32_t)col[mid] < bound) lo = mid + 1;
                else hi = mid;
            }
            below = (uint32_t)(lo - S.rp[r]) - (S.dpos[r] >= 0 && S.dpos[r] < lo ? 1u : 0u);
        }
        S.cum[c * TG + r] = below;
    }
    __syncthreads();
    for (int i = t; i < nch * TG; i += kBuildThreads) {
        const int c = i / TG, r = i % TG;
        uint32_t n = 0;
        if (r < nrows) {
            const uint32_t tot = (uint32_t)(S.rp[r + 1] - S.rp[r]) - (S.dpos[r] >= 0 ? 1u : 0u);
            n = (c + 1 < nch ? S.cum[(c + 1) * TG + r] : tot) - S.cum[c * TG + r];
        }
        S.counts[c * TG + r] = n;
    }
    __syncthreads();
    // T: per chunk, the row starts; chunk blocks start on 4-entry boundaries
    if (t < nch) {
        uint32_t run = 0;
        for (int r = 0; r < TG; ++r) {
            S.T[t * (TG + 1) + r] = run;
            run += QA ? (S.counts[t * TG + r] + 3u) & ~3u : S.counts[t * TG + r];
        }
        S.T[t * (TG + 1) + TG] = run;  // chunk size, made absolute below
    }
    __syncthreads();
    if (t == 0) {
        uint32_t b = 0;
        for (int c = 0; c < nch; ++c) {
            S.cbase[c] = b;
            b += (S.T[c * (TG + 1) + TG] + 3u) & ~3u;
        }
        S.cbase[nch] = b;
    }
    __syncthreads();
    for (int i = t; i < nch * (TG + 1); i += kBuildThreads) S.T[i] += S.cbase[i / (TG + 1)];
    __syncthreads();
}

// Pass 1: per group, the distinct lines, the chunks and the padded entries.
template <int TG, int LS, bool QA>
__global__ void __launch_bounds__(kBuildThreads)
k_tile_count(int64_t dim, const int64_t* __restrict__ rowptr, const int32_t* __restrict__ col,
             const uint16_t* __restrict__ code, int64_t* __restrict__ n_lines, int64_t* __restrict__ n_entries,
             int64_t* __restrict__ n_chunks, int* __restrict__ overflow)
{
    __shared__ TilePlanLds<TG> S;
    const int64_t g = blockIdx.x;
    tile_plan<TG, LS, QA>(S, dim, rowptr, col, code, g);
    if (threadIdx.x == 0) {
        n_lines[g] = S.n_lines;
        n_chunks[g] = S.n_chunks;
        n_entries[g] = S.flags ? 0 : S.cbase[S.n_chunks];
        if (S.flags) atomicOr(overflow, S.flags);
    }
}

// Pass 2: the line list, the chunk table (T per chunk, then the chunk line
// starts), the entries and the diagonal (W doubles per value: 1 real, 2
// complex; a complex line holds 8 values, so its word field uses 3 bits).
template <int TG, int W, bool QA>
__global__ void __launch_bounds__(kBuildThreads)
k_tile_fill(int64_t dim, const int64_t* __restrict__ rowptr, const int32_t* __restrict__ col,
            const uint16_t* __restrict__ code, const double* __restrict__ diag, const int64_t* __restrict__ base,
            int64_t groups, uint32_t* __restrict__ entries, uint32_t* __restrict__ lines,
            uint32_t* __restrict__ meta, double* __restrict__ tdiag, int* __restrict__ error)
{
    constexpr int LS = W == 2 ? 3 : 4;
    __shared__ TilePlanLds<TG> S;
    const int t = threadIdx.x;
    const int64_t g = blockIdx.x;
    tile_plan<TG, LS, QA>(S, dim, rowptr, col, code, g);
    const int64_t r0 = g * TG, r1 = min(r0 + TG, dim);
    const int nrows = (int)(r1 - r0);
    const int64_t eb = base[g], en = base[g + 1] - eb;
    const int64_t lb = base[groups + 1 + g], ln = base[groups + 1 + g + 1] - lb;
    const int64_t mb = base[2 * (groups + 1) + g], mn = base[2 * (groups + 1) + g + 1] - mb;
    const int L = S.n_lines, nch = S.n_chunks;
    if (S.flags || L != ln || (int64_t)nch * (TG + 2) + 1 != mn || (int64_t)S.cbase[nch] != en) {
        if (t == 0) atomicOr(error, 16 | S.flags);
        return;
    }
    for (int i = t; i < L; i += kBuildThreads) lines[lb + i] = S.sorted[i];
    for (int i = t; i < nch * (TG + 1); i += kBuildThreads) meta[mb + i] = S.T[i];
    for (int i = t; i <= nch; i += kBuildThreads) meta[mb + nch * (TG + 1) + i] = S.ls[i];
    // padding entries: kEmpty (never read, or read and skipped as a quad's tail)
    for (int c = 0; c < nch; ++c) {
        const uint32_t a = S.T[c * (TG + 1) + TG], b = S.cbase[c + 1];
        if (t < (int)(b - a)) entries[eb + a + t] = kEmpty;
    }
    if (QA)
        for (int i = t; i < nch * TG; i += kBuildThreads) {
            const int c = i / TG, r = i % TG;
            const uint32_t a = S.T[c * (TG + 1) + r] + S.counts[c * TG + r], b = S.T[c * (TG + 1) + r + 1];
            for (uint32_t p = a; p < b && (int64_t)p < en; ++p) entries[eb + p] = kEmpty;
        }
    // one wavefront per row, as in the plan's counts
    for (int r = t >> 6; r < nrows; r += kBuildThreads / 64) {
        const int64_t rb = S.rp[r + 1];
        for (int64_t jb = S.rp[r]; jb < rb; jb += 64 * kLU) {
            uint32_t qv[kLU], cv[kLU];
#pragma unroll
            for (int u = 0; u < kLU; ++u) {
                const int64_t j = min(jb + 64 * u + (t & 63), rb - 1);
                qv[u] = code[j];
                cv[u] = (uint32_t)col[j];
            }
#pragma unroll
            for (int u = 0; u < kLU; ++u) {
                const int64_t j = jb + 64 * u + (t & 63);
                const uint32_t q = qv[u];
                if (j >= rb || q == 0xFFFFu) continue;
                const uint32_t cj = cv[u];
                const int h = hash_find(S.keys, cj >> LS);
                if (h < 0) continue;  // flagged by the plan
                const int rank = S.rank_of[h];
                const int c = S.chunk_of[rank];
                const int64_t i_off = (j - S.rp[r]) - (S.dpos[r] >= 0 && S.dpos[r] < j ? 1 : 0);
                const int64_t pos = (int64_t)S.T[c * (TG + 1) + r] + (i_off - S.cum[c * TG + r]);
                if (pos < S.T[c * (TG + 1) + r] || pos >= S.T[c * (TG + 1) + r + 1] || pos >= en) {
                    atomicOr(error, 32);
                    continue;
                }
                const uint32_t code20 = ((q >> 15) << 19) | (q & 0x7FFFu);
                entries[eb + pos] = ((uint32_t)(rank - (int)S.ls[c]) << 24) | ((cj & ((1u << LS) - 1u)) << 20) | code20;
            }
        }
    }
    for (int i = t; i < nrows * W; i += kBuildThreads)
        tdiag[r0 * W + i] = S.dpos[i / W] >= 0 ? diag[r0 * W + i] : 0.0;
}

// y = H x on the tiled layout: one block of kSpmvThreads per group of TG rows.  The
// group's line ids and chunk table are staged in LDS once.  While chunk c is
// processed, chunk c+1's x lines (block-wide) and each lane's first kU entries
// of chunk c+1 are already in flight into registers (double buffering through
// the register file); only rows with more than kTPR * kU entries in a chunk
// issue further loads inside it.
//
// XT = float stages x in FP32 (a line's 16 values are 64 bytes) and keeps the
// FP64 dictionary, FP64 products and FP64 accumulation, so y is the same FP64
// sum as with XT = double for x rounded to FP32: the product is exact for
// FP32 inputs.
template <typename XT> struct XPiece;
template <> struct XPiece<double> {
    using P = double2;
    static constexpr int kWords = 2;
    HD_DEV static P zero() { return make_double2(0.0, 0.0); }
    HD_DEV static P load(const double* x, int64_t xi, int64_t dim, int aligned)
    {
        if (aligned && xi + 1 < dim) return *reinterpret_cast<const double2*>(x + xi);
        return make_double2(xi < dim ? x[xi] : 0.0, xi + 1 < dim ? x[xi + 1] : 0.0);
    }
};
template <> struct XPiece<float> {
    using P = float4;
    static constexpr int kWords = 4;
    HD_DEV static P zero() { return make_float4(0.f, 0.f, 0.f, 0.f); }
    HD_DEV static P load(const float* x, int64_t xi, int64_t dim, int aligned)
    {
        if (aligned && xi + 3 < dim) return *reinterpret_cast<const float4*>(x + xi);
        return make_float4(xi < dim ? x[xi] : 0.f, xi + 1 < dim ? x[xi + 1] : 0.f, xi + 2 < dim ? x[xi + 2] : 0.f,
                           xi + 3 < dim ? x[xi + 3] : 0.f);
    }
};

template <int TG, bool PK, typename XT>
__global__ void __launch_bounds__(kSpmvThreads)
k_spmv_tiled(int64_t dim, int64_t groups, const int64_t* __restrict__ base, const uint32_t* __restrict__ entries,
             const uint32_t* __restrict__ lines, const uint32_t* __restrict__ meta, const double* __restrict__ tdiag,
             const double* __restrict__ dict, int ndict, int x_aligned, int64_t win, const XT* __restrict__ x,
             double* __restrict__ y)
{
    using XP = XPiece<XT>;
    using Piece = typename XP::P;
    constexpr int PPL = kLineWords / XP::kWords;       // 16-byte pieces per line
    constexpr int PER = (kTC * PPL) / kSpmvThreads;     // 16-byte x pieces per thread per chunk
    static_assert(PER * kSpmvThreads == kTC * PPL, "pieces must divide evenly");
    constexpr int kMeta = kMaxChunks * (TG + 2) + 1;
    constexpr int kTPR = kSpmvThreads / TG;
    __shared__ XT xs[kTC * kLineWords];
    __shared__ uint32_t ids[kMaxLines];
    __shared__ uint32_t sm[kMeta];
    extern __shared__ double sdict[];
    for (int i = threadIdx.x; i < ndict; i += kSpmvThreads) sdict[i] = dict[i];
    const int64_t g = xcd_window(blockIdx.x, groups, win);
    const int64_t lb = base[groups + 1 + g], le = base[groups + 1 + g + 1];
    const int64_t mb = base[2 * (groups + 1) + g], me = base[2 * (groups + 1) + g + 1];
    const int nmeta = (int)min(me - mb, (int64_t)kMeta);
    const int nch = (nmeta - 1) / (TG + 2);
    // packed: 3 bytes per entry; a group starts at a multiple of 4 entries,
    // i.e. at dword 3 * base[g] / 4, and every (chunk, row) block holds a
    // multiple of 4 entries, read as one 12-byte item per 4 entries
    const uint32_t* ent = entries + (PK ? base[g] / 4 * 3 : base[g]);
    using Item = typename std::conditional<PK, uint3, uint32_t>::type;
    constexpr int kPer = PK ? 4 : 1;        // entries per item
    constexpr int kUI = kU / kPer;          // items each lane has in flight
    auto item_at = [&](uint32_t q) -> Item {
        if constexpr (PK) return *reinterpret_cast<const uint3*>(ent + 3u * q);
        else return ent[q];
    };
    auto no_item = []() -> Item {
        if constexpr (PK) return make_uint3(~0u, ~0u, ~0u);
        else return kEmpty;
    };
    const int r = threadIdx.x / kTPR, sub = threadIdx.x % kTPR;
    const int64_t row = g * TG + r;
    const int nl = (int)min(le - lb, (int64_t)kMaxLines);
    for (int i = threadIdx.x; i < nl; i += kSpmvThreads) ids[i] = lines[lb + i];
    for (int i = threadIdx.x; i < nmeta; i += kSpmvThreads) sm[i] = meta[mb + i];
    __syncthreads();
    const uint32_t* T = sm;
    const uint32_t* LS = sm + nch * (TG + 1);
    // two chunks in flight: while chunk c is consumed, the lines and first
    // entries of chunks c+1 and c+2 are loading into registers (3.36 against
    // 3.44 ms with one chunk ahead, 4-byte entries)
    constexpr int kAhead = 2;
    Piece preA[PER], preB[PER];
    Item eA[kUI], eB[kUI];
    auto fetch = [&](int c, Piece (&pre)[PER], Item (&e)[kUI]) {
        const int l0 = (int)LS[c], n_here = (int)LS[c + 1] - l0;
#pragma unroll
        for (int k = 0; k < PER; ++k) {
            const int i = threadIdx.x + k * kSpmvThreads;
            pre[k] = XP::zero();
            if (i < n_here * PPL) {
                const int64_t xi = (int64_t)ids[l0 + i / PPL] * kLineWords + (i % PPL) * XP::kWords;
                pre[k] = XP::load(x, xi, dim, x_aligned);
            }
        }
        const uint32_t j0 = T[c * (TG + 1) + r] / kPer + sub, je = T[c * (TG + 1) + r + 1] / kPer;
#pragma unroll
        for (int u = 0; u < kUI; ++u) {
            const uint32_t j = j0 + u * kTPR;
            e[u] = j < je ? item_at(j) : no_item();
        }
    };
    auto consume = [&](const Item (&b)[kUI], double& a) {
#pragma unroll
        for (int u = 0; u < kUI; ++u) {
            if constexpr (PK) {
                // four little-endian 3-byte entries in 12 bytes: slot 16-23,
                // word 12-15, sign 11, index 0-10; 0xFFFFFF pads a block
                const uint32_t w[4] = {b[u].x & 0xFFFFFFu, (b[u].x >> 24) | ((b[u].y & 0xFFFFu) << 8),
                                       (b[u].y >> 16) | ((b[u].z & 0xFFu) << 16), b[u].z >> 8};
#pragma unroll
                for (int k = 0; k < 4; ++k) {
                    if (w[k] == 0xFFFFFFu) continue;
                    double v = sdict[w[k] & 0x7FFu];
                    if (w[k] & 0x800u) v = -v;
                    a = fma(v, (double)xs[(w[k] >> 16) * kLineWords + ((w[k] >> 12) & 15u)], a);
                }
            } else {
                if (b[u] == kEmpty) continue;
                double v = sdict[b[u] & 0x7FFFFu];
                if (b[u] & 0x80000u) v = -v;
                a = fma(v, (double)xs[(b[u] >> 24) * kLineWords + ((b[u] >> 20) & 15u)], a);
            }
        }
    };
    double acc = 0.0;
    auto step = [&](int c, Piece (&pre)[PER], Item (&e)[kUI]) {
        __syncthreads();  // chunk c-1 consumed
#pragma unroll
        for (int k = 0; k < PER; ++k) {
            const int i = threadIdx.x + k * kSpmvThreads;
            *reinterpret_cast<Piece*>(&xs[(i / PPL) * kLineWords + (i % PPL) * XP::kWords]) = pre[k];
        }
        Item cur[kUI];
#pragma unroll
        for (int u = 0; u < kUI; ++u) cur[u] = e[u];
        __syncthreads();
        if (c + kAhead < nch) fetch(c + kAhead, pre, e);
        consume(cur, acc);
        // rows with more than kTPR * kU entries in this chunk
        const uint32_t je = T[c * (TG + 1) + r + 1] / kPer;
        for (uint32_t j0 = T[c * (TG + 1) + r] / kPer + sub + kTPR * kUI; j0 < je; j0 += kTPR * kUI) {
#pragma unroll
            for (int u = 0; u < kUI; ++u) {
                const uint32_t j = j0 + u * kTPR;
                cur[u] = j < je ? item_at(j) : no_item();
            }
            consume(cur, acc);
        }
    };
    if (nch > 0) fetch(0, preA, eA);
    if (kAhead == 2) {
        if (nch > 1) fetch(1, preB, eB);
        for (int c = 0; c < nch; c += 2) {
            step(c, preA, eA);
            if (c + 1 < nch) step(c + 1, preB, eB);
        }
    } else {
        for (int c = 0; c < nch; ++c) step(c, preA, eA);
    }
#pragma unroll
    for (int m = kTPR / 2; m > 0; m >>= 1) acc += __shfl_xor(acc, m, 64);
    if (row < dim && sub == 0) y[row] = fma(tdiag[row], (double)x[row], acc);
}

// y = H x for a complex (Hermitian) matrix on the tiled layout.  The same
// schedule as k_spmv_tiled with complex values: a 128-byte line holds 8 of
// them (entry word field bits 20-22), the staged lines are the same 16-byte
// pieces, and the dictionary (complex, up to 32767 values, 512 KiB) is read
// in place from L2 instead of LDS.
template <int TG>
__global__ void __launch_bounds__(kSpmvThreads)
k_spmv_tiled_c(int64_t dim, int64_t groups, const int64_t* __restrict__ base, const uint32_t* __restrict__ entries,
               const uint32_t* __restrict__ lines, const uint32_t* __restrict__ meta, const double2* __restrict__ tdiag,
               const double2* __restrict__ dict, int x_aligned, int64_t win, const double2* __restrict__ x2,
               double2* __restrict__ y)
{
    constexpr int PER = (kTC * 8) / kSpmvThreads;
    constexpr int kMeta = kMaxChunks * (TG + 2) + 1;
    constexpr int kTPR = kSpmvThreads / TG;
    __shared__ double xs[kTC * kLineWords];
    __shared__ uint32_t ids[kMaxLines];
    __shared__ uint32_t sm[kMeta];
    const double* x = reinterpret_cast<const double*>(x2);
    const int64_t dimd = 2 * dim;  // x as doubles
    const int64_t g = xcd_window(blockIdx.x, groups, win);
    const int64_t lb = base[groups + 1 + g], le = base[groups + 1 + g + 1];
    const int64_t mb = base[2 * (groups + 1) + g], me = base[2 * (groups + 1) + g + 1];
    const int nmeta = (int)min(me - mb, (int64_t)kMeta);
    const int nch = (nmeta - 1) / (TG + 2);
    const uint32_t* ent = entries + base[g];
    const int r = threadIdx.x / kTPR, sub = threadIdx.x % kTPR;
    const int64_t row = g * TG + r;
    const int nl = (int)min(le - lb, (int64_t)kMaxLines);
    for (int i = threadIdx.x; i < nl; i += kSpmvThreads) ids[i] = lines[lb + i];
    for (int i = threadIdx.x; i < nmeta; i += kSpmvThreads) sm[i] = meta[mb + i];
    __syncthreads();
    const uint32_t* T = sm;
    const uint32_t* LS = sm + nch * (TG + 1);
    // one chunk ahead: a second one fits 8 waves per SIMD only with 6 entries
    // per lane, and measured 0.953 against 0.962 ms (flux lattice 6x5, N=10)
    double2 preA[PER];
    uint32_t eA[kU];
    auto fetch = [&](int c, double2 (&pre)[PER], uint32_t (&e)[kU]) {
        const int l0 = (int)LS[c], n_here = (int)LS[c + 1] - l0;
#pragma unroll
        for (int k = 0; k < PER; ++k) {
            const int i = threadIdx.x + k * kSpmvThreads;
            pre[k] = make_double2(0.0, 0.0);
            if (i < n_here * 8) {
                const int64_t xi = (int64_t)ids[l0 + (i >> 3)] * kLineWords + (i & 7) * 2;
                if (x_aligned && xi + 1 < dimd) pre[k] = *reinterpret_cast<const double2*>(x + xi);
                else pre[k] = make_double2(xi < dimd ? x[xi] : 0.0, xi + 1 < dimd ? x[xi + 1] : 0.0);
            }
        }
        const uint32_t j0 = T[c * (TG + 1) + r] + sub, je = T[c * (TG + 1) + r + 1];
#pragma unroll
        for (int u = 0; u < kU; ++u) {
            const uint32_t j = j0 + u * kTPR;
            e[u] = j < je ? ent[j] : kEmpty;
        }
    };
    double2 acc = make_double2(0.0, 0.0);
    auto consume = [&](const uint32_t (&b)[kU]) {
#pragma unroll
        for (int u = 0; u < kU; ++u) {
            if (b[u] == kEmpty) continue;
            double2 v = dict[b[u] & 0x7FFFFu];
            if (b[u] & 0x80000u) v = make_double2(-v.x, -v.y);
            const double2 xv = *reinterpret_cast<const double2*>(&xs[(b[u] >> 24) * kLineWords + ((b[u] >> 20) & 7u) * 2]);
            acc.x = fma(v.x, xv.x, acc.x);
            acc.x = fma(-v.y, xv.y, acc.x);
            acc.y = fma(v.x, xv.y, acc.y);
            acc.y = fma(v.y, xv.x, acc.y);
        }
    };
    auto step = [&](int c, double2 (&pre)[PER], uint32_t (&e)[kU]) {
        __syncthreads();  // chunk c-1 consumed
#pragma unroll
        for (int k = 0; k < PER; ++k) {
            const int i = threadIdx.x + k * kSpmvThreads;
            *reinterpret_cast<double2*>(&xs[(i >> 3) * kLineWords + (i & 7) * 2]) = pre[k];
        }
        uint32_t cur[kU];
#pragma unroll
        for (int u = 0; u < kU; ++u) cur[u] = e[u];
        __syncthreads();
        if (c + 1 < nch) fetch(c + 1, pre, e);
        consume(cur);
        const uint32_t je = T[c * (TG + 1) + r + 1];
        for (uint32_t j0 = T[c * (TG + 1) + r] + sub + kTPR * kU; j0 < je; j0 += kTPR * kU) {
#pragma unroll
            for (int u = 0; u < kU; ++u) {
                const uint32_t j = j0 + u * kTPR;
                cur[u] = j < je ? ent[j] : kEmpty;
            }
            consume(cur);
        }
    };
    if (nch > 0) fetch(0, preA, eA);
    for (int c = 0; c < nch; ++c) step(c, preA, eA);
#pragma unroll
    for (int m = kTPR / 2; m > 0; m >>= 1) {
        acc.x += __shfl_xor(acc.x, m, 64);
        acc.y += __shfl_xor(acc.y, m, 64);
    }
    if (row < dim && sub == 0) {
        const double2 d = tdiag[row], xr = x2[row];
        double2 out;
        out.x = fma(d.x, xr.x, acc.x);
        out.x = fma(-d.y, xr.y, out.x);
        out.y = fma(d.x, xr.y, acc.y);
        out.y = fma(d.y, xr.x, out.y);
        y[row] = out;
    }
}

// Packs 4-byte entries (slot 24-31, word 20-23, sign 19, index 0-18) into
// 3-byte entries (slot 16-23, word 12-15, sign 11, index 0-10, little endian):
// four entries per thread, 16 bytes in, 12 bytes out.  The index fits 11 bits
// because the dictionary holds at most 2047 values.
__global__ void __launch_bounds__(256) k_tile_pack(const uint4* __restrict__ in, int64_t n4, uint32_t* __restrict__ out)
{
    const int64_t stride = (int64_t)gridDim.x * blockDim.x;
    for (int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; i < n4; i += stride) {
        const uint4 q = in[i];
        uint32_t p[4] = {q.x, q.y, q.z, q.w};
#pragma unroll
        for (int k = 0; k < 4; ++k)
            p[k] = ((p[k] >> 24) << 16) | (((p[k] >> 20) & 15u) << 12) | (((p[k] >> 19) & 1u) << 11) | (p[k] & 0x7FFu);
        out[3 * i] = p[0] | (p[1] << 24);
        out[3 * i + 1] = (p[1] >> 8) | (p[2] << 16);
        out[3 * i + 2] = (p[2] >> 16) | (p[3] << 8);
    }
}

}  // namespace

// Builds the layout with TG-row groups; HD_ERR_UNSUPPORTED when a group reads
// more than kMaxLines distinct lines.
// Chooses the block order of the tiled product (groups per XCD window) by
// timing each candidate on zero vectors, one warm and three timed launches
// each: the launch order, windows of 16, 64 and 256 groups per XCD, and one
// window over the whole grid.
static int tiled_pick_window(hd_hamiltonian* h)
{
    h->tg_win = 0;
    if (h->dim == 0) return HD_OK;
    double *x = nullptr, *y = nullptr;
    const size_t n = (size_t)h->dim * (h->is_complex ? 2 : 1);  // doubles per vector
    HD_CHECK(dalloc_t(&x, n, "window autotune x"));
    int st = dalloc_t(&y, n, "window autotune y");
    if (st != HD_OK) {
        dfree(x);
        return st;
    }
    HipErrors q;
    HD_ENQ(q, hipMemsetAsync(x, 0, n * sizeof(double), h->stream));
    st = q.status();
    const int64_t cands[] = {0, 16, 64, 256, (h->tg_groups + kXcds - 1) / kXcds};
    float best = 1e30f;
    int64_t best_w = 0;
    EventTimer timer;
    for (const int64_t w : cands) {
        if (st != HD_OK) break;
        h->tg_win = w;
        if ((st = tiled_launch(h, x, y, h->stream)) != HD_OK) break;
        timer.start(h->stream);
        for (int r = 0; r < 3 && st == HD_OK; ++r) st = tiled_launch(h, x, y, h->stream);
        const float ms = timer.stop(h->stream);
        if (st == HD_OK) st = timer.status("tiled window autotune");
        if (st == HD_OK && ms < best) {
            best = ms;
            best_w = w;
        }
    }
    h->tg_win = best_w;
    dfree(x);
    dfree(y);
    return st;
}

template <int TG>
static int tiled_build_tg(hd_hamiltonian* h, int entry_bytes)
{
    const hipStream_t st = h->stream;
    const int64_t G = (h->dim + TG - 1) / TG;
    EventTimer tm;
    tm.start(st);
    int64_t *d_nl = nullptr, *d_ne = nullptr, *d_nc = nullptr;
    int* d_flag = nullptr;
    int rc = dalloc_t(&d_nl, (size_t)G, "tile line counts");
    if (rc == HD_OK) rc = dalloc_t(&d_ne, (size_t)G, "tile entry counts");
    if (rc == HD_OK) rc = dalloc_t(&d_nc, (size_t)G, "tile chunk counts");
    if (rc == HD_OK) rc = dalloc_t(&d_flag, 2, "tile flags");
    HipErrors q;
    std::vector<int64_t> nl(G), ne(G), nc(G);
    int flags[2] = {0, 0};
    if (rc == HD_OK) {
        HD_ENQ(q, hipMemsetAsync(d_flag, 0, 2 * sizeof(int), st));
        if (h->is_complex)
            hipLaunchKernelGGL((k_tile_count<TG, 3, false>), dim3((unsigned)G), dim3(kBuildThreads), 0, st, h->dim,
                               h->d_rowptr, h->d_col, h->d_code, d_nl, d_ne, d_nc, d_flag);
        else if (entry_bytes == 3)
            hipLaunchKernelGGL((k_tile_count<TG, 4, true>), dim3((unsigned)G), dim3(kBuildThreads), 0, st, h->dim,
                               h->d_rowptr, h->d_col, h->d_code, d_nl, d_ne, d_nc, d_flag);
        else
            hipLaunchKernelGGL((k_tile_count<TG, 4, false>), dim3((unsigned)G), dim3(kBuildThreads), 0, st, h->dim,
                               h->d_rowptr, h->d_col, h->d_code, d_nl, d_ne, d_nc, d_flag);
        HD_ENQ(q, hipGetLastError());
        HD_ENQ(q, hipMemcpyAsync(nl.data(), d_nl, G * sizeof(int64_t), hipMemcpyDeviceToHost, st));
        HD_ENQ(q, hipMemcpyAsync(ne.data(), d_ne, G * sizeof(int64_t), hipMemcpyDeviceToHost, st));
        HD_ENQ(q, hipMemcpyAsync(nc.data(), d_nc, G * sizeof(int64_t), hipMemcpyDeviceToHost, st));
        HD_ENQ(q, hipMemcpyAsync(flags, d_flag, sizeof(int), hipMemcpyDeviceToHost, st));
        HD_ENQ(q, hipStreamSynchronize(st));
        rc = q.status();
    }
    if (rc == HD_OK && flags[0])
        rc = fail(HD_ERR_UNSUPPORTED, "a " + std::to_string(TG) + "-row group reads more than " +
                                          std::to_string(kMaxLines) + " distinct x lines (flags " +
                                          std::to_string(flags[0]) + ")");
    std::vector<int64_t> base(3 * (size_t)(G + 1));
    if (rc == HD_OK) {
        int64_t e = 0, l = 0, o = 0;
        for (int64_t g = 0; g < G; ++g) {
            base[g] = e;
            base[G + 1 + g] = l;
            base[2 * (G + 1) + g] = o;
            e += ne[g];  // a multiple of 4: every group starts 16-byte aligned
            l += nl[g];
            o += nc[g] * (TG + 2) + 1;
        }
        base[G] = e;
        base[2 * G + 1] = l;
        base[3 * G + 2] = o;
        h->tg_groups = G;
        h->tg_entries = e;
        h->tg_lines = l;
        h->tg_offs = o;
    }
    if (rc == HD_OK) rc = dalloc_t(&h->d_tg_base, base.size(), "tile bases");
    if (rc == HD_OK) rc = dalloc_t(&h->d_tg_entries, (size_t)std::max<int64_t>(h->tg_entries, 1), "tile entries");
    if (rc == HD_OK) rc = dalloc_t(&h->d_tg_lines, (size_t)std::max<int64_t>(h->tg_lines, 1), "tile lines");
    if (rc == HD_OK) rc = dalloc_t(&h->d_tg_offs, (size_t)std::max<int64_t>(h->tg_offs, 1), "tile chunk tables");
    const int W = h->is_complex ? 2 : 1;  // doubles per value
    if (rc == HD_OK) rc = dalloc_t(&h->d_tg_diag, (size_t)h->dim * W, "tile diagonal");
    if (rc == HD_OK) {
        HD_ENQ(q, hipMemcpyAsync(h->d_tg_base, base.data(), base.size() * sizeof(int64_t), hipMemcpyHostToDevice, st));
        HD_ENQ(q, hipMemsetAsync(d_flag + 1, 0, sizeof(int), st));
#define HD_TF(WV, QAV)                                                                                             \
    hipLaunchKernelGGL((k_tile_fill<TG, WV, QAV>), dim3((unsigned)G), dim3(kBuildThreads), 0, st, h->dim,            \
                       h->d_rowptr, h->d_col, h->d_code, static_cast<const double*>(h->d_diag), h->d_tg_base, G,     \
                       h->d_tg_entries, h->d_tg_lines, h->d_tg_offs, h->d_tg_diag, d_flag + 1)
        if (W == 2) HD_TF(2, false);
        else if (entry_bytes == 3) HD_TF(1, true);
        else HD_TF(1, false);
#undef HD_TF
        HD_ENQ(q, hipGetLastError());
        HD_ENQ(q, hipMemcpyAsync(flags + 1, d_flag + 1, sizeof(int), hipMemcpyDeviceToHost, st));
        HD_ENQ(q, hipStreamSynchronize(st));
        rc = q.status();
        if (rc == HD_OK && flags[1])
            rc = fail(HD_ERR_HIP, "tiled layout build found inconsistent counts (flags " + std::to_string(flags[1]) +
                                      ")");
    }
    if (rc == HD_OK && entry_bytes == 3 && h->tg_entries > 0) {
        // entries are a multiple of 4; two zero pad words let the SpMV read
        // the dword after any entry
        uint32_t* packed = nullptr;
        const int64_t n4 = h->tg_entries / 4;
        rc = dalloc_t(&packed, (size_t)(3 * n4 + 2), "packed tile entries");
        if (rc == HD_OK) {
            HD_ENQ(q, hipMemsetAsync(packed + 3 * n4, 0, 2 * sizeof(uint32_t), st));
            const int grid = (int)std::min<int64_t>((n4 + 255) / 256, 65536);
            hipLaunchKernelGGL(k_tile_pack, dim3((unsigned)grid), dim3(256), 0, st,
                               reinterpret_cast<const uint4*>(h->d_tg_entries), n4, packed);
            HD_ENQ(q, hipGetLastError());
            HD_ENQ(q, hipStreamSynchronize(st));
            rc = q.status();
            dfree(h->d_tg_entries);
            h->d_tg_entries = packed;
        }
    }
    h->t_tile = tm.stop(st);
    if (rc == HD_OK) rc = tm.status("tiled layout timing");
    dfree(d_nl);
    dfree(d_ne);
    dfree(d_nc);
    dfree(d_flag);
    if (rc != HD_OK) {
        tiled_free(h);
        return rc;
    }
    h->tg_rows = TG;
    h->tg_entry_bytes = entry_bytes;
    h->tg_bytes = (size_t)h->tg_entries * entry_bytes + (size_t)h->tg_lines * 4 + (size_t)h->tg_offs * 4 +
                  base.size() * 8 + (size_t)h->dim * 8 * W;
    h->bytes += h->tg_bytes;
    return tiled_pick_window(h);
}

int tiled_build(hd_hamiltonian* h, int rows_per_group, int entry_bytes)
{
    if (rows_per_group != 0 && rows_per_group != 64 && rows_per_group != 128)
        return fail(HD_ERR_ARG, "rows per group must be 0 (automatic), 64 or 128");
    if (entry_bytes != 0 && entry_bytes != 3 && entry_bytes != 4)
        return fail(HD_ERR_ARG, "entry bytes must be 0 (automatic), 3 or 4");
    if (entry_bytes == 3 && h->is_complex)
        return fail(HD_ERR_UNSUPPORTED, "3-byte entries are implemented for real matrices");
    if (entry_bytes == 3 && h->ndict > 2047)
        return fail(HD_ERR_UNSUPPORTED, "3-byte entries need a dictionary of at most 2047 values");
    // 3 where it applies: every (chunk, row) block is padded to 4 entries and
    // read as one 12-byte item per 4 entries, a quarter fewer bytes than
    // 4-byte entries (N=12, 2S=33: 3.09 against 3.42 ms per product; N=11,
    // 2S=30: 0.414 against 0.464 ms)
    if (entry_bytes == 0) entry_bytes = (!h->is_complex && h->ndict <= 2047) ? 3 : 4;
    if (!h->d_code) return fail(HD_ERR_ARG, "the tiled layout is built from the coded values (value_mask & 4)");
    if (h->storage != HD_STORAGE_FULL) return fail(HD_ERR_ARG, "the tiled layout needs full storage");
    if (h->ndict > (1 << 19)) return fail(HD_ERR_UNSUPPORTED, "dictionary too large for the tiled layout");
    if (!h->is_complex && (size_t)h->ndict * 8 + sizeof(double) * kTC * kLineWords +
            sizeof(uint32_t) * (kMaxLines + kMaxChunks * (128 + 2) + 1) >
        80 * 1024)
        return fail(HD_ERR_UNSUPPORTED, "dictionary too large for LDS next to the line buffer");
    tiled_free(h);
    if (h->dim == 0) return HD_OK;
    if (rows_per_group == 64) return tiled_build_tg<64>(h, entry_bytes);
    // 128-row groups stage fewer lines per entry; denser matrices fall back to 64.
    // (256-row groups stage 24% fewer lines per entry again, but measured no
    // faster: N=11, 2S=30, 0.410 against 0.414 ms with 3-byte entries, 0.507
    // against 0.464 ms with 4-byte entries)
    const int rc = tiled_build_tg<128>(h, entry_bytes);
    if (rc != HD_ERR_UNSUPPORTED || rows_per_group == 128) return rc;
    return tiled_build_tg<64>(h, entry_bytes);
}

void tiled_free(hd_hamiltonian* h)
{
    dfree(h->d_tg_base);
    dfree(h->d_tg_entries);
    dfree(h->d_tg_lines);
    dfree(h->d_tg_offs);
    dfree(h->d_tg_diag);
    h->d_tg_base = nullptr;
    h->d_tg_entries = nullptr;
    h->d_tg_lines = nullptr;
    h->d_tg_offs = nullptr;
    h->d_tg_diag = nullptr;
    h->bytes -= std::min(h->bytes, h->tg_bytes);  // the layout's share of the handle's HBM
    h->tg_bytes = 0;
    h->tg_groups = h->tg_entries = h->tg_lines = h->tg_offs = 0;
    h->tg_rows = 0;
    h->tg_entry_bytes = 4;
    h->tg_win = 0;
}

int tiled_launch(const hd_hamiltonian* h, const void* x, void* y, hipStream_t s, bool x32)
{
    if (!h->d_tg_base) return fail(HD_ERR_ARG, "no tiled layout (hd_ham_tile)");
    if (x32 && h->is_complex) return fail(HD_ERR_UNSUPPORTED, "the FP32-x tiled product is implemented for real matrices");
    if (h->dim == 0) return HD_OK;
    const int aligned = ((uintptr_t)x & 15) == 0 ? 1 : 0;
    if (h->is_complex) {
#define HD_TC(TG)                                                                                                 \
    hipLaunchKernelGGL((k_spmv_tiled_c<TG>), dim3((unsigned)h->tg_groups), dim3(kSpmvThreads), 0, s, h->dim,      \
                       h->tg_groups, h->d_tg_base, h->d_tg_entries, h->d_tg_lines, h->d_tg_offs,                  \
                       reinterpret_cast<const double2*>(h->d_tg_diag), static_cast<const double2*>(h->d_dict),     \
                       aligned, h->tg_win, static_cast<const double2*>(x), static_cast<double2*>(y))
        if (h->tg_rows == 128) HD_TC(128);
        else HD_TC(64);
#undef HD_TC
        HD_HIP(hipGetLastError());
        return HD_OK;
    }
    const size_t lds = sizeof(double) * (size_t)std::max(h->ndict, 1);
#define HD_TL(TG, PK, XT)                                                                                            \
    hipLaunchKernelGGL((k_spmv_tiled<TG, PK, XT>), dim3((unsigned)h->tg_groups), dim3(kSpmvThreads), lds, s, h->dim, \
                       h->tg_groups, h->d_tg_base, h->d_tg_entries, h->d_tg_lines, h->d_tg_offs, h->d_tg_diag,        \
                       static_cast<const double*>(h->d_dict), h->ndict, aligned, h->tg_win,                         \
                       static_cast<const XT*>(x), static_cast<double*>(y))
#define HD_TL_X(TG, PK)         \
    if (x32) HD_TL(TG, PK, float); \
    else HD_TL(TG, PK, double)
    const bool pk = h->tg_entry_bytes == 3;
    if (h->tg_rows == 128) {
        if (pk) HD_TL_X(128, true);
        else HD_TL_X(128, false);
    } else {
        if (pk) HD_TL_X(64, true);
        else HD_TL_X(64, false);
    }
#undef HD_TL_X
#undef HD_TL
    HD_HIP(hipGetLastError());
    return HD_OK;
}

// HBM bytes one tiled SpMV must move: entries, offsets, line lists, group
// bases, the dictionary, the diagonal, x once and y.
size_t tiled_bytes(const hd_hamiltonian* h)
{
    const size_t w = h->is_complex ? 16 : 8;  // bytes per value
    return (size_t)h->tg_entries * h->tg_entry_bytes + (size_t)h->tg_offs * 4 + (size_t)h->tg_lines * 4 +
           (size_t)(3 * (h->tg_groups + 1)) * 8 + (size_t)h->ndict * w + (size_t)h->dim * w * 3;
}

}  // namespace hd

using namespace hd;

extern "C" int hd_ham_tile(hd_hamiltonian* h, int32_t rows_per_group, int32_t entry_bytes)
{
    if (!h) return fail(HD_ERR_ARG, "hd_ham_tile: null handle");
    return tiled_build(h, rows_per_group, entry_bytes);
}

extern "C" int hd_ham_tile_info(const hd_hamiltonian* h, hd_tile_info_t* info)
{
    if (!h || !info) return fail(HD_ERR_ARG, "hd_ham_tile_info: null argument");
    if (!h->d_tg_base && h->dim > 0) return fail(HD_ERR_ARG, "no tiled layout (hd_ham_tile)");
    info->groups = h->tg_groups;
    info->entries = h->tg_entries;
    info->lines = h->tg_lines;
    info->chunks = h->tg_rows ? (h->tg_offs - h->tg_groups) / (h->tg_rows + 2) : 0;
    info->rows_per_group = h->tg_rows;
    info->spmv_bytes = (int64_t)tiled_bytes(h);
    info->build_ms = h->t_tile;
    info->entry_bytes = h->tg_entry_bytes;
    info->block_window = h->tg_win;
    return HD_OK;
}

extern "C" int hd_ham_tile_order(hd_hamiltonian* h, int64_t groups_per_xcd)
{
    if (!h) return fail(HD_ERR_ARG, "hd_ham_tile_order: null handle");
    if (!h->d_tg_base && h->dim > 0) return fail(HD_ERR_ARG, "no tiled layout (hd_ham_tile)");
    if (groups_per_xcd < -1) return fail(HD_ERR_ARG, "groups per XCD must be -1 (autotune), 0 or positive");
    if (groups_per_xcd == -1) return tiled_pick_window(h);
    h->tg_win = groups_per_xcd;
    return HD_OK;
}

extern "C" int hd_ham_device_tiled(const hd_hamiltonian* h, const int64_t** d_base, const uint32_t** d_entries,
                                   const uint32_t** d_lines, const uint32_t** d_offsets, const double** d_diag)
{
    if (!h) return fail(HD_ERR_ARG, "hd_ham_device_tiled: null handle");
    if (d_base) *d_base = h->d_tg_base;
    if (d_entries) *d_entries = h->d_tg_entries;
    if (d_lines) *d_lines = h->d_tg_lines;
    if (d_offsets) *d_offsets = h->d_tg_offs;
    if (d_diag) *d_diag = h->d_tg_diag;
    return HD_OK;
}
