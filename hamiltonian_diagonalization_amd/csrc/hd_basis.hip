// hd_basis.hip -- runtime helpers, the sector counting table, and the Fock
// basis kernels (unranking for GenerateFockSpace, ranking for
// FindFockStateIndex).
//
// Reference: src/basis/fermion_fock_basis.cpp:134-161 (FindFockStateIndex by
// binary search) and :241-354 (GenerateFockSpace: enumerate all C(Norb,N)
// bitstrings in ascending order and filter with the sector predicate).  Here
// the sector is a weight constraint, so its states are counted by a small
// dynamic-programming table and the i-th state is produced directly
// (combinatorial unranking), one thread per state, in the same ascending order.
#include <hip/hip_runtime.h>

#include <algorithm>
#include <cstring>
#include <mutex>

#include "hd_internal.hpp"

namespace hd {

static thread_local std::string g_err;

void set_error(const std::string& msg) { g_err = msg; }

int fail(int code, const std::string& msg)
{
    g_err = msg;
    return code;
}

int dalloc(void** p, size_t bytes, const char* what)
{
    *p = nullptr;
    if (bytes == 0) bytes = 16;
    hipError_t e = hipMalloc(p, bytes);
    if (e != hipSuccess) {
        (void)hipGetLastError();
        *p = nullptr;
        return fail(HD_ERR_NOMEM, std::string("hipMalloc of ") + std::to_string(bytes) +
                                      " bytes for " + what + " failed: " + hipGetErrorString(e));
    }
    return HD_OK;
}

void dfree(void* p)
{
    if (p) (void)hipFree(p);
}

int device_props(int* n_cu, size_t* max_lds)
{
    int dev = 0;
    HD_HIP(hipGetDevice(&dev));
    hipDeviceProp_t prop;
    HD_HIP(hipGetDeviceProperties(&prop, dev));
    if (n_cu) *n_cu = prop.multiProcessorCount;
    if (max_lds) *max_lds = prop.sharedMemPerBlock;
    return HD_OK;
}

// ------------------------------------------------------------ rank table

namespace {
struct GroupSpec {
    bool lz = false;
    int G = 1;
    int X = 1, Y = 1;  // finite: element (gx, gy) -> gx * Y + gy
    std::vector<int> w;  // weight of each orbital (encoded)
    int target = 0;
    bool empty = false;
};

int mod(int a, int n) { return ((a % n) + n) % n; }

int gsub(const GroupSpec& gs, int g, int w)
{
    if (gs.lz) return g - w;
    const int gx = g / gs.Y, gy = g % gs.Y, wx = w / gs.Y, wy = w % gs.Y;
    return mod(gx - wx, gs.X) * gs.Y + mod(gy - wy, gs.Y);
}

int gadd(const GroupSpec& gs, int g, int w)
{
    if (gs.lz) return g + w;
    const int gx = g / gs.Y, gy = g % gs.Y, wx = w / gs.Y, wy = w % gs.Y;
    return mod(gx + wx, gs.X) * gs.Y + mod(gy + wy, gs.Y);
}

int group_spec(const hd_sector& sec, int N, int Norb, GroupSpec& gs)
{
    gs.w.assign(Norb, 0);
    switch (sec.kind) {
    case HD_SECTOR_NONE:
        gs.G = 1;
        gs.target = 0;
        return HD_OK;
    case HD_SECTOR_LZ: {
        // TestAngularMomentumSector: sum_k (2k - maxLz) == lzTot  <=>  sum k = T.
        const long long tot = (long long)sec.a + (long long)N * sec.b;
        gs.lz = true;
        for (int k = 0; k < Norb; ++k) gs.w[k] = k;
        if (tot % 2 != 0 || tot < 0) {
            gs.empty = true;
            return HD_OK;
        }
        gs.target = (int)(tot / 2);
        const long long maxsum = (long long)N * (2LL * Norb - N - 1) / 2;
        if (gs.target > maxsum) gs.empty = true;
        return HD_OK;
    }
    case HD_SECTOR_LZ2: {
        // TestAngularMomentumSector2LL (two_level_angular_momentum_constraint.hpp:41-73)
        // on the combined state (first << N2) | second: second-level orbital j
        // has m = 2j - max, first-level orbital i has m = 2i - (max - 2), so
        // with weights w(j) = j and w(N2 + i) = i + 1 the condition is
        // sum m = 2 sum w - N max = lzTot.
        const int n1 = sec.c, n2 = Norb - sec.c;
        if (n1 <= 0 || n2 != n1 + 2)
            return fail(HD_ERR_ARG, "two-level sector: the orbitals must be n1 first-level plus n1 + 2 second-level");
        const long long tot = (long long)sec.a + (long long)N * sec.b;
        gs.lz = true;
        for (int j = 0; j < n2; ++j) gs.w[j] = j;
        for (int i = 0; i < n1; ++i) gs.w[n2 + i] = i + 1;
        if (tot % 2 != 0 || tot < 0) {
            gs.empty = true;
            return HD_OK;
        }
        gs.target = (int)(tot / 2);
        std::vector<int> ws(gs.w);
        std::sort(ws.begin(), ws.end());
        long long maxsum = 0;
        for (int k = 0; k < N; ++k) maxsum += ws[Norb - 1 - k];
        if (gs.target > maxsum) gs.empty = true;
        return HD_OK;
    }
    case HD_SECTOR_K2D:
    case HD_SECTOR_K1D: {
        int X, Y, kx, ky;
        if (sec.kind == HD_SECTOR_K2D) {
            kx = sec.a; ky = sec.b; X = sec.c; Y = sec.d;
        } else {
            kx = 0; ky = sec.a; X = 1; Y = sec.b;
        }
        if (X <= 0 || Y <= 0 || (long long)X * Y > 4096)
            return fail(HD_ERR_ARG, "momentum sector: grid dimensions must be positive and X*Y <= 4096");
        gs.X = X; gs.Y = Y; gs.G = X * Y;
        // TestLinearMomentumSector2/1: x = sum floor(k/dimY) mod dimX, y = sum k mod dimY.
        for (int k = 0; k < Norb; ++k) gs.w[k] = mod(k / Y, X) * Y + mod(k, Y);
        if (kx < 0 || kx >= X || ky < 0 || ky >= Y) {
            gs.empty = true;  // kState_t comparison can never match
            return HD_OK;
        }
        gs.target = kx * Y + ky;
        return HD_OK;
    }
    }
    return fail(HD_ERR_ARG, "unknown sector kind " + std::to_string(sec.kind));
}
}  // namespace

int make_rank_table(const hd_sector& sec, int N, int Norb, RankTable& rt)
{
    if (N <= 0 || Norb <= 0) return fail(HD_ERR_ARG, "MUST SPECIFY nbr particles > 0 and nbr orbitals > 0");
    if (N >= 64) return fail(HD_ERR_ARG, "CANNOT ADDRESS 64 PARTICLES OR MORE");
    if (N > Norb) return fail(HD_ERR_ARG, "NBR PARTICLES > NBR ORBITALS - NOT ALLOWED");
    if (Norb > 64) return fail(HD_ERR_ARG, "at most 64 orbitals fit a fock_t");
    GroupSpec gs;
    HD_CHECK(group_spec(sec, N, Norb, gs));
    rt = RankTable();
    rt.N = N;
    rt.Norb = Norb;
    rt.is_lz = gs.lz ? 1 : 0;
    rt.target = gs.target;
    const int NG = gs.lz ? gs.target + 1 : gs.G;
    rt.G = gs.lz ? 0 : gs.G;
    if (gs.empty) {
        rt.empty = true;
        rt.dim = 0;
        return HD_OK;
    }
    const int npc = (Norb + 1) * (N + 1);
    // dense DP over (p, c, g) in 64 bits
    std::vector<unsigned long long> W((size_t)(Norb + 1) * (N + 1) * NG, 0ull);
    auto at = [&](int p, int c, int g) -> unsigned long long& {
        return W[((size_t)p * (N + 1) + c) * NG + g];
    };
    at(0, 0, 0) = 1;
    for (int p = 0; p < Norb; ++p) {
        for (int c = 0; c <= N; ++c)
            for (int g = 0; g < NG; ++g) {
                unsigned long long v = at(p, c, g);
                if (c > 0) {
                    if (gs.lz) {
                        if (g - gs.w[p] >= 0) v += at(p, c - 1, g - gs.w[p]);
                    } else {
                        v += at(p, c - 1, gsub(gs, g, gs.w[p]));
                    }
                }
                at(p + 1, c, g) = v;
            }
    }
    rt.dim = at(Norb, N, gs.target);
    if (rt.dim == 0) {
        rt.empty = true;
        return HD_OK;
    }
    if (rt.dim >= (1ull << 31))
        return fail(HD_ERR_UNSUPPORTED, "sector dimension >= 2^31 exceeds int32 CRS column indices");
    // compact layout
    std::vector<int32_t> off(npc), gmm(2 * npc);
    std::vector<uint32_t> tab;
    for (int p = 0; p <= Norb; ++p)
        for (int c = 0; c <= N; ++c) {
            const int i = p * (N + 1) + c;
            int lo, hi;
            if (gs.lz) {
                // sums of the c smallest / largest weights among orbitals 0..p-1
                std::vector<int> ws(gs.w.begin(), gs.w.begin() + p);
                std::sort(ws.begin(), ws.end());
                lo = 0;
                hi = 0;
                for (int k = 0; k < c && k < p; ++k) {
                    lo += ws[k];
                    hi += ws[p - 1 - k];
                }
                hi = std::min(gs.target, hi);
                if (c > p) hi = lo - 1;
            } else {
                lo = 0;
                hi = (c <= p) ? gs.G - 1 : -1;
            }
            gmm[2 * i] = lo;
            gmm[2 * i + 1] = hi;
            off[i] = (int32_t)tab.size() - lo;
            for (int g = lo; g <= hi; ++g) {
                const unsigned long long v = at(p, c, g);
                if (v >= (1ull << 32))
                    return fail(HD_ERR_UNSUPPORTED, "sector counting table entry exceeds 32 bits");
                tab.push_back((uint32_t)v);
            }
            if (hi < lo) off[i] = 0;
        }
    std::vector<int32_t>& b = rt.blob;
    b.clear();
    b.push_back(N);
    b.push_back(Norb);
    b.push_back(rt.G);
    b.push_back(gs.target);
    b.push_back(rt.is_lz);
    b.push_back((int32_t)tab.size());
    b.insert(b.end(), off.begin(), off.end());
    b.insert(b.end(), gmm.begin(), gmm.end());
    if (gs.lz) {
        for (int k = 0; k < Norb; ++k) b.push_back(gs.w[k]);  // orbital weights
    } else {
        for (int g = 0; g < gs.G; ++g)
            for (int k = 0; k < Norb; ++k) b.push_back(gsub(gs, g, gs.w[k]));
    }
    const size_t base = b.size();
    b.resize(base + tab.size());
    std::memcpy(b.data() + base, tab.data(), tab.size() * sizeof(uint32_t));
    rt.empty = false;
    return HD_OK;
}

// Is weight(a)+weight(b) == weight(c)+weight(d) in the sector's group?
int sector_weight_delta_ok(const RankTable& rt, const hd_sector& sec, int a, int b, int c,
                           int d)
{
    if (sec.kind == HD_SECTOR_NONE) return 1;
    GroupSpec gs;
    if (group_spec(sec, rt.N, rt.Norb, gs) != HD_OK) return 0;
    const int lhs = gadd(gs, gadd(gs, 0, gs.w[a]), gs.w[b]);
    const int rhs = gadd(gs, gadd(gs, 0, gs.w[c]), gs.w[d]);
    return lhs == rhs;
}

int sector_pair_sums(const RankTable& rt, const hd_sector& sec, std::vector<int>& sums)
{
    const int K = rt.Norb;
    sums.assign((size_t)K * K, 0);
    if (sec.kind == HD_SECTOR_NONE) return HD_OK;
    GroupSpec gs;
    const int st = group_spec(sec, rt.N, rt.Norb, gs);
    if (st != HD_OK) return st;
    for (int a = 0; a < K; ++a)
        for (int b = 0; b < K; ++b) sums[(size_t)a * K + b] = gadd(gs, gadd(gs, 0, gs.w[a]), gs.w[b]);
    return HD_OK;
}

// ----------------------------------------------------------------- kernels

template <bool LZ>
__global__ void __launch_bounds__(256) k_unrank(const int32_t* __restrict__ blob, int blob_words,
                                                int64_t dim, uint64_t* __restrict__ out)
{
    extern __shared__ int32_t lds[];
    block_copy_words(lds, blob, blob_words);
    __syncthreads();
    const RankView v = rank_view(lds);
    const int64_t stride = (int64_t)gridDim.x * blockDim.x;
    for (int64_t r = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; r < dim; r += stride)
        out[r] = unrank_state<LZ>(v, (uint64_t)r);
}

template <bool LZ>
__global__ void __launch_bounds__(256) k_rank(const int32_t* __restrict__ blob, int blob_words,
                                              const uint64_t* __restrict__ states, int64_t n,
                                              uint64_t offset, uint64_t* __restrict__ idx)
{
    extern __shared__ int32_t lds[];
    block_copy_words(lds, blob, blob_words);
    __syncthreads();
    const RankView v = rank_view(lds);
    const int64_t stride = (int64_t)gridDim.x * blockDim.x;
    for (int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; i < n; i += stride) {
        uint64_t r;
        idx[i] = rank_state<LZ>(v, states[i], r) ? offset + r : ~0ull;
    }
}

static int lds_launch_ok(const void* fn, size_t lds)
{
    size_t max_lds = 0;
    HD_CHECK(device_props(nullptr, &max_lds));
    if (lds > 160 * 1024)
        return fail(HD_ERR_UNSUPPORTED, "sector counting table of " + std::to_string(lds) +
                                            " bytes does not fit in LDS");
    if (lds > 64 * 1024) {
        hipError_t e = hipFuncSetAttribute(fn, hipFuncAttributeMaxDynamicSharedMemorySize, (int)lds);
        if (e != hipSuccess) {
            (void)hipGetLastError();
            return fail(HD_ERR_UNSUPPORTED, std::string("LDS request of ") + std::to_string(lds) +
                                                " bytes refused: " + hipGetErrorString(e));
        }
    }
    return HD_OK;
}

static int grid_for(int64_t n, int block, int per_cu)
{
    int cu = 256;
    device_props(&cu, nullptr);
    const int64_t want = (n + block - 1) / block;
    return (int)std::max<int64_t>(1, std::min<int64_t>(want, (int64_t)cu * per_cu));
}

}  // namespace hd

using namespace hd;

// ------------------------------------------------------------------- C ABI

extern "C" {

const char* hd_last_error(void) { return g_err.c_str(); }

int hd_version(void) { return 10000; }

int hd_malloc(void** d_ptr, size_t bytes)
{
    if (!d_ptr) return fail(HD_ERR_ARG, "hd_malloc: null output");
    return dalloc(d_ptr, bytes, "hd_malloc");
}

int hd_free(void* d_ptr)
{
    HD_HIP(hipFree(d_ptr));
    return HD_OK;
}

int hd_memcpy_h2d(void* d, const void* h, size_t bytes)
{
    HD_HIP(hipMemcpy(d, h, bytes, hipMemcpyHostToDevice));
    return HD_OK;
}

int hd_memcpy_d2h(void* h, const void* d, size_t bytes)
{
    HD_HIP(hipMemcpy(h, d, bytes, hipMemcpyDeviceToHost));
    return HD_OK;
}

int hd_memset(void* d, int value, size_t bytes)
{
    HD_HIP(hipMemset(d, value, bytes));
    return HD_OK;
}

int hd_synchronize(void)
{
    HD_HIP(hipDeviceSynchronize());
    return HD_OK;
}

int hd_device_count(int* n)
{
    if (!n) return fail(HD_ERR_ARG, "null output");
    HD_HIP(hipGetDeviceCount(n));
    return HD_OK;
}

int hd_basis_create(const hd_sector* sector, uint32_t N, uint32_t Norb, hd_basis** out)
{
    if (!sector || !out) return fail(HD_ERR_ARG, "hd_basis_create: null argument");
    *out = nullptr;
    hd_basis* b = new hd_basis();
    b->sector = *sector;
    b->N = (int)N;
    b->Norb = (int)Norb;
    int st = make_rank_table(*sector, (int)N, (int)Norb, b->rt);
    if (st != HD_OK) {
        delete b;
        return st;
    }
    if (b->rt.empty) {
        delete b;
        return fail(HD_ERR_EMPTY, "FOCK BASIS IS ZERO DIMENSIONAL");
    }
    b->dim = (int64_t)b->rt.dim;
    if (hipGetDevice(&b->device) != hipSuccess) {
        delete b;
        return fail(HD_ERR_HIP, "hipGetDevice failed");
    }
    st = HD_OK;
    auto cleanup = [&]() {
        dfree(b->d_rank);
        dfree(b->d_states);
        if (b->stream) (void)hipStreamDestroy(b->stream);
        delete b;
    };
    if (hipStreamCreateWithFlags(&b->stream, hipStreamNonBlocking) != hipSuccess) {
        cleanup();
        return fail(HD_ERR_HIP, "stream creation failed");
    }
    const size_t lds = b->rt.bytes();
    if ((st = dalloc(reinterpret_cast<void**>(&b->d_rank), lds, "rank table")) != HD_OK ||
        (st = dalloc_t(&b->d_states, (size_t)b->dim, "basis states")) != HD_OK) {
        cleanup();
        return st;
    }
    const void* fn = b->rt.is_lz ? (const void*)k_unrank<true> : (const void*)k_unrank<false>;
    if ((st = lds_launch_ok(fn, lds)) != HD_OK) {
        cleanup();
        return st;
    }
    if (hipMemcpyAsync(b->d_rank, b->rt.blob.data(), lds, hipMemcpyHostToDevice, b->stream) !=
        hipSuccess) {
        cleanup();
        return fail(HD_ERR_HIP, "rank table upload failed");
    }
    const int grid = grid_for(b->dim, 256, lds > 80 * 1024 ? 1 : 4);
    if (b->rt.is_lz)
        hipLaunchKernelGGL(k_unrank<true>, dim3(grid), dim3(256), lds, b->stream, b->d_rank,
                           (int)b->rt.blob.size(), b->dim, b->d_states);
    else
        hipLaunchKernelGGL(k_unrank<false>, dim3(grid), dim3(256), lds, b->stream, b->d_rank,
                           (int)b->rt.blob.size(), b->dim, b->d_states);
    hipError_t e = hipGetLastError();
    if (e == hipSuccess) e = hipStreamSynchronize(b->stream);
    if (e != hipSuccess) {
        cleanup();
        return fail(HD_ERR_HIP, std::string("basis unrank kernel: ") + hipGetErrorString(e));
    }
    *out = b;
    return HD_OK;
}

int hd_basis_create_from_states(uint32_t N, uint32_t Norb, const uint64_t* h_states, int64_t dim, hd_basis** out)
{
    if (!out || (dim > 0 && !h_states)) return fail(HD_ERR_ARG, "hd_basis_create_from_states: null argument");
    *out = nullptr;
    if (N == 0 || Norb == 0) return fail(HD_ERR_ARG, "MUST SPECIFY nbr particles > 0 and nbr orbitals > 0");
    if (N >= 64) return fail(HD_ERR_ARG, "CANNOT ADDRESS 64 PARTICLES OR MORE");
    if (N > Norb) return fail(HD_ERR_ARG, "NBR PARTICLES > NBR ORBITALS - NOT ALLOWED");
    if (Norb > 64) return fail(HD_ERR_ARG, "at most 64 orbitals fit a fock_t");
    if (dim < 0) return fail(HD_ERR_ARG, "negative dimension");
    if (dim == 0) return fail(HD_ERR_EMPTY, "FOCK BASIS IS ZERO DIMENSIONAL");
    if (dim > (int64_t)0x7FFFFFFE) return fail(HD_ERR_UNSUPPORTED, "more states than int32 columns address");
    for (int64_t i = 0; i < dim; ++i) {
        const uint64_t s = h_states[i];
        if (__builtin_popcountll(s) != (int)N || (Norb < 64 && (s >> Norb) != 0))
            return fail(HD_ERR_ARG, "state " + std::to_string(i) + " does not have " + std::to_string(N) +
                                        " particles in " + std::to_string(Norb) + " orbitals");
        if (i > 0 && s <= h_states[i - 1]) return fail(HD_ERR_ARG, "states must be strictly ascending");
    }
    hd_basis* b = new hd_basis();
    b->sector = hd_sector{HD_SECTOR_LIST, 0, 0, 0, 0};
    b->N = (int)N;
    b->Norb = (int)Norb;
    b->dim = dim;
    auto cleanup = [&]() {
        dfree(b->d_rank);
        dfree(b->d_states);
        if (b->stream) (void)hipStreamDestroy(b->stream);
        delete b;
    };
    if (hipGetDevice(&b->device) != hipSuccess) {
        delete b;
        return fail(HD_ERR_HIP, "hipGetDevice failed");
    }
    if (hipStreamCreateWithFlags(&b->stream, hipStreamNonBlocking) != hipSuccess) {
        cleanup();
        return fail(HD_ERR_HIP, "stream creation failed");
    }
    int st = dalloc_t(&b->d_states, (size_t)dim, "basis states");
    if (st != HD_OK) {
        cleanup();
        return st;
    }
    // header (kRankList), then the states' address and count; the offset,
    // group and count sections of the other kinds are empty
    RankTable& rt = b->rt;
    rt.N = (int)N;
    rt.Norb = (int)Norb;
    rt.G = 0;
    rt.is_lz = kRankList;
    rt.dim = (uint64_t)dim;
    rt.empty = false;
    const uint64_t addr = reinterpret_cast<uint64_t>(b->d_states);
    rt.blob = {(int32_t)N, (int32_t)Norb, 0, 0, kRankList, 0, (int32_t)(uint32_t)addr, (int32_t)(uint32_t)(addr >> 32),
               (int32_t)(uint32_t)(uint64_t)dim, (int32_t)(uint32_t)((uint64_t)dim >> 32)};
    // rank_view lays the sections out after the header from N and Norb: pad
    // the blob so that every section pointer it forms stays inside it
    const size_t npc = ((size_t)Norb + 1) * ((size_t)N + 1);
    rt.blob.resize(std::max(rt.blob.size(), kRankHeader + 3 * npc + (size_t)Norb + 1), 0);
    if ((st = dalloc(reinterpret_cast<void**>(&b->d_rank), rt.bytes(), "rank table")) != HD_OK) {
        cleanup();
        return st;
    }
    if (hipMemcpyAsync(b->d_states, h_states, (size_t)dim * sizeof(uint64_t), hipMemcpyHostToDevice, b->stream) !=
            hipSuccess ||
        hipMemcpyAsync(b->d_rank, rt.blob.data(), rt.bytes(), hipMemcpyHostToDevice, b->stream) != hipSuccess ||
        hipStreamSynchronize(b->stream) != hipSuccess) {
        cleanup();
        return fail(HD_ERR_HIP, "state list upload failed");
    }
    *out = b;
    return HD_OK;
}

int hd_basis_destroy(hd_basis* b)
{
    if (!b) return HD_OK;
    dfree(b->d_rank);
    dfree(b->d_states);
    if (b->stream) (void)hipStreamDestroy(b->stream);
    delete b;
    return HD_OK;
}

int hd_basis_dim(const hd_basis* b, int64_t* dim)
{
    if (!b || !dim) return fail(HD_ERR_ARG, "hd_basis_dim: null argument");
    *dim = b->dim;
    return HD_OK;
}

int hd_basis_copy_states(const hd_basis* b, int64_t first, int64_t count, uint64_t* h_out)
{
    if (!b || !h_out || first < 0 || count < 0 || first + count > b->dim)
        return fail(HD_ERR_ARG, "hd_basis_copy_states: range outside the basis");
    HD_HIP(hipMemcpy(h_out, b->d_states + first, (size_t)count * sizeof(uint64_t),
                     hipMemcpyDeviceToHost));
    return HD_OK;
}

int hd_basis_device_states(const hd_basis* b, const uint64_t** d_states)
{
    if (!b || !d_states) return fail(HD_ERR_ARG, "hd_basis_device_states: null argument");
    *d_states = b->d_states;
    return HD_OK;
}

int hd_basis_find_index(const hd_basis* b, uint64_t offset, const uint64_t* d_states, int64_t n,
                        uint64_t* d_index)
{
    if (!b || (n > 0 && (!d_states || !d_index)) || n < 0)
        return fail(HD_ERR_ARG, "hd_basis_find_index: bad argument");
    if (n == 0) return HD_OK;
    const size_t lds = b->rt.bytes();
    const int grid = grid_for(n, 256, lds > 80 * 1024 ? 1 : 4);
    if (b->rt.is_lz) {
        HD_CHECK(lds_launch_ok((const void*)k_rank<true>, lds));
        hipLaunchKernelGGL(k_rank<true>, dim3(grid), dim3(256), lds, b->stream, b->d_rank,
                           (int)b->rt.blob.size(), d_states, n, offset, d_index);
    } else {
        HD_CHECK(lds_launch_ok((const void*)k_rank<false>, lds));
        hipLaunchKernelGGL(k_rank<false>, dim3(grid), dim3(256), lds, b->stream, b->d_rank,
                           (int)b->rt.blob.size(), d_states, n, offset, d_index);
    }
    HD_HIP(hipGetLastError());
    HD_HIP(hipStreamSynchronize(b->stream));
    return HD_OK;
}

}  // extern "C"
