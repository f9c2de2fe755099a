// hd_mlp.hip -- learned Lanczos start vector: an amplitude MLP evaluated and
// trained with BF16 MFMA kernels (FP32 accumulation, FP32 master weights).
//
// No reference counterpart: the reference's only warm-start hook is the ARPACK
// initial-vector file (src/utilities/wrappers/arpack_wrapper.hpp:753-774,
// --use-initial).  The network maps an occupation bitstring plus Hamiltonian
// parameters to (log|psi_s|, sign logit_s) for S amplitudes (the S lowest
// eigenvectors it was trained on; S = 1: the ground state):
//     features (n_bits occupations, n_params parameters, zero padded to 64)
//     -> 256 -> 256 -> 256 (SiLU) -> 2 S
//
// All layers run through one GEMM kernel, C[M x N] = A[M x K] . B[N x K]^T with
// both operands K-contiguous in BF16, FP32 accumulation in
// v_mfma_f32_32x32x16_bf16 (a 64-lane wavefront owns a 32x32 output tile; four
// waves make a 64x64 block tile).  Epilogues fuse bias + SiLU (forward) and the
// SiLU derivative (backward) and write BF16 activations in row-major and, for
// training, transposed layout so that every backward GEMM again sees
// K-contiguous operands.  MFMA operand layout (lane l, 0..63):
//     A: row l%32, k = 8*(l/32) .. +7      B: column l%32, k = 8*(l/32) .. +7
//     D: column l%32, register i -> row (i/4)*8 + (l/32)*4 + i%4
#include <hip/hip_runtime.h>

#include <algorithm>
#include <cmath>
#include <cstring>
#include <random>

#include "hd_internal.hpp"

typedef __bf16 bf16x8 __attribute__((ext_vector_type(8)));
typedef float f32x16 __attribute__((ext_vector_type(16)));

namespace hd {

constexpr int kIn = 64;      // padded input width
constexpr int kW = 256;      // hidden width
constexpr int kOutPad = 64;  // padded output width (2 used)

enum { EPI_STORE = 0, EPI_BIAS_SILU = 1, EPI_DSILU = 2, EPI_BIAS = 3 };

HD_DEV float silu(float z) { return z / (1.f + __expf(-z)); }
HD_DEV float dsilu(float z)
{
    const float s = 1.f / (1.f + __expf(-z));
    return s * (1.f + z * (1.f - s));
}

struct GemmArgs {
    int M, N, K;
    const __bf16* A; int lda;
    const __bf16* B; int ldb;
    float* C; int ldc;          // FP32 output (pre-activation z, or raw result)
    const float* bias;          // EPI_BIAS_SILU / EPI_BIAS
    const float* zin; int ldz;  // EPI_DSILU: pre-activation of the layer below
    __bf16* H; int ldh;         // BF16 output, row-major (nullable)
    __bf16* HT; int ldht;       // BF16 output, transposed (nullable)
};

// M, N multiples of 64, K a multiple of 16 (buffers are padded by the callers).
template <int EPI>
__global__ void __launch_bounds__(256) k_gemm(GemmArgs g)
{
    const int lane = threadIdx.x & 63, wave = threadIdx.x >> 6;
    const int m0 = blockIdx.y * 64 + (wave >> 1) * 32;
    const int n0 = blockIdx.x * 64 + (wave & 1) * 32;
    const int r = lane & 31, kh = (lane >> 5) * 8;
    const __bf16* pa = g.A + (size_t)(m0 + r) * g.lda + kh;
    const __bf16* pb = g.B + (size_t)(n0 + r) * g.ldb + kh;
    f32x16 acc = {};
    for (int k0 = 0; k0 < g.K; k0 += 16) {
        const bf16x8 a = *reinterpret_cast<const bf16x8*>(pa + k0);
        const bf16x8 b = *reinterpret_cast<const bf16x8*>(pb + k0);
        acc = __builtin_amdgcn_mfma_f32_32x32x16_bf16(a, b, acc, 0, 0, 0);
    }
    const int col = n0 + r;
#pragma unroll
    for (int i = 0; i < 16; ++i) {
        const int row = m0 + (i >> 2) * 8 + (lane >> 5) * 4 + (i & 3);
        float v = acc[i];
        if (EPI == EPI_BIAS_SILU || EPI == EPI_BIAS) v += g.bias[col];
        if (g.C) g.C[(size_t)row * g.ldc + col] = v;
        float h = v;
        if (EPI == EPI_BIAS_SILU) h = silu(v);
        if (EPI == EPI_DSILU) h = v * dsilu(g.zin[(size_t)row * g.ldz + col]);
        if (g.H) g.H[(size_t)row * g.ldh + col] = (__bf16)h;
        if (g.HT) g.HT[(size_t)col * g.ldht + row] = (__bf16)h;
    }
}

// Input features: occupations of n_bits orbitals, then n_params parameters.
__global__ void k_featurize(int64_t n, int64_t npad, const uint64_t* __restrict__ states,
                            const float* __restrict__ params, int per_sample, int n_bits, int n_params,
                            __bf16* __restrict__ H0, __bf16* __restrict__ H0T)
{
    const int64_t b = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (b >= npad) return;
    const bool live = b < n;
    const uint64_t s = live ? states[b] : 0ull;
    const float* p = params + (per_sample ? b * n_params : 0);
    for (int k = 0; k < kIn; ++k) {
        float f = 0.f;
        if (live) {
            if (k < n_bits) f = (float)((s >> k) & 1ull);
            else if (k < n_bits + n_params) f = p[k - n_bits];
        }
        H0[b * kIn + k] = (__bf16)f;
        if (H0T) H0T[(size_t)k * npad + b] = (__bf16)f;
    }
}

// loss and dL/dout for out = (log|psi_s|, sign logit_s), s < S, against
// targets (log|psi_s|, sign_s, weight_s) per state (3 S floats per sample);
// lambda weights the sign term.
__global__ void k_loss(int64_t n, int64_t npad, int S, const float* __restrict__ out, const float* __restrict__ tgt,
                       float lambda, float inv_n, __bf16* __restrict__ d4, __bf16* __restrict__ d4T,
                       float* __restrict__ partial)
{
    __shared__ float red[256];
    const int64_t b = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    float l = 0.f;
    if (b < npad) {
        for (int c = 0; c < kOutPad; ++c) {
            const int st = c >> 1;
            float g = 0.f;
            if (b < n && st < S) {
                const float o0 = out[b * kOutPad + 2 * st], o1 = out[b * kOutPad + 2 * st + 1];
                const float* t = tgt + (b * S + st) * 3;
                const float a = t[0], sg = t[1], w = t[2];
                const float e = o0 - a;
                const float m = -sg * o1;  // softplus(m) = log(1 + exp(-s o1))
                if (c & 1) {
                    g = -lambda * w * sg * (1.f / (1.f + __expf(-m))) * inv_n;
                } else {
                    const float sp = m > 20.f ? m : log1pf(__expf(m));
                    l += w * (e * e + lambda * sp) * inv_n;
                    g = 2.f * w * e * inv_n;
                }
            }
            d4[b * kOutPad + c] = (__bf16)g;
            d4T[(size_t)c * npad + b] = (__bf16)g;
        }
    }
    red[threadIdx.x] = l;
    __syncthreads();
    for (int o = 128; o > 0; o >>= 1) {
        if ((int)threadIdx.x < o) red[threadIdx.x] += red[threadIdx.x + o];
        __syncthreads();
    }
    if (threadIdx.x == 0) partial[blockIdx.x] = red[0];
}

// db[c] = sum_b dT[c][b] (one block per output channel, fixed-order tree).
__global__ void k_bias_grad(const __bf16* __restrict__ dT, int64_t npad, float* __restrict__ db)
{
    __shared__ float red[256];
    const int c = blockIdx.x;
    float s = 0.f;
    for (int64_t b = threadIdx.x; b < npad; b += 256) s += (float)dT[(size_t)c * npad + b];
    red[threadIdx.x] = s;
    __syncthreads();
    for (int o = 128; o > 0; o >>= 1) {
        if ((int)threadIdx.x < o) red[threadIdx.x] += red[threadIdx.x + o];
        __syncthreads();
    }
    if (threadIdx.x == 0) db[c] = red[0];
}

// Adam (Kingma and Ba, ICLR 2015) on the FP32 master copy.
__global__ void k_adam(int64_t n, float* __restrict__ w, const float* __restrict__ gr, float* __restrict__ m,
                       float* __restrict__ v, float lr, float b1, float b2, float c1, float c2, float eps)
{
    const int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= n) return;
    const float g = gr[i];
    const float mi = b1 * m[i] + (1.f - b1) * g;
    const float vi = b2 * v[i] + (1.f - b2) * g * g;
    m[i] = mi;
    v[i] = vi;
    w[i] -= lr * (mi / c1) / (sqrtf(vi / c2) + eps);
}

// BF16 copies W (rows x cols) and W^T from the FP32 master.
__global__ void k_to_bf16(int rows, int cols, const float* __restrict__ w, __bf16* __restrict__ wb,
                          __bf16* __restrict__ wbT)
{
    const int i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= rows * cols) return;
    const int r = i / cols, c = i % cols;
    const __bf16 v = (__bf16)w[i];
    wb[i] = v;
    wbT[(size_t)c * rows + r] = v;
}

// ---------------------------------------------------------------- inference
//
// The whole network for kIT states per workgroup in one launch: each of the
// kIT/32 waves owns 32 states (one MFMA row tile) and keeps their features and
// activations in LDS (32 x 256 BF16); the weights are staged through LDS in
// K chunks of kIKc shared by the waves.  Every output is the MFMA sequence of
// k_gemm (same 32x32x16 tiles, K ascending in steps of 16) with the same
// epilogues, so the outputs equal the layer-by-layer path bit for bit, while
// HBM sees only the states and the outputs.
constexpr int kIT = 128;            // states per workgroup
constexpr int kIWaves = kIT / 32;   // waves per workgroup, one 32-state strip each
constexpr int kIKc = 32;            // K chunk of the staged weights

struct InferArgs {
    int64_t n;
    const uint64_t* states;
    const float* params;  // n_params values, per state when per_sample
    int per_sample, n_bits, n_params;
    const __bf16* w[4];   // BF16 weights, padded shapes (out x in)
    const float* b[4];
    int n_out;            // live outputs per state: 2 per amplitude (log|psi|, sign logit)
    float* out;           // n x n_out
    float* partial_max;   // per workgroup and amplitude s: largest output 2s over its states
};

__global__ void __launch_bounds__(kIWaves * 64) k_mlp_infer(InferArgs g)
{
    __shared__ __bf16 act[kIT * kW];
    __shared__ __bf16 wc[kW * kIKc];
    __shared__ float red[kIWaves][32];
    const int lane = threadIdx.x & 63, wave = threadIdx.x >> 6;
    const int r = lane & 31, kh = (lane >> 5) * 8;
    const int64_t s0 = (int64_t)blockIdx.x * kIT + wave * 32;
    __bf16* strip = act + (size_t)wave * 32 * kW;
    {   // features of state s0 + r: lane halves build columns 0-31 and 32-63
        const int64_t b = s0 + r;
        const bool live = b < g.n;
        const uint64_t st = live ? g.states[b] : 0ull;
        const float* p = g.params + (g.per_sample && live ? b * g.n_params : 0);
        for (int k = (lane >> 5) * 32; k < (lane >> 5) * 32 + 32; ++k) {
            float f = 0.f;
            if (live) {
                if (k < g.n_bits) f = (float)((st >> k) & 1ull);
                else if (k < g.n_bits + g.n_params) f = p[k - g.n_bits];
            }
            strip[r * kW + k] = (__bf16)f;
        }
    }
    // hidden layers: 64 -> 256 -> 256 -> 256, bias + SiLU
    for (int l = 0; l < 3; ++l) {
        const int K = l == 0 ? kIn : kW;
        f32x16 acc[8];
#pragma unroll
        for (int t = 0; t < 8; ++t) acc[t] = f32x16{};
        for (int kc = 0; kc < K; kc += kIKc) {
            __syncthreads();  // the previous chunk is consumed
            for (int i = threadIdx.x; i < kW * (kIKc / 8); i += kIWaves * 64) {
                const int nr = i / (kIKc / 8), kp = (i % (kIKc / 8)) * 8;
                *reinterpret_cast<bf16x8*>(&wc[nr * kIKc + kp]) =
                    *reinterpret_cast<const bf16x8*>(g.w[l] + (size_t)nr * K + kc + kp);
            }
            __syncthreads();
#pragma unroll
            for (int k16 = 0; k16 < kIKc; k16 += 16) {
                const bf16x8 a = *reinterpret_cast<const bf16x8*>(&strip[r * kW + kc + k16 + kh]);
#pragma unroll
                for (int t = 0; t < 8; ++t) {
                    const bf16x8 bv = *reinterpret_cast<const bf16x8*>(&wc[(t * 32 + r) * kIKc + k16 + kh]);
                    acc[t] = __builtin_amdgcn_mfma_f32_32x32x16_bf16(a, bv, acc[t], 0, 0, 0);
                }
            }
        }
        // every read of this wave's strip has been consumed by the MFMAs above;
        // other waves never read it
#pragma unroll
        for (int t = 0; t < 8; ++t)
#pragma unroll
            for (int i = 0; i < 16; ++i) {
                const int row = (i >> 2) * 8 + (lane >> 5) * 4 + (i & 3);
                const int col = t * 32 + r;
                strip[row * kW + col] = (__bf16)silu(acc[t][i] + g.b[l][col]);
            }
    }
    // output layer: 256 -> the first 32 (padded) outputs, bias
    f32x16 acc = f32x16{};
    for (int kc = 0; kc < kW; kc += kIKc) {
        __syncthreads();
        for (int i = threadIdx.x; i < 32 * (kIKc / 8); i += kIWaves * 64) {
            const int nr = i / (kIKc / 8), kp = (i % (kIKc / 8)) * 8;
            *reinterpret_cast<bf16x8*>(&wc[nr * kIKc + kp]) =
                *reinterpret_cast<const bf16x8*>(g.w[3] + (size_t)nr * kW + kc + kp);
        }
        __syncthreads();
#pragma unroll
        for (int k16 = 0; k16 < kIKc; k16 += 16) {
            const bf16x8 a = *reinterpret_cast<const bf16x8*>(&strip[r * kW + kc + k16 + kh]);
            const bf16x8 bv = *reinterpret_cast<const bf16x8*>(&wc[r * kIKc + k16 + kh]);
            acc = __builtin_amdgcn_mfma_f32_32x32x16_bf16(a, bv, acc, 0, 0, 0);
        }
    }
    float mx = -3.0e38f;  // this lane's column (even columns are log-amplitudes)
#pragma unroll
    for (int i = 0; i < 16; ++i) {
        const int row = (i >> 2) * 8 + (lane >> 5) * 4 + (i & 3);
        const int64_t b = s0 + row;
        if (r < g.n_out && b < g.n) {
            const float v = acc[i] + g.b[3][r];
            g.out[b * g.n_out + r] = v;
            mx = fmaxf(mx, v);
        }
    }
    mx = fmaxf(mx, __shfl_xor(mx, 32, 64));  // the two half-waves hold the same column
    if (lane < 32) red[wave][lane] = mx;
    __syncthreads();
    if ((int)threadIdx.x < g.n_out / 2) {
        const int c = 2 * threadIdx.x;
        float m = red[0][c];
        for (int w = 1; w < kIWaves; ++w) m = fmaxf(m, red[w][c]);
        g.partial_max[(int64_t)blockIdx.x * (g.n_out / 2) + threadIdx.x] = m;
    }
}

// shift[s] = max over workgroups of partial[w * S + s] (one block per amplitude, fixed order)
__global__ void __launch_bounds__(256) k_max_states(int64_t nblk, int S, const float* __restrict__ partial,
                                                    float* __restrict__ shift)
{
    __shared__ float red[256];
    const int s = blockIdx.x;
    float m = -3.0e38f;
    for (int64_t i = threadIdx.x; i < nblk; i += 256) m = fmaxf(m, partial[i * S + s]);
    red[threadIdx.x] = m;
    __syncthreads();
    for (int o = 128; o > 0; o >>= 1) {
        if ((int)threadIdx.x < o) red[threadIdx.x] = fmaxf(red[threadIdx.x], red[threadIdx.x + o]);
        __syncthreads();
    }
    if (threadIdx.x == 0) shift[s] = red[0];
}

// part[blk * S + s] = sum over this block's states of exp(2 (o_2s - shift_s))
__global__ void __launch_bounds__(256) k_norm2_states(int64_t n, int S, const float* __restrict__ out,
                                                      const float* __restrict__ shift, double* __restrict__ part)
{
    __shared__ double red[256];
    for (int s = 0; s < S; ++s) {
        double acc = 0.0;
        for (int64_t b = (int64_t)blockIdx.x * 256 + threadIdx.x; b < n; b += (int64_t)gridDim.x * 256)
            acc += exp(2.0 * (double)(out[b * 2 * S + 2 * s] - shift[s]));
        red[threadIdx.x] = acc;
        __syncthreads();
        for (int o = 128; o > 0; o >>= 1) {
            if ((int)threadIdx.x < o) red[threadIdx.x] += red[threadIdx.x + o];
            __syncthreads();
        }
        if (threadIdx.x == 0) part[(int64_t)blockIdx.x * S + s] = red[0];
        __syncthreads();
    }
}

// inv[s] = 1 / sqrt(sum over blocks of part[blk * S + s]) (one thread per amplitude, block order)
__global__ void k_inv_norm_states(int nb, int S, const double* __restrict__ part, double* __restrict__ inv)
{
    const int s = threadIdx.x;
    if (s >= S) return;
    double t = 0.0;
    for (int b = 0; b < nb; ++b) t += part[(int64_t)b * S + s];
    inv[s] = t > 0.0 ? 1.0 / sqrt(t) : 0.0;
}

// psi_b = sum_s exp(o_2s - shift_s) sign(o_2s+1) * scale_s; with one amplitude
// the scale is 1 (psi = exp(o0 - max o0) sign(o1)), with several each
// amplitude is normalised to unit length before the sum
__global__ void k_amplitude(int64_t n, int S, const float* __restrict__ out, const float* __restrict__ shift,
                            const double* __restrict__ inv, double* __restrict__ psi)
{
    const int64_t b = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (b >= n) return;
    double v = 0.0;
    for (int s = 0; s < S; ++s) {
        const float o0 = out[b * 2 * S + 2 * s], o1 = out[b * 2 * S + 2 * s + 1];
        const double a = exp((double)(o0 - shift[s])) * (S > 1 ? inv[s] : 1.0);
        v += o1 >= 0.f ? a : -a;
    }
    psi[b] = v;
}

}  // namespace hd

using namespace hd;

// ------------------------------------------------------------------ handle

struct hd_mlp {
    int n_bits = 0, n_params = 0;
    int n_states = 1;  // amplitudes the network predicts (2 outputs each)
    hipStream_t stream = nullptr;
    // layer l: out_l x in_l ; l = 0..3 with padded shapes
    int rows[4] = {kW, kW, kW, kOutPad};
    int cols[4] = {kIn, kW, kW, kW};
    int live_rows[4] = {kW, kW, kW, 2};
    float* w[4] = {};     // FP32 master (padded shape, padding kept at zero)
    float* b[4] = {};
    __bf16* wb[4] = {};   // BF16 copy
    __bf16* wbT[4] = {};  // BF16 transposed copy
    float *mw[4] = {}, *vw[4] = {}, *mb[4] = {}, *vb[4] = {};  // Adam moments
    float *gw[4] = {}, *gb[4] = {};
    int64_t step = 0;
    // work buffers for a batch of `cap` rows
    int64_t cap = 0;
    __bf16 *h[4] = {}, *hT[4] = {};  // h[0] = features; h[l] = layer-l output (l=1..3)
    float* z[4] = {};                // pre-activations of layers 1..3 (z[1..3]); z[0] unused
    float* out = nullptr;            // npad x 64
    __bf16 *d[5] = {}, *dT[5] = {};  // deltas: d[4] output layer, d[3..1] hidden
    float* partial = nullptr;
    uint64_t* st_buf = nullptr;      // staging for host-provided batches
    float* pr_buf = nullptr;
    float* tg_buf = nullptr;
    int64_t st_cap = 0;
};

namespace {

void mlp_free_work(hd_mlp* m)
{
    for (int l = 0; l < 4; ++l) {
        dfree(m->h[l]); dfree(m->hT[l]); dfree(m->z[l]);
        m->h[l] = m->hT[l] = nullptr;
        m->z[l] = nullptr;
    }
    for (int l = 0; l < 5; ++l) {
        dfree(m->d[l]); dfree(m->dT[l]);
        m->d[l] = m->dT[l] = nullptr;
    }
    dfree(m->out); dfree(m->partial);
    m->out = nullptr;
    m->partial = nullptr;
    m->cap = 0;
}

int mlp_work(hd_mlp* m, int64_t npad, bool train)
{
    if (m->cap >= npad && (!train || m->hT[0])) return HD_OK;
    mlp_free_work(m);
    const int64_t c = std::max<int64_t>(npad, 64);
    HD_CHECK(dalloc_t(&m->h[0], (size_t)c * kIn, "MLP features"));
    for (int l = 1; l < 4; ++l) {
        HD_CHECK(dalloc_t(&m->h[l], (size_t)c * kW, "MLP activations"));
        HD_CHECK(dalloc_t(&m->z[l], (size_t)c * kW, "MLP pre-activations"));
    }
    HD_CHECK(dalloc_t(&m->out, (size_t)c * kOutPad, "MLP output"));
    HD_CHECK(dalloc_t(&m->partial, (size_t)(c / 256 + 1024), "MLP partials"));
    if (train) {
        HD_CHECK(dalloc_t(&m->hT[0], (size_t)c * kIn, "MLP features^T"));
        for (int l = 1; l < 4; ++l) HD_CHECK(dalloc_t(&m->hT[l], (size_t)c * kW, "MLP activations^T"));
        HD_CHECK(dalloc_t(&m->d[4], (size_t)c * kOutPad, "MLP output delta"));
        HD_CHECK(dalloc_t(&m->dT[4], (size_t)c * kOutPad, "MLP output delta^T"));
        for (int l = 1; l < 4; ++l) {
            HD_CHECK(dalloc_t(&m->d[l], (size_t)c * kW, "MLP deltas"));
            HD_CHECK(dalloc_t(&m->dT[l], (size_t)c * kW, "MLP deltas^T"));
        }
    }
    m->cap = c;
    return HD_OK;
}

int mlp_refresh_bf16(hd_mlp* m)
{
    for (int l = 0; l < 4; ++l) {
        const int n = m->rows[l] * m->cols[l];
        hipLaunchKernelGGL(k_to_bf16, dim3((n + 255) / 256), dim3(256), 0, m->stream, m->rows[l], m->cols[l], m->w[l],
                           m->wb[l], m->wbT[l]);
    }
    HD_HIP(hipGetLastError());
    return HD_OK;
}

void gemm(hipStream_t st, int epi, const GemmArgs& g)
{
    const dim3 grid(g.N / 64, g.M / 64);
    switch (epi) {
    case EPI_STORE: hipLaunchKernelGGL(k_gemm<EPI_STORE>, grid, dim3(256), 0, st, g); break;
    case EPI_BIAS_SILU: hipLaunchKernelGGL(k_gemm<EPI_BIAS_SILU>, grid, dim3(256), 0, st, g); break;
    case EPI_DSILU: hipLaunchKernelGGL(k_gemm<EPI_DSILU>, grid, dim3(256), 0, st, g); break;
    default: hipLaunchKernelGGL(k_gemm<EPI_BIAS>, grid, dim3(256), 0, st, g); break;
    }
}

GemmArgs ga(int M, int N, int K, const __bf16* A, int lda, const __bf16* B, int ldb)
{
    GemmArgs g{};
    g.M = M; g.N = N; g.K = K;
    g.A = A; g.lda = lda; g.B = B; g.ldb = ldb;
    return g;
}

InferArgs infer_args(const hd_mlp* m, int64_t n, const uint64_t* states, const float* params, int per_sample,
                     float* out, float* partial_max)
{
    InferArgs a{};
    a.n = n;
    a.states = states;
    a.params = params;
    a.per_sample = per_sample;
    a.n_bits = m->n_bits;
    a.n_params = m->n_params;
    for (int l = 0; l < 4; ++l) {
        a.w[l] = m->wb[l];
        a.b[l] = m->b[l];
    }
    a.n_out = m->live_rows[3];
    a.out = out;
    a.partial_max = partial_max;
    return a;
}

// forward over npad rows already featurized in h[0] (and hT[0] when training)
void mlp_forward_core(hd_mlp* m, int64_t npad, bool train)
{
    const int M = (int)npad;
    for (int l = 0; l < 3; ++l) {
        const int K = l == 0 ? kIn : kW;
        GemmArgs g = ga(M, kW, K, m->h[l], K, m->wb[l], m->cols[l]);
        g.C = m->z[l + 1]; g.ldc = kW;
        g.bias = m->b[l];
        g.H = m->h[l + 1]; g.ldh = kW;
        if (train) { g.HT = m->hT[l + 1]; g.ldht = M; }
        gemm(m->stream, EPI_BIAS_SILU, g);
    }
    GemmArgs g = ga(M, kOutPad, kW, m->h[3], kW, m->wb[3], kW);
    g.C = m->out; g.ldc = kOutPad;
    g.bias = m->b[3];
    gemm(m->stream, EPI_BIAS, g);
}

// gradients of the loss into gw/gb; returns the loss
int mlp_backward(hd_mlp* m, int64_t n, int64_t npad, const float* d_tgt, float lambda, float* loss)
{
    const hipStream_t st = m->stream;
    const int M = (int)npad;
    const int nblk = (int)((npad + 255) / 256);
    hipLaunchKernelGGL(k_loss, dim3(nblk), dim3(256), 0, st, n, npad, m->n_states, m->out, d_tgt, lambda, 1.f / (float)n,
                       m->d[4], m->dT[4], m->partial);
    // dW4 = d4^T h3 ; db4
    {
        GemmArgs g = ga(kOutPad, kW, M, m->dT[4], M, m->hT[3], M);
        g.C = m->gw[3]; g.ldc = kW;
        gemm(st, EPI_STORE, g);
        hipLaunchKernelGGL(k_bias_grad, dim3(kOutPad), dim3(256), 0, st, m->dT[4], npad, m->gb[3]);
    }
    // hidden layers l = 3, 2, 1: d_l = (d_{l+1} W_{l+1}) * silu'(z_l); dW_l = d_l^T h_{l-1}
    for (int l = 3; l >= 1; --l) {
        const int Kd = l == 3 ? kOutPad : kW;  // width of d_{l+1}
        GemmArgs g = ga(M, kW, Kd, m->d[l + 1 == 4 ? 4 : l + 1], Kd, m->wbT[l], m->rows[l]);
        g.zin = m->z[l]; g.ldz = kW;
        g.H = m->d[l]; g.ldh = kW;
        g.HT = m->dT[l]; g.ldht = M;
        gemm(st, EPI_DSILU, g);
        const int K_in = l == 1 ? kIn : kW;
        GemmArgs gw = ga(kW, K_in, M, m->dT[l], M, m->hT[l - 1], M);
        gw.C = m->gw[l - 1]; gw.ldc = K_in;
        gemm(st, EPI_STORE, gw);
        hipLaunchKernelGGL(k_bias_grad, dim3(kW), dim3(256), 0, st, m->dT[l], npad, m->gb[l - 1]);
    }
    HD_HIP(hipGetLastError());
    std::vector<float> part(nblk);
    HD_HIP(hipMemcpyAsync(part.data(), m->partial, sizeof(float) * nblk, hipMemcpyDeviceToHost, st));
    HD_HIP(hipStreamSynchronize(st));
    double s = 0.0;
    for (float p : part) s += p;
    if (loss) *loss = (float)s;
    return HD_OK;
}

int stage_batch(hd_mlp* m, const uint64_t* states, const float* params, const float* tgt, int64_t n, bool host,
                const uint64_t** ds, const float** dp, const float** dt, int per_sample)
{
    if (!host) {
        *ds = states;
        *dp = params;
        *dt = tgt;
        return HD_OK;
    }
    if (m->st_cap < n) {
        dfree(m->st_buf); dfree(m->pr_buf); dfree(m->tg_buf);
        m->st_buf = nullptr; m->pr_buf = nullptr; m->tg_buf = nullptr;
        HD_CHECK(dalloc_t(&m->st_buf, (size_t)n, "MLP staged states"));
        HD_CHECK(dalloc_t(&m->pr_buf, (size_t)n * std::max(m->n_params, 1), "MLP staged params"));
        HD_CHECK(dalloc_t(&m->tg_buf, (size_t)n * 3 * m->n_states, "MLP staged targets"));
        m->st_cap = n;
    }
    HD_HIP(hipMemcpyAsync(m->st_buf, states, sizeof(uint64_t) * n, hipMemcpyHostToDevice, m->stream));
    HD_HIP(hipMemcpyAsync(m->pr_buf, params, sizeof(float) * (per_sample ? n : 1) * m->n_params,
                          hipMemcpyHostToDevice, m->stream));
    if (tgt)
        HD_HIP(hipMemcpyAsync(m->tg_buf, tgt, sizeof(float) * n * 3 * m->n_states, hipMemcpyHostToDevice, m->stream));
    *ds = m->st_buf;
    *dp = m->pr_buf;
    *dt = m->tg_buf;
    return HD_OK;
}

int64_t pad64(int64_t n) { return (n + 63) / 64 * 64; }

int gradients_impl(hd_mlp* m, const uint64_t* states, const float* params, const float* tgt, int64_t n, float lambda,
                   int host, float* loss)
{
    if (n <= 0) return fail(HD_ERR_ARG, "empty batch");
    const int64_t npad = pad64(n);
    HD_CHECK(mlp_work(m, npad, true));
    const uint64_t* ds;
    const float *dp, *dt;
    HD_CHECK(stage_batch(m, states, params, tgt, n, host != 0, &ds, &dp, &dt, 1));
    hipLaunchKernelGGL(k_featurize, dim3((unsigned)((npad + 255) / 256)), dim3(256), 0, m->stream, n, npad, ds, dp, 1,
                       m->n_bits, m->n_params, m->h[0], m->hT[0]);
    mlp_forward_core(m, npad, true);
    return mlp_backward(m, n, npad, dt, lambda, loss);
}

}  // namespace

extern "C" {

int hd_mlp_create(int32_t n_bits, int32_t n_params, uint64_t seed, hd_mlp** out)
{
    return hd_mlp_create_states(n_bits, n_params, 1, seed, out);
}

int hd_mlp_create_states(int32_t n_bits, int32_t n_params, int32_t n_states, uint64_t seed, hd_mlp** out)
{
    if (!out || n_bits <= 0 || n_params < 0 || n_bits + n_params > kIn)
        return fail(HD_ERR_ARG, "hd_mlp_create: need 0 < n_bits and n_bits + n_params <= 64");
    if (n_states < 1 || n_states > 16) return fail(HD_ERR_ARG, "hd_mlp_create_states: 1 to 16 amplitudes");
    *out = nullptr;
    hd_mlp* m = new hd_mlp();
    m->n_bits = n_bits;
    m->n_params = n_params;
    m->n_states = n_states;
    m->live_rows[3] = 2 * n_states;
    if (hipStreamCreateWithFlags(&m->stream, hipStreamNonBlocking) != hipSuccess) {
        delete m;
        return fail(HD_ERR_HIP, "stream creation failed");
    }
    std::mt19937_64 rng(seed);
    int st = HD_OK;
    HipErrors q;
    for (int l = 0; l < 4 && st == HD_OK; ++l) {
        const int R = m->rows[l], C = m->cols[l];
        const size_t n = (size_t)R * C;
        std::vector<float> w(n, 0.f), bz(R, 0.f);
        // He-normal over the live fan-in (n_bits + n_params for the first layer)
        const int fan_in = l == 0 ? n_bits + n_params : kW;
        std::normal_distribution<float> nd(0.f, std::sqrt(2.f / (float)fan_in) * (l == 3 ? 0.1f : 1.f));
        for (int r = 0; r < m->live_rows[l]; ++r)
            for (int c = 0; c < (l == 0 ? fan_in : C); ++c) w[(size_t)r * C + c] = nd(rng);
        if ((st = dalloc_t(&m->w[l], n, "MLP weights")) != HD_OK) break;
        if ((st = dalloc_t(&m->b[l], R, "MLP biases")) != HD_OK) break;
        if ((st = dalloc_t(&m->wb[l], n, "MLP bf16 weights")) != HD_OK) break;
        if ((st = dalloc_t(&m->wbT[l], n, "MLP bf16 weights^T")) != HD_OK) break;
        if ((st = dalloc_t(&m->gw[l], n, "MLP weight grads")) != HD_OK) break;
        if ((st = dalloc_t(&m->gb[l], R, "MLP bias grads")) != HD_OK) break;
        for (float** p : {&m->mw[l], &m->vw[l]})
            if ((st = dalloc_t(p, n, "Adam moments")) != HD_OK) break;
        for (float** p : {&m->mb[l], &m->vb[l]})
            if ((st = dalloc_t(p, R, "Adam moments")) != HD_OK) break;
        if (st != HD_OK) break;
        // synchronous copies: w and bz are released at the end of this iteration
        HD_ENQ(q, hipMemcpy(m->w[l], w.data(), n * 4, hipMemcpyHostToDevice));
        HD_ENQ(q, hipMemcpy(m->b[l], bz.data(), (size_t)R * 4, hipMemcpyHostToDevice));
        HD_ENQ(q, hipMemsetAsync(m->mw[l], 0, n * 4, m->stream));
        HD_ENQ(q, hipMemsetAsync(m->vw[l], 0, n * 4, m->stream));
        HD_ENQ(q, hipMemsetAsync(m->mb[l], 0, (size_t)R * 4, m->stream));
        HD_ENQ(q, hipMemsetAsync(m->vb[l], 0, (size_t)R * 4, m->stream));
    }
    if (st == HD_OK) st = q.status();
    if (st == HD_OK) st = mlp_refresh_bf16(m);
    if (st == HD_OK && hipStreamSynchronize(m->stream) != hipSuccess) st = fail(HD_ERR_HIP, "MLP init failed");
    if (st != HD_OK) {
        const std::string msg = hd_last_error();
        hd_mlp_destroy(m);
        set_error(msg);
        return st;
    }
    *out = m;
    return HD_OK;
}

int hd_mlp_destroy(hd_mlp* m)
{
    if (!m) return HD_OK;
    mlp_free_work(m);
    for (int l = 0; l < 4; ++l) {
        for (void* p : {(void*)m->w[l], (void*)m->b[l], (void*)m->wb[l], (void*)m->wbT[l], (void*)m->gw[l],
                        (void*)m->gb[l], (void*)m->mw[l], (void*)m->vw[l], (void*)m->mb[l], (void*)m->vb[l]})
            dfree(p);
    }
    dfree(m->st_buf); dfree(m->pr_buf); dfree(m->tg_buf);
    if (m->stream) (void)hipStreamDestroy(m->stream);
    delete m;
    return HD_OK;
}

int hd_mlp_num_weights(const hd_mlp* m, int64_t* count)
{
    if (!m || !count) return fail(HD_ERR_ARG, "null argument");
    int64_t c = 0;
    for (int l = 0; l < 4; ++l) c += (int64_t)m->rows[l] * m->cols[l] + m->rows[l];
    *count = c;
    return HD_OK;
}

// Flat FP32 layout: for each layer, W (rows x cols, padded shape) then b.
static int copy_params(hd_mlp* m, float* h, bool to_host, bool grads)
{
    size_t off = 0;
    for (int l = 0; l < 4; ++l) {
        const size_t n = (size_t)m->rows[l] * m->cols[l];
        float* wsrc = grads ? m->gw[l] : m->w[l];
        float* bsrc = grads ? m->gb[l] : m->b[l];
        if (to_host) {
            HD_HIP(hipMemcpyAsync(h + off, wsrc, n * 4, hipMemcpyDeviceToHost, m->stream));
            HD_HIP(hipMemcpyAsync(h + off + n, bsrc, (size_t)m->rows[l] * 4, hipMemcpyDeviceToHost, m->stream));
        } else {
            HD_HIP(hipMemcpyAsync(wsrc, h + off, n * 4, hipMemcpyHostToDevice, m->stream));
            HD_HIP(hipMemcpyAsync(bsrc, h + off + n, (size_t)m->rows[l] * 4, hipMemcpyHostToDevice, m->stream));
        }
        off += n + m->rows[l];
    }
    if (!to_host && !grads) HD_CHECK(mlp_refresh_bf16(m));
    HD_HIP(hipStreamSynchronize(m->stream));
    return HD_OK;
}

int hd_mlp_get_weights(hd_mlp* m, float* h_out)
{
    if (!m || !h_out) return fail(HD_ERR_ARG, "null argument");
    return copy_params(m, h_out, true, false);
}

int hd_mlp_set_weights(hd_mlp* m, const float* h_in)
{
    if (!m || !h_in) return fail(HD_ERR_ARG, "null argument");
    return copy_params(m, const_cast<float*>(h_in), false, false);
}

int hd_mlp_forward(hd_mlp* m, const uint64_t* h_states, const float* h_params, int64_t n, float* h_out)
{
    if (!m || !h_states || !h_params || !h_out || n <= 0) return fail(HD_ERR_ARG, "hd_mlp_forward: bad argument");
    const uint64_t* ds;
    const float *dp, *dt;
    HD_CHECK(stage_batch(m, h_states, h_params, nullptr, n, true, &ds, &dp, &dt, 1));
    const int64_t nblk = (n + kIT - 1) / kIT;
    float *dout = nullptr, *dmax = nullptr;
    const int S = m->n_states;
    HD_CHECK(dalloc_t(&dout, (size_t)n * 2 * S, "MLP outputs"));
    int st = dalloc_t(&dmax, (size_t)nblk * S, "MLP output maxima");
    if (st == HD_OK) {
        InferArgs a = infer_args(m, n, ds, dp, 1, dout, dmax);
        hipLaunchKernelGGL(k_mlp_infer, dim3((unsigned)nblk), dim3(kIWaves * 64), 0, m->stream, a);
        HipErrors q;
        HD_ENQ(q, hipGetLastError());
        HD_ENQ(q, hipMemcpyAsync(h_out, dout, sizeof(float) * 2 * S * n, hipMemcpyDeviceToHost, m->stream));
        HD_ENQ(q, hipStreamSynchronize(m->stream));
        st = q.status();
    }
    dfree(dout);
    dfree(dmax);
    return st;
}

int hd_mlp_gradients(hd_mlp* m, const uint64_t* h_states, const float* h_params, const float* h_targets, int64_t n,
                     float lambda, float* loss, float* h_grads)
{
    if (!m || !h_states || !h_params || !h_targets || n <= 0) return fail(HD_ERR_ARG, "hd_mlp_gradients: bad argument");
    HD_CHECK(gradients_impl(m, h_states, h_params, h_targets, n, lambda, 1, loss));
    if (h_grads) HD_CHECK(copy_params(m, h_grads, true, true));
    return HD_OK;
}

int hd_mlp_train_step(hd_mlp* m, const uint64_t* d_states, const float* d_params, const float* d_targets, int64_t n,
                      float lambda, float lr, float* loss)
{
    if (!m || !d_states || !d_params || !d_targets || n <= 0) return fail(HD_ERR_ARG, "hd_mlp_train_step: bad argument");
    HD_CHECK(gradients_impl(m, d_states, d_params, d_targets, n, lambda, 0, loss));
    m->step += 1;
    const float b1 = 0.9f, b2 = 0.999f;
    const float c1 = 1.f - std::pow(b1, (float)m->step), c2 = 1.f - std::pow(b2, (float)m->step);
    for (int l = 0; l < 4; ++l) {
        const int64_t nw = (int64_t)m->rows[l] * m->cols[l];
        hipLaunchKernelGGL(k_adam, dim3((unsigned)((nw + 255) / 256)), dim3(256), 0, m->stream, nw, m->w[l], m->gw[l],
                           m->mw[l], m->vw[l], lr, b1, b2, c1, c2, 1e-8f);
        hipLaunchKernelGGL(k_adam, dim3((unsigned)((m->rows[l] + 255) / 256)), dim3(256), 0, m->stream,
                           (int64_t)m->rows[l], m->b[l], m->gb[l], m->mb[l], m->vb[l], lr, b1, b2, c1, c2, 1e-8f);
    }
    HD_HIP(hipGetLastError());
    return mlp_refresh_bf16(m);
}

int hd_mlp_start_vector(hd_mlp* m, const hd_basis* basis, const float* h_params, double* d_vec)
{
    if (!m || !basis || !h_params || !d_vec) return fail(HD_ERR_ARG, "hd_mlp_start_vector: bad argument");
    if (basis->Norb != m->n_bits) return fail(HD_ERR_ARG, "network input width differs from the orbital count");
    const int64_t n = basis->dim;
    if (n == 0) return HD_OK;
    const int S = m->n_states;
    const int64_t nblk = (n + kIT - 1) / kIT;
    const int nred = (int)std::min<int64_t>((n + 255) / 256, 1024);
    float *dpar = nullptr, *dout = nullptr, *dmax = nullptr;
    double* dnorm = nullptr;
    const int np = std::max(m->n_params, 1);
    int st = dalloc_t(&dpar, (size_t)np + S, "MLP parameters and shifts");
    if (st == HD_OK) st = dalloc_t(&dout, (size_t)n * 2 * S, "start vector outputs");
    if (st == HD_OK) st = dalloc_t(&dmax, (size_t)nblk * S, "start vector output maxima");
    if (st == HD_OK) st = dalloc_t(&dnorm, (size_t)(nred + 1) * S, "start vector norms");
    if (st == HD_OK) {
        // one pass of the fused network over the sector, then per amplitude the
        // largest log-amplitude (and, with several, the norm), then the sum
        HipErrors q;
        HD_ENQ(q, hipMemcpyAsync(dpar, h_params, sizeof(float) * m->n_params, hipMemcpyHostToDevice, m->stream));
        InferArgs a = infer_args(m, n, basis->d_states, dpar, 0, dout, dmax);
        hipLaunchKernelGGL(k_mlp_infer, dim3((unsigned)nblk), dim3(kIWaves * 64), 0, m->stream, a);
        HD_ENQ(q, hipGetLastError());
        float* shift = dpar + np;
        hipLaunchKernelGGL(k_max_states, dim3(S), dim3(256), 0, m->stream, nblk, S, dmax, shift);
        HD_ENQ(q, hipGetLastError());
        double* inv = dnorm + (size_t)nred * S;
        if (S > 1) {
            hipLaunchKernelGGL(k_norm2_states, dim3(nred), dim3(256), 0, m->stream, n, S, dout, shift, dnorm);
            hipLaunchKernelGGL(k_inv_norm_states, dim3(1), dim3(64), 0, m->stream, nred, S, dnorm, inv);
            HD_ENQ(q, hipGetLastError());
        }
        hipLaunchKernelGGL(k_amplitude, dim3((unsigned)((n + 255) / 256)), dim3(256), 0, m->stream, n, S, dout, shift,
                           inv, d_vec);
        HD_ENQ(q, hipGetLastError());
        HD_ENQ(q, hipStreamSynchronize(m->stream));
        st = q.status();
    }
    dfree(dpar);
    dfree(dout);
    dfree(dmax);
    dfree(dnorm);
    return st;
}

int hd_mlp_num_states(const hd_mlp* m, int32_t* n_states)
{
    if (!m || !n_states) return fail(HD_ERR_ARG, "hd_mlp_num_states: null argument");
    *n_states = m->n_states;
    return HD_OK;
}

// Raw GEMM check: C (M x N, FP32) = A (M x K, BF16 bits) . B (N x K, BF16 bits)^T
// through the MFMA kernel; M, N multiples of 64, K of 16 (host buffers).
int hd_mlp_gemm_check(int32_t M, int32_t N, int32_t K, const uint16_t* h_a, const uint16_t* h_b, float* h_c)
{
    if (M <= 0 || N <= 0 || K <= 0 || M % 64 || N % 64 || K % 16 || !h_a || !h_b || !h_c)
        return fail(HD_ERR_ARG, "hd_mlp_gemm_check: M, N multiples of 64 and K of 16");
    __bf16 *a = nullptr, *b = nullptr;
    float* c = nullptr;
    HD_CHECK(dalloc_t(&a, (size_t)M * K, "gemm A"));
    int st = dalloc_t(&b, (size_t)N * K, "gemm B");
    if (st == HD_OK) st = dalloc_t(&c, (size_t)M * N, "gemm C");
    if (st == HD_OK) {
        HipErrors q;
        HD_ENQ(q, hipMemcpy(a, h_a, (size_t)M * K * 2, hipMemcpyHostToDevice));
        HD_ENQ(q, hipMemcpy(b, h_b, (size_t)N * K * 2, hipMemcpyHostToDevice));
        st = q.status();
    }
    if (st == HD_OK) {
        GemmArgs g = ga(M, N, K, a, K, b, K);
        g.C = c; g.ldc = N;
        gemm(0, EPI_STORE, g);
        if (hipGetLastError() != hipSuccess || hipMemcpy(h_c, c, (size_t)M * N * 4, hipMemcpyDeviceToHost) != hipSuccess)
            st = fail(HD_ERR_HIP, "gemm check failed");
    }
    dfree(a); dfree(b); dfree(c);
    return st;
}

}  // extern "C"
