// hd_ofl.hip -- flux-lattice quartic terms from Bloch coefficients.
//
// Reference: InteractingOflModel::GenerateQuarticTerms
// (src/optical_flux_lattice/interacting_ofl_model/interacting_ofl_model.cpp:457-590)
// with ArrayConvolution2D (src/utilities/mathematics/array_convolution.hpp:136-230)
// and KahanAccumulation (src/utilities/mathematics/kahan_arithmetic.hpp:42-76).
// For every folded table entry (k4 <= k3, k2), k1 and the reciprocal-lattice
// shift (gx, gy) follow from FindK1; V is the Kahan sum over the flat index i
// of conv1[i+o3*cao+o4]*conv2[i+o1*cao+o2] + conv3[..]*conv4[..], where the four
// 2D convolutions pair D with conj(B) and conj(A) with C per spin.
//
// One 64-lane workgroup per entry.  The eight Bloch arrays (conj applied) sit
// in LDS up to 64 KiB (cutoff 10), else they are read from L2; each lane owns
// a flat index i of the current 64, evaluates the four convolution outputs
// that i pairs (each its own Kahan sum over the overlap, x outer, y inner, as
// the reference) and stores the two products in LDS; lane 0 then adds them in
// the reference's order.  Every operation is the reference's: complex products
// as (ac - bd, ad + bc) with each product rounded, no fused multiply-add (the
// pragma below), so the table is bit-identical to the sequential C++.
#include <hip/hip_runtime.h>

#include <algorithm>
#include <cmath>
#include <cstdlib>
#include <vector>

#include "hd_internal.hpp"

#pragma clang fp contract(off)

namespace hd {

namespace {

constexpr int kOflLanes = 64;

struct OflArgs {
    int dimX, dimY, K;
    int X, Y, n;     // Bloch table dimensions (2c+1) and X*Y
    int cX, cY;      // convolution grid (2X-1) x (2Y-1)
    int cao;         // GetArrayConvolution1DSize(X, Y) = X + Y - 1
    int64_t entries;
    const double2* bloch;  // [K][n][2]: band-0 block, spin fastest
    double2* out;          // [entries]
};

__device__ __forceinline__ double2 cmul(double2 a, double2 b)
{
    const double ac = a.x * b.x, bd = a.y * b.y, ad = a.x * b.y, bc = a.y * b.x;
    return make_double2(ac - bd, ad + bc);
}

__device__ __forceinline__ void kahan(double2& sum, double2& corr, double2 b)
{
    const double tr = b.x - corr.x, ti = b.y - corr.y;
    const double sr = sum.x + tr, si = sum.y + ti;
    corr.x = (sr - sum.x) - tr;
    corr.y = (si - sum.y) - ti;
    sum.x = sr;
    sum.y = si;
}

// Two ArrayConvolution2D outputs at the same flat index t (row stride cY):
// (right0, left0) and (right1, left1) share the overlap geometry, so their two
// Kahan chains run interleaved in one loop (each in the reference's order, x
// outer, y inner); zero outside [top, cX-bottom) x [lo, cY-ro).  Element q of
// an array is at [q * S]; CL / CR conjugate the left / right element as read.
template <int S, bool CL, bool CR>
__device__ void conv_pair(const double2* right0, const double2* left0, const double2* right1, const double2* left1,
                          const OflArgs& a, int t, int top, int lo, int bottom, int ro, double2& out0, double2& out1)
{
    const int i = t / a.cY, j = t - (t / a.cY) * a.cY;
    double2 s0 = make_double2(0.0, 0.0), c0 = make_double2(0.0, 0.0);
    double2 s1 = make_double2(0.0, 0.0), c1 = make_double2(0.0, 0.0);
    if (i >= top && i < a.cX - bottom && j >= lo && j < a.cY - ro) {
        const int kX = min(i + 1, a.cX - i), kY = min(j + 1, a.cY - j);
        const int lxo = a.X > i ? 0 : a.X - kX, rxo = a.X > i ? a.X - kX : 0;
        const int lyo = a.Y > j ? 0 : a.Y - kY, ryo = a.Y > j ? a.Y - kY : 0;
        for (int x = 0; x < kX; ++x)
            for (int y = 0; y < kY; ++y) {
                const int li = ((x + lxo) * a.Y + (y + lyo)) * S, ri = ((x + rxo) * a.Y + (y + ryo)) * S;
                double2 l0 = left0[li], r0 = right0[ri], l1 = left1[li], r1 = right1[ri];
                if (CL) {
                    l0.y = -l0.y;
                    l1.y = -l1.y;
                }
                if (CR) {
                    r0.y = -r0.y;
                    r1.y = -r1.y;
                }
                kahan(s0, c0, cmul(l0, r0));
                kahan(s1, c1, cmul(l1, r1));
            }
    }
    out0 = s0;
    out1 = s1;
}

// LDS_IN: the eight arrays are staged in LDS with the conjugates applied
// (8 (2c+1)^2 complex values, up to 64 KiB); otherwise they are read from the
// Bloch table (L2-resident) with spin stride 2.  In both cases 64 product
// pairs at a time go through LDS to lane 0, which adds them in order.
template <bool LDS_IN>
__global__ void __launch_bounds__(kOflLanes) k_ofl_vterms(OflArgs a)
{
    extern __shared__ double2 lds2[];
    const int n = a.n;
    double2* P12 = lds2;
    double2* P34 = P12 + kOflLanes;
    double2* Au = P34 + kOflLanes;  // LDS_IN: conj(c^{k1}) up, down, conj(c^{k2}), c^{k3}, c^{k4}
    double2* Ad = Au + n;
    double2* Bu = Ad + n;
    double2* Bd = Bu + n;
    double2* Cu = Bd + n;
    double2* Cd = Cu + n;
    double2* Du = Cd + n;
    double2* Dd = Du + n;
    const int cdim = a.cX * a.cY;
    const int lane = threadIdx.x;
    for (int64_t e = blockIdx.x; e < a.entries; e += gridDim.x) {
        // folded index e = (k4*K + k3 - k4(k4+1)/2)*K + k2, k3 >= k4
        const int K = a.K;
        const int64_t row = e / K;
        const int k2 = (int)(e - row * K);
        int k4 = 0;
        while (k4 + 1 < K && (int64_t)(k4 + 1) * K - (int64_t)(k4 + 1) * k4 / 2 <= row) ++k4;
        const int k3 = (int)(row - ((int64_t)k4 * K - (int64_t)k4 * (k4 + 1) / 2));
        // FindK1 (interacting_ofl_model.cpp:374-407)
        const int fx = k3 / a.dimY + k4 / a.dimY - k2 / a.dimY;
        const int fy = k3 % a.dimY + k4 % a.dimY - k2 % a.dimY;
        const int sx = ((fx % a.dimX) + a.dimX) % a.dimX, sy = ((fy % a.dimY) + a.dimY) % a.dimY;
        const int gx = (sx - fx) / a.dimX, gy = (sy - fy) / a.dimY;
        const int k1 = sx * a.dimY + sy;
        const double2* b1 = a.bloch + (size_t)k1 * 2 * n;
        const double2* b2 = a.bloch + (size_t)k2 * 2 * n;
        const double2* b3 = a.bloch + (size_t)k3 * 2 * n;
        const double2* b4 = a.bloch + (size_t)k4 * 2 * n;
        if (LDS_IN) {
            for (int i = lane; i < n; i += kOflLanes) {
                const double2 a0 = b1[2 * i], a1 = b1[2 * i + 1], c0 = b2[2 * i], c1 = b2[2 * i + 1];
                Au[i] = make_double2(a0.x, -a0.y);
                Ad[i] = make_double2(a1.x, -a1.y);
                Bu[i] = make_double2(c0.x, -c0.y);
                Bd[i] = make_double2(c1.x, -c1.y);
                Cu[i] = b3[2 * i];
                Cd[i] = b3[2 * i + 1];
                Du[i] = b4[2 * i];
                Dd[i] = b4[2 * i + 1];
            }
            __syncthreads();
        }
        const int o1 = max(-gx, 0), o2 = max(-gy, 0), o3 = max(gx, 0), o4 = max(gy, 0);
        const int s1 = o3 * a.cao + o4, s2 = o1 * a.cao + o2;
        const int lim = cdim - o1 * a.cao - o2 - o3 * a.cao - o4;
        double2 sum = make_double2(0.0, 0.0), corr = make_double2(0.0, 0.0);
        for (int base = 0; base < lim; base += kOflLanes) {
            const int i = base + lane;
            if (i < lim) {
                // conv1 = (Bdown, Dup), conv2 = (Cdown, Aup), conv3 = (Bup, Ddown), conv4 = (Cup, Adown)
                double2 c1, c2, c3, c4;
                if (LDS_IN) {
                    conv_pair<1, false, false>(Bd, Du, Bu, Dd, a, i + s1, 0, 0, 0, 0, c1, c3);
                    conv_pair<1, false, false>(Cd, Au, Cu, Ad, a, i + s2, o1, o2, o3, o4, c2, c4);
                } else {
                    conv_pair<2, false, true>(b2 + 1, b4, b2, b4 + 1, a, i + s1, 0, 0, 0, 0, c1, c3);
                    conv_pair<2, true, false>(b3 + 1, b1, b3, b1 + 1, a, i + s2, o1, o2, o3, o4, c2, c4);
                }
                P12[lane] = cmul(c1, c2);
                P34[lane] = cmul(c3, c4);
            }
            __syncthreads();
            if (lane == 0) {
                const int m = min(kOflLanes, lim - base);
                for (int q = 0; q < m; ++q) {
                    kahan(sum, corr, P12[q]);
                    kahan(sum, corr, P34[q]);
                }
            }
            __syncthreads();
        }
        if (lane == 0) a.out[e] = sum;
    }
}

// Hybrid Wannier basis (ChangeToWannierBasis, interacting_ofl_model.cpp:820-905):
// one thread per (k2, k3, k4, x1).  k1 = x1*dimY + ky1 with ky1 fixed by ky
// momentum (FindK1 :410-432); the sum runs over k4x, k3x, k2x in the
// reference's nesting (k2x fastest), q1 from the Bloch-basis array k table,
// V from the folded array (GetK1 / GetVkkkk, term_tables.cpp:108-175), each
// term V * conj(U1 * U2) * U3 * U4 multiplied left to right.
struct WannierArgs {
    int dimX, dimY, K;
    double tol;
    const uint16_t* kt;  // folded Bloch-basis k table
    const double2* vt;   // folded Bloch-basis V
    const double2* u;    // change of basis, K x K, row (x, ky), column (kx, ky')
    uint8_t* nk1;        // [K^3]
    uint8_t* k1;         // [K^3 * dimX]
    double2* v;          // [K^3 * dimX]
};

__global__ void __launch_bounds__(256) k_wannier(WannierArgs a)
{
    const int K = a.K, X = a.dimX, Y = a.dimY;
    const int64_t total = (int64_t)K * K * K * X;
    const int64_t stride = (int64_t)gridDim.x * blockDim.x;
    for (int64_t idx = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; idx < total; idx += stride) {
        const int64_t e = idx / X;
        const int x1 = (int)(idx - e * X);
        const int k2 = (int)(e % K), k3 = (int)((e / K) % K), k4 = (int)(e / ((int64_t)K * K));
        const int k2y = k2 % Y, k3y = k3 % Y, k4y = k4 % Y;
        const int k1y = (((k3y + k4y - k2y) % Y) + Y) % Y;
        const int k1 = x1 * Y + k1y;
        double2 sum = make_double2(0.0, 0.0);
        for (int k4x = 0; k4x < X; ++k4x)
            for (int k3x = 0; k3x < X; ++k3x)
                for (int k2x = 0; k2x < X; ++k2x) {
                    const int q2 = k2x * Y + k2y, q3 = k3x * Y + k3y, q4 = k4x * Y + k4y;
                    const int t1 = q4 > q3 ? q3 : q4, t2 = q4 > q3 ? q4 : q3;
                    const int64_t row = ((int64_t)t1 * K + t2 - (int64_t)t1 * (t1 + 1) / 2) * K;
                    const int q1 = a.kt[row + q2];
                    const double2 V = a.vt[row + (q4 > q3 ? q1 : q2)];
                    double2 p = cmul(a.u[(int64_t)k1 * K + q1], a.u[(int64_t)k2 * K + q2]);
                    p.y = -p.y;
                    double2 t = cmul(V, p);
                    t = cmul(t, a.u[(int64_t)k3 * K + q3]);
                    t = cmul(t, a.u[(int64_t)k4 * K + q4]);
                    sum.x = sum.x + t.x;
                    sum.y = sum.y + t.y;
                }
        a.k1[idx] = (uint8_t)k1;
        a.v[idx] = hypot(sum.x, sum.y) > a.tol ? sum : make_double2(0.0, 0.0);
        if (x1 == 0) a.nk1[e] = (uint8_t)X;
    }
}

}  // namespace

}  // namespace hd

using namespace hd;

extern "C" {

int hd_ofl_ktable(int32_t dim_x, int32_t dim_y, uint16_t* h_ktable, int32_t* h_gx, int32_t* h_gy)
{
    if (!h_ktable) return fail(HD_ERR_ARG, "hd_ofl_ktable: null table");
    if (dim_x <= 0 || dim_y <= 0 || (int64_t)dim_x * dim_y > 64)
        return fail(HD_ERR_ARG, "hd_ofl_ktable: dim_x * dim_y must be in 1..64");
    const int K = dim_x * dim_y;
    int64_t it = 0;
    for (int k4 = 0; k4 < K; ++k4)
        for (int k3 = k4; k3 < K; ++k3)
            for (int k2 = 0; k2 < K; ++k2, ++it) {
                const int fx = k3 / dim_y + k4 / dim_y - k2 / dim_y;
                const int fy = k3 % dim_y + k4 % dim_y - k2 % dim_y;
                const int sx = ((fx % dim_x) + dim_x) % dim_x, sy = ((fy % dim_y) + dim_y) % dim_y;
                h_ktable[it] = (uint16_t)(sx * dim_y + sy);
                if (h_gx) h_gx[it] = (sx - fx) / dim_x;
                if (h_gy) h_gy[it] = (sy - fy) / dim_y;
            }
    return HD_OK;
}

int hd_ofl_vtable(int32_t dim_x, int32_t dim_y, int32_t cutoff, const double* h_bloch, double* h_vtable,
                  double* kernel_ms)
{
    if (!h_bloch || !h_vtable) return fail(HD_ERR_ARG, "hd_ofl_vtable: null argument");
    if (dim_x <= 0 || dim_y <= 0 || (int64_t)dim_x * dim_y > 64)
        return fail(HD_ERR_ARG, "hd_ofl_vtable: dim_x * dim_y must be in 1..64");
    if (cutoff < 0 || cutoff > 64) return fail(HD_ERR_ARG, "hd_ofl_vtable: cutoff must be in 0..64");
    OflArgs a{};
    a.dimX = dim_x;
    a.dimY = dim_y;
    a.K = dim_x * dim_y;
    a.X = a.Y = 2 * cutoff + 1;
    a.n = a.X * a.Y;
    a.cX = 2 * a.X - 1;
    a.cY = 2 * a.Y - 1;
    a.cao = a.X + a.Y - 1;
    a.entries = (int64_t)a.K * a.K * (a.K + 1) / 2;
    const bool lds_in = sizeof(double2) * 8 * (size_t)a.n <= 64 * 1024;
    const size_t lds = sizeof(double2) * (2 * kOflLanes + (lds_in ? 8 * (size_t)a.n : 0));
    int n_cu = 0;
    size_t max_lds = 0;
    HD_CHECK(device_props(&n_cu, &max_lds));
    const size_t bloch_n = (size_t)a.K * 2 * a.n;
    double2 *d_bloch = nullptr, *d_out = nullptr;
    hipStream_t st = nullptr;
    if (hipStreamCreateWithFlags(&st, hipStreamNonBlocking) != hipSuccess)
        return fail(HD_ERR_HIP, "stream creation failed");
    int rc = dalloc_t(&d_bloch, bloch_n, "Bloch coefficients");
    if (rc == HD_OK) rc = dalloc_t(&d_out, (size_t)a.entries, "quartic terms");
    if (rc == HD_OK) {
        HipErrors q;
        EventTimer ev;
        HD_ENQ(q, hipMemcpyAsync(d_bloch, h_bloch, bloch_n * sizeof(double2), hipMemcpyHostToDevice, st));
        a.bloch = d_bloch;
        a.out = d_out;
        const int grid = (int)std::min<int64_t>(a.entries, (int64_t)n_cu * 64);
        ev.start(st);
        if (lds_in) hipLaunchKernelGGL(k_ofl_vterms<true>, dim3(grid), dim3(kOflLanes), lds, st, a);
        else hipLaunchKernelGGL(k_ofl_vterms<false>, dim3(grid), dim3(kOflLanes), lds, st, a);
        HD_ENQ(q, hipGetLastError());
        const float ms = ev.stop(st);
        HD_ENQ(q, hipMemcpyAsync(h_vtable, d_out, (size_t)a.entries * sizeof(double2), hipMemcpyDeviceToHost, st));
        HD_ENQ(q, hipStreamSynchronize(st));
        rc = q.status();
        if (rc == HD_OK) rc = ev.status("hd_ofl_vtable");
        if (kernel_ms) *kernel_ms = ms;
    }
    dfree(d_bloch);
    dfree(d_out);
    (void)hipStreamDestroy(st);
    return rc;
}

int hd_ofl_wannier(int32_t dim_x, int32_t dim_y, double offset_x, const uint16_t* h_ktable, const double* h_vtable,
                   const double* h_eps, const double* h_u, double tol, uint8_t* h_nk1, uint16_t* h_k1,
                   double* h_v, uint8_t* h_nq1, uint16_t* h_q1, double* h_e, double* kernel_ms)
{
    if (!h_ktable || !h_vtable || !h_u || !h_nk1 || !h_k1 || !h_v)
        return fail(HD_ERR_ARG, "hd_ofl_wannier: null argument");
    if (h_eps && (!h_nq1 || !h_q1 || !h_e)) return fail(HD_ERR_ARG, "hd_ofl_wannier: null quadratic output");
    if (dim_x <= 0 || dim_y <= 0 || (int64_t)dim_x * dim_y > 64)
        return fail(HD_ERR_ARG, "hd_ofl_wannier: dim_x * dim_y must be in 1..64");
    const int K = dim_x * dim_y;
    const int64_t folded = (int64_t)K * K * (K + 1) / 2;
    for (int64_t i = 0; i < folded; ++i)
        if (h_ktable[i] >= K) return fail(HD_ERR_ARG, "hd_ofl_wannier: k table label >= K");
    // quadratic terms on the host: the backward DFT over kx of E/dim_x per ky
    // (DiscreteFourierTransform1D with FFTW_BACKWARD), then the phase of
    // |x2 - x1| (interacting_ofl_model.cpp:771-807)
    if (h_eps) {
        const long double two_pi = 6.283185307179586476925286766559005768L;
        std::vector<double> out_re((size_t)K), out_im((size_t)K);
        for (int ky = 0; ky < dim_y; ++ky)
            for (int j = 0; j < dim_x; ++j) {
                long double re = 0.0L, im = 0.0L;
                for (int kx = 0; kx < dim_x; ++kx) {
                    const double ire = h_eps[2 * (kx * dim_y + ky)] / (double)dim_x;
                    const double iim = h_eps[2 * (kx * dim_y + ky) + 1] / (double)dim_x;
                    const long double ang = two_pi * (long double)((kx * j) % dim_x) / (long double)dim_x;
                    const long double c = cosl(ang), sn = sinl(ang);
                    re += c * ire - sn * iim;
                    im += c * iim + sn * ire;
                }
                out_re[(size_t)ky * dim_x + j] = (double)re;
                out_im[(size_t)ky * dim_x + j] = (double)im;
            }
        const double pi = 3.1415926535897932384626433832795028841972;
        for (int x2 = 0; x2 < dim_x; ++x2)
            for (int ky = 0; ky < dim_y; ++ky) {
                const int k2 = x2 * dim_y + ky;
                h_nq1[k2] = (uint8_t)dim_x;
                for (int x1 = 0; x1 < dim_x; ++x1) {
                    const int d = std::abs(x2 - x1);
                    double th = 2.0 * pi;
                    th = th * (double)d;
                    th = th * offset_x;
                    th = th / (double)dim_x;
                    const double c = std::cos(th), sn = std::sin(th);
                    const double ore = out_re[(size_t)ky * dim_x + d], oim = out_im[(size_t)ky * dim_x + d];
                    const double pr0 = c * ore, pr1 = sn * oim, pi0 = c * oim, pi1 = sn * ore;
                    const size_t slot = (size_t)k2 * dim_x + x1;
                    h_q1[slot] = (uint16_t)(x1 * dim_y + ky);
                    h_e[2 * slot] = pr0 - pr1;
                    h_e[2 * slot + 1] = pi0 + pi1;
                }
            }
    }
    int n_cu = 0;
    HD_CHECK(device_props(&n_cu, nullptr));
    const int64_t nslots = (int64_t)K * K * K * dim_x;
    uint16_t* d_kt = nullptr;
    double2 *d_vt = nullptr, *d_u = nullptr, *d_v = nullptr;
    uint8_t *d_nk1 = nullptr, *d_k1 = nullptr;
    hipStream_t st = nullptr;
    if (hipStreamCreateWithFlags(&st, hipStreamNonBlocking) != hipSuccess)
        return fail(HD_ERR_HIP, "stream creation failed");
    int rc = dalloc_t(&d_kt, (size_t)folded, "Bloch k table");
    if (rc == HD_OK) rc = dalloc_t(&d_vt, (size_t)folded, "Bloch V table");
    if (rc == HD_OK) rc = dalloc_t(&d_u, (size_t)K * K, "change of basis");
    if (rc == HD_OK) rc = dalloc_t(&d_v, (size_t)nslots, "Wannier V");
    if (rc == HD_OK) rc = dalloc_t(&d_nk1, (size_t)K * K * K, "Wannier list lengths");
    if (rc == HD_OK) rc = dalloc_t(&d_k1, (size_t)nslots, "Wannier k1 lists");
    std::vector<uint8_t> k1b;
    if (rc == HD_OK) {
        HipErrors q;
        EventTimer ev;
        HD_ENQ(q, hipMemcpyAsync(d_kt, h_ktable, (size_t)folded * 2, hipMemcpyHostToDevice, st));
        HD_ENQ(q, hipMemcpyAsync(d_vt, h_vtable, (size_t)folded * sizeof(double2), hipMemcpyHostToDevice, st));
        HD_ENQ(q, hipMemcpyAsync(d_u, h_u, (size_t)K * K * sizeof(double2), hipMemcpyHostToDevice, st));
        WannierArgs a{dim_x, dim_y, K, tol, d_kt, d_vt, d_u, d_nk1, d_k1, d_v};
        const int grid = (int)std::max<int64_t>(1, std::min<int64_t>((nslots + 255) / 256, (int64_t)n_cu * 16));
        ev.start(st);
        hipLaunchKernelGGL(k_wannier, dim3(grid), dim3(256), 0, st, a);
        HD_ENQ(q, hipGetLastError());
        const float ms = ev.stop(st);
        k1b.resize((size_t)nslots);
        HD_ENQ(q, hipMemcpyAsync(h_v, d_v, (size_t)nslots * sizeof(double2), hipMemcpyDeviceToHost, st));
        HD_ENQ(q, hipMemcpyAsync(h_nk1, d_nk1, (size_t)K * K * K, hipMemcpyDeviceToHost, st));
        HD_ENQ(q, hipMemcpyAsync(k1b.data(), d_k1, (size_t)nslots, hipMemcpyDeviceToHost, st));
        HD_ENQ(q, hipStreamSynchronize(st));
        rc = q.status();
        if (rc == HD_OK) rc = ev.status("hd_ofl_wannier");
        if (kernel_ms) *kernel_ms = ms;
    }
    if (rc == HD_OK)
        for (int64_t i = 0; i < nslots; ++i) h_k1[i] = k1b[(size_t)i];
    dfree(d_kt);
    dfree(d_vt);
    dfree(d_u);
    dfree(d_v);
    dfree(d_nk1);
    dfree(d_k1);
    (void)hipStreamDestroy(st);
    return rc;
}

}  // extern "C"
