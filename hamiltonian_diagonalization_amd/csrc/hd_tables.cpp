// hd_tables.cpp -- sphere term tables on the host (reference layer a8/a9).
//
// Reference: BaseGenerate4KTable and BaseTermsFromPseudopotentials
// (src/fqhe_sphere/pseudopotential_model/pseudopotential_model_base.hpp:62-229)
// with the Clebsch-Gordan lookup of src/utilities/wrappers/clebsch_gordan_wrapper.hpp:93-173
// (coefficients with |CG| <= 1e-12 are not stored and read as zero).  The
// reference obtains CG coefficients from a numerical SU(N) algorithm; here each
// coefficient is evaluated exactly with Racah's closed form in big-integer
// arithmetic and rounded once to double (correctly rounded: the square root of
// the exact rational CG^2 is checked against the squared midpoints of its
// neighbouring doubles), and the table is accumulated in the reference's
// floating-point operation order.
#include <algorithm>
#include <cmath>
#include <cstdint>
#include <map>
#include <mutex>
#include <string>
#include <vector>

#include "hd_internal.hpp"

namespace hd {
namespace {

// Unsigned big integer, little-endian base 2^32 limbs.
struct Big {
    std::vector<uint32_t> d;
    Big() {}
    explicit Big(uint64_t v)
    {
        while (v) {
            d.push_back((uint32_t)v);
            v >>= 32;
        }
    }
    bool zero() const { return d.empty(); }
    void trim()
    {
        while (!d.empty() && d.back() == 0) d.pop_back();
    }
    void mul_small(uint32_t m)
    {
        uint64_t carry = 0;
        for (auto& x : d) {
            const uint64_t p = (uint64_t)x * m + carry;
            x = (uint32_t)p;
            carry = p >> 32;
        }
        if (carry) d.push_back((uint32_t)carry);
        if (m == 0) d.clear();
    }
};

int cmp(const Big& a, const Big& b)
{
    if (a.d.size() != b.d.size()) return a.d.size() < b.d.size() ? -1 : 1;
    for (size_t i = a.d.size(); i-- > 0;)
        if (a.d[i] != b.d[i]) return a.d[i] < b.d[i] ? -1 : 1;
    return 0;
}

Big add(const Big& a, const Big& b)
{
    Big r;
    const size_t n = std::max(a.d.size(), b.d.size());
    r.d.resize(n);
    uint64_t carry = 0;
    for (size_t i = 0; i < n; ++i) {
        const uint64_t s = (uint64_t)(i < a.d.size() ? a.d[i] : 0) + (i < b.d.size() ? b.d[i] : 0) + carry;
        r.d[i] = (uint32_t)s;
        carry = s >> 32;
    }
    if (carry) r.d.push_back((uint32_t)carry);
    return r;
}

Big subtract(const Big& a, const Big& b)  // a >= b
{
    Big r;
    r.d.resize(a.d.size());
    int64_t borrow = 0;
    for (size_t i = 0; i < a.d.size(); ++i) {
        int64_t s = (int64_t)a.d[i] - (i < b.d.size() ? b.d[i] : 0) - borrow;
        borrow = s < 0;
        if (s < 0) s += (int64_t)1 << 32;
        r.d[i] = (uint32_t)s;
    }
    r.trim();
    return r;
}

Big mul(const Big& a, const Big& b)
{
    if (a.zero() || b.zero()) return Big();
    std::vector<uint64_t> acc(a.d.size() + b.d.size() + 1, 0);
    Big r;
    r.d.assign(a.d.size() + b.d.size(), 0);
    for (size_t i = 0; i < a.d.size(); ++i) {
        uint64_t carry = 0;
        for (size_t j = 0; j < b.d.size(); ++j) {
            const unsigned __int128 p = (unsigned __int128)a.d[i] * b.d[j] + r.d[i + j] + carry;
            r.d[i + j] = (uint32_t)p;
            carry = (uint64_t)(p >> 32);
        }
        size_t k = i + b.d.size();
        while (carry) {
            const uint64_t s = (uint64_t)r.d[k] + carry;
            r.d[k] = (uint32_t)s;
            carry = s >> 32;
            ++k;
        }
    }
    r.trim();
    return r;
}

// x = mantissa * 2^exp with the top 64 bits of the integer as mantissa.
long double to_ld(const Big& a, long& exp2)
{
    if (a.zero()) {
        exp2 = 0;
        return 0.0L;
    }
    const size_t n = a.d.size();
    long double m = 0.0L;
    const size_t take = std::min<size_t>(3, n);
    for (size_t i = 0; i < take; ++i) m = m * 4294967296.0L + (long double)a.d[n - 1 - i];
    exp2 = (long)(32 * (n - take));
    return m;
}

Big shl(const Big& a, long bits)  // a * 2^bits, bits >= 0
{
    if (a.zero() || bits == 0) return a;
    Big r;
    const size_t limbs = (size_t)(bits / 32);
    const int sh = (int)(bits % 32);
    r.d.assign(limbs, 0);
    uint32_t carry = 0;
    for (const uint32_t x : a.d) {
        r.d.push_back(sh ? (x << sh) | carry : x);
        carry = sh ? x >> (32 - sh) : 0;
    }
    if (carry) r.d.push_back(carry);
    r.trim();
    return r;
}

// Sign of num/den - ((a + b) / 2)^2 for positive finite doubles a, b, exactly:
// with a = ma 2^xa, b = mb 2^xb (53-bit integer mantissas) and e0 = min(xa, xb),
// ((a + b)/2)^2 = (ma 2^(xa-e0) + mb 2^(xb-e0))^2 2^(2 e0 - 2).
int cmp_square_of_mid(const Big& num, const Big& den, double a, double b)
{
    int ea, eb;
    const double fa = std::frexp(a, &ea), fb = std::frexp(b, &eb);
    const long xa = ea - 53, xb = eb - 53, e0 = std::min(xa, xb);
    const Big s = add(shl(Big((uint64_t)std::ldexp(fa, 53)), xa - e0), shl(Big((uint64_t)std::ldexp(fb, 53)), xb - e0));
    Big rhs = mul(den, mul(s, s));
    Big lhs = num;
    const long sh = 2 * e0 - 2;
    if (sh >= 0) rhs = shl(rhs, sh);
    else lhs = shl(lhs, -sh);
    return cmp(lhs, rhs);
}

// sqrt(num/den) rounded to the nearest double (ties to even): a long double
// first guess, moved by whole ulps while num/den lies beyond the square of the
// midpoint to a neighbour.
double sqrt_ratio_rounded(const Big& num, const Big& den)
{
    long en, ed;
    const long double mn = to_ld(num, en), md = to_ld(den, ed);
    double v = (double)std::sqrt(std::ldexp(mn / md, (int)(en - ed)));
    auto odd = [](double x) {
        int e;
        return (((uint64_t)std::ldexp(std::frexp(x, &e), 53)) & 1u) != 0;
    };
    for (;;) {
        const double lo = std::nextafter(v, 0.0);
        if (!(lo > 0.0)) break;
        const int c = cmp_square_of_mid(num, den, lo, v);
        if (c < 0 || (c == 0 && odd(v))) v = lo;
        else break;
    }
    for (;;) {
        const double hi = std::nextafter(v, HUGE_VAL);
        const int c = cmp_square_of_mid(num, den, v, hi);
        if (c > 0 || (c == 0 && odd(v))) v = hi;
        else break;
    }
    return v;
}

void mul_fact(Big& x, int n)  // x *= n!
{
    for (int i = 2; i <= n; ++i) x.mul_small((uint32_t)i);
}
void mul_range(Big& x, int lo, int hi)  // x *= (lo+1)(lo+2)...hi
{
    for (int i = lo + 1; i <= hi; ++i) x.mul_small((uint32_t)i);
}

// Exact <j1 m1; j2 m2 | J M> (all arguments doubled), rounded to double.
double clebsch_gordan(int j1, int m1, int j2, int m2, int J, int M)
{
    if (m1 + m2 != M || std::abs(m1) > j1 || std::abs(m2) > j2 || std::abs(M) > J) return 0.0;
    if (J > j1 + j2 || J < std::abs(j1 - j2) || ((j1 + j2 + J) & 1)) return 0.0;
    if (((j1 + m1) & 1) || ((j2 + m2) & 1) || ((J + M) & 1)) return 0.0;
    const int A = (j1 + j2 - J) / 2, B = (j1 - m1) / 2, Cc = (j2 + m2) / 2;
    const int D = (J - j2 + m1) / 2, Ee = (J - j1 - m2) / 2;
    const int kmin = std::max({0, -D, -Ee}), kmax = std::min({A, B, Cc});
    if (kmin > kmax) return 0.0;
    // S' = sum_k (-1)^k P / d_k with P = kmax! A! B! C! (D+kmax)! (E+kmax)!
    Big pos, neg;
    for (int k = kmin; k <= kmax; ++k) {
        Big t(1);
        mul_range(t, k, kmax);
        mul_range(t, A - k, A);
        mul_range(t, B - k, B);
        mul_range(t, Cc - k, Cc);
        mul_range(t, D + k, D + kmax);
        mul_range(t, Ee + k, Ee + kmax);
        if (k & 1) neg = add(neg, t);
        else pos = add(pos, t);
    }
    int sgn = cmp(pos, neg);
    if (sgn == 0) return 0.0;
    const Big S = sgn > 0 ? subtract(pos, neg) : subtract(neg, pos);
    Big P(1);
    mul_fact(P, kmax);
    mul_fact(P, A);
    mul_fact(P, B);
    mul_fact(P, Cc);
    mul_fact(P, D + kmax);
    mul_fact(P, Ee + kmax);
    // CG^2 = pref_num S^2 / (pref_den P^2)
    Big num(J + 1);
    mul_fact(num, (J + j1 - j2) / 2);
    mul_fact(num, (J - j1 + j2) / 2);
    mul_fact(num, (j1 + j2 - J) / 2);
    mul_fact(num, (J + M) / 2);
    mul_fact(num, (J - M) / 2);
    mul_fact(num, (j1 - m1) / 2);
    mul_fact(num, (j1 + m1) / 2);
    mul_fact(num, (j2 - m2) / 2);
    mul_fact(num, (j2 + m2) / 2);
    num = mul(num, mul(S, S));
    Big den(1);
    mul_fact(den, (j1 + j2 + J) / 2 + 1);
    den = mul(den, mul(P, P));
    const double v = sqrt_ratio_rounded(num, den);
    return sgn > 0 ? v : -v;
}

}  // namespace
}  // namespace hd

using namespace hd;

extern "C" {

int hd_sphere_ktable(int32_t max_lz, uint16_t* kt)
{
    if (!kt || max_lz < 0 || max_lz > 63) return fail(HD_ERR_ARG, "hd_sphere_ktable: bad argument");
    size_t it = 0;
    for (int m4 = max_lz; m4 >= -max_lz; m4 -= 2)
        for (int m3 = max_lz; m3 >= -max_lz; m3 -= 2)
            for (int m2 = max_lz; m2 >= -max_lz; m2 -= 2, ++it) {
                int m1 = m3 + m4 - m2;
                if (!(m1 <= max_lz && m1 >= -max_lz)) m1 = max_lz + 2;  // FindM1 error code
                kt[it] = (uint16_t)((-m1 + max_lz) / 2);                // ConvertStateLabel
            }
    return HD_OK;
}

int hd_sphere_vtable(int32_t max_lz, const double* vm, int32_t n_pseudo, double* vt)
{
    if (!vt || max_lz < 0 || max_lz > 63 || n_pseudo < 0 || (n_pseudo > 0 && !vm))
        return fail(HD_ERR_ARG, "hd_sphere_vtable: bad argument");
    if (n_pseudo > max_lz)
        return fail(HD_ERR_ARG, "TOO MANY 2-BODY PSEUDOPOTENTIAL COEFFICIENTS SPECIFIED");
    const int K = max_lz + 1;
    // CG(maxLz, m_a; maxLz, m_b | 2 maxLz - L, m_a + m_b) over (a, b, L/2 < K),
    // kept for the process per maxLz: a sweep over pseudopotentials reuses them
    static std::mutex cache_mutex;
    static std::map<int, std::vector<double>> caches;
    std::lock_guard<std::mutex> lock(cache_mutex);
    std::vector<double>& cache = caches[max_lz];
    if (cache.empty()) cache.assign((size_t)K * K * K, NAN);
    auto cg = [&](int ma, int mb, int Lh) {
        const int ia = (max_lz - ma) / 2, ib = (max_lz - mb) / 2;
        double& c = cache[((size_t)ia * K + ib) * K + Lh];
        if (std::isnan(c)) {
            const double v = clebsch_gordan(max_lz, ma, max_lz, mb, 2 * max_lz - 2 * Lh, ma + mb);
            c = std::fabs(v) > 1e-12 ? v : 0.0;
        }
        return c;
    };
    size_t it = 0;
    for (int m4 = max_lz; m4 >= -max_lz; m4 -= 2)
        for (int m3 = max_lz; m3 >= -max_lz; m3 -= 2)
            for (int m2 = max_lz; m2 >= -max_lz; m2 -= 2, ++it) {
                double total = 0.0;
                const int m1 = m3 + m4 - m2;
                if (m1 <= max_lz && m1 >= -max_lz) {
                    for (int L = 0; L < 2 * n_pseudo; L += 2) {
                        double temp = 0.0;
                        const int Lact = 2 * max_lz - L;
                        const int M = m1 + m2;
                        if (-Lact <= M && M <= Lact) temp += cg(m1, m2, L / 2) * cg(m3, m4, L / 2);
                        total -= vm[L / 2] * temp;
                    }
                }
                vt[it] = total;
            }
    return HD_OK;
}

}  // extern "C"
