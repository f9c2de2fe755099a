// hd_eigsym.cpp -- dense symmetric eigensolver for the projected matrices.
//
// Householder reduction to tridiagonal form followed by the implicit QL
// iteration with Wilkinson-type shifts (Wilkinson and Reinsch, Handbook for
// Automatic Computation II, 1971), eigenvectors accumulated throughout.
// O(n^3) with small constants: the block Lanczos projects onto up to 160
// columns, where cyclic Jacobi would cost tens of milliseconds per restart.
#include <algorithm>
#include <cmath>
#include <numeric>
#include <vector>

#include "hd_internal.hpp"

namespace hd {

// A (n x n, row-major, symmetric) is destroyed; w ascending, eigenvector i is
// column i of Z (row-major).  Returns false if the QL iteration does not
// converge (not observed for symmetric input).
bool sym_eig(int n, std::vector<double>& A, std::vector<double>& w, std::vector<double>& Z)
{
    auto a = [&](int i, int j) -> double& { return A[(size_t)i * n + j]; };
    Z.assign((size_t)n * n, 0.0);
    for (int i = 0; i < n; ++i) Z[(size_t)i * n + i] = 1.0;
    std::vector<double> v(n), p(n), q(n);
    // --- Householder: column k below the subdiagonal is reflected onto e_{k+1}
    for (int k = 0; k + 2 < n; ++k) {
        double norm = 0.0;
        for (int i = k + 1; i < n; ++i) norm += a(i, k) * a(i, k);
        norm = std::sqrt(norm);
        if (norm == 0.0) continue;
        const double x0 = a(k + 1, k);
        const double alpha = x0 > 0 ? -norm : norm;
        // v = x - alpha e1, normalised
        double vn = 0.0;
        for (int i = k + 1; i < n; ++i) {
            v[i] = a(i, k) - (i == k + 1 ? alpha : 0.0);
            vn += v[i] * v[i];
        }
        vn = std::sqrt(vn);
        if (vn == 0.0) continue;
        for (int i = k + 1; i < n; ++i) v[i] /= vn;
        // p = A22 v, K = v^T p, q = p - K v; A22 -= 2 (v q^T + q v^T)
        double K = 0.0;
        for (int i = k + 1; i < n; ++i) {
            double s = 0.0;
            for (int j = k + 1; j < n; ++j) s += a(i, j) * v[j];
            p[i] = s;
            K += v[i] * s;
        }
        for (int i = k + 1; i < n; ++i) q[i] = p[i] - K * v[i];
        for (int i = k + 1; i < n; ++i)
            for (int j = k + 1; j < n; ++j) a(i, j) -= 2.0 * (v[i] * q[j] + q[i] * v[j]);
        a(k + 1, k) = alpha;
        a(k, k + 1) = alpha;
        for (int i = k + 2; i < n; ++i) a(i, k) = a(k, i) = 0.0;
        // Z <- Z (I - 2 v v^T) on columns k+1..n-1
        for (int r = 0; r < n; ++r) {
            double s = 0.0;
            for (int j = k + 1; j < n; ++j) s += Z[(size_t)r * n + j] * v[j];
            s *= 2.0;
            for (int j = k + 1; j < n; ++j) Z[(size_t)r * n + j] -= s * v[j];
        }
    }
    std::vector<double> d(n), e(n, 0.0);
    for (int i = 0; i < n; ++i) d[i] = a(i, i);
    for (int i = 0; i + 1 < n; ++i) e[i] = a(i + 1, i);
    // --- implicit QL on (d, e); rotations applied to the columns of Z
    const double eps = 2.220446049250313e-16;
    for (int l = 0; l < n; ++l) {
        for (int iter = 0;; ++iter) {
            int m = l;
            for (; m + 1 < n; ++m) {
                const double dd = std::fabs(d[m]) + std::fabs(d[m + 1]);
                if (std::fabs(e[m]) <= eps * dd || std::fabs(e[m]) < 1e-300) break;
            }
            if (m == l) break;
            if (iter == 80) return false;
            // shift from the leading 2x2 block (the eigenvalue of it closer to d[l])
            double g = (d[l + 1] - d[l]) / (2.0 * e[l]);
            double r = std::hypot(g, 1.0);
            g = d[m] - d[l] + e[l] / (g + (g >= 0 ? r : -r));
            double s = 1.0, c = 1.0, pshift = 0.0;
            bool deflated = false;
            for (int i = m - 1; i >= l; --i) {
                const double f = s * e[i];
                const double bb = c * e[i];
                r = std::hypot(f, g);
                e[i + 1] = r;
                if (r == 0.0) {  // split: d[i+1] is converged
                    d[i + 1] -= pshift;
                    e[m] = 0.0;
                    deflated = true;
                    break;
                }
                s = f / r;
                c = g / r;
                g = d[i + 1] - pshift;
                r = (d[i] - g) * s + 2.0 * c * bb;
                pshift = s * r;
                d[i + 1] = g + pshift;
                g = c * r - bb;
                for (int k = 0; k < n; ++k) {
                    double& zi = Z[(size_t)k * n + i];
                    double& zi1 = Z[(size_t)k * n + i + 1];
                    const double t = zi1;
                    zi1 = s * zi + c * t;
                    zi = c * zi - s * t;
                }
            }
            if (deflated) continue;
            d[l] -= pshift;
            e[l] = g;
            e[m] = 0.0;
        }
    }
    std::vector<int> idx(n);
    std::iota(idx.begin(), idx.end(), 0);
    std::sort(idx.begin(), idx.end(), [&](int x, int y) { return d[x] < d[y]; });
    w.resize(n);
    std::vector<double> Zs((size_t)n * n);
    for (int c = 0; c < n; ++c) {
        w[c] = d[idx[c]];
        for (int r = 0; r < n; ++r) Zs[(size_t)r * n + c] = Z[(size_t)r * n + idx[c]];
    }
    Z.swap(Zs);
    return true;
}

}  // namespace hd

using namespace hd;

extern "C" int hd_sym_eig(int32_t n, const double* h_a, double* h_w, double* h_z)
{
    if (n <= 0 || !h_a || !h_w) return fail(HD_ERR_ARG, "hd_sym_eig: bad argument");
    std::vector<double> A(h_a, h_a + (size_t)n * n), w, Z;
    if (!sym_eig(n, A, w, Z)) return fail(HD_ERR_NOT_CONVERGED, "QL iteration did not converge");
    std::copy(w.begin(), w.end(), h_w);
    if (h_z) std::copy(Z.begin(), Z.end(), h_z);
    return HD_OK;
}
