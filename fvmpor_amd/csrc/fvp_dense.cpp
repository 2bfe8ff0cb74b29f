// Small dense linear algebra on the host for the deflated GMRES restarts
// (decomp's deflate, gmres.c:527-849, calls LAPACK dgesv, dgeevx and a QR for
// the same steps): LU solves with partial pivoting, least squares by
// Householder QR, eigenvalues of a general real matrix (Householder reduction
// to Hessenberg form, then Francis double-shift QR steps with deflation,
// Golub & Van Loan, Matrix Computations, algorithms 7.4.2 and 7.5.2),
// eigenvectors by inverse iteration, and modified Gram-Schmidt.  Matrices are
// row-major; sizes are a GMRES restart length (at most 64).
#include <algorithm>
#include <cmath>
#include <complex>
#include <limits>
#include <vector>

#include "fvp_dense.h"

namespace fvp {
namespace dense {

using cplx = std::complex<double>;

// x = A^-1 b (A n x n, destroyed); false for a singular matrix
bool lu_solve(int n, std::vector<double> A, std::vector<double> &b) {
    std::vector<int> piv(n);
    for (int k = 0; k < n; ++k) {
        int p = k;
        for (int i = k + 1; i < n; ++i)
            if (std::fabs(A[(size_t)i * n + k]) > std::fabs(A[(size_t)p * n + k])) p = i;
        if (A[(size_t)p * n + k] == 0.) return false;
        if (p != k) {
            for (int j = 0; j < n; ++j) std::swap(A[(size_t)p * n + j], A[(size_t)k * n + j]);
            std::swap(b[p], b[k]);
        }
        for (int i = k + 1; i < n; ++i) {
            const double l = A[(size_t)i * n + k] / A[(size_t)k * n + k];
            if (l == 0.) continue;
            for (int j = k + 1; j < n; ++j) A[(size_t)i * n + j] -= l * A[(size_t)k * n + j];
            b[i] -= l * b[k];
        }
    }
    for (int i = n - 1; i >= 0; --i) {
        double s = b[i];
        for (int j = i + 1; j < n; ++j) s -= A[(size_t)i * n + j] * b[j];
        b[i] = s / A[(size_t)i * n + i];
    }
    return true;
}

// y = argmin || c - A y || for A rows x cols (rows >= cols, full column rank)
bool lstsq(int rows, int cols, std::vector<double> A, std::vector<double> c, std::vector<double> &y) {
    for (int k = 0; k < cols; ++k) {
        double nx = 0.;
        for (int i = k; i < rows; ++i) nx += A[(size_t)i * cols + k] * A[(size_t)i * cols + k];
        nx = std::sqrt(nx);
        if (nx == 0.) return false;
        const double x0 = A[(size_t)k * cols + k];
        const double alpha = x0 > 0 ? -nx : nx;
        std::vector<double> v(rows - k);
        for (int i = k; i < rows; ++i) v[i - k] = A[(size_t)i * cols + k];
        v[0] -= alpha;
        double vv = 0.;
        for (double e : v) vv += e * e;
        if (vv > 0.) {
            for (int j = k; j < cols; ++j) {
                double s = 0.;
                for (int i = k; i < rows; ++i) s += v[i - k] * A[(size_t)i * cols + j];
                s *= 2. / vv;
                for (int i = k; i < rows; ++i) A[(size_t)i * cols + j] -= s * v[i - k];
            }
            double s = 0.;
            for (int i = k; i < rows; ++i) s += v[i - k] * c[i];
            s *= 2. / vv;
            for (int i = k; i < rows; ++i) c[i] -= s * v[i - k];
        }
    }
    y.assign(cols, 0.);
    for (int i = cols - 1; i >= 0; --i) {
        double s = c[i];
        for (int j = i + 1; j < cols; ++j) s -= A[(size_t)i * cols + j] * y[j];
        if (A[(size_t)i * cols + i] == 0.) return false;
        y[i] = s / A[(size_t)i * cols + i];
    }
    return true;
}

namespace {

// reflector I - tau v v^T, v = (1, v1, v2), mapping (x, y, z) onto the first axis
void house3(double x, double y, double z, double &v1, double &v2, double &tau) {
    const double nx = std::sqrt(x * x + y * y + z * z);
    if (nx == 0.) { v1 = v2 = tau = 0.; return; }
    const double alpha = x > 0 ? -nx : nx;
    const double d = x - alpha;
    v1 = y / d;
    v2 = z / d;
    tau = 2. / (1. + v1 * v1 + v2 * v2);
}

}  // namespace

// eigenvalues of a general real n x n matrix (destroyed): reduction to upper
// Hessenberg form by Householder reflectors, then Francis double-shift QR
// steps on the active window, deflating negligible subdiagonal entries; 1 x 1
// and 2 x 2 blocks give the eigenvalues.  False when the iteration stalls.
bool eig(int n, std::vector<double> H, std::vector<double> &wr, std::vector<double> &wi) {
    auto h = [&](int i, int j) -> double & { return H[(size_t)i * n + j]; };
    for (int k = 0; k + 2 < n; ++k) {
        double nx = 0.;
        for (int i = k + 1; i < n; ++i) nx += h(i, k) * h(i, k);
        nx = std::sqrt(nx);
        if (nx == 0.) continue;
        std::vector<double> v(n - k - 1);
        for (int i = k + 1; i < n; ++i) v[i - k - 1] = h(i, k);
        const double alpha = v[0] > 0 ? -nx : nx;
        v[0] -= alpha;
        double vv = 0.;
        for (double e : v) vv += e * e;
        if (vv == 0.) continue;
        for (int j = 0; j < n; ++j) {
            double s = 0.;
            for (int i = k + 1; i < n; ++i) s += v[i - k - 1] * h(i, j);
            s *= 2. / vv;
            for (int i = k + 1; i < n; ++i) h(i, j) -= s * v[i - k - 1];
        }
        for (int i = 0; i < n; ++i) {
            double s = 0.;
            for (int j = k + 1; j < n; ++j) s += h(i, j) * v[j - k - 1];
            s *= 2. / vv;
            for (int j = k + 1; j < n; ++j) h(i, j) -= s * v[j - k - 1];
        }
        for (int i = k + 2; i < n; ++i) h(i, k) = 0.;
    }
    wr.assign(n, 0.);
    wi.assign(n, 0.);
    double anorm = 0.;
    for (int i = 0; i < n; ++i)
        for (int j = std::max(0, i - 1); j < n; ++j) anorm += std::fabs(h(i, j));
    const double eps = std::numeric_limits<double>::epsilon();
    int hi = n - 1, iter = 0, total = 0;
    while (hi >= 0) {
        int l = hi;
        for (; l > 0; --l) {
            double s = std::fabs(h(l - 1, l - 1)) + std::fabs(h(l, l));
            if (s == 0.) s = anorm;
            if (std::fabs(h(l, l - 1)) <= eps * s) { h(l, l - 1) = 0.; break; }
        }
        if (l == hi) {
            wr[hi] = h(hi, hi);
            wi[hi] = 0.;
            --hi;
            iter = 0;
            continue;
        }
        if (l == hi - 1) {
            const double a = h(hi - 1, hi - 1), b = h(hi - 1, hi), c = h(hi, hi - 1), d = h(hi, hi);
            const double p = 0.5 * (a - d), q = p * p + b * c;
            if (q >= 0.) {
                const double z = p + std::copysign(std::sqrt(q), p);
                wr[hi - 1] = d + z;
                wr[hi] = z != 0. ? d - b * c / z : d + z;
            } else {
                wr[hi - 1] = wr[hi] = d + p;
                wi[hi - 1] = std::sqrt(-q);
                wi[hi] = -std::sqrt(-q);
            }
            hi -= 2;
            iter = 0;
            continue;
        }
        if (++total > 100 * n) return false;
        ++iter;
        double s, t;
        if (iter % 10 == 0) {  // exceptional shift after a run of slow steps
            const double e = std::fabs(h(hi, hi - 1)) + std::fabs(h(hi - 1, hi - 2));
            s = 1.5 * e;
            t = e * e;
        } else {
            s = h(hi - 1, hi - 1) + h(hi, hi);
            t = h(hi - 1, hi - 1) * h(hi, hi) - h(hi - 1, hi) * h(hi, hi - 1);
        }
        double x = h(l, l) * h(l, l) + h(l, l + 1) * h(l + 1, l) - s * h(l, l) + t;
        double y = h(l + 1, l) * (h(l, l) + h(l + 1, l + 1) - s);
        double z = h(l + 1, l) * h(l + 2, l + 1);
        for (int k = l; k <= hi - 2; ++k) {
            double v1, v2, tau;
            house3(x, y, z, v1, v2, tau);
            if (tau != 0.) {
                for (int j = std::max(l, k - 1); j <= hi; ++j) {
                    const double w = tau * (h(k, j) + v1 * h(k + 1, j) + v2 * h(k + 2, j));
                    h(k, j) -= w;
                    h(k + 1, j) -= w * v1;
                    h(k + 2, j) -= w * v2;
                }
                for (int i = l; i <= std::min(k + 3, hi); ++i) {
                    const double w = tau * (h(i, k) + v1 * h(i, k + 1) + v2 * h(i, k + 2));
                    h(i, k) -= w;
                    h(i, k + 1) -= w * v1;
                    h(i, k + 2) -= w * v2;
                }
            }
            x = h(k + 1, k);
            y = h(k + 2, k);
            if (k < hi - 2) z = h(k + 3, k);
        }
        double v1, v2, tau;
        house3(x, y, 0., v1, v2, tau);
        if (tau != 0.) {
            for (int j = std::max(l, hi - 2); j <= hi; ++j) {
                const double w = tau * (h(hi - 1, j) + v1 * h(hi, j));
                h(hi - 1, j) -= w;
                h(hi, j) -= w * v1;
            }
            for (int i = l; i <= hi; ++i) {
                const double w = tau * (h(i, hi - 1) + v1 * h(i, hi));
                h(i, hi - 1) -= w;
                h(i, hi) -= w * v1;
            }
        }
    }
    return true;
}

// eigenvector of A (n x n) for the eigenvalue lam by inverse iteration with
// a complex LU of (A - mu I), mu = lam moved off the spectrum by a relative
// 1e-10; unit 2-norm
bool eigvec(int n, const std::vector<double> &A, cplx lam, std::vector<cplx> &v) {
    double scale = 0.;
    for (double a : A) scale = std::max(scale, std::fabs(a));
    const cplx mu = lam + cplx(1e-10 * std::max(std::abs(lam), scale * 1e-3), 0.);
    std::vector<cplx> M((size_t)n * n);
    for (int i = 0; i < n; ++i)
        for (int j = 0; j < n; ++j) M[(size_t)i * n + j] = cplx(A[(size_t)i * n + j], 0.) - (i == j ? mu : cplx(0.));
    std::vector<int> piv(n);
    for (int k = 0; k < n; ++k) {
        int p = k;
        for (int i = k + 1; i < n; ++i)
            if (std::abs(M[(size_t)i * n + k]) > std::abs(M[(size_t)p * n + k])) p = i;
        piv[k] = p;
        if (p != k)
            for (int j = 0; j < n; ++j) std::swap(M[(size_t)p * n + j], M[(size_t)k * n + j]);
        if (std::abs(M[(size_t)k * n + k]) == 0.) M[(size_t)k * n + k] = cplx(1e-300, 0.);
        for (int i = k + 1; i < n; ++i) {
            const cplx l = M[(size_t)i * n + k] / M[(size_t)k * n + k];
            M[(size_t)i * n + k] = l;
            for (int j = k + 1; j < n; ++j) M[(size_t)i * n + j] -= l * M[(size_t)k * n + j];
        }
    }
    v.assign(n, cplx(1., 0.));
    for (int i = 0; i < n; ++i) v[i] = cplx(1. + 0.1 * std::sin(1.7 * i), 0.05 * std::cos(2.3 * i));
    for (int it = 0; it < 4; ++it) {
        // P v (the interchanges in order), then the unit lower and the upper triangle
        for (int k = 0; k < n; ++k)
            if (piv[k] != k) std::swap(v[k], v[piv[k]]);
        for (int i = 1; i < n; ++i)
            for (int k = 0; k < i; ++k) v[i] -= M[(size_t)i * n + k] * v[k];
        for (int i = n - 1; i >= 0; --i) {
            cplx s = v[i];
            for (int j = i + 1; j < n; ++j) s -= M[(size_t)i * n + j] * v[j];
            v[i] = s / M[(size_t)i * n + i];
        }
        double nv = 0.;
        for (const cplx &e : v) nv += std::norm(e);
        nv = std::sqrt(nv);
        if (!(nv > 0.) || !std::isfinite(nv)) return false;
        for (cplx &e : v) e /= nv;
    }
    return true;
}

// orthonormal columns (two passes of modified Gram-Schmidt) of X (rows x cols,
// column-major); false when a column is dependent on the previous ones
bool orthonormalise(int rows, int cols, std::vector<double> &X) {
    for (int c = 0; c < cols; ++c) {
        double *x = &X[(size_t)c * rows];
        for (int pass = 0; pass < 2; ++pass)
            for (int q = 0; q < c; ++q) {
                const double *y = &X[(size_t)q * rows];
                double s = 0.;
                for (int i = 0; i < rows; ++i) s += x[i] * y[i];
                for (int i = 0; i < rows; ++i) x[i] -= s * y[i];
            }
        double nx = 0.;
        for (int i = 0; i < rows; ++i) nx += x[i] * x[i];
        nx = std::sqrt(nx);
        if (!(nx > 1e-13)) return false;
        for (int i = 0; i < rows; ++i) x[i] /= nx;
    }
    return true;
}

}  // namespace dense
}  // namespace fvp
