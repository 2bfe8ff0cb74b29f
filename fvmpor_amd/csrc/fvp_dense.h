// Small dense linear algebra on the host (see fvp_dense.cpp).  Row-major
// matrices unless stated.
#pragma once

#include <complex>
#include <vector>

namespace fvp {
namespace dense {

// b := A^-1 b (partial pivoting); false for a singular matrix
bool lu_solve(int n, std::vector<double> A, std::vector<double> &b);
// y = argmin || c - A y ||, A rows x cols, full column rank
bool lstsq(int rows, int cols, std::vector<double> A, std::vector<double> c, std::vector<double> &y);
// eigenvalues (wr + i wi) of a general real n x n matrix
bool eig(int n, std::vector<double> H, std::vector<double> &wr, std::vector<double> &wi);
// unit eigenvector of A for the eigenvalue lam (inverse iteration)
bool eigvec(int n, const std::vector<double> &A, std::complex<double> lam, std::vector<std::complex<double>> &v);
// orthonormal columns of X (rows x cols, column-major), modified Gram-Schmidt twice
bool orthonormalise(int rows, int cols, std::vector<double> &X);

}  // namespace dense
}  // namespace fvp
