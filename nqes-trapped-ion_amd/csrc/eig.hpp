// eig.hpp -- eigen-decomposition of a small general complex matrix (host only).
//
// Replaces the two Eigen calls of the reference's loop: e_mat.eigenvalues()
// (rbm.cc:1186) and ComplexEigenSolver (rbm.cc:1198-1205), for K x K matrices.
// Method: Householder reduction to Hessenberg form, complex single-shift QR with
// Wilkinson shifts and deflation (Golub & Van Loan, Matrix Computations, 7.5), then
// eigenvectors by inverse iteration with the original matrix.
#pragma once

#include <complex>

namespace eigk {

using cplx = std::complex<double>;

// A: n x n row-major (n <= 16).  evals[n]; V (optional): n x n row-major, column i
// is a unit eigenvector for evals[i].  Returns false if QR did not converge.
bool eig(int n, const cplx *A, cplx *evals, cplx *V);

// Moore-Penrose pseudo-inverse of the n x n row-major A (n <= 16), the role of
// Eigen's completeOrthogonalDecomposition().pseudoInverse() at rbm.cc:547, 608, 1158:
// one-sided Jacobi SVD A = U S V^H (Hestenes), singular values below rtol * s_max
// treated as zero.  X: n x n row-major.  Returns the numerical rank.
int pinv(int n, const cplx *A, cplx *X, double rtol = 1e-12);

// determinant of the n x n row-major A (LU with partial pivoting)
cplx det(int n, const cplx *A);

} // namespace eigk
