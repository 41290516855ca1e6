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

} // namespace eigk
