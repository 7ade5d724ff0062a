// host.h — host-side data model of the backend: gate matrices, circuit validation,
// Kraus operators and operator strings. Pure C++; no device code.
#pragma once

#include <cstdint>
#include <string>
#include <vector>

#include "../../include/yaohip.h"

namespace yh {

struct cd {
  double re, im;
};
inline cd operator*(cd a, cd b) { return cd{a.re * b.re - a.im * b.im, a.re * b.im + a.im * b.re}; }
inline cd operator+(cd a, cd b) { return cd{a.re + b.re, a.im + b.im}; }
inline cd cconj(cd a) { return cd{a.re, -a.im}; }

using Mat = std::vector<cd>;  // row-major square matrix

// Gate::matrix (gate.rs:63-68, 284-393); returns the dimension.
int gate_matrix(const yh_op& op, Mat& m);
// Gate::is_diagonal (gate.rs:98-104).
bool gate_is_diagonal(const yh_op& op);
// Circuit::new rules for a qubit circuit (circuit.rs:150-272); returns a status.
int validate_circuit(int nbits, const yh_op* ops, size_t nops);
// NoiseChannel::kraus_operators (noise.rs:183-352) with parameter validation.
int kraus_operators(const yh_op& op, std::vector<Mat>& ops, int& dim);
int channel_num_qubits(const yh_op& op);
// is_diagonal_matrix (density_matrix.rs:125-129).
bool is_diagonal_matrix(const Mat& m, int d);

// Eigenvalues (ascending) of a Hermitian d x d matrix, as hermitian_eigenvalues
// (density_matrix.rs:227-242) computes them with faer; host-side, for reduced
// density matrices.
std::vector<double> hermitian_eigenvalues(const Mat& m, int d);

// One Pauli-product term, OperatorString::new applied (operator.rs:27-33).
struct Term {
  cd coeff;
  std::vector<std::pair<int, int>> ops;  // (site, op) sorted by site, I removed
};
std::vector<Term> read_terms(size_t nterms, const double* coeffs, const int32_t* term_len,
                             const int32_t* sites, const int32_t* ops);

// For single-site operator `op` (operator.rs:53-65): every row of its 2x2 matrix
// has at most one non-zero entry. For output bit v, src[v] is the input bit that
// feeds it and fac[v] its factor (0 when the row is empty).
void op_row_table(int op, int src[2], cd fac[2]);

}  // namespace yh
