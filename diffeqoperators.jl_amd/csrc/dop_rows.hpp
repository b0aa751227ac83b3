// dop_rows.hpp -- host-side construction of the per-axis row programs from the
// operators' own FP64 tables (the arithmetic of DiffEqOperators.jl's convolution kernels
// expressed as data: one (start, weights) program per output row).
#pragma once
#include <string>
#include <vector>

#include "../../include/dop.h"
#include "dop_plan.hpp"

namespace dop {

// One output row: du[r] = sum_k w[k] * P(J0 + k), P the padded line (0-based J).
struct RowProgram {
    int J0 = 0;
    std::vector<double> w;
};

// Interior rows [lo, hi) (0-based) of the reference's convolve_interior! loop for A.
void interior_rows(const dop_operator& A, int sem, int* lo, int* hi);

// Row program of row r (0-based) of operator A.  Returns false (and sets *err) if the
// reference would index outside the padded vector.
bool row_program(const dop_operator& A, int sem, int r, RowProgram* out, std::string* err);

// Checks that an operator's scalar fields are consistent with its constructor
// (derivative_operator.jl:85-92, 461-463) and that its tables are present.
bool validate_operator(const dop_operator& A, std::string* err);

struct AxisTables {
    AxisDev meta{};
    std::vector<double> w64;
    std::vector<float> w32;
    std::vector<int32_t> bJ0, bn;
    std::vector<double> bw;
    std::vector<double> lbc;   // per-line BCs, kLineBC per line (empty: one BC per axis)
};

// Merges all operators acting on one axis (sum in the given order) into row programs and
// splits them into regular and boundary rows.
bool build_axis(const std::vector<const dop_operator*>& ops, int sem, int n, AxisTables* out,
                std::string* err);

}  // namespace dop
