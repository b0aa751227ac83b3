/*
 * dop_oracle.h -- FP64 CPU restatement of DiffEqOperators.jl's operator-application
 * path (the checker, not the product).
 *
 * TEST INFRASTRUCTURE ONLY.  Only tests/, __graft_entry__.smoke() and bench.py's
 * cpu_baseline leg may load this library, and only as the checker / the timed CPU
 * baseline.  The product library (libdop.so) never links or calls it.
 *
 * Every function restates one piece of the reference (cited file:line, paths relative
 * to the DiffEqOperators.jl source tree).  Arrays are column-major (Julia order):
 * axis 0 is the fastest-varying one.  All arithmetic is IEEE binary64 with no FMA
 * contraction (built with -ffp-contract=off), so table generation reproduces the
 * reference's operation order.
 *
 * What pins it (tests/test_oracle_golden.py): the analytic matrices and values of the
 * reference's own tests (tests/golden/reference_fixtures.json), closed forms, invariants
 * (weights sum to 0, polynomial exactness, reflected boundaries, max-offside upwind ==
 * centered) and a planted mistake that must fail.  Parity unpinned (restated from the
 * source only; no test, fixture or closed form of the reference fixes them):
 *   - orc_mul_line_bpv for centered operators with non-constant coefficient vectors
 *     (the coefficient index of convolutions.jl:384);
 *   - orc_composite_mul_padded's transverse views 1:n on non-separable fields
 *     (derivative_operator_functions.jl:429-460, 780-800);
 *   - orc_bc.periodic == 2, the N-D PeriodicBC padding with the first / last planes
 *     (multi_dim_bc_operators.jl:221-252).
 */
#ifndef DOP_ORACLE_H
#define DOP_ORACLE_H

#ifdef __cplusplus
extern "C" {
#endif

#define ORC_OK 0
#define ORC_EINVAL (-1)
#define ORC_ENOMEM (-4)

/* coefficient-function kind of an operator: the reference's non-uniform upwind
 * interior kernel branches on `A.coeff_func isa Number && A.coeff_func >= 0`
 * (convolutions.jl:234-253) before it looks at per-row coefficient signs. */
#define ORC_COEFF_VECTOR 0   /* Vector or Function coeff_func: per-row sign */
#define ORC_COEFF_NUM_POS 1  /* coeff_func::Number >= 0 */
#define ORC_COEFF_NUM_NEG 2  /* coeff_func::Number <  0 */

/* One DerivativeOperator (derivative_operator.jl:14-29).  Table layouts (row-major):
 *   centered uniform:      stencil[SL], low[BPC][BSL], high[BPC][BSL]
 *   centered non-uniform:  stencil[len-2BPC][SL], low[BPC][BSL], high[BPC][BSL]
 *   upwind uniform:        stencil[SL], low[BPC][BSL], high[BPC+offside][BSL]
 *   upwind non-uniform:    stencil[2][len-2BPC][SL], low[2][BPC][BSL],
 *                          high[2][BPC+offside][BSL]   (index 0 = upwind, 1 = downwind)
 * coeff[len] is the operator's `coefficients` vector. */
typedef struct orc_op {
    int axis;       /* 0-based axis of differentiation (Julia's N - 1) */
    int wind;       /* 0 = CenteredDifference (Wind=false), 1 = UpwindDifference */
    int d, a;       /* derivative_order, approximation_order */
    int len;
    int SL, BSL, BPC, offside;
    int uniform;    /* dx isa Number */
    int coeff_kind; /* ORC_COEFF_* */
    const double* stencil;
    const double* low;
    const double* high;
    const double* coeff;
} orc_op;

/* One boundary condition along an axis.  Affine (bc_operators.jl:21-39): the ghosts are
 * l = a_l . u[1:nl] + b_l and r = a_r . u[end-nr+1:end] + b_r.  periodic = 1 is the atomic
 * PeriodicBC on a line, l = u[end], r = u[1] (bc_operators.jl:192); periodic = 2 is the
 * padding of a MultiDimDirectionalBC / ComposedMultiDimBC whose BCs are all PeriodicBC,
 * whose ghost planes are selectdim(u, D, 1) and selectdim(u, D, size(u, D)), i.e.
 * l = u[1], r = u[end] (multi_dim_bc_operators.jl:221-252).  The taps are ignored then. */
typedef struct orc_bc {
    int nl, nr;
    const double* a_l;
    double b_l;
    const double* a_r;
    double b_r;
    int periodic;
    /* N-D: one affine BC per line along the axis (MultiDimBC{dim} of an Array of atomic
     * BCs, multi_dim_bc_operators.jl:93-95; slice_rmul applies BC[I] to the line u[I, :, J],
     * :34), lines in column-major order over the other axes; NULL: this BC on every line */
    const struct orc_bc* lines;
} orc_bc;

/* fornberg.jl:7-62 */
int orc_calculate_weights(int order, double x0, const double* x, int n, double* out);

/* sizes of the tables of CenteredDifference / UpwindDifference */
int orc_centered_sizes(int d, int a, int* SL, int* BSL, int* BPC);
int orc_upwind_sizes(int d, int a, int offside, int* SL, int* BSL, int* BPC);

/* derivative_operator.jl:81-135 */
int orc_centered_uniform(int d, int a, double dx, int len,
                         double* stencil, double* low, double* high);
/* derivative_operator.jl:137-205 ; dx has len+1 entries */
int orc_centered_nonuniform(int d, int a, const double* dx, int len,
                            double* stencil, double* low, double* high);
/* derivative_operator.jl:454-512 */
int orc_upwind_uniform(int d, int a, double dx, int len, int offside,
                       double* stencil, double* low, double* high);
/* derivative_operator.jl:515-639 ; dx has len+1 entries */
int orc_upwind_nonuniform(int d, int a, const double* dx, int len, int offside,
                          double* stencil, double* low, double* high);

/* bc_operators.jl:26-39 ; a_l, a_r have `order` entries */
int orc_robin_uniform(double al, double bl, double cl, double ar, double br, double cr,
                      double dx, int order, double* a_l, double* b_l, double* a_r,
                      double* b_r);
/* bc_operators.jl:40-60 ; dx has ndx >= order+1 entries */
/* bc_operators.jl:85-170 -- GeneralBC: alpha_l[0] + alpha_l[1] u + sum_i alpha_l[i+1] u^(i) = 0
 * at each end (nl, nr coefficients); ndx == 0: uniform step dx[0], else dx[ndx].
 * a_l gets order + nl - 3 taps, a_r order + nr - 3. */
int orc_general(const double* alpha_l, int nl, const double* alpha_r, int nr, const double* dx,
                int ndx, int order, double* a_l, double* b_l, double* a_r, double* b_r);

int orc_robin_nonuniform(double al, double bl, double cl, double ar, double br, double cr,
                         const double* dx, int ndx, int order, double* a_l, double* b_l,
                         double* a_r, double* b_r);

/* 1-D mul! against a plain (already padded) vector x[len+2]: convolutions.jl:17-22 with
 * the AbstractVector kernels (:27-118 centered, :123-215 upwind uniform, :221-362
 * upwind non-uniform).  overwrite=0 adds into out (the `!overwrite * x_temp[i]` term). */
int orc_mul_line(const orc_op* op, const double* x, double* out, int overwrite);

/* 1-D mul! against a BoundaryPaddedVector(l, r, u[len]): convolutions.jl:367-730. */
int orc_mul_line_bpv(const orc_op* op, double l, double r, const double* u, double* out,
                     int overwrite);

/* N-D line driver _mul! (derivative_operator_functions.jl:27-69): out has size
 * out_size[ndim]; M has size M_size[ndim] with M_size[axis] == out_size[axis]+2 and every
 * other extent either equal or padded by 2 (then trimmed to 2:end-1). */
int orc_mul_nd(const orc_op* op, int ndim, const long long* out_size, double* out,
               const long long* M_size, const double* M, int overwrite);

/* GhostDerivativeOperator sum: out = sum_t (L_t * Q) * u (ghost_derivative_operator.jl:
 * 11-37, composite_operators.jl:64-65).  bcs[ndim] are the per-axis affine BCs of the
 * composed Q (multi_dim_bc_operators.jl:141-149, 178-192, 230-242).  ndim == 1 takes the
 * BoundaryPaddedVector kernels; ndim > 1 takes the AbstractVector kernels on the lines of
 * the ComposedBoundaryPaddedArray (boundary_padded_arrays.jl:158-180).  The per-op
 * results are summed left to right.  bpv_1d = 0 forces the line kernels in 1-D too. */
int orc_ghost_apply(int nops, const orc_op* ops, int ndim, const long long* size,
                    const orc_bc* bcs, const double* u, double* out, int bpv_1d);

/* mul!(x_temp, A::DiffEqOperatorCombination, M) for 2-D / 3-D padded M, the branch the
 * public constructors reach (derivative_operator_functions.jl:414-461, 768-801),
 * including its transverse views 1:n.  M has size out_size .+ 2. */
int orc_composite_mul_padded(int nops, const orc_op* ops, int ndim,
                             const long long* out_size, const double* M, double* out);

/* OpenMP threads used by the line drivers (1 = scalar, like the reference). */
void orc_set_threads(int n);
int orc_get_threads(void);

#ifdef __cplusplus
}
#endif
#endif
