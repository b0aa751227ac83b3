/*
 * dop_oracle.c -- FP64 CPU restatement of DiffEqOperators.jl's operator-application path.
 *
 * TEST INFRASTRUCTURE ONLY (the checker and the timed CPU baseline).  The product
 * library never links or calls this file.
 *
 * Each function follows the reference file:line it cites, keeping the reference's loop
 * bounds, branch conditions, index arithmetic and operation order.  Indices inside the
 * kernels are kept 1-based (the P(j) / U(j) macros) so that each bound reads exactly as
 * in the reference.  Built with -ffp-contract=off: no FMA is formed from a*b+c.
 */
#include "dop_oracle.h"

#include <math.h>
#include <stdlib.h>
#include <string.h>
#ifdef _OPENMP
#include <omp.h>
#endif

static int g_threads = 1;
void orc_set_threads(int n) { g_threads = n < 1 ? 1 : n; }
int orc_get_threads(void) { return g_threads; }

/* ------------------------------------------------------------------------------------ */
/* fornberg.jl:7-62 -- Fornberg (1988) weights, then the sum-to-zero correction at the
 * middle entry (:57-60).  C is N x (M+1), rows = points, cols = derivative order. */
int orc_calculate_weights(int order, double x0, const double* x, int N, double* out)
{
    if (order < 0 || N <= order) return ORC_EINVAL;
    const int M = order;
    double* C = (double*)calloc((size_t)N * (size_t)(M + 1), sizeof(double));
    if (!C) return ORC_ENOMEM;
#define CC(r, c) C[(size_t)(r) * (size_t)(M + 1) + (size_t)(c)]
    double c1 = 1.0;
    double c4 = x[0] - x0;
    CC(0, 0) = 1.0;
    for (int i = 1; i <= N - 1; ++i) {   /* reference i = 1:N-1, i1 = i+1 -> row i */
        int mn = i < M ? i : M;
        double c2 = 1.0;
        double c5 = c4;
        c4 = x[i] - x0;
        for (int j = 0; j <= i - 1; ++j) {   /* reference j = 0:i-1, j1 = j+1 -> row j */
            double c3 = x[i] - x[j];
            c2 *= c3;
            if (j == i - 1) {
                for (int s = mn; s >= 1; --s)
                    CC(i, s) = c1 * ((double)s * CC(i - 1, s - 1) - c5 * CC(i - 1, s)) / c2;
                CC(i, 0) = -c1 * c5 * CC(i - 1, 0) / c2;
            }
            for (int s = mn; s >= 1; --s)
                CC(j, s) = (c4 * CC(j, s) - (double)s * CC(j, s - 1)) / c3;
            CC(j, 0) = c4 * CC(j, 0) / c3;
        }
        c1 = c2;
    }
    for (int r = 0; r < N; ++r) out[r] = CC(r, M);
#undef CC
    free(C);
    if (order != 0) {
        double s = 0.0;   /* Base.sum of < 16 entries: sequential left-to-right */
        for (int r = 0; r < N; ++r) s += out[r];
        out[N / 2] -= s;
    }
    return ORC_OK;
}

/* ------------------------------------------------------------------------------------ */
/* derivative_operator.jl:85-92 (centered) and :461-463 (upwind) */
int orc_centered_sizes(int d, int a, int* SL, int* BSL, int* BPC)
{
    if (d < 0 || a < 1) return ORC_EINVAL;
    *SL = d + a - 1 + (d + a) % 2;
    *BSL = d + a;
    *BPC = *SL / 2 - 1;
    return ORC_OK;
}

int orc_upwind_sizes(int d, int a, int offside, int* SL, int* BSL, int* BPC)
{
    if (d < 0 || a < 1) return ORC_EINVAL;
    if (offside < 0 || offside > (d + a - 1) / 2) return ORC_EINVAL;   /* :458-459 */
    *SL = d + a;
    *BSL = d + a;
    *BPC = *BSL - 2 - offside;
    return ORC_OK;
}

/* derivative_operator.jl:81-135 -- uniform CenteredDifference */
int orc_centered_uniform(int d, int a, double dx, int len, double* stencil, double* low,
                         double* high)
{
    int SL, BSL, BPC;
    if (a <= 1) return ORC_EINVAL;   /* :84 */
    if (orc_centered_sizes(d, a, &SL, &BSL, &BPC) != ORC_OK || len < 1) return ORC_EINVAL;
    double xs[64], w[64];
    if (SL > 64 || BSL > 64) return ORC_EINVAL;
    const double scale = 1.0 / pow(dx, (double)d);   /* (1 / dx^derivative_order) */
    for (int k = 0; k < SL; ++k) xs[k] = (double)(k - SL / 2);   /* -SL÷2 : SL÷2 */
    if (orc_calculate_weights(d, 0.0, xs, SL, w) != ORC_OK) return ORC_EINVAL;
    for (int k = 0; k < SL; ++k) stencil[k] = scale * w[k];
    for (int k = 0; k < BSL; ++k) xs[k] = (double)k;   /* left_boundary_x = 0:BSL-1 */
    for (int i = 1; i <= BPC; ++i) {   /* L_boundary_deriv_spots = 1 : SL÷2-1 */
        if (orc_calculate_weights(d, (double)i, xs, BSL, w) != ORC_OK) return ORC_EINVAL;
        for (int k = 0; k < BSL; ++k) low[(size_t)(i - 1) * BSL + k] = scale * w[k];
    }
    /* :114-116 high = reverse(map(reverse, low * (-1)^d)) */
    const double sgn = (d % 2 == 0) ? 1.0 : -1.0;
    for (int i = 0; i < BPC; ++i)
        for (int k = 0; k < BSL; ++k)
            high[(size_t)i * BSL + k] = low[(size_t)(BPC - 1 - i) * BSL + (BSL - 1 - k)] * sgn;
    return ORC_OK;
}

/* derivative_operator.jl:137-205 -- non-uniform CenteredDifference (dx[len+1]) */
int orc_centered_nonuniform(int d, int a, const double* dx, int len, double* stencil,
                            double* low, double* high)
{
    int SL, BSL, BPC;
    if (orc_centered_sizes(d, a, &SL, &BSL, &BPC) != ORC_OK) return ORC_EINVAL;
    if (SL > 64 || BSL > 64 || len < 2 * BPC + 1 || len + 1 < BSL) return ORC_EINVAL;
    double sx[64], w[64], lx[64], hx[64];
    const int half = SL / 2;
    /* interior_x = BPC+2 : len+1-BPC (padded point index i), generate_coordinates :137-148 */
    int row = 0;
    for (int i = BPC + 2; i <= len + 1 - BPC; ++i, ++row) {
        /* stencil_x .= stencil_x .* zero(T): the centre entry stays +0.0 and every other
         * entry is overwritten below */
        for (int k = 0; k < SL; ++k) sx[k] = 0.0;
        for (int idx = 1; idx <= half; ++idx) {
            int s1 = idx + half;        /* index(idx, len) - 1 (0-based) */
            int s2 = -idx + half;       /* index(-idx, len) - 1 */
            sx[s1] = sx[s1 - 1] + dx[(i + idx - 1) - 1];
            sx[s2] = sx[s2 + 1] - dx[(i - idx) - 1];
        }
        if (orc_calculate_weights(d, 0.0, sx, SL, w) != ORC_OK) return ORC_EINVAL;
        for (int k = 0; k < SL; ++k) stencil[(size_t)row * SL + k] = w[k];
    }
    /* low_boundary_x = [0; cumsum(dx[1:BSL-1])] */
    lx[0] = 0.0;
    {
        double acc = 0.0;
        for (int m = 1; m < BSL; ++m) { acc += dx[m - 1]; lx[m] = acc; }
    }
    for (int i = 1; i <= BPC; ++i) {
        if (orc_calculate_weights(d, lx[i], lx, BSL, w) != ORC_OK) return ORC_EINVAL;
        for (int k = 0; k < BSL; ++k) low[(size_t)(i - 1) * BSL + k] = w[k];
    }
    /* high_boundary_x = cumsum(dx[end-BSL+1:end]) ; dx has len+1 entries */
    {
        double acc = 0.0;
        for (int m = 0; m < BSL; ++m) { acc += dx[(len + 1 - BSL) + m]; hx[m] = acc; }
    }
    /* for i in BPC:-1:1 -> x0 = high_boundary_x[end - i] */
    for (int j = 0; j < BPC; ++j) {
        int i = BPC - j;
        if (orc_calculate_weights(d, hx[BSL - 1 - i], hx, BSL, w) != ORC_OK) return ORC_EINVAL;
        for (int k = 0; k < BSL; ++k) high[(size_t)j * BSL + k] = w[k];
    }
    return ORC_OK;
}

/* derivative_operator.jl:454-512 -- uniform UpwindDifference */
int orc_upwind_uniform(int d, int a, double dx, int len, int offside, double* stencil,
                       double* low, double* high)
{
    int SL, BSL, BPC;
    if (orc_upwind_sizes(d, a, offside, &SL, &BSL, &BPC) != ORC_OK || len < 1) return ORC_EINVAL;
    if (SL > 64 || BPC < 0) return ORC_EINVAL;
    double xs[64], w[64];
    const double scale = 1.0 / pow(dx, (double)d);
    for (int k = 0; k < SL; ++k) xs[k] = (double)(k - offside);   /* (0-off):(SL-1-off) */
    if (orc_calculate_weights(d, 0.0, xs, SL, w) != ORC_OK) return ORC_EINVAL;
    for (int k = 0; k < SL; ++k) stencil[k] = scale * w[k];
    for (int k = 0; k < BSL; ++k) xs[k] = (double)k;   /* low_boundary_x = 0:BSL-1 */
    for (int i = 1; i <= BPC; ++i) {   /* L_boundary_deriv_spots = 1:(BSL-2-off) */
        if (orc_calculate_weights(d, (double)i, xs, BSL, w) != ORC_OK) return ORC_EINVAL;
        for (int k = 0; k < BSL; ++k) low[(size_t)(i - 1) * BSL + k] = scale * w[k];
    }
    /* high_boundary_x = 0:-1:-(BSL-1), R spots = -1:-1:-(BSL-2), scale (-1/dx)^d */
    const double hscale = pow(-1.0 / dx, (double)d);
    for (int k = 0; k < BSL; ++k) xs[k] = -(double)k;
    for (int m = 1; m <= BSL - 2; ++m) {
        if (orc_calculate_weights(d, -(double)m, xs, BSL, w) != ORC_OK) return ORC_EINVAL;
        for (int k = 0; k < BSL; ++k) high[(size_t)(m - 1) * BSL + k] = hscale * w[k];
    }
    return ORC_OK;
}

/* derivative_operator.jl:515-639 -- non-uniform UpwindDifference (dx[len+1]) */
int orc_upwind_nonuniform(int d, int a, const double* dx, int len, int offside,
                          double* stencil, double* low, double* high)
{
    int SL, BSL, BPC;
    if (orc_upwind_sizes(d, a, offside, &SL, &BSL, &BPC) != ORC_OK) return ORC_EINVAL;
    const int off = offside, stl = SL;
    if (BPC < 0 || len < 2 * BPC + 1 || len + 2 < BSL) return ORC_EINVAL;
    /* x = [0.0, cumsum(dx)...] has len+2 points; X(k) is 1-based */
    double* x = (double*)malloc(sizeof(double) * (size_t)(len + 2));
    if (!x) return ORC_ENOMEM;
    x[0] = 0.0;
    {
        double acc = 0.0;
        for (int k = 1; k <= len + 1; ++k) { acc += dx[k - 1]; x[k] = acc; }
    }
#define X(k) x[(k) - 1]
#define WGT(dst, x0, first, n) \
    do { if (orc_calculate_weights(d, (x0), &X(first), (n), (dst)) != ORC_OK) { free(x); return ORC_EINVAL; } } while (0)
    const int nin = len - 2 * BPC;
    const int nhi = BPC + off;
    /* low_boundary_coefs[1, :] (upwind), :531-554 */
    if (off == 0) {
        for (int i = 1; i <= BPC; ++i) WGT(&low[(size_t)(i - 1) * BSL], X(i + 1), i + 1, BSL);
    } else {
        int r = 0;
        for (int i = 1; i <= off - 1; ++i, ++r) WGT(&low[(size_t)r * BSL], X(i + 1), 1, BSL);
        for (int i = off; i <= BPC; ++i, ++r) WGT(&low[(size_t)r * BSL], X(i + 1), i + 1 - off, BSL);
    }
    /* low_boundary_coefs[2, :] (downwind), :555-561 */
    for (int i = 1; i <= BPC; ++i) WGT(&low[(size_t)(BPC + i - 1) * BSL], X(i + 1), 1, BSL);
    /* stencil_coefs, :565-579 */
    for (int i = BPC + 1; i <= len - BPC; ++i) {
        int r = i - BPC - 1;
        WGT(&stencil[(size_t)r * stl], X(i + 1), i + 1 - off, stl);
        WGT(&stencil[(size_t)(nin + r) * stl], X(i + 1), i - stl + 2 + off, stl);
    }
    /* high_boundary_coefs[1, :] (upwind), :582-588 */
    {
        int r = 0;
        for (int i = len - BPC + 1 - off; i <= len; ++i, ++r)
            WGT(&high[(size_t)r * BSL], X(i + 1), len - BSL + 3, BSL);
    }
    /* high_boundary_coefs[2, :] (downwind), :589-622 */
    {
        double* hd = &high[(size_t)nhi * BSL];
        int r = 0;
        if (off == 0) {
            for (int i = len - BPC + 1; i <= len; ++i, ++r)
                WGT(&hd[(size_t)r * BSL], X(i + 1), i - stl + 2, BSL);
        } else if (off == 1) {
            for (int i = len - BPC + 1 - off; i <= len - off + 1; ++i, ++r)
                WGT(&hd[(size_t)r * BSL], X(i + 1), i - stl + 2 + off, BSL);
        } else {
            for (int i = len - BPC + 1 - off; i <= len - off + 1; ++i, ++r)
                WGT(&hd[(size_t)r * BSL], X(i + 1), i - stl + 2 + off, BSL);
            for (int i = len - off + 2; i <= len; ++i, ++r)
                WGT(&hd[(size_t)r * BSL], X(i + 1), len - stl + 3, BSL);
        }
    }
#undef WGT
#undef X
    free(x);
    return ORC_OK;
}

/* ------------------------------------------------------------------------------------ */
/* bc_operators.jl:26-39 -- RobinBC on a uniform grid */
int orc_robin_uniform(double al, double bl, double cl, double ar, double br, double cr,
                      double dx, int order, double* a_l, double* b_l, double* a_r,
                      double* b_r)
{
    double xs[64], s[64];
    if (order < 1 || order + 1 > 64) return ORC_EINVAL;
    for (int k = 0; k <= order; ++k) xs[k] = (double)(k + 1);   /* one(T):order+1 */
    if (orc_calculate_weights(1, 1.0, xs, order + 1, s) != ORC_OK) return ORC_EINVAL;
    const double dl = al * dx / bl + s[0];
    const double dr = ar * dx / br - s[0];
    for (int k = 1; k <= order; ++k) a_l[k - 1] = -s[k] / dl;            /* -s[2:end] ./ ... */
    for (int k = 0; k < order; ++k) a_r[k] = s[order - k] / dr;          /* s[end:-1:2] ./ ... */
    *b_l = cl / (al + bl * s[0] / dx);
    *b_r = cr / (ar - br * s[0] / dx);
    return ORC_OK;
}

/* bc_operators.jl:40-60 -- RobinBC on a non-uniform grid */
int orc_robin_nonuniform(double al, double bl, double cl, double ar, double br, double cr,
                         const double* dx, int ndx, int order, double* a_l, double* b_l,
                         double* a_r, double* b_r)
{
    double xs[64], s[64];
    const int ns = order + 1;
    if (order < 1 || ns > 64 || ndx < ns) return ORC_EINVAL;
    for (int k = 0; k < ns; ++k) xs[k] = (double)(k + 1);
    const double* dx_l = dx;                 /* dx[1:ns] */
    const double* dx_r = dx + (ndx - ns);    /* dx[end-ns+1:end] */
    if (orc_calculate_weights(1, 1.0, xs, ns, s) != ORC_OK) return ORC_EINVAL;
    const double denom_l = al + bl * s[0] / dx_l[0];
    const double denom_r = ar - br * s[0] / dx_r[ns - 1];
    for (int k = 1; k < ns; ++k) a_l[k - 1] = -bl * s[k] / (denom_l * dx_l[k]);
    for (int m = 0; m < order; ++m) a_r[m] = br * s[ns - 1 - m] / (denom_r * dx_r[m]);
    *b_l = cl / denom_l;
    *b_r = cr / denom_r;
    return ORC_OK;
}

/* bc_operators.jl:85-170 -- GeneralBC.  Row i (1-based, i <= n-2) of S holds the weights of
 * the i-th derivative on the points 1..order+i (at point 1 on the left, at point order+i on
 * the right), zero-padded to order+n-2 columns and divided by dx^i (per column for a vector
 * step); the ghost solves the BC with the ghost's column (s0) split off. */
static double ipow(double x, int i)
{
    double r = x;
    for (int k = 1; k < i; ++k) r *= x;
    return r;
}

static int general_side(const double* alpha, int n, const double* dx, int ndx, int order,
                        int right, double* a, double* b)
{
    enum { MAXC = 32 };
    const int ncol = order + n - 2;   /* columns of S */
    if (n < 2 || order < 1 || ncol > MAXC || n - 2 > MAXC) return ORC_EINVAL;
    if (ndx > 0 && ndx < ncol) return ORC_EINVAL;
    double S[MAXC][MAXC];
    double xs[MAXC], w[MAXC];
    for (int i = 1; i <= n - 2; ++i) {
        const int m = order + i;
        for (int k = 0; k < m; ++k) xs[k] = (double)(k + 1);
        if (orc_calculate_weights(i, right ? (double)m : 1.0, xs, m, w) != ORC_OK) return ORC_EINVAL;
        for (int j = 0; j < ncol; ++j) {
            const double v = j < m ? w[j] : 0.0;
            double h;
            if (ndx == 0) h = ipow(dx[0], i);
            else h = ipow(right ? dx[ndx - 1 - j] : dx[j], i);   /* dx_r = reverse(dx[end-ncol+1:end]) */
            S[i - 1][j] = v / h;
        }
    }
    /* s0 = S[:, 1] (left) or S[:, end] (right); the rest Sl = S[:, 2:end] or Sr = S[:, end-1:-1:1] */
    double denom = alpha[1];
    for (int i = 0; i < n - 2; ++i) denom += alpha[i + 2] * S[i][right ? ncol - 1 : 0];
    const int nt = ncol - 1;
    double t[MAXC];
    for (int j = 0; j < nt; ++j) {
        const int col = right ? (ncol - 2 - j) : (j + 1);
        double acc = 0.0;
        for (int i = 0; i < n - 2; ++i) acc += alpha[i + 2] * S[i][col];
        t[j] = -acc / denom;
    }
    for (int j = 0; j < nt; ++j) a[j] = right ? t[nt - 1 - j] : t[j];   /* a_r = reverse(...) */
    *b = -alpha[0] / denom;
    return ORC_OK;
}

int orc_general(const double* alpha_l, int nl, const double* alpha_r, int nr, const double* dx,
                int ndx, int order, double* a_l, double* b_l, double* a_r, double* b_r)
{
    int r = general_side(alpha_l, nl, dx, ndx, order, 0, a_l, b_l);
    if (r == ORC_OK) r = general_side(alpha_r, nr, dx, ndx, order, 1, a_r, b_r);
    return r;
}

/* ------------------------------------------------------------------------------------ */
/* Line kernels.  P(j) is the 1-based padded vector, OUT(i) the 1-based output. */
#define P(j) x[(j) - 1]
#define OUT(i) out[(i) - 1]
#define STORE(i, v) \
    do { OUT(i) = overwrite ? ((v) + 0.0) : ((v) + OUT(i)); } while (0)

/* convolutions.jl:27-118 -- centered, AbstractVector */
static void centered_plain(const orc_op* A, const double* x, double* out, int overwrite)
{
    const int len = A->len, SL = A->SL, BSL = A->BSL, BPC = A->BPC;
    const double* coeff = A->coeff;
    const int mid = SL / 2;
    /* convolve_interior! :40-60 */
    for (int i = 1 + BPC; i <= len - BPC; ++i) {
        const double* st = A->uniform ? A->stencil : A->stencil + (size_t)(i - BPC - 1) * SL;
        const double c = coeff[i - 1];
        double acc = 0.0;
        for (int idx = 1; idx <= SL; ++idx) acc += c * st[idx - 1] * P(i - mid + idx);
        STORE(i, acc);
    }
    /* convolve_BC_left! :76-95 */
    for (int i = 1; i <= BPC; ++i) {
        const double* st = A->low + (size_t)(i - 1) * BSL;
        const double c = coeff[i - 1];
        double acc = 0.0;
        for (int idx = 1; idx <= BSL; ++idx) acc += c * st[idx - 1] * P(idx);
        STORE(i, acc);
    }
    /* convolve_BC_right! :97-118 (coeff[i], i in 1:BPC, as written) */
    const int xlen = len + 2;
    for (int i = 1; i <= BPC; ++i) {
        const double* st = A->high + (size_t)(i - 1) * BSL;
        const double c = coeff[i - 1];
        double acc = 0.0;
        for (int idx = BSL - 1; idx >= 0; --idx) acc += c * st[BSL - 1 - idx] * P(xlen - idx);
        STORE(len - BPC + i, acc);
    }
}

/* convolutions.jl:367-469 -- centered, BoundaryPaddedVector(l, r, u) */
static void centered_bpv(const orc_op* A, double l, double r, const double* u, double* out,
                         int overwrite)
{
    const int len = A->len, SL = A->SL, BSL = A->BSL, BPC = A->BPC;
    const double* coeff = A->coeff;
    const int mid = SL / 2 + 1;
#define U(j) u[(j) - 1]
    /* convolve_interior! :367-400 -- rows 2+BPC : len-BPC-1, coeff[i-BPC] */
    for (int i = 2 + BPC; i <= len - BPC - 1; ++i) {
        const double* st = A->uniform ? A->stencil : A->stencil + (size_t)(i - BPC - 1) * SL;
        const double c = coeff[i - BPC - 1];
        double acc = 0.0;
        for (int idx = 1; idx <= SL; ++idx) acc += c * st[idx - 1] * U((i - 1) - (mid - idx) + 1);
        STORE(i, acc);
    }
    /* convolve_BC_left! :402-434 */
    for (int i = 1; i <= BPC; ++i) {
        const double* st = A->low + (size_t)(i - 1) * BSL;
        const double c = coeff[i - 1];
        double acc = c * st[0] * l;
        for (int idx = 2; idx <= BSL; ++idx) acc += c * st[idx - 1] * U(idx - 1);
        STORE(i, acc);
    }
    {
        const int i = 1 + BPC;
        const double* st = A->uniform ? A->stencil : A->stencil + (size_t)(i - BPC - 1) * SL;
        const double c = coeff[i - BPC - 1];
        double acc = c * st[0] * l;
        for (int idx = 2; idx <= SL; ++idx) acc += c * st[idx - 1] * U((i - 1) - (mid - idx) + 1);
        STORE(i, acc);
    }
    /* convolve_BC_right! :436-469 */
    {
        const int i = len - BPC;
        const double* st = A->uniform ? A->stencil : A->stencil + (size_t)(i - BPC - 1) * SL;
        const double c = coeff[i - BPC - 1];
        double acc = c * st[SL - 1] * r;
        for (int idx = 1; idx <= SL - 1; ++idx) acc += c * st[idx - 1] * U((i - 1) - (mid - idx) + 1);
        STORE(i, acc);
    }
    const int bc_start = len - BPC;
    for (int i = 1; i <= BPC; ++i) {
        const double* st = A->high + (size_t)(i - 1) * BSL;
        const double c = coeff[bc_start + i - 1];
        double acc = c * st[BSL - 1] * r;
        for (int idx = BSL - 1; idx >= 1; --idx) acc += c * st[BSL - 1 - idx] * U(len - idx + 1);
        STORE(bc_start + i, acc);
    }
#undef U
}

/* convolutions.jl:123-215 -- upwind on a uniform grid, AbstractVector */
static void upwind_uniform_plain(const orc_op* A, const double* x, double* out, int overwrite)
{
    const int len = A->len, SL = A->SL, BSL = A->BSL, BPC = A->BPC, off = A->offside;
    const double* coeff = A->coeff;
    const double* s = A->stencil;
    const int xlen = len + 2;
    const double dsgn = (A->d % 2 == 0) ? 1.0 : -1.0;   /* (-1)^derivative_order */
    /* convolve_interior! :123-144 */
    for (int i = 1 + BPC; i <= len - BPC - off; ++i) {
        const double c = coeff[i - 1];
        double acc = 0.0;
        for (int idx = 1; idx <= SL; ++idx) {
            /* c >= 0: stencil ; else reverse(stencil) (d even) / -1*reverse(stencil) */
            const double w = (c >= 0) ? s[idx - 1] : dsgn * s[SL - idx];
            const double xv = (c < 0) ? P(i - SL + 1 + idx + off) : P(i + idx - off);
            acc += c * w * xv;
        }
        STORE(i, acc);
    }
    /* convolve_BC_left! :146-168 */
    for (int i = 1; i <= BPC; ++i) {
        const double c = coeff[i - 1];
        double acc = 0.0;
        if (c >= 0 && i + SL <= xlen && i >= off) {
            for (int idx = 1; idx <= SL; ++idx) acc += c * s[idx - 1] * P(i + idx - off);
        } else {
            const double* st = A->low + (size_t)(i - 1) * BSL;
            for (int idx = 1; idx <= BSL; ++idx) acc += c * st[idx - 1] * P(idx);
        }
        STORE(i, acc);
    }
    /* convolve_BC_right! :170-215 */
    for (int i = 1; i <= BPC + off; ++i) {
        const int row = len - BPC + i - off;
        const double c = coeff[row - 1];
        double acc = 0.0;
        if (c < 0 && xlen - SL - BPC + i >= 1 && i <= BPC + 1) {
            for (int idx = 1; idx <= SL; ++idx)
                acc += c * (dsgn * s[SL - idx]) * P(xlen - SL + idx - BPC + i - 1);
        } else if (c < 0 && xlen - SL - BPC + i >= 1 && i > BPC + 1) {
            const double* st = A->high + (size_t)(BPC + off + 1 - i - 1) * BSL;
            for (int idx = 1; idx <= BSL; ++idx)
                acc += c * (dsgn * st[BSL - idx]) * P(xlen - BSL + idx);
        } else if (c >= 0 && i < off + 1) {
            for (int idx = 1; idx <= SL; ++idx) acc += c * s[idx - 1] * P(xlen - SL + i + idx - off);
        } else {
            const double* st = A->high + (size_t)(i - 1) * BSL;
            for (int idx = 1; idx <= BSL; ++idx) acc += c * st[idx - 1] * P(xlen - BSL + idx);
        }
        STORE(row, acc);
    }
}

/* convolutions.jl:221-362 -- upwind on a non-uniform grid, AbstractVector.  coeff_kind
 * selects the `A.coeff_func isa Number` branches of convolve_interior! (:234-253). */
static void upwind_nonuniform_plain(const orc_op* A, const double* x, double* out,
                                    int overwrite, int coeff_kind)
{
    const int len = A->len, stl = A->SL, bstl = A->BSL, bpc = A->BPC, off = A->offside;
    const double* coeff = A->coeff;
    const int nin = len - 2 * bpc, nhi = bpc + off;
#define UP(r) (A->stencil + (size_t)((r) - 1) * stl)          /* stencil_coefs[1, r] */
#define DN(r) (A->stencil + (size_t)(nin + (r) - 1) * stl)    /* stencil_coefs[2, r] */
#define LOW(k, r) (A->low + (size_t)((k) - 1) * bpc * bstl + (size_t)((r) - 1) * bstl)
#define HIGH(k, r) (A->high + (size_t)((k) - 1) * nhi * bstl + (size_t)((r) - 1) * bstl)
    /* convolve_interior! :221-275 */
    for (int i = bpc + 1; i <= len - bpc - off; ++i) {
        const double c = coeff[i - 1];
        int fwd;
        if (coeff_kind == ORC_COEFF_NUM_POS) fwd = 1;
        else if (coeff_kind == ORC_COEFF_NUM_NEG) fwd = 0;
        else fwd = (c >= 0);
        double acc = 0.0;
        if (fwd) {
            const double* st = UP(i - bpc);
            for (int idx = 1; idx <= stl; ++idx) acc += c * st[idx - 1] * P(i + idx - off);
        } else {
            const double* st = DN(i - bpc);
            for (int idx = 1; idx <= stl; ++idx) acc += c * st[idx - 1] * P(i - stl + 1 + idx + off);
        }
        STORE(i, acc);
    }
    /* convolve_BC_left! :277-317 */
    for (int i = 1; i <= bpc; ++i) {
        const double c = coeff[i - 1];
        double acc = 0.0;
        if (c >= 0 && off == 0) {
            const double* st = LOW(1, i);
            for (int idx = 1; idx <= stl; ++idx) acc += c * st[idx - 1] * P(i + idx);
        } else if (c >= 0 && i < off) {
            const double* st = LOW(1, i);
            for (int idx = 1; idx <= stl; ++idx) acc += c * st[idx - 1] * P(idx);
        } else if (c >= 0 && i >= off) {
            const double* st = LOW(1, i);
            for (int idx = 1; idx <= stl; ++idx) acc += c * st[idx - 1] * P(i + idx - off);
        } else {
            const double* st = LOW(2, i);
            for (int idx = 1; idx <= bstl; ++idx) acc += c * st[idx - 1] * P(idx);
        }
        STORE(i, acc);
    }
    /* convolve_BC_right! :319-362 */
    for (int i = len - bpc + 1 - off; i <= len; ++i) {
        const double c = coeff[i - 1];
        double acc = 0.0;
        if (c < 0) {
            const double* st = HIGH(2, i - len + bpc + off);
            if (i <= len - off) {
                for (int idx = 1; idx <= stl; ++idx) acc += c * st[idx - 1] * P(i - stl + 1 + idx + off);
            } else {
                for (int idx = 1; idx <= stl; ++idx) acc += c * st[idx - 1] * P(len - stl + 2 + idx);
            }
        } else {
            if (i <= len - bpc) {
                const double* st = UP(i - bpc);
                for (int idx = 1; idx <= stl; ++idx) acc += c * st[idx - 1] * P(i - stl + 1 + idx + off);
            } else {
                const double* st = HIGH(1, i - len + bpc + off);
                for (int idx = 1; idx <= bstl; ++idx) acc += c * st[idx - 1] * P(len - bstl + 2 + idx);
            }
        }
        STORE(i, acc);
    }
#undef UP
#undef DN
#undef LOW
#undef HIGH
}
#undef P
#undef OUT
#undef STORE

/* Bounds check shared by the line drivers: every padded index a kernel touches must lie
 * in 1..len+2 (the reference's @turbo loops do not check).  Conservative: the row
 * windows of the interior and the one-sided boundary tables must fit. */
static int op_valid(const orc_op* A)
{
    if (!A || A->len < 1 || A->SL < 1 || A->BSL < 1 || A->BPC < 0) return 0;
    const int len = A->len;
    if (A->BSL > len + 2 || A->SL > len + 2) return 0;
    if (!A->wind) {
        if (len < 2 * A->BPC + 2) return 0;   /* first/last interior rows are distinct */
    } else {
        if (len < 2 * A->BPC + A->offside + 1) return 0;
        if (A->SL + A->BPC + A->offside > len + 2) return 0;
    }
    return 1;
}

int orc_mul_line(const orc_op* op, const double* x, double* out, int overwrite)
{
    if (!op_valid(op)) return ORC_EINVAL;
    if (!op->wind) centered_plain(op, x, out, overwrite);
    else if (op->uniform) upwind_uniform_plain(op, x, out, overwrite);
    else upwind_nonuniform_plain(op, x, out, overwrite, op->coeff_kind);
    return ORC_OK;
}

/* The BoundaryPaddedVector upwind kernels (convolutions.jl:479-730) read P(j) through the
 * same index expressions as the AbstractVector ones, substituting l / r at j = 1 / len+2;
 * the non-uniform interior (:597-631) has no `coeff_func isa Number` branch, so it always
 * takes the per-row sign. */
int orc_mul_line_bpv(const orc_op* op, double l, double r, const double* u, double* out,
                     int overwrite)
{
    if (!op_valid(op)) return ORC_EINVAL;
    if (!op->wind) {
        centered_bpv(op, l, r, u, out, overwrite);
        return ORC_OK;
    }
    const int len = op->len;
    double* x = (double*)malloc(sizeof(double) * (size_t)(len + 2));
    if (!x) return ORC_ENOMEM;
    x[0] = l;
    memcpy(x + 1, u, sizeof(double) * (size_t)len);
    x[len + 1] = r;
    if (op->uniform) upwind_uniform_plain(op, x, out, overwrite);
    else upwind_nonuniform_plain(op, x, out, overwrite, ORC_COEFF_VECTOR);
    free(x);
    return ORC_OK;
}

/* ------------------------------------------------------------------------------------ */
static long long prod_ll(const long long* s, int n)
{
    long long p = 1;
    for (int i = 0; i < n; ++i) p *= s[i];
    return p;
}

/* Linear offset of (multi-index `idx` with idx[axis] ignored) in an array of size `sz`,
 * adding `shift[k]` to each non-axis index. */
static long long line_base(int ndim, const long long* sz, const long long* idx,
                           const long long* shift, int axis)
{
    long long off = 0, stride = 1;
    for (int k = 0; k < ndim; ++k) {
        if (k != axis) off += (idx[k] + shift[k]) * stride;
        stride *= sz[k];
    }
    return off;
}

/* Apply op along its axis to every line of M (size msz), writing out (size osz).  The
 * transverse index into M is the out index plus mshift[k]. */
static int apply_lines(const orc_op* op, int ndim, const long long* osz, double* out,
                       const long long* msz, const double* M, const long long* mshift,
                       int overwrite)
{
    const int ax = op->axis;
    const int len = op->len;
    long long ostride = 1, mstride = 1;
    for (int k = 0; k < ax; ++k) { ostride *= osz[k]; mstride *= msz[k]; }
    const long long nlines = prod_ll(osz, ndim) / osz[ax];
    int err = ORC_OK;
#pragma omp parallel num_threads(g_threads)
    {
        double* xl = (double*)malloc(sizeof(double) * (size_t)(len + 2));
        double* ol = (double*)malloc(sizeof(double) * (size_t)len);
        long long idx[8];
#pragma omp for schedule(static)
        for (long long ln = 0; ln < nlines; ++ln) {
            if (!xl || !ol) { err = ORC_ENOMEM; continue; }
            /* decode the line number into the non-axis indices (column-major) */
            long long rem = ln;
            for (int k = 0; k < ndim; ++k) {
                if (k == ax) { idx[k] = 0; continue; }
                idx[k] = rem % osz[k];
                rem /= osz[k];
            }
            const long long ob = line_base(ndim, osz, idx, (const long long[8]){0}, ax);
            const long long mb = line_base(ndim, msz, idx, mshift, ax);
            for (int j = 0; j < len + 2; ++j) xl[j] = M[mb + (long long)j * mstride];
            if (!overwrite)
                for (int j = 0; j < len; ++j) ol[j] = out[ob + (long long)j * ostride];
            orc_mul_line(op, xl, ol, overwrite);
            for (int j = 0; j < len; ++j) out[ob + (long long)j * ostride] = ol[j];
        }
        free(xl);
        free(ol);
    }
    return err;
}

/* derivative_operator_functions.jl:27-69 */
int orc_mul_nd(const orc_op* op, int ndim, const long long* out_size, double* out,
               const long long* M_size, const double* M, int overwrite)
{
    if (!op_valid(op) || ndim < 1 || ndim > 8 || op->axis < 0 || op->axis >= ndim) return ORC_EINVAL;
    long long shift[8];
    for (int k = 0; k < ndim; ++k) {   /* :33-35, :40 */
        if (k == op->axis) {
            if (M_size[k] != out_size[k] + 2 || out_size[k] != op->len) return ORC_EINVAL;
            shift[k] = 0;
        } else if (M_size[k] == out_size[k]) {
            shift[k] = 0;
        } else if (M_size[k] == out_size[k] + 2) {
            shift[k] = 1;   /* minimal padding view 2:end-1, :50-57 */
        } else {
            return ORC_EINVAL;
        }
    }
    return apply_lines(op, ndim, out_size, out, M_size, M, shift, overwrite);
}

/* ghost_derivative_operator.jl:15-37 per op, composite_operators.jl:64-65 for the sum */
int orc_ghost_apply(int nops, const orc_op* ops, int ndim, const long long* size,
                    const orc_bc* bcs, const double* u, double* out, int bpv_1d)
{
    if (nops < 1 || ndim < 1 || ndim > 8) return ORC_EINVAL;
    const long long n = prod_ll(size, ndim);
    for (int t = 0; t < nops; ++t) {
        if (!op_valid(&ops[t]) || ops[t].axis < 0 || ops[t].axis >= ndim) return ORC_EINVAL;
        if (ops[t].len != size[ops[t].axis]) return ORC_EINVAL;
        const orc_bc* q = &bcs[ops[t].axis];
        if (q->nl < 0 || q->nr < 0 || q->nl > size[ops[t].axis] || q->nr > size[ops[t].axis])
            return ORC_EINVAL;
        if (q->lines) {
            const long long nl_lines = n / size[ops[t].axis];
            for (long long l = 0; l < nl_lines; ++l) {
                /* a line's own atomic BC: affine, or PeriodicBC (the line wraps) */
                const orc_bc* e = &q->lines[l];
                if (e->periodic != 0 && e->periodic != 1) return ORC_EINVAL;
                if (!e->periodic && (e->nl < 0 || e->nr < 0 || e->nl > size[ops[t].axis] ||
                                     e->nr > size[ops[t].axis]))
                    return ORC_EINVAL;
            }
        }
    }
    double* tmp = (double*)malloc(sizeof(double) * (size_t)n);
    if (!tmp) return ORC_ENOMEM;
    int err = ORC_OK;
    for (int t = 0; t < nops && err == ORC_OK; ++t) {
        const orc_op* op = &ops[t];
        double* dst = (t == 0) ? out : tmp;
        const int ax = op->axis;
        const int len = op->len;
        const orc_bc* q = &bcs[ax];
        long long stride = 1;
        for (int k = 0; k < ax; ++k) stride *= size[k];
        const long long nlines = n / size[ax];
#pragma omp parallel num_threads(g_threads)
        {
            double* xl = (double*)malloc(sizeof(double) * (size_t)(len + 2));
            double* ol = (double*)malloc(sizeof(double) * (size_t)len);
            long long idx[8];
#pragma omp for schedule(static)
            for (long long ln = 0; ln < nlines; ++ln) {
                if (!xl || !ol) { err = ORC_ENOMEM; continue; }
                long long rem = ln;
                for (int k = 0; k < ndim; ++k) {
                    if (k == ax) { idx[k] = 0; continue; }
                    idx[k] = rem % size[k];
                    rem /= size[k];
                }
                const long long zero[8] = {0};
                const long long b = line_base(ndim, size, idx, zero, ax);
                /* u[I, :, J] copied (multi_dim_bc_operators.jl:34) */
                for (int j = 0; j < len; ++j) xl[j + 1] = u[b + (long long)j * stride];
                /* Q*u_line: bc_operators.jl:188-191 (a_l . u[1:nl] + b_l), or the
                 * periodic wrap of :192 (u[end], u[1]) */
                double gl = 0.0, gr = 0.0;
                const orc_bc* ql = q->lines ? &q->lines[ln] : q;
                if (ql->periodic == 1) {
                    gl = xl[len];
                    gr = xl[1];
                } else if (ql->periodic == 2) {
                    gl = xl[1];
                    gr = xl[len];
                } else {
                    for (int k = 0; k < ql->nl; ++k) gl += ql->a_l[k] * xl[1 + k];
                    for (int k = 0; k < ql->nr; ++k) gr += ql->a_r[k] * xl[1 + len - ql->nr + k];
                    gl = gl + ql->b_l;
                    gr = gr + ql->b_r;
                }
                xl[0] = gl;
                xl[len + 1] = gr;
                if (ndim == 1 && bpv_1d) orc_mul_line_bpv(op, gl, gr, xl + 1, ol, 1);
                else orc_mul_line(op, xl, ol, 1);
                for (int j = 0; j < len; ++j) dst[b + (long long)j * stride] = ol[j];
            }
            free(xl);
            free(ol);
        }
        if (t > 0)
            for (long long p = 0; p < n; ++p) out[p] = out[p] + tmp[p];
    }
    free(tmp);
    return err;
}

/* derivative_operator_functions.jl:414-461 (2-D) and :768-801 (3-D) */
int orc_composite_mul_padded(int nops, const orc_op* ops, int ndim,
                             const long long* out_size, const double* M, double* out)
{
    if (nops < 1 || (ndim != 2 && ndim != 3)) return ORC_EINVAL;
    long long msz[3];
    for (int k = 0; k < ndim; ++k) msz[k] = out_size[k] + 2;
    int opdims[3] = {0, 0, 0};
    for (int t = 0; t < nops; ++t) {
        if (!op_valid(&ops[t]) || ops[t].axis < 0 || ops[t].axis >= ndim) return ORC_EINVAL;
        if (ops[t].len != out_size[ops[t].axis]) return ORC_EINVAL;
        if (ndim == 2) opdims[ops[t].axis == 0 ? 0 : 1] = 1;
    }
    for (int t = 0; t < nops; ++t) {
        const orc_op* op = &ops[t];
        const int overwrite = (t == 0);
        long long vsz[3], shift[3];
        int use_view;
        if (ndim == 2) use_view = (op->axis == 0) ? opdims[1] : opdims[0];
        else use_view = 1;
        for (int k = 0; k < ndim; ++k) {
            if (k == op->axis) { vsz[k] = msz[k]; shift[k] = 0; }
            else if (use_view) { vsz[k] = out_size[k]; shift[k] = 0; }   /* view 1:n */
            else { vsz[k] = msz[k]; shift[k] = 1; }                      /* _mul! trims 2:end-1 */
        }
        (void)vsz;
        int err = apply_lines(op, ndim, out_size, out, msz, M, shift, overwrite);
        if (err != ORC_OK) return err;
    }
    return ORC_OK;
}
