// dop_fast.cpp -- host side of the tiled interior kernels: eligibility, interior box,
// padded symmetric weights, shell decomposition and the tile-shape candidates.
#include "dop_fast.hpp"

#include <algorithm>

namespace dop {

namespace {

constexpr int kMaxK2D = 13;
constexpr int kMaxK3D = 9;   // 3-D stages 2*H halo rows through 8 waves: H <= 4

int64_t ceil_div(int64_t a, int64_t b) { return (a + b - 1) / b; }

int64_t shell_boxes(const PlanDev& P, const int32_t* lo, const int32_t* hi, ShellArgs* out)
{
    ShellArgs& S = *out;
    S = ShellArgs{};
    int64_t total = 0;
    for (int a = P.ndim - 1; a >= 0; --a) {
        for (int side = 0; side < 2; ++side) {
            int64_t org[3], ext[3];
            for (int b = 0; b < 3; ++b) {
                if (b > a) { org[b] = lo[b]; ext[b] = hi[b] - lo[b]; }      // inside box
                else if (b < a) { org[b] = 0; ext[b] = P.n[b]; }            // full
                else if (side == 0) { org[b] = 0; ext[b] = lo[b]; }
                else { org[b] = hi[b]; ext[b] = P.n[b] - hi[b]; }
            }
            const int64_t cnt = ext[0] * ext[1] * ext[2];
            if (cnt <= 0) continue;
            for (int b = 0; b < 3; ++b) {
                S.org[S.nbox][b] = org[b];
                S.ext[S.nbox][b] = ext[b];
            }
            S.start[S.nbox] = total;
            total += cnt;
            ++S.nbox;
        }
    }
    S.start[S.nbox] = total;
    return total;
}

}  // namespace

void build_fast(const HostPlan& H, FastPlan* fp)
{
    *fp = FastPlan{};
    const PlanDev& P = H.meta;
    if (!P.ghost || P.ndim < 1 || P.ndim > 3) return;
    const bool f64 = P.storage == kF64 && P.accum == kAccF64;
    const bool f32 = (P.storage == kF32 || P.storage == kBF16) &&
                     (P.accum == kAccF32 || P.accum == kAccF32Kahan || P.accum == kAccF32Wide);
    if (!f64 && !f32) return;
    // compensated modes: batched tiles only (single plans take the generic kernel)
    fp->kahan = (P.accum == kAccF32Kahan || P.accum == kAccF32Wide) ? 1 : 0;
    // one symmetric window for all axes
    int half = 0;
    for (int a = 0; a < P.ndim; ++a) {
        const AxisDev& X = P.ax[a];
        if (!X.active || X.rhi <= X.rlo || X.K < 1) return;
        half = std::max(half, std::max(-X.dmin, X.dmin + X.K - 1));
    }
    const int K = 2 * half + 1;
    if (K > (P.ndim <= 2 ? kMaxK2D : kMaxK3D)) return;
    FastArgs& A = fp->args;
    for (int a = 0; a < 3; ++a) A.n[a] = P.n[a];
    A.s1 = P.sdu[1];
    A.s2 = P.sdu[2];
    A.K = K;
    for (int a = 0; a < 3; ++a) {
        if (a >= P.ndim) { A.lo[a] = 0; A.hi[a] = 1; continue; }
        const AxisDev& X = P.ax[a];
        A.lo[a] = std::max(X.rlo, half);
        A.hi[a] = std::min(X.rhi, (int)P.n[a] - half);
        if (A.hi[a] <= A.lo[a]) return;
    }
    for (int a = 0; a < 3; ++a) {
        fp->nat_lo[a] = A.lo[a];
        fp->nat_hi[a] = A.hi[a];
    }
    // points per lane: 16-byte lane loads when rows stay aligned
    const int vx_native = (P.storage == kF64) ? 2 : 4;
    fp->VX = (P.n[0] % vx_native == 0) ? vx_native : 1;
    if (P.ndim >= 2) {
        // lanes store whole VX groups: the box's x edges are aligned to the widest VX of the
        // tile variants (the few columns this moves into the shell are evaluated there)
        const int al = (P.ndim == 2 && fp->VX > 1 && P.n[0] % (2 * fp->VX) == 0) ? 2 * fp->VX : fp->VX;
        A.lo[0] = (int)ceil_div(A.lo[0], al) * al;
        A.hi[0] = A.hi[0] / al * al;
        if (A.hi[0] <= A.lo[0]) return;
    }
    // padded weights (rows outside the natural box are zero).  The tiled kernels index every axis
    // per row unless all axes are uniform, so a uniform axis of a mixed plan is expanded to
    // one row per point.
    bool all_uniform = true;
    for (int a = 0; a < P.ndim; ++a) all_uniform = all_uniform && P.ax[a].uniform;
    for (int a = 0; a < P.ndim; ++a) {
        const AxisDev& X = P.ax[a];
        const AxisTables& T = H.axis[a];
        const int rows = all_uniform ? 1 : (int)P.n[a];
        fp->w64[a].assign((size_t)rows * K, 0.0);
        for (int r = 0; r < rows; ++r) {
            const int src = X.uniform ? 0 : r;
            if (!X.uniform && (r < fp->nat_lo[a] || r >= fp->nat_hi[a])) continue;
            for (int k = 0; k < X.K; ++k)
                fp->w64[a][(size_t)r * K + (X.dmin + k + half)] =
                    T.w64[X.uniform ? (size_t)k : (size_t)k * P.n[a] + src];
        }
        fp->w32[a].resize(fp->w64[a].size());
        for (size_t i = 0; i < fp->w64[a].size(); ++i) fp->w32[a][i] = (float)fp->w64[a][i];
        A.uniform[a] = X.uniform;
    }
    // shells: the domain minus the box, slabs from the slowest axis down
    fp->shell_total = shell_boxes(P, A.lo, A.hi, &fp->shell);
    {
        StripArgs& S = fp->strips;
        S = StripArgs{};
        S.nbox = fp->shell.nbox;
        S.vec_ok = (P.n[0] % kStripPts == 0) ? 1 : 0;
        int64_t tot = 0;
        for (int b = 0; b < S.nbox; ++b) {
            for (int k = 0; k < 3; ++k) {
                S.org[b][k] = fp->shell.org[b][k];
                S.ext[b][k] = fp->shell.ext[b][k];
            }
            S.nsx[b] = ceil_div(S.ext[b][0], kStripPts);
            S.xcol[b] = S.ext[b][0] < P.n[0] ? 1 : 0;
            S.start[b] = tot;
            tot += S.nsx[b] * S.ext[b][1] * S.ext[b][2];
        }
        S.start[S.nbox] = tot;
    }
    (void)shell_boxes(P, fp->nat_lo, fp->nat_hi, &fp->nat_shell);
    fp->kind = P.ndim;
}

// Candidates (autotuner).  2-D: chunk height {32,16,64,8} x prefetch {2,4} x XCD remap x
// points per lane {widest 16-byte, twice that}; 3-D: chunk depth {64,32} x block shape
// (waves, rows per wave) {(8,1), (16,1), (8,2), (16,2), (8,3)} x XCD remap x points per lane
// {widest, half}.  A block stages waves * rows y rows and computes all but the 2H halo rows.
namespace {
bool wide_ok(const FastPlan& fp) { return fp.VX > 1 && fp.args.n[0] % (2 * fp.VX) == 0; }

int gen1_count(const FastPlan& fp) { return fp.VX > 1 ? 40 : 20; }
int gen2_count(const FastPlan& fp) { return fp.VX > 1 ? 32 : 16; }
}  // namespace

int fast_variant_count(const FastPlan& fp)
{
    if (fp.kind == 1) return 1;
    if (fp.kind == 2) return wide_ok(fp) ? 32 : 16;
    if (fp.kind == 3) return gen1_count(fp) + gen2_count(fp);
    return 1;
}

bool fast_variant(const FastPlan& fp, int v, FastArgs* out)
{
    if (fp.kind < 2 || v < 0 || v >= fast_variant_count(fp)) return false;
    FastArgs A = fp.args;
    A.gen = 1;
    if (fp.kind == 3 && v >= gen1_count(fp)) {
        // edge-loaded segments: chunk depth {64,32} x block shape (waves, rows per wave)
        // {(8,1),(8,2),(16,1),(16,2)} x XCD remap x points per lane {widest, half}
        const int w = v - gen1_count(fp);
        A.gen = 2;
        A.xcd_remap = (w & 1) ? 0 : 1;
        static const int chunks[2] = {64, 32};
        A.chunk = chunks[(w >> 1) & 1];
        static const int shapes[4][2] = {{8, 1}, {8, 2}, {16, 1}, {16, 2}};
        A.W = shapes[(w >> 2) & 3][0];
        A.VY = shapes[(w >> 2) & 3][1];
        A.VX = (w >> 4) ? fp.VX / 2 : fp.VX;
        A.PF = 1;
        A.xbase = A.lo[0];
        A.nxseg = (int)ceil_div(A.hi[0] - A.lo[0], 64 * A.VX);
        A.nyblk = (int)ceil_div(A.hi[1] - A.lo[1], A.W * A.VY);
        A.gx = 0;
        A.nblocks = (int)(A.nxseg * A.nyblk * ceil_div(A.hi[2] - A.lo[2], A.chunk));
        *out = A;
        return true;
    }
    A.xcd_remap = (v & 1) ? 0 : 1;
    if (fp.kind == 2) {
        static const int chunks[4] = {32, 16, 64, 8};
        A.chunk = chunks[(v >> 1) & 3];
        A.PF = ((v >> 3) & 1) ? 4 : 2;
        A.VX = ((v >> 4) & 1) ? 2 * fp.VX : fp.VX;
        A.W = 4;
        A.VY = 1;
    } else {
        static const int chunks[2] = {64, 32};
        A.chunk = chunks[(v >> 1) & 1];
        A.PF = 1;
        static const int shapes[5][2] = {{8, 1}, {16, 1}, {8, 2}, {16, 2}, {8, 3}};
        const int sh = (v >> 2) % 5;
        A.W = shapes[sh][0];
        A.VY = shapes[sh][1];
        if (A.VY == 1 && 2 * (A.K / 2) >= 8) A.W = 16;   // a block must compute some rows
        A.VX = ((v >> 2) / 5) ? fp.VX / 2 : fp.VX;
    }
    {
        // segments of 64*VX points overlapping by 2*PADL (the x halo, whole lanes)
        const int H = A.K / 2;
        const int padl = (H + A.VX - 1) / A.VX * A.VX;
        A.xbase = A.lo[0] - padl;
        A.nxseg = (int)ceil_div(A.hi[0] - A.lo[0], 64 * A.VX - 2 * padl);
    }
    if (fp.kind == 2) {
        A.gx = (int)ceil_div(A.nxseg, 4);
        const int64_t nch = ceil_div(A.hi[1] - A.lo[1], A.chunk);
        A.nyblk = 0;
        A.nblocks = (int)(A.gx * nch);
    } else {
        // W waves stage W rows and compute the middle W - 2H
        A.nyblk = (int)ceil_div(A.hi[1] - A.lo[1], A.W * A.VY - 2 * (A.K / 2));
        const int64_t nzc = ceil_div(A.hi[2] - A.lo[2], A.chunk);
        A.gx = 0;
        A.nblocks = (int)(A.nxseg * A.nyblk * nzc);
    }
    *out = A;
    return true;
}

}  // namespace dop
