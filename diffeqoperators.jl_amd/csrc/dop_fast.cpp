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

// Whether a shell strip that touches x boundary rows finds every x tap and ghost it reads in
// the kStripStage values at its end of the line (left rows: [0, kStripStage), right rows:
// [n - kStripStage, n)): strips are kStripPts aligned, so their points lie below
// roundup(rlo) + kStripPts or at or above rhi - kStripPts.
bool strips_stage_ok(const PlanDev& P, const AxisTables& T)
{
    const AxisDev& X = P.ax[0];
    const int n = (int)P.n[0];
    if (!P.ghost || !X.active || n < 2 * kStripStage || n % kStripPts != 0 || X.periodic) return false;
    for (size_t l = 0; l < T.lbc.size(); l += kLineBC)
        if (T.lbc[l] < 0.0) return false;   // a periodic line reads the far end
    const int lend = std::min(n, (X.rlo + kStripPts - 1) / kStripPts * kStripPts + kStripPts);
    const int rbeg = std::max(0, X.rhi - kStripPts);
    auto within = [&](int r, int j) {
        return r < lend ? (j >= 0 && j < kStripStage) : (j >= n - kStripStage && j < n);
    };
    for (int r = 0; r < n; ++r) {
        if (r >= lend && r < rbeg) continue;
        if (r >= X.rlo && r < X.rhi) {
            for (int k = 0; k < X.K; ++k)
                if (!within(r, r + X.dmin + k)) return false;
            continue;
        }
        const int b = r < X.rlo ? r : X.rlo + (r - X.rhi);
        for (int k = 0; k < T.bn[b]; ++k) {
            const int J = T.bJ0[b] + k;
            if (J == 0) {
                for (int m = 0; m < X.nl; ++m)
                    if (!within(r, m)) return false;
            } else if (J == n + 1) {
                for (int m = 0; m < X.nr; ++m)
                    if (!within(r, n - X.nr + m)) return false;
            } else if (!within(r, J - 1)) {
                return false;
            }
        }
    }
    return true;
}

}  // namespace

namespace {

// The 3-D shell (the domain minus the box [lo, hi)) as work segments of the strip kernel; false
// if a segment would not fit its 32-bit counts or the tables.
bool build_strips(const PlanDev& P, const HostPlan& H, const int32_t* lo, const int32_t* hi,
                  StripArgs* out)
{
    StripArgs& S = *out;
    S = StripArgs{};
    const int32_t n[3] = {(int32_t)P.n[0], (int32_t)P.n[1], (int32_t)P.n[2]};
    const AxisDev& X0 = P.ax[0];
    // every item starts a 16-byte group of u and du (inside the array) when n0, the row and
    // plane strides, the transverse shifts and the x edges of the box are multiples of 4
    bool al = n[0] % kStripPts == 0 && lo[0] % kStripPts == 0 && hi[0] % kStripPts == 0;
    for (int a = 1; a < 3; ++a) al = al && P.ssrc[a] % kStripPts == 0 && P.sdu[a] % kStripPts == 0;
    for (int a = 0; a < 3; ++a) al = al && P.tshift[a] % kStripPts == 0;
    S.vec_ok = al ? 1 : 0;
    S.stage_ok = strips_stage_ok(P, H.axis[0]) ? 1 : 0;
    // x strips of four regular rows
    const int32_t ax = (X0.rlo + kStripPts - 1) / kStripPts * kStripPts;
    const int32_t bx = X0.rhi / kStripPts * kStripPts;
    const bool faces = S.vec_ok && n[0] >= kStripStage && ax < bx;
    const bool face[3] = {faces && S.stage_ok && !X0.lbc, faces && !P.ax[1].lbc, faces && !P.ax[2].lbc};

    int nbox = 0;
    auto add_seg = [&](int kind, int r0, int nr, int xend, int64_t items) {
        if (nr <= 0 || items <= 0) return true;
        if (S.nseg >= kMaxShellSegs || items >= ((int64_t)1 << 31)) return false;
        ShellSeg& G = S.seg[S.nseg++];
        G = ShellSeg{kind, r0, nr, xend, (int32_t)items, S.nblocks};
        const int64_t nb = (int64_t)nr * ceil_div(items, kStripBlock);
        if ((int64_t)S.nblocks + nb >= ((int64_t)1 << 31)) return false;
        S.nblocks += (int32_t)nb;
        return true;
    };
    auto add_box = [&](int32_t x0, int32_t x1, int32_t y0, int32_t y1, int32_t z0, int32_t z1) {
        if (x1 <= x0 || y1 <= y0 || z1 <= z0) return true;
        if (nbox >= kMaxStripBoxes) return false;
        const int b = nbox++;
        S.org[b][0] = x0; S.org[b][1] = y0; S.org[b][2] = z0;
        S.ext[b][0] = x1 - x0; S.ext[b][1] = y1 - y0; S.ext[b][2] = z1 - z0;
        S.nsx[b] = (int32_t)ceil_div(x1 - x0, kStripPts);
        S.zfirst[b] = S.ext[b][2] <= S.ext[b][1] ? 1 : 0;
        return add_seg(3, b, 1, 0, (int64_t)S.nsx[b] * S.ext[b][1] * S.ext[b][2]);
    };
    bool ok = true;
    // x sides: y and z inside the box
    S.flo[0][0] = lo[1]; S.fn[0][0] = hi[1] - lo[1];
    S.flo[0][1] = lo[2]; S.fn[0][1] = hi[2] - lo[2];
    // kind 4 (x faces in tiles of kXfTile^2 lines, one block each): the y and z windows must
    // reach no further than the kXfHalo border lines
    bool reach = true;
    for (int a = 1; a < 3; ++a)
        reach = reach && -P.ax[a].dmin <= kXfHalo && P.ax[a].dmin + P.ax[a].K - 1 <= kXfHalo;
    const bool tiles = face[0] && reach && kXfTile * kXfTile == kStripBlock;
    const int64_t xf_tiles = ceil_div(S.fn[0][0], kXfTile) * ceil_div(S.fn[0][1], kXfTile);
    for (int side = 0; side < 2; ++side) {
        const int32_t x0 = side ? hi[0] : 0, x1 = side ? n[0] : lo[0];
        if (tiles) {
            for (int32_t xs = x0; xs < x1; xs += kStripPts)
                ok = ok && add_seg(4, xs, 1, x1, xf_tiles * kStripBlock);
            S.nfirst = S.nblocks;
        } else if (face[0]) {
            ok = ok && add_seg(0, x0, (int32_t)ceil_div(x1 - x0, kStripPts), x1,
                               (int64_t)S.fn[0][0] * S.fn[0][1]);
        } else {
            ok = ok && add_box(x0, x1, lo[1], hi[1], lo[2], hi[2]);
        }
    }
    // z sides: all x and y
    S.flo[2][0] = ax; S.fn[2][0] = (bx - ax) / kStripPts;
    S.flo[2][1] = P.ax[1].rlo; S.fn[2][1] = P.ax[1].rhi - P.ax[1].rlo;
    for (int side = 0; side < 2; ++side) {
        const int32_t z0 = side ? hi[2] : 0, z1 = side ? n[2] : lo[2];
        if (face[2] && S.fn[2][1] > 0) {
            ok = ok && add_seg(2, z0, z1 - z0, 0, (int64_t)S.fn[2][0] * S.fn[2][1]);
            ok = ok && add_box(0, ax, 0, n[1], z0, z1) && add_box(bx, n[0], 0, n[1], z0, z1);
            ok = ok && add_box(ax, bx, 0, P.ax[1].rlo, z0, z1) && add_box(ax, bx, P.ax[1].rhi, n[1], z0, z1);
        } else {
            ok = ok && add_box(0, n[0], 0, n[1], z0, z1);
        }
    }
    // y sides: all x, z inside the box
    S.flo[1][0] = ax; S.fn[1][0] = (bx - ax) / kStripPts;
    S.flo[1][1] = lo[2]; S.fn[1][1] = hi[2] - lo[2];
    for (int side = 0; side < 2; ++side) {
        const int32_t y0 = side ? hi[1] : 0, y1 = side ? n[1] : lo[1];
        if (face[1]) {
            ok = ok && add_seg(1, y0, y1 - y0, 0, (int64_t)S.fn[1][0] * S.fn[1][1]);
            ok = ok && add_box(0, ax, y0, y1, lo[2], hi[2]) && add_box(bx, n[0], y0, y1, lo[2], hi[2]);
        } else {
            ok = ok && add_box(0, n[0], y0, y1, lo[2], hi[2]);
        }
    }
    return ok;
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
    // the Dot2 compensated mode: batched tiles only (single plans take the generic kernel);
    // f32 values summed in f64 (DOP_ACC_F32_WIDE) run the single-plan tiles too
    fp->kahan = (P.accum == kAccF32Kahan) ? 1 : 0;
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
    fp->f64 = P.storage == kF64;
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
    if (P.ndim == 3 && !build_strips(P, H, A.lo, A.hi, &fp->strips)) return;
    (void)shell_boxes(P, fp->nat_lo, fp->nat_hi, &fp->nat_shell);
    fp->kind = P.ndim;
}

// Candidates (autotuner).  2-D: chunk height {32,16,64,8} x prefetch {2,4} ({4,8}, FP64 {8,16} edge-loaded) x XCD remap x
// points per lane {widest 16-byte, twice that} x marches {all downwards, alternating} x
// segments {overlapping, edge-loaded (when the x half-width is at most the points per lane)}; 3-D:
// chunk depth {64,32} x block shape (waves, rows per wave) {(8,1), (16,1), (8,2), (16,2),
// (8,3)} x XCD remap x points per lane {widest, half}, then the edge-loaded forms.  A block
// stages waves * rows y rows and computes all but the 2H halo rows.
namespace {
bool wide_ok(const FastPlan& fp) { return fp.VX > 1 && fp.args.n[0] % (2 * fp.VX) == 0; }
// 2-D edge-loaded segments: one edge group per side at the narrowest points per lane
bool edge2_ok(const FastPlan& fp) { return fp.args.K / 2 <= fp.VX; }
int base2_count(const FastPlan& fp) { return wide_ok(fp) ? 64 : 32; }

int gen1_count(const FastPlan& fp) { return fp.VX > 1 ? 40 : 20; }
int gen2_count(const FastPlan& fp) { return fp.VX > 1 ? 32 : 16; }
}  // namespace

int fast_variant_count(const FastPlan& fp)
{
    if (fp.kind == 1) return 1;
    if (fp.kind == 2) return base2_count(fp) * (edge2_ok(fp) ? 2 : 1);
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
        // {(8,1),(8,2),(16,1),(8,3)} x XCD remap x points per lane {widest, half}
        const int w = v - gen1_count(fp);
        A.gen = 2;
        A.xcd_remap = (w & 1) ? 0 : 1;
        static const int chunks[2] = {64, 32};
        A.chunk = chunks[(w >> 1) & 1];
        static const int shapes[4][2] = {{8, 1}, {8, 2}, {16, 1}, {8, 3}};
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
    if (fp.kind == 2 && v >= base2_count(fp)) {
        A.gen = 2;
        v -= base2_count(fp);
    }
    A.xcd_remap = (v & 1) ? 0 : 1;
    if (fp.kind == 2) {
        static const int chunks[4] = {32, 16, 64, 8};
        A.chunk = chunks[(v >> 1) & 3];
        // edge-loaded rings carry H edge values per row, so they afford deeper prefetch (FP64
        // rows of 2 points per lane: 8 or 16 rows; 4 or 8 otherwise)
        const bool f64 = fp.f64 != 0;
        A.PF = A.gen == 2 ? (((v >> 3) & 1) ? (f64 ? 16 : 8) : (f64 ? 8 : 4))
                          : (((v >> 3) & 1) ? 4 : 2);
        // bit 4: twice the points per lane (when rows allow), next bit: alternating marches
        const bool wide = wide_ok(fp);
        A.VX = (wide && ((v >> 4) & 1)) ? 2 * fp.VX : fp.VX;
        A.alt = (v >> (wide ? 5 : 4)) & 1;
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
    if (A.gen == 2) {
        // segments of 64*VX points at multiples of 64*VX, not overlapping
        const int seg = 64 * A.VX;
        A.xbase = A.lo[0] / seg * seg;
        A.nxseg = (int)ceil_div(A.hi[0] - A.xbase, seg);
    } else {
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
