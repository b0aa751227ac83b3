// dop_fast.hpp -- launch descriptions of the tiled interior kernels (host-built, passed by
// value to the kernels) and of the generic "shell" launch that covers the rest.
#pragma once
#include <cstdint>
#include <vector>

#include "dop_build.hpp"
#include "dop_plan.hpp"

namespace dop {

constexpr int kMaxBoxes = 6;

// Points outside the interior box, as up to 6 axis-aligned boxes.
struct ShellArgs {
    int32_t nbox;
    int32_t pad_;
    int64_t org[kMaxBoxes][3];
    int64_t ext[kMaxBoxes][3];
    int64_t start[kMaxBoxes + 1];   // prefix counts; start[nbox] = total
};

// The shell as strips of up to kStripPts consecutive x points of one row (the 3-D shell
// kernel's work items): per box, its origin, extent, strips along x and prefix counts.
constexpr int kStripPts = 4;
struct StripArgs {
    int32_t nbox;
    int32_t vec_ok;                 // n0 % 4 == 0: strips starting at x % 4 == 0 are aligned
    int64_t org[kMaxBoxes][3];
    int64_t ext[kMaxBoxes][3];
    int64_t nsx[kMaxBoxes];         // strips along x
    int64_t start[kMaxBoxes + 1];   // prefix counts of strips; start[nbox] = total
    int32_t xcol[kMaxBoxes];        // 1: the box holds x boundary columns (stage the line end)
};

// The interior box and the per-axis weights of the tiled kernels.  Every active axis uses
// one symmetric odd window of K taps (narrower operator windows are padded with zero
// weights); rows/planes march with a register ring.
struct FastArgs {
    int64_t n[3];
    int64_t s1, s2;        // row / plane strides
    int32_t K;             // window taps on every axis
    int32_t VX;            // points per lane along x (16-, 8- or 4-byte lane loads)
    int32_t PF;            // rows / planes prefetched ahead in the register ring
    int32_t W;             // 3-D: waves per block
    int32_t VY;            // 3-D: rows staged per wave (a block stages W*VY rows)
    int32_t lo[3], hi[3];  // interior box [lo, hi) per axis
    int32_t xbase;         // x of the first segment's first point (VX-aligned; 3-D: may be
                           // negative, the segment's first PADL points are x halo)
    int32_t nxseg;         // x segments (64*VX points; 3-D: overlapping by 2*PADL) per row
    int32_t chunk;         // rows (2-D) / planes (3-D) per block
    int32_t nyblk;         // 3-D: y blocks of W rows
    int32_t gx;            // 2-D: blocks across x (4 segments each)
    int32_t nblocks;
    int32_t xcd_remap;     // keep neighbouring tiles on one XCD
    int32_t gen;           // 3-D: 1 overlapping segments (x halo from the segment's own end
                           // lanes), 2 edge-loaded segments (k_tiled3e)
    int32_t uniform[3];    // weights identical for every row of the axis
    const double* w64[3];  // [uniform ? 1 : n][K]
    const float* w32[3];
};

struct FastPlan {
    int kind = 0;          // 0: generic only, 1: 1-D window tables only (batched tiles),
                           // 2: 2-D tiled, 3: 3-D tiled
    int VX = 1;            // widest points-per-lane the row alignment allows
    int kahan = 0;         // compensated f32 accumulation: batched tiles only
    FastArgs args{};
    ShellArgs shell{};
    int64_t shell_total = 0;
    StripArgs strips{};    // the same shell as strips (3-D)
    // the box before the x edges are aligned for the single-plan kernels' vector stores, and
    // its shell (the batched tiles store single points and use these)
    int32_t nat_lo[3] = {0, 0, 0}, nat_hi[3] = {1, 1, 1};
    ShellArgs nat_shell{};
    std::vector<double> w64[3];   // host copies before upload
    std::vector<float> w32[3];
};

// Decides whether the plan has a tiled path and builds its description (host tables in
// fp->w64/w32; device pointers are filled by the caller after upload).
void build_fast(const HostPlan& H, FastPlan* fp);

// Launch-time geometry for variant `v` (autotuner candidates); returns false if `v` is
// not a valid variant.
bool fast_variant(const FastPlan& fp, int v, FastArgs* out);
int fast_variant_count(const FastPlan& fp);

}  // namespace dop
