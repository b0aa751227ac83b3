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

// The 3-D shell as work items of up to kStripPts consecutive x points of one row, in
// segments of whole blocks of the shell kernel:
//  - face segments (kind a = 0, 1, 2): the shell rows of axis a on one side of the box, over
//    the part of the face whose other two axes are regular rows (kind 1, 2: x strips of four
//    regular rows; kind 0: the strips at one x end of the lines).  Within a block every item
//    has the same row along a, so the row's program is uniform; the rows of a segment take
//    turns block by block (items sharing their taps run together);
//  - general segments (kind 3): one box of the rest (edges, ends of lines, per-line BCs, ...).
constexpr int kStripPts = 4;
constexpr int kStripBlock = 256;    // items per block of the shell kernel
constexpr int kMaxStripBoxes = 16;
constexpr int kMaxShellSegs = 32;
struct ShellSeg {
    int32_t kind;    // 0, 1, 2: face rows of that axis; 3: general strips of box r0; 4: the x
                     // face strip starting at x0 = r0 in tiles of kXfTile^2 lines (one block each)
    int32_t r0;      // face: first row (kind 0: first strip start x0); general: box
    int32_t nr;      // face: rows r0 .. r0 + nr - 1 (kind 0: strip starts 4 apart); general: 1
    int32_t xend;    // kind 0: end of the x range (its last strip may be narrower)
    int32_t items;   // work items per row (general: strips of the box)
    int32_t first;   // first block (in the order the XCDs take them)
};
struct StripArgs {
    int32_t nseg;
    int32_t nblocks;                 // blocks of the launch
    int32_t nfirst;                  // blocks of the leading kind-4 segments (taken in hardware
                                     // order; the rest XCD-contiguous)
    int32_t vec_ok;                  // every item starts an aligned group of 4 values of u and du
    int32_t stage_ok;                // every x tap and ghost of a strip touching x boundary rows
                                     // lies in the kStripStage values at its end of the line
    ShellSeg seg[kMaxShellSegs];
    // general boxes: origin, extent, strips along x, order (zfirst: x, z, y -- a box thin
    // along z -- instead of x, y, z)
    int32_t org[kMaxStripBoxes][3];
    int32_t ext[kMaxStripBoxes][3];
    int32_t nsx[kMaxStripBoxes];
    int32_t zfirst[kMaxStripBoxes];
    // faces of kind a: the item's coordinates (u fastest, v): kind 2 (x strip, y), kind 1
    // (x strip, z), kind 0 (y, z); flo = first x (strips) or row, fn = count
    int32_t flo[3][2];
    int32_t fn[3][2];
};
constexpr int kXfTile = 16;     // kind 4: lines per tile side (y and z); kXfTile^2 = block size
constexpr int kXfHalo = 4;      // kind 4: border lines staged on each side (>= the y and z reach)
constexpr int kStripStage = 16;

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
    int32_t gen;           // 1 overlapping segments (x halo from the segment's own end lanes),
                           // 2 edge-loaded segments (3-D: k_tiled3e; 2-D: k_tiled2d's EDG form)
    int32_t alt;           // 2-D: odd row chunks march upwards (towards their neighbours)
    int32_t uniform[3];    // weights identical for every row of the axis
    const double* w64[3];  // [uniform ? 1 : n][K]
    const float* w32[3];
};

struct FastPlan {
    int kind = 0;          // 0: generic only, 1: 1-D window tables only (batched tiles),
                           // 2: 2-D tiled, 3: 3-D tiled
    int VX = 1;            // widest points-per-lane the row alignment allows
    int f64 = 0;           // FP64 storage
    int kahan = 0;         // Dot2-compensated f32 accumulation: batched tiles only
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
