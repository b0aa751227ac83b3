// dop_plan.hpp -- device-side plan layout shared by the host builder and the kernels.
//
// An apply evaluates, for every output point p and every axis a that carries operators,
//     du[p] += sum_k W_a(row)[k] * P_a(line(p), J0(row) + k)
// where P_a is the padded line along a (padded index J: 0 = left ghost, 1..n = u,
// n+1 = right ghost) and the row program (J0, W) is the sum of the reference's per-row
// arithmetic over all operators on that axis.  Rows are split into
//   regular rows [rlo, rhi): window u[r+dmin .. r+dmin+K-1], no ghost, weights W[r][K]
//                            (a single row when every regular row has the same weights)
//   boundary rows (the rest): padded start bJ0, bn taps, weights bw[KB] (FP64 always)
#pragma once
#include <cstdint>

#ifndef DOP_HD
#if defined(__HIPCC__)
#define DOP_HD __host__ __device__
#else
#define DOP_HD
#endif
#endif

namespace dop {

// Element i of a plan table.  Every plan (PlanDev, its AxisDev and the tables they point to)
// lives in device global memory, so device code reads it with global rather than flat loads
// (flat loads also count against the LDS wait counter).
template <typename V>
DOP_HD inline V tab(const V* p, int64_t i)
{
#if defined(__HIP_DEVICE_COMPILE__)
    return ((const __attribute__((address_space(1))) V*)p)[i];
#else
    return p[i];
#endif
}

constexpr int kMaxBCTaps = 8;
constexpr int kMaxDim = 3;

enum Storage : int32_t { kF64 = 0, kF32 = 1, kBF16 = 2 };
enum Accum : int32_t { kAccF64 = 1, kAccF32 = 2, kAccF32Kahan = 3, kAccF32Wide = 4 };

struct AxisDev {
    int32_t n;        // extent along the axis
    int32_t active;   // carries at least one operator
    int32_t K;        // regular window width
    int32_t dmin;     // regular window start relative to the row (u index r + dmin)
    int32_t rlo, rhi; // regular rows [rlo, rhi)
    int32_t uniform;  // regular weights identical for all regular rows
    int32_t KB;       // max taps of a boundary row
    int32_t nb;       // number of boundary rows = rlo + (n - rhi)
    int32_t nl, nr;   // ghost taps
    int32_t periodic; // 0: affine ghosts; 1: ghost_l = u[n-1], ghost_r = u[0];
                      // 2: ghost_l = u[0], ghost_r = u[n-1]
    double al[kMaxBCTaps];
    double ar[kMaxBCTaps];
    double bl, br;
    const double* w64;    // uniform: [K]; else tap-major [K][n] (row r, tap k at k * n + r)
    const float* w32;     // same, rounded to f32
    const int32_t* bJ0;   // [nb] padded start
    const int32_t* bn;    // [nb] taps
    const double* bw;     // [nb][KB]
    // per-line BCs (nullptr: al/bl/ar/br above): line l at lbc + l * kLineBC, laid out as
    // nl, nr, al[8], bl, ar[8], br (nl = nr = kLinePeriodic: the line wraps, ghosts u[n-1],
    // u[0]); line of a point = sum over the other axes b of idx[b] * lstride[b]
    const double* lbc;
    int64_t lstride[kMaxDim];
};

constexpr int kLineBC = 2 * kMaxBCTaps + 4;
// nl (and nr) of a per-line entry whose line wraps (PeriodicBC on that line)
constexpr double kLinePeriodic = -1.0;

struct PlanDev {
    int32_t ndim;
    int32_t storage;
    int32_t accum;
    int32_t ghost;        // 1: GHOST input (joff = -1), 0: PADDED input (joff = 0)
    int64_t n[kMaxDim];   // output extents (1 beyond ndim)
    int64_t sdu[kMaxDim]; // output strides
    int64_t ssrc[kMaxDim];    // source strides (u or M)
    int64_t tshift[kMaxDim];  // per axis a: sum_{b != a} shift_ab * ssrc[b]
    int64_t total;        // number of output points
    AxisDev ax[kMaxDim];
};

}  // namespace dop
