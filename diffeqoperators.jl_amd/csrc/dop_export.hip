// dop_export.hip -- concretization of a plan (its CSR matrix and affine part) and the banded
// direct solve A \ u built on it.
//
// The reference concretizes an operator on the host (concretization.jl: dense, sparse and
// banded forms of L, of Q and of L*Q) and solves A \ u with a sparse LU of L*Q_L
// (ghost_derivative_operator.jl:39-58).  Here both run on the device from the plan's own row
// programs: one thread per output row enumerates and merges its entries (dop_concrete.hpp),
// a scan turns row counts into row pointers, and the solve factors the band of the matrix
// with partial pivoting (the LAPACK gbtf2 / gbtrs algorithms) in device memory.
#include <hip/hip_runtime.h>

#include <algorithm>
#include <mutex>
#include <string>

#include "../../include/dop.h"
#include "dop_concrete.hpp"
#include "dop_objects.hpp"

namespace dop {
namespace {

constexpr int kExpBlock = 256;

// ---------------------------------------------------------------------------------------
// rows of the matrix
template <int CAP, bool FILL>
__global__ __launch_bounds__(kExpBlock) void k_csr(const PlanDev* __restrict__ Pp,
                                                   int64_t* __restrict__ rp,
                                                   int64_t* __restrict__ col,
                                                   double* __restrict__ val,
                                                   double* __restrict__ bvec,
                                                   unsigned long long* __restrict__ band)
{
    const PlanDev& P = *Pp;
    uint64_t lo = 0, hi = 0;
    for (int64_t p = (int64_t)blockIdx.x * kExpBlock + threadIdx.x; p < P.total;
         p += (int64_t)gridDim.x * kExpBlock) {
        int64_t idx[3];
        point_index(P, p, idx);
        RowList<CAP> L;
        row_entries(P, idx, L);
        if (!FILL) {
            int c = 0;
            for (int i = 0; i < L.n; ++i) {
                if (L.val[i] == 0.0) continue;
                ++c;
                const int64_t d = p - L.col[i];
                if (d > 0 && (uint64_t)d > lo) lo = (uint64_t)d;
                if (d < 0 && (uint64_t)(-d) > hi) hi = (uint64_t)(-d);
            }
            rp[p + 1] = c;
        } else {
            int64_t o = rp[p];
            for (int i = 0; i < L.n; ++i) {
                if (L.val[i] == 0.0) continue;
                col[o] = L.col[i];
                val[o] = L.val[i];
                ++o;
            }
            if (bvec) bvec[p] = L.b;
        }
    }
    if (!FILL) {
        for (int off = 32; off > 0; off >>= 1) {
            const uint64_t l2 = __shfl_xor(lo, off), h2 = __shfl_xor(hi, off);
            lo = l2 > lo ? l2 : lo;
            hi = h2 > hi ? h2 : hi;
        }
        if ((threadIdx.x & 63) == 0) {
            if (lo) atomicMax(&band[0], (unsigned long long)lo);
            if (hi) atomicMax(&band[1], (unsigned long long)hi);
        }
    }
}

// ---------------------------------------------------------------------------------------
// in-place inclusive scan of int64 counts: tiles of kScanTile, then the tile totals, then
// the tile offsets added back
constexpr int kScanPer = 16;
constexpr int kScanTile = kExpBlock * kScanPer;

__device__ inline int64_t block_incl_scan(int64_t v, int64_t* sh)
{
    const int t = threadIdx.x;
    sh[t] = v;
    __syncthreads();
    for (int off = 1; off < kExpBlock; off <<= 1) {
        const int64_t add = t >= off ? sh[t - off] : 0;
        __syncthreads();
        sh[t] += add;
        __syncthreads();
    }
    const int64_t r = sh[t];
    __syncthreads();
    return r;
}

__global__ __launch_bounds__(kExpBlock) void k_scan_tiles(int64_t* __restrict__ a, int64_t n,
                                                          int64_t* __restrict__ part)
{
    __shared__ int64_t sh[kExpBlock];
    const int64_t base = (int64_t)blockIdx.x * kScanTile + (int64_t)threadIdx.x * kScanPer;
    int64_t loc[kScanPer];
    int64_t s = 0;
#pragma unroll
    for (int i = 0; i < kScanPer; ++i) {
        const int64_t j = base + i;
        s += j < n ? a[j] : 0;
        loc[i] = s;
    }
    const int64_t incl = block_incl_scan(s, sh);
    const int64_t excl = incl - s;
#pragma unroll
    for (int i = 0; i < kScanPer; ++i) {
        const int64_t j = base + i;
        if (j < n) a[j] = loc[i] + excl;
    }
    if (threadIdx.x == kExpBlock - 1) part[blockIdx.x] = incl;
}

// tile totals -> exclusive tile offsets (one block, any number of tiles)
__global__ __launch_bounds__(kExpBlock) void k_scan_parts(int64_t* __restrict__ part, int64_t n)
{
    __shared__ int64_t sh[kExpBlock];
    int64_t carry = 0;
    for (int64_t c0 = 0; c0 < n; c0 += kExpBlock) {
        const int64_t j = c0 + threadIdx.x;
        const int64_t v = j < n ? part[j] : 0;
        const int64_t incl = block_incl_scan(v, sh);
        if (j < n) part[j] = carry + incl - v;
        // the last thread's inclusive value is the chunk total
        if (threadIdx.x == kExpBlock - 1) sh[0] = incl;
        __syncthreads();
        carry += sh[0];
        __syncthreads();
    }
}

__global__ __launch_bounds__(kExpBlock) void k_scan_add(int64_t* __restrict__ a, int64_t n,
                                                        const int64_t* __restrict__ part)
{
    const int64_t off = part[blockIdx.x];
    const int64_t base = (int64_t)blockIdx.x * kScanTile;
    for (int i = threadIdx.x; i < kScanTile; i += kExpBlock) {
        const int64_t j = base + i;
        if (j < n) a[j] += off;
    }
}

// ---------------------------------------------------------------------------------------
// band storage (LAPACK layout): A(i, j) at ab[(kv + i - j) + j * ldab], kv = kl + ku,
// ldab = 2 kl + ku + 1; rows [0, kl) of every column hold the fill-in of the pivoting
__global__ __launch_bounds__(kExpBlock) void k_band_fill(const int64_t* __restrict__ rp,
                                                         const int64_t* __restrict__ col,
                                                         const double* __restrict__ val,
                                                         int64_t n, int64_t kv, int64_t ldab,
                                                         double* __restrict__ ab)
{
    for (int64_t i = (int64_t)blockIdx.x * kExpBlock + threadIdx.x; i < n;
         i += (int64_t)gridDim.x * kExpBlock) {
        for (int64_t e = rp[i]; e < rp[i + 1]; ++e) {
            const int64_t c = col[e];
            ab[(kv + i - c) + c * ldab] = val[e];
        }
    }
}

// LU with partial pivoting of a band matrix (the gbtf2 algorithm), one block: for every
// column, the pivot search over the kl sub-diagonal entries, the row interchange over the
// columns the interchange reaches, the multipliers, and the rank-1 update of the trailing
// band.  info = first zero pivot (1-based) or 0.
template <int NT>
__global__ __launch_bounds__(NT) void k_gbtf2(int64_t n, int64_t kl, int64_t ku, int64_t ldab,
                                             double* __restrict__ ab, int64_t* __restrict__ ipiv,
                                             int64_t* __restrict__ info)
{
    __shared__ double s_abs[NT];
    __shared__ int64_t s_idx[NT];
    __shared__ int64_t s_ju, s_jp;
    __shared__ double s_piv;
    const int t = threadIdx.x;
    const int64_t kv = kl + ku;
    if (t == 0) {
        s_ju = 0;
        info[0] = 0;
    }
    __syncthreads();
    for (int64_t j = 0; j < n; ++j) {
        double* cj = ab + j * ldab;
        const int64_t km = kl < n - 1 - j ? kl : n - 1 - j;
        // pivot: first entry of largest magnitude in rows j .. j+km
        double best = -1.0;
        int64_t bi = 0;
        for (int64_t i = t; i <= km; i += NT) {
            const double a = fabs(cj[kv + i]);
            if (a > best) {
                best = a;
                bi = i;
            }
        }
        s_abs[t] = best;
        s_idx[t] = bi;
        __syncthreads();
        for (int off = NT / 2; off > 0; off >>= 1) {
            if (t < off) {
                const double o = s_abs[t + off];
                const int64_t oi = s_idx[t + off];
                if (o > s_abs[t] || (o == s_abs[t] && oi < s_idx[t])) {
                    s_abs[t] = o;
                    s_idx[t] = oi;
                }
            }
            __syncthreads();
        }
        if (t == 0) {
            const int64_t jp = s_idx[0];
            s_jp = jp;
            ipiv[j] = j + jp;
            const double piv = cj[kv + jp];
            s_piv = piv;
            if (piv != 0.0) {
                const int64_t reach = j + ku + jp < n - 1 ? j + ku + jp : n - 1;
                if (reach > s_ju) s_ju = reach;
            } else if (info[0] == 0) {
                info[0] = j + 1;
            }
        }
        __syncthreads();
        if (s_piv == 0.0) continue;
        const int64_t jp = s_jp, ju = s_ju;
        if (jp != 0) {
            for (int64_t c = j + t; c <= ju; c += NT) {
                double* cc = ab + c * ldab;
                const double x = cc[kv + j - c];
                cc[kv + j - c] = cc[kv + j + jp - c];
                cc[kv + j + jp - c] = x;
            }
            __syncthreads();
        }
        if (km > 0) {
            const double r = 1.0 / cj[kv];
            for (int64_t i = 1 + t; i <= km; i += NT) cj[kv + i] *= r;
            __syncthreads();
            const int64_t w = ju - j;
            for (int64_t q = t; q < km * w; q += NT) {
                const int64_t i = 1 + q % km;
                const int64_t c = j + 1 + q / km;
                double* cc = ab + c * ldab;
                cc[kv + j + i - c] -= cj[kv + i] * cc[kv + j - c];
            }
            __syncthreads();
        }
    }
}

// x = A^-1 (rhs - b) from the factors (the gbtrs algorithm, no transpose): forward with the
// interchanges and L, then backward with the band of U.  One block per right-hand side.
template <int NT>
__global__ __launch_bounds__(NT) void k_gbtrs(int64_t n, int64_t kl, int64_t ku, int64_t ldab,
                                             const double* __restrict__ ab,
                                             const int64_t* __restrict__ ipiv,
                                             const double* __restrict__ bvec,
                                             const double* __restrict__ rhs,
                                             double* __restrict__ x)
{
    const int t = threadIdx.x;
    const int64_t kv = kl + ku;
    const double* r = rhs + (int64_t)blockIdx.x * n;
    double* y = x + (int64_t)blockIdx.x * n;
    for (int64_t i = t; i < n; i += NT) y[i] = r[i] - (bvec ? bvec[i] : 0.0);
    __syncthreads();
    if (kl > 0) {
        for (int64_t j = 0; j < n - 1; ++j) {
            const int64_t lm = kl < n - 1 - j ? kl : n - 1 - j;
            const int64_t l = ipiv[j];
            if (l != j) {
                if (t == 0) {
                    const double s = y[l];
                    y[l] = y[j];
                    y[j] = s;
                }
                __syncthreads();
            }
            const double yj = y[j];
            const double* cj = ab + j * ldab;
            for (int64_t i = 1 + t; i <= lm; i += NT) y[j + i] -= cj[kv + i] * yj;
            __syncthreads();
        }
    }
    for (int64_t j = n - 1; j >= 0; --j) {
        const double* cj = ab + j * ldab;
        if (t == 0) y[j] /= cj[kv];
        __syncthreads();
        const double yj = y[j];
        const int64_t m = kv < j ? kv : j;
        for (int64_t i = 1 + t; i <= m; i += NT) y[j - i] -= cj[kv - i] * yj;
        __syncthreads();
    }
}

// ---------------------------------------------------------------------------------------
// narrow bands (kl + ku <= kWaveBand): one wave per system, the working set in LDS.
// Step j of the factorization touches only columns j .. j+kv, so those kv+1 columns live in
// an LDS ring (column c in slot c mod (kv+1)); column j is final after step j and is written
// back while column j+kv+1 takes its slot.  The device-memory reads (the entering columns,
// the multipliers, pivots and vector entries of the solve) run kPipe steps ahead in a
// register queue, and the lanes of the one wave order their LDS traffic with wave-scope
// fences only, so no step waits for device memory.
constexpr int kWaveBand = 62;
constexpr int kPipe = 8;

__device__ inline void wave_sync()
{
    __builtin_amdgcn_fence(__ATOMIC_RELEASE, "wavefront");
    __builtin_amdgcn_wave_barrier();
    __builtin_amdgcn_fence(__ATOMIC_ACQUIRE, "wavefront");
}

__device__ inline double readlane_d(double v, int l)
{
    const uint64_t b = __builtin_bit_cast(uint64_t, v);
    const uint32_t lo = (uint32_t)__builtin_amdgcn_readlane((int)(uint32_t)b, l);
    const uint32_t hi = (uint32_t)__builtin_amdgcn_readlane((int)(uint32_t)(b >> 32), l);
    return __builtin_bit_cast(double, ((uint64_t)hi << 32) | lo);
}

// lane i <- lane i+1 (DPP wave_shl:1); lane 63 reads 0
__device__ inline double wave_shl1(double v)
{
    const uint64_t b = __builtin_bit_cast(uint64_t, v);
    const int lo = __builtin_amdgcn_update_dpp(0, (int)(uint32_t)b, 0x130, 0xf, 0xf, false);
    const int hi = __builtin_amdgcn_update_dpp(0, (int)(uint32_t)(b >> 32), 0x130, 0xf, 0xf, false);
    return __builtin_bit_cast(double, ((uint64_t)(uint32_t)hi << 32) | (uint32_t)lo);
}

// (a, i) of the lane with the largest a, the smallest i among equals, over the aligned groups
// of 2 * span lanes (span = the largest power of two <= m, m >= 1): lanes 0 .. m are in lane
// 0's group
__device__ inline void wave_argmax_first(double& a, int& i, int64_t m)
{
    for (int off = 1; off <= m && off < 64; off <<= 1) {
        const double oa = __shfl_xor(a, off);
        const int oi = __shfl_xor(i, off);
        if (oa > a || (oa == a && oi < i)) {
            a = oa;
            i = oi;
        }
    }
}

__global__ __launch_bounds__(64) void k_gbtf2_wave(int64_t n, int64_t kl, int64_t ku, int64_t ldab,
                                                   double* __restrict__ ab,
                                                   int64_t* __restrict__ ipiv,
                                                   int64_t* __restrict__ info)
{
    extern __shared__ double win[];   // (kv + 1) columns of ldab
    const int lane = threadIdx.x;
    const int64_t kv = kl + ku, W = kv + 1;
    for (int64_t c = 0; c < W && c < n; ++c)
        for (int64_t r = lane; r < ldab; r += 64) win[c * ldab + r] = ab[c * ldab + r];
    // queue slot s holds column W + s (then W + s + kPipe, ...): entries lane and lane + 64
    double q0[kPipe], q1[kPipe];
#pragma unroll
    for (int s = 0; s < kPipe; ++s) {
        const int64_t c = W + s;
        q0[s] = (c < n && lane < ldab) ? ab[c * ldab + lane] : 0.0;
        q1[s] = (c < n && lane + 64 < ldab) ? ab[c * ldab + lane + 64] : 0.0;
    }
    __syncthreads();
    int64_t ju = 0, first_zero = 0;
    for (int64_t jb = 0; jb < n; jb += kPipe) {
#pragma unroll
        for (int s = 0; s < kPipe; ++s) {
            const int64_t j = jb + s;
            if (j >= n) break;
            double* cj = win + (j % W) * ldab;
            const int64_t km = kl < n - 1 - j ? kl : n - 1 - j;
                const double v = lane <= km ? cj[kv + lane] : 0.0;
            double a = lane <= km ? fabs(v) : -1.0;
            int jpi = lane <= km ? lane : 0;
            wave_argmax_first(a, jpi, km);
            const int jr = __builtin_amdgcn_readfirstlane(jpi);
            const int64_t jp = jr <= km ? jr : km;
            const double piv = readlane_d(v, (int)jp);
            if (lane == 0) ipiv[j] = j + jp;
            if (piv != 0.0) {
                const int64_t reach = j + ku + jp < n - 1 ? j + ku + jp : n - 1;
                ju = reach > ju ? reach : ju;
                if (jp != 0) {
                    for (int64_t c = j + lane; c <= ju; c += 64) {
                        double* cc = win + (c % W) * ldab;
                        const double t = cc[kv + j - c];
                        cc[kv + j - c] = cc[kv + j + jp - c];
                        cc[kv + j + jp - c] = t;
                    }
                    wave_sync();
                }
                if (km > 0) {
                    const double rcp = 1.0 / cj[kv];
                    if (lane >= 1 && lane <= km) cj[kv + lane] *= rcp;
                    wave_sync();
                    const int64_t w = ju - j;
                    for (int64_t q = lane; q < km * w; q += 64) {
                        const int64_t i = 1 + q % km;
                        const int64_t c = j + 1 + q / km;
                        double* cc = win + (c % W) * ldab;
                        cc[kv + j + i - c] -= cj[kv + i] * cc[kv + j - c];
                    }
                    wave_sync();
                }
            } else if (first_zero == 0) {
                first_zero = j + 1;
            }
            // retire column j; column j + W (queued kPipe steps ago) takes its slot
            for (int64_t r = lane; r < ldab; r += 64) ab[j * ldab + r] = cj[r];
            wave_sync();
            if (j + W < n) {
                if (lane < ldab) cj[lane] = q0[s];
                if (lane + 64 < ldab) cj[lane + 64] = q1[s];
            }
            const int64_t c = j + W + kPipe;
            q0[s] = (c < n && lane < ldab) ? ab[c * ldab + lane] : 0.0;
            q1[s] = (c < n && lane + 64 < ldab) ? ab[c * ldab + lane + 64] : 0.0;
            wave_sync();
        }
    }
    if (lane == 0) info[0] = first_zero;
}

// gbtrs for narrow bands: one wave per right-hand side.  The live part of the vector sits in
// registers, lane t holding y[j+t] (forward, t <= kl) or y[j-t] (backward, t <= kv); a step is
// two lane reads, one FMA per lane and a one-lane DPP shift of the window.
__global__ __launch_bounds__(64) void k_gbtrs_wave(int64_t n, int64_t kl, int64_t ku, int64_t ldab,
                                                   const double* __restrict__ ab,
                                                   const int64_t* __restrict__ ipiv,
                                                   const double* __restrict__ bvec,
                                                   const double* __restrict__ rhs,
                                                   double* __restrict__ x)
{
    const int lane = threadIdx.x;
    const int64_t kv = kl + ku;
    const double* r = rhs + (int64_t)blockIdx.x * n;
    double* y = x + (int64_t)blockIdx.x * n;
    auto src = [&](int64_t i) { return r[i] - (bvec ? bvec[i] : 0.0); };
    if (kl > 0) {
        // forward, step j: y[j] <-> y[ipiv[j]], y[j+i] -= L(j+i, j) y[j] (i = 1 .. lm).
        // queue slot s of step j: ipiv[j] (lane 0), L(j+lane, j), y[j+kl+1] (lane kl)
        double v = (lane <= kl && lane < n) ? src(lane) : 0.0;
        int64_t ql[kPipe];
        double qm[kPipe], qy[kPipe], qb[kPipe];
        // Every lane loads (clamped, lane-dependent addresses), so that the loads stay vector
        // loads the queue keeps in flight (a one-lane load of a uniform address becomes a
        // scalar load, waited for on the spot); the queue holds the raw loaded values, so that
        // no arithmetic on them waits before the step that uses them.
        auto fetch = [&](int64_t j, int s) {
            const bool ok = j < n - 1;
            const int64_t lm = ok ? (kl < n - 1 - j ? kl : n - 1 - j) : 0;
            const int64_t jj = ok ? j : 0;
            const int64_t ip = jj + lane < n ? jj + lane : n - 1;
            ql[s] = ipiv[ip];
            qm[s] = (lane >= 1 && lane <= lm) ? ab[jj * ldab + kv + lane] : 0.0;
            const int64_t iy = jj + 1 + lane < n ? jj + 1 + lane : n - 1;
            qy[s] = r[iy];
            qb[s] = bvec ? bvec[iy] : 0.0;
        };
#pragma unroll
        for (int s = 0; s < kPipe; ++s) fetch(s, s);
        for (int64_t jb = 0; jb < n - 1; jb += kPipe) {
#pragma unroll
            for (int s = 0; s < kPipe; ++s) {
                const int64_t j = jb + s;
                if (j >= n - 1) break;
                const int64_t lm = kl < n - 1 - j ? kl : n - 1 - j;
                const int d = __builtin_amdgcn_readlane((int)ql[s], 0) - (int)j;   // ipiv[j] - j
                const double m = qm[s];
                const double ynew = qy[s] - qb[s];
                fetch(j + kPipe, s);
                const double y0 = readlane_d(v, 0);
                double yj = y0;
                if (d != 0) {
                    yj = readlane_d(v, d);
                    v = lane == d ? y0 : v;
                }
                if (lane == 0) y[j] = yj;
                if (lane >= 1 && lane <= lm) v -= m * yj;
                v = wave_shl1(v);
                if (lane == kl) v = ynew;   // lane kl loaded y[j+1+kl]
            }
        }
        if (lane == 0) y[n - 1] = v;
    } else {
        for (int64_t i = lane; i < n; i += 64) y[i] = src(i);
    }
    // the backward sweep reads what the forward sweep stored
    __syncthreads();
    // backward, step j: y[j] /= U(j, j), y[j-i] -= U(j-i, j) y[j] (i = 1 .. min(kv, j)).
    // queue slot s of step j: U(j-lane, j), y[j-kv-1] (lane kv)
    double v = (lane <= kv && n - 1 - lane >= 0) ? y[n - 1 - lane] : 0.0;
    double qu[kPipe], qv[kPipe];
    auto fetchb = [&](int64_t j, int s) {
        const bool ok = j >= 0;
        qu[s] = (ok && lane <= kv && lane <= j) ? ab[j * ldab + kv - lane] : 0.0;
        const int64_t iy = j - 1 - lane > 0 ? j - 1 - lane : 0;   // lane kv: y[j-kv-1]
        qv[s] = y[iy];
    };
#pragma unroll
    for (int s = 0; s < kPipe; ++s) fetchb(n - 1 - s, s);
    for (int64_t jb = n - 1; jb >= 0; jb -= kPipe) {
#pragma unroll
        for (int s = 0; s < kPipe; ++s) {
            const int64_t j = jb - s;
            if (j < 0) break;
            const double u = qu[s];
            const double ynew = qv[s];
            fetchb(j - kPipe, s);
            const double yj = readlane_d(v, 0) / readlane_d(u, 0);
            const int64_t m = kv < j ? kv : j;
            if (lane == 0) y[j] = yj;
            if (lane >= 1 && lane <= m) v -= u * yj;
            v = wave_shl1(v);
            if (lane == kv) v = ynew;
        }
    }
}

int grid_for(int64_t rows)
{
    int64_t g = (rows + kExpBlock - 1) / kExpBlock;
    return (int)std::max<int64_t>(1, std::min<int64_t>(g, 1 << 20));
}

template <int CAP>
hipError_t launch_csr(const dop_plan* p, bool fill, int64_t* rp, int64_t* col, double* val,
                      double* b, unsigned long long* band, hipStream_t s)
{
    const int g = grid_for(p->d.host.total);
    if (fill)
        hipLaunchKernelGGL((k_csr<CAP, true>), dim3(g), dim3(kExpBlock), 0, s, p->dev, rp, col, val,
                           b, band);
    else
        hipLaunchKernelGGL((k_csr<CAP, false>), dim3(g), dim3(kExpBlock), 0, s, p->dev, rp, col,
                           val, b, band);
    return hipGetLastError();
}

hipError_t csr_pass(const dop_plan* p, bool fill, int64_t* rp, int64_t* col, double* val,
                    double* b, unsigned long long* band, hipStream_t s)
{
    const int bound = row_entry_bound(p->d.host);
    if (bound <= 16) return launch_csr<16>(p, fill, rp, col, val, b, band, s);
    if (bound <= 32) return launch_csr<32>(p, fill, rp, col, val, b, band, s);
    if (bound <= 64) return launch_csr<64>(p, fill, rp, col, val, b, band, s);
    return launch_csr<128>(p, fill, rp, col, val, b, band, s);
}

}  // namespace
}  // namespace dop

namespace {
int efail(int code, const std::string& m) { return dop::set_error(code, m); }

#define EXP_HIP(x)                                                                         \
    do {                                                                                   \
        hipError_t e_ = (x);                                                               \
        if (e_ != hipSuccess) {                                                            \
            rc = efail(DOP_EHIP, std::string(#x) + ": " + hipGetErrorString(e_));          \
            goto done;                                                                     \
        }                                                                                  \
    } while (0)

constexpr int kMaxRowEntries = 128;

// Checks shared by both row-pointer exports.
int csr_rows_check(const dop_plan* plan, const int64_t* row_ptr, const int64_t* nnz)
{
    if (!plan || !row_ptr || !nnz) return efail(DOP_EINVAL, "null plan, row_ptr or nnz");
    if (dop::row_entry_bound(plan->d.host) > kMaxRowEntries)
        return efail(DOP_EUNSUPPORTED, "rows of this plan have more than 128 entries");
    return DOP_OK;
}

// Row counts, their scan into row_ptr[1..n] and the bandwidths.  row_ptr[0] and band[0..1]
// must already be zero on the stream; part holds one partial sum per scan tile.  Synchronizes
// the stream.
int csr_rows_into(const dop_plan* plan, int64_t* row_ptr, int64_t* nnz, int64_t* kl, int64_t* ku,
                  int64_t* part, unsigned long long* band, hipStream_t s)
{
    const int64_t n = plan->d.host.total;
    const int64_t ntiles = (n + dop::kScanTile - 1) / dop::kScanTile;
    int rc = DOP_OK;
    unsigned long long hb[2] = {0, 0};
    EXP_HIP(dop::csr_pass(plan, false, row_ptr, nullptr, nullptr, nullptr, band, s));
    if (ntiles > 0) {
        hipLaunchKernelGGL(dop::k_scan_tiles, dim3((unsigned)ntiles), dim3(dop::kExpBlock), 0, s,
                           row_ptr + 1, n, part);
        hipLaunchKernelGGL(dop::k_scan_parts, dim3(1), dim3(dop::kExpBlock), 0, s, part, ntiles);
        hipLaunchKernelGGL(dop::k_scan_add, dim3((unsigned)ntiles), dim3(dop::kExpBlock), 0, s,
                           row_ptr + 1, n, part);
        EXP_HIP(hipGetLastError());
    }
    EXP_HIP(hipMemcpyAsync(nnz, row_ptr + n, sizeof(int64_t), hipMemcpyDeviceToHost, s));
    EXP_HIP(hipMemcpyAsync(hb, band, sizeof(hb), hipMemcpyDeviceToHost, s));
    EXP_HIP(hipStreamSynchronize(s));
    if (kl) *kl = (int64_t)hb[0];
    if (ku) *ku = (int64_t)hb[1];
done:
    return rc;
}

// ---- scratch of dop_plan_csr_rows_noalloc ----
// One static device array holds the bandwidth pair (first 256 bytes) and one partial sum per
// scan tile, so the export allocates nothing.  Each GPU has its own copy of the array at its
// own address: the address is taken once per device with hipGetSymbolAddress and only that
// address is handed to kernels and copies (the array's name is not a device address in host
// code).  Every call on a GPU shares that GPU's copy: overlapping calls on one GPU race.
constexpr int64_t kNoAllocTiles = 65536;   // rows <= kNoAllocTiles * kScanTile = 2^28
constexpr size_t kScratchBandBytes = 256;
constexpr size_t kScratchBytes = kScratchBandBytes + sizeof(int64_t) * kNoAllocTiles;
constexpr int kMaxDevices = 64;
std::mutex g_scratch_mu;
void* g_scratch_ptr[kMaxDevices] = {};

}  // namespace

static __device__ __align__(256) unsigned char scratch[kScratchBytes];

namespace {

int scratch_base(void** out)
{
    int dev = 0;
    const hipError_t e = hipGetDevice(&dev);
    if (e != hipSuccess) return efail(DOP_EHIP, std::string("hipGetDevice: ") + hipGetErrorString(e));
    if (dev < 0 || dev >= kMaxDevices) return efail(DOP_EINVAL, "device index beyond 64");
    std::lock_guard<std::mutex> lk(g_scratch_mu);
    if (!g_scratch_ptr[dev]) {
        void* p = nullptr;
        const hipError_t g = hipGetSymbolAddress(&p, HIP_SYMBOL(scratch));
        if (g != hipSuccess || !p)
            return efail(DOP_EHIP, std::string("hipGetSymbolAddress(scratch): ") +
                                       hipGetErrorString(g));
        g_scratch_ptr[dev] = p;
    }
    *out = g_scratch_ptr[dev];
    return DOP_OK;
}

__global__ void k_clear_csr_start(unsigned long long* __restrict__ band, int64_t* __restrict__ row_ptr)
{
    if (threadIdx.x < 2) band[threadIdx.x] = 0;
    if (threadIdx.x == 0) row_ptr[0] = 0;
}

}  // namespace

struct dop_solver {
    int64_t n = 0, kl = 0, ku = 0, ldab = 0;
    double* ab = nullptr;
    int64_t* ipiv = nullptr;
    double* b = nullptr;
    ~dop_solver()
    {
        if (ab) (void)hipFree(ab);
        if (ipiv) (void)hipFree(ipiv);
        if (b) (void)hipFree(b);
    }
};

extern "C" {

int dop_plan_csr_rows(const dop_plan* plan, int64_t* row_ptr, int64_t* nnz, int64_t* kl,
                      int64_t* ku, void* stream)
{
    int rc = csr_rows_check(plan, row_ptr, nnz);
    if (rc != DOP_OK) return rc;
    const hipStream_t s = (hipStream_t)stream;
    const int64_t ntiles = (plan->d.host.total + dop::kScanTile - 1) / dop::kScanTile;
    int64_t* part = nullptr;
    unsigned long long* band = nullptr;
    EXP_HIP(hipMalloc((void**)&part, sizeof(int64_t) * std::max<int64_t>(ntiles, 1)));
    EXP_HIP(hipMalloc((void**)&band, 2 * sizeof(unsigned long long)));
    EXP_HIP(hipMemsetAsync(band, 0, 2 * sizeof(unsigned long long), s));
    EXP_HIP(hipMemsetAsync(row_ptr, 0, sizeof(int64_t), s));
    rc = csr_rows_into(plan, row_ptr, nnz, kl, ku, part, band, s);
done:
    if (part) (void)hipFree(part);
    if (band) (void)hipFree(band);
    return rc;
}

int dop_plan_csr_rows_noalloc(const dop_plan* plan, int64_t* row_ptr, int64_t* nnz, int64_t* kl,
                              int64_t* ku, void* stream)
{
    int rc = csr_rows_check(plan, row_ptr, nnz);
    if (rc != DOP_OK) return rc;
    const int64_t ntiles = (plan->d.host.total + dop::kScanTile - 1) / dop::kScanTile;
    if (ntiles > kNoAllocTiles)
        return efail(DOP_EBUDGET, "dop_plan_csr_rows_noalloc: the plan has " +
                                      std::to_string(plan->d.host.total) +
                                      " rows; its static scratch holds the scan of at most " +
                                      std::to_string(kNoAllocTiles * dop::kScanTile) + " rows");
    void* base = nullptr;
    if ((rc = scratch_base(&base)) != DOP_OK) return rc;
    unsigned long long* band = (unsigned long long*)base;
    int64_t* part = (int64_t*)((char*)base + kScratchBandBytes);
    const hipStream_t s = (hipStream_t)stream;
    // the scratch keeps the previous call's contents: the bandwidth maxima and row_ptr[0]
    // start at zero (a kernel, not a memset, so that a captured stream records no memset node)
    hipLaunchKernelGGL(k_clear_csr_start, dim3(1), dim3(64), 0, s, band, row_ptr);
    EXP_HIP(hipGetLastError());
    rc = csr_rows_into(plan, row_ptr, nnz, kl, ku, part, band, s);
done:
    return rc;
}

int dop_plan_csr_fill(const dop_plan* plan, const int64_t* row_ptr, int64_t* col, double* val,
                      double* b, void* stream)
{
    if (!plan || !row_ptr || !col || !val) return efail(DOP_EINVAL, "null plan or buffer");
    if (dop::row_entry_bound(plan->d.host) > kMaxRowEntries)
        return efail(DOP_EUNSUPPORTED, "rows of this plan have more than 128 entries");
    const hipError_t e = dop::csr_pass(plan, true, (int64_t*)row_ptr, col, val, b, nullptr,
                                       (hipStream_t)stream);
    if (e != hipSuccess) return efail(DOP_EHIP, std::string("csr fill: ") + hipGetErrorString(e));
    return DOP_OK;
}

int dop_solver_create(const dop_plan* plan, void* stream, dop_solver** out)
{
    if (!plan || !out) return efail(DOP_EINVAL, "null plan or out");
    *out = nullptr;
    const dop::PlanDev& P = plan->d.host;
    if (!P.ghost)
        return efail(DOP_EINVAL, "A \\ u needs a GhostDerivativeOperator (L*Q) plan");
    if (plan->in_elems != plan->out_elems)
        return efail(DOP_EDIM, "A \\ u needs a square operator");
    const hipStream_t s = (hipStream_t)stream;
    const int64_t n = P.total;
    int rc = DOP_OK;
    int64_t nnz = 0, kl = 0, ku = 0, hinfo = 0;
    int64_t *rp = nullptr, *col = nullptr, *info = nullptr;
    double* val = nullptr;
    dop_solver* S = new dop_solver();
    EXP_HIP(hipMalloc((void**)&rp, sizeof(int64_t) * (n + 1)));
    if ((rc = dop_plan_csr_rows(plan, rp, &nnz, &kl, &ku, stream)) != DOP_OK) goto done;
    {
        const int64_t ldab = 2 * kl + ku + 1;
        // the band with its fill-in rows must fit: 2^33 bytes
        if (ldab > ((int64_t)1 << 30) / std::max<int64_t>(n, 1)) {
            rc = efail(DOP_EUNSUPPORTED, "band of A too wide for the banded solve (" +
                                             std::to_string(kl) + " sub- and " +
                                             std::to_string(ku) + " super-diagonals)");
            goto done;
        }
        S->n = n;
        S->kl = kl;
        S->ku = ku;
        S->ldab = ldab;
    }
    EXP_HIP(hipMalloc((void**)&col, sizeof(int64_t) * std::max<int64_t>(nnz, 1)));
    EXP_HIP(hipMalloc((void**)&val, sizeof(double) * std::max<int64_t>(nnz, 1)));
    EXP_HIP(hipMalloc((void**)&S->b, sizeof(double) * n));
    EXP_HIP(hipMalloc((void**)&S->ab, sizeof(double) * S->ldab * n));
    EXP_HIP(hipMalloc((void**)&S->ipiv, sizeof(int64_t) * n));
    EXP_HIP(hipMalloc((void**)&info, sizeof(int64_t)));
    if ((rc = dop_plan_csr_fill(plan, rp, col, val, S->b, stream)) != DOP_OK) goto done;
    EXP_HIP(hipMemsetAsync(S->ab, 0, sizeof(double) * S->ldab * n, s));
    hipLaunchKernelGGL(dop::k_band_fill, dim3(dop::grid_for(n)), dim3(dop::kExpBlock), 0, s, rp,
                       col, val, n, kl + ku, S->ldab, S->ab);
    if (kl + ku <= dop::kWaveBand)
        hipLaunchKernelGGL(dop::k_gbtf2_wave, dim3(1), dim3(64),
                           (unsigned)(sizeof(double) * (kl + ku + 1) * S->ldab), s, n, kl, ku,
                           S->ldab, S->ab, S->ipiv, info);
    else if ((kl + 1) * (kl + ku + 1) >= 4096)
        hipLaunchKernelGGL(dop::k_gbtf2<1024>, dim3(1), dim3(1024), 0, s, n, kl, ku, S->ldab,
                           S->ab, S->ipiv, info);
    else
        hipLaunchKernelGGL(dop::k_gbtf2<256>, dim3(1), dim3(256), 0, s, n, kl, ku, S->ldab, S->ab,
                           S->ipiv, info);
    EXP_HIP(hipGetLastError());
    EXP_HIP(hipMemcpyAsync(&hinfo, info, sizeof(int64_t), hipMemcpyDeviceToHost, s));
    EXP_HIP(hipStreamSynchronize(s));
    if (hinfo != 0)
        rc = efail(DOP_ESINGULAR, "A is singular: zero pivot in column " + std::to_string(hinfo));
done:
    if (rp) (void)hipFree(rp);
    if (col) (void)hipFree(col);
    if (val) (void)hipFree(val);
    if (info) (void)hipFree(info);
    if (rc != DOP_OK) {
        delete S;
        return rc;
    }
    *out = S;
    return DOP_OK;
}

int dop_solver_band(const dop_solver* S, int64_t* n, int64_t* kl, int64_t* ku)
{
    if (!S) return efail(DOP_EINVAL, "null solver");
    if (n) *n = S->n;
    if (kl) *kl = S->kl;
    if (ku) *ku = S->ku;
    return DOP_OK;
}

int dop_solver_solve(const dop_solver* S, void* x, const void* rhs, int nrhs, void* stream)
{
    if (!S || !x || !rhs || nrhs < 0) return efail(DOP_EINVAL, "null solver or buffer");
    if (x == rhs) return efail(DOP_EINVAL, "x must not alias rhs");
    if (nrhs == 0) return DOP_OK;
    const hipStream_t s = (hipStream_t)stream;
    if (S->kl + S->ku <= dop::kWaveBand)
        hipLaunchKernelGGL(dop::k_gbtrs_wave, dim3(nrhs), dim3(64), 0, s, S->n, S->kl, S->ku,
                           S->ldab, S->ab, S->ipiv, S->b, (const double*)rhs, (double*)x);
    else if (S->kl + S->ku + 1 >= 1024)
        hipLaunchKernelGGL(dop::k_gbtrs<1024>, dim3(nrhs), dim3(1024), 0, s, S->n, S->kl, S->ku,
                           S->ldab, S->ab, S->ipiv, S->b, (const double*)rhs, (double*)x);
    else
        hipLaunchKernelGGL(dop::k_gbtrs<256>, dim3(nrhs), dim3(256), 0, s, S->n, S->kl, S->ku,
                           S->ldab, S->ab, S->ipiv, S->b, (const double*)rhs, (double*)x);
    const hipError_t e = hipGetLastError();
    if (e != hipSuccess) return efail(DOP_EHIP, std::string("solve: ") + hipGetErrorString(e));
    return DOP_OK;
}

void dop_solver_destroy(dop_solver* S) { delete S; }

}  // extern "C"
