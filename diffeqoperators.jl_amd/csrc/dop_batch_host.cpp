// dop_batch_host.cpp -- host side of the persistent batched launch: per-problem tile
// geometry and the work list.
#include <algorithm>
#include <tuple>

#include "dop_batch.hpp"

namespace dop {

namespace {
int64_t ceil_div(int64_t a, int64_t b) { return (a + b - 1) / b; }
}  // namespace

void batch_geometry(const DevicePlan& p, int chunk, BProb* B)
{
    const PlanDev& P = p.host;
    const FastPlan& f = p.fast;
    B->plan = p.dev;
    B->plan_accum = P.accum;
    B->chunk = chunk;
    // 2-D and 1-D tiles exist for windows up to 11 taps (the batched tile kernels'
    // instantiations); they cover the natural box (one point per lane, no alignment)
    B->tiled = (f.args.K <= kMaxBatchTileK) ? (f.kind == 2 ? 1 : f.kind == 1 ? 2 : 0) : 0;
    if (B->tiled) {
        const FastArgs& A = f.args;
        const int H = A.K / 2;
        B->K = A.K;
        for (int a = 0; a < 2; ++a) {
            B->lo[a] = f.nat_lo[a];
            B->hi[a] = f.nat_hi[a];
            B->uniform[a] = A.uniform[a];
            B->w64[a] = A.w64[a];
            B->w32[a] = A.w32[a];
        }
        B->nseg = (int)ceil_div(B->hi[0] - B->lo[0], 64 - 2 * H);
        B->gx = (int)ceil_div(B->nseg, 4);
        B->n0 = A.n[0];
        B->n1 = A.n[1];
        B->ys = A.s1;
        B->shell = f.nat_shell;
    } else {
        B->K = 0;
        ShellArgs& S = B->shell;
        S = ShellArgs{};
        S.nbox = 1;
        for (int k = 0; k < 3; ++k) {
            S.org[0][k] = 0;
            S.ext[0][k] = P.n[k];
        }
        S.start[0] = 0;
        S.start[1] = P.total;
    }
}

void batch_items(const std::vector<BProb>& probs, std::vector<BItem>* items,
                 std::vector<BGroup>* groups)
{
    // group key: kind, window, accumulator
    auto key = [&](const BItem& I) {
        const BProb& B = probs[I.prob];
        return std::make_tuple(I.kind, I.kind != 1 ? B.K : 0, B.plan_accum);
    };
    items->clear();
    for (size_t i = 0; i < probs.size(); ++i) {
        const BProb& B = probs[i];
        if (B.tiled == 1) {
            const int64_t nch = ceil_div(B.hi[1] - B.lo[1], B.chunk);
            for (int64_t ch = 0; ch < nch; ++ch)
                for (int xg = 0; xg < B.gx; ++xg) items->push_back({(int32_t)i, 0, xg, (int32_t)ch});
        } else if (B.tiled == 2) {
            for (int xg = 0; xg < B.gx; ++xg) items->push_back({(int32_t)i, 2, xg, 0});
        }
        const int64_t ns = ceil_div(B.shell.start[B.shell.nbox], kChunkPts);
        for (int64_t c = 0; c < ns; ++c) items->push_back({(int32_t)i, 1, (int32_t)c, 0});
    }
    std::stable_sort(items->begin(), items->end(),
                     [&](const BItem& a, const BItem& b) { return key(a) < key(b); });
    groups->clear();
    for (size_t k = 0; k < items->size(); ++k) {
        const BItem& I = (*items)[k];
        if (groups->empty() || key((*items)[groups->back().first]) != key(I)) {
            const auto t = key(I);
            groups->push_back({std::get<0>(t), std::get<1>(t), std::get<2>(t), (int32_t)k, 0});
        }
        ++groups->back().count;
    }
}

}  // namespace dop
