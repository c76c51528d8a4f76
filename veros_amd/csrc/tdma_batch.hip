// Batched TDMA over many independent grids of different shapes and depths.
//
// A plan is built once from the list of systems (the arrays of every grid
// stay where they are): systems are grouped by 16-level nz bucket (8 levels
// per lane in the segmented solve: S = bucket / 8) and each bucket's columns
// form one global column space. vx_tdma_plan_run launches one persistent
// grid-stride kernel per non-empty bucket; a group of kSegP lanes owns one
// global column at a time, finds its system by binary search over the
// bucket's column offsets, and runs the segmented solve of the single-grid
// kernels (tdma_seg.hpp) in float64 (mode 0) or in mixed precision with one
// refinement step (mode 1). vx_tdma_plan_tune times block-size / grid-size
// candidates per bucket on the plan's own data and keeps the fastest (a
// process-wide cache keyed by bucket, mode and size reuses earlier choices);
// vx_tdma_plan_graph_launch captures run() once into a hipGraph and replays
// it.
#include <algorithm>
#include <cstring>
#include <map>
#include <mutex>
#include <tuple>
#include <vector>

#include "tdma_seg.hpp"

struct vx_tdma_plan;

namespace vx {

struct SysDesc {
    TdmaArgs args;
    int64_t col0;  // first column of this system in its bucket's column space
    int32_t vec;   // 16-byte vector loads allowed
    int32_t _pad;
};

// One launch per bucket: S levels per lane, MODE 0 = float64, 1 = mixed with
// one refinement step. Every system's columns start at a multiple of 8 in
// the bucket's column space (its count padded to a multiple of 8), so the 8
// columns of a wave's tile always belong to one system: a wave finds it once
// and, when its arrays allow 16-byte copies, runs the LDS-staged segmented
// solve of the single-grid kernels; otherwise the register variant.
template <int S, int MODE>
__global__ __launch_bounds__(128) void tdma_seg_batch_kernel(const SysDesc* __restrict__ sys, int nsys,
                                                             int64_t ntiles) {
    extern __shared__ double seg_lds[];
    double* lds = seg_lds + (threadIdx.x / 64) * (seg_lds_wave_bytes(S) / 8);
    const int64_t wstride = (int64_t)gridDim.x * (blockDim.x / 64);
    const int lane = threadIdx.x & 63;
    const int s = lane % kSegP;
    const int c = lane / kSegP;
    // the tile index is the same in every lane of a wave: keep it (and so
    // the system search and descriptor loads) in scalar registers
    const int64_t w0 = (int64_t)blockIdx.x * (blockDim.x / 64) + __builtin_amdgcn_readfirstlane(threadIdx.x / 64);
    for (int64_t w = w0; w < ntiles; w += wstride) {
        const int64_t g0 = w * kSegCols;  // the tile's first column
        int lo = 0, hi = nsys - 1;
        while (lo < hi) {
            const int mid = (lo + hi + 1) >> 1;
            if (sys[mid].col0 <= g0) lo = mid;
            else hi = mid - 1;
        }
        lo = __builtin_amdgcn_readfirstlane(lo);
        const SysDesc& sd = sys[lo];
        const TdmaArgs p = sd.args;
        const int64_t col = g0 + c - sd.col0;
        SegLane L;
        L.s = s;
        L.lo = s * S;
        L.valid = col < p.ncol;
        L.base = 0;
        L.k0 = p.nz;
        if (L.valid) {
            const int64_t i = col / p.ny;
            const int64_t j = col - i * p.ny;
            L.base = i * p.sx + j * p.sy;
            const int kb = p.kbot[i * p.kx + j * p.ky];
            L.k0 = (kb >= 1 && kb <= p.nz) ? kb - 1 : p.nz;
        }
        double va[S], vb[S], vc[S], vd[S];
        if (sd.vec) {
            seg_rows_lds<S>(p, L, lds, va, vb, vc, vd);
        } else {
            seg_rows<S, false>(p, L, va, vb, vc, vd);
        }
        double x[S];
        if constexpr (MODE == 0) {
            seg_solve_f64<S>(L, va, vb, vc, vd);
#pragma unroll
            for (int t = 0; t < S; ++t) x[t] = vd[t];
        } else {
            seg_mixed_core<S, 1>(L, va, vb, vc, vd, x);
        }
        if (sd.vec) {
            seg_store_lds<S>(p, L, lds, x);
        } else if (L.valid) {
            seg_store<S, false>(p.x + L.base, L.lo, p.nz, x);
        }
    }
}

namespace {

constexpr int kBuckets[] = {16, 32, 48, 64, 80, 96, 112, 128};
constexpr int kNumBuckets = 8;
constexpr int kBlockSizes[] = {64, 128};  // one or two waves (LDS: seg_lds_wave_bytes(S) per wave)
constexpr int kGridFactors[] = {1, 2, 4, 8, 0};  // blocks per CU; 0 = one block per 256 columns

struct Bucket {
    int nzb = 0;
    int first = 0;       // index of the first system of this bucket in the device table
    int nsys = 0;
    int64_t total = 0;   // columns, each system's count padded to a multiple of 8
    int64_t cols = 0;    // columns without the padding
    int bs = 128;
    int factor = 0;
};

using TuneKey = std::tuple<int, int, int64_t, int>;  // nzb, mode, log2(columns), CUs
std::mutex g_tune_mutex;
std::map<TuneKey, std::pair<int, int>> g_tune_cache;

int ilog2(int64_t v) {
    int r = 0;
    while (v > 1) {
        v >>= 1;
        ++r;
    }
    return r;
}

template <int S, int MODE>
void launch_bs(int bs, const SysDesc* d, int nsys, int64_t total, unsigned grid, hipStream_t s) {
    const size_t lds = (size_t)seg_lds_wave_bytes(S) * (bs / 64);
    hipLaunchKernelGGL((tdma_seg_batch_kernel<S, MODE>), dim3(grid), dim3(bs), lds, s, d, nsys, total / kSegCols);
}

template <int MODE>
void launch_mode(int nzb, int bs, const SysDesc* d, int nsys, int64_t total, unsigned grid, hipStream_t s) {
    switch (nzb) {
        case 16: launch_bs<2, MODE>(bs, d, nsys, total, grid, s); break;
        case 32: launch_bs<4, MODE>(bs, d, nsys, total, grid, s); break;
        case 48: launch_bs<6, MODE>(bs, d, nsys, total, grid, s); break;
        case 64: launch_bs<8, MODE>(bs, d, nsys, total, grid, s); break;
        case 80: launch_bs<10, MODE>(bs, d, nsys, total, grid, s); break;
        case 96: launch_bs<12, MODE>(bs, d, nsys, total, grid, s); break;
        case 112: launch_bs<14, MODE>(bs, d, nsys, total, grid, s); break;
        default: launch_bs<16, MODE>(bs, d, nsys, total, grid, s); break;
    }
}

}  // namespace
}  // namespace vx

struct vx_tdma_plan {
    int mode = 0;
    int num_cu = 0;
    std::vector<vx::Bucket> buckets;
    vx::SysDesc* d_sys = nullptr;
    hipStream_t capture_stream = nullptr;
    hipGraph_t graph = nullptr;
    hipGraphExec_t exec = nullptr;
    // every non-empty system (tune checks that none solves in place)
    std::vector<vx_tdma_system> all;
};

namespace vx {
namespace {

unsigned grid_for(const vx_tdma_plan* P, const Bucket& b, int bs, int factor) {
    const int64_t full = (b.total * kSegP + bs - 1) / bs;  // one 8-column tile per wave
    if (factor == 0) return (unsigned)std::max<int64_t>(1, full);
    return (unsigned)std::max<int64_t>(1, std::min<int64_t>(full, (int64_t)factor * P->num_cu));
}

void launch_bucket(const vx_tdma_plan* P, const Bucket& b, int bs, int factor, hipStream_t s) {
    const SysDesc* d = P->d_sys + b.first;
    const unsigned grid = grid_for(P, b, bs, factor);
    if (P->mode == 0) launch_mode<0>(b.nzb, bs, d, b.nsys, b.total, grid, s);
    else launch_mode<1>(b.nzb, bs, d, b.nsys, b.total, grid, s);
}

}  // namespace
}  // namespace vx

extern "C" int vx_tdma_plan_create(const vx_tdma_system* systems, int32_t nsys, int32_t mode,
                                   vx_tdma_plan** out) {
    using namespace vx;
    const char* who = "vx_tdma_plan_create";
    if (!out) return fail(VX_ERR_INVALID_ARGUMENT, "%s: null output handle", who);
    *out = nullptr;
    if (!systems || nsys < 1) return fail(VX_ERR_INVALID_ARGUMENT, "%s: need at least one system", who);
    if (mode < 0 || mode > 1)
        return fail(VX_ERR_INVALID_ARGUMENT, "%s: mode must be 0 (fp64) or 1 (mixed, one refinement step)", who);
    std::vector<std::vector<SysDesc>> per(kNumBuckets);
    std::vector<vx_tdma_system> all;
    for (int q = 0; q < nsys; ++q) {
        const vx_tdma_system& y = systems[q];
        if (y.nz < 1 || y.nx < 0 || y.ny < 0)
            return fail(VX_ERR_INVALID_ARGUMENT, "%s: system %d has an invalid shape", who, q);
        if (y.nz > kMaxNz) return fail(VX_ERR_UNSUPPORTED, "%s: system %d has nz=%d > %d", who, q, y.nz, kMaxNz);
        if (y.nx * y.ny == 0) continue;
        if (!y.a || !y.b || !y.c || !y.d || !y.kbot || !y.x)
            return fail(VX_ERR_INVALID_ARGUMENT, "%s: system %d has a null array", who, q);
        if (y.sy < y.nz || y.sx < (y.ny - 1) * y.sy + y.nz)
            return fail(VX_ERR_INVALID_ARGUMENT, "%s: system %d strides overlap columns", who, q);
        SysDesc sd{};
        TdmaArgs& p = sd.args;
        p.a = y.a, p.b = y.b, p.c = y.c, p.d = y.d, p.b_edge = y.b_edge, p.d_edge = y.d_edge;
        p.kbot = y.kbot, p.x = y.x;
        p.ny = y.ny, p.ncol = y.nx * y.ny, p.sx = y.sx, p.sy = y.sy, p.kx = y.kbot_sx, p.ky = y.kbot_sy;
        p.nz = y.nz;
        sd.vec = (!(y.nz & 1) && !(y.sx & 1) && !(y.sy & 1) && aligned16(y.a) && aligned16(y.b) &&
                  aligned16(y.c) && aligned16(y.d) && aligned16(y.x))
                     ? 1
                     : 0;
        per[(y.nz + 15) / 16 - 1].push_back(sd);
        all.push_back(y);
    }
    auto* P = new vx_tdma_plan();
    P->mode = mode;
    P->all = std::move(all);
    int dev = 0;
    hipDeviceProp_t prop;
    if (hipGetDevice(&dev) != hipSuccess || hipGetDeviceProperties(&prop, dev) != hipSuccess) {
        delete P;
        return fail(VX_ERR_LAUNCH, "%s: cannot query the device", who);
    }
    P->num_cu = prop.multiProcessorCount;
    std::vector<SysDesc> table;
    for (int bk = 0; bk < kNumBuckets; ++bk) {
        if (per[bk].empty()) continue;
        Bucket b;
        b.nzb = kBuckets[bk];
        b.first = (int)table.size();
        b.nsys = (int)per[bk].size();
        int64_t col = 0;
        for (SysDesc& sd : per[bk]) {
            sd.col0 = col;
            col += (sd.args.ncol + kSegCols - 1) / kSegCols * kSegCols;
            b.cols += sd.args.ncol;
            table.push_back(sd);
        }
        b.total = col;
        {
            std::lock_guard<std::mutex> lk(g_tune_mutex);
            auto it = g_tune_cache.find(TuneKey{b.nzb, mode, ilog2(b.total), P->num_cu});
            if (it != g_tune_cache.end()) {
                b.bs = it->second.first;
                b.factor = it->second.second;
            }
        }
        P->buckets.push_back(b);
    }
    if (!table.empty()) {
        if (hipMalloc(&P->d_sys, table.size() * sizeof(SysDesc)) != hipSuccess) {
            delete P;
            return fail(VX_ERR_LAUNCH, "%s: hipMalloc of the system table failed", who);
        }
        if (hipMemcpy(P->d_sys, table.data(), table.size() * sizeof(SysDesc), hipMemcpyHostToDevice) != hipSuccess) {
            (void)hipFree(P->d_sys);
            delete P;
            return fail(VX_ERR_LAUNCH, "%s: copying the system table failed", who);
        }
    }
    *out = P;
    return VX_OK;
}

namespace vx {
namespace {
// Byte range [lo, hi) spanned by one (nx, ny, nz) float64 array of a system.
std::pair<uintptr_t, uintptr_t> span_of(const double* p, const vx_tdma_system& y) {
    const uintptr_t lo = reinterpret_cast<uintptr_t>(p);
    const int64_t last = (y.nx - 1) * y.sx + (y.ny - 1) * y.sy + y.nz;
    return {lo, lo + (uintptr_t)last * sizeof(double)};
}

// True when some system's x overlaps one of its (or another system's) inputs.
bool solves_in_place(const std::vector<vx_tdma_system>& sys) {
    for (const vx_tdma_system& y : sys) {
        const auto xs = span_of(y.x, y);
        for (const vx_tdma_system& z : sys) {
            for (const double* in : {z.a, z.b, z.c, z.d, z.b_edge, z.d_edge}) {
                if (!in) continue;
                const auto r = span_of(in, z);
                if (xs.first < r.second && r.first < xs.second) return true;
            }
        }
    }
    return false;
}

int run_plan(vx_tdma_plan* plan, hipStream_t s, const char* who) {
    for (const Bucket& b : plan->buckets) launch_bucket(plan, b, b.bs, b.factor, s);
    return check_launch(who);
}
}  // namespace
}  // namespace vx

extern "C" int vx_tdma_plan_run(vx_tdma_plan* plan, void* stream) {
    using namespace vx;
    if (!plan) return fail(VX_ERR_INVALID_ARGUMENT, "vx_tdma_plan_run: null plan");
    return run_plan(plan, (hipStream_t)stream, "vx_tdma_plan_run");
}

extern "C" int vx_tdma_plan_tune(vx_tdma_plan* plan, void* stream) {
    using namespace vx;
    const char* who = "vx_tdma_plan_tune";
    if (!plan) return fail(VX_ERR_INVALID_ARGUMENT, "%s: null plan", who);
    // tuning solves every bucket many times over: an in-place system (x
    // overlapping an input) would solve its own previous result
    if (solves_in_place(plan->all))
        return fail(VX_ERR_INVALID_ARGUMENT, "%s: a system's x overlaps an input (in-place solve); tuning would "
                    "overwrite that input", who);
    hipStream_t s = (hipStream_t)stream;
    hipEvent_t e0, e1;
    if (hipEventCreate(&e0) != hipSuccess || hipEventCreate(&e1) != hipSuccess)
        return fail(VX_ERR_LAUNCH, "%s: hipEventCreate failed", who);
    int rc = VX_OK;
    for (Bucket& b : plan->buckets) {
        float best = 1e30f;
        int best_bs = b.bs, best_f = b.factor;
        for (int bs : kBlockSizes) {
            for (int f : kGridFactors) {
                launch_bucket(plan, b, bs, f, s);  // warm-up
                float ms = 1e30f;
                for (int r = 0; r < 3; ++r) {
                    (void)hipEventRecord(e0, s);
                    launch_bucket(plan, b, bs, f, s);
                    (void)hipEventRecord(e1, s);
                    if (hipEventSynchronize(e1) != hipSuccess) {
                        rc = fail(VX_ERR_LAUNCH, "%s: %s", who, hipGetErrorString(hipGetLastError()));
                        break;
                    }
                    float t = 0.f;
                    (void)hipEventElapsedTime(&t, e0, e1);
                    ms = std::min(ms, t);
                }
                if (rc) break;
                if (ms < best) {
                    best = ms;
                    best_bs = bs;
                    best_f = f;
                }
            }
            if (rc) break;
        }
        if (rc) break;
        b.bs = best_bs;
        b.factor = best_f;
        std::lock_guard<std::mutex> lk(g_tune_mutex);
        g_tune_cache[TuneKey{b.nzb, plan->mode, ilog2(b.total), plan->num_cu}] = {b.bs, b.factor};
    }
    (void)hipEventDestroy(e0);
    (void)hipEventDestroy(e1);
    if (rc) return rc;
    // a graph captured before tuning used the old launch shapes
    if (plan->exec) {
        (void)hipGraphExecDestroy(plan->exec);
        plan->exec = nullptr;
    }
    if (plan->graph) {
        (void)hipGraphDestroy(plan->graph);
        plan->graph = nullptr;
    }
    return check_launch(who);
}

extern "C" int vx_tdma_plan_graph_launch(vx_tdma_plan* plan, void* stream) {
    using namespace vx;
    const char* who = "vx_tdma_plan_graph_launch";
    if (!plan) return fail(VX_ERR_INVALID_ARGUMENT, "%s: null plan", who);
    if (!plan->exec) {
        // capture on a private stream (the legacy default stream cannot be captured)
        if (!plan->capture_stream &&
            hipStreamCreateWithFlags(&plan->capture_stream, hipStreamNonBlocking) != hipSuccess)
            return fail(VX_ERR_LAUNCH, "%s: hipStreamCreate failed", who);
        if (hipStreamBeginCapture(plan->capture_stream, hipStreamCaptureModeThreadLocal) != hipSuccess)
            return fail(VX_ERR_LAUNCH, "%s: hipStreamBeginCapture failed", who);
        const int rc = run_plan(plan, plan->capture_stream, who);
        hipGraph_t g = nullptr;
        if (hipStreamEndCapture(plan->capture_stream, &g) != hipSuccess || !g)
            return fail(VX_ERR_LAUNCH, "%s: capture failed", who);
        if (rc) {
            (void)hipGraphDestroy(g);
            return rc;
        }
        if (hipGraphInstantiate(&plan->exec, g, nullptr, nullptr, 0) != hipSuccess) {
            (void)hipGraphDestroy(g);
            return fail(VX_ERR_LAUNCH, "%s: hipGraphInstantiate failed", who);
        }
        plan->graph = g;
    }
    if (hipGraphLaunch(plan->exec, (hipStream_t)stream) != hipSuccess)
        return fail(VX_ERR_LAUNCH, "%s: hipGraphLaunch failed: %s", who, hipGetErrorString(hipGetLastError()));
    return VX_OK;
}

extern "C" int vx_tdma_plan_info(const vx_tdma_plan* plan, int32_t* nbuckets, int64_t* total_columns,
                                 int32_t* choices, int32_t max_choices) {
    using namespace vx;
    if (!plan) return fail(VX_ERR_INVALID_ARGUMENT, "vx_tdma_plan_info: null plan");
    int64_t tot = 0;
    int q = 0;
    for (const Bucket& b : plan->buckets) {
        tot += b.cols;
        if (choices && q < max_choices) {
            choices[3 * q + 0] = b.nzb;
            choices[3 * q + 1] = b.bs;
            choices[3 * q + 2] = b.factor;
        }
        ++q;
    }
    if (nbuckets) *nbuckets = (int32_t)plan->buckets.size();
    if (total_columns) *total_columns = tot;
    return VX_OK;
}

extern "C" void vx_tdma_plan_destroy(vx_tdma_plan* plan) {
    if (!plan) return;
    if (plan->exec) (void)hipGraphExecDestroy(plan->exec);
    if (plan->graph) (void)hipGraphDestroy(plan->graph);
    if (plan->capture_stream) (void)hipStreamDestroy(plan->capture_stream);
    if (plan->d_sys) (void)hipFree(plan->d_sys);
    delete plan;
}
