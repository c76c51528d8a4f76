// Batched kbot-masked tridiagonal (Thomas) solves for gfx950.
//
// Semantics follow solve_implicit / solve_tridiagonal (veros/core/utilities.py:51-59,
// veros/core/operators.py:60-77,133-156): per column the wet suffix
// k in [kbot-1, nz) is one system, the sub-diagonal is 0 at the edge row
// k = kbot-1 (where b, d may be replaced by b_edge, d_edge) and the
// super-diagonal is 0 at the top row; x is 0 outside the wet suffix.
//
// Layout: (x, y, z) with z contiguous. One lane owns one column and walks it
// bottom-up; the elimination factors cp/dp live in VGPRs (the level loop is
// fully unrolled per nz bucket, so the register arrays are statically
// indexed). Coefficients are read in 16-byte pairs where alignment allows;
// each lane streams whole cache lines of its own column, so the L1 does the
// transposition and HBM sees full-line traffic.
#include "tdma_seg.hpp"

#include <cstdlib>

namespace vx {

template <int S, bool VEC>
__global__ __launch_bounds__(kSegBlock) void tdma_seg_f64_kernel(TdmaArgs p) {
    seg_f64<S, VEC>(p, seg_lane_of(p, S));
}

template <int S, bool VEC, int REFINE>
__global__ __launch_bounds__(kSegBlock) void tdma_seg_mixed_kernel(TdmaArgs p) {
    seg_mixed<S, VEC, REFINE>(p, seg_lane_of(p, S));
}

// LDS-staged segmented kernels (16-byte aligned operands only). Dynamic LDS:
// seg_lds_wave_bytes(S) per wave.
template <int S, bool VEC>
__global__ __launch_bounds__(kSegLdsBlock) void tdma_segl_f64_kernel(TdmaArgs p) {
    extern __shared__ double seg_lds[];
    double* lds = seg_lds + (threadIdx.x / 64) * (seg_lds_wave_bytes(S) / 8);
    seg_f64_lds<S>(p, seg_lane_of(p, S), lds);
}

template <int S, bool VEC, int REFINE>
__global__ __launch_bounds__(kSegLdsBlock) void tdma_segl_mixed_kernel(TdmaArgs p) {
    extern __shared__ double seg_lds[];
    double* lds = seg_lds + (threadIdx.x / 64) * (seg_lds_wave_bytes(S) / 8);
    seg_mixed_lds<S, REFINE>(p, seg_lane_of(p, S), lds);
}

template <int NZB, bool VEC>
__global__ __launch_bounds__(kTdmaBlock) void tdma_f64_kernel(TdmaArgs p) {
    ColumnRef cr;
    if (column_of(p, cr)) solve_f64_column<NZB, VEC>(p, cr);
}

template <int NZB, bool VEC, int REFINE>
__global__ __launch_bounds__(kTdmaBlock) void tdma_mixed_kernel(TdmaArgs p) {
    ColumnRef cr;
    if (column_of(p, cr)) solve_mixed_column<NZB, VEC, REFINE>(p, cr);
}

// ---------------------------------------------------------------------------
// Adjoint of the float64 solve: given x = A^-1 d and g = dL/dx, the
// transposed system A^T lam = g (sub-diagonal c[k-1], super-diagonal a[k+1]
// on the wet rows) is solved by Thomas, re-deriving the elimination from a, b,
// c instead of storing forward factors; then dL/dd = lam, dL/db = -lam*x,
// dL/da[k] = -lam[k]*x[k-1] (k above the edge), dL/dc[k] = -lam[k]*x[k+1]
// (k below the top). Entries the forward solve ignores (a at the edge, c at
// the top, dry cells, b/d at the edge when b_edge/d_edge replace them) get 0.
// ---------------------------------------------------------------------------
struct TdmaAdjArgs {
    TdmaArgs f;       // forward inputs; f.x is the forward solution
    const double* g;  // dL/dx
    double *ga, *gb, *gc, *gd;
    double *gbe, *gde;  // dL/db_edge, dL/dd_edge (nullable; used when b_edge / d_edge were given)
};

template <int NZB>
__global__ __launch_bounds__(kTdmaBlock) void tdma_adjoint_kernel(TdmaAdjArgs q) {
    const TdmaArgs& p = q.f;
    ColumnRef cr;
    if (!column_of(p, cr)) return;
    const int nz = p.nz;
    const int k0 = cr.k0;
    const int64_t o = cr.base;
    const bool has_be = p.b_edge != nullptr;
    const bool has_de = p.d_edge != nullptr;
    double cp[NZB], lm[NZB];
    double cpp = 0.0, dpp = 0.0;
#pragma unroll
    for (int k = 0; k < NZB; ++k) {
        if (k < nz && k >= k0) {
            const double sub = (k > k0) ? p.c[o + k - 1] : 0.0;
            const double diag = (k == k0 && has_be) ? p.b_edge[o + k] : p.b[o + k];
            const double sup = (k < nz - 1) ? p.a[o + k + 1] : 0.0;
            const double r = 1.0 / (diag - sub * cpp);
            cpp = sup * r;
            dpp = (q.g[o + k] - sub * dpp) * r;
        }
        cp[k] = cpp;
        lm[k] = dpp;
    }
    double ln = 0.0;
#pragma unroll
    for (int k = NZB - 1; k >= 0; --k) {
        if (k < nz && k >= k0) {
            ln = lm[k] - cp[k] * ln;
            lm[k] = ln;
        }
    }
    double xm = 0.0;  // x[k-1]
#pragma unroll
    for (int k = 0; k < NZB; ++k) {
        if (k < nz) {
            double ga = 0.0, gb = 0.0, gc = 0.0, gd = 0.0, gbe = 0.0, gde = 0.0;
            if (k >= k0) {
                const double lam = lm[k];
                const double xk = p.x[o + k];
                const bool edge = k == k0;
                if (edge && has_de) gde = lam;
                else gd = lam;
                if (edge && has_be) gbe = -lam * xk;
                else gb = -lam * xk;
                if (!edge) ga = -lam * xm;
                if (k < nz - 1) gc = -lam * p.x[o + k + 1];
                xm = xk;
            }
            q.ga[o + k] = ga;
            q.gb[o + k] = gb;
            q.gc[o + k] = gc;
            q.gd[o + k] = gd;
            if (q.gbe) q.gbe[o + k] = gbe;
            if (q.gde) q.gde[o + k] = gde;
        }
    }
}

// ---------------------------------------------------------------------------
// host side
// ---------------------------------------------------------------------------
namespace {

int validate_tdma(const TdmaArgs& p, int64_t nx, int64_t ny, int32_t nz, const char* who) {
    if (nx < 0 || ny < 0) return fail(VX_ERR_INVALID_ARGUMENT, "%s: negative shape (%lld, %lld)", who, (long long)nx, (long long)ny);
    if (nz < 1) return fail(VX_ERR_INVALID_ARGUMENT, "%s: nz must be >= 1 (got %d)", who, nz);
    if (nz > kMaxNz) return fail(VX_ERR_UNSUPPORTED, "%s: nz=%d exceeds the compiled maximum %d", who, nz, kMaxNz);
    if (nx * ny == 0) return VX_OK;
    if (!p.a || !p.b || !p.c || !p.d || !p.kbot || !p.x)
        return fail(VX_ERR_INVALID_ARGUMENT, "%s: null array pointer", who);
    if (p.sy < nz || p.sx < (ny - 1) * p.sy + nz)
        return fail(VX_ERR_INVALID_ARGUMENT, "%s: strides (sx=%lld, sy=%lld) overlap columns of depth %d", who,
                    (long long)p.sx, (long long)p.sy, nz);
    return VX_OK;
}

bool vec_ok(const TdmaArgs& p) {
    if ((p.nz & 1) || (p.sx & 1) || (p.sy & 1)) return false;
    return aligned16(p.a) && aligned16(p.b) && aligned16(p.c) && aligned16(p.d) && aligned16(p.x);
}

template <template <int, bool> class K>
int launch_bucket(const TdmaArgs& p, hipStream_t s, const char* who) {
    const int nzb = ((p.nz + 15) / 16) * 16;
    const bool vec = vec_ok(p);
    const dim3 grid((unsigned)((p.ncol + kTdmaBlock - 1) / kTdmaBlock));
    const dim3 block(kTdmaBlock);
#define VX_TDMA_CASE(N)                                                                      \
    case N:                                                                                  \
        if (vec) hipLaunchKernelGGL((K<N, true>::fn), grid, block, 0, s, p);       \
        else hipLaunchKernelGGL((K<N, false>::fn), grid, block, 0, s, p);          \
        break;
    switch (nzb) {
        VX_TDMA_CASE(16)
        VX_TDMA_CASE(32)
        VX_TDMA_CASE(48)
        VX_TDMA_CASE(64)
        VX_TDMA_CASE(80)
        VX_TDMA_CASE(96)
        VX_TDMA_CASE(112)
        VX_TDMA_CASE(128)
        default:
            return fail(VX_ERR_UNSUPPORTED, "%s: nz=%d", who, p.nz);
    }
#undef VX_TDMA_CASE
    return check_launch(who);
}

// Segmented kernels (kSegP lanes per column), instantiated per levels-per-lane S.
template <template <int, bool> class K>
int launch_seg(const TdmaArgs& p, hipStream_t s, const char* who) {
    const int S = seg_levels(p.nz);
    const bool vec = vec_ok(p);
    const int64_t threads = p.ncol * kSegP;
    const dim3 grid((unsigned)((threads + kSegBlock - 1) / kSegBlock));
    const dim3 block(kSegBlock);
#define VX_SEG_CASE(N)                                                            \
    case N:                                                                       \
        if (vec) hipLaunchKernelGGL((K<N, true>::fn), grid, block, 0, s, p);    \
        else hipLaunchKernelGGL((K<N, false>::fn), grid, block, 0, s, p);       \
        break;
    switch (S) {
        VX_SEG_CASE(2)
        VX_SEG_CASE(4)
        VX_SEG_CASE(6)
        VX_SEG_CASE(8)
        VX_SEG_CASE(10)
        VX_SEG_CASE(12)
        VX_SEG_CASE(14)
        VX_SEG_CASE(16)
        default:
            return fail(VX_ERR_UNSUPPORTED, "%s: nz=%d", who, p.nz);
    }
#undef VX_SEG_CASE
    return check_launch(who);
}

// LDS-staged segmented kernels: only with 16-byte aligned operands.
template <template <int, bool> class K>
int launch_segl(const TdmaArgs& p, hipStream_t s, const char* who) {
    const int S = seg_levels(p.nz);
    const int64_t threads = p.ncol * kSegP;
    const dim3 grid((unsigned)((threads + kSegLdsBlock - 1) / kSegLdsBlock));
    const dim3 block(kSegLdsBlock);
    const size_t lds = (size_t)seg_lds_wave_bytes(S) * (kSegLdsBlock / 64);
#define VX_SEGL_CASE(N)                                                     \
    case N:                                                                 \
        hipLaunchKernelGGL((K<N, true>::fn), grid, block, lds, s, p);       \
        break;
    switch (S) {
        VX_SEGL_CASE(2)
        VX_SEGL_CASE(4)
        VX_SEGL_CASE(6)
        VX_SEGL_CASE(8)
        VX_SEGL_CASE(10)
        VX_SEGL_CASE(12)
        VX_SEGL_CASE(14)
        VX_SEGL_CASE(16)
        default:
            return fail(VX_ERR_UNSUPPORTED, "%s: nz=%d", who, p.nz);
    }
#undef VX_SEGL_CASE
    return check_launch(who);
}

template <int N, bool V>
struct SeglF64K {
    static constexpr auto fn = tdma_segl_f64_kernel<N, V>;
};
template <int N, bool V>
struct SeglMixK0 {
    static constexpr auto fn = tdma_segl_mixed_kernel<N, V, 0>;
};
template <int N, bool V>
struct SeglMixK1 {
    static constexpr auto fn = tdma_segl_mixed_kernel<N, V, 1>;
};
template <int N, bool V>
struct SeglMixK2 {
    static constexpr auto fn = tdma_segl_mixed_kernel<N, V, 2>;
};

template <int N, bool V>
struct SegF64K {
    static constexpr auto fn = tdma_seg_f64_kernel<N, V>;
};
template <int N, bool V>
struct SegMixK0 {
    static constexpr auto fn = tdma_seg_mixed_kernel<N, V, 0>;
};
template <int N, bool V>
struct SegMixK1 {
    static constexpr auto fn = tdma_seg_mixed_kernel<N, V, 1>;
};
template <int N, bool V>
struct SegMixK2 {
    static constexpr auto fn = tdma_seg_mixed_kernel<N, V, 2>;
};

// Kernel family of the single-grid solves (VX_TDMA_KERNEL, read once):
// "lds" (default; segmented with LDS-staged coalesced copies when the operands
// are 16-byte aligned, else "segment"), "segment" (segmented, register loads),
// "column" (lane per column).
int tdma_kernel_family() {
    static const int fam = [] {
        const char* e = std::getenv("VX_TDMA_KERNEL");
        if (e != nullptr && std::string(e) == "column") return 0;
        if (e != nullptr && std::string(e) == "segment") return 1;
        return 2;
    }();
    return fam;
}
bool use_column_kernels() { return tdma_kernel_family() == 0; }
bool use_lds_kernels(const TdmaArgs& p) { return tdma_kernel_family() == 2 && vec_ok(p); }

template <int N, bool V>
struct F64K {
    static constexpr auto fn = tdma_f64_kernel<N, V>;
};
template <int N, bool V>
struct MixK0 {
    static constexpr auto fn = tdma_mixed_kernel<N, V, 0>;
};
template <int N, bool V>
struct MixK1 {
    static constexpr auto fn = tdma_mixed_kernel<N, V, 1>;
};
template <int N, bool V>
struct MixK2 {
    static constexpr auto fn = tdma_mixed_kernel<N, V, 2>;
};

TdmaArgs make_args(const double* a, const double* b, const double* c, const double* d, const double* b_edge,
                   const double* d_edge, const int32_t* kbot, int64_t kbot_sx, int64_t kbot_sy, double* x,
                   int64_t nx, int64_t ny, int32_t nz, int64_t sx, int64_t sy) {
    TdmaArgs p;
    p.a = a;
    p.b = b;
    p.c = c;
    p.d = d;
    p.b_edge = b_edge;
    p.d_edge = d_edge;
    p.kbot = kbot;
    p.x = x;
    p.ny = ny;
    p.ncol = nx * ny;
    p.sx = sx;
    p.sy = sy;
    p.kx = kbot_sx;
    p.ky = kbot_sy;
    p.nz = nz;
    return p;
}

}  // namespace
}  // namespace vx

extern "C" int vx_tdma_f64(void* stream, const double* a, const double* b, const double* c, const double* d,
                           const double* b_edge, const double* d_edge, const int32_t* kbot, int64_t kbot_sx,
                           int64_t kbot_sy, double* x, int64_t nx, int64_t ny, int32_t nz, int64_t sx, int64_t sy) {
    using namespace vx;
    TdmaArgs p = make_args(a, b, c, d, b_edge, d_edge, kbot, kbot_sx, kbot_sy, x, nx, ny, nz, sx, sy);
    int rc = validate_tdma(p, nx, ny, nz, "vx_tdma_f64");
    if (rc != VX_OK || p.ncol == 0) return rc;
    if (use_lds_kernels(p)) return launch_segl<SeglF64K>(p, (hipStream_t)stream, "vx_tdma_f64");
    if (!use_column_kernels()) return launch_seg<SegF64K>(p, (hipStream_t)stream, "vx_tdma_f64");
    return launch_bucket<F64K>(p, (hipStream_t)stream, "vx_tdma_f64");
}

extern "C" int vx_tdma_mixed(void* stream, const double* a, const double* b, const double* c, const double* d,
                             const double* b_edge, const double* d_edge, const int32_t* kbot, int64_t kbot_sx,
                             int64_t kbot_sy, double* x, int64_t nx, int64_t ny, int32_t nz, int64_t sx, int64_t sy,
                             int32_t refine_steps) {
    using namespace vx;
    if (refine_steps < 0 || refine_steps > 2)
        return fail(VX_ERR_INVALID_ARGUMENT, "vx_tdma_mixed: refine_steps must be 0, 1 or 2 (got %d)", refine_steps);
    TdmaArgs p = make_args(a, b, c, d, b_edge, d_edge, kbot, kbot_sx, kbot_sy, x, nx, ny, nz, sx, sy);
    int rc = validate_tdma(p, nx, ny, nz, "vx_tdma_mixed");
    if (rc != VX_OK || p.ncol == 0) return rc;
    if (use_lds_kernels(p)) {
        switch (refine_steps) {
            case 0: return launch_segl<SeglMixK0>(p, (hipStream_t)stream, "vx_tdma_mixed");
            case 1: return launch_segl<SeglMixK1>(p, (hipStream_t)stream, "vx_tdma_mixed");
            default: return launch_segl<SeglMixK2>(p, (hipStream_t)stream, "vx_tdma_mixed");
        }
    }
    if (!use_column_kernels()) {
        switch (refine_steps) {
            case 0: return launch_seg<SegMixK0>(p, (hipStream_t)stream, "vx_tdma_mixed");
            case 1: return launch_seg<SegMixK1>(p, (hipStream_t)stream, "vx_tdma_mixed");
            default: return launch_seg<SegMixK2>(p, (hipStream_t)stream, "vx_tdma_mixed");
        }
    }
    switch (refine_steps) {
        case 0: return launch_bucket<MixK0>(p, (hipStream_t)stream, "vx_tdma_mixed");
        case 1: return launch_bucket<MixK1>(p, (hipStream_t)stream, "vx_tdma_mixed");
        default: return launch_bucket<MixK2>(p, (hipStream_t)stream, "vx_tdma_mixed");
    }
}

extern "C" int vx_tdma_adjoint_f64(void* stream, const double* a, const double* b, const double* c,
                                   const double* b_edge, const double* d_edge, const int32_t* kbot,
                                   int64_t kbot_sx, int64_t kbot_sy, const double* x, const double* gx,
                                   double* ga, double* gb, double* gc, double* gd, double* gb_edge,
                                   double* gd_edge, int64_t nx, int64_t ny, int32_t nz, int64_t sx, int64_t sy) {
    using namespace vx;
    const char* who = "vx_tdma_adjoint_f64";
    TdmaArgs p = make_args(a, b, c, gx, b_edge, d_edge, kbot, kbot_sx, kbot_sy, const_cast<double*>(x), nx, ny,
                           nz, sx, sy);
    int rc = validate_tdma(p, nx, ny, nz, who);
    if (rc != VX_OK || p.ncol == 0) return rc;
    if (!gx || !ga || !gb || !gc || !gd) return fail(VX_ERR_INVALID_ARGUMENT, "%s: null gradient array", who);
    if ((b_edge != nullptr) != (gb_edge != nullptr) || (d_edge != nullptr) != (gd_edge != nullptr))
        return fail(VX_ERR_INVALID_ARGUMENT, "%s: gb_edge/gd_edge must be given exactly when b_edge/d_edge are",
                    who);
    TdmaAdjArgs q{p, gx, ga, gb, gc, gd, gb_edge, gd_edge};
    const int nzb = ((nz + 15) / 16) * 16;
    const dim3 grid((unsigned)((p.ncol + kTdmaBlock - 1) / kTdmaBlock)), block(kTdmaBlock);
    hipStream_t s = (hipStream_t)stream;
#define VX_ADJ_CASE(N) \
    case N: hipLaunchKernelGGL(tdma_adjoint_kernel<N>, grid, block, 0, s, q); break;
    switch (nzb) {
        VX_ADJ_CASE(16)
        VX_ADJ_CASE(32)
        VX_ADJ_CASE(48)
        VX_ADJ_CASE(64)
        VX_ADJ_CASE(80)
        VX_ADJ_CASE(96)
        VX_ADJ_CASE(112)
        VX_ADJ_CASE(128)
        default: return fail(VX_ERR_UNSUPPORTED, "%s: nz=%d", who, nz);
    }
#undef VX_ADJ_CASE
    return check_launch(who);
}
