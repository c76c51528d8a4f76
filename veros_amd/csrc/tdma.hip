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
#include "vx_common.hpp"

namespace vx {

struct TdmaArgs {
    const double* a;
    const double* b;
    const double* c;
    const double* d;
    const double* b_edge;   // nullable
    const double* d_edge;   // nullable
    const int32_t* kbot;
    double* x;
    int64_t ny, ncol, sx, sy, kx, ky;
    int32_t nz;
};

constexpr int kChunk = 8;       // levels per load batch
constexpr int kTdmaBlock = 256; // 4 waves, one column per lane

template <bool VEC>
__device__ __forceinline__ void load_levels(const double* __restrict__ col, int q, int nz, double (&v)[kChunk]) {
    if constexpr (VEC) {
#pragma unroll
        for (int t = 0; t < kChunk; t += 2) {
            if (q + t < nz) {
                const double2 w = *reinterpret_cast<const double2*>(col + q + t);
                v[t] = w.x;
                v[t + 1] = w.y;
            }
        }
    } else {
#pragma unroll
        for (int t = 0; t < kChunk; ++t) {
            if (q + t < nz) v[t] = col[q + t];
        }
    }
}

template <bool VEC, int NZB>
__device__ __forceinline__ void store_levels(double* __restrict__ col, int nz, const double (&v)[NZB]) {
    if constexpr (VEC) {
#pragma unroll
        for (int k = 0; k < NZB; k += 2) {
            if (k < nz) *reinterpret_cast<double2*>(col + k) = make_double2(v[k], v[k + 1]);
        }
    } else {
#pragma unroll
        for (int k = 0; k < NZB; ++k) {
            if (k < nz) col[k] = v[k];
        }
    }
}

// Column geometry shared by the kernels: base offset and first wet level
// (nz when the column has no wet level: kbot outside [1, nz]).
struct ColumnRef {
    int64_t base;
    int k0;
};

__device__ __forceinline__ bool column_of(const TdmaArgs& p, ColumnRef& cr) {
    const int64_t col = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (col >= p.ncol) return false;
    const int64_t i = col / p.ny;
    const int64_t j = col - i * p.ny;
    cr.base = i * p.sx + j * p.sy;
    const int kb = p.kbot[i * p.kx + j * p.ky];
    cr.k0 = (kb >= 1 && kb <= p.nz) ? kb - 1 : p.nz;
    return true;
}

// ---------------------------------------------------------------------------
// float64 Thomas: forward elimination in registers, back substitution in place.
// ---------------------------------------------------------------------------
template <int NZB, bool VEC>
__global__ __launch_bounds__(kTdmaBlock) void tdma_f64_kernel(TdmaArgs p) {
    ColumnRef cr;
    if (!column_of(p, cr)) return;
    const int nz = p.nz;
    const int k0 = cr.k0;
    const double* __restrict__ ca = p.a + cr.base;
    const double* __restrict__ cb = p.b + cr.base;
    const double* __restrict__ cc = p.c + cr.base;
    const double* __restrict__ cd = p.d + cr.base;
    const bool wet = k0 < nz;
    const double be = (wet && p.b_edge) ? p.b_edge[cr.base + k0] : 0.0;
    const double de = (wet && p.d_edge) ? p.d_edge[cr.base + k0] : 0.0;
    const bool has_be = p.b_edge != nullptr;
    const bool has_de = p.d_edge != nullptr;

    double cp[NZB], dp[NZB];
    double cpp = 0.0, dpp = 0.0;
#pragma unroll
    for (int q = 0; q < NZB; q += kChunk) {
        if (q < nz) {
            double va[kChunk] = {}, vb[kChunk] = {}, vc[kChunk] = {}, vd[kChunk] = {};
            if (q + kChunk > k0) {
                load_levels<VEC>(ca, q, nz, va);
                load_levels<VEC>(cb, q, nz, vb);
                load_levels<VEC>(cc, q, nz, vc);
                load_levels<VEC>(cd, q, nz, vd);
            }
#pragma unroll
            for (int t = 0; t < kChunk; ++t) {
                const int k = q + t;
                if (k < nz && k >= k0) {
                    const bool edge = k == k0;
                    const double ak = edge ? 0.0 : va[t];
                    const double bk = (edge && has_be) ? be : vb[t];
                    const double ck = (k == nz - 1) ? 0.0 : vc[t];
                    const double dk = (edge && has_de) ? de : vd[t];
                    const double r = 1.0 / (bk - ak * cpp);
                    cpp = ck * r;
                    dpp = (dk - ak * dpp) * r;
                }
                cp[k] = cpp;
                dp[k] = dpp;
            }
        }
    }
    double xn = 0.0;
#pragma unroll
    for (int k = NZB - 1; k >= 0; --k) {
        if (k < nz) {
            const double xk = (k >= k0) ? dp[k] - cp[k] * xn : 0.0;
            dp[k] = xk;
            xn = xk;
        }
    }
    store_levels<VEC, NZB>(p.x + cr.base, nz, dp);
}

// ---------------------------------------------------------------------------
// Mixed precision: float32 factorisation and sweeps, float64 residual
// r = d - A x and float64 accumulation of the corrections.
//   pass 1 : stream A, d; factor A in float32 (cp, 1/den, a/den); x0 = A32^-1 d
//   pass 2+: stream A, d again; r = d - A x (float64); z = A32^-1 r; x += z
// The correction passes re-read the coefficients: the float64 matrix does not
// fit on chip next to the float32 factors for a useful number of columns.
// With refine_steps = 2 the first corrected solution is kept as the float pair
// (x0, z1), whose double sum is exact.
// ---------------------------------------------------------------------------
template <int NZB, bool VEC, int REFINE>
__global__ __launch_bounds__(kTdmaBlock) void tdma_mixed_kernel(TdmaArgs p) {
    ColumnRef cr;
    if (!column_of(p, cr)) return;
    const int nz = p.nz;
    const int k0 = cr.k0;
    const double* __restrict__ ca = p.a + cr.base;
    const double* __restrict__ cb = p.b + cr.base;
    const double* __restrict__ cc = p.c + cr.base;
    const double* __restrict__ cd = p.d + cr.base;
    const bool wet = k0 < nz;
    const double be = (wet && p.b_edge) ? p.b_edge[cr.base + k0] : 0.0;
    const double de = (wet && p.d_edge) ? p.d_edge[cr.base + k0] : 0.0;
    const bool has_be = p.b_edge != nullptr;
    const bool has_de = p.d_edge != nullptr;

    // float32 factors: cp, 1/den, a/den; x0 = float32 solution
    float cpf[NZB], rif[NZB], arf[NZB], x0[NZB];
    // second float word of the solution (REFINE == 2 keeps x = x0 + z1)
    float z1[REFINE == 2 ? NZB : 1];
    {
        float cpp = 0.f, ypp = 0.f;
#pragma unroll
        for (int q = 0; q < NZB; q += kChunk) {
            if (q < nz) {
                double va[kChunk] = {}, vb[kChunk] = {}, vc[kChunk] = {}, vd[kChunk] = {};
                if (q + kChunk > k0) {
                    load_levels<VEC>(ca, q, nz, va);
                    load_levels<VEC>(cb, q, nz, vb);
                    load_levels<VEC>(cc, q, nz, vc);
                    load_levels<VEC>(cd, q, nz, vd);
                }
#pragma unroll
                for (int t = 0; t < kChunk; ++t) {
                    const int k = q + t;
                    if (k < nz && k >= k0) {
                        const bool edge = k == k0;
                        const float af = edge ? 0.f : (float)va[t];
                        const float bf = (float)((edge && has_be) ? be : vb[t]);
                        const float cf = (k == nz - 1) ? 0.f : (float)vc[t];
                        const float df = (float)((edge && has_de) ? de : vd[t]);
                        const float ri = __builtin_amdgcn_rcpf(bf - af * cpp);
                        cpp = cf * ri;
                        ypp = (df - af * ypp) * ri;
                        arf[k] = af * ri;
                        rif[k] = ri;
                    }
                    cpf[k] = cpp;
                    x0[k] = ypp;
                }
            }
        }
        float xn = 0.f;
#pragma unroll
        for (int k = NZB - 1; k >= 0; --k) {
            if (k < nz) {
                const float xk = (k >= k0) ? x0[k] - cpf[k] * xn : 0.f;
                x0[k] = xk;
                xn = xk;
            }
        }
    }
    if constexpr (REFINE == 0) {
        double out[NZB];
#pragma unroll
        for (int k = 0; k < NZB; ++k) out[k] = (double)x0[k];
        store_levels<VEC, NZB>(p.x + cr.base, nz, out);
        return;
    } else {
        if constexpr (REFINE == 2) {
#pragma unroll
            for (int k = 0; k < NZB; ++k) z1[k] = 0.f;
        }
        auto xval = [&](int k) -> double {
            if constexpr (REFINE == 2) return (double)x0[k] + (double)z1[k];
            else return (double)x0[k];
        };
#pragma unroll 1
        for (int step = 0; step < REFINE; ++step) {
            // residual r = d - A x in float64, forward substitution with the float32 factors
            float zz[NZB];
            float zpp = 0.f;
            double xprev = 0.0;
#pragma unroll
            for (int q = 0; q < NZB; q += kChunk) {
                if (q < nz) {
                    double va[kChunk] = {}, vb[kChunk] = {}, vc[kChunk] = {}, vd[kChunk] = {};
                    if (q + kChunk > k0) {
                        load_levels<VEC>(ca, q, nz, va);
                        load_levels<VEC>(cb, q, nz, vb);
                        load_levels<VEC>(cc, q, nz, vc);
                        load_levels<VEC>(cd, q, nz, vd);
                    }
#pragma unroll
                    for (int t = 0; t < kChunk; ++t) {
                        const int k = q + t;
                        if (k < nz && k >= k0) {
                            const bool edge = k == k0;
                            const double xk = xval(k);
                            const int ku = (k + 1 < NZB) ? k + 1 : NZB - 1;  // static after unrolling
                            const double xu = (k + 1 < nz) ? xval(ku) : 0.0;
                            const double ak = edge ? 0.0 : va[t];
                            const double bk = (edge && has_be) ? be : vb[t];
                            const double ck = (k == nz - 1) ? 0.0 : vc[t];
                            const double dk = (edge && has_de) ? de : vd[t];
                            const double r = dk - ak * xprev - bk * xk - ck * xu;
                            zpp = (float)r * rif[k] - arf[k] * zpp;
                            xprev = xk;
                        }
                        zz[k] = zpp;
                    }
                }
            }
            float zn = 0.f;
#pragma unroll
            for (int k = NZB - 1; k >= 0; --k) {
                if (k < nz) {
                    const float zk = (k >= k0) ? zz[k] - cpf[k] * zn : 0.f;
                    zz[k] = zk;
                    zn = zk;
                }
            }
            if (step == REFINE - 1) {
                double out[NZB];
#pragma unroll
                for (int k = 0; k < NZB; ++k) out[k] = xval(k) + (double)zz[k];
                store_levels<VEC, NZB>(p.x + cr.base, nz, out);
            } else if constexpr (REFINE == 2) {
                // x = x0 + z1 + zz, re-split into two float words (x0 + z1 is exact in float64)
#pragma unroll
                for (int k = 0; k < NZB; ++k) {
                    const double s = (double)x0[k] + (double)z1[k] + (double)zz[k];
                    const float hi = (float)s;
                    x0[k] = hi;
                    z1[k] = (float)(s - (double)hi);
                }
            }
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
    switch (refine_steps) {
        case 0: return launch_bucket<MixK0>(p, (hipStream_t)stream, "vx_tdma_mixed");
        case 1: return launch_bucket<MixK1>(p, (hipStream_t)stream, "vx_tdma_mixed");
        default: return launch_bucket<MixK2>(p, (hipStream_t)stream, "vx_tdma_mixed");
    }
}
