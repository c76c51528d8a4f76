// Device code shared by the single-system and batched TDMA kernels: the
// per-column float64 Thomas solve and the mixed-precision (float32 sweeps,
// float64 residual refinement) solve. See tdma.hip for the semantics.
#pragma once
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

// Store value(k) for every level k < nz without materialising the column.
template <bool VEC, int NZB, class F>
__device__ __forceinline__ void store_levels_fn(double* __restrict__ col, int nz, F value) {
    if constexpr (VEC) {
#pragma unroll
        for (int k = 0; k < NZB; k += 2) {
            if (k < nz) *reinterpret_cast<double2*>(col + k) = make_double2(value(k), value(k + 1));
        }
    } else {
#pragma unroll
        for (int k = 0; k < NZB; ++k) {
            if (k < nz) col[k] = value(k);
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
__device__ __forceinline__ void solve_f64_column(const TdmaArgs& p, const ColumnRef& cr) {
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
__device__ __forceinline__ void solve_mixed_column(const TdmaArgs& p, const ColumnRef& cr) {
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

    // float32 factors: cp and 1/den (a/den is re-formed from the re-streamed
    // a); x0 = float32 solution
    float cpf[NZB], rif[NZB], x0[NZB];
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
        store_levels_fn<VEC, NZB>(p.x + cr.base, nz, [&](int k) { return (double)x0[k]; });
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
                            zpp = (float)r * rif[k] - ((float)ak * rif[k]) * zpp;
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
                store_levels_fn<VEC, NZB>(p.x + cr.base, nz, [&](int k) { return xval(k) + (double)zz[k]; });
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

}  // namespace vx
