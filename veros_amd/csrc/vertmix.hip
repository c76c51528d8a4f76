// Implicit vertical mixing of temperature and salinity for gfx950 (float64):
// vertmix_tempsalt, veros/core/thermodynamics.py:248-300.
//
// The reference builds the tridiagonal system of each interior column from
// kappaH and the vertical grid (delta = dt / dzw * kappaH, a = -delta[k-1] /
// dzt, b = 1 + (delta[k] + delta[k-1]) / dzt, c = -delta / dzt, the edge row
// b_edge = 1 + delta / dzt), adds the surface forcing to the top level of the
// right-hand side, solves it with solve_implicit for temp and for salt, and
// then forms the tendencies (new - old) / dt over the whole array and copies
// the cyclic x halos.
//
// Here a group of kSegP lanes owns a column (the segmented solve of
// tdma_seg.hpp, lane s holding levels [s*S, s*S + S)): each lane forms its
// rows in registers, the group solves temperature and salinity with one
// factorisation, and each lane writes the new tracers on the wet levels and
// both tendencies on every level. The coefficients never touch memory: the
// kernel reads kappaH and the two tracer columns once and writes them and the
// tendencies once. With the time levels outermost and 16-byte aligned arrays
// the wave's 8 columns are copied into LDS and written back through it with
// coalesced 16-byte accesses (vertmix_segl_kernel); otherwise each lane
// loads and stores its own levels (vertmix_seg_kernel).
#include "tdma_seg.hpp"

namespace vx {
namespace {

struct VmixArgs {
    const int32_t* kbot;      // (X, Y)
    const double* dzt;        // (nz)
    const double* dzw;        // (nz)
    const double* kappaH;     // (X, Y, nz)
    double* tr[2];            // temp, salt (X, Y, nz, 3)
    const double* forc[2];    // forc_temp_surface, forc_salt_surface (X, Y)
    double* dvmix[2];         // dtemp_vmix, dsalt_vmix (X, Y, nz)
    int64_t X, Y;
    int32_t nz, taup1;
    double dt;
    // time levels of temp/salt: level l of cell c at l * ls + c * cs (cs = 3,
    // ls = 1: levels innermost; cs = 1, ls = X*Y*nz: levels outermost)
    int64_t cs, ls;
};

// The lane's column over the whole (X, Y) grid; only interior columns
// [2, X-2) x [2, Y-2) are solved (k0 = nz elsewhere, and on land).
__device__ __forceinline__ SegLane vmix_lane(const VmixArgs& p, int S, int64_t& col) {
    SegLane L;
    const int64_t gt = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    col = gt / kSegP;
    L.s = (int)(gt % kSegP);
    L.lo = L.s * S;
    L.valid = col < p.X * p.Y;
    if (!L.valid) col = 0;
    const int64_t i = col / p.Y, j = col - (col / p.Y) * p.Y;
    const bool interior = L.valid && i >= 2 && i < p.X - 2 && j >= 2 && j < p.Y - 2;
    const int kb = interior ? p.kbot[col] : 0;
    L.k0 = (kb >= 1 && kb <= p.nz) ? kb - 1 : p.nz;
    L.base = col * p.nz;
    return L;
}

// dzt (s[0, nz)) and dzw (s[kMaxNz, kMaxNz + nz)) in LDS; every thread of
// the block calls it. The rows read the spacings inside per-level selects,
// where a global load is waited for at once, level after level.
__device__ __forceinline__ void vmix_stage_dz(const VmixArgs& p, double* s) {
    for (int k = threadIdx.x; k < p.nz; k += blockDim.x) {
        s[k] = p.dzt[k];
        s[kMaxNz + k] = p.dzw[k];
    }
    __syncthreads();
}

// The rows of the lane's levels (veros/core/thermodynamics.py:267-277),
// identity rows off the wet suffix; z1, z2 get the surface forcing on the top
// level. KM(k) is kappaH at level k of the column, for k in [lo-1, lo+S); its
// value matters for k0 <= k < nz-1 only. zt, zw: dzt and dzw (in LDS).
template <int S, class KMF>
__device__ __forceinline__ void vmix_rows(const VmixArgs& p, const double* __restrict__ zt,
                                          const double* __restrict__ zw, const SegLane& L, KMF KM, double fT, double fS,
                                          double (&va)[S], double (&vb)[S], double (&vc)[S], double (&z1)[S],
                                          double (&z2)[S]) {
    const int nz = p.nz;
    const double dt = p.dt;
    const int kp = L.lo - 1;  // delta at the level below the segment
    double dlm = (kp >= L.k0 && kp < nz - 1) ? dt / zw[kp > 0 ? kp : 0] * KM(kp) : 0.0;
#pragma unroll
    for (int t = 0; t < S; ++t) {
        const int k = L.lo + t;
        const bool wet = k >= L.k0 && k < nz;
        const int kc = k < nz ? k : nz - 1;
        const double dz = zt[kc];
        const double dl = (wet && k < nz - 1) ? dt / zw[kc] * KM(k) : 0.0;
        const bool edge = k == L.k0;
        va[t] = (wet && !edge) ? -dlm / dz : 0.0;
        vb[t] = wet ? (edge ? 1 + dl / dz : 1 + (dl + dlm) / dz) : 1.0;
        vc[t] = (wet && k < nz - 1) ? -dl / dz : 0.0;
        if (k == nz - 1) {  // surface forcing (:277, :284)
            z1[t] = z1[t] + dt * fT / dz;
            z2[t] = z2[t] + dt * fS / dz;
        }
        z1[t] = wet ? z1[t] : 0.0;
        z2[t] = wet ? z2[t] : 0.0;
        dlm = dl;
    }
}

// Register variant: each lane reads and writes its own levels (any layout
// and alignment; VEC: 16-byte tendency stores, nz even and both tendency
// arrays 16-byte aligned).
template <int S, bool VEC>
__global__ __launch_bounds__(kSegBlock) void vertmix_seg_kernel(VmixArgs p) {
    __shared__ double sdz[2 * kMaxNz];
    vmix_stage_dz(p, sdz);
    int64_t col;
    const SegLane L = vmix_lane(p, S, col);
    const int nz = p.nz;
    const int64_t cs = p.cs;
    double* __restrict__ T = p.tr[0] + p.taup1 * p.ls;
    double* __restrict__ Sa = p.tr[1] + p.taup1 * p.ls;
    double km[S + 1], oT[S], oS[S];  // km[t]: kappaH at level lo-1+t
#pragma unroll
    for (int t = 0; t <= S; ++t) {
        const int k = L.lo - 1 + t;
        km[t] = (k >= L.k0 && k < nz) ? p.kappaH[L.base + k] : 0.0;
    }
#pragma unroll
    for (int t = 0; t < S; ++t) {
        const int k = L.lo + t;
        const bool w = k >= L.k0 && k < nz;
        oT[t] = w ? T[cs * (L.base + k)] : 0.0;
        oS[t] = w ? Sa[cs * (L.base + k)] : 0.0;
    }
    const bool solve = L.k0 < nz;
    const double fT = solve ? p.forc[0][col] : 0.0, fS = solve ? p.forc[1][col] : 0.0;
    double va[S], vb[S], vc[S], z1[S], z2[S];
#pragma unroll
    for (int t = 0; t < S; ++t) {
        z1[t] = oT[t];
        z2[t] = oS[t];
    }
    vmix_rows<S>(p, sdz, sdz + kMaxNz, L, [&](int k) { return km[k - L.lo + 1]; }, fT, fS, va, vb, vc, z1, z2);
    seg_solve2_f64<S>(L, va, vb, vc, z1, z2);
    if (!L.valid) return;
    const double dt = p.dt;
    double* __restrict__ dT = p.dvmix[0] + L.base;
    double* __restrict__ dS = p.dvmix[1] + L.base;
#pragma unroll
    for (int t = 0; t < S; ++t) {
        const int k = L.lo + t;
        const bool w = k >= L.k0 && k < nz;
        if (w) {
            T[cs * (L.base + k)] = z1[t];
            Sa[cs * (L.base + k)] = z2[t];
        }
        z1[t] = w ? (z1[t] - oT[t]) / dt : 0.0;  // (new - old) / dt; 0 where unchanged
        z2[t] = w ? (z2[t] - oS[t]) / dt : 0.0;
    }
    if (VEC) {
#pragma unroll
        for (int t = 0; t < S; t += 2) {
            const int k = L.lo + t;
            if (k < nz) {
                *reinterpret_cast<double2*>(dT + k) = make_double2(z1[t], z1[t + 1]);
                *reinterpret_cast<double2*>(dS + k) = make_double2(z2[t], z2[t + 1]);
            }
        }
    } else {
#pragma unroll
        for (int t = 0; t < S; ++t) {
            const int k = L.lo + t;
            if (k < nz) {
                dT[k] = z1[t];
                dS[k] = z2[t];
            }
        }
    }
}

// Stage the lane's levels v into its column's row of a wave image, then store
// the image's pairs selected by the plan, coalesced (a wave's LDS operations
// complete in order, so the image may be reused right after).
template <int S>
__device__ __forceinline__ void vmix_store_pairs(double* img, const SegCopyPlan<S>& cp, double* __restrict__ dst,
                                                 const SegLane& L, const double (&v)[S]) {
    double* row = img + ((threadIdx.x & 63) / kSegP) * (kSegP * S) + L.lo;
#pragma unroll
    for (int t = 0; t < S; t += 2) *reinterpret_cast<double2*>(row + t) = make_double2(v[t], v[t + 1]);
#pragma unroll
    for (int i = 0; i < SegCopyPlan<S>::NI; ++i) {
        if (cp.pred[i]) {
            const int pidx = i * 64 + (threadIdx.x & 63);
            *reinterpret_cast<double2*>(dst + cp.off[i]) = *reinterpret_cast<const double2*>(img + 2 * pidx);
        }
    }
}

// LDS bytes of one wave: temp and salt images of its 8 columns
__host__ __device__ constexpr int vmix_lds_wave_bytes(int S) { return 2 * kSegCols * kSegP * S * 8; }

// LDS variant (time levels outermost, nz even, kappaH, both tracer levels and
// both tendencies 16-byte aligned): the wave copies the wet pairs of its 8
// columns' temp and salt into LDS, each lane reads its levels there and loads
// its kappaH pairs itself (level lo-1 from the lane below), and the results go
// back through the two images: tracers on the pairs that touch the wet suffix
// (a pair's dry level gets its old value back), tendencies on every pair. Two
// images rather than three (kappaH staged too) fit six blocks per CU in LDS
// instead of five, which is what the registers allow (three waves per SIMD).
template <int S>
__global__ __launch_bounds__(kSegLdsBlock) void vertmix_segl_kernel(VmixArgs p) {
    extern __shared__ double vmix_lds[];
    __shared__ double sdz[2 * kMaxNz];
    vmix_stage_dz(p, sdz);
    constexpr int IMG = kSegCols * kSegP * S;
    double* lds = vmix_lds + (threadIdx.x / 64) * 2 * IMG;
    int64_t col;
    const SegLane L = vmix_lane(p, S, col);
    const int nz = p.nz;
    double* __restrict__ T = p.tr[0] + p.taup1 * p.ls;
    double* __restrict__ Sa = p.tr[1] + p.taup1 * p.ls;
    {
        SegCopyPlan<S> wet;
        seg_copy_plan<S>(L, threadIdx.x & 63, nz, true, wet);
        seg_lds_copy<S>(T, lds + 0 * IMG, wet);
        seg_lds_copy<S>(Sa, lds + 1 * IMG, wet);
    }
    // kappaH at levels lo-1 .. lo+S-1 (its value matters on the wet levels only)
    double kmw[S], km[S + 1];
#pragma unroll
    for (int t = 0; t < S; ++t) kmw[t] = 0.0;
    col_load<S, true>(p.kappaH + L.base, 1, L.lo, nz, [&](int k) { return k + 1 >= L.k0; }, kmw);
    const bool solve = L.k0 < nz;
    const double fT = solve ? p.forc[0][col] : 0.0, fS = solve ? p.forc[1][col] : 0.0;
    km[0] = grp_prev(L.s, kmw[S - 1]);
#pragma unroll
    for (int t = 0; t < S; ++t) km[t + 1] = kmw[t];
    seg_wait_vmem();
    double va[S], vb[S], vc[S], z1[S], z2[S];
    seg_lds_read<S>(lds + 0 * IMG, L, z1);
    seg_lds_read<S>(lds + 1 * IMG, L, z2);
    vmix_rows<S>(p, sdz, sdz + kMaxNz, L, [&](int k) { return km[k - L.lo + 1]; }, fT, fS, va, vb, vc, z1, z2);
    seg_solve2_f64<S>(L, va, vb, vc, z1, z2);
    // tracers: new values on the wet levels, the old ones (still in the
    // images) elsewhere
    const double dt = p.dt;
    double oT[S], oS[S];
    seg_lds_read<S>(lds + 0 * IMG, L, oT);
    seg_lds_read<S>(lds + 1 * IMG, L, oS);
#pragma unroll
    for (int t = 0; t < S; ++t) {
        const int k = L.lo + t;
        const bool w = k >= L.k0 && k < nz;
        const double nT = w ? z1[t] : oT[t], nS = w ? z2[t] : oS[t];
        z1[t] = nT;
        z2[t] = nS;
        oT[t] = w ? (nT - oT[t]) / dt : 0.0;  // (new - old) / dt; 0 where unchanged
        oS[t] = w ? (nS - oS[t]) / dt : 0.0;
    }
    // the copy plans are formed again here rather than kept through the solve
    SegCopyPlan<S> wet;
    seg_copy_plan<S>(L, threadIdx.x & 63, nz, true, wet);
    vmix_store_pairs<S>(lds + 0 * IMG, wet, T, L, z1);
    vmix_store_pairs<S>(lds + 1 * IMG, wet, Sa, L, z2);
    SegCopyPlan<S> all;
    seg_copy_plan<S>(L, threadIdx.x & 63, nz, false, all);
    vmix_store_pairs<S>(lds + 0 * IMG, all, p.dvmix[0], L, oT);
    vmix_store_pairs<S>(lds + 1 * IMG, all, p.dvmix[1], L, oS);
}

// enforce_boundaries on one process (veros/core/utilities.py:7-21): the x
// halos of level taup1, arr[-2:] = arr[2:4] and arr[:2] = arr[-4:-2].
__global__ __launch_bounds__(256) void vertmix_cyclic_kernel(VmixArgs p) {
    const int64_t plane = p.Y * p.nz;  // one x row of a level
    const int64_t n = 2 * 4 * plane;   // tracers x halo rows x row
    const int64_t e = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (e >= n) return;
    const int t = (int)(e / (4 * plane));
    const int64_t r = e - (int64_t)t * 4 * plane;
    const int h = (int)(r / plane);  // halo row 0, 1, X-2, X-1
    const int64_t off = r - (int64_t)h * plane;
    const int64_t dst = h < 2 ? h : p.X - 4 + h;
    const int64_t src = h < 2 ? p.X - 4 + h : 2 + (h - 2);
    double* __restrict__ tr = p.tr[t] + p.taup1 * p.ls;
    tr[p.cs * (dst * plane + off)] = tr[p.cs * (src * plane + off)];
}

template <int S>
void launch_vmix(const VmixArgs& p, bool lds, bool vec, hipStream_t s) {
    const int64_t threads = p.X * p.Y * kSegP;
    if (lds) {
        const size_t bytes = (size_t)vmix_lds_wave_bytes(S) * (kSegLdsBlock / 64);
        hipLaunchKernelGGL(vertmix_segl_kernel<S>, dim3((unsigned)((threads + kSegLdsBlock - 1) / kSegLdsBlock)),
                           dim3(kSegLdsBlock), bytes, s, p);
    } else if (vec) {
        hipLaunchKernelGGL((vertmix_seg_kernel<S, true>), dim3((unsigned)((threads + kSegBlock - 1) / kSegBlock)),
                           dim3(kSegBlock), 0, s, p);
    } else {
        hipLaunchKernelGGL((vertmix_seg_kernel<S, false>), dim3((unsigned)((threads + kSegBlock - 1) / kSegBlock)),
                           dim3(kSegBlock), 0, s, p);
    }
}

}  // namespace
}  // namespace vx

extern "C" int vx_vertmix_tempsalt_f64(void* stream, const vx_grid* grid, double dt_tracer, double* temp,
                                       double* salt, int32_t taup1, const double* kappaH,
                                       const double* forc_temp_surface, const double* forc_salt_surface,
                                       double* dtemp_vmix, double* dsalt_vmix, int32_t enable_cyclic_x) {
    using namespace vx;
    const char* who = "vx_vertmix_tempsalt_f64";
    if (!grid) return fail(VX_ERR_INVALID_ARGUMENT, "%s: null grid", who);
    if (grid->nx < 1 || grid->ny < 1 || grid->nz < 1)
        return fail(VX_ERR_INVALID_ARGUMENT, "%s: bad shape nx=%lld ny=%lld nz=%d", who, (long long)grid->nx,
                    (long long)grid->ny, grid->nz);
    if (grid->nz > kMaxNz) return fail(VX_ERR_UNSUPPORTED, "%s: nz=%d too deep (max %d)", who, grid->nz, kMaxNz);
    if (!grid->kbot || !grid->dzt || !grid->dzw) return fail(VX_ERR_INVALID_ARGUMENT, "%s: null grid array", who);
    if (taup1 < 0 || taup1 > 2) return fail(VX_ERR_INVALID_ARGUMENT, "%s: taup1 must be 0..2", who);
    if (!temp || !salt || !kappaH || !forc_temp_surface || !forc_salt_surface || !dtemp_vmix || !dsalt_vmix)
        return fail(VX_ERR_INVALID_ARGUMENT, "%s: null array pointer", who);
    if (temp == salt || dtemp_vmix == dsalt_vmix)
        return fail(VX_ERR_INVALID_ARGUMENT, "%s: temp/salt and their tendencies must be distinct arrays", who);
    if (!(dt_tracer > 0.0)) return fail(VX_ERR_INVALID_ARGUMENT, "%s: dt_tracer must be positive", who);
    // the halo copy runs as one parallel pass; its source and destination rows
    // are disjoint from nx = 4 on (the reference copies them one after the other)
    if (enable_cyclic_x && grid->nx < 4)
        return fail(VX_ERR_UNSUPPORTED, "%s: cyclic x needs nx >= 4 (nx=%lld)", who, (long long)grid->nx);
    VmixArgs p{};
    p.kbot = grid->kbot;
    p.dzt = grid->dzt;
    p.dzw = grid->dzw;
    p.kappaH = kappaH;
    p.tr[0] = temp;
    p.tr[1] = salt;
    p.forc[0] = forc_temp_surface;
    p.forc[1] = forc_salt_surface;
    p.dvmix[0] = dtemp_vmix;
    p.dvmix[1] = dsalt_vmix;
    p.X = grid->nx + 4;
    p.Y = grid->ny + 4;
    p.nz = grid->nz;
    p.taup1 = taup1;
    p.dt = dt_tracer;
    if (grid->time_layout != VX_LEVELS_INNER && grid->time_layout != VX_LEVELS_OUTER)
        return fail(VX_ERR_INVALID_ARGUMENT, "%s: time_layout must be VX_LEVELS_INNER or VX_LEVELS_OUTER", who);
    const bool outer = grid->time_layout == VX_LEVELS_OUTER;
    p.cs = outer ? 1 : 3;
    p.ls = outer ? p.X * p.Y * p.nz : 1;
    const bool vec = !(p.nz & 1) && aligned16(dtemp_vmix) && aligned16(dsalt_vmix);
    const bool lds = vec && outer && aligned16(kappaH) && aligned16(temp + taup1 * p.ls) &&
                     aligned16(salt + taup1 * p.ls);
    hipStream_t s = (hipStream_t)stream;
    switch (seg_levels(p.nz)) {
        case 2: launch_vmix<2>(p, lds, vec, s); break;
        case 4: launch_vmix<4>(p, lds, vec, s); break;
        case 6: launch_vmix<6>(p, lds, vec, s); break;
        case 8: launch_vmix<8>(p, lds, vec, s); break;
        case 10: launch_vmix<10>(p, lds, vec, s); break;
        case 12: launch_vmix<12>(p, lds, vec, s); break;
        case 14: launch_vmix<14>(p, lds, vec, s); break;
        default: launch_vmix<16>(p, lds, vec, s); break;
    }
    int rc = check_launch(who);
    if (rc) return rc;
    if (enable_cyclic_x) {
        const int64_t n = 2 * 4 * p.Y * p.nz;
        hipLaunchKernelGGL(vertmix_cyclic_kernel, dim3((unsigned)((n + 255) / 256)), dim3(256), 0, s, p);
        rc = check_launch(who);
    }
    return rc;
}
