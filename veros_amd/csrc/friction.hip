// Implicit vertical friction of the horizontal velocity for gfx950 (float64):
// implicit_vert_friction, veros/core/friction.py:90-205.
//
// For each velocity face column of [1, X-2) x [1, Y-2) the reference builds a
// tridiagonal system from kappaM averaged onto the face (delta = dt_mom / dzw
// * fxa * maskU[k+1] * maskU[k], a = -delta[k-1] / dzt, b = 1 + delta[k-1] /
// dzt + delta[k] / dzt, c = -delta / dzt, the edge row b_edge = 1 + delta /
// dzt), solves it with solve_implicit over the wet suffix k >= max(kbot,
// kbot_next) - 1, stores the result at taup1, forms du_mix / dv_mix = (new -
// tau) / dt_mom over the face range, and adds the dissipation of the solve,
// interpolated onto the T grid, to K_diss_v.
//
// Here a group of kSegP lanes owns a face column (the segmented solve of
// tdma_seg.hpp, lane s holding levels [s*S, s*S + S)): each lane forms its
// rows in registers from the face viscosity, the group solves, and each lane
// writes the new velocity on the wet levels and the tendency on every level,
// and forms the face's dissipation from the velocities still in registers
// (the level above from the next lane). A T cell's dissipation needs the u
// faces i-1 and i and the v faces j-1 and j, so a block owns a strip of T
// columns and walks a chunk of rows in order: the u faces of the previous row
// are kept in LDS, and the v faces of the row pass through LDS to the group
// north of them. No face dissipation goes through memory.
#include "tdma_seg.hpp"

namespace vx {
namespace {

struct FricArgs {
    const int32_t* kbot;      // (X, Y)
    const double* dzt;        // (nz)
    const double* dzw;        // (nz)
    const double* dxt;        // (X)
    const double* dxu;        // (X)
    const double* area_t;     // (X, Y)
    const double* area_v;     // (X, Y)
    const double* kappaM;     // (X, Y, nz)
    double* vel[2];           // u, v (X, Y, nz, 3)
    double* dmix[2];          // du_mix, dv_mix (X, Y, nz)
    double* K_diss_v;         // (X, Y, nz)
    int64_t X, Y;
    int32_t nz, tau, taup1;
    double dt;
    // time levels of u, v: level l of cell c at l * ls + c * cs (cs = 3,
    // ls = 1: levels innermost; cs = 1, ls = X*Y*nz: levels outermost)
    int64_t cs, ls;
};

// The face of component t (0: u, between T columns i and i+1; 1: v, between
// rows j and j+1) is wet at level k iff both T columns are wet there:
// k >= max(kbot, kbot_next) - 1 with both kbot > 0 (maskU / maskV as
// veros/core/numerics.py:195-221 builds them from kbot). The solve runs over
// k >= max(kbot, kbot_next) - 1 whenever that maximum is positive
// (create_water_masks of the max, veros/core/friction.py:108-109, 158-159), so
// a face with one dry side is solved with delta = 0 and keeps tau's value.
struct Face {
    int k0;      // first level of the solve (nz: no solve)
    bool both;   // both T columns wet somewhere: the face mask is k >= k0
};

__device__ inline Face face_of(int kb0, int kb1, int nz) {
    const int kss = kb0 > kb1 ? kb0 : kb1;
    Face f;
    f.k0 = (kss >= 1 && kss <= nz) ? kss - 1 : nz;
    f.both = kb0 > 0 && kb1 > 0;
    return f;
}

// One face column of component t (face range only when `live`): rows from
// the face viscosity (veros/core/friction.py:111-126, 161-176), the solve,
// with `store` the velocity at taup1 on the wet levels and the tendency on
// every level (:127-130, 178-181), and in ds the face's dissipation on every
// level (:137-153, 189-201): fxa * (x1[k+1] - x1[k]) / dzw[k] * mask[k+1] *
// mask[k] times (x0[k+1] - x0[k]) / dzw[k], x0 at tau and x1 at taup1, 0 on
// the top level. The masks are 0 or 1, so the products are taken as selects.
// Every lane of the wave calls it (the solve moves data across lanes). zt and
// zw hold dzt and dzw in LDS: read inside per-level selects, a global load
// would be waited for at once, level after level.
template <int S, bool VEC, bool PV>
__device__ __forceinline__ void fric_face(const FricArgs& p, const double* __restrict__ zt,
                                          const double* __restrict__ zw, SegLane L, int t, int64_t col, int64_t nb,
                                          bool live, bool both, bool store, double (&ds)[S]) {
    const int nz = p.nz;
    const int lo = L.lo;
    const int k0 = L.k0;
    const double dt = p.dt;
    const int64_t base = col * nz;
    // face viscosity on the solved levels
    double ka[S], kn[S], x0[S], x1[S];
#pragma unroll
    for (int q = 0; q < S; ++q) ka[q] = kn[q] = x0[q] = x1[q] = 0.0;
    col_load<S, PV>(p.kappaM + base, 1, lo, nz, [&](int k) { return k >= k0; }, ka);
    col_load<S, PV>(p.kappaM + nb * nz, 1, lo, nz, [&](int k) { return k >= k0; }, kn);
    // tau on every level, taup1 below the solve (kept there: du_mix reads it)
    const double* u0p = p.vel[t] + p.tau * p.ls + p.cs * base;
    double* u1p = p.vel[t] + p.taup1 * p.ls + p.cs * base;
    col_load<S, VEC>(u0p, p.cs, lo, nz, [&](int) { return live; }, x0);
    col_load<S, VEC>(u1p, p.cs, lo, nz, [&](int k) { return live && k < k0; }, x1);
    double fxa[S];
#pragma unroll
    for (int q = 0; q < S; ++q) fxa[q] = 0.5 * (ka[q] + kn[q]);
    // rows; delta at level lo-1 from the lane below
    double va[S], vb[S], vc[S], z[S];
    const double fxm = grp_prev(L.s, fxa[S - 1]);
    const int kp = lo - 1;
    double dlm = (both && kp >= k0 && kp < nz - 1) ? dt / zw[kp] * fxm : 0.0;
#pragma unroll
    for (int q = 0; q < S; ++q) {
        const int k = lo + q;
        const bool wet = k >= k0 && k < nz;
        const int kc = k < nz ? k : nz - 1;
        const double dz = zt[kc];
        const double dl = (both && wet && k < nz - 1) ? dt / zw[kc] * fxa[q] : 0.0;
        const bool edge = k == k0;
        va[q] = (wet && !edge) ? -dlm / dz : 0.0;
        vb[q] = wet ? (edge ? 1 + dl / dz : (1 + dlm / dz) + dl / dz) : 1.0;
        vc[q] = wet ? -dl / dz : 0.0;
        z[q] = wet ? x0[q] : 0.0;
        dlm = dl;
    }
    seg_solve_f64<S>(L, va, vb, vc, z);
#pragma unroll
    for (int q = 0; q < S; ++q) {
        const int k = lo + q;
        x1[q] = (k >= k0 && k < nz) ? z[q] : x1[q];
    }
    // level lo+S from the lane above (every lane takes part in the moves)
    const double x1n = grp_next(L.s, x1[0]);
    const double x0n = grp_next(L.s, x0[0]);
#pragma unroll
    for (int q = 0; q < S; ++q) {
        const int k = lo + q;
        const double a1 = q < S - 1 ? x1[q < S - 1 ? q + 1 : 0] : x1n;
        const double a0 = q < S - 1 ? x0[q < S - 1 ? q + 1 : 0] : x0n;
        const int kc = k < nz - 1 ? k : 0;
        const double dzw = zw[kc];
        ds[q] = (both && k >= k0 && k < nz - 1) ? ((a0 - x0[q]) * ((fxa[q] * (a1 - x1[q])) / dzw)) / dzw : 0.0;
    }
    if (!(live && store)) return;
    bool st[S];
    double dm[S];
#pragma unroll
    for (int q = 0; q < S; ++q) {
        const int k = lo + q;
        st[q] = k >= k0;
        dm[q] = (x1[q] - x0[q]) / dt;
    }
    // the levels below the solve hold taup1's own values, so a pair store
    // that covers one of them writes back what is there
    col_store_if<S, VEC>(u1p, p.cs, lo, nz, x1, st);
    col_store<S, PV>(p.dmix[t] + base, 1, lo, nz, dm);
}

// A block of 32 face-column groups owns a strip of 31 T columns j (group g at
// column jb - 1 + g; group 0 solves the v face south of the strip for group
// 1) and walks the rows of a chunk [ib, ie) in order, starting one row early
// for the u faces west of the chunk. Per row: the v face, then the u face of
// each column. The u-face dissipation of the previous row stays in LDS, each
// lane reading back only what it wrote; the v-face dissipation of the row
// goes through LDS to the group north of it (two buffers, one barrier per
// row). Each face is stored by exactly one block. K_diss_v +=
// ugrid_to_tgrid(diss_u), then += vgrid_to_tgrid(diss_v)
// (veros/core/friction.py:154-155, 203-204; veros/core/numerics.py:313-336):
// at T cell (i, j) the u faces i and i-1 weighted dxu / (2 dxt) for 2 <= i <
// X-2 (face i alone for i = 1), the v faces j and j-1 weighted area_v / (2
// area_t) for 2 <= j < Y-2 (face j alone for j = 1), formed as (K + u part)
// + v part on the levels
// below the top one (which dissipates nothing). Blocks are XCD-contiguous,
// strips fastest. Two waves per SIMD: 218 VGPRs at nz = 80.
// VEC: levels outermost, nz even and every array 16-byte aligned. PV (implied
// by VEC): nz even and the arrays without time levels (kappaM, du_mix, dv_mix,
// K_diss_v) 16-byte aligned, which then move in level pairs whatever the
// layout of u and v.
template <int S, bool VEC, bool PV>
__global__ __launch_bounds__(kSegBlock) __attribute__((amdgpu_waves_per_eu(2))) void friction_walk_kernel(FricArgs p, int rows, int64_t nstrips,
                                                                  int64_t nblocks) {
    __shared__ double dsu[S * kSegBlock];
    __shared__ double dsv[2][S * kSegBlock];
    const int64_t b = xcd_contiguous_block();
    if (b >= nblocks) return;
    const int tid = threadIdx.x;
    __shared__ double sdz[2 * kMaxNz];  // dzt, then dzw
    for (int k = tid; k < p.nz; k += kSegBlock) {
        sdz[k] = p.dzt[k];
        sdz[kMaxNz + k] = p.dzw[k];
    }
    __syncthreads();
    const double* zt = sdz;
    const double* zw = sdz + kMaxNz;
    const int g = tid / kSegP;
    const int64_t strip = b % nstrips, chunk = b / nstrips;
    const int64_t j = 31 * strip + g;  // the strip's T columns are [31 strip + 1, 31 strip + 32)
    const int64_t ib = 1 + (int64_t)rows * chunk;
    const int64_t ie = ib + rows < p.X - 2 ? ib + rows : p.X - 2;
    const bool jlive = j >= 1 && j <= p.Y - 3;
    const bool jown = g >= 1 && jlive;
    const int nz = p.nz;
    SegLane L;
    L.s = tid % kSegP;
    L.lo = L.s * S;
    L.valid = true;
#pragma unroll 1
    for (int64_t i = ib >= 2 ? ib - 1 : ib; i < ie; ++i) {
        // the grid spacings are loaded again in each face (kept across the
        // row loop they would double the registers of the solve)
        asm volatile("" ::: "memory");
        const bool pre = i < ib;
        const int64_t col = jlive ? i * p.Y + j : 0;
        L.base = col * nz;
        const int kb = jlive ? p.kbot[col] : 0;
        const int buf = (int)(i & 1);
        if (!pre) {
            const int64_t nb = jlive ? col + 1 : col;
            const Face f = face_of(kb, jlive ? p.kbot[nb] : 0, nz);
            L.k0 = f.k0;
            double ds[S];
            fric_face<S, VEC, PV>(p, zt, zw, L, 1, col, nb, jlive, f.both, jown, ds);
#pragma unroll
            for (int q = 0; q < S; ++q) dsv[buf][q * kSegBlock + tid] = ds[q];
        }
        double bu[S];
        asm volatile("" ::: "memory");
        {
            const bool ulive = jown;
            const int64_t nb = ulive ? col + p.Y : col;
            const Face f = face_of(ulive ? kb : 0, ulive ? p.kbot[nb] : 0, nz);
            L.k0 = f.k0;
            double ds[S];
            fric_face<S, VEC, PV>(p, zt, zw, L, 0, col, nb, ulive, f.both, !pre, ds);
            const double wi = p.dxu[i], wm = p.dxu[i - 1], den = 2 * p.dxt[i];
#pragma unroll
            for (int q = 0; q < S; ++q) {
                bu[q] = i >= 2 ? (wi * ds[q] + wm * dsu[q * kSegBlock + tid]) / den : ds[q];
                dsu[q * kSegBlock + tid] = ds[q];
            }
        }
        if (pre) continue;
        // K_diss_v and the weights are loaded before the barrier, which their
        // latency then overlaps
        double K[S];
#pragma unroll
        for (int q = 0; q < S; ++q) K[q] = 0.0;
        double* Kp = p.K_diss_v + L.base;
        double wj = 0.0, wjm = 0.0, dnj = 1.0;
        if (jown) {
            col_load<S, PV>(Kp, 1, L.lo, nz, [&](int k) { return k < nz - 1; }, K);
            wj = p.area_v[col], wjm = p.area_v[col - 1], dnj = 2 * p.area_t[col];
        }
        __syncthreads();
        if (jown) {
            bool st[S];
#pragma unroll
            for (int q = 0; q < S; ++q) {
                const double dv = dsv[buf][q * kSegBlock + tid];
                const double bv = j >= 2 ? (wj * dv + wjm * dsv[buf][q * kSegBlock + tid - kSegP]) / dnj : dv;
                st[q] = L.lo + q < nz - 1;
                K[q] = st[q] ? (K[q] + bu[q]) + bv : K[q];
            }
            col_store_if<S, PV>(Kp, 1, L.lo, nz, K, st);
        }
    }
}

#ifndef VX_FRIC_ROWS
#define VX_FRIC_ROWS 24  // rows per block of the walk kernel
#endif

template <int S>
void launch_fric_walk(const FricArgs& p, bool vec, bool pv, hipStream_t s) {
    const int rows = VX_FRIC_ROWS;
    const int64_t nstrips = (p.Y - 3 + 30) / 31, nchunks = (p.X - 3 + rows - 1) / rows;
    const int64_t nblocks = nstrips * nchunks;
    const dim3 grid((unsigned)((nblocks + 7) / 8 * 8));
    if (vec)
        hipLaunchKernelGGL((friction_walk_kernel<S, true, true>), grid, dim3(kSegBlock), 0, s, p, rows, nstrips,
                           nblocks);
    else if (pv)
        hipLaunchKernelGGL((friction_walk_kernel<S, false, true>), grid, dim3(kSegBlock), 0, s, p, rows, nstrips,
                           nblocks);
    else
        hipLaunchKernelGGL((friction_walk_kernel<S, false, false>), grid, dim3(kSegBlock), 0, s, p, rows, nstrips,
                           nblocks);
}

}  // namespace
}  // namespace vx

extern "C" int vx_implicit_vert_friction_f64(void* stream, const vx_grid* grid, double dt_mom, double* u, double* v,
                                             int32_t tau, int32_t taup1, const double* kappaM, const double* area_t,
                                             const double* area_v, double* du_mix, double* dv_mix,
                                             double* K_diss_v) {
    using namespace vx;
    const char* who = "vx_implicit_vert_friction_f64";
    if (!grid) return fail(VX_ERR_INVALID_ARGUMENT, "%s: null grid", who);
    if (grid->nx < 1 || grid->ny < 1 || grid->nz < 1)
        return fail(VX_ERR_INVALID_ARGUMENT, "%s: bad shape nx=%lld ny=%lld nz=%d", who, (long long)grid->nx,
                    (long long)grid->ny, grid->nz);
    if (grid->nz > kMaxNz) return fail(VX_ERR_UNSUPPORTED, "%s: nz=%d too deep (max %d)", who, grid->nz, kMaxNz);
    if (!grid->kbot || !grid->dzt || !grid->dzw || !grid->dxt || !grid->dxu)
        return fail(VX_ERR_INVALID_ARGUMENT, "%s: null grid array", who);
    if (tau < 0 || tau > 2 || taup1 < 0 || taup1 > 2 || tau == taup1)
        return fail(VX_ERR_INVALID_ARGUMENT, "%s: tau and taup1 must be distinct levels in 0..2", who);
    if (!u || !v || !kappaM || !area_t || !area_v || !du_mix || !dv_mix || !K_diss_v)
        return fail(VX_ERR_INVALID_ARGUMENT, "%s: null array pointer", who);
    if (u == v || du_mix == dv_mix || (const double*)K_diss_v == kappaM || K_diss_v == du_mix || K_diss_v == dv_mix)
        return fail(VX_ERR_INVALID_ARGUMENT, "%s: u/v, du_mix/dv_mix and K_diss_v must be distinct arrays", who);
    if (!(dt_mom > 0.0)) return fail(VX_ERR_INVALID_ARGUMENT, "%s: dt_mom must be positive", who);
    FricArgs p{};
    p.kbot = grid->kbot;
    p.dzt = grid->dzt;
    p.dzw = grid->dzw;
    p.dxt = grid->dxt;
    p.dxu = grid->dxu;
    p.area_t = area_t;
    p.area_v = area_v;
    p.kappaM = kappaM;
    p.vel[0] = u;
    p.vel[1] = v;
    p.dmix[0] = du_mix;
    p.dmix[1] = dv_mix;
    p.K_diss_v = K_diss_v;
    p.X = grid->nx + 4;
    p.Y = grid->ny + 4;
    p.nz = grid->nz;
    p.tau = tau;
    p.taup1 = taup1;
    p.dt = dt_mom;
    if (grid->time_layout != VX_LEVELS_INNER && grid->time_layout != VX_LEVELS_OUTER)
        return fail(VX_ERR_INVALID_ARGUMENT, "%s: time_layout must be VX_LEVELS_INNER or VX_LEVELS_OUTER", who);
    const bool outer = grid->time_layout == VX_LEVELS_OUTER;
    p.cs = outer ? 1 : 3;
    p.ls = outer ? p.X * p.Y * p.nz : 1;
    const bool pv = !(p.nz & 1) && aligned16(kappaM) && aligned16(du_mix) && aligned16(dv_mix) && aligned16(K_diss_v);
    const bool vec = pv && outer && aligned16(u) && aligned16(v);
    hipStream_t s = (hipStream_t)stream;
    switch (seg_levels(p.nz)) {
        case 2: launch_fric_walk<2>(p, vec, pv, s); break;
        case 4: launch_fric_walk<4>(p, vec, pv, s); break;
        case 6: launch_fric_walk<6>(p, vec, pv, s); break;
        case 8: launch_fric_walk<8>(p, vec, pv, s); break;
        case 10: launch_fric_walk<10>(p, vec, pv, s); break;
        case 12: launch_fric_walk<12>(p, vec, pv, s); break;
        case 14: launch_fric_walk<14>(p, vec, pv, s); break;
        default: launch_fric_walk<16>(p, vec, pv, s); break;
    }
    return check_launch(who);
}
