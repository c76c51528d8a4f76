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
// tdma_seg.hpp, lane s holding levels [s*S, s*S + S)) and treats its u face,
// then its v face: each lane forms its rows in registers from the face
// viscosity, the group solves, and each lane writes the new velocity on the
// wet levels, the tendency on every level, and the face's dissipation (from
// the velocities still in registers; the level above comes from the next
// lane) into a workspace. A second kernel adds the dissipation of the four
// faces around each T cell to K_diss_v in the reference's order, its blocks
// remapped so that each XCD walks a contiguous range of rows and the faces of
// row i-1 are still in its L2.
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
    double* diss[2];          // workspace: dissipation of the u and v faces (X, Y, nz)
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
// the velocity at taup1 on the wet levels and the tendency on every level
// (:127-130, 178-181), and the face's dissipation on every level (:137-153,
// 189-201): fxa * (x1[k+1] - x1[k]) / dzw[k] * mask[k+1] * mask[k] times
// (x0[k+1] - x0[k]) / dzw[k], x0 at tau and x1 at taup1, 0 on the top level.
// The masks are 0 or 1, so the products are taken as selects.
template <int S, bool VEC>
__device__ __forceinline__ void fric_face(const FricArgs& p, SegLane L, int t, int64_t col, int64_t nb, bool live,
                                          bool both) {
    const int nz = p.nz;
    const int lo = L.lo;
    const int k0 = L.k0;
    const double dt = p.dt;
    const int64_t base = col * nz;
    // face viscosity on the solved levels
    double ka[S], kn[S], x0[S], x1[S];
#pragma unroll
    for (int q = 0; q < S; ++q) ka[q] = kn[q] = x0[q] = x1[q] = 0.0;
    col_load<S, VEC>(p.kappaM + base, 1, lo, nz, [&](int k) { return k >= k0; }, ka);
    col_load<S, VEC>(p.kappaM + nb * nz, 1, lo, nz, [&](int k) { return k >= k0; }, kn);
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
    double dlm = (both && kp >= k0 && kp < nz - 1) ? dt / p.dzw[kp] * fxm : 0.0;
#pragma unroll
    for (int q = 0; q < S; ++q) {
        const int k = lo + q;
        const bool wet = k >= k0 && k < nz;
        const int kc = k < nz ? k : nz - 1;
        const double dz = p.dzt[kc];
        const double dl = (both && wet && k < nz - 1) ? dt / p.dzw[kc] * fxa[q] : 0.0;
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
    if (!live) return;
#pragma unroll
    for (int q = 0; q < S; ++q) {
        const int k = lo + q;
        if (k >= k0 && k < nz) u1p[p.cs * k] = x1[q];
    }
    double dm[S], ds[S];
#pragma unroll
    for (int q = 0; q < S; ++q) {
        const int k = lo + q;
        const double a1 = q < S - 1 ? x1[q < S - 1 ? q + 1 : 0] : x1n;
        const double a0 = q < S - 1 ? x0[q < S - 1 ? q + 1 : 0] : x0n;
        const int kc = k < nz - 1 ? k : 0;
        const double dzw = p.dzw[kc];
        ds[q] = (both && k >= k0 && k < nz - 1) ? ((a0 - x0[q]) * ((fxa[q] * (a1 - x1[q])) / dzw)) / dzw : 0.0;
        dm[q] = (x1[q] - x0[q]) / dt;
    }
    col_store<S, VEC>(p.dmix[t] + base, 1, lo, nz, dm);
    col_store<S, VEC>(p.diss[t] + base, 1, lo, nz, ds);
}

// VEC: levels outermost, nz even and every array 16-byte aligned.
template <int S, bool VEC>
__global__ __launch_bounds__(kSegBlock) void friction_seg_kernel(FricArgs p) {
    const int64_t gt = xcd_contiguous_block() * blockDim.x + threadIdx.x;
    int64_t col = gt / kSegP;
    SegLane L;
    L.s = (int)(gt % kSegP);
    L.lo = L.s * S;
    L.valid = col < p.X * p.Y;
    if (!L.valid) col = 0;
    const int64_t i = col / p.Y, j = col - (col / p.Y) * p.Y;
    const bool live = L.valid && i >= 1 && i <= p.X - 3 && j >= 1 && j <= p.Y - 3;
    const int kb = live ? p.kbot[col] : 0;
    L.base = col * p.nz;
#pragma unroll 1
    for (int t = 0; t < 2; ++t) {
        const int64_t nb = live ? (t == 0 ? col + p.Y : col + 1) : col;  // the T column across the face
        const Face f = face_of(kb, live ? p.kbot[nb] : 0, p.nz);
        L.k0 = f.k0;
        fric_face<S, VEC>(p, L, t, col, nb, live, f.both);
    }
}

// K_diss_v += ugrid_to_tgrid(diss_u), then += vgrid_to_tgrid(diss_v)
// (veros/core/friction.py:154-155, 203-204; veros/core/numerics.py:313-336):
// at T cell (i, j) the u faces i and i-1 weighted dxu / (2 dxt) for 2 <= i <
// X-2 (face i alone elsewhere), the v faces j and j-1 weighted area_v / (2
// area_t) for 2 <= j < Y-2. Thread per (T cell column, level) of [1, X-2) x
// [1, Y-2); the top level dissipates nothing. The grid is a multiple of 8
// blocks, remapped so that XCD x takes the x-th eighth of the cells.
__global__ __launch_bounds__(256) void friction_gather_kernel(FricArgs p) {
    const int nz1 = p.nz - 1;
    const int64_t e = xcd_contiguous_block() * blockDim.x + threadIdx.x;
    const int64_t ncols = (p.X - 3) * (p.Y - 3);  // T columns [1, X-2) x [1, Y-2)
    if (e >= ncols * nz1) return;
    const int64_t c = e / nz1;
    const int k = (int)(e - c * nz1);
    const int64_t i = 1 + c / (p.Y - 3), j = 1 + c % (p.Y - 3);
    const int64_t ij = i * p.Y + j;
    const int64_t cell = ij * p.nz + k;
    const double du_i = p.diss[0][cell];
    double bu = du_i;
    if (i >= 2) bu = (p.dxu[i] * du_i + p.dxu[i - 1] * p.diss[0][cell - p.Y * p.nz]) / (2 * p.dxt[i]);
    const double dv_j = p.diss[1][cell];
    double bv = dv_j;
    if (j >= 2) bv = (p.area_v[ij] * dv_j + p.area_v[ij - 1] * p.diss[1][cell - p.nz]) / (2 * p.area_t[ij]);
    double* K = p.K_diss_v + cell;
    *K = (*K + bu) + bv;
}

template <int S>
void launch_fric(const FricArgs& p, bool vec, hipStream_t s) {
    const int64_t threads = p.X * p.Y * kSegP;
    const dim3 grid(blocks_mult8(threads, kSegBlock));
    if (vec)
        hipLaunchKernelGGL((friction_seg_kernel<S, true>), grid, dim3(kSegBlock), 0, s, p);
    else
        hipLaunchKernelGGL((friction_seg_kernel<S, false>), grid, dim3(kSegBlock), 0, s, p);
}

int64_t fric_ws_elems(int64_t X, int64_t Y, int nz) { return X * Y * (int64_t)nz; }

}  // namespace
}  // namespace vx

extern "C" int64_t vx_implicit_vert_friction_workspace_bytes(const vx_grid* grid) {
    if (!grid || grid->nx < 1 || grid->ny < 1 || grid->nz < 1) return 0;
    const int64_t n = vx::fric_ws_elems(grid->nx + 4, grid->ny + 4, grid->nz);
    return 2 * ((n * 8 + 255) / 256 * 256);
}

extern "C" int vx_implicit_vert_friction_f64(void* stream, const vx_grid* grid, double dt_mom, double* u, double* v,
                                             int32_t tau, int32_t taup1, const double* kappaM, const double* area_t,
                                             const double* area_v, double* du_mix, double* dv_mix,
                                             double* K_diss_v, void* workspace) {
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
    if (!u || !v || !kappaM || !area_t || !area_v || !du_mix || !dv_mix || !K_diss_v || !workspace)
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
    const int64_t plane = (vx::fric_ws_elems(p.X, p.Y, p.nz) * 8 + 255) / 256 * 256;
    p.diss[0] = (double*)workspace;
    p.diss[1] = (double*)((char*)workspace + plane);
    const bool vec = outer && !(p.nz & 1) && aligned16(u) && aligned16(v) && aligned16(kappaM) &&
                     aligned16(du_mix) && aligned16(dv_mix) && aligned16(workspace);
    hipStream_t s = (hipStream_t)stream;
    switch (seg_levels(p.nz)) {
        case 2: launch_fric<2>(p, vec, s); break;
        case 4: launch_fric<4>(p, vec, s); break;
        case 6: launch_fric<6>(p, vec, s); break;
        case 8: launch_fric<8>(p, vec, s); break;
        case 10: launch_fric<10>(p, vec, s); break;
        case 12: launch_fric<12>(p, vec, s); break;
        case 14: launch_fric<14>(p, vec, s); break;
        default: launch_fric<16>(p, vec, s); break;
    }
    int rc = check_launch(who);
    if (rc) return rc;
    if (p.nz > 1) {
        const int64_t n = (p.X - 3) * (p.Y - 3) * (int64_t)(p.nz - 1);
        hipLaunchKernelGGL(friction_gather_kernel, dim3(blocks_mult8(n, 256)), dim3(256), 0, s, p);
        rc = check_launch(who);
    }
    return rc;
}
