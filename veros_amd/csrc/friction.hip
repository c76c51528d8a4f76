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
// interpolated onto the T grid, to K_diss_v. Here one lane per (face column,
// component) forms the rows in registers while it sweeps, and a second kernel
// forms the dissipation of the four faces around each T cell from the solved
// velocities and adds it to K_diss_v in one pass.
#include "vx_common.hpp"

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

constexpr int kFricBlock = 256;  // 128 face columns x 2 components
constexpr int kFricChunk = 8;    // levels per load batch (NZB is a multiple)

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

// Lane 2*col + t solves component t of face column col.
template <int NZB>
__global__ __launch_bounds__(kFricBlock) void friction_solve_kernel(FricArgs p) {
    const int64_t gid = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    const int t = (int)(gid & 1);
    const int64_t col = gid >> 1;
    if (col >= p.X * p.Y) return;
    const int64_t i = col / p.Y, j = col - i * p.Y;
    if (i < 1 || i > p.X - 3 || j < 1 || j > p.Y - 3) return;
    const int nz = p.nz;
    const int64_t nb = t == 0 ? col + p.Y : col + 1;  // the T column across the face
    const Face f = face_of(p.kbot[col], p.kbot[nb], nz);
    const int k0 = f.k0;
    const double dt = p.dt;
    const int64_t base = col * nz;
    const double* __restrict__ ka = p.kappaM + base;
    const double* __restrict__ kn = p.kappaM + nb * nz;
    double* __restrict__ ut = p.vel[t] + p.tau * p.ls;    // level tau of cell c at ut[cs c]
    double* __restrict__ u1 = p.vel[t] + p.taup1 * p.ls;  // level taup1
    double* __restrict__ dm = p.dmix[t] + base;

    // forward elimination with the rows formed on the fly
    // (veros/core/friction.py:111-126, veros/core/operators.py:60-77)
    double cp[NZB], dp[NZB];
    double cpp = 0.0, dpp = 0.0, dlm = 0.0;  // dlm = delta[k-1]
#pragma unroll
    for (int q = 0; q < NZB; q += kFricChunk) {
        double va[kFricChunk], vn[kFricChunk], vt[kFricChunk];
        const bool live = q < nz && q + kFricChunk > k0;
        if (live) {
#pragma unroll
            for (int u = 0; u < kFricChunk; ++u) {
                const int kk = q + u < nz ? q + u : nz - 1;
                va[u] = ka[kk];
                vn[u] = kn[kk];
                vt[u] = ut[p.cs * (base + kk)];
            }
        }
#pragma unroll
        for (int u = 0; u < kFricChunk; ++u) {
            const int k = q + u;
            if (live && k < nz && k >= k0) {
                const double fxa = 0.5 * (va[u] + vn[u]);
                const double dl = (f.both && k < nz - 1) ? dt / p.dzw[k] * fxa : 0.0;
                const double dz = p.dzt[k];
                const bool edge = k == k0;
                const double ak = edge ? 0.0 : -dlm / dz;
                const double bk = edge ? 1 + dl / dz : (1 + dlm / dz) + dl / dz;
                const double ck = -dl / dz;
                const double r = 1.0 / (bk - ak * cpp);
                cpp = ck * r;
                dpp = (vt[u] - ak * dpp) * r;
                dlm = dl;
            }
            cp[k] = cpp;
            dp[k] = dpp;
        }
    }
    // back substitution from the top, a chunk at a time; the velocity at
    // taup1 on the wet levels and du_mix = (taup1 - tau) / dt on every level
    // (:127-130), the tau values (and taup1 below the solve) re-read per chunk
    double xn = 0.0;
#pragma unroll
    for (int q = (NZB - 1) / kFricChunk * kFricChunk; q >= 0; q -= kFricChunk) {
        if (q < nz) {
            double vo[kFricChunk], vw[kFricChunk];
#pragma unroll
            for (int u = 0; u < kFricChunk; ++u) {
                const int64_t c3 = p.cs * (base + (q + u < nz ? q + u : nz - 1));
                vo[u] = ut[c3];
                vw[u] = q + u < k0 ? u1[c3] : 0.0;
            }
#pragma unroll
            for (int u = kFricChunk - 1; u >= 0; --u) {
                const int k = q + u;
                if (k < nz) {
                    double x = vw[u];
                    if (k >= k0) {
                        x = dp[k] - cp[k] * xn;
                        xn = x;
                        u1[p.cs * (base + k)] = x;
                    }
                    dm[k] = (x - vo[u]) / dt;
                }
            }
        }
    }
}

// The dissipation of one face at level k < nz-1 (veros/core/friction.py:
// 137-153, 189-201): fxa * (x1[k+1] - x1[k]) / dzw[k] * mask[k+1] * mask[k]
// times (x0[k+1] - x0[k]) / dzw[k], with x0 at tau and x1 at taup1. Faces
// outside [1, X-2) x [1, Y-2) dissipate nothing.
__device__ inline double face_diss(const FricArgs& p, int t, int64_t i, int64_t j, int k) {
    if (i < 1 || i > p.X - 3 || j < 1 || j > p.Y - 3) return 0.0;
    const int64_t col = i * p.Y + j;
    const int64_t nb = t == 0 ? col + p.Y : col + 1;
    const Face f = face_of(p.kbot[col], p.kbot[nb], p.nz);
    const double m = (f.both && k >= f.k0) ? 1.0 : 0.0;  // mask[k] * mask[k+1]
    const double fxa = 0.5 * (p.kappaM[col * p.nz + k] + p.kappaM[nb * p.nz + k]);
    const double* v = p.vel[t];
    const int64_t c0 = p.cs * (col * p.nz + k), c1 = c0 + p.cs;
    const int64_t l0 = p.tau * p.ls, l1 = p.taup1 * p.ls;
    const double dzw = p.dzw[k];
    const double flux = fxa * (v[c1 + l1] - v[c0 + l1]) / dzw * m * m;
    return (v[c1 + l0] - v[c0 + l0]) * flux / dzw;
}

// K_diss_v += ugrid_to_tgrid(diss_u), then += vgrid_to_tgrid(diss_v)
// (veros/core/friction.py:154-155, 203-204; veros/core/numerics.py:313-336):
// at T cell (i, j) the u faces i and i-1 weighted dxu / (2 dxt) for 2 <= i <
// X-2 (face i alone elsewhere), the v faces j and j-1 weighted area_v / (2
// area_t) for 2 <= j < Y-2. Thread per (T cell column, level); the top level
// dissipates nothing.
__global__ __launch_bounds__(256) void friction_diss_kernel(FricArgs p) {
    const int nz1 = p.nz - 1;
    const int64_t e = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    const int64_t ncols = (p.X - 3) * (p.Y - 3);  // T columns [1, X-2) x [1, Y-2)
    if (e >= ncols * nz1) return;
    const int64_t c = e / nz1;
    const int k = (int)(e - c * nz1);
    const int64_t i = 1 + c / (p.Y - 3), j = 1 + c % (p.Y - 3);
    const double du_i = face_diss(p, 0, i, j, k);
    double bu = du_i;
    if (i >= 2 && i <= p.X - 3) {
        const double du_m = face_diss(p, 0, i - 1, j, k);
        bu = (p.dxu[i] * du_i + p.dxu[i - 1] * du_m) / (2 * p.dxt[i]);
    }
    const double dv_j = face_diss(p, 1, i, j, k);
    double bv = dv_j;
    if (j >= 2 && j <= p.Y - 3) {
        const double dv_m = face_diss(p, 1, i, j - 1, k);
        bv = (p.area_v[i * p.Y + j] * dv_j + p.area_v[i * p.Y + j - 1] * dv_m) / (2 * p.area_t[i * p.Y + j]);
    }
    double* K = p.K_diss_v + (i * p.Y + j) * p.nz + k;
    *K = (*K + bu) + bv;
}

template <int NZB>
void launch_fric(const FricArgs& p, hipStream_t s) {
    const int64_t lanes = 2 * p.X * p.Y;
    hipLaunchKernelGGL(friction_solve_kernel<NZB>, dim3((unsigned)((lanes + kFricBlock - 1) / kFricBlock)),
                       dim3(kFricBlock), 0, s, p);
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
    hipStream_t s = (hipStream_t)stream;
    const int nzb = ((p.nz + 15) / 16) * 16;
    switch (nzb) {
        case 16: launch_fric<16>(p, s); break;
        case 32: launch_fric<32>(p, s); break;
        case 48: launch_fric<48>(p, s); break;
        case 64: launch_fric<64>(p, s); break;
        case 80: launch_fric<80>(p, s); break;
        case 96: launch_fric<96>(p, s); break;
        case 112: launch_fric<112>(p, s); break;
        default: launch_fric<128>(p, s); break;
    }
    int rc = check_launch(who);
    if (rc) return rc;
    if (p.nz > 1) {
        const int64_t n = (p.X - 3) * (p.Y - 3) * (int64_t)(p.nz - 1);
        hipLaunchKernelGGL(friction_diss_kernel, dim3((unsigned)((n + 255) / 256)), dim3(256), 0, s, p);
        rc = check_launch(who);
    }
    return rc;
}
