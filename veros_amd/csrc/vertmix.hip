// Implicit vertical mixing of temperature and salinity for gfx950 (float64):
// vertmix_tempsalt, veros/core/thermodynamics.py:248-300.
//
// The reference builds the tridiagonal system of each interior column from
// kappaH and the vertical grid (delta = dt / dzw * kappaH, a = -delta[k-1] /
// dzt, b = 1 + (delta[k] + delta[k-1]) / dzt, c = -delta / dzt, the edge row
// b_edge = 1 + delta / dzt), adds the surface forcing to the top level of the
// right-hand side, solves it with solve_implicit for temp and for salt, and
// then forms the tendencies (new - old) / dt over the whole array and copies
// the cyclic x halos. Here one lane per (column, tracer) forms the rows in
// registers while it sweeps, so the coefficients never touch memory (the
// reference's six coefficient passes and the separate tendency pass fold into
// the solve): it reads kappaH and the tracer column once and writes the tracer
// and the tendency once.
#include "vx_common.hpp"

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
    bool vec_dv;  // nz even and both tendency arrays 16-byte aligned
    // time levels of temp/salt: level l of cell c at l * ls + c * cs (cs = 3,
    // ls = 1: levels innermost; cs = 1, ls = X*Y*nz: levels outermost)
    int64_t cs, ls;
};

constexpr int kVmixBlock = 256;  // 128 columns x 2 tracers
constexpr int kVmixChunk = 8;    // levels per load batch (NZB is a multiple)

// Lane 2*col + t solves tracer t of column col (the two lanes of a column read
// the same kappaH addresses in the same instruction). Columns outside the
// interior [2, X-2) x [2, Y-2) and land columns only write their zero
// tendencies: the reference's (tr - tr) / dt there.
template <int NZB>
__global__ __launch_bounds__(kVmixBlock) void vertmix_kernel(VmixArgs p) {
    const int64_t gid = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    const int t = (int)(gid & 1);
    const int64_t col = gid >> 1;
    if (col >= p.X * p.Y) return;
    const int64_t i = col / p.Y, j = col - (col / p.Y) * p.Y;
    const int nz = p.nz;
    const int64_t base = col * nz;
    double* __restrict__ dv = p.dvmix[t] + base;
    const bool interior = i >= 2 && i < p.X - 2 && j >= 2 && j < p.Y - 2;
    const int kb = interior ? p.kbot[col] : 0;
    const int k0 = (kb >= 1 && kb <= nz) ? kb - 1 : nz;  // first wet level (nz: none)
    const double dt = p.dt;
    double* __restrict__ tr = p.tr[t] + p.taup1 * p.ls;  // level taup1 of cell c at tr[cs c]
    const double* __restrict__ kap = p.kappaH + base;
    const bool vec = p.vec_dv;

    // forward elimination with the rows formed on the fly
    // (veros/core/thermodynamics.py:267-276, veros/core/operators.py:60-77);
    // each chunk of kVmixChunk levels is loaded in one batch (levels past nz-1
    // re-read level nz-1, dry levels are read and ignored) before its sweep
    double cp[NZB], dp[NZB];
    double cpp = 0.0, dpp = 0.0, dlm = 0.0;  // dlm = delta[k-1]
    const double forc = k0 < nz ? p.forc[t][col] : 0.0;
#pragma unroll
    for (int q = 0; q < NZB; q += kVmixChunk) {
        double vk[kVmixChunk], vt[kVmixChunk];
        const bool live = q < nz && q + kVmixChunk > k0;
        if (live) {
#pragma unroll
            for (int u = 0; u < kVmixChunk; ++u) {
                const int kk = q + u < nz ? q + u : nz - 1;
                vk[u] = kap[kk];
                vt[u] = tr[p.cs * (base + kk)];
            }
        }
#pragma unroll
        for (int u = 0; u < kVmixChunk; ++u) {
            const int k = q + u;
            if (live && k < nz && k >= k0) {
                const double dl = (k < nz - 1) ? dt / p.dzw[k] * vk[u] : 0.0;
                const double dz = p.dzt[k];
                const bool edge = k == k0;
                const double ak = edge ? 0.0 : -dlm / dz;
                const double bk = edge ? 1 + dl / dz : 1 + (dl + dlm) / dz;
                const double ck = (k < nz - 1) ? -dl / dz : 0.0;
                double dk = vt[u];
                if (k == nz - 1) dk = dk + dt * forc / dz;  // surface forcing (:277, :284)
                const double r = 1.0 / (bk - ak * cpp);
                cpp = ck * r;
                dpp = (dk - ak * dpp) * r;
                dlm = dl;
            }
            cp[k] = cpp;
            dp[k] = dpp;
        }
    }
    // back substitution from the top, a chunk at a time; the tracer and its
    // tendency (:279-288), the old values re-read in one batch per chunk
    double xn = 0.0;
#pragma unroll
    for (int q = (NZB - 1) / kVmixChunk * kVmixChunk; q >= 0; q -= kVmixChunk) {
        if (q < nz) {
            double vo[kVmixChunk];
            const bool live = q + kVmixChunk > k0;
            if (live) {
#pragma unroll
                for (int u = 0; u < kVmixChunk; ++u) vo[u] = tr[p.cs * (base + (q + u < nz ? q + u : nz - 1))];
            }
            double dvv[kVmixChunk];
#pragma unroll
            for (int u = kVmixChunk - 1; u >= 0; --u) {
                const int k = q + u;
                dvv[u] = 0.0;
                if (k < nz && live && k >= k0) {
                    const double x = dp[k] - cp[k] * xn;
                    xn = x;
                    tr[p.cs * (base + k)] = x;
                    dvv[u] = (x - vo[u]) / dt;
                }
            }
            // the chunk's tendencies: 16-byte stores when the column start is
            // 16-byte aligned (nz even)
            if (vec) {
#pragma unroll
                for (int u = 0; u < kVmixChunk; u += 2) {
                    if (q + u < nz) *reinterpret_cast<double2*>(dv + q + u) = make_double2(dvv[u], dvv[u + 1]);
                }
            } else {
#pragma unroll
                for (int u = 0; u < kVmixChunk; ++u) {
                    if (q + u < nz) dv[q + u] = dvv[u];
                }
            }
        }
    }
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

template <int NZB>
void launch_vmix(const VmixArgs& p, hipStream_t s) {
    const int64_t lanes = 2 * p.X * p.Y;
    hipLaunchKernelGGL(vertmix_kernel<NZB>, dim3((unsigned)((lanes + kVmixBlock - 1) / kVmixBlock)),
                       dim3(kVmixBlock), 0, s, p);
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
    p.vec_dv = !(p.nz & 1) && aligned16(dtemp_vmix) && aligned16(dsalt_vmix);
    if (grid->time_layout != VX_LEVELS_INNER && grid->time_layout != VX_LEVELS_OUTER)
        return fail(VX_ERR_INVALID_ARGUMENT, "%s: time_layout must be VX_LEVELS_INNER or VX_LEVELS_OUTER", who);
    const bool outer = grid->time_layout == VX_LEVELS_OUTER;
    p.cs = outer ? 1 : 3;
    p.ls = outer ? p.X * p.Y * p.nz : 1;
    hipStream_t s = (hipStream_t)stream;
    const int nzb = ((p.nz + 15) / 16) * 16;
    switch (nzb) {
        case 16: launch_vmix<16>(p, s); break;
        case 32: launch_vmix<32>(p, s); break;
        case 48: launch_vmix<48>(p, s); break;
        case 64: launch_vmix<64>(p, s); break;
        case 80: launch_vmix<80>(p, s); break;
        case 96: launch_vmix<96>(p, s); break;
        case 112: launch_vmix<112>(p, s); break;
        default: launch_vmix<128>(p, s); break;
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
