// Isoneutral triad-tensor stencils for gfx950 (float64).
//
//   iso_pre_dry_kernel, iso_pre_kernel
//                        isoneutral_diffusion_pre  veros/core/isoneutral/isoneutral.py:18-229
//   iso_flux_kernel      _calc_tracer_fluxes       veros/core/isoneutral/diffusion.py:9-113
//   wet_scan_*_kernel    the wet-cell plan of a step (from kbot)
//   iso_step_wet_kernel  _calc_explicit_part, _calc_implicit_part, the tracer and
//                        dtr_iso updates and the energy branch, over the wet cells
//                                                  veros/core/isoneutral/diffusion.py:116-283,
//                                                  veros/core/diffusion.py:9-62
//   iso_adj_*_kernel     the adjoint of one isoneutral tracer step
//
// Layout: veros state arrays with 2-cell halos, z contiguous, k = 0 deepest.
// Cell-parallel kernels put consecutive lanes on consecutive levels of one
// column, so every stencil read of a neighbouring column is itself a
// contiguous run. Masks are derived from kbot exactly as calc_topo_kernel
// builds them (veros/core/numerics.py:195-221): maskT = kbot>0 & k>=kbot-1,
// maskU/V = min of the two adjacent maskT, maskW = min(maskT[k], maskT[k+1]).
#include <map>
#include <mutex>

#include "tdma_seg.hpp"

// Tuning knob (compile-time): minimum waves per SIMD requested for the pre
// kernel (caps its VGPRs).
#ifndef VX_PRE_MIN_WAVES
#define VX_PRE_MIN_WAVES 3
#endif

namespace vx {

constexpr double kEpsln = 1e-20;  // veros/core/isoneutral/isoneutral.py:28

struct Grid {
    const int32_t* kbot;
    const double *dxt, *dxu, *dyt, *dyu, *cost, *cosu, *dzt, *dzw;
    int32_t X, Y, nz;
    // time-level arrays (temp, salt, tracer, int_drhodX): level l of cell c at
    // l * ls + c * cs; cs = 3, ls = 1 for the reference's (X, Y, nz, 3)
    // layout, cs = 1, ls = X*Y*nz for levels stored one after another
    uint32_t cs, ls;
};

struct Prm {
    double drhodT, drhodS, K_iso_steep, iso_slopec, iso_dslope, dt_tracer, grav, rho_0;
};

__device__ __forceinline__ int64_t cidx(const Grid& g, int i, int j) { return (int64_t)i * g.Y + j; }
__device__ __forceinline__ int64_t idx3(const Grid& g, int i, int j, int k) { return cidx(g, i, j) * g.nz + k; }
__device__ __forceinline__ int kb_of(const Grid& g, int i, int j) { return g.kbot[cidx(g, i, j)]; }
__device__ __forceinline__ double mT(int kb, int k) { return (kb > 0 && kb - 1 <= k) ? 1.0 : 0.0; }
// Level k of a column with bottom index kb is a wet T cell (maskT = 1).
__device__ __forceinline__ bool wet(int kb, int k) { return kb > 0 && kb - 1 <= k; }

// Tracer value at time level `lev` of a time-level array.
__device__ __forceinline__ double tr3(const double* __restrict__ f, const Grid& g, int i, int j, int k, int lev) {
    return f[(uint32_t)idx3(g, i, j, k) * g.cs + (uint32_t)lev * g.ls];
}

// Masked face gradients (veros/core/isoneutral/isoneutral.py:46-95).
struct Fields {
    const double* __restrict__ T;
    const double* __restrict__ S;
    int tau;
};

// Blocks are dispatched round-robin over the 8 XCDs; remap so that each XCD
// walks a contiguous range of the domain and neighbouring rows share its L2.
// `nb` must be a multiple of 8.
__device__ __forceinline__ int64_t xcd_block(int64_t nb) {
    const int64_t b = blockIdx.x;
    return (b & 7) * (nb >> 3) + (b >> 3);
}

// 1/x from v_rcp_f64 refined by two Newton steps (about 1 ulp; finite
// non-zero x only).
__device__ __forceinline__ double rcp_nr(double x) {
    double r = __builtin_amdgcn_rcp(x);
    double e = fma(-x, r, 1.0);
    r = fma(r, e, r);
    e = fma(-x, r, 1.0);
    return fma(r, e, r);
}

// Masked face gradients (veros/core/isoneutral/isoneutral.py:46-95); `rd` is
// the reciprocal of the metric factor (dxu*cost, dyu or dzw).
struct Grad {
    double T, S;
};

__device__ __forceinline__ Grad grad_x(const Grid& g, const Fields& f, int i, int j, int k, double rd) {
    const double m = mT(kb_of(g, i, j), k) * mT(kb_of(g, i + 1, j), k);
    return {m * (tr3(f.T, g, i + 1, j, k, f.tau) - tr3(f.T, g, i, j, k, f.tau)) * rd,
            m * (tr3(f.S, g, i + 1, j, k, f.tau) - tr3(f.S, g, i, j, k, f.tau)) * rd};
}

__device__ __forceinline__ Grad grad_y(const Grid& g, const Fields& f, int i, int j, int k, double rd) {
    const double m = mT(kb_of(g, i, j), k) * mT(kb_of(g, i, j + 1), k);
    return {m * (tr3(f.T, g, i, j + 1, k, f.tau) - tr3(f.T, g, i, j, k, f.tau)) * rd,
            m * (tr3(f.S, g, i, j + 1, k, f.tau) - tr3(f.S, g, i, j, k, f.tau)) * rd};
}

// dT/dz at the top face of (i, j, k); 0 at the top level and below k = 0.
__device__ __forceinline__ Grad grad_z(const Grid& g, const Fields& f, int i, int j, int k, double rd) {
    if (k < 0 || k >= g.nz - 1) return {0.0, 0.0};
    const int kb = kb_of(g, i, j);
    const double m = mT(kb, k) * mT(kb, k + 1);
    return {m * (tr3(f.T, g, i, j, k + 1, f.tau) - tr3(f.T, g, i, j, k, f.tau)) * rd,
            m * (tr3(f.S, g, i, j, k + 1, f.tau) - tr3(f.S, g, i, j, k, f.tau)) * rd};
}

// exp(x) for |x| <= 64 by Cody-Waite reduction: x = n ln2 + r with ln2 split
// into a 32-bit-exact high part and a low part, |r| <= ln2/2, e^r from its
// Taylor series to degree 13 (truncation < 5e-18 relative), then 2^n.
// The coefficients are read once per kernel into an ExpCoef (uniform, so
// they stay in scalar registers and feed v_fma_f64 directly).
__constant__ double kExpTable[16] = {
    1.4426950408889634,                       // log2(e)
    6.93147180369123816490e-01,               // ln2, high part
    1.90821492927058770002e-10,               // ln2, low part
    1.0 / 6227020800.0, 1.0 / 479001600.0, 1.0 / 39916800.0, 1.0 / 3628800.0, 1.0 / 362880.0,
    1.0 / 40320.0, 1.0 / 5040.0, 1.0 / 720.0, 1.0 / 120.0, 1.0 / 24.0, 1.0 / 6.0, 0.5, 1.0};

struct ExpCoef {
    double c[16];
};

__device__ __forceinline__ ExpCoef load_exp_coef() {
    ExpCoef e;
#pragma unroll
    for (int q = 0; q < 16; ++q) e.c[q] = kExpTable[q];
    return e;
}

__device__ __forceinline__ double exp_reduced(const ExpCoef& e, double x) {
    const double n = rint(x * e.c[0]);
    double r = fma(-n, e.c[1], x);
    r = fma(-n, e.c[2], r);
    double q = e.c[3];
#pragma unroll
    for (int t = 4; t < 16; ++t) q = fma(q, r, e.c[t]);
    q = fma(q, r, 1.0);
    return ldexp(q, (int)n);
}

// dm_taper (veros/core/isoneutral/isoneutral.py:10-15):
//   0.5*(1 + tanh(u)) == 1/(1 + exp(-2u)),  u = (slopec - |s|)/dslope.
// In float64 tanh(u) rounds to exactly -1 / +1 for |u| > 55*ln(2)/2 (~19.06),
// which makes the reference taper exactly 0 / 1 there; those branches are
// taken without evaluating the exponential. `two_over_dslope` = 2/dslope.
__device__ __forceinline__ double taper(const Prm& p, const ExpCoef& ec, double two_over_dslope, double s) {
    const double x = (fabs(s) - p.iso_slopec) * two_over_dslope;  // -2u
    // branch-free: the exponential of the clamped argument, then the exact
    // end values by selection (a NaN slope stays NaN)
    const double t = rcp_nr(1.0 + exp_reduced(ec, fmin(fmax(x, -38.12), 38.12)));
    return x > 38.12 ? 0.0 : (x < -38.12 ? 1.0 : (x == x ? t : x));
}

// Horizontal and vertical density gradients of one triad, weighted by the
// drho/dX of one T point (the 0/1 mask is folded into the weights).
struct Dens {
    double drdT, drdS;
};
__device__ __forceinline__ Dens dens(const Prm& p, double mt) { return {mt * p.drhodT, mt * p.drhodS}; }
__device__ __forceinline__ double dot(const Dens& w, const Grad& gr) { return w.drdT * gr.T + w.drdS * gr.S; }
// 1 / (min(0, drho/dz) - epsln): the slope denominator
__device__ __forceinline__ double rden(double drodz) { return rcp_nr(fmin(0.0, drodz) - kEpsln); }

// ---------------------------------------------------------------------------
// Cell-parallel launch geometry: block = (nz levels) x (cpb columns along j);
// consecutive lanes are consecutive levels of one column, so every read of a
// neighbouring column is a contiguous run. Blocks are ordered in strips of
// kStripBlocks blocks along j, walked row by row along i, and the 1-D grid is
// remapped XCD-contiguously: the i-1 / i / i+1 rows of a strip are then in
// flight on the same XCD together and their shared reads hit its L2.
// ---------------------------------------------------------------------------
#ifndef VX_STRIP_BLOCKS
#define VX_STRIP_BLOCKS 16
#endif
constexpr int kStripBlocks = VX_STRIP_BLOCKS;
// the wet-cell pre kernel's blocks are wider (pre_columns): fewer per strip
// (4 x 24 columns at nz = 80; measured on 1440x720x80, 16 per strip made the
// fused pre + temperature/salinity call 1% slower, 8 5% slower)
#ifndef VX_PRE_STRIP_BLOCKS
#define VX_PRE_STRIP_BLOCKS 4
#endif
constexpr int kPreStripBlocks = VX_PRE_STRIP_BLOCKS;

struct CellIdx {
    int i, j, k;
    bool valid;
};

// The block's row i and first column jb when blocks own cpb columns each.
struct BlockCols {
    int i, j;
    bool valid;  // false for a padding block
};
__device__ __forceinline__ BlockCols block_cols(int i0, int ni, int j0, int nj, int cpb, int sb = kStripBlocks) {
    const int sw = sb * cpb;                          // strip width in columns
    const int nstrips = (nj + sw - 1) / sw;
    const int per_strip = ni * sb;
    const int L = (int)xcd_block(gridDim.x);
    const int sidx = L / per_strip;
    const int rem = L - sidx * per_strip;
    const int ii = rem / sb;
    const int b = rem - ii * sb;
    BlockCols r;
    r.i = i0 + ii;
    r.j = j0 + sidx * sw + b * cpb;
    r.valid = sidx < nstrips && r.j < j0 + nj;
    return r;
}
__device__ __forceinline__ CellIdx cell_of(int i0, int ni, int j0, int nj) {
    const BlockCols bc = block_cols(i0, ni, j0, nj, (int)blockDim.y);
    CellIdx r;
    r.i = bc.i;
    r.j = bc.j + (int)threadIdx.y;
    r.k = (int)threadIdx.x;
    r.valid = bc.valid && r.j < j0 + nj;
    return r;
}

__device__ __forceinline__ double m01(int kb, int k) { return (kb > 0 && kb - 1 <= k) ? 1.0 : 0.0; }

// ---------------------------------------------------------------------------
// isoneutral_diffusion_pre: thread per cell (i, j, k), i in [1, X-2), j in [1, Y-2).
// The 13 tracer points of the stencil are loaded once per field; all face
// gradients are formed from them in registers. Offsets are 32-bit element
// offsets (the host checks the arrays fit).
// ---------------------------------------------------------------------------
// Face-flux arithmetic (veros/core/isoneutral/diffusion.py:22-111), shared by
// the flux kernel and the fused pre + flux kernel so both produce the same
// values. A01/A23 are the (ip=0, kr=0..1) / (ip=1, kr=0..1) entries of Ai_ez or
// Ai_nz; t0* the tracer at the cell (u = k+1, d = k-1, clamped at the ends),
// t1* at the east or north neighbour.
__device__ __forceinline__ double east_flux_core(double diffloc, double2 A01, double2 A23, double t0, double t0u,
                                                 double t0d, double t1, double t1u, double t1d, double dzt,
                                                 double cost, double dxu, double K11) {
    double sumz = 0.0;
    sumz = sumz + diffloc * A01.x * (t0 - t0d);  // kr=0, ip=0
    sumz = sumz + diffloc * A23.x * (t1 - t1d);  // kr=0, ip=1
    sumz = sumz + diffloc * A01.y * (t0u - t0);  // kr=1, ip=0
    sumz = sumz + diffloc * A23.y * (t1u - t1);  // kr=1, ip=1
    return sumz / (4.0 * dzt) + (t1 - t0) / (cost * dxu) * K11;
}
__device__ __forceinline__ double north_flux_core(double diffloc, double2 A01, double2 A23, double t0, double t0u,
                                                  double t0d, double t1, double t1u, double t1d, double dzt,
                                                  double cosu, double dyu, double K22) {
    double sumz = 0.0;
    sumz = sumz + diffloc * A01.x * (t0 - t0d);
    sumz = sumz + diffloc * A23.x * (t1 - t1d);
    sumz = sumz + diffloc * A01.y * (t0u - t0);
    sumz = sumz + diffloc * A23.y * (t1u - t1);
    return cosu * (sumz / (4.0 * dzt) + (t1 - t0) / dyu * K22);
}
// tc/tw/te/ts/tn = the tracer at the cell and its west/east/south/north
// neighbours, levels k (0) and k+1 (1); cosu_s = cosu[j-1]
__device__ __forceinline__ double top_flux_core(double diffloc, double2 B01, double2 B23, double2 C01, double2 C23,
                                                double tc0, double tc1, double tw0, double tw1, double te0,
                                                double te1, double ts0, double ts1, double tn0, double tn1,
                                                double cost, double cosu, double cosu_s, double dxt, double dyt) {
    double sumx = 0.0;
    sumx = sumx + diffloc * B01.x / cost * (tc0 - tw0);  // ip=0, kr=0
    sumx = sumx + diffloc * B01.y / cost * (tc1 - tw1);  // ip=0, kr=1
    sumx = sumx + diffloc * B23.x / cost * (te0 - tc0);  // ip=1, kr=0
    sumx = sumx + diffloc * B23.y / cost * (te1 - tc1);  // ip=1, kr=1
    double sumy = 0.0;
    sumy = sumy + diffloc * C01.x * cosu_s * (tc0 - ts0);
    sumy = sumy + diffloc * C01.y * cosu_s * (tc1 - ts1);
    sumy = sumy + diffloc * C23.x * cosu * (tn0 - tc0);
    sumy = sumy + diffloc * C23.y * cosu * (tn1 - tc1);
    return sumx / (4 * dxt) + sumy / (4 * dyt * cost);
}

// Face-flux outputs of the fused pre + flux kernel (temperature, salinity).
struct PreFlux {
    double *fe[2], *fn[2], *ft[2];
};

// The dry cells of isoneutral_diffusion_pre. A dry cell has every output
// exactly zero: its own maskT factor is in maskU (Ai_ez, K_11), maskV (Ai_nz,
// K_22) and maskW (Ai_bx, Ai_by, K_33), which multiply each entry and each
// summed term. The dry cells of a column are its levels 0 .. kbot-2, one
// contiguous run per array, so a wave per column zero-fills those runs with
// lane-contiguous stores (full cache lines) over the index ranges the
// reference writes; iso_pre_kernel handles the wet cells.
// The zeros are stored nontemporal: nothing in the call reads them back, and
// streaming them past the caches leaves the L2 to the wet-cell kernels beside
// them (measured on 1440x720x80: pre 4.9 -> 4.6 ms, pre + the
// temperature/salinity step in one call 8.3 -> 7.5 ms; the same stores for the
// wet cells' tensor made pre slower).
typedef double vx_d2 __attribute__((ext_vector_type(2)));
__device__ __forceinline__ void zero_run(double* __restrict__ p, int n, int lane) {
    for (int m = lane; m < n; m += 64) __builtin_nontemporal_store(0.0, p + m);
}
// Ai entries 4 c + 2 ip + kr of the cells 0 .. nd-1 of a column; at level 0
// only the kr = 1 entries are written (the kr = 0 ones keep their values).
__device__ __forceinline__ void zero_ai_run(double* __restrict__ p, int nd, int lane, bool keep_level0_kr0) {
    const int n2 = 2 * nd;  // double2 pieces
    for (int m = lane; m < n2; m += 64) {
        if (keep_level0_kr0 && m < 2) __builtin_nontemporal_store(0.0, p + 2 * m + 1);
        else __builtin_nontemporal_store((vx_d2){0.0, 0.0}, reinterpret_cast<vx_d2*>(p + 2 * m));
    }
}

__global__ __launch_bounds__(256) void iso_pre_dry_kernel(Grid g, double* __restrict__ Aez, double* __restrict__ Anz,
                                                          double* __restrict__ Abx, double* __restrict__ Aby,
                                                          double* __restrict__ K11, double* __restrict__ K22,
                                                          double* __restrict__ K33) {
    const int lane = (int)(threadIdx.x & 63);
    const int64_t ni = g.X - 3, nj = g.Y - 3;
    const int64_t stride = (int64_t)gridDim.x * (blockDim.x >> 6);
    // grid-stride over the columns (the grid may be smaller than one wave per
    // column: a background fill beside other kernels)
    for (int64_t col = (int64_t)blockIdx.x * (blockDim.x >> 6) + (threadIdx.x >> 6); col < ni * nj; col += stride) {
        const int i = 1 + (int)(col / nj), j = 1 + (int)(col - (col / nj) * nj);
        const int nz = g.nz;
        const int kb = g.kbot[i * g.Y + j];
        const int nd = (kb >= 1 && kb <= nz) ? kb - 1 : nz;  // dry levels 0 .. nd-1
        if (nd == 0) continue;
        const uint32_t c0 = ((uint32_t)i * (uint32_t)g.Y + (uint32_t)j) * (uint32_t)nz;
        if (j >= 2) {
            zero_ai_run(Aez + 4u * c0, nd, lane, true);
            zero_run(K11 + c0, nd, lane);
        }
        if (i >= 2) {
            zero_ai_run(Anz + 4u * c0, nd, lane, true);
            zero_run(K22 + c0, nd, lane);
        }
        if (i >= 2 && j >= 2) {
            const int nt = nd < nz - 1 ? nd : nz - 1;  // top faces: levels below nz-1
            zero_ai_run(Abx + 4u * c0, nt, lane, false);
            zero_ai_run(Aby + 4u * c0, nt, lane, false);
            zero_run(K33 + c0, nt, lane);
        }
    }
}

// One wet cell (i, j, k) of isoneutral_diffusion_pre. FUSED: also the face
// fluxes of temperature and salinity at level tau from the tensor just formed
// (the flux kernel's work for the step that follows pre), into `fo`.
template <bool FUSED>
__device__ __forceinline__ void pre_wet_cell(const Grid& g, const Prm& p, const Fields& f,
                                             const double* __restrict__ K, double* __restrict__ Aez,
                                             double* __restrict__ Anz, double* __restrict__ Abx,
                                             double* __restrict__ Aby, double* __restrict__ K11,
                                             double* __restrict__ K22, double* __restrict__ K33, const PreFlux& fo,
                                             const int i, const int j, const int k) {
    const int nz = g.nz;
    const uint32_t sj = (uint32_t)nz, si = (uint32_t)g.Y * (uint32_t)nz;
    const uint32_t c = (uint32_t)i * si + (uint32_t)j * sj + (uint32_t)k;
    const uint32_t ce = c + si, cw = c - si, cn = c + sj, cs = c - sj;
    const bool hp = k < nz - 1, hm = k >= 1;
    const double* __restrict__ T = f.T + (uint32_t)f.tau * g.ls;
    const double* __restrict__ S = f.S + (uint32_t)f.tau * g.ls;
    const int kb_c = g.kbot[i * g.Y + j], kb_e = g.kbot[(i + 1) * g.Y + j], kb_w = g.kbot[(i - 1) * g.Y + j];
    const int kb_n = g.kbot[i * g.Y + j + 1], kb_s = g.kbot[i * g.Y + j - 1];

    // tracer values: c(enter), e(ast), w(est), n(orth), s(outh); p = k+1, m = k-1.
    // Loaded whether or not they are used (k+1 at the top and k-1 at the
    // bottom are cells of the next and previous columns, in bounds for every
    // cell pre visits, i in [1, X-3], j in [1, Y-3]): a load under a per-lane
    // condition is waited for at once.
    auto L2 = [&](uint32_t o, bool ok, double& t, double& sv) {
        const double tv = T[g.cs * o], svv = S[g.cs * o];
        t = ok ? tv : 0.0;
        sv = ok ? svv : 0.0;
    };
    double Tc, Sc, Tcp, Scp, Tcm, Scm, Te, Se, Tep, Sep, Tem, Sem, Tw, Sw, Twp, Swp;
    double Tn, Sn, Tnp, Snp, Tnm, Snm, Ts, Ss, Tsp, Ssp;
    L2(c, true, Tc, Sc);
    L2(c + 1, hp, Tcp, Scp);
    L2(c - 1, hm, Tcm, Scm);
    L2(ce, true, Te, Se);
    L2(ce + 1, hp, Tep, Sep);
    L2(ce - 1, hm, Tem, Sem);
    L2(cw, true, Tw, Sw);
    L2(cw + 1, hp, Twp, Swp);
    L2(cn, true, Tn, Sn);
    L2(cn + 1, hp, Tnp, Snp);
    L2(cn - 1, hm, Tnm, Snm);
    L2(cs, true, Ts, Ss);
    L2(cs + 1, hp, Tsp, Ssp);

    // Every other global read up front: reads placed inside the divergent
    // face branches below would each cost a separate memory round trip.
    const double dzw_k = g.dzw[k], dzw_l = g.dzw[hm ? k - 1 : k], dzt_k = g.dzt[k];
    const double dzw_m = hm ? dzw_l : 0.0;
    const double dxu_i = g.dxu[i], dxu_w = g.dxu[i - 1], dxt_i = g.dxt[i];
    const double cost_j = g.cost[j], dyu_j = g.dyu[j], dyu_s = g.dyu[j - 1], dyt_j = g.dyt[j];
    const double cosu_j = g.cosu[j], cosu_s = g.cosu[j - 1];
    const double Kc = K[c], Kcl = K[c - 1], Ke = K[ce], Kel = K[ce - 1], Kn = K[cn], Knl = K[cn - 1];
    const double Kcm = hm ? Kcl : 0.0, Kem = hm ? Kel : 0.0, Knm = hm ? Knl : 0.0;

    // reciprocal metric factors
    const double rz_c = hp ? rcp_nr(dzw_k) : 0.0;
    const double rz_m = hm ? rcp_nr(dzw_m) : 0.0;
    const double rx_c = rcp_nr(dxu_i * cost_j);
    const double ry_c = rcp_nr(dyu_j);

    // masked face gradients (veros/core/isoneutral/isoneutral.py:46-95)
    auto G = [](double m, double a1, double a0, double b1, double b0, double rd) -> Grad {
        return {m * (a1 - a0) * rd, m * (b1 - b0) * rd};
    };
    const double mc = m01(kb_c, k), mcp = m01(kb_c, k + 1), mcm = m01(kb_c, k - 1);
    const double me = m01(kb_e, k), mep = m01(kb_e, k + 1), mem = m01(kb_e, k - 1);
    const double mn = m01(kb_n, k), mnp = m01(kb_n, k + 1), mnm = m01(kb_n, k - 1);
    const Grad gz_c = hp ? G(mc * mcp, Tcp, Tc, Scp, Sc, rz_c) : Grad{0.0, 0.0};
    const Grad gz_m = hm ? G(mcm * mc, Tc, Tcm, Sc, Scm, rz_m) : Grad{0.0, 0.0};

    const double tds = 2.0 / p.iso_dslope;
    const ExpCoef ec = load_exp_coef();

    // Two triads belong to two faces of this cell: east (ip=0, kr=1) is the
    // top-face x triad (ip=1, kr=0) and north (jp=0, kr=1) the top-face y triad
    // (jp=1, kr=0); same density point, same gradients, so the same slope
    // and taper, formed once here.
    const Dens w_c = dens(p, mc);
    const double rd_c = rden(dot(w_c, gz_c));
    double sl_x = 0.0, tp_x = 0.0, sl_y = 0.0, tp_y = 0.0;
    if (j >= 2) {
        sl_x = -dot(w_c, G(mc * me, Te, Tc, Se, Sc, rx_c)) * rd_c;
        tp_x = taper(p, ec, tds, sl_x);
    }
    if (i >= 2) {
        sl_y = -dot(w_c, G(mc * mn, Tn, Tc, Sn, Sc, ry_c)) * rd_c;
        tp_y = taper(p, ec, tds, sl_y);
    }

    // --- east face (i+1/2, j, k), j in [2, Y-2)   (:97-132)
    if (j >= 2) {
        const double mU = mc * me;
        const Grad gx = G(mU, Te, Tc, Se, Sc, rx_c);
        const Grad gz_e = hp ? G(me * mep, Tep, Te, Sep, Se, rz_c) : Grad{0.0, 0.0};
        const Grad gz_em = hm ? G(mem * me, Te, Tem, Se, Sem, rz_m) : Grad{0.0, 0.0};
        const double diffloc = hm ? 0.25 * (Kc + Kcm + Ke + Kem) : 0.5 * (Kc + Ke);
        double sumz = 0.0;
        double out[4];
#pragma unroll
        for (int kr = 0; kr < 2; ++kr) {
#pragma unroll
            for (int ip = 0; ip < 2; ++ip) {
                out[ip * 2 + kr] = 0.0;
                if (kr == 0 && !hm) continue;  // levels k >= 1 - kr
                double sl = sl_x, tp = tp_x;
                if (ip != 0 || kr != 1) {
                    const Dens w = dens(p, ip ? me : mc);
                    const Grad& gz = ip ? (kr ? gz_e : gz_em) : (kr ? gz_c : gz_m);
                    sl = -dot(w, gx) * rden(dot(w, gz));
                    tp = taper(p, ec, tds, sl);
                }
                sumz += (kr ? dzw_k : dzw_m) * mU * fmax(p.K_iso_steep, diffloc * tp);
                out[ip * 2 + kr] = tp * sl * mU;
            }
        }
        if (hm) {
            *reinterpret_cast<double2*>(Aez + 4u * c) = make_double2(out[0], out[1]);
        } else {
            Aez[4u * c + 1] = out[1];  // (ip=0, kr=0) keeps its value at k = 0
        }
        if (hm) {
            *reinterpret_cast<double2*>(Aez + 4u * c + 2) = make_double2(out[2], out[3]);
        } else {
            Aez[4u * c + 3] = out[3];
        }
        const double k11v = sumz / (4.0 * dzt_k);
        K11[c] = k11v;
        if constexpr (FUSED) {
            // at level 0 the kr = 0 entries keep their values (their tracer
            // differences vanish there)
            const double a0 = hm ? out[0] : Aez[4u * c], a2 = hm ? out[2] : Aez[4u * c + 2];
            const double2 A01 = make_double2(a0, out[1]), A23 = make_double2(a2, out[3]);
            fo.fe[0][c] = east_flux_core(diffloc, A01, A23, Tc, hp ? Tcp : Tc, hm ? Tcm : Tc, Te, hp ? Tep : Te,
                                         hm ? Tem : Te, dzt_k, cost_j, dxu_i, k11v);
            fo.fe[1][c] = east_flux_core(diffloc, A01, A23, Sc, hp ? Scp : Sc, hm ? Scm : Sc, Se, hp ? Sep : Se,
                                         hm ? Sem : Se, dzt_k, cost_j, dxu_i, k11v);
        }
    }

    // --- north face (i, j+1/2, k), i in [2, X-2)   (:134-168)
    if (i >= 2) {
        const double mV = mc * mn;
        const Grad gy = G(mV, Tn, Tc, Sn, Sc, ry_c);
        const Grad gz_n = hp ? G(mn * mnp, Tnp, Tn, Snp, Sn, rz_c) : Grad{0.0, 0.0};
        const Grad gz_nm = hm ? G(mnm * mn, Tn, Tnm, Sn, Snm, rz_m) : Grad{0.0, 0.0};
        const double diffloc = hm ? 0.25 * (Kc + Kcm + Kn + Knm) : 0.5 * (Kc + Kn);
        double sumz = 0.0;
        double out[4];
#pragma unroll
        for (int kr = 0; kr < 2; ++kr) {
#pragma unroll
            for (int jp = 0; jp < 2; ++jp) {
                out[jp * 2 + kr] = 0.0;
                if (kr == 0 && !hm) continue;
                double sl = sl_y, tp = tp_y;
                if (jp != 0 || kr != 1) {
                    const Dens w = dens(p, jp ? mn : mc);
                    const Grad& gz = jp ? (kr ? gz_n : gz_nm) : (kr ? gz_c : gz_m);
                    sl = -dot(w, gy) * rden(dot(w, gz));
                    tp = taper(p, ec, tds, sl);
                }
                sumz += (kr ? dzw_k : dzw_m) * mV * fmax(p.K_iso_steep, diffloc * tp);
                out[jp * 2 + kr] = tp * sl * mV;
            }
        }
        if (hm) {
            *reinterpret_cast<double2*>(Anz + 4u * c) = make_double2(out[0], out[1]);
            *reinterpret_cast<double2*>(Anz + 4u * c + 2) = make_double2(out[2], out[3]);
        } else {
            Anz[4u * c + 1] = out[1];
            Anz[4u * c + 3] = out[3];
        }
        const double k22v = sumz / (4.0 * dzt_k);
        K22[c] = k22v;
        if constexpr (FUSED) {
            const double a0 = hm ? out[0] : Anz[4u * c], a2 = hm ? out[2] : Anz[4u * c + 2];
            const double2 A01 = make_double2(a0, out[1]), A23 = make_double2(a2, out[3]);
            fo.fn[0][c] = north_flux_core(diffloc, A01, A23, Tc, hp ? Tcp : Tc, hm ? Tcm : Tc, Tn, hp ? Tnp : Tn,
                                          hm ? Tnm : Tn, dzt_k, cosu_j, dyu_j, k22v);
            fo.fn[1][c] = north_flux_core(diffloc, A01, A23, Sc, hp ? Scp : Sc, hm ? Scm : Sc, Sn, hp ? Snp : Sn,
                                          hm ? Snm : Sn, dzt_k, cosu_j, dyu_j, k22v);
        }
    }

    // --- top face (i, j, k+1/2), i, j in [2, X/Y-2), k < nz-1   (:170-225)
    if (i >= 2 && j >= 2 && hp) {
        const int kb_w2 = kb_w, kb_s2 = kb_s;
        const double mW = mc * mcp;
        const double rx_w = rcp_nr(dxu_w * cost_j);
        const double ry_s = rcp_nr(dyu_s);
        const double mw = m01(kb_w2, k), mwp = m01(kb_w2, k + 1);
        const double ms = m01(kb_s2, k), msp = m01(kb_s2, k + 1);
        // horizontal gradients at (i-1+ip, k+kr) and (j-1+jp, k+kr)
        const Grad gx_w = G(mw * mc, Tc, Tw, Sc, Sw, rx_w);
        const Grad gx_wp = G(mwp * mcp, Tcp, Twp, Scp, Swp, rx_w);
        const Grad gx_c = G(mc * me, Te, Tc, Se, Sc, rx_c);
        const Grad gx_cp = G(mcp * mep, Tep, Tcp, Sep, Scp, rx_c);
        const Grad gy_s = G(ms * mc, Tc, Ts, Sc, Ss, ry_s);
        const Grad gy_sp = G(msp * mcp, Tcp, Tsp, Scp, Ssp, ry_s);
        const Grad gy_c = G(mc * mn, Tn, Tc, Sn, Sc, ry_c);
        const Grad gy_cp = G(mcp * mnp, Tnp, Tcp, Snp, Scp, ry_c);
        double sumx = 0.0, sumy = 0.0;
        double ox[4], oy[4];
#pragma unroll
        for (int kr = 0; kr < 2; ++kr) {
            const Dens w = dens(p, kr ? mcp : mc);
            const double rd = kr ? rden(dot(w, gz_c)) : rd_c;
#pragma unroll
            for (int ip = 0; ip < 2; ++ip) {
                double sl = sl_x, tp = tp_x;
                if (ip != 1 || kr != 0) {
                    const Grad& gx = ip ? (kr ? gx_cp : gx_c) : (kr ? gx_wp : gx_w);
                    sl = -dot(w, gx) * rd;
                    tp = taper(p, ec, tds, sl);
                }
                sumx += (ip ? dxu_i : dxu_w) * Kc * tp * (sl * sl) * mW;
                ox[ip * 2 + kr] = tp * sl * mW;
            }
#pragma unroll
            for (int jp = 0; jp < 2; ++jp) {
                double sl = sl_y, tp = tp_y;
                if (jp != 1 || kr != 0) {
                    const Grad& gy = jp ? (kr ? gy_cp : gy_c) : (kr ? gy_sp : gy_s);
                    sl = -dot(w, gy) * rd;
                    tp = taper(p, ec, tds, sl);
                }
                const double facty = jp ? cosu_j * dyu_j : cosu_s * dyu_s;
                sumy += facty * Kc * tp * (sl * sl) * mW;
                oy[jp * 2 + kr] = tp * sl * mW;
            }
        }
        *reinterpret_cast<double2*>(Abx + 4u * c) = make_double2(ox[0], ox[1]);
        *reinterpret_cast<double2*>(Abx + 4u * c + 2) = make_double2(ox[2], ox[3]);
        *reinterpret_cast<double2*>(Aby + 4u * c) = make_double2(oy[0], oy[1]);
        *reinterpret_cast<double2*>(Aby + 4u * c + 2) = make_double2(oy[2], oy[3]);
        K33[c] = sumx / (4 * dxt_i) + sumy / (4 * dyt_j * cost_j);
        if constexpr (FUSED) {
            const double2 B01 = make_double2(ox[0], ox[1]), B23 = make_double2(ox[2], ox[3]);
            const double2 C01 = make_double2(oy[0], oy[1]), C23 = make_double2(oy[2], oy[3]);
            fo.ft[0][c] = top_flux_core(Kc, B01, B23, C01, C23, Tc, Tcp, Tw, Twp, Te, Tep, Ts, Tsp, Tn, Tnp, cost_j,
                                        cosu_j, cosu_s, dxt_i, dyt_j);
            fo.ft[1][c] = top_flux_core(Kc, B01, B23, C01, C23, Sc, Scp, Sw, Swp, Se, Sep, Ss, Ssp, Sn, Snp, cost_j,
                                        cosu_j, cosu_s, dxt_i, dyt_j);
        }
    }
    if constexpr (FUSED) {
        if (i >= 2 && j >= 2 && !hp) {  // no flux through the top level
            fo.ft[0][c] = 0.0;
            fo.ft[1][c] = 0.0;
        }
    }
}

// The wet cells of isoneutral_diffusion_pre (the dry ones are written by
// iso_pre_dry_kernel). Block-local compaction: a block owns the columns
// (i, jb .. jb+ncb-1); their wet cells (the wet suffixes) are numbered in
// column order and dealt round-robin to the block's threads, so the slope work
// fills whole waves whatever the land fraction.
template <bool FUSED>
__global__ __launch_bounds__(256) __attribute__((amdgpu_waves_per_eu(FUSED ? 2 : VX_PRE_MIN_WAVES))) void iso_pre_kernel(
    Grid g, Prm p, Fields f, const double* __restrict__ K, double* __restrict__ Aez, double* __restrict__ Anz,
    double* __restrict__ Abx, double* __restrict__ Aby, double* __restrict__ K11, double* __restrict__ K22,
    double* __restrict__ K33, PreFlux fo, int ncb) {
    extern __shared__ int pre_counts[];  // prefix sums of the columns' wet counts
    const BlockCols bc = block_cols(1, g.X - 3, 1, g.Y - 3, ncb, kPreStripBlocks);
    if (!bc.valid) return;  // padding block (uniform)
    const int nz = g.nz, i = bc.i, jb = bc.j;
    const int tid = (int)threadIdx.x, nthr = (int)blockDim.x;
    int* sW = pre_counts;
    for (int q = tid; q < ncb; q += nthr) {
        const bool col_ok = jb + q <= g.Y - 3;
        const int kbq = col_ok ? g.kbot[i * g.Y + jb + q] : 0;
        sW[q + 1] = (kbq >= 1 && kbq <= nz) ? nz - kbq + 1 : 0;
    }
    __syncthreads();
    if (tid == 0) {
        sW[0] = 0;
        for (int q = 0; q < ncb; ++q) sW[q + 1] += sW[q];
    }
    __syncthreads();
    const int nwet = sW[ncb];
    for (int t = tid; t < nwet; t += nthr) {
        // the column q with sW[q] <= t < sW[q+1]: the last q with sW[q] <= t
        int lo = 0, hi = ncb - 1;
        while (lo < hi) {
            const int mid = (lo + hi + 1) >> 1;
            if (sW[mid] <= t) lo = mid;
            else hi = mid - 1;
        }
        const int k = nz - (sW[lo + 1] - sW[lo]) + (t - sW[lo]);  // wet suffix of column lo
        pre_wet_cell<FUSED>(g, p, f, K, Aez, Anz, Abx, Aby, K11, K22, K33, fo, i, jb + lo, k);
    }
}

// K_33[..., nz-1] = 0 over the whole plane (veros/core/isoneutral/isoneutral.py:225)
__global__ void k33_top_zero_kernel(Grid g, double* __restrict__ K33) {
    const int64_t t = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (t >= (int64_t)g.X * g.Y) return;
    K33[t * g.nz + g.nz - 1] = 0.0;
}

// ---------------------------------------------------------------------------
// Tracer step (veros/core/isoneutral/diffusion.py:9-283 with
// veros/core/diffusion.py:9-62), NTR tracers sharing one pass over the
// isoneutral tensor, covering (i, j) in [1, X-1) x [1, Y-1):
//   iso_flux_kernel (thread per cell): the face fluxes -> workspace (in the
//     fused pre + temperature/salinity call, iso_pre_kernel<true> instead)
//   iso_step_wet_kernel phase 1 (thread per wet cell): lateral dissipation on
//     T points, the flux divergence, rhs = tr[taup1] + dt*dtr, and the K33
//     rows -> LDS
//   phase 2 (lane per interior column and tracer): Thomas sweep in LDS
//   phase 3 (thread per wet cell): tracer, dtr_iso and P_diss updates
// ---------------------------------------------------------------------------
struct TracerArgs {
    double* tr;           // (X, Y, nz, 3)
    double* dtr;          // (X, Y, nz)
    const double* I;      // int_drhodX (X, Y, nz, 3), nullable
    double *fe, *fn, *ft; // workspace: flux_east, flux_north, flux_top (X, Y, nz)
};

struct StepArgs {
    TracerArgs t[2];
    int tau, taup1;
    // K1 = K_iso - K_skew (east/north faces), K2 = K_iso + K_skew (top faces);
    // a null pointer is a zero term (veros/core/isoneutral/diffusion.py:15-16,180-188)
    const double* Kiso;
    const double* Kskew;
    const double *Aez, *Anz, *Abx, *Aby, *K11, *K22, *K33;
    double* P;            // P_diss_iso / P_diss_skew, nullable (energy branch off)
};

struct Strides {
    uint32_t si, sj;
};

// Inputs of the face fluxes of one cell (i, j, k), read from global memory
// before any flux is formed. `K` is the one diffusivity of the step (K_iso for
// the isoneutral step, K_gm for the skew step) at c, c-1, the east column
// (e, e-1) and the north column (n, n-1).
struct FaceIn {
    double K_c, K_cm, K_e, K_em, K_n, K_nm;
    double2 ez01, ez23, nz01, nz23, bx01, bx23, by01, by23;
    double K11, K22;
    double dzt, cost, dxu, cosu, cosu_s, dyu, dxt, dyt;
};

// K1 = K_iso - K_skew (east/north faces) and K2 = K_iso + K_skew (top faces)
// (veros/core/isoneutral/diffusion.py:15-16,180-188); one of the two is zero.
template <bool ISO>
__device__ __forceinline__ double k1_of(double K) { return ISO ? K : 0.0 - K; }
template <bool ISO>
__device__ __forceinline__ double k2_of(double K) { return ISO ? K : 0.0 + K; }

// Face fluxes of NTR tracers (veros/core/isoneutral/diffusion.py:9-113):
// flux_east for (i, j) in [1, X-3] x [2, Y-3], flux_north for [2, X-3] x
// [1, Y-3], flux_top for [2, X-3] x [2, Y-3] (0 at the top level); the tracers
// share every read of the isoneutral tensor. Every use of a face flux is
// multiplied by the maskT of a cell next to the face (the divergence, the
// lateral dissipation and the vertical mixing term of that cell), so only the
// faces touching a wet cell are formed and written; the step kernel reads no
// other. The cells that need one are compacted like the wet-cell pre kernel's:
// a column (i, j) needs levels from the lowest of its own wet
// cells, the top face of the dry cell below them, and the wet cells of its east
// and north neighbours, up to nz-1 (one contiguous run); a block owns ncb
// columns of row i and deals their runs round-robin to its threads. The
// tracer stencil is read straight from global memory (every read issued
// before the first flux), so no lane of a dry region is launched.
__device__ __forceinline__ int first_wet(int kb, int nz) { return (kb >= 1 && kb <= nz) ? kb - 1 : nz; }
__device__ __forceinline__ int flux_run_start(const Grid& g, int i, int j) {
    const int nz = g.nz;
    const int fc = first_wet(g.kbot[i * g.Y + j], nz);
    int s0 = fc < nz ? (fc >= 1 ? fc - 1 : 0) : nz;             // top face of the dry cell below
    const int fe = j >= 2 ? first_wet(g.kbot[(i + 1) * g.Y + j], nz) : nz;
    const int fn = i >= 2 ? first_wet(g.kbot[i * g.Y + j + 1], nz) : nz;
    s0 = fe < s0 ? fe : s0;
    return fn < s0 ? fn : s0;
}

template <int NTR, bool ISO>
__global__ __launch_bounds__(256) void iso_flux_kernel(Grid g, StepArgs a, int ncb) {
    extern __shared__ int flux_counts[];  // prefix sums of the columns' run lengths
    const BlockCols bc = block_cols(1, g.X - 3, 1, g.Y - 3, ncb, kPreStripBlocks);
    if (!bc.valid) return;  // padding block (uniform)
    const int nz = g.nz, i = bc.i, jb = bc.j;
    const int tid = (int)threadIdx.x, nthr = (int)blockDim.x;
    int* sW = flux_counts;
    for (int q = tid; q < ncb; q += nthr) sW[q + 1] = jb + q <= g.Y - 3 ? nz - flux_run_start(g, i, jb + q) : 0;
    __syncthreads();
    if (tid == 0) {
        sW[0] = 0;
        for (int q = 0; q < ncb; ++q) sW[q + 1] += sW[q];
    }
    __syncthreads();
    const int ncell = sW[ncb];
    const uint32_t si = (uint32_t)g.Y * (uint32_t)nz, sj = (uint32_t)nz;
    const double* __restrict__ Kf = ISO ? a.Kiso : a.Kskew;
    for (int t = tid; t < ncell; t += nthr) {
        int q = 0, hi = ncb - 1;
        while (q < hi) {
            const int mid = (q + hi + 1) >> 1;
            if (sW[mid] <= t) q = mid;
            else hi = mid - 1;
        }
        const int j = jb + q;
        const int k = nz - (sW[q + 1] - sW[q]) + (t - sW[q]);
        const int kb = g.kbot[i * g.Y + j], kb_e = g.kbot[(i + 1) * g.Y + j], kb_n = g.kbot[i * g.Y + j + 1];
        const bool wc = wet(kb, k);
        const bool need_e = j >= 2 && (wc || wet(kb_e, k));
        const bool need_n = i >= 2 && (wc || wet(kb_n, k));
        const bool need_t = i >= 2 && j >= 2 && (wc || (k + 1 < nz && wet(kb, k + 1)));
        if (!(need_e || need_n || need_t)) continue;
        const uint32_t c = (uint32_t)i * si + (uint32_t)j * sj + (uint32_t)k;
        const bool hm = k >= 1, hp = k + 1 < nz;
        const bool top = need_t && hp;
        const uint32_t up = hp ? 1u : 0u, dn = hm ? 1u : 0u;  // pad_z_edges clamping
        // every global read before the first flux
        FaceIn in{};
        in.K_c = Kf[c];
        in.K_cm = hm ? Kf[c - 1] : 0.0;
        in.dzt = g.dzt[k];
        in.cost = g.cost[j];
        in.dxu = g.dxu[i];
        in.cosu = g.cosu[j];
        in.cosu_s = g.cosu[j - 1];
        in.dyu = g.dyu[j];
        in.dxt = g.dxt[i];
        in.dyt = g.dyt[j];
        if (need_e) {
            in.K_e = Kf[c + si];
            in.K_em = hm ? Kf[c + si - 1] : 0.0;
            in.ez01 = *reinterpret_cast<const double2*>(a.Aez + 4u * c);
            in.ez23 = *reinterpret_cast<const double2*>(a.Aez + 4u * c + 2);
            in.K11 = a.K11[c];
        }
        if (need_n) {
            in.K_n = Kf[c + sj];
            in.K_nm = hm ? Kf[c + sj - 1] : 0.0;
            in.nz01 = *reinterpret_cast<const double2*>(a.Anz + 4u * c);
            in.nz23 = *reinterpret_cast<const double2*>(a.Anz + 4u * c + 2);
            in.K22 = a.K22[c];
        }
        if (top) {
            in.bx01 = *reinterpret_cast<const double2*>(a.Abx + 4u * c);
            in.bx23 = *reinterpret_cast<const double2*>(a.Abx + 4u * c + 2);
            in.by01 = *reinterpret_cast<const double2*>(a.Aby + 4u * c);
            in.by23 = *reinterpret_cast<const double2*>(a.Aby + 4u * c + 2);
        }
        // tracer at level tau: centre and east/north columns at k, k+up, k-dn;
        // west/south at k, k+1 (top faces only)
        double Tc[NTR], Tcu[NTR], Tcd[NTR], Te[NTR], Teu[NTR], Ted[NTR], Tn[NTR], Tnu[NTR], Tnd[NTR];
        double Tw[NTR], Twu[NTR], Ts[NTR], Tsu[NTR];
#pragma unroll
        for (int r = 0; r < NTR; ++r) {
            const double* __restrict__ T = a.t[r].tr + (uint32_t)a.tau * g.ls;
            Tc[r] = T[g.cs * c];
            Tcu[r] = T[g.cs * (c + up)];
            Tcd[r] = T[g.cs * (c - dn)];
            Te[r] = Teu[r] = Ted[r] = Tn[r] = Tnu[r] = Tnd[r] = Tw[r] = Twu[r] = Ts[r] = Tsu[r] = 0.0;
            if (need_e || top) {
                Te[r] = T[g.cs * (c + si)];
                Teu[r] = T[g.cs * (c + si + up)];
            }
            if (need_e) Ted[r] = T[g.cs * (c + si - dn)];
            if (need_n || top) {
                Tn[r] = T[g.cs * (c + sj)];
                Tnu[r] = T[g.cs * (c + sj + up)];
            }
            if (need_n) Tnd[r] = T[g.cs * (c + sj - dn)];
            if (top) {
                Tw[r] = T[g.cs * (c - si)];
                Twu[r] = T[g.cs * (c - si + 1)];
                Ts[r] = T[g.cs * (c - sj)];
                Tsu[r] = T[g.cs * (c - sj + 1)];
            }
        }
        // flux_east (:22-47), flux_north (:49-77), flux_top (:79-111)
        const double dl_e = hm ? 0.25 * (k1_of<ISO>(in.K_c) + k1_of<ISO>(in.K_cm) + k1_of<ISO>(in.K_e) +
                                         k1_of<ISO>(in.K_em))
                               : 0.5 * (k1_of<ISO>(in.K_c) + k1_of<ISO>(in.K_e));
        const double dl_n = hm ? 0.25 * (k1_of<ISO>(in.K_c) + k1_of<ISO>(in.K_cm) + k1_of<ISO>(in.K_n) +
                                         k1_of<ISO>(in.K_nm))
                               : 0.5 * (k1_of<ISO>(in.K_c) + k1_of<ISO>(in.K_n));
#pragma unroll
        for (int r = 0; r < NTR; ++r) {
            if (need_e)
                a.t[r].fe[c] = east_flux_core(dl_e, in.ez01, in.ez23, Tc[r], Tcu[r], Tcd[r], Te[r], Teu[r], Ted[r],
                                              in.dzt, in.cost, in.dxu, in.K11);
            if (need_n)
                a.t[r].fn[c] = north_flux_core(dl_n, in.nz01, in.nz23, Tc[r], Tcu[r], Tcd[r], Tn[r], Tnu[r], Tnd[r],
                                               in.dzt, in.cosu, in.dyu, in.K22);
            if (need_t)
                a.t[r].ft[c] = top ? top_flux_core(k2_of<ISO>(in.K_c), in.bx01, in.bx23, in.by01, in.by23, Tc[r],
                                                   Tcu[r], Tw[r], Twu[r], Te[r], Teu[r], Ts[r], Tsu[r], Tn[r], Tnu[r],
                                                   in.cost, in.cosu, in.cosu_s, in.dxt, in.dyt)
                                   : 0.0;
        }
    }
}

// Workspace fluxes with their valid index ranges (0 elsewhere).
__device__ __forceinline__ double fe_at(const Grid& g, const TracerArgs& t, int i, int j, uint32_t o) {
    return (i >= 1 && i <= g.X - 3 && j >= 2 && j <= g.Y - 3) ? t.fe[o] : 0.0;
}
__device__ __forceinline__ double fn_at(const Grid& g, const TracerArgs& t, int i, int j, uint32_t o) {
    return (i >= 2 && i <= g.X - 3 && j >= 1 && j <= g.Y - 3) ? t.fn[o] : 0.0;
}


// ---------------------------------------------------------------------------
// Wet-cell tracer step. The wet columns of the step's range [1, X-1) x
// [1, Y-1) are ranked in the strip order of wet_col_ij (rank r, a column's wet
// cells being its levels kbot-1 .. nz-1), and their cells are numbered in the
// same order (column r holds cells o_r .. o_r + w_r - 1). The cells are cut
// into groups: group b takes the wet columns with
// b*C <= o_r < (b+1)*C, C = kWetThreads - nz + 1, so a group holds at most
// C + nz - 1 = kWetThreads cells (w_r <= nz). A block steps one group at a
// time with a thread per cell, so no lane is spent on a dry cell; the plan is
// built on the device from kbot by three small scan kernels.
// ---------------------------------------------------------------------------
constexpr int kWetThreads = 256;
constexpr int kScanColsPerThread = 8;
constexpr int kScanChunk = kWetThreads * kScanColsPerThread;  // columns per scan block

struct WetPlan {
    int* wetcol;  // [ncols]   enumeration index (wet_col_ij) of the wet column of rank r
    int* wetoff;  // [ncols]   o_r, the number of wet cells before it
    int* first;   // [ncols+2] first[b] = first rank of group b; first[ngroups] = number of wet columns
    int* nb;      // [12]      ngroups, wet columns, wet cells, -, then a group counter per XCD
    int2* chunk;  // [nchunks] (wet columns, wet cells) per scan chunk, then their exclusive sums
};

// The columns of the step range are enumerated in strips of kWetStrip
// columns along j, each strip walked row by row along i, so that the rows a
// group's stencil reaches (i-1, i+1) are stepped close in time and share an
// L2; the columns past Y-2 in the last strip are padding without wet cells.
#ifndef VX_WET_STRIP
#define VX_WET_STRIP 64
#endif
constexpr int kWetStrip = VX_WET_STRIP;
__host__ __device__ inline int64_t wet_enum_cols(int64_t X, int64_t Y) {
    return (Y - 2 + kWetStrip - 1) / kWetStrip * (X - 2) * kWetStrip;
}
__device__ __forceinline__ void wet_col_ij(const Grid& g, int col, int& i, int& j) {
    const int per = (g.X - 2) * kWetStrip;
    const int sidx = col / per;
    const int rem = col - sidx * per;
    const int ii = rem / kWetStrip;
    i = 1 + ii;
    j = 1 + sidx * kWetStrip + (rem - ii * kWetStrip);
}
__device__ __forceinline__ int wet_count(const Grid& g, int64_t col, int64_t ncols) {
    if (col >= ncols) return 0;
    int i, j;
    wet_col_ij(g, (int)col, i, j);
    if (j > g.Y - 2) return 0;
    const int kb = g.kbot[i * g.Y + j];
    return (kb >= 1 && kb <= g.nz) ? g.nz - kb + 1 : 0;
}

// In-place inclusive sum over the block (Hillis-Steele; every thread calls it).
__device__ __forceinline__ void block_scan2(int* s0, int* s1, int tid) {
    for (int st = 1; st < kWetThreads; st <<= 1) {
        const int a0 = tid >= st ? s0[tid - st] : 0, a1 = tid >= st ? s1[tid - st] : 0;
        __syncthreads();
        s0[tid] += a0;
        s1[tid] += a1;
        __syncthreads();
    }
}

// per scan chunk of kScanChunk columns: (wet columns, wet cells)
__global__ __launch_bounds__(kWetThreads) void wet_scan_count_kernel(Grid g, WetPlan w, int64_t ncols) {
    __shared__ int s0[kWetThreads], s1[kWetThreads];
    const int tid = (int)threadIdx.x;
    const int64_t base = (int64_t)blockIdx.x * kScanChunk + (int64_t)tid * kScanColsPerThread;
    int nc = 0, ncell = 0;
    for (int q = 0; q < kScanColsPerThread; ++q) {
        const int n = wet_count(g, base + q, ncols);
        nc += n > 0 ? 1 : 0;
        ncell += n;
    }
    s0[tid] = nc;
    s1[tid] = ncell;
    __syncthreads();
    block_scan2(s0, s1, tid);
    if (tid == kWetThreads - 1) w.chunk[blockIdx.x] = make_int2(s0[tid], s1[tid]);
}

// one block: exclusive sums of the chunk counts, and the totals
__global__ __launch_bounds__(kWetThreads) void wet_scan_chunks_kernel(WetPlan w, int nchunks) {
    __shared__ int s0[kWetThreads], s1[kWetThreads];
    __shared__ int carry[2];
    const int tid = (int)threadIdx.x;
    if (tid == 0) carry[0] = carry[1] = 0;
    __syncthreads();
    for (int base = 0; base < nchunks; base += kWetThreads) {
        const int q = base + tid;
        const int2 v = q < nchunks ? w.chunk[q] : make_int2(0, 0);
        s0[tid] = v.x;
        s1[tid] = v.y;
        __syncthreads();
        block_scan2(s0, s1, tid);
        const int c0 = carry[0], c1 = carry[1];
        if (q < nchunks) w.chunk[q] = make_int2(c0 + s0[tid] - v.x, c1 + s1[tid] - v.y);
        __syncthreads();
        if (tid == kWetThreads - 1) {
            carry[0] = c0 + s0[tid];
            carry[1] = c1 + s1[tid];
        }
        __syncthreads();
    }
    if (tid == 0) {
        w.nb[0] = 0;  // set by wet_scan_columns_kernel when there is a wet column
        w.nb[1] = carry[0];
        w.nb[2] = carry[1];
    }
    if (tid < 8) w.nb[4 + tid] = 0;  // the step kernel's group counters
}

// ranks, offsets and the group starts
__global__ __launch_bounds__(kWetThreads) void wet_scan_columns_kernel(Grid g, WetPlan w, int64_t ncols, int C) {
    __shared__ int s0[kWetThreads], s1[kWetThreads];
    const int tid = (int)threadIdx.x;
    const int64_t base = (int64_t)blockIdx.x * kScanChunk + (int64_t)tid * kScanColsPerThread;
    int n[kScanColsPerThread];
    int nc = 0, ncell = 0;
#pragma unroll
    for (int q = 0; q < kScanColsPerThread; ++q) {
        n[q] = wet_count(g, base + q, ncols);
        nc += n[q] > 0 ? 1 : 0;
        ncell += n[q];
    }
    s0[tid] = nc;
    s1[tid] = ncell;
    __syncthreads();
    block_scan2(s0, s1, tid);
    const int2 co = w.chunk[blockIdx.x];
    int r = co.x + s0[tid] - nc, o = co.y + s1[tid] - ncell;
    const int nwc = w.nb[1];
#pragma unroll
    for (int q = 0; q < kScanColsPerThread; ++q) {
        if (n[q] > 0) {
            w.wetcol[r] = (int)(base + q);
            w.wetoff[r] = o;
            if (r == 0) w.first[0] = 0;
            // the group boundary m = b*C inside (o, o + n]: the columns before
            // group b are ranks 0 .. r
            const int m = (o / C + 1) * C;
            if (m <= o + n[q]) w.first[m / C] = r + 1;
            if (r == nwc - 1) {
                const int ng = o / C + 1;
                w.first[ng] = nwc;
                w.nb[0] = ng;
            }
            ++r;
            o += n[q];
        }
    }
}

// Per-tracer LDS arrays of the tracer step: F (flux_top), D (lateral
// dissipation), R (rhs -> alpha -> new), Q (c' -> gamma), Wk (spike -> beta).
constexpr int kStepArrays = 5;

// One column's K33 system (rows a, b, c of n levels in LDS, right-hand side
// in r) solved in place into r by the kSegP lanes of a group, lane sg owning
// the rows [sg * S, sg * S + S) with S = max(2, ceil(n / kSegP)); rows from n
// on are identity rows. Each lane eliminates its interior rows with the spike
// from the row below (q: c', then gamma; w: the spike, then beta; r: dp, then
// alpha), the interface rows are solved across the group (iface_solve), and
// each lane writes its rows of x = alpha + beta L_{s-1} + gamma L_s. Every
// lane of the group must call it.
__device__ __forceinline__ void seg_lds_column(const double* a, const double* b, const double* c, double* r,
                                               double* q, double* w, int n, int sg) {
    const int S = max(2, (n + kSegP - 1) / kSegP);
    const int m0 = sg * S;
    auto A = [&](int m) { return m < n ? a[m] : 0.0; };
    auto B = [&](int m) { return m < n ? b[m] : 1.0; };
    auto C = [&](int m) { return m < n ? c[m] : 0.0; };
    auto D = [&](int m) { return m < n ? r[m] : 0.0; };
    // forward sweep over the interior rows t = 0 .. S-2 (c', spike, dp to LDS)
    double cp = 0.0, wp = 0.0, dp = 0.0;
    for (int t = 0; t < S - 1; ++t) {
        const int m = m0 + t;
        const double at = A(m);
        const double rr = seg_rcp(t == 0 ? B(m) : B(m) - at * cp);
        cp = C(m) * rr;
        wp = (t == 0 ? at : -at * wp) * rr;
        dp = (t == 0 ? D(m) : D(m) - at * dp) * rr;
        if (m < n) {
            q[m] = cp;
            w[m] = wp;
            r[m] = dp;
        }
    }
    // back substitution: alpha, beta, gamma of the interior rows (rows past
    // the run have c' = spike = dp = 0)
    const double alS = dp, beS = -wp, gaS = -cp;  // row S-2
    if (m0 + S - 2 < n) {
        r[m0 + S - 2] = alS;
        w[m0 + S - 2] = beS;
        q[m0 + S - 2] = gaS;
    }
    double al = alS, be = beS, ga = gaS;
    for (int t = S - 3; t >= 0; --t) {
        const int m = m0 + t;
        const bool real = m < n;
        const double cpt = real ? q[m] : 0.0, dpt = real ? r[m] : 0.0, wpt = real ? w[m] : 0.0;
        al = dpt - cpt * al;
        be = -wpt - cpt * be;
        ga = -cpt * ga;
        if (real) {
            r[m] = al;
            w[m] = be;
            q[m] = ga;
        }
    }
    // the interface row (row S-1) with the next segment's row 0 by DPP
    const int mL = m0 + S - 1;
    const double aL = A(mL), bL = B(mL), cL = C(mL), dL = D(mL);
    const double n_al = grp_next(sg, al), n_be = grp_next(sg, be), n_ga = grp_next(sg, ga);
    const double RA = aL * beS;
    const double RB = bL + aL * gaS + cL * n_be;
    const double RC = cL * n_ga;
    const double RD[1] = {dL - aL * alS - cL * n_al};
    double Ls[1], Lm[1];
    iface_solve<1>(sg, RA, RB, RC, RD, Ls, Lm);
    for (int t = 0; t < S - 1; ++t) {
        const int m = m0 + t;
        if (m < n) r[m] = r[m] + w[m] * Lm[0] + q[m] * Ls[0];
    }
    if (mL < n) r[mL] = Ls[0];
}

// Persistent grid (a multiple of 8 blocks): the blocks of XCD x (blockIdx % 8)
// share the groups [x, x+1) * ngroups / 8, so neighbouring rows share an L2,
// and take them one at a time from that XCD's counter.
// IMPLICIT: the K33 part (isoneutral step); else the explicit step only (skew).
// WETFLUX: the face fluxes exist for wet cells only (the fused pre formed them
// from the tensor it wrote, in which every face touching a dry cell carries a
// zero flux), so a face whose other cell is dry reads as zero; else they come
// from iso_flux_kernel, which writes every face of a wet cell, including the
// top face of the dry cell below the lowest wet one.
#ifndef VX_WET_WAVES
#define VX_WET_WAVES 5  // 96 VGPRs: five blocks per CU
#endif
template <int NTR, bool IMPLICIT, bool WETFLUX>
__global__ __launch_bounds__(kWetThreads) __attribute__((amdgpu_waves_per_eu(VX_WET_WAVES))) void iso_step_wet_kernel(
    Grid g, Prm p, StepArgs a, WetPlan w) {
    extern __shared__ double lds[];
    constexpr int T = kWetThreads;
    // LDS: rows A, B, C (shared by the tracers), then per tracer F (flux_top),
    // D (lateral dissipation), R (rhs -> alpha -> new), Q (c' -> gamma), Wk
    // (spike -> beta), each T cells; then the group's column table
    double* sA = lds;
    double* sB = lds + T;
    double* sC = lds + 2 * T;
    double* sT = lds + 3 * T;
    int* sOff = reinterpret_cast<int*>(lds + (3 + kStepArrays * NTR) * T);  // [T+1] first cell of each column
    int* sCol = sOff + (T + 1);                                   // [T]   its column index
    const int nz = g.nz;
    const int tid = (int)threadIdx.x;
    const int ngroups = w.nb[0];
    const int xcd = (int)(blockIdx.x & 7);
    const int gbeg = (int)((int64_t)ngroups * xcd / 8), gend = (int)((int64_t)ngroups * (xcd + 1) / 8);
    const Strides st{(uint32_t)g.Y * (uint32_t)nz, (uint32_t)nz};
    const bool energy = a.P != nullptr;
    const double dt = p.dt_tracer;
    __shared__ int sGrp;
    for (;;) {
        __syncthreads();  // the previous group's LDS (and sGrp) is no longer read
        if (tid == 0) sGrp = gbeg + atomicAdd(&w.nb[4 + xcd], 1);
        __syncthreads();
        const int grp = sGrp;
        if (grp >= gend) break;  // uniform: every block ends once its XCD's share is taken
        const int r0 = w.first[grp], r1 = w.first[grp + 1];
        const int ncolg = r1 - r0;
        const int o0 = w.wetoff[r0];
        if (tid < ncolg) {
            sOff[tid] = w.wetoff[r0 + tid] - o0;
            sCol[tid] = w.wetcol[r0 + tid];
        }
        if (tid == 0) sOff[ncolg] = (grp + 1 < ngroups ? w.wetoff[r1] : w.nb[2]) - o0;
        __syncthreads();
        const int ncell = sOff[ncolg];
        const bool have = tid < ncell;
        // this thread's cell: the column q with sOff[q] <= tid < sOff[q+1]
        int q = 0;
        if (have) {
            int hi = ncolg - 1;
            while (q < hi) {
                const int mid = (q + hi + 1) >> 1;
                if (sOff[mid] <= tid) q = mid;
                else hi = mid - 1;
            }
        }
        const int lo = tid;
        int i, j;
        wet_col_ij(g, have ? sCol[q] : 0, i, j);
        const int kb = have ? nz - (sOff[q + 1] - sOff[q]) + 1 : 1;
        const int k = kb - 1 + (tid - sOff[q]);
        const bool interior = have && i >= 2 && i <= g.X - 3 && j >= 2 && j <= g.Y - 3;
        const uint32_t c = (uint32_t)i * st.si + (uint32_t)j * st.sj + (uint32_t)k;
        bool west_ok = have, south_ok = have;
        if (WETFLUX && have) {
            west_ok = wet(g.kbot[(i - 1) * g.Y + j], k);
            south_ok = wet(g.kbot[i * g.Y + j - 1], k);
        }

        // ---- every global read first
        double fe_c[NTR], fe_w[NTR], fn_c[NTR], fn_s[NTR], ft_c[NTR], ft_b[NTR];
        double Ic[NTR], Ie[NTR], Iw[NTR], In[NTR], Is[NTR], Iu[NTR], trp[NTR], dtr0[NTR];
#pragma unroll
        for (int t = 0; t < NTR; ++t) {
            const TracerArgs& ta = a.t[t];
            fe_c[t] = fe_w[t] = fn_c[t] = fn_s[t] = ft_c[t] = ft_b[t] = 0.0;
            Ic[t] = Ie[t] = Iw[t] = In[t] = Is[t] = Iu[t] = trp[t] = dtr0[t] = 0.0;
            if (have) {
                fe_c[t] = fe_at(g, ta, i, j, c);
                fe_w[t] = west_ok ? fe_at(g, ta, i - 1, j, c - st.si) : 0.0;
                fn_c[t] = fn_at(g, ta, i, j, c);
                fn_s[t] = south_ok ? fn_at(g, ta, i, j - 1, c - st.sj) : 0.0;
                if (interior) {
                    trp[t] = ta.tr[g.cs * c + (uint32_t)a.taup1 * g.ls];
                    dtr0[t] = ta.dtr[c];
                    ft_c[t] = ta.ft[c];
                    // the top face of the dry cell below the lowest wet cell
                    if (!WETFLUX && k == kb - 1 && k >= 1) ft_b[t] = ta.ft[c - 1];
                }
                if (energy) {
                    const double* __restrict__ I = ta.I + (uint32_t)a.tau * g.ls;
                    Ic[t] = I[g.cs * c];
                    Ie[t] = I[g.cs * (c + st.si)];
                    Iw[t] = I[g.cs * (c - st.si)];
                    In[t] = I[g.cs * (c + st.sj)];
                    Is[t] = I[g.cs * (c - st.sj)];
                    if (interior && k < nz - 1) Iu[t] = I[g.cs * (c + 1)];
                }
            }
        }
        double K33k = 0.0, K33m = 0.0;
        if (IMPLICIT && interior) {
            K33k = a.K33[c];
            if (k >= kb) K33m = a.K33[c - 1];  // the edge row k = kbot-1 has no sub-diagonal
        }
        const double P0 = (energy && have) ? a.P[c] : 0.0;
        double dxt_i = 0.0, cost_j = 1.0, dyt_j = 0.0, dzt_k = 1.0, dzw_k = 1.0, dzw_lo = 1.0;
        if (have) {
            dxt_i = g.dxt[i];
            cost_j = g.cost[j];
            dyt_j = g.dyt[j];
            dzt_k = g.dzt[k];
            dzw_k = g.dzw[k];
            dzw_lo = g.dzw[k >= 1 ? k - 1 : 0];  // pad_z_edges(dzw)[k]
        }

        // ---- phase 1a: flux_top and lateral dissipation -> LDS, K33 rows -> LDS
        if (have) {
#pragma unroll
            for (int t = 0; t < NTR; ++t) {
                double* sTt = sT + t * kStepArrays * T;
                sTt[lo] = ft_c[t];
                if (energy) {
                    // compute_dissipation (veros/core/diffusion.py:13-33)
                    const double de = (Ie[t] - Ic[t]) * fe_c[t] + (Ic[t] - Iw[t]) * fe_w[t];
                    const double dn = (In[t] - Ic[t]) * fn_c[t] + (Ic[t] - Is[t]) * fn_s[t];
                    const double g05 = 0.5 * p.grav / p.rho_0;
                    sTt[T + lo] = g05 * de / (dxt_i * cost_j) + g05 * dn / (dyt_j * cost_j);
                }
            }
        }
        if (IMPLICIT && interior) {
            // rows of the K33 system (veros/core/isoneutral/diffusion.py:154-164); the
            // edge row k = kbot-1 takes b_edge and a = 0 (veros/core/operators.py:72)
            const double delta = (k < nz - 1) ? dt / dzw_k * K33k : 0.0;
            const double delta_m = (k >= 1) ? dt / dzw_lo * K33m : 0.0;
            const double dz = dzt_k;
            const bool edge = k == kb - 1;
            sA[lo] = edge ? 0.0 : -delta_m / dz;
            sB[lo] = edge ? 1 + delta / dz : 1 + (delta + delta_m) / dz;
            sC[lo] = (k < nz - 1) ? -delta / dz : 0.0;
        }
        __syncthreads();

        // ---- phase 1b: divergence and rhs (veros/core/isoneutral/diffusion.py:116-139,196-197)
        double dtrk[NTR], rhs[NTR];
#pragma unroll
        for (int t = 0; t < NTR; ++t) {
            dtrk[t] = rhs[t] = 0.0;
            if (interior) {
                double* sTt = sT + t * kStepArrays * T;
                // below the lowest wet cell: with the fused pre's fluxes the
                // top face there is zero (its maskW), else it was read above
                const double ftm = (k >= kb) ? sTt[lo - 1] : ft_b[t];
                const double hor = (fe_c[t] - fe_w[t]) / (cost_j * dxt_i) +
                                   (fn_c[t] - fn_s[t]) / (cost_j * dyt_j);
                const double vert = (k == 0) ? ft_c[t] / dzt_k : (ft_c[t] - ftm) / dzt_k;
                dtrk[t] = hor + vert;
                rhs[t] = trp[t] + dt * dtrk[t];
                sTt[2 * T + lo] = rhs[t];
            }
        }
        __syncthreads();

        // ---- phase 2: a group of kSegP lanes per (interior column, tracer)
        // solves the column's run in LDS by the segmented elimination
        // (tdma_seg.hpp), segments of S = max(2, ceil(n / kSegP)) rows, rows
        // past the run identity rows
        if (IMPLICIT) {
            const int sg = tid & (kSegP - 1);
            for (int task = tid / kSegP; task < ncolg * NTR; task += T / kSegP) {
                const int cq = task % ncolg, t = task / ncolg;
                int ci, cj;
                wet_col_ij(g, sCol[cq], ci, cj);
                if (ci < 2 || ci > g.X - 3 || cj < 2 || cj > g.Y - 3) continue;  // uniform in the group
                const int b0 = sOff[cq], n = sOff[cq + 1] - b0;
                double* sTt = sT + t * kStepArrays * T;
                seg_lds_column(sA + b0, sB + b0, sC + b0, sTt + 2 * T + b0, sTt + 3 * T + b0, sTt + 4 * T + b0, n,
                               sg);
            }
        }
        __syncthreads();

        // ---- phase 3
        if (have) {
            double Pk = P0;
#pragma unroll
            for (int t = 0; t < NTR; ++t) {
                const double* sTt = sT + t * kStepArrays * T;
                if (interior) {
                    const double newk = sTt[2 * T + lo];
                    a.t[t].tr[g.cs * c + (uint32_t)a.taup1 * g.ls] = newk;
                    a.t[t].dtr[c] = (dtr0[t] + dtrk[t]) + (newk - rhs[t]) / dt;
                }
                if (energy) {
                    // dissipation_on_wgrid (veros/core/diffusion.py:39-62)
                    const double d0 = sTt[T + lo];
                    double dw;
                    if (k == nz - 1) {
                        dw = d0;
                    } else {
                        const double d1 = sTt[T + lo + 1];
                        const double avg = 0.5 * (d0 + d1);
                        const bool edge = k == kb - 1;
                        dw = (avg + 0.5 * (d0 * dzw_lo / dzw_k)) * (edge ? 1.0 : 0.0) + avg * (edge ? 0.0 : 1.0);
                    }
                    Pk = Pk + dw;
                    // vertical mixing term (veros/core/isoneutral/diffusion.py:254-279)
                    if (interior && k < nz - 1) {
                        const double fxa = (-Iu[t] + Ic[t]) / dzw_k;
                        const double gr = -p.grav / p.rho_0;
                        if constexpr (IMPLICIT) {
                            const double dtrz = sTt[2 * T + lo + 1] - sTt[2 * T + lo];
                            Pk = Pk + gr * fxa * (ft_c[t] * 1.0 + K33k * dtrz / dzw_k * 1.0);
                        } else {
                            Pk = Pk + gr * fxa * ft_c[t] * 1.0;
                        }
                    }
                }
            }
            if (energy) a.P[c] = Pk;
        }
    }
}

// ---------------------------------------------------------------------------
// Adjoint of one isoneutral tracer step (iso=True, energy branch off) with the
// isoneutral tensor (Ai_*, K_11, K_22, K_33) held fixed:
//   T_new = M^-1 (T[taup1] + dt * D(T[tau], K_iso))   on interior wet cells,
//   T_new = T[taup1] + dt * D                          on interior dry cells,
//   T_new = T[taup1]                                   elsewhere,
// with D the flux divergence (linear in T[tau], and in K_iso through K1 = K2 =
// K_iso). Given mu = dL/dT_new:
//   adj_column : lam = M^-T mu (transposed K33 Thomas, rows re-derived from
//                K_33), g_taup1 = lam, nu = dt * lam on interior cells
//   adj_faces  : adjoints of flux_east / flux_north / flux_top from nu
//   adj_scatter: each face scatters (atomics) into dL/dT[tau] and dL/dK_iso
// ---------------------------------------------------------------------------
struct AdjArgs {
    const double* tr;       // forward input tracer (X, Y, nz, 3), level tau read
    int tau;
    const double* Kiso;
    const double *Aez, *Anz, *Abx, *Aby, *K11, *K22, *K33;
    const double* mu;       // dL/dT_new (X, Y, nz)
    double *g_tau, *g_taup1, *g_K;
    double *nu, *mfe, *mfn, *mft;  // workspace (X, Y, nz) each
};

// g_taup1 = mu, nu = 0 on interior cells below, g_tau = g_K = 0 everywhere
__global__ __launch_bounds__(256) void iso_adj_init_kernel(int64_t n, AdjArgs a) {
    const int64_t t = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (t >= n) return;
    a.g_taup1[t] = a.mu[t];
    a.g_tau[t] = 0.0;
    a.g_K[t] = 0.0;
    a.nu[t] = 0.0;
    a.mfe[t] = 0.0;
    a.mfn[t] = 0.0;
    a.mft[t] = 0.0;
}

template <int NZB>
__global__ __launch_bounds__(256) void iso_adj_column_kernel(Grid g, Prm p, AdjArgs a) {
    const int nz = g.nz;
    const int ny4 = g.Y - 4;
    const int64_t col = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (col >= (int64_t)(g.X - 4) * ny4) return;
    const int i = 2 + (int)(col / ny4);
    const int j = 2 + (int)(col - (int64_t)(i - 2) * ny4);
    const int kb = kb_of(g, i, j);
    const int k0 = (kb >= 1 && kb <= nz) ? kb - 1 : nz;
    const int64_t o = idx3(g, i, j, 0);
    const double dt = p.dt_tracer;
    // transposed rows: sub'[k] = c[k-1] = -delta[k-1]/dzt[k-1], sup'[k] = a[k+1] = -delta[k]/dzt[k+1]
    double cp[NZB], lm[NZB];
    double cpp = 0.0, dpp = 0.0, delta_m = 0.0;
#pragma unroll
    for (int k = 0; k < NZB; ++k) {
        if (k < nz && k >= k0) {
            const double delta = (k < nz - 1) ? dt / g.dzw[k] * a.K33[o + k] : 0.0;
            const bool edge = k == k0;
            const double dz = g.dzt[k];
            const double diag = edge ? 1 + delta / dz : 1 + (delta + delta_m) / dz;
            const double sub = edge ? 0.0 : -delta_m / g.dzt[k - 1];
            const double sup = (k < nz - 1) ? -delta / g.dzt[k + 1] : 0.0;
            const double r = 1.0 / (diag - sub * cpp);
            cpp = sup * r;
            dpp = (a.mu[o + k] - sub * dpp) * r;
            delta_m = delta;
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
#pragma unroll
    for (int k = 0; k < NZB; ++k) {
        if (k < nz) {
            const double lam = (k >= k0) ? lm[k] : a.mu[o + k];
            if (k >= k0) a.g_taup1[o + k] = lam;
            a.nu[o + k] = dt * lam;
        }
    }
}

// adjoints of the face fluxes: dtr = m*((fe - fe_w)/(cost dxt) + (fn - fn_s)/(cost dyt)) + m*(ft - ft_m)/dzt
__global__ __launch_bounds__(256) void iso_adj_faces_kernel(Grid g, AdjArgs a) {
    const CellIdx ci = cell_of(1, g.X - 3, 1, g.Y - 3);
    if (!ci.valid) return;
    const int i = ci.i, j = ci.j, k = ci.k;
    const int nz = g.nz;
    const uint32_t si = (uint32_t)g.Y * (uint32_t)nz, sj = (uint32_t)nz;
    const uint32_t c = (uint32_t)i * si + (uint32_t)j * sj + (uint32_t)k;
    auto inner = [&](int ii, int jj) { return ii >= 2 && ii <= g.X - 3 && jj >= 2 && jj <= g.Y - 3; };
    auto wnu = [&](int ii, int jj, int kk, uint32_t o) {
        return inner(ii, jj) ? a.nu[o] * m01(kb_of(g, ii, jj), kk) : 0.0;
    };
    if (j >= 2) {  // flux_east at (i, j, k): divergence of cells i and i+1
        a.mfe[c] = wnu(i, j, k, c) / (g.cost[j] * g.dxt[i]) - wnu(i + 1, j, k, c + si) / (g.cost[j] * g.dxt[i + 1]);
    }
    if (i >= 2) {  // flux_north at (i, j, k): cells j and j+1
        a.mfn[c] = wnu(i, j, k, c) / (g.cost[j] * g.dyt[j]) - wnu(i, j + 1, k, c + sj) / (g.cost[j + 1] * g.dyt[j + 1]);
    }
    if (i >= 2 && j >= 2 && k < nz - 1) {  // flux_top at (i, j, k): levels k and k+1
        a.mft[c] = wnu(i, j, k, c) / g.dzt[k] - wnu(i, j, k + 1, c + 1) / g.dzt[k + 1];
    }
}

__device__ __forceinline__ void adj_add(double* p, double v) {
    if (v != 0.0) unsafeAtomicAdd(p, v);
}

__global__ __launch_bounds__(256) void iso_adj_scatter_kernel(Grid g, AdjArgs a) {
    const CellIdx ci = cell_of(1, g.X - 3, 1, g.Y - 3);
    if (!ci.valid) return;
    const int i = ci.i, j = ci.j, k = ci.k;
    const int nz = g.nz;
    const uint32_t si = (uint32_t)g.Y * (uint32_t)nz, sj = (uint32_t)nz;
    const uint32_t c = (uint32_t)i * si + (uint32_t)j * sj + (uint32_t)k;
    const double* __restrict__ tr = a.tr + (uint32_t)a.tau * g.ls;
    const bool hm = k >= 1;
    const uint32_t up = (k + 1 < nz) ? 1u : 0u, dn = hm ? 1u : 0u;
    const double* K = a.Kiso;

    if (j >= 2) {  // flux_east (veros/core/isoneutral/diffusion.py:22-47)
        const double mu = a.mfe[c];
        if (mu != 0.0) {
            const uint32_t ce = c + si;
            const double diffloc = hm ? 0.25 * (K[c] + K[c - 1] + K[ce] + K[ce - 1]) : 0.5 * (K[c] + K[ce]);
            const double* A = a.Aez + 4u * c;  // (ip, kr) at 2 ip + kr
            const double q4 = 1.0 / (4.0 * g.dzt[k]);
            const double t0 = tr[g.cs * c], t0u = tr[g.cs * (c + up)], t0d = tr[g.cs * (c - dn)];
            const double t1 = tr[g.cs * ce], t1u = tr[g.cs * (ce + up)], t1d = tr[g.cs * (ce - dn)];
            // diffloc terms
            const double w00 = mu * diffloc * A[0] * q4, w01 = mu * diffloc * A[1] * q4;
            const double w10 = mu * diffloc * A[2] * q4, w11 = mu * diffloc * A[3] * q4;
            adj_add(a.g_tau + c, w00 - w01);
            adj_add(a.g_tau + c - dn, -w00);
            adj_add(a.g_tau + c + up, w01);
            adj_add(a.g_tau + ce, w10 - w11);
            adj_add(a.g_tau + ce - dn, -w10);
            adj_add(a.g_tau + ce + up, w11);
            // K_11 term
            const double wk = mu * a.K11[c] / (g.cost[j] * g.dxu[i]);
            adj_add(a.g_tau + ce, wk);
            adj_add(a.g_tau + c, -wk);
            // K_iso through diffloc
            const double S = (A[0] * (t0 - t0d) + A[2] * (t1 - t1d) + A[1] * (t0u - t0) + A[3] * (t1u - t1)) * q4;
            const double wK = mu * S * (hm ? 0.25 : 0.5);
            adj_add(a.g_K + c, wK);
            adj_add(a.g_K + ce, wK);
            if (hm) {
                adj_add(a.g_K + c - 1, wK);
                adj_add(a.g_K + ce - 1, wK);
            }
        }
    }
    if (i >= 2) {  // flux_north (:49-77)
        const double mu = a.mfn[c] * g.cosu[j];
        if (mu != 0.0) {
            const uint32_t cn = c + sj;
            const double diffloc = hm ? 0.25 * (K[c] + K[c - 1] + K[cn] + K[cn - 1]) : 0.5 * (K[c] + K[cn]);
            const double* A = a.Anz + 4u * c;
            const double q4 = 1.0 / (4.0 * g.dzt[k]);
            const double t0 = tr[g.cs * c], t0u = tr[g.cs * (c + up)], t0d = tr[g.cs * (c - dn)];
            const double t1 = tr[g.cs * cn], t1u = tr[g.cs * (cn + up)], t1d = tr[g.cs * (cn - dn)];
            const double w00 = mu * diffloc * A[0] * q4, w01 = mu * diffloc * A[1] * q4;
            const double w10 = mu * diffloc * A[2] * q4, w11 = mu * diffloc * A[3] * q4;
            adj_add(a.g_tau + c, w00 - w01);
            adj_add(a.g_tau + c - dn, -w00);
            adj_add(a.g_tau + c + up, w01);
            adj_add(a.g_tau + cn, w10 - w11);
            adj_add(a.g_tau + cn - dn, -w10);
            adj_add(a.g_tau + cn + up, w11);
            const double wk = mu * a.K22[c] / g.dyu[j];
            adj_add(a.g_tau + cn, wk);
            adj_add(a.g_tau + c, -wk);
            const double S = (A[0] * (t0 - t0d) + A[2] * (t1 - t1d) + A[1] * (t0u - t0) + A[3] * (t1u - t1)) * q4;
            const double wK = mu * S * (hm ? 0.25 : 0.5);
            adj_add(a.g_K + c, wK);
            adj_add(a.g_K + cn, wK);
            if (hm) {
                adj_add(a.g_K + c - 1, wK);
                adj_add(a.g_K + cn - 1, wK);
            }
        }
    }
    if (i >= 2 && j >= 2 && k < nz - 1) {  // flux_top (:79-111)
        const double mu = a.mft[c];
        if (mu != 0.0) {
            const double K2 = K[c];
            const double* B = a.Abx + 4u * c;
            const double* C = a.Aby + 4u * c;
            const uint32_t cw = c - si, ce = c + si, cs = c - sj, cn = c + sj;
            const double fx = 1.0 / (4 * g.dxt[i]) / g.cost[j];
            const double fy = 1.0 / (4 * g.dyt[j] * g.cost[j]);
            double S = 0.0;
#pragma unroll
            for (int kr = 0; kr < 2; ++kr) {
                // x: (ip=0) tr[i] - tr[i-1], (ip=1) tr[i+1] - tr[i], at level k + kr
                const double bx0 = B[kr] * fx, bx1 = B[2 + kr] * fx;
                const double cy0 = C[kr] * g.cosu[j - 1] * fy, cy1 = C[2 + kr] * g.cosu[j] * fy;
                const uint32_t ck = c + kr;
                S += bx0 * (tr[g.cs * ck] - tr[g.cs * (cw + kr)]) + bx1 * (tr[g.cs * (ce + kr)] - tr[g.cs * ck]) +
                     cy0 * (tr[g.cs * ck] - tr[g.cs * (cs + kr)]) + cy1 * (tr[g.cs * (cn + kr)] - tr[g.cs * ck]);
                const double m = mu * K2;
                adj_add(a.g_tau + ck, m * (bx0 - bx1 + cy0 - cy1));
                adj_add(a.g_tau + cw + kr, -m * bx0);
                adj_add(a.g_tau + ce + kr, m * bx1);
                adj_add(a.g_tau + cs + kr, -m * cy0);
                adj_add(a.g_tau + cn + kr, m * cy1);
            }
            adj_add(a.g_K + c, mu * S);
        }
    }
}

// ---------------------------------------------------------------------------
// host side
// ---------------------------------------------------------------------------
namespace {

int make_grid(const vx_grid* in, Grid& g, const char* who) {
    if (!in) return fail(VX_ERR_INVALID_ARGUMENT, "%s: null grid", who);
    if (in->nx < 1 || in->ny < 1 || in->nz < 2)
        return fail(VX_ERR_INVALID_ARGUMENT, "%s: need nx, ny >= 1 and nz >= 2 (got %lld, %lld, %d)", who,
                    (long long)in->nx, (long long)in->ny, in->nz);
    if (in->nz > kMaxNz) return fail(VX_ERR_UNSUPPORTED, "%s: nz=%d exceeds %d", who, in->nz, kMaxNz);
    if (in->nx + 4 > INT32_MAX / 8 || in->ny + 4 > INT32_MAX / 8)
        return fail(VX_ERR_UNSUPPORTED, "%s: horizontal size too large", who);
    if (!in->kbot || !in->dxt || !in->dxu || !in->dyt || !in->dyu || !in->cost || !in->cosu || !in->dzt || !in->dzw)
        return fail(VX_ERR_INVALID_ARGUMENT, "%s: null grid array", who);
    g.kbot = in->kbot;
    g.dxt = in->dxt;
    g.dxu = in->dxu;
    g.dyt = in->dyt;
    g.dyu = in->dyu;
    g.cost = in->cost;
    g.cosu = in->cosu;
    g.dzt = in->dzt;
    g.dzw = in->dzw;
    g.X = (int32_t)(in->nx + 4);
    g.Y = (int32_t)(in->ny + 4);
    if (in->time_layout != VX_LEVELS_INNER && in->time_layout != VX_LEVELS_OUTER)
        return fail(VX_ERR_INVALID_ARGUMENT, "%s: time_layout must be VX_LEVELS_INNER or VX_LEVELS_OUTER (got %d)",
                    who, in->time_layout);
    if ((in->nx + 4) * (in->ny + 4) * (int64_t)in->nz * 3 >= ((int64_t)1 << 32))
        return fail(VX_ERR_UNSUPPORTED, "%s: grid too large for 32-bit offsets", who);
    const bool outer = in->time_layout == VX_LEVELS_OUTER;
    g.cs = outer ? 1u : 3u;
    g.ls = outer ? (uint32_t)((in->nx + 4) * (in->ny + 4) * (int64_t)in->nz) : 1u;
    g.nz = in->nz;
    return VX_OK;
}

Prm make_prm(const vx_iso_params* q) {
    Prm p;
    p.drhodT = q->drhodT;
    p.drhodS = q->drhodS;
    p.K_iso_steep = q->K_iso_steep;
    p.iso_slopec = q->iso_slopec;
    p.iso_dslope = q->iso_dslope;
    p.dt_tracer = q->dt_tracer;
    p.grav = q->grav;
    p.rho_0 = q->rho_0;
    return p;
}

// A helper stream per device (default priority, non-blocking) with a fork and
// a join event, created on first use and kept for the process.
struct AuxStream {
    hipStream_t s = nullptr;
    hipEvent_t fork = nullptr, join = nullptr;
};
std::mutex& aux_mutex() {
    static std::mutex m;
    return m;
}
// Caller holds aux_mutex(). Null when the stream or events cannot be made.
AuxStream* aux_stream() {
    static std::map<int, AuxStream> per_device;
    int dev = 0;
    if (hipGetDevice(&dev) != hipSuccess) return nullptr;
    auto it = per_device.find(dev);
    if (it != per_device.end()) return &it->second;
    AuxStream a;
    if (hipStreamCreateWithFlags(&a.s, hipStreamNonBlocking) != hipSuccess) return nullptr;
    if (hipEventCreateWithFlags(&a.fork, hipEventDisableTiming) != hipSuccess ||
        hipEventCreateWithFlags(&a.join, hipEventDisableTiming) != hipSuccess)
        return nullptr;
    return &(per_device[dev] = a);
}

// Work launched on the returned stream runs after everything enqueued on `s`
// so far; `s` itself when no helper stream can be had. The fork record and
// the wait on it form one step under the lock, so calls from other host
// threads cannot slip their own fork in between.
hipStream_t fork_aux(hipStream_t s) {
    std::lock_guard<std::mutex> lk(aux_mutex());
    AuxStream* ax = aux_stream();
    if (ax && hipEventRecord(ax->fork, s) == hipSuccess && hipStreamWaitEvent(ax->s, ax->fork, 0) == hipSuccess)
        return ax->s;
    return s;
}
// Work enqueued on `s` from here on runs after everything enqueued on the
// helper stream `sd` (a no-op when `sd` is `s`).
int join_aux(hipStream_t s, hipStream_t sd, const char* who) {
    if (sd == s) return VX_OK;
    std::lock_guard<std::mutex> lk(aux_mutex());
    AuxStream* ax = aux_stream();
    if (!ax || hipEventRecord(ax->join, sd) != hipSuccess || hipStreamWaitEvent(s, ax->join, 0) != hipSuccess)
        return fail(VX_ERR_LAUNCH, "%s: joining the helper stream failed", who);
    return VX_OK;
}

// Compute units of the current device (0 if it cannot be queried), cached
// per device.
int device_cu_count() {
    static std::mutex m;
    static std::map<int, int> per_device;
    int dev = 0;
    if (hipGetDevice(&dev) != hipSuccess) return 0;
    std::lock_guard<std::mutex> lk(m);
    auto it = per_device.find(dev);
    if (it != per_device.end()) return it->second;
    int n = 0;
    if (hipDeviceGetAttribute(&n, hipDeviceAttributeMultiprocessorCount, dev) != hipSuccess) n = 0;
    return per_device[dev] = n;
}

#ifndef VX_PRE_CELLS
#define VX_PRE_CELLS 3840
#endif
// Columns per block of the wet-cell pre kernel: about VX_PRE_CELLS cells
// (48 columns at nz = 80), so that a block's wet cells fill its 256 threads
// for several rounds at typical land fractions and the partly filled last
// round wastes few lanes (on 1440x720x80 with 31% wet cells a block holds
// about 1180 wet cells: 5 rounds at 92% of the lanes, against 3 at 77% with
// 24 columns). Measured there, pre + the temperature/salinity step in one
// call: 240 cells per block 11.7 ms, 640 10.1 ms, 1920 9.8 ms (an earlier
// build); 1920 against 3840 cells, four runs each alternating on one box:
// 6.49 / 6.23 / 5.63 / 6.39 against 6.17 / 5.66 / 5.99 / 5.50 ms.
int pre_columns(int nz) {
    const int n = (VX_PRE_CELLS + nz - 1) / nz;
    return n < 1 ? 1 : (n > 64 ? 64 : n);
}

unsigned blocks_for(int64_t n, int bs) { return (unsigned)((n + bs - 1) / bs); }
unsigned blocks_xcd(int64_t n, int bs) { return (unsigned)(((n + bs - 1) / bs + 7) / 8 * 8); }

// block of nz levels x cpb columns (about 256 threads); 1-D grid of ni row
// groups, rounded up to a multiple of 8 for the XCD remap
dim3 cell_block(int nz) { return dim3(nz, nz >= 256 ? 1 : 256 / nz); }
dim3 cell_grid(int ni, int nj, dim3 blk, int sb = kStripBlocks) {
    const int64_t sw = (int64_t)sb * blk.y;
    const int64_t nstrips = (nj + sw - 1) / sw;
    return dim3((unsigned)((nstrips * ni * sb + 7) / 8 * 8));
}

}  // namespace
}  // namespace vx



namespace vx {
namespace {


// The wet-cell plan after the tracers' face fluxes in a step's workspace.
int64_t wet_plan_ncols(int64_t X, int64_t Y) { return wet_enum_cols(X, Y); }
int64_t wet_plan_nchunks(int64_t X, int64_t Y) { return (wet_plan_ncols(X, Y) + kScanChunk - 1) / kScanChunk; }
int64_t wet_plan_bytes(int64_t X, int64_t Y) {
    const int64_t ints = 3 * wet_plan_ncols(X, Y) + 2 + 12;
    return (ints * (int64_t)sizeof(int) + 7) / 8 * 8 + wet_plan_nchunks(X, Y) * (int64_t)sizeof(int2);
}
WetPlan wire_wet_plan(const Grid& g, void* ws, int ntr) {
    const int64_t ncols = wet_plan_ncols(g.X, g.Y);
    char* base = static_cast<char*>(ws) + 3 * ntr * (int64_t)g.X * g.Y * g.nz * (int64_t)sizeof(double);
    WetPlan w;
    int* ip = reinterpret_cast<int*>(base);
    w.wetcol = ip;
    w.wetoff = ip + ncols;
    w.first = ip + 2 * ncols;
    w.nb = ip + 3 * ncols + 2;
    const int64_t ints = 3 * ncols + 2 + 12;
    w.chunk = reinterpret_cast<int2*>(base + (ints * (int64_t)sizeof(int) + 7) / 8 * 8);
    return w;
}

// The wet-cell plan of the step from kbot (three small kernels on `s`).
int launch_wet_plan(const Grid& g, const WetPlan& wp, hipStream_t s, const char* who) {
    const int64_t ncols = wet_plan_ncols(g.X, g.Y);
    const int nchunks = (int)wet_plan_nchunks(g.X, g.Y);
    const int C = kWetThreads - g.nz + 1;  // >= nz as nz <= kMaxNz
    hipLaunchKernelGGL(wet_scan_count_kernel, dim3(nchunks), dim3(kWetThreads), 0, s, g, wp, ncols);
    int rc = check_launch(who);
    if (rc) return rc;
    hipLaunchKernelGGL(wet_scan_chunks_kernel, dim3(1), dim3(kWetThreads), 0, s, wp, nchunks);
    if ((rc = check_launch(who))) return rc;
    hipLaunchKernelGGL(wet_scan_columns_kernel, dim3(nchunks), dim3(kWetThreads), 0, s, g, wp, ncols, C);
    return check_launch(who);
}

#ifndef VX_WET_BPC
#define VX_WET_BPC 5
#endif
// The tracer step over the wet cells: persistent, VX_WET_BPC blocks per CU (as
// many as fit), a multiple of 8 in all (measured on 1440x720x80: 3 blocks per
// CU 9.0 ms for the fused pre + temperature/salinity call, 4 8.45 ms, 5 at
// five waves per SIMD 8.29 ms).
template <int NTR, bool IMPLICIT, bool WETFLUX>
int launch_wet_step(const Grid& g, const Prm& p, const StepArgs& a, const WetPlan& wp, hipStream_t s,
                    const char* who) {
    const int ncu = device_cu_count();
    const unsigned nblk = (unsigned)(((ncu > 0 ? ncu : 256) * VX_WET_BPC + 7) / 8 * 8);
    const size_t lds = (size_t)(3 + kStepArrays * NTR) * kWetThreads * sizeof(double) +
                       (2 * kWetThreads + 1) * sizeof(int);
    hipLaunchKernelGGL((iso_step_wet_kernel<NTR, IMPLICIT, WETFLUX>), dim3(nblk), dim3(kWetThreads), lds, s, g, p, a,
                       wp);
    return check_launch(who);
}

template <int NTR, bool IMPLICIT>
int launch_step(const Grid& g, const Prm& p, const StepArgs& a, void* workspace, hipStream_t s, const char* who) {
    if ((int64_t)g.X * g.Y * g.nz * 4 >= ((int64_t)1 << 32))
        return fail(VX_ERR_UNSUPPORTED, "%s: grid too large for 32-bit offsets", who);
    if (g.nz > kMaxNz) return fail(VX_ERR_UNSUPPORTED, "%s: nz=%d too deep", who, g.nz);
    const WetPlan wp = wire_wet_plan(g, workspace, NTR);
    int rc = launch_wet_plan(g, wp, s, who);
    if (rc) return rc;
    const int ncb = pre_columns(g.nz);
    hipLaunchKernelGGL((iso_flux_kernel<NTR, IMPLICIT>), cell_grid(g.X - 3, g.Y - 3, dim3(1, ncb), kPreStripBlocks),
                       dim3(256), (size_t)(ncb + 1) * sizeof(int), s, g, a, ncb);
    if ((rc = check_launch(who))) return rc;
    return launch_wet_step<NTR, IMPLICIT, false>(g, p, a, wp, s, who);
}

int check_common(const vx_grid* grid, const vx_iso_params* prm, int32_t tau, int32_t taup1, Grid& g,
                 const char* who) {
    int rc = make_grid(grid, g, who);
    if (rc) return rc;
    if (!prm) return fail(VX_ERR_INVALID_ARGUMENT, "%s: null params", who);
    if (tau < 0 || tau > 2 || taup1 < 0 || taup1 > 2)
        return fail(VX_ERR_INVALID_ARGUMENT, "%s: time levels must be 0..2", who);
    return VX_OK;
}

bool tensor_ok(const double* Ai_ez, const double* Ai_nz, const double* Ai_bx, const double* Ai_by,
               const double* K_11, const double* K_22) {
    return Ai_ez && Ai_nz && Ai_bx && Ai_by && K_11 && K_22;
}

// three flux arrays per tracer
// isoneutral_diffusion_pre on `s`: the dry-cell zero fill on a helper stream
// (forked from and joined back into `s`), the wet-cell kernel and the top
// level of K_33 on `s`; FUSED also writes the step's face fluxes into `fo`.
// The dry-cell zero fill (store-bound) and the wet-cell slopes (latency-bound)
// write disjoint cells: the zero fill runs on a helper stream forked from `s`.
// With `dry_stream` null it is joined back into `s` before returning; else the
// helper stream goes to *dry_stream and the caller joins it (join_aux) once
// nothing more it enqueues on `s` reads the dry cells' outputs.
template <bool FUSED>
int launch_pre(const Grid& g, const Prm& p, const Fields& f, const double* K_iso, double* Ai_ez, double* Ai_nz,
               double* Ai_bx, double* Ai_by, double* K_11, double* K_22, double* K_33, PreFlux fo, hipStream_t s,
               const char* who, hipStream_t* dry_stream = nullptr, bool use_aux = true) {
    int rc = VX_OK;
    const int64_t ncol = (int64_t)(g.X - 3) * (g.Y - 3);  // one wave per column, 4 per block
    // without the helper stream (use_aux false) the zero fill runs on `s`
    const hipStream_t sd = use_aux ? fork_aux(s) : s;
    if (dry_stream) *dry_stream = sd;
    // Blocks of the zero fill, in eighths of the CU count (0: one wave per
    // column). Beside the whole one call, fewer blocks than CUs fill in the
    // background at a modest store rate instead of flooding the memory system
    // ahead of the wet-cell kernels (measured on 1440x720x80 with the
    // nontemporal stores, two runs each: 96 blocks 6.34 / 5.61 ms, 128 blocks
    // 5.95 / 5.79, 160 blocks 5.80 / 5.73, 192 blocks 5.80 / 6.38, one block
    // per CU 6.90 / 6.61, two per CU 7.05, one wave per column 6.98). Pre
    // alone is joined at the end of pre, so there the fill must not start
    // late: one wave per column (4.6 ms; one block per CU gave 4.26 ms on one
    // box and a 5.1 ms median on another, where the fill's blocks were
    // dispatched behind the wet cells').
#ifndef VX_DRY_EIGHTHS_PRE
#define VX_DRY_EIGHTHS_PRE 0
#endif
#ifndef VX_DRY_EIGHTHS_FUSED
#define VX_DRY_EIGHTHS_FUSED 5
#endif
    // (on the caller's stream, without the helper, the fill runs before pre:
    // the whole grid of it)
    const unsigned eighths = (dry_stream && use_aux) ? VX_DRY_EIGHTHS_FUSED : VX_DRY_EIGHTHS_PRE;
    const unsigned dry_full = blocks_for(ncol, 4), ndry = eighths * (unsigned)device_cu_count() / 8;
    const unsigned dry_grid = ndry > 0 && ndry < dry_full ? ndry : dry_full;
    hipLaunchKernelGGL(iso_pre_dry_kernel, dim3(dry_grid), dim3(256), 0, sd, g, Ai_ez, Ai_nz, Ai_bx, Ai_by,
                       K_11, K_22, K_33);
    if ((rc = check_launch(who))) return rc;
    const int ncb = pre_columns(g.nz);
    const size_t pre_lds = (size_t)(ncb + 1) * sizeof(int);
    hipLaunchKernelGGL(iso_pre_kernel<FUSED>, cell_grid(g.X - 3, g.Y - 3, dim3(1, ncb), kPreStripBlocks), dim3(256),
                       pre_lds, s, g, p,
                       f, K_iso, Ai_ez, Ai_nz, Ai_bx, Ai_by, K_11, K_22, K_33, fo, ncb);
    if ((rc = check_launch(who))) return rc;
    hipLaunchKernelGGL(k33_top_zero_kernel, dim3(blocks_for((int64_t)g.X * g.Y, 256)), dim3(256), 0, s, g, K_33);
    if ((rc = check_launch(who))) return rc;
    return dry_stream ? VX_OK : join_aux(s, sd, who);
}

void wire_workspace(const Grid& g, StepArgs& a, int ntr, void* ws) {
    const int64_t n3 = (int64_t)g.X * g.Y * g.nz;
    double* w = static_cast<double*>(ws);
    for (int t = 0; t < ntr; ++t) {
        a.t[t].fe = w + (3 * t + 0) * n3;
        a.t[t].fn = w + (3 * t + 1) * n3;
        a.t[t].ft = w + (3 * t + 2) * n3;
    }
}
}  // namespace
}  // namespace vx

extern "C" int vx_iso_diffusion_f64(void* stream, const vx_grid* grid, const vx_iso_params* prm, double* tracer,
                                    int32_t tau, int32_t taup1, const double* K_iso, const double* Ai_ez,
                                    const double* Ai_nz, const double* Ai_bx, const double* Ai_by,
                                    const double* K_11, const double* K_22, const double* K_33, double* dtr_iso,
                                    const double* int_drhodX, double* P_diss, void* workspace) {
    using namespace vx;
    const char* who = "vx_iso_diffusion_f64";
    Grid g;
    int rc = check_common(grid, prm, tau, taup1, g, who);
    if (rc) return rc;
    if (!tracer || !K_iso || !tensor_ok(Ai_ez, Ai_nz, Ai_bx, Ai_by, K_11, K_22) || !K_33 || !dtr_iso)
        return fail(VX_ERR_INVALID_ARGUMENT, "%s: null array pointer", who);
    if ((P_diss == nullptr) != (int_drhodX == nullptr))
        return fail(VX_ERR_INVALID_ARGUMENT, "%s: P_diss and int_drhodX must both be given or both be NULL", who);
    StepArgs a{};
    if (!workspace) return fail(VX_ERR_INVALID_ARGUMENT, "%s: null workspace", who);
    a.t[0] = {tracer, dtr_iso, int_drhodX};
    wire_workspace(g, a, 1, workspace);
    a.tau = tau;
    a.taup1 = taup1;
    a.Kiso = K_iso;
    a.Kskew = nullptr;
    a.Aez = Ai_ez, a.Anz = Ai_nz, a.Abx = Ai_bx, a.Aby = Ai_by, a.K11 = K_11, a.K22 = K_22, a.K33 = K_33;
    a.P = P_diss;
    return launch_step<1, true>(g, make_prm(prm), a, workspace, (hipStream_t)stream, who);
}

extern "C" int vx_iso_diffusion_ts_f64(void* stream, const vx_grid* grid, const vx_iso_params* prm, double* temp,
                                       double* salt, int32_t tau, int32_t taup1, const double* K_iso,
                                       const double* Ai_ez, const double* Ai_nz, const double* Ai_bx,
                                       const double* Ai_by, const double* K_11, const double* K_22,
                                       const double* K_33, double* dtemp_iso, double* dsalt_iso,
                                       const double* int_drhodT, const double* int_drhodS, double* P_diss,
                                       void* workspace) {
    using namespace vx;
    const char* who = "vx_iso_diffusion_ts_f64";
    Grid g;
    int rc = check_common(grid, prm, tau, taup1, g, who);
    if (rc) return rc;
    if (!temp || !salt || !K_iso || !tensor_ok(Ai_ez, Ai_nz, Ai_bx, Ai_by, K_11, K_22) || !K_33 || !dtemp_iso ||
        !dsalt_iso)
        return fail(VX_ERR_INVALID_ARGUMENT, "%s: null array pointer", who);
    if (temp == salt) return fail(VX_ERR_INVALID_ARGUMENT, "%s: temp and salt must be distinct arrays", who);
    const bool e = P_diss != nullptr;
    if (e != (int_drhodT != nullptr) || e != (int_drhodS != nullptr))
        return fail(VX_ERR_INVALID_ARGUMENT, "%s: P_diss, int_drhodT and int_drhodS must all be given or all be NULL",
                    who);
    StepArgs a{};
    if (!workspace) return fail(VX_ERR_INVALID_ARGUMENT, "%s: null workspace", who);
    a.t[0] = {temp, dtemp_iso, int_drhodT};
    a.t[1] = {salt, dsalt_iso, int_drhodS};
    wire_workspace(g, a, 2, workspace);
    a.tau = tau;
    a.taup1 = taup1;
    a.Kiso = K_iso;
    a.Kskew = nullptr;
    a.Aez = Ai_ez, a.Anz = Ai_nz, a.Abx = Ai_bx, a.Aby = Ai_by, a.K11 = K_11, a.K22 = K_22, a.K33 = K_33;
    a.P = P_diss;
    return launch_step<2, true>(g, make_prm(prm), a, workspace, (hipStream_t)stream, who);
}

extern "C" int vx_iso_skew_diffusion_f64(void* stream, const vx_grid* grid, const vx_iso_params* prm,
                                         double* tracer, int32_t tau, int32_t taup1, const double* K_gm,
                                         const double* Ai_ez, const double* Ai_nz, const double* Ai_bx,
                                         const double* Ai_by, const double* K_11, const double* K_22,
                                         double* dtr_iso, const double* int_drhodX, double* P_diss_skew,
                                         void* workspace) {
    using namespace vx;
    const char* who = "vx_iso_skew_diffusion_f64";
    Grid g;
    int rc = check_common(grid, prm, tau, taup1, g, who);
    if (rc) return rc;
    if (!tracer || !K_gm || !tensor_ok(Ai_ez, Ai_nz, Ai_bx, Ai_by, K_11, K_22) || !dtr_iso)
        return fail(VX_ERR_INVALID_ARGUMENT, "%s: null array pointer", who);
    if ((P_diss_skew == nullptr) != (int_drhodX == nullptr))
        return fail(VX_ERR_INVALID_ARGUMENT, "%s: P_diss_skew and int_drhodX must both be given or both be NULL", who);
    StepArgs a{};
    if (!workspace) return fail(VX_ERR_INVALID_ARGUMENT, "%s: null workspace", who);
    a.t[0] = {tracer, dtr_iso, int_drhodX};
    wire_workspace(g, a, 1, workspace);
    a.tau = tau;
    a.taup1 = taup1;
    a.Kiso = nullptr;
    a.Kskew = K_gm;
    a.Aez = Ai_ez, a.Anz = Ai_nz, a.Abx = Ai_bx, a.Aby = Ai_by, a.K11 = K_11, a.K22 = K_22, a.K33 = nullptr;
    a.P = P_diss_skew;
    return launch_step<1, false>(g, make_prm(prm), a, workspace, (hipStream_t)stream, who);
}

extern "C" int64_t vx_iso_workspace_bytes(const vx_grid* grid, int32_t ntracers) {
    if (!grid || ntracers < 1 || ntracers > 2 || grid->nx < 1 || grid->ny < 1 || grid->nz < 1) return -1;
    return 3 * (int64_t)ntracers * (grid->nx + 4) * (grid->ny + 4) * (int64_t)grid->nz * (int64_t)sizeof(double) +
           vx::wet_plan_bytes(grid->nx + 4, grid->ny + 4);
}

extern "C" int64_t vx_iso_pre_diffusion_ts_workspace_bytes(const vx_grid* grid) {
    return vx_iso_workspace_bytes(grid, 2);
}

extern "C" int64_t vx_iso_adjoint_workspace_bytes(const vx_grid* grid) {
    if (!grid) return -1;
    return 4 * (grid->nx + 4) * (grid->ny + 4) * (int64_t)grid->nz * (int64_t)sizeof(double);
}

extern "C" int vx_iso_diffusion_adjoint_f64(void* stream, const vx_grid* grid, const vx_iso_params* prm,
                                            const double* tracer, int32_t tau, const double* K_iso,
                                            const double* Ai_ez, const double* Ai_nz, const double* Ai_bx,
                                            const double* Ai_by, const double* K_11, const double* K_22,
                                            const double* K_33, const double* g_new, double* g_tau,
                                            double* g_taup1, double* g_K_iso, void* workspace) {
    using namespace vx;
    const char* who = "vx_iso_diffusion_adjoint_f64";
    Grid g;
    int rc = check_common(grid, prm, tau, 0, g, who);
    if (rc) return rc;
    if (!tracer || !K_iso || !tensor_ok(Ai_ez, Ai_nz, Ai_bx, Ai_by, K_11, K_22) || !K_33 || !g_new || !g_tau ||
        !g_taup1 || !g_K_iso || !workspace)
        return fail(VX_ERR_INVALID_ARGUMENT, "%s: null array pointer", who);
    if ((int64_t)g.X * g.Y * g.nz * 4 >= ((int64_t)1 << 32))
        return fail(VX_ERR_UNSUPPORTED, "%s: grid too large for 32-bit offsets", who);
    const int64_t n3 = (int64_t)g.X * g.Y * g.nz;
    double* w = static_cast<double*>(workspace);
    AdjArgs a{tracer, tau, K_iso, Ai_ez, Ai_nz, Ai_bx, Ai_by, K_11, K_22, K_33, g_new, g_tau, g_taup1, g_K_iso,
              w, w + n3, w + 2 * n3, w + 3 * n3};
    hipStream_t s = (hipStream_t)stream;
    const Prm p = make_prm(prm);
    hipLaunchKernelGGL(iso_adj_init_kernel, dim3(blocks_for(n3, 256)), dim3(256), 0, s, n3, a);
    if ((rc = check_launch(who))) return rc;
    const int64_t ncol = (int64_t)(g.X - 4) * (g.Y - 4);
    const int nzb = ((g.nz + 15) / 16) * 16;
    const dim3 cgrid(blocks_for(ncol, 256)), cblk(256);
#define VX_ADJ_CASE(N) \
    case N: hipLaunchKernelGGL(iso_adj_column_kernel<N>, cgrid, cblk, 0, s, g, p, a); break;
    switch (nzb) {
        VX_ADJ_CASE(16)
        VX_ADJ_CASE(32)
        VX_ADJ_CASE(48)
        VX_ADJ_CASE(64)
        VX_ADJ_CASE(80)
        VX_ADJ_CASE(96)
        VX_ADJ_CASE(112)
        VX_ADJ_CASE(128)
        default: return fail(VX_ERR_UNSUPPORTED, "%s: nz=%d", who, g.nz);
    }
#undef VX_ADJ_CASE
    if ((rc = check_launch(who))) return rc;
    const dim3 blk = cell_block(g.nz);
    hipLaunchKernelGGL(iso_adj_faces_kernel, cell_grid(g.X - 3, g.Y - 3, blk), blk, 0, s, g, a);
    if ((rc = check_launch(who))) return rc;
    hipLaunchKernelGGL(iso_adj_scatter_kernel, cell_grid(g.X - 3, g.Y - 3, blk), blk, 0, s, g, a);
    return check_launch(who);
}

namespace vx {
namespace {

// The one call (pre, then the temperature and salinity steps) with the given
// workspace; `use_aux` false keeps every launch on `s` (no helper stream).
int pre_diffusion_ts(hipStream_t s, const vx_grid* grid, const vx_iso_params* prm, double* temp, double* salt,
                     int32_t tau, int32_t taup1, const double* K_iso, double* Ai_ez, double* Ai_nz, double* Ai_bx,
                     double* Ai_by, double* K_11, double* K_22, double* K_33, double* dtemp_iso, double* dsalt_iso,
                     const double* int_drhodT, const double* int_drhodS, double* P_diss, void* workspace,
                     bool use_aux, const char* who) {
    Grid g;
    int rc = check_common(grid, prm, tau, taup1, g, who);
    if (rc) return rc;
    if (!temp || !salt || !K_iso || !tensor_ok(Ai_ez, Ai_nz, Ai_bx, Ai_by, K_11, K_22) || !K_33 || !dtemp_iso ||
        !dsalt_iso)
        return fail(VX_ERR_INVALID_ARGUMENT, "%s: null array pointer", who);
    if (temp == salt) return fail(VX_ERR_INVALID_ARGUMENT, "%s: temp and salt must be distinct arrays", who);
    const bool e = P_diss != nullptr;
    if (e != (int_drhodT != nullptr) || e != (int_drhodS != nullptr))
        return fail(VX_ERR_INVALID_ARGUMENT, "%s: P_diss, int_drhodT and int_drhodS must all be given or all be NULL",
                    who);
    if (!workspace) return fail(VX_ERR_INVALID_ARGUMENT, "%s: null workspace", who);
    if ((int64_t)g.X * g.Y * g.nz * 4 >= ((int64_t)1 << 32))
        return fail(VX_ERR_UNSUPPORTED, "%s: grid too large for 32-bit offsets", who);
    if (g.nz > kMaxNz) return fail(VX_ERR_UNSUPPORTED, "%s: nz=%d too deep", who, g.nz);
    StepArgs a{};
    a.t[0] = {temp, dtemp_iso, int_drhodT};
    a.t[1] = {salt, dsalt_iso, int_drhodS};
    wire_workspace(g, a, 2, workspace);
    a.tau = tau;
    a.taup1 = taup1;
    a.Kiso = K_iso;
    a.Kskew = nullptr;
    a.Aez = Ai_ez, a.Anz = Ai_nz, a.Abx = Ai_bx, a.Aby = Ai_by, a.K11 = K_11, a.K22 = K_22, a.K33 = K_33;
    a.P = P_diss;
    const Prm p = make_prm(prm);
    const PreFlux fo{{a.t[0].fe, a.t[1].fe}, {a.t[0].fn, a.t[1].fn}, {a.t[0].ft, a.t[1].ft}};
    // the wet-cell plan of the step (reads kbot only)
    const WetPlan wp = wire_wet_plan(g, workspace, 2);
    if ((rc = launch_wet_plan(g, wp, s, who))) return rc;
    // The step kernel reads no output of a dry cell (its fluxes come from the
    // wet-cell kernel, and it reads K_33 only on wet cells), so the dry-cell
    // zero fill overlaps the whole call and is joined at its end.
    hipStream_t sd = s;
    rc = launch_pre<true>(g, p, Fields{temp, salt, tau}, K_iso, Ai_ez, Ai_nz, Ai_bx, Ai_by, K_11, K_22, K_33, fo, s,
                          who, &sd, use_aux);
    if (!rc) rc = launch_wet_step<2, true, true>(g, p, a, wp, s, who);
    const int rj = join_aux(s, sd, who);
    return rc ? rc : rj;
}

}  // namespace
}  // namespace vx

extern "C" int vx_iso_pre_diffusion_ts_f64(void* stream, const vx_grid* grid, const vx_iso_params* prm, double* temp,
                                           double* salt, int32_t tau, int32_t taup1, const double* K_iso,
                                           double* Ai_ez, double* Ai_nz, double* Ai_bx, double* Ai_by, double* K_11,
                                           double* K_22, double* K_33, double* dtemp_iso, double* dsalt_iso,
                                           const double* int_drhodT, const double* int_drhodS, double* P_diss,
                                           void* workspace) {
    return vx::pre_diffusion_ts((hipStream_t)stream, grid, prm, temp, salt, tau, taup1, K_iso, Ai_ez, Ai_nz, Ai_bx,
                                Ai_by, K_11, K_22, K_33, dtemp_iso, dsalt_iso, int_drhodT, int_drhodS, P_diss,
                                workspace, true, "vx_iso_pre_diffusion_ts_f64");
}

// ---------------------------------------------------------------------------
// The one call without a caller workspace: the face fluxes and the wet-cell
// plan live in `scratch`, a device array of the library (one per GPU, held
// until the process ends). Its device address is taken once per device with
// hipGetSymbolAddress; host code never passes the array's own name to a
// kernel. The entry point allocates nothing: no hipMalloc, no helper stream.
// ---------------------------------------------------------------------------
namespace {
constexpr int64_t kNoAllocScratchBytes = int64_t(192) << 20;  // 192 MiB
}
static __device__ __align__(256) unsigned char scratch[kNoAllocScratchBytes];

namespace vx {
namespace {
// Device address of `scratch` on the current device, looked up on the first
// call on that device (null and an error message if the lookup fails).
void* scratch_on_current_device(const char* who) {
    static std::mutex m;
    static std::map<int, void*> per_device;
    int dev = 0;
    if (hipGetDevice(&dev) != hipSuccess) {
        fail(VX_ERR_LAUNCH, "%s: hipGetDevice failed", who);
        return nullptr;
    }
    std::lock_guard<std::mutex> lk(m);
    auto it = per_device.find(dev);
    if (it != per_device.end()) return it->second;
    void* ptr = nullptr;
    const hipError_t e = hipGetSymbolAddress(&ptr, HIP_SYMBOL(scratch));
    if (e != hipSuccess || ptr == nullptr) {
        fail(VX_ERR_LAUNCH, "%s: hipGetSymbolAddress(scratch) failed: %s", who, hipGetErrorString(e));
        return nullptr;
    }
    per_device[dev] = ptr;
    return ptr;
}
}  // namespace
}  // namespace vx

extern "C" int vx_iso_pre_diffusion_ts_f64_noalloc(void* stream, const vx_grid* grid, const vx_iso_params* prm,
                                                   double* temp, double* salt, int32_t tau, int32_t taup1,
                                                   const double* K_iso, double* Ai_ez, double* Ai_nz, double* Ai_bx,
                                                   double* Ai_by, double* K_11, double* K_22, double* K_33,
                                                   double* dtemp_iso, double* dsalt_iso, const double* int_drhodT,
                                                   const double* int_drhodS, double* P_diss) {
    using namespace vx;
    const char* who = "vx_iso_pre_diffusion_ts_f64_noalloc";
    const int64_t need = vx_iso_pre_diffusion_ts_workspace_bytes(grid);
    if (need < 0) return fail(VX_ERR_INVALID_ARGUMENT, "%s: invalid grid", who);
    if (need > kNoAllocScratchBytes)
        return fail(VX_ERR_UNSUPPORTED,
                    "%s: the grid needs %lld bytes of scratch, above the library's fixed %lld (use "
                    "vx_iso_pre_diffusion_ts_f64 with a workspace)",
                    who, (long long)need, (long long)kNoAllocScratchBytes);
    void* ws = scratch_on_current_device(who);
    if (!ws) return VX_ERR_LAUNCH;
    return pre_diffusion_ts((hipStream_t)stream, grid, prm, temp, salt, tau, taup1, K_iso, Ai_ez, Ai_nz, Ai_bx,
                            Ai_by, K_11, K_22, K_33, dtemp_iso, dsalt_iso, int_drhodT, int_drhodS, P_diss, ws, false,
                            who);
}

extern "C" int vx_iso_pre_f64(void* stream, const vx_grid* grid, const vx_iso_params* prm, const double* temp,
                              const double* salt, int32_t tau, const double* K_iso, double* Ai_ez, double* Ai_nz,
                              double* Ai_bx, double* Ai_by, double* K_11, double* K_22, double* K_33) {
    using namespace vx;
    Grid g;
    int rc = make_grid(grid, g, "vx_iso_pre_f64");
    if (rc) return rc;
    if (!prm) return fail(VX_ERR_INVALID_ARGUMENT, "vx_iso_pre_f64: null params");
    if (tau < 0 || tau > 2) return fail(VX_ERR_INVALID_ARGUMENT, "vx_iso_pre_f64: tau must be 0..2");
    if (!temp || !salt || !K_iso || !Ai_ez || !Ai_nz || !Ai_bx || !Ai_by || !K_11 || !K_22 || !K_33)
        return fail(VX_ERR_INVALID_ARGUMENT, "vx_iso_pre_f64: null array pointer");
    hipStream_t s = (hipStream_t)stream;
    Fields f{temp, salt, tau};
    const Prm p = make_prm(prm);
    if ((int64_t)g.X * g.Y * g.nz * 4 >= ((int64_t)1 << 32))
        return fail(VX_ERR_UNSUPPORTED, "vx_iso_pre_f64: grid too large for 32-bit offsets");
    return launch_pre<false>(g, p, f, K_iso, Ai_ez, Ai_nz, Ai_bx, Ai_by, K_11, K_22, K_33, PreFlux{}, s,
                             "vx_iso_pre_f64");
}
