// Implicit vertical integration of TKE, EKE and internal-wave energy on
// gfx950: integrate_tke_kernel (veros/core/tke.py:123-321),
// integrate_eke_kernel (veros/core/eke.py:93-255) and integrate_idemix_kernel
// (veros/core/idemix.py:47-265), without their advection branches.
//
// One kernel per model. A group of kSegP lanes owns one (i, j) column of the
// whole (X, Y) grid and lane s holds levels [s*S, s*S + S) (tdma_seg.hpp):
// each lane sums the forcing terms, forms its rows of the implicit system
// (delta from the diffusivity at k and k+1, the dissipation on the diagonal,
// the edge row at kbot-1, the surface flux in the top rhs) in registers,
// the group solves the column by the segmented elimination, and each lane
// then writes its levels: the new energy at taup1 on the wet levels of the
// interior columns, the dissipation diagnostic over the whole array, the
// lateral diffusion of the tau level added on top (its fluxes read the four
// neighbouring columns), and for TKE the surface correction. Halo columns
// only produce the diagnostics. Masks come from kbot (maskT = maskW = wet
// suffix, maskU / maskV the products of the two adjacent maskT).
#include "tdma_seg.hpp"

namespace vx {

constexpr int kMaxTerms = 12;
// threads per block of the three kernels; a constant, since reading the
// block size at run time costs a memory round trip at the start of every wave
constexpr int kTurbBlock = 256;

struct TurbGrid {
    const int32_t* kbot;
    const double *dxt, *dxu, *dyt, *dyu, *cost, *cosu, *dzt, *dzw;
    int32_t X, Y, nz;
    int64_t cs, ls;  // time levels: level l of cell c at l * ls + c * cs
};

// forcing = sgn[0] * p[0] + sgn[1] * p[1] + ... , summed left to right
struct Terms {
    const double* p[kMaxTerms];
    double sgn[kMaxTerms];
    int n;
};

__device__ __forceinline__ double sum_terms(const Terms& T, int64_t c) {
    double f = 0.0;
    for (int q = 0; q < T.n; ++q) f = q == 0 ? T.sgn[0] * T.p[0][c] : f + T.sgn[q] * T.p[q][c];
    return f;
}

__device__ __forceinline__ bool wet_at(int kb, int k) { return kb > 0 && k >= kb - 1; }

// The column of a lane group, over the whole (X, Y) grid.
struct TurbCol {
    int i, j, s, lo;
    int64_t col;    // i * Y + j
    bool valid;     // inside the grid
    bool interior;  // 2 <= i < X-2, 2 <= j < Y-2
    int kb;         // kbot
    int k0;         // first solved level (nz: none)
};

__device__ __forceinline__ TurbCol turb_col(const TurbGrid& g, int S) {
    TurbCol t;
    const int64_t gt = xcd_contiguous_block() * kTurbBlock + threadIdx.x;
    const int64_t col = gt / kSegP;
    t.s = (int)(gt % kSegP);
    t.lo = t.s * S;
    t.valid = col < (int64_t)g.X * g.Y;
    t.col = t.valid ? col : 0;
    t.i = (int)(t.col / g.Y);
    t.j = (int)(t.col - (int64_t)t.i * g.Y);
    t.interior = t.valid && t.i >= 2 && t.i < g.X - 2 && t.j >= 2 && t.j < g.Y - 2;
    t.kb = t.valid ? g.kbot[t.col] : 0;
    t.k0 = (t.interior && t.kb >= 1 && t.kb <= g.nz) ? t.kb - 1 : g.nz;
    return t;
}

// Lateral diffusion of the tau level at (i, j, k) of an interior column
// (veros/core/tke.py:250-276; eke.py:168-200; idemix.py:186-218):
//   fe(i) = ce(i) * (F(i+1) - F(i)) / (cost dxu(i)) * maskU(i)
//   fn(j) = cn(j) * (F(j+1) - F(j)) / dyu(j) * maskV(j) * cosu(j)
//   div   = (fe(i) - fe(i-1)) / (cost dxt) + (fn(j) - fn(j-1)) / (cost dyt)
// F(ii, jj) and the coefficients come from functors of the cell offset.
template <class FV, class CE, class CN>
__device__ __forceinline__ double lateral_div(const TurbGrid& g, int i, int j, int k, FV F, CE ce, CN cn) {
    const int64_t Y = g.Y, nz = g.nz;
    const int64_t c = ((int64_t)i * Y + j) * nz + k;
    const int64_t cE = c + Y * nz, cW = c - Y * nz, cN = c + nz, cS = c - nz;
    const bool mC = wet_at(g.kbot[(int64_t)i * Y + j], k);
    const bool mE = wet_at(g.kbot[(int64_t)(i + 1) * Y + j], k), mW = wet_at(g.kbot[(int64_t)(i - 1) * Y + j], k);
    const bool mN = wet_at(g.kbot[(int64_t)i * Y + j + 1], k), mS = wet_at(g.kbot[(int64_t)i * Y + j - 1], k);
    const double cost = g.cost[j];
    const double Fc = F(c), Fe = F(cE), Fw = F(cW), Fn = F(cN), Fs = F(cS);
    const double fe = ce(c, cE) * (Fe - Fc) / (cost * g.dxu[i]) * ((mC && mE) ? 1.0 : 0.0);
    const double fw = ce(cW, c) * (Fc - Fw) / (cost * g.dxu[i - 1]) * ((mW && mC) ? 1.0 : 0.0);
    const double fn = cn(c, cN) * (Fn - Fc) / g.dyu[j] * ((mC && mN) ? 1.0 : 0.0) * g.cosu[j];
    const double fs = cn(cS, c) * (Fc - Fs) / g.dyu[j - 1] * ((mS && mC) ? 1.0 : 0.0) * g.cosu[j - 1];
    return (fe - fw) / (cost * g.dxt[i]) + (fn - fs) / (cost * g.dyt[j]);
}

__device__ __forceinline__ double2 ld2(const double* __restrict__ p) { return *reinterpret_cast<const double2*>(p); }

// lateral_div at the levels k and k+1 together: F2 and the coefficient
// functors return the two levels' values from 16-byte loads (levels
// outermost, k even), the arithmetic of each level that of lateral_div.
template <class FV2, class CE2, class CN2>
__device__ __forceinline__ double2 lateral_div_pair(const TurbGrid& g, int i, int j, int k, FV2 F, CE2 ce, CN2 cn) {
    const int64_t Y = g.Y, nz = g.nz;
    const int64_t c = ((int64_t)i * Y + j) * nz + k;
    const int64_t cE = c + Y * nz, cW = c - Y * nz, cN = c + nz, cS = c - nz;
    const int kC = g.kbot[(int64_t)i * Y + j], kE = g.kbot[(int64_t)(i + 1) * Y + j];
    const int kW = g.kbot[(int64_t)(i - 1) * Y + j], kN = g.kbot[(int64_t)i * Y + j + 1];
    const int kS = g.kbot[(int64_t)i * Y + j - 1];
    const double cost = g.cost[j], dxu = g.dxu[i], dxuw = g.dxu[i - 1], dyu = g.dyu[j], dyus = g.dyu[j - 1];
    const double cosu = g.cosu[j], cosus = g.cosu[j - 1], dxt = g.dxt[i], dyt = g.dyt[j];
    const double2 Fc = F(c), Fe = F(cE), Fw = F(cW), Fn = F(cN), Fs = F(cS);
    const double2 ceE = ce(c, cE), ceW = ce(cW, c), cnN = cn(c, cN), cnS = cn(cS, c);
    double r[2];
#pragma unroll
    for (int h = 0; h < 2; ++h) {
        const int kk = k + h;
        const bool mC = wet_at(kC, kk), mE = wet_at(kE, kk), mW = wet_at(kW, kk);
        const bool mN = wet_at(kN, kk), mS = wet_at(kS, kk);
        const double fc = h ? Fc.y : Fc.x, fe_ = h ? Fe.y : Fe.x, fw_ = h ? Fw.y : Fw.x;
        const double fn_ = h ? Fn.y : Fn.x, fs_ = h ? Fs.y : Fs.x;
        const double fe = (h ? ceE.y : ceE.x) * (fe_ - fc) / (cost * dxu) * ((mC && mE) ? 1.0 : 0.0);
        const double fw = (h ? ceW.y : ceW.x) * (fc - fw_) / (cost * dxuw) * ((mW && mC) ? 1.0 : 0.0);
        const double fn = (h ? cnN.y : cnN.x) * (fn_ - fc) / dyu * ((mC && mN) ? 1.0 : 0.0) * cosu;
        const double fs = (h ? cnS.y : cnS.x) * (fc - fs_) / dyus * ((mS && mC) ? 1.0 : 0.0) * cosus;
        r[h] = (fe - fw) / (cost * dxt) + (fn - fs) / (cost * dyt);
    }
    return make_double2(r[0], r[1]);
}

// The lateral term of the lane's levels where need(k) (0 elsewhere): per
// level with lateral_div, or with VEC per level pair with lateral_div_pair.
template <int S, bool VEC, class NEED, class L1, class L2>
__device__ __forceinline__ void lateral_levels(int lo, NEED need, L1 one, L2 pair, double (&lat)[S]) {
#pragma unroll
    for (int t = 0; t < S; ++t) lat[t] = 0.0;
    if constexpr (VEC) {
#pragma unroll
        for (int t = 0; t < S; t += 2) {
            const int k = lo + t;
            if (need(k) || need(k + 1)) {
                const double2 v = pair(k);
                lat[t] = v.x;
                lat[t + 1] = v.y;
            }
        }
    } else {
#pragma unroll
        for (int t = 0; t < S; ++t) {
            if (need(lo + t)) lat[t] = one(lo + t);
        }
    }
}

// dzt (s[0, nz)) and dzw (s[kMaxNz, kMaxNz + nz)) of the grid in LDS; every
// thread of the block calls it. The row loops read the spacings of their
// levels inside per-lane branches, where a global load is waited for at once:
// from LDS that wait is short.
__device__ __forceinline__ void stage_dz(const TurbGrid& g, double* s) {
    for (int k = threadIdx.x; k < g.nz; k += kTurbBlock) {
        s[k] = g.dzt[k];
        s[kMaxNz + k] = g.dzw[k];
    }
    __syncthreads();
}

// levels lo-1 .. lo+S of a (X, Y, nz) array at column `base` (0 outside
// [0, nz)), in two steps so that other loads can be issued between them:
// load_halo_issue loads the lane's levels (VEC: its own, in pairs; otherwise
// all S+2), and load_halo_finish takes lo-1 and lo+S from the lanes below
// and above (VEC; every lane must call it).
template <int S, bool VEC>
__device__ __forceinline__ void load_halo_issue(const double* __restrict__ p, int64_t base, int lo, int nz,
                                                double (&v)[S + 2]) {
    if constexpr (VEC) {
        double w[S];
#pragma unroll
        for (int t = 0; t < S; ++t) w[t] = 0.0;
        col_load<S, true>(p + base, 1, lo, nz, [](int) { return true; }, w);
        v[0] = v[S + 1] = 0.0;
#pragma unroll
        for (int t = 0; t < S; ++t) v[t + 1] = w[t];
    } else {
#pragma unroll
        for (int t = 0; t < S + 2; ++t) {
            const int k = lo - 1 + t;
            v[t] = (k >= 0 && k < nz) ? p[base + k] : 0.0;
        }
    }
}
template <int S, bool VEC>
__device__ __forceinline__ void load_halo_finish(int s, double (&v)[S + 2]) {
    if constexpr (VEC) {
        v[0] = grp_prev(s, v[S]);
        v[S + 1] = grp_next(s, v[1]);
    }
}

// levels lo .. lo+S-1 (0 from nz on)
template <int S, bool VEC>
__device__ __forceinline__ void load_lv(const double* __restrict__ p, int64_t base, int lo, int nz, double (&v)[S]) {
#pragma unroll
    for (int t = 0; t < S; ++t) v[t] = 0.0;
    col_load<S, VEC>(p + base, 1, lo, nz, [](int) { return true; }, v);
}

#ifndef VX_TERMS_GROUP
#define VX_TERMS_GROUP 3
#endif
// f (levels lo .. lo+S-1) plus the sum of the forcing terms where want(k),
// summed left to right as sum_terms; with `first` the sum replaces f. The
// terms are loaded VX_TERMS_GROUP at a time, all loads of a group before its
// first use, so a group costs one memory round trip instead of one per term.
template <int S, bool VEC, class W>
__device__ __forceinline__ void add_terms(const Terms& T, int64_t base, int lo, int nz, W want, bool first,
                                          double (&f)[S]) {
    constexpr int G = VX_TERMS_GROUP;
    for (int q0 = 0; q0 < T.n; q0 += G) {
        double v[G][S];
#pragma unroll
        for (int r = 0; r < G; ++r) {
#pragma unroll
            for (int t = 0; t < S; ++t) v[r][t] = 0.0;
            if (q0 + r < T.n) col_load<S, VEC>(T.p[q0 + r] + base, 1, lo, nz, want, v[r]);
        }
#pragma unroll
        for (int r = 0; r < G; ++r) {
            if (q0 + r < T.n) {
                const double sg = T.sgn[q0 + r];
                const bool set = first && q0 + r == 0;
#pragma unroll
                for (int t = 0; t < S; ++t) f[t] = set ? sg * v[r][t] : f[t] + sg * v[r][t];
            }
        }
    }
}

// sum of the forcing terms at levels lo .. lo+S-1 where want(k), summed
// left to right as sum_terms (0 elsewhere)
template <int S, bool VEC, class W>
__device__ __forceinline__ void load_terms(const Terms& T, int64_t base, int lo, int nz, W want, double (&f)[S]) {
#pragma unroll
    for (int t = 0; t < S; ++t) f[t] = 0.0;
    add_terms<S, VEC>(T, base, lo, nz, want, true, f);
}

// ---------------------------------------------------------------------------
// TKE (veros/core/tke.py:123-321)
// ---------------------------------------------------------------------------
struct TkeArgs {
    TurbGrid g;
    double* tke;  // time levels
    int tau, taup1;
    const double *kappaM, *sqrttke, *mxl, *forc_surf;
    Terms forc;
    double *surf_corr, *diss;
    double dt, alpha, c_eps, K_h;
    int hor;
};

// VEC (all three kernels): levels outermost, nz even and every array
// 16-byte aligned; the lanes then move their levels in 16-byte pairs. PV
// (implied by VEC): nz even and the arrays without time levels aligned, which
// then move in pairs whatever the layout of the energy's time levels.
template <int S, bool VEC, bool PV>
__global__ __launch_bounds__(kTurbBlock) void tke_kernel(TkeArgs a) {
    const TurbGrid& g = a.g;
    const TurbCol tc = turb_col(g, S);
    __shared__ double sdz[2 * kMaxNz];
    stage_dz(g, sdz);
    const double* __restrict__ zt = sdz;           // dzt
    const double* __restrict__ zw = sdz + kMaxNz;  // dzw
    const int nz = g.nz, lo = tc.lo, k0 = tc.k0;
    const int64_t base = tc.col * nz;
    const double* __restrict__ E0 = a.tke + a.tau * g.ls;
    double* __restrict__ E1 = a.tke + a.taup1 * g.ls;
    // VEC: the lateral diffusion's east and west neighbour columns (and, for
    // the wave's first and last groups, the south and north ones outside the
    // wave) are copied into the wave's LDS now, asynchronously and without
    // registers; the other north and south neighbours are the adjacent groups'
    // own columns, taken from their registers after the solve. Lane l's pair
    // of region r lands at lat_stage[wave][r][l], read back by lane l only.
    constexpr int NP = S / 2;
    __shared__ double lat_stage[VEC ? (kTurbBlock / 64) * 3 * NP * 128 : 1];
    double* const stg = lat_stage + (VEC ? (threadIdx.x / 64) * 3 * NP * 128 : 0);
    const int g8 = (int)((threadIdx.x & 63) / kSegP);  // group in the wave
    const bool lat_on = a.hor && tc.interior;
    int kE = 0, kW = 0, kX = 0;
    double m_cost = 1.0, m_dxu = 1.0, m_dxuw = 1.0, m_dyu = 1.0, m_dyus = 1.0, m_cosu = 0.0, m_cosus = 0.0,
           m_dxt = 1.0, m_dyt = 1.0;
    if constexpr (VEC) {
        if (a.hor) {
            const int64_t Y = g.Y;
            const int64_t dE = lat_on ? Y * nz : 0;
            const int64_t dX = lat_on ? (g8 == 0 ? -(int64_t)nz : (g8 == kSegCols - 1 ? (int64_t)nz : 0)) : 0;
            const double* pc = E0 + base;
            // only the pairs that touch the column's wet suffix (the term is
            // taken on wet levels only); the other lanes' slots are not read
            const int kw = tc.kb >= 1 && tc.kb <= nz ? tc.kb - 1 : nz;
#pragma unroll
            for (int t = 0; t < S; t += 2) {
                const int k = lo + t;
                if (lat_on && k < nz && k + 1 >= kw) {
                    __builtin_amdgcn_global_load_lds((const void*)(pc + dE + k),
                                                     (lds_void_ptr)(stg + (0 * NP + t / 2) * 128), 16, 0, 0);
                    __builtin_amdgcn_global_load_lds((const void*)(pc - dE + k),
                                                     (lds_void_ptr)(stg + (1 * NP + t / 2) * 128), 16, 0, 0);
                    __builtin_amdgcn_global_load_lds((const void*)(pc + dX + k),
                                                     (lds_void_ptr)(stg + (2 * NP + t / 2) * 128), 16, 0, 0);
                }
            }
            const int64_t cj = lat_on ? tc.col : (int64_t)Y + 1;  // column (1, 1) when unused
            const int i = (int)(cj / Y), j = (int)(cj - (cj / Y) * Y);
            kE = g.kbot[cj + Y], kW = g.kbot[cj - Y];
            kX = g.kbot[cj + (g8 == 0 ? -1 : 1)];
            m_cost = g.cost[j], m_dxu = g.dxu[i], m_dxuw = g.dxu[i - 1], m_dyu = g.dyu[j], m_dyus = g.dyu[j - 1];
            m_cosu = g.cosu[j], m_cosus = g.cosu[j - 1], m_dxt = g.dxt[i], m_dyt = g.dyt[j];
        }
    }
    double km[S + 2];
    load_halo_issue<S, PV>(a.kappaM, base, lo, nz, km);
    double sq[S], ml[S];
    load_lv<S, PV>(a.sqrttke, base, lo, nz, sq);
    load_lv<S, PV>(a.mxl, base, lo, nz, ml);
    // tau on the wet levels, taup1 on the others (kept there and read by the
    // diagnostic); VEC with lateral diffusion: tau on every column's wet
    // levels, halo columns included, for the neighbours' exchange
    double e0[S], e1[S], fo[S];
#pragma unroll
    for (int t = 0; t < S; ++t) e0[t] = e1[t] = 0.0;
    if (VEC && a.hor)
        col_load<S, VEC>(E0 + base * g.cs, g.cs, lo, nz, [&](int k) { return tc.valid && wet_at(tc.kb, k); }, e0);
    else
        col_load<S, VEC>(E0 + base * g.cs, g.cs, lo, nz, [&](int k) { return k >= k0; }, e0);
    col_load<S, VEC>(E1 + base * g.cs, g.cs, lo, nz, [&](int k) { return tc.valid && k < k0; }, e1);
    load_terms<S, PV>(a.forc, base, lo, nz, [&](int k) { return k >= k0; }, fo);
    const double fsurf = a.forc_surf[tc.col];  // loaded with the rest, used by the top row
    const double dt = a.dt;
    load_halo_finish<S, PV>(tc.s, km);  // after the other loads are issued
    double va[S], vb[S], vc[S], vd[S];
    auto delta = [&](int t) {  // delta at level lo - 1 + t
        const int k = lo - 1 + t;
        return (k >= 0 && k < nz - 1) ? dt / zt[k + 1] * a.alpha * 0.5 * (km[t] + km[t + 1]) : 0.0;
    };
#pragma unroll
    for (int t = 0; t < S; ++t) {
        const int k = lo + t;
        const bool w = tc.valid && k < nz && k >= k0;
        double A = 0.0, B = 1.0, C = 0.0, D = 0.0;
        if (w) {
            const double dk = delta(t + 1), dm = delta(t);
            const double dzw = zw[k];
            D = e0[t] + dt * fo[t];
            if (k == k0) {
                A = 0.0;
                B = 1 + dk / dzw + dt * a.c_eps / ml[t] * sq[t];
            } else if (k == nz - 1) {
                A = -dm / (0.5 * dzw);
                B = 1 + dm / (0.5 * dzw) + dt * a.c_eps / ml[t] * sq[t];
            } else {
                A = -dm / dzw;
                B = 1 + (dk + dm) / dzw + dt * a.c_eps * sq[t] / ml[t];
            }
            C = k < nz - 1 ? -dk / dzw : 0.0;
            if (k == nz - 1) D += dt * fsurf / (0.5 * dzw);
        }
        va[t] = A;
        vb[t] = B;
        vc[t] = C;
        vd[t] = D;
    }
    SegLane L;
    L.s = tc.s;
    L.lo = lo;
    L.k0 = k0;
    L.valid = tc.valid;
    L.base = base;
    seg_solve_f64<S>(L, va, vb, vc, vd);
    // the lateral diffusion of the tau level (:250-276)
    double lat[S];
    const double Kh = a.K_h;
    if constexpr (VEC) {
#pragma unroll
        for (int t = 0; t < S; ++t) lat[t] = 0.0;
        if (a.hor) {
            // the adjacent groups' columns (j-1 and j+1 inside the wave): every
            // lane takes part in the exchange
            const int lane = (int)(threadIdx.x & 63);
            const int kN_in = __shfl(tc.kb, (lane + kSegP) & 63), kS_in = __shfl(tc.kb, (lane - kSegP) & 63);
            double eN[S], eS[S];
#pragma unroll
            for (int t = 0; t < S; ++t) {
                eN[t] = __shfl(e0[t], (lane + kSegP) & 63);
                eS[t] = __shfl(e0[t], (lane - kSegP) & 63);
            }
            seg_wait_vmem();  // the staged neighbours have landed
            if (lat_on) {
                const int kC = tc.kb;
                const int kN = g8 == kSegCols - 1 ? kX : kN_in, kS = g8 == 0 ? kX : kS_in;
#pragma unroll
                for (int t = 0; t < S; ++t) {
                    const int kk = lo + t, h = t & 1;
                    const double* sp = stg + 2 * lane + h;
                    const double fe_ = sp[(0 * NP + t / 2) * 128], fw_ = sp[(1 * NP + t / 2) * 128];
                    const double fx_ = sp[(2 * NP + t / 2) * 128];
                    const double fn_ = g8 == kSegCols - 1 ? fx_ : eN[t], fs_ = g8 == 0 ? fx_ : eS[t];
                    const double fc = e0[t];
                    const bool mC = wet_at(kC, kk), mE = wet_at(kE, kk), mW = wet_at(kW, kk);
                    const bool mN = wet_at(kN, kk), mS = wet_at(kS, kk);
                    const double fe = Kh * (fe_ - fc) / (m_cost * m_dxu) * ((mC && mE) ? 1.0 : 0.0);
                    const double fw = Kh * (fc - fw_) / (m_cost * m_dxuw) * ((mW && mC) ? 1.0 : 0.0);
                    const double fn = Kh * (fn_ - fc) / m_dyu * ((mC && mN) ? 1.0 : 0.0) * m_cosu;
                    const double fs = Kh * (fc - fs_) / m_dyus * ((mS && mC) ? 1.0 : 0.0) * m_cosus;
                    const double r = (fe - fw) / (m_cost * m_dxt) + (fn - fs) / (m_cost * m_dyt);
                    lat[t] = (kk < nz && mC) ? r : 0.0;
                }
            }
        }
    }
    if (!tc.valid) return;
    if constexpr (!VEC) {
        lateral_levels<S, VEC>(
            lo, [&](int k) { return tc.interior && a.hor && k < nz && wet_at(tc.kb, k); },
            [&](int k) {
                return lateral_div(
                    g, tc.i, tc.j, k, [&](int64_t cc) { return E0[cc * g.cs]; },
                    [&](int64_t, int64_t) { return Kh; }, [&](int64_t, int64_t) { return Kh; });
            },
            [&](int k) {
                return lateral_div_pair(
                    g, tc.i, tc.j, k, [&](int64_t cc) { return ld2(E0 + cc); },
                    [&](int64_t, int64_t) { return make_double2(Kh, Kh); },
                    [&](int64_t, int64_t) { return make_double2(Kh, Kh); });
            },
            lat);
    }
    double dg[S];
    bool st[S];
#pragma unroll
    for (int t = 0; t < S; ++t) {
        const int k = lo + t;
        const bool w = k >= k0;
        double v = w ? vd[t] : e1[t];
        dg[t] = a.c_eps / ml[t] * sq[t] * v;  // tke_diss over the whole array (:230)
        bool store = tc.interior && w;
        if (tc.interior && k == nz - 1) {  // surface correction and clamp of every interior column (:235-244)
            a.surf_corr[tc.col] = v < 0.0 ? -v * 0.5 * zw[k] / dt : 0.0;
            v = fmax(0.0, v);
            store = true;
        }
        if (tc.interior && k < nz && a.hor && wet_at(tc.kb, k)) {
            v += dt * 1.0 * lat[t];
            store = true;
        }
        e1[t] = v;
        st[t] = store;
    }
    col_store<S, PV>(a.diss + base, 1, lo, nz, dg);
    col_store_if<S, VEC>(E1 + base * g.cs, g.cs, lo, nz, e1, st);
}

// ---------------------------------------------------------------------------
// EKE (veros/core/eke.py:93-255)
// ---------------------------------------------------------------------------
struct EkeArgs {
    TurbGrid g;
    double* eke;
    int tau, taup1;
    const double *kappaM, *sqrteke, *eke_len, *K_gm;
    Terms forc;
    double *diss_iw, *diss_tke;
    double dt, alpha, c_eps;
};

template <int S, bool VEC, bool PV>
__global__ __launch_bounds__(kTurbBlock) void eke_kernel(EkeArgs a) {
    const TurbGrid& g = a.g;
    const TurbCol tc = turb_col(g, S);
    __shared__ double sdz[2 * kMaxNz];
    stage_dz(g, sdz);
    const double* __restrict__ zt = sdz;           // dzt
    const double* __restrict__ zw = sdz + kMaxNz;  // dzw
    const int nz = g.nz, lo = tc.lo, k0 = tc.k0;
    const int64_t base = tc.col * nz;
    double km[S + 2];
    load_halo_issue<S, PV>(a.kappaM, base, lo, nz, km);
    double sq[S], el[S];
    load_lv<S, PV>(a.sqrteke, base, lo, nz, sq);
    load_lv<S, PV>(a.eke_len, base, lo, nz, el);
    double cint[S];  // eke_c_eps * sqrteke / eke_len * maskW (:121)
#pragma unroll
    for (int t = 0; t < S; ++t) cint[t] = a.c_eps * sq[t] / el[t] * (wet_at(tc.kb, lo + t) ? 1.0 : 0.0);
    const double* __restrict__ E0 = a.eke + a.tau * g.ls;
    double* __restrict__ E1 = a.eke + a.taup1 * g.ls;
    double e0[S], e1[S], fo[S];
#pragma unroll
    for (int t = 0; t < S; ++t) e0[t] = e1[t] = 0.0;
    col_load<S, VEC>(E0 + base * g.cs, g.cs, lo, nz, [&](int k) { return k >= k0; }, e0);
    col_load<S, VEC>(E1 + base * g.cs, g.cs, lo, nz, [&](int k) { return tc.valid && k < k0; }, e1);
    load_terms<S, PV>(a.forc, base, lo, nz, [&](int k) { return k >= k0; }, fo);
    const double dt = a.dt;
    load_halo_finish<S, PV>(tc.s, km);  // after the other loads are issued
    double va[S], vb[S], vc[S], vd[S];
    auto delta = [&](int t) {
        const int k = lo - 1 + t;
        return (k >= 0 && k < nz - 1) ? dt / zt[k + 1] * 0.5 * (km[t] + km[t + 1]) * a.alpha : 0.0;
    };
#pragma unroll
    for (int t = 0; t < S; ++t) {
        const int k = lo + t;
        const bool w = tc.valid && k < nz && k >= k0;
        double A = 0.0, B = 1.0, C = 0.0, D = 0.0;
        if (w) {
            const double dk = delta(t + 1), dm = delta(t);
            const double dzw = zw[k];
            D = e0[t] + dt * fo[t];
            if (k == k0) {
                B = 1 + dk / dzw + dt * cint[t];
            } else if (k == nz - 1) {
                A = -dm / (0.5 * dzw);
                B = 1 + dm / (0.5 * dzw) + dt * cint[t];
            } else {
                A = -dm / dzw;
                B = 1 + (dk + dm) / dzw + dt * cint[t];
            }
            C = k < nz - 1 ? -dk / dzw : 0.0;
        }
        va[t] = A;
        vb[t] = B;
        vc[t] = C;
        vd[t] = D;
    }
    SegLane L;
    L.s = tc.s;
    L.lo = lo;
    L.k0 = k0;
    L.valid = tc.valid;
    L.base = base;
    seg_solve_f64<S>(L, va, vb, vc, vd);
    if (!tc.valid) return;
    // the lateral diffusion of the tau level (:168-200)
    double lat[S];
    {
        const double* __restrict__ K = a.K_gm;
        auto ce2 = [&](int64_t c0, int64_t c1) {
            const double2 p0 = ld2(K + c0), p1 = ld2(K + c1);
            return make_double2(0.5 * fmax(500.0, p0.x + p1.x), 0.5 * fmax(500.0, p0.y + p1.y));
        };
        lateral_levels<S, VEC>(
            lo, [&](int k) { return tc.interior && k < nz && wet_at(tc.kb, k); },
            [&](int k) {
                return lateral_div(
                    g, tc.i, tc.j, k, [&](int64_t cc) { return E0[cc * g.cs]; },
                    [&](int64_t c0, int64_t c1) { return 0.5 * fmax(500.0, K[c0] + K[c1]); },
                    [&](int64_t c0, int64_t c1) { return 0.5 * fmax(500.0, K[c0] + K[c1]); });
            },
            [&](int k) {
                return lateral_div_pair(g, tc.i, tc.j, k, [&](int64_t cc) { return ld2(E0 + cc); }, ce2, ce2);
            },
            lat);
    }
    double di[S], zero[S];
    bool st[S];
#pragma unroll
    for (int t = 0; t < S; ++t) {
        const int k = lo + t;
        double v = k >= k0 ? vd[t] : e1[t];
        di[t] = cint[t] * v;  // (:151)
        zero[t] = 0.0;        // (:152)
        st[t] = false;
        if (tc.interior && k < nz && wet_at(tc.kb, k)) {
            v += dt * 1.0 * lat[t];
            st[t] = true;
        }
        e1[t] = v;
    }
    col_store<S, PV>(a.diss_iw + base, 1, lo, nz, di);
    col_store<S, PV>(a.diss_tke + base, 1, lo, nz, zero);
    col_store_if<S, VEC>(E1 + base * g.cs, g.cs, lo, nz, e1, st);
}

// ---------------------------------------------------------------------------
// IDEMIX (veros/core/idemix.py:47-265)
// ---------------------------------------------------------------------------
struct IdemixArgs {
    TurbGrid g;
    double* E;
    int tau, taup1;
    const double *c0, *v0, *alpha_c, *forc_bot, *forc_surf;
    Terms forc;       // before the optional move to the bottom / surface cells
    Terms forc_post;  // added afterwards (bottom friction)
    double* iw_diss;
    double dt, tau_v, tau_h, frac;
    int hor;
    int diss_mode;  // 0: as is; 1: column integral at the bottom cell; 2: split bottom / surface
};

template <int S, bool VEC, bool PV>
__global__ __launch_bounds__(kTurbBlock) void idemix_kernel(IdemixArgs a) {
    const TurbGrid& g = a.g;
    const TurbCol tc = turb_col(g, S);
    __shared__ double sdz[2 * kMaxNz];
    stage_dz(g, sdz);
    const double* __restrict__ zt = sdz;           // dzt
    const double* __restrict__ zw = sdz + kMaxNz;  // dzw
    const int nz = g.nz;
    const int64_t base = tc.col * nz;
    double c0[S + 2];
    load_halo_issue<S, PV>(a.c0, base, tc.lo, nz, c0);
    double al[S];
    load_lv<S, PV>(a.alpha_c, base, tc.lo, nz, al);
    const double* __restrict__ E0 = a.E + a.tau * g.ls;
    double* __restrict__ E1 = a.E + a.taup1 * g.ls;
    const double dt = a.dt;
    // the bottom and surface forcing of the column, used by the edge rows
    const double fbot = a.forc_bot[tc.col], fsurf = a.forc_surf[tc.col];
    // E_iw at tau on every level, at taup1 off the solve (kept there and read by the diagnostic)
    double e0[S], e1[S], mx[S];
#pragma unroll
    for (int t = 0; t < S; ++t) e0[t] = e1[t] = 0.0;
    col_load<S, VEC>(E0 + base * g.cs, g.cs, tc.lo, nz, [&](int) { return tc.valid; }, e0);
    col_load<S, VEC>(E1 + base * g.cs, g.cs, tc.lo, nz, [&](int k) { return tc.valid && k < tc.k0; }, e1);
#pragma unroll
    for (int t = 0; t < S; ++t) mx[t] = fmax(0.0, e0[t]);  // max(0, E_iw[tau]) (:118)
    // the forcing of the lane's levels (interior columns; wet levels feed the
    // solve, the column integral sums maskW-weighted levels)
    double fo[S];
    load_terms<S, PV>(a.forc, base, tc.lo, nz, [&](int k) { return tc.interior && wet_at(tc.kb, k); }, fo);
#pragma unroll
    for (int t = 0; t < S; ++t) fo[t] = (tc.interior && wet_at(tc.kb, tc.lo + t)) ? fo[t] : 0.0;
    if (a.diss_mode != 0) {  // (:70-103)
        double part = 0.0;
#pragma unroll
        for (int t = 0; t < S; ++t) {
            const int k = tc.lo + t;
            if (k < nz - 1) part += zw[k] * fo[t];
            else if (k == nz - 1) part += 0.5 * fo[t] * zw[k];
        }
        part += __shfl_xor(part, 1, kSegP);
        part += __shfl_xor(part, 2, kSegP);
        part += __shfl_xor(part, 4, kSegP);
        const double a_loc = part;
        const int ks = tc.kb - 1 > 0 ? tc.kb - 1 : 0;
#pragma unroll
        for (int t = 0; t < S; ++t) {
            const int k = tc.lo + t;
            double f = 0.0;
            if (k == ks) f = (a.diss_mode == 1 ? 1.0 : a.frac) * a_loc / zw[k];
            if (a.diss_mode == 2 && k == nz - 1) f = (1.0 - a.frac) * a_loc / (0.5 * zw[k]);
            fo[t] = f;
        }
        // the moved term applies to every interior column; rows of other
        // columns are never solved
    }
    // terms added afterwards, on the levels whose rows are solved
    add_terms<S, PV>(a.forc_post, base, tc.lo, nz, [&](int k) { return k >= tc.k0; }, false, fo);
    load_halo_finish<S, PV>(tc.s, c0);  // after the other loads are issued
    double va[S], vb[S], vc[S], vd[S];
    auto delta = [&](int t) {
        const int k = tc.lo - 1 + t;
        return (k >= 0 && k < nz - 1) ? dt * a.tau_v / zt[k + 1] * 0.5 * (c0[t] + c0[t + 1]) : 0.0;
    };
#pragma unroll
    for (int t = 0; t < S; ++t) {
        const int k = tc.lo + t;
        const bool w = tc.valid && k < nz && k >= tc.k0;
        double A = 0.0, B = 1.0, C = 0.0, D = 0.0;
        if (w) {
            const double dk = delta(t + 1), dm = delta(t);
            const double dzw = zw[k];
            const double diss = dt * al[t] * mx[t];
            const double cc = c0[t + 1];
            D = e0[t] + dt * fo[t];
            if (k == tc.k0) {
                B = 1 + dk / dzw * cc + diss;
                D = D + dt * fbot / dzw;  // d_edge (:157-159)
            } else if (k == nz - 1) {
                A = -dm / (0.5 * dzw) * c0[t];
                B = 1 + dm / (0.5 * dzw) * cc + diss;
                D += dt * fsurf / (0.5 * dzw);
            } else {
                A = -dm * c0[t] / dzw;
                B = 1 + dk * cc / dzw + dm * cc / dzw + diss;
            }
            C = k < nz - 1 ? -dk / dzw * c0[t + 2] : 0.0;
        }
        va[t] = A;
        vb[t] = B;
        vc[t] = C;
        vd[t] = D;
    }
    SegLane L;
    L.s = tc.s;
    L.lo = tc.lo;
    L.k0 = tc.k0;
    L.valid = tc.valid;
    L.base = base;
    seg_solve_f64<S>(L, va, vb, vc, vd);
    if (!tc.valid) return;
    // the lateral diffusion of the tau level (:186-218)
    double lat[S];
    {
        const double* __restrict__ V = a.v0;
        const double th = a.tau_h;
        auto ce2 = [&](int64_t c0_, int64_t c1) {
            const double2 p1 = ld2(V + c1), p0 = ld2(V + c0_);
            return make_double2(th * 0.5 * (p1.x + p0.x), th * 0.5 * (p1.y + p0.y));
        };
        lateral_levels<S, VEC>(
            tc.lo, [&](int k) { return tc.interior && k < nz && a.hor && wet_at(tc.kb, k); },
            [&](int k) {
                return lateral_div(
                    g, tc.i, tc.j, k, [&](int64_t cc) { return V[cc] * E0[cc * g.cs]; },
                    [&](int64_t c0_, int64_t c1) { return th * 0.5 * (V[c1] + V[c0_]); },
                    [&](int64_t c0_, int64_t c1) { return th * 0.5 * (V[c1] + V[c0_]); });
            },
            [&](int k) {
                return lateral_div_pair(
                    g, tc.i, tc.j, k,
                    [&](int64_t cc) {
                        const double2 v = ld2(V + cc), e = ld2(E0 + cc);
                        return make_double2(v.x * e.x, v.y * e.y);
                    },
                    ce2, ce2);
            },
            lat);
    }
    double di[S];
    bool st[S];
#pragma unroll
    for (int t = 0; t < S; ++t) {
        const int k = tc.lo + t;
        double v = k >= tc.k0 ? vd[t] : e1[t];
        di[t] = al[t] * mx[t] * v;  // (:174)
        st[t] = tc.interior && k >= tc.k0;
        if (tc.interior && k < nz && a.hor && wet_at(tc.kb, k)) {
            v += dt * 1.0 * lat[t];
            st[t] = true;
        }
        e1[t] = v;
    }
    col_store<S, PV>(a.iw_diss + base, 1, tc.lo, nz, di);
    col_store_if<S, VEC>(E1 + base * g.cs, g.cs, tc.lo, nz, e1, st);
}

namespace {

int make_turb_grid(const vx_grid* in, TurbGrid& g, const char* who) {
    if (!in) return fail(VX_ERR_INVALID_ARGUMENT, "%s: null grid", who);
    if (in->nx < 1 || in->ny < 1 || in->nz < 2)
        return fail(VX_ERR_INVALID_ARGUMENT, "%s: need nx, ny >= 1 and nz >= 2", who);
    if (in->nz > kMaxNz) return fail(VX_ERR_UNSUPPORTED, "%s: nz=%d exceeds %d", who, in->nz, kMaxNz);
    if (!in->kbot || !in->dxt || !in->dxu || !in->dyt || !in->dyu || !in->cost || !in->cosu || !in->dzt || !in->dzw)
        return fail(VX_ERR_INVALID_ARGUMENT, "%s: null grid array", who);
    if (in->time_layout != VX_LEVELS_INNER && in->time_layout != VX_LEVELS_OUTER)
        return fail(VX_ERR_INVALID_ARGUMENT, "%s: time_layout must be VX_LEVELS_INNER or VX_LEVELS_OUTER", who);
    g.kbot = in->kbot;
    g.dxt = in->dxt, g.dxu = in->dxu, g.dyt = in->dyt, g.dyu = in->dyu;
    g.cost = in->cost, g.cosu = in->cosu, g.dzt = in->dzt, g.dzw = in->dzw;
    g.X = (int32_t)(in->nx + 4);
    g.Y = (int32_t)(in->ny + 4);
    g.nz = in->nz;
    const bool outer = in->time_layout == VX_LEVELS_OUTER;
    g.cs = outer ? 1 : 3;
    g.ls = outer ? (int64_t)g.X * g.Y * g.nz : 1;
    return VX_OK;
}

int make_terms(const vx_terms* in, Terms& t, bool allow_empty, const char* who) {
    t.n = 0;
    if (!in) {
        if (allow_empty) return VX_OK;
        return fail(VX_ERR_INVALID_ARGUMENT, "%s: null forcing terms", who);
    }
    if (in->n < (allow_empty ? 0 : 1) || in->n > kMaxTerms)
        return fail(VX_ERR_INVALID_ARGUMENT, "%s: forcing needs 1..%d terms (got %d)", who, kMaxTerms, in->n);
    for (int q = 0; q < in->n; ++q) {
        if (!in->arr[q]) return fail(VX_ERR_INVALID_ARGUMENT, "%s: null forcing term %d", who, q);
        t.p[q] = in->arr[q];
        t.sgn[q] = in->sign[q];
    }
    t.n = in->n;
    return VX_OK;
}

template <class Args, template <int, bool, bool> class K>
int launch_turb(const Args& a, const TurbGrid& g, bool vec, bool pv, hipStream_t s, const char* who) {
    const int S = seg_levels(g.nz);
    const int64_t threads = (int64_t)g.X * g.Y * kSegP;
    const dim3 grid(blocks_mult8(threads, kTurbBlock)), block(kTurbBlock);
#define VX_TURB_CASE(N)                                                     \
    case N:                                                                 \
        if (vec)                                                            \
            hipLaunchKernelGGL((K<N, true, true>::fn), grid, block, 0, s, a); \
        else if (pv)                                                        \
            hipLaunchKernelGGL((K<N, false, true>::fn), grid, block, 0, s, a); \
        else                                                                \
            hipLaunchKernelGGL((K<N, false, false>::fn), grid, block, 0, s, a); \
        break;
    switch (S) {
        VX_TURB_CASE(2)
        VX_TURB_CASE(4)
        VX_TURB_CASE(6)
        VX_TURB_CASE(8)
        VX_TURB_CASE(10)
        VX_TURB_CASE(12)
        VX_TURB_CASE(14)
        VX_TURB_CASE(16)
        default: return fail(VX_ERR_UNSUPPORTED, "%s: nz=%d", who, g.nz);
    }
#undef VX_TURB_CASE
    return check_launch(who);
}

template <int N, bool V, bool P>
struct TkeK {
    static constexpr auto fn = tke_kernel<N, V, P>;
};
template <int N, bool V, bool P>
struct EkeK {
    static constexpr auto fn = eke_kernel<N, V, P>;
};
template <int N, bool V, bool P>
struct IdemixK {
    static constexpr auto fn = idemix_kernel<N, V, P>;
};

// the 16-byte pair variant: levels outermost, nz even, every array aligned
// level pairs for the arrays listed (nz even, every one 16-byte aligned)
bool turb_pairs(const TurbGrid& g, std::initializer_list<const void*> arrays, const Terms& t1, const Terms* t2) {
    if (g.nz & 1) return false;
    for (const void* p : arrays)
        if (!aligned16(p)) return false;
    for (int q = 0; q < t1.n; ++q)
        if (!aligned16(t1.p[q])) return false;
    if (t2)
        for (int q = 0; q < t2->n; ++q)
            if (!aligned16(t2->p[q])) return false;
    return true;
}

// the 16-byte pair variant for every array, the energy's time levels too
// (they must be stored outermost)
bool turb_vec(const TurbGrid& g, std::initializer_list<const void*> arrays, const Terms& t1, const Terms* t2) {
    return g.cs == 1 && turb_pairs(g, arrays, t1, t2);
}

bool levels_ok(int tau, int taup1) { return tau >= 0 && tau <= 2 && taup1 >= 0 && taup1 <= 2 && tau != taup1; }

}  // namespace
}  // namespace vx

extern "C" int vx_integrate_tke_f64(void* stream, const vx_grid* grid, const vx_tke_params* prm, double* tke,
                                    int32_t tau, int32_t taup1, const double* kappaM, const double* sqrttke,
                                    const double* mxl, const double* forc_tke_surface, const vx_terms* forc,
                                    double* tke_surf_corr, double* tke_diss) {
    using namespace vx;
    const char* who = "vx_integrate_tke_f64";
    TkeArgs a{};
    int rc = make_turb_grid(grid, a.g, who);
    if (rc) return rc;
    if (!prm) return fail(VX_ERR_INVALID_ARGUMENT, "%s: null parameters", who);
    if (!levels_ok(tau, taup1)) return fail(VX_ERR_INVALID_ARGUMENT, "%s: tau, taup1 must be distinct in 0..2", who);
    if (!tke || !kappaM || !sqrttke || !mxl || !forc_tke_surface || !tke_surf_corr || !tke_diss)
        return fail(VX_ERR_INVALID_ARGUMENT, "%s: null array pointer", who);
    if ((rc = make_terms(forc, a.forc, false, who))) return rc;
    if (!(prm->dt_mom > 0.0)) return fail(VX_ERR_INVALID_ARGUMENT, "%s: dt_mom must be positive", who);
    a.tke = tke, a.tau = tau, a.taup1 = taup1;
    a.kappaM = kappaM, a.sqrttke = sqrttke, a.mxl = mxl, a.forc_surf = forc_tke_surface;
    a.surf_corr = tke_surf_corr, a.diss = tke_diss;
    a.dt = prm->dt_mom, a.alpha = prm->alpha_tke, a.c_eps = prm->c_eps, a.K_h = prm->K_h_tke;
    a.hor = prm->hor_diffusion != 0;
    const bool vec = turb_vec(a.g, {tke, tke + tau * a.g.ls, tke + taup1 * a.g.ls, kappaM, sqrttke, mxl, tke_diss},
                              a.forc, nullptr);
    // the arrays without time levels in pairs whatever the layout
    const bool pv = vec || turb_pairs(a.g, {kappaM, sqrttke, mxl, tke_diss}, a.forc, nullptr);
    return launch_turb<TkeArgs, TkeK>(a, a.g, vec, pv, (hipStream_t)stream, who);
}

extern "C" int vx_integrate_eke_f64(void* stream, const vx_grid* grid, const vx_eke_params* prm, double* eke,
                                    int32_t tau, int32_t taup1, const double* kappaM, const double* sqrteke,
                                    const double* eke_len, const double* K_gm, const vx_terms* forc,
                                    double* eke_diss_iw, double* eke_diss_tke) {
    using namespace vx;
    const char* who = "vx_integrate_eke_f64";
    EkeArgs a{};
    int rc = make_turb_grid(grid, a.g, who);
    if (rc) return rc;
    if (!prm) return fail(VX_ERR_INVALID_ARGUMENT, "%s: null parameters", who);
    if (!levels_ok(tau, taup1)) return fail(VX_ERR_INVALID_ARGUMENT, "%s: tau, taup1 must be distinct in 0..2", who);
    if (!eke || !kappaM || !sqrteke || !eke_len || !K_gm || !eke_diss_iw || !eke_diss_tke)
        return fail(VX_ERR_INVALID_ARGUMENT, "%s: null array pointer", who);
    if ((rc = make_terms(forc, a.forc, false, who))) return rc;
    if (!(prm->dt_tracer > 0.0)) return fail(VX_ERR_INVALID_ARGUMENT, "%s: dt_tracer must be positive", who);
    a.eke = eke, a.tau = tau, a.taup1 = taup1;
    a.kappaM = kappaM, a.sqrteke = sqrteke, a.eke_len = eke_len, a.K_gm = K_gm;
    a.diss_iw = eke_diss_iw, a.diss_tke = eke_diss_tke;
    a.dt = prm->dt_tracer, a.alpha = prm->alpha_eke, a.c_eps = prm->eke_c_eps;
    const bool vec = turb_vec(a.g, {eke, eke + tau * a.g.ls, eke + taup1 * a.g.ls, kappaM, sqrteke, eke_len,
                                    K_gm, eke_diss_iw, eke_diss_tke}, a.forc, nullptr);
    const bool pv = vec || turb_pairs(a.g, {kappaM, sqrteke, eke_len, eke_diss_iw, eke_diss_tke}, a.forc, nullptr);
    return launch_turb<EkeArgs, EkeK>(a, a.g, vec, pv, (hipStream_t)stream, who);
}

extern "C" int vx_integrate_idemix_f64(void* stream, const vx_grid* grid, const vx_idemix_params* prm, double* E_iw,
                                       int32_t tau, int32_t taup1, const double* c0, const double* v0,
                                       const double* alpha_c, const double* forc_iw_bottom,
                                       const double* forc_iw_surface, const vx_terms* forc,
                                       const vx_terms* forc_post, double* iw_diss) {
    using namespace vx;
    const char* who = "vx_integrate_idemix_f64";
    IdemixArgs a{};
    int rc = make_turb_grid(grid, a.g, who);
    if (rc) return rc;
    if (!prm) return fail(VX_ERR_INVALID_ARGUMENT, "%s: null parameters", who);
    if (!levels_ok(tau, taup1)) return fail(VX_ERR_INVALID_ARGUMENT, "%s: tau, taup1 must be distinct in 0..2", who);
    if (!E_iw || !c0 || !v0 || !alpha_c || !forc_iw_bottom || !forc_iw_surface || !iw_diss)
        return fail(VX_ERR_INVALID_ARGUMENT, "%s: null array pointer", who);
    if ((rc = make_terms(forc, a.forc, false, who))) return rc;
    if ((rc = make_terms(forc_post, a.forc_post, true, who))) return rc;
    if (prm->eke_diss_mode < 0 || prm->eke_diss_mode > 2)
        return fail(VX_ERR_INVALID_ARGUMENT, "%s: eke_diss_mode must be 0, 1 or 2", who);
    if (!(prm->dt_tracer > 0.0)) return fail(VX_ERR_INVALID_ARGUMENT, "%s: dt_tracer must be positive", who);
    a.E = E_iw, a.tau = tau, a.taup1 = taup1;
    a.c0 = c0, a.v0 = v0, a.alpha_c = alpha_c, a.forc_bot = forc_iw_bottom, a.forc_surf = forc_iw_surface;
    a.iw_diss = iw_diss;
    a.dt = prm->dt_tracer, a.tau_v = prm->tau_v, a.tau_h = prm->tau_h, a.frac = prm->eke_diss_surfbot_frac;
    a.hor = prm->hor_diffusion != 0;
    a.diss_mode = prm->eke_diss_mode;
    const bool vec = turb_vec(a.g, {E_iw, E_iw + tau * a.g.ls, E_iw + taup1 * a.g.ls, c0, v0, alpha_c, iw_diss},
                              a.forc, &a.forc_post);
    const bool pv = vec || turb_pairs(a.g, {c0, alpha_c, iw_diss}, a.forc, &a.forc_post);
    return launch_turb<IdemixArgs, IdemixK>(a, a.g, vec, pv, (hipStream_t)stream, who);
}
