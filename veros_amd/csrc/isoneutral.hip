// Isoneutral triad-tensor stencils for gfx950 (float64).
//
//   iso_pre_kernel       isoneutral_diffusion_pre  veros/core/isoneutral/isoneutral.py:18-229
//   iso_flux_kernel      _calc_tracer_fluxes       veros/core/isoneutral/diffusion.py:9-113
//   iso_column_tile_kernel  _calc_explicit_part, _calc_implicit_part, the tracer and
//                        dtr_iso updates and the energy branch
//                                                  veros/core/isoneutral/diffusion.py:116-283,
//                                                  veros/core/diffusion.py:9-62
//
// Layout: veros state arrays with 2-cell halos, z contiguous, k = 0 deepest.
// Cell-parallel kernels put consecutive lanes on consecutive levels of one
// column, so every stencil read of a neighbouring column is itself a
// contiguous run. Masks are derived from kbot exactly as calc_topo_kernel
// builds them (veros/core/numerics.py:195-221): maskT = kbot>0 & k>=kbot-1,
// maskU/V = min of the two adjacent maskT, maskW = min(maskT[k], maskT[k+1]).
#include "vx_common.hpp"

namespace vx {

constexpr double kEpsln = 1e-20;  // veros/core/isoneutral/isoneutral.py:28

struct Grid {
    const int32_t* kbot;
    const double *dxt, *dxu, *dyt, *dyu, *cost, *cosu, *dzt, *dzw;
    int32_t X, Y, nz;
};

struct Prm {
    double drhodT, drhodS, K_iso_steep, iso_slopec, iso_dslope, dt_tracer, grav, rho_0;
};

__device__ __forceinline__ int64_t cidx(const Grid& g, int i, int j) { return (int64_t)i * g.Y + j; }
__device__ __forceinline__ int64_t idx3(const Grid& g, int i, int j, int k) { return cidx(g, i, j) * g.nz + k; }
__device__ __forceinline__ int kb_of(const Grid& g, int i, int j) { return g.kbot[cidx(g, i, j)]; }
__device__ __forceinline__ double mT(int kb, int k) { return (kb > 0 && kb - 1 <= k) ? 1.0 : 0.0; }

// Tracer value at time level `lev` of an (X, Y, nz, 3) array.
__device__ __forceinline__ double tr3(const double* __restrict__ f, const Grid& g, int i, int j, int k, int lev) {
    return f[idx3(g, i, j, k) * 3 + lev];
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
    if (x > 38.12) return 0.0;
    if (x < -38.12) return 1.0;
    return rcp_nr(1.0 + exp_reduced(ec, x));
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
// neighbouring column is a contiguous run. The 1-D grid holds ni x nbx
// blocks (rounded up to a multiple of 8) and is remapped XCD-contiguously.
// ---------------------------------------------------------------------------
struct CellIdx {
    int i, j, k;
    bool valid;
};

__device__ __forceinline__ CellIdx cell_of(int i0, int ni, int j0, int nj) {
    const int cpb = blockDim.y;
    const int nbx = (nj + cpb - 1) / cpb;
    const int L = (int)xcd_block(gridDim.x);
    CellIdx r;
    const int ii = L / nbx;
    const int bx = L - ii * nbx;
    r.i = i0 + ii;
    r.j = j0 + bx * cpb + (int)threadIdx.y;
    r.k = (int)threadIdx.x;
    r.valid = ii < ni && r.j < j0 + nj;
    return r;
}

__device__ __forceinline__ double m01(int kb, int k) { return (kb > 0 && kb - 1 <= k) ? 1.0 : 0.0; }

// ---------------------------------------------------------------------------
// isoneutral_diffusion_pre: thread per cell (i, j, k), i in [1, X-2), j in [1, Y-2).
// The 13 tracer points of the stencil are loaded once per field; all face
// gradients are formed from them in registers. Offsets are 32-bit element
// offsets (the host checks the arrays fit).
// ---------------------------------------------------------------------------
__global__ __launch_bounds__(256) void iso_pre_kernel(Grid g, Prm p, Fields f, const double* __restrict__ K,
                                                      double* __restrict__ Aez, double* __restrict__ Anz,
                                                      double* __restrict__ Abx, double* __restrict__ Aby,
                                                      double* __restrict__ K11, double* __restrict__ K22,
                                                      double* __restrict__ K33) {
    const CellIdx ci = cell_of(1, g.X - 3, 1, g.Y - 3);
    if (!ci.valid) return;
    const int i = ci.i, j = ci.j, k = ci.k;
    const int nz = g.nz;
    const uint32_t sj = (uint32_t)nz, si = (uint32_t)g.Y * (uint32_t)nz;
    const uint32_t c = (uint32_t)i * si + (uint32_t)j * sj + (uint32_t)k;
    const uint32_t ce = c + si, cw = c - si, cn = c + sj, cs = c - sj;
    const bool hp = k < nz - 1, hm = k >= 1;
    const double* __restrict__ T = f.T + f.tau;
    const double* __restrict__ S = f.S + f.tau;
    const int kb_c = g.kbot[i * g.Y + j], kb_e = g.kbot[(i + 1) * g.Y + j], kb_w = g.kbot[(i - 1) * g.Y + j];
    const int kb_n = g.kbot[i * g.Y + j + 1], kb_s = g.kbot[i * g.Y + j - 1];

    // tracer values: c(enter), e(ast), w(est), n(orth), s(outh); p = k+1, m = k-1
    auto L2 = [&](uint32_t o, bool ok, double& t, double& sv) {
        t = ok ? T[3u * o] : 0.0;
        sv = ok ? S[3u * o] : 0.0;
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

    // reciprocal metric factors
    const double rz_c = hp ? rcp_nr(g.dzw[k]) : 0.0;
    const double rz_m = hm ? rcp_nr(g.dzw[k - 1]) : 0.0;
    const double rx_c = rcp_nr(g.dxu[i] * g.cost[j]);
    const double ry_c = rcp_nr(g.dyu[j]);

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

    // --- east face (i+1/2, j, k), j in [2, Y-2)   (:97-132)
    if (j >= 2) {
        const double mU = mc * me;
        const Grad gx = G(mU, Te, Tc, Se, Sc, rx_c);
        const Grad gz_e = hp ? G(me * mep, Tep, Te, Sep, Se, rz_c) : Grad{0.0, 0.0};
        const Grad gz_em = hm ? G(mem * me, Te, Tem, Se, Sem, rz_m) : Grad{0.0, 0.0};
        const double diffloc = hm ? 0.25 * (K[c] + K[c - 1] + K[ce] + K[ce - 1]) : 0.5 * (K[c] + K[ce]);
        double sumz = 0.0;
        double out[4];
#pragma unroll
        for (int kr = 0; kr < 2; ++kr) {
#pragma unroll
            for (int ip = 0; ip < 2; ++ip) {
                out[ip * 2 + kr] = 0.0;
                if (kr == 0 && !hm) continue;  // levels k >= 1 - kr
                const Dens w = dens(p, ip ? me : mc);
                const Grad& gz = ip ? (kr ? gz_e : gz_em) : (kr ? gz_c : gz_m);
                const double sl = -dot(w, gx) * rden(dot(w, gz));
                const double tp = taper(p, ec, tds, sl);
                sumz += g.dzw[k - 1 + kr] * mU * fmax(p.K_iso_steep, diffloc * tp);
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
        K11[c] = sumz / (4.0 * g.dzt[k]);
    }

    // --- north face (i, j+1/2, k), i in [2, X-2)   (:134-168)
    if (i >= 2) {
        const double mV = mc * mn;
        const Grad gy = G(mV, Tn, Tc, Sn, Sc, ry_c);
        const Grad gz_n = hp ? G(mn * mnp, Tnp, Tn, Snp, Sn, rz_c) : Grad{0.0, 0.0};
        const Grad gz_nm = hm ? G(mnm * mn, Tn, Tnm, Sn, Snm, rz_m) : Grad{0.0, 0.0};
        const double diffloc = hm ? 0.25 * (K[c] + K[c - 1] + K[cn] + K[cn - 1]) : 0.5 * (K[c] + K[cn]);
        double sumz = 0.0;
        double out[4];
#pragma unroll
        for (int kr = 0; kr < 2; ++kr) {
#pragma unroll
            for (int jp = 0; jp < 2; ++jp) {
                out[jp * 2 + kr] = 0.0;
                if (kr == 0 && !hm) continue;
                const Dens w = dens(p, jp ? mn : mc);
                const Grad& gz = jp ? (kr ? gz_n : gz_nm) : (kr ? gz_c : gz_m);
                const double sl = -dot(w, gy) * rden(dot(w, gz));
                const double tp = taper(p, ec, tds, sl);
                sumz += g.dzw[k - 1 + kr] * mV * fmax(p.K_iso_steep, diffloc * tp);
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
        K22[c] = sumz / (4.0 * g.dzt[k]);
    }

    // --- top face (i, j, k+1/2), i, j in [2, X/Y-2), k < nz-1   (:170-225)
    if (i >= 2 && j >= 2 && hp) {
        const int kb_w2 = kb_w, kb_s2 = kb_s;
        const double mW = mc * mcp;
        const double Kc = K[c];
        const double rx_w = rcp_nr(g.dxu[i - 1] * g.cost[j]);
        const double ry_s = rcp_nr(g.dyu[j - 1]);
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
            const double rd = rden(dot(w, gz_c));
#pragma unroll
            for (int ip = 0; ip < 2; ++ip) {
                const Grad& gx = ip ? (kr ? gx_cp : gx_c) : (kr ? gx_wp : gx_w);
                const double sl = -dot(w, gx) * rd;
                const double tp = taper(p, ec, tds, sl);
                sumx += g.dxu[i - 1 + ip] * Kc * tp * (sl * sl) * mW;
                ox[ip * 2 + kr] = tp * sl * mW;
            }
#pragma unroll
            for (int jp = 0; jp < 2; ++jp) {
                const Grad& gy = jp ? (kr ? gy_cp : gy_c) : (kr ? gy_sp : gy_s);
                const double sl = -dot(w, gy) * rd;
                const double tp = taper(p, ec, tds, sl);
                const double facty = g.cosu[j - 1 + jp] * g.dyu[j - 1 + jp];
                sumy += facty * Kc * tp * (sl * sl) * mW;
                oy[jp * 2 + kr] = tp * sl * mW;
            }
        }
        *reinterpret_cast<double2*>(Abx + 4u * c) = make_double2(ox[0], ox[1]);
        *reinterpret_cast<double2*>(Abx + 4u * c + 2) = make_double2(ox[2], ox[3]);
        *reinterpret_cast<double2*>(Aby + 4u * c) = make_double2(oy[0], oy[1]);
        *reinterpret_cast<double2*>(Aby + 4u * c + 2) = make_double2(oy[2], oy[3]);
        K33[c] = sumx / (4 * g.dxt[i]) + sumy / (4 * g.dyt[j] * g.cost[j]);
    }
}

// K_33[..., nz-1] = 0 over the whole plane (veros/core/isoneutral/isoneutral.py:225)
__global__ void k33_top_zero_kernel(Grid g, double* __restrict__ K33) {
    const int64_t t = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (t >= (int64_t)g.X * g.Y) return;
    K33[t * g.nz + g.nz - 1] = 0.0;
}

// ---------------------------------------------------------------------------
// tracer fluxes: thread per cell, i in [1, X-2), j in [1, Y-2)
// fe is written for j in [2, Y-2), fn for i in [2, X-2), ft for i, j >= 2
// (ft = 0 at the top level). Other workspace entries are never read.
// ---------------------------------------------------------------------------
struct FluxArgs {
    const double* __restrict__ tr;
    int tau;
    // K1 = K_iso - K_skew (east/north faces), K2 = K_iso + K_skew (top faces),
    // either term absent when its pointer is null (veros/core/isoneutral/diffusion.py:15-16,180-188)
    const double* __restrict__ Kiso;
    const double* __restrict__ Kskew;
    const double *Aez, *Anz, *Abx, *Aby, *K11, *K22;
    double *fe, *fn, *ft;
};

__device__ __forceinline__ double k1_at(const FluxArgs& a, uint32_t o) {
    const double ki = a.Kiso ? a.Kiso[o] : 0.0;
    return a.Kskew ? ki - a.Kskew[o] : ki;
}
__device__ __forceinline__ double k2_at(const FluxArgs& a, uint32_t o) {
    const double ki = a.Kiso ? a.Kiso[o] : 0.0;
    return a.Kskew ? ki + a.Kskew[o] : ki;
}

__global__ __launch_bounds__(256) void iso_flux_kernel(Grid g, FluxArgs a) {
    const CellIdx ci = cell_of(1, g.X - 3, 1, g.Y - 3);
    if (!ci.valid) return;
    const int i = ci.i, j = ci.j, k = ci.k;
    const int nz = g.nz;
    const uint32_t sj = (uint32_t)nz, si = (uint32_t)g.Y * (uint32_t)nz;
    const uint32_t c = (uint32_t)i * si + (uint32_t)j * sj + (uint32_t)k;
    const uint32_t ce = c + si, cn = c + sj;
    const double* __restrict__ tr = a.tr + a.tau;
    const bool hm = k >= 1, hp = k + 1 < nz;
    // pad_z_edges clamping: tr[min(k+1, nz-1)], tr[max(k-1, 0)]
    const uint32_t up = hp ? 1u : 0u, dn = hm ? 1u : 0u;
    const double trc = tr[3u * c];

    if (j >= 2) {  // east face (:22-47)
        const double diffloc = hm ? 0.25 * (k1_at(a, c) + k1_at(a, c - 1) + k1_at(a, ce) + k1_at(a, ce - 1))
                                  : 0.5 * (k1_at(a, c) + k1_at(a, ce));
        const double2 A01 = *reinterpret_cast<const double2*>(a.Aez + 4u * c);      // (ip=0, kr=0), (ip=0, kr=1)
        const double2 A23 = *reinterpret_cast<const double2*>(a.Aez + 4u * c + 2);  // (ip=1, kr=0), (ip=1, kr=1)
        const double t0 = trc, t0u = tr[3u * (c + up)], t0d = tr[3u * (c - dn)];
        const double t1 = tr[3u * ce], t1u = tr[3u * (ce + up)], t1d = tr[3u * (ce - dn)];
        double sumz = 0.0;
        sumz = sumz + diffloc * A01.x * (t0 - t0d);   // kr=0, ip=0
        sumz = sumz + diffloc * A23.x * (t1 - t1d);   // kr=0, ip=1
        sumz = sumz + diffloc * A01.y * (t0u - t0);   // kr=1, ip=0
        sumz = sumz + diffloc * A23.y * (t1u - t1);   // kr=1, ip=1
        a.fe[c] = sumz / (4.0 * g.dzt[k]) + (t1 - t0) / (g.cost[j] * g.dxu[i]) * a.K11[c];
    }
    if (i >= 2) {  // north face (:49-77)
        const double diffloc = hm ? 0.25 * (k1_at(a, c) + k1_at(a, c - 1) + k1_at(a, cn) + k1_at(a, cn - 1))
                                  : 0.5 * (k1_at(a, c) + k1_at(a, cn));
        const double2 A01 = *reinterpret_cast<const double2*>(a.Anz + 4u * c);
        const double2 A23 = *reinterpret_cast<const double2*>(a.Anz + 4u * c + 2);
        const double t0 = trc, t0u = tr[3u * (c + up)], t0d = tr[3u * (c - dn)];
        const double t1 = tr[3u * cn], t1u = tr[3u * (cn + up)], t1d = tr[3u * (cn - dn)];
        double sumz = 0.0;
        sumz = sumz + diffloc * A01.x * (t0 - t0d);
        sumz = sumz + diffloc * A23.x * (t1 - t1d);
        sumz = sumz + diffloc * A01.y * (t0u - t0);
        sumz = sumz + diffloc * A23.y * (t1u - t1);
        a.fn[c] = g.cosu[j] * (sumz / (4.0 * g.dzt[k]) + (t1 - t0) / g.dyu[j] * a.K22[c]);
    }
    if (i >= 2 && j >= 2) {  // top face (:79-111)
        if (!hp) {
            a.ft[c] = 0.0;
        } else {
            const double diffloc = k2_at(a, c);
            const uint32_t cw = c - si, cs = c - sj;
            const double2 B01 = *reinterpret_cast<const double2*>(a.Abx + 4u * c);      // ip=0: kr=0,1
            const double2 B23 = *reinterpret_cast<const double2*>(a.Abx + 4u * c + 2);  // ip=1: kr=0,1
            const double2 C01 = *reinterpret_cast<const double2*>(a.Aby + 4u * c);
            const double2 C23 = *reinterpret_cast<const double2*>(a.Aby + 4u * c + 2);
            const double tc0 = trc, tc1 = tr[3u * (c + 1)];
            const double tw0 = tr[3u * cw], tw1 = tr[3u * (cw + 1)];
            const double te0 = tr[3u * ce], te1 = tr[3u * (ce + 1)];
            const double ts0 = tr[3u * cs], ts1 = tr[3u * (cs + 1)];
            const double tn0 = tr[3u * cn], tn1 = tr[3u * (cn + 1)];
            double sumx = 0.0;
            sumx = sumx + diffloc * B01.x / g.cost[j] * (tc0 - tw0);   // ip=0, kr=0
            sumx = sumx + diffloc * B01.y / g.cost[j] * (tc1 - tw1);   // ip=0, kr=1
            sumx = sumx + diffloc * B23.x / g.cost[j] * (te0 - tc0);   // ip=1, kr=0
            sumx = sumx + diffloc * B23.y / g.cost[j] * (te1 - tc1);   // ip=1, kr=1
            double sumy = 0.0;
            sumy = sumy + diffloc * C01.x * g.cosu[j - 1] * (tc0 - ts0);
            sumy = sumy + diffloc * C01.y * g.cosu[j - 1] * (tc1 - ts1);
            sumy = sumy + diffloc * C23.x * g.cosu[j] * (tn0 - tc0);
            sumy = sumy + diffloc * C23.y * g.cosu[j] * (tn1 - tc1);
            a.ft[c] = sumx / (4 * g.dxt[i]) + sumy / (4 * g.dyt[j] * g.cost[j]);
        }
    }
}

// ---------------------------------------------------------------------------
// Column-tile kernel: everything of the tracer step after the fluxes.
// Block = nz levels x TC columns (consecutive j of one i row), covering
// (i, j) in [1, X-1) x [1, Y-1).
//   phase 1 (thread per cell): lateral dissipation on T points
//     (veros/core/diffusion.py:9-36), flux divergence dtr
//     (veros/core/isoneutral/diffusion.py:116-139), rhs = tr[taup1] + dt*dtr
//     (:196-197) and delta = dt/dzw*K_33 (:154-157) -> LDS
//   phase 2 (lane per interior column): Thomas on the wet suffix, in LDS
//     (:158-169); the rhs slots become the new tracer (dry slots keep rhs)
//   phase 3 (thread per cell): tr[taup1] = new, dtr_iso += dtr + (new-rhs)/dt
//     (:196,204), P += dissipation_on_wgrid(...) (veros/core/diffusion.py:39-62)
//     + the vertical-mixing term (veros/core/isoneutral/diffusion.py:254-279)
// ---------------------------------------------------------------------------
struct ColArgs {
    double* tr;
    int tau, taup1;
    const double* K33;
    double* dtr;
    const double* I;   // int_drhodX (X, Y, nz, 3), nullable
    double* P;         // P_diss, nullable
    const double *fe, *fn, *ft;
};

// Face fluxes with their valid index ranges (0 elsewhere).
__device__ __forceinline__ double fe_at(const Grid& g, const ColArgs& a, int i, int j, uint32_t o) {
    return (i >= 1 && i <= g.X - 3 && j >= 2 && j <= g.Y - 3) ? a.fe[o] : 0.0;
}
__device__ __forceinline__ double fn_at(const Grid& g, const ColArgs& a, int i, int j, uint32_t o) {
    return (i >= 2 && i <= g.X - 3 && j >= 1 && j <= g.Y - 3) ? a.fn[o] : 0.0;
}

template <int TC>
__global__ __launch_bounds__(512) void iso_column_tile_kernel(Grid g, Prm p, ColArgs a, int implicit) {
    extern __shared__ double lds[];
    const int nz = g.nz;
    double* sD = lds;                 // [TC][nz] lateral dissipation on T points
    double* sR = lds + TC * nz;       // [TC][nz] rhs -> forward sweep -> new tracer
    double* sA = lds + 2 * TC * nz;   // [TC][nz] sub-diagonal
    double* sB = lds + 3 * TC * nz;   // [TC][nz] diagonal
    double* sC = lds + 4 * TC * nz;   // [TC][nz] super-diagonal -> cp

    const CellIdx ci = cell_of(1, g.X - 2, 1, g.Y - 2);
    const int i = ci.i, j = ci.j, k = ci.k;
    const int tc = (int)threadIdx.y;
    const bool valid = ci.valid;
    const bool interior = valid && i >= 2 && i <= g.X - 3 && j >= 2 && j <= g.Y - 3;
    const uint32_t sj = (uint32_t)nz, si = (uint32_t)g.Y * (uint32_t)nz;
    const uint32_t c = (uint32_t)i * si + (uint32_t)j * sj + (uint32_t)k;
    const int kb = valid ? g.kbot[i * g.Y + j] : 0;
    const bool energy = a.P != nullptr;
    const double dt = p.dt_tracer;

    // ---- phase 1
    double dtrk = 0.0, rhs = 0.0, ftk = 0.0, K33k = 0.0, Ic = 0.0;
    if (valid && energy) {
        const double* __restrict__ I = a.I + a.tau;
        Ic = I[3u * c];
        const double de = (I[3u * (c + si)] - Ic) * fe_at(g, a, i, j, c) +
                          (Ic - I[3u * (c - si)]) * fe_at(g, a, i - 1, j, c - si);
        const double dn = (I[3u * (c + sj)] - Ic) * fn_at(g, a, i, j, c) +
                          (Ic - I[3u * (c - sj)]) * fn_at(g, a, i, j - 1, c - sj);
        const double g05 = 0.5 * p.grav / p.rho_0;
        sD[tc * nz + k] = g05 * de / (g.dxt[i] * g.cost[j]) + g05 * dn / (g.dyt[j] * g.cost[j]);
    }
    if (interior) {
        const double mt = m01(kb, k);
        ftk = a.ft[c];
        const double hor = (a.fe[c] - a.fe[c - si]) / (g.cost[j] * g.dxt[i]) +
                           (a.fn[c] - a.fn[c - sj]) / (g.cost[j] * g.dyt[j]);
        const double vert = (k == 0) ? ftk / g.dzt[0] : (ftk - a.ft[c - 1]) / g.dzt[k];
        dtrk = mt * hor + mt * vert;
        rhs = a.tr[3u * c + a.taup1] + dt * dtrk;
        sR[tc * nz + k] = rhs;
        if (implicit) {
            // rows of the K33 system (veros/core/isoneutral/diffusion.py:154-164); the
            // edge row k = kbot-1 takes b_edge and a = 0 (veros/core/operators.py:72)
            K33k = a.K33[c];
            const double delta = (k < nz - 1) ? dt / g.dzw[k] * K33k : 0.0;
            const double delta_m = (k >= 1) ? dt / g.dzw[k - 1] * a.K33[c - 1] : 0.0;
            const double dz = g.dzt[k];
            const bool edge = k == kb - 1;
            sA[tc * nz + k] = edge ? 0.0 : -delta_m / dz;
            sB[tc * nz + k] = edge ? 1 + delta / dz : 1 + (delta + delta_m) / dz;
            sC[tc * nz + k] = (k < nz - 1) ? -delta / dz : 0.0;
        }
    }
    __syncthreads();

    // ---- phase 2: threads 0..TC-1 (first wave) sweep one column each
    if (implicit) {
        const int lin = (int)(threadIdx.y * blockDim.x + threadIdx.x);
        if (lin < TC) {
            const int cj = j - tc + lin;  // j - tc is the tile's first column
            const bool col_ok = i >= 2 && i <= g.X - 3 && cj >= 2 && cj <= g.Y - 3;
            const int kbl = col_ok ? g.kbot[i * g.Y + cj] : 0;
            if (col_ok && kbl >= 1 && kbl <= nz) {
                const int k0 = kbl - 1;
                double* R = sR + lin * nz;
                const double* Ar = sA + lin * nz;
                const double* Br = sB + lin * nz;
                double* Cc = sC + lin * nz;
                double cpp = 0.0, dpp = 0.0;
#pragma unroll 8
                for (int kk = k0; kk < nz; ++kk) {
                    const double ak = Ar[kk];
                    const double r = rcp_nr(Br[kk] - ak * cpp);
                    cpp = Cc[kk] * r;
                    dpp = (R[kk] - ak * dpp) * r;
                    Cc[kk] = cpp;
                    R[kk] = dpp;
                }
                double xn = 0.0;
#pragma unroll 8
                for (int kk = nz - 1; kk >= k0; --kk) {
                    xn = R[kk] - Cc[kk] * xn;
                    R[kk] = xn;
                }
            }
        }
        __syncthreads();
    }

    // ---- phase 3
    if (!valid) return;
    if (interior) {
        const double newk = sR[tc * nz + k];
        a.tr[3u * c + a.taup1] = newk;
        a.dtr[c] = (a.dtr[c] + dtrk) + (newk - rhs) / dt;
    }
    if (energy) {
        const double d0 = sD[tc * nz + k];
        const bool land = kb >= 1;  // create_water_masks(kbot): ks = kbot - 1 >= 0
        double dw;
        if (k == nz - 1) {
            dw = d0 * (land ? 1.0 : 0.0);
        } else {
            const double d1 = sD[tc * nz + k + 1];
            const double avg = 0.5 * (d0 + d1);
            const bool edge = land && k == kb - 1;
            const bool water = land && k > kb - 1;
            const double dzw_lo = g.dzw[k >= 1 ? k - 1 : 0];  // pad_z_edges(dzw)[k]
            dw = (avg + 0.5 * (d0 * dzw_lo / g.dzw[k])) * (edge ? 1.0 : 0.0) + avg * (water ? 1.0 : 0.0);
        }
        double Pk = a.P[c] + dw;
        if (interior && k < nz - 1) {
            const double mw = m01(kb, k) * m01(kb, k + 1);
            const double fxa = (-a.I[3u * (c + 1) + a.tau] + Ic) / g.dzw[k];
            const double gr = -p.grav / p.rho_0;
            if (implicit) {
                const double dtrz = sR[tc * nz + k + 1] - sR[tc * nz + k];
                Pk = Pk + gr * fxa * (ftk * mw + K33k * dtrz / g.dzw[k] * mw);
            } else {
                Pk = Pk + gr * fxa * ftk * mw;
            }
        }
        a.P[c] = Pk;
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

unsigned blocks_for(int64_t n, int bs) { return (unsigned)((n + bs - 1) / bs); }
unsigned blocks_xcd(int64_t n, int bs) { return (unsigned)(((n + bs - 1) / bs + 7) / 8 * 8); }

// block of nz levels x cpb columns (about 256 threads); 1-D grid of ni row
// groups, rounded up to a multiple of 8 for the XCD remap
dim3 cell_block(int nz) { return dim3(nz, nz >= 256 ? 1 : 256 / nz); }
dim3 cell_grid(int ni, int nj, dim3 blk) {
    const int64_t nbx = (nj + blk.y - 1) / blk.y;
    return dim3((unsigned)(((int64_t)ni * nbx + 7) / 8 * 8));
}

}  // namespace
}  // namespace vx

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
    const dim3 blk = cell_block(g.nz);
    hipLaunchKernelGGL(iso_pre_kernel, cell_grid(g.X - 3, g.Y - 3, blk), blk, 0, s, g, p, f, K_iso, Ai_ez, Ai_nz,
                       Ai_bx, Ai_by, K_11, K_22, K_33);
    if ((rc = check_launch("vx_iso_pre_f64"))) return rc;
    hipLaunchKernelGGL(k33_top_zero_kernel, dim3(blocks_for((int64_t)g.X * g.Y, 256)), dim3(256), 0, s, g, K_33);
    return check_launch("vx_iso_pre_f64 (K_33 top)");
}

extern "C" int64_t vx_iso_workspace_bytes(const vx_grid* grid) {
    if (!grid) return -1;
    return 3 * (grid->nx + 4) * (grid->ny + 4) * (int64_t)grid->nz * (int64_t)sizeof(double);
}

namespace vx {
namespace {

constexpr int kTileColumns = 4;

int launch_tile(const Grid& g, const Prm& p, const ColArgs& a, int implicit, hipStream_t s) {
    constexpr int TC = kTileColumns;
    if (g.nz * TC > 512) return fail(VX_ERR_UNSUPPORTED, "iso column tile: nz=%d too deep", g.nz);
    const dim3 blk(g.nz, TC);
    const size_t lds = (size_t)5 * TC * g.nz * sizeof(double);
    hipLaunchKernelGGL(iso_column_tile_kernel<TC>, cell_grid(g.X - 2, g.Y - 2, blk), blk, lds, s, g, p, a, implicit);
    return check_launch("iso_column_tile_kernel");
}

int iso_diffusion_common(void* stream, const vx_grid* grid, const vx_iso_params* prm, double* tracer, int32_t tau,
                         int32_t taup1, const double* Kiso, const double* Kskew, const double* Ai_ez, const double* Ai_nz,
                         const double* Ai_bx, const double* Ai_by, const double* K_11, const double* K_22,
                         const double* K_33, double* dtr_iso, const double* int_drhodX, double* P_diss,
                         void* workspace, int implicit, const char* who) {
    Grid g;
    int rc = make_grid(grid, g, who);
    if (rc) return rc;
    if (!prm) return fail(VX_ERR_INVALID_ARGUMENT, "%s: null params", who);
    if (tau < 0 || tau > 2 || taup1 < 0 || taup1 > 2)
        return fail(VX_ERR_INVALID_ARGUMENT, "%s: time levels must be 0..2", who);
    if (!tracer || (!Kiso && !Kskew) || !Ai_ez || !Ai_nz || !Ai_bx || !Ai_by || !K_11 || !K_22 || !dtr_iso || !workspace ||
        (implicit && !K_33))
        return fail(VX_ERR_INVALID_ARGUMENT, "%s: null array pointer", who);
    if ((P_diss == nullptr) != (int_drhodX == nullptr))
        return fail(VX_ERR_INVALID_ARGUMENT, "%s: P_diss and int_drhodX must both be given or both be NULL", who);
    hipStream_t s = (hipStream_t)stream;
    const Prm p = make_prm(prm);
    const int64_t n3 = (int64_t)g.X * g.Y * g.nz;
    double* ws = static_cast<double*>(workspace);
    FluxArgs fa{tracer, tau, Kiso, Kskew, Ai_ez, Ai_nz, Ai_bx, Ai_by, K_11, K_22, ws, ws + n3, ws + 2 * n3};
    if (n3 * 4 >= ((int64_t)1 << 32)) return fail(VX_ERR_UNSUPPORTED, "%s: grid too large for 32-bit offsets", who);
    const dim3 blk = cell_block(g.nz);
    hipLaunchKernelGGL(iso_flux_kernel, cell_grid(g.X - 3, g.Y - 3, blk), blk, 0, s, g, fa);
    if ((rc = check_launch(who))) return rc;
    ColArgs ca{tracer, tau, taup1, K_33, dtr_iso, int_drhodX, P_diss, ws, ws + n3, ws + 2 * n3};
    if ((rc = launch_tile(g, p, ca, implicit, s))) return rc;
    return VX_OK;
}
}  // namespace
}  // namespace vx

extern "C" int vx_iso_diffusion_f64(void* stream, const vx_grid* grid, const vx_iso_params* prm, double* tracer,
                                    int32_t tau, int32_t taup1, const double* K_iso, const double* Ai_ez,
                                    const double* Ai_nz, const double* Ai_bx, const double* Ai_by,
                                    const double* K_11, const double* K_22, const double* K_33, double* dtr_iso,
                                    const double* int_drhodX, double* P_diss, void* workspace) {
    if (!K_iso) return vx::fail(VX_ERR_INVALID_ARGUMENT, "vx_iso_diffusion_f64: null K_iso");
    return vx::iso_diffusion_common(stream, grid, prm, tracer, tau, taup1, K_iso, nullptr, Ai_ez, Ai_nz, Ai_bx, Ai_by,
                                    K_11, K_22, K_33, dtr_iso, int_drhodX, P_diss, workspace, 1,
                                    "vx_iso_diffusion_f64");
}

extern "C" int vx_iso_skew_diffusion_f64(void* stream, const vx_grid* grid, const vx_iso_params* prm,
                                         double* tracer, int32_t tau, int32_t taup1, const double* K_gm,
                                         const double* Ai_ez, const double* Ai_nz, const double* Ai_bx,
                                         const double* Ai_by, const double* K_11, const double* K_22,
                                         double* dtr_iso, const double* int_drhodX, double* P_diss_skew,
                                         void* workspace) {
    if (!K_gm) return vx::fail(VX_ERR_INVALID_ARGUMENT, "vx_iso_skew_diffusion_f64: null K_gm");
    return vx::iso_diffusion_common(stream, grid, prm, tracer, tau, taup1, nullptr, K_gm, Ai_ez, Ai_nz, Ai_bx,
                                    Ai_by, K_11, K_22, nullptr, dtr_iso, int_drhodX, P_diss_skew, workspace, 0,
                                    "vx_iso_skew_diffusion_f64");
}
