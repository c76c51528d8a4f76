// Isoneutral triad-tensor stencils for gfx950 (float64).
//
//   iso_pre_kernel       isoneutral_diffusion_pre  veros/core/isoneutral/isoneutral.py:18-229
//   iso_flux_kernel      _calc_tracer_fluxes       veros/core/isoneutral/diffusion.py:9-113
//   iso_explicit_kernel  _calc_explicit_part, the explicit tracer update and the lateral
//                        dissipation               veros/core/isoneutral/diffusion.py:116-139,196-197,
//                                                  veros/core/diffusion.py:9-62
//   iso_implicit_kernel  _calc_implicit_part       veros/core/isoneutral/diffusion.py:142-169,202-204
//   iso_post_kernel      vertical-mixing dissipation veros/core/isoneutral/diffusion.py:254-279
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
    const double* __restrict__ K1;   // K_iso - K_skew (east/north)
    const double* __restrict__ K2;   // K_iso + K_skew (top)
    const double *Aez, *Anz, *Abx, *Aby, *K11, *K22;
    double *fe, *fn, *ft;
};

__global__ __launch_bounds__(256) void iso_flux_kernel(Grid g, FluxArgs a) {
    const int nz = g.nz;
    const int ny3 = g.Y - 3;
    const int64_t t = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    const int64_t col = t / nz;
    if (col >= (int64_t)(g.X - 3) * ny3) return;
    const int k = (int)(t - col * nz);
    const int i = 1 + (int)(col / ny3);
    const int j = 1 + (int)(col - (int64_t)(i - 1) * ny3);
    const int64_t c = idx3(g, i, j, k);
    const double* __restrict__ tr = a.tr;
    const int lev = a.tau;
    const int kup = (k + 1 < nz) ? k + 1 : nz - 1;  // pad_z_edges clamping
    const int kdn = (k >= 1) ? k - 1 : 0;
    const double trc = tr3(tr, g, i, j, k, lev);

    if (j >= 2) {  // east face (:22-47)
        const double* K1 = a.K1;
        const double diffloc = (k >= 1) ? 0.25 * (K1[idx3(g, i, j, k)] + K1[idx3(g, i, j, k - 1)] +
                                                  K1[idx3(g, i + 1, j, k)] + K1[idx3(g, i + 1, j, k - 1)])
                                        : 0.5 * (K1[idx3(g, i, j, 0)] + K1[idx3(g, i + 1, j, 0)]);
        double sumz = 0.0;
#pragma unroll
        for (int kr = 0; kr < 2; ++kr) {
#pragma unroll
            for (int ip = 0; ip < 2; ++ip) {
                // tr[min(k+kr, nz-1)] - tr[max(k+kr-1, 0)]
                const double up = tr3(tr, g, i + ip, j, kr ? kup : k, lev);
                const double dn = tr3(tr, g, i + ip, j, kr ? k : kdn, lev);
                sumz = sumz + diffloc * a.Aez[c * 4 + ip * 2 + kr] * (up - dn);
            }
        }
        const double tre = tr3(tr, g, i + 1, j, k, lev);
        a.fe[c] = sumz / (4.0 * g.dzt[k]) + (tre - trc) / (g.cost[j] * g.dxu[i]) * a.K11[c];
    }
    if (i >= 2) {  // north face (:49-77)
        const double* K1 = a.K1;
        const double diffloc = (k >= 1) ? 0.25 * (K1[idx3(g, i, j, k)] + K1[idx3(g, i, j, k - 1)] +
                                                  K1[idx3(g, i, j + 1, k)] + K1[idx3(g, i, j + 1, k - 1)])
                                        : 0.5 * (K1[idx3(g, i, j, 0)] + K1[idx3(g, i, j + 1, 0)]);
        double sumz = 0.0;
#pragma unroll
        for (int kr = 0; kr < 2; ++kr) {
#pragma unroll
            for (int jp = 0; jp < 2; ++jp) {
                const double up = tr3(tr, g, i, j + jp, kr ? kup : k, lev);
                const double dn = tr3(tr, g, i, j + jp, kr ? k : kdn, lev);
                sumz = sumz + diffloc * a.Anz[c * 4 + jp * 2 + kr] * (up - dn);
            }
        }
        const double trn = tr3(tr, g, i, j + 1, k, lev);
        a.fn[c] = g.cosu[j] * (sumz / (4.0 * g.dzt[k]) + (trn - trc) / g.dyu[j] * a.K22[c]);
    }
    if (i >= 2 && j >= 2) {  // top face (:79-111)
        if (k == nz - 1) {
            a.ft[c] = 0.0;
        } else {
            const double diffloc = a.K2[c];
            double sumx = 0.0;
#pragma unroll
            for (int ip = 0; ip < 2; ++ip) {
#pragma unroll
                for (int kr = 0; kr < 2; ++kr) {
                    const double d = tr3(tr, g, i + ip, j, k + kr, lev) - tr3(tr, g, i - 1 + ip, j, k + kr, lev);
                    sumx = sumx + diffloc * a.Abx[c * 4 + ip * 2 + kr] / g.cost[j] * d;
                }
            }
            double sumy = 0.0;
#pragma unroll
            for (int jp = 0; jp < 2; ++jp) {
#pragma unroll
                for (int kr = 0; kr < 2; ++kr) {
                    const double d = tr3(tr, g, i, j + jp, k + kr, lev) - tr3(tr, g, i, j - 1 + jp, k + kr, lev);
                    sumy = sumy + diffloc * a.Aby[c * 4 + jp * 2 + kr] * g.cosu[j - 1 + jp] * d;
                }
            }
            a.ft[c] = sumx / (4 * g.dxt[i]) + sumy / (4 * g.dyt[j] * g.cost[j]);
        }
    }
}

// ---------------------------------------------------------------------------
// explicit part: thread per cell (i, j, k), (i, j) in [1, X-1) x [1, Y-1)
//  - all such cells (energy on): P_diss += dissipation_on_wgrid(compute_dissipation(...))
//    (veros/core/diffusion.py:9-62; diss is evaluated at k and k+1 here)
//  - interior cells: dtr = flux divergence (veros/core/isoneutral/diffusion.py:116-139),
//    dtr_iso += dtr, tr[taup1] += dt * dtr (:196-197)
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
__device__ __forceinline__ double fe_at(const Grid& g, const ColArgs& a, int i, int j, int k) {
    return (i >= 1 && i <= g.X - 3 && j >= 2 && j <= g.Y - 3) ? a.fe[idx3(g, i, j, k)] : 0.0;
}
__device__ __forceinline__ double fn_at(const Grid& g, const ColArgs& a, int i, int j, int k) {
    return (i >= 2 && i <= g.X - 3 && j >= 1 && j <= g.Y - 3) ? a.fn[idx3(g, i, j, k)] : 0.0;
}

// compute_dissipation at one T point (veros/core/diffusion.py:13-33)
__device__ __forceinline__ double diss_at(const Grid& g, const ColArgs& a, double g05, int i, int j, int k) {
    const double Ic = tr3(a.I, g, i, j, k, a.tau);
    const double de = (tr3(a.I, g, i + 1, j, k, a.tau) - Ic) * fe_at(g, a, i, j, k) +
                      (Ic - tr3(a.I, g, i - 1, j, k, a.tau)) * fe_at(g, a, i - 1, j, k);
    const double dn = (tr3(a.I, g, i, j + 1, k, a.tau) - Ic) * fn_at(g, a, i, j, k) +
                      (Ic - tr3(a.I, g, i, j - 1, k, a.tau)) * fn_at(g, a, i, j - 1, k);
    return g05 * de / (g.dxt[i] * g.cost[j]) + g05 * dn / (g.dyt[j] * g.cost[j]);
}

__global__ __launch_bounds__(256) void iso_explicit_kernel(Grid g, Prm p, ColArgs a) {
    const int nz = g.nz;
    const int ny2 = g.Y - 2;
    const int64_t t = xcd_block(gridDim.x) * blockDim.x + threadIdx.x;
    const int64_t col = t / nz;
    if (col >= (int64_t)(g.X - 2) * ny2) return;
    const int k = (int)(t - col * nz);
    const int i = 1 + (int)(col / ny2);
    const int j = 1 + (int)(col - (int64_t)(i - 1) * ny2);
    const int64_t c = idx3(g, i, j, k);
    const int kb = kb_of(g, i, j);

    if (a.P) {
        const double g05 = 0.5 * p.grav / p.rho_0;
        const double d0 = diss_at(g, a, g05, i, j, k);
        const bool land = kb >= 1;  // create_water_masks(kbot): ks = kbot - 1 >= 0
        double dw;
        if (k == nz - 1) {
            dw = d0 * (land ? 1.0 : 0.0);
        } else {
            const double d1 = diss_at(g, a, g05, i, j, k + 1);
            const double avg = 0.5 * (d0 + d1);
            const bool edge = land && k == kb - 1;
            const bool water = land && k > kb - 1;
            const double dzw_lo = g.dzw[k >= 1 ? k - 1 : 0];  // pad_z_edges(dzw)[k]
            dw = (avg + 0.5 * (d0 * dzw_lo / g.dzw[k])) * (edge ? 1.0 : 0.0) + avg * (water ? 1.0 : 0.0);
        }
        a.P[c] = a.P[c] + dw;
    }
    if (i < 2 || i > g.X - 3 || j < 2 || j > g.Y - 3) return;

    const double mt = mT(kb, k);
    const double ftk = a.ft[c];
    const double hor = (a.fe[c] - fe_at(g, a, i - 1, j, k)) / (g.cost[j] * g.dxt[i]) +
                       (a.fn[c] - fn_at(g, a, i, j - 1, k)) / (g.cost[j] * g.dyt[j]);
    const double vert = (k == 0) ? ftk / g.dzt[0] : (ftk - a.ft[c - 1]) / g.dzt[k];
    const double dtrk = mt * hor + mt * vert;
    a.dtr[c] = a.dtr[c] + dtrk;
    double* trp = a.tr + c * 3 + a.taup1;
    *trp = *trp + p.dt_tracer * dtrk;
}

// ---------------------------------------------------------------------------
// implicit K_33 solve (veros/core/isoneutral/diffusion.py:142-169,202-204):
// lane per interior column, Thomas in registers over the wet suffix;
// tr[taup1] <- solution, dtr_iso += (solution - rhs) / dt on wet cells
// (dry cells keep rhs, so their increment is 0).
// ---------------------------------------------------------------------------
template <int NZB, bool VEC>
__global__ __launch_bounds__(256) void iso_implicit_kernel(Grid g, Prm p, ColArgs a) {
    const int nz = g.nz;
    const int ny4 = g.Y - 4;
    const int64_t col = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (col >= (int64_t)(g.X - 4) * ny4) return;
    const int i = 2 + (int)(col / ny4);
    const int j = 2 + (int)(col - (int64_t)(i - 2) * ny4);
    const int kb = kb_of(g, i, j);
    if (kb < 1 || kb > nz) return;
    const int k0 = kb - 1;
    const int64_t base = idx3(g, i, j, 0);
    const double* __restrict__ K33 = a.K33 + base;
    double* __restrict__ trc = a.tr + base * 3 + a.taup1;
    double* __restrict__ dtr = a.dtr + base;
    const double dt = p.dt_tracer;

    double cp[NZB], dp[NZB];
    double cpp = 0.0, dpp = 0.0, delta_m = 0.0;
#pragma unroll
    for (int q = 0; q < NZB; q += 8) {
        if (q < nz && q + 8 > k0) {
            double kv[8] = {}, rv[8] = {};
            if constexpr (VEC) {
#pragma unroll
                for (int t = 0; t < 8; t += 2) {
                    if (q + t < nz) {
                        const double2 w = *reinterpret_cast<const double2*>(K33 + q + t);
                        kv[t] = w.x;
                        kv[t + 1] = w.y;
                    }
                }
            } else {
#pragma unroll
                for (int t = 0; t < 8; ++t)
                    if (q + t < nz) kv[t] = K33[q + t];
            }
#pragma unroll
            for (int t = 0; t < 8; ++t)
                if (q + t < nz && q + t >= k0) rv[t] = trc[(q + t) * 3];
#pragma unroll
            for (int t = 0; t < 8; ++t) {
                const int k = q + t;
                if (k < nz && k >= k0) {
                    const double delta = (k < nz - 1) ? dt / g.dzw[k] * kv[t] : 0.0;
                    const double dz = g.dzt[k];
                    const bool edge = k == k0;
                    const double ak = edge ? 0.0 : -delta_m / dz;
                    const double bk = edge ? 1 + delta / dz : 1 + (delta + delta_m) / dz;
                    const double ck = (k < nz - 1) ? -delta / dz : 0.0;
                    const double r = 1.0 / (bk - ak * cpp);
                    cpp = ck * r;
                    dpp = (rv[t] - ak * dpp) * r;
                    delta_m = delta;
                }
                cp[k] = cpp;
                dp[k] = dpp;
            }
        }
    }
    double xn = 0.0;
#pragma unroll
    for (int k = NZB - 1; k >= 0; --k) {
        if (k < nz && k >= k0) {
            const double xk = dp[k] - cp[k] * xn;
            const double rhs = trc[k * 3];
            trc[k * 3] = xk;
            dtr[k] = dtr[k] + (xk - rhs) / dt;
            xn = xk;
        }
    }
}

// ---------------------------------------------------------------------------
// vertical-mixing dissipation (veros/core/isoneutral/diffusion.py:254-279):
// thread per interior cell, k < nz-1, after the implicit solve.
// ---------------------------------------------------------------------------
__global__ __launch_bounds__(256) void iso_post_kernel(Grid g, Prm p, ColArgs a, int implicit) {
    const int nz1 = g.nz - 1;
    const int ny4 = g.Y - 4;
    const int64_t t = xcd_block(gridDim.x) * blockDim.x + threadIdx.x;
    const int64_t col = t / nz1;
    if (col >= (int64_t)(g.X - 4) * ny4) return;
    const int k = (int)(t - col * nz1);
    const int i = 2 + (int)(col / ny4);
    const int j = 2 + (int)(col - (int64_t)(i - 2) * ny4);
    const int64_t c = idx3(g, i, j, k);
    const int kb = kb_of(g, i, j);
    const double mw = mT(kb, k) * mT(kb, k + 1);
    const double fxa = (-tr3(a.I, g, i, j, k + 1, a.tau) + tr3(a.I, g, i, j, k, a.tau)) / g.dzw[k];
    const double gr = -p.grav / p.rho_0;
    if (implicit) {
        const double dtrz = a.tr[(c + 1) * 3 + a.taup1] - a.tr[c * 3 + a.taup1];
        a.P[c] = a.P[c] + gr * fxa * (a.ft[c] * mw + a.K33[c] * dtrz / g.dzw[k] * mw);
    } else {
        a.P[c] = a.P[c] + gr * fxa * a.ft[c] * mw;
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

int launch_implicit(const Grid& g, const Prm& p, const ColArgs& a, hipStream_t s) {
    const int64_t ncol = (int64_t)(g.X - 4) * (g.Y - 4);
    const dim3 grid(blocks_for(ncol, 256)), block(256);
    const int nzb = ((g.nz + 15) / 16) * 16;
    const bool vec = (g.nz % 2 == 0) && aligned16(a.K33);
#define VX_IMPL_CASE(N)                                                                   \
    case N:                                                                               \
        if (vec) hipLaunchKernelGGL((iso_implicit_kernel<N, true>), grid, block, 0, s, g, p, a); \
        else hipLaunchKernelGGL((iso_implicit_kernel<N, false>), grid, block, 0, s, g, p, a);    \
        break;
    switch (nzb) {
        VX_IMPL_CASE(16)
        VX_IMPL_CASE(32)
        VX_IMPL_CASE(48)
        VX_IMPL_CASE(64)
        VX_IMPL_CASE(80)
        VX_IMPL_CASE(96)
        VX_IMPL_CASE(112)
        VX_IMPL_CASE(128)
        default: return fail(VX_ERR_UNSUPPORTED, "iso implicit kernel: nz=%d", g.nz);
    }
#undef VX_IMPL_CASE
    return check_launch("iso_implicit_kernel");
}

int iso_diffusion_common(void* stream, const vx_grid* grid, const vx_iso_params* prm, double* tracer, int32_t tau,
                         int32_t taup1, const double* K1, const double* K2, const double* Ai_ez, const double* Ai_nz,
                         const double* Ai_bx, const double* Ai_by, const double* K_11, const double* K_22,
                         const double* K_33, double* dtr_iso, const double* int_drhodX, double* P_diss,
                         void* workspace, int implicit, const char* who) {
    Grid g;
    int rc = make_grid(grid, g, who);
    if (rc) return rc;
    if (!prm) return fail(VX_ERR_INVALID_ARGUMENT, "%s: null params", who);
    if (tau < 0 || tau > 2 || taup1 < 0 || taup1 > 2)
        return fail(VX_ERR_INVALID_ARGUMENT, "%s: time levels must be 0..2", who);
    if (!tracer || !K1 || !K2 || !Ai_ez || !Ai_nz || !Ai_bx || !Ai_by || !K_11 || !K_22 || !dtr_iso || !workspace ||
        (implicit && !K_33))
        return fail(VX_ERR_INVALID_ARGUMENT, "%s: null array pointer", who);
    if ((P_diss == nullptr) != (int_drhodX == nullptr))
        return fail(VX_ERR_INVALID_ARGUMENT, "%s: P_diss and int_drhodX must both be given or both be NULL", who);
    hipStream_t s = (hipStream_t)stream;
    const Prm p = make_prm(prm);
    const int64_t n3 = (int64_t)g.X * g.Y * g.nz;
    double* ws = static_cast<double*>(workspace);
    FluxArgs fa{tracer, tau, K1, K2, Ai_ez, Ai_nz, Ai_bx, Ai_by, K_11, K_22, ws, ws + n3, ws + 2 * n3};
    const int64_t n = (int64_t)(g.X - 3) * (g.Y - 3) * g.nz;
    hipLaunchKernelGGL(iso_flux_kernel, dim3(blocks_for(n, 256)), dim3(256), 0, s, g, fa);
    if ((rc = check_launch(who))) return rc;
    ColArgs ca{tracer, tau, taup1, K_33, dtr_iso, int_drhodX, P_diss, ws, ws + n3, ws + 2 * n3};
    const int64_t ne = (int64_t)(g.X - 2) * (g.Y - 2) * g.nz;
    hipLaunchKernelGGL(iso_explicit_kernel, dim3(blocks_xcd(ne, 256)), dim3(256), 0, s, g, p, ca);
    if ((rc = check_launch(who))) return rc;
    if (implicit && (rc = launch_implicit(g, p, ca, s))) return rc;
    if (P_diss) {
        const int64_t np_ = (int64_t)(g.X - 4) * (g.Y - 4) * (g.nz - 1);
        hipLaunchKernelGGL(iso_post_kernel, dim3(blocks_xcd(np_, 256)), dim3(256), 0, s, g, p, ca, implicit);
        if ((rc = check_launch(who))) return rc;
    }
    return VX_OK;
}
}  // namespace
}  // namespace vx

extern "C" int vx_iso_diffusion_f64(void* stream, const vx_grid* grid, const vx_iso_params* prm, double* tracer,
                                    int32_t tau, int32_t taup1, const double* K_iso, const double* Ai_ez,
                                    const double* Ai_nz, const double* Ai_bx, const double* Ai_by,
                                    const double* K_11, const double* K_22, const double* K_33, double* dtr_iso,
                                    const double* int_drhodX, double* P_diss, void* workspace) {
    return vx::iso_diffusion_common(stream, grid, prm, tracer, tau, taup1, K_iso, K_iso, Ai_ez, Ai_nz, Ai_bx, Ai_by,
                                    K_11, K_22, K_33, dtr_iso, int_drhodX, P_diss, workspace, 1,
                                    "vx_iso_diffusion_f64");
}
