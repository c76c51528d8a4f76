// Isoneutral triad-tensor stencils for gfx950 (float64).
//
//   iso_pre_kernel      isoneutral_diffusion_pre   veros/core/isoneutral/isoneutral.py:18-229
//   iso_flux_kernel     _calc_tracer_fluxes        veros/core/isoneutral/diffusion.py:9-113
//   iso_column_kernel   _calc_explicit_part + _calc_implicit_part + tracer update +
//                       the energy branch          veros/core/isoneutral/diffusion.py:116-283,
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

__device__ __forceinline__ void grad_x(const Grid& g, const Fields& f, int i, int j, int k, double& gT, double& gS) {
    const double m = mT(kb_of(g, i, j), k) * mT(kb_of(g, i + 1, j), k);
    const double den = g.dxu[i] * g.cost[j];
    gT = m * (tr3(f.T, g, i + 1, j, k, f.tau) - tr3(f.T, g, i, j, k, f.tau)) / den;
    gS = m * (tr3(f.S, g, i + 1, j, k, f.tau) - tr3(f.S, g, i, j, k, f.tau)) / den;
}

__device__ __forceinline__ void grad_y(const Grid& g, const Fields& f, int i, int j, int k, double& gT, double& gS) {
    const double m = mT(kb_of(g, i, j), k) * mT(kb_of(g, i, j + 1), k);
    const double den = g.dyu[j];
    gT = m * (tr3(f.T, g, i, j + 1, k, f.tau) - tr3(f.T, g, i, j, k, f.tau)) / den;
    gS = m * (tr3(f.S, g, i, j + 1, k, f.tau) - tr3(f.S, g, i, j, k, f.tau)) / den;
}

// dT/dz at the top face of (i, j, k); 0 at the top level.
__device__ __forceinline__ void grad_z(const Grid& g, const Fields& f, int i, int j, int k, double& gT, double& gS) {
    if (k >= g.nz - 1) {
        gT = 0.0;
        gS = 0.0;
        return;
    }
    const int kb = kb_of(g, i, j);
    const double m = mT(kb, k) * mT(kb, k + 1);
    const double den = g.dzw[k];
    gT = m * (tr3(f.T, g, i, j, k + 1, f.tau) - tr3(f.T, g, i, j, k, f.tau)) / den;
    gS = m * (tr3(f.S, g, i, j, k + 1, f.tau) - tr3(f.S, g, i, j, k, f.tau)) / den;
}

__device__ __forceinline__ double taper(const Prm& p, double s) {
    // dm_taper (veros/core/isoneutral/isoneutral.py:10-15)
    return 0.5 * (1.0 + tanh((-fabs(s) + p.iso_slopec) / p.iso_dslope));
}

// One triad: slope from a horizontal density gradient and a vertical one,
// both weighted by the same drho/dX at a T point (mask folded in).
__device__ __forceinline__ double slope(const Prm& p, double mt, double gxT, double gxS, double gzT, double gzS) {
    const double drdT = mt * p.drhodT;
    const double drdS = mt * p.drhodS;
    const double num = drdT * gxT + drdS * gxS;
    const double den = drdT * gzT + drdS * gzS;
    return -num / (fmin(0.0, den) - kEpsln);
}

// ---------------------------------------------------------------------------
// isoneutral_diffusion_pre: thread per cell (i, j, k), i in [1, X-2), j in [1, Y-2)
// ---------------------------------------------------------------------------
__global__ __launch_bounds__(256) void iso_pre_kernel(Grid g, Prm p, Fields f, const double* __restrict__ K,
                                                      double* __restrict__ Aez, double* __restrict__ Anz,
                                                      double* __restrict__ Abx, double* __restrict__ Aby,
                                                      double* __restrict__ K11, double* __restrict__ K22,
                                                      double* __restrict__ K33) {
    const int nz = g.nz;
    const int ny3 = g.Y - 3;
    const int64_t t = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    const int64_t col = t / nz;
    if (col >= (int64_t)(g.X - 3) * ny3) return;
    const int k = (int)(t - col * nz);
    const int i = 1 + (int)(col / ny3);
    const int j = 1 + (int)(col - (int64_t)(i - 1) * ny3);
    const int64_t c = idx3(g, i, j, k);
    const int kbc = kb_of(g, i, j);

    // gradients at this cell's faces, shared by the three face blocks
    double gzT_c, gzS_c, gzT_m, gzS_m;  // top faces of (i,j,k) and (i,j,k-1)
    grad_z(g, f, i, j, k, gzT_c, gzS_c);
    if (k >= 1) grad_z(g, f, i, j, k - 1, gzT_m, gzS_m);
    else { gzT_m = 0.0; gzS_m = 0.0; }

    // --- east face (i+1/2, j, k), j in [2, Y-2)   (:97-132)
    if (j >= 2) {
        const int kbe = kb_of(g, i + 1, j);
        const double mU = mT(kbc, k) * mT(kbe, k);
        double gxT, gxS;
        grad_x(g, f, i, j, k, gxT, gxS);
        double gzT_e, gzS_e, gzT_em = 0.0, gzS_em = 0.0;
        grad_z(g, f, i + 1, j, k, gzT_e, gzS_e);
        if (k >= 1) grad_z(g, f, i + 1, j, k - 1, gzT_em, gzS_em);
        const double diffloc = (k >= 1) ? 0.25 * (K[idx3(g, i, j, k)] + K[idx3(g, i, j, k - 1)] +
                                                  K[idx3(g, i + 1, j, k)] + K[idx3(g, i + 1, j, k - 1)])
                                        : 0.5 * (K[idx3(g, i, j, 0)] + K[idx3(g, i + 1, j, 0)]);
        double sumz = 0.0;
#pragma unroll
        for (int kr = 0; kr < 2; ++kr) {
            if (k < 1 - kr) continue;  // levels k >= 1 - kr
#pragma unroll
            for (int ip = 0; ip < 2; ++ip) {
                const double mt = mT(ip ? kbe : kbc, k);
                const double zT = ip ? (kr ? gzT_e : gzT_em) : (kr ? gzT_c : gzT_m);
                const double zS = ip ? (kr ? gzS_e : gzS_em) : (kr ? gzS_c : gzS_m);
                const double s = slope(p, mt, gxT, gxS, zT, zS);
                const double tp = taper(p, s);
                sumz += g.dzw[k - 1 + kr] * mU * fmax(p.K_iso_steep, diffloc * tp);
                Aez[c * 4 + ip * 2 + kr] = tp * s * mU;
            }
        }
        K11[c] = sumz / (4.0 * g.dzt[k]);
    }

    // --- north face (i, j+1/2, k), i in [2, X-2)   (:134-168)
    if (i >= 2) {
        const int kbn = kb_of(g, i, j + 1);
        const double mV = mT(kbc, k) * mT(kbn, k);
        double gyT, gyS;
        grad_y(g, f, i, j, k, gyT, gyS);
        double gzT_n, gzS_n, gzT_nm = 0.0, gzS_nm = 0.0;
        grad_z(g, f, i, j + 1, k, gzT_n, gzS_n);
        if (k >= 1) grad_z(g, f, i, j + 1, k - 1, gzT_nm, gzS_nm);
        const double diffloc = (k >= 1) ? 0.25 * (K[idx3(g, i, j, k)] + K[idx3(g, i, j, k - 1)] +
                                                  K[idx3(g, i, j + 1, k)] + K[idx3(g, i, j + 1, k - 1)])
                                        : 0.5 * (K[idx3(g, i, j, 0)] + K[idx3(g, i, j + 1, 0)]);
        double sumz = 0.0;
#pragma unroll
        for (int kr = 0; kr < 2; ++kr) {
            if (k < 1 - kr) continue;
#pragma unroll
            for (int jp = 0; jp < 2; ++jp) {
                const double mt = mT(jp ? kbn : kbc, k);
                const double zT = jp ? (kr ? gzT_n : gzT_nm) : (kr ? gzT_c : gzT_m);
                const double zS = jp ? (kr ? gzS_n : gzS_nm) : (kr ? gzS_c : gzS_m);
                const double s = slope(p, mt, gyT, gyS, zT, zS);
                const double tp = taper(p, s);
                sumz += g.dzw[k - 1 + kr] * mV * fmax(p.K_iso_steep, diffloc * tp);
                Anz[c * 4 + jp * 2 + kr] = tp * s * mV;
            }
        }
        K22[c] = sumz / (4.0 * g.dzt[k]);
    }

    // --- top face (i, j, k+1/2), i, j in [2, X/Y-2), k < nz-1   (:170-225)
    if (i >= 2 && j >= 2 && k < nz - 1) {
        const double mW = mT(kbc, k) * mT(kbc, k + 1);
        const double Kc = K[c];
        double sumx = 0.0, sumy = 0.0;
#pragma unroll
        for (int kr = 0; kr < 2; ++kr) {
            const double mt = mT(kbc, k + kr);
#pragma unroll
            for (int ip = 0; ip < 2; ++ip) {
                double gxT, gxS;
                grad_x(g, f, i - 1 + ip, j, k + kr, gxT, gxS);
                const double s = slope(p, mt, gxT, gxS, gzT_c, gzS_c);
                const double tp = taper(p, s);
                sumx += g.dxu[i - 1 + ip] * Kc * tp * (s * s) * mW;
                Abx[c * 4 + ip * 2 + kr] = tp * s * mW;
            }
#pragma unroll
            for (int jp = 0; jp < 2; ++jp) {
                double gyT, gyS;
                grad_y(g, f, i, j - 1 + jp, k + kr, gyT, gyS);
                const double s = slope(p, mt, gyT, gyS, gzT_c, gzS_c);
                const double tp = taper(p, s);
                const double facty = g.cosu[j - 1 + jp] * g.dyu[j - 1 + jp];
                sumy += facty * Kc * tp * (s * s) * mW;
                Aby[c * 4 + jp * 2 + kr] = tp * s * mW;
            }
        }
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
// column kernel: thread per column (i, j) in [1, X-1) x [1, Y-1).
//  - every such column: P_diss += dissipation_on_wgrid(compute_dissipation(...))
//  - interior columns additionally: explicit divergence, tracer and dtr_iso
//    update, implicit K_33 solve (Thomas in registers), the implicit part of
//    dtr_iso, and the vertical-mixing dissipation term.
// ---------------------------------------------------------------------------
struct ColArgs {
    double* tr;
    int tau, taup1;
    const double* K33;
    double* dtr;
    const double* I;   // int_drhodX (X, Y, nz, 3), nullable
    double* P;         // P_diss, nullable
    const double *fe, *fn, *ft;
    int implicit;      // 1: iso (implicit K33 solve), 0: skew (explicit only)
};

// Face fluxes with their valid index ranges (0 elsewhere).
__device__ __forceinline__ double fe_at(const Grid& g, const ColArgs& a, int i, int j, int k) {
    return (i >= 1 && i <= g.X - 3 && j >= 2 && j <= g.Y - 3) ? a.fe[idx3(g, i, j, k)] : 0.0;
}
__device__ __forceinline__ double fn_at(const Grid& g, const ColArgs& a, int i, int j, int k) {
    return (i >= 2 && i <= g.X - 3 && j >= 1 && j <= g.Y - 3) ? a.fn[idx3(g, i, j, k)] : 0.0;
}

template <int NZB>
__global__ __launch_bounds__(256) void iso_column_kernel(Grid g, Prm p, ColArgs a) {
    const int nz = g.nz;
    const int ny2 = g.Y - 2;
    const int64_t col = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (col >= (int64_t)(g.X - 2) * ny2) return;
    const int i = 1 + (int)(col / ny2);
    const int j = 1 + (int)(col - (int64_t)(i - 1) * ny2);
    const bool interior = i >= 2 && i <= g.X - 3 && j >= 2 && j <= g.Y - 3;
    const int kb = kb_of(g, i, j);
    const bool energy = a.P != nullptr;

    // ---- dissipation of lateral fluxes on W points (veros/core/diffusion.py:9-62)
    if (energy) {
        const double g05 = 0.5 * p.grav / p.rho_0;
        const double rx = g.dxt[i] * g.cost[j];
        const double ry = g.dyt[j] * g.cost[j];
        const bool land = kb >= 1;          // create_water_masks(kbot): ks = kbot - 1 >= 0
        double diss_next = 0.0;             // diss at k+1
        for (int k = nz - 1; k >= 0; --k) {
            const double Ic = tr3(a.I, g, i, j, k, a.tau);
            const double de = (tr3(a.I, g, i + 1, j, k, a.tau) - Ic) * fe_at(g, a, i, j, k) +
                              (Ic - tr3(a.I, g, i - 1, j, k, a.tau)) * fe_at(g, a, i - 1, j, k);
            const double dn = (tr3(a.I, g, i, j + 1, k, a.tau) - Ic) * fn_at(g, a, i, j, k) +
                              (Ic - tr3(a.I, g, i, j - 1, k, a.tau)) * fn_at(g, a, i, j - 1, k);
            const double diss = g05 * de / rx + g05 * dn / ry;
            double dw;
            if (k == nz - 1) {
                dw = diss * (land ? 1.0 : 0.0);
            } else {
                const double avg = 0.5 * (diss + diss_next);
                const bool edge = land && k == kb - 1;
                const bool water = land && k > kb - 1;
                const double dzw_lo = g.dzw[k >= 1 ? k - 1 : 0];  // pad_z_edges(dzw)[k]
                dw = (avg + 0.5 * (diss * dzw_lo / g.dzw[k])) * (edge ? 1.0 : 0.0) + avg * (water ? 1.0 : 0.0);
            }
            double* Pk = a.P + idx3(g, i, j, k);
            *Pk = *Pk + dw;
            diss_next = diss;
        }
    }
    if (!interior) return;

    const double dt = p.dt_tracer;
    const double cx = g.cost[j] * g.dxt[i];
    const double cy = g.cost[j] * g.dyt[j];
    const int k0 = (kb >= 1 && kb <= nz) ? kb - 1 : nz;  // first wet level of the implicit system
    double* __restrict__ trc = a.tr;

    // ---- explicit part + forward elimination of the implicit system
    double cp[NZB], dp[NZB];
    double cpp = 0.0, dpp = 0.0;
    double delta_m = 0.0;   // delta at k-1
    double ft_m = 0.0;      // flux_top at k-1
#pragma unroll
    for (int k = 0; k < NZB; ++k) {
        if (k < nz) {
            const int64_t ck = idx3(g, i, j, k);
            const double mt = mT(kb, k);
            const double ftk = a.ft[ck];
            const double hor = (a.fe[ck] - fe_at(g, a, i - 1, j, k)) / cx + (a.fn[ck] - fn_at(g, a, i, j - 1, k)) / cy;
            const double vert = (k == 0) ? ftk / g.dzt[0] : (ftk - ft_m) / g.dzt[k];
            const double dtrk = mt * hor + mt * vert;
            ft_m = ftk;
            a.dtr[ck] = a.dtr[ck] + dtrk;
            const double rhs = trc[ck * 3 + a.taup1] + dt * dtrk;
            trc[ck * 3 + a.taup1] = rhs;
            if (a.implicit) {
                const double delta = (k < nz - 1) ? dt / g.dzw[k] * a.K33[ck] : 0.0;
                if (k >= k0) {
                    const double dz = g.dzt[k];
                    const bool edge = k == k0;
                    const double ak = edge ? 0.0 : -delta_m / dz;
                    const double bk = edge ? 1 + delta / dz : 1 + (delta + delta_m) / dz;
                    const double ck2 = (k < nz - 1) ? -delta / dz : 0.0;
                    const double r = 1.0 / (bk - ak * cpp);
                    cpp = ck2 * r;
                    dpp = (rhs - ak * dpp) * r;
                }
                delta_m = delta;
            }
            cp[k] = cpp;
            dp[k] = dpp;
        }
    }
    if (!a.implicit) {
        if (energy) {
            const double gr = -p.grav / p.rho_0;
            for (int k = 0; k < nz - 1; ++k) {
                const int64_t ck = idx3(g, i, j, k);
                const double mw = mT(kb, k) * mT(kb, k + 1);
                const double fxa = (-tr3(a.I, g, i, j, k + 1, a.tau) + tr3(a.I, g, i, j, k, a.tau)) / g.dzw[k];
                a.P[ck] = a.P[ck] + gr * fxa * a.ft[ck] * mw;
            }
        }
        return;
    }

    // ---- back substitution, tracer/dtr_iso update, vertical dissipation term
    const double gr = -p.grav / p.rho_0;
    double xn = 0.0;       // new tracer at k+1
#pragma unroll
    for (int k = NZB - 1; k >= 0; --k) {
        if (k < nz) {
            const int64_t ck = idx3(g, i, j, k);
            double xk;
            if (k >= k0) {
                xk = dp[k] - cp[k] * xn;
                const double rhs = trc[ck * 3 + a.taup1];
                trc[ck * 3 + a.taup1] = xk;
                a.dtr[ck] = a.dtr[ck] + (xk - rhs) / dt;
            } else {
                xk = trc[ck * 3 + a.taup1];
            }
            if (energy && k < nz - 1) {
                const double mw = mT(kb, k) * mT(kb, k + 1);
                const double fxa = (-tr3(a.I, g, i, j, k + 1, a.tau) + tr3(a.I, g, i, j, k, a.tau)) / g.dzw[k];
                a.P[ck] = a.P[ck] + gr * fxa * (a.ft[ck] * mw + a.K33[ck] * (xn - xk) / g.dzw[k] * mw);
            }
            xn = xk;
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
    const int64_t n = (int64_t)(g.X - 3) * (g.Y - 3) * g.nz;
    hipLaunchKernelGGL(iso_pre_kernel, dim3(blocks_for(n, 256)), dim3(256), 0, s, g, p, f, K_iso, Ai_ez, Ai_nz,
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
int launch_column(const Grid& g, const Prm& p, const ColArgs& a, hipStream_t s) {
    const int64_t ncol = (int64_t)(g.X - 2) * (g.Y - 2);
    const dim3 grid(blocks_for(ncol, 256)), block(256);
    const int nzb = ((g.nz + 15) / 16) * 16;
#define VX_COL_CASE(N) \
    case N: hipLaunchKernelGGL(iso_column_kernel<N>, grid, block, 0, s, g, p, a); break;
    switch (nzb) {
        VX_COL_CASE(16)
        VX_COL_CASE(32)
        VX_COL_CASE(48)
        VX_COL_CASE(64)
        VX_COL_CASE(80)
        VX_COL_CASE(96)
        VX_COL_CASE(112)
        VX_COL_CASE(128)
        default: return fail(VX_ERR_UNSUPPORTED, "iso column kernel: nz=%d", g.nz);
    }
#undef VX_COL_CASE
    return check_launch("iso_column_kernel");
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
    ColArgs ca{tracer, tau, taup1, K_33, dtr_iso, int_drhodX, P_diss, ws, ws + n3, ws + 2 * n3, implicit};
    return launch_column(g, p, ca, s);
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
