// Segmented (partitioned) tridiagonal solve: kSegP lanes per column.
//
// The column's levels [0, P*S) are cut into P segments of S levels; lane s of
// the column's lane group holds levels [s*S, s*S + S) in registers. Rows
// outside the wet suffix [k0, nz) are identity rows (a = c = d = 0, b = 1), as
// in solve_tridiagonal_jax (veros/core/operators.py:133-136), so a single
// segmented system covers the whole column and its dry part solves to exact
// zeros.
//
// Per segment the last level is the interface unknown L_s. The S-1 interior
// rows are eliminated in the lane (forward sweep carrying the left spike from
// L_{s-1}, back substitution), which expresses every interior x as
//   x_t = alpha_t + beta_t * L_{s-1} + gamma_t * L_s.
// Substituting that into the interface rows gives a P x P tridiagonal system
// in L_0..L_{P-1}, one row per lane (its rows need alpha/beta/gamma of the
// next lane's first level); the group solves it by parallel cyclic reduction
// with DPP row shifts (in float64 followed by one refinement step of the
// interface rows, iface_solve; in the float32 factorisation of the mixed form
// the float64 refinement of the whole column recovers the lost digits), and
// each lane then reconstructs its own levels.
//
// Two ways in: the register variant loads the lane's S levels of a, b, c, d
// itself (any alignment); the LDS variant (16-byte aligned operands) copies
// the wave's 8 columns into LDS with coalesced global_load_lds_dwordx4 and
// stores x back coalesced through LDS. Either way all loads of a wave are in
// flight at once: one memory round trip per column.
#pragma once
#include "tdma_core.hpp"

namespace vx {

constexpr int kSegP = 8;              // lanes per column
constexpr int kSegBlock = 256;        // register variant: 32 columns per block
constexpr int kSegLdsBlock = 128;     // LDS variant: 2 waves, 16 columns per block
constexpr int kSegCols = 64 / kSegP;  // columns per wave

// levels per lane for a depth nz (even, >= 2)
__host__ __device__ constexpr int seg_levels(int nz) {
    const int s = (nz + kSegP - 1) / kSegP;
    return s < 2 ? 2 : s + (s & 1);
}

// LDS bytes of one wave in the LDS variant: a, b, c, d images of 8 columns
__host__ __device__ constexpr int seg_lds_wave_bytes(int S) { return 4 * kSegCols * kSegP * S * 8; }

struct SegLane {
    int64_t base;  // column offset in a, b, c, d, x
    int s;         // segment index in [0, kSegP)
    int lo;        // first level of the segment
    int k0;        // first wet level (nz when the column is dry)
    bool valid;    // the column exists
};

__device__ __forceinline__ SegLane seg_lane_of(const TdmaArgs& p, int S) {
    SegLane L;
    const int64_t gt = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    const int64_t col = gt / kSegP;
    L.s = (int)(gt % kSegP);
    L.lo = L.s * S;
    L.valid = col < p.ncol;
    L.base = 0;
    L.k0 = p.nz;
    if (L.valid) {
        const int64_t i = col / p.ny;
        const int64_t j = col - i * p.ny;
        L.base = i * p.sx + j * p.sy;
        const int kb = p.kbot[i * p.kx + j * p.ky];
        L.k0 = (kb >= 1 && kb <= p.nz) ? kb - 1 : p.nz;
    }
    return L;
}

// 1/x from v_rcp_f64 refined by two Newton steps (about 1 ulp; finite
// non-zero x only).
__device__ __forceinline__ double seg_rcp(double x) {
    double r = __builtin_amdgcn_rcp(x);
    double e = fma(-x, r, 1.0);
    r = fma(r, e, r);
    e = fma(-x, r, 1.0);
    return fma(r, e, r);
}
__device__ __forceinline__ float seg_rcp(float x) { return __builtin_amdgcn_rcpf(x); }

// Moves between the lanes of a group. A group's 8 lanes lie inside one
// 16-lane DPP row, so row shifts reach every lane of the group; values that
// come from the neighbouring group (or from outside the row, read as 0) are
// replaced by the callers.
constexpr int kDppShl = 0x100;  // row_shl:n  lane l reads lane l+n
constexpr int kDppShr = 0x110;  // row_shr:n  lane l reads lane l-n

template <int CTRL>
__device__ __forceinline__ float dpp_move(float v) {
    return __int_as_float(__builtin_amdgcn_update_dpp(0, __float_as_int(v), CTRL, 0xF, 0xF, true));
}
template <int CTRL>
__device__ __forceinline__ double dpp_move(double v) {
    const long long b = __double_as_longlong(v);
    const int lo = __builtin_amdgcn_update_dpp(0, (int)b, CTRL, 0xF, 0xF, true);
    const int hi = __builtin_amdgcn_update_dpp(0, (int)(b >> 32), CTRL, 0xF, 0xF, true);
    return __longlong_as_double(((long long)hi << 32) | (unsigned int)lo);
}

// The interface system, one row per lane of the group (A x_{s-1} + B x_s +
// C x_{s+1} = D), solved by parallel cyclic reduction over strides 1, 2, 4:
// each step folds the rows s-d and s+d into row s, so that row s couples
// only to s-2d and s+2d; rows outside the group act as identity rows (the
// rows s < d already have A = 0, those with s + d >= P have C = 0). After
// the last stride every row is diagonal. Each step's multipliers depend on
// the matrix only; PcrFactors keeps them for further right-hand sides.
template <class T>
struct PcrFactors {
    T al[3], ga[3];  // multipliers of the rows below and above, per stride
    T rb;            // 1 / the final diagonal
};

template <int D, class T>
__device__ __forceinline__ void pcr_step(int s, T& A, T& B, T& C, T& R, T& al, T& ga) {
    T Am = dpp_move<kDppShr + D>(A), Bm = dpp_move<kDppShr + D>(B), Cm = dpp_move<kDppShr + D>(C);
    T Rm = dpp_move<kDppShr + D>(R);
    T Ap = dpp_move<kDppShl + D>(A), Bp = dpp_move<kDppShl + D>(B), Cp = dpp_move<kDppShl + D>(C);
    T Rp = dpp_move<kDppShl + D>(R);
    if (s < D) {
        Am = T(0);
        Bm = T(1);
        Cm = T(0);
        Rm = T(0);
    }
    if (s + D >= kSegP) {
        Ap = T(0);
        Bp = T(1);
        Cp = T(0);
        Rp = T(0);
    }
    al = -A * seg_rcp(Bm);
    ga = -C * seg_rcp(Bp);
    A = al * Am;
    C = ga * Cp;
    B = B + al * Cm + ga * Ap;
    R = R + al * Rm + ga * Rp;
}

// Solves the group's interface system; returns this lane's unknown and
// fills F (when non-null) for pcr_resolve.
template <class T>
__device__ __forceinline__ T pcr_solve(int s, T A, T B, T C, T R, PcrFactors<T>* F) {
    T al, ga;
    pcr_step<1>(s, A, B, C, R, al, ga);
    if (F) F->al[0] = al, F->ga[0] = ga;
    pcr_step<2>(s, A, B, C, R, al, ga);
    if (F) F->al[1] = al, F->ga[1] = ga;
    pcr_step<4>(s, A, B, C, R, al, ga);
    if (F) F->al[2] = al, F->ga[2] = ga;
    const T rb = seg_rcp(B);
    if (F) F->rb = rb;
    return R * rb;
}

template <int D, class T>
__device__ __forceinline__ T pcr_rhs_step(int s, T R, T al, T ga) {
    T Rm = dpp_move<kDppShr + D>(R);
    T Rp = dpp_move<kDppShl + D>(R);
    if (s < D) Rm = T(0);
    if (s + D >= kSegP) Rp = T(0);
    return R + al * Rm + ga * Rp;
}

// Another right-hand side with the factors of pcr_solve.
template <class T>
__device__ __forceinline__ T pcr_resolve(int s, T R, const PcrFactors<T>& F) {
    R = pcr_rhs_step<1>(s, R, F.al[0], F.ga[0]);
    R = pcr_rhs_step<2>(s, R, F.al[1], F.ga[1]);
    R = pcr_rhs_step<4>(s, R, F.al[2], F.ga[2]);
    return R * F.rb;
}

// x of the lane above / below in the group (0 at the group's ends).
template <class T>
__device__ __forceinline__ T grp_next(int s, T v) {
    const T w = dpp_move<kDppShl + 1>(v);
    return s == kSegP - 1 ? T(0) : w;
}
template <class T>
__device__ __forceinline__ T grp_prev(int s, T v) {
    const T w = dpp_move<kDppShr + 1>(v);
    return s == 0 ? T(0) : w;
}

// The float64 interface system: parallel cyclic reduction, then, in a group
// where some interface row is not strongly diagonally dominant (|A| + |C| >
// |B| / 4), one step of iterative refinement (the float64 residual of the
// interface rows, neighbours by DPP, solved with the same factors and added).
// Cyclic reduction alone loses digits on strongly coupled columns, whose
// interface rows are far from diagonally dominant (a 16x12x20 friction case
// with large viscosity: 4e-12 relative in u, against 3e-14 after the
// refinement step, the accuracy of a sequential sweep of the interface rows);
// on strongly dominant rows it is as accurate as the sweep. The branch is
// taken by the whole wave when any of its groups needs the step, so every
// lane takes part in the moves. NR right-hand sides
// share the factors; Ls[q] is this lane's unknown, Lm[q] the one of the lane
// below (0 in lane 0).
template <int NR>
__device__ __forceinline__ void iface_solve(int s, double A, double B, double C, const double (&R)[NR],
                                            double (&Ls)[NR], double (&Lm)[NR]) {
    PcrFactors<double> F;
#pragma unroll
    for (int q = 0; q < NR; ++q) Ls[q] = q == 0 ? pcr_solve<double>(s, A, B, C, R[0], &F) : pcr_resolve(s, R[q], F);
    const bool weak = fabs(A) + fabs(C) > 0.25 * fabs(B);
    if (__any(weak)) {
        // the step is kept only in groups with such a row, so a column's
        // result does not depend on the columns that share its wave
        int gw = weak ? 1 : 0;
        gw |= __shfl_xor(gw, 1, kSegP);
        gw |= __shfl_xor(gw, 2, kSegP);
        gw |= __shfl_xor(gw, 4, kSegP);
#pragma unroll
        for (int q = 0; q < NR; ++q) {
            const double res = R[q] - (A * grp_prev(s, Ls[q]) + B * Ls[q] + C * grp_next(s, Ls[q]));
            const double ref = Ls[q] + pcr_resolve(s, res, F);
            Ls[q] = gw ? ref : Ls[q];
        }
    }
#pragma unroll
    for (int q = 0; q < NR; ++q) Lm[q] = grp_prev(s, Ls[q]);
}

// The masked system on the lane's levels: identity rows off the wet suffix,
// a = 0 (and b, d from b_edge, d_edge when given) at the edge row k0, c = 0
// at the top row nz-1. Values at rows off the wet suffix may be garbage on
// entry; they are replaced by selects.
template <int S>
__device__ __forceinline__ void seg_mask_rows(const TdmaArgs& p, const SegLane& L, double be, double de,
                                              double (&va)[S], double (&vb)[S], double (&vc)[S], double (&vd)[S]) {
    const int nz = p.nz;
#pragma unroll
    for (int t = 0; t < S; ++t) {
        const int k = L.lo + t;
        const bool wet = k >= L.k0 && k < nz;
        va[t] = (wet && k != L.k0) ? va[t] : 0.0;
        vb[t] = wet ? vb[t] : 1.0;
        vc[t] = (wet && k != nz - 1) ? vc[t] : 0.0;
        vd[t] = wet ? vd[t] : 0.0;
    }
    // edge-row replacements (uniform branches: the pointers are kernel
    // arguments); a dry column (k0 == nz) has no edge row
    const int ke = L.k0 < nz ? L.k0 - L.lo : -1;
    if (p.b_edge != nullptr) {
#pragma unroll
        for (int t = 0; t < S; ++t) vb[t] = (t == ke) ? be : vb[t];
    }
    if (p.d_edge != nullptr) {
#pragma unroll
        for (int t = 0; t < S; ++t) vd[t] = (t == ke) ? de : vd[t];
    }
}

// b_edge / d_edge at the edge row, loaded by the lane whose segment holds it.
__device__ __forceinline__ void seg_edge_values(const TdmaArgs& p, const SegLane& L, int S, double& be,
                                                double& de) {
    be = 0.0;
    de = 0.0;
    const bool edge_here = L.valid && L.k0 < p.nz && L.k0 >= L.lo && L.k0 < L.lo + S;
    if (edge_here && p.b_edge) be = p.b_edge[L.base + L.k0];
    if (edge_here && p.d_edge) de = p.d_edge[L.base + L.k0];
}

// ---------------------------------------------------------------------------
// float64 segmented solve, in place: on return z holds the lane's levels of
// x; va and vc are overwritten by the sweep's factors.
// ---------------------------------------------------------------------------
template <int S>
__device__ __forceinline__ void seg_solve_f64(const SegLane& L, double (&va)[S], const double (&vb)[S],
                                              double (&vc)[S], double (&z)[S]) {
    // a (spike w, then beta), c (c', then gamma) and z (dp, then alpha) are
    // transformed in place; every entry is read before it is overwritten, and
    // the interface level S-1 is left untouched.
    const double aL = va[S - 1], bL = vb[S - 1], cL = vc[S - 1], dL = z[S - 1];
    {
        const double r = seg_rcp(vb[0]);
        vc[0] *= r;
        va[0] *= r;
        z[0] *= r;
    }
#pragma unroll
    for (int t = 1; t < S - 1; ++t) {
        const double at = va[t];
        const double r = seg_rcp(vb[t] - at * vc[t - 1]);
        vc[t] *= r;
        va[t] = -at * va[t - 1] * r;
        z[t] = (z[t] - at * z[t - 1]) * r;
    }
    va[S - 2] = -va[S - 2];
    vc[S - 2] = -vc[S - 2];
#pragma unroll
    for (int t = S - 3; t >= 0; --t) {
        const double cp = vc[t];
        z[t] = z[t] - cp * z[t + 1];
        va[t] = -va[t] - cp * va[t + 1];
        vc[t] = -cp * vc[t + 1];
    }
    const double n_al = grp_next(L.s, z[0]);
    const double n_be = grp_next(L.s, va[0]);
    const double n_ga = grp_next(L.s, vc[0]);
    const double RA = aL * va[S - 2];
    const double RB = bL + aL * vc[S - 2] + cL * n_be;
    const double RC = cL * n_ga;
    const double RD[1] = {dL - aL * z[S - 2] - cL * n_al};
    double Ls[1], Lm[1];
    iface_solve<1>(L.s, RA, RB, RC, RD, Ls, Lm);
#pragma unroll
    for (int t = 0; t < S - 1; ++t) z[t] = z[t] + va[t] * Lm[0] + vc[t] * Ls[0];
    z[S - 1] = Ls[0];
}

// Two right-hand sides of one matrix (z1, z2 in, their solutions out); the
// sweep's factors are formed once. va and vc are overwritten as in
// seg_solve_f64.
template <int S>
__device__ __forceinline__ void seg_solve2_f64(const SegLane& L, double (&va)[S], const double (&vb)[S],
                                               double (&vc)[S], double (&z1)[S], double (&z2)[S]) {
    const double aL = va[S - 1], bL = vb[S - 1], cL = vc[S - 1], d1 = z1[S - 1], d2 = z2[S - 1];
    {
        const double r = seg_rcp(vb[0]);
        vc[0] *= r;
        va[0] *= r;
        z1[0] *= r;
        z2[0] *= r;
    }
#pragma unroll
    for (int t = 1; t < S - 1; ++t) {
        const double at = va[t];
        const double r = seg_rcp(vb[t] - at * vc[t - 1]);
        vc[t] *= r;
        va[t] = -at * va[t - 1] * r;
        z1[t] = (z1[t] - at * z1[t - 1]) * r;
        z2[t] = (z2[t] - at * z2[t - 1]) * r;
    }
    va[S - 2] = -va[S - 2];
    vc[S - 2] = -vc[S - 2];
#pragma unroll
    for (int t = S - 3; t >= 0; --t) {
        const double cp = vc[t];
        z1[t] = z1[t] - cp * z1[t + 1];
        z2[t] = z2[t] - cp * z2[t + 1];
        va[t] = -va[t] - cp * va[t + 1];
        vc[t] = -cp * vc[t + 1];
    }
    const double n_be = grp_next(L.s, va[0]);
    const double n_ga = grp_next(L.s, vc[0]);
    const double RA = aL * va[S - 2];
    const double RB = bL + aL * vc[S - 2] + cL * n_be;
    const double RC = cL * n_ga;
    const double RD[2] = {d1 - aL * z1[S - 2] - cL * grp_next(L.s, z1[0]),
                          d2 - aL * z2[S - 2] - cL * grp_next(L.s, z2[0])};
    double Ls[2], Lm[2];
    iface_solve<2>(L.s, RA, RB, RC, RD, Ls, Lm);
#pragma unroll
    for (int t = 0; t < S - 1; ++t) {
        z1[t] = z1[t] + va[t] * Lm[0] + vc[t] * Ls[0];
        z2[t] = z2[t] + va[t] * Lm[1] + vc[t] * Ls[1];
    }
    z1[S - 1] = Ls[0];
    z2[S - 1] = Ls[1];
}

// ---------------------------------------------------------------------------
// float32 factorisation of the lane's segment and of the group's interface
// system, formed once and reused by every float32 solve of the mixed form.
// ---------------------------------------------------------------------------
template <int S>
struct SegF32 {
    float af[S];       // a of every level
    float rd[S];       // 1 / den of the interior rows 0..S-2
    float cp[S];       // c' of the interior rows
    float be[S];       // beta of the interior rows
    float ga[S];       // gamma of the interior rows
    float aL, cL;      // interface row a, c
    PcrFactors<float> pcr;  // the group's interface system
};

template <int S>
__device__ __forceinline__ void seg_factor_f32(const SegLane& L, const double (&va)[S], const double (&vb)[S],
                                               const double (&vc)[S], SegF32<S>& F) {
#pragma unroll
    for (int t = 0; t < S; ++t) F.af[t] = (float)va[t];
    float w[S];
    {
        const float r = seg_rcp((float)vb[0]);
        F.rd[0] = r;
        F.cp[0] = (float)vc[0] * r;
        w[0] = F.af[0] * r;
    }
#pragma unroll
    for (int t = 1; t < S - 1; ++t) {
        const float r = seg_rcp((float)vb[t] - F.af[t] * F.cp[t - 1]);
        F.rd[t] = r;
        F.cp[t] = (float)vc[t] * r;
        w[t] = -F.af[t] * w[t - 1] * r;
    }
    F.be[S - 2] = -w[S - 2];
    F.ga[S - 2] = -F.cp[S - 2];
#pragma unroll
    for (int t = S - 3; t >= 0; --t) {
        F.be[t] = -w[t] - F.cp[t] * F.be[t + 1];
        F.ga[t] = -F.cp[t] * F.ga[t + 1];
    }
    F.aL = F.af[S - 1];
    F.cL = (float)vc[S - 1];
    const float n_be = grp_next(L.s, F.be[0]);
    const float n_ga = grp_next(L.s, F.ga[0]);
    const float RA = F.aL * F.be[S - 2];
    const float RB = (float)vb[S - 1] + F.aL * F.ga[S - 2] + F.cL * n_be;
    const float RC = F.cL * n_ga;
    (void)pcr_solve<float>(L.s, RA, RB, RC, 0.f, &F.pcr);
}

// z: rhs in, solution out (float32, the lane's levels).
template <int S>
__device__ __forceinline__ void seg_solve_f32(const SegLane& L, const SegF32<S>& F, float (&z)[S]) {
    z[0] *= F.rd[0];
#pragma unroll
    for (int t = 1; t < S - 1; ++t) z[t] = (z[t] - F.af[t] * z[t - 1]) * F.rd[t];
#pragma unroll
    for (int t = S - 3; t >= 0; --t) z[t] = z[t] - F.cp[t] * z[t + 1];
    const float n_al = grp_next(L.s, z[0]);
    const float RD = z[S - 1] - F.aL * z[S - 2] - F.cL * n_al;
    const float Ls = pcr_resolve<float>(L.s, RD, F.pcr);
    const float Lm = grp_prev(L.s, Ls);
#pragma unroll
    for (int t = 0; t < S - 1; ++t) z[t] = z[t] + F.be[t] * Lm + F.ga[t] * Ls;
    z[S - 1] = Ls;
}

// Mixed precision on the masked rows: float32 solve, then REFINE steps of the
// float64 residual r = d - A x solved again in float32 (same factors) and
// added in float64. On return x holds the lane's levels.
template <int S, int REFINE>
__device__ __forceinline__ void seg_mixed_core(const SegLane& L, const double (&va)[S], const double (&vb)[S],
                                               const double (&vc)[S], const double (&vd)[S], double (&x)[S]) {
    SegF32<S> F;
    seg_factor_f32<S>(L, va, vb, vc, F);
    float z[S];
#pragma unroll
    for (int t = 0; t < S; ++t) z[t] = (float)vd[t];
    seg_solve_f32<S>(L, F, z);
#pragma unroll
    for (int t = 0; t < S; ++t) x[t] = (double)z[t];
#pragma unroll 1
    for (int step = 0; step < REFINE; ++step) {
        const double xm = grp_prev(L.s, x[S - 1]);  // level lo-1, from the lane below
        const double xp = grp_next(L.s, x[0]);      // level lo+S, from the lane above
#pragma unroll
        for (int t = 0; t < S; ++t) {
            const double lo_x = t == 0 ? xm : x[t > 0 ? t - 1 : 0];
            const double hi_x = t == S - 1 ? xp : x[t < S - 1 ? t + 1 : S - 1];
            z[t] = (float)(vd[t] - va[t] * lo_x - vb[t] * x[t] - vc[t] * hi_x);
        }
        seg_solve_f32<S>(L, F, z);
#pragma unroll
        for (int t = 0; t < S; ++t) x[t] += (double)z[t];
    }
}

// ---------------------------------------------------------------------------
// Register variant: each lane loads its own levels (8-byte or, when VEC,
// 16-byte loads; VEC needs lo, nz and the column base even and aligned
// operands).
// ---------------------------------------------------------------------------
template <int S, bool VEC>
__device__ __forceinline__ void seg_load(const double* __restrict__ col, int lo, int nz, int k0, double (&v)[S]) {
    if constexpr (VEC) {
#pragma unroll
        for (int t = 0; t < S; t += 2) {
            const int k = lo + t;
            if (k < nz && k + 1 >= k0) {
                const double2 w = *reinterpret_cast<const double2*>(col + k);
                v[t] = w.x;
                v[t + 1] = w.y;
            }
        }
    } else {
#pragma unroll
        for (int t = 0; t < S; ++t) {
            const int k = lo + t;
            if (k < nz && k >= k0) v[t] = col[k];
        }
    }
}

template <int S, bool VEC>
__device__ __forceinline__ void seg_store(double* __restrict__ col, int lo, int nz, const double (&v)[S]) {
    if constexpr (VEC) {
#pragma unroll
        for (int t = 0; t < S; t += 2) {
            if (lo + t < nz) *reinterpret_cast<double2*>(col + lo + t) = make_double2(v[t], v[t + 1]);
        }
    } else {
#pragma unroll
        for (int t = 0; t < S; ++t) {
            if (lo + t < nz) col[lo + t] = v[t];
        }
    }
}

// Levels lo .. lo+S-1 of one column of an array whose level k sits at
// a[cs * k]: VEC uses 16-byte loads of level pairs (cs = 1, lo and nz even,
// 16-byte aligned column). Levels for which want(k) is false (or k >= nz)
// keep their values.
template <int S, bool VEC, class W>
__device__ __forceinline__ void col_load(const double* __restrict__ a, int64_t cs, int lo, int nz, W want,
                                         double (&v)[S]) {
    if constexpr (VEC) {
#pragma unroll
        for (int t = 0; t < S; t += 2) {
            const int k = lo + t;
            if (k < nz && (want(k) || want(k + 1))) {
                const double2 w = *reinterpret_cast<const double2*>(a + k);
                v[t] = w.x;
                v[t + 1] = w.y;
            }
        }
    } else {
#pragma unroll
        for (int t = 0; t < S; ++t) {
            const int k = lo + t;
            if (k < nz && want(k)) v[t] = a[cs * k];
        }
    }
}

// Stores levels lo .. lo+S-1 (k < nz) of one column (layout as col_load).
template <int S, bool VEC>
__device__ __forceinline__ void col_store(double* __restrict__ a, int64_t cs, int lo, int nz, const double (&v)[S]) {
    if constexpr (VEC) {
#pragma unroll
        for (int t = 0; t < S; t += 2) {
            if (lo + t < nz) *reinterpret_cast<double2*>(a + lo + t) = make_double2(v[t], v[t + 1]);
        }
    } else {
#pragma unroll
        for (int t = 0; t < S; ++t) {
            if (lo + t < nz) a[cs * (lo + t)] = v[t];
        }
    }
}

// Stores the levels with st[t] set; with VEC a pair is stored when either
// level is, so the other level's v must then hold the value already there.
template <int S, bool VEC>
__device__ __forceinline__ void col_store_if(double* __restrict__ a, int64_t cs, int lo, int nz, const double (&v)[S],
                                             const bool (&st)[S]) {
    if constexpr (VEC) {
#pragma unroll
        for (int t = 0; t < S; t += 2) {
            if (lo + t < nz && (st[t] || st[t + 1]))
                *reinterpret_cast<double2*>(a + lo + t) = make_double2(v[t], v[t + 1]);
        }
    } else {
#pragma unroll
        for (int t = 0; t < S; ++t) {
            if (lo + t < nz && st[t]) a[cs * (lo + t)] = v[t];
        }
    }
}

template <int S, bool VEC>
__device__ __forceinline__ void seg_rows(const TdmaArgs& p, const SegLane& L, double (&va)[S], double (&vb)[S],
                                         double (&vc)[S], double (&vd)[S]) {
    const int nz = p.nz;
#pragma unroll
    for (int t = 0; t < S; ++t) {
        va[t] = 0.0;
        vb[t] = 1.0;
        vc[t] = 0.0;
        vd[t] = 0.0;
    }
    if (L.valid && L.k0 < nz && L.lo + S > L.k0) {
        seg_load<S, VEC>(p.a + L.base, L.lo, nz, L.k0, va);
        seg_load<S, VEC>(p.b + L.base, L.lo, nz, L.k0, vb);
        seg_load<S, VEC>(p.c + L.base, L.lo, nz, L.k0, vc);
        seg_load<S, VEC>(p.d + L.base, L.lo, nz, L.k0, vd);
    }
    double be, de;
    seg_edge_values(p, L, S, be, de);
    seg_mask_rows<S>(p, L, be, de, va, vb, vc, vd);
}

template <int S, bool VEC>
__device__ __forceinline__ void seg_f64(const TdmaArgs& p, const SegLane& L) {
    double va[S], vb[S], vc[S], vd[S];
    seg_rows<S, VEC>(p, L, va, vb, vc, vd);
    seg_solve_f64<S>(L, va, vb, vc, vd);
    if (L.valid) seg_store<S, VEC>(p.x + L.base, L.lo, p.nz, vd);
}

template <int S, bool VEC, int REFINE>
__device__ __forceinline__ void seg_mixed(const TdmaArgs& p, const SegLane& L) {
    double va[S], vb[S], vc[S], vd[S];
    seg_rows<S, VEC>(p, L, va, vb, vc, vd);
    double x[S];
    seg_mixed_core<S, REFINE>(L, va, vb, vc, vd, x);
    if (L.valid) seg_store<S, VEC>(p.x + L.base, L.lo, p.nz, x);
}

// ---------------------------------------------------------------------------
// LDS variant. A wave owns the 8 columns of its lane groups; the columns' a,
// b, c, d are copied by the wave straight from global memory into its LDS
// region with global_load_lds_dwordx4 (no VGPRs for staging): lane l of the
// copy's instruction i moves the 16-byte pair p = 64 i + l of the wave's LDS
// image [column][P*S], so consecutive lanes read consecutive 16 bytes of a
// column and every instruction is a coalesced run. Only pairs that touch the
// wet suffix are copied. Each lane then reads its segment's levels from LDS,
// and the solution goes back the same way (segments written to LDS, pairs
// stored coalesced). LDS per wave: seg_lds_wave_bytes(S) (20 KiB at nz = 80).
// ---------------------------------------------------------------------------
typedef __attribute__((address_space(3))) void* lds_void_ptr;

__device__ __forceinline__ void seg_wait_vmem() {
    // s_waitcnt vmcnt(0): the wave's LDS copies have landed
    __builtin_amdgcn_s_waitcnt(0x0F70);
}

// The pairs a lane copies for its wave; offsets and predicates are formed
// once and shared by the four arrays. The column's base and first wet level
// come from its lane group by shuffles.
template <int S>
struct SegCopyPlan {
    static constexpr int PS = kSegP * S;
    static constexpr int NPAIR = kSegCols * PS / 2;
    static constexpr int NI = (NPAIR + 63) / 64;
    int64_t off[NI];
    bool pred[NI];
};

template <int S>
__device__ __forceinline__ void seg_copy_plan(const SegLane& L, int lane, int nz, bool wet_only,
                                              SegCopyPlan<S>& cp) {
    using P = SegCopyPlan<S>;
#pragma unroll
    for (int i = 0; i < P::NI; ++i) {
        const int pidx = i * 64 + lane;
        const int c = pidx / (P::PS / 2);
        const int k = 2 * (pidx - c * (P::PS / 2));
        const int src = (c < kSegCols ? c : 0) * kSegP;  // first lane of column c's group
        const int64_t b = __shfl(L.base, src);
        const int k0 = __shfl(L.k0, src);
        const bool v = __shfl((int)L.valid, src) != 0;
        cp.off[i] = b + k;
        cp.pred[i] = pidx < P::NPAIR && v && k < nz && (!wet_only || k + 1 >= k0);
    }
}

template <int S>
__device__ __forceinline__ void seg_lds_copy(const double* __restrict__ arr, double* lds_arr,
                                             const SegCopyPlan<S>& cp) {
#pragma unroll
    for (int i = 0; i < SegCopyPlan<S>::NI; ++i) {
        if (cp.pred[i])
            __builtin_amdgcn_global_load_lds((const void*)(arr + cp.off[i]), (lds_void_ptr)(lds_arr + 128 * i), 16,
                                             0, 0);
    }
}

template <int S>
__device__ __forceinline__ void seg_lds_read(const double* lds_arr, const SegLane& L, double (&v)[S]) {
    const double* src = lds_arr + ((threadIdx.x & 63) / kSegP) * (kSegP * S) + L.lo;
#pragma unroll
    for (int t = 0; t < S; t += 2) {
        const double2 w = *reinterpret_cast<const double2*>(src + t);
        v[t] = w.x;
        v[t + 1] = w.y;
    }
}

template <int S>
__device__ __forceinline__ void seg_rows_lds(const TdmaArgs& p, const SegLane& L, double* lds, double (&va)[S],
                                             double (&vb)[S], double (&vc)[S], double (&vd)[S]) {
    constexpr int IMG = kSegCols * kSegP * S;
    SegCopyPlan<S> cp;
    seg_copy_plan<S>(L, threadIdx.x & 63, p.nz, true, cp);
    seg_lds_copy<S>(p.a, lds + 0 * IMG, cp);
    seg_lds_copy<S>(p.b, lds + 1 * IMG, cp);
    seg_lds_copy<S>(p.c, lds + 2 * IMG, cp);
    seg_lds_copy<S>(p.d, lds + 3 * IMG, cp);
    double be, de;
    seg_edge_values(p, L, S, be, de);
    seg_wait_vmem();
    seg_lds_read<S>(lds + 0 * IMG, L, va);
    seg_lds_read<S>(lds + 1 * IMG, L, vb);
    seg_lds_read<S>(lds + 2 * IMG, L, vc);
    seg_lds_read<S>(lds + 3 * IMG, L, vd);
    seg_mask_rows<S>(p, L, be, de, va, vb, vc, vd);
}

// Store the lane's levels of x through LDS (the a image is free by now; a
// wave's LDS operations complete in order): segments written, then coalesced
// 16-byte pairs stored by consecutive lanes.
template <int S>
__device__ __forceinline__ void seg_store_lds(const TdmaArgs& p, const SegLane& L, double* lds, const double (&x)[S]) {
    constexpr int PS = kSegP * S;
    double* img = lds + ((threadIdx.x & 63) / kSegP) * PS + L.lo;
#pragma unroll
    for (int t = 0; t < S; t += 2) *reinterpret_cast<double2*>(img + t) = make_double2(x[t], x[t + 1]);
    SegCopyPlan<S> cp;
    seg_copy_plan<S>(L, threadIdx.x & 63, p.nz, false, cp);
#pragma unroll
    for (int i = 0; i < SegCopyPlan<S>::NI; ++i) {
        if (cp.pred[i]) {
            const int pidx = i * 64 + (threadIdx.x & 63);
            const double2 w = *reinterpret_cast<const double2*>(lds + 2 * pidx);
            *reinterpret_cast<double2*>(p.x + cp.off[i]) = w;
        }
    }
}

template <int S>
__device__ __forceinline__ void seg_f64_lds(const TdmaArgs& p, const SegLane& L, double* lds) {
    double va[S], vb[S], vc[S], vd[S];
    seg_rows_lds<S>(p, L, lds, va, vb, vc, vd);
    seg_solve_f64<S>(L, va, vb, vc, vd);
    seg_store_lds<S>(p, L, lds, vd);
}

template <int S, int REFINE>
__device__ __forceinline__ void seg_mixed_lds(const TdmaArgs& p, const SegLane& L, double* lds) {
    double va[S], vb[S], vc[S], vd[S];
    seg_rows_lds<S>(p, L, lds, va, vb, vc, vd);
    double x[S];
    seg_mixed_core<S, REFINE>(L, va, vb, vc, vd, x);
    seg_store_lds<S>(p, L, lds, x);
}

}  // namespace vx
