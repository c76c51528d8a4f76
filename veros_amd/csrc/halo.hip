// Two small steps either side of the mixing path on one GPU:
//
// * enforce_boundaries (veros/core/utilities.py:7-21) on one process: with
//   cyclic x, arr[-2:] = arr[2:4] and arr[:2] = arr[-4:-2] for an array whose
//   first axis is x (any trailing shape, given as one strided run per x row);
// * isoneutral_diag_streamfunction_kernel
//   (veros/core/isoneutral/isoneutral.py:232-255): the horizontal components
//   B1_gm, B2_gm of the eddy-driven streamfunction from K_gm and the slopes
//   Ai_nz, Ai_ez, with K_gm edge-padded in z (pad_z_edges,
//   veros/core/utilities.py:24-35: level -1 reads level 0).
#include "vx_common.hpp"

namespace vx {
namespace {

// One thread per element of the 4 halo rows; the source rows (2, 3 and
// X-4, X-3) are disjoint from the destination rows for X >= 8 (nx >= 4).
__global__ __launch_bounds__(256) void cyclic_x_kernel(double* __restrict__ a, int64_t X, int64_t row_elems,
                                                       int64_t row_stride, int64_t elem_stride) {
    const int64_t e = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (e >= 4 * row_elems) return;
    const int h = (int)(e / row_elems);  // halo row 0, 1, X-2, X-1
    const int64_t r = e - (int64_t)h * row_elems;
    const int64_t dst = h < 2 ? h : X - 4 + h;
    const int64_t src = h < 2 ? X - 4 + h : 2 + (h - 2);
    a[dst * row_stride + r * elem_stride] = a[src * row_stride + r * elem_stride];
}

struct SfArgs {
    const double *K, *Aez, *Anz;
    double *B1, *B2;
    int32_t X, Y, nz;
};

// Thread per (i, j, k) of the whole grid; B2_gm on [1, X-2) x [2, Y-2),
// B1_gm on [2, X-2) x [1, Y-2), every level.
__global__ __launch_bounds__(256) void streamfunction_kernel(SfArgs a) {
    const int64_t n = (int64_t)a.X * a.Y * a.nz;
    const int64_t c = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (c >= n) return;
    const int k = (int)(c % a.nz);
    const int64_t col = c / a.nz;
    const int i = (int)(col / a.Y), j = (int)(col % a.Y);
    const int km = k > 0 ? 1 : 0;  // pad_z_edges: level k-1, clamped at the bottom
    if (i >= 1 && i < a.X - 2 && j >= 2 && j < a.Y - 2) {
        const int64_t ce = c + (int64_t)a.Y * a.nz;
        const double diffloc = 0.25 * (a.K[c] + a.K[c - km] + a.K[ce] + a.K[ce - km]);
        const double* A = a.Aez + 4 * c;
        a.B2[c] = 0.25 * diffloc * (((A[0] + A[1]) + A[2]) + A[3]);
    }
    if (i >= 2 && i < a.X - 2 && j >= 1 && j < a.Y - 2) {
        const int64_t cn = c + a.nz;
        const double diffloc = 0.25 * (a.K[c] + a.K[c - km] + a.K[cn] + a.K[cn - km]);
        const double* A = a.Anz + 4 * c;
        a.B1[c] = -0.25 * diffloc * (((A[0] + A[1]) + A[2]) + A[3]);
    }
}

}  // namespace
}  // namespace vx

extern "C" int vx_enforce_boundaries_f64(void* stream, double* arr, int64_t X, int64_t row_elems,
                                         int64_t row_stride, int64_t elem_stride, int32_t enable_cyclic_x) {
    using namespace vx;
    const char* who = "vx_enforce_boundaries_f64";
    if (!enable_cyclic_x) return VX_OK;  // one process, no cyclic x: nothing to exchange
    if (!arr) return fail(VX_ERR_INVALID_ARGUMENT, "%s: null array", who);
    if (X < 8) return fail(VX_ERR_UNSUPPORTED, "%s: cyclic x needs X >= 8 (nx >= 4), got X=%lld", who, (long long)X);
    if (row_elems < 0 || elem_stride < 1 || row_stride < 1 || (row_elems > 0 && (row_elems - 1) * elem_stride >= row_stride))
        return fail(VX_ERR_INVALID_ARGUMENT, "%s: rows of %lld elements at stride %lld overlap (row stride %lld)", who,
                    (long long)row_elems, (long long)elem_stride, (long long)row_stride);
    if (row_elems == 0) return VX_OK;
    const int64_t n = 4 * row_elems;
    hipLaunchKernelGGL(cyclic_x_kernel, dim3((unsigned)((n + 255) / 256)), dim3(256), 0, (hipStream_t)stream, arr, X,
                       row_elems, row_stride, elem_stride);
    return check_launch(who);
}

extern "C" int vx_iso_diag_streamfunction_f64(void* stream, const vx_grid* grid, const double* K_gm,
                                              const double* Ai_ez, const double* Ai_nz, double* B1_gm,
                                              double* B2_gm) {
    using namespace vx;
    const char* who = "vx_iso_diag_streamfunction_f64";
    if (!grid) return fail(VX_ERR_INVALID_ARGUMENT, "%s: null grid", who);
    if (grid->nx < 1 || grid->ny < 1 || grid->nz < 1)
        return fail(VX_ERR_INVALID_ARGUMENT, "%s: bad shape", who);
    if (!K_gm || !Ai_ez || !Ai_nz || !B1_gm || !B2_gm) return fail(VX_ERR_INVALID_ARGUMENT, "%s: null array", who);
    SfArgs a{K_gm, Ai_ez, Ai_nz, B1_gm, B2_gm, (int32_t)(grid->nx + 4), (int32_t)(grid->ny + 4), grid->nz};
    const int64_t n = (int64_t)a.X * a.Y * a.nz;
    hipLaunchKernelGGL(streamfunction_kernel, dim3((unsigned)((n + 255) / 256)), dim3(256), 0, (hipStream_t)stream, a);
    return check_launch(who);
}
