// Internal declarations shared by the libpeec_hip translation units.
#pragma once
#include <hip/hip_runtime.h>
#include <cstdint>
#include <cstdio>
#include <string>
#include <vector>

#include "../../include/peec_hip.h"

#define PEEC_HIP_TRY(expr)                                                       \
    do {                                                                         \
        hipError_t _e = (expr);                                                  \
        if (_e != hipSuccess) {                                                  \
            peec_set_error("%s:%d: %s -> %s", __FILE__, __LINE__, #expr,         \
                           hipGetErrorString(_e));                               \
            return PEEC_ERR_HIP;                                                 \
        }                                                                        \
    } while (0)

void peec_set_error(const char *fmt, ...);

// ---------------------------------------------------------------- integrals
// Device rows: centers are N x 3 column-major (Julia layout) in device memory.
int peec_lp_row_device(hipStream_t s, const double src[3], const double *d_centers, int64_t N,
                       double sx, double sy, double sz, double sx2, double sy2, double sz2,
                       int dc, double scale, double *d_out);
int peec_p_row_device(hipStream_t s, const double src[3], const double *d_centers, int64_t N,
                      double sx, double sy, double sz, int caser, int caser2, double scale,
                      double *d_out);

// Circulant-grid rows computed straight from the grid description (the
// centre arithmetic of build_center_P_Voxels.jl / build_centers_Lp_with_air.jl
// is done on the device).  kind: 0..2 = Lp x/y/z, 3..8 = P 12,34,56,1234,1256,3456.
// variant 0: static rows of the quasi-static path; 1: static rows of the
// retarded (Rcc) path; 2: its centre distances
int peec_grid_row_device(hipStream_t s, int kind, const double origin[3], int Nx, int Ny, int Nz,
                         double sx, double sy, double sz, double scale, double *d_out,
                         int64_t *n_out, int dims_out[3], int variant = 0);

// ---------------------------------------------------------------- fft
struct FftPlan1D {
    int L = 0;
    int nstages = 0;
    int radix[16];
};
bool fft_make_plan(int L, FftPlan1D *p);
int fft_smooth_length(int nmin);
