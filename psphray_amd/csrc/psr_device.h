// psr_device.h -- device helpers shared by the trace and deposit kernels:
// table lookups of the reference (src/tabfun.c, src/atomic-rate.c,
// src/sph-depth.c) and wavefront reductions.
#pragma once
#include <hip/hip_runtime.h>

#include "psr_internal.h"

// clamped linear interpolation of src/tabfun.c:38-58
__device__ __forceinline__ double tab_lin(const double *xcol, const double *ycol, int nrows, double min,
                                          double step_inv, double v) {
    const long long idx = (long long)((v - min) * step_inv);
    if (idx < 0) return ycol[0];
    if (idx >= nrows - 1) return ycol[nrows - 1];
    const double left = xcol[idx], right = xcol[idx + 1];
    const double lw = (right - v), rw = (v - left);
    return (lw * ycol[idx] + rw * ycol[idx + 1]) * step_inv;
}

// cross section with the threshold cut of src/atomic-rate.c:313-318
__device__ __forceinline__ double xs_get(const DevTables &T, int id, double eng) {
    if (eng < T.xs_cut[id]) return 0.0;
    return tab_lin(T.xs, T.xs + (size_t)(id + 1) * T.xs_nrows, T.xs_nrows, T.xs_min, T.xs_step_inv, eng);
}

// global-memory view of a pointer read from device memory (the compiler
// cannot prove the address space and would emit flat loads, which count on
// both the vector-memory and the LDS wait counters)
template <typename T>
__device__ __forceinline__ const __attribute__((address_space(1))) T *gmem(const T *p) {
    return (const __attribute__((address_space(1))) T *)p;
}

// relaxed device-scope add without return through a global-memory pointer:
// a global (not flat) atomic, which counts on the vector-memory wait counter
// only, so the issuing wave's LDS waits do not wait for it
template <typename T>
__device__ __forceinline__ void gadd(T *p, T v) {
    __hip_atomic_fetch_add((__attribute__((address_space(1))) T *)p, v, __ATOMIC_RELAXED,
                           __HIP_MEMORY_SCOPE_AGENT);
}

// src/sph-depth.c:224-231: nearest-lower entry, 0 outside the table; the
// load is unconditional (clamped index) so no branch hides it
__device__ __forceinline__ double sph_depth(const DevTables &T, double r_h) {
    const int n = (int)(T.kernel_n * r_h);
    const bool in = n < T.kernel_n && n >= 0;
    const double v = gmem(T.kernel)[in ? n : 0];
    return in ? v : 0.0;
}

// bilinear TabFun2 lookup of src/tabfun.c:173-223
__device__ inline void tab2_get3(const DevTables &T, double x, double y, double v[3]) {
    // the three tables interleaved, [n1][n2][3]; each value with TabFun2's
    // interpolation (src/tabfun.c:173-223), the indices computed once
    const double *d = T.si_tab3;
    const int n1 = T.si_n1, n2 = T.si_n2;
    const int xind = (int)((x - T.si_min1) * T.si_step1_inv);
    const int yind = (int)((y - T.si_min2) * T.si_step2_inv);
    int linoffset = -1;
    if (xind < 0) linoffset = 0;
    if (xind >= n1 - 1) linoffset = (n1 - 1) * n2;
    if (linoffset != -1) {
        if (yind < 0) {
            for (int t = 0; t < 3; t++) v[t] = d[3 * (size_t)linoffset + t];
            return;
        }
        if (yind >= n2 - 1) {
            for (int t = 0; t < 3; t++) v[t] = d[3 * (size_t)(linoffset + n2 - 1) + t];
            return;
        }
        const double left = T.si_y[yind], right = T.si_y[yind + 1];
        for (int t = 0; t < 3; t++)
            v[t] = ((right - y) * d[3 * (size_t)(linoffset + yind) + t] +
                    (y - left) * d[3 * (size_t)(linoffset + yind + 1) + t]) *
                   T.si_step2_inv;
        return;
    }
    int lin = -1;
    if (yind < 0) lin = 0;
    if (yind >= n2 - 1) lin = n2 - 1;
    if (lin != -1) {
        const double left = T.si_x[xind], right = T.si_x[xind + 1];
        for (int t = 0; t < 3; t++)
            v[t] = ((right - x) * d[3 * ((size_t)xind * n2 + lin) + t] +
                    (x - left) * d[3 * ((size_t)(xind + 1) * n2 + lin) + t]) *
                   T.si_step1_inv;
        return;
    }
    const double x1 = T.si_x[xind], x2 = T.si_x[xind + 1];
    const double y1 = T.si_y[yind], y2 = T.si_y[yind + 1];
    const double fac = T.si_step1_inv * T.si_step2_inv;
    const size_t o11 = 3 * ((size_t)xind * n2 + yind), o21 = 3 * ((size_t)(xind + 1) * n2 + yind);
    for (int t = 0; t < 3; t++) {
        const double f11 = d[o11 + t], f12 = d[o11 + 3 + t];
        const double f21 = d[o21 + t], f22 = d[o21 + 3 + t];
        v[t] = fac * (f11 * (x2 - x) * (y2 - y) + f21 * (x - x1) * (y2 - y) + f12 * (x2 - x) * (y - y1) +
                      f22 * (x - x1) * (y - y1));
    }
}

__device__ __forceinline__ double wave_sum(double v) {
#pragma unroll
    for (int o = 32; o > 0; o >>= 1) v += __shfl_xor(v, o, 64);
    return v;
}
__device__ __forceinline__ unsigned long long wave_sum_u64(unsigned long long v) {
#pragma unroll
    for (int o = 32; o > 0; o >>= 1) v += __shfl_xor(v, o, 64);
    return v;
}
__device__ __forceinline__ int wave_sum_i(int v) {
#pragma unroll
    for (int o = 32; o > 0; o >>= 1) v += __shfl_xor(v, o, 64);
    return v;
}
