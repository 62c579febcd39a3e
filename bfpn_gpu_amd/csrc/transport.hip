// transport.hip — explicit MUSCL predictor-corrector y/z transport (Eq.18).
//
// Reference: muscl_paper_slope, spatial_flux_divergence_value and the
// flux / predictor / corrector kernels of SpatialTransportSolver::solveEq18
// (src/spatial_transport.cu:23-161, 202-262): four launches and three
// scratch arrays per call, a division per slope.
//
// MI355X design: one launch, in place, each value read once and written once.
//  * A work-group owns one energy group g and a chunk of MW consecutive
//    directions m (the contiguous axis) for the whole (y, z) plane, so no
//    other work-group touches its data and the update can be written in place.
//  * It marches along z with rolling windows in LDS: 5 rows of psi, 5 rows of
//    the predictor psi* and 3 rows of L(psi).  At step t it loads psi row t+4,
//    forms L(psi) and psi* = clip(psi - h L psi) on row t+2, and writes row t:
//    clip(psi - h/2 (L psi + L psi*)).  Rows outside [0, Nz] are clamped,
//    exactly like the reference's clamped stencil taps.
//  * The superbee slope is evaluated without a division:
//      a = R - C, b = C - L, ab > 0:
//      slope = sign(a) max(min(2|b|, |a|), min(|b|, 2|a|)),   else 0,
//    which equals forward * max(0, min(2 theta,1), min(theta,2)) with
//    theta = b/a, up to the rounding of that quotient; |a| < 1e-10 -> 0 as in
//    the reference.
//  * MW adapts to the plane width (64 / 32 / 16 directions for <= 24 / 48 /
//    more y nodes) so every z step gives the 256 threads ~1000 cells.

#include "kernels.hpp"

namespace bfpn {
namespace {

constexpr int kThreads = 256;
constexpr int kPsiRows = 5, kPredRows = 5, kLRows = 3;

template <typename T>
__device__ inline T superbee_slope(T left, T center, T right) {
    const T a = right - center;
    const T b = center - left;
    if (fabs(a) < T(1e-10) || !(a * b > T(0))) return T(0);
    const T aa = fabs(a), ab = fabs(b);
    const T mag = fmax(fmin(T(2) * ab, aa), fmin(ab, T(2) * aa));
    return a > T(0) ? mag : -mag;
}

// Upwind MUSCL face-flux difference along one axis (w >= 0: faces use the
// left cells' reconstructions; w < 0: the right cells').
template <typename T>
__device__ inline T flux_difference(T w, T m2, T m1, T c, T p1, T p2) {
    const bool pos = w >= T(0);
    // left face value and right face value of the reconstruction
    const T sl = pos ? superbee_slope(m2, m1, c) : superbee_slope(m1, c, p1);
    const T sr = pos ? superbee_slope(m1, c, p1) : superbee_slope(c, p1, p2);
    const T left = pos ? m1 + T(0.5) * sl : c - T(0.5) * sl;
    const T right = pos ? c + T(0.5) * sr : p1 - T(0.5) * sr;
    return w * right - w * left;
}

// L(row r) for the cell (iy, mm) from a ring of rows in LDS.
template <typename T, int MW>
__device__ inline T divergence(const T* __restrict__ rows, int ring, int r, int Nz, int iy, int Ny,
                               int mm, T wy, T wz, T inv_dy, T inv_dz, int row_stride) {
    auto at = [&](int rr, int yy) -> T {
        rr = min(max(rr, 0), Nz);
        return rows[(rr % ring) * row_stride + yy * MW + mm];
    };
    const T c = at(r, iy);
    const T ym2 = at(r, max(iy - 2, 0)), ym1 = at(r, max(iy - 1, 0));
    const T yp1 = at(r, min(iy + 1, Ny)), yp2 = at(r, min(iy + 2, Ny));
    const T zm2 = at(r - 2, iy), zm1 = at(r - 1, iy), zp1 = at(r + 1, iy), zp2 = at(r + 2, iy);
    return flux_difference(wy, ym2, ym1, c, yp1, yp2) * inv_dy +
           flux_difference(wz, zm2, zm1, c, zp1, zp2) * inv_dz;
}

// grid: (ceil(nang/MW), Ng); block: 256 threads; dynamic LDS (13 rows).
template <typename T, int MW>
__global__ __launch_bounds__(kThreads) void transport_march_kernel(
    T* __restrict__ F, const T* __restrict__ omend, int Ny, int Nz, int nang, T inv_dy, T inv_dz,
    T h, int preserve_sign) {
    extern __shared__ unsigned char lds_raw[];
    const int Ny1 = Ny + 1;
    const int row_stride = Ny1 * MW;
    T* psi = reinterpret_cast<T*>(lds_raw);
    T* pred = psi + kPsiRows * row_stride;
    T* Lrow = pred + kPredRows * row_stride;
    const int m0 = blockIdx.x * MW;
    const long long plane = static_cast<long long>(Ny1) * (Nz + 1) * nang;
    T* Fg = F + static_cast<long long>(blockIdx.y) * plane;
    const long long zstride = static_cast<long long>(Ny1) * nang;

    // MW divides the block size, so every thread keeps one direction mm and
    // walks the y nodes iy0, iy0 + kThreads/MW, ...
    static_assert(kThreads % MW == 0, "MW must divide the block size");
    constexpr int kYStep = kThreads / MW;
    const int mm = threadIdx.x % MW;
    const int iy0 = threadIdx.x / MW;
    const bool live = m0 + mm < nang;
    const int m = live ? m0 + mm : nang - 1;
    const T wy = omend[2 * m], wz = omend[2 * m + 1];

    auto load_row = [&](int r) {  // psi row r (0 <= r <= Nz) into the ring
        T* dst = psi + (r % kPsiRows) * row_stride;
        const T* src = Fg + r * zstride + m0 + mm;
        for (int iy = iy0; iy < Ny1; iy += kYStep)
            dst[iy * MW + mm] = live ? src[static_cast<long long>(iy) * nang] : T(0);
    };

    for (int r = 0; r <= min(Nz, 1); ++r) load_row(r);
    for (int t = -2; t <= Nz; ++t) {
        if (t + 4 <= Nz && t + 4 >= 2) load_row(t + 4);
        __syncthreads();
        const int rp = t + 2;
        if (rp >= 0 && rp <= Nz) {  // L(psi) and psi* on row rp
            for (int iy = iy0; iy < Ny1; iy += kYStep) {
                const int k = iy * MW + mm;
                const T L0 = divergence<T, MW>(psi, kPsiRows, rp, Nz, iy, Ny, mm, wy, wz, inv_dy,
                                               inv_dz, row_stride);
                const T v = psi[(rp % kPsiRows) * row_stride + k] - h * L0;
                pred[(rp % kPredRows) * row_stride + k] = preserve_sign ? v : fmax(v, T(0));
                Lrow[(rp % kLRows) * row_stride + k] = L0;
            }
        }
        __syncthreads();
        if (t >= 0 && live) {  // corrector on row t, written in place
            T* dst = Fg + t * zstride + m0 + mm;
            for (int iy = iy0; iy < Ny1; iy += kYStep) {
                const int k = iy * MW + mm;
                const T L1 = divergence<T, MW>(pred, kPredRows, t, Nz, iy, Ny, mm, wy, wz, inv_dy,
                                               inv_dz, row_stride);
                const T L0 = Lrow[(t % kLRows) * row_stride + k];
                const T v = psi[(t % kPsiRows) * row_stride + k] - T(0.5) * h * (L0 + L1);
                dst[static_cast<long long>(iy) * nang] = preserve_sign ? v : fmax(v, T(0));
            }
        }
        __syncthreads();
    }
}

template <typename T, int MW>
void launch_march(T* F, const T* omend, int Ny, int Nz, int Ng, int nang, double dy, double dz,
                  double h, bool preserve_sign, hipStream_t s) {
    const size_t lds = sizeof(T) * static_cast<size_t>(kPsiRows + kPredRows + kLRows) * (Ny + 1) * MW;
    BFPN_REQUIRE(lds <= 160 * 1024, "transport plane too wide for the LDS window");
    BFPN_HIP(hipFuncSetAttribute(reinterpret_cast<const void*>(&transport_march_kernel<T, MW>),
                                 hipFuncAttributeMaxDynamicSharedMemorySize, static_cast<int>(lds)));
    dim3 grid((nang + MW - 1) / MW, Ng);
    hipLaunchKernelGGL((transport_march_kernel<T, MW>), grid, dim3(kThreads), lds, s, F, omend, Ny,
                       Nz, nang, static_cast<T>(1.0 / dy), static_cast<T>(1.0 / dz),
                       static_cast<T>(h), preserve_sign ? 1 : 0);
}

}  // namespace

template <typename T>
void launch_transport_half(T* F, const T* omend, int Ny, int Nz, int Ng, int nang, double dy,
                           double dz, double h, bool preserve_sign, hipStream_t s) {
    if (Ng <= 0 || nang <= 0) return;
    BFPN_REQUIRE(Ny >= 1 && Nz >= 1, "transport needs Ny, Nz >= 1");
    const int Ny1 = Ny + 1;
    if (Ny1 <= 24)
        launch_march<T, 64>(F, omend, Ny, Nz, Ng, nang, dy, dz, h, preserve_sign, s);
    else if (Ny1 <= 48)
        launch_march<T, 32>(F, omend, Ny, Nz, Ng, nang, dy, dz, h, preserve_sign, s);
    else if (Ny1 <= 81)
        launch_march<T, 16>(F, omend, Ny, Nz, Ng, nang, dy, dz, h, preserve_sign, s);
    else
        launch_march<T, 8>(F, omend, Ny, Nz, Ng, nang, dy, dz, h, preserve_sign, s);
    BFPN_HIP(hipGetLastError());
}

template void launch_transport_half<float>(float*, const float*, int, int, int, int, double, double,
                                           double, bool, hipStream_t);
template void launch_transport_half<double>(double*, const double*, int, int, int, int, double,
                                            double, double, bool, hipStream_t);

}  // namespace bfpn
