// transport.hip — explicit MUSCL predictor-corrector y/z transport (Eq.18).
//
// Reference: muscl_paper_slope, spatial_flux_divergence_value and the
// flux / predictor / corrector kernels of SpatialTransportSolver::solveEq18
// (src/spatial_transport.cu:23-161, 202-262): four launches and three
// scratch arrays per call.
//
// MI355X design: two launches.  The predictor writes psi* = clip(psi - h L psi);
// the corrector recomputes L psi from the input instead of storing it and
// writes psi_new = clip(psi - h/2 (L psi + L psi*)) to a separate buffer, so no
// flux array ever touches HBM.  Threads run along the direction index m,
// which is the contiguous axis of the [g][s][m] layout, so every stencil tap
// (+-1, +-2 nodes in y or z) is a coalesced row read served mostly from L2.

#include "kernels.hpp"

namespace bfpn {
namespace {

template <typename T>
__device__ inline T limited_slope(T left, T center, T right) {
    const T forward = right - center;
    if (fabs(forward) < T(1e-10)) return T(0);
    const T theta = (center - left) / forward;
    const T eta = fmax(T(0), fmax(fmin(T(2) * theta, T(1)), fmin(theta, T(2))));
    return forward * eta;
}

// Upwind MUSCL face-flux difference along one axis for velocity w, given the
// five taps (clamped at the domain edge) around the centre c.
template <typename T>
__device__ inline T axis_flux_difference(T w, T m2, T m1, T c, T p1, T p2) {
    if (w >= T(0)) {
        const T left = w * (m1 + T(0.5) * limited_slope(m2, m1, c));
        const T right = w * (c + T(0.5) * limited_slope(m1, c, p1));
        return right - left;
    }
    const T left = w * (c - T(0.5) * limited_slope(m1, c, p1));
    const T right = w * (p1 - T(0.5) * limited_slope(c, p1, p2));
    return right - left;
}

template <typename T>
__device__ inline T flux_divergence(const T* __restrict__ in, long long idx, int i, int j, int Ny,
                                    int Nz, long long sy, long long sz, T wy, T wz, T dy,
                                    T dz) {
    const T c = in[idx];
    const T ym2 = in[idx + (max(0, i - 2) - i) * sy];
    const T ym1 = in[idx + (max(0, i - 1) - i) * sy];
    const T yp1 = in[idx + (min(Ny, i + 1) - i) * sy];
    const T yp2 = in[idx + (min(Ny, i + 2) - i) * sy];
    const T zm2 = in[idx + (max(0, j - 2) - j) * sz];
    const T zm1 = in[idx + (max(0, j - 1) - j) * sz];
    const T zp1 = in[idx + (min(Nz, j + 1) - j) * sz];
    const T zp2 = in[idx + (min(Nz, j + 2) - j) * sz];
    return axis_flux_difference(wy, ym2, ym1, c, yp1, yp2) / dy +
           axis_flux_difference(wz, zm2, zm1, c, zp1, zp2) / dz;
}

// grid: (ceil(nyz*nang/256), Ng)
template <typename T, bool CORRECT>
__global__ __launch_bounds__(256) void transport_kernel(const T* __restrict__ in,
                                                        const T* __restrict__ pred,
                                                        T* __restrict__ out,
                                                        const T* __restrict__ omend, int Ny,
                                                        int Nz, int nang, T dy, T dz, T h,
                                                        int preserve_sign) {
    const int Ny1 = Ny + 1;
    const int plane = Ny1 * (Nz + 1) * nang;
    const int r = blockIdx.x * 256 + threadIdx.x;
    if (r >= plane) return;
    const int m = r % nang;
    const int s = r / nang;
    const int i = s % Ny1;
    const int j = s / Ny1;
    const long long idx = static_cast<long long>(blockIdx.y) * plane + r;
    const long long sy = nang;
    const long long sz = static_cast<long long>(Ny1) * nang;
    const T wy = omend[2 * m], wz = omend[2 * m + 1];
    const T L0 = flux_divergence(in, idx, i, j, Ny, Nz, sy, sz, wy, wz, dy, dz);
    T v;
    if (CORRECT) {
        const T L1 = flux_divergence(pred, idx, i, j, Ny, Nz, sy, sz, wy, wz, dy, dz);
        v = in[idx] - T(0.5) * h * (L0 + L1);
    } else {
        v = in[idx] - h * L0;
    }
    out[idx] = preserve_sign ? v : fmax(v, T(0));
}

}  // namespace

template <typename T>
void launch_transport_half(const T* in, T* pred, T* out, const T* omend, int Ny, int Nz, int Ng,
                           int nang, double dy, double dz, double h, bool preserve_sign,
                           hipStream_t s) {
    const long long plane = static_cast<long long>(Ny + 1) * (Nz + 1) * nang;
    BFPN_REQUIRE(plane < (1LL << 31), "transport plane too large");
    if (plane == 0 || Ng <= 0) return;
    dim3 grid(static_cast<unsigned>((plane + 255) / 256), Ng);
    const T ty = static_cast<T>(dy), tz = static_cast<T>(dz);
    hipLaunchKernelGGL((transport_kernel<T, false>), grid, dim3(256), 0, s, in, nullptr, pred, omend,
                       Ny, Nz, nang, ty, tz, static_cast<T>(h), preserve_sign ? 1 : 0);
    hipLaunchKernelGGL((transport_kernel<T, true>), grid, dim3(256), 0, s, in, pred, out, omend, Ny,
                       Nz, nang, ty, tz, static_cast<T>(h), preserve_sign ? 1 : 0);
    BFPN_HIP(hipGetLastError());
}

template void launch_transport_half<float>(const float*, float*, float*, const float*, int, int,
                                           int, int, double, double, double, bool, hipStream_t);
template void launch_transport_half<double>(const double*, double*, double*, const double*, int,
                                            int, int, int, double, double, double, bool,
                                            hipStream_t);

}  // namespace bfpn
