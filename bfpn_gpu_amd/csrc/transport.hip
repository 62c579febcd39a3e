// transport.hip — explicit MUSCL predictor-corrector y/z transport (Eq.18).
//
// Reference: muscl_paper_slope, spatial_flux_divergence_value and the
// flux / predictor / corrector kernels of SpatialTransportSolver::solveEq18
// (src/spatial_transport.cu:23-161, 202-262): four launches and three
// scratch arrays per call, three limited slopes per axis and cell, a division
// per slope.
//
// MI355X design: one launch, in place, each value read once and written once.
//  * A work-group owns one energy group g and MW consecutive directions m
//    (the contiguous axis) for the whole (y, z) plane, so no other work-group
//    reads or writes its data and the update can be written in place.
//  * Threads own y columns (P of them each) and march along z.  The z window
//    of psi, psi* and L(psi) stays in registers, and the limited z slopes are
//    carried from row to row, so each row costs one new z slope per stage.
//    y neighbours come from LDS rings of 4 rows of psi and 4 rows of psi*.
//    Step t commits psi row t+2 (prefetched into registers one step earlier),
//    forms L(psi) and psi* = clip(psi - h L psi) on row t, and writes row t-2:
//    clip(psi - h/2 (L psi + L psi*)).  One barrier per z step.
//  * The flow direction is fixed per lane, so only the upwind reconstruction
//    is formed, in the lane's upwind frame: per stage two limited slopes in y
//    (the cell's own and its upwind neighbour's) and one in z.
//  * The superbee slope is evaluated without a division:
//      a = R - C, b = C - L, ab > 0 and |a| >= 1e-10:
//      slope = sign(a) min(max(|a|, |b|), 2 min(|a|, |b|)),   else 0,
//    which equals forward * max(0, min(2 theta, 1), min(theta, 2)) with
//    theta = b / a, up to the rounding of that quotient.
//  * Rows and columns outside the plane are clamped, exactly like the
//    reference's clamped stencil taps: the slope of a boundary cell is 0 and
//    the cell outside reads the boundary value with slope 0.

#include "kernels.hpp"

namespace bfpn {
namespace {

constexpr int kMaxThreads = 1024;

template <typename T>
__device__ inline T limited_slope(T a, T b) {  // a: forward, b: backward difference
    const T aa = fabs(a), ab = fabs(b);
    const T mag = fmin(fmax(aa, ab), T(2) * fmin(aa, ab));
    return (aa < T(1e-10) || !(a * b > T(0))) ? T(0) : copysign(mag, a);
}

// Limited slope from two differences of one orientation with the 1e-10
// test on `e` (the difference that is forward in the physical axis).  With
// a*b > 0 the sign of a is the sign of b, so the value is symmetric in (a, b).
template <typename T>
__device__ inline T upwind_slope(T a, T b, T e) {
    const T aa = fabs(a), ab = fabs(b);
    const T mag = fmin(fmax(aa, ab), T(2) * fmin(aa, ab));
    return (fabs(e) < T(1e-10) || !(a * b > T(0))) ? T(0) : copysign(mag, a);
}

// Flux divergence at one cell, per axis written in the lane's upwind frame:
//   w (F(+1/2) - F(-1/2)) = |w| (downwind face - upwind face),
// where the downwind face is the cell's own reconstruction and the upwind
// face the upwind neighbour's.  y: near = upwind neighbour, far = the one
// beyond it, down = downwind neighbour (LDS row); with A = down - c,
// B = c - near, C = near - far the two slopes are (A, B) and (B, C), the
// 1e-10 test going to whichever is forward in +y.  z: physical window
// (zm, zc, zp) and carried physical slopes of rows -1, 0, +1; hz = +-1/2
// with the sign of wz.
template <typename T>
__device__ inline T divergence(const T* row, int o_near, int o_far, int o_down, T zm, T zc, T zp,
                               T sz_m, T sz_0, T sz_p, T awy, T awz, T hz, bool py, bool pz,
                               T inv_dy, T inv_dz) {
    const T near = row[o_near], far = row[o_far], down = row[o_down];
    const T A = down - zc, B = zc - near, C = near - far;
    const T s_down = upwind_slope(A, B, py ? A : B);
    const T s_up = upwind_slope(B, C, py ? B : C);
    const T fy = awy * fma(T(0.5), s_down, zc) - awy * fma(T(0.5), s_up, near);
    const T fz = awz * fma(hz, sz_0, zc) - awz * fma(hz, pz ? sz_m : sz_p, pz ? zm : zp);
    return fy * inv_dy + fz * inv_dz;
}

// grid: (ceil(nang / MW), Ng), block: MW * R threads (R y columns per pass,
// P passes: the thread's columns are y = ry + j R).  LDS: 8 rows of
// (Ny+1) * MW values (two rings of 4 rows).
template <typename T, int MW, int P, bool CLIP>
__global__ __launch_bounds__(kMaxThreads) void transport_march_kernel(
    T* __restrict__ F, const T* __restrict__ omend, int Ny, int Nz, int nang, int R, T inv_dy,
    T inv_dz, T h, int tiles) {
    extern __shared__ unsigned char lds_raw[];
    // Blocks are dispatched round-robin over the 8 XCDs.  Direction chunks of
    // one group share 128-byte lines (a plane row is nang floats), so the
    // chunks of a group run consecutively on one XCD and share its L2:
    // block b -> logical tile (b mod 8) * (grid / 8) + b / 8.
    const int L = (blockIdx.x & 7) * (gridDim.x >> 3) + (blockIdx.x >> 3);
    if (L >= tiles) return;
    const int mchunks = (nang + MW - 1) / MW;
    const int g = L / mchunks;
    const int Ny1 = Ny + 1;
    const int rs = Ny1 * MW;  // LDS row stride (elements)
    T* psi_ring = reinterpret_cast<T*>(lds_raw);  // 4 rows of psi
    T* pred_ring = psi_ring + 4 * rs;             // 4 rows of psi*

    const int mm = threadIdx.x % MW, ry = threadIdx.x / MW;
    const int m0 = (L - g * mchunks) * MW;
    const bool live = m0 + mm < nang;
    const int m = live ? m0 + mm : nang - 1;
    const T wy = omend[2 * m], wz = omend[2 * m + 1];
    const bool py = wy >= T(0), pz = wz >= T(0);
    const T awy = fabs(wy), awz = fabs(wz), hz = pz ? T(0.5) : T(-0.5);
    const long long zstride = static_cast<long long>(Ny1) * nang;
    T* Fg = F + static_cast<long long>(g) * zstride * (Nz + 1) + m0 + mm;

    bool on[P];
    int goff[P], o0[P], o_near[P], o_far[P], o_down[P];
    T wpsi[P][5];   // psi rows t-2 .. t+2 (after the commit of step t)
    T wpred[P][4];  // psi* rows t-3 .. t (after the predictor of step t)
    T wL[P][3];     // L(psi) rows t-2 .. t
    T sz[P][2];     // z slopes of psi, rows t-1 and t (before step t)
    T szp[P][2];    // z slopes of psi*, rows t-3 and t-2 (before step t)
    T nxt[P];
#pragma unroll
    for (int j = 0; j < P; ++j) {
        const int y = ry + j * R;
        on[j] = y < Ny1;
        const int yc = min(y, Ny);
        goff[j] = yc * nang;
        auto lo = [&](int yy) { return (min(max(yy, 0), Ny)) * MW + mm; };
        o0[j] = lo(yc);
        o_near[j] = lo(py ? yc - 1 : yc + 1);
        o_far[j] = lo(py ? yc - 2 : yc + 2);
        o_down[j] = lo(py ? yc + 1 : yc - 1);
        const bool ld = on[j] && live;
        const T r0 = ld ? Fg[goff[j]] : T(0);
        const T r1 = ld ? Fg[zstride + goff[j]] : T(0);  // Nz >= 1
        if (on[j]) {
            psi_ring[o0[j]] = r0;
            psi_ring[rs + o0[j]] = r1;
        }
        wpsi[j][0] = wpsi[j][1] = wpsi[j][2] = wpsi[j][3] = r0;
        wpsi[j][4] = r1;
        nxt[j] = (ld && Nz >= 2) ? Fg[2 * zstride + goff[j]] : T(0);
        sz[j][0] = sz[j][1] = T(0);  // boundary slopes are 0 (clamped taps)
        szp[j][0] = szp[j][1] = T(0);
#pragma unroll
        for (int k = 0; k < 4; ++k) wpred[j][k] = T(0);
#pragma unroll
        for (int k = 0; k < 3; ++k) wL[j][k] = T(0);
    }

    for (int t = 0; t <= Nz + 2; ++t) {
        // commit psi row t+2 (clamped past Nz)
        const int rc = t + 2;
        T* crow = psi_ring + (rc & 3) * rs;
#pragma unroll
        for (int j = 0; j < P; ++j) {
            const T v = rc <= Nz ? nxt[j] : wpsi[j][4];
#pragma unroll
            for (int k = 0; k < 4; ++k) wpsi[j][k] = wpsi[j][k + 1];
            wpsi[j][4] = v;
            if (rc <= Nz && on[j]) crow[o0[j]] = v;
        }
        __syncthreads();
        if (t + 3 <= Nz) {
#pragma unroll
            for (int j = 0; j < P; ++j)
                nxt[j] = (on[j] && live) ? Fg[(t + 3) * zstride + goff[j]] : T(0);
        }
        if (t <= Nz) {  // predictor on row t
            const T* row = psi_ring + (t & 3) * rs;
            T* prow = pred_ring + (t & 3) * rs;
#pragma unroll
            for (int j = 0; j < P; ++j) {
                const T zm = wpsi[j][1], zc = wpsi[j][2], zp = wpsi[j][3];
                const T s_next = limited_slope(wpsi[j][4] - zp, zp - zc);
                T v = T(0), L0 = T(0);
                if (on[j]) {
                    L0 = divergence(row, o_near[j], o_far[j], o_down[j], zm, zc, zp, sz[j][0],
                                    sz[j][1], s_next, awy, awz, hz, py, pz, inv_dy, inv_dz);
                    v = zc - h * L0;
                    if (CLIP) v = fmax(v, T(0));
                    prow[o0[j]] = v;
                }
                sz[j][0] = sz[j][1];
                sz[j][1] = s_next;
                if (t == 0) {
#pragma unroll
                    for (int k = 0; k < 4; ++k) wpred[j][k] = v;  // rows < 0 clamp to row 0
                } else {
#pragma unroll
                    for (int k = 0; k < 3; ++k) wpred[j][k] = wpred[j][k + 1];
                    wpred[j][3] = v;
                }
                wL[j][0] = wL[j][1];
                wL[j][1] = wL[j][2];
                wL[j][2] = L0;
            }
        } else {  // rows past Nz clamp to row Nz
#pragma unroll
            for (int j = 0; j < P; ++j) {
#pragma unroll
                for (int k = 0; k < 3; ++k) wpred[j][k] = wpred[j][k + 1];
                wL[j][0] = wL[j][1];
                wL[j][1] = wL[j][2];
            }
        }
        if (t >= 2) {  // corrector on row t-2, written in place
            const int ro = t - 2;
            const T* row = pred_ring + (ro & 3) * rs;
            T* dst = Fg + ro * zstride;
#pragma unroll
            for (int j = 0; j < P; ++j) {
                const T zm = wpred[j][0], zc = wpred[j][1], zp = wpred[j][2];
                const T s_next = limited_slope(wpred[j][3] - zp, zp - zc);
                if (on[j]) {
                    const T L1 = divergence(row, o_near[j], o_far[j], o_down[j], zm, zc, zp,
                                            szp[j][0], szp[j][1], s_next, awy, awz, hz, py,
                                            pz, inv_dy, inv_dz);
                    T v = wpsi[j][0] - T(0.5) * h * (wL[j][0] + L1);
                    if (CLIP) v = fmax(v, T(0));
                    if (live) dst[goff[j]] = v;
                }
                szp[j][0] = szp[j][1];
                szp[j][1] = s_next;
            }
        }
    }
}

struct MarchShape {
    int MW, P, R, threads;
};

// Widest direction chunk (coalesced rows) whose plane fits one work-group
// with at most 2 (MW = 32) or 4 (MW = 16, 8) columns per thread and whose two
// LDS rings fit the 160 KB of a CU.
MarchShape march_shape(int Ny1, size_t elem) {
    for (int MW : {32, 16, 8}) {
        const int rows = kMaxThreads / MW;
        const int P = (Ny1 + rows - 1) / rows;
        if (P > (MW == 32 ? 2 : 4)) continue;
        if (8 * static_cast<size_t>(Ny1) * MW * elem > 160 * 1024) continue;
        const int R0 = (Ny1 + P - 1) / P;
        const int threads = (MW * R0 + 63) / 64 * 64;
        return MarchShape{MW, P, threads / MW, threads};
    }
    return MarchShape{0, 0, 0, 0};
}

template <typename T, int MW, int P>
void launch_march(T* F, const T* omend, int Ny, int Nz, int Ng, int nang, const MarchShape& sh,
                  double dy, double dz, double h, bool preserve_sign, hipStream_t s) {
    const size_t lds = sizeof(T) * 8 * static_cast<size_t>(Ny + 1) * MW;
    BFPN_REQUIRE(lds <= 160 * 1024, "transport plane too wide for the LDS rings");
    const long long tiles = static_cast<long long>((nang + MW - 1) / MW) * Ng;
    BFPN_REQUIRE(tiles < (1LL << 31) - 8, "transport grid too large");
    const dim3 grid(static_cast<unsigned>((tiles + 7) / 8 * 8));
    auto go = [&](auto kernel) {
        BFPN_HIP(hipFuncSetAttribute(reinterpret_cast<const void*>(kernel),
                                     hipFuncAttributeMaxDynamicSharedMemorySize,
                                     static_cast<int>(lds)));
        hipLaunchKernelGGL(kernel, grid, dim3(sh.threads), lds, s, F, omend, Ny, Nz, nang, sh.R,
                           static_cast<T>(1.0 / dy), static_cast<T>(1.0 / dz), static_cast<T>(h),
                           static_cast<int>(tiles));
    };
    if (preserve_sign)
        go(&transport_march_kernel<T, MW, P, false>);
    else
        go(&transport_march_kernel<T, MW, P, true>);
    BFPN_HIP(hipGetLastError());
}

}  // namespace

template <typename T>
void launch_transport_half(T* F, const T* omend, int Ny, int Nz, int Ng, int nang, double dy,
                           double dz, double h, bool preserve_sign, hipStream_t s) {
    if (Ng <= 0 || nang <= 0) return;
    BFPN_REQUIRE(Ny >= 1 && Nz >= 1, "transport needs Ny, Nz >= 1");
    const MarchShape sh = march_shape(Ny + 1, sizeof(T));
    BFPN_REQUIRE(sh.MW > 0, sizeof(T) == 4 ? "transport supports at most 512 y nodes in FP32"
                                           : "transport supports at most 320 y nodes in FP64");
#define BFPN_MARCH(MW_, P_)                                                                    \
    if (sh.MW == MW_ && sh.P == P_)                                                            \
        return launch_march<T, MW_, P_>(F, omend, Ny, Nz, Ng, nang, sh, dy, dz, h, preserve_sign, \
                                        s)
    BFPN_MARCH(32, 1);
    BFPN_MARCH(32, 2);
    BFPN_MARCH(16, 2);
    BFPN_MARCH(16, 3);
    BFPN_MARCH(16, 4);
    BFPN_MARCH(8, 2);
    BFPN_MARCH(8, 3);
    BFPN_MARCH(8, 4);
#undef BFPN_MARCH
    throw Error(BFPN_E_INVALID, "no transport launch shape for this plane");
}

template void launch_transport_half<float>(float*, const float*, int, int, int, int, double, double,
                                           double, bool, hipStream_t);
template void launch_transport_half<double>(double*, const double*, int, int, int, int, double,
                                            double, double, bool, hipStream_t);

}  // namespace bfpn
