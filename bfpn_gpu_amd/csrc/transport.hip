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
//  * It marches along z with rolling windows in LDS: 7 rows of psi, 6 rows of
//    the predictor psi* and 4 rows of L(psi).  At step t it commits psi row
//    t+4 (prefetched into registers one step earlier), forms L(psi) and
//    psi* = clip(psi - h L psi) on row t+2, and writes row t-1:
//    clip(psi - h/2 (L psi + L psi*)).  Rows outside [0, Nz] are clamped,
//    exactly like the reference's clamped stencil taps; one barrier per step.
//  * Planes wider than 24 nodes are cut into y strips of <= 40 nodes that
//    recompute a 4-node halo, so the LDS window stays ~50 KB (3 groups/CU).
//  * The superbee slope is evaluated without a division:
//      a = R - C, b = C - L, ab > 0:
//      slope = sign(a) max(min(2|b|, |a|), min(|b|, 2|a|)),   else 0,
//    which equals forward * max(0, min(2 theta,1), min(theta,2)) with
//    theta = b/a, up to the rounding of that quotient; |a| < 1e-10 -> 0 as in
//    the reference.
//  * MW = 32 directions (128-byte rows) for planes up to 24 nodes wide, 16
//    for strips, so each z step gives the 256 threads 500-800 cells.

#include "kernels.hpp"

namespace bfpn {
namespace {

constexpr int kThreads = 256;
// Ring depths (rows): psi 7, psi* 6, L(psi) 4.  With these depths the only
// LDS write issued before a step's barrier (the next psi row) can never alias
// a row that a slower thread of the previous step still reads, so one barrier
// per z step suffices.
constexpr int kPsiRing = 7, kPredRing = 6, kLRing = 4;
constexpr int kRows = kPsiRing + kPredRing + kLRing;
constexpr int kMaxPerThread = 3;  // y nodes per thread and row (W*MW <= 768)

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
    const T sl = pos ? superbee_slope(m2, m1, c) : superbee_slope(m1, c, p1);
    const T sr = pos ? superbee_slope(m1, c, p1) : superbee_slope(c, p1, p2);
    const T left = pos ? m1 + T(0.5) * sl : c - T(0.5) * sl;
    const T right = pos ? c + T(0.5) * sr : p1 - T(0.5) * sr;
    return w * right - w * left;
}

// L at one cell: z taps from five row bases, y taps at offsets within row r0.
template <typename T>
__device__ inline T divergence(const T* zm2, const T* zm1, const T* r0, const T* zp1,
                               const T* zp2, int c, int ym2, int ym1, int yp1, int yp2, T wy, T wz,
                               T inv_dy, T inv_dz) {
    const T v = r0[c];
    return flux_difference(wy, r0[ym2], r0[ym1], v, r0[yp1], r0[yp2]) * inv_dy +
           flux_difference(wz, zm2[c], zm1[c], v, zp1[c], zp2[c]) * inv_dz;
}

// grid: (strips * ceil(nang/MW), Ng).  A work-group owns one energy group, MW
// consecutive directions and the y strip [y0, y0+TY) of the plane (with a
// 4-node halo it recomputes), and marches along z: at step t it commits psi
// row t+4 (prefetched into registers during the previous step), forms L(psi)
// and psi* on row t+2 and writes row t-1 in place.
template <typename T, int MW>
__global__ __launch_bounds__(kThreads) void transport_march_kernel(
    T* __restrict__ F, const T* __restrict__ omend, int Ny, int Nz, int nang, int TY, int mchunks,
    T inv_dy, T inv_dz, T h, int preserve_sign) {
    extern __shared__ unsigned char lds_raw[];
    static_assert(kThreads % MW == 0, "MW must divide the block size");
    constexpr int kYStep = kThreads / MW;
    const int Ny1 = Ny + 1;
    const int strip = blockIdx.x / mchunks;
    const int m0 = (blockIdx.x - strip * mchunks) * MW;
    const int y0 = strip * TY, y1 = min(Ny1, y0 + TY);  // rows written
    const int ylo = max(0, y0 - 4), yhi = min(Ny1, y1 + 4);  // rows held
    const int plo = max(0, y0 - 2), phi = min(Ny1, y1 + 2);  // rows predicted
    const int W = yhi - ylo;
    const int rs = W * MW;  // LDS row stride
    T* psi = reinterpret_cast<T*>(lds_raw);
    T* pred = psi + kPsiRing * rs;
    T* Lr = pred + kPredRing * rs;

    const int mm = threadIdx.x % MW;
    const int iy0 = threadIdx.x / MW;
    const bool live = m0 + mm < nang;
    const int m = live ? m0 + mm : nang - 1;
    const T wy = omend[2 * m], wz = omend[2 * m + 1];
    const long long zstride = static_cast<long long>(Ny1) * nang;
    T* Fg = F + static_cast<long long>(blockIdx.y) * zstride * (Nz + 1) + m0 + mm;

    T nxt[kMaxPerThread];
    auto fetch = [&](int r) {
#pragma unroll
        for (int j = 0; j < kMaxPerThread; ++j) {
            const int iy = ylo + iy0 + j * kYStep;
            nxt[j] = (live && iy < yhi) ? Fg[r * zstride + static_cast<long long>(iy) * nang] : T(0);
        }
    };
    auto commit = [&](int r) {
        T* dst = psi + (r % kPsiRing) * rs;
#pragma unroll
        for (int j = 0; j < kMaxPerThread; ++j) {
            const int ly = iy0 + j * kYStep;
            if (ylo + ly < yhi) dst[ly * MW + mm] = nxt[j];
        }
    };
    auto clampz = [&](int r) { return min(max(r, 0), Nz); };
    auto yoff = [&](int iy) { return (min(max(iy, 0), Ny) - ylo) * MW + mm; };

    for (int r = 0; r <= min(Nz, 2); ++r) {
        fetch(r);
        commit(r);
    }
    for (int t = -2; t <= Nz + 1; ++t) {
        if (t + 4 >= 3 && t + 4 <= Nz) commit(t + 4);
        __syncthreads();
        if (t + 5 >= 3 && t + 5 <= Nz) fetch(t + 5);
        const int rp = t + 2;
        if (rp >= 0 && rp <= Nz) {  // L(psi) and psi* on row rp
            const T* b0 = psi + (rp % kPsiRing) * rs;
            const T* bm2 = psi + (clampz(rp - 2) % kPsiRing) * rs;
            const T* bm1 = psi + (clampz(rp - 1) % kPsiRing) * rs;
            const T* bp1 = psi + (clampz(rp + 1) % kPsiRing) * rs;
            const T* bp2 = psi + (clampz(rp + 2) % kPsiRing) * rs;
            T* pdst = pred + (rp % kPredRing) * rs;
            T* ldst = Lr + (rp % kLRing) * rs;
            for (int iy = plo + iy0; iy < phi; iy += kYStep) {
                const int c = yoff(iy);
                const T L0 = divergence(bm2, bm1, b0, bp1, bp2, c, yoff(iy - 2), yoff(iy - 1),
                                        yoff(iy + 1), yoff(iy + 2), wy, wz, inv_dy, inv_dz);
                const T v = b0[c] - h * L0;
                pdst[c] = preserve_sign ? v : fmax(v, T(0));
                ldst[c] = L0;
            }
        }
        const int ro = t - 1;
        if (ro >= 0 && ro <= Nz && live) {  // corrector on row ro, in place
            const T* b0 = pred + (ro % kPredRing) * rs;
            const T* bm2 = pred + (clampz(ro - 2) % kPredRing) * rs;
            const T* bm1 = pred + (clampz(ro - 1) % kPredRing) * rs;
            const T* bp1 = pred + (clampz(ro + 1) % kPredRing) * rs;
            const T* bp2 = pred + (clampz(ro + 2) % kPredRing) * rs;
            const T* lrow = Lr + (ro % kLRing) * rs;
            const T* prow = psi + (ro % kPsiRing) * rs;
            T* dst = Fg + ro * zstride;
            for (int iy = y0 + iy0; iy < y1; iy += kYStep) {
                const int c = yoff(iy);
                const T L1 = divergence(bm2, bm1, b0, bp1, bp2, c, yoff(iy - 2), yoff(iy - 1),
                                        yoff(iy + 1), yoff(iy + 2), wy, wz, inv_dy, inv_dz);
                const T v = prow[c] - T(0.5) * h * (lrow[c] + L1);
                dst[static_cast<long long>(iy) * nang] = preserve_sign ? v : fmax(v, T(0));
            }
        }
    }
}

template <typename T, int MW>
void launch_march(T* F, const T* omend, int Ny, int Nz, int Ng, int nang, int TY, double dy,
                  double dz, double h, bool preserve_sign, hipStream_t s) {
    const int Ny1 = Ny + 1;
    const int strips = (Ny1 + TY - 1) / TY;
    const int W = std::min(Ny1, TY + 8);
    BFPN_REQUIRE((W + kThreads / MW - 1) / (kThreads / MW) <= kMaxPerThread,
                 "transport strip too wide for the per-thread prefetch");
    const size_t lds = sizeof(T) * static_cast<size_t>(kRows) * W * MW;
    BFPN_REQUIRE(lds <= 160 * 1024, "transport strip too wide for the LDS window");
    BFPN_HIP(hipFuncSetAttribute(reinterpret_cast<const void*>(&transport_march_kernel<T, MW>),
                                 hipFuncAttributeMaxDynamicSharedMemorySize, static_cast<int>(lds)));
    const int mchunks = (nang + MW - 1) / MW;
    dim3 grid(strips * mchunks, Ng);
    hipLaunchKernelGGL((transport_march_kernel<T, MW>), grid, dim3(kThreads), lds, s, F, omend, Ny,
                       Nz, nang, TY, mchunks, static_cast<T>(1.0 / dy), static_cast<T>(1.0 / dz),
                       static_cast<T>(h), preserve_sign ? 1 : 0);
}

}  // namespace

template <typename T>
void launch_transport_half(T* F, const T* omend, int Ny, int Nz, int Ng, int nang, double dy,
                           double dz, double h, bool preserve_sign, hipStream_t s) {
    if (Ng <= 0 || nang <= 0) return;
    BFPN_REQUIRE(Ny >= 1 && Nz >= 1, "transport needs Ny, Nz >= 1");
    const int Ny1 = Ny + 1;
    if (Ny1 <= 24) {  // whole plane width, 32 directions per work-group
        launch_march<T, 32>(F, omend, Ny, Nz, Ng, nang, Ny1, dy, dz, h, preserve_sign, s);
    } else {  // strips of <= 40 interior nodes (+4 halo each side), 16 directions
        const int strips = (Ny1 + 39) / 40;
        const int TY = (Ny1 + strips - 1) / strips;
        launch_march<T, 16>(F, omend, Ny, Nz, Ng, nang, TY, dy, dz, h, preserve_sign, s);
    }
    BFPN_HIP(hipGetLastError());
}

template void launch_transport_half<float>(float*, const float*, int, int, int, int, double, double,
                                           double, bool, hipStream_t);
template void launch_transport_half<double>(double*, const double*, int, int, int, int, double,
                                            double, double, bool, hipStream_t);

}  // namespace bfpn
