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
//    theta = b / a, up to the rounding of that quotient; it is computed as
//    two signed medians of three (one instruction each in FP32).
//  * Rows and columns outside the plane are clamped, exactly like the
//    reference's clamped stencil taps: the slope of a boundary cell is 0 and
//    the cell outside reads the boundary value with slope 0.

#include "buffer.hpp"
#include "kernels.hpp"

#include <type_traits>

namespace bfpn {
namespace {

// Work-group size limit: 1024 threads in FP32; 768 in FP64, so that a
// thread may hold up to 168 VGPRs (its windows are twice as wide).
template <typename T>
constexpr int max_threads() { return sizeof(T) == 4 ? 1024 : 768; }

// LDS slots per y column (14 used, one pad: an odd pitch in units of MW)
constexpr int kSlots = 15;

// Superbee slope of two differences (either order), without the 1e-10 test:
//   ab > 0: sign(a) min(max(|a|, |b|), 2 min(|a|, |b|)),  else 0,
// which is the median of a, b and 2 m, with m = median(a, b, 0) the minmod
// of a and b: m = 0 when the signs differ (then the median is 0 too), and for
// a common sign the median of a, b, 2m is the expression above.  One
// median-of-three instruction each in FP32; FP64 uses
// median(x, y, z) = max(min(x, y), min(max(x, y), z)).
__device__ inline float superbee(float a, float b) {
    const float m = __builtin_amdgcn_fmed3f(a, b, 0.0f);
    return __builtin_amdgcn_fmed3f(a, b, 2.0f * m);
}
__device__ inline double superbee(double a, double b) {
    const double lo = fmin(a, b), hi = fmax(a, b);
    const double m = fmax(lo, fmin(hi, 0.0));
    return fmax(lo, fmin(hi, 2.0 * m));
}

// a: forward, b: backward difference (physical orientation); 0 when the
// forward difference is below 1e-10.
template <typename T>
__device__ inline T limited_slope(T a, T b) {
    return fabs(a) < T(1e-10) ? T(0) : superbee(a, b);
}

// z flux term of a cell in the lane's upwind frame (see divergence):
// hz (S_c - S_up) + (c - z_up) from the backward differences for wz >= 0 and
// the negated forward ones otherwise.
template <typename T>
__device__ inline T z_term(T dz_b, T dz_f, T dS_b, T dS_f, T hz, bool pz) {
    return fma(hz, pz ? dS_b : -dS_f, pz ? dz_b : -dz_f);
}

// Flux divergence at one cell, per axis written in the lane's upwind frame:
//   w (F(+1/2) - F(-1/2)) / d = (|w| / d) (downwind face - upwind face),
// where the downwind face is the cell's own reconstruction c + s_down/2 and
// the upwind face the upwind neighbour's, n + s_up/2, so the difference is
// (c - n) + (s_down - s_up)/2.  y: from the physical limited slopes of the
// cell (S_c) and of its upwind neighbour (S_n), the upwind frame's slopes are
// +-S (superbee is odd and symmetric, and the 1e-10 test goes to the
// physical forward difference either way), so the y term is
// (c - n) + hy (S_c - S_n) with hy = +-1/2 the sign of wy.  z: the same with
// the row window (zm, zc, zp) and the physical slopes of rows -1, 0, +1,
// given as differences: dz_b = zc - zm, dz_f = zp - zc, dS_b = sz_0 - sz_m,
// dS_f = sz_p - sz_0 (each is also the other row's difference, so the march
// forms it once); the upwind terms zc - z_up and sz_0 - sz_up are dz_b, dS_b
// for wz >= 0 and -dz_f, -dS_f (exact negations) otherwise.  hz = +-1/2 with
// the sign of wz; cy = |wy| / dy, cz = |wz| / dz.
// The z term of a row for wz < 0, fma(hz, -dS_f, -dz_f), is the negation of
// fma(hz, dS_b, dz_b) of the next row (its backward differences are this
// row's forward ones; rounding to nearest is odd), so the march forms
// U = fma(hz, dS_f, dz_f) once per row, for the next row's wz >= 0 lanes and
// this row's other lanes, and selects between U(row) and -U(row + 1).
template <typename T>
__device__ inline T divergence(T zc, T near, T S_c, T S_n, T fz, T cy, T cz, T hy) {
    const T fy = fma(hy, S_c - S_n, zc - near);
    return cy * fy + cz * fz;
}


// grid: (ceil(nang / MW), Ng), block: MW * R threads (R y columns per pass,
// P passes: the thread's columns are y = ry + j R).  LDS, stored y-major
// (column y holds its slots of MW values side by side, pitch 15 MW, odd, so
// the y columns of one wave fall on distinct banks; a slot is then a constant
// offset from the column):
//   slots 0-3   psi rows (row mod 4)        slots 8-11  y slopes of psi rows
//   slots 4-7   psi* rows (row mod 4)       slots 12-13 y slopes of psi* rows
// The physical y slope of a row is computed once per cell, two steps before
// the stage that needs it (after the barrier that publishes the row), kept
// in a register by its own lane and published in LDS for the downwind
// neighbour, which reads it instead of recomputing it.
//
// The z march is unrolled by 8: the psi window and the rows in flight share
// one register ring indexed by row mod 8 (row t+6 is loaded into the slot of
// row t-2, which step t no longer reads), the other windows are indexed by
// row mod 4 (slopes by row mod 2), so with the step's phase a compile-time
// constant, sliding a window is a renaming rather than a chain of moves, a
// loaded row is never copied (a copy would wait for the load), and the LDS
// ring slots are immediate offsets.  The corrector needs psi and L(psi) of
// its row only as psi - h/2 L(psi), kept per row.
template <typename T, int MW, int P, bool CLIP>
__global__ __launch_bounds__(max_threads<T>()) void transport_march_kernel(
    T* __restrict__ F, const T* __restrict__ omend, int Ny, int Nz, int nang, int R, T inv_dy,
    T inv_dz, T h, int tiles, const int* __restrict__ top) {
    extern __shared__ unsigned char lds_raw[];
    const int mchunks = (nang + MW - 1) / MW;
    // Only the groups up to the state's top hold nonzero values (the others
    // stay zero); their tiles are the first `active` and the grid's later
    // blocks exit at once.  Blocks are dispatched round-robin over the 8
    // XCDs.  Direction chunks of one group share 128-byte lines (a plane row
    // is nang floats), so the chunks of a group run consecutively on one XCD
    // and share its L2: block b -> logical tile (b mod 8) * (n8 / 8) + b / 8,
    // n8 = active rounded up to a multiple of 8.
    const int active = top ? min(tiles, (min(*top, tiles / mchunks - 1) + 1) * mchunks) : tiles;
    const int n8 = (active + 7) & ~7;
    if (static_cast<int>(blockIdx.x) >= n8) return;
    const int L = (blockIdx.x & 7) * (n8 >> 3) + (blockIdx.x >> 3);
    if (L >= active) return;
    const int g = L / mchunks;
    T* ring = reinterpret_cast<T*>(lds_raw);
    constexpr int YP = kSlots * MW;
    constexpr int kPsi = 0, kPred = 4, kSPsi = 8, kSPred = 12;  // first slot of each ring

    const int mm = threadIdx.x % MW, ry = threadIdx.x / MW;
    const int m0 = (L - g * mchunks) * MW;
    const bool live = m0 + mm < nang;
    // Lanes past the last direction (and columns past Ny) compute a copy of
    // the last one from in-bounds addresses, so loads and arithmetic need no
    // predicate.  Their LDS writes repeat the owner's value at the owner's
    // slot in the same step (a harmless duplicate), and their global stores
    // go to an offset past the end of the buffer resource, which the
    // hardware's range check drops: no store needs a branch.
    const int m = live ? m0 + mm : nang - 1;
    const T wy = omend[2 * m], wz = omend[2 * m + 1];
    const bool py = wy >= T(0), pz = wz >= T(0);
    const T cy = fabs(wy) * inv_dy, cz = fabs(wz) * inv_dz;
    const T hy = py ? T(0.5) : T(-0.5), hz = pz ? T(0.5) : T(-0.5);
    const T hh = T(0.5) * h;
    const long long zstride = static_cast<long long>(Ny + 1) * nang;
    const unsigned zbytes = static_cast<unsigned>(zstride * sizeof(T));  // one z row
    // the group's block from direction m0 on, as a buffer resource: uniform
    // row offsets, 32-bit lane offsets
    const __amdgpu_buffer_rsrc_t Fg =
        make_rsrc(F + static_cast<long long>(g) * zstride * (Nz + 1) + m0,
                   static_cast<unsigned>((zstride * (Nz + 1) - m0) * sizeof(T)));

    unsigned boff[P];  // byte offset of the lane's column in a plane row
    unsigned soff[P];  // the same for stores; kOutside (dropped) for copies
    int o0[P], o_near[P], o_m[P], o_p[P];  // LDS column: own, upwind, y-1, y+1 (clamped)
    T win[P][8];    // psi rows t-1 .. t+2 after the commit of step t, then
                    // rows t+3 .. t+6 in flight (slot: row mod 8)
    T wpred[P][4];  // psi*, rows t-3 .. t after the predictor of step t
    T wbase[P][4];  // psi - h/2 L(psi), rows t-2 .. t
    T wS[P][4];     // own y slope of psi, rows t .. t+2 after step t's slope pass
    T wSp[P][2];    // own y slope of psi*, rows t-2, t-1 after step t's slope pass
    T sz[P][2];     // z slopes of psi, rows t-1 and t before step t
    T szp[P][2];    // z slopes of psi*, rows t-3 and t-2 before step t
    T uz[P], uzp[P];  // z term U of the row the predictor / corrector works on next
#pragma unroll
    for (int j = 0; j < P; ++j) {
        const int y = ry + j * R;
        const int yc = min(y, Ny);
        boff[j] = static_cast<unsigned>((yc * nang + (m - m0)) * sizeof(T));
        soff[j] = (y <= Ny && live) ? boff[j] : kOutside;
        auto lo = [&](int yy) { return (min(max(yy, 0), Ny)) * YP + mm; };
        o0[j] = lo(yc);
        o_near[j] = lo(py ? yc - 1 : yc + 1);
        o_m[j] = lo(yc - 1);
        o_p[j] = lo(yc + 1);
        const T r0 = bload<T>(Fg, boff[j], 0);
        const T r1 = bload<T>(Fg, boff[j], zbytes);  // Nz >= 1
        ring[o0[j] + kPsi * MW] = r0;
        ring[o0[j] + (kPsi + 1) * MW] = r1;
        win[j][7] = win[j][0] = r0;  // row -1 clamps to row 0
        win[j][1] = r1;
        win[j][6] = T(0);
#pragma unroll
        for (int r = 2; r < 6; ++r)  // rows past Nz: copies of row Nz
            win[j][r] = bload<T>(Fg, boff[j], static_cast<unsigned>(min(r, Nz)) * zbytes);
#pragma unroll
        for (int k = 0; k < 2; ++k) sz[j][k] = szp[j][k] = wSp[j][k] = T(0);  // boundary: 0
        uz[j] = uzp[j] = T(0);  // (edge steps form their own z terms)
#pragma unroll
        for (int k = 0; k < 4; ++k) wpred[j][k] = wbase[j][k] = wS[j][k] = T(0);
    }
    __syncthreads();
    // y slopes of psi rows 0 and 1 (later rows: two steps ahead, in the march)
#pragma unroll
    for (int j = 0; j < P; ++j) {
#pragma unroll
        for (int r = 0; r < 2; ++r) {
            const T* row = ring + (kPsi + r) * MW;
            const T S = limited_slope(row[o_p[j]] - win[j][r], win[j][r] - row[o_m[j]]);
            wS[j][r] = S;
            ring[o0[j] + (kSPsi + r) * MW] = S;
        }
    }

    // Step t (phase k8 = t mod 8, k = t mod 4): commit psi row t+2; slopes of
    // psi row t+2 and psi* row t-1; predictor on row t; corrector on row t-2.
    // Every step runs the same branch-free code; near the ends of the march
    // (EDGE) selects apply the clamped rows and the boundary slopes, and the
    // global store is skipped (uniformly) before row 0 and past row Nz.  Past
    // the last row the predictor works on clamped copies and its result is
    // replaced by the clamp; steps past Nz+2 (padding the last block) store
    // nothing.
    auto step = [&](int t, auto phase, auto edge) {
        constexpr int k8 = decltype(phase)::value;  // t mod 8
        constexpr int k = k8 & 3;
        constexpr bool E = decltype(edge)::value;
        const int rc = t + 2;  // commit psi row t+2 (LDS ring slot of row t-2)
        {
            T* crow = ring + (kPsi + ((k + 2) & 3)) * MW;
#pragma unroll
            for (int j = 0; j < P; ++j) {
                if (E && rc > Nz) win[j][(k8 + 2) & 7] = win[j][(k8 + 1) & 7];
                crow[o0[j]] = win[j][(k8 + 2) & 7];
            }
        }
        __syncthreads();
        // Keep each step's arithmetic after its barrier: moved earlier, the z
        // differences of later rows would wait for rows still in flight.
        __builtin_amdgcn_sched_barrier(0);
        {
            // row t+6 into the register slot of row t-2: four rows in flight
            const unsigned so = static_cast<unsigned>(E ? min(t + 6, Nz) : t + 6) * zbytes;
#pragma unroll
            for (int j = 0; j < P; ++j) win[j][(k8 + 6) & 7] = bload<T>(Fg, boff[j], so);
        }
        // LDS reads of the step, issued together: y neighbours of psi row t+2
        // and psi* row t-1 (their slopes), upwind value and slope of psi row t
        // (predictor) and psi* row t-2 (corrector)
        T a_m[P], a_p[P], b_m[P], b_p[P], pn[P], pS[P], cn_[P], cS[P];
        {
            const T* psi2 = ring + (kPsi + ((k + 2) & 3)) * MW;    // psi row t+2
            const T* prd1 = ring + (kPred + ((k + 3) & 3)) * MW;   // psi* row t-1
            const T* psi0 = ring + (kPsi + k) * MW;                // psi row t
            const T* sps0 = ring + (kSPsi + k) * MW;               // its slopes
            const T* prd2 = ring + (kPred + ((k + 2) & 3)) * MW;   // psi* row t-2
            const T* spd2 = ring + (kSPred + (k & 1)) * MW;        // its slopes
#pragma unroll
            for (int j = 0; j < P; ++j) {
                a_m[j] = psi2[o_m[j]];
                a_p[j] = psi2[o_p[j]];
                b_m[j] = prd1[o_m[j]];
                b_p[j] = prd1[o_p[j]];
                pn[j] = psi0[o_near[j]];
                pS[j] = sps0[o_near[j]];
                cn_[j] = prd2[o_near[j]];
                cS[j] = spd2[o_near[j]];
            }
        }
        {  // y slopes of psi row t+2 and psi* row t-1, published for step t+2
            T* sp2 = ring + (kSPsi + ((k + 2) & 3)) * MW;
            T* sd1 = ring + (kSPred + ((k + 1) & 1)) * MW;
#pragma unroll
            for (int j = 0; j < P; ++j) {
                const T c2 = win[j][(k8 + 2) & 7], c1 = wpred[j][(k + 3) & 3];
                const T S2 = limited_slope(a_p[j] - c2, c2 - a_m[j]);
                const T S1 = limited_slope(b_p[j] - c1, c1 - b_m[j]);
                wS[j][(k + 2) & 3] = S2;
                wSp[j][(k + 1) & 1] = S1;
                sp2[o0[j]] = S2;
                sd1[o0[j]] = S1;
            }
        }
        {  // predictor on row t
            T* prow = ring + (kPred + k) * MW;
#pragma unroll
            for (int j = 0; j < P; ++j) {
                const T zm = win[j][(k8 + 7) & 7], zc = win[j][k8], zp = win[j][(k8 + 1) & 7];
                const T dz_b = zc - zm, dz_f = zp - zc;
                const T s_next = limited_slope(win[j][(k8 + 2) & 7] - zp, dz_f);
                const T sz_m = sz[j][(k + 1) & 1], sz_0 = sz[j][k & 1];
                const T dS_f = s_next - sz_0;
                const T un = fma(hz, dS_f, dz_f);  // U of row t+1
                const T fz = E ? z_term(dz_b, dz_f, sz_0 - sz_m, dS_f, hz, pz) : (pz ? uz[j] : -un);
                uz[j] = un;
                const T L0 = divergence(zc, pn[j], wS[j][k], pS[j], fz, cy, cz, hy);
                T v = zc - h * L0;
                if (CLIP) v = fmax(v, T(0));
                prow[o0[j]] = v;
                sz[j][(k + 1) & 1] = s_next;
                wbase[j][k] = zc - hh * L0;
                wpred[j][k] = (!E || t <= Nz) ? v : wpred[j][(k + 3) & 3];  // past Nz: row Nz
                if (E && k == 0) wpred[j][3] = t == 0 ? v : wpred[j][3];  // row -1: row 0
            }
        }
        {  // corrector on row t-2, written in place
            const unsigned so = static_cast<unsigned>(E ? max(t - 2, 0) : t - 2) * zbytes;
            const bool real = !E || (t >= 2 && t <= Nz + 2);
#pragma unroll
            for (int j = 0; j < P; ++j) {
                const T zm = wpred[j][(k + 1) & 3], zc = wpred[j][(k + 2) & 3],
                        zp = wpred[j][(k + 3) & 3];
                const T dz_b = zc - zm, dz_f = zp - zc;
                const T s_next = limited_slope(wpred[j][k] - zp, dz_f);
                const T sz_m = szp[j][(k + 1) & 1], sz_0 = szp[j][k & 1];
                const T dS_f = s_next - sz_0;
                const T un = fma(hz, dS_f, dz_f);  // U of row t-1
                const T fz = E ? z_term(dz_b, dz_f, sz_0 - sz_m, dS_f, hz, pz) : (pz ? uzp[j] : -un);
                const T L1 = divergence(zc, cn_[j], wSp[j][k & 1], cS[j], fz, cy, cz, hy);
                T v = wbase[j][(k + 2) & 3] - hh * L1;
                if (CLIP) v = fmax(v, T(0));
                if (real) bstore(v, Fg, soff[j], so);
                const T s_kept = (!E || t >= 2) ? s_next : T(0);  // rows <= 0: boundary
                szp[j][(k + 1) & 1] = s_kept;
                // the next row's U from the slope kept for it (at edge steps the
                // kept slope may be the boundary's 0)
                uzp[j] = E ? fma(hz, s_kept - sz_0, dz_f) : un;
            }
        }
    };
    // four steps from t0, t0 mod 8 = B
    auto block = [&](int t0, auto base, auto edge) {
        constexpr int B = decltype(base)::value;
        step(t0, std::integral_constant<int, B>{}, edge);
        step(t0 + 1, std::integral_constant<int, B + 1>{}, edge);
        step(t0 + 2, std::integral_constant<int, B + 2>{}, edge);
        step(t0 + 3, std::integral_constant<int, B + 3>{}, edge);
    };
    using I0 = std::integral_constant<int, 0>;
    using I4 = std::integral_constant<int, 4>;
    block(0, I0{}, std::true_type{});
    int t0 = 4;
    // interior: 2 <= t and t + 6 <= Nz for all eight steps
    for (; t0 + 13 <= Nz; t0 += 8) {
        block(t0, I4{}, std::false_type{});
        block(t0 + 4, I0{}, std::false_type{});
    }
    for (; t0 <= Nz + 2; t0 += 8) {
        block(t0, I4{}, std::true_type{});
        if (t0 + 4 > Nz + 2) break;
        block(t0 + 4, I0{}, std::true_type{});
    }
}

struct MarchShape {
    int MW, P, R, threads;
};

// Launch shape, the first of (MW, P) = (16, 1), (16, 2), (8, 2), (8, 3)
// whose work-group covers the plane and whose LDS rings fit the
// 160 KB of a CU.  Measured on MI355X (FP32, 400 directions): 16 directions
// per work-group beat 32 (smaller groups pack the CU's wave slots better:
// 235 vs 248 us at 21 x 21 nodes) and 8 (64-byte row segments: 3.41 vs
// 3.57 ms at 81 x 81 nodes with two columns per thread).  At 81 x 81 nodes
// the (16, 2) shape runs one work-group per CU; 12 or 8 directions with one
// column per thread fit two, but their 48- and 32-byte row segments cost
// more than the occupancy gains (3.30 and 3.78 vs 2.52 ms per launch).
MarchShape march_shape(int Ny1, size_t elem) {
    const int max_t = elem == 4 ? max_threads<float>() : max_threads<double>();
    constexpr int cand[4][2] = {{16, 1}, {16, 2}, {8, 2}, {8, 3}};
    for (const auto& c : cand) {
        const int MW = c[0], P = c[1];
        if (static_cast<long long>(max_t / MW) * P < Ny1) continue;
        if (static_cast<size_t>(kSlots) * Ny1 * MW * elem > 160 * 1024) continue;
        const int R0 = (Ny1 + P - 1) / P;
        const int threads = (MW * R0 + 63) / 64 * 64;
        return MarchShape{MW, P, threads / MW, threads};
    }
    return MarchShape{0, 0, 0, 0};
}

template <typename T, int MW, int P>
void launch_march(T* F, const T* omend, int Ny, int Nz, int Ng, int nang, const MarchShape& sh,
                  double dy, double dz, double h, bool preserve_sign, hipStream_t s,
                  const int* top) {
    const size_t lds = sizeof(T) * kSlots * static_cast<size_t>(Ny + 1) * MW;
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
                           static_cast<int>(tiles), top);
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
                           double dz, double h, bool preserve_sign, hipStream_t s,
                           const int* top) {
    if (Ng <= 0 || nang <= 0) return;
    BFPN_REQUIRE(Ny >= 1 && Nz >= 1, "transport needs Ny, Nz >= 1");
    const MarchShape sh = march_shape(Ny + 1, sizeof(T));
    // the LDS holds 15 slots of MW values per y column: 15 * (Ny+1) * 8 * elem <= 160 KB
    BFPN_REQUIRE(sh.MW > 0, sizeof(T) == 4 ? "transport supports at most 341 y nodes in FP32"
                                           : "transport supports at most 170 y nodes in FP64");
    // one group's block is addressed with 32-bit byte offsets
    BFPN_REQUIRE(static_cast<long long>(Ny + 1) * (Nz + 1) * nang * sizeof(T) < (1LL << 31),
                 "transport needs (Ny+1)(Nz+1)*Nmu*Nom values of one energy group below 2 GiB");
#define BFPN_MARCH(MW_, P_)                                                                    \
    if (sh.MW == MW_ && sh.P == P_)                                                            \
        return launch_march<T, MW_, P_>(F, omend, Ny, Nz, Ng, nang, sh, dy, dz, h, preserve_sign, \
                                        s, top)
    BFPN_MARCH(16, 1);
    BFPN_MARCH(16, 2);
    BFPN_MARCH(8, 2);
    BFPN_MARCH(8, 3);
#undef BFPN_MARCH
    throw Error(BFPN_E_INVALID, "no transport launch shape for this plane");
}

template void launch_transport_half<float>(float*, const float*, int, int, int, int, double, double,
                                           double, bool, hipStream_t, const int*);
template void launch_transport_half<double>(double*, const double*, int, int, int, int, double,
                                            double, double, bool, hipStream_t, const int*);

}  // namespace bfpn
