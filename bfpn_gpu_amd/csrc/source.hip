// source.hip — catastrophic secondary source Q1 from the primary state.
//
// Reference: secondarySourceFromPrimaryKernel (src/energy_solver.cu:287-333):
//   Q1[g,s,m] = dg * sum_{g'} sum_{m'} (K1[g,g'] max(Fp,0) + K2[g,g'] fp/3)[g',s,m'] Kv[g',m',m]
// evaluated per output cell as a (g', m') double loop (1.6k-4k terms/cell).
//
// MI355X design: the sum is factorised into two GEMMs,
//   1. angular, one per source group g':  A1[g'] = max(Fp[g'],0) . Kv[g'],
//      A2[g'] = fp[g'] . Kv[g']   (rows = transverse nodes, K = N = nang);
//   2. banded energy:  Q = E1 . A1 + E2 . A2  with E1 = dg K1 and E2 = dg K2/3,
//      where each 16-group output tile only visits the source groups of its
//      band [min beg, max end).
// FP32 state runs both on the matrix cores in split-BF16 with FP32
// accumulation (three MFMA issues per tile, ~2^-16 relative per product);
// the Kv and E operands are split once at set-up.  FP64 state keeps exact
// FP64 vector kernels (oracle-grade instantiation).  This is the reference's
// sum regrouped; only rounding order differs.

#include "kernels.hpp"

#include <algorithm>
#include <vector>
#include "mfma.hpp"

namespace bfpn {
namespace {

// ------------------------------------------------------------ FP32 / MFMA --
// Work-groups are dispatched round-robin over the 8 XCDs (each with its own
// L2).  Both GEMMs re-read their operands from several work-groups, so block
// ids are remapped to put the work-groups that share operands on one XCD,
// consecutively in dispatch order: block b runs logical tile
//     (b mod 8) * (grid / 8) + b / 8,   grid padded to a multiple of 8.
__device__ inline int xcd_tile(int b, int grid) { return (b & 7) * (grid >> 3) + (b >> 3); }
inline int xcd_grid(long long tiles) { return static_cast<int>((tiles + 7) / 8 * 8); }

// Angular GEMM.  Logical tile = (g, 64 transverse nodes, 64 directions); the
// 7 x 7 tiles of one source group g share X[g] and Kv[g] through one L2.
// Kv is thresholded (at 20 x 20 directions, water 230 MeV, 58 % of its
// 64 x 32 blocks are zero in the groups that have a kernel, all of them in
// the others), so each tile visits only its nonzero K tiles.
// 4 waves, each a 32x32 output block of both inputs.  Per 32-wide K step the
// work-group stages a 64(s) x 32(m') tile of both inputs -- split into BF16
// hi/lo once -- and the matching 64(m) x 32(m') tile of the pre-split
// transposed Kv in LDS; the next step's global loads are in flight while the
// matrix cores work on the current one.  VEC: nang % 8 == 0 (16-byte loads).
constexpr int kKPitch = 40;  // bf16 per LDS row (32 + 8 pad)

template <bool VEC>
__global__ __launch_bounds__(256) void source_angular_mfma(const float* __restrict__ Fp,
                                                           const float* __restrict__ fp,
                                                           const __bf16* __restrict__ kth,
                                                           const __bf16* __restrict__ ktl,
                                                           const int* __restrict__ kt_count,
                                                           const int* __restrict__ kt_list,
                                                           const int* __restrict__ glist,
                                                           const int* __restrict__ gcount,
                                                           float* __restrict__ A1,
                                                           float* __restrict__ A2, int Ng,
                                                           int nyz, int nang, int nap, int tiles,
                                                           const int* __restrict__ xtop) {
    __shared__ __attribute__((aligned(16))) __bf16 xs[4][64][kKPitch];  // X1 hi/lo, X2 hi/lo
    __shared__ __attribute__((aligned(16))) __bf16 ks[2][64][kKPitch];  // Kv^T hi/lo
    const int mtiles = (nang + 63) / 64, stiles = (nyz + 63) / 64;
    // Tiles run over the source groups of glist (tiles = its length x per_g).
    // Those above the primary's top are zero: their tiles (the last ones) are
    // skipped, and their A rows are never read.
    const int per_g = mtiles * stiles;
    const int active =
        xtop ? min(tiles, gcount[min(max(-1, *xtop), Ng - 1) + 1] * per_g) : tiles;
    const int n8 = (active + 7) & ~7;
    if (static_cast<int>(blockIdx.x) >= n8) return;
    const int L = xcd_tile(blockIdx.x, n8);
    if (L >= active) return;
    const int gi = L / per_g, g = glist[gi];
    const int rest = L - gi * per_g;
    const int s0 = (rest / mtiles) * 64, m0 = (rest % mtiles) * 64;
    const int wave = threadIdx.x >> 6, lane = threadIdx.x & 63;
    const int i = lane & 15, q = lane >> 4;
    const int ws = (wave & 1) * 32, wm = (wave >> 1) * 32;  // wave's sub-block
    const long long plane = static_cast<long long>(g) * nyz * nang;
    const long long kt_plane = static_cast<long long>(g) * nap * nap;
    const int lrow = threadIdx.x >> 2, lk = (threadIdx.x & 3) * 8;  // staging: row, 8 columns
    const int srow = s0 + lrow, mrow = m0 + lrow;
    f32x4 acc1[2][2], acc2[2][2];
#pragma unroll
    for (int a = 0; a < 2; ++a)
#pragma unroll
        for (int b = 0; b < 2; ++b) acc1[a][b] = acc2[a][b] = zero4();

    float v1[8], v2[8];  // next K step, in flight
    bf16x8 kh, kl;
    auto load = [&](int k0) {
        const long long base = plane + static_cast<long long>(srow) * nang + k0 + lk;
        if (VEC) {
            // Unmasked: rows past nyz read the last row (their products only
            // reach output rows that are never stored) and columns past nang
            // read the row's last 8 values (Kv^T is zero there).
            const long long cb = plane + static_cast<long long>(min(srow, nyz - 1)) * nang +
                                 min(k0 + lk, nang - 8);
            const float4 a0 = *reinterpret_cast<const float4*>(Fp + cb);
            const float4 a1 = *reinterpret_cast<const float4*>(Fp + cb + 4);
            const float4 b0 = *reinterpret_cast<const float4*>(fp + cb);
            const float4 b1 = *reinterpret_cast<const float4*>(fp + cb + 4);
            v1[0] = a0.x; v1[1] = a0.y; v1[2] = a0.z; v1[3] = a0.w;
            v1[4] = a1.x; v1[5] = a1.y; v1[6] = a1.z; v1[7] = a1.w;
            v2[0] = b0.x; v2[1] = b0.y; v2[2] = b0.z; v2[3] = b0.w;
            v2[4] = b1.x; v2[5] = b1.y; v2[6] = b1.z; v2[7] = b1.w;
        } else {
#pragma unroll
            for (int e = 0; e < 8; ++e) {
                const bool in = srow < nyz && k0 + lk + e < nang;
                v1[e] = in ? Fp[base + e] : 0.f;
                v2[e] = in ? fp[base + e] : 0.f;
            }
        }
        {  // rows past nap read the last row: their columns are never stored
            const long long off = kt_plane + static_cast<long long>(min(mrow, nap - 1)) * nap + k0 + lk;
            kh = *reinterpret_cast<const bf16x8*>(kth + off);
            kl = *reinterpret_cast<const bf16x8*>(ktl + off);
        }
    };
    // only the K tiles where this direction tile's Kv block is nonzero (the
    // skipped products are exact zeros)
    const int mt_idx = m0 / 64, ktiles = nap / 32;
    const int nk = kt_count[g * mtiles + mt_idx];
    const int* kl_ = kt_list + static_cast<long long>(g * mtiles + mt_idx) * ktiles;
    if (nk > 0) load(kl_[0] * 32);
    for (int j = 0; j < nk; ++j) {
        {  // stage: rows s0+lrow / m0+lrow, columns k0+lk .. +7
            float p1[8];
#pragma unroll
            for (int e = 0; e < 8; ++e) p1[e] = v1[e] > 0.f ? v1[e] : 0.f;
            const SplitFrag f1 = split8(p1), f2 = split8(v2);
            *reinterpret_cast<bf16x8*>(&xs[0][lrow][lk]) = f1.hi;
            *reinterpret_cast<bf16x8*>(&xs[1][lrow][lk]) = f1.lo;
            *reinterpret_cast<bf16x8*>(&xs[2][lrow][lk]) = f2.hi;
            *reinterpret_cast<bf16x8*>(&xs[3][lrow][lk]) = f2.lo;
            *reinterpret_cast<bf16x8*>(&ks[0][lrow][lk]) = kh;
            *reinterpret_cast<bf16x8*>(&ks[1][lrow][lk]) = kl;
        }
        __syncthreads();
        if (j + 1 < nk) load(kl_[j + 1] * 32);
        SplitFrag x1[2], x2[2], kf[2];
#pragma unroll
        for (int t = 0; t < 2; ++t) {
            const int r = ws + t * 16 + i, c = wm + t * 16 + i;
            x1[t].hi = *reinterpret_cast<const bf16x8*>(&xs[0][r][8 * q]);
            x1[t].lo = *reinterpret_cast<const bf16x8*>(&xs[1][r][8 * q]);
            x2[t].hi = *reinterpret_cast<const bf16x8*>(&xs[2][r][8 * q]);
            x2[t].lo = *reinterpret_cast<const bf16x8*>(&xs[3][r][8 * q]);
            kf[t].hi = *reinterpret_cast<const bf16x8*>(&ks[0][c][8 * q]);
            kf[t].lo = *reinterpret_cast<const bf16x8*>(&ks[1][c][8 * q]);
        }
#pragma unroll
        for (int mt = 0; mt < 2; ++mt)
#pragma unroll
            for (int nt = 0; nt < 2; ++nt) {
                acc1[mt][nt] = mma_split(x1[mt], kf[nt], acc1[mt][nt]);
                acc2[mt][nt] = mma_split(x2[mt], kf[nt], acc2[mt][nt]);
            }
        __syncthreads();
    }
    // Epilogue: each output tile passes through LDS (the staging buffers are
    // free after the loop's last barrier) so that 16 threads store a row's
    // 64 directions as one 256-byte run of 16-byte vectors, instead of each
    // MFMA fragment's four 64-byte pieces.
    constexpr int kSP = 68;  // floats per staged row (64 + 4: rows start on different banks)
    static_assert(64 * kSP * sizeof(float) <= sizeof(xs), "output tile must fit the X staging");
    float* stage = reinterpret_cast<float*>(&xs[0][0][0]);
    const int orow = threadIdx.x >> 4, ocol = (threadIdx.x & 15) * 4;  // store: 16 rows per pass
#pragma unroll
    for (int in = 0; in < 2; ++in) {
        const f32x4 (&acc)[2][2] = in == 0 ? acc1 : acc2;
        float* A = in == 0 ? A1 : A2;
#pragma unroll
        for (int mt = 0; mt < 2; ++mt)
#pragma unroll
            for (int nt = 0; nt < 2; ++nt)
#pragma unroll
                for (int r = 0; r < 4; ++r)
                    stage[(ws + mt * 16 + 4 * q + r) * kSP + wm + nt * 16 + i] = acc[mt][nt][r];
        __syncthreads();
#pragma unroll
        for (int pass = 0; pass < 4; ++pass) {
            const int row = pass * 16 + orow, sidx = s0 + row, m = m0 + ocol;
            if (sidx < nyz && m < nang) {
                const float* v = stage + row * kSP + ocol;
                const long long idx = plane + static_cast<long long>(sidx) * nang + m;
                if (VEC) {  // nang % 8 == 0: m + 3 < nang and 16-byte aligned
                    *reinterpret_cast<float4*>(A + idx) = make_float4(v[0], v[1], v[2], v[3]);
                } else {
#pragma unroll
                    for (int e = 0; e < 4; ++e)
                        if (m + e < nang) A[idx + e] = v[e];
                }
            }
        }
        __syncthreads();
    }
}

constexpr int kBandLanes = 128;  // lanes per work-group
constexpr int kBandRows = 64;    // output groups per work-group

// Banded energy GEMM.  Logical tile = (128 lanes, 64 output groups), the
// group blocks of one lane block consecutive (one XCD, one L2 for the A rows
// their bands share).  Q[g][lane] = E1.A1 + E2.A2 over the union of the
// 16-row tiles' bands; each 16-row tile skips K steps outside its own band.
// Per 32-wide K step the work-group stages in LDS the 32 x 128 tiles of A1
// and A2 (FP32) and the 64 x 32 tiles of E1 and E2 (BF16 hi/lo), the next
// step's tiles in flight in registers meanwhile.  Each wave forms the B
// fragments of one input at a time and reads E fragments one 16-row tile at a
// time from LDS (rows padded to 40 values against bank conflicts), so few
// fragments are live (three work-groups per CU).  The
// A tiles have no padding: the 16-column blocks of a row are XOR-swizzled by
// (row / 8) mod 4, so the fragment reads (rows 8q..8q+7 of one column, q =
// 0..3) fall on 64 distinct banks.  VEC: lanes % 4 == 0 (16-byte loads).
__device__ inline int band_swz(int r, int c) { return c ^ (((r >> 3) & 3) << 4); }

template <bool VEC>
__global__ __launch_bounds__(256) void source_band_mfma(
    const float* __restrict__ A1, const float* __restrict__ A2, const __bf16* __restrict__ e1h,
    const __bf16* __restrict__ e1l, const __bf16* __restrict__ e2h, const __bf16* __restrict__ e2l,
    const int2* __restrict__ tile_band, const unsigned* __restrict__ amask, float* __restrict__ Q,
    int Ng, int ngp, long long lanes, int tiles, const int* __restrict__ xtop,
    const int* __restrict__ qmap) {
    __shared__ __attribute__((aligned(16))) float bs[2][32][kBandLanes];        // A1, A2
    __shared__ __attribute__((aligned(16))) __bf16 es[4][kBandRows][kKPitch];   // E1 hi/lo, E2 hi/lo
    int rblocks = (Ng + kBandRows - 1) / kBandRows;
    // A rows above the primary's top (kmax) hold no data of this step and
    // read as zero; output group blocks above the source's top are skipped
    // (their values are never read).
    int kmax = Ng - 1, active = tiles;
    if (xtop) {
        kmax = max(-1, min(Ng - 1, *xtop));
        const int lblocks = tiles / rblocks;
        rblocks = (qmap[kmax + 1] + kBandRows) / kBandRows;  // blocks up to the source's top
        active = rblocks * lblocks;
    }
    const int n8 = (active + 7) & ~7;
    if (static_cast<int>(blockIdx.x) >= n8) return;
    const int L = xcd_tile(blockIdx.x, n8);
    if (L >= active) return;
    const int wave = threadIdx.x >> 6, lane = threadIdx.x & 63;
    const int i = lane & 15, q = lane >> 4;
    const int g_base = (L % rblocks) * kBandRows;
    const long long n_base = static_cast<long long>(L / rblocks) * kBandLanes;
    const int t0 = g_base / 16;
    int2 band[4];
    int kb = 1 << 30, ke = 0;
#pragma unroll
    for (int mt = 0; mt < 4; ++mt) {
        band[mt] = tile_band[t0 + mt];
        if (band[mt].y > band[mt].x) {
            kb = min(kb, band[mt].x);
            ke = max(ke, band[mt].y);
        }
    }
    ke = min(ke, kmax + 1);  // zero A rows contribute exact zeros
    f32x4 acc[4][2];
#pragma unroll
    for (int mt = 0; mt < 4; ++mt) acc[mt][0] = acc[mt][1] = zero4();

    // staging: A, 4 float4 per thread and input; E, one 64-byte row segment
    // (32 columns) of array ea = wave, row er = lane
    float4 r1[4], r2[4];
    bf16x8 re[4];
    const __bf16* esrc = wave == 0 ? e1h : wave == 1 ? e1l : wave == 2 ? e2h : e2l;
    // A rows of source groups outside amask (zero Kv, or in no band) are not
    // written by the angular GEMM and read as zero.
    auto load = [&](int k0) {
        const unsigned rows = amask[k0 >> 5];  // k0 % 32 == 0
#pragma unroll
        for (int j = 0; j < 4; ++j) {
            const int e4 = threadIdx.x + j * 256;
            const int r = e4 >> 5, c = (e4 & 31) * 4;
            const int gs = k0 + r;
            const bool has = gs <= kmax && ((rows >> r) & 1u);
            const long long ln = n_base + c;
            const long long idx = static_cast<long long>(gs) * lanes + ln;
            if (VEC) {
                r1[j] = r2[j] = make_float4(0.f, 0.f, 0.f, 0.f);
                if (has && ln < lanes) {
                    r1[j] = *reinterpret_cast<const float4*>(A1 + idx);
                    r2[j] = *reinterpret_cast<const float4*>(A2 + idx);
                }
            } else {
                float a[4], b[4];
#pragma unroll
                for (int e = 0; e < 4; ++e) {
                    const bool in = has && ln + e < lanes;
                    a[e] = in ? A1[idx + e] : 0.f;
                    b[e] = in ? A2[idx + e] : 0.f;
                }
                r1[j] = make_float4(a[0], a[1], a[2], a[3]);
                r2[j] = make_float4(b[0], b[1], b[2], b[3]);
            }
        }
        // rows g_base .. g_base+63 < erows and columns k0 .. k0+31 < ngp
        const __bf16* src = esrc + static_cast<long long>(g_base + lane) * ngp + k0;
#pragma unroll
        for (int c = 0; c < 4; ++c) re[c] = *reinterpret_cast<const bf16x8*>(src + 8 * c);
    };
    const int kstart = (kb / 32) * 32;
    if (kstart < ke) load(kstart);
    for (int k0 = kstart; k0 < ke; k0 += 32) {
#pragma unroll
        for (int j = 0; j < 4; ++j) {
            const int e4 = threadIdx.x + j * 256;
            const int r = e4 >> 5, c = band_swz(r, (e4 & 31) * 4);
            *reinterpret_cast<float4*>(&bs[0][r][c]) = r1[j];
            *reinterpret_cast<float4*>(&bs[1][r][c]) = r2[j];
        }
#pragma unroll
        for (int c = 0; c < 4; ++c) *reinterpret_cast<bf16x8*>(&es[wave][lane][8 * c]) = re[c];
        __syncthreads();
        if (k0 + 32 < ke) load(k0 + 32);
#pragma unroll
        for (int in = 0; in < 2; ++in) {  // E1.A1, then E2.A2 (the same order per accumulator)
            SplitFrag bf[2];
#pragma unroll
            for (int nt = 0; nt < 2; ++nt) {
                const int c = wave * 32 + nt * 16 + i;
                float v[8];
#pragma unroll
                for (int e = 0; e < 8; ++e) v[e] = bs[in][8 * q + e][band_swz(8 * q + e, c)];
                bf[nt] = split8(v);
            }
#pragma unroll
            for (int mt = 0; mt < 4; ++mt) {
                if (band[mt].y <= band[mt].x || k0 + 32 <= band[mt].x || k0 >= band[mt].y) continue;
                SplitFrag a;
                a.hi = *reinterpret_cast<const bf16x8*>(&es[2 * in][mt * 16 + i][8 * q]);
                a.lo = *reinterpret_cast<const bf16x8*>(&es[2 * in + 1][mt * 16 + i][8 * q]);
#pragma unroll
                for (int nt = 0; nt < 2; ++nt) acc[mt][nt] = mma_split(a, bf[nt], acc[mt][nt]);
            }
        }
        __syncthreads();
    }
#pragma unroll
    for (int mt = 0; mt < 4; ++mt)
#pragma unroll
        for (int nt = 0; nt < 2; ++nt) {
            const long long ln = n_base + wave * 32 + nt * 16 + i;
#pragma unroll
            for (int r = 0; r < 4; ++r) {
                const int g = g_base + mt * 16 + 4 * q + r;
                if (g < Ng && ln < lanes) Q[static_cast<long long>(g) * lanes + ln] = acc[mt][nt][r];
            }
        }
}

// Banded energy GEMM as one march over the source groups.  A work-group
// owns 128 lanes and walks the 32-wide source-group steps once, so every A
// row it needs is read once; the output tiles (16 groups) whose bands hold
// the current step -- at most kMarchSlots consecutive tiles, a host-checked
// window -- keep their accumulators in a register ring (tile t in slot
// t mod kMarchSlots), and a tile is stored when the march passes its band.
// Per step the work-group stages the 32 x 128 tiles of A1 and A2 (FP32,
// XOR-swizzled as in source_band_mfma) and the window's E rows (BF16 hi/lo of
// E1 and E2) in LDS, the latter in MFMA fragment order (16-byte chunk
// [slot][array][q][row]); the next step's tiles are in flight in registers.
// 8 waves, 16 lanes each.  Per output tile the sums run over the same steps
// in the same order, with the same operands, as in source_band_mfma.
constexpr int kMarchLanes = 128;
constexpr int kMarchSlots = 10;

template <bool VEC>
__global__ __launch_bounds__(512, 4) void source_band_march(
    const float* __restrict__ A1, const float* __restrict__ A2, const __bf16* __restrict__ e1h,
    const __bf16* __restrict__ e1l, const __bf16* __restrict__ e2h, const __bf16* __restrict__ e2l,
    const int2* __restrict__ tile_band, const int2* __restrict__ steps, int s_first, int s_last,
    const unsigned* __restrict__ amask, float* __restrict__ Q, int Ng, int ngp, long long lanes,
    int lblocks, const int* __restrict__ xtop, const int* __restrict__ qmap, int q_rows) {
    constexpr int W = kMarchSlots;
    __shared__ __attribute__((aligned(16))) float as[2][32][kMarchLanes];  // A1, A2
    __shared__ bf16x8 es[W * 256];  // [slot][array][q][row] 16-byte chunks
    const int n8 = (lblocks + 7) & ~7;
    if (static_cast<int>(blockIdx.x) >= n8) return;
    const int L = xcd_tile(blockIdx.x, n8);
    if (L >= lblocks) return;
    const long long n_base = static_cast<long long>(L) * kMarchLanes;
    // A rows above the primary's top (kmax) read as zero; output rows above
    // the source's top (gmax) are never read.
    int kmax = Ng - 1, gmax = Ng - 1;
    if (xtop) {
        kmax = max(-1, min(Ng - 1, *xtop));
        gmax = qmap[kmax + 1];
    }
    if (q_rows >= 0) gmax = min(gmax, q_rows);  // rows above are never read
    const int s_end = min(s_last, kmax < 0 ? 0 : kmax / 32 + 1);  // exclusive
    const int wave = threadIdx.x >> 6, lane = threadIdx.x & 63;
    const int i = lane & 15, q = lane >> 4;
    const long long ln_out = n_base + wave * 16 + i;  // this lane's output column
    f32x4 acc[W];
#pragma unroll
    for (int s = 0; s < W; ++s) acc[s] = zero4();

    float4 ra[4];   // next step's A: 2 float4 per input
    bf16x8 re[5];   // next step's E window: W * 256 / 512 chunks per thread
    auto load = [&](int k) {
        const int k0 = 32 * k;
        const unsigned rows = amask[k];
#pragma unroll
        for (int j = 0; j < 4; ++j) {
            const int in = j >> 1, e4 = threadIdx.x + (j & 1) * 512;
            const int r = e4 >> 5, c = (e4 & 31) * 4;
            const int gs = k0 + r;
            const bool has = gs <= kmax && ((rows >> r) & 1u);
            const long long ln = n_base + c;
            const float* A = in == 0 ? A1 : A2;
            const long long idx = static_cast<long long>(gs) * lanes + ln;
            if (VEC) {
                ra[j] = make_float4(0.f, 0.f, 0.f, 0.f);
                if (has && ln < lanes) ra[j] = *reinterpret_cast<const float4*>(A + idx);
            } else {
                float v[4];
#pragma unroll
                for (int e = 0; e < 4; ++e) v[e] = has && ln + e < lanes ? A[idx + e] : 0.f;
                ra[j] = make_float4(v[0], v[1], v[2], v[3]);
            }
        }
        const int2 w = steps[k];  // live tiles [x, y]
        const int r0 = w.x % W;
#pragma unroll
        for (int j = 0; j < 5; ++j) {
            const int c = threadIdx.x + 512 * j;
            const int slot = c >> 8, rem = c & 255;
            const int a = rem >> 6, cq = (rem >> 4) & 3, row = rem & 15;
            const int t = w.x + (slot - r0 + W) % W;
            re[j] = bf16x8{};
            if (t <= w.y) {
                const __bf16* E = a == 0 ? e1h : a == 1 ? e1l : a == 2 ? e2h : e2l;
                re[j] = *reinterpret_cast<const bf16x8*>(
                    E + static_cast<long long>(16 * t + row) * ngp + k0 + 8 * cq);
            }
        }
    };
    auto store_tile = [&](int t, const f32x4& v) {
#pragma unroll
        for (int r = 0; r < 4; ++r) {
            const int g = 16 * t + 4 * q + r;
            if (g < Ng && ln_out < lanes) Q[static_cast<long long>(g) * lanes + ln_out] = v[r];
        }
    };

    if (s_first < s_end) load(s_first);
    for (int k = s_first; k < s_end; ++k) {
#pragma unroll
        for (int j = 0; j < 4; ++j) {
            const int in = j >> 1, e4 = threadIdx.x + (j & 1) * 512;
            const int r = e4 >> 5, c = band_swz(r, (e4 & 31) * 4);
            *reinterpret_cast<float4*>(&as[in][r][c]) = ra[j];
        }
#pragma unroll
        for (int j = 0; j < 5; ++j) es[threadIdx.x + 512 * j] = re[j];
        __syncthreads();
        if (k + 1 < s_end) load(k + 1);
        const int2 w = steps[k];
        const int r0 = w.x % W;
        int tl[W];
        bool on[W], done[W];
#pragma unroll
        for (int s = 0; s < W; ++s) {  // the tile in each slot, and whether it works / retires
            tl[s] = w.x + (s - r0 + W) % W;
            on[s] = done[s] = false;
            if (tl[s] <= w.y) {
                const int2 b = tile_band[tl[s]];
                on[s] = b.y > b.x && b.x < 32 * k + 32 && b.y > 32 * k;
                done[s] = on[s] && min((b.y - 1) / 32, s_end - 1) == k;
            }
        }
#pragma unroll
        for (int in = 0; in < 2; ++in) {  // E1.A1, then E2.A2 (the same order per accumulator)
            SplitFrag bf;
            {
                const int c = wave * 16 + i;
                float v[8];
#pragma unroll
                for (int e = 0; e < 8; ++e) v[e] = as[in][8 * q + e][band_swz(8 * q + e, c)];
                bf = split8(v);
            }
#pragma unroll
            for (int s = 0; s < W; ++s) {
                if (!on[s]) continue;
                SplitFrag a;
                a.hi = es[s * 256 + (2 * in) * 64 + q * 16 + i];
                a.lo = es[s * 256 + (2 * in + 1) * 64 + q * 16 + i];
                acc[s] = mma_split(a, bf, acc[s]);
            }
        }
#pragma unroll
        for (int s = 0; s < W; ++s)
            if (done[s]) {
                store_tile(tl[s], acc[s]);
                acc[s] = zero4();
            }
        __syncthreads();
    }
    // tiles no step touched: zero rows up to the source's top
    const int tiles = min((gmax + 16) / 16, (Ng + 15) / 16);
    for (int t = 0; t < tiles; ++t) {
        const int2 b = tile_band[t];
        const bool marched = b.y > b.x && b.x < 32 * s_end && b.y > 32 * s_first && s_first < s_end;
        if (!marched) store_tile(t, zero4());
    }
}

// grid: (ktiles, mtiles, Ng)
__global__ __launch_bounds__(256) void kv_block_flags_kernel(const __bf16* __restrict__ kth,
                                                             unsigned char* __restrict__ flags,
                                                             int nap) {
    const int kt = blockIdx.x, mt = blockIdx.y, g = blockIdx.z;
    const unsigned short* k = reinterpret_cast<const unsigned short*>(kth);
    int nz = 0;
    for (int e = threadIdx.x; e < 64 * 32; e += 256) {
        const int r = mt * 64 + (e >> 5), c = kt * 32 + (e & 31);
        if (r < nap) nz |= (k[(static_cast<long long>(g) * nap + r) * nap + c] & 0x7fff) != 0;
    }
    nz = __syncthreads_or(nz);
    if (threadIdx.x == 0)
        flags[(static_cast<long long>(g) * gridDim.y + mt) * gridDim.x + kt] = nz ? 1 : 0;
}

// Kv[g][in][out] (FP32) -> transposed split copies KT[g][out][in] (BF16 hi/lo),
// zero padded to nap x nap.  grid: (ceil(nap*nap/256), Ng)
__global__ void split_transpose_kernel(const float* __restrict__ kv, __bf16* __restrict__ hi,
                                       __bf16* __restrict__ lo, int nang, int nap) {
    const int e = blockIdx.x * 256 + threadIdx.x;
    if (e >= nap * nap) return;
    const int out = e / nap, in = e - out * nap;
    const long long g = blockIdx.y;
    const float v = (out < nang && in < nang) ? kv[(g * nang + in) * nang + out] : 0.f;
    __bf16 h, l;
    split_bf16(v, h, l);
    hi[g * nap * nap + e] = h;
    lo[g * nap * nap + e] = l;
}

__global__ void split_kernel(const float* __restrict__ src, __bf16* __restrict__ hi,
                             __bf16* __restrict__ lo, long long n) {
    const long long e = static_cast<long long>(blockIdx.x) * 256 + threadIdx.x;
    if (e >= n) return;
    __bf16 h, l;
    split_bf16(src[e], h, l);
    hi[e] = h;
    lo[e] = l;
}

// ------------------------------------------------------------ FP64 vector --
constexpr int kTile = 64;
constexpr int kDepth = 16;

template <typename T>
__global__ __launch_bounds__(256) void source_angular_kernel(const T* __restrict__ Fp,
                                                             const T* __restrict__ fp,
                                                             const T* __restrict__ kv,
                                                             T* __restrict__ A1, T* __restrict__ A2,
                                                             int nyz, int nang,
                                                             const int* __restrict__ xtop,
                                                             const int* __restrict__ kcount,
                                                             const int* __restrict__ klist,
                                                             const int* __restrict__ need) {
    __shared__ T x1s[kDepth][kTile + 1];
    __shared__ T x2s[kDepth][kTile + 1];
    __shared__ T ks[kDepth][kTile + 1];
    const int g = blockIdx.z;
    if (xtop && g > *xtop) return;  // a zero source group: its A rows are never read
    if (need && !need[g]) return;   // A rows the banded sum does not read
    const int m0 = blockIdx.x * kTile;
    const int s0 = blockIdx.y * kTile;
    const long long plane = static_cast<long long>(g) * nyz * nang;
    const T* K = kv + static_cast<long long>(g) * nang * nang;
    const int tx = threadIdx.x & 15;
    const int ty = threadIdx.x >> 4;
    T acc1[4][4], acc2[4][4];
#pragma unroll
    for (int i = 0; i < 4; ++i)
#pragma unroll
        for (int j = 0; j < 4; ++j) acc1[i][j] = acc2[i][j] = T(0);
    // With block lists, only the kDepth-deep K blocks where this direction
    // tile's Kv block has a nonzero entry (the skipped products are zeros).
    const int ktiles = (nang + kDepth - 1) / kDepth;
    int nk = ktiles;
    const int* list = nullptr;
    if (kcount) {
        const long long t = static_cast<long long>(g) * gridDim.x + blockIdx.x;
        nk = kcount[t];
        list = klist + t * ktiles;
    }
    for (int kj = 0; kj < nk; ++kj) {
        const int k0 = (list ? list[kj] : kj) * kDepth;
        for (int e = threadIdx.x; e < kTile * kDepth; e += 256) {
            const int kk = e % kDepth, r = e / kDepth;
            const int s = s0 + r, k = k0 + kk;
            T v1 = T(0), v2 = T(0);
            if (s < nyz && k < nang) {
                const long long idx = plane + static_cast<long long>(s) * nang + k;
                v1 = Fp[idx] > T(0) ? Fp[idx] : T(0);
                v2 = fp[idx];
            }
            x1s[kk][r] = v1;
            x2s[kk][r] = v2;
        }
        for (int e = threadIdx.x; e < kTile * kDepth; e += 256) {
            const int c = e % kTile, kk = e / kTile;
            const int k = k0 + kk, m = m0 + c;
            ks[kk][c] = (k < nang && m < nang) ? K[static_cast<long long>(k) * nang + m] : T(0);
        }
        __syncthreads();
#pragma unroll
        for (int kk = 0; kk < kDepth; ++kk) {
            T a1[4], a2[4], b[4];
#pragma unroll
            for (int i = 0; i < 4; ++i) {
                a1[i] = x1s[kk][ty + 16 * i];
                a2[i] = x2s[kk][ty + 16 * i];
            }
#pragma unroll
            for (int j = 0; j < 4; ++j) b[j] = ks[kk][tx + 16 * j];
#pragma unroll
            for (int i = 0; i < 4; ++i)
#pragma unroll
                for (int j = 0; j < 4; ++j) {
                    acc1[i][j] += a1[i] * b[j];
                    acc2[i][j] += a2[i] * b[j];
                }
        }
        __syncthreads();
    }
#pragma unroll
    for (int i = 0; i < 4; ++i) {
        const int s = s0 + ty + 16 * i;
        if (s >= nyz) continue;
#pragma unroll
        for (int j = 0; j < 4; ++j) {
            const int m = m0 + tx + 16 * j;
            if (m < nang) {
                const long long idx = plane + static_cast<long long>(s) * nang + m;
                A1[idx] = acc1[i][j];
                A2[idx] = acc2[i][j];
            }
        }
    }
}

// Banded energy sum, FP64 vector path.  A work-group takes 256 lanes (one
// per thread) and 16 consecutive output groups, walks the union of their
// bands once and keeps the 32 sums (E1.A1, E2.A2 per group) in registers, so
// each A row is read once per 16 output groups rather than once per group.
// The E rows of the 16 groups are staged through LDS, 128 source groups at a
// time, and read as broadcasts.  Every sum runs over ascending source groups
// (a zero weight adds an exact zero), as in the reference's loop.
constexpr int kEnergyRows = 16;
constexpr int kEnergyChunk = 128;

template <typename T>
__global__ __launch_bounds__(256) void source_energy_kernel(
    const T* __restrict__ A1, const T* __restrict__ A2, const double* __restrict__ e1,
    const double* __restrict__ e2, const int* __restrict__ beg, const int* __restrict__ end,
    T* __restrict__ Q, int Ng, long long lanes, double dg, const int* __restrict__ xtop,
    const int* __restrict__ qmap, const int* __restrict__ need) {
    __shared__ double w1s[kEnergyRows][kEnergyChunk];
    __shared__ double w2s[kEnergyRows][kEnergyChunk];
    const int g0 = blockIdx.y * kEnergyRows;
    int kmax = Ng - 1, gmax = Ng - 1;
    if (xtop) {  // groups above the source's top are never read; A rows above kmax are zero
        kmax = max(-1, min(Ng - 1, *xtop));
        gmax = qmap[kmax + 1];
    }
    if (g0 > gmax) return;
    const int rows = min(kEnergyRows, gmax - g0 + 1);
    int kb = Ng, ke = 0;  // union of the rows' bands, below kmax + 1
    for (int j = 0; j < rows; ++j) {
        const int b = max(0, beg[g0 + j]), e = min(kmax + 1, end[g0 + j]);
        if (b < e) {
            kb = min(kb, b);
            ke = max(ke, e);
        }
    }
    const long long lane = static_cast<long long>(blockIdx.x) * 256 + threadIdx.x;
    const bool live = lane < lanes;
    double q1[kEnergyRows], q2[kEnergyRows];
#pragma unroll
    for (int j = 0; j < kEnergyRows; ++j) q1[j] = q2[j] = 0.0;
    for (int c0 = kb; c0 < ke; c0 += kEnergyChunk) {
        const int cn = min(kEnergyChunk, ke - c0);
        for (int e = threadIdx.x; e < kEnergyRows * kEnergyChunk; e += 256) {
            const int j = e / kEnergyChunk, c = e % kEnergyChunk;
            const bool in = j < rows && c < cn;
            const long long off = static_cast<long long>(g0 + j) * Ng + c0 + c;
            w1s[j][c] = in ? e1[off] : 0.0;
            w2s[j][c] = in ? e2[off] : 0.0;
        }
        __syncthreads();
        if (live) {
            for (int c = 0; c < cn; ++c) {
                if (need && !need[c0 + c]) continue;  // A row unwritten; its terms are zeros
                const long long idx = static_cast<long long>(c0 + c) * lanes + lane;
                const double a1 = static_cast<double>(A1[idx]);
                const double a2 = static_cast<double>(A2[idx]);
#pragma unroll
                for (int j = 0; j < kEnergyRows; ++j) {
                    q1[j] += w1s[j][c] * a1;
                    q2[j] += w2s[j][c] * a2;
                }
            }
        }
        __syncthreads();
    }
    if (!live) return;
#pragma unroll
    for (int j = 0; j < kEnergyRows; ++j)
        if (j < rows)
            Q[static_cast<long long>(g0 + j) * lanes + lane] =
                static_cast<T>(dg * (q1[j] + q2[j] / 3.0));
}

}  // namespace

int source_nap(int nang) { return (nang + 31) / 32 * 32; }

void source_angular_blocks(const std::vector<double>& kv, int Ng, int nang, std::vector<int>& count,
                           std::vector<int>& list) {
    const int mtiles = (nang + kTile - 1) / kTile, ktiles = (nang + kDepth - 1) / kDepth;
    count.assign(static_cast<size_t>(Ng) * mtiles, 0);
    list.assign(static_cast<size_t>(Ng) * mtiles * ktiles, 0);
    for (int g = 0; g < Ng; ++g)
        for (int mt = 0; mt < mtiles; ++mt) {
            const size_t t = static_cast<size_t>(g) * mtiles + mt;
            for (int kt = 0; kt < ktiles; ++kt) {
                bool nz = false;
                for (int k = kt * kDepth; k < std::min(nang, (kt + 1) * kDepth) && !nz; ++k) {
                    const double* row = kv.data() + (static_cast<size_t>(g) * nang + k) * nang;
                    for (int m = mt * kTile; m < std::min(nang, (mt + 1) * kTile); ++m)
                        if (row[m] != 0.0) {
                            nz = true;
                            break;
                        }
                }
                if (nz) list[t * ktiles + count[t]++] = kt;
            }
        }
}
std::vector<int> source_needed_groups(const std::vector<int>& count, const std::vector<int>& beg,
                                      const std::vector<int>& end, int Ng, int nang) {
    const int mtiles = (nang + kTile - 1) / kTile;
    std::vector<int> need(Ng, 0);
    for (int r = 0; r < Ng; ++r)
        for (int g = std::max(0, beg[r]); g < std::min(Ng, end[r]); ++g) need[g] = 1;
    for (int g = 0; g < Ng; ++g) {
        bool nz = false;
        for (int mt = 0; mt < mtiles && !nz; ++mt) nz = count[static_cast<size_t>(g) * mtiles + mt] > 0;
        need[g] = need[g] && nz;
    }
    return need;
}
int source_mtiles(int nang) { return (nang + 63) / 64; }
int source_march_slots() { return kMarchSlots; }
int source_ktiles(int nang) { return source_nap(nang) / 32; }

void launch_kv_block_flags(const __bf16* kth, unsigned char* flags, int Ng, int nang,
                           hipStream_t s) {
    const dim3 grid(source_ktiles(nang), source_mtiles(nang), Ng);
    hipLaunchKernelGGL(kv_block_flags_kernel, grid, dim3(256), 0, s, kth, flags, source_nap(nang));
    BFPN_HIP(hipGetLastError());
}
int source_ngp(int Ng) { return (Ng + 31) / 32 * 32; }
int source_erows(int Ng) { return (Ng + kBandRows - 1) / kBandRows * kBandRows; }

void launch_split_transpose(const float* kv, __bf16* hi, __bf16* lo, int Ng, int nang,
                            hipStream_t s) {
    const int nap = source_nap(nang);
    dim3 grid((nap * nap + 255) / 256, Ng);
    hipLaunchKernelGGL(split_transpose_kernel, grid, dim3(256), 0, s, kv, hi, lo, nang, nap);
    BFPN_HIP(hipGetLastError());
}

void launch_split(const float* src, __bf16* hi, __bf16* lo, long long n, hipStream_t s) {
    if (n <= 0) return;
    hipLaunchKernelGGL(split_kernel, dim3(static_cast<unsigned>((n + 255) / 256)), dim3(256), 0, s,
                       src, hi, lo, n);
    BFPN_HIP(hipGetLastError());
}

void launch_source_mfma(const float* Fp, const float* fp, float* A1, float* A2, float* Q,
                        const SourcePlan& plan, int Ng, int nyz, int nang, hipStream_t s,
                        hipEvent_t inputs_read, const int* xtop, const int* qmap, int q_rows) {
    BFPN_REQUIRE((xtop == nullptr) == (qmap == nullptr), "source tops need both xtop and qmap");
    const int nap = source_nap(nang);
    // tiles over the source groups whose A rows the band GEMM reads
    const long long t1 =
        static_cast<long long>(plan.ngroups) * ((nyz + 63) / 64) * ((nang + 63) / 64);
    BFPN_REQUIRE(t1 < (1LL << 31) - 8, "source GEMM grid too large");
    const int g1 = xcd_grid(t1);
    if (g1 > 0) {
        if (nang % 8 == 0)
            hipLaunchKernelGGL(source_angular_mfma<true>, dim3(g1), dim3(256), 0, s, Fp, fp,
                               plan.kth, plan.ktl, plan.kt_count, plan.kt_list, plan.glist,
                               plan.gcount, A1, A2, Ng, nyz, nang, nap, static_cast<int>(t1),
                               xtop);
        else
            hipLaunchKernelGGL(source_angular_mfma<false>, dim3(g1), dim3(256), 0, s, Fp, fp,
                               plan.kth, plan.ktl, plan.kt_count, plan.kt_list, plan.glist,
                               plan.gcount, A1, A2, Ng, nyz, nang, nap, static_cast<int>(t1),
                               xtop);
        BFPN_HIP(hipGetLastError());
    }
    if (inputs_read) BFPN_HIP(hipEventRecord(inputs_read, s));
    const long long lanes = static_cast<long long>(nyz) * nang;
    if (plan.march) {
        const long long lb = (lanes + kMarchLanes - 1) / kMarchLanes;
        BFPN_REQUIRE(lb < (1LL << 31) - 8, "source band grid too large");
        const int gm = xcd_grid(lb);
        if (lanes % 4 == 0)
            hipLaunchKernelGGL(source_band_march<true>, dim3(gm), dim3(512), 0, s, A1, A2,
                               plan.e1h, plan.e1l, plan.e2h, plan.e2l, plan.tile_band,
                               plan.march_steps, plan.march_first, plan.march_last, plan.amask, Q,
                               Ng, source_ngp(Ng), lanes, static_cast<int>(lb), xtop, qmap,
                               q_rows);
        else
            hipLaunchKernelGGL(source_band_march<false>, dim3(gm), dim3(512), 0, s, A1, A2,
                               plan.e1h, plan.e1l, plan.e2h, plan.e2l, plan.tile_band,
                               plan.march_steps, plan.march_first, plan.march_last, plan.amask, Q,
                               Ng, source_ngp(Ng), lanes, static_cast<int>(lb), xtop, qmap,
                               q_rows);
        BFPN_HIP(hipGetLastError());
        return;
    }
    const long long t2 =
        ((lanes + kBandLanes - 1) / kBandLanes) * ((Ng + kBandRows - 1) / kBandRows);
    BFPN_REQUIRE(t2 < (1LL << 31) - 8, "source band grid too large");
    const int g2 = xcd_grid(t2);
    if (lanes % 4 == 0)
        hipLaunchKernelGGL(source_band_mfma<true>, dim3(g2), dim3(256), 0, s, A1, A2, plan.e1h,
                           plan.e1l, plan.e2h, plan.e2l, plan.tile_band, plan.amask, Q, Ng,
                           source_ngp(Ng), lanes, static_cast<int>(t2), xtop, qmap);
    else
        hipLaunchKernelGGL(source_band_mfma<false>, dim3(g2), dim3(256), 0, s, A1, A2, plan.e1h,
                           plan.e1l, plan.e2h, plan.e2l, plan.tile_band, plan.amask, Q, Ng,
                           source_ngp(Ng), lanes, static_cast<int>(t2), xtop, qmap);
    BFPN_HIP(hipGetLastError());
}

template <typename T>
void launch_source_angular(const T* Fp, const T* fp, const T* kv, T* A1, T* A2, int Ng, int nyz,
                           int nang, hipStream_t s, const int* xtop, const int* kcount,
                           const int* klist, const int* need) {
    if (Ng <= 0 || nyz <= 0 || nang <= 0) return;
    dim3 grid((nang + kTile - 1) / kTile, (nyz + kTile - 1) / kTile, Ng);
    hipLaunchKernelGGL(source_angular_kernel<T>, grid, dim3(256), 0, s, Fp, fp, kv, A1, A2, nyz,
                       nang, xtop, kcount, klist, need);
    BFPN_HIP(hipGetLastError());
}

template <typename T>
void launch_source_energy(const T* A1, const T* A2, const double* e1, const double* e2,
                          const int* beg, const int* end, T* Q, int Ng, long long lanes, double dg,
                          hipStream_t s, const int* xtop, const int* qmap, const int* need) {
    if (Ng <= 0 || lanes <= 0) return;
    BFPN_REQUIRE((xtop == nullptr) == (qmap == nullptr), "source tops need both xtop and qmap");
    dim3 grid(static_cast<unsigned>((lanes + 255) / 256), (Ng + kEnergyRows - 1) / kEnergyRows);
    hipLaunchKernelGGL(source_energy_kernel<T>, grid, dim3(256), 0, s, A1, A2, e1, e2, beg, end, Q,
                       Ng, lanes, dg, xtop, qmap, need);
    BFPN_HIP(hipGetLastError());
}

template void launch_source_angular<float>(const float*, const float*, const float*, float*, float*,
                                           int, int, int, hipStream_t, const int*, const int*,
                                           const int*, const int*);
template void launch_source_angular<double>(const double*, const double*, const double*, double*,
                                            double*, int, int, int, hipStream_t, const int*,
                                            const int*, const int*, const int*);
template void launch_source_energy<float>(const float*, const float*, const double*, const double*,
                                          const int*, const int*, float*, int, long long, double,
                                          hipStream_t, const int*, const int*, const int*);
template void launch_source_energy<double>(const double*, const double*, const double*,
                                           const double*, const int*, const int*, double*, int,
                                           long long, double, hipStream_t, const int*, const int*,
                                           const int*);

}  // namespace bfpn
