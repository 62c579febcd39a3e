// angle.hip — Crank-Nicolson angular diffusion on the zero-boundary DST-I basis.
//
// Reference: AngleDiffusionSolver::solveEq19 (src/angle_diffusion.cu:420-483):
// factor build, scale, and four separable kernels (FwdU, FwdV*factor, InvV,
// InvU; src/angle_diffusion.cu:254-362), each a full pass over HBM with two
// scratch arrays.
//
// MI355X design: one launch, one read and one write of the state per call.
// The whole sandwich
//     Y = Psi Su^T,  Z = coeff_g (.) (Sv Y),  W = Sv^T Z,  Psi' = W Su
// runs per (g, s) plane of Nmu x Nom values, in place.  The CN factors
// (ratio * DST normalisation) are built once on the host per (g, direction).
// FP32 state: matrix cores, stages chained in registers (below).  FP64 state
// or planes wider than 32: vector kernel with the plane and bases in LDS.

#include "buffer.hpp"
#include "kernels.hpp"
#include "mfma.hpp"

namespace bfpn {
namespace {

// ---------------------------------------------------------- FP32 / MFMA ----
// One wave per (g, s) plane (each wave walks a contiguous run of planes), the
// plane zero padded to 32 x 32, every stage a 32x32x32 product on the matrix
// cores in split FP16 with FP32 accumulation (MT / NT = 16-row tiles along v
// and a): the state and the constant bases in two FP16 terms each, three
// products kept.  The DST runs every depth step on the same beam, so its
// error compounds: with two-term BF16 splits the FP32 IDD drifts ~1 % above
// the FP64 one over 1600 steps; two FP16 terms carry 22 significand bits,
// and the operator is about as exact as FP32 arithmetic.  FP16 needs a range: each
// plane is scaled by a power of two s so that max|X s| lies in [16, 32),
// the bases are stored times 16, so every intermediate stays below 2^14
// (|Y|, |W| <= 21 max|X| for the sine bases, |Z| <= 1.6 max|X| with the CN
// normalisation) and small values stay clear of the FP16 subnormals; the
// CN factors absorb 1/256 and the result is multiplied by 1/(256 s), all
// exact.
// The stages chain in registers, with no LDS: a stage's output tile, lane
// holding D[4q+r][i], is the next stage's K operand when the K axis is taken
// in the order  slot 8q + 4u + r  <->  index 16u + 4q + r  (chain_index), and
// the constant sine bases on the other side are stored with their columns in
// that order.  The sandwich is evaluated as
//     Y  = X Su^T            (A = X from HBM,   B = Su)
//     Z  = C (.) (Sv Y)      (A = Sv chained,   B = Y)
//     W^T = Z^T Sv           (A = Z^T,          B = Sv chained)
//     X'^T = Su^T W^T        (A = Su chained,   B = W)
// using the symmetry of the DST-I bases; X'^T leaves the last stage with 4
// consecutive elements of a plane row per lane and is stored straight back.
// V4: Nom % 4 == 0 -- rows are read and written as 16-byte vectors.
template <int MT, int NT, bool V4>
__global__ __launch_bounds__(256) void angle_mfma_kernel(float* __restrict__ F,
                                                         const float* __restrict__ factors,
                                                         AnglePlan plan, int planes, int per_wave,
                                                         int nyz, int Nmu, int Nom, float one,
                                                         const int* __restrict__ top) {
    // Only the planes of groups up to the state's top are nonzero (the
    // others stay zero): the first `active` planes, shared out evenly over
    // the grid's waves; blocks without any exit before touching LDS.
    const int active = top ? min(planes, (min(*top, planes / nyz - 1) + 1) * nyz) : planes;
    const int pw = top ? (active + gridDim.x * 4 - 1) / (gridDim.x * 4) : per_wave;
    if (static_cast<int>(blockIdx.x) * 4 * pw >= active) return;
    // The six constant basis arrays (Su, Su chained, Sv chained; hi, lo)
    // live in LDS, shared by the work-group's four waves and read per stage:
    // held in registers they would cost 48 VGPRs and one of the four waves
    // per SIMD.  Rows padded to 40 values (80 bytes) against bank conflicts.
    __shared__ __attribute__((aligned(16))) _Float16 cst[6][32][40];
    {
        const _Float16* src[6] = {plan.su_h, plan.su_l, plan.sup_h, plan.sup_l, plan.svp_h,
                                  plan.svp_l};
        for (int e = threadIdx.x; e < 6 * 32 * 4; e += blockDim.x) {  // 16-byte chunks
            const int a = e >> 7, r = (e >> 2) & 31, c = (e & 3) * 8;
            *reinterpret_cast<f16x8*>(&cst[a][r][c]) =
                *reinterpret_cast<const f16x8*>(src[a] + r * 32 + c);
        }
    }
    __syncthreads();
    const int lane = threadIdx.x & 63, i = lane & 15, q = lane >> 4;
    // (the wave's first plane, made visibly wave-uniform for the compiler)
    int p = __builtin_amdgcn_readfirstlane((blockIdx.x * 4 + (threadIdx.x >> 6)) * pw);
    const int p_end = min(active, p + pw);
    if (p >= p_end) return;
    const int na = Nmu * Nom;
    // The wave's run of planes as a buffer resource: a plane is a uniform
    // offset, the lane's elements fixed 32-bit offsets (V4: two 16-byte loads
    // per row tile, and one 16-byte store per output tile, kOutside where
    // the tile lies outside the plane).
    const int p0 = p;
    const __amdgpu_buffer_rsrc_t Fw =
        make_rsrc(F + static_cast<long long>(p0) * na,
                  static_cast<unsigned>(static_cast<long long>(p_end - p0) * na * sizeof(float)));
    unsigned ld_off[MT][2], st_off[MT][NT];
#pragma unroll
    for (int mt = 0; mt < MT; ++mt) {
#pragma unroll
        for (int h = 0; h < 2; ++h)
            ld_off[mt][h] = static_cast<unsigned>(
                (min(mt * 16 + i, Nmu - 1) * Nom + min(8 * q + 4 * h, max(Nom - 4, 0))) * sizeof(float));
#pragma unroll
        for (int nt = 0; nt < NT; ++nt) {
            const int v = mt * 16 + i, a0 = nt * 16 + 4 * q;
            st_off[mt][nt] = (v < Nmu && a0 < Nom) ? static_cast<unsigned>((v * Nom + a0) * sizeof(float))
                                                   : kOutside;
        }
    }
    // Fragment of constant `a` (0 Su, 1 Su chained, 2 Sv chained), tile t,
    // at compile-time offsets from the lane's base (row i, columns 8q..).
    auto frag = [&](const _Float16* base, int a, int t) {
        constexpr int kArr = 32 * 40, kTile = 16 * 40;
        return HalfFrag2{*reinterpret_cast<const f16x8*>(base + (2 * a) * kArr + t * kTile),
                         *reinterpret_cast<const f16x8*>(base + (2 * a + 1) * kArr + t * kTile)};
    };

    int g_cur = -1;
    float cf[MT][NT][4];
    int g = p / nyz, s_in = p - g * nyz;  // group and node of plane p (uniform)
    for (; p < p_end; ++p, (++s_in == nyz) ? (s_in = 0, ++g) : 0) {
        // an opaque zero per plane keeps the constant reads inside the loop
        // (`one`, a kernel argument equal to 1, keeps the residuals of the
        // splits in stages 2-4 single FMAs)
        int oz;
        asm volatile("v_mov_b32 %0, 0" : "=v"(oz));
        const _Float16* cb = &cst[0][i][8 * q] + oz;
        float* X = F + static_cast<long long>(p) * na;
        const unsigned so = static_cast<unsigned>((p - p0) * na * static_cast<int>(sizeof(float)));
        if (g != g_cur) {  // CN factors of this group, in the D layout, / 256
            g_cur = g;
            const float* C = factors + static_cast<long long>(g) * na;
            constexpr float kUnscale = 1.0f / (kAngleBasisScale * kAngleBasisScale);
#pragma unroll
            for (int mt = 0; mt < MT; ++mt)
#pragma unroll
                for (int nt = 0; nt < NT; ++nt)
#pragma unroll
                    for (int r = 0; r < 4; ++r) {
                        const int nu = mt * 16 + 4 * q + r, mu = nt * 16 + i;
                        cf[mt][nt][r] = (nu < Nmu && mu < Nom) ? C[nu * Nom + mu] * kUnscale : 0.f;
                    }
        }
        // stage 1: Y[v][mu] = sum_a X[v][a] Su[mu][a].  The padding rows and
        // columns of the 32 x 32 tile load clamped copies of plane values:
        // the bases' zero padding multiplies them by exact zeros (columns
        // a >= Nom here, rows v >= Nmu in stage 2), so no lane is masked.
        float x[MT][8];
#pragma unroll
        for (int mt = 0; mt < MT; ++mt) {
            const float* row = X + min(mt * 16 + i, Nmu - 1) * Nom;
            if (V4) {
#pragma unroll
                for (int h = 0; h < 2; ++h) {
                    const float4 t = bload4(Fw, ld_off[mt][h], so);
                    x[mt][4 * h] = t.x;
                    x[mt][4 * h + 1] = t.y;
                    x[mt][4 * h + 2] = t.z;
                    x[mt][4 * h + 3] = t.w;
                }
            } else {
#pragma unroll
                for (int e = 0; e < 8; ++e) x[mt][e] = row[min(8 * q + e, Nom - 1)];
            }
        }
        // the plane's scale s = 2^(4 - floor(log2 max|X|)) (biased exponent
        // clamped so that s and 1/s are normal numbers), uniform over the wave
        float mx = 0.f;
#pragma unroll
        for (int mt = 0; mt < MT; ++mt)
#pragma unroll
            for (int e = 0; e < 8; ++e) mx = fmaxf(mx, fabsf(x[mt][e]));
        const int ebits = __float_as_int(wave_max_nonneg(mx)) >> 23;
        const int sb = min(max(258 - ebits, 1), 253);
        const float sc = __int_as_float(sb << 23);
        const float unscale = __int_as_float((254 - sb) << 23) / (kAngleBasisScale * kAngleBasisScale);
        HalfFrag2 xa[MT];
#pragma unroll
        for (int mt = 0; mt < MT; ++mt) xa[mt] = split8_h<true>(x[mt], sc);
        f32x4 d1[MT][NT];
#pragma unroll
        for (int mt = 0; mt < MT; ++mt)
#pragma unroll
            for (int nt = 0; nt < NT; ++nt) d1[mt][nt] = mma_sc(xa[mt], frag(cb, 0, nt), zero4());
        // stage 2: Z[nu][mu] = C[nu][mu] sum_v Sv[nu][v] Y[v][mu]
        f32x4 d2[MT][NT];
#pragma unroll
        for (int nt = 0; nt < NT; ++nt) {
            float y[8];
#pragma unroll
            for (int r = 0; r < 4; ++r) {
                y[r] = d1[0][nt][r];
                y[4 + r] = MT > 1 ? d1[MT - 1][nt][r] : 0.f;
            }
            const HalfFrag2 b = split8_h<false>(y, one);
#pragma unroll
            for (int mt = 0; mt < MT; ++mt) {
                d2[mt][nt] = mma_cs(frag(cb, 2, mt), b, zero4());
#pragma unroll
                for (int r = 0; r < 4; ++r) d2[mt][nt][r] *= cf[mt][nt][r];
            }
        }
        // stage 3: W^T[mu][v] = sum_nu Z[nu][mu] Sv[nu][v]
        f32x4 d3[NT][MT];
#pragma unroll
        for (int nt = 0; nt < NT; ++nt) {
            float z[8];
#pragma unroll
            for (int r = 0; r < 4; ++r) {
                z[r] = d2[0][nt][r];
                z[4 + r] = MT > 1 ? d2[MT - 1][nt][r] : 0.f;
            }
            const HalfFrag2 a = split8_h<false>(z, one);
#pragma unroll
            for (int mt = 0; mt < MT; ++mt) d3[nt][mt] = mma_sc(a, frag(cb, 2, mt), zero4());
        }
        // stage 4, transposed: X'^T[a][v] = sum_mu Su[mu][a] W[v][mu] (operands
        // swapped), so each lane holds 4 consecutive a of one row v; in place
#pragma unroll
        for (int mt = 0; mt < MT; ++mt) {
            float w[8];
#pragma unroll
            for (int r = 0; r < 4; ++r) {
                w[r] = d3[0][mt][r];
                w[4 + r] = NT > 1 ? d3[NT - 1][mt][r] : 0.f;
            }
            const HalfFrag2 b = split8_h<false>(w, one);
            const int v = mt * 16 + i;
#pragma unroll
            for (int nt = 0; nt < NT; ++nt) {
                f32x4 d = mma_cs(frag(cb, 1, nt), b, zero4());
#pragma unroll
                for (int r = 0; r < 4; ++r) d[r] *= unscale;
                const int a0 = nt * 16 + 4 * q;
                if (V4) {
                    bstore4(make_float4(d[0], d[1], d[2], d[3]), Fw, st_off[mt][nt], so);
                } else {
#pragma unroll
                    for (int r = 0; r < 4; ++r)
                        if (v < Nmu && a0 + r < Nom) X[v * Nom + a0 + r] = d[r];
                }
            }
        }
    }
}

// --------------------------------------------------- vector (FP64, wide) ---

constexpr int kThreads = 256;
constexpr int kMaxAngles = 1024;  // Nmu*Nom per plane held in LDS

template <typename T>
__global__ __launch_bounds__(kThreads) void angle_kernel(T* __restrict__ F,
                                                         const T* __restrict__ factors,
                                                         const T* __restrict__ su,
                                                         const T* __restrict__ sv, long long planes,
                                                         int nyz, int Nmu, int Nom, int per_block,
                                                         const int* __restrict__ top) {
    extern __shared__ unsigned char lds_raw[];
    // planes of groups above the state's top are zero and stay zero
    const long long active =
        top ? min(planes, static_cast<long long>(min(*top, static_cast<int>(planes / nyz) - 1) + 1) * nyz)
            : planes;
    const long long p0 = static_cast<long long>(blockIdx.x) * per_block;
    if (p0 >= active) return;
    T* lds = reinterpret_cast<T*>(lds_raw);
    const int na = Nmu * Nom;
    T* Su = lds;                    // Nom*Nom, Su[mu][a]
    T* Sv = Su + Nom * Nom;         // Nmu*Nmu, Sv[nu][v]
    T* X = Sv + Nmu * Nmu;          // per_block * na
    T* Y = X + per_block * na;      // per_block * na
    for (int e = threadIdx.x; e < Nom * Nom; e += kThreads) Su[e] = su[e];
    for (int e = threadIdx.x; e < Nmu * Nmu; e += kThreads) Sv[e] = sv[e];
    const int np = static_cast<int>(min(static_cast<long long>(per_block), active - p0));
    const int n = np * na;
    T* base = F + p0 * na;
    for (int e = threadIdx.x; e < n; e += kThreads) X[e] = base[e];
    __syncthreads();
    // Y[p][v][mu] = sum_a X[p][v][a] Su[mu][a]
    for (int e = threadIdx.x; e < n; e += kThreads) {
        const int p = e / na, r = e - p * na, v = r / Nom, mu = r - v * Nom;
        const T* x = X + p * na + v * Nom;
        const T* s = Su + mu * Nom;
        T acc = T(0);
        for (int a = 0; a < Nom; ++a) acc += x[a] * s[a];
        Y[e] = acc;
    }
    __syncthreads();
    // X[p][nu][mu] = coeff[g][nu][mu] * sum_v Sv[nu][v] Y[p][v][mu]
    for (int e = threadIdx.x; e < n; e += kThreads) {
        const int p = e / na, r = e - p * na, nu = r / Nom, mu = r - nu * Nom;
        const T* y = Y + p * na + mu;
        const T* s = Sv + nu * Nmu;
        T acc = T(0);
        for (int v = 0; v < Nmu; ++v) acc += s[v] * y[v * Nom];
        const long long g = (p0 + p) / nyz;
        X[e] = acc * factors[g * na + r];
    }
    __syncthreads();
    // Y[p][v][mu] = sum_nu Sv[nu][v] X[p][nu][mu]
    for (int e = threadIdx.x; e < n; e += kThreads) {
        const int p = e / na, r = e - p * na, v = r / Nom, mu = r - v * Nom;
        const T* z = X + p * na + mu;
        T acc = T(0);
        for (int nu = 0; nu < Nmu; ++nu) acc += Sv[nu * Nmu + v] * z[nu * Nom];
        Y[e] = acc;
    }
    __syncthreads();
    // F[p][v][a] = sum_mu Y[p][v][mu] Su[mu][a]
    for (int e = threadIdx.x; e < n; e += kThreads) {
        const int p = e / na, r = e - p * na, v = r / Nom, a = r - v * Nom;
        const T* w = Y + p * na + v * Nom;
        T acc = T(0);
        for (int mu = 0; mu < Nom; ++mu) acc += w[mu] * Su[mu * Nom + a];
        base[e] = acc;
    }
}

}  // namespace

void launch_angle_mfma(float* F, const float* factors, const AnglePlan& plan, int Ng, int nyz,
                       int Nmu, int Nom, hipStream_t s, const int* top) {
    BFPN_REQUIRE(Nmu >= 1 && Nom >= 1 && Nmu <= 32 && Nom <= 32, "MFMA DST needs Nmu, Nom <= 32");
    const long long planes = static_cast<long long>(Ng) * nyz;
    if (planes == 0) return;
    BFPN_REQUIRE(planes * Nmu * Nom < (1LL << 31) && planes < (1LL << 30),
                 "MFMA DST indexes planes with 32-bit offsets");
    // 8192 waves (32 per CU) walking contiguous runs of planes
    const int waves = static_cast<int>(std::min<long long>(planes, 8192));
    const int per_wave = static_cast<int>((planes + waves - 1) / waves);
    const int blocks = (static_cast<int>((planes + per_wave - 1) / per_wave) + 3) / 4;
    const dim3 grid(blocks), block(256);
    const int mt = Nmu > 16 ? 2 : 1, nt = Nom > 16 ? 2 : 1;
#define BFPN_ANGLE(MT_, NT_, V4_)                                                              \
    hipLaunchKernelGGL((angle_mfma_kernel<MT_, NT_, V4_>), grid, block, 0, s, F, factors, plan,\
                       static_cast<int>(planes), per_wave, nyz, Nmu, Nom, 1.0f, top)
#define BFPN_ANGLE_V(MT_, NT_)                                                                 \
    if (Nom % 4 == 0) BFPN_ANGLE(MT_, NT_, true);                                              \
    else BFPN_ANGLE(MT_, NT_, false)
    if (mt == 2 && nt == 2) { BFPN_ANGLE_V(2, 2); }
    else if (mt == 2) { BFPN_ANGLE_V(2, 1); }
    else if (nt == 2) { BFPN_ANGLE_V(1, 2); }
    else { BFPN_ANGLE_V(1, 1); }
#undef BFPN_ANGLE_V
#undef BFPN_ANGLE
    BFPN_HIP(hipGetLastError());
}

template <typename T>
void launch_angle_full(T* F, const T* factors, const T* su, const T* sv, int Ng, int nyz, int Nmu,
                       int Nom, hipStream_t s, const int* top) {
    const int na = Nmu * Nom;
    BFPN_REQUIRE(na > 0 && na <= kMaxAngles, "Nmu*Nom must be in [1, 1024]");
    const long long planes = static_cast<long long>(Ng) * nyz;
    if (planes == 0) return;
    const int per_block = std::max(1, 1024 / na);
    const size_t lds = sizeof(T) * (static_cast<size_t>(Nom) * Nom + static_cast<size_t>(Nmu) * Nmu +
                                    2 * static_cast<size_t>(per_block) * na);
    const long long blocks = (planes + per_block - 1) / per_block;
    BFPN_REQUIRE(blocks < (1LL << 31), "too many angular planes");
    hipLaunchKernelGGL(angle_kernel<T>, dim3(static_cast<unsigned>(blocks)), dim3(kThreads), lds, s,
                       F, factors, su, sv, planes, nyz, Nmu, Nom, per_block, top);
    BFPN_HIP(hipGetLastError());
}

template void launch_angle_full<float>(float*, const float*, const float*, const float*, int, int,
                                       int, int, hipStream_t, const int*);
template void launch_angle_full<double>(double*, const double*, const double*, const double*, int,
                                        int, int, int, hipStream_t, const int*);

}  // namespace bfpn
