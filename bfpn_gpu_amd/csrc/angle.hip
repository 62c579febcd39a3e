// angle.hip — Crank-Nicolson angular diffusion on the zero-boundary DST-I basis.
//
// Reference: AngleDiffusionSolver::solveEq19 (src/angle_diffusion.cu:420-483):
// factor build, scale, and four separable kernels (FwdU, FwdV*factor, InvV,
// InvU; src/angle_diffusion.cu:254-362), each a full pass over HBM with two
// scratch arrays.
//
// MI355X design: one launch.  A work-group owns a batch of (g, s) planes
// (Nmu x Nom values each), keeps the planes and both sine bases in LDS and
// runs the whole sandwich
//     Y = Psi Su^T,  Z = coeff_g (.) (Sv Y),  W = Sv^T Z,  Psi' = W Su
// between LDS buffers, writing the plane back in place: one read and one
// write of the state per call.  The CN factors (ratio * DST normalisation)
// are built once on the host per (g, direction).

#include "kernels.hpp"
#include "mfma.hpp"

namespace bfpn {
namespace {

// ---------------------------------------------------------- FP32 / MFMA ----
// One wave per (g, s) plane (grid-stride over planes).  The plane, padded to
// 32 x 32 with zeros, lives in a per-wave LDS tile; each stage is a 32x32x32
// product on the matrix cores (split-BF16, FP32 accumulation), and the LDS
// tile converts the D layout of one stage into the A/B layout of the next.
// The DST-I bases are symmetric (S[k][j] = S[j][k]), so Y = X Su^T uses Su
// as the B operand and W = Sv^T Z uses Sv as the A operand, like Sv Y.
// MT / NT: 16-row tiles along v (Nmu) and a (Nom), 1 or 2 (compile time, so
// the fragment arrays stay in registers).
template <int MT, int NT>
__global__ __launch_bounds__(256) void angle_mfma_kernel(
    float* __restrict__ F, const float* __restrict__ factors, const __bf16* __restrict__ su_h,
    const __bf16* __restrict__ su_l, const __bf16* __restrict__ sv_h,
    const __bf16* __restrict__ sv_l, long long planes, int nyz, int Nmu, int Nom) {
    __shared__ float tile[4][32][33];
    const int wave = threadIdx.x >> 6, lane = threadIdx.x & 63;
    const int i = lane & 15, q = lane >> 4;
    float(*B)[33] = tile[wave];
    const int na = Nmu * Nom;
    constexpr int mt_n = MT;  // tiles along v / nu
    constexpr int nt_n = NT;  // tiles along a / mu
    SplitFrag su[2], sv[2];
#pragma unroll
    for (int t = 0; t < 2; ++t) {
        const int off = (t * 16 + i) * 32 + 8 * q;
        su[t].hi = *reinterpret_cast<const bf16x8*>(su_h + off);
        su[t].lo = *reinterpret_cast<const bf16x8*>(su_l + off);
        sv[t].hi = *reinterpret_cast<const bf16x8*>(sv_h + off);
        sv[t].lo = *reinterpret_cast<const bf16x8*>(sv_l + off);
    }
    for (int e = lane; e < 32 * 32; e += 64) B[e >> 5][e & 31] = 0.f;

    auto row_frag = [&](int row) {  // A operand: B[row][8q .. 8q+7]
        float v[8];
#pragma unroll
        for (int e = 0; e < 8; ++e) v[e] = B[row][8 * q + e];
        return split8(v);
    };
    auto col_frag = [&](int col) {  // B operand: B[8q .. 8q+7][col]
        float v[8];
#pragma unroll
        for (int e = 0; e < 8; ++e) v[e] = B[8 * q + e][col];
        return split8(v);
    };
    auto store = [&](const f32x4 (&d)[2][2]) {
#pragma unroll
        for (int mt = 0; mt < mt_n; ++mt)
#pragma unroll
            for (int nt = 0; nt < nt_n; ++nt)
#pragma unroll
                for (int r = 0; r < 4; ++r) B[mt * 16 + 4 * q + r][nt * 16 + i] = d[mt][nt][r];
    };

    for (long long p = static_cast<long long>(blockIdx.x) * 4 + wave; p < planes;
         p += static_cast<long long>(gridDim.x) * 4) {
        float* X = F + p * na;
        const float* C = factors + (p / nyz) * na;
        for (int e = lane; e < na; e += 64) {
            const int v = e / Nom;
            B[v][e - v * Nom] = X[e];
        }
        f32x4 d[2][2];
        // stage 1: Y[v][mu] = sum_a X[v][a] Su[mu][a]
        SplitFrag a[2], b[2];
#pragma unroll
        for (int mt = 0; mt < mt_n; ++mt) a[mt] = row_frag(mt * 16 + i);
#pragma unroll
        for (int mt = 0; mt < mt_n; ++mt)
#pragma unroll
            for (int nt = 0; nt < nt_n; ++nt) d[mt][nt] = mma_split(a[mt], su[nt], zero4());
        store(d);
        // stage 2: Z[nu][mu] = C[nu][mu] sum_v Sv[nu][v] Y[v][mu]
#pragma unroll
        for (int nt = 0; nt < nt_n; ++nt) b[nt] = col_frag(nt * 16 + i);
#pragma unroll
        for (int mt = 0; mt < mt_n; ++mt)
#pragma unroll
            for (int nt = 0; nt < nt_n; ++nt) {
                d[mt][nt] = mma_split(sv[mt], b[nt], zero4());
                const int mu = nt * 16 + i;
#pragma unroll
                for (int r = 0; r < 4; ++r) {
                    const int nu = mt * 16 + 4 * q + r;
                    d[mt][nt][r] *= (nu < Nmu && mu < Nom) ? C[nu * Nom + mu] : 0.f;
                }
            }
        store(d);
        // stage 3: W[v][mu] = sum_nu Sv[nu][v] Z[nu][mu]
#pragma unroll
        for (int nt = 0; nt < nt_n; ++nt) b[nt] = col_frag(nt * 16 + i);
#pragma unroll
        for (int mt = 0; mt < mt_n; ++mt)
#pragma unroll
            for (int nt = 0; nt < nt_n; ++nt) d[mt][nt] = mma_split(sv[mt], b[nt], zero4());
        store(d);
        // stage 4: X'[v][a] = sum_mu W[v][mu] Su[mu][a]
#pragma unroll
        for (int mt = 0; mt < mt_n; ++mt) a[mt] = row_frag(mt * 16 + i);
#pragma unroll
        for (int mt = 0; mt < mt_n; ++mt)
#pragma unroll
            for (int nt = 0; nt < nt_n; ++nt) d[mt][nt] = mma_split(a[mt], su[nt], zero4());
        store(d);
        for (int e = lane; e < na; e += 64) {
            const int v = e / Nom;
            X[e] = B[v][e - v * Nom];
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
                                                         int nyz, int Nmu, int Nom, int per_block) {
    extern __shared__ unsigned char lds_raw[];
    T* lds = reinterpret_cast<T*>(lds_raw);
    const int na = Nmu * Nom;
    T* Su = lds;                    // Nom*Nom, Su[mu][a]
    T* Sv = Su + Nom * Nom;         // Nmu*Nmu, Sv[nu][v]
    T* X = Sv + Nmu * Nmu;          // per_block * na
    T* Y = X + per_block * na;      // per_block * na
    for (int e = threadIdx.x; e < Nom * Nom; e += kThreads) Su[e] = su[e];
    for (int e = threadIdx.x; e < Nmu * Nmu; e += kThreads) Sv[e] = sv[e];
    const long long p0 = static_cast<long long>(blockIdx.x) * per_block;
    const int np = static_cast<int>(min(static_cast<long long>(per_block), planes - p0));
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
                       int Nmu, int Nom, hipStream_t s) {
    BFPN_REQUIRE(Nmu >= 1 && Nom >= 1 && Nmu <= 32 && Nom <= 32, "MFMA DST needs Nmu, Nom <= 32");
    const long long planes = static_cast<long long>(Ng) * nyz;
    if (planes == 0) return;
    const long long blocks = std::min<long long>((planes + 3) / 4, 4096);
    const dim3 grid(static_cast<unsigned>(blocks)), block(256);
    const int mt = Nmu > 16 ? 2 : 1, nt = Nom > 16 ? 2 : 1;
#define BFPN_ANGLE(MT_, NT_)                                                                     \
    hipLaunchKernelGGL((angle_mfma_kernel<MT_, NT_>), grid, block, 0, s, F, factors, plan.su_h,  \
                       plan.su_l, plan.sv_h, plan.sv_l, planes, nyz, Nmu, Nom)
    if (mt == 2 && nt == 2) BFPN_ANGLE(2, 2);
    else if (mt == 2) BFPN_ANGLE(2, 1);
    else if (nt == 2) BFPN_ANGLE(1, 2);
    else BFPN_ANGLE(1, 1);
#undef BFPN_ANGLE
    BFPN_HIP(hipGetLastError());
}

template <typename T>
void launch_angle_full(T* F, const T* factors, const T* su, const T* sv, int Ng, int nyz, int Nmu,
                       int Nom, hipStream_t s) {
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
                       F, factors, su, sv, planes, nyz, Nmu, Nom, per_block);
    BFPN_HIP(hipGetLastError());
}

template void launch_angle_full<float>(float*, const float*, const float*, const float*, int, int,
                                       int, int, hipStream_t);
template void launch_angle_full<double>(double*, const double*, const double*, const double*, int,
                                        int, int, int, hipStream_t);

}  // namespace bfpn
