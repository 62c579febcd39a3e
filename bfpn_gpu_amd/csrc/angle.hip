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

namespace bfpn {
namespace {

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
