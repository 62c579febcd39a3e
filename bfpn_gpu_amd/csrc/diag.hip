// diag.hip — per-step diagnostics on the resident state: the paper's Eq.25
// energy moments and the Y-Z dose plane at the current depth.
//
// Reference: energyMomentKernel (src/bfpm_solver.cu:311-360, one block per
// group, tree sum in shared memory) and spotDosePlaneKernel (433-475, one
// thread per transverse node looping over all groups and directions).
//
// MI355X design:
//  * moments: a 2D grid (lane chunks x groups) so that every CU has work
//    (a block per group alone would leave most of the GPU idle on small
//    grids), FP64 partial sums per block, then one fixed-order pass per
//    group: deterministic, independent of the launch shape;
//  * dose plane: a work-group per transverse node; its threads read the
//    node's direction rows (contiguous, nang values) of every group, so the
//    reads are coalesced where the reference's per-thread loops stride by
//    nang.  The Eq.24 weights are the solver's per-group dose weights.

#include "kernels.hpp"

namespace bfpn {
namespace {

constexpr int kThreads = 256;
constexpr int kPerThread = 16;  // lanes summed per thread and chunk

__device__ inline double block_reduce(double v, double* lds) {
    for (int off = 32; off > 0; off >>= 1) v += __shfl_xor(v, off);
    const int wave = threadIdx.x >> 6;
    if ((threadIdx.x & 63) == 0) lds[wave] = v;
    __syncthreads();
    double t = 0.0;
    if (threadIdx.x == 0)
        for (int w = 0; w < kThreads / 64; ++w) t += lds[w];
    __syncthreads();
    return t;
}

// grid (chunks, Ng): partial sums of psi^1 = F + F1 and psi^2 = f + f1 over
// one chunk of a group's lanes -> part[(g * chunks + c) * 2 + {0, 1}]
template <typename T>
__global__ __launch_bounds__(kThreads) void moment_partials_kernel(
    const T* __restrict__ F, const T* __restrict__ f, const T* __restrict__ F1,
    const T* __restrict__ f1, long long per, double* __restrict__ part) {
    __shared__ double lds[kThreads / 64];
    const long long base = static_cast<long long>(blockIdx.y) * per;
    const long long c0 = static_cast<long long>(blockIdx.x) * kThreads * kPerThread;
    double s1 = 0.0, s2 = 0.0;
#pragma unroll 4
    for (int k = 0; k < kPerThread; ++k) {
        const long long l = c0 + k * kThreads + threadIdx.x;
        if (l < per) {
            s1 += static_cast<double>(F[base + l]) + static_cast<double>(F1[base + l]);
            s2 += static_cast<double>(f[base + l]) + static_cast<double>(f1[base + l]);
        }
    }
    s1 = block_reduce(s1, lds);
    s2 = block_reduce(s2, lds);
    if (threadIdx.x == 0) {
        const long long o = (static_cast<long long>(blockIdx.y) * gridDim.x + blockIdx.x) * 2;
        part[o] = s1;
        part[o + 1] = s2;
    }
}

// one thread per (group, moment): the chunks in order
__global__ void moment_finish_kernel(const double* __restrict__ part, int chunks, int n,
                                     double weight, double* __restrict__ out) {
    const int i = blockIdx.x * blockDim.x + threadIdx.x;  // i = 2 g + moment
    if (i >= n) return;
    const int g = i >> 1, mom = i & 1;
    double s = 0.0;
    for (int c = 0; c < chunks; ++c) s += part[(static_cast<long long>(g) * chunks + c) * 2 + mom];
    out[i] = s * weight;
}

// one work-group per transverse node s
template <typename T>
__global__ __launch_bounds__(kThreads) void spot_plane_kernel(
    const T* __restrict__ F, const T* __restrict__ f, const T* __restrict__ F1,
    const T* __restrict__ f1, const EnergyCoef* __restrict__ coef, int Ng, int nyz, int nang,
    double scale, double* __restrict__ plane) {
    __shared__ double lds[kThreads / 64];
    const int s = blockIdx.x;
    double acc = 0.0;
    for (int g = 0; g < Ng; ++g) {
        const double w1 = coef[g].w1, w2 = coef[g].w2;
        const long long base = (static_cast<long long>(g) * nyz + s) * nang;
        for (int a = threadIdx.x; a < nang; a += kThreads) {
            const long long c = base + a;
            const double p1 = static_cast<double>(F[c]) + static_cast<double>(F1[c]);
            const double p2 = static_cast<double>(f[c]) + static_cast<double>(f1[c]);
            acc += p1 * w1 + p2 * w2;
        }
    }
    acc = block_reduce(acc, lds);
    if (threadIdx.x == 0) plane[s] = acc * scale;
}

}  // namespace

size_t moment_partial_count(int Ng, long long per) {
    const long long chunks = (per + kThreads * kPerThread - 1) / (kThreads * kPerThread);
    return static_cast<size_t>(chunks) * Ng * 2;
}

template <typename T>
void launch_energy_moments(const T* F, const T* f, const T* F1, const T* f1, int Ng, long long per,
                           double weight, double* partials, double* out, hipStream_t s) {
    if (Ng <= 0 || per <= 0) return;
    const long long chunks = (per + kThreads * kPerThread - 1) / (kThreads * kPerThread);
    BFPN_REQUIRE(chunks < (1LL << 31) && Ng < 65536, "moment grid too large");
    hipLaunchKernelGGL(moment_partials_kernel<T>, dim3(static_cast<unsigned>(chunks), Ng),
                       dim3(kThreads), 0, s, F, f, F1, f1, per, partials);
    BFPN_HIP(hipGetLastError());
    const int n = 2 * Ng;
    hipLaunchKernelGGL(moment_finish_kernel, dim3((n + 255) / 256), dim3(256), 0, s, partials,
                       static_cast<int>(chunks), n, weight, out);
    BFPN_HIP(hipGetLastError());
}

template <typename T>
void launch_spot_plane(const T* F, const T* f, const T* F1, const T* f1, const EnergyCoef* coef,
                       int Ng, int nyz, int nang, double scale, double* plane, hipStream_t s) {
    if (Ng <= 0 || nyz <= 0 || nang <= 0) return;
    hipLaunchKernelGGL(spot_plane_kernel<T>, dim3(nyz), dim3(kThreads), 0, s, F, f, F1, f1, coef,
                       Ng, nyz, nang, scale, plane);
    BFPN_HIP(hipGetLastError());
}

template void launch_energy_moments<float>(const float*, const float*, const float*, const float*,
                                           int, long long, double, double*, double*, hipStream_t);
template void launch_energy_moments<double>(const double*, const double*, const double*,
                                            const double*, int, long long, double, double*,
                                            double*, hipStream_t);
template void launch_spot_plane<float>(const float*, const float*, const float*, const float*,
                                       const EnergyCoef*, int, int, int, double, double*,
                                       hipStream_t);
template void launch_spot_plane<double>(const double*, const double*, const double*, const double*,
                                        const EnergyCoef*, int, int, int, double, double*,
                                        hipStream_t);

}  // namespace bfpn
