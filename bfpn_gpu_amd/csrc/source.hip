// source.hip — catastrophic secondary source Q1 from the primary state.
//
// Reference: secondarySourceFromPrimaryKernel (src/energy_solver.cu:287-333):
//   Q1[g,s,m] = dg * sum_{g'} sum_{m'} (K1[g,g'] max(Fp,0) + K2[g,g'] fp/3)[g',s,m'] Kv[g',m',m]
// evaluated per output cell as a (g', m') double loop (1.6k-4k terms/cell).
//
// MI355X design: the sum is factorised into
//   1. an angular contraction per source group, a dense GEMM
//        A1[g'] = max(Fp[g'],0) . Kv[g'],   A2[g'] = fp[g'] . Kv[g']
//      with rows = transverse nodes (nyz) and K = N = nang, tiled through LDS
//      (both products share the Kv tile); and
//   2. a banded energy sum Q = dg (sum K1 A1 + (sum K2 A2)/3) over the band
//      [beg_g, end_g) of non-zero source groups, coalesced over lanes.
// This is exactly the reference's sum regrouped; only rounding order differs.

#include "kernels.hpp"

namespace bfpn {
namespace {

constexpr int kTile = 64;   // output tile: 64 nodes x 64 directions
constexpr int kDepth = 16;  // contraction depth per LDS stage

// grid: (ceil(nang/64), ceil(nyz/64), Ng); 256 threads, each 4x4 outputs x2.
template <typename T>
__global__ __launch_bounds__(256) void source_angular_kernel(const T* __restrict__ Fp,
                                                             const T* __restrict__ fp,
                                                             const T* __restrict__ kv,
                                                             T* __restrict__ A1, T* __restrict__ A2,
                                                             int nyz, int nang) {
    __shared__ T x1s[kDepth][kTile + 1];
    __shared__ T x2s[kDepth][kTile + 1];
    __shared__ T ks[kDepth][kTile + 1];
    const int g = blockIdx.z;
    const int m0 = blockIdx.x * kTile;
    const int s0 = blockIdx.y * kTile;
    const long long plane = static_cast<long long>(g) * nyz * nang;
    const T* K = kv + static_cast<long long>(g) * nang * nang;
    const int tx = threadIdx.x & 15;  // output directions m0 + tx + 16*j
    const int ty = threadIdx.x >> 4;  // output nodes     s0 + ty + 16*i
    T acc1[4][4], acc2[4][4];
#pragma unroll
    for (int i = 0; i < 4; ++i)
#pragma unroll
        for (int j = 0; j < 4; ++j) acc1[i][j] = acc2[i][j] = T(0);

    for (int k0 = 0; k0 < nang; k0 += kDepth) {
        // stage x tiles: 64 nodes x 16 inputs (coalesced along the input index)
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
        // stage Kv tile: 16 inputs x 64 outputs (coalesced along outputs)
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

// grid: (ceil(lanes/256), Ng); Q[g][lane] over the band of row g.
template <typename T>
__global__ __launch_bounds__(256) void source_energy_kernel(
    const T* __restrict__ A1, const T* __restrict__ A2, const double* __restrict__ e1,
    const double* __restrict__ e2, const int* __restrict__ beg, const int* __restrict__ end,
    T* __restrict__ Q, int Ng, long long lanes, double dg) {
    const long long lane = static_cast<long long>(blockIdx.x) * 256 + threadIdx.x;
    const int g = blockIdx.y;
    if (lane >= lanes) return;
    const int b = max(0, beg[g]);
    const int e = min(Ng, end[g]);
    double q1 = 0.0, q2 = 0.0;
    const double* r1 = e1 + static_cast<long long>(g) * Ng;
    const double* r2 = e2 + static_cast<long long>(g) * Ng;
    for (int src = b; src < e; ++src) {
        const double w1 = r1[src], w2 = r2[src];
        if (w1 == 0.0 && w2 == 0.0) continue;
        const long long idx = static_cast<long long>(src) * lanes + lane;
        q1 += w1 * static_cast<double>(A1[idx]);
        q2 += w2 * static_cast<double>(A2[idx]);
    }
    Q[static_cast<long long>(g) * lanes + lane] = static_cast<T>(dg * (q1 + q2 / 3.0));
}

}  // namespace

template <typename T>
void launch_source_angular(const T* Fp, const T* fp, const T* kv, T* A1, T* A2, int Ng, int nyz,
                           int nang, hipStream_t s) {
    if (Ng <= 0 || nyz <= 0 || nang <= 0) return;
    dim3 grid((nang + kTile - 1) / kTile, (nyz + kTile - 1) / kTile, Ng);
    hipLaunchKernelGGL(source_angular_kernel<T>, grid, dim3(256), 0, s, Fp, fp, kv, A1, A2, nyz,
                       nang);
    BFPN_HIP(hipGetLastError());
}

template <typename T>
void launch_source_energy(const T* A1, const T* A2, const double* e1, const double* e2,
                          const int* beg, const int* end, T* Q, int Ng, long long lanes, double dg,
                          hipStream_t s) {
    if (Ng <= 0 || lanes <= 0) return;
    dim3 grid(static_cast<unsigned>((lanes + 255) / 256), Ng);
    hipLaunchKernelGGL(source_energy_kernel<T>, grid, dim3(256), 0, s, A1, A2, e1, e2, beg, end, Q,
                       Ng, lanes, dg);
    BFPN_HIP(hipGetLastError());
}

template void launch_source_angular<float>(const float*, const float*, const float*, float*, float*,
                                           int, int, int, hipStream_t);
template void launch_source_angular<double>(const double*, const double*, const double*, double*,
                                            double*, int, int, int, hipStream_t);
template void launch_source_energy<float>(const float*, const float*, const double*, const double*,
                                          const int*, const int*, float*, int, long long, double,
                                          hipStream_t);
template void launch_source_energy<double>(const double*, const double*, const double*,
                                           const double*, const int*, const int*, double*, int,
                                           long long, double, hipStream_t);

}  // namespace bfpn
