// misc.hip — dose reduction, deterministic final sums and beam initialisation.
//
// Reference: integratedDepthDoseKernel + reduceSumKernel + the 8-byte D2H per
// step (src/bfpm_solver.cu:199-309, 2132-2164) and the host-built Eq.23 beam
// copied H2D at init (src/bfpm_solver.cu:1978-2020).
//
// Here the per-step IDD normally comes out of the fused energy kernel; this
// stand-alone dose kernel serves the operator-level API.  Final sums run in
// one block in a fixed order, so results never depend on scheduling, and the
// IDD history stays on the device until the caller asks for it.

#include "kernels.hpp"

namespace bfpn {
namespace {

constexpr int kBlock = 256;

__device__ inline double block_sum256(double v, double* lds) {
    for (int off = 32; off > 0; off >>= 1) v += __shfl_xor(v, off);
    if ((threadIdx.x & 63) == 0) lds[threadIdx.x >> 6] = v;
    __syncthreads();
    double t = 0.0;
    if (threadIdx.x == 0)
        for (int w = 0; w < kBlock / 64; ++w) t += lds[w];
    return t;
}

// grid: (ceil(lanes/256)); each thread walks its lane through all groups.
template <typename T>
__global__ __launch_bounds__(kBlock) void dose_kernel(const T* __restrict__ F,
                                                      const T* __restrict__ f,
                                                      const T* __restrict__ F1,
                                                      const T* __restrict__ f1,
                                                      const DoseCoef* __restrict__ w,
                                                      const double* __restrict__ yzw, int Ng,
                                                      long long lanes, int nang,
                                                      double* __restrict__ partials) {
    __shared__ double lds[kBlock / 64];
    const long long lane = static_cast<long long>(blockIdx.x) * kBlock + threadIdx.x;
    double acc = 0.0;
    if (lane < lanes) {
        for (int g = 0; g < Ng; ++g) {
            const long long idx = g * lanes + lane;
            double p1 = static_cast<double>(F[idx]);
            double p2 = f ? static_cast<double>(f[idx]) : 0.0;
            if (F1) p1 += static_cast<double>(F1[idx]);
            if (f1) p2 += static_cast<double>(f1[idx]);
            acc += p1 * w[g].w1 + p2 * w[g].w2;
        }
        acc *= yzw[lane / nang];
    }
    const double t = block_sum256(acc, lds);
    if (threadIdx.x == 0) partials[blockIdx.x] = t;
}

__global__ __launch_bounds__(1024) void finish_sum_kernel(const double* __restrict__ a, int na,
                                                          const double* __restrict__ b, int nb,
                                                          double scale, double* __restrict__ out) {
    __shared__ double lds[1024];
    double s = 0.0;
    for (int k = threadIdx.x; k < na; k += 1024) s += a[k];
    if (b)
        for (int k = threadIdx.x; k < nb; k += 1024) s += b[k];
    lds[threadIdx.x] = s;
    __syncthreads();
    for (int off = 512; off > 0; off >>= 1) {
        if (threadIdx.x < off) lds[threadIdx.x] += lds[threadIdx.x + off];
        __syncthreads();
    }
    if (threadIdx.x == 0) *out = lds[0] * scale;
}

__global__ void sum_in_order_kernel(const double* __restrict__ v, int n, double* __restrict__ out) {
    double s = 0.0;
    for (int k = 0; k < n; ++k) s += v[k];
    *out = s;
}

// One node's directions are contiguous, so each block copies runs of nang
// values, in units of U (a 16-byte vector when nang allows, else one value);
// grid (ceil(n nang / (U kBlock)), Ng).
// MODE 0: gather, 1: scatter, 2: zero the listed nodes of dst.
template <typename T, typename U, int MODE>
__global__ __launch_bounds__(kBlock) void node_copy_kernel(T* __restrict__ dst, const T* __restrict__ src,
                                                           const int* __restrict__ nodes, int n,
                                                           long long nyz, int nang) {
    constexpr int V = sizeof(U) / sizeof(T);
    const long long idx = static_cast<long long>(blockIdx.x) * kBlock + threadIdx.x;  // unit
    const int nu = nang / V;                                                          // units per node
    const long long per = static_cast<long long>(n) * nu;
    if (idx >= per) return;
    const long long k = idx / nu, a = idx - k * nu, g = blockIdx.y;
    U* d = reinterpret_cast<U*>(dst);
    const U* s = reinterpret_cast<const U*>(src);
    const long long packed = g * per + idx;
    const long long full = (g * nyz + nodes[k]) * nu + a;
    if (MODE == 0)
        d[packed] = s[full];
    else if (MODE == 1)
        d[full] = s[packed];
    else
        d[full] = U{};
}

template <typename T>
__global__ __launch_bounds__(kBlock) void beam_kernel(T* __restrict__ F,
                                                      const double* __restrict__ space_angle,
                                                      const double* __restrict__ energy,
                                                      long long lanes) {
    const long long lane = static_cast<long long>(blockIdx.x) * kBlock + threadIdx.x;
    if (lane >= lanes) return;
    const double sa = space_angle[lane];
    const double e = energy[blockIdx.y];
    F[static_cast<long long>(blockIdx.y) * lanes + lane] = static_cast<T>(sa * e);
}

template <typename T>
__global__ __launch_bounds__(kBlock) void convert_kernel(const double* __restrict__ src,
                                                         T* __restrict__ dst, long long n) {
    const long long k = static_cast<long long>(blockIdx.x) * kBlock + threadIdx.x;
    if (k < n) dst[k] = static_cast<T>(src[k]);
}

// Plain copy (the traffic-counter calibration kernel and the reference rate
// for one read and one write of every byte): each work-group copies
// contiguous chunks of 1024 16-byte vectors, every thread four non-temporal
// loads in flight before their stores.  Of the variants measured on MI355X
// (grid-stride with one load in flight; chunks of 4 or 8 per thread, with
// and without non-temporal access; 2048 to 65536 work-groups; 1 and 4 GiB)
// this one was the fastest: 5.75-5.85 TB/s.
typedef float f32x4v_copy __attribute__((ext_vector_type(4)));
__global__ __launch_bounds__(kBlock) void copy16_kernel(const f32x4v_copy* __restrict__ src,
                                                        f32x4v_copy* __restrict__ dst,
                                                        long long n) {
    constexpr int U = 4;
    const long long pass = static_cast<long long>(gridDim.x) * kBlock * U;
    for (long long base = static_cast<long long>(blockIdx.x) * kBlock * U; base < n; base += pass) {
        f32x4v_copy v[U];
#pragma unroll
        for (int k = 0; k < U; ++k) {
            const long long i = base + k * kBlock + threadIdx.x;
            if (i < n) v[k] = __builtin_nontemporal_load(src + i);
        }
#pragma unroll
        for (int k = 0; k < U; ++k) {
            const long long i = base + k * kBlock + threadIdx.x;
            if (i < n) __builtin_nontemporal_store(v[k], dst + i);
        }
    }
}

// Per-group coefficient tables of the operator-level API, built on the device
// from device tables so that an operator call needs no host round trip.  Each
// mirrors the host builder in tables.cpp operation for operation, without
// contraction, so both give the same bits.
__device__ inline double max_like_std(double a, double b) { return (a < b) ? b : a; }  // std::max

__global__ __launch_bounds__(kBlock) void energy_coef_kernel(const double* __restrict__ S,
                                                             const double* __restrict__ sigma,
                                                             const double* __restrict__ Tc,
                                                             int straggling, double h, double dg,
                                                             double rho, int Ng,
                                                             EnergyCoef* __restrict__ out) {
#pragma clang fp contract(off)
    const int g = blockIdx.x * kBlock + threadIdx.x;
    if (g >= Ng) return;
    const double dg2 = max_like_std(dg * dg, kEps);
    const double sig = max_like_std(sigma[g], 0.0);
    const double Slo = max_like_std(S[g], 0.0);
    const double Shi = max_like_std(S[g + 1], 0.0);
    const bool st = straggling && Tc != nullptr;
    const double Tg = st ? max_like_std(Tc[g], 0.0) : 0.0;
    const double Thi = st && g + 1 < Ng ? max_like_std(Tc[g + 1], 0.0) : 0.0;
    const double Tlo = st && g > 0 ? max_like_std(Tc[g - 1], 0.0) : 0.0;
    const double inv_h = 1.0 / h, half_sig = 0.5 * sig;
    const double c_up = 0.5 * rho * Shi / dg, c_cur = 0.5 * rho * Slo / dg;
    const double inv_F = 1.0 / max_like_std(1.0 / h + 0.5 * rho * Slo / dg + 0.5 * sig, kEps);
    const double A = (1.5 * rho * Shi / dg) * inv_F;
    const double B = (0.5 * rho * Slo / dg) * inv_F;
    const double Cc = (0.5 * rho * Shi / dg) * inv_F;
    const double k_up = 1.5 * rho * Shi / dg, k_cur = 1.5 * rho * Slo / dg;
    const double k_F = 1.5 * rho * (Slo + Shi) / dg, k_f = 0.5 * rho * (Shi - Slo) / dg;
    const double H = k_up - A * 0.5 * rho * Shi / dg;
    const double coe = 1.0 / h + 0.5 * sig + 1.5 * rho * Slo / dg + 0.5 * rho * (Shi - Slo) / dg +
                       A * 0.5 * rho * Slo / dg;
    const double inv_coe = 1.0 / max_like_std(coe, kEps);
    const double a_hi = 0.5 * rho * Thi / dg2, a_lo = 0.5 * rho * Tlo / dg2, a_g = rho * Tg / dg2;
    const double cm[6] = {c_up + a_hi, -c_up, inv_h - half_sig - c_cur - a_g, c_cur, a_lo, 1.0};
    const double r[6] = {k_up - A * cm[0],
                         -k_up - A * cm[1],
                         (k_cur - k_F) - A * cm[2],
                         (inv_h - half_sig - k_cur - k_f) - A * cm[3],
                         -A * cm[4],
                         -A * cm[5]};
    EnergyCoef c;
#pragma unroll
    for (int i = 0; i < 6; ++i) {
        c.af[i] = inv_coe * r[i];
        c.aF[i] = B * c.af[i] + inv_F * cm[i];
    }
    c.af[6] = inv_coe * H;
    c.aF[6] = B * c.af[6] + Cc;
    c.w1 = c.w2 = 0.0;
    out[g] = c;
}

__global__ __launch_bounds__(kBlock) void dose_coef_kernel(const double* __restrict__ S,
                                                           const double* __restrict__ sigma,
                                                           const double* __restrict__ en,
                                                           double dg, int catastrophic, int Ng,
                                                           DoseCoef* __restrict__ out) {
#pragma clang fp contract(off)
    const int g = blockIdx.x * kBlock + threadIdx.x;
    if (g >= Ng) return;
    DoseCoef w;
    w.w1 = 0.5 * (S[g] + S[g + 1]) * dg;
    w.w2 = (S[g + 1] - S[g]) / 6.0 * dg;
    if (catastrophic) {
        const double sig = max_like_std(sigma[g], 0.0);
        w.w1 += 0.5 * (en[g] + en[g + 1]) * sig * dg;
        w.w2 += (en[g + 1] - en[g]) * sig * dg / 6.0;
    }
    out[g] = w;
}

// Transverse quadrature weight per node s = iz (Ny+1) + iy: 1, or the
// trapezoid's 1/2 per boundary axis.
__global__ __launch_bounds__(kBlock) void yz_weight_kernel(int Ny, int Nz, int trapezoidal,
                                                           double* __restrict__ out) {
    const int s = blockIdx.x * kBlock + threadIdx.x;
    if (s >= (Ny + 1) * (Nz + 1)) return;
    const int iz = s / (Ny + 1), iy = s - iz * (Ny + 1);
    double v = 1.0;
    if (trapezoidal) {
        if (iy == 0 || iy == Ny) v *= 0.5;
        if (iz == 0 || iz == Nz) v *= 0.5;
    }
    out[s] = v;
}

}  // namespace

void launch_copy16(const void* src, void* dst, size_t bytes, hipStream_t s) {
    const long long n = static_cast<long long>(bytes / 16);
    if (n == 0) return;
    const long long blocks = std::min<long long>((n + 4 * kBlock - 1) / (4 * kBlock), 65536);
    hipLaunchKernelGGL(copy16_kernel, dim3(static_cast<unsigned>(blocks)), dim3(kBlock), 0, s,
                       static_cast<const f32x4v_copy*>(src), static_cast<f32x4v_copy*>(dst), n);
    BFPN_HIP(hipGetLastError());
}

template <typename T>
void launch_convert(const double* src, T* dst, long long n, hipStream_t s) {
    if (n <= 0) return;
    hipLaunchKernelGGL(convert_kernel<T>, dim3(static_cast<unsigned>((n + kBlock - 1) / kBlock)),
                       dim3(kBlock), 0, s, src, dst, n);
    BFPN_HIP(hipGetLastError());
}
template void launch_convert<float>(const double*, float*, long long, hipStream_t);
template void launch_convert<double>(const double*, double*, long long, hipStream_t);

void launch_energy_coefs(const double* S, const double* sigma, const double* Tc, bool straggling,
                         double h, double dg, double rho, int Ng, EnergyCoef* out, hipStream_t s) {
    if (Ng <= 0) return;
    hipLaunchKernelGGL(energy_coef_kernel, dim3((Ng + kBlock - 1) / kBlock), dim3(kBlock), 0, s, S,
                       sigma, Tc, straggling ? 1 : 0, h, dg, rho, Ng, out);
    BFPN_HIP(hipGetLastError());
}

void launch_dose_coefs(const double* S, const double* sigma, const double* en, double dg,
                       bool catastrophic, int Ng, DoseCoef* out, hipStream_t s) {
    if (Ng <= 0) return;
    hipLaunchKernelGGL(dose_coef_kernel, dim3((Ng + kBlock - 1) / kBlock), dim3(kBlock), 0, s, S,
                       sigma, en, dg, catastrophic ? 1 : 0, Ng, out);
    BFPN_HIP(hipGetLastError());
}

void launch_yz_weights(int Ny, int Nz, bool trapezoidal, double* out, hipStream_t s) {
    const int n = (Ny + 1) * (Nz + 1);
    if (n <= 0) return;
    hipLaunchKernelGGL(yz_weight_kernel, dim3((n + kBlock - 1) / kBlock), dim3(kBlock), 0, s, Ny, Nz,
                       trapezoidal ? 1 : 0, out);
    BFPN_HIP(hipGetLastError());
}

int dose_blocks(int Ng, long long lanes) {
    (void)Ng;
    return static_cast<int>((lanes + kBlock - 1) / kBlock);
}

template <typename T>
void launch_dose(const T* F, const T* f, const T* F1, const T* f1, const DoseCoef* w,
                 const double* yz_weight, int Ng, long long lanes, int nang, double* partials,
                 hipStream_t s) {
    if (lanes <= 0) return;
    hipLaunchKernelGGL(dose_kernel<T>, dim3(dose_blocks(Ng, lanes)), dim3(kBlock), 0, s, F, f, F1,
                       f1, w, yz_weight, Ng, lanes, nang, partials);
    BFPN_HIP(hipGetLastError());
}

void launch_finish_sum(const double* a, int na, const double* b, int nb, double scale, double* out,
                       hipStream_t s) {
    hipLaunchKernelGGL(finish_sum_kernel, dim3(1), dim3(1024), 0, s, a, na, b, nb, scale, out);
    BFPN_HIP(hipGetLastError());
}

void launch_sum_in_order(const double* v, int n, double* out, hipStream_t s) {
    hipLaunchKernelGGL(sum_in_order_kernel, dim3(1), dim3(1), 0, s, v, n, out);
    BFPN_HIP(hipGetLastError());
}

template <typename T, int MODE>
static void node_copy(T* dst, const T* src, const int* nodes, int n, long long nyz, int nang, int Ng,
                      hipStream_t s) {
    typedef __attribute__((ext_vector_type(16 / sizeof(T)))) T vec_t;
    constexpr int V = 16 / sizeof(T);
    const bool vec = nang % V == 0 && reinterpret_cast<uintptr_t>(dst) % 16 == 0 &&
                     reinterpret_cast<uintptr_t>(src) % 16 == 0;  // (src is null when zeroing)
    const long long per = static_cast<long long>(n) * (vec ? nang / V : nang);
    if (per <= 0 || Ng <= 0) return;
    BFPN_REQUIRE(Ng <= 65535 && (per + kBlock - 1) / kBlock < (1LL << 31), "node copy grid too large");
    const dim3 grid(static_cast<unsigned>((per + kBlock - 1) / kBlock), static_cast<unsigned>(Ng));
    if (vec)
        hipLaunchKernelGGL((node_copy_kernel<T, vec_t, MODE>), grid, dim3(kBlock), 0, s, dst, src,
                           nodes, n, nyz, nang);
    else
        hipLaunchKernelGGL((node_copy_kernel<T, T, MODE>), grid, dim3(kBlock), 0, s, dst, src, nodes,
                           n, nyz, nang);
    BFPN_HIP(hipGetLastError());
}

template <typename T>
void launch_gather_nodes(T* dst, const T* src, const int* nodes, int n, long long nyz, int nang,
                         int Ng, hipStream_t s) {
    node_copy<T, 0>(dst, src, nodes, n, nyz, nang, Ng, s);
}

template <typename T>
void launch_scatter_nodes(T* dst, const T* src, const int* nodes, int n, long long nyz, int nang,
                          int Ng, hipStream_t s) {
    node_copy<T, 1>(dst, src, nodes, n, nyz, nang, Ng, s);
}

template <typename T>
void launch_zero_nodes(T* dst, const int* nodes, int n, long long nyz, int nang, int Ng,
                       hipStream_t s) {
    node_copy<T, 2>(dst, nullptr, nodes, n, nyz, nang, Ng, s);
}

template void launch_gather_nodes<float>(float*, const float*, const int*, int, long long, int, int,
                                         hipStream_t);
template void launch_gather_nodes<double>(double*, const double*, const int*, int, long long, int,
                                          int, hipStream_t);
template void launch_scatter_nodes<float>(float*, const float*, const int*, int, long long, int,
                                          int, hipStream_t);
template void launch_scatter_nodes<double>(double*, const double*, const int*, int, long long, int,
                                           int, hipStream_t);
template void launch_zero_nodes<float>(float*, const int*, int, long long, int, int, hipStream_t);
template void launch_zero_nodes<double>(double*, const int*, int, long long, int, int, hipStream_t);

template <typename T>
void launch_beam(T* F, const double* space_angle, const double* energy, int Ng, long long lanes,
                 hipStream_t s) {
    if (lanes <= 0 || Ng <= 0) return;
    dim3 grid(static_cast<unsigned>((lanes + kBlock - 1) / kBlock), Ng);
    hipLaunchKernelGGL(beam_kernel<T>, grid, dim3(kBlock), 0, s, F, space_angle, energy, lanes);
    BFPN_HIP(hipGetLastError());
}

template void launch_dose<float>(const float*, const float*, const float*, const float*,
                                 const DoseCoef*, const double*, int, long long, int, double*,
                                 hipStream_t);
template void launch_dose<double>(const double*, const double*, const double*, const double*,
                                  const DoseCoef*, const double*, int, long long, int, double*,
                                  hipStream_t);
template void launch_beam<float>(float*, const double*, const double*, int, long long, hipStream_t);
template void launch_beam<double>(double*, const double*, const double*, int, long long,
                                  hipStream_t);

}  // namespace bfpn
