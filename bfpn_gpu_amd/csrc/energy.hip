// energy.hip — Eq.15 DG/Crank-Nicolson energy half step (psi^1, psi^2).
//
// Reference: energyDgCnKernel, eq15 branch (src/energy_solver.cu:41-161) and
// the copy-back in EnergyDepositionSolver::solve (src/energy_solver.cu:497-505).
//
// MI355X design:
//  * one thread per phase-space lane (s, m); lanes are contiguous for a fixed
//    group g, so every load/store of the sweep is a fully coalesced row;
//  * the sweep runs in place: the old values of g+1 (needed by the
//    straggling and slope couplings) and the new Delta_{g+1} are carried in
//    registers, so the reference's scratch arrays and D2D copy-back vanish;
//  * group coefficients come from a host-built FP64 table (wave-uniform, so
//    they are scalar loads) and the recurrence runs in FP64 for any state
//    type;
//  * loads are software-prefetched PF groups ahead to hide HBM latency
//    behind the serial recurrence;
//  * the final E-half of a depth step also accumulates the Eq.24 dose and the
//    psi^2 proxy in FP64 (block partials, reduced deterministically later), so
//    the IDD costs no extra pass over the state.

#include "kernels.hpp"

namespace bfpn {
namespace {

constexpr int kBlock = 256;
constexpr int kPrefetch = 4;

__device__ inline double block_sum(double v, double* lds) {
    for (int off = 32; off > 0; off >>= 1) v += __shfl_xor(v, off);
    const int wave = threadIdx.x >> 6;
    if ((threadIdx.x & 63) == 0) lds[wave] = v;
    __syncthreads();
    double t = 0.0;
    if (threadIdx.x == 0)
        for (int w = 0; w < kBlock / 64; ++w) t += lds[w];
    __syncthreads();
    return t;
}

template <typename T, bool SRC, bool STRAG, bool DOSE>
__global__ __launch_bounds__(kBlock) void energy_half_kernel(
    T* __restrict__ F, T* __restrict__ f, const EnergyCoef* __restrict__ coef,
    const T* __restrict__ Q1, int Ng, long long lanes, const DoseCoef* __restrict__ dw,
    const double* __restrict__ yzw, int nang, double* __restrict__ partials, int proxy) {
    __shared__ double lds[kBlock / 64];
    const long long lane = static_cast<long long>(blockIdx.x) * kBlock + threadIdx.x;
    double acc_dose = 0.0, acc_proxy = 0.0;
    if (lane < lanes) {
        T* Fl = F + lane;
        T* fl = f + lane;
        const T* Ql = SRC ? Q1 + lane : nullptr;
        const double wyz = DOSE ? yzw[lane / nang] : 0.0;

        T cF[kPrefetch], cf[kPrefetch], cQ[kPrefetch];
#pragma unroll
        for (int k = 0; k < kPrefetch; ++k) {
            const int g = Ng - 1 - k;
            cF[k] = g >= 0 ? Fl[g * lanes] : T(0);
            cf[k] = g >= 0 ? fl[g * lanes] : T(0);
            if (SRC) cQ[k] = g >= 0 ? Ql[g * lanes] : T(0);
        }
        double P_hi = 0.0, R_hi = 0.0, hi_delta = 0.0;
        for (int g0 = Ng - 1; g0 >= 0; g0 -= kPrefetch) {
            T nF[kPrefetch], nf[kPrefetch], nQ[kPrefetch];
#pragma unroll
            for (int k = 0; k < kPrefetch; ++k) {
                const int g = g0 - kPrefetch - k;
                nF[k] = g >= 0 ? Fl[g * lanes] : T(0);
                nf[k] = g >= 0 ? fl[g * lanes] : T(0);
                if (SRC) nQ[k] = g >= 0 ? Ql[g * lanes] : T(0);
            }
#pragma unroll
            for (int k = 0; k < kPrefetch; ++k) {
                const int g = g0 - k;
                if (g >= 0) {
                    const EnergyCoef& c = coef[g];
                    const double P = static_cast<double>(cF[k]);
                    const double R = static_cast<double>(cf[k]);
                    const double up = P_hi - R_hi;
                    const double cur = P - R;
                    double strag = 0.0;
                    if (STRAG) {
                        const double P_lo =
                            g > 0 ? static_cast<double>(k + 1 < kPrefetch ? cF[k + 1] : nF[0]) : 0.0;
                        strag = c.a_hi * P_hi + c.a_lo * P_lo - c.a_g * P;
                    }
                    const double src = SRC ? static_cast<double>(cQ[k]) : 0.0;
                    const double common =
                        c.c_up * up - c.c_cur * cur - c.half_sig * P + P * c.inv_h + strag + src;
                    const double rhs4 = c.inv_F * common;
                    const double rhs1e = R * c.inv_h - c.half_sig * R;
                    const double rhs2e = c.k_up * up + c.k_cur * cur - c.k_F * P - c.k_f * R;
                    const double rhs = rhs1e + rhs2e - c.A * common + c.H * hi_delta;
                    const double fv = rhs * c.inv_coe;
                    const double Fv = c.B * fv + rhs4 + c.Cc * hi_delta;
                    Fl[g * lanes] = static_cast<T>(Fv);
                    fl[g * lanes] = static_cast<T>(fv);
                    if (DOSE) {
                        acc_dose += wyz * (Fv * dw[g].w1 + fv * dw[g].w2);
                        acc_proxy += fv;
                    }
                    P_hi = P;
                    R_hi = R;
                    hi_delta = Fv - fv;
                }
            }
#pragma unroll
            for (int k = 0; k < kPrefetch; ++k) {
                cF[k] = nF[k];
                cf[k] = nf[k];
                if (SRC) cQ[k] = nQ[k];
            }
        }
    }
    if (DOSE) {
        const double d = block_sum(acc_dose, lds);
        if (threadIdx.x == 0) partials[blockIdx.x] = d;
        if (proxy) {
            const double p = block_sum(acc_proxy, lds);
            if (threadIdx.x == 0) partials[gridDim.x + blockIdx.x] = p;
        }
    }
}

template <typename T, bool SRC, bool STRAG>
void dispatch_dose(T* F, T* f, const EnergyCoef* coef, const T* Q1, int Ng, long long lanes,
                   const DoseCoef* dw, const double* yzw, int nang, double* partials, bool proxy,
                   hipStream_t s) {
    const int blocks = energy_blocks(lanes);
    if (partials)
        hipLaunchKernelGGL((energy_half_kernel<T, SRC, STRAG, true>), dim3(blocks), dim3(kBlock), 0,
                           s, F, f, coef, Q1, Ng, lanes, dw, yzw, nang, partials, proxy ? 1 : 0);
    else
        hipLaunchKernelGGL((energy_half_kernel<T, SRC, STRAG, false>), dim3(blocks), dim3(kBlock), 0,
                           s, F, f, coef, Q1, Ng, lanes, dw, yzw, nang, partials, 0);
}

}  // namespace

int energy_blocks(long long lanes) { return static_cast<int>((lanes + kBlock - 1) / kBlock); }

template <typename T>
void launch_energy_half(T* F, T* f, const EnergyCoef* coef, const T* Q1, int Ng, long long lanes,
                        bool straggling, const DoseCoef* dose_coef, const double* yz_weight,
                        int nang, double* dose_partials, bool proxy, hipStream_t s) {
    if (lanes <= 0 || Ng <= 0) return;
    if (Q1) {
        if (straggling)
            dispatch_dose<T, true, true>(F, f, coef, Q1, Ng, lanes, dose_coef, yz_weight, nang,
                                         dose_partials, proxy, s);
        else
            dispatch_dose<T, true, false>(F, f, coef, Q1, Ng, lanes, dose_coef, yz_weight, nang,
                                          dose_partials, proxy, s);
    } else {
        if (straggling)
            dispatch_dose<T, false, true>(F, f, coef, Q1, Ng, lanes, dose_coef, yz_weight, nang,
                                          dose_partials, proxy, s);
        else
            dispatch_dose<T, false, false>(F, f, coef, Q1, Ng, lanes, dose_coef, yz_weight, nang,
                                           dose_partials, proxy, s);
    }
    BFPN_HIP(hipGetLastError());
}

template void launch_energy_half<float>(float*, float*, const EnergyCoef*, const float*, int,
                                        long long, bool, const DoseCoef*, const double*, int,
                                        double*, bool, hipStream_t);
template void launch_energy_half<double>(double*, double*, const EnergyCoef*, const double*, int,
                                         long long, bool, const DoseCoef*, const double*, int,
                                         double*, bool, hipStream_t);

}  // namespace bfpn
