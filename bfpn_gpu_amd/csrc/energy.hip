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
//  * loads run 16 groups (FP64 state: 8) ahead in a register ring to hide HBM
//    latency behind the serial recurrence (small grids leave only a few
//    waves per SIMD); per-group coefficients, dose weights included, come
//    through the scalar cache so they never queue behind the ring;
//  * the final E-half of a depth step also accumulates the Eq.24 dose and the
//    psi^2 proxy in FP64 (block partials, reduced deterministically later), so
//    the IDD costs no extra pass over the state.

#include "kernels.hpp"

#include <type_traits>

#include <utility>

namespace bfpn {
namespace {

constexpr int kBlock = 256;

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

__device__ inline int block_max(int v, int* lds) {
    for (int off = 32; off > 0; off >>= 1) v = max(v, __shfl_xor(v, off));
    const int wave = threadIdx.x >> 6;
    if ((threadIdx.x & 63) == 0) lds[wave] = v;
    __syncthreads();
    int t = -1;
    if (threadIdx.x == 0)
        for (int w = 0; w < kBlock / 64; ++w) t = max(t, lds[w]);
    __syncthreads();
    return t;
}

// Results below the smallest normal number of the state type are flushed to
// zero (what a flush-to-zero mode does).  Without it the upper energy tail of
// a beam never vanishes: a value at the smallest denormal, scaled by a decay
// factor above 1/2 each half step, rounds back to itself, so the top nonzero
// group could never fall as the beam slows.  The change is below 2^-126 of
// the state's scale in FP32 (2^-1022 in FP64).
template <typename T>
__device__ inline double flush_tiny(double v) {
    constexpr double tiny = sizeof(T) == 4 ? 1.1754943508222875e-38 : 2.2250738585072014e-308;
    return fabs(v) < tiny ? 0.0 : v;
}

// Top group of the state a sweep reads, and of a source made from a primary
// state with top *x (all groups when the tops are not tracked).
__device__ inline int state_top(const SweepTops& t, int Ng) { return t.in ? *t.in : Ng - 1; }
__device__ inline int source_top(const SweepTops& t, const int* x, int Ng) {
    const int top = (t.qmap && x) ? t.qmap[max(-1, min(Ng - 1, *x)) + 1] : Ng - 1;
    return max(-1, min(top, t.q_cap));
}
// Source row read at group g: rows above the source's top are masked where
// used, so their loads re-read the top row (an L2 hit) instead of HBM.
__device__ inline long long source_row(int g, int q_top) {
    return static_cast<long long>(max(min(g, q_top), 0));
}

// Presets the slots a later sweep writes (one thread of the grid).
__device__ inline void reset_tops(const SweepTops& t) {
    if (blockIdx.x == 0 && threadIdx.x == 0) {
        if (t.reset) *t.reset = -1;
        if (t.reset_x) *t.reset_x = -1;
    }
}

// Publishes a block's highest nonzero output group (-1: none).
__device__ inline void publish_top(int top, int* out, int* out2, int* lds) {
    const int b = block_max(top, lds);
    if (threadIdx.x == 0 && b >= 0) {
        if (out) atomicMax(out, b);
        if (out2) atomicMax(out2, b);
    }
}

// LPT adjacent lanes per thread (4 with FP32 state on large lane counts):
// 16-byte row loads and stores, four independent recurrences in flight per
// thread, one set of coefficient loads for four lanes.
template <typename T, int LPT>
struct Pack {
    T v[LPT];
};

// f(integral_constant<int, 0>), ..., f(integral_constant<int, N - 1>)
template <typename F, int... I>
__device__ inline void unroll_seq(F&& f, std::integer_sequence<int, I...>) {
    (f(std::integral_constant<int, I>{}), ...);
}
template <int N, typename F>
__device__ inline void unroll(F&& f) {
    unroll_seq(f, std::make_integer_sequence<int, N>{});
}

// Loads and stores are non-temporal: every element is touched once per sweep
// and nothing the sweep moves is read again before the next operator has
// streamed the whole state through the caches, so keeping it out of L2 only
// frees the cache for the loads in flight (on the paper grid the paired
// sweeps ran 12 % faster with non-temporal stores and 3 % more with
// non-temporal loads).
template <typename T, int LPT>
__device__ inline Pack<T, LPT> load_pack(const T* p) {
    Pack<T, LPT> r;
    if constexpr (LPT == 4) {
        static_assert(sizeof(T) == 4, "4-lane packs are FP32");
        typedef float f32x4v __attribute__((ext_vector_type(4)));
        const f32x4v q = __builtin_nontemporal_load(reinterpret_cast<const f32x4v*>(p));
        r.v[0] = q[0];
        r.v[1] = q[1];
        r.v[2] = q[2];
        r.v[3] = q[3];
    } else {
        static_assert(LPT == 1, "one or four lanes per thread");
        r.v[0] = __builtin_nontemporal_load(p);
    }
    return r;
}

template <typename T, int LPT>
__device__ inline void store_pack(T* p, const double (&v)[LPT]) {
    if constexpr (LPT == 4) {
        typedef float f32x4v __attribute__((ext_vector_type(4)));
        const f32x4v q = {static_cast<float>(v[0]), static_cast<float>(v[1]),
                          static_cast<float>(v[2]), static_cast<float>(v[3])};
        __builtin_nontemporal_store(q, reinterpret_cast<f32x4v*>(p));
    } else {
        __builtin_nontemporal_store(static_cast<T>(v[0]), p);
    }
}

// Prefetch depth D (groups in flight per thread).  The register ring has
// D + 1 slots and the sweep is unrolled by D + 1, so slot indices are
// compile-time constants: the step on group g (slot k) loads group g - D
// into slot k - 1, which the previous step has finished with.  No load ever
// targets a register whose old value is still in use, so the compiler never
// has to copy a loaded value -- a copy waits for the load and would drain
// the ring.
template <typename T, int LPT>
constexpr int prefetch_depth() { return LPT == 4 ? 8 : (sizeof(T) == 4 ? 16 : 8); }

// Lanes per thread for a sweep over `lanes` lanes of `elem`-byte state.
inline int lanes_per_thread(long long lanes, size_t elem) {
    return (elem == 4 && lanes % 4 == 0 && lanes >= (1LL << 18)) ? 4 : 1;
}

// The two linear forms of a group (f_new = af.x, F_new = aF.x over
// x = (P_hi, R_hi, P, R, P_lo, Q, Delta_hi)), wave-uniform: scalar loads.
struct Forms {
    double af[7], aF[7];
};

__device__ inline Forms load_forms(const EnergyCoef* coef, int g) {
    Forms r;
#pragma unroll
    for (int i = 0; i < 7; ++i) {
        r.af[i] = coef[g].af[i];
        r.aF[i] = coef[g].aF[i];
    }
    return r;
}

// One group of one lane's recurrence: input (P, R) at g, carried input
// (P_hi, R_hi) and output difference hd at g+1; returns the outputs (a: psi^2,
// b: psi^1) and moves the carries to g.  Inputs are state values, so the
// carries are held in the state type C (exactly).
template <bool SRC, bool STRAG, typename C>
__device__ inline void recur(const Forms& c, C P, C R, double P_lo, double src, C& P_hi, C& R_hi,
                             double& hd, double& a, double& b) {
    const double Ph = P_hi, Rh = R_hi, Pd = P, Rd = R;
    // explicit FMAs: every kernel that inlines this evaluates it identically
    // (left to the compiler, contraction could differ between kernels and
    // break the paired sweep's bit-identity with two single sweeps)
    a = fma(c.af[3], Rd, fma(c.af[2], Pd, fma(c.af[1], Rh, c.af[0] * Ph)));
    b = fma(c.aF[3], Rd, fma(c.aF[2], Pd, fma(c.aF[1], Rh, c.aF[0] * Ph)));
    if (STRAG) {
        a = fma(c.af[4], P_lo, a);
        b = fma(c.aF[4], P_lo, b);
    }
    if (SRC) {
        a = fma(c.af[5], src, a);
        b = fma(c.aF[5], src, b);
    }
    a = flush_tiny<C>(fma(c.af[6], hd, a));
    b = flush_tiny<C>(fma(c.aF[6], hd, b));
    P_hi = P;
    R_hi = R;
    hd = b - a;
}

// Eq.24 dose of one group's result, accumulated with explicit FMAs (the same
// operations in every kernel, see recur).
__device__ inline double add_dose(double acc, double F, double f, double w1, double w2) {
    return fma(f, w2, fma(F, w1, acc));
}

// Everything a group needs besides the state (recurrence coefficients and the
// dose weights) sits in one EnergyCoef read through the scalar cache, so the
// vector loads in flight are the state ring alone: the in-order load counter
// then never waits on the ring to deliver a coefficient.
template <typename T, int LPT, bool SRC, bool STRAG, bool DOSE>
__global__ __launch_bounds__(kBlock) void energy_half_kernel(
    T* __restrict__ F, T* __restrict__ f, const EnergyCoef* __restrict__ coef,
    const T* __restrict__ Q1, int Ng, int lanes, const double* __restrict__ yzw, int nang,
    double* __restrict__ partials, int proxy, SweepTops tops) {
    constexpr int D = prefetch_depth<T, LPT>(), RING = D + 1;
    __shared__ double lds[kBlock / 64];
    __shared__ int ldi[kBlock / 64];
    const int lane = (blockIdx.x * kBlock + threadIdx.x) * LPT;  // first of LPT lanes
    reset_tops(tops);
    // Groups above G0 are zero before and after the sweep: the state above
    // its top, the source above its own, and straggling lifts the result at
    // most one group.  Source values above the source's top are not read.
    const int q_top = SRC ? source_top(tops, tops.xa, Ng) : -1;
    const int G0 = min(Ng - 1, max(state_top(tops, Ng), q_top) + (STRAG ? 1 : 0));
    int top = -1;  // highest group with a nonzero result in this thread's lanes
    double acc_dose = 0.0, acc_proxy = 0.0;
    if (lane < lanes) {
        Pack<T, LPT> rF[RING], rf[RING], rQ[RING];
        // Loads are unconditional (groups below 0 re-read group 0 and are
        // never consumed) and full chunks run without branches, so the
        // compiler can count the loads in flight instead of draining them.
        // Row bases are wave-uniform (scalar), the lane a 32-bit offset.
        auto fetch = [&](int k, int g) {
            const long long base = static_cast<long long>(max(g, 0)) * lanes;
            rF[k] = load_pack<T, LPT>(F + base + lane);
            rf[k] = load_pack<T, LPT>(f + base + lane);
            if (SRC) rQ[k] = load_pack<T, LPT>(Q1 + source_row(g, q_top) * lanes + lane);
        };
#pragma unroll
        for (int k = 0; k < D; ++k) fetch(k, G0 - k);
        T P_hi[LPT], R_hi[LPT];
        double hi_delta[LPT], dose[LPT];
#pragma unroll
        for (int j = 0; j < LPT; ++j) {
            P_hi[j] = R_hi[j] = T(0);
            hi_delta[j] = dose[j] = 0.0;
        }
        // The linear forms run one group ahead through the scalar cache:
        // group g-1's are requested before group g's arithmetic (the
        // scheduling barrier keeps the request there), so their latency is
        // covered.  (Dose weights are read at use: two sets of forms plus
        // them would not fit the scalar registers.)
        Forms cn = load_forms(coef, max(G0, 0));
        auto step = [&](auto slot, int g) {  // group g sits in ring slot `slot`
            constexpr int k = decltype(slot)::value;
            const Forms c = cn;
            cn = load_forms(coef, max(g - 1, 0));
            __builtin_amdgcn_sched_barrier(0);
            fetch((k + D) % RING, g - D);
            const Pack<T, LPT> cF = rF[k], cf = rf[k], cQ = rQ[k];
            const Pack<T, LPT> lo = rF[(k + 1) % RING];  // old F at g-1
            const bool q_live = g <= q_top;
            double Fv[LPT], fv[LPT];
#pragma unroll
            for (int j = 0; j < LPT; ++j)
                recur<SRC, STRAG>(c, cF.v[j], cf.v[j], g > 0 ? static_cast<double>(lo.v[j]) : 0.0,
                                  SRC && q_live ? static_cast<double>(cQ.v[j]) : 0.0, P_hi[j],
                                  R_hi[j], hi_delta[j], fv[j], Fv[j]);
            const long long base = static_cast<long long>(g) * lanes;
            store_pack<T, LPT>(F + base + lane, Fv);
            store_pack<T, LPT>(f + base + lane, fv);
            bool nz = false;
#pragma unroll
            for (int j = 0; j < LPT; ++j)
                nz |= (static_cast<T>(Fv[j]) != T(0)) | (static_cast<T>(fv[j]) != T(0));
            top = max(top, nz ? g : -1);
            if (DOSE) {  // the lanes' transverse weights are applied after the sweep
                const double w1 = coef[g].w1, w2 = coef[g].w2;
#pragma unroll
                for (int j = 0; j < LPT; ++j) {
                    dose[j] = add_dose(dose[j], Fv[j], fv[j], w1, w2);
                    acc_proxy += fv[j];
                }
            }
        };
        int g0 = G0;
        for (; g0 >= RING - 1; g0 -= RING)
            unroll<RING>([&](auto k) { step(k, g0 - decltype(k)::value); });
        unroll<RING>([&](auto k) {  // tail: groups g0 .. 0 (fewer than RING)
            if (decltype(k)::value <= g0) step(k, g0 - decltype(k)::value);
        });
        // (loaded only now: a load issued before the ring would hold up its
        // first consumer behind every state load in flight)
        if (DOSE)
#pragma unroll
            for (int j = 0; j < LPT; ++j) acc_dose = fma(dose[j], yzw[(lane + j) / nang], acc_dose);
    }
    if (DOSE) {
        const double d = block_sum(acc_dose, lds);
        if (threadIdx.x == 0) partials[blockIdx.x] = d;
        if (proxy) {
            const double p = block_sum(acc_proxy, lds);
            if (threadIdx.x == 0) partials[gridDim.x + blockIdx.x] = p;
        }
    }
    if (tops.out) publish_top(top, tops.out, tops.out_x, ldi);
}

// Two consecutive half steps in one sweep: the final E-half of depth step n
// (source Qa, with the dose and proxy of step n) and the first E-half of step
// n+1 (source Qb), i.e. E(h) o E(h) over the same coefficients, reading and
// writing the state once instead of twice.  The second recurrence runs one
// group behind the first: at group g it needs the first's outputs at g+1 and
// g+2 and, for straggling, at g, which the first has just produced.  The
// first's outputs are rounded to the state type before the second uses them,
// exactly as if stored and reloaded, so the result is bit-identical to two
// launches.  XOUT also stores them (the state at the end of step n, which the
// secondary source of step n reads).
template <typename T, int LPT, bool SRC>
constexpr int pair_depth() {
    return LPT == 4 ? (SRC ? 4 : 6) : (sizeof(T) == 4 ? (SRC ? 10 : 14) : (SRC ? 4 : 6));
}

template <typename T, int LPT, bool SRC, bool STRAG, bool XOUT>
__global__ __launch_bounds__(kBlock) void energy_pair_kernel(
    T* __restrict__ F, T* __restrict__ f, const EnergyCoef* __restrict__ coef,
    const T* __restrict__ Qa, const T* __restrict__ Qb, T* __restrict__ XF, T* __restrict__ Xf,
    const unsigned* __restrict__ xmask, int Ng, int lanes, const double* __restrict__ yzw,
    int nang, double* __restrict__ partials, int proxy, SweepTops tops) {
    constexpr int D = pair_depth<T, LPT, SRC>(), RING = D + 1;
    __shared__ double lds[kBlock / 64];
    __shared__ int ldi[kBlock / 64];
    const int lane = (blockIdx.x * kBlock + threadIdx.x) * LPT;
    reset_tops(tops);
    // The first half's result is zero above t1, the second's above G0 (see
    // energy_half_kernel); both recurrences start at G0.
    const int qa_top = SRC ? source_top(tops, tops.xa, Ng) : -1;
    const int qb_top = SRC ? source_top(tops, tops.xb, Ng) : -1;
    const int t1 = max(state_top(tops, Ng), qa_top) + (STRAG ? 1 : 0);
    const int G0 = min(Ng - 1, max(t1, qb_top) + (STRAG ? 1 : 0));
    int top_a = -1, top_b = -1;  // highest nonzero group of each half's result
    double acc_dose = 0.0, acc_proxy = 0.0;
    if (lane < lanes && G0 >= 0) {
        Pack<T, LPT> rF[RING], rf[RING], rQa[RING], rQb[RING];
        // Qb (the second half's source) is refilled separately, at the end of
        // a step: the second half reads slot k - 1's Qb in the step that
        // refills that slot.
        auto fetch = [&](int k, int g) {
            const long long base = static_cast<long long>(max(g, 0)) * lanes;
            rF[k] = load_pack<T, LPT>(F + base + lane);
            rf[k] = load_pack<T, LPT>(f + base + lane);
            if (SRC) rQa[k] = load_pack<T, LPT>(Qa + source_row(g, qa_top) * lanes + lane);
        };
        auto fetch_qb = [&](int k, int g) {
            if (SRC) rQb[k] = load_pack<T, LPT>(Qb + source_row(g, qb_top) * lanes + lane);
        };
#pragma unroll
        for (int k = 0; k < D; ++k) {
            fetch(k, G0 - k);
            fetch_qb(k, G0 - k);
        }
        // first half: carried input at g+1, output difference; dose sums
        T P_hi[LPT], R_hi[LPT];
        double hd[LPT], dose[LPT];
        // second half, one group behind: its input (the first's output,
        // rounded to the state type) at g+1 and g+2, its output difference
        T P2[LPT], R2[LPT], P2_hi[LPT], R2_hi[LPT];
        double hd2[LPT];
#pragma unroll
        for (int j = 0; j < LPT; ++j) {
            P_hi[j] = R_hi[j] = P2[j] = R2[j] = P2_hi[j] = R2_hi[j] = T(0);
            hd[j] = dose[j] = hd2[j] = 0.0;
        }
        // second half at group g2, whose lower neighbour's input is P_lo2 and
        // whose source is q
        auto second = [&](int g2, const T (&P_lo2)[LPT], const Pack<T, LPT>& q) {
            // (requested here, not hoisted: three sets of forms live at once
            // would not fit the scalar registers)
            __builtin_amdgcn_sched_barrier(0);
            const Forms c2 = load_forms(coef, g2);
            const bool q_live = g2 <= qb_top;
            double Fv[LPT], fv[LPT];
#pragma unroll
            for (int j = 0; j < LPT; ++j)
                recur<SRC, STRAG>(c2, P2[j], R2[j], static_cast<double>(P_lo2[j]),
                                  SRC && q_live ? static_cast<double>(q.v[j]) : 0.0, P2_hi[j],
                                  R2_hi[j], hd2[j], fv[j], Fv[j]);
            const long long base = static_cast<long long>(g2) * lanes;
            store_pack<T, LPT>(F + base + lane, Fv);
            store_pack<T, LPT>(f + base + lane, fv);
            bool nz = false;
#pragma unroll
            for (int j = 0; j < LPT; ++j)
                nz |= (static_cast<T>(Fv[j]) != T(0)) | (static_cast<T>(fv[j]) != T(0));
            top_b = max(top_b, nz ? g2 : -1);
        };
        Forms cn = load_forms(coef, G0);
        // group g in ring slot `slot`; FIRST: g = Ng - 1, no second half behind it
        auto step = [&](auto slot, int g, auto first) {
            constexpr int k = decltype(slot)::value;
            const Forms c = cn;
            cn = load_forms(coef, max(g - 1, 0));
            __builtin_amdgcn_sched_barrier(0);
            fetch((k + D) % RING, g - D);
            const Pack<T, LPT> cF = rF[k], cf = rf[k], cQa = rQa[k];
            const Pack<T, LPT> lo = rF[(k + 1) % RING];
            const bool q_live = g <= qa_top;
            double Fv[LPT], fv[LPT];
            T tF[LPT];
#pragma unroll
            for (int j = 0; j < LPT; ++j)
                recur<SRC, STRAG>(c, cF.v[j], cf.v[j], g > 0 ? static_cast<double>(lo.v[j]) : 0.0,
                                  SRC && q_live ? static_cast<double>(cQa.v[j]) : 0.0, P_hi[j],
                                  R_hi[j], hd[j], fv[j], Fv[j]);
            {  // dose and proxy of step n from the first half's outputs
                const double w1 = coef[g].w1, w2 = coef[g].w2;
#pragma unroll
                for (int j = 0; j < LPT; ++j) {
                    dose[j] = add_dose(dose[j], Fv[j], fv[j], w1, w2);
                    acc_proxy += fv[j];
                }
            }
            // (only the groups the source reads, when xmask is given)
            if (XOUT && (!xmask || ((xmask[g >> 5] >> (g & 31)) & 1u))) {
                const long long base = static_cast<long long>(g) * lanes;
                store_pack<T, LPT>(XF + base + lane, Fv);
                store_pack<T, LPT>(Xf + base + lane, fv);
            }
            bool nz = false;
#pragma unroll
            for (int j = 0; j < LPT; ++j) {
                tF[j] = static_cast<T>(Fv[j]);
                nz |= (tF[j] != T(0)) | (static_cast<T>(fv[j]) != T(0));
            }
            top_a = max(top_a, nz ? g : -1);
            if (!decltype(first)::value) second(g + 1, tF, rQb[(k + D) % RING]);  // Qb at g+1
            fetch_qb((k + D) % RING, g - D);
#pragma unroll
            for (int j = 0; j < LPT; ++j) {  // the first's output at g: the second's next input
                P2[j] = tF[j];
                R2[j] = static_cast<T>(fv[j]);
            }
        };
        // group G0 (slot 0) on its own, so the loop body has no branch
        step(std::integral_constant<int, 0>{}, G0, std::true_type{});
        int g0 = G0 - 1;  // in slot 1
        for (; g0 >= RING - 1; g0 -= RING)
            unroll<RING>([&](auto i) {
                step(std::integral_constant<int, (decltype(i)::value + 1) % RING>{},
                     g0 - decltype(i)::value, std::false_type{});
            });
        unroll<RING>([&](auto i) {
            if (decltype(i)::value <= g0)
                step(std::integral_constant<int, (decltype(i)::value + 1) % RING>{},
                     g0 - decltype(i)::value, std::false_type{});
        });
        const T none[LPT] = {};
        Pack<T, LPT> q0{};
        if (SRC) q0 = load_pack<T, LPT>(Qb + lane);
        second(0, none, q0);  // group 0 has no lower neighbour
#pragma unroll
        for (int j = 0; j < LPT; ++j) acc_dose = fma(dose[j], yzw[(lane + j) / nang], acc_dose);
    }
    const double d = block_sum(acc_dose, lds);
    if (threadIdx.x == 0) partials[blockIdx.x] = d;
    if (proxy) {
        const double p = block_sum(acc_proxy, lds);
        if (threadIdx.x == 0) partials[gridDim.x + blockIdx.x] = p;
    }
    if (tops.out) publish_top(top_b, tops.out, nullptr, ldi);
    if (tops.out_x) publish_top(top_a, tops.out_x, nullptr, ldi);
}

// First-order implicit upwind energy step: the reference's legacy energy
// model (energyDgCnKernel, use_legacy branch, src/energy_solver.cu:88-106)
// and its lite-memory primary step (litePrimaryEnergyStepKernel,
// src/bfpm_solver.cu:600-622, psi^1 only: f == null).  One thread per lane,
// high to low energy, in place: the old psi^1 of g+1 (straggling) is carried
// in a register and the old psi^1 of g-1 is read before it is overwritten.
// Divisions as in the reference; results flushed like the eq15 sweeps.
template <typename T, bool SRC, bool DOSE>
__global__ __launch_bounds__(kBlock) void energy_upwind_kernel(
    T* __restrict__ F, T* __restrict__ f, const UpwindCoef* __restrict__ uc,
    const EnergyCoef* __restrict__ coef, const T* __restrict__ Q1, int Ng, int lanes,
    const double* __restrict__ yzw, int nang, double* __restrict__ partials, int proxy,
    int straggling, SweepTops tops) {
    __shared__ double lds[kBlock / 64];
    __shared__ int ldi[kBlock / 64];
    const int lane = blockIdx.x * kBlock + threadIdx.x;
    reset_tops(tops);
    const int q_top = SRC ? source_top(tops, tops.xa, Ng) : -1;
    const int G0 = min(Ng - 1, max(state_top(tops, Ng), q_top) + (straggling ? 1 : 0));
    int top = -1;
    double acc_dose = 0.0, acc_proxy = 0.0;
    if (lane < lanes && G0 >= 0) {
        double next_F = 0.0, next_f = 0.0, dose = 0.0;
        double F_hi = 0.0;  // old psi^1 at g+1 (zero above G0)
        double Fg = static_cast<double>(F[static_cast<long long>(G0) * lanes + lane]);
        for (int g = G0; g >= 0; --g) {
            const long long idx = static_cast<long long>(g) * lanes + lane;
            const double F_lo = g > 0 ? static_cast<double>(F[idx - lanes]) : 0.0;
            const double fg = f ? static_cast<double>(f[idx]) : 0.0;
            const UpwindCoef c = uc[g];
            const double src = (SRC && g <= q_top) ? static_cast<double>(Q1[idx]) : 0.0;
            double Fv = (fmax(Fg, 0.0) / c.h + c.c_hi * next_F + src) / c.denom;
            double fv = (fg / c.h + c.c_hi * next_f) / c.denom;
            if (c.strag > 0.0) {
                const double lap = fmax(F_hi, 0.0) - 2.0 * fmax(Fg, 0.0) + fmax(F_lo, 0.0);
                Fv += c.strag * lap;
            }
            const T sF = static_cast<T>(flush_tiny<T>(fmax(Fv, 0.0)));
            const T sf = static_cast<T>(flush_tiny<T>(fmax(fv, 0.0)));
            F[idx] = sF;
            if (f) f[idx] = sf;
            next_F = static_cast<double>(sF);
            next_f = static_cast<double>(sf);
            top = max(top, (sF != T(0) || sf != T(0)) ? g : -1);
            if (DOSE) {
                dose = add_dose(dose, next_F, next_f, coef[g].w1, coef[g].w2);
                acc_proxy += next_f;
            }
            F_hi = Fg;
            Fg = F_lo;
        }
        if (DOSE) acc_dose = dose * yzw[lane / nang];
    }
    if (DOSE) {
        const double d = block_sum(acc_dose, lds);
        if (threadIdx.x == 0) partials[blockIdx.x] = d;
        if (proxy) {
            const double p = block_sum(acc_proxy, lds);
            if (threadIdx.x == 0) partials[gridDim.x + blockIdx.x] = p;
        }
    }
    if (tops.out) publish_top(top, tops.out, tops.out_x, ldi);
}

template <typename T, int LPT, bool SRC, bool STRAG>
void dispatch_pair(T* F, T* f, const EnergyCoef* coef, const T* Qa, const T* Qb, T* XF, T* Xf,
                   const unsigned* xmask, int Ng, long long lanes, const double* yzw, int nang,
                   double* partials, bool proxy, hipStream_t s, const SweepTops& tops) {
    const int blocks = energy_blocks(lanes, sizeof(T));
    BFPN_REQUIRE(lanes < (1LL << 31), "energy sweep indexes lanes with 32-bit offsets");
    if (XF)
        hipLaunchKernelGGL((energy_pair_kernel<T, LPT, SRC, STRAG, true>), dim3(blocks),
                           dim3(kBlock), 0, s, F, f, coef, Qa, Qb, XF, Xf, xmask, Ng,
                           static_cast<int>(lanes), yzw, nang, partials, proxy ? 1 : 0, tops);
    else
        hipLaunchKernelGGL((energy_pair_kernel<T, LPT, SRC, STRAG, false>), dim3(blocks),
                           dim3(kBlock), 0, s, F, f, coef, Qa, Qb, XF, Xf, xmask, Ng,
                           static_cast<int>(lanes), yzw, nang, partials, proxy ? 1 : 0, tops);
}

template <typename T, bool SRC, bool STRAG>
void dispatch_pair_lpt(T* F, T* f, const EnergyCoef* coef, const T* Qa, const T* Qb, T* XF,
                       T* Xf, const unsigned* xmask, int Ng, long long lanes, const double* yzw,
                       int nang, double* partials, bool proxy, hipStream_t s,
                       const SweepTops& tops) {
    if constexpr (sizeof(T) == 4) {
        const int lpt = lanes_per_thread(lanes, sizeof(T));
        if (lpt == 4)
            return dispatch_pair<T, 4, SRC, STRAG>(F, f, coef, Qa, Qb, XF, Xf, xmask, Ng, lanes,
                                                   yzw, nang, partials, proxy, s, tops);
    }
    dispatch_pair<T, 1, SRC, STRAG>(F, f, coef, Qa, Qb, XF, Xf, xmask, Ng, lanes, yzw, nang,
                                    partials, proxy, s, tops);
}

template <typename T, int LPT, bool SRC, bool STRAG>
void dispatch_dose(T* F, T* f, const EnergyCoef* coef, const T* Q1, int Ng, long long lanes,
                   const double* yzw, int nang, double* partials, bool proxy, hipStream_t s,
                   const SweepTops& tops) {
    const int blocks = energy_blocks(lanes, sizeof(T));
    BFPN_REQUIRE(lanes < (1LL << 31), "energy sweep indexes lanes with 32-bit offsets");
    if (partials)
        hipLaunchKernelGGL((energy_half_kernel<T, LPT, SRC, STRAG, true>), dim3(blocks),
                           dim3(kBlock), 0, s, F, f, coef, Q1, Ng, static_cast<int>(lanes), yzw,
                           nang, partials, proxy ? 1 : 0, tops);
    else
        hipLaunchKernelGGL((energy_half_kernel<T, LPT, SRC, STRAG, false>), dim3(blocks),
                           dim3(kBlock), 0, s, F, f, coef, Q1, Ng, static_cast<int>(lanes), yzw,
                           nang, partials, 0, tops);
}

template <typename T, bool SRC, bool STRAG>
void dispatch_lpt(T* F, T* f, const EnergyCoef* coef, const T* Q1, int Ng, long long lanes,
                  const double* yzw, int nang, double* partials, bool proxy, hipStream_t s,
                  const SweepTops& tops) {
    if constexpr (sizeof(T) == 4) {
        const int lpt = lanes_per_thread(lanes, sizeof(T));
        if (lpt == 4)
            return dispatch_dose<T, 4, SRC, STRAG>(F, f, coef, Q1, Ng, lanes, yzw, nang, partials,
                                                   proxy, s, tops);
    }
    dispatch_dose<T, 1, SRC, STRAG>(F, f, coef, Q1, Ng, lanes, yzw, nang, partials, proxy, s, tops);
}

}  // namespace

int energy_blocks(long long lanes, size_t elem) {
    const long long per_block = static_cast<long long>(kBlock) * lanes_per_thread(lanes, elem);
    return static_cast<int>((lanes + per_block - 1) / per_block);
}

template <typename T>
void launch_energy_half(T* F, T* f, const EnergyCoef* coef, const T* Q1, int Ng, long long lanes,
                        bool straggling, const double* yz_weight, int nang, double* dose_partials,
                        bool proxy, hipStream_t s, const SweepTops& tops) {
    if (lanes <= 0 || Ng <= 0) return;
    if (Q1) {
        if (straggling)
            dispatch_lpt<T, true, true>(F, f, coef, Q1, Ng, lanes, yz_weight, nang, dose_partials,
                                         proxy, s, tops);
        else
            dispatch_lpt<T, true, false>(F, f, coef, Q1, Ng, lanes, yz_weight, nang, dose_partials,
                                          proxy, s, tops);
    } else {
        if (straggling)
            dispatch_lpt<T, false, true>(F, f, coef, Q1, Ng, lanes, yz_weight, nang, dose_partials,
                                          proxy, s, tops);
        else
            dispatch_lpt<T, false, false>(F, f, coef, Q1, Ng, lanes, yz_weight, nang,
                                           dose_partials, proxy, s, tops);
    }
    BFPN_HIP(hipGetLastError());
}

template <typename T>
void launch_energy_pair(T* F, T* f, const EnergyCoef* coef, const T* Qa, const T* Qb, int Ng,
                        long long lanes, bool straggling, const double* yz_weight, int nang,
                        double* dose_partials, bool proxy, T* XF, T* Xf, hipStream_t s,
                        const SweepTops& tops, const unsigned* xmask) {
    if (lanes <= 0 || Ng <= 0) return;
    BFPN_REQUIRE(dose_partials, "the paired energy sweep needs dose partials");
    BFPN_REQUIRE((Qa != nullptr) == (Qb != nullptr), "both half steps need a source, or neither");
    BFPN_REQUIRE((XF != nullptr) == (Xf != nullptr), "the intermediate state needs both arrays");
    const auto run = [&](auto src, auto strag) {
        dispatch_pair_lpt<T, decltype(src)::value, decltype(strag)::value>(
            F, f, coef, Qa, Qb, XF, Xf, xmask, Ng, lanes, yz_weight, nang, dose_partials, proxy,
            s, tops);
    };
    using Y = std::true_type;
    using N = std::false_type;
    if (Qa) {
        if (straggling) run(Y{}, Y{});
        else run(Y{}, N{});
    } else {
        if (straggling) run(N{}, Y{});
        else run(N{}, N{});
    }
    BFPN_HIP(hipGetLastError());
}

template <typename T>
void launch_energy_upwind(T* F, T* f, const UpwindCoef* uc, const EnergyCoef* coef, const T* Q1,
                          int Ng, long long lanes, bool straggling, const double* yz_weight,
                          int nang, double* dose_partials, bool proxy, hipStream_t s,
                          const SweepTops& tops) {
    if (lanes <= 0 || Ng <= 0) return;
    BFPN_REQUIRE(lanes < (1LL << 31), "energy sweep indexes lanes with 32-bit offsets");
    const int blocks = static_cast<int>((lanes + kBlock - 1) / kBlock);
    const int st = straggling ? 1 : 0, px = proxy ? 1 : 0;
    const int n = static_cast<int>(lanes);
    if (Q1 && dose_partials)
        hipLaunchKernelGGL((energy_upwind_kernel<T, true, true>), dim3(blocks), dim3(kBlock), 0, s,
                           F, f, uc, coef, Q1, Ng, n, yz_weight, nang, dose_partials, px, st, tops);
    else if (Q1)
        hipLaunchKernelGGL((energy_upwind_kernel<T, true, false>), dim3(blocks), dim3(kBlock), 0, s,
                           F, f, uc, coef, Q1, Ng, n, yz_weight, nang, dose_partials, 0, st, tops);
    else if (dose_partials)
        hipLaunchKernelGGL((energy_upwind_kernel<T, false, true>), dim3(blocks), dim3(kBlock), 0, s,
                           F, f, uc, coef, Q1, Ng, n, yz_weight, nang, dose_partials, px, st, tops);
    else
        hipLaunchKernelGGL((energy_upwind_kernel<T, false, false>), dim3(blocks), dim3(kBlock), 0,
                           s, F, f, uc, coef, Q1, Ng, n, yz_weight, nang, dose_partials, 0, st,
                           tops);
    BFPN_HIP(hipGetLastError());
}

template void launch_energy_upwind<float>(float*, float*, const UpwindCoef*, const EnergyCoef*,
                                          const float*, int, long long, bool, const double*, int,
                                          double*, bool, hipStream_t, const SweepTops&);
template void launch_energy_upwind<double>(double*, double*, const UpwindCoef*, const EnergyCoef*,
                                           const double*, int, long long, bool, const double*, int,
                                           double*, bool, hipStream_t, const SweepTops&);

template void launch_energy_pair<float>(float*, float*, const EnergyCoef*, const float*,
                                        const float*, int, long long, bool, const double*, int,
                                        double*, bool, float*, float*, hipStream_t,
                                        const SweepTops&, const unsigned*);
template void launch_energy_pair<double>(double*, double*, const EnergyCoef*, const double*,
                                         const double*, int, long long, bool, const double*, int,
                                         double*, bool, double*, double*, hipStream_t,
                                         const SweepTops&, const unsigned*);

template void launch_energy_half<float>(float*, float*, const EnergyCoef*, const float*, int,
                                        long long, bool, const double*, int, double*, bool,
                                        hipStream_t, const SweepTops&);
template void launch_energy_half<double>(double*, double*, const EnergyCoef*, const double*, int,
                                         long long, bool, const double*, int, double*, bool,
                                         hipStream_t, const SweepTops&);

}  // namespace bfpn
