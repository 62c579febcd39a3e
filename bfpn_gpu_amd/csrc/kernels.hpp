// kernels.hpp — launch wrappers of the HIP kernels (all stream-ordered).
#pragma once

#include "common.hpp"

#include <vector>

namespace bfpn {

// Active energy groups.  Above the highest nonzero group ("top") of a beam
// every group of the state is exactly zero: the energy sweep moves particles
// down in energy (straggling spreads them at most one group up per half
// step), the secondary source emits at or below its source group, and
// transport and angular diffusion never mix groups.  The operators therefore
// skip the groups above the top -- each result they would compute there is
// an exact zero -- and the energy sweep measures the exact top of what it
// writes.  Tops are device ints (-1: no nonzero group), so no host round
// trip is needed; a null pointer means all Ng groups.  A sweep presets (to
// -1) the slots the next sweep writes, so tops rotate over small rings.
struct SweepTops {
    const int* in = nullptr;    // top of the state before the sweep
    const int* xa = nullptr;    // top of the primary state source Qa was made from
    const int* xb = nullptr;    // the same for Qb (paired sweeps)
    const int* qmap = nullptr;  // top of a source from its primary's top t: qmap[t + 1]
    int* out = nullptr;         // receives the top of the result (preset to -1)
    int* out_x = nullptr;       // paired: top of the state between the halves; single: copy of out
    int* reset = nullptr;       // set to -1 by the sweep (the slot a later sweep writes)
    int* reset_x = nullptr;
    // Source rows above q_cap are zero whatever the state (the output groups
    // with no band): they are neither read nor, by the context, written.
    int q_cap = 1 << 30;
};

// Eq.15 energy half step, in place.  coef: Ng device EnergyCoef.  Q1 may be
// null.  When dose != null each block writes FP64 partials:
//   dose[block]        = sum w_yz(s) (F*w1[g] + f*w2[g])  over its lanes
//   dose[nblk + block] = sum f                           (proxy), if proxy
// The lanes handled by a block are [block*256, block*256+256).
template <typename T>
void launch_energy_half(T* F, T* f, const EnergyCoef* coef, const T* Q1, int Ng, long long lanes,
                        bool straggling, const double* yz_weight,
                        int nang, double* dose_partials, bool proxy, hipStream_t s,
                        const SweepTops& tops = SweepTops{});
// Dose partials written by a sweep over `lanes` lanes of `elem`-byte state.
int energy_blocks(long long lanes, size_t elem);

// Two consecutive energy half steps in one sweep (end of depth step n with
// source Qa and the dose/proxy partials of step n, start of step n+1 with
// source Qb); XF/Xf (optional) receive the state between them, at every group
// or, with xmask (device, bit g % 32 of word g / 32), at the groups it sets.
template <typename T>
void launch_energy_pair(T* F, T* f, const EnergyCoef* coef, const T* Qa, const T* Qb, int Ng,
                        long long lanes, bool straggling, const double* yz_weight, int nang,
                        double* dose_partials, bool proxy, T* XF, T* Xf, hipStream_t s,
                        const SweepTops& tops = SweepTops{}, const unsigned* xmask = nullptr);
// First-order upwind energy step, in place (legacy energy model; the
// lite-memory step with f == null).  uc: Ng UpwindCoef; coef supplies the
// dose weights w1/w2 when dose_partials != null (partials as above).
template <typename T>
void launch_energy_upwind(T* F, T* f, const UpwindCoef* uc, const EnergyCoef* coef, const T* Q1,
                          int Ng, long long lanes, bool straggling, const double* yz_weight,
                          int nang, double* dose_partials, bool proxy, hipStream_t s,
                          const SweepTops& tops = SweepTops{});
// Block count of launch_energy_upwind's partials.
inline int upwind_blocks(long long lanes) { return static_cast<int>((lanes + 255) / 256); }

// Split-BF16 operands of the FP32 source (matrix cores), built once per beam.
struct SourcePlan {
    const __bf16* kth = nullptr;  // Kv transposed [g][out][in], nap x nap, hi part
    const __bf16* ktl = nullptr;  //                                      lo part
    const __bf16* e1h = nullptr;  // dg*K1  [erows][ngp], hi / lo
    const __bf16* e1l = nullptr;
    const __bf16* e2h = nullptr;  // dg*K2/3
    const __bf16* e2l = nullptr;
    const int2* tile_band = nullptr;  // [min beg, max end) per 16-group tile
    // Kv block sparsity of the angular GEMM: per (g, 64-direction tile) the
    // 32-wide K tiles holding a nonzero entry, ascending.
    const int* kt_count = nullptr;    // [Ng][mtiles]
    const int* kt_list = nullptr;     // [Ng][mtiles][ktiles] (32-bit: scalar loads)
    // Source groups whose A rows the band GEMM reads: inside some output
    // group's band and with a nonzero Kv (the others' A rows are zero or never
    // read, so the angular GEMM skips them).
    const int* glist = nullptr;        // [ngroups] ascending
    const int* gcount = nullptr;       // [Ng + 1]: entries of glist below g
    const unsigned* amask = nullptr;   // [ngp / 32]: bit g % 32 of word g / 32 set for glist
    int ngroups = 0;
    // Band GEMM as one march over the source groups (when every 32-wide step's
    // window of live 16-group output tiles fits source_march_slots()): per
    // step the first and last tile whose band holds it (empty: x > y), and
    // the steps [march_first, march_last) that hold any.
    const int2* march_steps = nullptr;  // [ngp / 32]
    int march_first = 0, march_last = 0;
    bool march = false;
};
int source_march_slots();  // output tiles a march step may hold at once
int source_nap(int nang);    // nang padded to 32
int source_ngp(int Ng);      // Ng padded to 32 (E columns)
int source_erows(int Ng);    // Ng padded to 64 (E rows); tile_band has erows/16 entries
void launch_split_transpose(const float* kv, __bf16* hi, __bf16* lo, int Ng, int nang,
                            hipStream_t s);
void launch_split(const float* src, __bf16* hi, __bf16* lo, long long n, hipStream_t s);
// Plain streaming copy, bytes % 16 == 0.
void launch_copy16(const void* src, void* dst, size_t bytes, hipStream_t s);
// flags[g][mt][kt] = 1 if the 64 x 32 block (directions mt, K tile kt) of the
// split transposed Kv holds a nonzero entry.
void launch_kv_block_flags(const __bf16* kth, unsigned char* flags, int Ng, int nang, hipStream_t s);
int source_mtiles(int nang);  // 64-direction output tiles of the angular GEMM
int source_ktiles(int nang);  // 32-wide K tiles
// FP32 source on the matrix cores: A1/A2 scratch, Q output.
// `inputs_read` (optional) is recorded on s once the angular GEMM, the only
// reader of Fp and fp, has been issued.  xtop/qmap (optional, device; both or
// neither): the primary's top group and the source-top map of SweepTops;
// rows of Q above the source's top are then left unwritten.
// q_rows (optional, >= 0): rows of Q above it need not be written (the
// caller never reads them: they hold no source).
void launch_source_mfma(const float* Fp, const float* fp, float* A1, float* A2, float* Q,
                        const SourcePlan& plan, int Ng, int nyz, int nang, hipStream_t s,
                        hipEvent_t inputs_read = nullptr, const int* xtop = nullptr,
                        const int* qmap = nullptr, int q_rows = -1);

// Secondary source (FP64 vector path): A1/A2 angular contraction, banded sum.
// kcount/klist (optional, device; from source_angular_blocks): per (g,
// 64-direction tile) the 16-deep K blocks of Kv with a nonzero entry.
// need (optional, device, [Ng]; from source_needed_groups): the source groups
// whose A rows the banded sum reads; the others are neither written nor read
// (their contributions are exact zeros).
template <typename T>
void launch_source_angular(const T* Fp, const T* fp, const T* kv, T* A1, T* A2, int Ng, int nyz,
                           int nang, hipStream_t s, const int* xtop = nullptr,
                           const int* kcount = nullptr, const int* klist = nullptr,
                           const int* need = nullptr);
// Host: the nonzero K blocks of Kv ([g][in][out], FP64) for launch_source_angular.
void source_angular_blocks(const std::vector<double>& kv, int Ng, int nang, std::vector<int>& count,
                           std::vector<int>& list);
// Host: need[g] = 1 if source group g lies in some output group's band
// [beg, end) and its Kv (count from source_angular_blocks) is not all zero.
std::vector<int> source_needed_groups(const std::vector<int>& count, const std::vector<int>& beg,
                                      const std::vector<int>& end, int Ng, int nang);
template <typename T>
void launch_source_energy(const T* A1, const T* A2, const double* e1, const double* e2,
                          const int* beg, const int* end, T* Q, int Ng, long long lanes, double dg,
                          hipStream_t s, const int* xtop = nullptr, const int* qmap = nullptr,
                          const int* need = nullptr);

// MUSCL predictor-corrector transport half step, in place (no scratch).
// top (optional, device): groups above it are zero and are skipped.
template <typename T>
void launch_transport_half(T* F, const T* omend, int Ny, int Nz, int Ng, int nang, double dy,
                           double dz, double h, bool preserve_sign, hipStream_t s,
                           const int* top = nullptr);

// CN angular diffusion on the DST-I basis, in place.  factors: Ng*nang (T),
// su: Nom*Nom, sv: Nmu*Nmu (T).
template <typename T>
void launch_angle_full(T* F, const T* factors, const T* su, const T* sv, int Ng, int nyz, int Nmu,
                       int Nom, hipStream_t s, const int* top = nullptr);

// Sine bases of the MFMA DST: 16 x the basis (a power of two, exact), split
// into two FP16 terms (hi + lo) from the FP64 values, 32x32 zero padded,
// row-major: Su as is, and Su and Sv with their columns in the
// register-chaining order of the DST kernel.
constexpr float kAngleBasisScale = 16.0f;
struct AnglePlan {
    const _Float16* su_h = nullptr;
    const _Float16* su_l = nullptr;
    const _Float16* sup_h = nullptr;
    const _Float16* sup_l = nullptr;
    const _Float16* svp_h = nullptr;
    const _Float16* svp_l = nullptr;
};
// K slot 8q + 4u + r of a chained MFMA operand <-> matrix index 16u + 4q + r.
constexpr int chain_index(int slot) { return 16 * ((slot & 7) >> 2) + 4 * (slot >> 3) + (slot & 3); }
// FP32 DST on the matrix cores (Nmu, Nom <= 32), in place.
void launch_angle_mfma(float* F, const float* factors, const AnglePlan& plan, int Ng, int nyz,
                       int Nmu, int Nom, hipStream_t s, const int* top = nullptr);

// Eq.24 dose of (F+F1, f+f1) into per-block partials (count = dose_blocks()).
template <typename T>
void launch_dose(const T* F, const T* f, const T* F1, const T* f1, const DoseCoef* w,
                 const double* yz_weight, int Ng, long long lanes, int nang, double* partials,
                 hipStream_t s);
int dose_blocks(int Ng, long long lanes);

// Coefficient tables built on the device from device tables (same values as
// the host builders energy_coefs, dose_coefs and yz_weights).
void launch_energy_coefs(const double* S, const double* sigma, const double* Tc, bool straggling,
                         double h, double dg, double rho, int Ng, EnergyCoef* out, hipStream_t s);
void launch_dose_coefs(const double* S, const double* sigma, const double* en, double dg,
                       bool catastrophic, int Ng, DoseCoef* out, hipStream_t s);
void launch_yz_weights(int Ny, int Nz, bool trapezoidal, double* out, hipStream_t s);

// out[slot] = scale * sum(a[0:na]) + scale * sum(b[0:nb])  (b may be null), one block.
void launch_finish_sum(const double* a, int na, const double* b, int nb, double scale, double* out,
                       hipStream_t s);
// *out = ((0 + v[0]) + v[1]) + ... + v[n-1], in that order.
void launch_sum_in_order(const double* v, int n, double* out, hipStream_t s);

// Node subsets of a state [g][node][direction] (Ng x nyz x nang): gather
// packs the nodes listed in `nodes` (n of them) into dst [g][k][direction];
// scatter writes them back.
template <typename T>
void launch_gather_nodes(T* dst, const T* src, const int* nodes, int n, long long nyz, int nang,
                         int Ng, hipStream_t s);
template <typename T>
void launch_scatter_nodes(T* dst, const T* src, const int* nodes, int n, long long nyz, int nang,
                          int Ng, hipStream_t s);
// Sets the listed nodes of a state to zero.
template <typename T>
void launch_zero_nodes(T* dst, const int* nodes, int n, long long nyz, int nang, int Ng,
                       hipStream_t s);

// F[g][s][m] = space_angle[s*nang+m] * energy[g]
template <typename T>
void launch_beam(T* F, const double* space_angle, const double* energy, int Ng, long long lanes,
                 hipStream_t s);

}  // namespace bfpn

namespace bfpn {
// dst[k] = (T) src[k], k < n (device to device)
template <typename T>
void launch_convert(const double* src, T* dst, long long n, hipStream_t s);
// Eq.25 energy moments of psi^1 = F + F1 and psi^2 = f + f1 per group (out:
// 2 Ng doubles, weight = du dv dy dz); partials: moment_partial_count doubles.
size_t moment_partial_count(int Ng, long long lanes);
template <typename T>
void launch_energy_moments(const T* F, const T* f, const T* F1, const T* f1, int Ng, long long lanes,
                           double weight, double* partials, double* out, hipStream_t s);

// Y-Z dose plane (nyz doubles): per node, the Eq.24 dose weights of `coef`
// applied to psi^1 and psi^2, summed over groups and directions, times scale.
template <typename T>
void launch_spot_plane(const T* F, const T* f, const T* F1, const T* f1, const EnergyCoef* coef,
                       int Ng, int nyz, int nang, double scale, double* plane, hipStream_t s);

}  // namespace bfpn
