// solver.hip — shared physics tables, per-beam contexts, the Strang depth
// march, and the context/operator C ABI.
//
// Reference orchestration: BFPnSolver (src/bfpm_solver.cu:1171-1568,
// 2166-2267, 3238-3502).  Its loop body per depth step is
//   stepPrimary:   E(dt/2) -> T(dt/2) -> A(dt) -> T(dt/2) -> E(dt/2)     on (F, f)
//   stepSecondary: the same order on (F1, f1) with the source from the primary
//                  (only when step > 0 and the secondary is enabled)
//   IDD (Eq.24) of F+F1, f+f1 and the proxy dg*sum(f), recorded at x=(k+1)dt,
// with a host sync after every operator and two scalar D2H copies per step.
//
// Here a step is stream-ordered with no host sync, on two HIP streams when
// the secondary is on: the primary march runs on the context's stream; the
// source, the secondary march and the IDD reduction on a second one, which
// waits for the primary of its step.  The primary of the next step waits only
// until the source's angular GEMM (its last reader) has been issued and run,
// so it overlaps the band GEMM and the secondary march of the previous step
// (kernels of the two streams fill each other's tails and latency gaps).
// Transport and angular diffusion update each state array in
// place, the IDD and proxy are accumulated inside the final energy half steps
// and reduced into a device history array, and one source evaluation per step
// feeds both secondary energy half steps (the primary is unchanged in
// between, so the reference's two evaluations are identical).  Tables that do
// not depend on the beam live in a Physics object shared between contexts.

#include <hip/hip_runtime.h>

#include <algorithm>
#include <cmath>
#include <cstdio>
#include <cstring>
#include <memory>
#include <mutex>
#include <sstream>
#include <string>
#include <type_traits>
#include <vector>

#include "device.hpp"
#include "kernels.hpp"
#include "solver.hpp"
#include "tables.hpp"

namespace bfpn {

namespace {
thread_local std::string g_last_error;
}

void set_last_error(const std::string& msg) { g_last_error = msg; }

std::vector<double> yz_weights(int Ny, int Nz, bool trapezoidal) {
    std::vector<double> w(static_cast<size_t>(Ny + 1) * (Nz + 1), 1.0);
    if (trapezoidal)
        for (int iz = 0; iz <= Nz; ++iz)
            for (int iy = 0; iy <= Ny; ++iy) {
                double v = 1.0;
                if (iy == 0 || iy == Ny) v *= 0.5;
                if (iz == 0 || iz == Nz) v *= 0.5;
                w[static_cast<size_t>(iz) * (Ny + 1) + iy] = v;
            }
    return w;
}

void SourcePlanBuffers::build(const float* kv_f32, const std::vector<double>& e1,
                              const std::vector<double>& e2, const std::vector<int>& beg,
                              const std::vector<int>& end, int Ng, int nang, double dg,
                              hipStream_t s) {
    const int nap = source_nap(nang), ngp = source_ngp(Ng), erows = source_erows(Ng);
    const size_t kt = static_cast<size_t>(Ng) * nap * nap;
    kth.ensure(kt * sizeof(__bf16));
    ktl.ensure(kt * sizeof(__bf16));
    launch_split_transpose(kv_f32, kth.as<__bf16>(), ktl.as<__bf16>(), Ng, nang, s);
    {  // nonzero K tiles per (g, direction tile)
        const int mtiles = source_mtiles(nang), ktiles = source_ktiles(nang);
        const size_t nb = static_cast<size_t>(Ng) * mtiles * ktiles;
        scratch.ensure(nb);
        launch_kv_block_flags(kth.as<__bf16>(), scratch.as<unsigned char>(), Ng, nang, s);
        std::vector<unsigned char> flags(nb);
        BFPN_HIP(hipMemcpyAsync(flags.data(), scratch.as<void>(), nb, hipMemcpyDeviceToHost, s));
        BFPN_HIP(hipStreamSynchronize(s));
        std::vector<int> count(static_cast<size_t>(Ng) * mtiles, 0);
        std::vector<int> list(nb, 0);
        for (size_t b = 0; b < count.size(); ++b)
            for (int k = 0; k < ktiles; ++k)
                if (flags[b * ktiles + k]) list[b * ktiles + count[b]++] = k;
        upload(kt_count, count);
        upload(kt_list, list);
        // source groups the band GEMM reads with a nonzero Kv
        std::vector<char> in_band(Ng, 0);
        for (int r = 0; r < Ng; ++r)
            for (int g = std::max(0, beg[r]); g < std::min(Ng, end[r]); ++g) in_band[g] = 1;
        std::vector<int> gl, gc(static_cast<size_t>(Ng) + 1, 0);
        std::vector<unsigned> am(static_cast<size_t>(ngp) / 32, 0u);
        for (int g = 0; g < Ng; ++g) {
            bool nz = false;
            for (int mt = 0; mt < mtiles && !nz; ++mt) nz = count[static_cast<size_t>(g) * mtiles + mt] > 0;
            gc[g + 1] = gc[g];
            if (nz && in_band[g]) {
                gl.push_back(g);
                ++gc[g + 1];
                am[g / 32] |= 1u << (g % 32);
            }
        }
        upload(glist, gl.empty() ? std::vector<int>{0} : gl);
        upload(gcount, gc);
        upload(amask, am);
        plan.ngroups = static_cast<int>(gl.size());
    }
    const size_t ne = static_cast<size_t>(erows) * ngp;
    std::vector<float> E1(ne, 0.f), E2(ne, 0.f);
    for (int r = 0; r < Ng; ++r)
        for (int c = 0; c < Ng; ++c) {
            const size_t src = static_cast<size_t>(r) * Ng + c, dst = static_cast<size_t>(r) * ngp + c;
            E1[dst] = static_cast<float>(dg * e1[src]);
            E2[dst] = static_cast<float>(dg * e2[src] / 3.0);
        }
    for (DeviceBuffer* b : {&e1h, &e1l, &e2h, &e2l}) b->ensure(ne * sizeof(__bf16));
    upload(scratch, E1);
    launch_split(scratch.as<float>(), e1h.as<__bf16>(), e1l.as<__bf16>(), static_cast<long long>(ne), s);
    BFPN_HIP(hipStreamSynchronize(s));
    upload(scratch, E2);
    launch_split(scratch.as<float>(), e2h.as<__bf16>(), e2l.as<__bf16>(), static_cast<long long>(ne), s);
    std::vector<int2> tiles(erows / 16);
    for (size_t t = 0; t < tiles.size(); ++t) {
        int lo = 1 << 30, hi = 0;
        for (int r = static_cast<int>(t) * 16; r < static_cast<int>(t) * 16 + 16 && r < Ng; ++r)
            if (end[r] > beg[r]) {
                lo = std::min(lo, beg[r]);
                hi = std::max(hi, end[r]);
            }
        tiles[t] = hi > 0 ? make_int2(lo, hi) : make_int2(0, 0);
    }
    upload(band, tiles);
    {  // per 32-wide source step: the window of output tiles whose band holds it
        const int nsteps = ngp / 32;
        std::vector<int2> st(nsteps, make_int2(1, 0));
        int first = nsteps, last = 0, widest = 0;
        for (int k = 0; k < nsteps; ++k) {
            for (int t = 0; t < static_cast<int>(tiles.size()); ++t) {
                const int2 b = tiles[t];
                if (b.y > b.x && b.x < 32 * k + 32 && b.y > 32 * k) {
                    if (st[k].x > st[k].y) st[k] = make_int2(t, t);
                    st[k].x = std::min(st[k].x, t);
                    st[k].y = std::max(st[k].y, t);
                }
            }
            if (st[k].x <= st[k].y) {
                first = std::min(first, k);
                last = k + 1;
                widest = std::max(widest, st[k].y - st[k].x + 1);
            }
        }
        upload(steps, st);
        plan.march_first = std::min(first, last);
        plan.march_last = last;
        plan.march = widest <= source_march_slots();
    }
    BFPN_HIP(hipStreamSynchronize(s));
    scratch.release();
    plan.kth = kth.as<__bf16>();
    plan.ktl = ktl.as<__bf16>();
    plan.e1h = e1h.as<__bf16>();
    plan.e1l = e1l.as<__bf16>();
    plan.e2h = e2h.as<__bf16>();
    plan.e2l = e2l.as<__bf16>();
    plan.tile_band = band.as<int2>();
    plan.kt_count = kt_count.as<int>();
    plan.kt_list = kt_list.as<int>();
    plan.glist = glist.as<int>();
    plan.gcount = gcount.as<int>();
    plan.amask = amask.as<unsigned>();
    plan.march_steps = steps.as<int2>();
}


namespace {
// Round to the nearest FP16 (ties to even, subnormals kept), as a bit pattern.
uint16_t half_bits(double v) {
    const _Float16 h = static_cast<_Float16>(static_cast<float>(v));
    uint16_t b;
    std::memcpy(&b, &h, 2);
    return b;
}
double half_value(uint16_t b) {
    _Float16 h;
    std::memcpy(&h, &b, 2);
    return static_cast<double>(h);
}
}  // namespace

void AnglePlanBuffers::build(int Nmu, int Nom, hipStream_t) {
    // Su, Su (chained columns), Sv (chained columns); each zero padded to
    // 32x32, scaled by kAngleBasisScale (exact) and split into FP16 hi + lo
    // from the FP64 basis.  (The scale keeps the smallest basis values' lo
    // terms clear of the FP16 subnormal range.)
    std::vector<double> pad(3 * 32 * 32, 0.0);
    const auto su = sine_basis(Nom), sv = sine_basis(Nmu);
    for (int r = 0; r < 32; ++r)
        for (int c = 0; c < 32; ++c) {
            const int cc = chain_index(c);
            if (r < Nom && c < Nom) pad[r * 32 + c] = su[r * Nom + c];
            if (r < Nom && cc < Nom) pad[1024 + r * 32 + c] = su[r * Nom + cc];
            if (r < Nmu && cc < Nmu) pad[2048 + r * 32 + c] = sv[r * Nmu + cc];
        }
    std::vector<uint16_t> vh(pad.size()), vl(pad.size());
    for (size_t k = 0; k < pad.size(); ++k) {
        const double x = kAngleBasisScale * pad[k];
        vh[k] = half_bits(x);
        vl[k] = half_bits(x - half_value(vh[k]));
    }
    upload(h, vh);
    upload(l, vl);
    const _Float16* H = h.as<_Float16>();
    const _Float16* L = l.as<_Float16>();
    plan.su_h = H;
    plan.su_l = L;
    plan.sup_h = H + 1024;
    plan.sup_l = L + 1024;
    plan.svp_h = H + 2048;
    plan.svp_l = L + 2048;
}

void validate_params(const bfpn_grid& g, const bfpn_phys& p) {
    BFPN_REQUIRE(g.Nx > 0 && g.Ny > 0 && g.Nz > 0 && g.Ng > 1 && g.Nmu > 0 && g.Nom > 0,
                 "grid sizes must be positive (Ng > 1)");
    BFPN_REQUIRE(g.Nmu * g.Nom <= 1024, "Nmu*Nom must be <= 1024");
    BFPN_REQUIRE(g.Lx > 0 && g.Ly > 0 && g.Lz > 0 && g.Lu > 0 && g.Lv > 0 && g.Lg > 0,
                 "domain lengths must be positive");
    BFPN_REQUIRE(p.density > 0 && p.energy_density > 0, "density must be positive");
    BFPN_REQUIRE(p.sig_E > 0, "sigma-e must be positive");
}

std::string physics_key(const bfpn_grid& g, const bfpn_phys& p, const std::string& data_path,
                        int precision) {
    std::ostringstream k;
    k.precision(17);
    k << precision << '|' << data_path << '|' << g.Nx << ',' << g.Ny << ',' << g.Nz << ',' << g.Ng
      << ',' << g.Nmu << ',' << g.Nom << '|' << g.Lx << ',' << g.Ly << ',' << g.Lz << ',' << g.Lu
      << ',' << g.Lv << ',' << g.Lg << '|' << p.density << ',' << p.energy_density << '|'
      << p.material_name << '|' << p.primary_only << p.eq15_straggling << p.energy_only
      << p.trapezoidal_yz << p.dose_catastrophic_loss << p.calibrate_cross_sections
      << p.legacy_energy << p.lite_memory;
    return k.str();
}

namespace {

enum Phase { kEnergy = 0, kTransport = 1, kAngle = 2, kSource = 3, kDose = 4, kWait = 5 };

// Top group of the secondary source made from a primary state whose top is
// t: entry t + 1 (t = -1 .. Ng-1) is the highest output group whose kernel
// row has a nonzero entry at a source group <= t (-1: none).
std::vector<int> source_top_map(const SourceKernels& ker, int Ng) {
    std::vector<int> first(Ng, -1), map(static_cast<size_t>(Ng) + 1, -1);
    for (int g = 0; g < Ng; ++g) {
        const int b = std::max(0, ker.beg[g]), e = std::min(Ng, ker.end[g]);
        if (b < e) first[b] = std::max(first[b], g);
    }
    for (int t = 0; t < Ng; ++t) map[t + 1] = std::max(map[t], first[t]);
    return map;
}

// Highest group with a nonzero beam value (-1: none).
int beam_top(const BeamFactors& b) {
    for (int g = static_cast<int>(b.energy.size()) - 1; g >= 0; --g)
        if (b.energy[g] != 0.0) return g;
    return -1;
}

template <typename T>
struct Physics final : PhysicsBase {
    DeviceBuffer coef, dose_w, yzw, om_dev, factors, su, sv, kv, e1, e2, beg, end, qmap, ucoef;
    DeviceBuffer kcount, klist, need;  // FP64 source: nonzero Kv blocks, groups read
    DeviceBuffer xmask64;              // FP64: the groups read, as bits (FP32: plan.amask)
    bool legacy = false, lite = false;  // upwind energy model; lite-memory primary step
    double density = 1.0;               // g/cm^3: the IDD and spot-plane scale
    int qcap = -1;                      // highest output group with a source band (-1: none)
    SourcePlanBuffers plan;
    AnglePlanBuffers angle_plan;

    Physics(const bfpn_grid& g, const bfpn_phys& p, const std::string& data_path, int prec) {
        precision = prec;
        grid = g;
        d = compute_deltas(g);
        nyz = (g.Ny + 1) * (g.Nz + 1);
        nang = g.Nmu * g.Nom;
        lanes = static_cast<long long>(nyz) * nang;
        cells = static_cast<size_t>(g.Ng) * lanes;
        secondary = !p.primary_only && !p.energy_only && !p.lite_memory;
        legacy = p.legacy_energy != 0;
        lite = p.lite_memory != 0;
        density = p.density;
        key = physics_key(g, p, data_path, prec);

        // host tables (initialize, src/bfpm_solver.cu:1435-1568)
        mat = load_material_tables(data_path, g.Ng, p.calibrate_cross_sections != 0);
        S = stopping_power_table(p.material_name, g.Ng, g.Lg, d.dg);
        Tc = straggling_table(g.Ng);
        sig_tr = transport_xs_table(g.Ng, d.dg);
        en = energy_grid(g.Ng, d.dg);
        om = direction_table(g, d);
        ker = build_source_kernels(g.Ng, d.dg, nang, mat.mu, mat.sigma, om);

        hipStream_t s = nullptr;
        BFPN_HIP(hipStreamCreateWithFlags(&s, hipStreamNonBlocking));
        try {
            auto ec = energy_coefs(S, mat.sigma, Tc.data(), p.eq15_straggling != 0, 0.5 * d.dt, d.dg,
                                   p.energy_density, g.Ng);
            const auto dc = dose_coefs(S, mat.sigma, en, d.dg, p.dose_catastrophic_loss != 0, g.Ng);
            for (int k = 0; k < g.Ng; ++k) {  // the closing E-half of a step accumulates the IDD
                ec[k].w1 = dc[k].w1;
                ec[k].w2 = dc[k].w2;
            }
            upload(coef, ec);
            upload(dose_w, dc);
            if (legacy || lite)  // lite: one full step on psi^1 without density or straggling
                upload(ucoef, upwind_coefs(S, mat.sigma, Tc.data(),
                                           !lite && p.eq15_straggling != 0,
                                           lite ? d.dt : 0.5 * d.dt, d.dg,
                                           lite ? 1.0 : p.energy_density, g.Ng));
            upload(yzw, yz_weights(g.Ny, g.Nz, p.trapezoidal_yz != 0));
            upload(om_dev, cast_vector<T>(om));
            upload(factors, cast_vector<T>(angle_factors(sig_tr.data(), g.Ng, g.Nmu, g.Nom, d.du,
                                                         d.dv, d.dt, p.density)));
            upload(su, cast_vector<T>(sine_basis(g.Nom)));
            upload(sv, cast_vector<T>(sine_basis(g.Nmu)));
            if (std::is_same_v<T, float> &&
                AnglePlanBuffers::usable(g.Nmu, g.Nom, static_cast<long long>(g.Ng) * nyz))
                angle_plan.build(g.Nmu, g.Nom, s);
            if (secondary) {
                for (int r = 0; r < g.Ng; ++r)
                    if (ker.end[r] > ker.beg[r]) qcap = r;
                upload(qmap, source_top_map(ker, g.Ng));
                upload(kv, cast_vector<T>(ker.v));
                if constexpr (std::is_same_v<T, float>) {
                    plan.build(kv.as<float>(), ker.e1, ker.e2, ker.beg, ker.end, g.Ng, nang, d.dg, s);
                    kv.release();
                } else {
                    upload(e1, ker.e1);
                    upload(e2, ker.e2);
                    upload(beg, ker.beg);
                    upload(end, ker.end);
                    std::vector<int> cnt, lst;
                    source_angular_blocks(ker.v, g.Ng, nang, cnt, lst);
                    upload(kcount, cnt);
                    upload(klist, lst);
                    const std::vector<int> nd = source_needed_groups(cnt, ker.beg, ker.end, g.Ng,
                                                                     nang);
                    upload(need, nd);
                    std::vector<unsigned> bits((g.Ng + 31) / 32, 0u);
                    for (int k = 0; k < g.Ng; ++k)
                        if (nd[k]) bits[k / 32] |= 1u << (k % 32);
                    upload(xmask64, bits);
                }
            }
            BFPN_HIP(hipStreamSynchronize(s));
        } catch (...) {
            (void)hipStreamDestroy(s);
            throw;
        }
        BFPN_HIP(hipStreamDestroy(s));
    }

    size_t device_bytes() const override {
        return coef.bytes() + dose_w.bytes() + yzw.bytes() + ucoef.bytes() + om_dev.bytes() +
               factors.bytes() +
               su.bytes() + sv.bytes() + kv.bytes() + e1.bytes() + e2.bytes() + beg.bytes() +
               end.bytes() + qmap.bytes() + kcount.bytes() + klist.bytes() + need.bytes() +
               xmask64.bytes() +
               plan.bytes() +
               angle_plan.h.bytes() +
               angle_plan.l.bytes();
    }
};

template <typename T>
class Context final : public bfpn_ctx {
public:
    Context(std::shared_ptr<Physics<T>> P, const bfpn_grid& g, const bfpn_phys& p) : P_(std::move(P)) {
        grid = g;
        phys = p;
        physics = P_;
        d = P_->d;
        cells = P_->cells;
        const BeamFactors beam = beam_factors(g, p, d, P_->om);
        BFPN_HIP(hipStreamCreateWithFlags(&stream, hipStreamNonBlocking));
        done_stream = stream;
        if (P_->secondary) {
            BFPN_HIP(hipStreamCreateWithFlags(&stream2_, hipStreamNonBlocking));
            done_stream = stream2_;
            BFPN_HIP(hipEventCreateWithFlags(&ev_prim_, hipEventDisableTiming));
            BFPN_HIP(hipEventCreateWithFlags(&ev_src_, hipEventDisableTiming));
            BFPN_HIP(hipEventCreateWithFlags(&ev_tail_, hipEventDisableTiming));
        }
        const size_t state_bytes = cells * sizeof(T);
        // lite-memory keeps psi^1 of the primary only
        for (int k = 0; k < (P_->lite ? 1 : 4); ++k) state_[k].ensure(state_bytes);
        if (P_->secondary) {
            a1_.ensure(state_bytes);
            a2_.ensure(state_bytes);
            for (auto& q : q_) q.ensure(state_bytes);
            for (auto& x : x_) x.ensure(state_bytes);
        }
        for (int k = 0; k < 4; ++k) cur_[k] = state_[k].as<T>();
        eb_ = (P_->legacy || P_->lite) ? upwind_blocks(P_->lanes)
                                       : energy_blocks(P_->lanes, sizeof(T));
        partial_p_.ensure(4 * 2 * eb_ * sizeof(double));  // four slots: steps rotate
        partial_s_.ensure(eb_ * sizeof(double));
        grow_history(std::max(16, g.Nx));
        // active-group tops: primary ring, secondary ring, end-of-step primary ring
        tops_.ensure(kTopSlots * sizeof(int));
        init_tops(beam_top(beam), -1);
        // Eq.23 beam on the device (the reference builds a full host array)
        DeviceBuffer sa, ee;
        upload(sa, beam.space_angle);
        upload(ee, beam.energy);
        launch_beam<T>(cur_[0], sa.as<double>(), ee.as<double>(), g.Ng, P_->lanes, stream);
        for (int k = 1; k < 4; ++k)
            if (cur_[k]) BFPN_HIP(hipMemsetAsync(cur_[k], 0, state_bytes, stream));
        BFPN_HIP(hipStreamSynchronize(stream));
    }

    ~Context() override {
        if (stream2_) {
            (void)hipStreamSynchronize(stream2_);
            (void)hipStreamDestroy(stream2_);
        }
        for (hipEvent_t e : {ev_prim_, ev_src_, ev_tail_})
            if (e) (void)hipEventDestroy(e);
        if (stream) {
            (void)hipStreamSynchronize(stream);
            (void)hipStreamDestroy(stream);
        }
    }

    size_t device_bytes() const override {
        size_t b = a1_.bytes() + a2_.bytes() + q_[0].bytes() + q_[1].bytes() + x_[0].bytes() +
                   x_[1].bytes() + hist_.bytes() + partial_p_.bytes() +
                   partial_s_.bytes();
        for (const auto& s : state_) b += s.bytes();
        for (size_t k = 1; k < layers_.size(); ++k) b += layers_[k]->device_bytes();
        for (const auto& c : cbuf_) b += c.bytes();
        for (const auto& c : cbuf2_) b += c.bytes();
        for (const auto& c : xm_) b += c.bytes();
        for (const auto& c : qm_) b += c.bytes();
        b += map_nodes_.bytes() + map_yzw_.bytes() + map_tmp_.bytes();
        for (const auto& m : mats_)
            if (m != P_) b += m->device_bytes();
        return b;
    }

    // Depth layers: layer k's tables serve the steps whose midpoint depth
    // lies below x_end[k] (the last layer: any depth).  The tables must share
    // this context's grid and flags; the first layer is the one it was built on.
    void set_layers(std::vector<std::shared_ptr<Physics<T>>> layers, std::vector<double> x_end) {
        BFPN_REQUIRE(!layers.empty() && layers.front() == P_ && x_end.size() == layers.size(),
                     "layer tables do not match the context");
        layers_ = std::move(layers);
        x_end_ = std::move(x_end);
    }

    // Material map: the node-local operators (energy, angular diffusion,
    // source, dose) of node s in depth slab j use the tables of material
    // map[j nyz + s]; slab j covers the depth steps whose midpoint lies below
    // slab_end[j] (the last: any depth).  Each operator runs per material with
    // the same kernels as a homogeneous run: the slab's largest material in
    // place over every node (the others' nodes copied out and zeroed first,
    // so it gives them nothing), the others on their nodes packed together,
    // then written back; transport, which does not depend on the material,
    // runs on the whole state.  Streams and active-group
    // skipping as in a homogeneous run (each material's sweep publishes into
    // the same top slots).
    void set_map(std::vector<std::shared_ptr<Physics<T>>> mats, std::vector<double> slab_end,
                 const std::vector<unsigned char>& map) {
        BFPN_REQUIRE(!mats.empty() && !slab_end.empty(), "a material map needs materials and slabs");
        for (const auto& m : mats)
            BFPN_REQUIRE(!m->legacy && !m->lite,
                         "material maps support the eq15 energy model with the full state only");
        const long long nyz = P_->nyz;
        const int nang = P_->nang, Ng = grid.Ng;
        BFPN_REQUIRE(map.size() == slab_end.size() * static_cast<size_t>(nyz),
                     "the map needs one material index per node and slab");
        const auto w = yz_weights(grid.Ny, grid.Nz, phys.trapezoidal_yz != 0);
        std::vector<int> lists;
        std::vector<double> weights;
        long long max_nodes = 0, pslot = 0;
        for (size_t j = 0; j < slab_end.size(); ++j) {
            std::vector<std::vector<int>> by(mats.size());
            for (long long s = 0; s < nyz; ++s) {
                const unsigned m = map[j * nyz + s];
                BFPN_REQUIRE(m < mats.size(), "map entry names no material");
                by[m].push_back(static_cast<int>(s));
            }
            MapSlab sl;
            for (size_t m = 0; m < by.size(); ++m) {
                if (by[m].empty()) continue;
                sl.mat.push_back(static_cast<int>(m));
                sl.count.push_back(static_cast<int>(by[m].size()));
                sl.first.push_back(static_cast<int>(lists.size() - j * nyz));
                for (int s : by[m]) {
                    lists.push_back(s);
                    weights.push_back(w[s]);
                }
            }
            // the material with the most nodes runs in place over every node
            // (its sweeps cover all lanes); the others run packed
            for (size_t k = 1; k < sl.count.size(); ++k)
                if (sl.count[k] > sl.count[sl.major]) sl.major = static_cast<int>(k);
            long long moff = 0;
            int poff = 0;
            for (size_t k = 0; k < sl.mat.size(); ++k) {
                const bool maj = static_cast<int>(k) == sl.major;
                const long long nodes_k = maj ? nyz : sl.count[k];
                const int eb = energy_blocks(nodes_k * nang, sizeof(T));
                sl.eb.push_back(eb);
                sl.poff.push_back(poff);
                poff += 2 * eb;
                sl.moff.push_back(maj ? -1 : moff);
                if (!maj) moff += static_cast<long long>(Ng) * sl.count[k] * nang;
            }
            max_nodes = std::max<long long>(max_nodes, moff);  // packed values of the minority
            pslot = std::max<long long>(pslot, poff);
            slabs_.push_back(std::move(sl));
        }
        upload(map_nodes_, lists);
        upload(map_yzw_, weights);
        const size_t packed_bytes = static_cast<size_t>(max_nodes) * sizeof(T);
        for (auto& b : cbuf_) b.ensure(packed_bytes);
        if (P_->secondary) {  // the secondary's stream packs into its own pair
            for (auto& b : cbuf2_) b.ensure(packed_bytes);
            for (auto& b : xm_) b.ensure(packed_bytes);  // packed end-of-step copies and sources
            for (auto& b : qm_) b.ensure(packed_bytes);
        }
        pslot_ = static_cast<int>(pslot);
        partial_p_.ensure(4 * static_cast<size_t>(pslot) * sizeof(double));
        partial_s_.ensure(static_cast<size_t>(pslot) * sizeof(double));
        map_tmp_.ensure(2 * mats.size() * sizeof(double));
        mats_ = std::move(mats);
        slab_end_ = std::move(slab_end);
    }

    int steps_done() const override { return step_; }

    // Eq.25 moments of the current state (computeEnergyMoments,
    // src/bfpm_solver.cu:3504-3526): [g][psi^1, psi^2] x du dv dy dz.
    void energy_moments(double* host) override {
        if (!mats_.empty())
            throw Error(BFPN_E_UNSUPPORTED, "energy moments are not available with a material map");
        if (P_->lite)
            throw Error(BFPN_E_UNSUPPORTED, "energy moments need psi^2 and the secondary "
                                            "(not kept with --lite-memory)");
        sync_all();
        const long long per = static_cast<long long>(P_->lanes);
        diag_part_.ensure(moment_partial_count(grid.Ng, per) * sizeof(double));
        diag_out_.ensure(2 * static_cast<size_t>(grid.Ng) * sizeof(double));
        launch_energy_moments<T>(cur_[0], cur_[1], cur_[2], cur_[3], grid.Ng, per,
                                 d.du * d.dv * d.dy * d.dz, diag_part_.as<double>(),
                                 diag_out_.as<double>(), stream);
        BFPN_HIP(hipStreamSynchronize(stream));
        BFPN_HIP(hipMemcpy(host, diag_out_.as<double>(), 2 * static_cast<size_t>(grid.Ng) * sizeof(double),
                           hipMemcpyDeviceToHost));
    }

    // Y-Z dose plane of the current state (computeSpotDosePlane,
    // src/bfpm_solver.cu:3560-3585): node s = iz (Ny+1) + iy.
    void spot_plane(double* host) override {
        if (!mats_.empty())
            throw Error(BFPN_E_UNSUPPORTED, "spot planes are not available with a material map");
        if (P_->lite)
            throw Error(BFPN_E_UNSUPPORTED, "spot planes need psi^2 and the secondary "
                                            "(not kept with --lite-memory)");
        sync_all();
        const int nyz = static_cast<int>(P_->nyz);
        diag_out_.ensure(static_cast<size_t>(nyz) * sizeof(double));
        launch_spot_plane<T>(cur_[0], cur_[1], cur_[2], cur_[3], P_->coef.template as<EnergyCoef>(),
                             grid.Ng, nyz, P_->nang, d.du * d.dv / P_->density,
                             diag_out_.as<double>(), stream);
        BFPN_HIP(hipStreamSynchronize(stream));
        BFPN_HIP(hipMemcpy(host, diag_out_.as<double>(), static_cast<size_t>(nyz) * sizeof(double),
                           hipMemcpyDeviceToHost));
    }

    void advance(int nsteps) override {
        BFPN_REQUIRE(nsteps >= 0, "nsteps must be >= 0");
        if (nsteps == 0) return;
        if (step_ + nsteps > hist_cap_) grow_history(std::max(2 * hist_cap_, step_ + nsteps));
        for (auto& ev : events_) ev.clear();
        if (stream2_) {  // work a caller ordered on done_stream precedes these steps
            BFPN_HIP(hipEventRecord(ev_tail_, stream2_));
            BFPN_HIP(hipStreamWaitEvent(stream, ev_tail_, 0));
        }
        for (int k = 0; k < nsteps; ++k) {
            bool first = k == 0, last = k == nsteps - 1;
            if (!mats_.empty()) {  // a paired sweep never spans two slabs
                const int j = slab_of(step_);
                step_mapped(first || j != slab_of(step_ - 1), last || j != slab_of(step_ + 1));
                continue;
            }
            if (!layers_.empty()) {  // a paired sweep never spans two materials
                const int L = layer_of(step_);
                P_ = layers_[L];
                first = first || L != layer_of(step_ - 1);
                last = last || L != layer_of(step_ + 1);
            }
            if (P_->lite)
                lite_step();
            else  // the upwind model has no paired sweep: every step stands alone
                step(first || P_->legacy, last || P_->legacy);
        }
        if (stream2_ && !overlap) {  // all on `stream`: done_stream still ends last
            BFPN_HIP(hipEventRecord(ev_tail_, stream));
            BFPN_HIP(hipStreamWaitEvent(stream2_, ev_tail_, 0));
        }
        if (profiling) collect_profile();
    }

    void history(double* idd, double* proxy, int cap, int* n) override {
        sync_all();
        const int cnt = std::min(cap, step_);
        if (cnt > 0) {
            if (idd)
                BFPN_HIP(hipMemcpy(idd, hist_.as<double>(), cnt * sizeof(double), hipMemcpyDeviceToHost));
            if (proxy)
                BFPN_HIP(hipMemcpy(proxy, hist_.as<double>() + hist_cap_, cnt * sizeof(double),
                                   hipMemcpyDeviceToHost));
        }
        if (n) *n = step_;
    }

    void read_state(int which, double* host, size_t count) override {
        BFPN_REQUIRE(which >= 0 && which < 4, "state index must be 0..3");
        BFPN_REQUIRE(count == cells, "state count mismatch");
        sync_all();
        if (!cur_[which]) {  // not kept (lite-memory): identically zero
            std::fill(host, host + cells, 0.0);
            return;
        }
        std::vector<T> tmp(cells);
        BFPN_HIP(hipMemcpy(tmp.data(), cur_[which], cells * sizeof(T), hipMemcpyDeviceToHost));
        for (size_t k = 0; k < cells; ++k) host[k] = static_cast<double>(tmp[k]);
    }

    void write_state(int which, const double* host, size_t count) override {
        BFPN_REQUIRE(which >= 0 && which < 4, "state index must be 0..3");
        BFPN_REQUIRE(count == cells, "state count mismatch");
        if (!cur_[which])
            throw Error(BFPN_E_UNSUPPORTED, "this state array is not kept with --lite-memory");
        std::vector<T> tmp(host, host + cells);
        sync_all();
        BFPN_HIP(hipMemcpy(cur_[which], tmp.data(), cells * sizeof(T), hipMemcpyHostToDevice));
        // the written values may reach any group: that species' top is Ng-1
        const int all = grid.Ng - 1;
        BFPN_HIP(hipMemcpy(top_slot(which < 2 ? p_cur_ : 3 + s_cur_), &all, sizeof(int),
                           hipMemcpyHostToDevice));
    }

    // Active-group skipping (on by default): bit-identical either way.
    // Turning it on makes every top conservative (Ng-1) until the next sweeps
    // have measured them.
    void set_group_skip(bool on) override {
        if (on && !skip_) {
            sync_all();
            init_tops(grid.Ng - 1, grid.Ng - 1);
        }
        skip_ = on;
    }

    void active_tops(int* primary, int* secondary) override {
        sync_all();
        int h[kTopSlots];
        BFPN_HIP(hipMemcpy(h, tops_.as<int>(), sizeof(h), hipMemcpyDeviceToHost));
        if (primary) *primary = skip_ ? h[p_cur_] : grid.Ng - 1;
        if (secondary) *secondary = skip_ ? h[3 + s_cur_] : grid.Ng - 1;
    }

private:
    // Tops (see SweepTops): slots 0-2 the primary's ring, 3-5 the
    // secondary's, 6-13 the primary state at the end of step n (slot
    // 6 + n%8), which the source of step n and the secondary sweeps using that
    // source (steps n and n+1) read.  A sweep reads its species' current
    // slot, writes the next and presets the one after; the closing primary
    // sweep of step n writes slot 6 + n%8 and presets 6 + (n+1)%8.  The last
    // readers of that preset slot, step n-7's, are long done: the primary of
    // step n-1 has waited for the source of step n-2, which the secondary
    // stream runs after the sweeps of step n-3.
    static constexpr int kTopSlots = 14;
    // Layer of depth step n: the one holding the step's midpoint (n + 1/2) dt.
    int layer_of(int n) const {
        const double x = (n + 0.5) * d.dt;
        for (size_t k = 0; k + 1 < layers_.size(); ++k)
            if (x < x_end_[k]) return static_cast<int>(k);
        return static_cast<int>(layers_.size()) - 1;
    }
    int* top_slot(int k) { return tops_.as<int>() + k; }
    int* x_top(int n) { return top_slot(6 + (n & 7)); }
    const int* cur_top(int species) {
        return skip_ ? top_slot(species == 0 ? p_cur_ : 3 + s_cur_) : nullptr;
    }
    // Host-side reset: current tops of the two species, the rest preset
    // (the slot the next closing sweep writes: -1; the others conservative).
    void init_tops(int primary, int secondary) {
        std::vector<int> h(kTopSlots, -1);
        h[p_cur_] = primary;
        h[3 + s_cur_] = secondary;
        for (int k = 0; k < 8; ++k) h[6 + k] = grid.Ng - 1;
        h[6 + (step_ & 7)] = -1;
        BFPN_HIP(hipMemcpy(tops_.as<int>(), h.data(), h.size() * sizeof(int),
                           hipMemcpyHostToDevice));
    }
    // Tops of the next energy sweep of `species`, which advances its ring:
    // sources made at steps xa / xb (-1: none), end-of-step top written for
    // step x_out (-1: none).  Empty when skipping is off.
    SweepTops sweep_tops(int species, int xa, int xb, int x_out) {
        SweepTops t;
        t.q_cap = P_->qcap;  // static: rows above hold no source
        if (!skip_) return t;
        int& cur = species == 0 ? p_cur_ : s_cur_;
        int* ring = top_slot(3 * species);
        t.in = ring + cur;
        t.out = ring + (cur + 1) % 3;
        t.reset = ring + (cur + 2) % 3;
        cur = (cur + 1) % 3;
        if (xa >= 0) {
            t.xa = x_top(xa);
            t.qmap = P_->qmap.template as<int>();
        }
        if (xb >= 0) t.xb = x_top(xb);
        if (x_out >= 0) {
            t.out_x = x_top(x_out);
            t.reset_x = x_top(x_out + 1);
        }
        return t;
    }

    void sync_all() {
        BFPN_HIP(hipStreamSynchronize(stream));
        if (stream2_) BFPN_HIP(hipStreamSynchronize(stream2_));
    }

    void grow_history(int cap) {
        sync_all();  // the IDD reduction writes the history on the second stream
        DeviceBuffer nb;
        nb.ensure(2 * static_cast<size_t>(cap) * sizeof(double));
        BFPN_HIP(hipMemsetAsync(nb.as<void>(), 0, nb.bytes(), stream));
        if (hist_cap_ > 0 && step_ > 0) {
            BFPN_HIP(hipMemcpyAsync(nb.as<double>(), hist_.as<double>(), step_ * sizeof(double),
                                    hipMemcpyDeviceToDevice, stream));
            BFPN_HIP(hipMemcpyAsync(nb.as<double>() + cap, hist_.as<double>() + hist_cap_,
                                    step_ * sizeof(double), hipMemcpyDeviceToDevice, stream));
        }
        BFPN_HIP(hipStreamSynchronize(stream));
        hist_.swap(nb);
        hist_cap_ = cap;
    }

    // Event on stream s after an operator call; the interval since the
    // previous event on s is charged to phase p together with the call's
    // algorithmic bytes (kWait: the interval ends a wait for the other stream
    // and is not charged).  With both streams busy, a kernel's interval is its
    // duration while sharing the GPU.
    void mark(Phase p, double bytes, hipStream_t s) {
        if (!profiling) return;
        hipEvent_t e;
        BFPN_HIP(hipEventCreate(&e));
        BFPN_HIP(hipEventRecord(e, s));
        events_[s == stream ? 0 : 1].push_back({e, p});
        if (p == kWait) return;
        phase_bytes[p] += bytes;
        phase_calls[p] += bytes > 0.0 ? 1 : 0;
    }

    void collect_profile() {
        sync_all();
        for (auto& ev : events_) {
            for (size_t k = 1; k < ev.size(); ++k) {
                if (ev[k].second == kWait) continue;
                float ms = 0.f;
                BFPN_HIP(hipEventElapsedTime(&ms, ev[k - 1].first, ev[k].first));
                phase_seconds[ev[k].second] += ms * 1e-3;
            }
            for (auto& e : ev) (void)hipEventDestroy(e.first);
            ev.clear();
        }
    }

    double array_bytes() const { return static_cast<double>(cells) * sizeof(T); }

    // One energy half step of a species (source Q or none; dose partials or none).
    void energy(int species, const T* Q, double* partials, bool proxy, hipStream_t s,
                const SweepTops& tops) {
        if (P_->legacy) {
            launch_energy_upwind<T>(cur_[2 * species], cur_[2 * species + 1],
                                    P_->ucoef.template as<UpwindCoef>(),
                                    P_->coef.template as<EnergyCoef>(), Q, grid.Ng, P_->lanes,
                                    phys.eq15_straggling != 0, P_->yzw.template as<double>(),
                                    P_->nang, partials, proxy, s, tops);
            mark(kEnergy, array_bytes() * (Q ? 5 : 4), s);
            return;
        }
        launch_energy_half<T>(cur_[2 * species], cur_[2 * species + 1], P_->coef.template as<EnergyCoef>(),
                              Q, grid.Ng, P_->lanes, phys.eq15_straggling != 0,
                              P_->yzw.template as<double>(), P_->nang, partials, proxy, s, tops);
        mark(kEnergy, array_bytes() * (Q ? 5 : 4), s);
    }

    // The end half of one depth step and the start half of the next in one
    // sweep (sources Qa, Qb or none; the dose of the first); X (optional)
    // receives the state at the end of the step.
    void energy_pair(int species, const T* Qa, const T* Qb, double* partials, bool proxy, T* const* X,
                     hipStream_t s, const SweepTops& tops) {
        launch_energy_pair<T>(cur_[2 * species], cur_[2 * species + 1], P_->coef.template as<EnergyCoef>(),
                              Qa, Qb, grid.Ng, P_->lanes, phys.eq15_straggling != 0,
                              P_->yzw.template as<double>(), P_->nang, partials, proxy,
                              X ? X[0] : nullptr, X ? X[1] : nullptr, s, tops, x_rows());
        mark(kEnergy, array_bytes() * (4 + (Qa ? 2 : 0) + (X ? 2 : 0)), s);
    }

    // The groups of the end-of-step copy that the source reads (the others are
    // never written): the source plan's list as bits.
    const unsigned* x_rows() const {
        if constexpr (std::is_same_v<T, float>)
            return P_->plan.plan.amask;
        else
            return P_->xmask64.template as<unsigned>();
    }

    void transport(int which, bool preserve_sign, hipStream_t s) {
        launch_transport_half<T>(cur_[which], P_->om_dev.template as<T>(), grid.Ny, grid.Nz, grid.Ng,
                                 P_->nang, d.dy, d.dz, 0.5 * d.dt, preserve_sign, s,
                                 cur_top(which / 2));
        mark(kTransport, array_bytes() * 2, s);
    }

    void angle(int which, hipStream_t s) {
        if constexpr (std::is_same_v<T, float>) {
            if (P_->angle_plan.plan.su_h) {
                launch_angle_mfma(cur_[which], P_->factors.template as<float>(), P_->angle_plan.plan,
                                  grid.Ng, P_->nyz, grid.Nmu, grid.Nom, s, cur_top(which / 2));
                mark(kAngle, array_bytes() * 2, s);
                return;
            }
        }
        launch_angle_full<T>(cur_[which], P_->factors.template as<T>(), P_->su.template as<T>(),
                             P_->sv.template as<T>(), grid.Ng, P_->nyz, grid.Nmu, grid.Nom, s,
                             cur_top(which / 2));
        mark(kAngle, array_bytes() * 2, s);
    }

    // Secondary source Q from the primary state (Fp, fp) on stream s;
    // `inputs_read` (optional) is recorded on s once (Fp, fp) have been read
    // (after the angular GEMM).
    // With skipping on, `n` names the step whose end-of-step primary top
    // bounds the source groups.
    void source(const T* Fp, const T* fp, T* Q, hipStream_t s, hipEvent_t inputs_read, int n) {
        const int* xt = skip_ ? x_top(n) : nullptr;
        const int* qm = skip_ ? P_->qmap.template as<int>() : nullptr;
        if constexpr (std::is_same_v<T, float>) {
            launch_source_mfma(Fp, fp, a1_.as<float>(), a2_.as<float>(), Q, P_->plan.plan,
                               grid.Ng, P_->nyz, P_->nang, s, inputs_read, xt, qm, P_->qcap);
        } else {
            const int* need = P_->need.template as<int>();
            launch_source_angular<T>(Fp, fp, P_->kv.template as<T>(), a1_.as<T>(), a2_.as<T>(),
                                     grid.Ng, P_->nyz, P_->nang, s, xt,
                                     P_->kcount.template as<int>(), P_->klist.template as<int>(),
                                     need);
            if (inputs_read) BFPN_HIP(hipEventRecord(inputs_read, s));
            launch_source_energy<T>(a1_.as<T>(), a2_.as<T>(), P_->e1.template as<double>(),
                                    P_->e2.template as<double>(), P_->beg.template as<int>(),
                                    P_->end.template as<int>(), Q, grid.Ng, P_->lanes, d.dg, s,
                                    xt, qm, need);
        }
        mark(kSource, array_bytes() * 7, s);  // 2 in, 2 out; 2 in, 1 out
    }

    // One lite-memory depth step (src/bfpm_solver.cu:2191-2212): the upwind
    // energy step over dt on psi^1 (stepLitePrimary, 3396-3416) and the IDD of
    // psi^1 alone (computeIntegratedDepthDoseLite, 3528-3549); the proxy
    // history stays zero as in the reference.
    void lite_step() {
        const int n = step_;
        double* pp = partial_p_.as<double>() + (n & 3) * 2 * eb_;
        mark(kWait, 0.0, stream);  // opens the step's first interval
        launch_energy_upwind<T>(cur_[0], nullptr, P_->ucoef.template as<UpwindCoef>(),
                                P_->coef.template as<EnergyCoef>(), nullptr, grid.Ng, P_->lanes,
                                false, P_->yzw.template as<double>(), P_->nang, pp, false, stream,
                                sweep_tops(0, -1, -1, n));
        mark(kEnergy, array_bytes() * 2, stream);
        launch_finish_sum(pp, eb_, nullptr, 0, d.du * d.dv * d.dy * d.dz / P_->density,
                          hist_.as<double>() + n, stream);
        mark(kDose, 0.0, stream);
        ++step_;
    }

    // One depth step (BFPnSolver::solve loop body, src/bfpm_solver.cu:2213-2242)
    // of an advance() call (`first`, `last`: its first and last step).
    //
    // Energy half steps: inside a call, the end half of step n and the start
    // half of step n+1 run as one paired sweep, so a call of k steps makes
    // k + 1 energy sweeps per species instead of 2k; the call's first step
    // starts and its last step ends with a single half step, so between calls
    // the state is the reference's state after the last step.
    //
    // Streams: the primary on `stream` (sp); the source, the secondary and the
    // IDD reduction on stream2_ (ss).  The primary state at the end of step n
    // is what the source of step n reads: the paired primary sweep also
    // writes it to x_ (inside a call), or it is the state itself (end of a
    // call).  The primary waits for that source's angular GEMM before it next
    // overwrites what the GEMM reads: x_ at its next paired sweep, or the state
    // at the start of the next call.  The secondary's end half of step n needs
    // the source of step n+1, so it runs, paired with the start half of n+1,
    // after that source; the IDD of step n is reduced then.  Sources alternate
    // between q_[0] and q_[1]; the primary dose partials rotate over four slots
    // (the reduction of step n precedes, on ss, the source of step n+2, which
    // the primary waits for before it writes slot n again at step n+4).
    void step(bool first, bool last) {
        const bool preserve_primary = phys.no_spatial_clipping != 0;
        const int eb = eb_;
        const int n = step_;
        double* pp = partial_p_.as<double>() + (n & 3) * 2 * eb;
        double* ps = partial_s_.as<double>();
        const bool two = stream2_ && overlap;
        hipStream_t sp = stream, ss = two ? stream2_ : stream;
        const bool secondary = P_->secondary;
        T* X[2] = {secondary ? x_[0].as<T>() : nullptr, secondary ? x_[1].as<T>() : nullptr};
        auto wait_source = [&] {
            if (src_pending_) {
                BFPN_HIP(hipStreamWaitEvent(sp, ev_src_, 0));
                src_pending_ = false;
            }
        };
        // ---- primary: stepPrimary / stepPrimaryEnergyOnly (src/bfpm_solver.cu:3238-3333)
        if (first) {
            wait_source();
            mark(kWait, 0.0, sp);  // opens the first interval
            energy(0, nullptr, nullptr, false, sp, sweep_tops(0, -1, -1, -1));
        }
        if (!phys.energy_only) {
            if (!phys.no_transport) {
                transport(0, preserve_primary, sp);
                transport(1, true, sp);
            }
            if (!phys.no_angle) {
                angle(0, sp);
                angle(1, sp);
            }
            if (!phys.no_transport) {
                transport(0, preserve_primary, sp);
                transport(1, true, sp);
            }
        }
        if (last) {
            energy(0, nullptr, pp, true, sp, sweep_tops(0, -1, -1, n));
        } else {
            wait_source();  // x_ is about to be rewritten
            if (!first) mark(kWait, 0.0, sp);
            energy_pair(0, nullptr, nullptr, pp, true, secondary ? X : nullptr, sp,
                        sweep_tops(0, -1, -1, n));
        }
        const T* Fp = last ? cur_[0] : X[0];  // primary state at the end of step n
        const T* fp = last ? cur_[1] : X[1];
        if (two) {
            BFPN_HIP(hipEventRecord(ev_prim_, sp));
            BFPN_HIP(hipStreamWaitEvent(ss, ev_prim_, 0));
            mark(kWait, 0.0, ss);
        }
        // ---- secondary: stepSecondary (src/bfpm_solver.cu:3335-3394), step n > 0
        const double scale = d.du * d.dv * d.dy * d.dz / P_->density;
        double* hist = hist_.as<double>();
        auto finish = [&](int m, bool with_secondary) {  // IDD and proxy of step m
            double* pm = partial_p_.as<double>() + (m & 3) * 2 * eb;
            launch_finish_sum(pm, eb, with_secondary ? ps : nullptr, eb, scale, hist + m, ss);
            launch_finish_sum(pm + eb, eb, nullptr, 0, d.dg, hist + hist_cap_ + m, ss);
            mark(kDose, 0.0, ss);
        };
        const bool sec = secondary && n > 0;
        if (sec) {
            T* Q = q_[n & 1].as<T>();
            source(Fp, fp, Q, ss, two ? ev_src_ : nullptr, n);
            src_pending_ = two;
            if (sec_open_) {  // end half of step n-1 with the start half of step n
                energy_pair(1, q_[(n - 1) & 1].as<T>(), Q, ps, false, nullptr, ss,
                            sweep_tops(1, n - 1, n, -1));
                finish(n - 1, true);
            } else {
                energy(1, Q, nullptr, false, ss, sweep_tops(1, n, -1, -1));
            }
            transport(2, false, ss);  // secondary psi^1 is always clipped
            transport(3, true, ss);
            angle(2, ss);
            angle(3, ss);
            transport(2, false, ss);
            transport(3, true, ss);
            if (last) {
                energy(1, Q, ps, false, ss, sweep_tops(1, n, -1, -1));
                finish(n, true);
                sec_open_ = false;
            } else {
                sec_open_ = true;  // its end half waits for the source of step n+1
            }
        } else {
            finish(n, false);  // no secondary in step n: its IDD is the primary's
        }
        ++step_;
    }

    // Slab of depth step n under a material map (the one holding its midpoint).
    int slab_of(int n) const {
        const double x = (n + 0.5) * d.dt;
        for (size_t k = 0; k + 1 < slab_end_.size(); ++k)
            if (x < slab_end_[k]) return static_cast<int>(k);
        return static_cast<int>(slab_end_.size()) - 1;
    }

    // One depth step under a material map: step() with every node-local
    // operator run per material, on the same two streams and events (each
    // stream packs into its own scratch).  The slab's largest material runs
    // in place over the whole state, the others on their packed nodes (end-
    // of-step copies and sources too: packed at moff[k], the largest's full
    // size); the dose partials of material k sit at poff[k] of a slot
    // ([dose: eb][proxy: eb]), summed per material and then over the
    // materials in order.
    void step_mapped(bool first, bool last) {
        const int n = step_, j = slab_of(n);
        const MapSlab& sl = slabs_[j];
        const int nk = static_cast<int>(sl.mat.size());
        const int Ng = grid.Ng;
        // streams as in step(): the primary on `stream`, the source and the
        // secondary on stream2_; each species packs into its own scratch
        const bool two = stream2_ && overlap;
        hipStream_t sp = stream, ss = two ? stream2_ : stream;
        hipStream_t st[2] = {sp, ss};
        T* scratch[2][2] = {{cbuf_[0].as<T>(), cbuf_[1].as<T>()},
                            {cbuf2_[0].as<T>(), cbuf2_[1].as<T>()}};
        const bool secondary = P_->secondary;
        const bool preserve_primary = phys.no_spatial_clipping != 0;
        const bool strag = phys.eq15_straggling != 0;
        // The material with the most nodes (the only one, in a slab of one
        // material) runs in place over the whole state; before it does, the
        // other materials' nodes are copied out packed and set to zero, so
        // its sweeps give them nothing (no dose, no proxy, no source, no top),
        // and afterwards their own packed results are written back.
        const int km = sl.major;
        auto use = [&](int k) { P_ = mats_[sl.mat[k]]; };
        auto minor = [&](int k) { return k != km; };
        auto wait_source = [&] {
            if (src_pending_) {
                BFPN_HIP(hipStreamWaitEvent(sp, ev_src_, 0));
                src_pending_ = false;
            }
        };
        auto nodes = [&](int k) { return map_nodes_.as<int>() + static_cast<long long>(j) * P_->nyz + sl.first[k]; };
        auto nodes_of = [&](int k) { return static_cast<long long>(minor(k) ? sl.count[k] : P_->nyz); };
        auto lanes = [&](int k) { return nodes_of(k) * P_->nang; };
        auto weights = [&](int k) {
            return minor(k) ? map_yzw_.as<double>() + static_cast<long long>(j) * P_->nyz + sl.first[k]
                            : P_->yzw.template as<double>();
        };
        auto part = [&](double* base, int k) { return base ? base + sl.poff[k] : nullptr; };
        auto xbuf = [&](int i, int k) { return minor(k) ? xm_[i].as<T>() + sl.moff[k] : x_[i].as<T>(); };
        auto qbuf = [&](int slot, int k) {
            return minor(k) ? qm_[slot].as<T>() + sl.moff[k] : q_[slot].as<T>();
        };
        // operand i of material k on stream set sc: its packed copy, or the state itself
        auto arr = [&](int sc, int i, int k, int w) {
            return minor(k) ? scratch[sc][i] + sl.moff[k] : cur_[w];
        };
        auto split = [&](int sc, const int* ws, int nw) {  // copy out, then zero, the minority nodes
            for (int k = 0; k < nk; ++k) {
                if (!minor(k)) continue;
                use(k);
                for (int i = 0; i < nw; ++i)
                    launch_gather_nodes<T>(arr(sc, i, k, ws[i]), cur_[ws[i]], nodes(k), sl.count[k],
                                           P_->nyz, P_->nang, Ng, st[sc]);
                for (int i = 0; i < nw; ++i)
                    launch_zero_nodes<T>(cur_[ws[i]], nodes(k), sl.count[k], P_->nyz, P_->nang, Ng, st[sc]);
            }
        };
        auto merge = [&](int sc, const int* ws, int nw) {
            for (int k = 0; k < nk; ++k) {
                if (!minor(k)) continue;
                use(k);
                for (int i = 0; i < nw; ++i)
                    launch_scatter_nodes<T>(cur_[ws[i]], arr(sc, i, k, ws[i]), nodes(k), sl.count[k],
                                            P_->nyz, P_->nang, Ng, st[sc]);
            }
        };
        // The tops of one sweep (sweep_tops advances the species' ring once);
        // each material's launch publishes into the same slots (atomic
        // maxima), with its own source-top map and cap.
        auto own = [&](SweepTops t) {
            t.q_cap = P_->qcap;
            if (t.qmap) t.qmap = P_->qmap.template as<int>();
            return t;
        };
        // single energy half of species sc; qs: the source slot or -1
        auto half = [&](int sc, int qs, double* partials, bool proxy, const SweepTops& tops) {
            const int ws[2] = {2 * sc, 2 * sc + 1};
            split(sc, ws, 2);
            for (int k = 0; k < nk; ++k) {
                use(k);
                launch_energy_half<T>(arr(sc, 0, k, ws[0]), arr(sc, 1, k, ws[1]),
                                      P_->coef.template as<EnergyCoef>(), qs >= 0 ? qbuf(qs, k) : nullptr,
                                      Ng, lanes(k), strag, weights(k), P_->nang, part(partials, k), proxy,
                                      st[sc], own(tops));
            }
            merge(sc, ws, 2);
            mark(kEnergy, array_bytes() * (qs >= 0 ? 5 : 4), st[sc]);
        };
        // paired energy sweep; xout: the end-of-step copies (primary)
        auto pair = [&](int sc, int qa, int qb, double* partials, bool proxy, bool xout,
                        const SweepTops& tops) {
            const int ws[2] = {2 * sc, 2 * sc + 1};
            split(sc, ws, 2);
            for (int k = 0; k < nk; ++k) {
                use(k);
                launch_energy_pair<T>(arr(sc, 0, k, ws[0]), arr(sc, 1, k, ws[1]),
                                      P_->coef.template as<EnergyCoef>(), qa >= 0 ? qbuf(qa, k) : nullptr,
                                      qb >= 0 ? qbuf(qb, k) : nullptr, Ng, lanes(k), strag, weights(k),
                                      P_->nang, part(partials, k), proxy, xout ? xbuf(0, k) : nullptr,
                                      xout ? xbuf(1, k) : nullptr, st[sc], own(tops), x_rows());
            }
            merge(sc, ws, 2);
            mark(kEnergy, array_bytes() * (4 + (qa >= 0 ? 2 : 0) + (xout ? 2 : 0)), st[sc]);
        };
        auto tr = [&](int w, bool preserve) {
            hipStream_t s = st[w / 2];
            launch_transport_half<T>(cur_[w], P_->om_dev.template as<T>(), grid.Ny, grid.Nz, Ng,
                                     P_->nang, d.dy, d.dz, 0.5 * d.dt, preserve, s, cur_top(w / 2));
            mark(kTransport, array_bytes() * 2, s);
        };
        auto ang = [&](int w) {
            const int sc = w / 2;
            split(sc, &w, 1);
            for (int k = 0; k < nk; ++k) {
                use(k);
                T* a = arr(sc, 0, k, w);
                const int planes_nodes = static_cast<int>(nodes_of(k));
                bool done = false;
                if constexpr (std::is_same_v<T, float>) {
                    if (P_->angle_plan.plan.su_h) {
                        launch_angle_mfma(a, P_->factors.template as<float>(), P_->angle_plan.plan, Ng,
                                          planes_nodes, grid.Nmu, grid.Nom, st[sc], cur_top(sc));
                        done = true;
                    }
                }
                if (!done)
                    launch_angle_full<T>(a, P_->factors.template as<T>(), P_->su.template as<T>(),
                                         P_->sv.template as<T>(), Ng, planes_nodes, grid.Nmu, grid.Nom,
                                         st[sc], cur_top(sc));
            }
            merge(sc, &w, 1);
            mark(kAngle, array_bytes() * 2, st[sc]);
        };
        // the source of step n for material k from (Fp, fp) into Q, on ss
        auto src = [&](int k, const T* Fp, const T* fp, T* Q) {
            const int* xt = skip_ ? x_top(n) : nullptr;
            const int* qm = skip_ ? P_->qmap.template as<int>() : nullptr;
            const int nn = static_cast<int>(nodes_of(k));
            if constexpr (std::is_same_v<T, float>) {
                launch_source_mfma(Fp, fp, a1_.as<float>(), a2_.as<float>(), Q, P_->plan.plan, Ng, nn,
                                   P_->nang, ss, nullptr, xt, qm, P_->qcap);
            } else {
                const int* need = P_->need.template as<int>();
                launch_source_angular<T>(Fp, fp, P_->kv.template as<T>(), a1_.as<T>(), a2_.as<T>(), Ng,
                                         nn, P_->nang, ss, xt, P_->kcount.template as<int>(),
                                         P_->klist.template as<int>(), need);
                launch_source_energy<T>(a1_.as<T>(), a2_.as<T>(), P_->e1.template as<double>(),
                                        P_->e2.template as<double>(), P_->beg.template as<int>(),
                                        P_->end.template as<int>(), Q, Ng, lanes(k), d.dg, ss, xt, qm,
                                        need);
            }
        };
        double* hist = hist_.as<double>();
        double* tmp = map_tmp_.as<double>();
        const int nm = static_cast<int>(mats_.size());
        auto finish = [&](int m, bool with_secondary) {  // IDD and proxy of step m (same slab)
            hipStream_t s = ss;
            double* pm = partial_p_.as<double>() + static_cast<size_t>(m & 3) * pslot_;
            double* ps = partial_s_.as<double>();
            for (int k = 0; k < nk; ++k) {
                const double scale = d.du * d.dv * d.dy * d.dz / mats_[sl.mat[k]]->density;
                launch_finish_sum(pm + sl.poff[k], sl.eb[k], with_secondary ? ps + sl.poff[k] : nullptr,
                                  sl.eb[k], scale, tmp + k, s);
                launch_finish_sum(pm + sl.poff[k] + sl.eb[k], sl.eb[k], nullptr, 0, d.dg, tmp + nm + k, s);
            }
            launch_sum_in_order(tmp, nk, hist + m, s);
            launch_sum_in_order(tmp + nm, nk, hist + hist_cap_ + m, s);
            mark(kDose, 0.0, s);
        };
        double* pp = partial_p_.as<double>() + static_cast<size_t>(n & 3) * pslot_;
        if (first) {
            wait_source();
            mark(kWait, 0.0, sp);
            half(0, -1, nullptr, false, sweep_tops(0, -1, -1, -1));
        }
        if (!phys.energy_only) {
            if (!phys.no_transport) {
                tr(0, preserve_primary);
                tr(1, true);
            }
            if (!phys.no_angle) {
                ang(0);
                ang(1);
            }
            if (!phys.no_transport) {
                tr(0, preserve_primary);
                tr(1, true);
            }
        }
        if (last) {
            half(0, -1, pp, true, sweep_tops(0, -1, -1, n));
        } else {
            wait_source();  // the end-of-step copies are about to be rewritten
            if (!first) mark(kWait, 0.0, sp);
            pair(0, -1, -1, pp, true, secondary, sweep_tops(0, -1, -1, n));
        }
        if (two) {
            BFPN_HIP(hipEventRecord(ev_prim_, sp));
            BFPN_HIP(hipStreamWaitEvent(ss, ev_prim_, 0));
            mark(kWait, 0.0, ss);
        }
        const bool sec = secondary && n > 0;
        if (sec) {
            const int qs = n & 1;
            for (int k = 0; k < nk; ++k) {  // the source of step n, on ss
                use(k);
                const T* a = xbuf(0, k);
                const T* b = xbuf(1, k);
                if (last) {  // from the state; otherwise from the end-of-step copy
                    a = arr(1, 0, k, 0);
                    b = arr(1, 1, k, 1);
                    if (minor(k)) {
                        launch_gather_nodes<T>(arr(1, 0, k, 0), cur_[0], nodes(k), sl.count[k], P_->nyz,
                                               P_->nang, Ng, ss);
                        launch_gather_nodes<T>(arr(1, 1, k, 1), cur_[1], nodes(k), sl.count[k], P_->nyz,
                                               P_->nang, Ng, ss);
                    }
                }
                src(k, a, b, qbuf(qs, k));
            }
            for (int k = 0; k < nk; ++k)  // the in-place source holds nothing at the other nodes
                if (minor(k)) {
                    use(k);
                    launch_zero_nodes<T>(q_[qs].as<T>(), nodes(k), sl.count[k], P_->nyz, P_->nang, Ng, ss);
                }
            if (two) BFPN_HIP(hipEventRecord(ev_src_, ss));  // its inputs are read
            src_pending_ = two;
            mark(kSource, array_bytes() * 7, ss);
            if (sec_open_) {
                pair(1, (n - 1) & 1, qs, partial_s_.as<double>(), false, false,
                     sweep_tops(1, n - 1, n, -1));
                finish(n - 1, true);
            } else {
                half(1, qs, nullptr, false, sweep_tops(1, n, -1, -1));
            }
            tr(2, false);
            tr(3, true);
            ang(2);
            ang(3);
            tr(2, false);
            tr(3, true);
            if (last) {
                half(1, qs, partial_s_.as<double>(), false, sweep_tops(1, n, -1, -1));
                finish(n, true);
                sec_open_ = false;
            } else {
                sec_open_ = true;
            }
        } else {
            finish(n, false);
        }
        ++step_;
    }

    struct MapSlab {                     // one depth slab of a material map
        std::vector<int> mat, count, first;  // per material present: index, nodes, first list entry
        int major = 0;                   // the material run in place over every node
        std::vector<long long> moff;     // the others: element offset of their packed arrays
        std::vector<int> eb, poff;       // dose blocks of its sweeps, offset of its partials
    };

    std::shared_ptr<Physics<T>> P_;  // tables of the current (or last) step
    std::vector<std::shared_ptr<Physics<T>>> layers_;  // empty: one material
    std::vector<double> x_end_;
    std::vector<std::shared_ptr<Physics<T>>> mats_;  // material map (empty: none)
    std::vector<double> slab_end_;
    std::vector<MapSlab> slabs_;
    DeviceBuffer map_nodes_, map_yzw_, map_tmp_, cbuf_[2], cbuf2_[2];  // node lists and weights, slab-major
    DeviceBuffer xm_[2], qm_[2];  // packed end-of-step copies and sources of the minority materials
    int pslot_ = 0;                                           // doubles per partial slot
    DeviceBuffer state_[4], a1_, a2_, q_[2], x_[2];
    DeviceBuffer diag_part_, diag_out_;  // diagnostics scratch, allocated on first use
    DeviceBuffer partial_p_, partial_s_, hist_, tops_;
    T* cur_[4] = {nullptr, nullptr, nullptr, nullptr};
    int step_ = 0;
    int p_cur_ = 0, s_cur_ = 0;  // current slots of the primary / secondary top rings
    int eb_ = 0;                 // blocks (dose partials) of a closing energy sweep
    bool skip_ = true;           // active-group skipping
    int hist_cap_ = 0;
    std::vector<std::pair<hipEvent_t, Phase>> events_[2];  // per stream (profiling)
    hipStream_t stream2_ = nullptr;  // source + secondary; null without a secondary
    hipEvent_t ev_prim_ = nullptr, ev_src_ = nullptr, ev_tail_ = nullptr;
    bool src_pending_ = false;  // the primary must wait for ev_src_ before its next write
    bool sec_open_ = false;     // the secondary's last step still lacks its end half
};

}  // namespace

std::shared_ptr<PhysicsBase> make_physics(const bfpn_grid& g, const bfpn_phys& p,
                                          const std::string& data_path, int precision) {
    validate_params(g, p);
    if (precision == BFPN_FP32) return std::make_shared<Physics<float>>(g, p, data_path, precision);
    if (precision == BFPN_FP64) return std::make_shared<Physics<double>>(g, p, data_path, precision);
    throw Error(BFPN_E_INVALID, "precision must be BFPN_FP32 or BFPN_FP64");
}

bfpn_ctx* make_context(std::shared_ptr<PhysicsBase> physics, const bfpn_grid& g,
                       const bfpn_phys& p) {
    validate_params(g, p);
    if (physics->precision == BFPN_FP32)
        return new Context<float>(std::static_pointer_cast<Physics<float>>(physics), g, p);
    return new Context<double>(std::static_pointer_cast<Physics<double>>(physics), g, p);
}

void validate_layers(const std::vector<LayerSpec>& layers) {
    BFPN_REQUIRE(!layers.empty(), "at least one layer is needed");
    for (size_t k = 0; k < layers.size(); ++k) {
        const LayerSpec& l = layers[k];
        BFPN_REQUIRE(l.density > 0, "layer density must be positive");
        BFPN_REQUIRE(l.material.size() < sizeof(bfpn_phys{}.material_name),
                     "material name too long");
        if (k + 1 == layers.size()) break;  // the last layer's depth is not used
        BFPN_REQUIRE(l.x_end > 0, "layer depths must be positive");
        BFPN_REQUIRE(k == 0 || l.x_end > layers[k - 1].x_end, "layer depths must increase");
    }
}

bfpn_phys layer_phys(const bfpn_phys& p, const LayerSpec& l) {
    bfpn_phys q = p;
    std::snprintf(q.material_name, sizeof(q.material_name), "%s", l.material.c_str());
    q.density = q.energy_density = l.density;
    return q;
}

bfpn_ctx* make_layered_context(const bfpn_grid& g, const bfpn_phys& p,
                               const std::vector<LayerSpec>& layers, int precision) {
    validate_layers(layers);
    std::vector<std::shared_ptr<PhysicsBase>> tables;
    for (const LayerSpec& l : layers)
        tables.push_back(make_physics(g, layer_phys(p, l), l.data_path, precision));
    return make_layered_context(g, p, layers, tables);
}

bfpn_ctx* make_layered_context(const bfpn_grid& g, const bfpn_phys& p,
                               const std::vector<LayerSpec>& layers,
                               const std::vector<std::shared_ptr<PhysicsBase>>& tables) {
    validate_layers(layers);
    BFPN_REQUIRE(tables.size() == layers.size() && !tables.empty(), "one table set per layer");
    const int precision = tables.front()->precision;
    std::vector<double> x_end;
    for (size_t k = 0; k < layers.size(); ++k) {
        BFPN_REQUIRE(tables[k]->precision == precision, "layer tables of one precision");
        x_end.push_back(layers[k].x_end);
    }
    std::unique_ptr<bfpn_ctx> ctx(make_context(tables.front(), g, p));
    if (precision == BFPN_FP32) {
        std::vector<std::shared_ptr<Physics<float>>> t;
        for (auto& b : tables) t.push_back(std::static_pointer_cast<Physics<float>>(b));
        static_cast<Context<float>*>(ctx.get())->set_layers(std::move(t), std::move(x_end));
    } else {
        std::vector<std::shared_ptr<Physics<double>>> t;
        for (auto& b : tables) t.push_back(std::static_pointer_cast<Physics<double>>(b));
        static_cast<Context<double>*>(ctx.get())->set_layers(std::move(t), std::move(x_end));
    }
    return ctx.release();
}

void validate_map(const std::vector<LayerSpec>& materials, const std::vector<double>& slab_end,
                  const std::vector<unsigned char>& map, long long nyz) {
    BFPN_REQUIRE(!materials.empty() && materials.size() <= 256, "a map needs 1 to 256 materials");
    for (const LayerSpec& m : materials) {
        BFPN_REQUIRE(m.density > 0, "material density must be positive");
        BFPN_REQUIRE(m.material.size() < sizeof(bfpn_phys{}.material_name), "material name too long");
    }
    BFPN_REQUIRE(!slab_end.empty(), "a map needs at least one depth slab");
    for (size_t k = 0; k + 1 < slab_end.size(); ++k) {
        BFPN_REQUIRE(slab_end[k] > 0, "slab depths must be positive");
        BFPN_REQUIRE(k == 0 || slab_end[k] > slab_end[k - 1], "slab depths must increase");
    }
    BFPN_REQUIRE(map.size() == slab_end.size() * static_cast<size_t>(nyz),
                 "the map needs (Ny+1)(Nz+1) material indices per slab");
    for (unsigned char m : map) BFPN_REQUIRE(m < materials.size(), "map entry names no material");
}

size_t map_bytes(const bfpn_grid& g, const bfpn_phys& p, const std::vector<unsigned char>& map,
                 int nmat, int precision) {
    const size_t nyz = static_cast<size_t>(g.Ny + 1) * (g.Nz + 1);
    size_t minority = 0;  // the most nodes outside the largest material, over the slabs
    for (size_t j = 0; nyz && j + 1 <= map.size() / nyz; ++j) {
        std::vector<size_t> count(static_cast<size_t>(std::max(nmat, 1)), 0);
        for (size_t s = 0; s < nyz; ++s) ++count[std::min<size_t>(map[j * nyz + s], count.size() - 1)];
        minority = std::max(minority, nyz - *std::max_element(count.begin(), count.end()));
    }
    const bool sec = !p.primary_only && !p.energy_only && !p.lite_memory;
    const size_t word = precision == BFPN_FP64 ? 8 : 4;
    return (sec ? 8 : 2) * minority * g.Ng * g.Nmu * g.Nom * word;
}

bfpn_ctx* make_mapped_context(const bfpn_grid& g, const bfpn_phys& p,
                              const std::vector<LayerSpec>& materials,
                              const std::vector<double>& slab_end,
                              const std::vector<unsigned char>& map, int precision) {
    validate_params(g, p);
    validate_map(materials, slab_end, map, static_cast<long long>(g.Ny + 1) * (g.Nz + 1));
    BFPN_REQUIRE(!p.legacy_energy && !p.lite_memory,
                 "material maps support the eq15 energy model with the full state only");
    std::vector<std::shared_ptr<PhysicsBase>> tables;
    for (const LayerSpec& m : materials)
        tables.push_back(make_physics(g, layer_phys(p, m), m.data_path, precision));
    return make_mapped_context(g, p, materials, slab_end, map, tables);
}

bfpn_ctx* make_mapped_context(const bfpn_grid& g, const bfpn_phys& p,
                              const std::vector<LayerSpec>& materials,
                              const std::vector<double>& slab_end,
                              const std::vector<unsigned char>& map,
                              const std::vector<std::shared_ptr<PhysicsBase>>& tables) {
    validate_params(g, p);
    validate_map(materials, slab_end, map, static_cast<long long>(g.Ny + 1) * (g.Nz + 1));
    BFPN_REQUIRE(!p.legacy_energy && !p.lite_memory,
                 "material maps support the eq15 energy model with the full state only");
    BFPN_REQUIRE(tables.size() == materials.size(), "one table set per material");
    const int precision = tables.front()->precision;
    for (const auto& t : tables) BFPN_REQUIRE(t->precision == precision, "map tables of one precision");
    std::unique_ptr<bfpn_ctx> ctx(make_context(tables.front(), g, p));
    if (precision == BFPN_FP32) {
        std::vector<std::shared_ptr<Physics<float>>> t;
        for (auto& b : tables) t.push_back(std::static_pointer_cast<Physics<float>>(b));
        static_cast<Context<float>*>(ctx.get())->set_map(std::move(t), slab_end, map);
    } else {
        std::vector<std::shared_ptr<Physics<double>>> t;
        for (auto& b : tables) t.push_back(std::static_pointer_cast<Physics<double>>(b));
        static_cast<Context<double>*>(ctx.get())->set_map(std::move(t), slab_end, map);
    }
    return ctx.release();
}

size_t context_bytes(const bfpn_grid& g, const bfpn_phys& p, int precision) {
    const size_t lanes = static_cast<size_t>(g.Ny + 1) * (g.Nz + 1) * g.Nmu * g.Nom;
    const size_t cells = lanes * g.Ng;
    const size_t word = precision == BFPN_FP64 ? 8 : 4;
    const bool sec = !p.primary_only && !p.energy_only && !p.lite_memory;
    return cells * word * (p.lite_memory ? 1 : sec ? 10 : 4) + 9 * ((lanes + 255) / 256) * 8 +
           2 * static_cast<size_t>(std::max(16, g.Nx)) * 8;
}

size_t physics_bytes(const bfpn_grid& g, const bfpn_phys& p, int precision) {
    const size_t nang = static_cast<size_t>(g.Nmu) * g.Nom;
    const bool sec = !p.primary_only && !p.energy_only && !p.lite_memory;
    if (!sec) return static_cast<size_t>(g.Ng) * nang * 8 + (1 << 20);
    if (precision == BFPN_FP64) return static_cast<size_t>(g.Ng) * nang * nang * 8 + (1 << 22);
    const size_t nap = source_nap(static_cast<int>(nang));
    return static_cast<size_t>(g.Ng) * nap * nap * 4 +                           // split Kv
           4 * static_cast<size_t>(source_erows(g.Ng)) * source_ngp(g.Ng) * 2 +  // split E
           static_cast<size_t>(g.Ng) * nang * nang * 4 + (1 << 22);              // Kv during set-up
}

}  // namespace bfpn

using namespace bfpn;

// ============================================================ workspace ===
// Scratch and cached operands of the operator-level API.  Per call, the
// coefficient tables are rebuilt on the device from the caller's device
// tables (stream-ordered, no host round trip).  Two operand sets are built on
// the host and kept: the secondary source's operands (keyed by the kernel
// table pointers and sizes; bfpn_secondary_source_prepare_* builds them) and
// the CN factors of the angular step (keyed by the values of sig_tr and the
// scalars).  A workspace orders its scratch on the stream of each call.
struct bfpn_workspace {
    struct SourceKey {
        const void *e1 = nullptr, *e2 = nullptr, *kv = nullptr, *beg = nullptr, *end = nullptr;
        int Ng = 0, nang = 0, precision = -1;
        double dg = 0.0;
        bool operator==(const SourceKey& o) const {
            return e1 == o.e1 && e2 == o.e2 && kv == o.kv && beg == o.beg && end == o.end &&
                   Ng == o.Ng && nang == o.nang && precision == o.precision && dg == o.dg;
        }
    };
    DeviceBuffer coef, partial, a1, a2, factors, su, sv, om, dose_w, yzw, kvT;
    SourcePlanBuffers plan;
    AnglePlanBuffers angle_plan;
    SourceKey source_key;
    bool source_ready = false;
    std::vector<double> angle_key;  // precision, grid scalars and sig_tr of the cached factors
    int builds = 0;                 // host-side builds of cached operands

    // Grows `b` to at least `bytes`.  Earlier calls on `s` may still read the
    // old buffer, so it is freed only after they have finished.
    static void grow(DeviceBuffer& b, size_t bytes, hipStream_t s) {
        if (bytes <= b.bytes()) return;
        BFPN_HIP(hipStreamSynchronize(s));
        b.ensure(bytes);
    }
    size_t device_bytes() const {
        size_t n = plan.bytes() + angle_plan.h.bytes() + angle_plan.l.bytes();
        for (const DeviceBuffer* b : {&coef, &partial, &a1, &a2, &factors, &su, &sv, &om, &dose_w,
                                      &yzw, &kvT})
            n += b->bytes();
        return n;
    }
};

namespace {

template <typename T>
void op_energy(bfpn_workspace* ws, T* F, T* f, const double* S, const double* sigma,
               const double* Tc, double h, double dg, double rho, int Ng, int nyz, int nang,
               const T* Q1, const T* Q2, int straggling, hipStream_t s) {
    BFPN_REQUIRE(ws && F && f && S && sigma, "null argument");
    BFPN_REQUIRE(Ng > 0 && nyz > 0 && nang > 0 && h > 0, "invalid sizes");
    if (Q2) throw Error(BFPN_E_UNSUPPORTED, "a psi^2 source (Q2) is not supported");
    const bool st = straggling != 0 && Tc != nullptr;
    bfpn_workspace::grow(ws->coef, static_cast<size_t>(Ng) * sizeof(EnergyCoef), s);
    launch_energy_coefs(S, sigma, Tc, st, h, dg, rho, Ng, ws->coef.as<EnergyCoef>(), s);
    launch_energy_half<T>(F, f, ws->coef.as<EnergyCoef>(), Q1, Ng,
                          static_cast<long long>(nyz) * nang, st, nullptr, nang, nullptr, false, s);
}

// Builds the source operands of (e1, e2, kv, beg, end) into the workspace;
// synchronizes the stream (host copies of the tables are needed).
template <typename T>
void prepare_source(bfpn_workspace* ws, const double* e1, const double* e2, const double* kv,
                    const int* beg, const int* end, int Ng, int nang, double dg, hipStream_t s) {
    BFPN_REQUIRE(ws && e1 && e2 && kv && beg && end, "null argument");
    BFPN_REQUIRE(Ng > 0 && nang > 0, "invalid sizes");
    bfpn_workspace::SourceKey key;
    key.e1 = e1;
    key.e2 = e2;
    key.kv = kv;
    key.beg = beg;
    key.end = end;
    key.Ng = Ng;
    key.nang = nang;
    key.dg = dg;
    key.precision = std::is_same_v<T, float> ? BFPN_FP32 : BFPN_FP64;
    BFPN_HIP(hipStreamSynchronize(s));
    ws->source_ready = false;
    if constexpr (std::is_same_v<T, float>) {
        const long long nk = static_cast<long long>(Ng) * nang * nang;
        ws->kvT.ensure(nk * sizeof(float));
        launch_convert<float>(kv, ws->kvT.as<float>(), nk, s);
        std::vector<double> h1(static_cast<size_t>(Ng) * Ng), h2(h1.size());
        std::vector<int> hb(Ng), he(Ng);
        BFPN_HIP(hipMemcpy(h1.data(), e1, h1.size() * sizeof(double), hipMemcpyDeviceToHost));
        BFPN_HIP(hipMemcpy(h2.data(), e2, h2.size() * sizeof(double), hipMemcpyDeviceToHost));
        BFPN_HIP(hipMemcpy(hb.data(), beg, Ng * sizeof(int), hipMemcpyDeviceToHost));
        BFPN_HIP(hipMemcpy(he.data(), end, Ng * sizeof(int), hipMemcpyDeviceToHost));
        BFPN_HIP(hipStreamSynchronize(s));
        ws->plan.build(ws->kvT.as<float>(), h1, h2, hb, he, Ng, nang, dg, s);
        ws->kvT.release();
    }
    ws->source_key = key;
    ws->source_ready = true;
    ++ws->builds;
}

template <typename T>
void op_source(bfpn_workspace* ws, const T* Fp, const T* fp, T* Q1, const double* e1,
               const double* e2, const double* kv, const int* beg, const int* end, int Ng, int nyz,
               int nang, double dg, hipStream_t s) {
    BFPN_REQUIRE(ws && Fp && fp && Q1 && e1 && e2 && kv && beg && end, "null argument");
    BFPN_REQUIRE(Ng > 0 && nyz > 0 && nang > 0, "invalid sizes");
    bfpn_workspace::SourceKey key;
    key.e1 = e1;
    key.e2 = e2;
    key.kv = kv;
    key.beg = beg;
    key.end = end;
    key.Ng = Ng;
    key.nang = nang;
    key.dg = dg;
    key.precision = std::is_same_v<T, float> ? BFPN_FP32 : BFPN_FP64;
    if (!ws->source_ready || !(ws->source_key == key))
        prepare_source<T>(ws, e1, e2, kv, beg, end, Ng, nang, dg, s);
    const size_t n = static_cast<size_t>(Ng) * nyz * nang;
    bfpn_workspace::grow(ws->a1, n * sizeof(T), s);
    bfpn_workspace::grow(ws->a2, n * sizeof(T), s);
    if constexpr (std::is_same_v<T, float>) {
        launch_source_mfma(Fp, fp, ws->a1.as<float>(), ws->a2.as<float>(), Q1, ws->plan.plan, Ng,
                           nyz, nang, s);
    } else {
        launch_source_angular<T>(Fp, fp, kv, ws->a1.as<T>(), ws->a2.as<T>(), Ng, nyz, nang, s);
        launch_source_energy<T>(ws->a1.as<T>(), ws->a2.as<T>(), e1, e2, beg, end, Q1, Ng,
                                static_cast<long long>(nyz) * nang, dg, s);
    }
}

template <typename T>
void op_transport(bfpn_workspace* ws, T* F, const double* omend, int Ny, int Nz, int Ng, int nang,
                  double dy, double dz, double h, int preserve, hipStream_t s) {
    BFPN_REQUIRE(ws && F && omend, "null argument");
    BFPN_REQUIRE(Ny > 0 && Nz > 0 && Ng > 0 && nang > 0, "invalid sizes");
    bfpn_workspace::grow(ws->om, 2 * nang * sizeof(T), s);
    launch_convert<T>(omend, ws->om.as<T>(), 2 * nang, s);
    launch_transport_half<T>(F, ws->om.as<T>(), Ny, Nz, Ng, nang, dy, dz, h, preserve != 0, s);
}

template <typename T>
void op_angle(bfpn_workspace* ws, T* F, const double* sig_tr_host, int nyz, int Ng, int Nmu,
              int Nom, double du, double dv, double dt, double rho, hipStream_t s) {
    BFPN_REQUIRE(ws && F && sig_tr_host, "null argument");
    BFPN_REQUIRE(nyz > 0 && Ng > 0 && Nmu > 0 && Nom > 0 && Nmu * Nom <= 1024, "invalid sizes");
    const bool mfma = std::is_same_v<T, float> &&
                      AnglePlanBuffers::usable(Nmu, Nom, static_cast<long long>(Ng) * nyz);
    std::vector<double> key = {std::is_same_v<T, float> ? 0.0 : 1.0, double(Ng), double(Nmu),
                               double(Nom), du, dv, dt, rho};
    key.insert(key.end(), sig_tr_host, sig_tr_host + Ng);
    if (key != ws->angle_key) {  // new tables: the buffers are rewritten after earlier calls
        BFPN_HIP(hipStreamSynchronize(s));
        ws->angle_key.clear();
        upload(ws->factors, cast_vector<T>(angle_factors(sig_tr_host, Ng, Nmu, Nom, du, dv, dt, rho)));
        if (mfma) {
            ws->angle_plan.build(Nmu, Nom, s);
        } else {
            upload(ws->su, cast_vector<T>(sine_basis(Nom)));
            upload(ws->sv, cast_vector<T>(sine_basis(Nmu)));
        }
        ws->angle_key.swap(key);
        ++ws->builds;
    }
    if constexpr (std::is_same_v<T, float>) {
        if (mfma) {
            launch_angle_mfma(F, ws->factors.as<float>(), ws->angle_plan.plan, Ng, nyz, Nmu, Nom, s);
            return;
        }
    }
    launch_angle_full<T>(F, ws->factors.as<T>(), ws->su.as<T>(), ws->sv.as<T>(), Ng, nyz, Nmu, Nom, s);
}

template <typename T>
void op_idd(bfpn_workspace* ws, const T* F, const T* f, const T* F1, const T* f1, const double* S,
            const double* sigma, const double* en, int Ng, int Ny, int Nz, int nang, double dg,
            double du, double dv, double dy, double dz, double rho, int trap, int cat, double* out,
            hipStream_t s) {
    BFPN_REQUIRE(ws && F && S && sigma && en && out, "null argument");
    BFPN_REQUIRE(Ng > 0 && Ny > 0 && Nz > 0 && nang > 0 && rho > 0, "invalid sizes");
    const long long lanes = static_cast<long long>(Ny + 1) * (Nz + 1) * nang;
    const int nb = dose_blocks(Ng, lanes);
    bfpn_workspace::grow(ws->dose_w, static_cast<size_t>(Ng) * sizeof(DoseCoef), s);
    bfpn_workspace::grow(ws->yzw, static_cast<size_t>(Ny + 1) * (Nz + 1) * sizeof(double), s);
    bfpn_workspace::grow(ws->partial, (nb + 1) * sizeof(double), s);
    launch_dose_coefs(S, sigma, en, dg, cat != 0, Ng, ws->dose_w.as<DoseCoef>(), s);
    launch_yz_weights(Ny, Nz, trap != 0, ws->yzw.as<double>(), s);
    double* part = ws->partial.as<double>();
    launch_dose<T>(F, f, F1, f1, ws->dose_w.as<DoseCoef>(), ws->yzw.as<double>(), Ng, lanes, nang, part, s);
    launch_finish_sum(part, nb, nullptr, 0, du * dv * dy * dz / rho, part + nb, s);
    BFPN_HIP(hipMemcpyAsync(out, part + nb, sizeof(double), hipMemcpyDeviceToHost, s));
    BFPN_HIP(hipStreamSynchronize(s));
}

hipStream_t as_stream(void* s) { return static_cast<hipStream_t>(s); }

}  // namespace

// ---------------------------------------------------- allocation-free IDD ---
// bfpn_idd_f32_noalloc keeps the scratch of the IDD reduction (per-group dose
// coefficients, y/z weights, per-block partial sums and the final sum) in
// this file-scope device array instead of a workspace.  Each GPU holds its
// own copy until the process ends; every call on a GPU shares it, so calls
// that overlap on one GPU race.  Every byte a call reads from it is written
// earlier in the same call (no clearing is needed).
constexpr size_t kNoAllocScratchBytes = size_t(2) << 20;
__device__ __align__(256) unsigned char scratch[kNoAllocScratchBytes];

namespace {

std::mutex g_scratch_mu;
std::vector<unsigned char*> g_scratch_ptr;  // device address of `scratch`, per device

// The device address of `scratch` on the current device, looked up once per
// device (with that device current).  The array's host-side name is not a
// device address and is never passed to a kernel or copy.
unsigned char* noalloc_scratch() {
    int dev = 0;
    BFPN_HIP(hipGetDevice(&dev));
    std::lock_guard<std::mutex> lk(g_scratch_mu);
    if (g_scratch_ptr.size() <= static_cast<size_t>(dev)) g_scratch_ptr.resize(dev + 1, nullptr);
    if (!g_scratch_ptr[dev]) {
        void* p = nullptr;
        BFPN_HIP(hipGetSymbolAddress(&p, HIP_SYMBOL(scratch)));
        BFPN_REQUIRE(p != nullptr, "no device address for the IDD scratch array");
        g_scratch_ptr[dev] = static_cast<unsigned char*>(p);
    }
    return g_scratch_ptr[dev];
}

size_t round256(size_t b) { return (b + 255) & ~size_t(255); }

void op_idd_noalloc(const float* F, const float* f, const float* F1, const float* f1,
                    const double* S, const double* sigma, const double* en, int Ng, int Ny,
                    int Nz, int nang, double dg, double du, double dv, double dy, double dz,
                    double rho, int trap, int cat, double* out, hipStream_t s) {
    BFPN_REQUIRE(F && S && sigma && en && out, "null argument");
    BFPN_REQUIRE(Ng > 0 && Ny > 0 && Nz > 0 && nang > 0 && rho > 0, "invalid sizes");
    const long long nyz = static_cast<long long>(Ny + 1) * (Nz + 1);
    const long long lanes = nyz * nang;
    const int nb = dose_blocks(Ng, lanes);
    const size_t w_bytes = round256(static_cast<size_t>(Ng) * sizeof(DoseCoef));
    const size_t yz_bytes = round256(static_cast<size_t>(nyz) * sizeof(double));
    const size_t part_bytes = (static_cast<size_t>(nb) + 1) * sizeof(double);
    const size_t need = w_bytes + yz_bytes + part_bytes;
    if (need > kNoAllocScratchBytes)
        throw Error(BFPN_E_INVALID,
                    "bfpn_idd_f32_noalloc: this grid needs " + std::to_string(need) +
                        " bytes of scratch, more than its fixed " +
                        std::to_string(kNoAllocScratchBytes) + " (use bfpn_idd_f32)");
    unsigned char* base = noalloc_scratch();
    auto* w = reinterpret_cast<DoseCoef*>(base);
    auto* yzw = reinterpret_cast<double*>(base + w_bytes);
    auto* part = reinterpret_cast<double*>(base + w_bytes + yz_bytes);
    launch_dose_coefs(S, sigma, en, dg, cat != 0, Ng, w, s);
    launch_yz_weights(Ny, Nz, trap != 0, yzw, s);
    launch_dose<float>(F, f, F1, f1, w, yzw, Ng, lanes, nang, part, s);
    launch_finish_sum(part, nb, nullptr, 0, du * dv * dy * dz / rho, part + nb, s);
    BFPN_HIP(hipMemcpyAsync(out, part + nb, sizeof(double), hipMemcpyDeviceToHost, s));
    BFPN_HIP(hipStreamSynchronize(s));
}

}  // namespace

// ================================================================ C ABI ===
extern "C" {

BFPN_API const char* bfpn_last_error(void) { return g_last_error.c_str(); }
BFPN_API const char* bfpn_version(void) { return "bfpn-mi355x 0.2 (gfx950)"; }

BFPN_API bfpn_status bfpn_device_count(int* count) {
    return guarded([&] {
        BFPN_REQUIRE(count, "null argument");
        BFPN_HIP(hipGetDeviceCount(count));
    });
}

BFPN_API bfpn_status bfpn_device_info(int device, char* name, size_t name_len, size_t* total,
                                      int* cus) {
    return guarded([&] {
        hipDeviceProp_t prop;
        BFPN_HIP(hipGetDeviceProperties(&prop, device));
        if (name && name_len) {
            std::strncpy(name, prop.name, name_len - 1);
            name[name_len - 1] = 0;
        }
        if (total) *total = prop.totalGlobalMem;
        if (cus) *cus = prop.multiProcessorCount;
    });
}

BFPN_API bfpn_status bfpn_malloc(void** p, size_t bytes) {
    return guarded([&] {
        BFPN_REQUIRE(p, "null argument");
        hipError_t e = hipMalloc(p, bytes);
        if (e != hipSuccess) {
            *p = nullptr;
            (void)hipGetLastError();
            throw Error(BFPN_E_NOMEM, std::string("hipMalloc failed: ") + hipGetErrorString(e));
        }
    });
}
BFPN_API bfpn_status bfpn_free(void* p) { return guarded([&] { BFPN_HIP(hipFree(p)); }); }
BFPN_API bfpn_status bfpn_memcpy_h2d(void* dst, const void* src, size_t bytes) {
    return guarded([&] { BFPN_HIP(hipMemcpy(dst, src, bytes, hipMemcpyHostToDevice)); });
}
BFPN_API bfpn_status bfpn_memcpy_d2h(void* dst, const void* src, size_t bytes) {
    return guarded([&] { BFPN_HIP(hipMemcpy(dst, src, bytes, hipMemcpyDeviceToHost)); });
}
BFPN_API bfpn_status bfpn_memset(void* p, int v, size_t bytes) {
    return guarded([&] {
        BFPN_HIP(hipMemset(p, v, bytes));
        BFPN_HIP(hipDeviceSynchronize());
    });
}
BFPN_API bfpn_status bfpn_device_copy(void* dst, const void* src, size_t bytes) {
    return guarded([&] {
        BFPN_REQUIRE(dst && src && bytes % 16 == 0, "device copy needs 16-byte multiples");
        launch_copy16(src, dst, bytes, nullptr);
        BFPN_HIP(hipDeviceSynchronize());
    });
}
BFPN_API bfpn_status bfpn_synchronize(void) {
    return guarded([&] { BFPN_HIP(hipDeviceSynchronize()); });
}
BFPN_API bfpn_status bfpn_mem_info(size_t* free_b, size_t* total_b) {
    return guarded([&] { BFPN_HIP(hipMemGetInfo(free_b, total_b)); });
}

BFPN_API bfpn_status bfpn_workspace_create(bfpn_workspace** ws) {
    return guarded([&] {
        BFPN_REQUIRE(ws, "null argument");
        *ws = new bfpn_workspace();
    });
}
BFPN_API void bfpn_workspace_destroy(bfpn_workspace* ws) { delete ws; }

BFPN_API bfpn_status bfpn_energy_half_f32(bfpn_workspace* ws, float* F, float* f, const double* S,
                                          const double* sigma, const double* T, double h, double dg,
                                          double rho, int Ng, int nyz, int nang, const float* Q1,
                                          const float* Q2, int straggling, void* stream) {
    return guarded([&] { op_energy<float>(ws, F, f, S, sigma, T, h, dg, rho, Ng, nyz, nang, Q1, Q2, straggling, as_stream(stream)); });
}
BFPN_API bfpn_status bfpn_energy_half_f64(bfpn_workspace* ws, double* F, double* f, const double* S,
                                          const double* sigma, const double* T, double h, double dg,
                                          double rho, int Ng, int nyz, int nang, const double* Q1,
                                          const double* Q2, int straggling, void* stream) {
    return guarded([&] { op_energy<double>(ws, F, f, S, sigma, T, h, dg, rho, Ng, nyz, nang, Q1, Q2, straggling, as_stream(stream)); });
}

BFPN_API bfpn_status bfpn_secondary_source_f32(bfpn_workspace* ws, const float* Fp, const float* fp,
                                               float* Q1, const double* e1, const double* e2,
                                               const double* kv, const int* beg, const int* end,
                                               int Ng, int nyz, int nang, double dg, void* stream) {
    return guarded([&] { op_source<float>(ws, Fp, fp, Q1, e1, e2, kv, beg, end, Ng, nyz, nang, dg, as_stream(stream)); });
}
BFPN_API bfpn_status bfpn_secondary_source_f64(bfpn_workspace* ws, const double* Fp, const double* fp,
                                               double* Q1, const double* e1, const double* e2,
                                               const double* kv, const int* beg, const int* end,
                                               int Ng, int nyz, int nang, double dg, void* stream) {
    return guarded([&] { op_source<double>(ws, Fp, fp, Q1, e1, e2, kv, beg, end, Ng, nyz, nang, dg, as_stream(stream)); });
}

BFPN_API bfpn_status bfpn_secondary_source_prepare_f32(bfpn_workspace* ws, const double* e1,
                                                       const double* e2, const double* kv,
                                                       const int* beg, const int* end, int Ng,
                                                       int nang, double dg, void* stream) {
    return guarded([&] { prepare_source<float>(ws, e1, e2, kv, beg, end, Ng, nang, dg, as_stream(stream)); });
}
BFPN_API bfpn_status bfpn_secondary_source_prepare_f64(bfpn_workspace* ws, const double* e1,
                                                       const double* e2, const double* kv,
                                                       const int* beg, const int* end, int Ng,
                                                       int nang, double dg, void* stream) {
    return guarded([&] { prepare_source<double>(ws, e1, e2, kv, beg, end, Ng, nang, dg, as_stream(stream)); });
}

BFPN_API bfpn_status bfpn_workspace_stats(const bfpn_workspace* ws, int* builds, size_t* bytes) {
    return guarded([&] {
        BFPN_REQUIRE(ws, "null workspace");
        if (builds) *builds = ws->builds;
        if (bytes) *bytes = ws->device_bytes();
    });
}

BFPN_API bfpn_status bfpn_transport_half_f32(bfpn_workspace* ws, float* F, const double* omend,
                                             int Ny, int Nz, int Ng, int nang, double dy, double dz,
                                             double h, int preserve_sign, void* stream) {
    return guarded([&] { op_transport<float>(ws, F, omend, Ny, Nz, Ng, nang, dy, dz, h, preserve_sign, as_stream(stream)); });
}
BFPN_API bfpn_status bfpn_transport_half_f64(bfpn_workspace* ws, double* F, const double* omend,
                                             int Ny, int Nz, int Ng, int nang, double dy, double dz,
                                             double h, int preserve_sign, void* stream) {
    return guarded([&] { op_transport<double>(ws, F, omend, Ny, Nz, Ng, nang, dy, dz, h, preserve_sign, as_stream(stream)); });
}

BFPN_API bfpn_status bfpn_angle_full_f32(bfpn_workspace* ws, float* F, const double* sig_tr,
                                         int nyz, int Ng, int Nmu, int Nom, double du, double dv,
                                         double dt, double rho, void* stream) {
    return guarded([&] { op_angle<float>(ws, F, sig_tr, nyz, Ng, Nmu, Nom, du, dv, dt, rho, as_stream(stream)); });
}
BFPN_API bfpn_status bfpn_angle_full_f64(bfpn_workspace* ws, double* F, const double* sig_tr,
                                         int nyz, int Ng, int Nmu, int Nom, double du, double dv,
                                         double dt, double rho, void* stream) {
    return guarded([&] { op_angle<double>(ws, F, sig_tr, nyz, Ng, Nmu, Nom, du, dv, dt, rho, as_stream(stream)); });
}

BFPN_API bfpn_status bfpn_idd_f32(bfpn_workspace* ws, const float* F, const float* f, const float* F1,
                                  const float* f1, const double* S, const double* sigma,
                                  const double* en, int Ng, int Ny, int Nz, int nang, double dg,
                                  double du, double dv, double dy, double dz, double rho, int trap,
                                  int cat, double* idd, void* stream) {
    return guarded([&] { op_idd<float>(ws, F, f, F1, f1, S, sigma, en, Ng, Ny, Nz, nang, dg, du, dv, dy, dz, rho, trap, cat, idd, as_stream(stream)); });
}
BFPN_API bfpn_status bfpn_idd_f64(bfpn_workspace* ws, const double* F, const double* f,
                                  const double* F1, const double* f1, const double* S,
                                  const double* sigma, const double* en, int Ng, int Ny, int Nz,
                                  int nang, double dg, double du, double dv, double dy, double dz,
                                  double rho, int trap, int cat, double* idd, void* stream) {
    return guarded([&] { op_idd<double>(ws, F, f, F1, f1, S, sigma, en, Ng, Ny, Nz, nang, dg, du, dv, dy, dz, rho, trap, cat, idd, as_stream(stream)); });
}
BFPN_API bfpn_status bfpn_idd_f32_noalloc(const float* F, const float* f, const float* F1,
                                          const float* f1, const double* S, const double* sigma,
                                          const double* en, int Ng, int Ny, int Nz, int nang,
                                          double dg, double du, double dv, double dy, double dz,
                                          double rho, int trap, int cat, double* idd,
                                          void* stream) {
    return guarded([&] { op_idd_noalloc(F, f, F1, f1, S, sigma, en, Ng, Ny, Nz, nang, dg, du, dv, dy, dz, rho, trap, cat, idd, as_stream(stream)); });
}

BFPN_API bfpn_status bfpn_create(const bfpn_grid* grid, const bfpn_phys* phys, const char* data_path,
                                 int precision, bfpn_ctx** out) {
    return guarded([&] {
        BFPN_REQUIRE(grid && phys && out, "null argument");
        *out = nullptr;
        validate_params(*grid, *phys);
        // Memory preflight (preflightMemory, src/bfpm_solver.cu:1321-1433)
        size_t free_b = 0, total_b = 0;
        BFPN_HIP(hipMemGetInfo(&free_b, &total_b));
        const size_t need = context_bytes(*grid, *phys, precision) +
                            physics_bytes(*grid, *phys, precision) + (64ull << 20);
        if (need > free_b)
            throw Error(BFPN_E_NOMEM,
                        "Requested grid needs about " + std::to_string(need >> 20) +
                            " MiB of device memory, but only " + std::to_string(free_b >> 20) +
                            " MiB is free (Ny=" + std::to_string(grid->Ny) + ", Nz=" +
                            std::to_string(grid->Nz) + ", Ng=" + std::to_string(grid->Ng) +
                            ", Nmu=" + std::to_string(grid->Nmu) + ", Nom=" + std::to_string(grid->Nom) + ")");
        auto physics = make_physics(*grid, *phys, data_path ? data_path : "./data", precision);
        *out = make_context(physics, *grid, *phys);
    });
}

BFPN_API bfpn_status bfpn_create_layered(const bfpn_grid* grid, const bfpn_phys* phys,
                                         const bfpn_layer* layers, int nlayers, int precision,
                                         bfpn_ctx** out) {
    return guarded([&] {
        BFPN_REQUIRE(grid && phys && layers && out && nlayers > 0, "null argument");
        *out = nullptr;
        validate_params(*grid, *phys);
        std::vector<LayerSpec> spec;
        for (int k = 0; k < nlayers; ++k) {
            BFPN_REQUIRE(layers[k].material && layers[k].data_path, "null layer field");
            spec.push_back({layers[k].x_end, layers[k].material, layers[k].density,
                            layers[k].data_path});
        }
        validate_layers(spec);  // before any device call
        size_t free_b = 0, total_b = 0;
        BFPN_HIP(hipMemGetInfo(&free_b, &total_b));
        const size_t need = context_bytes(*grid, *phys, precision) +
                            nlayers * physics_bytes(*grid, *phys, precision) + (64ull << 20);
        if (need > free_b)
            throw Error(BFPN_E_NOMEM, "Requested grid and " + std::to_string(nlayers) +
                                          " layers need about " + std::to_string(need >> 20) +
                                          " MiB of device memory, but only " +
                                          std::to_string(free_b >> 20) + " MiB is free");
        *out = make_layered_context(*grid, *phys, spec, precision);
    });
}

BFPN_API bfpn_status bfpn_create_mapped(const bfpn_grid* grid, const bfpn_phys* phys,
                                        const bfpn_layer* materials, int nmat,
                                        const double* slab_end, int nslabs,
                                        const unsigned char* map, int precision, bfpn_ctx** out) {
    return guarded([&] {
        BFPN_REQUIRE(grid && phys && materials && slab_end && map && out && nmat > 0 && nslabs > 0,
                     "null argument");
        *out = nullptr;
        validate_params(*grid, *phys);
        std::vector<LayerSpec> mats;
        for (int k = 0; k < nmat; ++k) {
            BFPN_REQUIRE(materials[k].material && materials[k].data_path, "null material field");
            mats.push_back({0.0, materials[k].material, materials[k].density, materials[k].data_path});
        }
        const long long nyz = static_cast<long long>(grid->Ny + 1) * (grid->Nz + 1);
        std::vector<double> ends(slab_end, slab_end + nslabs);
        std::vector<unsigned char> m(map, map + nslabs * nyz);
        validate_map(mats, ends, m, nyz);  // before any device call
        BFPN_REQUIRE(!phys->legacy_energy && !phys->lite_memory,
                     "material maps support the eq15 energy model with the full state only");
        size_t free_b = 0, total_b = 0;
        BFPN_HIP(hipMemGetInfo(&free_b, &total_b));
        const size_t need = context_bytes(*grid, *phys, precision) +
                            map_bytes(*grid, *phys, m, nmat, precision) +
                            nmat * physics_bytes(*grid, *phys, precision) + (64ull << 20);
        if (need > free_b)
            throw Error(BFPN_E_NOMEM, "Requested grid and " + std::to_string(nmat) +
                                          " materials need about " + std::to_string(need >> 20) +
                                          " MiB of device memory, but only " +
                                          std::to_string(free_b >> 20) + " MiB is free");
        *out = make_mapped_context(*grid, *phys, mats, ends, m, precision);
    });
}

BFPN_API void bfpn_destroy(bfpn_ctx* ctx) { delete ctx; }

BFPN_API bfpn_status bfpn_stopping_power(const char* material, int Ng, double Lg, double* S) {
    return guarded([&] {
        BFPN_REQUIRE(S && Ng > 0 && Lg > 0, "invalid arguments");
        const auto v = stopping_power_table(material ? material : "", Ng, Lg, Lg / Ng);
        std::copy(v.begin(), v.end(), S);
    });
}

BFPN_API bfpn_status bfpn_get_deltas(const bfpn_ctx* ctx, bfpn_deltas* out) {
    return guarded([&] {
        BFPN_REQUIRE(ctx && out, "null argument");
        *out = ctx->d;
    });
}

BFPN_API size_t bfpn_cells(const bfpn_ctx* ctx) { return ctx ? ctx->cells : 0; }
BFPN_API size_t bfpn_device_bytes(const bfpn_ctx* ctx) {
    return ctx ? ctx->device_bytes() + ctx->physics->device_bytes() : 0;
}
BFPN_API int bfpn_num_steps(const bfpn_ctx* ctx, double t_final) {
    return ctx ? static_cast<int>(std::ceil(t_final / ctx->d.dt - 1e-12)) : -1;
}

BFPN_API bfpn_status bfpn_get_table(const bfpn_ctx* ctx, int which, double* out, size_t count) {
    return guarded([&] {
        BFPN_REQUIRE(ctx && out, "null argument");
        const PhysicsBase& P = *ctx->physics;
        std::vector<double> tmp;
        const std::vector<double>* v = nullptr;
        switch (which) {
            case 0: v = &P.S; break;
            case 1: v = &P.mat.sigma; break;
            case 2: v = &P.Tc; break;
            case 3: v = &P.sig_tr; break;
            case 4: v = &P.en; break;
            case 5: v = &P.ker.e1; break;
            case 6: v = &P.ker.e2; break;
            case 7: v = &P.ker.v; break;
            case 8: v = &P.om; break;
            case 9: v = &P.mat.mu; break;
            case 10: tmp.assign(P.ker.beg.begin(), P.ker.beg.end()); v = &tmp; break;
            case 11: tmp.assign(P.ker.end.begin(), P.ker.end.end()); v = &tmp; break;
            default: throw Error(BFPN_E_INVALID, "unknown table index");
        }
        BFPN_REQUIRE(count == v->size(), "table size mismatch: expected " + std::to_string(v->size()));
        std::memcpy(out, v->data(), count * sizeof(double));
    });
}

BFPN_API bfpn_status bfpn_advance(bfpn_ctx* ctx, int nsteps) {
    return guarded([&] {
        BFPN_REQUIRE(ctx, "null context");
        ctx->advance(nsteps);
    });
}

BFPN_API bfpn_status bfpn_history(bfpn_ctx* ctx, double* idd, double* proxy, int cap, int* n) {
    return guarded([&] {
        BFPN_REQUIRE(ctx, "null context");
        ctx->history(idd, proxy, cap, n);
    });
}

BFPN_API bfpn_status bfpn_run(bfpn_ctx* ctx, double t_final, double* idd, double* proxy, int cap,
                              int* n) {
    return guarded([&] {
        BFPN_REQUIRE(ctx, "null context");
        ctx->advance(static_cast<int>(std::ceil(t_final / ctx->d.dt - 1e-12)));
        ctx->history(idd, proxy, cap, n);
    });
}

BFPN_API bfpn_status bfpn_read_state(bfpn_ctx* ctx, int which, double* host, size_t count) {
    return guarded([&] {
        BFPN_REQUIRE(ctx && host, "null argument");
        ctx->read_state(which, host, count);
    });
}

BFPN_API bfpn_status bfpn_energy_moments(bfpn_ctx* ctx, double* moments, int capacity) {
    return guarded([&] {
        BFPN_REQUIRE(ctx && moments, "null argument");
        BFPN_REQUIRE(capacity >= 2 * ctx->grid.Ng, "energy moments need 2*Ng values");
        ctx->energy_moments(moments);
    });
}

BFPN_API bfpn_status bfpn_spot_plane(bfpn_ctx* ctx, double* plane, size_t count) {
    return guarded([&] {
        BFPN_REQUIRE(ctx && plane, "null argument");
        BFPN_REQUIRE(count == static_cast<size_t>(ctx->grid.Ny + 1) * (ctx->grid.Nz + 1),
                     "spot plane needs (Ny+1)*(Nz+1) values");
        ctx->spot_plane(plane);
    });
}

BFPN_API bfpn_status bfpn_write_state(bfpn_ctx* ctx, int which, const double* host, size_t count) {
    return guarded([&] {
        BFPN_REQUIRE(ctx && host, "null argument");
        ctx->write_state(which, host, count);
    });
}

BFPN_API void* bfpn_stream(bfpn_ctx* ctx) { return ctx ? static_cast<void*>(ctx->done_stream) : nullptr; }

BFPN_API bfpn_status bfpn_set_overlap(bfpn_ctx* ctx, int on) {
    return guarded([&] {
        BFPN_REQUIRE(ctx, "null context");
        ctx->overlap = on != 0;
    });
}

BFPN_API bfpn_status bfpn_set_group_skip(bfpn_ctx* ctx, int on) {
    return guarded([&] {
        BFPN_REQUIRE(ctx, "null context");
        ctx->set_group_skip(on != 0);
    });
}

BFPN_API bfpn_status bfpn_active_groups(bfpn_ctx* ctx, int* primary_top, int* secondary_top) {
    return guarded([&] {
        BFPN_REQUIRE(ctx, "null context");
        ctx->active_tops(primary_top, secondary_top);
    });
}

BFPN_API bfpn_status bfpn_set_profiling(bfpn_ctx* ctx, int on) {
    return guarded([&] {
        BFPN_REQUIRE(ctx, "null context");
        ctx->profiling = on != 0;
        for (int k = 0; k < 5; ++k) {
            ctx->phase_seconds[k] = 0.0;
            ctx->phase_bytes[k] = 0.0;
            ctx->phase_calls[k] = 0;
        }
    });
}

BFPN_API bfpn_status bfpn_phase_stats(bfpn_ctx* ctx, double* seconds5, int* calls5,
                                      double* bytes5) {
    return guarded([&] {
        BFPN_REQUIRE(ctx, "null context");
        for (int k = 0; k < 5; ++k) {
            if (seconds5) seconds5[k] = ctx->phase_seconds[k];
            if (calls5) calls5[k] = ctx->phase_calls[k];
            if (bytes5) bytes5[k] = ctx->phase_bytes[k];
        }
    });
}

}  // extern "C"
