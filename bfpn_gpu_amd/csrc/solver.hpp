// solver.hpp — per-beam contexts over shared physics tables (internal API).
#pragma once

#include <hip/hip_runtime.h>

#include <memory>
#include <string>
#include <vector>

#include "common.hpp"
#include "tables.hpp"

namespace bfpn {

// Everything a depth march reads that does not depend on the beam itself:
// host FP64 tables, the per-group eq15 coefficients, CN factors, dose
// weights, direction table and the source operands.  Beams with the same
// grid, material tables, density, step and flags share one instance.
struct PhysicsBase {
    virtual ~PhysicsBase() = default;
    virtual size_t device_bytes() const = 0;
    std::string key;
    int precision = BFPN_FP32;
    bfpn_grid grid{};
    bfpn_deltas d{};
    int nyz = 0, nang = 0;
    long long lanes = 0;
    size_t cells = 0;
    bool secondary = true;
    MaterialTables mat;
    std::vector<double> S, Tc, sig_tr, en, om;
    SourceKernels ker;
};

std::string physics_key(const bfpn_grid& g, const bfpn_phys& p, const std::string& data_path,
                        int precision);
std::shared_ptr<PhysicsBase> make_physics(const bfpn_grid& g, const bfpn_phys& p,
                                          const std::string& data_path, int precision);
// Device bytes of a context's own arrays (state, source scratch, partials),
// excluding the shared physics.
size_t context_bytes(const bfpn_grid& g, const bfpn_phys& p, int precision);
size_t physics_bytes(const bfpn_grid& g, const bfpn_phys& p, int precision);

void validate_params(const bfpn_grid& g, const bfpn_phys& p);

// One depth layer of a layered medium: material tables and density up to
// depth x_end (cm).
struct LayerSpec {
    double x_end;
    std::string material;
    double density;
    std::string data_path;
};
void validate_layers(const std::vector<LayerSpec>& layers);

}  // namespace bfpn

// One beam: its state in HBM, history, stream; the opaque handle of the C ABI.
struct bfpn_ctx {
    virtual ~bfpn_ctx() = default;
    virtual void advance(int nsteps) = 0;
    virtual void history(double* idd, double* proxy, int cap, int* n) = 0;
    virtual void read_state(int which, double* host, size_t count) = 0;
    virtual void write_state(int which, const double* host, size_t count) = 0;
    virtual size_t device_bytes() const = 0;
    virtual int steps_done() const = 0;
    virtual void energy_moments(double* host) = 0;  // 2 Ng values
    virtual void spot_plane(double* host) = 0;      // (Ny+1)(Nz+1) values
    virtual void set_group_skip(bool on) = 0;
    virtual void active_tops(int* primary, int* secondary) = 0;

    bfpn_grid grid{};
    bfpn_phys phys{};
    bfpn_deltas d{};
    size_t cells = 0;
    hipStream_t stream = nullptr;       // primary march
    hipStream_t done_stream = nullptr;  // work issued by advance() ends on this stream
    bool overlap = true;                // secondary on its own stream (if any)
    bool profiling = false;
    double phase_seconds[5] = {0, 0, 0, 0, 0};
    double phase_bytes[5] = {0, 0, 0, 0, 0};
    int phase_calls[5] = {0, 0, 0, 0, 0};
    std::shared_ptr<bfpn::PhysicsBase> physics;
};

namespace bfpn {
bfpn_ctx* make_context(std::shared_ptr<PhysicsBase> physics, const bfpn_grid& g,
                       const bfpn_phys& p);
// A context marching through depth layers (tables built per layer; the beam
// from `p`), or over table sets built already (one per layer, from
// layer_phys(p, layer) and the layer's data path).
bfpn_ctx* make_layered_context(const bfpn_grid& g, const bfpn_phys& p,
                               const std::vector<LayerSpec>& layers, int precision);
bfpn_ctx* make_layered_context(const bfpn_grid& g, const bfpn_phys& p,
                               const std::vector<LayerSpec>& layers,
                               const std::vector<std::shared_ptr<PhysicsBase>>& tables);
// The physics parameters of one layer: `p` with the layer's material and
// density (both density and energy_density).
bfpn_phys layer_phys(const bfpn_phys& p, const LayerSpec& l);
// A context under a material map (materials as layers, x_end unused; depth
// slabs ending at slab_end; map[j (Ny+1)(Nz+1) + s] the material of node s in
// slab j), with tables built here or given (one set per material, from
// layer_phys(p, material) and its data path).
void validate_map(const std::vector<LayerSpec>& materials, const std::vector<double>& slab_end,
                  const std::vector<unsigned char>& map, long long nyz);
// Device bytes a material map adds to a context: packed copies of the nodes
// of every material but the largest one of each slab (two per stream, and the
// end-of-step copies and sources with a secondary).
size_t map_bytes(const bfpn_grid& g, const bfpn_phys& p, const std::vector<unsigned char>& map,
                 int nmat, int precision);
bfpn_ctx* make_mapped_context(const bfpn_grid& g, const bfpn_phys& p,
                              const std::vector<LayerSpec>& materials,
                              const std::vector<double>& slab_end,
                              const std::vector<unsigned char>& map, int precision);
bfpn_ctx* make_mapped_context(const bfpn_grid& g, const bfpn_phys& p,
                              const std::vector<LayerSpec>& materials,
                              const std::vector<double>& slab_end,
                              const std::vector<unsigned char>& map,
                              const std::vector<std::shared_ptr<PhysicsBase>>& tables);
}  // namespace bfpn
