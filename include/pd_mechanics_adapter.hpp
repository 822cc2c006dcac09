// pd_mechanics_adapter.hpp -- header-only C++ adapter that keeps the reference's
// mechanics signatures and runs them on lib2dpdgpu (pd_hip.h).
//
// A driver written against EricYJA/2D_PD_GPU's
//   compute_velocity_GPU_host(...)   include/mechanics.cuh:867-881 (called at src/main.cu:628)
//   computeVelocity(...)             src/main_seq.cpp:396-398       (called at src/main_seq.cpp:698)
//   computeDamageStatus(...)         src/main_seq.cpp:473, include/mechanics.h:24
//                                    (called at src/main_seq.cpp:683)
// recompiles unchanged against this header: the functions are templates over the
// driver's own `Particle` (include/SetupParticleSystem.h:11-22) and `matrix`
// (include/matrix.h:7-26) types, so no reference header is needed here.
//
// Semantics kept from the reference GPU path:
//   * families come from Particle::neighbors through globalLocalIDmap; the
//     Neighborslist and bondDamage arguments are not read
//     (include/mechanics.cuh:926-944);
//   * velocity / acce / netF are overwritten in place with double-rounded values
//     (include/mechanics.cuh:886-908, :1113-1121);
//   * compute_velocity_GPU_host stores acce = netF / rho (include/mechanics.cuh:863),
//     computeVelocity (the CPU path) leaves acce untouched.
// Differences, by design:
//   * the current bond vector is y_l - y_i (the CPU semantics of src/main_seq.cpp:105),
//     not the reference kernel's y_l - X_i (include/mechanics.cuh:165);
//   * the device context persists across calls while the mesh is unchanged
//     (the reference allocates, uploads and frees everything per call);
//   * the displacement u = y - X is formed in long double before rounding, so
//     the bond strain is free of the cancellation of forming y in double.
#pragma once

#include <cmath>
#include <cstdint>
#include <cstring>
#include <stdexcept>
#include <string>
#include <unordered_map>
#include <vector>

#include "pd_hip.h"

namespace pd_adapter {

struct Cache {
    pd_context* ctx = nullptr;
    int n = -1;
    uint64_t sig = 0;
    pd_params p{};
    ~Cache()
    {
        if (ctx) pd_destroy(ctx);
    }
};

inline Cache& cache()
{
    static Cache c;
    return c;
}

inline void check(int rc, pd_context* ctx)
{
    if (rc != PD_OK) throw std::runtime_error(std::string("lib2dpdgpu: ") + pd_last_error(ctx));
}

// FNV-1a over everything the context is built from: each particle's family
// (neighbour ids as mapped to local indices, -1 where the map has none) and the
// exact bits of its initial position (x and y, as the double the context
// stores) and volume.  A mesh that differs in any of them gets a new context.
inline uint64_t fnv_mix(uint64_t h, uint64_t x) { return (h ^ x) * 1099511628211ULL; }

inline uint64_t double_bits(double d)
{
    uint64_t b;
    std::memcpy(&b, &d, sizeof b);
    return b;
}

template <class ParticleT>
uint64_t mesh_signature(const std::vector<ParticleT>& ps, const std::unordered_map<int, int>& globalLocalIDmap)
{
    uint64_t h = 1469598103934665603ULL;
    for (const auto& p : ps) {
        h = fnv_mix(h, (uint64_t)p.neighbors.size());
        for (int nb : p.neighbors) {
            auto it = globalLocalIDmap.find(nb);
            h = fnv_mix(h, (uint64_t)(uint32_t)(it == globalLocalIDmap.end() ? -1 : it->second));
        }
        h = fnv_mix(h, double_bits((double)p.initialPositions[0]));
        h = fnv_mix(h, double_bits((double)p.initialPositions[1]));
        h = fnv_mix(h, double_bits((double)p.volume));
    }
    return h;
}

// The cached context for this mesh; created (or re-created) when the mesh or
// the parameters change.  geometry_only: any material will do (damage needs
// only the families and the weights), so an existing context for the same mesh
// and weight parameters is kept.
template <class ParticleT>
pd_context* mesh_context(const pd_params& p, const std::vector<ParticleT>& allParticles,
                         const std::unordered_map<int, int>& globalLocalIDmap, bool geometry_only)
{
    const int n = (int)allParticles.size();
    Cache& c = cache();
    const uint64_t sig = mesh_signature(allParticles, globalLocalIDmap);
    const bool same_mesh = c.ctx && c.n == n && c.sig == sig && c.p.n1 == p.n1 && c.p.n2 == p.n2 &&
                           c.p.horizon == p.horizon && c.p.dx == p.dx;
    const bool same = same_mesh && (geometry_only || (c.p.E == p.E && c.p.nu == p.nu &&
                                                      c.p.rho == p.rho && c.p.dt == p.dt));
    if (!same) {
        if (c.ctx) pd_destroy(c.ctx);
        c.ctx = nullptr;
        std::vector<double> X(2 * (size_t)n), V(n);
        std::vector<int> row(n + 1, 0), col;
        for (int i = 0; i < n; ++i) {
            const ParticleT& q = allParticles[i];
            X[2 * i] = (double)q.initialPositions[0];
            X[2 * i + 1] = (double)q.initialPositions[1];
            V[i] = (double)q.volume;
            for (int g : q.neighbors) {
                auto it = globalLocalIDmap.find(g);
                if (it != globalLocalIDmap.end()) col.push_back(it->second);
            }
            row[i + 1] = (int)col.size();
        }
        check(pd_create(&p, n, X.data(), V.data(), row.data(), col.data(), &c.ctx), nullptr);
        c.n = n;
        c.sig = sig;
        c.p = p;
    }
    return c.ctx;
}

// u = y - X, formed in long double before rounding.
template <class ParticleT>
std::vector<double> displacement(const std::vector<ParticleT>& allParticles)
{
    const int n = (int)allParticles.size();
    std::vector<double> u(2 * (size_t)n);
    for (int i = 0; i < n; ++i)
        for (int d = 0; d < 2; ++d)
            u[2 * i + d] = (double)(allParticles[i].currentPositions[d] -
                                    allParticles[i].initialPositions[d]);
    return u;
}

// Shared body of both velocity entry points.
template <class ParticleT, class MatrixT>
void compute_velocity(double n1, double n2, double horizon, double dx, double massDensity,
                      MatrixT& StiffnessTensor, double stepSize,
                      std::vector<std::vector<long double>>& velocity,
                      std::vector<std::vector<long double>>& acce,
                      std::vector<std::vector<long double>>& netF,
                      std::vector<ParticleT>& allParticles,
                      const std::unordered_map<int, int>& globalLocalIDmap, bool write_acce)
{
    const int n = (int)allParticles.size();
    // plane stress C = E/(1-nu^2) [[1, nu, 0], [nu, 1, 0], [0, 0, 1-nu]] (src/main_seq.cpp:180-201)
    const double c00 = (double)StiffnessTensor.elements[0][0];
    const double nu = (double)StiffnessTensor.elements[0][1] / c00;
    pd_params p{};
    p.ndim = 2;
    p.precision = PD_FP64;
    p.n1 = n1;
    p.n2 = n2;
    p.horizon = horizon;
    p.dx = dx;
    p.rho = massDensity;
    p.dt = stepSize;
    p.E = c00 * (1.0 - nu * nu);
    p.nu = nu;
    p.s0 = 0.0;  // the reference never breaks bonds
    p.kappa_tol = 0.0;

    pd_context* ctx = mesh_context(p, allParticles, globalLocalIDmap, false);
    std::vector<double> u = displacement(allParticles);
    std::vector<double> v(2 * (size_t)n), a(2 * (size_t)n), f(2 * (size_t)n);
    for (int i = 0; i < n; ++i)
        for (int d = 0; d < 2; ++d) {
            v[2 * i + d] = (double)velocity[i][d];
            a[2 * i + d] = (double)acce[i][d];
            f[2 * i + d] = (double)netF[i][d];
        }
    check(pd_set_state(ctx, u.data(), v.data(), a.data(), f.data()), ctx);
    check(pd_compute_velocity(ctx, write_acce ? 1 : 0), ctx);
    check(pd_get_state(ctx, nullptr, v.data(), a.data(), f.data()), ctx);
    for (int i = 0; i < n; ++i)
        for (int d = 0; d < 2; ++d) {
            velocity[i][d] = (long double)v[2 * i + d];
            acce[i][d] = write_acce ? (long double)a[2 * i + d] : (long double)(double)acce[i][d];
            netF[i][d] = (long double)f[2 * i + d];
        }
}

// computeDamageStatus body: the caller owns the per-bond damage and thresholds
// (bondDamage / bondDamageThreshold, indexed [particle][position in its
// neighbour list]); they are loaded, updated at the particles' current
// positions with the reference's Mazars constants (src/main_seq.cpp:480), and
// written back, with Particle::damageStatus.
template <class ParticleT>
void compute_damage(double n1, double n2, double dx, double horizon,
                    std::vector<std::vector<double>>& bondDamageThreshold,
                    std::vector<ParticleT>& localParticles,
                    std::vector<std::vector<double>>& bondDamage,
                    const std::unordered_map<int, int>& globalLocalIDmap)
{
    const int n = (int)localParticles.size();
    pd_params p{};
    p.ndim = 2;
    p.precision = PD_FP64;
    p.n1 = n1;
    p.n2 = n2;
    p.horizon = horizon;
    p.dx = dx;
    p.rho = p.dt = p.E = 1.0;  // not used by the damage update
    p.nu = 0.0;
    pd_context* ctx = mesh_context(p, localParticles, globalLocalIDmap, true);
    // the bond order of the context: neighbours found in the id map, in order
    std::vector<double> d, k;
    std::vector<int> cnt(n, 0);
    for (int i = 0; i < n; ++i) {
        int j = 0;
        for (int g : localParticles[i].neighbors) {
            if (globalLocalIDmap.find(g) == globalLocalIDmap.end()) continue;
            d.push_back(bondDamage.at(i).at(j));
            k.push_back(bondDamageThreshold.at(i).at(j));
            ++j;
        }
        cnt[i] = j;
    }
    pd_damage_params dp{PD_DAMAGE_REFERENCE, 1.0, 0.87, 20000.0, 0.65, 2150.0};
    check(pd_set_damage(ctx, &dp), ctx);
    check(pd_set_damage_state(ctx, d.data(), k.data()), ctx);
    std::vector<double> u = displacement(localParticles);
    check(pd_set_state(ctx, u.data(), nullptr, nullptr, nullptr), ctx);
    check(pd_compute_damage(ctx), ctx);
    std::vector<double> st(n);
    check(pd_get_damage(ctx, d.data(), st.data(), k.data()), ctx);
    // the caller owns the state: stop updating it in later force evaluations
    dp.mode = PD_DAMAGE_OFF;
    check(pd_set_damage(ctx, &dp), ctx);
    size_t e = 0;
    for (int i = 0; i < n; ++i) {
        for (int j = 0; j < cnt[i]; ++j, ++e) {
            bondDamage[i][j] = d[e];
            bondDamageThreshold[i][j] = k[e];
        }
        // damageStatus divides by the neighbour count (src/main_seq.cpp:593)
        const size_t m = localParticles[i].neighbors.size();
        localParticles[i].damageStatus = m ? st[i] * cnt[i] / (double)m : 0.0;
    }
}

}  // namespace pd_adapter

// include/mechanics.cuh:867-881
template <class ParticleT, class MatrixT>
void compute_velocity_GPU_host(int ndim, double n1, double n2, double horizon, double dx,
                               double massDensity, MatrixT& StiffnessTensor, double stepSize,
                               std::vector<std::vector<long double>>& velocity,
                               std::vector<std::vector<ParticleT*>>& Neighborslist,
                               std::vector<std::vector<long double>>& acce,
                               std::vector<std::vector<long double>>& netF,
                               std::vector<ParticleT>& allParticles,
                               const std::unordered_map<int, int>& globalLocalIDmap,
                               std::vector<std::vector<double>>& bondDamage)
{
    (void)Neighborslist;
    (void)bondDamage;
    if (ndim != 2) throw std::runtime_error("lib2dpdgpu: only ndim == 2 is supported");
    pd_adapter::compute_velocity(n1, n2, horizon, dx, massDensity, StiffnessTensor, stepSize,
                                 velocity, acce, netF, allParticles, globalLocalIDmap, true);
}

// src/main_seq.cpp:396-398
template <class ParticleT, class MatrixT>
void computeVelocity(int ndim, double n1, double n2, double horizon, double dx,
                     long double massDensity, MatrixT& StiffnessTensor, long double stepSize,
                     std::vector<std::vector<long double>>& velocity,
                     std::vector<std::vector<ParticleT*>>& Neighborslist,
                     std::vector<std::vector<long double>>& acce,
                     std::vector<std::vector<long double>>& netF,
                     std::vector<ParticleT>& localParticles,
                     const std::unordered_map<int, int>& globalLocalIDmap,
                     std::vector<std::vector<double>>& bondDamage)
{
    (void)Neighborslist;
    (void)bondDamage;
    if (ndim != 2) throw std::runtime_error("lib2dpdgpu: only ndim == 2 is supported");
    pd_adapter::compute_velocity(n1, n2, horizon, dx, (double)massDensity, StiffnessTensor,
                                 (double)stepSize, velocity, acce, netF, localParticles,
                                 globalLocalIDmap, false);
}

// src/main_seq.cpp:473 (include/mechanics.h:24)
template <class ParticleT>
void computeDamageStatus(int ndim, double n1, double n2, double dx, double horizon,
                         std::vector<std::vector<double>>& bondDamageThreshold,
                         std::vector<ParticleT>& localParticles,
                         const std::vector<std::vector<ParticleT*>>& Neighborslist,
                         std::vector<std::vector<double>>& bondDamage,
                         const std::unordered_map<int, int>& globalLocalIDmap)
{
    (void)Neighborslist;
    if (ndim != 2) throw std::runtime_error("lib2dpdgpu: only ndim == 2 is supported");
    pd_adapter::compute_damage(n1, n2, dx, horizon, bondDamageThreshold, localParticles, bondDamage,
                               globalLocalIDmap);
}
