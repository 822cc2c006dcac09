// pd_mechanics_adapter.hpp -- header-only C++ adapter that keeps the reference's
// mechanics signatures and runs them on lib2dpdgpu (pd_hip.h).
//
// A driver written against EricYJA/2D_PD_GPU's
//   compute_velocity_GPU_host(...)   include/mechanics.cuh:867-881 (called at src/main.cu:628)
//   computeVelocity(...)             src/main_seq.cpp:396-398       (called at src/main_seq.cpp:698)
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

template <class ParticleT>
uint64_t mesh_signature(const std::vector<ParticleT>& ps)
{
    uint64_t h = 1469598103934665603ULL;
    for (const auto& p : ps) {
        h = (h ^ (uint64_t)p.neighbors.size()) * 1099511628211ULL;
        for (int nb : p.neighbors) h = (h ^ (uint64_t)(uint32_t)nb) * 1099511628211ULL;
        h = (h ^ (uint64_t)std::llround((double)p.initialPositions[0] * 1e9)) * 1099511628211ULL;
    }
    return h;
}

// Shared body of both entry points.
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

    Cache& c = cache();
    const uint64_t sig = mesh_signature(allParticles);
    const bool same = c.ctx && c.n == n && c.sig == sig && c.p.E == p.E && c.p.nu == p.nu &&
                      c.p.n1 == p.n1 && c.p.n2 == p.n2 && c.p.horizon == p.horizon &&
                      c.p.dx == p.dx && c.p.rho == p.rho && c.p.dt == p.dt;
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
    std::vector<double> u(2 * (size_t)n), v(2 * (size_t)n), a(2 * (size_t)n), f(2 * (size_t)n);
    for (int i = 0; i < n; ++i)
        for (int d = 0; d < 2; ++d) {
            const ParticleT& q = allParticles[i];
            u[2 * i + d] = (double)(q.currentPositions[d] - q.initialPositions[d]);
            v[2 * i + d] = (double)velocity[i][d];
            a[2 * i + d] = (double)acce[i][d];
            f[2 * i + d] = (double)netF[i][d];
        }
    check(pd_set_state(c.ctx, u.data(), v.data(), a.data(), f.data()), c.ctx);
    check(pd_compute_velocity(c.ctx, write_acce ? 1 : 0), c.ctx);
    check(pd_get_state(c.ctx, nullptr, v.data(), a.data(), f.data()), c.ctx);
    for (int i = 0; i < n; ++i)
        for (int d = 0; d < 2; ++d) {
            velocity[i][d] = (long double)v[2 * i + d];
            acce[i][d] = write_acce ? (long double)a[2 * i + d] : (long double)(double)acce[i][d];
            netF[i][d] = (long double)f[2 * i + d];
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
