// pd_context.h -- the context object shared by the host units of lib2dpdgpu
// (pd_context.hip: creation, state, steps; pd_io.hip: mesh and result files),
// with the error and allocation helpers they use.
#pragma once

#include <hip/hip_runtime.h>

#include <algorithm>
#include <cstdarg>
#include <cstdint>
#include <cstdio>
#include <string>
#include <thread>
#include <vector>

#include "pd_hip.h"
#include "pd_internal.h"

inline thread_local std::string g_err;

struct pd_context {
    pd_params p;
    int n = 0;
    long nb = 0;
    int maxM = 0;
    int lpf = 32;
    int bpl = 1;
    int mf = 0;              // family contractions on the matrix cores
    int n2_is_10 = 1;
    int dev = 0;
    std::vector<int> perm;   // internal -> caller
    std::vector<int> iperm;  // caller -> internal
    std::vector<long> bmap;  // internal bond -> caller bond
    std::vector<int> row_host, rev_host;  // internal CSR
    std::vector<double> X_caller;         // 2n, caller order
    std::vector<double> V_int;            // internal order
    // device buffers
    double2 *X = nullptr, *u = nullptr, *v = nullptr, *a = nullptr, *f = nullptr, *fown = nullptr;
    double2* v0 = nullptr;
    double *V = nullptr, *E = nullptr;
    void* kinv = nullptr;    // 3 per bond, double (FP64) or float (mixed)
    double* pf1 = nullptr;   // the uniform plane-stress factor (one value)
    // Mazars bond damage
    pd_damage_params dmgp = {PD_DAMAGE_OFF, 0, 0, 0, 0, 0};
    double *kappa = nullptr, *dmg = nullptr, *dstat = nullptr;
    int *row = nullptr, *col = nullptr, *rev = nullptr, *bc = nullptr, *ids = nullptr;
    long ids_cap = 0;
    unsigned char* broken = nullptr;
    void* W = nullptr;
    unsigned int* counters = nullptr;
    double2* stage = nullptr;           // pinned host staging for one particle vector
    bool stage_failed = false;          // pinned allocation refused: pageable staging
    // mixed-precision FP64 guard (PD_OPT_MIXED_GUARD)
    unsigned char* famflag = nullptr;   // per family, maintained by every K assembly
    int* famlist = nullptr;             // guarded families collected by one bond launch
    unsigned int* famcnt = nullptr;     // list entries (cleared by the gather)
    long guard_cond = 1000;             // 0: off
    int n_groups = 0;
    bool has_bc = false;
    // execution
    hipStream_t stream = nullptr;
    bool own_stream = false;
    int use_graph = 1, profile = 0, wpb = 4, xcd = 8, lpf_opt = 0;
    int check_finite = 1;    // pd_step reports a non-finite state (PD_OPT_CHECK_FINITE)
    int persist = 0;  // persistent bond kernel: workgroups per CU (0 = one wave per group)
    int n_cu = 256;
    hipGraph_t graph = nullptr;
    hipGraphExec_t gexec = nullptr;
    bool graph_valid = false;
    hipEvent_t ev0 = nullptr, ev1 = nullptr;
    std::vector<hipEvent_t> ev_pool;  // profiling: start/stop pairs
    std::vector<int> ev_phase;
    size_t ev_used = 0;
    double phase_ms[4] = {0, 0, 0, 0};
    long phase_calls[4] = {0, 0, 0, 0};
    long steps = 0;
    long n_singular = 0;
    std::string err;
};

// ---------------------------------------------------------------------------
// helpers
// ---------------------------------------------------------------------------
inline int fail(pd_context* c, int code, const char* fmt, ...)
{
    char buf[512];
    va_list ap;
    va_start(ap, fmt);
    vsnprintf(buf, sizeof(buf), fmt, ap);
    va_end(ap);
    if (c) c->err = buf;
    g_err = buf;
    return code;
}

#define HIPCHK(c, call)                                                                     \
    do {                                                                                    \
        hipError_t e_ = (call);                                                             \
        if (e_ != hipSuccess)                                                               \
            return fail((c), PD_ERR_HIP, "%s failed: %s", #call, hipGetErrorString(e_));   \
    } while (0)

template <typename F>
void parallel_for(long n, F fn)
{
    unsigned nt = std::max(1u, std::min(16u, std::thread::hardware_concurrency()));
    if (n < 4096 || nt == 1) {
        fn(0L, n);
        return;
    }
    std::vector<std::thread> th;
    long chunk = (n + nt - 1) / nt;
    for (unsigned t = 0; t < nt; ++t) {
        long b = t * chunk, e = std::min(n, b + chunk);
        if (b >= e) break;
        th.emplace_back([=]() { fn(b, e); });
    }
    for (auto& t : th) t.join();
}

inline uint64_t spread32(uint64_t x)
{
    x &= 0xffffffffULL;
    x = (x | (x << 16)) & 0x0000ffff0000ffffULL;
    x = (x | (x << 8)) & 0x00ff00ff00ff00ffULL;
    x = (x | (x << 4)) & 0x0f0f0f0f0f0f0f0fULL;
    x = (x | (x << 2)) & 0x3333333333333333ULL;
    x = (x | (x << 1)) & 0x5555555555555555ULL;
    return x;
}

template <typename T>
int dalloc(pd_context* c, T** p, size_t count)
{
    *p = nullptr;
    if (count == 0) count = 1;
    hipError_t e = hipMalloc((void**)p, sizeof(T) * count);
    if (e != hipSuccess)
        return fail(c, PD_ERR_NOMEM, "hipMalloc of %zu bytes failed: %s", sizeof(T) * count,
                    hipGetErrorString(e));
    return PD_OK;
}

// Mesh arrays in the internal order (X, V, row, col, rev) for c->n, c->nb.
int alloc_mesh(pd_context* c);
// pd_create_from_points: families, Morton order, internal CSR and reverse
// bonds built on the device (pd_setup.hip); sets c->nb, c->maxM and the host
// mirrors.
int build_mesh_on_device(pd_context* c, const double* X, const double* V);
