// pd_static.hip -- pd_compute_velocity_static: one force evaluation and
// velocity update of an FP64 mesh held in the caller's device buffers, with
// every intermediate in a fixed device workspace of this file, so that the
// call allocates nothing (for code under HIP graph capture, serving loops and
// harnesses that may not call hipMalloc).
//
// The arithmetic is pd_compute_velocity's on a context of the same mesh: the
// same bond kernel (tile (32,1) or (64,1), VALU contractions), K^-1 assembled
// by the same shape kernel, the same gather, and each family's bonds taken in
// the same ascending order of their reference lengths.  Particles stay in the
// caller's order (no space-filling-curve reorder), which changes no family
// sum.
//
// Workspace (`scratch`, one copy per GPU, held until the process ends):
//   rev   int32   per bond      reverse-bond index (length-ordered families)
//   kinv  double3 per bond      K^-1
//   W     double2 per bond      reverse-slot force terms
//   col   int32   per bond      neighbours, each family in length order
//   bmap  int32   per bond      the caller's bond of each length-ordered bond
//   brk   uint8   per bond      broken flags in the length order
//   fown  double2 per particle  own-bond force sums
//   misc  counters (4 x uint32) and the plane-stress factor
// Every call on a GPU uses that GPU's workspace: calls that overlap on one GPU
// (other host threads, other streams) race and must be serialised by the
// caller.  Each call rebuilds everything it reads from the workspace.
#include <hip/hip_runtime.h>

#include <mutex>

#include "pd_context.h"

using namespace pd;

namespace {

constexpr long kMaxParticles = PD_STATIC_MAX_PARTICLES;
constexpr long kMaxBonds = PD_STATIC_MAX_BONDS;
constexpr size_t kRevOff = 0;
constexpr size_t kKinvOff = kRevOff + 4 * (size_t)kMaxBonds;
constexpr size_t kWOff = kKinvOff + 24 * (size_t)kMaxBonds;
constexpr size_t kColOff = kWOff + 16 * (size_t)kMaxBonds;
constexpr size_t kBmapOff = kColOff + 4 * (size_t)kMaxBonds;
constexpr size_t kBrkOff = kBmapOff + 4 * (size_t)kMaxBonds;
constexpr size_t kFownOff = kBrkOff + (size_t)kMaxBonds;
constexpr size_t kMiscOff = kFownOff + 16 * (size_t)kMaxParticles;
constexpr size_t kScratchBytes = kMiscOff + 256;
static_assert(kKinvOff % 256 == 0 && kWOff % 256 == 0 && kColOff % 256 == 0 && kBmapOff % 256 == 0 &&
                  kBrkOff % 256 == 0 && kFownOff % 256 == 0 && kMiscOff % 256 == 0,
              "workspace sections stay 256-byte aligned");

// The workspace.  In host code `scratch` names a host-side stand-in, not
// device memory: only the per-device address from hipGetSymbolAddress
// (workspace()) may be handed to kernels or copies.
static __device__ __align__(256) unsigned char scratch[kScratchBytes];

constexpr int kMaxDevices = 64;
std::mutex g_mu;
void* g_ptr[kMaxDevices] = {};

// The device address of this GPU's workspace, looked up once per device with
// that device current.
int workspace(int dev, unsigned char** out)
{
    if (dev < 0 || dev >= kMaxDevices)
        return fail(nullptr, PD_ERR_UNSUPPORTED, "device %d: at most %d devices", dev, kMaxDevices);
    std::lock_guard<std::mutex> lk(g_mu);
    if (!g_ptr[dev]) {
        void* p = nullptr;
        hipError_t e = hipGetSymbolAddress(&p, HIP_SYMBOL(scratch));
        if (e != hipSuccess || !p)
            return fail(nullptr, PD_ERR_HIP, "hipGetSymbolAddress of the workspace failed: %s",
                        hipGetErrorString(e));
        g_ptr[dev] = p;
    }
    *out = static_cast<unsigned char*>(g_ptr[dev]);
    return PD_OK;
}

// Per particle i: its family in ascending order of the bonds' reference
// lengths (ties in the caller's order): the neighbours, the caller's bond of
// each, and its broken flag.  Families of more than 64 bonds are left to the
// mesh check.
__global__ void __launch_bounds__(256) k_static_order(int n, const int* __restrict__ row,
                                                      const int* __restrict__ col,
                                                      const unsigned char* __restrict__ broken,
                                                      const double2* __restrict__ X, int* __restrict__ col_o,
                                                      int* __restrict__ bmap, unsigned char* __restrict__ brk_o)
{
    const int i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= n) return;
    const int b0 = row[i], m = row[i + 1] - b0;
    if (m > 64) return;
    int cj[64], ord[64];
    for (int e = 0; e < m; ++e) cj[e] = col[b0 + e];
    family_length_order(m, cj, X, n, X[i], ord);
    for (int r = 0; r < m; ++r) {
        col_o[b0 + r] = cj[ord[r]];
        bmap[b0 + r] = b0 + ord[r];
        brk_o[b0 + r] = broken[b0 + ord[r]];
    }
}

// The broken flags back in the caller's order.
__global__ void __launch_bounds__(256) k_static_broken_out(long nb, const int* __restrict__ bmap,
                                                           const unsigned char* __restrict__ brk_o,
                                                           unsigned char* __restrict__ broken)
{
    const long b = (long)blockIdx.x * blockDim.x + threadIdx.x;
    if (b < nb) broken[bmap[b]] = brk_o[b];
}

// Per particle i: the reverse bond of each of its bonds (i -> j finds i in
// row j), the mesh checks (neighbour in range and not i, a reverse found,
// found once), and, by thread 0, the counters cleared and the plane-stress
// factor stored.  A failed check sets counters[2] and points the bond at its own
// slot, so that no later kernel reads or writes out of bounds.
__global__ void __launch_bounds__(256) k_static_prep(int n, const int* __restrict__ row,
                                                     const int* __restrict__ col, int* __restrict__ rev,
                                                     unsigned int* __restrict__ counters, double* pf,
                                                     double pf_value, int max_family)
{
    const int i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i == 0) {
        counters[0] = 0u;
        counters[1] = 0u;
        counters[3] = 0u;
        *pf = pf_value;
    }
    if (i >= n) return;
    bool bad = row[i + 1] - row[i] > max_family;
    for (int b = row[i]; b < row[i + 1]; ++b) {
        const int j = col[b];
        int r = -1, hits = 0;
        if (j >= 0 && j < n && j != i) {
            for (int e = row[j]; e < row[j + 1]; ++e)
                if (col[e] == i) {
                    if (r < 0) r = e;
                    ++hits;
                }
        }
        // a neighbour listed twice finds the same reverse twice: its second
        // copy is also a duplicate entry of row i
        for (int e = row[i]; e < b && hits == 1; ++e)
            if (col[e] == j) hits = 2;
        if (r < 0 || hits != 1) {
            bad = true;
            r = b;
        }
        rev[b] = r;
    }
    if (bad) counters[2] = 1u;
}

__global__ void k_static_clear_flag(unsigned int* counters) { counters[2] = 0u; }

}  // namespace

extern "C" int pd_compute_velocity_static(const pd_params* p, int n, long nb, int max_family,
                                          const double* X, const double* V, const int* row_ptr,
                                          const int* col, unsigned char* broken, const double* u,
                                          double* v, double* a, double* f, int write_acceleration,
                                          void* stream)
{
    if (!p) return fail(nullptr, PD_ERR_INVALID, "params is NULL");
    if (p->ndim != 2) return fail(nullptr, PD_ERR_UNSUPPORTED, "only ndim == 2 is supported (got %d)", p->ndim);
    if (p->precision != PD_FP64)
        return fail(nullptr, PD_ERR_UNSUPPORTED, "pd_compute_velocity_static runs PD_FP64 only");
    if (!(p->dx > 0) || !(p->horizon > 0) || !(p->rho > 0) || !(p->dt > 0) || !(p->E > 0) ||
        !(p->nu > -1.0 && p->nu < 1.0) || !(p->n1 >= 0) || !(p->n2 >= 0))
        return fail(nullptr, PD_ERR_INVALID, "invalid material or discretisation parameters");
    if (n <= 0 || n > kMaxParticles)
        return fail(nullptr, PD_ERR_INVALID, "n = %d particles: the workspace holds 1 .. %ld", n, kMaxParticles);
    if (nb < 0 || nb > kMaxBonds)
        return fail(nullptr, PD_ERR_INVALID, "nb = %ld bonds: the workspace holds at most %ld", nb, kMaxBonds);
    if (max_family < 0 || max_family > 64)
        return fail(nullptr, PD_ERR_UNSUPPORTED, "family of %d bonds: at most 64 are supported", max_family);
    if (!X || !V || !row_ptr || !u || !v || !a || !f || (nb > 0 && (!col || !broken)))
        return fail(nullptr, PD_ERR_INVALID, "NULL device array");
    const void* v2[] = {X, u, v, a, f};
    for (const void* q : v2)
        if (reinterpret_cast<uintptr_t>(q) % 16 != 0)
            return fail(nullptr, PD_ERR_INVALID, "X, u, v, a and f must be 16-byte aligned (double2)");
    hipStream_t s = static_cast<hipStream_t>(stream);
    int dev = 0;
    if (hipGetDevice(&dev) != hipSuccess) return fail(nullptr, PD_ERR_HIP, "no HIP device available");
    unsigned char* ws = nullptr;
    if (int rc = workspace(dev, &ws)) return rc;
    int* rev = reinterpret_cast<int*>(ws + kRevOff);
    int* col_o = reinterpret_cast<int*>(ws + kColOff);
    int* bmap = reinterpret_cast<int*>(ws + kBmapOff);
    unsigned char* brk_o = ws + kBrkOff;
    unsigned int* counters = reinterpret_cast<unsigned int*>(ws + kMiscOff);
    double* pf = reinterpret_cast<double*>(ws + kMiscOff + 64);

    // mesh checks and reverse bonds; outside graph capture the call waits for
    // them and reports a bad mesh before any force is evaluated
    hipStreamCaptureStatus cap = hipStreamCaptureStatusNone;
    if (hipStreamIsCapturing(s, &cap) != hipSuccess)
        return fail(nullptr, PD_ERR_HIP, "hipStreamIsCapturing failed");
    hipLaunchKernelGGL(k_static_clear_flag, dim3(1), dim3(1), 0, s, counters);
    if (nb > 0)
        hipLaunchKernelGGL(k_static_order, dim3((n + 255) / 256), dim3(256), 0, s, n, row_ptr, col, broken,
                           reinterpret_cast<const double2*>(X), col_o, bmap, brk_o);
    hipLaunchKernelGGL(k_static_prep, dim3((n + 255) / 256), dim3(256), 0, s, n, row_ptr, col_o, rev, counters,
                       pf, p->E / (1.0 - p->nu * p->nu), max_family);
    hipError_t e = hipGetLastError();
    if (e != hipSuccess) return fail(nullptr, PD_ERR_HIP, "mesh check launch failed: %s", hipGetErrorString(e));
    if (cap == hipStreamCaptureStatusNone) {
        unsigned int flag = 0;
        e = hipMemcpyAsync(&flag, counters + 2, sizeof(flag), hipMemcpyDeviceToHost, s);
        e = e != hipSuccess ? e : hipStreamSynchronize(s);
        if (e != hipSuccess) return fail(nullptr, PD_ERR_HIP, "mesh check failed: %s", hipGetErrorString(e));
        if (flag)
            return fail(nullptr, PD_ERR_INVALID,
                        "families must be symmetric, with neighbours in range, not the particle itself and "
                        "listed once");
    }

    BondArgs ba = {};
    ba.n = n;
    ba.xcd = 8;
    ba.X = reinterpret_cast<const double2*>(X);
    ba.V = V;
    ba.row = row_ptr;
    ba.col = nb > 0 ? col_o : rev;   // the bond kernel reads entry 0 for idle lanes
    ba.rev = rev;
    ba.u = reinterpret_cast<const double2*>(u);
    ba.broken = nb > 0 ? brk_o : reinterpret_cast<unsigned char*>(ws + kMiscOff + 128);
    ba.kinv = ws + kKinvOff;
    ba.kinv_f32 = 0;
    ba.W = ws + kWOff;
    ba.fown = reinterpret_cast<double2*>(ws + kFownOff);
    ba.pf = pf;
    ba.pf_mask = 0;
    ba.nu = p->nu;
    ba.A = p->n1 / (p->horizon * p->dx);
    ba.n2 = p->n2;
    ba.s0 = p->s0;
    ba.brk_c = (1.0 + p->s0) * (1.0 + p->s0);
    ba.ktol = p->kappa_tol > 0.0 ? p->kappa_tol : 1e-10;
    ba.counters = counters;
    ba.illr = 1e300;
    const int lpf = max_family <= 32 ? 32 : 64;
    const int n2_is_10 = p->n2 == 10.0;
    ba.force_k = 1;
    e = launch_shape(ba, lpf, 1, n2_is_10, 4, s);
    ba.force_k = 0;
    e = e != hipSuccess ? e : launch_bond(ba, 0, lpf, 1, 0, 0, n2_is_10, 4, 0, s);
    if (e == hipSuccess && nb > 0) {
        hipLaunchKernelGGL(k_static_broken_out, dim3((unsigned)((nb + 255) / 256)), dim3(256), 0, s, nb, bmap,
                           brk_o, broken);
        e = hipGetLastError();
    }
    GatherArgs g = {};
    g.n = n;
    g.xcd = 8;
    g.row = row_ptr;
    g.fown = ba.fown;
    g.W = ba.W;
    g.v = reinterpret_cast<double2*>(v);
    g.a = reinterpret_cast<double2*>(a);
    g.f = reinterpret_cast<double2*>(f);
    g.rho = p->rho;
    g.dt = p->dt;
    g.write_a = write_acceleration ? 1 : 0;
    g.max_family = max_family;
    g.famcnt = nullptr;
    e = e != hipSuccess ? e : launch_gather(g, 0, s, nullptr);
    if (e != hipSuccess) return fail(nullptr, PD_ERR_HIP, "kernel launch failed: %s", hipGetErrorString(e));
    return PD_OK;
}
