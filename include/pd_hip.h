/*
 * pd_hip.h -- C ABI of lib2dpdgpu, the MI355X (gfx950) implementation of the
 * explicit bond-associated NOSB peridynamics timestep of EricYJA/2D_PD_GPU.
 *
 * Plain C, POD arguments, status-code returns, no exceptions across the ABI.
 * All arrays are host arrays in the CALLER's particle order and the caller's
 * bond (CSR) order; the library reorders particles along a space-filling curve
 * internally and keeps every buffer resident in device memory for the life of
 * the context.  One context per host thread.
 *
 * Reference entry points replaced (EricYJA/2D_PD_GPU):
 *   computeVelocity        src/main_seq.cpp:396-470, include/mechanics.h:21-23,
 *                          compute_velocity_GPU_host include/mechanics.cuh:867-1136
 *                            -> pd_compute_velocity
 *   updatePositions        src/main_seq.cpp:386-393, include/mechanics.h:20
 *                            -> pd_update_positions
 *   applyVelocityBC        src/main_seq.cpp:28-36,  include/mechanics.h:9
 *                            -> pd_apply_velocity_bc
 *   applyFixedBC           src/main_seq.cpp:38-46,  include/mechanics.h:10
 *                            -> pd_apply_fixed_bc
 *   getStiffnessTensor     src/main_seq.cpp:180-201 -> pd_params {E, nu}
 *   buildLocalNeighborlist src/main_seq.cpp:49-68, findNeighbor
 *                          src/SetupParticleSystem.cpp:100-136
 *                            -> pd_create (CSR) / pd_create_padded (the GPU
 *                               path's padded layout, include/mechanics.cuh:
 *                               1033-1046) / pd_build_families
 *   computeDamageStatus    src/main_seq.cpp:473-598, include/mechanics.h:24
 *                            -> pd_set_damage / pd_compute_damage / pd_get_damage
 *   step loop              src/main_seq.cpp:680-715 -> pd_step
 *   readMeshFile, storeVelocity, outputLocalParticlesPositions,
 *   outputGatheredPositions -> pd_read_mesh, pd_write_velocity,
 *                              pd_write_positions, pd_write_gathered
 * A header-only C++ adapter keeping the reference signatures is in
 * include/pd_mechanics_adapter.hpp.
 */
#ifndef PD_HIP_H
#define PD_HIP_H

#include <stddef.h>

#ifdef __cplusplus
extern "C" {
#endif

/* ---- status codes ---------------------------------------------------------- */
#define PD_OK               0
#define PD_ERR_INVALID      1  /* bad argument or inconsistent mesh            */
#define PD_ERR_HIP          2  /* a HIP runtime call failed                    */
#define PD_ERR_UNSUPPORTED  3  /* e.g. family larger than 64 bonds, ndim != 2  */
#define PD_ERR_NOMEM        4  /* device or host allocation failed             */
#define PD_ERR_NOT_FINITE   5  /* a non-finite value appeared in the state     */

/* ---- precision ------------------------------------------------------------- */
#define PD_FP64   0  /* every bond quantity in FP64                             */
#define PD_MIXED  1  /* FP32 bond arithmetic (F, sigma, T, force state);
                        FP64 K assembly/inversion and FP64 u, v, a, f           */

/* ---- boundary codes (pd_set_boundary) --------------------------------------- */
#define PD_BC_FREE  (-1)  /* no constraint                                       */
#define PD_BC_FIXED (-2)  /* applyFixedBC: u = 0 (y = X) after the position update */
                          /* g >= 0: applyVelocityBC with velocity group g       */

/* ---- options (pd_set_option) ----------------------------------------------- */
#define PD_OPT_USE_GRAPH     1  /* 1: pd_step replays a captured hipGraph (default 1) */
#define PD_OPT_PROFILE       2  /* 1: time each kernel phase with HIP events          */
#define PD_OPT_WAVES_PER_WG  3  /* families-per-wavefront tile: waves per workgroup   */
#define PD_OPT_XCD_REMAP     4  /* 1: map consecutive families to one XCD (default 1) */
#define PD_OPT_TILE          5  /* family tile: 10000*matrix_cores + 100*lanes_per_family
                                   + bonds_per_lane: 3201, 6401, 1602, 1603 (and 3202
                                   in mixed precision); matrix-core contractions
                                   13201, 16401, 11602, 11603 (and 13202 in mixed
                                   precision); matrix-core weights and contractions
                                   23201, 26401, 21603 (PD_FP64, n2 > 0); rank-form
                                   weights 31602, 31603 (and 41602, 41603 with
                                   matrix-core contractions in mixed precision;
                                   n2 > 0);
                                   0 = 3201 or 6401 by the largest family            */
#define PD_OPT_PERSISTENT    6  /* bond kernel as a persistent grid of value workgroups
                                   per CU that prefetch their next family group
                                   (0 = one wave per family group, the default)      */
#define PD_OPT_CHECK_FINITE  7  /* 1: pd_step returns PD_ERR_NOT_FINITE, naming the
                                   first particle, when u, v or f is not finite after
                                   its steps (default 1; one reduction per call)     */
#define PD_OPT_MIXED_GUARD   8  /* PD_MIXED only: condition-number bound c of the
                                   FP64 guard.  A family with an alive bond whose K
                                   has cond(K) >= c is evaluated again in FP64 after
                                   the FP32 bond kernel, every step (FP32 sums err
                                   by about 6e-8 cond(K) along K's weak direction).
                                   Default 1000; 0 turns the guard off              */

/* ---- Mazars bond damage (pd_set_damage) ------------------------------------ */
#define PD_DAMAGE_OFF        0
#define PD_DAMAGE_REFERENCE  1  /* the law as computeDamageStatus writes it: the
                                   running threshold kappa_ik in d_t, d_c
                                   (src/main_seq.cpp:579-582)                     */
#define PD_DAMAGE_MAZARS     2  /* the published law: the initial threshold kappa0 */

typedef struct pd_damage_params {
    int    mode;      /* PD_DAMAGE_* */
    double kappa0;    /* initial threshold (reference: tensileStrength / E, src/main_seq.cpp:660) */
    double At, Bt;    /* tension law (reference: 0.87, 20000, src/main_seq.cpp:480) */
    double Ac, Bc;    /* compression law (reference: 0.65, 2150) */
} pd_damage_params;

typedef struct pd_params {
    int    ndim;       /* must be 2 */
    int    precision;  /* PD_FP64 or PD_MIXED */
    double n1, n2;     /* weight exponents (reference: 10, 10) */
    double horizon;    /* delta in units of dx */
    double dx;         /* grid spacing */
    double rho;        /* mass density */
    double dt;         /* step size */
    double E;          /* Young's modulus (per-particle values: pd_set_youngs_modulus) */
    double nu;         /* Poisson ratio (plane stress) */
    double s0;         /* critical stretch; <= 0 disables bond breaking */
    double kappa_tol;  /* K_ik treated as singular when det <= kappa_tol*tr^2; 0 -> 1e-10 */
} pd_params;

typedef struct pd_stats {
    int    n;              /* particles */
    long   nb;             /* directed bonds */
    long   n_broken;       /* directed bonds currently broken */
    long   n_singular;     /* bonds whose K was singular at the last assembly */
    int    max_family;     /* largest family */
    int    lanes_per_family;
    int    bonds_per_lane;
    int    persistent;     /* PD_OPT_PERSISTENT in effect */
    int    waves_per_wg;   /* PD_OPT_WAVES_PER_WG in effect */
    long   steps;          /* steps run by pd_step since creation */
    double phase_ms[4];    /* profiled time: positions, bonds, gather, other */
    long   phase_calls[4];
    int    matrix_cores;   /* 1: family contractions on the matrix cores (MFMA) */
    long   n_ill_conditioned; /* alive bonds whose K has condition number >= PD_ILL_CONDITIONED
                                 now (counted from the stored K^-1 on every pd_get_stats) */
    long   n_guarded_families; /* PD_MIXED: families the FP64 guard re-evaluates each step */
} pd_stats;

/* Condition number from which a shape tensor counts as ill-conditioned.  The
   mixed-precision step forms F - I = D K^-1 from FP32 sums, whose relative
   error along K's weak direction grows as about 6e-8 * cond(K).  The FP64
   guard (PD_OPT_MIXED_GUARD, bound 1000 by default) re-evaluates the families
   holding such bonds in FP64; with the guard off, PD_MIXED can drift from
   PD_FP64 by more than 1e-4 once bonds reach this bound. */
#define PD_ILL_CONDITIONED 1e4

typedef struct pd_context pd_context;

/* Last error message; ctx may be NULL for errors raised before a context
 * existed (thread-local). */
const char* pd_last_error(const pd_context* ctx);

/* Library / device identification ("gfx950 lib2dpdgpu x.y"). */
const char* pd_version(void);

/* Create a context.  X: 2n initial positions (xy interleaved), V: n volumes,
 * families as CSR (row_ptr n+1, col nb) in the caller's particle ids.  Families
 * must be symmetric (j in H_i <=> i in H_j) and hold at most 64 bonds. */
int pd_create(const pd_params* p, int n, const double* X, const double* V,
              const int* row_ptr, const int* col, pd_context** out);

/* Same, from the reference GPU path's padded layout: nbr[n*cap] padded with
 * INT_MAX, nbr_count[n] (include/mechanics.cuh:1033-1046). */
int pd_create_padded(const pd_params* p, int n, const double* X, const double* V,
                     const int* nbr, const int* nbr_count, int cap, pd_context** out);

/* Same, from particles alone: the families (findNeighbor's rule, cutoff
 * |xi| < 1.01*dx*horizon, src/SetupParticleSystem.cpp:100-136, neighbours in
 * ascending id order) are built on the device from a cell list, with the
 * space-filling-curve order, internal CSR and reverse-bond index; the result
 * equals pd_create with pd_build_families' families.  pd_get_families returns
 * them (they define the caller's bond order of every per-bond array). */
int pd_create_from_points(const pd_params* p, int n, const double* X, const double* V,
                          pd_context** out);
/* The context's families in the caller's order: row_ptr[n+1] and, when col is
 * not NULL, col[nb] (nb = row_ptr[n]). */
int pd_get_families(pd_context* ctx, int* row_ptr, int* col);

void pd_destroy(pd_context* ctx);

/* Work on the caller's HIP stream (hipStream_t); NULL = the context's own. */
int pd_set_stream(pd_context* ctx, void* hip_stream);
int pd_synchronize(pd_context* ctx);
int pd_set_option(pd_context* ctx, int option, long value);

/* Per-particle Young's modulus (n values, caller order). */
int pd_set_youngs_modulus(pd_context* ctx, const double* E);

/* State in caller order, double[2n] each; NULL pointers are skipped.
 * u = y - X (displacement), v velocity, a acceleration, f net force density. */
int pd_set_state(pd_context* ctx, const double* u, const double* v, const double* a,
                 const double* f);
int pd_get_state(pd_context* ctx, double* u, double* v, double* a, double* f);
/* Current positions y (double[2n]); u is formed as y - X. */
int pd_set_positions(pd_context* ctx, const double* y);
int pd_get_positions(pd_context* ctx, double* y);

/* ---- phase calls mirroring the reference ------------------------------------ */
/* applyVelocityBC: v = v0, a = 0 for ids (caller ids). */
int pd_apply_velocity_bc(pd_context* ctx, int count, const int* ids, const double* v0);
/* applyFixedBC: y = X (u = 0) for ids. */
int pd_apply_fixed_bc(pd_context* ctx, int count, const int* ids);
/* updatePositions: a = f/rho; u += v dt + 0.5 a dt^2, all particles. */
int pd_update_positions(pd_context* ctx);
/* computeVelocity: bond breaking (when s0 > 0), force density f_i =
 * sum_j (T_ij - T_ji) V_j, then v += 0.5 (a + f/rho) dt.  write_acceleration
 * != 0 also stores a = f/rho, as the reference GPU kernel does
 * (include/mechanics.cuh:863); the CPU path does not. */
int pd_compute_velocity(pd_context* ctx, int write_acceleration);

/* ---- Mazars bond damage -----------------------------------------------------
 * computeDamageStatus (src/main_seq.cpp:473-598; include/mechanics.h:24):
 * per bond, the principal strains of eps_ik = sym(F_ik) - I give
 * eps_eq = sqrt(sum <l>+^2); past the bond's threshold kappa_ik the damage
 * d_ik = alpha_t d_t + alpha_c d_c is recomputed and kappa_ik = eps_eq;
 * damageStatus_i = sum_k d_ik / |H_i|.  A broken bond counts as d = 1.
 * Damage does not feed back into the stress (as in the reference,
 * src/main_seq.cpp:254).
 * pd_set_damage: mode PD_DAMAGE_OFF stops the updates; any other mode resets
 * kappa = kappa0 and d = 0.  While damage is on, every force evaluation
 * (pd_compute_velocity, pd_step) also updates it at the positions it sees,
 * fused into the bond kernel; pd_compute_damage updates it alone (no forces
 * needed, no bond breaking).  An update at unchanged positions changes
 * nothing, so the reference's order (damage at the start of a step, at the
 * previous step's positions) gives the same state as the fused update one
 * evaluation earlier. */
int pd_set_damage(pd_context* ctx, const pd_damage_params* d);
int pd_compute_damage(pd_context* ctx);
/* Per-bond damage and threshold (caller CSR order, nb each) and per-particle
 * damageStatus (caller order, n); NULL pointers are skipped. */
int pd_get_damage(pd_context* ctx, double* bond_damage, double* status, double* kappa);
/* Load per-bond damage and threshold (caller CSR order; NULL skips), e.g. the
 * caller-owned bondDamage / bondDamageThreshold of src/main_seq.cpp:643-676.
 * Damage must be on. */
int pd_set_damage_state(pd_context* ctx, const double* bond_damage, const double* kappa);

/* ---- whole steps ------------------------------------------------------------ */
/* Boundary program for pd_step: code[n] (caller order) with PD_BC_FREE,
 * PD_BC_FIXED or a velocity group g in [0, n_groups); v0[2*n_groups]. */
int pd_set_boundary(pd_context* ctx, const int* code, int n_groups, const double* v0);
/* nsteps explicit steps: velocity BC -> updatePositions -> fixed BC ->
 * bond breaking -> computeVelocity (src/main_seq.cpp:686-699 order). */
int pd_step(pd_context* ctx, int nsteps);
/* Times each launch-shape variant for `steps` steps on the context's own
 * state (restored afterwards) and keeps the fastest. */
int pd_autotune(pd_context* ctx, int steps);

/* ---- bonds and diagnostics -------------------------------------------------- */
/* Broken mask per directed bond in the caller's CSR order (nb bytes). */
int pd_get_broken(pd_context* ctx, unsigned char* mask);
int pd_set_broken(pd_context* ctx, const unsigned char* mask);
/* Force state T_ij per directed bond (double[2*nb], caller CSR order) from the
 * last force evaluation. */
int pd_get_bond_forces(pd_context* ctx, double* T);
int pd_get_stats(pd_context* ctx, pd_stats* out);
void pd_reset_profile(pd_context* ctx);

/* ---- mesh and result files in the reference's formats ------------------------
 * Text through a C++ stream as the reference writes it; digits <= 0 keeps the
 * stream default (6 significant digits, the reference's), 17 reproduces every
 * double.  ids: the particles' global ids in caller order (NULL: 0..n-1). */
/* readMeshFile (src/SetupParticleSystem.cpp:11-60): a header line, then
 * "x y area" per line.  reference_notch != 0 drops |x| > 17.5 && |y| < 2
 * (:33), the reference's two edge notches.  X (2*cap) and V (cap) may both be
 * NULL to count.  Returns the particle count, or -1 (pd_last_error(NULL)). */
long pd_read_mesh(const char* path, int reference_notch, double* X, double* V, long cap);
/* storeVelocity (include/util.h:20-39): "vx vy" per particle. */
int pd_write_velocity(pd_context* ctx, const char* path, int digits);
/* outputLocalParticlesPositions (src/SetupParticleSystem.cpp:646-658):
 * header "ID, x, y, ux, uy ", then "id,x,y,ux,uy" (initial x, y; u = y - X). */
int pd_write_positions(pd_context* ctx, const char* path, const int* ids, int digits);
/* outputGatheredPositions (src/SetupParticleSystem.cpp:552-627), the rank-0
 * file of a one-GPU run: "ID, x, y, ux, uy, damage" with damageStatus (0 when
 * damage was never enabled). */
int pd_write_gathered(pd_context* ctx, const char* path, const int* ids, int digits);

/* ---- checkpoint / resume (no counterpart in the reference) -------------------
 * Binary file of the state the next step reads: u, v, a, f, the broken mask,
 * the damage state and the step count, in the caller's order, with a hash of
 * the mesh.  A context of the same mesh and precision that loads it continues
 * bit for bit like the run that wrote it. */
int pd_save_checkpoint(pd_context* ctx, const char* path);
int pd_load_checkpoint(pd_context* ctx, const char* path);

/* ---- diagnostics (no counterpart in the reference) ---------------------------
 * Times a plain device-to-device copy kernel that reads and writes `bytes`
 * bytes, `iters` times on the current device; *gbps = 2*bytes*iters / time.
 * Gives the achievable HBM rate next to the roofline and a kernel of known
 * traffic to scale HBM counter readings (FETCH_SIZE / WRITE_SIZE) against. */
int pd_probe_copy(long bytes, int iters, double* gbps);

/* ---- allocation-free force evaluation (no counterpart in the reference) -----
 * For callers that cannot allocate device memory: code under HIP graph
 * capture, tight serving loops, harnesses that forbid hipMalloc.  One step of
 * pd_compute_velocity (bond breaking when s0 > 0, force density f, velocity
 * update v += 1/2 (a + f/rho) dt, and a = f/rho when write_acceleration),
 * PD_FP64 only, on a mesh and state held entirely in the caller's device
 * buffers, in the caller's particle order:
 *   X, u, v, a, f   double[2n] (xy interleaved, 16-byte aligned); V double[n];
 *   row_ptr int[n+1], col int[nb] (symmetric families, as pd_create takes);
 *   broken  uint8[nb], in/out (bonds that break are set to 1).
 * nb = row_ptr[n] and max_family (the largest family) are passed by the host.
 * Every intermediate (reverse bonds, K^-1, the force terms) lives in a fixed
 * per-GPU workspace inside the library (about 29 MiB, held until the process
 * ends), so the call allocates nothing; it works for n <= PD_STATIC_MAX_PARTICLES
 * and nb <= PD_STATIC_MAX_BONDS and rejects larger meshes.  Work is enqueued on
 * `stream` (hipStream_t, NULL = the null stream).  Outside stream capture the
 * call first checks the families on the device (symmetric, in range, no
 * self bond, no repeats) and waits for that check; under capture it does not
 * wait, and a mesh the call has not accepted before must not be passed.
 * Calls that overlap on one GPU share its workspace and race: serialise them. */
#define PD_STATIC_MAX_PARTICLES 131072
#define PD_STATIC_MAX_BONDS 524288
int pd_compute_velocity_static(const pd_params* p, int n, long nb, int max_family,
                               const double* X, const double* V, const int* row_ptr,
                               const int* col, unsigned char* broken, const double* u,
                               double* v, double* a, double* f, int write_acceleration,
                               void* stream);

/* ---- setup helpers (host side; no counterpart in the hot path) -------------- */
/* Families with the reference cutoff |xi| < 1.01*dx*horizon
 * (src/SetupParticleSystem.cpp:113) from a cell list, neighbours in ascending
 * id order like findNeighbor.  row_ptr[n+1]; col has room for cap entries.
 * Returns nb (>= 0) or -1 when cap is too small / bad input. */
long pd_build_families(int n, const double* X, double dx, double horizon,
                       int* row_ptr, int* col, long cap);

/* Pre-broken notch bonds: mask[b] = 1 when the segment X_i -> X_j of bond b
 * properly crosses one of nseg segments seg[4s..4s+3] = (x0, y0, x1, y1).
 * The reference makes notches by deleting particles
 * (src/SetupParticleSystem.cpp:33); the configs pre-break crossing bonds.
 * Symmetric in i, j.  Returns the number of marked directed bonds. */
long pd_notch_mask(int n, const double* X, const int* row_ptr, const int* col,
                   int nseg, const double* seg, unsigned char* mask);

#ifdef __cplusplus
}
#endif
#endif
