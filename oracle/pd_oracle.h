/*
 * pd_oracle.h -- CPU oracle for the bond-associated NOSB peridynamics step.
 *
 * TEST INFRASTRUCTURE ONLY.  Only tests/, __graft_entry__.smoke() and the
 * cpu_baseline leg of bench.py may load this library, and only as the checker
 * or the timed CPU baseline.  The product (lib2dpdgpu, include/pd_hip.h) never
 * links, loads or calls it.
 *
 * Two restatements of the reference algorithm live here:
 *
 *  ref_cpu_*   A literal restatement of src/main_seq.cpp (EricYJA/2D_PD_GPU):
 *              the same O(N*M^3) loop nest of computeVelocity
 *              (src/main_seq.cpp:396-470) -> computeForceDensityStates
 *              (:271-380) -> computeShapeTensors (:71-156) ->
 *              computeStressTensor (:204-268), in x87 long double with the
 *              reference's double/long-double promotions, and the matrix
 *              arithmetic of src/matrix.cpp:19-129 (unguarded inverse2D).
 *              Single particle range per call so it can be timed on a bounded
 *              sample.  Parity: pinned against output/velocity_step_0_after_
 *              seq_ref.txt (Hassa mesh, 6 printed digits) and against
 *              or_forces_fp64 on small meshes.
 *
 *  or_*        The restructured O(N*M^2) FP64 oracle in displacement form
 *              (u = y - X, bond term du = u_l - u_i).  K_ik and N_ik do not
 *              depend on j (src/main_seq.cpp:325), the weight is symmetric, so
 *              each family is evaluated once.  Adds the project's bond-breaking
 *              rule (critical stretch, both directions, FP64) and the
 *              scale-relative singular-K rule, which the reference lacks.
 *              Parity: pinned against ref_cpu_* (tiny meshes, all digits the
 *              formulations share), against the Hassa fixtures, and by the
 *              closed forms in tests/test_oracle.py (affine patch test, rigid
 *              motion, momentum balance, weight special cases, Verlet closed
 *              form).
 */
#ifndef PD_ORACLE_H
#define PD_ORACLE_H

#ifdef __cplusplus
extern "C" {
#endif

/* Material / discretisation parameters shared by every or_* routine. */
typedef struct {
    double n1, n2;     /* weight exponents (src/main_seq.cpp:609) */
    double horizon;    /* delta in units of dx                     */
    double dx;         /* grid spacing                              */
    double rho;        /* mass density                              */
    double dt;         /* step size                                 */
    double E;          /* Young's modulus (used when E_arr == NULL) */
    double nu;         /* Poisson ratio                             */
    double s0;         /* critical stretch; <= 0 disables breaking  */
    double kappa_tol;  /* singular K: det(K) <= kappa_tol*tr(K)^2   */
} or_params;

/* Weight omega_i(j,k) of src/main_seq.cpp:103-118 for bonds xj, xk (initial
 * bond vectors from the family centre). */
double or_weight(const or_params* p, double xj0, double xj1, double xk0, double xk1);

/* Reverse-bond index: rev[b] is the bond (j->i) for b = (i->j).  Returns 0 on
 * success, -1 when some bond has no reverse (asymmetric families). */
int or_build_rev(int n, const int* row, const int* col, int* rev);

/* Force density f_i = sum_j (T_ij - T_ji) V_j for displacement u (2n).
 * broken: per-bond flag (may be NULL).  E_arr: per-particle E (may be NULL).
 * T_out (2*nb, may be NULL) receives T_ij per bond.  Kinv_out (3*nb, may be
 * NULL) receives the symmetric K_ik^-1 (xx, xy, yy).
 * bugcompat != 0 reproduces the reference GPU kernel's current-bond vector
 * y_l - X_i (include/mechanics.cuh:165): du_l := u_l instead of u_l - u_i.
 * Returns the number of bonds whose K was treated as singular. */
long or_forces_fp64(const or_params* p, int n, const double* X, const double* V,
                    const double* E_arr, const int* row, const int* col,
                    const int* rev, const unsigned char* broken, const double* u,
                    int bugcompat, double* f_out, double* T_out, double* Kinv_out,
                    int nthreads);

/* Stretch of bond (i->j): (|xi + du| - |xi|) / |xi|, evaluated in FP64 with
 * no fused operations, identical for both directions of a bond. */
double or_bond_stretch(double xi0, double xi1, double du0, double du1);

/* Breaking test of the project rule s > s0, evaluated as
 * |xi + du|^2 > c |xi|^2 with c = (1 + s0)^2, every operation rounded on its
 * own (identical bits for (i->j) and (j->i)). */
int or_bond_breaks(double xi0, double xi1, double du0, double du1, double c);

/* Mark bonds whose stretch exceeds s0 as broken (irreversible).  Returns the
 * number of newly broken directed bonds. */
long or_break_bonds(const or_params* p, int n, const double* X, const int* row,
                    const int* col, const double* u, unsigned char* broken);

/* Full explicit steps (project step order, reference semantics for the BC
 * handling of src/main_seq.cpp:680-699):
 *   velocity BC -> a = f/rho; u += v dt + 0.5 a dt^2 -> fixed BC ->
 *   bond breaking -> forces -> v += 0.5 (a + f/rho) dt.
 * bc[i]: -1 free, -2 fixed (u reset to 0), g >= 0 velocity group with
 * prescribed velocity v0[2g..2g+1]. */
long or_step_fp64(const or_params* p, int n, const double* X, const double* V,
                  const double* E_arr, const int* row, const int* col, const int* rev,
                  const int* bc, const double* v0, double* u, double* v, double* a,
                  double* f, unsigned char* broken, int nsteps, int nthreads);

/* ---------------- Mazars bond damage (computeDamageStatus) ----------------- *
 * Restatement of src/main_seq.cpp:473-598.  Per bond (i, k): eps = sym(F) - I
 * (here sym(F - I) in displacement form), principal strains l1 <= l2 of the
 * 2x2 symmetric eps (closed form: the reference calls Eigen's
 * SelfAdjointEigenSolver, :523-526, Eigen 3.x, absent from this image),
 * eps_eq = sqrt(sum <l>+^2) (:529-536).  When eps_eq > kappa_ik (:540):
 *   tension part  eps_t = sum_{l>0} l v v^T, compression part eps_c =
 *   sum_{l<0} l v v^T (:546-561; zero-initialised here: the reference leaves
 *   strain_t / strain_c uninitialised, :509, :513), their ascending spectra
 *   t_i, c_i (:564-568), alpha_t = sum_{t_i>0} t_i (t_i + c_i) / eps_eq^2,
 *   alpha_c = sum_{c_i>0} c_i (t_i + c_i) / eps_eq^2 (:571-578),
 *   d_t = 1 - kr (1 - A_t) / eps_eq - A_t / exp(B_t (eps_eq - kr)),
 *   d_c likewise (:581-582), d = alpha_t d_t + alpha_c d_c, kappa_ik = eps_eq.
 * kr is the running threshold kappa_ik in mode OR_DAMAGE_REFERENCE (what
 * :579-582 compute) and the initial kappa0 in mode OR_DAMAGE_MAZARS (the
 * published model).  damageStatus_i = sum_k d_ik / |H_i| (:593).
 * Project rule for bond breaking (which the reference lacks): a broken bond
 * has d = 1 and is not updated.
 * Parity unpinned against the reference: no reference file holds damage
 * values.  Pinned by closed forms in tests/test_oracle.py (principal strains
 * of rotated tensors against numpy.linalg.eigvalsh, uniaxial / biaxial /
 * mixed-sign eps_eq, idempotence at fixed strain, the affine patch test). */
#define OR_DAMAGE_REFERENCE 1
#define OR_DAMAGE_MAZARS    2

typedef struct {
    int mode;          /* OR_DAMAGE_REFERENCE or OR_DAMAGE_MAZARS */
    double kappa0;     /* initial threshold (reference: tensileStrength / E, :660) */
    double At, Bt, Ac, Bc;  /* :480: 0.87, 20000, 0.65, 2150 */
} or_damage_params;

/* Principal values l1 <= l2 of [[a, b], [b, c]]. */
void or_sym2_eig(double a, double b, double c, double* l1, double* l2);

/* One bond's update for strain (e11, e22, e12); returns 1 when eps_eq > kappa
 * (kappa and d updated), else 0. */
int or_mazars_update(const or_damage_params* dp, double e11, double e22, double e12,
                     double* kappa, double* d);

/* or_forces_fp64 plus the damage update of every bond at displacement u:
 * kappa, dmg (nb each) updated in place, status (n) = sum_k d_ik / |H_i|.
 * Returns the number of bonds whose damage changed. */
long or_forces_damage_fp64(const or_params* p, const or_damage_params* dp, int n,
                           const double* X, const double* V, const double* E_arr,
                           const int* row, const int* col, const int* rev,
                           const unsigned char* broken, const double* u, double* f_out,
                           double* T_out, double* kappa, double* dmg, double* status,
                           int nthreads);

/* or_step_fp64 with the damage update after each step's force evaluation (at
 * the same positions). */
long or_step_damage_fp64(const or_params* p, const or_damage_params* dp, int n,
                         const double* X, const double* V, const double* E_arr,
                         const int* row, const int* col, const int* rev, const int* bc,
                         const double* v0, double* u, double* v, double* a, double* f,
                         unsigned char* broken, double* kappa, double* dmg, double* status,
                         int nsteps, int nthreads);

/* ---------------- literal long-double restatement of main_seq -------------- */

/* Parameters exactly as the reference passes them: n1, n2, horizon, dx and
 * E, nu are double at the call sites (src/main_seq.cpp:396, :180); massDensity
 * and stepSize are long double. */
typedef struct {
    double n1, n2, horizon, dx;
    double E, nu;
    long double rho, dt;
} ref_params;

/* computeVelocity (src/main_seq.cpp:396-470) for particles [i0, i1):
 * X, y: 2n initial / current positions; V: n volumes; families as CSR in the
 * reference's neighbour order.  Updates v[2i..] and f[2i..] (netF) for
 * i in [i0, i1); reads a.  nthreads > 1 splits the range (results identical,
 * each particle is independent). */
void ref_cpu_compute_velocity(const ref_params* p, int n, const long double* X,
                              const long double* y, const long double* V,
                              const int* row, const int* col, long double* v,
                              const long double* a, long double* f, int i0, int i1,
                              int nthreads);

/* updatePositions (src/main_seq.cpp:386-393). */
void ref_cpu_update_positions(const ref_params* p, int n, long double* y,
                              const long double* v, long double* a,
                              const long double* f);

/* applyVelocityBC (src/main_seq.cpp:28-36). */
void ref_cpu_apply_velocity_bc(int cnt, const int* ids, const long double* v0,
                               long double* v, long double* a);

/* applyFixedBC (src/main_seq.cpp:38-46). */
void ref_cpu_apply_fixed_bc(int cnt, const int* ids, const long double* X,
                            long double* y);

/* findNeighbor (src/SetupParticleSystem.cpp:100-136), all-pairs in long
 * double with the cutoff |xi| < 1.01*dx*horizon.  Writes CSR (row: n+1,
 * col: capacity cap) in ascending neighbour order.  Returns nb, or -1 when
 * cap is too small. */
long ref_find_neighbor(int n, const long double* X, long double dx,
                       long double horizon, int* row, int* col, long cap);

/* 2x2 long-double matrix helpers of src/matrix.cpp, exposed so that they can
 * be checked bit-for-bit against the reference's own matrix.cpp. */
void ref_mat_inverse2d(const long double* a, long double* out);
void ref_mat_mul(const long double* a, const long double* b, long double* out);

#ifdef __cplusplus
}
#endif
#endif
