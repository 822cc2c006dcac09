/*
 * pd_oracle.c -- CPU oracle for the bond-associated NOSB peridynamics step.
 *
 * TEST INFRASTRUCTURE ONLY (see pd_oracle.h).  Built with plain gcc -O2 and no
 * -march flags, so double arithmetic is SSE2 with no fused multiply-add, as in
 * a default build of the reference.
 */
#include "pd_oracle.h"

#include <math.h>
#include <stdlib.h>
#include <string.h>
#ifdef _OPENMP
#include <omp.h>
#endif

/* ------------------------------------------------------------------------ */
/* Weight omega_i(j,k), src/main_seq.cpp:103-118 (also :300-314).            */
/* ------------------------------------------------------------------------ */
double or_weight(const or_params* p, double xj0, double xj1, double xk0, double xk1)
{
    double lj = sqrt(xj0 * xj0 + xj1 * xj1);
    double lk = sqrt(xk0 * xk0 + xk1 * xk1);
    double num = xj0 * xk0 + xj1 * xk1;
    double c = num / (lj * lk);
    if (c > 1.0) c = 1.0; else if (c < -1.0) c = -1.0;
    double ratio = fabs(lj - lk) / (p->horizon * p->dx);
    return exp(-p->n1 * ratio) * pow(0.5 + 0.5 * c, p->n2);
}

int or_build_rev(int n, const int* row, const int* col, int* rev)
{
    int ok = 0;
    for (int i = 0; i < n; ++i) {
        for (int b = row[i]; b < row[i + 1]; ++b) {
            int j = col[b];
            int r = -1;
            for (int c = row[j]; c < row[j + 1]; ++c)
                if (col[c] == i) { r = c; break; }
            rev[b] = r;
            if (r < 0) ok = -1;
        }
    }
    return ok;
}

/* Stretch with every operation rounded separately (no contraction), so that
 * (i->j) and (j->i) give the same bits: xi and du flip sign exactly. */
double or_bond_stretch(double xi0, double xi1, double du0, double du1)
{
    volatile double e0 = xi0 + du0;
    volatile double e1 = xi1 + du1;
    volatile double p0 = xi0 * xi0;
    volatile double p1 = xi1 * xi1;
    volatile double q0 = e0 * e0;
    volatile double q1 = e1 * e1;
    double len = sqrt(p0 + p1);
    double cur = sqrt(q0 + q1);
    return (cur - len) / len;
}

/* s > s0  <=>  |xi + du|^2 > (1 + s0)^2 |xi|^2 (both sides positive); evaluated
 * with every operation rounded on its own and c = (1 + s0)^2 formed once, so
 * both directions of a bond agree bit for bit. */
int or_bond_breaks(double xi0, double xi1, double du0, double du1, double c)
{
    volatile double e0 = xi0 + du0;
    volatile double e1 = xi1 + du1;
    volatile double q0 = e0 * e0;
    volatile double q1 = e1 * e1;
    volatile double p0 = xi0 * xi0;
    volatile double p1 = xi1 * xi1;
    volatile double cur2 = q0 + q1;
    volatile double len2 = p0 + p1;
    volatile double lim = c * len2;
    return cur2 > lim;
}

long or_break_bonds(const or_params* p, int n, const double* X, const int* row,
                    const int* col, const double* u, unsigned char* broken)
{
    long nnew = 0;
    if (!(p->s0 > 0.0)) return 0;
    const double c = (1.0 + p->s0) * (1.0 + p->s0);
    for (int i = 0; i < n; ++i) {
        for (int b = row[i]; b < row[i + 1]; ++b) {
            if (broken[b]) continue;
            int j = col[b];
            if (or_bond_breaks(X[2 * j] - X[2 * i], X[2 * j + 1] - X[2 * i + 1],
                               u[2 * j] - u[2 * i], u[2 * j + 1] - u[2 * i + 1], c)) {
                broken[b] = 1;
                ++nnew;
            }
        }
    }
    return nnew;
}

/* ------------------------------------------------------------------------ */
/* FP64 displacement-form forces.                                            */
/*   K_ik = sum_l w(k,l) V_l xi_l (x) xi_l                                   */
/*   D_ik = sum_l w(k,l) V_l du_l (x) xi_l          (N_ik = K_ik + D_ik)      */
/*   F_ik - I = D_ik K_ik^-1,  eps = sym(F - I),  sigma = C eps              */
/*   T_ij = (1/V_H) [sum_k w(j,k) V_k sigma_ik K_ik^-1] xi_ij                */
/*   f_i  = sum_j (T_ij - T_ji) V_j                                          */
/* following src/main_seq.cpp:71-156, :204-268, :271-380, :396-470.         */
/* ------------------------------------------------------------------------ */
typedef struct {
    double* xi;   /* 2m */
    double* len;  /* m  */
    double* du;   /* 2m */
    double* Vl;   /* m  */
    double* w;    /* m*m */
    double* Q;    /* 4m, V_k sigma_k K_k^-1 */
    unsigned char* alive;
    int cap;
} fam_ws;

static void ws_reserve(fam_ws* ws, int m)
{
    if (m <= ws->cap) return;
    free(ws->xi); free(ws->len); free(ws->du); free(ws->Vl); free(ws->w); free(ws->Q);
    free(ws->alive);
    ws->xi = (double*)malloc(sizeof(double) * 2 * m);
    ws->len = (double*)malloc(sizeof(double) * m);
    ws->du = (double*)malloc(sizeof(double) * 2 * m);
    ws->Vl = (double*)malloc(sizeof(double) * m);
    ws->w = (double*)malloc(sizeof(double) * (size_t)m * m);
    ws->Q = (double*)malloc(sizeof(double) * 4 * m);
    ws->alive = (unsigned char*)malloc(m);
    ws->cap = m;
}

static void ws_free(fam_ws* ws)
{
    free(ws->xi); free(ws->len); free(ws->du); free(ws->Vl); free(ws->w); free(ws->Q);
    free(ws->alive);
}

/* Plane-stress stiffness of src/main_seq.cpp:180-201 applied to the Voigt
 * strain [e11, e22, e12] (tensor shear), giving sigma12 = (1-nu) pf e12. */
static void plane_stress(double E, double nu, double e11, double e22, double e12,
                         double* s11, double* s22, double* s12)
{
    double pf = E / (1.0 - nu * nu);
    *s11 = pf * e11 + pf * nu * e22;
    *s22 = pf * nu * e11 + pf * e22;
    *s12 = pf * (1.0 - nu) * e12;
}

/* ------------------------------------------------------------------------ */
/* Mazars damage, src/main_seq.cpp:473-598 (see pd_oracle.h).                */
/* ------------------------------------------------------------------------ */
void or_sym2_eig(double a, double b, double c, double* l1, double* l2)
{
    double m = 0.5 * (a + c);
    double h = 0.5 * (a - c);
    double r = sqrt(h * h + b * b);
    *l1 = m - r;
    *l2 = m + r;
}

int or_mazars_update(const or_damage_params* dp, double e11, double e22, double e12,
                     double* kappa, double* d)
{
    double l[2];
    or_sym2_eig(e11, e12, e22, &l[0], &l[1]);
    double s2 = 0.0, eq = 0.0;
    for (int i = 0; i < 2; ++i)
        if (l[i] > 0.0) s2 += l[i] * l[i];
    if (s2 > 0.0) eq = sqrt(s2);
    if (!(eq > *kappa)) return 0;
    /* eps_t and eps_c share eps's principal directions: their ascending
     * spectra are (<l1>+, <l2>+) and (<l1>-, <l2>-). */
    double t[2], c[2];
    for (int i = 0; i < 2; ++i) {
        t[i] = l[i] > 0.0 ? l[i] : 0.0;
        c[i] = l[i] < 0.0 ? l[i] : 0.0;
    }
    double at = 0.0, ac = 0.0;
    for (int i = 0; i < 2; ++i) {
        if (t[i] > 0.0) at += t[i] * (t[i] + c[i]) / (eq * eq);
        if (c[i] > 0.0) ac += c[i] * (t[i] + c[i]) / (eq * eq);
    }
    double kr = dp->mode == OR_DAMAGE_MAZARS ? dp->kappa0 : *kappa;
    double dt = 1.0 - kr * (1.0 - dp->At) / eq - dp->At / exp(dp->Bt * (eq - kr));
    double dc = 1.0 - kr * (1.0 - dp->Ac) / eq - dp->Ac / exp(dp->Bc * (eq - kr));
    *d = at * dt + ac * dc;
    *kappa = eq;
    return 1;
}

typedef struct {
    const or_damage_params* dp;
    double* kappa;
    double* dmg;
    double* status;
    long changed;
} dmg_ws;

static long family_T(const or_params* p, int i, const double* X, const double* V,
                     const double* E_arr, const int* row, const int* col,
                     const unsigned char* broken, const double* u, int bugcompat,
                     double* T_out, double* Kinv_out, fam_ws* ws, dmg_ws* dw)
{
    int beg = row[i];
    int m = row[i + 1] - beg;
    long nsing = 0;
    ws_reserve(ws, m > 0 ? m : 1);
    double E = E_arr ? E_arr[i] : p->E;
    for (int b = 0; b < m; ++b) {
        int j = col[beg + b];
        ws->xi[2 * b] = X[2 * j] - X[2 * i];
        ws->xi[2 * b + 1] = X[2 * j + 1] - X[2 * i + 1];
        ws->len[b] = sqrt(ws->xi[2 * b] * ws->xi[2 * b] + ws->xi[2 * b + 1] * ws->xi[2 * b + 1]);
        if (bugcompat) {
            ws->du[2 * b] = u[2 * j];
            ws->du[2 * b + 1] = u[2 * j + 1];
        } else {
            ws->du[2 * b] = u[2 * j] - u[2 * i];
            ws->du[2 * b + 1] = u[2 * j + 1] - u[2 * i + 1];
        }
        ws->Vl[b] = V[j];
        ws->alive[b] = broken ? !broken[beg + b] : 1;
    }
    /* symmetric weight table */
    for (int k = 0; k < m; ++k)
        for (int l = 0; l < m; ++l)
            ws->w[(size_t)k * m + l] = or_weight(p, ws->xi[2 * k], ws->xi[2 * k + 1],
                                                 ws->xi[2 * l], ws->xi[2 * l + 1]);
    double VH = 0.0;
    for (int k = 0; k < m; ++k) {
        double* Q = ws->Q + 4 * k;
        Q[0] = Q[1] = Q[2] = Q[3] = 0.0;
        if (Kinv_out) {
            Kinv_out[3 * (size_t)(beg + k)] = 0.0;
            Kinv_out[3 * (size_t)(beg + k) + 1] = 0.0;
            Kinv_out[3 * (size_t)(beg + k) + 2] = 0.0;
        }
        if (!ws->alive[k]) {
            if (dw && dw->dmg[beg + k] != 1.0) {  /* broken: fully damaged */
                dw->dmg[beg + k] = 1.0;
                dw->changed++;
            }
            continue;
        }
        VH += ws->Vl[k];
        double K00 = 0, K01 = 0, K11 = 0, D00 = 0, D01 = 0, D10 = 0, D11 = 0;
        for (int l = 0; l < m; ++l) {
            if (!ws->alive[l]) continue;
            double wv = ws->w[(size_t)k * m + l] * ws->Vl[l];
            double x0 = ws->xi[2 * l], x1 = ws->xi[2 * l + 1];
            double d0 = ws->du[2 * l], d1 = ws->du[2 * l + 1];
            K00 += wv * x0 * x0;
            K01 += wv * x0 * x1;
            K11 += wv * x1 * x1;
            D00 += wv * d0 * x0;
            D01 += wv * d0 * x1;
            D10 += wv * d1 * x0;
            D11 += wv * d1 * x1;
        }
        double det = K00 * K11 - K01 * K01;
        double tr = K00 + K11;
        if (!(det > p->kappa_tol * tr * tr)) { ++nsing; continue; }
        double i00 = K11 / det, i01 = -K01 / det, i11 = K00 / det;
        if (Kinv_out) {
            Kinv_out[3 * (size_t)(beg + k)] = i00;
            Kinv_out[3 * (size_t)(beg + k) + 1] = i01;
            Kinv_out[3 * (size_t)(beg + k) + 2] = i11;
        }
        double e11, e22, e12;
        if (bugcompat) {
            /* position form of the reference GPU kernel: N = K + D, F = N K^-1,
             * eps = sym(F) - I (include/mechanics.cuh:232-268) */
            double N00 = K00 + D00, N01 = K01 + D01, N10 = K01 + D10, N11 = K11 + D11;
            double F00 = N00 * i00 + N01 * i01, F01 = N00 * i01 + N01 * i11;
            double F10 = N10 * i00 + N11 * i01, F11 = N10 * i01 + N11 * i11;
            e11 = 0.5 * (F00 + F00) - 1.0;
            e22 = 0.5 * (F11 + F11) - 1.0;
            e12 = 0.5 * (F01 + F10);
        } else {
            double G00 = D00 * i00 + D01 * i01, G01 = D00 * i01 + D01 * i11;
            double G10 = D10 * i00 + D11 * i01, G11 = D10 * i01 + D11 * i11;
            e11 = G00;
            e22 = G11;
            e12 = 0.5 * (G01 + G10);
        }
        if (dw)
            dw->changed += or_mazars_update(dw->dp, e11, e22, e12, dw->kappa + beg + k,
                                            dw->dmg + beg + k);
        double s11, s22, s12;
        plane_stress(E, p->nu, e11, e22, e12, &s11, &s22, &s12);
        double Vk = ws->Vl[k];
        Q[0] = (s11 * i00 + s12 * i01) * Vk;
        Q[1] = (s11 * i01 + s12 * i11) * Vk;
        Q[2] = (s12 * i00 + s22 * i01) * Vk;
        Q[3] = (s12 * i01 + s22 * i11) * Vk;
    }
    for (int jj = 0; jj < m; ++jj) {
        double t0 = 0.0, t1 = 0.0;
        if (ws->alive[jj] && VH > 0.0) {
            double S00 = 0, S01 = 0, S10 = 0, S11 = 0;
            for (int k = 0; k < m; ++k) {
                double w = ws->w[(size_t)jj * m + k];
                const double* Q = ws->Q + 4 * k;
                S00 += w * Q[0];
                S01 += w * Q[1];
                S10 += w * Q[2];
                S11 += w * Q[3];
            }
            double x0 = ws->xi[2 * jj], x1 = ws->xi[2 * jj + 1];
            t0 = (S00 * x0 + S01 * x1) / VH;
            t1 = (S10 * x0 + S11 * x1) / VH;
        }
        T_out[2 * (size_t)(beg + jj)] = t0;
        T_out[2 * (size_t)(beg + jj) + 1] = t1;
    }
    if (dw) {
        double sum = 0.0;
        for (int k = 0; k < m; ++k) sum += dw->dmg[beg + k];
        dw->status[i] = m > 0 ? sum / m : 0.0;
    }
    return nsing;
}

static long forces_impl(const or_params* p, int n, const double* X, const double* V,
                        const double* E_arr, const int* row, const int* col,
                        const int* rev, const unsigned char* broken, const double* u,
                        int bugcompat, double* f_out, double* T_out, double* Kinv_out,
                        const or_damage_params* dp, double* kappa, double* dmg,
                        double* status, long* changed, int nthreads);

long or_forces_fp64(const or_params* p, int n, const double* X, const double* V,
                    const double* E_arr, const int* row, const int* col,
                    const int* rev, const unsigned char* broken, const double* u,
                    int bugcompat, double* f_out, double* T_out, double* Kinv_out,
                    int nthreads)
{
    return forces_impl(p, n, X, V, E_arr, row, col, rev, broken, u, bugcompat, f_out, T_out,
                       Kinv_out, NULL, NULL, NULL, NULL, NULL, nthreads);
}

long or_forces_damage_fp64(const or_params* p, const or_damage_params* dp, int n,
                           const double* X, const double* V, const double* E_arr,
                           const int* row, const int* col, const int* rev,
                           const unsigned char* broken, const double* u, double* f_out,
                           double* T_out, double* kappa, double* dmg, double* status,
                           int nthreads)
{
    long changed = 0;
    forces_impl(p, n, X, V, E_arr, row, col, rev, broken, u, 0, f_out, T_out, NULL, dp, kappa,
                dmg, status, &changed, nthreads);
    return changed;
}

static long forces_impl(const or_params* p, int n, const double* X, const double* V,
                        const double* E_arr, const int* row, const int* col,
                        const int* rev, const unsigned char* broken, const double* u,
                        int bugcompat, double* f_out, double* T_out, double* Kinv_out,
                        const or_damage_params* dp, double* kappa, double* dmg,
                        double* status, long* changed, int nthreads)
{
    long nb = row[n];
    double* T = T_out ? T_out : (double*)malloc(sizeof(double) * 2 * (size_t)(nb > 0 ? nb : 1));
    int* rv = (int*)rev;
    if (!rv) {
        rv = (int*)malloc(sizeof(int) * (size_t)(nb > 0 ? nb : 1));
        or_build_rev(n, row, col, rv);
    }
    long nsing = 0, nchg = 0;
#ifdef _OPENMP
    if (nthreads < 1) nthreads = 1;
#pragma omp parallel num_threads(nthreads) reduction(+ : nsing, nchg)
#endif
    {
        fam_ws ws;
        memset(&ws, 0, sizeof(ws));
        dmg_ws dw = {dp, kappa, dmg, status, 0};
#ifdef _OPENMP
#pragma omp for schedule(dynamic, 64)
#endif
        for (int i = 0; i < n; ++i)
            nsing += family_T(p, i, X, V, E_arr, row, col, broken, u, bugcompat, T, Kinv_out, &ws,
                              dp ? &dw : NULL);
        nchg += dw.changed;
        ws_free(&ws);
    }
    if (changed) *changed = nchg;
#ifdef _OPENMP
#pragma omp parallel for num_threads(nthreads) schedule(static)
#endif
    for (int i = 0; i < n; ++i) {
        double f0 = 0.0, f1 = 0.0;
        for (int b = row[i]; b < row[i + 1]; ++b) {
            int r = rv[b];
            double Vj = V[col[b]];
            double tr0 = r >= 0 ? T[2 * (size_t)r] : 0.0;
            double tr1 = r >= 0 ? T[2 * (size_t)r + 1] : 0.0;
            f0 += (T[2 * (size_t)b] - tr0) * Vj;
            f1 += (T[2 * (size_t)b + 1] - tr1) * Vj;
        }
        f_out[2 * i] = f0;
        f_out[2 * i + 1] = f1;
    }
    if (!T_out) free(T);
    if (rv != rev) free(rv);
    return nsing;
}

long or_step_fp64(const or_params* p, int n, const double* X, const double* V,
                  const double* E_arr, const int* row, const int* col, const int* rev,
                  const int* bc, const double* v0, double* u, double* v, double* a,
                  double* f, unsigned char* broken, int nsteps, int nthreads)
{
    return or_step_damage_fp64(p, NULL, n, X, V, E_arr, row, col, rev, bc, v0, u, v, a, f,
                               broken, NULL, NULL, NULL, nsteps, nthreads);
}

long or_step_damage_fp64(const or_params* p, const or_damage_params* dp, int n,
                         const double* X, const double* V, const double* E_arr,
                         const int* row, const int* col, const int* rev, const int* bc,
                         const double* v0, double* u, double* v, double* a, double* f,
                         unsigned char* broken, double* kappa, double* dmg, double* status,
                         int nsteps, int nthreads)
{
    long nb = row[n];
    double* fn = (double*)malloc(sizeof(double) * 2 * (size_t)(n > 0 ? n : 1));
    double* T = (double*)malloc(sizeof(double) * 2 * (size_t)(nb > 0 ? nb : 1));
    long nbroken = 0;
    double dt = p->dt, rho = p->rho;
    double dt2 = dt * dt;
    for (int s = 0; s < nsteps; ++s) {
        for (int i = 0; i < n; ++i) {
            int g = bc ? bc[i] : -1;
            if (g >= 0) {
                v[2 * i] = v0[2 * g];
                v[2 * i + 1] = v0[2 * g + 1];
            }
            for (int d = 0; d < 2; ++d) {
                a[2 * i + d] = f[2 * i + d] / rho;
                volatile double t1 = v[2 * i + d] * dt;
                volatile double h = 0.5 * a[2 * i + d];
                volatile double t2 = h * dt2;
                volatile double inc = t1 + t2;
                u[2 * i + d] = u[2 * i + d] + inc;
            }
            if (g == -2) { u[2 * i] = 0.0; u[2 * i + 1] = 0.0; }
        }
        if (broken) nbroken += or_break_bonds(p, n, X, row, col, u, broken);
        forces_impl(p, n, X, V, E_arr, row, col, rev, broken, u, 0, fn, T, NULL, dp, kappa, dmg,
                    status, NULL, nthreads);
        for (int i = 0; i < n; ++i) {
            for (int d = 0; d < 2; ++d) {
                double an = fn[2 * i + d] / rho;
                volatile double s2 = a[2 * i + d] + an;
                volatile double h = 0.5 * s2;
                volatile double inc = h * dt;
                v[2 * i + d] = v[2 * i + d] + inc;
                f[2 * i + d] = fn[2 * i + d];
            }
        }
    }
    free(fn);
    free(T);
    return nbroken;
}

/* ------------------------------------------------------------------------ */
/* Literal long-double restatement of src/main_seq.cpp.                      */
/* ------------------------------------------------------------------------ */

/* matrix::inverse2D, src/matrix.cpp:116-129 (no singularity guard). */
void ref_mat_inverse2d(const long double* e, long double* inv)
{
    long double det = e[0] * e[3] - e[1] * e[2];
    inv[0] = e[3] / det;
    inv[1] = -e[1] / det;
    inv[2] = -e[2] / det;
    inv[3] = e[0] / det;
}

/* matrix::timeMatrix, src/matrix.cpp:69-83. */
void ref_mat_mul(const long double* a, const long double* b, long double* out)
{
    long double r[4] = {0.0L, 0.0L, 0.0L, 0.0L};
    for (int i = 0; i < 2; ++i)
        for (int j = 0; j < 2; ++j)
            for (int k = 0; k < 2; ++k)
                r[2 * i + j] += a[2 * i + k] * b[2 * k + j];
    memcpy(out, r, sizeof(r));
}

typedef struct {
    const ref_params* p;
    const long double* X;
    const long double* y;
    const long double* V;
    const int* row;
    const int* col;
    long double C[9]; /* getStiffnessTensor, src/main_seq.cpp:180-201 */
} ref_ctx;

/* Bond quantities shared by computeShapeTensors and computeForceDensityStates:
 * length2 is double accumulating long double squares (pow(long double, 2)),
 * numerator is double accumulating long double products. */
static double ref_weight(const ref_ctx* c, const long double* bij, double length,
                         const long double* bnb, double* lengthNb_out)
{
    double lengthNb2 = 0.0, numerator = 0.0;
    for (int d = 0; d < 2; ++d) {
        lengthNb2 = (double)((long double)lengthNb2 + bnb[d] * bnb[d]);
        numerator = (double)((long double)numerator + bij[d] * bnb[d]);
    }
    double lengthNb = sqrt(lengthNb2);
    double cosAngle = numerator / (length * lengthNb);
    if (cosAngle > 1.0) cosAngle = 1.0; else if (cosAngle < -1.0) cosAngle = -1.0;
    double lengthRatio = fabs(length - lengthNb) / (c->p->horizon * c->p->dx);
    if (lengthNb_out) *lengthNb_out = lengthNb;
    return exp(-c->p->n1 * lengthRatio) * pow(0.5 + 0.5 * cosAngle, c->p->n2);
}

static double ref_length(const long double* b)
{
    double length2 = 0.0;
    for (int d = 0; d < 2; ++d) length2 = (double)((long double)length2 + b[d] * b[d]);
    return sqrt(length2);
}

/* computeShapeTensors, src/main_seq.cpp:71-156. */
static void ref_shape(const ref_ctx* c, int beg, int m, int pi, int pj, long double* K,
                      long double* N)
{
    long double bij[2];
    for (int d = 0; d < 2; ++d) bij[d] = c->X[2 * pj + d] - c->X[2 * pi + d];
    double length = ref_length(bij);
    for (int q = 0; q < 4; ++q) { K[q] = 0.0L; N[q] = 0.0L; }
    for (int t = 0; t < m; ++t) {
        int nb = c->col[beg + t];
        long double bnb[2], bcur[2];
        for (int d = 0; d < 2; ++d) {
            bnb[d] = c->X[2 * nb + d] - c->X[2 * pi + d];
            bcur[d] = c->y[2 * nb + d] - c->y[2 * pi + d];
        }
        double weight = ref_weight(c, bij, length, bnb, NULL);
        long double vol = c->V[nb];
        for (int k = 0; k < 2; ++k)
            for (int l = 0; l < 2; ++l) {
                K[2 * k + l] += (long double)weight * bnb[k] * bnb[l] * vol;
                N[2 * k + l] += (long double)weight * bcur[k] * bnb[l] * vol;
            }
    }
}

/* computeStressTensor, src/main_seq.cpp:204-268 (d = 0). */
static void ref_stress(const ref_ctx* c, const long double* K, const long double* N,
                       long double* S)
{
    long double Kinv[4], F[4], st[4];
    ref_mat_inverse2d(K, Kinv);
    ref_mat_mul(N, Kinv, F);
    /* (F^T).matrixAdd(F), timeScalar(0.5), matrixSub(I) */
    for (int i = 0; i < 2; ++i)
        for (int j = 0; j < 2; ++j) {
            long double s = F[2 * j + i] + F[2 * i + j];
            s = s * 0.5L;
            st[2 * i + j] = s - (i == j ? 1.0L : 0.0L);
        }
    long double sv[3] = {st[0], st[3], st[1]};
    long double out[3];
    for (int i = 0; i < 3; ++i) {
        long double r = 0.0L;
        for (int j = 0; j < 3; ++j) r += c->C[3 * i + j] * sv[j];
        out[i] = r;
    }
    S[0] = out[0];
    S[1] = out[2];
    S[2] = out[2];
    S[3] = out[1];
}

/* computeForceDensityStates, src/main_seq.cpp:271-380: T_ij over H_i. */
static void ref_force_state(const ref_ctx* c, int pi, int pj, long double* Tv)
{
    int beg = c->row[pi], m = c->row[pi + 1] - beg;
    long double Tm[4] = {0.0L, 0.0L, 0.0L, 0.0L};
    long double bij[2];
    for (int d = 0; d < 2; ++d) bij[d] = c->X[2 * pj + d] - c->X[2 * pi + d];
    double length = ref_length(bij);
    double horizonVolume = 0.0;
    for (int t = 0; t < m; ++t) {
        int nb = c->col[beg + t];
        long double bnb[2];
        for (int d = 0; d < 2; ++d) bnb[d] = c->X[2 * nb + d] - c->X[2 * pi + d];
        double weight = ref_weight(c, bij, length, bnb, NULL);
        long double K[4], N[4], S[4], Kinv[4], SK[4];
        ref_shape(c, beg, m, pi, nb, K, N);
        ref_stress(c, K, N, S);
        ref_mat_inverse2d(K, Kinv);
        ref_mat_mul(S, Kinv, SK);
        long double sc = (long double)weight * c->V[nb];
        for (int q = 0; q < 4; ++q) Tm[q] = Tm[q] + SK[q] * sc;
        horizonVolume = (double)((long double)horizonVolume + c->V[nb]);
    }
    long double inv = (long double)(1.0 / horizonVolume);
    for (int i = 0; i < 2; ++i) {
        long double r = 0.0L;
        for (int j = 0; j < 2; ++j) r += (Tm[2 * i + j] * inv) * bij[j];
        Tv[i] = r;
    }
}

void ref_cpu_compute_velocity(const ref_params* p, int n, const long double* X,
                              const long double* y, const long double* V,
                              const int* row, const int* col, long double* v,
                              const long double* a, long double* f, int i0, int i1,
                              int nthreads)
{
    ref_ctx c;
    c.p = p; c.X = X; c.y = y; c.V = V; c.row = row; c.col = col;
    {
        double pf = p->E / (1 - pow(p->nu, 2));
        double vals[9] = {1.0, p->nu, 0.0, p->nu, 1.0, 0.0, 0.0, 0.0, 1 - p->nu};
        for (int q = 0; q < 9; ++q) c.C[q] = (long double)vals[q] * (long double)pf;
    }
    if (i0 < 0) i0 = 0;
    if (i1 > n) i1 = n;
    (void)nthreads;
#ifdef _OPENMP
    if (nthreads < 1) nthreads = 1;
#pragma omp parallel for num_threads(nthreads) schedule(dynamic, 4)
#endif
    for (int i = i0; i < i1; ++i) {
        long double net[2] = {0.0L, 0.0L};
        for (int b = row[i]; b < row[i + 1]; ++b) {
            int j = col[b];
            long double fij[2], fji[2];
            ref_force_state(&c, i, j, fij);
            ref_force_state(&c, j, i, fji);
            for (int d = 0; d < 2; ++d) net[d] += (fij[d] - fji[d]) * V[j];
        }
        for (int d = 0; d < 2; ++d) {
            f[2 * i + d] = net[d];
            long double an = f[2 * i + d] / p->rho;
            v[2 * i + d] += 0.5L * (a[2 * i + d] + an) * p->dt;
        }
    }
}

void ref_cpu_update_positions(const ref_params* p, int n, long double* y,
                              const long double* v, long double* a,
                              const long double* f)
{
    long double dt2 = powl(p->dt, 2);
    for (int i = 0; i < n; ++i)
        for (int d = 0; d < 2; ++d) {
            a[2 * i + d] = f[2 * i + d] / p->rho;
            y[2 * i + d] += v[2 * i + d] * p->dt + 0.5L * a[2 * i + d] * dt2;
        }
}

void ref_cpu_apply_velocity_bc(int cnt, const int* ids, const long double* v0,
                               long double* v, long double* a)
{
    for (int t = 0; t < cnt; ++t) {
        int i = ids[t];
        for (int d = 0; d < 2; ++d) {
            v[2 * i + d] = v0[d];
            a[2 * i + d] = 0.0L;
        }
    }
}

void ref_cpu_apply_fixed_bc(int cnt, const int* ids, const long double* X,
                            long double* y)
{
    for (int t = 0; t < cnt; ++t) {
        int i = ids[t];
        y[2 * i] = X[2 * i];
        y[2 * i + 1] = X[2 * i + 1];
    }
}

long ref_find_neighbor(int n, const long double* X, long double dx,
                       long double horizon, int* row, int* col, long cap)
{
    long nb = 0;
    long double cut = (long double)1.01 * dx * horizon;
    for (int i = 0; i < n; ++i) {
        row[i] = (int)nb;
        for (int j = 0; j < n; ++j) {
            if (i == j) continue;
            long double d2 = 0.0L;
            for (int k = 0; k < 2; ++k) {
                long double r = X[2 * j + k] - X[2 * i + k];
                d2 += r * r;
            }
            if (sqrtl(d2) < cut) {
                if (nb >= cap) return -1;
                col[nb++] = j;
            }
        }
    }
    row[n] = (int)nb;
    return nb;
}
