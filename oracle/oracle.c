/*
 * oracle.c -- CPU restatement of TreESaP's explicit time-stepping arithmetic.
 *
 * TEST INFRASTRUCTURE ONLY. This file is the checker that the HIP library is
 * compared against (tests/, the smoke entry point and the CPU-baseline leg of
 * bench.py). It is never linked into, loaded by or called from the product
 * library, and the product path has no CPU fallback that routes through it.
 *
 * Every function follows the reference file:line it cites (paths relative to
 * the treesap source tree). The arithmetic keeps the reference's operation
 * order term by term; build it with -ffp-contract=off so no FMA contraction
 * changes the rounding (the reference's CPU build has no -ffast-math either).
 *
 * Pinning (how each part is checked without a runnable reference, since the
 * reference needs AMReX, which is absent; see the header of oracle.py):
 *   weights  : Fornberg tables + polynomial exactness            (pinned)
 *   KO       : binomial closed form + Fourier symbol              (pinned)
 *   stencils : polynomial exactness, Fourier symbols               (pinned)
 *   RK4/Euler: single-Fourier-mode closed form G(k)^N              (pinned)
 *   ABL sigma: brute-force distance formula on tiny grids          (pinned)
 *   ABL term : RK4 polynomial of the damped affine ODE where sigma
 *              is constant over a step's reach (bulk and corners)  (pinned)
 *   ghosts   : brute-force periodic image / Dirichlet constants    (pinned)
 *   tensor C3 RHS (not in the reference): Fourier closed form      (pinned)
 */
#include <math.h>
#include <omp.h>
#include <stdint.h>
#include <stdlib.h>
#include <string.h>

#define OR_MAX_COMP 12

/* ------------------------------------------------------------------------- */
/* cexprmath (util/cexprmath.hpp)                                            */
/* ------------------------------------------------------------------------- */

/* util/cexprmath.hpp:32-37 */
static int64_t or_factorial(int n) {
    int64_t res = 1;
    for (int i = 2; i <= n; ++i) res *= i;
    return res;
}

/* util/cexprmath.hpp:46-59 */
int64_t or_binom(int n, int k) {
    if (k < 0 || k > n) return 0;
    if (k == 0 || k == n) return 1;
    if (k > n / 2) k = n - k;
    int64_t res = 1;
    for (int i = 1; i <= k; ++i) res = (res * (n - i + 1)) / i;
    return res;
}

/* util/cexprmath.hpp:68-77 */
static int64_t or_power(int64_t n, int k) {
    int64_t result = 1;
    while (k > 0) {
        if (k & 1) result *= n;
        n *= n;
        k >>= 1;
    }
    return result;
}

/* util/cexprmath.hpp:85 */
static int or_alt(int k) { return (k & 1) ? -1 : 1; }

/* util/cexprmath.hpp:99-105: left-to-right sum of 1/k^m in double */
double or_harmonic(int n, int m) {
    double sum = 0.0;
    for (int k = 1; k <= n; k++) sum += 1.0 / (double)or_power(k, m);
    return sum;
}

/* ------------------------------------------------------------------------- */
/* stencil weights (util/stencil_impl.hpp:56-140)                            */
/* ------------------------------------------------------------------------- */

/* util/stencil_impl.hpp:56-71 */
double or_w1(int n, int p) {
    int ka = p > 0 ? p : -p;
    if (p == 0) return 0;
    int64_t num = (int64_t)or_alt(ka + 1) * or_power(or_factorial(n), 2);
    int64_t den = (int64_t)ka * or_factorial(n - ka) * or_factorial(n + ka);
    double w = (double)num / (double)den;
    return p > 0 ? w : -w;
}

/* util/stencil_impl.hpp:91-104 */
double or_w2(int n, int p) {
    int k = p > 0 ? p : -p;
    if (p == 0) return -2.0 * or_harmonic(n, 2);
    int64_t num = (int64_t)(2 * or_alt(k + 1)) * or_power(or_factorial(n), 2);
    int64_t den = (int64_t)(k * k) * or_factorial(n - k) * or_factorial(n + k);
    return (double)num / (double)den;
}

/* util/stencil_impl.hpp:124-140 */
double or_w3(int n, int p) {
    int ka = p > 0 ? p : -p;
    if (p == 0) return 0;
    int64_t num = (int64_t)(6 * or_alt(ka + 1)) * or_power(or_factorial(n), 2);
    int64_t den = or_power(ka, 3) * or_factorial(n - ka) * or_factorial(n + ka);
    double w = ((1.0 - (double)(ka * ka) * or_harmonic(n, 2)) * (double)num / (double)den);
    return p > 0 ? w : -w;
}

/* KO coefficient as the reference forms it, util/stencil.hpp:96-101:
 * s == 0 -> -binom(2r, r); s > 0 -> -alternating_sign(s) * binom(2r, r+s). */
int64_t or_ko_coef(int r, int s) {
    if (s == 0) return -or_binom(2 * r, r);
    return (int64_t)(-or_alt(s)) * or_binom(2 * r, r + s);
}

/* ------------------------------------------------------------------------- */
/* solver state                                                               */
/* ------------------------------------------------------------------------- */

typedef struct {
    int system;          /* 0 (phi,Pi) wave, 1 (phi,u_a) wave, 2 nonlinear (phi,pi), 3 tensor (h,K) */
    int n[3];            /* valid cells per axis */
    int stencil_n;       /* derivative half-width */
    int ghost;           /* ghost width G */
    int ko_order;        /* 0: no KO decorator */
    double ko_strength;
    double prob_lo[3];
    double prob_hi[3];
    int periodic[3];
    int bc;              /* 0: NoOpBC, 1: DirichletBC(target) */
    int abl_thickness;   /* 0: no ABL decorator */
    double abl_max;
    int abl_profile;     /* 0: s, 1: s^2, 2: s^3 */
    double target[OR_MAX_COMP];
    double gamma;        /* nonlinear coefficient for system 2 */
    double dt;
} or_desc;

typedef struct {
    or_desc d;
    int nc, G;
    int64_t gx, gy, gz, gcells;
    double dx[3];
    double *reg[4];
    /* register roles (SingleLevelIntegrator.hpp:182-183,
       integrators/SingleLevelRK4Integrator.hpp:162-166) */
    int r_old, r_new, r_arg, r_argn;
    double time;
    long steps;
} or_solver;

static int sys_ncomp(int system) {
    switch (system) {
    case 0: return 2;
    case 1: return 4;
    case 2: return 2;
    case 3: return 12;
    default: return -1;
    }
}

int or_system_ncomp(int system) { return sys_ncomp(system); }

#define AT(s, r, i, j, k, c)                                                                   \
    ((s)->reg[r][((int64_t)(i) + (s)->G) + ((int64_t)(j) + (s)->G) * (s)->gx +                  \
                 ((int64_t)(k) + (s)->G) * (s)->gx * (s)->gy + (int64_t)(c) * (s)->gcells])

or_solver *or_create(const or_desc *d) {
    int nc = sys_ncomp(d->system);
    if (nc < 0) return NULL;
    if (d->stencil_n < 1 || d->stencil_n > 8) return NULL;
    if (d->ghost < d->stencil_n || d->ghost < d->ko_order) return NULL;
    for (int a = 0; a < 3; ++a)
        if (d->n[a] < 1 || (d->periodic[a] && d->ghost > d->n[a])) return NULL;
    or_solver *s = (or_solver *)calloc(1, sizeof(or_solver));
    s->d = *d;
    s->nc = nc;
    s->G = d->ghost;
    s->gx = d->n[0] + 2 * s->G;
    s->gy = d->n[1] + 2 * s->G;
    s->gz = d->n[2] + 2 * s->G;
    s->gcells = s->gx * s->gy * s->gz;
    /* Geometry::CellSize = (prob_hi - prob_lo) / n_cells */
    for (int a = 0; a < 3; ++a) s->dx[a] = (d->prob_hi[a] - d->prob_lo[a]) / (double)d->n[a];
    for (int r = 0; r < 4; ++r) {
        s->reg[r] = (double *)calloc((size_t)(s->gcells * nc), sizeof(double));
        if (!s->reg[r]) return NULL;
    }
    s->r_old = 0;
    s->r_new = 1;
    s->r_arg = 2;
    s->r_argn = 3;
    return s;
}

void or_destroy(or_solver *s) {
    if (!s) return;
    for (int r = 0; r < 4; ++r) free(s->reg[r]);
    free(s);
}

double or_time(const or_solver *s) { return s->time; }
long or_steps(const or_solver *s) { return s->steps; }
int64_t or_register_size(const or_solver *s) { return s->gcells * s->nc; }

/* valid-region state in [c][k][j][i] order */
void or_set_state(or_solver *s, const double *v) {
    int nx = s->d.n[0], ny = s->d.n[1], nz = s->d.n[2];
    for (int c = 0; c < s->nc; ++c)
        for (int k = 0; k < nz; ++k)
            for (int j = 0; j < ny; ++j)
                for (int i = 0; i < nx; ++i)
                    AT(s, s->r_old, i, j, k, c) = v[(((int64_t)c * nz + k) * ny + j) * nx + i];
}

void or_get_state(const or_solver *s, double *v) {
    int nx = s->d.n[0], ny = s->d.n[1], nz = s->d.n[2];
    for (int c = 0; c < s->nc; ++c)
        for (int k = 0; k < nz; ++k)
            for (int j = 0; j < ny; ++j)
                for (int i = 0; i < nx; ++i)
                    v[(((int64_t)c * nz + k) * ny + j) * nx + i] = AT(s, s->r_old, i, j, k, c);
}

/* whole ghosted register r (role index: 0 old, 1 new, 2 arg, 3 arg_next) */
void or_get_register(const or_solver *s, int role, double *out) {
    int r = role == 0 ? s->r_old : role == 1 ? s->r_new : role == 2 ? s->r_arg : s->r_argn;
    memcpy(out, s->reg[r], sizeof(double) * (size_t)(s->gcells * s->nc));
}

void or_set_register(or_solver *s, int role, const double *in) {
    int r = role == 0 ? s->r_old : role == 1 ? s->r_new : role == 2 ? s->r_arg : s->r_argn;
    memcpy(s->reg[r], in, sizeof(double) * (size_t)(s->gcells * s->nc));
}

/* ------------------------------------------------------------------------- */
/* ghost fill                                                                 */
/* ------------------------------------------------------------------------- */

/* SingleLevelIntegrator.hpp:153-163 -> FillPatchSingleLevel(mf,t,{&mf},{t},...):
 * FillBoundary copies periodic images of valid cells into ghosts; then the
 * physical-boundary functor runs on ghost cells outside the domain along a
 * non-periodic axis. DirichletBC (pde-model/bcs/DirichletBC.hpp:43-65) writes
 * all ncomp boundary values there; NoOpBC (pde-model/bcs/NoOpBC.hpp:41-46)
 * writes nothing. */
static void fill_ghosts(or_solver *s, int r) {
    const int G = s->G;
    const int nx = s->d.n[0], ny = s->d.n[1], nz = s->d.n[2];
    const int n[3] = {nx, ny, nz};
#pragma omp parallel for schedule(static)
    for (int k = -G; k < nz + G; ++k)
        for (int j = -G; j < ny + G; ++j)
            for (int i = -G; i < nx + G; ++i) {
                int iv[3] = {i, j, k};
                int inside = (i >= 0 && i < nx && j >= 0 && j < ny && k >= 0 && k < nz);
                if (inside) continue;
                int phys = 0;
                int src[3];
                for (int a = 0; a < 3; ++a) {
                    int v = iv[a];
                    if (v < 0 || v >= n[a]) {
                        if (s->d.periodic[a]) {
                            v = v < 0 ? v + n[a] : v - n[a];
                        } else {
                            phys = 1;
                        }
                    }
                    src[a] = v;
                }
                if (phys) {
                    if (s->d.bc == 1)
                        for (int c = 0; c < s->nc; ++c) AT(s, r, i, j, k, c) = s->d.target[c];
                } else {
                    for (int c = 0; c < s->nc; ++c) AT(s, r, i, j, k, c) = AT(s, r, src[0], src[1], src[2], c);
                }
            }
}

void or_fill_ghosts(or_solver *s, int role) {
    int r = role == 0 ? s->r_old : role == 1 ? s->r_new : role == 2 ? s->r_arg : s->r_argn;
    fill_ghosts(s, r);
}

/* ------------------------------------------------------------------------- */
/* stencils (util/stencil_impl.hpp:206-378, util/stencil.hpp:77-104)         */
/* ------------------------------------------------------------------------- */

static inline void step_dir(int dir, int s, int *di, int *dj, int *dk) {
    *di = dir == 0 ? s : 0;
    *dj = dir == 1 ? s : 0;
    *dk = dir == 2 ? s : 0;
}

/* util/stencil_impl.hpp:206-222 */
static double d1(const or_solver *s, int r, int n, int dir, int i, int j, int k, int c) {
    double grid_fac = 1.0 / s->dx[dir];
    double st = 0.0;
    for (int q = 1; q <= n; ++q) {
        int di, dj, dk;
        step_dir(dir, q, &di, &dj, &dk);
        st += or_w1(n, q) * (AT(s, r, i + di, j + dj, k + dk, c) - AT(s, r, i - di, j - dj, k - dk, c));
    }
    return grid_fac * st;
}

/* util/stencil_impl.hpp:254-268 (unmixed) and :269-285 (mixed, outer dir0 of inner dir1) */
static double d2(const or_solver *s, int r, int n, int dir0, int dir1, int i, int j, int k, int c) {
    if (dir0 == dir1) {
        double grid_fac = 1.0 / (s->dx[dir0] * s->dx[dir1]);
        double st = or_w2(n, 0) * AT(s, r, i, j, k, c);
        for (int q = 1; q <= n; ++q) {
            int di, dj, dk;
            step_dir(dir0, q, &di, &dj, &dk);
            st += or_w2(n, q) * (AT(s, r, i + di, j + dj, k + dk, c) + AT(s, r, i - di, j - dj, k - dk, c));
        }
        return grid_fac * st;
    }
    double grid_fac = 1.0 / s->dx[dir0];
    double st = 0.0;
    for (int q = 1; q <= n; ++q) {
        int di, dj, dk;
        step_dir(dir0, q, &di, &dj, &dk);
        st += or_w1(n, q) * (d1(s, r, n, dir1, i + di, j + dj, k + dk, c) - d1(s, r, n, dir1, i - di, j - dj, k - dk, c));
    }
    return grid_fac * st;
}

/* util/stencil_impl.hpp:318-377 */
static double d3(const or_solver *s, int r, int n, int dir0, int dir1, int dir2, int i, int j, int k, int c) {
    if (dir0 == dir1 && dir1 == dir2) {
        double grid_fac = 1.0 / (s->dx[dir0] * s->dx[dir1] * s->dx[dir2]);
        double st = 0.0;
        for (int q = 1; q <= n; ++q) {
            int di, dj, dk;
            step_dir(dir0, q, &di, &dj, &dk);
            st += or_w3(n, q) * (AT(s, r, i + di, j + dj, k + dk, c) - AT(s, r, i - di, j - dj, k - dk, c));
        }
        return grid_fac * st;
    } else if (dir0 == dir1 || dir0 == dir2 || dir1 == dir2) {
        int single, dbl;
        if (dir0 == dir1) {
            single = dir2;
            dbl = dir0;
        } else if (dir0 == dir2) {
            single = dir1;
            dbl = dir0;
        } else {
            single = dir0;
            dbl = dir1;
        }
        double grid_fac = 1.0 / s->dx[single];
        double st = 0.0;
        for (int q = 1; q <= n; ++q) {
            int di, dj, dk;
            step_dir(single, q, &di, &dj, &dk);
            st += or_w1(n, q) *
                  (d2(s, r, n, dbl, dbl, i + di, j + dj, k + dk, c) - d2(s, r, n, dbl, dbl, i - di, j - dj, k - dk, c));
        }
        return grid_fac * st;
    } else {
        double grid_fac = 1.0 / s->dx[dir0];
        double st = 0.0;
        for (int q = 1; q <= n; ++q) {
            int di, dj, dk;
            step_dir(dir0, q, &di, &dj, &dk);
            st += or_w1(n, q) *
                  (d2(s, r, n, dir1, dir2, i + di, j + dj, k + dk, c) - d2(s, r, n, dir1, dir2, i - di, j - dj, k - dk, c));
        }
        return grid_fac * st;
    }
}

/* util/stencil.hpp:77-104 */
static double ko(const or_solver *s, int r, int order, int dir, int i, int j, int k, int c) {
    double norm = 1.0 / (double)or_power(2, 2 * order);
    double grid_fac = 1.0 / s->dx[dir];
    double result = (double)(-or_binom(2 * order, order)) * AT(s, r, i, j, k, c);
    for (int q = 1; q <= order; ++q) {
        int di, dj, dk;
        step_dir(dir, q, &di, &dj, &dk);
        result += (double)((int64_t)(-or_alt(q)) * or_binom(2 * order, order + q)) *
                  (AT(s, r, i + di, j + dj, k + dk, c) + AT(s, r, i - di, j - dj, k - dk, c));
    }
    return result * norm * grid_fac;
}

/* ------------------------------------------------------------------------- */
/* ABL damping (pde-model/rhs-decorators/ABLDecorator.hpp:58-132)             */
/* ------------------------------------------------------------------------- */

static double abl_profile(int id, double x) {
    switch (id) {
    case 0: return x;
    case 1: return x * x;
    default: return x * x * x;
    }
}

static inline int clampi(int v, int lo, int hi) { return (v < lo) ? lo : (hi < v) ? hi : v; }

/* ABLDecorator.hpp:85-90 (bulk = domain grown by -thickness) and :115-132 */
double or_abl_sigma(const or_solver *s, int i, int j, int k) {
    int T = s->d.abl_thickness;
    int lo[3] = {T, T, T};
    int hi[3] = {s->d.n[0] - 1 - T, s->d.n[1] - 1 - T, s->d.n[2] - 1 - T};
    int iv[3] = {i, j, k};
    if (iv[0] >= lo[0] && iv[0] <= hi[0] && iv[1] >= lo[1] && iv[1] <= hi[1] && iv[2] >= lo[2] && iv[2] <= hi[2])
        return 0.0;
    int cv[3];
    for (int a = 0; a < 3; ++a) cv[a] = iv[a] - clampi(iv[a], lo[a], hi[a]);
    double norm = sqrt((double)(cv[0] * cv[0] + cv[1] * cv[1] + cv[2] * cv[2]));
    double inv_thickness = 1.0 / (double)T;
    double x = inv_thickness * norm;
    double scaled = (x < 0.0) ? 0.0 : (1.0 < x) ? 1.0 : x;
    return s->d.abl_max * abl_profile(s->d.abl_profile, scaled);
}

/* ------------------------------------------------------------------------- */
/* user right-hand sides                                                      */
/* ------------------------------------------------------------------------- */

/* symmetric rank-2 alias, pde-model/tensor-alias/SymRank2Alias.hpp:41-47 */
static inline int sym2(int a, int b) {
    int row = a > b ? a : b;
    int col = a > b ? b : a;
    return (row * (row + 1)) / 2 + col;
}
/* SymRank2Alias.hpp:55-62 */
static inline void sym2_restore(int rel, int *a, int *b) {
    int row = 0;
    while ((row + 1) * (row + 2) / 2 <= rel) ++row;
    *a = row;
    *b = rel - (row * (row + 1)) / 2;
}

typedef struct {
    double pi; /* NonLinearRHS Workspace (examples/02-temporaries/NonLinearRHS.hpp:34-36) */
} workspace_t;

static double user_rhs(const or_solver *s, int r, int c, int i, int j, int k, workspace_t *ws) {
    const int n = s->d.stencil_n;
    switch (s->d.system) {
    case 0: /* linear (phi, Pi) wave: phi_t = Pi; Pi_t = contract_a D_aa phi */
        if (c == 0) return AT(s, r, i, j, k, 1);
        {
            double result = 0.0; /* util/tensor_operations.hpp:49-52 */
            result += d2(s, r, n, 0, 0, i, j, k, 0);
            result += d2(s, r, n, 1, 1, i, j, k, 0);
            result += d2(s, r, n, 2, 2, i, j, k, 0);
            return result;
        }
    case 1: /* examples/01-wave-eqn/WaveEqnRHS.hpp:37-45 */
        if (c == 0) {
            double result = 0.0;
            result += d1(s, r, n, 0, i, j, k, 1);
            result += d1(s, r, n, 1, i, j, k, 2);
            result += d1(s, r, n, 2, i, j, k, 3);
            return result;
        }
        return d1(s, r, n, c - 1, i, j, k, 0);
    case 2: /* examples/02-temporaries/NonLinearRHS.hpp:39-71 */
        if (c == 0) {
            ws->pi = AT(s, r, i, j, k, 1);
            return ws->pi;
        } else {
            double g[3];
            for (int a = 0; a < 3; ++a) g[a] = d1(s, r, n, a, i, j, k, 0);
            double grad_sq = 0.0;
            grad_sq += g[0] * g[0];
            grad_sq += g[1] * g[1];
            grad_sq += g[2] * g[2];
            double lap = 0.0;
            lap += d2(s, r, n, 0, 0, i, j, k, 0);
            lap += d2(s, r, n, 1, 1, i, j, k, 0);
            lap += d2(s, r, n, 2, 2, i, j, k, 0);
            return (lap - s->d.gamma * ws->pi * grad_sq);
        }
    case 3: /* linearised ADM (h_ij, K_ij), Sym2 aliases at components 0 and 6 */
        if (c < 6) {
            return -2.0 * AT(s, r, i, j, k, 6 + c);
        } else {
            int a, b;
            sym2_restore(c - 6, &a, &b);
            /* symmetrise<a,b>(f), f<p,q> = contract_m D_m D_p h_{mq}  (tensor_operations.hpp:123-125) */
            double f_ab = 0.0, f_ba = 0.0;
            for (int m = 0; m < 3; ++m) f_ab += d2(s, r, n, m, a, i, j, k, sym2(m, b));
            for (int m = 0; m < 3; ++m) f_ba += d2(s, r, n, m, b, i, j, k, sym2(m, a));
            double sym = 0.5 * (f_ab + f_ba);
            double tdd = 0.0; /* contract_m D_a D_b h_mm */
            for (int m = 0; m < 3; ++m) tdd += d2(s, r, n, a, b, i, j, k, sym2(m, m));
            double lap = 0.0; /* contract_m D_m D_m h_ab */
            for (int m = 0; m < 3; ++m) lap += d2(s, r, n, m, m, i, j, k, sym2(a, b));
            return sym - 0.5 * (tdd + lap);
        }
    }
    return NAN;
}

/* decorated RHS: ABLDecorator<KODecorator<User>>
 * KODecorator.hpp:58-70: user + strength * (((0 + KO_x) + KO_y) + KO_z)
 * ABLDecorator.hpp:102-108: inner - sigma(iv) * (arr(iv,c) - target[c]) */
static double rhs(const or_solver *s, int r, int c, int i, int j, int k, workspace_t *ws) {
    double v = user_rhs(s, r, c, i, j, k, ws);
    if (s->d.ko_order > 0) {
        double ko_term = 0.0;
        ko_term += ko(s, r, s->d.ko_order, 0, i, j, k, c);
        ko_term += ko(s, r, s->d.ko_order, 1, i, j, k, c);
        ko_term += ko(s, r, s->d.ko_order, 2, i, j, k, c);
        v = v + s->d.ko_strength * ko_term;
    }
    if (s->d.abl_thickness > 0) v = v - or_abl_sigma(s, i, j, k) * (AT(s, r, i, j, k, c) - s->d.target[c]);
    return v;
}

/* evaluate the decorated RHS of register `role` (after filling its ghosts) into out[c][k][j][i] */
void or_eval_rhs(or_solver *s, int role, double *out) {
    int r = role == 0 ? s->r_old : role == 1 ? s->r_new : role == 2 ? s->r_arg : s->r_argn;
    fill_ghosts(s, r);
    int nx = s->d.n[0], ny = s->d.n[1], nz = s->d.n[2];
#pragma omp parallel for schedule(static)
    for (int k = 0; k < nz; ++k)
        for (int j = 0; j < ny; ++j)
            for (int i = 0; i < nx; ++i) {
                workspace_t ws = {0};
                for (int c = 0; c < s->nc; ++c)
                    out[(((int64_t)c * nz + k) * ny + j) * nx + i] = rhs(s, r, c, i, j, k, &ws);
            }
}

/* point derivative of component c of the old state (ghosts are filled first):
 * nd = 1..3 directions d0,d1,d2 (util/stencil.hpp:42-52), or nd = 0 -> KO order d0 along d1 */
double or_point_derivative(or_solver *s, int nd, int d0, int d1_, int d2_, int i, int j, int k, int c) {
    int r = s->r_old;
    int n = s->d.stencil_n;
    if (nd == 1) return d1(s, r, n, d0, i, j, k, c);
    if (nd == 2) return d2(s, r, n, d0, d1_, i, j, k, c);
    if (nd == 3) return d3(s, r, n, d0, d1_, d2_, i, j, k, c);
    return ko(s, r, d0, d1_, i, j, k, c);
}

/* ------------------------------------------------------------------------- */
/* integrators                                                                */
/* ------------------------------------------------------------------------- */

/* integrators/SingleLevelRK4Integrator.hpp:71-74 */
static const double a_frac[4] = {0.0, 0.5, 0.5, 1.0};
static const double b_frac[4] = {1.0 / 6.0, 1.0 / 3.0, 1.0 / 3.0, 1.0 / 6.0};

/* integrators/SingleLevelRK4Integrator.hpp:115-157 */
static void rk4_stage(or_solver *s, int stage, double dt) {
    const int is_first = stage == 0, is_last = stage == 3;
    double a_next = dt * a_frac[(stage + 1) % 4];
    double time_arg = s->time + dt * a_frac[stage];
    double b = dt * b_frac[stage];
    (void)time_arg; /* the registered systems are autonomous */
    if (!is_first) {
        int t = s->r_argn;
        s->r_argn = s->r_arg;
        s->r_arg = t;
        fill_ghosts(s, s->r_arg);
    }
    const int init = s->r_old, acc = s->r_new, arg = s->r_arg, argn = s->r_argn;
    const int src = is_first ? init : arg;
    int nx = s->d.n[0], ny = s->d.n[1], nz = s->d.n[2];
#pragma omp parallel for schedule(static)
    for (int k = 0; k < nz; ++k)
        for (int j = 0; j < ny; ++j)
            for (int i = 0; i < nx; ++i) {
                workspace_t ws = {0};
                for (int c = 0; c < s->nc; ++c) {
                    double st = rhs(s, src, c, i, j, k, &ws);
                    if (is_first)
                        AT(s, acc, i, j, k, c) = AT(s, init, i, j, k, c) + b * st;
                    else
                        AT(s, acc, i, j, k, c) += b * st;
                    if (!is_last) AT(s, argn, i, j, k, c) = AT(s, init, i, j, k, c) + a_next * st;
                }
            }
}

/* Integrator.hpp:91-106 with SingleLevelIntegrator.hpp:102-125 and
 * integrators/SingleLevelRK4Integrator.hpp:95-101 */
void or_rk4_steps(or_solver *s, int nsteps) {
    for (int n = 0; n < nsteps; ++n) {
        double dt = s->d.dt;
        fill_ghosts(s, s->r_old); /* pre_step */
        for (int st = 0; st < 4; ++st) rk4_stage(s, st, dt);
        s->steps++;
        s->time += dt;
        int t = s->r_old; /* post_step: swap(state_old, state_new) */
        s->r_old = s->r_new;
        s->r_new = t;
    }
}

/* integrators/SingleLevelEulerIntegrator.hpp:66-88 */
void or_euler_steps(or_solver *s, int nsteps) {
    for (int n = 0; n < nsteps; ++n) {
        double dt = s->d.dt;
        fill_ghosts(s, s->r_old);
        const int old = s->r_old, nw = s->r_new;
        int nx = s->d.n[0], ny = s->d.n[1], nz = s->d.n[2];
#pragma omp parallel for schedule(static)
        for (int k = 0; k < nz; ++k)
            for (int j = 0; j < ny; ++j)
                for (int i = 0; i < nx; ++i) {
                    workspace_t ws = {0};
                    for (int c = 0; c < s->nc; ++c)
                        AT(s, nw, i, j, k, c) = AT(s, old, i, j, k, c) + dt * rhs(s, old, c, i, j, k, &ws);
                }
        s->steps++;
        s->time += dt;
        s->r_old = nw;
        s->r_new = old;
    }
}

/* ------------------------------------------------------------------------- */
/* discrete adjoint of one RK4 step, linear (phi, Pi) system                  */
/* ------------------------------------------------------------------------- */

/* Transpose of the decorated (phi, Pi) operator with homogeneous ghosts (zero
 * Dirichlet values, periodic images): D_aa and the KO stencil are symmetric,
 * the ABL term is diagonal, so L^T = [[KO - s, Lap], [I, KO - s]]. Same
 * operation order as the forward RHS above. */
static double adj_rhs(const or_solver *s, int r, int c, int i, int j, int k) {
    const int n = s->d.stencil_n;
    double v;
    if (c == 0) {
        double result = 0.0;
        result += d2(s, r, n, 0, 0, i, j, k, 1);
        result += d2(s, r, n, 1, 1, i, j, k, 1);
        result += d2(s, r, n, 2, 2, i, j, k, 1);
        v = result;
    } else {
        v = AT(s, r, i, j, k, 0);
    }
    if (s->d.ko_order > 0) {
        double ko_term = 0.0;
        ko_term += ko(s, r, s->d.ko_order, 0, i, j, k, c);
        ko_term += ko(s, r, s->d.ko_order, 1, i, j, k, c);
        ko_term += ko(s, r, s->d.ko_order, 2, i, j, k, c);
        v = v + s->d.ko_strength * ko_term;
    }
    if (s->d.abl_thickness > 0) v = v - or_abl_sigma(s, i, j, k) * (AT(s, r, i, j, k, c) - 0.0);
    return v;
}

/* ghost fill with homogeneous boundary values (adjoint registers) */
static void fill_ghosts_homogeneous(or_solver *s, int r) {
    double saved[OR_MAX_COMP];
    int saved_bc = s->d.bc;
    memcpy(saved, s->d.target, sizeof(saved));
    memset(s->d.target, 0, sizeof(s->d.target));
    s->d.bc = 1;
    fill_ghosts(s, r);
    memcpy(s->d.target, saved, sizeof(saved));
    s->d.bc = saved_bc;
}

/* lambda <- G^T lambda for one RK4 step G (integrators/SingleLevelRK4Integrator.hpp:95-157):
 *   mu_3 = b_3 lambda';  nu_j = L^T mu_j;
 *   mu_{j-1} = b_{j-1} lambda' + a_j nu_j  (j = 3, 2, 1);
 *   lambda = (((lambda' + nu_3) + nu_2) + nu_1) + nu_0,
 * with b_j = dt * b_fraction[j], a_j = dt * a_fraction[j]. Uses all four
 * registers of the solver as scratch; lambda is the valid region [c][k][j][i]. */
int or_adjoint_step(or_solver *s, double *lam) {
    if (s->d.system != 0) return -1;
    const double dt = s->d.dt;
    const int nx = s->d.n[0], ny = s->d.n[1], nz = s->d.n[2];
    const int LP = 0, ACC = 1, MA = 2, MB = 3; /* lambda', acc, mu ping-pong */
    for (int c = 0; c < 2; ++c)
        for (int k = 0; k < nz; ++k)
            for (int j = 0; j < ny; ++j)
                for (int i = 0; i < nx; ++i) {
                    double v = lam[(((int64_t)c * nz + k) * ny + j) * nx + i];
                    AT(s, LP, i, j, k, c) = v;
                    AT(s, MA, i, j, k, c) = (dt * b_frac[3]) * v;
                }
    int src = MA, dst = MB;
    for (int stage = 3; stage >= 0; --stage) {
        fill_ghosts_homogeneous(s, src);
        const double bprev = stage > 0 ? dt * b_frac[stage - 1] : 0.0;
        const double astage = dt * a_frac[stage];
#pragma omp parallel for schedule(static)
        for (int k = 0; k < nz; ++k)
            for (int j = 0; j < ny; ++j)
                for (int i = 0; i < nx; ++i)
                    for (int c = 0; c < 2; ++c) {
                        double nu = adj_rhs(s, src, c, i, j, k);
                        double lp = AT(s, LP, i, j, k, c);
                        if (stage == 3)
                            AT(s, ACC, i, j, k, c) = lp + nu;
                        else
                            AT(s, ACC, i, j, k, c) = AT(s, ACC, i, j, k, c) + nu;
                        if (stage > 0) AT(s, dst, i, j, k, c) = bprev * lp + astage * nu;
                    }
        int t = src;
        src = dst;
        dst = t;
    }
    for (int c = 0; c < 2; ++c)
        for (int k = 0; k < nz; ++k)
            for (int j = 0; j < ny; ++j)
                for (int i = 0; i < nx; ++i)
                    lam[(((int64_t)c * nz + k) * ny + j) * nx + i] = AT(s, ACC, i, j, k, c);
    return 0;
}

/* ------------------------------------------------------------------------- */
/* discrete adjoint of one RK4 step, NonLinear (phi, pi) system               */
/* ------------------------------------------------------------------------- */

/* examples/02-temporaries/NonLinearRHS.hpp:52-71: F_phi = pi, F_pi = Lap phi -
 * gamma pi |grad phi|^2 (+ KO, ABL). Its Jacobian at a state (phi_b, pi_b) is
 * the (phi, Pi) operator L plus N with
 *   (N v)_pi = -gamma (v_pi |g|^2 + 2 pi_b g . D1(v_phi)),  g = D1(phi_b),
 * so J^T mu = L^T mu + N^T mu,
 *   (N^T mu)_pi  = -(gamma |g|^2) mu_pi,
 *   (N^T mu)_phi = -sum_a gf1_a sum_s w1(n,s) (W_a(i - s e_a) - W_a(i + s e_a)),
 *   W_a = ((2 gamma) pi_b) g_a mu_pi  (zero outside the valid box along
 *   non-periodic axes; periodic images along periodic ones).
 * Each reverse stage j applies L^T as or_adjoint_step does, then adds N^T at
 * the stage argument a_j of the forward step (a_0 = u, a_j = u + a_j k_{j-1})
 * to nu_j in both of its uses. */
static double nl_w_at(const double *W, const or_solver *s, int a, int i, int j, int k) {
    int idx[3] = {i, j, k};
    for (int d = 0; d < 3; ++d) {
        if (idx[d] < 0 || idx[d] >= s->d.n[d]) {
            if (!s->d.periodic[d]) return 0.0;
            idx[d] = (idx[d] % s->d.n[d] + s->d.n[d]) % s->d.n[d];
        }
    }
    return W[(((int64_t)a * s->d.n[2] + idx[2]) * s->d.n[1] + idx[1]) * s->d.n[0] + idx[0]];
}

/* The forward step's stage arguments a_0..a_3 from the state u (valid region),
 * ghosts filled as the forward stages read them: st[q] = register of a_q. */
static void nl_stage_args(const or_solver *s, const double *u, double **st) {
    or_solver *f = or_create(&s->d);
    or_set_state(f, u);
    const double dt = s->d.dt;
    const size_t bytes = sizeof(double) * (size_t)f->gcells * f->nc;
    fill_ghosts(f, f->r_old);
    memcpy(st[0], f->reg[f->r_old], bytes);
    for (int q = 0; q < 4; ++q) {
        rk4_stage(f, q, dt);
        if (q > 0) memcpy(st[q], f->reg[f->r_arg], bytes); /* the stage's source, ghosts filled */
    }
    or_destroy(f);
}

int or_adjoint_step_nl(or_solver *s, double *lam, const double *u) {
    if (s->d.system != 2) return -1;
    const double dt = s->d.dt, gamma = s->d.gamma;
    const int nx = s->d.n[0], ny = s->d.n[1], nz = s->d.n[2], n = s->d.stencil_n;
    const int64_t cells = (int64_t)nx * ny * nz;
    const size_t rbytes = sizeof(double) * (size_t)s->gcells * s->nc;
    double *st[4];
    for (int q = 0; q < 4; ++q) st[q] = (double *)malloc(rbytes);
    double *W = (double *)malloc(sizeof(double) * (size_t)cells * 4);
    nl_stage_args(s, u, st);
    const int LP = 0, ACC = 1, MA = 2, MB = 3;
    for (int c = 0; c < 2; ++c)
        for (int k = 0; k < nz; ++k)
            for (int j = 0; j < ny; ++j)
                for (int i = 0; i < nx; ++i) {
                    double v = lam[(((int64_t)c * nz + k) * ny + j) * nx + i];
                    AT(s, LP, i, j, k, c) = v;
                    AT(s, MA, i, j, k, c) = (dt * b_frac[3]) * v;
                }
    int src = MA, dst = MB;
    for (int stage = 3; stage >= 0; --stage) {
        fill_ghosts_homogeneous(s, src);
        const double bprev = stage > 0 ? dt * b_frac[stage - 1] : 0.0;
        const double astage = dt * a_frac[stage];
#pragma omp parallel for schedule(static)
        for (int k = 0; k < nz; ++k)
            for (int j = 0; j < ny; ++j)
                for (int i = 0; i < nx; ++i)
                    for (int c = 0; c < 2; ++c) {
                        double nu = adj_rhs(s, src, c, i, j, k);
                        double lp = AT(s, LP, i, j, k, c);
                        if (stage == 3)
                            AT(s, ACC, i, j, k, c) = lp + nu;
                        else
                            AT(s, ACC, i, j, k, c) = AT(s, ACC, i, j, k, c) + nu;
                        if (stage > 0) AT(s, dst, i, j, k, c) = bprev * lp + astage * nu;
                    }
        /* N^T at the stage argument: W and the pi part from mu (src) and a_stage */
        double *areg = st[stage];
#pragma omp parallel for schedule(static)
        for (int k = 0; k < nz; ++k)
            for (int j = 0; j < ny; ++j)
                for (int i = 0; i < nx; ++i) {
                    const int64_t q = ((int64_t)k * ny + j) * nx + i;
                    double g[3];
                    for (int a = 0; a < 3; ++a) {
                        double sum = 0.0;
                        for (int t = 1; t <= n; ++t) {
                            int di, dj, dk;
                            step_dir(a, t, &di, &dj, &dk);
                            const double fp = areg[((int64_t)(i + di) + s->G) + ((int64_t)(j + dj) + s->G) * s->gx +
                                                   ((int64_t)(k + dk) + s->G) * s->gx * s->gy];
                            const double fm = areg[((int64_t)(i - di) + s->G) + ((int64_t)(j - dj) + s->G) * s->gx +
                                                   ((int64_t)(k - dk) + s->G) * s->gx * s->gy];
                            sum += or_w1(n, t) * (fp - fm);
                        }
                        g[a] = (1.0 / s->dx[a]) * sum;
                    }
                    double gsq = g[0] * g[0];
                    gsq += g[1] * g[1];
                    gsq += g[2] * g[2];
                    const double pib = areg[((int64_t)i + s->G) + ((int64_t)j + s->G) * s->gx +
                                            ((int64_t)k + s->G) * s->gx * s->gy + s->gcells];
                    const double mupi = AT(s, src, i, j, k, 1);
                    const double tp = (2.0 * gamma) * pib;
                    for (int a = 0; a < 3; ++a) W[a * cells + q] = (tp * g[a]) * mupi;
                    W[3 * cells + q] = -((gamma * gsq) * mupi);
                }
#pragma omp parallel for schedule(static)
        for (int k = 0; k < nz; ++k)
            for (int j = 0; j < ny; ++j)
                for (int i = 0; i < nx; ++i) {
                    double cphi = 0.0;
                    for (int a = 0; a < 3; ++a) {
                        double sum = 0.0;
                        for (int t = 1; t <= n; ++t) {
                            int di, dj, dk;
                            step_dir(a, t, &di, &dj, &dk);
                            sum += or_w1(n, t) * (nl_w_at(W, s, a, i - di, j - dj, k - dk) -
                                                  nl_w_at(W, s, a, i + di, j + dj, k + dk));
                        }
                        cphi += (1.0 / s->dx[a]) * sum;
                    }
                    const double c_phi = -cphi;
                    const double c_pi = W[3 * cells + ((int64_t)k * ny + j) * nx + i];
                    AT(s, ACC, i, j, k, 0) = AT(s, ACC, i, j, k, 0) + c_phi;
                    AT(s, ACC, i, j, k, 1) = AT(s, ACC, i, j, k, 1) + c_pi;
                    if (stage > 0) {
                        AT(s, dst, i, j, k, 0) = AT(s, dst, i, j, k, 0) + astage * c_phi;
                        AT(s, dst, i, j, k, 1) = AT(s, dst, i, j, k, 1) + astage * c_pi;
                    }
                }
        int t = src;
        src = dst;
        dst = t;
    }
    for (int c = 0; c < 2; ++c)
        for (int k = 0; k < nz; ++k)
            for (int j = 0; j < ny; ++j)
                for (int i = 0; i < nx; ++i)
                    lam[(((int64_t)c * nz + k) * ny + j) * nx + i] = AT(s, ACC, i, j, k, c);
    for (int q = 0; q < 4; ++q) free(st[q]);
    free(W);
    return 0;
}

/* OpenMP threads used by the stepping loops (CPU-baseline timing) */
void or_set_threads(int n) { omp_set_num_threads(n > 0 ? n : 1); }
int or_max_threads(void) { return omp_get_max_threads(); }
