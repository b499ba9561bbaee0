/*
 * mpm_oracle.c -- CPU restatement of DiffStir's MLS-MPM stirring substep and its adjoint.
 *
 * TEST INFRASTRUCTURE ONLY.  This file is the checker for the HIP path: only tests/,
 * __graft_entry__.smoke() and bench.py's cpu_baseline leg load it.  The product
 * library (diffstir_amd/csrc) never links, calls or falls back to it.
 *
 * Compiled three times from this one source (oracle/Makefile): REAL=double is the FP64
 * oracle; REAL=float is an FP32 restatement, used to size FP32 rounding drift; REAL=float
 * with ORACLE_ATOMIC is the reference's execution structure (one parallel loop over the
 * particles of each species with atomic adds into a dense grid, as Taichi runs sf:104-196 on
 * ti.cpu), used only as the CPU baseline of the benchmark.  Parameters, masses, the loss and
 * every accumulation of the parameter gradient are double in every build.
 *
 * Threads (OpenMP).  Without ORACLE_ATOMIC the results do not depend on the thread count:
 * gathers and per-node work are independent; scatters run over groups of particles with one
 * stencil-base column (bx, bz), the nine residue classes (bx mod 3, bz mod 3) one after the
 * other, so groups running at once never write the same node and every node receives its
 * contributions in an order fixed by the data; the stick-force and parameter-gradient sums
 * are formed per particle and added in particle order.
 *
 * Reference semantics followed (file:line in the DiffStir reference,
 * sf = src/simulator/simulation_functions.py, init = src/simulator/initialization.py):
 *   or_bspline        quadratic B-spline with truncating base          sf:108-110, sf:231-233
 *   p2g_*             F_tmp=(I+dt C)F, J=det F_tmp, rheology, stress,  sf:39-42, sf:104-196
 *                     APIC scatter of three species
 *   visc              Carreau / Cross / Herschel-Bulkley viscosity    sf:119-146
 *   gridop_node       momentum->velocity, cylindrical slip wall,      sf:199-225
 *                     box bounds (top face zeroes all components)
 *   g2p_*             gather, advect liquid, kinematic stick with dv,  sf:228-290
 *                     container C from the grid
 *   forces            F(s) = m_stick/dt * sum_p dv[s,p]               sf:330-339, fcam:16-19
 * The adjoint (or_step_adjoint) is a hand derivation of the same step: it plays the role
 * of the reference's Taichi autodiff (sf:351-365 driven by ti.ad.Tape, gdm:103-107).
 *
 * Parity pins (tests/test_oracle_*.py): B-spline partition of unity and moment identities,
 * mass/momentum conservation of P2G, exact reconstruction of an affine grid field by G2P,
 * closed-form viscosities, brute-force enumeration of every grid_op branch, the stick
 * centripetal-force special case, the trajectory transform against the reference's
 * output/stick_positions.npy, an independent PyTorch FP64 autograd restatement of one
 * substep (adjoint), and central finite differences of whole rollouts.
 */
#include <math.h>
#include <stdlib.h>
#include <string.h>
#ifdef _OPENMP
#include <omp.h>
#endif

#ifndef REAL
#define REAL double
#endif

#if defined(ORACLE_ATOMIC)
#define OR_SYM(name) ora32_##name
#elif defined(ORACLE_F32)
#define OR_SYM(name) or32_##name
#else
#define OR_SYM(name) or_##name
#endif

/* grid accumulation: atomic in the reference-structured build, plain (race-free by the
 * column colouring) otherwise */
#ifdef ORACLE_ATOMIC
#define GRID_ADD(dst, val) do { REAL _v = (val); _Pragma("omp atomic") (dst) += _v; } while (0)
#else
#define GRID_ADD(dst, val) ((dst) += (val))
#endif

typedef struct {
    int n_grid;        /* nodes per axis */
    int bound;         /* box-boundary width in nodes (init:15) */
    int model;         /* 0 Carreau, 1 Cross, 2 Herschel-Bulkley (sf:120,129,138) */
    int k_from_param;  /* 1: bulk modulus is params[3] (HB only); 0: use k_bulk */
    double dx, inv_dx, dt;
    double gravity;    /* signed y acceleration added to liquid velocity (sf:152) */
    double k_bulk;     /* reference: 20000 hard-coded (sf:112) */
    double wall_r, wall_cx, wall_cz;  /* cylindrical slip wall (sf:205-210) */
    double m_liq, vol_liq, m_stick, m_cont;
} or_scene;

typedef struct {
    int nl, ns, nc;
    REAL *lx, *lv, *lC, *lF;   /* liquid: x[nl*3], v[nl*3], C[nl*9], F[nl*9] (row-major 3x3) */
    REAL *sx, *sv, *sC;        /* stick */
    REAL *cx, *cv, *cC;        /* container */
} or_state;

typedef struct {
    REAL *lx, *lv, *lC, *lF;   /* adjoints of liquid x, v, C, F */
    REAL *sC;                  /* adjoint of stick C */
    REAL *cC;                  /* adjoint of container C */
} or_adj;

enum { OR_OK = 0, OR_OUT_OF_GRID = 1, OR_NONFINITE = 2, OR_NOMEM = 3, OR_BADARG = 4 };

/* ------------------------------------------------------------------ helpers */

static int or_bspline(const or_scene* sc, const REAL* xp, int b[3], REAL fx[3],
                      REAL w[3][3], REAL dw[3][3])
{
    for (int a = 0; a < 3; a++) {
        REAL X = xp[a] * (REAL)sc->inv_dx;
        if (!isfinite((double)X)) return OR_NONFINITE;
        /* sf:108: ti.cast(x*inv_dx - 0.5, ti.i32) truncates toward zero */
        b[a] = (int)(X - (REAL)0.5);
        fx[a] = X - (REAL)b[a];
        REAL f = fx[a];
        w[a][0] = (REAL)0.5 * ((REAL)1.5 - f) * ((REAL)1.5 - f);
        w[a][1] = (REAL)0.75 - (f - (REAL)1) * (f - (REAL)1);
        w[a][2] = (REAL)0.5 * (f - (REAL)0.5) * (f - (REAL)0.5);
        dw[a][0] = f - (REAL)1.5;
        dw[a][1] = (REAL)-2 * (f - (REAL)1);
        dw[a][2] = f - (REAL)0.5;
        if (b[a] < 0 || b[a] + 2 >= sc->n_grid) return OR_OUT_OF_GRID;
    }
    return OR_OK;
}

static inline long nidx(const or_scene* sc, int i, int j, int k)
{
    long n = sc->n_grid;
    return ((long)i * n + j) * n + k;
}

static REAL det3(const REAL* M)
{
    return M[0] * (M[4] * M[8] - M[5] * M[7])
         - M[1] * (M[3] * M[8] - M[5] * M[6])
         + M[2] * (M[3] * M[7] - M[4] * M[6]);
}

/* cofactor matrix: d det(M) / dM */
static void cof3(const REAL* M, REAL* K)
{
    K[0] = M[4] * M[8] - M[5] * M[7];
    K[1] = M[5] * M[6] - M[3] * M[8];
    K[2] = M[3] * M[7] - M[4] * M[6];
    K[3] = M[2] * M[7] - M[1] * M[8];
    K[4] = M[0] * M[8] - M[2] * M[6];
    K[5] = M[1] * M[6] - M[0] * M[7];
    K[6] = M[1] * M[5] - M[2] * M[4];
    K[7] = M[2] * M[3] - M[0] * M[5];
    K[8] = M[0] * M[4] - M[1] * M[3];
}

/* viscosity eta(gamma; params) and its partials.  sf:119-146 */
static double visc(int model, const double* th, double g, double* d_dg, double* d_dth)
{
    double eta = 0.0;
    double dg = 0.0, dth[4] = {0, 0, 0, 0};
    if (model == 0) {            /* Carreau: eta_inf + (eta0-eta_inf)(1+(lam g)^2)^((n-1)/2) */
        double e0 = th[0], ei = th[1], n = th[2], lam = th[3];
        double lg = lam * g;
        double q = 1.0 + lg * lg;
        double e = (n - 1.0) / 2.0;
        double qe = pow(q, e);
        eta = ei + (e0 - ei) * qe;
        double dqe_dq = e * pow(q, e - 1.0);
        dth[0] = qe;
        dth[1] = 1.0 - qe;
        dth[2] = (e0 - ei) * qe * log(q) * 0.5;
        dth[3] = (e0 - ei) * dqe_dq * 2.0 * lg * g;
        dg = (e0 - ei) * dqe_dq * 2.0 * lg * lam;
    } else if (model == 1) {     /* Cross: eta_inf + (eta0-eta_inf)/(1+(lam g)^(n-1)) */
        double e0 = th[0], ei = th[1], n = th[2], lam = th[3];
        double lg = lam * g;
        double r = pow(lg, n - 1.0);
        double D = 1.0 + r;
        eta = ei + (e0 - ei) * (1.0 / D);
        double dr_dlg = (n - 1.0) * pow(lg, n - 2.0);
        double c = -(e0 - ei) / (D * D);
        dth[0] = 1.0 / D;
        dth[1] = 1.0 - 1.0 / D;
        dth[2] = c * r * log(lg);
        dth[3] = c * dr_dlg * g;
        dg = c * dr_dlg * lam;
    } else {                     /* Herschel-Bulkley: tau_y/g + [K g^(n-1) if > 0] */
        double K = th[0], ty = th[1], n = th[2];
        double gn = pow(g, n - 1.0);
        double de = K * gn;
        eta = ty / g;
        dth[1] = 1.0 / g;
        dg = -ty / (g * g);
        if (de > 0) {
            eta += de;
            dth[0] = gn;
            dth[2] = de * log(g);
            dg += K * (n - 1.0) * pow(g, n - 2.0);
        }
    }
    if (d_dg) *d_dg = dg;
    if (d_dth) for (int i = 0; i < 4; i++) d_dth[i] = dth[i];
    return eta;
}

static double bulk_k(const or_scene* sc, const double* th)
{
    return sc->k_from_param ? th[3] : sc->k_bulk;
}

/* ------------------------------------------------------------------ forward */

/* Liquid particle p2g quantities shared by the forward and the adjoint. */
typedef struct {
    int b[3];
    REAL fx[3], w[3][3], dw[3][3];
    REAL Fn[9];      /* (I + dt C) F */
    REAL J, gamma, eta;
    REAL S[9];       /* stress before scaling */
    REAL A[9];       /* affine */
    REAL u[3];       /* v + dt g */
} liq_pre;

static int liquid_pre(const or_scene* sc, const double* th, const REAL* x, const REAL* v,
                      const REAL* C, const REAL* F, liq_pre* q)
{
    int rc = or_bspline(sc, x, q->b, q->fx, q->w, q->dw);
    if (rc) return rc;
    REAL dt = (REAL)sc->dt;
    REAL Mx[9];
    for (int i = 0; i < 3; i++)
        for (int j = 0; j < 3; j++) Mx[i * 3 + j] = (i == j ? (REAL)1 : (REAL)0) + dt * C[i * 3 + j];
    for (int i = 0; i < 3; i++)
        for (int j = 0; j < 3; j++) {
            REAL s = 0;
            for (int k = 0; k < 3; k++) s += Mx[i * 3 + k] * F[k * 3 + j];
            q->Fn[i * 3 + j] = s;
        }
    q->J = det3(q->Fn);
    REAL nc = 0;
    for (int i = 0; i < 9; i++) nc += C[i] * C[i];
    q->gamma = (REAL)sqrt((double)nc) + (REAL)1e-6;
    q->eta = (REAL)visc(sc->model, th, (double)q->gamma, NULL, NULL);
    REAL k = (REAL)bulk_k(sc, th);
    REAL J = q->J;
    REAL pr = (k / (REAL)2) * (J - ((REAL)1 / J) - (REAL)1) * J;
    REAL alpha = (REAL)(-sc->dt * sc->vol_liq * 4.0 * sc->inv_dx * sc->inv_dx);
    REAL m = (REAL)sc->m_liq;
    for (int i = 0; i < 3; i++)
        for (int j = 0; j < 3; j++) {
            REAL s = q->eta * (REAL)0.5 * (C[i * 3 + j] + C[j * 3 + i]) + (i == j ? pr : (REAL)0);
            q->S[i * 3 + j] = s;
            q->A[i * 3 + j] = alpha * s + m * C[i * 3 + j];
        }
    q->u[0] = v[0];
    q->u[1] = v[1] + (REAL)(sc->dt * sc->gravity);
    q->u[2] = v[2];
    return OR_OK;
}

static void scatter(const or_scene* sc, REAL* P, const int b[3], const REAL fx[3],
                    const REAL w[3][3], REAL m, const REAL* u, const REAL* A)
{
    REAL dx = (REAL)sc->dx;
    for (int i = 0; i < 3; i++)
        for (int j = 0; j < 3; j++)
            for (int k = 0; k < 3; k++) {
                REAL dp[3] = {((REAL)i - fx[0]) * dx, ((REAL)j - fx[1]) * dx, ((REAL)k - fx[2]) * dx};
                REAL W = w[0][i] * w[1][j] * w[2][k];
                REAL* g = P + 4 * nidx(sc, b[0] + i, b[1] + j, b[2] + k);
                for (int a = 0; a < 3; a++)
                    GRID_ADD(g[a], W * (m * u[a] + A[a * 3 + 0] * dp[0] + A[a * 3 + 1] * dp[1] + A[a * 3 + 2] * dp[2]));
                GRID_ADD(g[3], W * m);
            }
}

/* ------------------------------------------------------------------ parallel loops */

/* Particles of one species grouped by stencil-base column g = bx * n + bz (a counting sort,
 * index order within a group).  Checks every position first: returns the error of the lowest
 * failing particle index, as a sequential loop would. */
typedef struct {
    int np;
    int* start;   /* [n*n + 1] */
    int* list;    /* [np] */
} colgroups;

static void colgroups_free(colgroups* g)
{
    free(g->start);
    free(g->list);
    g->start = g->list = NULL;
}

static int colgroups_build(const or_scene* sc, int np, const REAL* X, colgroups* g)
{
    int n = sc->n_grid;
    long ng = (long)n * n;
    g->np = np;
    g->start = (int*)calloc(ng + 1, sizeof(int));
    g->list = (int*)malloc(sizeof(int) * (np > 0 ? np : 1));
    int* key = (int*)malloc(sizeof(int) * (np > 0 ? np : 1));
    if (!g->start || !g->list || !key) { free(key); colgroups_free(g); return OR_NOMEM; }
    int first_bad = np;
    #pragma omp parallel for reduction(min : first_bad)
    for (int p = 0; p < np; p++) {
        int b[3];
        REAL fx[3], w[3][3], dw[3][3];
        if (or_bspline(sc, X + 3 * p, b, fx, w, dw)) { if (p < first_bad) first_bad = p; key[p] = 0; }
        else key[p] = b[0] * n + b[2];
    }
    if (first_bad < np) {
        int b[3];
        REAL fx[3], w[3][3], dw[3][3];
        int rc = or_bspline(sc, X + 3 * first_bad, b, fx, w, dw);
        free(key);
        colgroups_free(g);
        return rc;
    }
    for (int p = 0; p < np; p++) g->start[key[p] + 1]++;
    for (long i = 0; i < ng; i++) g->start[i + 1] += g->start[i];
    int* fill = (int*)malloc(sizeof(int) * (ng > 0 ? ng : 1));
    if (!fill) { free(key); colgroups_free(g); return OR_NOMEM; }
    memcpy(fill, g->start, sizeof(int) * ng);
    for (int p = 0; p < np; p++) g->list[fill[key[p]]++] = p;
    free(fill);
    free(key);
    return OR_OK;
}

typedef void (*particle_fn)(void* ctx, int p);

/* Run fn on every particle of a scatter.  ORACLE_ATOMIC: one parallel loop over the particles
 * (fn adds atomically).  Otherwise: the nine column colours in turn, the groups of one colour
 * in parallel (a group writes node columns [bx, bx+2] x [bz, bz+2] only). */
static void run_scatter(const or_scene* sc, const colgroups* g, particle_fn fn, void* ctx)
{
#ifdef ORACLE_ATOMIC
    (void)sc;
    #pragma omp parallel for schedule(static)
    for (int p = 0; p < g->np; p++) fn(ctx, p);
#else
    int n = sc->n_grid;
    for (int c = 0; c < 9; c++) {
        int cx = c / 3, cz = c % 3;
        int nx = (n - cx + 2) / 3, nz = (n - cz + 2) / 3;
        #pragma omp parallel for schedule(dynamic, 1) collapse(2)
        for (int ix = 0; ix < nx; ix++)
            for (int iz = 0; iz < nz; iz++) {
                long gi = (long)(cx + 3 * ix) * n + (cz + 3 * iz);
                for (int t = g->start[gi]; t < g->start[gi + 1]; t++) fn(ctx, g->list[t]);
            }
    }
#endif
}

/* P2G of the liquid: sf:104-163 */
typedef struct {
    const or_scene* sc;
    const double* th;
    const or_state* s;
    REAL* P;
    REAL* Fnew;
} p2g_liq_ctx;

static void p2g_liq_one(void* vc, int p)
{
    p2g_liq_ctx* c = (p2g_liq_ctx*)vc;
    const or_state* s = c->s;
    liq_pre q;
    liquid_pre(c->sc, c->th, s->lx + 3 * p, s->lv + 3 * p, s->lC + 9 * p, s->lF + 9 * p, &q);
    if (c->Fnew) memcpy(c->Fnew + 9 * p, q.Fn, sizeof(REAL) * 9);
    scatter(c->sc, c->P, q.b, q.fx, q.w, (REAL)c->sc->m_liq, q.u, q.A);
}

/* P2G of the stick and container: sf:165-196 (affine = m C, no stress, no gravity) */
typedef struct {
    const or_scene* sc;
    const REAL *X, *V, *Cm;
    REAL m;
    REAL* P;
} p2g_rigid_ctx;

static void p2g_rigid_one(void* vc, int p)
{
    p2g_rigid_ctx* c = (p2g_rigid_ctx*)vc;
    int b[3];
    REAL fx[3], w[3][3], dw[3][3], A[9];
    or_bspline(c->sc, c->X + 3 * p, b, fx, w, dw);
    for (int i = 0; i < 9; i++) A[i] = c->m * c->Cm[9 * p + i];
    scatter(c->sc, c->P, b, fx, w, c->m, c->V + 3 * p, A);
}

/* P2G for all species into P (n^3 x 4: momentum xyz, mass).  Writes F_new for liquid. */
static int p2g_all(const or_scene* sc, const double* th, const or_state* s, REAL* P, REAL* Fnew)
{
    long nn = (long)sc->n_grid * sc->n_grid * sc->n_grid;
    memset(P, 0, sizeof(REAL) * 4 * nn);
    colgroups g;
    int rc = colgroups_build(sc, s->nl, s->lx, &g);
    if (rc) return rc;
    p2g_liq_ctx lc = {sc, th, s, P, Fnew};
    run_scatter(sc, &g, p2g_liq_one, &lc);
    colgroups_free(&g);
    for (int sp = 0; sp < 2; sp++) {
        p2g_rigid_ctx rc2 = {sc, sp ? s->cx : s->sx, sp ? s->cv : s->sv, sp ? s->cC : s->sC,
                             (REAL)(sp ? sc->m_cont : sc->m_stick), P};
        rc = colgroups_build(sc, sp ? s->nc : s->ns, rc2.X, &g);
        if (rc) return rc;
        run_scatter(sc, &g, p2g_rigid_one, &rc2);
        colgroups_free(&g);
    }
    return OR_OK;
}

/* Per-node grid update; records which branches fired so the adjoint can replay them. */
typedef struct {
    REAL inv_m, v0[3];
    int wall;
    REAL nrm[3];
    int mask[3];   /* 1 = component zeroed by a box bound */
} node_rec;

static void gridop_node(const or_scene* sc, const REAL* Pn, int i, int j, int k, REAL* vout, node_rec* r)
{
    REAL inv_m = (REAL)1 / (Pn[3] + (REAL)1e-10);
    REAL v[3] = {inv_m * Pn[0], inv_m * Pn[1], inv_m * Pn[2]};
    node_rec rr;
    rr.inv_m = inv_m;
    rr.v0[0] = v[0]; rr.v0[1] = v[1]; rr.v0[2] = v[2];
    rr.wall = 0;
    rr.mask[0] = rr.mask[1] = rr.mask[2] = 0;
    REAL dx = (REAL)sc->dx;
    REAL d[3] = {(REAL)i * dx - (REAL)sc->wall_cx, 0, (REAL)k * dx - (REAL)sc->wall_cz};
    REAL r2 = (REAL)sc->wall_r * (REAL)sc->wall_r;
    REAL dd = d[0] * d[0] + d[1] * d[1] + d[2] * d[2];
    if (dd > r2) {
        REAL inv = (REAL)1 / (REAL)sqrt((double)dd);
        rr.wall = 1;
        for (int a = 0; a < 3; a++) rr.nrm[a] = d[a] * inv;
        REAL vn = v[0] * rr.nrm[0] + v[1] * rr.nrm[1] + v[2] * rr.nrm[2];
        for (int a = 0; a < 3; a++) v[a] -= rr.nrm[a] * vn;
    }
    int n = sc->n_grid, bd = sc->bound;
    if (i < bd && v[0] < 0) { v[0] = 0; rr.mask[0] = 1; }
    if (i > n - bd && v[0] > 0) { v[0] = 0; rr.mask[0] = 1; }
    if (k < bd && v[2] < 0) { v[2] = 0; rr.mask[2] = 1; }
    if (k > n - bd && v[2] > 0) { v[2] = 0; rr.mask[2] = 1; }
    if (j < bd && v[1] < 0) { v[1] = 0; rr.mask[1] = 1; }
    if (j > n - bd && v[1] > 0) { v[0] = v[1] = v[2] = 0; rr.mask[0] = rr.mask[1] = rr.mask[2] = 1; }
    vout[0] = v[0]; vout[1] = v[1]; vout[2] = v[2];
    if (r) *r = rr;
}

static void gridop_all(const or_scene* sc, const REAL* P, REAL* Vout)
{
    int n = sc->n_grid;
    #pragma omp parallel for schedule(static)
    for (int i = 0; i < n; i++)
        for (int j = 0; j < n; j++)
            for (int k = 0; k < n; k++) {
                long id = nidx(sc, i, j, k);
                gridop_node(sc, P + 4 * id, i, j, k, Vout + 3 * id, NULL);
            }
}

static void gather(const or_scene* sc, const REAL* Vout, const int b[3], const REAL fx[3],
                   const REAL w[3][3], REAL* nv, REAL* nC)
{
    for (int a = 0; a < 3; a++) nv[a] = 0;
    for (int a = 0; a < 9; a++) nC[a] = 0;
    REAL c4 = (REAL)4 * (REAL)sc->inv_dx;
    for (int i = 0; i < 3; i++)
        for (int j = 0; j < 3; j++)
            for (int k = 0; k < 3; k++) {
                REAL d[3] = {(REAL)i - fx[0], (REAL)j - fx[1], (REAL)k - fx[2]};
                REAL W = w[0][i] * w[1][j] * w[2][k];
                const REAL* g = Vout + 3 * nidx(sc, b[0] + i, b[1] + j, b[2] + k);
                for (int a = 0; a < 3; a++) {
                    nv[a] += W * g[a];
                    for (int c = 0; c < 3; c++) nC[a * 3 + c] += c4 * W * g[a] * d[c];
                }
            }
}

/* First failing particle index (np if none) of a species, and its error. */
static int first_bad_particle(const or_scene* sc, int np, const REAL* X, int* rc)
{
    int first = np;
    #pragma omp parallel for reduction(min : first)
    for (int p = 0; p < np; p++) {
        int b[3];
        REAL fx[3], w[3][3], dw[3][3];
        if (or_bspline(sc, X + 3 * p, b, fx, w, dw) && p < first) first = p;
    }
    *rc = OR_OK;
    if (first < np) {
        int b[3];
        REAL fx[3], w[3][3], dw[3][3];
        *rc = or_bspline(sc, X + 3 * first, b, fx, w, dw);
    }
    return first;
}

/* G2P for all species from state `s` into `o`; dvsum receives sum_p dv over stick particles. */
static int g2p_all(const or_scene* sc, const REAL* Vout, const or_state* s, or_state* o,
                   const double* traj_f, const double* traj_f1, double* dvsum)
{
    REAL dt = (REAL)sc->dt;
    int rc;
    first_bad_particle(sc, s->nl, s->lx, &rc);
    if (rc) return rc;
    first_bad_particle(sc, s->ns, s->sx, &rc);
    if (rc) return rc;
    first_bad_particle(sc, s->nc, s->cx, &rc);
    if (rc) return rc;
    #pragma omp parallel for schedule(static)
    for (int p = 0; p < s->nl; p++) {
        int b[3];
        REAL fx[3], w[3][3], dw[3][3], nv[3], nC[9];
        or_bspline(sc, s->lx + 3 * p, b, fx, w, dw);
        gather(sc, Vout, b, fx, w, nv, nC);
        for (int a = 0; a < 3; a++) {
            o->lv[3 * p + a] = nv[a];
            o->lx[3 * p + a] = s->lx[3 * p + a] + dt * nv[a];
        }
        memcpy(o->lC + 9 * p, nC, sizeof(REAL) * 9);
    }
    REAL vk[3];
    for (int a = 0; a < 3; a++) vk[a] = (REAL)((REAL)traj_f1[a] - (REAL)traj_f[a]) / dt;
    double* dvp = (double*)malloc(sizeof(double) * 3 * (s->ns > 0 ? s->ns : 1));
    if (!dvp) return OR_NOMEM;
    #pragma omp parallel for schedule(static)
    for (int p = 0; p < s->ns; p++) {
        int b[3];
        REAL fx[3], w[3][3], dw[3][3], nv[3], nC[9];
        or_bspline(sc, s->sx + 3 * p, b, fx, w, dw);
        gather(sc, Vout, b, fx, w, nv, nC);
        for (int a = 0; a < 3; a++) {
            dvp[3 * p + a] = (double)(vk[a] - nv[a]);
            o->sv[3 * p + a] = vk[a];
            o->sx[3 * p + a] = s->sx[3 * p + a] + vk[a] * dt;
        }
        memcpy(o->sC + 9 * p, nC, sizeof(REAL) * 9);
    }
    double acc[3] = {0, 0, 0};
    for (int p = 0; p < s->ns; p++)
        for (int a = 0; a < 3; a++) acc[a] += dvp[3 * p + a];
    free(dvp);
    #pragma omp parallel for schedule(static)
    for (int p = 0; p < s->nc; p++) {
        int b[3];
        REAL fx[3], w[3][3], dw[3][3], nv[3], nC[9];
        or_bspline(sc, s->cx + 3 * p, b, fx, w, dw);
        gather(sc, Vout, b, fx, w, nv, nC);
        for (int a = 0; a < 3; a++) {
            o->cv[3 * p + a] = 0;
            o->cx[3 * p + a] = s->cx[3 * p + a];
        }
        memcpy(o->cC + 9 * p, nC, sizeof(REAL) * 9);
    }
    if (dvsum) for (int a = 0; a < 3; a++) dvsum[a] = acc[a];
    return OR_OK;
}

static long grid_nodes(const or_scene* sc)
{
    return (long)sc->n_grid * sc->n_grid * sc->n_grid;
}

/* One substep s -> s+1 (sf:341-348).  `in` and `out` must not alias. */
int OR_SYM(step)(const or_scene* sc, const double* th, const or_state* in, or_state* out,
                 const double* traj_f, const double* traj_f1, double* dvsum)
{
    long nn = grid_nodes(sc);
    REAL* P = (REAL*)malloc(sizeof(REAL) * 4 * nn);
    REAL* V = (REAL*)malloc(sizeof(REAL) * 3 * nn);
    if (!P || !V) { free(P); free(V); return OR_NOMEM; }
    int rc = p2g_all(sc, th, in, P, out->lF);
    if (!rc) {
        gridop_all(sc, P, V);
        rc = g2p_all(sc, V, in, out, traj_f, traj_f1, dvsum);
    }
    free(P);
    free(V);
    return rc;
}

/* ------------------------------------------------------------------ adjoint */

/* G2P adjoint of one liquid particle: scatter into Vb, partial x-bar (sf:230-248 reversed) */
typedef struct {
    const or_scene* sc;
    const or_state* s;
    const or_adj* ob;
    const REAL* V;
    REAL* Vb;
    REAL* xb;
} g2p_adj_ctx;

static void g2p_adj_liq_one(void* vc, int p)
{
    g2p_adj_ctx* c = (g2p_adj_ctx*)vc;
    const or_scene* sc = c->sc;
    REAL dt = (REAL)sc->dt, inv_dx = (REAL)sc->inv_dx;
    REAL c4 = (REAL)4 * inv_dx;
    int b[3];
    REAL fx[3], w[3][3], dw[3][3];
    or_bspline(sc, c->s->lx + 3 * p, b, fx, w, dw);
    const REAL* Cb = c->ob->lC + 9 * p;
    REAL vt[3], fxb[3] = {0, 0, 0};
    for (int a = 0; a < 3; a++) vt[a] = c->ob->lv[3 * p + a] + dt * c->ob->lx[3 * p + a];
    for (int i = 0; i < 3; i++)
        for (int j = 0; j < 3; j++)
            for (int k = 0; k < 3; k++) {
                REAL d[3] = {(REAL)i - fx[0], (REAL)j - fx[1], (REAL)k - fx[2]};
                REAL W = w[0][i] * w[1][j] * w[2][k];
                REAL dW[3] = {dw[0][i] * w[1][j] * w[2][k], w[0][i] * dw[1][j] * w[2][k],
                              w[0][i] * w[1][j] * dw[2][k]};
                long id = nidx(sc, b[0] + i, b[1] + j, b[2] + k);
                const REAL* g = c->V + 3 * id;
                REAL cd[3];
                for (int a = 0; a < 3; a++)
                    cd[a] = Cb[a * 3 + 0] * d[0] + Cb[a * 3 + 1] * d[1] + Cb[a * 3 + 2] * d[2];
                for (int a = 0; a < 3; a++) GRID_ADD(c->Vb[3 * id + a], W * (vt[a] + c4 * cd[a]));
                REAL Wb = g[0] * vt[0] + g[1] * vt[1] + g[2] * vt[2]
                        + c4 * (g[0] * cd[0] + g[1] * cd[1] + g[2] * cd[2]);
                for (int cc = 0; cc < 3; cc++) {
                    REAL ctg = Cb[0 * 3 + cc] * g[0] + Cb[1 * 3 + cc] * g[1] + Cb[2 * 3 + cc] * g[2];
                    fxb[cc] += Wb * dW[cc] - c4 * W * ctg;
                }
            }
    for (int a = 0; a < 3; a++) c->xb[3 * p + a] = c->ob->lx[3 * p + a] + fxb[a] * inv_dx;
}

/* G2P adjoint of one stick (dv = v_kin - sum W v_grid, sf:250-270) or container particle */
typedef struct {
    const or_scene* sc;
    const REAL* X;
    const REAL* CB;
    const double* dvbar;   /* NULL for the container */
    REAL* Vb;
} g2p_adj_rigid_ctx;

static void g2p_adj_rigid_one(void* vc, int p)
{
    g2p_adj_rigid_ctx* c = (g2p_adj_rigid_ctx*)vc;
    const or_scene* sc = c->sc;
    REAL c4 = (REAL)4 * (REAL)sc->inv_dx;
    int b[3];
    REAL fx[3], w[3][3], dw[3][3];
    or_bspline(sc, c->X + 3 * p, b, fx, w, dw);
    const REAL* Cb = c->CB + 9 * p;
    for (int i = 0; i < 3; i++)
        for (int j = 0; j < 3; j++)
            for (int k = 0; k < 3; k++) {
                REAL d[3] = {(REAL)i - fx[0], (REAL)j - fx[1], (REAL)k - fx[2]};
                REAL W = w[0][i] * w[1][j] * w[2][k];
                long id = nidx(sc, b[0] + i, b[1] + j, b[2] + k);
                for (int a = 0; a < 3; a++) {
                    REAL cd = Cb[a * 3 + 0] * d[0] + Cb[a * 3 + 1] * d[1] + Cb[a * 3 + 2] * d[2];
                    REAL t = c4 * cd;
                    if (c->dvbar) t -= (REAL)c->dvbar[a];
                    GRID_ADD(c->Vb[3 * id + a], W * t);
                }
            }
}

/* P2G adjoint of one liquid particle: gather Pbar, stress and rheology adjoint; writes the
 * particle's share of dL/dparams into thp[4] */
static void p2g_adj_liq_one(const or_scene* sc, const double* th, const or_state* s, const or_adj* ob,
                            const REAL* Pb, const REAL* xb, or_adj* ib, int p, double* thp)
{
    REAL dx = (REAL)sc->dx, inv_dx = (REAL)sc->inv_dx, dt = (REAL)sc->dt;
    REAL m = (REAL)sc->m_liq;
    REAL alpha = (REAL)(-sc->dt * sc->vol_liq * 4.0 * sc->inv_dx * sc->inv_dx);
    double kk = bulk_k(sc, th);
    liq_pre q;
    const REAL* C = s->lC + 9 * p;
    const REAL* F = s->lF + 9 * p;
    liquid_pre(sc, th, s->lx + 3 * p, s->lv + 3 * p, C, F, &q);
    REAL ub[3] = {0, 0, 0}, Ab[9] = {0}, fxb[3] = {0, 0, 0};
    for (int i = 0; i < 3; i++)
        for (int j = 0; j < 3; j++)
            for (int k = 0; k < 3; k++) {
                REAL dp[3] = {((REAL)i - q.fx[0]) * dx, ((REAL)j - q.fx[1]) * dx, ((REAL)k - q.fx[2]) * dx};
                REAL W = q.w[0][i] * q.w[1][j] * q.w[2][k];
                REAL dW[3] = {q.dw[0][i] * q.w[1][j] * q.w[2][k], q.w[0][i] * q.dw[1][j] * q.w[2][k],
                              q.w[0][i] * q.w[1][j] * q.dw[2][k]};
                long id = nidx(sc, q.b[0] + i, q.b[1] + j, q.b[2] + k);
                const REAL* g = Pb + 4 * id;
                REAL h = g[3];
                REAL mom[3];
                for (int a = 0; a < 3; a++)
                    mom[a] = m * q.u[a] + q.A[a * 3] * dp[0] + q.A[a * 3 + 1] * dp[1] + q.A[a * 3 + 2] * dp[2];
                for (int a = 0; a < 3; a++) {
                    ub[a] += W * m * g[a];
                    for (int c = 0; c < 3; c++) Ab[a * 3 + c] += W * g[a] * dp[c];
                }
                REAL Wb = g[0] * mom[0] + g[1] * mom[1] + g[2] * mom[2] + h * m;
                for (int c = 0; c < 3; c++) {
                    REAL atg = q.A[0 * 3 + c] * g[0] + q.A[1 * 3 + c] * g[1] + q.A[2 * 3 + c] * g[2];
                    fxb[c] += Wb * dW[c] - dx * W * atg;
                }
            }
    REAL* oxb = ib->lx + 3 * p;
    REAL* ovb = ib->lv + 3 * p;
    REAL* oCb = ib->lC + 9 * p;
    REAL* oFb = ib->lF + 9 * p;
    REAL Fnb[9];
    for (int a = 0; a < 9; a++) Fnb[a] = ob->lF[9 * p + a];   /* adjoint of F[s+1] = F_tmp */
    for (int a = 0; a < 3; a++) oxb[a] = xb[3 * p + a] + fxb[a] * inv_dx;
    for (int a = 0; a < 3; a++) ovb[a] = ub[a];
    REAL Cb[9], Sb[9];
    for (int a = 0; a < 9; a++) { Cb[a] = m * Ab[a]; Sb[a] = alpha * Ab[a]; }
    /* S = eta sym(C) + I (k/2)(J - 1/J - 1) J */
    double etab = 0;
    for (int a = 0; a < 3; a++)
        for (int c = 0; c < 3; c++)
            etab += (double)Sb[a * 3 + c] * 0.5 * ((double)C[a * 3 + c] + (double)C[c * 3 + a]);
    for (int a = 0; a < 3; a++)
        for (int c = 0; c < 3; c++)
            Cb[a * 3 + c] += q.eta * (REAL)0.5 * (Sb[a * 3 + c] + Sb[c * 3 + a]);
    double trS = (double)Sb[0] + (double)Sb[4] + (double)Sb[8];
    double J = (double)q.J;
    double Jb = trS * (kk / 2.0) * (2.0 * J - 1.0);
    double deta_dg, deta_dth[4];
    visc(sc->model, th, (double)q.gamma, &deta_dg, deta_dth);
    for (int a = 0; a < 4; a++) thp[a] = etab * deta_dth[a];
    if (sc->k_from_param) thp[3] += trS * 0.5 * (J - 1.0 / J - 1.0) * J;
    double gb = etab * deta_dg;
    double n2 = 0;
    for (int a = 0; a < 9; a++) n2 += (double)C[a] * (double)C[a];
    if (n2 > 0) {
        double nrm = sqrt(n2);
        for (int a = 0; a < 9; a++) Cb[a] += (REAL)(gb * (double)C[a] / nrm);
    }
    REAL K[9];
    cof3(q.Fn, K);
    for (int a = 0; a < 9; a++) Fnb[a] += (REAL)Jb * K[a];
    /* Fn = (I + dt C) F */
    for (int a = 0; a < 3; a++)
        for (int c = 0; c < 3; c++) {
            REAL sF = 0, sC2 = 0;
            for (int k = 0; k < 3; k++) {
                REAL Mka = (k == a ? (REAL)1 : (REAL)0) + dt * C[k * 3 + a];
                sF += Mka * Fnb[k * 3 + c];           /* (I+dtC)^T Fnb */
                sC2 += Fnb[a * 3 + k] * F[c * 3 + k]; /* Fnb F^T */
            }
            oFb[a * 3 + c] = sF;
            Cb[a * 3 + c] += dt * sC2;
        }
    memcpy(oCb, Cb, sizeof(REAL) * 9);
}

/* Adjoint of one substep.  Given the state at s and the adjoints of the state at s+1
 * (`ob`) and of the stick's summed dv at s+1 (`dvbar`, the same vector for every stick
 * particle), writes the adjoints of the state at s into `ib` and adds dL/dparams into
 * th_bar.  Stick and container positions/velocities are kinematic (independent of the
 * parameters), so only their C adjoints are carried.  ob and ib may alias. */
int OR_SYM(step_adjoint)(const or_scene* sc, const double* th, const or_state* s,
                         const double* traj_f, const double* traj_f1,
                         const or_adj* ob, const double* dvbar, or_adj* ib, double* th_bar)
{
    (void)traj_f; (void)traj_f1;
    long nn = grid_nodes(sc);
    int n = sc->n_grid;
    int nl1 = s->nl > 0 ? s->nl : 1;
    REAL* P = (REAL*)malloc(sizeof(REAL) * 4 * nn);
    REAL* V = (REAL*)malloc(sizeof(REAL) * 3 * nn);
    REAL* Vb = (REAL*)calloc(3 * nn, sizeof(REAL));
    REAL* Pb = (REAL*)malloc(sizeof(REAL) * 4 * nn);
    REAL* xb = (REAL*)malloc(sizeof(REAL) * 3 * nl1);
    double* thp = (double*)malloc(sizeof(double) * 4 * nl1);
    colgroups g = {0, NULL, NULL};
    int rc = OR_OK;
    if (!P || !V || !Vb || !Pb || !xb || !thp) { rc = OR_NOMEM; goto done; }

    /* recompute the grid of this step (sf:353-358) */
    rc = p2g_all(sc, th, s, P, NULL);
    if (rc) goto done;
    gridop_all(sc, P, V);

    /* --- G2P adjoint: scatter into Vb, partial x-bar for liquid */
    {
        g2p_adj_ctx c = {sc, s, ob, V, Vb, xb};
        rc = colgroups_build(sc, s->nl, s->lx, &g);
        if (rc) goto done;
        run_scatter(sc, &g, g2p_adj_liq_one, &c);
        colgroups_free(&g);
    }
    for (int sp = 0; sp < 2; sp++) {
        g2p_adj_rigid_ctx c = {sc, sp ? s->cx : s->sx, sp ? ob->cC : ob->sC, sp ? NULL : dvbar, Vb};
        rc = colgroups_build(sc, sp ? s->nc : s->ns, c.X, &g);
        if (rc) goto done;
        run_scatter(sc, &g, g2p_adj_rigid_one, &c);
        colgroups_free(&g);
    }

    /* --- grid-op adjoint: Vb -> (Pbar momentum, Pbar mass) */
    #pragma omp parallel for schedule(static)
    for (int i = 0; i < n; i++)
        for (int j = 0; j < n; j++)
            for (int k = 0; k < n; k++) {
                long id = nidx(sc, i, j, k);
                node_rec r;
                REAL vo[3];
                gridop_node(sc, P + 4 * id, i, j, k, vo, &r);
                REAL gg[3] = {Vb[3 * id], Vb[3 * id + 1], Vb[3 * id + 2]};
                for (int a = 0; a < 3; a++) if (r.mask[a]) gg[a] = 0;
                if (r.wall) {
                    REAL gn = gg[0] * r.nrm[0] + gg[1] * r.nrm[1] + gg[2] * r.nrm[2];
                    for (int a = 0; a < 3; a++) gg[a] -= r.nrm[a] * gn;
                }
                for (int a = 0; a < 3; a++) Pb[4 * id + a] = r.inv_m * gg[a];
                Pb[4 * id + 3] = -r.inv_m * (gg[0] * r.v0[0] + gg[1] * r.v0[1] + gg[2] * r.v0[2]);
            }

    /* --- P2G adjoint: gather Pbar (positions were checked by the scatter's grouping) */
    #pragma omp parallel for schedule(static)
    for (int p = 0; p < s->nl; p++) p2g_adj_liq_one(sc, th, s, ob, Pb, xb, ib, p, thp + 4 * p);
    for (int p = 0; p < s->nl; p++)
        for (int a = 0; a < 4; a++) th_bar[a] += thp[4 * p + a];
    for (int sp = 0; sp < 2; sp++) {
        int np = sp ? s->nc : s->ns;
        const REAL* X = sp ? s->cx : s->sx;
        REAL* CB = sp ? ib->cC : ib->sC;
        REAL m = (REAL)(sp ? sc->m_cont : sc->m_stick);
        REAL dx = (REAL)sc->dx;
        #pragma omp parallel for schedule(static)
        for (int p = 0; p < np; p++) {
            int b[3];
            REAL fx[3], w[3][3], dw[3][3];
            or_bspline(sc, X + 3 * p, b, fx, w, dw);
            REAL Cb[9] = {0};
            for (int i = 0; i < 3; i++)
                for (int j = 0; j < 3; j++)
                    for (int k = 0; k < 3; k++) {
                        REAL dp[3] = {((REAL)i - fx[0]) * dx, ((REAL)j - fx[1]) * dx, ((REAL)k - fx[2]) * dx};
                        REAL W = w[0][i] * w[1][j] * w[2][k];
                        const REAL* gp = Pb + 4 * nidx(sc, b[0] + i, b[1] + j, b[2] + k);
                        for (int a = 0; a < 3; a++)
                            for (int c = 0; c < 3; c++) Cb[a * 3 + c] += W * gp[a] * dp[c];
                    }
            for (int a = 0; a < 9; a++) CB[9 * p + a] = m * Cb[a];
        }
    }
done:
    colgroups_free(&g);
    free(P); free(V); free(Vb); free(Pb); free(xb); free(thp);
    return rc;
}

/* ------------------------------------------------------------------ rollouts */

typedef struct {
    or_state st;
    REAL* buf;
} state_buf;

static int state_alloc(state_buf* b, int nl, int ns, int nc)
{
    size_t tot = (size_t)nl * 24 + (size_t)ns * 15 + (size_t)nc * 15;
    b->buf = (REAL*)calloc(tot ? tot : 1, sizeof(REAL));
    if (!b->buf) return OR_NOMEM;
    REAL* q = b->buf;
    b->st.nl = nl; b->st.ns = ns; b->st.nc = nc;
    b->st.lx = q; q += 3 * (size_t)nl;
    b->st.lv = q; q += 3 * (size_t)nl;
    b->st.lC = q; q += 9 * (size_t)nl;
    b->st.lF = q; q += 9 * (size_t)nl;
    b->st.sx = q; q += 3 * (size_t)ns;
    b->st.sv = q; q += 3 * (size_t)ns;
    b->st.sC = q; q += 9 * (size_t)ns;
    b->st.cx = q; q += 3 * (size_t)nc;
    b->st.cv = q; q += 3 * (size_t)nc;
    b->st.cC = q; q += 9 * (size_t)nc;
    return OR_OK;
}

static void state_copy(or_state* d, const or_state* s)
{
    memcpy(d->lx, s->lx, sizeof(REAL) * 3 * s->nl);
    memcpy(d->lv, s->lv, sizeof(REAL) * 3 * s->nl);
    memcpy(d->lC, s->lC, sizeof(REAL) * 9 * s->nl);
    memcpy(d->lF, s->lF, sizeof(REAL) * 9 * s->nl);
    memcpy(d->sx, s->sx, sizeof(REAL) * 3 * s->ns);
    memcpy(d->sv, s->sv, sizeof(REAL) * 3 * s->ns);
    memcpy(d->sC, s->sC, sizeof(REAL) * 9 * s->ns);
    memcpy(d->cx, s->cx, sizeof(REAL) * 3 * s->nc);
    memcpy(d->cv, s->cv, sizeof(REAL) * 3 * s->nc);
    memcpy(d->cC, s->cC, sizeof(REAL) * 9 * s->nc);
}

/* Forward rollout of nsteps substeps from `init` (the state at step 0).
 * traj has nsteps+1 rows; forces has nsteps+1 rows (forces[0] = 0, sf:330-336).
 * If `fin` is given it receives the state at step nsteps. */
int OR_SYM(rollout)(const or_scene* sc, const double* th, const or_state* init, const double* traj,
                    int nsteps, double* forces, or_state* fin)
{
    state_buf a, b;
    if (state_alloc(&a, init->nl, init->ns, init->nc)) return OR_NOMEM;
    if (state_alloc(&b, init->nl, init->ns, init->nc)) { free(a.buf); return OR_NOMEM; }
    state_copy(&a.st, init);
    or_state* cur = &a.st;
    or_state* nxt = &b.st;
    double fs = sc->m_stick / sc->dt;
    forces[0] = forces[1] = forces[2] = 0;
    int rc = OR_OK;
    for (int s = 0; s < nsteps; s++) {
        double dv[3];
        rc = OR_SYM(step)(sc, th, cur, nxt, traj + 3 * s, traj + 3 * (s + 1), dv);
        if (rc) break;
        for (int a2 = 0; a2 < 3; a2++) forces[3 * (s + 1) + a2] = fs * dv[a2];
        or_state* t = cur; cur = nxt; nxt = t;
    }
    if (!rc && fin) state_copy(fin, cur);
    free(a.buf);
    free(b.buf);
    return rc;
}

/* Forward rollout plus adjoint with checkpoints every `ckpt` steps and segment recompute.
 * dLdF has nsteps+1 rows (adjoint of forces); th_bar[4] receives dL/dparams (overwritten). */
int OR_SYM(rollout_grad)(const or_scene* sc, const double* th, const or_state* init, const double* traj,
                         int nsteps, int ckpt, const double* dLdF, double* forces, double* th_bar)
{
    if (ckpt < 1) return OR_BADARG;
    int nl = init->nl, ns = init->ns, nc = init->nc;
    int ncp = nsteps / ckpt + 1;
    state_buf* cp = (state_buf*)calloc(ncp, sizeof(state_buf));
    state_buf* seg = (state_buf*)calloc(ckpt + 1, sizeof(state_buf));
    int rc = OR_OK;
    if (!cp || !seg) { free(cp); free(seg); return OR_NOMEM; }
    for (int i = 0; i < ncp && !rc; i++) rc = state_alloc(&cp[i], nl, ns, nc);
    for (int i = 0; i <= ckpt && !rc; i++) rc = state_alloc(&seg[i], nl, ns, nc);
    state_buf ab;
    memset(&ab, 0, sizeof ab);
    if (!rc) rc = state_alloc(&ab, nl, ns, nc);
    if (rc) goto out;

    /* forward with checkpoints */
    {
        double fs = sc->m_stick / sc->dt;
        forces[0] = forces[1] = forces[2] = 0;
        state_copy(&seg[0].st, init);
        or_state* cur = &seg[0].st;
        or_state* nxt = &seg[1].st;
        for (int s = 0; s < nsteps; s++) {
            if (s % ckpt == 0) state_copy(&cp[s / ckpt].st, cur);
            double dv[3];
            rc = OR_SYM(step)(sc, th, cur, nxt, traj + 3 * s, traj + 3 * (s + 1), dv);
            if (rc) goto out;
            for (int a = 0; a < 3; a++) forces[3 * (s + 1) + a] = fs * dv[a];
            or_state* t = cur; cur = nxt; nxt = t;
        }
    }
    /* adjoint, last segment first */
    {
        for (int a = 0; a < 4; a++) th_bar[a] = 0;
        or_adj A = {ab.st.lx, ab.st.lv, ab.st.lC, ab.st.lF, ab.st.sC, ab.st.cC};
        double fs = sc->m_stick / sc->dt;
        for (int c = (nsteps - 1) / ckpt; c >= 0; c--) {
            int s0 = c * ckpt;
            int s1 = s0 + ckpt < nsteps ? s0 + ckpt : nsteps;
            state_copy(&seg[0].st, &cp[c].st);
            for (int s = s0; s < s1 - 1; s++) {
                rc = OR_SYM(step)(sc, th, &seg[s - s0].st, &seg[s - s0 + 1].st, traj + 3 * s, traj + 3 * (s + 1), NULL);
                if (rc) goto out;
            }
            for (int s = s1 - 1; s >= s0; s--) {
                double dvb[3];
                for (int a = 0; a < 3; a++) dvb[a] = fs * dLdF[3 * (s + 1) + a];
                rc = OR_SYM(step_adjoint)(sc, th, &seg[s - s0].st, traj + 3 * s, traj + 3 * (s + 1), &A, dvb, &A, th_bar);
                if (rc) goto out;
            }
        }
    }
out:
    if (cp) for (int i = 0; i < ncp; i++) free(cp[i].buf);
    if (seg) for (int i = 0; i <= ckpt; i++) free(seg[i].buf);
    free(ab.buf);
    free(cp);
    free(seg);
    return rc;
}

/* Single grid-node update, exposed for the brute-force branch test of grid_op. */
void OR_SYM(gridop_one)(const or_scene* sc, const REAL* Pn, int i, int j, int k, REAL* vout)
{
    gridop_node(sc, Pn, i, j, k, vout, NULL);
}

/* Viscosity and partials, exposed for closed-form tests. */
double OR_SYM(eta)(int model, const double* th, double gamma, double* d_dg, double* d_dth)
{
    return visc(model, th, gamma, d_dg, d_dth);
}

/* P2G only (grid momentum+mass), exposed for the conservation tests. */
int OR_SYM(p2g)(const or_scene* sc, const double* th, const or_state* s, REAL* P)
{
    return p2g_all(sc, th, s, P, NULL);
}

/* G2P gather of a given velocity grid for arbitrary positions, exposed for the
 * affine-field reconstruction test: nv[n*3], nC[n*9]. */
int OR_SYM(gather)(const or_scene* sc, const REAL* Vout, int n, const REAL* x, REAL* nv, REAL* nC)
{
    int b[3];
    REAL fx[3], w[3][3], dw[3][3];
    for (int p = 0; p < n; p++) {
        int rc = or_bspline(sc, x + 3 * p, b, fx, w, dw);
        if (rc) return rc;
        gather(sc, Vout, b, fx, w, nv + 3 * p, nC + 9 * p);
    }
    return OR_OK;
}

/* Threads of the OpenMP loops for the calls that follow (n <= 0: leave the current setting);
 * returns the number the next parallel loop will use (1 without OpenMP). */
int OR_SYM(set_threads)(int n)
{
#ifdef _OPENMP
    if (n > 0) omp_set_num_threads(n);
    return omp_get_max_threads();
#else
    (void)n;
    return 1;
#endif
}
