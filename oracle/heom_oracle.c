/*
 * heom_oracle.c -- CPU FP64 restatement of the reference ETO-HEOM propagation.
 *
 * TEST INFRASTRUCTURE ONLY (the checker, and the timed CPU baseline "port"):
 * nothing in the product library links or calls this file.
 *
 * Every function names the reference lines it follows (paths relative to the
 * reference repository root).  Arithmetic follows the reference operation by
 * operation where the order is visible in its source (field, time grid, RK4
 * in-place axpy sequence); GSL's zgemm/zaxpy are restated as the textbook CBLAS
 * loops (beta handling, then i-k-j accumulation of alpha*A[i][k]*B[k][j] with
 * zero terms skipped), which is how the CBLAS reference implementation orders
 * them.  Rounding-order differences against GSL are ~1e-16 relative.
 */
#include "heom_oracle.h"

#include <math.h>
#include <stdlib.h>
#include <string.h>

typedef struct { double re, im; } cx;

static inline cx cmul(cx a, cx b) { cx r = {a.re * b.re - a.im * b.im, a.re * b.im + a.im * b.re}; return r; }
static inline cx cadd(cx a, cx b) { cx r = {a.re + b.re, a.im + b.im}; return r; }
static inline cx cmulr(cx a, double x) { cx r = {a.re * x, a.im * x}; return r; }

static const cx ONE = {1.0, 0.0};
static const cx ZERO = {0.0, 0.0};
static const cx MINUS_ONE = {-1.0, 0.0};
static const cx ICNT = {0.0, 1.0};
static const cx MINUS_ICNT = {0.0, -1.0};

/* C = alpha*A*B + beta*C, all n x n row-major (gsl_blas_zgemm NoTrans/NoTrans) */
static void zgemm(int n, cx alpha, const cx* A, const cx* B, cx beta, cx* C) {
    if (beta.re == 0.0 && beta.im == 0.0) {
        for (int i = 0; i < n * n; i++) C[i] = ZERO;
    } else if (!(beta.re == 1.0 && beta.im == 0.0)) {
        for (int i = 0; i < n * n; i++) C[i] = cmul(beta, C[i]);
    }
    for (int i = 0; i < n; i++) {
        for (int k = 0; k < n; k++) {
            cx t = cmul(alpha, A[i * n + k]);
            if (t.re == 0.0 && t.im == 0.0) continue;
            for (int j = 0; j < n; j++) C[i * n + j] = cadd(C[i * n + j], cmul(t, B[k * n + j]));
        }
    }
}

/* y += alpha*x over len complex entries (cblas_zaxpy) */
static void zaxpy(long len, cx alpha, const cx* x, cx* y) {
    for (long i = 0; i < len; i++) y[i] = cadd(y[i], cmul(alpha, x[i]));
}

/* ---- hierarchy (utilize.cpp:3-36) ------------------------------------- */

static void dfs(int d, int L, unsigned char* cur, int pos, int sum, unsigned char* out, long cap, long* m) {
    if (pos == d) {
        if (sum <= L) {
            if (*m < cap && out) memcpy(out + (*m) * d, cur, (size_t)d);
            (*m)++;
        }
        return;
    }
    for (int c = 0; c <= L; c++) {
        if (sum + c <= L) {
            cur[pos] = (unsigned char)c;
            dfs(d, L, cur, pos + 1, sum + c, out, cap, m);
        }
    }
}

long oh_build_ado(int d, int L, unsigned char* out, long cap) {
    if (d < 0 || L < 0 || d > 255) return -1;
    unsigned char cur[256];
    long m = 0;
    if (d == 0) return 1;
    dfs(d, L, cur, 0, 0, out, cap, &m);
    return m;
}

long oh_num_ado(int d, int L) { return oh_build_ado(d, L, NULL, 0); }

/* string -> index map (utilize.cpp:20-24); open addressing on an FNV-1a hash */
typedef struct {
    int d;
    long cap;        /* power of two */
    long* slot;      /* ADO index or -1 */
    const unsigned char* ado;
} ado_map;

static unsigned long long fnv(const unsigned char* s, int d) {
    unsigned long long h = 1469598103934665603ULL;
    for (int i = 0; i < d; i++) { h ^= s[i]; h *= 1099511628211ULL; }
    return h;
}

static int map_init(ado_map* mp, const unsigned char* ado, long M, int d) {
    mp->d = d;
    mp->ado = ado;
    mp->cap = 1;
    while (mp->cap < 2 * M + 2) mp->cap <<= 1;
    mp->slot = (long*)malloc(sizeof(long) * (size_t)mp->cap);
    if (!mp->slot) return -1;
    for (long i = 0; i < mp->cap; i++) mp->slot[i] = -1;
    for (long i = 0; i < M; i++) {
        long h = (long)(fnv(ado + i * d, d) & (unsigned long long)(mp->cap - 1));
        while (mp->slot[h] >= 0) h = (h + 1) & (mp->cap - 1);
        mp->slot[h] = i;
    }
    return 0;
}

static long map_find(const ado_map* mp, const unsigned char* key) {
    long h = (long)(fnv(key, mp->d) & (unsigned long long)(mp->cap - 1));
    while (mp->slot[h] >= 0) {
        if (memcmp(mp->ado + mp->slot[h] * mp->d, key, (size_t)mp->d) == 0) return mp->slot[h];
        h = (h + 1) & (mp->cap - 1);
    }
    return -1;
}

/* ---- solver context -------------------------------------------------- */

typedef struct {
    const oh_problem* p;
    int N, d;
    long M, nn;
    unsigned char* ado;
    ado_map map;
    cx** S;          /* K one-hot matrices */
} ctx_t;

static void ctx_free(ctx_t* c) {
    free(c->ado);
    free(c->map.slot);
    if (c->S) {
        for (int j = 0; j < c->p->K; j++) free(c->S[j]);
        free(c->S);
    }
}

static int ctx_init(ctx_t* c, const oh_problem* p) {
    memset(c, 0, sizeof(*c));
    c->p = p;
    c->N = p->N;
    c->d = p->K * p->Km;
    c->nn = (long)p->N * p->N;
    if (p->N <= 0 || p->K < 0 || p->Km < 0 || p->L < 0 || c->d > 255) return -1;
    c->M = oh_num_ado(c->d, p->L);
    c->ado = (unsigned char*)malloc((size_t)(c->M * (c->d > 0 ? c->d : 1)));
    if (!c->ado) return -2;
    oh_build_ado(c->d, p->L, c->ado, c->M);
    if (map_init(&c->map, c->ado, c->M, c->d)) return -2;
    c->S = (cx**)calloc((size_t)(p->K > 0 ? p->K : 1), sizeof(cx*));
    for (int j = 0; j < p->K; j++) {
        int s = j + p->site_offset;
        if (s < 0 || s >= p->N) return -3;
        c->S[j] = (cx*)calloc((size_t)c->nn, sizeof(cx));
        c->S[j][s * p->N + s] = ONE;   /* param.cpp:190-195 (2DES), :310-315 (dynamics) */
    }
    return 0;
}

static cx getc_(const double* a, long i) { cx r = {a[2 * i], a[2 * i + 1]}; return r; }

/* total_ADO_dynamics_Ht, dynamics.cpp:117-179 (total_ADO_dynamics :53-113 is the same with V=H) */
static void rhs(const ctx_t* c, const cx* V, cx* const* rho, cx** drho) {
    const oh_problem* p = c->p;
    const int N = c->N, K = p->K, Km = p->Km, d = c->d;
    unsigned char tmp[256];
    for (long i = 0; i < c->M; i++) {
        const unsigned char* n = c->ado + i * d;
        /* L_s: drho = i rho V - i V rho   (:126-127) */
        zgemm(N, ICNT, rho[i], V, ZERO, drho[i]);
        zgemm(N, MINUS_ICNT, V, rho[i], ONE, drho[i]);
        /* damping -(sum n_p gamma_p) rho   (:129-140) */
        if (i != 0) {
            cx x = ZERO;
            for (int j = 0; j < K; j++)
                for (int k = 0; k < Km; k++) {
                    int off = j + k * K;
                    if ((int)n[off] == 0) continue;
                    cx y = cmul(MINUS_ONE, cmulr(getc_(p->gamma, j * Km + k), (double)(int)n[off]));
                    x = cadd(x, y);
                }
            zaxpy(c->nn, x, rho[i], drho[i]);
        }
        /* up couplings  x ( -i rho_{n+} S + i S rho_{n+} ),  x = sqrt((n_p+1)|alpha|)   (:143-156) */
        for (int j = 0; j < K; j++)
            for (int k = 0; k < Km; k++) {
                int off = j + k * K;
                if ((int)n[off] >= p->L) continue;
                memcpy(tmp, n, (size_t)d);
                tmp[off] += 1;
                long m = map_find(&c->map, tmp);
                if (m >= 0) {
                    cx a = getc_(p->alpha, j * Km + k);
                    double x = sqrt(((double)n[off] + 1) * hypot(a.re, a.im));
                    zgemm(N, cmulr(MINUS_ICNT, x), rho[m], c->S[j], ONE, drho[i]);
                    zgemm(N, cmulr(ICNT, x), c->S[j], rho[m], ONE, drho[i]);
                }
            }
        /* down couplings  y ( i alpha S rho_{n-} - i alpha~ rho_{n-} S ),  y = sqrt(n_p/|alpha|)   (:160-179) */
        if (i != 0) {
            for (int j = 0; j < K; j++)
                for (int k = 0; k < Km; k++) {
                    int off = j + k * K;
                    if ((int)n[off] <= 0) continue;
                    memcpy(tmp, n, (size_t)d);
                    tmp[off] -= 1;
                    long m = map_find(&c->map, tmp);
                    if (m >= 0) {
                        cx a = getc_(p->alpha, j * Km + k);
                        cx at = getc_(p->alpha_t, j * Km + k);
                        double y = sqrt((double)n[off] / hypot(a.re, a.im));
                        cx x = cmul(ICNT, cmulr(a, y));
                        zgemm(N, x, c->S[j], rho[m], ONE, drho[i]);
                        x = cmul(MINUS_ICNT, cmulr(at, y));
                        zgemm(N, x, rho[m], c->S[j], ONE, drho[i]);
                    }
                }
        }
    }
}

int oh_rhs(const oh_problem* p, const double* V, const double* rho, double* drho) {
    ctx_t c;
    int rc = ctx_init(&c, p);
    if (rc) { ctx_free(&c); return rc; }
    cx** r = (cx**)malloc(sizeof(cx*) * (size_t)c.M);
    cx** dr = (cx**)malloc(sizeof(cx*) * (size_t)c.M);
    for (long i = 0; i < c.M; i++) {
        r[i] = (cx*)(void*)(rho + 2 * i * c.nn);
        dr[i] = (cx*)(void*)(drho + 2 * i * c.nn);
    }
    rhs(&c, (const cx*)(const void*)V, r, dr);
    free(r);
    free(dr);
    ctx_free(&c);
    return 0;
}

/* build_V_matrix / build_V_dagger_matrix, TD_hamiltonian.cpp:5-31: returns e = E(t) e^{+-i w (t - tau0)} */
static cx pulse_scalar(const oh_problem* p, float t, int idx, int dagger) {
    const double E0 = p->pulse[idx][0], tau0 = p->pulse[idx][1], taup = p->pulse[idx][2], w0 = p->pulse[idx][3];
    double E = E0 * exp(-2.77 * (t - tau0) * (t - tau0) / taup / taup);
    double th = dagger ? -w0 * (t - tau0) : w0 * (t - tau0);
    cx e = {E * cos(th), E * sin(th)};   /* gsl_complex_polar */
    return e;
}

/* V = H - active pulses, dynamics.cpp:348-363 (X[1], X[2] transposed, :305-308) */
static void build_V(const oh_problem* p, float t, int nv1, int nv2, int nv3, cx* V) {
    const int N = p->N;
    const long nn = (long)N * N;
    for (long i = 0; i < nn; i++) V[i] = getc_(p->H, i);
    const int nv[3] = {nv1, nv2, nv3};
    for (int q = 0; q < 3; q++) {
        if (!(nv[q] == 1 && fabs(t - p->pulse[q][1]) <= p->pulse[q][2] * 2)) continue;
        cx e = pulse_scalar(p, t, q, q != 0);
        const double* X = p->X + 2 * nn * q;
        for (int a = 0; a < N; a++)
            for (int b = 0; b < N; b++) {
                cx x = (q == 0) ? getc_(X, (long)a * N + b) : getc_(X, (long)b * N + a);
                cx v = cmul(x, e);  /* gsl_matrix_complex_scale */
                V[a * N + b].re -= v.re;
                V[a * N + b].im -= v.im;
            }
    }
}

int oh_field(const oh_problem* p, float t, int nv1, int nv2, int nv3, double* V) {
    build_V(p, t, nv1, nv2, nv3, (cx*)(void*)V);
    return 0;
}

/* state helpers */
static cx** alloc_state(long M, long nn) {
    cx** s = (cx**)malloc(sizeof(cx*) * (size_t)M);
    for (long i = 0; i < M; i++) s[i] = (cx*)calloc((size_t)nn, sizeof(cx));
    return s;
}
static void free_state(cx** s, long M) {
    for (long i = 0; i < M; i++) free(s[i]);
    free(s);
}

/* the in-place RK4 step of dynamics.cpp:385-418 (identical in :256-289) */
static void rk4_step(const ctx_t* c, const cx* V, cx** rho, cx** k1, cx** k2, cx** k3, cx** k4, float step_size) {
    const double b[4] = {1.0 / 6, 2.0 / 6, 2.0 / 6, 1.0 / 6};
    const long M = c->M, nn = c->nn;
    cx x;
    rhs(c, V, rho, k1);
    x.re = step_size * 0.5; x.im = 0.0;
    for (long j = 0; j < M; j++) zaxpy(nn, x, k1[j], rho[j]);
    rhs(c, V, rho, k2);
    x = cmulr(x, -1);
    for (long j = 0; j < M; j++) zaxpy(nn, x, k1[j], rho[j]);

    x.re = step_size * 0.5; x.im = 0.0;
    for (long j = 0; j < M; j++) zaxpy(nn, x, k2[j], rho[j]);
    rhs(c, V, rho, k3);
    x = cmulr(x, -1);
    for (long j = 0; j < M; j++) zaxpy(nn, x, k2[j], rho[j]);

    x.re = step_size; x.im = 0.0;
    for (long j = 0; j < M; j++) zaxpy(nn, x, k3[j], rho[j]);
    rhs(c, V, rho, k4);
    x = cmulr(x, -1);
    for (long j = 0; j < M; j++) zaxpy(nn, x, k3[j], rho[j]);

    x.re = step_size * b[0]; x.im = 0.0;
    for (long j = 0; j < M; j++) zaxpy(nn, x, k1[j], rho[j]);
    x.re = step_size * b[1];
    for (long j = 0; j < M; j++) zaxpy(nn, x, k2[j], rho[j]);
    x.re = step_size * b[2];
    for (long j = 0; j < M; j++) zaxpy(nn, x, k3[j], rho[j]);
    x.re = step_size * b[3];
    for (long j = 0; j < M; j++) zaxpy(nn, x, k4[j], rho[j]);
}

static void init_rho(const ctx_t* c, cx** rho, const double* rho_init) {
    for (long i = 0; i < c->M; i++) {
        if (rho_init) {
            memcpy(rho[i], rho_init + 2 * i * c->nn, sizeof(cx) * (size_t)c->nn);
        } else {
            for (long e = 0; e < c->nn; e++) rho[i][e] = ZERO;
        }
    }
    if (!rho_init) rho[0][0] = ONE;   /* reset_rho, param.cpp:350-353 */
}

int oh_2des(const oh_problem* p, int nv1, int nv2, int nv3, const double* rho_init,
            long max_steps, double* P_out, long cap, long* n_points, double* rho_out) {
    ctx_t c;
    int rc = ctx_init(&c, p);
    if (rc || !p->X) { ctx_free(&c); return rc ? rc : -4; }
    const int N = c.N;
    const long M = c.M, nn = c.nn;
    cx** rho = alloc_state(M, nn);
    cx** k1 = alloc_state(M, nn);
    cx** k2 = alloc_state(M, nn);
    cx** k3 = alloc_state(M, nn);
    cx** k4 = alloc_state(M, nn);
    cx* V = (cx*)calloc((size_t)nn, sizeof(cx));
    cx* P = (cx*)calloc((size_t)nn, sizeof(cx));
    init_rho(&c, rho, rho_init);
    const cx* X3 = (const cx*)(const void*)(p->X + 2 * nn * 3);

    /* time grid, dynamics.cpp:344-347 and :418 (float arithmetic) */
    float nstep = (p->t_end - p->t_start) / p->step_size;
    float t = p->t_start;
    int print_step = (int)(p->print_step / p->step_size);
    long np = 0;
    if (print_step <= 0) rc = -5;
    for (int i = 0; rc == 0 && i <= nstep; i++) {
        if (max_steps >= 0 && i >= max_steps) break;
        build_V(p, t, nv1, nv2, nv3, V);
        /* readout before the step (:367-381) */
        if (t >= p->pulse[2][1] && i % print_step == 0) {
            zgemm(N, ONE, rho[0], X3, ZERO, P);
            cx tr = ZERO;
            for (int j = 0; j < N; j++) tr = cadd(tr, P[j * N + j]);
            if (np < cap && P_out) { P_out[2 * np] = tr.re; P_out[2 * np + 1] = tr.im; }
            np++;
        }
        rk4_step(&c, V, rho, k1, k2, k3, k4, p->step_size);
        t += p->step_size;
    }
    if (n_points) *n_points = np;
    if (rho_out)
        for (long i = 0; i < M; i++) memcpy(rho_out + 2 * i * nn, rho[i], sizeof(cx) * (size_t)nn);
    free_state(rho, M); free_state(k1, M); free_state(k2, M); free_state(k3, M); free_state(k4, M);
    free(V); free(P);
    ctx_free(&c);
    return rc;
}

int oh_dynamics(const oh_problem* p, const int* sites, int nsites, const double* rho_init,
                long max_steps, double* pop_out, long cap, long* n_points, double* rho_out) {
    ctx_t c;
    int rc = ctx_init(&c, p);
    if (rc) { ctx_free(&c); return rc; }
    const int N = c.N;
    const long M = c.M, nn = c.nn;
    /* site checks, dynamics.cpp:206-215 (error code instead of exit) */
    if (nsites > N) rc = -6;
    for (int s = 0; rc == 0 && s < nsites; s++)
        if (sites[s] > N || sites[s] < 1) rc = -6;  /* site 0 would read row -1 in the reference */
    cx** rho = alloc_state(M, nn);
    cx** k1 = alloc_state(M, nn);
    cx** k2 = alloc_state(M, nn);
    cx** k3 = alloc_state(M, nn);
    cx** k4 = alloc_state(M, nn);
    init_rho(&c, rho, rho_init);
    const cx* H = (const cx*)(const void*)p->H;
    if (pop_out)
        for (long e = 0; e < (long)nsites * cap; e++) pop_out[e] = 0.0;

    float nstep = (p->t_end - p->t_start) / p->step_size;
    float t = p->t_start;
    int print_step = (int)(p->print_step / p->step_size);
    long ct = 0;
    if (rc == 0 && print_step <= 0) rc = -5;
    for (int i = 0; rc == 0 && i <= nstep; i++) {
        if (max_steps >= 0 && i >= max_steps) break;
        if (i % print_step == 0) {   /* dynamics.cpp:246-253 */
            for (int s = 0; s < nsites; s++) {
                cx r = rho[0][(sites[s] - 1) * N + (sites[s] - 1)];
                if (ct < cap && pop_out) pop_out[(long)s * cap + ct] += r.re;
            }
            ct++;
        }
        rk4_step(&c, H, rho, k1, k2, k3, k4, p->step_size);
        t += p->step_size;
    }
    (void)t;
    if (n_points) *n_points = ct;
    if (rho_out)
        for (long i = 0; i < M; i++) memcpy(rho_out + 2 * i * nn, rho[i], sizeof(cx) * (size_t)nn);
    free_state(rho, M); free_state(k1, M); free_state(k2, M); free_state(k3, M); free_state(k4, M);
    ctx_free(&c);
    return rc;
}
