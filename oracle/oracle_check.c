/*
 * oracle_check.c -- TEST INFRASTRUCTURE ONLY: a driver of the oracle
 * (heom_oracle.c) for its address / undefined-behaviour sanitizer build
 * (`make sanitize`, run by tests/test_sanitizers.py).  It runs the ADO
 * enumeration, one RHS, a 2DES propagation through three pulses (both with the
 * default initial state and a given one) and a population-dynamics propagation of
 * a small dimer, and checks that every result is finite.  A sanitizer finding
 * aborts with a nonzero status.
 */
#include <math.h>
#include <stdio.h>
#include <stdlib.h>
#include <string.h>

#include "heom_oracle.h"

static int finite_all(const double* v, long n) {
    for (long i = 0; i < n; i++)
        if (!isfinite(v[i])) return 0;
    return 1;
}

int main(void) {
    enum { N = 4, K = 2, KM = 2, L = 3 };
    const double hb = 5308.0;
    double H[2 * N * N], X[2 * 4 * N * N], al[2 * K * KM], alt[2 * K * KM], ga[2 * K * KM];
    memset(H, 0, sizeof H);
    memset(X, 0, sizeof X);
    const double e[N] = {0.0, 12000.0, 12100.0, 24100.0};
    for (int i = 0; i < N; i++) H[2 * (i * N + i)] = e[i] / hb;
    H[2 * (1 * N + 2)] = H[2 * (2 * N + 1)] = 100.0 / hb;
    H[2 * (1 * N + 3)] = H[2 * (3 * N + 1)] = 100.0 / hb;
    for (int q = 0; q < 4; q++) {
        X[2 * (q * N * N + 0 * N + 1)] = 1.0;
        X[2 * (q * N * N + 0 * N + 2)] = 0.5;
        X[2 * (q * N * N + 2 * N + 3)] = 1.0;
        X[2 * (q * N * N + 1 * N + 3)] = 0.5;
    }
    for (int j = 0; j < K * KM; j++) {
        al[2 * j] = 3.3e-5;
        al[2 * j + 1] = -1.5e-5;
        alt[2 * j] = 3.3e-5;
        alt[2 * j + 1] = 1.5e-5;
        ga[2 * j] = 0.023;
        ga[2 * j + 1] = (j & 1) ? -0.004 : 0.004;
    }
    oh_problem p;
    memset(&p, 0, sizeof p);
    p.N = N;
    p.K = K;
    p.Km = KM;
    p.L = L;
    p.site_offset = 1;
    p.H = H;
    p.X = X;
    p.alpha = al;
    p.alpha_t = alt;
    p.gamma = ga;
    for (int q = 0; q < 3; q++) {
        p.pulse[q][0] = 1.0 / hb;
        p.pulse[q][1] = 5.0 + 10.0 * q;
        p.pulse[q][2] = 3.0;
        p.pulse[q][3] = 12050.0 / hb;
    }
    p.t_start = 0.0f;
    p.t_end = 60.0f;
    p.step_size = 0.5f;
    p.print_step = 0.5f;

    const long M = oh_num_ado(K * KM, L);
    unsigned char* ado = malloc((size_t)M * K * KM);
    if (oh_build_ado(K * KM, L, ado, M) != M) return 2;
    const long S = M * N * N;
    double* rho = calloc((size_t)2 * S, sizeof(double));
    double* drho = calloc((size_t)2 * S, sizeof(double));
    double* out = calloc((size_t)2 * S, sizeof(double));
    for (long i = 0; i < 2 * S; i++) rho[i] = 1e-3 * (double)((i * 7919) % 101 - 50);
    double V[2 * N * N];
    if (oh_field(&p, 7.0f, 1, 1, 1, V) || oh_rhs(&p, V, rho, drho) || !finite_all(drho, 2 * S)) return 3;

    double P[2 * 200];
    long np = 0;
    if (oh_2des(&p, 1, 1, 1, NULL, -1, P, 200, &np, out) || np <= 0 || !finite_all(P, 2 * np) ||
        !finite_all(out, 2 * S))
        return 4;
    if (oh_2des(&p, 1, 0, 1, rho, 40, P, 200, &np, out) || !finite_all(out, 2 * S)) return 5;
    p.site_offset = 0;
    const int sites[2] = {1, 2};
    double pop[2 * 200];
    if (oh_dynamics(&p, sites, 2, NULL, -1, pop, 200, &np, out) || np <= 0 || !finite_all(pop, 2 * np)) return 6;
    free(ado);
    free(rho);
    free(drho);
    free(out);
    printf("oracle_check ok: %ld ADOs, %ld readouts\n", M, np);
    return 0;
}
