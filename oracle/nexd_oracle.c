/*
 * TEST INFRASTRUCTURE ONLY -- the CPU oracle.  A plain-C FP64 restatement of
 * the reference's time loop, element by element, in the reference's own
 * algorithm (81 mat-vecs per element for the volume term, explicit 9x9 flux
 * matrices VT = T*(A+|A|)*T^-1, LIFT mat-vecs, separate RK pass).  Used only
 * by tests/, the smoke check and bench.py's cpu_baseline leg.
 *
 * Followed line by line:
 *   time loop            src/timeloopMod.f90:184-255 (initial state, RK choice),
 *                        :257-282 (flux matrices, anelasticvar), :357-503 (loop)
 *   fluxes               src/waveMod.f90:29-91 (elastic/anelastic fluxes),
 *                        :94-267 (A+|A|, T, T^-1), :269-405 (exact Riemann fluxes)
 *   Runge-Kutta          src/rungekuttaMod.f90:27-120; timeint 4 is the
 *                        low-storage RK4 of Carpenter and Kennedy (no
 *                        counterpart in the reference)
 *   energy               src/timeloopMod.f90:490-516
 *   PML switch-off       src/timeloopMod.f90:532-539 (at a given step; the
 *                        trigger rule itself is not restated here)
 */
#include "nexd_oracle.h"

#include <stdlib.h>
#include <string.h>

#define I3(a, b, c, nb, nc) (((size_t)(a) * (nb) + (b)) * (nc) + (c))

/* ---- flux matrices (src/waveMod.f90:94-267) --------------------------- */
static void get_apa(double vp, double vs, double rho, double lam, double mu, double a[9][9]) {
    memset(a, 0, sizeof(double) * 81);
    a[0][0] = vp;  a[0][6] = -(lam + 2 * mu);
    a[1][0] = (lam * vp) / (lam + 2 * mu); a[1][6] = -lam;
    a[2][0] = (lam * vp) / (lam + 2 * mu); a[2][6] = -lam;
    a[3][3] = vs;  a[3][7] = -mu;
    a[5][5] = vs;  a[5][8] = -mu;
    a[6][0] = -1 / rho; a[6][6] = vp;
    a[7][3] = -1 / rho; a[7][7] = vs;
    a[8][5] = -1 / rho; a[8][8] = vs;
}

static void get_anelastic_apa(double vp, double vs, double rho, double a[6][9]) {
    memset(a, 0, sizeof(double) * 54);
    a[0][0] = 1.0 / (vp * rho); a[0][6] = -1.0;
    a[3][3] = 1.0 / (2 * vs * rho); a[3][7] = -1.0 / 2.0;
    a[5][5] = 1.0 / (2 * vs * rho); a[5][8] = -1.0 / 2.0;
}

static void get_t(const double *n, const double *s, const double *t, double T[9][9]) {
    memset(T, 0, sizeof(double) * 81);
    for (int r = 0; r < 3; ++r) {
        T[r][0] = n[r] * n[r]; T[r][1] = s[r] * s[r]; T[r][2] = t[r] * t[r];
        T[r][3] = 2 * n[r] * s[r]; T[r][4] = 2 * s[r] * t[r]; T[r][5] = 2 * n[r] * t[r];
    }
    const int pr[3][3] = {{3, 1, 0}, {4, 2, 1}, {5, 2, 0}};
    for (int k = 0; k < 3; ++k) {
        int row = pr[k][0], a = pr[k][1], b = pr[k][2];
        T[row][0] = n[a] * n[b]; T[row][1] = s[a] * s[b]; T[row][2] = t[a] * t[b];
        T[row][3] = n[a] * s[b] + n[b] * s[a];
        T[row][4] = s[a] * t[b] + s[b] * t[a];
        T[row][5] = n[a] * t[b] + n[b] * t[a];
    }
    for (int r = 0; r < 3; ++r) { T[6 + r][6] = n[r]; T[6 + r][7] = s[r]; T[6 + r][8] = t[r]; }
}

static void get_inv_t(const double *n, const double *s, const double *t, double T[9][9]) {
    memset(T, 0, sizeof(double) * 81);
    const double *v3[3] = {n, s, t};
    for (int row = 0; row < 3; ++row) {
        const double *v = v3[row];
        T[row][0] = v[0] * v[0]; T[row][1] = v[1] * v[1]; T[row][2] = v[2] * v[2];
        T[row][3] = 2 * v[1] * v[0]; T[row][4] = 2 * v[2] * v[1]; T[row][5] = 2 * v[2] * v[0];
    }
    const double *u6[3] = {n, s, n}, *w6[3] = {s, t, t};
    for (int k = 0; k < 3; ++k) {
        const double *u = u6[k], *w = w6[k];
        int row = 3 + k;
        T[row][0] = u[0] * w[0]; T[row][1] = u[1] * w[1]; T[row][2] = u[2] * w[2];
        T[row][3] = u[1] * w[0] + u[0] * w[1];
        T[row][4] = u[2] * w[1] + u[1] * w[2];
        T[row][5] = u[2] * w[0] + u[0] * w[2];
    }
    for (int c = 0; c < 3; ++c) { T[6][6 + c] = n[c]; T[7][6 + c] = s[c]; T[8][6 + c] = t[c]; }
}

static void matmul(int m, int k, int n, const double *a, const double *b, double *c) {
    /* c[m][n] = a[m][k] * b[k][n], summing over k in order */
    for (int i = 0; i < m; ++i)
        for (int j = 0; j < n; ++j) {
            double acc = 0.0;
            for (int l = 0; l < k; ++l) acc += a[i * k + l] * b[l * n + j];
            c[i * n + j] = acc;
        }
}

typedef struct {
    double vt[4][9][9], vtf[4][9][9];
    double avt[4][6][9], avtf[4][6][9];
} elem_mats;

static void build_matrices(const oracle_problem *p, int e, elem_mats *m) {
    double apa[9][9], aapa[6][9], T[9][9], iT[9][9], tmp[9][9], tmp2[9][9], freem[9][9];
    double apaf[9][9], aapaf[6][9], atmp[6][9];
    memset(freem, 0, sizeof(freem));
    freem[0][0] = -2; freem[3][3] = -2; freem[5][5] = -2;
    get_apa(p->vpu[e], p->vsu[e], p->rho[e], p->lambdau[e], p->muu[e], apa);
    get_anelastic_apa(p->vpu[e], p->vsu[e], p->rho[e], aapa);
    matmul(9, 9, 9, &apa[0][0], &freem[0][0], &apaf[0][0]);
    matmul(6, 9, 9, &aapa[0][0], &freem[0][0], &aapaf[0][0]);
    for (int f = 0; f < 4; ++f) {
        const double *n = p->nnormal + I3(e, f, 0, 4, 3);
        const double *s = p->stangential + I3(e, f, 0, 4, 3);
        const double *t = p->ttangential + I3(e, f, 0, 4, 3);
        get_t(n, s, t, T);
        get_inv_t(n, s, t, iT);
        matmul(9, 9, 9, &apa[0][0], &iT[0][0], &tmp[0][0]);
        matmul(9, 9, 9, &T[0][0], &tmp[0][0], &m->vt[f][0][0]);
        matmul(9, 9, 9, &apaf[0][0], &iT[0][0], &tmp2[0][0]);
        matmul(9, 9, 9, &T[0][0], &tmp2[0][0], &m->vtf[f][0][0]);
        if (p->attenuation) {
            double aT[6][6];
            for (int i = 0; i < 6; ++i) for (int j = 0; j < 6; ++j) aT[i][j] = T[i][j];
            matmul(6, 9, 9, &aapa[0][0], &iT[0][0], &atmp[0][0]);
            matmul(6, 6, 9, &aT[0][0], &atmp[0][0], &m->avt[f][0][0]);
            matmul(6, 9, 9, &aapaf[0][0], &iT[0][0], &atmp[0][0]);
            matmul(6, 6, 9, &aT[0][0], &atmp[0][0], &m->avtf[f][0][0]);
        }
    }
}

/* ---- RK formulas (src/rungekuttaMod.f90:90-120) -------------------------- */
static inline double rk2(double a, double an, double dadt, double dt, int irk) {
    return irk == 1 ? a + dadt * dt : 0.5 * (an + a + dadt * dt);
}
static inline double rk3(double a, double an, double dadt, double dt, int irk, double w13, double w23) {
    if (irk == 1) return a + dadt * dt;
    if (irk == 2) return 0.75 * an + 0.25 * (a + dadt * dt);
    return w13 * an + w23 * (a + dadt * dt);
}

/* Low-storage RK4 (timeint = 4; no counterpart in the reference, whose
 * schemes are Euler, TVD RK2 and TVD RK3, src/timeloopMod.f90:230-248):
 * the five-stage fourth-order 2N-storage scheme of Carpenter and Kennedy
 * (NASA TM-109112, 1994), coefficients as published there.  Stage k:
 *     res = A_k*res + dt*dydt;   y = y + B_k*res       (A_1 = 0) */
static const double LSRK_A[5] = {0.0, -567301805773.0 / 1357537059087.0, -2404267990393.0 / 2016746695238.0,
                                 -3550918686646.0 / 2091501179385.0, -1275806237668.0 / 842570457699.0};
static const double LSRK_B[5] = {1432997174477.0 / 9575080441755.0, 5161836677717.0 / 13612068292357.0,
                                 1720146321549.0 / 2090206949498.0, 3134564353537.0 / 4481467310338.0,
                                 2277821191437.0 / 14882151754819.0};
static inline double lsrk(double y, double *res, double dydt, double dt, int irk) {
    const double r = (irk == 1 ? 0.0 : LSRK_A[irk - 1] * *res) + dt * dydt;
    *res = r;
    return y + LSRK_B[irk - 1] * r;
}

/* the stage arithmetic of every scheme on the scalar ODE y' = lam*y, for
 * checks of the schemes themselves (amplification polynomial, order) */
int oracle_rk_scalar(int timeint, int compat, double lam, double dt, int nsteps, double y0, double *y_out) {
    const int nrk = timeint == 1 ? 1 : (timeint == 2 ? 2 : (timeint == 3 ? 3 : 5));
    const double w13 = (compat & 1) ? (double)(1.0f / 3.0f) : 1.0 / 3.0;
    const double w23 = (compat & 1) ? (double)(2.0f / 3.0f) : 2.0 / 3.0;
    if (timeint < 1 || timeint > 4) return -1;
    double y = y0, res = 0.0;
    for (int it = 0; it < nsteps; ++it) {
        const double yn = y;
        for (int irk = 1; irk <= nrk; ++irk) {
            const double f = lam * y;
            if (timeint == 4) y = lsrk(y, &res, f, dt, irk);
            else if (timeint == 3) y = rk3(y, yn, f, dt, irk, w13, w23);
            else y = rk2(y, yn, f, dt, irk);
        }
    }
    *y_out = y;
    return 0;
}

/* elastic fluxes (src/waveMod.f90:29-63) of one element, (Np,9) each */
static void elastic_fluxes(const double *qm, size_t nglob, int base, int np, const double efv[4],
                           double *F, double *G, double *H) {
    for (int j = 0; j < np; ++j) {
        const double q1 = qm[0 * nglob + base + j], q2 = qm[1 * nglob + base + j], q3 = qm[2 * nglob + base + j];
        const double q4 = qm[3 * nglob + base + j], q5 = qm[4 * nglob + base + j], q6 = qm[5 * nglob + base + j];
        const double q7 = qm[6 * nglob + base + j], q8 = qm[7 * nglob + base + j], q9 = qm[8 * nglob + base + j];
        double *f = F + j * 9, *g = G + j * 9, *h = H + j * 9;
        f[0] = efv[0] * q7; f[1] = efv[1] * q7; f[2] = efv[1] * q7; f[3] = efv[2] * q8; f[4] = 0.0;
        f[5] = efv[2] * q9; f[6] = efv[3] * q1; f[7] = efv[3] * q4; f[8] = efv[3] * q6;
        g[0] = efv[1] * q8; g[1] = efv[0] * q8; g[2] = efv[1] * q8; g[3] = efv[2] * q7; g[4] = efv[2] * q9;
        g[5] = 0.0; g[6] = efv[3] * q4; g[7] = efv[3] * q2; g[8] = efv[3] * q5;
        h[0] = efv[1] * q9; h[1] = efv[1] * q9; h[2] = efv[0] * q9; h[3] = 0.0; h[4] = efv[2] * q8;
        h[5] = efv[2] * q7; h[6] = efv[3] * q6; h[7] = efv[3] * q5; h[8] = efv[3] * q3;
    }
}

static void pml_modify(const oracle_problem *p, const double *fpm, const double *gpm, const double *hpm,
                       size_t nglob, int base, double *F, double *G, double *H) {
    for (int i = 0; i < 9; ++i)
        for (int j = 0; j < p->np; ++j) {
            size_t g = base + j;
            F[j * 9 + i] = (fpm[i * nglob + g] + F[j * 9 + i]) / p->kx[g];
            G[j * 9 + i] = (gpm[i * nglob + g] + G[j * 9 + i]) / p->ky[g];
            H[j * 9 + i] = (hpm[i * nglob + g] + H[j * 9 + i]) / p->kz[g];
        }
}

int oracle_run(const oracle_problem *p, int nsteps, double *seismo, double *energy, double *q_out) {
    const int np = p->np, npf = p->npf, nmb = p->nmb, nelem = p->nelem;
    const size_t nglob = (size_t)nelem * np;
    const int nfp = 4 * npf;
    const double dt = p->dt;
    const int nrk = p->timeint == 1 ? 1 : (p->timeint == 2 ? 2 : (p->timeint == 3 ? 3 : 5));
    const int wrk = p->timeint == 3 ? 2 : (p->timeint == 4 ? 4 : 1);
    const double w13 = (p->compat & 1) ? (double)(1.0f / 3.0f) : 1.0 / 3.0;
    const double w23 = (p->compat & 1) ? (double)(2.0f / 3.0f) : 2.0 / 3.0;
    if (p->timeint < 1 || p->timeint > 4) return -1;

    double *q = malloc(sizeof(double) * 9 * nglob), *qn = malloc(sizeof(double) * 9 * nglob);
    double *qm = malloc(sizeof(double) * 9 * nglob), *rq = malloc(sizeof(double) * 9 * nglob);
    double *ux = malloc(sizeof(double) * 3 * nglob), *ax = malloc(sizeof(double) * 3 * nglob);
    size_t nth = (size_t)6 * nmb * nglob;
    double *th = NULL, *thn = NULL, *thm = NULL, *arq = NULL;
    double *fp = NULL, *fpn = NULL, *fpm = NULL, *epml = NULL;
    elem_mats *mats = malloc(sizeof(elem_mats) * (size_t)nelem);
    if (!q || !qn || !qm || !rq || !ux || !ax || !mats) return -2;
    /* the default-REAL literal 1e-24 widened to REAL(8) (src/timeloopMod.f90:184-195) */
    const double tiny = (double)1e-24f;
    for (size_t i = 0; i < 9 * nglob; ++i) { q[i] = tiny; rq[i] = tiny; qn[i] = tiny; }
    for (size_t i = 0; i < 3 * nglob; ++i) { ux[i] = tiny; ax[i] = tiny; }
    if (p->attenuation) {
        th = malloc(sizeof(double) * nth); thn = malloc(sizeof(double) * nth);
        thm = malloc(sizeof(double) * nth); arq = malloc(sizeof(double) * nth);
        for (size_t i = 0; i < nth; ++i) { th[i] = tiny; arq[i] = tiny; thn[i] = 0.0; }
    }
    if (p->pml) {
        fp = calloc(27 * nglob, sizeof(double)); fpn = calloc(27 * nglob, sizeof(double));
        fpm = calloc(27 * nglob, sizeof(double)); epml = malloc(sizeof(double) * 6 * nglob);
    }
    /* low-storage RK residuals (timeint 4) */
    double *resq = NULL, *resth = NULL, *resfp = NULL;
    if (wrk == 4) {
        resq = calloc(9 * nglob, sizeof(double));
        if (p->attenuation) resth = calloc(nth, sizeof(double));
        if (p->pml) resfp = calloc(27 * nglob, sizeof(double));
    }
    int *pmlcheck = malloc(sizeof(int) * (size_t)nelem);
    for (int e = 0; e < nelem; ++e) pmlcheck[e] = p->pml ? p->pmlcheck[e] : 0;

    #pragma omp parallel for schedule(static)
    for (int e = 0; e < nelem; ++e) build_matrices(p, e, &mats[e]);

    double last_f[9 * 64], last_g[9 * 64], last_h[9 * 64];

    for (int it = 0; it < nsteps; ++it) {
        /* receivers: state before step it (src/timeloopMod.f90:358-373) */
        for (int r = 0; r < p->nrec; ++r) {
            double acc[9] = {0, 0, 0, 0, 0, 0, 0, 0, 0};
            const int base = p->recelem[r] * np;
            for (int j = 0; j < np; ++j) {
                const double w = p->rec_int[(size_t)r * np + j];
                const size_t g = base + j;
                acc[0] = acc[0] + ux[g] * w; acc[1] = acc[1] + ux[nglob + g] * w; acc[2] = acc[2] + ux[2 * nglob + g] * w;
                acc[3] = acc[3] + q[6 * nglob + g] * w; acc[4] = acc[4] + q[7 * nglob + g] * w;
                acc[5] = acc[5] + q[8 * nglob + g] * w;
                acc[6] = acc[6] + ax[g] * w; acc[7] = acc[7] + ax[nglob + g] * w; acc[8] = acc[8] + ax[2 * nglob + g] * w;
            }
            for (int k = 0; k < 9; ++k) seismo[I3(r, k, it, 9, p->nt)] = acc[k];
        }
        for (int irk = 1; irk <= nrk; ++irk) {
            memcpy(qm, q, sizeof(double) * 9 * nglob);
            if (irk == 1) memcpy(qn, q, sizeof(double) * 9 * nglob);
            if (p->attenuation) {
                memcpy(thm, th, sizeof(double) * nth);
                if (irk == 1) memcpy(thn, th, sizeof(double) * nth);
            }
            if (p->pml) {
                memcpy(fpm, fp, sizeof(double) * 27 * nglob);
                if (irk == 1) memcpy(fpn, fp, sizeof(double) * 27 * nglob);
            }
            const double *fpm_ = fpm, *gpm_ = fpm ? fpm + 9 * nglob : NULL, *hpm_ = fpm ? fpm + 18 * nglob : NULL;

            #pragma omp parallel for schedule(static)
            for (int e = 0; e < nelem; ++e) {
                double F[64 * 9], G[64 * 9], H[64 * 9], flux[4 * 21 * 9], aflux[4 * 21 * 6];
                double af[64 * 18], ag[64 * 18], ah[64 * 18];
                const int base = e * np;
                const double efv[4] = {-(p->lambdau[e] + 2 * p->muu[e]), -p->lambdau[e], -p->muu[e], -1.0 / p->rho[e]};
                elastic_fluxes(qm, nglob, base, np, efv, F, G, H);
                if (pmlcheck[e]) pml_modify(p, fpm_, gpm_, hpm_, nglob, base, F, G, H);
                /* volume term (src/timeloopMod.f90:407-420) */
                for (int i = 0; i < 9; ++i)
                    for (int j = 0; j < np; ++j) {
                        double dFdr = 0, dFds = 0, dFdt = 0, dGdr = 0, dGds = 0, dGdt = 0, dHdr = 0, dHds = 0, dHdt = 0;
                        for (int m = 0; m < np; ++m) {
                            const double dr = p->dr[j * np + m], ds = p->ds[j * np + m], dtt = p->dt_[j * np + m];
                            dFdr += dr * F[m * 9 + i]; dFds += ds * F[m * 9 + i]; dFdt += dtt * F[m * 9 + i];
                            dGdr += dr * G[m * 9 + i]; dGds += ds * G[m * 9 + i]; dGdt += dtt * G[m * 9 + i];
                            dHdr += dr * H[m * 9 + i]; dHds += ds * H[m * 9 + i]; dHdt += dtt * H[m * 9 + i];
                        }
                        const size_t g = base + j;
                        rq[i * nglob + g] = (p->rx[g] * dFdr + p->sx[g] * dFds + p->tx[g] * dFdt) +
                                            (p->ry[g] * dGdr + p->sy[g] * dGds + p->ty[g] * dGdt) +
                                            (p->rz[g] * dHdr + p->sz[g] * dHds + p->tz[g] * dHdt);
                    }
                /* exact Riemann fluxes (src/waveMod.f90:269-336) */
                const elem_mats *mm = &mats[e];
                for (int is = 0; is < 4; ++is) {
                    const int nb = p->neighbor[e * 4 + is], fc = p->face[e * 4 + is];
                    for (int i = 0; i < npf; ++i) {
                        const size_t gt = p->ibt[I3(e, is, i, 4, npf)];
                        double dq[9];
                        const double (*M)[9];
                        double sgn = 1.0;
                        if (nb >= 0) {
                            const size_t gn = p->ibn[I3(e, is, i, 4, npf)];
                            for (int k = 0; k < 9; ++k) dq[k] = qm[k * nglob + gn] - qm[k * nglob + gt];
                            M = mm->vt[is];
                        } else if (fc == -1) {
                            for (int k = 0; k < 9; ++k) dq[k] = qm[k * nglob + gt];
                            M = mm->vtf[is];
                        } else {
                            for (int k = 0; k < 9; ++k) dq[k] = qm[k * nglob + gt];
                            M = mm->vt[is];
                            sgn = -1.0;
                        }
                        for (int k = 0; k < 9; ++k) {
                            double acc = 0.0;
                            for (int l = 0; l < 9; ++l) acc += M[k][l] * dq[l];
                            flux[(is * npf + i) * 9 + k] = sgn * acc;
                        }
                        if (p->attenuation) {
                            const double (*A)[9] = (nb >= 0 || fc == -2) ? mm->avt[is] : mm->avtf[is];
                            for (int k = 0; k < 6; ++k) {
                                double acc = 0.0;
                                for (int l = 0; l < 9; ++l) acc += A[k][l] * dq[l];
                                aflux[(is * npf + i) * 6 + k] = sgn * acc;
                            }
                        }
                    }
                }
                if (p->attenuation) {
                    /* anelastic fluxes and memory-variable RHS (src/timeloopMod.f90:426-461) */
                    memset(af, 0, sizeof(double) * np * 18);
                    memset(ag, 0, sizeof(double) * np * 18);
                    memset(ah, 0, sizeof(double) * np * 18);
                    for (int i = 0; i < nmb; ++i)
                        for (int j = 0; j < np; ++j) {
                            const size_t g = base + j;
                            const double v1 = qm[6 * nglob + g], v2 = qm[7 * nglob + g], v3 = qm[8 * nglob + g];
                            const double w = p->wl[i];
                            const double t1a = -0.5 * w * v1, t1b = -0.5 * w * v2, t1c = -0.5 * w * v3;
                            const double t2a = -w * v1, t2b = -w * v2, t2c = -w * v3;
                            af[j * 18 + i * 6 + 0] = t2a; af[j * 18 + i * 6 + 3] = t1b; af[j * 18 + i * 6 + 5] = t1c;
                            ag[j * 18 + i * 6 + 1] = t2b; ag[j * 18 + i * 6 + 3] = t1a; ag[j * 18 + i * 6 + 4] = t1c;
                            ah[j * 18 + i * 6 + 2] = t2c; ah[j * 18 + i * 6 + 4] = t1b; ah[j * 18 + i * 6 + 5] = t1a;
                        }
                    for (int c = 0; c < 6 * nmb; ++c)
                        for (int j = 0; j < np; ++j) {
                            double dFdr = 0, dFds = 0, dFdt = 0, dGdr = 0, dGds = 0, dGdt = 0, dHdr = 0, dHds = 0, dHdt = 0;
                            for (int m = 0; m < np; ++m) {
                                const double dr = p->dr[j * np + m], ds = p->ds[j * np + m], dtt = p->dt_[j * np + m];
                                dFdr += dr * af[m * 18 + c]; dFds += ds * af[m * 18 + c]; dFdt += dtt * af[m * 18 + c];
                                dGdr += dr * ag[m * 18 + c]; dGds += ds * ag[m * 18 + c]; dGdt += dtt * ag[m * 18 + c];
                                dHdr += dr * ah[m * 18 + c]; dHds += ds * ah[m * 18 + c]; dHdt += dtt * ah[m * 18 + c];
                            }
                            const size_t g = base + j;
                            arq[(size_t)c * nglob + g] = (p->rx[g] * dFdr + p->sx[g] * dFds + p->tx[g] * dFdt) +
                                                         (p->ry[g] * dGdr + p->sy[g] * dGds + p->ty[g] * dGdt) +
                                                         (p->rz[g] * dHdr + p->sz[g] * dHds + p->tz[g] * dHdt);
                        }
                    int c = 0;
                    for (int i = 0; i < nmb; ++i)
                        for (int jj = 0; jj < 6; ++jj, ++c)
                            for (int j = 0; j < np; ++j) {
                                double acc = 0.0;
                                for (int k = 0; k < nfp; ++k)
                                    acc += p->lift[j * nfp + k] * (p->fscale[(size_t)e * nfp + k] * p->wl[i] * aflux[k * 6 + jj] / 2.0);
                                const size_t g = base + j;
                                arq[(size_t)c * nglob + g] = arq[(size_t)c * nglob + g] + acc - p->wl[i] * thm[((size_t)i * 6 + jj) * nglob + g];
                            }
                    for (int i = 0; i < nmb; ++i) {
                        const double *av = p->anelasticvar + ((size_t)e * nmb + i) * 3;
                        for (int j = 0; j < np; ++j) {
                            const size_t g = base + j;
                            const double *t0 = thm + ((size_t)i * 6) * nglob + g;
                            const double th1 = t0[0], th2 = t0[nglob], th3 = t0[2 * nglob];
                            const double th4 = t0[3 * nglob], th5 = t0[4 * nglob], th6 = t0[5 * nglob];
                            rq[0 * nglob + g] = rq[0 * nglob + g] + av[2] * th1 + av[0] * (th2 + th3);
                            rq[1 * nglob + g] = rq[1 * nglob + g] + av[2] * th2 + av[0] * (th3 + th1);
                            rq[2 * nglob + g] = rq[2 * nglob + g] + av[2] * th3 + av[0] * (th1 + th2);
                            rq[3 * nglob + g] = rq[3 * nglob + g] + av[1] * th4;
                            rq[4 * nglob + g] = rq[4 * nglob + g] + av[1] * th5;
                            rq[5 * nglob + g] = rq[5 * nglob + g] + av[1] * th6;
                        }
                    }
                }
                /* LIFT (src/timeloopMod.f90:463-465) */
                for (int i = 0; i < 9; ++i)
                    for (int j = 0; j < np; ++j) {
                        double acc = 0.0;
                        for (int k = 0; k < nfp; ++k)
                            acc += p->lift[j * nfp + k] * (p->fscale[(size_t)e * nfp + k] * flux[k * 9 + i] / 2.0);
                        rq[i * nglob + base + j] = rq[i * nglob + base + j] + acc;
                    }
                /* sources (src/timeloopMod.f90:467-475) */
                for (int s = 0; s < p->nsrc; ++s) {
                    if (p->srcelem[s] != e) continue;
                    if (p->srctype[s] == 0) {
                        for (int c2 = 0; c2 < 3; ++c2)
                            for (int j = 0; j < np; ++j)
                                rq[(6 + c2) * nglob + base + j] += p->src_array[(((size_t)it * p->nsrc + s) * 3 + c2) * np + j];
                    } else {
                        for (int c2 = 0; c2 < 6; ++c2)
                            for (int j = 0; j < np; ++j)
                                rq[c2 * nglob + base + j] += p->src_array_m[(((size_t)it * p->nsrc + s) * 6 + c2) * np + j];
                    }
                }
            }
            /* the buffer the PML update reads: the last element's fluxes (H1),
             * or nothing when every element uses its own (not compat) */
            if (p->pml) {
                const int e = nelem - 1, base = e * np;
                const double efv[4] = {-(p->lambdau[e] + 2 * p->muu[e]), -p->lambdau[e], -p->muu[e], -1.0 / p->rho[e]};
                elastic_fluxes(qm, nglob, base, np, efv, last_f, last_g, last_h);
                if (pmlcheck[e]) pml_modify(p, fpm_, gpm_, hpm_, nglob, base, last_f, last_g, last_h);
            }
            /* Runge-Kutta (src/rungekuttaMod.f90:27-88) */
            #pragma omp parallel for schedule(static)
            for (size_t i = 0; i < 9 * nglob; ++i)
                q[i] = wrk == 4 ? lsrk(q[i], &resq[i], rq[i], dt, irk)
                                : (wrk == 1 ? rk2(q[i], qn[i], rq[i], dt, irk) : rk3(q[i], qn[i], rq[i], dt, irk, w13, w23));
            if (p->attenuation) {
                #pragma omp parallel for schedule(static)
                for (size_t i = 0; i < nth; ++i)
                    th[i] = wrk == 4 ? lsrk(th[i], &resth[i], arq[i], dt, irk)
                                     : (wrk == 1 ? rk2(th[i], thn[i], arq[i], dt, irk)
                                                 : rk3(th[i], thn[i], arq[i], dt, irk, w13, w23));
            }
            if (p->pml) {
                #pragma omp parallel for schedule(static)
                for (int e = 0; e < nelem; ++e) {
                    if (!pmlcheck[e]) continue;
                    double ownF[64 * 9], ownG[64 * 9], ownH[64 * 9];
                    const double *uf = last_f, *ug = last_g, *uh = last_h;
                    if (!(p->compat & 2)) {
                        const double efv[4] = {-(p->lambdau[e] + 2 * p->muu[e]), -p->lambdau[e], -p->muu[e], -1.0 / p->rho[e]};
                        elastic_fluxes(qm, nglob, e * np, np, efv, ownF, ownG, ownH);
                        pml_modify(p, fpm_, gpm_, hpm_, nglob, e * np, ownF, ownG, ownH);
                        uf = ownF; ug = ownG; uh = ownH;
                    }
                    for (int i = 0; i < 9; ++i)
                        for (int j = 0; j < np; ++j) {
                            const size_t g = (size_t)e * np + j;
                            double *f = fp + i * nglob + g, *gg = fp + (9 + i) * nglob + g, *h = fp + (18 + i) * nglob + g;
                            const double fn = fpn[i * nglob + g], gn = fpn[(9 + i) * nglob + g], hn = fpn[(18 + i) * nglob + g];
                            const double df = -p->alphax[g] * (*f) - p->ddx[g] / p->kx[g] * ((*f) + uf[j * 9 + i]);
                            const double dg = -p->alphay[g] * (*gg) - p->ddy[g] / p->ky[g] * ((*gg) + ug[j * 9 + i]);
                            const double dh = -p->alphaz[g] * (*h) - p->ddz[g] / p->kz[g] * ((*h) + uh[j * 9 + i]);
                            if (wrk == 4) {
                                *f = lsrk(*f, &resfp[i * nglob + g], df, dt, irk);
                                *gg = lsrk(*gg, &resfp[(9 + i) * nglob + g], dg, dt, irk);
                                *h = lsrk(*h, &resfp[(18 + i) * nglob + g], dh, dt, irk);
                            } else if (wrk == 1) {
                                *f = rk2(*f, fn, df, dt, irk); *gg = rk2(*gg, gn, dg, dt, irk); *h = rk2(*h, hn, dh, dt, irk);
                            } else {
                                *f = rk3(*f, fn, df, dt, irk, w13, w23); *gg = rk3(*gg, gn, dg, dt, irk, w13, w23);
                                *h = rk3(*h, hn, dh, dt, irk, w13, w23);
                            }
                        }
                }
            }
            if (irk == nrk) {
                /* post-step (src/timeloopMod.f90:482-503) */
                for (size_t g = 0; g < nglob; ++g) {
                    ux[g] = ux[g] + dt * q[6 * nglob + g];
                    ux[nglob + g] = ux[nglob + g] + dt * q[7 * nglob + g];
                    ux[2 * nglob + g] = ux[2 * nglob + g] + dt * q[8 * nglob + g];
                    ax[g] = rq[6 * nglob + g]; ax[nglob + g] = rq[7 * nglob + g]; ax[2 * nglob + g] = rq[8 * nglob + g];
                }
                if (p->pml) {
                    for (int e = 0; e < nelem; ++e) {
                        const double mu = p->mu[e], la = p->lambda[e];
                        const double temp1 = 1. / (4 * mu * mu + 6 * mu * la), temp2 = 2 * mu;
                        for (int j = 0; j < np; ++j) {
                            const size_t g = (size_t)e * np + j;
                            const double q1 = q[g], q2 = q[nglob + g], q3 = q[2 * nglob + g];
                            epml[g] = (2 * (la + mu) * q1 - la * (q2 + q3)) / temp1;
                            epml[nglob + g] = (2 * (la + mu) * q2 - la * (q1 + q3)) / temp1;
                            epml[2 * nglob + g] = (2 * (la + mu) * q3 - la * (q1 + q1)) / temp1;
                            epml[3 * nglob + g] = q[3 * nglob + g] / temp2;
                            epml[4 * nglob + g] = q[4 * nglob + g] / temp2;
                            epml[5 * nglob + g] = q[5 * nglob + g] / temp2;
                        }
                    }
                }
            }
        }
        if (p->pml && energy) {
            /* energy (src/timeloopMod.f90:508-516); the STA/LTA trigger that
             * follows it divides by uninitialised locals and is not restated */
            double en = 0.0;
            for (int e = 0; e < nelem; ++e) {
                const size_t b = (size_t)e * np;
                double terms[9];
                for (int c = 0; c < 9; ++c) {
                    const double *x = (c >= 6) ? q + (size_t)c * nglob + b : q + (size_t)c * nglob + b;
                    const double *y = (c >= 6) ? q + (size_t)c * nglob + b : epml + (size_t)c * nglob + b;
                    double dot = 0.0;
                    for (int i = 0; i < np; ++i) {
                        double mv = 0.0;
                        for (int k = 0; k < np; ++k) mv += p->mass[i * np + k] * y[k];
                        dot += x[i] * mv;
                    }
                    terms[c] = dot;
                }
                en = en + 0.5 * p->jac[b] *
                     (p->rho[e] * terms[6] + p->rho[e] * terms[7] + p->rho[e] * terms[8] + terms[0] + terms[1] + terms[2] +
                      2 * (terms[3] + terms[4] + terms[5]));
            }
            energy[it] = en;
        }
        /* the trigger's switch-off: pmlcheck(:)=.false. after the step
         * (src/timeloopMod.f90:532-539); the energy stays on */
        if (p->pml_off_after > 0 && it + 1 == p->pml_off_after)
            for (int e = 0; e < nelem; ++e) pmlcheck[e] = 0;
    }
    if (q_out) memcpy(q_out, q, sizeof(double) * 9 * nglob);
    free(q); free(qn); free(qm); free(rq); free(ux); free(ax); free(mats); free(pmlcheck);
    free(th); free(thn); free(thm); free(arq); free(fp); free(fpn); free(fpm); free(epml);
    free(resq); free(resth); free(resfp);
    return 0;
}
