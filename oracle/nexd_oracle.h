/*
 * TEST INFRASTRUCTURE ONLY.  The oracle is a plain-C, FP64 restatement of
 * the reference's time loop and is used solely by tests/, by the smoke
 * check and by bench.py's cpu_baseline leg, as the checker.  No product code
 * links, loads or calls it.
 *
 * Layouts (0-based indices, C row-major unless stated):
 *   node arrays          [nglob], nglob = nelem*Np, node = e*Np + j
 *   q, fprime, ...       [9][nglob]         (Fortran q(nglob,9))
 *   theta                [nMB][6][nglob]    (Fortran theta(nglob,6,nMB))
 *   dr, ds, dt, mass     [Np][Np]           D[i][j] = Dr(i+1,j+1)
 *   lift                 [Np][4*NpF]
 *   fscale               [nelem][4*NpF]     (Fortran fscale(4NpF,nelem))
 *   ibt, ibn             [nelem][4][NpF]    0-based node ids, ibn<0 on boundaries
 *   neighbor, face       [nelem][4]         neighbor<0 on boundaries;
 *                                           face: >=0 neighbour face, -1 free, -2 absorbing
 *   nnormal, st, tt      [nelem][4][3]
 *   anelasticvar         [nelem][nMB][3]
 *   src_array            [nt][nsrc][3][Np]  (Fortran srcArray(Np,3,nsrc,nt))
 *   src_array_m          [nt][nsrc][6][Np]
 *   rec_int              [nrec][Np]
 *   seismo (output)      [nrec][9][nt]      ux uy uz vx vy vz ax ay az
 */
#ifndef NEXD_ORACLE_H
#define NEXD_ORACLE_H

#ifdef __cplusplus
extern "C" {
#endif

typedef struct oracle_problem {
    int nelem, np, npf, nmb;
    int nt, timeint, attenuation, pml;   /* timeint 1..3 as the reference; 4: low-storage RK4 */
    int compat;          /* bit 0: RK3 float weights (H3); bit 1: PML uses last element's flux (H1) */
    double dt;
    const double *dr, *ds, *dt_, *lift, *mass;
    const double *rx, *ry, *rz, *sx, *sy, *sz, *tx, *ty, *tz, *jac;
    const double *fscale;
    const int *neighbor, *face, *ibt, *ibn;
    const double *nnormal, *stangential, *ttangential;
    const double *vpu, *vsu, *rho, *lambdau, *muu, *mu, *lambda;
    const double *wl, *anelasticvar;
    const int *pmlcheck;
    const double *ddx, *ddy, *ddz, *alphax, *alphay, *alphaz, *kx, *ky, *kz;
    int nsrc;
    const int *srcelem, *srctype;
    const double *src_array, *src_array_m;
    int nrec;
    const int *recelem;
    const double *rec_int;
    int pml_off_after;   /* 0: PML stays on; k > 0: the PML elements are switched off after
                            step k (1-based), as the reference's STA/LTA trigger does
                            (src/timeloopMod.f90:532-539) */
} oracle_problem;

/* Runs steps [0, nsteps) of the time loop from the reference's initial
 * state (every field 1e-24).  Writes seismograms for those steps into
 * seismo[nrec][9][nt] (only columns < nsteps are touched), the per-step
 * energy (PML runs) into energy[nt] if non-NULL, and the final q into
 * q_out[9][nglob] if non-NULL.  Returns 0 on success. */
int oracle_run(const oracle_problem *p, int nsteps, double *seismo, double *energy, double *q_out);

/* The stage arithmetic of timeint 1..4 (Euler, TVD RK2, TVD RK3, the
 * five-stage low-storage RK4 of Carpenter and Kennedy) applied to the scalar
 * ODE y' = lam*y for nsteps steps from y0; *y_out receives y.  Used to check
 * the schemes themselves. */
int oracle_rk_scalar(int timeint, int compat, double lam, double dt, int nsteps, double y0, double *y_out);

#ifdef __cplusplus
}
#endif
#endif
