// Stage kernels of the time loop, shared by the single-shot FP64 path and the
// batched FP32/FP64 path.
//
// One kernel launch advances one Runge-Kutta stage for a list of elements:
// it fuses what the reference does in its element loop and its separate RK
// pass (src/timeloopMod.f90:395-480, src/rungekuttaMod.f90:27-120):
//   volume term      derivatives of the fluxes through Dr/Ds/Dt, metric terms
//                    applied per node (src/timeloopMod.f90:407-420)
//   face term        upwind flux T*(A+|A|)*T^-1 applied factored to the jump
//                    (src/waveMod.f90:269-336), lifted through LIFT
//                    (src/timeloopMod.f90:463-465)
//   attenuation      memory-variable RHS and the stress coupling
//                    (src/timeloopMod.f90:426-461)
//   PML              flux modification and the ADE update of f',g',h'
//                    (src/timeloopMod.f90:399-405, src/rungekuttaMod.f90:52-63)
//   sources          srcArray/srcArrayM of the step (src/timeloopMod.f90:467-475)
//   RK update        into a separate output buffer (double buffering replaces
//                    the reference's qm/qn copies, src/timeloopMod.f90:376-391)
//   post-step        receiver-element displacement/acceleration and the energy
//                    partials on the last stage (src/timeloopMod.f90:482-516)
//
// Layout in HBM: fields are structure-of-arrays, element-major within a field
// (the reference's q(nglob,9) column-major), so the Np nodes of an element are
// contiguous in every field and a block of consecutive elements is one
// contiguous run per field.  Node phases map threads to (element-in-block,
// node); the volume and LIFT products run on the matrix cores (16x16x4 MFMA)
// over all four waves of the block.
#pragma once
#include <hip/hip_runtime.h>

namespace nexd {

template <int ORDER>
struct Dims {
    static constexpr int NP = (ORDER + 1) * (ORDER + 2) * (ORDER + 3) / 6;
    static constexpr int NPF = (ORDER + 1) * (ORDER + 2) / 2;
    static constexpr int NFP = 4 * NPF;
};

constexpr int NMB = 3;
constexpr unsigned kXcds = 8;   // accelerator complex dies of an MI355X

template <typename R>
struct StageArgs {
    int nglob;
    int nshot;                  // shots advanced together (1 for the single-shot path)
    int mat_stride;             // per-shot stride of matc (0: shared), in R
    int anel_stride;            // per-shot stride of anel (0: shared)
    int pmlc_stride_ng;         // per-shot stride of pmlc in units of nglob (0 or 9)
    int src_factored;           // 1: sources as spatial x STF; 0: srcArray/srcArrayM
    int nt;
    const int *elist;           // elements handled by this launch
    int nlist;
    int irk;                    // 1-based stage
    int scheme;                 // 1 Euler, 2 RK2, 3 RK3, 4 low-storage RK4
    R lsa, lsb;                 // low-storage RK4: A and B of this stage
    int last_stage;             // 1 on the stage that completes a step
    R dt;
    const R *dt_elem;           // [nelem] the step of each element (instances of different
                                // step packed in one mesh) or null: dt for all
    R w13, w23;                 // RK3 final-stage weights
    const R *qin;               // [9][nglob] stage input
    const R *qbase;             // [9][nglob] step-start state (RK base); low-storage RK4: the residual
    R *qout;                    // [9][nglob]
    const R *thin;              // [18][nglob]
    const R *thbase;
    R *thout;
    const R *pin;               // [27][nglob] PML f',g',h'
    const R *pbase;
    R *pout;
    const R *last_flux;         // [27][NP] fluxes of the last element (compat)
    int pml_compat;
    const R *D;                 // Dr/Ds/Dt as MFMA A-operand fragments [RT][KS][64]
    const R *DK;                // [Dr|Ds|Dt] as MFMA A-operand fragments [RTV][3KS][64] (PML)
    const R *LT;                // LIFT as MFMA A-operand fragments [MTL][NFP/4][64]
    const int *fmask;           // [4][NPF] 0-based
    const R *metric;            // [9][nglob] rx ry rz sx sy sz tx ty tz
    const R *fscale;            // [nelem][NFP]
    const R *frame;             // [nelem][4][9] n s t
    const R *matc;              // [nelem][8] vp vs rho lam mu
    const int *nbr;             // [nelem][4][NPF] neighbour node or -1
    const int *ftype;           // [nelem][4] >=0 interior, -1 free, -2 absorbing
    const R *anel;              // [nelem][NMB][3]
    R wl[NMB];
    const R *pmlc;              // [9][nglob] alpha xyz, d/k xyz, k xyz
    const int *src_head;        // [nelem] first source of the element or -1
    const int *src_next;        // [nsrc]
    const int *srctype;         // [nsrc]
    const R *src_array;         // [nt][nsrc][3][NP]
    const R *src_array_m;       // [nt][nsrc][6][NP]
    const R *src_spatial;       // [shot][nsrc][6][NP]  factored sources
    const R *src_stf;           // [shot][nsrc][nt]
    int nsrc;                   // sources per shot
    const int *d_it;            // current step (0-based), device-resident
    const int *rec_slot;        // [nelem] receiver-element slot or -1
    int nslot;
    R *u_slot;                  // [shot][nslot][3][NP]
    R *a_slot;                  // [shot][nslot][3][NP]
    R *u_full;                  // [3][nglob] or null
    double *blk_energy;         // [shot][blk_stride] energy partials (this list's part) or null
    int blk_stride;             // per-shot stride of blk_energy: element blocks of a whole step
    const R *mass;              // mass matrix as MFMA A fragments [node tile][k-step][64] (energy)
    const R *jac;               // [nglob]
    const R *emat;              // [nelem][2] mu, lambda for the energy
    int emat_stride;            // per-shot stride of emat (0: shared)
    const int *pml_on;          // [shot] 0 once a shot's PML is switched off; null: on
    // compensated state (float pairs hi + lo for the 9 wavefield components
    // and the receiver-element displacement); null when not compensated
    const R *qlo_in;            // [9][nglob] low parts of qin
    const R *qlo_base;          // [9][nglob] low parts of qbase
    R *qlo_out;                 // [9][nglob]
    R *u_slot_lo;               // [shot][nslot][3][NP]
    // low-storage RK4: the residual buffers, updated in place (the same
    // memory as qbase, thbase, pbase); null for the other schemes
    R *qres, *thres, *pres;
};

// ---------------------------------------------------------------------------
// upwind flux of one face node: out = T (A+|A|) T^-1 dq, the operator the
// reference assembles as VT (src/timeloopMod.f90:262-271); free surfaces
// apply diag(-2,0,0,-2,0,-2,0,0,0) between A+|A| and T^-1.
// ---------------------------------------------------------------------------
template <typename R>
__device__ __forceinline__ void upwind_flux(const R *fr, R vp, R vs, R rho, R lam, R mu, bool free,
                                            const R dq[9], R out[9], R aout[6], bool att) {
    const R n0 = fr[0], n1 = fr[1], n2 = fr[2];
    const R s0 = fr[3], s1 = fr[4], s2 = fr[5];
    const R t0 = fr[6], t1 = fr[7], t2 = fr[8];
    // rows 1, 4, 6, 7, 8, 9 of T^-1 dq (the others meet zero columns of A+|A|)
    R w1 = n0 * n0 * dq[0] + n1 * n1 * dq[1] + n2 * n2 * dq[2] + 2 * n1 * n0 * dq[3] + 2 * n2 * n1 * dq[4] +
           2 * n2 * n0 * dq[5];
    R w4 = n0 * s0 * dq[0] + n1 * s1 * dq[1] + n2 * s2 * dq[2] + (n1 * s0 + n0 * s1) * dq[3] +
           (n2 * s1 + n1 * s2) * dq[4] + (n2 * s0 + n0 * s2) * dq[5];
    R w6 = n0 * t0 * dq[0] + n1 * t1 * dq[1] + n2 * t2 * dq[2] + (n1 * t0 + n0 * t1) * dq[3] +
           (n2 * t1 + n1 * t2) * dq[4] + (n2 * t0 + n0 * t2) * dq[5];
    R w7 = n0 * dq[6] + n1 * dq[7] + n2 * dq[8];
    R w8 = s0 * dq[6] + s1 * dq[7] + s2 * dq[8];
    R w9 = t0 * dq[6] + t1 * dq[7] + t2 * dq[8];
    if (free) {
        w1 = -2 * w1; w4 = -2 * w4; w6 = -2 * w6;
        w7 = 0; w8 = 0; w9 = 0;
    }
    const R l2m = lam + 2 * mu;
    const R z1 = vp * w1 - l2m * w7;
    const R z2 = (lam * vp) / l2m * w1 - lam * w7;   // z3 == z2
    const R z4 = vs * w4 - mu * w8;
    const R z6 = vs * w6 - mu * w9;
    const R irho = R(1) / rho;
    const R z7 = -irho * w1 + vp * w7;
    const R z8 = -irho * w4 + vs * w8;
    const R z9 = -irho * w6 + vs * w9;
    // T z (z5 == 0)
    out[0] = n0 * n0 * z1 + (s0 * s0 + t0 * t0) * z2 + 2 * n0 * s0 * z4 + 2 * n0 * t0 * z6;
    out[1] = n1 * n1 * z1 + (s1 * s1 + t1 * t1) * z2 + 2 * n1 * s1 * z4 + 2 * n1 * t1 * z6;
    out[2] = n2 * n2 * z1 + (s2 * s2 + t2 * t2) * z2 + 2 * n2 * s2 * z4 + 2 * n2 * t2 * z6;
    out[3] = n1 * n0 * z1 + (s1 * s0 + t1 * t0) * z2 + (n1 * s0 + n0 * s1) * z4 + (n1 * t0 + n0 * t1) * z6;
    out[4] = n2 * n1 * z1 + (s2 * s1 + t2 * t1) * z2 + (n2 * s1 + n1 * s2) * z4 + (n2 * t1 + n1 * t2) * z6;
    out[5] = n2 * n0 * z1 + (s2 * s0 + t2 * t0) * z2 + (n2 * s0 + n0 * s2) * z4 + (n2 * t0 + n0 * t2) * z6;
    out[6] = n0 * z7 + s0 * z8 + t0 * z9;
    out[7] = n1 * z7 + s1 * z8 + t1 * z9;
    out[8] = n2 * z7 + s2 * z8 + t2 * z9;
    if (att) {
        // anelastic part: T(1:6,1:6) * A_anel * T^-1 dq
        // (src/waveMod.f90:146-159, src/timeloopMod.f90:270-271)
        const R a1 = w1 / (vp * rho) - w7;
        const R a4 = w4 / (2 * vs * rho) - w8 / 2;
        const R a6 = w6 / (2 * vs * rho) - w9 / 2;
        aout[0] = n0 * n0 * a1 + 2 * n0 * s0 * a4 + 2 * n0 * t0 * a6;
        aout[1] = n1 * n1 * a1 + 2 * n1 * s1 * a4 + 2 * n1 * t1 * a6;
        aout[2] = n2 * n2 * a1 + 2 * n2 * s2 * a4 + 2 * n2 * t2 * a6;
        aout[3] = n1 * n0 * a1 + (n1 * s0 + n0 * s1) * a4 + (n1 * t0 + n0 * t1) * a6;
        aout[4] = n2 * n1 * a1 + (n2 * s1 + n1 * s2) * a4 + (n2 * t1 + n1 * t2) * a6;
        aout[5] = n2 * n0 * a1 + (n2 * s0 + n0 * s2) * a4 + (n2 * t0 + n0 * t2) * a6;
    }
}

// One Runge-Kutta stage update (src/rungekuttaMod.f90:90-120):
//   stage 1 (and Euler):   a + dt*rhs
//   TVD RK2, stage 2:      ((base + a) + dt*rhs) / 2
//   TVD RK3, stage 2:      3/4 base + 1/4 (a + dt*rhs)
//   TVD RK3, stage 3:      w13 base + w23 (a + dt*rhs)
// and the five-stage low-storage RK4 of Carpenter and Kennedy (NASA
// TM-109112, 1994; no counterpart in the reference), whose base is the
// residual of the previous stage:
//   res_new = A_k res + dt*rhs  (A_1 = 0),   a_new = a + B_k res_new
// The stage is uniform over a launch, so its weights are chosen once
// (rk_stage) and every combined value is the same straight-line code: both
// forms are computed and one is selected.  A branch per value would split a
// node's independent updates into separate basic blocks, which the compiler
// then cannot interleave, and each update's latency would add up.  dt*rhs
// and a + dt*rhs are rounded separately; the weighted sum is one FMA (exact
// for stage 1: 0 * base + x).  WEIGHTED takes the weighted form for TVD
// RK2 stage 2 as well (1/2 base + 1/2 (a + dt*rhs), equal up to rounding),
// which saves four operations per value; the FP32 memory variables and
// fields use it.
// fused multiply-add in the working precision (__builtin_fma is the double
// one: on floats it would widen, fuse in FP64 and round back)
__device__ __forceinline__ float fma_r(float a, float b, float c) { return __builtin_fmaf(a, b, c); }
__device__ __forceinline__ double fma_r(double a, double b, double c) { return __builtin_fma(a, b, c); }

// ---------------------------------------------------------------------------
// Field planes through buffer resources.  A field is a plane of ng values;
// the plane's base and byte size make one buffer descriptor (built in
// scalar registers at each access) and the node index is a 32-bit byte
// offset shared by every plane, so a batch of loads or stores needs no
// 64-bit address arithmetic in vector registers.  The hardware checks the
// bound: a read past the plane returns 0 and a write is dropped.  Planes
// are at most 4 GiB (the host checks ng * sizeof(R)).
// ---------------------------------------------------------------------------
template <typename R>
__device__ __forceinline__ __amdgpu_buffer_rsrc_t plane_rsrc(const R *plane, unsigned nbytes) {
    return __builtin_amdgcn_make_buffer_rsrc(const_cast<R *>(plane), (short)0, (int)nbytes, 0x00020000);
}
// BUF selects the buffer form; without it the same accesses are plain
// global loads and stores (each stage kernel instantiation picks the form
// measured faster for it).
template <bool BUF = true, typename R>
__device__ __forceinline__ R pload(const R *plane, unsigned nbytes, unsigned idx) {
    if constexpr (!BUF) {
        (void)nbytes;
        return plane[idx];
    } else {
        const __amdgpu_buffer_rsrc_t r = plane_rsrc(plane, nbytes);
        if constexpr (sizeof(R) == 4) return __builtin_bit_cast(R, __builtin_amdgcn_raw_buffer_load_b32(r, (int)(idx * 4u), 0, 0));
        else return __builtin_bit_cast(R, __builtin_amdgcn_raw_buffer_load_b64(r, (int)(idx * 8u), 0, 0));
    }
}
template <bool BUF = true, typename R>
__device__ __forceinline__ void pstore(R *plane, unsigned nbytes, unsigned idx, R v) {
    if constexpr (!BUF) {
        (void)nbytes;
        plane[idx] = v;
    } else {
        const __amdgpu_buffer_rsrc_t r = plane_rsrc(plane, nbytes);
        if constexpr (sizeof(R) == 4) __builtin_amdgcn_raw_buffer_store_b32(__builtin_bit_cast(unsigned, v), r, (int)(idx * 4u), 0, 0);
        else {
            typedef unsigned u32x2 __attribute__((ext_vector_type(2)));
            __builtin_amdgcn_raw_buffer_store_b64(__builtin_bit_cast(u32x2, v), r, (int)(idx * 8u), 0, 0);
        }
    }
}

// Carpenter and Kennedy's five-stage fourth-order 2N-storage coefficients
// (NASA TM-109112, 1994), as published
constexpr double kLsrkA[5] = {0.0, -567301805773.0 / 1357537059087.0, -2404267990393.0 / 2016746695238.0,
                              -3550918686646.0 / 2091501179385.0, -1275806237668.0 / 842570457699.0};
constexpr double kLsrkB[5] = {1432997174477.0 / 9575080441755.0, 5161836677717.0 / 13612068292357.0,
                              1720146321549.0 / 2090206949498.0, 3134564353537.0 / 4481467310338.0,
                              2277821191437.0 / 14882151754819.0};

template <typename R>
struct RkStage {
    R wb, wx;    // wb * base + wx * (a + dt*rhs)
    bool avg;    // TVD RK2 stage 2
    bool ls;     // low-storage RK4
    R la, lb;    // its A and B of the stage
};
template <typename R>
__device__ __forceinline__ RkStage<R> rk_stage(int scheme, int irk, R w13, R w23, R lsa, R lsb) {
    if (scheme == 4) return {R(0), R(1), false, true, irk == 1 ? R(0) : lsa, lsb};
    if (irk == 1) return {R(0), R(1), false, false, R(0), R(0)};
    if (scheme == 2) return {R(0.5), R(0.5), true, false, R(0), R(0)};
    if (irk == 2) return {R(0.75), R(0.25), false, false, R(0), R(0)};
    return {w13, w23, false, false, R(0), R(0)};
}
template <typename R, bool WEIGHTED = false>
__device__ __forceinline__ R rk_combine(const RkStage<R> &s, R a, R base, R rhs, R dt) {
#pragma clang fp contract(off)
    const R k = rhs * dt;
    const R x = a + k;
    const R z = fma_r(s.wb, base, s.wx * x);
    if constexpr (WEIGHTED) return z;
    const R y = ((base + a) + k) * R(0.5);
    return s.avg ? y : z;
}
// One stage of either kind: the low-storage form when LS (`res` is the
// previous residual on entry and the new one on return; A_1 = 0, so the
// stage-1 residual need only be finite), rk_combine otherwise (`res` is the
// RK base and stays as it is).
template <typename R, bool LS, bool WEIGHTED = false>
__device__ __forceinline__ R rk_update(const RkStage<R> &s, R a, R &res, R rhs, R dt) {
#pragma clang fp contract(off)
    if constexpr (LS) {
        const R r = s.la * res + rhs * dt;
        res = r;
        return a + s.lb * r;
    } else {
        return rk_combine<R, WEIGHTED>(s, a, res, rhs, dt);
    }
}

// ---------------------------------------------------------------------------
// Compensated RK combination: the state is carried as an unevaluated pair
// hi + lo, and the stage sum of rk_combine is formed with error-free
// transformations (TwoSum: Knuth, TAOCP vol. 2, 4.2.2; FastTwoSum: Dekker
// 1971; TwoProduct through one FMA), so each step's rounding of the stored
// state goes into lo instead of being lost.  The right-hand side is computed
// from hi only; its own rounding is relative to dt*rhs, not to the state.
// ---------------------------------------------------------------------------
template <typename R>
__device__ __forceinline__ void two_sum(R a, R b, R &s, R &e) {
#pragma clang fp contract(off)
    s = a + b;
    const R bb = s - a;
    e = (a - (s - bb)) + (b - bb);
}
template <typename R>
__device__ __forceinline__ void fast_two_sum(R a, R b, R &s, R &e) {   // |a| >= |b| or a == 0
#pragma clang fp contract(off)
    s = a + b;
    e = b - (s - a);
}
template <typename R>
__device__ __forceinline__ void two_prod(R a, R b, R &p, R &e) {
#pragma clang fp contract(off)
    p = a * b;
    e = fma_r(a, b, -p);
}
template <typename R>
__device__ __forceinline__ void rk_combine_comp(int scheme, int irk, R a, R alo, R base, R blo, R rhs, R dt,
                                                R w13, R w23, R &hi, R &lo) {
#pragma clang fp contract(off)
    const R k = rhs * dt;
    R s, e, h, l;
    if (irk == 1) {                                  // a + k
        two_sum(a, k, s, e);
        h = s; l = e + alo;
    } else if (scheme == 2) {                        // (base + a + k) / 2
        R s1, e1;
        two_sum(base, a, s1, e1);
        two_sum(s1, k, s, e);
        h = R(0.5) * s; l = R(0.5) * ((e1 + e) + (blo + alo));
    } else {                                         // c1*base + c2*(a + k)
        const R c1 = irk == 2 ? R(0.75) : w13, c2 = irk == 2 ? R(0.25) : w23;
        R p1, ep1, p2, ep2;
        two_sum(a, k, s, e);
        two_prod(c1, base, p1, ep1);
        two_prod(c2, s, p2, ep2);
        two_sum(p1, p2, h, l);
        l = l + (ep1 + ep2) + (c1 * blo + c2 * (e + alo));
    }
    if (h == R(0)) { hi = l; lo = 0; return; }
    fast_two_sum(h, l, hi, lo);
}

// elastic fluxes F,G,H of one node (src/waveMod.f90:29-63), 9 columns each
template <typename R>
__device__ __forceinline__ void elastic_flux_node(const R q[9], R ef0, R ef1, R ef2, R ef3, R F[9], R G[9], R H[9]) {
    F[0] = ef0 * q[6]; F[1] = ef1 * q[6]; F[2] = ef1 * q[6]; F[3] = ef2 * q[7]; F[4] = 0;
    F[5] = ef2 * q[8]; F[6] = ef3 * q[0]; F[7] = ef3 * q[3]; F[8] = ef3 * q[5];
    G[0] = ef1 * q[7]; G[1] = ef0 * q[7]; G[2] = ef1 * q[7]; G[3] = ef2 * q[6]; G[4] = ef2 * q[8];
    G[5] = 0; G[6] = ef3 * q[3]; G[7] = ef3 * q[1]; G[8] = ef3 * q[4];
    H[0] = ef1 * q[8]; H[1] = ef1 * q[8]; H[2] = ef0 * q[8]; H[3] = 0; H[4] = ef2 * q[7];
    H[5] = ef2 * q[6]; H[6] = ef3 * q[5]; H[7] = ef3 * q[4]; H[8] = ef3 * q[2];
}

// flux columns kept for PML elements: F without column 5, G without 6,
// H without 4 (those are identically zero, src/waveMod.f90:37-62)
__device__ constexpr int kPmlCols = 24;

// Block shape of one stage kernel instantiation.  Node threads (element-in-
// block, node) do the loads, the face term, LIFT and the update; the volume
// and LIFT products run on the matrix cores over all four waves.
//
// Volume term.  A = [Dr;Ds;Dt] (row tiles of (node, operator) rows, no
// padding row) times the block's columns, held as fragments in registers: the 9 fields outside
// the PML; in FP64 PML elements the 24 modified fluxes (F_i for d/dx, G_i
// for d/dy, H_i for d/dz) and, with attenuation, the 3 velocities.  The
// metric of each output node is applied to the three derivatives in the
// epilogue, as the reference does.
//
// FP32 PML elements use the contravariant form: on a straight-sided
// tetrahedron the metric is constant, so the sum of the flux derivatives is
//   rhs_i = Dr C_r,i + Ds C_s,i + Dt C_t,i,  C_o,i = o_x F_i + o_y G_i + o_z H_i,
// one product of A = [Dr|Ds|Dt] (Np x 3Np) with 9 columns per element
// instead of three operators on 24 columns; with attenuation 6 more columns
// give the strain rates the memory variables use, dv_l/dx_k + dv_k/dx_l =
// sum_o D_o (o_k v_l + o_l v_k) (one term on the diagonal).  Each wave
// holds one column tile's values at every k-step in registers and streams
// the operator.  The metric the reference
// computes varies between the nodes of an element at the level of single-
// precision rounding (its geometry is single precision), which this form
// averages over the element's nodes: a relative change near 1e-7, below the
// FP32 path's own error but above the FP64 path's parity bound.
// VAR selects one of the block shapes the autotuner times (FP32 only):
// 0 sizes a block for four blocks per CU, 1 for two, which doubles its LDS
// share (more elements per block) and its register budget.
template <int ORDER, bool ATT, bool PML, typename R, int VAR = 0>
struct Shape {
    static constexpr int NP = Dims<ORDER>::NP, NPF = Dims<ORDER>::NPF, NFP = Dims<ORDER>::NFP;
    static constexpr bool KC = PML && sizeof(R) == 4;          // contravariant volume term
    // volume columns per element
    static constexpr int NCOL = KC ? (ATT ? 15 : 9) : (PML ? (ATT ? 27 : 24) : 9);
    // volume values per node in LDS (row stride; odd for the contravariant
    // form so the node threads' row reads fall in different banks)
    static constexpr int NV = KC ? (NCOL | 1) : (PML ? (ATT ? 33 : 24) : 27);
    static constexpr int NFC = ATT ? 15 : 9;                    // face values per face node
    // LDS row lengths, odd so that threads reading or writing one value of
    // consecutive rows (one row per lane) fall in different banks
    static constexpr int SFN = NFP | 1;                         // face values: row of face nodes
    static constexpr int sln(int eb) { return (eb * NP) | 1; }  // lifted values: row of block nodes
    // rows of the first LDS region: q, and in the PML the 24 modified fluxes
    // or, contravariant form, per operator o the 9 fluxes C_o,i and the
    // metric row o_x o_y o_z (one stride for both kinds of column)
    static constexpr int ROWS_A = KC ? 9 + 36 : (PML ? 33 : 9);
    // four waves per block: the node phases use EB*NP threads of them, the
    // volume term all four
    static constexpr int NTHR = 256;
    // blocks per compute unit the shape is sized for (160 KiB of LDS per CU):
    // two FP64 blocks or four FP32 blocks, i.e. 2 or 4 waves per SIMD
    static constexpr int BPC = sizeof(R) == 8 ? 2 : (VAR ? 2 : 4);
    static constexpr int LDS_BYTES = 160 * 1024 / BPC;
    // waves per SIMD at that occupancy (4 SIMDs per CU): the second
    // __launch_bounds__ argument, which sets the register budget
    static constexpr int WPS = BPC * (NTHR / 64) / 4;
    // LDS regions, in reals:
    //   a (from 0): q (and the PML fluxes); the volume values [NT][NV]
    //     reuse it once the columns are in registers or, in the
    //     contravariant form, after the volume product
    //   face values [EB][NFC][SFN] after a, until they are in registers;
    //     the lifted values [16][sln] alias them (contravariant form) or
    //     follow the largest of a and the volume values
    //   metric [9][NT] after all of them, from the first phase to the end of
    //     the volume product (in the first region in the contravariant form)
    //   the energy values and the block reduction alias everything at the end
    static constexpr int off_face(int eb) { return ROWS_A * eb * NP; }
    static constexpr int off_lift(int eb) {
        const int nt = eb * NP;
        if (KC) return off_face(eb);
        return ROWS_A * nt > nt * NV ? ROWS_A * nt : nt * NV;
    }
    // the face region also holds the last element's 27 fluxes in phase 1
    static constexpr int face_reals(int eb) {
        return eb * NFC * SFN > (PML ? 27 * NP : 0) ? eb * NFC * SFN : 27 * NP;
    }
    static constexpr int off_metric(int eb) {
        const int f = off_face(eb) + face_reals(eb);
        const int l = off_lift(eb) + 16 * sln(eb);
        return f > l ? f : l;
    }
    static constexpr int lds_reals(int eb) {
        const int m = off_metric(eb) + (KC ? 0 : 9 * eb * NP);
        const int e = 24 * eb * NP + 4 * eb;    // energy values and per-element coefficients
        return m > e ? m : e;
    }
    // elements per block: at most one node per thread and 16 tiles of
    // columns, within the LDS share (the element ids take EB ints besides)
    static constexpr int pick_eb() {
        int eb = NTHR / NP < 256 / NCOL ? NTHR / NP : 256 / NCOL;
        if (eb < 1) eb = 1;
        while (eb > 1 && lds_reals(eb) * (int)sizeof(R) + 4 * eb > LDS_BYTES) --eb;
        return eb;
    }
    static constexpr int EB = pick_eb();
    static constexpr int NT = EB * NP;
    // FP32 PML shapes whose node threads fit in the first three waves run
    // the face term on the fourth wave while the node threads update the
    // PML fields (the fourth wave has no node of its own); the last
    // element's staged fluxes then live in operator 0's flux rows, which
    // are written only after that update
    static constexpr bool FW = KC && NT <= 3 * 64 && NTHR == 4 * 64 && EB * NFP <= 4 * 64 && EB * 5 <= 64 &&
                               27 * NP + EB * (4 * 9 + 5) <= 9 * NT;
    static constexpr int NCT = EB * NCOL;
    static constexpr int LDS = lds_reals(EB);
    static constexpr int OFF_F = off_face(EB);
    static constexpr int OFF_M = off_metric(EB);
    static constexpr int OFF_L = off_lift(EB);
    static_assert(!KC || NT * NV <= ROWS_A * NT, "volume values must fit the first region");
    static constexpr int NW = NTHR / 64;             // waves
    static constexpr int KS = (NP + 3) / 4;          // k-steps over one operator's columns
    static constexpr int CTL = (NCT + 15) / 16;      // column tiles
    // [Dr;Ds;Dt]: column tiles per wave; row tiles packed so that lane
    // group g (of 4) owns nodes g, g+4, ... with three operator rows each
    static constexpr int CPW = (CTL + NW - 1) / NW;
    static constexpr int RT = (3 * ((NP + 3) / 4) + 3) / 4;
    // contravariant form: node tiles of 16, k-steps over [Dr|Ds|Dt]; each
    // wave takes column tiles w, w+NW, ... against all node tiles
    static constexpr int RTV = (NP + 15) / 16;
    static constexpr int KSV = 3 * KS;
    // LIFT on the matrix cores: LIFT (node tiles of 16) times each element's
    // face values (16 columns: NFC values and padding), k-steps of 4 face nodes
    static constexpr int MTL = (NP + 15) / 16;       // node tiles
    static constexpr int KSL = NFP / 4;              // k-steps
    static constexpr int EPW = (EB + NW - 1) / NW;   // elements per wave
};

// one 16x16x4 MFMA, FP32 or FP64 (the accumulator layouts differ; the
// operator fragments are ordered to match, see mfma_operator_fragments)
template <typename R>
struct Mfma;
template <>
struct Mfma<float> {
    typedef float v4 __attribute__((ext_vector_type(4)));
    static __device__ __forceinline__ v4 fma(float a, float b, v4 c) {
        return __builtin_amdgcn_mfma_f32_16x16x4f32(a, b, c, 0, 0, 0);
    }
};
template <>
struct Mfma<double> {
    typedef double v4 __attribute__((ext_vector_type(4)));
    static __device__ __forceinline__ v4 fma(double a, double b, v4 c) {
        return __builtin_amdgcn_mfma_f64_16x16x4f64(a, b, c, 0, 0, 0);
    }
};

// LS: the low-storage RK4 instantiation (residuals read and written in
// place of the RK base); the TVD schemes use LS = false.  VAR: the block
// shape (Shape).
template <int ORDER, bool ATT, bool PML, typename R, bool LS, int VAR = 0>
__global__ void __launch_bounds__((Shape<ORDER, ATT, PML, R, VAR>::NTHR), (Shape<ORDER, ATT, PML, R, VAR>::WPS))
stage_kernel(StageArgs<R> a) {
    using S = Shape<ORDER, ATT, PML, R, VAR>;
    constexpr int NP = S::NP, NPF = S::NPF, NFP = S::NFP, EB = S::EB, NT = S::NT;
    constexpr int NCOL = S::NCOL, NV = S::NV, NFC = S::NFC, NTHR = S::NTHR;
    constexpr int SFN = S::SFN, SLN = S::sln(S::EB);
    __shared__ __align__(16) R lds[S::LDS];
    __shared__ int sElem[EB];             // the block's element ids, for the face phase
    constexpr bool KC = S::KC;
    // field planes through buffer descriptors, where measured faster: with
    // PML or attenuation, in the two-blocks-per-CU shape and at N = 5; the
    // plain elastic kernels of the four-blocks-per-CU shape up to N = 4 ran
    // 10-25 % slower with them (their few loads were scheduled in series)
    constexpr bool BUF = PML || ATT || VAR == 1 || ORDER >= 5;
    R *sQ = lds;                          // [9][NT]
    R *sX = lds + 9 * NT;                 // [24][NT]    PML flux columns (FP64)
    R *sC = lds + 9 * NT;                 // [3][12][NT] C_o,i and o_k, contravariant form (FP32 PML)
    R *sF = lds + S::OFF_F;               // [EB][NFC][SFN]  face values, face-node rows
    R *sV = lds;                          // [NT][NV]    volume values, after the volume product
    R *sM = lds + S::OFF_M;               // [9][NT]     metric of the block's nodes
    R *sL = lds + S::OFF_L;               // [16][SLN]   lifted face values, node rows
    static_assert(NTHR % 64 == 0, "the energy block sum assumes whole waves");

    const int t = threadIdx.x;
    const size_t ng = (size_t)a.nglob;
    const int it_step = *a.d_it;
    // The hardware deals consecutive workgroups round-robin to the XCDs, each
    // with its own L2.  The logical block index gives each XCD one
    // contiguous range instead (a bijection for any grid), so neighbouring
    // element groups, whose traces the face phase gathers, share an L2.
    // Shots vary fastest within a range, so the blocks of one element group
    // run together and share its geometry.
    const unsigned nb = gridDim.x, xcd = blockIdx.x % kXcds, rank = blockIdx.x / kXcds;
    const unsigned per = nb / kXcds, extra = nb % kXcds;
    const unsigned lb = xcd * per + (xcd < extra ? xcd : extra) + rank;
    const int shot = (int)(lb % (unsigned)a.nshot);
    const int eblk = (int)(lb / (unsigned)a.nshot);
    const R *qin = a.qin + (size_t)shot * 9 * ng;
    const R *qbase = a.qbase + (size_t)shot * 9 * ng;
    R *qout = a.qout + (size_t)shot * 9 * ng;
    const R *matc = a.matc + (size_t)shot * a.mat_stride;

    // ---- node-thread identity ------------------------------------------------
    const bool node_thr = t < NT;
    const int eb = node_thr ? t / NP : 0, j = node_thr ? t - (t / NP) * NP : 0;
    const int idx = eblk * EB + eb;
    const bool valid = node_thr && idx < a.nlist;
    // The last element's fluxes (PML, compat), the same for every element,
    // are staged once per block in the face-value region of LDS, which is
    // free until the face phase (in the FW shapes, where the face wave
    // writes that region during the PML update, in operator 0's flux rows);
    // they are loaded together with the element id, so their round trip
    // overlaps that one instead of following the phase-1 batch.
    R *sLast = S::FW ? sC : lds + S::OFF_F;   // [27][NP] during phase 1
    static_assert(!PML || 27 * NP <= S::face_reals(EB), "last-element fluxes must fit the face region");
    constexpr int NLS = PML ? (27 * NP + NTHR - 1) / NTHR : 1;
    R lst[NLS];
    if constexpr (PML) {
        if (a.pml_compat) {
            const R *lastf = a.last_flux + (size_t)shot * 27 * NP;
#pragma unroll
            for (int r = 0; r < NLS; ++r) {
                const int i = t + r * NTHR;
                lst[r] = lastf[i < 27 * NP ? i : 0];
            }
        }
    }
    const int e = valid ? a.elist[idx] : 0;
    if constexpr (PML) {
        if (a.pml_compat) {
#pragma unroll
            for (int r = 0; r < NLS; ++r) {
                const int i = t + r * NTHR;
                if (i < 27 * NP) sLast[i] = lst[r];
            }
        }
    }
    const size_t g = (size_t)e * NP + j;
    const unsigned g32 = (unsigned)e * NP + j;             // node index (planes are < 4 GiB)
    const unsigned pnb = (unsigned)(ng * sizeof(R));       // plane size in bytes
    const unsigned nel = (unsigned)(ng / NP);              // elements of the mesh (per-element arrays)
    // the RK base planes: stage 1 does not use the base (its weight is 0, or
    // A_1 = 0 for the low-storage residual), so their descriptors get size 0
    // there and the loads return 0 without touching memory
    const unsigned pnb_base = a.irk == 1 ? 0u : pnb;
    if (node_thr && j == 0) sElem[eb] = e;
    // the element's step (one value per element; the same for a whole mesh
    // unless instances of different steps share the launch)
    const R dt = a.dt_elem ? a.dt_elem[e] : a.dt;
    // first source of the element and its receiver slot, loaded with the
    // first batch instead of at the assembly (read only where valid)
    int src_first = -1, rec_slot_e = -1;

    // The face wave (FW shapes: the fourth wave, which has no node) makes
    // the first round trip of the face term before phase 1: the element ids
    // of its face nodes, then their neighbour indices, face types, fscale
    // and own nodes (staged in operator 1's flux rows) and the block's face
    // frames and materials (in operator 0's flux rows, after the last
    // element's fluxes).  Those rows are written only after the PML update,
    // and nothing of this is held in registers across phase 1.
    constexpr int NFN = EB * NFP, FIT = S::FW ? (NFN + 63) / 64 : 1;
    constexpr int NFR = EB * 4 * 9, FRL = S::FW ? (NFR + 63) / 64 : 1;
    static_assert(!S::FW || 4 * NFN <= 9 * NT, "face-node data must fit operator 1's flux rows");
    const bool fwave = S::FW && (t >> 6) == S::NW - 1;
    R *sFr = sC + 27 * NP;                    // [EB*4][9] face frames (FW shapes)
    R *sMc = sFr + NFR;                       // [EB][5] materials
    R *sFn = sC + 12 * NT;                    // [4][NFN] neighbour node, face type, own node (bits), fscale
    if constexpr (S::FW) {
        if (fwave) {
            const int l = t & 63;
            int fel[EB];
#pragma unroll
            for (int b = 0; b < EB; ++b) fel[b] = eblk * EB + b < a.nlist ? a.elist[eblk * EB + b] : 0;
            int fnb[FIT], fty[FIT], fown[FIT];
            R ffs[FIT], frv[FRL], mcv;
#pragma unroll
            for (int i = 0; i < FIT; ++i) {
                const int q = l + 64 * i, qq = q < NFN ? q : 0;
                const int fb = qq / NFP, k = qq - fb * NFP, fi = k / NPF, ii = k - fi * NPF;
                int fe = fel[0];
#pragma unroll
                for (int b = 1; b < EB; ++b) fe = fb == b ? fel[b] : fe;
                fown[i] = fb * NP + pload<BUF>(a.fmask, 4u * NPF * 4u, (unsigned)(fi * NPF + ii));
                fty[i] = pload<BUF>(a.ftype, nel * 16u, (unsigned)fe * 4 + fi);
                fnb[i] = pload<BUF>(a.nbr, nel * (unsigned)(NFP * 4), ((unsigned)fe * 4 + fi) * NPF + ii);
                ffs[i] = pload<BUF>(a.fscale, nel * (unsigned)(NFP * sizeof(R)), (unsigned)fe * NFP + k);
            }
#pragma unroll
            for (int r = 0; r < FRL; ++r) {
                const int v = l + 64 * r, vv = v < NFR ? v : 0;
                const int face = vv / 9, c = vv - face * 9, fb = face / 4;
                int fe = fel[0];
#pragma unroll
                for (int b = 1; b < EB; ++b) fe = fb == b ? fel[b] : fe;
                frv[r] = pload<BUF>(a.frame, nel * (unsigned)(36 * sizeof(R)), ((unsigned)fe * 4 + face % 4) * 9 + c);
            }
            {
                const int ll = l < EB * 5 ? l : 0, fb = ll / 5, c = ll - fb * 5;
                int fe = fel[0];
#pragma unroll
                for (int b = 1; b < EB; ++b) fe = fb == b ? fel[b] : fe;
                mcv = pload<BUF>(matc, nel * 8u * (unsigned)sizeof(R), (unsigned)fe * 8 + c);
            }
#pragma unroll
            for (int i = 0; i < FIT; ++i) {
                const int q = l + 64 * i;
                if (q < NFN) {
                    sFn[q] = __builtin_bit_cast(R, fnb[i]);
                    sFn[NFN + q] = __builtin_bit_cast(R, fty[i]);
                    sFn[2 * NFN + q] = __builtin_bit_cast(R, fown[i]);
                    sFn[3 * NFN + q] = ffs[i];
                }
            }
#pragma unroll
            for (int r = 0; r < FRL; ++r)
                if (l + 64 * r < NFR) sFr[l + 64 * r] = frv[r];
            if (l < EB * 5) sMc[l] = mcv;
        }
    }

    // Phase 1 issues every load of the node threads in one batch: q, metric
    // and material and (PML) the coefficients and all 27 auxiliaries and RK
    // bases.  No load is guarded by a branch: padding threads read element
    // 0, and in stage 1 the RK base loads return 0 (their descriptors have
    // size 0; the value is not used).
    R qs[9];
    R ef0 = 0, ef1 = 0, ef2 = 0, ef3 = 0, rho = 1;
    R qv[9], mv[9], m2 = 1, m3 = 0, m4 = 0;
    R pc[PML ? 9 : 1], cur[PML ? 27 : 1], bse[PML ? 27 : 1];
    const R *pin = a.pin + (size_t)shot * 27 * ng;
    const R *pbase = a.pbase + (size_t)shot * 27 * ng;
    R *pout = a.pout + (size_t)shot * 27 * ng;
    if (node_thr) {
        src_first = a.src_head[(size_t)shot * (a.nglob / NP) + e];
        rec_slot_e = a.rec_slot[e];
#pragma unroll
        for (int f = 0; f < 9; ++f) qv[f] = pload<BUF>(qin + f * ng, pnb, g32);
#pragma unroll
        for (int k = 0; k < 9; ++k) mv[k] = pload<BUF>(a.metric + k * ng, pnb, g32);
        const unsigned mb = nel * 8u * (unsigned)sizeof(R);
        m2 = pload<BUF>(matc, mb, (unsigned)e * 8 + 2);
        m3 = pload<BUF>(matc, mb, (unsigned)e * 8 + 3);
        m4 = pload<BUF>(matc, mb, (unsigned)e * 8 + 4);
        if constexpr (PML) {
            const R *pmlc = a.pmlc + (size_t)shot * a.pmlc_stride_ng * ng;
#pragma unroll
            for (int k = 0; k < 9; ++k) pc[k] = pload<BUF>(pmlc + k * ng, pnb, g32);
#pragma unroll
            for (int c = 0; c < 27; ++c) cur[c] = pload<BUF>(pin + c * ng, pnb, g32);
#pragma unroll
            for (int c = 0; c < 27; ++c) bse[c] = pload<BUF>(pbase + c * ng, pnb_base, g32);
        }
    }
    if constexpr (PML) __syncthreads();   // the staged fluxes are in LDS
    // Node threads past the end of the list (last block only) carry element
    // 0's values through the node phases: finite, confined to their own
    // columns, never stored.
    if (node_thr) {
#pragma unroll
        for (int f = 0; f < 9; ++f) {
            qs[f] = qv[f];
            sQ[f * NT + t] = qs[f];
        }
#pragma unroll
        for (int k = 0; k < 9; ++k) (KC ? sC + ((k / 3) * 12 + 9 + k % 3) * NT : sM + k * NT)[t] = mv[k];
        rho = m2;
        ef0 = -(m3 + 2 * m4); ef1 = -m3; ef2 = -m4; ef3 = R(-1.0) / rho;
    }
    if constexpr (S::FW) __syncthreads();   // q is in LDS for the face wave
    R cf[KC ? 27 : 1];
    if (fwave) {
        if constexpr (S::FW) {
            // ---- face term on the face wave, while the node threads update
            // the PML fields: the upwind flux of every face node of the
            // block into sF, as the face phase below computes it; the
            // second round trip, the neighbour traces
            const int l = t & 63;
            int fnb[FIT], fty[FIT], fown[FIT];
            R ffs[FIT];
#pragma unroll
            for (int i = 0; i < FIT; ++i) {
                const int q = l + 64 * i, qq = q < NFN ? q : 0;
                fnb[i] = __builtin_bit_cast(int, sFn[qq]);
                fty[i] = __builtin_bit_cast(int, sFn[NFN + qq]);
                fown[i] = __builtin_bit_cast(int, sFn[2 * NFN + qq]);
                ffs[i] = sFn[3 * NFN + qq];
            }
            R fqn[FIT][9];
#pragma unroll
            for (int i = 0; i < FIT; ++i) {
                const unsigned gn = fnb[i] >= 0 ? (unsigned)fnb[i] : 0u;
#pragma unroll
                for (int f = 0; f < 9; ++f) fqn[i][f] = pload<BUF>(qin + f * ng, pnb, gn);
            }
#pragma unroll
            for (int i = 0; i < FIT; ++i) {
                const int q = l + 64 * i;
                if (q < NFN) {
                    const int fb = q / NFP, k = q - fb * NFP, fi = k / NPF;
                    const bool fvalid = eblk * EB + fb < a.nlist;
                    R dq[9], fr[9];
#pragma unroll
                    for (int f = 0; f < 9; ++f) {
                        const R qo = sQ[f * NT + fown[i]];
                        dq[f] = fty[i] >= 0 ? fqn[i][f] - qo : qo;
                    }
#pragma unroll
                    for (int c = 0; c < 9; ++c) fr[c] = sFr[(fb * 4 + fi) * 9 + c];
                    const R *mc = sMc + fb * 5;
                    R fl[9], af[6];
                    upwind_flux<R>(fr, mc[0], mc[1], mc[2], mc[3], mc[4], fty[i] == -1, dq, fl, af, ATT);
                    const R sgn = fty[i] == -2 ? R(-1) : R(1);
                    R *dst = sF + fb * NFC * SFN + k;
#pragma unroll
                    for (int f = 0; f < 9; ++f) dst[f * SFN] = fvalid ? ffs[i] * (sgn * fl[f]) / 2 : R(0);
                    if constexpr (ATT) {
#pragma unroll
                        for (int f = 0; f < 6; ++f) dst[(9 + f) * SFN] = fvalid ? ffs[i] * (sgn * af[f]) / 2 : R(0);
                    }
                }
            }
        }
    } else if (node_thr) {
        if constexpr (PML) {
            // modified fluxes (src/timeloopMod.f90:399-405) and the ADE update
            // of f',g',h' (src/rungekuttaMod.f90:52-63); the update needs only
            // stage-start values, so it is done here and the auxiliaries die
            R Fm[9], Gm[9], Hm[9];
            elastic_flux_node(qs, ef0, ef1, ef2, ef3, Fm, Gm, Hm);
            const RkStage<R> rs = rk_stage<R>(a.scheme, a.irk, a.w13, a.w23, a.lsa, a.lsb);
            // a shot whose PML is switched off (src/timeloopMod.f90:532-539)
            // keeps its PML elements in this launch with the modification
            // and the ADE update made neutral: auxiliaries read as 0,
            // alpha = d/k = 0, 1/k = 1, so the fluxes are the elastic ones
            const bool on = !a.pml_on || a.pml_on[shot] != 0;
            // F_4, G_5 and H_3 are identically zero (src/waveMod.f90:37-62) and
            // left out of the volume columns.  Contravariant form: C_o,i = o_x
            // F_i + o_y G_i + o_z H_i, o = r, s, t with the node's metric
            // (mv = rx ry rz sx sy sz tx ty tz)
            int col = 0;
#pragma unroll
            for (int d = 0; d < 3; ++d) {
                R pv[9];
                const R al = on ? pc[d] : R(0), dk = on ? pc[3 + d] : R(0);
                // (f' + F) / k as a product with 1/k: one division per
                // direction instead of nine, within an ulp of the quotient
                const R rk = on ? R(1) / pc[6 + d] : R(1);
                R *FF = d == 0 ? Fm : (d == 1 ? Gm : Hm);
                const int skip = d == 0 ? 4 : (d == 1 ? 5 : 3);
#pragma unroll
                for (int i = 0; i < 9; ++i) {
                    const int c = d * 9 + i;
                    const R c0 = on ? cur[c] : R(0);
                    const R mod = (c0 + FF[i]) * rk;
                    if constexpr (KC) {
#pragma unroll
                        for (int o = 0; o < 3; ++o) {
                            const R m = mv[3 * o + d];
                            if (d == 0) cf[o * 9 + i] = i != skip ? m * mod : R(0);
                            else if (i != skip) cf[o * 9 + i] = cf[o * 9 + i] + m * mod;
                        }
                    } else if (i != skip) {
                        sX[(col++) * NT + t] = mod;
                    }
                    // read unconditionally and selected (the region is
                    // stale without compat): no branch between components
                    const R lastv = sLast[c * NP + j];
                    const R u = a.pml_compat ? lastv : mod;
                    const R dadt = -al * c0 - dk * (c0 + u);
                    pv[i] = rk_update<R, LS>(rs, c0, bse[c], dadt, dt);
                }
                // one guarded block of stores per direction (a guard per
                // store leaves waits for every earlier store at the joins)
                if (valid) {
#pragma unroll
                    for (int i = 0; i < 9; ++i) pstore<BUF>(pout + (d * 9 + i) * ng, pnb, g32, pv[i]);
                }
                if constexpr (LS) {
                    if (valid) {
                        R *pres = a.pres + (size_t)shot * 27 * ng;
#pragma unroll
                        for (int i = 0; i < 9; ++i) pstore<BUF>(pres + (d * 9 + i) * ng, pnb, g32, bse[d * 9 + i]);
                    }
                }
                // the nine updates of a direction interleave; a scheduling
                // fence per direction keeps the working set in registers
                __builtin_amdgcn_sched_barrier(0);
            }
        }
    }
    // the staged fluxes (and the face wave's frames) are read: the flux rows
    // may be written
    if constexpr (S::FW) __syncthreads();
    if constexpr (KC) {
        if (node_thr) {
#pragma unroll
            for (int c = 0; c < 27; ++c) sC[((c / 9) * 12 + c % 9) * NT + t] = cf[c];
        }
    }
    __syncthreads();

    // LIFT fragments of node tile 0 (operator data, from L2): issued before
    // the face term, whose latency covers theirs
    const R *lfr = a.LT + (t & 63);
    R lfa[S::KSL];
#pragma unroll
    for (int ks = 0; ks < S::KSL; ++ks) lfa[ks] = lfr[ks * 64];

    // ---- face term: upwind flux per face node into sF (done above by the
    // face wave in the FW shapes) ---------------------------------------------
    for (int it = t; it < (S::FW ? 0 : EB * NFP); it += NTHR) {
        const int fb = it / NFP, k = it - fb * NFP;
        const int fidx = eblk * EB + fb;
        // value f of this face node at dst[f * SFN]: consecutive threads
        // write consecutive words
        R *dst = sF + fb * NFC * SFN + k;
        if (fidx >= a.nlist) {
#pragma unroll
            for (int f = 0; f < NFC; ++f) dst[f * SFN] = 0;
            continue;
        }
        const int fe = sElem[fb];
        const int fi = k / NPF, i = k - fi * NPF;
        const int own = fb * NP + pload<BUF>(a.fmask, 4u * NPF * 4u, (unsigned)(fi * NPF + i));
        const int ft = pload<BUF>(a.ftype, nel * 16u, (unsigned)fe * 4 + fi);
        // the neighbour trace is loaded unconditionally (node 0 stands in on
        // boundary faces) so the nine loads issue together
        const int nb = pload<BUF>(a.nbr, nel * (unsigned)(NFP * 4), ((unsigned)fe * 4 + fi) * NPF + i);
        const unsigned gn = nb >= 0 ? (unsigned)nb : 0u;
        R dq[9];
#pragma unroll
        for (int f = 0; f < 9; ++f) {
            const R qn = pload<BUF>(qin + f * ng, pnb, gn), qo = sQ[f * NT + own];
            dq[f] = ft >= 0 ? qn - qo : qo;
        }
        const unsigned mb = nel * 8u * (unsigned)sizeof(R);
        R mc[5], fr[9];
#pragma unroll
        for (int c = 0; c < 5; ++c) mc[c] = pload<BUF>(matc, mb, (unsigned)fe * 8 + c);
#pragma unroll
        for (int c = 0; c < 9; ++c) fr[c] = pload<BUF>(a.frame, nel * (unsigned)(36 * sizeof(R)), ((unsigned)fe * 4 + fi) * 9 + c);
        R fl[9], af[6];
        upwind_flux<R>(fr, mc[0], mc[1], mc[2], mc[3], mc[4], ft == -1, dq, fl, af, ATT);
        const R sgn = ft == -2 ? R(-1) : R(1);
        const R fs = pload<BUF>(a.fscale, nel * (unsigned)(NFP * sizeof(R)), (unsigned)fe * NFP + k);
#pragma unroll
        for (int f = 0; f < 9; ++f) dst[f * SFN] = fs * (sgn * fl[f]) / 2;
        if constexpr (ATT) {
#pragma unroll
            for (int f = 0; f < 6; ++f) dst[(9 + f) * SFN] = fs * (sgn * af[f]) / 2;
        }
    }
    __syncthreads();

    // The memory variables and anelastic coefficients of the assembly are
    // loaded in one batch at the assembly (loading them before the
    // matrix-core phase measured neutral and costs registers).
    R th[ATT ? 18 : 1], avc[ATT ? 9 : 1];
    auto load_memory_variables = [&]() {
        if constexpr (ATT) {
            if (node_thr) {
                const R *thin = a.thin + (size_t)shot * 18 * ng;
                const R *av = a.anel + (size_t)shot * a.anel_stride;
#pragma unroll
                for (int c = 0; c < 18; ++c) {
                    const R v = pload<BUF>(thin + c * ng, pnb, g32);
                    th[c] = valid ? v : R(0);
                }
#pragma unroll
                for (int c = 0; c < 9; ++c) {
                    const R v = pload<BUF>(av, nel * (unsigned)(NMB * 3 * sizeof(R)), (unsigned)e * (NMB * 3) + c);
                    avc[c] = valid ? v : R(0);
                }
            }
        }
    };

    // ---- operand fragments into registers ------------------------------------
    constexpr int NW = S::NW, CTL = S::CTL, CPW = S::CPW, KS = S::KS, RT = S::RT, NCT = S::NCT;
    constexpr int MTL = S::MTL, KSL = S::KSL, EPW = S::EPW;
    const int lane = t & 63;
    const int w = __builtin_amdgcn_readfirstlane(t >> 6);
    // volume with [Dr;Ds;Dt]: each wave takes column tiles w, w+NW, ... of
    // the block's EB*NCOL columns (a q field, a PML flux column or, with PML
    // and attenuation, a velocity column): B fragment = column 16*tile + l%16
    // at node k = 4*ks + l/16
    R xb[KC ? 1 : CPW][KS];
    if constexpr (!KC) {
#pragma unroll
        for (int c = 0; c < CPW; ++c) {
            const int col = (w + c * NW) * 16 + (lane & 15);
            const bool cok = w + c * NW < CTL && col < NCT;
            const int ceb = cok ? col / NCOL : 0, cc = cok ? col - (col / NCOL) * NCOL : 0;
            const R *colp;
            if constexpr (PML) colp = cc < 24 ? sX + cc * NT : sQ + (6 + cc - 24) * NT;
            else colp = sQ + cc * NT;
            colp += ceb * NP;
#pragma unroll
            for (int ks = 0; ks < KS; ++ks) {
                const int m = ks * 4 + (lane >> 4);
                const R v = colp[m < NP ? m : 0];
                xb[c][ks] = (cok && m < NP) ? v : R(0);
            }
        }
    }
    // LIFT: wave w takes elements NW-1-w, 2NW-1-w, ... (the last waves have
    // the fewest volume tiles): B fragment = face value l%16 (of NFC) at face
    // node 4*ks + l/16
    R fb[EPW][KSL];
#pragma unroll
    for (int c = 0; c < EPW; ++c) {
        const int el = (NW - 1 - w) + c * NW;
        const int f = lane & 15;
        const bool ok = el < EB && f < NFC;
        const R *fp = sF + ((ok ? el : 0) * NFC + (ok ? f : 0)) * SFN + (lane >> 4);
#pragma unroll
        for (int ks = 0; ks < KSL; ++ks) {
            const R v = fp[ks * 4];
            fb[c][ks] = ok ? v : R(0);
        }
    }
    // the face values are dead from here (the contravariant form's lifted
    // values reuse them); otherwise the columns are in registers and the
    // volume values reuse their region
    __syncthreads();

    // ---- LIFT on the matrix cores: sL[f][node] = sum_k LIFT(node,k) F(k,f)
    // (src/timeloopMod.f90:463-465); the face values already carry fscale/2 --
    R vfa[KC ? 1 : KS];   // volume operator fragments of row tile 0
    // contravariant form: the leading chunks (3 k-steps of every node tile)
    // of [Dr|Ds|Dt] for the first column tile
    constexpr int KPRE = KC ? (S::RTV <= 3 ? 2 : 1) : 0;
    R kfa[KC ? KPRE * 3 * S::RTV : 1];
    {
        using V4 = typename Mfma<R>::v4;
        // node tiles outermost: each tile's fragments load once for all of
        // the wave's elements, tile mt+1's while mt computes
        R lfb[KSL];
#pragma unroll
        for (int mt = 0; mt < MTL; ++mt) {
            R *lc = (mt & 1) ? lfb : lfa;
            R *ln = (mt & 1) ? lfa : lfb;
            if (mt + 1 < MTL) {
#pragma unroll
                for (int ks = 0; ks < KSL; ++ks) ln[ks] = lfr[((mt + 1) * KSL + ks) * 64];
            } else if constexpr (!KC) {
                // the last tile computes while the volume term's first
                // operator fragments load
#pragma unroll
                for (int ks = 0; ks < KS; ++ks) vfa[ks] = a.D[lane + ks * 64];
            } else {
#pragma unroll
                for (int p = 0; p < KPRE; ++p)
#pragma unroll
                    for (int h = 0; h < 3; ++h)
#pragma unroll
                        for (int rt = 0; rt < S::RTV; ++rt)
                            kfa[(p * 3 + h) * S::RTV + rt] =
                                a.DK[lane + (rt * S::KSV + (p < S::KSV / 3 ? p : 0) * 3 + h) * 64];
            }
#pragma unroll
            for (int c = 0; c < EPW; ++c) {
                const int el = (NW - 1 - w) + c * NW;
                if (el >= EB) break;
                V4 acc = {0, 0, 0, 0};
#pragma unroll
                for (int ks = 0; ks < KSL; ++ks) acc = Mfma<R>::fma(lc[ks], fb[c][ks], acc);
                // accumulator element v: node row 4*(l/16)+v (FP32) or
                // 4v + l/16 (FP64) of the tile, face value l%16
#pragma unroll
                for (int v = 0; v < 4; ++v) {
                    const int node = mt * 16 + (sizeof(R) == 8 ? 4 * v + (lane >> 4) : 4 * (lane >> 4) + v);
                    if (node < NP) sL[(lane & 15) * SLN + el * NP + node] = acc[v];
                }
            }
        }
    }

    if constexpr (KC) {
        // ---- volume term, contravariant form: [Dr|Ds|Dt] x contravariant
        // columns on the matrix cores (src/timeloopMod.f90:399-420) -----------
        // Column 16*tile + l%16 is (element ce, kind cc): cc < 9 the fluxes
        // C_o,cc; else strain rate s = cc - 9 (ex ey ez exy eyz exz), whose
        // column is o_k1 v_c1 (+ o_k2 v_c2 off the diagonal).  At k-step ks
        // its B value is node 4*(ks % KS) + l/16 of operator block ks / KS.
        using V4 = typename Mfma<R>::v4;
        constexpr int RTV = S::RTV, KSV = S::KSV;
        constexpr int CPK = (CTL + NW - 1) / NW;   // column tiles per wave
        constexpr int OST = 12 * NT;               // LDS stride between operator blocks
        constexpr int CH = 3;                      // k-steps per chunk of operator fragments
        constexpr int PREF = RTV <= 3 ? 3 : 2;    // chunks in flight (registers allow 3 up to 3 node tiles)
        constexpr int NCH = KSV / CH;              // KSV = 3*KS is a multiple
        V4 acc[CPK][RTV];
        const R *afr = a.DK + lane;
#pragma unroll
        for (int c = 0; c < CPK; ++c) {
            const int ct = w + c * NW;
            if (ct >= CTL) break;
            const int col = ct * 16 + (lane & 15);
            const bool cok = col < NCT;
            const int ce = cok ? col / NCOL : 0, cc = cok ? col - (col / NCOL) * NCOL : 0;
            const bool flux = !ATT || cc < 9;
            const int st = flux ? 0 : cc - 9;
            const int k1 = st < 3 ? st : (st == 4 ? 1 : 0), c1 = st < 3 ? st : (st == 3 ? 1 : 2);
            const int k2 = st == 3 ? 1 : 2, c2 = st == 4 ? 1 : 0;
            const bool two = st >= 3;
            const int r1 = 9 * NT + (flux ? cc : 9 + k1) * NT + ce * NP, q1 = (6 + c1) * NT + ce * NP;
            const int r2 = 9 * NT + (9 + k2) * NT + ce * NP, q2 = (6 + c2) * NT + ce * NP;
            // the lane's column at every k-step
            R bf[KSV];
#pragma unroll
            for (int ks = 0; ks < KSV; ++ks) {
                const int o = ks / KS, m = 4 * (ks - o * KS) + (lane >> 4);
                const int mm = m < NP ? m : 0;   // the operator is zero beyond Np
                const R x1 = lds[r1 + o * OST + mm];
                if constexpr (ATT) {
                    const R y1 = lds[q1 + mm], x2 = lds[r2 + o * OST + mm], y2 = lds[q2 + mm];
                    const R e = x1 * y1 + (two ? x2 * y2 : R(0));
                    bf[ks] = flux ? x1 : e;
                } else {
                    bf[ks] = x1;
                }
            }
#pragma unroll
            for (int rt = 0; rt < RTV; ++rt) acc[c][rt] = V4{0, 0, 0, 0};
            // operator fragments of every node tile in chunks of CH k-steps,
            // the next chunk loading while this one computes
            constexpr int NB = PREF;   // chunks in flight
            static_assert(NB - 1 == KPRE && CH == 3, "leading chunks are loaded during the LIFT");
            R abuf[NB][CH][RTV];
#pragma unroll
            for (int p = 0; p < NB - 1; ++p)
#pragma unroll
                for (int h = 0; h < CH; ++h)
#pragma unroll
                    for (int rt = 0; rt < RTV; ++rt)
                        abuf[p][h][rt] = c == 0 ? kfa[(p * CH + h) * RTV + rt]
                                                : afr[(rt * KSV + (p < NCH ? p : 0) * CH + h) * 64];
#pragma unroll
            for (int ch = 0; ch < NCH; ++ch) {
                if (ch + NB - 1 < NCH) {
#pragma unroll
                    for (int h = 0; h < CH; ++h)
#pragma unroll
                        for (int rt = 0; rt < RTV; ++rt)
                            abuf[(ch + NB - 1) % NB][h][rt] = afr[(rt * KSV + (ch + NB - 1) * CH + h) * 64];
                }
#pragma unroll
                for (int h = 0; h < CH; ++h)
#pragma unroll
                    for (int rt = 0; rt < RTV; ++rt)
                        acc[c][rt] = Mfma<R>::fma(abuf[ch % NB][h][rt], bf[ch * CH + h], acc[c][rt]);
                __builtin_amdgcn_sched_barrier(0);
            }
        }
        __syncthreads();   // q, the fluxes and the metric are dead: sV reuses them
        // accumulator element v: node row 4*(l/16)+v (FP32) or 4v + l/16
        // (FP64) of node tile rt, column l%16 of the column tile
#pragma unroll
        for (int c = 0; c < CPK; ++c) {
            const int ct = w + c * NW;
            if (ct >= CTL) break;
            const int col = ct * 16 + (lane & 15);
            if (col < NCT) {
                const int ce = col / NCOL, cc = col - ce * NCOL;
#pragma unroll
                for (int rt = 0; rt < RTV; ++rt)
#pragma unroll
                    for (int v = 0; v < 4; ++v) {
                        const int node = rt * 16 + (sizeof(R) == 8 ? 4 * v + (lane >> 4) : 4 * (lane >> 4) + v);
                        if (node < NP) sV[(ce * NP + node) * NV + cc] = acc[c][rt][v];
                    }
            }
        }
    } else {
        // ---- volume term: [Dr;Ds;Dt] x columns on the matrix cores, then the
        // metric of each node (src/timeloopMod.f90:407-420) --------------------
        // Accumulator element v of row tile rt is row q = 4rt + v of the
        // lane's group g: node g + 4*(q/3), operator q%3.  A node's three
        // derivatives arrive in order, possibly across two tiles; the metric
        // is applied when the third arrives.  The tiles are unrolled so that
        // this bookkeeping is static.
        using V4 = typename Mfma<R>::v4;
        const R *dfr = a.D + lane;
        const int lg = lane >> 4;
        // operator fragments double-buffered: row tile rt+1 loads while rt
        // computes (vector loads return in order)
        R *fa = vfa, fb2[KS];
        R part[CPW][3];
#pragma unroll
        for (int rt = 0; rt < RT; ++rt) {
            R *af = (rt & 1) ? fb2 : fa;
            R *an = (rt & 1) ? fa : fb2;
            if (rt + 1 < RT) {
#pragma unroll
                for (int ks = 0; ks < KS; ++ks) an[ks] = dfr[((rt + 1) * KS + ks) * 64];
            }
#pragma unroll
            for (int c = 0; c < CPW; ++c) {
                if (w + c * NW >= CTL) break;
                V4 acc = {0, 0, 0, 0};
#pragma unroll
                for (int ks = 0; ks < KS; ++ks) acc = Mfma<R>::fma(af[ks], xb[c][ks], acc);
                const int col = (w + c * NW) * 16 + (lane & 15);
#pragma unroll
                for (int v = 0; v < 4; ++v) {
                    const int q = 4 * rt + v, op = q % 3;
                    part[c][op] = acc[v];
                    if (op != 2) continue;
                    // column 16*tile + l%16, node jr: its three derivatives
                    const int jr = lg + 4 * (q / 3);
                    if (col < NCT && jr < NP) {
                        const int ceb = col / NCOL, cc = col - ceb * NCOL;
                        const R dr = part[c][0], ds = part[c][1], dt = part[c][2];
                        // which metric row(s) the column needs: PML flux columns
                        // one direction, q / velocity columns all three
                        const int kind = PML ? (cc < 8 ? 0 : (cc < 16 ? 1 : (cc < 24 ? 2 : 3))) : 3;
                        const R *met = sM + ceb * NP;
                        R *out = sV + (ceb * NP + jr) * NV;
                        if (kind < 3) {
                            const R m0 = met[(kind + 0) * NT + jr], m1 = met[(kind + 3) * NT + jr],
                                    m2 = met[(kind + 6) * NT + jr];
                            out[cc] = m0 * dr + m1 * ds + m2 * dt;
                        } else {
                            const R rx = met[0 * NT + jr], ry = met[1 * NT + jr], rz = met[2 * NT + jr];
                            const R sx = met[3 * NT + jr], sy = met[4 * NT + jr], sz = met[5 * NT + jr];
                            const R tx = met[6 * NT + jr], ty = met[7 * NT + jr], tz = met[8 * NT + jr];
                            const int base = PML ? 24 + (cc - 24) * 3 : cc * 3;
                            out[base + 0] = rx * dr + sx * ds + tx * dt;
                            out[base + 1] = ry * dr + sy * ds + ty * dt;
                            out[base + 2] = rz * dr + sz * ds + tz * dt;
                        }
                    }
                }
            }
            __builtin_amdgcn_sched_barrier(0);
        }
    }
    __syncthreads();

    // ---- assemble the right-hand side (node threads) -----------------------
    load_memory_variables();
    R thb[ATT ? 18 : 1], qb[9];
    if (node_thr) {
        if constexpr (ATT) {
            const R *thbase = a.thbase + (size_t)shot * 18 * ng;
            if (a.irk == 1) {
#pragma unroll
                for (int c = 0; c < 18; ++c) thb[c] = th[c];
            } else {
#pragma unroll
                for (int c = 0; c < 18; ++c) {
                    const R v = pload<BUF>(thbase + c * ng, pnb, g32);
                    thb[c] = valid ? v : R(0);
                }
            }
        }
        if (a.irk == 1) {
#pragma unroll
            for (int f = 0; f < 9; ++f) qb[f] = qs[f];
        } else {
#pragma unroll
            for (int f = 0; f < 9; ++f) {
                const R v = pload<BUF>(qbase + f * ng, pnb, g32);
                qb[f] = valid ? v : R(0);
            }
        }
    }
    R lf[NFC];
    if (node_thr) {
#pragma unroll
        for (int f = 0; f < NFC; ++f) lf[f] = sL[f * SLN + t];
    }
    R rhs[9];
    R gv[9];   // velocity gradients: gv[c*3+k] = d v_c / d x_k
    R sr[6];   // strain rates ex ey ez exy eyz exz (contravariant form)
    if (node_thr) {
        const R *v = sV + t * NV;
        if constexpr (!PML) {
            R dx[9], dy[9], dz[9];
#pragma unroll
            for (int f = 0; f < 9; ++f) { dx[f] = v[f * 3]; dy[f] = v[f * 3 + 1]; dz[f] = v[f * 3 + 2]; }
            rhs[0] = ef0 * dx[6] + ef1 * dy[7] + ef1 * dz[8];
            rhs[1] = ef1 * dx[6] + ef0 * dy[7] + ef1 * dz[8];
            rhs[2] = ef1 * dx[6] + ef1 * dy[7] + ef0 * dz[8];
            rhs[3] = ef2 * (dx[7] + dy[6]);
            rhs[4] = ef2 * (dy[8] + dz[7]);
            rhs[5] = ef2 * (dx[8] + dz[6]);
            rhs[6] = ef3 * (dx[0] + dy[3] + dz[5]);
            rhs[7] = ef3 * (dx[3] + dy[1] + dz[4]);
            rhs[8] = ef3 * (dx[5] + dy[4] + dz[2]);
#pragma unroll
            for (int c = 0; c < 3; ++c) { gv[c * 3] = dx[6 + c]; gv[c * 3 + 1] = dy[6 + c]; gv[c * 3 + 2] = dz[6 + c]; }
        } else {
            if constexpr (KC) {
#pragma unroll
                for (int i = 0; i < 9; ++i) rhs[i] = v[i];
                if constexpr (ATT) {
#pragma unroll
                    for (int c = 0; c < 6; ++c) sr[c] = v[9 + c];
                }
            } else {
#pragma unroll
                for (int i = 0; i < 9; ++i) {
                    const R vF = i == 4 ? R(0) : v[i < 4 ? i : i - 1];
                    const R vG = i == 5 ? R(0) : v[8 + (i < 5 ? i : i - 1)];
                    const R vH = i == 3 ? R(0) : v[16 + (i < 3 ? i : i - 1)];
                    rhs[i] = vF + vG + vH;
                }
                if constexpr (ATT) {
#pragma unroll
                    for (int c = 0; c < 9; ++c) gv[c] = v[24 + c];
                }
            }
        }
    }

    // ---- attenuation (src/timeloopMod.f90:426-461), one mechanism at a time --
    if constexpr (ATT) {
        if (node_thr) {
            R ex, ey, ez, exy, eyz, exz;
            if constexpr (KC) {
                ex = sr[0]; ey = sr[1]; ez = sr[2]; exy = sr[3]; eyz = sr[4]; exz = sr[5];
            } else {
                ex = gv[0]; ey = gv[4]; ez = gv[8];
                exy = gv[3] + gv[1]; eyz = gv[7] + gv[5]; exz = gv[6] + gv[2];
            }
            R *thout = a.thout + (size_t)shot * 18 * ng;
            const RkStage<R> rs = rk_stage<R>(a.scheme, a.irk, a.w13, a.w23, a.lsa, a.lsb);
#pragma unroll
            for (int i = 0; i < NMB; ++i) {
                const R w = a.wl[i];
                const R vol[6] = {-w * ex, -w * ey, -w * ez, R(-0.5) * w * exy, R(-0.5) * w * eyz, R(-0.5) * w * exz};
                R *tm = th + i * 6;
                const R a1 = avc[i * 3 + 0], a2 = avc[i * 3 + 1], a3 = avc[i * 3 + 2];
                rhs[0] = rhs[0] + a3 * tm[0] + a1 * (tm[1] + tm[2]);
                rhs[1] = rhs[1] + a3 * tm[1] + a1 * (tm[2] + tm[0]);
                rhs[2] = rhs[2] + a3 * tm[2] + a1 * (tm[0] + tm[1]);
                rhs[3] = rhs[3] + a2 * tm[3];
                rhs[4] = rhs[4] + a2 * tm[4];
                rhs[5] = rhs[5] + a2 * tm[5];
                // the stage-start values are used up: th takes the new ones
#pragma unroll
                for (int c = 0; c < 6; ++c) {
                    const int k = i * 6 + c;
                    const R arq = vol[c] + w * lf[9 + c] - w * th[k];
                    th[k] = rk_update<R, LS, sizeof(R) == 4>(rs, th[k], thb[k], arq, dt);
                }
            }
            // one guarded block of stores
            if (valid) {
#pragma unroll
                for (int k = 0; k < 18; ++k) pstore<BUF>(thout + k * ng, pnb, g32, th[k]);
            }
            if constexpr (LS) {
                if (valid) {
                    R *thres = a.thres + (size_t)shot * 18 * ng;
#pragma unroll
                    for (int k = 0; k < 18; ++k) pstore<BUF>(thres + k * ng, pnb, g32, thb[k]);
                }
            }
        }
    }

    R qnew[9];
    if (node_thr) {
#pragma unroll
        for (int f = 0; f < 9; ++f) rhs[f] = rhs[f] + lf[f];
        // sources (src/timeloopMod.f90:467-475)
        if (valid) {
            const int *next = a.src_next + (size_t)shot * a.nsrc;
            const int *stype = a.srctype + (size_t)shot * a.nsrc;
            for (int s = src_first; s >= 0; s = next[s]) {
                if (a.src_factored) {
                    // spatial operand times the source time function of the step
                    const R *sp = a.src_spatial + (((size_t)shot * a.nsrc + s) * 6) * NP + j;
                    const R amp = a.src_stf[((size_t)shot * a.nsrc + s) * a.nt + it_step];
                    if (stype[s] == 0) {
                        rhs[6] += sp[0] * amp; rhs[7] += sp[NP] * amp; rhs[8] += sp[2 * NP] * amp;
                    } else {
#pragma unroll
                        for (int c = 0; c < 6; ++c) rhs[c] += sp[c * NP] * amp;
                    }
                } else if (stype[s] == 0) {
                    const R *sp = a.src_array + (((size_t)it_step * a.nsrc + s) * 3) * NP + j;
                    rhs[6] += sp[0]; rhs[7] += sp[NP]; rhs[8] += sp[2 * NP];
                } else {
                    const R *sp = a.src_array_m + (((size_t)it_step * a.nsrc + s) * 6) * NP + j;
#pragma unroll
                    for (int c = 0; c < 6; ++c) rhs[c] += sp[c * NP];
                }
            }
        }
        // Runge-Kutta update
        if (!LS && a.qlo_out) {
            const size_t so = (size_t)shot * 9 * ng;
            R alo[9], blo[9];
#pragma unroll
            for (int f = 0; f < 9; ++f) {
                const R v = pload<BUF>(a.qlo_in + so + f * ng, pnb, g32);
                alo[f] = valid ? v : R(0);
            }
#pragma unroll
            for (int f = 0; f < 9; ++f) {
                const R v = pload<BUF>(a.qlo_base + so + f * ng, pnb_base, g32);
                blo[f] = (a.irk == 1 || !valid) ? alo[f] : v;
            }
#pragma unroll
            for (int f = 0; f < 9; ++f) {
                R lo;
                rk_combine_comp<R>(a.scheme, a.irk, qs[f], alo[f], qb[f], blo[f], rhs[f], dt, a.w13, a.w23, qnew[f],
                                   lo);
                if (valid) { pstore<BUF>(qout + f * ng, pnb, g32, qnew[f]); pstore<BUF>(a.qlo_out + so + f * ng, pnb, g32, lo); }
            }
        } else {
            const RkStage<R> rs = rk_stage<R>(a.scheme, a.irk, a.w13, a.w23, a.lsa, a.lsb);
#pragma unroll
            for (int f = 0; f < 9; ++f) qnew[f] = rk_update<R, LS, sizeof(R) == 4>(rs, qs[f], qb[f], rhs[f], dt);
            if (valid) {
#pragma unroll
                for (int f = 0; f < 9; ++f) pstore<BUF>(qout + f * ng, pnb, g32, qnew[f]);
            }
            if constexpr (LS) {
                if (valid) {
                    R *qres = a.qres + (size_t)shot * 9 * ng;
#pragma unroll
                    for (int f = 0; f < 9; ++f) pstore<BUF>(qres + f * ng, pnb, g32, qb[f]);
                }
            }
        }
    }

    // ---- post-step (src/timeloopMod.f90:482-516) ---------------------------
    if (a.last_stage) {
        if (valid) {
            if (a.u_full) {
#pragma unroll
                for (int c = 0; c < 3; ++c) a.u_full[c * ng + g] = a.u_full[c * ng + g] + dt * qnew[6 + c];
            }
            const int slot = rec_slot_e;
            if (slot >= 0) {
                const size_t so = (size_t)shot * a.nslot + slot;
#pragma unroll
                for (int c = 0; c < 3; ++c) {
                    R *u = a.u_slot + (so * 3 + c) * NP + j;
                    if (a.u_slot_lo) {
                        // displacement by forward Euler, compensated (Kahan)
                        R *ul = a.u_slot_lo + (so * 3 + c) * NP + j;
                        R s, e;
                        two_sum(*u, dt * qnew[6 + c], s, e);
                        fast_two_sum(s, e + *ul, *u, *ul);
                    } else {
                        *u = *u + dt * qnew[6 + c];
                    }
                    a.a_slot[(so * 3 + c) * NP + j] = rhs[6 + c];
                }
            }
        }
        if (a.blk_energy) {
            // energy partial of this block (src/timeloopMod.f90:490-516), summed
            // in FP64.  Per element e and field pair: sum over nodes of
            // x . (M y), with y the velocities and the reference's strain
            // columns and x the velocities and stresses.  The products M y run
            // on the matrix cores: M (node tiles of 16, k-steps of 4 nodes,
            // fragments from a.mass) times the block's EB*9 columns y, one
            // column per lane; each lane then dots its column with x and
            // scales by its element's coefficient.  The reference's normal
            // "strains" are stress combinations divided by temp1 = 1/(4 mu^2 +
            // 6 mu lam), that is multiplied by ce = 4 mu^2 + 6 mu lam; with
            // stresses of a few 1e10 the products leave the FP32 range, so FP32
            // carries the combinations through the mass product and applies
            // ce in FP64.  FP64 follows the reference.
            constexpr bool F32 = sizeof(R) == 4;
            constexpr int NKE = (NP + 3) / 4, MTE = (NP + 15) / 16;   // k-steps, node tiles of M
            constexpr int NCE = EB * 9, CTE = (NCE + 15) / 16;          // columns, column tiles
            __syncthreads();
            R *sE = lds;                  // [15][NT] y (rows 6..14 are the columns' values)
            R *sX = lds + 15 * NT;        // [9][NT]  x = q after the step
            R *sEl = lds + 24 * NT;       // [EB][4]  rho, mu, lambda, J of each element
            static_assert(24 * NT + 4 * EB <= S::LDS, "energy values must fit the LDS");
            if (node_thr) {
                R lam_e = 0, mu_e = 0;
                if (valid) {
                    const R *em = a.emat + (size_t)shot * a.emat_stride + (size_t)e * 2;
                    mu_e = em[0]; lam_e = em[1];
                }
                const R temp1 = R(1.) / (4 * mu_e * mu_e + 6 * mu_e * lam_e);
                const R temp2 = 2 * mu_e;
                R y[15];
#pragma unroll
                for (int c = 0; c < 9; ++c) y[c] = qnew[c];
                const R c0 = 2 * (lam_e + mu_e) * qnew[0] - lam_e * (qnew[1] + qnew[2]);
                const R c1 = 2 * (lam_e + mu_e) * qnew[1] - lam_e * (qnew[0] + qnew[2]);
                const R c2 = 2 * (lam_e + mu_e) * qnew[2] - lam_e * (qnew[0] + qnew[0]);
                y[9] = F32 ? c0 : c0 / temp1;
                y[10] = F32 ? c1 : c1 / temp1;
                y[11] = F32 ? c2 : c2 / temp1;
                y[12] = qnew[3] / temp2; y[13] = qnew[4] / temp2; y[14] = qnew[5] / temp2;
#pragma unroll
                for (int c = 6; c < 15; ++c) sE[c * NT + t] = valid ? y[c] : R(0);
#pragma unroll
                for (int c = 0; c < 9; ++c) sX[c * NT + t] = valid ? qnew[c] : R(0);
                if (j == 0) {
                    R *el = sEl + eb * 4;
                    el[0] = valid ? rho : R(0);
                    el[1] = mu_e;
                    el[2] = lam_e;
                    el[3] = valid ? a.jac[(size_t)e * NP] : R(0);
                }
            }
            __syncthreads();
            double part = 0;
            {
                using V4 = typename Mfma<R>::v4;
                constexpr int NWE = NTHR / 64, CPWE = (CTE + NWE - 1) / NWE;   // column tiles per wave
                const int lane = t & 63, w = __builtin_amdgcn_readfirstlane(t >> 6);
#pragma unroll
                for (int cw = 0; cw < CPWE; ++cw) {
                    const int ct = w + cw * NWE;
                    if (ct >= CTE) break;
                    // column kind 0..2: velocities (x = v, y = v), 3..5: normal
                    // (x = sigma_ii, y = strain combination), 6..8: shear
                    const int col = ct * 16 + (lane & 15);
                    const bool cok = col < NCE;
                    const int ce = cok ? col / 9 : 0, kind = cok ? col - (col / 9) * 9 : 0;
                    const int xr = kind < 3 ? 6 + kind : kind - 3;
                    const R *yc = sE + (6 + kind) * NT + ce * NP;
                    const R *xc = sX + xr * NT + ce * NP;
                    R bfr[NKE];
#pragma unroll
                    for (int ks = 0; ks < NKE; ++ks) {
                        const int m = 4 * ks + (lane >> 4);
                        bfr[ks] = yc[m < NP ? m : 0];   // M is zero beyond Np
                    }
                    double dsum = 0;
#pragma unroll
                    for (int mt = 0; mt < MTE; ++mt) {
                        R afr[NKE];
#pragma unroll
                        for (int ks = 0; ks < NKE; ++ks) afr[ks] = a.mass[(mt * NKE + ks) * 64 + lane];
                        V4 acc = {0, 0, 0, 0};
#pragma unroll
                        for (int ks = 0; ks < NKE; ++ks) acc = Mfma<R>::fma(afr[ks], bfr[ks], acc);
#pragma unroll
                        for (int v = 0; v < 4; ++v) {
                            const int node = mt * 16 + (sizeof(R) == 8 ? 4 * v + (lane >> 4) : 4 * (lane >> 4) + v);
                            if (node < NP) dsum += (double)(xc[node] * acc[v]);
                        }
                    }
                    const R *el = sEl + ce * 4;
                    const double mu = el[1], lam = el[2];
                    const double coef = kind < 3 ? (double)el[0] : (kind < 6 ? (F32 ? 4 * mu * mu + 6 * mu * lam : 1.0)
                                                                               : 2.0);
                    part += cok ? 0.5 * (double)el[3] * coef * dsum : 0.0;
                }
            }
            // block sum in a fixed order: within each wave, then over the waves
#pragma unroll
            for (int off = 32; off > 0; off >>= 1) part += __shfl_down(part, off, 64);
            __shared__ double sWave[NTHR / 64];
            if ((t & 63) == 0) sWave[t >> 6] = part;
            __syncthreads();
            if (t == 0) {
                double sum = 0;
#pragma unroll
                for (int i = 0; i < NTHR / 64; ++i) sum += sWave[i];
                a.blk_energy[(size_t)shot * a.blk_stride + eblk] = sum;
            }
        }
    }
}

// Fluxes of the last element for the PML ADE in compat mode (H1): one block
// per shot, one thread per node of element nelem-1.
template <int ORDER, typename R>
__global__ void last_flux_kernel(const R *qin, const R *pin, const R *pmlc, const R *matc, int last, int last_is_pml,
                                 int nglob, int mat_stride, int pmlc_stride_ng, R *out) {
    constexpr int NP = Dims<ORDER>::NP;
    const int j = threadIdx.x;
    const int shot = blockIdx.x;
    if (j >= NP) return;
    const size_t ng = (size_t)nglob, g = (size_t)last * NP + j;
    qin += (size_t)shot * 9 * ng;
    pin += (size_t)shot * 27 * ng;
    pmlc += (size_t)shot * pmlc_stride_ng * ng;
    matc += (size_t)shot * mat_stride;
    out += (size_t)shot * 27 * NP;
    R q[9];
#pragma unroll
    for (int f = 0; f < 9; ++f) q[f] = qin[f * ng + g];
    const R *mc = matc + (size_t)last * 8;
    const R lam = mc[3], mu = mc[4], rho = mc[2];
    R F[9], G[9], H[9];
    elastic_flux_node<R>(q, -(lam + 2 * mu), -lam, -mu, R(-1.0) / rho, F, G, H);
    if (last_is_pml) {
        const R kx = pmlc[6 * ng + g], ky = pmlc[7 * ng + g], kz = pmlc[8 * ng + g];
#pragma unroll
        for (int i = 0; i < 9; ++i) {
            F[i] = (pin[i * ng + g] + F[i]) / kx;
            G[i] = (pin[(9 + i) * ng + g] + G[i]) / ky;
            H[i] = (pin[(18 + i) * ng + g] + H[i]) / kz;
        }
    }
#pragma unroll
    for (int i = 0; i < 9; ++i) {
        out[i * NP + j] = F[i];
        out[(9 + i) * NP + j] = G[i];
        out[(18 + i) * NP + j] = H[i];
    }
}

// Receiver sampling (src/timeloopMod.f90:358-373) of row `row_offset + *d_it`
// for every shot, each shot's energy of the step (sum of the stage kernels'
// block partials in a fixed order), and the step-counter advance.
//
// Grid (nshot, chunks) of 1024-thread blocks: block (s, c) samples values
// 1024 c .. 1024 c + 1023 of shot s's nrec x 9 receiver values, and block
// (s, 0) also sums shot s's energy (the same order for any grid).  Every
// block reads the step counter first; with more than one block the last
// block to finish (counted in *d_done, which it puts back to 0) advances it,
// so no block can read the next step's value.  *d_done starts at 0 and is
// only used when the grid has more than one block.
template <int ORDER, typename R>
__global__ void __launch_bounds__(1024)
step_end_kernel(const R *q, const R *u_slot, const R *a_slot, const R *rec_int, const int *recelem,
                const int *rec_slot_of, int nrec, int nslot, int nglob, int nt, int row_offset,
                int advance, R *seis, const double *blk_energy, int nblk, double *energy_out, int *d_it,
                unsigned *d_done) {
    constexpr int NP = Dims<ORDER>::NP;
    __shared__ int s_it;
    __shared__ double s_red[1024];
    if (threadIdx.x == 0) s_it = *d_it;
    __syncthreads();
    const int it = s_it;
    const int row = it + row_offset;
    const int shot = (int)blockIdx.x;
    if (row < nt) {
        const R *qs = q + (size_t)shot * 9 * nglob;
        for (int rk = blockIdx.y * blockDim.x + threadIdx.x; rk < nrec * 9; rk += gridDim.y * blockDim.x) {
            const int r = rk / 9, k = rk - r * 9;
            const int e = recelem[r];
            const size_t so = (size_t)shot * nslot + rec_slot_of[r];
            const R *wt = rec_int + (size_t)r * NP;
            R acc = 0;
            for (int jn = 0; jn < NP; ++jn) {
                R v;
                if (k < 3) v = u_slot[(so * 3 + k) * NP + jn];
                else if (k < 6) v = qs[(size_t)(6 + k - 3) * nglob + (size_t)e * NP + jn];
                else v = a_slot[(so * 3 + (k - 6)) * NP + jn];
                acc = acc + v * wt[jn];
            }
            seis[(((size_t)shot * 9 + k) * nt + row) * nrec + r] = acc;
        }
    }
    if (advance && blk_energy && blockIdx.y == 0) {
        // energy_out[shot][nt]: the shot's partials [shot][nblk] in a fixed order
        const double *pb = blk_energy + (size_t)shot * nblk;
        // eight independent partial sums, so eight loads are in flight per
        // round trip; combined in a fixed order
        constexpr int U = 8;
        double acc[U];
#pragma unroll
        for (int u = 0; u < U; ++u) acc[u] = 0;
        int b = threadIdx.x;
        const int stride = blockDim.x;
        for (; b + (U - 1) * stride < nblk; b += U * stride) {
#pragma unroll
            for (int u = 0; u < U; ++u) acc[u] += pb[b + u * stride];
        }
        for (; b < nblk; b += stride) acc[0] += pb[b];
        s_red[threadIdx.x] = ((acc[0] + acc[1]) + (acc[2] + acc[3])) + ((acc[4] + acc[5]) + (acc[6] + acc[7]));
        __syncthreads();
        for (int s = blockDim.x / 2; s > 0; s >>= 1) {
            if ((int)threadIdx.x < s) s_red[threadIdx.x] += s_red[threadIdx.x + s];
            __syncthreads();
        }
        if (threadIdx.x == 0 && it < nt) energy_out[(size_t)shot * nt + it] = s_red[0];
    }
    if (!advance) return;
    __syncthreads();
    if (threadIdx.x != 0) return;
    const unsigned nblocks = gridDim.x * gridDim.y;
    if (nblocks == 1) {
        *d_it = it + 1;
        return;
    }
    // this block's read of the counter is done (before the barrier); the
    // last arrival advances it
    __threadfence();
    if (atomicAdd(d_done, 1u) == nblocks - 1) {
        atomicExch(d_done, 0u);
        *d_it = it + 1;
    }
}

}  // namespace nexd
