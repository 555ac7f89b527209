// libnexd_gpu.so: the C-ABI of include/nexd_gpu.h over the HIP stage kernels.
//
// One context = one HIP stream on one device holding the whole mesh.  The
// uploads convert the reference's column-major, 1-based arrays into the
// device layout of plain_kernels.hpp; nexd_run replays a captured HIP graph
// of one time step (all RK stages plus the step-end kernel) once per step.
#include <hip/hip_runtime.h>

#include <algorithm>

#include <cmath>
#include <cstdio>
#include <cstring>
#include <mutex>
#include <string>
#include <utility>
#include <vector>

#include "../../include/nexd_gpu.h"
#include "host_layout.hpp"
#include "plain_kernels.hpp"
#include "stage_launch.hpp"

using nexd::Dims;
using nexd::StageArgs;

namespace {

thread_local std::string g_create_error;

template <typename T>
struct DevBuf {
    T *p = nullptr;
    size_t n = 0;
};

}  // namespace

struct nexd_ctx {
    nexd_config cfg{};
    int np = 0, npf = 0, nfp = 0, nrk = 0;
    size_t nglob = 0;
    hipStream_t stream = nullptr;
    std::string err;
    std::vector<std::pair<void *, size_t>> allocs;   // device buffers and their sizes in bytes
    // wavefield
    double *q[3] = {nullptr, nullptr, nullptr};
    double *qres = nullptr, *thres = nullptr, *pres = nullptr;   // low-storage RK4 residuals
    double *th[2] = {nullptr, nullptr};
    double *paux[2] = {nullptr, nullptr};
    double *u_full = nullptr;
    // operators and mesh
    double *D = nullptr, *LT = nullptr, *mass = nullptr;
    int *fmask = nullptr;
    std::vector<int> fmask_host;
    double *metric = nullptr, *fscale = nullptr, *frame = nullptr, *jac = nullptr;
    int *nbr = nullptr, *ftype = nullptr;
    double *matc = nullptr, *emat = nullptr, *anel = nullptr, *pmlc = nullptr;
    double wl[nexd::NMB] = {0, 0, 0};
    std::vector<int> pmlcheck;
    int *elist_plain = nullptr, *elist_pml = nullptr;
    int n_plain = 0, n_pml = 0;
    bool pml_active = false;
    // sources / receivers
    int *src_head = nullptr, *src_next = nullptr, *srctype = nullptr;
    double *src_array = nullptr, *src_array_m = nullptr;
    int *recelem = nullptr, *rec_slot = nullptr, *rec_slot_of = nullptr;
    double *rec_int = nullptr, *u_slot = nullptr, *a_slot = nullptr;
    int nslot = 0;
    double *seis = nullptr;
    // per step
    double *last_flux = nullptr, *blk_energy = nullptr, *energy = nullptr;
    int nblk_total = 0;
    int *d_it = nullptr;
    unsigned *d_done = nullptr;   // blocks of the step-end launch that finished
    int step = 0;                 // next step, 0-based
    bool have_ops = false, have_mesh = false, have_mat = false, have_att = false, have_pml = false,
         have_src = false, have_rec = false, initialized = false;
    hipGraphExec_t graph = nullptr;
    hipEvent_t ev0 = nullptr, ev1 = nullptr;
    float last_ms = 0.f;
};

namespace {

int fail(nexd_ctx *c, int code, const std::string &msg) {
    if (c) c->err = msg;
    else g_create_error = msg;
    return code;
}

#define HIPCHK(ctx, expr)                                                                   \
    do {                                                                                    \
        hipError_t _e = (expr);                                                             \
        if (_e != hipSuccess)                                                               \
            return fail(ctx, NEXD_E_HIP, std::string(#expr ": ") + hipGetErrorString(_e));   \
    } while (0)

template <typename T>
int dalloc(nexd_ctx *c, T **p, size_t n) {
    void *v = nullptr;
    if (n == 0) n = 1;
    hipError_t e = hipMalloc(&v, n * sizeof(T));
    if (e != hipSuccess) return fail(c, NEXD_E_HIP, std::string("hipMalloc: ") + hipGetErrorString(e));
    c->allocs.push_back({v, n * sizeof(T)});
    *p = static_cast<T *>(v);
    return 0;
}

// A device buffer of at least n elements at *p: the existing one when it is
// large enough, else a new one (the old one is freed; a captured step graph
// holding it is dropped by every caller of this function).
template <typename T>
int ensure(nexd_ctx *c, T **p, size_t n) {
    if (n == 0) n = 1;
    if (*p) {
        for (size_t i = 0; i < c->allocs.size(); ++i) {
            if (c->allocs[i].first != (void *)*p) continue;
            if (c->allocs[i].second >= n * sizeof(T)) return 0;
            HIPCHK(c, hipDeviceSynchronize());
            (void)hipFree(c->allocs[i].first);
            c->allocs.erase(c->allocs.begin() + (long)i);
            break;
        }
        *p = nullptr;
    }
    return dalloc(c, p, n);
}

template <typename T>
int upload(nexd_ctx *c, T **p, const std::vector<T> &h) {
    int rc = ensure(c, p, h.size());
    if (rc) return rc;
    HIPCHK(c, hipMemcpy(*p, h.data(), h.size() * sizeof(T), hipMemcpyHostToDevice));
    return 0;
}

void drop_graph(nexd_ctx *c) {
    if (c->graph) {
        (void)hipGraphExecDestroy(c->graph);
        c->graph = nullptr;
    }
}

// ---------------------------------------------------------------------------
// stage launches (stage_launch.hpp: the kernels are instantiated in the
// stage_*.hip translation units)
// ---------------------------------------------------------------------------
int nblocks(int order, int n, bool att, bool pml) { return nexd::stage_blocks<double>(order, n, att, pml); }

void dispatch_stage(int order, const StageArgs<double> &a, bool att, bool pml, hipStream_t s) {
    if (a.scheme == NEXD_LSRK4) nexd::launch_stage<double, true, 0>(order, att, pml, a, s);
    else nexd::launch_stage<double, false, 0>(order, att, pml, a, s);
}

// Where a step's kernels keep their scratch: the energy block partials
// (nblk_total values) and the last element's 27 fluxes (27*Np values).
// nexd_run uses the context's buffers, nexd_run_noalloc a static array.
struct StepScratch {
    double *blk_energy;
    double *last_flux;
};

template <int ORDER>
void launch_last_flux(nexd_ctx *c, const double *qin, const double *pin, double *out) {
    const int last = c->cfg.nelem - 1;
    hipLaunchKernelGGL((nexd::last_flux_kernel<ORDER, double>), dim3(1), dim3(64), 0, c->stream, qin, pin, c->pmlc,
                       c->matc, last, c->pmlcheck[last] ? 1 : 0, (int)c->nglob, 0, 0, out);
}

void dispatch_last_flux(nexd_ctx *c, const double *qin, const double *pin, double *out) {
    switch (c->cfg.order) {
        case 1: launch_last_flux<1>(c, qin, pin, out); break;
        case 2: launch_last_flux<2>(c, qin, pin, out); break;
        case 3: launch_last_flux<3>(c, qin, pin, out); break;
        case 4: launch_last_flux<4>(c, qin, pin, out); break;
        default: launch_last_flux<5>(c, qin, pin, out); break;
    }
}

// The reference computes the energy after every step of a run with PML
// (par%set_pml), also after its trigger has switched the PML elements off
// (src/timeloopMod.f90:482-516): it does not depend on pml_active.
bool energy_on(const nexd_ctx *c) { return c->cfg.energy && c->cfg.pml && c->have_pml; }

template <int ORDER>
void launch_step_end(nexd_ctx *c, int row_offset, int advance, const double *blk_energy) {
    const double *be = energy_on(c) ? blk_energy : nullptr;
    const int chunks = c->cfg.nrec > 0 ? (c->cfg.nrec * 9 + 1023) / 1024 : 1;
    hipLaunchKernelGGL((nexd::step_end_kernel<ORDER, double>), dim3(1, chunks), dim3(1024), 0, c->stream, c->q[0],
                       c->u_slot, c->a_slot, c->rec_int, c->recelem, c->rec_slot_of, c->cfg.nrec, c->nslot,
                       (int)c->nglob, c->cfg.nt, row_offset, advance, c->seis, be, c->nblk_total, c->energy, c->d_it,
                       c->d_done);
}

void dispatch_step_end(nexd_ctx *c, int row_offset, int advance, const double *blk_energy) {
    switch (c->cfg.order) {
        case 1: launch_step_end<1>(c, row_offset, advance, blk_energy); break;
        case 2: launch_step_end<2>(c, row_offset, advance, blk_energy); break;
        case 3: launch_step_end<3>(c, row_offset, advance, blk_energy); break;
        case 4: launch_step_end<4>(c, row_offset, advance, blk_energy); break;
        default: launch_step_end<5>(c, row_offset, advance, blk_energy); break;
    }
}
void dispatch_step_end(nexd_ctx *c, int row_offset, int advance) { dispatch_step_end(c, row_offset, advance, c->blk_energy); }

StageArgs<double> base_args(nexd_ctx *c) {
    const bool f32w = (c->cfg.compat & NEXD_COMPAT_RK3_FLOAT_WEIGHTS) != 0;
    StageArgs<double> a{};
    a.nglob = (int)c->nglob;
    a.nshot = 1;
    a.mat_stride = 0;
    a.anel_stride = 0;
    a.pmlc_stride_ng = 0;
    a.src_factored = 0;
    a.nt = c->cfg.nt;
    a.nslot = c->nslot;
    a.scheme = c->cfg.timeint;
    a.dt = c->cfg.dt;
    a.dt_elem = nullptr;
    a.w13 = f32w ? (double)(1.0f / 3.0f) : 1.0 / 3.0;
    a.w23 = f32w ? (double)(2.0f / 3.0f) : 2.0 / 3.0;
    a.pml_compat = (c->cfg.compat & NEXD_COMPAT_PML_LAST_ELEMENT) ? 1 : 0;
    a.last_flux = c->last_flux;
    a.D = c->D; a.DK = nullptr; a.LT = c->LT; a.fmask = c->fmask;   // FP64: no contravariant form
    a.metric = c->metric; a.fscale = c->fscale; a.frame = c->frame; a.matc = c->matc;
    a.nbr = c->nbr; a.ftype = c->ftype; a.anel = c->anel;
    for (int i = 0; i < nexd::NMB; ++i) a.wl[i] = c->wl[i];
    a.pmlc = c->pmlc;
    a.src_head = c->src_head; a.src_next = c->src_next; a.srctype = c->srctype;
    a.src_array = c->src_array; a.src_array_m = c->src_array_m; a.nsrc = c->cfg.nsrc;
    a.d_it = c->d_it;
    a.rec_slot = c->rec_slot; a.u_slot = c->u_slot; a.a_slot = c->a_slot; a.u_full = c->u_full;
    a.mass = c->mass; a.jac = c->jac; a.emat = c->emat; a.emat_stride = 0;
    a.blk_stride = c->nblk_total;
    a.pml_on = nullptr;
    return a;
}

// enqueue one full time step on the context stream
// buffer roles of stage irk: RK2 ping-pongs Q0 -> Q1 -> Q0, RK3 runs
// Q0 -> Q1 -> Q2 -> Q0; the step-start state Q0 is the RK base throughout
void stage_buffers(nexd_ctx *c, int irk, StageArgs<double> &a) {
    const int scheme = c->cfg.timeint;
    double *qin, *qout, *thin, *thout, *pin, *pout;
    if (scheme == NEXD_EULER) {
        qin = c->q[0]; qout = c->q[1]; thin = c->th[0]; thout = c->th[0]; pin = c->paux[0]; pout = c->paux[0];
    } else if (scheme == NEXD_RK2) {
        if (irk == 1) { qin = c->q[0]; qout = c->q[1]; thin = c->th[0]; thout = c->th[1]; pin = c->paux[0]; pout = c->paux[1]; }
        else { qin = c->q[1]; qout = c->q[0]; thin = c->th[1]; thout = c->th[0]; pin = c->paux[1]; pout = c->paux[0]; }
    } else if (scheme == NEXD_RK3) {
        if (irk == 1) { qin = c->q[0]; qout = c->q[1]; thin = c->th[0]; thout = c->th[1]; pin = c->paux[0]; pout = c->paux[1]; }
        else if (irk == 2) { qin = c->q[1]; qout = c->q[2]; thin = c->th[1]; thout = c->th[1]; pin = c->paux[1]; pout = c->paux[1]; }
        else { qin = c->q[2]; qout = c->q[0]; thin = c->th[1]; thout = c->th[0]; pin = c->paux[1]; pout = c->paux[0]; }
    } else {
        // low-storage RK4: q rotates 0 -> 1 -> 2 -> 1 -> 2 -> 0 (the face
        // term reads neighbours of the stage input); memory variables and
        // PML auxiliaries, read only at their own node, update in place
        static const int from[5] = {0, 1, 2, 1, 2}, to[5] = {1, 2, 1, 2, 0};
        qin = c->q[from[irk - 1]]; qout = c->q[to[irk - 1]];
        thin = thout = c->th[0]; pin = pout = c->paux[0];
    }
    a.irk = irk;
    a.last_stage = irk == c->nrk ? 1 : 0;
    const bool ls = scheme == NEXD_LSRK4;
    a.qin = qin; a.qbase = ls ? c->qres : c->q[0]; a.qout = qout;
    a.thin = thin; a.thbase = ls ? c->thres : c->th[0]; a.thout = thout;
    a.pin = pin; a.pbase = ls ? c->pres : c->paux[0]; a.pout = pout;
    a.qres = ls ? c->qres : nullptr; a.thres = ls ? c->thres : nullptr; a.pres = ls ? c->pres : nullptr;
    a.lsa = ls ? nexd::kLsrkA[irk - 1] : 0.0;
    a.lsb = ls ? nexd::kLsrkB[irk - 1] : 0.0;
}

// one full time step on the context stream.  The interior and PML lists are
// launched one after the other on the one stream: each fills the device, and
// a captured fork onto a second stream keeps the buffers its kernels use
// allocated after the graph and the buffers are freed (ROCm 7.2).
void enqueue_step(nexd_ctx *c, const StepScratch &sc) {
    const int scheme = c->cfg.timeint;
    const bool att = c->cfg.attenuation != 0;
    const bool pml = c->have_pml && c->pml_active && c->n_pml > 0;
    StageArgs<double> a = base_args(c);
    a.last_flux = sc.last_flux;
    const int nrk = c->nrk;
    for (int irk = 1; irk <= nrk; ++irk) {
        stage_buffers(c, irk, a);
        const double *qin = a.qin, *pin = a.pin;
        // energy partials: the plain list's blocks first, then the PML list's
        // (every element is in the plain list once the PML is switched off)
        const bool energy = a.last_stage && energy_on(c);
        // elements without PML (every element when PML is off), then the PML
        // list: both read the stage input and write disjoint elements and
        // energy partials
        a.elist = c->elist_plain;
        a.nlist = c->n_plain;
        a.blk_energy = energy ? sc.blk_energy : nullptr;
        dispatch_stage(c->cfg.order, a, att, false, c->stream);
        if (pml) {
            if (a.pml_compat) dispatch_last_flux(c, qin, pin, sc.last_flux);
            a.elist = c->elist_pml;
            a.nlist = c->n_pml;
            a.blk_energy = energy ? sc.blk_energy + nblocks(c->cfg.order, c->n_plain, att, false) : nullptr;
            dispatch_stage(c->cfg.order, a, att, true, c->stream);
        }
        if (scheme == NEXD_EULER)
            (void)hipMemcpyAsync(c->q[0], c->q[1], sizeof(double) * 9 * c->nglob, hipMemcpyDeviceToDevice, c->stream);
    }
    dispatch_step_end(c, 1, 1, sc.blk_energy);
}

int rebuild_lists(nexd_ctx *c) {
    // elements without PML first (all elements when PML is inactive), PML ones second
    std::vector<int> plain, pmll;
    const bool pml = c->have_pml && c->pml_active;
    for (int e = 0; e < c->cfg.nelem; ++e) {
        if (pml && c->pmlcheck[e]) pmll.push_back(e);
        else if (pml) plain.push_back(e);
    }
    if (!pml) {
        plain.resize(c->cfg.nelem);
        for (int e = 0; e < c->cfg.nelem; ++e) plain[e] = e;
    }
    c->n_plain = (int)plain.size();
    c->n_pml = (int)pmll.size();
    if (c->n_plain) {
        HIPCHK(c, hipMemcpy(c->elist_plain, plain.data(), sizeof(int) * plain.size(), hipMemcpyHostToDevice));
    }
    if (c->n_pml) {
        HIPCHK(c, hipMemcpy(c->elist_pml, pmll.data(), sizeof(int) * pmll.size(), hipMemcpyHostToDevice));
    }
    const bool att = c->cfg.attenuation != 0;
    c->nblk_total = nblocks(c->cfg.order, c->n_plain, att, false) + nblocks(c->cfg.order, c->n_pml, att, true);
    drop_graph(c);
    return 0;
}

// ---------------------------------------------------------------------------
// Scratch of nexd_run_noalloc: one static array per GPU, resolved once per
// device to its device address (hipGetSymbolAddress) at context creation.
// Every nexd_run_noalloc on a GPU shares that GPU's array.
// ---------------------------------------------------------------------------
constexpr size_t kScratchReals = (size_t)1 << 20;   // 8 MiB of doubles per GPU
constexpr int kMaxDevices = 64;
std::mutex g_scratch_mu;
double *g_scratch[kMaxDevices] = {};

}  // namespace

__device__ __align__(256) static double scratch[kScratchReals];

namespace {

// the device address of `scratch` on the current device (called with that
// device current); 0 on success
int resolve_scratch(nexd_ctx *c) {
    const int dev = c->cfg.device;
    if (dev >= kMaxDevices) return fail(c, NEXD_E_UNSUPPORTED, "device ordinal beyond the scratch table");
    std::lock_guard<std::mutex> lock(g_scratch_mu);
    if (g_scratch[dev]) return 0;
    void *p = nullptr;
    const hipError_t e = hipGetSymbolAddress(&p, HIP_SYMBOL(scratch));
    if (e != hipSuccess || !p)
        return fail(c, NEXD_E_HIP, std::string("hipGetSymbolAddress(scratch): ") + hipGetErrorString(e));
    g_scratch[dev] = static_cast<double *>(p);
    return 0;
}

}  // namespace

// ===========================================================================
// C-ABI
// ===========================================================================
extern "C" {

int nexd_create(const nexd_config *cfg, nexd_ctx **out) {
    if (!cfg || !out) return fail(nullptr, NEXD_E_ARG, "null argument");
    *out = nullptr;
    if (cfg->order < 1 || cfg->order > 5) return fail(nullptr, NEXD_E_UNSUPPORTED, "order must be 1..5");
    if (cfg->nelem <= 0 || cfg->nt <= 0 || cfg->nsrc < 0 || cfg->nrec < 0)
        return fail(nullptr, NEXD_E_ARG, "nelem, nt must be positive; nsrc, nrec non-negative");
    if (cfg->timeint < 1 || cfg->timeint > 4)
        return fail(nullptr, NEXD_E_UNSUPPORTED, "timeint must be 1, 2, 3 or 4 (low-storage RK4)");
    if (!(cfg->dt > 0)) return fail(nullptr, NEXD_E_ARG, "dt must be positive");
    int ndev = 0;
    if (hipGetDeviceCount(&ndev) != hipSuccess || ndev == 0) return fail(nullptr, NEXD_E_HIP, "no HIP device");
    if (cfg->device < 0 || cfg->device >= ndev) return fail(nullptr, NEXD_E_ARG, "device ordinal out of range");
    nexd_ctx *c = new nexd_ctx();
    c->cfg = *cfg;
    c->np = (cfg->order + 1) * (cfg->order + 2) * (cfg->order + 3) / 6;
    c->npf = (cfg->order + 1) * (cfg->order + 2) / 2;
    c->nfp = 4 * c->npf;
    c->nglob = (size_t)cfg->nelem * c->np;
    // the kernels address a field plane (nglob values) with 32-bit byte offsets
    const size_t per_elem = std::max<size_t>(std::max<size_t>(c->np, c->nfp), 36);   // field plane, fscale, frames
    if ((size_t)cfg->nelem * per_elem * sizeof(double) > 0xFFFFFFFFull) {
        g_create_error = "mesh too large: a field plane (nelem*Np) or per-element face array (nelem*max(4NpF, 36)) "
                         "of doubles must stay below 4 GiB";
        delete c;
        return NEXD_E_UNSUPPORTED;
    }
    c->nrk = cfg->timeint == 1 ? 1 : (cfg->timeint == 2 ? 2 : (cfg->timeint == 3 ? 3 : 5));
    auto bail = [&](int rc) { g_create_error = c->err; nexd_destroy(c); return rc; };
    if (hipSetDevice(cfg->device) != hipSuccess) return bail(fail(c, NEXD_E_HIP, "hipSetDevice failed"));
    if (hipStreamCreate(&c->stream) != hipSuccess) return bail(fail(c, NEXD_E_HIP, "hipStreamCreate failed"));
    if (hipEventCreate(&c->ev0) != hipSuccess || hipEventCreate(&c->ev1) != hipSuccess)
        return bail(fail(c, NEXD_E_HIP, "hipEventCreate failed"));
    {
        const int rs = resolve_scratch(c);
        if (rs) return bail(rs);
    }
    const size_t ng = c->nglob;
    int rc = 0;
    for (int i = 0; i < 3 && !rc; ++i) rc = dalloc(c, &c->q[i], 9 * ng);
    if (!rc && cfg->attenuation) (rc = dalloc(c, &c->th[0], 18 * ng)) || (rc = dalloc(c, &c->th[1], 18 * ng));
    if (!rc && cfg->pml)
        (rc = dalloc(c, &c->paux[0], 27 * ng)) || (rc = dalloc(c, &c->paux[1], 27 * ng)) ||
            (rc = dalloc(c, &c->pmlc, 9 * ng)) || (rc = dalloc(c, &c->last_flux, 27 * c->np));
    if (!rc && cfg->track_displacement) rc = dalloc(c, &c->u_full, 3 * ng);
    if (!rc && cfg->timeint == NEXD_LSRK4) {
        rc = dalloc(c, &c->qres, 9 * ng);
        if (!rc && cfg->attenuation) rc = dalloc(c, &c->thres, 18 * ng);
        if (!rc && cfg->pml) rc = dalloc(c, &c->pres, 27 * ng);
    }
    if (!rc)
        (rc = dalloc(c, &c->elist_plain, cfg->nelem)) || (rc = dalloc(c, &c->elist_pml, cfg->nelem)) ||
            (rc = dalloc(c, &c->d_it, 1)) || (rc = dalloc(c, &c->d_done, 1)) ||
            (rc = dalloc(c, &c->seis, (size_t)9 * cfg->nt * (cfg->nrec > 0 ? cfg->nrec : 1))) ||
            (rc = dalloc(c, &c->energy, cfg->nt)) || (rc = dalloc(c, &c->rec_slot, cfg->nelem)) ||
            (rc = dalloc(c, &c->src_head, cfg->nelem)) || (rc = dalloc(c, &c->blk_energy, (size_t)cfg->nelem + 64));
    if (rc) return bail(rc);
    c->pmlcheck.assign(cfg->nelem, 0);
    c->pml_active = cfg->pml != 0;
    // defaults: no sources, no receivers
    std::vector<int> minus(cfg->nelem, -1);
    if ((rc = upload(c, &c->src_head, minus)) || (rc = upload(c, &c->rec_slot, minus))) return bail(rc);
    // the step-end launch's block counter starts at 0 (each launch leaves it at 0)
    const unsigned zero_done = 0;
    if (hipMemcpy(c->d_done, &zero_done, sizeof(unsigned), hipMemcpyHostToDevice) != hipSuccess)
        return bail(fail(c, NEXD_E_HIP, "hipMemcpy failed"));
    if (!cfg->attenuation) c->have_att = true;
    if (!cfg->pml) c->have_pml = false;
    if (cfg->nsrc == 0) c->have_src = true;
    if (cfg->nrec == 0) c->have_rec = true;
    if ((rc = rebuild_lists(c))) return bail(rc);
    *out = c;
    return NEXD_OK;
}

int nexd_upload_operators(nexd_ctx *c, const double *dr, const double *ds, const double *dt, const double *lift,
                          const double *mass, const int *fmask) {
    if (!c) return fail(c, NEXD_E_ARG, "null context");
    nexd::OpsHost o;
    const std::string err = nexd::convert_operators(c->np, c->npf, dr, ds, dt, lift, mass, fmask, o);
    if (!err.empty()) return fail(c, NEXD_E_ARG, err);
    int rc;
    if ((rc = upload(c, &c->D, nexd::mfma_operator_fragments(o.D, c->np, true))) ||
        (rc = upload(c, &c->LT, nexd::mfma_lift_fragments(o.LT, c->np, 4 * c->npf))) || (rc = upload(c, &c->mass, nexd::mfma_mass_fragments(o.M, c->np))) ||
        (rc = upload(c, &c->fmask, o.fm)))
        return rc;
    c->fmask_host = o.fm;
    c->have_ops = true;
    drop_graph(c);
    return NEXD_OK;
}

int nexd_upload_mesh(nexd_ctx *c, const int *ibool, const double *rx, const double *ry, const double *rz,
                     const double *sx, const double *sy, const double *sz, const double *tx, const double *ty,
                     const double *tz, const double *jacobian, const double *fscale, const int *neighbor,
                     const int *face, const int *ibt, const int *ibn, const double *nnormal,
                     const double *stangential, const double *ttangential) {
    if (!c) return fail(c, NEXD_E_ARG, "null context");
    if (!c->have_ops) return fail(c, NEXD_E_STATE, "upload operators before the mesh");
    const double *met[9] = {rx, ry, rz, sx, sy, sz, tx, ty, tz};
    nexd::MeshHost m;
    const std::string err = nexd::convert_mesh(c->cfg.nelem, c->np, c->npf, c->fmask_host, ibool, met, jacobian,
                                               fscale, neighbor, face, ibt, ibn, nnormal, stangential,
                                               ttangential, m);
    if (!err.empty())
        return err[0] == '!' ? fail(c, NEXD_E_UNSUPPORTED, err.substr(1)) : fail(c, NEXD_E_ARG, err);
    int rc;
    if ((rc = upload(c, &c->metric, m.metric)) || (rc = upload(c, &c->jac, m.jac)) ||
        (rc = upload(c, &c->fscale, m.fscale)) || (rc = upload(c, &c->frame, m.frame)) ||
        (rc = upload(c, &c->nbr, m.nbr)) || (rc = upload(c, &c->ftype, m.ftype)))
        return rc;
    c->have_mesh = true;
    drop_graph(c);
    return NEXD_OK;
}

int nexd_upload_material(nexd_ctx *c, const double *vpu, const double *vsu, const double *rho, const double *lambdau,
                         const double *muu, const double *mu, const double *lambda) {
    if (!c || !vpu || !vsu || !rho || !lambdau || !muu || !mu || !lambda) return fail(c, NEXD_E_ARG, "null material");
    const int nel = c->cfg.nelem;
    std::vector<double> mc((size_t)nel * 8, 0.0), em((size_t)nel * 2);
    for (int e = 0; e < nel; ++e) {
        if (!(rho[e] > 0)) return fail(c, NEXD_E_ARG, "density must be positive");
        mc[e * 8 + 0] = vpu[e]; mc[e * 8 + 1] = vsu[e]; mc[e * 8 + 2] = rho[e];
        mc[e * 8 + 3] = lambdau[e]; mc[e * 8 + 4] = muu[e];
        em[e * 2] = mu[e]; em[e * 2 + 1] = lambda[e];
    }
    int rc;
    if ((rc = upload(c, &c->matc, mc)) || (rc = upload(c, &c->emat, em))) return rc;
    c->have_mat = true;
    drop_graph(c);
    return NEXD_OK;
}

int nexd_upload_anelastic(nexd_ctx *c, const double *wl, const double *anelasticvar) {
    if (!c || !wl || !anelasticvar) return fail(c, NEXD_E_ARG, "null anelastic array");
    if (!c->cfg.attenuation) return fail(c, NEXD_E_STATE, "context created without attenuation");
    for (int i = 0; i < nexd::NMB; ++i) c->wl[i] = wl[i];
    std::vector<double> av(anelasticvar, anelasticvar + (size_t)c->cfg.nelem * nexd::NMB * 3);
    int rc = upload(c, &c->anel, av);
    if (rc) return rc;
    c->have_att = true;
    drop_graph(c);
    return NEXD_OK;
}

int nexd_upload_pml(nexd_ctx *c, const int *pmlcheck, const double *ddx, const double *ddy, const double *ddz,
                    const double *alphax, const double *alphay, const double *alphaz, const double *kx,
                    const double *ky, const double *kz) {
    if (!c || !pmlcheck || !ddx || !ddy || !ddz || !alphax || !alphay || !alphaz || !kx || !ky || !kz)
        return fail(c, NEXD_E_ARG, "null PML array");
    if (!c->cfg.pml) return fail(c, NEXD_E_STATE, "context created without PML");
    std::vector<double> pc;
    nexd::pml_coefficients(c->nglob, ddx, ddy, ddz, alphax, alphay, alphaz, kx, ky, kz, pc);
    for (int e = 0; e < c->cfg.nelem; ++e) c->pmlcheck[e] = pmlcheck[e] ? 1 : 0;
    int rc = upload(c, &c->pmlc, pc);
    if (rc) return rc;
    c->have_pml = true;
    return rebuild_lists(c);
}

int nexd_upload_sources(nexd_ctx *c, const int *srcelem, const int *srctype, const double *src_array,
                        const double *src_array_m) {
    if (!c) return fail(c, NEXD_E_ARG, "null context");
    const int ns = c->cfg.nsrc, np = c->np, nel = c->cfg.nelem;
    if (ns == 0) { c->have_src = true; return NEXD_OK; }
    if (!srcelem || !srctype) return fail(c, NEXD_E_ARG, "null source array");
    std::vector<int> head(nel, -1), next(ns, -1), st(ns);
    bool any_f = false, any_m = false;
    for (int s = ns - 1; s >= 0; --s) {
        const int e = srcelem[s] - 1;
        if (e < 0 || e >= nel) return fail(c, NEXD_E_ARG, "srcelem out of range");
        if (srctype[s] != 0 && srctype[s] != 1) return fail(c, NEXD_E_ARG, "srctype must be 0 or 1");
        st[s] = srctype[s];
        any_f |= srctype[s] == 0;
        any_m |= srctype[s] == 1;
        next[s] = head[e];
        head[e] = s;
    }
    if ((any_f && !src_array) || (any_m && !src_array_m)) return fail(c, NEXD_E_ARG, "missing source array");
    const size_t n3 = (size_t)np * 3 * ns * c->cfg.nt, n6 = (size_t)np * 6 * ns * c->cfg.nt;
    std::vector<double> a3 = src_array ? std::vector<double>(src_array, src_array + n3) : std::vector<double>(1, 0.0);
    std::vector<double> a6 = src_array_m ? std::vector<double>(src_array_m, src_array_m + n6) : std::vector<double>(1, 0.0);
    int rc;
    if ((rc = upload(c, &c->src_head, head)) || (rc = upload(c, &c->src_next, next)) ||
        (rc = upload(c, &c->srctype, st)) || (rc = upload(c, &c->src_array, a3)) ||
        (rc = upload(c, &c->src_array_m, a6)))
        return rc;
    c->have_src = true;
    drop_graph(c);
    return NEXD_OK;
}

int nexd_upload_receivers(nexd_ctx *c, const int *recelem, const double *rec_int) {
    if (!c) return fail(c, NEXD_E_ARG, "null context");
    const int nr = c->cfg.nrec, np = c->np, nel = c->cfg.nelem;
    if (nr == 0) { c->have_rec = true; return NEXD_OK; }
    if (!recelem || !rec_int) return fail(c, NEXD_E_ARG, "null receiver array");
    std::vector<int> slot_of_elem(nel, -1), slot_of_rec(nr), re(nr);
    int nslot = 0;
    for (int r = 0; r < nr; ++r) {
        const int e = recelem[r] - 1;
        if (e < 0 || e >= nel) return fail(c, NEXD_E_ARG, "recelem out of range");
        if (slot_of_elem[e] < 0) slot_of_elem[e] = nslot++;
        slot_of_rec[r] = slot_of_elem[e];
        re[r] = e;
    }
    std::vector<double> w(rec_int, rec_int + (size_t)np * nr);
    int rc;
    // the receiver-element slots carry state (displacement) from step to
    // step: they can change only before the first step
    if (c->step > 0) return fail(c, NEXD_E_STATE, "receivers cannot change after steps have run (reset first)");
    if ((rc = ensure(c, &c->u_slot, (size_t)nslot * 3 * np)) || (rc = ensure(c, &c->a_slot, (size_t)nslot * 3 * np)))
        return rc;
    c->initialized = false;   // the slots get the initial state at the next run
    c->nslot = nslot;
    if ((rc = upload(c, &c->rec_slot, slot_of_elem)) || (rc = upload(c, &c->rec_slot_of, slot_of_rec)) ||
        (rc = upload(c, &c->recelem, re)) || (rc = upload(c, &c->rec_int, w)))
        return rc;
    c->have_rec = true;
    drop_graph(c);
    return NEXD_OK;
}

int nexd_reset(nexd_ctx *c) {
    if (!c) return fail(c, NEXD_E_ARG, "null context");
    const size_t ng = c->nglob;
    // the reference assigns the default-REAL literal 1e-24 (src/timeloopMod.f90:184-195)
    const double tiny_v = (double)1e-24f;
    std::vector<double> tiny(9 * ng, tiny_v);
    for (int i = 0; i < 3; ++i) HIPCHK(c, hipMemcpy(c->q[i], tiny.data(), sizeof(double) * 9 * ng, hipMemcpyHostToDevice));
    if (c->cfg.attenuation) {
        std::vector<double> t18(18 * ng, tiny_v);
        for (int i = 0; i < 2; ++i) HIPCHK(c, hipMemcpy(c->th[i], t18.data(), sizeof(double) * 18 * ng, hipMemcpyHostToDevice));
    }
    if (c->cfg.pml) {
        std::vector<double> z(27 * ng, 0.0);
        for (int i = 0; i < 2; ++i) HIPCHK(c, hipMemcpy(c->paux[i], z.data(), sizeof(double) * 27 * ng, hipMemcpyHostToDevice));
    }
    if (c->u_full) HIPCHK(c, hipMemcpy(c->u_full, tiny.data(), sizeof(double) * 3 * ng, hipMemcpyHostToDevice));
    if (c->qres) {
        const std::vector<double> z(27 * ng, 0.0);
        HIPCHK(c, hipMemcpy(c->qres, z.data(), sizeof(double) * 9 * ng, hipMemcpyHostToDevice));
        if (c->thres) HIPCHK(c, hipMemcpy(c->thres, z.data(), sizeof(double) * 18 * ng, hipMemcpyHostToDevice));
        if (c->pres) HIPCHK(c, hipMemcpy(c->pres, z.data(), sizeof(double) * 27 * ng, hipMemcpyHostToDevice));
    }
    if (c->nslot) {
        std::vector<double> t(3 * (size_t)c->nslot * c->np, tiny_v);
        HIPCHK(c, hipMemcpy(c->u_slot, t.data(), sizeof(double) * t.size(), hipMemcpyHostToDevice));
        HIPCHK(c, hipMemcpy(c->a_slot, t.data(), sizeof(double) * t.size(), hipMemcpyHostToDevice));
    }
    std::vector<double> zs((size_t)9 * c->cfg.nt * (c->cfg.nrec > 0 ? c->cfg.nrec : 1), 0.0);
    HIPCHK(c, hipMemcpy(c->seis, zs.data(), sizeof(double) * zs.size(), hipMemcpyHostToDevice));
    std::vector<double> ze(c->cfg.nt, 0.0);
    HIPCHK(c, hipMemcpy(c->energy, ze.data(), sizeof(double) * ze.size(), hipMemcpyHostToDevice));
    int zero = 0;
    HIPCHK(c, hipMemcpy(c->d_it, &zero, sizeof(int), hipMemcpyHostToDevice));
    HIPCHK(c, hipMemcpy(c->d_done, &zero, sizeof(unsigned), hipMemcpyHostToDevice));
    c->step = 0;
    c->initialized = true;
    return NEXD_OK;
}

int nexd_run(nexd_ctx *c, int it_begin, int it_end, double *energy) {
    if (!c) return fail(c, NEXD_E_ARG, "null context");
    if (!c->have_ops || !c->have_mesh || !c->have_mat || !c->have_att || (c->cfg.pml && !c->have_pml) ||
        !c->have_src || !c->have_rec)
        return fail(c, NEXD_E_STATE, "upload operators, mesh, material, anelastic/PML, sources and receivers first");
    if (!c->initialized) {
        int rc = nexd_reset(c);
        if (rc) return rc;
    }
    if (it_begin != c->step + 1) return fail(c, NEXD_E_ARG, "it_begin must be the next step");
    if (it_end < it_begin || it_end > c->cfg.nt) return fail(c, NEXD_E_ARG, "it_end out of range");
    HIPCHK(c, hipSetDevice(c->cfg.device));
    if (c->step == 0) dispatch_step_end(c, 0, 0);   // row 1: the initial state
    if (!c->graph) {
        hipGraph_t g = nullptr;
        HIPCHK(c, hipStreamBeginCapture(c->stream, hipStreamCaptureModeThreadLocal));
        enqueue_step(c, StepScratch{c->blk_energy, c->last_flux});
        HIPCHK(c, hipStreamEndCapture(c->stream, &g));
        hipError_t e = hipGraphInstantiate(&c->graph, g, nullptr, nullptr, 0);
        (void)hipGraphDestroy(g);
        if (e != hipSuccess) { c->graph = nullptr; return fail(c, NEXD_E_HIP, std::string("hipGraphInstantiate: ") + hipGetErrorString(e)); }
    }
    HIPCHK(c, hipEventRecord(c->ev0, c->stream));
    for (int it = it_begin; it <= it_end; ++it) HIPCHK(c, hipGraphLaunch(c->graph, c->stream));
    HIPCHK(c, hipEventRecord(c->ev1, c->stream));
    HIPCHK(c, hipStreamSynchronize(c->stream));
    HIPCHK(c, hipGetLastError());
    HIPCHK(c, hipEventElapsedTime(&c->last_ms, c->ev0, c->ev1));
    if (energy && energy_on(c))
        HIPCHK(c, hipMemcpy(energy, c->energy + (it_begin - 1), sizeof(double) * (it_end - it_begin + 1), hipMemcpyDeviceToHost));
    c->step = it_end;
    return NEXD_OK;
}

int nexd_run_noalloc(nexd_ctx *c, int it_begin, int it_end, double *energy) {
    if (!c) return fail(c, NEXD_E_ARG, "null context");
    if (!c->have_ops || !c->have_mesh || !c->have_mat || !c->have_att || (c->cfg.pml && !c->have_pml) ||
        !c->have_src || !c->have_rec)
        return fail(c, NEXD_E_STATE, "upload operators, mesh, material, anelastic/PML, sources and receivers first");
    if (!c->initialized) return fail(c, NEXD_E_STATE, "call nexd_reset (or nexd_set_state) before nexd_run_noalloc");
    if (it_begin != c->step + 1) return fail(c, NEXD_E_ARG, "it_begin must be the next step");
    if (it_end < it_begin || it_end > c->cfg.nt) return fail(c, NEXD_E_ARG, "it_end out of range");
    // scratch: the energy block partials, then the last element's fluxes
    const size_t need = (size_t)c->nblk_total + 27 * (size_t)c->np;
    if (need > kScratchReals)
        return fail(c, NEXD_E_UNSUPPORTED, "nexd_run_noalloc: the mesh needs more scratch than the static 2^20 values "
                                           "(use nexd_run)");
    double *base = g_scratch[c->cfg.device];
    if (!base) return fail(c, NEXD_E_STATE, "scratch not resolved for this device");
    const StepScratch sc{base, base + c->nblk_total};
    HIPCHK(c, hipSetDevice(c->cfg.device));
    if (c->step == 0) dispatch_step_end(c, 0, 0, sc.blk_energy);   // row 1: the initial state
    HIPCHK(c, hipEventRecord(c->ev0, c->stream));
    for (int it = it_begin; it <= it_end; ++it) enqueue_step(c, sc);
    HIPCHK(c, hipEventRecord(c->ev1, c->stream));
    HIPCHK(c, hipStreamSynchronize(c->stream));
    HIPCHK(c, hipGetLastError());
    HIPCHK(c, hipEventElapsedTime(&c->last_ms, c->ev0, c->ev1));
    if (energy && energy_on(c))
        HIPCHK(c, hipMemcpy(energy, c->energy + (it_begin - 1), sizeof(double) * (it_end - it_begin + 1), hipMemcpyDeviceToHost));
    c->step = it_end;
    return NEXD_OK;
}

int nexd_set_pml_active(nexd_ctx *c, int active) {
    if (!c) return fail(c, NEXD_E_ARG, "null context");
    if (!c->cfg.pml && active) return fail(c, NEXD_E_STATE, "context created without PML");
    c->pml_active = active != 0;
    return rebuild_lists(c);
}

int nexd_get_seismograms(nexd_ctx *c, double *ux, double *uy, double *uz, double *vx, double *vy, double *vz,
                         double *ax, double *ay, double *az) {
    if (!c) return fail(c, NEXD_E_ARG, "null context");
    double *outs[9] = {ux, uy, uz, vx, vy, vz, ax, ay, az};
    const size_t n = (size_t)c->cfg.nt * c->cfg.nrec;
    for (int k = 0; k < 9; ++k) {
        if (!outs[k]) continue;
        HIPCHK(c, hipMemcpy(outs[k], c->seis + k * n, sizeof(double) * n, hipMemcpyDeviceToHost));
    }
    return NEXD_OK;
}

int nexd_get_state(nexd_ctx *c, double *q, double *ux, double *uy, double *uz) {
    if (!c) return fail(c, NEXD_E_ARG, "null context");
    const size_t ng = c->nglob;
    if (q) HIPCHK(c, hipMemcpy(q, c->q[0], sizeof(double) * 9 * ng, hipMemcpyDeviceToHost));
    double *u[3] = {ux, uy, uz};
    for (int k = 0; k < 3; ++k) {
        if (!u[k]) continue;
        if (!c->u_full) return fail(c, NEXD_E_STATE, "context does not track displacement");
        HIPCHK(c, hipMemcpy(u[k], c->u_full + k * ng, sizeof(double) * ng, hipMemcpyDeviceToHost));
    }
    return NEXD_OK;
}

int nexd_set_state(nexd_ctx *c, int it, const double *q, const double *theta, const double *fprime,
                   const double *gprime, const double *hprime, const double *ux, const double *uy,
                   const double *uz) {
    if (!c) return fail(c, NEXD_E_ARG, "null context");
    if (!c->have_ops || !c->have_mesh || !c->have_rec || !c->have_src)
        return fail(c, NEXD_E_STATE, "upload operators, mesh, sources and receivers first");
    if (it < 0 || it >= c->cfg.nt) return fail(c, NEXD_E_ARG, "it out of range");
    if (theta && !c->cfg.attenuation) return fail(c, NEXD_E_ARG, "theta given without attenuation");
    if ((fprime || gprime || hprime) && !c->cfg.pml) return fail(c, NEXD_E_ARG, "PML fields given without PML");
    if (!c->initialized) {
        int rc = nexd_reset(c);
        if (rc) return rc;
    }
    HIPCHK(c, hipSetDevice(c->cfg.device));
    const size_t ng = c->nglob;
    // device buffer 0 holds the step-start state in the reference's own
    // column-major layouts: q(nglob,9), theta(nglob,6,nMB), f'(nglob,9)...
    if (q) HIPCHK(c, hipMemcpy(c->q[0], q, sizeof(double) * 9 * ng, hipMemcpyHostToDevice));
    if (theta) HIPCHK(c, hipMemcpy(c->th[0], theta, sizeof(double) * 18 * ng, hipMemcpyHostToDevice));
    const double *pf[3] = {fprime, gprime, hprime};
    for (int k = 0; k < 3; ++k)
        if (pf[k]) HIPCHK(c, hipMemcpy(c->paux[0] + (size_t)k * 9 * ng, pf[k], sizeof(double) * 9 * ng, hipMemcpyHostToDevice));
    const double *u[3] = {ux, uy, uz};
    for (int k = 0; k < 3; ++k) {
        if (!u[k]) continue;
        if (c->u_full) HIPCHK(c, hipMemcpy(c->u_full + k * ng, u[k], sizeof(double) * ng, hipMemcpyHostToDevice));
    }
    if (c->nslot && (ux || uy || uz)) {
        // receiver-element displacement slots from the given fields
        std::vector<int> slot_of_elem(c->cfg.nelem, -1);
        HIPCHK(c, hipMemcpy(slot_of_elem.data(), c->rec_slot, sizeof(int) * c->cfg.nelem, hipMemcpyDeviceToHost));
        std::vector<double> us((size_t)3 * c->nslot * c->np);
        HIPCHK(c, hipMemcpy(us.data(), c->u_slot, sizeof(double) * us.size(), hipMemcpyDeviceToHost));
        for (int e = 0; e < c->cfg.nelem; ++e) {
            const int sl = slot_of_elem[e];
            if (sl < 0) continue;
            for (int k = 0; k < 3; ++k)
                if (u[k])
                    for (int j = 0; j < c->np; ++j) us[((size_t)sl * 3 + k) * c->np + j] = u[k][(size_t)e * c->np + j];
        }
        HIPCHK(c, hipMemcpy(c->u_slot, us.data(), sizeof(double) * us.size(), hipMemcpyHostToDevice));
    }
    HIPCHK(c, hipMemcpy(c->d_it, &it, sizeof(int), hipMemcpyHostToDevice));
    c->step = it;
    // receiver row it+1 is the state before step it+1: the state just set
    dispatch_step_end(c, 0, 0);
    HIPCHK(c, hipStreamSynchronize(c->stream));
    HIPCHK(c, hipGetLastError());
    return NEXD_OK;
}

int nexd_last_run_ms(nexd_ctx *c, double *ms) {
    if (!c || !ms) return fail(c, NEXD_E_ARG, "null argument");
    *ms = c->last_ms;
    return NEXD_OK;
}

int nexd_time_stage_kernel(nexd_ctx *c, int list, int nrep, double *avg_ms, int *nelem_launch) {
    if (!c || !avg_ms || nrep <= 0 || (list != 0 && list != 1)) return fail(c, NEXD_E_ARG, "bad argument");
    if (!c->initialized) return fail(c, NEXD_E_STATE, "run or reset before timing");
    const bool pml = c->have_pml && c->pml_active && c->n_pml > 0;
    if (list == 1 && !pml) return fail(c, NEXD_E_STATE, "no PML element list");
    if (c->nrk == 1) return fail(c, NEXD_E_UNSUPPORTED, "timing needs a multi-stage scheme");
    StageArgs<double> a = base_args(c);
    a.blk_energy = nullptr;
    a.elist = list == 0 ? c->elist_plain : c->elist_pml;
    a.nlist = list == 0 ? c->n_plain : c->n_pml;
    HIPCHK(c, hipSetDevice(c->cfg.device));
    // every RK stage in turn, as the time loop runs them (stage 1 first as a
    // warm-up); the receiver-element updates of the last stage are part of
    // it.  The last element's fluxes are computed once, before the timed
    // launches: their values do not change the work.
    const bool att = c->cfg.attenuation != 0;
    // the last stage writes its energy partials, as in the time loop
    double *en = energy_on(c) ? c->blk_energy + (list == 1 ? nblocks(c->cfg.order, c->n_plain, att, false) : 0)
                              : nullptr;
    auto one = [&](int irk) {
        stage_buffers(c, irk, a);
        a.blk_energy = a.last_stage ? en : nullptr;
        dispatch_stage(c->cfg.order, a, att, list == 1, c->stream);
    };
    stage_buffers(c, 1, a);
    if (list == 1 && a.pml_compat) dispatch_last_flux(c, a.qin, a.pin, c->last_flux);
    one(1);
    HIPCHK(c, hipEventRecord(c->ev0, c->stream));
    for (int r = 0; r < nrep; ++r)
        for (int irk = 1; irk <= c->nrk; ++irk) one(irk);
    HIPCHK(c, hipEventRecord(c->ev1, c->stream));
    HIPCHK(c, hipStreamSynchronize(c->stream));
    HIPCHK(c, hipGetLastError());
    float ms = 0.f;
    HIPCHK(c, hipEventElapsedTime(&ms, c->ev0, c->ev1));
    nrep *= c->nrk;
    *avg_ms = ms / nrep;
    if (nelem_launch) *nelem_launch = a.nlist;
    return NEXD_OK;
}

void nexd_destroy(nexd_ctx *c) {
    if (!c) return;
    if (c->stream) (void)hipStreamSynchronize(c->stream);
    drop_graph(c);
    for (auto &a : c->allocs) (void)hipFree(a.first);
    if (c->ev0) (void)hipEventDestroy(c->ev0);
    if (c->ev1) (void)hipEventDestroy(c->ev1);
    if (c->stream) (void)hipStreamDestroy(c->stream);
    delete c;
}

const char *nexd_last_error(const nexd_ctx *c) { return c ? c->err.c_str() : g_create_error.c_str(); }

}  // extern "C"
