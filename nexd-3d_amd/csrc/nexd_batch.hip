// Batched shots: the nexd_batch_* part of include/nexd_gpu.h.
//
// nshot shots on one mesh share operators, geometry and receivers and
// advance together: every stage launch covers all shots, with the shot index
// varying fastest across blocks so that the blocks of one element group run
// back to back and reuse its geometry from L2.  Arithmetic is FP32 or FP64
// (the same kernels as the single-shot path, instantiated for the type).
#include <hip/hip_runtime.h>

#include <algorithm>

#include <cstring>
#include <map>
#include <memory>
#include <tuple>
#include <string>
#include <vector>

#include "../../include/nexd_gpu.h"
#include "host_layout.hpp"
#include "plain_kernels.hpp"
#include "stage_launch.hpp"

namespace {

thread_local std::string g_batch_create_error;

struct BatchBase {
    virtual ~BatchBase() = default;
    virtual int group_prepare(hipGraph_t *out) = 0;
    virtual int next_step() const = 0;
    virtual void set_steps_done(int s) = 0;
    nexd_batch_config cfg{};
    std::string err;
};

template <typename R>
std::vector<R> cvt(const std::vector<double> &v) {
    return std::vector<R>(v.begin(), v.end());
}

template <typename R>
struct Batch : BatchBase {
    int np = 0, npf = 0, nfp = 0, nrk = 0, nslot = 0;
    size_t ng = 0;
    hipStream_t stream = nullptr;
    hipEvent_t ev0 = nullptr, ev1 = nullptr;
    std::vector<void *> allocs;
    R *q[3] = {nullptr, nullptr, nullptr};
    R *qres = nullptr, *thres = nullptr, *pres = nullptr;   // low-storage RK4 residuals
    R *qlo[3] = {nullptr, nullptr, nullptr};   // low parts (compensated state)
    R *u_slot_lo = nullptr;
    R *th[2] = {nullptr, nullptr};
    R *paux[2] = {nullptr, nullptr};
    R *D = nullptr, *DK = nullptr, *LT = nullptr, *mass = nullptr, *metric = nullptr, *fscale = nullptr, *frame = nullptr;
    R *jac = nullptr;
    int *fmask = nullptr, *nbr = nullptr, *ftype = nullptr;
    std::vector<int> fmask_host;
    R *matc = nullptr, *anel = nullptr, *pmlc = nullptr, *emat = nullptr;
    double wl[nexd::NMB] = {0, 0, 0};
    std::vector<int> pmlcheck;
    int *elist_plain = nullptr, *elist_pml = nullptr;
    int n_plain = 0, n_pml = 0;
    int *src_head = nullptr, *src_next = nullptr, *srctype = nullptr;
    R *src_spatial = nullptr, *src_stf = nullptr;
    int *recelem = nullptr, *rec_slot = nullptr, *rec_slot_of = nullptr;
    R *rec_int = nullptr, *u_slot = nullptr, *a_slot = nullptr, *seis = nullptr, *last_flux = nullptr;
    int *d_it = nullptr;
    unsigned *d_done = nullptr;   // blocks of the step-end launch that finished
    int *pml_on = nullptr;                    // [shot] 1 while the shot's PML is on
    std::vector<int> pml_on_host;
    double *blk_energy = nullptr, *energy = nullptr;   // [shot][nblk_total], [shot][nt] (FP64 sums)
    int nblk_total = 0;
    int tile = 0;                             // block shape of the stage launches (autotuner)
    R *dt_elem = nullptr;                     // per-element step, or null (cfg.dt everywhere)
    int step = 0;
    bool have_ops = false, have_mesh = false, initialized = false;
    std::vector<char> have_mat, have_att, have_pml, have_src;
    bool have_rec = false;
    hipGraphExec_t graph = nullptr;
    float last_ms = 0.f;

    ~Batch() override {
        if (stream) (void)hipStreamSynchronize(stream);
        if (graph) (void)hipGraphExecDestroy(graph);
        for (void *p : allocs) (void)hipFree(p);
        if (ev0) (void)hipEventDestroy(ev0);
        if (ev1) (void)hipEventDestroy(ev1);
        if (stream) (void)hipStreamDestroy(stream);
    }

    int fail(int code, const std::string &m) {
        err = m;
        return code;
    }

    template <typename T>
    int dalloc(T **p, size_t n) {
        void *v = nullptr;
        if (n == 0) n = 1;
        hipError_t e = hipMalloc(&v, n * sizeof(T));
        if (e != hipSuccess) return fail(NEXD_E_HIP, std::string("hipMalloc: ") + hipGetErrorString(e));
        allocs.push_back(v);
        *p = static_cast<T *>(v);
        return 0;
    }

    template <typename T>
    int put(T *dst, const std::vector<T> &h) {
        hipError_t e = hipMemcpy(dst, h.data(), h.size() * sizeof(T), hipMemcpyHostToDevice);
        return e == hipSuccess ? 0 : fail(NEXD_E_HIP, std::string("hipMemcpy: ") + hipGetErrorString(e));
    }

    template <typename T>
    int upload(T **p, const std::vector<T> &h) {
        if (!*p) {
            int rc = dalloc(p, h.size());
            if (rc) return rc;
        }
        return put(*p, h);
    }

    void drop_graph() {
        if (graph) {
            (void)hipGraphExecDestroy(graph);
            graph = nullptr;
        }
    }

    int nmat_sets() const { return cfg.shared_material ? 1 : cfg.nshot; }

    int init() {
        const nexd_batch_config &c = cfg;
        np = (c.order + 1) * (c.order + 2) * (c.order + 3) / 6;
        npf = (c.order + 1) * (c.order + 2) / 2;
        nfp = 4 * npf;
        ng = (size_t)c.nelem * np;
        // the kernels address a field plane (ng values) with 32-bit byte offsets
        const size_t per_elem = std::max<size_t>(std::max<size_t>((size_t)np, (size_t)nfp), 36);
        if ((size_t)c.nelem * per_elem * sizeof(R) > 0xFFFFFFFFull)
            return fail(NEXD_E_UNSUPPORTED, "mesh too large: a field plane (nelem*Np) or per-element face array "
                                            "(nelem*max(4NpF, 36)) must stay below 4 GiB");
        nrk = c.timeint == 1 ? 1 : (c.timeint == 2 ? 2 : (c.timeint == 3 ? 3 : 5));
        if (hipSetDevice(c.device) != hipSuccess) return fail(NEXD_E_HIP, "hipSetDevice failed");
        if (hipStreamCreate(&stream) != hipSuccess) return fail(NEXD_E_HIP, "hipStreamCreate failed");
        if (hipEventCreate(&ev0) != hipSuccess || hipEventCreate(&ev1) != hipSuccess)
            return fail(NEXD_E_HIP, "hipEventCreate failed");
        const size_t S = (size_t)c.nshot;
        int rc = 0;
        for (int i = 0; i < 3 && !rc; ++i) rc = dalloc(&q[i], S * 9 * ng);
        for (int i = 0; i < 3 && !rc && c.compensated; ++i) rc = dalloc(&qlo[i], S * 9 * ng);
        if (!rc && c.attenuation) (rc = dalloc(&th[0], S * 18 * ng)) || (rc = dalloc(&th[1], S * 18 * ng));
        if (!rc && c.timeint == NEXD_LSRK4) {
            rc = dalloc(&qres, S * 9 * ng);
            if (!rc && c.attenuation) rc = dalloc(&thres, S * 18 * ng);
            if (!rc && c.pml) rc = dalloc(&pres, S * 27 * ng);
        }
        if (!rc && c.pml)
            (rc = dalloc(&paux[0], S * 27 * ng)) || (rc = dalloc(&paux[1], S * 27 * ng)) ||
                (rc = dalloc(&pmlc, (size_t)nmat_sets() * 9 * ng)) || (rc = dalloc(&last_flux, S * 27 * np));
        if (!rc && c.pml)
            (rc = dalloc(&pml_on, S)) || (rc = dalloc(&blk_energy, S * ((size_t)c.nelem + 64))) ||
                (rc = dalloc(&energy, S * c.nt));
        if (!rc)
            (rc = dalloc(&matc, (size_t)nmat_sets() * c.nelem * 8)) || (rc = dalloc(&emat, (size_t)nmat_sets() * c.nelem * 2)) ||
                (rc = dalloc(&anel, (size_t)nmat_sets() * c.nelem * nexd::NMB * 3)) ||
                (rc = dalloc(&elist_plain, c.nelem)) || (rc = dalloc(&elist_pml, c.nelem)) || (rc = dalloc(&d_it, 1)) || (rc = dalloc(&d_done, 1)) ||
                (rc = dalloc(&seis, S * 9 * c.nt * (c.nrec > 0 ? c.nrec : 1))) ||
                (rc = dalloc(&rec_slot, c.nelem)) || (rc = dalloc(&src_head, S * c.nelem)) ||
                (rc = dalloc(&src_next, S * (c.nsrc > 0 ? c.nsrc : 1))) ||
                (rc = dalloc(&srctype, S * (c.nsrc > 0 ? c.nsrc : 1))) ||
                (rc = dalloc(&src_spatial, S * (c.nsrc > 0 ? c.nsrc : 1) * 6 * np)) ||
                (rc = dalloc(&src_stf, S * (c.nsrc > 0 ? c.nsrc : 1) * c.nt));
        if (rc) return rc;
        pmlcheck.assign(c.nelem, 0);
        have_mat.assign(nmat_sets(), 0);
        have_att.assign(nmat_sets(), c.attenuation ? 0 : 1);
        have_pml.assign(nmat_sets(), c.pml ? 0 : 1);
        have_src.assign(c.nshot, c.nsrc ? 0 : 1);
        have_rec = c.nrec == 0;
        std::vector<int> minus(S * c.nelem, -1);
        if ((rc = put(src_head, minus)) || (rc = put(rec_slot, std::vector<int>(c.nelem, -1)))) return rc;
        // the step-end launch's block counter starts at 0 (each launch leaves it at 0)
        if ((rc = put(d_done, std::vector<unsigned>(1, 0u)))) return rc;
        pml_on_host.assign(S, 1);
        if (c.pml && (rc = put(pml_on, pml_on_host))) return rc;
        return rebuild_lists();
    }

    // element blocks of one list launch (per shot)
    int nblocks(int n, bool pml) const {
        return nexd::stage_blocks<R>(cfg.order, n, cfg.attenuation != 0, pml, tile);
    }

    int rebuild_lists() {
        std::vector<int> plain, pmll;
        const bool pml = cfg.pml != 0;
        for (int e = 0; e < cfg.nelem; ++e) (pml && pmlcheck[e] ? pmll : plain).push_back(e);
        n_plain = (int)plain.size();
        n_pml = (int)pmll.size();
        int rc;
        if (n_plain && (rc = put(elist_plain, plain))) return rc;
        if (n_pml && (rc = put(elist_pml, pmll))) return rc;
        nblk_total = nblocks(n_plain, false) + nblocks(n_pml, true);
        drop_graph();
        return 0;
    }

    using TileKey = std::tuple<int, int, int, int, int, int, int, int>;
    static std::map<TileKey, std::pair<int, std::vector<double>>> &tile_cache() {
        static std::map<TileKey, std::pair<int, std::vector<double>>> c;
        return c;
    }

    // the reference computes the energy after every step of a run with PML
    // (src/timeloopMod.f90:482-516)
    bool energy_on() const { return cfg.energy && cfg.pml; }

    int check_shot(int shot) {
        if (cfg.shared_material ? shot != 0 : (shot < 0 || shot >= cfg.nshot))
            return fail(NEXD_E_ARG, "shot out of range (shared materials use shot 0)");
        return 0;
    }

    nexd::StageArgs<R> base_args() {
        const bool f32w = (cfg.compat & NEXD_COMPAT_RK3_FLOAT_WEIGHTS) != 0;
        nexd::StageArgs<R> a{};
        a.nglob = (int)ng;
        a.nshot = cfg.nshot;
        a.mat_stride = cfg.shared_material ? 0 : cfg.nelem * 8;
        a.anel_stride = cfg.shared_material ? 0 : cfg.nelem * nexd::NMB * 3;
        a.pmlc_stride_ng = cfg.shared_material ? 0 : 9;
        a.src_factored = 1;
        a.nt = cfg.nt;
        a.nslot = nslot;
        a.scheme = cfg.timeint;
        a.dt = (R)cfg.dt;
        a.dt_elem = dt_elem;
        a.w13 = f32w ? (R)(1.0f / 3.0f) : (R)(1.0 / 3.0);
        a.w23 = f32w ? (R)(2.0f / 3.0f) : (R)(2.0 / 3.0);
        a.pml_compat = (cfg.compat & NEXD_COMPAT_PML_LAST_ELEMENT) ? 1 : 0;
        a.last_flux = last_flux;
        a.D = D; a.DK = DK; a.LT = LT; a.fmask = fmask;
        a.metric = metric; a.fscale = fscale; a.frame = frame; a.matc = matc;
        a.nbr = nbr; a.ftype = ftype; a.anel = anel;
        for (int i = 0; i < nexd::NMB; ++i) a.wl[i] = (R)wl[i];
        a.pmlc = pmlc;
        a.src_head = src_head; a.src_next = src_next; a.srctype = srctype;
        a.src_spatial = src_spatial; a.src_stf = src_stf; a.nsrc = cfg.nsrc;
        a.d_it = d_it;
        a.rec_slot = rec_slot; a.u_slot = u_slot; a.a_slot = a_slot; a.u_full = nullptr;
        a.blk_energy = nullptr;
        a.blk_stride = nblk_total;
        a.mass = mass; a.jac = jac; a.emat = emat;
        a.emat_stride = cfg.shared_material ? 0 : cfg.nelem * 2;
        a.pml_on = cfg.pml ? pml_on : nullptr;
        return a;
    }

    void dispatch(const nexd::StageArgs<R> &a, bool pml) { dispatch(a, pml, stream); }
    void dispatch(const nexd::StageArgs<R> &a, bool pml, hipStream_t s) {
        const bool att = cfg.attenuation != 0;
        if (a.scheme == NEXD_LSRK4) {
            nexd::launch_stage<R, true, 0>(cfg.order, att, pml, a, s);
        } else if (tile == 1) {
            if constexpr (sizeof(R) == 4) nexd::launch_stage<float, false, 1>(cfg.order, att, pml, a, s);
        } else {
            nexd::launch_stage<R, false, 0>(cfg.order, att, pml, a, s);
        }
    }

    int nvariants() const { return nexd::stage_variants<R>(cfg.timeint == NEXD_LSRK4); }

    int set_tile(int t) {
        if (t < 0 || t >= nvariants()) return fail(NEXD_E_ARG, "no such block shape for this precision and scheme");
        tile = t;
        nblk_total = nblocks(n_plain, false) + nblocks(n_pml, true);
        drop_graph();
        return 0;
    }

    // Times every block shape on this batch's own element lists (one round
    // of all RK stages per list, nrep rounds after a warm-up) and keeps the
    // fastest.  Choices are remembered per (order, attenuation, PML,
    // precision, shots, elements, PML elements) for the process, so
    // batches of the same kind reuse them without timing.  The timing
    // overwrites the state: the batch is reset afterwards.
    int autotune(int nrep, int *choice, double *ms_out) {
        const int nv = nvariants();
        const auto key = std::make_tuple(cfg.order, cfg.attenuation, cfg.pml, cfg.precision, cfg.nshot, cfg.nelem,
                                         n_pml, cfg.timeint * 2 + (cfg.compensated ? 1 : 0));
        std::vector<double> ms(nv, 0.0);
        int best = 0;
        auto &cache = tile_cache();
        auto hit = cache.find(key);
        if (hit != cache.end()) {
            best = hit->second.first;
            ms = hit->second.second;
        } else if (nrk > 1) {
            if (!initialized) {
                int rc = reset();
                if (rc) return rc;
            }
            for (int v = 0; v < nv; ++v) {
                int rc = set_tile(v);
                if (rc) return rc;
                for (int list = 0; list < 2; ++list) {
                    if ((list == 0 && n_plain == 0) || (list == 1 && (!cfg.pml || n_pml == 0))) continue;
                    double t = 0;
                    int n = 0;
                    if ((rc = time_stage(list, nrep, &t, &n))) return rc;
                    ms[v] += t;
                }
                if (ms[v] < ms[best]) best = v;
            }
            cache[key] = std::make_pair(best, ms);
        }
        int rc = set_tile(best);
        if (rc) return rc;
        if ((rc = reset())) return rc;
        if (choice) *choice = best;
        if (ms_out)
            for (int v = 0; v < nv; ++v) ms_out[v] = ms[v];
        return 0;
    }

    template <int ORDER>
    void last_flux_impl(const R *qin, const R *pin) {
        const int last = cfg.nelem - 1;
        hipLaunchKernelGGL((nexd::last_flux_kernel<ORDER, R>), dim3(cfg.nshot), dim3(64), 0, stream, qin, pin, pmlc,
                           matc, last, pmlcheck[last] ? 1 : 0, (int)ng, cfg.shared_material ? 0 : cfg.nelem * 8,
                           cfg.shared_material ? 0 : 9, last_flux);
    }

    void last_flux_dispatch(const R *qin, const R *pin) {
        switch (cfg.order) {
            case 1: last_flux_impl<1>(qin, pin); break;
            case 2: last_flux_impl<2>(qin, pin); break;
            case 3: last_flux_impl<3>(qin, pin); break;
            case 4: last_flux_impl<4>(qin, pin); break;
            default: last_flux_impl<5>(qin, pin); break;
        }
    }

    template <int ORDER>
    void step_end_impl(int row_offset, int advance) {
        // one block per shot and 1024 of its receiver values
        const int chunks = cfg.nrec > 0 ? (cfg.nrec * 9 + 1023) / 1024 : 1;
        hipLaunchKernelGGL((nexd::step_end_kernel<ORDER, R>), dim3(cfg.nshot, chunks), dim3(1024), 0, stream, q[0],
                           u_slot, a_slot, rec_int, recelem, rec_slot_of, cfg.nrec, nslot, (int)ng, cfg.nt, row_offset,
                           advance, seis, energy_on() ? (const double *)blk_energy : (const double *)nullptr,
                           nblk_total, energy, d_it, d_done);
    }

    void step_end(int row_offset, int advance) {
        switch (cfg.order) {
            case 1: step_end_impl<1>(row_offset, advance); break;
            case 2: step_end_impl<2>(row_offset, advance); break;
            case 3: step_end_impl<3>(row_offset, advance); break;
            case 4: step_end_impl<4>(row_offset, advance); break;
            default: step_end_impl<5>(row_offset, advance); break;
        }
    }

    void set_stage_buffers(nexd::StageArgs<R> &a, int irk) {
        const int scheme = cfg.timeint;
        R *qin, *qout, *thin, *thout, *pin, *pout;
        if (scheme == NEXD_EULER) {
            qin = q[0]; qout = q[1]; thin = th[0]; thout = th[0]; pin = paux[0]; pout = paux[0];
        } else if (scheme == NEXD_RK2) {
            if (irk == 1) { qin = q[0]; qout = q[1]; thin = th[0]; thout = th[1]; pin = paux[0]; pout = paux[1]; }
            else { qin = q[1]; qout = q[0]; thin = th[1]; thout = th[0]; pin = paux[1]; pout = paux[0]; }
        } else if (scheme == NEXD_RK3) {
            if (irk == 1) { qin = q[0]; qout = q[1]; thin = th[0]; thout = th[1]; pin = paux[0]; pout = paux[1]; }
            else if (irk == 2) { qin = q[1]; qout = q[2]; thin = th[1]; thout = th[1]; pin = paux[1]; pout = paux[1]; }
            else { qin = q[2]; qout = q[0]; thin = th[1]; thout = th[0]; pin = paux[1]; pout = paux[0]; }
        } else {
            // low-storage RK4: q rotates 0 -> 1 -> 2 -> 1 -> 2 -> 0; memory
            // variables and PML auxiliaries update in place
            static const int from[5] = {0, 1, 2, 1, 2}, to[5] = {1, 2, 1, 2, 0};
            qin = q[from[irk - 1]]; qout = q[to[irk - 1]];
            thin = thout = th[0]; pin = pout = paux[0];
        }
        a.irk = irk;
        a.last_stage = irk == nrk ? 1 : 0;
        const bool ls = scheme == NEXD_LSRK4;
        a.qin = qin; a.qbase = ls ? qres : q[0]; a.qout = qout;
        a.thin = thin; a.thbase = ls ? thres : th[0]; a.thout = thout;
        a.pin = pin; a.pbase = ls ? pres : paux[0]; a.pout = pout;
        a.qres = ls ? qres : nullptr; a.thres = ls ? thres : nullptr; a.pres = ls ? pres : nullptr;
        a.lsa = ls ? (R)nexd::kLsrkA[irk - 1] : R(0);
        a.lsb = ls ? (R)nexd::kLsrkB[irk - 1] : R(0);
        if (cfg.compensated) {
            auto lo = [&](const R *p) { return p == q[0] ? qlo[0] : (p == q[1] ? qlo[1] : qlo[2]); };
            a.qlo_in = lo(qin); a.qlo_base = qlo[0]; a.qlo_out = lo(qout);
            a.u_slot_lo = u_slot_lo;
        }
    }

    void enqueue_step() {
        nexd::StageArgs<R> a = base_args();
        const bool pml = cfg.pml && n_pml > 0;
        for (int irk = 1; irk <= nrk; ++irk) {
            set_stage_buffers(a, irk);
            // the interior list, then the PML list, on the one stream (a
            // captured fork onto a second stream keeps the buffers its
            // kernels use allocated after the graph and the buffers are
            // freed, ROCm 7.2; each list fills the device anyway)
            // energy partials of the last stage: [shot][plain blocks, PML blocks]
            const bool en = a.last_stage && energy_on();
            a.elist = elist_plain;
            a.nlist = n_plain;
            a.blk_energy = en ? blk_energy : nullptr;
            dispatch(a, false, stream);
            if (pml) {
                if (a.pml_compat) last_flux_dispatch(a.qin, a.pin);
                a.elist = elist_pml;
                a.nlist = n_pml;
                a.blk_energy = en ? blk_energy + nblocks(n_plain, false) : nullptr;
                dispatch(a, true);
            }
            if (cfg.timeint == NEXD_EULER) {
                (void)hipMemcpyAsync(q[0], q[1], sizeof(R) * 9 * ng * cfg.nshot, hipMemcpyDeviceToDevice, stream);
                if (cfg.compensated)
                    (void)hipMemcpyAsync(qlo[0], qlo[1], sizeof(R) * 9 * ng * cfg.nshot, hipMemcpyDeviceToDevice,
                                         stream);
            }
        }
        step_end(1, 1);
    }

    bool ready() const {
        for (char c : have_mat) if (!c) return false;
        for (char c : have_att) if (!c) return false;
        for (char c : have_pml) if (!c) return false;
        for (char c : have_src) if (!c) return false;
        return have_ops && have_mesh && have_rec;
    }

    int reset() {
        const size_t S = cfg.nshot;
        const R tiny = (R)(double)1e-24f;   // the reference's default-REAL 1e-24
        for (int i = 0; i < 3; ++i) {
            int rc = put(q[i], std::vector<R>(S * 9 * ng, tiny));
            if (rc) return rc;
            if (qlo[i] && (rc = put(qlo[i], std::vector<R>(S * 9 * ng, R(0))))) return rc;
        }
        if (cfg.attenuation)
            for (int i = 0; i < 2; ++i) {
                int rc = put(th[i], std::vector<R>(S * 18 * ng, tiny));
                if (rc) return rc;
            }
        if (cfg.pml)
            for (int i = 0; i < 2; ++i) {
                int rc = put(paux[i], std::vector<R>(S * 27 * ng, R(0)));
                if (rc) return rc;
            }
        if (qres) {
            int rc;
            if ((rc = put(qres, std::vector<R>(S * 9 * ng, R(0))))) return rc;
            if (thres && (rc = put(thres, std::vector<R>(S * 18 * ng, R(0))))) return rc;
            if (pres && (rc = put(pres, std::vector<R>(S * 27 * ng, R(0))))) return rc;
        }
        if (nslot) {
            std::vector<R> t(S * nslot * 3 * np, tiny);
            int rc;
            if ((rc = put(u_slot, t)) || (rc = put(a_slot, t))) return rc;
            if (u_slot_lo && (rc = put(u_slot_lo, std::vector<R>(S * nslot * 3 * np, R(0))))) return rc;
        }
        int rc = put(seis, std::vector<R>(S * 9 * cfg.nt * (cfg.nrec > 0 ? cfg.nrec : 1), R(0)));
        if (rc) return rc;
        if (energy && (rc = put(energy, std::vector<double>(S * cfg.nt, 0.0)))) return rc;
        if ((rc = put(d_it, std::vector<int>(1, 0))) || (rc = put(d_done, std::vector<unsigned>(1, 0u)))) return rc;
        step = 0;
        initialized = true;
        return 0;
    }

    int run(int it_begin, int it_end) {
        if (!ready()) return fail(NEXD_E_STATE, "upload operators, mesh, materials, PML, sources and receivers first");
        if (!initialized) {
            int rc = reset();
            if (rc) return rc;
        }
        if (it_begin != step + 1) return fail(NEXD_E_ARG, "it_begin must be the next step");
        if (it_end < it_begin || it_end > cfg.nt) return fail(NEXD_E_ARG, "it_end out of range");
        if (hipSetDevice(cfg.device) != hipSuccess) return fail(NEXD_E_HIP, "hipSetDevice failed");
        if (step == 0) step_end(0, 0);
        if (!graph) {
            hipGraph_t g = nullptr;
            if (hipStreamBeginCapture(stream, hipStreamCaptureModeThreadLocal) != hipSuccess)
                return fail(NEXD_E_HIP, "hipStreamBeginCapture failed");
            enqueue_step();
            if (hipStreamEndCapture(stream, &g) != hipSuccess) return fail(NEXD_E_HIP, "hipStreamEndCapture failed");
            hipError_t e = hipGraphInstantiate(&graph, g, nullptr, nullptr, 0);
            (void)hipGraphDestroy(g);
            if (e != hipSuccess) {
                graph = nullptr;
                return fail(NEXD_E_HIP, std::string("hipGraphInstantiate: ") + hipGetErrorString(e));
            }
        }
        if (hipEventRecord(ev0, stream) != hipSuccess) return fail(NEXD_E_HIP, "hipEventRecord failed");
        for (int it = it_begin; it <= it_end; ++it)
            if (hipGraphLaunch(graph, stream) != hipSuccess) return fail(NEXD_E_HIP, "hipGraphLaunch failed");
        if (hipEventRecord(ev1, stream) != hipSuccess) return fail(NEXD_E_HIP, "hipEventRecord failed");
        hipError_t e = hipStreamSynchronize(stream);
        if (e != hipSuccess) return fail(NEXD_E_HIP, std::string("hipStreamSynchronize: ") + hipGetErrorString(e));
        e = hipGetLastError();
        if (e != hipSuccess) return fail(NEXD_E_HIP, std::string("kernel: ") + hipGetErrorString(e));
        (void)hipEventElapsedTime(&last_ms, ev0, ev1);
        step = it_end;
        return 0;
    }

    // ---- group membership: the graph of one step, without instantiating it,
    // for a parent graph that advances several batches together
    int group_prepare(hipGraph_t *out) override {
        if (!ready()) return fail(NEXD_E_STATE, "upload operators, mesh, materials, PML, sources and receivers first");
        if (!initialized) {
            int rc = reset();
            if (rc) return rc;
        }
        if (hipSetDevice(cfg.device) != hipSuccess) return fail(NEXD_E_HIP, "hipSetDevice failed");
        if (step == 0) {
            step_end(0, 0);   // receiver row 1: the initial state
            if (hipStreamSynchronize(stream) != hipSuccess) return fail(NEXD_E_HIP, "hipStreamSynchronize failed");
        }
        if (hipStreamBeginCapture(stream, hipStreamCaptureModeThreadLocal) != hipSuccess)
            return fail(NEXD_E_HIP, "hipStreamBeginCapture failed");
        enqueue_step();
        if (hipStreamEndCapture(stream, out) != hipSuccess) return fail(NEXD_E_HIP, "hipStreamEndCapture failed");
        return 0;
    }

    int next_step() const override { return step + 1; }
    void set_steps_done(int s) override { step = s; }

    int time_stage(int list, int nrep, double *avg_ms, int *nel) {
        if (!initialized) return fail(NEXD_E_STATE, "run or reset before timing");
        if (nrk == 1) return fail(NEXD_E_UNSUPPORTED, "timing needs a multi-stage scheme");
        const bool pml = cfg.pml && n_pml > 0;
        if (list == 1 && !pml) return fail(NEXD_E_STATE, "no PML element list");
        nexd::StageArgs<R> a = base_args();
        a.blk_energy = nullptr;
        a.elist = list == 0 ? elist_plain : elist_pml;
        a.nlist = list == 0 ? n_plain : n_pml;
        // every RK stage in turn, as the time loop runs them, after a stage-1
        // warm-up; the last element's fluxes are computed once beforehand
        set_stage_buffers(a, 1);
        if (list == 1 && a.pml_compat) last_flux_dispatch(a.qin, a.pin);
        dispatch(a, list == 1);
        if (hipEventRecord(ev0, stream) != hipSuccess) return fail(NEXD_E_HIP, "hipEventRecord failed");
        // the last stage writes its energy partials, as in the time loop
        double *en = energy_on() ? blk_energy + (list == 1 ? nblocks(n_plain, false) : 0) : nullptr;
        for (int r = 0; r < nrep; ++r)
            for (int irk = 1; irk <= nrk; ++irk) {
                set_stage_buffers(a, irk);
                a.blk_energy = a.last_stage ? en : nullptr;
                dispatch(a, list == 1);
            }
        if (hipEventRecord(ev1, stream) != hipSuccess) return fail(NEXD_E_HIP, "hipEventRecord failed");
        if (hipStreamSynchronize(stream) != hipSuccess) return fail(NEXD_E_HIP, "hipStreamSynchronize failed");
        float ms = 0.f;
        (void)hipEventElapsedTime(&ms, ev0, ev1);
        nrep *= nrk;
        *avg_ms = ms / nrep;
        if (nel) *nel = a.nlist;
        return 0;
    }
};

template <typename R>
Batch<R> *as(nexd_batch *b) {
    return static_cast<Batch<R> *>(reinterpret_cast<BatchBase *>(b));
}

BatchBase *base(nexd_batch *b) { return reinterpret_cast<BatchBase *>(b); }

#define DISPATCH(b, call)                                                                   \
    (base(b)->cfg.precision == NEXD_FP32 ? as<float>(b)->call : as<double>(b)->call)

template <typename R>
int upload_ops(Batch<R> *b, const double *dr, const double *ds, const double *dt, const double *lift,
               const double *mass, const int *fmask) {
    nexd::OpsHost o;
    const std::string err = nexd::convert_operators(b->np, b->npf, dr, ds, dt, lift, mass, fmask, o);
    if (!err.empty()) return b->fail(NEXD_E_ARG, err);
    int rc;
    if ((rc = b->upload(&b->D, cvt<R>(nexd::mfma_operator_fragments(o.D, b->np, sizeof(R) == 8)))) ||
        (rc = b->upload(&b->DK, cvt<R>(nexd::mfma_operator_fragments_k(o.D, b->np)))) ||
        (rc = b->upload(&b->LT, cvt<R>(nexd::mfma_lift_fragments(o.LT, b->np, 4 * b->npf)))) ||
        (rc = b->upload(&b->mass, cvt<R>(nexd::mfma_mass_fragments(o.M, b->np)))) || (rc = b->upload(&b->fmask, o.fm)))
        return rc;
    b->fmask_host = o.fm;
    b->have_ops = true;
    b->drop_graph();
    return 0;
}

template <typename R>
int upload_mesh(Batch<R> *b, const int *ibool, const double *const met[9], const double *jac, const double *fscale,
                const int *neighbor, const int *face, const int *ibt, const int *ibn, const double *nn,
                const double *st, const double *tt) {
    if (!b->have_ops) return b->fail(NEXD_E_STATE, "upload operators before the mesh");
    nexd::MeshHost m;
    const std::string err = nexd::convert_mesh(b->cfg.nelem, b->np, b->npf, b->fmask_host, ibool, met, jac, fscale,
                                               neighbor, face, ibt, ibn, nn, st, tt, m);
    if (!err.empty())
        return err[0] == '!' ? b->fail(NEXD_E_UNSUPPORTED, err.substr(1)) : b->fail(NEXD_E_ARG, err);
    int rc;
    if ((rc = b->upload(&b->metric, cvt<R>(m.metric))) || (rc = b->upload(&b->jac, cvt<R>(m.jac))) ||
        (rc = b->upload(&b->fscale, cvt<R>(m.fscale))) || (rc = b->upload(&b->frame, cvt<R>(m.frame))) ||
        (rc = b->upload(&b->nbr, m.nbr)) || (rc = b->upload(&b->ftype, m.ftype)))
        return rc;
    b->have_mesh = true;
    b->drop_graph();
    return 0;
}

template <typename R>
int upload_material(Batch<R> *b, int shot, const double *vpu, const double *vsu, const double *rho,
                    const double *lambdau, const double *muu, const double *mu, const double *lambda) {
    int rc = b->check_shot(shot);
    if (rc) return rc;
    if (!vpu || !vsu || !rho || !lambdau || !muu || !mu || !lambda) return b->fail(NEXD_E_ARG, "null material");
    const int nel = b->cfg.nelem;
    std::vector<R> mc((size_t)nel * 8, R(0)), em((size_t)nel * 2);
    for (int e = 0; e < nel; ++e) {
        if (!(rho[e] > 0)) return b->fail(NEXD_E_ARG, "density must be positive");
        mc[e * 8 + 0] = (R)vpu[e]; mc[e * 8 + 1] = (R)vsu[e]; mc[e * 8 + 2] = (R)rho[e];
        mc[e * 8 + 3] = (R)lambdau[e]; mc[e * 8 + 4] = (R)muu[e];
        em[e * 2] = (R)mu[e]; em[e * 2 + 1] = (R)lambda[e];
    }
    if ((rc = b->put(b->matc + (size_t)shot * nel * 8, mc)) || (rc = b->put(b->emat + (size_t)shot * nel * 2, em)))
        return rc;
    b->have_mat[shot] = 1;
    b->drop_graph();
    return 0;
}

template <typename R>
int upload_anelastic(Batch<R> *b, int shot, const double *wl, const double *av) {
    int rc = b->check_shot(shot);
    if (rc) return rc;
    if (!b->cfg.attenuation) return b->fail(NEXD_E_STATE, "batch created without attenuation");
    if (!wl || !av) return b->fail(NEXD_E_ARG, "null anelastic array");
    bool any_before = false;
    for (char c : b->have_att) any_before |= c != 0;
    for (int i = 0; i < nexd::NMB; ++i)
        if (any_before && b->wl[i] != wl[i])
            return b->fail(NEXD_E_ARG, "relaxation frequencies must be the same for every shot");
    for (int i = 0; i < nexd::NMB; ++i) b->wl[i] = wl[i];
    const size_t n = (size_t)b->cfg.nelem * nexd::NMB * 3;
    std::vector<R> v(av, av + n);
    if ((rc = b->put(b->anel + (size_t)shot * n, v))) return rc;
    b->have_att[shot] = 1;
    b->drop_graph();
    return 0;
}

template <typename R>
int upload_pml(Batch<R> *b, int shot, const int *pmlcheck, const double *const prof[9]) {
    int rc = b->check_shot(shot);
    if (rc) return rc;
    if (!b->cfg.pml) return b->fail(NEXD_E_STATE, "batch created without PML");
    if (!pmlcheck) return b->fail(NEXD_E_ARG, "null PML array");
    for (int k = 0; k < 9; ++k)
        if (!prof[k]) return b->fail(NEXD_E_ARG, "null PML array");
    // the mask is a mesh property: checked against the shots uploaded so
    // far before it replaces the stored one
    std::vector<int> mask(b->cfg.nelem);
    bool changed = false;
    for (int e = 0; e < b->cfg.nelem; ++e) {
        mask[e] = pmlcheck[e] ? 1 : 0;
        if (mask[e] != b->pmlcheck[e]) changed = true;
    }
    bool any_before = false;
    for (char c : b->have_pml) any_before |= c != 0;
    if (changed && any_before) return b->fail(NEXD_E_ARG, "pmlcheck must be the same for every shot");
    b->pmlcheck.swap(mask);
    std::vector<double> pc;
    nexd::pml_coefficients(b->ng, prof[0], prof[1], prof[2], prof[3], prof[4], prof[5], prof[6], prof[7], prof[8], pc);
    if ((rc = b->put(b->pmlc + (size_t)shot * 9 * b->ng, cvt<R>(pc)))) return rc;
    b->have_pml[shot] = 1;
    return b->rebuild_lists();
}

template <typename R>
int upload_sources(Batch<R> *b, int shot, const int *srcelem, const int *srctype, const double *spatial,
                   const double *stf) {
    const int ns = b->cfg.nsrc, nel = b->cfg.nelem, np = b->np, nt = b->cfg.nt;
    if (shot < 0 || shot >= b->cfg.nshot) return b->fail(NEXD_E_ARG, "shot out of range");
    if (ns == 0) { b->have_src[shot] = 1; return 0; }
    if (!srcelem || !srctype || !spatial || !stf) return b->fail(NEXD_E_ARG, "null source array");
    std::vector<int> head(nel, -1), next(ns, -1), st(ns);
    for (int s = ns - 1; s >= 0; --s) {
        const int e = srcelem[s] - 1;
        if (e < 0 || e >= nel) return b->fail(NEXD_E_ARG, "srcelem out of range");
        if (srctype[s] != 0 && srctype[s] != 1) return b->fail(NEXD_E_ARG, "srctype must be 0 or 1");
        st[s] = srctype[s];
        next[s] = head[e];
        head[e] = s;
    }
    std::vector<R> sp(spatial, spatial + (size_t)ns * 6 * np);
    // stf arrives as (nt,nsrc) column-major, i.e. [nsrc][nt] in memory
    std::vector<R> sf((size_t)ns * nt);
    for (int s = 0; s < ns; ++s)
        for (int it = 0; it < nt; ++it) sf[(size_t)s * nt + it] = (R)stf[(size_t)s * nt + it];
    int rc;
    if ((rc = b->put(b->src_head + (size_t)shot * nel, head)) || (rc = b->put(b->src_next + (size_t)shot * ns, next)) ||
        (rc = b->put(b->srctype + (size_t)shot * ns, st)) ||
        (rc = b->put(b->src_spatial + (size_t)shot * ns * 6 * np, sp)) ||
        (rc = b->put(b->src_stf + (size_t)shot * ns * nt, sf)))
        return rc;
    b->have_src[shot] = 1;
    b->drop_graph();
    return 0;
}

template <typename R>
int upload_receivers(Batch<R> *b, const int *recelem, const double *rec_int) {
    const int nr = b->cfg.nrec, np = b->np, nel = b->cfg.nelem;
    if (nr == 0) { b->have_rec = true; return 0; }
    if (!recelem || !rec_int) return b->fail(NEXD_E_ARG, "null receiver array");
    if (b->u_slot) return b->fail(NEXD_E_STATE, "receivers already uploaded");
    std::vector<int> slot_of_elem(nel, -1), slot_of_rec(nr), re(nr);
    int nslot = 0;
    for (int r = 0; r < nr; ++r) {
        const int e = recelem[r] - 1;
        if (e < 0 || e >= nel) return b->fail(NEXD_E_ARG, "recelem out of range");
        if (slot_of_elem[e] < 0) slot_of_elem[e] = nslot++;
        slot_of_rec[r] = slot_of_elem[e];
        re[r] = e;
    }
    int rc;
    const size_t S = b->cfg.nshot;
    if ((rc = b->dalloc(&b->u_slot, S * nslot * 3 * np)) || (rc = b->dalloc(&b->a_slot, S * nslot * 3 * np))) return rc;
    if (b->cfg.compensated && (rc = b->dalloc(&b->u_slot_lo, S * nslot * 3 * np))) return rc;
    b->nslot = nslot;
    std::vector<R> w(rec_int, rec_int + (size_t)np * nr);
    if ((rc = b->put(b->rec_slot, slot_of_elem)) || (rc = b->upload(&b->rec_slot_of, slot_of_rec)) ||
        (rc = b->upload(&b->recelem, re)) || (rc = b->upload(&b->rec_int, w)))
        return rc;
    b->have_rec = true;
    b->drop_graph();
    return 0;
}

template <typename R>
int get_seis(Batch<R> *b, int shot, double *const outs[9]) {
    if (shot < 0 || shot >= b->cfg.nshot) return b->fail(NEXD_E_ARG, "shot out of range");
    const size_t n = (size_t)b->cfg.nt * b->cfg.nrec;
    std::vector<R> h(n);
    for (int k = 0; k < 9; ++k) {
        if (!outs[k]) continue;
        hipError_t e = hipMemcpy(h.data(), b->seis + ((size_t)shot * 9 + k) * n, sizeof(R) * n, hipMemcpyDeviceToHost);
        if (e != hipSuccess) return b->fail(NEXD_E_HIP, std::string("hipMemcpy: ") + hipGetErrorString(e));
        for (size_t i = 0; i < n; ++i) outs[k][i] = (double)h[i];
    }
    return 0;
}

template <typename R>
int upload_element_dt(Batch<R> *b, const double *dt_elem) {
    std::vector<R> v(b->cfg.nelem);
    for (int e = 0; e < b->cfg.nelem; ++e) {
        if (!(dt_elem[e] > 0)) return b->fail(NEXD_E_ARG, "element steps must be positive");
        v[e] = (R)dt_elem[e];
    }
    int rc = b->upload(&b->dt_elem, v);
    if (rc) return rc;
    b->drop_graph();
    return 0;
}

template <typename R>
int get_state(Batch<R> *b, int shot, double *q) {
    if (shot < 0 || shot >= b->cfg.nshot || !q) return b->fail(NEXD_E_ARG, "shot out of range or null output");
    const size_t n = 9 * b->ng;
    std::vector<R> h(n);
    hipError_t e = hipMemcpy(h.data(), b->q[0] + (size_t)shot * n, sizeof(R) * n, hipMemcpyDeviceToHost);
    if (e != hipSuccess) return b->fail(NEXD_E_HIP, std::string("hipMemcpy: ") + hipGetErrorString(e));
    for (size_t i = 0; i < n; ++i) q[i] = (double)h[i];
    return 0;
}

template <typename R>
int get_energy(Batch<R> *b, int shot, double *out) {
    if (shot < 0 || shot >= b->cfg.nshot || !out) return b->fail(NEXD_E_ARG, "shot out of range or null output");
    if (!b->energy_on()) return b->fail(NEXD_E_STATE, "batch created without energy or without PML");
    hipError_t e = hipMemcpy(out, b->energy + (size_t)shot * b->cfg.nt, sizeof(double) * b->cfg.nt, hipMemcpyDeviceToHost);
    if (e != hipSuccess) return b->fail(NEXD_E_HIP, std::string("hipMemcpy: ") + hipGetErrorString(e));
    return 0;
}

template <typename R>
int set_pml_active(Batch<R> *b, int shot, int active) {
    if (shot < 0 || shot >= b->cfg.nshot) return b->fail(NEXD_E_ARG, "shot out of range");
    if (!b->cfg.pml) return b->fail(NEXD_E_STATE, "batch created without PML");
    b->pml_on_host[shot] = active ? 1 : 0;
    hipError_t e = hipMemcpy(b->pml_on + shot, &b->pml_on_host[shot], sizeof(int), hipMemcpyHostToDevice);
    if (e != hipSuccess) return b->fail(NEXD_E_HIP, std::string("hipMemcpy: ") + hipGetErrorString(e));
    return 0;
}

}  // namespace

// A group of batches advanced together: each member's step graph becomes a
// child node of one parent graph, with no dependencies between members, so
// their launches overlap on the device; one graph launch per step.
struct nexd_group {
    std::vector<BatchBase *> members;
    hipStream_t stream = nullptr;
    hipEvent_t ev0 = nullptr, ev1 = nullptr;
    hipGraphExec_t exec = nullptr;
    int step = 0, nt = 0, device = 0;
    float last_ms = 0.f;
    std::string err;
    int fail(int code, const std::string &m) { err = m; return code; }
    ~nexd_group() {
        if (stream) (void)hipStreamSynchronize(stream);
        if (exec) (void)hipGraphExecDestroy(exec);
        if (ev0) (void)hipEventDestroy(ev0);
        if (ev1) (void)hipEventDestroy(ev1);
        if (stream) (void)hipStreamDestroy(stream);
    }
};

static std::string g_group_create_error;

extern "C" {

int nexd_group_create(nexd_batch *const *members, int n, nexd_group **out) {
    if (!members || n <= 0 || !out) { g_group_create_error = "null argument or empty group"; return NEXD_E_ARG; }
    *out = nullptr;
    std::unique_ptr<nexd_group> g(new nexd_group());
    for (int i = 0; i < n; ++i) {
        if (!members[i]) { g_group_create_error = "null member"; return NEXD_E_ARG; }
        for (int k = 0; k < i; ++k)
            if (members[k] == members[i]) { g_group_create_error = "a batch appears twice"; return NEXD_E_ARG; }
        BatchBase *m = base(members[i]);
        if (i == 0) { g->nt = m->cfg.nt; g->device = m->cfg.device; g->step = m->next_step() - 1; }
        if (m->cfg.nt != g->nt || m->cfg.device != g->device || m->next_step() - 1 != g->step) {
            g_group_create_error = "members must share nt, device and the current step";
            return NEXD_E_ARG;
        }
        g->members.push_back(m);
    }
    if (hipSetDevice(g->device) != hipSuccess || hipStreamCreate(&g->stream) != hipSuccess ||
        hipEventCreate(&g->ev0) != hipSuccess || hipEventCreate(&g->ev1) != hipSuccess) {
        g_group_create_error = "HIP stream/event creation failed";
        return NEXD_E_HIP;
    }
    hipGraph_t parent = nullptr;
    if (hipGraphCreate(&parent, 0) != hipSuccess) { g_group_create_error = "hipGraphCreate failed"; return NEXD_E_HIP; }
    for (BatchBase *m : g->members) {
        hipGraph_t child = nullptr;
        int rc = m->group_prepare(&child);
        if (rc) { g_group_create_error = "member: " + m->err; (void)hipGraphDestroy(parent); return rc; }
        hipGraphNode_t node;
        hipError_t e = hipGraphAddChildGraphNode(&node, parent, nullptr, 0, child);
        (void)hipGraphDestroy(child);
        if (e != hipSuccess) {
            g_group_create_error = std::string("hipGraphAddChildGraphNode: ") + hipGetErrorString(e);
            (void)hipGraphDestroy(parent);
            return NEXD_E_HIP;
        }
    }
    hipError_t e = hipGraphInstantiate(&g->exec, parent, nullptr, nullptr, 0);
    (void)hipGraphDestroy(parent);
    if (e != hipSuccess) {
        g->exec = nullptr;
        g_group_create_error = std::string("hipGraphInstantiate: ") + hipGetErrorString(e);
        return NEXD_E_HIP;
    }
    *out = g.release();
    return NEXD_OK;
}

int nexd_group_run(nexd_group *g, int it_begin, int it_end) {
    if (!g) return NEXD_E_ARG;
    if (it_begin != g->step + 1) return g->fail(NEXD_E_ARG, "it_begin must be the next step");
    if (it_end < it_begin || it_end > g->nt) return g->fail(NEXD_E_ARG, "it_end out of range");
    if (hipSetDevice(g->device) != hipSuccess) return g->fail(NEXD_E_HIP, "hipSetDevice failed");
    for (BatchBase *m : g->members)
        if (m->next_step() != it_begin) return g->fail(NEXD_E_STATE, "a member was advanced outside the group");
    if (hipEventRecord(g->ev0, g->stream) != hipSuccess) return g->fail(NEXD_E_HIP, "hipEventRecord failed");
    for (int it = it_begin; it <= it_end; ++it)
        if (hipGraphLaunch(g->exec, g->stream) != hipSuccess) return g->fail(NEXD_E_HIP, "hipGraphLaunch failed");
    if (hipEventRecord(g->ev1, g->stream) != hipSuccess) return g->fail(NEXD_E_HIP, "hipEventRecord failed");
    hipError_t e = hipStreamSynchronize(g->stream);
    if (e != hipSuccess) return g->fail(NEXD_E_HIP, std::string("hipStreamSynchronize: ") + hipGetErrorString(e));
    e = hipGetLastError();
    if (e != hipSuccess) return g->fail(NEXD_E_HIP, std::string("kernel: ") + hipGetErrorString(e));
    (void)hipEventElapsedTime(&g->last_ms, g->ev0, g->ev1);
    for (BatchBase *m : g->members) m->set_steps_done(it_end);
    g->step = it_end;
    return NEXD_OK;
}

int nexd_group_last_run_ms(nexd_group *g, double *ms) {
    if (!g || !ms) return NEXD_E_ARG;
    *ms = g->last_ms;
    return NEXD_OK;
}

void nexd_group_destroy(nexd_group *g) { delete g; }

const char *nexd_group_last_error(const nexd_group *g) { return g ? g->err.c_str() : g_group_create_error.c_str(); }

int nexd_batch_create(const nexd_batch_config *cfg, nexd_batch **out) {
    if (!cfg || !out) { g_batch_create_error = "null argument"; return NEXD_E_ARG; }
    *out = nullptr;
    if (cfg->order < 1 || cfg->order > 5) { g_batch_create_error = "order must be 1..5"; return NEXD_E_UNSUPPORTED; }
    if (cfg->nelem <= 0 || cfg->nt <= 0 || cfg->nshot <= 0 || cfg->nsrc < 0 || cfg->nrec < 0) {
        g_batch_create_error = "nelem, nt, nshot must be positive; nsrc, nrec non-negative";
        return NEXD_E_ARG;
    }
    if (cfg->timeint < 1 || cfg->timeint > 4) {
        g_batch_create_error = "timeint must be 1, 2, 3 or 4 (low-storage RK4)";
        return NEXD_E_UNSUPPORTED;
    }
    if (cfg->timeint == NEXD_LSRK4 && cfg->compensated) {
        g_batch_create_error = "the compensated stage sum is built for the TVD schemes only";
        return NEXD_E_UNSUPPORTED;
    }
    if (cfg->precision != NEXD_FP32 && cfg->precision != NEXD_FP64) {
        g_batch_create_error = "precision must be NEXD_FP32 or NEXD_FP64";
        return NEXD_E_ARG;
    }
    if (!(cfg->dt > 0)) { g_batch_create_error = "dt must be positive"; return NEXD_E_ARG; }
    int ndev = 0;
    if (hipGetDeviceCount(&ndev) != hipSuccess || ndev == 0) { g_batch_create_error = "no HIP device"; return NEXD_E_HIP; }
    if (cfg->device < 0 || cfg->device >= ndev) { g_batch_create_error = "device ordinal out of range"; return NEXD_E_ARG; }
    BatchBase *b;
    int rc;
    if (cfg->precision == NEXD_FP32) {
        auto *p = new Batch<float>();
        p->cfg = *cfg;
        rc = p->init();
        b = p;
    } else {
        auto *p = new Batch<double>();
        p->cfg = *cfg;
        rc = p->init();
        b = p;
    }
    if (rc) {
        g_batch_create_error = b->err;
        delete b;
        return rc;
    }
    *out = reinterpret_cast<nexd_batch *>(b);
    return NEXD_OK;
}

int nexd_batch_upload_operators(nexd_batch *b, const double *dr, const double *ds, const double *dt,
                                const double *lift, const double *mass, const int *fmask) {
    if (!b) return NEXD_E_ARG;
    return base(b)->cfg.precision == NEXD_FP32 ? upload_ops(as<float>(b), dr, ds, dt, lift, mass, fmask)
                                               : upload_ops(as<double>(b), dr, ds, dt, lift, mass, fmask);
}

int nexd_batch_upload_mesh(nexd_batch *b, const int *ibool, const double *rx, const double *ry, const double *rz,
                           const double *sx, const double *sy, const double *sz, const double *tx, const double *ty,
                           const double *tz, const double *jacobian, const double *fscale, const int *neighbor,
                           const int *face, const int *ibt, const int *ibn, const double *nnormal,
                           const double *stangential, const double *ttangential) {
    if (!b) return NEXD_E_ARG;
    const double *met[9] = {rx, ry, rz, sx, sy, sz, tx, ty, tz};
    return base(b)->cfg.precision == NEXD_FP32
               ? upload_mesh(as<float>(b), ibool, met, jacobian, fscale, neighbor, face, ibt, ibn, nnormal, stangential, ttangential)
               : upload_mesh(as<double>(b), ibool, met, jacobian, fscale, neighbor, face, ibt, ibn, nnormal, stangential, ttangential);
}

int nexd_batch_upload_material(nexd_batch *b, int shot, const double *vpu, const double *vsu, const double *rho,
                               const double *lambdau, const double *muu, const double *mu, const double *lambda) {
    if (!b) return NEXD_E_ARG;
    return base(b)->cfg.precision == NEXD_FP32
               ? upload_material(as<float>(b), shot, vpu, vsu, rho, lambdau, muu, mu, lambda)
               : upload_material(as<double>(b), shot, vpu, vsu, rho, lambdau, muu, mu, lambda);
}

int nexd_batch_upload_anelastic(nexd_batch *b, int shot, const double *wl, const double *anelasticvar) {
    if (!b) return NEXD_E_ARG;
    return base(b)->cfg.precision == NEXD_FP32 ? upload_anelastic(as<float>(b), shot, wl, anelasticvar)
                                               : upload_anelastic(as<double>(b), shot, wl, anelasticvar);
}

int nexd_batch_upload_pml(nexd_batch *b, int shot, const int *pmlcheck, const double *ddx, const double *ddy,
                          const double *ddz, const double *alphax, const double *alphay, const double *alphaz,
                          const double *kx, const double *ky, const double *kz) {
    if (!b) return NEXD_E_ARG;
    const double *prof[9] = {ddx, ddy, ddz, alphax, alphay, alphaz, kx, ky, kz};
    return base(b)->cfg.precision == NEXD_FP32 ? upload_pml(as<float>(b), shot, pmlcheck, prof)
                                               : upload_pml(as<double>(b), shot, pmlcheck, prof);
}

int nexd_batch_upload_sources(nexd_batch *b, int shot, const int *srcelem, const int *srctype, const double *spatial,
                              const double *stf) {
    if (!b) return NEXD_E_ARG;
    return base(b)->cfg.precision == NEXD_FP32 ? upload_sources(as<float>(b), shot, srcelem, srctype, spatial, stf)
                                               : upload_sources(as<double>(b), shot, srcelem, srctype, spatial, stf);
}

int nexd_batch_upload_receivers(nexd_batch *b, const int *recelem, const double *rec_int) {
    if (!b) return NEXD_E_ARG;
    return base(b)->cfg.precision == NEXD_FP32 ? upload_receivers(as<float>(b), recelem, rec_int)
                                               : upload_receivers(as<double>(b), recelem, rec_int);
}

int nexd_batch_reset(nexd_batch *b) {
    if (!b) return NEXD_E_ARG;
    return DISPATCH(b, reset());
}

int nexd_batch_run(nexd_batch *b, int it_begin, int it_end) {
    if (!b) return NEXD_E_ARG;
    return DISPATCH(b, run(it_begin, it_end));
}

int nexd_batch_get_seismograms(nexd_batch *b, int shot, double *ux, double *uy, double *uz, double *vx, double *vy,
                               double *vz, double *ax, double *ay, double *az) {
    if (!b) return NEXD_E_ARG;
    double *const outs[9] = {ux, uy, uz, vx, vy, vz, ax, ay, az};
    return base(b)->cfg.precision == NEXD_FP32 ? get_seis(as<float>(b), shot, outs) : get_seis(as<double>(b), shot, outs);
}

int nexd_batch_autotune(nexd_batch *b, int nrep, int *tile, double *ms) {
    if (!b || nrep <= 0) return NEXD_E_ARG;
    return DISPATCH(b, autotune(nrep, tile, ms));
}

int nexd_batch_set_tile(nexd_batch *b, int tile) {
    if (!b) return NEXD_E_ARG;
    return DISPATCH(b, set_tile(tile));
}

int nexd_batch_upload_element_dt(nexd_batch *b, const double *dt_elem) {
    if (!b || !dt_elem) return NEXD_E_ARG;
    return base(b)->cfg.precision == NEXD_FP32 ? upload_element_dt(as<float>(b), dt_elem)
                                               : upload_element_dt(as<double>(b), dt_elem);
}

int nexd_batch_get_state(nexd_batch *b, int shot, double *q) {
    if (!b) return NEXD_E_ARG;
    return base(b)->cfg.precision == NEXD_FP32 ? get_state(as<float>(b), shot, q) : get_state(as<double>(b), shot, q);
}

int nexd_batch_get_energy(nexd_batch *b, int shot, double *energy) {
    if (!b) return NEXD_E_ARG;
    return base(b)->cfg.precision == NEXD_FP32 ? get_energy(as<float>(b), shot, energy)
                                               : get_energy(as<double>(b), shot, energy);
}

int nexd_batch_set_pml_active(nexd_batch *b, int shot, int active) {
    if (!b) return NEXD_E_ARG;
    return base(b)->cfg.precision == NEXD_FP32 ? set_pml_active(as<float>(b), shot, active)
                                               : set_pml_active(as<double>(b), shot, active);
}

int nexd_batch_last_run_ms(nexd_batch *b, double *ms) {
    if (!b || !ms) return NEXD_E_ARG;
    *ms = DISPATCH(b, last_ms);
    return 0;
}

int nexd_batch_time_stage_kernel(nexd_batch *b, int list, int nrep, double *avg_ms, int *nelem_launch) {
    if (!b || !avg_ms || nrep <= 0 || (list != 0 && list != 1)) return NEXD_E_ARG;
    return DISPATCH(b, time_stage(list, nrep, avg_ms, nelem_launch));
}

void nexd_batch_destroy(nexd_batch *b) {
    if (b) delete base(b);
}

const char *nexd_batch_last_error(const nexd_batch *b) {
    return b ? reinterpret_cast<const BatchBase *>(b)->err.c_str() : g_batch_create_error.c_str();
}

}  // extern "C"
