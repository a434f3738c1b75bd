// afid_ensemble_api.hpp -- host side of the batched 2-D ensemble
// (include/afid_ensemble.h).  Included at the end of afid_gpu.hip: every run
// is an afid_ctx on the ensemble's stream (grid, metrics, Poisson factors and
// fields as afid_gpu_init builds them), and the ensemble drives them all with
// k_ens_substep (afid_ensemble.hpp).
#include "afid_ensemble.h"
#include "afid_ensemble.hpp"

#include <numeric>
#include <queue>

struct afid_ens {
    std::string err;
    hipStream_t stream = nullptr;
    int device = 0, n_cu = 256, max_lds = 65536;
    int nruns = 0, nsst = 3, ch = 1, nzmax = 1;
    std::vector<afid_ctx *> runs;
    std::vector<afid_run_control> ctl;
    std::vector<int> host_cur;           // buffer set each context's pointers currently name
    std::vector<DevBuf> bufs;
    int64_t bytes = 0;
    EnsRun *d_runs = nullptr;
    EnsState *d_state = nullptr;
    int *d_id_off = nullptr, *d_id_run = nullptr;      // one run per workgroup (prepare kernel)
    int *d_sched_off = nullptr, *d_sched_run = nullptr;
    int sched_groups = 0;
    afid_ens_tuning tune{1024, 0, 8, 0, 0.0};
    size_t lds = 0;
    // phased form: work items per phase and the per-run factor sets
    bool mixed = false;                          // AFID_MIXED runs (phased form only)
    unsigned long long *d_rslot = nullptr;       // per run: residual slots of the mixed form's sweeps
    int max_sweeps = 1;
    EnsItem *d_items_col = nullptr, *d_items_mode = nullptr, *d_items_line = nullptr, *d_items_cell = nullptr;
    int n_col = 0, n_mode = 0, n_line = 0, n_cell = 0, max_line = 0;
    // per sweep: some run's residual still exceeds its tolerance (device flags)
    unsigned long long *d_any = nullptr;
    double *d_fac = nullptr;
    size_t lds_line = 0;
    hipGraphExec_t graph = nullptr;
    int graph_steps = 0;
};

namespace {

constexpr int kEnsGraphSteps = 10;   // RK3 steps per captured graph

int ens_err(afid_ens *e, int code, const std::string &m) {
    if (e) e->err = m;
    return code;
}
#define ENSCHK(e, expr)                                                         \
    do {                                                                        \
        hipError_t e_ = (expr);                                                 \
        if (e_ != hipSuccess) return ens_err((e), AFID_E_HIP, std::string(#expr) + ": " + hipGetErrorString(e_)); \
    } while (0)

int ens_alloc(afid_ens *e, void **p, size_t bytes) {
    bytes = std::max<size_t>(bytes, 256);
    if (hipMalloc(p, bytes) != hipSuccess) return ens_err(e, AFID_E_MEM, "hipMalloc(" + std::to_string(bytes) + ") failed");
    e->bufs.push_back({*p, bytes});
    e->bytes += int64_t(bytes);
    return AFID_OK;
}

// work estimate of a run (cells; the transforms and tridiagonals scale alike)
double ens_cost(const afid_ctx *c) { return double(c->nx) * c->nz; }

// the substep kernel for (threads, rows per lane)
template <int NT>
int ens_launch_nt(afid_ens *e, int ch, dim3 grid, int ns) {
    const size_t lds = e->lds;
    switch (ch) {
#define AFID_ENS_CASE(CH)                                                                                          \
    case CH:                                                                                                       \
        hipLaunchKernelGGL((k_ens_substep<NT, CH>), grid, dim3(NT), lds, e->stream, e->d_runs, e->d_state,       \
                           e->d_sched_off, e->d_sched_run, ns, e->tune.lines);                                    \
        break;
        AFID_ENS_CASE(1) AFID_ENS_CASE(2) AFID_ENS_CASE(3) AFID_ENS_CASE(4)
#undef AFID_ENS_CASE
    default: return ens_err(e, AFID_E_ARG, "ensemble: nx above 256 is not supported");
    }
    hipError_t err = hipGetLastError();
    if (err != hipSuccess) return ens_err(e, AFID_E_HIP, std::string("launch of k_ens_substep failed: ") + hipGetErrorString(err));
    return AFID_OK;
}

template <int NT>
const void *ens_kernel_ptr(int ch) {
    switch (ch) {
    case 1: return reinterpret_cast<const void *>(k_ens_substep<NT, 1>);
    case 2: return reinterpret_cast<const void *>(k_ens_substep<NT, 2>);
    case 3: return reinterpret_cast<const void *>(k_ens_substep<NT, 3>);
    default: return reinterpret_cast<const void *>(k_ens_substep<NT, 4>);
    }
}

int ens_launched(afid_ens *e, const char *what) {
    hipError_t err = hipGetLastError();
    if (err != hipSuccess) return ens_err(e, AFID_E_HIP, std::string("launch of ") + what + " failed: " + hipGetErrorString(err));
    return AFID_OK;
}

// the phased form of substep ns: factors, columns, then per Poisson sweep
// forward transform, modes, inverse transform, projection (and in the mixed
// form the FP64 residual that gates the next sweep), then finish
template <int CH, typename HT, typename R>
int ens_substep_phased_t(afid_ens *e, int ns) {
    int rc;
    const dim3 blk(ENSP_NT);
    hipLaunchKernelGGL((k_ensp_factors<CH>), dim3(unsigned(e->nruns)), dim3(192), 0, e->stream, e->d_runs, e->d_state, ns, e->d_fac,
                       e->d_any);
    if ((rc = ens_launched(e, "k_ensp_factors"))) return rc;
    hipLaunchKernelGGL((k_ensp_columns<CH, HT>), dim3(unsigned(e->n_col)), blk, 0, e->stream, e->d_runs, e->d_state, e->d_items_col,
                       ns, e->d_fac);
    if ((rc = ens_launched(e, "k_ensp_columns"))) return rc;
    const int sweeps = e->mixed ? e->max_sweeps : 1;
    for (int sw = 1; sw <= sweeps; ++sw) {
        hipLaunchKernelGGL(k_ensp_transform<R>, dim3(unsigned(e->n_line)), blk, e->lds_line, e->stream, e->d_runs, e->d_state,
                           e->d_items_line, ns, e->tune.lines, 0, sw, e->d_any);
        if ((rc = ens_launched(e, "k_ensp_transform"))) return rc;
        hipLaunchKernelGGL((k_ensp_modes<CH, R>), dim3(unsigned(e->n_mode)), blk, 0, e->stream, e->d_runs, e->d_state, e->d_items_mode, sw,
                           e->d_any);
        if ((rc = ens_launched(e, "k_ensp_modes"))) return rc;
        hipLaunchKernelGGL(k_ensp_transform<R>, dim3(unsigned(e->n_line)), blk, e->lds_line, e->stream, e->d_runs, e->d_state,
                           e->d_items_line, ns, e->tune.lines, 1, sw, e->d_any);
        if ((rc = ens_launched(e, "k_ensp_transform"))) return rc;
        hipLaunchKernelGGL(k_ensp_project<R>, dim3(unsigned(e->n_cell)), blk, 0, e->stream, e->d_runs, e->d_state, e->d_items_cell, ns, sw,
                           e->d_any);
        if ((rc = ens_launched(e, "k_ensp_project"))) return rc;
        if (e->mixed) {
            hipLaunchKernelGGL(k_ensp_resid, dim3(unsigned(e->n_cell)), blk, 0, e->stream, e->d_runs, e->d_state, e->d_items_cell, ns, sw,
                               e->d_any);
            if ((rc = ens_launched(e, "k_ensp_resid"))) return rc;
        }
    }
    hipLaunchKernelGGL(k_ensp_finish, dim3(unsigned(e->nruns)), dim3(ENSP_FIN_NT), 0, e->stream, e->d_runs, e->d_state, ns);
    return ens_launched(e, "k_ensp_finish");
}
template <int CH>
int ens_substep_phased_ch(afid_ens *e, int ns) {
    return e->mixed ? ens_substep_phased_t<CH, float, float>(e, ns) : ens_substep_phased_t<CH, double, double>(e, ns);
}

int ens_substep(afid_ens *e, int ns) {
    if (e->tune.phased) {
        switch (e->ch) {
        case 1: return ens_substep_phased_ch<1>(e, ns);
        case 2: return ens_substep_phased_ch<2>(e, ns);
        case 3: return ens_substep_phased_ch<3>(e, ns);
        default: return ens_substep_phased_ch<4>(e, ns);
        }
    }
    const dim3 grid(unsigned(e->sched_groups));
    switch (e->tune.threads) {
    case 256: return ens_launch_nt<256>(e, e->ch, grid, ns);
    case 512: return ens_launch_nt<512>(e, e->ch, grid, ns);
    default: return ens_launch_nt<1024>(e, e->ch, grid, ns);
    }
}

// LDS bytes of a launch shape (0 if it does not fit)
size_t ens_lds_bytes(const afid_ens *e, int nt, int lines) {
    size_t col = 0;
    switch (nt * 10 + e->ch) {
#define AFID_ENS_LDS(NT, CH) case NT * 10 + CH: col = size_t(ens_col_lds<NT, CH>()); break;
        AFID_ENS_LDS(256, 1) AFID_ENS_LDS(256, 2) AFID_ENS_LDS(256, 3) AFID_ENS_LDS(256, 4)
        AFID_ENS_LDS(512, 1) AFID_ENS_LDS(512, 2) AFID_ENS_LDS(512, 3) AFID_ENS_LDS(512, 4)
        AFID_ENS_LDS(1024, 1) AFID_ENS_LDS(1024, 2) AFID_ENS_LDS(1024, 3) AFID_ENS_LDS(1024, 4)
#undef AFID_ENS_LDS
    default: return 0;
    }
    size_t fft = 0;
    for (const afid_ctx *c : e->runs) {
        if (c->nz > ENS_FFT_MAXV * nt) return 0;
        const int L = ens_fft_lines(c->nz, lines, nt);
        fft = std::max(fft, size_t(L) * size_t(line_stride<double>(c->nz)) * sizeof(cpx<double>));
    }
    const size_t b = std::max(std::max(col, fft), size_t(2 * nt / 64) * sizeof(double));
    return b <= size_t(e->max_lds) ? b : 0;
}

// largest-first static schedule of the runs over g workgroups
int ens_set_schedule(afid_ens *e, int g) {
    g = std::max(1, std::min(g, e->nruns));
    std::vector<int> order(size_t(e->nruns));
    std::iota(order.begin(), order.end(), 0);
    std::stable_sort(order.begin(), order.end(), [&](int a, int b) { return ens_cost(e->runs[size_t(a)]) > ens_cost(e->runs[size_t(b)]); });
    std::vector<std::vector<int>> lists((size_t)g);
    using Item = std::pair<double, int>;
    std::priority_queue<Item, std::vector<Item>, std::greater<Item>> load;
    for (int w = 0; w < g; ++w) load.push({0.0, w});
    for (int r : order) {
        Item it = load.top();
        load.pop();
        lists[size_t(it.second)].push_back(r);
        load.push({it.first + ens_cost(e->runs[size_t(r)]), it.second});
    }
    std::vector<int> off(size_t(g) + 1, 0), run;
    for (int w = 0; w < g; ++w) {
        off[size_t(w) + 1] = off[size_t(w)] + int(lists[size_t(w)].size());
        run.insert(run.end(), lists[size_t(w)].begin(), lists[size_t(w)].end());
    }
    ENSCHK(e, hipStreamSynchronize(e->stream));
    ENSCHK(e, hipMemcpy(e->d_sched_off, off.data(), sizeof(int) * off.size(), hipMemcpyHostToDevice));
    ENSCHK(e, hipMemcpy(e->d_sched_run, run.data(), sizeof(int) * run.size(), hipMemcpyHostToDevice));
    e->sched_groups = g;
    return AFID_OK;
}

int ens_drop_graph(afid_ens *e) {
    if (e->graph) (void)hipGraphExecDestroy(e->graph);
    e->graph = nullptr;
    e->graph_steps = 0;
    return AFID_OK;
}

// line-batch work items of the phased transforms for a lines cap
int ens_set_line_items(afid_ens *e, int lines) {
    std::vector<EnsItem> it;
    size_t lds = 0;
    for (int r = 0; r < e->nruns; ++r) {
        const afid_ctx *c = e->runs[size_t(r)];
        if (c->nz > ENS_FFT_MAXV * ENSP_NT) return ens_err(e, AFID_E_ARG, "ensemble: nz too large for the phased transforms");
        // the cell loops divide by nxm with one multiply, exact while cells * nxm <= 2^32
        if ((long long)c->nxm * c->nz * c->nxm > (1LL << 32)) return ens_err(e, AFID_E_ARG, "ensemble: run too large");
        const int L = ens_fft_lines(c->nz, lines, ENSP_NT), nlines = (c->nxm + 1) / 2;
        for (int l0 = 0; l0 < nlines; l0 += L) it.push_back(EnsItem{r, l0});
        const size_t line = e->mixed ? size_t(line_stride<float>(c->nz)) * sizeof(cpx<float>)
                                     : size_t(line_stride<double>(c->nz)) * sizeof(cpx<double>);
        lds = std::max(lds, size_t(L) * line);
    }
    if (int(it.size()) > e->max_line) return ens_err(e, AFID_E_ARG, "ensemble: line items exceed their table");
    if (lds > size_t(e->max_lds)) return ens_err(e, AFID_E_ARG, "ensemble tuning: transform lines do not fit the LDS");
    ENSCHK(e, hipStreamSynchronize(e->stream));
    ENSCHK(e, hipMemcpy(e->d_items_line, it.data(), sizeof(EnsItem) * it.size(), hipMemcpyHostToDevice));
    e->n_line = int(it.size());
    e->lds_line = lds;
    const void *kt = e->mixed ? reinterpret_cast<const void *>(k_ensp_transform<float>)
                              : reinterpret_cast<const void *>(k_ensp_transform<double>);
    ENSCHK(e, hipFuncSetAttribute(kt, hipFuncAttributeMaxDynamicSharedMemorySize, int(std::max<size_t>(lds, 256))));
    return AFID_OK;
}

int ens_apply_tuning(afid_ens *e, const afid_ens_tuning &t) {
    if (t.phased) {
        if (t.lines < 1) return ens_err(e, AFID_E_ARG, "ensemble tuning: lines must be >= 1");
        int rc = ens_set_line_items(e, t.lines);
        if (rc) return rc;
        e->tune.phased = 1;
        e->tune.lines = t.lines;
        e->tune.threads = ENSP_NT;
        e->tune.workgroups = 0;
        ens_drop_graph(e);
        return AFID_OK;
    }
    if (e->mixed) return ens_err(e, AFID_E_ARG, "ensemble tuning: the mixed form runs in the phased form only");
    e->tune.phased = 0;
    if (t.threads != 256 && t.threads != 512 && t.threads != 1024)
        return ens_err(e, AFID_E_ARG, "ensemble tuning: threads must be 256, 512 or 1024");
    if (t.lines < 1) return ens_err(e, AFID_E_ARG, "ensemble tuning: lines must be >= 1");
    const size_t lds = ens_lds_bytes(e, t.threads, t.lines);
    if (lds == 0) return ens_err(e, AFID_E_ARG, "ensemble tuning: the launch shape does not fit the LDS or the transform size");
    e->tune.threads = t.threads;
    e->tune.lines = t.lines;
    e->tune.workgroups = std::max(1, std::min(t.workgroups > 0 ? t.workgroups : e->nruns, e->nruns));
    e->lds = lds;
    const void *k = nullptr;
    switch (t.threads) {
    case 256: k = ens_kernel_ptr<256>(e->ch); break;
    case 512: k = ens_kernel_ptr<512>(e->ch); break;
    default: k = ens_kernel_ptr<1024>(e->ch); break;
    }
    ENSCHK(e, hipFuncSetAttribute(k, hipFuncAttributeMaxDynamicSharedMemorySize, int(lds)));
    ens_drop_graph(e);
    return ens_set_schedule(e, e->tune.workgroups);
}

int ens_steps_eager(afid_ens *e, int nsteps) {
    for (int s = 0; s < nsteps; ++s)
        for (int ns = 1; ns <= e->nsst; ++ns) {
            int rc = ens_substep(e, ns);
            if (rc) return rc;
        }
    return AFID_OK;
}

int ens_capture(afid_ens *e) {
    ens_drop_graph(e);
    hipGraph_t g = nullptr;
    ENSCHK(e, hipStreamBeginCapture(e->stream, hipStreamCaptureModeThreadLocal));
    int rc = ens_steps_eager(e, kEnsGraphSteps);
    hipError_t ec = hipStreamEndCapture(e->stream, &g);
    if (rc) { if (g) (void)hipGraphDestroy(g); return rc; }
    if (ec != hipSuccess) return ens_err(e, AFID_E_HIP, std::string("graph capture: ") + hipGetErrorString(ec));
    ec = hipGraphInstantiate(&e->graph, g, nullptr, nullptr, 0);
    (void)hipGraphDestroy(g);
    if (ec != hipSuccess) { e->graph = nullptr; return ens_err(e, AFID_E_HIP, std::string("graph instantiate: ") + hipGetErrorString(ec)); }
    e->graph_steps = kEnsGraphSteps;
    return AFID_OK;
}

// make run r's context name the buffer set its device state says is current
int ens_sync_run(afid_ens *e, int r) {
    EnsState s;
    ENSCHK(e, hipMemcpyAsync(&s, e->d_state + r, sizeof(s), hipMemcpyDeviceToHost, e->stream));
    ENSCHK(e, hipStreamSynchronize(e->stream));
    afid_ctx *c = e->runs[size_t(r)];
    if (s.cur != e->host_cur[size_t(r)]) {
        std::swap(c->f[AFID_VX], c->alt[0]);
        std::swap(c->f[AFID_VZ], c->alt[2]);
        std::swap(c->f[AFID_TEMP], c->alt[3]);
        e->host_cur[size_t(r)] = s.cur;
    }
    return AFID_OK;
}

int ens_prepare(afid_ens *e, int top) {
    hipLaunchKernelGGL(k_ens_prepare<256>, dim3(unsigned(e->nruns)), dim3(256), 0, e->stream, e->d_runs, e->d_state,
                       e->d_id_off, e->d_id_run, top);
    hipError_t err = hipGetLastError();
    if (err != hipSuccess) return ens_err(e, AFID_E_HIP, std::string("launch of k_ens_prepare failed: ") + hipGetErrorString(err));
    return AFID_OK;
}

int ens_reset_state(afid_ens *e) {
    std::vector<EnsState> st(size_t(e->nruns));
    for (int r = 0; r < e->nruns; ++r) {
        EnsState &s = st[size_t(r)];
        s = EnsState{};
        s.dt = e->ctl[size_t(r)].dt;
        s.cur = e->host_cur[size_t(r)];
    }
    ENSCHK(e, hipStreamSynchronize(e->stream));
    ENSCHK(e, hipMemcpy(e->d_state, st.data(), sizeof(EnsState) * st.size(), hipMemcpyHostToDevice));
    return ens_prepare(e, 1);
}

}  // namespace

extern "C" {

int afid_ens_init(const afid_config *cfgs, const afid_run_control *ctl, int32_t nruns, afid_ens **out) {
    if (!cfgs || !ctl || !out || nruns < 1) return AFID_E_ARG;
    *out = nullptr;
    afid_ens *e = new afid_ens();
    *out = e;     // returned even on failure so afid_ens_last_error can report; caller finalizes
    if (hipGetDevice(&e->device) != hipSuccess) return ens_err(e, AFID_E_HIP, "no HIP device");
    if (hipStreamCreate(&e->stream) != hipSuccess) return ens_err(e, AFID_E_HIP, "hipStreamCreate failed");
    {
        int v = 0;
        if (hipDeviceGetAttribute(&v, hipDeviceAttributeMultiprocessorCount, e->device) == hipSuccess && v > 0) e->n_cu = v;
        if (hipDeviceGetAttribute(&v, hipDeviceAttributeMaxSharedMemoryPerBlock, e->device) == hipSuccess && v > 0)
            e->max_lds = std::min(v, 160 * 1024);
    }
    e->nruns = nruns;
    e->nsst = cfgs[0].nsst;
    e->mixed = cfgs[0].precision == AFID_MIXED;
    int maxnx = 1;
    for (int r = 0; r < nruns; ++r) {
        const afid_config &g = cfgs[r];
        const std::string who = "ensemble run " + std::to_string(r) + ": ";
        if (g.nym != 1) return ens_err(e, AFID_E_ARG, who + "the ensemble holds 2-D runs (nym = 1)");
        if (g.multires || g.salinity || g.phasefield || g.moist)
            return ens_err(e, AFID_E_ARG, who + "no refined grid, salinity, phase field or moisture");
        if (g.precision != cfgs[0].precision) return ens_err(e, AFID_E_ARG, who + "every run needs the same precision");
        if (g.nsst != e->nsst) return ens_err(e, AFID_E_ARG, who + "every run needs the same nsst");
        if (g.gaxis == 2) return ens_err(e, AFID_E_ARG, who + "gravity along y has no effect in a 2-D run (gaxis 1 or 3)");
        if (!(ctl[r].dt > 0.0) || !(ctl[r].limit_cfl > 0.0)) return ens_err(e, AFID_E_ARG, who + "DT and CFLMAX must be positive");
        if (g.nxm + 1 > 256) return ens_err(e, AFID_E_ARG, who + "nxm above 255 is not supported by the ensemble");
        afid_ctx *c = nullptr;
        int rc = ctx_create(&g, e->stream, &c);
        if (rc) {
            std::string m = who + (c ? c->err : std::string("set-up failed"));
            if (c) afid_gpu_finalize(c);
            return ens_err(e, rc, m);
        }
        e->runs.push_back(c);
        e->bytes += c->bytes;
        maxnx = std::max(maxnx, c->nx);
        e->nzmax = std::max(e->nzmax, c->nz);
    }
    e->ctl.assign(ctl, ctl + nruns);
    e->host_cur.assign(size_t(nruns), 0);
    e->ch = (maxnx + 63) / 64;

    // run table
    std::vector<EnsRun> tab((size_t)nruns);
    for (int r = 0; r < nruns; ++r) {
        afid_ctx *c = e->runs[size_t(r)];
        EnsRun &R = tab[size_t(r)];
        R = EnsRun{};
        R.p = make_kp(c);
        R.vxb[0] = c->f[AFID_VX]; R.vxb[1] = c->alt[0];
        R.vzb[0] = c->f[AFID_VZ]; R.vzb[1] = c->alt[2];
        R.tb[0] = c->f[AFID_TEMP]; R.tb[1] = c->alt[3];
        if (e->mixed) {
            R.pff = c->pff;
            R.specf = static_cast<cpx<float> *>(c->spec);
            R.planf = c->plan_zf;
            R.div_tol = c->div_tol;
            R.max_sweeps = c->max_sweeps;
            e->max_sweeps = std::max(e->max_sweeps, c->max_sweeps);
        } else {
            R.pf = c->pois;
            R.spec = static_cast<cpx<double> *>(c->spec);
            R.plan = c->plan_z;
            R.max_sweeps = 1;
        }
        for (int s = 0; s < 3; ++s) { R.gam[s] = c->gam[s]; R.rom[s] = c->rom[s]; R.alm[s] = c->alm[s]; }
        const afid_run_control &k = ctl[r];
        R.dtmin = k.dtmin; R.dtmax = k.dtmax > 0.0 ? k.dtmax : k.dt; R.limit_cfl = k.limit_cfl;
        R.tmax = k.tmax > 0.0 ? k.tmax : HUGE_VAL; R.resid = k.resid > 0.0 ? k.resid : HUGE_VAL;
        R.variable_dt = k.variable_dt ? 1 : 0; R.ntst = k.ntst; R.nsst = c->c.nsst;
    }
    int rc;
    if ((rc = ens_alloc(e, reinterpret_cast<void **>(&e->d_rslot), sizeof(unsigned long long) * 4 * size_t(nruns)))) return rc;
    ENSCHK(e, hipMemset(e->d_rslot, 0, sizeof(unsigned long long) * 4 * size_t(nruns)));
    if ((rc = ens_alloc(e, reinterpret_cast<void **>(&e->d_any), sizeof(unsigned long long) * 4))) return rc;
    ENSCHK(e, hipMemset(e->d_any, 0, sizeof(unsigned long long) * 4));
    if (e->mixed)
        for (int r = 0; r < nruns; ++r) tab[size_t(r)].rslot = e->d_rslot + 4 * r;
    if ((rc = ens_alloc(e, reinterpret_cast<void **>(&e->d_runs), sizeof(EnsRun) * tab.size()))) return rc;
    if ((rc = ens_alloc(e, reinterpret_cast<void **>(&e->d_state), sizeof(EnsState) * size_t(nruns)))) return rc;
    if ((rc = ens_alloc(e, reinterpret_cast<void **>(&e->d_id_off), sizeof(int) * size_t(nruns + 1)))) return rc;
    if ((rc = ens_alloc(e, reinterpret_cast<void **>(&e->d_id_run), sizeof(int) * size_t(nruns)))) return rc;
    if ((rc = ens_alloc(e, reinterpret_cast<void **>(&e->d_sched_off), sizeof(int) * size_t(nruns + 1)))) return rc;
    if ((rc = ens_alloc(e, reinterpret_cast<void **>(&e->d_sched_run), sizeof(int) * size_t(nruns)))) return rc;
    {
        // phased form: fixed work-item tables (columns, modes, cells) and the
        // line table at its largest (one line pair per item)
        std::vector<EnsItem> col, mode, cell;
        int lines = 0;
        for (int r = 0; r < nruns; ++r) {
            const afid_ctx *c = e->runs[size_t(r)];
            for (int i = 0; i < c->nz; i += ENSP_NT / 64) col.push_back(EnsItem{r, i});
            for (int m = 0; m <= c->nz / 2; m += ENSP_NT / 64) mode.push_back(EnsItem{r, m});
            for (int q = 0; q < c->nxm * c->nz; q += ENSP_CELLS) cell.push_back(EnsItem{r, q});
            lines += (c->nxm + 1) / 2;
        }
        e->n_col = int(col.size()); e->n_mode = int(mode.size()); e->n_cell = int(cell.size()); e->max_line = lines;
        if ((rc = ens_alloc(e, reinterpret_cast<void **>(&e->d_items_col), sizeof(EnsItem) * col.size()))) return rc;
        if ((rc = ens_alloc(e, reinterpret_cast<void **>(&e->d_items_mode), sizeof(EnsItem) * mode.size()))) return rc;
        if ((rc = ens_alloc(e, reinterpret_cast<void **>(&e->d_items_cell), sizeof(EnsItem) * cell.size()))) return rc;
        if ((rc = ens_alloc(e, reinterpret_cast<void **>(&e->d_items_line), sizeof(EnsItem) * size_t(lines)))) return rc;
        if ((rc = ens_alloc(e, reinterpret_cast<void **>(&e->d_fac), sizeof(double) * size_t(nruns) * 12 * 64 * e->ch))) return rc;
        ENSCHK(e, hipStreamSynchronize(e->stream));
        ENSCHK(e, hipMemcpy(e->d_items_col, col.data(), sizeof(EnsItem) * col.size(), hipMemcpyHostToDevice));
        ENSCHK(e, hipMemcpy(e->d_items_mode, mode.data(), sizeof(EnsItem) * mode.size(), hipMemcpyHostToDevice));
        ENSCHK(e, hipMemcpy(e->d_items_cell, cell.data(), sizeof(EnsItem) * cell.size(), hipMemcpyHostToDevice));
    }
    ENSCHK(e, hipStreamSynchronize(e->stream));
    ENSCHK(e, hipMemcpy(e->d_runs, tab.data(), sizeof(EnsRun) * tab.size(), hipMemcpyHostToDevice));
    std::vector<int> ioff(size_t(nruns) + 1), irun((size_t)nruns);
    std::iota(ioff.begin(), ioff.end(), 0);
    std::iota(irun.begin(), irun.end(), 0);
    ENSCHK(e, hipMemcpy(e->d_id_off, ioff.data(), sizeof(int) * ioff.size(), hipMemcpyHostToDevice));
    ENSCHK(e, hipMemcpy(e->d_id_run, irun.data(), sizeof(int) * irun.size(), hipMemcpyHostToDevice));
    // default launch shape: 1024 threads, one run per workgroup, 8 lines
    // (the mixed form: phased, 2 lines)
    afid_ens_tuning t = e->mixed ? afid_ens_tuning{ENSP_NT, 0, 2, 1, 0.0} : afid_ens_tuning{1024, nruns, 8, 0, 0.0};
    if ((rc = ens_apply_tuning(e, t))) return rc;
    if ((rc = ens_reset_state(e))) return rc;
    ENSCHK(e, hipStreamSynchronize(e->stream));
    return AFID_OK;
}

int afid_ens_finalize(afid_ens *e) {
    if (!e) return AFID_OK;
    if (e->stream) (void)hipStreamSynchronize(e->stream);
    ens_drop_graph(e);
    for (afid_ctx *c : e->runs) afid_gpu_finalize(c);
    for (auto &b : e->bufs) (void)hipFree(b.p);
    if (e->stream) (void)hipStreamDestroy(e->stream);
    delete e;
    return AFID_OK;
}

const char *afid_ens_last_error(const afid_ens *e) { return e ? e->err.c_str() : "null ensemble"; }
int64_t afid_ens_device_bytes(const afid_ens *e) { return e ? e->bytes : 0; }

static int ens_field_ok(afid_ens *e, int32_t run, int32_t field) {
    if (!e || run < 0 || run >= e->nruns) return ens_err(e, AFID_E_ARG, "ensemble: bad run index");
    switch (field) {
    case AFID_VX: case AFID_VY: case AFID_VZ: case AFID_TEMP: case AFID_PR: case AFID_RUX: case AFID_RUZ:
    case AFID_RUTEMP: case AFID_DPH: case AFID_TEMPBP: case AFID_TEMPTP: return AFID_OK;
    default: return ens_err(e, AFID_E_ARG, "ensemble: field not held by a 2-D run");
    }
}

int afid_ens_upload(afid_ens *e, int32_t run, int32_t field, const double *host, const int64_t lb[3], const int64_t ext[3]) {
    int rc = ens_field_ok(e, run, field);
    if (rc) return rc;
    if ((rc = ens_sync_run(e, run))) return rc;
    afid_ctx *c = e->runs[size_t(run)];
    if (field == AFID_VY) {
        const int64_t n = ext[0] * ext[1] * ext[2];
        for (int64_t q = 0; q < n; ++q)
            if (host[q] != 0.0) return ens_err(e, AFID_E_ARG, "ensemble: v is identically zero in a 2-D run");
        return AFID_OK;
    }
    ens_drop_graph(e);
    if ((rc = afid_upload(c, field, host, lb, ext))) return ens_err(e, rc, c->err);
    return AFID_OK;
}

int afid_ens_download(afid_ens *e, int32_t run, int32_t field, double *host, const int64_t lb[3], const int64_t ext[3]) {
    int rc = ens_field_ok(e, run, field);
    if (rc) return rc;
    if ((rc = ens_sync_run(e, run))) return rc;
    afid_ctx *c = e->runs[size_t(run)];
    if ((rc = afid_download(c, field, host, lb, ext))) return ens_err(e, rc, c->err);
    return AFID_OK;
}

int afid_ens_create_initial_conditions(afid_ens *e, uint64_t seed0, double eps) {
    if (!e) return AFID_E_ARG;
    for (int r = 0; r < e->nruns; ++r) {
        int rc = ens_sync_run(e, r);
        if (rc) return rc;
        afid_ctx *c = e->runs[size_t(r)];
        if ((rc = afid_create_initial_conditions(c, seed0 + uint64_t(r), eps))) return ens_err(e, rc, c->err);
    }
    return ens_reset_state(e);
}

int afid_ens_reset(afid_ens *e) {
    if (!e) return AFID_E_ARG;
    return ens_reset_state(e);
}

int afid_ens_step(afid_ens *e, int32_t nsteps, int32_t graph) {
    if (!e || nsteps < 0) return AFID_E_ARG;
    int rc;
    if (graph) {
        if (!e->graph && (rc = ens_capture(e))) return rc;
        while (nsteps >= e->graph_steps) {
            ENSCHK(e, hipGraphLaunch(e->graph, e->stream));
            nsteps -= e->graph_steps;
        }
    }
    return ens_steps_eager(e, nsteps);
}

int afid_ens_step_timed(afid_ens *e, int32_t nsteps, int32_t graph, double *ms) {
    if (!e || !ms) return AFID_E_ARG;
    int rc = AFID_OK;
    if (graph && !e->graph && (rc = ens_capture(e))) return rc;    // capture outside the timed region
    hipEvent_t a = nullptr, b = nullptr;
    ENSCHK(e, hipEventCreate(&a));
    if (hipEventCreate(&b) != hipSuccess) { (void)hipEventDestroy(a); return ens_err(e, AFID_E_HIP, "hipEventCreate failed"); }
    hipError_t he = hipEventRecord(a, e->stream);
    if (he == hipSuccess) rc = afid_ens_step(e, nsteps, graph);
    if (he == hipSuccess && rc == AFID_OK) he = hipEventRecord(b, e->stream);
    if (he == hipSuccess && rc == AFID_OK) he = hipEventSynchronize(b);
    float t = 0.f;
    if (he == hipSuccess && rc == AFID_OK) he = hipEventElapsedTime(&t, a, b);
    (void)hipEventDestroy(a);
    (void)hipEventDestroy(b);
    if (rc) return rc;
    if (he != hipSuccess) return ens_err(e, AFID_E_HIP, std::string("timed step: ") + hipGetErrorString(he));
    *ms = double(t);
    return AFID_OK;
}

int afid_ens_run_state(afid_ens *e, int32_t run, afid_run_state *out) {
    if (!e || !out || run < 0 || run >= e->nruns) return ens_err(e, AFID_E_ARG, "ensemble: bad run index");
    EnsState s;
    ENSCHK(e, hipMemcpyAsync(&s, e->d_state + run, sizeof(s), hipMemcpyDeviceToHost, e->stream));
    ENSCHK(e, hipStreamSynchronize(e->stream));
    out->time = s.time; out->dt = s.dt; out->cfl = s.cfl; out->max_div = s.dmax;
    out->ntime = s.ntime; out->status = s.status; out->steps = s.steps;
    return AFID_OK;
}

int afid_ens_global_quantities(afid_ens *e, int32_t run, double out[4]) {
    if (!e || !out || run < 0 || run >= e->nruns) return ens_err(e, AFID_E_ARG, "ensemble: bad run index");
    int rc = ens_sync_run(e, run);
    if (rc) return rc;
    afid_ctx *c = e->runs[size_t(run)];
    if ((rc = afid_global_quantities(c, out))) return ens_err(e, rc, c->err);
    return AFID_OK;
}

int afid_ens_set_tuning(afid_ens *e, const afid_ens_tuning *t) {
    if (!e || !t) return AFID_E_ARG;
    ENSCHK(e, hipStreamSynchronize(e->stream));
    return ens_apply_tuning(e, *t);
}

int afid_ens_get_tuning(afid_ens *e, afid_ens_tuning *t) {
    if (!e || !t) return AFID_E_ARG;
    *t = e->tune;
    return AFID_OK;
}

// Candidates: workgroup size x transform lines per batch x persistent grid
// size (one run per workgroup, or the runs dealt over the CUs).  Each is
// timed over `steps` RK3 steps after one untimed step; the fields, the
// driver state and the buffer parities are saved before and restored after.
int afid_ens_autotune(afid_ens *e, int32_t steps, afid_ens_tuning *best, afid_ens_tuning *log, int32_t maxlog,
                      int32_t *nlog) {
    if (!e || steps < 1) return AFID_E_ARG;
    ENSCHK(e, hipStreamSynchronize(e->stream));
    // snapshot
    std::vector<std::pair<void *, size_t>> arrays;      // (array, bytes)
    for (afid_ctx *c : e->runs) {
        const size_t nb = size_t(c->nz) * c->ny * c->ldx;
        for (double *a : {c->f[AFID_VX], c->alt[0], c->f[AFID_VZ], c->alt[2], c->f[AFID_TEMP], c->alt[3], c->f[AFID_PR]})
            arrays.push_back({a, sizeof(double) * nb});
        for (int f : {AFID_DPH, AFID_RUX, AFID_RUZ, AFID_RUTEMP}) {
            if (c->g[f]) arrays.push_back({c->g[f], sizeof(float) * nb});
            else arrays.push_back({c->f[f], sizeof(double) * nb});
        }
    }
    size_t total = 0;
    for (auto &a : arrays) total += a.second;
    unsigned char *save = nullptr;
    ENSCHK(e, hipMalloc(reinterpret_cast<void **>(&save), total));
    std::vector<EnsState> st_save(size_t(e->nruns));
    auto copy_all = [&](bool to_save) -> hipError_t {
        size_t off = 0;
        for (auto &a : arrays) {
            hipError_t r = to_save ? hipMemcpyAsync(save + off, a.first, a.second, hipMemcpyDeviceToDevice, e->stream)
                                   : hipMemcpyAsync(a.first, save + off, a.second, hipMemcpyDeviceToDevice, e->stream);
            if (r != hipSuccess) return r;
            off += a.second;
        }
        return hipSuccess;
    };
    hipError_t herr = copy_all(true);
    if (herr == hipSuccess)
        herr = hipMemcpyAsync(st_save.data(), e->d_state, sizeof(EnsState) * st_save.size(), hipMemcpyDeviceToHost, e->stream);
    if (herr == hipSuccess) herr = hipStreamSynchronize(e->stream);
    if (herr != hipSuccess) { (void)hipFree(save); return ens_err(e, AFID_E_HIP, std::string("autotune snapshot: ") + hipGetErrorString(herr)); }

    const afid_ens_tuning keep = e->tune;
    std::vector<int> groups = {e->nruns};
    for (int g : {e->n_cu, 2 * e->n_cu, e->nruns / 2})
        if (g >= 1 && g < e->nruns && std::find(groups.begin(), groups.end(), g) == groups.end()) groups.push_back(g);
    afid_ens_tuning win{0, 0, 0, 0, HUGE_VAL};
    int n = 0, rc = AFID_OK;
    hipEvent_t a = nullptr, b = nullptr;
    (void)hipEventCreate(&a);
    (void)hipEventCreate(&b);
    // candidate shapes: the phased form over its line caps, then the
    // persistent form over threads x lines x grid
    std::vector<afid_ens_tuning> cands;
    for (int lines : {1, 2, 4, 8})
        if (lines == 1 || ens_fft_lines(e->nzmax, lines, ENSP_NT) != ens_fft_lines(e->nzmax, lines / 2, ENSP_NT))
            cands.push_back(afid_ens_tuning{ENSP_NT, 0, lines, 1, 0.0});
    for (int nt : {256, 512, 1024})
        for (int lines : {2, 4, 8, 16}) {
            if (e->mixed) break;                          // the mixed form runs phased only
            if (ens_lds_bytes(e, nt, lines) == 0) continue;
            if (lines > 2 && ens_fft_lines(e->nzmax, lines, nt) == ens_fft_lines(e->nzmax, lines / 2, nt)) continue;
            for (int g : groups) cands.push_back(afid_ens_tuning{nt, g, lines, 0, 0.0});
        }
    {
        for (afid_ens_tuning t : cands) {
            {
                if ((rc = ens_apply_tuning(e, t))) break;
                if ((rc = ens_steps_eager(e, 1))) break;
                (void)hipEventRecord(a, e->stream);
                if ((rc = ens_steps_eager(e, steps))) break;
                (void)hipEventRecord(b, e->stream);
                (void)hipEventSynchronize(b);
                float ms = 0.f;
                (void)hipEventElapsedTime(&ms, a, b);
                t.ms_per_step = double(ms) / steps;
                t.workgroups = e->tune.workgroups;
                t.threads = e->tune.threads;
                if (log && n < maxlog) log[n] = t;
                ++n;
                if (t.ms_per_step < win.ms_per_step) win = t;
                // restore the state for the next candidate
                herr = copy_all(false);
                if (herr == hipSuccess)
                    herr = hipMemcpyAsync(e->d_state, st_save.data(), sizeof(EnsState) * st_save.size(), hipMemcpyHostToDevice, e->stream);
                if (herr == hipSuccess) herr = hipStreamSynchronize(e->stream);
                if (herr != hipSuccess) { rc = ens_err(e, AFID_E_HIP, std::string("autotune restore: ") + hipGetErrorString(herr)); break; }
            }
        }
    }
    (void)hipEventDestroy(a);
    (void)hipEventDestroy(b);
    (void)hipFree(save);
    if (nlog) *nlog = n;
    if (rc) { (void)ens_apply_tuning(e, keep); return rc; }
    if (n == 0) { (void)ens_apply_tuning(e, keep); return ens_err(e, AFID_E_ARG, "autotune: no launch shape fits"); }
    if ((rc = ens_apply_tuning(e, win))) return rc;
    e->tune.ms_per_step = win.ms_per_step;
    if (best) *best = e->tune;
    return AFID_OK;
}

int afid_ens_synchronize(afid_ens *e) {
    if (!e) return AFID_E_ARG;
    ENSCHK(e, hipStreamSynchronize(e->stream));
    return AFID_OK;
}

}  // extern "C"
