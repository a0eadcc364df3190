// Batched ensemble form (included into swm_capi.hip): cubes of different resolutions, each
// with its own context (geometry, work buffers) and time step, advanced together.  One RK4
// step of the whole ensemble = 4 stage launches (ens_stage_kernel), each walking the
// concatenated tile list of all members; the launches of one step and of RK4_GRAPH_STEPS steps
// are captured once in hipGraphs and replayed.  The tile shape, the on-the-fly metric and the
// grid are chosen by swm_ens_autotune, which times the stage-2 launch of the whole ensemble
// (persistent grids of 1 or 4 blocks per resident slot, or one block per task).

struct swm_ens {
    std::vector<swm_ctx *> m;
    std::vector<void *> state;
    int prec = 64, order = 42;
    hipStream_t stream = nullptr;
    int num_cu = 0;
    int tx = 64, ty = 8, bps = 1;  // tile shape (swm::tile_shape_ok), blocks per resident slot (0: one per task)
    bool otf = false;
    void *tables[4] = {nullptr, nullptr, nullptr, nullptr};  // EnsMember<T>[m] per RK stage
    int2 *tasks = nullptr;  // per task of the launch: (member, task index within the member)
    int ntiles = 0, ntiles_cap = 0;
    std::vector<double> dt;
    std::vector<unsigned> hs_gen;  // the members' orography generations the tables were built with
    bool tables_ok = false;
    hipGraphExec_t graph = nullptr, graph_multi = nullptr;  // one step; RK4_GRAPH_STEPS steps
    hipEvent_t order_ev = nullptr;  // orders the members' own streams around ensemble steps
};

namespace {

void ens_drop_graph(swm_ens *e) {
    if (e->graph) { (void)hipGraphExecDestroy(e->graph); e->graph = nullptr; }
    if (e->graph_multi) { (void)hipGraphExecDestroy(e->graph_multi); e->graph_multi = nullptr; }
}

void free_ens(swm_ens *e) {
    if (!e) return;
    ens_drop_graph(e);
    for (void *t : e->tables)
        if (t) (void)hipFree(t);
    if (e->tasks) (void)hipFree(e->tasks);
    if (e->order_ev) (void)hipEventDestroy(e->order_ev);
    delete e;
}

template <typename T>
void fill_stage_table(swm_ens *e, int stage, std::vector<swm::EnsMember<T>> &tab) {
    tab.resize(e->m.size());
    for (size_t k = 0; k < e->m.size(); ++k) {
        swm_ctx *c = e->m[k];
        const double dt = e->dt[k], c1 = 0.5 * dt, c2 = dt / 6.0, c3 = dt / 3.0;
        swm::StageArgs<T> a{};
        a.hs = (const T *)c->hs;
        // the same buffers and coefficients as enqueue_rk4_step (rk4_mod.f90:28-40)
        switch (stage) {
        case 0: a.y = (const T *)e->state[k]; a.out = (T *)c->yA; a.acc = (T *)c->acc; a.cy = (T)c1; a.cacc = (T)c2; break;
        case 1: a.y = (const T *)c->yA; a.out = (T *)c->yB; a.v0 = (const T *)e->state[k]; a.acc = (T *)c->acc;
                a.cy = (T)c1; a.cacc = (T)c3; break;
        case 2: a.y = (const T *)c->yB; a.out = (T *)c->yA; a.v0 = (const T *)e->state[k]; a.acc = (T *)c->acc;
                a.cy = (T)dt; a.cacc = (T)c3; break;
        default: a.y = (const T *)c->yA; a.out = (T *)e->state[k]; a.acc = (T *)c->acc; a.cy = (T)c2; a.cacc = (T)0; break;
        }
        tab[k].g = make_geo<T>(c);
        tab[k].a = a;
    }
}

int ens_build_tables(swm_ens *e) {
    const size_t M = e->m.size();
    HIP_TRY(hipStreamSynchronize(e->stream));  // no launch of the ensemble still reads the tables
    for (int s = 0; s < 4; ++s) {
        if (e->prec == 64) {
            std::vector<swm::EnsMember<double>> tab;
            fill_stage_table<double>(e, s, tab);
            if (!e->tables[s]) HIP_TRY(hipMalloc(&e->tables[s], M * sizeof(tab[0])));
            HIP_TRY(hipMemcpy(e->tables[s], tab.data(), M * sizeof(tab[0]), hipMemcpyHostToDevice));
        } else {
            std::vector<swm::EnsMember<float>> tab;
            fill_stage_table<float>(e, s, tab);
            if (!e->tables[s]) HIP_TRY(hipMalloc(&e->tables[s], M * sizeof(tab[0])));
            HIP_TRY(hipMemcpy(e->tables[s], tab.data(), M * sizeof(tab[0]), hipMemcpyHostToDevice));
        }
    }
    // task table for the current tile shape: each member's tiles, then its edge segments.  The
    // launch gives XCD x the x-th eighth of the table (xcd_tile), so the members are dealt to 8
    // bins, largest task count first into the bin with the fewest tasks, and the bins are
    // concatenated: every XCD then gets about the same mix of member sizes (the same for an
    // ensemble of 8 k members of each size) instead of the smallest members on the first XCDs
    // and the largest on the last.
    std::vector<int> cnt(M), order(M);
    for (size_t k = 0; k < M; ++k) {
        const int n = e->m[k]->n;
        cnt[k] = 6 * swm::tiles_along(n, e->tx) * swm::tiles_along(n, e->ty) + swm::edge_segs(n, true);
        order[k] = (int)k;
    }
    std::stable_sort(order.begin(), order.end(), [&](int a, int b) { return cnt[a] > cnt[b]; });
    std::vector<std::vector<int>> bins(8);
    std::vector<long long> load(8, 0);
    for (int k : order) {
        const int b = (int)(std::min_element(load.begin(), load.end()) - load.begin());
        bins[b].push_back(k);
        load[b] += cnt[k];
    }
    std::vector<int2> tk;
    for (const auto &bin : bins)
        for (int k : bin)
            for (int t = 0; t < cnt[k]; ++t) tk.push_back(make_int2(k, t));
    if (e->tasks && (int)tk.size() > e->ntiles_cap) { HIP_TRY(hipFree(e->tasks)); e->tasks = nullptr; }
    if (!e->tasks) {
        HIP_TRY(hipMalloc((void **)&e->tasks, tk.size() * sizeof(int2)));
        e->ntiles_cap = (int)tk.size();
    }
    HIP_TRY(hipMemcpy(e->tasks, tk.data(), tk.size() * sizeof(int2), hipMemcpyHostToDevice));
    HIP_TRY(hipDeviceSynchronize());  // the non-blocking stream does not wait for null-stream copies
    e->ntiles = (int)tk.size();
    e->hs_gen.resize(M);
    for (size_t k = 0; k < M; ++k) e->hs_gen[k] = e->m[k]->hs_gen;
    e->tables_ok = true;
    return SWM_OK;
}

// a member's orography was replaced (swm_set_orography frees the old buffer) since the tables
// were built: they hold a stale pointer
bool ens_tables_stale(const swm_ens *e) {
    if (!e->tables_ok || e->hs_gen.size() != e->m.size()) return true;
    for (size_t k = 0; k < e->m.size(); ++k)
        if (e->m[k]->hs_gen != e->hs_gen[k]) return true;
    return false;
}

int ens_launch(swm_ens *e, int stage) {
    swm::EnsLaunch r{};
    r.prec = e->prec;
    r.mode = stage == 0 ? swm::MODE_STAGE1 : stage == 3 ? swm::MODE_STAGE4 : swm::MODE_STAGE23;
    r.tx = e->tx;
    r.ty = e->ty;
    r.otf = e->otf && e->prec == 64;
    r.stream = e->stream;
    r.members = e->tables[stage];
    r.tasks = e->tasks;
    r.nmembers = (int)e->m.size();
    r.ntiles = e->ntiles;
    r.num_cu = e->num_cu;
    r.blocks_per_slot = e->bps;
    hipError_t err;
    switch (e->order) {
    case 42: err = swm::launch_ens_o42(r); break;
    case 43: err = swm::launch_ens_o43(r); break;
    default: err = swm::launch_ens_o63(r); break;
    }
    if (err != hipSuccess) return fail(SWM_EHIP, std::string("ensemble stage launch: ") + hipGetErrorString(err));
    return SWM_OK;
}

// make `to` wait for the work enqueued so far on `from`
int ens_order(swm_ens *e, hipStream_t from, hipStream_t to) {
    if (from == to) return SWM_OK;
    HIP_TRY(hipEventRecord(e->order_ev, from));
    HIP_TRY(hipStreamWaitEvent(to, e->order_ev, 0));
    return SWM_OK;
}

int ens_step(swm_ens *e) {
    int rc;
    for (int s = 0; s < 4; ++s)
        if ((rc = ens_launch(e, s))) return rc;
    return SWM_OK;
}

}  // namespace

extern "C" int swm_ens_create(swm_ctx *const *members, void *const *states, int nmembers, swm_ens **out) {
    if (!members || !states || !out || nmembers < 1) return fail(SWM_EARG, "bad argument");
    *out = nullptr;
    for (int k = 0; k < nmembers; ++k)
        if (members[k]) HIP_TRY(hipStreamSynchronize(members[k]->stream));
    swm_ens *e = new swm_ens();
    e->prec = members[0]->prec;
    e->order = members[0]->grad_order;
    if (e->order != 42 && e->order != 43 && e->order != 63) {
        delete e;
        return fail(SWM_EARG, "the ensemble covers the SBP families 42, 43 and 63");
    }
    for (int k = 0; k < nmembers; ++k) {
        if (!members[k] || !states[k]) { delete e; return fail(SWM_EARG, "NULL member or state"); }
        if (members[k]->prec != e->prec || members[k]->grad_order != e->order || members[k]->div_kind != e->order) {
            delete e;
            return fail(SWM_EARG, "ensemble members must share precision and operator family");
        }
        if (members[k]->n <= 16) { delete e; return fail(SWM_EARG, "ensemble members need N >= 16"); }
        for (int j = 0; j < k; ++j)
            if (members[j] == members[k] || states[j] == states[k]) {
                delete e;
                return fail(SWM_EARG, "each ensemble member needs its own context and state");
            }
        e->m.push_back(members[k]);
        e->state.push_back(states[k]);
    }
    e->stream = members[0]->stream;
    int dev = 0;
    hipDeviceProp_t prop;
    if (hipGetDevice(&dev) != hipSuccess || hipGetDeviceProperties(&prop, dev) != hipSuccess) {
        delete e;
        return fail(SWM_EHIP, "device query failed");
    }
    e->num_cu = prop.multiProcessorCount;
    e->dt.assign(nmembers, 0.0);
    if (hipEventCreateWithFlags(&e->order_ev, hipEventDisableTiming) != hipSuccess) {
        delete e;
        return fail(SWM_EHIP, "event creation failed");
    }
    *out = e;
    return SWM_OK;
}

extern "C" int swm_ens_destroy(swm_ens *e) {
    if (!e) return fail(SWM_EARG, "ensemble is NULL");
    (void)hipStreamSynchronize(e->stream);
    free_ens(e);
    return SWM_OK;
}

extern "C" int swm_ens_set_tile_shape(swm_ens *e, int tile_cols, int tile_rows, int metric_on_the_fly,
                                      int blocks_per_slot) {
    if (!e) return fail(SWM_EARG, "ensemble is NULL");
    if (!swm::tile_shape_ok(tile_cols, tile_rows))
        return fail(SWM_EARG, "tile shape must be 64 x 8, 32 x 16, 16 x 32 or 32 x 8 points");
    if (metric_on_the_fly && e->prec != 64) return fail(SWM_EARG, "metric on the fly is FP64 only");
    if (blocks_per_slot < 0 || blocks_per_slot > 64)
        return fail(SWM_EARG, "blocks_per_slot must be in [0, 64] (0: one block per task)");
    for (swm_ctx *c : e->m)
        if (c->n <= tile_rows) return fail(SWM_EARG, "tile_rows must be smaller than every member's N+1");
    HIP_TRY(hipStreamSynchronize(e->stream));
    ens_drop_graph(e);
    e->tx = tile_cols;
    e->ty = tile_rows;
    e->otf = metric_on_the_fly != 0;
    e->bps = blocks_per_slot;
    e->tables_ok = false;
    return SWM_OK;
}

extern "C" int swm_ens_get_tile_shape(const swm_ens *e, int *tile_cols, int *tile_rows, int *metric_on_the_fly,
                                      int *blocks_per_slot) {
    if (!e) return fail(SWM_EARG, "ensemble is NULL");
    if (tile_cols) *tile_cols = e->tx;
    if (tile_rows) *tile_rows = e->ty;
    if (metric_on_the_fly) *metric_on_the_fly = e->otf ? 1 : 0;
    if (blocks_per_slot) *blocks_per_slot = e->bps;
    return SWM_OK;
}

extern "C" int swm_ens_set_tiling(swm_ens *e, int tile_rows, int metric_on_the_fly, int blocks_per_slot) {
    if (!e) return fail(SWM_EARG, "ensemble is NULL");
    if (tile_rows != 8 && tile_rows != 16 && tile_rows != 32) return fail(SWM_EARG, "tile_rows must be 8, 16 or 32");
    return swm_ens_set_tile_shape(e, swm::tile_width(tile_rows), tile_rows, metric_on_the_fly, blocks_per_slot);
}

extern "C" int swm_ens_get_tiling(const swm_ens *e, int *tile_rows, int *metric_on_the_fly, int *blocks_per_slot) {
    return swm_ens_get_tile_shape(e, nullptr, tile_rows, metric_on_the_fly, blocks_per_slot);
}

extern "C" int swm_ens_rk4_steps(swm_ens *e, const double *dt, int nsteps) {
    if (!e || !dt) return fail(SWM_EARG, "NULL argument");
    if (nsteps < 0) return fail(SWM_EARG, "nsteps must be >= 0");
    bool same = true;
    for (size_t k = 0; k < e->m.size(); ++k) same = same && e->dt[k] == dt[k];
    if (!same || ens_tables_stale(e)) {
        HIP_TRY(hipStreamSynchronize(e->stream));
        ens_drop_graph(e);
        e->dt.assign(dt, dt + e->m.size());
        int rc = ens_build_tables(e);
        if (rc) return rc;
    }
    if (nsteps == 0) return SWM_OK;
    int rc0;
    for (swm_ctx *c : e->m)
        if ((rc0 = ens_order(e, c->stream, e->stream))) return rc0;
    if (!e->graph && !e->graph_multi) {
        int rc = ens_step(e);  // eager first step: sets up the launch configuration
        if (rc) return rc;
        --nsteps;
    }
    // whole runs of RK4_GRAPH_STEPS steps replay the multi-step graph, the rest the one-step graph
    const int nmulti = nsteps / RK4_GRAPH_STEPS, nsingle = nsteps - nmulti * RK4_GRAPH_STEPS;
    for (int which = 0; which < 2; ++which) {
        hipGraphExec_t &ge = which == 0 ? e->graph : e->graph_multi;
        const int per = which == 0 ? 1 : RK4_GRAPH_STEPS;
        if (ge || (which == 0 ? nsingle : nmulti) == 0) continue;
        hipGraph_t g = nullptr;
        HIP_TRY(hipStreamBeginCapture(e->stream, hipStreamCaptureModeThreadLocal));
        int rc = SWM_OK;
        for (int k = 0; k < per && !rc; ++k) rc = ens_step(e);
        hipError_t err = hipStreamEndCapture(e->stream, &g);
        if (rc) { if (g) (void)hipGraphDestroy(g); return rc; }
        if (err != hipSuccess) return fail(SWM_EHIP, std::string("graph capture: ") + hipGetErrorString(err));
        err = hipGraphInstantiate(&ge, g, nullptr, nullptr, 0);
        (void)hipGraphDestroy(g);
        if (err != hipSuccess) { ge = nullptr; return fail(SWM_EHIP, std::string("graph instantiate: ") + hipGetErrorString(err)); }
    }
    for (int s = 0; s < nmulti; ++s) HIP_TRY(hipGraphLaunch(e->graph_multi, e->stream));
    for (int s = 0; s < nsingle; ++s) HIP_TRY(hipGraphLaunch(e->graph, e->stream));
    // later work on a member's own stream (downloads, diagnostics) follows the ensemble steps
    for (swm_ctx *c : e->m)
        if ((rc0 = ens_order(e, e->stream, c->stream))) return rc0;
    return SWM_OK;
}

// Autotuner: times `reps` stage-2 launches of the whole ensemble for every tile shape (64 x 8,
// 32 x 16, 16 x 32, 32 x 8), metric source (FP64: read or on the fly) and grid (a persistent
// grid of 1 or 4 blocks per resident slot, or one block per task); keeps the fastest.  The stage-2 launch reads the work buffers and writes
// yB / acc of every member, so the states are unchanged.
extern "C" int swm_ens_autotune(swm_ens *e, int reps, double *best_us) {
    if (!e || reps < 1) return fail(SWM_EARG, "bad argument");
    if (e->dt[0] == 0.0) {
        // coefficients do not matter for timing; tables need some dt
        for (auto &d : e->dt) d = 1.0;
    }
    struct Cand { int tx, ty; bool otf; int bps; };
    std::vector<Cand> cands;
    int min_n = 1 << 30;
    for (swm_ctx *c : e->m) min_n = std::min(min_n, c->n);
    for (auto sh : {std::make_pair(64, 8), std::make_pair(32, 16), std::make_pair(16, 32), std::make_pair(32, 8)}) {
        if (min_n <= sh.second) continue;  // a tile may not span both y-edges of a panel
        for (int otf = 0; otf < (e->prec == 64 ? 2 : 1); ++otf)
            for (int bps : {1, 4, 0}) cands.push_back({sh.first, sh.second, otf != 0, bps});
    }
    hipEvent_t e0, e1;
    HIP_TRY(hipEventCreate(&e0));
    HIP_TRY(hipEventCreate(&e1));
    double best = 1e300;
    Cand pick = cands[0];
    int rc = SWM_OK;
    for (const Cand &cd : cands) {
        ens_drop_graph(e);
        e->tx = cd.tx; e->ty = cd.ty; e->otf = cd.otf; e->bps = cd.bps;
        if ((rc = ens_build_tables(e))) break;
        for (int r = 0; r < 2 && !rc; ++r) rc = ens_launch(e, 1);
        if (rc) break;
        if (hipEventRecord(e0, e->stream) != hipSuccess) { rc = fail(SWM_EHIP, "event record failed"); break; }
        for (int r = 0; r < reps && !rc; ++r) rc = ens_launch(e, 1);
        if (rc) break;
        float ms = 0.f;
        if (hipEventRecord(e1, e->stream) != hipSuccess || hipEventSynchronize(e1) != hipSuccess ||
            hipEventElapsedTime(&ms, e0, e1) != hipSuccess) { rc = fail(SWM_EHIP, "event timing failed"); break; }
        const double us = 1e3 * ms / reps;
        if (us < best) { best = us; pick = cd; }
    }
    (void)hipEventDestroy(e0);
    (void)hipEventDestroy(e1);
    if (rc) return rc;
    e->tx = pick.tx; e->ty = pick.ty; e->otf = pick.otf; e->bps = pick.bps;
    e->tables_ok = false;
    if (best_us) *best_us = best;
    return SWM_OK;
}

// Measurement helper: `reps` back-to-back stage-2 launches of the whole ensemble.
extern "C" int swm_ens_bench_stage(swm_ens *e, int reps) {
    if (!e || reps < 0) return fail(SWM_EARG, "bad argument");
    if (ens_tables_stale(e)) {
        if (e->dt[0] == 0.0)
            for (auto &d : e->dt) d = 1.0;
        int rc = ens_build_tables(e);
        if (rc) return rc;
    }
    for (int r = 0; r < reps; ++r) {
        int rc = ens_launch(e, 1);
        if (rc) return rc;
    }
    return SWM_OK;
}
