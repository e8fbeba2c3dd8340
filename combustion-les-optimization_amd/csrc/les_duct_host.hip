// les_duct_host.hip - host side of the duct engine: the run pool (arenas,
// first-fit cell and tile ranges, run descriptors), the per-step launch
// sequence and its captured hipGraphs, the side stream, the autotuner, slot
// transfers, the single-run context and the batch ABI (include/les_gpu.h).
// Included by les_engine.hip after les_duct_kernels.hip (one translation unit).
// host code below keeps IEEE operation order
#pragma clang fp contract(off)

// ---------------------------------------------------------------------------
// host side
// ---------------------------------------------------------------------------
#define DHIPC(call)                                                                        \
    do {                                                                                   \
        hipError_t e_ = (call);                                                            \
        if (e_ != hipSuccess) {                                                            \
            snprintf(c->err, sizeof(c->err), "%s failed: %s (%s:%d)", #call,               \
                     hipGetErrorString(e_), __FILE__, __LINE__);                           \
            return LES_ERR_HIP;                                                            \
        }                                                                                  \
    } while (0)

#define DLAUNCH_CHECK() DHIPC(hipGetLastError())

// Calls f with the tile width as a compile-time constant.
template <typename F>
static int tk_dispatch(int tk, F&& f)
{
    switch (tk) {
    case 0: return f(std::integral_constant<int, 0>());
    case 16: return f(std::integral_constant<int, 16>());
    case 32: return f(std::integral_constant<int, 32>());
    case 128: return f(std::integral_constant<int, 128>());
    default: return f(std::integral_constant<int, 64>());
    }
}

static bool tk_valid(int tk) { return tk == 0 || tk == 16 || tk == 32 || tk == 64 || tk == 128; }

// Tiles of one run: TK > 0 covers TK x (256/TK) cells of an x-plane, flat
// tiles 256 consecutive cells of an x-plane.
static void run_tiles(int tk, int nx, int ny, int nz, int& tiles_j, int& tiles_k, long long& ntiles)
{
    if (tk == 0) {
        tiles_j = 1;
        tiles_k = (int)(((long long)ny * nz + 255) / 256);
    } else {
        const int tj = 256 / tk;
        tiles_j = (ny + tj - 1) / tj;
        tiles_k = (nz + tk - 1) / tk;
    }
    ntiles = (long long)nx * tiles_j * tiles_k;
}

// First-fit allocator of ranges in [0, cap).
struct RangeAlloc {
    long long cap = 0;
    std::vector<std::pair<long long, long long>> free_;   // (start, length), sorted
    void reset(long long c) { cap = c; free_.clear(); free_.push_back({0, c}); }
    bool alloc(long long n, long long& start)
    {
        for (size_t a = 0; a < free_.size(); ++a)
            if (free_[a].second >= n) {
                start = free_[a].first;
                free_[a].first += n;
                free_[a].second -= n;
                if (free_[a].second == 0) free_.erase(free_.begin() + a);
                return true;
            }
        return false;
    }
    void release(long long start, long long n)
    {
        size_t a = 0;
        while (a < free_.size() && free_[a].first < start) ++a;
        free_.insert(free_.begin() + a, {start, n});
        // merge neighbours
        for (size_t b = 0; b + 1 < free_.size();) {
            if (free_[b].first + free_[b].second == free_[b + 1].first) {
                free_[b].second += free_[b + 1].second;
                free_.erase(free_.begin() + b + 1);
            } else ++b;
        }
    }
    long long largest() const
    {
        long long m = 0;
        for (auto& f : free_) m = f.second > m ? f.second : m;
        return m;
    }
};

// groups of runs in a pool's stage launches: one per DUCT_GROUP_BYTES of a
// field's values (12288 FP32 tiles), at most DUCT_GROUPS_MAX
#define DUCT_GROUP_BYTES (12288LL * 256 * 4)
#define DUCT_GROUPS_MAX 4

struct DuctState {
    int precision = LES_PREC_FP64;
    size_t esize = 8;
    int max_slots = 1;
    long long pool_cells = 0;
    int tile_cap = 0;
    int nx_max = 0, ny_max = 0, nz_max = 0, nvort_max = 0, nzb = 0;
    int point_cap = 1024;
    int probe_cap = 0;
    int stats_cells = 0;
    int cus = 256;
    // launch configuration (autotuned)
    int tk = 64;
    int grid = 8;                  // tile high-water mark rounded up to a multiple of 8
    // the stage launches in groups of runs (tile ranges split at run
    // boundaries), each group's three stages before the next group's
    int group_max = DUCT_GROUPS_MAX;   // LES_GROUPS=n caps the group count (1: one group)
    int ngroups = 1;
    int gtb[DUCT_GROUPS_MAX] = {}, gtn[DUCT_GROUPS_MAX] = {}, ggrid[DUCT_GROUPS_MAX] = {};
    int xcd_map = 1;
    int zigzag = 0;                // stencil launches alternate their sweep direction (XCD_REVERSE)
    double tuned_ms = 0.0;
    Pool P;
    SlotDesc* d_desc = nullptr;
    SlotDyn* d_dyn = nullptr;
    int* d_tile_slot = nullptr;
    int* d_ntile = nullptr;
    std::vector<SlotDesc> desc;
    std::vector<int> tile_slot;
    std::vector<int> taken;        // slot admitted (its state may not be set yet)
    std::vector<int> host_step;
    std::vector<int> point_base;   // host step index of the oldest unfetched point sample
    std::vector<int> probe_base;   // same for the probe ring
    std::vector<std::vector<les_probe_row>> probe_rows;   // drained, not yet fetched (batch)
    std::vector<std::vector<double>> probe_x;             // station x per run
    RangeAlloc cells, tiles;
    int ntile_hwm = 0;
    double* stage64 = nullptr;
    double* mon_part = nullptr;
    int max_reinit_iters = 0;
    int reinit_odd = 0;            // some active run reinitialises with an odd sweep count
    int max_proj_iters = 0;
    // [0] plain step, [1] step with reinitialisation, [2] DUCT_GRAPH_STEPS plain steps
    hipGraphExec_t gexec[3] = {nullptr, nullptr, nullptr};
    bool graph_ok[3] = {false, false, false};
    int use_graph = 1;
    // side stream for the inflow turbulence, which runs beside the RK stages,
    // forked from and joined back into the main stream
    hipStream_t aux = nullptr;
    hipEvent_t evf[1] = {nullptr}, evj[1] = {nullptr};   // fork / join of the inflow branch
};

static void grid_dims(const les_params& p, GridDims& g)
{
    g.nx = p.nx; g.ny = p.ny; g.nz = p.nz;
    g.n = (long long)p.nx * p.ny * p.nz;
    g.sx = (long long)p.ny * p.nz;
    g.sy = p.nz;
    g.dx = p.lx / (double)p.nx;
    g.dy = p.ly / (double)p.ny;
    g.dz = p.lz / (double)p.nz;
}

static DuctConsts make_duct_consts(const les_params& p)
{
    DuctConsts k;
    memset(&k, 0, sizeof(k));
    k.rho_u_Tu = p.rho_unburnt * p.inlet_temperature;
    k.Tu = p.inlet_temperature; k.Tb = p.burnt_temperature;
    k.U = p.inlet_velocity; k.intensity = p.turbulence_intensity;
    k.wc = p.filter_weight_center;
    k.cs_u = p.cs_unburned; k.cs_b = p.cs_burnt; k.cs_f = p.cs_flame_boost;
    k.s_l = p.s_l; k.nu_unburnt = p.nu_unburnt_approx; k.d3 = p.fractal_d3;
    k.tau = p.burnt_temperature / p.inlet_temperature - 1.0;
    k.pr_t = p.pr_t; k.sc_t = p.sc_t; k.prandtl = p.prandtl;
    k.accel_strength = p.flame_accel_strength; k.accel_width = p.flame_accel_width;
    k.accel_center = p.flame_accel_center; k.weight_x = p.weight_x; k.max_force = p.max_flame_force;
    k.relaxation = p.relaxation_coeff;
    k.courant = p.courant; k.courant_diff = p.courant_diffusion;
    k.fixed_dt = p.fixed_dt;
    k.table_t0 = p.table_t0; k.table_dt = p.table_dt; k.table_n = p.table_n;
    k.scheme = p.advection;
    k.nvort = p.num_vortices;
    k.reinit_interval = p.reinit_interval; k.reinit_iters = p.reinit_iters;
    k.proj_iters = p.proj_iters;
    k.stats_start = p.stats_start_step; k.stats_per_cell = p.stats_per_cell;
    k.L = p.turbulent_length_scale;
    k.dx4 = pow(p.lx / (double)p.nx, 4.0);
    k.ly = p.ly; k.lz = p.lz;
    k.seed = (unsigned long long)p.seed;
    GridDims g;
    grid_dims(p, g);
    const double dd[3] = {g.dx, g.dy, g.dz};
    for (int d = 0; d < 3; ++d) {
        k.rc[DUCT_RC_ID + d] = 1.0 / dd[d];
        k.rc[DUCT_RC_HID + d] = 0.5 / dd[d];
        k.rc[DUCT_RC_ID2 + d] = 1.0 / (dd[d] * dd[d]);
    }
    k.rc[DUCT_RC_IRUT] = 1.0 / k.rho_u_Tu;
    k.rc[DUCT_RC_IDEN] = 1.0 / (6.0 + k.wc);
    k.rc[DUCT_RC_ISC] = 0.5 / k.sc_t;
    k.rc[DUCT_RC_IPR] = 0.5 / k.pr_t;
    k.rc[DUCT_RC_INVD] = 1.0 / (2.0 / (g.dx * g.dx) + 2.0 / (g.dy * g.dy) + 2.0 / (g.dz * g.dz));
    k.rc[DUCT_RC_ITDT] = 1.0 / k.table_dt;
    k.rc[DUCT_RC_IPRAN] = 1.0 / k.prandtl;
    k.rc[DUCT_RC_IW2] = 1.0 / (2.0 * k.accel_width * k.accel_width);
    k.rc[DUCT_RC_DF] = k.nu_unburnt / (k.s_l + LES_SMALL);
    k.rc[DUCT_RC_ITSL] = (k.tau * k.s_l > LES_SMALL) ? 1.0 / (k.tau * k.s_l) : 0.0;
    k.rc[DUCT_RC_ILY] = 1.0 / k.ly;
    k.rc[DUCT_RC_ILZ] = 1.0 / k.lz;
    for (int r = 0; r < DUCT_NRC; ++r) k.rcf[r] = (float)k.rc[r];
    return k;
}

static void* dalloc_bytes(les_gpu* c, size_t bytes)
{
    void* p = nullptr;
    if (hipMalloc(&p, bytes + 256) != hipSuccess) return nullptr;
    c->allocs.push_back(p);
    (void)hipMemset(p, 0, bytes + 256);
    return p;
}

// Allocates the pool of one context.  cfg limits come from the batch config
// (a single-run context is a pool of one slot sized to its grid).
static int pool_create(les_gpu* c, const les_batch_config& cfg)
{
    DuctState* d = new (std::nothrow) DuctState();
    if (!d) return LES_ERR_HIP;
    c->d = d;
    if (hipStreamCreateWithFlags(&d->aux, hipStreamNonBlocking) != hipSuccess) return LES_ERR_HIP;
    for (int a = 0; a < 1; ++a)
        if (hipEventCreateWithFlags(&d->evf[a], hipEventDisableTiming) != hipSuccess ||
            hipEventCreateWithFlags(&d->evj[a], hipEventDisableTiming) != hipSuccess)
            return LES_ERR_HIP;
    d->precision = cfg.precision;
    d->esize = (cfg.precision == LES_PREC_FP32) ? 4 : 8;
    if (const char* fg = getenv("LES_GROUPS")) d->group_max = std::max(1, std::min(DUCT_GROUPS_MAX, atoi(fg)));
    d->max_slots = cfg.max_slots;
    d->pool_cells = cfg.pool_cells;
    d->nx_max = cfg.max_nx; d->ny_max = cfg.max_ny; d->nz_max = cfg.max_nz;
    d->nvort_max = cfg.max_vortices > 0 ? cfg.max_vortices : 1;
    d->point_cap = cfg.point_cap > 0 ? cfg.point_cap : 1024;
    d->probe_cap = cfg.probe_cap > 0 ? cfg.probe_cap : 0;
    d->stats_cells = cfg.stats_per_cell;
    d->nzb = (d->nz_max + 255) / 256;
    d->tk = cfg.tile_k;
    int dev = 0;
    hipDeviceProp_t prop;
    if (hipGetDevice(&dev) == hipSuccess && hipGetDeviceProperties(&prop, dev) == hipSuccess)
        d->cus = prop.multiProcessorCount;
    d->xcd_map = cfg.xcd_map & 7;
    d->zigzag = (cfg.xcd_map & XCD_REVERSE) ? 1 : 0;
    // tiles of all resident runs, for any tile width: per run at most
    // n/256 + nx*ny/2 + nx*nz/16 + nx (ceil bounds of each tile shape)
    {
        long long cap = d->pool_cells / 256 +
                        (long long)d->max_slots * d->nx_max * (d->ny_max / 2 + d->nz_max / 16 + 2) + 64;
        if (cap > (1LL << 30)) return LES_ERR_PARAM;
        d->tile_cap = (int)cap;
    }
    const size_t nb = (size_t)d->pool_cells * d->esize;
    Pool& P = d->P;
    memset(&P, 0, sizeof(P));
    for (int f = 0; f < DNF; ++f)
        if (!(P.A[f] = dalloc_bytes(c, nb)) || !(P.B[f] = dalloc_bytes(c, nb)) || !(P.F[f] = dalloc_bytes(c, nb)))
            return LES_ERR_HIP;
    if (!(P.nut = dalloc_bytes(c, nb)) || !(P.st = dalloc_bytes(c, nb))) return LES_ERR_HIP;
    if (d->stats_cells)
        for (int a = 0; a < 10; ++a)
            if (!(P.stat[a] = (double*)dalloc_bytes(c, 8 * (size_t)d->pool_cells))) return LES_ERR_HIP;
    // the viscosity table (the kernels need no cp: alpha = lambda / (rho cp)
    // with lambda = mu cp / Pr is mu / (rho Pr))
    const int tn = cfg.table_n;
    std::vector<double> mu(tn);
    for (int t = 0; t < tn; ++t) {
        double T = cfg.table_t0 + t * cfg.table_dt;
        mu[t] = 1.8e-5 * pow(T / 300.0, 0.7);
    }
    P.mu_tab = (double*)dalloc_bytes(c, 8 * (size_t)tn);
    if (!P.mu_tab) return LES_ERR_HIP;
    DHIPC(hipMemcpy(P.mu_tab, mu.data(), 8 * (size_t)tn, hipMemcpyHostToDevice));
    const size_t S = (size_t)d->max_slots;
    P.eddy = (double*)dalloc_bytes(c, 8 * 7 * S * d->nvort_max);
    P.inlet = (double*)dalloc_bytes(c, 8 * 3 * S * d->ny_max * d->nz_max);
    P.prof = (double*)dalloc_bytes(c, 8 * 10 * S * d->nx_max);
    P.points = (double*)dalloc_bytes(c, 8 * 4 * DUCT_NPTS * S * d->point_cap);
    P.rho_part = (double*)dalloc_bytes(c, 8 * (size_t)d->tile_cap);
    P.vpart = (double*)dalloc_bytes(c, 8 * (size_t)d->tile_cap);
    P.dpart = (double*)dalloc_bytes(c, 8 * 2 * 3 * (size_t)d->tile_cap);
    P.bcpart = (double*)dalloc_bytes(c, 8 * S * d->ny_max * d->nzb);
    P.spart = (double*)dalloc_bytes(c, 8 * 10 * S * d->nx_max * STAT_CHUNKS);
    P.ppart = (double*)dalloc_bytes(c, 8 * 2 * (size_t)d->tile_cap);
    P.sdue = (int*)dalloc_bytes(c, 4 * (size_t)S);
    P.dtpart = (double*)dalloc_bytes(c, 8 * 3 * DT_BLOCKS_MAX * S);
    P.dtcnt = (unsigned*)dalloc_bytes(c, 4 * S);
    d->d_desc = (SlotDesc*)dalloc_bytes(c, sizeof(SlotDesc) * S);
    d->d_dyn = (SlotDyn*)dalloc_bytes(c, sizeof(SlotDyn) * S);
    d->d_tile_slot = (int*)dalloc_bytes(c, 4 * (size_t)d->tile_cap);
    d->d_ntile = (int*)dalloc_bytes(c, 64);
    d->mon_part = (double*)dalloc_bytes(c, 8 * 6 * 1024);
    if (!P.eddy || !P.inlet || !P.prof || !P.points || !P.rho_part || !P.vpart || !P.dpart || !P.bcpart ||
        !P.spart || !P.ppart || !P.sdue || !P.dtpart || !P.dtcnt || !d->d_desc || !d->d_dyn || !d->d_tile_slot || !d->d_ntile || !d->mon_part)
        return LES_ERR_HIP;
    if (d->precision == LES_PREC_FP32 && !(d->stage64 = (double*)dalloc_bytes(c, 8 * (size_t)d->pool_cells)))
        return LES_ERR_HIP;
    P.desc = d->d_desc; P.dyn = d->d_dyn; P.tile_slot = d->d_tile_slot; P.ntile = d->d_ntile;
    P.tile_cap = d->tile_cap;
    P.ny_max = d->ny_max; P.nz_max = d->nz_max; P.nx_max = d->nx_max; P.nzb = d->nzb;
    P.point_cap = d->point_cap;
    P.probe_cap = d->probe_cap;
    P.sem_split = (d->nvort_max + SEM_CHUNK - 1) / SEM_CHUNK;
    if (P.sem_split < 1) P.sem_split = 1;
    if ((long long)S * P.sem_split > 65535) {
        snprintf(c->err, sizeof(c->err), "max_slots x eddy chunks exceeds the launch grid (65535)");
        return LES_ERR_PARAM;
    }
    if (!(P.sem_part = (double*)dalloc_bytes(c, 8 * S * P.sem_split * 3 * (size_t)d->ny_max * d->nz_max)))
        return LES_ERR_HIP;
    if (d->probe_cap > 0 &&
        !(P.probe = (double*)dalloc_bytes(c, 8 * S * d->probe_cap * LES_NPROBE_MAX * (size_t)PROBE_STRIDE(P))))
        return LES_ERR_HIP;
    d->desc.assign(S, SlotDesc());
    for (auto& x : d->desc) memset(&x, 0, sizeof(x));
    d->tile_slot.assign(d->tile_cap, -1);
    d->taken.assign(S, 0);
    d->host_step.assign(S, 0);
    d->point_base.assign(S, 0);
    d->probe_base.assign(S, 0);
    d->probe_rows.assign(S, std::vector<les_probe_row>());
    d->probe_x.assign(S, std::vector<double>(LES_NPROBE_MAX, 0.0));
    d->cells.reset(d->pool_cells);
    d->tiles.reset(d->tile_cap);
    std::vector<int> neg(d->tile_cap, -1);
    DHIPC(hipMemcpy(d->d_tile_slot, neg.data(), 4 * (size_t)d->tile_cap, hipMemcpyHostToDevice));
    return LES_OK;
}

// The stencil grid: the tile high-water mark rounded up to whole rounds of
// the XCD mapping (8 blocks, 8 chunks of tiles, or whole planes of slot 0's
// layout for the slabs, so that the slab permutation covers every tile).
static void pool_set_grid(DuctState* d)
{
    int r = 8 * xcd_chunk(d->xcd_map);
    if ((d->xcd_map & 7) == 4) {
        const int tpp = d->desc.empty() ? 0 : d->desc[0].tiles_j * d->desc[0].tiles_k;
        if (tpp > 0 && tpp % 8 == 0) r = tpp;
    }
    const int grid = d->ntile_hwm > 0 ? (d->ntile_hwm + r - 1) / r * r : r;
    if (grid != d->grid) { d->grid = grid; d->graph_ok[0] = d->graph_ok[1] = d->graph_ok[2] = false; }
    // groups for a pool of several runs large enough that each group's fields
    // are a fraction of the whole pool's: one group per DUCT_GROUP_BYTES of a
    // field (at most group_max), split at the run boundaries nearest the even
    // split points (runs' tile ranges are disjoint, so no run straddles a
    // boundary of another)
    const int hwm = d->ntile_hwm;
    std::vector<int> bnd;
    int nact = 0;
    for (const SlotDesc& D : d->desc) {
        if (!D.active) continue;
        ++nact;
        for (int c : {D.tile0, D.tile0 + D.ntiles})
            if (c > 0 && c < hwm) bnd.push_back(c);
    }
    int want = (int)std::min<long long>(d->group_max, (long long)hwm * 256 * d->esize / DUCT_GROUP_BYTES);
    if (nact < 2) want = 1;
    int cut[DUCT_GROUPS_MAX + 1] = {0};
    int ng = 1;
    for (int g = 1; g < want; ++g) {
        const long long target = (long long)hwm * g / want;
        int best = -1;
        for (int c : bnd)
            if (c > cut[ng - 1] && (best < 0 || std::llabs(c - target) < std::llabs(best - target))) best = c;
        if (best > 0) cut[ng++] = best;
    }
    cut[ng] = hwm;
    int tb[DUCT_GROUPS_MAX] = {}, tn[DUCT_GROUPS_MAX] = {}, gg[DUCT_GROUPS_MAX] = {};
    if (ng > 1)
        for (int g = 0; g < ng; ++g) {
            tb[g] = cut[g]; tn[g] = cut[g + 1] - cut[g];
            gg[g] = (tn[g] + r - 1) / r * r;
        }
    bool same = ng == d->ngroups;
    for (int g = 0; g < DUCT_GROUPS_MAX; ++g)
        same = same && tb[g] == d->gtb[g] && tn[g] == d->gtn[g] && gg[g] == d->ggrid[g];
    if (!same) d->graph_ok[0] = d->graph_ok[1] = d->graph_ok[2] = false;
    d->ngroups = ng;
    for (int g = 0; g < DUCT_GROUPS_MAX; ++g) { d->gtb[g] = tb[g]; d->gtn[g] = tn[g]; d->ggrid[g] = gg[g]; }
}

// Uploads the descriptor table, tile list and tile high-water mark.
static int pool_sync_tables(les_gpu* c)
{
    DuctState* d = c->d;
    DHIPC(hipStreamSynchronize(c->stream));     // no step may be reading the tables
    int hwm = 0;
    for (int t = 0; t < d->tile_cap; ++t) if (d->tile_slot[t] >= 0) hwm = t + 1;
    d->ntile_hwm = hwm;
    pool_set_grid(d);
    DHIPC(hipMemcpyAsync(d->d_desc, d->desc.data(), sizeof(SlotDesc) * d->max_slots, hipMemcpyHostToDevice, c->stream));
    DHIPC(hipMemcpyAsync(d->d_tile_slot, d->tile_slot.data(), 4 * (size_t)d->tile_cap, hipMemcpyHostToDevice, c->stream));
    DHIPC(hipMemcpyAsync(d->d_ntile, &d->ntile_hwm, 4, hipMemcpyHostToDevice, c->stream));
    DHIPC(hipStreamSynchronize(c->stream));
    int maxit = 0, maxpj = 0, odd = 0;
    for (auto& x : d->desc) {
        if (x.active && x.k.reinit_interval > 0 && x.k.reinit_iters > maxit) maxit = x.k.reinit_iters;
        if (x.active && x.k.reinit_interval > 0 && (x.k.reinit_iters & 1)) odd = 1;
        if (x.active && x.k.proj_iters > maxpj) maxpj = x.k.proj_iters;
    }
    if (maxit != d->max_reinit_iters || maxpj != d->max_proj_iters || odd != d->reinit_odd)
        d->graph_ok[0] = d->graph_ok[1] = d->graph_ok[2] = false;
    d->reinit_odd = odd;
    d->max_reinit_iters = maxit;
    d->max_proj_iters = maxpj;
    return LES_OK;
}

// Admits run p into a free slot (first-fit cells and tiles); the slot starts
// empty (state zero) until init or upload.
static int pool_admit(les_gpu* c, const les_params& p, int n_steps, int* slot_out)
{
    DuctState* d = c->d;
    if (p.mode != LES_MODE_DUCT || p.integrator != LES_INT_SSPRK3 ||
        (p.advection != LES_ADV_WENO5 && p.advection != LES_ADV_MUSCL_MINMOD))
        return LES_ERR_PARAM;
    if (p.precision != d->precision || p.nx > d->nx_max || p.ny > d->ny_max || p.nz > d->nz_max ||
        p.num_vortices > d->nvort_max || p.reinit_iters < 0 || p.reinit_iters > 64 || p.proj_iters < 0 ||
        p.proj_iters > 256)
        return LES_ERR_ADMISSION;
    // kernels address a run's field with 32-bit byte offsets (RunArr)
    if ((long long)p.nx * p.ny * p.nz * (long long)d->esize >= (1ll << 32)) return LES_ERR_PARAM;
    int s = -1;
    for (int a = 0; a < d->max_slots; ++a) if (!d->taken[a]) { s = a; break; }
    if (s < 0) return LES_ERR_ADMISSION;
    GridDims g;
    grid_dims(p, g);
    int tiles_j, tiles_k;
    long long ntiles;
    run_tiles(d->tk, p.nx, p.ny, p.nz, tiles_j, tiles_k, ntiles);
    long long c0 = 0, t0 = 0;
    if (!d->cells.alloc(g.n, c0)) return LES_ERR_ADMISSION;
    if (!d->tiles.alloc(ntiles, t0)) { d->cells.release(c0, g.n); return LES_ERR_ADMISSION; }
    SlotDesc& D = d->desc[s];
    memset(&D, 0, sizeof(D));
    D.g = g;
    D.off = c0;
    D.poff = (long long)s * 3 * d->ny_max * d->nz_max;
    D.xoff = (long long)s * 10 * d->nx_max;
    D.eoff = (long long)s * d->nvort_max;
    D.ptoff = (long long)s * 4 * DUCT_NPTS * d->point_cap;
    D.tile0 = (int)t0; D.ntiles = (int)ntiles; D.tiles_j = tiles_j; D.tiles_k = tiles_k;
    D.active = 0;               // set once the run has a state (init or upload)
    D.n_steps = n_steps;
    int t = 0;
    for (int a = 0; a < 4; ++a)
        for (int b = 0; b < 4; ++b)
            for (int e = 0; e < 4; ++e) {
                int ii = (a + 1) * p.nx / 5, jj = (2 * b + 1) * p.ny / 8, kk = (2 * e + 1) * p.nz / 8;
                D.point_idx[t++] = (int)(((long long)ii * p.ny + jj) * p.nz + kk);
            }
    D.npts = DUCT_NPTS;
    D.k = make_duct_consts(p);
    D.nprobe = p.num_probe_x;
    for (int n = 0; n < p.num_probe_x; ++n) {
        D.probe_i[n] = (int)(p.probe_x[n] / g.dx);
        d->probe_x[s][n] = p.probe_x[n];
    }
    d->probe_rows[s].clear();
    D.proff = (long long)s * d->probe_cap * LES_NPROBE_MAX * (2 + 6LL * d->ny_max);
    for (long long a = 0; a < ntiles; ++a) d->tile_slot[t0 + a] = s;
    d->host_step[s] = 0;
    d->point_base[s] = 0;
    d->probe_base[s] = 0;
    d->taken[s] = 1;
    SlotDyn y;
    memset(&y, 0, sizeof(y));
    DHIPC(hipMemcpyAsync(d->d_dyn + s, &y, sizeof(y), hipMemcpyHostToDevice, c->stream));
    int e = pool_sync_tables(c);
    if (e != LES_OK) return e;
    *slot_out = s;
    return LES_OK;
}

static int pool_release(les_gpu* c, int s)
{
    DuctState* d = c->d;
    if (s < 0 || s >= d->max_slots || !d->taken[s]) return LES_ERR_PARAM;
    SlotDesc& D = d->desc[s];
    d->cells.release(D.off, D.g.n);
    d->tiles.release(D.tile0, D.ntiles);
    for (int a = 0; a < D.ntiles; ++a) d->tile_slot[D.tile0 + a] = -1;
    D.active = 0;
    d->taken[s] = 0;
    return pool_sync_tables(c);
}

// ---- launches ----
#define DUCT_GRAPH_STEPS 4   // steps of the multi-step captured graph
template <typename real, int TK>
static int stage_launch(les_gpu* c, int stage, real* const* Lout, int group = -1)
{
    DuctState* d = c->d;
    // a group's launches carry its tile range and cover only its grid
    Pool P = d->P;
    int G = d->grid;
    if (group >= 0) { P.tb = d->gtb[group]; P.tn = d->gtn[group]; G = d->ggrid[group]; }
    const int qin_B = stage > 0, qout_B = stage < 2;
    // zigzag: successive launches alternate the direction in which each XCD
    // walks its tiles (three per stage, so the pattern flips from stage to
    // stage), so each starts on the tiles whose outputs the previous one wrote
    // last (still in that XCD's L2); a tuned choice (it pays for a pool of many
    // small runs, not for one large run)
    const int zz = d->zigzag ? XCD_REVERSE : 0;
    const int zf = (stage & 1) ? zz : 0;
    hipLaunchKernelGGL((k_duct_filter<real, TK>), dim3(G), dim3(256), 0, c->stream, P, qin_B, d->xcd_map | zf);
    hipLaunchKernelGGL((k_duct_sgs<real, TK>), dim3(G), dim3(256), 0, c->stream, P, stage, d->xcd_map | (zz ^ zf));
    DLAUNCH_CHECK();
    RhsArgs<real> a;
    a.qin_is_B = qin_B; a.qout_is_B = qout_B; a.stage = stage;
    a.ca = (stage == 1) ? 0.75 : (1.0 / 3.0);
    a.cb = (stage == 1) ? 0.25 : (2.0 / 3.0);
    a.track_vel = (stage == 2);
    a.xcd_map = d->xcd_map | zf;
    for (int f = 0; f < DNF; ++f) a.Lout[f] = Lout ? Lout[f] : nullptr;
    if (c->timing) DHIPC(hipEventRecord(c->ev[0], c->stream));
    hipLaunchKernelGGL((k_duct_rhs_all<real, TK>), dim3(G), dim3(256), 0, c->stream, P, a);
    DLAUNCH_CHECK();
    if (c->timing) {
        DHIPC(hipEventRecord(c->ev[1], c->stream));
        DHIPC(hipEventSynchronize(c->ev[1]));
        float ms = 0.f;
        DHIPC(hipEventElapsedTime(&ms, c->ev[0], c->ev[1]));
        c->rhs_ms_sum += ms;
        c->rhs_launches++;
    }
    return LES_OK;
}

// One engine step of every active, unfinished slot; all launches on c->stream.
// Forks the side stream off the main stream (f) or joins it back (j); in the
// per-kernel timing mode everything stays on the main stream.
static int side_fork(les_gpu* c, int f)
{
    DuctState* d = c->d;
    if (c->timing) return LES_OK;
    DHIPC(hipEventRecord(d->evf[f], c->stream));
    DHIPC(hipStreamWaitEvent(d->aux, d->evf[f], 0));
    return LES_OK;
}

static int side_join(les_gpu* c, int f)
{
    DuctState* d = c->d;
    if (c->timing) return LES_OK;
    DHIPC(hipEventRecord(d->evj[f], d->aux));
    DHIPC(hipStreamWaitEvent(c->stream, d->evj[f], 0));
    return LES_OK;
}

template <typename real, int TK>
static int step_launches(les_gpu* c, bool reinit)
{
    DuctState* d = c->d;
    const int S = d->max_slots;
    hipStream_t side = c->timing ? c->stream : d->aux;
    {
        // blocks per slot for the DT maxima: one per 1024 tiles of the largest run
        const long long tl = (long long)d->nx_max * d->ny_max * d->nz_max / 256;
        const int nb = (int)std::max<long long>(1, std::min<long long>(DT_BLOCKS_MAX, (tl + 1023) / 1024));
        hipLaunchKernelGGL(k_duct_dt, dim3(nb, S), dim3(256), 0, c->stream, d->P, nb);
    }
    // the inflow turbulence (eddies -> SEM plane) is read only by the boundary
    // update after the stages, so it runs beside them, forked after the DT.
    // The first stage's launches are captured before the side chain's: a
    // captured step keeps the branch captured first on the queue of the node
    // it forks from, and with the side chain captured first the stages moved
    // to a second queue (about 10 us more before the first filter and 4 us
    // more at the join, per step)
    int e = side_fork(c, 0);
    if (e != LES_OK) return e;
    for (int gs = 0; gs < 3 * d->ngroups; ++gs) {
        e = stage_launch<real, TK>(c, gs % 3, nullptr, d->ngroups > 1 ? gs / 3 : -1);
        if (e != LES_OK) return e;
        if (gs != 0) continue;
        hipLaunchKernelGGL(k_duct_eddy_update, dim3((d->nvort_max + 255) / 256, S), dim3(256), 0, side, d->P);
        hipLaunchKernelGGL(k_duct_sem_plane,
                           dim3((d->ny_max + SEM_PATCH - 1) / SEM_PATCH * ((d->nz_max + SEM_PATCH - 1) / SEM_PATCH) *
                                    d->P.sem_split, S),
                           dim3(256), 0, side, d->P);
        if ((long long)d->ny_max * d->nz_max >= 8192)
            hipLaunchKernelGGL(k_duct_sem_norm<1024>, dim3(3, S), dim3(1024), 0, side, d->P);
        else
            hipLaunchKernelGGL(k_duct_sem_norm<256>, dim3(3, S), dim3(256), 0, side, d->P);
        DLAUNCH_CHECK();
    }
    if (d->max_proj_iters > 0) {
        hipLaunchKernelGGL((k_duct_proj_div<real, TK>), dim3(d->grid), dim3(256), 0, c->stream, d->P, d->xcd_map);
        for (int it = 0; it < d->max_proj_iters; ++it)
            hipLaunchKernelGGL((k_duct_proj_sweep<real, TK>), dim3(d->grid), dim3(256), 0, c->stream, d->P, it,
                               d->xcd_map);
        hipLaunchKernelGGL((k_duct_proj_correct<real, TK>), dim3(d->grid), dim3(256), 0, c->stream, d->P, d->xcd_map);
        hipLaunchKernelGGL(k_duct_proj_final, dim3(S), dim3(1024), 0, c->stream, d->P);
        DLAUNCH_CHECK();
    }
    if ((e = side_join(c, 0)) != LES_OK) return e;
    hipLaunchKernelGGL((k_duct_bc<real>), dim3(d->nzb, d->ny_max, S), dim3(256), 0, c->stream, d->P);
    DLAUNCH_CHECK();
    if (reinit) {
        // one launch per sweep (ping-pong), and the move back into phi where
        // some run has an odd sweep count
        for (int it = 0; it < d->max_reinit_iters; ++it)
            hipLaunchKernelGGL((k_duct_reinit<real, TK>), dim3(d->grid), dim3(256), 0, c->stream, d->P, 1, it,
                               d->xcd_map);
        if (d->reinit_odd)
            hipLaunchKernelGGL((k_duct_reinit<real, TK>), dim3(d->grid), dim3(256), 0, c->stream, d->P, 2, 0,
                               d->xcd_map);
        DLAUNCH_CHECK();
    }
    // outputs of the new state (statistics chunks, probe rows, point samples)
    // in one launch, then the plane sums and the step end in another (the
    // outputs as separate launches, or as a side branch, cost more than their
    // own work: each short launch and each join added its own latency)
    {
        const long long plm = (long long)d->ny_max * d->nz_max;
        const int sch = (int)std::min<long long>(STAT_CHUNKS, std::max<long long>(1, (plm + STAT_CHUNK_CELLS - 1) / STAT_CHUNK_CELLS));
        const int nout = sch * d->nx_max * S + (d->probe_cap > 0 ? LES_NPROBE_MAX * S : 0) + S;
        hipLaunchKernelGGL((k_duct_outputs<real>), dim3(nout), dim3(256), 0, c->stream, d->P, sch, S);
        const int fx = (d->nx_max * 10 + 255) / 256;
        hipLaunchKernelGGL(k_duct_step_close, dim3(fx * S + (S + 255) / 256), dim3(256), 0, c->stream, d->P, fx, S);
    }
    DLAUNCH_CHECK();
    return LES_OK;
}

static int pool_step_launches(les_gpu* c, bool reinit)
{
    return tk_dispatch(c->d->tk, [&](auto K) {
        return c->d->precision == LES_PREC_FP32 ? step_launches<float, decltype(K)::value>(c, reinit)
                                                : step_launches<double, decltype(K)::value>(c, reinit);
    });
}

// Advances the pool by n engine steps: graph replay, one graph with and one
// without the reinitialisation launches (chosen from the host's step mirror).
static int pool_steps(les_gpu* c, int n)
{
    DuctState* d = c->d;
    // the reinitialisation is due in the step m steps from now for some
    // unfinished run
    auto reinit_due = [&](int m) {
        for (int s = 0; s < d->max_slots; ++s) {
            const SlotDesc& D = d->desc[s];
            if (!D.active || (D.n_steps > 0 && d->host_step[s] + m >= D.n_steps)) continue;
            if (D.k.reinit_interval > 0 && (d->host_step[s] + m + 1) % D.k.reinit_interval == 0) return true;
        }
        return false;
    };
    for (int a = 0; a < n;) {
        const bool reinit = reinit_due(0);
        // runs of DUCT_GRAPH_STEPS plain steps replay one graph of that many
        // steps (one graph launch, and its start-up gap, instead of several)
        int gi = reinit ? 1 : 0, steps = 1;
        if (!reinit && n - a >= DUCT_GRAPH_STEPS) {
            bool plain = true;
            for (int m = 1; m < DUCT_GRAPH_STEPS && plain; ++m) plain = !reinit_due(m);
            if (plain) { gi = 2; steps = DUCT_GRAPH_STEPS; }
        }
        bool done = false;
        if (d->use_graph && !c->timing) {
            if (!d->graph_ok[gi]) {
                if (d->gexec[gi]) { (void)hipGraphExecDestroy(d->gexec[gi]); d->gexec[gi] = nullptr; }
                if (hipStreamBeginCapture(c->stream, hipStreamCaptureModeThreadLocal) == hipSuccess) {
                    int e = LES_OK;
                    for (int m = 0; m < steps && e == LES_OK; ++m) e = pool_step_launches(c, reinit);
                    hipGraph_t gr = nullptr;
                    hipError_t ce = hipStreamEndCapture(c->stream, &gr);
                    if (e == LES_OK && ce == hipSuccess && gr &&
                        hipGraphInstantiate(&d->gexec[gi], gr, nullptr, nullptr, 0) == hipSuccess)
                        d->graph_ok[gi] = true;
                    if (gr) (void)hipGraphDestroy(gr);
                }
                (void)hipGetLastError();
                if (!d->graph_ok[gi]) d->use_graph = 0;
            }
            if (d->graph_ok[gi]) {
                DHIPC(hipGraphLaunch(d->gexec[gi], c->stream));
                done = true;
            }
        }
        if (!done) {
            for (int m = 0; m < steps; ++m) {
                int e = pool_step_launches(c, reinit);
                if (e != LES_OK) return e;
            }
        }
        for (int m = 0; m < steps; ++m)
            for (int s = 0; s < d->max_slots; ++s) {
                const SlotDesc& D = d->desc[s];
                if (!D.active || (D.n_steps > 0 && d->host_step[s] >= D.n_steps)) continue;
                d->host_step[s]++;
            }
        a += steps;
    }
    return LES_OK;
}

// ---- state exchange of one slot (fields rho*u, rho*v, rho*w, phi, T as double) ----
static int slot_xfer(les_gpu* c, int s, void* arena, double* host, bool upload)
{
    DuctState* d = c->d;
    const SlotDesc& D = d->desc[s];
    const size_t n = (size_t)D.g.n;
    char* dev = (char*)arena + (size_t)D.off * d->esize;
    if (d->precision == LES_PREC_FP64) {
        if (upload) DHIPC(hipMemcpyAsync(dev, host, n * 8, hipMemcpyHostToDevice, c->stream));
        else DHIPC(hipMemcpyAsync(host, dev, n * 8, hipMemcpyDeviceToHost, c->stream));
        return LES_OK;
    }
    double* tmp = d->stage64;
    int nb = (int)((n + 255) / 256);
    if (upload) {
        DHIPC(hipMemcpyAsync(tmp, host, n * 8, hipMemcpyHostToDevice, c->stream));
        hipLaunchKernelGGL(k_to_real<float>, dim3(nb), dim3(256), 0, c->stream, (long long)n, tmp, (float*)dev);
    } else {
        hipLaunchKernelGGL(k_to_double<float>, dim3(nb), dim3(256), 0, c->stream, (long long)n, (const float*)dev, tmp);
        DHIPC(hipMemcpyAsync(host, tmp, n * 8, hipMemcpyDeviceToHost, c->stream));
    }
    DHIPC(hipStreamSynchronize(c->stream));
    return LES_OK;
}

// After new state in slot s: DT maxima of the state (|u|), no diffusive history.
static int slot_state_changed(les_gpu* c, int s)
{
    DuctState* d = c->d;
    const SlotDesc& D = d->desc[s];
    DHIPC(hipMemsetAsync(d->P.bcpart + (size_t)s * d->ny_max * d->nzb, 0, 8 * (size_t)d->ny_max * d->nzb, c->stream));
    tk_dispatch(d->tk, [&](auto K) {
        constexpr int TK = decltype(K)::value;
        if (d->precision == LES_PREC_FP32)
            hipLaunchKernelGGL((k_duct_velmax<float, TK>), dim3(d->grid), dim3(256), 0, c->stream, d->P, s, d->xcd_map);
        else
            hipLaunchKernelGGL((k_duct_velmax<double, TK>), dim3(d->grid), dim3(256), 0, c->stream, d->P, s, d->xcd_map);
        return 0;
    });
    DLAUNCH_CHECK();
    DHIPC(hipStreamSynchronize(c->stream));
    (void)D;
    return LES_OK;
}

// New state of slot s (the step counter is kept, as the oracle's set_state
// keeps its step); the run becomes active if it was only admitted.
static int slot_upload(les_gpu* c, int s, const double* const* f, int nf)
{
    DuctState* d = c->d;
    if (nf < DNF) { snprintf(c->err, sizeof(c->err), "duct mode state has 5 fields"); return LES_ERR_PARAM; }
    if (!d->desc[s].active) {
        d->desc[s].active = 1;
        int e = pool_sync_tables(c);
        if (e != LES_OK) return e;
    }
    for (int a = 0; a < DNF; ++a) {
        int e = slot_xfer(c, s, d->P.A[a], (double*)f[a], true);
        if (e != LES_OK) return e;
    }
    return slot_state_changed(c, s);
}

// initial state of a slot: plug flow at U, planar flame at flame_x0 (host
// computed: runs once per run), Gaussian eddies, cleared statistics
static int slot_init(les_gpu* c, int s, const les_params& p)
{
    DuctState* d = c->d;
    const SlotDesc& D = d->desc[s];
    const size_t n = (size_t)D.g.n;
    std::vector<double> h(5 * n);
    const double U = p.inlet_velocity, Tu = p.inlet_temperature, Tb = p.burnt_temperature;
    for (int i = 0; i < p.nx; ++i) {
        double phi = i * D.g.dx - p.flame_x0;
        double cc = 0.5 * (1.0 + tanh(phi / (2.0 * D.g.dx)));
        double T = Tu + cc * (Tb - Tu);
        double rho = p.rho_unburnt * Tu / T;
        for (size_t q = (size_t)i * p.ny * p.nz; q < (size_t)(i + 1) * p.ny * p.nz; ++q) {
            h[q] = rho * U; h[n + q] = 0.0; h[2 * n + q] = 0.0; h[3 * n + q] = phi; h[4 * n + q] = T;
        }
    }
    SlotDyn y;
    memset(&y, 0, sizeof(y));
    DHIPC(hipMemcpyAsync(d->d_dyn + s, &y, sizeof(y), hipMemcpyHostToDevice, c->stream));
    d->host_step[s] = 0;
    d->point_base[s] = 0;
    d->probe_base[s] = 0;
    const double* f[5] = {&h[0], &h[n], &h[2 * n], &h[3 * n], &h[4 * n]};
    int e = slot_upload(c, s, f, 5);
    if (e != LES_OK) return e;
    if (D.k.nvort > 0) {
        hipLaunchKernelGGL(k_duct_eddy_init, dim3((D.k.nvort + 255) / 256), dim3(256), 0, c->stream, d->P, s);
        DLAUNCH_CHECK();
    }
    DHIPC(hipMemsetAsync(d->P.inlet + D.poff, 0, 8 * 3 * (size_t)D.g.ny * D.g.nz, c->stream));
    DHIPC(hipMemsetAsync(d->P.prof + D.xoff, 0, 8 * 10 * (size_t)D.g.nx, c->stream));
    if (d->P.stat[0])
        for (int a = 0; a < 10; ++a) DHIPC(hipMemsetAsync(d->P.stat[a] + D.off, 0, 8 * n, c->stream));
    DHIPC(hipStreamSynchronize(c->stream));
    return LES_OK;
}

template <typename real>
static int slot_monitor_t(les_gpu* c, int s, les_step_report* rep)
{
    DuctState* d = c->d;
    const SlotDesc& D = d->desc[s];
    int nb = (int)((D.g.n + 255) / 256);
    if (nb > 1024) nb = 1024;
    hipLaunchKernelGGL((k_duct_monitor<real>), dim3(nb), dim3(256), 0, c->stream, d->P, s, d->mon_part);
    DLAUNCH_CHECK();
    std::vector<double> h((size_t)nb * 6);
    SlotDyn y;
    DHIPC(hipMemcpyAsync(h.data(), d->mon_part, h.size() * 8, hipMemcpyDeviceToHost, c->stream));
    DHIPC(hipMemcpyAsync(&y, d->d_dyn + s, sizeof(y), hipMemcpyDeviceToHost, c->stream));
    DHIPC(hipStreamSynchronize(c->stream));
    double tm = 0, mnT = 1e9, mxT = -1e9, mnG = 1e9, mxG = -1e9, bad = 0;
    for (int b = 0; b < nb; ++b) {
        const double* q = &h[6 * (size_t)b];
        tm += q[0]; mnT = fmin(mnT, q[1]); mxT = fmax(mxT, q[2]); mnG = fmin(mnG, q[3]); mxG = fmax(mxG, q[4]);
        bad = fmax(bad, q[5]);
    }
    rep->step = y.step - 1;
    rep->dt = y.dt; rep->max_velocity = y.maxvel;
    rep->total_mass = tm; rep->min_T = mnT; rep->max_T = mxT; rep->min_G = mnG; rep->max_G = mxG;
    rep->min_Y_fuel = 0.0; rep->nonfinite = bad > 0.0;
    rep->proj_residual = y.proj_res; rep->proj_residual0 = y.proj_res0;
    return LES_OK;
}

static int slot_monitor(les_gpu* c, int s, les_step_report* rep)
{
    return c->d->precision == LES_PREC_FP32 ? slot_monitor_t<float>(c, s, rep) : slot_monitor_t<double>(c, s, rep);
}

// ---------------------------------------------------------------------------
// autotuner: times each (tile width, XCD mapping) candidate on a scratch pool
// filled with replicas of the probe run (as many as the pool's slots and cells
// hold, so the timed launches have the pool's occupancy) and keeps the fastest,
// then (pools) times the winner with and without alternating sweep
// directions.  tune_tile / tune_xcd: whether the tile width / the XCD mapping
// is tuned (the pool's own is kept otherwise; the XCD mapping does not change
// results, the tile width only groups the per-tile partial sums of rho);
// sweeps: whether alternating sweeps are tried.
// ---------------------------------------------------------------------------
static int pool_autotune(les_gpu* c, const les_params& probe, bool tune_tile, bool tune_xcd, bool sweeps)
{
    DuctState* d = c->d;
    // one candidate's fastest of three 5-step samples (the candidates differ
    // by a few per cent, about the spread of one short sample) on a scratch
    // pool of replicas of the probe run; rc LES_OK with ms 0 when the scratch
    // pool does not fit beside the real one
    auto time_candidate = [&](int tk, int xmap, float* ms_out) -> int {
        *ms_out = 0.f;
        const long long pc = (long long)probe.nx * probe.ny * probe.nz;
        long long reps = d->pool_cells / pc;
        if (reps > d->max_slots) reps = d->max_slots;
        if (reps < 1) reps = 1;
        les_batch_config cfg;
        memset(&cfg, 0, sizeof(cfg));
        cfg.max_slots = (int)reps;
        cfg.pool_cells = reps * pc;
        cfg.max_nx = probe.nx; cfg.max_ny = probe.ny; cfg.max_nz = probe.nz;
        cfg.max_vortices = probe.num_vortices > 0 ? probe.num_vortices : 1;
        cfg.precision = d->precision;
        cfg.tile_k = tk;
        cfg.xcd_map = xmap;
        cfg.table_t0 = probe.table_t0; cfg.table_dt = probe.table_dt; cfg.table_n = probe.table_n;
        cfg.point_cap = 16;
        les_gpu* t = new (std::nothrow) les_gpu();
        if (!t) return LES_ERR_HIP;
        t->p = probe;
        t->stream = c->stream;
        int e = pool_create(t, cfg);
        if (e != LES_OK) {
            // skipped (the pool keeps its configured launch shape if no
            // candidate fits)
            t->stream = nullptr;
            les_gpu_destroy(t);
            (void)hipGetLastError();
            return LES_OK;
        }
        for (long long r = 0; r < reps && e == LES_OK; ++r) {
            int slot = -1;
            e = pool_admit(t, probe, 0, &slot);
            if (e == LES_OK) e = slot_init(t, slot, probe);
        }
        if (e == LES_OK) e = pool_steps(t, 3);
        float ms = 0.f;
        for (int rep = 0; rep < 3 && e == LES_OK; ++rep) {
            float m = 0.f;
            if (hipEventRecord(c->ev[2], c->stream) != hipSuccess) break;
            e = pool_steps(t, 5);
            (void)hipEventRecord(c->ev[3], c->stream);
            (void)hipEventSynchronize(c->ev[3]);
            (void)hipEventElapsedTime(&m, c->ev[2], c->ev[3]);
            if (m > 0 && (ms == 0.f || m < ms)) ms = m;
        }
        (void)hipStreamSynchronize(c->stream);
        t->stream = nullptr;
        les_gpu_destroy(t);
        if (e != LES_OK) {
            snprintf(c->err, sizeof(c->err), "autotune candidate tile_k=%d xcd_map=%d failed (%d)", tk, xmap, e);
            return e;
        }
        *ms_out = ms;
        return LES_OK;
    };
    int tks[5] = {0, 16, 32, 64, 128};
    if (!tune_tile) tks[0] = d->tk;
    const int zz = d->zigzag ? XCD_REVERSE : 0;
    double best = 1e30;
    int btk = d->tk, bx = d->xcd_map;
    // two passes over the candidates, each kept at its fastest sample: the
    // first candidates of a single pass ran on a GPU not yet at its working
    // clocks and lost to the later ones by more than the candidates differ
    for (int pass = 0; pass < 2; ++pass)
        for (int a = 0; a < (tune_tile ? 5 : 1); ++a)
            for (int x = 0; x < (tune_xcd ? 5 : 1); ++x) {
                const int xm = tune_xcd ? x : d->xcd_map;
                float ms = 0.f;
                const int e = time_candidate(tks[a], xm | zz, &ms);
                if (e != LES_OK) return e;
                if (ms > 0 && ms < best) {
                    best = ms;
                    btk = tks[a];
                    bx = xm;
                }
            }
    int bz = d->zigzag;
    if (sweeps && best < 1e29) {
        // the winner with and without alternating sweep directions, each
        // timed twice in turn
        double bs[2] = {1e30, 1e30};
        for (int pass = 0; pass < 2; ++pass)
            for (int z = 0; z < 2; ++z) {
                float ms = 0.f;
                const int e = time_candidate(btk, bx | (z ? XCD_REVERSE : 0), &ms);
                if (e != LES_OK) return e;
                if (ms > 0 && ms < bs[z]) bs[z] = ms;
            }
        if (bs[0] < 1e29 || bs[1] < 1e29) {
            bz = bs[1] < bs[0] ? 1 : 0;
            best = bs[bz];
        }
    }
    d->tk = btk;
    d->xcd_map = bx;
    d->zigzag = bz;
    pool_set_grid(d);
    d->tuned_ms = best / 5.0;
    d->graph_ok[0] = d->graph_ok[1] = d->graph_ok[2] = false;
    return LES_OK;
}

// ---------------------------------------------------------------------------
// downloads of one slot
// ---------------------------------------------------------------------------
static int slot_download_field(les_gpu* c, int s, const char* name, double* out)
{
    DuctState* d = c->d;
    const SlotDesc& D = d->desc[s];
    struct { const char* nm; void* p; } tab[] = {
        {"u_f", d->P.F[0]}, {"v_f", d->P.F[1]}, {"w_f", d->P.F[2]}, {"T_f", d->P.F[3]}, {"phi_f", d->P.F[4]},
        {"nu_t", d->P.nut}, {"ST", d->P.st},
        {"rhou", d->P.A[0]}, {"rhov", d->P.A[1]}, {"rhow", d->P.A[2]}, {"phi", d->P.A[3]}, {"T", d->P.A[4]}};
    for (auto& t : tab)
        if (strcmp(t.nm, name) == 0) {
            int e = slot_xfer(c, s, t.p, out, false);
            if (e != LES_OK) return e;
            DHIPC(hipStreamSynchronize(c->stream));
            return LES_OK;
        }
    const size_t PL = (size_t)D.g.ny * D.g.nz;
    struct { const char* nm; const double* p; size_t cnt; } dt[] = {
        {"inlet", d->P.inlet + D.poff, 3 * PL}, {"eddies", d->P.eddy + 7 * D.eoff, 7 * (size_t)D.k.nvort},
        {"profiles", d->P.prof + D.xoff, 10 * (size_t)D.g.nx}};
    for (auto& t : dt)
        if (strcmp(t.nm, name) == 0) {
            DHIPC(hipMemcpyAsync(out, t.p, t.cnt * 8, hipMemcpyDeviceToHost, c->stream));
            DHIPC(hipStreamSynchronize(c->stream));
            return LES_OK;
        }
    if (strcmp(name, "dyn") == 0) {          // dt, mean density, max |u|, time of the last step
        SlotDyn y;
        DHIPC(hipMemcpyAsync(&y, d->d_dyn + s, sizeof(y), hipMemcpyDeviceToHost, c->stream));
        DHIPC(hipStreamSynchronize(c->stream));
        out[0] = y.dt; out[1] = y.rho_avg; out[2] = y.maxvel; out[3] = y.time;
        return LES_OK;
    }
    if (strncmp(name, "stat", 4) == 0 && d->P.stat[0]) {
        int a = atoi(name + 4);
        if (a < 0 || a > 9) return LES_ERR_PARAM;
        DHIPC(hipMemcpyAsync(out, d->P.stat[a] + D.off, 8 * (size_t)D.g.n, hipMemcpyDeviceToHost, c->stream));
        DHIPC(hipStreamSynchronize(c->stream));
        return LES_OK;
    }
    snprintf(c->err, sizeof(c->err), "unknown duct field %s", name);
    return LES_ERR_PARAM;
}

static int slot_download(les_gpu* c, int s, double* const* f, int nf)
{
    for (int a = 0; a < nf && a < DNF; ++a) {
        int e = slot_xfer(c, s, c->d->P.A[a], f[a], false);
        if (e != LES_OK) return e;
    }
    DHIPC(hipStreamSynchronize(c->stream));
    return LES_OK;
}

// fixed-point series since the last fetch: out[step][point][u, v, w, G]; the
// ring keeps the newest point_cap steps (the sample of step t sits at t % cap)
static int slot_point_fetch(les_gpu* c, int s, double* out, int max_steps, int* n_steps)
{
    DuctState* d = c->d;
    const int total = d->host_step[s] - d->point_base[s];
    const int avail = total < d->point_cap ? total : d->point_cap;
    const int n = avail < max_steps ? avail : max_steps;
    const int first = d->host_step[s] - avail;
    const double* ring = d->P.points + d->desc[s].ptoff;
    for (int a = 0; a < n; ++a) {
        const int r = (first + a) % d->point_cap;
        const int np = d->desc[s].npts;
        DHIPC(hipMemcpyAsync(out + (size_t)a * np * 4, ring + (size_t)r * DUCT_NPTS * 4, (size_t)np * 4 * 8,
                             hipMemcpyDeviceToHost, c->stream));
    }
    DHIPC(hipStreamSynchronize(c->stream));
    d->point_base[s] = d->host_step[s];
    *n_steps = n;
    return LES_OK;
}

// probe rows recorded since the last drain (the newest probe_cap steps), in
// the row format of the reference's probe CSV (SAMPLING_RANGE 1.0 m, :280)
static int slot_probe_drain(les_gpu* c, int s, const double* probe_x, std::vector<les_probe_row>& out)
{
    DuctState* d = c->d;
    const SlotDesc& D = d->desc[s];
    if (d->probe_cap <= 0 || D.nprobe <= 0) { d->probe_base[s] = d->host_step[s]; return LES_OK; }
    const int total = d->host_step[s] - d->probe_base[s];
    const int avail = total < d->probe_cap ? total : d->probe_cap;
    const int first = d->host_step[s] - avail;
    const size_t stride = 2 + 6 * (size_t)d->ny_max;
    std::vector<double> h((size_t)LES_NPROBE_MAX * stride);
    for (int a = 0; a < avail; ++a) {
        const int step = first + a;
        const double* src = d->P.probe + D.proff + (size_t)(step % d->probe_cap) * LES_NPROBE_MAX * stride;
        DHIPC(hipMemcpyAsync(h.data(), src, h.size() * 8, hipMemcpyDeviceToHost, c->stream));
        DHIPC(hipStreamSynchronize(c->stream));
        for (int n = 0; n < D.nprobe; ++n) {
            const double* e = &h[(size_t)n * stride];
            if (e[0] != 1.0) continue;
            for (int j = 0; j < D.g.ny; ++j) {
                const double* r = e + 2 + 6 * j;
                double yp = j * D.g.dy, yr = yp - e[1];
                if (fabs(yr) > 1.0) continue;
                les_probe_row pr;
                pr.timestep = step; pr.station = n; pr.x_location = probe_x[n];
                pr.y_flame = e[1]; pr.y_physical = yp; pr.y_relative = yr;
                pr.u = r[0]; pr.v = r[1]; pr.w = r[2]; pr.G = r[3]; pr.nu_t = r[4]; pr.S_mag = r[5];
                out.push_back(pr);
            }
        }
    }
    d->probe_base[s] = d->host_step[s];
    return LES_OK;
}

// ---------------------------------------------------------------------------
// single-run duct context: a pool of one slot sized to its grid
// ---------------------------------------------------------------------------
int duct_create(les_gpu* c)
{
    const les_params& p = c->p;
    if (p.integrator != LES_INT_SSPRK3 || (p.advection != LES_ADV_WENO5 && p.advection != LES_ADV_MUSCL_MINMOD))
        return LES_ERR_PARAM;
    les_batch_config cfg;
    les_batch_config_default(&cfg);
    cfg.max_slots = 1;
    cfg.pool_cells = (long long)p.nx * p.ny * p.nz;
    cfg.max_nx = p.nx; cfg.max_ny = p.ny; cfg.max_nz = p.nz;
    cfg.max_vortices = p.num_vortices > 0 ? p.num_vortices : 1;
    cfg.precision = p.precision;
    // the widest tile row that divides nz (no idle lanes), else flat tiles
    cfg.tile_k = (p.nz % 64 == 0) ? 64 : ((p.nz % 32 == 0) ? 32 : ((p.nz % 16 == 0) ? 16 : LES_TILE_FLAT));
    cfg.xcd_map = 1;        // tuned below
    cfg.point_cap = 4096;
    cfg.probe_cap = p.num_probe_x > 0 ? 64 : 0;
    cfg.stats_per_cell = (p.stats_start_step >= 0 && p.stats_per_cell) ? 1 : 0;
    cfg.table_t0 = p.table_t0; cfg.table_dt = p.table_dt; cfg.table_n = p.table_n;
    int e = pool_create(c, cfg);
    if (e != LES_OK) return e;
    // which XCD mapping is faster depends on the grid and precision (FP64 at
    // 128^3 prefers round-robin tiles, smaller grids and FP32 contiguous
    // ranges per XCD); results do not depend on it, so each run times the four
    // before it is admitted (a failed trial, e.g. no memory for the scratch
    // pool, keeps the default), unless LES_XCD_MAP (0..4) fixes it.  The tile
    // width is not timed: it groups the per-tile partial sums of rho, so a
    // timed choice would make a run's rounding depend on the trial; it is the
    // widest row that divides nz, unless LES_TILE_K (0, 16, 32, 64, 128) fixes it
    const char* fx = getenv("LES_XCD_MAP");
    const char* ft = getenv("LES_TILE_K");
    const bool fixed_x = fx && fx[0] >= '0' && fx[0] <= '4' && fx[1] == 0;
    if (ft && ft[0] && tk_valid(atoi(ft))) c->d->tk = atoi(ft);
    if (fixed_x) {
        c->d->xcd_map = fx[0] - '0';
    } else if (pool_autotune(c, p, false, true, false) != LES_OK) {
        c->d->xcd_map = 1;
        c->err[0] = 0;
    }
    int slot = -1;
    e = pool_admit(c, p, 0, &slot);
    if (e != LES_OK) {
        snprintf(c->err, sizeof(c->err), "duct run does not fit its own pool (status %d)", e);
        return e;
    }
    pool_set_grid(c->d);
    c->attr_h.assign((size_t)c->g.n, 3);
    return LES_OK;
}

int duct_launch_config(const les_gpu* c, int* tile_k, int* xcd_map)
{
    if (tile_k) *tile_k = c->d->tk;
    if (xcd_map) *xcd_map = c->d->xcd_map;
    return LES_OK;
}

void duct_destroy(les_gpu* c)
{
    if (!c->d) return;
    for (auto& g : c->d->gexec)
        if (g) (void)hipGraphExecDestroy(g);
    if (c->d->aux) {
        (void)hipStreamSynchronize(c->d->aux);
        (void)hipStreamDestroy(c->d->aux);
    }
    for (int a = 0; a < 1; ++a) {
        if (c->d->evf[a]) (void)hipEventDestroy(c->d->evf[a]);
        if (c->d->evj[a]) (void)hipEventDestroy(c->d->evj[a]);
    }
    delete c->d;
    c->d = nullptr;
}

void duct_params_changed(les_gpu* c)
{
    if (!c->d) return;
    c->d->desc[0].k = make_duct_consts(c->p);
    (void)pool_sync_tables(c);
}

int duct_init_state(les_gpu* c)
{
    int e = slot_init(c, 0, c->p);
    if (e != LES_OK) return e;
    c->step = 0;
    c->stat_count = 0;
    c->probe.clear();
    return LES_OK;
}

int duct_upload_state(les_gpu* c, const double* const* f, int nf) { return slot_upload(c, 0, f, nf); }
int duct_download_state(les_gpu* c, double* const* f, int nf) { return slot_download(c, 0, f, nf); }
int duct_download_field(les_gpu* c, const char* name, double* out) { return slot_download_field(c, 0, name, out); }
int duct_point_fetch(les_gpu* c, double* out, int max_steps, int* n_steps)
{
    return slot_point_fetch(c, 0, out, max_steps, n_steps);
}

int duct_npoints(const les_gpu* c) { return c->d->desc[0].npts; }

int duct_set_points(les_gpu* c, const int* cells, int n)
{
    SlotDesc& D = c->d->desc[0];
    for (int a = 0; a < n; ++a) D.point_idx[a] = cells[a];
    D.npts = n;
    c->d->point_base[0] = c->d->host_step[0];
    return pool_sync_tables(c);
}

int duct_step(les_gpu* c, int n_steps, les_step_report* rep)
{
    DuctState* d = c->d;
    const int st0 = c->p.stats_start_step;
    for (int a = 0; a < n_steps; ++a)
        if (st0 >= 0 && d->host_step[0] + a >= st0) c->stat_count++;
    // probe rows are drained into the context's row list at least every
    // probe_cap steps (the ring holds that many); the step time excludes the
    // drains
    const int chunk = d->probe_cap > 0 ? d->probe_cap : (n_steps > 0 ? n_steps : 1);
    double ms_sum = 0.0;
    for (int done = 0; done < n_steps;) {
        const int n = n_steps - done < chunk ? n_steps - done : chunk;
        DHIPC(hipEventRecord(c->ev[2], c->stream));
        int e = pool_steps(c, n);
        if (e != LES_OK) return e;
        DHIPC(hipEventRecord(c->ev[3], c->stream));
        DHIPC(hipEventSynchronize(c->ev[3]));
        float ms = 0.f;
        DHIPC(hipEventElapsedTime(&ms, c->ev[2], c->ev[3]));
        ms_sum += ms;
        done += n;
        if (d->probe_cap > 0) {
            e = slot_probe_drain(c, 0, c->p.probe_x, c->probe);
            if (e != LES_OK) return e;
        }
    }
    c->step_ms = n_steps > 0 ? ms_sum / n_steps : 0.0;
    c->step += n_steps;
    int e = LES_OK;
    if (rep) {
        e = slot_monitor(c, 0, rep);
        if (e != LES_OK) return e;
        rep->step = c->step - 1;
        if (rep->nonfinite) {
            snprintf(c->err, sizeof(c->err), "non-finite value in the state at step %d", rep->step);
            return LES_ERR_NAN;
        }
    }
    return LES_OK;
}

// Testing hook: the RHS L(q) of the current state (a stage-0 evaluation whose
// combined output goes to the dead stage buffers B).
int duct_rhs(les_gpu* c, double* const* out, int n_out)
{
    if (n_out < DNF) return LES_ERR_PARAM;
    DuctState* d = c->d;
    const SlotDesc& D = d->desc[0];
    const size_t n = (size_t)D.g.n;
    std::vector<void*> Lbuf(DNF, nullptr);
    for (int f = 0; f < DNF; ++f) {
        DHIPC(hipMalloc(&Lbuf[f], n * d->esize));
        DHIPC(hipMemsetAsync(Lbuf[f], 0, n * d->esize, c->stream));   // boundary planes read as 0
    }
    int e = tk_dispatch(d->tk, [&](auto K) {
        constexpr int TK = decltype(K)::value;
        if (d->precision == LES_PREC_FP32) return stage_launch<float, TK>(c, 0, (float* const*)Lbuf.data());
        return stage_launch<double, TK>(c, 0, (double* const*)Lbuf.data());
    });
    // slot 0 starts at cell 0 of the arenas, so slot-local buffers xfer as slot 0
    for (int f = 0; f < DNF && e == LES_OK; ++f) e = slot_xfer(c, 0, (char*)Lbuf[f] - (size_t)D.off * d->esize, out[f], false);
    DHIPC(hipStreamSynchronize(c->stream));
    for (int f = 0; f < DNF; ++f) (void)hipFree(Lbuf[f]);
    return e;
}

// ---------------------------------------------------------------------------
// The same RHS hook without device allocation.  The five RHS fields are staged
// in `scratch`, a file-scope device array of LES_RHS_NOALLOC_MAX_BYTES that
// each GPU holds from module load to process end.  In host code the array's
// name is only a host-side stand-in, so its device address is looked up once
// per device (hipGetSymbolAddress with that device current) and only that
// address is handed to kernels and copies.  Every call on one GPU shares the
// array: the staged fields are cleared by a fill kernel at the start of each
// call, and calls that overlap on one GPU race.
// ---------------------------------------------------------------------------
static __device__ __align__(256) unsigned char scratch[LES_RHS_NOALLOC_MAX_BYTES];

static constexpr int kScratchDevices = 64;
static void* g_scratch_dev[kScratchDevices];   // device address of `scratch` per device
static std::mutex g_scratch_mu;

// Device address of `scratch` on the device of c's stream.
static int scratch_address(les_gpu* c, unsigned char** out)
{
    int dev = -1;
    DHIPC(hipStreamGetDevice(c->stream, &dev));
    if (dev < 0 || dev >= kScratchDevices) {
        snprintf(c->err, sizeof(c->err), "device %d outside the %d devices the RHS scratch tracks", dev, kScratchDevices);
        return LES_ERR_STATE;
    }
    std::lock_guard<std::mutex> lock(g_scratch_mu);
    if (!g_scratch_dev[dev]) {
        int cur = -1;
        DHIPC(hipGetDevice(&cur));
        if (cur != dev) DHIPC(hipSetDevice(dev));
        void* p = nullptr;
        hipError_t e = hipGetSymbolAddress(&p, HIP_SYMBOL(scratch));
        if (cur != dev) (void)hipSetDevice(cur);
        if (e != hipSuccess || !p) {
            snprintf(c->err, sizeof(c->err), "hipGetSymbolAddress(scratch) failed on device %d: %s", dev,
                     hipGetErrorString(e));
            return LES_ERR_HIP;
        }
        g_scratch_dev[dev] = p;
    }
    *out = (unsigned char*)g_scratch_dev[dev];
    return LES_OK;
}

__global__ void k_zero_words(long long n_words, unsigned int* dst)
{
    for (long long q = (long long)blockIdx.x * blockDim.x + threadIdx.x; q < n_words;
         q += (long long)gridDim.x * blockDim.x)
        dst[q] = 0u;
}

static size_t align256(size_t b) { return (b + 255) & ~(size_t)255; }

// Bytes of scratch the duct RHS of this context stages (five fields, each
// starting on a 256-byte boundary).
static size_t duct_rhs_scratch_bytes(const les_gpu* c)
{
    const DuctState* d = c->d;
    return DNF * align256((size_t)d->desc[0].g.n * d->esize);
}

int duct_rhs_noalloc(les_gpu* c, double* const* out, int n_out)
{
    if (n_out < DNF) return LES_ERR_PARAM;
    DuctState* d = c->d;
    const SlotDesc& D = d->desc[0];
    const size_t n = (size_t)D.g.n;
    const size_t need = duct_rhs_scratch_bytes(c);
    if (need > (size_t)LES_RHS_NOALLOC_MAX_BYTES) {
        snprintf(c->err, sizeof(c->err),
                 "les_gpu_rhs_noalloc: a %zu-cell grid stages %zu bytes, above the %zu-byte static scratch "
                 "(use les_gpu_rhs)", n, need, (size_t)LES_RHS_NOALLOC_MAX_BYTES);
        return LES_ERR_PARAM;
    }
    unsigned char* base = nullptr;
    int e = scratch_address(c, &base);
    if (e != LES_OK) return e;
    // boundary planes are not written by the RHS kernels and read as 0
    const long long words = (long long)(need / 4);
    const int zb = (int)std::min<long long>((words + 255) / 256, 4096);
    hipLaunchKernelGGL(k_zero_words, dim3(zb), dim3(256), 0, c->stream, words, (unsigned int*)base);
    DLAUNCH_CHECK();
    void* Lbuf[DNF];
    for (int f = 0; f < DNF; ++f) Lbuf[f] = base + (size_t)f * align256(n * d->esize);
    e = tk_dispatch(d->tk, [&](auto K) {
        constexpr int TK = decltype(K)::value;
        if (d->precision == LES_PREC_FP32) return stage_launch<float, TK>(c, 0, (float* const*)Lbuf);
        return stage_launch<double, TK>(c, 0, (double* const*)Lbuf);
    });
    for (int f = 0; f < DNF && e == LES_OK; ++f) e = slot_xfer(c, 0, (char*)Lbuf[f] - (size_t)D.off * d->esize, out[f], false);
    DHIPC(hipStreamSynchronize(c->stream));
    return e;
}

// ---------------------------------------------------------------------------
// batch ABI
// ---------------------------------------------------------------------------
struct les_batch {
    les_gpu* c = nullptr;
};

static int batch_config_check(const les_batch_config* cfg)
{
    if (!cfg || cfg->max_slots < 1 || cfg->max_slots > 8192 || cfg->pool_cells < 1 || cfg->max_nx < 6 ||
        cfg->max_ny < 1 || cfg->max_nz < 1 || cfg->max_vortices < 0 || !tk_valid(cfg->tile_k) ||
        (cfg->precision != LES_PREC_FP64 && cfg->precision != LES_PREC_FP32) || cfg->point_cap < 1 ||
        cfg->table_n < 2 || cfg->table_dt <= 0.0 || cfg->xcd_map < 0 || (cfg->xcd_map & ~15) || (cfg->xcd_map & 7) > 4)
        return LES_ERR_PARAM;
    return LES_OK;
}

extern "C" int les_batch_pool_bytes(const les_batch_config* cfg, size_t* bytes)
{
    return abi_guard(nullptr, [&]() -> int {
        if (batch_config_check(cfg) != LES_OK || !bytes) return LES_ERR_PARAM;
        const size_t es = cfg->precision == LES_PREC_FP32 ? 4 : 8;
        const size_t n = (size_t)cfg->pool_cells, S = (size_t)cfg->max_slots;
        const size_t tiles = n / 256 + S * cfg->max_nx * (cfg->max_ny / 2 + cfg->max_nz / 16 + 2) + 64;
        size_t b = 17 * n * es;                                   // A, B, F (5 each), nu_t, S_T
        if (cfg->stats_per_cell) b += 80 * n;
        if (cfg->precision == LES_PREC_FP32) b += 8 * n;           // FP64 staging of one run
        b += tiles * (8 + 8 + 48 + 16 + 4);                       // rho, |u|, nu/alpha, residual partials, tile list
        const size_t nv = cfg->max_vortices > 0 ? cfg->max_vortices : 1;
        b += S * 8 * (7 * nv + 3 * (size_t)cfg->max_ny * cfg->max_nz + 10 * (size_t)cfg->max_nx +
                      4 * (size_t)DUCT_NPTS * cfg->point_cap + (size_t)cfg->max_ny * ((cfg->max_nz + 255) / 256) +
                      10 * (size_t)cfg->max_nx * STAT_CHUNKS);
        b += S * (sizeof(SlotDesc) + sizeof(SlotDyn) + 8 + 24 * DT_BLOCKS_MAX) + 8 * (size_t)cfg->table_n + 8 * 6 * 1024;
        b += S * 8 * 3 * (size_t)cfg->max_ny * cfg->max_nz * ((nv + SEM_CHUNK - 1) / SEM_CHUNK);   // SEM partials
        if (cfg->probe_cap > 0)                                   // flame-region probe rings
            b += S * 8 * (size_t)cfg->probe_cap * LES_NPROBE_MAX * (2 + 6 * (size_t)cfg->max_ny);
        *bytes = b;
        return LES_OK;
    });
}

extern "C" int les_batch_create(const les_batch_config* cfg, const les_params* probe, les_batch** out)
{
    return abi_guard(nullptr, [&]() -> int {
        if (!out) return LES_ERR_PARAM;
        *out = nullptr;
        if (batch_config_check(cfg) != LES_OK) return LES_ERR_PARAM;
        if (cfg->autotune && (!probe || les_params_validate(probe) != LES_OK || probe->mode != LES_MODE_DUCT))
            return LES_ERR_PARAM;
        les_batch* b = new (std::nothrow) les_batch();
        if (!b) return LES_ERR_HIP;
        les_gpu* c = new (std::nothrow) les_gpu();
        if (!c) { delete b; return LES_ERR_HIP; }
        b->c = c;
        if (probe) c->p = *probe;
        else les_params_default(&c->p, LES_PRESET_ENSEMBLE);
        c->p.precision = cfg->precision;
        les_mixture_synthetic(&c->mix);
        int rc = LES_OK;
        if (hipStreamCreateWithFlags(&c->stream, hipStreamNonBlocking) != hipSuccess) rc = LES_ERR_HIP;
        for (int a = 0; a < 4 && rc == LES_OK; ++a)
            if (hipEventCreate(&c->ev[a]) != hipSuccess) rc = LES_ERR_HIP;
        if (rc == LES_OK) rc = pool_create(c, *cfg);
        if (rc == LES_OK && cfg->autotune) rc = pool_autotune(c, *probe, true, true, true);
        if (rc != LES_OK) {
            fprintf(stderr, "les_batch_create: %s\n", c->err);
            les_gpu_destroy(c);
            delete b;
            return rc;
        }
        *out = b;
        return LES_OK;
    });
}

extern "C" void les_batch_destroy(les_batch* b)
{
    if (!b) return;
    les_gpu_destroy(b->c);
    delete b;
}

extern "C" const char* les_batch_last_error(const les_batch* b) { return b && b->c ? b->c->err : "null context"; }

extern "C" int les_batch_tuning(const les_batch* b, int* tile_k, int* grid_blocks, int* xcd_map, double* tuned_ms)
{
    return abi_guard(b ? b->c : nullptr, [&]() -> int {
        if (!b) return LES_ERR_PARAM;
        const DuctState* d = b->c->d;
        if (tile_k) *tile_k = d->tk;
        if (grid_blocks) *grid_blocks = d->grid;
        if (xcd_map) *xcd_map = d->xcd_map | (d->zigzag ? XCD_REVERSE : 0);
        if (tuned_ms) *tuned_ms = d->tuned_ms;
        return LES_OK;
    });
}

extern "C" int les_batch_groups(const les_batch* b, int cap, int* ngroups, int* tile_base, int* tile_count)
{
    return abi_guard(b ? b->c : nullptr, [&]() -> int {
        if (!b || !ngroups || cap < 0 || (cap > 0 && (!tile_base || !tile_count))) return LES_ERR_PARAM;
        const DuctState* d = b->c->d;
        *ngroups = d->ngroups;
        for (int g = 0; g < std::min(cap, d->ngroups); ++g) {
            tile_base[g] = d->ngroups > 1 ? d->gtb[g] : 0;
            tile_count[g] = d->ngroups > 1 ? d->gtn[g] : d->ntile_hwm;
        }
        return LES_OK;
    });
}

static int batch_slot_ok(les_batch* b, int slot)
{
    if (!b || slot < 0 || slot >= b->c->d->max_slots || !b->c->d->taken[slot]) {
        if (b) snprintf(b->c->err, sizeof(b->c->err), "slot %d is not admitted", slot);
        return 0;
    }
    return 1;
}

extern "C" int les_batch_admit(les_batch* b, const les_params* p, int n_steps, int init, int* slot)
{
    return abi_guard(b ? b->c : nullptr, [&]() -> int {
        if (!b || !p || !slot || n_steps < 0) return LES_ERR_PARAM;
        les_gpu* c = b->c;
        *slot = -1;
        if (les_params_validate(p) != LES_OK || p->mode != LES_MODE_DUCT) {
            snprintf(c->err, sizeof(c->err), "batched runs are duct-mode runs with valid parameters");
            return LES_ERR_PARAM;
        }
        const DuctState* d = c->d;
        if (p->table_t0 != c->p.table_t0 || p->table_dt != c->p.table_dt || p->table_n != c->p.table_n ||
            (p->stats_per_cell && !d->stats_cells)) {
            snprintf(c->err, sizeof(c->err), "run needs a property table or per-cell statistics the pool lacks");
            return LES_ERR_ADMISSION;
        }
        int e = pool_admit(c, *p, n_steps, slot);
        if (e == LES_ERR_ADMISSION) {
            snprintf(c->err, sizeof(c->err), "no free slot, cells or tiles for a %dx%dx%d run", p->nx, p->ny, p->nz);
            return e;
        }
        if (e != LES_OK) return e;
        if (init) {
            e = slot_init(c, *slot, *p);
            if (e != LES_OK) { (void)pool_release(c, *slot); *slot = -1; return e; }
        }
        return LES_OK;
    });
}

extern "C" int les_batch_release(les_batch* b, int slot)
{
    return abi_guard(b ? b->c : nullptr, [&]() -> int {
        if (!batch_slot_ok(b, slot)) return LES_ERR_PARAM;
        return pool_release(b->c, slot);
    });
}

extern "C" int les_batch_step(les_batch* b, int n_steps, double* ms_per_step)
{
    return abi_guard(b ? b->c : nullptr, [&]() -> int {
        if (!b || n_steps < 0) return LES_ERR_PARAM;
        les_gpu* c = b->c;
        DuctState* d = c->d;
        // each run's probe ring holds probe_cap steps: drain every resident
        // run's ring into its row list at least that often (the step time
        // excludes the drains)
        const int chunk = d->probe_cap > 0 ? d->probe_cap : (n_steps > 0 ? n_steps : 1);
        double ms_sum = 0.0;
        for (int done = 0; done < n_steps;) {
            const int n = n_steps - done < chunk ? n_steps - done : chunk;
            DHIPC(hipEventRecord(c->ev[2], c->stream));
            int e = pool_steps(c, n);
            if (e != LES_OK) return e;
            DHIPC(hipEventRecord(c->ev[3], c->stream));
            DHIPC(hipEventSynchronize(c->ev[3]));
            float ms = 0.f;
            DHIPC(hipEventElapsedTime(&ms, c->ev[2], c->ev[3]));
            ms_sum += ms;
            done += n;
            if (d->probe_cap > 0 && done < n_steps)
                for (int s = 0; s < d->max_slots; ++s)
                    if (d->desc[s].active) {
                        e = slot_probe_drain(c, s, d->probe_x[s].data(), d->probe_rows[s]);
                        if (e != LES_OK) return e;
                    }
        }
        c->step_ms = n_steps > 0 ? ms_sum / n_steps : 0.0;
        if (ms_per_step) *ms_per_step = c->step_ms;
        return LES_OK;
    });
}

extern "C" int les_batch_slot_info(les_batch* b, int slot, les_slot_info* info)
{
    return abi_guard(b ? b->c : nullptr, [&]() -> int {
        if (!b || !info || slot < 0 || slot >= b->c->d->max_slots) return LES_ERR_PARAM;
        les_gpu* c = b->c;
        DuctState* d = c->d;
        memset(info, 0, sizeof(*info));
        if (!d->taken[slot]) return LES_OK;
        const SlotDesc& D = d->desc[slot];
        SlotDyn y;
        DHIPC(hipMemcpyAsync(&y, d->d_dyn + slot, sizeof(y), hipMemcpyDeviceToHost, c->stream));
        DHIPC(hipStreamSynchronize(c->stream));
        info->active = D.active;
        info->done = y.done;
        info->step = y.step;
        info->n_steps = D.n_steps;
        info->time = y.time;
        info->dt = y.dt;
        info->max_velocity = y.maxvel;
        info->nx = D.g.nx; info->ny = D.g.ny; info->nz = D.g.nz;
        info->cells = D.g.n;
        return LES_OK;
    });
}

extern "C" int les_batch_poll(les_batch* b, int* slots, int max_slots, int* n_done)
{
    return abi_guard(b ? b->c : nullptr, [&]() -> int {
        if (!b || !n_done || (max_slots > 0 && !slots)) return LES_ERR_PARAM;
        const DuctState* d = b->c->d;
        int n = 0;
        for (int s = 0; s < d->max_slots; ++s) {
            const SlotDesc& D = d->desc[s];
            if (d->taken[s] && D.active && D.n_steps > 0 && d->host_step[s] >= D.n_steps) {
                if (n < max_slots) slots[n] = s;
                ++n;
            }
        }
        *n_done = n;
        return LES_OK;
    });
}

extern "C" int les_batch_upload(les_batch* b, int slot, const double* const* fields, int n_fields)
{
    return abi_guard(b ? b->c : nullptr, [&]() -> int {
        if (!batch_slot_ok(b, slot) || !fields) return LES_ERR_PARAM;
        return slot_upload(b->c, slot, fields, n_fields);
    });
}

extern "C" int les_batch_download(les_batch* b, int slot, double* const* fields, int n_fields)
{
    return abi_guard(b ? b->c : nullptr, [&]() -> int {
        if (!batch_slot_ok(b, slot) || !fields) return LES_ERR_PARAM;
        return slot_download(b->c, slot, fields, n_fields);
    });
}

extern "C" int les_batch_download_field(les_batch* b, int slot, const char* name, double* out)
{
    return abi_guard(b ? b->c : nullptr, [&]() -> int {
        if (!batch_slot_ok(b, slot) || !name || !out) return LES_ERR_PARAM;
        return slot_download_field(b->c, slot, name, out);
    });
}

extern "C" int les_batch_report(les_batch* b, int slot, les_step_report* rep)
{
    return abi_guard(b ? b->c : nullptr, [&]() -> int {
        if (!batch_slot_ok(b, slot) || !rep) return LES_ERR_PARAM;
        return slot_monitor(b->c, slot, rep);
    });
}

extern "C" int les_batch_probe_fetch(les_batch* b, int slot, les_probe_row* rows, int max_rows, int* n_rows)
{
    return abi_guard(b ? b->c : nullptr, [&]() -> int {
        if (!batch_slot_ok(b, slot) || !rows || !n_rows || max_rows < 0) return LES_ERR_PARAM;
        les_gpu* c = b->c;
        DuctState* d = c->d;
        std::vector<les_probe_row>& v = d->probe_rows[slot];
        int e = slot_probe_drain(c, slot, d->probe_x[slot].data(), v);
        if (e != LES_OK) return e;
        const int n = (int)v.size() < max_rows ? (int)v.size() : max_rows;
        if (n > 0) memcpy(rows, v.data(), (size_t)n * sizeof(les_probe_row));
        v.erase(v.begin(), v.begin() + n);
        *n_rows = n;
        return LES_OK;
    });
}

extern "C" int les_batch_point_fetch(les_batch* b, int slot, double* out, int max_steps, int* n_steps)
{
    return abi_guard(b ? b->c : nullptr, [&]() -> int {
        if (!batch_slot_ok(b, slot) || !out || !n_steps || max_steps < 0) return LES_ERR_PARAM;
        return slot_point_fetch(b->c, slot, out, max_steps, n_steps);
    });
}

extern "C" int les_batch_xface_copy(les_batch* b, int src_slot, int src_plane, int dst_slot, int dst_plane,
                                    int planes, int reps, double* ms_per_copy)
{
    return abi_guard(b ? b->c : nullptr, [&]() -> int {
        if (!batch_slot_ok(b, src_slot) || !batch_slot_ok(b, dst_slot)) return LES_ERR_PARAM;
        les_gpu* c = b->c;
        DuctState* d = c->d;
        const GridDims& gs = d->desc[src_slot].g;
        const GridDims& gd = d->desc[dst_slot].g;
        const bool overlap = src_slot == dst_slot && src_plane < dst_plane + planes && dst_plane < src_plane + planes;
        if (planes < 1 || reps < 1 || gs.ny != gd.ny || gs.nz != gd.nz || src_plane < 0 || dst_plane < 0 ||
            src_plane + planes > gs.nx || dst_plane + planes > gd.nx || overlap) {
            snprintf(c->err, sizeof(c->err), "x-face copy of %d planes from %d of slot %d (%dx%dx%d) to %d of slot %d "
                     "(%dx%dx%d) is out of range or overlaps", planes, src_plane, src_slot, gs.nx, gs.ny, gs.nz,
                     dst_plane, dst_slot, gd.nx, gd.ny, gd.nz);
            return LES_ERR_PARAM;
        }
        const long long plane = (long long)gs.ny * gs.nz, wpe = (long long)d->esize / 4;
        const long long sw = (d->desc[src_slot].off + src_plane * plane) * wpe;
        const long long dw = (d->desc[dst_slot].off + dst_plane * plane) * wpe;
        const long long words = planes * plane * wpe;
        const long long nb = (words + 255) / 256;
        const int blocks = (int)(nb < 2048 ? nb : 2048);
        DHIPC(hipEventRecord(c->ev[2], c->stream));
        for (int r = 0; r < reps; ++r)
            hipLaunchKernelGGL(k_duct_xface_copy, dim3(blocks), dim3(256), 0, c->stream, d->P, sw, dw, words);
        DLAUNCH_CHECK();
        DHIPC(hipEventRecord(c->ev[3], c->stream));
        DHIPC(hipEventSynchronize(c->ev[3]));
        float ms = 0.f;
        DHIPC(hipEventElapsedTime(&ms, c->ev[2], c->ev[3]));
        if (ms_per_copy) *ms_per_copy = ms / reps;
        return slot_state_changed(c, dst_slot);
    });
}

// ---------------------------------------------------------------------------
// resume support of a single-run duct context (slot 0)
// ---------------------------------------------------------------------------
int duct_upload_field(les_gpu* c, const char* name, const double* data)
{
    DuctState* d = c->d;
    const SlotDesc& D = d->desc[0];
    if (strcmp(name, "profiles") == 0) {
        DHIPC(hipMemcpyAsync(d->P.prof + D.xoff, data, 8 * 10 * (size_t)D.g.nx, hipMemcpyHostToDevice, c->stream));
    } else if (strcmp(name, "eddies") == 0) {
        DHIPC(hipMemcpyAsync(d->P.eddy + 7 * D.eoff, data, 8 * 7 * (size_t)D.k.nvort, hipMemcpyHostToDevice, c->stream));
    } else if (strncmp(name, "stat", 4) == 0 && d->P.stat[0]) {
        int a = atoi(name + 4);
        if (a < 0 || a > 9) return LES_ERR_PARAM;
        DHIPC(hipMemcpyAsync(d->P.stat[a] + D.off, data, 8 * (size_t)D.g.n, hipMemcpyHostToDevice, c->stream));
    } else {
        snprintf(c->err, sizeof(c->err), "duct field %s cannot be uploaded", name);
        return LES_ERR_PARAM;
    }
    DHIPC(hipStreamSynchronize(c->stream));
    return LES_OK;
}

struct AuxDuct {
    char magic[8];
    int mode, step, stat_count, nvort, nx, pad;
    double time, max_nu, max_alpha;
};

static const char DUCT_AUX_MAGIC[8] = {'L', 'E', 'S', 'A', 'U', 'X', '1', 0};

int duct_export_aux(les_gpu* c, void* buf, size_t* bytes)
{
    DuctState* d = c->d;
    const SlotDesc& D = d->desc[0];
    const size_t ne = 7 * (size_t)D.k.nvort, np = 10 * (size_t)D.g.nx;
    const size_t need = sizeof(AuxDuct) + (ne + np) * 8;
    if (!buf || *bytes < need) { *bytes = need; return buf ? LES_ERR_PARAM : LES_OK; }
    AuxDuct a;
    memset(&a, 0, sizeof(a));
    memcpy(a.magic, DUCT_AUX_MAGIC, 8);
    a.mode = LES_MODE_DUCT; a.step = c->step; a.stat_count = c->stat_count; a.nvort = D.k.nvort; a.nx = D.g.nx;
    SlotDyn y;
    std::vector<double> dp(6 * (size_t)D.ntiles);
    DHIPC(hipMemcpyAsync(&y, d->d_dyn, sizeof(y), hipMemcpyDeviceToHost, c->stream));
    for (int stg = 0; stg < 3; ++stg)
        DHIPC(hipMemcpyAsync(dp.data() + 2 * (size_t)stg * D.ntiles, d->P.dpart + 2 * ((size_t)stg * d->tile_cap + D.tile0),
                             16 * (size_t)D.ntiles, hipMemcpyDeviceToHost, c->stream));
    char* o = (char*)buf + sizeof(a);
    DHIPC(hipMemcpyAsync(o, d->P.eddy + 7 * D.eoff, ne * 8, hipMemcpyDeviceToHost, c->stream));
    DHIPC(hipMemcpyAsync(o + ne * 8, d->P.prof + D.xoff, np * 8, hipMemcpyDeviceToHost, c->stream));
    DHIPC(hipStreamSynchronize(c->stream));
    // the diffusive DT history: maxima of nu and alpha over the last step's stages
    double mn = 0.0, ma = 0.0;
    for (size_t t = 0; t < 3 * (size_t)D.ntiles; ++t) { mn = fmax(mn, dp[2 * t]); ma = fmax(ma, dp[2 * t + 1]); }
    a.time = y.time; a.max_nu = mn; a.max_alpha = ma;
    memcpy(buf, &a, sizeof(a));
    *bytes = need;
    return LES_OK;
}

// after les_gpu_upload_state (which recomputes the velocity maxima of the new
// state and clears the diffusive history): restores step, run time, eddies,
// profiles and the nu/alpha maxima the next DT needs
int duct_import_aux(les_gpu* c, const void* buf, size_t bytes)
{
    DuctState* d = c->d;
    const SlotDesc& D = d->desc[0];
    AuxDuct a;
    if (bytes < sizeof(a)) { snprintf(c->err, sizeof(c->err), "aux blob too short"); return LES_ERR_PARAM; }
    memcpy(&a, buf, sizeof(a));
    const size_t ne = 7 * (size_t)D.k.nvort, np = 10 * (size_t)D.g.nx;
    if (memcmp(a.magic, DUCT_AUX_MAGIC, 8) != 0 || a.mode != LES_MODE_DUCT || a.nvort != D.k.nvort || a.nx != D.g.nx ||
        bytes != sizeof(a) + (ne + np) * 8 || a.step < 0) {
        snprintf(c->err, sizeof(c->err), "aux blob does not match this context");
        return LES_ERR_PARAM;
    }
    const char* in = (const char*)buf + sizeof(a);
    DHIPC(hipMemcpyAsync(d->P.eddy + 7 * D.eoff, in, ne * 8, hipMemcpyHostToDevice, c->stream));
    DHIPC(hipMemcpyAsync(d->P.prof + D.xoff, in + ne * 8, np * 8, hipMemcpyHostToDevice, c->stream));
    SlotDyn y;
    DHIPC(hipMemcpyAsync(&y, d->d_dyn, sizeof(y), hipMemcpyDeviceToHost, c->stream));
    DHIPC(hipStreamSynchronize(c->stream));
    y.step = a.step;
    y.time = a.time;
    y.done = 0;
    const double h[2] = {a.max_nu, a.max_alpha};
    DHIPC(hipMemcpyAsync(d->d_dyn, &y, sizeof(y), hipMemcpyHostToDevice, c->stream));
    DHIPC(hipMemcpyAsync(d->P.dpart + 2 * (size_t)D.tile0, h, 16, hipMemcpyHostToDevice, c->stream));
    DHIPC(hipStreamSynchronize(c->stream));
    d->host_step[0] = a.step;
    d->point_base[0] = a.step;
    d->probe_base[0] = a.step;
    c->step = a.step;
    c->stat_count = a.stat_count;
    c->probe.clear();
    return LES_OK;
}

int duct_set_step(les_gpu* c, int step)
{
    DuctState* d = c->d;
    SlotDyn y;
    DHIPC(hipMemcpyAsync(&y, d->d_dyn, sizeof(y), hipMemcpyDeviceToHost, c->stream));
    DHIPC(hipStreamSynchronize(c->stream));
    y.step = step;
    y.done = 0;
    DHIPC(hipMemcpyAsync(d->d_dyn, &y, sizeof(y), hipMemcpyHostToDevice, c->stream));
    DHIPC(hipStreamSynchronize(c->stream));
    d->host_step[0] = d->point_base[0] = d->probe_base[0] = step;
    c->step = step;
    return LES_OK;
}

int duct_init_sem(les_gpu* c)
{
    DuctState* d = c->d;
    const SlotDesc& D = d->desc[0];
    if (D.k.nvort > 0) {
        hipLaunchKernelGGL(k_duct_eddy_init, dim3((D.k.nvort + 255) / 256), dim3(256), 0, c->stream, d->P, 0);
        DLAUNCH_CHECK();
    }
    DHIPC(hipStreamSynchronize(c->stream));
    return LES_OK;
}
