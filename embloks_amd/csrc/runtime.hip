// runtime.hip - device handle, resident scenes and the single-scene C ABI
// (the GpuSim replacement, crates/embloks-gpu/src/sim.rs:180-850).

#include "runtime.hpp"

#include "gemm.hpp"

#include <algorithm>
#include <cmath>
#include <cstdlib>
#include <cstring>

// Partial sums of emb_sim_diagnostics_static, one array per GPU for the life of
// the process: DIAG_BLOCKS (residual max, energy) pairs from k_diagnostics, then
// the two reduced values. Host code only ever uses its device address from
// hipGetSymbolAddress (EmbDevice::scratch), never this name.
constexpr int DIAG_BLOCKS = 1024;
static __device__ __align__(256) double scratch[2 * DIAG_BLOCKS + 2];

// The block partials reduced in block order, as emb_sim_diagnostics reduces them
// on the host: residual max and energy sum into p[2 * blocks], p[2 * blocks + 1].
__global__ void k_diag_reduce(double *p, int blocks) {
    if (threadIdx.x != 0 || blockIdx.x != 0) return;
    double res = 0.0, en = 0.0;
    for (int b = 0; b < blocks; b++) {
        res = fmax(res, p[2 * b]);
        en += p[2 * b + 1];
    }
    p[2 * blocks] = res;
    p[2 * blocks + 1] = en;
}

namespace emb {

static thread_local std::string g_error;

void set_error(const std::string &msg) { g_error = msg; }
int fail(int code, const std::string &msg) {
    g_error = msg;
    return code;
}

namespace {

constexpr double PI = 3.14159265358979323846;

struct Bump {
    char *base;
    size_t off = 0;
    template <class T> T *take(size_t n) {
        off = (off + 255) & ~(size_t)255;
        T *p = base ? reinterpret_cast<T *>(base + off) : nullptr;
        off += std::max<size_t>(n, 1) * sizeof(T);
        return p;
    }
};

struct Counts {
    int cols, rows, nx, ny;
    int64_t n;
    int n_stacks, n_stack_cells, n_bats, n_ac;
    int max_substeps;
};

// Every per-scene array in one arena (sizes follow the reference's regions,
// sim.rs:230-297, 449-466).
void layout(Bump &a, const Counts &c, DevScene &d, DevScene **d_scene, int32_t **d_prefix, SceneCounters **ctr,
            float **tab) {
    const size_t nx = c.nx, ny = c.ny, nodes = nx * ny, cells = (size_t)c.cols * c.rows, n = (size_t)c.n;
    const size_t pnodes = nx * (size_t)poisson_pitch(c.ny);   // the Poisson arrays' rows are padded
    d.ex = a.take<float>((nx + 1) * ny);
    d.ey = a.take<float>(nx * (ny + 1));
    d.bz = a.take<float>((nx + 1) * (ny + 1));
    d.ex2 = a.take<float>((nx + 1) * ny);
    d.ey2 = a.take<float>(nx * (ny + 1));
    d.bz2 = a.take<float>((nx + 1) * (ny + 1));
    d.rho = a.take<float>(nodes);
    d.jx = a.take<float>(nodes);
    d.jy = a.take<float>(nodes);
    d.w1 = a.take<float>(2 * pnodes);
    d.w2 = a.take<float>(2 * pnodes);
    d.phi = a.take<float>(2 * pnodes);
    d.damp_ex_x = a.take<float>(nx + 1);
    d.damp_ex_y = a.take<float>(ny);
    d.damp_ey_x = a.take<float>(nx);
    d.damp_ey_y = a.take<float>(ny + 1);
    d.damp_bz_x = a.take<float>(nx + 1);
    d.damp_bz_y = a.take<float>(ny + 1);
    d.inv_lap = a.take<float>(pnodes);
    d.cxy = a.take<int32_t>(n);
    d.fx = a.take<float>(n);
    d.fy = a.take<float>(n);
    d.vx = a.take<float>(n);
    d.vy = a.take<float>(n);
    d.scell = a.take<int32_t>(n);
    d.snap = a.take<snap4>(n);
    d.pforce = a.take<float2>(n);
    d.order = a.take<int32_t>(n);
    d.ecell = a.take<int32_t>(n);
    d.erank = a.take<int32_t>(n);
    d.count = a.take<int32_t>(cells);
    d.lstart = a.take<int32_t>(cells);
    d.toff = a.take<int32_t>((cells + SCAN_TILE - 1) / SCAN_TILE);
    d.toff_pre = a.take<int32_t>((cells + SCAN_TILE - 1) / SCAN_TILE);
    d.cstart = a.take<int32_t>(cells + 1);
    const size_t ptiles = (size_t)((c.cols + PUSH_TILE - 1) / PUSH_TILE) * ((c.rows + PUSH_TILE - 1) / PUSH_TILE);
    d.tile_dense = a.take<uint8_t>(ptiles);
    d.dense_tiles = a.take<int32_t>(ptiles);
    d.heavy_cells = a.take<int32_t>(n / HEAVY_CELL + 1);
    d.cur = a.take<long long>(2 * cells);
    d.cur_smooth = a.take<float>(2 * cells);
    d.display = a.take<float>(3 * cells);
    d.static_rho = a.take<long long>(cells);
    d.info = a.take<CellInfo>(cells);
    d.density = a.take<float>(cells);
    d.ion_pot = a.take<float>(cells);
    d.bats = a.take<BatteryCell>(c.n_bats);
    d.stacks = a.take<StackInfo>(c.n_stacks);
    d.stack_cells = a.take<int32_t>(c.n_stack_cells);
    d.gains = a.take<float>(c.n_stacks);
    d.has_gain = a.take<int32_t>(c.n_stacks);
    d.trim_part = a.take<float>(2 * (size_t)c.n_stacks * TRIM_BLOCKS);
    d.stack_read = a.take<float>(3 * (size_t)c.n_stacks);
    d.rest_sync = a.take<int32_t>(rest_sync_len(c.max_substeps));
    *tab = a.take<float>((size_t)TICK_TABLE * (c.n_ac + c.n_stacks));
    d.phase_tab = *tab;
    d.target_tab = *tab ? *tab + (size_t)TICK_TABLE * c.n_ac : nullptr;
    *ctr = a.take<SceneCounters>(1);
    d.ctr = *ctr;
    *d_scene = a.take<DevScene>(1);
    *d_prefix = a.take<int32_t>(2 * KC_COUNT);
}

template <class T> hipError_t put(T *dst, const std::vector<T> &src, hipStream_t st) {
    if (src.empty()) return hipSuccess;
    return hipMemcpyAsync(dst, src.data(), src.size() * sizeof(T), hipMemcpyHostToDevice, st);
}
template <class T> hipError_t put(const T *dst, const std::vector<T> &src, hipStream_t st) {
    return put(const_cast<T *>(dst), src, st);
}

// The DST-I matrix of length n and its halves on the device, built once per
// device and shared, with a reference taken; the caller holds dev->lock.
int sine_locked(EmbDevice *dev, int n, const SineSet **out) {
    auto it = dev->sine.find(n);
    if (it == dev->sine.end()) {
        const std::vector<float> full = sine_matrix(n);
        const SineHalves hv = sine_halves(n);
        SineSet set;
        auto upload = [](float **p, const std::vector<float> &v) {
            if (v.empty()) return hipSuccess;
            hipError_t e = hipMalloc(p, v.size() * sizeof(float));
            return e == hipSuccess ? hipMemcpy(*p, v.data(), v.size() * sizeof(float), hipMemcpyHostToDevice) : e;
        };
        hipError_t e = upload(&set.full, full);
        if (e == hipSuccess) e = upload(&set.e, hv.se);
        if (e == hipSuccess) e = upload(&set.e_t, hv.se_t);
        if (e == hipSuccess) e = upload(&set.o, hv.so);
        if (e != hipSuccess) {
            for (float *p : {set.full, set.e, set.e_t, set.o}) (void)hipFree(p);
            return fail(EMB_E_DEVICE, std::string("sine matrices: ") + hipGetErrorString(e));
        }
        set.bytes = sine_set_bytes(n);
        dev->sine_bytes += set.bytes;
        it = dev->sine.emplace(n, set).first;
    }
    it->second.refs++;
    it->second.last_use = ++dev->sine_clock;
    *out = &it->second;
    return EMB_OK;
}

void release_locked(EmbDevice *dev, int n) {
    auto it = dev->sine.find(n);
    if (it == dev->sine.end() || it->second.refs <= 0) return;
    it->second.refs--;
    it->second.last_use = ++dev->sine_clock;
}

void trim_locked(EmbDevice *dev, size_t keep) {
    while (dev->sine_bytes > keep) {
        auto victim = dev->sine.end();
        for (auto it = dev->sine.begin(); it != dev->sine.end(); ++it)
            if (it->second.refs == 0 && (victim == dev->sine.end() || it->second.last_use < victim->second.last_use))
                victim = it;
        if (victim == dev->sine.end()) return;
        for (float *p : {victim->second.full, victim->second.e, victim->second.e_t, victim->second.o}) (void)hipFree(p);
        dev->sine_bytes -= victim->second.bytes;
        dev->sine.erase(victim);
    }
}

// The close-range tables of the settings on the device, uploaded once per
// parameter set and kept; the caller holds dev->lock.
int close_tables_locked(EmbDevice *dev, const EmbSettings &s, DeviceTables &tab) {
    const auto key = std::make_tuple(s.electron_softening, (int)s.smoothing_passes, s.electron_cutoff);
    auto it = dev->tables.find(key);
    if (it == dev->tables.end()) {
        DeviceTables t;
        if (s.electron_cutoff > 0.0) {
            auto h = find_tables(s.electron_softening, s.smoothing_passes, s.electron_cutoff);
            if (!h) return fail(EMB_E_INVALID, "close-range tables not built for these settings");
            std::vector<float> g(h->g.begin(), h->g.end()), sq(h->square.begin(), h->square.end());
            g.push_back(0.0f);   // two zero samples past the end: where pairs beyond the cutoff read
            g.push_back(0.0f);
            EMB_HIP(hipMalloc(&t.g, g.size() * sizeof(float)));
            EMB_HIP(hipMalloc(&t.square, sq.size() * sizeof(float)));
            EMB_HIP(hipMemcpy(t.g, g.data(), g.size() * sizeof(float), hipMemcpyHostToDevice));
            EMB_HIP(hipMemcpy(t.square, sq.data(), sq.size() * sizeof(float), hipMemcpyHostToDevice));
            // Segment i of the lerp as intercept and slope in the table
            // coordinate t: g(t) = (g[i] - i d) + d t with d = g[i+1] - g[i], so
            // the pair loop needs no fraction of t (formed in FP64, rounded once;
            // within 1.6e-7 of max |g r| of the exact lerp, as the fraction form).
            // Entries PAIR_SAMPLES and up are (0, 0), where pairs at or beyond the
            // cutoff are clamped.
            std::vector<double> gp(h->g.begin(), h->g.end());
            gp.push_back(0.0);
            gp.push_back(0.0);
            gp.push_back(0.0);
            std::vector<float2> gl(gp.size() - 1);
            for (size_t i = 0; i + 1 < gp.size(); i++) {
                const double d = gp[i + 1] - gp[i];
                gl[i] = i < (size_t)PAIR_SAMPLES ? make_float2((float)(gp[i] - (double)i * d), (float)d)
                                                 : make_float2(0.0f, 0.0f);
            }
            EMB_HIP(hipMalloc(&t.g_lerp, gl.size() * sizeof(float2)));
            EMB_HIP(hipMemcpy(t.g_lerp, gl.data(), gl.size() * sizeof(float2), hipMemcpyHostToDevice));
            std::vector<float2> sf(h->square.size() / 3);
            for (size_t i = 0; i < sf.size(); i++)
                sf[i] = make_float2((float)h->square[3 * i + 1], (float)h->square[3 * i + 2]);
            EMB_HIP(hipMalloc(&t.square_f2, sf.size() * sizeof(float2)));
            EMB_HIP(hipMemcpy(t.square_f2, sf.data(), sf.size() * sizeof(float2), hipMemcpyHostToDevice));
            std::vector<float> is(h->ion_sum.begin(), h->ion_sum.end());
            EMB_HIP(hipMalloc(&t.ion_sum, is.size() * sizeof(float)));
            EMB_HIP(hipMemcpy(t.ion_sum, is.data(), is.size() * sizeof(float), hipMemcpyHostToDevice));
            t.square_n = h->square_n;
            t.half = h->half;
            t.reach = h->reach;
            t.ion_a0 = h->ion_a0;
        }
        it = dev->tables.emplace(key, t).first;
    }
    tab = it->second;
    return EMB_OK;
}

// Tables and sine sets of a scene of nx x ny nodes; on success the scene holds a
// reference to both sine lengths.
int shared_tables(EmbDevice *dev, const EmbSettings &s, int nx, int ny, DevScene &d, DeviceTables &tab) {
    std::lock_guard<std::mutex> guard(dev->lock);
    int rc = close_tables_locked(dev, s, tab);
    if (rc != EMB_OK) return rc;
    const SineSet *sx = nullptr, *sy = nullptr;
    rc = sine_locked(dev, nx, &sx);
    if (rc != EMB_OK) return rc;
    rc = sine_locked(dev, ny, &sy);
    if (rc != EMB_OK) {
        release_locked(dev, nx);
        return rc;
    }
    d.sx_e = sx->e;
    d.sx_o = sx->o;
    d.sy_et = sy->e_t;
    d.sy_o = sy->o;
    return EMB_OK;
}

}  // namespace

int device_sine(EmbDevice *dev, int n, const SineSet **out) {
    std::lock_guard<std::mutex> guard(dev->lock);
    return sine_locked(dev, n, out);
}

void release_sine(EmbDevice *dev, int n) {
    std::lock_guard<std::mutex> guard(dev->lock);
    release_locked(dev, n);
    // Idle sets stay for the next scene of the same length, up to a cap.
    size_t idle = 0;
    for (auto &kv : dev->sine)
        if (kv.second.refs == 0) idle += kv.second.bytes;
    if (idle > SINE_IDLE_BYTES) trim_locked(dev, dev->sine_bytes - (idle - SINE_IDLE_BYTES));
}

size_t sine_set_bytes(int n) {
    const size_t h0 = (size_t)sine_half_pitch((n + 1) / 2), h1 = (size_t)sine_half_pitch(n / 2);
    return ((size_t)n * n + 2 * h0 * h0 + h1 * h1) * sizeof(float);
}

size_t sine_missing_bytes(EmbDevice *dev, int nx, int ny) {
    std::lock_guard<std::mutex> guard(dev->lock);
    size_t b = 0;
    if (!dev->sine.count(nx)) b += sine_set_bytes(nx);
    if (ny != nx && !dev->sine.count(ny)) b += sine_set_bytes(ny);
    return b;
}

size_t sine_held_bytes(EmbDevice *dev) {
    std::lock_guard<std::mutex> guard(dev->lock);
    return dev->sine_bytes;
}

void trim_sine(EmbDevice *dev, size_t keep) {
    std::lock_guard<std::mutex> guard(dev->lock);
    trim_locked(dev, keep);
}

size_t scene_bytes(const EmbSettings &s, int cols, int rows, int64_t n, int n_stacks, int n_stack_cells, int n_bats,
                   int n_ac) {
    Counts c{cols, rows, cols + 2 * s.absorbing_margin, rows + 2 * s.absorbing_margin, n, n_stacks, n_stack_cells, n_bats, n_ac,
             s.max_substeps};
    Bump a{nullptr};
    DevScene d{};
    DevScene *ds;
    int32_t *dp;
    SceneCounters *ct;
    float *tab;
    layout(a, c, d, &ds, &dp, &ct, &tab);
    return a.off;
}

void scene_blocks(const DevScene &d, int out[KC_COUNT]) {
    const long long cells = (long long)d.cols * d.rows;
    auto cdiv = [](long long a, long long b) { return (int)((a + b - 1) / b); };
    out[KC_ELECTRONS] = std::max(1, cdiv(d.n, 256));   // (k_scatter<true> runs the scene's top-level scan even with no electrons)
    out[KC_SCAN] = cdiv(cells, SCAN_TILE);
    out[KC_CELLS] = cdiv(cells, 256);
    out[KC_SMOOTH] = cdiv(d.cols + 2 * d.passes, DEP_TILE) * cdiv(d.rows + 2 * d.passes, DEP_TILE);
    out[KC_FDTD] = cdiv(cdiv(d.ny + 1, FDTD_STRIP), FDTD_WAVES) * cdiv(d.nx + 1, fdtd_rows(d.nx, d.ny));
    out[KC_GEMM] = dst_gemm_blocks(d.nx, d.ny);
    out[KC_GEMM_SMALL] = dst_gemm_small_blocks(d.nx, d.ny);
    out[KC_NODES] = cdiv((long long)(d.nx + 1) * (d.ny + 1), 256);
    out[KC_STACKS] = std::max(1, d.n_stacks) * TRIM_BLOCKS;
    out[KC_SCENE] = 1;
    // (the tiled push walks the sort's list of dense tiles: blocks enough for
    // four rounds of the chip's resident blocks, never more than tiles)
    out[KC_PUSH] = std::min(PUSH_BLOCKS_MAX, cdiv(d.cols, PUSH_TILE) * cdiv(d.rows, PUSH_TILE));
    out[KC_HEAVY] = (int)std::min<long long>(256, std::max<long long>(1, d.n / 16384));
    out[KC_REST] = (int)std::min<long long>(256, std::max<long long>(1, cdiv(d.n, 1024)));
}

int make_resident(EmbDevice *dev, const EmbSettings &s, HostScene &&hs_in, hipStream_t st, Resident &r) {
    r.dev = dev;
    r.s = s;
    r.hs = std::move(hs_in);
    HostScene &hs = r.hs;
    const Counts c{hs.cols, hs.rows, hs.nx, hs.ny, hs.n, (int)hs.stacks.size(), (int)hs.stack_cells.size(), (int)hs.bats.size(),
                   (int)hs.ac_freq.size(), s.max_substeps};
    Bump dry{nullptr};
    DevScene tmp{};
    DevScene *dsp;
    int32_t *dpp;
    SceneCounters *ctp;
    float *tabp;
    layout(dry, c, tmp, &dsp, &dpp, &ctp, &tabp);
    r.bytes = dry.off;
    EMB_HIP(hipMalloc(&r.arena, r.bytes));
    EMB_HIP(hipMemsetAsync(r.arena, 0, r.bytes, st));
    Bump a{static_cast<char *>(r.arena)};
    DevScene &d = r.ds;
    std::memset(&d, 0, sizeof(d));
    SceneCounters *ctr;
    float *tab;
    layout(a, c, d, &r.d_scene, &r.d_prefix, &ctr, &tab);
    d.cols = hs.cols;
    d.rows = hs.rows;
    d.nx = hs.nx;
    d.ny = hs.ny;
    d.pny = poisson_pitch(hs.ny);
    d.margin = hs.margin;
    d.n = (int32_t)hs.n;
    d.n_stacks = (int32_t)hs.stacks.size();
    d.fdtd_substeps = hs.fdtd_substeps;
    d.max_substeps = s.max_substeps;
    d.passes = s.smoothing_passes;
    d.n_ac = (int32_t)hs.ac_freq.size();
    DeviceTables tables;
    int rc = shared_tables(dev, s, hs.nx, hs.ny, d, tables);
    if (rc != EMB_OK) return rc;
    r.sine_nx = hs.nx;
    r.sine_ny = hs.ny;
    d.close_on = s.electron_cutoff > 0.0 ? 1 : 0;
    d.reach = tables.reach;
    d.square_n = tables.square_n;
    d.g_table = tables.g;
    d.square = tables.square;
    d.square_f2 = reinterpret_cast<const lerp2 *>(tables.square_f2);
    d.ion_sum = tables.ion_sum;
    d.g_lerp = tables.g_lerp;
    d.ion_a0 = tables.ion_a0;
    const double q = hs.q, mass = hs.mass, k = s.coulomb_k;
    d.dt = (float)s.dt;
    d.max_step = (float)s.max_step;
    d.restitution = (float)s.wall_restitution;
    d.q = (float)q;
    d.qm = (float)(q / mass);
    d.inv_mass = (float)(1.0 / mass);
    d.k_pair = (float)(k * q * q);
    d.k_ion_q = (float)(k * s.cell_charge * q);
    d.k_q = (float)(k * q);
    d.cutoff2 = (float)(s.electron_cutoff * s.electron_cutoff);
    d.g_scale = s.electron_cutoff > 0.0 ? (float)(4096.0 / (s.electron_cutoff * s.electron_cutoff)) : 0.0f;
    d.pair_s2 = s.electron_cutoff > 0.0 ? (float)(4096.0 / (s.electron_cutoff * s.electron_cutoff) / 281474976710656.0) : 0.0f;
    d.k_pair_fx = (float)(k * q * q / 16777216.0);
    d.alpha_current = (float)std::min(s.dt / s.current_smoothing_time, 1.0);
    d.alpha_display = s.display_smoothing_time > 0.0 ? (float)(1.0 - std::exp(-s.dt / s.display_smoothing_time)) : 1.0f;
    d.half = (float)tables.half;
    d.square_samples = 32.0f;
    d.c2h = (float)(hs.h_fdtd * s.speed_of_light * s.speed_of_light);
    d.half_h = (float)(0.5 * hs.h_fdtd);
    d.jscale = (float)(0.5 * (4.0 * PI * k * hs.h_fdtd));
    d.four_pi_k = (float)(4.0 * PI * k);
    d.trim_rate = (float)(s.dt / s.battery_trim_time);
    d.gain_low = (float)s.battery_gain_low;
    d.gain_high = (float)s.battery_gain_high;
    const double smooth = std::pow(4.0, 2 * s.smoothing_passes);
    const double unit = s.cell_charge / (double)s.carriers_per_cell;
    d.rho_from_fixed = (float)(unit / RHO_SCALE / smooth);
    d.j_from_fixed = (float)(q / J_SCALE / smooth);
    d.i_from_fixed = (float)(q / J_SCALE);
    EMB_HIP(put(d.damp_ex_x, hs.damp[0], st));
    EMB_HIP(put(d.damp_ex_y, hs.damp[1], st));
    EMB_HIP(put(d.damp_ey_x, hs.damp[2], st));
    EMB_HIP(put(d.damp_ey_y, hs.damp[3], st));
    EMB_HIP(put(d.damp_bz_x, hs.damp[4], st));
    EMB_HIP(put(d.damp_bz_y, hs.damp[5], st));
    EMB_HIP(put(d.inv_lap, hs.inv_lap, st));
    EMB_HIP(put(d.cxy, hs.cxy, st));
    EMB_HIP(put(d.fx, hs.fx, st));
    EMB_HIP(put(d.fy, hs.fy, st));
    EMB_HIP(put(d.vx, hs.vx, st));
    EMB_HIP(put(d.vy, hs.vy, st));
    EMB_HIP(put(d.static_rho, hs.static_rho, st));
    EMB_HIP(put(d.info, hs.info, st));
    EMB_HIP(put(d.density, hs.density, st));
    EMB_HIP(put(d.ion_pot, hs.ion_pot, st));
    EMB_HIP(put(d.bats, hs.bats, st));
    EMB_HIP(put(d.stacks, hs.stack_info, st));
    EMB_HIP(put(d.stack_cells, hs.stack_cells, st));
    EMB_HIP(put(d.gains, hs.gains, st));
    EMB_HIP(put(d.has_gain, hs.has_gain, st));
    EMB_HIP(hipMemcpyAsync(r.d_scene, &d, sizeof(DevScene), hipMemcpyHostToDevice, st));
    int blocks[KC_COUNT];
    scene_blocks(d, blocks);
    std::vector<int32_t> prefix(2 * KC_COUNT);
    for (int k2 = 0; k2 < KC_COUNT; k2++) {
        prefix[2 * k2] = 0;
        prefix[2 * k2 + 1] = blocks[k2];
        r.batch.total[k2] = blocks[k2];
    }
    EMB_HIP(put(r.d_prefix, prefix, st));
    r.batch.scenes = r.d_scene;
    r.batch.prefix = r.d_prefix;
    r.batch.n_scenes = 1;
    r.batch.even_fdtd = r.ds.fdtd_substeps % 2 == 0;
    r.batch.top_small = (r.ds.cols * r.ds.rows + SCAN_TILE - 1) / SCAN_TILE <= 256;
    r.tab_floats = (size_t)TICK_TABLE * (hs.ac_freq.size() + hs.stacks.size());
    if (r.tab_floats > 0) {
        EMB_HIP(hipHostMalloc((void **)&r.pin_tab, r.tab_floats * sizeof(float), hipHostMallocDefault));
        std::fill(r.pin_tab, r.pin_tab + r.tab_floats, 0.0f);
        EMB_HIP(hipEventCreateWithFlags(&r.tab_event, hipEventDisableTiming));
        EMB_HIP(hipEventRecord(r.tab_event, st));
    }
    // The copies above read pageable host vectors: let them finish before the
    // host copies of the electrons are dropped.
    EMB_HIP(hipStreamSynchronize(st));
    hs.cxy = {};
    hs.fx = hs.fy = hs.vx = hs.vy = {};
    return EMB_OK;
}

void free_resident(Resident &r) {
    if (r.sine_nx > 0) {
        release_sine(r.dev, r.sine_nx);
        release_sine(r.dev, r.sine_ny);
        r.sine_nx = r.sine_ny = 0;
    }
    if (r.tab_event) (void)hipEventSynchronize(r.tab_event), (void)hipEventDestroy(r.tab_event);
    if (r.pin_tab) (void)hipHostFree(r.pin_tab);
    if (r.arena) (void)hipFree(r.arena);
    if (r.dep_dump) (void)hipFree(r.dep_dump);
    r.tab_event = nullptr;
    r.pin_tab = nullptr;
    r.arena = nullptr;
    r.dep_dump = nullptr;
}

int enqueue_tick_tables(Resident &r, int count, hipStream_t st) {
    const int n_ac = (int)r.hs.ac_freq.size(), n_st = (int)r.hs.stacks.size();
    if (r.tab_floats > 0) EMB_HIP(hipEventSynchronize(r.tab_event));
    for (int j = 0; j < count; j++) {
        const uint64_t row = (r.tick + (uint64_t)j) % TICK_TABLE;
        const double t = r.time;
        // AC phase frac(f t) in FP64, so FP32 keeps sin(2 pi f t) exact at any t
        // (the reference keeps t mod 1000 for the same reason, sim.rs:30-32, 614-622).
        for (int k = 0; k < n_ac; k++) {
            const double ph = r.hs.ac_freq[(size_t)k] * t;
            r.pin_tab[row * n_ac + k] = (float)(ph - std::floor(ph));
        }
        for (int k = 0; k < n_st; k++)
            r.pin_tab[(size_t)TICK_TABLE * n_ac + row * n_st + k] = (float)r.hs.stacks[(size_t)k].emf_at(t);
        r.time += r.s.dt;   // Simulation::step: time += dt (simulation.rs:70)
    }
    r.tick += (uint64_t)count;
    if (r.tab_floats > 0) {
        EMB_HIP(hipMemcpyAsync(const_cast<float *>(r.ds.phase_tab), r.pin_tab, r.tab_floats * sizeof(float),
                               hipMemcpyHostToDevice, st));
        EMB_HIP(hipEventRecord(r.tab_event, st));
    }
    return EMB_OK;
}

// Electron positions as FP64 cell + fraction (exact).
int read_positions(const DevScene &d, double *pos) {
    const size_t n = (size_t)d.n;
    std::vector<int32_t> c(n);
    std::vector<float> fx(n), fy(n);
    if (n) {
        EMB_HIP(hipMemcpy(c.data(), d.cxy, n * 4, hipMemcpyDeviceToHost));
        EMB_HIP(hipMemcpy(fx.data(), d.fx, n * 4, hipMemcpyDeviceToHost));
        EMB_HIP(hipMemcpy(fy.data(), d.fy, n * 4, hipMemcpyDeviceToHost));
    }
    for (size_t i = 0; i < n; i++) {
        pos[2 * i] = (double)cell_x(c[i]) + (double)fx[i];
        pos[2 * i + 1] = (double)cell_y(c[i]) + (double)fy[i];
    }
    return EMB_OK;
}

}  // namespace emb

// =====================================================================================
// C ABI
// =====================================================================================

using namespace emb;

struct EmbSim {
    EmbDevice *dev = nullptr;
    hipStream_t stream = nullptr;
    Resident r;
    double *diag = nullptr;
    float2 *force = nullptr;   // emb_sim_close_force output, grown to the electron count on use
    size_t force_cap = 0;      // electrons `force` holds (an upload may raise the count)
    float *probe = nullptr;    // emb_sim_probe output: 32 words + 3 x 256 counts, allocated on first use
    // One tick of this scene captured as a HIP graph on first step (kernels only:
    // the per-tick tables are copied before it), dropped when an upload replaces
    // the resident scene.
    hipGraph_t graph = nullptr;
    hipGraphExec_t graph_exec = nullptr;
    // The charge's potential (phi's channel 0) is solved every tick only in a
    // scene with battery stacks; otherwise on the first read after a tick
    // (read_potential, the probe), from the right-hand side the tick left.
    bool pot_fresh = true;
};

namespace {

int fresh_potential(EmbSim *sim) {
    if (sim->pot_fresh) return EMB_OK;
    EMB_HIP(launch_dst_gemms(sim->r.batch, sim->stream, DST_POT));
    sim->pot_fresh = true;
    return EMB_OK;
}

}  // namespace

namespace {

void drop_tick_graph(EmbSim *sim) {
    if (sim->graph_exec) (void)hipGraphExecDestroy(sim->graph_exec);
    if (sim->graph) (void)hipGraphDestroy(sim->graph);
    sim->graph_exec = nullptr;
    sim->graph = nullptr;
}

int capture_tick_graph(EmbSim *sim) {
    EMB_HIP(hipStreamBeginCapture(sim->stream, hipStreamCaptureModeThreadLocal));
    const hipError_t e = tick_kernels(sim->r.batch, sim->r.ds.fdtd_substeps, sim->stream);
    hipGraph_t g = nullptr;
    const hipError_t e2 = hipStreamEndCapture(sim->stream, &g);
    if (e != hipSuccess || e2 != hipSuccess) {
        if (g) (void)hipGraphDestroy(g);
        return fail(EMB_E_DEVICE, std::string("tick graph capture: ") + hipGetErrorString(e != hipSuccess ? e : e2));
    }
    sim->graph = g;
    const hipError_t e3 = hipGraphInstantiate(&sim->graph_exec, g, nullptr, nullptr, 0);
    if (e3 != hipSuccess) {
        drop_tick_graph(sim);
        return fail(EMB_E_DEVICE, std::string("tick graph instantiate: ") + hipGetErrorString(e3));
    }
    return EMB_OK;
}

}  // namespace

extern "C" {

const char *emb_last_error(void) { return g_error.c_str(); }

void emb_settings_default(EmbSettings *s) {
    std::memset(s, 0, sizeof(*s));
    s->dt = 1.0 / 30.0;
    s->cell_charge = 4.0;
    s->carriers_per_cell = 8;
    s->battery_density = 4;
    s->carrier_charge_to_mass = 1.0;
    s->drag_per_resistance = 1.0;
    s->battery_resistance = 0.1;
    s->battery_trim_time = 0.25;
    s->battery_gain_low = 0.5;
    s->battery_gain_high = 4.0;
    s->current_smoothing_time = 2.0;
    s->wall_restitution = 0.3;
    s->max_step = 0.4;
    s->max_substeps = 16;
    s->smoothing_passes = 2;
    s->coulomb_k = 1.0;
    s->speed_of_light = 8.0;
    s->absorbing_margin = 16;
    s->absorbing_strength = 3.0;
    s->electron_softening = 0.25;
    s->electron_cutoff = 3.0;
    s->display_smoothing_time = 0.5;
}

int emb_device_open(int ordinal, EmbDevice **out) {
    EMB_GUARD({
        if (!out) return fail(EMB_E_INVALID, "null output");
        int count = 0;
        EMB_HIP(hipGetDeviceCount(&count));
        if (ordinal < 0 || ordinal >= count) return fail(EMB_E_INVALID, "no such device");
        EMB_HIP(hipSetDevice(ordinal));
        hipDeviceProp_t p;
        EMB_HIP(hipGetDeviceProperties(&p, ordinal));
        if (std::strncmp(p.gcnArchName, "gfx950", 6) != 0)
            return fail(EMB_E_DEVICE, std::string("built for gfx950, device is ") + p.gcnArchName);
        // This device's address of the diagnostics scratch array (the device is current).
        void *scr = nullptr;
        EMB_HIP(hipGetSymbolAddress(&scr, HIP_SYMBOL(scratch)));
        auto *d = new EmbDevice();
        d->ordinal = ordinal;
        d->n_cu = p.multiProcessorCount;
        d->scratch = static_cast<double *>(scr);
        *out = d;
        return EMB_OK;
    })
}

int emb_device_close(EmbDevice *dev) {
    if (!dev) return EMB_OK;
    (void)hipSetDevice(dev->ordinal);
    for (auto &kv : dev->sine)
        for (float *p : {kv.second.full, kv.second.e, kv.second.e_t, kv.second.o}) (void)hipFree(p);
    for (auto &kv : dev->tables) {
        if (kv.second.g) (void)hipFree(kv.second.g);
        if (kv.second.square) (void)hipFree(kv.second.square);
        if (kv.second.ion_sum) (void)hipFree(kv.second.ion_sum);
        if (kv.second.g_lerp) (void)hipFree(kv.second.g_lerp);
        if (kv.second.square_f2) (void)hipFree(kv.second.square_f2);
    }
    delete dev;
    return EMB_OK;
}

int64_t emb_scene_footprint(const EmbSettings *s, int32_t cols, int32_t rows, int64_t n_electrons) {
    if (!s || cols <= 0 || rows <= 0 || n_electrons < 0) return -1;
    // Stacks and battery cells are small; budget a generous fixed share for them.
    return (int64_t)scene_bytes(*s, cols, rows, n_electrons, 64, 4096, 256, 256) + (1 << 20);
}

int emb_sim_create(EmbDevice *dev, const EmbSettings *settings, const EmbScene *scene, EmbSim **out) {
    EMB_GUARD({
        if (!dev || !settings || !scene || !out) return fail(EMB_E_INVALID, "null argument");
        int trc = ensure_tables(dev, settings->electron_softening, settings->smoothing_passes, settings->electron_cutoff);
        if (trc != EMB_OK) return trc;
        HostScene hs;
        const std::string why = prepare_scene(*settings, *scene, hs);
        if (!why.empty()) return fail(EMB_E_INVALID, why);
        EMB_HIP(hipSetDevice(dev->ordinal));
        auto *sim = new EmbSim();
        sim->dev = dev;
        hipError_t e = hipStreamCreateWithFlags(&sim->stream, hipStreamNonBlocking);
        if (e != hipSuccess) {
            delete sim;
            return fail(EMB_E_DEVICE, std::string("hipStreamCreate: ") + hipGetErrorString(e));
        }
        int rc = make_resident(dev, *settings, std::move(hs), sim->stream, sim->r);
        if (rc == EMB_OK) {
            sim->r.time = scene->time;
            e = setup_kernels(sim->r.batch, sim->stream);
            if (e == hipSuccess) e = finish_sync_kernel(sim->r.batch, sim->stream);   // display = fields
            if (e == hipSuccess) e = hipStreamSynchronize(sim->stream);
            if (e != hipSuccess) rc = fail(EMB_E_DEVICE, std::string("setup: ") + hipGetErrorString(e));
        }
        if (rc != EMB_OK) {
            const std::string keep = g_error;
            free_resident(sim->r);
            (void)hipStreamDestroy(sim->stream);
            delete sim;
            g_error = keep;
            return rc;
        }
        *out = sim;
        return EMB_OK;
    })
}

int emb_sim_upload(EmbSim *sim, const EmbScene *scene, int reset_display) {
    EMB_GUARD({
        if (!sim || !scene) return fail(EMB_E_INVALID, "null argument");
        if (scene->cols != sim->r.hs.cols || scene->rows != sim->r.hs.rows)
            return fail(EMB_E_INVALID, "upload must keep the grid shape");
        HostScene hs;
        const std::string why = prepare_scene(sim->r.s, *scene, hs);
        if (!why.empty()) return fail(EMB_E_INVALID, why);
        EMB_HIP(hipSetDevice(sim->dev->ordinal));
        Resident fresh;
        int rc = make_resident(sim->dev, sim->r.s, std::move(hs), sim->stream, fresh);
        if (rc != EMB_OK) {
            free_resident(fresh);
            return rc;
        }
        const DevScene &o = sim->r.ds, &d = fresh.ds;
        const size_t nx = o.nx, ny = o.ny;
        // The fields carry on (GpuSim::upload keeps them, sim.rs:391-395).
        EMB_HIP(hipMemcpyAsync(d.ex, o.ex, (nx + 1) * ny * 4, hipMemcpyDeviceToDevice, sim->stream));
        EMB_HIP(hipMemcpyAsync(d.ey, o.ey, nx * (ny + 1) * 4, hipMemcpyDeviceToDevice, sim->stream));
        EMB_HIP(hipMemcpyAsync(d.bz, o.bz, (nx + 1) * (ny + 1) * 4, hipMemcpyDeviceToDevice, sim->stream));
        EMB_HIP(hipMemcpyAsync(&d.ctr->tick, &o.ctr->tick, sizeof(uint32_t), hipMemcpyDeviceToDevice, sim->stream));
        // So do the display and the smoothed current (the grid shape is unchanged).
        const size_t cells = (size_t)o.cols * o.rows;
        EMB_HIP(hipMemcpyAsync(d.display, o.display, 3 * cells * 4, hipMemcpyDeviceToDevice, sim->stream));
        EMB_HIP(hipMemcpyAsync(d.cur_smooth, o.cur_smooth, 2 * cells * 4, hipMemcpyDeviceToDevice, sim->stream));
        fresh.time = sim->r.time;
        fresh.tick = sim->r.tick;
        EMB_HIP(setup_kernels(fresh.batch, sim->stream));
        if (reset_display) EMB_HIP(finish_sync_kernel(fresh.batch, sim->stream));
        EMB_HIP(hipStreamSynchronize(sim->stream));
        drop_tick_graph(sim);   // it holds the old scene's descriptors
        free_resident(sim->r);
        sim->r = std::move(fresh);
        sim->pot_fresh = true;   // set-up solved both channels
        return EMB_OK;
    })
}

int emb_sim_step(EmbSim *sim, int ticks) {
    EMB_GUARD({
        if (!sim || ticks < 0) return fail(EMB_E_INVALID, "bad argument");
        EMB_HIP(hipSetDevice(sim->dev->ordinal));
        // EMBLOKS_TICK_GRAPH=0 launches the same kernels one by one (for tools
        // that trace launches but not graph replays).
        static const bool graphs = [] {
            const char *v = std::getenv("EMBLOKS_TICK_GRAPH");
            return !(v && v[0] == '0');
        }();
        if (graphs && ticks > 0 && !sim->graph_exec) {
            const int rc = capture_tick_graph(sim);
            if (rc != EMB_OK) return rc;
        }
        while (ticks > 0) {
            const int chunk = std::min(ticks, TICK_TABLE);
            int rc = enqueue_tick_tables(sim->r, chunk, sim->stream);
            if (rc != EMB_OK) return rc;
            for (int t = 0; t < chunk; t++) {
                if (graphs) EMB_HIP(hipGraphLaunch(sim->graph_exec, sim->stream));
                else EMB_HIP(tick_kernels(sim->r.batch, sim->r.ds.fdtd_substeps, sim->stream));
            }
            if (chunk > 0) sim->pot_fresh = sim->r.ds.n_stacks > 0;
            ticks -= chunk;
        }
        return EMB_OK;
    })
}

int emb_sim_sync(EmbSim *sim) {
    if (!sim) return fail(EMB_E_INVALID, "null sim");
    EMB_HIP(hipStreamSynchronize(sim->stream));
    // k_rest's wait limit (a queue fault, never expected) leaves a mark instead of
    // a hang; a charge grid past the fixed-point headroom leaves another.
    int32_t flags[2] = {0, 0};
    EMB_HIP(hipMemcpy(flags, &sim->r.ds.ctr->rest_error, sizeof(flags), hipMemcpyDeviceToHost));
    const std::string why = numeric_fault(flags[0], flags[1]);
    if (!why.empty()) return fail(EMB_E_NUMERIC, why);
    return EMB_OK;
}

int emb_sim_dims(EmbSim *sim, EmbDims *out) {
    if (!sim || !out) return fail(EMB_E_INVALID, "null argument");
    const HostScene &h = sim->r.hs;
    std::memset(out, 0, sizeof(*out));
    out->cols = h.cols;
    out->rows = h.rows;
    out->margin = h.margin;
    out->nx = h.nx;
    out->ny = h.ny;
    out->n_stacks = (int32_t)h.stacks.size();
    out->n_electrons = h.n;
    out->fdtd_substeps = h.fdtd_substeps;
    out->time = sim->r.time;
    int32_t sub = 0;
    EMB_HIP(hipStreamSynchronize(sim->stream));
    EMB_HIP(hipMemcpy(&sub, &sim->r.ds.ctr->substeps, sizeof(sub), hipMemcpyDeviceToHost));
    out->last_substeps = sub;
    return EMB_OK;
}

static int copy_back(EmbSim *sim, const float *src, size_t n, double *dst, size_t stride = 1, size_t offset = 0) {
    std::vector<float> tmp(n);
    if (n) EMB_HIP(hipMemcpy(tmp.data(), src, n * sizeof(float), hipMemcpyDeviceToHost));
    for (size_t i = 0; i < n; i++) dst[i * stride + offset] = tmp[i];
    (void)sim;
    return EMB_OK;
}

int emb_sim_read_state(EmbSim *sim, double *pos, double *vel, double *gains) {
    EMB_GUARD({
        if (!sim) return fail(EMB_E_INVALID, "null sim");
        EMB_HIP(hipStreamSynchronize(sim->stream));
        const DevScene &d = sim->r.ds;
        const size_t n = (size_t)d.n;
        int rc = EMB_OK;
        if (pos && rc == EMB_OK) rc = read_positions(d, pos);
        if (vel && rc == EMB_OK) rc = copy_back(sim, d.vx, n, vel, 2, 0);
        if (vel && rc == EMB_OK) rc = copy_back(sim, d.vy, n, vel, 2, 1);
        if (rc != EMB_OK) return rc;
        if (gains && d.n_stacks > 0) {
            std::vector<float> g((size_t)d.n_stacks);
            std::vector<int32_t> has((size_t)d.n_stacks);
            EMB_HIP(hipMemcpy(g.data(), d.gains, g.size() * 4, hipMemcpyDeviceToHost));
            EMB_HIP(hipMemcpy(has.data(), d.has_gain, has.size() * 4, hipMemcpyDeviceToHost));
            for (size_t k = 0; k < g.size(); k++) gains[k] = has[k] ? (double)g[k] : std::nan("");
        }
        return EMB_OK;
    })
}

int emb_sim_read_yee(EmbSim *sim, double *ex, double *ey, double *bz) {
    EMB_GUARD({
        if (!sim) return fail(EMB_E_INVALID, "null sim");
        EMB_HIP(hipStreamSynchronize(sim->stream));
        const DevScene &d = sim->r.ds;
        const size_t nx = d.nx, ny = d.ny;
        int rc = EMB_OK;
        if (ex) rc = copy_back(sim, d.ex, (nx + 1) * ny, ex);
        if (ey && rc == EMB_OK) rc = copy_back(sim, d.ey, nx * (ny + 1), ey);
        if (bz && rc == EMB_OK) rc = copy_back(sim, d.bz, (nx + 1) * (ny + 1), bz);
        return rc;
    })
}

int emb_sim_write_yee(EmbSim *sim, const double *ex, const double *ey, const double *bz) {
    EMB_GUARD({
        if (!sim) return fail(EMB_E_INVALID, "null sim");
        EMB_HIP(hipStreamSynchronize(sim->stream));
        const DevScene &d = sim->r.ds;
        const size_t nx = d.nx, ny = d.ny;
        auto up = [&](float *dst, const double *src, size_t n) -> int {
            if (!src) return EMB_OK;
            std::vector<float> tmp(src, src + n);
            EMB_HIP(hipMemcpy(dst, tmp.data(), n * 4, hipMemcpyHostToDevice));
            return EMB_OK;
        };
        int rc = up(d.ex, ex, (nx + 1) * ny);
        if (rc == EMB_OK) rc = up(d.ey, ey, nx * (ny + 1));
        if (rc == EMB_OK) rc = up(d.bz, bz, (nx + 1) * (ny + 1));
        return rc;
    })
}

int emb_sim_read_potential(EmbSim *sim, double *potential) {
    EMB_GUARD({
        if (!sim || !potential) return fail(EMB_E_INVALID, "null argument");
        EMB_HIP(hipSetDevice(sim->dev->ordinal));
        const int rc = fresh_potential(sim);
        if (rc != EMB_OK) return rc;
        EMB_HIP(hipStreamSynchronize(sim->stream));
        const DevScene &d = sim->r.ds;
        // Rows of phi are padded to d.pny floats.
        std::vector<float> tmp((size_t)d.nx * d.pny);
        EMB_HIP(hipMemcpy(tmp.data(), d.phi, tmp.size() * sizeof(float), hipMemcpyDeviceToHost));
        for (size_t i = 0; i < (size_t)d.nx; i++)
            for (size_t j = 0; j < (size_t)d.ny; j++) potential[i * d.ny + j] = tmp[i * d.pny + j];
        return EMB_OK;
    })
}

int emb_sim_read_display(EmbSim *sim, double *e, double *b) {
    EMB_GUARD({
        if (!sim || !e || !b) return fail(EMB_E_INVALID, "null argument");
        EMB_HIP(hipStreamSynchronize(sim->stream));
        const DevScene &d = sim->r.ds;
        const size_t cells = (size_t)d.cols * d.rows;
        std::vector<float> v(3 * cells);
        EMB_HIP(hipMemcpy(v.data(), d.display, v.size() * 4, hipMemcpyDeviceToHost));
        for (size_t c = 0; c < cells; c++) {
            e[2 * c] = v[c];
            e[2 * c + 1] = v[cells + c];
            b[c] = v[2 * cells + c];
        }
        return EMB_OK;
    })
}

int emb_sim_read_smoothed_current(EmbSim *sim, double *current) {
    EMB_GUARD({
        if (!sim || !current) return fail(EMB_E_INVALID, "null argument");
        EMB_HIP(hipStreamSynchronize(sim->stream));
        const DevScene &d = sim->r.ds;
        const size_t cells = (size_t)d.cols * d.rows;
        std::vector<float> v(2 * cells);
        EMB_HIP(hipMemcpy(v.data(), d.cur_smooth, v.size() * 4, hipMemcpyDeviceToHost));
        for (size_t c = 0; c < cells; c++) {
            current[2 * c] = v[c];
            current[2 * c + 1] = v[cells + c];
        }
        return EMB_OK;
    })
}

int emb_sim_set_current(EmbSim *sim, const double *current) {
    EMB_GUARD({
        if (!sim || !current) return fail(EMB_E_INVALID, "null argument");
        const DevScene &d = sim->r.ds;
        const size_t cells = (size_t)d.cols * d.rows;
        std::vector<float> v(2 * cells);
        for (size_t c = 0; c < cells; c++) {
            v[c] = (float)current[2 * c];
            v[cells + c] = (float)current[2 * c + 1];
        }
        EMB_HIP(hipSetDevice(sim->dev->ordinal));
        EMB_HIP(hipMemcpyAsync(d.cur_smooth, v.data(), v.size() * 4, hipMemcpyHostToDevice, sim->stream));
        EMB_HIP(hipStreamSynchronize(sim->stream));
        return EMB_OK;
    })
}

int emb_sim_read_current(EmbSim *sim, double *current) {
    EMB_GUARD({
        if (!sim || !current) return fail(EMB_E_INVALID, "null argument");
        EMB_HIP(hipStreamSynchronize(sim->stream));
        const DevScene &d = sim->r.ds;
        const size_t cells = (size_t)d.cols * d.rows;
        std::vector<long long> tmp(2 * cells);
        EMB_HIP(hipMemcpy(tmp.data(), d.cur, tmp.size() * 8, hipMemcpyDeviceToHost));
        const double scale = sim->r.hs.q / J_SCALE;
        for (size_t c = 0; c < cells; c++) {
            current[2 * c] = (double)tmp[c] * scale;
            current[2 * c + 1] = (double)tmp[cells + c] * scale;
        }
        return EMB_OK;
    })
}

int emb_sim_read_charge_fixed(EmbSim *sim, int64_t *rho) {
    EMB_GUARD({
        if (!sim || !rho) return fail(EMB_E_INVALID, "null argument");
        Resident &r = sim->r;
        const size_t cells = (size_t)r.ds.cols * r.ds.rows;
        EMB_HIP(hipStreamSynchronize(sim->stream));
        if (!r.dep_dump) EMB_HIP(hipMalloc(&r.dep_dump, 3 * cells * sizeof(long long)));
        r.ds.dep_dump = r.dep_dump;
        EMB_HIP(hipMemcpyAsync(r.d_scene, &r.ds, sizeof(DevScene), hipMemcpyHostToDevice, sim->stream));
        // Re-running the deposit on the unchanged electrons is idempotent.
        EMB_HIP(deposit_kernel(r.batch, sim->stream));
        r.ds.dep_dump = nullptr;
        EMB_HIP(hipMemcpyAsync(r.d_scene, &r.ds, sizeof(DevScene), hipMemcpyHostToDevice, sim->stream));
        EMB_HIP(hipStreamSynchronize(sim->stream));
        std::vector<long long> tmp(3 * cells);
        EMB_HIP(hipMemcpy(tmp.data(), r.dep_dump, tmp.size() * 8, hipMemcpyDeviceToHost));
        std::memcpy(rho, tmp.data(), cells * 8);
        return EMB_OK;
    })
}

int emb_sim_probe(EmbSim *sim, int32_t c, int32_t r, int32_t arrow_stride, EmbProbe *out) {
    EMB_GUARD({
        if (!sim || !out) return fail(EMB_E_INVALID, "null argument");
        Resident &res = sim->r;
        EMB_HIP(hipSetDevice(sim->dev->ordinal));
        constexpr int WORDS = 32, BINS = 256;
        if (!sim->probe) EMB_HIP(hipMalloc(&sim->probe, (WORDS + 3 * BINS) * sizeof(float)));
        const int cols = res.hs.cols, rows = res.hs.rows;
        const int cell = (c >= 0 && r >= 0 && c < cols && r < rows) ? c * rows + r : -1;
        const double c2 = res.s.speed_of_light * res.s.speed_of_light;
        const double inv_mu0 = c2 / (4.0 * M_PI * res.s.coulomb_k);   // sim.rs:262
        unsigned *hist = reinterpret_cast<unsigned *>(sim->probe + WORDS);
        const int rc = fresh_potential(sim);   // the probe reads the potential
        if (rc != EMB_OK) return rc;
        EMB_HIP(probe_kernel(res.d_scene, cell, arrow_stride, (float)c2, (float)inv_mu0, sim->probe, hist, sim->stream));
        std::vector<float> w(WORDS);
        std::memset(out, 0, sizeof(*out));
        EMB_HIP(hipMemcpyAsync(w.data(), sim->probe, WORDS * sizeof(float), hipMemcpyDeviceToHost, sim->stream));
        EMB_HIP(hipMemcpyAsync(out->histograms, hist, 3 * BINS * sizeof(unsigned), hipMemcpyDeviceToHost, sim->stream));
        EMB_HIP(hipStreamSynchronize(sim->stream));
        out->valid = w[0] != 0.0f;
        out->electrons = (int32_t)w[1];
        out->potential = w[2];
        out->e[0] = w[3];
        out->e[1] = w[4];
        out->b = w[5];
        out->s[0] = w[6];
        out->s[1] = w[7];
        out->current[0] = w[8];
        out->current[1] = w[9];
        out->wire_current = w[10];
        out->wire_width = (int32_t)w[11];
        out->stack_state = (int32_t)w[12];
        out->v_low = w[13];
        out->v_high = w[14];
        out->stack_current = w[15];
        out->gain = w[16];
        out->mean_e_in_wires = w[17];
        out->has_mean_e = w[18] > 0.0f;
        return EMB_OK;
    })
}

int emb_sim_diagnostics(EmbSim *sim, double *gauss_residual, double *electron_charge, double *field_energy) {
    EMB_GUARD({
        if (!sim) return fail(EMB_E_INVALID, "null sim");
        Resident &r = sim->r;
        const int blocks = DIAG_BLOCKS;
        if (!sim->diag) EMB_HIP(hipMalloc(&sim->diag, 2 * blocks * sizeof(double)));
        EMB_HIP(diagnostics_kernel(r.d_scene, r.s.speed_of_light * r.s.speed_of_light, sim->diag, blocks, sim->stream));
        std::vector<double> out(2 * blocks);
        EMB_HIP(hipStreamSynchronize(sim->stream));
        EMB_HIP(hipMemcpy(out.data(), sim->diag, out.size() * 8, hipMemcpyDeviceToHost));
        double res = 0.0, en = 0.0;
        for (int b = 0; b < blocks; b++) {
            res = std::max(res, out[2 * b]);
            en += out[2 * b + 1];
        }
        if (gauss_residual) *gauss_residual = res;
        if (field_energy) *field_energy = en;
        if (electron_charge) {
            const size_t cells = (size_t)r.ds.cols * r.ds.rows;
            std::vector<int64_t> rho(cells);
            int rc = emb_sim_read_charge_fixed(sim, rho.data());
            if (rc != EMB_OK) return rc;
            // In electron charges: the cells sum to exactly -N 2^44, which passes
            // the int64 range beyond 2^19 electrons, so the sum is 128-bit.
            __int128 total = 0;
            for (auto v : rho) total += v;
            *electron_charge = (double)total / RHO_SCALE * (r.s.cell_charge / r.s.carriers_per_cell);
        }
        return EMB_OK;
    })
}

int emb_sim_diagnostics_static(EmbSim *sim, double *gauss_residual, double *field_energy) {
    if (!sim) return fail(EMB_E_INVALID, "null sim");
    EmbDevice *dev = sim->dev;
    if (!dev->scratch) return fail(EMB_E_STATE, "device has no diagnostics scratch address");
    // k_diagnostics is grid-stride over any grid with DIAG_BLOCKS blocks: every
    // scene fits the fixed scratch, and every partial it holds is rewritten here.
    EMB_HIP(hipSetDevice(dev->ordinal));
    Resident &r = sim->r;
    EMB_HIP(diagnostics_kernel(r.d_scene, r.s.speed_of_light * r.s.speed_of_light, dev->scratch, DIAG_BLOCKS, sim->stream));
    hipLaunchKernelGGL(k_diag_reduce, dim3(1), dim3(64), 0, sim->stream, dev->scratch, DIAG_BLOCKS);
    EMB_HIP(hipGetLastError());
    double out[2] = {0.0, 0.0};
    EMB_HIP(hipMemcpyAsync(out, dev->scratch + 2 * DIAG_BLOCKS, sizeof(out), hipMemcpyDeviceToHost, sim->stream));
    EMB_HIP(hipStreamSynchronize(sim->stream));
    if (gauss_residual) *gauss_residual = out[0];
    if (field_energy) *field_energy = out[1];
    return EMB_OK;
}

int emb_sim_kernel_times(EmbSim *sim, int ticks, double *avg_ms, int capacity) {
    EMB_GUARD({
        if (!sim || ticks < 1) return fail(EMB_E_INVALID, "bad argument");
        if (!avg_ms) return TICK_PHASES;
        EMB_HIP(hipSetDevice(sim->dev->ordinal));
        std::vector<hipEvent_t> ev((size_t)ticks * (TICK_PHASES + 1));
        for (auto &e : ev) EMB_HIP(hipEventCreate(&e));
        int rc = EMB_OK;
        for (int done = 0; done < ticks && rc == EMB_OK;) {
            const int chunk = std::min(ticks - done, TICK_TABLE);
            rc = enqueue_tick_tables(sim->r, chunk, sim->stream);
            for (int t = 0; t < chunk && rc == EMB_OK; t++) {
                hipError_t e = tick_kernels(sim->r.batch, sim->r.ds.fdtd_substeps, sim->stream,
                                            &ev[(size_t)(done + t) * (TICK_PHASES + 1)]);
                if (e != hipSuccess) rc = fail(EMB_E_DEVICE, hipGetErrorString(e));
            }
            done += chunk;
            sim->pot_fresh = sim->r.ds.n_stacks > 0;
        }
        if (rc == EMB_OK) {
            hipError_t e = hipStreamSynchronize(sim->stream);
            if (e != hipSuccess) rc = fail(EMB_E_DEVICE, hipGetErrorString(e));
        }
        std::vector<double> sum(TICK_PHASES, 0.0);
        for (int t = 0; t < ticks && rc == EMB_OK; t++)
            for (int p = 0; p < TICK_PHASES; p++) {
                float ms = 0.0f;
                hipError_t e = hipEventElapsedTime(&ms, ev[(size_t)t * (TICK_PHASES + 1) + p], ev[(size_t)t * (TICK_PHASES + 1) + p + 1]);
                if (e != hipSuccess) rc = fail(EMB_E_DEVICE, hipGetErrorString(e));
                sum[(size_t)p] += ms;
            }
        for (auto &e : ev) (void)hipEventDestroy(e);
        if (rc != EMB_OK) return rc;
        for (int p = 0; p < std::min(capacity, TICK_PHASES); p++) avg_ms[p] = sum[(size_t)p] / ticks;
        return TICK_PHASES;
    })
}

const char *emb_tick_phase_name(int phase) {
    return (phase >= 0 && phase < TICK_PHASES) ? TICK_PHASE_NAMES[phase] : nullptr;
}

int emb_sim_close_force(EmbSim *sim, int32_t repeats, double *force, double *avg_ms) {
    EMB_GUARD({
        if (!sim) return fail(EMB_E_INVALID, "null argument");
        if (repeats < 1) return fail(EMB_E_INVALID, "repeats must be >= 1");
        if (sim->r.ds.close_on == 0) return fail(EMB_E_STATE, "the close-range force is off (electron_cutoff = 0)");
        const size_t n = (size_t)sim->r.ds.n;
        EMB_HIP(hipSetDevice(sim->dev->ordinal));
        if (n > sim->force_cap) {
            // An upload can add electrons after an earlier call sized the buffer.
            EMB_HIP(hipStreamSynchronize(sim->stream));
            if (sim->force) EMB_HIP(hipFree(sim->force));
            sim->force = nullptr;
            sim->force_cap = 0;
            EMB_HIP(hipMalloc(&sim->force, n * sizeof(float2)));
            sim->force_cap = n;
        }
        hipEvent_t e0, e1;
        EMB_HIP(hipEventCreate(&e0));
        EMB_HIP(hipEventCreate(&e1));
        hipError_t err = hipEventRecord(e0, sim->stream);
        for (int r = 0; r < repeats && err == hipSuccess; r++) err = close_force_kernel(sim->r.batch, sim->force, sim->stream);
        if (err == hipSuccess) err = hipEventRecord(e1, sim->stream);
        if (err == hipSuccess) err = hipEventSynchronize(e1);
        float ms = 0.0f;
        if (err == hipSuccess) err = hipEventElapsedTime(&ms, e0, e1);
        (void)hipEventDestroy(e0);
        (void)hipEventDestroy(e1);
        EMB_HIP(err);
        if (avg_ms) *avg_ms = (double)ms / repeats;
        if (force && n > 0) {
            std::vector<float2> h(n);
            EMB_HIP(hipMemcpy(h.data(), sim->force, n * sizeof(float2), hipMemcpyDeviceToHost));
            for (size_t i = 0; i < n; i++) {
                force[2 * i] = h[i].x;
                force[2 * i + 1] = h[i].y;
            }
        }
        return EMB_OK;
    })
}

int emb_sim_destroy(EmbSim *sim) {
    if (!sim) return EMB_OK;
    (void)hipSetDevice(sim->dev->ordinal);
    (void)hipStreamSynchronize(sim->stream);
    drop_tick_graph(sim);
    free_resident(sim->r);
    if (sim->diag) (void)hipFree(sim->diag);
    if (sim->probe) (void)hipFree(sim->probe);
    if (sim->force) (void)hipFree(sim->force);
    (void)hipStreamDestroy(sim->stream);
    delete sim;
    return EMB_OK;
}

}  // extern "C"
