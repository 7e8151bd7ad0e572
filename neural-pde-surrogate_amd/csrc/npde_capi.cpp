// Host side of libnpde.so: the C-ABI declared in include/npde.h.
//
// Replaces the reference's in-process plugin API (solver2d.hpp:57-103 Solver; grid2d.hpp:30-66
// make_grid / apply_bcs; main2d.cpp:224-246 run loop; trajectory.hpp:31-51 block average) with
// opaque handles over device-resident state.  No call exits or throws: errors come back as
// npde_status with a message in npde_last_error().
#include <hip/hip_runtime.h>

#include <algorithm>
#include <chrono>
#include <cmath>
#include <cstdio>
#include <cstring>
#include <fstream>
#include <map>
#include <mutex>
#include <sstream>
#include <string>
#include <vector>

#include "../../include/npde.h"
#include "euler_numerics.hpp"
#include "npde_device.hpp"

namespace npde {
hipError_t launch_stage(int scheme, int lim, int stage, bool corr, bool fast, int waves_per_block, int mode,
                        int row_waves, int compact, int seam_rows, unsigned lds_reserve, const InstDesc* desc,
                        Ctl* ctl, const Task* tasks, int ntasks, Partial* part, int* ndone, hipStream_t s);
hipError_t stage_occupancy(int scheme, int lim, bool corr, bool fast, int waves_per_block, int mode, int row_waves,
                           int seam_rows, unsigned lds_reserve, int* workgroups);
hipError_t launch_wavespeed(int waves_per_block, int mode, int row_waves, const InstDesc* desc, Ctl* ctl,
                            const Task* tasks, int ntasks, Partial* part, int* ndone, hipStream_t s);
hipError_t launch_frame(const InstDesc* desc, Ctl* ctl, const Partial* part, const FTask* ftasks, int nftasks,
                        int complete, int hold, int* ndone, hipStream_t s);
hipError_t launch_block_average(const InstDesc* desc, int inst, int m, int nyc, float* out, hipStream_t s);
hipError_t launch_finish(const InstDesc* desc, Ctl* ctl, const Partial* part, int n, int mode, int* ndone,
                         hipStream_t s);
enum : int { kAfterStep = 0, kAfterWavespeed = 1 };
hipError_t launch_prep_host_step(Ctl* ctl, int n, const double* dts, hipStream_t s);
hipError_t launch_flux_net(const InstDesc* desc, Ctl* ctl, const NetTask* tasks, int ntasks, int n, int* counters,
                           const void* frags, const float* bias, float out_scale, unsigned* peak, hipStream_t s);
hipError_t launch_mfma32_probe(const float* a, const float* b, float* d, hipStream_t s);
hipError_t launch_begin_run(Ctl* ctl, int n, double t_start, double t_end, int nout, int record, long long max_new_steps,
                            int* ndone, hipStream_t s);
hipError_t launch_selftest_arith(const double* x, const double* y, long long n, double* qf, double* qi, double* rt,
                                 hipStream_t s);
}  // namespace npde

using namespace npde;

// Captured device loops come in sizes of 1, 2, 4, ... iterations up to steps_per_graph (the last size is
// steps_per_graph itself), one graph per size: level l runs min(2^l, steps_per_graph) iterations.
constexpr int kGraphLevels = 32;

// Flux-network runs: the network kernel's launch period in device-loop iterations (npde_tuning::net_period).
constexpr int kNetPeriod = 4;

struct npde_solver {
    int n = 0;
    int flux = kHllc, lim = kLimVanLeer;
    double cfl = 0.5;
    std::vector<npde_grid_desc> grids;
    std::vector<InstDesc> hdesc;
    std::vector<Ctl> hctl;
    std::vector<Task> htasks;
    std::vector<FTask> hftasks;
    std::vector<size_t> state_off;  // offset of each instance's U (U1 follows after 4 planes)
    InstDesc* ddesc = nullptr;
    Ctl* dctl = nullptr;
    Task* dtasks = nullptr;
    double* dedge = nullptr;  // edge columns of every instance (seam mode)
    size_t edge_cap = 0;      // doubles allocated at dedge
    FTask* dftasks = nullptr;
    Partial* dpart = nullptr;
    int* dndone = nullptr;
    double* ddts = nullptr;
    double* dstate = nullptr;
    float* dtraj = nullptr;
    double* dftimes = nullptr;
    size_t traj_bytes = 0;
    std::vector<int> traj_sizes;
    int traj_nout = -1;
    bool has_corr = false;
    bool fast = false;  // NPDE_ARITH_FAST: the fast-arithmetic stage kernels
    hipStream_t stream = nullptr;
    hipEvent_t ev0 = nullptr, ev1 = nullptr;
    // per-kernel profile (npde_set_profiling): an event pair around each launch, read after the stream syncs
    bool prof_on = false;
    std::vector<hipEvent_t> prof_events;  // pool: pairs (start, end) of the recorded launches
    std::vector<int> prof_kind;           // kind of each recorded launch not yet read
    double prof_ms[NPDE_PROF_KINDS] = {};
    long long prof_n[NPDE_PROF_KINDS] = {};
    npde_tuning tune{0, 4, 32, 1, 1, 1, 0, kNetPeriod};
    int lds_per_cu = 65536;    // LDS bytes per CU of the handle's device (queried at creation)
    int layout = kStrips;  // StageMode of the current task list
    int row_waves = 0;   // row groups: wavefronts per band
    int band = 0;        // band height of the current task list
    // instances whose tasks are in the current task list: all of them, except during the tail of a
    // device-resident run, when the list is rebuilt for the instances still stepping (retile_tail)
    std::vector<char> listed;
    bool subset = false;
    // [1]: iterations with the frame kernels of a recording run, [0]: without; captured on first use, all
    // sharing the keys below
    hipGraphExec_t graphs[2][kGraphLevels] = {};
    int graph_key_frames = -1;  // nftasks the cached graphs were captured with (0: no frame kernel)
    int graph_key_steps = 0;
    int graph_key_corr = -1;
    int graph_key_rows = -1, graph_key_wpb = -1;
    int graph_key_fast = -1;
    int graph_key_net = -1;
    // learned flux-correction network run at every frame landing (npde_set_flux_network): its weights and
    // biases (the caller's device buffers), the bound of its factors, the correction planes it writes (8 per
    // instance: face left of each cell, then face below it) and its tiles
    const void* net_frags = nullptr;
    const float* net_bias = nullptr;
    float net_scale = 0.f;
    double* dnetcorr = nullptr;
    size_t netcorr_doubles = 0;
    NetTask* dnettasks = nullptr;
    int nnettasks = 0;
    unsigned* dnetpeak = nullptr;  // per instance: the largest density seen at an evaluation (float bits)
    std::string err;
    double last_ms = 0.0;
    int device = 0;            // the GPU the handle's buffers and stream live on (current at creation)
    float* scratch = nullptr;  // that GPU's copy of the static staging array (npde_block_average_noalloc)
};

// Flux-network runs switch the network off for an instance whose CFL signal-speed sum grows beyond this
// multiple of its first frame's (npde_set_flux_network_limit).  Of 512 held-out random-Riemann instances
// at 64^2, the corrected flows of the eight slowest reach 1.4-2.2x where they stay healthy; the two that
// ran away (a cell drained toward vacuum at finite pressure) reached 3.5x and 8x and took 1,502 and 1,723
// steps against 821 and 695 plain, holding the whole frame-synchronous batch back.  At 2.5 four
// instances switch off, the batch's density error is unchanged (0.0519) and its rollout takes 0.19
// instead of 0.27 s.
constexpr double kDefaultNetRateLimit = 2.5;

// Staging buffer of npde_block_average_noalloc: a file-scope device array, so that the call allocates
// nothing.  Every GPU the library runs on holds its own copy, at its own address, until the process
// ends; the host-side name `scratch` is not a device address, so each handle keeps its device's address
// (looked up once per device by bind_scratch) and passes that to kernels and copies.
constexpr size_t kNoAllocFloats = size_t(4) * 512 * 512;  // 4 MiB: four 512 x 512 float32 planes
static __device__ __align__(256) float scratch[kNoAllocFloats];

namespace {
std::mutex g_scratch_mu;
std::map<int, float*> g_scratch_of_device;
}  // namespace

namespace {

npde_status fail(npde_solver* s, npde_status st, const std::string& msg) {
    if (s) s->err = msg;
    return st;
}

#define HIPCHK(s, call)                                                                              \
    do {                                                                                             \
        hipError_t e_ = (call);                                                                      \
        if (e_ != hipSuccess)                                                                        \
            return fail((s), NPDE_ERR_DEVICE, std::string(#call ": ") + hipGetErrorString(e_));     \
    } while (0)

// Address of the static staging array on the handle's device, current during npde_batch_create: looked
// up once per device (hipGetSymbolAddress) and shared by every handle on it.
npde_status bind_scratch(npde_solver* s) {
    std::lock_guard<std::mutex> lock(g_scratch_mu);
    const auto it = g_scratch_of_device.find(s->device);
    if (it != g_scratch_of_device.end()) {
        s->scratch = it->second;
        return NPDE_OK;
    }
    void* p = nullptr;
    HIPCHK(s, hipGetSymbolAddress(&p, HIP_SYMBOL(scratch)));
    if (!p) return fail(s, NPDE_ERR_DEVICE, "no device address for the staging array");
    s->scratch = static_cast<float*>(p);
    g_scratch_of_device.emplace(s->device, s->scratch);
    return NPDE_OK;
}

size_t plane_stride(const npde_grid_desc& g) {
    const size_t cells = size_t(g.nx) * g.ny;
    return (cells + 63) / 64 * 64;  // 512-byte aligned planes
}

bool valid_grid(const npde_grid_desc& g, std::string& why) {
    if (g.nx <= 0 || g.ny <= 0) { why = "nx and ny must be positive"; return false; }
    if (double(g.nx) * g.ny >= 2147483647.0) { why = "nx*ny must fit in int32"; return false; }
    if (!(g.x_max > g.x_min) || !(g.y_max > g.y_min)) { why = "empty domain"; return false; }
    if ((g.bc_x != 0 && g.bc_x != 1) || (g.bc_y != 0 && g.bc_y != 1)) { why = "bc must be 0 or 1"; return false; }
    return true;
}

// Band height: the tuned value, or (rows_per_task <= 0) the tallest band in [4, 64] ([2, 64] for flux-
// corrected runs) that still gives about 2048 tasks over the listed instances, so that small grids are not bound by one wavefront
// marching many rows, while bands stay tall enough not to recompute their 2 + 2 edge rows too often
// (measured best or within 5% of it from a single 512^2 instance to 64 x 512^2).  When even 4-row
// bands leave fewer than 512 tasks, the GPU is far from full and 2-row bands are faster (a single
// 256^2 instance: 20.6 against 25.9 us per step).  Bands shorter than 16 rows are doubled while 1024
// tasks remain.
int band_rows(const npde_solver* s, int strip_cols) {
    if (s->tune.rows_per_task > 0) return s->tune.rows_per_task;
    long long rows_total = 0;
    for (int k = 0; k < s->n; ++k)
        if (s->listed[k]) rows_total += (long long)((s->grids[k].nx + strip_cols - 1) / strip_cols) * s->grids[k].ny;
    // flux-corrected stages (learned increments) do more work on every row of a band's serial march, so
    // they take bands down to 2 rows while fewer than 2048 tasks result (64 x 64^2 learned rollout: 33.7
    // against 36.7 ms with 4-row bands; the plain rollout of the same batch gains nothing)
    const int rmin = s->has_corr ? 2 : 4;
    int r = 64;
    while (r > rmin && rows_total / r < 2048) r /= 2;
    if (r == 4 && rows_total / 4 < 512) return 2;
    // bands shorter than 16 rows spend too much of their march on the 2 + 2 edge rows: double them while
    // 1024 tasks remain (1 x 1024^2: 82.4 -> 79.7 us per step; 2 x 512^2: 53.5 -> 50.9; 4 x 512^2 and
    // 16 x 256^2: ~1% faster); from 16 rows on, doubling measured 3-4% slower (2 x 1024^2, 8 x 512^2)
    if (!s->has_corr && r < 16 && rows_total / (2 * r) >= 1024) r *= 2;
    return r;
}

// Row groups apply when every instance is periodic in x with the same width 64*W: then one workgroup
// of W wavefronts covers a band and no halo column is recomputed.  row_groups = 1 (the default) uses
// them for W = 1 (64 columns), where they measure slightly faster than seams (18.1 against 17.8 G
// cell-updates/s on 4096 x 64^2); for W = 2 .. 4 seams measure as fast or faster, and at W = 8 a
// workgroup of 8 wavefronts with 135-164 VGPRs is alone on its CU and waits at every barrier.
// row_groups = 2 allows W <= 8.
int row_group_waves_for(const npde_solver* s, const npde_tuning& tune) {
    if (tune.row_groups <= 0 || s->n == 0) return 0;
    int nx = -1;
    for (int k = 0; k < s->n; ++k) {
        if (!s->listed[k]) continue;
        const auto& g = s->grids[k];
        if (nx < 0) nx = g.nx;
        if (g.nx != nx || g.bc_x != NPDE_BC_PERIODIC) return 0;
    }
    const int wmax = tune.row_groups >= 2 ? kMaxRowWaves : 1;
    if (nx <= 0 || nx % kWave != 0 || nx / kWave > wmax) return 0;
    return nx / kWave;
}
int row_group_waves(const npde_solver* s) { return row_group_waves_for(s, s->tune); }

// Seam mode applies when every instance's width is a multiple of 64, row groups do not apply and bands
// are at most 64 rows (each wavefront keeps its band's seam solves in LDS): each wavefront then owns a
// 64-column strip and solves the two strip boundaries of its band itself before its march, so no halo
// column is recomputed and no wavefront waits for another.  seams = 1 (the default) uses it for
// batches of at least kSeamMinCells cells: the seam solves lengthen every wavefront's serial path,
// which a batch too small to fill the GPU pays in full (a single 256^2 instance: 38 against 28 us per
// step in strips), while from 2^18 cells on seams are ahead (a single 512^2 instance: 37.5 against
// 45.9 us); seams = 2 uses it at any size.
constexpr long long kSeamMinCells = 1 << 18;
bool seams_apply_for(const npde_solver* s, const npde_tuning& tune) {
    if (tune.seams <= 0 || s->n == 0 || row_group_waves_for(s, tune) > 0) return false;
    if (tune.rows_per_task > 64) return false;
    long long cells = 0;
    for (int k = 0; k < s->n; ++k) {
        if (!s->listed[k]) continue;
        const auto& g = s->grids[k];
        if (g.nx % kWave != 0) return false;
        cells += (long long)g.nx * g.ny;
    }
    return tune.seams >= 2 || cells >= kSeamMinCells;
}

// Task list and partial slots.  Strips and seams: one task and one partial per wavefront.  Row groups:
// one task per band (workgroup), W partials per task; D.task0 / D.ntasks index partial slots in every
// form.
void build_tasks(npde_solver* s) {
    s->htasks.clear();
    s->row_waves = row_group_waves(s);
    s->layout = s->row_waves > 0 ? kRowGroups : (seams_apply_for(s, s->tune) ? kSeams : kStrips);
    const int W = s->row_waves;
    const int cols = s->layout == kStrips ? kStripCols : kWave;
    const int rows = band_rows(s, cols);
    s->band = rows;
    for (int k = 0; k < s->n; ++k) {
        InstDesc& D = s->hdesc[k];
        const int first = int(s->htasks.size());
        if (!s->listed[k]) {  // not stepping in this list: no tasks, no partials
            D.task0 = W > 0 ? first * W : first;
            D.ntasks = 0;
            continue;
        }
        if (W > 0) {
            for (int j0 = 0; j0 < D.ny; j0 += rows) s->htasks.push_back({k, 0, j0, std::min(j0 + rows, D.ny)});
            D.task0 = first * W;
            D.ntasks = (int(s->htasks.size()) - first) * W;
        } else {
            const int strips = (D.nx + cols - 1) / cols;
            for (int j0 = 0; j0 < D.ny; j0 += rows)
                for (int st = 0; st < strips; ++st) s->htasks.push_back({k, st * cols, j0, std::min(j0 + rows, D.ny)});
            D.task0 = first;
            D.ntasks = int(s->htasks.size()) - first;
        }
    }
}

size_t partial_slots(const npde_solver* s) { return s->htasks.size() * size_t(s->row_waves > 0 ? s->row_waves : 1); }

npde_status upload_tasks(npde_solver* s) {
    build_tasks(s);
    if (s->dtasks) (void)hipFree(s->dtasks);
    if (s->dpart) (void)hipFree(s->dpart);
    s->dtasks = nullptr;
    s->dpart = nullptr;
    // seam mode, per instance: the edge columns of U and of U1, each a left and a right array of
    // (nx/64) x ny x kEdgeRecord doubles
    auto edge_doubles = [](const InstDesc& D) { return size_t(2 * kEdgeRecord) * (D.nx / kWave) * D.ny; };
    size_t edge_need = 0;
    if (s->layout == kSeams)
        for (const auto& D : s->hdesc) edge_need += 2 * edge_doubles(D);
    if (edge_need > s->edge_cap) {
        if (s->dedge) (void)hipFree(s->dedge);
        s->dedge = nullptr;
        s->edge_cap = 0;
        HIPCHK(s, hipMalloc(&s->dedge, edge_need * sizeof(double)));
        s->edge_cap = edge_need;
    }
    size_t off = 0;
    for (auto& D : s->hdesc) {
        D.edge[0] = D.edge[1] = nullptr;
        if (s->layout != kSeams) continue;
        D.edge[0] = s->dedge + off;
        D.edge[1] = D.edge[0] + edge_doubles(D);
        off += 2 * edge_doubles(D);
    }
    HIPCHK(s, hipMalloc(&s->dtasks, sizeof(Task) * s->htasks.size()));
    HIPCHK(s, hipMalloc(&s->dpart, sizeof(Partial) * partial_slots(s)));
    HIPCHK(s, hipMemcpy(s->dtasks, s->htasks.data(), sizeof(Task) * s->htasks.size(), hipMemcpyHostToDevice));
    HIPCHK(s, hipMemcpy(s->ddesc, s->hdesc.data(), sizeof(InstDesc) * s->n, hipMemcpyHostToDevice));
    return NPDE_OK;
}

npde_status pull_ctl(npde_solver* s) {
    HIPCHK(s, hipStreamSynchronize(s->stream));
    HIPCHK(s, hipMemcpy(s->hctl.data(), s->dctl, sizeof(Ctl) * s->n, hipMemcpyDeviceToHost));
    return NPDE_OK;
}

npde_status push_ctl(npde_solver* s) {
    HIPCHK(s, hipMemcpy(s->dctl, s->hctl.data(), sizeof(Ctl) * s->n, hipMemcpyHostToDevice));
    return NPDE_OK;
}

void drop_graph(npde_solver* s) {
    for (auto& row : s->graphs)
        for (auto& g : row) {
            if (g) (void)hipGraphExecDestroy(g);
            g = nullptr;
        }
    s->graph_key_frames = -1;
}

int loop_iters(const npde_solver* s, int level) {
    const long long K = std::max(1, s->tune.steps_per_graph);
    return int(std::min(1LL << level, K));
}

// The device-loop size for a run that expects `steps` more steps of its slowest instance (0: unknown).
// Each launch ends in a host round trip, and iterations past every instance's last step still launch
// kernels that return at once, so: the next power of two when that wastes at most two iterations,
// otherwise the largest power of two that does not overshoot (the remainder takes further launches).
int loop_level(const npde_solver* s, long long steps) {
    const int K = std::max(1, s->tune.steps_per_graph);
    int top = 0;
    while (loop_iters(s, top) < K) ++top;
    if (steps <= 0 || steps >= K) return top;
    int up = 0;
    while ((1LL << up) < steps) ++up;
    if ((1LL << up) - steps <= 2) return std::min(up, top);
    return std::min(up - 1, top);
}

// Steps the slowest stepping instance still needs to reach its end time at its current CFL step, from the
// host copy of the run state: fresh after a launch of this run (then each instance's own end time, step
// budget and done flag count), or left from the previous call at the start of a run (then t_end and
// max_new_steps are the new run's).  dt changes as the state evolves, so this only sizes launches and
// never changes a result.  0 = unknown.
long long expected_steps(const npde_solver* s, bool fresh, double t_end, long long max_new_steps) {
    long long est = 0;
    for (int k = 0; k < s->n; ++k) {
        if (!s->listed[k]) continue;
        const Ctl& c = s->hctl[k];
        if (fresh && c.done) continue;
        const InstDesc& D = s->hdesc[k];
        const double rate = c.sx / D.dx + c.sy / D.dy;
        if (!(rate > 0.0) || !std::isfinite(rate)) return 0;
        const double need = std::ceil(((fresh ? c.t_end : t_end) - c.t) * rate / D.cfl);
        long long st = need >= 1.0 ? (need < 1e12 ? (long long)need : (long long)1e12) : 1;
        if (fresh) st = std::min(st, std::max(1LL, c.step_limit - c.steps));
        else if (max_new_steps >= 0) st = std::min(st, std::max(1LL, max_new_steps));
        est = std::max(est, st);
    }
    return est;
}

// Steps until the first stepping instance of a recording run lands on its next frame time (its target
// T(k), main2d.cpp:224-246) at its current CFL step, from the run state read after the last launch:
// 0 when a frame waits to be written, -1 when unknown.  Launches shorter than that can leave the frame
// kernels out.  An instance that lands during such a launch anyway keeps its frame pending and takes no
// step until a launch with the frame kernels writes it, so the estimate never changes a result.
long long steps_to_frame(const npde_solver* s) {
    long long best = -1;
    for (int k = 0; k < s->n; ++k) {
        if (!s->listed[k]) continue;
        const Ctl& c = s->hctl[k];
        if (c.done) continue;
        if (c.pending >= 0 || c.hold) return 0;  // a frame (or the batch's network call) waits
        if (!c.record || c.nout <= 0) return -1;
        const double T = c.k == 0 ? c.t_start : c.t_start + (c.t_end - c.t_start) * double(c.k) / double(c.nout);
        const InstDesc& D = s->hdesc[k];
        const double rate = c.sx / D.dx + c.sy / D.dy;
        if (!(rate > 0.0) || !std::isfinite(rate)) return -1;
        const double need = std::ceil((T - c.t) * rate / D.cfl);
        const long long st = need >= 1.0 ? (need < 1e12 ? (long long)need : (long long)1e12) : 1;
        best = best < 0 ? st : std::min(best, st);
    }
    return best;
}

// The task list for the instances in mask (band height and layout chosen for them), and back to all.
npde_status list_instances(npde_solver* s, const std::vector<char>& mask) {
    s->listed = mask;
    s->subset = std::find(mask.begin(), mask.end(), 0) != mask.end();
    drop_graph(s);
    return upload_tasks(s);
}
npde_status list_all(npde_solver* s) {
    if (!s->subset) return NPDE_OK;
    return list_instances(s, std::vector<char>(s->n, 1));
}

// to_primitive on the host (same operations as the device), for error messages
Prim host_prim(const double* c) { return prim_of(c[0], c[1], c[2], c[3]); }

npde_status fill_bad(npde_solver* s, npde_bad_cell* bad) {
    for (int k = 0; k < s->n; ++k) {
        const Ctl& c = s->hctl[k];
        if (c.error || (c.mode == kModeHost && c.bad != kNoBad)) {
            const InstDesc& D = s->hdesc[k];
            const int i = c.bad % D.nx, j = c.bad / D.nx;
            double cell[4];
            for (int q = 0; q < 4; ++q)
                (void)hipMemcpy(&cell[q], D.U + size_t(q) * D.plane + c.bad, sizeof(double), hipMemcpyDeviceToHost);
            const Prim w = host_prim(cell);
            if (bad) *bad = {k, i, j, w.d, w.p};
            char buf[256];
            std::snprintf(buf, sizeof buf, "Non-physical state in instance %d, cell (%d, %d): rho = %g, p = %g", k,
                          i, j, w.d, w.p);
            return fail(s, NPDE_ERR_NONPHYSICAL, buf);
        }
    }
    return NPDE_OK;
}

// The occupancy cap (tune.waves_per_simd > 0): the LDS bytes each stage workgroup holds so that a CU fits
// at most 4 * cap / (wavefronts per workgroup) workgroups -- at most `cap` wavefronts on each of its 4
// SIMDs -- or 0.  w workgroups fit and w + 1 do not when each holds floor(LDS / w) rounded down to 512
// bytes (for w <= 17).  Reservations above kMaxLaunchLds are refused by npde_set_tuning.
constexpr unsigned kMaxLaunchLds = 65536;
unsigned lds_reserve_for(const npde_solver* s, const npde_tuning& t, int row_waves) {
    if (t.waves_per_simd <= 0) return 0;
    const int wg = row_waves > 0 ? row_waves : t.waves_per_block;
    const int fit = std::max(1, 4 * t.waves_per_simd / wg);
    return unsigned(s->lds_per_cu / fit) & ~511u;
}
unsigned lds_reserve(const npde_solver* s) { return lds_reserve_for(s, s->tune, s->row_waves); }

// Profiling: read the recorded launches' event pairs into the per-kind totals.
npde_status prof_harvest(npde_solver* s) {
    const size_t n = s->prof_kind.size();
    if (n == 0) return NPDE_OK;
    HIPCHK(s, hipEventSynchronize(s->prof_events[2 * n - 1]));
    for (size_t i = 0; i < n; ++i) {
        float ms = 0.f;
        HIPCHK(s, hipEventElapsedTime(&ms, s->prof_events[2 * i], s->prof_events[2 * i + 1]));
        s->prof_ms[s->prof_kind[i]] += ms;
        s->prof_n[s->prof_kind[i]] += 1;
    }
    s->prof_kind.clear();
    return NPDE_OK;
}

// A kernel launch (launch() returns hipError_t), between two events when profiling is on.  Profiled runs
// launch on the stream, never under graph capture.
constexpr size_t kProfBatch = 4096;  // recorded launches before the host reads them
template <class F>
npde_status timed_launch(npde_solver* s, int kind, F&& launch) {
    if (!s->prof_on) {
        HIPCHK(s, launch());
        return NPDE_OK;
    }
    if (s->prof_kind.size() >= kProfBatch) {
        npde_status st = prof_harvest(s);
        if (st != NPDE_OK) return st;
    }
    const size_t i = s->prof_kind.size();
    while (s->prof_events.size() < 2 * i + 2) {
        hipEvent_t e = nullptr;
        HIPCHK(s, hipEventCreate(&e));
        s->prof_events.push_back(e);
    }
    HIPCHK(s, hipEventRecord(s->prof_events[2 * i], s->stream));
    HIPCHK(s, launch());
    HIPCHK(s, hipEventRecord(s->prof_events[2 * i + 1], s->stream));
    s->prof_kind.push_back(kind);
    return NPDE_OK;
}

// One stage launch.  In seam mode the wavefronts' seam solves read the stage input's edge columns from
// the edge arrays, except in stage 1 of the first step of a launch sequence (first_step: a host call, a
// graph, an iteration loop), whose input the caller may have changed since a stage 2 last wrote them.
// With a flux network, stage 1 always reads the state itself: an instance's first step of a run waits
// at its first frame for the network, which need not run in a launch's first iteration, so that step
// can come later in the launch, before any stage 2 has written the edge arrays of the new state.
npde_status launch_stage_kernel(npde_solver* s, int stage, bool first_step) {
    const int compact = (stage == 2 || (!first_step && s->net_frags == nullptr)) ? 1 : 0;
    return timed_launch(s, stage == 1 ? NPDE_PROF_STAGE1 : NPDE_PROF_STAGE2, [&] {
        return launch_stage(s->flux, s->lim, stage, s->has_corr, s->fast, s->tune.waves_per_block, s->layout,
                            s->row_waves, compact, s->band, lds_reserve(s), s->ddesc, s->dctl, s->dtasks,
                            int(s->htasks.size()), s->dpart, s->dndone, s->stream);
    });
}

npde_status launch_finish_kernel(npde_solver* s, int mode) {
    return timed_launch(s, NPDE_PROF_FINISH, [&] {
        return launch_finish(s->ddesc, s->dctl, s->dpart, s->n, mode, s->dndone, s->stream);
    });
}

// One RK2 step of every stepping instance: two stage launches, then the per-instance reduction of the
// fused wave-speed partials and the run-state decision.
npde_status launch_step_pair(npde_solver* s, bool first_step) {
    for (int stage = 1; stage <= 2; ++stage) {
        npde_status st = launch_stage_kernel(s, stage, first_step);
        if (st != NPDE_OK) return st;
    }
    return launch_finish_kernel(s, kAfterStep);
}

npde_status launch_wavespeed_and_finish(npde_solver* s) {
    npde_status st = timed_launch(s, NPDE_PROF_WAVESPEED, [&] {
        return launch_wavespeed(s->tune.waves_per_block, s->layout, s->row_waves, s->ddesc, s->dctl, s->dtasks,
                                int(s->htasks.size()), s->dpart, s->dndone, s->stream);
    });
    if (st != NPDE_OK) return st;
    return launch_finish_kernel(s, kAfterWavespeed);
}

// `iters` iterations of the device loop.  Without frames: stage 1, stage 2, k_finish.  With frames (nft
// frame tasks): the frame kernel, which completes the previous iteration's step (in place of its k_finish;
// not in the first iteration, whose predecessor is a k_finish: a step it decided is not taken yet) and
// writes the frames of the instances that landed, [the network,] stage 1, stage 2 -- and one k_finish after
// the last iteration, so that every launch sequence ends with its steps completed.
npde_status launch_iterations(npde_solver* s, int iters, int nft) {
    for (int it = 0; it < iters; ++it) {
        if (nft > 0) {
            const int hold = s->net_frags ? 1 : 0;
            npde_status st = timed_launch(s, NPDE_PROF_FRAME, [&] {
                return launch_frame(s->ddesc, s->dctl, s->dpart, s->dftasks, nft, it > 0 ? 1 : 0, hold, s->dndone,
                                    s->stream);
            });
            // flux-network run: an instance that reached its frame holds there; once every running instance
            // holds, the network turns their states into the next interval's factors (one launch for the
            // batch), and its last workgroup lets them all move on.  The network kernel is launched in every
            // net_period-th iteration, counted back from the launch's last one (so every launch, of any
            // length, can release a held batch): a batch that holds waits at most net_period - 1 iterations
            // in which nothing steps, and most iterations are spared a launch that would find the gate shut.
            const int period = std::max(1, s->tune.net_period);
            if (st == NPDE_OK && hold && (iters - 1 - it) % period == 0)
                st = timed_launch(s, NPDE_PROF_NETWORK, [&] {
                    return launch_flux_net(s->ddesc, s->dctl, s->dnettasks, s->nnettasks, s->n, s->dndone,
                                           s->net_frags, s->net_bias, s->net_scale, s->dnetpeak, s->stream);
                });
            for (int stage = 1; stage <= 2 && st == NPDE_OK; ++stage) st = launch_stage_kernel(s, stage, it == 0);
            if (st != NPDE_OK) return st;
        } else {
            npde_status st = launch_step_pair(s, it == 0);
            if (st != NPDE_OK) return st;
        }
    }
    if (nft > 0 && iters > 0) return launch_finish_kernel(s, kAfterStep);
    return NPDE_OK;
}

// frames: capture the iterations with the frame kernels (nft > 0) or without them.
npde_status ensure_graph(npde_solver* s, int nft, int level, bool frames) {
    const int K = std::max(1, s->tune.steps_per_graph);
    if (!(s->graph_key_frames == nft && s->graph_key_steps == K && s->graph_key_corr == int(s->has_corr) &&
          s->graph_key_rows == s->tune.rows_per_task && s->graph_key_wpb == s->tune.waves_per_block &&
          s->graph_key_fast == int(s->fast) && s->graph_key_net == int(s->net_frags != nullptr))) {
        drop_graph(s);
        s->graph_key_frames = nft;
        s->graph_key_steps = K;
        s->graph_key_corr = int(s->has_corr);
        s->graph_key_rows = s->tune.rows_per_task;
        s->graph_key_wpb = s->tune.waves_per_block;
        s->graph_key_fast = int(s->fast);
        s->graph_key_net = int(s->net_frags != nullptr);
    }
    hipGraphExec_t& slot = s->graphs[frames ? 1 : 0][level];
    if (slot) return NPDE_OK;
    hipGraph_t g = nullptr;
    HIPCHK(s, hipStreamBeginCapture(s->stream, hipStreamCaptureModeThreadLocal));
    npde_status st = launch_iterations(s, loop_iters(s, level), frames ? nft : 0);
    hipError_t e = hipStreamEndCapture(s->stream, &g);
    if (st != NPDE_OK) {
        if (g) (void)hipGraphDestroy(g);
        return st;
    }
    if (e != hipSuccess) return fail(s, NPDE_ERR_DEVICE, std::string("graph capture: ") + hipGetErrorString(e));
    e = hipGraphInstantiate(&slot, g, nullptr, nullptr, 0);
    (void)hipGraphDestroy(g);
    if (e != hipSuccess) {
        slot = nullptr;
        return fail(s, NPDE_ERR_DEVICE, std::string("graph instantiate: ") + hipGetErrorString(e));
    }
    return NPDE_OK;
}

// Tail of a device-resident run: most instances have finished, and the bands chosen for the whole batch
// leave the rest latency-bound (4 x 256^2 instances step in 36 us with 4-row bands, 232 us with 64-row
// bands).  Rebuild the task list -- band height and layout chosen again -- for the instances still
// stepping, when their remaining work is worth a new graph: more than two graph launches at their current
// dt.  listed_n becomes the number still stepping either way, so the next look comes when half of them
// have finished.  Every layout gives the same bits.  The host copy of the run state is the one read
// after the last launch; the graphs of the new list are captured when first launched.
npde_status retile_tail(npde_solver* s, int& listed_n) {
    std::vector<char> mask(s->n, 0);
    int active = 0;
    double remaining = 0.0;
    for (int k = 0; k < s->n; ++k) {
        const Ctl& c = s->hctl[k];
        if (c.done) continue;
        mask[k] = 1;
        ++active;
        double steps = c.dt > 0.0 ? (c.t_end - c.t) / c.dt : 0.0;
        steps = std::min(steps, double(c.step_limit - c.steps));
        remaining = std::max(remaining, steps);
    }
    listed_n = active;
    const int K = std::max(1, s->tune.steps_per_graph);
    if (active == 0 || remaining < 2.0 * K) return NPDE_OK;
    return list_instances(s, mask);
}

// Device-resident loop shared by npde_advance_to and npde_run_trajectory.
npde_status auto_run(npde_solver* s, double t_start, double t_end, int nout, bool record, npde_bad_cell* bad,
                     long long max_new_steps = -1) {
    npde_status st = list_all(s);
    if (st != NPDE_OK) return st;
    const int nft = record ? int(s->hftasks.size()) : 0;
    // The first launch is sized from the run state the host already holds (the end of the previous call),
    // the later ones from the run state read back after each launch.  The run state is set on the device:
    // no host round trip before the run.  The first launch's graph is captured on an idle stream.
    long long expect = expected_steps(s, false, t_end, max_new_steps);
    int level = loop_level(s, expect);
    bool frames = nft > 0;  // the frame kernels in this launch
    const bool graphs = s->tune.use_graph && !s->prof_on;  // a profiled run launches on the stream
    if (graphs && (st = ensure_graph(s, nft, level, frames)) != NPDE_OK) return st;
    HIPCHK(s, launch_begin_run(s->dctl, s->n, t_start, t_end, nout, record ? 1 : 0, max_new_steps, s->dndone,
                               s->stream));
    HIPCHK(s, hipEventRecord(s->ev0, s->stream));
    if ((st = launch_wavespeed_and_finish(s)) != NPDE_OK) return st;
    int done = 0;
    int listed_n = s->n;  // instances in the current task list
    // A state whose wave speed is infinite gives dt = 0 and would never reach its target (the reference
    // loops forever there too); bound the loop instead.
    long long iters = 0;
    for (;;) {
        if (iters >= 50000000LL) return fail(s, NPDE_ERR_STATE, "run did not finish within 5e7 steps");
        if (graphs) {
            if ((st = ensure_graph(s, nft, level, frames)) != NPDE_OK) return st;
            HIPCHK(s, hipGraphLaunch(s->graphs[frames ? 1 : 0][level], s->stream));
        } else if ((st = launch_iterations(s, loop_iters(s, level), frames ? nft : 0)) != NPDE_OK) {
            return st;
        }
        iters += loop_iters(s, level);
        HIPCHK(s, hipMemcpyAsync(&done, s->dndone, sizeof(int), hipMemcpyDeviceToHost, s->stream));
        HIPCHK(s, hipMemcpyAsync(s->hctl.data(), s->dctl, sizeof(Ctl) * s->n, hipMemcpyDeviceToHost, s->stream));
        HIPCHK(s, hipStreamSynchronize(s->stream));
        if ((st = prof_harvest(s)) != NPDE_OK) return st;
        if (done >= s->n) break;
        if (2 * (s->n - done) <= listed_n && listed_n > 1 && (st = retile_tail(s, listed_n)) != NPDE_OK)
            return st;
        expect = expected_steps(s, true, t_end, max_new_steps);
        level = loop_level(s, expect);
        frames = nft > 0;
        if (frames) {
            // Every stepping instance is at least 8 steps from its next frame time (one stays near one
            // almost always in a large batch, never in a single-instance run): run the steps before it
            // without the frame kernel and its finish, 3 launches per step instead of 5.
            const long long tf = steps_to_frame(s);
            if (tf >= 8) {
                frames = false;
                level = loop_level(s, std::min(expect > 0 ? expect : tf - 2, tf - 2));
            }
        }
    }
    HIPCHK(s, hipEventRecord(s->ev1, s->stream));
    HIPCHK(s, hipEventSynchronize(s->ev1));
    float ms = 0.f;
    HIPCHK(s, hipEventElapsedTime(&ms, s->ev0, s->ev1));
    s->last_ms = ms;
    if ((st = list_all(s)) != NPDE_OK) return st;
    return fill_bad(s, bad);
}

npde_status set_host_mode(npde_solver* s) {
    npde_status st = pull_ctl(s);
    if (st != NPDE_OK) return st;
    for (auto& c : s->hctl) {
        c.mode = kModeHost;
        c.done = 0;
        c.error = 0;
        c.active = 0;
        c.pending = -1;
        c.hold = 0;
    }
    return push_ctl(s);
}

void fill_ghosts(const npde_grid_desc& g, double* A) {
    // A: (ny+4) x (nx+4) x 4, physical cells already set; x ghosts on physical rows, then y ghosts
    // over all columns (so the corners are filled too), as grid2d.hpp:43-66 does.
    const int ng = 2, W = g.nx + 2 * ng;
    auto at = [&](int i, int j) { return A + (size_t(j + ng) * W + (i + ng)) * 4; };
    for (int j = 0; j < g.ny; ++j)
        for (int k = 1; k <= ng; ++k) {
            const int il = g.bc_x == kPeriodic ? g.nx - k : 0;
            const int ir = g.bc_x == kPeriodic ? k - 1 : g.nx - 1;
            std::memcpy(at(-k, j), at(il, j), 32);
            std::memcpy(at(g.nx - 1 + k, j), at(ir, j), 32);
        }
    for (int i = -ng; i < g.nx + ng; ++i)
        for (int k = 1; k <= ng; ++k) {
            const int jb = g.bc_y == kPeriodic ? g.ny - k : 0;
            const int jt = g.bc_y == kPeriodic ? k - 1 : g.ny - 1;
            std::memcpy(at(i, -k), at(i, jb), 32);
            std::memcpy(at(i, g.ny - 1 + k), at(i, jt), 32);
        }
}

}  // namespace

extern "C" {

npde_status npde_batch_create(int n, const npde_grid_desc* grids, int flux, int limiter, double cfl,
                              npde_solver** out) {
    if (!out || !grids || n <= 0) return NPDE_ERR_INVALID;
    *out = nullptr;
    if (flux != kRusanov && flux != kHllc) return NPDE_ERR_INVALID;
    if (limiter < kLimNone || limiter > kLimVanLeer) return NPDE_ERR_INVALID;
    if (!(cfl > 0.0)) return NPDE_ERR_INVALID;
    std::string why;
    for (int k = 0; k < n; ++k)
        if (!valid_grid(grids[k], why)) return NPDE_ERR_INVALID;
    auto* s = new npde_solver;
    s->n = n;
    s->flux = flux;
    s->lim = limiter;
    s->cfl = cfl;
    s->grids.assign(grids, grids + n);
    s->hdesc.resize(n);
    s->listed.assign(n, 1);
    s->hctl.assign(n, Ctl{});
    s->state_off.resize(n);
    size_t total = 0;
    for (int k = 0; k < n; ++k) {
        s->state_off[k] = total;
        total += 8 * plane_stride(grids[k]);
    }
    auto cleanup = [&](npde_status st) {
        npde_destroy(s);
        return st;
    };
    if (hipGetDevice(&s->device) != hipSuccess) return cleanup(NPDE_ERR_DEVICE);
    if (hipDeviceGetAttribute(&s->lds_per_cu, hipDeviceAttributeMaxSharedMemoryPerMultiprocessor, s->device) !=
            hipSuccess || s->lds_per_cu <= 0)
        return cleanup(NPDE_ERR_DEVICE);
    if (bind_scratch(s) != NPDE_OK) return cleanup(NPDE_ERR_DEVICE);
    if (hipStreamCreateWithFlags(&s->stream, hipStreamNonBlocking) != hipSuccess) return cleanup(NPDE_ERR_DEVICE);
    if (hipEventCreate(&s->ev0) != hipSuccess || hipEventCreate(&s->ev1) != hipSuccess)
        return cleanup(NPDE_ERR_DEVICE);
    if (hipMalloc(&s->dstate, total * sizeof(double)) != hipSuccess) return cleanup(NPDE_ERR_NOMEM);
    if (hipMemset(s->dstate, 0, total * sizeof(double)) != hipSuccess) return cleanup(NPDE_ERR_DEVICE);
    if (hipMalloc(&s->ddesc, sizeof(InstDesc) * n) != hipSuccess || hipMalloc(&s->dctl, sizeof(Ctl) * n) != hipSuccess ||
        hipMalloc(&s->dndone, sizeof(int) * kRunCounters) != hipSuccess || hipMalloc(&s->ddts, sizeof(double) * n) != hipSuccess)
        return cleanup(NPDE_ERR_NOMEM);
    for (int k = 0; k < n; ++k) {
        const npde_grid_desc& g = grids[k];
        InstDesc& D = s->hdesc[k];
        std::memset(&D, 0, sizeof D);
        const size_t ps = plane_stride(g);
        D.U = s->dstate + s->state_off[k];
        D.U1 = D.U + 4 * ps;
        D.plane = (long long)ps;
        D.nx = g.nx;
        D.ny = g.ny;
        D.bc_x = g.bc_x;
        D.bc_y = g.bc_y;
        D.dx = (g.x_max - g.x_min) / g.nx;  // make_grid (grid2d.hpp:37-38)
        D.dy = (g.y_max - g.y_min) / g.ny;
        D.cx = 1.0 / D.dx;
        D.cy = 1.0 / D.dy;
        D.cfl = cfl;
        D.net_rate_limit = kDefaultNetRateLimit;
        Ctl& c = s->hctl[k];
        std::memset(&c, 0, sizeof c);
        c.bad = kNoBad;
        c.pending = -1;
        c.mode = kModeHost;
    }
    if (upload_tasks(s) != NPDE_OK) return cleanup(NPDE_ERR_DEVICE);
    if (push_ctl(s) != NPDE_OK) return cleanup(NPDE_ERR_DEVICE);
    *out = s;
    return NPDE_OK;
}

npde_status npde_create(const npde_grid_desc* grid, int flux, int limiter, double cfl, npde_solver** out) {
    return npde_batch_create(1, grid, flux, limiter, cfl, out);
}

npde_status npde_destroy(npde_solver* s) {
    if (!s) return NPDE_OK;
    if (s->stream) (void)hipStreamSynchronize(s->stream);
    drop_graph(s);
    for (void* p : {(void*)s->ddesc, (void*)s->dctl, (void*)s->dtasks, (void*)s->dedge, (void*)s->dnetcorr,
                    (void*)s->dnettasks, (void*)s->dnetpeak,
                    (void*)s->dftasks, (void*)s->dpart, (void*)s->dndone, (void*)s->ddts, (void*)s->dstate,
                    (void*)s->dtraj, (void*)s->dftimes})
        if (p) (void)hipFree(p);
    for (hipEvent_t e : s->prof_events) (void)hipEventDestroy(e);
    if (s->ev0) (void)hipEventDestroy(s->ev0);
    if (s->ev1) (void)hipEventDestroy(s->ev1);
    if (s->stream) (void)hipStreamDestroy(s->stream);
    delete s;
    return NPDE_OK;
}

int npde_num_instances(const npde_solver* s) { return s ? s->n : 0; }

const char* npde_last_error(const npde_solver* s) { return s ? s->err.c_str() : "null handle"; }

npde_status npde_set_state(npde_solver* s, int inst, const double* U, int layout) {
    if (!s || !U || inst < 0 || inst >= s->n) return fail(s, NPDE_ERR_INVALID, "bad instance or null state");
    const npde_grid_desc& g = s->grids[inst];
    const size_t cells = size_t(g.nx) * g.ny;
    const InstDesc& D = s->hdesc[inst];
    npde_status st = pull_ctl(s);
    if (st != NPDE_OK) return st;
    if (layout == NPDE_LAYOUT_PLANES) {
        for (int q = 0; q < 4; ++q)
            HIPCHK(s, hipMemcpy(D.U + q * D.plane, U + q * cells, cells * sizeof(double), hipMemcpyHostToDevice));
    } else if (layout == NPDE_LAYOUT_GHOSTED) {
        std::vector<double> tmp(4 * cells);
        const int W = g.nx + 4;
        for (int j = 0; j < g.ny; ++j)
            for (int i = 0; i < g.nx; ++i)
                for (int q = 0; q < 4; ++q)
                    tmp[q * cells + size_t(j) * g.nx + i] = U[(size_t(j + 2) * W + (i + 2)) * 4 + q];
        for (int q = 0; q < 4; ++q)
            HIPCHK(s, hipMemcpy(D.U + q * D.plane, tmp.data() + q * cells, cells * sizeof(double),
                                hipMemcpyHostToDevice));
    } else {
        return fail(s, NPDE_ERR_INVALID, "unknown layout");
    }
    Ctl& c = s->hctl[inst];
    c.t = 0.0;
    c.steps = 0;
    c.active = 0;
    c.bad = kNoBad;
    c.error = 0;
    c.done = 0;
    c.hold = 0;
    return push_ctl(s);
}

npde_status npde_get_state(npde_solver* s, int inst, double* U, int layout) {
    if (!s || !U || inst < 0 || inst >= s->n) return fail(s, NPDE_ERR_INVALID, "bad instance or null state");
    HIPCHK(s, hipStreamSynchronize(s->stream));
    const npde_grid_desc& g = s->grids[inst];
    const size_t cells = size_t(g.nx) * g.ny;
    const InstDesc& D = s->hdesc[inst];
    if (layout == NPDE_LAYOUT_PLANES) {
        for (int q = 0; q < 4; ++q)
            HIPCHK(s, hipMemcpy(U + q * cells, D.U + q * D.plane, cells * sizeof(double), hipMemcpyDeviceToHost));
        return NPDE_OK;
    }
    if (layout != NPDE_LAYOUT_GHOSTED) return fail(s, NPDE_ERR_INVALID, "unknown layout");
    std::vector<double> tmp(4 * cells);
    for (int q = 0; q < 4; ++q)
        HIPCHK(s, hipMemcpy(tmp.data() + q * cells, D.U + q * D.plane, cells * sizeof(double), hipMemcpyDeviceToHost));
    const int W = g.nx + 4;
    std::fill(U, U + size_t(W) * (g.ny + 4) * 4, 0.0);
    for (int j = 0; j < g.ny; ++j)
        for (int i = 0; i < g.nx; ++i)
            for (int q = 0; q < 4; ++q) U[(size_t(j + 2) * W + (i + 2)) * 4 + q] = tmp[q * cells + size_t(j) * g.nx + i];
    fill_ghosts(g, U);
    return NPDE_OK;
}

npde_status npde_set_state_device(npde_solver* s, int inst, const double* dU) {
    if (!s || !dU || inst < 0 || inst >= s->n) return fail(s, NPDE_ERR_INVALID, "bad instance or null state");
    const npde_grid_desc& g = s->grids[inst];
    const size_t cells = size_t(g.nx) * g.ny;
    const InstDesc& D = s->hdesc[inst];
    npde_status st = pull_ctl(s);
    if (st != NPDE_OK) return st;
    // The handle's stream does not wait for the legacy stream, so the caller must have finished writing dU
    // (npde.h); the copy is complete when the call returns, so the caller may free or reuse dU at once.
    HIPCHK(s, hipMemcpy2DAsync(D.U, D.plane * sizeof(double), dU, cells * sizeof(double), cells * sizeof(double), 4,
                               hipMemcpyDeviceToDevice, s->stream));
    HIPCHK(s, hipStreamSynchronize(s->stream));
    Ctl& c = s->hctl[inst];
    c.t = 0.0;
    c.steps = 0;
    c.active = 0;
    c.bad = kNoBad;
    c.error = 0;
    c.done = 0;
    c.hold = 0;
    return push_ctl(s);
}

npde_status npde_get_state_device(npde_solver* s, int inst, double* dU) {
    if (!s || !dU || inst < 0 || inst >= s->n) return fail(s, NPDE_ERR_INVALID, "bad instance or null state");
    const npde_grid_desc& g = s->grids[inst];
    const size_t cells = size_t(g.nx) * g.ny;
    const InstDesc& D = s->hdesc[inst];
    HIPCHK(s, hipMemcpy2DAsync(dU, cells * sizeof(double), D.U, D.plane * sizeof(double), cells * sizeof(double), 4,
                               hipMemcpyDeviceToDevice, s->stream));
    HIPCHK(s, hipStreamSynchronize(s->stream));
    return NPDE_OK;
}

npde_status npde_get_time(npde_solver* s, int inst, double* t, long long* steps) {
    if (!s || inst < 0 || inst >= s->n) return fail(s, NPDE_ERR_INVALID, "bad instance");
    npde_status st = pull_ctl(s);
    if (st != NPDE_OK) return st;
    if (t) *t = s->hctl[inst].t;
    if (steps) *steps = s->hctl[inst].steps;
    return NPDE_OK;
}

npde_status npde_get_times(npde_solver* s, double* t, long long* steps) {
    if (!s || (!t && !steps)) return NPDE_ERR_INVALID;
    npde_status st = pull_ctl(s);
    if (st != NPDE_OK) return st;
    for (int k = 0; k < s->n; ++k) {
        if (t) t[k] = s->hctl[k].t;
        if (steps) steps[k] = s->hctl[k].steps;
    }
    return NPDE_OK;
}

npde_status npde_state_pointer(npde_solver* s, int inst, double** U, long long* plane, long long* instance_stride) {
    if (!s || inst < 0 || inst >= s->n || !U) return fail(s, NPDE_ERR_INVALID, "bad arguments");
    *U = s->hdesc[inst].U;
    if (plane) *plane = s->hdesc[inst].plane;
    if (instance_stride)
        *instance_stride = inst + 1 < s->n ? (long long)(s->state_off[inst + 1] - s->state_off[inst])
                                           : 8 * s->hdesc[inst].plane;
    return NPDE_OK;
}

npde_status npde_compute_dt(npde_solver* s, double* dt, npde_bad_cell* bad) {
    if (!s || !dt) return fail(s, NPDE_ERR_INVALID, "null argument");
    npde_status st = list_all(s);
    if (st == NPDE_OK) st = set_host_mode(s);
    if (st != NPDE_OK) return st;
    if ((st = launch_wavespeed_and_finish(s)) != NPDE_OK) return st;
    if ((st = pull_ctl(s)) != NPDE_OK) return st;
    for (int k = 0; k < s->n; ++k) {
        const Ctl& c = s->hctl[k];
        const InstDesc& D = s->hdesc[k];
        dt[k] = D.cfl / (c.sx / D.dx + c.sy / D.dy);  // solver2d.hpp:82
    }
    return fill_bad(s, bad);
}

// One host-driven step with the given dts, timed into s->last_ms.  first_step: the first of a sequence
// of such steps (the caller may have changed the state before it).
static npde_status host_step(npde_solver* s, const double* dt, bool first_step) {
    if (!s || !dt) return fail(s, NPDE_ERR_INVALID, "null argument");
    npde_status st = list_all(s);
    if (st == NPDE_OK) st = set_host_mode(s);
    if (st != NPDE_OK) return st;
    HIPCHK(s, hipMemcpyAsync(s->ddts, dt, sizeof(double) * s->n, hipMemcpyHostToDevice, s->stream));
    HIPCHK(s, launch_prep_host_step(s->dctl, s->n, s->ddts, s->stream));
    HIPCHK(s, hipEventRecord(s->ev0, s->stream));
    if ((st = launch_step_pair(s, first_step)) != NPDE_OK) return st;
    HIPCHK(s, hipEventRecord(s->ev1, s->stream));
    HIPCHK(s, hipEventSynchronize(s->ev1));
    float ms = 0.f;
    HIPCHK(s, hipEventElapsedTime(&ms, s->ev0, s->ev1));
    s->last_ms = ms;
    return NPDE_OK;
}

npde_status npde_step(npde_solver* s, const double* dt) { return host_step(s, dt, true); }

npde_status npde_advance_to(npde_solver* s, double t_target, npde_bad_cell* bad) {
    if (!s) return NPDE_ERR_INVALID;
    return auto_run(s, t_target, t_target, 0, false, bad);
}

npde_status npde_advance_steps(npde_solver* s, long long nsteps, double t_target, npde_bad_cell* bad) {
    if (!s || nsteps < 0) return fail(s, NPDE_ERR_INVALID, "bad step count");
    return auto_run(s, t_target, t_target, 0, false, bad, nsteps);
}

npde_status npde_time_stages(npde_solver* s, int nsteps, double* ms_stage1, double* ms_stage2) {
    if (!s || nsteps <= 0 || !ms_stage1 || !ms_stage2) return fail(s, NPDE_ERR_INVALID, "bad arguments");
    npde_status st = list_all(s);
    if (st != NPDE_OK) return st;
    // The steps of npde_advance_steps toward an end time they never reach, so each takes the CFL dt of its
    // own state, launched one at a time with events around the two stage kernels.  The first opens the
    // launch sequence, as in the device loop.
    const double t_open = 1e300;
    HIPCHK(s, launch_begin_run(s->dctl, s->n, t_open, t_open, 0, 0, nsteps, s->dndone, s->stream));
    if ((st = launch_wavespeed_and_finish(s)) != NPDE_OK) return st;
    hipEvent_t e[3] = {};
    hipError_t he = hipSuccess;
    for (auto& x : e)
        if (he == hipSuccess) he = hipEventCreate(&x);
    double acc[2] = {0.0, 0.0};
    for (int k = 0; k < nsteps && st == NPDE_OK && he == hipSuccess; ++k) {
        for (int stage = 1; stage <= 2 && st == NPDE_OK && he == hipSuccess; ++stage) {
            he = hipEventRecord(e[stage - 1], s->stream);
            if (he == hipSuccess) st = launch_stage_kernel(s, stage, k == 0);
        }
        if (st == NPDE_OK && he == hipSuccess) he = hipEventRecord(e[2], s->stream);
        if (st == NPDE_OK && he == hipSuccess) st = launch_finish_kernel(s, kAfterStep);
        if (st == NPDE_OK && he == hipSuccess) he = hipEventSynchronize(e[2]);
        for (int q = 0; q < 2 && st == NPDE_OK && he == hipSuccess; ++q) {
            float m = 0.f;
            he = hipEventElapsedTime(&m, e[q], e[q + 1]);
            acc[q] += m;
        }
    }
    for (auto& x : e)
        if (x) (void)hipEventDestroy(x);
    if (st != NPDE_OK) return st;
    if (he != hipSuccess) return fail(s, NPDE_ERR_DEVICE, std::string("timing: ") + hipGetErrorString(he));
    HIPCHK(s, hipMemcpyAsync(s->hctl.data(), s->dctl, sizeof(Ctl) * s->n, hipMemcpyDeviceToHost, s->stream));
    HIPCHK(s, hipStreamSynchronize(s->stream));
    *ms_stage1 = acc[0] / nsteps;
    *ms_stage2 = acc[1] / nsteps;
    return fill_bad(s, nullptr);
}

npde_status npde_run_trajectory(npde_solver* s, double t_start, double t_end, int nout, const int* sizes, int n_sizes,
                                npde_bad_cell* bad) {
    if (!s || nout <= 0 || n_sizes < 0 || n_sizes > kMaxLevels || (n_sizes > 0 && !sizes))
        return fail(s, NPDE_ERR_INVALID, "bad run arguments");
    if (!(t_start >= 0.0 && t_start < t_end)) return fail(s, NPDE_ERR_INVALID, "need 0 <= t_start < t_end");
    // trajectory levels: m must divide nx, nx/m must divide ny (trajectory.hpp:22-29)
    size_t total = 0;
    std::vector<std::vector<size_t>> off(s->n, std::vector<size_t>(kMaxLevels, 0));
    for (int k = 0; k < s->n; ++k) {
        const npde_grid_desc& g = s->grids[k];
        for (int l = 0; l < n_sizes; ++l) {
            const int m = sizes[l];
            if (m <= 0 || g.nx % m != 0)
                return fail(s, NPDE_ERR_INVALID, "save size " + std::to_string(m) + " does not divide nx = " +
                                                     std::to_string(g.nx));
            const int f = g.nx / m;
            if (g.ny % f != 0)
                return fail(s, NPDE_ERR_INVALID, "save size " + std::to_string(m) + " does not divide ny evenly");
            off[k][l] = total;
            total += size_t(nout + 1) * 4 * m * (g.ny / f);
        }
    }
    // the same levels and frame count as the last run: keep its buffers, frame tasks and captured graph
    if (s->dftasks && s->dftimes && s->traj_nout == nout && int(s->traj_sizes.size()) == n_sizes &&
        std::equal(s->traj_sizes.begin(), s->traj_sizes.end(), sizes))
        return auto_run(s, t_start, t_end, nout, true, bad);
    if (s->dtraj) (void)hipFree(s->dtraj);
    s->dtraj = nullptr;
    s->traj_bytes = total * sizeof(float);
    if (total) {
        if (hipMalloc(&s->dtraj, s->traj_bytes) != hipSuccess) return fail(s, NPDE_ERR_NOMEM, "trajectory buffer");
        HIPCHK(s, hipMemset(s->dtraj, 0, s->traj_bytes));
    }
    s->traj_sizes.assign(sizes, sizes + n_sizes);
    s->traj_nout = nout;
    if (s->dftimes) (void)hipFree(s->dftimes);
    s->dftimes = nullptr;
    if (hipMalloc(&s->dftimes, sizeof(double) * size_t(nout + 1) * s->n) != hipSuccess)
        return fail(s, NPDE_ERR_NOMEM, "frame times");
    HIPCHK(s, hipMemset(s->dftimes, 0, sizeof(double) * size_t(nout + 1) * s->n));
    s->hftasks.clear();
    for (int k = 0; k < s->n; ++k) {
        InstDesc& D = s->hdesc[k];
        D.nlevels = n_sizes;
        D.ftimes = s->dftimes + size_t(k) * (nout + 1);
        D.ftask0 = int(s->hftasks.size());
        for (int l = 0; l < n_sizes; ++l) {
            const int m = sizes[l], f = D.nx / m, cells = m * (D.ny / f);
            D.lvl_m[l] = m;
            D.lvl_f[l] = f;
            D.traj[l] = s->dtraj + off[k][l];
            for (int c0 = 0; c0 < cells; c0 += 256) s->hftasks.push_back({k, l, c0, 0});
        }
        D.nftasks = int(s->hftasks.size()) - D.ftask0;
    }
    // an instance with no levels still needs one (empty) frame task so its frames complete
    for (int k = 0; k < s->n && n_sizes == 0; ++k) {
        InstDesc& D = s->hdesc[k];
        D.ftask0 = int(s->hftasks.size());
        s->hftasks.push_back({k, 0, 1 << 30, 0});
        D.nftasks = 1;
        D.lvl_m[0] = 1;
        D.lvl_f[0] = D.nx;
    }
    if (s->dftasks) (void)hipFree(s->dftasks);
    s->dftasks = nullptr;
    HIPCHK(s, hipMalloc(&s->dftasks, sizeof(FTask) * s->hftasks.size()));
    HIPCHK(s, hipMemcpy(s->dftasks, s->hftasks.data(), sizeof(FTask) * s->hftasks.size(), hipMemcpyHostToDevice));
    HIPCHK(s, hipMemcpy(s->ddesc, s->hdesc.data(), sizeof(InstDesc) * s->n, hipMemcpyHostToDevice));
    drop_graph(s);
    return auto_run(s, t_start, t_end, nout, true, bad);
}

npde_status npde_get_trajectory(npde_solver* s, int inst, int m, float* out) {
    if (!s || !out || inst < 0 || inst >= s->n) return fail(s, NPDE_ERR_INVALID, "bad arguments");
    if (!s->dtraj) return fail(s, NPDE_ERR_STATE, "no trajectory recorded");
    const InstDesc& D = s->hdesc[inst];
    for (int l = 0; l < D.nlevels; ++l)
        if (D.lvl_m[l] == m) {
            const size_t n = size_t(s->traj_nout + 1) * 4 * m * (D.ny / D.lvl_f[l]);
            HIPCHK(s, hipMemcpy(out, D.traj[l], n * sizeof(float), hipMemcpyDeviceToHost));
            return NPDE_OK;
        }
    return fail(s, NPDE_ERR_INVALID, "size not recorded");
}

npde_status npde_get_frame_times(npde_solver* s, int inst, double* t) {
    if (!s || !t || inst < 0 || inst >= s->n) return fail(s, NPDE_ERR_INVALID, "bad arguments");
    if (s->traj_nout < 0 || !s->dftimes) return fail(s, NPDE_ERR_STATE, "no trajectory recorded");
    HIPCHK(s, hipMemcpy(t, s->hdesc[inst].ftimes, sizeof(double) * (s->traj_nout + 1), hipMemcpyDeviceToHost));
    return NPDE_OK;
}

npde_status npde_block_average(npde_solver* s, int inst, int m, float* out) {
    if (!s || !out || inst < 0 || inst >= s->n) return fail(s, NPDE_ERR_INVALID, "bad arguments");
    const npde_grid_desc& g = s->grids[inst];
    if (m <= 0 || g.nx % m != 0 || g.ny % (g.nx / m) != 0) return fail(s, NPDE_ERR_INVALID, "size does not divide");
    const int nyc = g.ny / (g.nx / m);
    float* d = nullptr;
    const size_t n = size_t(4) * m * nyc;
    if (hipMalloc(&d, n * sizeof(float)) != hipSuccess) return fail(s, NPDE_ERR_NOMEM, "block average buffer");
    hipError_t e = launch_block_average(s->ddesc, inst, m, nyc, d, s->stream);
    if (e == hipSuccess) e = hipMemcpyAsync(out, d, n * sizeof(float), hipMemcpyDeviceToHost, s->stream);
    if (e == hipSuccess) e = hipStreamSynchronize(s->stream);
    (void)hipFree(d);
    if (e != hipSuccess) return fail(s, NPDE_ERR_DEVICE, hipGetErrorString(e));
    return NPDE_OK;
}

npde_status npde_block_average_noalloc(npde_solver* s, int inst, int m, float* out) {
    if (!s || !out || inst < 0 || inst >= s->n) return fail(s, NPDE_ERR_INVALID, "bad arguments");
    const npde_grid_desc& g = s->grids[inst];
    if (m <= 0 || g.nx % m != 0 || g.ny % (g.nx / m) != 0) return fail(s, NPDE_ERR_INVALID, "size does not divide");
    const int nyc = g.ny / (g.nx / m);
    const size_t n = size_t(4) * m * nyc;
    if (n > kNoAllocFloats)
        return fail(s, NPDE_ERR_INVALID, "block average of " + std::to_string(n) + " values exceeds the " +
                                             std::to_string(kNoAllocFloats) + " of the static staging array");
    if (!s->scratch) return fail(s, NPDE_ERR_STATE, "no staging array on this handle's device");
    // k_block_average writes every one of the n values before the copy reads them: nothing to clear
    hipError_t e = launch_block_average(s->ddesc, inst, m, nyc, s->scratch, s->stream);
    if (e == hipSuccess) e = hipMemcpyAsync(out, s->scratch, n * sizeof(float), hipMemcpyDeviceToHost, s->stream);
    if (e == hipSuccess) e = hipStreamSynchronize(s->stream);
    if (e != hipSuccess) return fail(s, NPDE_ERR_DEVICE, hipGetErrorString(e));
    return NPDE_OK;
}

npde_status npde_set_flux_correction(npde_solver* s, int inst, const double* corr_x, const double* corr_y) {
    if (!s || inst < 0 || inst >= s->n) return fail(s, NPDE_ERR_INVALID, "bad instance");
    if ((corr_x == nullptr) != (corr_y == nullptr)) return fail(s, NPDE_ERR_INVALID, "give both corrections or none");
    InstDesc& D = s->hdesc[inst];
    D.corr_x = corr_x;
    D.corr_y = corr_y;
    bool any = false;
    for (const auto& d : s->hdesc) any = any || d.corr_x != nullptr;
    if (corr_x && D.plane != (long long)(size_t(D.nx) * D.ny)) {
        D.corr_x = D.corr_y = nullptr;
        return fail(s, NPDE_ERR_INVALID, "flux correction needs nx*ny to be a multiple of 64");
    }
    const bool changed = s->has_corr != any;
    s->has_corr = any;
    drop_graph(s);
    if (changed) return upload_tasks(s);  // the automatic band height depends on it; uploads the descriptors
    HIPCHK(s, hipMemcpy(s->ddesc, s->hdesc.data(), sizeof(InstDesc) * s->n, hipMemcpyHostToDevice));
    return NPDE_OK;
}

npde_status npde_set_flux_network(npde_solver* s, const void* weights, const float* biases, double out_scale) {
    if (!s) return NPDE_ERR_INVALID;
    HIPCHK(s, hipStreamSynchronize(s->stream));
    // off: the instances' corrections that pointed at the network's planes go, and with them the planes
    if (!weights) {
        for (auto& D : s->hdesc)
            if (s->dnetcorr && D.corr_x >= s->dnetcorr && D.corr_x < s->dnetcorr + s->netcorr_doubles)
                D.corr_x = D.corr_y = nullptr;
        if (s->dnetcorr) (void)hipFree(s->dnetcorr);
        if (s->dnettasks) (void)hipFree(s->dnettasks);
        if (s->dnetpeak) (void)hipFree(s->dnetpeak);
        s->dnetpeak = nullptr;
        s->dnetcorr = nullptr;
        s->netcorr_doubles = 0;
        s->dnettasks = nullptr;
        s->nnettasks = 0;
        s->net_frags = nullptr;
        s->net_bias = nullptr;
        bool any = false;
        for (const auto& d : s->hdesc) any = any || d.corr_x != nullptr;
        const bool changed = s->has_corr != any;
        s->has_corr = any;
        drop_graph(s);
        if (changed) return upload_tasks(s);
        HIPCHK(s, hipMemcpy(s->ddesc, s->hdesc.data(), sizeof(InstDesc) * s->n, hipMemcpyHostToDevice));
        return NPDE_OK;
    }
    if (!biases || !(out_scale >= 0.0)) return fail(s, NPDE_ERR_INVALID, "need biases and a non-negative bound");
    for (const auto& g : s->grids)
        if (g.bc_x != NPDE_BC_PERIODIC || g.bc_y != NPDE_BC_PERIODIC || g.nx % 16 || g.ny % 16)
            return fail(s, NPDE_ERR_INVALID, "the flux network needs periodic instances with nx and ny multiples of 16");
    if (!s->dnetcorr) {
        size_t total = 0;
        for (const auto& D : s->hdesc) total += size_t(8) * D.plane;
        HIPCHK(s, hipMalloc(&s->dnetcorr, total * sizeof(double)));
        s->netcorr_doubles = total;
        HIPCHK(s, hipMemset(s->dnetcorr, 0, total * sizeof(double)));
        std::vector<NetTask> tasks;
        size_t off = 0;
        for (int k = 0; k < s->n; ++k) {
            InstDesc& D = s->hdesc[k];
            D.corr_x = s->dnetcorr + off;
            D.corr_y = s->dnetcorr + off + 4 * D.plane;
            off += size_t(8) * D.plane;
            for (int y0 = 0; y0 < D.ny; y0 += 16)
                for (int x0 = 0; x0 < D.nx; x0 += 16) tasks.push_back({k, x0, y0, 0});
        }
        HIPCHK(s, hipMalloc(&s->dnetpeak, sizeof(unsigned) * s->n));
        HIPCHK(s, hipMemset(s->dnetpeak, 0, sizeof(unsigned) * s->n));
        HIPCHK(s, hipMalloc(&s->dnettasks, sizeof(NetTask) * tasks.size()));
        HIPCHK(s, hipMemcpy(s->dnettasks, tasks.data(), sizeof(NetTask) * tasks.size(), hipMemcpyHostToDevice));
        s->nnettasks = int(tasks.size());
    }
    s->net_frags = weights;
    s->net_bias = biases;
    s->net_scale = float(out_scale);
    const bool changed = !s->has_corr;
    s->has_corr = true;
    drop_graph(s);
    if (changed) return upload_tasks(s);
    HIPCHK(s, hipMemcpy(s->ddesc, s->hdesc.data(), sizeof(InstDesc) * s->n, hipMemcpyHostToDevice));
    return NPDE_OK;
}

npde_status npde_set_flux_network_limit(npde_solver* s, double ratio) {
    if (!s) return NPDE_ERR_INVALID;
    if (!(ratio >= 0.0) || ratio == HUGE_VAL) return fail(s, NPDE_ERR_INVALID, "need a finite limit >= 0");
    HIPCHK(s, hipStreamSynchronize(s->stream));
    for (auto& D : s->hdesc) D.net_rate_limit = ratio;
    HIPCHK(s, hipMemcpy(s->ddesc, s->hdesc.data(), sizeof(InstDesc) * s->n, hipMemcpyHostToDevice));
    return NPDE_OK;
}

npde_status npde_flux_network_switched_off(npde_solver* s, int* off) {
    if (!s || !off) return NPDE_ERR_INVALID;
    npde_status st = pull_ctl(s);
    if (st != NPDE_OK) return st;
    for (int k = 0; k < s->n; ++k) off[k] = s->hctl[k].net_off;
    return NPDE_OK;
}

npde_status npde_flux_network_eval(npde_solver* s) {
    if (!s) return NPDE_ERR_INVALID;
    if (!s->net_frags) return fail(s, NPDE_ERR_STATE, "no flux network set");
    npde_status st = timed_launch(s, NPDE_PROF_NETWORK, [&] {
        return launch_flux_net(s->ddesc, s->dctl, s->dnettasks, s->nnettasks, s->n, nullptr, s->net_frags,
                               s->net_bias, s->net_scale, s->dnetpeak, s->stream);
    });
    if (st != NPDE_OK) return st;
    HIPCHK(s, hipStreamSynchronize(s->stream));
    return NPDE_OK;
}

npde_status npde_flux_network_peak(npde_solver* s, float* peak, int reset) {
    if (!s) return NPDE_ERR_INVALID;
    if (!s->dnetpeak) return fail(s, NPDE_ERR_STATE, "no flux network set");
    HIPCHK(s, hipStreamSynchronize(s->stream));
    if (peak) HIPCHK(s, hipMemcpy(peak, s->dnetpeak, sizeof(float) * s->n, hipMemcpyDeviceToHost));
    if (reset) HIPCHK(s, hipMemset(s->dnetpeak, 0, sizeof(unsigned) * s->n));
    return NPDE_OK;
}

npde_status npde_flux_network_planes(npde_solver* s, int inst, double** corr_x, double** corr_y) {
    if (!s || inst < 0 || inst >= s->n || !corr_x || !corr_y) return fail(s, NPDE_ERR_INVALID, "bad arguments");
    if (!s->dnetcorr) return fail(s, NPDE_ERR_STATE, "no flux network set");
    *corr_x = const_cast<double*>(s->hdesc[inst].corr_x);
    *corr_y = const_cast<double*>(s->hdesc[inst].corr_y);
    return NPDE_OK;
}

npde_status npde_get_tuning(const npde_solver* s, npde_tuning* t) {
    if (!s || !t) return NPDE_ERR_INVALID;
    *t = s->tune;
    return NPDE_OK;
}

npde_status npde_set_tuning(npde_solver* s, const npde_tuning* t) {
    if (!s || !t) return NPDE_ERR_INVALID;
    if (t->rows_per_task < 0 || !(t->waves_per_block == 1 || t->waves_per_block == 2 || t->waves_per_block == 4) ||
        t->steps_per_graph < 1 || t->waves_per_simd < 0 || t->waves_per_simd > 8 || t->net_period < 1 ||
        t->net_period > 64)
        return fail(s, NPDE_ERR_INVALID, "bad tuning");
    // an occupancy cap must be reachable with a launch's LDS (checked for the layout it gives; row groups take
    // their workgroup size from the width)
    if (t->waves_per_simd > 0 && lds_reserve_for(s, *t, row_group_waves_for(s, *t)) > kMaxLaunchLds)
        return fail(s, NPDE_ERR_INVALID, "occupancy cap below what a workgroup's LDS can enforce");
    const bool layout_changed = t->rows_per_task != s->tune.rows_per_task || t->row_groups != s->tune.row_groups ||
                                t->seams != s->tune.seams;
    // (the occupancy cap only changes the launches' LDS reservation: the captured graphs are dropped below)
    s->tune = *t;
    drop_graph(s);
    if (s->subset) return list_all(s);
    if (layout_changed) return upload_tasks(s);
    return NPDE_OK;
}

npde_status npde_set_arithmetic(npde_solver* s, int arith) {
    if (!s) return NPDE_ERR_INVALID;
    if (arith != NPDE_ARITH_EXACT && arith != NPDE_ARITH_FAST) return fail(s, NPDE_ERR_INVALID, "bad arithmetic mode");
    s->fast = arith == NPDE_ARITH_FAST;
    drop_graph(s);
    return NPDE_OK;
}

npde_status npde_get_arithmetic(const npde_solver* s, int* arith) {
    if (!s || !arith) return NPDE_ERR_INVALID;
    *arith = s->fast ? NPDE_ARITH_FAST : NPDE_ARITH_EXACT;
    return NPDE_OK;
}

npde_status npde_get_task_layout(const npde_solver* s, int* layout) {
    if (!s || !layout) return NPDE_ERR_INVALID;
    static_assert(int(NPDE_TASKS_STRIPS) == kStrips && int(NPDE_TASKS_ROW_GROUPS) == kRowGroups &&
                      int(NPDE_TASKS_SEAMS) == kSeams,
                  "npde_task_layout mirrors StageMode");
    *layout = s->layout;
    return NPDE_OK;
}

npde_status npde_set_profiling(npde_solver* s, int on) {
    if (!s) return NPDE_ERR_INVALID;
    npde_status st = prof_harvest(s);  // launches recorded before the switch
    if (st != NPDE_OK) return st;
    s->prof_on = on != 0;
    if (s->prof_on)
        for (int k = 0; k < NPDE_PROF_KINDS; ++k) {
            s->prof_ms[k] = 0.0;
            s->prof_n[k] = 0;
        }
    return NPDE_OK;
}

npde_status npde_get_profile(npde_solver* s, double* ms, long long* launches) {
    if (!s) return NPDE_ERR_INVALID;
    npde_status st = prof_harvest(s);
    if (st != NPDE_OK) return st;
    for (int k = 0; k < NPDE_PROF_KINDS; ++k) {
        if (ms) ms[k] = s->prof_ms[k];
        if (launches) launches[k] = s->prof_n[k];
    }
    return NPDE_OK;
}

npde_status npde_get_occupancy(const npde_solver* s, int* waves_per_simd) {
    if (!s || !waves_per_simd) return NPDE_ERR_INVALID;
    int wg = 0;
    const hipError_t e = stage_occupancy(s->flux, s->lim, s->has_corr, s->fast, s->tune.waves_per_block, s->layout,
                                         s->row_waves, s->band, lds_reserve(s), &wg);
    if (e != hipSuccess) return NPDE_ERR_DEVICE;
    const int waves = s->row_waves > 0 ? s->row_waves : s->tune.waves_per_block;
    *waves_per_simd = wg * waves / 4;
    return NPDE_OK;
}

npde_status npde_last_elapsed_ms(const npde_solver* s, double* ms) {
    if (!s || !ms) return NPDE_ERR_INVALID;
    *ms = s->last_ms;
    return NPDE_OK;
}

void* npde_stream(npde_solver* s) { return s ? (void*)s->stream : nullptr; }

npde_status npde_selftest_arith(const double* x, const double* y, long long n, double* q_fast, double* q_ieee,
                                double* root) {
    if (!x || !y || !q_fast || !q_ieee || !root || n <= 0) return NPDE_ERR_INVALID;
    double* d = nullptr;
    if (hipMalloc(&d, sizeof(double) * 5 * size_t(n)) != hipSuccess) return NPDE_ERR_NOMEM;
    const size_t b = sizeof(double) * size_t(n);
    hipError_t e = hipMemcpy(d, x, b, hipMemcpyHostToDevice);
    if (e == hipSuccess) e = hipMemcpy(d + n, y, b, hipMemcpyHostToDevice);
    if (e == hipSuccess) e = launch_selftest_arith(d, d + n, n, d + 2 * n, d + 3 * n, d + 4 * n, nullptr);
    if (e == hipSuccess) e = hipDeviceSynchronize();
    if (e == hipSuccess) e = hipMemcpy(q_fast, d + 2 * n, b, hipMemcpyDeviceToHost);
    if (e == hipSuccess) e = hipMemcpy(q_ieee, d + 3 * n, b, hipMemcpyDeviceToHost);
    if (e == hipSuccess) e = hipMemcpy(root, d + 4 * n, b, hipMemcpyDeviceToHost);
    (void)hipFree(d);
    return e == hipSuccess ? NPDE_OK : NPDE_ERR_DEVICE;
}

int npde_mfma32_probe(const float* a, const float* b, float* d, void* stream) {
    if (!a || !b || !d) return 1;
    return launch_mfma32_probe(a, b, d, (hipStream_t)stream) == hipSuccess ? 0 : 3;
}

// Built-in initial conditions of the driver (main2d.cpp:52-135), evaluated with the host libm.
npde_status npde_builtin_problem(const char* test, int n, npde_grid_desc* grid, double* t_end, double* U) {
    if (!test || !grid || !t_end || n <= 0) return NPDE_ERR_INVALID;
    const std::string name(test);
    const double PI = 3.14159265358979323846;
    npde_grid_desc g{};
    if (name == "sod_x" || name == "sod_y") {
        const int m = 4;
        g = name == "sod_x" ? npde_grid_desc{n, m, 0.0, 1.0, 0.0, double(m) / n, kTransmissive, kTransmissive}
                            : npde_grid_desc{m, n, 0.0, double(m) / n, 0.0, 1.0, kTransmissive, kTransmissive};
        *t_end = 0.2;
    } else if (name == "lax_liu3") {
        g = {n, n, 0.0, 1.0, 0.0, 1.0, kTransmissive, kTransmissive};
        *t_end = 0.3;
    } else if (name == "sine2d") {
        g = {n, n, 0.0, 1.0, 0.0, 1.0, kPeriodic, kPeriodic};
        *t_end = 1.0;
    } else if (name == "kh") {
        g = {n, n, 0.0, 1.0, 0.0, 1.0, kPeriodic, kPeriodic};
        *t_end = 2.0;
    } else {
        return NPDE_ERR_INVALID;
    }
    *grid = g;
    if (!U) return NPDE_OK;
    const double dx = (g.x_max - g.x_min) / g.nx, dy = (g.y_max - g.y_min) / g.ny;
    const size_t cells = size_t(g.nx) * g.ny;
    for (int j = 0; j < g.ny; ++j)
        for (int i = 0; i < g.nx; ++i) {
            const double x = g.x_min + (i + 0.5) * dx, y = g.y_min + (j + 0.5) * dy;  // Grid::xc / yc
            Prim w{};
            if (name == "sod_x" || name == "sod_y") {
                const double pos = name == "sod_x" ? x : y;
                w = pos < 0.5 ? Prim{1.0, 0.0, 0.0, 1.0} : Prim{0.125, 0.0, 0.0, 0.1};
            } else if (name == "lax_liu3") {  // Lax & Liu configuration 3
                const bool right = x > 0.5, top = y > 0.5;
                w = top ? (right ? Prim{1.5, 0.0, 0.0, 1.5} : Prim{0.5323, 1.206, 0.0, 0.3})
                        : (right ? Prim{0.5323, 0.0, 1.206, 0.3} : Prim{0.138, 1.206, 1.206, 0.029});
            } else if (name == "sine2d") {
                w = {1.0 + 0.2 * std::sin(2.0 * PI * (x + y)), 1.0, 1.0, 1.0};
            } else {  // kh: two exponential shear layers (McNally, Lyra & Passy 2012)
                const double r1 = 1.0, r2 = 2.0, rm = 0.5 * (r1 - r2);
                const double u1 = 0.5, u2 = -0.5, um = 0.5 * (u1 - u2);
                const double L = 0.025;
                double rho, u;
                if (y < 0.25) {
                    rho = r1 - rm * std::exp((y - 0.25) / L);
                    u = u1 - um * std::exp((y - 0.25) / L);
                } else if (y < 0.5) {
                    rho = r2 + rm * std::exp((0.25 - y) / L);
                    u = u2 + um * std::exp((0.25 - y) / L);
                } else if (y < 0.75) {
                    rho = r2 + rm * std::exp((y - 0.75) / L);
                    u = u2 + um * std::exp((y - 0.75) / L);
                } else {
                    rho = r1 - rm * std::exp((0.75 - y) / L);
                    u = u1 - um * std::exp((0.75 - y) / L);
                }
                w = {rho, u, 0.01 * std::sin(4.0 * PI * x), 2.5};
            }
            const Cons c = cons_of(w);
            const size_t k = size_t(j) * g.nx + i;
            U[k] = c.d;
            U[cells + k] = c.m;
            U[2 * cells + k] = c.n;
            U[3 * cells + k] = c.e;
        }
    return NPDE_OK;
}

// Autotuner: times candidate (task layout, rows_per_task, waves_per_block) on the handle's own instances
// with host-driven steps of the current CFL dt, restoring the state after every trial, then occupancy
// caps below the best candidate's natural occupancy (fewer resident wavefronts per SIMD: less cache
// contention between them, less latency hiding).  Every candidate computes the same bits (the per-cell
// arithmetic does not depend on the launch shape); only the time differs.
npde_status npde_autotune(npde_solver* s, const char* cache_path, npde_tuning* chosen) {
    if (!s) return NPDE_ERR_INVALID;
    std::ostringstream key;
    key << "flux" << s->flux << "_lim" << s->lim << (s->fast ? "_fast" : "") << "_n" << s->n;
    {
        std::map<std::string, int> sizes;
        for (const auto& g : s->grids) sizes[std::to_string(g.nx) + "x" + std::to_string(g.ny) + "b" +
                                             std::to_string(g.bc_x) + std::to_string(g.bc_y)] += 1;
        for (const auto& kv : sizes) key << "_" << kv.first << "*" << kv.second;
    }
    if (cache_path) {
        std::ifstream in(cache_path);
        std::string line;
        while (std::getline(in, line)) {  // "key rows waves_per_block [row_groups [seams [waves_per_simd]]]"
            std::istringstream ls(line);
            std::string k;
            npde_tuning t = s->tune;
            if (!(ls >> k >> t.rows_per_task >> t.waves_per_block) || k != key.str()) continue;
            // a line without the layout fields keeps the handle's current layout choices
            int v = 0;
            if (ls >> v) {
                t.row_groups = v;
                if (ls >> v) {
                    t.seams = v;
                    t.waves_per_simd = ls >> v ? v : 0;
                }
            }
            npde_status st = npde_set_tuning(s, &t);
            if (st == NPDE_OK && chosen) *chosen = s->tune;
            return st;
        }
    }
    // save state and run state
    size_t total = 0;
    for (int k = 0; k < s->n; ++k) total += 4 * size_t(s->hdesc[k].plane);
    double* save = nullptr;
    if (hipMalloc(&save, total * sizeof(double)) != hipSuccess) return fail(s, NPDE_ERR_NOMEM, "autotune buffer");
    npde_status st = pull_ctl(s);
    std::vector<Ctl> ctl_saved = s->hctl;
    auto copy_state = [&](bool to_save) {
        size_t off = 0;
        for (int k = 0; k < s->n; ++k) {
            const size_t n4 = 4 * size_t(s->hdesc[k].plane);
            if (to_save) (void)hipMemcpy(save + off, s->hdesc[k].U, n4 * sizeof(double), hipMemcpyDeviceToDevice);
            else (void)hipMemcpy(s->hdesc[k].U, save + off, n4 * sizeof(double), hipMemcpyDeviceToDevice);
            off += n4;
        }
    };
    copy_state(true);
    std::vector<double> dts(s->n);
    if (st == NPDE_OK) st = npde_compute_dt(s, dts.data(), nullptr);
    const npde_tuning orig = s->tune;
    npde_tuning best = orig;
    double best_ms = 1e30;
    int max_ny = 1;
    for (const auto& g : s->grids) max_ny = std::max(max_ny, g.ny);
    const int rows_c[] = {2, 4, 8, 16, 32, 64, 128};
    const int wpb_c[] = {1, 2, 4};
    // task layouts that apply to this batch, as (row_groups, seams): row groups up to W = 8, seam
    // strips, plain strips
    npde_tuning probe = orig;
    probe.row_groups = 2;
    std::vector<std::pair<int, int>> layouts;
    if (row_group_waves_for(s, probe) > 0) layouts.push_back({2, 0});
    probe.row_groups = 0;
    probe.seams = 2;
    if (seams_apply_for(s, probe)) layouts.push_back({0, 2});
    layouts.push_back({0, 0});
    // the fastest of 5 timed steps (after one warm-up), each the second step of a sequence as in the
    // device loop: robust to a stray slow step
    auto trial = [&](const npde_tuning& t) {
        if ((st = npde_set_tuning(s, &t)) != NPDE_OK) return;
        double ms_min = 1e30;
        for (int rep = 0; rep < 6 && st == NPDE_OK; ++rep) {
            copy_state(false);
            st = host_step(s, dts.data(), true);
            if (st == NPDE_OK) st = host_step(s, dts.data(), false);
            if (rep > 0) ms_min = std::min(ms_min, s->last_ms);
        }
        if (st == NPDE_OK && ms_min < best_ms) {
            best_ms = ms_min;
            best = t;
        }
    };
    for (const auto& lay : layouts)
    for (int r : rows_c) {
        if (st != NPDE_OK) break;
        if (r > max_ny && r != rows_c[0]) continue;
        if (lay.second && r > 64) continue;  // seam bands keep their seam solves in LDS: at most 64 rows
        for (int w : wpb_c) {
            if (st != NPDE_OK) break;
            if (lay.first && w != wpb_c[0]) continue;  // a row group's workgroup size is fixed by the width
            npde_tuning t = orig;
            t.rows_per_task = r;
            t.waves_per_block = w;
            t.row_groups = lay.first;
            t.seams = lay.second;
            t.waves_per_simd = 0;
            trial(t);
        }
    }
    // occupancy caps below the best candidate's own occupancy (caps an LDS reservation cannot enforce at
    // this workgroup size are refused by npde_set_tuning and skipped)
    int natural = 0;
    if (st == NPDE_OK && (st = npde_set_tuning(s, &best)) == NPDE_OK) st = npde_get_occupancy(s, &natural);
    const npde_tuning uncapped = best;
    for (int cap = natural - 1; cap >= 1 && st == NPDE_OK; --cap) {
        npde_tuning t = uncapped;
        t.waves_per_simd = cap;
        if (npde_set_tuning(s, &t) != NPDE_OK) continue;
        trial(t);
    }
    copy_state(false);
    (void)hipFree(save);
    if (st != NPDE_OK) {
        (void)npde_set_tuning(s, &orig);
        s->hctl = ctl_saved;
        (void)push_ctl(s);
        return st;
    }
    st = npde_set_tuning(s, &best);
    s->hctl = ctl_saved;
    if (st == NPDE_OK) st = push_ctl(s);
    if (st == NPDE_OK && cache_path) {
        std::ofstream out(cache_path, std::ios::app);
        out << key.str() << " " << best.rows_per_task << " " << best.waves_per_block << " " << best.row_groups
            << " " << best.seams << " " << best.waves_per_simd << "\n";
    }
    if (chosen) *chosen = s->tune;
    return st;
}

}  // extern "C"
