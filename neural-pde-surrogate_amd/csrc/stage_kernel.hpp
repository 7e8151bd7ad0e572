// The SSP-RK2 stage kernel of the 2D Euler finite-volume step, shared by the exact build
// (euler_kernels.hip) and the fast-arithmetic build (euler_kernels_fast.hip); the arithmetic comes
// from euler_numerics.hpp, whose namespace (npde::exact or npde::fast) this file's kernels join.
//
// Stage kernel design (one wavefront = a strip of columns x a band of rows; the strip layout is
// described first, the row-group and seam layouts below):
//   * each lane owns one column; in the strip layout lanes 0,1 and 62,63 are the 2-column halo of a
//     60-column strip;
//   * the wavefront marches up its band one row per iteration, keeping a 3-row window of primitives
//     in registers, so every cell's primitives, limited slopes and four face-side states
//     (W +- dW/2 -> conserved -> primitive + sound speed) are computed exactly once per stage;
//   * x neighbours come from the adjacent lanes by DPP wavefront shifts; each lane solves the
//     Riemann problem at its left x face, and receives its right-face flux from the next lane;
//   * each y face is solved once, by the lane below/above it, carried to the next row in registers;
//   * loads of one row are four fully coalesced 512-byte reads (one per conserved plane);
//   * ghost cells are virtual: stencil indices outside the grid are remapped (periodic wrap or
//     transmissive clamp) before the load;
//   * row-group mode (ROWS): for periodic-x instances whose width is 64*W columns (W <= 8), one
//     workgroup of W wavefronts covers a whole band of rows and every lane owns a column; the edge
//     lanes' x neighbours come from the adjacent wavefronts of the workgroup (periodically) through
//     a few LDS slots, read as the `old` operand of the DPP shift, so no halo column is recomputed;
//     two workgroup barriers per row order the exchanges (by default the host uses it for W = 1 only,
//     where it beats the seam mode below; an 8-wavefront workgroup at these register counts is alone
//     on its CU and waits at each barrier);
//   * seam mode (SEAMS): for widths that are a multiple of 64, every lane owns a column of a 64-column
//     strip.  Before its march, the wavefront solves the two strip boundaries ("seams") of its band
//     in parallel (one lane per seam and row) into LDS: the primitives of the outside neighbour
//     column and the x flux through the seam.  During the march lanes 0 and 63 take the neighbour's
//     primitives as the `old` operand of the DPP shift, lane 0 its left-face flux and lane 63 its
//     right-face flux from there.  The seam solves read the four columns around each seam from
//     small edge arrays the previous stage kernel wrote (its lanes 0, 1, 62, 63), not from the full
//     state.  The same functions on the same operands produce every value, so the bits are those of
//     the other modes, and no wavefront waits for another;
//   * workgroups are renumbered so that each XCD runs a contiguous range of the task list, which
//     keeps neighbouring tasks (overlapping halos) in one L2;
//   * the row loop is unrolled so that the carried rows are renamed instead of copied;
//   * stage 2 fuses the next step's dt reduction into its epilogue (to_primitive, positivity check,
//     max(|u|+a), max(|v|+a) of U^{n+1}) and writes one partial per task; k_finish (one wavefront per
//     instance) reduces them and advances that instance's run state.
#pragma once
#include <hip/hip_runtime.h>

#include "euler_numerics.hpp"
#include "npde_device.hpp"

namespace npde {
inline namespace NPDE_ARITH_NS {

// ---------------------------------------------------------------------------------------------
// cross-lane moves: lane l receives lane l-1 (from_prev) or lane l+1 (from_next).  bound_ctrl makes
// the lane without a source (0 or 63, a halo lane whose value is never used) read 0, so the move
// needs no initialised destination (update_dpp with an `old` operand costs one extra v_mov per dword).
__device__ __forceinline__ int dpp_prev(int x) {
    return __builtin_amdgcn_mov_dpp(x, 0x138 /* wave_shr:1 */, 0xF, 0xF, true);
}
__device__ __forceinline__ int dpp_next(int x) {
    return __builtin_amdgcn_mov_dpp(x, 0x130 /* wave_shl:1 */, 0xF, 0xF, true);
}
// the same shifts with the lane that has no source (lane 0 for from_prev, 63 for from_next) taking
// `old` instead: the neighbour value that lives in another wavefront
__device__ __forceinline__ int dpp_prev_or(int x, int old) {
    return __builtin_amdgcn_update_dpp(old, x, 0x138 /* wave_shr:1 */, 0xF, 0xF, false);
}
__device__ __forceinline__ int dpp_next_or(int x, int old) {
    return __builtin_amdgcn_update_dpp(old, x, 0x130 /* wave_shl:1 */, 0xF, 0xF, false);
}
__device__ __forceinline__ double from_prev_or(double x, double old) {
    return __hiloint2double(dpp_prev_or(__double2hiint(x), __double2hiint(old)),
                            dpp_prev_or(__double2loint(x), __double2loint(old)));
}
__device__ __forceinline__ double from_next_or(double x, double old) {
    return __hiloint2double(dpp_next_or(__double2hiint(x), __double2hiint(old)),
                            dpp_next_or(__double2loint(x), __double2loint(old)));
}
__device__ __forceinline__ Prim from_prev_or(const Prim& w, const double* o) {
    return {from_prev_or(w.d, o[0]), from_prev_or(w.u, o[1]), from_prev_or(w.v, o[2]), from_prev_or(w.p, o[3])};
}
__device__ __forceinline__ Prim from_next_or(const Prim& w, const double* o) {
    return {from_next_or(w.d, o[0]), from_next_or(w.u, o[1]), from_next_or(w.v, o[2]), from_next_or(w.p, o[3])};
}
__device__ __forceinline__ Side from_prev_or(const Side& s, const double* o) {
    return {from_prev_or(s.d, o[0]),  from_prev_or(s.mn, o[1]), from_prev_or(s.mt, o[2]),
            from_prev_or(s.e, o[3]),  from_prev_or(s.un, o[4]), from_prev_or(s.ut, o[5]),
            from_prev_or(s.p, o[6]),  from_prev_or(s.a, o[7]),  from_prev_or(s.rd, o[8])};
}
__device__ __forceinline__ Cons from_next_or(const Cons& c, const double* o) {
    return {from_next_or(c.d, o[0]), from_next_or(c.m, o[1]), from_next_or(c.n, o[2]), from_next_or(c.e, o[3])};
}
__device__ __forceinline__ void put(double* o, const Prim& w) { o[0] = w.d; o[1] = w.u; o[2] = w.v; o[3] = w.p; }
__device__ __forceinline__ void put(double* o, const Cons& c) { o[0] = c.d; o[1] = c.m; o[2] = c.n; o[3] = c.e; }
__device__ __forceinline__ void put(double* o, const Side& s) {
    o[0] = s.d; o[1] = s.mn; o[2] = s.mt; o[3] = s.e; o[4] = s.un; o[5] = s.ut; o[6] = s.p; o[7] = s.a; o[8] = s.rd;
}

__device__ __forceinline__ double from_prev(double x) {
    return __hiloint2double(dpp_prev(__double2hiint(x)), dpp_prev(__double2loint(x)));
}
__device__ __forceinline__ double from_next(double x) {
    return __hiloint2double(dpp_next(__double2hiint(x)), dpp_next(__double2loint(x)));
}
__device__ __forceinline__ Prim from_prev(const Prim& w) {
    return {from_prev(w.d), from_prev(w.u), from_prev(w.v), from_prev(w.p)};
}
__device__ __forceinline__ Prim from_next(const Prim& w) {
    return {from_next(w.d), from_next(w.u), from_next(w.v), from_next(w.p)};
}
__device__ __forceinline__ Side from_prev(const Side& s) {
    return {from_prev(s.d),  from_prev(s.mn), from_prev(s.mt), from_prev(s.e),
            from_prev(s.un), from_prev(s.ut), from_prev(s.p),  from_prev(s.a),
            from_prev(s.rd)};
}
__device__ __forceinline__ Cons from_next(const Cons& c) {
    return {from_next(c.d), from_next(c.m), from_next(c.n), from_next(c.e)};
}

__device__ __forceinline__ double wave_max(double x) {
    for (int o = 32; o > 0; o >>= 1) x = tmax(x, __shfl_xor(x, o, 64));
    return x;
}
__device__ __forceinline__ int wave_min(int x) {
    for (int o = 32; o > 0; o >>= 1) {
        const int y = __shfl_xor(x, o, 64);
        x = y < x ? y : x;
    }
    return x;
}

// State pointers live in the instance descriptor, so the compiler cannot see that they address
// global memory; casting to the global address space gives global_load/global_store (scalar base +
// lane offset) instead of flat accesses.
typedef __attribute__((address_space(1))) const double gcdouble;
typedef __attribute__((address_space(1))) double gdouble;
struct alignas(16) DPair { double a, b; };  // 16-byte loads and stores of edge-array entries
typedef __attribute__((address_space(1))) const DPair gcpair;
typedef __attribute__((address_space(1))) DPair gpair;

__device__ __forceinline__ Cons load4(const double* B, long long off, long long P) {
    const gcdouble* g = (const gcdouble*)B;
    return {g[off], g[off + P], g[off + 2 * P], g[off + 3 * P]};
}

__device__ __forceinline__ void store4(double* B, long long off, long long P, const Cons& c) {
    gdouble* g = (gdouble*)B;
    g[off] = c.d;
    g[off + P] = c.m;
    g[off + 2 * P] = c.n;
    g[off + 3 * P] = c.e;
}

__device__ __forceinline__ Prim prim_of(const Cons& c) { return prim_of(c.d, c.m, c.n, c.e); }

// Workgroups are dispatched to the 8 XCDs round-robin (workgroup b runs on XCD b % 8), and each XCD has
// its own L2.  Renumbering them so that every XCD receives one contiguous range of the task list keeps
// neighbouring tasks, whose halo rows and columns overlap, in the same L2.  A bijection of [0, nb).
__device__ __forceinline__ int xcd_block(int b, int nb) {
    constexpr int kXcds = 8;
    const int per = nb / kXcds, rem = nb % kXcds;
    const int x = b % kXcds, k = b / kXcds;
    return x < rem ? x * (per + 1) + k : rem * (per + 1) + (x - rem) * per + k;
}

// This task's maxima (all lanes) -> its partial slot.
__device__ __forceinline__ void publish_partial(Partial* __restrict__ part, int tid, int lane, double sx,
                                                double sy, int bad) {
    sx = wave_max(sx);
    sy = wave_max(sy);
    bad = wave_min(bad);
    if (lane == 0) {
        part[tid].sx = sx;
        part[tid].sy = sy;
        part[tid].bad = bad;
    }
}

// ---------------------------------------------------------------------------------------------
// Seam mode: the face left of column i0 (a multiple of 64, 0 <= i0 <= nx) in row j, solved as the strip
// lanes beside it would solve it: the primitives of columns i0-2 .. i0+1 (ghosts remapped), the limited
// slopes of columns i0-1 and i0, their facing edge sides, the Riemann solve and the face's learned
// increment.  The four cells come from the edge arrays E written by the previous stage kernel
// (compact), or from the state V itself.  Gives the primitives of the cells left (wl) and right (wr) of
// the face and its flux F.
template <int SCHEME, int LIM>
__device__ __forceinline__ void seam_face(const InstDesc& D, const double* V, const double* E, bool compact,
                                          bool corr, int i0, int j, Prim& wl, Prim& wr, Cons& F) {
    const int nx = D.nx, ny = D.ny;
    const long long P = D.plane, r = (long long)j * nx;
    const long long R = (long long)(nx / kWave) * ny * kEdgeRecord;  // right arrays follow the left ones
    auto cell = [&](int i) -> Cons {
        const int c = remap(i, nx, D.bc_x);
        if (!compact) return load4(V, r + c, P);
        const int st = c / kWave, e = c % kWave;  // e is 0, 1 (left edge) or 62, 63 (right edge)
        const gcpair* p = (const gcpair*)(E + (e < 2 ? 0 : R) + ((long long)st * ny + j) * kEdgeRecord + (e & 1) * 4);
        return {p[0].a, p[0].b, p[1].a, p[1].b};
    };
    const Prim w0 = prim_of(cell(i0 - 2));
    wl = prim_of(cell(i0 - 1));
    wr = prim_of(cell(i0));
    const Prim w3 = prim_of(cell(i0 + 1));
    const Side rt = edge_side(wl, slope<LIM>(w0, wl, wr), +0.5);
    const Side lf = edge_side(wr, slope<LIM>(wl, wr, w3), -0.5);
    F = flux_x<SCHEME>(rt, lf);
    if (corr && D.corr_x != nullptr && D.corr_y != nullptr) {
        // the increment of the face left of column i0 (the right boundary face wraps when periodic), scaled
        // by the face's flux scale from the two cells beside it
        const int cxi = i0 < nx ? i0 : (D.bc_x == kPeriodic ? remap(i0, nx, kPeriodic) : -1);
        if (cxi >= 0)
            F = add_scaled(F, load4(D.corr_x, r + cxi, P),
                           corr_scale(wl.d, sound(wl.d, wl.p), wr.d, sound(wr.d, wr.p)));
    }
}

// seam mode: LDS doubles per wavefront and band row (for each of the strip's two seams, the outside
// neighbour's primitives and the seam's flux); bands are at most kMaxSeamRows rows
constexpr int kSeamLds = 16;
constexpr int kMaxSeamRows = 64;
inline unsigned seam_lds_bytes(int mode, int waves_per_block, int seam_rows) {
    return mode == kSeams ? unsigned(waves_per_block * seam_rows * kSeamLds * sizeof(double)) : 0u;
}

template <int SCHEME, int LIM, int STAGE, bool CORR, int MODE>
__global__ __launch_bounds__(kMaxRowWaves * kWave) void k_stage(const InstDesc* __restrict__ desc,
                                                                Ctl* __restrict__ ctl,
                                                                const Task* __restrict__ tasks, int ntasks,
                                                                Partial* __restrict__ part,
                                                                int* __restrict__ ndone, int compact,
                                                                int seam_rows) {
    constexpr bool ROWS = MODE == kRowGroups;
    constexpr bool SEAMS = MODE == kSeams;
    const int lane = threadIdx.x & (kWave - 1);
    const int wv = __builtin_amdgcn_readfirstlane(threadIdx.x / kWave);  // wavefront in the workgroup
    const int nw = blockDim.x / kWave;
    // strips: one task per wavefront; row groups: one task (band) per workgroup, one partial per wavefront
    const int blk = xcd_block(int(blockIdx.x), int(gridDim.x));
    const int task = ROWS ? blk : __builtin_amdgcn_readfirstlane(blk * nw + wv);
    const int tid = ROWS ? __builtin_amdgcn_readfirstlane(blk * nw + wv) : task;
    if (task >= ntasks) return;
    const Task tk = tasks[task];
    Ctl* c = ctl + tk.inst;
    if (!c->active) return;
    const double dt = c->dt;
    const InstDesc& D = desc[tk.inst];
    const int nx = D.nx, ny = D.ny, bcy = D.bc_y;
    const long long P = D.plane;
    const double cx = D.cx, cy = D.cy;

    // this lane's column, and where it is loaded from (in range for every lane)
    const int i = ROWS ? wv * kWave + lane : (SEAMS ? tk.i0 + lane : tk.i0 - kHalo + lane);
    const int ci = (ROWS || SEAMS) ? i : remap(i, nx, D.bc_x);
    const bool owner = ROWS || SEAMS || (lane >= kHalo && lane < kWave - kHalo && i < nx);

    gpair* eout = nullptr;  // seams, lanes 0, 1, 62, 63: where this lane's output column goes in the edge arrays
    if constexpr (SEAMS) {
        const long long strips = nx / kWave;
        eout = (gpair*)(D.edge[STAGE == 1 ? 1 : 0] + (lane >= kWave / 2 ? strips * ny * kEdgeRecord : 0) +
                        (long long)(tk.i0 / kWave) * ny * kEdgeRecord + (lane & 1) * 4);
    }

    // row groups: edge values exchanged between the wavefronts of the workgroup.  s_w holds the edge
    // lanes' primitives of a row (double-buffered by row parity), s_rt lane 63's right-face side,
    // s_fl lane 0's left-face flux.
    __shared__ double s_w[2][2][kMaxRowWaves][5];  // primitives, then (flux corrections) the sound speed
    __shared__ double s_rt[kMaxRowWaves][9];
    __shared__ double s_fl[kMaxRowWaves][4];
    const int wl = ROWS ? (wv + nw - 1) % nw : 0;  // wavefront holding the columns to the left
    const int wr = ROWS ? (wv + 1) % nw : 0;       // ... and to the right (periodic)

    const double* __restrict__ V = (STAGE == 1) ? D.U : D.U1;  // stencil input
    double* O = (STAGE == 1) ? D.U1 : D.U;                     // output (stage 2: in place on U)

    // seams: before the march, lanes 0-31 solve the seam left of the strip and lanes 32-63 the seam right
    // of it, 32 band rows at a time, into this wavefront's LDS rows: the outside neighbour's primitives
    // (column i0-1 for lane 0, i0+64 for lane 63) and the seam's flux.  Only this wavefront reads them,
    // and a wavefront's LDS accesses complete in order, so no barrier is needed.
    double* sseam = nullptr;
    if constexpr (SEAMS) {
        extern __shared__ double s_dyn[];
        sseam = s_dyn + wv * seam_rows * kSeamLds;
        const int side = lane >> 5;
        for (int r0 = tk.j0; r0 < tk.j1; r0 += kWave / 2) {
            const int j = r0 + (lane & (kWave / 2 - 1));
            if (j < tk.j1) {
                Prim wl, wr;
                Cons F;
                seam_face<SCHEME, LIM>(D, V, D.edge[STAGE == 1 ? 0 : 1], compact != 0, CORR, tk.i0 + side * kWave, j,
                                       wl, wr, F);
                const Prim& w = side ? wr : wl;
                double* o = sseam + (j - tk.j0) * kSeamLds + side * (kSeamLds / 2);
                o[0] = w.d; o[1] = w.u; o[2] = w.v; o[3] = w.p;
                o[4] = F.d; o[5] = F.m; o[6] = F.n; o[7] = F.e;
            }
        }
        __builtin_amdgcn_wave_barrier();
    }

    auto row = [&](int j) -> long long { return (long long)remap(j, ny, bcy) * nx + ci; };

    // learned flux increments (coarse solver only): face left of column i, face below row r
    int cxi = -1;
    // an instance of a corrected batch without increments of its own reads none (null pointers)
    const bool has_corr = CORR && D.corr_x != nullptr && D.corr_y != nullptr;
    if constexpr (CORR) cxi = (i >= 0 && i < nx) ? i : (D.bc_x == kPeriodic ? remap(i, nx, kPeriodic) : -1);
    if (!has_corr) cxi = -1;
    auto corr_x = [&](int r) -> Cons {
        if (cxi < 0) return {0.0, 0.0, 0.0, 0.0};
        return load4(D.corr_x, (long long)r * nx + cxi, P);
    };
    auto corr_y = [&](int r) -> Cons {  // r in [0, ny]
        int rr = r < ny ? r : (D.bc_y == kPeriodic ? 0 : -1);
        if (rr < 0 || !has_corr) return {0.0, 0.0, 0.0, 0.0};
        return load4(D.corr_y, (long long)rr * nx + ci, P);
    };

    const int j0 = tk.j0, j1 = tk.j1;

    // prologue: rows j0-2 .. j0+1, the y face below row j0, and the top side of row j0
    Cons va = load4(V, row(j0), P);
    Cons vb = load4(V, row(j0 + 1), P);
    const Prim w0 = prim_of(load4(V, row(j0 - 2), P));
    const Prim w1 = prim_of(load4(V, row(j0 - 1), P));
    Prim wa = prim_of(va);
    Prim wb = prim_of(vb);
    // flux corrections: the sound speeds of the carried rows, for the faces' flux scales
    double aa = 0.0, ab = 0.0;
    if constexpr (CORR) {
        aa = sound(wa.d, wa.p);
        ab = sound(wb.d, wb.p);
    }
    Side top = edge_side(w1, slope<LIM>(w0, w1, wa), +0.5);
    Cons glo;
    {
        const Prim dy = slope<LIM>(w1, wa, wb);
        glo = flux_y<SCHEME>(top, edge_side(wa, dy, -0.5));
        if constexpr (CORR) glo = add_scaled(glo, corr_y(j0), corr_scale(w1.d, sound(w1.d, w1.p), wa.d, aa));
        top = edge_side(wa, dy, +0.5);
    }
    if constexpr (ROWS) {  // edge primitives (and sound speeds) of row j0 for the first x slopes
        if (lane == 0) {
            put(s_w[j0 & 1][0][wv], wa);
            if constexpr (CORR) s_w[j0 & 1][0][wv][4] = aa;
        }
        if (lane == kWave - 1) {
            put(s_w[j0 & 1][1][wv], wa);
            if constexpr (CORR) s_w[j0 & 1][1][wv][4] = aa;
        }
        __syncthreads();
    }

    double sx = 0.0, sy = 0.0;
    int bad = kNoBad;

    // software pipeline: the rows an iteration consumes were loaded one iteration earlier, so HBM
    // latency overlaps a whole row of arithmetic instead of stalling the wavefront
    Cons vn = load4(V, row(j0 + 2), P);
    Cons ub{};
    if constexpr (STAGE == 2) ub = load4(O, (long long)j0 * nx + ci, P);  // U, point-wise

    auto row_step = [&](const int f) __attribute__((always_inline)) {
        const Cons vq = load4(V, row(f + 3), P);
        Cons uq{};
        if constexpr (STAGE == 2) uq = load4(O, (long long)(f + 1 < ny ? f + 1 : f) * nx + ci, P);
        const Prim wn = prim_of(vn);
        double an = 0.0;
        if constexpr (CORR) an = sound(wn.d, wn.p);

        // y: slope of row f+1, its two y-face sides, and the face between rows f and f+1
        const Prim dyb = slope<LIM>(wa, wb, wn);
        Cons ghi = flux_y<SCHEME>(top, edge_side(wb, dyb, -0.5));
        if constexpr (CORR) ghi = add_scaled(ghi, corr_y(f + 1), corr_scale(wa.d, aa, wb.d, ab));
        const Side topn = edge_side(wb, dyb, +0.5);

        // x: slope of this cell from the neighbouring lanes, left-face flux here, right-face flux from lane+1
        Prim dxs, wp;  // wp: the cell to the left
        double ap = 0.0;
        if constexpr (ROWS) {
            wp = from_prev_or(wa, s_w[f & 1][1][wl]);
            if constexpr (CORR) ap = from_prev_or(aa, s_w[f & 1][1][wl][4]);
            dxs = slope<LIM>(wp, wa, from_next_or(wa, s_w[f & 1][0][wr]));
        } else if constexpr (SEAMS) {
            const double* q = sseam + (f - j0) * kSeamLds + (lane >> 5) * (kSeamLds / 2);
            const double o[4] = {q[0], q[1], q[2], q[3]};
            wp = from_prev_or(wa, o);
            // lane 0's left face is the seam, whose flux (increment included) comes from the seam solve
            if constexpr (CORR) ap = from_prev(aa);
            dxs = slope<LIM>(wp, wa, from_next_or(wa, o));
        } else {
            wp = from_prev(wa);
            if constexpr (CORR) ap = from_prev(aa);
            dxs = slope<LIM>(wp, wa, from_next(wa));
        }
        const Side lf = edge_side(wa, dxs, -0.5);
        const Side rt = edge_side(wa, dxs, +0.5);
        Side rl;  // right-face side of the column to the left
        if constexpr (ROWS) {
            if (lane == kWave - 1) put(s_rt[wv], rt);
            __syncthreads();
            rl = from_prev_or(rt, s_rt[wl]);
        } else {
            rl = from_prev(rt);
        }
        Cons fl = flux_x<SCHEME>(rl, lf);
        if constexpr (CORR) fl = add_scaled(fl, corr_x(f), corr_scale(wp.d, ap, wa.d, aa));
        Cons fr;  // left-face flux of the column to the right
        if constexpr (ROWS) {
            // with it, the edge primitives of the next row (the slots of this row's parity were
            // read before the barrier above, so the other parity is free)
            if (lane == 0) {
                put(s_fl[wv], fl);
                put(s_w[(f + 1) & 1][0][wv], wb);
                if constexpr (CORR) s_w[(f + 1) & 1][0][wv][4] = ab;
            }
            if (lane == kWave - 1) {
                put(s_w[(f + 1) & 1][1][wv], wb);
                if constexpr (CORR) s_w[(f + 1) & 1][1][wv][4] = ab;
            }
            __syncthreads();
            fr = from_next_or(fl, s_fl[wr]);
        } else if constexpr (SEAMS) {
            // lane 0's left face is the left seam (its own Riemann solve had no left neighbour)
            const double* q = sseam + (f - j0) * kSeamLds + (lane >> 5) * (kSeamLds / 2);
            const double F[4] = {q[4], q[5], q[6], q[7]};
            if (lane == 0) fl = {F[0], F[1], F[2], F[3]};
            fr = from_next_or(fl, F);
        } else {
            fr = from_next(fl);
        }

        // L = -(F_right - F_left)/dx - (G_top - G_bottom)/dy, x first then y (solver2d.hpp:139-178)
        Cons L;
        L.d = (0.0 + cx * fl.d) - cx * fr.d;
        L.m = (0.0 + cx * fl.m) - cx * fr.m;
        L.n = (0.0 + cx * fl.n) - cx * fr.n;
        L.e = (0.0 + cx * fl.e) - cx * fr.e;
        L.d = (L.d + cy * glo.d) - cy * ghi.d;
        L.m = (L.m + cy * glo.m) - cy * ghi.m;
        L.n = (L.n + cy * glo.n) - cy * ghi.n;
        L.e = (L.e + cy * glo.e) - cy * ghi.e;

        Cons out;
        if constexpr (STAGE == 1) {  // U1 = U + dt L(U)
            out = {va.d + dt * L.d, va.m + dt * L.m, va.n + dt * L.n, va.e + dt * L.e};
        } else {  // U = 1/2 ((U + U1) + dt L(U1))
            out = {0.5 * (ub.d + va.d + dt * L.d), 0.5 * (ub.m + va.m + dt * L.m),
                   0.5 * (ub.n + va.n + dt * L.n), 0.5 * (ub.e + va.e + dt * L.e)};
        }
        if (owner) store4(O, (long long)f * nx + ci, P, out);
        if constexpr (SEAMS) {  // the strip's edge columns, for the next stage's seam prologues
            if (lane < 2 || lane >= kWave - 2) {
                gpair* e = eout + f * (kEdgeRecord / 2);
                e[0].a = out.d;
                e[0].b = out.m;
                e[1].a = out.n;
                e[1].b = out.e;
            }
        }

        if constexpr (STAGE == 2) {  // wave speeds of U^{n+1} for the next dt (solver2d.hpp:67-83)
            const double rd = recip(out.d);
            const Prim w = prim_of(out.d, out.m, out.n, out.e, rd);
            const bool badc = !(w.d > 0.0) || !(w.p > 0.0);
            const double a = sound(w.d, w.p, rd);
            if (owner) {
                if (badc) {
                    const int idx = f * nx + i;
                    bad = idx < bad ? idx : bad;
                } else {
                    sx = tmax(sx, dabs(w.u) + a);
                    sy = tmax(sy, dabs(w.v) + a);
                }
            }
        }

        glo = ghi;
        top = topn;
        va = vb;
        vb = vn;
        wa = wb;
        wb = wn;
        if constexpr (CORR) {
            aa = ab;
            ab = an;
        }
        vn = vq;
        if constexpr (STAGE == 2) ub = uq;
    };

    // unrolled so that the carried rows are renamed instead of copied each iteration (stage 1 keeps
    // four waves per SIMD at x2; stage 2 needs more registers anyway and takes x4)
    constexpr int kUnroll = STAGE == 1 ? 2 : 4;
    int f = j0;
    for (; f + kUnroll <= j1; f += kUnroll) {
#pragma unroll
        for (int u = 0; u < kUnroll; ++u) row_step(f + u);
    }
    for (; f < j1; ++f) row_step(f);

    if constexpr (STAGE == 2) {
        publish_partial(part, tid, lane, sx, sy, bad);
    }
}

// ---------------------------------------------------------------------------------------------
// launch table
using StageFn = void (*)(const InstDesc*, Ctl*, const Task*, int, Partial*, int*, int, int);

template <int S, int L, int M>
constexpr StageFn stage_fn(int stage, bool corr) {
    return stage == 1 ? (corr ? k_stage<S, L, 1, true, M> : k_stage<S, L, 1, false, M>)
                      : (corr ? k_stage<S, L, 2, true, M> : k_stage<S, L, 2, false, M>);
}

template <int M>
StageFn pick_stage_t(int scheme, int lim, int stage, bool corr) {
    if (scheme == kRusanov) {
        switch (lim) {
            case kLimNone: return stage_fn<kRusanov, kLimNone, M>(stage, corr);
            case kLimUnlimited: return stage_fn<kRusanov, kLimUnlimited, M>(stage, corr);
            case kLimMinmod: return stage_fn<kRusanov, kLimMinmod, M>(stage, corr);
            default: return stage_fn<kRusanov, kLimVanLeer, M>(stage, corr);
        }
    }
    switch (lim) {
        case kLimNone: return stage_fn<kHllc, kLimNone, M>(stage, corr);
        case kLimUnlimited: return stage_fn<kHllc, kLimUnlimited, M>(stage, corr);
        case kLimMinmod: return stage_fn<kHllc, kLimMinmod, M>(stage, corr);
        default: return stage_fn<kHllc, kLimVanLeer, M>(stage, corr);
    }
}

// The seam-layout kernels are instantiated only in their own translation units (euler_kernels_seams.hip,
// euler_kernels_fast_seams.hip, which define NPDE_SEAMS_UNIT), compiled with the ILP-oriented machine
// scheduler: it shortens their launches by 0.5-1% at unchanged register counts, but raises the other
// layouts' (strips' stage 1 from 126 to 178 VGPRs, 4 to 2 wavefronts per SIMD).
StageFn pick_stage_seams(int scheme, int lim, int stage, bool corr);

// Dynamic LDS of a stage launch: the seam rows (seam bytes), or under an occupancy cap (reserve > 0: LDS
// bytes a workgroup must hold in total) what tops the kernel's static LDS up to the reservation, if more.
inline hipError_t stage_dynamic_lds(StageFn fn, unsigned seam, unsigned reserve, unsigned& bytes) {
    bytes = seam;
    if (reserve == 0) return hipSuccess;
    hipFuncAttributes a{};
    const hipError_t e = hipFuncGetAttributes(&a, reinterpret_cast<const void*>(fn));
    if (e != hipSuccess) return e;
    const unsigned fixed = unsigned(a.sharedSizeBytes);
    if (reserve > fixed && reserve - fixed > bytes) bytes = reserve - fixed;
    return hipSuccess;
}

#if defined(NPDE_SEAMS_UNIT)
StageFn pick_stage_seams(int scheme, int lim, int stage, bool corr) {
    return pick_stage_t<kSeams>(scheme, lim, stage, corr);
}
#else
// mode: StageMode (strips, row groups, seams)
StageFn pick_stage(int scheme, int lim, int stage, bool corr, int mode) {
    switch (mode) {
        case kRowGroups: return pick_stage_t<kRowGroups>(scheme, lim, stage, corr);
        case kSeams: return pick_stage_seams(scheme, lim, stage, corr);
        default: return pick_stage_t<kStrips>(scheme, lim, stage, corr);
    }
}
#endif

}  // namespace NPDE_ARITH_NS
}  // namespace npde
