// Device-resident descriptors shared by the HIP kernels and the host side of the C-ABI.
//
// HBM layout, per instance: the state U and the stage buffer U1 are each four planes
// (rho, rho*u, rho*v, E) of nx*ny binary64 cells, row-major with j as the row index and no ghost
// cells (ghosts are virtual: a stencil index outside the grid is remapped to the interior cell the
// reference's ghost fill would copy, grid2d.hpp:43-66).  Instances of a batch live back to back in
// one allocation, so a batch of 64 x 512^2 is 1.07 GB of U + U1.
//
// Work decomposition: a "task" is one wavefront's share of one RK stage -- a strip of columns and a
// band of rows, marched upward row by row.  Tasks of all instances form one list; a stage launch
// covers the whole list, and waves of instances that are not stepping leave at once.  Three forms
// (StageMode):
//   strips     60 interior columns per wavefront (64 lanes minus a 2-column halo on each side);
//   row groups periodic x, width 64*W: a task is a band of all the columns, run by one workgroup of
//              W wavefronts that exchange edge columns through LDS, one partial per wavefront;
//   seams      width a multiple of 64: 64 owned columns per wavefront; each wavefront first solves the
//              two strip boundaries ("seams") of its band, from edge columns the previous stage wrote.
#pragma once
#include <cstdint>

namespace npde {

constexpr int kWave = 64;
constexpr int kHalo = 2;
constexpr int kStripCols = kWave - 2 * kHalo;  // 60 interior columns per wavefront
constexpr int kMaxRowWaves = 8;                  // row-group mode: widths up to 8 * 64 = 512 columns
constexpr int kMaxLevels = 4;                  // trajectory resolutions per run
constexpr int kNoBad = 0x7fffffff;             // "no non-physical cell" marker

enum Mode : int { kModeHost = 0, kModeAuto = 1 };
enum StageMode : int { kStrips = 0, kRowGroups = 1, kSeams = 2 };

// Edge columns (seam mode): the stage kernels also write the conserved values of columns 0, 1 (left
// edge) and 62, 63 (right edge) of every 64-column strip, so that the seam solves of the next stage read
// them contiguously instead of gathering four columns per row from the full state.  Per state buffer
// (U or U1) and instance: a left array and a right array, each (nx/64) x ny x 8 doubles; column e of
// the pair (0 or 1), variable q at [(strip*ny + j)*8 + e*4 + q].
constexpr int kEdgeRecord = 8;

struct InstDesc {
    double* U;        // 4 planes: state (stage-1 stencil input, stage-2 output)
    double* U1;       // 4 planes: stage buffer (stage-1 output, stage-2 stencil input)
    long long plane;  // cells per plane = nx*ny
    int nx, ny, bc_x, bc_y;
    double dx, dy, cx, cy, cfl;  // cx = 1/dx, cy = 1/dy (the reference multiplies by 1/dx)
    int task0, ntasks;           // this instance's slice of the stage task list
    int ftask0, nftasks;         // this instance's slice of the frame task list
    const double* corr_x;        // optional learned x-face flux increments (4 planes), or null
    const double* corr_y;        // optional learned y-face flux increments (4 planes), or null
    double* edge[2];             // seam mode: edge columns of U [0] and U1 [1] (left array, then right), else null
    int nlevels;
    int lvl_m[kMaxLevels];       // coarse size (columns) of each trajectory level
    int lvl_f[kMaxLevels];       // coarsening factor nx / m
    float* traj[kMaxLevels];     // (frames, 4, ny/f, m) float32 per level, or null
    double* ftimes;              // time at which each frame was written (nout + 1), or null
    double net_rate_limit;       // flux-network runs: switch the network off for the instance when its CFL
                                 // signal-speed sum exceeds this multiple of its value at the first frame (0: never)
};

// Mutable per-instance run state.  Written by k_finish (one wavefront per instance), by the last frame
// workgroup of an instance's landing (k_frame), by the last workgroup of a flux-network launch (releasing
// the instances that held), or by host set-up between launches; read by every wavefront of the
// following launches.
struct Ctl {
    double t, dt, t_start, t_end;
    double sx, sy;        // max |u|+a and max |v|+a of the current state
    long long steps;
    long long step_limit; // stop (done) instead of stepping once steps reaches this
    int k;                // index of the current landing target (0 = t_start, k = frame k)
    int nout;             // frames after t_start
    int active;           // a step with `dt` is in flight
    int mode;             // kModeHost (host supplies dt) or kModeAuto (device decides)
    int record;           // write block-averaged frames when a target is reached
    int pending;          // frame index waiting to be written, or -1
    int done;             // all targets reached (or failed)
    int error;            // 1: non-physical state met when a step was needed
    int bad;              // row-major index of the first non-physical cell, or kNoBad
    int hold;             // flux-network runs: landed on a frame, waiting for the batch's network call
    int fdone;            // frame workgroups of the pending frame that have written their part (0 between launches)
    int net_off;          // flux-network runs: the network is switched off for this instance (zero factors)
    double rate0;         // flux-network runs: sx/dx + sy/dy at the run's first frame landing
};

// Device counters of a run, one int each: instances done; (flux-network runs) instances holding at a
// frame landing -- the network runs, and the holding instances move on, once every instance that is not
// done holds -- and workgroups of that network launch that have finished (its last one releases them).
enum RunCounter : int { kCountDone = 0, kCountHeld = 1, kCountNetGroups = 2, kRunCounters = 3 };

struct Partial {
    double sx, sy;
    int bad;
    int pad;
};

struct Task {  // one wavefront of one RK stage
    int inst;
    int i0;  // first interior column of the strip
    int j0;  // first row of the band
    int j1;  // one past the last row of the band
};

struct NetTask {  // one workgroup of the flux-correction network: a 16 x 16 tile of one instance
    int inst;
    int x0, y0;  // first column and row of the tile
    int pad;
};

struct FTask {  // one workgroup of the frame (block-average) kernel
    int inst;
    int level;
    int c0;   // first coarse cell (row-major) handled by the workgroup
    int pad;
};

}  // namespace npde
