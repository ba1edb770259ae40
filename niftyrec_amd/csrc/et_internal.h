// et_internal.h -- declarations shared by the HIP kernels (et_kernels.hip) and the
// C++ host layer (et_host.cpp) of lib_et_array_interface.so.
//
// Device data layouts
// -------------------
// ABI layout   [nx, ny, nz] x fastest: idx = x + nx*(y + ny*z)   (caller buffers)
// YF layout    axial-fastest, zero-padded source volumes read by the sweep kernels:
//              idx = (gy + y) + yp*((x + 1) + xp*(z + 1)),
//              xp = nx + 2, zp = nz + 2, yp = gy + roundup(ny, ET_T) + gy.
//              Everything outside the logical volume is zero, so trilinear
//              corners that fall one voxel outside the grid read zeros.
// U layout     rotated-frame per-view volumes written by the sweep kernels and
//              pulled by the backprojection kernels:
//              idx = v*ustride + y + yu*((x' + 1) + xp*(z' + 1)), yu = roundup(ny, ET_T),
//              with a zero border of one voxel in x' and z'.
// Sinogram     [nx, ny, V] x fastest (the ABI layout), partial projections
//              [D][V][ny][nx].
#pragma once
#include <stddef.h>
#include <stdint.h>

#define ET_T 16      // axial rows per slab in the sweep and pull kernels
#define ET_GY 16     // guard rows of mu in the YF layout (and of every volume off the fast path)
#define ET_YF_SLACK 16  // zero floats after the last column of a packed mu (its halo windows' tail)
#define ET_RMAX 10   // largest PSF radius handled inside the sweep kernels (K <= 21)
#define ET_VIEW_REC 40  // floats per view record (see EtViewRec)
#define ET_MAX_NX 512   // widest volume: sweep blocks hold one thread per rotated column (512 threads)

// Per-view record in device memory (ET_VIEW_REC floats):
//   [0..11]  forward 3x4 matrix M (row-major): sample position = M (x,y,z,1)
//   [12..23] reference backprojection matrix (negated angles, ZYX order)
//   [24..32] inverse of the 3x3 linear part of M (row-major), used to bound the
//            candidate search of the exact-adjoint gather
//   [33..39] unused
enum EtSweepMode {
    ET_MODE_PROJECT = 0,   // p = sum_z PSF(act') e^{-A} (+ bg e^{-A_tot})
    ET_MODE_BPU = 1,       // u = PSF(s e^{-A})  (kernels flipped by the host for the adjoint)
    ET_MODE_GRAD_REF = 2,  // u = PSF(s sum_{k>=z} act' e^{-A})        (reverse sweep)
    ET_MODE_GRAD_MU = 3,   // u = -s (sum_{k>=z} PSF(act') e^{-A} + bg e^{-A_tot})  (reverse sweep)
    ET_MODE_MUSUM = 4,     // per-chunk sums of mu' along each ray
};

enum EtSource {
    ET_SRC_SINO = 0,   // s read from a [V][ny][nx] array
    ET_SRC_RATIO = 1,  // s = ratio(meas, sum_d part[d]) with the MLEM epsilon rule
    ET_SRC_ONES = 2,   // s = 1 (sensitivity)
};

struct EtSweepArgs {
    int nx, ny, nz;
    int yp, xp, gy;          // YF layout of act (row stride, columns, guard rows)
    int gmu;                 // guard rows of mu in the same layout
    int yu, ustride;         // U layout of the output volumes (per view stride)
    int V, D, zlen;          // views in launch, depth chunks, z' per chunk
    int slabs;               // ceil(ny / ET_T)
    const float *views;      // [V][ET_VIEW_REC]
    const float *act;        // YF
    const float *mu;         // YF or null
    const float *psf;        // [nz][2K] row kernel, column kernel (null when R == 0)
    int K;
    float bg_act, bg_mu;     // padding used when resampling act / mu
    float bg_line;           // background term of the line integral
    const float *astart;     // [V][D][ny][nx] attenuation at the start of each chunk (D > 1)
    const float *atot;       // [V][ny][nx] total attenuation along each ray (reverse sweeps)
    const float *atot_lo;    // its low parts: atot + atot_lo is the compensated sum
    float *part;             // PROJECT: [D][V][ny][nx]; MUSUM: [V][D][ny][nx]
    float *part_lo;          // MUSUM: low parts of the compensated chunk sums
    int truncate;            // PROJECT with D == 1: clamp negatives
    int src;                 // EtSource
    const float *sino;       // ET_SRC_SINO: [V][ny][nx]
    const float *meas;       // ET_SRC_RATIO: measured counts [V][ny][nx]
    const float *rpart;      // ET_SRC_RATIO: partial projections [D_r][V][ny][nx]
    int rD;
    float eps;
    float *u;                // U layout outputs
    float *att;              // attenuation factors e^{-A(x',y,z')} of the views, [V][nz][yu][nthr] with
                             // nthr = roundup(nx, 64) (the block width); see et_launch_sweep's has_mu
                             // (row-lane kernels: [V][nz][nx][yu])
    int yl_nbuf;             // row-lane sweeps: LDS planes (2: alternate, one barrier per plane; 1: two)
    int yl_tiles;            // row-lane projector: column tiles per plane (1: the whole width per block)
};

enum EtEpilogue {
    ET_EPI_ABI_STORE = 0,    // out_abi  = acc (optionally + previous, truncated at the end)
    ET_EPI_YF_STORE = 1,     // out_yf   = acc (optionally + previous)
    ET_EPI_OSEM = 2,         // x_yf    *= max(acc,0) / sens_yf   (sens already clamped)
};

struct EtPullArgs {
    int nx, ny, nz;
    int yu, xp, ustride;     // U layout of the inputs
    int V;
    const float *views;      // [V][ET_VIEW_REC]
    const float *u;
    float bg;                // padding of the reference inverse-rotation pull
    int adjoint;             // 0: reference pull, 1: exact transpose of the forward pull
    int epilogue;            // EtEpilogue
    int accumulate;          // add the previous content of the output
    int truncate;            // clamp negatives after this launch
    float *out;              // ABI or YF output
    int yp, gy;              // YF layout of out / x / sens
    float *x;                // ET_EPI_OSEM: activity (YF), updated in place
    const float *sens;       // ET_EPI_OSEM: clamped sensitivity (YF)
    const float *u2;         // exact adjoint pull of two volumes of the same views
    float *out2;             // (u into out, u2 into out2; null otherwise)
};

// MLEM ratio of one problem from its depth-chunk partial projections.
struct EtRatioArgs {
    const float *meas;       // [V][ny][nx]
    const float *part;       // [D][V][ny][nx]
    int D;
    long long VP;            // V * ny * nx
    float eps;
    float *out;
};

// ---- launchers (et_kernels.hip) -------------------------------------------------
// All return hipSuccess or the launch error; all are asynchronous on `stream`.
typedef struct ihipStream_t *et_stream_t;
typedef struct ihipEvent_t *et_event_t;

// Kernel-timed projector launches: a profiling scope offers an event pair for
// the next row-lane projector launch of this thread (et_offer_launch_events);
// that launch records its own dispatch start and end in them
// (hipExtLaunchKernelGGL) and takes them (et_claim_launch_events).
// et_withdraw_launch_events clears the offer and returns whether a launch took it.
void et_offer_launch_events(et_event_t start, et_event_t stop);
bool et_claim_launch_events(et_event_t *start, et_event_t *stop);
bool et_withdraw_launch_events();

// has_mu: 0 no attenuation; 1 attenuation from the resampled a.mu; 2 (PROJECT)
// as 1, and the attenuation factors stored in a.att; 3 (PROJECT, BPU) the
// factors read from a.att instead of resampling mu.
int et_launch_sweep(int mode, int R, int has_mu, int has_act, const EtSweepArgs &a,
                    et_stream_t stream);
int et_launch_pull(const EtPullArgs &a, et_stream_t stream);
int et_launch_pull_general(const EtPullArgs &a, et_stream_t stream);
int et_launch_pack_yf(const float *abi, float *yf, int nx, int ny, int nz, int yp, int gy,
                      et_stream_t stream);
int et_launch_unpack_yf(const float *yf, float *abi, int nx, int ny, int nz, int yp, int gy,
                        et_stream_t stream);
int et_launch_resample_general(const float *src, int nx, int ny, int nz, const float *M12,
                               float bg, float *dst, int dst_yf, int yp, int gy,
                               et_stream_t stream);
// Attenuation at the start of every chunk (astart, rounded) and along whole
// rays (atot + atot_lo), compensated sums of the chunk sums csum + csum_lo.
int et_launch_chunk_scan(const float *csum, const float *csum_lo, float *astart, float *atot, float *atot_lo,
                         int V, int D, int P, et_stream_t stream);
int et_launch_reduce_parts(const float *part, int D, int V, int P, float *sino, int truncate,
                           et_stream_t stream);
int et_launch_bg_only(const float *atot, const float *atot_lo, float bg, int n, float *sino, int truncate,
                      et_stream_t stream);
int et_launch_conv_axis(const float *src, float *dst, int nx, int ny, int nz, long sx, long sy,
                        long sz, long off, int axis, const float *sep, int K, int flip,
                        et_stream_t stream);
int et_launch_truncate(float *buf, size_t n, et_stream_t stream);
int et_launch_fill(float *buf, size_t n, float value, et_stream_t stream);
int et_launch_clamp_min(float *buf, size_t n, float lo, et_stream_t stream);
int et_launch_yf_clamp_min(float *yf, int nx, int ny, int nz, int yp, int gy, float lo,
                           et_stream_t stream);
int et_launch_em_update_flat(float *x, const float *b, const float *s, size_t n, et_stream_t stream);
int et_launch_gather_planes(float *dst, const float *src, const int *order, int V, size_t P,
                            et_stream_t stream);
int et_launch_ratio_parts(const float *meas, const float *part, int D, size_t VP, float eps, float *out,
                          et_stream_t stream);
int et_launch_ratio(const float *meas, const float *proj, float *out, size_t n, float eps,
                    et_stream_t stream);
// Grouped launches over several problems.  args: device array of n argument
// records; tab: device array of n int4 {first block, gx, gy, 0} with first
// blocks ascending multiples of 8 (each problem's gx*gy blocks, padded to a
// multiple of 8); blocks: total grid.  Sweeps: project or BPU mode, every
// problem with PSF radius R, the same mu presence and nthr = roundup(nx, 64)
// threads; gx = ceil(ny / et_sweep_slab_rows(mode, R)), gy = V*D.  Pull:
// gx = ceil(nx*nz / 256), gy = ceil(ny / ET_T).  Ratio: gx = ceil(VP/256), gy = 1.
int et_launch_sweep_group(int mode, int R, int has_mu, int nthr, const EtSweepArgs *args, const void *tab, int n,
                          int blocks, et_stream_t stream);
int et_launch_pull_group(int epilogue, const EtPullArgs *args, const void *tab, int n, int blocks,
                         et_stream_t stream);
int et_launch_ratio_group(const EtRatioArgs *args, const void *tab, int n, int blocks, et_stream_t stream);
// Partial line integrals of one rotated view (ABI layouts): part[z] = sum of
// what lies beyond plane z, sino = the whole ray; act or mu may be null.
int et_launch_partial_line(const float *act, const float *mu, const float *bgimg, float bg, int nx, int ny, int nz,
                           float *sino, float *part, et_stream_t stream);
// Content fingerprints of p0 and (when not null) p1, n floats each: the keys
// under which the caches remember the attenuation map and the activity.  The
// launch stores et_fingerprint_blocks(n) partial sums per array in parts (p0's,
// then p1's; at most 2 * ET_FP_BLOCKS); a key is the sum modulo 2^64 of its
// array's partials.
#define ET_FP_BLOCKS 1024
int et_fingerprint_blocks(size_t n);
int et_launch_fingerprints(const float *p0, const float *p1, size_t n, unsigned long long *parts,
                           et_stream_t stream);
// Row-lane kernels of the OSEM subset loop (et_yl.h).  Grouped launches only
// (tab as et_launch_sweep_group): sweeps kind 0 projection, 1 rotated-frame
// backprojection, 3 exact attenuation gradient (reads the projection's partial
// sums of the same chunking), gx = depth chunks, gy = views, nthr = 16 * W
// threads, `lds` bytes of LDS; kinds 2 and 4 the tiled projection and the tiled
// gradient (R > 0), gx = depth chunks x tiles (tile fastest), 512 threads, NR
// unused; pull gx = ceil(nx * nz / 16),
// gy = 1, 256 threads.  RPL (rows per lane) in {2, 4, 6, 8}, NR (column
// rounds) in {1, 2, 3}.
// Tiled row-lane projector: columns per tile, LDS bytes of a block (two
// planes of 64 + 2R columns of rows + 4 floats).
#define ET_YL_TILE 64
inline constexpr size_t et_yl_tile_lds_bytes(int R, int rows)
{
    return (size_t)2 * (ET_YL_TILE + 2 * R) * (rows + 4) * sizeof(float);
}
int et_launch_yl_sweep(int kind, int R, int RPL, int NR, int nthr, size_t lds, const EtSweepArgs *args,
                       const void *tab, int n, int blocks, et_stream_t stream);
int et_launch_yl_pull(int epilogue, int RPL, const EtPullArgs *args, const void *tab, int n, int blocks,
                      et_stream_t stream);
// Attenuation factors of a.V views for the sweep kernels' OSEM cache: a.att
// [v][z'][yu][roundup(nx, 64)] = e^{-A} from a.mu (YF); 16-row slabs.
int et_launch_att_factors(const EtSweepArgs &a, et_stream_t stream);
// Exact adjoint pull of the row-lane rotated-frame volumes into a.out (YF layout,
// accumulate / truncate as ET_EPI_YF_STORE); grid as the row-lane pull.  nsrc 2:
// a.u into a.out and a.u2 into a.out2 with one set of candidate weights.
int et_launch_yl_pull_adj(int RPL, int nsrc, const EtPullArgs *args, const void *tab, int n, int blocks,
                          et_stream_t stream);
// Attenuation factors of a.V views into a.att (row-lane layout) from a.mu (YF).
int et_launch_yl_factors(int RPL, const EtSweepArgs &a, et_stream_t stream);
// sino[v][y][x] = sum over d < D of part[d][v][x][y] (compensated), truncated
// at 0 if asked: row-lane partial sums (V views, rows per column) in the ABI layout.
int et_launch_yl_parts_to_sino(const float *part, int D, int V, int nx, int ny, int rows, int truncate, float *sino,
                               et_stream_t stream);
// dst[v][x][y] = src[order[v]][y][x] (y < ny), 0 for ny <= y < rows (order null: v).
int et_launch_gather_planes_t(float *dst, const float *src, const int *order, int V, int nx, int ny, int rows,
                              et_stream_t stream);
// Axial rows per block of a sweep launch (mode, PSF radius).
int et_sweep_slab_rows(int mode, int R);

// Fisher information of a grid of points (reference emission-lib/_et.cpp:531-746).
struct EtFisherArgs {
    const int *coords;   // [V][n][3] grid points rotated into each camera frame (truncated)
    const int *coords0;  // [n][3] grid points in the image frame (prior)
    const float *sino;   // [V][sny][snx] projection of the activity
    const float *psf;    // [planes][py][px]
    float *fisher;       // [n][n]
    float *prior;        // [n][n] or null
    int n, V, snx, sny, px, py, nz;
    int prior_rule;      // 1: |d| <= 1 with the diagonal; 3: |d| <= 3 without it
};
// Every pair i <= j of grid points in one thread: F[i][j] = F[j][i] (and the prior).
int et_launch_fisher_pairs(const EtFisherArgs &a, et_stream_t stream);
