// et_host.cpp -- C++ host layer of lib_et_array_interface.so.
//
// Validates and expands the NiftyRec array-interface arguments, builds the
// per-view rotation records with the reference's single-precision arithmetic,
// owns device workspaces (plans), selects the fast (axial rotation) or general
// path and sequences the HIP kernels of et_kernels.hip on one stream.  Every
// HIP call is checked; errors are returned as niftyrec status codes and the
// process is never terminated.
#include <hip/hip_runtime.h>

#include <algorithm>
#include <cmath>
#include <cstdio>
#include <cstdlib>
#include <cstring>
#include <mutex>
#include <new>
#include <string>
#include <map>
#include <memory>
#include <tuple>
#include <vector>

#include "../../include/et_array_interface.h"
#include "et_internal.h"

#pragma clang fp contract(off)

#define ET_VERSION "niftyrec-mi355x 0.1 (gfx950)"

namespace {

int g_device = 0;
std::mutex g_device_mutex;

// ---------------------------------------------------------------------------
// Status handling
// ---------------------------------------------------------------------------

struct EtError {
    int code;
};

inline void check_hip(hipError_t e, int code)
{
    if (e != hipSuccess) throw EtError{code};
}
inline void check_launch(int e)
{
    if (e != (int)hipSuccess) throw EtError{niftyrec_error_kernel};
}

// Selects the library's device; returns niftyrec_error_nogpuinstalled when
// there is no usable HIP device.
int select_device()
{
    int count = 0;
    if (hipGetDeviceCount(&count) != hipSuccess || count <= 0) return niftyrec_error_nogpuinstalled;
    int dev;
    {
        std::lock_guard<std::mutex> lk(g_device_mutex);
        dev = g_device;
    }
    if (dev < 0 || dev >= count) return niftyrec_error_parameters;
    if (hipSetDevice(dev) != hipSuccess) return niftyrec_error_nogpuinstalled;
    return niftyrec_success;
}

// ---------------------------------------------------------------------------
// Rotation records (single precision, reference operation order)
// ---------------------------------------------------------------------------

typedef float Mat4[16];

// reg-lib/_reg_affineTransformation.cpp:18-30: R[i][j] = A[i][0]B[0][j] + ... + A[i][3]B[3][j].
void mat_mul(const float *A, const float *B, float *R)
{
    float t[16];
    for (int i = 0; i < 4; i++)
        for (int j = 0; j < 4; j++)
            t[4 * i + j] = A[4 * i + 0] * B[0 * 4 + j] + A[4 * i + 1] * B[1 * 4 + j] +
                           A[4 * i + 2] * B[2 * 4 + j] + A[4 * i + 3] * B[3 * 4 + j];
    std::memcpy(R, t, sizeof(t));
}

// Centred rotation in one of six axis orders, emission-lib/_et_common.cpp:35-160.
void rotation_matrix(float *M, float ax, float ay, float az, float cx, float cy, float cz, int order)
{
    const float sx = std::sin(ax), cxs = std::cos(ax), sy = std::sin(ay), cys = std::cos(ay);
    const float sz = std::sin(az), czs = std::cos(az);
    const float Rz[16] = {czs, -sz, 0, 0, sz, czs, 0, 0, 0, 0, 1, 0, 0, 0, 0, 1};
    const float Ry[16] = {cys, 0, sy, 0, 0, 1, 0, 0, -sy, 0, cys, 0, 0, 0, 0, 1};
    const float Rx[16] = {1, 0, 0, 0, 0, cxs, -sx, 0, 0, sx, cxs, 0, 0, 0, 0, 1};
    const float Tm[16] = {0, 0, 0, cx, 0, 0, 0, cy, 0, 0, 0, cz, 0, 0, 0, 1};
    float T1[16];
    switch (order) {
    case 1: mat_mul(Rz, Rx, T1); mat_mul(Ry, T1, M); break;
    case 2: mat_mul(Rx, Ry, T1); mat_mul(Rz, T1, M); break;
    case 3: mat_mul(Rz, Ry, T1); mat_mul(Rx, T1, M); break;
    case 4: mat_mul(Rx, Rz, T1); mat_mul(Ry, T1, M); break;
    case 5: mat_mul(Ry, Rz, T1); mat_mul(Rx, T1, M); break;
    default: mat_mul(Ry, Rx, T1); mat_mul(Rz, T1, M); break;
    }
    float Tr[16];
    mat_mul(M, Tm, Tr);
    M[3] = cx - Tr[3];
    M[7] = cy - Tr[7];
    M[11] = cz - Tr[11];
}

// True when y' maps exactly to y: rows 1 = (0,1,0,0) and no y coupling.
bool is_axial(const float *M)
{
    return M[1] == 0.f && M[9] == 0.f && M[4] == 0.f && M[5] == 1.f && M[6] == 0.f && M[7] == 0.f;
}

void inverse3(const float *M, float *out9)
{
    double a[9] = {M[0], M[1], M[2], M[4], M[5], M[6], M[8], M[9], M[10]};
    double c00 = a[4] * a[8] - a[5] * a[7], c01 = a[5] * a[6] - a[3] * a[8], c02 = a[3] * a[7] - a[4] * a[6];
    double det = a[0] * c00 + a[1] * c01 + a[2] * c02;
    if (det == 0.0) det = 1e-30;
    double inv[9] = {c00 / det,
                     (a[2] * a[7] - a[1] * a[8]) / det,
                     (a[1] * a[5] - a[2] * a[4]) / det,
                     c01 / det,
                     (a[0] * a[8] - a[2] * a[6]) / det,
                     (a[2] * a[3] - a[0] * a[5]) / det,
                     c02 / det,
                     (a[1] * a[6] - a[0] * a[7]) / det,
                     (a[0] * a[4] - a[1] * a[3]) / det};
    for (int i = 0; i < 9; i++) out9[i] = (float)inv[i];
}

void make_record(const float *M, const float *Mi, float *rec)
{
    std::memset(rec, 0, sizeof(float) * ET_VIEW_REC);
    std::memcpy(rec, M, sizeof(float) * 12);
    std::memcpy(rec + 12, Mi, sizeof(float) * 12);
    inverse3(M, rec + 24);
}

void identity_record(float *rec)
{
    float I[16] = {1, 0, 0, 0, 0, 1, 0, 0, 0, 0, 1, 0, 0, 0, 0, 1};
    make_record(I, I, rec);
}

// Expanded cameras (3 x V, [axis*V + v]) and centres (3 x V).
struct Cameras {
    int V = 0;
    std::vector<float> ang, ctr;
};

// Expansion of the cameras argument, emission-lib/_et_array_interface.cpp:396-440:
// one column -> rotation about axis 1; six columns carry explicit centres.
int expand_cameras(const float *cameras, int V, int cols, const int *ref_size, Cameras &out)
{
    if (!(cols == 1 || cols == 3 || cols == 6) || V <= 0) return niftyrec_error_parameters;
    out.V = V;
    out.ang.assign(3 * (size_t)V, 0.f);
    out.ctr.assign(3 * (size_t)V, 0.f);
    if (cols == 1)
        for (int c = 0; c < V; c++) out.ang[V + c] = cameras[c];
    else
        std::memcpy(out.ang.data(), cameras, sizeof(float) * 3 * V);
    if (cols == 6)
        std::memcpy(out.ctr.data(), cameras + 3 * V, sizeof(float) * 3 * V);
    else
        for (int c = 0; c < V; c++)
            for (int k = 0; k < 3; k++) out.ctr[k * V + c] = (float)((ref_size[k] - 1) / 2.0);
    return niftyrec_success;
}

// Smallest odd multiple of 16 floats (64 bytes) not below n.
int odd_segments(int n)
{
    int s = (n + 15) / 16;
    if (s % 2 == 0) s++;
    return 16 * s;
}


// ---------------------------------------------------------------------------
// Device buffers
// ---------------------------------------------------------------------------

struct DevBuf {
    float *p = nullptr;
    size_t n = 0;
    DevBuf() = default;
    DevBuf(const DevBuf &) = delete;
    DevBuf &operator=(const DevBuf &) = delete;
    ~DevBuf() { release(); }
    void release()
    {
        if (p) (void)hipFree(p);
        p = nullptr;
        n = 0;
    }
    // Grows to at least `count` floats; new memory is zero-filled when `zero`,
    // by a kernel on `stream`, so the fill is ordered before the stream's later
    // work whatever kind of stream the caller uses.
    float *ensure(size_t count, bool zero, hipStream_t stream = nullptr)
    {
        if (count <= n && p) return p;
        release();
        if (hipMalloc((void **)&p, std::max<size_t>(count, 1) * sizeof(float)) != hipSuccess) {
            p = nullptr;
            throw EtError{niftyrec_error_allocgpu};
        }
        n = count;
        if (zero) check_launch(et_launch_fill(p, count, 0.f, (et_stream_t)stream));
        return p;
    }
};

}  // namespace

// ---------------------------------------------------------------------------
// Plan
// ---------------------------------------------------------------------------

// One grouped launch of the row-lane OSEM kernels: argument records and block
// table at byte offsets of a device table buffer.
struct YlLaunch {
    int kind;                 // YL_PROJECT, YL_RATIO, YL_BPU, YL_PULL_OSEM, YL_PULL_STORE, YL_GRAD, YL_PULL_ADJ(2)
    int R, RPL, NR, nthr;     // sweep template and block shape (pull: RPL)
    size_t lds;               // dynamic LDS bytes (sweeps)
    size_t args_off, tab_off;
    int n, blocks;
    double bytes;             // algorithmic bytes (the profiling scopes' figure)
};
enum {
    YL_PROJECT = 0, YL_RATIO = 1, YL_BPU = 2, YL_PULL_OSEM = 3, YL_PULL_STORE = 4, YL_GRAD = 5, YL_PULL_ADJ = 6,
    YL_PULL_ADJ2 = 7  // exact adjoint pull of two volumes (u, u2) into two outputs
};
// Table slots of the single operations (each keeps its last upload).
enum { YL_OP_PROJECT = 0, YL_OP_BACKPROJECT = 1, YL_OP_GRADIENT = 2, YL_OP_BACKWARD = 3, YL_OP_SLOTS = 4 };

struct et_plan {
    int device = 0;
    int nx = 0, ny = 0, nz = 0, V = 0;
    int K = 0, R = 0;
    float bg = 0.f, bga = 0.f;
    bool fast = true;
    std::vector<float> recs;        // V * ET_VIEW_REC
    std::vector<float> sep_h;       // nz * 2K, convolution kernels
    DevBuf d_recs, d_ident, d_sep, d_sep_flip;
    // layouts
    int gy = ET_GY, gmu = ET_GY, yp = 0, xp = 0, zp = 0, yu = 0, slabs = 0;
    size_t yf_elems = 0, ustride = 0, P = 0, nvox = 0;
    // workspaces
    DevBuf act_yf, mu_yf, rot_act, rot_mu, u, part, ratio, csum, csum_lo, astart, atot, atot_lo, tmp, tmp2, em_sino,
        em_vol, att;
    // OSEM caches (valid while osem_valid; keyed on the subset count, the
    // chunking, eps and the attenuation map's key: its content fingerprint, or
    // its address when mu_check is off)
    DevBuf x_yf, sens, osem_recs, osem_meas, osem_E, osem_order;
    bool osem_valid = false;
    bool osem_yl = false;           // the cached data are the row-lane kernels'
    int osem_S = 0, osem_D = 0, osem_Dp = 0, osem_Db = 0;
    bool osem_mu_yf = false;
    std::vector<int> osem_start;    // subset s = views osem_start[s] .. osem_start[s+1]-1 (subset order)
    const float *osem_cur_mu = nullptr;
    float osem_cur_eps = 0.f;
    bool osem_has_mu = false;
    unsigned long long osem_mu_key = 0;
    float osem_eps = 0.f;
    int mu_check = 1;               // 1: fingerprint mu's content on every OSEM call
    // row-lane OSEM kernels (et_yl.h): fast-path plans with attenuation whose
    // rotated plane fits the LDS of one block; rows = yu = 16 * yl_rpl
    bool yl = false;
    int yl_rpl = 0, yl_nr = 0, yl_w = 0, yl_nbuf = 0;  // rows per lane, column rounds, columns per round, LDS planes
    size_t yl_lds = 0, yl_lds_bpu = 0;                 // LDS bytes of a projector / backprojection block
    int yl_tiles = 1;                                  // projector column tiles (> 1: the tiled projector)
    int yl_tile_n = 1;                                 // tiles of the tiled projector where it applies, else 1
    size_t yl_lds_full = 0, yl_lds_tile = 0;           // LDS of a full-width / tiled projector block
    DevBuf osem_tables;                                // argument records and block tables of the OSEM launches
    // row-lane kernels for the single operations of a created plan (projection,
    // backprojection, exact attenuation gradient with attenuation): factors of
    // every view in view order, cached on the attenuation map's key
    bool yl_ops = false;
    DevBuf fac, sino_t, op_acc, op_acc2;  // (op_acc, op_acc2: pull accumulators, YF layout)
    DevBuf op_tables[YL_OP_SLOTS];        // per operation: its last table, and that table's bytes
    std::vector<char> op_tables_h[YL_OP_SLOTS];
    bool fac_valid = false;
    unsigned long long fac_key = 0;
    // the last projection's partial sums [fwd_D][V][x'][rows] and the content
    // keys of its activity and map: an exact gradient of the same pair seeds
    // its reverse sweeps from them instead of projecting again
    DevBuf fwd_part;
    bool fwd_valid = false;
    unsigned long long fwd_act = 0, fwd_mu = 0;
    int fwd_D = 0, fwd_tiles = 0;  // (fwd_D: the depth chunks the projection asked yl_chunks for)
    std::vector<struct YlLaunch> yl_launches;          // per subset: projection, ratio, backprojection, pull
    std::vector<const void *> osem_sig;                // buffers yl_launches refer to
    DevBuf fp_dev;                  // fingerprint partial sums (two arrays of ET_FP_BLOCKS)
    unsigned long long *fp_host = nullptr;  // their pinned host copy
    // tuning
    int tune_D = 0, tune_vb = 0;
    // profiling
    unsigned prof = 0;              // bit c set: launches of profiling category c are timed
    int prof_every = 1;             // of those, every prof_every-th launch (counted in prof_seen)
    long long prof_seen[ET_PROF_COUNT] = {0};
    std::vector<hipEvent_t> ev_pool;  // events reused by the profiling scopes
    struct Ev {
        int cat;
        double bytes;
        hipEvent_t a, b;
    };
    std::vector<Ev> events;
    double prof_ms[ET_PROF_COUNT] = {0}, prof_bytes[ET_PROF_COUNT] = {0};
    int prof_n[ET_PROF_COUNT] = {0};
    hipStream_t st = nullptr;

    ~et_plan()
    {
        if (fp_host) (void)hipHostFree(fp_host);
        for (auto &e : events) {
            (void)hipEventDestroy(e.a);
            (void)hipEventDestroy(e.b);
        }
        for (hipEvent_t e : ev_pool) (void)hipEventDestroy(e);
    }
};

namespace {

int nthreads_for(int nx) { return ((nx + 63) / 64) * 64; }

int build_plan(et_plan *p, int nx, int ny, int nz, const Cameras &cams, const float *psf, int psf_h,
               float bg, float bga)
{
    p->nx = nx;
    p->ny = ny;
    p->nz = nz;
    p->V = cams.V;
    p->bg = bg;
    p->bga = bga;
    p->K = (psf && psf_h > 0) ? psf_h : 0;
    p->R = p->K > 0 ? (p->K - 1) / 2 : 0;
    p->recs.assign((size_t)cams.V * ET_VIEW_REC, 0.f);
    p->fast = true;
    for (int v = 0; v < cams.V; v++) {
        float M[16], Mi[16];
        const int V = cams.V;
        rotation_matrix(M, cams.ang[v], cams.ang[V + v], cams.ang[2 * V + v], cams.ctr[v], cams.ctr[V + v],
                        cams.ctr[2 * V + v], 0);
        rotation_matrix(Mi, -cams.ang[v], -cams.ang[V + v], -cams.ang[2 * V + v], cams.ctr[v], cams.ctr[V + v],
                        cams.ctr[2 * V + v], 5);
        make_record(M, Mi, &p->recs[(size_t)v * ET_VIEW_REC]);
        if (!is_axial(M) || !is_axial(Mi)) p->fast = false;
    }
    // the fast kernels sample with zero padding; a non-zero background takes the
    // general path, whose resampler reproduces the reference padding rule
    if (p->R > ET_RMAX || bg != 0.f || bga != 0.f) p->fast = false;
    // layouts
    p->slabs = (ny + ET_T - 1) / ET_T;
    p->xp = nx + 2;
    p->zp = nz + 2;
    // Row-lane OSEM kernels: rows = 16 * RPL >= ny with RPL in {2, 4, 6, 8},
    // at most 3 rounds of up to 64 columns, and one or two LDS planes of
    // (nx + 2R) x (rows + 4) floats within the device's per-block LDS.
    p->yl = false;
    if (p->fast) {
        const int rows32 = ((ny + 31) / 32) * 32;
        const int rpl = rows32 / 16, nr = (nx + 63) / 64;
        if (rpl <= 8 && nr <= 3) {
            const size_t plane = (size_t)(nx + 2 * p->R) * (rows32 + 4) * sizeof(float);
            int dev = 0, lim = 0;
            if (hipGetDevice(&dev) != hipSuccess ||
                hipDeviceGetAttribute(&lim, hipDeviceAttributeMaxSharedMemoryPerBlock, dev) != hipSuccess)
                lim = 0;
            const int nbuf = p->R == 0 ? 1 : (2 * plane <= (size_t)lim ? 2 : (plane <= (size_t)lim ? 1 : 0));
            if (nbuf > 0) {
                p->yl = true;
                p->yl_rpl = rpl;
                p->yl_nr = nr;
                p->yl_w = (((nx + nr - 1) / nr) + 3) / 4 * 4;
                p->yl_nbuf = nbuf;
                p->yl_lds = p->yl_lds_full = p->R == 0 ? 0 : nbuf * plane;
                p->yl_lds_bpu = p->R == 0 ? 0 : plane;
                // tiled projector where the volume is wider than one tile and
                // two tiled blocks fit a CU's LDS (the point of tiling)
                const size_t tl = et_yl_tile_lds_bytes(p->R, rows32);
                int per_cu = 0;
                if (hipDeviceGetAttribute(&per_cu, hipDeviceAttributeMaxSharedMemoryPerMultiprocessor, dev) !=
                    hipSuccess)
                    per_cu = 0;
                if (p->R > 0 && nx > ET_YL_TILE && 2 * tl <= (size_t)per_cu) {
                    p->yl_tile_n = p->yl_tiles = (nx + ET_YL_TILE - 1) / ET_YL_TILE;
                    p->yl_lds = p->yl_lds_tile = tl;
                }
            }
        }
    }
    const int yrows = p->yl ? 16 * p->yl_rpl : p->slabs * ET_T;  // axial rows held per column
    // YF layout.  Guard rows: the activity-like volumes are gathered in windows
    // of rows y0 - R4 .. y0 + TS + R4 - 1 (R4 = R rounded up to 4); with R4
    // guard rows every 16-row slab's window starts on a 64-byte segment
    // boundary (a 32-row window touches two segments instead of three).
    // Attenuation is mostly read in runs of rows y0 .. y0 + TS - 1, aligned
    // with 16 guard rows; its windows with R4 halo rows may run up to R4 rows
    // into the next column's (zero) guard rows, and its buffer carries
    // ET_YF_SLACK zero floats after the last column for them.  Row stride: one
    // for both (one column offset addresses either) and an odd number of
    // segments: column strides that are multiples of 128 bytes concentrate the
    // gathers' neighbouring columns on fewer memory channels (the projector ran
    // 12% slower at 160 floats than at 144 or 176).  Plans off the fast path
    // use 16 guard rows and tail rows throughout.
    const int rows = yrows;
    p->gmu = ET_GY;
    if (p->fast) {
        const int r4 = (p->R + 3) & ~3;
        p->gy = r4;
        p->yp = odd_segments(std::max(r4 + rows + r4, ET_GY + rows));
    } else {
        p->gy = ET_GY;
        p->yp = odd_segments(ET_GY + rows + ET_GY);
    }
    p->yu = rows;
    p->yf_elems = (size_t)p->yp * p->xp * p->zp;
    p->ustride = (size_t)p->yu * p->xp * p->zp;
    p->P = (size_t)nx * ny;
    p->nvox = (size_t)nx * ny * nz;
    // device copies
    p->d_recs.ensure(p->recs.size(), false);
    check_hip(hipMemcpy(p->d_recs.p, p->recs.data(), sizeof(float) * p->recs.size(), hipMemcpyHostToDevice),
              niftyrec_error_transfergpu);
    std::vector<float> ident(ET_VIEW_REC);
    identity_record(ident.data());
    p->d_ident.ensure(ET_VIEW_REC, false);
    check_hip(hipMemcpy(p->d_ident.p, ident.data(), sizeof(float) * ET_VIEW_REC, hipMemcpyHostToDevice),
              niftyrec_error_transfergpu);
    if (p->K > 0) {
        // PSF separation, emission-lib/_et.cpp:192-215: row kernel = central row,
        // column kernel = central column / centre value.
        const int h = p->K, c = (h - 1) / 2;
        p->sep_h.assign((size_t)nz * 2 * h, 0.f);
        std::vector<float> flip((size_t)nz * 2 * h);
        for (int n = 0; n < nz; n++) {
            const float *pl = psf + (size_t)h * h * n;
            const float norm = pl[c * h + c];
            for (int i = 0; i < h; i++) {
                p->sep_h[(size_t)2 * h * n + i] = pl[c * h + i];
                p->sep_h[(size_t)2 * h * n + h + i] = pl[c + i * h] / norm;
            }
            for (int i = 0; i < h; i++) {
                flip[(size_t)2 * h * n + i] = p->sep_h[(size_t)2 * h * n + (h - 1 - i)];
                flip[(size_t)2 * h * n + h + i] = p->sep_h[(size_t)2 * h * n + h + (h - 1 - i)];
            }
        }
        p->d_sep.ensure(p->sep_h.size(), false);
        p->d_sep_flip.ensure(flip.size(), false);
        check_hip(hipMemcpy(p->d_sep.p, p->sep_h.data(), sizeof(float) * p->sep_h.size(), hipMemcpyHostToDevice),
                  niftyrec_error_transfergpu);
        check_hip(hipMemcpy(p->d_sep_flip.p, flip.data(), sizeof(float) * flip.size(), hipMemcpyHostToDevice),
                  niftyrec_error_transfergpu);
    }
    return niftyrec_success;
}

// ---- profiling --------------------------------------------------------------

hipEvent_t pool_event(et_plan *p)
{
    if (!p->ev_pool.empty()) {
        hipEvent_t e = p->ev_pool.back();
        p->ev_pool.pop_back();
        return e;
    }
    hipEvent_t e = nullptr;
    return hipEventCreate(&e) == hipSuccess ? e : nullptr;
}

// Times the launches issued during its lifetime when category `cat` is
// selected (et_plan_profile_select); the events come from the plan's pool.
// Profiling scope: events recorded on the stream before and after the scope's
// launch; for a projection, a second pair is offered to the launch, and a
// row-lane projector takes it and records its own dispatch start and end (the
// kernel's duration without the stream's gaps), which is then used instead.
struct ProfScope {
    et_plan *p;
    int cat;
    double bytes;
    hipEvent_t a = nullptr, b = nullptr;
    hipEvent_t ka = nullptr, kb = nullptr;  // offered to the launch
    ProfScope(et_plan *p_, int cat_, double bytes_) : p(p_), cat(cat_), bytes(bytes_)
    {
        if (!p || !((p->prof >> cat) & 1u)) return;
        if (p->prof_seen[cat]++ % p->prof_every != 0) return;
        a = pool_event(p);
        b = pool_event(p);
        if (!a || !b) {
            if (a) p->ev_pool.push_back(a);
            if (b) p->ev_pool.push_back(b);
            a = b = nullptr;
            return;
        }
        if (cat == ET_PROF_PROJECT) {
            ka = pool_event(p);
            kb = pool_event(p);
            if (ka && kb) {
                et_offer_launch_events(ka, kb);
            } else {
                if (ka) p->ev_pool.push_back(ka);
                if (kb) p->ev_pool.push_back(kb);
                ka = kb = nullptr;
            }
        }
        (void)hipEventRecord(a, p->st);
    }
    ~ProfScope()
    {
        if (!a) return;
        const bool kernel_timed = ka && et_withdraw_launch_events();
        if (kernel_timed) {
            p->events.push_back({cat, bytes, ka, kb});
            p->ev_pool.push_back(a);
            p->ev_pool.push_back(b);
            return;
        }
        if (ka) {
            p->ev_pool.push_back(ka);
            p->ev_pool.push_back(kb);
        }
        (void)hipEventRecord(b, p->st);
        p->events.push_back({cat, bytes, a, b});
    }
};

void prof_collect(et_plan *p)
{
    if (p->events.empty()) return;
    (void)hipStreamSynchronize(p->st);
    for (auto &e : p->events) {
        float ms = 0.f;
        if (hipEventElapsedTime(&ms, e.a, e.b) == hipSuccess) {
            p->prof_ms[e.cat] += ms;
            p->prof_bytes[e.cat] += e.bytes;
            p->prof_n[e.cat] += 1;
        }
        p->ev_pool.push_back(e.a);
        p->ev_pool.push_back(e.b);
    }
    p->events.clear();
}

// ---- sweep helpers ------------------------------------------------------------

EtSweepArgs base_args(et_plan *p)
{
    EtSweepArgs a;
    std::memset(&a, 0, sizeof(a));
    a.nx = p->nx;
    a.ny = p->ny;
    a.nz = p->nz;
    a.yp = p->yp;
    a.xp = p->xp;
    a.gy = p->gy;
    a.gmu = p->gmu;
    a.yu = p->yu;
    a.ustride = (int)p->ustride;
    a.slabs = p->slabs;
    a.K = p->K;
    a.psf = p->K > 0 ? p->d_sep.p : nullptr;
    a.bg_act = p->bg;
    a.bg_mu = p->bga;
    a.bg_line = p->bg;
    a.eps = 0.f;
    a.D = 1;
    a.zlen = p->nz;
    return a;
}

// Depth chunks so that a launch over `views` views fills the device
// (about 2 waves per SIMD of 256 CUs), each chunk at least 16 planes deep:
// more, shorter chunks cost more in per-block set-up than they gain.
int choose_D(et_plan *p, int views)
{
    if (p->tune_D > 0) return std::min(p->tune_D, p->nz);
    const int waves_per_block = nthreads_for(p->nx) / 64;
    const int ts = et_sweep_slab_rows(ET_MODE_PROJECT, p->R);
    const long slabs = (p->ny + ts - 1) / ts;
    const long target = 2L * 1024;
    const long waves = slabs * views * waves_per_block;
    int D = 1;
    while (waves * D < target && p->nz / (D * 2) >= 16) D *= 2;
    return D;
}

int view_batch(et_plan *p)
{
    if (p->tune_vb > 0) return std::min(p->tune_vb, p->V);
    const size_t bytes = p->ustride * sizeof(float);
    long vb = (long)((256ull << 20) / std::max<size_t>(bytes, 1));
    return (int)std::max(1L, std::min<long>(vb, p->V));
}

// Views per launch of the gradient sweeps.  They run whole rays (one depth
// chunk: the reverse sum needs the ray's total), so a launch has only
// slabs x views blocks: at least two per CU are taken, with rotated-frame
// volumes of up to 4 GiB (256^3: 32 views, 2.2 GB), where view_batch's 256 MiB
// would give 3 views (48 blocks).
int grad_view_batch(et_plan *p)
{
    if (p->tune_vb > 0) return std::min(p->tune_vb, p->V);
    int dev = 0, cus = 0;
    if (hipGetDevice(&dev) != hipSuccess ||
        hipDeviceGetAttribute(&cus, hipDeviceAttributeMultiprocessorCount, dev) != hipSuccess || cus <= 0)
        cus = 256;
    const long want = (2L * cus + p->slabs - 1) / std::max(p->slabs, 1);
    const long cap = (long)((4ull << 30) / std::max<size_t>(p->ustride * sizeof(float), 1));
    const long vb = std::max<long>(view_batch(p), std::min(want, cap));
    return (int)std::max(1L, std::min<long>(vb, p->V));
}

void pack(et_plan *p, const float *abi, DevBuf &yf)
{
    const bool is_mu = &yf == &p->mu_yf;  // attenuation: own guard rows, zero slack after the last column
    yf.ensure(p->yf_elems + (is_mu ? ET_YF_SLACK : 0), true, p->st);
    ProfScope ps(p, ET_PROF_OTHER, 8.0 * p->nvox);
    const int guard = is_mu ? p->gmu : p->gy;
    check_launch(et_launch_pack_yf(abi, yf.p, p->nx, p->ny, p->nz, p->yp, guard, (et_stream_t)p->st));
}

// Attenuation at the start of every depth chunk (D > 1) and along whole rays
// for `nv` views whose records start at `recs`; results in p->astart / p->atot.
void mu_prefix(et_plan *p, const float *recs, int nv, int D, const float *mu_yf, bool want_start, bool want_tot,
               float *astart_out = nullptr, float *atot_out = nullptr)
{
    EtSweepArgs a = base_args(p);
    a.views = recs;
    a.V = nv;
    a.D = D;
    a.zlen = (p->nz + D - 1) / D;
    a.mu = mu_yf;
    p->csum.ensure((size_t)nv * D * p->P, false);
    p->csum_lo.ensure((size_t)nv * D * p->P, false);
    a.part = p->csum.p;
    a.part_lo = p->csum_lo.p;
    {
        ProfScope ps(p, ET_PROF_MUSUM, (double)nv * (4.0 * p->nvox + 8.0 * D * p->P));
        check_launch(et_launch_sweep(ET_MODE_MUSUM, 0, 1, 0, a, (et_stream_t)p->st));
    }
    float *as = nullptr, *at = nullptr, *atl = nullptr;
    if (want_start) as = astart_out ? astart_out : p->astart.ensure((size_t)nv * D * p->P, false);
    if (want_tot) {
        at = atot_out ? atot_out : p->atot.ensure((size_t)nv * p->P, false);
        atl = p->atot_lo.ensure((size_t)nv * p->P, false);
    }
    ProfScope ps(p, ET_PROF_OTHER, 0);
    check_launch(et_launch_chunk_scan(p->csum.p, p->csum_lo.p, as, at, atl, nv, D, (int)p->P, (et_stream_t)p->st));
}

// Floats of the attenuation-factor buffer of nv views ([nv][nz][yu][roundup(nx, 64)]).
size_t att_elems(const et_plan *p, int nv)
{
    return (size_t)nv * p->nz * p->yu * (size_t)nthreads_for(p->nx);
}

// Arguments of a projection sweep of nv views into `part` ([D][nv][ny][nx]).
EtSweepArgs project_args(et_plan *p, const float *recs, int nv, const float *act_yf, const float *mu_yf, float *part,
                         int truncate, int D, const float *astart, float *att = nullptr)
{
    EtSweepArgs a = base_args(p);
    a.views = recs;
    a.V = nv;
    a.D = D;
    a.zlen = (p->nz + D - 1) / D;
    a.act = act_yf;
    a.mu = mu_yf;
    a.astart = astart;
    a.part = part;
    a.truncate = truncate;
    a.att = mu_yf ? att : nullptr;
    return a;
}

// Fast-path projection of nv views (records recs) into sino [nv][ny][nx].
// att (with mu): the attenuation factors of the views, stored there by this
// projection (att_read false) or read from there instead of resampling mu.
void project_fast(et_plan *p, const float *recs, int nv, const float *act_yf, const float *mu_yf, float *sino,
                  int truncate, int D, const float *astart, float *parts_out = nullptr, float *att = nullptr,
                  bool att_read = false)
{
    float *part = parts_out;
    if (!part) part = (D == 1) ? sino : p->part.ensure((size_t)D * nv * p->P, false);
    const EtSweepArgs a =
        project_args(p, recs, nv, act_yf, mu_yf, part, (D == 1 && !parts_out) ? truncate : 0, D, astart, att);
    {
        const double bytes = (double)nv * ((mu_yf ? 8.0 : 4.0) * p->nvox + 4.0 * p->P);
        ProfScope ps(p, ET_PROF_PROJECT, bytes);
        const int hm = mu_yf ? (a.att ? (att_read ? 3 : 2) : 1) : 0;
        check_launch(et_launch_sweep(ET_MODE_PROJECT, p->R, hm, 1, a, (et_stream_t)p->st));
    }
    if (D > 1 && !parts_out) {
        ProfScope ps(p, ET_PROF_OTHER, 0);
        check_launch(et_launch_reduce_parts(part, D, nv, (int)p->P, sino, truncate, (et_stream_t)p->st));
    }
}

// Arguments of a rotated-frame backprojection sweep (see bpu_fast).
EtSweepArgs bpu_args(et_plan *p, const float *recs, int nv, const float *mu_yf, int src, const float *sino,
                     const float *meas, const float *rpart, int rD, float eps, int D, const float *astart, bool flip,
                     float *u_out, float *att = nullptr)
{
    EtSweepArgs a = base_args(p);
    a.views = recs;
    a.V = nv;
    a.D = D;
    a.zlen = (p->nz + D - 1) / D;
    a.mu = mu_yf;
    a.astart = astart;
    a.src = src;
    a.sino = sino;
    a.meas = meas;
    a.rpart = rpart;
    a.rD = rD;
    a.eps = eps;
    a.u = u_out;
    a.att = mu_yf ? att : nullptr;
    if (p->K > 0) a.psf = flip ? p->d_sep_flip.p : p->d_sep.p;
    return a;
}

// u for nv views: PSF(s e^{-A}) (flip: correlation kernels).
// att: attenuation factors of these views (the mu resampling is skipped), or null.
void bpu_fast(et_plan *p, const float *recs, int nv, const float *mu_yf, int src, const float *sino,
              const float *meas, const float *rpart, int rD, float eps, int D, const float *astart, bool flip,
              float *u_out, float *att = nullptr)
{
    const EtSweepArgs a =
        bpu_args(p, recs, nv, mu_yf, src, sino, meas, rpart, rD, eps, D, astart, flip, u_out, att);
    const double bytes = (double)nv * ((mu_yf ? 8.0 : 4.0) * p->nvox + 4.0 * p->P);
    ProfScope ps(p, ET_PROF_BPU, bytes);
    check_launch(et_launch_sweep(ET_MODE_BPU, p->R, mu_yf ? (a.att ? 3 : 1) : 0, 0, a, (et_stream_t)p->st));
}

EtPullArgs pull_args(et_plan *p, const float *recs, int nv, const float *u, bool adjoint, int epi, bool accumulate,
                     bool truncate, float *out, float *x, const float *sens)
{
    EtPullArgs a;
    std::memset(&a, 0, sizeof(a));
    a.nx = p->nx;
    a.ny = p->ny;
    a.nz = p->nz;
    a.yu = p->yu;
    a.xp = p->xp;
    a.ustride = (int)p->ustride;
    a.V = nv;
    a.views = recs;
    a.u = u;
    a.bg = p->bg;
    a.adjoint = adjoint;
    a.epilogue = epi;
    a.accumulate = accumulate;
    a.truncate = truncate;
    a.out = out;
    a.yp = p->yp;
    a.gy = p->gy;
    a.x = x;
    a.sens = sens;
    return a;
}

void pull_fast(et_plan *p, const float *recs, int nv, const float *u, bool adjoint, int epi, bool accumulate,
               bool truncate, float *out, float *x = nullptr, const float *sens = nullptr)
{
    const EtPullArgs a = pull_args(p, recs, nv, u, adjoint, epi, accumulate, truncate, out, x, sens);
    const double bytes = (double)nv * 4.0 * p->nvox + (accumulate ? 8.0 : 4.0) * p->nvox;
    ProfScope ps(p, ET_PROF_PULL, bytes);
    check_launch(et_launch_pull(a, (et_stream_t)p->st));
}

// ---- general path: one view at a time ------------------------------------------

// Rotated copy of an ABI volume for view v in the YF layout (+ PSF when the
// radius is too large for the sweep kernels).
void rotate_general(et_plan *p, const float *abi, int v, float bg, DevBuf &dst, bool psf, bool flip)
{
    dst.ensure(p->yf_elems, true, p->st);
    {
        ProfScope ps(p, ET_PROF_OTHER, 8.0 * p->nvox);
        check_launch(et_launch_resample_general(abi, p->nx, p->ny, p->nz, p->d_recs.p + (size_t)v * ET_VIEW_REC, bg,
                                                dst.p, 1, p->yp, p->gy, (et_stream_t)p->st));
    }
    if (psf && p->K > 0 && p->R > ET_RMAX) {
        // row (x) pass then column (y) pass, emission-lib/_et_convolveSeparable2D.cpp:43-75
        p->tmp.ensure(p->yf_elems, true, p->st);
        const long sx = p->yp, sy = 1, sz = (long)p->yp * p->xp, off = p->gy + (long)p->yp * (1 + p->xp);
        const float *sep = flip ? p->d_sep_flip.p : p->d_sep.p;
        check_launch(et_launch_conv_axis(dst.p, p->tmp.p, p->nx, p->ny, p->nz, sx, sy, sz, off, 0, sep, p->K, 0,
                                         (et_stream_t)p->st));
        check_launch(et_launch_conv_axis(p->tmp.p, dst.p, p->nx, p->ny, p->nz, sx, sy, sz, off, 1, sep, p->K, 0,
                                         (et_stream_t)p->st));
    }
}

// Sweep radius used on the general path (0 when the PSF is applied separately).
int general_R(et_plan *p) { return p->R > ET_RMAX ? 0 : p->R; }

void conv_u_general(et_plan *p, float *u, bool flip)
{
    if (!(p->K > 0 && p->R > ET_RMAX)) return;
    p->tmp2.ensure(p->ustride, true, p->st);
    const long sx = p->yu, sy = 1, sz = (long)p->yu * p->xp, off = (long)p->yu * (1 + p->xp);
    const float *sep = flip ? p->d_sep_flip.p : p->d_sep.p;
    if (!flip) {
        check_launch(et_launch_conv_axis(u, p->tmp2.p, p->nx, p->ny, p->nz, sx, sy, sz, off, 0, sep, p->K, 0,
                                         (et_stream_t)p->st));
        check_launch(et_launch_conv_axis(p->tmp2.p, u, p->nx, p->ny, p->nz, sx, sy, sz, off, 1, sep, p->K, 0,
                                         (et_stream_t)p->st));
    } else {
        check_launch(et_launch_conv_axis(u, p->tmp2.p, p->nx, p->ny, p->nz, sx, sy, sz, off, 1, sep, p->K, 0,
                                         (et_stream_t)p->st));
        check_launch(et_launch_conv_axis(p->tmp2.p, u, p->nx, p->ny, p->nz, sx, sy, sz, off, 0, sep, p->K, 0,
                                         (et_stream_t)p->st));
    }
}

void pull_general(et_plan *p, int v, const float *u, bool adjoint, bool accumulate, bool truncate, float *out,
                  int epi)
{
    EtPullArgs a;
    std::memset(&a, 0, sizeof(a));
    a.nx = p->nx;
    a.ny = p->ny;
    a.nz = p->nz;
    a.yu = p->yu;
    a.xp = p->xp;
    a.ustride = (int)p->ustride;
    a.V = 1;
    a.views = p->d_recs.p + (size_t)v * ET_VIEW_REC;
    a.u = u;
    a.bg = p->bg;
    a.adjoint = adjoint;
    a.epilogue = epi;
    a.accumulate = accumulate;
    a.truncate = truncate;
    a.out = out;
    a.yp = p->yp;
    a.gy = p->gy;
    ProfScope ps(p, ET_PROF_PULL, 4.0 * p->nvox + (accumulate ? 8.0 : 4.0) * p->nvox);
    check_launch(et_launch_pull_general(a, (et_stream_t)p->st));
}

// ---------------------------------------------------------------------------
// Operations (device pointers, ABI layouts)
// ---------------------------------------------------------------------------

// Row-lane versions (defined with the row-lane kernels' host code); false when
// the plan does not take them.
bool yl_op_project(et_plan *p, const float *act, const float *act_yf, const float *mu, const float *mu_yf, float *sino,
                   int truncate);
bool yl_op_backproject(et_plan *p, const float *sino, const float *mu, const float *mu_yf, float *out, bool adjoint,
                       int truncate);
bool yl_op_gradient(et_plan *p, const float *sino, const float *act, const float *act_yf, const float *mu,
                    const float *mu_yf, float *out, int truncate);
bool yl_op_backward(et_plan *p, const float *sino, const float *act, const float *act_yf, const float *mu,
                    const float *mu_yf, float *out_act, float *out_mu, int truncate);

void op_project(et_plan *p, const float *act, const float *mu, float *sino, int truncate)
{
    const float *act_yf = nullptr, *mu_yf = nullptr;
    if (p->fast) {
        if (act) {
            pack(p, act, p->act_yf);
            act_yf = p->act_yf.p;
        }
        if (mu) {
            pack(p, mu, p->mu_yf);
            mu_yf = p->mu_yf.p;
        }
        if (!act) {
            mu_prefix(p, p->d_recs.p, p->V, 1, mu_yf, false, true);
            ProfScope ps(p, ET_PROF_OTHER, 0);
            check_launch(et_launch_bg_only(p->atot.p, p->atot_lo.p, p->bg, (int)(p->V * p->P), sino, truncate,
                                           (et_stream_t)p->st));
            return;
        }
        if (yl_op_project(p, act, act_yf, mu, mu_yf, sino, truncate)) return;
        const int D = mu ? choose_D(p, p->V) : std::max(1, choose_D(p, p->V));
        const float *astart = nullptr;
        if (mu && D > 1) {
            mu_prefix(p, p->d_recs.p, p->V, D, mu_yf, true, false);
            astart = p->astart.p;
        }
        project_fast(p, p->d_recs.p, p->V, act_yf, mu_yf, sino, truncate, D, astart);
        return;
    }
    // general path
    const int R = general_R(p);
    for (int v = 0; v < p->V; v++) {
        float *sv = sino + (size_t)v * p->P;
        if (mu) rotate_general(p, mu, v, p->bga, p->rot_mu, false, false);
        if (!act) {
            mu_prefix(p, p->d_ident.p, 1, 1, p->rot_mu.p, false, true);
            check_launch(et_launch_bg_only(p->atot.p, p->atot_lo.p, p->bg, (int)p->P, sv, truncate,
                                           (et_stream_t)p->st));
            continue;
        }
        rotate_general(p, act, v, p->bg, p->rot_act, true, false);
        EtSweepArgs a = base_args(p);
        a.views = p->d_ident.p;
        a.V = 1;
        a.act = p->rot_act.p;
        a.mu = mu ? p->rot_mu.p : nullptr;
        a.bg_act = 0.f;
        a.bg_mu = 0.f;
        a.part = sv;
        a.truncate = truncate;
        ProfScope ps(p, ET_PROF_PROJECT, (mu ? 8.0 : 4.0) * p->nvox + 4.0 * p->P);
        check_launch(et_launch_sweep(ET_MODE_PROJECT, R, mu != nullptr, 1, a, (et_stream_t)p->st));
    }
}

void op_backproject(et_plan *p, const float *sino, const float *mu, float *out, bool adjoint, int truncate)
{
    if (p->fast) {
        const float *mu_yf = nullptr;
        if (mu) {
            pack(p, mu, p->mu_yf);
            mu_yf = p->mu_yf.p;
        }
        if (yl_op_backproject(p, sino, mu, mu_yf, out, adjoint, truncate)) return;
        const int vb = view_batch(p);
        p->u.ensure((size_t)vb * p->ustride, true, p->st);
        for (int v0 = 0; v0 < p->V; v0 += vb) {
            const int nv = std::min(vb, p->V - v0);
            const float *recs = p->d_recs.p + (size_t)v0 * ET_VIEW_REC;
            const int D = choose_D(p, nv);
            const float *astart = nullptr;
            if (mu && D > 1) {
                mu_prefix(p, recs, nv, D, mu_yf, true, false);
                astart = p->astart.p;
            }
            bpu_fast(p, recs, nv, mu_yf, ET_SRC_SINO, sino + (size_t)v0 * p->P, nullptr, nullptr, 0, 0.f, D, astart,
                     adjoint, p->u.p);
            const bool last = v0 + nv >= p->V;
            pull_fast(p, recs, nv, p->u.p, adjoint, ET_EPI_ABI_STORE, v0 > 0, last && truncate, out);
        }
        return;
    }
    const int R = general_R(p);
    p->u.ensure(p->ustride, true, p->st);
    for (int v = 0; v < p->V; v++) {
        if (mu) rotate_general(p, mu, v, p->bga, p->rot_mu, false, false);
        EtSweepArgs a = base_args(p);
        a.views = p->d_ident.p;
        a.V = 1;
        a.mu = mu ? p->rot_mu.p : nullptr;
        a.bg_mu = 0.f;
        a.src = ET_SRC_SINO;
        a.sino = sino + (size_t)v * p->P;
        a.u = p->u.p;
        if (p->K > 0) a.psf = adjoint ? p->d_sep_flip.p : p->d_sep.p;
        {
            ProfScope ps(p, ET_PROF_BPU, (mu ? 8.0 : 4.0) * p->nvox + 4.0 * p->P);
            check_launch(et_launch_sweep(ET_MODE_BPU, R, mu != nullptr, 0, a, (et_stream_t)p->st));
        }
        conv_u_general(p, p->u.p, adjoint);
        pull_general(p, v, p->u.p, adjoint, v > 0, truncate && v == p->V - 1, out, ET_EPI_ABI_STORE);
    }
}

void op_gradient(et_plan *p, const float *sino, const float *act, const float *mu, float *out, bool exact,
                 int truncate)
{
    const int mode = exact ? ET_MODE_GRAD_MU : ET_MODE_GRAD_REF;
    if (p->fast) {
        pack(p, act, p->act_yf);
        pack(p, mu, p->mu_yf);
        if (exact && yl_op_gradient(p, sino, act, p->act_yf.p, mu, p->mu_yf.p, out, truncate)) return;
        const int vb = grad_view_batch(p);
        p->u.ensure((size_t)vb * p->ustride, true, p->st);
        for (int v0 = 0; v0 < p->V; v0 += vb) {
            const int nv = std::min(vb, p->V - v0);
            const float *recs = p->d_recs.p + (size_t)v0 * ET_VIEW_REC;
            mu_prefix(p, recs, nv, 1, p->mu_yf.p, false, true);
            EtSweepArgs a = base_args(p);
            a.views = recs;
            a.V = nv;
            a.act = p->act_yf.p;
            a.mu = p->mu_yf.p;
            a.atot = p->atot.p;
            a.atot_lo = p->atot_lo.p;
            a.src = ET_SRC_SINO;
            a.sino = sino + (size_t)v0 * p->P;
            a.u = p->u.p;
            {
                ProfScope ps(p, ET_PROF_GRAD, (double)nv * (12.0 * p->nvox + 4.0 * p->P));
                check_launch(et_launch_sweep(mode, p->R, 1, 1, a, (et_stream_t)p->st));
            }
            const bool last = v0 + nv >= p->V;
            pull_fast(p, recs, nv, p->u.p, exact, ET_EPI_ABI_STORE, v0 > 0, last && truncate, out);
        }
        return;
    }
    const int R = general_R(p);
    p->u.ensure(p->ustride, true, p->st);
    for (int v = 0; v < p->V; v++) {
        rotate_general(p, mu, v, p->bga, p->rot_mu, false, false);
        rotate_general(p, act, v, p->bg, p->rot_act, exact, false);
        mu_prefix(p, p->d_ident.p, 1, 1, p->rot_mu.p, false, true);
        EtSweepArgs a = base_args(p);
        a.views = p->d_ident.p;
        a.V = 1;
        a.act = p->rot_act.p;
        a.mu = p->rot_mu.p;
        a.bg_act = 0.f;
        a.bg_mu = 0.f;
        a.atot = p->atot.p;
        a.atot_lo = p->atot_lo.p;
        a.src = ET_SRC_SINO;
        a.sino = sino + (size_t)v * p->P;
        a.u = p->u.p;
        {
            ProfScope ps(p, ET_PROF_GRAD, 12.0 * p->nvox + 4.0 * p->P);
            check_launch(et_launch_sweep(mode, R, 1, 1, a, (et_stream_t)p->st));
        }
        if (!exact) conv_u_general(p, p->u.p, false);
        pull_general(p, v, p->u.p, exact, v > 0, truncate && v == p->V - 1, out, ET_EPI_ABI_STORE);
    }
}

// Both gradients of <sino, project(act, mu)>: the exact adjoint backprojection
// of sino (activity) and the exact attenuation gradient (mu).  Row-lane plans
// after a projection of the same pair take one pass (yl_op_backward); otherwise
// the two operations in turn.
void op_backward(et_plan *p, const float *sino, const float *act, const float *mu, float *out_act, float *out_mu,
                 int truncate)
{
    if (p->fast && p->yl_ops && p->yl) {
        pack(p, act, p->act_yf);
        pack(p, mu, p->mu_yf);
        if (yl_op_backward(p, sino, act, p->act_yf.p, mu, p->mu_yf.p, out_act, out_mu, truncate)) return;
    }
    op_backproject(p, sino, mu, out_act, true, truncate);
    op_gradient(p, sino, act, mu, out_mu, true, truncate);
}

size_t align256(size_t x) { return (x + 255) & ~(size_t)255; }

struct TableBytes {
    std::vector<char> bytes;
    template <class T>
    size_t put(const std::vector<T> &v)
    {
        const size_t off = align256(bytes.size());
        bytes.resize(off + v.size() * sizeof(T));
        std::memcpy(bytes.data() + off, v.data(), v.size() * sizeof(T));
        return off;
    }
};

// Block table of a grouped launch: {first block, gx, gy, 0} per problem, each
// range padded to a multiple of 8 blocks.
struct BlockTable {
    std::vector<int> t;
    int total = 0;
    void add(int gx, int gy)
    {
        t.insert(t.end(), {total, gx, gy, 0});
        total += ((gx * gy + 7) / 8) * 8;
    }
};

// MLEM update restricted to views [v0, v0+nv) of the plan with every volume in
// the YF layout; the reference update of _et_array_interface.cpp:1080-1131.
// sens_yf must hold the clamped sensitivity of these views.
// The projection runs in Dp depth chunks and the rotated-frame backprojection
// in Db (each with its own attenuation prefix, null without mu or for one chunk).
void em_subset_update(et_plan *p, const float *recs, int nv, float *x_yf, const float *mu_yf, const float *meas,
                      const float *sens_yf, float eps, int Dp, int Db, float *att)
{
    p->u.ensure((size_t)nv * p->ustride, true, p->st);
    float *parts = p->part.ensure((size_t)Dp * nv * p->P, false);
    float *ratio = p->ratio.ensure((size_t)nv * p->P, false);
    project_fast(p, recs, nv, x_yf, mu_yf, nullptr, 1, Dp, nullptr, parts, att, true);
    {
        ProfScope ps(p, ET_PROF_OTHER, (Dp + 2.0) * 4.0 * nv * p->P);
        check_launch(et_launch_ratio_parts(meas, parts, Dp, (size_t)nv * p->P, eps, ratio, (et_stream_t)p->st));
    }
    bpu_fast(p, recs, nv, mu_yf, ET_SRC_SINO, ratio, nullptr, nullptr, 0, 0.f, Db, nullptr, false, p->u.p, att);
    pull_fast(p, recs, nv, p->u.p, false, ET_EPI_OSEM, false, false, nullptr, x_yf, sens_yf);
}

// Sensitivity of views [recs, nv): reference backprojection of ones, truncated,
// then clamped below at eps (_et_array_interface.cpp:1112-1121), into sens_yf.
// att: the views' attenuation factors (with mu).
void em_sensitivity(et_plan *p, const float *recs, int nv, const float *mu_yf, float eps, int D, float *att,
                    float *sens_yf)
{
    p->u.ensure((size_t)nv * p->ustride, true, p->st);
    bpu_fast(p, recs, nv, mu_yf, ET_SRC_ONES, nullptr, nullptr, nullptr, 0, 0.f, D, nullptr, false, p->u.p, att);
    pull_fast(p, recs, nv, p->u.p, false, ET_EPI_YF_STORE, false, true, sens_yf);
    ProfScope ps(p, ET_PROF_OTHER, 0);
    check_launch(et_launch_yf_clamp_min(sens_yf, p->nx, p->ny, p->nz, p->yp, p->gy, eps, (et_stream_t)p->st));
}

// General-path MLEM step over all views of the plan (ABI layouts), the
// reference update of _et_array_interface.cpp:1080-1131.
void em_step_general(et_plan *p, float *x, const float *meas, const float *mu, float eps)
{
    const size_t Ps = p->P * p->V;
    float *proj = p->em_sino.ensure(Ps, false);
    op_project(p, x, mu, proj, 1);
    {
        ProfScope ps(p, ET_PROF_OTHER, 0);
        check_launch(et_launch_ratio(meas, proj, proj, Ps, eps, (et_stream_t)p->st));
    }
    float *b = p->em_vol.ensure(2 * p->nvox, false);
    float *s = b + p->nvox;
    op_backproject(p, proj, mu, b, false, 1);
    check_launch(et_launch_fill(proj, Ps, 1.f, (et_stream_t)p->st));
    op_backproject(p, proj, mu, s, false, 1);
    check_launch(et_launch_clamp_min(s, p->nvox, eps, (et_stream_t)p->st));
    check_launch(et_launch_em_update_flat(x, b, s, p->nvox, (et_stream_t)p->st));
}

// Cached attenuation factors of the OSEM views from view v0 (subset order), or
// null without attenuation.
float *osem_att(et_plan *p, int v0)
{
    if (!p->osem_mu_yf) return nullptr;
    return p->osem_E.p + (size_t)v0 * att_elems(p, 1);
}

// Depth chunking of the projection (mode ET_MODE_PROJECT) or rotated-frame
// backprojection (ET_MODE_BPU) sweep for nv views, timed on the plan's stream:
// candidates 4..32 chunks (at least 4 planes each), three launches each after
// one warm-up, in three interleaved rounds; the candidate with the fastest
// round is kept.  A backprojection candidate is timed behind the projection
// (Dp chunks) and ratio of the same views, as in the subset update.  More chunks mean more, shorter blocks
// (better balance over the CUs) at the cost of more partial sums.  Inputs are
// the plan's own work buffers; the values computed are discarded.
// Depth chunking without timing, for one-shot calls: the smallest power of two
// (up to 32, at least 4 planes per chunk) whose launch holds about four waves
// per SIMD of the device.
int launch_depth_chunks(const et_plan *p, int nv)
{
    int dev = 0, cus = 0;
    if (hipGetDevice(&dev) != hipSuccess ||
        hipDeviceGetAttribute(&cus, hipDeviceAttributeMultiprocessorCount, dev) != hipSuccess || cus <= 0)
        cus = 256;
    const long waves_per_chunk = (long)p->slabs * nv * (nthreads_for(p->nx) / 64);
    int D = 1;
    while (D < 32 && p->nz / (2 * D) >= 4 && waves_per_chunk * D < 16L * cus) D *= 2;
    return D;
}

int tune_depth_chunks(et_plan *p, int mode, const float *recs, int nv, const float *mu_yf, int Dp = 0)
{
    std::vector<int> cand;
    for (int D = 4; D <= 32; D *= 2)
        if (p->nz / D >= 4) cand.push_back(D);
    if (cand.empty()) return 1;
    hipEvent_t e0, e1;
    check_hip(hipEventCreate(&e0), niftyrec_error_kernel);
    check_hip(hipEventCreate(&e1), niftyrec_error_kernel);
    const unsigned prof = p->prof;
    p->prof = 0;
    const float *act = mu_yf ? mu_yf : p->act_yf.ensure(p->yf_elems, true, p->st);
    // candidates interleaved over rounds; each keeps its fastest round, so a
    // transient slowdown of one round does not decide the choice
    constexpr int kRounds = 3;
    std::vector<float> best_of(cand.size(), 1e30f);
    for (int round = 0; round < kRounds; round++) {
        for (size_t ci = 0; ci < cand.size(); ci++) {
            const int D = cand[ci];
            const float *as = nullptr;  // both sweeps read the factors: no attenuation prefix
            float *parts = p->part.ensure((size_t)std::max(D, Dp) * nv * p->P, false);
            p->u.ensure((size_t)nv * p->ustride, true, p->st);
            // as in em_subset_update: both sweeps read the attenuation factors
            float *att = mu_yf ? p->att.ensure(att_elems(p, nv), true, p->st) : nullptr;
            float *ratio = p->ratio.ensure((size_t)nv * p->P, false);
            auto run = [&]() {
                if (mode == ET_MODE_PROJECT) {
                    // with the ratio kernel, whose reads grow with the number of partial sums
                    project_fast(p, recs, nv, act, mu_yf, nullptr, 1, D, as, parts, att, true);
                    check_launch(et_launch_ratio_parts(parts, parts, D, (size_t)nv * p->P, 1e-6f, ratio,
                                                       (et_stream_t)p->st));
                } else {
                    // behind the chosen projection and its ratio, reading that ratio, as in
                    // the subset update (the cache state the backprojection meets there)
                    project_fast(p, recs, nv, act, mu_yf, nullptr, 1, Dp, as, parts, att, true);
                    check_launch(et_launch_ratio_parts(parts, parts, Dp, (size_t)nv * p->P, 1e-6f, ratio,
                                                       (et_stream_t)p->st));
                    bpu_fast(p, recs, nv, mu_yf, ET_SRC_SINO, ratio, nullptr, nullptr, 0, 0.f, D, as, false, p->u.p,
                             att);
                }
            };
            run();
            check_hip(hipEventRecord(e0, p->st), niftyrec_error_kernel);
            for (int r = 0; r < 3; r++) run();
            check_hip(hipEventRecord(e1, p->st), niftyrec_error_kernel);
            check_hip(hipEventSynchronize(e1), niftyrec_error_kernel);
            float ms = 0.f;
            check_hip(hipEventElapsedTime(&ms, e0, e1), niftyrec_error_kernel);
            best_of[ci] = std::min(best_of[ci], ms);
        }
    }
    (void)hipEventDestroy(e0);
    (void)hipEventDestroy(e1);
    p->prof = prof;
    return cand[std::min_element(best_of.begin(), best_of.end()) - best_of.begin()];
}

// Fast-path OSEM in three stages: osem_prepare (subset caches built, with the
// attenuation packed, or reused, measured planes gathered in subset order, x packed
// into the YF layout), osem_subset (one subset update on the YF iterate) and
// osem_finish (x unpacked).  The subset layout stays in the plan between them.
// Key of the attenuation map for the OSEM caches: mu_key_launch queues its
// content fingerprint (copied to pinned host memory) on the plan's stream;
// mu_key_read returns it once the stream has reached that point (or mu's
// address when the plan's mu_check is off and `content` is not asked for, 0
// without attenuation).
// Queues the fingerprints of a and (when not null) b, n = nvox floats each,
// and the copy of their partial sums to pinned host memory.
void fp_launch(et_plan *p, const float *a, const float *b)
{
    p->fp_dev.ensure(2 * 2 * ET_FP_BLOCKS, false);  // (floats: two arrays of 8-byte partials)
    if (!p->fp_host && hipHostMalloc((void **)&p->fp_host, 2 * ET_FP_BLOCKS * sizeof(unsigned long long),
                                     hipHostMallocDefault) != hipSuccess) {
        p->fp_host = nullptr;
        throw EtError{niftyrec_error_alloccpu};
    }
    const int nb = et_fingerprint_blocks(p->nvox);
    unsigned long long *d = reinterpret_cast<unsigned long long *>(p->fp_dev.p);
    ProfScope ps(p, ET_PROF_OTHER, (b ? 8.0 : 4.0) * p->nvox);
    check_launch(et_launch_fingerprints(a, b, p->nvox, d, (et_stream_t)p->st));
    check_hip(hipMemcpyAsync(p->fp_host, d, (b ? 2 : 1) * nb * sizeof(unsigned long long), hipMemcpyDeviceToHost,
                             p->st),
              niftyrec_error_transfergpu);
}

// Key of array k (0: a, 1: b) of the last fp_launch, once the stream has
// reached it: the sum modulo 2^64 of its partials.
unsigned long long fp_key(const et_plan *p, int k)
{
    const int nb = et_fingerprint_blocks(p->nvox);
    unsigned long long key = 0;
    for (int i = 0; i < nb; i++) key += p->fp_host[(size_t)k * nb + i];
    return key;
}

void mu_key_launch(et_plan *p, const float *mu, bool content = false)
{
    if (!mu || !(p->mu_check || content)) return;
    fp_launch(p, mu, nullptr);
}

unsigned long long mu_key_read(const et_plan *p, const float *mu, bool content = false)
{
    if (!mu) return 0;
    if (!(p->mu_check || content)) return (unsigned long long)(uintptr_t)mu;
    return fp_key(p, 0);
}


// ---- row-lane OSEM path (et_yl.h) ------------------------------------------------

// Depth chunking of a row-lane sweep: the requested count rounded so that every
// chunk holds at least one plane.
void yl_chunks(const et_plan *p, int D, int &Deff, int &zlen)
{
    D = std::max(1, std::min(D, p->nz));
    zlen = (p->nz + D - 1) / D;
    Deff = (p->nz + zlen - 1) / zlen;
}

size_t yl_att_elems(const et_plan *p, int nv) { return (size_t)nv * p->nz * p->nx * p->yu; }
size_t yl_plane_elems(const et_plan *p) { return (size_t)p->nx * p->yu; }  // [x'][rows]

// Arguments of a row-lane sweep over nv views (records recs, factors att).
EtSweepArgs yl_sweep_args(const et_plan *p, const float *recs, int nv, int D, const float *att)
{
    EtSweepArgs a = base_args(const_cast<et_plan *>(p));
    int Deff, zlen;
    yl_chunks(p, D, Deff, zlen);
    a.views = recs;
    a.V = nv;
    a.D = Deff;
    a.zlen = zlen;
    a.att = const_cast<float *>(att);
    a.yl_nbuf = p->yl_nbuf;
    a.yl_tiles = p->yl_tiles;
    return a;
}

// Blocks of the row-lane pulls (reference and exact adjoint): 4 x 4 tiles of
// image columns in super-tiles of 8 x 8 tiles.
int yl_pull_tiles(const et_plan *p) { return (((p->nx + 3) / 4 + 7) / 8) * (((p->nz + 3) / 4 + 7) / 8) * 64; }

// Grouped launches of row-lane OSEM work over several plans; the block shapes
// and templates come from each plan.  Sweeps share a launch when they share
// (R, RPL, NR, block size); pulls when they share RPL.
struct YlTable {
    TableBytes tb;
    std::vector<YlLaunch> L;

    template <class Args, class Key, class Fill>
    void grouped(int kind, const std::vector<et_plan *> &plans, Key key, Fill fill)
    {
        std::map<std::tuple<int, int, int, int>, std::vector<int>> groups;
        for (size_t i = 0; i < plans.size(); i++) groups[key(plans[i])].push_back((int)i);
        for (auto &g : groups) {
            std::vector<Args> args;
            BlockTable bt;
            size_t lds = 0;
            double bytes = 0;
            for (int i : g.second) {
                int gx = 0, gy = 0;
                args.push_back(fill(i, gx, gy, bytes));
                bt.add(gx, gy);
                const et_plan *q = plans[i];
                // tiled projector (key NR 0, R > 0), full-width sweeps, backprojection
                const bool tiled = (kind == YL_PROJECT || kind == YL_GRAD) && std::get<2>(g.first) == 0 && q->R > 0;
                lds = std::max(lds, kind == YL_BPU ? q->yl_lds_bpu : (tiled ? q->yl_lds_tile : q->yl_lds_full));
            }
            const auto &k = g.first;
            L.push_back({kind, std::get<0>(k), std::get<1>(k), std::get<2>(k), std::get<3>(k), lds, tb.put(args),
                         tb.put(bt.t), (int)args.size(), bt.total, bytes});
        }
    }

    static std::tuple<int, int, int, int> sweep_key(const et_plan *p)
    {
        return std::make_tuple(p->R, p->yl_rpl, p->yl_nr, 16 * p->yl_w);
    }
    static std::tuple<int, int, int, int> pull_key(const et_plan *p) { return std::make_tuple(0, p->yl_rpl, 0, 256); }
    // projector launches: the tiled kernel is keyed with NR = 0
    static std::tuple<int, int, int, int> project_key(const et_plan *p)
    {
        return p->yl_tiles > 1 ? std::make_tuple(p->R, p->yl_rpl, 0, 512) : sweep_key(p);
    }

    // Projection of each plan's views [v0, v0 + nv) of its OSEM records into its
    // partial sums (depth chunks Dp of the plan).
    void project(const std::vector<et_plan *> &plans, const std::vector<int> &v0, const std::vector<int> &nv,
                 const std::vector<int> &Dp)
    {
        grouped<EtSweepArgs>(YL_PROJECT, plans, project_key, [&](int i, int &gx, int &gy, double &bytes) {
            et_plan *p = plans[i];
            bytes += (double)nv[i] * (8.0 * p->nvox + 4.0 * p->P);  // activity and mu (factors) per view, sinogram
            EtSweepArgs a = yl_sweep_args(p, p->osem_recs.p + (size_t)v0[i] * ET_VIEW_REC, nv[i], Dp[i],
                                          p->osem_E.p + (size_t)v0[i] * p->nz * yl_plane_elems(p));
            a.act = p->x_yf.p;
            a.part = p->part.p;
            gx = a.D * p->yl_tiles;
            gy = nv[i];
            return a;
        });
    }

    void ratio(const std::vector<et_plan *> &plans, const std::vector<int> &v0, const std::vector<int> &nv,
               const std::vector<int> &Dp)
    {
        std::vector<EtRatioArgs> ra;
        BlockTable bt;
        double bytes = 0;
        for (size_t i = 0; i < plans.size(); i++) {
            et_plan *p = plans[i];
            int Deff, zlen;
            yl_chunks(p, Dp[i], Deff, zlen);
            const long long VP = (long long)nv[i] * (long long)yl_plane_elems(p);
            ra.push_back({p->osem_meas.p + (size_t)v0[i] * yl_plane_elems(p), p->part.p, Deff, VP, p->osem_cur_eps,
                          p->ratio.p});
            bt.add((int)((VP + 255) / 256), 1);
            bytes += (Deff + 2.0) * 4.0 * nv[i] * p->P;
        }
        L.push_back({YL_RATIO, 0, 0, 0, 256, 0, tb.put(ra), tb.put(bt.t), (int)ra.size(), bt.total, bytes});
    }

    // Rotated-frame backprojection of the ratio (ones = false) or of ones.
    void bpu(const std::vector<et_plan *> &plans, const std::vector<int> &v0, const std::vector<int> &nv,
             const std::vector<int> &Db, bool ones)
    {
        grouped<EtSweepArgs>(YL_BPU, plans, sweep_key, [&](int i, int &gx, int &gy, double &bytes) {
            et_plan *p = plans[i];
            bytes += (double)nv[i] * (8.0 * p->nvox + 4.0 * p->P);  // factors read, u write, sinogram read
            EtSweepArgs a = yl_sweep_args(p, p->osem_recs.p + (size_t)v0[i] * ET_VIEW_REC, nv[i], Db[i],
                                          p->osem_E.p + (size_t)v0[i] * p->nz * yl_plane_elems(p));
            a.src = ones ? ET_SRC_ONES : ET_SRC_SINO;
            a.sino = p->ratio.p;
            a.u = p->u.p;
            gx = a.D;
            gy = nv[i];
            return a;
        });
    }

    // Pull of the rotated-frame volumes with the OSEM update (sens_out null) or
    // the truncated store of the backprojection into sens_out[i].
    void pull(const std::vector<et_plan *> &plans, const std::vector<int> &v0, const std::vector<int> &nv,
              const std::vector<const float *> &sens, const std::vector<float *> *sens_out)
    {
        grouped<EtPullArgs>(sens_out ? YL_PULL_STORE : YL_PULL_OSEM, plans, pull_key,
                            [&](int i, int &gx, int &gy, double &bytes) {
            et_plan *p = plans[i];
            bytes += (double)nv[i] * 4.0 * p->nvox + (sens_out ? 8.0 : 16.0) * p->nvox;
            const float *r = p->osem_recs.p + (size_t)v0[i] * ET_VIEW_REC;
            EtPullArgs a = sens_out ? pull_args(p, r, nv[i], p->u.p, false, ET_EPI_YF_STORE, false, true,
                                                (*sens_out)[i], nullptr, nullptr)
                                    : pull_args(p, r, nv[i], p->u.p, false, ET_EPI_OSEM, false, false, nullptr,
                                                p->x_yf.p, sens[i]);
            gx = yl_pull_tiles(p);
            gy = 1;
            return a;
        });
    }
};

// Issues launches [from, to) of a table uploaded to `dev`.
void yl_issue(const float *dev, const std::vector<YlLaunch> &L, size_t from, size_t to, hipStream_t st,
              et_plan *prof = nullptr)
{
    const char *base = reinterpret_cast<const char *>(dev);
    for (size_t k = from; k < to; k++) {
        const YlLaunch &l = L[k];
        const void *args = base + l.args_off, *tab = base + l.tab_off;
        const et_stream_t s = (et_stream_t)st;
        switch (l.kind) {
        case YL_PROJECT:
        case YL_BPU:
        case YL_GRAD: {
            ProfScope ps(prof, l.kind == YL_PROJECT ? ET_PROF_PROJECT : (l.kind == YL_BPU ? ET_PROF_BPU : ET_PROF_GRAD),
                         l.bytes);
            const int k = l.kind == YL_BPU ? 1 : (l.kind == YL_GRAD ? (l.NR == 0 ? 4 : 3) : (l.NR == 0 ? 2 : 0));
            check_launch(et_launch_yl_sweep(k, l.R, l.RPL, l.NR, l.nthr, l.lds, static_cast<const EtSweepArgs *>(args),
                                            tab, l.n, l.blocks, s));
            break;
        }
        case YL_RATIO: {
            ProfScope ps(prof, ET_PROF_OTHER, l.bytes);
            check_launch(et_launch_ratio_group(static_cast<const EtRatioArgs *>(args), tab, l.n, l.blocks, s));
            break;
        }
        case YL_PULL_ADJ:
        case YL_PULL_ADJ2: {
            ProfScope ps(prof, ET_PROF_PULL, l.bytes);
            check_launch(et_launch_yl_pull_adj(l.RPL, l.kind == YL_PULL_ADJ2 ? 2 : 1,
                                               static_cast<const EtPullArgs *>(args), tab, l.n, l.blocks, s));
            break;
        }
        default: {
            ProfScope ps(prof, ET_PROF_PULL, l.bytes);
            check_launch(et_launch_yl_pull(l.kind == YL_PULL_OSEM ? ET_EPI_OSEM : ET_EPI_YF_STORE, l.RPL,
                                           static_cast<const EtPullArgs *>(args), tab, l.n, l.blocks, s));
        }
        }
    }
}

// Uploads a table into buf (synchronously with respect to the stream).
void yl_upload(DevBuf &buf, const YlTable &t, hipStream_t st)
{
    buf.ensure((align256(t.tb.bytes.size()) + 3) / 4, false);
    check_hip(hipMemcpyAsync(buf.p, t.tb.bytes.data(), t.tb.bytes.size(), hipMemcpyHostToDevice, st),
              niftyrec_error_transfergpu);
    check_hip(hipStreamSynchronize(st), niftyrec_error_transfergpu);
}

// ---- row-lane single operations -------------------------------------------------
// A created plan on the row-lane kernels with attenuation projects,
// backprojects (reference or exact adjoint) and forms the exact attenuation
// gradient with them: the factors e^{-A} of every view are cached (on the
// attenuation map's key, as the OSEM caches), the activity is gathered by the
// projector, and the rotated-frame volumes go to the sweep kernels' pulls
// (same u layout).

// Content fingerprints of the map mu and (when given) the activity act, with
// one stream synchronisation.  The single operations key their caches on
// content whatever the plan's mu_check (which keys the OSEM caches): they may
// see other maps between OSEM calls, and a map rewritten in place is caught.
void content_keys(et_plan *p, const float *mu, const float *act, unsigned long long &kmu, unsigned long long &kact)
{
    kmu = kact = 0;
    if (!mu && !act) return;
    fp_launch(p, mu ? mu : act, mu ? act : nullptr);  // one launch for both
    check_hip(hipStreamSynchronize(p->st), niftyrec_error_kernel);
    if (mu) kmu = fp_key(p, 0);
    if (act) kact = fp_key(p, mu ? 1 : 0);
}

// Factors of every view in view order ([v][z'][x'][rows]) for the map mu (mu_yf
// its YF copy; key its content key), rebuilt when the key changes.
const float *yl_view_factors(et_plan *p, const float *mu_yf, unsigned long long key)
{
    if (p->fac_valid && p->fac_key == key) return p->fac.p;
    p->fac_valid = false;
    p->fac.ensure(yl_att_elems(p, p->V), false);
    EtSweepArgs a = yl_sweep_args(p, p->d_recs.p, p->V, 1, p->fac.p);
    a.mu = mu_yf;
    {
        ProfScope ps(p, ET_PROF_OTHER, (double)p->V * 8.0 * p->nvox);
        check_launch(et_launch_yl_factors(p->yl_rpl, a, (et_stream_t)p->st));
    }
    p->fac_valid = true;
    p->fac_key = key;
    return p->fac.p;
}

// Depth chunks of a row-lane launch of `per_chunk` blocks per chunk: the fewest
// (powers of two, chunks of at least 4 planes) that give two blocks per CU.
int yl_op_chunks(const et_plan *p, long per_chunk)
{
    int dev = 0, cus = 0;
    if (hipGetDevice(&dev) != hipSuccess ||
        hipDeviceGetAttribute(&cus, hipDeviceAttributeMultiprocessorCount, dev) != hipSuccess || cus <= 0)
        cus = 256;
    int D = 1;
    while (D < 32 && p->nz / (2 * D) >= 4 && per_chunk * D < 2L * cus) D *= 2;
    return D;
}

// The plan's sinogram [V][ny][nx] in the row-lane layout [V][x'][rows].
float *yl_sino_t(et_plan *p, const float *sino)
{
    float *t = p->sino_t.ensure((size_t)p->V * yl_plane_elems(p), false);
    ProfScope ps(p, ET_PROF_OTHER, 8.0 * p->V * p->P);
    check_launch(et_launch_gather_planes_t(t, sino, nullptr, p->V, p->nx, p->ny, p->yu, (et_stream_t)p->st));
    return t;
}

// Issues a single operation's table from its slot.  A table equal to the
// slot's last upload (the same launches over the same buffers and sizes:
// repeated calls) is on the device already and is not uploaded again (no
// stream synchronisation); the comparison covers the device bytes and each
// launch's kernel choice and offsets.
void yl_run(et_plan *p, const YlTable &t, int slot)
{
    std::vector<char> sig = t.tb.bytes;
    for (const YlLaunch &l : t.L) {
        const long long f[] = {l.kind, l.R, l.RPL, l.NR, l.nthr, (long long)l.lds, (long long)l.args_off,
                               (long long)l.tab_off, l.n, l.blocks};
        const char *c = reinterpret_cast<const char *>(f);
        sig.insert(sig.end(), c, c + sizeof(f));
    }
    DevBuf &buf = p->op_tables[slot];
    std::vector<char> &last = p->op_tables_h[slot];
    if (!buf.p || last != sig) {
        last.clear();  // (an upload that throws leaves the slot unmatched)
        yl_upload(buf, t, p->st);
        last.swap(sig);
    }
    yl_issue(buf.p, t.L, 0, t.L.size(), p->st, p);
}

// Adds to t the pull of nv views' rotated-frame volumes (p->u) into the YF
// accumulator acc: the reference pull or the exact adjoint; accumulate onto acc
// after the first batch, truncate at the last.
void yl_pull_into(YlTable &t, et_plan *p, const float *recs, int nv, bool adjoint, bool accumulate, bool truncate,
                  float *acc)
{
    t.grouped<EtPullArgs>(adjoint ? YL_PULL_ADJ : YL_PULL_STORE, {p}, YlTable::pull_key,
                          [&](int, int &gx, int &gy, double &bytes) {
                              bytes += (double)nv * 4.0 * p->nvox + (accumulate ? 8.0 : 4.0) * p->nvox;
                              gx = yl_pull_tiles(p);
                              gy = 1;
                              return pull_args(p, recs, nv, p->u.p, adjoint, ET_EPI_YF_STORE, accumulate, truncate,
                                               acc, nullptr, nullptr);
                          });
}

bool yl_op_project(et_plan *p, const float *act, const float *act_yf, const float *mu, const float *mu_yf, float *sino,
                   int truncate)
{
    if (!p->yl_ops || !p->yl || !act_yf || !mu) return false;
    unsigned long long kmu, kact;
    content_keys(p, mu, act, kmu, kact);
    const float *fac = yl_view_factors(p, mu_yf, kmu);
    const int D = yl_op_chunks(p, (long)p->V * p->yl_tiles);
    int Deff, zlen;
    yl_chunks(p, D, Deff, zlen);
    p->fwd_valid = false;
    float *part = p->fwd_part.ensure((size_t)Deff * p->V * yl_plane_elems(p), false);
    YlTable t;
    t.grouped<EtSweepArgs>(YL_PROJECT, {p}, YlTable::project_key, [&](int, int &gx, int &gy, double &bytes) {
        bytes += (double)p->V * (8.0 * p->nvox + 4.0 * p->P);
        EtSweepArgs a = yl_sweep_args(p, p->d_recs.p, p->V, D, fac);
        a.act = const_cast<float *>(act_yf);
        a.part = part;
        gx = a.D * p->yl_tiles;
        gy = p->V;
        return a;
    });
    yl_run(p, t, YL_OP_PROJECT);
    ProfScope ps(p, ET_PROF_OTHER, (Deff + 1.0) * 4.0 * p->V * p->P);
    check_launch(
        et_launch_yl_parts_to_sino(part, Deff, p->V, p->nx, p->ny, p->yu, truncate, sino, (et_stream_t)p->st));
    p->fwd_valid = true;
    p->fwd_act = kact;
    p->fwd_mu = kmu;
    p->fwd_D = D;  // the chunking as requested: yl_chunks maps it to the same chunks again
    p->fwd_tiles = p->yl_tiles;
    return true;
}

bool yl_op_backproject(et_plan *p, const float *sino, const float *mu, const float *mu_yf, float *out, bool adjoint,
                       int truncate)
{
    if (!p->yl_ops || !p->yl || !mu) return false;
    unsigned long long kmu, kact;
    content_keys(p, mu, nullptr, kmu, kact);
    const float *fac = yl_view_factors(p, mu_yf, kmu);
    const size_t pe = yl_plane_elems(p);
    const float *st = yl_sino_t(p, sino);
    const int vb = view_batch(p);
    p->u.ensure((size_t)vb * p->ustride, true, p->st);
    float *acc = p->op_acc.ensure(p->yf_elems, true, p->st);
    YlTable t;  // every batch in one table: the batches reuse u in stream order
    for (int v0 = 0; v0 < p->V; v0 += vb) {
        const int nv = std::min(vb, p->V - v0);
        const int D = yl_op_chunks(p, nv);
        const float *recs = p->d_recs.p + (size_t)v0 * ET_VIEW_REC;
        t.grouped<EtSweepArgs>(YL_BPU, {p}, YlTable::sweep_key, [&](int, int &gx, int &gy, double &bytes) {
            bytes += (double)nv * (8.0 * p->nvox + 4.0 * p->P);
            EtSweepArgs a = yl_sweep_args(p, recs, nv, D, fac + (size_t)v0 * p->nz * pe);
            a.src = ET_SRC_SINO;
            a.sino = st + (size_t)v0 * pe;
            a.u = p->u.p;
            if (p->K > 0 && adjoint) a.psf = p->d_sep_flip.p;  // correlation: the adjoint of the PSF
            gx = a.D;
            gy = nv;
            return a;
        });
        yl_pull_into(t, p, recs, nv, adjoint, v0 > 0, v0 + nv >= p->V && truncate, acc);
    }
    yl_run(p, t, YL_OP_BACKPROJECT);
    ProfScope ps(p, ET_PROF_OTHER, 8.0 * p->nvox);
    check_launch(et_launch_unpack_yf(acc, out, p->nx, p->ny, p->nz, p->yp, p->gy, (et_stream_t)p->st));
    return true;
}

// Exact attenuation gradient: per view batch, the projection's partial sums
// (the plan's projector kind, the gradient sweep's chunking), the gradient
// sweep of the same kind (tiled or full width) and the exact adjoint pull.
bool yl_op_gradient(et_plan *p, const float *sino, const float *act, const float *act_yf, const float *mu,
                    const float *mu_yf, float *out, int truncate)
{
    if (!p->yl_ops || !p->yl) return false;
    unsigned long long kmu, kact;
    content_keys(p, mu, act, kmu, kact);
    const float *fac = yl_view_factors(p, mu_yf, kmu);
    const size_t pe = yl_plane_elems(p);
    const float *st = yl_sino_t(p, sino);
    const int tiles = p->yl_tiles;
    // the last projection's partial sums, when it projected this activity
    // through this map with this projector: all views in one batch at its
    // chunking, no projection pass
    const bool reuse = p->fwd_valid && p->fwd_act == kact && p->fwd_mu == kmu && p->fwd_tiles == tiles;
    const int vb = reuse ? p->V : view_batch(p);
    p->u.ensure((size_t)vb * p->ustride, true, p->st);
    float *acc = p->op_acc.ensure(p->yf_elems, true, p->st);
    // every batch in one table (the batches reuse part and u in stream order);
    // the partial sums take the most chunks any batch uses
    float *part = nullptr;
    if (reuse) {
        part = p->fwd_part.p;
    } else {
        int Dmax = 1;
        for (int v0 = 0; v0 < p->V; v0 += vb) {
            int Deff, zlen;
            yl_chunks(p, yl_op_chunks(p, (long)std::min(vb, p->V - v0) * tiles), Deff, zlen);
            Dmax = std::max(Dmax, Deff);
        }
        part = p->part.ensure((size_t)Dmax * vb * pe, false);
    }
    YlTable t;
    for (int v0 = 0; v0 < p->V; v0 += vb) {
        const int nv = std::min(vb, p->V - v0);
        const int D = reuse ? p->fwd_D : yl_op_chunks(p, (long)nv * tiles);
        const float *recs = p->d_recs.p + (size_t)v0 * ET_VIEW_REC;
        auto args = [&](int &gx, int &gy, double &bytes) {
            bytes += (double)nv * (8.0 * p->nvox + 4.0 * p->P);
            EtSweepArgs a = yl_sweep_args(p, recs, nv, D, fac + (size_t)v0 * p->nz * pe);
            a.act = const_cast<float *>(act_yf);
            a.part = part;
            a.sino = st + (size_t)v0 * pe;
            a.u = p->u.p;
            gx = a.D * tiles;
            gy = nv;
            return a;
        };
        // the tiled kernels where the plan projects with them (key NR 0)
        if (!reuse)
            t.grouped<EtSweepArgs>(YL_PROJECT, {p}, YlTable::project_key,
                                   [&](int, int &gx, int &gy, double &bytes) { return args(gx, gy, bytes); });
        t.grouped<EtSweepArgs>(YL_GRAD, {p}, YlTable::project_key,
                               [&](int, int &gx, int &gy, double &bytes) { return args(gx, gy, bytes); });
        yl_pull_into(t, p, recs, nv, true, v0 > 0, v0 + nv >= p->V && truncate, acc);
    }
    yl_run(p, t, YL_OP_GRADIENT);
    ProfScope ps(p, ET_PROF_OTHER, 8.0 * p->nvox);
    check_launch(et_launch_unpack_yf(acc, out, p->nx, p->ny, p->nz, p->yp, p->gy, (et_stream_t)p->st));
    return true;
}

// Both gradients in one pass over the views, after a projection of the same
// activity through the same map (the forward partial sums are reused, as
// yl_op_gradient does): the adjoint backprojection sweep writes its rotated-frame
// volumes into the first half of u, the gradient sweep into the second, and one
// exact adjoint pull (YL_PULL_ADJ2), whose candidate weights serve both, adds
// them up.  Each output equals its operation's on its own with all views in one
// batch (the view batch set to the view count).  False (nothing issued) without
// the forward sums or when the two volumes of every view would exceed 4 GiB.
bool yl_op_backward(et_plan *p, const float *sino, const float *act, const float *act_yf, const float *mu,
                    const float *mu_yf, float *out_act, float *out_mu, int truncate)
{
    if (!p->yl_ops || !p->yl || !act || !mu) return false;
    const int V = p->V;
    if (2.0 * V * p->ustride * sizeof(float) > 4.0 * (1ull << 30)) return false;
    unsigned long long kmu, kact;
    content_keys(p, mu, act, kmu, kact);
    const int tiles = p->yl_tiles;
    if (!(p->fwd_valid && p->fwd_act == kact && p->fwd_mu == kmu && p->fwd_tiles == tiles)) return false;
    const float *fac = yl_view_factors(p, mu_yf, kmu);
    const float *st = yl_sino_t(p, sino);
    p->u.ensure((size_t)2 * V * p->ustride, true, p->st);
    float *ub = p->u.p, *ug = p->u.p + (size_t)V * p->ustride;
    float *acc_a = p->op_acc.ensure(p->yf_elems, true, p->st);
    float *acc_m = p->op_acc2.ensure(p->yf_elems, true, p->st);
    const float *recs = p->d_recs.p;
    YlTable t;
    const int Db = yl_op_chunks(p, V);
    t.grouped<EtSweepArgs>(YL_BPU, {p}, YlTable::sweep_key, [&](int, int &gx, int &gy, double &bytes) {
        bytes += (double)V * (8.0 * p->nvox + 4.0 * p->P);
        EtSweepArgs a = yl_sweep_args(p, recs, V, Db, fac);
        a.src = ET_SRC_SINO;
        a.sino = st;
        a.u = ub;
        if (p->K > 0) a.psf = p->d_sep_flip.p;  // correlation: the adjoint of the PSF
        gx = a.D;
        gy = V;
        return a;
    });
    t.grouped<EtSweepArgs>(YL_GRAD, {p}, YlTable::project_key, [&](int, int &gx, int &gy, double &bytes) {
        bytes += (double)V * (8.0 * p->nvox + 4.0 * p->P);
        EtSweepArgs a = yl_sweep_args(p, recs, V, p->fwd_D, fac);
        a.act = const_cast<float *>(act_yf);
        a.part = p->fwd_part.p;
        a.sino = st;
        a.u = ug;
        gx = a.D * tiles;
        gy = V;
        return a;
    });
    t.grouped<EtPullArgs>(YL_PULL_ADJ2, {p}, YlTable::pull_key, [&](int, int &gx, int &gy, double &bytes) {
        bytes += (double)V * 8.0 * p->nvox + 8.0 * p->nvox;
        gx = yl_pull_tiles(p);
        gy = 1;
        EtPullArgs a = pull_args(p, recs, V, ub, true, ET_EPI_YF_STORE, false, truncate != 0, acc_a, nullptr, nullptr);
        a.u2 = ug;
        a.out2 = acc_m;
        return a;
    });
    yl_run(p, t, YL_OP_BACKWARD);
    ProfScope ps(p, ET_PROF_OTHER, 16.0 * p->nvox);
    check_launch(et_launch_unpack_yf(acc_a, out_act, p->nx, p->ny, p->nz, p->yp, p->gy, (et_stream_t)p->st));
    check_launch(et_launch_unpack_yf(acc_m, out_mu, p->nx, p->ny, p->nz, p->yp, p->gy, (et_stream_t)p->st));
    return true;
}

// Projector kind of a row-lane plan: the tiled kernel (where it applies) or the
// full-width one.
void yl_use_tiles(et_plan *p, bool tiled)
{
    tiled = tiled && p->yl_tile_n > 1;
    p->yl_tiles = tiled ? p->yl_tile_n : 1;
    p->yl_lds = tiled ? p->yl_lds_tile : p->yl_lds_full;
}

// Depth chunking of the row-lane projection (mode ET_MODE_PROJECT, timed with
// the ratio that reads its partial sums) or backprojection (ET_MODE_BPU, timed
// behind the chosen projection and ratio) on the first subset: candidates 8,
// 16, 32 chunks and the count that puts one block per CU, three launches each
// after a warm-up, in three interleaved rounds; the fastest round decides.
// Where the tiled projector applies, the projection candidates are timed with
// both projector kinds, and the plan keeps the kind of the fastest.
int yl_tune_chunks(et_plan *p, int mode, int nv, int Dp = 0)
{
    int cus = 0, dev = 0;
    if (hipGetDevice(&dev) != hipSuccess ||
        hipDeviceGetAttribute(&cus, hipDeviceAttributeMultiprocessorCount, dev) != hipSuccess || cus <= 0)
        cus = 256;
    std::vector<int> chunks;
    for (int D : {8, 16, 32, cus / std::max(nv, 1)}) {
        int Deff, zlen;
        yl_chunks(p, D, Deff, zlen);
        if (Deff >= 1 && std::find(chunks.begin(), chunks.end(), Deff) == chunks.end()) chunks.push_back(Deff);
    }
    // (depth chunks, tiled projector); the backprojection keeps the projector chosen
    const bool both = mode == ET_MODE_PROJECT && p->yl_tile_n > 1;
    const bool tiled_now = p->yl_tiles > 1;
    std::vector<std::pair<int, bool>> cand;
    for (int D : chunks) {
        cand.push_back({D, mode == ET_MODE_PROJECT ? p->yl_tile_n > 1 : tiled_now});
        if (both) cand.push_back({D, false});
    }
    const std::vector<et_plan *> one{p};
    const std::vector<int> v0{0}, nvv{nv};
    hipEvent_t e0, e1;
    check_hip(hipEventCreate(&e0), niftyrec_error_kernel);
    check_hip(hipEventCreate(&e1), niftyrec_error_kernel);
    const unsigned prof = p->prof;
    p->prof = 0;
    std::vector<float> best(cand.size(), 1e30f);
    DevBuf tables;
    for (int round = 0; round < 3; round++) {
        for (size_t ci = 0; ci < cand.size(); ci++) {
            const int D = cand[ci].first;
            yl_use_tiles(p, cand[ci].second);
            YlTable t;
            const std::vector<int> dp{mode == ET_MODE_PROJECT ? D : Dp}, db{D};
            t.project(one, v0, nvv, dp);
            t.ratio(one, v0, nvv, dp);
            if (mode == ET_MODE_BPU) t.bpu(one, v0, nvv, db, false);
            yl_upload(tables, t, p->st);
            auto run = [&]() { yl_issue(tables.p, t.L, 0, t.L.size(), p->st); };
            run();
            check_hip(hipEventRecord(e0, p->st), niftyrec_error_kernel);
            for (int r = 0; r < 3; r++) run();
            check_hip(hipEventRecord(e1, p->st), niftyrec_error_kernel);
            check_hip(hipEventSynchronize(e1), niftyrec_error_kernel);
            float ms = 0.f;
            check_hip(hipEventElapsedTime(&ms, e0, e1), niftyrec_error_kernel);
            best[ci] = std::min(best[ci], ms);
        }
    }
    (void)hipEventDestroy(e0);
    (void)hipEventDestroy(e1);
    p->prof = prof;
    const auto &c = cand[std::min_element(best.begin(), best.end()) - best.begin()];
    yl_use_tiles(p, c.second);
    return c.first;
}

// Device buffers the row-lane OSEM launch table refers to.  A single
// operation of the plan may grow (reallocate) some of them between OSEM calls;
// the table is then rebuilt.
std::vector<const void *> yl_osem_sig(const et_plan *p)
{
    return {p->x_yf.p, p->osem_meas.p, p->part.p, p->ratio.p, p->u.p, p->sens.p, p->osem_recs.p, p->osem_E.p};
}

// Launch table of one row-lane OSEM iteration: per subset projection, ratio,
// backprojection, pull with the update.
void yl_osem_tables(et_plan *p, int S, const std::vector<int> &start)
{
    const std::vector<et_plan *> one{p};
    YlTable t;
    for (int s = 0; s < S; s++) {
        const std::vector<int> v0{start[s]}, nv{start[s + 1] - start[s]}, dp{p->osem_Dp}, db{p->osem_Db};
        const std::vector<const float *> sens{p->sens.p + (size_t)s * p->yf_elems};
        t.project(one, v0, nv, dp);
        t.ratio(one, v0, nv, dp);
        t.bpu(one, v0, nv, db, false);
        t.pull(one, v0, nv, sens, nullptr);
    }
    yl_upload(p->osem_tables, t, p->st);
    p->yl_launches = t.L;
    p->osem_sig = yl_osem_sig(p);
}

// Row-lane caches of an OSEM rebuild: attenuation factors of every view (one
// launch), chunkings, subset sensitivities, and the launch table of one
// iteration.  The measured planes and the iterate are set by the caller.
void yl_osem_build(et_plan *p, const float *mu_yf, int S, float eps, const std::vector<int> &start)
{
    const int maxsub = (p->V + S - 1) / S;
    // every buffer the launch tables refer to is allocated before they are built
    p->x_yf.ensure(p->yf_elems, true, p->st);
    p->osem_meas.ensure((size_t)p->V * yl_plane_elems(p), true, p->st);
    p->osem_E.ensure(yl_att_elems(p, p->V), false);
    {
        EtSweepArgs a = yl_sweep_args(p, p->osem_recs.p, p->V, 1, p->osem_E.p);
        a.mu = mu_yf;
        ProfScope ps(p, ET_PROF_OTHER, (double)p->V * 8.0 * p->nvox);
        check_launch(et_launch_yl_factors(p->yl_rpl, a, (et_stream_t)p->st));
    }
    p->u.ensure((size_t)maxsub * p->ustride, true, p->st);
    p->ratio.ensure((size_t)maxsub * yl_plane_elems(p), true, p->st);
    p->part.ensure((size_t)p->nz * maxsub * yl_plane_elems(p), false);
    p->osem_cur_eps = eps;
    int Dp = p->tune_D, Db = p->tune_D;
    yl_use_tiles(p, true);  // the tiled projector unless timing prefers the full-width one
    if (p->tune_D <= 0) {
        Dp = yl_tune_chunks(p, ET_MODE_PROJECT, start[1] - start[0]);
        Db = yl_tune_chunks(p, ET_MODE_BPU, start[1] - start[0], Dp);
    }
    int zlen;
    yl_chunks(p, Dp, p->osem_Dp, zlen);
    yl_chunks(p, Db, p->osem_Db, zlen);
    // sensitivities: backprojection of ones, truncated, clamped below at eps
    p->sens.ensure((size_t)S * p->yf_elems, true, p->st);
    const std::vector<et_plan *> one{p};
    for (int s = 0; s < S; s++) {
        YlTable t;
        const std::vector<int> v0{start[s]}, nv{start[s + 1] - start[s]}, db{p->osem_Db};
        std::vector<float *> out{p->sens.p + (size_t)s * p->yf_elems};
        t.bpu(one, v0, nv, db, true);
        t.pull(one, v0, nv, {}, &out);
        yl_upload(p->osem_tables, t, p->st);
        yl_issue(p->osem_tables.p, t.L, 0, t.L.size(), p->st, p);
        ProfScope ps(p, ET_PROF_OTHER, 0);
        check_launch(et_launch_yf_clamp_min(out[0], p->nx, p->ny, p->nz, p->yp, p->gy, eps, (et_stream_t)p->st));
    }
    yl_osem_tables(p, S, start);
}

void osem_prepare(et_plan *p, float *x, const float *meas, const float *mu, int S, float eps,
                  unsigned long long mu_key, bool allow_yl = true)
{
    if (S <= 0 || S > p->V || !p->fast) throw EtError{niftyrec_error_parameters};
    // the row-lane kernels need the attenuation factors (a map) and a plan whose
    // plane fits their block
    const bool use_yl = allow_yl && p->yl && mu != nullptr;
    // subset s: views s, s+S, ...; records and measured planes reordered by subset
    std::vector<int> order, start(S + 1, 0);
    for (int s = 0; s < S; s++) {
        start[s] = (int)order.size();
        for (int v = s; v < p->V; v += S) order.push_back(v);
    }
    start[S] = (int)order.size();
    const int maxsub = (p->V + S - 1) / S;
    const int D = choose_D(p, maxsub);
    const bool rebuild = !p->osem_valid || p->osem_S != S || p->osem_D != D || p->osem_has_mu != (mu != nullptr) ||
                         p->osem_mu_key != mu_key || p->osem_eps != eps || p->osem_yl != use_yl;
    // mu in the YF layout is read only to build the caches (the subset loop
    // reads the cached factors), so it is packed only for a rebuild
    if (mu && rebuild) pack(p, mu, p->mu_yf);
    const float *mu_yf = mu ? p->mu_yf.p : nullptr;
    if (rebuild) {
        std::vector<float> recs((size_t)p->V * ET_VIEW_REC);
        for (int i = 0; i < p->V; i++)
            std::memcpy(&recs[(size_t)i * ET_VIEW_REC], &p->recs[(size_t)order[i] * ET_VIEW_REC],
                        sizeof(float) * ET_VIEW_REC);
        p->osem_recs.ensure(recs.size(), false);
        check_hip(hipMemcpyAsync(p->osem_recs.p, recs.data(), sizeof(float) * recs.size(), hipMemcpyHostToDevice,
                                 p->st),
                  niftyrec_error_transfergpu);
        p->osem_order.ensure(order.size(), false);
        check_hip(hipMemcpyAsync(p->osem_order.p, order.data(), sizeof(int) * order.size(), hipMemcpyHostToDevice,
                                 p->st),
                  niftyrec_error_transfergpu);
        p->osem_valid = false;  // until the rebuild completes
        if (use_yl) {
            p->osem_mu_yf = true;
            yl_osem_build(p, mu_yf, S, eps, start);
            check_hip(hipStreamSynchronize(p->st), niftyrec_error_kernel);
            p->osem_S = S;
            p->osem_D = D;
            p->osem_has_mu = true;
            p->osem_mu_key = mu_key;
            p->osem_eps = eps;
            p->osem_yl = true;
            p->osem_valid = true;
        }
    }
    if (use_yl) {
        // measured data in subset order, transposed to [v][x][rows]
        p->osem_meas.ensure((size_t)p->V * yl_plane_elems(p), false);
        {
            ProfScope ps(p, ET_PROF_OTHER, 8.0 * p->V * p->P);
            check_launch(et_launch_gather_planes_t(p->osem_meas.p, meas, reinterpret_cast<const int *>(p->osem_order.p),
                                                   p->V, p->nx, p->ny, p->yu, (et_stream_t)p->st));
        }
        pack(p, x, p->x_yf);
        // a single operation reallocated a buffer of the cached table
        if (yl_osem_sig(p) != p->osem_sig) yl_osem_tables(p, S, start);
        p->osem_start = start;
        p->osem_cur_mu = mu_yf;
        p->osem_cur_eps = eps;
        return;
    }
    if (rebuild) {
        // chunking: fixed by et_plan_set_tuning, else timed on the first subset
        int Dp = D, Db = D;
        if (p->tune_D <= 0) {
            Dp = tune_depth_chunks(p, ET_MODE_PROJECT, p->osem_recs.p, start[1], mu_yf);
            Db = tune_depth_chunks(p, ET_MODE_BPU, p->osem_recs.p, start[1], mu_yf, Dp);
        }
        p->osem_Dp = Dp;
        p->osem_Db = Db;
        p->osem_mu_yf = mu_yf != nullptr;
        p->sens.ensure((size_t)S * p->yf_elems, true, p->st);
        // attenuation factors of every view, once per attenuation map
        if (mu_yf) p->osem_E.ensure(att_elems(p, p->V), false);
        for (int s = 0; s < S; s++) {
            const int nv = start[s + 1] - start[s];
            const float *r = p->osem_recs.p + (size_t)start[s] * ET_VIEW_REC;
            float *E = osem_att(p, start[s]);
            if (mu_yf) {
                EtSweepArgs a = base_args(p);
                a.views = r;
                a.V = nv;
                a.mu = mu_yf;
                a.att = E;
                ProfScope ps(p, ET_PROF_OTHER, (double)nv * 8.0 * p->nvox);
                check_launch(et_launch_att_factors(a, (et_stream_t)p->st));
            }
            em_sensitivity(p, r, nv, mu_yf, eps, Db, E, p->sens.p + (size_t)s * p->yf_elems);
        }
        check_hip(hipStreamSynchronize(p->st), niftyrec_error_kernel);
        p->osem_S = S;
        p->osem_D = D;
        p->osem_has_mu = mu != nullptr;
        p->osem_mu_key = mu_key;
        p->osem_eps = eps;
        p->osem_yl = false;
        p->osem_valid = true;
    }
    // measured data in subset order
    p->osem_meas.ensure((size_t)p->V * p->P, false);
    {
        ProfScope ps(p, ET_PROF_OTHER, 8.0 * p->V * p->P);
        check_launch(et_launch_gather_planes(p->osem_meas.p, meas, reinterpret_cast<const int *>(p->osem_order.p),
                                             p->V, p->P, (et_stream_t)p->st));
    }
    pack(p, x, p->x_yf);
    // work buffers of the largest subset (allocated here, outside any capture)
    p->u.ensure((size_t)maxsub * p->ustride, true, p->st);
    p->part.ensure((size_t)p->osem_Dp * maxsub * p->P, false);
    p->ratio.ensure((size_t)maxsub * p->P, false);
    p->osem_start = start;
    p->osem_cur_mu = mu_yf;
    p->osem_cur_eps = eps;
}

// Device work of subset s (views start[s] .. start[s+1]-1 in subset order).
void osem_subset(et_plan *p, int s)
{
    if (p->osem_yl) {
        yl_issue(p->osem_tables.p, p->yl_launches, 4 * (size_t)s, 4 * (size_t)s + 4, p->st, p);
        return;
    }
    const std::vector<int> &start = p->osem_start;
    const int nv = start[s + 1] - start[s];
    const float *r = p->osem_recs.p + (size_t)start[s] * ET_VIEW_REC;
    em_subset_update(p, r, nv, p->x_yf.p, p->osem_cur_mu, p->osem_meas.p + (size_t)start[s] * p->P,
                     p->sens.p + (size_t)s * p->yf_elems, p->osem_cur_eps, p->osem_Dp, p->osem_Db,
                     osem_att(p, start[s]));
}

void osem_finish(et_plan *p, float *x)
{
    ProfScope ps(p, ET_PROF_OTHER, 8.0 * p->nvox);
    check_launch(et_launch_unpack_yf(p->x_yf.p, x, p->nx, p->ny, p->nz, p->yp, p->gy, (et_stream_t)p->st));
}

void op_osem(et_plan *p, float *x, const float *meas, const float *mu, int S, int iters, float eps)
{
    if (S <= 0 || S > p->V) throw EtError{niftyrec_error_parameters};
    if (!p->fast) {
        if (S != 1) throw EtError{niftyrec_error_parameters};
        for (int it = 0; it < iters; it++) em_step_general(p, x, meas, mu, eps);
        return;
    }
    mu_key_launch(p, mu);
    if (mu && p->mu_check) check_hip(hipStreamSynchronize(p->st), niftyrec_error_kernel);
    osem_prepare(p, x, meas, mu, S, eps, mu_key_read(p, mu));
    for (int it = 0; it < iters; it++)
        for (int s = 0; s < S; s++) osem_subset(p, s);
    osem_finish(p, x);
}

// ---------------------------------------------------------------------------
// Host-pointer calls: upload, run on a private stream, download.
// ---------------------------------------------------------------------------

struct HostCall {
    hipStream_t st = nullptr;
    std::vector<float *> bufs;
    ~HostCall()
    {
        if (st) (void)hipStreamSynchronize(st);
        for (float *b : bufs) (void)hipFree(b);
        if (st) (void)hipStreamDestroy(st);
    }
    void init()
    {
        check_hip(hipStreamCreateWithFlags(&st, hipStreamNonBlocking), niftyrec_error_unspecified);
    }
    float *alloc(size_t n)
    {
        float *p = nullptr;
        if (hipMalloc((void **)&p, std::max<size_t>(n, 1) * sizeof(float)) != hipSuccess)
            throw EtError{niftyrec_error_allocgpu};
        bufs.push_back(p);
        return p;
    }
    float *upload(const float *h, size_t n)
    {
        float *d = alloc(n);
        check_hip(hipMemcpyAsync(d, h, n * sizeof(float), hipMemcpyHostToDevice, st), niftyrec_error_transfergpu);
        return d;
    }
    void download(float *h, const float *d, size_t n)
    {
        check_hip(hipMemcpyAsync(h, d, n * sizeof(float), hipMemcpyDeviceToHost, st), niftyrec_error_transfergpu);
        check_hip(hipStreamSynchronize(st), niftyrec_error_kernel);
    }
};

template <class F>
int guarded(F &&f)
{
    try {
        return f();
    } catch (const EtError &e) {
        return e.code;
    } catch (const std::bad_alloc &) {
        return niftyrec_error_alloccpu;
    } catch (...) {
        return niftyrec_error_unspecified;
    }
}

bool is_absent(const int *s) { return s == nullptr || (s[0] == 0 && s[1] == 0 && s[2] == 0); }

int project_host(float *activity, int *activity_size, float *sinogram, int *sinogram_size, float *cameras,
                 int *cameras_size, float *psf, int *psf_size, float *attenuation, int *attenuation_size,
                 float background, float background_attenuation, int truncate)
{
    // Validation, emission-lib/_et_array_interface.cpp:357-393.
    const bool no_act = is_absent(activity_size) || activity == nullptr;
    const bool no_psf = is_absent(psf_size) || psf == nullptr;
    const bool no_mu = is_absent(attenuation_size) || attenuation == nullptr;
    if (no_act && no_mu) return niftyrec_error_parameters;
    if (no_act && activity_size) {  // reference quirk: activity_size takes the attenuation size (:360)
        activity_size[0] = attenuation_size[0];
        activity_size[1] = attenuation_size[1];
        activity_size[2] = attenuation_size[2];
    }
    const int *vs = no_act ? attenuation_size : activity_size;
    const int nx = vs[0], ny = vs[1], nz = vs[2];
    const int V = cameras_size[0], cols = cameras_size[1];
    if (!(cols == 1 || cols == 3 || cols == 6)) return niftyrec_error_parameters;
    if (nx <= 0 || ny <= 0 || nz <= 0 || V <= 0) return niftyrec_error_parameters;
    if (!no_psf) {
        if (psf_size[0] % 2 != 1 || psf_size[1] % 2 != 1 || psf_size[2] != nx) return niftyrec_error_parameters;
        if (psf_size[0] != psf_size[1] || nz != nx) return niftyrec_error_parameters;
    }
    if (!no_act && !no_mu && (attenuation_size[0] != nx || attenuation_size[1] != ny || attenuation_size[2] != nz))
        return niftyrec_error_parameters;
    if (sinogram_size && (sinogram_size[0] != nx || sinogram_size[1] != ny || sinogram_size[2] != V))
        return niftyrec_error_parameters;
    if (nx > ET_MAX_NX) return niftyrec_error_parameters;
    int st = select_device();
    if (st) return st;
    return guarded([&]() -> int {
        Cameras cams;
        int rc = expand_cameras(cameras, V, cols, vs, cams);
        if (rc) return rc;
        et_plan plan;
        HostCall hc;
        hc.init();
        plan.st = hc.st;
        rc = build_plan(&plan, nx, ny, nz, cams, no_psf ? nullptr : psf, no_psf ? 0 : psf_size[0], background,
                        background_attenuation);
        if (rc) return rc;
        const size_t n = (size_t)nx * ny * nz;
        const float *d_act = no_act ? nullptr : hc.upload(activity, n);
        const float *d_mu = no_mu ? nullptr : hc.upload(attenuation, n);
        float *d_sino = hc.alloc((size_t)nx * ny * V);
        op_project(&plan, d_act, d_mu, d_sino, truncate);
        hc.download(sinogram, d_sino, (size_t)nx * ny * V);
        return niftyrec_success;
    });
}

// Fisher information of a grid of points, from an activity (its truncated
// projection first, from_projection false) or from a given projection.
// Reference emission-lib/_et_array_interface.cpp:1135-1559 (validation and
// camera expansion) and emission-lib/_et.cpp:531-746 (et_fisher_grid).
//   vol_size   the activity size (from an activity) or the grid size
//              (bkpr_size, from a projection), [nx, ny, nz]
//   input      activity [nx, ny, nz] or projection [snx, sny, V] (input_size)
//   grid       [nx, ny, nz]: 0, or the 1-based index of a grid point
//   fisher, prior  [n, n], n = fisher_size[0]; prior may be null
// The prior rule follows the reference's dispatch on the GPU flag: the CUDA
// rule (|d| <= 3 off the diagonal, _et.cpp:1633-1650) for GPU != 0, the CPU
// rule (|d| <= 1 with the diagonal, _et.cpp:703-719) for GPU == 0.
int fisher_host(bool from_projection, const float *input, const int *input_size, const int *vol_size,
                const float *cameras, const int *cameras_size, const float *psf, const int *psf_size,
                const float *grid, float *fisher, float *prior, const int *fisher_size, const float *attenuation,
                const int *attenuation_size, float background, float background_attenuation, int GPU)
{
    // Validation in the reference's order and with its status (1) where the
    // reference checks (:1146-1199, :1356-1410).
    const bool no_psf = is_absent(psf_size) || psf == nullptr;
    const bool no_mu = is_absent(attenuation_size) || attenuation == nullptr;
    const int V = cameras_size[0], cols = cameras_size[1];
    if (!(cols == 1 || cols == 3 || cols == 6)) return niftyrec_error_unspecified;
    const bool dims2 = vol_size[2] == 1;
    if (dims2) {
        if (vol_size[0] != vol_size[1]) return niftyrec_error_unspecified;
        if (!no_psf && (psf_size[0] % 2 != 1 || psf_size[1] % 2 != 1)) return niftyrec_error_unspecified;
    } else {
        if (!from_projection && vol_size[0] != vol_size[2]) return niftyrec_error_unspecified;
        if (!no_psf && (psf_size[0] % 2 != 1 || psf_size[1] % 2 != 1 || psf_size[2] != vol_size[0]))
            return niftyrec_error_unspecified;
    }
    // Beyond the reference's checks (where it would read out of bounds or
    // dereference a null PSF): a PSF is required and covers every depth of the
    // grid, the projection holds every camera, the sizes are positive.
    const int nx = vol_size[0], ny = vol_size[1], nz = vol_size[2], n = fisher_size[0];
    if (no_psf || V <= 0 || nx <= 0 || ny <= 0 || nz <= 0 || n < 0 || psf_size[2] < nz)
        return niftyrec_error_parameters;
    if (from_projection && (input_size[0] <= 0 || input_size[1] <= 0 || input_size[2] != V))
        return niftyrec_error_parameters;
    if (!from_projection) {
        // the projection runs through this library's projector: a 3-D [N, m, N]
        // volume with a square PSF and a matching attenuation map
        if (dims2 || psf_size[0] != psf_size[1] || nx > ET_MAX_NX) return niftyrec_error_parameters;
        if (!no_mu && (attenuation_size[0] != nx || attenuation_size[1] != ny || attenuation_size[2] != nz))
            return niftyrec_error_parameters;
    }
    Cameras cams;
    // centres: the activity size, or [N, m, N] from the projection's size (:1412-1414)
    const int ref_size_proj[3] = {from_projection ? input_size[0] : 0, from_projection ? input_size[1] : 0,
                                  from_projection ? input_size[0] : 0};
    int rc = expand_cameras(cameras, V, cols, from_projection ? ref_size_proj : vol_size, cams);
    if (rc) return rc;
    // grid points (:583-602), x-, then y-, then z-major scan
    std::vector<int> c0(3 * (size_t)std::max(n, 1), -1);
    for (int x = 0; x < nx; x++)
        for (int y = 0; y < ny; y++)
            for (int z = 0; z < nz; z++) {
                const float g = grid[(size_t)z * ny * nx + (size_t)y * nx + x];
                if (g == 0.f) continue;
                const int k = (int)(g - 1);
                if (k < 0 || k >= n || c0[3 * (size_t)k] != -1) return 2;  // ET_ERROR_BADGRID
                c0[3 * (size_t)k] = x;
                c0[3 * (size_t)k + 1] = y;
                c0[3 * (size_t)k + 2] = z;
            }
    if (n == 0) return niftyrec_success;
    // grid points in each camera frame: XYZ rotation about the camera's centre,
    // M (x, y, z, 1) left to right in float, truncated (:617-623)
    std::vector<int> rc3((size_t)V * n * 3);
    for (int cam = 0; cam < V; cam++) {
        float M[16];
        rotation_matrix(M, cams.ang[cam], cams.ang[V + cam], cams.ang[2 * V + cam], cams.ctr[cam],
                        cams.ctr[V + cam], cams.ctr[2 * V + cam], 0);
        for (int k = 0; k < n; k++) {
            const float p0 = (float)c0[3 * (size_t)k], p1 = (float)c0[3 * (size_t)k + 1],
                        p2 = (float)c0[3 * (size_t)k + 2];
            int *o = &rc3[((size_t)cam * n + k) * 3];
            for (int r = 0; r < 3; r++) o[r] = (int)(M[4 * r] * p0 + M[4 * r + 1] * p1 + M[4 * r + 2] * p2 + M[4 * r + 3]);
        }
    }
    int st = select_device();
    if (st) return st;
    return guarded([&]() -> int {
        HostCall hc;
        hc.init();
        const size_t npsf = (size_t)psf_size[0] * psf_size[1] * psf_size[2];
        EtFisherArgs a;
        a.n = n;
        a.V = V;
        a.px = psf_size[0];
        a.py = psf_size[1];
        a.nz = nz;
        a.prior_rule = GPU ? 3 : 1;
        a.psf = hc.upload(psf, npsf);
        a.coords = reinterpret_cast<const int *>(hc.upload(reinterpret_cast<const float *>(rc3.data()), rc3.size()));
        a.coords0 = reinterpret_cast<const int *>(hc.upload(reinterpret_cast<const float *>(c0.data()), 3 * (size_t)n));
        if (from_projection) {
            a.snx = input_size[0];
            a.sny = input_size[1];
            a.sino = hc.upload(input, (size_t)a.snx * a.sny * V);
        } else {
            // the activity's projection, truncated (_et.cpp:563)
            et_plan plan;
            plan.st = hc.st;
            int r = build_plan(&plan, nx, ny, nz, cams, psf, psf_size[0], background, background_attenuation);
            if (r) return r;
            const size_t nv = (size_t)nx * ny * nz;
            const float *d_act = hc.upload(input, nv);
            const float *d_mu = no_mu ? nullptr : hc.upload(attenuation, nv);
            float *d_sino = hc.alloc((size_t)nx * ny * V);
            op_project(&plan, d_act, d_mu, d_sino, 1);
            a.snx = nx;
            a.sny = ny;
            a.sino = d_sino;
        }
        a.fisher = hc.alloc((size_t)n * n);
        a.prior = prior ? hc.alloc((size_t)n * n) : nullptr;
        check_launch(et_launch_fisher_pairs(a, (et_stream_t)hc.st));
        hc.download(fisher, a.fisher, (size_t)n * n);
        if (prior) hc.download(prior, a.prior, (size_t)n * n);
        return niftyrec_success;
    });
}

int backproject_host(float *sino, int *sino_size, float *bkpr, int *bkpr_size, float *cameras, int *cameras_size,
                     float *psf, int *psf_size, float *attenuation, int *attenuation_size, float background,
                     float background_attenuation, int truncate, bool adjoint)
{
    // Validation, emission-lib/_et_array_interface.cpp:565-600.
    const bool no_psf = is_absent(psf_size) || psf == nullptr;
    const bool no_mu = is_absent(attenuation_size) || attenuation == nullptr;
    const int V = cameras_size[0], cols = cameras_size[1];
    if (!(cols == 1 || cols == 3 || cols == 6)) return niftyrec_error_parameters;
    if (sino_size[2] != V || V <= 0) return niftyrec_error_parameters;
    const int nx = bkpr_size[0], ny = bkpr_size[1], nz = bkpr_size[2];
    if (nx != sino_size[0] || ny != sino_size[1] || nz <= 0 || nx <= 0 || ny <= 0) return niftyrec_error_parameters;
    if (!no_psf) {
        if (psf_size[0] % 2 != 1 || psf_size[1] % 2 != 1 || psf_size[2] != sino_size[0])
            return niftyrec_error_parameters;
        if (psf_size[0] != psf_size[1] || nz != nx) return niftyrec_error_parameters;
    }
    if (!no_mu && (attenuation_size[0] != nx || attenuation_size[1] != ny || attenuation_size[2] != nz))
        return niftyrec_error_parameters;
    if (nx > ET_MAX_NX) return niftyrec_error_parameters;
    int st = select_device();
    if (st) return st;
    return guarded([&]() -> int {
        // centres from activity size [N, m, N] (_et_array_interface.cpp:598-600)
        const int act_size[3] = {sino_size[0], sino_size[1], sino_size[0]};
        Cameras cams;
        int rc = expand_cameras(cameras, V, cols, act_size, cams);
        if (rc) return rc;
        et_plan plan;
        HostCall hc;
        hc.init();
        plan.st = hc.st;
        rc = build_plan(&plan, nx, ny, nz, cams, no_psf ? nullptr : psf, no_psf ? 0 : psf_size[0], background,
                        background_attenuation);
        if (rc) return rc;
        const size_t n = (size_t)nx * ny * nz;
        const float *d_sino = hc.upload(sino, (size_t)nx * ny * V);
        const float *d_mu = no_mu ? nullptr : hc.upload(attenuation, n);
        float *d_out = hc.alloc(n);
        op_backproject(&plan, d_sino, d_mu, d_out, adjoint, truncate);
        hc.download(bkpr, d_out, n);
        return niftyrec_success;
    });
}

// One MLEM update on host arrays with cameras [3 x V] (the EM drivers).
int em_step_host(float *activity, int size_x, int size_y, const float *sino, int V, const float *cams3,
                 const float *attenuation, const float *psf, int psf_h, float bg, float bga, float eps)
{
    if (size_x > ET_MAX_NX) return niftyrec_error_parameters;
    int st = select_device();
    if (st) return st;
    return guarded([&]() -> int {
        const int nx = size_x, ny = size_y, nz = size_x;
        const int vs[3] = {nx, ny, nz};
        Cameras cams;
        int rc = expand_cameras(cams3, V, 3, vs, cams);
        if (rc) return rc;
        et_plan plan;
        HostCall hc;
        hc.init();
        plan.st = hc.st;
        rc = build_plan(&plan, nx, ny, nz, cams, psf, psf ? psf_h : 0, bg, bga);
        if (rc) return rc;
        // one update: the chunking comes from the launch size, not from timing
        plan.tune_D = launch_depth_chunks(&plan, V);
        const size_t n = (size_t)nx * ny * nz;
        float *d_x = hc.upload(activity, n);
        const float *d_meas = hc.upload(sino, (size_t)nx * ny * V);
        const float *d_mu = attenuation ? hc.upload(attenuation, n) : nullptr;
        op_osem(&plan, d_x, d_meas, d_mu, 1, 1, eps);
        hc.download(activity, d_x, n);
        return niftyrec_success;
    });
}

// Camera subset of et_array_osem_step, emission-lib/_et_array_interface.cpp:981-1009:
// n_cameras / subset_order cameras drawn with replacement from glibc rand()
// (two calls per camera, R = base + fine/scale in single precision); an index
// equal to n_cameras (the reference's clamp) is clamped to n_cameras - 1.
std::vector<int> random_subset(int n_cameras, int nsub)
{
    std::vector<int> idx(nsub);
    const float scale = (float)(RAND_MAX + 1.);
    for (int c = 0; c < nsub; c++) {
        const float base = rand() / scale;
        const float fine = rand() / scale;
        const float Rr = base + fine / scale;
        unsigned int i = (unsigned int)(n_cameras * Rr);
        if (i >= (unsigned int)n_cameras) i = n_cameras - 1;
        idx[c] = (int)i;
    }
    return idx;
}

// Cameras [3 x n] of the views idx of a [3 x V] camera array.
std::vector<float> camera_subset(const float *cams3, int V, const std::vector<int> &idx)
{
    const int n = (int)idx.size();
    std::vector<float> out((size_t)3 * n);
    for (int c = 0; c < n; c++)
        for (int axis = 0; axis < 3; axis++) out[c + axis * n] = cams3[idx[c] + axis * V];
    return out;
}

enum EmDriver { EM_MLEM = 0, EM_OSEM_ORDERED = 1, EM_OSEM_RANDOM = 2 };

// Device-resident EM drivers (the loops of emission-lib/_et_array_interface.cpp:940-960
// and :1058-1076 without the per-step uploads): the data, attenuation and
// iterate stay in device memory for every iteration, one plan's workspaces
// and caches serve every step.
//   EM_MLEM          `iterations` MLEM updates over all cameras (sensitivity once);
//   EM_OSEM_ORDERED  `iterations` passes over `param` ordered subsets (subset s:
//                    cameras s, s + S, ...), the update of et_plan_osem;
//   EM_OSEM_RANDOM   `iterations` et_array_osem_step updates, each on
//                    n_cameras / `param` cameras drawn with replacement from rand().
// The activity is reset to 1 first, as the reference drivers do.
int em_driver_host(int kind, float *activity, int size_x, int size_y, const float *sino, int V, const float *cams3,
                   const float *attenuation, const float *psf, int psf_h, float bg, float bga, float eps, int param,
                   int iterations)
{
    if (size_x <= 0 || size_y <= 0 || V <= 0 || iterations < 0) return niftyrec_error_parameters;
    if (size_x > ET_MAX_NX) return niftyrec_error_parameters;
    if (kind == EM_OSEM_ORDERED && (param <= 0 || param > V)) return niftyrec_error_parameters;
    if (kind == EM_OSEM_RANDOM && (param <= 0 || V / param <= 0)) return niftyrec_error_parameters;
    // The random subsets are drawn before any device work, so the draws follow
    // the caller's rand() state alone (the GPU runtime's start-up can call
    // rand() itself).  The reference's CPU path draws them step by step with no
    // other rand() call in between: the same sequence.
    std::vector<std::vector<int>> draws;
    if (kind == EM_OSEM_RANDOM)
        for (int it = 0; it < iterations; it++) draws.push_back(random_subset(V, V / param));
    int st = select_device();
    if (st) return st;
    return guarded([&]() -> int {
        const int nx = size_x, ny = size_y, nz = size_x;
        const int vs[3] = {nx, ny, nz};
        const size_t n = (size_t)nx * ny * nz, P = (size_t)nx * ny;
        HostCall hc;
        hc.init();
        et_plan plan;
        plan.st = hc.st;
        std::vector<float> ones(n, 1.0f);
        float *d_x = hc.upload(ones.data(), n);
        const float *d_meas = hc.upload(sino, P * V);
        const float *d_mu = attenuation ? hc.upload(attenuation, n) : nullptr;
        if (kind != EM_OSEM_RANDOM) {
            Cameras cams;
            int rc = expand_cameras(cams3, V, 3, vs, cams);
            if (rc) return rc;
            rc = build_plan(&plan, nx, ny, nz, cams, psf, psf ? psf_h : 0, bg, bga);
            if (rc) return rc;
            if (kind == EM_OSEM_ORDERED && param > 1 && !plan.fast) {
                // off the fast path: one plan per subset, reference update on each
                std::vector<std::unique_ptr<et_plan>> sub;
                for (int s = 0; s < param; s++) {
                    std::vector<int> idx;
                    for (int v = s; v < V; v += param) idx.push_back(v);
                    const std::vector<float> c = camera_subset(cams3, V, idx);
                    Cameras cs;
                    if ((rc = expand_cameras(c.data(), (int)idx.size(), 3, vs, cs))) return rc;
                    sub.emplace_back(new et_plan());
                    sub.back()->st = hc.st;
                    if ((rc = build_plan(sub.back().get(), nx, ny, nz, cs, psf, psf ? psf_h : 0, bg, bga))) return rc;
                }
                float *d_msub = hc.alloc(P * ((V + param - 1) / param));
                for (int it = 0; it < iterations; it++)
                    for (int s = 0; s < param; s++) {
                        int k = 0;
                        for (int v = s; v < V; v += param, k++)
                            check_hip(hipMemcpyAsync(d_msub + k * P, d_meas + (size_t)v * P, P * sizeof(float),
                                                     hipMemcpyDeviceToDevice, hc.st),
                                      niftyrec_error_transfergpu);
                        em_step_general(sub[s].get(), d_x, d_msub, d_mu, eps);
                    }
            } else {
                // chunking from the launch size, as the one-step entry points
                plan.tune_D = launch_depth_chunks(&plan, kind == EM_MLEM ? V : (V + param - 1) / param);
                op_osem(&plan, d_x, d_meas, d_mu, kind == EM_MLEM ? 1 : param, iterations, eps);
            }
        } else {
            const int nsub = V / param;
            float *d_msub = hc.alloc(P * nsub);
            int *d_idx = reinterpret_cast<int *>(hc.alloc(nsub));
            for (int it = 0; it < iterations; it++) {
                const std::vector<int> &idx = draws[it];
                const std::vector<float> c = camera_subset(cams3, V, idx);
                Cameras cs;
                int rc = expand_cameras(c.data(), nsub, 3, vs, cs);
                if (rc) return rc;
                // the same plan object each step: its workspaces are reused, its
                // OSEM caches rebuilt for the new cameras
                rc = build_plan(&plan, nx, ny, nz, cs, psf, psf ? psf_h : 0, bg, bga);
                if (rc) return rc;
                plan.osem_valid = false;
                plan.tune_D = launch_depth_chunks(&plan, nsub);
                check_hip(hipMemcpyAsync(d_idx, idx.data(), sizeof(int) * nsub, hipMemcpyHostToDevice, hc.st),
                          niftyrec_error_transfergpu);
                check_launch(et_launch_gather_planes(d_msub, d_meas, d_idx, nsub, P, (et_stream_t)hc.st));
                // the host copy of idx must outlive the transfer
                check_hip(hipStreamSynchronize(hc.st), niftyrec_error_kernel);
                op_osem(&plan, d_x, d_msub, d_mu, 1, 1, eps);
            }
        }
        hc.download(activity, d_x, n);
        return niftyrec_success;
    });
}

thread_local char t_msg[2048];

}  // namespace

// ===========================================================================
// Batches of plans: one OSEM iteration over many problems as a short list of
// grouped launches (problems with the same PSF radius, block width and
// attenuation presence share a launch), captured once into a hipGraph.
// ===========================================================================

struct et_batch {
    int device = 0;
    std::vector<et_plan *> plans;
    hipStream_t own = nullptr;        // used when the caller passes the null stream
    DevBuf tables;                    // argument records and block tables of every launch
    struct Launch {
        int kind;                     // 0 projection, 1 ratio, 2 rotated-frame backprojection, 3 pull + update
        int R, nthr, has_mu;
        size_t args_off, tab_off;     // byte offsets into `tables`
        int n, blocks;
    };
    std::vector<Launch> launches;     // sweep-kernel plans' launches of one OSEM iteration, in order
    std::vector<YlLaunch> yl_launches;  // row-lane plans' launches (table yl_tables), group after group
    DevBuf yl_tables;
    std::vector<size_t> yl_group_begin;  // group g: yl_launches [yl_group_begin[g], yl_group_begin[g + 1])
    std::vector<size_t> subset_begin;   // subset s: launches [subset_begin[s], subset_begin[s + 1])
    // Side streams and events of the fork / join that runs the row-lane groups'
    // chains next to each other (each group's launches only depend on its own).
    std::vector<hipStream_t> side;
    std::vector<hipEvent_t> ev;         // ev[0]: fork; ev[1 + k]: join of side stream k
    int S = 0;
    std::vector<const void *> key;    // buffers and chunkings the launch list refers to
    hipGraph_t graph = nullptr;
    hipGraphExec_t exec = nullptr;
    hipStream_t graph_stream = nullptr;

    void drop_graph()
    {
        if (exec) (void)hipGraphExecDestroy(exec);
        if (graph) (void)hipGraphDestroy(graph);
        exec = nullptr;
        graph = nullptr;
        graph_stream = nullptr;
    }
    ~et_batch()
    {
        drop_graph();
        for (hipStream_t x : side) (void)hipStreamDestroy(x);
        for (hipEvent_t e : ev) (void)hipEventDestroy(e);
        if (own) (void)hipStreamDestroy(own);
    }
};

namespace {

std::vector<const void *> batch_key(const et_batch *b, int S, float eps)
{
    std::vector<const void *> k;
    k.push_back(reinterpret_cast<const void *>((uintptr_t)S));
    uint32_t e;
    std::memcpy(&e, &eps, 4);
    k.push_back(reinterpret_cast<const void *>((uintptr_t)e));
    for (const et_plan *p : b->plans) {
        k.insert(k.end(), {p->x_yf.p, p->osem_meas.p, p->part.p, p->ratio.p, p->u.p, p->sens.p, p->osem_recs.p,
                           p->osem_E.p, p->osem_cur_mu, p->d_sep.p});
        k.push_back(reinterpret_cast<const void *>((uintptr_t)(p->osem_Dp * 1024 + p->osem_Db)));
        k.push_back(reinterpret_cast<const void *>((uintptr_t)p->osem_yl));
    }
    return k;
}

// Launch lists of one OSEM iteration.  Plans whose cached data are the
// row-lane kernels' take those (yl_launches, one table; one chain of all
// subsets per sweep template, yl_group_begin), the others the sweep kernels
// (launches, subset by subset, subset_begin): per subset, grouped
// projections, the ratios, grouped rotated-frame backprojections, then the
// pulls with the multiplicative update, with the same argument records as a
// plan run alone, so a batch result is bit-identical to its plans' own
// et_plan_osem.
void batch_build(et_batch *b, int S, hipStream_t st)
{
    std::vector<et_plan *> ylp, oldp;
    for (et_plan *p : b->plans) (p->osem_yl ? ylp : oldp).push_back(p);
    b->subset_begin.assign((size_t)S + 1, 0);
    // row-lane plans: one group per sweep template (their sweeps share
    // launches); a group's chain holds all its subsets in order and depends on
    // no other group's
    std::map<std::tuple<int, int, int, int>, std::vector<et_plan *>> ygroups;
    for (et_plan *p : ylp) ygroups[YlTable::sweep_key(p)].push_back(p);
    YlTable t;
    b->yl_group_begin.clear();
    for (auto &g : ygroups) {
        b->yl_group_begin.push_back(t.L.size());
        const std::vector<et_plan *> &gp = g.second;
        const size_t n = gp.size();
        for (int s = 0; s < S; s++) {
            std::vector<int> v0(n), nv(n), dp(n), db(n);
            std::vector<const float *> sens(n);
            for (size_t i = 0; i < n; i++) {
                const et_plan *p = gp[i];
                v0[i] = p->osem_start[s];
                nv[i] = p->osem_start[s + 1] - v0[i];
                dp[i] = p->osem_Dp;
                db[i] = p->osem_Db;
                sens[i] = p->sens.p + (size_t)s * p->yf_elems;
            }
            t.project(gp, v0, nv, dp);
            t.ratio(gp, v0, nv, dp);
            t.bpu(gp, v0, nv, db, false);
            t.pull(gp, v0, nv, sens, nullptr);
        }
    }
    b->yl_group_begin.push_back(t.L.size());
    // sweep-kernel plans
    std::map<std::tuple<int, int, int>, std::vector<int>> groups;
    for (size_t i = 0; i < oldp.size(); i++) {
        const et_plan *p = oldp[i];
        groups[std::make_tuple(p->R, nthreads_for(p->nx), p->osem_cur_mu ? 1 : 0)].push_back((int)i);
    }
    TableBytes tb;
    std::vector<et_batch::Launch> L;
    for (int s = 0; s < S; s++) {
        b->subset_begin[s] = L.size();
        if (oldp.empty()) continue;
        for (int mode : {ET_MODE_PROJECT, ET_MODE_BPU}) {
            if (mode == ET_MODE_BPU) {
                std::vector<EtRatioArgs> ra;
                BlockTable bt;
                for (et_plan *p : oldp) {
                    const int v0 = p->osem_start[s], nv = p->osem_start[s + 1] - v0;
                    const long long VP = (long long)nv * (long long)p->P;
                    ra.push_back({p->osem_meas.p + (size_t)v0 * p->P, p->part.p, p->osem_Dp, VP, p->osem_cur_eps,
                                  p->ratio.p});
                    bt.add((int)((VP + 255) / 256), 1);
                }
                L.push_back({1, 0, 256, 0, tb.put(ra), tb.put(bt.t), (int)ra.size(), bt.total});
            }
            for (auto &g : groups) {
                std::vector<EtSweepArgs> args;
                BlockTable bt;
                for (int i : g.second) {
                    et_plan *p = oldp[i];
                    const int v0 = p->osem_start[s], nv = p->osem_start[s + 1] - v0;
                    const float *r = p->osem_recs.p + (size_t)v0 * ET_VIEW_REC;
                    const int ts = et_sweep_slab_rows(mode, p->R);
                    if (mode == ET_MODE_PROJECT) {
                        args.push_back(project_args(p, r, nv, p->x_yf.p, p->osem_cur_mu, p->part.p, 0, p->osem_Dp,
                                                    nullptr, osem_att(p, v0)));
                        bt.add((p->ny + ts - 1) / ts, nv * p->osem_Dp);
                    } else {
                        args.push_back(bpu_args(p, r, nv, p->osem_cur_mu, ET_SRC_SINO, p->ratio.p, nullptr, nullptr, 0,
                                                0.f, p->osem_Db, nullptr, false, p->u.p, osem_att(p, v0)));
                        bt.add((p->ny + ts - 1) / ts, nv * p->osem_Db);
                    }
                }
                // with mu: both sweeps read the cached attenuation factors (has_mu 3)
                const int hm = std::get<2>(g.first) ? 3 : 0;
                L.push_back({mode == ET_MODE_PROJECT ? 0 : 2, std::get<0>(g.first), std::get<1>(g.first), hm,
                             tb.put(args), tb.put(bt.t), (int)args.size(), bt.total});
            }
        }
        std::vector<EtPullArgs> pa;
        BlockTable bt;
        for (et_plan *p : oldp) {
            const int v0 = p->osem_start[s], nv = p->osem_start[s + 1] - v0;
            const float *r = p->osem_recs.p + (size_t)v0 * ET_VIEW_REC;
            pa.push_back(pull_args(p, r, nv, p->u.p, false, ET_EPI_OSEM, false, false, nullptr, p->x_yf.p,
                                   p->sens.p + (size_t)s * p->yf_elems));
            bt.add((int)(((size_t)p->nx * p->nz + 255) / 256), (p->ny + ET_T - 1) / ET_T);
        }
        L.push_back({3, 0, 256, 0, tb.put(pa), tb.put(bt.t), (int)pa.size(), bt.total});
    }
    b->subset_begin[S] = L.size();
    b->drop_graph();
    if (!tb.bytes.empty()) {
        b->tables.ensure((align256(tb.bytes.size()) + 3) / 4, false);
        check_hip(hipMemcpyAsync(b->tables.p, tb.bytes.data(), tb.bytes.size(), hipMemcpyHostToDevice, st),
                  niftyrec_error_transfergpu);
        check_hip(hipStreamSynchronize(st), niftyrec_error_transfergpu);
    }
    if (!t.tb.bytes.empty()) yl_upload(b->yl_tables, t, st);
    b->launches = L;
    b->yl_launches = t.L;
    b->S = S;
}

// Side streams (at most ET_BATCH_SIDE, default priority) and the events of
// the fork / join, made once per batch.
constexpr int ET_BATCH_SIDE = 4;

void batch_streams(et_batch *b, int groups)
{
    const int want = std::min(groups, ET_BATCH_SIDE);
    while ((int)b->side.size() < want) {
        hipStream_t x = nullptr;
        check_hip(hipStreamCreateWithFlags(&x, hipStreamNonBlocking), niftyrec_error_kernel);
        b->side.push_back(x);
    }
    while (b->ev.size() < b->side.size() + 1) {
        hipEvent_t e = nullptr;
        check_hip(hipEventCreateWithFlags(&e, hipEventDisableTiming), niftyrec_error_kernel);
        b->ev.push_back(e);
    }
}

// One OSEM iteration of every plan.  The sweep-kernel plans' launches go on
// `st`; with more than one row-lane group, group g's chain goes on side stream
// g % side.size() after a fork from `st`, and `st` joins every side stream at
// the end (under capture: parallel branches of the graph), so the groups'
// launches overlap instead of each waiting out the tail of the one before.
void batch_issue(const et_batch *b, hipStream_t st)
{
    const int G = (int)b->yl_group_begin.size() - 1;
    const bool fork = G > 1 && !b->side.empty();
    if (fork) {
        check_hip(hipEventRecord(b->ev[0], st), niftyrec_error_kernel);
        for (hipStream_t x : b->side) check_hip(hipStreamWaitEvent(x, b->ev[0], 0), niftyrec_error_kernel);
    }
    for (int g = 0; g < G; g++)
        yl_issue(b->yl_tables.p, b->yl_launches, b->yl_group_begin[g], b->yl_group_begin[g + 1],
                 fork ? b->side[g % b->side.size()] : st);
    if (fork) {
        for (size_t k = 0; k < b->side.size(); k++) {
            check_hip(hipEventRecord(b->ev[1 + k], b->side[k]), niftyrec_error_kernel);
            check_hip(hipStreamWaitEvent(st, b->ev[1 + k], 0), niftyrec_error_kernel);
        }
    }
    const char *base = reinterpret_cast<const char *>(b->tables.p);
    for (int sub = 0; sub < b->S; sub++) {
        for (size_t k = b->subset_begin[sub]; k < b->subset_begin[sub + 1]; k++) {
            const auto &l = b->launches[k];
            const void *args = base + l.args_off, *tab = base + l.tab_off;
            const et_stream_t s = (et_stream_t)st;
            switch (l.kind) {
            case 0:
            case 2:
                check_launch(et_launch_sweep_group(l.kind == 0 ? ET_MODE_PROJECT : ET_MODE_BPU, l.R, l.has_mu, l.nthr,
                                                   static_cast<const EtSweepArgs *>(args), tab, l.n, l.blocks, s));
                break;
            case 1:
                check_launch(et_launch_ratio_group(static_cast<const EtRatioArgs *>(args), tab, l.n, l.blocks, s));
                break;
            default:
                check_launch(et_launch_pull_group(ET_EPI_OSEM, static_cast<const EtPullArgs *>(args), tab, l.n,
                                                  l.blocks, s));
            }
        }
    }
}

void op_batch_osem(et_batch *b, float *const *act, const float *const *sino, const float *const *mu, int S,
                   int iters, float eps, bool use_graph, hipStream_t st)
{
    // attenuation keys of every plan, read after one synchronisation
    bool any_fp = false;
    for (size_t i = 0; i < b->plans.size(); i++) {
        et_plan *p = b->plans[i];
        p->st = st;
        const float *m = mu ? mu[i] : nullptr;
        mu_key_launch(p, m);
        any_fp = any_fp || (m && p->mu_check);
    }
    if (any_fp) check_hip(hipStreamSynchronize(st), niftyrec_error_kernel);
    for (size_t i = 0; i < b->plans.size(); i++) {
        et_plan *p = b->plans[i];
        const float *m = mu ? mu[i] : nullptr;
        osem_prepare(p, act[i], sino[i], m, S, eps, mu_key_read(p, m));
    }
    const std::vector<const void *> key = batch_key(b, S, eps);
    if (key != b->key || b->subset_begin.empty()) {
        batch_build(b, S, st);
        batch_streams(b, (int)b->yl_group_begin.size() - 1);
        b->key = key;
    }
    if (use_graph && (!b->exec || b->graph_stream != st)) {
        b->drop_graph();
        check_hip(hipStreamBeginCapture(st, hipStreamCaptureModeThreadLocal), niftyrec_error_kernel);
        try {
            batch_issue(b, st);
        } catch (...) {
            hipGraph_t g = nullptr;
            (void)hipStreamEndCapture(st, &g);
            if (g) (void)hipGraphDestroy(g);
            throw;
        }
        check_hip(hipStreamEndCapture(st, &b->graph), niftyrec_error_kernel);
        check_hip(hipGraphInstantiate(&b->exec, b->graph, nullptr, nullptr, 0), niftyrec_error_kernel);
        b->graph_stream = st;
    }
    for (int it = 0; it < iters; it++) {
        if (use_graph)
            check_hip(hipGraphLaunch(b->exec, st), niftyrec_error_kernel);
        else
            batch_issue(b, st);
    }
    for (size_t i = 0; i < b->plans.size(); i++) osem_finish(b->plans[i], act[i]);
}

}  // namespace

// ===========================================================================
// Exported C ABI
// ===========================================================================

// ---------------------------------------------------------------------------
// Projection without allocation.  Its only temporaries, the YF copies of the
// activity and of the attenuation map, live in `scratch`, a static device
// array: every GPU the library runs on holds its own copy until the process
// ends.  Host code never passes `scratch` itself (in host code the name is a
// host-side stand-in, not a device address): the first call on a device, with
// that device current, takes the array's device address from
// hipGetSymbolAddress and keeps it for that device.  Calls share the array, so
// calls that overlap on one GPU (different streams, or different host threads)
// race; each call clears what it needs zeroed (with a fill kernel, which graph
// capture records as a kernel node) before packing.
// ---------------------------------------------------------------------------
namespace {

constexpr size_t ET_NOALLOC_SCRATCH_BYTES = size_t(32) << 20;
static __device__ __align__(256) unsigned char scratch[ET_NOALLOC_SCRATCH_BYTES];

constexpr int ET_NOALLOC_MAX_DEVICES = 64;
std::mutex g_scratch_mu;
void *g_scratch_dev[ET_NOALLOC_MAX_DEVICES] = {};

// Device address of this device's `scratch` (the device must be current).
float *scratch_of(int device)
{
    if (device < 0 || device >= ET_NOALLOC_MAX_DEVICES) throw EtError{niftyrec_error_parameters};
    std::lock_guard<std::mutex> lk(g_scratch_mu);
    if (!g_scratch_dev[device]) {
        void *ptr = nullptr;
        if (hipGetSymbolAddress(&ptr, HIP_SYMBOL(scratch)) != hipSuccess || !ptr)
            throw EtError{niftyrec_error_nogpuinstalled};
        g_scratch_dev[device] = ptr;
    }
    return static_cast<float *>(g_scratch_dev[device]);
}

size_t round256(size_t bytes) { return (bytes + 255) / 256 * 256; }

void op_project_noalloc(et_plan *p, const float *act, const float *mu, float *sino, int truncate)
{
    // axial plans with zero backgrounds (the fast path) only: the general
    // path's rotated copies would not fit
    if (!p->fast || !act || !sino) throw EtError{niftyrec_error_parameters};
    const size_t act_bytes = round256(p->yf_elems * sizeof(float));
    const size_t mu_elems = p->yf_elems + ET_YF_SLACK;
    const size_t need = act_bytes + (mu ? round256(mu_elems * sizeof(float)) : 0);
    if (need > ET_NOALLOC_SCRATCH_BYTES) throw EtError{niftyrec_error_parameters};
    float *base = scratch_of(p->device);
    float *act_yf = base, *mu_yf = mu ? base + act_bytes / sizeof(float) : nullptr;
    const et_stream_t st = (et_stream_t)p->st;
    // guard rows and borders of the YF layouts read as zeros
    check_launch(et_launch_fill(act_yf, mu ? act_bytes / sizeof(float) + mu_elems : p->yf_elems, 0.f, st));
    check_launch(et_launch_pack_yf(act, act_yf, p->nx, p->ny, p->nz, p->yp, p->gy, st));
    if (mu) check_launch(et_launch_pack_yf(mu, mu_yf, p->nx, p->ny, p->nz, p->yp, p->gmu, st));
    // one depth chunk: the sweep writes the sinogram itself (no partial sums)
    const EtSweepArgs a = project_args(p, p->d_recs.p, p->V, act_yf, mu_yf, sino, truncate, 1, nullptr, nullptr);
    check_launch(et_launch_sweep(ET_MODE_PROJECT, p->R, mu ? 1 : 0, 1, a, st));
}

}  // namespace

extern "C" {

int status_success(int *status)
{
    *status = STATUS_SUCCESS;
    return STATUS_SUCCESS;
}
int status_io_error(int *status)
{
    *status = STATUS_IO_ERROR;
    return STATUS_SUCCESS;
}
int status_initialisation_error(int *status)
{
    *status = STATUS_INITIALISATION_ERROR;
    return STATUS_SUCCESS;
}
int status_parameter_error(int *status)
{
    *status = STATUS_PARAMETER_ERROR;
    return STATUS_SUCCESS;
}
int status_unhandled_error(int *status)
{
    *status = STATUS_UNHANDLED_ERROR;
    return STATUS_SUCCESS;
}
int echo(int *in, int *out)
{
    *out = *in;
    return STATUS_SUCCESS;
}

char *niftyrec_error_msg(int status)
{
    const char *m;
    switch (status) {
    case niftyrec_success: m = "niftyrec: OK.\n"; break;
    case niftyrec_error_unspecified: m = "niftyrec: ERROR - unspecified error.\n"; break;
    case niftyrec_error_parameters: m = "niftyrec: ERROR - inconsistent parameters.\n"; break;
    case niftyrec_error_kernel: m = "niftyrec: ERROR - gpu kernel execution error.\n"; break;
    case niftyrec_error_allocgpu: m = "niftyrec: ERROR - GPU memory allocation.\n"; break;
    case niftyrec_error_alloccpu: m = "niftyrec: ERROR - cpu memory allocation.\n"; break;
    case niftyrec_error_transfergpu: m = "niftyrec: ERROR - transferring memory between host and GPU.\n"; break;
    case niftyrec_error_nogpubuilt: m = "niftyrec: ERROR - built without GPU support.\n"; break;
    case niftyrec_error_nogpuinstalled: m = "niftyrec: ERROR - no GPUs installed in the system.\n"; break;
    default: m = "niftyrec: unknown status.\n"; break;
    }
    std::snprintf(t_msg, sizeof(t_msg), "%s", m);
    return t_msg;
}

const char *et_library_version(void) { return ET_VERSION; }

int et_array_list_gpus(int *gpu_count, int *gpus_info_array)
{
    int count = 0;
    if (hipGetDeviceCount(&count) != hipSuccess || count <= 0) {
        *gpu_count = 0;
        return niftyrec_error_nogpuinstalled;
    }
    const int cmax = std::min(count, 16);
    for (int d = 0; d < cmax; d++) {
        hipDeviceProp_t prop;
        if (hipGetDeviceProperties(&prop, d) != hipSuccess) return niftyrec_error_unspecified;
        // {id, gflops proxy (CUs * clock), CUs, clock kHz, MB}, reference emission-lib/_et.cpp:2004-2018
        gpus_info_array[5 * d + 0] = d;
        gpus_info_array[5 * d + 1] = prop.multiProcessorCount * prop.clockRate;
        gpus_info_array[5 * d + 2] = prop.multiProcessorCount;
        gpus_info_array[5 * d + 3] = prop.clockRate;
        gpus_info_array[5 * d + 4] = (int)std::floor(prop.totalGlobalMem / 1000000.0);
    }
    *gpu_count = count;
    return niftyrec_success;
}

int et_array_set_gpu(int id)
{
    int count = 0;
    if (hipGetDeviceCount(&count) != hipSuccess || count <= 0) return niftyrec_error_nogpuinstalled;
    if (id < 0 || id >= count) return niftyrec_error_parameters;
    std::lock_guard<std::mutex> lk(g_device_mutex);
    g_device = id;
    return niftyrec_success;
}

int et_array_set_gpu_pointer(int *id) { return et_array_set_gpu(*id); }

int et_array_isinstalled(void) { return 0; }

int et_array_reset_gpu(void)
{
    int st = select_device();
    if (st) return st;
    return hipDeviceSynchronize() == hipSuccess ? niftyrec_success : niftyrec_error_kernel;
}

unsigned short int et_array_is_block_multiple(unsigned short int size)
{
    (void)size;
    return 1;
}

unsigned short int et_array_get_block_size(void) { return 1; }

int SPECT_project_parallelholes(float *activity, int *activity_size, float *sinogram, int *sinogram_size,
                                float *cameras, int *cameras_size, float *psf, int *psf_size, float *attenuation,
                                int *attenuation_size, float *Background, float *Background_attenuation, int *Gpu,
                                int *Truncate_negative_values)
{
    (void)Gpu;
    return project_host(activity, activity_size, sinogram, sinogram_size, cameras, cameras_size, psf, psf_size,
                        attenuation, attenuation_size, *Background, *Background_attenuation,
                        *Truncate_negative_values);
}

int SPECT_backproject_parallelholes(float *sino, int *sino_size, float *bkpr, int *bkpr_size, float *cameras,
                                    int *cameras_size, float *psf, int *psf_size, float *attenuation,
                                    int *attenuation_size, float *Background, float *Background_attenuation, int *Gpu,
                                    int *Truncate_negative_values)
{
    (void)Gpu;
    return backproject_host(sino, sino_size, bkpr, bkpr_size, cameras, cameras_size, psf, psf_size, attenuation,
                            attenuation_size, *Background, *Background_attenuation, *Truncate_negative_values, false);
}

int et_array_project(float *activity, int *activity_size, float *sinogram, int *sinogram_size, float *cameras,
                     int *cameras_size, float *psf, int *psf_size, float *attenuation, int *attenuation_size,
                     float background, float background_attenuation, int GPU, int truncate_negative_values)
{
    (void)GPU;
    return project_host(activity, activity_size, sinogram, sinogram_size, cameras, cameras_size, psf, psf_size,
                        attenuation, attenuation_size, background, background_attenuation, truncate_negative_values);
}

int et_array_fisher_grid(float *activity_ptr, int *activity_size, float *cameras, int *cameras_size, float *psf,
                         int *psf_size, float *grid_ptr, float *fisher_ptr, float *fisher_prior_ptr, int *fisher_size,
                         float *attenuation, int *attenuation_size, float epsilon, float background,
                         float background_attenuation, int GPU)
{
    (void)epsilon;  // unused by the reference too (its use is commented out, _et.cpp:576-578)
    if (!activity_ptr || !activity_size || !cameras || !cameras_size || !psf_size || !grid_ptr || !fisher_ptr ||
        !fisher_size)
        return niftyrec_error_parameters;
    return fisher_host(false, activity_ptr, activity_size, activity_size, cameras, cameras_size, psf, psf_size,
                       grid_ptr, fisher_ptr, fisher_prior_ptr, fisher_size, attenuation, attenuation_size, background,
                       background_attenuation, GPU);
}

int et_array_fisher_grid_projection(float *sinogram_ptr, int *sinogram_size, int *bkpr_size, float *cameras,
                                    int *cameras_size, float *psf, int *psf_size, float *grid_ptr, float *fisher_ptr,
                                    float *fisher_prior_ptr, int *fisher_size, float *attenuation,
                                    int *attenuation_size, float epsilon, float background,
                                    float background_attenuation, int GPU)
{
    (void)epsilon;
    (void)attenuation;  // the projection is given: attenuation and backgrounds play no part
    (void)attenuation_size;
    if (!sinogram_ptr || !sinogram_size || !bkpr_size || !cameras || !cameras_size || !psf_size || !grid_ptr ||
        !fisher_ptr || !fisher_size)
        return niftyrec_error_parameters;
    return fisher_host(true, sinogram_ptr, sinogram_size, bkpr_size, cameras, cameras_size, psf, psf_size, grid_ptr,
                       fisher_ptr, fisher_prior_ptr, fisher_size, nullptr, nullptr, background,
                       background_attenuation, GPU);
}

int et_array_backproject(float *sino, int *sino_size, float *bkpr, int *bkpr_size, float *cameras, int *cameras_size,
                         float *psf, int *psf_size, float *attenuation, int *attenuation_size, float background,
                         float background_attenuation, int GPU, int truncate_negative_values)
{
    (void)GPU;
    return backproject_host(sino, sino_size, bkpr, bkpr_size, cameras, cameras_size, psf, psf_size, attenuation,
                            attenuation_size, background, background_attenuation, truncate_negative_values, false);
}

int et_array_gradient_attenuation(float *sino, int *sino_size, float *activity, int *activity_size, float *gradient,
                                  int *gradient_size, float *cameras, int *cameras_size, float *psf, int *psf_size,
                                  float *attenuation, int *attenuation_size, float background,
                                  float background_attenuation, int GPU, int truncate_negative_values)
{
    (void)GPU;
    // Validation, emission-lib/_et_array_interface.cpp:727-780.  The 2-D branch
    // (sinogram [N, n_cameras, 1] with more than one camera) describes a volume
    // of one axial row: the reference wraps the sinogram as
    // [gradient_size[0], gradient_size[1], n_cameras] (:830-834), so the gradient
    // must be [N, 1, N] and the N * n_cameras sinogram values are read as
    // [N, 1, n_cameras].  Its PSF is the 3-D one, [h, h, N]; the [h, N]
    // spelling of the 2-D check (:752) is refused.
    const bool no_psf = is_absent(psf_size) || psf == nullptr;
    const int V = cameras_size[0], cols = cameras_size[1];
    if (!(cols == 1 || cols == 3 || cols == 6) || V <= 0) return niftyrec_error_parameters;
    const bool two_d = sino_size[2] == 1 && V > 1;
    if (two_d) {
        if (sino_size[1] != V) return niftyrec_error_parameters;
    } else {
        if (sino_size[2] != V) return niftyrec_error_parameters;
    }
    if (!no_psf && (psf_size[0] % 2 != 1 || psf_size[1] % 2 != 1 || psf_size[2] != sino_size[0]))
        return niftyrec_error_parameters;
    const int nx = gradient_size[0], ny = gradient_size[1], nz = gradient_size[2];
    if (nx != sino_size[0] || ny != (two_d ? 1 : sino_size[1]) || nz <= 0) return niftyrec_error_parameters;
    if (activity == nullptr || attenuation == nullptr) return niftyrec_error_parameters;
    if (!no_psf && (psf_size[0] != psf_size[1] || nz != nx)) return niftyrec_error_parameters;
    if (activity_size && (activity_size[0] != nx || activity_size[1] != ny || activity_size[2] != nz))
        return niftyrec_error_parameters;
    if (attenuation_size && !is_absent(attenuation_size) &&
        (attenuation_size[0] != nx || attenuation_size[1] != ny || attenuation_size[2] != nz))
        return niftyrec_error_parameters;
    if (nx > ET_MAX_NX) return niftyrec_error_parameters;
    int st = select_device();
    if (st) return st;
    return guarded([&]() -> int {
        Cameras cams;
        int rc = expand_cameras(cameras, V, cols, activity_size ? activity_size : gradient_size, cams);
        if (rc) return rc;
        et_plan plan;
        HostCall hc;
        hc.init();
        plan.st = hc.st;
        rc = build_plan(&plan, nx, ny, nz, cams, no_psf ? nullptr : psf, no_psf ? 0 : psf_size[0], background,
                        background_attenuation);
        if (rc) return rc;
        const size_t n = (size_t)nx * ny * nz;
        const float *d_sino = hc.upload(sino, (size_t)nx * ny * V);
        const float *d_act = hc.upload(activity, n);
        const float *d_mu = hc.upload(attenuation, n);
        float *d_out = hc.alloc(n);
        op_gradient(&plan, d_sino, d_act, d_mu, d_out, false, truncate_negative_values);
        hc.download(gradient, d_out, n);
        return niftyrec_success;
    });
}

int et_array_rotate(float *image, int *size, float *rotated_image, float *angles, float *centers, float background,
                    int axis_order, int GPU)
{
    (void)GPU;
    if (size[0] <= 0 || size[1] <= 0 || size[2] <= 0) return niftyrec_error_parameters;
    int st = select_device();
    if (st) return st;
    return guarded([&]() -> int {
        float M[16];
        rotation_matrix(M, angles[0], angles[1], angles[2], centers[0], centers[1], centers[2], axis_order);
        HostCall hc;
        hc.init();
        const size_t n = (size_t)size[0] * size[1] * size[2];
        const float *d_in = hc.upload(image, n);
        float *d_M = hc.upload(M, 12);
        float *d_out = hc.alloc(n);
        check_launch(et_launch_resample_general(d_in, size[0], size[1], size[2], d_M, background, d_out, 0, 0, 0,
                                                (et_stream_t)hc.st));
        hc.download(rotated_image, d_out, n);
        return niftyrec_success;
    });
}

int et_array_affine(float *image, int *image_size, float *transformed_image, int *transformed_image_size,
                    float *affine, int *affine_size, float background, int GPU)
{
    (void)GPU;
    (void)affine_size;
    if (image_size[0] <= 0 || image_size[1] <= 0 || image_size[2] <= 0) return niftyrec_error_parameters;
    if (transformed_image_size && (transformed_image_size[0] != image_size[0] ||
                                   transformed_image_size[1] != image_size[1] ||
                                   transformed_image_size[2] != image_size[2]))
        return niftyrec_error_parameters;
    int st = select_device();
    if (st) return st;
    return guarded([&]() -> int {
        HostCall hc;
        hc.init();
        const size_t n = (size_t)image_size[0] * image_size[1] * image_size[2];
        const float *d_in = hc.upload(image, n);
        float *d_M = hc.upload(affine, 12);
        float *d_out = hc.alloc(n);
        check_launch(et_launch_resample_general(d_in, image_size[0], image_size[1], image_size[2], d_M, background,
                                                d_out, 0, 0, 0, (et_stream_t)hc.st));
        hc.download(transformed_image, d_out, n);
        return niftyrec_success;
    });
}

// Partial projection, reference _et_array_interface.cpp:1561-1787 and the
// CUDA path _et.cpp:2106-2440 (the reference CPU path is a stub).  Per view:
// activity and attenuation resampled on the general path (pads background,
// background_attenuation), activity blurred by the separated PSF, partial line
// integrals (k_partial_line), then the partial sums pulled back to the image
// frame (negated angles, ZYX order, pad background) when do_rotate_partial.
// partialsum: [nx, ny, nz, n_cameras] host memory.
int et_array_project_partial(float *activity, int *activity_size, float *sinogram, int *sinogram_size,
                             float *partialsum, int *partialsum_size, float *cameras, int *cameras_size, float *psf,
                             int *psf_size, float *attenuation, int *attenuation_size, float background,
                             float *background_image, float background_attenuation, int GPU,
                             int truncate_negative_values, int do_rotate_partial)
{
    (void)GPU;
    (void)sinogram_size;
    const bool no_act = activity == nullptr || is_absent(activity_size);
    const bool no_mu = attenuation == nullptr || is_absent(attenuation_size);
    const bool no_psf = psf == nullptr || is_absent(psf_size);
    if (no_act && no_mu) return niftyrec_error_parameters;
    if (no_act && activity_size) {
        // the reference reports the attenuation size through activity_size (:1578-1582)
        activity_size[0] = attenuation_size[0];
        activity_size[1] = attenuation_size[1];
        activity_size[2] = attenuation_size[2];
    }
    const int *vs = no_act ? attenuation_size : activity_size;
    const int nx = vs[0], ny = vs[1], nz = vs[2];
    if (nx <= 0 || ny <= 0 || nz <= 0) return niftyrec_error_parameters;
    if (!no_act && !no_mu &&
        (attenuation_size[0] != nx || attenuation_size[1] != ny || attenuation_size[2] != nz))
        return niftyrec_error_parameters;
    const int V = cameras_size[0], cols = cameras_size[1];
    if (!(cols == 1 || cols == 3 || cols == 6) || V <= 0) return niftyrec_error_parameters;
    int h = 0;
    if (!no_psf) {
        h = psf_size[0];
        if (h % 2 != 1 || psf_size[1] % 2 != 1 || psf_size[2] != nx || psf_size[0] != psf_size[1])
            return niftyrec_error_parameters;
    }
    if (partialsum_size && (partialsum_size[0] != nx || partialsum_size[1] != ny || partialsum_size[2] != nz ||
                            partialsum_size[3] != V))
        return niftyrec_error_parameters;
    int st = select_device();
    if (st) return st;
    return guarded([&]() -> int {
        const int dims[3] = {nx, ny, nz};
        Cameras cams;
        int rc = expand_cameras(cameras, V, cols, dims, cams);
        if (rc) return rc;
        HostCall hc;
        hc.init();
        const et_stream_t s = (et_stream_t)hc.st;
        const size_t n = (size_t)nx * ny * nz, P = (size_t)nx * ny;
        // forward and inverse (reference pull) matrices of every view
        std::vector<float> mats((size_t)V * 24);
        for (int v = 0; v < V; v++) {
            float M[16];
            rotation_matrix(M, cams.ang[v], cams.ang[V + v], cams.ang[2 * V + v], cams.ctr[v], cams.ctr[V + v],
                            cams.ctr[2 * V + v], 0);
            std::memcpy(&mats[(size_t)24 * v], M, sizeof(float) * 12);
            rotation_matrix(M, -cams.ang[v], -cams.ang[V + v], -cams.ang[2 * V + v], cams.ctr[v], cams.ctr[V + v],
                            cams.ctr[2 * V + v], 5);
            std::memcpy(&mats[(size_t)24 * v + 12], M, sizeof(float) * 12);
        }
        const float *d_M = hc.upload(mats.data(), mats.size());
        const float *d_act = no_act ? nullptr : hc.upload(activity, n);
        const float *d_mu = no_mu ? nullptr : hc.upload(attenuation, n);
        const float *d_bgimg = background_image ? hc.upload(background_image, P) : nullptr;
        const float *d_sep = nullptr;
        if (h > 0) {
            const int c = (h - 1) / 2;
            std::vector<float> sep((size_t)nz * 2 * h);
            for (int z = 0; z < nz; z++) {
                const float *pl = psf + (size_t)h * h * z;
                for (int i = 0; i < h; i++) {
                    sep[(size_t)2 * h * z + i] = pl[c * h + i];
                    sep[(size_t)2 * h * z + h + i] = pl[c + i * h] / pl[c * h + c];
                }
            }
            d_sep = hc.upload(sep.data(), sep.size());
        }
        float *rot_act = no_act ? nullptr : hc.alloc(n);
        float *tmp = h > 0 ? hc.alloc(n) : nullptr;
        float *rot_mu = no_mu ? nullptr : hc.alloc(n);
        float *part = hc.alloc(n);
        float *pout = do_rotate_partial ? hc.alloc(n) : part;
        float *d_sino = hc.alloc(P * V);
        for (int v = 0; v < V; v++) {
            const float *Mf = d_M + (size_t)24 * v, *Mi = Mf + 12;
            if (d_act) {
                check_launch(et_launch_resample_general(d_act, nx, ny, nz, Mf, background, rot_act, 0, 0, 0, s));
                if (d_sep) {
                    check_launch(et_launch_conv_axis(rot_act, tmp, nx, ny, nz, 1, nx, (long)P, 0, 0, d_sep, h, 0, s));
                    check_launch(et_launch_conv_axis(tmp, rot_act, nx, ny, nz, 1, nx, (long)P, 0, 1, d_sep, h, 0, s));
                }
            }
            if (d_mu)
                check_launch(
                    et_launch_resample_general(d_mu, nx, ny, nz, Mf, background_attenuation, rot_mu, 0, 0, 0, s));
            check_launch(et_launch_partial_line(rot_act, rot_mu, d_bgimg, background, nx, ny, nz, d_sino + P * v,
                                                part, s));
            if (do_rotate_partial)
                check_launch(et_launch_resample_general(part, nx, ny, nz, Mi, background, pout, 0, 0, 0, s));
            hc.download(partialsum + n * v, pout, n);
        }
        if (truncate_negative_values) check_launch(et_launch_truncate(d_sino, P * V, s));
        hc.download(sinogram, d_sino, P * V);
        return niftyrec_success;
    });
}

int et_array_convolve(float *image, int *image_size, float *out, int *out_size, float *psf, int *psf_size, int GPU)
{
    (void)GPU;
    const int nx = image_size[0], ny = image_size[1], nz = image_size[2];
    if (nx <= 0 || ny <= 0 || nz <= 0) return niftyrec_error_parameters;
    if (out_size && (out_size[0] != nx || out_size[1] != ny || out_size[2] != nz)) return niftyrec_error_parameters;
    if (psf_size[0] % 2 != 1 || psf_size[0] != psf_size[1] || psf_size[2] != nz) return niftyrec_error_parameters;
    int st = select_device();
    if (st) return st;
    return guarded([&]() -> int {
        HostCall hc;
        hc.init();
        const int h = psf_size[0], c = (h - 1) / 2;
        std::vector<float> sep((size_t)nz * 2 * h);
        for (int n = 0; n < nz; n++) {
            const float *pl = psf + (size_t)h * h * n;
            const float norm = pl[c * h + c];
            for (int i = 0; i < h; i++) {
                sep[(size_t)2 * h * n + i] = pl[c * h + i];
                sep[(size_t)2 * h * n + h + i] = pl[c + i * h] / norm;
            }
        }
        const size_t n = (size_t)nx * ny * nz;
        const float *d_in = hc.upload(image, n);
        const float *d_sep = hc.upload(sep.data(), sep.size());
        float *d_tmp = hc.alloc(n);
        float *d_out = hc.alloc(n);
        check_launch(et_launch_conv_axis(d_in, d_tmp, nx, ny, nz, 1, nx, (long)nx * ny, 0, 0, d_sep, h, 0,
                                         (et_stream_t)hc.st));
        check_launch(et_launch_conv_axis(d_tmp, d_out, nx, ny, nz, 1, nx, (long)nx * ny, 0, 1, d_sep, h, 0,
                                         (et_stream_t)hc.st));
        hc.download(out, d_out, n);
        return niftyrec_success;
    });
}

int et_array_calculate_size_psf(unsigned int *psf_size_x, unsigned int *psf_size_y, float fwhm_pixels_dist0,
                                float sensitivity0, float dist0, float fwhm_pixels_dist1, float sensitivity1,
                                float dist1)
{
    (void)fwhm_pixels_dist0; (void)sensitivity0; (void)dist0;
    (void)fwhm_pixels_dist1; (void)sensitivity1; (void)dist1;
    *psf_size_x = 5;
    *psf_size_y = 5;
    return 0;
}

int et_array_make_psf(float *psf_data, unsigned int psf_size_x, unsigned int psf_size_y, float fwhm_pixels_dist0,
                      float sensitivity0, float dist0, float fwhm_pixels_dist1, float sensitivity1, float dist1,
                      unsigned int N_psf_planes)
{
    unsigned int sx, sy;
    et_array_calculate_size_psf(&sx, &sy, fwhm_pixels_dist0, sensitivity0, dist0, fwhm_pixels_dist1, sensitivity1,
                                dist1);
    if (psf_size_x != sx || psf_size_y != sy) return 1;
    for (size_t i = 0; i < (size_t)psf_size_x * psf_size_y * N_psf_planes; i++) psf_data[i] = 1.f;
    return 0;
}

// Inclusive angular endpoints, emission-lib/_et_array_interface.cpp:923-937.
int et_array_make_cameras(float *cameras_data, float firstcamera_deg, float lastcamera_deg, unsigned int n_cameras,
                          unsigned int rotation_axis)
{
    const float PI_F = 3.141592653589793f;
    for (unsigned int cam = 0; cam < n_cameras; cam++) {
        const float angle =
            (firstcamera_deg + (lastcamera_deg - firstcamera_deg) * cam / (n_cameras - 1)) / 180.0f * PI_F;
        for (unsigned int axis = 0; axis < 3; axis++)
            cameras_data[cam + axis * n_cameras] = (axis == rotation_axis) ? angle : 0.0f;
    }
    return 0;
}

int et_array_mlem_step(float *activity_data, int size_x, int size_y, float *sinogram_data, int n_cameras,
                       float *cameras_data, int use_attenuation, float *attenuation_data, int use_psf,
                       float *psf_data, unsigned int psf_size_x, unsigned int psf_size_y, float background,
                       float background_attenuation, float epsilon, int use_gpu)
{
    (void)use_gpu;
    if (size_x <= 0 || size_y <= 0 || n_cameras <= 0) return niftyrec_error_parameters;
    if (use_psf && (psf_size_x % 2 != 1 || psf_size_x != psf_size_y)) return niftyrec_error_parameters;
    if (size_x > ET_MAX_NX) return niftyrec_error_parameters;
    return em_step_host(activity_data, size_x, size_y, sinogram_data, n_cameras, cameras_data,
                        use_attenuation ? attenuation_data : nullptr, use_psf ? psf_data : nullptr,
                        (int)psf_size_x, background, background_attenuation, epsilon);
}

// Random camera subset drawn with replacement from glibc rand(), as
// emission-lib/_et_array_interface.cpp:981-1009 does (two rand() calls per
// camera, R = base + fine/scale in single precision); an index equal to
// n_cameras (the reference's clamp) is clamped to n_cameras - 1 here.
int et_array_osem_step(float *activity_data, int size_x, int size_y, unsigned int subset_order, float *sinogram_data,
                       int n_cameras, float *cameras_data, int use_attenuation, float *attenuation_data, int use_psf,
                       float *psf_data, unsigned int psf_size_x, unsigned int psf_size_y, float background,
                       float background_attenuation, float epsilon, int use_gpu)
{
    (void)use_gpu;
    if (size_x <= 0 || size_y <= 0 || n_cameras <= 0 || subset_order == 0) return niftyrec_error_parameters;
    const int nsub = n_cameras / (int)subset_order;
    if (nsub <= 0) return niftyrec_error_parameters;
    if (use_psf && (psf_size_x % 2 != 1 || psf_size_x != psf_size_y)) return niftyrec_error_parameters;
    const size_t P = (size_t)size_x * size_y;
    const std::vector<int> idx = random_subset(n_cameras, nsub);
    const std::vector<float> cams = camera_subset(cameras_data, n_cameras, idx);
    std::vector<float> sino_sub(P * nsub);
    for (int c = 0; c < nsub; c++)
        std::memcpy(&sino_sub[(size_t)c * P], sinogram_data + (size_t)idx[c] * P, sizeof(float) * P);
    return em_step_host(activity_data, size_x, size_y, sino_sub.data(), nsub, cams.data(),
                        use_attenuation ? attenuation_data : nullptr, use_psf ? psf_data : nullptr,
                        (int)psf_size_x, background, background_attenuation, epsilon);
}

int et_array_mlem(float *activity_data, unsigned int size_x, unsigned int size_y, float *sinogram_data, int n_cameras,
                  float firstcamera, float lastcamera, unsigned int rotation_axis, int iterations, int use_attenuation,
                  float *attenuation_data, int use_psf, float *psf_data, unsigned int psf_size_x,
                  unsigned int psf_size_y, float background, float background_attenuation, float epsilon, int use_gpu)
{
    (void)use_gpu;
    if (n_cameras <= 0 || rotation_axis > 2) return niftyrec_error_parameters;
    if (use_psf && (psf_size_x % 2 != 1 || psf_size_x != psf_size_y)) return niftyrec_error_parameters;
    std::vector<float> cams((size_t)3 * n_cameras);
    et_array_make_cameras(cams.data(), firstcamera, lastcamera, n_cameras, rotation_axis);
    return em_driver_host(EM_MLEM, activity_data, (int)size_x, (int)size_y, sinogram_data, n_cameras, cams.data(),
                          use_attenuation ? attenuation_data : nullptr, use_psf ? psf_data : nullptr, (int)psf_size_x,
                          background, background_attenuation, epsilon, 1, iterations);
}

int et_array_osem(float *activity_data, unsigned int size_x, unsigned int size_y, unsigned int subset_order,
                  float *sinogram_data, int n_cameras, float firstcamera, float lastcamera,
                  unsigned int rotation_axis, int iterations, int use_attenuation, float *attenuation_data,
                  int use_psf, float *psf_data, unsigned int psf_size_x, unsigned int psf_size_y, float background,
                  float background_attenuation, float epsilon, int use_gpu)
{
    (void)use_gpu;
    if (n_cameras <= 0 || rotation_axis > 2 || subset_order == 0) return niftyrec_error_parameters;
    if (use_psf && (psf_size_x % 2 != 1 || psf_size_x != psf_size_y)) return niftyrec_error_parameters;
    std::vector<float> cams((size_t)3 * n_cameras);
    et_array_make_cameras(cams.data(), firstcamera, lastcamera, n_cameras, rotation_axis);
    return em_driver_host(EM_OSEM_RANDOM, activity_data, (int)size_x, (int)size_y, sinogram_data, n_cameras,
                          cams.data(), use_attenuation ? attenuation_data : nullptr, use_psf ? psf_data : nullptr,
                          (int)psf_size_x, background, background_attenuation, epsilon, (int)subset_order,
                          iterations);
}

int et_array_osem_ordered(float *activity_data, unsigned int size_x, unsigned int size_y, unsigned int n_subsets,
                          float *sinogram_data, int n_cameras, float firstcamera, float lastcamera,
                          unsigned int rotation_axis, int iterations, int use_attenuation, float *attenuation_data,
                          int use_psf, float *psf_data, unsigned int psf_size_x, unsigned int psf_size_y,
                          float background, float background_attenuation, float epsilon, int use_gpu)
{
    (void)use_gpu;
    if (n_cameras <= 0 || rotation_axis > 2 || n_subsets == 0) return niftyrec_error_parameters;
    if (use_psf && (psf_size_x % 2 != 1 || psf_size_x != psf_size_y)) return niftyrec_error_parameters;
    std::vector<float> cams((size_t)3 * n_cameras);
    et_array_make_cameras(cams.data(), firstcamera, lastcamera, n_cameras, rotation_axis);
    return em_driver_host(EM_OSEM_ORDERED, activity_data, (int)size_x, (int)size_y, sinogram_data, n_cameras,
                          cams.data(), use_attenuation ? attenuation_data : nullptr, use_psf ? psf_data : nullptr,
                          (int)psf_size_x, background, background_attenuation, epsilon, (int)n_subsets, iterations);
}

// ---- device-resident plans ------------------------------------------------------

int et_plan_create(et_plan **plan, int nx, int ny, int nz, const float *cameras, int n_cameras, int camera_cols,
                   const float *psf, int psf_h, float background, float background_attenuation)
{
    if (!plan) return niftyrec_error_parameters;
    *plan = nullptr;
    if (nx <= 0 || ny <= 0 || nz <= 0 || nx > ET_MAX_NX || n_cameras <= 0) return niftyrec_error_parameters;
    if (psf && (psf_h % 2 != 1 || nz != nx)) return niftyrec_error_parameters;
    int st = select_device();
    if (st) return st;
    return guarded([&]() -> int {
        const int vs[3] = {nx, ny, nz};
        Cameras cams;
        int rc = expand_cameras(cameras, n_cameras, camera_cols, vs, cams);
        if (rc) return rc;
        et_plan *p = new et_plan();
        (void)hipGetDevice(&p->device);
        try {
            rc = build_plan(p, nx, ny, nz, cams, psf, psf ? psf_h : 0, background, background_attenuation);
        } catch (...) {
            delete p;
            throw;
        }
        if (rc) {
            delete p;
            return rc;
        }
        p->yl_ops = p->yl;
        *plan = p;
        return niftyrec_success;
    });
}

int et_plan_destroy(et_plan *plan)
{
    if (!plan) return niftyrec_error_parameters;
    (void)hipSetDevice(plan->device);
    if (plan->st) (void)hipStreamSynchronize(plan->st);
    delete plan;
    return niftyrec_success;
}

#define ET_PLAN_CALL(plan, stream, body)                                        \
    if (!plan) return niftyrec_error_parameters;                                \
    if (hipSetDevice(plan->device) != hipSuccess) return niftyrec_error_nogpuinstalled; \
    plan->st = reinterpret_cast<hipStream_t>(stream);                           \
    return guarded([&]() -> int { body; return niftyrec_success; });

int et_plan_project(et_plan *plan, const float *act, const float *mu, float *sino, int truncate, void *stream)
{
    if (!act && !mu) return niftyrec_error_parameters;
    ET_PLAN_CALL(plan, stream, op_project(plan, act, mu, sino, truncate))
}

int et_plan_project_noalloc(et_plan *plan, const float *act, const float *mu, float *sino, int truncate,
                            void *stream)
{
    ET_PLAN_CALL(plan, stream, op_project_noalloc(plan, act, mu, sino, truncate))
}

int et_plan_backproject(et_plan *plan, const float *sino, const float *mu, float *out, int adjoint, int truncate,
                        void *stream)
{
    ET_PLAN_CALL(plan, stream, op_backproject(plan, sino, mu, out, adjoint != 0, truncate))
}

int et_plan_gradient_attenuation(et_plan *plan, const float *sino, const float *act, const float *mu, float *out,
                                 int exact, int truncate, void *stream)
{
    if (!act || !mu) return niftyrec_error_parameters;
    ET_PLAN_CALL(plan, stream, op_gradient(plan, sino, act, mu, out, exact != 0, truncate))
}

int et_plan_project_backward(et_plan *plan, const float *sino, const float *act, const float *mu, float *out_act,
                             float *out_mu, int truncate, void *stream)
{
    if (!sino || !act || !mu || !out_act || !out_mu) return niftyrec_error_parameters;
    const void *in[3] = {sino, act, mu};
    for (const void *q : in)
        if (q == out_act || q == out_mu) return niftyrec_error_parameters;
    if (out_act == out_mu) return niftyrec_error_parameters;
    ET_PLAN_CALL(plan, stream, op_backward(plan, sino, act, mu, out_act, out_mu, truncate))
}

int et_plan_osem(et_plan *plan, float *act, const float *sino, const float *mu, int subsets, int iterations,
                 float epsilon, void *stream)
{
    ET_PLAN_CALL(plan, stream, op_osem(plan, act, sino, mu, subsets, iterations, epsilon))
}

int et_plan_set_tuning(et_plan *plan, int depth_chunks, int view_batch_)
{
    if (!plan) return niftyrec_error_parameters;
    plan->tune_D = depth_chunks;
    plan->tune_vb = view_batch_;
    plan->osem_valid = false;  // cached OSEM data depend on the chunking
    return niftyrec_success;
}

int et_plan_get_device(const et_plan *plan, int *device)
{
    if (!plan || !device) return niftyrec_error_parameters;
    *device = plan->device;
    return niftyrec_success;
}

int et_plan_invalidate(et_plan *plan)
{
    if (!plan) return niftyrec_error_parameters;
    plan->osem_valid = false;
    plan->fac_valid = false;
    plan->fwd_valid = false;
    return niftyrec_success;
}

int et_plan_set_mu_check(et_plan *plan, int check)
{
    if (!plan) return niftyrec_error_parameters;
    plan->mu_check = check != 0;
    plan->osem_valid = false;  // the key changes meaning
    return niftyrec_success;
}

int et_plan_get_tuning(const et_plan *plan, int *depth_chunks_project, int *depth_chunks_backproject)
{
    if (!plan || !depth_chunks_project || !depth_chunks_backproject) return niftyrec_error_parameters;
    const bool cached = plan->osem_valid;
    *depth_chunks_project = cached ? plan->osem_Dp : 0;
    *depth_chunks_backproject = cached ? plan->osem_Db : 0;
    return niftyrec_success;
}

int et_plan_get_projector(const et_plan *plan, int *tiles)
{
    if (!plan || !tiles) return niftyrec_error_parameters;
    *tiles = plan->osem_valid && plan->osem_yl ? plan->yl_tiles : 0;
    return niftyrec_success;
}

int et_plan_profile(et_plan *plan, int enable)
{
    return et_plan_profile_select(plan, enable ? (1u << ET_PROF_COUNT) - 1u : 0u);
}

int et_plan_profile_select(et_plan *plan, unsigned int categories)
{
    if (!plan) return niftyrec_error_parameters;
    prof_collect(plan);
    plan->prof = categories & ((1u << ET_PROF_COUNT) - 1u);
    for (int i = 0; i < ET_PROF_COUNT; i++) {
        plan->prof_ms[i] = 0;
        plan->prof_bytes[i] = 0;
        plan->prof_n[i] = 0;
        plan->prof_seen[i] = 0;
    }
    return niftyrec_success;
}

int et_plan_profile_sample(et_plan *plan, int every)
{
    if (!plan || every < 1) return niftyrec_error_parameters;
    plan->prof_every = every;
    for (int i = 0; i < ET_PROF_COUNT; i++) plan->prof_seen[i] = 0;
    return niftyrec_success;
}

int et_plan_profile_read(et_plan *plan, int category, int *launches, double *total_ms, double *bytes)
{
    if (!plan || category < 0 || category >= ET_PROF_COUNT) return niftyrec_error_parameters;
    prof_collect(plan);
    if (launches) *launches = plan->prof_n[category];
    if (total_ms) *total_ms = plan->prof_ms[category];
    if (bytes) *bytes = plan->prof_bytes[category];
    return niftyrec_success;
}

int et_batch_create(et_batch **batch, et_plan *const *plans, int n_plans)
{
    if (!batch || !plans || n_plans <= 0) return niftyrec_error_parameters;
    *batch = nullptr;
    for (int i = 0; i < n_plans; i++)
        if (!plans[i] || !plans[i]->fast || plans[i]->device != plans[0]->device) return niftyrec_error_parameters;
    if (hipSetDevice(plans[0]->device) != hipSuccess) return niftyrec_error_nogpuinstalled;
    return guarded([&]() -> int {
        et_batch *b = new et_batch();
        b->device = plans[0]->device;
        b->plans.assign(plans, plans + n_plans);
        if (hipStreamCreateWithFlags(&b->own, hipStreamNonBlocking) != hipSuccess) {
            delete b;
            return niftyrec_error_kernel;
        }
        *batch = b;
        return niftyrec_success;
    });
}

int et_batch_destroy(et_batch *batch)
{
    if (!batch) return niftyrec_error_parameters;
    (void)hipSetDevice(batch->device);
    delete batch;
    return niftyrec_success;
}

int et_batch_osem(et_batch *batch, float *const *act, const float *const *sino, const float *const *mu, int subsets,
                  int iterations, float epsilon, int use_graph, void *stream)
{
    if (!batch || !act || !sino || iterations < 0) return niftyrec_error_parameters;
    for (size_t i = 0; i < batch->plans.size(); i++)
        if (!act[i] || !sino[i] || subsets <= 0 || subsets > batch->plans[i]->V) return niftyrec_error_parameters;
    if (hipSetDevice(batch->device) != hipSuccess) return niftyrec_error_nogpuinstalled;
    return guarded([&]() -> int {
        hipStream_t st = reinterpret_cast<hipStream_t>(stream);
        if (!st) {
            // the null stream cannot be captured: order after it, run on the batch's own stream, finish
            check_hip(hipStreamSynchronize(nullptr), niftyrec_error_kernel);
            st = batch->own;
        }
        op_batch_osem(batch, act, sino, mu, subsets, iterations, epsilon, use_graph != 0, st);
        if (st == batch->own) check_hip(hipStreamSynchronize(st), niftyrec_error_kernel);
        return niftyrec_success;
    });
}

}  // extern "C"
