// et_kernels.hip -- HIP kernels of the SPECT projector for gfx950 (CDNA4).
//
// Fast path (every view a rotation about the axial y axis, which is what
// one-column cameras produce, reference emission-lib/_et_array_interface.cpp:414-419):
// the sample position of rotated voxel (x', y, z') is (px(x',z'), y, pz(x',z')),
// so the bilinear weights and corner addresses are computed once per (x', z')
// and reused for a whole run of axial rows, which are contiguous in the YF
// layout.  One workgroup owns a (view, axial slab of ET_T rows, depth chunk) and
// sweeps depth z': resample -> depth-dependent separable PSF (column pass in
// registers, row pass through LDS) -> attenuation e^{-A} with A the running
// inclusive sum of the resampled mu -> accumulate.  Nothing rotated is ever
// stored for the projection.
//
// General path (any other rotation, a PSF wider than ET_RMAX, or a non-zero
// background, whose padding rule k_resample_general reproduces): the rotated
// volumes are materialised per view and swept with an identity view record.
//
// No MFMA: nothing here is a dense contraction.  Kernels never allocate.
#include <hip/hip_runtime.h>
#include <cstring>

#include "et_internal.h"

namespace {

// Keeps the scheduler from hoisting the loads of every 4-row chunk of a run
// ahead of the arithmetic (which multiplies the register footprint and
// halves occupancy); latency is hidden by the other waves of the SIMD instead.
#ifndef ET_NO_CHUNK_FENCE
#define ET_CHUNK_FENCE() __builtin_amdgcn_sched_barrier(0)
#else
#define ET_CHUNK_FENCE()
#endif

constexpr int T = ET_T;

// ------------------------------------------------------------------------------
// Arithmetic helpers
// ------------------------------------------------------------------------------

// Sample position with the reference's rounding: ((m0*x + m1*y) + m2*z) + m3,
// reg-lib/_reg_affineTransformation.cpp:42-49.  For fast views m1 == 0 so the
// y term is an exact +-0 and is dropped.
__device__ __forceinline__ float pos2(float m0, float x, float m2, float z, float m3)
{
#pragma clang fp contract(off)
    return (m0 * x + m2 * z) + m3;
}

__device__ __forceinline__ float pos3(const float *m, float x, float y, float z)
{
#pragma clang fp contract(off)
    return ((m[0] * x + m[1] * y) + m[2] * z) + m[3];
}

// e^{-A} for A >= 0.  -A log2(e) = x + r with x the rounded product and r its
// residual (exact fma residual plus the low part of log2(e)), |r| <= 2^-24 |x|;
// e^{-A} = 2^x 2^r = 2^x (1 + r ln 2) (the dropped term is below 1e-12
// relative for A < 60), with 2^x from
// the hardware exp2 (v_exp_f32, 1 ulp).  Six VALU operations.
__device__ __forceinline__ float exp_neg(float A)
{
    const float L2E = 1.44269502162933349609375f;   // float(log2 e)
    const float L2E_LO = 1.925963033500011e-08f;    // log2(e) - L2E
    const float LN2 = 0.693147180559945309f;
    const float x = -A * L2E;
    float r = fmaf(-A, L2E, -x);
    r = fmaf(-A, L2E_LO, r);
    return __builtin_amdgcn_exp2f(x) * fmaf(r, LN2, 1.f);
}

// Loads and stores through the global address space.  Pointers read from an
// argument record in memory (the grouped launches) are generic to the
// compiler, which would otherwise emit flat accesses.
// (The qualifiers apply to the device compilation; the host pass, which
// never runs these functions, sees plain pointers.)
#ifdef __HIP_DEVICE_COMPILE__
#define ET_GLOBAL_AS __attribute__((address_space(1)))
#define ET_CONST_AS __attribute__((address_space(4)))
#else
#define ET_GLOBAL_AS
#define ET_CONST_AS
#endif
typedef ET_GLOBAL_AS const float g_cfloat;
typedef ET_GLOBAL_AS float g_float;
typedef ET_GLOBAL_AS const float4 g_cfloat4;
typedef ET_GLOBAL_AS float4 g_float4;
__device__ __forceinline__ float gload(const float *p) { return *(g_cfloat *)p; }
__device__ __forceinline__ float4 gload4(const float *p) { return *(g_cfloat4 *)p; }
__device__ __forceinline__ void gstore(float *p, float v) { *(g_float *)p = v; }
__device__ __forceinline__ void gstore4(float *p, float4 v) { *(g_float4 *)p = v; }
// Read-only, block-uniform data (view records, PSF taps, argument records)
// read through the constant address space: scalar loads.
typedef ET_CONST_AS const float c_cfloat;
__device__ __forceinline__ float cload(const float *p) { return *(c_cfloat *)p; }

// Bilinear sample in the (x, z) plane shared by a run of axial rows.
struct Samp {
    int off;                        // element offset of corner (x0, z0) in a padded layout
    float w00, w01, w10, w11;       // weights of (x0,z0), (x0+1,z0), (x0,z0+1), (x0+1,z0+1)
    float xb0, xb1;                 // x basis (kept for the background rule)
    bool vx0, vx1, vz0, vz1;        // corner validity against the logical grid
};

// px, pz: sample position; strides dx, dz of the padded layout whose padded
// index of grid coordinate c is c + 1 (so corners one voxel outside read 0).
// Weight rule: reference reg-lib/_reg_resampling.cpp:345-364.
__device__ __forceinline__ Samp make_samp(float px, float pz, int nx, int nz, int dx, int dz)
{
    Samp s;
    float fx = floorf(px), fz = floorf(pz);
    float rx = px - fx, rz = pz - fz;
    s.xb0 = 1.f - rx;
    s.xb1 = rx;
    float zb0 = 1.f - rz, zb1 = rz;
    s.vx0 = (fx >= 0.f) && (fx <= (float)(nx - 1));
    s.vx1 = (fx >= -1.f) && (fx <= (float)(nx - 2));
    s.vz0 = (fz >= 0.f) && (fz <= (float)(nz - 1));
    s.vz1 = (fz >= -1.f) && (fz <= (float)(nz - 2));
    bool inside = (fx >= -1.f) && (fx <= (float)(nx - 1)) && (fz >= -1.f) && (fz <= (float)(nz - 1));
    int x0 = inside ? (int)fx : -1;
    int z0 = inside ? (int)fz : -1;
    float m = inside ? 1.f : 0.f;
    s.off = dx * (x0 + 1) + dz * (z0 + 1);
    s.w00 = m * (s.xb0 * zb0);
    s.w01 = m * (s.xb1 * zb0);
    s.w10 = m * (s.xb0 * zb1);
    s.w11 = m * (s.xb1 * zb1);
    return s;
}

// Bilinear samples of NA consecutive rows (starting SA rows into a LA-long
// aligned run at p00 of every corner) into out[NA], mapped through F(j, value).
// Corner-major: the whole run of one corner (one or two cache lines per lane)
// is loaded back to back and folded into out[] before the next corner, so a
// line is fetched from L2 once while it is hot in L1; the fence keeps the
// scheduler from hoisting the next corner's loads (register footprint).
template <int NA, int LA, int SA, class F>
__device__ __forceinline__ void bilinear_rows(const float *__restrict__ p00, int dx, int dz, const Samp &s, F f,
                                              float *out)
{
    const float *cp[4] = {p00, p00 + dx, p00 + dz, p00 + dz + dx};
    const float w[4] = {s.w00, s.w01, s.w10, s.w11};
#pragma unroll
    for (int c = 0; c < 4; c++) {
        float run[LA];
#pragma unroll
        for (int k = 0; k < LA / 4; k++) {
            if (4 * k + 3 < SA || 4 * k >= SA + NA) continue;
            const float4 t = gload4(cp[c] + 4 * k);
            run[4 * k] = t.x;
            run[4 * k + 1] = t.y;
            run[4 * k + 2] = t.z;
            run[4 * k + 3] = t.w;
        }
#pragma unroll
        for (int j = 0; j < NA; j++) out[j] = (c == 0) ? run[j + SA] * w[0] : fmaf(run[j + SA], w[c], out[j]);
        ET_CHUNK_FENCE();
    }
#pragma unroll
    for (int j = 0; j < NA; j++) out[j] = f(j, out[j]);
}

// Axial (column) pass of the PSF in registers:
//   yc[i] = sum_{j=0}^{2R} in[i+j] * kc[2R-j],  i < TS
// (convolution out[y] = sum_k in[y+k] kc[R-k], _et_convolveSeparable2D.cpp:28-39).
template <int TS, int R>
__device__ __forceinline__ void column_pass(const float *in, const float *__restrict__ kc, float *yc)
{
#pragma unroll
    for (int i = 0; i < TS; i++) {
        if constexpr (R == 0) {
            yc[i] = in[i];
        } else {
            float t = 0.f;
#pragma unroll
            for (int j = 0; j <= 2 * R; j++) t = fmaf(in[i + j], cload(kc + 2 * R - j), t);
            yc[i] = t;
        }
    }
}

// Resample NA rows and apply the column pass (see bilinear_rows / column_pass).
template <int TS, int R, int NA, int LA, int SA, class F>
__device__ __forceinline__ void resample_column(const float *__restrict__ p00, int dx, int dz, const Samp &s,
                                                const float *__restrict__ kc, F f, float *yc)
{
    float rows[NA];
    bilinear_rows<NA, LA, SA>(p00, dx, dz, s, f, rows);
    column_pass<TS, R>(rows, kc, yc);
}

// out[j] = bilinear sample of rows SKIP + j of a LOADED-long aligned run, j < NROWS,
// mapped through F(j, value) (corner-major, see bilinear_rows).
template <int NROWS, int LOADED, int SKIP, class F>
__device__ __forceinline__ void bilinear_map(const float *__restrict__ p00, int dx, int dz, const Samp &s, F f)
{
    float rows[NROWS];
    bilinear_rows<NROWS, LOADED, SKIP>(p00, dx, dz, s, [](int, float v) { return v; }, rows);
#pragma unroll
    for (int j = 0; j < NROWS; j++) f(j, rows[j]);
}

// Orders LDS accesses of the lanes of one wave (compiler and memory fence at
// wavefront scope; the LDS executes one wave's instructions in order).
__device__ __forceinline__ void wave_sync()
{
    __builtin_amdgcn_fence(__ATOMIC_RELEASE, "wavefront");
    __builtin_amdgcn_wave_barrier();
    __builtin_amdgcn_fence(__ATOMIC_ACQUIRE, "wavefront");
}

// Sample table of one wave for coop_rows (structure of arrays: offset bits,
// w00, w01, w10, w11; 64 entries each).
constexpr int ET_COOP_TAB = 5 * 64;

// Largest number of owners whose runs of LA rows fit a stage of `floats`
// floats (row stride LA + 4), rounded down to whole rounds of 64/(LA/4) owners.
__host__ __device__ constexpr int coop_owners(int LA, int floats)
{
    return (floats / (LA + 4)) >= 64 ? 64 : ((floats / (LA + 4)) / (64 / (LA / 4))) * (64 / (LA / 4));
}

// Wave-cooperative form of bilinear_rows; every lane of the wave must call it.
// A lane-private gather touches a different cache line in every lane of every
// load instruction, and the L1 processes those one lane at a time.  Here each
// lane publishes its sample (corner offset, four weights) in `tab`; then, round
// by round, the CH = LA/4 consecutive lanes serving one owner load the CH
// consecutive 16-byte chunks of that owner's four corner runs (coalesced),
// combine them with the owner's weights and store the owner's bilinear rows to
// `stage` (OPP owners at a time, row stride LA + 4), from which the owners read
// their runs back.  `stage` and `tab` belong to the calling wave alone.  A
// round's records are read after the previous round's stage write (they share
// LDS), which keeps one round of loads in flight: issuing every round's loads
// together was slower for both the projector and the pull.
template <int NA, int LA, int SA, int OPP, class F>
__device__ __forceinline__ void coop_rows(const float *__restrict__ base, int dx, int dz, const Samp &s,
                                          float *tab, float *stage, F f, float *out)
{
    constexpr int CH = LA / 4;
    constexpr int RPR = 64 / CH;
    constexpr int RS = LA + 4;
    constexpr int PH = (64 + OPP - 1) / OPP;
    constexpr int RPP = (OPP + RPR - 1) / RPR;
    static_assert(CH <= 64 && OPP >= 1 && OPP <= 64, "bad cooperative shape");
    const int lane = threadIdx.x & 63;
    tab[lane] = __int_as_float(s.off);
    tab[64 + lane] = s.w00;
    tab[128 + lane] = s.w01;
    tab[192 + lane] = s.w10;
    tab[256 + lane] = s.w11;
    wave_sync();
    const int k = lane % CH, tl = lane / CH;
    const bool serve = (tl < RPR) && (4 * k + 3 >= SA) && (4 * k < SA + NA);
    float run[LA];
#pragma unroll
    for (int ph = 0; ph < PH; ph++) {
        int off[RPP];
        float w[RPP][4];
#pragma unroll
        for (int r = 0; r < RPP; r++) {
            const int tt = r * RPR + tl;
            const int t = ph * OPP + tt;
            if (serve && tt < OPP && t < 64) {
                off[r] = __float_as_int(tab[t]);
                w[r][0] = tab[64 + t];
                w[r][1] = tab[128 + t];
                w[r][2] = tab[192 + t];
                w[r][3] = tab[256 + t];
                const float *q = base + off[r] + 4 * k;
                const float w00 = w[r][0], w01 = w[r][1], w10 = w[r][2], w11 = w[r][3];
                const float4 a = gload4(q);
                const float4 b = gload4(q + dx);
                const float4 c = gload4(q + dz);
                const float4 d = gload4(q + dz + dx);
                float4 v;
                v.x = a.x * w00 + b.x * w01 + c.x * w10 + d.x * w11;
                v.y = a.y * w00 + b.y * w01 + c.y * w10 + d.y * w11;
                v.z = a.z * w00 + b.z * w01 + c.z * w10 + d.z * w11;
                v.w = a.w * w00 + b.w * w01 + c.w * w10 + d.w * w11;
                *reinterpret_cast<float4 *>(stage + tt * RS + 4 * k) = v;
            }
        }
        wave_sync();
        if (lane / OPP == ph) {
            const float *mine = stage + (lane - ph * OPP) * RS;
#pragma unroll
            for (int qd = 0; qd < CH; qd++) {
                if (4 * qd + 3 < SA || 4 * qd >= SA + NA) continue;
                const float4 v = *reinterpret_cast<const float4 *>(mine + 4 * qd);
                run[4 * qd] = v.x;
                run[4 * qd + 1] = v.y;
                run[4 * qd + 2] = v.z;
                run[4 * qd + 3] = v.w;
            }
        }
        wave_sync();
    }
#pragma unroll
    for (int j = 0; j < NA; j++) out[j] = f(j, run[j + SA]);
}

// Row (x) pass through LDS for a block whose thread tx owns column x' = tx:
// out[i] = sum_j buf[tx + j][i] * kr[2R-j].  buf holds (nthr + 2R) columns of
// TS+4 floats with R zero columns on each side.  One barrier between the
// write and the reads; the caller keeps the next write from racing these reads
// (a block barrier before it, or a second buffer).
template <int TS, int R>
__device__ __forceinline__ void row_pass(float *buf, int tx, const float *yc, const float *__restrict__ kr,
                                         float *out)
{
    constexpr int L = TS + 4;
    float4 *w4 = reinterpret_cast<float4 *>(buf + (tx + R) * L);
#pragma unroll
    for (int q = 0; q < TS / 4; q++)
        w4[q] = make_float4(yc[4 * q], yc[4 * q + 1], yc[4 * q + 2], yc[4 * q + 3]);
    __syncthreads();
#pragma unroll
    for (int i = 0; i < TS; i++) out[i] = 0.f;
#pragma unroll
    for (int j = 0; j <= 2 * R; j++) {
        const float kj = cload(kr + 2 * R - j);
        const float4 *r4 = reinterpret_cast<const float4 *>(buf + (tx + j) * L);
#pragma unroll
        for (int q = 0; q < TS / 4; q++) {
            const float4 t = r4[q];
            out[4 * q + 0] = fmaf(t.x, kj, out[4 * q + 0]);
            out[4 * q + 1] = fmaf(t.y, kj, out[4 * q + 1]);
            out[4 * q + 2] = fmaf(t.z, kj, out[4 * q + 2]);
            out[4 * q + 3] = fmaf(t.w, kj, out[4 * q + 3]);
        }
    }
}

struct FastView {
    float m00, m02, m03, m20, m22, m23;
};

// Separated kernels of depth plane z (row kernel then column kernel); the PSF
// pointer is null when there is no PSF, so no address is formed for R == 0.
template <int R>
__device__ __forceinline__ const float *psf_plane(const float *psf, int z)
{
    if constexpr (R == 0)
        return nullptr;
    else
        return psf + (size_t)z * 2 * (2 * R + 1);
}

__device__ __forceinline__ FastView load_fast_view(const float *rec)
{
    FastView f;
    f.m00 = cload(rec + 0);
    f.m02 = cload(rec + 2);
    f.m03 = cload(rec + 3);
    f.m20 = cload(rec + 8);
    f.m22 = cload(rec + 10);
    f.m23 = cload(rec + 11);
    return f;
}

__device__ __forceinline__ Samp fast_sample(const FastView &fv, float fxp, int z, int nx, int nz, int dx, int dz)
{
    const float fz = (float)z;
    const float px = pos2(fv.m00, fxp, fv.m02, fz, fv.m03);
    const float pz = pos2(fv.m20, fxp, fv.m22, fz, fv.m23);
    return make_samp(px, pz, nx, nz, dx, dz);
}

// Sum over d < D of part[d * stride + i], added in order of d.  The loads go out
// eight at a time: a loop that adds each value as it arrives waits out one
// memory latency per term.
__device__ __forceinline__ float sum_parts(const float *part, int D, size_t stride, size_t i)
{
    float p = 0.f;
    int d = 0;
    for (; d + 8 <= D; d += 8) {
        float v[8];
#pragma unroll
        for (int k = 0; k < 8; k++) v[k] = gload(part + (size_t)(d + k) * stride + i);
#pragma unroll
        for (int k = 0; k < 8; k++) p += v[k];
    }
    for (; d < D; d++) p += gload(part + (size_t)d * stride + i);
    return p;
}

// Input projection value s(v, y, x) for the backprojection sweeps.
__device__ __forceinline__ float load_source(const EtSweepArgs &a, int v, int y, int x)
{
    const size_t P = (size_t)a.nx * a.ny;
    const size_t r = (size_t)y * a.nx + x;
    if (a.src == ET_SRC_ONES) return 1.f;
    if (a.src == ET_SRC_SINO) return gload(a.sino + (size_t)v * P + r);
    // ET_SRC_RATIO: reference _et_array_interface.cpp:1098-1108 (projection
    // truncated at 0, r = p < eps ? y/eps : y/p).
    float p = sum_parts(a.rpart, a.rD, (size_t)a.V * P, (size_t)v * P + r);
    if (p < 0.f) p = 0.f;
    const float m = gload(a.meas + (size_t)v * P + r);
    return (p < a.eps) ? m / a.eps : m / p;
}

// Stores TS consecutive rows of column (x', z') of view v in the U layout.
template <int TS>
__device__ __forceinline__ void store_u(const EtSweepArgs &a, int v, int tx, int z, int y0, const float *val)
{
    float *dst = a.u + (size_t)v * a.ustride + (size_t)a.yu * ((tx + 1) + (size_t)a.xp * (z + 1)) + y0;
#pragma unroll
    for (int q = 0; q < TS / 4; q++)
        gstore4(dst + 4 * q, make_float4(val[4 * q], val[4 * q + 1], val[4 * q + 2], val[4 * q + 3]));
}

// Coalesced form of store_u for the waves of a sweep block with a PSF: the
// wave's TS-row runs go through `stage` (its own slice of the row-pass buffer,
// row stride TS+4, idle until the next row-pass write) so that TS/4
// consecutive lanes store the aligned float4 chunks of one run.  Every lane of
// the wave calls it; columns x' >= nx store nothing.
template <int TS>
__device__ __forceinline__ void store_u_coop(const EtSweepArgs &a, int v, int xbase, int z, int y0, float *stage,
                                             const float *val)
{
    constexpr int L = TS + 4, CHS = TS / 4, CPR = 64 / CHS;
    const int lane = threadIdx.x & 63;
    float4 *w4 = reinterpret_cast<float4 *>(stage + lane * L);
#pragma unroll
    for (int q = 0; q < CHS; q++) w4[q] = make_float4(val[4 * q], val[4 * q + 1], val[4 * q + 2], val[4 * q + 3]);
    wave_sync();
    const int k = lane % CHS, tl = lane / CHS;
    float *uz = a.u + (size_t)v * a.ustride + (size_t)a.yu * a.xp * (z + 1) + y0 + 4 * k;
#pragma unroll
    for (int r = 0; r < CHS; r++) {
        const int t = r * CPR + tl;
        const int xq = xbase + t;
        if (xq < a.nx)
            gstore4(uz + (size_t)a.yu * (xq + 1), *reinterpret_cast<const float4 *>(stage + t * L + 4 * k));
    }
    wave_sync();
}

// Common geometry of a sweep block.
struct SweepCtx {
    int tx, nthr, v, d, y0, zb, ze, dx, dz, bufsz;
    bool xin;
    size_t P;
    float fxp;
    FastView fv;
};

// Logical block of a launch: linear id (the dispatcher deals linear ids to the
// 8 XCDs round-robin), coordinates and grid.  A single-problem launch uses its
// own blockIdx/gridDim; a grouped launch maps its blocks onto the problems'
// grids (each problem's block range starts at a multiple of 8, so lin & 7 is
// the hardware XCD in both cases).
struct Blk {
    int lin, x, y, gx, gy;
};

__device__ __forceinline__ Blk blk_here()
{
    Blk b;
    b.x = blockIdx.x;
    b.y = blockIdx.y;
    b.gx = gridDim.x;
    b.gy = gridDim.y;
    b.lin = b.x + b.gx * b.y;
    return b;
}

// Grouped launches: entry j of `tab` = {first block, gx, gy, 0}, first blocks
// ascending and multiples of 8.  Finds the problem of this block (binary
// search; uniform per block) and its logical block; false for the padding
// blocks at the end of a problem's range.
__device__ __forceinline__ bool group_block(const int4 *__restrict__ tab, int n, Blk &b, int &j)
{
    const int L = blockIdx.x;
    int lo = 0, hi = n - 1;
    while (lo < hi) {
        const int mid = (lo + hi + 1) >> 1;
        if (tab[mid].x <= L) lo = mid;
        else hi = mid - 1;
    }
    j = __builtin_amdgcn_readfirstlane(lo);   // uniform: keeps the argument record in SGPRs
    const int4 t = tab[j];
    const int local = L - t.x;
    if (local >= t.y * t.z) return false;
    b.lin = local;
    b.gx = t.y;
    b.gy = t.z;
    b.x = local % t.y;
    b.y = local / t.y;
    return true;
}

// XCD-aware block order.  Each XCD has its own L2.  For a grid of S fast x N
// slow blocks with N a multiple of 8, block lin is given fast index
// (lin/8) % S and slow index ((lin/8) / S) * 8 + lin % 8: each XCD then runs
// all S fast blocks of one slow index back to back, so the cache lines they
// share (neighbouring axial slabs, halo rows) are fetched once per XCD.
// Otherwise the launch order (fast0, slow0) is kept.  Returns (fast, slow).
__device__ __forceinline__ int2 xcd_order(const Blk &b, int S, int N, int fast0, int slow0)
{
    if ((N & 7) != 0) return make_int2(fast0, slow0);
    const int seq = b.lin >> 3;
    return make_int2(seq % S, (seq / S) * 8 + (b.lin & 7));
}

// Contiguous form: XCD k runs the slow indices [k N/8, (k+1) N/8) in order, each
// with its S fast blocks back to back, so neighbouring slow indices (which share
// cache lines) meet in one L2.
__device__ __forceinline__ int2 xcd_order_contig(const Blk &b, int S, int N, int fast0, int slow0)
{
    if ((N & 7) != 0) return make_int2(fast0, slow0);
    const int seq = b.lin >> 3;
    return make_int2(seq % S, (b.lin & 7) * (N >> 3) + seq / S);
}

template <int TS, int R>
__device__ __forceinline__ SweepCtx sweep_ctx(const EtSweepArgs &a, bool chunked, const Blk &b)
{
    SweepCtx c;
    c.tx = threadIdx.x;
    c.nthr = blockDim.x;
    const int2 fs = xcd_order(b, b.gx, b.gy, b.x, b.y);
    const int vd = fs.y;
    c.v = chunked ? vd / a.D : vd;
    c.d = chunked ? vd - c.v * a.D : 0;
    c.y0 = fs.x * TS;
    c.zb = chunked ? c.d * a.zlen : 0;
    c.ze = chunked ? min(a.nz, c.zb + a.zlen) : a.nz;
    c.xin = c.tx < a.nx;
    c.P = (size_t)a.nx * a.ny;
    c.bufsz = (c.nthr + 2 * R) * (TS + 4);
    c.fv = load_fast_view(a.views + (size_t)c.v * ET_VIEW_REC);
    c.dx = a.yp;
    c.dz = a.yp * a.xp;
    c.fxp = (float)c.tx;
    return c;
}

template <int R>
__device__ __forceinline__ void zero_lds(float *lds, int n, int tx, int nthr)
{
    if constexpr (R > 0) {
        for (int i = tx; i < n; i += nthr) lds[i] = 0.f;
        __syncthreads();
    }
}

// ------------------------------------------------------------------------------
// Sweep kernels (fast path)
// ------------------------------------------------------------------------------

// grid: (slabs, V*D); block: nthr >= nx threads, thread tx owns column x' = tx.
// PROJECT: partial projection of chunk d written to part[d][v][y][x]
// (the sinogram itself when D == 1).  Reference emission-lib/_et.cpp:218-281.
// AM (with HAS_MU): 0 attenuation from the resampled mu; 1 the same, and the
// factors e^{-A} of every plane stored in a.att; 2 the factors read from a.att
// (mu is not read).
template <int TS, int R, bool HAS_MU, bool COOP, int AM>
__device__ __forceinline__ void project_body(const EtSweepArgs &a, const Blk &b)
{
    constexpr int K = 2 * R + 1;
    constexpr int R4 = (R + 3) & ~3;
    constexpr int NA = TS + 2 * R, LA = TS + 2 * R4, SA = R4 - R;
    extern __shared__ float lds[];
    const SweepCtx c = sweep_ctx<TS, R>(a, true, b);
    zero_lds<R>(lds, c.bufsz, c.tx, c.nthr);
    // One row-pass buffer.  Each plane starts with a block barrier (the row
    // pass of the previous plane has been read by every wave); until its own
    // row-pass write, a wave stages cooperative rows in its own slice of the
    // buffer.  Without a PSF the stages follow the sample tables.
    constexpr int L = TS + 4;
    constexpr int OWN = 64 * L;
    constexpr int OPA = R > 0 ? coop_owners(LA, OWN) : 64, OPM = R > 0 ? coop_owners(TS, OWN) : 64;
    const int wv = c.tx >> 6;
    float *tab = lds + (R > 0 ? c.bufsz : 0) + wv * ET_COOP_TAB;
    float *stage = R > 0 ? lds + (64 * wv + R) * L : lds + (c.nthr >> 6) * ET_COOP_TAB + wv * 64 * (LA + 4);

    float A[TS], acc[TS];
#pragma unroll
    for (int i = 0; i < TS; i++) {
        A[i] = 0.f;
        acc[i] = 0.f;
    }
    if (HAS_MU && AM != 2 && c.d > 0 && c.xin) {
        const float *as = a.astart + ((size_t)c.v * a.D + c.d) * c.P;
#pragma unroll
        for (int i = 0; i < TS; i++) {
            const int y = c.y0 + i;
            A[i] = (y < a.ny) ? gload(as + (size_t)y * a.nx + c.tx) : 0.f;
        }
    }

    for (int z = c.zb; z < c.ze; z++) {
        const Samp s = fast_sample(c.fv, c.fxp, z, a.nx, a.nz, c.dx, c.dz);
        const float *kz = psf_plane<R>(a.psf, z);  // never read when R == 0
        if (R > 0 && z > c.zb) __syncthreads();
        float yc[TS];
        if constexpr (COOP) {
            float rows[NA];
            coop_rows<NA, LA, SA, OPA>(a.act + (a.gy + c.y0 - R4), c.dx, c.dz, s, tab, stage,
                                       [](int, float v) { return v; }, rows);
            column_pass<TS, R>(rows, kz + K, yc);
        } else {
            resample_column<TS, R, NA, LA, SA>(a.act + s.off + (a.gy + c.y0 - R4), c.dx, c.dz, s, kz + K,
                                               [](int, float v) { return v; }, yc);
        }
        if constexpr (HAS_MU && AM != 2) {
            // attenuation up to and including plane z (used after the row pass)
            auto cum = [&](int i, float m) {
                A[i] += m;
                return 0.f;
            };
            if constexpr (COOP) {
                float dummy[TS];
                coop_rows<TS, TS, 0, OPM>(a.mu + (a.gmu + c.y0), c.dx, c.dz, s, tab, stage, cum, dummy);
            } else {
                bilinear_map<TS, TS, 0>(a.mu + s.off + (a.gmu + c.y0), c.dx, c.dz, s,
                                        [&](int i, float m) { A[i] += m; });
            }
        }
        float h[TS];
        if constexpr (R == 0) {
#pragma unroll
            for (int i = 0; i < TS; i++) h[i] = yc[i];
        } else {
            if (!c.xin) {
#pragma unroll
                for (int i = 0; i < TS; i++) yc[i] = 0.f;
            }
            row_pass<TS, R>(lds, c.tx, yc, kz, h);
        }
        // factors of plane z in the padded layout att[v][z][yu][nthr]
        // (coalesced over x'; every lane stores or loads)
        float *ez = a.att + (((size_t)c.v * a.nz + z) * a.yu + c.y0) * c.nthr + c.tx;
        if constexpr (HAS_MU && AM == 1) {
#pragma unroll
            for (int i = 0; i < TS; i++) {
                const float e = exp_neg(A[i]);
                acc[i] = fmaf(h[i], e, acc[i]);
                gstore(ez + i * c.nthr, e);
            }
        } else if constexpr (HAS_MU && AM == 2) {
#pragma unroll
            for (int i = 0; i < TS; i++) acc[i] = fmaf(h[i], gload(ez + i * c.nthr), acc[i]);
        } else {
#pragma unroll
            for (int i = 0; i < TS; i++) acc[i] = HAS_MU ? fmaf(h[i], exp_neg(A[i]), acc[i]) : acc[i] + h[i];
        }
    }
    if (c.d == a.D - 1) {
        if constexpr (HAS_MU && AM == 2) {
            // e^{-A} along the whole ray = the factor of the last plane
            if (a.bg_line != 0.f) {
                const float *el = a.att + (((size_t)c.v * a.nz + a.nz - 1) * a.yu + c.y0) * c.nthr + c.tx;
#pragma unroll
                for (int i = 0; i < TS; i++) acc[i] += a.bg_line * el[i * c.nthr];
            }
        } else {
#pragma unroll
            for (int i = 0; i < TS; i++) acc[i] += HAS_MU ? a.bg_line * exp_neg(A[i]) : a.bg_line;
        }
    }
    if (c.xin) {
        float *out = a.part + ((size_t)c.d * a.V + c.v) * c.P;
#pragma unroll
        for (int i = 0; i < TS; i++) {
            const int y = c.y0 + i;
            if (y < a.ny) {
                float val = acc[i];
                if (a.truncate && val < 0.f) val = 0.f;
                gstore(out + (size_t)y * a.nx + c.tx, val);
            }
        }
    }
}

// BPU: u(x', y, z') = PSF_z'( s(x', y) e^{-A(x', y, z')} ) for the slab rows,
// reference emission-lib/_et.cpp:441-471 (line backprojection, then PSF).
// ATT: the attenuation factors come from a.att (written by the projection of
// the same views) instead of resampling mu.
template <int TS, int R, bool HAS_MU, bool COOP, bool ATT>
__device__ __forceinline__ void bpu_body(const EtSweepArgs &a, const Blk &b)
{
    constexpr int K = 2 * R + 1;
    constexpr int R4 = (R + 3) & ~3;
    constexpr int NA = TS + 2 * R, LA = TS + 2 * R4, SA = R4 - R;
    extern __shared__ float lds[];
    const SweepCtx c = sweep_ctx<TS, R>(a, true, b);
    zero_lds<R>(lds, c.bufsz, c.tx, c.nthr);
    // One row-pass buffer with cooperative staging in the wave's own slice of
    // it, as in k_project.
    constexpr int L = TS + 4;
    constexpr int OPA = R > 0 ? coop_owners(LA, 64 * L) : 64;
    const int wv = c.tx >> 6;
    float *tab = lds + (R > 0 ? c.bufsz : 0) + wv * ET_COOP_TAB;
    float *stage = R > 0 ? lds + (64 * wv + R) * L : lds + (c.nthr >> 6) * ET_COOP_TAB + wv * 64 * (LA + 4);

    float sv[NA], A[NA];
#pragma unroll
    for (int j = 0; j < NA; j++) {
        const int y = c.y0 - R + j;
        const bool ok = c.xin && y >= 0 && y < a.ny;
        sv[j] = ok ? load_source(a, c.v, y, c.tx) : 0.f;
        A[j] = (HAS_MU && !ATT && ok && c.d > 0) ? gload(a.astart + ((size_t)c.v * a.D + c.d) * c.P + (size_t)y * a.nx + c.tx)
                                                 : 0.f;
    }

    // with a PSF, plane z's output is stored during plane z+1, after its
    // barrier, through the wave's idle row-buffer slice (coalesced)
    float outp[TS];
    float *own = lds + (64 * wv + R) * L;
    const int xbase = c.tx & ~63;
    for (int z = c.zb; z < c.ze; z++) {
        const Samp s = fast_sample(c.fv, c.fxp, z, a.nx, a.nz, c.dx, c.dz);
        const float *kz = psf_plane<R>(a.psf, z);  // never read when R == 0
        if (R > 0 && z > c.zb) {
            __syncthreads();
            store_u_coop<TS>(a, c.v, xbase, z - 1, c.y0, own, outp);
        }
        float yc[TS];
        if constexpr (HAS_MU && ATT) {
            // rows outside the volume have s = 0: their factor is read from an
            // in-range row and not used
            const float *ez = a.att + ((size_t)c.v * a.nz + z) * a.yu * c.nthr + c.tx;
            float rows[NA];
#pragma unroll
            for (int j = 0; j < NA; j++) {
                const int yl = min(max(c.y0 - R + j, 0), a.ny - 1);
                rows[j] = sv[j] * gload(ez + yl * c.nthr);
            }
            column_pass<TS, R>(rows, kz + K, yc);
        } else if constexpr (HAS_MU) {
            auto wfun = [&](int j, float m) {
                A[j] += m;
                return sv[j] * exp_neg(A[j]);
            };
            if constexpr (COOP) {
                float rows[NA];
                coop_rows<NA, LA, SA, OPA>(a.mu + (a.gmu + c.y0 - R4), c.dx, c.dz, s, tab, stage, wfun, rows);
                column_pass<TS, R>(rows, kz + K, yc);
            } else {
                resample_column<TS, R, NA, LA, SA>(a.mu + s.off + (a.gmu + c.y0 - R4), c.dx, c.dz, s, kz + K, wfun,
                                                   yc);
            }
        } else if constexpr (R == 0) {
#pragma unroll
            for (int i = 0; i < TS; i++) yc[i] = sv[i];
        } else {
#pragma unroll
            for (int i = 0; i < TS; i++) {
                float t = 0.f;
#pragma unroll
                for (int j = 0; j <= 2 * R; j++) t = fmaf(sv[i + j], cload(kz + K + 2 * R - j), t);
                yc[i] = t;
            }
        }
        if constexpr (R == 0) {
            if (c.xin) store_u<TS>(a, c.v, c.tx, z, c.y0, yc);
        } else {
            row_pass<TS, R>(lds, c.tx, yc, kz, outp);
        }
    }
    if constexpr (R > 0) {
        if (c.ze > c.zb) {
            __syncthreads();
            store_u_coop<TS>(a, c.v, xbase, c.ze - 1, c.y0, own, outp);
        }
    }
}

// Reverse sweeps for the attenuation gradients (one block per view and slab,
// all of z').  A(z') = A_tot - sum_{k>z'} mu'(k), A_tot precomputed per ray.
// Kernels of the projector and rotated-frame backprojection sweeps: one
// problem (grid (slabs, views x chunks)), or a group of problems sharing the
// template parameters and block size (1-D grid over the padded block ranges).
template <int TS, int R, bool HAS_MU, bool COOP, int AM>
__global__ void __launch_bounds__(512) k_project(EtSweepArgs a)
{
    project_body<TS, R, HAS_MU, COOP, AM>(a, blk_here());
}

template <int TS, int R, bool HAS_MU, bool COOP, int AM>
__global__ void __launch_bounds__(512) k_project_group(const EtSweepArgs *__restrict__ args,
                                                       const int4 *__restrict__ tab, int n)
{
    Blk b;
    int j;
    if (!group_block(tab, n, b, j)) return;
    const EtSweepArgs a = ((ET_CONST_AS const EtSweepArgs *)args)[j];
    project_body<TS, R, HAS_MU, COOP, AM>(a, b);
}

template <int TS, int R, bool HAS_MU, bool COOP, bool ATT>
__global__ void __launch_bounds__(512) k_bpu(EtSweepArgs a)
{
    bpu_body<TS, R, HAS_MU, COOP, ATT>(a, blk_here());
}

template <int TS, int R, bool HAS_MU, bool COOP, bool ATT>
__global__ void __launch_bounds__(512) k_bpu_group(const EtSweepArgs *__restrict__ args,
                                                   const int4 *__restrict__ tab, int n)
{
    Blk b;
    int j;
    if (!group_block(tab, n, b, j)) return;
    const EtSweepArgs a = ((ET_CONST_AS const EtSweepArgs *)args)[j];
    bpu_body<TS, R, HAS_MU, COOP, ATT>(a, b);
}

// GRAD_REF: u = PSF( s * sum_{k>=z'} act'(k) e^{-A(k)} ), the reference kernel
//   emission-lib_gpu/_et_attenuation_gradient_gpu_kernels.cu:15-42 followed by the
//   PSF of emission-lib/_et.cpp:1829-1843.
template <int TS, int R>
__global__ void __launch_bounds__(512) k_grad_ref(EtSweepArgs a)
{
    constexpr int K = 2 * R + 1;
    constexpr int R4 = (R + 3) & ~3;
    constexpr int NA = TS + 2 * R, LA = TS + 2 * R4, SA = R4 - R;
    extern __shared__ float lds[];
    const SweepCtx c = sweep_ctx<TS, R>(a, false, blk_here());
    zero_lds<R>(lds, 2 * c.bufsz, c.tx, c.nthr);
    const int ys = c.y0 - R;

    float sv[NA], At[NA], suf[NA], tail[NA];
#pragma unroll
    for (int j = 0; j < NA; j++) {
        const int y = ys + j;
        const bool ok = c.xin && y >= 0 && y < a.ny;
        sv[j] = ok ? load_source(a, c.v, y, c.tx) : 0.f;
        At[j] = ok ? a.atot[(size_t)c.v * c.P + (size_t)y * a.nx + c.tx] : 0.f;
        suf[j] = 0.f;
        tail[j] = 0.f;
    }
    for (int z = a.nz - 1; z >= 0; z--) {
        const Samp s = fast_sample(c.fv, c.fxp, z, a.nx, a.nz, c.dx, c.dz);
        const float *kz = psf_plane<R>(a.psf, z);  // never read when R == 0
        // mu rows first: tail needs A(z') before this plane's mu joins the suffix
        float Az[NA];
        bilinear_map<NA, LA, SA>(a.mu + s.off + (a.gmu + c.y0 - R4), c.dx, c.dz, s, [&](int j, float m) {
            Az[j] = At[j] - suf[j];
            suf[j] += m;
        });
        float yc[TS];
        resample_column<TS, R, NA, LA, SA>(a.act + s.off + (a.gy + c.y0 - R4), c.dx, c.dz, s, kz + K,
                                           [&](int j, float v) {
                                               tail[j] = fmaf(v, exp_neg(Az[j]), tail[j]);
                                               return sv[j] * tail[j];
                                           },
                                           yc);
        float out[TS];
        if constexpr (R == 0) {
#pragma unroll
            for (int i = 0; i < TS; i++) out[i] = yc[i];
        } else {
            row_pass<TS, R>(lds + (z & 1) * c.bufsz, c.tx, yc, kz, out);
        }
        if (c.xin) store_u<TS>(a, c.v, c.tx, z, c.y0, out);
    }
}

// GRAD_MU: u = -s * ( sum_{k>=z'} PSF(act')(k) e^{-A(k)} + bg e^{-A_tot} ), the
// exact derivative of the projection with respect to mu'(z') (not in the reference).
template <int TS, int R>
__global__ void __launch_bounds__(512) k_grad_mu(EtSweepArgs a)
{
    constexpr int K = 2 * R + 1;
    constexpr int R4 = (R + 3) & ~3;
    constexpr int NA = TS + 2 * R, LA = TS + 2 * R4, SA = R4 - R;
    extern __shared__ float lds[];
    const SweepCtx c = sweep_ctx<TS, R>(a, false, blk_here());
    zero_lds<R>(lds, 2 * c.bufsz, c.tx, c.nthr);

    float sv[TS], At[TS], suf[TS], tail[TS];
#pragma unroll
    for (int i = 0; i < TS; i++) {
        const int y = c.y0 + i;
        const bool ok = c.xin && y < a.ny;
        sv[i] = ok ? load_source(a, c.v, y, c.tx) : 0.f;
        At[i] = ok ? a.atot[(size_t)c.v * c.P + (size_t)y * a.nx + c.tx] : 0.f;
        suf[i] = 0.f;
        tail[i] = a.bg_line * exp_neg(At[i]);
    }
    for (int z = a.nz - 1; z >= 0; z--) {
        const Samp s = fast_sample(c.fv, c.fxp, z, a.nx, a.nz, c.dx, c.dz);
        const float *kz = psf_plane<R>(a.psf, z);  // never read when R == 0
        float h[TS];
        {
            float yc[TS];
            resample_column<TS, R, NA, LA, SA>(a.act + s.off + (a.gy + c.y0 - R4), c.dx, c.dz, s, kz + K,
                                               [](int, float v) { return v; }, yc);
            if constexpr (R == 0) {
#pragma unroll
                for (int i = 0; i < TS; i++) h[i] = yc[i];
            } else {
                if (!c.xin) {
#pragma unroll
                    for (int i = 0; i < TS; i++) yc[i] = 0.f;
                }
                row_pass<TS, R>(lds + (z & 1) * c.bufsz, c.tx, yc, kz, h);
            }
        }
        float out[TS];
        bilinear_map<TS, TS, 0>(a.mu + s.off + (a.gmu + c.y0), c.dx, c.dz, s, [&](int i, float m) {
            const float Az = At[i] - suf[i];
            suf[i] += m;
            tail[i] = fmaf(h[i], exp_neg(Az), tail[i]);
            out[i] = -sv[i] * tail[i];
        });
        if (c.xin) store_u<TS>(a, c.v, c.tx, z, c.y0, out);
    }
}

// MUSUM: part[v][d][y][x] = sum over chunk d of mu'(x, y, z').
template <int TS>
__global__ void __launch_bounds__(512) k_musum(EtSweepArgs a)
{
    const SweepCtx c = sweep_ctx<TS, 0>(a, true, blk_here());
    if (!c.xin) return;
    float acc[TS];
#pragma unroll
    for (int i = 0; i < TS; i++) acc[i] = 0.f;
    for (int z = c.zb; z < c.ze; z++) {
        const Samp s = fast_sample(c.fv, c.fxp, z, a.nx, a.nz, c.dx, c.dz);
        bilinear_map<TS, TS, 0>(a.mu + s.off + (a.gmu + c.y0), c.dx, c.dz, s, [&](int i, float m) { acc[i] += m; });
    }
    float *out = a.part + ((size_t)c.v * a.D + c.d) * c.P;
#pragma unroll
    for (int i = 0; i < TS; i++) {
        const int y = c.y0 + i;
        if (y < a.ny) out[(size_t)y * a.nx + c.tx] = acc[i];
    }
}

// ------------------------------------------------------------------------------
// Pull kernels (fast path): accumulate the rotated-frame volumes of V views into
// the output grid.  Thread = output column (x, z) and a run of T axial rows.
// ------------------------------------------------------------------------------

// Reference pull: inverse rotation (negated angles, ZYX order), trilinear
// pull, accumulation in view order (emission-lib/_et.cpp:474-498).  The fast
// path runs with zero padding only; a non-zero background takes the general path.
__device__ __forceinline__ void pull_ref_view(const EtPullArgs &a, const float *rec, const float *uv,
                                              int x, int z, int y0, float *tab, float *stage, float *acc)
{
    const float qx = pos2(cload(rec + 12), (float)x, cload(rec + 14), (float)z, cload(rec + 15));
    const float qz = pos2(cload(rec + 20), (float)x, cload(rec + 22), (float)z, cload(rec + 23));
    const int dx = a.yu, dz = a.yu * a.xp;
    const Samp s = make_samp(qx, qz, a.nx, a.nz, dx, dz);
    float dummy[T];
    coop_rows<T, T, 0, 64>(uv + y0, dx, dz, s, tab, stage, [&](int i, float v) { acc[i] += v; return 0.f; }, dummy);
}

// Exact adjoint of the forward fast-path pull: gathers from every rotated-frame
// column (x', z') whose forward sample has (x, z) among its corners, with the
// forward weight.  Candidates are bounded through the inverse linear map;
// membership and weights are recomputed with the forward arithmetic, so this is
// the transpose of exactly what k_project / k_grad sample.
__device__ __forceinline__ void pull_adj_view(const EtPullArgs &a, const float *rec, const float *uv,
                                              int x, int z, int y0, float *acc)
{
    const float f00 = cload(rec + 0), f02 = cload(rec + 2), f03 = cload(rec + 3), f20 = cload(rec + 8),
                f22 = cload(rec + 10), f23 = cload(rec + 11);
    const float i00 = cload(rec + 24), i02 = cload(rec + 26), i20 = cload(rec + 30), i22 = cload(rec + 32);
    const float rx = (float)x - f03, rz = (float)z - f23;
    const float cx = i00 * rx + i02 * rz, cz = i20 * rx + i22 * rz;
    const float hx = fabsf(i00) + fabsf(i02) + 0.01f, hz = fabsf(i20) + fabsf(i22) + 0.01f;
    const int xlo = max(0, (int)ceilf(cx - hx)), xhi = min(a.nx - 1, (int)floorf(cx + hx));
    const int zlo = max(0, (int)ceilf(cz - hz)), zhi = min(a.nz - 1, (int)floorf(cz + hz));
    const float fxo = (float)x, fzo = (float)z;
    for (int xc = xlo; xc <= xhi; xc++) {
        for (int zc = zlo; zc <= zhi; zc++) {
            const float px = pos2(f00, (float)xc, f02, (float)zc, f03);
            const float pz = pos2(f20, (float)xc, f22, (float)zc, f23);
            const float fx = floorf(px), fz = floorf(pz);
            const float rxx = px - fx, rzz = pz - fz;
            float wx = (fx == fxo) ? 1.f - rxx : ((fx == fxo - 1.f) ? rxx : 0.f);
            float wz = (fz == fzo) ? 1.f - rzz : ((fz == fzo - 1.f) ? rzz : 0.f);
            const float w = wx * wz;
            if (w != 0.f) {
                const float4 *q = reinterpret_cast<const float4 *>(
                    uv + (size_t)a.yu * ((xc + 1) + (size_t)a.xp * (zc + 1)) + y0);
#pragma unroll
                for (int k = 0; k < T / 4; k++) {
                    float4 t = q[k];
                    acc[4 * k + 0] = fmaf(t.x, w, acc[4 * k + 0]);
                    acc[4 * k + 1] = fmaf(t.y, w, acc[4 * k + 1]);
                    acc[4 * k + 2] = fmaf(t.z, w, acc[4 * k + 2]);
                    acc[4 * k + 3] = fmaf(t.w, w, acc[4 * k + 3]);
                }
            }
        }
    }
}

// Coalesced YF epilogue of the reference pull: the wave's 64 columns of T
// rows go through its LDS stage so that 4 consecutive lanes handle the 4
// aligned float4 chunks of one column run (YF runs start at multiples of 16
// floats).  Padding rows (y >= ny) keep their content, so the zero border of
// the YF layout survives.  Every lane of the wave must call it.
template <int EPI>
__device__ __forceinline__ void yf_epilogue_coop(const EtPullArgs &a, const float *acc, float *stage,
                                                 int cbase, int y0)
{
    static_assert(T == 16, "epilogue assumes 16-row runs");
    constexpr int RS = T + 4;
    const int lane = threadIdx.x & 63;
#pragma unroll
    for (int q = 0; q < T / 4; q++)
        *reinterpret_cast<float4 *>(stage + lane * RS + 4 * q) =
            make_float4(acc[4 * q], acc[4 * q + 1], acc[4 * q + 2], acc[4 * q + 3]);
    wave_sync();
    const int k = lane & 3;
    const int ncol = a.nx * a.nz;
#pragma unroll
    for (int r = 0; r < 4; r++) {
        const int t = 16 * r + (lane >> 2);
        const int c = cbase + t;
        if (c < ncol) {
            const int x = c % a.nx, z = c / a.nx;
            const size_t base = (a.gy + y0 + 4 * k) + (size_t)a.yp * ((x + 1) + (size_t)(a.nx + 2) * (z + 1));
            const float4 b = *reinterpret_cast<const float4 *>(stage + t * RS + 4 * k);
            const float bv[4] = {b.x, b.y, b.z, b.w};
            float *dst = EPI == ET_EPI_OSEM ? a.x : a.out;
            float *d4 = dst + base;
            const float4 o = (EPI == ET_EPI_OSEM || a.accumulate) ? gload4(d4) : make_float4(0.f, 0.f, 0.f, 0.f);
            const float ov[4] = {o.x, o.y, o.z, o.w};
            float sv[4] = {1.f, 1.f, 1.f, 1.f};
            if constexpr (EPI == ET_EPI_OSEM) {
                const float4 s4 = gload4(a.sens + base);
                sv[0] = s4.x; sv[1] = s4.y; sv[2] = s4.z; sv[3] = s4.w;
            }
            float nv[4];
#pragma unroll
            for (int e = 0; e < 4; e++) {
                const bool live = y0 + 4 * k + e < a.ny;
                float val;
                if constexpr (EPI == ET_EPI_OSEM) {
                    const float bb = bv[e] < 0.f ? 0.f : bv[e];
                    val = ov[e] * (bb / sv[e]);
                } else {
                    val = a.accumulate ? ov[e] + bv[e] : bv[e];
                    if (a.truncate && val < 0.f) val = 0.f;
                }
                nv[e] = live ? val : ov[e];
            }
            gstore4(d4, make_float4(nv[0], nv[1], nv[2], nv[3]));
        }
    }
}

// Thread = output column (x, z) with T axial rows; the reference variant keeps
// every wave converged (columns past the grid sample a clamped column and
// store nothing) because its sampler is wave-cooperative.
template <bool ADJ, int EPI>
__device__ __forceinline__ void pull_body(const EtPullArgs &a, const Blk &b)
{
    constexpr int SCR = ET_COOP_TAB + 64 * (T + 4);
    __shared__ float scr[ADJ ? 4 : 4 * SCR];
    // fast index: axial slab (neighbouring slabs share cache lines of u); slow:
    // column blocks, contiguous per XCD (neighbouring blocks share the
    // bilinear neighbours of their samples)
    const int2 fs = xcd_order_contig(b, b.gy, b.gx, b.y, b.x);
    const int cbase = fs.y * blockDim.x;
    const int c0 = cbase + threadIdx.x;
    const bool cin = c0 < a.nx * a.nz;
    if (ADJ && !cin) return;
    const int c = cin ? c0 : a.nx * a.nz - 1;
    const int x = c % a.nx, z = c / a.nx;
    const int y0 = fs.x * T;
    float *tab = scr + (threadIdx.x >> 6) * SCR;
    float acc[T];
#pragma unroll
    for (int i = 0; i < T; i++) acc[i] = 0.f;
    for (int v = 0; v < a.V; v++) {
        const float *rec = a.views + (size_t)v * ET_VIEW_REC;
        const float *uv = a.u + (size_t)v * a.ustride;
        if (ADJ)
            pull_adj_view(a, rec, uv, x, z, y0, acc);
        else
            pull_ref_view(a, rec, uv, x, z, y0, tab, tab + ET_COOP_TAB, acc);
    }
    if (!cin && (ADJ || EPI == ET_EPI_ABI_STORE)) return;
    if constexpr (EPI == ET_EPI_ABI_STORE) {
#pragma unroll
        for (int i = 0; i < T; i++) {
            const int y = y0 + i;
            if (y < a.ny) {
                float *o = a.out + x + (size_t)a.nx * (y + (size_t)a.ny * z);
                float val = a.accumulate ? *o + acc[i] : acc[i];
                if (a.truncate && val < 0.f) val = 0.f;
                *o = val;
            }
        }
    } else if constexpr (!ADJ) {
        yf_epilogue_coop<EPI>(a, acc, tab + ET_COOP_TAB, cbase + (threadIdx.x & ~63), y0);
    } else if constexpr (EPI == ET_EPI_YF_STORE) {
        float *o = a.out + (a.gy + y0) + (size_t)a.yp * ((x + 1) + (size_t)(a.nx + 2) * (z + 1));
#pragma unroll
        for (int i = 0; i < T; i++) {
            const int y = y0 + i;
            if (y < a.ny) {
                float val = a.accumulate ? o[i] + acc[i] : acc[i];
                if (a.truncate && val < 0.f) val = 0.f;
                o[i] = val;
            }
        }
    } else {
        const size_t base = (a.gy + y0) + (size_t)a.yp * ((x + 1) + (size_t)(a.nx + 2) * (z + 1));
#pragma unroll
        for (int i = 0; i < T; i++) {
            const int y = y0 + i;
            if (y < a.ny) {
                float b = acc[i] < 0.f ? 0.f : acc[i];
                a.x[base + i] = a.x[base + i] * (b / a.sens[base + i]);
            }
        }
    }
}

// Pull kernels: one problem (grid (column blocks, axial slabs)) or a group of
// problems (1-D grid over the padded block ranges).
template <bool ADJ, int EPI>
__global__ void __launch_bounds__(256) k_pull(EtPullArgs a)
{
    pull_body<ADJ, EPI>(a, blk_here());
}

template <bool ADJ, int EPI>
__global__ void __launch_bounds__(256) k_pull_group(const EtPullArgs *__restrict__ args, const int4 *__restrict__ tab,
                                                    int n)
{
    Blk b;
    int j;
    if (!group_block(tab, n, b, j)) return;
    const EtPullArgs a = ((ET_CONST_AS const EtPullArgs *)args)[j];
    pull_body<ADJ, EPI>(a, b);
}

// ------------------------------------------------------------------------------
// General path
// ------------------------------------------------------------------------------

struct AbiAccess {
    const float *p;
    int nx, ny;
    __device__ float operator()(int X, int Y, int Z) const
    {
        return p[X + (size_t)nx * (Y + (size_t)ny * Z)];
    }
};
struct UAccess {
    const float *p;
    int yu, xp;
    __device__ float operator()(int X, int Y, int Z) const
    {
        return p[Y + (size_t)yu * ((X + 1) + (size_t)xp * (Z + 1))];
    }
};

// Trilinear pull with the reference's loop structure, borders and background
// rule, reg-lib/_reg_resampling.cpp:305-450.
template <class Acc>
__device__ float trilinear_ref(const Acc &src, int nx, int ny, int nz, float px, float py, float pz,
                               float bg)
{
    float p[3] = {px, py, pz};
    float pv[3], b0[3], b1[3];
#pragma unroll
    for (int i = 0; i < 3; i++) {
        pv[i] = floorf(p[i]);
        float rel = p[i] - pv[i];
        if (rel < 0.f) rel = 0.f;
        b0[i] = 1.f - rel;
        b1[i] = rel;
    }
    const float n[3] = {(float)nx, (float)ny, (float)nz};
    float intensity = 0.f, bgw = 0.f;
    for (int c = 0; c < 2; c++) {
        const float Z = pv[2] + c;
        if (Z >= 0.f && Z < n[2]) {
            float yt = 0.f;
            for (int b = 0; b < 2; b++) {
                const float Y = pv[1] + b;
                if (Y >= 0.f && Y < n[1]) {
                    float xt = 0.f;
                    for (int aa = 0; aa < 2; aa++) {
                        const float X = pv[0] + aa;
                        if (X >= 0.f && X < n[0])
                            xt += src((int)X, (int)Y, (int)Z) * (aa ? b1[0] : b0[0]);
                        else if (bg != 0.f)
                            bgw += bg * (aa ? b1[0] : b0[0]);
                    }
                    yt += xt * (b ? b1[1] : b0[1]);
                } else if (bg != 0.f)
                    bgw += bg * (b ? b1[0] : b0[0]);
            }
            intensity += yt * (c ? b1[2] : b0[2]);
        } else if (bg != 0.f)
            bgw += bg * (c ? b1[0] : b0[0]);
    }
    if (bg != 0.f) intensity += bgw;
    return intensity;
}

// dst(v) = src(M v) for every voxel v of the grid; dst in YF (dst_yf) or ABI layout.
__global__ void __launch_bounds__(256) k_resample_general(const float *__restrict__ src, int nx,
                                                          int ny, int nz, const float *__restrict__ M,
                                                          float bg, float *__restrict__ dst, int dst_yf,
                                                          int yp, int gy)
{
    const size_t i = (size_t)blockIdx.x * blockDim.x + threadIdx.x;
    const size_t n = (size_t)nx * ny * nz;
    if (i >= n) return;
    int x, y, z;
    if (dst_yf) {  // thread order follows the destination: y fastest
        y = (int)(i % ny);
        size_t r = i / ny;
        x = (int)(r % nx);
        z = (int)(r / nx);
    } else {
        x = (int)(i % nx);
        size_t r = i / nx;
        y = (int)(r % ny);
        z = (int)(r / ny);
    }
    const float fx = (float)x, fy = (float)y, fz = (float)z;
    const float px = pos3(M, fx, fy, fz), py = pos3(M + 4, fx, fy, fz), pz = pos3(M + 8, fx, fy, fz);
    AbiAccess acc{src, nx, ny};
    const float val = trilinear_ref(acc, nx, ny, nz, px, py, pz, bg);
    if (dst_yf)
        dst[(gy + y) + (size_t)yp * ((x + 1) + (size_t)(nx + 2) * (z + 1))] = val;
    else
        dst[x + (size_t)nx * (y + (size_t)ny * z)] = val;
}

// General pull of one view into the output (accumulating in place).
// Reference mode: inverse matrix (record[12..23]) with the reference borders.
// Adjoint mode: transpose of k_resample_general's pull with the forward matrix.
template <int EPI>
__global__ void __launch_bounds__(256) k_pull_general(EtPullArgs a)
{
    const size_t i = (size_t)blockIdx.x * blockDim.x + threadIdx.x;
    const size_t n = (size_t)a.nx * a.ny * a.nz;
    if (i >= n) return;
    const int x = (int)(i % a.nx);
    const size_t r = i / a.nx;
    const int y = (int)(r % a.ny);
    const int z = (int)(r / a.ny);
    const float *rec = a.views;
    const float *uv = a.u;
    float val = 0.f;
    if (!a.adjoint) {
        const float *Mi = rec + 12;
        const float fx = (float)x, fy = (float)y, fz = (float)z;
        UAccess acc{uv, a.yu, a.xp};
        val = trilinear_ref(acc, a.nx, a.ny, a.nz, pos3(Mi, fx, fy, fz), pos3(Mi + 4, fx, fy, fz),
                            pos3(Mi + 8, fx, fy, fz), a.bg);
    } else {
        const float *F = rec;
        const float *Fi = rec + 24;
        const float rv[3] = {(float)x - F[3], (float)y - F[7], (float)z - F[11]};
        int lo[3], hi[3];
        const int nn[3] = {a.nx, a.ny, a.nz};
#pragma unroll
        for (int k = 0; k < 3; k++) {
            const float c = Fi[3 * k] * rv[0] + Fi[3 * k + 1] * rv[1] + Fi[3 * k + 2] * rv[2];
            const float h = fabsf(Fi[3 * k]) + fabsf(Fi[3 * k + 1]) + fabsf(Fi[3 * k + 2]) + 0.01f;
            lo[k] = max(0, (int)ceilf(c - h));
            hi[k] = min(nn[k] - 1, (int)floorf(c + h));
        }
        const float o[3] = {(float)x, (float)y, (float)z};
        for (int zc = lo[2]; zc <= hi[2]; zc++)
            for (int yc = lo[1]; yc <= hi[1]; yc++)
                for (int xc = lo[0]; xc <= hi[0]; xc++) {
                    const float p[3] = {pos3(F, (float)xc, (float)yc, (float)zc),
                                        pos3(F + 4, (float)xc, (float)yc, (float)zc),
                                        pos3(F + 8, (float)xc, (float)yc, (float)zc)};
                    float w = 1.f;
#pragma unroll
                    for (int k = 0; k < 3; k++) {
                        const float f = floorf(p[k]), rr = p[k] - f;
                        w *= (f == o[k]) ? 1.f - rr : ((f == o[k] - 1.f) ? rr : 0.f);
                    }
                    if (w != 0.f) val = fmaf(w, uv[yc + (size_t)a.yu * ((xc + 1) + (size_t)a.xp * (zc + 1))], val);
                }
    }
    if constexpr (EPI == ET_EPI_ABI_STORE) {
        float *o = a.out + i;
        float res = a.accumulate ? *o + val : val;
        if (a.truncate && res < 0.f) res = 0.f;
        *o = res;
    } else {
        float *o = a.out + (a.gy + y) + (size_t)a.yp * ((x + 1) + (size_t)(a.nx + 2) * (z + 1));
        float res = a.accumulate ? *o + val : val;
        if (a.truncate && res < 0.f) res = 0.f;
        *o = res;
    }
}

// One axis of the depth-dependent separable convolution with a runtime radius
// (general path, PSF radius beyond ET_RMAX).  axis 0: row kernel along x,
// axis 1: column kernel along y.  flip: correlation (adjoint).
__global__ void __launch_bounds__(256) k_conv_axis(const float *__restrict__ src, float *__restrict__ dst,
                                                   int nx, int ny, int nz, long sx, long sy, long sz,
                                                   long off, int axis, const float *__restrict__ sep,
                                                   int K, int flip)
{
    const size_t i = (size_t)blockIdx.x * blockDim.x + threadIdx.x;
    const size_t n = (size_t)nx * ny * nz;
    if (i >= n) return;
    const int x = (int)(i % nx);
    const size_t r = i / nx;
    const int y = (int)(r % ny);
    const int z = (int)(r / ny);
    const int R = (K - 1) / 2;
    const float *ker = sep + (size_t)z * 2 * K + (axis ? K : 0);
    const long base = off + x * sx + y * sy + z * sz;
    const int pos = axis ? y : x, len = axis ? ny : nx;
    const long st = axis ? sy : sx;
    float s = 0.f;
    for (int k = -R; k <= R; k++) {
        const int d = pos + k;
        if (d >= 0 && d < len) s += src[base + (long)k * st] * ker[flip ? R + k : R - k];
    }
    dst[base] = s;
}

// ------------------------------------------------------------------------------
// Layout and elementwise kernels
// ------------------------------------------------------------------------------

// ABI [z][y][x] -> YF interior [z][x][y] through a 32x32 tile per z plane.
__global__ void __launch_bounds__(256) k_pack_yf(const float *__restrict__ abi, float *__restrict__ yf,
                                                 int nx, int ny, int nz, int yp, int gy)
{
    __shared__ float tile[32][33];
    const int z = blockIdx.z;
    const int xb = blockIdx.x * 32, yb = blockIdx.y * 32;
    const int tx = threadIdx.x & 31, ty = threadIdx.x >> 5;  // 32 x 8
    for (int k = ty; k < 32; k += 8) {
        const int x = xb + tx, y = yb + k;
        tile[k][tx] = (x < nx && y < ny) ? abi[x + (size_t)nx * (y + (size_t)ny * z)] : 0.f;
    }
    __syncthreads();
    for (int k = ty; k < 32; k += 8) {
        const int y = yb + tx, x = xb + k;
        if (x < nx && y < ny) yf[(gy + y) + (size_t)yp * ((x + 1) + (size_t)(nx + 2) * (z + 1))] = tile[tx][k];
    }
}

__global__ void __launch_bounds__(256) k_unpack_yf(const float *__restrict__ yf, float *__restrict__ abi,
                                                   int nx, int ny, int nz, int yp, int gy)
{
    __shared__ float tile[32][33];
    const int z = blockIdx.z;
    const int xb = blockIdx.x * 32, yb = blockIdx.y * 32;
    const int tx = threadIdx.x & 31, ty = threadIdx.x >> 5;
    for (int k = ty; k < 32; k += 8) {
        const int y = yb + tx, x = xb + k;
        tile[k][tx] = (x < nx && y < ny) ? yf[(gy + y) + (size_t)yp * ((x + 1) + (size_t)(nx + 2) * (z + 1))] : 0.f;
    }
    __syncthreads();
    for (int k = ty; k < 32; k += 8) {
        const int x = xb + tx, y = yb + k;
        if (x < nx && y < ny) abi[x + (size_t)nx * (y + (size_t)ny * z)] = tile[tx][k];
    }
}

// astart[v][d] = sum_{d' < d} csum[v][d'], atot[v] = sum_d csum[v][d].
__global__ void __launch_bounds__(256) k_chunk_scan(const float *__restrict__ csum, float *__restrict__ astart,
                                                    float *__restrict__ atot, int V, int D, int P)
{
    const size_t i = (size_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= (size_t)V * P) return;
    const int v = (int)(i / P);
    const size_t r = i % P;
    float run = 0.f;
    for (int d = 0; d < D; d++) {
        const size_t k = ((size_t)v * D + d) * P + r;
        if (astart) astart[k] = run;
        run += csum[k];
    }
    if (atot) atot[i] = run;
}

__global__ void __launch_bounds__(256) k_reduce_parts(const float *__restrict__ part, int D, size_t VP,
                                                      float *__restrict__ sino, int truncate)
{
    const size_t i = (size_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= VP) return;
    float s = sum_parts(part, D, VP, i);
    if (truncate && s < 0.f) s = 0.f;
    sino[i] = s;
}

__global__ void __launch_bounds__(256) k_bg_only(const float *__restrict__ atot, float bg, size_t n,
                                                 float *__restrict__ sino, int truncate)
{
    const size_t i = (size_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= n) return;
    float s = bg * exp_neg(atot[i]);
    if (truncate && s < 0.f) s = 0.f;
    sino[i] = s;
}

__global__ void __launch_bounds__(256) k_fill(float *buf, size_t n, float value)
{
    const size_t i = (size_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (i < n) buf[i] = value;
}

__global__ void __launch_bounds__(256) k_clamp_min(float *buf, size_t n, float lo)
{
    const size_t i = (size_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (i < n && buf[i] < lo) buf[i] = lo;
}

__global__ void __launch_bounds__(256) k_ratio(const float *__restrict__ meas, const float *__restrict__ proj,
                                               float *__restrict__ out, size_t n, float eps)
{
    const size_t i = (size_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= n) return;
    const float p = proj[i];
    out[i] = (p < eps) ? meas[i] / eps : meas[i] / p;
}

// YF interior only: f(i) for each logical voxel.
template <class F>
__global__ void __launch_bounds__(256) k_yf_map(int nx, int ny, int nz, int yp, int gy, F f)
{
    const size_t i = (size_t)blockIdx.x * blockDim.x + threadIdx.x;
    const size_t n = (size_t)nx * ny * nz;
    if (i >= n) return;
    const int y = (int)(i % ny);
    const size_t r = i / ny;
    const int x = (int)(r % nx);
    const int z = (int)(r / nx);
    f((gy + y) + (size_t)yp * ((x + 1) + (size_t)(nx + 2) * (z + 1)));
}

struct ClampMinOp {
    float *p;
    float lo;
    __device__ void operator()(size_t k) const
    {
        if (p[k] < lo) p[k] = lo;
    }
};

__global__ void __launch_bounds__(256) k_em_update_flat(float *__restrict__ x, const float *__restrict__ b,
                                                        const float *__restrict__ s, size_t n)
{
    const size_t i = (size_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= n) return;
    const float bb = b[i] < 0.f ? 0.f : b[i];
    x[i] = x[i] * (bb / s[i]);
}

// dst[i] = src[order[i]] for planes of P floats (views reordered by subset).
__global__ void __launch_bounds__(256) k_gather_planes(float *__restrict__ dst, const float *__restrict__ src,
                                                      const int *__restrict__ order, int V, size_t P)
{
    const size_t i = (size_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= (size_t)V * P) return;
    const int v = (int)(i / P);
    dst[i] = src[(size_t)order[v] * P + (i % P)];
}

// OSEM ratio from partial projections, reference _et_array_interface.cpp:1098-1108:
// p = max(sum_d part[d], 0); r = p < eps ? y / eps : y / p.
// Partial line integrals of one view in the rotated frame (reference
// emission-lib_gpu/_et_line_integral_attenuated_gpu_kernels.cu:57-103): along
// each ray from the far end, part[z] = S, then S = (S + act[z]) e^{-mu[z]}
// (act or mu may be absent), S starting at bg (times bgimg); sino = final S.
// ABI layouts; one thread per detector pixel (coalesced over x).
__global__ void __launch_bounds__(256) k_partial_line(const float *__restrict__ act, const float *__restrict__ mu,
                                                     const float *__restrict__ bgimg, float bg, int nx, int ny,
                                                     int nz, float *__restrict__ sino, float *__restrict__ part)
{
    const size_t P = (size_t)nx * ny;
    const size_t r = (size_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (r >= P) return;
    float s = bg;
    if (bgimg) s = s * bgimg[r];
    for (int z = nz - 1; z >= 0; z--) {
        const size_t i = r + P * z;
        part[i] = s;
        if (act && mu)
            s = (s + act[i]) * exp_neg(mu[i]);
        else if (act)
            s = s + act[i];
        else
            s = s * exp_neg(mu[i]);
    }
    sino[r] = s;
}

// Grouped MLEM ratio over several problems (1-D grid; tab entry j =
// {first block, blocks, 1, 0}).
__global__ void __launch_bounds__(256) k_ratio_parts_group(const EtRatioArgs *__restrict__ args,
                                                          const int4 *__restrict__ tab, int n)
{
    Blk b;
    int j;
    if (!group_block(tab, n, b, j)) return;
    const EtRatioArgs a = args[j];
    const size_t i = (size_t)b.lin * blockDim.x + threadIdx.x;
    if (i >= (size_t)a.VP) return;
    float p = sum_parts(a.part, a.D, (size_t)a.VP, i);
    if (p < 0.f) p = 0.f;
    a.out[i] = (p < a.eps) ? a.meas[i] / a.eps : a.meas[i] / p;
}

__global__ void __launch_bounds__(256) k_ratio_parts(const float *__restrict__ meas, const float *__restrict__ part,
                                                    int D, size_t VP, float eps, float *__restrict__ out)
{
    const size_t i = (size_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= VP) return;
    float p = sum_parts(part, D, VP, i);
    if (p < 0.f) p = 0.f;
    out[i] = (p < eps) ? meas[i] / eps : meas[i] / p;
}

inline unsigned nblk(size_t n, unsigned b) { return (unsigned)((n + b - 1) / b); }

}  // namespace

// ------------------------------------------------------------------------------
// Launchers
// ------------------------------------------------------------------------------

// Slab height (axial rows per block) of each sweep kernel: taller slabs
// amortise the PSF halo, shorter ones keep the register footprint down.
static int slab_rows(int mode, int R)
{
    switch (mode) {
    case ET_MODE_PROJECT: return 16;
    case ET_MODE_BPU: return R <= 7 ? 16 : 8;
    case ET_MODE_GRAD_REF: return R <= 2 ? 8 : 4;
    case ET_MODE_GRAD_MU: return R <= 6 ? 16 : 8;
    default: return 16;
    }
}

int et_sweep_slab_rows(int mode, int R) { return slab_rows(mode, R); }

// LDS of a sweep launch: row-pass buffers (one for the projector and the
// rotated-frame backprojection, two for the gradient sweeps) and, for the
// cooperative variants, the per-wave sample tables (plus stages without a PSF).
static size_t sweep_lds_bytes(int R, int TS, int nthr, bool coop, int nbuf)
{
    const int R4 = (R + 3) & ~3;
    size_t f = R > 0 ? (size_t)nbuf * (nthr + 2 * R) * (TS + 4) : 0;
    if (coop) f += (size_t)(nthr / 64) * (ET_COOP_TAB + (R > 0 ? 0 : 64 * (TS + 2 * R4 + 4)));
    return f * sizeof(float);
}

static size_t lds_limit()
{
    static size_t lim = 0;
    if (!lim) {
        int dev = 0, v = 0;
        if (hipGetDevice(&dev) == hipSuccess &&
            hipDeviceGetAttribute(&v, hipDeviceAttributeMaxSharedMemoryPerBlock, dev) == hipSuccess && v > 0)
            lim = (size_t)v;
        else
            lim = 65536;
    }
    return lim;
}

template <class Kern>
static void allow_lds(Kern k, size_t bytes)
{
    if (bytes > 65536) (void)hipFuncSetAttribute(reinterpret_cast<const void *>(k),
                                                 hipFuncAttributeMaxDynamicSharedMemorySize, (int)bytes);
}

// Grouped launch description (device tables; see group_block).
struct GroupLaunch {
    const EtSweepArgs *args;
    const int4 *tab;
    int n, blocks, nthr;
};

template <int TS, int R>
static int launch_sweep_ts(int mode, int has_mu, const EtSweepArgs &a, hipStream_t st,
                           const GroupLaunch *grp = nullptr)
{
    const int nthr = grp ? grp->nthr : ((a.nx + 63) / 64) * 64;
    const int slabs = (a.ny + TS - 1) / TS;
    const bool single = mode == ET_MODE_PROJECT || mode == ET_MODE_BPU;
    const bool coop = sweep_lds_bytes(R, TS, nthr, true, 1) <= lds_limit();
    const bool uses = mode == ET_MODE_PROJECT || (mode == ET_MODE_BPU && has_mu == 1);
    const size_t lds = sweep_lds_bytes(R, TS, nthr, coop && uses, single ? 1 : 2);
#define ET_LP(mu, cp, at)                                                                                  \
    {                                                                                                      \
        if (grp) {                                                                                         \
            allow_lds(k_project_group<TS, R, mu, cp, at>, lds);                                            \
            hipLaunchKernelGGL((k_project_group<TS, R, mu, cp, at>), dim3(grp->blocks), dim3(nthr), lds, st, \
                               grp->args, grp->tab, grp->n);                                               \
        } else {                                                                                           \
            allow_lds(k_project<TS, R, mu, cp, at>, lds);                                                  \
            hipLaunchKernelGGL((k_project<TS, R, mu, cp, at>), g, dim3(nthr), lds, st, a);                 \
        }                                                                                                  \
    }
#define ET_LB(mu, cp, at)                                                                                  \
    {                                                                                                      \
        if (grp) {                                                                                         \
            allow_lds(k_bpu_group<TS, R, mu, cp, at>, lds);                                                \
            hipLaunchKernelGGL((k_bpu_group<TS, R, mu, cp, at>), dim3(grp->blocks), dim3(nthr), lds, st,     \
                               grp->args, grp->tab, grp->n);                                               \
        } else {                                                                                           \
            allow_lds(k_bpu<TS, R, mu, cp, at>, lds);                                                      \
            hipLaunchKernelGGL((k_bpu<TS, R, mu, cp, at>), g, dim3(nthr), lds, st, a);                     \
        }                                                                                                  \
    }
    // only the slab heights slab_rows() selects are instantiated per mode
    if (mode == ET_MODE_PROJECT) {
        if constexpr (TS == 16) {
            dim3 g(slabs, a.V * a.D);
            if (has_mu == 3) {
                if (coop) ET_LP(true, true, 2) else ET_LP(true, false, 2)
            } else if (has_mu == 2) {
                if (coop) ET_LP(true, true, 1) else ET_LP(true, false, 1)
            } else if (has_mu) {
                if (coop) ET_LP(true, true, 0) else ET_LP(true, false, 0)
            } else {
                if (coop) ET_LP(false, true, 0) else ET_LP(false, false, 0)
            }
        } else {
            return (int)hipErrorInvalidValue;
        }
    } else if (mode == ET_MODE_BPU) {
        if constexpr (TS >= 8) {
            dim3 g(slabs, a.V * a.D);
            if (has_mu == 3) {
                ET_LB(true, false, true)
            } else if (has_mu) {
                if (coop) ET_LB(true, true, false) else ET_LB(true, false, false)
            } else {
                ET_LB(false, false, false)
            }
        } else {
            return (int)hipErrorInvalidValue;
        }
    } else if (grp) {
        return (int)hipErrorInvalidValue;
    } else if (mode == ET_MODE_GRAD_REF) {
        hipLaunchKernelGGL((k_grad_ref<TS, R>), dim3(slabs, a.V), dim3(nthr), lds, st, a);
    } else if (mode == ET_MODE_GRAD_MU) {
        hipLaunchKernelGGL((k_grad_mu<TS, R>), dim3(slabs, a.V), dim3(nthr), lds, st, a);
    } else {
        return (int)hipErrorInvalidValue;
    }
#undef ET_LP
#undef ET_LB
    return (int)hipGetLastError();
}

template <int R>
static int launch_sweep_r(int mode, int has_mu, const EtSweepArgs &a, hipStream_t st,
                          const GroupLaunch *grp = nullptr)
{
    switch (slab_rows(mode, R)) {
    case 16: return launch_sweep_ts<16, R>(mode, has_mu, a, st, grp);
    case 8: return launch_sweep_ts<8, R>(mode, has_mu, a, st, grp);
    default: return launch_sweep_ts<4, R>(mode, has_mu, a, st, grp);
    }
}

int et_launch_sweep_group(int mode, int R, int has_mu, int nthr, const EtSweepArgs *args, const void *tab, int n,
                          int blocks, et_stream_t stream)
{
    if (mode != ET_MODE_PROJECT && mode != ET_MODE_BPU) return (int)hipErrorInvalidValue;
    if (n <= 0 || blocks <= 0 || nthr <= 0 || nthr > 512 || (nthr & 63)) return (int)hipErrorInvalidValue;
    hipStream_t st = reinterpret_cast<hipStream_t>(stream);
    GroupLaunch g{args, reinterpret_cast<const int4 *>(tab), n, blocks, nthr};
    EtSweepArgs dummy;
    memset(&dummy, 0, sizeof(dummy));
    switch (R) {
#define ET_R_CASE(r) \
    case r: return launch_sweep_r<r>(mode, has_mu, dummy, st, &g);
        ET_R_CASE(0) ET_R_CASE(1) ET_R_CASE(2) ET_R_CASE(3) ET_R_CASE(4) ET_R_CASE(5)
        ET_R_CASE(6) ET_R_CASE(7) ET_R_CASE(8) ET_R_CASE(9) ET_R_CASE(10)
#undef ET_R_CASE
    default: return (int)hipErrorInvalidValue;
    }
}

int et_launch_pull_group(int epilogue, const EtPullArgs *args, const void *tab, int n, int blocks,
                         et_stream_t stream)
{
    if (n <= 0 || blocks <= 0) return (int)hipErrorInvalidValue;
    hipStream_t st = reinterpret_cast<hipStream_t>(stream);
    const int4 *t = reinterpret_cast<const int4 *>(tab);
    if (epilogue == ET_EPI_OSEM)
        hipLaunchKernelGGL((k_pull_group<false, ET_EPI_OSEM>), dim3(blocks), dim3(256), 0, st, args, t, n);
    else if (epilogue == ET_EPI_YF_STORE)
        hipLaunchKernelGGL((k_pull_group<false, ET_EPI_YF_STORE>), dim3(blocks), dim3(256), 0, st, args, t, n);
    else
        return (int)hipErrorInvalidValue;
    return (int)hipGetLastError();
}

int et_launch_partial_line(const float *act, const float *mu, const float *bgimg, float bg, int nx, int ny, int nz,
                           float *sino, float *part, et_stream_t stream)
{
    if (!act && !mu) return (int)hipErrorInvalidValue;
    const size_t P = (size_t)nx * ny;
    hipLaunchKernelGGL(k_partial_line, dim3(nblk(P, 256)), dim3(256), 0, reinterpret_cast<hipStream_t>(stream), act,
                       mu, bgimg, bg, nx, ny, nz, sino, part);
    return (int)hipGetLastError();
}

int et_launch_ratio_group(const EtRatioArgs *args, const void *tab, int n, int blocks, et_stream_t stream)
{
    if (n <= 0 || blocks <= 0) return (int)hipErrorInvalidValue;
    hipLaunchKernelGGL(k_ratio_parts_group, dim3(blocks), dim3(256), 0, reinterpret_cast<hipStream_t>(stream), args,
                       reinterpret_cast<const int4 *>(tab), n);
    return (int)hipGetLastError();
}

int et_launch_sweep(int mode, int R, int has_mu, int has_act, const EtSweepArgs &a, et_stream_t stream)
{
    (void)has_act;
    hipStream_t st = reinterpret_cast<hipStream_t>(stream);
    if (mode == ET_MODE_MUSUM) {
        const int nthr = ((a.nx + 63) / 64) * 64;
        hipLaunchKernelGGL(k_musum<16>, dim3((a.ny + 15) / 16, a.V * a.D), dim3(nthr), 0, st, a);
        return (int)hipGetLastError();
    }
    switch (R) {
#define ET_R_CASE(r) \
    case r: return launch_sweep_r<r>(mode, has_mu, a, st);
        ET_R_CASE(0) ET_R_CASE(1) ET_R_CASE(2) ET_R_CASE(3) ET_R_CASE(4) ET_R_CASE(5)
        ET_R_CASE(6) ET_R_CASE(7) ET_R_CASE(8) ET_R_CASE(9) ET_R_CASE(10)
#undef ET_R_CASE
    default: return (int)hipErrorInvalidValue;
    }
}


int et_launch_pull(const EtPullArgs &a, et_stream_t stream)
{
    hipStream_t st = reinterpret_cast<hipStream_t>(stream);
    dim3 g(nblk((size_t)a.nx * a.nz, 256), (a.ny + T - 1) / T);
#define ET_PULL(adj, epi) hipLaunchKernelGGL((k_pull<adj, epi>), g, dim3(256), 0, st, a)
    if (a.adjoint) {
        if (a.epilogue == ET_EPI_ABI_STORE) ET_PULL(true, ET_EPI_ABI_STORE);
        else if (a.epilogue == ET_EPI_YF_STORE) ET_PULL(true, ET_EPI_YF_STORE);
        else ET_PULL(true, ET_EPI_OSEM);
    } else {
        if (a.epilogue == ET_EPI_ABI_STORE) ET_PULL(false, ET_EPI_ABI_STORE);
        else if (a.epilogue == ET_EPI_YF_STORE) ET_PULL(false, ET_EPI_YF_STORE);
        else ET_PULL(false, ET_EPI_OSEM);
    }
#undef ET_PULL
    return (int)hipGetLastError();
}

int et_launch_pull_general(const EtPullArgs &a, et_stream_t stream)
{
    hipStream_t st = reinterpret_cast<hipStream_t>(stream);
    const size_t n = (size_t)a.nx * a.ny * a.nz;
    if (a.epilogue == ET_EPI_ABI_STORE)
        hipLaunchKernelGGL((k_pull_general<ET_EPI_ABI_STORE>), dim3(nblk(n, 256)), dim3(256), 0, st, a);
    else if (a.epilogue == ET_EPI_YF_STORE)
        hipLaunchKernelGGL((k_pull_general<ET_EPI_YF_STORE>), dim3(nblk(n, 256)), dim3(256), 0, st, a);
    else
        return (int)hipErrorInvalidValue;
    return (int)hipGetLastError();
}

int et_launch_pack_yf(const float *abi, float *yf, int nx, int ny, int nz, int yp, int gy, et_stream_t stream)
{
    dim3 g((nx + 31) / 32, (ny + 31) / 32, nz);
    hipLaunchKernelGGL(k_pack_yf, g, dim3(256), 0, reinterpret_cast<hipStream_t>(stream), abi, yf, nx, ny, nz, yp, gy);
    return (int)hipGetLastError();
}

int et_launch_unpack_yf(const float *yf, float *abi, int nx, int ny, int nz, int yp, int gy, et_stream_t stream)
{
    dim3 g((nx + 31) / 32, (ny + 31) / 32, nz);
    hipLaunchKernelGGL(k_unpack_yf, g, dim3(256), 0, reinterpret_cast<hipStream_t>(stream), yf, abi, nx, ny, nz, yp, gy);
    return (int)hipGetLastError();
}

int et_launch_resample_general(const float *src, int nx, int ny, int nz, const float *M12, float bg, float *dst,
                               int dst_yf, int yp, int gy, et_stream_t stream)
{
    const size_t n = (size_t)nx * ny * nz;
    hipLaunchKernelGGL(k_resample_general, dim3(nblk(n, 256)), dim3(256), 0, reinterpret_cast<hipStream_t>(stream),
                       src, nx, ny, nz, M12, bg, dst, dst_yf, yp, gy);
    return (int)hipGetLastError();
}

int et_launch_chunk_scan(const float *csum, float *astart, float *atot, int V, int D, int P, et_stream_t stream)
{
    hipLaunchKernelGGL(k_chunk_scan, dim3(nblk((size_t)V * P, 256)), dim3(256), 0,
                       reinterpret_cast<hipStream_t>(stream), csum, astart, atot, V, D, P);
    return (int)hipGetLastError();
}

int et_launch_reduce_parts(const float *part, int D, int V, int P, float *sino, int truncate, et_stream_t stream)
{
    const size_t VP = (size_t)V * P;
    hipLaunchKernelGGL(k_reduce_parts, dim3(nblk(VP, 256)), dim3(256), 0, reinterpret_cast<hipStream_t>(stream),
                       part, D, VP, sino, truncate);
    return (int)hipGetLastError();
}

int et_launch_bg_only(const float *atot, float bg, int n, float *sino, int truncate, et_stream_t stream)
{
    hipLaunchKernelGGL(k_bg_only, dim3(nblk((size_t)n, 256)), dim3(256), 0, reinterpret_cast<hipStream_t>(stream),
                       atot, bg, (size_t)n, sino, truncate);
    return (int)hipGetLastError();
}

int et_launch_conv_axis(const float *src, float *dst, int nx, int ny, int nz, long sx, long sy, long sz, long off,
                        int axis, const float *sep, int K, int flip, et_stream_t stream)
{
    const size_t n = (size_t)nx * ny * nz;
    hipLaunchKernelGGL(k_conv_axis, dim3(nblk(n, 256)), dim3(256), 0, reinterpret_cast<hipStream_t>(stream), src,
                       dst, nx, ny, nz, sx, sy, sz, off, axis, sep, K, flip);
    return (int)hipGetLastError();
}

int et_launch_truncate(float *buf, size_t n, et_stream_t stream) { return et_launch_clamp_min(buf, n, 0.f, stream); }

int et_launch_fill(float *buf, size_t n, float value, et_stream_t stream)
{
    hipLaunchKernelGGL(k_fill, dim3(nblk(n, 256)), dim3(256), 0, reinterpret_cast<hipStream_t>(stream), buf, n, value);
    return (int)hipGetLastError();
}

int et_launch_clamp_min(float *buf, size_t n, float lo, et_stream_t stream)
{
    hipLaunchKernelGGL(k_clamp_min, dim3(nblk(n, 256)), dim3(256), 0, reinterpret_cast<hipStream_t>(stream), buf, n, lo);
    return (int)hipGetLastError();
}

int et_launch_yf_clamp_min(float *yf, int nx, int ny, int nz, int yp, int gy, float lo, et_stream_t stream)
{
    const size_t n = (size_t)nx * ny * nz;
    hipLaunchKernelGGL((k_yf_map<ClampMinOp>), dim3(nblk(n, 256)), dim3(256), 0, reinterpret_cast<hipStream_t>(stream),
                       nx, ny, nz, yp, gy, ClampMinOp{yf, lo});
    return (int)hipGetLastError();
}

int et_launch_ratio(const float *meas, const float *proj, float *out, size_t n, float eps, et_stream_t stream)
{
    hipLaunchKernelGGL(k_ratio, dim3(nblk(n, 256)), dim3(256), 0, reinterpret_cast<hipStream_t>(stream), meas, proj,
                       out, n, eps);
    return (int)hipGetLastError();
}

int et_launch_em_update_flat(float *x, const float *b, const float *s, size_t n, et_stream_t stream)
{
    hipLaunchKernelGGL(k_em_update_flat, dim3(nblk(n, 256)), dim3(256), 0, reinterpret_cast<hipStream_t>(stream), x, b,
                       s, n);
    return (int)hipGetLastError();
}

int et_launch_gather_planes(float *dst, const float *src, const int *order, int V, size_t P, et_stream_t stream)
{
    hipLaunchKernelGGL(k_gather_planes, dim3(nblk((size_t)V * P, 256)), dim3(256), 0,
                       reinterpret_cast<hipStream_t>(stream), dst, src, order, V, P);
    return (int)hipGetLastError();
}

int et_launch_ratio_parts(const float *meas, const float *part, int D, size_t VP, float eps, float *out,
                          et_stream_t stream)
{
    hipLaunchKernelGGL(k_ratio_parts, dim3(nblk(VP, 256)), dim3(256), 0, reinterpret_cast<hipStream_t>(stream), meas,
                       part, D, VP, eps, out);
    return (int)hipGetLastError();
}
