// et_device.h -- device helpers shared by the HIP translation units of
// lib_et_array_interface.so: address spaces, the bilinear sampler of the fast
// path, the PSF passes, the wave-cooperative gather and the block mappings.
// Everything lives in an anonymous namespace: each translation unit that
// includes it compiles its own copy.
#pragma once
#include <hip/hip_runtime.h>

#include <algorithm>
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

// Error-free transformation of a sum (Knuth's TwoSum): s + e == a + b exactly,
// s = fl(a + b), for any a, b.  Six VALU operations; built without
// floating-point contraction (the file's contract setting is not relied on).
__device__ __forceinline__ void two_sum(float a, float b, float &s, float &e)
{
#pragma clang fp contract(off)
    s = a + b;
    const float bb = s - a;
    e = (a - (s - bb)) + (b - bb);
}

// Compensated running sum (hi, lo): hi + lo carries the sum of every term to
// about twice the working precision (Neumaier/TwoSum accumulation).
__device__ __forceinline__ void comp_add(float &hi, float &lo, float x)
{
    float s, e;
    two_sum(hi, x, s, e);
    hi = s;
    lo += e;
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
typedef ET_GLOBAL_AS const float2 g_cfloat2;
typedef ET_GLOBAL_AS float2 g_float2;
__device__ __forceinline__ float gload(const float *p) { return *(g_cfloat *)p; }
__device__ __forceinline__ float4 gload4(const float *p) { return *(g_cfloat4 *)p; }
__device__ __forceinline__ void gstore(float *p, float v) { *(g_float *)p = v; }
__device__ __forceinline__ void gstore4(float *p, float4 v) { *(g_float4 *)p = v; }
// Streaming loads (non-temporal: data read once, kept from displacing the
// caches' reused lines).
typedef float et_f4v __attribute__((ext_vector_type(4)));
typedef float et_f2v __attribute__((ext_vector_type(2)));
typedef ET_GLOBAL_AS const et_f4v g_cf4v;
typedef ET_GLOBAL_AS const et_f2v g_cf2v;
__device__ __forceinline__ float4 gload4_stream(const float *p)
{
    const et_f4v v = __builtin_nontemporal_load((g_cf4v *)p);
    return make_float4(v.x, v.y, v.z, v.w);
}
__device__ __forceinline__ float2 gload2_stream(const float *p)
{
    const et_f2v v = __builtin_nontemporal_load((g_cf2v *)p);
    return make_float2(v.x, v.y);
}
// Streaming (non-temporal) stores: data written once and not read again soon,
// kept from displacing the caches' reused lines.
typedef ET_GLOBAL_AS et_f4v g_f4v;
typedef ET_GLOBAL_AS et_f2v g_f2v;
__device__ __forceinline__ void gstore4_stream(float *p, float4 v)
{
    const et_f4v t = {v.x, v.y, v.z, v.w};
    __builtin_nontemporal_store(t, (g_f4v *)p);
}
__device__ __forceinline__ void gstore2_stream(float *p, float2 v)
{
    const et_f2v t = {v.x, v.y};
    __builtin_nontemporal_store(t, (g_f2v *)p);
}
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

// Sum over d < D of part[d * stride + i], added in order of d with TwoSum
// compensation.  The loads go out eight at a time: a loop that adds each value
// as it arrives waits out one memory latency per term.
__device__ __forceinline__ float sum_parts(const float *part, int D, size_t stride, size_t i)
{
    float p = 0.f, lo = 0.f;
    int d = 0;
    for (; d + 8 <= D; d += 8) {
        float v[8];
#pragma unroll
        for (int k = 0; k < 8; k++) v[k] = gload(part + (size_t)(d + k) * stride + i);
#pragma unroll
        for (int k = 0; k < 8; k++) comp_add(p, lo, v[k]);
    }
    for (; d < D; d++) comp_add(p, lo, gload(part + (size_t)d * stride + i));
    return p + lo;
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

}  // namespace
