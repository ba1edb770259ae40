// et_sweep.h -- the depth-sweep kernels of the fast path (projection,
// rotated-frame backprojection, attenuation gradients, mu sums) and their
// launchers.  The launchers are templates over the PSF radius R; each radius
// is instantiated in its own translation unit (et_sweep_r.hip, compiled once
// per radius with -DET_SWEEP_R=R) so that the library builds in parallel.
#pragma once
#include <hip/hip_ext.h>

#include "et_device.h"
#include "et_yl.h"

namespace {

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

    // Compensated sums along the ray: the attenuation A as a TwoSum (hi, lo)
    // pair, the projection acc with Kahan's correction accc (true sum about
    // acc - accc).
    float A[TS], Al[TS], acc[TS], accc[TS];
#pragma unroll
    for (int i = 0; i < TS; i++) {
        A[i] = 0.f;
        Al[i] = 0.f;
        acc[i] = 0.f;
        accc[i] = 0.f;
    }
    auto kahan = [&](int i, float h, float e) {
        const float y = fmaf(h, e, -accc[i]);
        const float t = acc[i] + y;
        accc[i] = (t - acc[i]) - y;
        acc[i] = t;
    };
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
                comp_add(A[i], Al[i], m);
                return 0.f;
            };
            if constexpr (COOP) {
                float dummy[TS];
                coop_rows<TS, TS, 0, OPM>(a.mu + (a.gmu + c.y0), c.dx, c.dz, s, tab, stage, cum, dummy);
            } else {
                bilinear_map<TS, TS, 0>(a.mu + s.off + (a.gmu + c.y0), c.dx, c.dz, s,
                                        [&](int i, float m) { comp_add(A[i], Al[i], m); });
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
                const float e = exp_neg(A[i] + Al[i]);
                kahan(i, h[i], e);
                gstore(ez + i * c.nthr, e);
            }
        } else if constexpr (HAS_MU && AM == 2) {
#pragma unroll
            for (int i = 0; i < TS; i++) kahan(i, h[i], gload(ez + i * c.nthr));
        } else {
#pragma unroll
            for (int i = 0; i < TS; i++) kahan(i, h[i], HAS_MU ? exp_neg(A[i] + Al[i]) : 1.f);
        }
    }
    if (c.d == a.D - 1) {
        if constexpr (HAS_MU && AM == 2) {
            // e^{-A} along the whole ray = the factor of the last plane
            if (a.bg_line != 0.f) {
                const float *el = a.att + (((size_t)c.v * a.nz + a.nz - 1) * a.yu + c.y0) * c.nthr + c.tx;
#pragma unroll
                for (int i = 0; i < TS; i++) kahan(i, a.bg_line, el[i * c.nthr]);
            }
        } else {
#pragma unroll
            for (int i = 0; i < TS; i++) kahan(i, a.bg_line, HAS_MU ? exp_neg(A[i] + Al[i]) : 1.f);
        }
    }
    if (c.xin) {
        float *out = a.part + ((size_t)c.d * a.V + c.v) * c.P;
#pragma unroll
        for (int i = 0; i < TS; i++) {
            const int y = c.y0 + i;
            if (y < a.ny) {
                float val = acc[i] - accc[i];
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

    // A(z') = A_tot - sum_{k>z'} mu'(k) with both terms as compensated (hi, lo)
    // pairs: near the detector A is small against A_tot, and a plain float
    // difference would carry an absolute error of about eps * A_tot there
    float sv[NA], At[NA], Atl[NA], suf[NA], sufl[NA], tail[NA];
#pragma unroll
    for (int j = 0; j < NA; j++) {
        const int y = ys + j;
        const bool ok = c.xin && y >= 0 && y < a.ny;
        sv[j] = ok ? load_source(a, c.v, y, c.tx) : 0.f;
        const size_t k = (size_t)c.v * c.P + (size_t)y * a.nx + c.tx;
        At[j] = ok ? a.atot[k] : 0.f;
        Atl[j] = ok ? a.atot_lo[k] : 0.f;
        suf[j] = 0.f;
        sufl[j] = 0.f;
        tail[j] = 0.f;
    }
    for (int z = a.nz - 1; z >= 0; z--) {
        const Samp s = fast_sample(c.fv, c.fxp, z, a.nx, a.nz, c.dx, c.dz);
        const float *kz = psf_plane<R>(a.psf, z);  // never read when R == 0
        // mu rows first: tail needs A(z') before this plane's mu joins the suffix
        float Az[NA];
        bilinear_map<NA, LA, SA>(a.mu + s.off + (a.gmu + c.y0 - R4), c.dx, c.dz, s, [&](int j, float m) {
            Az[j] = (At[j] - suf[j]) + (Atl[j] - sufl[j]);
            comp_add(suf[j], sufl[j], m);
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

    // A(z') = A_tot - sum_{k>z'} mu'(k) in compensated (hi, lo) pairs, as in k_grad_ref
    float sv[TS], At[TS], Atl[TS], suf[TS], sufl[TS], tail[TS];
#pragma unroll
    for (int i = 0; i < TS; i++) {
        const int y = c.y0 + i;
        const bool ok = c.xin && y < a.ny;
        sv[i] = ok ? load_source(a, c.v, y, c.tx) : 0.f;
        const size_t k = (size_t)c.v * c.P + (size_t)y * a.nx + c.tx;
        At[i] = ok ? a.atot[k] : 0.f;
        Atl[i] = ok ? a.atot_lo[k] : 0.f;
        suf[i] = 0.f;
        sufl[i] = 0.f;
        tail[i] = a.bg_line * exp_neg(At[i] + Atl[i]);
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
            const float Az = (At[i] - suf[i]) + (Atl[i] - sufl[i]);
            comp_add(suf[i], sufl[i], m);
            tail[i] = fmaf(h[i], exp_neg(Az), tail[i]);
            out[i] = -sv[i] * tail[i];
        });
        if (c.xin) store_u<TS>(a, c.v, c.tx, z, c.y0, out);
    }
}

// Attenuation factors of the sweep kernels' OSEM cache, att[v][z'][yu][nthr] =
// e^{-A(x', y, z')}, A the compensated (TwoSum) inclusive sum of the resampled
// mu along depth from the detector: what the AM = 1 projection stores, without
// its projection.  Block: one slab of TS rows of one view, thread = rotated
// column; grid (slabs, V).
template <int TS>
__global__ void __launch_bounds__(512) k_att_factors(EtSweepArgs a)
{
    const SweepCtx c = sweep_ctx<TS, 0>(a, false, blk_here());
    float A[TS], Al[TS];
#pragma unroll
    for (int i = 0; i < TS; i++) {
        A[i] = 0.f;
        Al[i] = 0.f;
    }
    for (int z = 0; z < a.nz; z++) {
        const Samp s = fast_sample(c.fv, c.fxp, z, a.nx, a.nz, c.dx, c.dz);
        bilinear_map<TS, TS, 0>(a.mu + s.off + (a.gmu + c.y0), c.dx, c.dz, s,
                                [&](int i, float m) { comp_add(A[i], Al[i], m); });
        float *ez = a.att + (((size_t)c.v * a.nz + z) * a.yu + c.y0) * c.nthr + c.tx;
#pragma unroll
        for (int i = 0; i < TS; i++) gstore(ez + i * c.nthr, exp_neg(A[i] + Al[i]));
    }
}

// MUSUM: part[v][d][y][x] = sum over chunk d of mu'(x, y, z'), accumulated as
// a compensated (hi, lo) pair; the low parts go to part_lo (same layout).
template <int TS>
__global__ void __launch_bounds__(512) k_musum(EtSweepArgs a)
{
    const SweepCtx c = sweep_ctx<TS, 0>(a, true, blk_here());
    if (!c.xin) return;
    float acc[TS], lo[TS];
#pragma unroll
    for (int i = 0; i < TS; i++) {
        acc[i] = 0.f;
        lo[i] = 0.f;
    }
    for (int z = c.zb; z < c.ze; z++) {
        const Samp s = fast_sample(c.fv, c.fxp, z, a.nx, a.nz, c.dx, c.dz);
        bilinear_map<TS, TS, 0>(a.mu + s.off + (a.gmu + c.y0), c.dx, c.dz, s,
                                [&](int i, float m) { comp_add(acc[i], lo[i], m); });
    }
    const size_t off = ((size_t)c.v * a.D + c.d) * c.P;
#pragma unroll
    for (int i = 0; i < TS; i++) {
        const int y = c.y0 + i;
        if (y < a.ny) {
            a.part[off + (size_t)y * a.nx + c.tx] = acc[i];
            a.part_lo[off + (size_t)y * a.nx + c.tx] = lo[i];
        }
    }
}

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

// Row-lane sweeps (et_yl.h) of PSF radius R: kind 0 projection, 1 rotated-frame
// backprojection, 2 tiled projection, 3 exact attenuation gradient, 4 its tiled
// variant; a grouped launch of `blocks` blocks of nthr = 16 * W threads
// with `lds` bytes of dynamic LDS (two planes of the widest problem).
// A row-lane projector launch, timed by its own dispatch (start and end
// recorded in the events a profiling scope offered) where there is an offer.
#define ET_YL_PROJECTOR_LAUNCH(kern)                                                                           \
    {                                                                                                          \
        et_event_t ea, eb;                                                                                     \
        if (et_claim_launch_events(&ea, &eb))                                                                  \
            hipExtLaunchKernelGGL(kern, dim3(blocks), dim3(nthr), (uint32_t)lds, st, ea, eb, 0u, args, tab, n); \
        else                                                                                                   \
            hipLaunchKernelGGL(kern, dim3(blocks), dim3(nthr), lds, st, args, tab, n);                         \
    }

template <int R>
static int launch_yl_r(int kind, int RPL, int NR, int nthr, size_t lds, const EtSweepArgs *args, const int4 *tab,
                       int n, int blocks, hipStream_t st)
{
    if (kind == 2 || kind == 4) {  // tiled projector, tiled attenuation gradient (R > 0)
#define ET_YLT(rpl)                                                                                               \
        if (RPL == rpl) {                                                                                         \
            if (kind == 2) {                                                                                      \
                allow_lds(k_yl_project_tile<R, rpl>, lds);                                                        \
                ET_YL_PROJECTOR_LAUNCH((k_yl_project_tile<R, rpl>))                                               \
            } else {                                                                                              \
                allow_lds(k_yl_grad_tile<R, rpl>, lds);                                                           \
                hipLaunchKernelGGL((k_yl_grad_tile<R, rpl>), dim3(blocks), dim3(nthr), lds, st, args, tab, n);    \
            }                                                                                                     \
            return (int)hipGetLastError();                                                                        \
        }
        if constexpr (R > 0) {
            ET_YLT(2) ET_YLT(4) ET_YLT(6) ET_YLT(8)
        }
#undef ET_YLT
        return (int)hipErrorInvalidValue;
    }
#define ET_YL(rpl, nr)                                                                                            \
    if (RPL == rpl && NR == nr) {                                                                                 \
        if (kind == 0) {                                                                                          \
            allow_lds(k_yl_project<R, rpl, nr>, lds);                                                             \
            ET_YL_PROJECTOR_LAUNCH((k_yl_project<R, rpl, nr>))                                                    \
        } else if (kind == 3) {                                                                                   \
            allow_lds(k_yl_grad<R, rpl, nr>, lds);                                                                \
            hipLaunchKernelGGL((k_yl_grad<R, rpl, nr>), dim3(blocks), dim3(nthr), lds, st, args, tab, n);         \
        } else {                                                                                                  \
            allow_lds(k_yl_bpu<R, rpl, nr>, lds);                                                                 \
            hipLaunchKernelGGL((k_yl_bpu<R, rpl, nr>), dim3(blocks), dim3(nthr), lds, st, args, tab, n);          \
        }                                                                                                         \
        return (int)hipGetLastError();                                                                            \
    }
    ET_YL(2, 1) ET_YL(2, 2) ET_YL(2, 3) ET_YL(4, 1) ET_YL(4, 2) ET_YL(4, 3)
    ET_YL(6, 1) ET_YL(6, 2) ET_YL(6, 3) ET_YL(8, 1) ET_YL(8, 2) ET_YL(8, 3)
#undef ET_YL
    return (int)hipErrorInvalidValue;
}
#undef ET_YL_PROJECTOR_LAUNCH

// Per-radius launchers (et_sweep_r.hip): sweep of `mode` with PSF radius R,
// one problem (grp == nullptr) or a grouped launch; row-lane sweeps.
#define ET_SWEEP_DECL(r) int et_sweep_launch_r##r(int mode, int has_mu, const EtSweepArgs &a, hipStream_t st, const GroupLaunch *grp);
ET_SWEEP_DECL(0) ET_SWEEP_DECL(1) ET_SWEEP_DECL(2) ET_SWEEP_DECL(3) ET_SWEEP_DECL(4) ET_SWEEP_DECL(5)
ET_SWEEP_DECL(6) ET_SWEEP_DECL(7) ET_SWEEP_DECL(8) ET_SWEEP_DECL(9) ET_SWEEP_DECL(10)
#undef ET_SWEEP_DECL
#define ET_YL_DECL(r) int et_yl_launch_r##r(int kind, int RPL, int NR, int nthr, size_t lds, const EtSweepArgs *args, \
                                            const int4 *tab, int n, int blocks, hipStream_t st);
ET_YL_DECL(0) ET_YL_DECL(1) ET_YL_DECL(2) ET_YL_DECL(3) ET_YL_DECL(4) ET_YL_DECL(5)
ET_YL_DECL(6) ET_YL_DECL(7) ET_YL_DECL(8) ET_YL_DECL(9) ET_YL_DECL(10)
#undef ET_YL_DECL
