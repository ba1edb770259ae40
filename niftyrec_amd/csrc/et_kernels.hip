// et_kernels.hip -- HIP kernels of the SPECT projector for gfx950 (CDNA4):
// the pull (backprojection) kernels, the general path, the layout and
// elementwise kernels and every launcher.  The depth-sweep kernels are in
// et_sweep.h (instantiated per PSF radius by et_sweep_r.hip).
//
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
#include "et_sweep.h"

namespace {

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

// astart[v][d] = sum_{d' < d} csum[v][d'], atot[v] = sum_d csum[v][d], with
// csum + csum_lo the compensated chunk sums; the scan is compensated as well and
// the whole-ray total kept as a (hi, lo) pair.
__global__ void __launch_bounds__(256) k_chunk_scan(const float *__restrict__ csum, const float *__restrict__ csum_lo,
                                                    float *__restrict__ astart, float *__restrict__ atot,
                                                    float *__restrict__ atot_lo, int V, int D, int P)
{
    const size_t i = (size_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= (size_t)V * P) return;
    const int v = (int)(i / P);
    const size_t r = i % P;
    float hi = 0.f, lo = 0.f;
    for (int d = 0; d < D; d++) {
        const size_t k = ((size_t)v * D + d) * P + r;
        if (astart) astart[k] = hi + lo;
        comp_add(hi, lo, csum[k]);
        lo += csum_lo[k];
    }
    if (atot) atot[i] = hi;
    if (atot_lo) atot_lo[i] = lo;
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

__global__ void __launch_bounds__(256) k_bg_only(const float *__restrict__ atot, const float *__restrict__ atot_lo,
                                                 float bg, size_t n, float *__restrict__ sino, int truncate)
{
    const size_t i = (size_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= n) return;
    float s = bg * exp_neg(atot[i] + atot_lo[i]);
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

// Content fingerprint of n floats: the sum modulo 2^64 of a 64-bit mix of
// (bits, index) per element (order-independent: any partition into blocks
// gives the same sum).  Any change of one element changes the sum.
__device__ __forceinline__ unsigned long long fp_mix(unsigned long long h)
{
    h ^= h >> 33;
    h *= 0xff51afd7ed558ccdull;
    h ^= h >> 33;
    h *= 0xc4ceb9fe1a85ec53ull;
    h ^= h >> 33;
    return h;
}

// Fingerprints of one or two arrays of n floats (p0, p1): blocks [0, nb) hash
// p0, blocks [nb, 2 nb) p1, and block b stores its partial sum in parts[b] (no
// atomics, nothing to clear first); an array's key is the sum modulo 2^64 of
// its nb partials.  16-byte loads where both arrays are 16-byte aligned (VEC4).
template <bool VEC4>
__global__ void __launch_bounds__(256) k_fingerprint(const unsigned *__restrict__ p0, const unsigned *__restrict__ p1,
                                                     size_t n, int nb, unsigned long long *__restrict__ parts)
{
    constexpr unsigned long long G = 0x9e3779b97f4a7c15ull;
    const int arr = (int)blockIdx.x / nb, lb = (int)blockIdx.x - arr * nb;
    const unsigned *__restrict__ p = arr ? p1 : p0;
    unsigned long long h = 0;
    const size_t t0 = (size_t)lb * blockDim.x + threadIdx.x, stride = (size_t)nb * blockDim.x;
    size_t done = 0;
    if constexpr (VEC4) {
        const size_t n4 = n / 4;
        const uint4 *p4 = reinterpret_cast<const uint4 *>(p);
        for (size_t j = t0; j < n4; j += stride) {
            const uint4 v = p4[j];
            const size_t i = 4 * j;
            h += fp_mix(((unsigned long long)v.x << 32) ^ (i * G)) +
                 fp_mix(((unsigned long long)v.y << 32) ^ ((i + 1) * G)) +
                 fp_mix(((unsigned long long)v.z << 32) ^ ((i + 2) * G)) +
                 fp_mix(((unsigned long long)v.w << 32) ^ ((i + 3) * G));
        }
        done = 4 * n4;
    }
    for (size_t i = done + t0; i < n; i += stride) h += fp_mix(((unsigned long long)p[i] << 32) ^ (i * G));
    for (int o = 32; o >= 1; o >>= 1) h += __shfl_xor(h, o, 64);
    __shared__ unsigned long long part[4];
    if ((threadIdx.x & 63) == 0) part[threadIdx.x >> 6] = h;
    __syncthreads();
    if (threadIdx.x == 0) parts[blockIdx.x] = part[0] + part[1] + part[2] + part[3];
}

// dst[v][x][y] = src[order[v]][y][x] for y < ny, 0 for ny <= y < rows: the
// measured planes of the row-lane kernels, gathered in subset order and
// transposed (through a 32 x 33 LDS tile); order null: view order.  grid
// (ceil(nx/32), ceil(rows/32), V).
__global__ void __launch_bounds__(256) k_gather_planes_t(float *__restrict__ dst, const float *__restrict__ src,
                                                         const int *__restrict__ order, int nx, int ny, int rows)
{
    __shared__ float tile[32][33];
    const int v = blockIdx.z;
    const int xb = blockIdx.x * 32, yb = blockIdx.y * 32;
    const int tx = threadIdx.x & 31, ty = threadIdx.x >> 5;
    const float *sp = src + (size_t)(order ? order[v] : v) * nx * ny;
    for (int k = ty; k < 32; k += 8) {
        const int x = xb + tx, y = yb + k;
        tile[k][tx] = (x < nx && y < ny) ? sp[x + (size_t)nx * y] : 0.f;
    }
    __syncthreads();
    float *dp = dst + (size_t)v * nx * rows;
    for (int k = ty; k < 32; k += 8) {
        const int y = yb + tx, x = xb + k;
        if (x < nx && y < rows) dp[y + (size_t)rows * x] = tile[tx][k];
    }
}

// sino[v][y][x] = sum over d < D of part[d][v][x][y] (row-lane partial sums,
// rows per column), added in order of d with TwoSum compensation, truncated at
// 0 if asked: the projection of the row-lane projector in the ABI layout,
// transposed through a 32 x 33 LDS tile.  grid (ceil(nx/32), ceil(ny/32), V).
__global__ void __launch_bounds__(256) k_yl_parts_to_sino(const float *__restrict__ part, int D, int V, int nx,
                                                          int ny, int rows, int truncate, float *__restrict__ sino)
{
    __shared__ float tile[32][33];
    const int v = blockIdx.z;
    const int xb = blockIdx.x * 32, yb = blockIdx.y * 32;
    const int tx = threadIdx.x & 31, ty = threadIdx.x >> 5;
    const size_t stride = (size_t)V * nx * rows;
    for (int k = ty; k < 32; k += 8) {
        const int x = xb + k, y = yb + tx;
        tile[k][tx] = (x < nx && y < ny) ? sum_parts(part, D, stride, ((size_t)v * nx + x) * rows + y) : 0.f;
    }
    __syncthreads();
    float *sp = sino + (size_t)v * nx * ny;
    for (int k = ty; k < 32; k += 8) {
        const int y = yb + k, x = xb + tx;
        if (x < nx && y < ny) {
            float val = tile[tx][k];
            if (truncate && val < 0.f) val = 0.f;
            sp[x + (size_t)nx * y] = val;
        }
    }
}

inline unsigned nblk(size_t n, unsigned b) { return (unsigned)((n + b - 1) / b); }

}  // namespace


int et_sweep_slab_rows(int mode, int R) { return slab_rows(mode, R); }

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
    case r: return et_sweep_launch_r##r(mode, has_mu, dummy, st, &g);
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
    case r: return et_sweep_launch_r##r(mode, has_mu, a, st, nullptr);
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

int et_launch_chunk_scan(const float *csum, const float *csum_lo, float *astart, float *atot, float *atot_lo,
                         int V, int D, int P, et_stream_t stream)
{
    hipLaunchKernelGGL(k_chunk_scan, dim3(nblk((size_t)V * P, 256)), dim3(256), 0,
                       reinterpret_cast<hipStream_t>(stream), csum, csum_lo, astart, atot, atot_lo, V, D, P);
    return (int)hipGetLastError();
}

int et_launch_reduce_parts(const float *part, int D, int V, int P, float *sino, int truncate, et_stream_t stream)
{
    const size_t VP = (size_t)V * P;
    hipLaunchKernelGGL(k_reduce_parts, dim3(nblk(VP, 256)), dim3(256), 0, reinterpret_cast<hipStream_t>(stream),
                       part, D, VP, sino, truncate);
    return (int)hipGetLastError();
}

int et_launch_bg_only(const float *atot, const float *atot_lo, float bg, int n, float *sino, int truncate,
                      et_stream_t stream)
{
    hipLaunchKernelGGL(k_bg_only, dim3(nblk((size_t)n, 256)), dim3(256), 0, reinterpret_cast<hipStream_t>(stream),
                       atot, atot_lo, bg, (size_t)n, sino, truncate);
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

// ------------------------------------------------------------------------------
// Fisher information of a grid of points.  Thread (i, j), j >= i, of a 16 x 16
// tile of the upper triangle (tiles below the diagonal exit at once) sums, for
// every camera, the products of the two points' PSFs over the overlap of
// their boxes on the detector, weighted by the projection:
//   F[i][j] = sum_cam sum_box PSF_j * PSF_i * sino
// in the reference's order (cameras, then box x, then box y, each product
// (PSF_j * PSF_i) * sino added to the running sum without contraction), so
// the result matches the reference CPU loop bit for bit.  Reference
// emission-lib/_et.cpp:627-698 (offsets of the two PSFs in the box are
// max(ix, jx) - ix etc., the four cases of :655-679); the prior :703-719
// (rule 1) and :1633-1650 (rule 3); symmetrisation :722-731.
// ------------------------------------------------------------------------------
__global__ void __launch_bounds__(256) k_fisher_pairs(EtFisherArgs a)
{
#pragma clang fp contract(off)
    if (blockIdx.x < blockIdx.y) return;  // tile below the diagonal
    const int i = blockIdx.y * 16 + (threadIdx.x >> 4), j = blockIdx.x * 16 + (threadIdx.x & 15);
    if (i >= a.n || j >= a.n || j < i) return;
    const int hx = (a.px - 1) / 2, hy = (a.py - 1) / 2;
    float f = 0.f;
    for (int cam = 0; cam < a.V; cam++) {
        const int *ci = a.coords + ((size_t)cam * a.n + i) * 3, *cj = a.coords + ((size_t)cam * a.n + j) * 3;
        const int ix = ci[0], iy = ci[1], iz = ci[2], jx = cj[0], jy = cj[1], jz = cj[2];
        const int bx = a.px - abs(jx - ix), by = a.py - abs(jy - iy);
        if (!(bx > 0 && by > 0 && iz >= 0 && jz >= 0 && iz < a.nz && jz < a.nz)) continue;
        const int mx = max(ix, jx), my = max(iy, jy);
        const float *Pi = a.psf + (size_t)iz * a.px * a.py + (my - iy) * a.px + (mx - ix);
        const float *Pj = a.psf + (size_t)jz * a.px * a.py + (my - jy) * a.px + (mx - jx);
        const float *s = a.sino + (size_t)cam * a.snx * a.sny;
        for (int x = 0; x < bx; x++) {
            const int qx = mx - hx + x;
            if (qx < 0 || qx >= a.snx) continue;  // every y of this x lies outside
            for (int y = 0; y < by; y++) {
                const int qy = my - hy + y;
                if (qy >= 0 && qy < a.sny) f = f + (Pj[y * a.px + x] * Pi[y * a.px + x]) * s[(size_t)qy * a.snx + qx];
            }
        }
    }
    a.fisher[(size_t)i * a.n + j] = f;
    a.fisher[(size_t)j * a.n + i] = f;
    if (a.prior) {
        const int *gi = a.coords0 + (size_t)i * 3, *gj = a.coords0 + (size_t)j * 3;
        const int lim = a.prior_rule;
        const bool near = abs(gi[0] - gj[0]) <= lim && abs(gi[1] - gj[1]) <= lim && abs(gi[2] - gj[2]) <= lim;
        const float p = (near && !(lim == 3 && i == j)) ? 1.f : 0.f;
        a.prior[(size_t)i * a.n + j] = p;
        a.prior[(size_t)j * a.n + i] = p;
    }
}

int et_launch_fisher_pairs(const EtFisherArgs &a, et_stream_t stream)
{
    const int t = (a.n + 15) / 16;
    hipLaunchKernelGGL(k_fisher_pairs, dim3(t, t), dim3(256), 0, reinterpret_cast<hipStream_t>(stream), a);
    return (int)hipGetLastError();
}

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

namespace {
struct LaunchEvents {
    et_event_t start = nullptr, stop = nullptr;
    bool taken = false;
};
thread_local LaunchEvents g_launch_events;
}  // namespace

void et_offer_launch_events(et_event_t start, et_event_t stop) { g_launch_events = {start, stop, false}; }

bool et_claim_launch_events(et_event_t *start, et_event_t *stop)
{
    LaunchEvents &e = g_launch_events;
    if (!e.start || !e.stop || e.taken) return false;
    *start = e.start;
    *stop = e.stop;
    e.taken = true;
    return true;
}

bool et_withdraw_launch_events()
{
    const bool taken = g_launch_events.taken;
    g_launch_events = LaunchEvents{};
    return taken;
}

int et_fingerprint_blocks(size_t n) { return (int)std::max<size_t>(1, std::min<size_t>(nblk(n, 1024), ET_FP_BLOCKS)); }

int et_launch_fingerprints(const float *p0, const float *p1, size_t n, unsigned long long *parts, et_stream_t stream)
{
    if (!p0) return (int)hipErrorInvalidValue;
    hipStream_t st = reinterpret_cast<hipStream_t>(stream);
    const int nb = et_fingerprint_blocks(n);
    const unsigned *q0 = reinterpret_cast<const unsigned *>(p0), *q1 = reinterpret_cast<const unsigned *>(p1);
    const bool vec4 = (reinterpret_cast<uintptr_t>(p0) & 15) == 0 && (reinterpret_cast<uintptr_t>(p1) & 15) == 0;
    const dim3 grid(p1 ? 2 * nb : nb);
    if (vec4)
        hipLaunchKernelGGL(k_fingerprint<true>, grid, dim3(256), 0, st, q0, q1, n, nb, parts);
    else
        hipLaunchKernelGGL(k_fingerprint<false>, grid, dim3(256), 0, st, q0, q1, n, nb, parts);
    return (int)hipGetLastError();
}

int et_launch_yl_sweep(int kind, int R, int RPL, int NR, int nthr, size_t lds, const EtSweepArgs *args,
                       const void *tab, int n, int blocks, et_stream_t stream)
{
    if (n <= 0 || blocks <= 0 || nthr <= 0 || nthr > 1024 || (nthr & 63)) return (int)hipErrorInvalidValue;
    hipStream_t st = reinterpret_cast<hipStream_t>(stream);
    const int4 *t = reinterpret_cast<const int4 *>(tab);
    switch (R) {
#define ET_R_CASE(r) \
    case r: return et_yl_launch_r##r(kind, RPL, NR, nthr, lds, args, t, n, blocks, st);
        ET_R_CASE(0) ET_R_CASE(1) ET_R_CASE(2) ET_R_CASE(3) ET_R_CASE(4) ET_R_CASE(5)
        ET_R_CASE(6) ET_R_CASE(7) ET_R_CASE(8) ET_R_CASE(9) ET_R_CASE(10)
#undef ET_R_CASE
    default: return (int)hipErrorInvalidValue;
    }
}

int et_launch_yl_pull(int epilogue, int RPL, const EtPullArgs *args, const void *tab, int n, int blocks,
                      et_stream_t stream)
{
    if (n <= 0 || blocks <= 0) return (int)hipErrorInvalidValue;
    hipStream_t st = reinterpret_cast<hipStream_t>(stream);
    const int4 *t = reinterpret_cast<const int4 *>(tab);
#define ET_YLP(rpl)                                                                                          \
    if (RPL == rpl) {                                                                                        \
        if (epilogue == ET_EPI_OSEM)                                                                         \
            hipLaunchKernelGGL((k_yl_pull<rpl, ET_EPI_OSEM>), dim3(blocks), dim3(256), 0, st, args, t, n);    \
        else if (epilogue == ET_EPI_YF_STORE)                                                                \
            hipLaunchKernelGGL((k_yl_pull<rpl, ET_EPI_YF_STORE>), dim3(blocks), dim3(256), 0, st, args, t, n); \
        else                                                                                                 \
            return (int)hipErrorInvalidValue;                                                                \
        return (int)hipGetLastError();                                                                       \
    }
    ET_YLP(2) ET_YLP(4) ET_YLP(6) ET_YLP(8)
#undef ET_YLP
    return (int)hipErrorInvalidValue;
}

int et_launch_att_factors(const EtSweepArgs &a, et_stream_t stream)
{
    const int nthr = ((a.nx + 63) / 64) * 64;
    if (nthr > 512 || a.V <= 0) return (int)hipErrorInvalidValue;
    hipLaunchKernelGGL(k_att_factors<16>, dim3((a.ny + 15) / 16, a.V), dim3(nthr), 0,
                       reinterpret_cast<hipStream_t>(stream), a);
    return (int)hipGetLastError();
}

int et_launch_yl_pull_adj(int RPL, int nsrc, const EtPullArgs *args, const void *tab, int n, int blocks,
                          et_stream_t stream)
{
    if (n <= 0 || blocks <= 0 || (nsrc != 1 && nsrc != 2)) return (int)hipErrorInvalidValue;
    hipStream_t st = reinterpret_cast<hipStream_t>(stream);
    const int4 *t = reinterpret_cast<const int4 *>(tab);
#define ET_YLA(rpl)                                                                                  \
    if (RPL == rpl) {                                                                                \
        if (nsrc == 2)                                                                               \
            hipLaunchKernelGGL((k_yl_pull_adj<rpl, 2>), dim3(blocks), dim3(256), 0, st, args, t, n); \
        else                                                                                         \
            hipLaunchKernelGGL((k_yl_pull_adj<rpl, 1>), dim3(blocks), dim3(256), 0, st, args, t, n); \
        return (int)hipGetLastError();                                                               \
    }
    ET_YLA(2) ET_YLA(4) ET_YLA(6) ET_YLA(8)
#undef ET_YLA
    return (int)hipErrorInvalidValue;
}

int et_launch_yl_factors(int RPL, const EtSweepArgs &a, et_stream_t stream)
{
    hipStream_t st = reinterpret_cast<hipStream_t>(stream);
    const dim3 g((a.nx + 15) / 16, a.V);
#define ET_YLF(rpl)                                                                   \
    if (RPL == rpl) {                                                                 \
        hipLaunchKernelGGL((k_yl_factors<rpl>), g, dim3(256), 0, st, a);              \
        return (int)hipGetLastError();                                                \
    }
    ET_YLF(2) ET_YLF(4) ET_YLF(6) ET_YLF(8)
#undef ET_YLF
    return (int)hipErrorInvalidValue;
}

int et_launch_yl_parts_to_sino(const float *part, int D, int V, int nx, int ny, int rows, int truncate, float *sino,
                               et_stream_t stream)
{
    if (D <= 0 || V <= 0 || nx <= 0 || ny <= 0 || rows < ny) return (int)hipErrorInvalidValue;
    dim3 g((nx + 31) / 32, (ny + 31) / 32, V);
    hipLaunchKernelGGL(k_yl_parts_to_sino, g, dim3(256), 0, reinterpret_cast<hipStream_t>(stream), part, D, V, nx, ny,
                       rows, truncate, sino);
    return (int)hipGetLastError();
}

int et_launch_gather_planes_t(float *dst, const float *src, const int *order, int V, int nx, int ny, int rows,
                              et_stream_t stream)
{
    dim3 g((nx + 31) / 32, (rows + 31) / 32, V);
    hipLaunchKernelGGL(k_gather_planes_t, g, dim3(256), 0, reinterpret_cast<hipStream_t>(stream), dst, src, order, nx,
                       ny, rows);
    return (int)hipGetLastError();
}
