// et_yl.h -- row-lane kernels of the OSEM subset loop for gfx950 (CDNA4).
//
// Lane layout: threadIdx.x = 16 * slot + lane.  A 16-lane DPP row (one slot)
// holds one rotated column (x', z') with its whole axial extent: lane l owns
// the RPL consecutive rows RPL*l .. RPL*l + RPL - 1 of the plan's
// rows = 16 * RPL >= ny (rows beyond ny are zero in every source).  An axial
// rotation leaves y alone, so the bilinear sample of (x', z') is the same for
// every lane of the row, and each of its four corner loads is one contiguous
// run of rows floats of the YF layout: coalesced by construction.
//
// PSF: the axial (column) pass takes the R rows beyond a lane's run from its
// DPP neighbours (row_shr / row_shl with bound control, which reads 0 past the
// ends of the row: the zero border of the reference convolution); the x'
// (row) pass goes through LDS.  A sweep block owns one (view, depth chunk) and
// all nx columns of each plane (NR rounds of W = blockDim / 16 columns), so the
// row pass needs no halo recomputation: the LDS plane keeps R zero columns on
// either side.  The projector alternates two LDS planes between depth planes:
// one block barrier per plane orders every write after the reads of the same
// buffer (a wave reaches plane z + 2's write only after every wave has passed
// plane z + 1's barrier, which follows its plane-z reads); with one plane
// (a.yl_nbuf == 1, planes too large for two) a second barrier follows the
// reads.  The backprojection always uses one plane and two barriers, so that
// two of its blocks fit a CU.
//
// Layouts (rows = 16 * RPL, y fastest everywhere):
//   factors   att[v][z'][x'][rows]  e^{-A(x', y, z')}, A the inclusive sum of mu'
//   partials  part[d][v][x'][rows]  projection of depth chunk d
//   measured counts and ratio       [v][x'][rows]
//   U         u + v*ustride + y + rows*((x'+1) + xp*(z'+1))  (zero border in x', z')
// Activity, attenuation and sensitivity are in the YF layout of the plan, whose
// columns hold at least `rows` rows.
#pragma once
#include "et_device.h"

namespace {

// Value held by lane (l - K) (dpp_shr) or (l + K) (dpp_shl) of the same 16-lane
// row; 0 where that lane lies outside the row.
template <int K>
__device__ __forceinline__ float dpp_shr(float v)
{
    static_assert(K >= 1 && K <= 15, "row shift");
    return __int_as_float(__builtin_amdgcn_update_dpp(0, __float_as_int(v), 0x110 + K, 0xf, 0xf, true));
}
template <int K>
__device__ __forceinline__ float dpp_shl(float v)
{
    static_assert(K >= 1 && K <= 15, "row shift");
    return __int_as_float(__builtin_amdgcn_update_dpp(0, __float_as_int(v), 0x100 + K, 0xf, 0xf, true));
}

// RPL consecutive floats (RPL even; 16-byte vectors when RPL % 4 == 0).
template <int RPL>
__device__ __forceinline__ void yl_load(const float *p, float *v)
{
    if constexpr (RPL % 4 == 0) {
#pragma unroll
        for (int q = 0; q < RPL / 4; q++) {
            const float4 t = gload4(p + 4 * q);
            v[4 * q] = t.x;
            v[4 * q + 1] = t.y;
            v[4 * q + 2] = t.z;
            v[4 * q + 3] = t.w;
        }
    } else {
#pragma unroll
        for (int q = 0; q < RPL / 2; q++) {
            const float2 t = *(g_cfloat2 *)(p + 2 * q);
            v[2 * q] = t.x;
            v[2 * q + 1] = t.y;
        }
    }
}

// yl_load through streaming loads (the attenuation factors: read once per
// subset, and larger than the caches).
template <int RPL>
__device__ __forceinline__ void yl_load_stream(const float *p, float *v)
{
    if constexpr (RPL % 4 == 0) {
#pragma unroll
        for (int q = 0; q < RPL / 4; q++) {
            const float4 t = gload4_stream(p + 4 * q);
            v[4 * q] = t.x;
            v[4 * q + 1] = t.y;
            v[4 * q + 2] = t.z;
            v[4 * q + 3] = t.w;
        }
    } else {
#pragma unroll
        for (int q = 0; q < RPL / 2; q++) {
            const float2 t = gload2_stream(p + 2 * q);
            v[2 * q] = t.x;
            v[2 * q + 1] = t.y;
        }
    }
}

// yl_store through streaming stores (the gradient's rotated-frame volumes:
// larger than the caches, read back by the pull only after the whole sweep).
template <int RPL>
__device__ __forceinline__ void yl_store_stream(float *p, const float *v)
{
    if constexpr (RPL % 4 == 0) {
#pragma unroll
        for (int q = 0; q < RPL / 4; q++)
            gstore4_stream(p + 4 * q, make_float4(v[4 * q], v[4 * q + 1], v[4 * q + 2], v[4 * q + 3]));
    } else {
#pragma unroll
        for (int q = 0; q < RPL / 2; q++) gstore2_stream(p + 2 * q, make_float2(v[2 * q], v[2 * q + 1]));
    }
}

template <int RPL>
__device__ __forceinline__ void yl_store(float *p, const float *v)
{
    if constexpr (RPL % 4 == 0) {
#pragma unroll
        for (int q = 0; q < RPL / 4; q++)
            gstore4(p + 4 * q, make_float4(v[4 * q], v[4 * q + 1], v[4 * q + 2], v[4 * q + 3]));
    } else {
#pragma unroll
        for (int q = 0; q < RPL / 2; q++) *(g_float2 *)(p + 2 * q) = make_float2(v[2 * q], v[2 * q + 1]);
    }
}

template <int RPL>
__device__ __forceinline__ void lds_store(float *p, const float *v)
{
    if constexpr (RPL % 4 == 0) {
#pragma unroll
        for (int q = 0; q < RPL / 4; q++)
            *reinterpret_cast<float4 *>(p + 4 * q) = make_float4(v[4 * q], v[4 * q + 1], v[4 * q + 2], v[4 * q + 3]);
    } else {
#pragma unroll
        for (int q = 0; q < RPL / 2; q++) *reinterpret_cast<float2 *>(p + 2 * q) = make_float2(v[2 * q], v[2 * q + 1]);
    }
}

template <int RPL>
__device__ __forceinline__ void lds_load(const float *p, float *v)
{
    if constexpr (RPL % 4 == 0) {
#pragma unroll
        for (int q = 0; q < RPL / 4; q++) {
            const float4 t = *reinterpret_cast<const float4 *>(p + 4 * q);
            v[4 * q] = t.x;
            v[4 * q + 1] = t.y;
            v[4 * q + 2] = t.z;
            v[4 * q + 3] = t.w;
        }
    } else {
#pragma unroll
        for (int q = 0; q < RPL / 2; q++) {
            const float2 t = *reinterpret_cast<const float2 *>(p + 2 * q);
            v[2 * q] = t.x;
            v[2 * q + 1] = t.y;
        }
    }
}

// Bilinear sample of the lane's RPL rows; p00 points at the lane's first row of
// corner (x0, z0), dx and dz are the column strides in x and z.  Same
// arithmetic order as the sweep kernels' gathers: c00 w00 + c01 w01 + c10 w10
// + c11 w11.
template <int RPL>
__device__ __forceinline__ void yl_bilinear(const float *p00, int dx, int dz, const Samp &s, float *out)
{
    float c[4][RPL];
    yl_load<RPL>(p00, c[0]);
    yl_load<RPL>(p00 + dx, c[1]);
    yl_load<RPL>(p00 + dz, c[2]);
    yl_load<RPL>(p00 + dz + dx, c[3]);
#pragma unroll
    for (int i = 0; i < RPL; i++) out[i] = c[0][i] * s.w00 + c[1][i] * s.w01 + c[2][i] * s.w10 + c[3][i] * s.w11;
}

// The R rows on either side of a lane's run, from its row neighbours:
// ext[R - t] = row -t (lane l - ceil(t / RPL)), ext[R + RPL + t] = row RPL + t
// (lane l + 1 + t / RPL).  The DPP controls are immediates, hence the recursion.
template <int R, int RPL, int T>
struct YlHalo {
    __device__ static __forceinline__ void fill(const float *in, float *ext)
    {
        if constexpr (T < R) {
            constexpr int tb = T + 1;
            constexpr int kb = (tb + RPL - 1) / RPL;
            ext[R - tb] = dpp_shr<kb>(in[kb * RPL - tb]);
            ext[R + RPL + T] = dpp_shl<1 + T / RPL>(in[T % RPL]);
            YlHalo<R, RPL, T + 1>::fill(in, ext);
        }
    }
};

// Axial PSF pass of the lane's rows:
//   out[i] = sum_{j=0}^{2R} in[i + j - R] * kc[2R - j]
// (the convolution of _et_convolveSeparable2D.cpp:28-39, in the order of the
// sweep kernels' column_pass).  Every lane of the 16-lane row must be active.
template <int R, int RPL>
__device__ __forceinline__ void yl_column_pass(const float *in, const float *__restrict__ kc, float *out)
{
    if constexpr (R == 0) {
#pragma unroll
        for (int i = 0; i < RPL; i++) out[i] = in[i];
    } else {
        float ext[RPL + 2 * R];
#pragma unroll
        for (int t = 0; t < RPL; t++) ext[R + t] = in[t];
        YlHalo<R, RPL, 0>::fill(in, ext);
#pragma unroll
        for (int i = 0; i < RPL; i++) {
            float t = 0.f;
#pragma unroll
            for (int j = 0; j <= 2 * R; j++) t = fmaf(ext[i + j], cload(kc + 2 * R - j), t);
            out[i] = t;
        }
    }
}

// x' (row) pass of the lane's rows of column xc from an LDS plane whose column
// c + R holds column c (R zero columns on either side), column stride CS:
//   out[i] = sum_{j=0}^{2R} plane[xc + j - R][row i] * kr[2R - j]
template <int R, int RPL>
__device__ __forceinline__ void yl_row_pass(const float *plane, int CS, int xc, int row0,
                                            const float *__restrict__ kr, float *out)
{
#pragma unroll
    for (int i = 0; i < RPL; i++) out[i] = 0.f;
#pragma unroll
    for (int j = 0; j <= 2 * R; j++) {
        const float kj = cload(kr + 2 * R - j);
        float t[RPL];
        lds_load<RPL>(plane + (xc + j) * CS + row0, t);
#pragma unroll
        for (int i = 0; i < RPL; i++) out[i] = fmaf(t[i], kj, out[i]);
    }
}

// Floats of one LDS plane of a row-lane sweep block (column stride rows + 4;
// rows = 16 RPL is a compile-time constant of each kernel, so the LDS column
// offsets are instruction immediates).
__host__ __device__ constexpr int yl_plane_floats(int nx, int R, int rows) { return (nx + 2 * R) * (rows + 4); }

// Zeroes both LDS planes of the block.
__device__ __forceinline__ void yl_zero_lds(float *lds, int n)
{
    for (int i = 4 * (int)threadIdx.x; i < n; i += 4 * (int)blockDim.x)
        *reinterpret_cast<float4 *>(lds + i) = make_float4(0.f, 0.f, 0.f, 0.f);
    __syncthreads();
}

// Samples of a slot's columns computed once per 16-lane row: lane k of the row
// computes the sample of plane z0 + k / NR, round k % NR (16 / NR planes per
// batch), and each round takes its offset and weights from that lane with
// cross-lane permutes (five LDS-crossbar moves instead of the ~25 VALU
// operations of a sample, which every lane of the row would otherwise repeat).
struct YlSamp {
    int off;
    float w00, w01, w10, w11;
};

__device__ __forceinline__ YlSamp yl_samp_of(const Samp &s) { return {s.off, s.w00, s.w01, s.w10, s.w11}; }

// The sample held by lane src of this lane's 16-lane row.
__device__ __forceinline__ Samp yl_samp_from(const YlSamp &h, int src)
{
    const int addr = (((int)threadIdx.x & ~15) + src) * 4;
    Samp s;
    s.off = __builtin_amdgcn_ds_bpermute(addr, h.off);
    s.w00 = __int_as_float(__builtin_amdgcn_ds_bpermute(addr, __float_as_int(h.w00)));
    s.w01 = __int_as_float(__builtin_amdgcn_ds_bpermute(addr, __float_as_int(h.w01)));
    s.w10 = __int_as_float(__builtin_amdgcn_ds_bpermute(addr, __float_as_int(h.w10)));
    s.w11 = __int_as_float(__builtin_amdgcn_ds_bpermute(addr, __float_as_int(h.w11)));
    return s;
}

// ------------------------------------------------------------------------------
// Projection with cached attenuation factors.  Block (d, v): depth chunk d of
// view v, all columns.  part[d][v][x'][rows] = sum over the chunk's planes of
// PSF_z'(act')(x', y, z') e^{-A(x', y, z')}, accumulated with Kahan's
// correction (+ bg e^{-A_tot} in the last chunk).  Reference
// emission-lib/_et.cpp:218-281 with _et_line_integral_attenuated.cpp:26-36.
// ------------------------------------------------------------------------------
template <int R, int RPL, int NR>
__device__ __forceinline__ void yl_project_body(const EtSweepArgs &a, const Blk &b)
{
    constexpr int K = 2 * R + 1;
    extern __shared__ float lds[];
    const int lane = threadIdx.x & 15, slot = threadIdx.x >> 4, W = blockDim.x >> 4;
    const int d = b.x, v = b.y;
    const int zb = d * a.zlen, ze = min(a.nz, zb + a.zlen);
    constexpr int rows = 16 * RPL, CS = rows + 4;  // == a.yu
    const int row0 = RPL * lane;
    const int dx = a.yp, dz = a.yp * a.xp;
    const int PF = yl_plane_floats(a.nx, R, rows);
    const FastView fv = load_fast_view(a.views + (size_t)v * ET_VIEW_REC);
    if constexpr (R > 0) yl_zero_lds(lds, a.yl_nbuf * PF);
    const float *act = a.act + a.gy + row0;

    float acc[NR][RPL], cmp[NR][RPL];
#pragma unroll
    for (int r = 0; r < NR; r++)
#pragma unroll
        for (int i = 0; i < RPL; i++) {
            acc[r][i] = 0.f;
            cmp[r][i] = 0.f;
        }
    auto kahan = [&](int r, int i, float h, float e) {
        const float y = fmaf(h, e, -cmp[r][i]);
        const float t = acc[r][i] + y;
        cmp[r][i] = (t - acc[r][i]) - y;
        acc[r][i] = t;
    };

    constexpr int PB = 16 / NR;  // planes per sample batch
    YlSamp sb;
    int kb = 0;                  // plane within the batch
    for (int z = zb; z < ze; z++, kb = kb + 1 == PB ? 0 : kb + 1) {
        const float *kz = psf_plane<R>(a.psf, z);  // never read when R == 0
        if (kb == 0)  // (planes and columns past the ends give samples no round takes)
            sb = yl_samp_of(fast_sample(fv, (float)((lane % NR) * W + slot), z + lane / NR, a.nx, a.nz, dx, dz));
        // the plane's attenuation factors first: their latency passes under the
        // gather and the column pass (every wave of the block reaches the
        // accumulation at the same time, after the barrier)
        float e[NR][RPL];
#pragma unroll
        for (int r = 0; r < NR; r++) {
            const int xc = r * W + slot;
            if (xc < a.nx) yl_load_stream<RPL>(a.att + (((size_t)v * a.nz + z) * a.nx + xc) * rows + row0, e[r]);
        }
        if constexpr (R > 0) {
            float *plane = lds + (a.yl_nbuf == 2 ? (z & 1) * PF : 0);
#pragma unroll
            for (int r = 0; r < NR; r++) {
                const int xc = r * W + slot;
                if (xc < a.nx) {
                    const Samp s = yl_samp_from(sb, kb * NR + r);
                    float bl[RPL], cp[RPL];
                    yl_bilinear<RPL>(act + s.off, dx, dz, s, bl);
                    yl_column_pass<R, RPL>(bl, kz + K, cp);
                    lds_store<RPL>(plane + (xc + R) * CS + row0, cp);
                }
            }
            __syncthreads();
#pragma unroll
            for (int r = 0; r < NR; r++) {
                const int xc = r * W + slot;
                if (xc < a.nx) {
                    float h[RPL];
                    yl_row_pass<R, RPL>(plane, CS, xc, row0, kz, h);
#pragma unroll
                    for (int i = 0; i < RPL; i++) kahan(r, i, h[i], e[r][i]);
                }
            }
            if (a.yl_nbuf == 1) __syncthreads();  // one plane: the next writes wait for these reads
        } else {
#pragma unroll
            for (int r = 0; r < NR; r++) {
                const int xc = r * W + slot;
                if (xc < a.nx) {
                    const Samp s = yl_samp_from(sb, kb * NR + r);
                    float h[RPL];
                    yl_bilinear<RPL>(act + s.off, dx, dz, s, h);
#pragma unroll
                    for (int i = 0; i < RPL; i++) kahan(r, i, h[i], e[r][i]);
                }
            }
        }
    }
#pragma unroll
    for (int r = 0; r < NR; r++) {
        const int xc = r * W + slot;
        if (xc >= a.nx) continue;
        if (d == a.D - 1 && a.bg_line != 0.f) {
            // e^{-A} along the whole ray = the factor of the last plane
            float e[RPL];
            yl_load_stream<RPL>(a.att + (((size_t)v * a.nz + a.nz - 1) * a.nx + xc) * rows + row0, e);
#pragma unroll
            for (int i = 0; i < RPL; i++) kahan(r, i, a.bg_line, e[i]);
        }
        float out[RPL];
#pragma unroll
        for (int i = 0; i < RPL; i++) out[i] = acc[r][i] - cmp[r][i];
        yl_store<RPL>(a.part + (((size_t)d * a.V + v) * a.nx + xc) * rows + row0, out);
    }
}

// ------------------------------------------------------------------------------
// Exact attenuation gradient in the rotated frame, with cached factors:
//   u(x', y, z') = -s(x', y) ( sum_{k >= z'} PSF_k(act')(x', y, k) e^{-A(k)}
//                              + bg e^{-A_tot} )
// (the derivative of the projection with respect to mu'(z'); the sweep kernels'
// k_grad_mu forms the same sum from the attenuation suffix).  Block (d, v)
// sweeps its chunk's planes from the far end toward the detector; the sum of
// the later chunks comes from the projector's partial sums of the same
// chunking, part[d'][v][x'][rows] (the last chunk's holds the bg term), added
// with TwoSum in order of d', and the running sum is Kahan-compensated.  s is
// in the row-lane layout [v][x'][rows]; u as the backprojection's.
// ------------------------------------------------------------------------------
template <int R, int RPL, int NR>
__device__ __forceinline__ void yl_grad_body(const EtSweepArgs &a, const Blk &b)
{
    constexpr int K = 2 * R + 1;
    extern __shared__ float lds[];
    const int lane = threadIdx.x & 15, slot = threadIdx.x >> 4, W = blockDim.x >> 4;
    const int d = b.x, v = b.y;
    const int zb = d * a.zlen, ze = min(a.nz, zb + a.zlen);
    constexpr int rows = 16 * RPL, CS = rows + 4;  // == a.yu
    const int row0 = RPL * lane;
    const int dx = a.yp, dz = a.yp * a.xp;
    const int PF = yl_plane_floats(a.nx, R, rows);
    const FastView fv = load_fast_view(a.views + (size_t)v * ET_VIEW_REC);
    if constexpr (R > 0) yl_zero_lds(lds, a.yl_nbuf * PF);
    const float *act = a.act + a.gy + row0;
    float *uv = a.u + (size_t)v * a.ustride + row0;

    float acc[NR][RPL], cmp[NR][RPL];
#pragma unroll
    for (int r = 0; r < NR; r++) {
        const int xc = r * W + slot;
#pragma unroll
        for (int i = 0; i < RPL; i++) {
            acc[r][i] = 0.f;
            cmp[r][i] = 0.f;
        }
        if (xc >= a.nx) continue;
        float hi[RPL], lo[RPL];
#pragma unroll
        for (int i = 0; i < RPL; i++) {
            hi[i] = 0.f;
            lo[i] = 0.f;
        }
        for (int d2 = d + 1; d2 < a.D; d2++) {
            float t[RPL];
            yl_load<RPL>(a.part + (((size_t)d2 * a.V + v) * a.nx + xc) * rows + row0, t);
#pragma unroll
            for (int i = 0; i < RPL; i++) comp_add(hi[i], lo[i], t[i]);
        }
        if (d == a.D - 1 && a.bg_line != 0.f) {
            float e[RPL];
            yl_load_stream<RPL>(a.att + (((size_t)v * a.nz + a.nz - 1) * a.nx + xc) * rows + row0, e);
#pragma unroll
            for (int i = 0; i < RPL; i++) comp_add(hi[i], lo[i], a.bg_line * e[i]);
        }
#pragma unroll
        for (int i = 0; i < RPL; i++) {
            acc[r][i] = hi[i];
            cmp[r][i] = -lo[i];  // the running sum is acc - cmp
        }
    }
    auto kahan = [&](int r, int i, float h, float e) {
        const float y = fmaf(h, e, -cmp[r][i]);
        const float t = acc[r][i] + y;
        cmp[r][i] = (t - acc[r][i]) - y;
        acc[r][i] = t;
    };
    // s is read again at every plane (a cache hit after the first) rather than
    // held in registers across the sweep
    auto store = [&](int r, int xc, int z) {
        float sv[RPL], out[RPL];
        yl_load<RPL>(a.sino + ((size_t)v * a.nx + xc) * rows + row0, sv);
#pragma unroll
        for (int i = 0; i < RPL; i++) out[i] = -sv[i] * (acc[r][i] - cmp[r][i]);
        yl_store_stream<RPL>(uv + (size_t)rows * ((xc + 1) + (size_t)a.xp * (z + 1)), out);
    };

    constexpr int PB = 16 / NR;  // planes per sample batch (descending)
    YlSamp sb;
    int kb = 0;
    for (int z = ze - 1; z >= zb; z--, kb = kb + 1 == PB ? 0 : kb + 1) {
        const float *kz = psf_plane<R>(a.psf, z);  // never read when R == 0
        if (kb == 0)  // (planes before the chunk and columns past the end give samples no round takes)
            sb = yl_samp_of(fast_sample(fv, (float)((lane % NR) * W + slot), z - lane / NR, a.nx, a.nz, dx, dz));
        float e[NR][RPL];
#pragma unroll
        for (int r = 0; r < NR; r++) {
            const int xc = r * W + slot;
            if (xc < a.nx) yl_load_stream<RPL>(a.att + (((size_t)v * a.nz + z) * a.nx + xc) * rows + row0, e[r]);
        }
        if constexpr (R > 0) {
            float *plane = lds + (a.yl_nbuf == 2 ? (z & 1) * PF : 0);
#pragma unroll
            for (int r = 0; r < NR; r++) {
                const int xc = r * W + slot;
                if (xc < a.nx) {
                    const Samp s = yl_samp_from(sb, kb * NR + r);
                    float bl[RPL], cp[RPL];
                    yl_bilinear<RPL>(act + s.off, dx, dz, s, bl);
                    yl_column_pass<R, RPL>(bl, kz + K, cp);
                    lds_store<RPL>(plane + (xc + R) * CS + row0, cp);
                }
            }
            __syncthreads();
#pragma unroll
            for (int r = 0; r < NR; r++) {
                const int xc = r * W + slot;
                if (xc < a.nx) {
                    float h[RPL];
                    yl_row_pass<R, RPL>(plane, CS, xc, row0, kz, h);
#pragma unroll
                    for (int i = 0; i < RPL; i++) kahan(r, i, h[i], e[r][i]);
                    store(r, xc, z);
                }
            }
            if (a.yl_nbuf == 1) __syncthreads();  // one plane: the next writes wait for these reads
        } else {
#pragma unroll
            for (int r = 0; r < NR; r++) {
                const int xc = r * W + slot;
                if (xc < a.nx) {
                    const Samp s = yl_samp_from(sb, kb * NR + r);
                    float h[RPL];
                    yl_bilinear<RPL>(act + s.off, dx, dz, s, h);
#pragma unroll
                    for (int i = 0; i < RPL; i++) kahan(r, i, h[i], e[r][i]);
                    store(r, xc, z);
                }
            }
        }
    }
}

// ------------------------------------------------------------------------------
// Tiled projection: as yl_project_body, but block (tile, d, v) of 512 threads
// (32 lane rows) owns YL_TILE = 64 columns of each plane, two per row, and
// also computes the R columns on either side of its tile (a third round of
// 2R rows) so that its x' pass needs no other block's columns.  Two LDS planes
// of (64 + 2R) columns alternate as in the full-width projector; a block then
// needs under half the LDS and two blocks share a CU, so one block's barriers
// and gathers overlap the other's.  The x' pass runs first, through LDS, and
// the axial pass follows in registers on the owned columns only, so the halo
// columns are blended but not convolved; the separable passes commute, and
// the results differ from the full-width kernel's only by rounding.
// ------------------------------------------------------------------------------
#define YL_TILE ET_YL_TILE

template <int R, int RPL>
__device__ __forceinline__ void yl_project_tile_body(const EtSweepArgs &a, const Blk &b)
{
    constexpr int K = 2 * R + 1, W = 32, NR = 2, NRR = NR + 1;
    static_assert(W * NR == YL_TILE && 2 * R <= W, "tile shape");
    extern __shared__ float lds[];
    const int lane = threadIdx.x & 15, slot = threadIdx.x >> 4;
    // the T tiles of one (view, depth chunk) on one XCD, dispatched together:
    // each tile's halo columns are the neighbour's owned columns, in the same L2
    const int T = a.yl_tiles;
    const int2 o = xcd_order(b, T, a.D * a.V, b.x % T, b.x / T + a.D * b.y);
    const int tile = o.x, d = o.y % a.D, v = o.y / a.D;
    const int x_lo = tile * YL_TILE;
    const int zb = d * a.zlen, ze = min(a.nz, zb + a.zlen);
    constexpr int rows = 16 * RPL, CS = rows + 4;  // == a.yu
    constexpr int PF = (YL_TILE + 2 * R) * CS;     // one LDS plane: LDS column c holds column x_lo - R + c
    const int row0 = RPL * lane;
    const int dx = a.yp, dz = a.yp * a.xp;
    const FastView fv = load_fast_view(a.views + (size_t)v * ET_VIEW_REC);
    yl_zero_lds(lds, 2 * PF);
    const float *act = a.act + a.gy + row0;
    // column of this slot in column pass r: owned columns in passes r = 0 and
    // r = 1, the tile's halo columns in pass r = 2 (slots 0 .. 2R - 1); -1 where none
    auto col = [&](int r) -> int {
        if (r < NR) return x_lo + r * W + slot;
        if (slot < R) return x_lo - R + slot;
        return slot < 2 * R ? x_lo + YL_TILE + slot - R : -1;
    };
    auto valid = [&](int xc) { return xc >= 0 && xc < a.nx; };

    float acc[NR][RPL], cmp[NR][RPL];
#pragma unroll
    for (int r = 0; r < NR; r++)
#pragma unroll
        for (int i = 0; i < RPL; i++) {
            acc[r][i] = 0.f;
            cmp[r][i] = 0.f;
        }
    auto kahan = [&](int r, int i, float h, float e) {
        const float y = fmaf(h, e, -cmp[r][i]);
        const float t = acc[r][i] + y;
        cmp[r][i] = (t - acc[r][i]) - y;
        acc[r][i] = t;
    };

    constexpr int PB = 16 / NRR;  // planes per sample batch
    YlSamp sb;
    int kb = 0;
    for (int z = zb; z < ze; z++, kb = kb + 1 == PB ? 0 : kb + 1) {
        const float *kz = psf_plane<R>(a.psf, z);
        if (kb == 0)  // (planes past the chunk and absent columns give samples no round takes)
            sb = yl_samp_of(fast_sample(fv, (float)col(lane % NRR), z + lane / NRR, a.nx, a.nz, dx, dz));
        // the owned columns' factors first (their latency passes under the gathers)
        float e[NR][RPL];
#pragma unroll
        for (int r = 0; r < NR; r++) {
            const int xc = col(r);
            if (valid(xc)) yl_load_stream<RPL>(a.att + (((size_t)v * a.nz + z) * a.nx + xc) * rows + row0, e[r]);
        }
        float *plane = lds + (z & 1) * PF;
#pragma unroll
        for (int r = 0; r < NRR; r++) {
            const int xc = col(r);
            if (valid(xc)) {
                const Samp s = yl_samp_from(sb, kb * NRR + r);
                float bl[RPL];
                yl_bilinear<RPL>(act + s.off, dx, dz, s, bl);
                lds_store<RPL>(plane + (xc - x_lo + R) * CS + row0, bl);
            }
        }
        __syncthreads();
#pragma unroll
        for (int r = 0; r < NR; r++) {
            const int xc = col(r);
            if (valid(xc)) {
                float t[RPL], h[RPL];
                yl_row_pass<R, RPL>(plane, CS, xc - x_lo, row0, kz, t);
                yl_column_pass<R, RPL>(t, kz + K, h);
#pragma unroll
                for (int i = 0; i < RPL; i++) kahan(r, i, h[i], e[r][i]);
            }
        }
    }
#pragma unroll
    for (int r = 0; r < NR; r++) {
        const int xc = col(r);
        if (!valid(xc)) continue;
        if (d == a.D - 1 && a.bg_line != 0.f) {
            float e[RPL];
            yl_load_stream<RPL>(a.att + (((size_t)v * a.nz + a.nz - 1) * a.nx + xc) * rows + row0, e);
#pragma unroll
            for (int i = 0; i < RPL; i++) kahan(r, i, a.bg_line, e[i]);
        }
        float out[RPL];
#pragma unroll
        for (int i = 0; i < RPL; i++) out[i] = acc[r][i] - cmp[r][i];
        yl_store<RPL>(a.part + (((size_t)d * a.V + v) * a.nx + xc) * rows + row0, out);
    }
}

// ------------------------------------------------------------------------------
// Tiled exact attenuation gradient: yl_grad_body's reverse sweep on the tiled
// projector's blocks (tile, d, v) of 512 threads, two per CU; the later
// chunks' sums come from the tiled projector's partial sums of the same
// chunking.  Halo columns are blended only; the x' pass runs first, then the
// axial pass on the owned columns.
// ------------------------------------------------------------------------------
template <int R, int RPL>
__device__ __forceinline__ void yl_grad_tile_body(const EtSweepArgs &a, const Blk &b)
{
    constexpr int K = 2 * R + 1, W = 32, NR = 2, NRR = NR + 1;
    static_assert(W * NR == YL_TILE && 2 * R <= W, "tile shape");
    extern __shared__ float lds[];
    const int lane = threadIdx.x & 15, slot = threadIdx.x >> 4;
    const int T = a.yl_tiles;
    const int2 o = xcd_order(b, T, a.D * a.V, b.x % T, b.x / T + a.D * b.y);
    const int tile = o.x, d = o.y % a.D, v = o.y / a.D;
    const int x_lo = tile * YL_TILE;
    const int zb = d * a.zlen, ze = min(a.nz, zb + a.zlen);
    constexpr int rows = 16 * RPL, CS = rows + 4;  // == a.yu
    constexpr int PF = (YL_TILE + 2 * R) * CS;     // LDS column c holds column x_lo - R + c
    const int row0 = RPL * lane;
    const int dx = a.yp, dz = a.yp * a.xp;
    const FastView fv = load_fast_view(a.views + (size_t)v * ET_VIEW_REC);
    yl_zero_lds(lds, 2 * PF);
    const float *act = a.act + a.gy + row0;
    float *uv = a.u + (size_t)v * a.ustride + row0;
    auto col = [&](int r) -> int {
        if (r < NR) return x_lo + r * W + slot;
        if (slot < R) return x_lo - R + slot;
        return slot < 2 * R ? x_lo + YL_TILE + slot - R : -1;
    };
    auto valid = [&](int xc) { return xc >= 0 && xc < a.nx; };

    float acc[NR][RPL], cmp[NR][RPL];
#pragma unroll
    for (int r = 0; r < NR; r++) {
        const int xc = col(r);
#pragma unroll
        for (int i = 0; i < RPL; i++) {
            acc[r][i] = 0.f;
            cmp[r][i] = 0.f;
        }
        if (!valid(xc)) continue;
        float hi[RPL], lo[RPL];
#pragma unroll
        for (int i = 0; i < RPL; i++) {
            hi[i] = 0.f;
            lo[i] = 0.f;
        }
        for (int d2 = d + 1; d2 < a.D; d2++) {
            float t[RPL];
            yl_load<RPL>(a.part + (((size_t)d2 * a.V + v) * a.nx + xc) * rows + row0, t);
#pragma unroll
            for (int i = 0; i < RPL; i++) comp_add(hi[i], lo[i], t[i]);
        }
        if (d == a.D - 1 && a.bg_line != 0.f) {
            float e[RPL];
            yl_load_stream<RPL>(a.att + (((size_t)v * a.nz + a.nz - 1) * a.nx + xc) * rows + row0, e);
#pragma unroll
            for (int i = 0; i < RPL; i++) comp_add(hi[i], lo[i], a.bg_line * e[i]);
        }
#pragma unroll
        for (int i = 0; i < RPL; i++) {
            acc[r][i] = hi[i];
            cmp[r][i] = -lo[i];
        }
    }
    auto kahan = [&](int r, int i, float h, float e) {
        const float y = fmaf(h, e, -cmp[r][i]);
        const float t = acc[r][i] + y;
        cmp[r][i] = (t - acc[r][i]) - y;
        acc[r][i] = t;
    };

    constexpr int PB = 16 / NRR;  // planes per sample batch (descending)
    YlSamp sb;
    int kb = 0;
    for (int z = ze - 1; z >= zb; z--, kb = kb + 1 == PB ? 0 : kb + 1) {
        const float *kz = psf_plane<R>(a.psf, z);
        if (kb == 0)  // (planes before the chunk and absent columns give samples no round takes)
            sb = yl_samp_of(fast_sample(fv, (float)col(lane % NRR), z - lane / NRR, a.nx, a.nz, dx, dz));
        float e[NR][RPL];
#pragma unroll
        for (int r = 0; r < NR; r++) {
            const int xc = col(r);
            if (valid(xc)) yl_load_stream<RPL>(a.att + (((size_t)v * a.nz + z) * a.nx + xc) * rows + row0, e[r]);
        }
        float *plane = lds + (z & 1) * PF;
#pragma unroll
        for (int r = 0; r < NRR; r++) {
            const int xc = col(r);
            if (valid(xc)) {
                const Samp s = yl_samp_from(sb, kb * NRR + r);
                float bl[RPL];
                yl_bilinear<RPL>(act + s.off, dx, dz, s, bl);
                lds_store<RPL>(plane + (xc - x_lo + R) * CS + row0, bl);
            }
        }
        __syncthreads();
#pragma unroll
        for (int r = 0; r < NR; r++) {
            const int xc = col(r);
            if (valid(xc)) {
                float t[RPL], h[RPL], sv[RPL], out[RPL];
                yl_row_pass<R, RPL>(plane, CS, xc - x_lo, row0, kz, t);
                yl_column_pass<R, RPL>(t, kz + K, h);
#pragma unroll
                for (int i = 0; i < RPL; i++) kahan(r, i, h[i], e[r][i]);
                yl_load<RPL>(a.sino + ((size_t)v * a.nx + xc) * rows + row0, sv);
#pragma unroll
                for (int i = 0; i < RPL; i++) out[i] = -sv[i] * (acc[r][i] - cmp[r][i]);
                yl_store_stream<RPL>(uv + (size_t)rows * ((xc + 1) + (size_t)a.xp * (z + 1)), out);
            }
        }
    }
}

// ------------------------------------------------------------------------------
// Rotated-frame backprojection with cached attenuation factors:
// u(x', y, z') = PSF_z'( s(x', y) e^{-A(x', y, z')} ), s the ratio [v][x'][rows]
// (ET_SRC_SINO) or 1 on the rows below ny (ET_SRC_ONES).  Reference
// emission-lib/_et.cpp:441-471 (line backprojection, then the PSF).  One LDS
// plane and two barriers per plane: the small footprint (one plane, no
// registers held across planes) lets two blocks share a CU.
// ------------------------------------------------------------------------------
template <int R, int RPL, int NR>
__device__ __forceinline__ void yl_bpu_body(const EtSweepArgs &a, const Blk &b)
{
    constexpr int K = 2 * R + 1;
    extern __shared__ float lds[];
    const int lane = threadIdx.x & 15, slot = threadIdx.x >> 4, W = blockDim.x >> 4;
    const int d = b.x, v = b.y;
    const int zb = d * a.zlen, ze = min(a.nz, zb + a.zlen);
    constexpr int rows = 16 * RPL, CS = rows + 4;  // == a.yu
    const int row0 = RPL * lane;
    const int PF = yl_plane_floats(a.nx, R, rows);
    if constexpr (R > 0) yl_zero_lds(lds, PF);

    float sv[NR][RPL];
#pragma unroll
    for (int r = 0; r < NR; r++) {
        const int xc = r * W + slot;
        if (xc >= a.nx) continue;
        if (a.src == ET_SRC_ONES) {
#pragma unroll
            for (int i = 0; i < RPL; i++) sv[r][i] = (row0 + i < a.ny) ? 1.f : 0.f;
        } else {
            yl_load<RPL>(a.sino + ((size_t)v * a.nx + xc) * rows + row0, sv[r]);
        }
    }
    float *uv = a.u + (size_t)v * a.ustride + row0;
    for (int z = zb; z < ze; z++) {
        const float *kz = psf_plane<R>(a.psf, z);  // never read when R == 0
#pragma unroll
        for (int r = 0; r < NR; r++) {
            const int xc = r * W + slot;
            if (xc < a.nx) {
                float w[RPL], e[RPL];
                yl_load_stream<RPL>(a.att + (((size_t)v * a.nz + z) * a.nx + xc) * rows + row0, e);
#pragma unroll
                for (int i = 0; i < RPL; i++) w[i] = sv[r][i] * e[i];
                if constexpr (R > 0) {
                    float cp[RPL];
                    yl_column_pass<R, RPL>(w, kz + K, cp);
                    lds_store<RPL>(lds + (xc + R) * CS + row0, cp);
                } else {
                    yl_store<RPL>(uv + (size_t)rows * ((xc + 1) + (size_t)a.xp * (z + 1)), w);
                }
            }
        }
        if constexpr (R > 0) {
            __syncthreads();
#pragma unroll
            for (int r = 0; r < NR; r++) {
                const int xc = r * W + slot;
                if (xc < a.nx) {
                    float o[RPL];
                    yl_row_pass<R, RPL>(lds, CS, xc, row0, kz, o);
                    yl_store<RPL>(uv + (size_t)rows * ((xc + 1) + (size_t)a.xp * (z + 1)), o);
                }
            }
            __syncthreads();  // the next plane's writes wait for these reads
        }
    }
}

template <int R, int RPL, int NR>
__global__ void __launch_bounds__(1024) k_yl_project(const EtSweepArgs *__restrict__ args, const int4 *__restrict__ tab,
                                                     int n)
{
    Blk b;
    int j;
    if (!group_block(tab, n, b, j)) return;
    const EtSweepArgs a = ((ET_CONST_AS const EtSweepArgs *)args)[j];
    yl_project_body<R, RPL, NR>(a, b);
}

template <int R, int RPL>
__global__ void __launch_bounds__(512, 4) k_yl_grad_tile(const EtSweepArgs *__restrict__ args,
                                                       const int4 *__restrict__ tab, int n)
{
    Blk b;
    int j;
    if (!group_block(tab, n, b, j)) return;
    const EtSweepArgs a = ((ET_CONST_AS const EtSweepArgs *)args)[j];
    yl_grad_tile_body<R, RPL>(a, b);
}

// Two blocks per CU: the second launch-bounds argument is the minimum number
// of waves per SIMD (4: at most 128 VGPRs).
template <int R, int RPL>
__global__ void __launch_bounds__(512, 4) k_yl_project_tile(const EtSweepArgs *__restrict__ args,
                                                          const int4 *__restrict__ tab, int n)
{
    Blk b;
    int j;
    if (!group_block(tab, n, b, j)) return;
    const EtSweepArgs a = ((ET_CONST_AS const EtSweepArgs *)args)[j];
    yl_project_tile_body<R, RPL>(a, b);
}

template <int R, int RPL, int NR>
__global__ void __launch_bounds__(1024) k_yl_grad(const EtSweepArgs *__restrict__ args, const int4 *__restrict__ tab,
                                                  int n)
{
    Blk b;
    int j;
    if (!group_block(tab, n, b, j)) return;
    const EtSweepArgs a = ((ET_CONST_AS const EtSweepArgs *)args)[j];
    yl_grad_body<R, RPL, NR>(a, b);
}

// One LDS plane per block; where the kernel needs at most 64 VGPRs (8 waves
// per SIMD: every instantiation with NR <= 2, such as 54 at R = 5, RPL = 8),
// two blocks share a CU and the barriers of one block's planes overlap the
// other's gathers and stores.  Not forced by the launch bounds: the widest
// instantiations (NR = 3, R >= 7) would spill.
template <int R, int RPL, int NR>
__global__ void __launch_bounds__(1024) k_yl_bpu(const EtSweepArgs *__restrict__ args, const int4 *__restrict__ tab,
                                                 int n)
{
    Blk b;
    int j;
    if (!group_block(tab, n, b, j)) return;
    const EtSweepArgs a = ((ET_CONST_AS const EtSweepArgs *)args)[j];
    yl_bpu_body<R, RPL, NR>(a, b);
}

// ------------------------------------------------------------------------------
// Attenuation factors of every view (no PSF): att[v][z'][x'][rows] =
// e^{-A(x', y, z')}, A the inclusive sum along depth of the resampled mu,
// accumulated as a compensated (hi, lo) pair.  Block: 16 columns of one view,
// all planes.  grid (ceil(nx / 16), V), 256 threads.
// ------------------------------------------------------------------------------
template <int RPL>
__device__ __forceinline__ void yl_factor_body(const EtSweepArgs &a)
{
    const int lane = threadIdx.x & 15, slot = threadIdx.x >> 4;
    const int xc = blockIdx.x * 16 + slot, v = blockIdx.y;
    if (xc >= a.nx) return;  // the whole 16-lane row
    constexpr int rows = 16 * RPL;  // == a.yu
    const int row0 = RPL * lane;
    const int dx = a.yp, dz = a.yp * a.xp;
    const FastView fv = load_fast_view(a.views + (size_t)v * ET_VIEW_REC);
    const float *mu = a.mu + a.gmu + row0;
    float A[RPL], Al[RPL];
#pragma unroll
    for (int i = 0; i < RPL; i++) {
        A[i] = 0.f;
        Al[i] = 0.f;
    }
    for (int z = 0; z < a.nz; z++) {
        const Samp s = fast_sample(fv, (float)xc, z, a.nx, a.nz, dx, dz);
        float m[RPL], e[RPL];
        yl_bilinear<RPL>(mu + s.off, dx, dz, s, m);
#pragma unroll
        for (int i = 0; i < RPL; i++) {
            comp_add(A[i], Al[i], m[i]);
            e[i] = exp_neg(A[i] + Al[i]);
        }
        yl_store<RPL>(a.att + (((size_t)v * a.nz + z) * a.nx + xc) * rows + row0, e);
    }
}

template <int RPL>
__global__ void __launch_bounds__(256) k_yl_factors(EtSweepArgs a)
{
    yl_factor_body<RPL>(a);
}

// ------------------------------------------------------------------------------
// Pull of the rotated-frame volumes of a.V views into the image frame (the
// reference inverse-rotation pull: negated angles in ZYX order, trilinear with
// zero padding, views accumulated in order; emission-lib/_et.cpp:474-498), with
// the OSEM update x *= max(b, 0) / sens (ET_EPI_OSEM) or the truncated store of
// b (ET_EPI_YF_STORE) on the rows below ny.  Block: a 4 x 4 tile of image
// columns (x, z), x fastest; tiles numbered in 8 x 8-tile super-tiles (x-major
// within and between them), contiguous per XCD, so that an XCD covers a
// compact image region (32 x 32 columns per super-tile).  256 threads; grid
// 64 * ceil(nx/32) * ceil(nz/32), blocks past the image exit at once.
// ------------------------------------------------------------------------------
template <int RPL, int EPI>
__device__ __forceinline__ void yl_pull_body(const EtPullArgs &a, const Blk &b)
{
    const int lane = threadIdx.x & 15, slot = threadIdx.x >> 4;
    const int cb = ((b.gx & 7) == 0) ? (b.lin & 7) * (b.gx >> 3) + (b.lin >> 3) : b.lin;
    const int ntx = (a.nx + 3) >> 2, nsx = (ntx + 7) >> 3;
    const int sup = cb >> 6, t = cb & 63;
    const int x = ((sup % nsx) * 8 + (t & 7)) * 4 + (slot & 3), z = ((sup / nsx) * 8 + (t >> 3)) * 4 + (slot >> 2);
    if (x >= a.nx || z >= a.nz) return;
    constexpr int rows = 16 * RPL;  // == a.yu
    const int row0 = RPL * lane;
    const int dx = rows, dz = rows * a.xp;
    float acc[RPL];
#pragma unroll
    for (int i = 0; i < RPL; i++) acc[i] = 0.f;
    YlSamp sb;
    for (int v = 0; v < a.V; v++) {
        if ((v & 15) == 0) {
            // lane k of the row: the sample of view v + k (views past the
            // subset give samples no view takes)
            const int vk = min(v + lane, a.V - 1);
            const float *rec = a.views + (size_t)vk * ET_VIEW_REC;
            const float qx = pos2(gload(rec + 12), (float)x, gload(rec + 14), (float)z, gload(rec + 15));
            const float qz = pos2(gload(rec + 20), (float)x, gload(rec + 22), (float)z, gload(rec + 23));
            sb = yl_samp_of(make_samp(qx, qz, a.nx, a.nz, dx, dz));
        }
        const Samp s = yl_samp_from(sb, v & 15);
        float t[RPL];
        yl_bilinear<RPL>(a.u + (size_t)v * a.ustride + row0 + s.off, dx, dz, s, t);
#pragma unroll
        for (int i = 0; i < RPL; i++) acc[i] += t[i];
    }
    const size_t base = (a.gy + row0) + (size_t)a.yp * ((x + 1) + (size_t)(a.nx + 2) * (z + 1));
    float o[RPL];
    if constexpr (EPI == ET_EPI_OSEM) {
        float xv[RPL], sv[RPL];
        yl_load<RPL>(a.x + base, xv);
        yl_load<RPL>(a.sens + base, sv);
#pragma unroll
        for (int i = 0; i < RPL; i++) {
            const float bb = acc[i] < 0.f ? 0.f : acc[i];
            o[i] = (row0 + i < a.ny) ? xv[i] * (bb / sv[i]) : xv[i];
        }
        yl_store<RPL>(a.x + base, o);
    } else {
        float ov[RPL];
        yl_load<RPL>(a.out + base, ov);
#pragma unroll
        for (int i = 0; i < RPL; i++) {
            float val = a.accumulate ? ov[i] + acc[i] : acc[i];
            if (a.truncate && val < 0.f) val = 0.f;
            o[i] = (row0 + i < a.ny) ? val : ov[i];
        }
        yl_store<RPL>(a.out + base, o);
    }
}

// ------------------------------------------------------------------------------
// Exact adjoint of the row-lane and sweep projectors' gather: image column
// (x, z) takes, from every rotated-frame column (x', z') whose forward sample
// has (x, z) among its corners, u(x', z') times the forward weight of that
// corner, summed over views in order and candidates x'-major (the order of the
// sweep kernels' adjoint pull, so each element sees the same FMA sequence).
// The candidates lie in a box of at most 3 x 3 columns (bounded through the
// inverse map); lane k of the row computes candidate k's weight and offset,
// the row's non-zero ones are listed in LDS in candidate order (ballot, rank by
// popcount), and the row takes them four at a time, the four loads in flight
// together.  Stores as ET_EPI_YF_STORE (accumulate, truncate; rows below ny).
// Block: a 4 x 4 tile of image columns (x fastest), as the reference pull's.
// 256 threads; grid ceil(nx/4) * ceil(nz/4).
// NS = 2: the same candidate lists pull a.u into a.out and a.u2 into a.out2,
// two hits at a time (as many loads in flight as four hits of one volume);
// each output sees the FMA sequence of a pull of its volume alone.
// ------------------------------------------------------------------------------
template <int RPL, int NS>
__device__ __forceinline__ void yl_pull_adj_body(const EtPullArgs &a, const Blk &b)
{
    static_assert(NS == 1 || NS == 2, "one or two volumes");
    constexpr int HB = 4 / NS;  // hits per batch
    const int lane = threadIdx.x & 15, slot = threadIdx.x >> 4;
    const int cb = ((b.gx & 7) == 0) ? (b.lin & 7) * (b.gx >> 3) + (b.lin >> 3) : b.lin;
    const int ntx = (a.nx + 3) >> 2, nsx = (ntx + 7) >> 3;
    const int sup = cb >> 6, t = cb & 63;
    const int x = ((sup % nsx) * 8 + (t & 7)) * 4 + (slot & 3), z = ((sup / nsx) * 8 + (t >> 3)) * 4 + (slot >> 2);
    if (x >= a.nx || z >= a.nz) return;  // (the whole 16-lane row)
    constexpr int rows = 16 * RPL;  // == a.yu
    const int row0 = RPL * lane;
    const float fxo = (float)x, fzo = (float)z;
    __shared__ float hit_w[16][16];  // per row: the weights and offsets of its non-zero candidates
    __shared__ int hit_o[16][16];
    const int rsh = (int)threadIdx.x & 48;  // this row's first lane within the wave
    float acc[NS][RPL];
#pragma unroll
    for (int s = 0; s < NS; s++)
#pragma unroll
        for (int i = 0; i < RPL; i++) acc[s][i] = 0.f;
    for (int v = 0; v < a.V; v++) {
        const float *rec = a.views + (size_t)v * ET_VIEW_REC;
        const float *uv[NS];
        uv[0] = a.u + (size_t)v * a.ustride + row0;
        if constexpr (NS == 2) uv[1] = a.u2 + (size_t)v * a.ustride + row0;
        const float f00 = cload(rec + 0), f02 = cload(rec + 2), f03 = cload(rec + 3), f20 = cload(rec + 8),
                    f22 = cload(rec + 10), f23 = cload(rec + 11);
        const float i00 = cload(rec + 24), i02 = cload(rec + 26), i20 = cload(rec + 30), i22 = cload(rec + 32);
        const float rx = fxo - f03, rz = fzo - f23;
        const float cx = i00 * rx + i02 * rz, cz = i20 * rx + i22 * rz;
        const float hx = fabsf(i00) + fabsf(i02) + 0.01f, hz = fabsf(i20) + fabsf(i22) + 0.01f;
        const int xlo = max(0, (int)ceilf(cx - hx)), xhi = min(a.nx - 1, (int)floorf(cx + hx));
        const int zlo = max(0, (int)ceilf(cz - hz)), zhi = min(a.nz - 1, (int)floorf(cz + hz));
        const int bz = zhi - zlo + 1;
        const int nc = (xhi >= xlo && bz > 0) ? (xhi - xlo + 1) * bz : 0;
        for (int k0 = 0; k0 < nc; k0 += 16) {
            float wk = 0.f;
            int ok = 0;
            const int kk = k0 + lane;
            if (kk < nc) {
                const int xc = xlo + kk / bz, zc = zlo + kk % bz;
                const float px = pos2(f00, (float)xc, f02, (float)zc, f03);
                const float pz = pos2(f20, (float)xc, f22, (float)zc, f23);
                const float fx = floorf(px), fz = floorf(pz);
                const float rxx = px - fx, rzz = pz - fz;
                const float wx = (fx == fxo) ? 1.f - rxx : ((fx == fxo - 1.f) ? rxx : 0.f);
                const float wz = (fz == fzo) ? 1.f - rzz : ((fz == fzo - 1.f) ? rzz : 0.f);
                wk = wx * wz;
                ok = rows * ((xc + 1) + a.xp * (zc + 1));
            }
            const unsigned mask =
                (unsigned)(__builtin_amdgcn_ballot_w64(wk != 0.f) >> rsh) & 0xffffu;  // the row's hits
            if (wk != 0.f) {
                const int rank = __builtin_popcount(mask & ((1u << lane) - 1u));
                hit_w[slot][rank] = wk;
                hit_o[slot][rank] = ok;
            }
            const int nh = __builtin_popcount(mask);  // the same for the whole row
            for (int j = 0; j < nh; j += HB) {
                float w[HB], t[HB][NS][RPL];
#pragma unroll
                for (int q = 0; q < HB; q++) {
                    const int h = j + q < nh ? j + q : j;  // past the list: reload hit j, weight unused
                    w[q] = hit_w[slot][h];
                    const int o = hit_o[slot][h];
#pragma unroll
                    for (int s = 0; s < NS; s++) yl_load<RPL>(uv[s] + o, t[q][s]);
                }
#pragma unroll
                for (int q = 0; q < HB; q++) {
                    if (j + q < nh) {
#pragma unroll
                        for (int s = 0; s < NS; s++)
#pragma unroll
                            for (int i = 0; i < RPL; i++) acc[s][i] = fmaf(t[q][s][i], w[q], acc[s][i]);
                    }
                }
            }
        }
    }
    const size_t base = (a.gy + row0) + (size_t)a.yp * ((x + 1) + (size_t)(a.nx + 2) * (z + 1));
#pragma unroll
    for (int s = 0; s < NS; s++) {
        float *out = s == 0 ? a.out : a.out2;
        float ov[RPL], o[RPL];
        yl_load<RPL>(out + base, ov);
#pragma unroll
        for (int i = 0; i < RPL; i++) {
            float val = a.accumulate ? ov[i] + acc[s][i] : acc[s][i];
            if (a.truncate && val < 0.f) val = 0.f;
            o[i] = (row0 + i < a.ny) ? val : ov[i];
        }
        yl_store<RPL>(out + base, o);
    }
}

template <int RPL, int NS>
__global__ void __launch_bounds__(256) k_yl_pull_adj(const EtPullArgs *__restrict__ args,
                                                     const int4 *__restrict__ tab, int n)
{
    Blk b;
    int j;
    if (!group_block(tab, n, b, j)) return;
    const EtPullArgs a = ((ET_CONST_AS const EtPullArgs *)args)[j];
    yl_pull_adj_body<RPL, NS>(a, b);
}

template <int RPL, int EPI>
__global__ void __launch_bounds__(256) k_yl_pull(const EtPullArgs *__restrict__ args, const int4 *__restrict__ tab,
                                                 int n)
{
    Blk b;
    int j;
    if (!group_block(tab, n, b, j)) return;
    const EtPullArgs a = ((ET_CONST_AS const EtPullArgs *)args)[j];
    yl_pull_body<RPL, EPI>(a, b);
}

}  // namespace
