/*
 * dn_kernels.hip -- kernels of the learned fluence denoiser (3D U-Net),
 * forward and backward, for gfx950.
 *
 * Tensors are BF16 in NDHWC order (channels innermost) with an explicit
 * channel stride, so a tensor may be a channel slice of a wider buffer (the
 * U-Net's skip concatenations are written in place).  All sums accumulate in
 * FP32; master weights, gradients of weights and the optimiser state are
 * FP32.
 *
 * 3x3x3 convolutions (stride 1, zero padding 1) are implicit GEMMs on the
 * matrix cores with v_mfma_f32_16x16x32_bf16:
 *  - forward / input gradient: M = output voxels, N = output channels,
 *    K = 27 taps x input channels (tap-major).  A wave owns 16 voxels and
 *    every output channel; a lane's A fragment is 8 consecutive channels of
 *    one tap of one voxel (one 16-byte load), its B fragments 8 consecutive
 *    K entries of one output channel's weight row.  The input gradient is
 *    the same kernel over the output gradient with the weights transposed
 *    and the taps mirrored; the ReLU derivative is applied as a mask on the
 *    A operand.
 *  - weight gradient: per tap, M = output channels, N = input channels,
 *    K = voxels.  A workgroup stages 32 voxels of the output gradient and of
 *    the tap-shifted input in LDS, transposed so that every fragment is one
 *    16-byte LDS read, and accumulates its voxel range; partial sums per
 *    range are reduced by a second kernel (no atomics).
 *
 * Operand layout of the 16x16x32 BF16 MFMA (wave64): lane l supplies
 * A[row = l%16][k = 8*(l/16) .. +7] and B[k = 8*(l/16) .. +7][col = l%16];
 * it receives D[row = 4*(l/16) + i][col = l%16], i = 0..3.
 */
#include <hip/hip_runtime.h>
#include <stdint.h>

#include <algorithm>
#include <mutex>

namespace dnk {

typedef __bf16 bf16x8 __attribute__((ext_vector_type(8)));
typedef float f32x4 __attribute__((ext_vector_type(4)));

__device__ __forceinline__ float bf2f(uint16_t h) { return __uint_as_float((uint32_t)h << 16); }

/* FP32 -> BF16, round to nearest even (gfx950's v_cvt_pk_bf16_f32; NaN
 * stays NaN) */
__device__ __forceinline__ uint16_t f2bf(float f) { return __builtin_bit_cast(uint16_t, (__bf16)f); }

union Frag {
    uint4 u;
    uint16_t h[8];
    bf16x8 v;
};

__device__ __forceinline__ Frag load8(const uint16_t* p) {
    Frag f;
    f.u = *(const uint4*)p;
    return f;
}

__device__ __forceinline__ Frag zero8() {
    Frag f;
    f.u = uint4{0u, 0u, 0u, 0u};
    return f;
}

/* the halves of a BF16 pair that are > 0 (1 .. 0x7F80: positive, not NaN)
 * as 0xFFFF, the others as 0 */
__device__ __forceinline__ uint32_t keep_mask2(uint32_t m) {
    const uint32_t lo = ((m & 0xFFFFu) - 1u) < 0x7F80u ? 0x0000FFFFu : 0u;
    const uint32_t hi = ((m >> 16) - 1u) < 0x7F80u ? 0xFFFF0000u : 0u;
    return lo | hi;
}

/* dz, dy, dx of tap t = 9*(dz+1) + 3*(dy+1) + (dx+1) */
__device__ __forceinline__ void tap_offsets(int t, int& dz, int& dy, int& dx) {
    dz = t / 9 - 1;
    dy = (t / 3) % 3 - 1;
    dx = t % 3 - 1;
}

/* Workgroups are dispatched to the 8 XCDs round-robin by linear id, and
 * each XCD has its own L2.  This bijection gives XCD x the contiguous
 * logical range [x*q + min(x, r), ...) of the n ids (q = n / 8, r = n % 8),
 * so neighbouring tiles, which read the same input rows, share one L2. */
__device__ __forceinline__ unsigned int xcd_logical(unsigned int h, unsigned int n) {
    const unsigned int x = h & 7u, j = h >> 3, q = n >> 3, r = n & 7u;
    return x * q + (x < r ? x : r) + j;
}

/* ---- 3x3x3 convolution, forward (and input gradient) --------------------
 * y[m][co] = act(bias[co] + sum_{t,ci} x[m + off_t][ci] * w[co][t*cin + ci])
 * where x is multiplied by (mask > 0) channel-wise when mask is given.
 * w rows are kpad long (27*cin rounded up to 32, zero padded).
 *
 * A wave computes MT tiles of 16 consecutive voxels against all NT tiles of
 * 16 output channels, so every weight fragment it loads feeds MT MFMAs and
 * every input fragment NT.  The fragments of the next 32-wide K chunk are
 * loaded before the MFMAs of the current one (register double buffering),
 * so their latency overlaps the matrix work.  A lane tracks the tap and
 * channel of its 8-channel slice of the chunk incrementally (cin is a
 * multiple of 8, so a slice never straddles two taps). */
template <int MT, int NT>
struct ConvFrags {
    Frag a[MT];
    Frag b[NT];
};

template <int MT, int NT>
__device__ __forceinline__ void conv_load(ConvFrags<MT, NT>& f, const uint16_t* __restrict__ x, int xcs,
                                          const uint16_t* __restrict__ mask, int mcs, const uint16_t* __restrict__ w,
                                          int kpad, int k0, int t, int ci, const long long* mv, const int* cx,
                                          const int* cy, const int* cz, int D, int H, int W, int r) {
#pragma unroll
    for (int j = 0; j < NT; j++) f.b[j] = load8(w + (long long)(j * 16 + r) * kpad + k0);
    int dz = 0, dy = 0, dx = 0;
    tap_offsets(t, dz, dy, dx);
    const long long off = ((long long)dz * H + dy) * W + dx;
#pragma unroll
    for (int i = 0; i < MT; i++) {
        f.a[i] = zero8();
        const int zz = cz[i] + dz, yy = cy[i] + dy, xx = cx[i] + dx;
        if (t < 27 && mv[i] >= 0 && (unsigned)zz < (unsigned)D && (unsigned)yy < (unsigned)H && (unsigned)xx < (unsigned)W) {
            const long long v = mv[i] + off;
            f.a[i] = load8(x + v * xcs + ci);
            if (mask) {
                const Frag mk = load8(mask + v * mcs + ci);
#pragma unroll
                for (int q = 0; q < 8; q++)
                    if (!(bf2f(mk.h[q]) > 0.f)) f.a[i].h[q] = 0;
            }
        }
    }
}

template <int MT, int NT>
__global__ void __launch_bounds__(256) conv3_fwd(const uint16_t* __restrict__ x, int xcs, const uint16_t* __restrict__ mask,
                                                 int mcs, const uint16_t* __restrict__ w, int kpad,
                                                 const float* __restrict__ bias, uint16_t* __restrict__ y, int ycs, int nb,
                                                 int D, int H, int W, int cin, int cout, int relu) {
    const int lane = threadIdx.x & 63, wave = threadIdx.x >> 6;
    const int r = lane & 15, g = lane >> 4;
    const long long mtot = (long long)nb * D * H * W;
    const unsigned int blk = xcd_logical(blockIdx.x, gridDim.x);
    const long long mbase = ((long long)blk * 4 + wave) * (16 * MT);
    long long mv[MT];
    int cx[MT], cy[MT], cz[MT];
#pragma unroll
    for (int i = 0; i < MT; i++) {
        const long long m = mbase + 16 * i + r;
        mv[i] = m < mtot ? m : -1;
        cx[i] = cy[i] = cz[i] = 0;
        if (m < mtot) {
            long long q = m;
            cx[i] = (int)(q % W); q /= W;
            cy[i] = (int)(q % H); q /= H;
            cz[i] = (int)(q % D);
        }
    }
    f32x4 acc[MT][NT];
#pragma unroll
    for (int i = 0; i < MT; i++)
#pragma unroll
        for (int j = 0; j < NT; j++) acc[i][j] = f32x4{0.f, 0.f, 0.f, 0.f};
    const int nch = kpad / 32;
    int t = (8 * g) / cin, ci = 8 * g - t * cin;
    ConvFrags<MT, NT> cur, nxt;
    conv_load<MT, NT>(cur, x, xcs, mask, mcs, w, kpad, 8 * g, t, ci, mv, cx, cy, cz, D, H, W, r);
    for (int kc = 0; kc < nch; kc++) {
        int tn = t, cn = ci + 32;
        while (cn >= cin) {
            cn -= cin;
            tn++;
        }
        if (kc + 1 < nch)
            conv_load<MT, NT>(nxt, x, xcs, mask, mcs, w, kpad, (kc + 1) * 32 + 8 * g, tn, cn, mv, cx, cy, cz, D, H, W, r);
#pragma unroll
        for (int i = 0; i < MT; i++)
#pragma unroll
            for (int j = 0; j < NT; j++)
                acc[i][j] = __builtin_amdgcn_mfma_f32_16x16x32_bf16(cur.a[i].v, cur.b[j].v, acc[i][j], 0, 0, 0);
        cur = nxt;
        t = tn;
        ci = cn;
    }
#pragma unroll
    for (int j = 0; j < NT; j++) {
        const int co = j * 16 + r;
        const float bv = bias ? bias[co] : 0.f;
#pragma unroll
        for (int i = 0; i < MT; i++) {
#pragma unroll
            for (int q = 0; q < 4; q++) {
                const long long mr = mbase + 16 * i + 4 * g + q;
                if (mr < mtot && co < cout) {
                    float v = acc[i][j][q] + bv;
                    if (relu) v = v > 0.f ? v : 0.f;
                    y[mr * ycs + co] = f2bf(v);
                }
            }
        }
    }
}

/* ---- 3x3x3 convolution with an LDS halo tile (cout <= 32) ------------------
 * Same result as conv3_fwd.  A workgroup owns 16 x 4 x 4 output voxels
 * (x, y, z); wave wv the z-slice wv, its M tile i the row y = i.  The input
 * halo of the tile (18 x 6 x 6 voxels) is staged in LDS for CG channels at
 * a time (zero outside the grid, masked when a mask is given), so each
 * input voxel is read from memory once per channel group instead of once
 * per tap, and every MFMA operand is one 16-byte LDS read.  The voxel
 * stride in LDS is an odd number of 16-byte units (CG + 8 or CG + 16), so
 * the 16 lanes of a row, which read 16 consecutive voxels, hit distinct
 * banks.  The weights of the channel group (16*NT output channels x 27
 * taps x CG channels) are staged in LDS with it, rows padded the same way,
 * so no MFMA waits on a global load.  K is ordered (channel group, tap,
 * 8-channel slice); lane group g takes slice-tap pair 4*kc + g of chunk kc,
 * padded with zeros past 27 taps. */
template <int CG, int NT>
__global__ void __launch_bounds__(256) conv3_fwd_halo(const uint16_t* __restrict__ x, int xcs,
                                                      const uint16_t* __restrict__ mask, int mcs,
                                                      const uint16_t* __restrict__ w, int kpad,
                                                      const float* __restrict__ bias, uint16_t* __restrict__ y, int ycs,
                                                      int nb, int D, int H, int W, int cin, int cout, int relu,
                                                      const uint16_t* __restrict__ xu, int xucs, int cu) {
    constexpr int TX = 16, TY = 4, TZ = 4, HX = TX + 2, HY = TY + 2, HZ = TZ + 2;
    constexpr int CS = (CG / 8) % 2 ? CG + 16 : CG + 8; /* LDS channel stride per voxel: odd 16-byte units */
    constexpr int SPT = CG / 8;          /* 8-channel slices per tap */
    constexpr int NPAIR = 27 * SPT;
    constexpr int NCH = (NPAIR + 3) / 4; /* K chunks per channel group */
    constexpr int WS = 27 * CG + ((27 * CG / 8) % 2 ? 0 : 8); /* LDS weight row: odd 16-byte units */
    __shared__ __attribute__((aligned(16))) uint16_t tile[HX * HY * HZ * CS];
    __shared__ __attribute__((aligned(16))) uint16_t wtile[16 * NT * WS];
    const int lane = threadIdx.x & 63, wv = threadIdx.x >> 6;
    const int r = lane & 15, g = lane >> 4;
    const int nbx = (W + TX - 1) / TX, nby = (H + TY - 1) / TY, nbz = (D + TZ - 1) / TZ;
    unsigned int b = xcd_logical(blockIdx.x, gridDim.x);
    const int bx = (int)(b % nbx); b /= nbx;
    const int by = (int)(b % nby); b /= nby;
    const int bz = (int)(b % nbz);
    const int bn = (int)(b / nbz);
    const int x0 = bx * TX, y0 = by * TY, z0 = bz * TZ;
    f32x4 acc[4][NT];
#pragma unroll
    for (int i = 0; i < 4; i++)
#pragma unroll
        for (int j = 0; j < NT; j++) acc[i][j] = f32x4{0.f, 0.f, 0.f, 0.f};
    /* LDS element offsets of this lane's operands, fixed for the kernel:
     * A of chunk kc (M tile 0; tile i adds i*HX*CS, an immediate) and B of
     * chunk 0 (chunk kc adds 32*kc: its slice-tap pair is 4*kc + g, and a
     * weight row holds the pairs in order, 8 channels each).  A pair past
     * the 27 taps gets A = 0 and a B offset inside the row. */
    int aoff[NCH];
    const int abase = (((wv + 1) * HY + 1) * HX + r + 1) * CS;
#pragma unroll
    for (int kc = 0; kc < NCH; kc++) {
        const int pr = 4 * kc + g;
        const int t = pr / SPT, sl = pr - t * SPT;
        int dz, dy, dx;
        tap_offsets(t < 27 ? t : 26, dz, dy, dx);
        aoff[kc] = t < 27 ? abase + ((dz * HY + dy) * HX + dx) * CS + 8 * sl : -1;
    }
    const int bbase = r * WS + 8 * g;
    for (int c0 = 0; c0 < cin; c0 += CG) {
        /* all of a lane's global loads are issued before any LDS store, so
         * their latencies overlap */
        constexpr int NH = HX * HY * HZ * SPT, NW = 16 * NT * NPAIR;
        constexpr int LH = (NH + 255) / 256, LW = (NW + 255) / 256;
        Frag hv[LH], wf[LW];
        /* halo voxel (hx, hy, hz) is grid voxel (x0 + hx - 1, ...): element
         * offsets are the halo origin's (64-bit, once) plus a 32-bit offset
         * within the slab of HZ planes (the launcher checks that it fits) */
        /* channels [0, cu) of a concatenated input come from xu, the
         * coarse tensor that a 2x nearest upsampling would produce them
         * from; the rest from x (channel c - cu) */
        const bool up = c0 < cu;
        const long long org = (((long long)bn * D + (z0 - 1)) * H + (y0 - 1)) * W + (x0 - 1);
        const long long xorg = org * xcs + (c0 - cu), morg = org * mcs + c0;
        /* coarse voxel of halo voxel h along an axis: origin/2 + floor((h-1)/2)
         * (tile origins are even) */
        const int D2 = D / 2, H2 = H / 2, W2 = W / 2;
        const long long uorg = ((((long long)bn * D2 + z0 / 2) * H2 + y0 / 2) * W2 + x0 / 2) * xucs + c0;
#pragma unroll
        for (int k = 0; k < LH; k++) {
            const int e = threadIdx.x + 256 * k;
            hv[k] = zero8();
            if (e < NH) {
                const int sl = e % SPT, vox = e / SPT;
                const int hx = vox % HX, hy = (vox / HX) % HY, hz = vox / (HX * HY);
                const int gx = x0 + hx - 1, gy = y0 + hy - 1, gz = z0 + hz - 1;
                if ((unsigned)gx < (unsigned)W && (unsigned)gy < (unsigned)H && (unsigned)gz < (unsigned)D) {
                    if (up) {
                        const int ru = (((hz - 1) >> 1) * H2 + ((hy - 1) >> 1)) * W2 + ((hx - 1) >> 1);
                        hv[k] = load8(xu + (uorg + (long long)ru * xucs + 8 * sl));
                        continue;
                    }
                    const int rel = (hz * H + hy) * W + hx;
                    hv[k] = load8(x + (xorg + rel * xcs + 8 * sl));
                    if (mask) {
                        const Frag mk = load8(mask + (morg + rel * mcs + 8 * sl));
#pragma unroll
                        for (int q = 0; q < 8; q++)
                            if (!(bf2f(mk.h[q]) > 0.f)) hv[k].h[q] = 0;
                    }
                }
            }
        }
#pragma unroll
        for (int k = 0; k < LW; k++) {
            const int e = threadIdx.x + 256 * k;
            if (e < NW) {
                const int co = e / NPAIR, pq = e - co * NPAIR;
                const int t = pq / SPT, sl = pq - t * SPT;
                wf[k] = load8(w + (long long)co * kpad + t * cin + c0 + 8 * sl);
            }
        }
        __syncthreads(); /* the previous channel group's reads of the tiles are done */
#pragma unroll
        for (int k = 0; k < LH; k++) {
            const int e = threadIdx.x + 256 * k;
            if (e < NH) *(uint4*)(tile + (e / SPT) * CS + 8 * (e % SPT)) = hv[k].u;
        }
#pragma unroll
        for (int k = 0; k < LW; k++) {
            const int e = threadIdx.x + 256 * k;
            if (e < NW) {
                const int co = e / NPAIR, pq = e - co * NPAIR;
                const int t = pq / SPT, sl = pq - t * SPT;
                *(uint4*)(wtile + co * WS + t * CG + 8 * sl) = wf[k].u;
            }
        }
        __syncthreads();
#pragma unroll
        for (int kc = 0; kc < NCH; kc++) {
            Frag bf[NT], af[4];
            if (4 * kc + 4 <= NPAIR) {
                /* every lane's pair is a real tap (known at compile time) */
#pragma unroll
                for (int j = 0; j < NT; j++) bf[j].u = *(const uint4*)(wtile + j * 16 * WS + bbase + 32 * kc);
#pragma unroll
                for (int i = 0; i < 4; i++) af[i].u = *(const uint4*)(tile + aoff[kc] + i * HX * CS);
            } else {
                const bool ok = aoff[kc] >= 0;
#pragma unroll
                for (int j = 0; j < NT; j++) bf[j].u = *(const uint4*)(wtile + j * 16 * WS + bbase + (ok ? 32 * kc : 0));
#pragma unroll
                for (int i = 0; i < 4; i++) {
                    af[i].u = *(const uint4*)(tile + (ok ? aoff[kc] : abase) + i * HX * CS);
                    if (!ok) af[i] = zero8();
                }
            }
#pragma unroll
            for (int i = 0; i < 4; i++)
#pragma unroll
                for (int j = 0; j < NT; j++)
                    acc[i][j] = __builtin_amdgcn_mfma_f32_16x16x32_bf16(af[i].v, bf[j].v, acc[i][j], 0, 0, 0);
        }
    }
    /* output: rows y0..y0+3 of z-slice z0+wv; lane (r, g) holds channel
     * 16j + r of voxels x0 + 4g + q.  Strides between the 16 outputs are
     * wave-uniform, and a tile inside the grid needs no bounds checks. */
    const int oz = z0 + wv;
    const bool inside = x0 + TX <= W && y0 + TY <= H && oz < D;
    const long long rowstride = (long long)W * ycs;
    uint16_t* yb = y + ((((long long)bn * D + oz) * H + y0) * W + x0 + 4 * g) * ycs;
#pragma unroll
    for (int j = 0; j < NT; j++) {
        const int co = j * 16 + r;
        const float bv = bias ? bias[co] : 0.f;
#pragma unroll
        for (int i = 0; i < 4; i++) {
#pragma unroll
            for (int q = 0; q < 4; q++) {
                if (inside || (x0 + 4 * g + q < W && y0 + i < H && oz < D)) {
                    float v = acc[i][j][q] + bv;
                    if (relu) v = v > 0.f ? v : 0.f;
                    yb[i * rowstride + q * ycs + co] = f2bf(v);
                }
            }
        }
    }
}

/* wt[ci][t'*cout + co] = w[co][t*cin + ci], t' = 26 - t (mirrored taps), rows
 * of wt padded with zeros to kpadt */
__global__ void weight_flip(const uint16_t* __restrict__ w, int kpad, uint16_t* __restrict__ wt, int kpadt, int cin, int cout) {
    const long long n = (long long)cin * kpadt;
    for (long long i = (long long)blockIdx.x * blockDim.x + threadIdx.x; i < n; i += (long long)gridDim.x * blockDim.x) {
        const int ci = (int)(i / kpadt), k = (int)(i % kpadt);
        uint16_t v = 0;
        if (k < 27 * cout) {
            const int tp = k / cout, co = k - tp * cout;
            v = w[(long long)co * kpad + (26 - tp) * cin + ci];
        }
        wt[i] = v;
    }
}

/* ---- weight gradient ----------------------------------------------------
 * partial[s][t][co][ci] = sum over voxels v of split s of
 *   dy[v][co] * (mask[v][co] > 0) * x[v + off_t][ci]
 * grid = (27, nsplit), 256 threads; TPW output tiles (16x16) per wave. */
template <int TPW>
__global__ void __launch_bounds__(256) conv3_wgrad(const uint16_t* __restrict__ x, int xcs, const uint16_t* __restrict__ dy,
                                                   int dycs, const uint16_t* __restrict__ mask, int mcs,
                                                   float* __restrict__ partial, int vps, int nb, int D, int H, int W, int cin,
                                                   int cout) {
    extern __shared__ uint4 smraw[];
    uint16_t* sm = (uint16_t*)smraw;
    const int cinp = (cin + 15) & ~15;
    uint16_t* sdy = sm;               /* [cout][32] */
    uint16_t* sx = sm + cout * 32;    /* [cinp][32] */
    /* the 27 taps of one voxel range read the same data: keep them on one XCD */
    const unsigned int lid = xcd_logical(blockIdx.x + 27u * blockIdx.y, 27u * gridDim.y);
    const int t = (int)(lid % 27u), s = (int)(lid / 27u);
    const int lane = threadIdx.x & 63, wave = threadIdx.x >> 6;
    const int r = lane & 15, g = lane >> 4;
    int dz, dyo, dxo;
    tap_offsets(t, dz, dyo, dxo);
    const long long mtot = (long long)nb * D * H * W;
    const long long v0 = (long long)s * vps;
    const long long v1 = (v0 + vps < mtot) ? v0 + vps : mtot;
    const int tco = cout / 16, tci = cinp / 16, ntile = tco * tci;
    f32x4 acc[TPW];
#pragma unroll
    for (int j = 0; j < TPW; j++) acc[j] = f32x4{0.f, 0.f, 0.f, 0.f};
    for (long long vb = v0; vb < v1; vb += 32) {
        /* stage dy (masked) transposed: sdy[co][vi] */
        for (int e = threadIdx.x; e < 32 * (cout / 8); e += 256) {
            const int vi = e / (cout / 8), c8 = (e % (cout / 8)) * 8;
            const long long v = vb + vi;
            Frag f = zero8();
            if (v < v1) {
                f = load8(dy + v * dycs + c8);
                if (mask) {
                    const Frag mk = load8(mask + v * mcs + c8);
#pragma unroll
                    for (int i = 0; i < 8; i++)
                        if (!(bf2f(mk.h[i]) > 0.f)) f.h[i] = 0;
                }
            }
#pragma unroll
            for (int i = 0; i < 8; i++) sdy[(c8 + i) * 32 + vi] = f.h[i];
        }
        /* stage tap-shifted x transposed: sx[ci][vi], zero outside */
        for (int e = threadIdx.x; e < 32 * (cinp / 8); e += 256) {
            const int vi = e / (cinp / 8), c8 = (e % (cinp / 8)) * 8;
            const long long v = vb + vi;
            Frag f = zero8();
            if (v < v1 && c8 < cin) {
                long long q = v;
                const int vx = (int)(q % W); q /= W;
                const int vy = (int)(q % H); q /= H;
                const int vz = (int)(q % D);
                const long long vn = q / D;
                const int zz = vz + dz, yy = vy + dyo, xx = vx + dxo;
                if ((unsigned)zz < (unsigned)D && (unsigned)yy < (unsigned)H && (unsigned)xx < (unsigned)W)
                    f = load8(x + (((vn * D + zz) * H + yy) * W + xx) * xcs + c8);
            }
#pragma unroll
            for (int i = 0; i < 8; i++) sx[(c8 + i) * 32 + vi] = f.h[i];
        }
        __syncthreads();
#pragma unroll
        for (int j = 0; j < TPW; j++) {
            const int tile = wave + 4 * j;
            if (tile < ntile) {
                const int ico = tile / tci, ici = tile % tci;
                Frag a, b;
                a.u = *(const uint4*)(sdy + (ico * 16 + r) * 32 + 8 * g);
                b.u = *(const uint4*)(sx + (ici * 16 + r) * 32 + 8 * g);
                acc[j] = __builtin_amdgcn_mfma_f32_16x16x32_bf16(a.v, b.v, acc[j], 0, 0, 0);
            }
        }
        __syncthreads();
    }
    float* out = partial + ((long long)s * 27 + t) * cout * cin;
#pragma unroll
    for (int j = 0; j < TPW; j++) {
        const int tile = wave + 4 * j;
        if (tile < ntile) {
            const int ico = tile / tci, ici = tile % tci;
            const int ci = ici * 16 + r;
#pragma unroll
            for (int i = 0; i < 4; i++) {
                const int co = ico * 16 + 4 * g + i;
                if (ci < cin) out[(long long)co * cin + ci] = acc[j][i];
            }
        }
    }
}

/* dw[co][t*cin + ci] = sum_s partial[s][t][co][ci]; padding entries of dw
 * (k >= 27*cin) are set to zero */
/* ---- weight gradient, 16 to 64 output channels, all taps per staging ---
 * Same partial sums as conv3_wgrad (partial[s][tap][co][ci]) for cout =
 * 16 * NCO; workgroup (s, y) takes input tiles y*NCI .. y*NCI + NCI - 1
 * (16 channels each) of the segments of split s.  The work unit is a segment of 32 voxels along x
 * (one x-row of one z-plane).  For a segment the workgroup stages, per
 * 16-channel input tile, the 9 neighbouring input rows (dz, dy in -1..1)
 * transposed to voxel-contiguous order, in three copies shifted by dx =
 * -1, 0, 1, so that the B operand of every tap is one aligned 16-byte LDS
 * read; the masked output gradient is staged once, transposed the same
 * way, as the A operand.  Wave w accumulates taps w, w+4, ... (7 at most)
 * over all segments of its workgroup: one staging feeds 27 MFMAs per input
 * tile instead of one.  Channel rows in LDS are SV = 40 voxels apart (5
 * 16-byte units, odd), so the 16 lanes that read 16 channels hit distinct
 * banks.  Transposition: a thread loads 8 channels of 8 consecutive voxels
 * and writes 8 rows of 8 voxels (v_perm pairs). */
template <int NCI, int NCO>
__global__ void __launch_bounds__(256) conv3_wgrad_rows(const uint16_t* __restrict__ x, int xcs,
                                                        const uint16_t* __restrict__ dy, int dycs,
                                                        const uint16_t* __restrict__ mask, int mcs,
                                                        float* __restrict__ partial, int segs, int nb, int D, int H,
                                                        int W, int cin) {
    constexpr int SV = 40;
    __shared__ __attribute__((aligned(16))) uint16_t xs[3 * 9 * 16 * SV]; /* [dx][row][ci][v] */
    __shared__ __attribute__((aligned(16))) uint16_t gs[16 * NCO * SV];   /* [co][v] */
    const int lane = threadIdx.x & 63, wv = threadIdx.x >> 6;
    const int r = lane & 15, g = lane >> 4;
    const int c0 = (int)blockIdx.y * NCI; /* first input tile of this workgroup */
    const int nxs = (W + 31) / 32;
    const long long nseg = (long long)nb * D * H * nxs;
    const long long sb = (long long)blockIdx.x * segs;
    const long long se = sb + segs < nseg ? sb + segs : nseg;
    f32x4 acc[7][NCI][NCO];
#pragma unroll
    for (int j = 0; j < 7; j++)
#pragma unroll
        for (int c = 0; c < NCI; c++)
#pragma unroll
            for (int o = 0; o < NCO; o++) acc[j][c][o] = f32x4{0.f, 0.f, 0.f, 0.f};
    /* staging items of this thread: x: (row, 8-voxel block, channel slice,
     * dx copy) = 9 * 4 * 2 * 3 = 216; dy: (block, slice) = 4 * 2 * NCO.
     * The work items are (segment, input tile) pairs; the global loads of
     * the next item are issued before the MFMAs of the current one, so
     * their latency overlaps the matrix work.  A mask is applied as the
     * output gradient arrives, which waits for its loads: callers that
     * train apply it beforehand (mcx_dn_relu_mask) and pass none. */
    const int it = threadIdx.x;
    const bool xthr = it < 216, gthr = it >= 216 && it < 216 + 8 * NCO;
    const int xrow = it / 24, xb = (it / 6) % 4, xsl = (it / 3) % 2, xdx = it % 3;
    const int gb = (it - 216) / (2 * NCO), gsl = (it - 216) % (2 * NCO);
    const long long nitem = (se > sb ? se - sb : 0) * NCI;
    Frag in[8];
    auto fetch = [&](long long item) {
        long long q = sb + item / NCI;
        const int c = (int)(item % NCI);
        const int x0 = (int)(q % nxs) * 32; q /= nxs;
        const int yy = (int)(q % H); q /= H;
        const int zz = (int)(q % D);
        const int nn = (int)(q / D);
        if (xthr) {
            const int gz = zz + xrow / 3 - 1, gy = yy + xrow % 3 - 1;
            const int ch = 16 * (c0 + c) + 8 * xsl;
#pragma unroll
            for (int k = 0; k < 8; k++) {
                const int gx = x0 + 8 * xb + k + xdx - 1;
                in[k] = zero8();
                if ((unsigned)gx < (unsigned)W && (unsigned)gy < (unsigned)H && (unsigned)gz < (unsigned)D && ch < cin)
                    in[k] = load8(x + ((((long long)nn * D + gz) * H + gy) * W + gx) * xcs + ch);
            }
        } else if (gthr && c == 0) {
#pragma unroll
            for (int k = 0; k < 8; k++) {
                const int gx = x0 + 8 * gb + k;
                in[k] = zero8();
                if (gx < W) {
                    const long long v = (((long long)nn * D + zz) * H + yy) * W + gx;
                    in[k] = load8(dy + v * dycs + 8 * gsl);
                    if (mask) {
                        const Frag m = load8(mask + v * mcs + 8 * gsl);
                        in[k].u.x &= keep_mask2(m.u.x);
                        in[k].u.y &= keep_mask2(m.u.y);
                        in[k].u.z &= keep_mask2(m.u.z);
                        in[k].u.w &= keep_mask2(m.u.w);
                    }
                }
            }
        }
    };
    if (nitem > 0) fetch(0);
    for (long long item = 0; item < nitem; item++) {
        const int c = (int)(item % NCI);
        __syncthreads(); /* the previous item's reads are done */
        if (xthr) {
#pragma unroll
            for (int ci = 0; ci < 8; ci++) {
                Frag o;
#pragma unroll
                for (int k = 0; k < 8; k++) o.h[k] = in[k].h[ci];
                *(uint4*)(xs + ((xdx * 9 + xrow) * 16 + 8 * xsl + ci) * SV + 8 * xb) = o.u;
            }
        } else if (gthr && c == 0) {
#pragma unroll
            for (int co = 0; co < 8; co++) {
                Frag o;
#pragma unroll
                for (int k = 0; k < 8; k++) o.h[k] = in[k].h[co];
                *(uint4*)(gs + (8 * gsl + co) * SV + 8 * gb) = o.u;
            }
        }
        __syncthreads();
        if (item + 1 < nitem) fetch(item + 1);
        Frag a[NCO];
#pragma unroll
        for (int o = 0; o < NCO; o++) a[o].u = *(const uint4*)(gs + (16 * o + r) * SV + 8 * g);
#pragma unroll
        for (int j = 0; j < 7; j++) {
            const int t = wv + 4 * j;
            if (t < 27) {
                const int row = (t / 9) * 3 + (t / 3) % 3, dxi = t % 3;
                Frag b;
                b.u = *(const uint4*)(xs + ((dxi * 9 + row) * 16 + r) * SV + 8 * g);
#pragma unroll
                for (int o = 0; o < NCO; o++)
#pragma unroll
                    for (int cc = 0; cc < NCI; cc++)
                        if (cc == c) acc[j][cc][o] = __builtin_amdgcn_mfma_f32_16x16x32_bf16(a[o].v, b.v, acc[j][cc][o], 0, 0, 0);
            }
        }
    }
    /* lane (r, g) holds D[co = 16o + 4g + i][ci = 16c + r] of each of its taps */
#pragma unroll
    for (int j = 0; j < 7; j++) {
        const int t = wv + 4 * j;
        if (t >= 27) continue;
        float* out = partial + ((long long)blockIdx.x * 27 + t) * 16 * NCO * cin;
#pragma unroll
        for (int c = 0; c < NCI; c++) {
            const int ci = 16 * (c0 + c) + r;
            if (ci >= cin) continue;
#pragma unroll
            for (int o = 0; o < NCO; o++)
#pragma unroll
                for (int i = 0; i < 4; i++) out[(16 * o + 4 * g + i) * cin + ci] = acc[j][c][o][i];
        }
    }
}

/* dw[co][k] = sum over the nsplit partial sums; 1024 threads per 64
 * outputs: 16 groups each add every 16th split, then one thread adds the
 * 16 group sums in order (deterministic) */
__global__ void __launch_bounds__(1024) wgrad_reduce(const float* __restrict__ partial, int nsplit, float* __restrict__ dw,
                                                     int kpad, int cin, int cout) {
    __shared__ float red[16][64];
    const long long n = (long long)cout * kpad;
    const long long per = 27LL * cout * cin;
    const int el = threadIdx.x & 63, sg = threadIdx.x >> 6;
    const long long i = (long long)blockIdx.x * 64 + el;
    float sum = 0.f;
    if (i < n) {
        const int co = (int)(i / kpad), k = (int)(i % kpad);
        if (k < 27 * cin) {
            const int t = k / cin, ci = k - t * cin;
            const long long off = ((long long)t * cout + co) * cin + ci;
            for (int s = sg; s < nsplit; s += 16) sum += partial[s * per + off];
        }
    }
    red[sg][el] = sum;
    __syncthreads();
    if (sg == 0 && i < n) {
        float tot = 0.f;
        for (int q = 0; q < 16; q++) tot += red[q][el];
        dw[i] = tot;
    }
}

/* db[co] partial sums over voxel ranges: part[s][co].  Each thread reads
 * 8 channels (one 16-byte load) of one voxel per iteration; cout/8 channel
 * groups x 256/(cout/8) voxel lanes per workgroup. */
__global__ void bias_grad(const uint16_t* __restrict__ dy, int dycs, const uint16_t* __restrict__ mask, int mcs,
                          float* __restrict__ part, long long nvox, int vps, int cout) {
    __shared__ float red[256][9];
    const int s = blockIdx.x;
    const long long v0 = (long long)s * vps;
    const long long v1 = (v0 + vps < nvox) ? v0 + vps : nvox;
    const int ng = cout / 8, lanes = 256 / ng;
    const int g = threadIdx.x % ng, vl = threadIdx.x / ng;
    float acc[8];
#pragma unroll
    for (int k = 0; k < 8; k++) acc[k] = 0.f;
    for (long long v = v0 + vl; v < v1; v += lanes) {
        Frag f = load8(dy + v * dycs + 8 * g);
        if (mask) {
            const Frag mk = load8(mask + v * mcs + 8 * g);
#pragma unroll
            for (int k = 0; k < 8; k++)
                if (!(bf2f(mk.h[k]) > 0.f)) f.h[k] = 0;
        }
#pragma unroll
        for (int k = 0; k < 8; k++) acc[k] += bf2f(f.h[k]);
    }
#pragma unroll
    for (int k = 0; k < 8; k++) red[threadIdx.x][k] = acc[k];
    __syncthreads();
    if (threadIdx.x < cout) {
        const int gg = threadIdx.x / 8, k = threadIdx.x % 8;
        float t = 0.f;
        for (int l = 0; l < lanes; l++) t += red[l * ng + gg][k];
        part[(long long)s * cout + threadIdx.x] = t;
    }
}

/* out[c] = sum over the nrow rows of part[row][c]; 1024 threads per 64
 * columns: 16 groups each add every 16th row, then one thread adds the 16
 * group sums in order (deterministic) */
__global__ void __launch_bounds__(1024) sum_rows(const float* __restrict__ part, int nrow, int ncol, float* __restrict__ out) {
    __shared__ float red[16][64];
    const int el = threadIdx.x & 63, sg = threadIdx.x >> 6;
    const int c = blockIdx.x * 64 + el;
    float t = 0.f;
    if (c < ncol)
        for (int s = sg; s < nrow; s += 16) t += part[(long long)s * ncol + c];
    red[sg][el] = t;
    __syncthreads();
    if (sg == 0 && c < ncol) {
        float tot = 0.f;
        for (int q = 0; q < 16; q++) tot += red[q][el];
        out[c] = tot;
    }
}

/* ---- 2x2x2 max pooling --------------------------------------------------- */
__global__ void maxpool_fwd(const uint16_t* __restrict__ x, int xcs, uint16_t* __restrict__ y, int ycs,
                            uint8_t* __restrict__ arg, int nb, int D, int H, int W, int c) {
    const int d2 = D / 2, h2 = H / 2, w2 = W / 2;
    const long long n = (long long)nb * d2 * h2 * w2 * c;
    for (long long i = (long long)blockIdx.x * blockDim.x + threadIdx.x; i < n; i += (long long)gridDim.x * blockDim.x) {
        const int ch = (int)(i % c);
        long long q = i / c;
        const int ox = (int)(q % w2); q /= w2;
        const int oy = (int)(q % h2); q /= h2;
        const int oz = (int)(q % d2);
        const long long on = q / d2;
        float best = -3.0e38f;
        int bi = 0;
        for (int k = 0; k < 8; k++) {
            const int zz = 2 * oz + (k >> 2), yy = 2 * oy + ((k >> 1) & 1), xx = 2 * ox + (k & 1);
            const float v = bf2f(x[(((on * D + zz) * H + yy) * W + xx) * xcs + ch]);
            if (v > best) { best = v; bi = k; }
        }
        y[(i / c) * ycs + ch] = f2bf(best);
        arg[i] = (uint8_t)bi;
    }
}

/* dx = dy routed to the window's maximum, zero elsewhere (every input
 * voxel is written) */
__global__ void maxpool_bwd(const uint16_t* __restrict__ dy, int dycs, const uint8_t* __restrict__ arg,
                            uint16_t* __restrict__ dx, int dxcs, int nb, int D, int H, int W, int c) {
    const int d2 = D / 2, h2 = H / 2, w2 = W / 2;
    const long long n = (long long)nb * D * H * W * c;
    for (long long i = (long long)blockIdx.x * blockDim.x + threadIdx.x; i < n; i += (long long)gridDim.x * blockDim.x) {
        const int ch = (int)(i % c);
        long long q = i / c;
        const int xx = (int)(q % W); q /= W;
        const int yy = (int)(q % H); q /= H;
        const int zz = (int)(q % D);
        const long long nn = q / D;
        const long long o = ((nn * d2 + zz / 2) * h2 + yy / 2) * w2 + xx / 2;
        const int k = ((zz & 1) << 2) | ((yy & 1) << 1) | (xx & 1);
        const uint16_t v = (arg[o * c + ch] == k) ? dy[o * dycs + ch] : (uint16_t)0;
        dx[(i / c) * dxcs + ch] = v;
    }
}

/* ---- nearest-neighbour x2 upsampling -------------------------------------
 * forward: y (2D x 2H x 2W) from x (D x H x W); backward: dx = sum of the
 * 8 children's gradients */
__global__ void upsample_fwd(const uint16_t* __restrict__ x, int xcs, uint16_t* __restrict__ y, int ycs, int nb, int D,
                             int H, int W, int c) {
    const int D2 = 2 * D, H2 = 2 * H, W2 = 2 * W;
    const long long n = (long long)nb * D2 * H2 * W2 * c;
    for (long long i = (long long)blockIdx.x * blockDim.x + threadIdx.x; i < n; i += (long long)gridDim.x * blockDim.x) {
        const int ch = (int)(i % c);
        long long q = i / c;
        const int xx = (int)(q % W2); q /= W2;
        const int yy = (int)(q % H2); q /= H2;
        const int zz = (int)(q % D2);
        const long long nn = q / D2;
        y[(i / c) * ycs + ch] = x[(((nn * D + zz / 2) * H + yy / 2) * W + xx / 2) * xcs + ch];
    }
}

__global__ void upsample_bwd(const uint16_t* __restrict__ dy, int dycs, uint16_t* __restrict__ dx, int dxcs, int nb, int D,
                             int H, int W, int c) {
    const long long n = (long long)nb * D * H * W * c;
    const int H2 = 2 * H, W2 = 2 * W;
    for (long long i = (long long)blockIdx.x * blockDim.x + threadIdx.x; i < n; i += (long long)gridDim.x * blockDim.x) {
        const int ch = (int)(i % c);
        long long q = i / c;
        const int xx = (int)(q % W); q /= W;
        const int yy = (int)(q % H); q /= H;
        const int zz = (int)(q % D);
        const long long nn = q / D;
        float s = 0.f;
        for (int k = 0; k < 8; k++) {
            const int z2 = 2 * zz + (k >> 2), y2 = 2 * yy + ((k >> 1) & 1), x2 = 2 * xx + (k & 1);
            s += bf2f(dy[(((nn * 2 * D + z2) * H2 + y2) * W2 + x2) * dycs + ch]);
        }
        dx[(i / c) * dxcs + ch] = f2bf(s);
    }
}

/* ---- 8-channel (16-byte) forms of the elementwise kernels, used when
 * channel counts, strides and pointers allow.  I is the index type:
 * uint32_t whenever every element index and address offset fits (the
 * common case: 32-bit division by the grid extents is a short sequence,
 * 64-bit division a long one), else long long. ------------------------- */
__device__ __forceinline__ void store8(uint16_t* p, const Frag& f) { *(uint4*)p = f.u; }

template <typename I>
__global__ void maxpool_fwd8(const uint16_t* __restrict__ x, int xcs, uint16_t* __restrict__ y, int ycs,
                             uint8_t* __restrict__ arg, int nb, int D, int H, int W, int c) {
    const int d2 = D / 2, h2 = H / 2, w2 = W / 2, c8 = c / 8;
    const I n = (I)nb * d2 * h2 * w2 * c8;
    for (I i = (I)blockIdx.x * blockDim.x + threadIdx.x; i < n; i += (I)gridDim.x * blockDim.x) {
        const int ch = (int)(i % c8) * 8;
        const I ov = i / c8;
        I q = ov;
        const int ox = (int)(q % w2); q /= w2;
        const int oy = (int)(q % h2); q /= h2;
        const int oz = (int)(q % d2);
        const I on = q / d2;
        float best[8];
        uint8_t bi[8];
#pragma unroll
        for (int j = 0; j < 8; j++) { best[j] = -3.0e38f; bi[j] = 0; }
        for (int k = 0; k < 8; k++) {
            const int zz = 2 * oz + (k >> 2), yy = 2 * oy + ((k >> 1) & 1), xx = 2 * ox + (k & 1);
            const Frag f = load8(x + (((on * D + zz) * H + yy) * W + xx) * xcs + ch);
#pragma unroll
            for (int j = 0; j < 8; j++) {
                const float v = bf2f(f.h[j]);
                if (v > best[j]) { best[j] = v; bi[j] = (uint8_t)k; }
            }
        }
        Frag o;
#pragma unroll
        for (int j = 0; j < 8; j++) o.h[j] = f2bf(best[j]);
        store8(y + ov * ycs + ch, o);
        uint2 a;
        a.x = bi[0] | (bi[1] << 8) | (bi[2] << 16) | ((uint32_t)bi[3] << 24);
        a.y = bi[4] | (bi[5] << 8) | (bi[6] << 16) | ((uint32_t)bi[7] << 24);
        *(uint2*)(arg + ov * c + ch) = a;
    }
}

template <typename I>
__global__ void maxpool_bwd8(const uint16_t* __restrict__ dy, int dycs, const uint8_t* __restrict__ arg,
                             uint16_t* __restrict__ dx, int dxcs, int nb, int D, int H, int W, int c) {
    const int d2 = D / 2, h2 = H / 2, w2 = W / 2, c8 = c / 8;
    const I n = (I)nb * D * H * W * c8;
    for (I i = (I)blockIdx.x * blockDim.x + threadIdx.x; i < n; i += (I)gridDim.x * blockDim.x) {
        const int ch = (int)(i % c8) * 8;
        const I v = i / c8;
        I q = v;
        const int xx = (int)(q % W); q /= W;
        const int yy = (int)(q % H); q /= H;
        const int zz = (int)(q % D);
        const I nn = q / D;
        const I o = ((nn * d2 + zz / 2) * h2 + yy / 2) * w2 + xx / 2;
        const int k = ((zz & 1) << 2) | ((yy & 1) << 1) | (xx & 1);
        const uint2 a = *(const uint2*)(arg + o * c + ch);
        const Frag g = load8(dy + o * dycs + ch);
        Frag r;
#pragma unroll
        for (int j = 0; j < 8; j++) {
            const uint32_t w = (j < 4) ? a.x : a.y;
            r.h[j] = (((w >> (8 * (j & 3))) & 0xFFu) == (uint32_t)k) ? g.h[j] : (uint16_t)0;
        }
        store8(dx + v * dxcs + ch, r);
    }
}

template <typename I>
__global__ void upsample_fwd8(const uint16_t* __restrict__ x, int xcs, uint16_t* __restrict__ y, int ycs, int nb, int D,
                              int H, int W, int c) {
    const int D2 = 2 * D, H2 = 2 * H, W2 = 2 * W, c8 = c / 8;
    const I n = (I)nb * D2 * H2 * W2 * c8;
    for (I i = (I)blockIdx.x * blockDim.x + threadIdx.x; i < n; i += (I)gridDim.x * blockDim.x) {
        const int ch = (int)(i % c8) * 8;
        const I v = i / c8;
        I q = v;
        const int xx = (int)(q % W2); q /= W2;
        const int yy = (int)(q % H2); q /= H2;
        const int zz = (int)(q % D2);
        const I nn = q / D2;
        store8(y + v * ycs + ch, load8(x + (((nn * D + zz / 2) * H + yy / 2) * W + xx / 2) * xcs + ch));
    }
}

template <typename I>
__global__ void upsample_bwd8(const uint16_t* __restrict__ dy, int dycs, uint16_t* __restrict__ dx, int dxcs, int nb, int D,
                              int H, int W, int c) {
    const int H2 = 2 * H, W2 = 2 * W, c8 = c / 8;
    const I n = (I)nb * D * H * W * c8;
    for (I i = (I)blockIdx.x * blockDim.x + threadIdx.x; i < n; i += (I)gridDim.x * blockDim.x) {
        const int ch = (int)(i % c8) * 8;
        const I v = i / c8;
        I q = v;
        const int xx = (int)(q % W); q /= W;
        const int yy = (int)(q % H); q /= H;
        const int zz = (int)(q % D);
        const I nn = q / D;
        float s8[8];
#pragma unroll
        for (int j = 0; j < 8; j++) s8[j] = 0.f;
        for (int k = 0; k < 8; k++) {
            const int z2 = 2 * zz + (k >> 2), y2 = 2 * yy + ((k >> 1) & 1), x2 = 2 * xx + (k & 1);
            const Frag f = load8(dy + (((nn * 2 * D + z2) * H2 + y2) * W2 + x2) * dycs + ch);
#pragma unroll
            for (int j = 0; j < 8; j++) s8[j] += bf2f(f.h[j]);
        }
        Frag r;
#pragma unroll
        for (int j = 0; j < 8; j++) r.h[j] = f2bf(s8[j]);
        store8(dx + v * dxcs + ch, r);
    }
}

template <typename I>
__global__ void add_bf16_8(const uint16_t* __restrict__ a, int acs, const uint16_t* __restrict__ b, int bcs,
                           uint16_t* __restrict__ out, int ocs, I nvox, int c) {
    const int c8 = c / 8;
    const I n = nvox * c8;
    for (I i = (I)blockIdx.x * blockDim.x + threadIdx.x; i < n; i += (I)gridDim.x * blockDim.x) {
        const I v = i / c8;
        const int ch = (int)(i % c8) * 8;
        const Frag fa = load8(a + v * acs + ch), fb = load8(b + v * bcs + ch);
        Frag r;
#pragma unroll
        for (int j = 0; j < 8; j++) r.h[j] = f2bf(bf2f(fa.h[j]) + bf2f(fb.h[j]));
        store8(out + v * ocs + ch, r);
    }
}

/* g[v][ch] = 0 where y[v][ch] is not > 0 (the ReLU derivative applied to
 * an output gradient in place), 8 channels per thread */
template <typename I>
__global__ void relu_mask_8(uint16_t* __restrict__ g, int gcs, const uint16_t* __restrict__ y, int ycs, I nvox, int c) {
    const int c8 = c / 8;
    const I n = nvox * c8;
    for (I i = (I)blockIdx.x * blockDim.x + threadIdx.x; i < n; i += (I)gridDim.x * blockDim.x) {
        const I v = i / c8;
        const int ch = (int)(i % c8) * 8;
        Frag fg = load8(g + v * gcs + ch);
        const Frag fy = load8(y + v * ycs + ch);
        fg.u.x &= keep_mask2(fy.u.x);
        fg.u.y &= keep_mask2(fy.u.y);
        fg.u.z &= keep_mask2(fy.u.z);
        fg.u.w &= keep_mask2(fy.u.w);
        store8(g + v * gcs + ch, fg);
    }
}

__global__ void relu_mask(uint16_t* __restrict__ g, int gcs, const uint16_t* __restrict__ y, int ycs, long long nvox, int c) {
    const long long n = nvox * c;
    for (long long i = (long long)blockIdx.x * blockDim.x + threadIdx.x; i < n; i += (long long)gridDim.x * blockDim.x) {
        const long long v = i / c;
        const int ch = (int)(i % c);
        if (!(bf2f(y[v * ycs + ch]) > 0.f)) g[v * gcs + ch] = 0;
    }
}

/* out = a + b (BF16, FP32 add) */
__global__ void add_bf16(const uint16_t* __restrict__ a, int acs, const uint16_t* __restrict__ b, int bcs,
                         uint16_t* __restrict__ out, int ocs, long long nvox, int c) {
    const long long n = nvox * c;
    for (long long i = (long long)blockIdx.x * blockDim.x + threadIdx.x; i < n; i += (long long)gridDim.x * blockDim.x) {
        const long long v = i / c;
        const int ch = (int)(i % c);
        out[v * ocs + ch] = f2bf(bf2f(a[v * acs + ch]) + bf2f(b[v * bcs + ch]));
    }
}

/* ---- output head: 1x1x1 conv to one channel + residual + MSE -------------
 * pred[v] = b + sum_c f[v][c] w[c] + res[v*rcs]; with a target, also
 * dpred[v] = 2 (pred - target) / nvox and per-block partial sums of the
 * squared error. */
__global__ void head_fwd(const uint16_t* __restrict__ f, int fcs, const float* __restrict__ w, const float* __restrict__ b,
                         const uint16_t* __restrict__ res, int rcs, const float* __restrict__ target, float* __restrict__ pred,
                         float* __restrict__ dpred, float* __restrict__ lpart, long long nvox, int c) {
    __shared__ float red[256];
    float l = 0.f;
    const bool vec = ((fcs | c) & 7) == 0 && ((uintptr_t)f & 15u) == 0;
    for (long long v = (long long)blockIdx.x * blockDim.x + threadIdx.x; v < nvox; v += (long long)gridDim.x * blockDim.x) {
        float s = b[0];
        if (vec) {
            /* 8 channels per 16-byte load, summed in channel order */
            for (int k = 0; k < c; k += 8) {
                const Frag q = load8(f + v * fcs + k);
#pragma unroll
                for (int j = 0; j < 8; j++) s += bf2f(q.h[j]) * w[k + j];
            }
        } else {
            for (int k = 0; k < c; k++) s += bf2f(f[v * fcs + k]) * w[k];
        }
        s += bf2f(res[v * rcs]);
        pred[v] = s;
        if (target) {
            /* a NaN target marks a voxel outside the loss */
            const float t = target[v];
            const float e = (t == t) ? s - t : 0.f;
            dpred[v] = 2.f * e / (float)nvox;
            l += e * e;
        }
    }
    red[threadIdx.x] = l;
    __syncthreads();
    for (int s2 = 128; s2 > 0; s2 >>= 1) {
        if (threadIdx.x < s2) red[threadIdx.x] += red[threadIdx.x + s2];
        __syncthreads();
    }
    if (threadIdx.x == 0 && lpart) lpart[blockIdx.x] = red[0];
}

/* df[v][c] = dpred[v] * w[c] (BF16); per-block partial sums of
 * dpred[v] * f[v][c] (dw) and dpred[v] (db): part[blk][c], part[blk][c] */
__global__ void head_bwd(const uint16_t* __restrict__ f, int fcs, const float* __restrict__ w, const float* __restrict__ dpred,
                         uint16_t* __restrict__ df, int dfcs, float* __restrict__ part, long long nvox, int c) {
    __shared__ float red[256][17];
    float acc[17];
    for (int k = 0; k <= c; k++) acc[k] = 0.f;
    for (long long v = (long long)blockIdx.x * blockDim.x + threadIdx.x; v < nvox; v += (long long)gridDim.x * blockDim.x) {
        const float d = dpred[v];
        for (int k = 0; k < c; k++) {
            acc[k] += d * bf2f(f[v * fcs + k]);
            df[v * dfcs + k] = f2bf(d * w[k]);
        }
        acc[c] += d;
    }
    for (int k = 0; k <= c; k++) red[threadIdx.x][k] = acc[k];
    __syncthreads();
    if (threadIdx.x <= c) {
        float t = 0.f;
        for (int i = 0; i < 256; i++) t += red[i][threadIdx.x];
        part[(long long)blockIdx.x * (c + 1) + threadIdx.x] = t;
    }
}

/* Network input features of a CW fluence map (grid units) on a label
 * volume, BF16 [v][8]: scaled log fluence (negative fluence counts as 0),
 * then (tissue voxels only) log mua
 * + 4, log mus, 10 (g - 0.8), 10 (n - 1.4) from the per-voxel media table
 * tab[label] = {mua, mus, g, n}, the tissue flag, two zeros.  flipz mirrors
 * the z axis (the output voxel (x, y, z) takes input (x, y, D-1-z)). */
__global__ void features(const float* __restrict__ phi, const uint8_t* __restrict__ vol, const float* __restrict__ tab,
                         int ntab, uint16_t* __restrict__ f, int nx, int ny, int nz, int flipz, float eps, float shift,
                         float scale) {
    const long long n = (long long)nx * ny * nz;
    for (long long o = (long long)blockIdx.x * blockDim.x + threadIdx.x; o < n; o += (long long)gridDim.x * blockDim.x) {
        /* output index o is x-slowest ([x][y][z]); the engine's arrays are
         * x-fastest ([z][y][x]) */
        const int z = (int)(o % nz), y = (int)((o / nz) % ny), x = (int)(o / ((long long)nz * ny));
        const int zs = flipz ? nz - 1 - z : z;
        const long long i = ((long long)zs * ny + y) * nx + x;
        const int lab = vol[i];
        float v[8];
        v[0] = (logf(fmaxf(phi[i], 0.f) + eps) + shift) / scale;
        const bool t = lab > 0 && lab < ntab;
        const float* m = tab + 4 * (t ? lab : 0);
        v[1] = t ? logf(fmaxf(m[0], 1e-6f)) + 4.f : 0.f;
        v[2] = t ? logf(fmaxf(m[1], 1e-6f)) : 0.f;
        v[3] = t ? (m[2] - 0.8f) * 10.f : 0.f;
        v[4] = t ? (m[3] - 1.4f) * 10.f : 0.f;
        v[5] = t ? 1.f : 0.f;
        v[6] = 0.f;
        v[7] = 0.f;
#pragma unroll
        for (int k = 0; k < 8; k++) f[o * 8 + k] = f2bf(v[k]);
    }
}

/* ---- signal box of a map (the region the network runs on) -------------
 * roi_max: out[k] = max(0, max over the nvox values of map k); grid (n, K),
 * out zeroed by the caller (non-negative floats order as their bits). */
__global__ void roi_max(const float* __restrict__ maps, long long nvox, float* __restrict__ out) {
    const float* m = maps + (long long)blockIdx.y * nvox;
    float v = 0.f;
    const long long n4 = nvox / 4;
    for (long long i = (long long)blockIdx.x * blockDim.x + threadIdx.x; i < n4; i += (long long)gridDim.x * blockDim.x) {
        const float4 q = ((const float4*)m)[i];
        v = fmaxf(v, fmaxf(fmaxf(q.x, q.y), fmaxf(q.z, q.w)));
    }
    for (long long i = n4 * 4 + (long long)blockIdx.x * blockDim.x + threadIdx.x; i < nvox; i += (long long)gridDim.x * blockDim.x)
        v = fmaxf(v, m[i]);
    for (int o = 32; o > 0; o >>= 1) v = fmaxf(v, __shfl_xor(v, o));
    if ((threadIdx.x & 63) == 0) atomicMax((int*)out + blockIdx.y, __float_as_int(v));
}

/* roi_rows: voxels of map k above thr[k]: rowcount[k][z*ny + y] per x row
 * and xcount[k][x] (added; zeroed by the caller).  One wave per row at a
 * time, each lane taking 4 consecutive x per 256-voxel stretch (nx a
 * multiple of 4, at most 1024); the row count is the popcount of the
 * wave's ballots, the x counts stay in registers until the end. */
__global__ void __launch_bounds__(256) roi_rows(const float* __restrict__ maps, int nx, int ny, int nz,
                                                const float* __restrict__ thr, unsigned int* __restrict__ xcount,
                                                unsigned int* __restrict__ rowcount) {
    const int k = blockIdx.y, lane = threadIdx.x & 63;
    const long long nrow = (long long)ny * nz;
    const float t = thr[k];
    const float* m = maps + (long long)k * nrow * nx;
    unsigned int xc[16];
#pragma unroll
    for (int j = 0; j < 16; j++) xc[j] = 0u;
    const long long wid = ((long long)blockIdx.x * blockDim.x + threadIdx.x) >> 6;
    const long long nw = ((long long)gridDim.x * blockDim.x) >> 6;
    for (long long row = wid; row < nrow; row += nw) {
        const float* r = m + row * nx;
        unsigned int cnt = 0u;
#pragma unroll
        for (int j = 0; j < 4; j++) {
            const int x = 256 * j + 4 * lane;
            float4 q = make_float4(0.f, 0.f, 0.f, 0.f);
            const bool in = x < nx;
            if (in) q = *(const float4*)(r + x);
            const bool a0 = in && q.x > t, a1 = in && q.y > t, a2 = in && q.z > t, a3 = in && q.w > t;
            xc[4 * j] += a0;
            xc[4 * j + 1] += a1;
            xc[4 * j + 2] += a2;
            xc[4 * j + 3] += a3;
            cnt += (unsigned int)(__popcll(__ballot(a0)) + __popcll(__ballot(a1)) + __popcll(__ballot(a2)) +
                                  __popcll(__ballot(a3)));
        }
        if (lane == 0) rowcount[(long long)k * nrow + row] = cnt;
    }
#pragma unroll
    for (int j = 0; j < 16; j++) {
        const int x = 256 * (j / 4) + 4 * lane + (j & 3);
        if (x < nx && xc[j]) atomicAdd(xcount + (long long)k * nx + x, xc[j]);
    }
}

/* fluence from a predicted scaled log fluence, back in the engine's
 * x-fastest order (with the same optional z mirror) */
__global__ void inv_features(const float* __restrict__ pred, float* __restrict__ phi, int nx, int ny, int nz, int flipz,
                             float eps, float shift, float scale) {
    const long long n = (long long)nx * ny * nz;
    for (long long o = (long long)blockIdx.x * blockDim.x + threadIdx.x; o < n; o += (long long)gridDim.x * blockDim.x) {
        const int z = (int)(o % nz), y = (int)((o / nz) % ny), x = (int)(o / ((long long)nz * ny));
        const int zs = flipz ? nz - 1 - z : z;
        const long long i = ((long long)zs * ny + y) * nx + x;
        phi[i] = fmaxf(expf(pred[o] * scale - shift) - eps, 0.f);
    }
}

/* Adam (Kingma & Ba) on FP32 master weights; optional BF16 copy */
__global__ void adam(float* __restrict__ p, float* __restrict__ m, float* __restrict__ v, const float* __restrict__ g,
                     uint16_t* __restrict__ pbf, long long n, float lr, float b1, float b2, float eps, float c1, float c2) {
    for (long long i = (long long)blockIdx.x * blockDim.x + threadIdx.x; i < n; i += (long long)gridDim.x * blockDim.x) {
        const float gi = g[i];
        const float mi = b1 * m[i] + (1.f - b1) * gi;
        const float vi = b2 * v[i] + (1.f - b2) * gi * gi;
        m[i] = mi;
        v[i] = vi;
        const float pi = p[i] - lr * (mi / c1) / (sqrtf(vi / c2) + eps);
        p[i] = pi;
        if (pbf) pbf[i] = f2bf(pi);
    }
}

__global__ void f32_to_bf16(const float* __restrict__ a, uint16_t* __restrict__ b, long long n) {
    for (long long i = (long long)blockIdx.x * blockDim.x + threadIdx.x; i < n; i += (long long)gridDim.x * blockDim.x)
        b[i] = f2bf(a[i]);
}

}  // namespace dnk

/* ---- C ABI ---------------------------------------------------------------
 * Every entry point validates its shapes and alignment on the host and
 * returns 0, or a negative code (-22 bad argument, otherwise -hipError). */
using namespace dnk;

static int dn_grid(long long n, int cap = 65536) {
    long long g = (n + 255) / 256;
    if (g < 1) g = 1;
    if (g > cap) g = cap;
    return (int)g;
}

static int dn_ret() {
    hipError_t e = hipGetLastError();
    return e == hipSuccess ? 0 : -(int)e;
}

static bool al16(const void* p) { return ((uintptr_t)p & 15u) == 0; }

/* every element offset of nvox voxels with channel strides up to
 * max(cs1, cs2) fits in 32 bits */
static bool fits32(long long nvox, int cs1, int cs2) {
    const long long cs = cs1 > cs2 ? cs1 : cs2;
    return nvox > 0 && cs > 0 && nvox <= (long long)0xFFFFFFFFu / (cs + 8);
}

/* 8 channels per thread: channel count and strides multiples of 8, both
 * pointers 16-byte aligned */
static bool vec8(int c, int s1, int s2, const void* p1, const void* p2) {
    return c % 8 == 0 && s1 % 8 == 0 && s2 % 8 == 0 && al16(p1) && al16(p2);
}

extern "C" int mcx_dn_conv3_fwd(const void* x, int xcs, const void* mask, int mcs, const void* w, int kpad, const float* bias,
                                void* y, int ycs, int nb, int d, int h, int wd, int cin, int cout, int relu, void* stream) {
    if (!x || !w || !y || nb <= 0 || d <= 0 || h <= 0 || wd <= 0) return -22;
    if (cin <= 0 || cin % 8 || cout <= 0 || cout % 16 || xcs < cin || xcs % 8 || ycs < cout) return -22;
    if (kpad < 27 * cin || kpad % 32 || !al16(x) || !al16(w) || (mask && (!al16(mask) || mcs % 8 || mcs < cin))) return -22;
    const long long m = (long long)nb * d * h * wd;
    hipStream_t st = (hipStream_t)stream;
    const uint16_t *xp = (const uint16_t*)x, *mp = (const uint16_t*)mask, *wp = (const uint16_t*)w;
    uint16_t* yp = (uint16_t*)y;
    const long long slab = 6LL * h * wd * std::max(xcs, mask ? mcs : 0); /* 32-bit halo offsets */
    if (cout <= 32 && (cin == 8 || cin % 16 == 0) && slab < 0x7fffffffLL) {
        /* halo-tiled kernel: 16 x 4 x 4 output voxels per workgroup */
        const long long nblk = (long long)((wd + 15) / 16) * ((h + 3) / 4) * ((d + 3) / 4) * nb;
        if (nblk > 0x7fffffffLL) return -22;
#define DN_HALO(CG, NT)                                                                                            \
    hipLaunchKernelGGL((conv3_fwd_halo<CG, NT>), dim3((unsigned int)nblk), dim3(256), 0, st, xp, xcs, mp, mcs, wp, kpad, \
                       bias, yp, ycs, nb, d, h, wd, cin, cout, relu, (const uint16_t*)nullptr, 0, 0)
        if (cin == 8) {
            if (cout == 16) DN_HALO(8, 1);
            else DN_HALO(8, 2);
        } else {
            if (cout == 16) DN_HALO(16, 1);
            else DN_HALO(16, 2);
        }
        return dn_ret();
    }
    if ((cout == 48 || cout == 64) && slab < 0x7fffffffLL) {
        /* 48 or 64 output channels: 8-channel groups keep the weights
         * (27.6 KB at 64 channels) and the halo within 64 KB of LDS */
        const long long nblk = (long long)((wd + 15) / 16) * ((h + 3) / 4) * ((d + 3) / 4) * nb;
        if (nblk > 0x7fffffffLL) return -22;
        if (cout == 48) DN_HALO(8, 3);
        else DN_HALO(8, 4);
#undef DN_HALO
        return dn_ret();
    }
    /* voxel tiles per wave: more where the weight fragments are few */
#define DN_FWD(MT, NT)                                                                                              \
    hipLaunchKernelGGL((conv3_fwd<MT, NT>), dim3((unsigned int)((m + 64 * MT - 1) / (64 * MT))), dim3(256), 0, st, xp, \
                       xcs, mp, mcs, wp, kpad, bias, yp, ycs, nb, d, h, wd, cin, cout, relu)
    switch (cout / 16) {
        case 1: DN_FWD(4, 1); break;
        case 2: DN_FWD(4, 2); break;
        case 3: DN_FWD(2, 3); break;
        case 4: DN_FWD(2, 4); break;
        case 6: DN_FWD(2, 6); break;
        case 8: DN_FWD(2, 8); break;
        case 12: DN_FWD(1, 12); break;
        default: return -22;
    }
#undef DN_FWD
    return dn_ret();
}

/* y = act(conv3(cat(up2(xu), x))): the convolution of a channel
 * concatenation whose first cu channels are the 2x nearest upsampling of
 * xu (dimensions d/2, h/2, wd/2, channel stride xucs) and whose other
 * cin - cu channels are x (dimensions d, h, wd).  Same result as
 * mcx_dn_upsample_fwd into the first cu channels of a cin-channel buffer,
 * x copied into the rest, then mcx_dn_conv3_fwd, without writing the
 * concatenation.  Output channels at most 32; cu and cin - cu multiples
 * of 16; d, h, wd even. */
extern "C" int mcx_dn_conv3_fwd_cat(const void* xu, int xucs, int cu, const void* x, int xcs, const void* w, int kpad,
                                    const float* bias, void* y, int ycs, int nb, int d, int h, int wd, int cin, int cout,
                                    int relu, void* stream) {
    if (!xu || !x || !w || !y || nb <= 0 || d <= 0 || h <= 0 || wd <= 0 || d % 2 || h % 2 || wd % 2) return -22;
    if (cu <= 0 || cu % 16 || cin <= cu || (cin - cu) % 16 || cout <= 0 || cout > 32 || cout % 16) return -22;
    if (xucs < cu || xucs % 8 || xcs < cin - cu || xcs % 8 || ycs < cout) return -22;
    if (kpad < 27 * cin || kpad % 32 || !al16(x) || !al16(xu) || !al16(w)) return -22;
    const long long slab = 6LL * h * wd * xcs;
    const long long nblk = (long long)((wd + 15) / 16) * ((h + 3) / 4) * ((d + 3) / 4) * nb;
    if (slab >= 0x7fffffffLL || nblk > 0x7fffffffLL) return -22;
    hipStream_t st = (hipStream_t)stream;
    const uint16_t *xp = (const uint16_t*)x, *up = (const uint16_t*)xu, *wp = (const uint16_t*)w;
    uint16_t* yp = (uint16_t*)y;
    if (cout == 16)
        hipLaunchKernelGGL((conv3_fwd_halo<16, 1>), dim3((unsigned int)nblk), dim3(256), 0, st, xp, xcs,
                           (const uint16_t*)nullptr, 0, wp, kpad, bias, yp, ycs, nb, d, h, wd, cin, cout, relu, up, xucs, cu);
    else
        hipLaunchKernelGGL((conv3_fwd_halo<16, 2>), dim3((unsigned int)nblk), dim3(256), 0, st, xp, xcs,
                           (const uint16_t*)nullptr, 0, wp, kpad, bias, yp, ycs, nb, d, h, wd, cin, cout, relu, up, xucs, cu);
    return dn_ret();
}

extern "C" int mcx_dn_weight_flip(const void* w, int kpad, void* wt, int kpadt, int cin, int cout, void* stream) {
    if (!w || !wt || kpad < 27 * cin || kpadt < 27 * cout || cin <= 0 || cout <= 0) return -22;
    const long long n = (long long)cin * kpadt;
    hipLaunchKernelGGL(weight_flip, dim3(dn_grid(n)), dim3(256), 0, (hipStream_t)stream, (const uint16_t*)w, kpad, (uint16_t*)wt,
                       kpadt, cin, cout);
    return dn_ret();
}

/* layers whose weight gradient runs as conv3_wgrad_rows: 16, 32 or 64
 * output channels and rows of at least 32 voxels (shorter rows leave
 * most of each 32-voxel segment empty; the per-tap kernel packs voxels
 * regardless of row length and wins there) */
static bool wgrad_rows_ok(int cin, int cout, int wd) {
    return (cout == 16 || cout == 32 || cout == 64) && wd >= 32 && cin > 0;
}

extern "C" int mcx_dn_conv3_wgrad(const void* x, int xcs, const void* dy, int dycs, const void* mask, int mcs, float* partial,
                                  int nsplit, int nb, int d, int h, int wd, int cin, int cout, void* stream) {
    if (!x || !dy || !partial || nsplit <= 0 || nsplit > 65535 || nb <= 0 || d <= 0 || h <= 0 || wd <= 0) return -22;
    if (cin <= 0 || cin % 8 || cout <= 0 || cout % 16 || xcs % 8 || dycs % 8 || xcs < cin || dycs < cout) return -22;
    if (!al16(x) || !al16(dy) || (mask && (!al16(mask) || mcs % 8 || mcs < cout))) return -22;
    const long long m = (long long)nb * d * h * wd;
    long long vps = (m + nsplit - 1) / nsplit;
    vps = (vps + 31) / 32 * 32;
    const int cinp = (cin + 15) & ~15;
    const int ntile = (cout / 16) * (cinp / 16);
    const int tpw = (ntile + 3) / 4;
    const size_t shmem = (size_t)(cout + cinp) * 32 * 2;
    if (shmem > 65536) return -22;
    hipStream_t st = (hipStream_t)stream;
    const uint16_t *xp = (const uint16_t*)x, *dp = (const uint16_t*)dy, *mp = (const uint16_t*)mask;
    if (wgrad_rows_ok(cin, cout, wd)) {
        /* all 27 taps per staging of 32-voxel x-row segments; input tiles
         * split over blockIdx.y */
        const long long nseg = (long long)nb * d * h * ((wd + 31) / 32);
        const long long segs = (nseg + nsplit - 1) / nsplit;
        if (segs > 0x7fffffffLL) return -22;
        const int tiles = cinp / 16;
#define DN_WGR(NCI, NCO)                                                                                        \
    hipLaunchKernelGGL((conv3_wgrad_rows<NCI, NCO>), dim3((unsigned int)nsplit, (unsigned int)(tiles / NCI)), dim3(256), 0, \
                       st, xp, xcs, dp, dycs, mp, mcs, partial, (int)segs, nb, d, h, wd, cin)
        if (cout == 16)
            DN_WGR(1, 1);
        else if (cout == 32 && tiles % 2 == 0)
            DN_WGR(2, 2);
        else if (cout == 32)
            DN_WGR(1, 2);
        else
            DN_WGR(1, 4);
#undef DN_WGR
        return dn_ret();
    }
    const dim3 grid(27, (unsigned int)nsplit);
#define DN_WG(T) hipLaunchKernelGGL(conv3_wgrad<T>, grid, dim3(256), shmem, st, xp, xcs, dp, dycs, mp, mcs, partial, (int)vps, nb, d, h, wd, cin, cout)
    if (tpw <= 1) DN_WG(1);
    else if (tpw <= 2) DN_WG(2);
    else if (tpw <= 4) DN_WG(4);
    else if (tpw <= 8) DN_WG(8);
    else if (tpw <= 16) DN_WG(16);
    else if (tpw <= 24) DN_WG(24);
    else return -22;
#undef DN_WG
    return dn_ret();
}

extern "C" int mcx_dn_wgrad_reduce(const float* partial, int nsplit, float* dw, int kpad, int cin, int cout, void* stream) {
    if (!partial || !dw || nsplit <= 0 || kpad < 27 * cin || cin <= 0 || cout <= 0) return -22;
    const long long nel = (long long)cout * kpad;
    hipLaunchKernelGGL(wgrad_reduce, dim3((unsigned int)((nel + 63) / 64)), dim3(1024), 0, (hipStream_t)stream, partial, nsplit,
                       dw, kpad, cin, cout);
    return dn_ret();
}

/* part must hold nsplit*cout floats; db receives cout sums */
extern "C" int mcx_dn_bias_grad(const void* dy, int dycs, const void* mask, int mcs, float* part, int nsplit, float* db,
                                long long nvox, int cout, void* stream) {
    if (!dy || !part || !db || nsplit <= 0 || nsplit > 65535 || cout <= 0 || cout > 256 || cout % 8 || 256 % (cout / 8) ||
        nvox <= 0 || dycs % 8 || !al16(dy) || (mask && (mcs % 8 || !al16(mask))))
        return -22;
    const int vps = (int)((nvox + nsplit - 1) / nsplit);
    hipStream_t st = (hipStream_t)stream;
    hipLaunchKernelGGL(bias_grad, dim3(nsplit), dim3(256), 0, st, (const uint16_t*)dy, dycs, (const uint16_t*)mask, mcs, part, nvox,
                       vps, cout);
    hipLaunchKernelGGL(sum_rows, dim3((cout + 63) / 64), dim3(1024), 0, st, (const float*)part, nsplit, cout, db);
    return dn_ret();
}

/* ---- bias gradient without allocation ------------------------------------
 * The per-range partial sums of mcx_dn_bias_grad live in this fixed device
 * array instead of a caller workspace, for callers that cannot allocate.
 * Each GPU holds its own copy until the process ends; host code reaches it
 * only through the address hipGetSymbolAddress returns for the current
 * device (the array's name is a host-side stand-in). */
#define MCX_DN_SCRATCH_FLOATS 65536
#define MCX_DN_MAX_DEVICES 64
static __device__ __align__(256) float scratch[MCX_DN_SCRATCH_FLOATS];
static float* g_scratch_dev[MCX_DN_MAX_DEVICES];
static std::mutex g_scratch_mu;

/* this device's copy of `scratch`, looked up once per device; NULL on error */
static float* scratch_ptr(int dev) {
    std::lock_guard<std::mutex> lk(g_scratch_mu);
    if (!g_scratch_dev[dev]) {
        void* p = nullptr;
        if (hipGetSymbolAddress(&p, HIP_SYMBOL(scratch)) != hipSuccess || !p) return nullptr;
        g_scratch_dev[dev] = (float*)p;
    }
    return g_scratch_dev[dev];
}

extern "C" int mcx_dn_bias_grad_noalloc(const void* dy, int dycs, const void* mask, int mcs, int nsplit, float* db, long long nvox,
                                        int cout, void* stream) {
    if (!dy || !db || nsplit <= 0 || nsplit > 65535 || cout <= 0 || cout > 256 || cout % 8 || 256 % (cout / 8) || nvox <= 0 ||
        dycs % 8 || !al16(dy) || (mask && (mcs % 8 || !al16(mask))))
        return -22;
    if ((long long)nsplit * cout > MCX_DN_SCRATCH_FLOATS) return -7; /* E2BIG: beyond the fixed scratch array */
    int dev = 0;
    if (hipGetDevice(&dev) != hipSuccess || dev < 0 || dev >= MCX_DN_MAX_DEVICES) return -22;
    float* part = scratch_ptr(dev);
    if (!part) return -(int)hipErrorInvalidSymbol;
    const int vps = (int)((nvox + nsplit - 1) / nsplit);
    hipStream_t st = (hipStream_t)stream;
    /* bias_grad writes every one of the nsplit x cout partial sums that
     * sum_rows reads: nothing to clear */
    hipLaunchKernelGGL(bias_grad, dim3(nsplit), dim3(256), 0, st, (const uint16_t*)dy, dycs, (const uint16_t*)mask, mcs, part, nvox,
                       vps, cout);
    hipLaunchKernelGGL(sum_rows, dim3((cout + 63) / 64), dim3(1024), 0, st, (const float*)part, nsplit, cout, db);
    return dn_ret();
}

/* launch an 8-channel kernel with 32-bit indices when they fit */
#define DN_LAUNCH8(K, small, n8, ...)                                                                          \
    do {                                                                                                       \
        if (small)                                                                                             \
            hipLaunchKernelGGL(K<uint32_t>, dim3(dn_grid(n8)), dim3(256), 0, (hipStream_t)stream, __VA_ARGS__); \
        else                                                                                                   \
            hipLaunchKernelGGL(K<long long>, dim3(dn_grid(n8)), dim3(256), 0, (hipStream_t)stream, __VA_ARGS__); \
    } while (0)

extern "C" int mcx_dn_maxpool_fwd(const void* x, int xcs, void* y, int ycs, unsigned char* arg, int nb, int d, int h, int wd,
                                  int c, void* stream) {
    if (!x || !y || !arg || d % 2 || h % 2 || wd % 2 || c <= 0 || xcs < c || ycs < c) return -22;
    const long long n = (long long)nb * (d / 2) * (h / 2) * (wd / 2) * c;
    if (vec8(c, xcs, ycs, x, y) && ((uintptr_t)arg & 7u) == 0)
        DN_LAUNCH8(maxpool_fwd8, fits32((long long)nb * d * h * wd, xcs, ycs), n / 8, (const uint16_t*)x, xcs, (uint16_t*)y, ycs,
                   (uint8_t*)arg, nb, d, h, wd, c);
    else
        hipLaunchKernelGGL(maxpool_fwd, dim3(dn_grid(n)), dim3(256), 0, (hipStream_t)stream, (const uint16_t*)x, xcs, (uint16_t*)y,
                           ycs, (uint8_t*)arg, nb, d, h, wd, c);
    return dn_ret();
}

extern "C" int mcx_dn_maxpool_bwd(const void* dy, int dycs, const unsigned char* arg, void* dx, int dxcs, int nb, int d, int h,
                                  int wd, int c, void* stream) {
    if (!dy || !dx || !arg || d % 2 || h % 2 || wd % 2 || c <= 0 || dycs < c || dxcs < c) return -22;
    const long long n = (long long)nb * d * h * wd * c;
    if (vec8(c, dycs, dxcs, dy, dx) && ((uintptr_t)arg & 7u) == 0)
        DN_LAUNCH8(maxpool_bwd8, fits32((long long)nb * d * h * wd, dycs, dxcs), n / 8, (const uint16_t*)dy, dycs,
                   (const uint8_t*)arg, (uint16_t*)dx, dxcs, nb, d, h, wd, c);
    else
        hipLaunchKernelGGL(maxpool_bwd, dim3(dn_grid(n)), dim3(256), 0, (hipStream_t)stream, (const uint16_t*)dy, dycs,
                           (const uint8_t*)arg, (uint16_t*)dx, dxcs, nb, d, h, wd, c);
    return dn_ret();
}

/* d, h, wd: the low-resolution dimensions */
extern "C" int mcx_dn_upsample_fwd(const void* x, int xcs, void* y, int ycs, int nb, int d, int h, int wd, int c, void* stream) {
    if (!x || !y || c <= 0 || xcs < c || ycs < c) return -22;
    const long long n = (long long)nb * 8 * d * h * wd * c;
    if (vec8(c, xcs, ycs, x, y))
        DN_LAUNCH8(upsample_fwd8, fits32((long long)nb * 8 * d * h * wd, xcs, ycs), n / 8, (const uint16_t*)x, xcs, (uint16_t*)y,
                   ycs, nb, d, h, wd, c);
    else
        hipLaunchKernelGGL(upsample_fwd, dim3(dn_grid(n)), dim3(256), 0, (hipStream_t)stream, (const uint16_t*)x, xcs, (uint16_t*)y,
                           ycs, nb, d, h, wd, c);
    return dn_ret();
}

extern "C" int mcx_dn_upsample_bwd(const void* dy, int dycs, void* dx, int dxcs, int nb, int d, int h, int wd, int c, void* stream) {
    if (!dy || !dx || c <= 0 || dycs < c || dxcs < c) return -22;
    const long long n = (long long)nb * d * h * wd * c;
    if (vec8(c, dycs, dxcs, dy, dx))
        DN_LAUNCH8(upsample_bwd8, fits32((long long)nb * 8 * d * h * wd, dycs, dxcs), n / 8, (const uint16_t*)dy, dycs,
                   (uint16_t*)dx, dxcs, nb, d, h, wd, c);
    else
        hipLaunchKernelGGL(upsample_bwd, dim3(dn_grid(n)), dim3(256), 0, (hipStream_t)stream, (const uint16_t*)dy, dycs,
                           (uint16_t*)dx, dxcs, nb, d, h, wd, c);
    return dn_ret();
}

extern "C" int mcx_dn_add(const void* a, int acs, const void* b, int bcs, void* out, int ocs, long long nvox, int c, void* stream) {
    if (!a || !b || !out || c <= 0 || nvox <= 0) return -22;
    if (vec8(c, acs, bcs, a, b) && ocs % 8 == 0 && al16(out))
        DN_LAUNCH8(add_bf16_8, fits32(nvox, acs, bcs > ocs ? bcs : ocs), nvox * c / 8, (const uint16_t*)a, acs,
                   (const uint16_t*)b, bcs, (uint16_t*)out, ocs, nvox, c);
    else
        hipLaunchKernelGGL(add_bf16, dim3(dn_grid(nvox * c)), dim3(256), 0, (hipStream_t)stream, (const uint16_t*)a, acs,
                           (const uint16_t*)b, bcs, (uint16_t*)out, ocs, nvox, c);
    return dn_ret();
}
extern "C" int mcx_dn_roi_max(const float* maps, int k, long long nvox, float* out, void* stream) {
    if (!maps || !out || k <= 0 || k > 65535 || nvox <= 0 || !al16(maps) || (nvox & 3)) return -22;
    const long long per = (nvox / 4 + 255) / 256;
    const unsigned int nbx = (unsigned int)std::min<long long>(per, 256);
    hipLaunchKernelGGL(roi_max, dim3(nbx, (unsigned int)k), dim3(256), 0, (hipStream_t)stream, maps, nvox, out);
    return dn_ret();
}

extern "C" int mcx_dn_roi_counts(const float* maps, int k, int nx, int ny, int nz, const float* thr,
                                 unsigned int* xcount, unsigned int* rowcount, void* stream) {
    if (!maps || !thr || !xcount || !rowcount || k <= 0 || k > 65535 || nx <= 0 || ny <= 0 || nz <= 0) return -22;
    if ((nx & 3) || nx > 1024 || !al16(maps)) return -22;
    const long long nrow = (long long)ny * nz;
    const unsigned int nbx = (unsigned int)std::min<long long>((nrow + 3) / 4, 64);
    hipLaunchKernelGGL(roi_rows, dim3(nbx, (unsigned int)k), dim3(256), 0, (hipStream_t)stream, maps, nx, ny, nz, thr,
                       xcount, rowcount);
    return dn_ret();
}

extern "C" int mcx_dn_relu_mask(void* g, int gcs, const void* y, int ycs, long long nvox, int c, void* stream) {
    if (!g || !y || c <= 0 || nvox <= 0 || gcs < c || ycs < c) return -22;
    if (vec8(c, gcs, ycs, g, y))
        DN_LAUNCH8(relu_mask_8, fits32(nvox, gcs, ycs), nvox * c / 8, (uint16_t*)g, gcs, (const uint16_t*)y, ycs, nvox, c);
    else
        hipLaunchKernelGGL(relu_mask, dim3(dn_grid(nvox * c)), dim3(256), 0, (hipStream_t)stream, (uint16_t*)g, gcs,
                           (const uint16_t*)y, ycs, nvox, c);
    return dn_ret();
}
#undef DN_LAUNCH8

/* lpart: gridsize floats (gridsize returned through *nblocks) */
extern "C" int mcx_dn_head_fwd(const void* f, int fcs, const float* w, const float* b, const void* res, int rcs,
                               const float* target, float* pred, float* dpred, float* lpart, int nblocks, long long nvox, int c,
                               void* stream) {
    if (!f || !w || !b || !res || !pred || c <= 0 || c > 16 || nvox <= 0 || nblocks <= 0) return -22;
    if (target && (!dpred || !lpart)) return -22;
    hipLaunchKernelGGL(head_fwd, dim3(nblocks), dim3(256), 0, (hipStream_t)stream, (const uint16_t*)f, fcs, w, b,
                       (const uint16_t*)res, rcs, target, pred, dpred, lpart, nvox, c);
    return dn_ret();
}

/* part: nblocks*(c+1) floats; dwb receives c weight sums then the bias sum */
extern "C" int mcx_dn_head_bwd(const void* f, int fcs, const float* w, const float* dpred, void* df, int dfcs, float* part,
                               int nblocks, float* dwb, long long nvox, int c, void* stream) {
    if (!f || !w || !dpred || !df || !part || !dwb || c <= 0 || c > 16 || nblocks <= 0 || nvox <= 0) return -22;
    hipStream_t st = (hipStream_t)stream;
    hipLaunchKernelGGL(head_bwd, dim3(nblocks), dim3(256), 0, st, (const uint16_t*)f, fcs, w, dpred, (uint16_t*)df, dfcs, part,
                       nvox, c);
    hipLaunchKernelGGL(sum_rows, dim3((c + 1 + 63) / 64), dim3(1024), 0, st, (const float*)part, nblocks, c + 1, dwb);
    return dn_ret();
}

extern "C" int mcx_dn_adam(float* p, float* m, float* v, const float* g, void* pbf16, long long n, float lr, float beta1,
                           float beta2, float eps, int step, void* stream) {
    if (!p || !m || !v || !g || n <= 0 || step <= 0) return -22;
    const float c1 = 1.f - powf(beta1, (float)step), c2 = 1.f - powf(beta2, (float)step);
    hipLaunchKernelGGL(adam, dim3(dn_grid(n, 4096)), dim3(256), 0, (hipStream_t)stream, p, m, v, g, (uint16_t*)pbf16, n, lr, beta1,
                       beta2, eps, c1, c2);
    return dn_ret();
}

extern "C" int mcx_dn_f32_to_bf16(const float* a, void* b, long long n, void* stream) {
    if (!a || !b || n <= 0) return -22;
    hipLaunchKernelGGL(f32_to_bf16, dim3(dn_grid(n, 4096)), dim3(256), 0, (hipStream_t)stream, a, (uint16_t*)b, n);
    return dn_ret();
}

/* phi: FP32 CW fluence in the engine's x-fastest order; vol: uint8 labels
 * (same order); tab: ntab x {mua, mus, g, n} per voxel; f: BF16 features
 * [nx][ny][nz][8] (the network's NDHWC order with D = x) */
extern "C" int mcx_dn_features(const float* phi, const unsigned char* vol, const float* tab, int ntab, void* f, int nx, int ny,
                               int nz, int flipz, float eps, float shift, float scale, void* stream) {
    if (!phi || !vol || !tab || !f || ntab <= 0 || nx <= 0 || ny <= 0 || nz <= 0 || scale == 0.f) return -22;
    const long long n = (long long)nx * ny * nz;
    hipLaunchKernelGGL(features, dim3(dn_grid(n)), dim3(256), 0, (hipStream_t)stream, phi, (const uint8_t*)vol, tab, ntab,
                       (uint16_t*)f, nx, ny, nz, flipz, eps, shift, scale);
    return dn_ret();
}

extern "C" int mcx_dn_inv_features(const float* pred, float* phi, int nx, int ny, int nz, int flipz, float eps, float shift,
                                   float scale, void* stream) {
    if (!pred || !phi || nx <= 0 || ny <= 0 || nz <= 0) return -22;
    const long long n = (long long)nx * ny * nz;
    hipLaunchKernelGGL(inv_features, dim3(dn_grid(n)), dim3(256), 0, (hipStream_t)stream, pred, phi, nx, ny, nz, flipz, eps,
                       shift, scale);
    return dn_ret();
}
