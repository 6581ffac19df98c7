/*
 * mcx_amd_denoise.h -- kernels of the learned fluence denoiser (3D U-Net):
 * forward, backward and optimiser steps, BF16 activations with FP32
 * accumulation on the MI355X matrix cores.
 *
 * None of these has a counterpart in the reference (which ships no learned
 * denoiser); they serve the denoised low-photon path beside
 * mcx_run_simulation / mcx_run_batch.
 *
 * Conventions: BF16 tensors are raw 16-bit words in NDHWC order (channels
 * innermost); *cs arguments are channel strides (elements between
 * consecutive voxels), so a pointer may address a channel slice of a wider
 * buffer.  BF16 pointers handed to the 3x3x3 convolution kernels must be
 * 16-byte aligned with channel counts, strides and slice offsets multiples
 * of 8.  Conv weights are BF16 [cout][kpad], k = tap*cin + ci, tap =
 * 9*(dz+1) + 3*(dy+1) + (dx+1), kpad = 27*cin rounded up to 32 (zero
 * padded).  `stream` is a hipStream_t (NULL: default stream).  Every
 * function returns 0, -22 for an invalid argument, or -hipError.
 */
#ifndef MCX_AMD_DENOISE_H
#define MCX_AMD_DENOISE_H

#ifdef __cplusplus
extern "C" {
#endif

/* y = act(bias + conv3x3x3(x * (mask > 0))), act = ReLU when relu != 0;
 * mask (NULL: none) has x's layout.  cout multiple of 16. */
int mcx_dn_conv3_fwd(const void* x, int xcs, const void* mask, int mcs, const void* w, int kpad, const float* bias,
                     void* y, int ycs, int nb, int d, int h, int wd, int cin, int cout, int relu, void* stream);
/* y = act(bias + conv3x3x3(cat(up2(xu), x))): the first cu input channels
 * are the 2x nearest upsampling of xu (d/2 x h/2 x wd/2 voxels, stride
 * xucs), the other cin - cu are x; the concatenation is never written.
 * cout 16 or 32, cu and cin - cu multiples of 16, d, h, wd even. */
int mcx_dn_conv3_fwd_cat(const void* xu, int xucs, int cu, const void* x, int xcs, const void* w, int kpad,
                         const float* bias, void* y, int ycs, int nb, int d, int h, int wd, int cin, int cout, int relu,
                         void* stream);
/* wt[ci][(26 - tap)*cout + co] = w[co][tap*cin + ci]: the weights of the
 * input-gradient convolution (run with mcx_dn_conv3_fwd) */
int mcx_dn_weight_flip(const void* w, int kpad, void* wt, int kpadt, int cin, int cout, void* stream);
/* partial[s][tap][co][ci] over nsplit voxel ranges of
 * dy[v][co] * (mask[v][co] > 0) * x[v + offset(tap)][ci] */
int mcx_dn_conv3_wgrad(const void* x, int xcs, const void* dy, int dycs, const void* mask, int mcs, float* partial,
                       int nsplit, int nb, int d, int h, int wd, int cin, int cout, void* stream);
/* dw[co][tap*cin + ci] (FP32, kpad columns, padding zeroed) = sum over s */
int mcx_dn_wgrad_reduce(const float* partial, int nsplit, float* dw, int kpad, int cin, int cout, void* stream);
/* db[co] = sum_v dy[v][co] * (mask[v][co] > 0); part: nsplit*cout floats */
int mcx_dn_bias_grad(const void* dy, int dycs, const void* mask, int mcs, float* part, int nsplit, float* db,
                     long long nvox, int cout, void* stream);
/* mcx_dn_bias_grad without a workspace argument and without allocating:
 * the nsplit x cout partial sums live in a fixed device array of 65536
 * floats (256 KB on each GPU the library runs on, held until the process
 * ends).  No counterpart in the reference; for callers that cannot
 * allocate (code under HIP graph capture, tight serving or training loops).
 * Returns -7 when nsplit * cout exceeds 65536.  Calls that overlap on one
 * GPU share the array and race; each call writes every partial sum it
 * reads, so nothing carries over from the previous call. */
int mcx_dn_bias_grad_noalloc(const void* dy, int dycs, const void* mask, int mcs, int nsplit, float* db, long long nvox,
                             int cout, void* stream);
/* 2x2x2 max pooling (d, h, wd even: input dims); arg: window index of the
 * maximum per output element */
int mcx_dn_maxpool_fwd(const void* x, int xcs, void* y, int ycs, unsigned char* arg, int nb, int d, int h, int wd, int c,
                       void* stream);
int mcx_dn_maxpool_bwd(const void* dy, int dycs, const unsigned char* arg, void* dx, int dxcs, int nb, int d, int h, int wd,
                       int c, void* stream);
/* nearest-neighbour x2 upsampling; d, h, wd: low-resolution dims */
int mcx_dn_upsample_fwd(const void* x, int xcs, void* y, int ycs, int nb, int d, int h, int wd, int c, void* stream);
int mcx_dn_upsample_bwd(const void* dy, int dycs, void* dx, int dxcs, int nb, int d, int h, int wd, int c, void* stream);
/* Signal box of k maps of nvox floats each (x fastest): out[k] = max(0,
 * the map's maximum), out zeroed by the caller; nvox a multiple of 4 and
 * maps 16-byte aligned */
int mcx_dn_roi_max(const float* maps, int k, long long nvox, float* out, void* stream);
/* voxels of map k above thr[k]: rowcount[k][z*ny + y] per x row and
 * xcount[k][x] added (zeroed by the caller); nx a multiple of 4, at most
 * 1024 */
int mcx_dn_roi_counts(const float* maps, int k, int nx, int ny, int nz, const float* thr, unsigned int* xcount,
                      unsigned int* rowcount, void* stream);
/* g[v][ch] = 0 where y[v][ch] is not > 0, in place: the ReLU derivative
 * applied once to an output gradient, after which the weight, bias and
 * input gradients take no mask (same results as passing y as their mask) */
int mcx_dn_relu_mask(void* g, int gcs, const void* y, int ycs, long long nvox, int c, void* stream);
/* out = a + b */
int mcx_dn_add(const void* a, int acs, const void* b, int bcs, void* out, int ocs, long long nvox, int c, void* stream);
/* output head: pred[v] = b + sum_c f[v][c] w[c] + res[v*rcs] (FP32); with a
 * target also dpred = 2 (pred - target) / nvox and lpart[blk] = partial sums
 * of the squared error over nblocks blocks; NaN targets are left out */
int mcx_dn_head_fwd(const void* f, int fcs, const float* w, const float* b, const void* res, int rcs, const float* target,
                    float* pred, float* dpred, float* lpart, int nblocks, long long nvox, int c, void* stream);
/* df[v][c] = dpred[v] w[c]; dwb[0..c-1] = sum_v dpred f, dwb[c] = sum_v dpred;
 * part: nblocks*(c+1) floats */
int mcx_dn_head_bwd(const void* f, int fcs, const float* w, const float* dpred, void* df, int dfcs, float* part, int nblocks,
                    float* dwb, long long nvox, int c, void* stream);
/* Adam step `step` (1-based) on FP32 p with moments m, v; pbf16 (NULL: none)
 * receives the BF16 copy of p */
int mcx_dn_adam(float* p, float* m, float* v, const float* g, void* pbf16, long long n, float lr, float beta1, float beta2,
                float eps, int step, void* stream);
int mcx_dn_f32_to_bf16(const float* a, void* b, long long n, void* stream);
/* network input features of a CW fluence map: phi (FP32) and vol (uint8
 * labels) in the engine's x-fastest order, tab = ntab x {mua, mus, g, n}
 * per voxel; f = BF16 [nx][ny][nz][8]: (log(phi + eps) + shift) / scale,
 * log mua + 4, log mus, 10 (g - 0.8), 10 (n - 1.4), tissue flag, 0, 0.
 * flipz mirrors the z axis. */
int mcx_dn_features(const float* phi, const unsigned char* vol, const float* tab, int ntab, void* f, int nx, int ny, int nz,
                    int flipz, float eps, float shift, float scale, void* stream);
/* phi (x-fastest) = max(exp(pred * scale - shift) - eps, 0) from a
 * prediction in the [nx][ny][nz] order of the features */
int mcx_dn_inv_features(const float* pred, float* phi, int nx, int ny, int nz, int flipz, float eps, float shift, float scale,
                        void* stream);

#ifdef __cplusplus
}
#endif

#endif
