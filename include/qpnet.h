/*
 * qpnet.h -- C ABI of the learned corrector seed's kernels in libqpslice.so:
 * the forward and backward passes of a U-Net of 3x3 convolutions on the
 * MFMA units (BF16 operands, FP32 accumulation), its loss and its optimiser.
 *
 * No counterpart in the reference, which seeds each slice's B-perp corrector
 * with the previous slice's field (simulation_class.f03:676-710).  These
 * entry points serve a caller that trains and runs the network
 * (quickpic-opensource_amd/qpnet.py); the network changes only how many
 * corrector passes reach a residual tolerance, never the converged answer.
 *
 * Tensors are device pointers.  Activations are BF16 bit patterns (uint16_t)
 * in NHWC order with a channel stride cs (so a concatenation is two writers
 * of one buffer at different channel offsets: pass the offset pointer).
 * 3x3 weights are BF16 [tap][cout][cin] with tap = 3 (dy+1) + (dx+1); biases,
 * gradients of weights and losses are FP32.  Channel counts are multiples of
 * 16; cout <= 192.  stream is a hipStream_t (0: the null stream).  Every call
 * enqueues on the stream and returns 0 or a negative status.
 */
#ifndef QPNET_H
#define QPNET_H

#include <stddef.h>
#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

/* out = relu?(conv3x3(x * (mask > 0 if mask), w) + bias) (+ out if accum) */
int qp_net_conv3x3(void *stream, const uint16_t *x, const uint16_t *mask, int H, int W,
                   int cs_in, int cin, const uint16_t *w, const float *bias, uint16_t *out,
                   int cs_out, int cout, int relu, int accum);
/* dw[tap][cout][cin] = sum_p dy'[p][co] x[p+off][ci], dy' = dy * (yout > 0);
 * db[co] = sum_p dy'[p][co] (db may be null); part: nchunk * 9 * cout * cin
 * floats of scratch; chunks are summed in order, so the result is
 * independent of scheduling */
int qp_net_conv3x3_wgrad(void *stream, const uint16_t *x, int cs_in, int cin,
                         const uint16_t *dy, const uint16_t *yout, int cs_out, int cout, int H,
                         int W, int nchunk, float *part, float *dw, float *db);
/* wt[tap][cin][cout] = w[8 - tap][cout][cin] (kernel of the data gradient) */
int qp_net_flip_weights(void *stream, const uint16_t *w, int cout, int cin, uint16_t *wt);
/* 2x2 average pooling (H, W even) and its gradient (din = / += dout / 4) */
int qp_net_pool2(void *stream, const uint16_t *in, int cs_in, uint16_t *out, int cs_out, int H,
                 int W, int C);
int qp_net_pool2_bwd(void *stream, const uint16_t *dout, int cs_out, uint16_t *din, int cs_in,
                     int H, int W, int C, int accum);
/* nearest 2x upsampling of an H x W map and its gradient */
int qp_net_up2(void *stream, const uint16_t *in, int cs_in, uint16_t *out, int cs_out, int H,
               int W, int C);
int qp_net_up2_bwd(void *stream, const uint16_t *dout, int cs_out, uint16_t *din, int cs_in,
                   int H, int W, int C);
/* nc FP32 planes (plane elements apart, row pitch ld) times scale[c] -> BF16
 * NHWC with cs channels, channels >= nc zero */
int qp_net_pack_input(void *stream, const float *planes, int nc, size_t plane, int ld,
                      const float *scale, int H, int W, uint16_t *out, int cs);
/* Block means of nc <= 8 real-space planes: out[c][Y][X] (FP32, H x W) is
 * the mean of planes[c] over the f x f nodes (fY+i, fX+j); plane c has row
 * pitch ld[c] elements and is FP64 when bit c of dbl is set, else FP32.  The
 * network runs on this coarse grid (f = 1: a plain copy). */
int qp_net_block_mean(void *stream, int nc, const void *const *planes, const int *ld, int dbl,
                      int f, int H, int W, float *out);
/* The network's correction in real space on ny1 x nx1 nodes (2 components):
 * dst[c][y][x] = 2 a[c][y][x] - b[c][y][x] + s up(corr[c])(y, x), where up
 * interpolates the coarse H x W correction bilinearly between coarse pixel
 * centres (clamped at the edges) on the nodes y < f H, x < f W and is zero on
 * the others; a and b may both be NULL (taken as zero).  a, b: FP32 planes
 * with pitch ld_ab, plane stride pl_ab; dst: FP64 if dst_double else FP32,
 * pitch ld_d, plane stride pl_d. */
int qp_net_seed_apply(void *stream, const float *corr, int H, int W, int f, float s,
                      const float *a, const float *b, int ld_ab, size_t pl_ab, void *dst,
                      int dst_double, int ld_d, size_t pl_d, int ny1, int nx1);
/* Corrector seed in k space over n elements (FP64 if dbl, else FP32), the
 * library's extrapolated seed plus a correction: e = epw; with extrap,
 * epw = 2 e - hist (+ corr if not NULL); then hist = e. */
int qp_net_seed_combine(void *stream, void *epw, void *hist, const void *corr, size_t n,
                        int extrap, int dbl);
/* 1x1 output layer in FP32: o[c][p] = b[c] + sum_k w[c][k] h[p][k] */
int qp_net_head(void *stream, const uint16_t *h, int cs, int cin, const float *w,
                const float *b, int cout, size_t npix, float *o);
/* its gradient: dh (BF16), dwb = [dw (cout x cin) | db (cout)] per output row
 * c as (cin + 1) floats; part: nchunk * cout * (cin + 1) floats */
int qp_net_head_bwd(void *stream, const uint16_t *h, int cs, int cin, const float *w, int cout,
                    size_t npix, const float *dout, uint16_t *dh, int nchunk, float *part,
                    float *dwb);
/* mean squared error: grad = 2 (o - t) / n, per-block FP64 partial sums of
 * (o - t)^2 in partial[0..nblk) */
int qp_net_mse(void *stream, const float *o, const float *t, size_t n, float *grad,
               double *partial, int nblk);
/* Adam step (step >= 1) on FP32 master weights; wb (optional) gets their
 * BF16 copy */
int qp_net_adam(void *stream, float *wm, const float *g, float *m, float *v, size_t n, float lr,
                float b1, float b2, float eps, int step, uint16_t *wb);
int qp_net_to_bf16(void *stream, const float *a, size_t n, uint16_t *b);

/* HIP-graph capture of a fixed launch sequence (the network's forward pass
 * on fixed buffers): begin capturing the stream's work, end it into an
 * executable graph, replay it on a stream, destroy it.  The captured
 * sequence contains kernel launches only (no memset or copy nodes). */
int qp_net_graph_begin(void *stream);
int qp_net_graph_end(void *stream, void **graph_exec);
int qp_net_graph_launch(void *graph_exec, void *stream);
int qp_net_graph_destroy(void *graph_exec);

#ifdef __cplusplus
}
#endif

#endif
