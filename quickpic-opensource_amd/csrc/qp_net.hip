// Kernels of the learned corrector seed: a U-Net of 3x3 convolutions whose
// forward and backward passes run on the MFMA units in BF16 with FP32
// accumulation.
//
// Activations are BF16 (raw 16-bit patterns) in NHWC order with a channel
// stride, so a concatenation is two layers writing disjoint channel ranges
// of one buffer.  Weights of a 3x3 convolution are BF16 [tap][cout][cin],
// tap = 3 (dy+1) + dx + 1; biases and every gradient accumulator are FP32.
// Channel counts are multiples of 16 (the input layer is zero-padded).
//
// MFMA v_mfma_f32_16x16x16bf16_1k, one wave: lane l holds A[l%16][4(l/16)+i],
// B[4(l/16)+i][l%16] and D[4(l/16)+i][l%16], i = 0..3.
#include "qp_common.h"
#include "../../include/qpnet.h"

namespace qp {

typedef short v4s __attribute__((ext_vector_type(4)));
typedef float v4f __attribute__((ext_vector_type(4)));

__device__ __forceinline__ uint16_t f2bf(float f)
{
    // round to nearest even (inputs are finite)
    uint32_t u = __float_as_uint(f);
    u += 0x7fffu + ((u >> 16) & 1u);
    return (uint16_t)(u >> 16);
}
__device__ __forceinline__ float bf2f(uint16_t h) { return __uint_as_float((uint32_t)h << 16); }

// 4 consecutive BF16 channels as an MFMA operand (8-byte aligned)
__device__ __forceinline__ v4s load4(const uint16_t *p)
{
    const uint2 u = *reinterpret_cast<const uint2 *>(p);
    v4s r;
    r[0] = (short)(u.x & 0xffff);
    r[1] = (short)(u.x >> 16);
    r[2] = (short)(u.y & 0xffff);
    r[3] = (short)(u.y >> 16);
    return r;
}

// ---------------------------------------------------------------------------
// 3x3 convolution, 'same' zero padding:
//   y[p][co] = bias[co] + sum_{tap, ci} x'[p + off(tap)][ci] w[tap][co][ci]
// x' = x, or x * (mask > 0) when mask is given (the ReLU derivative of the
// layer that produced x, for the data gradient).  out = relu?(y) as BF16,
// added to out when accum.  A workgroup computes 64 pixels of one row; each
// wave 16 pixels and all cout (up to 192 = 12 MFMA column tiles).
// ---------------------------------------------------------------------------
constexpr int MAXNT = 12;

__global__ void __launch_bounds__(256)
k_conv3x3(const uint16_t *__restrict__ x, const uint16_t *__restrict__ mask, int H, int W,
          int cs_in, int Cin, const uint16_t *__restrict__ w, const float *__restrict__ bias,
          uint16_t *out, int cs_out, int Cout, int relu, int accum)
{
    const int lane = threadIdx.x & 63, wave = threadIdx.x >> 6;
    const int y = blockIdx.y;
    const int px0 = blockIdx.x * 64 + wave * 16;      // this wave's first pixel in the row
    const int NT = Cout >> 4;
    const int m = lane & 15, q = lane >> 4;           // A row (pixel) and k quad
    v4f acc[MAXNT];
#pragma unroll
    for (int t = 0; t < MAXNT; ++t) acc[t] = v4f{0.f, 0.f, 0.f, 0.f};
    for (int tap = 0; tap < 9; ++tap) {
        const int dy = tap / 3 - 1, dx = tap % 3 - 1;
        const int yy = y + dy, xx = px0 + m + dx;
        const bool in = yy >= 0 && yy < H && xx >= 0 && xx < W;
        const size_t pix = in ? ((size_t)yy * W + xx) * cs_in : 0;
        for (int c0 = 0; c0 < Cin; c0 += 16) {
            v4s a = {0, 0, 0, 0};
            if (in) {
                a = load4(x + pix + c0 + 4 * q);
                if (mask) {
                    const v4s mk = load4(mask + pix + c0 + 4 * q);
#pragma unroll
                    for (int i = 0; i < 4; ++i)
                        if (!(bf2f((uint16_t)mk[i]) > 0.f)) a[i] = 0;
                }
            }
            const uint16_t *wt = w + ((size_t)tap * Cout + m) * Cin + c0 + 4 * q;
#pragma unroll
            for (int t = 0; t < MAXNT; ++t) {
                if (t < NT) {
                    const v4s b = load4(wt + (size_t)t * 16 * Cin);
                    acc[t] = __builtin_amdgcn_mfma_f32_16x16x16bf16_1k(a, b, acc[t], 0, 0, 0);
                }
            }
        }
    }
    // D[pixel 4q+i][cout 16t + m]
#pragma unroll
    for (int t = 0; t < MAXNT; ++t) {
        if (t >= NT) continue;
        const int co = 16 * t + m;
        const float bv = bias ? bias[co] : 0.f;
#pragma unroll
        for (int i = 0; i < 4; ++i) {
            const int xp = px0 + 4 * q + i;
            if (xp >= W) continue;
            float v = acc[t][i] + bv;
            if (relu) v = v > 0.f ? v : 0.f;
            uint16_t *o = out + ((size_t)y * W + xp) * cs_out + co;
            if (accum) v += bf2f(*o);
            *o = f2bf(v);
        }
    }
}

// ---------------------------------------------------------------------------
// Weight gradient of a 3x3 convolution:
//   dw[tap][co][ci] = sum_p dy'[p][co] x[p + off(tap)][ci]
// dy' = dy * (yout > 0) (ReLU derivative from the layer's own output).  The
// reduction over pixels is the MFMA K dimension.  Workgroup (co tile, ci
// tile, tap) x chunk: its 4 waves take 16-pixel groups of the chunk in turn;
// the 4 wave sums are added in wave order and written to
// part[chunk][tap][co][ci]; k_sum_chunks adds the chunks in order.
// ---------------------------------------------------------------------------
__global__ void __launch_bounds__(256)
k_conv3x3_wgrad(const uint16_t *__restrict__ x, int cs_in, int Cin,
                const uint16_t *__restrict__ dy, const uint16_t *__restrict__ yout, int cs_out,
                int Cout, int H, int W, int nchunk, float *part)
{
    __shared__ float red[4][256];
    const int lane = threadIdx.x & 63, wave = threadIdx.x >> 6;
    const int ntc = Cin >> 4;
    const int cot = blockIdx.x / ntc, cit = blockIdx.x % ntc;
    const int tap = blockIdx.y, chunk = blockIdx.z;
    const int dyo = tap / 3 - 1, dxo = tap % 3 - 1;
    const int npix = H * W;
    const int per = (npix + nchunk - 1) / nchunk;
    const int p0 = chunk * per, p1 = min(npix, p0 + per);
    const int m = lane & 15, q = lane >> 4;
    v4f acc = {0.f, 0.f, 0.f, 0.f};
    for (int g = p0 + wave * 16; g < p1; g += 64) {
        // A[co = m][k = pixel g + 4q + i], B[k = pixel g + 4q + i][ci = m]
        v4s a, b;
#pragma unroll
        for (int i = 0; i < 4; ++i) {
            const int p = g + 4 * q + i;
            short av = 0, bv = 0;
            if (p < p1) {
                const size_t po = (size_t)p * cs_out + cot * 16 + m;
                const uint16_t d = dy[po];
                if (bf2f(yout[po]) > 0.f) av = (short)d;
                const int py = p / W, pxx = p - py * W;
                const int sy = py + dyo, sx = pxx + dxo;
                if (sy >= 0 && sy < H && sx >= 0 && sx < W)
                    bv = (short)x[((size_t)sy * W + sx) * cs_in + cit * 16 + m];
            }
            a[i] = av;
            b[i] = bv;
        }
        acc = __builtin_amdgcn_mfma_f32_16x16x16bf16_1k(a, b, acc, 0, 0, 0);
    }
#pragma unroll
    for (int i = 0; i < 4; ++i) red[wave][lane * 4 + i] = acc[i];
    __syncthreads();
    if (wave == 0) {
#pragma unroll
        for (int i = 0; i < 4; ++i) {
            const float s = ((red[0][lane * 4 + i] + red[1][lane * 4 + i]) + red[2][lane * 4 + i])
                            + red[3][lane * 4 + i];
            const int co = cot * 16 + 4 * q + i, ci = cit * 16 + m;
            part[(((size_t)chunk * 9 + tap) * Cout + co) * Cin + ci] = s;
        }
    }
}

// bias gradient: db[co] = sum_p dy'[p][co], per chunk (same chunks)
__global__ void k_bias_grad(const uint16_t *__restrict__ dy, const uint16_t *__restrict__ yout,
                            int cs_out, int Cout, int npix, int nchunk, float *part)
{
    const int co = threadIdx.x, chunk = blockIdx.x;
    if (co >= Cout) return;
    const int per = (npix + nchunk - 1) / nchunk;
    const int p0 = chunk * per, p1 = min(npix, p0 + per);
    float s = 0.f;
    for (int p = p0; p < p1; ++p) {
        const size_t o = (size_t)p * cs_out + co;
        if (bf2f(yout[o]) > 0.f) s += bf2f(dy[o]);
    }
    part[(size_t)chunk * Cout + co] = s;
}

// out[i] = sum over chunks (in order) of part[chunk][i]
__global__ void k_sum_chunks(const float *__restrict__ part, size_t n, int nchunk, float *out)
{
    for (size_t i = blockIdx.x * (size_t)blockDim.x + threadIdx.x; i < n;
         i += (size_t)gridDim.x * blockDim.x) {
        float s = 0.f;
        for (int c = 0; c < nchunk; ++c) s += part[(size_t)c * n + i];
        out[i] = s;
    }
}

// w'[tap][ci][co] = w[8 - tap][co][ci]: the data gradient is the forward
// convolution of dy with the flipped, transposed kernel
__global__ void k_flip_weights(const uint16_t *__restrict__ w, int Cout, int Cin, uint16_t *wt)
{
    const size_t n = (size_t)9 * Cout * Cin;
    for (size_t i = blockIdx.x * (size_t)blockDim.x + threadIdx.x; i < n;
         i += (size_t)gridDim.x * blockDim.x) {
        const int tap = (int)(i / ((size_t)Cin * Cout));
        const int r = (int)(i % ((size_t)Cin * Cout));
        const int ci = r / Cout, co = r % Cout;
        wt[i] = w[((size_t)(8 - tap) * Cout + co) * Cin + ci];
    }
}

// ---------------------------------------------------------------------------
// pooling, upsampling, input and output layers
// ---------------------------------------------------------------------------
// 2x2 average: out (H/2 x W/2) from in (H x W), C channels
__global__ void k_pool2(const uint16_t *__restrict__ in, int cs_in, uint16_t *out, int cs_out,
                        int H, int W, int C)
{
    const int Ho = H / 2, Wo = W / 2;
    const size_t n = (size_t)Ho * Wo * C;
    for (size_t i = blockIdx.x * (size_t)blockDim.x + threadIdx.x; i < n;
         i += (size_t)gridDim.x * blockDim.x) {
        const int c = (int)(i % C);
        const size_t p = i / C;
        const int yo = (int)(p / Wo), xo = (int)(p % Wo);
        const size_t b = ((size_t)(2 * yo) * W + 2 * xo) * cs_in + c;
        const float s = (bf2f(in[b]) + bf2f(in[b + cs_in])) +
                        (bf2f(in[b + (size_t)W * cs_in]) + bf2f(in[b + (size_t)W * cs_in + cs_in]));
        out[p * cs_out + c] = f2bf(0.25f * s);
    }
}

// its gradient: din += 0.25 dout (each fine pixel from its coarse pixel)
__global__ void k_pool2_bwd(const uint16_t *__restrict__ dout, int cs_out, uint16_t *din,
                            int cs_in, int H, int W, int C, int accum)
{
    const size_t n = (size_t)H * W * C;
    for (size_t i = blockIdx.x * (size_t)blockDim.x + threadIdx.x; i < n;
         i += (size_t)gridDim.x * blockDim.x) {
        const int c = (int)(i % C);
        const size_t p = i / C;
        const int y = (int)(p / W), x = (int)(p % W);
        float v = 0.25f * bf2f(dout[((size_t)(y / 2) * (W / 2) + x / 2) * cs_out + c]);
        uint16_t *o = din + p * cs_in + c;
        if (accum) v += bf2f(*o);
        *o = f2bf(v);
    }
}

// nearest-neighbour 2x upsampling: out (2H x 2W) from in (H x W)
__global__ void k_up2(const uint16_t *__restrict__ in, int cs_in, uint16_t *out, int cs_out,
                      int H, int W, int C)
{
    const int Ho = 2 * H, Wo = 2 * W;
    const size_t n = (size_t)Ho * Wo * C;
    for (size_t i = blockIdx.x * (size_t)blockDim.x + threadIdx.x; i < n;
         i += (size_t)gridDim.x * blockDim.x) {
        const int c = (int)(i % C);
        const size_t p = i / C;
        const int y = (int)(p / Wo), x = (int)(p % Wo);
        out[p * cs_out + c] = in[((size_t)(y / 2) * W + x / 2) * cs_in + c];
    }
}

// its gradient: din (H x W) = sum of the 4 fine pixels of dout (2H x 2W)
__global__ void k_up2_bwd(const uint16_t *__restrict__ dout, int cs_out, uint16_t *din,
                          int cs_in, int H, int W, int C)
{
    const size_t n = (size_t)H * W * C;
    for (size_t i = blockIdx.x * (size_t)blockDim.x + threadIdx.x; i < n;
         i += (size_t)gridDim.x * blockDim.x) {
        const int c = (int)(i % C);
        const size_t p = i / C;
        const int y = (int)(p / W), x = (int)(p % W);
        const size_t b = ((size_t)(2 * y) * (2 * W) + 2 * x) * cs_out + c;
        const size_t r = (size_t)(2 * W) * cs_out;
        const float s = (bf2f(dout[b]) + bf2f(dout[b + cs_out])) + (bf2f(dout[b + r]) + bf2f(dout[b + r + cs_out]));
        din[p * cs_in + c] = f2bf(s);
    }
}

// input layer: f32 planes [c][H][W] (pitch ld) scaled by scale[c] -> BF16
// NHWC with cs channels (channels >= nc are zero)
__global__ void k_pack_input(const float *__restrict__ planes, int nc, size_t plane, int ld,
                             const float *__restrict__ scale, int H, int W, uint16_t *out, int cs)
{
    const size_t n = (size_t)H * W * cs;
    for (size_t i = blockIdx.x * (size_t)blockDim.x + threadIdx.x; i < n;
         i += (size_t)gridDim.x * blockDim.x) {
        const int c = (int)(i % cs);
        const size_t p = i / cs;
        const int y = (int)(p / W), x = (int)(p % W);
        const float v = c < nc ? planes[c * plane + (size_t)y * ld + x] * scale[c] : 0.f;
        out[i] = f2bf(v);
    }
}

struct PlaneSet {
    const void *p[8];
    int ld[8];
};

__global__ void k_block_mean(PlaneSet ps, int nc, int dbl, int f, int H, int W, float *out)
{
    const size_t n = (size_t)nc * H * W;
    const float inv = 1.f / (float)(f * f);
    for (size_t i = blockIdx.x * (size_t)blockDim.x + threadIdx.x; i < n;
         i += (size_t)gridDim.x * blockDim.x) {
        const int X = (int)(i % W);
        const size_t r = i / W;
        const int Y = (int)(r % H), c = (int)(r / H);
        const size_t o = (size_t)f * Y * ps.ld[c] + (size_t)f * X;
        float s = 0.f;
        if ((dbl >> c) & 1) {
            const double *q = (const double *)ps.p[c] + o;
            for (int a = 0; a < f; ++a)
                for (int b = 0; b < f; ++b) s += (float)q[(size_t)a * ps.ld[c] + b];
        } else {
            const float *q = (const float *)ps.p[c] + o;
            for (int a = 0; a < f; ++a)
                for (int b = 0; b < f; ++b) s += q[(size_t)a * ps.ld[c] + b];
        }
        out[i] = s * inv;
    }
}

// bilinear weights of fine node t (0 <= t < f n) between coarse centres
// f T + (f - 1) / 2, clamped at the first and last centre
__device__ __forceinline__ void lin1(int t, int f, int n, int &i0, int &i1, float &w1)
{
    const float u = ((float)t - 0.5f * (float)(f - 1)) / (float)f;
    if (u <= 0.f) { i0 = i1 = 0; w1 = 0.f; return; }
    if (u >= (float)(n - 1)) { i0 = i1 = n - 1; w1 = 0.f; return; }
    i0 = (int)u;
    i1 = i0 + 1;
    w1 = u - (float)i0;
}

template <typename R>
__global__ void k_seed_apply(const float *__restrict__ corr, int H, int W, int f, float s,
                             const float *__restrict__ a, const float *__restrict__ b, int ld_ab,
                             size_t pl_ab, R *dst, int ld_d, size_t pl_d, int ny1, int nx1)
{
    const size_t n = (size_t)2 * ny1 * nx1;
    for (size_t i = blockIdx.x * (size_t)blockDim.x + threadIdx.x; i < n;
         i += (size_t)gridDim.x * blockDim.x) {
        const int x = (int)(i % nx1);
        const size_t r = i / nx1;
        const int y = (int)(r % ny1), c = (int)(r / ny1);
        const size_t ab = c * pl_ab + (size_t)y * ld_ab + x;
        float v = a ? 2.f * a[ab] - b[ab] : 0.f;
        if (y < f * H && x < f * W) {
            int y0, y1, x0, x1;
            float wy, wx;
            lin1(y, f, H, y0, y1, wy);
            lin1(x, f, W, x0, x1, wx);
            const float *q = corr + (size_t)c * H * W;
            const float top = (1.f - wx) * q[(size_t)y0 * W + x0] + wx * q[(size_t)y0 * W + x1];
            const float bot = (1.f - wx) * q[(size_t)y1 * W + x0] + wx * q[(size_t)y1 * W + x1];
            v += s * ((1.f - wy) * top + wy * bot);
        }
        dst[c * pl_d + (size_t)y * ld_d + x] = (R)v;
    }
}

template <typename R>
__global__ void k_seed_combine(R *epw, R *hist, const R *__restrict__ corr, size_t n, int extrap)
{
    for (size_t i = blockIdx.x * (size_t)blockDim.x + threadIdx.x; i < n;
         i += (size_t)gridDim.x * blockDim.x) {
        const R e = epw[i];
        if (extrap) epw[i] = corr ? R(2) * e - hist[i] + corr[i] : R(2) * e - hist[i];
        hist[i] = e;
    }
}

// output layer (1x1, cin -> cout <= 4) in FP32: o[c][p] = b[c] + sum_k w[c][k] h[p][k]
__global__ void k_head(const uint16_t *__restrict__ h, int cs, int cin, const float *__restrict__ w,
                       const float *__restrict__ b, int cout, size_t npix, float *o)
{
    for (size_t p = blockIdx.x * (size_t)blockDim.x + threadIdx.x; p < npix;
         p += (size_t)gridDim.x * blockDim.x)
        for (int c = 0; c < cout; ++c) {
            float s = b[c];
            for (int k = 0; k < cin; ++k) s += w[c * cin + k] * bf2f(h[p * cs + k]);
            o[c * npix + p] = s;
        }
}

// head backward: dh[p][k] = sum_c w[c][k] do[c][p] (BF16); per-chunk partial
// weight / bias gradients part[chunk][c][k] and part[chunk][c][cin]
__global__ void k_head_bwd(const uint16_t *__restrict__ h, int cs, int cin,
                           const float *__restrict__ w, int cout, size_t npix,
                           const float *__restrict__ dout, uint16_t *dh, int nchunk, float *part)
{
    const int chunk = blockIdx.x;
    const size_t per = (npix + nchunk - 1) / nchunk;
    const size_t p0 = chunk * per, p1 = min(npix, p0 + per);
    // weight partials: thread t < cout*(cin+1) owns one (c, k) sum over the chunk
    const int t = threadIdx.x;
    if (t < cout * (cin + 1)) {
        const int c = t / (cin + 1), k = t % (cin + 1);
        float s = 0.f;
        for (size_t p = p0; p < p1; ++p)
            s += dout[c * npix + p] * (k < cin ? bf2f(h[p * cs + k]) : 1.f);
        part[(size_t)chunk * cout * (cin + 1) + t] = s;
    }
    for (size_t p = p0 + threadIdx.x; p < p1; p += blockDim.x)
        for (int k = 0; k < cin; ++k) {
            float s = 0.f;
            for (int c = 0; c < cout; ++c) s += w[c * cin + k] * dout[c * npix + p];
            dh[p * cs + k] = f2bf(s);
        }
}

// mean squared error over n values and its gradient 2 (o - t) / n; the loss
// sum goes to per-block partials (fixed-order sum afterwards)
__global__ void k_mse(const float *__restrict__ o, const float *__restrict__ t, size_t n,
                      float *grad, double *partial)
{
    __shared__ double s[256];
    double acc = 0.0;
    for (size_t i = blockIdx.x * (size_t)blockDim.x + threadIdx.x; i < n;
         i += (size_t)gridDim.x * blockDim.x) {
        const float d = o[i] - t[i];
        acc += (double)d * d;
        grad[i] = 2.f * d / (float)n;
    }
    s[threadIdx.x] = acc;
    __syncthreads();
    for (int k = 128; k > 0; k >>= 1) {
        if (threadIdx.x < k) s[threadIdx.x] += s[threadIdx.x + k];
        __syncthreads();
    }
    if (threadIdx.x == 0) partial[blockIdx.x] = s[0];
}

// Adam step on FP32 master weights, refreshing their BF16 copy (if any)
__global__ void k_adam(float *wm, const float *__restrict__ g, float *mo, float *ve, size_t n,
                       float lr, float b1, float b2, float eps, float c1, float c2, uint16_t *wb)
{
    for (size_t i = blockIdx.x * (size_t)blockDim.x + threadIdx.x; i < n;
         i += (size_t)gridDim.x * blockDim.x) {
        const float gi = g[i];
        const float mi = b1 * mo[i] + (1.f - b1) * gi;
        const float vi = b2 * ve[i] + (1.f - b2) * gi * gi;
        mo[i] = mi;
        ve[i] = vi;
        const float wv = wm[i] - lr * (mi / c1) / (sqrtf(vi / c2) + eps);
        wm[i] = wv;
        if (wb) wb[i] = f2bf(wv);
    }
}

__global__ void k_to_bf16(const float *__restrict__ a, size_t n, uint16_t *b)
{
    for (size_t i = blockIdx.x * (size_t)blockDim.x + threadIdx.x; i < n;
         i += (size_t)gridDim.x * blockDim.x)
        b[i] = f2bf(a[i]);
}

static dim3 g1(size_t n)
{
    size_t b = (n + 255) / 256;
    if (b > 16384) b = 16384;
    return dim3((unsigned)(b ? b : 1));
}

}  // namespace qp

using namespace qp;

#define NET_CHECK(cond, ...) QP_REQUIRE(cond, QP_ERR_ARG, __VA_ARGS__)

extern "C" {

int qp_net_conv3x3(void *stream_, const uint16_t *x, const uint16_t *mask, int H, int W,
                   int cs_in, int cin, const uint16_t *w, const float *bias, uint16_t *out,
                   int cs_out, int cout, int relu, int accum)
{
    hipStream_t st = (hipStream_t)stream_;
    NET_CHECK(x && w && out && H > 0 && W > 0, "qp_net_conv3x3: bad argument");
    NET_CHECK(cin % 16 == 0 && cout % 16 == 0 && cout <= 16 * MAXNT && cin <= cs_in &&
                  cout <= cs_out && cs_in % 4 == 0,
              "qp_net_conv3x3: channels cin %d cout %d (multiples of 16, cout <= 192)", cin, cout);
    dim3 grid((W + 63) / 64, H);
    hipLaunchKernelGGL(k_conv3x3, grid, dim3(256), 0, st, x, mask, H, W, cs_in, cin, w, bias,
                       out, cs_out, cout, relu, accum);
    QP_LAUNCH_CHECK();
    return QP_OK;
}

int qp_net_conv3x3_wgrad(void *stream_, const uint16_t *x, int cs_in, int cin,
                         const uint16_t *dy, const uint16_t *yout, int cs_out, int cout, int H,
                         int W, int nchunk, float *part, float *dw, float *db)
{
    hipStream_t st = (hipStream_t)stream_;
    NET_CHECK(x && dy && yout && part && dw && nchunk > 0, "qp_net_conv3x3_wgrad: bad argument");
    NET_CHECK(cin % 16 == 0 && cout % 16 == 0, "qp_net_conv3x3_wgrad: channels");
    dim3 grid((cout / 16) * (cin / 16), 9, nchunk);
    hipLaunchKernelGGL(k_conv3x3_wgrad, grid, dim3(256), 0, st, x, cs_in, cin, dy, yout, cs_out,
                       cout, H, W, nchunk, part);
    QP_LAUNCH_CHECK();
    const size_t n = (size_t)9 * cout * cin;
    hipLaunchKernelGGL(k_sum_chunks, g1(n), dim3(256), 0, st, part, n, nchunk, dw);
    QP_LAUNCH_CHECK();
    if (db) {
        hipLaunchKernelGGL(k_bias_grad, dim3(nchunk), dim3(cout < 64 ? 64 : cout), 0, st, dy,
                           yout, cs_out, cout, H * W, nchunk, part);
        QP_LAUNCH_CHECK();
        hipLaunchKernelGGL(k_sum_chunks, g1(cout), dim3(256), 0, st, part, (size_t)cout, nchunk,
                           db);
        QP_LAUNCH_CHECK();
    }
    return QP_OK;
}

int qp_net_flip_weights(void *stream_, const uint16_t *w, int cout, int cin, uint16_t *wt)
{
    hipStream_t st = (hipStream_t)stream_;
    NET_CHECK(w && wt, "qp_net_flip_weights: null");
    hipLaunchKernelGGL(k_flip_weights, g1((size_t)9 * cout * cin), dim3(256), 0, st, w, cout,
                       cin, wt);
    QP_LAUNCH_CHECK();
    return QP_OK;
}

int qp_net_pool2(void *stream_, const uint16_t *in, int cs_in, uint16_t *out, int cs_out, int H,
                 int W, int C)
{
    hipStream_t st = (hipStream_t)stream_;
    hipLaunchKernelGGL(k_pool2, g1((size_t)H * W * C / 4), dim3(256), 0, st, in, cs_in, out,
                       cs_out, H, W, C);
    QP_LAUNCH_CHECK();
    return QP_OK;
}

int qp_net_pool2_bwd(void *stream_, const uint16_t *dout, int cs_out, uint16_t *din, int cs_in,
                     int H, int W, int C, int accum)
{
    hipStream_t st = (hipStream_t)stream_;
    hipLaunchKernelGGL(k_pool2_bwd, g1((size_t)H * W * C), dim3(256), 0, st, dout, cs_out, din,
                       cs_in, H, W, C, accum);
    QP_LAUNCH_CHECK();
    return QP_OK;
}

int qp_net_up2(void *stream_, const uint16_t *in, int cs_in, uint16_t *out, int cs_out, int H,
               int W, int C)
{
    hipStream_t st = (hipStream_t)stream_;
    hipLaunchKernelGGL(k_up2, g1((size_t)4 * H * W * C), dim3(256), 0, st, in, cs_in, out, cs_out,
                       H, W, C);
    QP_LAUNCH_CHECK();
    return QP_OK;
}

int qp_net_up2_bwd(void *stream_, const uint16_t *dout, int cs_out, uint16_t *din, int cs_in,
                   int H, int W, int C)
{
    hipStream_t st = (hipStream_t)stream_;
    hipLaunchKernelGGL(k_up2_bwd, g1((size_t)H * W * C), dim3(256), 0, st, dout, cs_out, din,
                       cs_in, H, W, C);
    QP_LAUNCH_CHECK();
    return QP_OK;
}

int qp_net_pack_input(void *stream_, const float *planes, int nc, size_t plane, int ld,
                      const float *scale, int H, int W, uint16_t *out, int cs)
{
    hipStream_t st = (hipStream_t)stream_;
    hipLaunchKernelGGL(k_pack_input, g1((size_t)H * W * cs), dim3(256), 0, st, planes, nc, plane,
                       ld, scale, H, W, out, cs);
    QP_LAUNCH_CHECK();
    return QP_OK;
}

int qp_net_block_mean(void *stream_, int nc, const void *const *planes, const int *ld, int dbl,
                      int f, int H, int W, float *out)
{
    hipStream_t st = (hipStream_t)stream_;
    NET_CHECK(nc >= 1 && nc <= 8 && planes && ld && out && f >= 1 && H > 0 && W > 0,
              "qp_net_block_mean: bad argument");
    PlaneSet ps{};
    for (int c = 0; c < nc; ++c) {
        NET_CHECK(planes[c] && ld[c] >= f * W, "qp_net_block_mean: plane %d", c);
        ps.p[c] = planes[c];
        ps.ld[c] = ld[c];
    }
    hipLaunchKernelGGL(k_block_mean, g1((size_t)nc * H * W), dim3(256), 0, st, ps, nc, dbl, f, H,
                       W, out);
    QP_LAUNCH_CHECK();
    return QP_OK;
}

int qp_net_seed_apply(void *stream_, const float *corr, int H, int W, int f, float s,
                      const float *a, const float *b, int ld_ab, size_t pl_ab, void *dst,
                      int dst_double, int ld_d, size_t pl_d, int ny1, int nx1)
{
    hipStream_t st = (hipStream_t)stream_;
    NET_CHECK(corr && dst && H > 0 && W > 0 && f >= 1 && ny1 > 0 && nx1 > 0 && ld_d >= nx1 &&
                  pl_d >= (size_t)ld_d * ny1 && (!a) == (!b) &&
                  (!a || (ld_ab >= nx1 && pl_ab >= (size_t)ld_ab * ny1)),
              "qp_net_seed_apply: bad argument");
    const size_t n = (size_t)2 * ny1 * nx1;
    if (dst_double)
        hipLaunchKernelGGL(k_seed_apply<double>, g1(n), dim3(256), 0, st, corr, H, W, f, s, a, b,
                           ld_ab, pl_ab, (double *)dst, ld_d, pl_d, ny1, nx1);
    else
        hipLaunchKernelGGL(k_seed_apply<float>, g1(n), dim3(256), 0, st, corr, H, W, f, s, a, b,
                           ld_ab, pl_ab, (float *)dst, ld_d, pl_d, ny1, nx1);
    QP_LAUNCH_CHECK();
    return QP_OK;
}

int qp_net_seed_combine(void *stream_, void *epw, void *hist, const void *corr, size_t n,
                        int extrap, int dbl)
{
    hipStream_t st = (hipStream_t)stream_;
    NET_CHECK(epw && hist && n > 0, "qp_net_seed_combine: bad argument");
    if (dbl)
        hipLaunchKernelGGL(k_seed_combine<double>, g1(n), dim3(256), 0, st, (double *)epw,
                           (double *)hist, (const double *)corr, n, extrap);
    else
        hipLaunchKernelGGL(k_seed_combine<float>, g1(n), dim3(256), 0, st, (float *)epw,
                           (float *)hist, (const float *)corr, n, extrap);
    QP_LAUNCH_CHECK();
    return QP_OK;
}

int qp_net_graph_begin(void *stream_)
{
    QP_HIP(hipStreamBeginCapture((hipStream_t)stream_, hipStreamCaptureModeThreadLocal));
    return QP_OK;
}

int qp_net_graph_end(void *stream_, void **graph_exec)
{
    NET_CHECK(graph_exec, "qp_net_graph_end: null output");
    hipGraph_t g = nullptr;
    QP_HIP(hipStreamEndCapture((hipStream_t)stream_, &g));
    hipGraphExec_t ge = nullptr;
    const hipError_t e = hipGraphInstantiate(&ge, g, nullptr, nullptr, 0);
    (void)hipGraphDestroy(g);
    QP_HIP(e);
    *graph_exec = (void *)ge;
    return QP_OK;
}

int qp_net_graph_launch(void *graph_exec, void *stream_)
{
    NET_CHECK(graph_exec, "qp_net_graph_launch: null graph");
    QP_HIP(hipGraphLaunch((hipGraphExec_t)graph_exec, (hipStream_t)stream_));
    return QP_OK;
}

int qp_net_graph_destroy(void *graph_exec)
{
    if (graph_exec) QP_HIP(hipGraphExecDestroy((hipGraphExec_t)graph_exec));
    return QP_OK;
}

int qp_net_head(void *stream_, const uint16_t *h, int cs, int cin, const float *w,
                const float *b, int cout, size_t npix, float *o)
{
    hipStream_t st = (hipStream_t)stream_;
    hipLaunchKernelGGL(k_head, g1(npix), dim3(256), 0, st, h, cs, cin, w, b, cout, npix, o);
    QP_LAUNCH_CHECK();
    return QP_OK;
}

int qp_net_head_bwd(void *stream_, const uint16_t *h, int cs, int cin, const float *w, int cout,
                    size_t npix, const float *dout, uint16_t *dh, int nchunk, float *part,
                    float *dwb)
{
    hipStream_t st = (hipStream_t)stream_;
    NET_CHECK(cout * (cin + 1) <= 256, "qp_net_head_bwd: head too large");
    hipLaunchKernelGGL(k_head_bwd, dim3(nchunk), dim3(256), 0, st, h, cs, cin, w, cout, npix,
                       dout, dh, nchunk, part);
    QP_LAUNCH_CHECK();
    const size_t n = (size_t)cout * (cin + 1);
    hipLaunchKernelGGL(k_sum_chunks, g1(n), dim3(256), 0, st, part, n, nchunk, dwb);
    QP_LAUNCH_CHECK();
    return QP_OK;
}

int qp_net_mse(void *stream_, const float *o, const float *t, size_t n, float *grad,
               double *partial, int nblk)
{
    hipStream_t st = (hipStream_t)stream_;
    hipLaunchKernelGGL(k_mse, dim3(nblk), dim3(256), 0, st, o, t, n, grad, partial);
    QP_LAUNCH_CHECK();
    return QP_OK;
}

int qp_net_adam(void *stream_, float *wm, const float *g, float *m, float *v, size_t n, float lr,
                float b1, float b2, float eps, int step, uint16_t *wb)
{
    hipStream_t st = (hipStream_t)stream_;
    const float c1 = 1.f - powf(b1, (float)step), c2 = 1.f - powf(b2, (float)step);
    hipLaunchKernelGGL(k_adam, g1(n), dim3(256), 0, st, wm, g, m, v, n, lr, b1, b2, eps, c1, c2,
                       wb);
    QP_LAUNCH_CHECK();
    return QP_OK;
}

int qp_net_to_bf16(void *stream_, const float *a, size_t n, uint16_t *b)
{
    hipStream_t st = (hipStream_t)stream_;
    hipLaunchKernelGGL(k_to_bf16, g1(n), dim3(256), 0, st, a, n, b);
    QP_LAUNCH_CHECK();
    return QP_OK;
}

}  // extern "C"
