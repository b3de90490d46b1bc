// fg_mlp.hip -- per-pixel Monte-Carlo error model: features and a small MLP
// (16 inputs, 12 used -> 64 -> 64 -> 1, ReLU) with BF16 weights and FP32
// accumulation on the gfx950 matrix cores, forward and backward.
//
// No counterpart in the reference.  The model predicts log(s), s the
// per-sample standard deviation sqrt(p (1 - p)) of a pixel's coverage
// indicator, from local statistics of the intensity plane; the sample budget
// of the adaptive renderer is proportional to s.
//
// Layout.  Parameters are one f32 array (fg_mlp.h offsets): W1 [16][64]
// (rows 12..15 zero), b1 [64], W2 [64][64] (input-major), b2 [64], w3 [64],
// b3.  Each workgroup converts them to BF16 in LDS, in the orders the matrix
// instructions read: W1T [64][16], W2T [64][64] (output-major) and W2 [64][64].
// Features are BF16 [n][16].  A wave works on 16 rows at a time:
//   layer 1: Z1 = X W1 + b1      4 x v_mfma_f32_16x16x16_bf16 (one per 16 outputs)
//   layer 2: Z2 = H1 W2 + b2     8 x v_mfma_f32_16x16x32_bf16, H1 = relu(Z1) in BF16
//                                 through LDS (accumulator rows -> operand rows)
//   layer 3: y = H2 . w3 + b3     FP32 dot products (a 64-vector), H2 = relu(Z2)
// Backward of the loss per row (fused, per tile), either the squared error to
// a target t, L = (y - t)^2, g = dL/dy = 2 (y - t), or the Gaussian negative
// log-likelihood of a residual t under the predicted log std y,
// L = y + t^2 e^(-2y) / 2, g = 1 - t^2 e^(-2y) (the exponent clamped to
// [-80, 80]); then with g scaled by 1 / n:
//   dZ2 = g w3 * [Z2 > 0], dZ1 = (dZ2 W2^T) * [Z1 > 0]
//   dW3 = H2^T g, dW2 = H1^T dZ2, dW1 = X^T dZ1 (MFMA over the 16 rows), biases
//   by sums.  Products feeding a matrix instruction use BF16 operands (dZ2,
//   dZ1, H1, X); their accumulations are FP32.  The CPU reference in the
//   tests rounds the same operands to BF16 at the same points.
// Matrix operand layouts (wave64): 16x16x16: lane l holds A[l%16][4(l/16)+0..3],
// B[4(l/16)+0..3][l%16]; 16x16x32: A[l%16][8(l/16)+0..7], B[8(l/16)+0..7][l%16];
// both: D[4(l/16)+0..3][l%16].
#include <hip/hip_runtime.h>

#include <algorithm>

#include "fg_kernels.h"
#include "fg_mlp.h"

namespace fg {

typedef short v4s __attribute__((ext_vector_type(4)));
typedef float v4f __attribute__((ext_vector_type(4)));
typedef __bf16 v8bf __attribute__((ext_vector_type(8)));

// f32 -> bf16 bits, round to nearest even (NaN kept quiet).
__device__ __forceinline__ uint16_t bf16_bits(float f)
{
    uint32_t u = __float_as_uint(f);
    if ((u & 0x7fffffffu) > 0x7f800000u) return (uint16_t)((u >> 16) | 0x40u);
    u += 0x7fffu + ((u >> 16) & 1u);
    return (uint16_t)(u >> 16);
}
__device__ __forceinline__ float bf16_float(uint16_t b) { return __uint_as_float((uint32_t)b << 16); }

struct MlpLds {
    uint16_t w1t[64 * 16];  // [j][k]
    uint16_t w2t[64 * 64];  // [j_out][k_in]
    uint16_t w2[64 * 64];   // [k_in][j_out]
    float b1[64], b2[64], w3[64], b3;
    uint16_t h[4][16 * 64];  // per wave: a 16 x 64 BF16 tile (H1, then dZ2)
    uint16_t x[4][16 * 16];  // per wave: the feature tile
};

__device__ void mlp_load(const float *P, MlpLds &s)
{
    for (int i = threadIdx.x; i < 64 * 16; i += blockDim.x) {
        const int j = i / 16, k = i % 16;
        s.w1t[i] = bf16_bits(P[kMlpW1 + k * 64 + j]);
    }
    for (int i = threadIdx.x; i < 64 * 64; i += blockDim.x) {
        const int j = i / 64, k = i % 64;
        s.w2t[i] = bf16_bits(P[kMlpW2 + k * 64 + j]);
        s.w2[i] = bf16_bits(P[kMlpW2 + i]);
    }
    for (int i = threadIdx.x; i < 64; i += blockDim.x) {
        s.b1[i] = P[kMlpB1 + i];
        s.b2[i] = P[kMlpB2 + i];
        s.w3[i] = bf16_float(bf16_bits(P[kMlpW3 + i]));
    }
    if (threadIdx.x == 0) s.b3 = P[kMlpB3];
    __syncthreads();
}

// Forward of one 16-row tile for the calling wave.  z1/z2 receive the
// pre-activations in accumulator layout (4 column blocks x 4 rows), y the
// row outputs (lanes of each 16-lane group hold the same 4 rows).
__device__ __forceinline__ void mlp_tile_forward(const MlpLds &s, uint16_t *hbuf, const uint16_t *xt, v4f z1[4],
                                                 v4f z2[4], float y[4])
{
    const int lane = threadIdx.x & 63, r = lane & 15, g = lane >> 4;
    // layer 1: A = X[r][4g..4g+3], B = W1[4g..][cb*16 + r]
    v4s a1;
    for (int t = 0; t < 4; ++t) a1[t] = (short)xt[r * 16 + 4 * g + t];
    for (int cb = 0; cb < 4; ++cb) {
        v4s b;
        for (int t = 0; t < 4; ++t) b[t] = (short)s.w1t[(cb * 16 + r) * 16 + 4 * g + t];
        v4f c = {0.0f, 0.0f, 0.0f, 0.0f};
        c = __builtin_amdgcn_mfma_f32_16x16x16bf16_1k(a1, b, c, 0, 0, 0);
        const float bias = s.b1[cb * 16 + r];
        for (int i = 0; i < 4; ++i) {
            c[i] += bias;
            hbuf[(4 * g + i) * 64 + cb * 16 + r] = bf16_bits(fmaxf(c[i], 0.0f));
        }
        z1[cb] = c;
    }
    // layer 2: A = H1[r][kc*32 + 8g..+7], B = W2[kc*32 + 8g..][cb*16 + r]
    v8bf a2[2];
    for (int kc = 0; kc < 2; ++kc)
        for (int t = 0; t < 8; ++t) a2[kc][t] = __builtin_bit_cast(__bf16, hbuf[r * 64 + kc * 32 + 8 * g + t]);
    for (int cb = 0; cb < 4; ++cb) {
        v4f c = {0.0f, 0.0f, 0.0f, 0.0f};
        for (int kc = 0; kc < 2; ++kc) {
            v8bf b;
            for (int t = 0; t < 8; ++t)
                b[t] = __builtin_bit_cast(__bf16, s.w2t[(cb * 16 + r) * 64 + kc * 32 + 8 * g + t]);
            c = __builtin_amdgcn_mfma_f32_16x16x32_bf16(a2[kc], b, c, 0, 0, 0);
        }
        const float bias = s.b2[cb * 16 + r];
        for (int i = 0; i < 4; ++i) c[i] += bias;
        z2[cb] = c;
    }
    // layer 3: y[row] = sum_j relu(Z2[row][j]) w3[j] + b3; lane r of group g
    // holds columns r, 16 + r, 32 + r, 48 + r of rows 4g..4g+3, so the sum
    // runs over its four columns, then over the 16 lanes of the group
    for (int i = 0; i < 4; ++i) {
        float acc = 0.0f;
        for (int cb = 0; cb < 4; ++cb) acc += fmaxf(z2[cb][i], 0.0f) * s.w3[cb * 16 + r];
        for (int o = 1; o < 16; o <<= 1) acc += __shfl_xor(acc, o, 64);
        y[i] = acc + s.b3;
    }
}

// Forward pass for inference, without LDS: the layers are computed
// transposed, Z1^T = W1^T X^T (16x16x16: A = a 16-output block of W1^T,
// B = X^T) and Z2^T = W2^T H1^T (16x16x32), so that a lane's accumulator
// holds row r = lane % 16 of the activations, columns cb*16 + 4g + i
// (g = lane / 16, i = 0..3).  Those are exactly the B operand layer 2 needs
// (row r, eight k per lane) under a fixed permutation of k within each
// 32-chunk: logical k = 8g + t is H1 column kc*32 + 4g + t (t < 4) or
// kc*32 + 16 + 4g + t - 4 (t >= 4); W2^T's operand is loaded in the same
// order, so the products are unchanged.  Weights, biases and w3 live in
// registers for the whole kernel; each lane reads its row's four features
// (8 bytes) from HBM.  Layer 3 sums the lane's 16 columns, then across the
// four lanes of the row.
__device__ __forceinline__ int mlp_h1_column(int kc, int g, int t)
{
    return kc * 32 + (t < 4 ? 4 * g + t : 16 + 4 * g + (t - 4));
}

__global__ __launch_bounds__(256) void mlp_forward_kernel(const float *P, const uint16_t *X, long long n, float *out)
{
    const int lane = threadIdx.x & 63, wid = threadIdx.x >> 6, r = lane & 15, g = lane >> 4;
    v4s a1[4];
    v8bf a2[4][2];
    float b1[4][4], b2[4][4], w3[4][4];
    for (int cb = 0; cb < 4; ++cb) {
        for (int t = 0; t < 4; ++t) a1[cb][t] = (short)bf16_bits(P[kMlpW1 + (4 * g + t) * 64 + cb * 16 + r]);
        for (int kc = 0; kc < 2; ++kc)
            for (int t = 0; t < 8; ++t)
                a2[cb][kc][t] =
                    __builtin_bit_cast(__bf16, bf16_bits(P[kMlpW2 + mlp_h1_column(kc, g, t) * 64 + cb * 16 + r]));
        for (int i = 0; i < 4; ++i) {
            const int j = cb * 16 + 4 * g + i;
            b1[cb][i] = P[kMlpB1 + j];
            b2[cb][i] = P[kMlpB2 + j];
            w3[cb][i] = bf16_float(bf16_bits(P[kMlpW3 + j]));
        }
    }
    const float b3 = P[kMlpB3];
    const long long tiles = (n + 15) / 16;
    const long long stride = (long long)gridDim.x * 4;
    for (long long tt = (long long)blockIdx.x * 4 + wid; tt < tiles; tt += stride) {
        const long long row = tt * 16 + r;
        v4s x = {0, 0, 0, 0};
        if (row < n) {
            const uint2 v = *reinterpret_cast<const uint2 *>(X + row * 16 + 4 * g);
            x[0] = (short)(v.x & 0xFFFFu);
            x[1] = (short)(v.x >> 16);
            x[2] = (short)(v.y & 0xFFFFu);
            x[3] = (short)(v.y >> 16);
        }
        // layer 1 (transposed) and the B operand of layer 2
        v8bf h[2];
        for (int cb = 0; cb < 4; ++cb) {
            v4f c = {0.0f, 0.0f, 0.0f, 0.0f};
            c = __builtin_amdgcn_mfma_f32_16x16x16bf16_1k(a1[cb], x, c, 0, 0, 0);
            for (int i = 0; i < 4; ++i)
                h[cb >> 1][(cb & 1) * 4 + i] = __builtin_bit_cast(__bf16, bf16_bits(fmaxf(c[i] + b1[cb][i], 0.0f)));
        }
        // layer 2 (transposed) and layer 3
        float acc = 0.0f;
        for (int cb = 0; cb < 4; ++cb) {
            v4f c = {0.0f, 0.0f, 0.0f, 0.0f};
            for (int kc = 0; kc < 2; ++kc) c = __builtin_amdgcn_mfma_f32_16x16x32_bf16(a2[cb][kc], h[kc], c, 0, 0, 0);
            for (int i = 0; i < 4; ++i) acc += fmaxf(c[i] + b2[cb][i], 0.0f) * w3[cb][i];
        }
        acc += __shfl_xor(acc, 16, 64);
        acc += __shfl_xor(acc, 32, 64);
        if (g == 0 && row < n) out[row] = acc + b3;
    }
}

// Fused forward + backward of the loss (loss_kind 0: squared error, 1:
// Gaussian negative log-likelihood of the residual) over rows [0, n) with
// loss scale 1 / n_total; gradients are added to grad (kMlpParams f32) and
// the unscaled loss sum to *loss.
__global__ __launch_bounds__(256) void mlp_grad_kernel(const float *P, const uint16_t *X, const float *target,
                                                       long long n, float inv_total, int loss_kind, float *grad,
                                                       float *loss)
{
    __shared__ MlpLds s;
    __shared__ float s_grad[kMlpParams];
    __shared__ float s_loss;
    mlp_load(P, s);
    for (int i = threadIdx.x; i < kMlpParams; i += blockDim.x) s_grad[i] = 0.0f;
    if (threadIdx.x == 0) s_loss = 0.0f;
    __syncthreads();
    const int lane = threadIdx.x & 63, wid = threadIdx.x >> 6, r = lane & 15, g = lane >> 4;
    v4f dw2[4][4], dw1[4];
    for (int a = 0; a < 4; ++a) {
        dw1[a] = (v4f){0.0f, 0.0f, 0.0f, 0.0f};
        for (int b = 0; b < 4; ++b) dw2[a][b] = (v4f){0.0f, 0.0f, 0.0f, 0.0f};
    }
    float db1[4] = {0, 0, 0, 0}, db2[4] = {0, 0, 0, 0}, dw3[4] = {0, 0, 0, 0}, db3 = 0.0f, lsum = 0.0f;
    const long long tiles = (n + 15) / 16;
    for (long long t = (long long)blockIdx.x * 4 + wid; t < tiles; t += (long long)gridDim.x * 4) {
        uint16_t *xt = s.x[wid];
        uint16_t *hb = s.h[wid];
        for (int i = lane; i < 256; i += 64) {
            const long long row = t * 16 + i / 16;
            xt[i] = row < n ? X[row * 16 + (i % 16)] : (uint16_t)0;
        }
        v4f z1[4], z2[4];
        float y[4];
        mlp_tile_forward(s, hb, xt, z1, z2, y);
        // output gradient (rows past n contribute nothing)
        float gr[4];
        for (int i = 0; i < 4; ++i) {
            const long long row = t * 16 + 4 * g + i;
            if (row >= n) {
                gr[i] = 0.0f;
                continue;
            }
            const float t = target[row];
            if (loss_kind == 1) {
                const float e = expf(fminf(fmaxf(-2.0f * y[i], -80.0f), 80.0f)) * t * t;
                gr[i] = (1.0f - e) * inv_total;
                if (r == 0) lsum += y[i] + 0.5f * e;
            } else {
                const float d = y[i] - t;
                gr[i] = 2.0f * d * inv_total;
                if (r == 0) lsum += d * d;
            }
        }
        // layer 3 and dZ2 (accumulator layout), dW2 = H1^T dZ2 on MFMA
        v4f dz2[4];
        for (int cb = 0; cb < 4; ++cb) {
            const float w = s.w3[cb * 16 + r];
            for (int i = 0; i < 4; ++i) {
                const float h2 = fmaxf(z2[cb][i], 0.0f);
                dw3[cb] += gr[i] * h2;
                dz2[cb][i] = z2[cb][i] > 0.0f ? gr[i] * w : 0.0f;
                db2[cb] += dz2[cb][i];
            }
        }
        if (r == 0)
            for (int i = 0; i < 4; ++i) db3 += gr[i];
        // dZ2 as BF16 B operand (rows 4g.., column r) and A^T operand via LDS
        v4s bz[4];
        for (int cb = 0; cb < 4; ++cb)
            for (int i = 0; i < 4; ++i) bz[cb][i] = (short)bf16_bits(dz2[cb][i]);
        // A = H1^T[kb*16 + r][rows 4g..4g+3] from the H1 tile (row-major)
        for (int kb = 0; kb < 4; ++kb) {
            v4s a;
            for (int i = 0; i < 4; ++i) a[i] = (short)hb[(4 * g + i) * 64 + kb * 16 + r];
            for (int cb = 0; cb < 4; ++cb)
                dw2[kb][cb] = __builtin_amdgcn_mfma_f32_16x16x16bf16_1k(a, bz[cb], dw2[kb][cb], 0, 0, 0);
        }
        // dH1 = dZ2 W2^T: A = dZ2[r][kc*32 + 8g..] through LDS, B[b][a] = W2[a][b]
        for (int cb = 0; cb < 4; ++cb)
            for (int i = 0; i < 4; ++i) hb[(4 * g + i) * 64 + cb * 16 + r] = (uint16_t)bz[cb][i];
        v8bf az[2];
        for (int kc = 0; kc < 2; ++kc)
            for (int q = 0; q < 8; ++q) az[kc][q] = __builtin_bit_cast(__bf16, hb[r * 64 + kc * 32 + 8 * g + q]);
        v4f dz1[4];
        for (int cb = 0; cb < 4; ++cb) {
            v4f c = {0.0f, 0.0f, 0.0f, 0.0f};
            for (int kc = 0; kc < 2; ++kc) {
                v8bf b;
                for (int q = 0; q < 8; ++q)
                    b[q] = __builtin_bit_cast(__bf16, s.w2[(cb * 16 + r) * 64 + kc * 32 + 8 * g + q]);
                c = __builtin_amdgcn_mfma_f32_16x16x32_bf16(az[kc], b, c, 0, 0, 0);
            }
            for (int i = 0; i < 4; ++i) {
                c[i] = z1[cb][i] > 0.0f ? c[i] : 0.0f;
                db1[cb] += c[i];
            }
            dz1[cb] = c;
        }
        // dW1 = X^T dZ1: A = X^T[r][rows 4g..] (feature r), B = dZ1 (BF16)
        v4s ax;
        for (int i = 0; i < 4; ++i) ax[i] = (short)xt[(4 * g + i) * 16 + r];
        for (int cb = 0; cb < 4; ++cb) {
            v4s b;
            for (int i = 0; i < 4; ++i) b[i] = (short)bf16_bits(dz1[cb][i]);
            dw1[cb] = __builtin_amdgcn_mfma_f32_16x16x16bf16_1k(ax, b, dw1[cb], 0, 0, 0);
        }
    }
    // reduce: accumulator blocks hold D[4g+i][r]
    for (int kb = 0; kb < 4; ++kb)
        for (int cb = 0; cb < 4; ++cb)
            for (int i = 0; i < 4; ++i)
                atomicAdd(&s_grad[kMlpW2 + (kb * 16 + 4 * g + i) * 64 + cb * 16 + r], dw2[kb][cb][i]);
    for (int cb = 0; cb < 4; ++cb)
        for (int i = 0; i < 4; ++i) atomicAdd(&s_grad[kMlpW1 + (4 * g + i) * 64 + cb * 16 + r], dw1[cb][i]);
    for (int cb = 0; cb < 4; ++cb) {
        atomicAdd(&s_grad[kMlpB1 + cb * 16 + r], db1[cb]);
        atomicAdd(&s_grad[kMlpB2 + cb * 16 + r], db2[cb]);
        atomicAdd(&s_grad[kMlpW3 + cb * 16 + r], dw3[cb]);
    }
    if (r == 0) {
        atomicAdd(&s_grad[kMlpB3], db3);
        atomicAdd(&s_loss, lsum);
    }
    __syncthreads();
    for (int i = threadIdx.x; i < kMlpParams; i += blockDim.x)
        if (s_grad[i] != 0.0f) atomicAdd(&grad[i], s_grad[i]);
    if (threadIdx.x == 0) atomicAdd(loss, s_loss);
}

// Adam (Kingma & Ba 2015) with bias correction; rows 12..15 of W1 stay 0.
__global__ void mlp_adam_kernel(float *P, const float *grad, float *m, float *v, float lr, float b1, float b2,
                                float eps, float bc1, float bc2)
{
    const int i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= kMlpParams) return;
    if (i >= kMlpW1 + 12 * 64 && i < kMlpW1 + 16 * 64) {
        P[i] = 0.0f;
        return;
    }
    const float gi = grad[i];
    const float mi = b1 * m[i] + (1.0f - b1) * gi;
    const float vi = b2 * v[i] + (1.0f - b2) * gi * gi;
    m[i] = mi;
    v[i] = vi;
    P[i] -= lr * (mi / bc1) / (sqrtf(vi / bc2) + eps);
}

// Features of output pixel (x, y) of a plane: the normalised intensity u of
// the input pixel it maps to (floor((x + 0.5) / zoom), clamped), the mean and
// variance of u over its 3x3 neighbourhood (edges clamped), the gradient
// magnitude (central differences), u (1 - u), the radius law, filter sigma and
// zoom (scaled to order one), and the channel one-hot.
__global__ __launch_bounds__(256) void feature_kernel(const float *u, int in_w, int in_h, int out_w, int out_h,
                                                      float inv_zoom, float f_radius, float f_sd, float f_sigma,
                                                      float f_zoom, int channel, uint16_t *X)
{
    const long long p = (long long)blockIdx.x * 256 + threadIdx.x;
    if (p >= (long long)out_w * out_h) return;
    const int y = (int)(p / out_w), x = (int)(p - (long long)y * out_w);
    const int ix = min(max((int)floorf(((float)x + 0.5f) * inv_zoom), 0), in_w - 1);
    const int iy = min(max((int)floorf(((float)y + 0.5f) * inv_zoom), 0), in_h - 1);
    auto at = [&](int dx, int dy) {
        const int xx = min(max(ix + dx, 0), in_w - 1), yy = min(max(iy + dy, 0), in_h - 1);
        return u[(size_t)yy * in_w + xx];
    };
    float s = 0.0f, s2 = 0.0f;
    for (int dy = -1; dy <= 1; ++dy)
        for (int dx = -1; dx <= 1; ++dx) {
            const float v = at(dx, dy);
            s += v;
            s2 += v * v;
        }
    const float mean = s * (1.0f / 9.0f);
    const float var = fmaxf(s2 * (1.0f / 9.0f) - mean * mean, 0.0f);
    const float gx = 0.5f * (at(1, 0) - at(-1, 0)), gy = 0.5f * (at(0, 1) - at(0, -1));
    const float c = at(0, 0);
    float f[16] = {c,       mean,       sqrtf(var), sqrtf(gx * gx + gy * gy), c * (1.0f - c), f_radius, f_sd,
                   f_sigma, f_zoom,     channel == 0 ? 1.0f : 0.0f, channel == 1 ? 1.0f : 0.0f,
                   channel == 2 ? 1.0f : 0.0f, 0.0f, 0.0f, 0.0f, 0.0f};
    uint16_t *dst = X + p * 16;
    for (int i = 0; i < 16; ++i) dst[i] = bf16_bits(f[i]);
}

// Sample budget n = clamp(ceil(c * s), nmin, nmax) per pixel, from s =
// exp(predicted log s) (pred != nullptr) or, for the allocator without the
// model, s = sqrt(u (1 - u)) from feature 4 of X.
// With dc != nullptr the scale is *dc (written by the scale search).
__global__ __launch_bounds__(256) void budget_kernel(const float *pred, const uint16_t *X, long long n, float c,
                                                     const float *dc, int nmin, int nmax, uint16_t *budget)
{
    const long long i = (long long)blockIdx.x * 256 + threadIdx.x;
    if (i >= n) return;
    if (dc) c = *dc;
    const float s = pred ? expf(pred[i]) : sqrtf(fmaxf(bf16_float(X[i * 16 + 4]), 0.0f));
    const float v = ceilf(c * s);
    const int b = v >= (float)nmax ? nmax : (v <= (float)nmin ? nmin : (int)v);
    budget[i] = (uint16_t)b;
}

hipError_t launch_budget(const float *pred, const uint16_t *X, long long n, float c, const float *dc, int nmin,
                         int nmax, uint16_t *budget, hipStream_t s)
{
    hipLaunchKernelGGL(budget_kernel, dim3((unsigned)((n + 255) / 256)), dim3(256), 0, s, pred, X, n, c, dc, nmin,
                       nmax, budget);
    return hipGetLastError();
}

// ---- the budget scale search on the device ----
// Per-sample stds s_i = exp(pred[i * step]) * s_mult over the strided
// subsample i < m.  The smallest c whose budgets clamp(ceil(c s), nmin, nmax)
// have mean predicted squared error s^2 (1/n - 1/n_ref) <= target, by rounds
// of kScaleProbes geometric candidates between lo and hi (each round keeps the
// bracket of the first candidate that meets the target).
constexpr int kScaleProbes = 64;

__device__ __forceinline__ float scale_probe(float lo, float hi, int j)
{
    const float a = logf(lo), b = logf(hi);
    return expf(a + (b - a) * ((float)j / (float)(kScaleProbes - 1)));
}

// State: lo, hi, c, done (floats), the bits of the smallest positive s (a
// positive float orders as its bits do), then the probe sums.
__global__ void scale_clear_kernel(float *state, double *sums)
{
    if (threadIdx.x < kScaleProbes) sums[threadIdx.x] = 0.0;
    if (threadIdx.x == 0) reinterpret_cast<unsigned int *>(state)[4] = 0x7F800000u;  // +inf
}

__global__ __launch_bounds__(256) void scale_min_kernel(const float *pred, long long step, long long m, float s_mult,
                                                        float *state)
{
    float mn = INFINITY;
    for (long long i = (long long)blockIdx.x * 256 + threadIdx.x; i < m; i += (long long)gridDim.x * 256) {
        const float s = expf(pred[i * step]) * s_mult;
        if (s > 0.0f) mn = fminf(mn, s);
    }
    for (int o = 32; o > 0; o >>= 1) mn = fminf(mn, __shfl_xor(mn, o, 64));
    if ((threadIdx.x & 63) == 0 && mn < INFINITY)
        atomicMin(reinterpret_cast<unsigned int *>(state) + 4, __float_as_uint(mn));
}

__global__ void scale_setup_kernel(int nmax, float *state)
{
    if (threadIdx.x == 0) {
        const float mn = __uint_as_float(reinterpret_cast<const unsigned int *>(state)[4]);
        const float smin = mn < INFINITY ? fmaxf(mn, 1e-6f) : 1.0f;
        state[0] = 1e-2f;
        state[1] = 2.0f * (float)nmax / smin;
        state[2] = state[1];
        state[3] = 0.0f;
    }
}

__global__ __launch_bounds__(256) void scale_sums_kernel(const float *pred, long long step, long long m, float s_mult,
                                                         int nmin, int nmax, float inv_nref, const float *state,
                                                         double *sums)
{
    __shared__ float s_c[kScaleProbes];
    __shared__ float s_part[256 / 64][kScaleProbes];
    if (state[3] != 0.0f) return;
    if (threadIdx.x < kScaleProbes) s_c[threadIdx.x] = scale_probe(state[0], state[1], threadIdx.x);
    __syncthreads();
    float acc[kScaleProbes];
#pragma unroll
    for (int j = 0; j < kScaleProbes; ++j) acc[j] = 0.0f;
    for (long long i = (long long)blockIdx.x * 256 + threadIdx.x; i < m; i += (long long)gridDim.x * 256) {
        const float s = expf(pred[i * step]) * s_mult;
        const float s2 = s * s;
#pragma unroll
        for (int j = 0; j < kScaleProbes; ++j) {
            const float v = ceilf(s_c[j] * s);
            const float nj = fminf(fmaxf(v, (float)nmin), (float)nmax);
            acc[j] += s2 * (1.0f / nj - inv_nref);
        }
    }
    const int lane = threadIdx.x & 63, wid = threadIdx.x >> 6;
#pragma unroll
    for (int j = 0; j < kScaleProbes; ++j) {
        float v = acc[j];
        for (int o = 32; o > 0; o >>= 1) v += __shfl_xor(v, o, 64);
        if (lane == 0) s_part[wid][j] = v;
    }
    __syncthreads();
    if (threadIdx.x < kScaleProbes) {
        double t = 0.0;
        for (int w = 0; w < 256 / 64; ++w) t += (double)s_part[w][threadIdx.x];
        atomicAdd(&sums[threadIdx.x], t);
    }
}

__global__ void scale_pick_kernel(long long m, double target, int last, float *state, double *sums)
{
    __shared__ int s_first;
    if (threadIdx.x == 0) s_first = kScaleProbes;
    __syncthreads();
    if (state[3] == 0.0f && threadIdx.x < kScaleProbes && sums[threadIdx.x] / (double)m <= target)
        atomicMin(&s_first, (int)threadIdx.x);
    __syncthreads();
    if (threadIdx.x == 0 && state[3] == 0.0f) {
        const float lo = state[0], hi = state[1];
        const int k = s_first;
        if (k == kScaleProbes) {
            state[2] = scale_probe(lo, hi, kScaleProbes - 1);
            state[3] = 1.0f;
        } else if (k == 0) {
            state[2] = scale_probe(lo, hi, 0);
            state[3] = 1.0f;
        } else {
            state[0] = scale_probe(lo, hi, k - 1);
            state[1] = scale_probe(lo, hi, k);
            state[2] = state[1];
            if (last) state[3] = 1.0f;
        }
    }
    __syncthreads();
    if (threadIdx.x < kScaleProbes) sums[threadIdx.x] = 0.0;
}

__global__ void scale_out_kernel(const float *state, float s_mult, float *out)
{
    if (threadIdx.x == 0) {
        out[0] = state[2];
        out[1] = state[2] * s_mult;
    }
}

size_t budget_scale_state_bytes() { return 32 + sizeof(double) * kScaleProbes; }

hipError_t launch_budget_scale(const float *pred, long long n, float s_mult, double target, int nmin, int nmax,
                               int n_ref, int rounds, long long max_rows, void *state_buf, float *out, hipStream_t s)
{
    if (n < 1 || rounds < 1 || max_rows < 1) return hipErrorInvalidValue;
    const long long step = std::max<long long>(1, n / max_rows);
    const long long m = (n + step - 1) / step;
    float *state = static_cast<float *>(state_buf);
    double *sums = reinterpret_cast<double *>(static_cast<unsigned char *>(state_buf) + 32);
    const unsigned grid = (unsigned)std::min<long long>((m + 255) / 256, 1024);
    hipLaunchKernelGGL(scale_clear_kernel, dim3(1), dim3(kScaleProbes), 0, s, state, sums);
    hipError_t e = hipGetLastError();
    if (e == hipSuccess) {
        hipLaunchKernelGGL(scale_min_kernel, dim3(std::min(grid, 128u)), dim3(256), 0, s, pred, step, m, s_mult,
                           state);
        e = hipGetLastError();
    }
    if (e == hipSuccess) {
        hipLaunchKernelGGL(scale_setup_kernel, dim3(1), dim3(64), 0, s, nmax, state);
        e = hipGetLastError();
    }
    // each thread of the sums kernel folds several rows before the 64-probe
    // reduction across its wave, which costs about as much as 8 rows
    const unsigned sums_grid = (unsigned)std::min<long long>((m + 2047) / 2048, 128);
    for (int r = 0; r < rounds && e == hipSuccess; ++r) {
        hipLaunchKernelGGL(scale_sums_kernel, dim3(std::max(sums_grid, 1u)), dim3(256), 0, s, pred, step, m, s_mult,
                           nmin, nmax, 1.0f / (float)n_ref, state, sums);
        e = hipGetLastError();
        if (e != hipSuccess) break;
        hipLaunchKernelGGL(scale_pick_kernel, dim3(1), dim3(kScaleProbes), 0, s, m, target, r == rounds - 1 ? 1 : 0,
                           state, sums);
        e = hipGetLastError();
    }
    if (e == hipSuccess) {
        hipLaunchKernelGGL(scale_out_kernel, dim3(1), dim3(64), 0, s, state, s_mult, out);
        e = hipGetLastError();
    }
    return e;
}

hipError_t launch_mlp_forward(const float *P, const uint16_t *X, long long n, float *out, hipStream_t s)
{
    // waves loop over 16-row tiles.  Each wave first loads the whole network
    // into registers (138 VGPRs: 3 waves per SIMD), so the grid is one
    // resident round of waves (3 workgroups of 4 waves per CU on 256 CUs)
    // and each wave amortises that load over many tiles.
    const long long tiles = (n + 15) / 16;
    const unsigned grid = (unsigned)std::min<long long>((tiles + 3) / 4, 768);
    hipLaunchKernelGGL(mlp_forward_kernel, dim3(std::max(grid, 1u)), dim3(256), 0, s, P, X, n, out);
    return hipGetLastError();
}

hipError_t launch_mlp_grad(const float *P, const uint16_t *X, const float *target, long long n, float inv_total,
                           int loss_kind, float *grad, float *loss, hipStream_t s)
{
    const long long tiles = (n + 15) / 16;
    const unsigned grid = (unsigned)std::min<long long>((tiles + 3) / 4, 1024);
    hipLaunchKernelGGL(mlp_grad_kernel, dim3(std::max(grid, 1u)), dim3(256), 0, s, P, X, target, n, inv_total,
                       loss_kind, grad, loss);
    return hipGetLastError();
}

hipError_t launch_mlp_adam(float *P, const float *grad, float *m, float *v, float lr, float b1, float b2, float eps,
                           float bc1, float bc2, hipStream_t s)
{
    hipLaunchKernelGGL(mlp_adam_kernel, dim3((kMlpParams + 255) / 256), dim3(256), 0, s, P, grad, m, v, lr, b1, b2,
                       eps, bc1, bc2);
    return hipGetLastError();
}

hipError_t launch_features(const float *u, int in_w, int in_h, int out_w, int out_h, float zoom, float radius,
                           float sd, float sigma, int channel, uint16_t *X, hipStream_t s)
{
    const long long n = (long long)out_w * out_h;
    hipLaunchKernelGGL(feature_kernel, dim3((unsigned)((n + 255) / 256)), dim3(256), 0, s, u, in_w, in_h, out_w, out_h,
                       1.0f / zoom, radius / 0.3f, sd / 0.15f, sigma / 2.0f, zoom / 4.0f, channel, X);
    return hipGetLastError();
}

}  // namespace fg
