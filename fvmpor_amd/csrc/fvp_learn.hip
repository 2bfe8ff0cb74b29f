// Mesh message-passing network on gfx950: BF16 weights and activations, FP32
// accumulation on the matrix cores (v_mfma_f32_32x32x8bf16_1k), training
// (forward, backward, Adam) and inference.
//
//   H0 = bf16(relu(X0 We + be));  per layer: A = bf16(mean_N(H)),
//   Z = H W1 + A W2 + b,  H = bf16(H + relu(Z));   y = H Wd + bd
//
// on the node-adjacency graph of a mesh (64 features per node).  The matrix
// products are N x K by K x 64 (K = 16 or 64): one wavefront computes 32 rows x
// 64 columns as two 32 x 32 MFMA tiles, the weights staged transposed in LDS.
// Weight gradients (K x 64, a reduction over the N rows) use the same
// instruction with the rows as the reduction dimension.  No counterpart in the
// reference: it supplies the Newton initial iterate of the integrator.
#include <hip/hip_runtime.h>

#include <cmath>
#include <cstring>
#include <vector>

#include "fvp_internal.h"

using fvp::DevBuf;
using fvp::set_error;

typedef unsigned short bf16_t;
typedef short short4_t __attribute__((ext_vector_type(4)));
typedef float float16_t __attribute__((ext_vector_type(16)));

namespace {

constexpr int kF = 16;      // input features (padded)
constexpr int kH = 64;      // hidden width
constexpr int kMaxLayers = 8;

__device__ __forceinline__ float bf2f(bf16_t b) { return __uint_as_float((unsigned)b << 16); }
__device__ __forceinline__ bf16_t f2bf(float f) {  // round to nearest even (finite inputs)
    unsigned u = __float_as_uint(f);
    u += 0x7FFFu + ((u >> 16) & 1u);
    return (bf16_t)(u >> 16);
}

// ---- forward / backward products -------------------------------------------------
enum { EPI_RELU = 0, EPI_RESID = 1, EPI_ADD = 2, EPI_BF16 = 3 };

// Y (N x 64) from X1 (N x K1) W1 (K1 x 64) [+ X2 (N x K2) W2 (K2 x 64)] [+ bias]:
//   EPI_RELU:  Yb = bf16(relu(acc + bias))
//   EPI_RESID: Zb = bf16(acc + bias), Yb = bf16(Hin + relu(acc + bias))
//   EPI_ADD:   Yf = Yf + acc                          (FP32)
//   EPI_BF16:  Yb = bf16(acc)
// K1, K2 multiples of 8, at most 64.  256 threads = 4 wavefronts x 32 rows.
template <int EPI>
__global__ void __launch_bounds__(256) mp_gemm(int N, const bf16_t *__restrict__ X1, int K1,
                                               const bf16_t *__restrict__ W1, const bf16_t *__restrict__ X2, int K2,
                                               const bf16_t *__restrict__ W2, const float *__restrict__ bias,
                                               const bf16_t *__restrict__ Hin, bf16_t *__restrict__ Zb,
                                               bf16_t *__restrict__ Yb, float *__restrict__ Yf) {
    __shared__ __align__(16) bf16_t wt[2][kH * kH];  // transposed weights: wt[n * K + k]
    for (int i = threadIdx.x; i < K1 * kH; i += blockDim.x) {
        const int k = i / kH, n = i % kH;
        wt[0][n * K1 + k] = W1[i];
    }
    if (X2)
        for (int i = threadIdx.x; i < K2 * kH; i += blockDim.x) {
            const int k = i / kH, n = i % kH;
            wt[1][n * K2 + k] = W2[i];
        }
    __syncthreads();
    const int lane = threadIdx.x & 63, wave = threadIdx.x >> 6;
    const int r0 = blockIdx.x * 128 + wave * 32;
    const int i = lane & 31, kq = (lane >> 5) * 4;
    const int row = r0 + i;
    float16_t acc0 = {0}, acc1 = {0};
    for (int src = 0; src < (X2 ? 2 : 1); ++src) {
        const bf16_t *X = src ? X2 : X1;
        const int K = src ? K2 : K1;
        const bf16_t *w = wt[src];
        for (int kb = 0; kb < K; kb += 8) {
            // 4 consecutive BF16 values per lane: one 8-byte load (K and kb + kq are multiples of 4)
            short4_t a = {0, 0, 0, 0};
            if (row < N) a = *reinterpret_cast<const short4_t *>(X + (size_t)row * K + kb + kq);
            const short4_t b0 = *reinterpret_cast<const short4_t *>(w + i * K + kb + kq);
            const short4_t b1 = *reinterpret_cast<const short4_t *>(w + (32 + i) * K + kb + kq);
            acc0 = __builtin_amdgcn_mfma_f32_32x32x8bf16_1k(a, b0, acc0, 0, 0, 0);
            acc1 = __builtin_amdgcn_mfma_f32_32x32x8bf16_1k(a, b1, acc1, 0, 0, 0);
        }
    }
    // D layout: lane holds column lane % 32; register v holds row (v / 4) * 8 + (lane / 32) * 4 + v % 4
#pragma unroll
    for (int v = 0; v < 16; ++v) {
        const int m = (v >> 2) * 8 + (lane >> 5) * 4 + (v & 3);
        const int rr = r0 + m;
        if (rr >= N) continue;
#pragma unroll
        for (int h = 0; h < 2; ++h) {
            const int n = (lane & 31) + 32 * h;
            const float x = h ? acc1[v] : acc0[v];
            const size_t o = (size_t)rr * kH + n;
            if (EPI == EPI_RELU) Yb[o] = f2bf(fmaxf(x + bias[n], 0.f));
            else if (EPI == EPI_RESID) {
                const float z = x + bias[n];
                Zb[o] = f2bf(z);
                Yb[o] = f2bf(bf2f(Hin[o]) + fmaxf(z, 0.f));
            } else if (EPI == EPI_ADD) Yf[o] += x;
            else Yb[o] = f2bf(x);
        }
    }
}

// Block partials of G (K x 64) = X^T D over a chunk of rows: X (N x K) and
// D (N x 64) in BF16, the rows the MFMA reduction dimension.  K = 16 or 64;
// a 32-feature tile with K = 16 is padded with zeros.  part: gridDim.x x K x 64.
__global__ void __launch_bounds__(256) mp_wgrad(int N, int rows_per_block, const bf16_t *__restrict__ X, int K,
                                                const bf16_t *__restrict__ D, float *__restrict__ part) {
    __shared__ float red[4][kH * kH];
    const int lane = threadIdx.x & 63, wave = threadIdx.x >> 6;
    const int mt = K > 32 ? 2 : 1;  // feature tiles of 32
    float16_t acc[2][2];
#pragma unroll
    for (int a = 0; a < 2; ++a)
#pragma unroll
        for (int b = 0; b < 2; ++b) acc[a][b] = float16_t{0};
    const int rb = blockIdx.x * rows_per_block, re = min(N, rb + rows_per_block);
    const int m0 = lane & 31, kq = (lane >> 5) * 4;
    for (int r = rb + wave * 8; r < re; r += 32) {
        short4_t bfr[2];
#pragma unroll
        for (int h = 0; h < 2; ++h) {
            short s[4];
#pragma unroll
            for (int j = 0; j < 4; ++j) {
                const int row = r + kq + j;
                s[j] = row < re ? (short)D[(size_t)row * kH + m0 + 32 * h] : (short)0;
            }
            bfr[h] = short4_t{s[0], s[1], s[2], s[3]};
        }
#pragma unroll
        for (int t = 0; t < 2; ++t) {
            if (t >= mt) break;
            short s[4];
            const int m = m0 + 32 * t;
#pragma unroll
            for (int j = 0; j < 4; ++j) {
                const int row = r + kq + j;
                s[j] = (row < re && m < K) ? (short)X[(size_t)row * K + m] : (short)0;
            }
            const short4_t afr = {s[0], s[1], s[2], s[3]};
            acc[t][0] = __builtin_amdgcn_mfma_f32_32x32x8bf16_1k(afr, bfr[0], acc[t][0], 0, 0, 0);
            acc[t][1] = __builtin_amdgcn_mfma_f32_32x32x8bf16_1k(afr, bfr[1], acc[t][1], 0, 0, 0);
        }
    }
    // per-wavefront tiles into LDS, then the four wavefronts summed
#pragma unroll
    for (int t = 0; t < 2; ++t) {
        if (t >= mt) break;
#pragma unroll
        for (int h = 0; h < 2; ++h)
#pragma unroll
            for (int v = 0; v < 16; ++v) {
                const int m = 32 * t + (v >> 2) * 8 + (lane >> 5) * 4 + (v & 3);
                const int n = (lane & 31) + 32 * h;
                if (m < K) red[wave][m * kH + n] = acc[t][h][v];
            }
    }
    __syncthreads();
    for (int idx = threadIdx.x; idx < K * kH; idx += blockDim.x)
        part[(size_t)blockIdx.x * K * kH + idx] = red[0][idx] + red[1][idx] + red[2][idx] + red[3][idx];
}

// out[i] += sum over the blocks b of chunk blockIdx.y of part[b * n + i]
// (out zeroed beforehand; one FP32 atomic per output and chunk)
__global__ void sum_parts(int nparts, int n, int per, const float *__restrict__ part, float *__restrict__ out) {
    const int i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= n) return;
    const int b0 = blockIdx.y * per, b1 = min(nparts, b0 + per);
    float s = 0.f;
    for (int b = b0; b < b1; ++b) s += part[(size_t)b * n + i];
    atomicAdd(&out[i], s);
}

// column sums of D (N x 64, BF16) into block partials part[b * 64 + n]
__global__ void __launch_bounds__(256) mp_colsum(int N, int rows_per_block, const bf16_t *__restrict__ D,
                                                 float *__restrict__ part) {
    __shared__ float red[4][kH];
    const int n = threadIdx.x & 63, q = threadIdx.x >> 6;
    const int rb = blockIdx.x * rows_per_block, re = min(N, rb + rows_per_block);
    float s = 0.f;
    for (int r = rb + q; r < re; r += 4) s += bf2f(D[(size_t)r * kH + n]);
    red[q][n] = s;
    __syncthreads();
    if (threadIdx.x < kH) part[(size_t)blockIdx.x * kH + n] = red[0][n] + red[1][n] + red[2][n] + red[3][n];
}

// neighbour mean: out_i = bf16((1/deg_i) sum_j H_j); transposed (scale_src):
// outf_i (+)= sum_j G_j / deg_j (FP32 out).  8 lanes per node, 8 features
// (one 16-byte load per neighbour) each.
template <bool TRANSPOSE>
__global__ void __launch_bounds__(256) mp_aggregate(int N, const int *__restrict__ ia, const int *__restrict__ ja,
                                                    const bf16_t *__restrict__ H, bf16_t *__restrict__ outb,
                                                    float *__restrict__ outf) {
    const long gid = blockIdx.x * (long)blockDim.x + threadIdx.x;
    const int i = (int)(gid >> 3), f = (int)(gid & 7) * 8;
    if (i >= N) return;
    float s[8] = {0.f, 0.f, 0.f, 0.f, 0.f, 0.f, 0.f, 0.f};
    const int k1 = ia[i + 1];
    for (int k = ia[i]; k < k1; ++k) {
        const int j = ja[k];
        const uint4 hv = *reinterpret_cast<const uint4 *>(H + (size_t)j * kH + f);
        const float w = TRANSPOSE ? 1.f / (float)max(1, ia[j + 1] - ia[j]) : 1.f;
        const unsigned u[4] = {hv.x, hv.y, hv.z, hv.w};
#pragma unroll
        for (int q = 0; q < 4; ++q) {
            s[2 * q] += w * __uint_as_float(u[q] << 16);
            s[2 * q + 1] += w * __uint_as_float(u[q] & 0xFFFF0000u);
        }
    }
    if (TRANSPOSE) {
        float4 *o = reinterpret_cast<float4 *>(outf + (size_t)i * kH + f);
        float4 a = o[0], b = o[1];
        a.x += s[0]; a.y += s[1]; a.z += s[2]; a.w += s[3];
        b.x += s[4]; b.y += s[5]; b.z += s[6]; b.w += s[7];
        o[0] = a;
        o[1] = b;
    } else {
        const float inv = 1.f / (float)max(1, k1 - ia[i]);
        unsigned u[4];
#pragma unroll
        for (int q = 0; q < 4; ++q)
            u[q] = (unsigned)f2bf(s[2 * q] * inv) | ((unsigned)f2bf(s[2 * q + 1] * inv) << 16);
        *reinterpret_cast<uint4 *>(outb + (size_t)i * kH + f) = make_uint4(u[0], u[1], u[2], u[3]);
    }
}

// y_i = H_i . Wd + bd  (FP32); 8 lanes per node, 8 features each, summed
// across the 8 lanes
__global__ void __launch_bounds__(256) mp_decode(int N, const bf16_t *__restrict__ H, const bf16_t *__restrict__ wd,
                                                 const float *__restrict__ bd, float *__restrict__ y) {
    const long gid = blockIdx.x * (long)blockDim.x + threadIdx.x;
    const int i = (int)(gid >> 3), f = (int)(gid & 7) * 8;
    float s = 0.f;
    if (i < N) {
        const uint4 hv = *reinterpret_cast<const uint4 *>(H + (size_t)i * kH + f);
        const unsigned u[4] = {hv.x, hv.y, hv.z, hv.w};
#pragma unroll
        for (int q = 0; q < 4; ++q) {
            s += __uint_as_float(u[q] << 16) * bf2f(wd[f + 2 * q]);
            s += __uint_as_float(u[q] & 0xFFFF0000u) * bf2f(wd[f + 2 * q + 1]);
        }
    }
    s += __shfl_xor(s, 1, 8);
    s += __shfl_xor(s, 2, 8);
    s += __shfl_xor(s, 4, 8);
    if (i < N && (gid & 7) == 0) y[i] = s + bd[0];
}

// loss partials and dy = 2 (y - t) / N
__global__ void __launch_bounds__(256) mp_loss(int N, const float *__restrict__ y, const float *__restrict__ t,
                                               float *__restrict__ dy, double *__restrict__ part) {
    __shared__ double red[256];
    const int i = blockIdx.x * blockDim.x + threadIdx.x;
    double e = 0.;
    if (i < N) {
        const float d = y[i] - t[i];
        dy[i] = 2.f * d / (float)N;
        e = (double)d * d;
    }
    red[threadIdx.x] = e;
    __syncthreads();
    for (int s = 128; s > 0; s >>= 1) {
        if (threadIdx.x < s) red[threadIdx.x] += red[threadIdx.x + s];
        __syncthreads();
    }
    if (threadIdx.x == 0) part[blockIdx.x] = red[0];
}

// decoder backward: dH_i = dy_i * Wd (FP32); partials of dWd = H^T dy and dbd = sum dy
__global__ void __launch_bounds__(256) mp_decode_bwd(int N, const bf16_t *__restrict__ H, const bf16_t *__restrict__ wd,
                                                     const float *__restrict__ dy, float *__restrict__ dH,
                                                     float *__restrict__ part) {
    __shared__ float red[kH + 1][4];
    const int f = threadIdx.x & 63, q = threadIdx.x >> 6;
    const int rows = (N + gridDim.x - 1) / gridDim.x;
    const int rb = blockIdx.x * rows, re = min(N, rb + rows);
    float s = 0.f, sb = 0.f;
    const float w = bf2f(wd[f]);
    for (int i = rb + q; i < re; i += 4) {
        const float d = dy[i];
        dH[(size_t)i * kH + f] = d * w;
        s += bf2f(H[(size_t)i * kH + f]) * d;
        if (f == 0) sb += d;
    }
    red[f][q] = s;
    if (f == 0) red[kH][q] = sb;
    __syncthreads();
    if (threadIdx.x <= kH) {
        const int g = threadIdx.x;
        part[(size_t)blockIdx.x * (kH + 1) + g] = red[g][0] + red[g][1] + red[g][2] + red[g][3];
    }
}

// dZ = bf16(dH * (Z > 0)) with Z given in BF16 (or as the encoder output H0 > 0)
__global__ void mp_relu_bwd(long n, const float *__restrict__ dH, const bf16_t *__restrict__ Z, bf16_t *__restrict__ dZ) {
    const long i = blockIdx.x * (long)blockDim.x + threadIdx.x;
    if (i >= n) return;
    dZ[i] = bf2f(Z[i]) > 0.f ? f2bf(dH[i]) : (bf16_t)0;
}

__global__ void f32_to_bf16(long n, const float *__restrict__ x, bf16_t *__restrict__ y) {
    const long i = blockIdx.x * (long)blockDim.x + threadIdx.x;
    if (i < n) y[i] = f2bf(x[i]);
}

// Adam on the FP32 master parameters, then their BF16 copy
__global__ void mp_adam(long n, float lr, float b1, float b2, float eps, float c1, float c2, float *__restrict__ p,
                        float *__restrict__ m, float *__restrict__ v, const float *__restrict__ g,
                        bf16_t *__restrict__ pb) {
    const long i = blockIdx.x * (long)blockDim.x + threadIdx.x;
    if (i >= n) return;
    const float gi = g[i];
    const float mi = b1 * m[i] + (1.f - b1) * gi, vi = b2 * v[i] + (1.f - b2) * gi * gi;
    m[i] = mi;
    v[i] = vi;
    const float pi = p[i] - lr * (mi / c1) / (sqrtf(vi / c2) + eps);
    p[i] = pi;
    pb[i] = f2bf(pi);
}

// 64 x 64 transpose of the BF16 weights (the backward products use W^T)
__global__ void bf16_transpose64(const bf16_t *__restrict__ w, bf16_t *__restrict__ wt) {
    const int i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i < kH * kH) wt[(i % kH) * kH + i / kH] = w[i];
}

inline int nb(long n, int bs) { return (int)((n + bs - 1) / bs); }

}  // namespace

struct fvp_mpnn {
    int device = 0, N = 0, L = 3;
    hipStream_t stream = nullptr, own_stream = nullptr;
    DevBuf<int> ia, ja;
    // parameters (FP32 master, BF16 copy, Adam moments, gradients), one flat layout:
    // We (16 x 64), be (64), per layer W1 (64 x 64), W2 (64 x 64), b (64), Wd (64), bd (1)
    long np = 0;
    long o_We = 0, o_be = 0, o_W1[kMaxLayers], o_W2[kMaxLayers], o_b[kMaxLayers], o_Wd = 0, o_bd = 0;
    DevBuf<float> p, m, v, g;
    DevBuf<unsigned short> pb, wt1, wt2;  // BF16 copy, transposed W1 / W2 per layer
    long step = 0;
    // activations
    DevBuf<unsigned short> X0, H, A, Z, dZ, dA;  // H: (L+1) x N x 64; A, Z: L x N x 64
    DevBuf<float> y, dy, dH, part;
    DevBuf<double> lpart;
    int nparts = 0;
};

static void layout(fvp_mpnn *m) {
    long o = 0;
    m->o_We = o; o += kF * kH;
    m->o_be = o; o += kH;
    for (int l = 0; l < m->L; ++l) {
        m->o_W1[l] = o; o += kH * kH;
        m->o_W2[l] = o; o += kH * kH;
        m->o_b[l] = o; o += kH;
    }
    m->o_Wd = o; o += kH;
    m->o_bd = o; o += 1;
    m->np = o;
}

static int refresh_bf16(fvp_mpnn *m) {
    hipStream_t s = m->stream;
    f32_to_bf16<<<nb(m->np, 256), 256, 0, s>>>(m->np, m->p.p, m->pb.p);
    for (int l = 0; l < m->L; ++l) {
        bf16_transpose64<<<nb(kH * kH, 256), 256, 0, s>>>(m->pb.p + m->o_W1[l], m->wt1.p + (size_t)l * kH * kH);
        bf16_transpose64<<<nb(kH * kH, 256), 256, 0, s>>>(m->pb.p + m->o_W2[l], m->wt2.p + (size_t)l * kH * kH);
    }
    FVP_HIP(hipGetLastError());
    return 0;
}

// forward on X0 (already BF16 in m->X0); keeps the activations; y into m->y
static int forward_impl(fvp_mpnn *m) {
    hipStream_t s = m->stream;
    const int N = m->N;
    const size_t NH = (size_t)N * kH;
    const int g = nb(N, 128);
    const bf16_t *pb = m->pb.p;
    mp_gemm<EPI_RELU><<<g, 256, 0, s>>>(N, m->X0.p, kF, pb + m->o_We, nullptr, 0, nullptr, m->p.p + m->o_be, nullptr,
                                        nullptr, m->H.p, nullptr);
    for (int l = 0; l < m->L; ++l) {
        const bf16_t *Hl = m->H.p + l * NH;
        bf16_t *Al = m->A.p + l * NH;
        mp_aggregate<false><<<nb((long)N * 8, 256), 256, 0, s>>>(N, m->ia.p, m->ja.p, Hl, Al, nullptr);
        mp_gemm<EPI_RESID><<<g, 256, 0, s>>>(N, Hl, kH, pb + m->o_W1[l], Al, kH, pb + m->o_W2[l], m->p.p + m->o_b[l],
                                             Hl, m->Z.p + l * NH, m->H.p + (l + 1) * NH, nullptr);
    }
    mp_decode<<<nb((long)N * 8, 256), 256, 0, s>>>(N, m->H.p + m->L * NH, pb + m->o_Wd, m->p.p + m->o_bd, m->y.p);
    FVP_HIP(hipGetLastError());
    return 0;
}

constexpr int kRowsPerBlock = 512;

// partial sums of nparts blocks reduced into out (zeroed by the caller)
static void reduce_parts(fvp_mpnn *m, int nparts, int n, float *out) {
    const int chunks = std::min(16, nparts), per = (nparts + chunks - 1) / chunks;
    sum_parts<<<dim3(nb(n, 256), chunks), 256, 0, m->stream>>>(nparts, n, per, m->part.p, out);
}

static int wgrad(fvp_mpnn *m, const bf16_t *X, int K, const bf16_t *D, float *out) {
    const int nbk = nb(m->N, kRowsPerBlock);
    mp_wgrad<<<nbk, 256, 0, m->stream>>>(m->N, kRowsPerBlock, X, K, D, m->part.p);
    reduce_parts(m, nbk, K * kH, out);
    return 0;
}

static int colsum(fvp_mpnn *m, const bf16_t *D, float *out) {
    const int nbk = nb(m->N, kRowsPerBlock);
    mp_colsum<<<nbk, 256, 0, m->stream>>>(m->N, kRowsPerBlock, D, m->part.p);
    reduce_parts(m, nbk, kH, out);
    return 0;
}

static int backward_impl(fvp_mpnn *m) {
    hipStream_t s = m->stream;
    const int N = m->N, L = m->L;
    const size_t NH = (size_t)N * kH;
    const int g = nb(N, 128);
    float *G = m->g.p;
    FVP_HIP(hipMemsetAsync(G, 0, sizeof(float) * m->np, s));  // every gradient below accumulates
    // decoder
    const int nbd = std::min(1024, nb(N, 256));
    mp_decode_bwd<<<nbd, 256, 0, s>>>(N, m->H.p + L * NH, m->pb.p + m->o_Wd, m->dy.p, m->dH.p, m->part.p);
    reduce_parts(m, nbd, kH + 1, G + m->o_Wd);  // Wd (64) then bd (1): adjacent
    for (int l = L - 1; l >= 0; --l) {
        const bf16_t *Hin = m->H.p + l * NH, *Al = m->A.p + l * NH, *Zl = m->Z.p + l * NH;
        mp_relu_bwd<<<nb((long)NH, 256), 256, 0, s>>>((long)NH, m->dH.p, Zl, m->dZ.p);
        wgrad(m, Hin, kH, m->dZ.p, G + m->o_W1[l]);
        wgrad(m, Al, kH, m->dZ.p, G + m->o_W2[l]);
        colsum(m, m->dZ.p, G + m->o_b[l]);
        // dA = bf16(dZ W2^T); dH += dZ W1^T + aggregate^T(dA)
        mp_gemm<EPI_BF16><<<g, 256, 0, s>>>(N, m->dZ.p, kH, m->wt2.p + (size_t)l * kH * kH, nullptr, 0, nullptr,
                                            nullptr, nullptr, nullptr, m->dA.p, nullptr);
        mp_gemm<EPI_ADD><<<g, 256, 0, s>>>(N, m->dZ.p, kH, m->wt1.p + (size_t)l * kH * kH, nullptr, 0, nullptr,
                                           nullptr, nullptr, nullptr, nullptr, m->dH.p);
        mp_aggregate<true><<<nb((long)N * 8, 256), 256, 0, s>>>(N, m->ia.p, m->ja.p, m->dA.p, nullptr, m->dH.p);
    }
    // encoder: relu mask from H0 > 0
    mp_relu_bwd<<<nb((long)NH, 256), 256, 0, s>>>((long)NH, m->dH.p, m->H.p, m->dZ.p);
    wgrad(m, m->X0.p, kF, m->dZ.p, G + m->o_We);
    colsum(m, m->dZ.p, G + m->o_be);
    FVP_HIP(hipGetLastError());
    return 0;
}

namespace fvp {
int mpnn_nodes(const fvp_mpnn *m) { return m ? m->N : -1; }
}  // namespace fvp

// ---- C ABI ---------------------------------------------------------------------------
extern "C" int fvp_mpnn_create(int device, int n_nodes, const int *ia, const int *ja, int layers,
                               unsigned long long seed, fvp_mpnn **out) {
    if (!ia || !ja || !out || n_nodes <= 0) { set_error("null argument"); return -1; }
    if (layers < 1 || layers > kMaxLayers) { set_error("layers must be 1..8"); return -1; }
    FVP_HIP(hipSetDevice(device));
    fvp_mpnn *m = new fvp_mpnn();
    m->device = device;
    m->N = n_nodes;
    m->L = layers;
    layout(m);
    if (hipStreamCreateWithFlags(&m->own_stream, hipStreamNonBlocking) != hipSuccess) {
        delete m;
        set_error("stream creation failed");
        return -1;
    }
    m->stream = m->own_stream;
    const size_t NH = (size_t)n_nodes * kH;
    // He / Glorot-like initial weights from a fixed-seed generator, zero biases
    std::vector<float> p0(m->np, 0.f);
    unsigned long long st = seed * 6364136223846793005ULL + 1442695040888963407ULL;
    auto gauss = [&]() {
        auto u = [&]() {
            st = st * 6364136223846793005ULL + 1442695040888963407ULL;
            return ((st >> 11) + 0.5) * (1.0 / 9007199254740992.0);
        };
        const double a = u(), b = u();
        return std::sqrt(-2.0 * std::log(a)) * std::cos(6.283185307179586 * b);
    };
    for (long k = 0; k < kF * kH; ++k) p0[m->o_We + k] = (float)(gauss() * std::sqrt(2.0 / kF));
    for (int l = 0; l < layers; ++l)
        for (long k = 0; k < kH * kH; ++k) {
            p0[m->o_W1[l] + k] = (float)(gauss() * std::sqrt(1.0 / kH));
            p0[m->o_W2[l] + k] = (float)(gauss() * std::sqrt(1.0 / kH));
        }
    for (long k = 0; k < kH; ++k) p0[m->o_Wd + k] = (float)(gauss() * std::sqrt(1.0 / kH));
    const int nparts = std::max(nb(n_nodes, kRowsPerBlock), 1024) * (kF > kH ? kF : kH) * kH;
    if (m->ia.upload(ia, n_nodes + 1) || m->ja.upload(ja, ia[n_nodes]) || m->p.upload(p0.data(), m->np) ||
        m->m.alloc(m->np) || m->v.alloc(m->np) || m->g.alloc(m->np) || m->pb.alloc(m->np) ||
        m->wt1.alloc((size_t)layers * kH * kH) || m->wt2.alloc((size_t)layers * kH * kH) ||
        m->X0.alloc((size_t)n_nodes * kF) || m->H.alloc((layers + 1) * NH) || m->A.alloc(layers * NH) ||
        m->Z.alloc(layers * NH) || m->dZ.alloc(NH) || m->dA.alloc(NH) || m->y.alloc(n_nodes) ||
        m->dy.alloc(n_nodes) || m->dH.alloc(NH) || m->part.alloc(nparts) || m->lpart.alloc(nb(n_nodes, 256))) {
        fvp_mpnn_destroy(m);
        return -1;
    }
    FVP_HIP(hipMemset(m->m.p, 0, sizeof(float) * m->np));
    FVP_HIP(hipMemset(m->v.p, 0, sizeof(float) * m->np));
    FVP_HIP(hipMemset(m->g.p, 0, sizeof(float) * m->np));
    if (refresh_bf16(m)) { fvp_mpnn_destroy(m); return -1; }
    FVP_HIP(hipStreamSynchronize(m->stream));
    *out = m;
    return 0;
}

extern "C" int fvp_mpnn_destroy(fvp_mpnn *m) {
    if (!m) return 0;
    (void)hipSetDevice(m->device);
    if (m->stream) (void)hipStreamSynchronize(m->stream);
    m->ia.release(); m->ja.release(); m->p.release(); m->m.release(); m->v.release(); m->g.release();
    m->pb.release(); m->wt1.release(); m->wt2.release(); m->X0.release(); m->H.release(); m->A.release();
    m->Z.release(); m->dZ.release(); m->dA.release(); m->y.release(); m->dy.release(); m->dH.release();
    m->part.release(); m->lpart.release();
    if (m->own_stream) (void)hipStreamDestroy(m->own_stream);
    delete m;
    return 0;
}

extern "C" long fvp_mpnn_num_params(const fvp_mpnn *m) { return m ? m->np : -1; }

extern "C" int fvp_mpnn_set_stream(fvp_mpnn *m, void *stream) {
    if (!m) { set_error("null model"); return -1; }
    FVP_HIP(hipSetDevice(m->device));
    FVP_HIP(hipStreamSynchronize(m->stream));
    m->stream = stream ? (hipStream_t)stream : m->own_stream;
    return 0;
}

extern "C" int fvp_mpnn_get_params(fvp_mpnn *m, float *host) {
    if (!m || !host) { set_error("null argument"); return -1; }
    FVP_HIP(hipSetDevice(m->device));
    FVP_HIP(hipStreamSynchronize(m->stream));
    FVP_HIP(hipMemcpy(host, m->p.p, sizeof(float) * m->np, hipMemcpyDeviceToHost));
    return 0;
}

extern "C" int fvp_mpnn_set_params(fvp_mpnn *m, const float *host) {
    if (!m || !host) { set_error("null argument"); return -1; }
    FVP_HIP(hipSetDevice(m->device));
    FVP_HIP(hipStreamSynchronize(m->stream));
    FVP_HIP(hipMemcpy(m->p.p, host, sizeof(float) * m->np, hipMemcpyHostToDevice));
    if (refresh_bf16(m)) return -1;
    FVP_HIP(hipStreamSynchronize(m->stream));
    return 0;
}

extern "C" int fvp_mpnn_get_grads(fvp_mpnn *m, float *host) {
    if (!m || !host) { set_error("null argument"); return -1; }
    FVP_HIP(hipSetDevice(m->device));
    FVP_HIP(hipStreamSynchronize(m->stream));
    FVP_HIP(hipMemcpy(host, m->g.p, sizeof(float) * m->np, hipMemcpyDeviceToHost));
    return 0;
}

// x: N x 16 FP32 features (device); y: N outputs (device, FP32)
extern "C" int fvp_mpnn_forward(fvp_mpnn *m, const float *x, float *y) {
    if (!m || !x || !y) { set_error("null argument"); return -1; }
    FVP_HIP(hipSetDevice(m->device));
    f32_to_bf16<<<nb((long)m->N * kF, 256), 256, 0, m->stream>>>((long)m->N * kF, x, m->X0.p);
    if (forward_impl(m)) return -1;
    FVP_HIP(hipMemcpyAsync(y, m->y.p, sizeof(float) * m->N, hipMemcpyDeviceToDevice, m->stream));
    return 0;
}

// forward + backward of the mean squared error against t (device, N); the
// gradients stay on the device (fvp_mpnn_get_grads, fvp_mpnn_adam)
extern "C" int fvp_mpnn_loss_grad(fvp_mpnn *m, const float *x, const float *t, double *loss) {
    if (!m || !x || !t) { set_error("null argument"); return -1; }
    FVP_HIP(hipSetDevice(m->device));
    const int N = m->N;
    f32_to_bf16<<<nb((long)N * kF, 256), 256, 0, m->stream>>>((long)N * kF, x, m->X0.p);
    if (forward_impl(m)) return -1;
    const int nl = nb(N, 256);
    mp_loss<<<nl, 256, 0, m->stream>>>(N, m->y.p, t, m->dy.p, m->lpart.p);
    if (backward_impl(m)) return -1;
    if (loss) {
        std::vector<double> lp(nl);
        FVP_HIP(hipMemcpyAsync(lp.data(), m->lpart.p, sizeof(double) * nl, hipMemcpyDeviceToHost, m->stream));
        FVP_HIP(hipStreamSynchronize(m->stream));
        double s = 0.;
        for (double v : lp) s += v;
        *loss = s / N;
    }
    return 0;
}

extern "C" int fvp_mpnn_adam(fvp_mpnn *m, float lr) {
    if (!m) { set_error("null model"); return -1; }
    FVP_HIP(hipSetDevice(m->device));
    ++m->step;
    const float b1 = 0.9f, b2 = 0.999f;
    const float c1 = 1.f - (float)std::pow(b1, (double)m->step), c2 = 1.f - (float)std::pow(b2, (double)m->step);
    mp_adam<<<nb(m->np, 256), 256, 0, m->stream>>>(m->np, lr, b1, b2, 1e-8f, c1, c2, m->p.p, m->m.p, m->v.p, m->g.p,
                                                   m->pb.p);
    for (int l = 0; l < m->L; ++l) {
        bf16_transpose64<<<nb(kH * kH, 256), 256, 0, m->stream>>>(m->pb.p + m->o_W1[l], m->wt1.p + (size_t)l * kH * kH);
        bf16_transpose64<<<nb(kH * kH, 256), 256, 0, m->stream>>>(m->pb.p + m->o_W2[l], m->wt2.p + (size_t)l * kH * kH);
    }
    FVP_HIP(hipGetLastError());
    return 0;
}

extern "C" int fvp_mpnn_synchronize(fvp_mpnn *m) {
    if (!m) { set_error("null model"); return -1; }
    FVP_HIP(hipStreamSynchronize(m->stream));
    return 0;
}
