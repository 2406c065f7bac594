// Kernels of the pose-proposal network (CDNA4 / gfx950): bf16 weights and
// activations in NHWC order, fp32 accumulation.
//
// The convolutions are implicit GEMMs on v_mfma_f32_16x16x16_bf16 with the
// weights as the A operand and the activations as B, so that every operand is
// one 8-byte load per lane and every result one 8-byte store:
//   A[i = co][k = ci]   lane l: co = l%16, ci = 4(l/16)+v  -> w[co][tap][ci..ci+3]
//   B[k = ci][j = px]   lane l: px = l%16, ci = 4(l/16)+v  -> x[px + tap offset][ci..ci+3]
//   D[i = co][j = px]   lane l: co = 4(l/16)+v, px = l%16  -> y[px][co..co+3]
// A wavefront owns 64 output channels x 64 pixels (16 accumulator blocks); a
// workgroup is 4 wavefronts along the pixels.  The same kernel computes the
// input gradient of a layer from the flipped weights.  The weight gradient is
// a reduction over pixels and runs on the vector ALUs from LDS tiles, split
// over pixel ranges into partial sums that a second kernel adds in a fixed
// order.  No reference counterpart: the reference has no learned component.
#include "isac_internal.h"
#include "../../include/isac_posenet.h"

#include <algorithm>
#include <cfloat>
#include <cmath>
#include <climits>
#include <cstdio>

namespace isac {
namespace {

typedef short short4_t __attribute__((ext_vector_type(4)));
typedef float float4_t __attribute__((ext_vector_type(4)));

constexpr int kWave = 64;

__device__ __forceinline__ float bf2f(unsigned short h) { return __uint_as_float((unsigned)h << 16); }

// fp32 -> bf16, round to nearest even (NaN stays NaN).
__device__ __forceinline__ unsigned short f2bf(float x) {
    unsigned u = __float_as_uint(x);
    if ((u & 0x7fffffffu) > 0x7f800000u) return (unsigned short)((u >> 16) | 0x40u);
    u += 0x7fffu + ((u >> 16) & 1u);
    return (unsigned short)(u >> 16);
}

__device__ __forceinline__ uint2 pack4(float a, float b, float c, float d) {
    return make_uint2((unsigned)f2bf(a) | ((unsigned)f2bf(b) << 16),
                      (unsigned)f2bf(c) | ((unsigned)f2bf(d) << 16));
}

__device__ __forceinline__ float4_t mfma_bf16(uint2 a, uint2 b, float4_t c) {
    return __builtin_amdgcn_mfma_f32_16x16x16bf16_1k(__builtin_bit_cast(short4_t, a),
                                                      __builtin_bit_cast(short4_t, b), c, 0, 0, 0);
}

// ------------------------------------------------------------ standardize
__global__ __launch_bounds__(256) void k_standardize(const float *__restrict__ img, int D,
                                                     unsigned short *__restrict__ x0) {
    __shared__ double red[2][4];
    const int i = blockIdx.x, tid = threadIdx.x, lane = tid & 63, wave = tid >> 6;
    const float *src = img + (size_t)i * D * D;
    double s = 0.0, q = 0.0;
    for (int k = tid; k < D * D; k += 256) {
        const double v = (double)src[k];
        s += v;
        q += v * v;
    }
#pragma unroll
    for (int off = 32; off >= 1; off >>= 1) {
        s += __shfl_xor(s, off, kWave);
        q += __shfl_xor(q, off, kWave);
    }
    if (lane == 0) {
        red[0][wave] = s;
        red[1][wave] = q;
    }
    __syncthreads();
    s = red[0][0] + red[0][1] + red[0][2] + red[0][3];
    q = red[1][0] + red[1][1] + red[1][2] + red[1][3];
    const double nn = (double)D * D;
    const double mu = s / nn;
    const double sd = sqrt(fmax(q / nn - mu * mu, 0.0));
    const float fm = (float)mu, fi = (float)(1.0 / (sd + 1e-6));
    unsigned short *dst = x0 + (size_t)i * D * D;
    for (int k = tid; k < D * D; k += 256) dst[k] = f2bf((src[k] - fm) * fi);
}

// ------------------------------------------------------------ conv1 (Cin = 1)
// One thread per pixel computes 32 channels; weights (bf16-rounded) and bias in LDS.
__global__ __launch_bounds__(256) void k_conv1(const unsigned short *__restrict__ x0, int n, int H,
                                               int W, const float *__restrict__ master,
                                               unsigned short *__restrict__ y) {
    constexpr int CO = 32;
    __shared__ float w[9][CO];
    __shared__ float b[CO];
    const int tid = threadIdx.x;
    for (int e = tid; e < CO * 9; e += 256) {
        const int co = e / 9, t = e % 9;
        w[t][co] = bf2f(f2bf(master[co * 10 + t]));
    }
    if (tid < CO) b[tid] = master[tid * 10 + 9];
    __syncthreads();
    const long long P = (long long)blockIdx.x * 256 + tid;
    if (P >= (long long)n * H * W) return;
    const int img = (int)(P / ((long long)H * W));
    const int rem = (int)(P - (long long)img * H * W);
    const int yy = rem / W, xx = rem % W;
    float xin[9];
#pragma unroll
    for (int t = 0; t < 9; ++t) {
        const int sy = yy + t / 3 - 1, sx = xx + t % 3 - 1;
        xin[t] = (sy >= 0 && sy < H && sx >= 0 && sx < W)
                     ? bf2f(x0[((size_t)img * H + sy) * W + sx])
                     : 0.0f;
    }
    uint4 *dst = reinterpret_cast<uint4 *>(y + (size_t)P * CO);
#pragma unroll
    for (int c8 = 0; c8 < CO / 8; ++c8) {
        float o[8];
#pragma unroll
        for (int e = 0; e < 8; ++e) {
            float acc = 0.0f;
#pragma unroll
            for (int t = 0; t < 9; ++t) acc += w[t][c8 * 8 + e] * xin[t];
            o[e] = fmaxf(acc + b[c8 * 8 + e], 0.0f);
        }
        const uint2 lo = pack4(o[0], o[1], o[2], o[3]), hi = pack4(o[4], o[5], o[6], o[7]);
        dst[c8] = make_uint4(lo.x, lo.y, hi.x, hi.y);
    }
}

// conv1 with the 2x2 max pool fused (inference): one thread per pooled pixel
// computes its four conv1 outputs exactly as k_conv1 does (bf16-rounded after
// bias and ReLU) and keeps the first maximum, so the result equals
// k_maxpool2(k_conv1(x)).
__global__ __launch_bounds__(256) void k_conv1_pool(const unsigned short *__restrict__ x0, int n,
                                                    int H, int W, const float *__restrict__ master,
                                                    unsigned short *__restrict__ a) {
    constexpr int CO = 32;
    __shared__ float w[9][CO];
    __shared__ float b[CO];
    const int tid = threadIdx.x;
    for (int e = tid; e < CO * 9; e += 256) {
        const int co = e / 9, t = e % 9;
        w[t][co] = bf2f(f2bf(master[co * 10 + t]));
    }
    if (tid < CO) b[tid] = master[tid * 10 + 9];
    __syncthreads();
    const int Ho = H / 2, Wo = W / 2;
    const long long Q = (long long)blockIdx.x * 256 + tid;
    if (Q >= (long long)n * Ho * Wo) return;
    const int img = (int)(Q / ((long long)Ho * Wo));
    const int rem = (int)(Q - (long long)img * Ho * Wo);
    const int yo = rem / Wo, xo = rem % Wo;
    float xin[4][4];   // the 4x4 input patch under the window's four 3x3 stencils
#pragma unroll
    for (int r = 0; r < 4; ++r)
#pragma unroll
        for (int c = 0; c < 4; ++c) {
            const int sy = 2 * yo + r - 1, sx = 2 * xo + c - 1;
            xin[r][c] = (sy >= 0 && sy < H && sx >= 0 && sx < W)
                            ? bf2f(x0[((size_t)img * H + sy) * W + sx])
                            : 0.0f;
        }
    uint4 *dst = reinterpret_cast<uint4 *>(a + (size_t)Q * CO);
#pragma unroll
    for (int c8 = 0; c8 < CO / 8; ++c8) {
        float o[8];
#pragma unroll
        for (int e = 0; e < 8; ++e) {
            float m = 0.0f;
#pragma unroll
            for (int q = 0; q < 4; ++q) {
                const int qy = q / 2, qx = q % 2;
                float acc = 0.0f;
#pragma unroll
                for (int t = 0; t < 9; ++t) acc += w[t][c8 * 8 + e] * xin[qy + t / 3][qx + t % 3];
                const float v = bf2f(f2bf(fmaxf(acc + b[c8 * 8 + e], 0.0f)));
                if (q == 0 || v > m) m = v;
            }
            o[e] = m;
        }
        const uint2 lo = pack4(o[0], o[1], o[2], o[3]), hi = pack4(o[4], o[5], o[6], o[7]);
        dst[c8] = make_uint4(lo.x, lo.y, hi.x, hi.y);
    }
}

// ------------------------------------------------------------ MFMA convolution
// POOL (inference; bias, ReLU, bf16 output): the pixel index runs over 2x2
// pooling windows x their 4 pixels (the last row / column of an odd map, which
// no window uses, is skipped), so the four pixels of a window sit in lanes l,
// l^1, l^2, l^3 of the accumulator layout; the epilogue rounds each value to
// bf16 as the unpooled kernel does, takes the first maximum across those lanes
// and stores the pooled map [n][H/2][W/2][Cout].
template <int KS, bool OUT_F32, bool POOL = false>
__global__ __launch_bounds__(256) void k_conv_mfma(const unsigned short *__restrict__ x, int n,
                                                   int H, int W, int Cin,
                                                   const unsigned short *__restrict__ w,
                                                   const float *__restrict__ bias, int bias_stride,
                                                   int Cout, int relu, void *__restrict__ yv) {
    const int lane = threadIdx.x & 63, wave = threadIdx.x >> 6;
    const int li = lane & 15, lk = (lane >> 4) * 4;
    const int Ho = H / 2, Wo = W / 2;
    const long long NP = POOL ? (long long)n * Ho * Wo * 4 : (long long)n * H * W;
    const long long pbase = (long long)blockIdx.x * 256 + wave * 64;
    const int cb0 = blockIdx.y * 64;
    if (pbase >= NP) return;   // uniform over the wavefront; no barriers below
    int pimg[4], py[4], px[4];
    bool pval[4];
#pragma unroll
    for (int pb = 0; pb < 4; ++pb) {
        const long long P = pbase + pb * 16 + li;
        pval[pb] = P < NP;
        const long long Pc = pval[pb] ? P : 0;
        if (POOL) {
            const long long win = Pc >> 2;
            const int sub = (int)(Pc & 3);
            pimg[pb] = (int)(win / ((long long)Ho * Wo));
            const int rem = (int)(win - (long long)pimg[pb] * Ho * Wo);
            py[pb] = 2 * (rem / Wo) + (sub >> 1);
            px[pb] = 2 * (rem % Wo) + (sub & 1);
        } else {
            pimg[pb] = (int)(Pc / ((long long)H * W));
            const int rem = (int)(Pc - (long long)pimg[pb] * H * W);
            py[pb] = rem / W;
            px[pb] = rem % W;
        }
    }
    bool cval[4];
    const unsigned short *wrow[4];
#pragma unroll
    for (int cb = 0; cb < 4; ++cb) {
        const int co = cb0 + cb * 16 + li;
        cval[cb] = cb0 + cb * 16 < Cout;   // uniform: Cout is a multiple of 16
        wrow[cb] = w + (size_t)min(co, Cout - 1) * KS * KS * Cin + lk;
    }
    float4_t acc[4][4];
#pragma unroll
    for (int cb = 0; cb < 4; ++cb)
#pragma unroll
        for (int pb = 0; pb < 4; ++pb) acc[cb][pb] = float4_t{0.0f, 0.0f, 0.0f, 0.0f};
    // k-steps (tap, 16 input channels), software-pipelined: the operands of the
    // next step load while the MFMAs of the current one run.  A tap moves every
    // pixel by the same offset, so the activation address is the pixel's base
    // plus a uniform term; the bounds test is per lane.
    const unsigned short *base[4];
#pragma unroll
    for (int pb = 0; pb < 4; ++pb)
        base[pb] = x + (((size_t)pimg[pb] * H + py[pb]) * W + px[pb]) * Cin + lk;
    const int NK = KS * KS * (Cin / 16);
    int t_n = 0, c_n = 0;   // next step to load
    auto load_step = [&](uint2 (&a)[4], uint2 (&b)[4]) {
        const int dy = t_n / KS - KS / 2, dx = t_n % KS - KS / 2;
        const long long off = ((long long)dy * W + dx) * Cin + c_n;
#pragma unroll
        for (int pb = 0; pb < 4; ++pb) {
            const int sy = py[pb] + dy, sx = px[pb] + dx;
            const bool ok = pval[pb] && sy >= 0 && sy < H && sx >= 0 && sx < W;
            b[pb] = ok ? *reinterpret_cast<const uint2 *>(base[pb] + off) : make_uint2(0u, 0u);
        }
#pragma unroll
        for (int cb = 0; cb < 4; ++cb)
            a[cb] = cval[cb] ? *reinterpret_cast<const uint2 *>(wrow[cb] + (size_t)t_n * Cin + c_n)
                             : make_uint2(0u, 0u);
        c_n += 16;
        if (c_n == Cin) {
            c_n = 0;
            ++t_n;
        }
    };
    auto mma = [&](const uint2 (&a)[4], const uint2 (&b)[4]) {
#pragma unroll
        for (int cb = 0; cb < 4; ++cb) {
            if (!cval[cb]) continue;
#pragma unroll
            for (int pb = 0; pb < 4; ++pb) acc[cb][pb] = mfma_bf16(a[cb], b[pb], acc[cb][pb]);
        }
    };
    uint2 a0[4], b0[4], a1[4], b1[4];
    load_step(a0, b0);
    for (int ks = 0; ks < NK; ks += 2) {
        if (ks + 1 < NK) load_step(a1, b1);
        mma(a0, b0);
        if (ks + 2 < NK) load_step(a0, b0);
        if (ks + 1 < NK) mma(a1, b1);
    }
#pragma unroll
    for (int cb = 0; cb < 4; ++cb) {
        if (!cval[cb]) continue;
        const int co = cb0 + cb * 16 + lk;
        float bb[4] = {0.0f, 0.0f, 0.0f, 0.0f};
        if (bias) {
#pragma unroll
            for (int v = 0; v < 4; ++v) bb[v] = bias[(size_t)(co + v) * bias_stride];
        }
#pragma unroll
        for (int pb = 0; pb < 4; ++pb) {
            const long long P = pbase + pb * 16 + li;
            float o[4];
#pragma unroll
            for (int v = 0; v < 4; ++v) {
                o[v] = acc[cb][pb][v] + bb[v];
                if (relu) o[v] = fmaxf(o[v], 0.0f);
            }
            if (POOL) {
                // window = lanes l..l|3; first maximum in window order (sub 0..3)
#pragma unroll
                for (int v = 0; v < 4; ++v) {
                    float m = bf2f(f2bf(o[v]));
                    const float m1 = __shfl_xor(m, 1, kWave);   // sub ^ 1
                    const float m2 = __shfl_xor(m, 2, kWave);   // sub ^ 2
                    const float m3 = __shfl_xor(m, 3, kWave);   // sub ^ 3
                    // lane with sub 0 holds subs 0,1,2,3 in m, m1, m2, m3
                    float r = m;
                    if (m1 > r) r = m1;
                    if (m2 > r) r = m2;
                    if (m3 > r) r = m3;
                    o[v] = r;
                }
                if (!pval[pb] || (li & 3)) continue;
                unsigned short *y = static_cast<unsigned short *>(yv) + (size_t)(P >> 2) * Cout + co;
                *reinterpret_cast<uint2 *>(y) = pack4(o[0], o[1], o[2], o[3]);
                continue;
            }
            if (!pval[pb]) continue;
            if (OUT_F32) {
                float *y = static_cast<float *>(yv) + (size_t)P * Cout + co;
                *reinterpret_cast<float4 *>(y) = make_float4(o[0], o[1], o[2], o[3]);
            } else {
                unsigned short *y = static_cast<unsigned short *>(yv) + (size_t)P * Cout + co;
                *reinterpret_cast<uint2 *>(y) = pack4(o[0], o[1], o[2], o[3]);
            }
        }
    }
}

// ------------------------------------------------------------ banded pooled convolution
// Inference form of k_conv_mfma<3, false, true> for Cout in {64, 128, 256}:
// workgroup = (image, band of P pooled rows).  The band's input rows
// 2 r0 - 1 .. 2 r1 and columns -1 .. 2 Wo are staged in LDS once (zeros outside
// the image; pixel stride Cin + 8 and row stride RS elements, RS = 64 mod 128,
// which keeps the B reads of a window group nearly free of bank conflicts), so
// the nine taps read it from LDS instead of global memory.  Wavefront w takes
// co block w / (4 / ncb) and every (4 / ncb)-th group of 4 windows, 4 groups at a
// time (64 output channels x 64 pixels, 16 accumulator tiles).  The k-steps,
// their order and each tile's MFMA chain are those of k_conv_mfma, and so is
// the epilogue: the pooled map is bit-identical to it.
constexpr int kBandLdsMax = 40 * 1024;

// G groups of 4 windows (16 pixels each) of one wavefront x 64 output channels.
// G is a template parameter so every MFMA of the chain is unconditional and the
// accumulators stay in their registers across the k-loop.
template <int G>
__device__ __forceinline__ void band_chunk(const unsigned short *__restrict__ band,
                                           const unsigned short *const (&wrow)[4],
                                           const float *__restrict__ bias, int bias_stride, int Cin,
                                           int Cout, int RS, int Cs, int Wo, int nwin, int cb0,
                                           int gfirst, int gstride, size_t win0,
                                           unsigned short *__restrict__ y) {
    const int lane = threadIdx.x & 63, li = lane & 15, lk = (lane >> 4) * 4, sub = li & 3;
    bool pval[G];
    int pix[G];
#pragma unroll
    for (int gi = 0; gi < G; ++gi) {
        const int win = 4 * (gfirst + gi * gstride) + (li >> 2);
        pval[gi] = win < nwin;
        const int wc = pval[gi] ? win : 0;
        const int pr = wc / Wo, pcl = wc - pr * Wo;
        pix[gi] = (2 * pr + (sub >> 1) + 1) * RS + (2 * pcl + (sub & 1) + 1) * Cs + lk;
    }
    float4_t acc[4][G];
#pragma unroll
    for (int ct = 0; ct < 4; ++ct)
#pragma unroll
        for (int gi = 0; gi < G; ++gi) acc[ct][gi] = float4_t{0.0f, 0.0f, 0.0f, 0.0f};
    // k-step s = (tap t, channels c .. c + 15) reads weights at k = 16 s; the
    // weights of step s + 2 load while step s runs (indices clamped at the end,
    // so the loop body has no branch and the accumulators keep their registers)
    const int NK = 9 * (Cin / 16);
    uint2 a0[4], a1[4];
#pragma unroll
    for (int ct = 0; ct < 4; ++ct) {
        a0[ct] = *reinterpret_cast<const uint2 *>(wrow[ct]);
        a1[ct] = *reinterpret_cast<const uint2 *>(wrow[ct] + 16 * min(1, NK - 1));
    }
    int t = 0, c = 0;
    for (int ks = 0; ks < NK; ++ks) {
        uint2 a2[4];
        const int sn = min(ks + 2, NK - 1);
#pragma unroll
        for (int ct = 0; ct < 4; ++ct) a2[ct] = *reinterpret_cast<const uint2 *>(wrow[ct] + 16 * sn);
        const int toff = (t / 3 - 1) * RS + (t % 3 - 1) * Cs + c;
        uint2 b[G];
#pragma unroll
        for (int gi = 0; gi < G; ++gi) b[gi] = *reinterpret_cast<const uint2 *>(band + pix[gi] + toff);
#pragma unroll
        for (int gi = 0; gi < G; ++gi)
#pragma unroll
            for (int ct = 0; ct < 4; ++ct) acc[ct][gi] = mfma_bf16(a0[ct], b[gi], acc[ct][gi]);
        c += 16;
        if (c == Cin) {
            c = 0;
            ++t;
        }
#pragma unroll
        for (int ct = 0; ct < 4; ++ct) {
            a0[ct] = a1[ct];
            a1[ct] = a2[ct];
        }
    }
#pragma unroll
    for (int ct = 0; ct < 4; ++ct) {
        const int co = cb0 + ct * 16 + lk;
        float bb[4] = {0.0f, 0.0f, 0.0f, 0.0f};
        if (bias) {
#pragma unroll
            for (int v = 0; v < 4; ++v) bb[v] = bias[(size_t)(co + v) * bias_stride];
        }
#pragma unroll
        for (int gi = 0; gi < G; ++gi) {
            float o[4];
#pragma unroll
            for (int v = 0; v < 4; ++v) {
                // window = lanes l..l|3; first maximum in window order (sub 0..3)
                const float m = bf2f(f2bf(fmaxf(acc[ct][gi][v] + bb[v], 0.0f)));
                const float m1 = __shfl_xor(m, 1, kWave);
                const float m2 = __shfl_xor(m, 2, kWave);
                const float m3 = __shfl_xor(m, 3, kWave);
                float r = m;
                if (m1 > r) r = m1;
                if (m2 > r) r = m2;
                if (m3 > r) r = m3;
                o[v] = r;
            }
            if (!pval[gi] || sub) continue;
            const int win = 4 * (gfirst + gi * gstride) + (li >> 2);
            *reinterpret_cast<uint2 *>(y + (win0 + win) * Cout + co) = pack4(o[0], o[1], o[2], o[3]);
        }
    }
}

__global__ __launch_bounds__(256) void k_conv_pool_band(const unsigned short *__restrict__ x, int H,
                                                        int W, int Cin,
                                                        const unsigned short *__restrict__ w,
                                                        const float *__restrict__ bias,
                                                        int bias_stride, int Cout, int P, int nband,
                                                        int RS, unsigned short *__restrict__ y) {
    extern __shared__ __align__(16) unsigned short band[];
    const int Ho = H / 2, Wo = W / 2, Wl = 2 * Wo + 2, Cs = Cin + 8, C8 = Cin / 8;
    const int img = blockIdx.x / nband, bnd = blockIdx.x - img * nband;
    const int r0 = bnd * P, r1 = min(Ho, r0 + P), nrow = 2 * (r1 - r0) + 2;
    const int tid = threadIdx.x;
    const unsigned short *src = x + (size_t)img * H * W * Cin;
    for (int i = tid; i < nrow * Wl * C8; i += 256) {
        const int c8 = i % C8, pc = i / C8;
        const int row = pc / Wl, col = pc - row * Wl;
        const int sy = 2 * r0 - 1 + row, sx = col - 1;
        uint4 v = make_uint4(0u, 0u, 0u, 0u);
        if (sy >= 0 && sy < H && sx >= 0 && sx < W)
            v = *reinterpret_cast<const uint4 *>(src + ((size_t)sy * W + sx) * Cin + c8 * 8);
        *reinterpret_cast<uint4 *>(band + row * RS + col * Cs + c8 * 8) = v;
    }
    __syncthreads();

    const int wave = tid >> 6, li = tid & 15;
    const int nwin = (r1 - r0) * Wo, ngrp = (nwin + 3) / 4;
    const int wpc = 4 / (Cout / 64);   // wavefronts per co block
    const int cb0 = (wave / wpc) * 64, h = wave % wpc;
    const int K = 9 * Cin;
    const unsigned short *wrow[4];
#pragma unroll
    for (int ct = 0; ct < 4; ++ct) wrow[ct] = w + (size_t)(cb0 + ct * 16 + li) * K + (tid & 63) / 16 * 4;
    const size_t win0 = ((size_t)img * Ho + r0) * Wo;
    // this wavefront's groups: h, h + wpc, h + 2 wpc, ... taken 4 at a time
    const int mine = h < ngrp ? (ngrp - h + wpc - 1) / wpc : 0;
    for (int c = 0; c < mine; c += 4) {
        const int gfirst = h + c * wpc;
        switch (min(4, mine - c)) {
        case 4:
            band_chunk<4>(band, wrow, bias, bias_stride, Cin, Cout, RS, Cs, Wo, nwin, cb0, gfirst, wpc,
                          win0, y);
            break;
        case 3:
            band_chunk<3>(band, wrow, bias, bias_stride, Cin, Cout, RS, Cs, Wo, nwin, cb0, gfirst, wpc,
                          win0, y);
            break;
        case 2:
            band_chunk<2>(band, wrow, bias, bias_stride, Cin, Cout, RS, Cs, Wo, nwin, cb0, gfirst, wpc,
                          win0, y);
            break;
        default:
            band_chunk<1>(band, wrow, bias, bias_stride, Cin, Cout, RS, Cs, Wo, nwin, cb0, gfirst, wpc,
                          win0, y);
        }
    }
}

// ------------------------------------------------------------ pooled convolution, resident weights
// Inference form for (Cin 32, Cout 64/128/256) and (Cin 64, Cout 32/64/128):
// persistent workgroups; each wavefront loads the weights of its CB blocks of 32
// output channels for all NK = 9 Cin / 16 k-steps into registers once (CB NK 4
// VGPRs = 144) and keeps them for every (image, band) item the workgroup takes.
// Per item the input band is staged in LDS as in k_conv_pool_band, and each
// wavefront runs 32 pixels (8 windows) at a time through the fully unrolled
// k-steps on v_mfma_f32_32x32x16_bf16 (B operand = one 16-byte LDS read):
//   A[i = co][k]   lane l: co = l%32, k = 16 s + 8(l/32) + e      -> w[co][16 s + ..]
//   B[k][j = px]   lane l: px = l%32, k = 16 s + 8(l/32) + e      -> band[px + tap][c + ..]
//   D[i = co][j]   lane l: co = 8(v/4) + 4(l/32) + v%4, px = l%32
// One 32x32x16 step gives the bits of the 16x16x16 step over the same 16 k
// (checked on the device for all four 16x16 quarters), and the k-steps run in
// the order of k_conv_mfma, so the pooled map is bit-identical to it.
typedef __bf16 bf16x8_t __attribute__((ext_vector_type(8)));
typedef float float16_t __attribute__((ext_vector_type(16)));

__device__ __forceinline__ float16_t mfma32_bf16(uint4 a, uint4 b, float16_t c) {
    return __builtin_amdgcn_mfma_f32_32x32x16_bf16(__builtin_bit_cast(bf16x8_t, a),
                                                   __builtin_bit_cast(bf16x8_t, b), c, 0, 0, 0);
}

constexpr int kResidentLdsMax = 63 * 1024;   // dynamic band; with the static bias copy <= 64 KB
constexpr int kResidentCoutMax = 128;

// First maximum (in lane order) of the four lanes of a quad, valid in lane 4k:
// the pairs (4k, 4k+1) and (4k+2, 4k+3) keep their first maximum, then the pair
// of those, through two DPP quad permutations (lane ^ 1, lane ^ 2).  Same pick
// as comparing the four values in order with '>'.
__device__ __forceinline__ float quad_first_max(float m) {
    const float m1 = __int_as_float(__builtin_amdgcn_update_dpp(0, __float_as_int(m), 0xB1, 0xF, 0xF, false));
    const float p = m1 > m ? m1 : m;
    const float p2 = __int_as_float(__builtin_amdgcn_update_dpp(0, __float_as_int(p), 0x4E, 0xF, 0xF, false));
    return p2 > p ? p2 : p;
}

template <int CIN, int CB>
__global__ __launch_bounds__(256, CIN == 32 ? 4 : 2) void k_conv_pool_resident(
    const unsigned short *__restrict__ x, int H, int W, const unsigned short *__restrict__ w,
    const float *__restrict__ bias, int bias_stride, int Cout, int P, int nband, int nitems, int RS,
    unsigned short *__restrict__ y) {
    constexpr int NK = 9 * CIN / 16, CPT = CIN / 16, Cs = CIN + 8, C8 = CIN / 8, K = 9 * CIN;
    extern __shared__ __align__(16) unsigned short band[];
    __shared__ __align__(16) float sbias[kResidentCoutMax];
    const int tid = threadIdx.x, lane = tid & 63, wave = tid >> 6;
    for (int i = tid; i < Cout; i += 256) sbias[i] = bias ? bias[(size_t)i * bias_stride] : 0.0f;
    const int ncg = Cout / (32 * CB), wpg = 4 / ncg;   // co groups, wavefronts per co group
    const int co0 = (wave / wpg) * 32 * CB, wi = wave % wpg;
    uint4 wa[CB][NK];
#pragma unroll
    for (int cb = 0; cb < CB; ++cb)
#pragma unroll
        for (int s = 0; s < NK; ++s)
            wa[cb][s] = *reinterpret_cast<const uint4 *>(
                w + (size_t)(co0 + cb * 32 + (lane & 31)) * K + 16 * s + 8 * (lane >> 5));
    const int Ho = H / 2, Wo = W / 2, Wl = 2 * Wo + 2;
    const int j = lane & 31, sub = j & 3;
    for (int item = blockIdx.x; item < nitems; item += gridDim.x) {
        const int img = item / nband, bnd = item - img * nband;
        const int r0 = bnd * P, r1 = min(Ho, r0 + P), nrow = 2 * (r1 - r0) + 2;
        const unsigned short *src = x + (size_t)img * H * W * CIN;
        __syncthreads();   // the previous item's band is no longer read
        for (int i = tid; i < nrow * Wl * C8; i += 256) {
            const int c8 = i % C8, pc = i / C8;
            const int row = pc / Wl, col = pc - row * Wl;
            const int sy = 2 * r0 - 1 + row, sx = col - 1;
            uint4 v = make_uint4(0u, 0u, 0u, 0u);
            if (sy >= 0 && sy < H && sx >= 0 && sx < W)
                v = *reinterpret_cast<const uint4 *>(src + ((size_t)sy * W + sx) * CIN + c8 * 8);
            *reinterpret_cast<uint4 *>(band + row * RS + col * Cs + c8 * 8) = v;
        }
        __syncthreads();
        const int nwin = (r1 - r0) * Wo, nblk = (nwin + 7) / 8;
        const size_t win0 = ((size_t)img * Ho + r0) * Wo;
        for (int b = wi; b < nblk; b += wpg) {
            const int win = 8 * b + (j >> 2);
            const bool pv = win < nwin;
            const int wc = pv ? win : 0;
            const int pr = wc / Wo, pcl = wc - pr * Wo;
            // top-left tap of the lane's pixel (band row / column of input pixel - 1)
            const unsigned short *pb =
                band + (2 * pr + (sub >> 1)) * RS + (2 * pcl + (sub & 1)) * Cs + 8 * (lane >> 5);
            float16_t acc[CB];
#pragma unroll
            for (int cb = 0; cb < CB; ++cb)
#pragma unroll
                for (int v = 0; v < 16; ++v) acc[cb][v] = 0.0f;
#pragma unroll
            for (int s = 0; s < NK; ++s) {
                const int t = s / CPT, c = (s % CPT) * 16;
                const uint4 bv = *reinterpret_cast<const uint4 *>(pb + (t / 3) * RS + (t % 3) * Cs + c);
#pragma unroll
                for (int cb = 0; cb < CB; ++cb) acc[cb] = mfma32_bf16(wa[cb][s], bv, acc[cb]);
            }
#pragma unroll
            for (int cb = 0; cb < CB; ++cb) {
#pragma unroll
                for (int q = 0; q < 4; ++q) {
                    const int co = co0 + cb * 32 + 8 * q + 4 * (lane >> 5);
                    const float4 bb = *reinterpret_cast<const float4 *>(sbias + co);
                    const float bv[4] = {bb.x, bb.y, bb.z, bb.w};
                    float o[4];
#pragma unroll
                    for (int e = 0; e < 4; ++e)   // window = the lanes of a quad (sub 0..3)
                        o[e] = quad_first_max(bf2f(f2bf(fmaxf(acc[cb][4 * q + e] + bv[e], 0.0f))));
                    if (pv && sub == 0)
                        *reinterpret_cast<uint2 *>(y + (win0 + win) * Cout + co) =
                            pack4(o[0], o[1], o[2], o[3]);
                }
            }
        }
    }
}

// ------------------------------------------------------------ 3x3 convolution, resident weights
// Unpooled 3x3 convolution (pad 1, bias, optional ReLU, bf16 output) for Cin 32 /
// 64 / 128 and Cout 16 / 32 / 64: persistent workgroups; each wavefront keeps
// the A operands of its 16 output channels for all NK = 9 Cin / 16 k-steps in
// registers (NK x 2 VGPRs) and streams 16-pixel blocks (row-major over a band
// of rows, input band with halo staged in LDS as in the pooled forms) through
// the fully unrolled k-steps on v_mfma_f32_16x16x16_bf16, in the order of
// k_conv_mfma: the output is bit-identical to it.
template <int CIN>
__global__ __launch_bounds__(256, 2) void k_conv_resident(
    const unsigned short *__restrict__ x, int H, int W, const unsigned short *__restrict__ w,
    const float *__restrict__ bias, int bias_stride, int Cout, int relu, int TH, int nband,
    int nitems, int RS, unsigned short *__restrict__ y) {
    constexpr int NK = 9 * CIN / 16, CPT = CIN / 16, Cs = CIN + 8, C8 = CIN / 8, K = 9 * CIN;
    extern __shared__ __align__(16) unsigned short band[];
    __shared__ __align__(16) float sbias[kResidentCoutMax];
    const int tid = threadIdx.x, lane = tid & 63, wave = tid >> 6;
    for (int i = tid; i < Cout; i += 256) sbias[i] = bias ? bias[(size_t)i * bias_stride] : 0.0f;
    const int ncb = Cout / 16, wpc = 4 / ncb;   // co blocks of 16, wavefronts per co block
    const int co0 = (wave / wpc) * 16, wi = wave % wpc;
    const int li = lane & 15, lk = (lane >> 4) * 4;
    uint2 wa[NK];
#pragma unroll
    for (int s = 0; s < NK; ++s)
        wa[s] = *reinterpret_cast<const uint2 *>(w + (size_t)(co0 + li) * K + 16 * s + lk);
    const int Wl = W + 2;
    for (int item = blockIdx.x; item < nitems; item += gridDim.x) {
        const int img = item / nband, bnd = item - img * nband;
        const int r0 = bnd * TH, r1 = min(H, r0 + TH), nrow = (r1 - r0) + 2;
        const unsigned short *src = x + (size_t)img * H * W * CIN;
        __syncthreads();   // the previous item's band is no longer read
        for (int i = tid; i < nrow * Wl * C8; i += 256) {
            const int c8 = i % C8, pc = i / C8;
            const int row = pc / Wl, col = pc - row * Wl;
            const int sy = r0 - 1 + row, sx = col - 1;
            uint4 v = make_uint4(0u, 0u, 0u, 0u);
            if (sy >= 0 && sy < H && sx >= 0 && sx < W)
                v = *reinterpret_cast<const uint4 *>(src + ((size_t)sy * W + sx) * CIN + c8 * 8);
            *reinterpret_cast<uint4 *>(band + row * RS + col * Cs + c8 * 8) = v;
        }
        __syncthreads();
        const int npx = (r1 - r0) * W, nblk = (npx + 15) / 16;
        const size_t px0 = ((size_t)img * H + r0) * W;
        for (int b = wi; b < nblk; b += wpc) {
            const int px = 16 * b + li;
            const bool pv = px < npx;
            const int pc = pv ? px : 0;
            const int pr = pc / W, pcl = pc - pr * W;
            // top-left tap of the lane's pixel (band row / column of input pixel - 1)
            const unsigned short *pb = band + pr * RS + pcl * Cs + lk;
            float4_t acc = float4_t{0.0f, 0.0f, 0.0f, 0.0f};
#pragma unroll
            for (int s = 0; s < NK; ++s) {
                const int t = s / CPT, c = (s % CPT) * 16;
                const uint2 bv = *reinterpret_cast<const uint2 *>(pb + (t / 3) * RS + (t % 3) * Cs + c);
                acc = mfma_bf16(wa[s], bv, acc);
            }
            const int co = co0 + lk;
            const float4 bb = *reinterpret_cast<const float4 *>(sbias + co);
            float o[4] = {acc[0] + bb.x, acc[1] + bb.y, acc[2] + bb.z, acc[3] + bb.w};
            if (relu) {
#pragma unroll
                for (int v = 0; v < 4; ++v) o[v] = fmaxf(o[v], 0.0f);
            }
            if (pv) *reinterpret_cast<uint2 *>(y + (px0 + px) * Cout + co) = pack4(o[0], o[1], o[2], o[3]);
        }
    }
}

// ------------------------------------------------------------ pooling
// One thread per (output pixel, 8 channels).
__global__ void k_maxpool2(const unsigned short *__restrict__ y, int n, int H, int W, int C,
                           unsigned short *__restrict__ a) {
    const int Ho = H / 2, Wo = W / 2, C8 = C / 8;
    const long long idx = (long long)blockIdx.x * blockDim.x + threadIdx.x;
    if (idx >= (long long)n * Ho * Wo * C8) return;
    const int c8 = (int)(idx % C8);
    long long r = idx / C8;
    const int xo = (int)(r % Wo);
    r /= Wo;
    const int yo = (int)(r % Ho);
    const int img = (int)(r / Ho);
    float m[8];
    unsigned short h[8];
#pragma unroll
    for (int q = 0; q < 4; ++q) {
        const int sy = 2 * yo + q / 2, sx = 2 * xo + q % 2;
        const uint4 v = *reinterpret_cast<const uint4 *>(y + (((size_t)img * H + sy) * W + sx) * C + c8 * 8);
        const unsigned vv[4] = {v.x, v.y, v.z, v.w};
#pragma unroll
        for (int e = 0; e < 8; ++e) {
            const unsigned short u = (unsigned short)(vv[e / 2] >> (16 * (e & 1)));
            const float f = bf2f(u);
            if (q == 0 || f > m[e]) {
                m[e] = f;
                h[e] = u;
            }
        }
    }
    uint4 o;
    o.x = h[0] | ((unsigned)h[1] << 16);
    o.y = h[2] | ((unsigned)h[3] << 16);
    o.z = h[4] | ((unsigned)h[5] << 16);
    o.w = h[6] | ((unsigned)h[7] << 16);
    *reinterpret_cast<uint4 *>(a + (((size_t)img * Ho + yo) * Wo + xo) * C + c8 * 8) = o;
}

// One thread per (input pixel, 8 channels): the gradient lands on the first
// maximum of its window when that maximum is positive (ReLU before the pool).
__global__ void k_maxpool2_bwd(const unsigned short *__restrict__ y,
                               const unsigned short *__restrict__ da, int n, int H, int W, int C,
                               unsigned short *__restrict__ dz) {
    const int Ho = H / 2, Wo = W / 2, C8 = C / 8;
    const long long idx = (long long)blockIdx.x * blockDim.x + threadIdx.x;
    if (idx >= (long long)n * H * W * C8) return;
    const int c8 = (int)(idx % C8);
    long long r = idx / C8;
    const int xi = (int)(r % W);
    r /= W;
    const int yi = (int)(r % H);
    const int img = (int)(r / H);
    const int yo = yi / 2, xo = xi / 2;
    unsigned out[4] = {0u, 0u, 0u, 0u};
    if (yo < Ho && xo < Wo) {
        const int me = (yi & 1) * 2 + (xi & 1);
        float m[8];
        int arg[8];
        float mine[8];
#pragma unroll
        for (int q = 0; q < 4; ++q) {
            const int sy = 2 * yo + q / 2, sx = 2 * xo + q % 2;
            const uint4 v = *reinterpret_cast<const uint4 *>(y + (((size_t)img * H + sy) * W + sx) * C + c8 * 8);
            const unsigned vv[4] = {v.x, v.y, v.z, v.w};
#pragma unroll
            for (int e = 0; e < 8; ++e) {
                const float f = bf2f((unsigned short)(vv[e / 2] >> (16 * (e & 1))));
                if (q == 0 || f > m[e]) {
                    m[e] = f;
                    arg[e] = q;
                }
                if (q == me) mine[e] = f;
            }
        }
        const uint4 g = *reinterpret_cast<const uint4 *>(da + (((size_t)img * Ho + yo) * Wo + xo) * C + c8 * 8);
        const unsigned gg[4] = {g.x, g.y, g.z, g.w};
#pragma unroll
        for (int e = 0; e < 8; ++e) {
            const unsigned short u = (unsigned short)(gg[e / 2] >> (16 * (e & 1)));
            const unsigned short o = (arg[e] == me && mine[e] > 0.0f) ? u : (unsigned short)0;
            out[e / 2] |= (unsigned)o << (16 * (e & 1));
        }
    }
    *reinterpret_cast<uint4 *>(dz + (((size_t)img * H + yi) * W + xi) * C + c8 * 8) =
        make_uint4(out[0], out[1], out[2], out[3]);
}

__global__ void k_relu_bwd(const unsigned short *__restrict__ y, const unsigned short *__restrict__ dy,
                           size_t count, unsigned short *__restrict__ dz) {
    const size_t i = (size_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= count) return;
    dz[i] = bf2f(y[i]) > 0.0f ? dy[i] : (unsigned short)0;
}

// ------------------------------------------------------------ weight gradient
// dW[co][k] = sum over pixels P of dz[P][co] x_col[P][k], k = tap Cin + ci (k == K
// is the bias column, input 1), on the matrix cores with the pixels as the
// contraction.  Tile: 64 output channels x 64 columns k per workgroup; wavefront
// (wm, wn) owns the 32 x 32 quadrant (co 32 wm.., k 32 wn..) on
// v_mfma_f32_32x32x16_bf16.  Per stage of 64 pixels both operands are staged in
// LDS transposed to [column][pixel], two pixels per dword, so a lane's 8
// consecutive pixels are one 16-byte read (A: co = l % 32, pixels 16 s + 8 (l / 32)
// ..; B: k = l % 32, the same pixels).  Products of bf16 values are exact in FP32.
// Blocks along z take pixel ranges [z chunk, (z + 1) chunk) and write partial
// sums part[z][Cout][K + 1], added in a fixed order by k_sum_parts.
constexpr int kWgPix = 64;              // pixels per stage
constexpr int kWgRow = kWgPix + 8;      // bf16 per LDS row: rows 8 apart fall 32 banks apart

__global__ __launch_bounds__(256) void k_wgrad(const unsigned short *__restrict__ x,
                                               const unsigned short *__restrict__ dz, int n, int H,
                                               int W, int Cin, int Cout, int KS, long long chunk,
                                               float *__restrict__ part) {
    __shared__ __align__(16) unsigned short sd[64 * kWgRow];
    __shared__ __align__(16) unsigned short sx[64 * kWgRow];
    const int tid = threadIdx.x, lane = tid & 63, wave = tid >> 6;
    const int K = KS * KS * Cin, K1 = K + 1;
    const int k0 = blockIdx.x * 64, c0 = blockIdx.y * 64;
    const long long NP = (long long)n * H * W;
    const long long p0 = (long long)blockIdx.z * chunk;
    const long long p1 = min(NP, p0 + chunk);
    const int wm = wave >> 1, wn = wave & 1;
    // staging: pixel pair pp (stage pixels 2 pp, 2 pp + 1), column group cg (8 columns)
    const int pp = tid & 31, cg = tid >> 5;
    const int cb = c0 + 8 * cg, kb = k0 + 8 * cg;
    const bool dvec = (Cout % 8 == 0) && cb + 8 <= Cout;
    const bool xvec = (Cin % 8 == 0) && kb + 8 <= K;   // 8 channels of one tap
    float16_t acc = {};
    for (long long pc = p0; pc < p1; pc += kWgPix) {
        unsigned dw[8], xw[8];   // (pixel 2 pp, pixel 2 pp + 1) per column
#pragma unroll
        for (int e = 0; e < 8; ++e) dw[e] = xw[e] = 0u;
#pragma unroll
        for (int h = 0; h < 2; ++h) {
            const long long P = pc + 2 * pp + h;
            if (P >= p1) continue;
            unsigned short d8[8], x8[8];
            if (dvec) {
                const uint4 v = *reinterpret_cast<const uint4 *>(dz + (size_t)P * Cout + cb);
                const unsigned vv[4] = {v.x, v.y, v.z, v.w};
#pragma unroll
                for (int e = 0; e < 8; ++e) d8[e] = (unsigned short)(vv[e / 2] >> (16 * (e & 1)));
            } else {
#pragma unroll
                for (int e = 0; e < 8; ++e)
                    d8[e] = (cb + e < Cout) ? dz[(size_t)P * Cout + cb + e] : (unsigned short)0;
            }
            const int img = (int)(P / ((long long)H * W));
            const int rem = (int)(P - (long long)img * H * W);
            const int py = rem / W, px = rem - py * W;
            if (xvec) {
                const int t = kb / Cin, ci = kb - t * Cin;
                const int sy = py + t / KS - KS / 2, sx2 = px + t % KS - KS / 2;
                uint4 v = make_uint4(0u, 0u, 0u, 0u);
                if (sy >= 0 && sy < H && sx2 >= 0 && sx2 < W)
                    v = *reinterpret_cast<const uint4 *>(x + (((size_t)img * H + sy) * W + sx2) * Cin + ci);
                const unsigned vv[4] = {v.x, v.y, v.z, v.w};
#pragma unroll
                for (int e = 0; e < 8; ++e) x8[e] = (unsigned short)(vv[e / 2] >> (16 * (e & 1)));
            } else {
#pragma unroll
                for (int e = 0; e < 8; ++e) {
                    const int k = kb + e;
                    unsigned short xv = 0;
                    if (k == K) {
                        xv = 0x3f80;   // bf16 1.0: the bias column
                    } else if (k < K) {
                        const int t = k / Cin, ci = k - t * Cin;
                        const int sy = py + t / KS - KS / 2, sx2 = px + t % KS - KS / 2;
                        if (sy >= 0 && sy < H && sx2 >= 0 && sx2 < W)
                            xv = x[(((size_t)img * H + sy) * W + sx2) * Cin + ci];
                    }
                    x8[e] = xv;
                }
            }
#pragma unroll
            for (int e = 0; e < 8; ++e) {
                dw[e] |= (unsigned)d8[e] << (16 * h);
                xw[e] |= (unsigned)x8[e] << (16 * h);
            }
        }
        __syncthreads();   // the previous stage's operand reads are done
#pragma unroll
        for (int e = 0; e < 8; ++e) {
            *reinterpret_cast<unsigned *>(sd + (8 * cg + e) * kWgRow + 2 * pp) = dw[e];
            *reinterpret_cast<unsigned *>(sx + (8 * cg + e) * kWgRow + 2 * pp) = xw[e];
        }
        __syncthreads();
#pragma unroll
        for (int s = 0; s < kWgPix / 16; ++s) {
            const int off = 16 * s + 8 * (lane >> 5);
            const uint4 a = *reinterpret_cast<const uint4 *>(sd + (32 * wm + (lane & 31)) * kWgRow + off);
            const uint4 b = *reinterpret_cast<const uint4 *>(sx + (32 * wn + (lane & 31)) * kWgRow + off);
            acc = mfma32_bf16(a, b, acc);
        }
    }
    // D: co = 8 (v / 4) + 4 (l / 32) + v % 4 of the quadrant, k = l % 32
    float *out = part + (size_t)blockIdx.z * Cout * K1;
    const int k = k0 + 32 * wn + (lane & 31);
#pragma unroll
    for (int v = 0; v < 16; ++v) {
        const int co = c0 + 32 * wm + 8 * (v / 4) + 4 * (lane >> 5) + (v & 3);
        if (co < Cout && k < K1) out[(size_t)co * K1 + k] = acc[v];
    }
}

__global__ void k_sum_parts(const float *__restrict__ part, int nparts, size_t count,
                            float *__restrict__ g) {
    const size_t i = (size_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= count) return;
    float s = 0.0f;
    for (int z = 0; z < nparts; ++z) s += part[(size_t)z * count + i];
    g[i] = s;
}

// ------------------------------------------------------------ loss
// One wavefront per particle.
__device__ __forceinline__ float wave_max(float v) {
#pragma unroll
    for (int off = 32; off >= 1; off >>= 1) v = fmaxf(v, __shfl_xor(v, off, kWave));
    return v;
}
__device__ __forceinline__ float wave_sum(float v) {
#pragma unroll
    for (int off = 32; off >= 1; off >>= 1) v += __shfl_xor(v, off, kWave);
    return v;
}

// Soft target of class c of head h given the label: psi classes (mirror-major,
// n_psi bins) get a circular Gaussian over bins of the labelled mirror, shift
// classes (x-major grid, ny per row) a 2-D Gaussian over grid steps; sigma 0
// gives the one-hot target.  Unnormalised.
__device__ __forceinline__ float soft_target(int h, int c, int lab, int h1, float sp, float ss,
                                             int ny) {
    if (h == 1 && sp > 0.0f) {
        const int np = h1 / 2;
        if (c / np != lab / np) return 0.0f;
        int d = abs(c % np - lab % np);
        d = min(d, np - d);
        return expf(-0.5f * (float)(d * d) / (sp * sp));
    }
    if (h == 2 && ss > 0.0f && ny > 0) {
        const float dx = (float)(c / ny - lab / ny), dy = (float)(c % ny - lab % ny);
        return expf(-0.5f * (dx * dx + dy * dy) / (ss * ss));
    }
    return c == lab ? 1.0f : 0.0f;
}

__global__ __launch_bounds__(256) void k_softmax_xent(const float *__restrict__ logits, int n, int ld,
                                                      int h0, int h1, int h2,
                                                      const int *__restrict__ labels, float sp,
                                                      float ss, int ny, float *__restrict__ loss,
                                                      unsigned short *__restrict__ dlogits) {
    const int lane = threadIdx.x & 63;
    const int i = blockIdx.x * 4 + (threadIdx.x >> 6);
    if (i >= n) return;
    const float *z = logits + (size_t)i * ld;
    unsigned short *dz = dlogits + (size_t)i * ld;
    const int off[3] = {0, h0, h0 + h1}, sz[3] = {h0, h1, h2};
    const float inv_n = 1.0f / (float)n;
    float total = 0.0f;
    for (int h = 0; h < 3; ++h) {
        const int lab = min(max(labels[(size_t)i * 3 + h], 0), sz[h] - 1);
        float m = -FLT_MAX, tsum = 0.0f;
        for (int c = lane; c < sz[h]; c += 64) {
            m = fmaxf(m, z[off[h] + c]);
            tsum += soft_target(h, c, lab, h1, sp, ss, ny);
        }
        m = wave_max(m);
        tsum = wave_sum(tsum);
        float se = 0.0f;
        for (int c = lane; c < sz[h]; c += 64) se += expf(z[off[h] + c] - m);
        se = wave_sum(se);
        const float lse = m + logf(se);
        float l = 0.0f;
        for (int c = lane; c < sz[h]; c += 64) {
            const float t = soft_target(h, c, lab, h1, sp, ss, ny) / tsum;
            const float pr = expf(z[off[h] + c] - lse);
            l += t * (lse - z[off[h] + c]);
            dz[off[h] + c] = f2bf((pr - t) * inv_n);
        }
        total += wave_sum(l);
    }
    for (int c = h0 + h1 + h2 + lane; c < ld; c += 64) dz[c] = 0;
    if (lane == 0) loss[i] = total;
}

// ------------------------------------------------------------ optimiser
__global__ void k_adamw(float *__restrict__ p, const float *__restrict__ g, float *__restrict__ m,
                        float *__restrict__ v, const unsigned char *__restrict__ mask, size_t count,
                        float lr, float b1, float b2, float eps, float wd, float c1, float c2) {
    const size_t i = (size_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= count) return;
    if (mask && !mask[i]) return;
    const float gi = g[i];
    const float mi = b1 * m[i] + (1.0f - b1) * gi;
    const float vi = b2 * v[i] + (1.0f - b2) * gi * gi;
    m[i] = mi;
    v[i] = vi;
    const float mh = mi / c1, vh = vi / c2;
    p[i] -= lr * (mh / (sqrtf(vh) + eps) + wd * p[i]);
}

__global__ void k_pack_weights(const float *__restrict__ master, int Cout, int Cin, int KS,
                               unsigned short *__restrict__ w, unsigned short *__restrict__ wflip) {
    const int K = KS * KS * Cin;
    const size_t i = (size_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= (size_t)Cout * K) return;
    const int co = (int)(i / K), k = (int)(i % K);
    const unsigned short h = f2bf(master[(size_t)co * (K + 1) + k]);
    w[i] = h;
    if (wflip) {
        const int t = k / Cin, ci = k % Cin;
        const int tf = KS * KS - 1 - t;
        wflip[((size_t)ci * KS * KS + tf) * Cout + co] = h;
    }
}

// ------------------------------------------------------------ proposals
// Wave top-KT of one head: values v[j] at index lane + 64 j (j < J), -inf padded.
// Lane k (k < KT) receives the k-th best value and index (value descending,
// index ascending); index -1 when the head has fewer than k+1 entries.
template <int J, int KT>
__device__ __forceinline__ void head_top(float (&v)[J], int lane, float &top_v, int &top_i) {
    top_v = -FLT_MAX;
    top_i = -1;
    for (int r = 0; r < KT; ++r) {
        float bv = -FLT_MAX;
        int bi = INT_MAX;
#pragma unroll
        for (int j = 0; j < J; ++j) {
            const int idx = lane + 64 * j;
            if (v[j] > bv || (v[j] == bv && idx < bi)) {
                bv = v[j];
                bi = idx;
            }
        }
#pragma unroll
        for (int off = 32; off >= 1; off >>= 1) {
            const float ov = __shfl_xor(bv, off, kWave);
            const int oi = __shfl_xor(bi, off, kWave);
            if (ov > bv || (ov == bv && oi < bi)) {
                bv = ov;
                bi = oi;
            }
        }
        if (lane == r && bi != INT_MAX && bv > -FLT_MAX) {
            top_v = bv;
            top_i = bi;
        }
#pragma unroll
        for (int j = 0; j < J; ++j)
            if (lane + 64 * j == bi) v[j] = -FLT_MAX;
    }
}

// One wavefront per particle.  Candidate triples: top-8 references x top-16 psi
// classes x top-16 shifts (heads up to 512 / 256 / 64 entries); score = sum of
// the three logits.  Greedy selection of k triples, best first (ties to the
// lower (ref rank, psi rank, shift rank)); with psi_r >= 0 every selection
// suppresses the candidates of the same reference and mirror within psi_r psi
// bins (circular) and shift_r grid steps in x and y (grid rows of ny shifts).
__global__ __launch_bounds__(256) void k_proposals(const float *__restrict__ logits, int n, int ld,
                                                   int n_ref, int n_psi, int n_shift, int k,
                                                   int psi_r, int shift_r, int ny, int min_refs,
                                                   int4 *__restrict__ prop, float *__restrict__ conf) {
    constexpr int KR = 8, KP = 16, KS = 16, PER = KR * KP * KS / 64;   // 32 per lane
    const int lane = threadIdx.x & 63;
    const int i = blockIdx.x * 4 + (threadIdx.x >> 6);
    if (i >= n) return;
    const float *z = logits + (size_t)i * ld;
    float rv[8], pv[4], sv[1];
#pragma unroll
    for (int j = 0; j < 8; ++j) rv[j] = lane + 64 * j < n_ref ? z[lane + 64 * j] : -FLT_MAX;
#pragma unroll
    for (int j = 0; j < 4; ++j) pv[j] = lane + 64 * j < 2 * n_psi ? z[n_ref + lane + 64 * j] : -FLT_MAX;
    sv[0] = lane < n_shift ? z[n_ref + 2 * n_psi + lane] : -FLT_MAX;
    // softmax of the reference head (for the coverage estimate)
    float mx = -FLT_MAX;
#pragma unroll
    for (int j = 0; j < 8; ++j) mx = fmaxf(mx, rv[j]);
    mx = wave_max(mx);
    float se = 0.0f;
#pragma unroll
    for (int j = 0; j < 8; ++j) se += lane + 64 * j < n_ref ? expf(rv[j] - mx) : 0.0f;
    se = wave_sum(se);
    float covered = 0.0f;   // lane 0: softmax mass of the distinct proposed references
    int seen[8];
    int n_seen = 0;
    float tr, tp, ts;
    int ir, ip, is;
    head_top<8, KR>(rv, lane, tr, ir);
    head_top<4, KP>(pv, lane, tp, ip);
    head_top<1, KS>(sv, lane, ts, is);
    // lane l holds candidates c = l * PER + j: ref rank c / 256, psi rank (c / 16) % 16,
    // shift rank c % 16 (so c orders candidates lexicographically by rank)
    float sc[PER];
    int cr[PER], cp[PER], cs[PER];
#pragma unroll
    for (int j = 0; j < PER; ++j) {
        const int c = lane * PER + j;
        const int a = c / (KP * KS), b = (c / KS) % KP, d = c % KS;
        cr[j] = __shfl(ir, a, kWave);
        cp[j] = __shfl(ip, b, kWave);
        cs[j] = __shfl(is, d, kWave);
        const float v = __shfl(tr, a, kWave) + __shfl(tp, b, kWave) + __shfl(ts, d, kWave);
        sc[j] = (cr[j] >= 0 && cp[j] >= 0 && cs[j] >= 0) ? v : -INFINITY;
    }
    for (int r = 0; r < k; ++r) {
        // the first min_refs picks take the best remaining triple of reference
        // rank r (its candidates live in lanes 8r .. 8r+7); otherwise, or when
        // that reference has none left, the best remaining triple overall
        int bi = INT_MAX;
        for (int pass = (r < min_refs && r < KR) ? 0 : 1; pass < 2 && bi == INT_MAX; ++pass) {
            float bv = -FLT_MAX;
            if (pass == 1 || (lane >> 3) == r) {
#pragma unroll
                for (int j = 0; j < PER; ++j) {
                    const int c = lane * PER + j;
                    if (sc[j] > bv || (sc[j] == bv && c < bi)) {
                        bv = sc[j];
                        bi = c;
                    }
                }
            }
#pragma unroll
            for (int off = 32; off >= 1; off >>= 1) {
                const float ov = __shfl_xor(bv, off, kWave);
                const int oi = __shfl_xor(bi, off, kWave);
                if (ov > bv || (ov == bv && oi < bi)) {
                    bv = ov;
                    bi = oi;
                }
            }
        }
        if (bi == INT_MAX) {   // fewer valid triples than k
            if (lane == 0) prop[(size_t)i * k + r] = make_int4(-1, 0, 0, 0);
            continue;
        }
        const int sa = bi / (KP * KS), sb = (bi / KS) % KP, sd = bi % KS;
        const int ref = __shfl(ir, sa, kWave), pc = __shfl(ip, sb, kWave), sh = __shfl(is, sd, kWave);
        if (lane == 0) {
            prop[(size_t)i * k + r] = make_int4(ref, pc / n_psi, pc % n_psi, sh);
            bool dup = false;
            for (int q = 0; q < n_seen; ++q) dup = dup || seen[q] == ref;
            if (!dup && n_seen < 8) {
                seen[n_seen++] = ref;
                covered += expf(z[ref] - mx);
            }
        }
#pragma unroll
        for (int j = 0; j < PER; ++j) {
            const int c = lane * PER + j;
            bool sup = c == bi;
            if (psi_r >= 0 && cr[j] == ref && cp[j] / n_psi == pc / n_psi) {
                int db = abs(cp[j] % n_psi - pc % n_psi);
                db = min(db, n_psi - db);
                const bool near_s = abs(cs[j] / ny - sh / ny) <= shift_r && abs(cs[j] % ny - sh % ny) <= shift_r;
                sup = sup || (db <= psi_r && near_s);
            }
            if (sup) sc[j] = -INFINITY;
        }
    }
    if (conf && lane == 0) conf[i] = se > 0.0f ? covered / se : 0.0f;
}

int fail(int code, const char *msg) {
    report_error(code, msg);
    return code;
}

int launched(const char *what) {
    const hipError_t e = hipGetLastError();
    if (e == hipSuccess) return 0;
    char buf[256];
    snprintf(buf, sizeof(buf), "%s launch failed: %s", what, hipGetErrorString(e));
    report_error(1, buf);
    return 1;
}

unsigned blocks_for(size_t count, int threads) { return (unsigned)((count + threads - 1) / threads); }

}  // namespace
}  // namespace isac

using namespace isac;

extern "C" {

int isac_nn_standardize(const float *img, int n, int D, uint16_t *x0, void *stream) {
    if (n < 0 || D < 1 || (n > 0 && (!img || !x0))) return fail(2, "isac_nn_standardize: bad arguments");
    if (n == 0) return 0;
    hipLaunchKernelGGL(k_standardize, dim3(n), dim3(256), 0, (hipStream_t)stream, img, D, x0);
    return launched("isac_nn_standardize");
}

int isac_nn_conv1_fwd(const uint16_t *x0, int n, int H, int W, const float *master, int Cout,
                      uint16_t *y, void *stream) {
    if (Cout != 32 || n < 0 || H < 1 || W < 1 || (n > 0 && (!x0 || !master || !y)))
        return fail(2, "isac_nn_conv1_fwd: bad arguments (Cout must be 32)");
    const size_t np = (size_t)n * H * W;
    if (np == 0) return 0;
    hipLaunchKernelGGL(k_conv1, dim3(blocks_for(np, 256)), dim3(256), 0, (hipStream_t)stream, x0, n, H,
                       W, master, y);
    return launched("isac_nn_conv1_fwd");
}

int isac_nn_conv_fwd(const uint16_t *x, int n, int H, int W, int Cin, const uint16_t *w,
                     const float *bias, int bias_stride, int Cout, int ks, int relu, int out_f32,
                     void *y, void *stream) {
    if ((ks != 1 && ks != 3) || Cin < 16 || Cin % 16 || Cout < 16 || Cout % 16 || n < 0 || H < 1 ||
        W < 1 || (n > 0 && (!x || !w || !y)))
        return fail(2, "isac_nn_conv_fwd: bad arguments (Cin, Cout multiples of 16; ks 1 or 3)");
    const size_t np = (size_t)n * H * W;
    if (np == 0) return 0;
    hipStream_t st = (hipStream_t)stream;
    if (ks == 3 && !out_f32 && (Cin == 32 || Cin == 64 || Cin == 128) &&
        (Cout == 16 || Cout == 32 || Cout == 64)) {
        // resident-weight form: band height with the fewest wavefront rounds of
        // 16 pixels, plus half a round per band for its staging
        const int Cs = Cin + 8, Wl = W + 2;
        const int RS = Wl * Cs + ((64 - Wl * Cs % 128) + 128) % 128;
        const int rmax = kResidentLdsMax / (RS * 2) - 2;
        if (rmax >= 1) {
            const int wpc = 4 / (Cout / 16);
            int TH = 1;
            double best = 1e30;
            for (int t = 1; t <= std::min(rmax, H); ++t) {
                double cost = 0.0;
                for (int r0 = 0; r0 < H; r0 += t) {
                    const int nblk = (std::min(t, H - r0) * W + 15) / 16;
                    cost += (nblk + wpc - 1) / wpc + 0.5;
                }
                if (cost <= best) {
                    best = cost;
                    TH = t;
                }
            }
            const int nband = (H + TH - 1) / TH;
            const size_t lds = (size_t)(TH + 2) * RS * 2;
            const long long items = (long long)n * nband;
            if (items > INT_MAX) return fail(2, "isac_nn_conv_fwd: too many images");
            const void *fn = Cin == 32   ? reinterpret_cast<const void *>(&k_conv_resident<32>)
                             : Cin == 64 ? reinterpret_cast<const void *>(&k_conv_resident<64>)
                                         : reinterpret_cast<const void *>(&k_conv_resident<128>);
            int dev = 0, cus = 0, per_cu = 0;
            if (hipGetDevice(&dev) != hipSuccess ||
                hipDeviceGetAttribute(&cus, hipDeviceAttributeMultiprocessorCount, dev) != hipSuccess ||
                hipOccupancyMaxActiveBlocksPerMultiprocessor(&per_cu, fn, 256, lds) != hipSuccess ||
                per_cu < 1)
                return fail(3, "isac_nn_conv_fwd: occupancy query failed");
            const unsigned grid = (unsigned)std::min<long long>(items, (long long)cus * per_cu);
            const unsigned short *xw = x;
            unsigned short *yw = static_cast<unsigned short *>(y);
            if (Cin == 32)
                hipLaunchKernelGGL(k_conv_resident<32>, dim3(grid), dim3(256), lds, st, xw, H, W, w, bias,
                                   bias_stride, Cout, relu, TH, nband, (int)items, RS, yw);
            else if (Cin == 64)
                hipLaunchKernelGGL(k_conv_resident<64>, dim3(grid), dim3(256), lds, st, xw, H, W, w, bias,
                                   bias_stride, Cout, relu, TH, nband, (int)items, RS, yw);
            else
                hipLaunchKernelGGL(k_conv_resident<128>, dim3(grid), dim3(256), lds, st, xw, H, W, w, bias,
                                   bias_stride, Cout, relu, TH, nband, (int)items, RS, yw);
            return launched("isac_nn_conv_fwd");
        }
    }
    dim3 grid(blocks_for(np, 256), (unsigned)((Cout + 63) / 64));
    if (ks == 3 && !out_f32)
        hipLaunchKernelGGL((k_conv_mfma<3, false>), grid, dim3(256), 0, st, x, n, H, W, Cin, w, bias,
                           bias_stride, Cout, relu, y);
    else if (ks == 3)
        hipLaunchKernelGGL((k_conv_mfma<3, true>), grid, dim3(256), 0, st, x, n, H, W, Cin, w, bias,
                           bias_stride, Cout, relu, y);
    else if (!out_f32)
        hipLaunchKernelGGL((k_conv_mfma<1, false>), grid, dim3(256), 0, st, x, n, H, W, Cin, w, bias,
                           bias_stride, Cout, relu, y);
    else
        hipLaunchKernelGGL((k_conv_mfma<1, true>), grid, dim3(256), 0, st, x, n, H, W, Cin, w, bias,
                           bias_stride, Cout, relu, y);
    return launched("isac_nn_conv_fwd");
}

int isac_nn_conv1_pool_fwd(const uint16_t *x0, int n, int H, int W, const float *master, int Cout,
                           uint16_t *a, void *stream) {
    if (Cout != 32 || n < 0 || H < 2 || W < 2 || (n > 0 && (!x0 || !master || !a)))
        return fail(2, "isac_nn_conv1_pool_fwd: bad arguments (Cout must be 32)");
    const size_t nq = (size_t)n * (H / 2) * (W / 2);
    if (nq == 0) return 0;
    hipLaunchKernelGGL(k_conv1_pool, dim3(blocks_for(nq, 256)), dim3(256), 0, (hipStream_t)stream, x0,
                       n, H, W, master, a);
    return launched("isac_nn_conv1_pool_fwd");
}

int isac_nn_conv_pool_fwd(const uint16_t *x, int n, int H, int W, int Cin, const uint16_t *w,
                          const float *bias, int bias_stride, int Cout, uint16_t *a, void *stream) {
    if (Cin < 16 || Cin % 16 || Cout < 16 || Cout % 16 || n < 0 || H < 2 || W < 2 ||
        (n > 0 && (!x || !w || !a)))
        return fail(2, "isac_nn_conv_pool_fwd: bad arguments (Cin, Cout multiples of 16)");
    const size_t np = (size_t)n * (H / 2) * (W / 2) * 4;
    if (np == 0) return 0;
    // banded form: largest band of pooled rows whose input rows fit kBandLdsMax,
    // then the bands evened out over the pooled rows
    const int Ho = H / 2, Cs = Cin + 8;
    const int RS = (2 * (W / 2) + 2) * Cs + ((64 - (2 * (W / 2) + 2) * Cs % 128) + 128) % 128;
    const int cb = 1;   // co blocks of 32 per wavefront in the resident form
    const int ncg = Cout % (32 * cb) ? 0 : Cout / (32 * cb);
    const int pres = (kResidentLdsMax / (RS * 2) - 2) / 2;
    if ((Cin == 32 || Cin == 64) && (ncg == 1 || ncg == 2 || ncg == 4) && Cout <= kResidentCoutMax &&
        pres >= 1) {
        // band height: fewest wavefront rounds of 8 windows over the image, plus
        // half a round per band for its staging
        const int wpg = 4 / ncg, Wo = W / 2;
        int P = 1;
        double best = 1e30;
        for (int p = 1; p <= std::min(pres, Ho); ++p) {
            double cost = 0.0;
            for (int r0 = 0; r0 < Ho; r0 += p) {
                const int nblk = (std::min(p, Ho - r0) * Wo + 7) / 8;
                cost += (nblk + wpg - 1) / wpg + 0.5;
            }
            if (cost <= best) {
                best = cost;
                P = p;
            }
        }
        const int nband = (Ho + P - 1) / P;
        const size_t lds = (size_t)(2 * P + 2) * RS * 2;
        const long long items = (long long)n * nband;
        if (items > INT_MAX) return fail(2, "isac_nn_conv_pool_fwd: too many images");
        const void *fn = Cin == 32 ? reinterpret_cast<const void *>(&k_conv_pool_resident<32, 1>)
                                   : reinterpret_cast<const void *>(&k_conv_pool_resident<64, 1>);
        int dev = 0, cus = 0, per_cu = 0;
        if (hipGetDevice(&dev) != hipSuccess ||
            hipDeviceGetAttribute(&cus, hipDeviceAttributeMultiprocessorCount, dev) != hipSuccess ||
            hipOccupancyMaxActiveBlocksPerMultiprocessor(&per_cu, fn, 256, lds) != hipSuccess ||
            per_cu < 1)
            return fail(3, "isac_nn_conv_pool_fwd: occupancy query failed");
        const unsigned grid = (unsigned)std::min<long long>(items, (long long)cus * per_cu);
        if (Cin == 32)
            hipLaunchKernelGGL((k_conv_pool_resident<32, 1>), dim3(grid), dim3(256), lds,
                               (hipStream_t)stream, x, H, W, w, bias, bias_stride, Cout, P, nband,
                               (int)items, RS, a);
        else
            hipLaunchKernelGGL((k_conv_pool_resident<64, 1>), dim3(grid), dim3(256), lds,
                               (hipStream_t)stream, x, H, W, w, bias, bias_stride, Cout, P, nband,
                               (int)items, RS, a);
        return launched("isac_nn_conv_pool_fwd");
    }
    const int pmax = (kBandLdsMax / (RS * 2) - 2) / 2;
    if ((Cout == 64 || Cout == 128 || Cout == 256) && pmax >= 1) {
        const int nb0 = (Ho + pmax - 1) / pmax, P = (Ho + nb0 - 1) / nb0, nband = (Ho + P - 1) / P;
        const size_t lds = (size_t)(2 * P + 2) * RS * 2;
        hipLaunchKernelGGL(k_conv_pool_band, dim3((unsigned)((size_t)n * nband)), dim3(256), lds,
                           (hipStream_t)stream, x, H, W, Cin, w, bias, bias_stride, Cout, P, nband, RS,
                           a);
        return launched("isac_nn_conv_pool_fwd");
    }
    dim3 grid(blocks_for(np, 256), (unsigned)((Cout + 63) / 64));
    hipLaunchKernelGGL((k_conv_mfma<3, false, true>), grid, dim3(256), 0, (hipStream_t)stream, x, n, H, W,
                       Cin, w, bias, bias_stride, Cout, 1, a);
    return launched("isac_nn_conv_pool_fwd");
}

int isac_nn_maxpool2(const uint16_t *y, int n, int H, int W, int C, uint16_t *a, void *stream) {
    if (C % 8 || C < 8 || n < 0 || H < 2 || W < 2 || (n > 0 && (!y || !a)))
        return fail(2, "isac_nn_maxpool2: bad arguments");
    const size_t cnt = (size_t)n * (H / 2) * (W / 2) * (C / 8);
    if (cnt == 0) return 0;
    hipLaunchKernelGGL(k_maxpool2, dim3(blocks_for(cnt, 256)), dim3(256), 0, (hipStream_t)stream, y, n,
                       H, W, C, a);
    return launched("isac_nn_maxpool2");
}

int isac_nn_maxpool2_bwd(const uint16_t *y, const uint16_t *da, int n, int H, int W, int C,
                         uint16_t *dz, void *stream) {
    if (C % 8 || C < 8 || n < 0 || H < 2 || W < 2 || (n > 0 && (!y || !da || !dz)))
        return fail(2, "isac_nn_maxpool2_bwd: bad arguments");
    const size_t cnt = (size_t)n * H * W * (C / 8);
    if (cnt == 0) return 0;
    hipLaunchKernelGGL(k_maxpool2_bwd, dim3(blocks_for(cnt, 256)), dim3(256), 0, (hipStream_t)stream, y,
                       da, n, H, W, C, dz);
    return launched("isac_nn_maxpool2_bwd");
}

int isac_nn_relu_bwd(const uint16_t *y, const uint16_t *dy, size_t count, uint16_t *dz, void *stream) {
    if (count && (!y || !dy || !dz)) return fail(2, "isac_nn_relu_bwd: NULL pointer");
    if (!count) return 0;
    hipLaunchKernelGGL(k_relu_bwd, dim3(blocks_for(count, 256)), dim3(256), 0, (hipStream_t)stream, y, dy,
                       count, dz);
    return launched("isac_nn_relu_bwd");
}

int isac_nn_conv_wgrad(const uint16_t *x, const uint16_t *dz, int n, int H, int W, int Cin,
                       int Cout, int ks, float *g, float *part, size_t part_bytes, void *stream) {
    if ((ks != 1 && ks != 3) || Cin < 1 || Cout < 1 || n < 1 || H < 1 || W < 1 || !x || !dz || !g || !part)
        return fail(2, "isac_nn_conv_wgrad: bad arguments");
    const int K1 = ks * ks * Cin + 1;
    const size_t per = (size_t)Cout * K1 * sizeof(float);
    const long long NP = (long long)n * H * W;
    long long nparts = (long long)(part_bytes / per);
    nparts = std::min<long long>(nparts, std::max<long long>(1, NP / 512));
    nparts = std::min<long long>(nparts, 512);
    if (nparts < 1) return fail(2, "isac_nn_conv_wgrad: scratch smaller than one partial sum");
    const long long chunk = ((NP + nparts - 1) / nparts + kWgPix - 1) / kWgPix * kWgPix;
    nparts = (NP + chunk - 1) / chunk;
    hipStream_t st = (hipStream_t)stream;
    dim3 grid((unsigned)((K1 + 63) / 64), (unsigned)((Cout + 63) / 64), (unsigned)nparts);
    hipLaunchKernelGGL(k_wgrad, grid, dim3(256), 0, st, x, dz, n, H, W, Cin, Cout, ks, chunk, part);
    if (int e = launched("isac_nn_conv_wgrad")) return e;
    const size_t cnt = (size_t)Cout * K1;
    hipLaunchKernelGGL(k_sum_parts, dim3(blocks_for(cnt, 256)), dim3(256), 0, st, part, (int)nparts, cnt, g);
    return launched("isac_nn_conv_wgrad (sum)");
}

int isac_nn_softmax_xent(const float *logits, int n, int ld, int h0, int h1, int h2,
                         const int *labels, float sigma_psi, float sigma_shift, int shift_ny,
                         float *loss, uint16_t *dlogits, void *stream) {
    if (n < 0 || h0 < 1 || h1 < 1 || h2 < 1 || ld < h0 + h1 + h2 || sigma_psi < 0.0f ||
        sigma_shift < 0.0f || (sigma_psi > 0.0f && h1 % 2) ||
        (sigma_shift > 0.0f && (shift_ny < 1 || h2 % shift_ny)) ||
        (n > 0 && (!logits || !labels || !loss || !dlogits)))
        return fail(2, "isac_nn_softmax_xent: bad arguments");
    if (n == 0) return 0;
    hipLaunchKernelGGL(k_softmax_xent, dim3((unsigned)((n + 3) / 4)), dim3(256), 0, (hipStream_t)stream,
                       logits, n, ld, h0, h1, h2, labels, sigma_psi, sigma_shift, shift_ny, loss,
                       dlogits);
    return launched("isac_nn_softmax_xent");
}

int isac_nn_adamw(float *p, const float *g, float *m, float *v, const uint8_t *mask, size_t count,
                  float lr, float beta1, float beta2, float eps, float wd, int t, void *stream) {
    if (t < 1 || (count && (!p || !g || !m || !v))) return fail(2, "isac_nn_adamw: bad arguments");
    if (!count) return 0;
    const float c1 = 1.0f - powf(beta1, (float)t), c2 = 1.0f - powf(beta2, (float)t);
    hipLaunchKernelGGL(k_adamw, dim3(blocks_for(count, 256)), dim3(256), 0, (hipStream_t)stream, p, g, m,
                       v, mask, count, lr, beta1, beta2, eps, wd, c1, c2);
    return launched("isac_nn_adamw");
}

int isac_nn_pack_weights(const float *master, int Cout, int Cin, int ks, uint16_t *w,
                         uint16_t *wflip, void *stream) {
    if (Cout < 1 || Cin < 1 || (ks != 1 && ks != 3) || !master || !w)
        return fail(2, "isac_nn_pack_weights: bad arguments");
    const size_t cnt = (size_t)Cout * ks * ks * Cin;
    hipLaunchKernelGGL(k_pack_weights, dim3(blocks_for(cnt, 256)), dim3(256), 0, (hipStream_t)stream,
                       master, Cout, Cin, ks, w, wflip);
    return launched("isac_nn_pack_weights");
}

int isac_nn_proposals(const float *logits, int n, int ld, int n_ref, int n_psi, int n_shift, int k,
                      int psi_r, int shift_r, int shift_ny, int min_refs, int *prop, float *conf,
                      void *stream) {
    if (n < 0 || k < 1 || k > 8 || n_ref < 1 || n_ref > 512 || n_psi < 1 || 2 * n_psi > 256 ||
        n_shift < 1 || n_shift > 64 || ld < n_ref + 2 * n_psi + n_shift ||
        (psi_r >= 0 && (shift_r < 0 || shift_ny < 1 || n_shift % shift_ny)) ||
        (n > 0 && (!logits || !prop)))
        return fail(2, "isac_nn_proposals: bad arguments (refs <= 512, psi classes <= 256, shifts <= 64, k <= 8)");
    if (n == 0) return 0;
    hipLaunchKernelGGL(k_proposals, dim3((unsigned)((n + 3) / 4)), dim3(256), 0, (hipStream_t)stream,
                       logits, n, ld, n_ref, n_psi, n_shift, k, psi_r, shift_r, shift_ny > 0 ? shift_ny : 1,
                       min_refs, reinterpret_cast<int4 *>(prop), conf);
    return launched("isac_nn_proposals");
}

}  // extern "C"
