// Mixed-precision (MFMA) form of the multi-reference search for CDNA4 (gfx950).
//
// Ring contraction: FP16 or BF16 operands on v_mfma_f32_4x4x4 (16 blocks), FP32
// accumulation.  Each instruction batches 16 spectrum slots (one per 4x4 block):
// block rows = 2 shift rows x (re, im) of the particle spectrum, block columns =
// 2 references x (w*Re, w*Im), K = 4 rings.  The 4x4 block holds the four real
// products P1..P4 of 4 (row, reference) pairs at that slot.
//
// Psi transform: the 256-point inverse DFT of Z = C0 + i C1 runs on
// v_mfma_f32_16x16x16_f16 as DFT16 (over g, k = b + 16 g) -> twiddle
// e^{+2 pi i b t1 / 256} -> DFT16 (over b); the accumulator layout of the first
// stage (lane l: b = 4(l/16)+v, t1 = l%16) is exactly the A-operand layout of the
// second, so no data moves between the two stages.  Output: lane l holds
// z[t1 + 16 t2] for t1 = 4(l/16)+v, t2 = l%16.
//
// Scaling: particle operands are normalised spectra (the DC correction already
// applied); reference operands carry the ring weight times 1/R ("per-ring"
// scaling, undone by multiplying the accumulator by R).  Before the FP16 psi
// transform each (row, reference) pair is scaled by a power of two chosen from
// the Cauchy-Schwarz bound |C[k]| <= ||S||_F ||wR||_F so that |Z| <= 2048 and the
// DFT16 intermediate stays below the FP16 maximum.
//
// The results are approximate (FP16/BF16 rounding); candidates are re-scored
// exactly in FP32 by k_ccf_cand4 (kernels.hip) and near-ties in FP64 here.
#include "isac_internal.h"

#include <climits>

namespace isac {
namespace {

typedef _Float16 half4_t __attribute__((ext_vector_type(4)));
typedef short short4_t __attribute__((ext_vector_type(4)));
typedef float float4_t __attribute__((ext_vector_type(4)));

constexpr int kWave = 64;

__device__ __forceinline__ void wave_lds_sync() {
    __builtin_amdgcn_fence(__ATOMIC_RELEASE, "wavefront");
    __builtin_amdgcn_wave_barrier();
    __builtin_amdgcn_fence(__ATOMIC_ACQUIRE, "wavefront");
}

__device__ __forceinline__ unsigned short f32_to_bf16(float x) {
    unsigned u = __float_as_uint(x);
    u += 0x7fffu + ((u >> 16) & 1u);  // round to nearest even
    return (unsigned short)(u >> 16);
}

__device__ __forceinline__ unsigned short f32_to_f16(float x) {
    _Float16 h = (_Float16)x;
    return __builtin_bit_cast(unsigned short, h);
}

__device__ __forceinline__ int xcd_remap(int b, int nb) {
    const int q = nb >> 3, r = nb & 7;
    const int xcd = b & 7, idx = b >> 3;
    return (xcd < r) ? xcd * (q + 1) + idx : r * (q + 1) + (xcd - r) * q + idx;
}

__device__ __forceinline__ void wave_argmax(float &v, int &t) {
#pragma unroll
    for (int off = 32; off >= 1; off >>= 1) {
        const float ov = __shfl_xor(v, off, kWave);
        const int ot = __shfl_xor(t, off, kWave);
        if (ov > v || (ov == v && ot < t)) {
            v = ov;
            t = ot;
        }
    }
}

// FP32 spectra -> MFMA operand layout [row][q][slot][part][4 rings] (16-bit), rings
// padded with zeros to 4Q.  Particles: normalisation applied (g, DC correction).
// References: the input is already normalised and ring-weighted; scaled by 1/R.
// Also writes the Frobenius norm of the (unscaled) spectrum of each row.
template <bool BF16>
__global__ __launch_bounds__(256) void k_spec_to_mixed(const float2 *__restrict__ spec,
                                                       const float2 *__restrict__ norm, int n_rows,
                                                       int R, int Q, int is_ref,
                                                       ushort4 *__restrict__ out,
                                                       float *__restrict__ row_norm) {
    const int lane = threadIdx.x & 63, wave = threadIdx.x >> 6;
    const int row = blockIdx.x * 4 + wave;
    if (row >= n_rows) return;
    const float2 nm = is_ref ? make_float2(0.0f, 1.0f) : norm[row];
    const float post = is_ref ? 1.0f / (float)R : 1.0f;
    // one pass over the rings: slots lane and lane + 64 share one 16-byte word
    // (slot_pos); the squared norm of each slot half accumulates on its own
    double ssh[2] = {0.0, 0.0};
    for (int q = 0; q < Q; ++q) {
        unsigned short re[2][4], im[2][4];
#pragma unroll
        for (int e = 0; e < 4; ++e) {
            const int i = 4 * q + e;
            float4 v4 = make_float4(0.0f, 0.0f, 0.0f, 0.0f);
            if (i < R) v4 = reinterpret_cast<const float4 *>(spec + ((size_t)row * R + i) * kSlots)[lane];
#pragma unroll
            for (int h = 0; h < 2; ++h) {
                float2 x = h ? make_float2(v4.z, v4.w) : make_float2(v4.x, v4.y);
                if (i < R && !is_ref) {
                    x.x = (lane == 0 && h == 0) ? (x.x - nm.x) * nm.y : x.x * nm.y;
                    x.y *= nm.y;
                }
                ssh[h] += (double)x.x * x.x + (double)x.y * x.y;
                x.x *= post;
                x.y *= post;
                re[h][e] = BF16 ? f32_to_bf16(x.x) : f32_to_f16(x.x);
                im[h][e] = BF16 ? f32_to_bf16(x.y) : f32_to_f16(x.y);
            }
        }
#pragma unroll
        for (int h = 0; h < 2; ++h) {
            ushort4 *o = out + (((size_t)row * Q + q) * kSlots + lane + 64 * h) * 2;
            o[0] = make_ushort4(re[h][0], re[h][1], re[h][2], re[h][3]);
            o[1] = make_ushort4(im[h][0], im[h][1], im[h][2], im[h][3]);
        }
    }
    double ss = ssh[0] + ssh[1];
#pragma unroll
    for (int off = 32; off >= 1; off >>= 1) ss += __shfl_xor(ss, off, kWave);
    if (lane == 0) row_norm[row] = (float)sqrt(ss);
}

template <bool BF16>
__device__ __forceinline__ float4_t mfma4(uint2 a, uint2 b, float4_t c) {
    if (BF16)
        return __builtin_amdgcn_mfma_f32_4x4x4bf16_1k(__builtin_bit_cast(short4_t, a),
                                                      __builtin_bit_cast(short4_t, b), c, 0, 0, 0);
    return __builtin_amdgcn_mfma_f32_4x4x4f16(__builtin_bit_cast(half4_t, a),
                                              __builtin_bit_cast(half4_t, b), c, 0, 0, 0);
}

__device__ __forceinline__ float4_t mfma16(half4_t a, half4_t b, float4_t c) {
    return __builtin_amdgcn_mfma_f32_16x16x16f16(a, b, c, 0, 0, 0);
}

__device__ __forceinline__ half4_t to_h4(float a, float b, float c, float d) {
    half4_t h;
    h[0] = (_Float16)a;
    h[1] = (_Float16)b;
    h[2] = (_Float16)c;
    h[3] = (_Float16)d;
    return h;
}

// One workgroup = one particle x 16 references (4 wavefronts x 4 references).
// For each block of 4 shift rows the particle operands are staged in LDS in
// double-buffered chunks of kQChunk ring-quads with direct global->LDS loads
// (one wavefront instruction moves 1 KB = 64 slots x (re, im) of one row);
// reference operands come through the vector L1.
constexpr int kQChunk = 3;

// Position of Z-form slot u (0..63) of pair pr in a wavefront's zre / zim: pairs
// 65 apart and bit 1 of slots 32..63 flipped.
__device__ __forceinline__ int zpos(int pr, int u) { return pr * 65 + (u ^ (((u >> 5) & 1) << 1)); }

template <bool BF16>
__global__ __launch_bounds__(256, 2) void k_mref_mixed(
    const uint2 *__restrict__ pop, const float *__restrict__ pnorm, const uint2 *__restrict__ rop,
    const float *__restrict__ rnorm, int n_img, int S, int Q, int N, int G, int GB, float rscale,
    PeakRec *__restrict__ rec) {
    // staged particle operands: [buf][qq][row 0..3][slot 0..127][part] uint2
    // rows padded by 16 uint2 (128 bytes): the two rows one MFMA operand read
    // touches fall in different bank halves
    __shared__ uint2 abuf[2][kQChunk][4][kSlots * 2 + 16];
    // [wave][pair * 65 + zslot(b*16+g)] fp16 bits, 4 per ushort4 (stride 65 and the
    // slot swizzle keep the Z-form writes and the stage-1 reads free of bank conflicts)
    __shared__ ushort4 zre[4][4 * 65];
    __shared__ ushort4 zim[4][4 * 65];
    __shared__ float4 zc[4][64];        // psi samples of one mirror for neighbour capture
    __shared__ float nbuf[4][4][2][8];
    const int lane = threadIdx.x & 63, wave = __builtin_amdgcn_readfirstlane(threadIdx.x >> 6);
    const int lb = xcd_remap(blockIdx.x, gridDim.x);
    const int p = lb / GB;
    if (p >= n_img) return;   // uniform over the workgroup
    const int g4 = (lb % GB) * 4 + wave;
    const bool active = g4 < G;

    int nref[4];
    bool nval[4];
    float rn[4];
#pragma unroll
    for (int c = 0; c < 4; ++c) {
        const int n = g4 * 4 + c;
        nval[c] = active && n < N;
        nref[c] = min(n, N - 1);
        rn[c] = rnorm[nref[c]];
    }
    // DFT16 operands (inverse): lane l holds B[k = 4(l/16)+v][j = l%16]
    half4_t Bc, Bs, Bns;
    float tw_re[4], tw_im[4];
    {
        const int j = lane & 15, kb = 4 * (lane >> 4);
#pragma unroll
        for (int v = 0; v < 4; ++v) {
            const int m = ((kb + v) * j) & 15;
            float sn, cs;
            __sincosf(6.283185307179586f * (float)m / 16.0f, &sn, &cs);
            Bc[v] = (_Float16)cs;
            Bs[v] = (_Float16)sn;
            Bns[v] = (_Float16)(-sn);
            // stage-2 twiddle for b = kb + v, t1 = j: e^{+2 pi i b t1 / 256}
            const float a = 6.283185307179586f * (float)(((kb + v) * j) & 255) / 256.0f;
            __sincosf(a, &tw_im[v], &tw_re[v]);
        }
    }
    float bv[4][2];
    int bf[4][2];
#pragma unroll
    for (int c = 0; c < 4; ++c) {
        bv[c][0] = bv[c][1] = -3.40282e+38f;
        bf[c][0] = bf[c][1] = INT_MAX;
    }
    const int blk = lane >> 2, jj = lane & 3;
    const size_t row_stride = (size_t)Q * kSlots * 2;   // uint2 per (image, shift)
    const int nqc = (Q + kQChunk - 1) / kQChunk;
    const uint2 *brow[2];
#pragma unroll
    for (int cp = 0; cp < 2; ++cp)
        brow[cp] = rop + (size_t)nref[2 * cp + (jj >> 1)] * row_stride + (jj & 1);

    for (int s0 = 0; s0 < S; s0 += 4) {
        const uint2 *rows[4];
#pragma unroll
        for (int r = 0; r < 4; ++r)
            rows[r] = pop + ((size_t)p * S + min(s0 + r, S - 1)) * row_stride;
        // one chunk = kQChunk ring-quads x 4 rows x 2 halves (1 KB each)
        auto load_chunk = [&](int qc, int buf) {
            for (int u = wave; u < kQChunk * 4 * 2; u += 4) {
                const int qq = u >> 3, r = (u >> 1) & 3, half = u & 1;
                const int q = qc * kQChunk + qq;
                if (q >= Q) continue;
                __builtin_amdgcn_global_load_lds(
                    (const void *)(rows[r] + (size_t)q * kSlots * 2 + half * kSlots + 2 * lane),
                    (__attribute__((address_space(3))) void *)(&abuf[buf][qq][r][half * kSlots]),
                    16, 0, 0);
            }
        };
        float4_t acc[8][2][2];
#pragma unroll
        for (int g = 0; g < 8; ++g)
#pragma unroll
            for (int a = 0; a < 2; ++a)
#pragma unroll
                for (int b = 0; b < 2; ++b) acc[g][a][b] = float4_t{0.f, 0.f, 0.f, 0.f};
        load_chunk(0, 0);
        __builtin_amdgcn_s_waitcnt(0);
        __syncthreads();
        for (int qc = 0; qc < nqc; ++qc) {
            const int buf = qc & 1;
            if (qc + 1 < nqc) load_chunk(qc + 1, buf ^ 1);
            const int q1 = min(Q, (qc + 1) * kQChunk);
            for (int q = qc * kQChunk; q < q1; ++q) {
                const int qq = q - qc * kQChunk;
                const uint2 *ar0 = &abuf[buf][qq][jj >> 1][(jj & 1)];
                const uint2 *ar1 = &abuf[buf][qq][2 + (jj >> 1)][(jj & 1)];
                // the quad's reference operands are all issued before its first LDS
                // read (the compiler waits for every outstanding global load there),
                // so the quad pays one round trip instead of one per slot block
                uint2 bq[8][2];
#pragma unroll
                for (int g = 0; g < 8; ++g) {
                    const size_t off = ((size_t)q * kSlots + 16 * g + blk) * 2;
                    bq[g][0] = brow[0][off];
                    bq[g][1] = brow[1][off];
                }
                __builtin_amdgcn_sched_barrier(0);
#pragma unroll
                for (int g = 0; g < 8; ++g) {
                    const int so = (16 * g + blk) * 2;
                    const uint2 a0 = ar0[so], a1 = ar1[so];
                    const uint2 b0 = bq[g][0], b1 = bq[g][1];
                    acc[g][0][0] = mfma4<BF16>(a0, b0, acc[g][0][0]);
                    acc[g][0][1] = mfma4<BF16>(a0, b1, acc[g][0][1]);
                    acc[g][1][0] = mfma4<BF16>(a1, b0, acc[g][1][0]);
                    acc[g][1][1] = mfma4<BF16>(a1, b1, acc[g][1][1]);
                }
            }
            __builtin_amdgcn_s_waitcnt(0);
            __syncthreads();
        }
        if (!active) continue;
        // psi stage, four (row, reference) pairs per round: pair (ri, ci) of row pair rp
        // and reference pair cp; lane j = 2ci + bpart writes Z-forms (bpart 0) or the
        // Hermitian partners (bpart 1) of its 8 slots 16g + blk.
#pragma unroll
        for (int rp = 0; rp < 2; ++rp) {
#pragma unroll
            for (int cp = 0; cp < 2; ++cp) {
                const int ci = jj >> 1, bpart = jj & 1;
                float sc[2];
#pragma unroll
                for (int ri = 0; ri < 2; ++ri) {
                    const int s = min(s0 + 2 * rp + ri, S - 1);
                    const float bound = pnorm[(size_t)p * S + s] * rn[2 * cp + ci];
                    int e;
                    frexpf(2048.0f / fmaxf(bound, 1e-30f), &e);
                    sc[ri] = ldexpf(1.0f, e - 1);
                }
#pragma unroll
                for (int gh = 0; gh < 2; ++gh) {   // slots g = 4gh .. 4gh+3
                    unsigned short hre[2][4], him[2][4];
#pragma unroll
                    for (int gg = 0; gg < 4; ++gg) {
                        const int g = 4 * gh + gg;
                        const float4_t d = acc[g][rp][cp];
                        float4_t o;
                        o[0] = __shfl_xor(d[0], 1, kWave);
                        o[1] = __shfl_xor(d[1], 1, kWave);
                        o[2] = __shfl_xor(d[2], 1, kWave);
                        o[3] = __shfl_xor(d[3], 1, kWave);
#pragma unroll
                        for (int ri = 0; ri < 2; ++ri) {
                            const float P1 = (bpart ? o[2 * ri] : d[2 * ri]) * rscale;
                            const float P3 = (bpart ? o[2 * ri + 1] : d[2 * ri + 1]) * rscale;
                            const float P4 = (bpart ? d[2 * ri] : o[2 * ri]) * rscale;
                            const float P2 = (bpart ? d[2 * ri + 1] : o[2 * ri + 1]) * rscale;
                            const float A = P1 + P2, B = P4 - P3, Cc = P1 - P2, Dd = -(P3 + P4);
                            float zr, zi;
                            if (bpart == 0) {
                                zr = A - Dd;
                                zi = B + Cc;
                                if (g == 0 && blk == 0) zr = zi = P1;   // Z[0] = C[0](1 + i)
                            } else {
                                zr = A + Dd;
                                zi = Cc - B;
                                if (g == 0 && blk == 0) zr = zi = P2;   // Z[128] = C[128](1 + i)
                            }
                            hre[ri][gg] = f32_to_f16(zr * sc[ri]);
                            him[ri][gg] = f32_to_f16(zi * sc[ri]);
                        }
                    }
                    // Z[k], k = 16 g + blk -> half index b*16 + g' (b = k%16, g' = k/16):
                    //   bpart 0: blk*16 + g;  bpart 1 (partner 256-k): blk >= 1 ->
                    //   (16-blk)*16 + 15-g; blk == 0 -> 16-g, with the Nyquist term (g=0) at 8.
#pragma unroll
                    for (int ri = 0; ri < 2; ++ri) {
                        const int pr = 2 * ri + ci;
                        if (bpart == 0) {
                            zre[wave][zpos(pr, blk * 4 + gh)] = make_ushort4(hre[ri][0], hre[ri][1], hre[ri][2], hre[ri][3]);
                            zim[wave][zpos(pr, blk * 4 + gh)] = make_ushort4(him[ri][0], him[ri][1], him[ri][2], him[ri][3]);
                        } else if (blk) {
                            // g = 4gh..4gh+3 -> g' = 15-g: gh 0 -> 15..12 (ushort4 3), gh 1 -> 11..8 (2)
                            const int u = (16 - blk) * 4 + 3 - gh;
                            zre[wave][zpos(pr, u)] = make_ushort4(hre[ri][3], hre[ri][2], hre[ri][1], hre[ri][0]);
                            zim[wave][zpos(pr, u)] = make_ushort4(him[ri][3], him[ri][2], him[ri][1], him[ri][0]);
                        } else {
                            // blk 0: g' = 16-g (g >= 1), Nyquist (g = 0) at 8:
                            // gh 0: g 0..3 -> 8, 15, 14, 13 ; gh 1: g 4..7 -> 12, 11, 10, 9
                            // slots 2 and 3 (< 32): zpos leaves them in place
                            unsigned short *zr16 = reinterpret_cast<unsigned short *>(&zre[wave][pr * 65]);
                            unsigned short *zi16 = reinterpret_cast<unsigned short *>(&zim[wave][pr * 65]);
#pragma unroll
                            for (int gg = 0; gg < 4; ++gg) {
                                const int g = 4 * gh + gg;
                                const int idx = g ? 16 - g : 8;
                                zr16[idx] = hre[ri][gg];
                                zi16[idx] = him[ri][gg];
                            }
                        }
                    }
                }
                wave_lds_sync();
#pragma unroll
                for (int pr = 0; pr < 4; ++pr) {
                    const int ri = pr >> 1, c = 2 * cp + (pr & 1);
                    const int s = s0 + 2 * rp + ri;
                    if (s >= S || !nval[c]) continue;
                    const int ci2 = pr & 1;
                    const float inv = 1.0f / __shfl(sc[ri], 2 * ci2, kWave);
                    // stage 1: rows b = l%16, K = g = 4(l/16)+v (re plane, then im plane)
                    const ushort4 ar = zre[wave][zpos(pr, (lane & 15) * 4 + (lane >> 4))];
                    const ushort4 ai = zim[wave][zpos(pr, (lane & 15) * 4 + (lane >> 4))];
                    const half4_t Are = __builtin_bit_cast(half4_t, ar);
                    const half4_t Aim = __builtin_bit_cast(half4_t, ai);
                    const float4_t z4 = {0.f, 0.f, 0.f, 0.f};
                    float4_t y_re = mfma16(Are, Bc, z4);
                    y_re = mfma16(Aim, Bns, y_re);
                    float4_t y_im = mfma16(Are, Bs, z4);
                    y_im = mfma16(Aim, Bc, y_im);
                    float tr[4], ti[4];
#pragma unroll
                    for (int v = 0; v < 4; ++v) {
                        tr[v] = y_re[v] * tw_re[v] - y_im[v] * tw_im[v];
                        ti[v] = y_re[v] * tw_im[v] + y_im[v] * tw_re[v];
                    }
                    const half4_t A2r = to_h4(tr[0], tr[1], tr[2], tr[3]);
                    const half4_t A2i = to_h4(ti[0], ti[1], ti[2], ti[3]);
                    float4_t c0 = mfma16(A2r, Bc, z4);
                    c0 = mfma16(A2i, Bns, c0);
                    float4_t c1 = mfma16(A2r, Bs, z4);
                    c1 = mfma16(A2i, Bc, c1);
                    // lane l holds t = 4(l/16) + v + 16 (l%16)
                    const int tb = 4 * (lane >> 4) + 16 * (lane & 15);
#pragma unroll
                    for (int m = 0; m < 2; ++m) {
                        float x[4];
#pragma unroll
                        for (int v = 0; v < 4; ++v) x[v] = (m ? c1[v] : c0[v]) * inv;
                        float bvl = x[0];
                        int bt = tb;
#pragma unroll
                        for (int v = 1; v < 4; ++v)
                            if (x[v] > bvl) {
                                bvl = x[v];
                                bt = tb + v;
                            }
                        if (!__any(bvl > bv[c][m])) continue;
                        wave_argmax(bvl, bt);
                        const int flat = s * kRingLen + bt;
                        if (bvl > bv[c][m] || (bvl == bv[c][m] && flat < bf[c][m])) {
                            bv[c][m] = bvl;
                            bf[c][m] = flat;
                            wave_lds_sync();
                            zc[wave][(lane >> 4) + 4 * (lane & 15)] = make_float4(x[0], x[1], x[2], x[3]);
                            wave_lds_sync();
                            const float *zf = reinterpret_cast<const float *>(zc[wave]);
                            if (lane < 7) nbuf[wave][c][m][lane] = zf[(bt + lane - 3) & 255];
                            wave_lds_sync();
                        }
                    }
                }
                wave_lds_sync();
            }
        }
    }
#pragma unroll
    for (int c = 0; c < 4; ++c) {
        if (!nval[c]) continue;
        const int m = (bv[c][1] > bv[c][0]) ? 1 : 0;
        const int f = (bf[c][m] == INT_MAX) ? 0 : bf[c][m];   // never updated (all values NaN)
        PeakRec *o = rec + (size_t)p * N + nref[c];
        if (lane == 0) {
            o->val = bv[c][m];
            o->m = m;
            o->s = f / kRingLen;
            o->t = f % kRingLen;
            o->pad = 0;
        }
        if (lane < 7) o->x[lane] = nbuf[wave][c][m][lane];
    }
}

}  // namespace

hipError_t launch_spec_to_mixed(const float2 *spec, const float2 *norm, int n_rows, int n_ring,
                                int is_ref, bool bf16, void *out, float *row_norm, hipStream_t st) {
    if (n_rows <= 0) return hipSuccess;
    const int Q = (n_ring + 3) / 4;
    const dim3 grid((n_rows + 3) / 4);
    if (bf16)
        hipLaunchKernelGGL((k_spec_to_mixed<true>), grid, dim3(256), 0, st, spec, norm, n_rows,
                           n_ring, Q, is_ref, (ushort4 *)out, row_norm);
    else
        hipLaunchKernelGGL((k_spec_to_mixed<false>), grid, dim3(256), 0, st, spec, norm, n_rows,
                           n_ring, Q, is_ref, (ushort4 *)out, row_norm);
    return hipGetLastError();
}

hipError_t launch_mref_mixed(const void *pop, const float *pnorm, const void *rop,
                             const float *rnorm, int n_img, int n_shift, int n_ring, int n_ref,
                             bool bf16, PeakRec *rec, hipStream_t st) {
    if (n_img <= 0 || n_ref <= 0) return hipSuccess;
    const int Q = (n_ring + 3) / 4;
    const int G = (n_ref + 3) / 4;
    const int GB = (G + 3) / 4;
    const long long blocks = (long long)n_img * GB;
    const float rscale = (float)n_ring;
    if (bf16)
        hipLaunchKernelGGL((k_mref_mixed<true>), dim3((unsigned)blocks), dim3(256), 0, st,
                           (const uint2 *)pop, pnorm, (const uint2 *)rop, rnorm, n_img, n_shift, Q,
                           n_ref, G, GB, rscale, rec);
    else
        hipLaunchKernelGGL((k_mref_mixed<false>), dim3((unsigned)blocks), dim3(256), 0, st,
                           (const uint2 *)pop, pnorm, (const uint2 *)rop, rnorm, n_img, n_shift, Q,
                           n_ref, G, GB, rscale, rec);
    return hipGetLastError();
}

}  // namespace isac
