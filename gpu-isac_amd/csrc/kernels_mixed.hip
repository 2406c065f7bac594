// Mixed-precision (MFMA) form of the multi-reference search for CDNA4 (gfx950).
//
// Ring contraction as a batched GEMM, one per spectrum slot k.  With the
// particle operand a = sr + si, b = sr - si (normalised spectrum S of one shift
// row) and the reference operand u = rr + ri, v = rr - ri (normalised,
// ring-weighted spectrum R), the four real products over the rings are exactly
// the inputs of the psi transform:
//     Z[k]       = C0[k] + i C1[k]                         = (a.u) + i (b.u)
//     Z[256 - k] = conj(C0[k]) + i conj(C1[k])              = (b.v) + i (a.v)
// where C0 = sum_i w_i R conj(S), C1 = sum_i w_i conj(R S) are the reference's
// two mirror cross-spectra (cuda/gpu_aln_noref.cu:1508-1536) and z = IDFT(Z)
// holds c0[t] + i c1[t], the psi CCF of both mirrors (:2393, :2459).  Slot 0
// packs the two real bins (X[0], X[128]): a = S[0], b = S[128], u = R[0],
// v = R[128], so Z[0] = (a.u)(1 + i) and Z[128] = (b.v)(1 + i).
//
// Search kernel (k_mref_gemm): workgroup = 16 particles x 16 references, four
// wavefronts.  For each block of 16 consecutive (particle, shift) rows:
//  * contraction: per slot one 32 x 32 output tile (rows = 16 rows x {a, b},
//    columns = 16 references x {u, v}) on v_mfma_f32_32x32x16 (+ 32x32x8 for
//    the last rings), K = rings; each wavefront computes a quarter of the slots;
//  * the tiles are converted to FP16 Z values and transposed through LDS into
//    the psi operand layout, which every wavefront keeps in registers for its
//    64 (row, reference) pairs (4 VGPRs per pair);
//  * psi transform per pair: the 256-point inverse DFT as DFT16 -> twiddle ->
//    DFT16 on v_mfma_f32_16x16x32_f16 (k = 16 k1 + k2, t = t1 + 16 t2); the
//    first stage's accumulator layout is the second stage's B operand;
//  * search: each psi sample carries its (shift, sample) index in the low
//    mantissa bits (first-max-wins order: a smaller index gets a larger code),
//    so the running maximum per (reference, mirror) is one v_max3 per three
//    samples; a wave-wide reduction per particle decodes the winner.
// The values are approximate (FP16/BF16 operands, FP16 psi transform, and the
// index bits perturb a value by at most 2^-(23 - index bits) relative); the
// best candidates are re-scored exactly in FP32 (k_ccf_cand_mref, kernels.hip) and
// near-ties in FP64 (kernels.hip).  The approximate records carry the peak
// sample only (no sub-sample refinement).
//
// Scaling: the particle operands are the normalised spectra times a global
// power of two c_p; each reference's operands are scaled by a power of two c_n
// so that its Frobenius norm lies in [0.7, 1.42).  For a normalised row the
// Frobenius norm of (a, b) is at most 256 sqrt(R (R + 1)), so |Z|, and the
// DFT16 intermediates, are bounded by c_p 256 sqrt(R (R + 1)) 1.42, which c_p
// keeps below 32768 (FP16 max 65504).  Results are multiplied by 1 / (c_p c_n).
#include "isac_internal.h"

#include <cfloat>
#include <climits>

namespace isac {
namespace {

typedef _Float16 half8_t __attribute__((ext_vector_type(8)));
typedef _Float16 half4_t __attribute__((ext_vector_type(4)));
typedef __bf16 bf16x8_t __attribute__((ext_vector_type(8)));
typedef short short4_t __attribute__((ext_vector_type(4)));
typedef float float16_t __attribute__((ext_vector_type(16)));
typedef float float4_t __attribute__((ext_vector_type(4)));
typedef _Float16 h2_t __attribute__((ext_vector_type(2)));

constexpr int kWave = 64;

__device__ __forceinline__ unsigned short f32_to_bf16(float x) {
    unsigned u = __float_as_uint(x);
    u += 0x7fffu + ((u >> 16) & 1u);  // round to nearest even
    return (unsigned short)(u >> 16);
}

__device__ __forceinline__ unsigned short f32_to_h(float x) {
    return __builtin_bit_cast(unsigned short, (_Float16)x);
}

template <bool BF16>
__device__ __forceinline__ unsigned short to16(float x) {
    return BF16 ? f32_to_bf16(x) : f32_to_h(x);
}

// two FP32 values as an FP16 pair (one v_cvt_pk_f16_f32, round to nearest even)
__device__ __forceinline__ unsigned pack_h2(float lo, float hi) {
    typedef _Float16 hp_t __attribute__((ext_vector_type(2)));
    const hp_t v = {(_Float16)lo, (_Float16)hi};
    return __builtin_bit_cast(unsigned, v);
}

template <bool BF16>
__device__ __forceinline__ float16_t mma32x16(uint4 a, uint4 b, float16_t c) {
    if (BF16)
        return __builtin_amdgcn_mfma_f32_32x32x16_bf16(__builtin_bit_cast(bf16x8_t, a),
                                                       __builtin_bit_cast(bf16x8_t, b), c, 0, 0, 0);
    return __builtin_amdgcn_mfma_f32_32x32x16_f16(__builtin_bit_cast(half8_t, a),
                                                  __builtin_bit_cast(half8_t, b), c, 0, 0, 0);
}

template <bool BF16>
__device__ __forceinline__ float16_t mma32x8(uint2 a, uint2 b, float16_t c) {
    if (BF16)
        return __builtin_amdgcn_mfma_f32_32x32x8bf16_1k(__builtin_bit_cast(short4_t, a),
                                                        __builtin_bit_cast(short4_t, b), c, 0, 0, 0);
    return __builtin_amdgcn_mfma_f32_32x32x8f16(__builtin_bit_cast(half4_t, a),
                                                __builtin_bit_cast(half4_t, b), c, 0, 0, 0);
}

__device__ __forceinline__ float4_t mma16x32(uint4 a, uint4 b, float4_t c) {
    return __builtin_amdgcn_mfma_f32_16x16x32_f16(__builtin_bit_cast(half8_t, a),
                                                  __builtin_bit_cast(half8_t, b), c, 0, 0, 0);
}

// max of three finite values without the NaN-quieting copies that fmaxf inserts
// for operands assembled from bits (the embedded psi values are finite by the
// scaling; -inf only initialises the running maxima)
__device__ __forceinline__ float vmax3(float a, float b, float c) {
    float d;
    asm("v_max3_f32 %0, %1, %2, %3" : "=v"(d) : "v"(a), "v"(b), "v"(c));
    return d;
}

__device__ __forceinline__ int xcd_remap(int b, int nb) {
    const int q = nb >> 3, r = nb & 7;
    const int xcd = b & 7, idx = b >> 3;
    return (xcd < r) ? xcd * (q + 1) + idx : r * (q + 1) + (xcd - r) * q + idx;
}

// ----------------------------------------------------------- operand layout
// Rows (particle shift rows or references) are contiguous: row r, slot k at
// r * row_bytes + k * slot_bytes, slot_bytes = 4 K_pad (two parts, K_pad rings,
// 16-bit).  Inside a slot, each K-step of 16 rings is 64 bytes: part (a / b or
// u / v) x half h, 16 bytes = rings 16 ks + 8 h .. + 7; a tail of 8 rings is 32
// bytes: part x h, 8 bytes = rings 16 ks + 4 h .. + 3.  An MFMA operand of slot
// k, step ks for a block of 16 rows is then lane l = 2 idx + part + 32 h ->
// row idx's 16 (or 8) bytes at part / h: four lanes read 64 contiguous bytes.
struct GemmGeom {
    int full16;      // number of 16-ring steps
    int tail;        // 0, 8 or 16 rings in the last step
    int slot_bytes;  // 4 K_pad
};

// Write the 2 x K_pad operand values of one row at slot k: val(part, i) for i < R,
// zero beyond.  row_base points at the row.
template <bool BF16, class F>
__device__ __forceinline__ void write_operand(unsigned char *row_base, int k, int R, GemmGeom g,
                                              F val) {
    unsigned char *base = row_base + (size_t)k * g.slot_bytes;
    const int n16 = g.full16 + (g.tail == 16 ? 1 : 0);
    for (int ks = 0; ks < n16; ++ks) {
#pragma unroll
        for (int part = 0; part < 2; ++part) {
#pragma unroll
            for (int h = 0; h < 2; ++h) {
                unsigned short e[8];
#pragma unroll
                for (int j = 0; j < 8; ++j) {
                    const int i = 16 * ks + 8 * h + j;
                    e[j] = to16<BF16>(i < R ? val(part, i) : 0.0f);
                }
                uint4 v;
                v.x = e[0] | ((unsigned)e[1] << 16);
                v.y = e[2] | ((unsigned)e[3] << 16);
                v.z = e[4] | ((unsigned)e[5] << 16);
                v.w = e[6] | ((unsigned)e[7] << 16);
                *reinterpret_cast<uint4 *>(base + ks * 64 + part * 32 + h * 16) = v;
            }
        }
    }
    if (g.tail == 8) {
#pragma unroll
        for (int part = 0; part < 2; ++part) {
#pragma unroll
            for (int h = 0; h < 2; ++h) {
                unsigned short e[4];
#pragma unroll
                for (int j = 0; j < 4; ++j) {
                    const int i = 16 * g.full16 + 4 * h + j;
                    e[j] = to16<BF16>(i < R ? val(part, i) : 0.0f);
                }
                uint2 v;
                v.x = e[0] | ((unsigned)e[1] << 16);
                v.y = e[2] | ((unsigned)e[3] << 16);
                *reinterpret_cast<uint2 *>(base + g.full16 * 64 + part * 16 + h * 8) = v;
            }
        }
    }
}

// Particle operands from the raw FP32 spectra: thread = (operand row, slot).
// Operand row p * S_pad + s holds spectra row p * S + s; rows with s >= S or
// p >= n_img (padding) are zero.
template <bool BF16>
__global__ __launch_bounds__(256) void k_frag_rows(const float2 *__restrict__ spec,
                                                   const float2 *__restrict__ norm, int n_img,
                                                   int S, int S_pad, int n_rows_pad, int R,
                                                   GemmGeom g, float cp,
                                                   unsigned char *__restrict__ frag) {
    const long long tid = (long long)blockIdx.x * blockDim.x + threadIdx.x;
    const int orow = (int)(tid / kSlots), k = (int)(tid % kSlots);
    if (orow >= n_rows_pad) return;
    unsigned char *rowp = frag + (size_t)orow * kSlots * g.slot_bytes;
    const int p = orow / S_pad, s = orow - p * S_pad;
    if (p >= n_img || s >= S) {
        write_operand<BF16>(rowp, k, R, g, [](int, int) { return 0.0f; });
        return;
    }
    const int row = p * S + s;
    const float2 nm = norm[row];
    const float gs = nm.y * cp;
    const float2 *src = spec + (size_t)row * R * kSlots + slot_pos(k);
    write_operand<BF16>(rowp, k, R, g, [&](int part, int i) {
        const float2 x = src[(size_t)i * kSlots];
        if (k == 0) return part ? x.y * gs : (x.x - nm.x) * gs;   // S[0] - mu L, S[128]
        return part ? (x.x - x.y) * gs : (x.x + x.y) * gs;
    });
}

// Per-reference scale: c_n = 2^-round(log2 ||(u, v)||_F).  Block = one reference.
__global__ __launch_bounds__(256) void k_ref_scale(const float2 *__restrict__ rspec, int R,
                                                   float cp, float *__restrict__ cn,
                                                   float *__restrict__ rinv) {
    const int n = blockIdx.x;
    double ss = 0.0;
    for (int e = threadIdx.x; e < R * kSlots; e += blockDim.x) {
        const float2 x = rspec[(size_t)n * R * kSlots + e];
        const int k = e % kSlots;   // slot_pos(0) == 0: position 0 is slot 0
        if (k == 0) ss += (double)x.x * x.x + (double)x.y * x.y;
        else ss += 2.0 * ((double)x.x * x.x + (double)x.y * x.y);
    }
    __shared__ double red[256];
    red[threadIdx.x] = ss;
    __syncthreads();
    for (int s = 128; s > 0; s >>= 1) {
        if ((int)threadIdx.x < s) red[threadIdx.x] += red[threadIdx.x + s];
        __syncthreads();
    }
    if (threadIdx.x == 0) {
        const double f = sqrt(red[0]);
        const float c = (f > 0.0) ? ldexpf(1.0f, -(int)lrint(log2(f))) : 1.0f;
        cn[n] = c;
        rinv[n] = 1.0f / (c * cp);
    }
}

// Reference operands from the normalised, ring-weighted spectra: thread = (ref, slot);
// references at or beyond n_ref (padding of the last block) are zero.
template <bool BF16>
__global__ __launch_bounds__(256) void k_frag_refs(const float2 *__restrict__ rspec, int n_ref,
                                                   int n_ref_pad, int R, GemmGeom g,
                                                   const float *__restrict__ cn,
                                                   unsigned char *__restrict__ frag) {
    const int tid = blockIdx.x * blockDim.x + threadIdx.x;
    const int n = tid / kSlots, k = tid % kSlots;
    if (n >= n_ref_pad) return;
    unsigned char *rowp = frag + (size_t)n * kSlots * g.slot_bytes;
    if (n >= n_ref) {
        write_operand<BF16>(rowp, k, R, g, [](int, int) { return 0.0f; });
        return;
    }
    const float c = cn[n];
    const float2 *src = rspec + (size_t)n * R * kSlots + slot_pos(k);
    write_operand<BF16>(rowp, k, R, g, [&](int part, int i) {
        const float2 x = src[(size_t)i * kSlots];
        if (k == 0) return part ? x.y * c : x.x * c;              // R[0], R[128]
        return part ? (x.x - x.y) * c : (x.x + x.y) * c;
    });
}

// ------------------------------------------------------------ search kernel
// psi operand layout of a pair (16x16x32 A operand, K = (k1, re/im)): lane l =
// k2 + 16 lg holds in register j the Z value at k1 = k1_of(lg, j).  A chunk c of
// 32 slots fills register j = c of every lane except lane 32 (k2 = 0, lg 2),
// whose four values (Z[224], Z[192], Z[160], Z[128]) go through a separate buffer.
__host__ __device__ constexpr int k1_of(int lg, int j) {
    return lg == 0 ? 2 * j : lg == 1 ? 2 * j + 1 : lg == 2 ? 14 - 2 * j : 15 - 2 * j;
}

// Workgroup: 8 wavefronts (two per SIMD); wavefront w owns the pairs (row r,
// reference 2 w + q), r < 16, q < 2, and computes the contraction tiles of 16
// of the 128 slots of each row block (4 per chunk of 32).
constexpr int kGemmWaves = 8;
constexpr int kGemmThreads = 64 * kGemmWaves;
constexpr int kQW = 16 / kGemmWaves;   // references per wavefront
constexpr int kRing = 2;                // slots of operand fragments in flight per wavefront

// LDS swizzle of a pair's 64 consumer lanes: owner wave and bit 1 of the row
// (spreads the 32 same-position writes of a tile over distinct banks).
__device__ __forceinline__ int zswz(int wo, int row) { return wo | (((row >> 1) & 1) << 3); }

// One slot's operand fragments: two 16-ring steps and a tail of 8 rings (the
// layout of any R <= 40; larger R loops over further 16-ring steps).
struct Frag {
    uint4 a0, b0, a1, b1;
    uint2 a8, b8;
};

// Buffer resource of a byte range (raw buffer loads: 32-bit lane offset in a VGPR,
// uniform offset in an SGPR, no 64-bit per-lane addresses).
__device__ __forceinline__ __amdgpu_buffer_rsrc_t byte_rsrc(const unsigned char *base,
                                                           unsigned bytes) {
    return __builtin_amdgcn_make_buffer_rsrc(const_cast<unsigned char *>(base), (short)0,
                                             (int)bytes, 0x00020000);
}

__device__ __forceinline__ uint4 bload16(__amdgpu_buffer_rsrc_t r, unsigned voff, unsigned soff) {
    return __builtin_bit_cast(uint4, __builtin_amdgcn_raw_buffer_load_b128(r, voff, soff, 0));
}

__device__ __forceinline__ uint2 bload8(__amdgpu_buffer_rsrc_t r, unsigned voff, unsigned soff) {
    return __builtin_bit_cast(uint2, __builtin_amdgcn_raw_buffer_load_b64(r, voff, soff, 0));
}

// One slot's fragments: pa / pb = uniform byte offsets of the slot in the particle
// and reference resources.
__device__ __forceinline__ void load_frag(Frag &f, __amdgpu_buffer_rsrc_t ra, unsigned pa,
                                          __amdgpu_buffer_rsrc_t rbr, unsigned pb, unsigned off16,
                                          unsigned off8, GemmGeom g) {
    if (g.full16 > 0) {
        f.a0 = bload16(ra, off16, pa);
        f.b0 = bload16(rbr, off16, pb);
    }
    if (g.full16 > 1 || (g.full16 == 1 && g.tail == 16)) {
        f.a1 = bload16(ra, off16, pa + 64);
        f.b1 = bload16(rbr, off16, pb + 64);
    }
    if (g.tail == 8) {
        f.a8 = bload8(ra, off8, pa + g.full16 * 64);
        f.b8 = bload8(rbr, off8, pb + g.full16 * 64);
    }
}

template <bool BF16>
__device__ __forceinline__ float16_t contract(const Frag &f, __amdgpu_buffer_rsrc_t ra, unsigned pa,
                                              __amdgpu_buffer_rsrc_t rbr, unsigned pb,
                                              unsigned off16, GemmGeom g) {
    // each step from a zero accumulator, summed on the VALU (no MFMA reads another
    // one's result as its accumulator input)
    const float16_t z0 = {};
    float16_t acc = {};
    if (g.full16 > 0) acc = mma32x16<BF16>(f.a0, f.b0, z0);
    if (g.full16 > 1 || (g.full16 == 1 && g.tail == 16)) acc += mma32x16<BF16>(f.a1, f.b1, z0);
    // rings beyond 32 (R > 40): further steps, loaded here
    for (int ks = 2; ks < g.full16 + (g.tail == 16 ? 1 : 0); ++ks)
        acc += mma32x16<BF16>(bload16(ra, off16, pa + ks * 64), bload16(rbr, off16, pb + ks * 64), z0);
    if (g.tail == 8) acc += mma32x8<BF16>(f.a8, f.b8, z0);
    return acc;
}

template <bool BF16>
__global__ __launch_bounds__(kGemmThreads, 1) void k_mref_gemm(
    const unsigned char *__restrict__ pfrag, const unsigned char *__restrict__ rfrag,
    const float *__restrict__ rinv, int P, int S, int S_pad, int N, GemmGeom g, int nbits,
    PeakRec *__restrict__ rec) {
    // [buffer][owner wave][row][consumer lane ^ swizzle] -> the owner's kQW pair
    // values (u32 = FP16 (re, im) each)
    __shared__ uint2 zc[2][kGemmWaves][16][64];
    // [row-block parity][owner wave][row * kQW + q] -> lane 32's four values
    __shared__ uint4 zl[2][kGemmWaves][16 * kQW];
    const int lane = threadIdx.x & 63;
    const int w = __builtin_amdgcn_readfirstlane(threadIdx.x >> 6);
    // 1D grid, logical block = particle group x reference block (fastest): the
    // reference blocks of a particle group are consecutive and share an XCD / L2
    const int lb = xcd_remap(blockIdx.x, gridDim.x);
    const int NB = (N + 15) / 16;
    const int pg = lb / NB, nb = lb % NB;
    const unsigned row_bytes = (unsigned)kSlots * g.slot_bytes;
    const unsigned blk_bytes = 16 * row_bytes;
    // the workgroup's particle rows and its reference block as buffer resources
    const __amdgpu_buffer_rsrc_t rsp = byte_rsrc(pfrag + (size_t)pg * S_pad * blk_bytes,
                                                 (unsigned)S_pad * blk_bytes);
    const __amdgpu_buffer_rsrc_t rsr = byte_rsrc(rfrag + (size_t)nb * blk_bytes, blk_bytes);
    // this lane's operand bytes in a block: row (l % 32) / 2, part l % 2, half l / 32
    // (32-bit: a block is at most 16 * 128 * 4 * 128 bytes)
    const unsigned off16 = ((lane & 31) >> 1) * row_bytes + (lane & 1) * 32 + (lane >> 5) * 16;
    const unsigned off8 = ((lane & 31) >> 1) * row_bytes + (lane & 1) * 16 + (lane >> 5) * 8;

    // ---- constant MFMA operands of the psi transform, kept in LDS (read per row
    // block: their registers serve the fragment pipeline during the contraction)
    __shared__ uint4 pcst[4][64];     // B1 re, B1 im, A2 re, A2 im per lane
    __shared__ float4 ptw[2][64];     // twiddle cos, sin per lane
    if (w == 0) {
        const int c16 = lane & 15, lg = lane >> 4;
        unsigned short e1r[8], e1i[8], e2r[8], e2i[8];
#pragma unroll
        for (int e = 0; e < 8; ++e) {
            const int part = e & 1;
            // stage 1: B[(k1, part)][t1], t1 = c16
            const int k1 = k1_of(lg, e >> 1);
            float s1, c1;
            sincospif((float)((k1 * c16) & 15) / 8.0f, &s1, &c1);
            e1r[e] = f32_to_h(part ? -s1 : c1);
            e1i[e] = f32_to_h(part ? c1 : s1);
            // stage 2: A[t2][(k2, part)], t2 = c16, k2 = 4 lg + e / 2
            const int k2 = 4 * lg + (e >> 1);
            float s2, c2;
            sincospif((float)((k2 * c16) & 15) / 8.0f, &s2, &c2);
            e2r[e] = f32_to_h(part ? -s2 : c2);
            e2i[e] = f32_to_h(part ? c2 : s2);
        }
        auto pk = [](const unsigned short *e) {
            return make_uint4(e[0] | ((unsigned)e[1] << 16), e[2] | ((unsigned)e[3] << 16),
                              e[4] | ((unsigned)e[5] << 16), e[6] | ((unsigned)e[7] << 16));
        };
        pcst[0][lane] = pk(e1r);
        pcst[1][lane] = pk(e1i);
        pcst[2][lane] = pk(e2r);
        pcst[3][lane] = pk(e2i);
        // twiddle e^{+2 pi i k2 t1 / 256} of the stage-1 outputs (k2 = 4 lg + v, t1 = c16)
        float twc[4], tws[4];
#pragma unroll
        for (int v = 0; v < 4; ++v)
            sincospif((float)(((4 * lg + v) * c16) & 255) / 128.0f, &tws[v], &twc[v]);
        ptw[0][lane] = make_float4(twc[0], twc[1], twc[2], twc[3]);
        ptw[1][lane] = make_float4(tws[0], tws[1], tws[2], tws[3]);
    }
    __syncthreads();

    // ---- search state: running maximum per (reference, mirror), index in the low bits
    const unsigned lowmask = (1u << nbits) - 1u;
    const int code_top = 4 * S - 1;
    // rmA: the particle of the block's first row; rmB: the next particle, whose
    // rows may start inside the block (S_pad >= 16: at most two per block)
    float rmA[kQW][2], rmB[kQW][2];
#pragma unroll
    for (int q = 0; q < kQW; ++q) rmA[q][0] = rmA[q][1] = rmB[q][0] = rmB[q][1] = -INFINITY;

    uint4 za[16 * kQW];
#pragma unroll
    for (int i = 0; i < 16 * kQW; ++i) za[i] = make_uint4(0u, 0u, 0u, 0u);

    unsigned *zc32 = reinterpret_cast<unsigned *>(&zc[0][0][0][0]);
    unsigned *zl32 = reinterpret_cast<unsigned *>(&zl[0][0][0]);

    // fragment pipeline: a ring of kRing slots.  An entry is refilled with the slot
    // kRing positions later (wrapping into the next row block) right after its
    // contraction, so kRing - 1 to kRing slots of operands are in flight.
    Frag fr[kRing];
    auto slot_of = [](int i, int wv) { return 32 * (i >> 2) + 4 * wv + (i & 3); };   // i-th of 16
#pragma unroll
    for (int s0 = 0; s0 < kRing; ++s0) {
        const unsigned k = slot_of(s0, w);
        load_frag(fr[s0], rsp, k * g.slot_bytes, rsr, k * g.slot_bytes, off16, off8, g);
    }

    for (int jb = 0; jb < S_pad; ++jb) {
        const size_t rb = (size_t)pg * S_pad + jb;
        const unsigned pblk = (unsigned)jb * blk_bytes;
        const int par = jb & 1;
        // opaque copies: what the slots and lanes select is recomputed per row block
        // instead of being kept live across the loop (16 slots x lane values)
        int wv = w, ln = lane;
        asm volatile("" : "+s"(wv));
        asm volatile("" : "+v"(ln));
        const int col = ln & 31, nl = col >> 1, uv = col & 1, hw = ln >> 5;
        const int wo = nl / kQW, qo = nl % kQW;
        // every row this lane writes has (row >> 1) & 1 == hw: one swizzle per lane
        const int sw = wo | (hw << 3);
        // lane part of the LDS addresses (u32 units); rows add compile-time offsets
        const int zc_lane = ((wo * 16 + 2 * hw) * 64) * kQW + qo;
        const int zl_lane = ((wo * (16 * kQW)) + 2 * hw * kQW + qo) * 4;
#pragma unroll
        for (int c = 0; c < 4; ++c) {
            const int buf = c & 1;
            // ---- produce: slots 32 c + 4 w + s
#pragma unroll
            for (int s = 0; s < 4; ++s) {
                const int i = 4 * c + s;
                const int k = slot_of(i, wv);
                const float16_t acc = contract<BF16>(fr[i % kRing], rsp, pblk + k * g.slot_bytes,
                                                     rsr, k * g.slot_bytes, off16, g);
                // refill this ring entry: slot i + kRing of this row block or the next
                {
                    const bool wrap = (i + kRing >= 16);
                    if (!wrap || jb + 1 < S_pad) {
                        const unsigned pn = wrap ? pblk + blk_bytes : pblk;
                        const unsigned kn = slot_of((i + kRing) & 15, wv);
                        load_frag(fr[i % kRing], rsp, pn + kn * g.slot_bytes, rsr,
                                  kn * g.slot_bytes, off16, off8, g);
                    }
                }
                // consumer position of this lane's Z values (uniform over half the lanes)
                const int k1 = k >> 4, k2 = k & 15;
                int lp = 0, late = -1;
                if (uv == 0) lp = k2 + 16 * (k1 & 1);
                else if (k2 != 0) lp = (16 - k2) + 16 * (2 + ((15 - k1) & 1));
                else if (k == 0) late = 3;                 // Z[128]
                else if ((k1 & 1) == 0) late = k1 / 2 - 1; // Z[256 - 16 k1], k1 even
                else lp = 48;                              // Z[256 - 16 k1], k1 odd
                unsigned *dst = (late >= 0)
                    ? zl32 + par * (kGemmWaves * 16 * kQW * 4) + zl_lane + late
                    : zc32 + buf * (kGemmWaves * 16 * 64 * kQW) + zc_lane + (lp ^ sw) * kQW;
                const int rstride = (late >= 0) ? kQW * 4 : 64 * kQW;   // u32 per row
#pragma unroll
                for (int q = 0; q < 8; ++q) {
                    const int rq = 4 * (q >> 1) + (q & 1);   // row - 2 hw
                    const float x0 = acc[2 * q], x1 = acc[2 * q + 1];   // (a.col, b.col)
                    float zr, zi;
                    if (uv == 0) {
                        zr = x0;
                        zi = (k == 0) ? x0 : x1;
                    } else {
                        zr = x1;
                        zi = (k == 0) ? x1 : x0;
                    }
                    dst[rq * rstride] = pack_h2(zr, zi);
                }
            }
            __syncthreads();
            // ---- consume: register c of this wavefront's pairs (row r, reference q)
#pragma unroll
            for (int r = 0; r < 16; ++r) {
                const uint2 v = zc[buf][w][r][lane ^ zswz(w, r)];
                const unsigned vv[2] = {v.x, v.y};
#pragma unroll
                for (int q = 0; q < kQW; ++q) {
                    if (c == 0) za[kQW * r + q].x = vv[q];
                    if (c == 1) za[kQW * r + q].y = vv[q];
                    if (c == 2) za[kQW * r + q].z = vv[q];
                    if (c == 3) za[kQW * r + q].w = vv[q];
                }
            }
        }
        if (lane == 32) {
#pragma unroll
            for (int i = 0; i < 16 * kQW; ++i) za[i] = zl[par][w][i];
        }

        // ---- psi transform and search, row by row
        const uint4 b1r = pcst[0][lane], b1i = pcst[1][lane];
        const uint4 a2r = pcst[2][lane], a2i = pcst[3][lane];
        const float4 twc4 = ptw[0][lane], tws4 = ptw[1][lane];
        const float twc[4] = {twc4.x, twc4.y, twc4.z, twc4.w};
        const float tws[4] = {tws4.x, tws4.y, tws4.z, tws4.w};
        const long long row0 = (long long)rb * 16;
        const int p_first = (int)(row0 / S_pad);
        const int sg0 = (int)(row0 - (long long)p_first * S_pad);
#pragma unroll
        for (int r = 0; r < 16; ++r) {
            // one row at a time: the scheduler must not interleave rows (their
            // temporaries on top of the pair registers would spill)
            __builtin_amdgcn_sched_barrier(0);
            // S_pad >= 16: the block's rows span at most two particles
            const bool second = sg0 + r >= S_pad;
            const int p = p_first + (second ? 1 : 0);
            const int sg = sg0 + r - (second ? S_pad : 0);
            if (p < P && sg < S) {
                const float4_t z4 = {0.f, 0.f, 0.f, 0.f};
#pragma unroll
                for (int q = 0; q < kQW; ++q) {
                    const uint4 a1 = za[kQW * r + q];
                    const float4_t yr = mma16x32(a1, b1r, z4);
                    const float4_t yi = mma16x32(a1, b1i, z4);
                    unsigned bb[4];
#pragma unroll
                    for (int v = 0; v < 4; ++v) {
                        const float tr = yr[v] * twc[v] - yi[v] * tws[v];
                        const float ti = yr[v] * tws[v] + yi[v] * twc[v];
                        bb[v] = pack_h2(tr, ti);
                    }
                    const uint4 b2 = make_uint4(bb[0], bb[1], bb[2], bb[3]);
                    const float4_t zr = mma16x32(a2r, b2, z4);
                    const float4_t zi = mma16x32(a2i, b2, z4);
#pragma unroll
                    for (int m = 0; m < 2; ++m) {
                        const float4_t x = m ? zi : zr;
                        float e[4];
#pragma unroll
                        for (int v = 0; v < 4; ++v) {
                            const unsigned code = (unsigned)(code_top - (4 * sg + v));
                            e[v] = __uint_as_float((__float_as_uint(x[v]) & ~lowmask) | code);
                        }
                        const float t3 = vmax3(e[0], e[1], e[2]);
                        if (second) rmB[q][m] = vmax3(rmB[q][m], t3, e[3]);
                        else rmA[q][m] = vmax3(rmA[q][m], t3, e[3]);
                    }
                }
            }
        }
        // particle p_first complete in this block: decode its records, shift the state
        const long long last_row = (long long)(p_first + 1) * S_pad - 1;
        if (last_row < row0 + 16) {
            const int p = p_first;
#pragma unroll
            for (int q = 0; q < kQW; ++q) {
                const int n = nb * 16 + kQW * w + q;
                float bv[2];
                int bt[2], bs[2];
#pragma unroll
                for (int m = 0; m < 2; ++m) {
                    float v = rmA[q][m];
                    const unsigned u = __float_as_uint(v);
                    const int idx = code_top - (int)(u & lowmask);
                    int t = (lane & 15) + 64 * (lane >> 4) + 16 * (idx & 3);
                    int s_ = idx >> 2;
                    if (!(v > -INFINITY)) {   // no finite value in any sample
                        t = INT_MAX;
                        s_ = 0;
                    }
#pragma unroll
                    for (int off = 32; off >= 1; off >>= 1) {
                        const float ov = __shfl_xor(v, off, kWave);
                        const int ot = __shfl_xor(t, off, kWave);
                        const int os = __shfl_xor(s_, off, kWave);
                        if (ov > v || (ov == v && ot < t)) {
                            v = ov;
                            t = ot;
                            s_ = os;
                        }
                    }
                    bv[m] = __uint_as_float(__float_as_uint(v) & ~lowmask);
                    bt[m] = (t == INT_MAX) ? 0 : t;
                    bs[m] = min(max(s_, 0), S - 1);
                    rmA[q][m] = rmB[q][m];
                    rmB[q][m] = -INFINITY;
                }
                if (p < P && n < N && lane == 0) {
                    const int m = (bv[1] > bv[0]) ? 1 : 0;
                    const float val = bv[m] * rinv[n];
                    PeakRec o;
                    o.val = val;
                    o.m = m;
                    o.s = bs[m];
                    o.t = bt[m];
#pragma unroll
                    for (int e = 0; e < 7; ++e) o.x[e] = val;
                    o.pad = 0;
                    rec[(size_t)p * N + n] = o;
                }
            }
        }
    }
}

// ------------------------------------- search kernel, reference-resident tiles
// k_mref_rt (rings <= 48): workgroup = 16 particles x 8 references, eight
// wavefronts.  Wavefront w holds the operand fragments of the 8 references for
// its 16 slots (k = 8 i + w) in registers for the whole workgroup, so per block
// of 16 (particle, shift) rows only the particle operands are read.  Per block:
//  * contraction: per slot two 16 x 16 tiles (rows = 8 particle rows x {a, b},
//    columns = 8 references x {u, v}) on v_mfma_f32_16x16x32 over rings 0..31
//    and v_mfma_f32_16x16x16 over a further 16-ring step and / or the 8-ring
//    tail; particle fragments stream through a ring of kRtRing slots;
//  * every Z value of the block (16 rows x 8 references x 256 bins, FP16 re/im,
//    128 KB) goes to LDS at its psi operand position, so the psi transform of a
//    pair loads its whole A operand with one ds_read_b128 per lane;
//  * psi transform and search as in k_mref_gemm, wavefront w on reference w.
// The v part of slot k is stored as (a.v, b.v) = (Im, Re) of Z[256 - k]; the
// constant stage-1 matrix swaps re / im for the lane groups that hold those
// bins (2 and 3), so no lane selects between its two products.
constexpr int kRtWaves = 8;
constexpr int kRtThreads = 64 * kRtWaves;
constexpr int kRtRefs = 8;
constexpr int kRtRing = 4;   // divides the 16 slots of a wavefront: entry i % kRtRing across blocks
constexpr int kRtMaxRings = 48;   // one 16x16x32 step plus one 16-ring or 8-ring 16x16x16 step

struct RtFrag {   // one slot's operand fragment of 8 rows (or references) x {parts}
    uint4 m;      // 16x16x32 step, rings 0..31
    uint2 a;      // 16x16x16 step over 16-ring step 2 * NX32
    uint2 b;      // 16x16x16 step over the 8-ring tail (lane groups 0, 1)
};

template <bool BF16>
__device__ __forceinline__ float4_t mmag32(uint4 a, uint4 b, float4_t c) {
    if (BF16)
        return __builtin_amdgcn_mfma_f32_16x16x32_bf16(__builtin_bit_cast(bf16x8_t, a),
                                                       __builtin_bit_cast(bf16x8_t, b), c, 0, 0, 0);
    return __builtin_amdgcn_mfma_f32_16x16x32_f16(__builtin_bit_cast(half8_t, a),
                                                  __builtin_bit_cast(half8_t, b), c, 0, 0, 0);
}

template <bool BF16>
__device__ __forceinline__ float4_t mmag16(uint2 a, uint2 b, float4_t c) {
    if (BF16)
        return __builtin_amdgcn_mfma_f32_16x16x16bf16_1k(__builtin_bit_cast(short4_t, a),
                                                         __builtin_bit_cast(short4_t, b), c, 0, 0, 0);
    return __builtin_amdgcn_mfma_f32_16x16x16f16(__builtin_bit_cast(half4_t, a),
                                                 __builtin_bit_cast(half4_t, b), c, 0, 0, 0);
}

// psi operand position of bin kk (1..255, or 128): lane k2 + 16 lg, register j
// with k1_of(lg, j) = kk >> 4; returned as 4 * lane + j
__device__ __forceinline__ int psi_pos(int kk) {
    const int k1 = kk >> 4, k2 = kk & 15;
    const int lg = (k1 < 8) ? (k1 & 1) : 2 + (k1 & 1);
    const int j = (k1 < 8) ? (k1 >> 1) : ((14 + (k1 & 1) - k1) >> 1);
    return 4 * (k2 + 16 * lg) + j;
}

template <bool BF16, int KP>
__global__ __launch_bounds__(kRtThreads, 1) void k_mref_rt(
    const unsigned char *__restrict__ pfrag, const unsigned char *__restrict__ rfrag,
    const float *__restrict__ rinv, int P, int S, int S_pad, int N, int nbits,
    PeakRec *__restrict__ rec) {
    // operand geometry from the padded ring count KP (8, 16, ..., 48)
    constexpr int FULL16 = KP / 16, TAIL = KP % 16;                 // 16-ring steps, 8-ring tail
    constexpr int NKS = FULL16;                                      // KP = 16 NKS + TAIL
    constexpr int NX32 = NKS >= 2 ? 1 : 0, X16A = NKS & 1, X16B = TAIL ? 1 : 0;
    constexpr unsigned SB = 4 * KP, RB = kSlots * SB, BB = 16 * RB;  // slot, row, block bytes
    // Z of the block: [reference][row][psi lane ^ reference] -> 4 registers of FP16 (re, im)
    __shared__ uint4 zs[kRtRefs * 16 * 64];
    __shared__ uint4 pcst[4][64];     // stage-1 B (re, im outputs), stage-2 A (re, im)
    __shared__ uint2 ptw[4][64];      // twiddle (c, s) and (-s, c) as FP16 pairs, per v
    const int lane = threadIdx.x & 63;
    const int w = __builtin_amdgcn_readfirstlane(threadIdx.x >> 6);
    const int lb = xcd_remap(blockIdx.x, gridDim.x);
    const int NB = (N + kRtRefs - 1) / kRtRefs;
    const int pg = lb / NB, nb = lb % NB;
    const __amdgpu_buffer_rsrc_t rsp = byte_rsrc(pfrag + (size_t)pg * S_pad * BB,
                                                 (unsigned)S_pad * BB);
    const __amdgpu_buffer_rsrc_t rsr = byte_rsrc(rfrag + (size_t)nb * kRtRefs * RB, kRtRefs * RB);
    // operand lanes: row (or reference) idx = (l % 16) / 2, part l % 2, K group gq = l / 16;
    // the wavefront's slot offset w * SB is folded into the lane offsets
    const int gq = lane >> 4, idx = (lane & 15) >> 1, part = lane & 1;
    const unsigned vb = idx * RB + w * SB;
    const unsigned o32 = vb + (gq >> 1) * 64 + part * 32 + (gq & 1) * 16;
    const unsigned o16a = vb + (2 * NX32) * 64 + part * 32 + (gq >> 1) * 16 + (gq & 1) * 8;
    const unsigned o16b = vb + FULL16 * 64 + part * 16 + (gq & 1) * 8;

    if (w == 0) {
        const int c16 = lane & 15, lg = lane >> 4;
        unsigned short e1r[8], e1i[8], e2r[8], e2i[8];
#pragma unroll
        for (int e = 0; e < 8; ++e) {
            // lane groups 2, 3 hold (Im, Re) pairs: their K rows of B1 are swapped
            const int p1 = (e & 1) ^ (lg >= 2 ? 1 : 0);
            const int k1 = k1_of(lg, e >> 1);
            float s1, c1;
            sincospif((float)((k1 * c16) & 15) / 8.0f, &s1, &c1);
            e1r[e] = f32_to_h(p1 ? -s1 : c1);
            e1i[e] = f32_to_h(p1 ? c1 : s1);
            const int p2 = e & 1;
            const int k2 = 4 * lg + (e >> 1);
            float s2, c2;
            sincospif((float)((k2 * c16) & 15) / 8.0f, &s2, &c2);
            e2r[e] = f32_to_h(p2 ? -s2 : c2);
            e2i[e] = f32_to_h(p2 ? c2 : s2);
        }
        auto pk = [](const unsigned short *e) {
            return make_uint4(e[0] | ((unsigned)e[1] << 16), e[2] | ((unsigned)e[3] << 16),
                              e[4] | ((unsigned)e[5] << 16), e[6] | ((unsigned)e[7] << 16));
        };
        pcst[0][lane] = pk(e1r);
        pcst[1][lane] = pk(e1i);
        pcst[2][lane] = pk(e2r);
        pcst[3][lane] = pk(e2i);
#pragma unroll
        for (int v = 0; v < 4; ++v) {
            float ts, tc;
            sincospif((float)(((4 * lg + v) * c16) & 255) / 128.0f, &ts, &tc);
            ptw[v][lane] = make_uint2(pack_h2(tc, ts), pack_h2(-ts, tc));
        }
    }

    // resident reference fragments of this wavefront's slots k = 8 i + w; the tail
    // step's K rows 8..15 (lane groups 2, 3) do not exist and stay zero, which also
    // cancels whatever the particle operand holds there
    RtFrag rf[16];
#pragma unroll
    for (int i = 0; i < 16; ++i) {
        if (NX32) rf[i].m = bload16(rsr, o32, 8 * i * SB);
        if (X16A) rf[i].a = bload8(rsr, o16a, 8 * i * SB);
        if (X16B) {
            const uint2 t = bload8(rsr, o16b, 8 * i * SB);
            rf[i].b = (gq < 2) ? t : make_uint2(0u, 0u);
        }
    }

    // LDS position (u32) of this lane's Z values of slot i: Z[k] (u lanes) or
    // Z[256 - k] (v lanes) of row 2 gq (+ 8 T + pp), reference q, at the psi lane
    // swizzled by the reference
    const int q = (lane & 15) >> 1, uv = lane & 1;
    unsigned dz[16];
#pragma unroll
    for (int i = 0; i < 16; ++i) {
        const int k = 8 * i + w;
        const int pos = uv ? psi_pos(k == 0 ? 128 : 256 - k) : psi_pos(k);
        dz[i] = (unsigned)(((q * 16 + 2 * gq) * 64) * 4 + (pos ^ (q << 2))) * 4u;
    }

    // particle fragments of one slot for the block's two 8-row tiles
    struct PF {
        RtFrag t[2];
    };
    auto pload = [&](PF &f, unsigned blk_off, int i) {
#pragma unroll
        for (int T = 0; T < 2; ++T) {
            const unsigned so = blk_off + 8 * i * SB + T * 8 * RB;
            if (NX32) f.t[T].m = bload16(rsp, o32, so);
            if (X16A) f.t[T].a = bload8(rsp, o16a, so);
            if (X16B) f.t[T].b = bload8(rsp, o16b, so);
        }
    };
    PF pf[kRtRing];
#pragma unroll
    for (int s0 = 0; s0 < kRtRing; ++s0) pload(pf[s0], 0u, s0);

    __syncthreads();   // the constant tables
    const uint4 b1r = pcst[0][lane], b1i = pcst[1][lane];
    const uint4 a2r = pcst[2][lane], a2i = pcst[3][lane];
    uint2 tw[4];
#pragma unroll
    for (int v = 0; v < 4; ++v) tw[v] = ptw[v][lane];

    // search state of reference nb * 8 + w: running maxima per mirror of the open
    // particle; the index code sits in the low nbits of each value
    const int code_top = 4 * S - 1;
    unsigned keep = ~((1u << nbits) - 1u);
    asm volatile("" : "+v"(keep));   // a VGPR operand: one bit-select per value
    float rm0 = -INFINITY, rm1 = -INFINITY;
    const int n_ref = nb * kRtRefs + w;
    const float rsc = (n_ref < N) ? rinv[n_ref] : 0.0f;
    int p_open = 16 * pg;                                        // first particle of the group
    int sg_open = 0;                                             // shift row of block row 0
    char *zsb = reinterpret_cast<char *>(zs);
    const uint4 *zp = zs + w * 16 * 64 + (lane ^ w);   // reference w, row 0, this lane

    // wave reduction of the open particle's maxima -> its record
    auto flush = [&](int p) {
        float bv[2];
        int bt[2], bs[2];
#pragma unroll
        for (int m = 0; m < 2; ++m) {
            float v = m ? rm1 : rm0;
            const unsigned u = __float_as_uint(v);
            const int ix = code_top - (int)(u & ~keep);
            int t = (lane & 15) + 64 * (lane >> 4) + 16 * (ix & 3);
            int s_ = ix >> 2;
            if (!(v > -INFINITY)) {   // no finite value in any sample
                t = INT_MAX;
                s_ = 0;
            }
#pragma unroll
            for (int off = 32; off >= 1; off >>= 1) {
                const float ov = __shfl_xor(v, off, kWave);
                const int ot = __shfl_xor(t, off, kWave);
                const int os = __shfl_xor(s_, off, kWave);
                if (ov > v || (ov == v && ot < t)) {
                    v = ov;
                    t = ot;
                    s_ = os;
                }
            }
            bv[m] = __uint_as_float(__float_as_uint(v) & keep);
            bt[m] = (t == INT_MAX) ? 0 : t;
            bs[m] = min(max(s_, 0), S - 1);
        }
        if (p < P && n_ref < N && lane == 0) {
            const int m = (bv[1] > bv[0]) ? 1 : 0;
            const float val = bv[m] * rsc;
            PeakRec o;
            o.val = val;
            o.m = m;
            o.s = bs[m];
            o.t = bt[m];
#pragma unroll
            for (int e = 0; e < 7; ++e) o.x[e] = val;
            o.pad = 0;
            rec[(size_t)p * N + n_ref] = o;
        }
        rm0 = rm1 = -INFINITY;
    };

    // psi transform of row r (reference w) and the running maxima
    auto psi_row = [&](int r, int sg) {
        const float4_t z4 = {0.f, 0.f, 0.f, 0.f};
        const uint4 a1 = zp[r * 64];
        const float4_t yr = mma16x32(a1, b1r, z4);
        const float4_t yi = mma16x32(a1, b1i, z4);
        unsigned bb[4];
#pragma unroll
        for (int v = 0; v < 4; ++v) {
            // (yr + i yi) e^{+2 pi i k2 t1 / 256} in FP16 pairs: yr (c, s) + yi (-s, c)
            const h2_t y0 = {(_Float16)yr[v], (_Float16)yr[v]};
            const h2_t y1 = {(_Float16)yi[v], (_Float16)yi[v]};
            const h2_t o = y0 * __builtin_bit_cast(h2_t, tw[v].x) +
                           y1 * __builtin_bit_cast(h2_t, tw[v].y);
            bb[v] = __builtin_bit_cast(unsigned, o);
        }
        const uint4 b2 = make_uint4(bb[0], bb[1], bb[2], bb[3]);
        const float4_t zr = mma16x32(a2r, b2, z4);
        const float4_t zi = mma16x32(a2i, b2, z4);
        const unsigned c0 = (unsigned)(code_top - 4 * sg);
        float e[2][4];
#pragma unroll
        for (int v = 0; v < 4; ++v) {
            const unsigned cv = c0 - v;
            e[0][v] = __uint_as_float((__float_as_uint(zr[v]) & keep) | (cv & ~keep));
            e[1][v] = __uint_as_float((__float_as_uint(zi[v]) & keep) | (cv & ~keep));
        }
        rm0 = vmax3(rm0, vmax3(e[0][0], e[0][1], e[0][2]), e[0][3]);
        rm1 = vmax3(rm1, vmax3(e[1][0], e[1][1], e[1][2]), e[1][3]);
    };

    // rows r0 .. r0 + n - 1 with shift rows sg0 .. : two rows per step, so their
    // MFMA -> VALU -> MFMA chains interleave
    auto psi_rows = [&](int r0, int n, int sg0) {
        int r = 0;
        for (; r + 2 <= n; r += 2) {
            psi_row(r0 + r, sg0 + r);
            psi_row(r0 + r + 1, sg0 + r + 1);
        }
        if (r < n) psi_row(r0 + r, sg0 + r);
    };

    for (int jb = 0; jb < S_pad; ++jb) {
        const unsigned pblk = (unsigned)jb * BB;
        // ---- contraction: this wavefront's 16 slots for both 8-row tiles
#pragma unroll
        for (int i = 0; i < 16; ++i) {
            const int s = i % kRtRing;
            // every step from a zero accumulator, summed on the VALU: no MFMA reads
            // another one's result as its accumulator input (chained 16x16x32 and
            // 16x16x16 accumulators, with the tiles' registers interleaved by the
            // scheduler, gave wrong sums)
            const float4_t z0 = {0.f, 0.f, 0.f, 0.f};
            float4_t acc[2];
#pragma unroll
            for (int T = 0; T < 2; ++T) {
                float4_t a = z0;
                if (NX32) a = mmag32<BF16>(pf[s].t[T].m, rf[i].m, z0);
                if (X16A) {
                    const float4_t t = mmag16<BF16>(pf[s].t[T].a, rf[i].a, z0);
                    a = NX32 ? a + t : t;
                }
                if (X16B) {
                    const float4_t t = mmag16<BF16>(pf[s].t[T].b, rf[i].b, z0);
                    a = (NX32 || X16A) ? a + t : t;
                }
                acc[T] = a;
            }
            if (i + kRtRing < 16) pload(pf[s], pblk, i + kRtRing);
            else if (jb + 1 < S_pad) pload(pf[s], pblk + BB, i + kRtRing - 16);
#pragma unroll
            for (int T = 0; T < 2; ++T)
#pragma unroll
                for (int pp = 0; pp < 2; ++pp)
                    *reinterpret_cast<unsigned *>(zsb + dz[i] + (8 * T + pp) * 1024) =
                        pack_h2(acc[T][2 * pp], acc[T][2 * pp + 1]);
            if (i == 0 && w == 0) {
                // slot 0 packs the real bins: Z[0] = (a.u)(1 + i) (u lanes) and
                // Z[128] = (b.v)(1 + i) (v lanes)
#pragma unroll
                for (int T = 0; T < 2; ++T)
#pragma unroll
                    for (int pp = 0; pp < 2; ++pp) {
                        const float x = uv ? acc[T][2 * pp + 1] : acc[T][2 * pp];
                        *reinterpret_cast<unsigned *>(zsb + dz[0] + (8 * T + pp) * 1024) =
                            pack_h2(x, x);
                    }
            }
        }
        __syncthreads();

        // ---- psi transform and search: reference w, rows 0..15.  S_pad >= 16:
        // the rows split at most once between the open particle and the next
        const int split = min(16, S_pad - sg_open);
        const int rows1 = (p_open < P) ? min(split, max(S - sg_open, 0)) : 0;
        psi_rows(0, rows1, sg_open);
        if (split < 16) {
            flush(p_open);
            ++p_open;
            const int rows2 = (p_open < P) ? min(16 - split, S) : 0;
            psi_rows(split, rows2, 0);
            sg_open = 16 - split;
        } else {
            sg_open += 16;
            if (sg_open == S_pad) {
                flush(p_open);
                ++p_open;
                sg_open = 0;
            }
        }
        __syncthreads();   // the next block's contraction overwrites zs
    }
}


}  // namespace

GemmLayout gemm_layout(int n_ring) {
    GemmLayout L;
    L.full16 = n_ring / 16;
    const int rem = n_ring % 16;
    L.tail = rem == 0 ? 0 : (rem <= 8 ? 8 : 16);
    L.k_pad = 16 * L.full16 + L.tail;
    L.slot_bytes = 4 * L.k_pad;
    L.row_bytes = (size_t)kSlots * L.slot_bytes;
    L.block_bytes = 16 * L.row_bytes;
    // |Z| <= c_p 256 sqrt(R (R + 1)) * 1.42 must stay below 32768
    const double bound = 256.0 * sqrt((double)n_ring * (n_ring + 1)) * 1.42;
    L.cp = ldexpf(1.0f, (int)floor(log2(32768.0 / bound)));
    return L;
}

int gemm_index_bits(int n_shift) {
    int b = 0;
    while ((1 << b) < n_shift) ++b;
    return 2 + b;
}

int gemm_rows_per_particle(int n_shift) { return n_shift < 16 ? 16 : n_shift; }

size_t gemm_particle_operand_bytes(int n_img, int n_shift, int n_ring) {
    const size_t rows = (size_t)((n_img + 15) / 16) * 16 * gemm_rows_per_particle(n_shift);
    return rows / 16 * gemm_layout(n_ring).block_bytes;
}

hipError_t launch_frag_rows(const float2 *spec, const float2 *norm, int n_img, int n_shift,
                            int n_ring, bool bf16, void *frag, hipStream_t st) {
    if (n_img <= 0) return hipSuccess;
    const GemmLayout L = gemm_layout(n_ring);
    const GemmGeom g{L.full16, L.tail, L.slot_bytes};
    const int S_pad = gemm_rows_per_particle(n_shift);
    const int n_rows_pad = (n_img + 15) / 16 * 16 * S_pad;
    const long long threads = (long long)n_rows_pad * kSlots;
    const dim3 grid((unsigned)((threads + 255) / 256));
    if (bf16)
        hipLaunchKernelGGL((k_frag_rows<true>), grid, dim3(256), 0, st, spec, norm, n_img, n_shift,
                           S_pad, n_rows_pad, n_ring, g, L.cp, (unsigned char *)frag);
    else
        hipLaunchKernelGGL((k_frag_rows<false>), grid, dim3(256), 0, st, spec, norm, n_img,
                           n_shift, S_pad, n_rows_pad, n_ring, g, L.cp, (unsigned char *)frag);
    return hipGetLastError();
}

hipError_t launch_frag_refs(const float2 *rspec, int n_ref, int n_ring, bool bf16, float *cn,
                            float *rinv, void *frag, hipStream_t st) {
    if (n_ref <= 0) return hipSuccess;
    const GemmLayout L = gemm_layout(n_ring);
    const GemmGeom g{L.full16, L.tail, L.slot_bytes};
    hipLaunchKernelGGL(k_ref_scale, dim3(n_ref), dim3(256), 0, st, rspec, n_ring, L.cp, cn, rinv);
    const int n_pad = (n_ref + 15) / 16 * 16;
    const dim3 grid((unsigned)((n_pad * kSlots + 255) / 256));
    if (bf16)
        hipLaunchKernelGGL((k_frag_refs<true>), grid, dim3(256), 0, st, rspec, n_ref, n_pad,
                           n_ring, g, cn, (unsigned char *)frag);
    else
        hipLaunchKernelGGL((k_frag_refs<false>), grid, dim3(256), 0, st, rspec, n_ref, n_pad,
                           n_ring, g, cn, (unsigned char *)frag);
    return hipGetLastError();
}

hipError_t launch_mref_gemm(const void *pfrag, const void *rfrag, const float *rinv, int n_img,
                            int n_shift, int n_ring, int n_ref, bool bf16, PeakRec *rec,
                            hipStream_t st) {
    if (n_img <= 0 || n_ref <= 0) return hipSuccess;
    const GemmLayout L = gemm_layout(n_ring);
    const GemmGeom g{L.full16, L.tail, L.slot_bytes};
    const int nbits = gemm_index_bits(n_shift);
    if (nbits > 12) return hipErrorInvalidValue;   // more than 1024 shifts
    const int S_pad = gemm_rows_per_particle(n_shift);
    if (n_ring <= kRtMaxRings) {
        // reference-resident tiles; k_pad = 8, 16, ..., 48 selects the instantiation
        const dim3 grid((unsigned)(((n_img + 15) / 16) * ((n_ref + kRtRefs - 1) / kRtRefs)));
        const unsigned char *pf = (const unsigned char *)pfrag;
        const unsigned char *rf = (const unsigned char *)rfrag;
#define ISAC_RT_LAUNCH(B, KP)                                                                   \
    hipLaunchKernelGGL((k_mref_rt<B, KP>), grid, dim3(kRtThreads), 0, st, pf, rf, rinv, n_img,  \
                       n_shift, S_pad, n_ref, nbits, rec)
        switch (L.k_pad * 2 + (bf16 ? 1 : 0)) {
        case 16: ISAC_RT_LAUNCH(false, 8); break;
        case 17: ISAC_RT_LAUNCH(true, 8); break;
        case 32: ISAC_RT_LAUNCH(false, 16); break;
        case 33: ISAC_RT_LAUNCH(true, 16); break;
        case 48: ISAC_RT_LAUNCH(false, 24); break;
        case 49: ISAC_RT_LAUNCH(true, 24); break;
        case 64: ISAC_RT_LAUNCH(false, 32); break;
        case 65: ISAC_RT_LAUNCH(true, 32); break;
        case 80: ISAC_RT_LAUNCH(false, 40); break;
        case 81: ISAC_RT_LAUNCH(true, 40); break;
        case 96: ISAC_RT_LAUNCH(false, 48); break;
        case 97: ISAC_RT_LAUNCH(true, 48); break;
        default: return hipErrorInvalidValue;
        }
#undef ISAC_RT_LAUNCH
        return hipGetLastError();
    }
    const dim3 grid((unsigned)(((n_img + 15) / 16) * ((n_ref + 15) / 16)));
    if (bf16)
        hipLaunchKernelGGL((k_mref_gemm<true>), grid, dim3(kGemmThreads), 0, st,
                           (const unsigned char *)pfrag, (const unsigned char *)rfrag, rinv, n_img,
                           n_shift, S_pad, n_ref, g, nbits, rec);
    else
        hipLaunchKernelGGL((k_mref_gemm<false>), grid, dim3(kGemmThreads), 0, st,
                           (const unsigned char *)pfrag, (const unsigned char *)rfrag, rinv, n_img,
                           n_shift, S_pad, n_ref, g, nbits, rec);
    return hipGetLastError();
}

}  // namespace isac
