// HIP kernels of the polar-ring 2D alignment path for CDNA4 (gfx950, wave64).
//
// Data layout in HBM:
//   images       [n][D][D] float32, row-major (y, x)
//   spec         [n][S][R][128] float2: raw ring spectra of every image at every
//                grid shift, packed: slot 0 = (X[0], X[128]), slot k = X[k], k=1..127
//   norm         [n][S] float2 = (mu*L, 1/(sd+1e-7)) of the radius-weighted
//                normalisation (reference cuda/gpu_aln_noref.cu:1294-1351)
//   rspec        [N][R][128] float2: reference spectra, normalised and weighted (i+1)
//   rec          PeakRec per (particle, reference)
// The reference's psi-CCF table (particles x refs x shifts x 2 x (L+2) floats,
// cuda/gpu_aln_noref.cu:2343-2461) is never materialised: the contraction over
// rings, the psi inverse FFT of both mirrors and the first-max-wins search run
// fused in registers / LDS of one wavefront.
#include "isac_internal.h"

#include <cfloat>
#include <climits>
#include <cmath>
#include <type_traits>

namespace isac {
namespace {

constexpr int kWave = 64;

__device__ __forceinline__ void wave_lds_sync() {
    __builtin_amdgcn_fence(__ATOMIC_RELEASE, "wavefront");
    __builtin_amdgcn_wave_barrier();
    __builtin_amdgcn_fence(__ATOMIC_ACQUIRE, "wavefront");
}

__device__ __forceinline__ float2 f2(float x, float y) { return make_float2(x, y); }

__device__ __forceinline__ float2 cmul(float2 a, float2 b) {
    return f2(a.x * b.x - a.y * b.y, a.x * b.y + a.y * b.x);
}

__device__ __forceinline__ float2 shfl2(float2 v, int src) {
    return f2(__shfl(v.x, src, kWave), __shfl(v.y, src, kWave));
}

// Radix-4 butterfly: forward Y_m = sum_r v_r e^{-2 pi i rm/4}; INV uses e^{+}.
template <bool INV>
__device__ __forceinline__ void radix4(float2 (&v)[4]) {
    const float2 a0 = f2(v[0].x + v[2].x, v[0].y + v[2].y);
    const float2 a1 = f2(v[0].x - v[2].x, v[0].y - v[2].y);
    const float2 a2 = f2(v[1].x + v[3].x, v[1].y + v[3].y);
    const float2 a3 = f2(v[1].x - v[3].x, v[1].y - v[3].y);
    v[0] = f2(a0.x + a2.x, a0.y + a2.y);
    v[2] = f2(a0.x - a2.x, a0.y - a2.y);
    if (!INV) {  // a1 -/+ i a3
        v[1] = f2(a1.x + a3.y, a1.y - a3.x);
        v[3] = f2(a1.x - a3.y, a1.y + a3.x);
    } else {
        v[1] = f2(a1.x - a3.y, a1.y + a3.x);
        v[3] = f2(a1.x + a3.y, a1.y - a3.x);
    }
}

// Position of element a of a wave's 256-float2 exchange buffer: a ^ ((a >> 2) & 31).
// Every access pattern of fft256_wave (4 l + q, l + 64 q, and the two stage
// writes) then hits 32 distinct bank pairs per half-wavefront.
__device__ __forceinline__ int fswz(int a) { return a ^ ((a >> 2) & 31); }

// Twiddles of the three Stockham stages, in the order the stages read them
// (entry (q - 1) NS + k of stage NS is e^{-2 pi i q k (64 / NS) / 256}): the
// reads of a stage are consecutive or broadcast, so free of bank conflicts.
constexpr int kFftTw = 3 * (4 + 16 + 64);   // 252 entries
__device__ __forceinline__ void fill_fft_twiddles(float2 *tw, const float2 *twg, int tid,
                                                  int nthreads) {
    for (int i = tid; i < kFftTw; i += nthreads) {
        const int NS = i < 12 ? 4 : i < 60 ? 16 : 64;
        const int j = i - (NS == 4 ? 0 : NS == 16 ? 12 : 60);
        const int q = j / NS + 1, k = j % NS;
        tw[i] = twg[q * k * (64 / NS)];
    }
}

// One Stockham radix-4 stage (Govindaraju et al., SC'08) with sub-transform size NS.
template <bool INV, int NS, bool WRITE>
__device__ __forceinline__ void stockham_stage(float2 (&v)[4], const float2 *in, float2 *out,
                                               const float2 *tw, int lane) {
#pragma unroll
    for (int q = 0; q < 4; ++q) v[q] = in[fswz(lane + 64 * q)];
    const int k = lane & (NS - 1);
    constexpr int off = NS == 4 ? 0 : NS == 16 ? 12 : 60;
#pragma unroll
    for (int q = 1; q < 4; ++q) {
        float2 w = tw[off + (q - 1) * NS + k];
        if (INV) w.y = -w.y;
        v[q] = cmul(v[q], w);
    }
    radix4<INV>(v);
    if (WRITE) {
        const int base = (lane - k) * 4 + k;
#pragma unroll
        for (int q = 0; q < 4; ++q) out[fswz(base + q * NS)] = v[q];
    }
}

// 256-point complex FFT by one wavefront.  In: v[q] = x[lane + 64 q].
// Out: v[q] = X[lane + 64 q] (natural order).  b0/b1: 256-float2 wave-private
// LDS buffers (swizzled, fswz); tw: the stage twiddles of fill_fft_twiddles, in
// LDS.  Unnormalised.
template <bool INV>
__device__ __forceinline__ void fft256_wave(float2 (&v)[4], float2 *b0, float2 *b1,
                                            const float2 *tw, int lane) {
    wave_lds_sync();
    radix4<INV>(v);
#pragma unroll
    for (int q = 0; q < 4; ++q) b0[fswz(4 * lane + q)] = v[q];
    wave_lds_sync();
    stockham_stage<INV, 4, true>(v, b0, b1, tw, lane);
    wave_lds_sync();
    stockham_stage<INV, 16, true>(v, b1, b0, tw, lane);
    wave_lds_sync();
    stockham_stage<INV, 64, false>(v, b0, nullptr, tw, lane);
}

__device__ __forceinline__ float bilinear(const float *img, int D, float x, float y) {
    const float x0 = floorf(x), y0 = floorf(y);
    const float fx = x - x0, fy = y - y0;
    const int ix = (int)x0, iy = (int)y0;
    const int xa = min(max(ix, 0), D - 1), xb = min(max(ix + 1, 0), D - 1);
    const int ya = min(max(iy, 0), D - 1), yb = min(max(iy + 1, 0), D - 1);
    const float v00 = img[ya * D + xa], v01 = img[ya * D + xb];
    const float v10 = img[yb * D + xa], v11 = img[yb * D + xb];
    return (1.0f - fy) * ((1.0f - fx) * v00 + fx * v01) + fy * ((1.0f - fx) * v10 + fx * v11);
}

// The same sample from an LDS copy of the image with a one-pixel border that
// repeats the edge (row / column c of the copy holds pixel clamp(c - 1)): the
// left / upper tap is clamped to -1 .. D - 1 and both taps of a row are adjacent
// words (one ds_read2 per row), with exactly the values and weights above.
__device__ __forceinline__ float bilinear_bordered(const float *pimg, int D, float x, float y) {
    const float x0 = floorf(x), y0 = floorf(y);
    const float fx = x - x0, fy = y - y0;
    const int cx = min(max((int)x0, -1), D - 1) + 1, cy = min(max((int)y0, -1), D - 1) + 1;
    const float *q = pimg + cy * (D + 2) + cx;
    const float v00 = q[0], v01 = q[1], v10 = q[D + 2], v11 = q[D + 3];
    return (1.0f - fy) * ((1.0f - fx) * v00 + fx * v01) + fy * ((1.0f - fx) * v10 + fx * v11);
}

// Blocks are dispatched round-robin over the 8 XCDs; remap so that consecutive
// logical blocks (which share particle spectra) run on the same XCD and L2.
__device__ __forceinline__ int xcd_remap(int b, int nb) {
    const int q = nb >> 3, r = nb & 7;
    const int xcd = b & 7, idx = b >> 3;
    return (xcd < r) ? xcd * (q + 1) + idx : r * (q + 1) + (xcd - r) * q + idx;
}

// Slots lane and lane + 64 of one ring row (adjacent in memory, slot_pos): one
// 16-byte load.
__device__ __forceinline__ void load_slots(const float2 *row, int lane, float2 &lo, float2 &hi) {
    const float4 v = reinterpret_cast<const float4 *>(row)[lane];
    lo = make_float2(v.x, v.y);
    hi = make_float2(v.z, v.w);
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

// ------------------------------------------------------------------ stage A
// Block = 4 wavefronts on one image (staged once in LDS); wavefront w handles
// grid shifts w, w+4, w+8, ...
// A wavefront samples two rings at a time (ring a real part, ring b imaginary
// part), runs one 256-point FFT and unpacks both real spectra.  The normalisation
// sums (radius-weighted sum and sum of squares) accumulate in double.
// MASKED: only the shifts whose bit is set in smask[p] (S <= 64) are computed;
// the other rows of spec / norm are left as they were.
template <bool MASKED, int G = 4>
__global__ __launch_bounds__(256) void k_polar_spectra(
    const float *__restrict__ imgs, int D, ShiftView acc, const int2 *__restrict__ shift_i,
    int S, const float2 *__restrict__ tmpl, int R, const float2 *__restrict__ twg,
    float2 *__restrict__ spec, float2 *__restrict__ norm,
    const unsigned long long *__restrict__ smask) {
    extern __shared__ float4 smem4[];
    __shared__ int slist[64];
    __shared__ int scnt;
    float2 *tw = reinterpret_cast<float2 *>(smem4);
    float2 *fb = tw + 256;
    float *img = reinterpret_cast<float *>(fb + 4 * 512);
    const int tid = threadIdx.x, lane = tid & 63, wave = __builtin_amdgcn_readfirstlane(tid >> 6);
    const int p = blockIdx.x;
    if (MASKED && smask[p] == 0ull) return;   // uniform over the workgroup
    const float *src = imgs + (size_t)p * D * D;
    for (int i = tid; i < (D + 2) * (D + 2); i += 256) {   // bordered copy (bilinear_bordered)
        const int r = i / (D + 2), c = i - r * (D + 2);
        img[i] = src[min(max(r - 1, 0), D - 1) * D + min(max(c - 1, 0), D - 1)];
    }
    fill_fft_twiddles(tw, twg, tid, 256);
    if (MASKED && tid == 0) {
        const unsigned long long mk = smask[p];
        int c = 0;
        for (int b = 0; b < S && b < 64; ++b)
            if ((mk >> b) & 1ull) slist[c++] = b;
        scnt = c;
    }
    __syncthreads();
    const int n_rows = MASKED ? scnt : S;
    float ax = 0.0f, ay = 0.0f;
    if (acc.base) {
        ax = acc.base[(size_t)p * acc.stride];
        ay = acc.base[(size_t)p * acc.stride + 1];
    }
    float2 *b0 = fb + wave * 512;
    float2 *b1 = b0 + 256;
    const int pl = (64 - lane) & 63;
    // rows j0, j0 + 4, ..., j0 + 4 (G - 1) of this wavefront share each ring
    // pair's template samples (loaded once per group instead of once per row)
    for (int j0 = wave; j0 < n_rows; j0 += 4 * G) {
        int sg[G];
        float cxg[G], cyg[G];
        double sA[G], sQ[G];
#pragma unroll
        for (int g = 0; g < G; ++g) {
            const int j = min(j0 + 4 * g, n_rows - 1);
            sg[g] = MASKED ? slist[j] : j;
            const int2 si = shift_i[sg[g]];
            cxg[g] = (float)(D / 2 + si.x) + ax;
            cyg[g] = (float)(D / 2 + si.y) + ay;
            sA[g] = 0.0;
            sQ[g] = 0.0;
        }
        for (int i = 0; i < R; i += 2) {
            const bool hb = (i + 1) < R;
            const double wa = (double)(i + 1), wb = (double)(i + 2);
            float2 ta[4], tb[4];
#pragma unroll
            for (int q = 0; q < 4; ++q) {
                ta[q] = tmpl[i * kRingLen + lane + 64 * q];
                tb[q] = hb ? tmpl[(i + 1) * kRingLen + lane + 64 * q] : f2(0.0f, 0.0f);
            }
#pragma unroll
            for (int g = 0; g < G; ++g) {
                if (j0 + 4 * g >= n_rows) break;   // uniform over the wavefront
                float2 v[4];
#pragma unroll
                for (int q = 0; q < 4; ++q) {
                    const float va = bilinear_bordered(img, D, cxg[g] + ta[q].x, cyg[g] + ta[q].y);
                    const float vb = hb ? bilinear_bordered(img, D, cxg[g] + tb[q].x, cyg[g] + tb[q].y) : 0.0f;
                    v[q] = f2(va, vb);
                    sA[g] += wa * (double)va + wb * (double)vb;
                    sQ[g] += wa * (double)va * (double)va + wb * (double)vb * (double)vb;
                }
                fft256_wave<false>(v, b0, b1, tw, lane);
                // v[q] = Z[lane + 64 q].  Partners Z[256-k]: k=lane -> lane pl, q=3;
                // k=lane+64 -> lane pl, q=2 (lane 0: Z[0] itself and Z[192] = own q=3).
                const float2 u2 = shfl2(v[2], pl), u3 = shfl2(v[3], pl);
                const float2 zl = v[0], zlp = lane ? u3 : v[0];
                const float2 zh = v[1], zhp = lane ? u2 : v[3];
                float2 al = f2(0.5f * (zl.x + zlp.x), 0.5f * (zl.y - zlp.y));
                float2 bl = f2(0.5f * (zl.y + zlp.y), 0.5f * (zlp.x - zl.x));
                const float2 ah = f2(0.5f * (zh.x + zhp.x), 0.5f * (zh.y - zhp.y));
                const float2 bh = f2(0.5f * (zh.y + zhp.y), 0.5f * (zhp.x - zh.x));
                if (lane == 0) {  // pack Nyquist: A[128] = Re Z[128], B[128] = Im Z[128]
                    al = f2(zl.x, v[2].x);
                    bl = f2(zl.y, v[2].y);
                }
                // slots lane and lane + 64 are adjacent (slot_pos): one 16-byte store per ring
                float4 *oa = reinterpret_cast<float4 *>(spec + (((size_t)p * S + sg[g]) * (size_t)R + i) * kSlots);
                oa[lane] = make_float4(al.x, al.y, ah.x, ah.y);
                if (hb) oa[kSlots / 2 + lane] = make_float4(bl.x, bl.y, bh.x, bh.y);
            }
        }
#pragma unroll
        for (int g = 0; g < G; ++g) {
            if (j0 + 4 * g >= n_rows) break;
            double a2 = sA[g], q2 = sQ[g];
#pragma unroll
            for (int off = 32; off >= 1; off >>= 1) {
                a2 += __shfl_xor(a2, off, kWave);
                q2 += __shfl_xor(q2, off, kWave);
            }
            if (lane == 0) {
                const double nn = (double)kRingLen * 0.5 * (double)R * (double)(R + 1);
                const double mu = a2 / nn;
                const double sd = sqrt(fabs(q2 / nn - mu * mu));
                norm[(size_t)p * S + sg[g]] = f2((float)(mu * kRingLen), (float)(1.0 / (sd + 1e-7)));
            }
        }
    }  // row groups
}

// Stage A of the mixed-precision search: the normalised ring spectra of every
// image at every grid shift, written directly as 16-bit MFMA operand rows (the
// GEMM layout of kernels_mixed.hip): row p * S_pad + s holds, per slot, the
// parts a = g (Re S + Im S) c_p and b = g (Re S - Im S) c_p over the rings (slot
// 0: a = g (S[0] - mu L) c_p, b = g S[128] c_p), K padded with zeros.  Rows
// s >= S and images p >= n_img (padding of the last group of 16) are zero.
// Workgroup = one image, staged in LDS as (image - m) / s with its own mean and
// standard deviation (so the raw spectra of a row fit FP16 and the DC term needs
// no large cancellation); wavefront w computes rows w, w + 4, ...  A row's raw
// parts (Re + Im, Re - Im) stay in registers as packed FP16 (2 slots x R rings
// per lane) until the row's normalisation sums are complete; the exact row
// normalisation is applied in FP32 at the final conversion, and every slot is
// one contiguous store.  Sampling positions are the centre plus (i + 1)(cos,
// sin) of the lane's four angles; the normalisation sums add FP32 partial sums
// of four samples in FP64.
template <bool BF16>
__device__ __forceinline__ unsigned short frag16(float x) {
    if (BF16) {
        unsigned u = __float_as_uint(x);
        u += 0x7fffu + ((u >> 16) & 1u);
        return (unsigned short)(u >> 16);
    }
    return __builtin_bit_cast(unsigned short, (_Float16)x);
}

__device__ __forceinline__ unsigned h2pack(float lo, float hi) {
    return (unsigned)__builtin_bit_cast(unsigned short, (_Float16)lo) |
           ((unsigned)__builtin_bit_cast(unsigned short, (_Float16)hi) << 16);
}
__device__ __forceinline__ float h2lo(unsigned v) {
    return (float)__builtin_bit_cast(_Float16, (unsigned short)(v & 0xffffu));
}
__device__ __forceinline__ float h2hi(unsigned v) {
    return (float)__builtin_bit_cast(_Float16, (unsigned short)(v >> 16));
}

constexpr int kFragMaxRings = 40;
constexpr int kFragBorder = 8;   // border of k_polar_frag's image copy, pixels

template <bool BF16>
__global__ __launch_bounds__(256, 2) void k_polar_frag(
    const float *__restrict__ imgs, int n_img, int D, ShiftView acc,
    const int2 *__restrict__ shift_i, int S, int S_pad, int R, int full16, int tail,
    int slot_bytes, float cp, const float2 *__restrict__ twg, unsigned char *__restrict__ frag) {
    extern __shared__ float4 smem4[];
    __shared__ double red[2][4];
    float2 *tw = reinterpret_cast<float2 *>(smem4);
    float2 *fb = tw + 256;
    float *img = reinterpret_cast<float *>(fb + 4 * 512);
    const int tid = threadIdx.x, lane = tid & 63, wave = __builtin_amdgcn_readfirstlane(tid >> 6);
    const int p = blockIdx.x;
    const size_t row_bytes = (size_t)kSlots * slot_bytes;
    unsigned char *prow = frag + (size_t)p * S_pad * row_bytes;
    const int n_rows = (p < n_img) ? S : 0;
    // zero rows: shifts S .. S_pad - 1 (and every row of a padding image)
    {
        const size_t zb = (size_t)(S_pad - n_rows) * row_bytes;
        uint4 *z = reinterpret_cast<uint4 *>(prow + (size_t)n_rows * row_bytes);
        for (size_t i = tid; i < zb / 16; i += 256) z[i] = make_uint4(0u, 0u, 0u, 0u);
    }
    if (n_rows == 0) return;   // uniform over the workgroup
    const float *src = imgs + (size_t)p * D * D;
    // image mean and standard deviation (FP64)
    double s1 = 0.0, s2 = 0.0;
    for (int i = tid; i < D * D; i += 256) {
        const double v = src[i];
        s1 += v;
        s2 += v * v;
    }
#pragma unroll
    for (int off = 32; off >= 1; off >>= 1) {
        s1 += __shfl_xor(s1, off, kWave);
        s2 += __shfl_xor(s2, off, kWave);
    }
    if (lane == 0) {
        red[0][wave] = s1;
        red[1][wave] = s2;
    }
    __syncthreads();
    const double t1 = red[0][0] + red[0][1] + red[0][2] + red[0][3];
    const double t2 = red[1][0] + red[1][1] + red[1][2] + red[1][3];
    const double im = t1 / (double)(D * D);
    double isd = sqrt(fmax(t2 / (double)(D * D) - im * im, 0.0));
    if (!(isd > 1e-30)) isd = 1.0;
    const float mimg = (float)im, rs = (float)(1.0 / isd);
    // standardised copy with a kFragBorder-pixel border that repeats the edge
    // (cell c of a row holds pixel clamp(c - kFragBorder))
    const int W = D + 2 * kFragBorder;
    for (int i = tid; i < W * W; i += 256) {
        const int r = i / W, c = i - r * W;
        img[i] = (src[min(max(r - kFragBorder, 0), D - 1) * D + min(max(c - kFragBorder, 0), D - 1)] -
                  mimg) * rs;
    }
    fill_fft_twiddles(tw, twg, tid, 256);
    __syncthreads();
    float ax = 0.0f, ay = 0.0f;
    if (acc.base) {
        ax = acc.base[(size_t)p * acc.stride];
        ay = acc.base[(size_t)p * acc.stride + 1];
    }
    float cq[4], sq[4];
#pragma unroll
    for (int q = 0; q < 4; ++q) sincospif((float)(lane + 64 * q) / 128.0f, &sq[q], &cq[q]);
    float2 *b0 = fb + wave * 512;
    float2 *b1 = b0 + 256;
    const int pl = (64 - lane) & 63;
    const float Wf = (float)W;

    // bilinear sample at (X, Y) of the bordered copy (coordinates include the
    // border).  INSIDE: the caller guarantees 0 <= floor(X), floor(Y) <= W - 2,
    // so the cell index is one exact FP32 multiply-add.  Otherwise the cell is
    // clamped to the copy, which reads the same edge values as clamping each tap
    // to the image.
    auto tap = [&](float X, float Y, auto inside_tag) __attribute__((always_inline)) {
        const float x0 = floorf(X), y0 = floorf(Y);
        const float fx = X - x0, fy = Y - y0;
        const float *q;
        if constexpr (decltype(inside_tag)::value) {
            q = img + (int)fmaf(y0, Wf, x0);
        } else {
            const int cx = min(max((int)x0, 0), W - 2), cy = min(max((int)y0, 0), W - 2);
            q = img + (int)__umul24((unsigned)cy, (unsigned)W) + cx;
        }
        const float v00 = q[0], v01 = q[1], v10 = q[W], v11 = q[W + 1];
        const float t0 = fmaf(fx, v01 - v00, v00), t1 = fmaf(fx, v11 - v10, v10);
        return fmaf(fy, t1 - t0, t0);
    };

    // ring pair (i, i + 1) of one shift row from the lane's samples of both
    // rings: normalisation sums, one complex FFT, both spectra packed as FP16
    // (Re + Im, Re - Im); slot 0: (S[0], S[128])
    auto ring_pair = [&](int i, const float (&va)[4], const float (&vb)[4],
                         unsigned (&sp)[kFragMaxRings][2], double &sA, double &sQ)
        __attribute__((always_inline)) {
        const float ra = (float)(i + 1), rb = (float)(i + 2);
        float2 v[4];
        float pa = 0.f, qa = 0.f, pb = 0.f, qb = 0.f;
#pragma unroll
        for (int q = 0; q < 4; ++q) {
            v[q] = f2(va[q], vb[q]);
            pa += va[q];
            qa = fmaf(va[q], va[q], qa);
            pb += vb[q];
            qb = fmaf(vb[q], vb[q], qb);
        }
        sA += (double)ra * pa + (double)rb * pb;
        sQ += (double)ra * qa + (double)rb * qb;
        fft256_wave<false>(v, b0, b1, tw, lane);
        const float2 u2 = shfl2(v[2], pl), u3 = shfl2(v[3], pl);
        const float2 zl = v[0], zlp = lane ? u3 : v[0];
        const float2 zh = v[1], zhp = lane ? u2 : v[3];
        const float2 al = f2(0.5f * (zl.x + zlp.x), 0.5f * (zl.y - zlp.y));
        const float2 bl = f2(0.5f * (zl.y + zlp.y), 0.5f * (zlp.x - zl.x));
        const float2 ah = f2(0.5f * (zh.x + zhp.x), 0.5f * (zh.y - zhp.y));
        const float2 bh = f2(0.5f * (zh.y + zhp.y), 0.5f * (zhp.x - zh.x));
        // slot 0 (lane 0) keeps the two real bins (S[0], S[128]) instead of
        // (Re + Im, Re - Im): selected, not branched
        const unsigned na = h2pack(al.x + al.y, al.x - al.y), za = h2pack(zl.x, v[2].x);
        sp[i][0] = lane ? na : za;
        if (i + 1 < kFragMaxRings) {
            const unsigned nb = h2pack(bl.x + bl.y, bl.x - bl.y), zb = h2pack(zl.y, v[2].y);
            sp[i + 1][0] = lane ? nb : zb;
        }
        sp[i][1] = h2pack(ah.x + ah.y, ah.x - ah.y);
        if (i + 1 < kFragMaxRings) sp[i + 1][1] = h2pack(bh.x + bh.y, bh.x - bh.y);
    };

    // row j's normalisation and 16-bit operand stores
    auto store_row = [&](int j, const unsigned (&sp)[kFragMaxRings][2], double sA, double sQ)
        __attribute__((always_inline)) {
#pragma unroll
        for (int off = 32; off >= 1; off >>= 1) {
            sA += __shfl_xor(sA, off, kWave);
            sQ += __shfl_xor(sQ, off, kWave);
        }
        // row statistics of the standardised samples; the row's normalisation of
        // the original image is g = 1 / (isd sd + 1e-7), applied to spectra / rs
        const double nn = (double)kRingLen * 0.5 * (double)R * (double)(R + 1);
        const double mu = sA / nn;
        const double sd = sqrt(fabs(sQ / nn - mu * mu));
        const float gf = (float)(isd / (isd * sd + 1e-7)) * cp;
        const float muL = (float)(mu * kRingLen);
        unsigned char *rowp = prow + (size_t)j * row_bytes;
#pragma unroll
        for (int h = 0; h < 2; ++h) {   // slot lane (h = 0) and lane + 64 (h = 1)
            const int k = lane + 64 * h;
            unsigned char *dst = rowp + (size_t)k * slot_bytes;
            const float dc = (k == 0) ? muL : 0.0f;
#pragma unroll
            for (int ks = 0; ks < (kFragMaxRings + 15) / 16; ++ks) {
                if (ks < full16 + (tail == 16 ? 1 : 0)) {
#pragma unroll
                    for (int pt = 0; pt < 2; ++pt)
#pragma unroll
                        for (int hh = 0; hh < 2; ++hh) {
                            unsigned e[4];
#pragma unroll
                            for (int jj = 0; jj < 4; ++jj) {
                                unsigned short lo = 0, hi = 0;
                                const int i0 = 16 * ks + 8 * hh + 2 * jj;
                                if (i0 < kFragMaxRings && i0 < R) {
                                    const unsigned v = sp[i0][h];
                                    lo = frag16<BF16>(pt ? h2hi(v) * gf : (h2lo(v) - dc) * gf);
                                }
                                if (i0 + 1 < kFragMaxRings && i0 + 1 < R) {
                                    const unsigned v = sp[i0 + 1][h];
                                    hi = frag16<BF16>(pt ? h2hi(v) * gf : (h2lo(v) - dc) * gf);
                                }
                                e[jj] = lo | ((unsigned)hi << 16);
                            }
                            *reinterpret_cast<uint4 *>(dst + ks * 64 + pt * 32 + hh * 16) =
                                make_uint4(e[0], e[1], e[2], e[3]);
                        }
                }
            }
            if (tail == 8) {
#pragma unroll
                for (int pt = 0; pt < 2; ++pt)
#pragma unroll
                    for (int hh = 0; hh < 2; ++hh) {
                        unsigned e[2];
#pragma unroll
                        for (int jj = 0; jj < 2; ++jj) {
                            unsigned short lo = 0, hi = 0;
                            // full16 <= 2 when R <= 40: ring indices below are static per full16
#pragma unroll
                            for (int f16 = 0; f16 <= 2; ++f16) {
                                if (f16 != full16) continue;
                                const int i0 = 16 * f16 + 4 * hh + 2 * jj;
                                if (i0 < kFragMaxRings && i0 < R) {
                                    const unsigned v = sp[i0][h];
                                    lo = frag16<BF16>(pt ? h2hi(v) * gf : (h2lo(v) - dc) * gf);
                                }
                                if (i0 + 1 < kFragMaxRings && i0 + 1 < R) {
                                    const unsigned v = sp[i0 + 1][h];
                                    hi = frag16<BF16>(pt ? h2hi(v) * gf : (h2lo(v) - dc) * gf);
                                }
                            }
                            e[jj] = lo | ((unsigned)hi << 16);
                        }
                        *reinterpret_cast<uint2 *>(dst + full16 * 64 + pt * 16 + hh * 8) =
                            make_uint2(e[0], e[1]);
                    }
            }
        }
    };

    // all rings of row j: sample position centre + shift + (i + 1)(cos, sin)
    auto row = [&](int j, float cx, float cy, auto inside_tag) __attribute__((always_inline)) {
        unsigned sp[kFragMaxRings][2];
        double sA = 0.0, sQ = 0.0;
#pragma unroll
        for (int i = 0; i < kFragMaxRings; i += 2) {
            if (i < R) {   // uniform
                const bool hb = (i + 1) < R;
                const float ra = (float)(i + 1), rb = (float)(i + 2);
                float va[4], vb[4];
#pragma unroll
                for (int q = 0; q < 4; ++q) {
                    // ring i + 2 of an odd R is sampled (inside the checked extent) and
                    // then zeroed, so the loop body has no branch per sample
                    va[q] = tap(fmaf(ra, cq[q], cx), fmaf(ra, sq[q], cy), inside_tag);
                    const float t = tap(fmaf(rb, cq[q], cx), fmaf(rb, sq[q], cy), inside_tag);
                    vb[q] = hb ? t : 0.0f;
                }
                ring_pair(i, va, vb, sp, sA, sQ);
            }
        }
        store_row(j, sp, sA, sQ);
    };

    for (int j = wave; j < n_rows; j += 4) {
        const int2 sh = shift_i[j];
        const float cx = (float)(D / 2 + kFragBorder + sh.x) + ax;
        const float cy = (float)(D / 2 + kFragBorder + sh.y) + ay;
        // uniform: every cell of the row's rings lies in the bordered copy (a
        // one-pixel margin covers the rounding of the positions)
        const float ext = (float)(R + 1), hi = (float)(W - 2);
        if (cx - ext >= 0.0f && cy - ext >= 0.0f && cx + ext <= hi && cy + ext <= hi)
            row(j, cx, cy, std::true_type{});
        else
            row(j, cx, cy, std::false_type{});
    }
}

__global__ __launch_bounds__(128) void k_ref_finalize(float2 *__restrict__ rspec,
                                                      const float2 *__restrict__ rnorm, int R,
                                                      float *__restrict__ rsum0) {
    const int n = blockIdx.x, k = threadIdx.x;
    const float2 nm = rnorm[n];
    double acc0 = 0.0;
    for (int i = 0; i < R; ++i) {
        float2 *px = rspec + ((size_t)n * R + i) * kSlots + slot_pos(k);
        float2 x = *px;
        const float gw = nm.y * (float)(i + 1);
        if (k == 0) {
            x.x = (x.x - nm.x) * gw;
            x.y = x.y * gw;
            acc0 += (double)x.x;
        } else {
            x.x *= gw;
            x.y *= gw;
        }
        *px = x;
    }
    if (k == 0) rsum0[n] = (float)acc0;
}

// ------------------------------------------------------------------ stage B
// Psi stage of one (shift row, reference) pair.  Q / P hold the four real ring
// products P1..P4 (sr*rr, si*ri, si*rr, sr*ri summed over rings with the ring
// weight folded into the reference) of spectrum slots `lane` and `lane+64`:
// C0 = R conj S = (P1+P2, P4-P3), C1 = conj(R S) = (P1-P2, -(P3+P4)).  The pair
// becomes one 256-point inverse FFT of Z = C0 + i C1 (Hermitian-extended), so
// z[t] = c0[t] + i c1[t] holds the psi CCF of both mirrors.  The running
// first-max over (mirror, shift, psi) of this reference (bv, bf) and the seven
// samples around it (nb, LDS) are updated; the 6-step wave argmax only runs
// when some lane beats the running maximum (rows arrive in increasing shift
// order, so an equal value can never displace the earlier maximum).
//
// MASKED: only samples whose bit (m * 4 + q) is set in `allow` (lane's samples
// lane + 64 q of mirror m) take part in the maximum; the seven neighbours stay
// the unmasked CCF values.
template <bool SINGLE, bool MASKED = false>
__device__ __forceinline__ void psi_update(const float (&Q)[4], const float (&P)[4], float2 nm,
                                           float rs0, int s, int lane, float2 *b0, float2 *b1,
                                           const float2 *tw, float *zb, float (*nb)[8],
                                           float (&bv)[2], int (&bf)[2], unsigned allow = 0xffu) {
    const int pl = (64 - lane) & 63;
    const float gs = nm.y;
    float A = P[0] + P[1], B = P[3] - P[2], Cc = P[0] - P[1], Dd = -(P[2] + P[3]);
    const float2 zhi = f2(gs * (A - Dd), gs * (B + Cc));
    const float2 yhi = f2(gs * (A + Dd), gs * (Cc - B));
    A = Q[0] + Q[1];
    B = Q[3] - Q[2];
    Cc = Q[0] - Q[1];
    Dd = -(Q[2] + Q[3]);
    float2 zlo = f2(gs * (A - Dd), gs * (B + Cc));
    const float2 ylo = f2(gs * (A + Dd), gs * (Cc - B));
    float2 nyq = f2(0.0f, 0.0f);
    if (lane == 0) {  // packed slot 0: DC (with the mean correction) and Nyquist
        const float c0 = gs * (Q[0] - nm.x * rs0);
        const float cn = gs * Q[1];
        zlo = f2(c0, c0);
        nyq = f2(cn, cn);
    }
    const float2 syhi = shfl2(yhi, pl), sylo = shfl2(ylo, pl);
    float2 v[4];
    v[0] = zlo;
    v[1] = zhi;
    v[2] = lane ? syhi : nyq;
    v[3] = lane ? sylo : yhi;
    fft256_wave<true>(v, b0, b1, tw, lane);
#pragma unroll
    for (int m = 0; m < 2; ++m) {
        float bvl = m ? v[0].y : v[0].x;
        if (MASKED && !((allow >> (m * 4)) & 1u)) bvl = -INFINITY;
        int bt = lane;
#pragma unroll
        for (int q = 1; q < 4; ++q) {
            float x = m ? v[q].y : v[q].x;
            if (MASKED && !((allow >> (m * 4 + q)) & 1u)) x = -INFINITY;
            if (x > bvl) {
                bvl = x;
                bt = lane + 64 * q;
            }
        }
        const bool sentinel = SINGLE && m == 0 && s == 0 && bf[0] == 0 && bv[0] == 0.0f;
        if (!__any(bvl > bv[m]) && !sentinel) continue;
        wave_argmax(bvl, bt);
        const int flat = s * kRingLen + bt;
        const bool upd = bvl > bv[m] || (bvl == bv[m] && flat < bf[m]);
        if (upd || sentinel) {
            const int tc = upd ? bt : 0;
            if (upd) {
                bv[m] = bvl;
                bf[m] = flat;
            }
            wave_lds_sync();
#pragma unroll
            for (int q = 0; q < 4; ++q) zb[lane + 64 * q] = m ? v[q].y : v[q].x;
            wave_lds_sync();
            if (lane < 7) nb[m][lane] = zb[(tc + lane - 3) & 255];
            wave_lds_sync();
        }
    }
}

__device__ __forceinline__ void write_rec(PeakRec *o, const float (&bv)[2], const int (&bf)[2],
                                          const float (*nb)[8], int lane) {
    const int m = (bv[1] > bv[0]) ? 1 : 0;
    const int f = (bf[m] == INT_MAX) ? 0 : bf[m];   // never updated (all values NaN)
    if (lane == 0) {
        o->val = bv[m];
        o->m = m;
        o->s = f / kRingLen;
        o->t = f % kRingLen;
        o->pad = 0;
    }
    if (lane < 7) o->x[lane] = nb[m][lane];
}

// Items (particle, reference): one wavefront per item walks the shifts four at a
// time (4 rows x 1 reference).  Item i uses particle i / K and reference
// ref_of[i * ref_stride].  ZERO_INIT: the search starts from value 0 at flat index
// 0 like the reference's row argmax for pre-alignment / reference-free
// (cuda/gpu_aln_noref.cu:1638-1664); otherwise from -FLT_MAX like the
// multi-reference wizard, giving exactly the FP32 multi-reference kernel's
// arithmetic for that pair (used to re-score candidates).
template <bool ZERO_INIT>
__global__ __launch_bounds__(256) void k_ccf_items(
    const float2 *__restrict__ spec, const float2 *__restrict__ norm,
    const float2 *__restrict__ rspec, const float *__restrict__ rsum0,
    const int *__restrict__ ref_of, int ref_stride, int K, int n_items, int S, int R, int N,
    const float2 *__restrict__ twg, PeakRec *__restrict__ rec, const int *__restrict__ pidx) {
    __shared__ float2 tw[256];
    __shared__ float2 fbuf[4][512];
    __shared__ float zbuf[4][256];
    __shared__ float nbuf[4][2][8];
    const int tid = threadIdx.x, lane = tid & 63, wave = __builtin_amdgcn_readfirstlane(tid >> 6);
    fill_fft_twiddles(tw, twg, tid, 256);
    __syncthreads();
    const int item = blockIdx.x * 4 + wave;
    if (item >= n_items) return;
    const int p = pidx ? pidx[item / K] : item / K;   // spectra row of the item's particle
    float bv[2];
    int bf[2];
    bv[0] = bv[1] = ZERO_INIT ? 0.0f : -3.40282e+38f;
    bf[0] = bf[1] = ZERO_INIT ? 0 : INT_MAX;
    const int n_raw = ref_of[(size_t)item * ref_stride];
    if (!ZERO_INIT && n_raw < 0) {   // no reference for this item: an empty record
        if (lane < 8) nbuf[wave][0][lane] = 0.0f;
        write_rec(rec + item, bv, bf, nbuf[wave], lane);
        return;
    }
    // the host validates ref ids; the clamp keeps every read in bounds
    const int n = min(max(n_raw, 0), N - 1);
    const float rs0 = rsum0[n];
    const float2 *rb0 = rspec + (size_t)n * R * kSlots;
    const size_t img_stride = (size_t)S * R * kSlots;
    for (int s0 = 0; s0 < S; s0 += 4) {
        float acc[4][2][4];
#pragma unroll
        for (int r = 0; r < 4; ++r)
#pragma unroll
            for (int h = 0; h < 2; ++h)
#pragma unroll
                for (int e = 0; e < 4; ++e) acc[r][h][e] = 0.0f;
        const float2 *arow[4];
#pragma unroll
        for (int r = 0; r < 4; ++r)
            arow[r] = spec + (size_t)p * img_stride + (size_t)min(s0 + r, S - 1) * R * kSlots;
        for (int i = 0; i < R; ++i) {
            float2 b0v, b1v;
            load_slots(rb0 + i * kSlots, lane, b0v, b1v);
            float2 a0s[4], a1s[4];
#pragma unroll
            for (int r = 0; r < 4; ++r) load_slots(arow[r] + i * kSlots, lane, a0s[r], a1s[r]);
            __builtin_amdgcn_sched_barrier(0);   // all of the ring's loads in flight before the FMAs
#pragma unroll
            for (int r = 0; r < 4; ++r) {
                const float2 a0 = a0s[r], a1 = a1s[r];
                acc[r][0][0] += b0v.x * a0.x;
                acc[r][0][1] += b0v.y * a0.y;
                acc[r][0][2] += b0v.x * a0.y;
                acc[r][0][3] += b0v.y * a0.x;
                acc[r][1][0] += b1v.x * a1.x;
                acc[r][1][1] += b1v.y * a1.y;
                acc[r][1][2] += b1v.x * a1.y;
                acc[r][1][3] += b1v.y * a1.x;
            }
        }
#pragma unroll
        for (int r = 0; r < 4; ++r) {
            const int s = s0 + r;
            if (s >= S) break;
            psi_update<ZERO_INIT>(acc[r][0], acc[r][1], norm[(size_t)p * S + s], rs0, s, lane,
                                  fbuf[wave], fbuf[wave] + 256, tw, zbuf[wave], nbuf[wave], bv, bf);
        }
    }
    write_rec(rec + item, bv, bf, nbuf[wave], lane);
}

// Top-K references of every particle by approximate peak value (descending; ties
// to the lower reference index).  One thread per particle.
template <int K>
__global__ void k_topk(const PeakRec *__restrict__ rec, int n_img, int N, int *__restrict__ cand) {
    const int p = blockIdx.x * blockDim.x + threadIdx.x;
    if (p >= n_img) return;
    float v[K];
    int id[K];
#pragma unroll
    for (int k = 0; k < K; ++k) {
        v[k] = -INFINITY;
        id[k] = 0;
    }
    for (int n = 0; n < N; ++n) {
        const float x = rec[(size_t)p * N + n].val;
        if (!(x > v[K - 1])) continue;
        float cv = x;
        int ci = n;
#pragma unroll
        for (int k = 0; k < K; ++k) {
            if (cv > v[k]) {
                const float tv = v[k];
                const int ti = id[k];
                v[k] = cv;
                id[k] = ci;
                cv = tv;
                ci = ti;
            }
        }
    }
    const int kk = min(K, N);
#pragma unroll
    for (int k = 0; k < K; ++k) cand[(size_t)p * K + k] = (k < kk) ? id[k] : id[0];
}

// Best exact candidate per particle: highest FP32 peak (ties to the lower
// reference index); flags a near-tie when the best distinct candidate is within
// tie_rel of it.  best[p] = candidate slot, second[p] = runner-up slot (-1 none).
__global__ void k_best_candidate(const PeakRec *__restrict__ rec, const int *__restrict__ cand,
                                 int n_img, int K, float tie_rel, int *__restrict__ best,
                                 int *__restrict__ second, int *__restrict__ tie) {
    const int p = blockIdx.x * blockDim.x + threadIdx.x;
    if (p >= n_img) return;
    int b = -1, s2 = -1;
    for (int k = 0; k < K; ++k) {
        const int n = cand[(size_t)p * K + k];
        bool dup = false;
        for (int j = 0; j < k; ++j) dup |= cand[(size_t)p * K + j] == n;
        if (dup) continue;
        const float x = rec[(size_t)p * K + k].val;
        if (b < 0) {
            b = k;
            continue;
        }
        const float xb = rec[(size_t)p * K + b].val;
        const int nb = cand[(size_t)p * K + b];
        if (x > xb || (x == xb && n < nb)) {
            s2 = b;
            b = k;
        } else if (s2 < 0) {
            s2 = k;
        } else {
            const float x2 = rec[(size_t)p * K + s2].val;
            const int n2 = cand[(size_t)p * K + s2];
            if (x > x2 || (x == x2 && n < n2)) s2 = k;
        }
    }
    best[p] = b;
    second[p] = s2;
    int t = 0;
    if (s2 >= 0) {
        const float xb = rec[(size_t)p * K + b].val, x2 = rec[(size_t)p * K + s2].val;
        t = (xb - x2) <= tie_rel * fabsf(xb);
    }
    tie[p] = t;
}

// Near-tie re-scoring in FP64: for a flagged particle, one wavefront per candidate
// (best, runner-up) recomputes C0/C1 over rings in double from the FP32 spectra
// and evaluates the psi CCF of its winning mirror / shift at t*-1, t*, t*+1 by
// direct summation: c[t] = C[0] + C[128] (-1)^t + 2 sum_k Re(C[k] e^{2 pi i k t/256}).
// val64[2p + j] = the largest of the three (j = 0 best, 1 runner-up).
__global__ __launch_bounds__(256) void k_rescore_f64(
    const float2 *__restrict__ spec, const float2 *__restrict__ norm,
    const float2 *__restrict__ rspec, const PeakRec *__restrict__ rec,
    const int *__restrict__ cand, const int *__restrict__ best, const int *__restrict__ second,
    const int *__restrict__ tie, int n_img, int K, int S, int R, int N, double *__restrict__ val64) {
    const int lane = threadIdx.x & 63;
    const int w = blockIdx.x * 4 + (threadIdx.x >> 6);
    const int p = w >> 1, j = w & 1;
    if (p >= n_img || !tie[p]) return;
    const int slot_k = j ? second[p] : best[p];
    if (slot_k < 0) return;
    const PeakRec r = rec[(size_t)p * K + slot_k];
    const int n = min(max(cand[(size_t)p * K + slot_k], 0), N - 1);
    const int s = min(max(r.s, 0), S - 1);
    const float2 nm = norm[(size_t)p * S + s];
    const double muL = nm.x, g = nm.y;
    double P[2][4] = {{0, 0, 0, 0}, {0, 0, 0, 0}};
    const float2 *a = spec + ((size_t)p * S + s) * R * kSlots;
    const float2 *b = rspec + (size_t)n * R * kSlots;
    for (int i = 0; i < R; ++i) {
#pragma unroll
        for (int h = 0; h < 2; ++h) {
            const int k = lane + 64 * h;
            const float2 av = a[i * kSlots + slot_pos(k)], bv = b[i * kSlots + slot_pos(k)];
            const double x = ((k == 0) ? ((double)av.x - muL) : (double)av.x) * g;
            const double y = (double)av.y * g;
            P[h][0] += (double)bv.x * x;
            P[h][1] += (double)bv.y * y;
            P[h][2] += (double)bv.x * y;
            P[h][3] += (double)bv.y * x;
        }
    }
    double best_v = -INFINITY;
    for (int dt = -1; dt <= 1; ++dt) {
        const int t = (r.t + dt + 256) & 255;
        double sum = 0.0;
#pragma unroll
        for (int h = 0; h < 2; ++h) {
            const int k = lane + 64 * h;
            const double c0r = P[h][0] + P[h][1], c0i = P[h][3] - P[h][2];
            const double c1r = P[h][0] - P[h][1], c1i = -(P[h][2] + P[h][3]);
            const double cr = r.m ? c1r : c0r, ci = r.m ? c1i : c0i;
            if (k == 0) {  // DC and Nyquist (slot 0 packs both)
                sum += P[h][0] + ((t & 1) ? -P[h][1] : P[h][1]);
            } else {
                double sn, cs;
                sincospi((double)(2 * k * t % 512) / 256.0, &sn, &cs);
                sum += 2.0 * (cr * cs - ci * sn);
            }
        }
#pragma unroll
        for (int off = 32; off >= 1; off >>= 1) sum += __shfl_xor(sum, off, 64);
        best_v = fmax(best_v, sum);
    }
    if (lane == 0) val64[2 * p + j] = best_v;
}

// Swap best and runner-up of flagged particles when FP64 ranks them the other way
// (equal values: the lower reference index wins).
__global__ void k_resolve_ties(const int *__restrict__ cand, const double *__restrict__ val64,
                               const int *__restrict__ tie, int n_img, int K, int *__restrict__ best,
                               int *__restrict__ second) {
    const int p = blockIdx.x * blockDim.x + threadIdx.x;
    if (p >= n_img || !tie[p] || second[p] < 0) return;
    const double vb = val64[2 * p], vs = val64[2 * p + 1];
    const int nb = cand[(size_t)p * K + best[p]], ns = cand[(size_t)p * K + second[p]];
    if (vs > vb || (vs == vb && ns < nb)) {
        const int t = best[p];
        best[p] = second[p];
        second[p] = t;
    }
}

// Multi-reference: one workgroup = one particle x NW groups of 4 references (one
// group per wavefront).  The workgroup walks the shifts four at a time; the four
// shift rows' ring spectra are staged cooperatively in LDS in double-buffered
// chunks of kRingChunk rings, so only the references' spectra come through the
// vector L1.  Lane l accumulates, over rings, P1..P4 of slots l and l+64 for
// 4 rows x 4 references (128 accumulators), then each (row, reference) pair goes
// through psi_update.
constexpr int kRingChunk = 4;

// RW shift rows per group x RB references per wavefront (RW RB = 16: 128 accumulators),
// rows staged CH rings at a time.
template <int NW, int RW = 4, int RB = 4, int CH = kRingChunk>
struct MrefLds {
    float2 tw[256];
    float4 abuf4[2][RW * CH * kSlots / 2];   // [buf][ring][row][slot pairs]
    float2 fbuf[NW][512];
    float zbuf[NW][256];
    float nbuf[NW][RB][2][8];
};

// Work of one workgroup: particle p of job j, reference groups gb*NW .. gb*NW+NW-1
// (the twiddles must already be in sh.tw), rows staged CH rings at a time.  CAND:
// the candidate form of MrefJob, p is the list index; every slot of an active
// group gets a record (an empty one for a negative id), and each listed
// reference's record is the one the exhaustive form computes for it.
template <int NW, int RW = 4, int RB = 4, int CH = kRingChunk, bool CAND = false>
__device__ __forceinline__ void mref_block(const MrefJob &j, int p, int gb,
                                           MrefLds<NW, RW, RB, CH> &sh) {
    static_assert(RW * RB == 16, "128 accumulators per lane");
    static_assert(CH % 2 == 0, "rings are taken in pairs within a chunk");
    constexpr int CH4 = RW * CH * kSlots / 2;   // float4 per staged chunk
    const int S = j.S, R = j.R, N = j.N;
    const int tid = threadIdx.x, lane = tid & 63;
    // uniform over the wavefront: the reference indices and spectrum addresses
    // built from it live in scalar registers
    const int wave = __builtin_amdgcn_readfirstlane(tid >> 6);
    const int g = gb * NW + wave;
    const bool active = g < j.G;
    const int ps = (CAND && j.pidx) ? j.pidx[p] : p;   // spectra row
    int nref[RB];
    bool nval[RB];
#pragma unroll
    for (int c = 0; c < RB; ++c) {
        if (CAND) {
            // the host validates candidate ids; the clamp keeps every read in bounds
            const int raw = active ? j.cand[(size_t)p * j.K + g * RB + c] : -1;
            nval[c] = raw >= 0;
            nref[c] = min(max(raw, 0), N - 1);
            if (active && !nval[c] && lane < 8) sh.nbuf[wave][c][0][lane] = 0.0f;
        } else {
            const int n = g * RB + c;
            nval[c] = active && n < N;
            nref[c] = min(n, N - 1);
        }
    }
    // candidate form: a wavefront whose four slots are all empty only helps stage
    // the rows, and a workgroup whose slots are all empty skips the rows
    bool wlive = true, wg_live = true;
    if (CAND) {
        wlive = false;
#pragma unroll
        for (int c = 0; c < RB; ++c) wlive = wlive || nval[c];
        wg_live = false;
        for (int w = 0; w < NW; ++w) {
            const int gw = gb * NW + w;
            if (gw >= j.G) break;
            for (int c = 0; c < RB; ++c) wg_live = wg_live || j.cand[(size_t)p * j.K + gw * RB + c] >= 0;
        }
    }
    float bv[RB][2];
    int bf[RB][2];
    float rs0[RB];
#pragma unroll
    for (int c = 0; c < RB; ++c) {
        rs0[c] = j.rsum0[nref[c]];
        bv[c][0] = bv[c][1] = -3.40282e+38f;
        bf[c][0] = bf[c][1] = INT_MAX;
    }
    const size_t img_stride = (size_t)S * R * kSlots;
    const int nchunk = (R + CH - 1) / CH;
    const float2 *abuf = reinterpret_cast<const float2 *>(&sh.abuf4[0][0]);

    for (int s0 = 0; wg_live && s0 < S; s0 += RW) {
        const float4 *rows4[RW];
#pragma unroll
        for (int r = 0; r < RW; ++r)
            rows4[r] = reinterpret_cast<const float4 *>(j.spec + (size_t)ps * img_stride +
                                                        (size_t)min(s0 + r, S - 1) * R * kSlots);
        // Stage one chunk (4 rows x CH rings, 1 KB per row-ring) with direct
        // global->LDS loads: one wavefront instruction moves one row-ring, each lane
        // 16 bytes to LDS base + 16*lane.
        auto load_chunk = [&](int ch, int buf) {
            for (int rr = wave; rr < RW * CH; rr += NW) {
                const int ii = rr / RW, r = rr % RW;
                const int ring = ch * CH + ii;
                if (ring >= R) continue;
                __builtin_amdgcn_global_load_lds(
                    (const void *)(rows4[r] + ring * (kSlots / 2) + lane),
                    (__attribute__((address_space(3))) void *)(&sh.abuf4[buf][(ii * RW + r) * (kSlots / 2)]),
                    16, 0, 0);
            }
        };
        float acc[RW][RB][2][4];
#pragma unroll
        for (int r = 0; r < RW; ++r)
#pragma unroll
            for (int c = 0; c < RB; ++c)
#pragma unroll
                for (int h = 0; h < 2; ++h)
#pragma unroll
                    for (int e = 0; e < 4; ++e) acc[r][c][h][e] = 0.0f;
        // Reference spectra of ring i + 1 load while ring i's FMAs run (ping-pong
        // registers bA / bB, rings taken in pairs), and the next chunk's rows are
        // staged from the chunk's last ring on: every load is issued after the
        // LDS reads of the ring in flight, so the waits the compiler places
        // before the next LDS reads find them a ring of FMAs old.
        auto load_refs = [&](int i, float2 (&b)[RB][2]) {
            if (CAND && !wlive) return;
#pragma unroll
            for (int c = 0; c < RB; ++c)
                load_slots(j.rspec + ((size_t)nref[c] * R + i) * kSlots, lane, b[c][0], b[c][1]);
        };
        auto ring = [&](int ii, int buf, const float2 (&b)[RB][2], auto &&issue) {
            if (CAND && !wlive) {   // uniform over the wavefront: its share of the staging only
                issue();
                return;
            }
            float2 a[RW][2];
#pragma unroll
            for (int r = 0; r < RW; ++r) {
                const float2 *ar = abuf + (size_t)buf * (2 * CH4) + (ii * RW + r) * kSlots;
                load_slots(ar, lane, a[r][0], a[r][1]);
            }
            issue();
            __builtin_amdgcn_sched_barrier(0);
#pragma unroll
            for (int r = 0; r < RW; ++r)
#pragma unroll
                for (int c = 0; c < RB; ++c)
#pragma unroll
                    for (int h = 0; h < 2; ++h) {
                        acc[r][c][h][0] += b[c][h].x * a[r][h].x;
                        acc[r][c][h][1] += b[c][h].y * a[r][h].y;
                        acc[r][c][h][2] += b[c][h].x * a[r][h].y;
                        acc[r][c][h][3] += b[c][h].y * a[r][h].x;
                    }
        };
        float2 bA[RB][2], bB[RB][2];
        load_chunk(0, 0);
        load_refs(0, bA);
        __builtin_amdgcn_s_waitcnt(0);
        __syncthreads();
        for (int ch = 0; ch < nchunk; ++ch) {
            const int buf = ch & 1;
            const int i0 = ch * CH, i1 = min(R, i0 + CH);
            // CH is even, so only the last chunk can end on an unpaired ring
            for (int i = i0; i < i1; i += 2) {
                const bool last_pair = i + 2 >= i1;
                ring(i - i0, buf, bA, [&] {
                    if (i + 1 < R) load_refs(i + 1, bB);
                    if (last_pair && i + 1 >= i1 && ch + 1 < nchunk) load_chunk(ch + 1, buf ^ 1);
                });
                if (i + 1 < i1)
                    ring(i + 1 - i0, buf, bB, [&] {
                        if (i + 2 < R) load_refs(i + 2, bA);
                        if (last_pair && ch + 1 < nchunk) load_chunk(ch + 1, buf ^ 1);
                    });
            }
            __builtin_amdgcn_s_waitcnt(0);
            __syncthreads();
        }
        if (!active) continue;
#pragma unroll
        for (int r = 0; r < RW; ++r) {
            const int s = s0 + r;
            if (s >= S) break;
            const float2 nm = j.norm[(size_t)ps * S + s];
#pragma unroll
            for (int c = 0; c < RB; ++c) {
                if (!nval[c]) continue;
                psi_update<false>(acc[r][c][0], acc[r][c][1], nm, rs0[c], s, lane, sh.fbuf[wave],
                                  sh.fbuf[wave] + 256, sh.tw, sh.zbuf[wave], sh.nbuf[wave][c],
                                  bv[c], bf[c]);
            }
        }
    }
#pragma unroll
    for (int c = 0; c < RB; ++c) {
        if (CAND) {
            if (active)
                write_rec(j.rec + (size_t)p * j.K + g * RB + c, bv[c], bf[c], sh.nbuf[wave][c], lane);
        } else if (nval[c]) {
            write_rec(j.rec + (size_t)p * N + nref[c], bv[c], bf[c], sh.nbuf[wave][c], lane);
        }
    }
}

// Multi-reference: one workgroup = one particle x NW groups of 4 references (one
// group per wavefront).  The workgroup walks the shifts four at a time; the four
// shift rows' ring spectra are staged cooperatively in LDS in double-buffered
// chunks of kRingChunk rings, so only the references' spectra come through the
// vector L1.  Lane l accumulates, over rings, P1..P4 of slots l and l+64 for
// 4 rows x 4 references (128 accumulators), then each (row, reference) pair goes
// through psi_update.
template <int NW, int RW = 4, int RB = 4>
__global__ __launch_bounds__(NW * 64, 2) void k_ccf_mref(MrefJob j, const float2 *__restrict__ twg) {
    __shared__ MrefLds<NW, RW, RB> sh;
    fill_fft_twiddles(sh.tw, twg, threadIdx.x, NW * 64);
    __syncthreads();
    const int lb = xcd_remap(blockIdx.x, gridDim.x);
    const int p = lb / j.GB;
    if (p >= j.n_img) return;  // uniform over the workgroup
    mref_block<NW, RW, RB>(j, p, lb % j.GB, sh);
}

// Candidate re-scoring (the candidate form of MrefJob) by the exhaustive kernel's
// workgroup code, so every record equals k_ccf_mref's for that (particle,
// reference): G = K / 4 groups of four listed references, NW per workgroup, GB
// workgroups per list.  Rows are staged two rings at a time, so that as many
// workgroups fit the LDS as the registers allow (two wavefronts per SIMD).
template <int NW>
__global__ __launch_bounds__(NW * 64, 2) void k_ccf_cand_mref(MrefJob j, const float2 *__restrict__ twg) {
    __shared__ MrefLds<NW, 4, 4, 2> sh;
    fill_fft_twiddles(sh.tw, twg, threadIdx.x, NW * 64);
    __syncthreads();
    const int p = blockIdx.x / j.GB;
    if (p >= j.n_img) return;  // uniform over the workgroup
    mref_block<NW, 4, 4, 2, true>(j, p, blockIdx.x - p * j.GB, sh);
}

// Persistent form over many jobs (batched subsets): work items (job, particle,
// reference block) sit in 8 queues, one per XCD (items of one particle share a
// queue, so its spectra stay in one L2); each workgroup takes items from the
// queue of the XCD it runs on (blockIdx % 8) until the queue is exhausted.  The
// claim is a vector atomic by one lane, broadcast through LDS.  Every workgroup
// leaves the loop once its queue is empty.  (The one-wavefront tile asks for no
// occupancy target: its registers leave one wavefront per SIMD either way.)
template <int NW>
__global__ __launch_bounds__(NW * 64, NW == 1 ? 1 : 2) void k_ccf_mref_queue(
    const MrefJob *__restrict__ jobs, const int4 *__restrict__ items, const int *__restrict__ qbeg,
    int *__restrict__ counters, const float2 *__restrict__ twg) {
    __shared__ MrefLds<NW> sh;
    __shared__ int claim;
    fill_fft_twiddles(sh.tw, twg, threadIdx.x, NW * 64);
    const int q = blockIdx.x & 7;
    const int q0 = qbeg[q], q1 = qbeg[q + 1];
    for (;;) {
        __syncthreads();
        if (threadIdx.x == 0) claim = atomicAdd(&counters[q], 1);
        __syncthreads();
        const int idx = q0 + claim;
        if (idx >= q1) break;
        const int4 w = items[idx];
        mref_block<NW>(jobs[w.x], w.y, w.z, sh);
    }
}

__global__ void k_reset_counters(int *__restrict__ c, int n) {
    const int i = threadIdx.x;
    if (i < n) c[i] = 0;
}

// ------------------------------------------------------------------ decode
// Angle of a multi-reference peak, as the reference wizard computes it
// (cuda/gpu_aln_noref.cu:1777-1793): float coefficients from double
// expressions, the offset in double, the angle stored as float.
__device__ __forceinline__ float wizard_angle(const PeakRec &r) {
    const float *x = r.x;
    const float c2 = (float)(49. * x[0] + 6. * x[1] - 21. * x[2] - 32. * x[3] - 27. * x[4] -
                             6. * x[5] + 31. * x[6]);
    const float c3 = (float)(5. * x[0] - 3. * x[2] - 4. * x[3] - 3. * x[4] + 5. * x[6]);
    const float step = (float)(360.0 / (double)kRingLen);
    float angle = step * (float)r.t;
    if (c3 != 0.0f) angle = (float)((double)angle + (double)step * ((double)c2 / (2.0 * (double)c3) - 4.0));
    angle = (float)(360.0 - (double)angle);
    if (r.m == 1) angle = (float)((double)angle + 180.0);
    if (angle >= 360.0f) angle = (float)((double)angle - 360.0);
    return angle;
}

// ------------------------------------------------------------------ narrowed search
// Proposals (ref, mirror, psi bin, shift) -> work items.  Proposals with the
// same (ref, shift) share one item whose mask is the union of their (mirror,
// psi) windows: bit m*256 + t for the psi samples t within psi_half degrees of
// the bin centre (b + 0.5) * 360 / n_bins.  smask[p] marks every shift inside
// the (2 hw + 1)^2 windows (for the masked polar pass).  Invalid proposals
// (ref or shift out of range) are dropped.  One thread per particle.
__global__ void k_prop_items(const int4 *__restrict__ prop, int n_img, int K, int N, int nx,
                             int ny, int hw, int n_bins, float psi_half, PropItem *__restrict__ items,
                             unsigned long long *__restrict__ smask) {
    const int p = blockIdx.x * blockDim.x + threadIdx.x;
    if (p >= n_img) return;
    PropItem it[kMaxProposals];
    int cnt = 0;
    unsigned long long mk = 0ull;
    const double step = 360.0 / (double)kRingLen, bw = 360.0 / (double)n_bins;
    for (int k = 0; k < K; ++k) {
        const int4 q = prop[(size_t)p * K + k];
        const int ref = q.x, m = q.y, b = q.z, sh = q.w;
        if (ref < 0 || ref >= N || m < 0 || m > 1 || b < 0 || b >= n_bins || sh < 0 || sh >= nx * ny)
            continue;
        int j = 0;
        while (j < cnt && !(it[j].ref == ref && it[j].s == sh)) ++j;
        if (j == cnt) {
            it[j].ref = ref;
            it[j].s = sh;
            for (int w = 0; w < 16; ++w) it[j].mask[w] = 0u;
            ++cnt;
        }
        const double c = ((double)b + 0.5) * bw;
        const int t0 = (int)ceil((c - (double)psi_half) / step);
        const int t1 = (int)floor((c + (double)psi_half) / step);
        for (int t = t0; t <= t1 && t < t0 + kRingLen; ++t) {
            const int tt = ((t % kRingLen) + kRingLen) % kRingLen;
            const int bit = m * kRingLen + tt;
            it[j].mask[bit >> 5] |= 1u << (bit & 31);
        }
        const int ix0 = sh / ny, iy0 = sh % ny;
        for (int ix = max(0, ix0 - hw); ix <= min(nx - 1, ix0 + hw); ++ix)
            for (int iy = max(0, iy0 - hw); iy <= min(ny - 1, iy0 + hw); ++iy)
                mk |= 1ull << (ix * ny + iy);
    }
    for (int j = 0; j < K; ++j) {
        PropItem o;
        if (j < cnt) {
            o = it[j];
        } else {
            o.ref = -1;
            o.s = 0;
            for (int w = 0; w < 16; ++w) o.mask[w] = 0u;
        }
        items[(size_t)p * K + j] = o;
    }
    smask[p] = mk;
}

// One wavefront per item: the shifts of the item's (2 hw + 1)^2 window (in
// increasing grid order) three at a time against its reference, with exactly
// the arithmetic of the exhaustive kernels (ring contraction, psi FFT), the
// maximum taken over the item's (mirror, psi) mask only.  Workgroup = one
// particle's K <= 8 items (K wavefronts), so the spectrum rows its windows share
// are read through one CU and one L2.
__global__ __launch_bounds__(512) void k_ccf_prop(
    const float2 *__restrict__ spec, const float2 *__restrict__ norm,
    const float2 *__restrict__ rspec, const float *__restrict__ rsum0,
    const PropItem *__restrict__ items, int K, int n_items, int S, int R, int N, int nx, int ny,
    int hw, const float2 *__restrict__ twg, PeakRec *__restrict__ rec) {
    __shared__ float2 tw[256];
    __shared__ float2 fbuf[kMaxProposals][512];
    __shared__ float zbuf[kMaxProposals][256];
    __shared__ float nbuf[kMaxProposals][2][8];
    const int tid = threadIdx.x, lane = tid & 63, wave = tid >> 6;
    fill_fft_twiddles(tw, twg, tid, blockDim.x);
    __syncthreads();
    const int item = blockIdx.x * K + wave;
    if (item >= n_items) return;
    const int p = item / K;
    const PropItem &it = items[item];
    float bv[2];
    int bf[2];
    bv[0] = bv[1] = -3.40282e+38f;
    bf[0] = bf[1] = INT_MAX;
    const int n = it.ref;
    if (n >= 0 && n < N) {
        unsigned allow = 0u;
#pragma unroll
        for (int m = 0; m < 2; ++m)
#pragma unroll
            for (int q = 0; q < 4; ++q) {
                const unsigned wd = it.mask[m * 8 + 2 * q + (lane >> 5)];
                allow |= ((wd >> (lane & 31)) & 1u) << (m * 4 + q);
            }
        // window rows j = 0..nr-1 in increasing grid order
        const int ix0 = it.s / ny, iy0 = it.s % ny;
        const int ixa = max(0, ix0 - hw), ixb = min(nx - 1, ix0 + hw);
        const int iya = max(0, iy0 - hw), iyb = min(ny - 1, iy0 + hw);
        const int wy = iyb - iya + 1, nr = (ixb - ixa + 1) * wy;
        auto row = [&](int j) { return (ixa + j / wy) * ny + iya + j % wy; };
        const float rs0 = rsum0[n];
        const float2 *rb0 = rspec + (size_t)n * R * kSlots;
        const size_t img_stride = (size_t)S * R * kSlots;
        for (int r0 = 0; r0 < nr; r0 += 3) {
            float acc[3][2][4];
#pragma unroll
            for (int r = 0; r < 3; ++r)
#pragma unroll
                for (int h = 0; h < 2; ++h)
#pragma unroll
                    for (int e = 0; e < 4; ++e) acc[r][h][e] = 0.0f;
            const float2 *arow[3];
#pragma unroll
            for (int r = 0; r < 3; ++r)
                arow[r] = spec + (size_t)p * img_stride + (size_t)row(min(r0 + r, nr - 1)) * R * kSlots;
            for (int i = 0; i < R; ++i) {
                float2 b0v, b1v;
                load_slots(rb0 + i * kSlots, lane, b0v, b1v);
                float2 a0s[3], a1s[3];
#pragma unroll
                for (int r = 0; r < 3; ++r) load_slots(arow[r] + i * kSlots, lane, a0s[r], a1s[r]);
                __builtin_amdgcn_sched_barrier(0);   // all of the ring's loads in flight before the FMAs
#pragma unroll
                for (int r = 0; r < 3; ++r) {
                    const float2 a0 = a0s[r], a1 = a1s[r];
                    acc[r][0][0] += b0v.x * a0.x;
                    acc[r][0][1] += b0v.y * a0.y;
                    acc[r][0][2] += b0v.x * a0.y;
                    acc[r][0][3] += b0v.y * a0.x;
                    acc[r][1][0] += b1v.x * a1.x;
                    acc[r][1][1] += b1v.y * a1.y;
                    acc[r][1][2] += b1v.x * a1.y;
                    acc[r][1][3] += b1v.y * a1.x;
                }
            }
#pragma unroll
            for (int r = 0; r < 3; ++r) {
                if (r0 + r >= nr) break;
                const int s = row(r0 + r);
                psi_update<false, true>(acc[r][0], acc[r][1], norm[(size_t)p * S + s], rs0, s, lane,
                                        fbuf[wave], fbuf[wave] + 256, tw, zbuf[wave], nbuf[wave], bv,
                                        bf, allow);
            }
        }
    }
    write_rec(rec + item, bv, bf, nbuf[wave], lane);
}

// Is (m, s, t) of reference `ref` admitted by one of the particle's items?
__device__ __forceinline__ bool admitted(const PropItem *it, int K, int ref, int m, int s, int t,
                                         int ny, int hw) {
    const int ix = s / ny, iy = s % ny;
    const int bit = m * kRingLen + t;
    for (int k = 0; k < K; ++k) {
        if (it[k].ref != ref) continue;
        const int cx = it[k].s / ny, cy = it[k].s % ny;
        if (abs(ix - cx) <= hw && abs(iy - cy) <= hw && ((it[k].mask[bit >> 5] >> (bit & 31)) & 1u))
            return true;
    }
    return false;
}

// Edge test of a maximum (m, s, t) of reference `ref`: true when one of its
// neighbours (psi sample t +- 1, or the next shift in x or y inside the grid) is
// admitted by no item of that reference, i.e. the windows may have cut the
// search short.
__device__ __forceinline__ bool on_edge(const PropItem *it, int K, int ref, int m, int s, int t,
                                        int nx, int ny, int hw) {
    const int ix = s / ny, iy = s % ny;
    bool in = admitted(it, K, ref, m, s, (t + 1) & (kRingLen - 1), ny, hw) &&
              admitted(it, K, ref, m, s, (t + kRingLen - 1) & (kRingLen - 1), ny, hw);
    if (ix > 0) in = in && admitted(it, K, ref, m, s - ny, t, ny, hw);
    if (ix < nx - 1) in = in && admitted(it, K, ref, m, s + ny, t, ny, hw);
    if (iy > 0) in = in && admitted(it, K, ref, m, s - 1, t, ny, hw);
    if (iy < ny - 1) in = in && admitted(it, K, ref, m, s + 1, t, ny, hw);
    return !in;
}

__device__ __forceinline__ void decode_best(const PeakRec &r, int ref, int p, ShiftView acc,
                                            const float2 *shifts_f, int n_shift, size_t pp,
                                            int *best_ref, float *best_val, float *best_peak) {
    float ax = 0.0f, ay = 0.0f;
    if (acc.base) {
        ax = acc.base[(size_t)p * acc.stride];
        ay = acc.base[(size_t)p * acc.stride + 1];
    }
    const float2 sh = shifts_f[min(max(r.s, 0), n_shift - 1)];
    best_ref[pp] = ref;
    best_val[pp] = r.val;
    best_peak[4 * pp + 0] = wizard_angle(r);
    best_peak[4 * pp + 1] = ax + sh.x;
    best_peak[4 * pp + 2] = ay + sh.y;
    best_peak[4 * pp + 3] = r.m ? 1.0f : 0.0f;
}

// Best item per particle (value, then lower reference, mirror, shift, psi
// sample), decoded like the multi-reference wizard into the outputs at p_off + p;
// edge[p] = 1 when it failed the edge test (with edge_test) or no item had a
// valid sample.
__global__ void k_best_prop(const PeakRec *__restrict__ rec, const PropItem *__restrict__ items,
                            int n_img, int K, int nx, int ny, int hw, int edge_test, ShiftView acc,
                            const float2 *__restrict__ shifts_f, int n_shift, int p_off,
                            int *__restrict__ best_ref, float *__restrict__ best_val,
                            float *__restrict__ best_peak, int *__restrict__ edge) {
    const int p = blockIdx.x * blockDim.x + threadIdx.x;
    if (p >= n_img) return;
    int bk = -1;
    for (int k = 0; k < K; ++k) {
        const PropItem &it = items[(size_t)p * K + k];
        const PeakRec &r = rec[(size_t)p * K + k];
        if (it.ref < 0 || !(r.val > -3.40282e+38f)) continue;
        if (bk < 0) {
            bk = k;
            continue;
        }
        const PeakRec &b = rec[(size_t)p * K + bk];
        const int br = items[(size_t)p * K + bk].ref;
        const bool better = r.val > b.val ||
                            (r.val == b.val &&
                             (it.ref < br || (it.ref == br && (r.m < b.m || (r.m == b.m &&
                              (r.s < b.s || (r.s == b.s && r.t < b.t)))))));
        if (better) bk = k;
    }
    const size_t pp = (size_t)p_off + p;
    if (bk < 0) {
        best_ref[pp] = -1;
        best_val[pp] = -3.40282e+38f;
        best_peak[4 * pp] = best_peak[4 * pp + 1] = best_peak[4 * pp + 2] = best_peak[4 * pp + 3] = 0.0f;
        edge[p] = 1;
        return;
    }
    PeakRec r = rec[(size_t)p * K + bk];
    r.m = min(max(r.m, 0), 1);
    r.s = min(max(r.s, 0), n_shift - 1);
    r.t &= kRingLen - 1;
    const PropItem *it = items + (size_t)p * K;
    const int ref = it[bk].ref;
    decode_best(r, ref, p, acc, shifts_f, n_shift, pp, best_ref, best_val, best_peak);
    edge[p] = (edge_test && on_edge(it, K, ref, r.m, r.s, r.t, nx, ny, hw)) ? 1 : 0;
}

// Second level: the distinct valid references among particle idx[j]'s
// proposals -> cand[j][0..K), padded with -1 (reference 0 if none is valid).
__global__ void k_prop_refs(const int4 *__restrict__ prop, const int *__restrict__ idx, int n, int K,
                            int N, int *__restrict__ cand) {
    const int j = blockIdx.x * blockDim.x + threadIdx.x;
    if (j >= n) return;
    const int p = idx[j];
    int c[kMaxProposals];
    int cnt = 0;
    for (int k = 0; k < K; ++k) {
        const int r = prop[(size_t)p * K + k].x;
        if (r < 0 || r >= N) continue;
        bool seen = false;
        for (int q = 0; q < cnt; ++q) seen = seen || c[q] == r;
        if (!seen) c[cnt++] = r;
    }
    if (cnt == 0) c[cnt++] = 0;
    for (int k = 0; k < K; ++k) cand[(size_t)j * K + k] = k < cnt ? c[k] : -1;
}

// Second-level shifts of a batch: for the particles flagged 1 (flag[p]), every
// shift the first level did not compute (done[p]); none for the others.
__global__ void k_level2_mask(const int *__restrict__ flag, int n, int S,
                              const unsigned long long *__restrict__ done,
                              unsigned long long *__restrict__ smask) {
    const int p = blockIdx.x * blockDim.x + threadIdx.x;
    if (p >= n) return;
    const unsigned long long full = S >= 64 ? ~0ull : ((1ull << S) - 1ull);
    smask[p] = flag[p] == 1 ? (full & ~done[p]) : 0ull;
}

// Best of the second-level records rec[j][K] (value, then lower reference) of
// batch particle idx[j], decoded into its outputs at p_off + idx[j];
// fallback = value below threshold.
__global__ void k_best_refs_level2(const PeakRec *__restrict__ rec, const int *__restrict__ cand,
                                   int n, int K, ShiftView acc,
                                   const float2 *__restrict__ shifts_f, int n_shift,
                                   const int *__restrict__ idx, int p_off, float threshold,
                                   int *__restrict__ best_ref, float *__restrict__ best_val,
                                   float *__restrict__ best_peak, int *__restrict__ fallback) {
    const int j = blockIdx.x * blockDim.x + threadIdx.x;
    if (j >= n) return;
    int bk = 0;   // entry 0 always holds a reference
    for (int k = 1; k < K; ++k) {
        if (cand[(size_t)j * K + k] < 0) continue;
        const float v = rec[(size_t)j * K + k].val, bv = rec[(size_t)j * K + bk].val;
        if (v > bv || (v == bv && cand[(size_t)j * K + k] < cand[(size_t)j * K + bk])) bk = k;
    }
    const PeakRec r = rec[(size_t)j * K + bk];
    const int p = idx[j];
    float2 a = f2(0.0f, 0.0f);
    if (acc.base) a = f2(acc.base[(size_t)p * acc.stride], acc.base[(size_t)p * acc.stride + 1]);
    const float2 sh = shifts_f[min(max(r.s, 0), n_shift - 1)];
    const size_t pp = (size_t)p_off + p;
    best_ref[pp] = cand[(size_t)j * K + bk];
    best_val[pp] = r.val;
    best_peak[4 * pp + 0] = wizard_angle(r);
    best_peak[4 * pp + 1] = a.x + sh.x;
    best_peak[4 * pp + 2] = a.y + sh.y;
    best_peak[4 * pp + 3] = r.m ? 1.0f : 0.0f;
    fallback[pp] = (r.val >= threshold) ? 0 : 1;
}

// fallback = 2 when the proposal confidence is below conf_min (straight to the
// exhaustive search), else 1 for an edge maximum or a value below the threshold
// (NaN included), else 0.
__global__ void k_fallback_flags(const float *__restrict__ best_val, const int *__restrict__ edge,
                                 const float *__restrict__ conf, float conf_min, int n_img,
                                 float threshold, int p_off, int *__restrict__ fallback) {
    const int p = blockIdx.x * blockDim.x + threadIdx.x;
    if (p >= n_img) return;
    const size_t pp = (size_t)p_off + p;
    int f = (edge[p] || !(best_val[pp] >= threshold)) ? 1 : 0;
    if (conf && !(conf[pp] >= conf_min)) f = 2;
    fallback[pp] = f;
}

// Indices of the particles whose flag lies in [lo, hi] (any order).
__global__ void k_compact(const int *__restrict__ flag, int n, int lo, int hi, int *__restrict__ idx,
                          int *__restrict__ count) {
    const int p = blockIdx.x * blockDim.x + threadIdx.x;
    if (p >= n || flag[p] < lo || flag[p] > hi) return;
    idx[atomicAdd(count, 1)] = p;
}

// Copies the images (and accumulated shifts) of the listed particles.
__global__ void k_gather(const float *__restrict__ imgs, int D, ShiftView acc,
                         const int *__restrict__ idx, int n, float *__restrict__ out,
                         float2 *__restrict__ out_acc) {
    const int j = blockIdx.x;
    if (j >= n) return;
    const int p = idx[j];
    const float *src = imgs + (size_t)p * D * D;
    float *dst = out + (size_t)j * D * D;
    for (int i = threadIdx.x; i < D * D; i += blockDim.x) dst[i] = src[i];
    if (threadIdx.x == 0)
        out_acc[j] = acc.base ? f2(acc.base[(size_t)p * acc.stride], acc.base[(size_t)p * acc.stride + 1])
                              : f2(0.0f, 0.0f);
}

// Exhaustive records rec[j][N] of gathered particle j -> best reference (first
// maximum over references), decoded into the outputs of particle idx[j0 + j].
__global__ void k_best_refs_scatter(const PeakRec *__restrict__ rec, int n, int N,
                                    const float2 *__restrict__ acc, const float2 *__restrict__ shifts_f,
                                    int n_shift, const int *__restrict__ idx, int *__restrict__ best_ref,
                                    float *__restrict__ best_val, float *__restrict__ best_peak) {
    const int j = blockIdx.x * blockDim.x + threadIdx.x;
    if (j >= n) return;
    int bn = 0;
    for (int k = 1; k < N; ++k)
        if (rec[(size_t)j * N + k].val > rec[(size_t)j * N + bn].val) bn = k;
    const PeakRec r = rec[(size_t)j * N + bn];
    const float2 a = acc[j];
    const float2 sh = shifts_f[min(max(r.s, 0), n_shift - 1)];
    const size_t pp = (size_t)idx[j];
    best_ref[pp] = bn;
    best_val[pp] = r.val;
    best_peak[4 * pp + 0] = wizard_angle(r);
    best_peak[4 * pp + 1] = a.x + sh.x;
    best_peak[4 * pp + 2] = a.y + sh.y;
    best_peak[4 * pp + 3] = r.m ? 1.0f : 0.0f;
}

// Best candidate per particle -> reference index, peak value, {angle, sx, sy, mirror}.
__global__ void k_decode_best(const PeakRec *__restrict__ rec, const int *__restrict__ cand,
                              const int *__restrict__ best, const int *__restrict__ tie, int n_img,
                              int K, ShiftView acc, const float2 *__restrict__ shifts_f,
                              int n_shift, int p_off, int *__restrict__ best_ref,
                              float *__restrict__ best_val, float *__restrict__ best_peak,
                              int *__restrict__ tie_out, const int *__restrict__ idx) {
    const int p = blockIdx.x * blockDim.x + threadIdx.x;
    if (p >= n_img) return;
    const int k = max(best[p], 0);
    const PeakRec r = rec[(size_t)p * K + k];
    float ax = 0.0f, ay = 0.0f;
    if (acc.base) {
        ax = acc.base[(size_t)p * acc.stride];
        ay = acc.base[(size_t)p * acc.stride + 1];
    }
    const float2 sh = shifts_f[min(max(r.s, 0), n_shift - 1)];
    // output slot: p_off + p, or idx[p_off + p] (a scatter to the caller's particles)
    const size_t pp = idx ? (size_t)idx[(size_t)p_off + p] : (size_t)p_off + p;
    best_ref[pp] = cand[(size_t)p * K + k];
    best_val[pp] = r.val;
    best_peak[4 * pp + 0] = wizard_angle(r);
    best_peak[4 * pp + 1] = ax + sh.x;
    best_peak[4 * pp + 2] = ay + sh.y;
    best_peak[4 * pp + 3] = r.m ? 1.0f : 0.0f;
    if (tie_out) tie_out[pp] = tie[p];
}

// Multi-reference decode (reference wizard, cuda/gpu_aln_noref.cu:1758-1793, with
// its float/double mix) and output scatter (:517-540).
__global__ void k_decode_mref(const PeakRec *__restrict__ rec, int n_img, int N, ShiftView acc,
                              const float2 *__restrict__ shifts_f, int n_shift, float *__restrict__ peak,
                              int peak_row, int p_off, float *__restrict__ d, int d_pitch) {
    const int idx = blockIdx.x * blockDim.x + threadIdx.x;
    if (idx >= n_img * N) return;
    const int p = idx / N, n = idx % N;
    const PeakRec r = rec[idx];
    const float angle = wizard_angle(r);
    float ax = 0.0f, ay = 0.0f;
    if (acc.base) {
        ax = acc.base[(size_t)p * acc.stride];
        ay = acc.base[(size_t)p * acc.stride + 1];
    }
    const float2 sh = shifts_f[min(max(r.s, 0), n_shift - 1)];
    const size_t pp = (size_t)p_off + p;
    d[(size_t)n * d_pitch + pp] = r.val;
    float *pk = peak + (size_t)n * peak_row + 4 * pp;
    pk[0] = angle;
    pk[1] = ax + sh.x;
    pk[2] = ay + sh.y;
    pk[3] = r.m ? 1.0f : 0.0f;
}

// Pre-align / reference-free decode (cuda/gpu_aln_noref.cu:2486-2519, FP64
// parabola :2558-2581).
__global__ void k_decode_single(const PeakRec *__restrict__ rec, int cnt, DevParam *__restrict__ prm,
                                const float2 *__restrict__ shifts_f, int n_shift, float lim) {
    const int i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= cnt) return;
    const PeakRec r = rec[i];
    DevParam q = prm[i];
    double x[7];
#pragma unroll
    for (int k = 0; k < 7; ++k) x[k] = (double)r.x[k];
    const double c2 = 49. * x[0] + 6. * x[1] - 21. * x[2] - 32. * x[3] - 27. * x[4] - 6. * x[5] + 31. * x[6];
    const double c3 = 5. * x[0] - 3. * x[2] - 4. * x[3] - 3. * x[4] + 5. * x[6];
    const double step = 360.0 / (double)kRingLen;
    double interp = step * (double)r.t;
    if (c3 != 0.0) interp = interp + step * (c2 / (2.0 * c3) - 4.0);
    q.mirror = (r.m == 1);
    const float2 sh = shifts_f[min(max(r.s, 0), n_shift - 1)];
    q.shift_x += sh.x;
    q.shift_y += sh.y;
    q.shift_x = fminf(fmaxf(q.shift_x, -lim), lim);
    q.shift_y = fminf(fmaxf(q.shift_y, -lim), lim);
    q.angle = (float)(360.0 - interp);
    if (q.mirror) {
        q.angle = (float)((double)q.angle + 180.0);
        if (q.angle >= 360.0f) q.angle = (float)((double)q.angle - 360.0);
    }
    prm[i] = q;
}

// ---------------------------------------------------- reference-free support
// cu_transform_batch (cuda/gpu_aln_noref.cu:1558-1588): output pixel -> mirror
// (x -> D - x) -> rotate by +angle about D/2 -> + accumulated shift -> bilinear.
__global__ __launch_bounds__(256) void k_transform(const float *__restrict__ imgs, int D,
                                                   const DevParam *__restrict__ prm,
                                                   float *__restrict__ out) {
    const int p = blockIdx.x;
    const DevParam q = prm[p];
    const float c = (float)(D / 2);
    const float ang = (3.141592653589793f / 180.0f) * q.angle;
    const float sa = sinf(ang), ca = cosf(ang);
    const float *img = imgs + (size_t)p * D * D;
    for (int idx = threadIdx.x; idx < D * D; idx += blockDim.x) {
        const int x = idx % D, y = idx / D;
        float sx = q.mirror ? (float)(D - x) : (float)x;
        float sy = (float)y;
        sx -= c;
        sy -= c;
        const float tx = sx;
        sx = tx * ca - sy * sa;
        sy = tx * sa + sy * ca;
        sx += c;
        sy += c;
        out[(size_t)p * D * D + idx] = bilinear(img, D, sx + q.shift_x, sy + q.shift_y);
    }
}

// cu_average_batch (cuda/gpu_aln_noref.cu:1605-1621).
__global__ __launch_bounds__(256) void k_class_average(const float *__restrict__ tr, int D,
                                                       const int *__restrict__ cid_idx,
                                                       float *__restrict__ refs) {
    const int k = blockIdx.x;
    const int a = cid_idx[k], b = cid_idx[k + 1];
    if (b <= a) return;  // a class without particles keeps its reference
    for (int pix = threadIdx.x; pix < D * D; pix += blockDim.x) {
        float s = 0.0f;
        for (int i = a; i < b; ++i) s += tr[(size_t)i * D * D + pix];
        refs[(size_t)k * D * D + pix] = s / (float)(b - a);
    }
}

// ------------------------------------------------------ pre-alignment loop
// The average of every particle transformed with its current parameters (the
// image transform of k_transform), without materialising the transformed stack:
// per particle the rotation (cos, sin), shifts and mirror; partial sums over
// chunks of kPreChunk particles per pixel; then the chunks in order, divided by P.
constexpr int kPreChunk = 256;

__global__ __launch_bounds__(256) void k_pre_tp(const DevParam *__restrict__ prm, int P,
                                                float4 *__restrict__ tp, int *__restrict__ mir) {
    const int p = blockIdx.x * blockDim.x + threadIdx.x;
    if (p >= P) return;
    const DevParam q = prm[p];
    const float ang = (3.141592653589793f / 180.0f) * q.angle;   // as k_transform
    tp[p] = make_float4(cosf(ang), sinf(ang), q.shift_x, q.shift_y);
    mir[p] = q.mirror ? 1 : 0;
}

__global__ __launch_bounds__(256) void k_pre_partial(const float *__restrict__ imgs, int D, int P,
                                                     const float4 *__restrict__ tp,
                                                     const int *__restrict__ mir,
                                                     float *__restrict__ part) {
    const int idx = blockIdx.x * blockDim.x + threadIdx.x;
    const int c = blockIdx.y;
    if (idx >= D * D) return;
    const int x = idx % D, y = idx / D;
    const float cen = (float)(D / 2);
    const int a = c * kPreChunk, b = min(a + kPreChunk, P);
    float s = 0.0f;
    for (int p = a; p < b; ++p) {
        const float4 t = tp[p];
        float sx = mir[p] ? (float)(D - x) : (float)x;
        float sy = (float)y;
        sx -= cen;
        sy -= cen;
        const float tx = sx;
        sx = tx * t.x - sy * t.y;
        sy = tx * t.y + sy * t.x;
        sx += cen;
        sy += cen;
        s += bilinear(imgs + (size_t)p * D * D, D, sx + t.z, sy + t.w);
    }
    part[(size_t)c * D * D + idx] = s;
}

__global__ __launch_bounds__(256) void k_pre_finish(const float *__restrict__ part, int D, int P,
                                                    int n_chunks, float *__restrict__ avg) {
    const int idx = blockIdx.x * blockDim.x + threadIdx.x;
    if (idx >= D * D) return;
    double s = 0.0;
    for (int c = 0; c < n_chunks; ++c) s += (double)part[(size_t)c * D * D + idx];
    avg[idx] = (float)(s / (double)P);
}

// Centring shift of the average from the particles' shifts after the driver's
// rotation (bin/sp_isac_applications.py:1664-1690): sx_sum = sum of +-shift_x'
// (minus for mirrored particles), sy_sum = sum of shift_y'; the average is then
// moved by (-sx_sum / P, -sy_sum / P) (:1547-1552).  One workgroup, FP64 sums.
__global__ __launch_bounds__(256) void k_pre_center(const DevParam *__restrict__ prm, int P,
                                                    float2 *__restrict__ shift) {
    __shared__ double rx[256], ry[256];
    double sx = 0.0, sy = 0.0;
    for (int p = threadIdx.x; p < P; p += blockDim.x) {
        const DevParam q = prm[p];
        const double a = (double)q.angle * (M_PI / 180.0);
        const double c = cos(a), sn = -sin(a);
        const double nx = -(double)q.shift_x, ny = -(double)q.shift_y;
        const double px = nx * c - ny * sn, py = nx * sn + ny * c;
        sx += q.mirror ? -px : px;
        sy += py;
    }
    rx[threadIdx.x] = sx;
    ry[threadIdx.x] = sy;
    __syncthreads();
    for (int o = 128; o > 0; o >>= 1) {
        if ((int)threadIdx.x < o) {
            rx[threadIdx.x] += rx[threadIdx.x + o];
            ry[threadIdx.x] += ry[threadIdx.x + o];
        }
        __syncthreads();
    }
    if (threadIdx.x == 0) shift[0] = make_float2((float)(-rx[0] / P), (float)(-ry[0] / P));
}

// ------------------------------------------------------- reference update
// Device form of the multi-reference driver's reference update
// (bin/sp_isac2_gpu.py:1146-1232): every particle is transformed with the
// parameters of its class's peak and added to that class's average, the average
// shift of each class centres the new reference, and each particle's shifts are
// re-expressed about that centre.

// Per particle i of class k = cls[i]: the peak (angle, sx, sy, mirror) of
// peak_list[k][4i..4i+3] as the library writes it.  tp[i] = (cos, sin, sx, sy) for
// the image transform of k_transform; post[i] = the shifts after the driver's
// rotation (bin/sp_isac2_gpu.py:1016-1022, in FP32 like the driver's numpy code).
// A class id outside [0, N) sets *bad and gives the particle neutral values.
__global__ void k_upd_params(const float *__restrict__ peak, int P, int N,
                             const int *__restrict__ cls, float4 *__restrict__ tp,
                             int *__restrict__ mir, float2 *__restrict__ post,
                             float4 *__restrict__ raw, int *__restrict__ bad) {
    const int i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= P) return;
    const int k = cls[i];
    if (k < 0 || k >= N) {
        atomicOr(bad, 1);
        tp[i] = make_float4(1.0f, 0.0f, 0.0f, 0.0f);
        mir[i] = 0;
        post[i] = f2(0.0f, 0.0f);
        raw[i] = make_float4(0.0f, 0.0f, 0.0f, 0.0f);
        return;
    }
    const float *q = peak + (size_t)k * 4 * P + 4 * (size_t)i;
    const float a = q[0], sx = q[1], sy = q[2], m = q[3];
    const float ang = (3.141592653589793f / 180.0f) * a;   // as k_transform
    tp[i] = make_float4(cosf(ang), sinf(ang), sx, sy);
    mir[i] = (m != 0.0f) ? 1 : 0;
    raw[i] = make_float4(a, sx, sy, m);
    const float rad = a * (float)(3.141592653589793 / 180.0);
    const float c = cosf(rad), sn = -sinf(rad);
    const float sxn = -sx, syn = -sy;
    post[i] = f2(sxn * c - syn * sn, sxn * sn + syn * c);
}

// Block k: the particles of class k in ascending order.  pass 0 counts them into
// cnt[k]; pass 1 writes them to members[off[k] ...].
template <int PASS>
__global__ __launch_bounds__(256) void k_class_members(const int *__restrict__ cls, int P,
                                                       int *__restrict__ cnt,
                                                       const int *__restrict__ off,
                                                       int *__restrict__ members) {
    __shared__ int wsum[4];
    const int k = blockIdx.x, tid = threadIdx.x, lane = tid & 63, w = tid >> 6;
    int base = PASS ? off[k] : 0;
    for (int p0 = 0; p0 < P; p0 += 256) {
        const int i = p0 + tid;
        const bool hit = i < P && cls[i] == k;
        const unsigned long long m = __ballot(hit);
        if (lane == 0) wsum[w] = __popcll(m);
        __syncthreads();
        int before = 0, total = 0;
#pragma unroll
        for (int v = 0; v < 4; ++v) {
            before += (v < w) ? wsum[v] : 0;
            total += wsum[v];
        }
        if (PASS && hit) {
            const unsigned long long lt = (lane == 0) ? 0ull : (m & (~0ull >> (64 - lane)));
            members[base + before + __popcll(lt)] = i;
        }
        base += total;
        __syncthreads();
    }
    if (!PASS && tid == 0) cnt[k] = base;
}

// Exclusive scan of cnt[0..N) into off[0..N] (one block).
__global__ void k_scan_counts(const int *__restrict__ cnt, int N, int *__restrict__ off) {
    if (threadIdx.x != 0) return;
    int s = 0;
    for (int k = 0; k < N; ++k) {
        off[k] = s;
        s += cnt[k];
    }
    off[N] = s;
}

// One thread per class: the shift sums of the driver (mirrored particles enter
// with -sx), summed in double in ascending particle order, rounded to float as the
// driver's MPI_FLOAT reduction does, divided by the member count in double.
// neg[k] = -(avg_x, avg_y) (the centring shift of the new reference); avg_out and
// members_out (either may be NULL) receive (avg_x, avg_y) as floats and the counts.
__global__ void k_class_shift(const int *__restrict__ off, const int *__restrict__ members, int N,
                              const float2 *__restrict__ post, const int *__restrict__ mir,
                              double2 *__restrict__ avg, float2 *__restrict__ neg,
                              float2 *__restrict__ avg_out, int *__restrict__ members_out) {
    const int k = blockIdx.x * blockDim.x + threadIdx.x;
    if (k >= N) return;
    const int a = off[k], b = off[k + 1];
    double sx = 0.0, sy = 0.0;
    for (int j = a; j < b; ++j) {
        const int i = members[j];
        const float2 q = post[i];
        sx += mir[i] ? -(double)q.x : (double)q.x;
        sy += (double)q.y;
    }
    double ax = 0.0, ay = 0.0;
    if (b > a) {
        ax = (double)(float)sx / (double)(b - a);
        ay = (double)(float)sy / (double)(b - a);
    }
    avg[k] = make_double2(ax, ay);
    neg[k] = f2((float)-ax, (float)-ay);
    if (avg_out) avg_out[k] = f2((float)ax, (float)ay);
    if (members_out) members_out[k] = b - a;
}

// Per particle: (angle, sx -/+ avg_x, sy - avg_y, mirror) in the driver's
// convention (bin/sp_isac2_gpu.py:1200-1208).
__global__ void k_upd_params_out(const int *__restrict__ cls, int P, int N,
                                 const float4 *__restrict__ raw, const float2 *__restrict__ post,
                                 const int *__restrict__ mir, const double2 *__restrict__ avg,
                                 float4 *__restrict__ out) {
    const int i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= P) return;
    const int k = cls[i];
    if (k < 0 || k >= N) {
        out[i] = make_float4(0.0f, 0.0f, 0.0f, 0.0f);
        return;
    }
    const double2 g = avg[k];
    const float2 q = post[i];
    const double x = mir[i] ? (double)q.x + g.x : (double)q.x - g.x;
    out[i] = make_float4(raw[i].x, (float)x, (float)((double)q.y - g.y), raw[i].w);
}

// Block (tile, k): pixels of class k's average.  Each member image is transformed
// as k_transform does and added in ascending particle order; the sum is scaled
// by 1 / members (a class without members gives zeros).
__global__ __launch_bounds__(256) void k_class_transform_average(
    const float *__restrict__ imgs, int D, const int *__restrict__ off,
    const int *__restrict__ members, const float4 *__restrict__ tp, const int *__restrict__ mir,
    float *__restrict__ refs) {
    const int k = blockIdx.y;
    const int idx = blockIdx.x * blockDim.x + threadIdx.x;
    if (idx >= D * D) return;
    const int a = off[k], b = off[k + 1];
    const float c = (float)(D / 2);
    const int x = idx % D, y = idx / D;
    float s = 0.0f;
    for (int j = a; j < b; ++j) {
        const int i = members[j];
        const float4 q = tp[i];
        float sx = mir[i] ? (float)(D - x) : (float)x;
        float sy = (float)y;
        sx -= c;
        sy -= c;
        const float tx = sx;
        sx = tx * q.x - sy * q.y;
        sy = tx * q.y + sy * q.x;
        sx += c;
        sy += c;
        s += bilinear(imgs + (size_t)i * D * D, D, sx + q.z, sy + q.w);
    }
    refs[(size_t)k * D * D + idx] = (b > a) ? s * (1.0f / (float)(b - a)) : 0.0f;
}

// Tangent filter, three separable passes over a half-plane spectrum
// F[n][ky][kx], kx = 0..D/2 (cuda/gpu_aln_noref.cu:1007-1037, :2203-2266).
__global__ void k_row_dft(const float *__restrict__ refs, int D, float2 *__restrict__ F,
                          const float2 *__restrict__ tw) {
    const int n = blockIdx.y, y = blockIdx.x, Dh = D / 2 + 1;
    const float *row = refs + ((size_t)n * D + y) * D;
    for (int kx = threadIdx.x; kx < Dh; kx += blockDim.x) {
        float re = 0.0f, im = 0.0f;
        for (int x = 0; x < D; ++x) {
            const float2 w = tw[(x * kx) % D];
            re += row[x] * w.x;
            im += row[x] * w.y;
        }
        F[((size_t)n * D + y) * Dh + kx] = f2(re, im);
    }
}

// Column pass: forward DFT over y, the tanh low-pass, and (shift != NULL) the phase
// ramp exp(-2 pi i (kx dx + ky dy) / D) that moves image n's content by
// shift[n] = (dx, dy) (signed ky; ky = D/2 taken as +D/2), then the inverse DFT.
// abs_freq = 0: the reference's frequency coordinates (x = kx / (D/2+1) / 2,
// y = min(ky, D-ky) / D); 1: absolute frequency in cycles per pixel
// (x = kx / D, y = min(ky, D-ky) / D).
__global__ void k_col_filter(float2 *__restrict__ F, int D, const float2 *__restrict__ tw,
                             float cutoff, float falloff, const float2 *__restrict__ shift,
                             int abs_freq) {
    extern __shared__ float2 col[];
    const int n = blockIdx.y, kx = blockIdx.x, Dh = D / 2 + 1;
    float2 *base = F + (size_t)n * D * Dh + kx;
    const float cc = 3.141592653589793f / (2.0f * falloff * cutoff);
    const float xx = abs_freq ? (float)kx / (float)D : (float)kx / (float)Dh * 0.5f;
    const float half = (float)D / 2.0f;
    const float2 sh = shift ? shift[n] : f2(0.0f, 0.0f);
    for (int ky = threadIdx.x; ky < D; ky += blockDim.x) {
        float re = 0.0f, im = 0.0f;
        for (int y = 0; y < D; ++y) {
            const float2 v = base[(size_t)y * Dh];
            const float2 w = tw[(y * ky) % D];
            re += v.x * w.x - v.y * w.y;
            im += v.x * w.y + v.y * w.x;
        }
        float yy = ((float)ky < half) ? (float)ky : (float)(D - ky);
        yy = yy / half * 0.5f;
        const float dd = sqrtf(xx * xx + yy * yy);
        // cutoff <= 0: no low-pass (the phase ramp alone)
        const float h = cutoff > 0.0f ? 0.5f * (tanhf(cc * (dd + cutoff)) - tanhf(cc * (dd - cutoff)))
                                      : 1.0f;
        re *= h;
        im *= h;
        if (shift) {
            const int kys = (2 * ky > D) ? ky - D : ky;
            float t = ((float)kx * sh.x + (float)kys * sh.y) / (float)D;   // turns
            t -= rintf(t);
            float sn, cs;
            sincospif(-2.0f * t, &sn, &cs);
            const float r2 = re * cs - im * sn;
            im = re * sn + im * cs;
            re = r2;
        }
        col[ky] = f2(re, im);
    }
    __syncthreads();
    for (int y = threadIdx.x; y < D; y += blockDim.x) {
        float re = 0.0f, im = 0.0f;
        for (int ky = 0; ky < D; ++ky) {
            const float2 v = col[ky];
            const float2 w = tw[(y * ky) % D];  // conj for the inverse
            re += v.x * w.x + v.y * w.y;
            im += v.y * w.x - v.x * w.y;
        }
        base[(size_t)y * Dh] = f2(re, im);
    }
}

// Row C2R.  gain > 0 scales every image by gain; otherwise the reference's
// scaling: 1 / (D*D)^2 for image 0 only (quirk of cuda/gpu_aln_noref.cu:2235-2262).
__global__ void k_row_c2r(float *__restrict__ refs, int D, const float2 *__restrict__ F,
                          const float2 *__restrict__ tw, float gain_all) {
    const int n = blockIdx.y, y = blockIdx.x, Dh = D / 2 + 1;
    const float2 *row = F + ((size_t)n * D + y) * Dh;
    const float dd = (float)D * (float)D;
    const float gain = gain_all > 0.0f ? gain_all : ((n == 0) ? 1.0f / (dd * dd) : 1.0f);
    for (int x = threadIdx.x; x < D; x += blockDim.x) {
        float s = 0.0f;
        for (int kx = 0; kx < Dh; ++kx) {
            const float2 v = row[kx];
            const float2 w = tw[(x * kx) % D];  // e^{+i}: (w.x, -w.y)
            const float re = v.x * w.x + v.y * w.y;
            const bool edge = (kx == 0) || ((D % 2 == 0) && kx == D / 2);
            s += edge ? re : 2.0f * re;
        }
        refs[((size_t)n * D + y) * D + x] = s * gain;
    }
}

}  // namespace

// ---------------------------------------------------------------- launchers

hipError_t prepare_launches() {
    static bool lds_raised = false;
    if (!lds_raised) {
        const int lds = (int)((256 + 4 * 512) * sizeof(float2) +
                              (kMaxImgDim + 2) * (kMaxImgDim + 2) * sizeof(float));
        hipError_t e = hipFuncSetAttribute((const void *)k_polar_spectra<false>,
                                           hipFuncAttributeMaxDynamicSharedMemorySize, lds);
        if (e != hipSuccess) return e;
        e = hipFuncSetAttribute((const void *)k_polar_spectra<true>,
                                hipFuncAttributeMaxDynamicSharedMemorySize, lds);
        if (e != hipSuccess) return e;
        const int lds_frag = (int)((256 + 4 * 512) * sizeof(float2) +
                                   (kMaxImgDim + 2 * kFragBorder) * (kMaxImgDim + 2 * kFragBorder) *
                                       sizeof(float));
        e = hipFuncSetAttribute((const void *)k_polar_frag<false>,
                                hipFuncAttributeMaxDynamicSharedMemorySize, lds_frag);
        if (e != hipSuccess) return e;
        e = hipFuncSetAttribute((const void *)k_polar_frag<true>,
                                hipFuncAttributeMaxDynamicSharedMemorySize, lds_frag);
        if (e != hipSuccess) return e;
        lds_raised = true;
    }
    return hipSuccess;
}

hipError_t launch_polar_spectra(const float *imgs, int n_img, int img_dim, ShiftView acc,
                                const int2 *shift_i, int n_shift, const float2 *tmpl, int n_ring,
                                const float2 *twiddle, float2 *spec, float2 *norm,
                                hipStream_t st) {
    if (n_img <= 0 || n_shift <= 0) return hipSuccess;
    const size_t lds = (256 + 4 * 512) * sizeof(float2) + (size_t)(img_dim + 2) * (img_dim + 2) * sizeof(float);
    if (hipError_t e = prepare_launches()) return e;
    dim3 grid(n_img);
    hipLaunchKernelGGL(k_polar_spectra<false>, grid, dim3(256), lds, st, imgs, img_dim, acc,
                       shift_i, n_shift, tmpl, n_ring, twiddle, spec, norm, nullptr);
    return hipGetLastError();
}

bool polar_frag_supported(int n_ring) { return n_ring >= 1 && n_ring <= kFragMaxRings; }

hipError_t launch_polar_frag(const float *imgs, int n_img, int img_dim, ShiftView acc,
                             const int2 *shift_i, int n_shift, int n_ring, bool bf16,
                             const float2 *twiddle, void *frag, hipStream_t st) {
    if (n_img <= 0 || n_shift <= 0) return hipSuccess;
    if (!polar_frag_supported(n_ring)) return hipErrorInvalidValue;
    const GemmLayout L = gemm_layout(n_ring);
    const int S_pad = gemm_rows_per_particle(n_shift);
    const size_t lds = (256 + 4 * 512) * sizeof(float2) +
                       (size_t)(img_dim + 2 * kFragBorder) * (img_dim + 2 * kFragBorder) * sizeof(float);
    if (hipError_t e = prepare_launches()) return e;
    const dim3 grid((unsigned)((n_img + 15) / 16 * 16));
    if (bf16)
        hipLaunchKernelGGL(k_polar_frag<true>, grid, dim3(256), lds, st, imgs, n_img, img_dim, acc,
                           shift_i, n_shift, S_pad, n_ring, L.full16, L.tail, L.slot_bytes, L.cp,
                           twiddle, (unsigned char *)frag);
    else
        hipLaunchKernelGGL(k_polar_frag<false>, grid, dim3(256), lds, st, imgs, n_img, img_dim, acc,
                           shift_i, n_shift, S_pad, n_ring, L.full16, L.tail, L.slot_bytes, L.cp,
                           twiddle, (unsigned char *)frag);
    return hipGetLastError();
}

hipError_t launch_polar_spectra_masked(const float *imgs, int n_img, int img_dim, ShiftView acc,
                                       const int2 *shift_i, int n_shift, const float2 *tmpl,
                                       int n_ring, const float2 *twiddle, float2 *spec,
                                       float2 *norm, const unsigned long long *smask,
                                       hipStream_t st) {
    if (n_img <= 0 || n_shift <= 0) return hipSuccess;
    if (n_shift > 64) return hipErrorInvalidValue;
    const size_t lds = (256 + 4 * 512) * sizeof(float2) + (size_t)(img_dim + 2) * (img_dim + 2) * sizeof(float);
    if (hipError_t e = prepare_launches()) return e;
    hipLaunchKernelGGL(k_polar_spectra<true>, dim3(n_img), dim3(256), lds, st, imgs, img_dim, acc,
                       shift_i, n_shift, tmpl, n_ring, twiddle, spec, norm, smask);
    return hipGetLastError();
}

hipError_t launch_ref_finalize(float2 *rspec, const float2 *rnorm, int n_ref, int n_ring,
                               float *rsum0, hipStream_t st) {
    if (n_ref <= 0) return hipSuccess;
    hipLaunchKernelGGL(k_ref_finalize, dim3(n_ref), dim3(kSlots), 0, st, rspec, rnorm, n_ring, rsum0);
    return hipGetLastError();
}

// Tile of the single-job search: kMrefNW wavefronts per workgroup, each with
// kMrefRW shift rows x kMrefRB references.
constexpr int kMrefNW = 4, kMrefRW = 4, kMrefRB = 4;

template <int NW>
static hipError_t launch_mref_nw(const MrefJob &j, const float2 *tw, hipStream_t st) {
    const long long blocks = (long long)j.n_img * j.GB;
    hipLaunchKernelGGL((k_ccf_mref<NW, kMrefRW, kMrefRB>), dim3((unsigned)blocks), dim3(NW * 64), 0,
                       st, j, tw);
    return hipGetLastError();
}

hipError_t launch_ccf_mref(const float2 *spec, const float2 *norm, const float2 *rspec,
                           const float *rsum0, int n_img, int n_shift, int n_ring, int n_ref,
                           const float2 *twiddle, PeakRec *rec, hipStream_t st) {
    if (n_img <= 0 || n_ref <= 0) return hipSuccess;
    MrefJob j = mref_job(spec, norm, rspec, rsum0, rec, n_img, n_shift, n_ring, n_ref, kMrefNW);
    j.G = (n_ref + kMrefRB - 1) / kMrefRB;   // groups of kMrefRB references, one per wavefront
    j.GB = (j.G + kMrefNW - 1) / kMrefNW;
    return launch_mref_nw<kMrefNW>(j, twiddle, st);
}

MrefJob mref_job(const float2 *spec, const float2 *norm, const float2 *rspec, const float *rsum0,
                 PeakRec *rec, int n_img, int n_shift, int n_ring, int n_ref, int nw) {
    MrefJob j;
    j.spec = spec;
    j.norm = norm;
    j.rspec = rspec;
    j.rsum0 = rsum0;
    j.rec = rec;
    j.n_img = n_img;
    j.S = n_shift;
    j.R = n_ring;
    j.N = n_ref;
    j.G = (n_ref + 3) / 4;
    j.GB = (j.G + nw - 1) / nw;
    return j;
}

int mref_queue_blocks_per_cu(int nw) {
    int n = 0;
    hipError_t e = hipErrorInvalidValue;
    if (nw == 1) e = hipOccupancyMaxActiveBlocksPerMultiprocessor(&n, k_ccf_mref_queue<1>, 64, 0);
    if (nw == 2) e = hipOccupancyMaxActiveBlocksPerMultiprocessor(&n, k_ccf_mref_queue<2>, 128, 0);
    if (nw == 4) e = hipOccupancyMaxActiveBlocksPerMultiprocessor(&n, k_ccf_mref_queue<4>, 256, 0);
    return e == hipSuccess ? n : 0;
}

hipError_t launch_ccf_mref_queue(const MrefJob *jobs, const int4 *items, const int *qbeg,
                                 int *counters, int nw, int blocks, const float2 *twiddle,
                                 hipStream_t st) {
    if (nw != 1 && nw != 2 && nw != 4) return hipErrorInvalidValue;
    hipLaunchKernelGGL(k_reset_counters, dim3(1), dim3(64), 0, st, counters, 8);
    hipError_t e = hipGetLastError();
    if (e != hipSuccess) return e;
    if (nw == 1)
        hipLaunchKernelGGL((k_ccf_mref_queue<1>), dim3(blocks), dim3(64), 0, st, jobs, items, qbeg,
                           counters, twiddle);
    else if (nw == 2)
        hipLaunchKernelGGL((k_ccf_mref_queue<2>), dim3(blocks), dim3(128), 0, st, jobs, items, qbeg,
                           counters, twiddle);
    else
        hipLaunchKernelGGL((k_ccf_mref_queue<4>), dim3(blocks), dim3(256), 0, st, jobs, items, qbeg,
                           counters, twiddle);
    return hipGetLastError();
}

hipError_t launch_ccf_single(const float2 *spec, const float2 *norm, const float2 *rspec,
                             const float *rsum0, const int *ref_of, int ref_stride, int n_img,
                             int n_shift, int n_ring, int n_ref, const float2 *twiddle,
                             PeakRec *rec, hipStream_t st) {
    if (n_img <= 0) return hipSuccess;
    const int blocks = (n_img + 3) / 4;
    hipLaunchKernelGGL(k_ccf_items<true>, dim3(blocks), dim3(256), 0, st, spec, norm, rspec, rsum0,
                       ref_of, ref_stride, 1, n_img, n_shift, n_ring, n_ref, twiddle, rec, nullptr);
    return hipGetLastError();
}

hipError_t launch_ccf_candidates(const float2 *spec, const float2 *norm, const float2 *rspec,
                                 const float *rsum0, const int *cand, int K, int n_img,
                                 int n_shift, int n_ring, int n_ref, const float2 *twiddle,
                                 PeakRec *rec, hipStream_t st, const int *pidx) {
    const int items = n_img * K;
    if (items <= 0) return hipSuccess;
    if (K == 4 || K == 8) {   // one workgroup per list, one wavefront per 4 slots
        MrefJob j = mref_job(spec, norm, rspec, rsum0, rec, n_img, n_shift, n_ring, n_ref, 1);
        j.cand = cand;
        j.pidx = pidx;
        j.K = K;
        j.G = K / 4;
        j.GB = 1;
        const dim3 grid((unsigned)n_img);
        if (K == 4)
            hipLaunchKernelGGL(k_ccf_cand_mref<1>, grid, dim3(64), 0, st, j, twiddle);
        else
            hipLaunchKernelGGL(k_ccf_cand_mref<2>, grid, dim3(128), 0, st, j, twiddle);
        return hipGetLastError();
    }
    hipLaunchKernelGGL(k_ccf_items<false>, dim3((items + 3) / 4), dim3(256), 0, st, spec, norm,
                       rspec, rsum0, cand, 1, K, items, n_shift, n_ring, n_ref, twiddle, rec, pidx);
    return hipGetLastError();
}

hipError_t launch_topk(const PeakRec *rec, int n_img, int n_ref, int K, int *cand, hipStream_t st) {
    if (n_img <= 0) return hipSuccess;
    if (K != 8) return hipErrorInvalidValue;
    hipLaunchKernelGGL(k_topk<8>, dim3((n_img + 127) / 128), dim3(128), 0, st, rec, n_img, n_ref, cand);
    return hipGetLastError();
}

hipError_t launch_best_candidate(const PeakRec *rec, const int *cand, int n_img, int K,
                                 float tie_rel, int *best, int *second, int *tie, hipStream_t st) {
    if (n_img <= 0) return hipSuccess;
    hipLaunchKernelGGL(k_best_candidate, dim3((n_img + 127) / 128), dim3(128), 0, st, rec, cand,
                       n_img, K, tie_rel, best, second, tie);
    return hipGetLastError();
}

hipError_t launch_rescore_f64(const float2 *spec, const float2 *norm, const float2 *rspec,
                              const PeakRec *rec, const int *cand, int *best, int *second,
                              const int *tie, int n_img, int K, int n_shift, int n_ring,
                              int n_ref, double *val64, hipStream_t st) {
    if (n_img <= 0) return hipSuccess;
    const int waves = 2 * n_img;
    hipLaunchKernelGGL(k_rescore_f64, dim3((waves + 3) / 4), dim3(256), 0, st, spec, norm, rspec,
                       rec, cand, best, second, tie, n_img, K, n_shift, n_ring, n_ref, val64);
    hipError_t e = hipGetLastError();
    if (e != hipSuccess) return e;
    hipLaunchKernelGGL(k_resolve_ties, dim3((n_img + 127) / 128), dim3(128), 0, st, cand, val64,
                       tie, n_img, K, best, second);
    return hipGetLastError();
}

hipError_t launch_decode_best(const PeakRec *rec, const int *cand, const int *best,
                              const int *tie, int n_img, int K, ShiftView acc,
                              const float2 *shifts_f, int n_shift, int p_off, int *best_ref,
                              float *best_val, float *best_peak, int *tie_out, hipStream_t st,
                              const int *idx) {
    if (n_img <= 0) return hipSuccess;
    hipLaunchKernelGGL(k_decode_best, dim3((n_img + 127) / 128), dim3(128), 0, st, rec, cand, best,
                       tie, n_img, K, acc, shifts_f, n_shift, p_off, best_ref, best_val, best_peak,
                       tie_out, idx);
    return hipGetLastError();
}

hipError_t launch_decode_mref(const PeakRec *rec, int n_img, int n_ref, ShiftView acc,
                              const float2 *shifts_f, int n_shift, float *peak, int peak_row,
                              int p_off, float *d, int d_pitch, hipStream_t st) {
    const long long n = (long long)n_img * n_ref;
    if (n <= 0) return hipSuccess;
    hipLaunchKernelGGL(k_decode_mref, dim3((unsigned)((n + 255) / 256)), dim3(256), 0, st, rec,
                       n_img, n_ref, acc, shifts_f, n_shift, peak, peak_row, p_off, d, d_pitch);
    return hipGetLastError();
}

hipError_t launch_decode_single(const PeakRec *rec, int cnt, DevParam *prm,
                                const float2 *shifts_f, int n_shift, float shift_limit,
                                hipStream_t st) {
    if (cnt <= 0) return hipSuccess;
    hipLaunchKernelGGL(k_decode_single, dim3((cnt + 255) / 256), dim3(256), 0, st, rec, cnt, prm,
                       shifts_f, n_shift, shift_limit);
    return hipGetLastError();
}

hipError_t launch_transform(const float *imgs, int n, int img_dim, const DevParam *prm,
                            float *out, hipStream_t st) {
    if (n <= 0) return hipSuccess;
    hipLaunchKernelGGL(k_transform, dim3(n), dim3(256), 0, st, imgs, img_dim, prm, out);
    return hipGetLastError();
}

size_t pre_average_scratch_floats(int P, int D) {
    return (size_t)((P + kPreChunk - 1) / kPreChunk) * D * D;
}

hipError_t launch_pre_average(const float *imgs, int P, int D, const DevParam *prm, float4 *tp,
                              int *mir, float *part, float *avg, hipStream_t st) {
    if (P <= 0) return hipSuccess;
    const int nc = (P + kPreChunk - 1) / kPreChunk;
    hipLaunchKernelGGL(k_pre_tp, dim3((P + 255) / 256), dim3(256), 0, st, prm, P, tp, mir);
    hipLaunchKernelGGL(k_pre_partial, dim3((D * D + 255) / 256, nc), dim3(256), 0, st, imgs, D, P,
                       tp, mir, part);
    hipLaunchKernelGGL(k_pre_finish, dim3((D * D + 255) / 256), dim3(256), 0, st, part, D, P, nc,
                       avg);
    return hipGetLastError();
}

hipError_t launch_pre_center(const DevParam *prm, int P, float2 *shift, hipStream_t st) {
    hipLaunchKernelGGL(k_pre_center, dim3(1), dim3(256), 0, st, prm, P, shift);
    return hipGetLastError();
}

hipError_t launch_class_average(const float *tr, int img_dim, const int *cid_idx, int n_cls,
                                float *refs, hipStream_t st) {
    if (n_cls <= 0) return hipSuccess;
    hipLaunchKernelGGL(k_class_average, dim3(n_cls), dim3(256), 0, st, tr, img_dim, cid_idx, refs);
    return hipGetLastError();
}

static hipError_t filter_passes(float *refs, int n, int D, float cutoff, float falloff,
                                const float2 *shift, int abs_freq, float gain, float2 *work,
                                const float2 *tw_d, hipStream_t st) {
    if (n <= 0) return hipSuccess;
    const int Dh = D / 2 + 1;
    hipLaunchKernelGGL(k_row_dft, dim3(D, n), dim3(128), 0, st, refs, D, work, tw_d);
    hipError_t e = hipGetLastError();
    if (e != hipSuccess) return e;
    hipLaunchKernelGGL(k_col_filter, dim3(Dh, n), dim3(128), D * sizeof(float2), st, work, D,
                       tw_d, cutoff, falloff, shift, abs_freq);
    e = hipGetLastError();
    if (e != hipSuccess) return e;
    hipLaunchKernelGGL(k_row_c2r, dim3(D, n), dim3(128), 0, st, refs, D, work, tw_d, gain);
    return hipGetLastError();
}

hipError_t launch_tanl_filter(float *refs, int n, int img_dim, float cutoff, float falloff,
                              float2 *work, const float2 *tw_d, hipStream_t st) {
    return filter_passes(refs, n, img_dim, cutoff, falloff, nullptr, 0, -1.0f, work, tw_d, st);
}

hipError_t launch_filter_shift(float *refs, int n, int img_dim, float cutoff, float falloff,
                               const float2 *shift, float2 *work, const float2 *tw_d,
                               hipStream_t st) {
    const float dd = (float)img_dim * (float)img_dim;
    return filter_passes(refs, n, img_dim, cutoff, falloff, shift, 1, 1.0f / dd, work, tw_d, st);
}

hipError_t launch_update_references(const float *imgs, int P, int D, const float *peak,
                                   const int *cls, int N, float cutoff, float falloff,
                                   UpdateScratch w, float *refs, float4 *params, hipStream_t st) {
    if (P <= 0 || N <= 0) return hipSuccess;
    const int gp = (P + 255) / 256;
    hipLaunchKernelGGL(k_upd_params, dim3(gp), dim3(256), 0, st, peak, P, N, cls, w.tp, w.mir,
                       w.post, w.raw, w.bad);
    hipLaunchKernelGGL(k_class_members<0>, dim3(N), dim3(256), 0, st, cls, P, w.cnt, w.off,
                       w.members);
    hipLaunchKernelGGL(k_scan_counts, dim3(1), dim3(64), 0, st, w.cnt, N, w.off);
    hipLaunchKernelGGL(k_class_members<1>, dim3(N), dim3(256), 0, st, cls, P, w.cnt, w.off,
                       w.members);
    hipLaunchKernelGGL(k_class_shift, dim3((N + 63) / 64), dim3(64), 0, st, w.off, w.members, N,
                       w.post, w.mir, w.avg, w.neg, w.avg_out, w.members_out);
    hipLaunchKernelGGL(k_upd_params_out, dim3(gp), dim3(256), 0, st, cls, P, N, w.raw, w.post,
                       w.mir, w.avg, params);
    hipLaunchKernelGGL(k_class_transform_average, dim3((D * D + 255) / 256, N), dim3(256), 0, st,
                       imgs, D, w.off, w.members, w.tp, w.mir, refs);
    hipError_t e = hipGetLastError();
    if (e != hipSuccess) return e;
    return launch_filter_shift(refs, N, D, cutoff, falloff, w.neg, w.work, w.tw, st);
}

hipError_t launch_prop_items(const int *prop, int n_img, int K, int N, int nx, int ny, int hw,
                             int n_bins, float psi_half, PropItem *items,
                             unsigned long long *smask, hipStream_t st) {
    if (n_img <= 0) return hipSuccess;
    if (K < 1 || K > kMaxProposals || nx * ny > 64 || hw < 0 || hw > 2) return hipErrorInvalidValue;
    hipLaunchKernelGGL(k_prop_items, dim3((n_img + 127) / 128), dim3(128), 0, st,
                       reinterpret_cast<const int4 *>(prop), n_img, K, N, nx, ny, hw, n_bins, psi_half,
                       items, smask);
    return hipGetLastError();
}

hipError_t launch_ccf_prop(const float2 *spec, const float2 *norm, const float2 *rspec,
                           const float *rsum0, const PropItem *items, int K, int n_img, int n_shift,
                           int n_ring, int n_ref, int nx, int ny, int hw, const float2 *twiddle,
                           PeakRec *rec, hipStream_t st) {
    const int n_items = n_img * K;
    if (n_items <= 0) return hipSuccess;
    if (K < 1 || K > kMaxProposals) return hipErrorInvalidValue;
    hipLaunchKernelGGL(k_ccf_prop, dim3(n_img), dim3(64 * K), 0, st, spec, norm, rspec, rsum0,
                       items, K, n_items, n_shift, n_ring, n_ref, nx, ny, hw, twiddle, rec);
    return hipGetLastError();
}

hipError_t launch_best_prop(const PeakRec *rec, const PropItem *items, int n_img, int K, int nx,
                            int ny, int hw, int edge_test, ShiftView acc, const float2 *shifts_f,
                            int n_shift, int p_off, int *best_ref, float *best_val,
                            float *best_peak, int *edge, hipStream_t st) {
    if (n_img <= 0) return hipSuccess;
    if (nx * ny != n_shift || ny < 1) return hipErrorInvalidValue;
    hipLaunchKernelGGL(k_best_prop, dim3((n_img + 127) / 128), dim3(128), 0, st, rec, items, n_img, K,
                       nx, ny, hw, edge_test, acc, shifts_f, n_shift, p_off, best_ref, best_val,
                       best_peak, edge);
    return hipGetLastError();
}

hipError_t launch_prop_refs(const int *prop, const int *idx, int n, int K, int N, int *cand,
                            hipStream_t st) {
    if (n <= 0) return hipSuccess;
    if (K < 1 || K > kMaxProposals) return hipErrorInvalidValue;
    hipLaunchKernelGGL(k_prop_refs, dim3((n + 127) / 128), dim3(128), 0, st,
                       reinterpret_cast<const int4 *>(prop), idx, n, K, N, cand);
    return hipGetLastError();
}

hipError_t launch_level2_mask(const int *flag, int n, int S, const unsigned long long *done,
                              unsigned long long *smask, hipStream_t st) {
    if (n <= 0) return hipSuccess;
    if (S > 64) return hipErrorInvalidValue;
    hipLaunchKernelGGL(k_level2_mask, dim3((n + 255) / 256), dim3(256), 0, st, flag, n, S, done, smask);
    return hipGetLastError();
}

hipError_t launch_best_refs_level2(const PeakRec *rec, const int *cand, int n, int K, ShiftView acc,
                                   const float2 *shifts_f, int n_shift, const int *idx, int p_off,
                                   float threshold, int *best_ref, float *best_val,
                                   float *best_peak, int *fallback, hipStream_t st) {
    if (n <= 0) return hipSuccess;
    hipLaunchKernelGGL(k_best_refs_level2, dim3((n + 127) / 128), dim3(128), 0, st, rec, cand, n, K,
                       acc, shifts_f, n_shift, idx, p_off, threshold, best_ref, best_val, best_peak,
                       fallback);
    return hipGetLastError();
}

hipError_t launch_fallback_flags(const float *best_val, const int *edge, const float *conf,
                                 float conf_min, int n_img, float threshold, int p_off,
                                 int *fallback, hipStream_t st) {
    if (n_img <= 0) return hipSuccess;
    hipLaunchKernelGGL(k_fallback_flags, dim3((n_img + 255) / 256), dim3(256), 0, st, best_val, edge,
                       conf, conf_min, n_img, threshold, p_off, fallback);
    return hipGetLastError();
}

hipError_t launch_compact(const int *flag, int n, int lo, int hi, int *idx, int *count,
                          hipStream_t st) {
    if (n <= 0) return hipSuccess;
    hipLaunchKernelGGL(k_reset_counters, dim3(1), dim3(64), 0, st, count, 1);
    hipLaunchKernelGGL(k_compact, dim3((n + 255) / 256), dim3(256), 0, st, flag, n, lo, hi, idx, count);
    return hipGetLastError();
}

hipError_t launch_gather(const float *imgs, int D, ShiftView acc, const int *idx, int n, float *out,
                         float2 *out_acc, hipStream_t st) {
    if (n <= 0) return hipSuccess;
    hipLaunchKernelGGL(k_gather, dim3(n), dim3(256), 0, st, imgs, D, acc, idx, n, out, out_acc);
    return hipGetLastError();
}

hipError_t launch_best_refs_scatter(const PeakRec *rec, int n, int N, const float2 *acc,
                                    const float2 *shifts_f, int n_shift, const int *idx,
                                    int *best_ref, float *best_val, float *best_peak,
                                    hipStream_t st) {
    if (n <= 0) return hipSuccess;
    hipLaunchKernelGGL(k_best_refs_scatter, dim3((n + 127) / 128), dim3(128), 0, st, rec, n, N, acc,
                       shifts_f, n_shift, idx, best_ref, best_val, best_peak);
    return hipGetLastError();
}

}  // namespace isac
