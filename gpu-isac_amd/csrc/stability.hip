// Multi-alignment stability test of ISAC classes, one workgroup per class.
//
// The driver (bin/sp_isac2_gpu.py:1481-1491) calls, per class,
//     stable_set, mirror_consistent_rate, err = multi_align_stability(
//         aln_param[k], 0.0, 10000.0, thld_err, False, last_ring * 2)
// on the L alignments {angle, sx, sy, mirror} of each member from L repeated
// reference-free runs (:1381-1393).  This kernel computes the same quantities:
//  1. mirror consistency: each member is consistent with exactly one pattern of
//     per-run mirror flips (f_j = m_j xor m_{L-1}); the pattern held by most
//     members wins (the smallest pattern index among equal counts, the order of
//     a brute-force scan over patterns), its members are the consistent set;
//  2. starting global transforms G_j = (alpha_j, t_j, f_j) from the mean angle
//     and shift differences to the last run; members whose pixel error is above
//     3 err_thld are left out of the fit;
//  3. the fit: the mean squared pixel error
//         e_i = r^2 (1 - |mean_j e^{i theta'_ij}|) + var_j(s'x_ij) + var_j(s'y_ij)
//     is minimised over the global transforms.  For given rotations the
//     optimal translations equalise the runs' mean shifts (closed form), so
//     only the L - 1 rotations z_j = e^{i alpha_j} are searched.  With
//     a_ij = e^{i phi_ij} (flipped angles) and b_ij the flipped, centred shifts
//     as complex numbers, minimising the error is maximising
//         F(z) = mean_i [ r^2 |sum_j z_j a_ij| / L + |sum_j z_j b_ij|^2 / L^2 ],
//     a sum of convex functions of each z_j; replacing each by its tangent at
//     the current point gives a lower bound whose maximum over |z_j| = 1 is
//     z_j = conj(c_j) / |c_j|,  c_j = sum_i [ r^2/L conj(u_i) a_ij + 2/L^2 conj(Q_i) b_ij ]
//     (u_i = S_i / |S_i|, S_i and Q_i the current sums): a monotone
//     minorise-maximise update, swept over j until the rotations stop moving;
//  4. pixel errors of every consistent member under the fitted transforms, the
//     class error sqrt(mean e_i), the stable members (pixel error <= err_thld,
//     class error <= grp_err_thld) and each member's average alignment in the
//     last run's frame.
// All arithmetic in FP64.
#include "../../include/gpu_aln_pack.h"
#include "isac_internal.h"

#include <cmath>

namespace isac {
namespace {

constexpr int kStabThreads = 256;
constexpr int kStabWaves = kStabThreads / 64;
constexpr int kStabMaxSweeps = 2000;

__device__ __forceinline__ double2 cmul(double2 a, double2 b) {
    return make_double2(a.x * b.x - a.y * b.y, a.x * b.y + a.y * b.x);
}

__device__ __forceinline__ double2 cexpi_deg(double deg) {
    double s, c;
    sincospi(deg / 180.0, &s, &c);
    return make_double2(c, s);
}

// sum over the workgroup of one double2 per thread; every thread gets the sum
__device__ double2 block_sum2(double2 v, double2 *red) {
#pragma unroll
    for (int off = 32; off >= 1; off >>= 1) {
        v.x += __shfl_xor(v.x, off, 64);
        v.y += __shfl_xor(v.y, off, 64);
    }
    const int w = threadIdx.x >> 6;
    __syncthreads();
    if ((threadIdx.x & 63) == 0) red[w] = v;
    __syncthreads();
    double2 t = make_double2(0.0, 0.0);
#pragma unroll
    for (int k = 0; k < kStabWaves; ++k) {
        t.x += red[k].x;
        t.y += red[k].y;
    }
    return t;
}

// per-member quantities of one class: param (run j, member i) at
// prm[((size_t)j * n_total + i0 + i) * 4]
struct ClassView {
    const double *prm;
    int n_total, i0, n, L;
    __device__ const double *at(int j, int i) const {
        return prm + ((size_t)j * n_total + i0 + i) * 4;
    }
};

// flipped angle (degrees) and flipped shift of run j, member i
__device__ __forceinline__ void member(const ClassView &cv, int j, int i, int f, double &phi,
                                       double2 &s) {
    const double *p = cv.at(j, i);
    const double th = f ? -p[0] : p[0];
    phi = th;
    s = make_double2(f ? -p[1] : p[1], p[2]);
}

// squared pixel error of member i under transforms (z_j, t_j), flips f; also its
// average alignment (angle degrees, shift) in the last run's frame
__device__ double pixel_error(const ClassView &cv, int i, const double2 *z, const double2 *t,
                              unsigned flips, double r, double *avg) {
    const int L = cv.L;
    double2 ua = make_double2(0.0, 0.0), ms = make_double2(0.0, 0.0);
    double ss = 0.0;
    for (int j = 0; j < L; ++j) {
        double phi;
        double2 s;
        member(cv, j, i, (flips >> j) & 1, phi, s);
        const double2 e = cmul(z[j], cexpi_deg(phi));
        ua.x += e.x;
        ua.y += e.y;
        const double2 sp = make_double2(cmul(z[j], s).x + t[j].x, cmul(z[j], s).y + t[j].y);
        ms.x += sp.x;
        ms.y += sp.y;
        ss += sp.x * sp.x + sp.y * sp.y;
    }
    const double invL = 1.0 / L;
    const double rho = sqrt(ua.x * ua.x + ua.y * ua.y) * invL;
    const double mx = ms.x * invL, my = ms.y * invL;
    const double var = ss * invL - (mx * mx + my * my);
    if (avg) {
        double ang = atan2(ua.y, ua.x) * (180.0 / M_PI);
        ang = fmod(ang, 360.0);
        if (ang < 0.0) ang += 360.0;
        avg[0] = ang;
        avg[1] = mx;
        avg[2] = my;
    }
    return r * r * (1.0 - rho) + var;
}

// translations that equalise the runs' mean shifts under rotations z (means over
// the members with sel(i)): t_j = mean s'_{L-1} - z_j mean s_j, t_{L-1} = 0
template <class Sel>
__device__ void fit_translations(const ClassView &cv, unsigned flips, const double2 *z, double2 *t,
                                 double2 *red, double cnt, Sel sel) {
    __shared__ double2 smean[16];
    for (int j = 0; j < cv.L; ++j) {
        double2 acc = make_double2(0.0, 0.0);
        for (int i = threadIdx.x; i < cv.n; i += kStabThreads)
            if (sel(i)) {
                double phi;
                double2 s;
                member(cv, j, i, (flips >> j) & 1, phi, s);
                acc.x += s.x;
                acc.y += s.y;
            }
        acc = block_sum2(acc, red);
        if (threadIdx.x == 0) smean[j] = make_double2(acc.x / cnt, acc.y / cnt);
    }
    __syncthreads();
    if (threadIdx.x == 0) {
        const double2 last = smean[cv.L - 1];
        for (int j = 0; j < cv.L; ++j) {
            const double2 u = cmul(z[j], smean[j]);
            t[j] = (j == cv.L - 1) ? make_double2(0.0, 0.0) : make_double2(last.x - u.x, last.y - u.y);
        }
    }
    __syncthreads();
}

__global__ __launch_bounds__(kStabThreads) void k_stability(
    const double *__restrict__ prm, int n_total, const int *__restrict__ off, int L,
    float mir_thld, float grp_thld, float err_thld, double r, float *__restrict__ pix,
    float *__restrict__ avg, int *__restrict__ stable, float *__restrict__ cls,
    float *__restrict__ gout) {
    const int c = blockIdx.x;
    int i0 = off[c], i1 = off[c + 1];
    if (i0 < 0 || i1 < i0 || i1 > n_total) i0 = i1 = 0;   // a malformed range: an empty class
    ClassView cv{prm, n_total, i0, i1 - i0, L};
    const int n = cv.n;
    const int npat = 1 << (L - 1);
    __shared__ int hist[1 << (kStabMaxRuns - 1)];
    __shared__ double2 red[kStabWaves];
    __shared__ double2 z[kStabMaxRuns], t[kStabMaxRuns];
    __shared__ int s_best, s_count, s_kept;
    __shared__ double s_err;

    // default outputs: not consistent
    for (int i = threadIdx.x; i < n; i += kStabThreads) {
        pix[cv.i0 + i] = -1.0f;
        stable[cv.i0 + i] = 0;
        for (int e = 0; e < 4; ++e) avg[(size_t)(cv.i0 + i) * 4 + e] = 0.0f;
    }
    for (int p = threadIdx.x; p < npat; p += kStabThreads) hist[p] = 0;
    __syncthreads();

    // ---- 1. mirror consistency: the flip pattern of each member
    auto pattern = [&](int i) {
        const unsigned ml = cv.at(L - 1, i)[3] != 0.0 ? 1u : 0u;
        unsigned idx = 0;
        for (int j = 0; j < L - 1; ++j) {
            const unsigned f = (cv.at(j, i)[3] != 0.0 ? 1u : 0u) ^ ml;
            idx |= f << (L - 2 - j);
        }
        return (int)idx;
    };
    for (int i = threadIdx.x; i < n; i += kStabThreads) atomicAdd(&hist[pattern(i)], 1);
    __syncthreads();
    if (threadIdx.x == 0) {
        int best = 0, cnt = -1;
        for (int p = 0; p < npat; ++p)
            if (hist[p] > cnt) {
                cnt = hist[p];
                best = p;
            }
        s_best = best;
        s_count = cnt;
    }
    __syncthreads();
    const int best = s_best, count = s_count;
    unsigned flips = 0;   // bit j = flip of run j (pattern bit L - 2 - j)
    for (int j = 0; j < L - 1; ++j) flips |= (((unsigned)best >> (L - 2 - j)) & 1u) << j;
    const double rate = n > 0 ? (double)count / n : 0.0;
    auto write_class = [&](double err) {
        if (threadIdx.x == 0) {
            cls[c * 2 + 0] = (float)rate;
            cls[c * 2 + 1] = (float)err;
            for (int j = 0; j < L; ++j) {
                float *g = gout + ((size_t)c * L + j) * 4;
                const double2 zz = z[j];
                g[0] = (float)(atan2(zz.y, zz.x) * (180.0 / M_PI));
                g[1] = (float)t[j].x;
                g[2] = (float)t[j].y;
                g[3] = (float)((flips >> j) & 1u);
            }
        }
    };
    if (threadIdx.x < kStabMaxRuns) {
        z[threadIdx.x] = make_double2(1.0, 0.0);
        t[threadIdx.x] = make_double2(0.0, 0.0);
    }
    __syncthreads();
    if (!(rate > (double)mir_thld) || count < 2) {
        write_class(-1.0);
        return;
    }
    auto consistent = [&](int i) { return pattern(i) == best; };

    // ---- 2. starting transforms from the consistent members
    for (int j = 0; j < L - 1; ++j) {
        double2 acc = make_double2(0.0, 0.0);
        for (int i = threadIdx.x; i < n; i += kStabThreads)
            if (consistent(i)) {
                double phi, phl;
                double2 s, sl;
                member(cv, j, i, (flips >> j) & 1, phi, s);
                member(cv, L - 1, i, 0, phl, sl);
                const double2 e = cexpi_deg(phl - phi);
                acc.x += e.x;
                acc.y += e.y;
            }
        acc = block_sum2(acc, red);
        if (threadIdx.x == 0) {
            const double h = sqrt(acc.x * acc.x + acc.y * acc.y);
            z[j] = h > 0.0 ? make_double2(acc.x / h, acc.y / h) : make_double2(1.0, 0.0);
        }
    }
    __syncthreads();
    fit_translations(cv, flips, z, t, red, (double)count, consistent);
    // members left out of the fit
    int kept_local = 0;
    for (int i = threadIdx.x; i < n; i += kStabThreads)
        if (consistent(i)) {
            const double e0 = pixel_error(cv, i, z, t, flips, r, nullptr);
            const bool keep = sqrt(fmax(e0, 0.0)) <= 3.0 * (double)err_thld;
            stable[cv.i0 + i] = keep ? 2 : 0;   // 2: in the fit (reset below)
            kept_local += keep;
        }
    {
        const double2 k2 = block_sum2(make_double2((double)kept_local, 0.0), red);
        if (threadIdx.x == 0) s_kept = (int)(k2.x + 0.5);
    }
    __syncthreads();
    const int nkeep = s_kept;
    auto kept = [&](int i) { return stable[cv.i0 + i] == 2; };
    if (nkeep < 2) {
        for (int i = threadIdx.x; i < n; i += kStabThreads) stable[cv.i0 + i] = 0;
        write_class(-1.0);
        return;
    }

    // ---- 3. rotations by minorise-maximise sweeps (translations in closed form)
    __shared__ double2 sbar[kStabMaxRuns];
    for (int j = 0; j < L; ++j) {
        double2 acc = make_double2(0.0, 0.0);
        for (int i = threadIdx.x; i < n; i += kStabThreads)
            if (kept(i)) {
                double phi;
                double2 s;
                member(cv, j, i, (flips >> j) & 1, phi, s);
                acc.x += s.x;
                acc.y += s.y;
            }
        acc = block_sum2(acc, red);
        if (threadIdx.x == 0) sbar[j] = make_double2(acc.x / nkeep, acc.y / nkeep);
    }
    __syncthreads();
    const double invL = 1.0 / L, wa = r * r * invL, wb = 2.0 * invL * invL;
    __shared__ double s_move;
    for (int sweep = 0; sweep < kStabMaxSweeps; ++sweep) {
        double move = 0.0;
        for (int j = 0; j < L - 1; ++j) {
            double2 acc = make_double2(0.0, 0.0);
            for (int i = threadIdx.x; i < n; i += kStabThreads)
                if (kept(i)) {
                    double2 S = make_double2(0.0, 0.0), Q = make_double2(0.0, 0.0);
                    double2 aj = make_double2(0.0, 0.0), bj = make_double2(0.0, 0.0);
                    for (int k = 0; k < L; ++k) {
                        double phi;
                        double2 s;
                        member(cv, k, i, (flips >> k) & 1, phi, s);
                        const double2 a = cexpi_deg(phi);
                        const double2 b = make_double2(s.x - sbar[k].x, s.y - sbar[k].y);
                        const double2 za = cmul(z[k], a), zb = cmul(z[k], b);
                        S.x += za.x;
                        S.y += za.y;
                        Q.x += zb.x;
                        Q.y += zb.y;
                        if (k == j) {
                            aj = a;
                            bj = b;
                        }
                    }
                    const double h = sqrt(S.x * S.x + S.y * S.y);
                    const double2 u = h > 0.0 ? make_double2(S.x / h, S.y / h) : make_double2(0.0, 0.0);
                    // conj(u) a_j and conj(Q) b_j
                    acc.x += wa * (u.x * aj.x + u.y * aj.y) + wb * (Q.x * bj.x + Q.y * bj.y);
                    acc.y += wa * (u.x * aj.y - u.y * aj.x) + wb * (Q.x * bj.y - Q.y * bj.x);
                }
            acc = block_sum2(acc, red);
            if (threadIdx.x == 0) {
                const double h = sqrt(acc.x * acc.x + acc.y * acc.y);
                if (h > 0.0) {
                    const double2 zn = make_double2(acc.x / h, -acc.y / h);   // conj(c) / |c|
                    const double2 d = cmul(zn, make_double2(z[j].x, -z[j].y));
                    move = fmax(move, fabs(atan2(d.y, d.x)));
                    z[j] = zn;
                }
            }
            __syncthreads();
        }
        if (threadIdx.x == 0) s_move = move;
        __syncthreads();
        if (s_move < 1e-14) break;
    }
    fit_translations(cv, flips, z, t, red, (double)nkeep, kept);

    // ---- 4. pixel errors, class error, stable members, average alignments
    double esum = 0.0;
    for (int i = threadIdx.x; i < n; i += kStabThreads)
        if (consistent(i)) esum += fmax(pixel_error(cv, i, z, t, flips, r, nullptr), 0.0);
    {
        const double2 es = block_sum2(make_double2(esum, 0.0), red);
        if (threadIdx.x == 0) s_err = sqrt(es.x / count);
    }
    __syncthreads();
    const double err = s_err;
    for (int i = threadIdx.x; i < n; i += kStabThreads) {
        int ok = 0;
        if (consistent(i)) {
            double a3[3];
            const double pe = sqrt(fmax(pixel_error(cv, i, z, t, flips, r, a3), 0.0));
            pix[cv.i0 + i] = (float)pe;
            float *o = avg + (size_t)(cv.i0 + i) * 4;
            o[0] = (float)a3[0];
            o[1] = (float)a3[1];
            o[2] = (float)a3[2];
            o[3] = cv.at(L - 1, i)[3] != 0.0 ? 1.0f : 0.0f;
            ok = (err <= (double)grp_thld && pe <= (double)err_thld) ? 1 : 0;
        }
        stable[cv.i0 + i] = ok;
    }
    write_class(err);
}

}  // namespace

hipError_t launch_stability(const double *prm, int n_total, const int *class_off, int n_class,
                            int L, float mir_thld, float grp_thld, float err_thld, double d,
                            float *pix, float *avg, int *stable, float *cls, float *gout,
                            hipStream_t st) {
    if (n_class <= 0) return hipSuccess;
    if (L < 2 || L > kStabMaxRuns) return hipErrorInvalidValue;
    hipLaunchKernelGGL(k_stability, dim3(n_class), dim3(kStabThreads), 0, st, prm, n_total,
                       class_off, L, mir_thld, grp_thld, err_thld, d / 2.0, pix, avg, stable, cls,
                       gout);
    return hipGetLastError();
}

}  // namespace isac
