// Equal-size group assignment of particles to classes from the alignment
// scores d[n][p] (the consumer of multi_ref_align's d, bin/sp_isac2_gpu.py:1119-1144,
// where EMAN2's Util.assign_groups does it on the host).
//
// Contract (restated from the reference's call site and its comment there):
// pairs (class n, particle p) are taken in descending score order (ties: lower
// class, then lower particle) and a particle joins its pair's class when it has
// no class yet and the class holds fewer than cap = n_img / n_ref particles.
// With one strict order on the pairs, that greedy result is the unique stable
// assignment, and particle-proposing deferred acceptance finds it in parallel:
// every free particle proposes to its best class not yet tried; every class
// keeps its cap best proposers (score descending, particle ascending) and
// releases the rest; repeat until no free particle has a class left.
//
// Output (n_img + 1 ints): class k's particles (ascending) at [k*cap, (k+1)*cap),
// then the r = n_img - n_ref*cap particles no class took (ascending), then one
// class id for those r particles, the class of the highest remaining pair
// (0 when r = 0).  NaN scores count as -infinity.
#include "isac_internal.h"
#include "../../include/gpu_aln_pack.h"

#include <cfloat>
#include <climits>
#include <cmath>
#include <cstdio>
#include <vector>

namespace isac {
namespace {

constexpr int kWave = 64;
constexpr int kSortN = 4096;          // LDS sort capacity of one class step
constexpr int kMaxCap = 2048;         // largest class size supported

__device__ __forceinline__ float score(const float *d, int pitch, int n, int p) {
    const float v = d[(size_t)n * pitch + p];
    return v == v ? v : -INFINITY;
}

// (v1, a1) before (v2, a2): larger value first, then smaller index.
__device__ __forceinline__ bool before(float v1, int a1, float v2, int a2) {
    return v1 > v2 || (v1 == v2 && a1 < a2);
}

// One wavefront per particle: a free particle proposes to its best class after
// the last one it tried (in the order value desc, class asc).
__global__ __launch_bounds__(256) void k_propose(const float *__restrict__ d, int pitch, int N,
                                                 int P, int *__restrict__ status,
                                                 int *__restrict__ last_n, float *__restrict__ last_v,
                                                 int *__restrict__ cnt, int *__restrict__ props,
                                                 int *__restrict__ n_props) {
    const int lane = threadIdx.x & 63;
    const int p = blockIdx.x * 4 + (threadIdx.x >> 6);
    if (p >= P) return;
    if (status[p] != 0) return;   // held or exhausted (uniform over the wavefront)
    const int ln = last_n[p];
    const float lv = last_v[p];
    float bv = -INFINITY;
    int bn = INT_MAX;
    for (int n = lane; n < N; n += kWave) {
        const float v = score(d, pitch, n, p);
        const bool after = ln < 0 || before(lv, ln, v, n);   // strictly after the last try
        if (after && (bn == INT_MAX || before(v, n, bv, bn))) {
            bv = v;
            bn = n;
        }
    }
#pragma unroll
    for (int off = 32; off >= 1; off >>= 1) {
        const float ov = __shfl_xor(bv, off, kWave);
        const int on = __shfl_xor(bn, off, kWave);
        if (on != INT_MAX && (bn == INT_MAX || before(ov, on, bv, bn))) {
            bv = ov;
            bn = on;
        }
    }
    if (lane != 0) return;
    if (bn == INT_MAX) {
        status[p] = 2;   // every class has been tried
        return;
    }
    last_n[p] = bn;
    last_v[p] = bv;
    status[p] = 1;
    const int slot = atomicAdd(&cnt[bn], 1);
    props[(size_t)bn * P + slot] = p;
    atomicAdd(n_props, 1);
}

// Smallest power of two >= x (x >= 1).
__device__ __forceinline__ int pow2_at_least(int x) {
    int n = 1;
    while (n < x) n <<= 1;
    return n;
}

// Bitonic sort of the first n (a power of two <= kSortN) (value, index) pairs in
// LDS, order `before`.
__device__ void lds_sort(float *v, int *a, int n, int tid, int nthreads) {
    for (int k = 2; k <= n; k <<= 1) {
        for (int j = k >> 1; j > 0; j >>= 1) {
            for (int i = tid; i < n; i += nthreads) {
                const int l = i ^ j;
                if (l > i) {
                    const bool up = (i & k) == 0;
                    const bool swap = up ? before(v[l], a[l], v[i], a[i]) : before(v[i], a[i], v[l], a[l]);
                    if (swap) {
                        const float tv = v[i];
                        v[i] = v[l];
                        v[l] = tv;
                        const int ta = a[i];
                        a[i] = a[l];
                        a[l] = ta;
                    }
                }
            }
            __syncthreads();
        }
    }
}

// One workgroup per class: merge the holders (LDS [0, h)) with this round's
// proposers in chunks (LDS [cap, kSortN)), keep the cap best, release the rest.
__global__ __launch_bounds__(256) void k_accept(const float *__restrict__ d, int pitch, int P,
                                                int cap, int *__restrict__ cnt,
                                                const int *__restrict__ props,
                                                int *__restrict__ holders, int *__restrict__ hold_cnt,
                                                int *__restrict__ status) {
    __shared__ float sv[kSortN];
    __shared__ int sa[kSortN];
    const int c = blockIdx.x, tid = threadIdx.x;
    const int m = cnt[c];
    if (m == 0) return;   // uniform
    int h = hold_cnt[c];
    for (int i = tid; i < cap; i += 256) {
        const int p = i < h ? holders[(size_t)c * cap + i] : -1;
        sv[i] = p >= 0 ? score(d, pitch, c, p) : -INFINITY;
        sa[i] = p >= 0 ? p : INT_MAX;
    }
    const int chunk = kSortN - cap;
    for (int q0 = 0; q0 < m; q0 += chunk) {
        const int q1 = min(m, q0 + chunk);
        const int n = pow2_at_least(cap + (q1 - q0));   // <= kSortN
        for (int i = cap + tid; i < n; i += 256) {
            const int q = q0 + (i - cap);
            const int p = q < q1 ? props[(size_t)c * P + q] : -1;
            sv[i] = p >= 0 ? score(d, pitch, c, p) : -INFINITY;
            sa[i] = p >= 0 ? p : INT_MAX;
        }
        __syncthreads();
        lds_sort(sv, sa, n, tid, 256);
        // valid entries now lead; those past cap are released
        for (int i = cap + tid; i < n; i += 256)
            if (sa[i] != INT_MAX) status[sa[i]] = 0;
        h = min(cap, h + (q1 - q0));
        __syncthreads();
    }
    for (int i = tid; i < h; i += 256) holders[(size_t)c * cap + i] = sa[i];
    if (tid == 0) {
        hold_cnt[c] = h;
        cnt[c] = 0;
    }
}

// Output, part 1: block c writes class c's holders ascending at out[c*cap ...].
__global__ __launch_bounds__(256) void k_class_out(int cap, const int *__restrict__ holders,
                                                   const int *__restrict__ hold_cnt,
                                                   int *__restrict__ out) {
    __shared__ float sv[kSortN];
    __shared__ int sa[kSortN];
    const int c = blockIdx.x, tid = threadIdx.x;
    const int h = hold_cnt[c];
    const int n = pow2_at_least(cap);   // cap <= kMaxCap <= kSortN
    for (int i = tid; i < n; i += 256) {
        // sort by particle ascending: value = -particle, index = particle
        const int p = i < h ? holders[(size_t)c * cap + i] : INT_MAX;
        sv[i] = p == INT_MAX ? -INFINITY : -(float)p;
        sa[i] = p;
    }
    __syncthreads();
    lds_sort(sv, sa, n, tid, 256);
    for (int i = tid; i < cap; i += 256) out[(size_t)c * cap + i] = i < h ? sa[i] : -1;
}

// Output, part 2 (one workgroup): the unplaced particles ascending after the
// classes, then the class of the highest remaining pair.
__global__ __launch_bounds__(256) void k_assign_output(const float *__restrict__ d, int pitch, int N,
                                                       int P, int cap, const int *__restrict__ status,
                                                       int *__restrict__ out) {
    __shared__ int sa[256];
    __shared__ int base;
    __shared__ float rv[256];
    __shared__ int rn[256], rp[256];
    const int tid = threadIdx.x;
    // unplaced particles, ascending (ordered block scan over chunks of 256)
    if (tid == 0) base = N * cap;
    __syncthreads();
    float bestv = -INFINITY;
    int bestn = INT_MAX, bestp = INT_MAX;
    for (int p0 = 0; p0 < P; p0 += 256) {
        const int p = p0 + tid;
        const int f = (p < P && status[p] != 1) ? 1 : 0;
        sa[tid] = f;
        __syncthreads();
        for (int off = 1; off < 256; off <<= 1) {   // inclusive scan
            const int x = tid >= off ? sa[tid - off] : 0;
            __syncthreads();
            sa[tid] += x;
            __syncthreads();
        }
        if (f) {
            out[base + sa[tid] - 1] = p;
            for (int n = 0; n < N; ++n) {
                const float v = score(d, pitch, n, p);
                if (bestn == INT_MAX || before(v, n, bestv, bestn) ||
                    (v == bestv && n == bestn && p < bestp)) {
                    bestv = v;
                    bestn = n;
                    bestp = p;
                }
            }
        }
        __syncthreads();
        if (tid == 255) base += sa[255];
        __syncthreads();
    }
    // highest remaining pair (value desc, class asc, particle asc)
    rv[tid] = bestv;
    rn[tid] = bestn;
    rp[tid] = bestp;
    __syncthreads();
    for (int off = 128; off >= 1; off >>= 1) {
        if (tid < off) {
            const float ov = rv[tid + off];
            const int on = rn[tid + off], op = rp[tid + off];
            const bool better = on != INT_MAX &&
                                (rn[tid] == INT_MAX || ov > rv[tid] ||
                                 (ov == rv[tid] && (on < rn[tid] || (on == rn[tid] && op < rp[tid]))));
            if (better) {
                rv[tid] = ov;
                rn[tid] = on;
                rp[tid] = op;
            }
        }
        __syncthreads();
    }
    if (tid == 0) out[P] = rn[0] == INT_MAX ? 0 : rn[0];
}

__global__ void k_assign_init(int *status, int *last_n, float *last_v, int P, int *hold_cnt,
                              int *cnt, int N) {
    const int i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i < P) {
        status[i] = 0;
        last_n[i] = -1;
        last_v[i] = INFINITY;
    }
    if (i < N) {
        hold_cnt[i] = 0;
        cnt[i] = 0;
    }
}

__global__ void k_zero1(int *x) {
    if (threadIdx.x == 0) *x = 0;
}

struct Buf {
    void *p = nullptr;
    ~Buf() {
        if (p) (void)hipFree(p);
    }
    template <class T>
    T *as() { return static_cast<T *>(p); }
};

int err(int code, const char *msg) {
    report_error(code, msg);
    return code;
}

}  // namespace

int assign_groups(const float *d_d, int N, int P, int pitch, int *d_out, hipStream_t st) {
    if (N < 1 || P < N || pitch < P || !d_d || !d_out)
        return err(2, "assign_groups: need 1 <= n_ref <= n_img <= d_pitch and device buffers");
    const int cap = P / N;
    if (cap > kMaxCap)
        return err(4, "assign_groups: classes of more than 2048 particles are not supported");
    Buf status, last_n, last_v, cnt, props, holders, hold_cnt, n_props;
    hipError_t e = hipSuccess;
    auto alloc = [&](Buf &b, size_t bytes) {
        if (e == hipSuccess) e = hipMalloc(&b.p, bytes > 0 ? bytes : 4);
    };
    alloc(status, (size_t)P * 4);
    alloc(last_n, (size_t)P * 4);
    alloc(last_v, (size_t)P * 4);
    alloc(cnt, (size_t)N * 4);
    alloc(props, (size_t)N * P * 4);
    alloc(holders, (size_t)N * cap * 4);
    alloc(hold_cnt, (size_t)N * 4);
    alloc(n_props, 4);
    if (e != hipSuccess) return err(1, "assign_groups: device allocation failed");
    const int mx = P > N ? P : N;
    hipLaunchKernelGGL(k_assign_init, dim3((mx + 255) / 256), dim3(256), 0, st, status.as<int>(),
                       last_n.as<int>(), last_v.as<float>(), P, hold_cnt.as<int>(), cnt.as<int>(), N);
    // every round makes at least one proposal until none is left; a particle
    // proposes to each class at most once, so N * P + 1 rounds bound the loop.
    // A round without proposals changes nothing (no particle is free), so rounds
    // are queued kRoundsPerCheck at a time and the last one's count is read back.
    constexpr int kRoundsPerCheck = 8;
    const long long max_rounds = (long long)N * P + 1;
    for (long long round = 0; round < max_rounds; round += kRoundsPerCheck) {
        for (int r = 0; r < kRoundsPerCheck; ++r) {
            hipLaunchKernelGGL(k_zero1, dim3(1), dim3(64), 0, st, n_props.as<int>());
            hipLaunchKernelGGL(k_propose, dim3((P + 3) / 4), dim3(256), 0, st, d_d, pitch, N, P,
                               status.as<int>(), last_n.as<int>(), last_v.as<float>(),
                               cnt.as<int>(), props.as<int>(), n_props.as<int>());
            hipLaunchKernelGGL(k_accept, dim3(N), dim3(256), 0, st, d_d, pitch, P, cap,
                               cnt.as<int>(), props.as<int>(), holders.as<int>(),
                               hold_cnt.as<int>(), status.as<int>());
        }
        int made = 0;
        if (hipMemcpyAsync(&made, n_props.p, 4, hipMemcpyDeviceToHost, st) != hipSuccess ||
            hipStreamSynchronize(st) != hipSuccess)
            return err(1, "assign_groups: proposal round failed");
        if (made == 0) break;
    }
    hipLaunchKernelGGL(k_class_out, dim3(N), dim3(256), 0, st, cap, holders.as<int>(),
                       hold_cnt.as<int>(), d_out);
    hipLaunchKernelGGL(k_assign_output, dim3(1), dim3(256), 0, st, d_d, pitch, N, P, cap,
                       status.as<int>(), d_out);
    if (hipGetLastError() != hipSuccess || hipStreamSynchronize(st) != hipSuccess)
        return err(1, "assign_groups: launch failed");
    return 0;
}

}  // namespace isac

extern "C" {

int isac_assign_groups_device(const float *d_d, unsigned int n_ref, unsigned int n_img,
                              unsigned int d_pitch, int *d_out, void *stream) {
    return isac::assign_groups(d_d, (int)n_ref, (int)n_img, (int)d_pitch, d_out,
                               (hipStream_t)stream);
}

int isac_assign_groups(const float *h_d, unsigned int n_ref, unsigned int n_img, int *h_out) {
    if (!h_d || !h_out || n_ref < 1 || n_img < n_ref) {
        isac::report_error(2, "isac_assign_groups: need host buffers and 1 <= n_ref <= n_img");
        return 2;
    }
    void *dd = nullptr, *dout = nullptr;
    const size_t nd = (size_t)n_ref * n_img * sizeof(float), no = ((size_t)n_img + 1) * sizeof(int);
    if (hipMalloc(&dd, nd) != hipSuccess || hipMalloc(&dout, no) != hipSuccess ||
        hipMemcpy(dd, h_d, nd, hipMemcpyHostToDevice) != hipSuccess) {
        if (dd) (void)hipFree(dd);
        if (dout) (void)hipFree(dout);
        isac::report_error(1, "isac_assign_groups: device buffers failed");
        return 1;
    }
    int rc = isac::assign_groups(static_cast<const float *>(dd), (int)n_ref, (int)n_img, (int)n_img,
                                 static_cast<int *>(dout), nullptr);
    if (rc == 0 && hipMemcpy(h_out, dout, no, hipMemcpyDeviceToHost) != hipSuccess) {
        isac::report_error(1, "isac_assign_groups: copy back failed");
        rc = 1;
    }
    (void)hipFree(dd);
    (void)hipFree(dout);
    return rc;
}

}  // extern "C"
