// C-ABI host layer of gpu_aln_pack.so for MI355X: process-global context,
// device memory, and the reference's entry points (see include/gpu_aln_pack.h).
//
// Context lifetime follows cuda/gpu_aln_noref.cu:31-48, :86-180: *_init sets the
// device (fixed for the process), builds the polar template and the shift grid
// and allocates the library-owned AlignParam records; gpu_clear frees
// everything but keeps the device id.
#include "../../include/gpu_aln_pack.h"
#include "isac_internal.h"

#include <algorithm>
#include <cmath>
#include <cstdarg>
#include <cstdio>
#include <cstdlib>
#include <cstring>
#include <climits>
#include <functional>
#include <limits>
#include <map>
#include <mutex>
#include <string>
#include <thread>
#include <vector>

using isac::DevParam;
using isac::PeakRec;
using isac::ShiftView;
using isac::kRingLen;
using isac::kSlots;

static_assert(sizeof(AlignParam) == 24, "AlignParam must keep the reference layout");
static_assert(sizeof(AlignConfig) == 40, "AlignConfig must keep the reference layout");
static_assert(offsetof(AlignParam, shift_x) == 8 && offsetof(AlignParam, mirror) == 20, "layout");

namespace {

// ------------------------------------------------------------------ errors
int g_err = 0;
char g_errmsg[512] = {0};

void set_error(int code, const char *fmt, ...) {
    char buf[512];
    va_list ap;
    va_start(ap, fmt);
    vsnprintf(buf, sizeof(buf), fmt, ap);
    va_end(ap);
    fprintf(stderr, "gpu_aln_pack: error %d: %s\n", code, buf);
    fflush(stderr);
    if (g_err == 0) {
        g_err = code;
        snprintf(g_errmsg, sizeof(g_errmsg), "%s", buf);
    }
    const char *ex = getenv("ISAC_EXIT_ON_ERROR");
    if (ex && ex[0] == '1') exit(EXIT_FAILURE);
}

enum : int {
    kErrHip = 1,
    kErrArg = 2,
    kErrState = 3,
    kErrUnsupported = 4,
};

#define HIP_OK(call)                                                                      \
    do {                                                                                  \
        hipError_t e_ = (call);                                                           \
        if (e_ != hipSuccess) {                                                           \
            set_error(kErrHip, "%s failed: %s (%s:%d)", #call, hipGetErrorString(e_),     \
                      __FILE__, __LINE__);                                                \
            return kErrHip;                                                               \
        }                                                                                 \
    } while (0)

// ------------------------------------------------------------ device buffers
struct DevBuf {
    void *p = nullptr;
    size_t cap = 0;
    hipError_t reserve(size_t bytes) {
        if (bytes <= cap && p) return hipSuccess;
        if (p) (void)hipFree(p);
        p = nullptr;
        cap = 0;
        hipError_t e = hipMalloc(&p, std::max<size_t>(bytes, 256));
        if (e == hipSuccess) cap = std::max<size_t>(bytes, 256);
        return e;
    }
    void release() {
        if (p) (void)hipFree(p);
        p = nullptr;
        cap = 0;
    }
    template <class T>
    T *as() const { return static_cast<T *>(p); }
};

// Batched multi-reference run: concatenated per-subset tables, the group
// workspace, work-queue tables and the captured graph (kept for replay).
struct BatchWs {
    DevBuf shifts_i, shifts_f, tmpl, rspec, rnorm, rsum0;   // per-subset tables
    DevBuf spec, norm, rec;                                // group workspace
    DevBuf jobs, items, qbeg, counters;                    // work queues of every group
    hipGraph_t graph = nullptr;
    hipGraphExec_t exec = nullptr;
    std::vector<unsigned char> key;          // descriptor bytes the graph was captured for
    std::map<std::string, int> tuned;        // group shape signature -> waves per workgroup
    int groups = 0, nodes = 0, replayed = 0, nw = 0, blocks = 0;
    float cand_ms[3] = {0.0f, 0.0f, 0.0f};
    void drop_graph() {
        if (exec) (void)hipGraphExecDestroy(exec);
        if (graph) (void)hipGraphDestroy(graph);
        exec = nullptr;
        graph = nullptr;
        key.clear();
    }
    void release() {
        drop_graph();
        DevBuf *all[] = {&shifts_i, &shifts_f, &tmpl, &rspec, &rnorm, &rsum0, &spec, &norm,
                         &rec, &jobs, &items, &qbeg, &counters};
        for (DevBuf *b : all) b->release();
        tuned.clear();
    }
};

// Work buffers shared by every entry point (grow-only until gpu_clear).
struct Workspace {
    int dev = -1;
    hipStream_t stream = nullptr;
    hipStream_t copy_stream = nullptr;   // host-to-device copies overlapped with the search
    int tmpl_rings = 0;
    DevBuf tmpl, twiddle, zero_shift, shifts_i, shifts_f;
    DevBuf spec, norm, rspec, rnorm, rsum0, rec;
    DevBuf peak, dvals, acc;
    DevBuf mix_p, mix_pn, mix_r, mix_rn;   // 16-bit MFMA operands and row norms
    DevBuf cand, rec_c, best, second, tie, val64;   // candidate re-scoring
    DevBuf items, smask, fb_idx, fb_count, fb_img, fb_acc;   // narrowed search + fallback
    DevBuf edge, smask2;                                     // edge flags, level-2 shifts
    DevBuf pre_tp, pre_mir, pre_part, pre_avg, pre_shift, pre_ref_of, pre_work, pre_twd;
    int pre_twd_dim = 0;
    std::vector<float2> shifts_uploaded;
    BatchWs batch;
    void release() {
        batch.release();
        DevBuf *all[] = {&tmpl, &twiddle, &zero_shift, &shifts_i, &shifts_f, &spec, &norm,
                         &rspec, &rnorm, &rsum0, &rec, &peak, &dvals, &acc,
                         &mix_p, &mix_pn, &mix_r, &mix_rn, &cand, &rec_c, &best, &second,
                         &tie, &val64, &items, &smask, &fb_idx, &fb_count, &fb_img, &fb_acc,
                         &edge, &smask2, &pre_tp, &pre_mir, &pre_part, &pre_avg, &pre_shift,
                         &pre_ref_of, &pre_work, &pre_twd};
        for (DevBuf *b : all) b->release();
        tmpl_rings = 0;
        pre_twd_dim = 0;
        shifts_uploaded.clear();
        if (stream) (void)hipStreamDestroy(stream);
        stream = nullptr;
        if (copy_stream) (void)hipStreamDestroy(copy_stream);
        copy_stream = nullptr;
    }
};

// Process-global alignment context (cuda/gpu_aln_noref.cu:29-48).
struct Context {
    int device = -1;                       // fixed once set (cuda/gpu_aln_noref.cu:103)
    const AlignConfig *cfg = nullptr;      // retained by pointer
    std::vector<float2> shifts;            // shift grid (x-major)
    AlignParam *h_param = nullptr;         // pinned host, returned to the caller
    size_t n_param = 0;
    enum Mode { kNone, kMref, kPre, kRefFree } mode = kNone;
    // pre-align / reference-free device data
    DevBuf sbj_img, ref_img, prm, tr, cid_idx, ref_of, twd, tanl_work;
    unsigned sbj_cap = 0, ref_cap = 0, sbj_count = 0, ref_count = 0;
    int twd_dim = 0;
    // multi_ref_align host staging
    float *h_stage = nullptr;
    size_t h_stage_cap = 0;
    // profiling
    bool prof = false;
    int prof_ccf_n = 0, prof_polar_n = 0;
    double prof_ccf_ms = 0.0, prof_polar_ms = 0.0;
};

Context g;
Workspace ws;

// -------------------------------------------------------------- grid helpers
// cuda/gpu_aln_common.cu:64-86: float loop, x-major.
std::vector<float2> make_shift_grid(float rx, float ry, float step) {
    std::vector<float2> v;
    if (!(step > 0.0f)) return v;
    // a grid beyond any caller's limit (more than 65536 shifts) comes back with
    // 65537 entries, which every shape check rejects, instead of being enumerated
    // (a step far below the range's float spacing would never end the loop)
    const double nx = 2.0 * (double)rx / step + 2.0, ny = 2.0 * (double)ry / step + 2.0;
    if (nx * ny > 65537.0) return std::vector<float2>(65537);
    for (float sx = -rx; sx <= rx; sx += step)
        for (float sy = -ry; sy <= ry; sy += step) v.push_back(make_float2(sx, sy));
    return v;
}

// cuda/gpu_aln_common.cu:53-59 (same float operations).
std::vector<float2> make_template(int R) {
    std::vector<float2> t((size_t)R * kRingLen);
    for (int i = 0; i < R; ++i)
        for (int j = 0; j < kRingLen; ++j) {
            const float angle = (float)((double)((float)j / (float)kRingLen) * 3.14159265358979 * 2);
            t[(size_t)i * kRingLen + j] =
                make_float2(cosf(angle) * (float)(i + 1) * (float)R / (float)R,
                            sinf(angle) * (float)(i + 1) * (float)R / (float)R);
        }
    return t;
}

int ensure_device_ws(int dev) {
    int cur = -1;
    HIP_OK(hipGetDevice(&cur));
    if (ws.dev != cur) {
        ws.release();
        ws.dev = cur;
    }
    if (!ws.stream) HIP_OK(hipStreamCreateWithFlags(&ws.stream, hipStreamNonBlocking));
    (void)dev;
    if (!ws.twiddle.p) {
        std::vector<float2> tw(256);
        for (int m = 0; m < 256; ++m) {
            const double a = -2.0 * M_PI * (double)m / 256.0;
            tw[m] = make_float2((float)cos(a), (float)sin(a));
        }
        HIP_OK(ws.twiddle.reserve(256 * sizeof(float2)));
        HIP_OK(hipMemcpy(ws.twiddle.p, tw.data(), 256 * sizeof(float2), hipMemcpyHostToDevice));
        HIP_OK(ws.zero_shift.reserve(sizeof(int2)));
        HIP_OK(hipMemset(ws.zero_shift.p, 0, sizeof(int2)));
    }
    return 0;
}

int ensure_template(int R) {
    if (ws.tmpl_rings == R && ws.tmpl.p) return 0;
    std::vector<float2> t = make_template(R);
    HIP_OK(ws.tmpl.reserve(t.size() * sizeof(float2)));
    HIP_OK(hipMemcpy(ws.tmpl.p, t.data(), t.size() * sizeof(float2), hipMemcpyHostToDevice));
    ws.tmpl_rings = R;
    return 0;
}

int ensure_shifts(const std::vector<float2> &sh) {
    if (sh.size() == ws.shifts_uploaded.size() &&
        std::equal(sh.begin(), sh.end(), ws.shifts_uploaded.begin(),
                   [](float2 a, float2 b) { return a.x == b.x && a.y == b.y; }))
        return 0;
    std::vector<int2> si(sh.size());
    for (size_t i = 0; i < sh.size(); ++i)  // the kernel argument is int (quirk Q1)
        si[i] = make_int2((int)sh[i].x, (int)sh[i].y);
    HIP_OK(ws.shifts_i.reserve(si.size() * sizeof(int2)));
    HIP_OK(ws.shifts_f.reserve(sh.size() * sizeof(float2)));
    HIP_OK(hipMemcpy(ws.shifts_i.p, si.data(), si.size() * sizeof(int2), hipMemcpyHostToDevice));
    HIP_OK(hipMemcpy(ws.shifts_f.p, sh.data(), sh.size() * sizeof(float2), hipMemcpyHostToDevice));
    ws.shifts_uploaded = sh;
    return 0;
}

int check_shape(int D, int R, int L, size_t S) {
    if (L != kRingLen) {
        set_error(kErrUnsupported, "ring_len %d unsupported (this build computes 256-sample rings)", L);
        return kErrUnsupported;
    }
    if (D < 2 || D > isac::kMaxImgDim) {
        set_error(kErrUnsupported, "img_dim %d outside [2, %d]", D, isac::kMaxImgDim);
        return kErrUnsupported;
    }
    if (R < 1 || R > isac::kMaxRings) {
        set_error(kErrArg, "ring_num %d outside [1, %d]", R, isac::kMaxRings);
        return kErrArg;
    }
    if (S < 1 || S > 65536) {
        set_error(kErrArg, "shift grid of %zu shifts (need 1..65536)", S);
        return kErrArg;
    }
    return 0;
}

size_t free_device_bytes() {
    size_t fr = 0, tot = 0;
    if (hipMemGetInfo(&fr, &tot) != hipSuccess) return 0;
    return fr;
}

// Particle-batch size for the spectra workspace.
int batch_for(size_t per_item, size_t fixed_bytes, float mem_frac, int total) {
    double frac = (mem_frac > 0.0f && mem_frac <= 1.0f) ? mem_frac : 0.9;
    size_t have = (size_t)(frac * (double)(free_device_bytes() + ws.spec.cap + ws.rec.cap + ws.norm.cap));
    const size_t cap = (size_t)48 << 30;  // 20 000 particles of 76 px in one batch
    size_t budget = std::min(have > fixed_bytes ? have - fixed_bytes : 0, cap);
    size_t b = per_item ? budget / per_item : (size_t)total;
    b = std::max<size_t>(b, 1);
    return (int)std::min<size_t>(b, (size_t)total);
}

struct ProfScope {
    hipEvent_t a = nullptr, b = nullptr;
    bool on;
    hipStream_t st;
    explicit ProfScope(bool enabled, hipStream_t s) : on(enabled), st(s) {
        if (on) {
            (void)hipEventCreate(&a);
            (void)hipEventCreate(&b);
            (void)hipEventRecord(a, st);
        }
    }
    double finish() {
        if (!on) return 0.0;
        (void)hipEventRecord(b, st);
        (void)hipEventSynchronize(b);
        float ms = 0.0f;
        (void)hipEventElapsedTime(&ms, a, b);
        (void)hipEventDestroy(a);
        (void)hipEventDestroy(b);
        return ms;
    }
};

// Reference spectra (centre D//2, no shift), normalised and ring-weighted.
int ref_spectra(const float *d_ref, int N, int D, int R, hipStream_t st) {
    HIP_OK(ws.rspec.reserve((size_t)N * R * kSlots * sizeof(float2)));
    HIP_OK(ws.rnorm.reserve((size_t)N * sizeof(float2)));
    HIP_OK(ws.rsum0.reserve((size_t)N * sizeof(float)));
    HIP_OK(isac::launch_polar_spectra(d_ref, N, D, ShiftView{nullptr, 0}, ws.zero_shift.as<int2>(),
                                      1, ws.tmpl.as<float2>(), R, ws.twiddle.as<float2>(),
                                      ws.rspec.as<float2>(), ws.rnorm.as<float2>(), st));
    HIP_OK(isac::launch_ref_finalize(ws.rspec.as<float2>(), ws.rnorm.as<float2>(), N, R,
                                     ws.rsum0.as<float>(), st));
    return 0;
}

// Device-resident multi-reference pipeline.
// before_batch(p0, n), when given, runs on the host before a particle batch's
// launches are queued (the host-pointer entry copies that batch's images there).
// max_batch > 0 caps the particles per batch below what memory allows.
int run_mref(const float *d_sbj, const float *d_ref, ShiftView acc, int P, int N, int D, int R,
             const std::vector<float2> &shifts, float *d_peak, int peak_row, float *d_d,
             int d_pitch, float mem_frac, hipStream_t st,
             const std::function<int(int, int)> &before_batch = {}, int max_batch = 0) {
    const int S = (int)shifts.size();
    if (int e = check_shape(D, R, kRingLen, shifts.size())) return e;
    if (P <= 0 || N <= 0) return 0;
    if (int e = ensure_template(R)) return e;
    if (int e = ensure_shifts(shifts)) return e;
    if (int e = ref_spectra(d_ref, N, D, R, st)) return e;
    const size_t per = (size_t)S * R * kSlots * sizeof(float2) + (size_t)S * sizeof(float2) +
                       (size_t)N * sizeof(PeakRec);
    int pb = batch_for(per, 0, mem_frac, P);
    if (max_batch > 0) pb = std::min(pb, max_batch);
    HIP_OK(ws.spec.reserve((size_t)pb * S * R * kSlots * sizeof(float2)));
    HIP_OK(ws.norm.reserve((size_t)pb * S * sizeof(float2)));
    HIP_OK(ws.rec.reserve((size_t)pb * N * sizeof(PeakRec)));
    for (int p0 = 0; p0 < P; p0 += pb) {
        const int n = std::min(pb, P - p0);
        if (before_batch)
            if (int e = before_batch(p0, n)) return e;
        ShiftView a = acc;
        if (a.base) a.base += (size_t)p0 * a.stride;
        {
            ProfScope ps(g.prof, st);
            HIP_OK(isac::launch_polar_spectra(d_sbj + (size_t)p0 * D * D, n, D, a,
                                              ws.shifts_i.as<int2>(), S, ws.tmpl.as<float2>(), R,
                                              ws.twiddle.as<float2>(), ws.spec.as<float2>(),
                                              ws.norm.as<float2>(), st));
            if (g.prof) { g.prof_polar_ms += ps.finish(); g.prof_polar_n++; }
        }
        {
            ProfScope ps(g.prof, st);
            HIP_OK(isac::launch_ccf_mref(ws.spec.as<float2>(), ws.norm.as<float2>(),
                                         ws.rspec.as<float2>(), ws.rsum0.as<float>(), n, S, R, N,
                                         ws.twiddle.as<float2>(), ws.rec.as<PeakRec>(), st));
            if (g.prof) { g.prof_ccf_ms += ps.finish(); g.prof_ccf_n++; }
        }
        HIP_OK(isac::launch_decode_mref(ws.rec.as<PeakRec>(), n, N, a, ws.shifts_f.as<float2>(), S,
                                        d_peak, peak_row, p0, d_d, d_pitch, st));
    }
    HIP_OK(hipStreamSynchronize(st));
    return 0;
}

// Device-resident mixed-precision multi-reference pipeline: FP16/BF16 MFMA ring
// contraction and FP16 MFMA psi transform; writes the approximate d / peak_list.
struct BestOut {
    int *ref = nullptr;      // [P] best reference (exact FP32 re-scoring of the top candidates)
    float *val = nullptr;    // [P] its peak value
    float *peak = nullptr;   // [P][4] angle, sx, sy, mirror
    int *tie = nullptr;      // [P] 1 when the runner-up was within tie_rel (re-scored in FP64)
    float tie_rel = 1e-5f;
    const int *idx = nullptr;   // [P] or NULL: particle p's outputs go to slot idx[p]
};

constexpr int kCandidates = 8;

int run_mref_mixed(const float *d_sbj, const float *d_ref, ShiftView acc, int P, int N, int D,
                   int R, const std::vector<float2> &shifts, bool bf16, float *d_peak,
                   int peak_row, float *d_d, int d_pitch, float mem_frac, hipStream_t st,
                   const BestOut *bo = nullptr) {
    const int S = (int)shifts.size();
    if (int e = check_shape(D, R, kRingLen, shifts.size())) return e;
    if (P <= 0 || N <= 0) return 0;
    if (int e = ensure_template(R)) return e;
    if (int e = ensure_shifts(shifts)) return e;
    if (int e = ref_spectra(d_ref, N, D, R, st)) return e;
    if (S > 1024) {
        set_error(kErrUnsupported, "mixed-precision search: %d shifts (at most 1024)", S);
        return kErrUnsupported;
    }
    const isac::GemmLayout gl = isac::gemm_layout(R);
    const int NB = (N + 15) / 16;
    HIP_OK(ws.mix_r.reserve((size_t)NB * gl.block_bytes));
    HIP_OK(ws.mix_rn.reserve((size_t)N * sizeof(float)));   // power-of-two scale per reference
    HIP_OK(ws.mix_pn.reserve((size_t)N * sizeof(float)));   // 1 / (c_p c_n) per reference
    HIP_OK(isac::launch_frag_refs(ws.rspec.as<float2>(), N, R, bf16, ws.mix_rn.as<float>(),
                                  ws.mix_pn.as<float>(), ws.mix_r.p, st));
    const size_t op_part = gl.block_bytes / 16 * (size_t)isac::gemm_rows_per_particle(S);
    // operands straight from the images when the rings fit the direct kernel; the
    // FP32 spectra are then only needed for the exact re-scoring of the best form
    const bool direct = isac::polar_frag_supported(R);
    const bool need_spec = bo != nullptr || !direct;
    const size_t per = (need_spec ? (size_t)S * R * kSlots * sizeof(float2) + (size_t)S * sizeof(float2) : 0) +
                       op_part + (size_t)N * sizeof(PeakRec) +
                       kCandidates * (sizeof(PeakRec) + sizeof(int)) + 16 * sizeof(int);
    const int pb = batch_for(per, (size_t)16 * op_part, mem_frac, P);
    if (bo) {
        HIP_OK(ws.cand.reserve((size_t)pb * kCandidates * sizeof(int)));
        HIP_OK(ws.rec_c.reserve((size_t)pb * kCandidates * sizeof(PeakRec)));
        HIP_OK(ws.best.reserve((size_t)pb * sizeof(int)));
        HIP_OK(ws.second.reserve((size_t)pb * sizeof(int)));
        HIP_OK(ws.tie.reserve((size_t)pb * sizeof(int)));
        HIP_OK(ws.val64.reserve((size_t)pb * 2 * sizeof(double)));
    }
    if (need_spec) {
        HIP_OK(ws.spec.reserve((size_t)pb * S * R * kSlots * sizeof(float2)));
        HIP_OK(ws.norm.reserve((size_t)pb * S * sizeof(float2)));
    }
    HIP_OK(ws.mix_p.reserve(isac::gemm_particle_operand_bytes(pb, S, R)));
    HIP_OK(ws.rec.reserve((size_t)pb * N * sizeof(PeakRec)));
    for (int p0 = 0; p0 < P; p0 += pb) {
        const int n = std::min(pb, P - p0);
        ShiftView a = acc;
        if (a.base) a.base += (size_t)p0 * a.stride;
        if (need_spec)
            HIP_OK(isac::launch_polar_spectra(d_sbj + (size_t)p0 * D * D, n, D, a,
                                              ws.shifts_i.as<int2>(), S, ws.tmpl.as<float2>(), R,
                                              ws.twiddle.as<float2>(), ws.spec.as<float2>(),
                                              ws.norm.as<float2>(), st));
        if (direct)
            HIP_OK(isac::launch_polar_frag(d_sbj + (size_t)p0 * D * D, n, D, a,
                                           ws.shifts_i.as<int2>(), S, R, bf16,
                                           ws.twiddle.as<float2>(), ws.mix_p.p, st));
        else
            HIP_OK(isac::launch_frag_rows(ws.spec.as<float2>(), ws.norm.as<float2>(), n, S, R, bf16,
                                          ws.mix_p.p, st));
        {
            ProfScope ps(g.prof, st);
            HIP_OK(isac::launch_mref_gemm(ws.mix_p.p, ws.mix_r.p, ws.mix_pn.as<float>(), n, S, R,
                                          N, bf16, ws.rec.as<PeakRec>(), st));
            if (g.prof) { g.prof_ccf_ms += ps.finish(); g.prof_ccf_n++; }
        }
        if (d_peak)
            HIP_OK(isac::launch_decode_mref(ws.rec.as<PeakRec>(), n, N, a, ws.shifts_f.as<float2>(),
                                            S, d_peak, peak_row, p0, d_d, d_pitch, st));
        if (bo) {
            const int K = kCandidates;
            HIP_OK(isac::launch_topk(ws.rec.as<PeakRec>(), n, N, K, ws.cand.as<int>(), st));
            HIP_OK(isac::launch_ccf_candidates(ws.spec.as<float2>(), ws.norm.as<float2>(),
                                               ws.rspec.as<float2>(), ws.rsum0.as<float>(),
                                               ws.cand.as<int>(), K, n, S, R, N,
                                               ws.twiddle.as<float2>(), ws.rec_c.as<PeakRec>(), st));
            HIP_OK(isac::launch_best_candidate(ws.rec_c.as<PeakRec>(), ws.cand.as<int>(), n, K,
                                               bo->tie_rel, ws.best.as<int>(), ws.second.as<int>(),
                                               ws.tie.as<int>(), st));
            HIP_OK(isac::launch_rescore_f64(ws.spec.as<float2>(), ws.norm.as<float2>(),
                                            ws.rspec.as<float2>(), ws.rec_c.as<PeakRec>(),
                                            ws.cand.as<int>(), ws.best.as<int>(),
                                            ws.second.as<int>(), ws.tie.as<int>(), n, K, S, R, N,
                                            ws.val64.as<double>(), st));
            HIP_OK(isac::launch_decode_best(ws.rec_c.as<PeakRec>(), ws.cand.as<int>(),
                                            ws.best.as<int>(), ws.tie.as<int>(), n, K, a,
                                            ws.shifts_f.as<float2>(), S, p0, bo->ref, bo->val,
                                            bo->peak, bo->tie, st, bo->idx));
        }
    }
    HIP_OK(hipStreamSynchronize(st));
    return 0;
}

// Narrowed search around proposals, in three levels.  Particle p's proposals
// d_prop[p][K] = {ref, mirror, psi bin, shift index}.
//  1. each proposal is searched over the +-hw shift window and the psi samples
//     within psi_half degrees of its bin centre;
//  2. particles whose maximum is below `threshold` or (edge_fb) on the edge of
//     the admitted region get, with `level2`, an exact search of every shift and
//     psi over their distinct proposed references;
//  3. particles still below the threshold (or flagged, without level 2) get the
//     exhaustive search.
// Every level uses the exhaustive kernels' arithmetic for the pairs it covers.
int run_proposal(const float *d_sbj, const float *d_ref, ShiftView acc, int P, int N, int D, int R,
                 const std::vector<float2> &shifts, int nx, int ny, const int *d_prop, int K,
                 int n_bins, float psi_half, int hw, float threshold, int edge_fb, int level2,
                 const float *conf, float conf_min, int *best_ref, float *best_val,
                 float *best_peak, int *fallback, unsigned *n_fb_out, unsigned *n_l2_out,
                 hipStream_t st, int fallback_form = 0) {
    const int S = (int)shifts.size();
    if (int e = check_shape(D, R, kRingLen, shifts.size())) return e;
    if (S > 64 || nx * ny != S) {
        set_error(kErrUnsupported, "narrowed search needs a shift grid of at most 64 shifts (got %d)", S);
        return kErrUnsupported;
    }
    if (n_fb_out) *n_fb_out = 0;
    if (n_l2_out) *n_l2_out = 0;
    if (P <= 0 || N <= 0) return 0;
    if (int e = ensure_template(R)) return e;
    if (int e = ensure_shifts(shifts)) return e;
    if (int e = ref_spectra(d_ref, N, D, R, st)) return e;
    const size_t spec_row = (size_t)S * R * kSlots * sizeof(float2);
    const size_t per = spec_row + (size_t)S * sizeof(float2) +
                       (size_t)K * (sizeof(isac::PropItem) + sizeof(PeakRec)) + sizeof(uint64_t) +
                       sizeof(int);
    const int pb = batch_for(per, (size_t)P * 2 * sizeof(int), 0.9f, P);
    HIP_OK(ws.spec.reserve((size_t)pb * spec_row));
    HIP_OK(ws.norm.reserve((size_t)pb * S * sizeof(float2)));
    HIP_OK(ws.items.reserve((size_t)pb * K * sizeof(isac::PropItem)));
    HIP_OK(ws.rec.reserve((size_t)pb * K * sizeof(PeakRec)));
    HIP_OK(ws.smask.reserve((size_t)pb * sizeof(uint64_t)));
    HIP_OK(ws.edge.reserve((size_t)pb * sizeof(int)));
    HIP_OK(ws.smask2.reserve((size_t)pb * sizeof(uint64_t)));
    HIP_OK(ws.cand.reserve((size_t)pb * K * sizeof(int)));
    HIP_OK(ws.fb_idx.reserve((size_t)P * sizeof(int)));
    HIP_OK(ws.fb_count.reserve(sizeof(int)));
    unsigned n_l2 = 0;
    // level 1
    for (int p0 = 0; p0 < P; p0 += pb) {
        const int n = std::min(pb, P - p0);
        ShiftView a = acc;
        if (a.base) a.base += (size_t)p0 * a.stride;
        HIP_OK(isac::launch_prop_items(d_prop + (size_t)p0 * K * 4, n, K, N, nx, ny, hw, n_bins,
                                       psi_half, ws.items.as<isac::PropItem>(),
                                       ws.smask.as<unsigned long long>(), st));
        HIP_OK(isac::launch_polar_spectra_masked(d_sbj + (size_t)p0 * D * D, n, D, a,
                                                 ws.shifts_i.as<int2>(), S, ws.tmpl.as<float2>(), R,
                                                 ws.twiddle.as<float2>(), ws.spec.as<float2>(),
                                                 ws.norm.as<float2>(),
                                                 ws.smask.as<unsigned long long>(), st));
        {
            ProfScope ps(g.prof, st);
            HIP_OK(isac::launch_ccf_prop(ws.spec.as<float2>(), ws.norm.as<float2>(),
                                         ws.rspec.as<float2>(), ws.rsum0.as<float>(),
                                         ws.items.as<isac::PropItem>(), K, n, S, R, N, nx, ny, hw,
                                         ws.twiddle.as<float2>(), ws.rec.as<PeakRec>(), st));
            if (g.prof) { g.prof_ccf_ms += ps.finish(); g.prof_ccf_n++; }
        }
        HIP_OK(isac::launch_best_prop(ws.rec.as<PeakRec>(), ws.items.as<isac::PropItem>(), n, K, nx,
                                      ny, hw, edge_fb, a, ws.shifts_f.as<float2>(), S, p0, best_ref,
                                      best_val, best_peak, ws.edge.as<int>(), st));
        HIP_OK(isac::launch_fallback_flags(best_val, ws.edge.as<int>(), conf, conf_min, n, threshold,
                                           p0, fallback, st));
        if (!level2) continue;
        // level 2 inside the batch: the missing spectra rows, then every shift and
        // psi of the distinct proposed references of the particles flagged 1
        int n2 = 0;
        HIP_OK(isac::launch_compact(fallback + p0, n, 1, 1, ws.fb_idx.as<int>(), ws.fb_count.as<int>(),
                                    st));
        HIP_OK(hipMemcpyAsync(&n2, ws.fb_count.p, sizeof(int), hipMemcpyDeviceToHost, st));
        HIP_OK(hipStreamSynchronize(st));
        if (n2 == 0) continue;
        n_l2 += (unsigned)n2;
        HIP_OK(isac::launch_prop_refs(d_prop + (size_t)p0 * K * 4, ws.fb_idx.as<int>(), n2, K, N,
                                      ws.cand.as<int>(), st));
        HIP_OK(isac::launch_level2_mask(fallback + p0, n, S, ws.smask.as<unsigned long long>(),
                                        ws.smask2.as<unsigned long long>(), st));
        HIP_OK(isac::launch_polar_spectra_masked(d_sbj + (size_t)p0 * D * D, n, D, a,
                                                 ws.shifts_i.as<int2>(), S, ws.tmpl.as<float2>(), R,
                                                 ws.twiddle.as<float2>(), ws.spec.as<float2>(),
                                                 ws.norm.as<float2>(),
                                                 ws.smask2.as<unsigned long long>(), st));
        HIP_OK(isac::launch_ccf_candidates(ws.spec.as<float2>(), ws.norm.as<float2>(),
                                           ws.rspec.as<float2>(), ws.rsum0.as<float>(),
                                           ws.cand.as<int>(), K, n2, S, R, N, ws.twiddle.as<float2>(),
                                           ws.rec.as<PeakRec>(), st, ws.fb_idx.as<int>()));
        HIP_OK(isac::launch_best_refs_level2(ws.rec.as<PeakRec>(), ws.cand.as<int>(), n2, K, a,
                                             ws.shifts_f.as<float2>(), S, ws.fb_idx.as<int>(), p0,
                                             threshold, best_ref, best_val, best_peak, fallback, st));
    }
    if (n_l2_out) *n_l2_out = n_l2;
    auto flagged = [&](int lo, int hi, int &count) -> int {
        HIP_OK(isac::launch_compact(fallback, P, lo, hi, ws.fb_idx.as<int>(), ws.fb_count.as<int>(), st));
        HIP_OK(hipMemcpyAsync(&count, ws.fb_count.p, sizeof(int), hipMemcpyDeviceToHost, st));
        HIP_OK(hipStreamSynchronize(st));
        return 0;
    };
    const size_t per_fb = (size_t)D * D * sizeof(float) + sizeof(float2) + spec_row +
                          (size_t)S * sizeof(float2) + (size_t)std::max(N, K) * sizeof(PeakRec) +
                          (size_t)K * sizeof(int);
    int n_fb = 0;
    if (int e = flagged(1, INT_MAX, n_fb)) return e;
    // level 3, MFMA form: the mixed-precision search with exact FP32 re-scoring of
    // each particle's 8 best references (FP64 for near-ties), results scattered to
    // the flagged particles
    if (n_fb > 0 && fallback_form != 0) {
        const int fb = batch_for((size_t)D * D * sizeof(float) + sizeof(float2), 0, 0.25f, n_fb);
        HIP_OK(ws.fb_img.reserve((size_t)fb * D * D * sizeof(float)));
        HIP_OK(ws.fb_acc.reserve((size_t)fb * sizeof(float2)));
        for (int j0 = 0; j0 < n_fb; j0 += fb) {
            const int m = std::min(fb, n_fb - j0);
            const int *idx = ws.fb_idx.as<int>() + j0;
            HIP_OK(isac::launch_gather(d_sbj, D, acc, idx, m, ws.fb_img.as<float>(),
                                       ws.fb_acc.as<float2>(), st));
            BestOut bo;
            bo.ref = best_ref;
            bo.val = best_val;
            bo.peak = best_peak;
            bo.idx = idx;
            if (int e = run_mref_mixed(ws.fb_img.as<float>(), d_ref, ShiftView{ws.fb_acc.as<float>(), 2},
                                       m, N, D, R, shifts, fallback_form == 2, nullptr, 0, nullptr,
                                       0, 0.9f, st, &bo))
                return e;
        }
        if (n_fb_out) *n_fb_out = (unsigned)n_fb;
        return 0;
    }
    // level 3: the exhaustive search
    if (n_fb > 0) {
        const int fb = batch_for(per_fb, 0, 0.9f, n_fb);
        HIP_OK(ws.fb_img.reserve((size_t)fb * D * D * sizeof(float)));
        HIP_OK(ws.fb_acc.reserve((size_t)fb * sizeof(float2)));
        HIP_OK(ws.spec.reserve((size_t)fb * spec_row));
        HIP_OK(ws.norm.reserve((size_t)fb * S * sizeof(float2)));
        HIP_OK(ws.rec.reserve((size_t)fb * N * sizeof(PeakRec)));
        for (int j0 = 0; j0 < n_fb; j0 += fb) {
            const int m = std::min(fb, n_fb - j0);
            const int *idx = ws.fb_idx.as<int>() + j0;
            HIP_OK(isac::launch_gather(d_sbj, D, acc, idx, m, ws.fb_img.as<float>(),
                                       ws.fb_acc.as<float2>(), st));
            HIP_OK(isac::launch_polar_spectra(ws.fb_img.as<float>(), m, D,
                                              ShiftView{ws.fb_acc.as<float>(), 2},
                                              ws.shifts_i.as<int2>(), S, ws.tmpl.as<float2>(), R,
                                              ws.twiddle.as<float2>(), ws.spec.as<float2>(),
                                              ws.norm.as<float2>(), st));
            HIP_OK(isac::launch_ccf_mref(ws.spec.as<float2>(), ws.norm.as<float2>(),
                                         ws.rspec.as<float2>(), ws.rsum0.as<float>(), m, S, R, N,
                                         ws.twiddle.as<float2>(), ws.rec.as<PeakRec>(), st));
            HIP_OK(isac::launch_best_refs_scatter(ws.rec.as<PeakRec>(), m, N, ws.fb_acc.as<float2>(),
                                                  ws.shifts_f.as<float2>(), S, idx, best_ref,
                                                  best_val, best_peak, st));
        }
        HIP_OK(hipStreamSynchronize(st));
    }
    if (n_fb_out) *n_fb_out = (unsigned)n_fb;
    return 0;
}

// Device-resident single-reference pipeline (pre-align / reference-free):
// particles d_sbj[0..cnt), accumulated shifts and outputs in prm[0..cnt),
// reference index of particle i at ref_of[i * ref_stride].
int run_single(const float *d_sbj, DevParam *prm, const int *ref_of, int ref_stride, int cnt,
               const float *d_ref, int N, int D, int R, const std::vector<float2> &shifts,
               float shift_limit, hipStream_t st, bool sync = true) {
    const int S = (int)shifts.size();
    if (int e = check_shape(D, R, kRingLen, shifts.size())) return e;
    if (cnt <= 0) return 0;
    if (N <= 0) {
        set_error(kErrState, "no reference has been fetched");
        return kErrState;
    }
    if (int e = ensure_template(R)) return e;
    if (int e = ensure_shifts(shifts)) return e;
    if (int e = ref_spectra(d_ref, N, D, R, st)) return e;
    const size_t per = (size_t)S * R * kSlots * sizeof(float2) + (size_t)S * sizeof(float2) +
                       sizeof(PeakRec);
    const int pb = batch_for(per, 0, 0.9f, cnt);
    HIP_OK(ws.spec.reserve((size_t)pb * S * R * kSlots * sizeof(float2)));
    HIP_OK(ws.norm.reserve((size_t)pb * S * sizeof(float2)));
    HIP_OK(ws.rec.reserve((size_t)pb * sizeof(PeakRec)));
    for (int p0 = 0; p0 < cnt; p0 += pb) {
        const int n = std::min(pb, cnt - p0);
        ShiftView a{&prm[p0].shift_x, (int)(sizeof(DevParam) / sizeof(float))};
        HIP_OK(isac::launch_polar_spectra(d_sbj + (size_t)p0 * D * D, n, D, a,
                                          ws.shifts_i.as<int2>(), S, ws.tmpl.as<float2>(), R,
                                          ws.twiddle.as<float2>(), ws.spec.as<float2>(),
                                          ws.norm.as<float2>(), st));
        HIP_OK(isac::launch_ccf_single(ws.spec.as<float2>(), ws.norm.as<float2>(),
                                       ws.rspec.as<float2>(), ws.rsum0.as<float>(),
                                       ref_of + (size_t)p0 * ref_stride, ref_stride, n, S, R, N,
                                       ws.twiddle.as<float2>(), ws.rec.as<PeakRec>(), st));
        HIP_OK(isac::launch_decode_single(ws.rec.as<PeakRec>(), n, prm + p0,
                                          ws.shifts_f.as<float2>(), S, shift_limit, st));
    }
    if (sync) HIP_OK(hipStreamSynchronize(st));
    return 0;
}

// ------------------------------------------------------------ batched subsets
// Many independent multi-reference subsets (uneven particle count, box, reference
// count, rings, shift range) in one captured graph.  Particles are cut into
// memory groups whose spectra fit the workspace; each group is: the polar
// launches of its slices, one persistent work-queue kernel over every
// (slice, particle, reference block) item, and the decode launches.  Every
// kernel is the one the single-subset path runs, so results are bit-identical.
struct BatchSub {
    const IsacMrefSubset *u;
    int P, N, D, R, S;
    size_t sh_off, tm_off, rs_off, rn_off;   // element offsets into the batch tables
    size_t per;                              // workspace bytes per particle
};
struct BatchSlice {
    int sub, p0, n;
    size_t spec_off, norm_off, rec_off;      // element offsets into the group workspace
};
struct BatchGroup {
    std::vector<BatchSlice> sl;
    size_t spec = 0, norm = 0, rec = 0;      // workspace elements used
    int nw = 4, blocks = 0;
    size_t job_off = 0, item_off = 0, q_off = 0;
    int n_items = 0;
};

// Items (slice, particle, reference block, 0) of a group for `nw` waves per
// workgroup; particle i of the group goes to XCD queue i % 8 with all its
// blocks.  `limit[k]` caps the particles taken from slice k (sampling).
void build_queue(const std::vector<BatchSub> &subs, const BatchGroup &gr, int nw,
                 const std::vector<int> *limit, std::vector<int4> &items, int qbeg[9]) {
    std::vector<int4> q[8];
    int running = 0;
    for (size_t k = 0; k < gr.sl.size(); ++k) {
        const BatchSlice &sl = gr.sl[k];
        const int G = (subs[sl.sub].N + 3) / 4, GB = (G + nw - 1) / nw;
        const int n = limit ? std::min(sl.n, (*limit)[k]) : sl.n;
        for (int i = 0; i < n; ++i, ++running)
            for (int b = 0; b < GB; ++b) q[running & 7].push_back(make_int4((int)k, i, b, 0));
    }
    items.clear();
    for (int x = 0; x < 8; ++x) {
        qbeg[x] = (int)items.size();
        items.insert(items.end(), q[x].begin(), q[x].end());
    }
    qbeg[8] = (int)items.size();
}

std::vector<isac::MrefJob> group_jobs(const std::vector<BatchSub> &subs, const BatchGroup &gr,
                                      int nw) {
    BatchWs &b = ws.batch;
    std::vector<isac::MrefJob> jobs;
    for (const BatchSlice &sl : gr.sl) {
        const BatchSub &u = subs[sl.sub];
        jobs.push_back(isac::mref_job(b.spec.as<float2>() + sl.spec_off,
                                      b.norm.as<float2>() + sl.norm_off,
                                      b.rspec.as<float2>() + u.rs_off, b.rsum0.as<float>() + u.rn_off,
                                      b.rec.as<PeakRec>() + sl.rec_off, sl.n, u.S, u.R, u.N, nw));
    }
    return jobs;
}

ShiftView subset_acc(const BatchSub &u, int p0) {
    ShiftView a{u.u->d_acc_xy, 2};
    if (a.base) a.base += (size_t)p0 * 2;
    return a;
}

int batch_polar(const std::vector<BatchSub> &subs, const BatchSlice &sl, int n, hipStream_t st) {
    BatchWs &b = ws.batch;
    const BatchSub &u = subs[sl.sub];
    HIP_OK(isac::launch_polar_spectra(u.u->d_sbj + (size_t)sl.p0 * u.D * u.D, n, u.D,
                                      subset_acc(u, sl.p0), b.shifts_i.as<int2>() + u.sh_off, u.S,
                                      b.tmpl.as<float2>() + u.tm_off, u.R, ws.twiddle.as<float2>(),
                                      b.spec.as<float2>() + sl.spec_off,
                                      b.norm.as<float2>() + sl.norm_off, st));
    return 0;
}

// Chooses waves per workgroup (1, 2 or 4) for a group by timing the persistent
// kernel on a sample of the group's particles (about four rounds of resident
// wavefronts); the winner is cached by the group's shape signature.
int tune_group(const std::vector<BatchSub> &subs, BatchGroup &gr, int n_cu, bool report,
               hipStream_t st) {
    BatchWs &b = ws.batch;
    std::string sig;
    for (const BatchSlice &sl : gr.sl) {
        const BatchSub &u = subs[sl.sub];
        sig += std::to_string(u.R) + "," + std::to_string(u.S) + "," + std::to_string(u.N) + ";";
    }
    const int cand[3] = {1, 2, 4};
    int bpc[3];
    for (int c = 0; c < 3; ++c) bpc[c] = isac::mref_queue_blocks_per_cu(cand[c]);
    auto grid_for = [&](int c) { return std::max(8, (bpc[c] * n_cu + 7) / 8 * 8); };
    auto it = b.tuned.find(sig);
    if (it != b.tuned.end()) {
        gr.nw = it->second;
        gr.blocks = grid_for(gr.nw == 1 ? 0 : gr.nw == 2 ? 1 : 2);
        return 0;
    }
    // sample: particles from every slice in proportion, ~4 rounds of resident waves
    long long wave_items = 0;
    for (const BatchSlice &sl : gr.sl) wave_items += (long long)sl.n * ((subs[sl.sub].N + 3) / 4);
    const long long want = 4LL * 8 * n_cu;
    std::vector<int> limit;
    for (const BatchSlice &sl : gr.sl) {
        const long long m = wave_items > want ? ((long long)sl.n * want + wave_items - 1) / wave_items
                                              : sl.n;
        limit.push_back((int)std::max<long long>(1, std::min<long long>(m, sl.n)));
    }
    for (size_t k = 0; k < gr.sl.size(); ++k)
        if (int e = batch_polar(subs, gr.sl[k], limit[k], st)) return e;
    hipEvent_t e0, e1;
    HIP_OK(hipEventCreate(&e0));
    HIP_OK(hipEventCreate(&e1));
    float best_ms = 0.0f;
    int best = 4;
    float ms_c[3] = {0.0f, 0.0f, 0.0f};
    int rc = 0;
    for (int c = 0; c < 3 && !rc; ++c) {
        if (bpc[c] <= 0) continue;
        std::vector<isac::MrefJob> jobs = group_jobs(subs, gr, cand[c]);
        std::vector<int4> items;
        int qb[9];
        build_queue(subs, gr, cand[c], &limit, items, qb);
        if (b.jobs.reserve(jobs.size() * sizeof(isac::MrefJob)) != hipSuccess ||
            b.items.reserve(std::max<size_t>(items.size(), 1) * sizeof(int4)) != hipSuccess ||
            b.qbeg.reserve(9 * sizeof(int)) != hipSuccess ||
            b.counters.reserve(8 * sizeof(int)) != hipSuccess) {
            rc = kErrHip;
            break;
        }
        (void)hipMemcpy(b.jobs.p, jobs.data(), jobs.size() * sizeof(isac::MrefJob), hipMemcpyHostToDevice);
        (void)hipMemcpy(b.items.p, items.data(), items.size() * sizeof(int4), hipMemcpyHostToDevice);
        (void)hipMemcpy(b.qbeg.p, qb, 9 * sizeof(int), hipMemcpyHostToDevice);
        for (int rep = 0; rep < 2 && !rc; ++rep) {
            (void)hipEventRecord(e0, st);
            if (isac::launch_ccf_mref_queue(b.jobs.as<isac::MrefJob>(), b.items.as<int4>(),
                                            b.qbeg.as<int>(), b.counters.as<int>(), cand[c],
                                            grid_for(c), ws.twiddle.as<float2>(), st) != hipSuccess)
                rc = kErrHip;
            (void)hipEventRecord(e1, st);
            if (hipEventSynchronize(e1) != hipSuccess) rc = kErrHip;
        }
        (void)hipEventElapsedTime(&ms_c[c], e0, e1);
        if (!rc && (best_ms == 0.0f || ms_c[c] < best_ms)) {
            best_ms = ms_c[c];
            best = cand[c];
        }
    }
    (void)hipEventDestroy(e0);
    (void)hipEventDestroy(e1);
    if (rc) {
        set_error(kErrHip, "batched tuning launch failed");
        return rc;
    }
    gr.nw = best;
    gr.blocks = grid_for(best == 1 ? 0 : best == 2 ? 1 : 2);
    b.tuned[sig] = best;
    if (report) std::copy(ms_c, ms_c + 3, b.cand_ms);
    return 0;
}

// Group workspace cap: 64 GiB (a config-4 batch of 182 053 particles in 6 groups,
// one persistent launch each), or ISAC_BATCH_WORKSPACE_MB when set; the group
// budget is also bounded by the free device memory.
size_t batch_cap() {
    size_t cap = (size_t)64 << 30;
    if (const char *env = getenv("ISAC_BATCH_WORKSPACE_MB")) {
        const long long mb = atoll(env);
        if (mb > 0) cap = (size_t)mb << 20;
    }
    return cap;
}

// Validates the descriptors, builds the per-subset tables and the groups.
int batch_plan(const IsacMrefSubset *d, unsigned n, std::vector<BatchSub> &subs,
               std::vector<BatchGroup> &groups, hipStream_t st) {
    BatchWs &b = ws.batch;
    std::vector<int2> shi;
    std::vector<float2> shf, tm;
    size_t rs = 0, rn = 0;
    for (unsigned i = 0; i < n; ++i) {
        const IsacMrefSubset &x = d[i];
        std::vector<float2> sh = make_shift_grid(x.shift_rng_x, x.shift_rng_y, x.shift_step);
        if (int e = check_shape((int)x.img_dim, (int)x.ring_num, kRingLen, sh.size())) return e;
        if (x.n_sbj == 0 || x.n_ref == 0) continue;
        if (!x.d_sbj || !x.d_ref || !x.d_peak_list || !x.d_d || x.d_pitch < x.n_sbj ||
            x.n_sbj > (1u << 30) || x.n_ref > (1u << 20)) {
            set_error(kErrArg, "subset %u: bad pointers or sizes (d_pitch < n_sbj?)", i);
            return kErrArg;
        }
        BatchSub u;
        u.u = &x;
        u.P = (int)x.n_sbj;
        u.N = (int)x.n_ref;
        u.D = (int)x.img_dim;
        u.R = (int)x.ring_num;
        u.S = (int)sh.size();
        u.sh_off = shf.size();
        u.tm_off = tm.size();
        u.rs_off = rs;
        u.rn_off = rn;
        u.per = (size_t)u.S * u.R * kSlots * sizeof(float2) + (size_t)u.S * sizeof(float2) +
                (size_t)u.N * sizeof(PeakRec);
        for (const float2 &v : sh) {   // the kernel argument is int (quirk Q1)
            shi.push_back(make_int2((int)v.x, (int)v.y));
            shf.push_back(v);
        }
        std::vector<float2> t = make_template(u.R);
        tm.insert(tm.end(), t.begin(), t.end());
        rs += (size_t)u.N * u.R * kSlots;
        rn += (size_t)u.N;
        subs.push_back(u);
    }
    if (subs.empty()) return 0;
    HIP_OK(b.shifts_i.reserve(shi.size() * sizeof(int2)));
    HIP_OK(b.shifts_f.reserve(shf.size() * sizeof(float2)));
    HIP_OK(b.tmpl.reserve(tm.size() * sizeof(float2)));
    HIP_OK(b.rspec.reserve(rs * sizeof(float2)));
    HIP_OK(b.rnorm.reserve(rn * sizeof(float2)));
    HIP_OK(b.rsum0.reserve(rn * sizeof(float)));
    HIP_OK(hipMemcpy(b.shifts_i.p, shi.data(), shi.size() * sizeof(int2), hipMemcpyHostToDevice));
    HIP_OK(hipMemcpy(b.shifts_f.p, shf.data(), shf.size() * sizeof(float2), hipMemcpyHostToDevice));
    HIP_OK(hipMemcpy(b.tmpl.p, tm.data(), tm.size() * sizeof(float2), hipMemcpyHostToDevice));

    // groups: subsets by work per particle, largest first, cut to the budget
    std::vector<int> order(subs.size());
    for (size_t i = 0; i < order.size(); ++i) order[i] = (int)i;
    std::stable_sort(order.begin(), order.end(), [&](int a, int c) {
        return (long long)subs[a].S * subs[a].R * subs[a].N > (long long)subs[c].S * subs[c].R * subs[c].N;
    });
    size_t max_per = 0;
    for (const BatchSub &u : subs) max_per = std::max(max_per, u.per);
    const size_t have = (size_t)(0.9 * (double)(free_device_bytes() + b.spec.cap + b.norm.cap + b.rec.cap));
    const size_t budget = std::max(std::min(have, batch_cap()), max_per);
    BatchGroup cur;
    size_t used = 0;
    for (int si : order) {
        const BatchSub &u = subs[si];
        int p0 = 0;
        while (p0 < u.P) {
            int fit = (int)std::min<size_t>((budget - used) / u.per, (size_t)(u.P - p0));
            if (fit <= 0) {
                groups.push_back(cur);
                cur = BatchGroup();
                used = 0;
                continue;
            }
            BatchSlice sl{si, p0, fit, cur.spec, cur.norm, cur.rec};
            cur.sl.push_back(sl);
            cur.spec += (size_t)fit * u.S * u.R * kSlots;
            cur.norm += (size_t)fit * u.S;
            cur.rec += (size_t)fit * u.N;
            used += (size_t)fit * u.per;
            p0 += fit;
        }
    }
    if (!cur.sl.empty()) groups.push_back(cur);
    size_t spec = 0, norm = 0, rec = 0;
    for (const BatchGroup &gr : groups) {
        spec = std::max(spec, gr.spec);
        norm = std::max(norm, gr.norm);
        rec = std::max(rec, gr.rec);
    }
    HIP_OK(b.spec.reserve(spec * sizeof(float2)));
    HIP_OK(b.norm.reserve(norm * sizeof(float2)));
    HIP_OK(b.rec.reserve(rec * sizeof(PeakRec)));
    (void)st;
    return 0;
}

// Records the whole batch on `st` (being captured).
int batch_record(const std::vector<BatchSub> &subs, const std::vector<BatchGroup> &groups,
                 hipStream_t st) {
    BatchWs &b = ws.batch;
    for (const BatchSub &u : subs) {
        HIP_OK(isac::launch_polar_spectra(u.u->d_ref, u.N, u.D, ShiftView{nullptr, 0},
                                          ws.zero_shift.as<int2>(), 1, b.tmpl.as<float2>() + u.tm_off,
                                          u.R, ws.twiddle.as<float2>(), b.rspec.as<float2>() + u.rs_off,
                                          b.rnorm.as<float2>() + u.rn_off, st));
        HIP_OK(isac::launch_ref_finalize(b.rspec.as<float2>() + u.rs_off,
                                         b.rnorm.as<float2>() + u.rn_off, u.N, u.R,
                                         b.rsum0.as<float>() + u.rn_off, st));
    }
    for (size_t gi = 0; gi < groups.size(); ++gi) {
        const BatchGroup &gr = groups[gi];
        for (const BatchSlice &sl : gr.sl)
            if (int e = batch_polar(subs, sl, sl.n, st)) return e;
        HIP_OK(isac::launch_ccf_mref_queue(b.jobs.as<isac::MrefJob>() + gr.job_off,
                                           b.items.as<int4>() + gr.item_off,
                                           b.qbeg.as<int>() + gr.q_off, b.counters.as<int>() + 8 * gi,
                                           gr.nw, gr.blocks, ws.twiddle.as<float2>(), st));
        for (const BatchSlice &sl : gr.sl) {
            const BatchSub &u = subs[sl.sub];
            HIP_OK(isac::launch_decode_mref(b.rec.as<PeakRec>() + sl.rec_off, sl.n, u.N,
                                            subset_acc(u, sl.p0), b.shifts_f.as<float2>() + u.sh_off,
                                            u.S, u.u->d_peak_list, 4 * u.P, sl.p0, u.u->d_d,
                                            (int)u.u->d_pitch, st));
        }
    }
    return 0;
}

int run_batched_graph(const IsacMrefSubset *d, unsigned n, hipStream_t st) {
    BatchWs &b = ws.batch;
    const unsigned char *kb = reinterpret_cast<const unsigned char *>(d);
    std::vector<unsigned char> key(kb, kb + (size_t)n * sizeof(IsacMrefSubset));
    const size_t cap = batch_cap();
    const unsigned char *cb = reinterpret_cast<const unsigned char *>(&cap);
    key.insert(key.end(), cb, cb + sizeof(cap));
    if (b.exec && key == b.key) {
        HIP_OK(hipGraphLaunch(b.exec, st));
        HIP_OK(hipStreamSynchronize(st));
        b.replayed = 1;
        return 0;
    }
    b.drop_graph();
    b.replayed = 0;
    std::vector<BatchSub> subs;
    std::vector<BatchGroup> groups;
    if (int e = batch_plan(d, n, subs, groups, ws.stream)) return e;
    b.groups = (int)groups.size();
    b.nodes = 0;
    if (subs.empty()) return 0;
    HIP_OK(isac::prepare_launches());
    int dev = 0, n_cu = 0;
    HIP_OK(hipGetDevice(&dev));
    HIP_OK(hipDeviceGetAttribute(&n_cu, hipDeviceAttributeMultiprocessorCount, dev));
    // reference spectra first (the tuning sample needs them)
    for (const BatchSub &u : subs) {
        HIP_OK(isac::launch_polar_spectra(u.u->d_ref, u.N, u.D, ShiftView{nullptr, 0},
                                          ws.zero_shift.as<int2>(), 1, b.tmpl.as<float2>() + u.tm_off,
                                          u.R, ws.twiddle.as<float2>(), b.rspec.as<float2>() + u.rs_off,
                                          b.rnorm.as<float2>() + u.rn_off, ws.stream));
        HIP_OK(isac::launch_ref_finalize(b.rspec.as<float2>() + u.rs_off,
                                         b.rnorm.as<float2>() + u.rn_off, u.N, u.R,
                                         b.rsum0.as<float>() + u.rn_off, ws.stream));
    }
    for (size_t gi = 0; gi < groups.size(); ++gi)
        if (int e = tune_group(subs, groups[gi], n_cu, gi == 0, ws.stream)) return e;
    b.nw = groups[0].nw;
    b.blocks = groups[0].blocks;
    // work-queue tables of every group
    std::vector<isac::MrefJob> jobs;
    std::vector<int4> items;
    std::vector<int> qbeg;
    for (BatchGroup &gr : groups) {
        std::vector<isac::MrefJob> j = group_jobs(subs, gr, gr.nw);
        std::vector<int4> it;
        int qb[9];
        build_queue(subs, gr, gr.nw, nullptr, it, qb);
        gr.job_off = jobs.size();
        gr.item_off = items.size();
        gr.q_off = qbeg.size();
        gr.n_items = (int)it.size();
        jobs.insert(jobs.end(), j.begin(), j.end());
        items.insert(items.end(), it.begin(), it.end());
        qbeg.insert(qbeg.end(), qb, qb + 9);
    }
    HIP_OK(b.jobs.reserve(jobs.size() * sizeof(isac::MrefJob)));
    HIP_OK(b.items.reserve(std::max<size_t>(items.size(), 1) * sizeof(int4)));
    HIP_OK(b.qbeg.reserve(qbeg.size() * sizeof(int)));
    HIP_OK(b.counters.reserve(groups.size() * 8 * sizeof(int)));
    HIP_OK(hipMemcpy(b.jobs.p, jobs.data(), jobs.size() * sizeof(isac::MrefJob), hipMemcpyHostToDevice));
    HIP_OK(hipMemcpy(b.items.p, items.data(), items.size() * sizeof(int4), hipMemcpyHostToDevice));
    HIP_OK(hipMemcpy(b.qbeg.p, qbeg.data(), qbeg.size() * sizeof(int), hipMemcpyHostToDevice));
    HIP_OK(hipStreamSynchronize(ws.stream));
    // capture
    HIP_OK(hipStreamBeginCapture(ws.stream, hipStreamCaptureModeRelaxed));
    const int rec_rc = batch_record(subs, groups, ws.stream);
    hipGraph_t graph = nullptr;
    const hipError_t ec = hipStreamEndCapture(ws.stream, &graph);
    if (rec_rc || ec != hipSuccess) {
        if (graph) (void)hipGraphDestroy(graph);
        if (!rec_rc) set_error(kErrHip, "graph capture failed: %s", hipGetErrorString(ec));
        return rec_rc ? rec_rc : kErrHip;
    }
    b.graph = graph;
    size_t nodes = 0;
    (void)hipGraphGetNodes(graph, nullptr, &nodes);
    b.nodes = (int)nodes;
    HIP_OK(hipGraphInstantiate(&b.exec, graph, nullptr, nullptr, 0));
    b.key = key;
    HIP_OK(hipGraphLaunch(b.exec, st));
    HIP_OK(hipStreamSynchronize(st));
    return 0;
}

// cuda/gpu_aln_noref.cu:86-115
int alignment_init(const AlignConfig *cfg, unsigned dev) {
    if (!cfg) {
        set_error(kErrArg, "NULL AlignConfig");
        return kErrArg;
    }
    if (g.device != -1 && g.device != (int)dev) {
        set_error(kErrState, "device %u requested but this process already uses device %d", dev,
                  g.device);
        return kErrState;
    }
    if (g.mode != Context::kNone) {
        set_error(kErrState, "a previous *_init was not followed by gpu_clear()");
        return kErrState;
    }
    HIP_OK(hipSetDevice((int)dev));
    g.device = (int)dev;
    g.cfg = cfg;
    g.shifts = make_shift_grid(cfg->shift_rng_x, cfg->shift_rng_y, cfg->shift_step);
    if (g.shifts.empty()) {
        set_error(kErrArg, "empty shift grid (step %g)", (double)cfg->shift_step);
        return kErrArg;
    }
    if (std::fmod(cfg->shift_rng_x, cfg->shift_step) != 0.0f ||
        std::fmod(cfg->shift_rng_y, cfg->shift_step) != 0.0f) {
        // cuda/gpu_aln_noref.cu:2380-2381 asserts this
        set_error(kErrArg, "shift range (%g, %g) is not a multiple of the step %g",
                  (double)cfg->shift_rng_x, (double)cfg->shift_rng_y, (double)cfg->shift_step);
        return kErrArg;
    }
    return ensure_device_ws((int)dev);
}

int alloc_params(size_t n, int ref_id) {
    if (g.h_param) (void)hipHostFree(g.h_param);
    g.h_param = nullptr;
    g.n_param = 0;
    if (n == 0) return 0;
    HIP_OK(hipHostMalloc((void **)&g.h_param, n * sizeof(AlignParam), hipHostMallocDefault));
    for (size_t i = 0; i < n; ++i) {
        g.h_param[i].sbj_id = -1;
        g.h_param[i].ref_id = ref_id;
        g.h_param[i].shift_x = 0.0f;
        g.h_param[i].shift_y = 0.0f;
        g.h_param[i].angle = 0.0f;
        g.h_param[i].mirror = false;
    }
    g.n_param = n;
    return 0;
}

// Host image list checks and the pinned staging buffer of the host-pointer entries.
int check_image_list(const float **imgs, unsigned n) {
    if (n == 0) return 0;
    if (!imgs) {
        set_error(kErrArg, "NULL image pointer array");
        return kErrArg;
    }
    for (unsigned i = 0; i < n; ++i)
        if (!imgs[i]) {
            set_error(kErrArg, "NULL image pointer at index %u", i);
            return kErrArg;
        }
    return 0;
}

int reserve_stage(size_t need) {
    if (g.h_stage_cap < need) {
        if (g.h_stage) (void)hipHostFree(g.h_stage);
        g.h_stage = nullptr;
        g.h_stage_cap = 0;
        HIP_OK(hipHostMalloc((void **)&g.h_stage, need, hipHostMallocDefault));
        g.h_stage_cap = need;
    }
    return 0;
}

// Gathers images i0 .. i0+n-1 into the staging buffer (same offsets), split over
// up to 8 host threads for large ranges.
void gather_images(const float **imgs, unsigned i0, unsigned n, size_t img_bytes) {
    char *dst = reinterpret_cast<char *>(g.h_stage);
    auto part = [&](unsigned a, unsigned b) {
        for (unsigned i = a; i < b; ++i) memcpy(dst + (size_t)i * img_bytes, imgs[i], img_bytes);
    };
    const unsigned hw = std::max(1u, std::min(8u, std::thread::hardware_concurrency()));
    const unsigned nt = (size_t)n * img_bytes < ((size_t)8 << 20) ? 1u : std::min(hw, n);
    if (nt <= 1) {
        part(i0, i0 + n);
        return;
    }
    std::vector<std::thread> th;
    for (unsigned t = 0; t < nt; ++t)
        th.emplace_back(part, i0 + (unsigned)((size_t)n * t / nt), i0 + (unsigned)((size_t)n * (t + 1) / nt));
    for (auto &x : th) x.join();
}

int upload_images(DevBuf &dst, const float **imgs, unsigned n, int D, hipStream_t st) {
    if (n == 0) return 0;
    if (int e = check_image_list(imgs, n)) return e;
    const size_t img_bytes = (size_t)D * D * sizeof(float);
    const size_t need = (size_t)n * img_bytes;
    if (int e = reserve_stage(need)) return e;
    gather_images(imgs, 0, n, img_bytes);
    HIP_OK(dst.reserve(need));
    HIP_OK(hipMemcpyAsync(dst.p, g.h_stage, need, hipMemcpyHostToDevice, st));
    HIP_OK(hipStreamSynchronize(st));
    return 0;
}

int upload_params(size_t first, size_t cnt, hipStream_t st) {
    HIP_OK(g.prm.reserve(std::max<size_t>(cnt, 1) * sizeof(DevParam)));
    if (cnt == 0) return 0;
    HIP_OK(hipMemcpyAsync(g.prm.p, g.h_param + first, cnt * sizeof(AlignParam),
                          hipMemcpyHostToDevice, st));
    return 0;
}

int download_params(size_t first, size_t cnt, hipStream_t st) {
    if (cnt == 0) return 0;
    HIP_OK(hipMemcpyAsync(g.h_param + first, g.prm.p, cnt * sizeof(AlignParam),
                          hipMemcpyDeviceToHost, st));
    HIP_OK(hipStreamSynchronize(st));
    return 0;
}

bool require_mode(Context::Mode m, const char *fn) {
    if (g.mode != m || !g.cfg) {
        set_error(kErrState, "%s called without the matching *_init", fn);
        return false;
    }
    return true;
}

size_t spec_bytes_per_particle(const AlignConfig *c) {
    const size_t S = make_shift_grid(c->shift_rng_x, c->shift_rng_y, c->shift_step).size();
    return S * c->ring_num * kSlots * sizeof(float2) + S * sizeof(float2) + sizeof(PeakRec);
}

// --------------------------------------------------- pre-alignment loop (device)
// bin/sp_isac_applications.py:1486-1693 with every iteration on the device: per
// shift step the grid of (range, step) (reset_shifts); per iteration the average
// of the particles transformed with their current parameters, optionally moved
// by the driver's average-shift centring (center == -1, from the parameters of
// the previous alignment) and low-passed (cutoff > 0), then every particle
// aligned to it (pre_align_run); the last iteration of the last step only builds
// the average (:1563).  Nothing returns to the host between iterations.
int run_pre_align_loop(const float *d_imgs, int P, int D, int R, const float *ranges,
                       const float *steps, int n_steps, int max_iter, int center, float cutoff,
                       float falloff, DevParam *prm, float *d_average, hipStream_t st) {
    if (P <= 0 || n_steps <= 0 || max_iter <= 0) return 0;
    // every step's grid is checked before any work, so an invalid step leaves the
    // caller's parameters untouched
    std::vector<std::vector<float2>> grids(n_steps);
    for (int n = 0; n < n_steps; ++n) {
        grids[n] = make_shift_grid(ranges[n], ranges[n], steps[n]);
        if (int e = check_shape(D, R, kRingLen, grids[n].size())) return e;
    }
    HIP_OK(ws.pre_tp.reserve((size_t)P * sizeof(float4)));
    HIP_OK(ws.pre_mir.reserve((size_t)P * sizeof(int)));
    HIP_OK(ws.pre_part.reserve(isac::pre_average_scratch_floats(P, D) * sizeof(float)));
    HIP_OK(ws.pre_avg.reserve((size_t)D * D * sizeof(float)));
    HIP_OK(ws.pre_shift.reserve(sizeof(float2)));
    HIP_OK(ws.pre_ref_of.reserve((size_t)P * sizeof(int)));
    HIP_OK(ws.pre_work.reserve((size_t)D * (D / 2 + 1) * sizeof(float2)));
    HIP_OK(hipMemsetAsync(ws.pre_ref_of.p, 0, (size_t)P * sizeof(int), st));   // one reference
    if (ws.pre_twd_dim != D) {
        std::vector<float2> tw(D);
        for (int m = 0; m < D; ++m) {
            const double a = -2.0 * M_PI * (double)m / (double)D;
            tw[m] = make_float2((float)cos(a), (float)sin(a));
        }
        HIP_OK(ws.pre_twd.reserve(D * sizeof(float2)));
        HIP_OK(hipMemcpy(ws.pre_twd.p, tw.data(), D * sizeof(float2), hipMemcpyHostToDevice));
        ws.pre_twd_dim = D;
    }
    float *avg = ws.pre_avg.as<float>();
    const float lim = (float)(D - R - 2);
    const int total = n_steps * max_iter;
    int it_glob = 0;
    for (int n = 0; n < n_steps; ++n) {
        const std::vector<float2> &sh = grids[n];
        for (int it = 0; it < max_iter; ++it) {
            ++it_glob;
            HIP_OK(isac::launch_pre_average(d_imgs, P, D, prm, ws.pre_tp.as<float4>(),
                                            ws.pre_mir.as<int>(), ws.pre_part.as<float>(), avg, st));
            // the driver's sums start at zero; later they come from the last alignment
            const bool shift_avg = (center == -1) && it_glob > 1;
            if (shift_avg) HIP_OK(isac::launch_pre_center(prm, P, ws.pre_shift.as<float2>(), st));
            if (shift_avg || cutoff > 0.0f)
                HIP_OK(isac::launch_filter_shift(avg, 1, D, cutoff, falloff,
                                                 shift_avg ? ws.pre_shift.as<float2>() : nullptr,
                                                 ws.pre_work.as<float2>(), ws.pre_twd.as<float2>(),
                                                 st));
            if (it_glob != total)
                if (int e = run_single(d_imgs, prm, ws.pre_ref_of.as<int>(), 1, P, avg, 1, D, R, sh,
                                       lim, st, false))
                    return e;
        }
    }
    if (d_average)
        HIP_OK(hipMemcpyAsync(d_average, avg, (size_t)D * D * sizeof(float), hipMemcpyDeviceToDevice,
                              st));
    HIP_OK(hipStreamSynchronize(st));
    return 0;
}

// ------------------------------------------------- allocation-free entry point
// isac_mref_align_device_noalloc works in one fixed device array per GPU
// (`scratch` below, kNoAllocBytes): tables, reference spectra and one particle
// batch of spectra and records at a time.  Its device address is looked up once
// per device with hipGetSymbolAddress (the array's name in host code is only a
// host-side handle) and kept in g_scratch_dev.
constexpr size_t kNoAllocBytes = ISAC_NOALLOC_SCRATCH_BYTES;
constexpr int kNoAllocMaxDevices = 64;
std::mutex g_scratch_mu;
unsigned char *g_scratch_dev[kNoAllocMaxDevices] = {};

size_t align256(size_t b) { return (b + 255) & ~(size_t)255; }

// Byte offsets of the regions of the scratch array for one call.
struct NoAllocLayout {
    size_t twiddle, tmpl, zero_shift, shifts_i, shifts_f, rspec, rnorm, rsum0;
    size_t spec, norm, rec;   // one particle batch of pb particles
    size_t end;
    int pb;
};

// Places the regions; pb = the largest batch that fits (0 when not even one does).
NoAllocLayout noalloc_layout(int N, int R, int S, int P) {
    NoAllocLayout L{};
    size_t o = 0;
    auto take = [&](size_t bytes) { const size_t at = o; o += align256(bytes); return at; };
    L.twiddle = take(256 * sizeof(float2));
    L.tmpl = take((size_t)R * kRingLen * sizeof(float2));
    L.zero_shift = take(sizeof(int2));
    L.shifts_i = take((size_t)S * sizeof(int2));
    L.shifts_f = take((size_t)S * sizeof(float2));
    L.rspec = take((size_t)N * R * kSlots * sizeof(float2));
    L.rnorm = take((size_t)N * sizeof(float2));
    L.rsum0 = take((size_t)N * sizeof(float));
    const size_t fixed = o;
    auto batch_bytes = [&](size_t pb) {
        return align256(pb * S * R * kSlots * sizeof(float2)) + align256(pb * S * sizeof(float2)) +
               align256(pb * N * sizeof(PeakRec));
    };
    size_t pb = 0;
    if (fixed < kNoAllocBytes) {
        const size_t per = (size_t)S * R * kSlots * sizeof(float2) + (size_t)S * sizeof(float2) +
                           (size_t)N * sizeof(PeakRec);
        pb = std::min<size_t>((kNoAllocBytes - fixed) / per, (size_t)P);
        while (pb > 0 && fixed + batch_bytes(pb) > kNoAllocBytes) --pb;   // alignment slack
    }
    L.pb = (int)pb;
    L.spec = take(pb * S * R * kSlots * sizeof(float2));
    L.norm = take(pb * S * sizeof(float2));
    L.rec = take(pb * N * sizeof(PeakRec));
    L.end = o;
    return L;
}

}  // namespace

// The allocation-free entry point's device memory (one copy per GPU, from module
// load until the process ends).
static __device__ __align__(256) unsigned char scratch[kNoAllocBytes];

namespace {

// Device address of `scratch` on the current device (looked up on the first call
// there, with that device current).
int scratch_on_current_device(unsigned char **out) {
    int dev = -1;
    HIP_OK(hipGetDevice(&dev));
    if (dev < 0 || dev >= kNoAllocMaxDevices) {
        set_error(kErrUnsupported, "device %d: the allocation-free entry supports devices 0..%d",
                  dev, kNoAllocMaxDevices - 1);
        return kErrUnsupported;
    }
    std::lock_guard<std::mutex> lock(g_scratch_mu);
    if (!g_scratch_dev[dev]) {
        void *p = nullptr;
        HIP_OK(hipGetSymbolAddress(&p, HIP_SYMBOL(scratch)));
        if (!p) {
            set_error(kErrHip, "hipGetSymbolAddress returned a null address for the scratch array");
            return kErrHip;
        }
        g_scratch_dev[dev] = static_cast<unsigned char *>(p);
    }
    *out = g_scratch_dev[dev];
    return 0;
}

// run_mref's pipeline in the scratch array: the same kernels on the same inputs
// (results bit-identical to isac_mref_align_device), particle batches of at most
// NoAllocLayout::pb particles, no device allocation, tables uploaded per call.
int run_mref_noalloc(const float *d_sbj, const float *d_ref, ShiftView acc, int P, int N, int D,
                     int R, const std::vector<float2> &shifts, float *d_peak, int peak_row,
                     float *d_d, int d_pitch, hipStream_t st) {
    const int S = (int)shifts.size();
    unsigned char *base = nullptr;
    if (int e = scratch_on_current_device(&base)) return e;
    const NoAllocLayout L = noalloc_layout(N, R, S, P);
    if (L.pb < 1) {
        set_error(kErrUnsupported,
                  "isac_mref_align_device_noalloc: %d references x %d rings x %d shifts need more "
                  "than its %zu-byte scratch array for one particle",
                  N, R, S, kNoAllocBytes);
        return kErrUnsupported;
    }
    std::vector<float2> tw(256);
    for (int m = 0; m < 256; ++m) {   // the twiddle table of ensure_device_ws
        const double a = -2.0 * M_PI * (double)m / 256.0;
        tw[m] = make_float2((float)cos(a), (float)sin(a));
    }
    const std::vector<float2> tm = make_template(R);
    std::vector<int2> si(shifts.size());
    for (size_t i = 0; i < shifts.size(); ++i) si[i] = make_int2((int)shifts[i].x, (int)shifts[i].y);
    const int2 zero{0, 0};
    auto at = [&](size_t off) { return static_cast<void *>(base + off); };
    HIP_OK(hipMemcpyAsync(at(L.twiddle), tw.data(), 256 * sizeof(float2), hipMemcpyHostToDevice, st));
    HIP_OK(hipMemcpyAsync(at(L.tmpl), tm.data(), tm.size() * sizeof(float2), hipMemcpyHostToDevice,
                          st));
    HIP_OK(hipMemcpyAsync(at(L.zero_shift), &zero, sizeof(int2), hipMemcpyHostToDevice, st));
    HIP_OK(hipMemcpyAsync(at(L.shifts_i), si.data(), si.size() * sizeof(int2),
                          hipMemcpyHostToDevice, st));
    HIP_OK(hipMemcpyAsync(at(L.shifts_f), shifts.data(), shifts.size() * sizeof(float2),
                          hipMemcpyHostToDevice, st));
    const float2 *twd = static_cast<const float2 *>(at(L.twiddle));
    const float2 *tmpl = static_cast<const float2 *>(at(L.tmpl));
    float2 *rspec = static_cast<float2 *>(at(L.rspec));
    float2 *rnorm = static_cast<float2 *>(at(L.rnorm));
    float *rsum0 = static_cast<float *>(at(L.rsum0));
    float2 *spec = static_cast<float2 *>(at(L.spec));
    float2 *norm = static_cast<float2 *>(at(L.norm));
    PeakRec *rec = static_cast<PeakRec *>(at(L.rec));
    const int2 *shi = static_cast<const int2 *>(at(L.shifts_i));
    const float2 *shf = static_cast<const float2 *>(at(L.shifts_f));
    HIP_OK(isac::launch_polar_spectra(d_ref, N, D, ShiftView{nullptr, 0},
                                      static_cast<const int2 *>(at(L.zero_shift)), 1, tmpl, R, twd,
                                      rspec, rnorm, st));
    HIP_OK(isac::launch_ref_finalize(rspec, rnorm, N, R, rsum0, st));
    for (int p0 = 0; p0 < P; p0 += L.pb) {
        const int n = std::min(L.pb, P - p0);
        ShiftView a = acc;
        if (a.base) a.base += (size_t)p0 * a.stride;
        HIP_OK(isac::launch_polar_spectra(d_sbj + (size_t)p0 * D * D, n, D, a, shi, S, tmpl, R, twd,
                                          spec, norm, st));
        HIP_OK(isac::launch_ccf_mref(spec, norm, rspec, rsum0, n, S, R, N, twd, rec, st));
        HIP_OK(isac::launch_decode_mref(rec, n, N, a, shf, S, d_peak, peak_row, p0, d_d, d_pitch,
                                        st));
    }
    // the host tables above must outlive their copies; the call returns on completion
    HIP_OK(hipStreamSynchronize(st));
    return 0;
}

}  // namespace

namespace isac {
void report_error(int code, const char *msg) { set_error(code, "%s", msg); }
}  // namespace isac

// ======================================================================= ABI
extern "C" {

int gpu_last_error(void) { return g_err; }
const char *gpu_last_error_string(void) { return g_errmsg; }
void gpu_clear_error(void) {
    g_err = 0;
    g_errmsg[0] = 0;
}

void print_gpu_info(unsigned int device_idx) {
    int drv = 0, rt = 0;
    (void)hipDriverGetVersion(&drv);
    (void)hipRuntimeGetVersion(&rt);
    hipDeviceProp_t pr;
    if (hipGetDeviceProperties(&pr, (int)device_idx) != hipSuccess) {
        set_error(kErrHip, "hipGetDeviceProperties(%u) failed", device_idx);
        return;
    }
    printf("HIP driver  version: %d\n", drv);
    printf("HIP runtime version: %d\n", rt);
    printf("HIP device[%u] info for %s (%s):\n", device_idx, pr.name, pr.gcnArchName);
    printf("No. of compute units: %d\n", pr.multiProcessorCount);
    printf("____[ Memory ]____\n");
    printf("Total global memory: %.3f GB\n", (double)pr.totalGlobalMem / (1024.0 * 1024.0 * 1024.0));
    printf("Shared memory (LDS) per block: %zu bytes\n", (size_t)pr.sharedMemPerBlock);
    printf("L2 cache: %d bytes\n", pr.l2CacheSize);
    printf("____[ Grid ]____\n");
    printf("Maximum grid dim:   (%d, %d, %d)\n", pr.maxGridSize[0], pr.maxGridSize[1], pr.maxGridSize[2]);
    printf("Max. no. of threads per block: %d\n", pr.maxThreadsPerBlock);
    printf("Wavefront size: %d\n", pr.warpSize);
    fflush(stdout);
}

int get_texture_limit(unsigned int img_dim) {
    if (img_dim == 0) return 0;
    // references whose spectra ([rings <= img_dim/2][128] float2) keep 32-bit offsets
    const long long per = (long long)std::max(1u, img_dim / 2) * kSlots;
    return (int)std::min<long long>((long long)std::numeric_limits<int>::max() / per, 1 << 20);
}

void gpu_clear(void) {
    if (g.h_param) (void)hipHostFree(g.h_param);
    g.h_param = nullptr;
    g.n_param = 0;
    if (g.h_stage) (void)hipHostFree(g.h_stage);
    g.h_stage = nullptr;
    g.h_stage_cap = 0;
    DevBuf *bufs[] = {&g.sbj_img, &g.ref_img, &g.prm, &g.tr, &g.cid_idx, &g.ref_of, &g.twd, &g.tanl_work};
    for (DevBuf *b : bufs) b->release();
    g.sbj_cap = g.ref_cap = g.sbj_count = g.ref_count = 0;
    g.twd_dim = 0;
    g.shifts.clear();
    g.cfg = nullptr;
    g.mode = Context::kNone;
    ws.release();
    // the device id stays fixed (cuda/gpu_aln_noref.cu:139-180 leaves CUDA_DEVICE_ID set)
}

void reset_shifts(float shift_range, float shift_step) {
    if (g.shifts.empty()) {
        set_error(kErrState, "reset_shifts called before *_init");
        return;
    }
    std::vector<float2> ns = make_shift_grid(shift_range, shift_range, shift_step);
    if (ns.size() != g.shifts.size()) {
        set_error(kErrArg, "reset_shifts(%g, %g) gives %zu shifts, the context holds %zu",
                  (double)shift_range, (double)shift_step, ns.size(), g.shifts.size());
        return;
    }
    g.shifts = ns;
}

// ----------------------------------------------------------- multi-reference
AlignParam *multi_ref_align_init(const AlignConfig *aln_cfg, unsigned int cuda_device_id) {
    if (alignment_init(aln_cfg, cuda_device_id)) return nullptr;
    if (alloc_params(aln_cfg->sbj_num, -1)) return nullptr;
    g.mode = Context::kMref;
    return g.h_param;
}

void multi_ref_align(const float **sbj_data, const float **ref_data, float *isac_peak_list,
                     float *isac_d, unsigned int isac_d_pitch, float cuda_mem_request) {
    if (!require_mode(Context::kMref, "multi_ref_align")) return;
    const AlignConfig *c = g.cfg;
    const int P = (int)c->sbj_num, N = (int)c->ref_num, D = (int)c->img_dim, R = (int)c->ring_num;
    if (check_shape(D, R, (int)c->ring_len, g.shifts.size())) return;
    if (P == 0 || N == 0) return;
    if (!isac_peak_list || !isac_d || isac_d_pitch < (unsigned)P) {
        set_error(kErrArg, "bad output buffers (d_pitch %u < sbj_num %d?)", isac_d_pitch, P);
        return;
    }
    hipStream_t st = ws.stream;
    if (check_image_list(sbj_data, (unsigned)P)) return;
    if (upload_images(g.ref_img, ref_data, (unsigned)N, D, st)) return;
    if (upload_params(0, (size_t)P, st)) return;
    // particle images: each batch is gathered into pinned memory and copied on a
    // second stream just before its launches are queued, and the search stream
    // waits for that copy only; so the host gathers batch b + 1, and the copy
    // engine moves it, while the device searches batch b. Batches of at most a
    // quarter of the particles keep that overlap when memory would allow one
    // batch: only the first quarter's gather and copy are exposed.
    const size_t img_bytes = (size_t)D * D * sizeof(float);
    if (reserve_stage((size_t)P * img_bytes)) return;
    if (g.sbj_img.reserve((size_t)P * img_bytes) != hipSuccess ||
        (!ws.copy_stream && hipStreamCreateWithFlags(&ws.copy_stream, hipStreamNonBlocking) != hipSuccess)) {
        set_error(kErrHip, "particle image buffer or copy stream: allocation failed");
        return;
    }
    std::vector<hipEvent_t> copied;
    const std::function<int(int, int)> stage_batch = [&](int p0, int n) -> int {
        gather_images(sbj_data, (unsigned)p0, (unsigned)n, img_bytes);
        hipEvent_t ev = nullptr;
        HIP_OK(hipEventCreateWithFlags(&ev, hipEventDisableTiming));
        copied.push_back(ev);
        HIP_OK(hipMemcpyAsync(g.sbj_img.as<char>() + (size_t)p0 * img_bytes,
                              reinterpret_cast<char *>(g.h_stage) + (size_t)p0 * img_bytes,
                              (size_t)n * img_bytes, hipMemcpyHostToDevice, ws.copy_stream));
        HIP_OK(hipEventRecord(ev, ws.copy_stream));
        HIP_OK(hipStreamWaitEvent(st, ev, 0));
        return 0;
    };
    struct EventsGone {   // after the search (or a failure) the copies are waited for
        std::vector<hipEvent_t> &e;
        hipStream_t cs;
        ~EventsGone() {
            (void)hipStreamSynchronize(cs);
            for (hipEvent_t x : e) (void)hipEventDestroy(x);
        }
    } events_gone{copied, ws.copy_stream};
    if (ws.peak.reserve((size_t)N * 4 * P * sizeof(float)) != hipSuccess ||
        ws.dvals.reserve((size_t)N * P * sizeof(float)) != hipSuccess) {
        set_error(kErrHip, "output buffer allocation failed");
        return;
    }
    ShiftView acc{&g.prm.as<DevParam>()[0].shift_x, (int)(sizeof(DevParam) / sizeof(float))};
    if (run_mref(g.sbj_img.as<float>(), g.ref_img.as<float>(), acc, P, N, D, R, g.shifts,
                 ws.peak.as<float>(), 4 * P, ws.dvals.as<float>(), P, cuda_mem_request, st,
                 stage_batch, (P + 3) / 4))
        return;
    if (hipMemcpy(isac_peak_list, ws.peak.p, (size_t)N * 4 * P * sizeof(float),
                  hipMemcpyDeviceToHost) != hipSuccess ||
        hipMemcpy2D(isac_d, (size_t)isac_d_pitch * sizeof(float), ws.dvals.p,
                    (size_t)P * sizeof(float), (size_t)P * sizeof(float), (size_t)N,
                    hipMemcpyDeviceToHost) != hipSuccess) {
        set_error(kErrHip, "copying results to the host failed");
        return;
    }
}

int isac_mref_align_device(const float *d_sbj, const float *d_ref, const float *d_acc_xy,
                           unsigned int n_sbj, unsigned int n_ref, unsigned int img_dim,
                           unsigned int ring_num, unsigned int ring_len, float shift_step,
                           float shift_rng_x, float shift_rng_y, float *d_peak_list, float *d_d,
                           unsigned int d_pitch, void *stream) {
    if (int e = ensure_device_ws(-1)) return e;
    std::vector<float2> sh = make_shift_grid(shift_rng_x, shift_rng_y, shift_step);
    if (int e = check_shape((int)img_dim, (int)ring_num, (int)ring_len, sh.size())) return e;
    if (n_sbj == 0 || n_ref == 0) return 0;
    if (!d_sbj || !d_ref || !d_peak_list || !d_d || d_pitch < n_sbj) {
        set_error(kErrArg, "isac_mref_align_device: bad pointers or d_pitch < n_sbj");
        return kErrArg;
    }
    hipStream_t st = (hipStream_t)stream;   // NULL: the null stream, ordered with the caller's work
    ShiftView acc{d_acc_xy, 2};
    return run_mref(d_sbj, d_ref, acc, (int)n_sbj, (int)n_ref, (int)img_dim, (int)ring_num, sh,
                    d_peak_list, 4 * (int)n_sbj, d_d, (int)d_pitch, 0.9f, st);
}

int isac_mref_align_device_noalloc(const float *d_sbj, const float *d_ref, const float *d_acc_xy,
                                   unsigned int n_sbj, unsigned int n_ref, unsigned int img_dim,
                                   unsigned int ring_num, unsigned int ring_len, float shift_step,
                                   float shift_rng_x, float shift_rng_y, float *d_peak_list,
                                   float *d_d, unsigned int d_pitch, void *stream) {
    std::vector<float2> sh = make_shift_grid(shift_rng_x, shift_rng_y, shift_step);
    if (int e = check_shape((int)img_dim, (int)ring_num, (int)ring_len, sh.size())) return e;
    if (n_sbj == 0 || n_ref == 0) return 0;
    if (!d_sbj || !d_ref || !d_peak_list || !d_d || d_pitch < n_sbj) {
        set_error(kErrArg, "isac_mref_align_device_noalloc: bad pointers or d_pitch < n_sbj");
        return kErrArg;
    }
    ShiftView acc{d_acc_xy, 2};
    return run_mref_noalloc(d_sbj, d_ref, acc, (int)n_sbj, (int)n_ref, (int)img_dim,
                            (int)ring_num, sh, d_peak_list, 4 * (int)n_sbj, d_d, (int)d_pitch,
                            (hipStream_t)stream);
}

int isac_pre_align_loop_device(const float *d_imgs, unsigned int n_img, unsigned int img_dim,
                               unsigned int ring_num, const float *shift_ranges,
                               const float *shift_steps, unsigned int n_steps,
                               unsigned int max_iter, int center, float cutoff, float falloff,
                               void *d_aln_param, float *d_average, void *stream) {
    if (int e = ensure_device_ws(-1)) return e;
    if (n_img == 0 || n_steps == 0 || max_iter == 0) return 0;
    if (!d_imgs || !shift_ranges || !shift_steps || !d_aln_param) {
        set_error(kErrArg, "isac_pre_align_loop_device: null pointer");
        return kErrArg;
    }
    if (center != 0 && center != -1) {
        set_error(kErrUnsupported, "isac_pre_align_loop_device: center %d (0 or -1)", center);
        return kErrUnsupported;
    }
    if (cutoff > 0.0f && !(falloff > 0.0f)) {
        set_error(kErrArg, "isac_pre_align_loop_device: falloff must be > 0 with a cutoff");
        return kErrArg;
    }
    return run_pre_align_loop(d_imgs, (int)n_img, (int)img_dim, (int)ring_num, shift_ranges,
                              shift_steps, (int)n_steps, (int)max_iter, center, cutoff, falloff,
                              static_cast<DevParam *>(d_aln_param), d_average,
                              (hipStream_t)stream);
}

int isac_multi_align_stability_device(const double *d_params, unsigned int n_total,
                                      const int *d_class_off, unsigned int n_class,
                                      unsigned int n_runs, float mir_stab_thld, float grp_err_thld,
                                      float err_thld, float diameter, float *d_pix, float *d_avg,
                                      int *d_stable, float *d_class, float *d_global, void *stream) {
    if (n_class == 0) return 0;
    if (n_runs < 2 || n_runs > (unsigned)isac::kStabMaxRuns) {
        set_error(kErrArg, "isac_multi_align_stability_device: %u runs (need 2..%d)", n_runs,
                  isac::kStabMaxRuns);
        return kErrArg;
    }
    if (!d_params || !d_class_off || !d_pix || !d_avg || !d_stable || !d_class || !d_global) {
        set_error(kErrArg, "isac_multi_align_stability_device: null pointer");
        return kErrArg;
    }
    hipStream_t st = (hipStream_t)stream;
    HIP_OK(isac::launch_stability(d_params, (int)n_total, d_class_off, (int)n_class, (int)n_runs,
                                  mir_stab_thld, grp_err_thld, err_thld, (double)diameter, d_pix,
                                  d_avg, d_stable, d_class, d_global, st));
    HIP_OK(hipStreamSynchronize(st));
    return 0;
}

int isac_mref_align_mixed_device(const float *d_sbj, const float *d_ref, const float *d_acc_xy,
                                 unsigned int n_sbj, unsigned int n_ref, unsigned int img_dim,
                                 unsigned int ring_num, unsigned int ring_len, float shift_step,
                                 float shift_rng_x, float shift_rng_y, int bf16, float *d_peak_list,
                                 float *d_d, unsigned int d_pitch, void *stream) {
    if (int e = ensure_device_ws(-1)) return e;
    std::vector<float2> sh = make_shift_grid(shift_rng_x, shift_rng_y, shift_step);
    if (int e = check_shape((int)img_dim, (int)ring_num, (int)ring_len, sh.size())) return e;
    if (n_sbj == 0 || n_ref == 0) return 0;
    if (!d_sbj || !d_ref || !d_peak_list || !d_d || d_pitch < n_sbj) {
        set_error(kErrArg, "isac_mref_align_mixed_device: bad pointers or d_pitch < n_sbj");
        return kErrArg;
    }
    hipStream_t st = (hipStream_t)stream;   // NULL: the null stream, ordered with the caller's work
    ShiftView acc{d_acc_xy, 2};
    return run_mref_mixed(d_sbj, d_ref, acc, (int)n_sbj, (int)n_ref, (int)img_dim, (int)ring_num,
                          sh, bf16 != 0, d_peak_list, 4 * (int)n_sbj, d_d, (int)d_pitch, 0.9f, st);
}

int isac_mref_best_mixed_device(const float *d_sbj, const float *d_ref, const float *d_acc_xy,
                                unsigned int n_sbj, unsigned int n_ref, unsigned int img_dim,
                                unsigned int ring_num, unsigned int ring_len, float shift_step,
                                float shift_rng_x, float shift_rng_y, int bf16, float tie_rel,
                                float *d_peak_list, float *d_d, int *d_best_ref, float *d_best_val,
                                float *d_best_peak, int *d_tie, void *stream) {
    if (int e = ensure_device_ws(-1)) return e;
    std::vector<float2> sh = make_shift_grid(shift_rng_x, shift_rng_y, shift_step);
    if (int e = check_shape((int)img_dim, (int)ring_num, (int)ring_len, sh.size())) return e;
    if (n_sbj == 0 || n_ref == 0) return 0;
    if (!d_sbj || !d_ref || !d_peak_list || !d_d || !d_best_ref || !d_best_val || !d_best_peak ||
        !d_tie) {
        set_error(kErrArg, "isac_mref_best_mixed_device: NULL pointer");
        return kErrArg;
    }
    hipStream_t st = (hipStream_t)stream;   // NULL: the null stream, ordered with the caller's work
    ShiftView acc{d_acc_xy, 2};
    BestOut bo;
    bo.ref = d_best_ref;
    bo.val = d_best_val;
    bo.peak = d_best_peak;
    bo.tie = d_tie;
    bo.tie_rel = tie_rel;
    return run_mref_mixed(d_sbj, d_ref, acc, (int)n_sbj, (int)n_ref, (int)img_dim, (int)ring_num,
                          sh, bf16 != 0, d_peak_list, 4 * (int)n_sbj, d_d, (int)n_sbj, 0.9f, st,
                          &bo);
}

int isac_mref_proposal_device(const float *d_sbj, const float *d_ref, const float *d_acc_xy,
                              unsigned int n_sbj, unsigned int n_ref, unsigned int img_dim,
                              unsigned int ring_num, unsigned int ring_len, float shift_step,
                              float shift_rng_x, float shift_rng_y, const int *d_prop,
                              unsigned int n_prop, unsigned int n_psi_bins, float psi_half_deg,
                              unsigned int shift_half, float threshold, int edge_fallback,
                              int level2, const float *d_conf, float conf_min,
                              int fallback_form, int *d_best_ref, float *d_best_val,
                              float *d_best_peak, int *d_fallback, unsigned int *n_level2,
                              unsigned int *n_fallback, void *stream) {
    if (int e = ensure_device_ws(-1)) return e;
    std::vector<float2> sh = make_shift_grid(shift_rng_x, shift_rng_y, shift_step);
    if (int e = check_shape((int)img_dim, (int)ring_num, (int)ring_len, sh.size())) return e;
    if (n_fallback) *n_fallback = 0;
    if (n_level2) *n_level2 = 0;
    if (n_sbj == 0 || n_ref == 0) return 0;
    if (!d_sbj || !d_ref || !d_prop || !d_best_ref || !d_best_val || !d_best_peak || !d_fallback ||
        n_prop < 1 || n_prop > (unsigned)isac::kMaxProposals || n_psi_bins < 1 || n_psi_bins > 256 ||
        !(psi_half_deg > 0.0f && psi_half_deg <= 180.0f) || shift_half > 2 ||
        n_sbj > (1u << 30) || n_ref > (1u << 20) || fallback_form < 0 || fallback_form > 2) {
        set_error(kErrArg, "isac_mref_proposal_device: bad arguments (1..8 proposals, psi bins "
                           "1..256, 0 < psi_half <= 180, shift_half <= 2, fallback form 0..2)");
        return kErrArg;
    }
    int ny = 0;
    while (ny < (int)sh.size() && sh[ny].x == sh[0].x) ++ny;
    const int nx = ny ? (int)sh.size() / ny : 0;
    hipStream_t st = (hipStream_t)stream;   // NULL: the null stream, ordered with the caller's work
    return run_proposal(d_sbj, d_ref, ShiftView{d_acc_xy, 2}, (int)n_sbj, (int)n_ref, (int)img_dim,
                        (int)ring_num, sh, nx, ny, d_prop, (int)n_prop, (int)n_psi_bins, psi_half_deg,
                        (int)shift_half, threshold, edge_fallback != 0, level2 != 0, d_conf,
                        conf_min, d_best_ref, d_best_val, d_best_peak, d_fallback, n_fallback,
                        n_level2, st, fallback_form);
}

int isac_mref_align_batched_device(const IsacMrefSubset *subsets, unsigned int n_subsets,
                                   int mode, void *stream) {
    if (int e = ensure_device_ws(-1)) return e;
    if (n_subsets == 0) return 0;
    if (!subsets) {
        set_error(kErrArg, "isac_mref_align_batched_device: NULL descriptor array");
        return kErrArg;
    }
    hipStream_t st = (hipStream_t)stream;   // NULL: the null stream, ordered with the caller's work
    if (mode == ISAC_BATCH_GRAPH) {
        HIP_OK(hipStreamSynchronize(st));   // inputs written on the caller's stream
        return run_batched_graph(subsets, n_subsets, st);
    }
    if (mode != ISAC_BATCH_PER_SUBSET) {
        set_error(kErrArg, "isac_mref_align_batched_device: unknown mode %d", mode);
        return kErrArg;
    }
    for (unsigned i = 0; i < n_subsets; ++i) {
        const IsacMrefSubset &x = subsets[i];
        const int e = isac_mref_align_device(x.d_sbj, x.d_ref, x.d_acc_xy, x.n_sbj, x.n_ref,
                                             x.img_dim, x.ring_num, kRingLen, x.shift_step,
                                             x.shift_rng_x, x.shift_rng_y, x.d_peak_list, x.d_d,
                                             x.d_pitch, stream);
        if (e) return e;
    }
    return 0;
}

void isac_batched_report(int *n_groups, int *n_nodes, int *replayed, int *nw, int *blocks,
                         float *cand_ms3) {
    const BatchWs &b = ws.batch;
    if (n_groups) *n_groups = b.groups;
    if (n_nodes) *n_nodes = b.nodes;
    if (replayed) *replayed = b.replayed;
    if (nw) *nw = b.nw;
    if (blocks) *blocks = b.blocks;
    if (cand_ms3) std::copy(b.cand_ms, b.cand_ms + 3, cand_ms3);
}

void isac_profile_enable(int on) {
    g.prof = on != 0;
    g.prof_ccf_n = g.prof_polar_n = 0;
    g.prof_ccf_ms = g.prof_polar_ms = 0.0;
}

void isac_profile_read(int *launches, double *total_ms, int *polar_launches, double *polar_total_ms) {
    if (launches) *launches = g.prof_ccf_n;
    if (total_ms) *total_ms = g.prof_ccf_ms;
    if (polar_launches) *polar_launches = g.prof_polar_n;
    if (polar_total_ms) *polar_total_ms = g.prof_polar_ms;
}

// -------------------------------------------------------------- pre-alignment
AlignParam *pre_align_init(unsigned int num_particles, const AlignConfig *aln_cfg,
                           unsigned int cuda_device_id) {
    if (alignment_init(aln_cfg, cuda_device_id)) return nullptr;
    if (alloc_params(num_particles, 0)) return nullptr;
    g.sbj_cap = aln_cfg->sbj_num;
    g.ref_cap = aln_cfg->ref_num;
    const size_t ib = (size_t)aln_cfg->img_dim * aln_cfg->img_dim * sizeof(float);
    if (g.sbj_img.reserve(std::max<size_t>(1, g.sbj_cap) * ib) != hipSuccess ||
        g.ref_img.reserve(std::max<size_t>(1, g.ref_cap) * ib) != hipSuccess) {
        set_error(kErrHip, "pre_align_init: image buffer allocation failed");
        return nullptr;
    }
    g.mode = Context::kPre;
    return g.h_param;
}

int pre_align_size_check(unsigned int num_particles, const AlignConfig *cfg,
                         unsigned int cuda_device_id, float request, bool verbose) {
    if (!cfg) return 0;
    if (hipSetDevice((int)cuda_device_id) != hipSuccess) return 0;
    const double mb = 1024.0 * 1024.0;
    const size_t ib = (size_t)cfg->img_dim * cfg->img_dim * sizeof(float);
    const size_t params = (size_t)num_particles * sizeof(AlignParam) * 2;
    const size_t sbj = (size_t)cfg->sbj_num * ib, ref = (size_t)cfg->ref_num * ib;
    // the spectra workspace is batched internally; require room for 64 particles
    const size_t work = 64 * spec_bytes_per_particle(cfg) +
                        (size_t)std::max(1u, cfg->ref_num) * cfg->ring_num * kSlots * sizeof(float2);
    const size_t total = params + sbj + ref + work;
    const size_t avail = free_device_bytes();
    if (verbose) {
        printf("GPU[%u] SIZE CHECK\n  in/out param: %.1f MB\n  sbj images: %.1f MB\n  ref images: %.1f MB\n"
               "  min. workspace: %.1f MB\nTOTAL: %.1f MB of %.1f MB x %.2f\n",
               cuda_device_id, params / mb, sbj / mb, ref / mb, work / mb, total / mb, avail / mb,
               (double)request);
    }
    return ((double)avail * (double)request >= (double)total) ? 1 : 0;
}

void pre_align_fetch(const float **img_data, unsigned int img_num, const char *batch_type) {
    if (!require_mode(Context::kPre, "pre_align_fetch")) return;
    const int D = (int)g.cfg->img_dim;
    if (!batch_type) {
        set_error(kErrArg, "pre_align_fetch: NULL batch type");
        return;
    }
    if (strcmp(batch_type, "sbj_batch") == 0) {
        if (img_num > g.sbj_cap) {
            set_error(kErrArg, "pre_align_fetch: %u images exceed the batch capacity %u", img_num, g.sbj_cap);
            return;
        }
        if (upload_images(g.sbj_img, img_data, img_num, D, ws.stream)) return;
        g.sbj_count = img_num;
    } else if (strcmp(batch_type, "ref_batch") == 0) {
        if (img_num > std::max(1u, g.ref_cap)) {
            set_error(kErrArg, "pre_align_fetch: %u references exceed the capacity %u", img_num, g.ref_cap);
            return;
        }
        if (upload_images(g.ref_img, img_data, img_num, D, ws.stream)) return;
        g.ref_count = img_num;
    } else {
        set_error(kErrArg, "fetch_data() :: Unknown batch type '%s' specified.", batch_type);
    }
}

void pre_align_run(int start_idx, int stop_idx) {
    if (!require_mode(Context::kPre, "pre_align_run")) return;
    const AlignConfig *c = g.cfg;
    if (start_idx < 0 || stop_idx < start_idx || (size_t)start_idx > g.n_param) {
        set_error(kErrArg, "pre_align_run(%d, %d): bad range for %zu particles", start_idx, stop_idx, g.n_param);
        return;
    }
    // rows decoded: the fetched batch, not beyond stop_idx (cuda/gpu_aln_noref.cu:2493-2496)
    size_t cnt = std::min<size_t>(g.sbj_count, (size_t)(stop_idx - start_idx));
    cnt = std::min(cnt, g.n_param - (size_t)start_idx);
    if (cnt == 0) return;
    hipStream_t st = ws.stream;
    // accumulated shifts of particles start..start+cnt; the reference index comes
    // from records 0..cnt-1 (ccf_mult is called with data index 0, :762)
    if (g.ref_of.reserve(cnt * sizeof(int)) != hipSuccess) {
        set_error(kErrHip, "pre_align_run: allocation failed");
        return;
    }
    std::vector<int> refs(cnt);
    for (size_t i = 0; i < cnt; ++i) {
        refs[i] = g.h_param[i].ref_id;
        if (refs[i] < 0 || refs[i] >= (int)std::max(1u, g.ref_count)) {
            set_error(kErrArg, "pre_align_run: particle %zu has ref_id %d", i, refs[i]);
            return;
        }
    }
    if (hipMemcpy(g.ref_of.p, refs.data(), cnt * sizeof(int), hipMemcpyHostToDevice) != hipSuccess) {
        set_error(kErrHip, "pre_align_run: copy failed");
        return;
    }
    if (upload_params((size_t)start_idx, cnt, st)) return;
    const float lim = (float)(c->img_dim - c->ring_num - 2);
    if (run_single(g.sbj_img.as<float>(), g.prm.as<DevParam>(), g.ref_of.as<int>(), 1, (int)cnt,
                   g.ref_img.as<float>(), (int)g.ref_count, (int)c->img_dim, (int)c->ring_num,
                   g.shifts, lim, st))
        return;
    (void)download_params((size_t)start_idx, cnt, st);
}

// ------------------------------------------------------ reference-free (2D)
AlignParam *ref_free_alignment_2D_init(const AlignConfig *aln_cfg, const float **sbj_data_list,
                                       const float **ref_data_list, const int *sbj_cid_list,
                                       unsigned int cuda_device_id) {
    if (alignment_init(aln_cfg, cuda_device_id)) return nullptr;
    const unsigned P = aln_cfg->sbj_num, N = aln_cfg->ref_num;
    if (!sbj_cid_list && P) {
        set_error(kErrArg, "NULL class id list");
        return nullptr;
    }
    // class index runs (cuda/gpu_aln_noref.cu:825-835)
    std::vector<int> cid_idx(N + 1, (int)P);
    int cid = -1;
    unsigned idx = 0;
    for (unsigned i = 0; i < P; ++i) {
        const int ci = sbj_cid_list[i];
        if (ci < 0 || ci >= (int)N) {
            set_error(kErrArg, "particle %u has class id %d outside [0, %u)", i, ci, N);
            return nullptr;
        }
        if (ci != cid) {
            if (idx >= N) {
                set_error(kErrArg, "class ids form more than %u contiguous runs", N);
                return nullptr;
            }
            cid_idx[idx++] = (int)i;
            cid = ci;
        }
    }
    if (alloc_params(P, 0)) return nullptr;
    for (unsigned i = 0; i < P; ++i) g.h_param[i].ref_id = sbj_cid_list[i];
    hipStream_t st = ws.stream;
    if (g.cid_idx.reserve((N + 1) * sizeof(int)) != hipSuccess ||
        hipMemcpy(g.cid_idx.p, cid_idx.data(), (N + 1) * sizeof(int), hipMemcpyHostToDevice) != hipSuccess) {
        set_error(kErrHip, "ref_free_alignment_2D_init: allocation failed");
        return nullptr;
    }
    if (upload_images(g.sbj_img, sbj_data_list, P, (int)aln_cfg->img_dim, st)) return nullptr;
    if (upload_images(g.ref_img, ref_data_list, N, (int)aln_cfg->img_dim, st)) return nullptr;
    g.sbj_count = g.sbj_cap = P;
    g.ref_count = g.ref_cap = N;
    g.mode = Context::kRefFree;
    return g.h_param;
}

int ref_free_alignment_2D_size_check(const AlignConfig *cfg, unsigned int cuda_device_id,
                                     float request, bool verbose) {
    if (!cfg) return 0;
    if (hipSetDevice((int)cuda_device_id) != hipSuccess) return 0;
    const double mb = 1024.0 * 1024.0;
    const size_t ib = (size_t)cfg->img_dim * cfg->img_dim * sizeof(float);
    const size_t params = (size_t)cfg->sbj_num * sizeof(AlignParam) * 2 + (cfg->ref_num + 1) * sizeof(int);
    const size_t imgs = (size_t)cfg->sbj_num * ib * 2 + (size_t)cfg->ref_num * ib;  // + transformed copy
    const size_t work = 64 * spec_bytes_per_particle(cfg) +
                        (size_t)cfg->ref_num * cfg->ring_num * kSlots * sizeof(float2) +
                        (size_t)cfg->ref_num * cfg->img_dim * (cfg->img_dim / 2 + 1) * sizeof(float2);
    const size_t total = params + imgs + work;
    const size_t avail = free_device_bytes();
    if (verbose) {
        printf("GPU[%u] SIZE CHECK\n  in/out param: %.1f MB\n  images: %.1f MB\n  min. workspace: %.1f MB\n"
               "TOTAL: %.1f MB of %.1f MB x %.2f\n",
               cuda_device_id, params / mb, imgs / mb, work / mb, total / mb, avail / mb, (double)request);
    }
    return ((double)avail * (double)request >= (double)total) ? 1 : 0;
}

void ref_free_alignment_2D(void) {
    if (!require_mode(Context::kRefFree, "ref_free_alignment_2D")) return;
    const AlignConfig *c = g.cfg;
    const int P = (int)g.sbj_count, N = (int)g.ref_count, D = (int)c->img_dim, R = (int)c->ring_num;
    if (check_shape(D, R, (int)c->ring_len, g.shifts.size())) return;
    hipStream_t st = ws.stream;
    for (int i = 0; i < P; ++i)
        if (g.h_param[i].ref_id < 0 || g.h_param[i].ref_id >= N) {
            set_error(kErrArg, "ref_free_alignment_2D: particle %d has ref_id %d outside [0, %d)", i,
                      g.h_param[i].ref_id, N);
            return;
        }
    if (upload_params(0, (size_t)P, st)) return;
    DevParam *prm = g.prm.as<DevParam>();
    const float lim = (float)(c->img_dim - c->ring_num - 2);
    if (run_single(g.sbj_img.as<float>(), prm, &prm[0].ref_id, (int)(sizeof(DevParam) / sizeof(int)),
                   P, g.ref_img.as<float>(), N, D, R, g.shifts, lim, st))
        return;
    if (g.tr.reserve((size_t)std::max(P, 1) * D * D * sizeof(float)) != hipSuccess) {
        set_error(kErrHip, "ref_free_alignment_2D: allocation failed");
        return;
    }
    hipError_t e = isac::launch_transform(g.sbj_img.as<float>(), P, D, prm, g.tr.as<float>(), st);
    if (e == hipSuccess)
        e = isac::launch_class_average(g.tr.as<float>(), D, g.cid_idx.as<int>(), N,
                                       g.ref_img.as<float>(), st);
    if (e != hipSuccess) {
        set_error(kErrHip, "ref_free_alignment_2D: kernel launch failed: %s", hipGetErrorString(e));
        return;
    }
    (void)download_params(0, (size_t)P, st);
}

void ref_free_alignment_2D_filter_references(float cutoff_freq, float falloff) {
    if (!require_mode(Context::kRefFree, "ref_free_alignment_2D_filter_references")) return;
    const int D = (int)g.cfg->img_dim, N = (int)g.ref_count;
    if (N <= 0) return;
    hipStream_t st = ws.stream;
    if (g.twd_dim != D) {
        std::vector<float2> tw(D);
        for (int m = 0; m < D; ++m) {
            const double a = -2.0 * M_PI * (double)m / (double)D;
            tw[m] = make_float2((float)cos(a), (float)sin(a));
        }
        if (g.twd.reserve(D * sizeof(float2)) != hipSuccess ||
            hipMemcpy(g.twd.p, tw.data(), D * sizeof(float2), hipMemcpyHostToDevice) != hipSuccess) {
            set_error(kErrHip, "filter: twiddle upload failed");
            return;
        }
        g.twd_dim = D;
    }
    if (g.tanl_work.reserve((size_t)N * D * (D / 2 + 1) * sizeof(float2)) != hipSuccess) {
        set_error(kErrHip, "filter: allocation failed");
        return;
    }
    hipError_t e = isac::launch_tanl_filter(g.ref_img.as<float>(), N, D, cutoff_freq, falloff,
                                            g.tanl_work.as<float2>(), g.twd.as<float2>(), st);
    if (e == hipSuccess) e = hipStreamSynchronize(st);
    if (e != hipSuccess) set_error(kErrHip, "filter failed: %s", hipGetErrorString(e));
}

}  // extern "C"
