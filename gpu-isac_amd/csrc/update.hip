// Reference update of the multi-reference iteration on the device: the step the
// reference's driver runs on the host after assign_groups
// (bin/sp_isac2_gpu.py:1146-1232), with the image-domain kernels of the
// reference-free path (transform cuda/gpu_aln_noref.cu:1539-1590, class
// average :1592-1622, tanh filter :1007-1037).
//
//   for every particle i of class k (cls[i]):
//     (alpha, sx, sy, m) = peak_list[k][4i .. 4i+3]; (sx', sy') = the driver's
//     rotation of (sx, sy) (bin/sp_isac2_gpu.py:1016-1022)
//     sum_x[k] += m ? -sx' : sx'; sum_y[k] += sy'; members[k] += 1
//     ref[k] += image i transformed with (alpha, sx, sy, m)
//   avg[k] = sum[k] / members[k]
//   params[i] = (alpha, m ? sx' + avg_x : sx' - avg_x, sy' - avg_y, m)
//   ref[k] = fshift(filt_tanl(ref[k] / members[k], fl, ff), -avg_x, -avg_y)
//
// Differences from the driver, which runs EMAN2 (absent here, so these stay
// unpinned against it): the transform is the reference GPU's own bilinear
// image transform, not EMAN2's rot_shift2D; the tanh filter uses absolute
// frequency sqrt(kx^2 + ky^2) / D; the centring shift is a Fourier phase ramp
// applied in the filter's frequency pass.  A class without members gives a zero
// image and zero average shift (the driver would divide by zero).
#include "isac_internal.h"
#include "../../include/gpu_aln_pack.h"

#include <cmath>
#include <vector>

namespace isac {
namespace {

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

int update_references(const float *imgs, int P, int D, const float *peak, const int *cls, int N,
                      float cutoff, float falloff, float *refs, float *params, int *members_out,
                      float *shift_avg_out, hipStream_t st) {
    if (P < 1 || N < 1 || D < 2 || D > kMaxImgDim || !imgs || !peak || !cls || !refs || !params)
        return err(2, "isac_update_references_device: need device buffers, n_img >= 1, n_ref >= 1 "
                      "and 2 <= img_dim <= 128");
    if (!(cutoff > 0.0f) || !(falloff > 0.0f))
        return err(2, "isac_update_references_device: cutoff and falloff must be positive");
    const int Dh = D / 2 + 1;
    Buf tp, mir, post, raw, cnt, off, mem, avg, neg, work, tw, bad;
    hipError_t e = hipSuccess;
    auto alloc = [&](Buf &b, size_t bytes) {
        if (e == hipSuccess) e = hipMalloc(&b.p, bytes);
    };
    alloc(tp, (size_t)P * sizeof(float4));
    alloc(mir, (size_t)P * sizeof(int));
    alloc(post, (size_t)P * sizeof(float2));
    alloc(raw, (size_t)P * sizeof(float4));
    alloc(cnt, (size_t)N * sizeof(int));
    alloc(off, ((size_t)N + 1) * sizeof(int));
    alloc(mem, (size_t)P * sizeof(int));
    alloc(avg, (size_t)N * sizeof(double2));
    alloc(neg, (size_t)N * sizeof(float2));
    alloc(work, (size_t)N * D * Dh * sizeof(float2));
    alloc(tw, (size_t)D * sizeof(float2));
    alloc(bad, sizeof(int));
    if (e != hipSuccess) return err(1, "isac_update_references_device: device allocation failed");
    std::vector<float2> twh(D);
    for (int m = 0; m < D; ++m) {
        const double a = -2.0 * M_PI * (double)m / (double)D;
        twh[m] = make_float2((float)cos(a), (float)sin(a));
    }
    if (hipMemcpy(tw.p, twh.data(), D * sizeof(float2), hipMemcpyHostToDevice) != hipSuccess ||
        hipMemsetAsync(bad.p, 0, sizeof(int), st) != hipSuccess)
        return err(1, "isac_update_references_device: upload failed");
    UpdateScratch w{tp.as<float4>(), mir.as<int>(), post.as<float2>(), raw.as<float4>(),
                    cnt.as<int>(), off.as<int>(), mem.as<int>(), avg.as<double2>(),
                    neg.as<float2>(), work.as<float2>(), tw.as<float2>(), bad.as<int>(),
                    reinterpret_cast<float2 *>(shift_avg_out), members_out};
    e = launch_update_references(imgs, P, D, peak, cls, N, cutoff, falloff, w, refs,
                                 reinterpret_cast<float4 *>(params), st);
    if (e != hipSuccess) return err(1, "isac_update_references_device: launch failed");
    int flag = 0;
    if (hipMemcpyAsync(&flag, bad.p, sizeof(int), hipMemcpyDeviceToHost, st) != hipSuccess ||
        hipStreamSynchronize(st) != hipSuccess)
        return err(1, "isac_update_references_device: execution failed");
    if (flag) return err(2, "isac_update_references_device: a class id is outside [0, n_ref)");
    return 0;
}

}  // namespace isac

extern "C" {

int isac_update_references_device(const float *d_imgs, unsigned int n_img, unsigned int img_dim,
                                  const float *d_peak_list, const int *d_class, unsigned int n_ref,
                                  float cutoff, float falloff, float *d_refs, float *d_params,
                                  int *d_members, float *d_shift_avg, void *stream) {
    return isac::update_references(d_imgs, (int)n_img, (int)img_dim, d_peak_list, d_class,
                                   (int)n_ref, cutoff, falloff, d_refs, d_params, d_members,
                                   d_shift_avg, (hipStream_t)stream);
}

}  // extern "C"
