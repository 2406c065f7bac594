// Private interface between the HIP kernels (kernels.hip) and the C-ABI
// host layer (gpu_aln_pack.hip).  Not installed; not part of the ABI.
#pragma once

#include <hip/hip_runtime.h>
#include <cstddef>
#include <cstdint>

namespace isac {

constexpr int kRingLen = 256;   // samples per ring (fixed by every caller)
constexpr int kSlots = 128;     // packed spectrum slots per ring: slot 0 = (X[0], X[128])

// Position of spectrum slot k in a ring row of kSlots float2: slots k and k + 64
// are adjacent, so a lane that owns slots l and l + 64 reads both with one
// 16-byte access (float4 l of the row).
__host__ __device__ constexpr int slot_pos(int k) { return 2 * (k & 63) + (k >> 6); }
constexpr int kMaxImgDim = 128; // image staged whole in LDS by the polar kernel
constexpr int kMaxRings = 127;  // rings must fit inside a 256-sample ring FFT budget

// Same memory layout as the reference's AlignParam (cuda/gpu_aln_common.h:78-85).
struct DevParam {
    int sbj_id;
    int ref_id;
    float shift_x;
    float shift_y;
    float angle;
    bool mirror;
};
static_assert(sizeof(DevParam) == 24, "AlignParam layout");

// Best psi-CCF peak of one (particle, reference) pair over (mirror, shift, psi):
// raw maximum, its mirror / shift index / psi sample, and the seven samples
// around the peak in the winning (mirror, shift) slot for the parabola.
struct PeakRec {
    float val;
    int m;
    int s;
    int t;
    float x[7];
    int pad;
};
static_assert(sizeof(PeakRec) == 48, "PeakRec layout");

// Narrowed-search work item: a reference, the centre of a shift window and the
// (mirror, psi sample) mask, bit m * 256 + t.  ref < 0: empty.
constexpr int kMaxProposals = 8;
struct PropItem {
    int ref;
    int s;
    unsigned mask[16];
};
static_assert(sizeof(PropItem) == 72, "PropItem layout");

// Strided view of per-image accumulated shifts: x at base[i*stride], y at +1.
// float2 arrays use stride 2; DevParam arrays use base = &p[0].shift_x, stride 6.
struct ShiftView {
    const float *base;
    int stride;
};

// One multi-reference job (a particle batch of one subset) for the FP32 kernel.
// The candidate form (cand set) scores K listed references per particle instead
// of all N: list i holds cand[i*K .. i*K+K-1] (negative: an empty slot), takes
// its spectra from row pidx[i] (or i when pidx is null), and fills rec[i*K + k].
struct MrefJob {
    const float2 *spec;     // [n_img][S][R][kSlots] raw particle spectra
    const float2 *norm;     // [n_img][S]
    const float2 *rspec;    // [N][R][kSlots] normalised, ring-weighted reference spectra
    const float *rsum0;     // [N]
    PeakRec *rec;           // [n_img][N], or [lists][K] in the candidate form
    int n_img, S, R, N, G, GB;   // G reference groups of 4, GB workgroups per particle
    const int *cand = nullptr;   // candidate form: [lists][K] reference ids
    const int *pidx = nullptr;   // candidate form: spectra row of each list, or null
    int K = 0;
};

MrefJob mref_job(const float2 *spec, const float2 *norm, const float2 *rspec, const float *rsum0,
                 PeakRec *rec, int n_img, int n_shift, int n_ring, int n_ref, int nw);

// Persistent multi-reference kernel over many jobs: items (job, particle, block)
// in 8 per-XCD queues [qbeg[q], qbeg[q+1]); counters (8 ints) are reset by a
// kernel first (no memset node, safe inside captured graphs).
// nw (waves per workgroup) is 1, 2 or 4; blocks should be a multiple of 8.
hipError_t launch_ccf_mref_queue(const MrefJob *jobs, const int4 *items, const int *qbeg,
                                 int *counters, int nw, int blocks, const float2 *twiddle,
                                 hipStream_t st);
int mref_queue_blocks_per_cu(int nw);   // occupancy of the queue kernel (0 on error)

// Polar spectra of the shifts marked in smask[img] only (n_shift <= 64).
hipError_t launch_polar_spectra_masked(const float *imgs, int n_img, int img_dim, ShiftView acc,
                                       const int2 *shift_i, int n_shift, const float2 *tmpl,
                                       int n_ring, const float2 *twiddle, float2 *spec,
                                       float2 *norm, const unsigned long long *smask,
                                       hipStream_t st);

// Narrowed search: prop [n_img][K] int4 (ref, mirror, psi bin, shift) -> items
// [n_img][K] and shift masks; shift grid nx x ny (x-major), window +-hw shifts.
hipError_t launch_prop_items(const int *prop, int n_img, int K, int N, int nx, int ny, int hw,
                             int n_bins, float psi_half, PropItem *items,
                             unsigned long long *smask, hipStream_t st);
hipError_t launch_ccf_prop(const float2 *spec, const float2 *norm, const float2 *rspec,
                           const float *rsum0, const PropItem *items, int K, int n_img, int n_shift,
                           int n_ring, int n_ref, int nx, int ny, int hw, const float2 *twiddle,
                           PeakRec *rec, hipStream_t st);
hipError_t launch_best_prop(const PeakRec *rec, const PropItem *items, int n_img, int K, int nx,
                            int ny, int hw, int edge_test, ShiftView acc, const float2 *shifts_f,
                            int n_shift, int p_off, int *best_ref, float *best_val,
                            float *best_peak, int *edge, hipStream_t st);
// Second level, inside a particle batch: exact search of all shifts and psi over
// the distinct proposed references (cand [n][K]) of batch particles idx[0..n);
// the missing spectra rows come from a masked polar pass (k_level2_mask); the
// best reference is decoded into the outputs at p_off + idx[j] with fallback =
// value below threshold.
hipError_t launch_prop_refs(const int *prop, const int *idx, int n, int K, int N, int *cand,
                            hipStream_t st);
hipError_t launch_level2_mask(const int *flag, int n, int S, const unsigned long long *done,
                              unsigned long long *smask, hipStream_t st);
hipError_t launch_best_refs_level2(const PeakRec *rec, const int *cand, int n, int K, ShiftView acc,
                                   const float2 *shifts_f, int n_shift, const int *idx, int p_off,
                                   float threshold, int *best_ref, float *best_val,
                                   float *best_peak, int *fallback, hipStream_t st);
hipError_t launch_fallback_flags(const float *best_val, const int *edge, const float *conf,
                                 float conf_min, int n_img, float threshold, int p_off,
                                 int *fallback, hipStream_t st);
// Exhaustive fallback: compact flagged particles, gather them, scatter the best
// reference of each back to its index.
hipError_t launch_compact(const int *flag, int n, int lo, int hi, int *idx, int *count,
                          hipStream_t st);
hipError_t launch_gather(const float *imgs, int D, ShiftView acc, const int *idx, int n, float *out,
                         float2 *out_acc, hipStream_t st);
hipError_t launch_best_refs_scatter(const PeakRec *rec, int n, int N, const float2 *acc,
                                    const float2 *shifts_f, int n_shift, const int *idx,
                                    int *best_ref, float *best_val, float *best_peak,
                                    hipStream_t st);

// Records an error for gpu_last_error() (codes: 1 HIP, 2 argument, 3 state, 4 unsupported).
void report_error(int code, const char *msg);

// One-time launch attributes (dynamic LDS limit); call before any stream capture.
hipError_t prepare_launches();

// Stage A: normalised-later ring spectra of every image at every grid shift.
// spec[img][s][ring][kSlots] float2 (raw, unnormalised), norm[img][s] = (mu*L, 1/(sd+1e-7)).
hipError_t launch_polar_spectra(const float *imgs, int n_img, int img_dim, ShiftView acc,
                                const int2 *shift_i, int n_shift, const float2 *tmpl,
                                int n_ring, const float2 *twiddle, float2 *spec, float2 *norm,
                                hipStream_t st);

// Normalise and ring-weight reference spectra in place (n_shift == 1 layout);
// rsum0[n] = sum_i w_i R'_i[0] (needed for the particle DC correction).
hipError_t launch_ref_finalize(float2 *rspec, const float2 *rnorm, int n_ref, int n_ring,
                               float *rsum0, hipStream_t st);

// Stage B, multi-reference: every particle x every reference; rec[p*n_ref + n].
hipError_t launch_ccf_mref(const float2 *spec, const float2 *norm, const float2 *rspec,
                           const float *rsum0, int n_img, int n_shift, int n_ring, int n_ref,
                           const float2 *twiddle, PeakRec *rec, hipStream_t st);

// Stage B, one reference per particle (pre-alignment / reference-free):
// ref index of image i at ref_of[i*ref_stride]; rec[i].
hipError_t launch_ccf_single(const float2 *spec, const float2 *norm, const float2 *rspec,
                             const float *rsum0, const int *ref_of, int ref_stride, int n_img,
                             int n_shift, int n_ring, int n_ref, const float2 *twiddle,
                             PeakRec *rec, hipStream_t st);

// Exact FP32 re-scoring of K candidate references per particle (ref cand[p*K+k];
// a negative entry gives an empty record), same arithmetic as launch_ccf_mref;
// rec[p*K + k].
// pidx (or NULL): spectra row of item i's particle is pidx[i / K] instead of i / K.
hipError_t launch_ccf_candidates(const float2 *spec, const float2 *norm, const float2 *rspec,
                                 const float *rsum0, const int *cand, int K, int n_img,
                                 int n_shift, int n_ring, int n_ref, const float2 *twiddle,
                                 PeakRec *rec, hipStream_t st, const int *pidx = nullptr);

// Top-K (K = 8) references per particle by approximate peak value.
hipError_t launch_topk(const PeakRec *rec, int n_img, int n_ref, int K, int *cand, hipStream_t st);

// Best / runner-up exact candidate per particle and near-tie flags.
hipError_t launch_best_candidate(const PeakRec *rec, const int *cand, int n_img, int K,
                                 float tie_rel, int *best, int *second, int *tie, hipStream_t st);

// FP64 re-scoring of flagged near-ties; may swap best and runner-up.
hipError_t launch_rescore_f64(const float2 *spec, const float2 *norm, const float2 *rspec,
                              const PeakRec *rec, const int *cand, int *best, int *second,
                              const int *tie, int n_img, int K, int n_shift, int n_ring,
                              int n_ref, double *val64, hipStream_t st);

// Per-particle best candidate -> outputs at particle p_off + p (idx != NULL: at
// particle idx[p_off + p]); tie_out may be NULL.
hipError_t launch_decode_best(const PeakRec *rec, const int *cand, const int *best,
                              const int *tie, int n_img, int K, ShiftView acc,
                              const float2 *shifts_f, int n_shift, int p_off, int *best_ref,
                              float *best_val, float *best_peak, int *tie_out, hipStream_t st,
                              const int *idx = nullptr);

// Multi-reference decode into the caller's d / peak_list layout.
hipError_t launch_decode_mref(const PeakRec *rec, int n_img, int n_ref, ShiftView acc,
                              const float2 *shifts_f, int n_shift, float *peak, int peak_row,
                              int p_off, float *d, int d_pitch, hipStream_t st);

// Pre-align / reference-free decode: update AlignParam records in place.
hipError_t launch_decode_single(const PeakRec *rec, int cnt, DevParam *prm,
                                const float2 *shifts_f, int n_shift, float shift_limit,
                                hipStream_t st);

// Apply found parameters to images: out[i] = T_i(imgs[i]).
hipError_t launch_transform(const float *imgs, int n, int img_dim, const DevParam *prm,
                            float *out, hipStream_t st);

// Pre-alignment loop: avg = mean over P particles of T_p(imgs[p]) (the transform of
// launch_transform), tp / mir hold P entries, part pre_average_scratch_floats(P, D);
// shift[0] = the centring move (-sx_sum / P, -sy_sum / P) of the driver's
// average-shift centring.
size_t pre_average_scratch_floats(int P, int D);
hipError_t launch_pre_average(const float *imgs, int P, int D, const DevParam *prm, float4 *tp,
                              int *mir, float *part, float *avg, hipStream_t st);
hipError_t launch_pre_center(const DevParam *prm, int P, float2 *shift, hipStream_t st);

// Per-class mean of transformed images into refs.
hipError_t launch_class_average(const float *tr, int img_dim, const int *cid_idx, int n_cls,
                                float *refs, hipStream_t st);

// Tangent low-pass filter of refs in place; work holds n*img_dim*(img_dim/2+1) float2,
// tw_d holds img_dim twiddles e^{-2 pi i m / img_dim}.
hipError_t launch_tanl_filter(float *refs, int n, int img_dim, float cutoff, float falloff,
                              float2 *work, const float2 *tw_d, hipStream_t st);

// Tanh low-pass with absolute frequency (cycles per pixel) and the phase ramp that
// moves image n's content by shift[n]; output scaled by 1 / (D*D) (a proper inverse).
// cutoff <= 0: no low-pass, the phase ramp alone.
hipError_t launch_filter_shift(float *refs, int n, int img_dim, float cutoff, float falloff,
                               const float2 *shift, float2 *work, const float2 *tw_d,
                               hipStream_t st);

// Device scratch of the reference update (P particles, N classes, D pixels).
struct UpdateScratch {
    float4 *tp;       // [P] transform parameters (cos, sin, sx, sy)
    int *mir;         // [P]
    float2 *post;     // [P] shifts after the driver's rotation
    float4 *raw;      // [P] (angle, sx, sy, mirror) as read from the peak list
    int *cnt;         // [N]
    int *off;         // [N + 1]
    int *members;     // [P] particles grouped by class, ascending within a class
    double2 *avg;     // [N] average shift per class
    float2 *neg;      // [N] centring shift per class (-avg)
    float2 *work;     // [N][D][D/2 + 1] spectra
    const float2 *tw; // [D] twiddles e^{-2 pi i m / D}
    int *bad;         // one flag: a class id outside [0, N)
    float2 *avg_out;  // [N] average shift per class as floats, or NULL
    int *members_out; // [N] member count per class, or NULL
};

// Reference update: class averages of the transformed particles, tanh filter and
// centring shift into refs [N][D][D]; per-particle parameters (angle, sx, sy,
// mirror) about the new centres into params [P].
hipError_t launch_update_references(const float *imgs, int P, int D, const float *peak,
                                   const int *cls, int N, float cutoff, float falloff,
                                   UpdateScratch w, float *refs, float4 *params, hipStream_t st);

// Mixed-precision (MFMA) multi-reference search.  Operands are FP16 / BF16 rows
// (particle shift rows, or references): per spectrum slot the two parts over
// K = rings padded to k_pad, grouped in the MFMA K-steps; blocks of 16 rows.
struct GemmLayout {
    int full16;          // 16-ring K-steps
    int tail;            // rings of the last step: 0, 8 or 16
    int k_pad;           // 16 full16 + tail
    int slot_bytes;      // 4 k_pad: one slot of one row
    size_t row_bytes;    // kSlots slot_bytes
    size_t block_bytes;  // 16 rows
    float cp;            // power-of-two scale of the particle operands
};
GemmLayout gemm_layout(int n_ring);
int gemm_index_bits(int n_shift);

// Particle operand rows: particle p, shift s at row p * S_pad + s with
// S_pad = gemm_rows_per_particle(S) = max(S, 16) (so a block of 16 rows holds
// rows of at most two particles); padding rows are zero.
int gemm_rows_per_particle(int n_shift);
size_t gemm_particle_operand_bytes(int n_img, int n_shift, int n_ring);
// Particle operands from the raw FP32 spectra [n_img][n_shift] + normalisation.
hipError_t launch_frag_rows(const float2 *spec, const float2 *norm, int n_img, int n_shift,
                            int n_ring, bool bf16, void *frag, hipStream_t st);
// Particle operands straight from the images (no FP32 spectra): the mixed
// search's stage A, for n_ring <= 40 (polar_frag_supported).
bool polar_frag_supported(int n_ring);
hipError_t launch_polar_frag(const float *imgs, int n_img, int img_dim, ShiftView acc,
                             const int2 *shift_i, int n_shift, int n_ring, bool bf16,
                             const float2 *twiddle, void *frag, hipStream_t st);
// Reference operands (normalised, ring-weighted spectra), padded to a multiple
// of 16 references; cn[n] = the power-of-two scale of reference n, rinv[n] =
// 1 / (c_p cn[n]) (the factor that undoes both scales).
hipError_t launch_frag_refs(const float2 *rspec, int n_ref, int n_ring, bool bf16, float *cn,
                            float *rinv, void *frag, hipStream_t st);
// Approximate peaks of every (particle, reference) pair, rec[p*n_ref + n].
hipError_t launch_mref_gemm(const void *pfrag, const void *rfrag, const float *rinv, int n_img,
                            int n_shift, int n_ring, int n_ref, bool bf16, PeakRec *rec,
                            hipStream_t st);

// Multi-alignment stability test, one workgroup per class (stability.hip).
// prm [L][n_total][4] FP64 {angle deg, sx, sy, mirror}; class c = members
// [class_off[c], class_off[c+1]).  Outputs per member: pix (-1 when not mirror
// consistent or the class is rejected), avg [4], stable (0/1); per class: cls [2]
// = {mirror-consistent rate, error (-1 when rejected)}, gout [L][4] = fitted
// {alpha deg, tx, ty, flip} per run.
constexpr int kStabMaxRuns = 12;
hipError_t launch_stability(const double *prm, int n_total, const int *class_off, int n_class,
                            int L, float mir_thld, float grp_thld, float err_thld, double d,
                            float *pix, float *avg, int *stable, float *cls, float *gout,
                            hipStream_t st);

}  // namespace isac
