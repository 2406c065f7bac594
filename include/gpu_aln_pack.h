/*
 * gpu_aln_pack.h -- C ABI of the MI355X (gfx950) polar-ring 2D alignment library.
 *
 * Drop-in for the reference's gpu_aln_pack.so: the two structs have the
 * reference's exact memory layout and the first fourteen functions keep the
 * reference's names, argument meaning, lifetime rules and signatures, so the
 * reference's ctypes drivers bind them unchanged.  The functions after them
 * have no counterpart in the reference and say so.
 *
 * Citations are to the ucdavis/GPU-ISAC tree.
 *
 * Errors: the reference prints and exit()s on any device error
 * (cuda/gpu_aln_common.cu:104-118) and assert()s on bad arguments.  This
 * library prints a message to stderr, records an error code readable through
 * gpu_last_error(), and returns; with ISAC_EXIT_ON_ERROR=1 in the
 * environment it exits(1) like the reference.
 *
 * Threading: one process-global context, not reentrant, one device per
 * process (cuda/gpu_aln_noref.cu:31-48, :103).  Every call is synchronous:
 * results are complete on return.
 */
#ifndef ISAC_AMD_GPU_ALN_PACK_H
#define ISAC_AMD_GPU_ALN_PACK_H

#include <stdbool.h>
#include <stddef.h>

#ifdef __cplusplus
extern "C" {
#endif

/* Search configuration; layout of cuda/gpu_aln_common.h:62-76 (40 bytes on LP64).
 * Retained BY POINTER from *_init until gpu_clear (cuda/gpu_aln_noref.cu:108):
 * the caller keeps it, and the aln_type string, alive. */
typedef struct AlignConfig {
    unsigned int sbj_num;      /* number of subject (particle) images        */
    unsigned int ref_num;      /* number of references / classes             */
    unsigned int img_dim;      /* images are img_dim x img_dim float32       */
    unsigned int ring_num;     /* polar rings at radius 1..ring_num          */
    unsigned int ring_len;     /* samples per ring; must be 256 (all callers) */
    float        shift_step;   /* translational search step                  */
    float        shift_rng_x;  /* search -rng..+rng in x                     */
    float        shift_rng_y;  /* search -rng..+rng in y                     */
    const char  *aln_type;     /* "pre_aln", "multigroup_aln" or "mref_aln"  */
} AlignConfig;

/* Per-particle parameters; layout of cuda/gpu_aln_common.h:78-85 (24 bytes). */
typedef struct AlignParam {
    int   sbj_id;
    int   ref_id;
    float shift_x;
    float shift_y;
    float angle;
    bool  mirror;
} AlignParam;

/* ------------------------------------------------ reference entry points -- */

/* Prints the properties of device `device_idx` (cuda/gpu_aln_common.cu:180-209). */
void print_gpu_info(unsigned int device_idx);

/* Number of img_dim x img_dim references one reference batch may hold
 * (cuda/gpu_aln_common.cu:233-237).  This library has no texture-height cap;
 * it returns the count whose polar spectra stay within 32-bit element
 * offsets.  Callers pass a second, ignored argument
 * (bin/sp_isac_applications.py:153-154); the C ABI tolerates it. */
int get_texture_limit(unsigned int img_dim);

/* Frees every resource of the context; the device id stays fixed
 * (cuda/gpu_aln_noref.cu:139-180). */
void gpu_clear(void);

/* Rebuilds the shift grid with x and y range `shift_range`; the number of
 * shifts must not change (cuda/gpu_aln_noref.cu:117-137). */
void reset_shifts(float shift_range, float shift_step);

/* Multi-reference alignment set-up (cuda/gpu_aln_noref.cu:311-324).  Returns
 * aln_cfg->sbj_num host-readable/-writable AlignParam records owned by the
 * library until gpu_clear; callers write the accumulated shift_x/shift_y
 * before multi_ref_align (bin/sp_isac2_gpu.py:979-983). */
AlignParam *multi_ref_align_init(const AlignConfig *aln_cfg, unsigned int cuda_device_id);

/* Aligns every particle with every reference (cuda/gpu_aln_noref.cu:326-554).
 * sbj_data: sbj_num host pointers to C-contiguous img_dim^2 float32 images;
 * ref_data: ref_num such pointers.  Writes, for particle p and reference n,
 *   isac_d[n * isac_d_pitch + p]            = maximum CCF value,
 *   isac_peak_list[n * 4*sbj_num + 4p + 0..3] = {angle, shift_x, shift_y, mirror}
 * where shift = accumulated shift + the winning grid shift.  isac_d_pitch is
 * the caller's global particle count; only columns 0..sbj_num-1 of each row
 * are written.  cuda_mem_request is the fraction of free device memory the
 * call may use (0.9 in every caller). */
void multi_ref_align(const float **sbj_data, const float **ref_data,
                     float *isac_peak_list, float *isac_d,
                     unsigned int isac_d_pitch, float cuda_mem_request);

/* Pre-alignment set-up for num_particles particles against one reference
 * (cuda/gpu_aln_noref.cu:565-609).  aln_cfg->sbj_num is the batch capacity. */
AlignParam *pre_align_init(unsigned int num_particles, const AlignConfig *aln_cfg,
                           unsigned int cuda_device_id);

/* 1 if a pre-alignment batch of cfg->sbj_num images fits in request * free
 * device memory, else 0 (cuda/gpu_aln_noref.cu:611-721).  Declared int so
 * the value is fully zero-extended for callers that read it as a C int
 * (bin/sp_isac_applications.py:1458 uses ctypes' default c_int restype). */
int pre_align_size_check(unsigned int num_particles, const AlignConfig *cfg,
                         unsigned int cuda_device_id, float request, bool verbose);

/* Copies img_num host images into the "sbj_batch" or "ref_batch" buffer
 * (cuda/gpu_aln_noref.cu:723-740). */
void pre_align_fetch(const float **img_data, unsigned int img_num, const char *batch_type);

/* Aligns the fetched subject batch (global particle indices start_idx ..
 * stop_idx-1) to the fetched reference and updates their AlignParam records
 * in place (cuda/gpu_aln_noref.cu:742-769, :2463-2520). */
void pre_align_run(int start_idx, int stop_idx);

/* Reference-free (multigroup) set-up: copies all particles and the initial
 * class references; sbj_cid_list gives each particle's class and must list
 * each class as one contiguous run (cuda/gpu_aln_noref.cu:782-846). */
AlignParam *ref_free_alignment_2D_init(const AlignConfig *aln_cfg, const float **sbj_data_list,
                                       const float **ref_data_list, const int *sbj_cid_list,
                                       unsigned int cuda_device_id);

/* 1 if a reference-free run of cfg fits in request * free memory, else 0
 * (cuda/gpu_aln_noref.cu:848-962).  int return: see pre_align_size_check. */
int ref_free_alignment_2D_size_check(const AlignConfig *cfg, unsigned int cuda_device_id,
                                     float request, bool verbose);

/* One reference-free iteration: align each particle to its class reference,
 * update its AlignParam, re-average the transformed particles into new
 * class references (cuda/gpu_aln_noref.cu:964-996). */
void ref_free_alignment_2D(void);

/* Tangent low-pass filter of the current class references
 * (cuda/gpu_aln_noref.cu:998-1003, :2203-2266). */
void ref_free_alignment_2D_filter_references(float cutoff_freq, float falloff);

/* ------------------------------------------------------------------ extras */

/* No counterpart in the reference.  Error code of the first failure since
 * the last gpu_clear_error() (0 = none) and its message. */
int gpu_last_error(void);
const char *gpu_last_error_string(void);
void gpu_clear_error(void);

/* No counterpart in the reference (device-resident form of multi_ref_align,
 * for callers whose images already live in device memory, e.g. the
 * benchmark and the Python host layer).  All pointers are device pointers:
 *   d_sbj [n_sbj][img_dim][img_dim], d_ref [n_ref][img_dim][img_dim] float32,
 *   d_acc_xy [n_sbj][2] accumulated shifts (NULL = zero),
 *   d_peak_list [n_ref][4*n_sbj], d_d [n_ref][d_pitch].
 * `stream` is a hipStream_t; the work is queued on it after whatever the
 * caller queued there before (NULL = the null stream), and the call returns
 * after it has completed.  Returns 0 or an error code. */
int isac_mref_align_device(const float *d_sbj, const float *d_ref, const float *d_acc_xy,
                           unsigned int n_sbj, unsigned int n_ref, unsigned int img_dim,
                           unsigned int ring_num, unsigned int ring_len,
                           float shift_step, float shift_rng_x, float shift_rng_y,
                           float *d_peak_list, float *d_d, unsigned int d_pitch,
                           void *stream);

/* No counterpart in the reference: for callers that cannot allocate device
 * memory (code under HIP graph capture, tight serving loops, harnesses that
 * forbid hipMalloc).  isac_mref_align_device without any device allocation or
 * free: same arguments, same kernels, bit-identical results.  It works in one
 * static device array of ISAC_NOALLOC_SCRATCH_BYTES per GPU (held from library
 * load until the process ends), which must hold the tables, the reference
 * spectra (n_ref * ring_num * 1 KiB) and the spectra of at least one particle
 * (shifts * ring_num * 1 KiB); larger jobs return error 4.  Particles are
 * processed in batches that fit.  Calls that overlap on one GPU (other threads,
 * other streams) share the array and race: serialise them.  Queued on `stream`;
 * returns after completion. */
#define ISAC_NOALLOC_SCRATCH_BYTES 67108864 /* 64 MiB */
int isac_mref_align_device_noalloc(const float *d_sbj, const float *d_ref, const float *d_acc_xy,
                                   unsigned int n_sbj, unsigned int n_ref, unsigned int img_dim,
                                   unsigned int ring_num, unsigned int ring_len,
                                   float shift_step, float shift_rng_x, float shift_rng_y,
                                   float *d_peak_list, float *d_d, unsigned int d_pitch,
                                   void *stream);

/* No counterpart in the reference (mixed-precision form of
 * isac_mref_align_device): the ring contraction runs on FP16 (bf16 = 0) or BF16
 * (bf16 = 1) MFMA with FP32 accumulation and the psi transform on FP16 MFMA.
 * Writes approximate d / peak_list for every pair (same layout as
 * isac_mref_align_device); FP32 re-scoring of candidates is separate. */
int isac_mref_align_mixed_device(const float *d_sbj, const float *d_ref, const float *d_acc_xy,
                                 unsigned int n_sbj, unsigned int n_ref, unsigned int img_dim,
                                 unsigned int ring_num, unsigned int ring_len,
                                 float shift_step, float shift_rng_x, float shift_rng_y, int bf16,
                                 float *d_peak_list, float *d_d, unsigned int d_pitch,
                                 void *stream);

/* No counterpart in the reference.  Mixed-precision search plus exact choice of
 * each particle's best reference: the eight references with the highest
 * approximate peaks are re-scored in FP32 (same arithmetic as
 * isac_mref_align_device) and, when the best two are within tie_rel
 * (relative), re-scored in FP64.  Outputs (device): approximate d [n_ref][n_sbj]
 * and peak_list [n_ref][4 n_sbj] for every pair; best reference, its peak value,
 * {angle, shift_x, shift_y, mirror} and the near-tie flag per particle. */
int isac_mref_best_mixed_device(const float *d_sbj, const float *d_ref, const float *d_acc_xy,
                                unsigned int n_sbj, unsigned int n_ref, unsigned int img_dim,
                                unsigned int ring_num, unsigned int ring_len, float shift_step,
                                float shift_rng_x, float shift_rng_y, int bf16, float tie_rel,
                                float *d_peak_list, float *d_d, int *d_best_ref, float *d_best_val,
                                float *d_best_peak, int *d_tie, void *stream);

/* No counterpart in the reference (its search is always exhaustive,
 * cuda/gpu_aln_noref.cu:326-554).  Narrowed multi-reference search around
 * proposals, with exact fallbacks; all pointers are device pointers except
 * n_level2 / n_fallback (host, may be NULL).  d_prop [n_sbj][n_prop] holds int4
 * {reference, mirror, psi bin, shift index} (n_prop <= 8; psi bin b covers
 * [b, b+1) * 360 / n_psi_bins degrees of the CCF's psi; shift index into the
 * x-major grid of isac_mref_align_device, at most 64 shifts; entries out of
 * range are ignored).  Three levels, each with the arithmetic of
 * isac_mref_align_device for the pairs it covers:
 *  1. each proposal is searched over the shifts within shift_half grid steps
 *     and the psi samples within psi_half_deg of its bin centre (a window
 *     that holds the exhaustive optimum returns it bit for bit);
 *  2. particles whose best value is below `threshold`, or (edge_fallback != 0)
 *     whose maximum lies on the edge of the admitted region (a psi or shift
 *     neighbour that no window of its reference admits), get with level2 != 0
 *     every shift and psi of their distinct proposed references (*n_level2
 *     particles);
 *  3. particles still below the threshold, or flagged when level2 == 0, get
 *     the exhaustive search (fallback = 1, *n_fallback particles); so do,
 *     directly, the particles whose d_conf (or NULL: none) is below conf_min
 *     (fallback = 2), e.g. the network's own estimate that its proposals hold
 *     the true reference.  fallback_form selects that search: 0 = the FP32
 *     search of isac_mref_align_device; 1 (FP16) or 2 (BF16) = the MFMA search
 *     with each particle's 8 best references re-scored in FP32 and near-ties
 *     in FP64 (isac_mref_best_mixed_device), which equals form 0 whenever the
 *     FP32 optimum is among the 8.
 * Outputs per particle: best reference, its peak value, {angle, shift_x,
 * shift_y, mirror} in d_best_peak [n_sbj][4], the fallback flag.  Returns after
 * the work has completed. */
int isac_mref_proposal_device(const float *d_sbj, const float *d_ref, const float *d_acc_xy,
                              unsigned int n_sbj, unsigned int n_ref, unsigned int img_dim,
                              unsigned int ring_num, unsigned int ring_len, float shift_step,
                              float shift_rng_x, float shift_rng_y, const int *d_prop,
                              unsigned int n_prop, unsigned int n_psi_bins, float psi_half_deg,
                              unsigned int shift_half, float threshold, int edge_fallback,
                              int level2, const float *d_conf, float conf_min,
                              int fallback_form, int *d_best_ref, float *d_best_val,
                              float *d_best_peak, int *d_fallback, unsigned int *n_level2,
                              unsigned int *n_fallback, void *stream);

/* No counterpart in the reference (the reference aligns one subset per
 * multi_ref_align call, cuda/gpu_aln_noref.cu:326-554, driven once per ISAC
 * subset by bin/sp_isac2_gpu.py:979-1003).  One subset of a batched
 * multi-reference run; every pointer is a device pointer and the outputs have
 * the layout of isac_mref_align_device (ring_len is 256). */
typedef struct IsacMrefSubset {
    const float *d_sbj;        /* [n_sbj][img_dim][img_dim]                  */
    const float *d_ref;        /* [n_ref][img_dim][img_dim]                  */
    const float *d_acc_xy;     /* [n_sbj][2] accumulated shifts, NULL = zero  */
    float *d_peak_list;        /* [n_ref][4*n_sbj]                           */
    float *d_d;                /* [n_ref][d_pitch]                           */
    unsigned int n_sbj, n_ref, img_dim, ring_num;
    float shift_step, shift_rng_x, shift_rng_y;
    unsigned int d_pitch;
} IsacMrefSubset;

#define ISAC_BATCH_PER_SUBSET 0  /* one launch sequence (and synchronisation) per subset */
#define ISAC_BATCH_GRAPH 1       /* every subset in one captured HIP graph around one
                                    persistent work-queue kernel per memory group */

/* No counterpart in the reference.  Aligns n_subsets independent subsets of
 * uneven particle count, box size, reference count, ring count and shift range;
 * results are bit-identical to one isac_mref_align_device call per subset.  In
 * ISAC_BATCH_GRAPH mode the tile shape (waves per workgroup) of each memory
 * group is chosen by timing the candidates on a sample of the group, and the
 * captured graph is kept and replayed while the descriptor array is unchanged
 * (same pointers and shapes; the data behind them may change).  Returns after
 * the work has completed. */
int isac_mref_align_batched_device(const IsacMrefSubset *subsets, unsigned int n_subsets,
                                   int mode, void *stream);

/* No counterpart in the reference.  What the last ISAC_BATCH_GRAPH call did:
 * memory groups, graph nodes, whether the graph was replayed from the cache,
 * and for the first group the chosen waves per workgroup, the persistent grid
 * size and the timed sample (ms) for 1, 2 and 4 waves. */
void isac_batched_report(int *n_groups, int *n_nodes, int *replayed, int *nw, int *blocks,
                         float *cand_ms3);

/* Equal-size group assignment from the scores d [n_ref][d_pitch] (columns
 * 0..n_img-1): the step that consumes multi_ref_align's d
 * (bin/sp_isac2_gpu.py:1119-1144), there EMAN2's Util.assign_groups on the
 * host.  Pairs (class, particle) in descending score order (ties: lower class,
 * then lower particle) place a particle that has no class yet in a class that
 * holds fewer than cap = n_img / n_ref particles.  Output (n_img + 1 ints):
 * class k's particles (ascending) at [k*cap, (k+1)*cap), then the particles no
 * class took (ascending), then one class id for them (the class of their
 * highest pair; 0 when there are none), the layout the driver reads.
 * NaN scores count as -infinity; cap <= 2048.  Device form: device pointers,
 * queued on `stream`, returns after completion; host form: host arrays. */
int isac_assign_groups_device(const float *d_d, unsigned int n_ref, unsigned int n_img,
                              unsigned int d_pitch, int *d_out, void *stream);
int isac_assign_groups(const float *h_d, unsigned int n_ref, unsigned int n_img, int *h_out);

/* Reference update after the assignment: the step the driver runs on the host
 * with EMAN2 (bin/sp_isac2_gpu.py:1146-1232), on the device.  Inputs: d_imgs
 * [n_img][img_dim][img_dim], d_peak_list [n_ref][4*n_img] as multi_ref_align /
 * isac_mref_align_device write it (before the driver's shift rotation), d_class
 * [n_img] the class of each particle (0..n_ref-1).  For class k: every member is
 * transformed with its peak of class k (the reference-free path's image
 * transform, cuda/gpu_aln_noref.cu:1539-1590) and averaged in ascending particle
 * order; the average shift of the members (their rotated shifts,
 * bin/sp_isac2_gpu.py:1016-1022, mirrored members with -sx) is
 * (avg_x, avg_y); d_refs[k] = the average low-passed with the tanh filter
 * (cutoff, falloff in absolute frequency, cycles per pixel) and moved by
 * (-avg_x, -avg_y) with a Fourier phase ramp.  d_params [n_img][4] = (angle,
 * sx -/+ avg_x (+ for mirrored), sy - avg_y, mirror) in the driver's
 * convention (bin/sp_isac2_gpu.py:1200-1208).  d_members [n_ref] and
 * d_shift_avg [n_ref][2] (either may be NULL) receive the member counts and
 * (avg_x, avg_y).  A class without members gets a zero image.  img_dim <= 128.
 * Queued on `stream`; returns after completion (2 if a class id is outside
 * [0, n_ref)). */
int isac_update_references_device(const float *d_imgs, unsigned int n_img, unsigned int img_dim,
                                  const float *d_peak_list, const int *d_class, unsigned int n_ref,
                                  float cutoff, float falloff, float *d_refs, float *d_params,
                                  int *d_members, float *d_shift_avg, void *stream);

/* Pre-alignment driver loop on the device: bin/sp_isac_applications.py:1486-1693
 * (the N_step x max_iter iterations of ali2d_base_gpu_isac_CLEAN, whose device
 * step is pre_align_fetch / pre_align_run, cuda/gpu_aln_noref.cu:565-769) with no
 * host round trip between iterations.  d_imgs [n_img][img_dim][img_dim] (device,
 * mean-subtracted by the caller as the driver does); shift_ranges / shift_steps
 * (host, n_steps each) replace the driver's xrng / step lists (reset_shifts per
 * step); d_aln_param: n_img AlignParam records in device memory, the starting
 * parameters on entry (the driver starts from zero) and the final ones on
 * return, in the convention of pre_align_init's records.  Per iteration: the
 * average of every particle transformed with its current parameters (the
 * reference-free path's image transform in place of EMAN2's sum_oe); with
 * center = -1 the average is moved by minus the mean of the particles' shifts
 * after the driver's rotation (its "average shift" centring, :1547-1552); with
 * cutoff > 0 it is low-passed with the tanh filter (absolute frequency; a
 * stand-in for SPHIRE's ref_ali2d user function, which the reference does not
 * ship); then every particle is aligned to the average and its shift
 * accumulated (pre_align_run).  The last iteration builds the average only
 * (:1563), which goes to d_average (device, may be NULL).  Queued on `stream`;
 * returns after completion. */
int isac_pre_align_loop_device(const float *d_imgs, unsigned int n_img, unsigned int img_dim,
                               unsigned int ring_num, const float *shift_ranges,
                               const float *shift_steps, unsigned int n_steps,
                               unsigned int max_iter, int center, float cutoff, float falloff,
                               void *d_aln_param, float *d_average, void *stream);

/* Multi-alignment stability test of ISAC classes on the device: the step the
 * driver runs per class on the host with SPHIRE's
 * sp_pixel_error.multi_align_stability(aln_param[k], 0.0, 10000.0, thld_err,
 * False, last_ring*2) (bin/sp_isac2_gpu.py:1481-1491; the module is not in the
 * reference tree).  d_params [n_runs][n_total][4] FP64 {angle deg, sx, sy,
 * mirror} of every member in every repeated run (the driver's aln_param[k],
 * :1381-1393, classes concatenated); class c holds members
 * [d_class_off[c], d_class_off[c+1]) (a range outside [0, n_total] or
 * decreasing is taken as an empty class).  Per class: the mirror-flip pattern held
 * by most members selects the mirror-consistent members; one global rigid
 * transform per run (the last run fixed) is fitted to minimise the mean squared
 * pixel error over a circle of the given diameter, leaving out members above
 * 3 err_thld at the start; then per member: d_pix = pixel error (-1 when not
 * mirror consistent or when the class is rejected: consistent rate <=
 * mir_stab_thld or fewer than 2 members to fit), d_avg [4] = average
 * {angle, sx, sy, mirror} in the last run's frame, d_stable = 1 when
 * d_pix <= err_thld and the class error <= grp_err_thld; per class: d_class
 * [2] = {mirror-consistent rate, class error sqrt(mean pix^2) or -1},
 * d_global [n_runs][4] = fitted {alpha deg, tx, ty, flip}.  2 <= n_runs <= 12.
 * Queued on `stream`; returns after completion. */
int isac_multi_align_stability_device(const double *d_params, unsigned int n_total,
                                      const int *d_class_off, unsigned int n_class,
                                      unsigned int n_runs, float mir_stab_thld, float grp_err_thld,
                                      float err_thld, float diameter, float *d_pix, float *d_avg,
                                      int *d_stable, float *d_class, float *d_global, void *stream);

/* No counterpart in the reference.  When enabled, the library brackets every
 * launch of its contraction+psi kernel with HIP events on the stream it
 * launches on, and accumulates launch count and total milliseconds. */
void isac_profile_enable(int on);
void isac_profile_read(int *launches, double *total_ms, int *polar_launches,
                       double *polar_total_ms);

#ifdef __cplusplus
}
#endif
#endif
