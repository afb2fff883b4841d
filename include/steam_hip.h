/*
 * libsteam_hip — C ABI of the MI355X (gfx950) field-generation hot path of
 * STEAM: world-keyed Philox noise + CMS extremal-Levy variates, turbulon
 * kernel synthesis, the periodic-x/y zero-padded-z FFT overlap-add
 * convolution, the flux-cascade advance and the periodic trilinear regrid.
 *
 * Conventions
 *   - Every pointer argument named *_dev is a device pointer.  Host pointers
 *     are only the small descriptor structs and ledgers passed by address.
 *   - Every call is asynchronous on the given hipStream_t (NULL = default
 *     stream) unless its comment says it synchronises.
 *   - Every call returns a status (STEAM_OK = 0).  steam_last_error() gives a
 *     thread-local message for the last failing call on the calling thread.
 *     The Python layer maps the codes to the reference's exception types:
 *       STEAM_EINVAL   -> ValueError
 *       STEAM_ENOMEM   -> MemoryError
 *       STEAM_EDEVICE  -> RuntimeError
 *       STEAM_ENOTIMPL -> NotImplementedError
 *   - Arrays are C order (x slowest, z fastest), float32 unless stated.
 *
 * Reference interfaces replaced (paths relative to the reference repo):
 *   steam/noise.py, steam/turbulons.py, steam/utils.py, steam/simulate.py,
 *   steam/streaming.py — cited per entry point below.
 */
#ifndef STEAM_HIP_H
#define STEAM_HIP_H

#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

typedef struct ihipStream_t* steam_stream_t; /* == hipStream_t */

#define STEAM_OK 0
#define STEAM_EINVAL 1
#define STEAM_ENOMEM 2
#define STEAM_EDEVICE 3
#define STEAM_ENOTIMPL 4

/* Library version (major*10000 + minor*100 + patch). */
int steam_version(void);
/* Message of the last failed call on this thread ("" if none). */
const char* steam_last_error(void);

/* ------------------------------------------------------------------------
 * (a) keyed noise
 * ---------------------------------------------------------------------- */

/* Philox4x32-10 uniforms over an index box, both lanes in C order over
 * (ix, iy, iz).  Replaces keyed_uniforms (steam/noise.py:103-130) with the
 * counter/key layout of _uniform_pair (steam/noise.py:88-100).  ix/iy/iz are
 * int64 device arrays of world indices; lane0/lane1 hold nx*ny*nz floats. */
int steam_keyed_uniforms(const int64_t* ix_dev, int64_t nx,
                         const int64_t* iy_dev, int64_t ny,
                         const int64_t* iz_dev, int64_t nz, uint32_t key0,
                         uint32_t key1, uint32_t key2, float* lane0_dev,
                         float* lane1_dev, steam_stream_t stream);

/* Raw Philox4x32-10 blocks: counters (n x 4 uint32, device), key (k0, k1);
 * out receives n x 4 uint32.  The round function of _philox4x32_10
 * (steam/noise.py:67-85), exposed for known-answer checks. */
int steam_philox4x32_10(const uint32_t* ctr_dev, int64_t n, uint32_t key0,
                        uint32_t key1, uint32_t* out_dev, steam_stream_t stream);

/* CMS transform of two uniform lanes, in place on phi (R is consumed too).
 * Replaces _levy_from_uniforms (steam/simulate.py:441-479). */
int steam_levy_from_uniforms(double alpha, float* phi_dev, float* r_dev,
                             int64_t n, steam_stream_t stream);

/* Where an array sits in its class's world grid (steam/simulate.py:520-551)
 * plus the sub-lattice factors. */
typedef struct steam_region {
  uint32_t key[3];        /* class key (noise.class_key) */
  int64_t origin[3];      /* world index of local (0,0,0) */
  int64_t world_shape[3]; /* world grid; x/y wrap modulo it, z never wraps */
  int32_t sparsity[3];    /* (s_x, s_y, s_z) */
} steam_region;

/* Sparse keyed Levy field: draw at every centre (world index divisible by s,
 * x/y wrapped first), 0 elsewhere.  Replaces _keyed_sparse_levy
 * (steam/simulate.py:554-598); out is (nx, ny, nz). */
int steam_keyed_sparse_levy(int64_t nx, int64_t ny, int64_t nz,
                            const steam_region* region, double alpha,
                            float* out_dev, steam_stream_t stream);

/* Unit-mean multiplier noise exp(gamma)-1 (0 off-centre), the construction
 * shared by _advance_flux (steam/simulate.py:2146-2161) and
 * _multiplier_noise (steam/streaming.py:822-847).  per_class_scale and shift
 * are the float32 values the reference casts. */
typedef struct steam_noise_params {
  steam_region region;
  double alpha;
  float per_class_scale;
  float shift;
  int32_t n_zero;
  int32_t zero_bottom;
  int32_t zero_top;
} steam_noise_params;

int steam_multiplier_noise(int64_t nx, int64_t ny, int64_t nz,
                           const steam_noise_params* noise, float* out_dev,
                           steam_stream_t stream);

/* ------------------------------------------------------------------------
 * (b) turbulon kernel
 * ---------------------------------------------------------------------- */

#define STEAM_SHAPE_MEXICAN_HAT 0
#define STEAM_SHAPE_MORLET_OMEGA0_6 1
#define STEAM_SHAPE_POWER_LAW 2 /* steam_turbulon_envelope_power only */

/* Half-widths ceil(support_factor*k/d) of the kernel (host-side, no GPU). */
int steam_envelope_halfwidths(double k, double dx, double dy, double dz,
                              double support_factor, int64_t half[3]);

/* Centred turbulon kernel of shape (2hx+1, 2hy+1, 2hz+1) with the
 * Gaussian-weighted zero-sum correction; the two correction sums are taken
 * in float64.  Replaces _turbulon_envelope (steam/simulate.py:2503-2556) and
 * the shape registry (steam/turbulons.py:20-48).  Synchronises. */
int steam_turbulon_envelope(double k, double dx, double dy, double dz,
                            double support_factor, int shape_id,
                            float* out_dev, steam_stream_t stream);

/* The power-law turbulon T = (1 + rho)^(-p/2), rho = r^2 / (k/pi)^2: regular
 * at the centre, |r|^(-p) beyond the core (p = d/alpha for a generator
 * kernel, d - H for a fractional integration), with the same support and
 * Gaussian-weighted zero-sum correction as steam_turbulon_envelope.  No
 * counterpart in the reference: an addition to its shape registry
 * (steam/turbulons.py:20), whose entries are the Mexican hat and Morlet
 * shapes.  0 < exponent < 16.  Synchronises. */
int steam_turbulon_envelope_power(double k, double dx, double dy, double dz,
                                  double support_factor, double exponent,
                                  float* out_dev, steam_stream_t stream);

/* steam_turbulon_envelope without allocating: no counterpart in the
 * reference.  It is for callers that cannot allocate device memory (code
 * under HIP graph capture, tight serving loops, harnesses that forbid
 * hipMalloc).  The same kernels and grid, so the same taps bit for bit; the
 * float64 partial sums of the correction (at most 4096 workgroups) and its
 * ratio live in one library-owned device array of 65,792 bytes per GPU,
 * held until the process ends (the grid never needs more; a larger need
 * would be refused with STEAM_EINVAL).  Asynchronous: it does not
 * synchronise.  Calls on one stream run in order; calls that overlap on one
 * GPU (other streams or host threads) race on the shared array. */
int steam_turbulon_envelope_noalloc(double k, double dx, double dy, double dz,
                                    double support_factor, int shape_id,
                                    float* out_dev, steam_stream_t stream);

/* ------------------------------------------------------------------------
 * (c) FFT overlap-add convolution: periodic x/y, zero-padded z, centred
 *     'same' output (steam/utils.py:222-242, 293-427).
 * ---------------------------------------------------------------------- */

typedef struct steam_conv_plan steam_conv_plan;

#define STEAM_PREC_PLAIN 0 /* float32 FFT, float32 overlap-add */
#define STEAM_PREC_KAHAN 1 /* float32 FFT, compensated overlap-add */

/* Plan the operator for field (nx, ny, nz) and kernel (kx, ky, kz).  n_fft is
 * the z transform length (even; 0 = let the planner choose).  A kernel wider
 * than the field in x/y is folded as fold_kernel_to_field does. */
int steam_conv_plan_create(int64_t nx, int64_t ny, int64_t nz, int64_t kx,
                           int64_t ky, int64_t kz, int64_t n_fft,
                           int precision, steam_conv_plan** plan);
int steam_conv_plan_destroy(steam_conv_plan* plan);

/* The z transform lengths a plan may use for (nz, kz), best model cost
 * first (spectral planes, favouring lengths with a register z plan); the
 * first is what n_fft = 0 selects.  Writes up to max_out values.  The
 * reference picks its block length from a memory budget instead
 * (steam/utils.py:35-131); an autotuner times the leading candidates. */
int steam_conv_nfft_candidates(int64_t nz, int64_t kz, int64_t* out,
                               int32_t max_out, int32_t* count);

/* The half lengths n_fft / 2 that have a compiled register z plan (a
 * single-block transform whose lines are held in registers); writes up to
 * max_out values in increasing order.  No counterpart in the reference,
 * which transforms with torch.fft (steam/utils.py:372-417). */
int steam_conv_zplans(int32_t* out, int32_t max_out, int32_t* count);

/* The x/y line lengths, other than the powers of two 64..2048 of the radix-8
 * register passes, that have a compiled mixed-radix register plan (lines
 * held in registers, exchanged through LDS); writes up to max_out values in
 * increasing order.  Others take the LDS-staged Stockham passes.  No
 * counterpart in the reference, which transforms with torch.fft
 * (steam/utils.py:372-417). */
int steam_conv_xyplans(int32_t* out, int32_t max_out, int32_t* count);

typedef struct steam_conv_info {
  int64_t n_fft;           /* z transform length */
  int64_t n_blocks;        /* overlap-add blocks along z */
  int64_t block_len;       /* input planes per block */
  int64_t workspace_bytes; /* caller-allocated scratch for execute */
  int64_t spectrum_bytes;  /* caller-allocated kernel spectrum */
  int64_t spectral_elems;  /* complex values per spectral pass */
} steam_conv_info;
int steam_conv_plan_info(const steam_conv_plan* plan, steam_conv_info* info);

/* Kernel spectrum of kernel_dev (kx, ky, kz) into spectrum_dev. */
int steam_conv_set_kernel(const steam_conv_plan* plan, const float* kernel_dev,
                          void* spectrum_dev, void* workspace_dev,
                          steam_stream_t stream);

/* out = field (*) kernel.  in_dev and out_dev may alias. */
int steam_conv_execute(const steam_conv_plan* plan, const float* in_dev,
                       float* out_dev, const void* spectrum_dev,
                       void* workspace_dev, steam_stream_t stream);

/* ------------------------------------------------------------------------
 * (d) flux-cascade advance (steam/simulate.py:2066-2228)
 * ---------------------------------------------------------------------- */

/* The realized means of one advance as additive (total, count) pairs, plus
 * the clip count (steam/ledger.py:170-202). */
typedef struct steam_flux_ledger {
  double entering_total;
  int64_t entering_count;
  double noise_abs_total;
  int64_t noise_abs_count;
  double realized_total;
  int64_t realized_count;
  int64_t n_clipped;
  float rescale;   /* float32(entering_mean / realized_mean) */
  int32_t flatten; /* 1 when the realized mean was 0 (flux set flat) */
} steam_flux_ledger;

#define STEAM_MEASURE 0 /* take every reduction, fill *ledger */
#define STEAM_APPLY 1   /* inject the reductions from *ledger */

#define STEAM_DEFER_RESCALE 1 /* leave the new rescale pending (in the ledger) */
#define STEAM_CHAIN_PENDING 2 /* read the pending rescale from the ledger the
                                 previous advance left in this workspace */
#define STEAM_CLIP_IN_WINDOW 4 /* count clipped cells inside the window only
                                  (the streamed passes count a tile's interior,
                                  steam/streaming.py:1233-1234; the in-RAM
                                  advance counts the whole array,
                                  steam/simulate.py:2198) */
#define STEAM_UNIT_FLUX 16 /* the entering flux is 1 everywhere: flux_dev is
                              written, never read (a served tile's first
                              class; no fill of the buffer beforehand).  In
                              apply mode the recorded rescale is then stored
                              by the advance itself (no rescale pass) */
#define STEAM_SK_RAW 8 /* leave noise * entering flux in sk_dev, without the
                          1/<|noise|> scale: the streamed passes measure the
                          pattern on the raw product before the class's mean
                          |noise| is known (steam/streaming.py:1166-1196) */

/* Multiplier-noise field of one advance, (nx, ny, nz) floats into out_dev
 * (the VALU-bound half of the advance).  It does not read the flux, so a
 * caller may compute it ahead on another stream and pass it to
 * steam_flux_advance as noise_dev. */
int steam_flux_noise(const steam_conv_plan* plan, const steam_noise_params* noise,
                     float* out_dev, steam_stream_t stream);

/* One class advance of flux (nx, ny, nz) in place
 * (steam/simulate.py:2066-2228):
 *   f = sigma_in * flux                 (pending rescale applied on read)
 *   noise = multiplier noise            (regenerated from counters)
 *   flux  = max(f + kernel (*) (noise f), 0)
 *   flux *= rescale                      (unless STEAM_DEFER_RESCALE)
 * window = {x0, x1, y0, y1} restricts the volume means and the noise mean to
 * the domain (steam/simulate.py:2052-2063); NULL = whole array.
 * sk_dev (optional) receives S_k = noise f / <|noise|>.  noise_dev
 * (optional) is the class's noise field from steam_flux_noise; NULL
 * generates it here.  flags: see above.  Synchronises iff ledger != NULL in
 * measure mode (the ledger is copied back to the host). */
int steam_flux_advance(const steam_conv_plan* plan, const void* spectrum_dev,
                       void* workspace_dev, float* flux_dev, float sigma_in,
                       const steam_noise_params* noise, const float* noise_dev,
                       const int64_t* window, int mode,
                       steam_flux_ledger* ledger, float* sk_dev, int flags,
                       steam_stream_t stream);

/* Device-side copy of the last advance's ledger (inside workspace_dev). */
int steam_flux_ledger_read(const steam_conv_plan* plan,
                           const void* workspace_dev, steam_flux_ledger* out,
                           steam_stream_t stream);

/* ------------------------------------------------------------------------
 * Batched advance: n independent instances (realisations or tiles of
 * different sizes, alpha and noise scale), each exactly the
 * steam_flux_advance of its own plan (measure mode, or apply-only with a
 * recorded ledger), advanced by one launch per pass and compiled kernel
 * variant.  The reference advances instances one call at a
 * time (steam/simulate.py:2066-2228 per class and realisation); the batch
 * returns, per instance, the same result as that call.
 * ---------------------------------------------------------------------- */
typedef struct steam_batch steam_batch;

typedef struct steam_batch_item {
  const steam_conv_plan* plan;
  const void* spectrum_dev;   /* kernel spectrum of this instance's plan */
  void* workspace_dev;        /* this instance's own workspace */
  float* flux_dev;            /* (nx, ny, nz), advanced in place */
  const float* noise_dev;     /* precomputed noise field, or NULL */
  float* sk_dev;              /* optional S_k output, or NULL */
  steam_noise_params noise;
  int64_t window[4];          /* x0, x1, y0, y1; all zero = whole field */
  float sigma_in;             /* pending rescale applied on read */
  int32_t flags;              /* STEAM_DEFER_RESCALE | STEAM_CHAIN_PENDING |
                                 STEAM_CLIP_IN_WINDOW | STEAM_UNIT_FLUX */
  const steam_flux_ledger* ledger; /* NULL: measure; else apply-only with these
                                      reductions (copied at creation) */
} steam_batch_item;

/* Group the instances by kernel variant and stage their descriptors
 * (pinned host copy + device copy).  Allocates.  An item with a ledger runs
 * the apply-only advance (STEAM_APPLY) of steam_flux_advance. */
int steam_batch_create(const steam_batch_item* items, int32_t n,
                       steam_batch** batch);
int steam_batch_destroy(steam_batch* batch);

/* Replace instance i's keyed-noise parameters for the next advance (waits
 * for the previous advance's descriptor upload to finish). */
int steam_batch_set_noise(steam_batch* batch, int32_t i,
                          const steam_noise_params* noise);

/* Store only a box of instance i's advanced flux, dense (x, y, z order) in
 * out_dev, for its next advances; the flux grid itself is then not written.
 * origin[3] and shape[3] are the box's corner and extent in the instance's
 * grid.  Only for an apply-mode instance with a unit entering flux
 * (STEAM_UNIT_FLUX), no S_k output and a register z plan: a served tile,
 * whose interior is the box, so no copy out of the padded grid follows the
 * advance.  out_dev NULL stores the whole grid again.  Waits for the
 * previous advance's descriptor upload, as steam_batch_set_noise does.  No
 * counterpart in the reference, whose tiles are cut from whole fields
 * (steam/streaming.py:822-847). */
int steam_batch_set_crop(steam_batch* batch, int32_t i, float* out_dev,
                         const int64_t* origin, const int64_t* shape);

/* The noise fields of the instances created with a noise_dev buffer, for
 * their next advance: every such instance's multiplier noise (its current
 * keyed-noise parameters) written into its buffer, one launch per variant,
 * after a descriptor upload; graph-capturable like steam_batch_advance.  A
 * caller runs it on another stream ahead of the advance that reads the
 * buffers, so the VALU-bound noise of one step overlaps the memory-bound
 * transforms of the step before.  No counterpart in the reference (which
 * draws each instance's noise inside its own advance,
 * steam/simulate.py:2146-2161). */
int steam_batch_noise(steam_batch* batch, steam_stream_t stream);

/* Mark, on stream, the point after a replay of a captured
 * steam_batch_advance.  A replayed graph copies the pinned descriptors when
 * the replay runs, not when it was captured, so the next
 * steam_batch_set_noise must not rewrite them before that copy: it waits for
 * the last marked point.  Call it after every replay (eager advances mark
 * themselves).  No counterpart in the reference, which has no batched or
 * captured advance (it advances one instance per call,
 * steam/simulate.py:2066-2228). */
int steam_batch_mark_replay(steam_batch* batch, steam_stream_t stream);

/* Kernel launches per advance (descriptor upload included) and the bytes of
 * descriptors uploaded per advance. */
int steam_batch_info(const steam_batch* batch, int32_t* launches,
                     int64_t* descriptor_bytes);

/* One measure-mode advance of every instance on stream: a host-to-device
 * descriptor copy followed by the pass launches; nothing else is enqueued
 * and nothing is allocated, so the call may be captured into a hipGraph
 * (replays read the pinned descriptor copy as it is at replay time).  Each
 * instance's ledger stays in its workspace (steam_flux_ledger_read). */
int steam_batch_advance(steam_batch* batch, steam_stream_t stream);

/* steam_batch_advance without its descriptor upload, for a batch whose
 * every instance reads a noise buffer: the advance runs on the descriptors
 * that the batch's last steam_batch_noise uploaded (the caller orders that
 * call before this one, and calls steam_batch_set_noise on the batch only
 * after it).  A pipelined caller then writes a batch's staging copy for only
 * one copy per step, its noise upload, so replays need not wait for a whole
 * step before the next step's keys are written.  No counterpart in the
 * reference, which has no batched or pipelined advance (one instance per
 * call, steam/simulate.py:2066-2228). */
int steam_batch_advance_after_noise(steam_batch* batch, steam_stream_t stream);

/* ------------------------------------------------------------------------
 * Scalar increments of the cascade (steam/simulate.py:1924-2029): the joint
 * pattern, per-level window reductions and the bounded amplitude-preserving
 * add.  Fields (nx, ny, nz) float32; a level is the (x, y) plane at one z.
 * ---------------------------------------------------------------------- */

/* np.gradient's z differences in float32 (steam/simulate.py:2709-2741):
 * uniform spacing uses (f[z+1]-f[z-1])/two_dz, otherwise the three-term
 * a f[z-1] + b f[z] + c f[z+1] (device arrays of nz-2); one-sided ends. */
typedef struct steam_zdiff {
  int32_t uniform;
  float two_dz;
  float dz_first;
  float dz_last;
  const float* a_dev;
  const float* b_dev;
  const float* c_dev;
} steam_zdiff;

/* out = W * S_k * g per cell: W the advective weight |grad_h phi| +
 * aspect(z) |d phi/dz| of the running field phi = pert + mean(z)
 * (_advective_weight, steam/simulate.py:2787-2868), g the bound taper
 * clip(min(phi - min, max - phi) / b(z), 0, 1) (_bound_taper, :2905-2930),
 * in the reference's float32 operation order.  sk_dev = NULL gives the
 * advective weight W alone (b_z_dev unused). */
int steam_scalar_pattern(const float* pert_dev, const float* mean_z_dev,
                         const float* sk_dev, const float* b_z_dev,
                         const float* aspect_z_dev, const steam_zdiff* zdiff,
                         int64_t nx, int64_t ny, int64_t nz, double dx, double dy,
                         double phi_min, double phi_max, float* out_dev,
                         steam_stream_t stream);

/* The joint pattern without forming the float32 running field (the
 * precision workload): horizontal differences of pert alone, the vertical
 * one of pert plus dmean_z[z] (the mean profile's z-derivative, formed in
 * float64 by the caller), the taper from pert + mlo_z[z] and mhi_z[z] - pert
 * with mlo = mean - phi_min and mhi = phi_max - mean formed in float64.  The
 * float64 transform of _advective_weight / _bound_taper
 * (steam/simulate.py:2787-2930) to within the rounding of each difference;
 * no counterpart in the reference, whose float32 form is
 * steam_scalar_pattern. */
int steam_scalar_pattern_precise(const float* pert_dev, const float* dmean_z_dev,
                                 const float* mlo_z_dev, const float* mhi_z_dev,
                                 const float* sk_dev, const float* b_z_dev,
                                 const float* aspect_z_dev, const steam_zdiff* zdiff,
                                 int64_t nx, int64_t ny, int64_t nz, double dx, double dy,
                                 float* out_dev, steam_stream_t stream);

/* w = w / (level_mean > 0 ? level_mean : 1) * c(z), the product-norm apply
 * (steam/simulate.py:1986-1990); level_mean and c are float32 per level. */
int steam_pattern_normalize(float* w_dev, const float* level_mean_dev,
                            const float* c_z_dev, int64_t n, int64_t nz,
                            steam_stream_t stream);

#define STEAM_LV_SUM 0     /* sum f */
#define STEAM_LV_ABS 1     /* sum |f| */
#define STEAM_LV_NONZERO 2 /* count f != 0 */
#define STEAM_LV_CLAMP 3   /* sum clamp(f - float(mu[z]), lo, hi) */
#define STEAM_LV_PHI 4     /* sum (f + mean[z]) */
#define STEAM_LV_MIN 5     /* min (f + mean[z]) over the whole level */
#define STEAM_LV_MAX 6     /* max (f + mean[z]) over the whole level */
/* the field in float64, e = (double)f + (double)mean[z] */
#define STEAM_LV_E_SUM 7   /* sum e */
#define STEAM_LV_E_MIN 8   /* min e over the whole level */
#define STEAM_LV_E_MAX 9   /* max e over the whole level */
#define STEAM_LV_E_CLAMP 10 /* sum clamp(e - mu[z], lo, hi) in float64 */

/* Device scratch bytes steam_level_reduce needs for this field shape. */
int64_t steam_level_scratch_bytes(int64_t nx, int64_t ny, int64_t nz);

/* Per-level reduction over window {x0, x1, y0, y1} (NULL = whole level) into
 * out_dev[nz] (float64, device), in a fixed order (runs of consecutive window
 * lines, then the runs in index order; the order depends on the window's
 * size only): the level sums of the product norm (PatternLedger,
 * steam/ledger.py:83-108), of the bounded add and of the projection.  The
 * bounded add's and the projection's bisections evaluate exactly these sums
 * at each of their midpoints. */
int steam_level_reduce(const float* f_dev, const float* mean_z_dev,
                       const double* mu_dev, int64_t nx, int64_t ny, int64_t nz,
                       const int64_t* window, int op, double lo, double hi,
                       double* out_dev, void* scratch_dev, int64_t scratch_bytes,
                       steam_stream_t stream);

/* The per-level scalars one bounded-add solve realized
 * (BoundedAddLedger, steam/ledger.py:145-190); host arrays (device arrays
 * with replay == 2), demean and scale row-major (nz, n_rescale). */
typedef struct steam_bounded_add_ledger {
  double* a0;
  double* demean;
  double* scale;
  int32_t* n_loop;
  int32_t* n_scale;
  double* final_demean;
  double* mu;
} steam_bounded_add_ledger;

int64_t steam_bounded_add_workspace_bytes(int64_t nx, int64_t ny, int64_t nz);

/* Add increment inc to pert under the three deposit conditions per level
 * (zero window mean, window mean |increment| preserved, bounds respected):
 * demean / clip / rescale loop, then a 60-step bisection of the zero-mean
 * shift, as _bounded_amplitude_add's GPU path (steam/simulate.py:2985-3111).
 * replay == 1 applies the recorded ledger with no reductions; replay == 0
 * records it into host arrays (synchronising once, for the copy); replay ==
 * 2 records it into device arrays in stream order, with no
 * synchronisation.  With record_applied, inc is left holding the increment
 * actually added. */
int steam_bounded_add(float* pert_dev, const float* mean_z_dev, float* inc_dev,
                      int64_t nx, int64_t ny, int64_t nz, const int64_t* window,
                      double phi_min, double phi_max, int n_rescale,
                      int record_applied, steam_bounded_add_ledger* ledger,
                      int replay, void* workspace_dev, int64_t workspace_bytes,
                      steam_stream_t stream);

/* Compensation deficit update, deficit += float(coef[z] * inc) with
 * inc = after - before (before may be NULL: inc = after).  Replaces the
 * deficit accumulations of steam/simulate.py:1893-1919 (flux, with the
 * state before and after the advance) and 2015-2016 (scalars, with the
 * recorded applied increment). */
int steam_deficit_accumulate(float* deficit_dev, const float* after_dev,
                             const float* before_dev, const float* coef_z_dev,
                             int64_t n, int64_t nz, steam_stream_t stream);

/* Device workspace bytes steam_project_onto_bounds and
 * steam_compose_output need for this shape. */
int64_t steam_project_workspace_bytes(int64_t nx, int64_t ny, int64_t nz);

/* Projection onto [phi_min, phi_max] preserving each level's window mean
 * (_project_onto_bounds, steam/simulate.py:3417-3500, in the float64 form
 * of its GPU dispatch _project_onto_bounds_cuda, 3286-3414): levels of e = pert + mean (float64) with any cell out of
 * bounds become float(clamp(e - mu, lo, hi) - mean) with mu from a 60-step
 * float64 bisection; other levels are left bit-identical.  mu[nz] (host)
 * receives the solved constants, NaN for untouched levels; with replay it
 * supplies them (apply-only, no reduction).  Synchronizes the stream. */
int steam_project_onto_bounds(float* pert_dev, const float* mean_z_dev, int64_t nx,
                              int64_t ny, int64_t nz, const int64_t* window,
                              double phi_min, double phi_max, double* mu, int replay,
                              void* workspace_dev, int64_t workspace_bytes,
                              steam_stream_t stream);

/* The written scalar field (_compose_output, steam/simulate.py:3558-3599):
 * with a deficit, deficit += state, projected as above, out =
 * clamp(deficit + mean, float(lo), float(hi)) (out may be deficit); with
 * deficit NULL, out = clamp(state + mean) and no projection is taken. */
int steam_compose_output(const float* state_dev, float* deficit_dev,
                         const float* mean_z_dev, int64_t nx, int64_t ny, int64_t nz,
                         const int64_t* window, double phi_min, double phi_max,
                         double* mu, int replay, float* out_dev, void* workspace_dev,
                         int64_t workspace_bytes, steam_stream_t stream);

/* Device workspace bytes steam_compose_flux_output needs for this shape. */
int64_t steam_compose_flux_workspace_bytes(int64_t nx, int64_t ny, int64_t nz);

/* The written flux (_compose_flux_output, steam/simulate.py:3602-3649):
 * deficit = max(deficit + state, 0), rescaled by float(entering mean /
 * realized mean) over the window (or filled with the entering mean when
 * nothing is left).  sums[4] = {entering total, entering count, realized
 * total, realized count} (host; out, or in with replay).  Synchronizes the
 * stream when measuring. */
int steam_compose_flux_output(const float* state_dev, float* deficit_dev, int64_t nx,
                              int64_t ny, int64_t nz, const int64_t* window,
                              double* sums, int replay, void* workspace_dev,
                              int64_t workspace_bytes, steam_stream_t stream);

/* ------------------------------------------------------------------------
 * (d) regrid: periodic trilinear upsample (steam/utils.py:435-534) and its
 *     tiled window form (steam/streaming.py:132-242).
 * ---------------------------------------------------------------------- */

/* dst[x, y, z] for target cells [x0, x0+wx) x [y0, y0+wy) x [0, tz) of the
 * zoom of src (sx, sy, sz) to (tx, ty, tz), each source value first
 * multiplied by float32 scale.  periodic = {px, py, pz}.  dst is (wx, wy, tz).
 */
int steam_regrid_trilinear(const float* src_dev, int64_t sx, int64_t sy,
                           int64_t sz, int64_t tx, int64_t ty, int64_t tz,
                           int64_t x0, int64_t wx, int64_t y0, int64_t wy,
                           const int32_t* periodic, float scale,
                           float* dst_dev, steam_stream_t stream);

/* Elementwise flux *= scale (or = flat when flatten) over n values. */
int steam_scale_inplace(float* data_dev, int64_t n, float scale,
                        steam_stream_t stream);

/* Deterministic float64 sum of a float32 array over a window of its first
 * two axes (x0,x1,y0,y1; NULL = whole): total and element count. */
int steam_sum_f64(const float* data_dev, int64_t nx, int64_t ny, int64_t nz,
                  const int64_t* window, double* total_out,
                  steam_stream_t stream);

/* ------------------------------------------------------------------------
 * instrumentation: per-kernel-family timing of this library's own launches,
 * by HIP events recorded on each launch's stream.
 * ---------------------------------------------------------------------- */

#define STEAM_KF_ZFWD 0     /* z forward pass (noise + r2c) */
#define STEAM_KF_YPASS 1    /* y passes (forward and inverse) */
#define STEAM_KF_XPASS 2    /* x pass: forward, spectral multiply, inverse */
#define STEAM_KF_ZINV 3     /* z inverse pass (c2r, overlap-add, epilogue) */
#define STEAM_KF_FINALIZE 4 /* ledger reduction */
#define STEAM_KF_POST 5     /* rescale / S_k scale */
#define STEAM_KF_NOISE 6    /* multiplier-noise field */
#define STEAM_KF_COUNT 7

/* Compute-only VALU ceiling probe: workgroups of 256 threads, each lane
 * running 8 independent float32 fma chains for iters iterations; writes
 * workgroups * 256 floats to out_dev and the number of wave-level fma
 * instructions issued to *wave_fmas.  Time it with events on stream to get
 * the measured issue rate of vector instructions, the ceiling of the
 * VALU-bound noise kernel.  No counterpart in the reference. */
int steam_valu_probe(int32_t workgroups, int32_t iters, float* out_dev,
                     int64_t* wave_fmas, steam_stream_t stream);

/* Start (on != 0) or stop recording; starting clears earlier records. */
int steam_profile_enable(int on);
/* Synchronise on the recorded events and report one family's launch count
 * and summed duration in milliseconds. */
int steam_profile_read(int family, int64_t* launches, double* total_ms);

#ifdef __cplusplus
}
#endif

#endif /* STEAM_HIP_H */
