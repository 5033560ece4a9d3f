/*
 * psx.h - C ABI of the MI355X-native acceleration-search core (libpsx.so).
 *
 * This is the drop-in boundary below peasoup's C++ transform classes.  Each
 * entry point replaces one function of the reference's kernel layer
 * (include/kernels/kernels.h) or one transform method, cited per function as
 * file:line of the reference source tree.  Conventions:
 *   - every function returns an int status (PSX_OK == 0); on failure
 *     psx_last_error_string() describes the error of the calling thread;
 *   - device pointers are plain pointers from psx_malloc (or any HIP
 *     allocation on the current device); sizes are element counts;
 *   - psx_stream is a hipStream_t (NULL = the default stream); kernel entry
 *     points enqueue work and return without synchronising unless stated;
 *   - no entry point below the "allocating" section allocates device memory.
 */
#ifndef PSX_H
#define PSX_H

#include <stddef.h>
#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

/* Layout-identical to cufftComplex (interleaved float re, im). */
typedef struct psx_c32 {
  float x, y;
} psx_c32;

typedef void* psx_stream; /* hipStream_t */
typedef void* psx_event;  /* hipEvent_t */

enum {
  PSX_OK = 0,
  PSX_ERR_ARG = 1,         /* bad argument (size, null pointer, range) */
  PSX_ERR_HIP = 2,         /* HIP runtime error */
  PSX_ERR_ALLOC = 3,       /* device allocation failed */
  PSX_ERR_OVERFLOW = 4,    /* output capacity exceeded (result truncated) */
  PSX_ERR_UNSUPPORTED = 5, /* size or option not supported */
  PSX_ERR_BUDGET = 6,      /* request exceeds the memory budget */
  PSX_ERR_STATE = 7,       /* object in the wrong state */
  PSX_ERR_TIMEOUT = 8      /* a bounded device wait gave up */
};

/* Error text of the last failing call on this thread (ErrorChecker,
 * include/utils/exceptions.hpp:13-74, which throws instead). */
const char* psx_last_error_string(void);
int psx_version(void);

/* ------------------------------------------------------------------ */
/* Device, memory and stream plumbing (utils.hpp:23-63 Utils::*)        */
/* ------------------------------------------------------------------ */
int psx_device_count(int* count);
int psx_set_device(int device);
int psx_get_device(int* device);
/* name (NUL-terminated, truncated to cap), HBM bytes, compute units. */
int psx_device_info(int device, char* name, size_t cap, uint64_t* hbm_bytes, int* cu_count, int* arch_gfx);
int psx_malloc(void** ptr, uint64_t bytes);         /* Utils::device_malloc, utils.hpp:23-27 */
int psx_free(void* ptr);                             /* Utils::device_free, utils.hpp:35-39 */
int psx_host_malloc(void** ptr, uint64_t bytes);    /* Utils::host_malloc (pinned), utils.hpp:29-33 */
int psx_host_free(void* ptr);                        /* Utils::host_free, utils.hpp:41-45 */
int psx_memcpy_h2d(void* dst, const void* src, uint64_t bytes, psx_stream s); /* utils.hpp:47-51 */
int psx_memcpy_d2h(void* dst, const void* src, uint64_t bytes, psx_stream s); /* utils.hpp:53-57 */
int psx_memcpy_d2d(void* dst, const void* src, uint64_t bytes, psx_stream s); /* utils.hpp:59-63 */
int psx_stream_create(psx_stream* s);
int psx_stream_destroy(psx_stream s);
int psx_stream_synchronize(psx_stream s);
int psx_device_synchronize(void);
int psx_event_create(psx_event* e);
int psx_event_destroy(psx_event e);
int psx_event_record(psx_event e, psx_stream s);
int psx_event_synchronize(psx_event e);
int psx_event_elapsed_ms(psx_event start, psx_event stop, float* ms);
int psx_mem_info(uint64_t* free_bytes, uint64_t* total_bytes);

/* ------------------------------------------------------------------ */
/* Per-stage kernels (the reference kernel layer, kernels.h:104-296)   */
/* ------------------------------------------------------------------ */

/* TimeDomainResampler::resampleII (resampler.hpp:26-31) ->
 * device_resampleII (src/kernels.cu:355-369): out[i] = in[rn(i + i*af*(i-n))],
 * af = double(float(acc*tsamp))/(2c).  Requires |af|*n < 1 (PSX_ERR_ARG). */
int psx_resample(const float* in, float* out, uint64_t n, float acc, float tsamp, psx_stream s);
/* TimeDomainResampler::resample (resampler.hpp:20-24) -> device_resample
 * (src/kernels.cu:371-386): out[i] = in[rn(i + af*((i-n/2)^2 - (n/2)^2))]. */
int psx_resample_v1(const float* in, float* out, uint64_t n, float acc, float tsamp, psx_stream s);

/* SpectrumFormer::form (spectrumformer.hpp:16-22) -> device_form_power_series
 * way 0 (src/kernels.cu:216-234,294-311): A[k] = |X[k]|. */
int psx_form_amplitude(const psx_c32* X, float* A, uint64_t nb, psx_stream s);
/* SpectrumFormer::form_interpolated (spectrumformer.hpp:8-14) -> way 1
 * (src/kernels.cu:238-259): A[k] = sqrt(max(|X_k|^2, |X_k - X_{k-1}|^2/2)). */
int psx_form_interbin(const psx_c32* X, float* A, uint64_t nb, psx_stream s);

/* median_scrunch5 (src/kernels.cu:974-1008, kernels.h:280-282). */
int psx_median_scrunch5(const float* in, uint64_t count, float* out, psx_stream s);
/* linear_stretch (src/kernels.cu:1026-1038, kernels.h:284-287). */
int psx_linear_stretch(const float* in, uint64_t in_count, float* out, uint64_t out_count, psx_stream s);
/* device_divide_c_by_f (src/kernels.cu:1040-1061, kernels.h:290-294). */
int psx_divide_c_by_f(psx_c32* X, const float* f, uint64_t nb, psx_stream s);
/* Zapper::zap (include/transforms/birdiezapper.hpp:60-71) ->
 * device_zap_birdies (src/kernels.cu:1063-1096): for each birdie (device
 * arrays freqs/widths, Hz), bins [floor((f-w)/bin_width), ceil((f+w)/bin_width))
 * clamped to [0, nb-1) are set to 1+0i. */
int psx_zap_birdies(psx_c32* X, const float* freqs, const float* widths, uint32_t nbirdies, float bin_width,
                    uint64_t nb, psx_stream s);
/* Dereddener::calculate_median (dereddener.hpp:41-62): running median of A
 * into median (nb), using work (>= psx_deredden_work_floats(nb) floats). */
uint64_t psx_deredden_work_floats(uint64_t nb);
int psx_deredden_median(const float* A, uint64_t nb, double bin_width, float boundary_5_freq,
                        float boundary_25_freq, float* median, float* work, psx_stream s);
/* No counterpart in the reference: psx_deredden_median for callers that
 * cannot allocate device memory (code under HIP graph capture, serving
 * loops, harnesses that forbid hipMalloc).  It allocates nothing: the
 * scrunched medians live in a static device array of
 * psx_deredden_work_floats(psx_deredden_noalloc_max_bins()) floats (33 MB)
 * held by each device until the process ends.  nb above
 * psx_deredden_noalloc_max_bins() (2^25 + 1, a 2^26-sample series) is
 * rejected with PSX_ERR_ARG.  Calls share that array per device, so calls
 * that overlap on one device race; the results equal psx_deredden_median's. */
uint64_t psx_deredden_noalloc_max_bins(void);
int psx_deredden_median_noalloc(const float* A, uint64_t nb, double bin_width, float boundary_5_freq,
                                float boundary_25_freq, float* median, psx_stream s);

/* GPU_mean / GPU_rms (src/kernels.cu:427-461) as used by stats::stats
 * (stats.hpp:25-33).  Synchronises s; results on the host.  work must hold
 * psx_reduce_work_bytes() bytes of device memory. */
uint64_t psx_reduce_work_bytes(void);
int psx_mean_rms(const float* x, uint64_t n, uint64_t first, float* mean, float* rms, void* work,
                 psx_stream s);
/* device_normalise (src/kernels.cu:496-521): p = (p - mean)/sigma. */
int psx_normalise(float* p, float mean, float sigma, uint64_t n, psx_stream s);
/* GPU_remove_baseline (src/kernels.cu:463-479); synchronises s per pass. */
int psx_remove_baseline(float* x, uint64_t n, void* work, psx_stream s);
/* GPU_fill (src/kernels.cu:482-488): x[start, end) = value. */
int psx_fill(float* x, uint64_t start, uint64_t end, float value, psx_stream s);
/* device_conversion (src/kernels.cu:1184-1200): out[i] = (float)in[i]. */
enum { PSX_U8 = 0, PSX_I8 = 1, PSX_U32 = 2, PSX_F32 = 3 };
int psx_convert_to_f32(const void* in, int in_type, float* out, uint64_t n, psx_stream s);

/* HarmonicFolder::fold (harmonicfolder.hpp:26-38) -> device_harmonic_sum
 * (src/kernels.cu:34-100,200-209).  folds is a HOST array of nfolds device
 * pointers (1 <= nfolds <= 5); no pointer table is copied to the device. */
int psx_harmonic_sum(const float* P, float* const* folds, uint32_t nfolds, uint64_t nb, psx_stream s);

/* device_find_peaks (src/kernels.cu:391-423): stable (ascending-index)
 * compaction of (i, p[i]) with p[i] > thr over i in [start, end).  Writes at
 * most cap pairs to d_idx/d_snr and the full count to *d_count (device).
 * work must hold psx_find_peaks_work_bytes(end - start) bytes. */
uint64_t psx_find_peaks_work_bytes(uint64_t n);
int psx_find_peaks(const float* p, int64_t start, int64_t end, float thr, int32_t* d_idx, float* d_snr,
                   uint32_t cap, uint32_t* d_count, void* work, psx_stream s);

/* ------------------------------------------------------------------ */
/* FFT (CuFFTerR2C / CuFFTerC2R, include/transforms/ffter.hpp:54-110)   */
/* ------------------------------------------------------------------ */
typedef struct psx_fft_plan psx_fft_plan;
enum { PSX_FFT_R2C = 0, PSX_FFT_C2R = 1 };
/* Allocating: the plan owns its twiddle tables and work buffer.  n is the real
 * length (even; 2^k or 3*2^k, 2^4 <= n <= 2^28). */
int psx_fft_plan_create(uint64_t n, int kind, uint32_t batch, psx_fft_plan** plan);
/* The same plan with a chosen decomposition: npass (1-3) passes of lengths
 * lens[q] (product n/2; pass 0 strided, the last contiguous) and column
 * tiles cols[q] (0 = the widest that fits).  cuFFT picks its own
 * decomposition (ffter.hpp:60-70 takes none); this is how the autotuner's
 * schedules (psx_batch_opts.fft_len/fft_cols) are checked as plain FFTs. */
int psx_fft_plan_create_decomp(uint64_t n, int kind, uint32_t batch, int npass, const int* lens, const int* cols,
                               psx_fft_plan** plan);
int psx_fft_plan_destroy(psx_fft_plan* plan);
/* R2C: in = batch x n float, out = batch x (n/2+1) psx_c32 (unnormalised).
 * C2R: in = batch x (n/2+1) psx_c32, out = batch x n float (scaled by n).
 * in and out may not alias. */
int psx_fft_exec(psx_fft_plan* plan, const void* in, void* out, psx_stream s);
/* Host-only check (no device work): builds the pass geometry of a plan
 * (npass = 0: default decomposition, else lens/cols per pass) and verifies
 * that every launch reads and writes inside its buffers, for the plain
 * transform and for the fused batched search. */
int psx_fft_layout_check(uint64_t n, int kind, uint32_t batch, int npass, const int* lens, const int* cols);
/* Number of kernel launches one exec enqueues (for profiling). */
int psx_fft_plan_describe(const psx_fft_plan* plan, char* buf, size_t cap);

/* ------------------------------------------------------------------ */
/* Synthetic dedispersed series (deterministic, counter-based RNG)      */
/* ------------------------------------------------------------------ */
typedef struct psx_pulsar {
  double period;   /* s */
  double accel;    /* m/s^2 */
  double snr;      /* target folded S/N */
  double phase0;   /* turns */
  double duty;     /* FWHM as a fraction of the period */
} psx_pulsar;
/* out[i] = N(0,1) (Philox4x32-10, key (seed, trial)) + red noise (random
 * walk, red_amp x white amplitude at 0.1 Hz; 0 disables) + pulsars, minus the
 * mean.  work must hold psx_synth_work_bytes(n) bytes. */
uint64_t psx_synth_work_bytes(uint64_t n);
int psx_synth_series(float* out, uint64_t n, double tsamp, uint64_t seed, uint32_t trial, double red_amp,
                     const psx_pulsar* pulsars, int npulsars, void* work, psx_stream s);

/* ------------------------------------------------------------------ */
/* Whole-DM-trial search drivers (Worker level)                        */
/* ------------------------------------------------------------------ */
/* The CmdLineOptions fields the reference Worker reads
 * (include/utils/cmdline.hpp:77-256; pipeline_multi.cu:111-149,412-421). */
typedef struct psx_search_args {
  uint64_t size;          /* FFT length N (--fft_size) */
  float tsamp;            /* sampling time, s */
  float min_snr;          /* -m/--min_snr */
  float min_freq;         /* --min_freq, Hz */
  float max_freq;         /* --max_freq, Hz */
  uint32_t nharmonics;    /* -n/--nharmonics, 1..5 */
  float freq_tol;         /* --freq_tol */
  float max_harm;         /* --max_harm_match */
  float acc_start;        /* --acc_start, m/s^2 */
  float acc_end;          /* --acc_end, m/s^2 */
  float acc_tol;          /* --acc_tol */
  float acc_pulse_width;  /* --acc_pulse_width, microseconds */
  float cdm;              /* --cdm */
  float cfreq;            /* centre frequency, Hz (get_cfreq()*1e6) */
  float ch_bw;            /* channel bandwidth, Hz (get_foff()*1e6) */
  int32_t min_gap;        /* PeakFinder min_gap (30) */
  uint32_t nbirdies;      /* birdies to zap after dereddening (--zapfile); 0 = none */
  const float* birdie_freq;   /* host, nbirdies frequencies (Hz); read during the call */
  const float* birdie_width;  /* host, nbirdies half-widths (Hz) */
} psx_search_args;

/* CandidatePOD fields (include/data_types/candidates.hpp:10-17) plus the
 * associated-candidate count (Candidate::count_assoc, :42-49).
 *
 * Candidate lists are written as a forest: each top-level candidate is
 * followed by the candidates associated with it, in pre-order (the order of
 * Candidate::collect_candidates, :78-84, which is also the order of the
 * detections in candidates.peasoup), and nassoc counts the records of a
 * record's own subtree that follow it.  Top-level candidates are therefore
 * at i0 = 0, i1 = i0 + 1 + nassoc[i0], ... */
typedef struct psx_candidate {
  float dm;
  int32_t dm_idx;
  float acc;
  int32_t nh;
  float snr;
  float freq;
  int32_t nassoc;
  int32_t reserved;
} psx_candidate;

/* Worker::start (src/pipeline_multi.cu:111-303) over ndm host series of
 * nsamps samples each (row-major, DispersionTrials layout), one call per
 * kernel as the reference does.  accs (naccs values) overrides the
 * AccelerationPlan list when non-NULL.  Writes the acceleration-distilled
 * candidates of every DM in DM order, as a forest with their associated
 * candidates (*count = records; PSX_ERR_OVERFLOW if it exceeds cap).  seconds[0..2] (optional) = prologue, accel loop, distill. */
int psx_worker_search(const psx_search_args* args, const float* series, uint64_t nsamps, const float* dm_list,
                      uint32_t ndm, const float* accs, uint32_t naccs, psx_candidate* out, uint32_t cap,
                      uint32_t* count, double* seconds);

/* Options of the batched search (psx_batch_search).  Zero fields take
 * defaults; fft_npass > 0 forces an FFT decomposition (autotuner). */
typedef struct psx_batch_opts {
  uint32_t batch;          /* trials per fused batch (0: 64 up to 2^26 samples,
                              32 at 2^27, 16 at 2^28) */
  uint32_t cand_capacity;  /* above-threshold bins per batch (default 2^20) */
  int32_t use_graph;       /* launch each batch as one captured hipGraph */
  int32_t fft_npass;       /* 0 = default decomposition */
  int32_t fft_len[3];      /* pass lengths (product = n/2) */
  int32_t fft_cols[3];     /* column tile per pass */
  int32_t spectrum_layout; /* PSX_SPECTRUM_*: how the last FFT pass hands over the spectra */
  int32_t keep_engine;     /* keep the engine for the next call with the same arguments and
                              options (psx_batch_release frees it) */
  int32_t harmonic_pass;   /* PSX_HARMONIC_*: form of the harmonic-sum pass */
} psx_batch_opts;

/* Spectrum hand-over of the batched path.  AUTO: a multi-pass plan whose
 * last pass has a contiguous-row kernel writes contiguous rows and the
 * spectrum pass reads that transposed layout; NATURAL: always natural bin
 * order; ROWS: the transposed layout, an error when the plan has none.  All
 * three give identical spectra for the same FFT decomposition. */
enum { PSX_SPECTRUM_AUTO = 0, PSX_SPECTRUM_NATURAL = 1, PSX_SPECTRUM_ROWS = 2 };

/* Harmonic-sum pass of the batched path (harmonic_sum_kernel + PeakFinder,
 * src/kernels.cu:34-100,391-423).  BOUND: a pooled bfloat16 copy of
 * each spectrum (every entry rounded up from the larger of two bins) is
 * summed in the reference's addend order; since every addend and every
 * rounding step only grows, those sums bound the exact fold sums from
 * above, and only bins whose bound crosses the threshold are summed exactly.
 * EXACT: every addend of every bin gathered from the spectrum.  Both give
 * identical candidates; BOUND moves about a third of the memory traffic. */
enum { PSX_HARMONIC_BOUND = 0, PSX_HARMONIC_EXACT = 1 };

/* Same contract as psx_worker_search, computed by the batched path: all DM
 * prologues, then every (DM, acceleration) trial through fused batches
 * (resample+FFT, untangle+interbin+normalise, harmonic sums+threshold), with
 * the Worker's per-trial HarmonicDistiller and per-DM AccelerationDistiller.
 * series is a device pointer when series_on_device != 0.  seconds[0..2]
 * (optional) = prologues, trials, total. */
int psx_batch_search(const psx_search_args* args, const psx_batch_opts* opts, const float* series,
                     int series_on_device, uint64_t nsamps, const float* dm_list, uint32_t ndm, const float* accs,
                     uint32_t naccs, psx_candidate* out, uint32_t cap, uint32_t* count, double* seconds);

/* Frees the engine a psx_batch_search call with keep_engine left behind
 * (its device buffers, FFT plans and launch graphs; one engine is kept per
 * process, replaced by the next keeping call with other arguments).  No
 * counterpart in the reference, whose Worker builds its transforms per DM
 * (pipeline_multi.cu:111-145); for pipelines that search many segments of
 * one shape. */
int psx_batch_release(void);

/* Times each launch of one fused batch (min(naccs, batch) trials over the
 * first series of d_series, after its prologue) with HIP events on the batch
 * stream, averaged over reps.  stage_ms[0..*nstage) = the FFT passes (the
 * first with the fused resampling), untangle+interbin+normalise, harmonic
 * sums+threshold; desc receives the FFT decomposition. */
int psx_batch_profile(const psx_search_args* args, const psx_batch_opts* opts, const float* d_series,
                      uint64_t nsamps, const float* accs, uint32_t naccs, int reps, double* stage_ms, int* nstage,
                      char* desc, size_t cap);

/* Self-check of the bound-filtered harmonic pass's input (no counterpart
 * in the reference): runs the prologue of d_series and the FFT and spectrum
 * stages of one batch (min(naccs, batch) trials), has the exact spectra P
 * written from the same FFT output, and checks on the host every bound
 * entry against max(P[2u], P[2u+1]) and the fundamental's candidates against
 * the bins above threshold.  result[0] = entries checked, [1] = entries below
 * their pair's maximum (invalid bounds: 0 when correct), [2] = fundamental
 * bins above threshold on the host, [3] = candidates the device appended
 * (equal to [2] when correct), [4] = bins of the harmonic window whose bound
 * sums cross at some fold (the bins summed exactly), [5] = bins of the
 * window, [6] = valid entries above the tight bound (the pair maximum
 * rounded up to bfloat16; the row-layout pass bounds each value from
 * approximate arithmetic and may land one step higher).  result holds 7
 * values.  For tests and for validating a new FFT decomposition. */
int psx_batch_bound_audit(const psx_search_args* args, const psx_batch_opts* opts, const float* d_series,
                          uint64_t nsamps, const float* accs, uint32_t naccs, uint64_t* result);

/* ------------------------------------------------------------------ */
/* Serving: continuous batching of segment-search requests             */
/* ------------------------------------------------------------------ */
typedef struct psx_server psx_server;
/* A tuned schedule for one FFT length (the autotuner's output). */
typedef struct psx_schedule {
  uint64_t size;      /* FFT length it applies to */
  uint32_t batch;     /* trials per fused batch (0 = the server's batch) */
  int32_t fft_npass;  /* 0 = default decomposition */
  int32_t fft_len[3];
  int32_t fft_cols[3];
} psx_schedule;

typedef struct psx_server_config {
  uint64_t hbm_budget_bytes;  /* resident series + engine buffers never exceed this */
  uint32_t batch;             /* trials per fused batch (0: by length, as in
                                 psx_batch_opts) */
  uint32_t max_active;        /* admitted requests at once (0 = budget only) */
  int32_t use_graph;          /* graph-launched batches */
  uint32_t nschedules;        /* per-length schedules (copied at create; may be 0) */
  const psx_schedule* schedules;
  int32_t dispatch;           /* PSX_DISPATCH_STREAMS or PSX_DISPATCH_MIXED_GRAPH */
} psx_server_config;
/* How batches of different FFT lengths reach the device: one stream per
 * length, each launching its own batch graphs (the lengths overlap as the
 * streams allow), or one graph launch per mixed batch, whose child graphs
 * are one batch of every length that has trials waiting. */
enum { PSX_DISPATCH_STREAMS = 0, PSX_DISPATCH_MIXED_GRAPH = 1 };

/* One beam: ndm dedispersed series of nsamps samples (host or device memory,
 * valid until the request completes) searched with args. */
typedef struct psx_request {
  psx_search_args args;
  const float* series;
  int32_t series_on_device;
  uint64_t nsamps;
  const float* dm_list;
  uint32_t ndm;
  const float* accs;  /* optional acceleration list (else AccelerationPlan) */
  uint32_t naccs;
} psx_request;

enum { PSX_REQ_WAITING = 0, PSX_REQ_ACTIVE = 1, PSX_REQ_DONE = 2, PSX_REQ_FAILED = 3 };
typedef struct psx_request_status {
  int32_t state;
  int32_t error;
  uint32_t ncand;
  uint64_t ntrials;
  double t_submit, t_admit, t_done; /* seconds on the library's monotonic clock */
} psx_request_status;

typedef struct psx_server_stats {
  uint64_t used_bytes, budget_bytes;
  uint32_t waiting, active, engines;
  uint64_t completed, failed, trials_done;
} psx_server_counters;

int psx_server_create(const psx_server_config* cfg, psx_server** server);
/* Non-blocking.  PSX_ERR_BUDGET when the request alone cannot fit. */
int psx_server_submit(psx_server* server, const psx_request* req, uint64_t* id);
int psx_server_poll(psx_server* server, uint64_t id, psx_request_status* status);
int psx_server_wait(psx_server* server, uint64_t id, double timeout_s, psx_request_status* status);
/* Candidates of a completed request (per DM, accel-distilled, DM order, as a
 * forest; status.ncand = records); releases the request. */
int psx_server_result(psx_server* server, uint64_t id, psx_candidate* out, uint32_t cap, uint32_t* count);
int psx_server_stats(psx_server* server, psx_server_counters* stats);
/* Completes queued and running requests, then frees everything. */
int psx_server_destroy(psx_server* server);

/* ------------------------------------------------------------------ */
/* Global post-processing and output files                             */
/* ------------------------------------------------------------------ */
/* Arguments of the program-level distilling and scoring
 * (src/pipeline_multi.cu:406-407,547-555).  cfreq, foff and bw are in the
 * filterbank header's units (MHz), as CandidateScorer receives them. */
typedef struct psx_final_args {
  float freq_tol;  /* --freq_tol (DMDistiller, HarmonicDistiller) */
  float max_harm;  /* --max_harm (HarmonicDistiller) */
  float tsamp;     /* Filterbank::get_tsamp */
  float cfreq;     /* Filterbank::get_cfreq */
  float foff;      /* Filterbank::get_foff */
  float bw;        /* |foff| * nchans */
  int32_t limit;   /* --limit: candidates kept after scoring */
  int32_t reserved;
} psx_final_args;

/* Candidate fields set by CandidateScorer::score (include/transforms/
 * scorer.hpp:77-81) and the association counts overview.xml reports. */
typedef struct psx_cand_score {
  float ddm_count_ratio;
  float ddm_snr_ratio;
  int32_t is_adjacent;
  int32_t is_physical;
  int32_t nassoc;   /* Candidate::count_assoc */
  int32_t ndirect;  /* Candidate::assoc.size() */
} psx_cand_score;

/* main()'s post-processing of the candidates of every DM trial
 * (src/pipeline_multi.cu:547-555): DMDistiller(freq_tol, true), then
 * HarmonicDistiller(freq_tol, max_harm, true, false), then
 * CandidateScorer(tsamp, cfreq, foff, bw).score_all, then the first `limit`.
 * in: the concatenated forests of the searches (DM order).  out: the final
 * candidates as a forest (*count records).  scores (optional, scap entries):
 * one per final top-level candidate (*nfinal of them).  Host only. */
int psx_finalize(const psx_candidate* in, uint32_t nin, const psx_final_args* args, psx_candidate* out,
                 uint32_t cap, uint32_t* count, psx_cand_score* scores, uint32_t scap, uint32_t* nfinal);

/* A folded profile (nints x nbins floats, row-major) for candidates.peasoup. */
typedef struct psx_fold {
  int32_t nbins;
  int32_t nints;
  const float* data;
} psx_fold;

/* CandidateFileWriter::write_binary (include/utils/output_stats.hpp:261-298):
 * per top-level candidate of the forest, an optional "FOLD" record (the four
 * characters, int32 nbins, int32 nints, nbins*nints float32), then int32
 * ndets and the ndets = 1 + nassoc CandidatePOD records (24 bytes each) of the
 * candidate and its associates in pre-order.  Native byte order, as the
 * reference writes it.  folds (optional): one per top-level candidate
 * (nbins*nints == 0: no FOLD record).  offsets (optional): each candidate's
 * byte offset (CandidateFileWriter::byte_mapping).  Both arrays hold
 * ntop_cap entries; *ntop = top-level candidates written. */
int psx_write_candidates(const char* path, const psx_candidate* forest, uint32_t n, const psx_fold* folds,
                         int64_t* offsets, uint32_t ntop_cap, uint32_t* ntop);

/* ------------------------------------------------------------------ */
/* Upstream of the search: SIGPROC filterbanks and dedispersion        */
/* ------------------------------------------------------------------ */
/* SIGPROC header values (include/data_types/header.hpp:171-215). */
typedef struct psx_sigproc_header {
  char source_name[80];
  char rawdatafile[80];
  double az_start, za_start, src_raj, src_dej, tstart, tsamp, period, fch1, foff, refdm;
  int32_t nchans, telescope_id, machine_id, data_type, ibeam, nbeams, nbits, barycentric, pulsarcentric, nbins,
      nsamples, nifs, npuls;
  uint8_t signed_data;
  uint8_t reserved[3];
  uint32_t size; /* header bytes */
} psx_sigproc_header;

/* read_header (include/data_types/header.hpp:339-403): keyword/value pairs
 * between HEADER_START and HEADER_END; nsamples from the file size when the
 * header has none.  PSX_ERR_ARG if the file is not a SIGPROC filterbank. */
int psx_sigproc_read_header(const char* path, psx_sigproc_header* header);
/* SigprocFilterbank::load_gulp (include/data_types/filterbank.hpp:313-333):
 * the packed bytes of samples [start, start + nsamps) into out (host,
 * nsamps * nchans * nbits / 8 bytes); *nread = whole samples read. */
int psx_sigproc_read_samples(const char* path, const psx_sigproc_header* header, uint64_t start, uint64_t nsamps,
                             uint8_t* out, uint64_t* nread);
/* Filterbank::get_cfreq (filterbank.hpp:231-237), float arithmetic. */
float psx_sigproc_cfreq(const psx_sigproc_header* header);

/* Dedisperser::generate_dm_list (include/transforms/dedisperser.hpp:62-70)
 * -> libdedisp dedisp_generate_dm_list: DMs from dm_start while below dm_end,
 * each step keeping the smearing within tol of the pulse width (us).
 * tsamp, f0 (fch1), df (foff) in the float precision the plan holds them. */
int psx_dedisp_dm_list(float dm_start, float dm_end, float tsamp, float pulse_width_us, float f0, float df,
                       uint32_t nchans, float tol, float* out, uint32_t cap, uint32_t* count);
/* libdedisp's per-channel delay in samples per unit DM:
 * 4.15e3/tsamp * (1/(f0 + c df)^2 - 1/f0^2). */
int psx_dedisp_delay_table(uint32_t nchans, float tsamp, float f0, float df, float* table);
/* Samples lost at the end of a dedispersed series: the delay of the last
 * channel at the largest DM of the list, rounded to nearest. */
uint64_t psx_dedisp_max_delay(const float* dm_list, uint32_t ndm, const float* table, uint32_t nchans);
/* Device work bytes of psx_dedisperse. */
uint64_t psx_dedisp_work_bytes(uint64_t nsamps_in, uint32_t nchans, uint32_t ndm);
/* Dedisperser::dedisperse (dedisperser.hpp:105-186) -> libdedisp
 * dedisp_execute, direct algorithm, float output: for each DM of dm_list
 * (host) and t < nsamps_out = nsamps_in - max_delay,
 *   out[dm * out_stride + t] = sum over unmasked channels c of
 *       v(t + rn(dm * table[c]), c) / ((2^nbits - 1) * nchans)
 * where v is the unpacked nbits sample (nbits 1, 2, 4 or 8, lowest bits
 * first within a byte).  d_in: packed samples on the device; killmask (host,
 * may be NULL): 0 drops a channel. */
int psx_dedisperse(const uint8_t* d_in, uint64_t nsamps_in, uint32_t nchans, uint32_t nbits, const float* dm_list,
                   uint32_t ndm, const int32_t* killmask, float tsamp, float f0, float df, float* d_out,
                   uint64_t out_stride, void* d_work, psx_stream s);

#ifdef __cplusplus
}
#endif
#endif /* PSX_H */
