/*
 * oracle.h - CPU restatement of peasoup's per-DM-trial acceleration search.
 *
 * TEST INFRASTRUCTURE ONLY.  This library is the checker that the HIP
 * product is compared against.  Only tests/, __graft_entry__.smoke() and
 * bench.py's cpu_baseline leg may load it; the product library never links
 * or calls it.
 *
 * Every function cites the reference file:line whose behaviour it restates.
 * Expression types follow the reference exactly (float arithmetic with double
 * literals where the reference has them).  Host-side reference code is built
 * by g++ on x86-64 without FMA contraction, so this file is compiled with
 * -ffp-contract=off; device-side expressions that nvcc contracts by default
 * are written with an explicit fma() (see each function).
 *
 * Parity pins (how each function is checked, tests/test_oracle_*.py):
 *   orc_accel_list            golden: example_output/overview.xml:91-10850
 *   orc_resample1/2           closed form, a=0 identity, resample vs resampleII
 *                             on the i%451 pattern of src/resampling_test.cpp
 *   orc_fft_r2c / _c2r        numpy.fft.rfft/irfft, brute-force DFT, Parseval
 *   orc_form_*                closed forms on hand-made spectra
 *   orc_median_scrunch5       brute-force sort median; counts 1..4
 *   orc_linear_stretch        closed form (float index arithmetic restated)
 *   orc_calculate_median      composition of the two above + zone boundaries
 *   orc_stats / normalise     numpy float64 reductions
 *   orc_harmonic_sum          brute force; comb known answer
 *                             (src/harmonic_sum_test.cpp:21-27)
 *   orc_find_candidates       brute-force threshold scan + clustering cases
 *   orc_*_distill             oracle/_ref build of the reference distiller.hpp
 *   orc_zap_birdies           hand-computed bin ranges (exact binary fractions)
 *   orc_dm_list               golden: the DM trials of the reference's example
 *                             candidates.peasoup (tutorial.fil, 59 values)
 *   orc_delay_table           closed form (cold-plasma delay in samples)
 *   orc_dedisperse            known answers: a dispersed pulse lands in one
 *                             sample at its DM; hand-packed sub-byte samples
 *   orc_score / orc_finalize  oracle/_ref build of the reference scorer.hpp
 *                             and distiller.hpp chained as main() chains them;
 *                             hand-computed scorer cases
 */
#pragma once
#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

/* Flat candidate record: the CandidatePOD fields (include/data_types/candidates.hpp:10-17)
 * plus the number of associated candidates (Candidate::count_assoc, :44-51). */
typedef struct {
  float dm;
  int32_t dm_idx;
  float acc;
  int32_t nh;
  float snr;
  float freq;
  int32_t nassoc;
  int32_t _pad;
} orc_cand;

typedef struct {
  uint64_t size;          /* FFT length N */
  float tsamp;            /* seconds */
  float dm;
  int32_t dm_idx;
  float min_snr;          /* PeakFinder threshold */
  float min_freq;         /* Hz */
  float max_freq;         /* Hz */
  int32_t nharmonics;     /* number of harmonic folds F */
  float freq_tol;         /* distiller tolerance */
  float max_harm;         /* HarmonicDistiller max_harm */
  int32_t min_gap;        /* PeakFinder min_gap (30) */
  float boundary_5_freq;  /* Dereddener zone boundaries (0.05, 0.5) */
  float boundary_25_freq;
  int32_t nbirdies;       /* zap list (Zapper): 0 = none */
  const float* birdie_freq;
  const float* birdie_width;
  int32_t fast_fft;       /* 0: exact double DFT (the checker); 1: the float
                             four-step FFT of fastfft.cpp (CPU baseline timing) */
} orc_search_params;

/* include/utils/utils.hpp:143-233 */
int orc_accel_list(float acc_lo, float acc_hi, float tol, float pulse_width,
                   uint64_t nsamps, float tsamp, float cfreq, float ch_bw,
                   float dm, float cdm, float* out, int cap);
float orc_accel_step(float acc_lo, float acc_hi, float tol, float pulse_width,
                     uint64_t nsamps, float tsamp, float cfreq, float ch_bw,
                     float dm, float cdm);

/* src/kernels.cu:321-324,342-369 (resampleII) and :315-318,327-339,371-386 (resample) */
uint64_t orc_resample2_index(uint64_t i, uint64_t n, float acc, float tsamp);
uint64_t orc_resample2_index_nofma(uint64_t i, uint64_t n, float acc, float tsamp);
void orc_resample2(const float* in, float* out, uint64_t n, float acc, float tsamp);
void orc_resample1(const float* in, float* out, uint64_t n, float acc, float tsamp);

/* include/transforms/ffter.hpp:54-110 (cuFFT R2C / C2R, unnormalised) */
int orc_fft_r2c(const float* in, float* out, uint64_t n);
int orc_fft_r2c_f64(const double* in, double* out, uint64_t n);
int orc_fft_c2r(const float* in, float* out, uint64_t n);
int orc_fft_c2c_f64(double* data, uint64_t m, int sign);
/* the float four-step FFT of the CPU baseline (fastfft.cpp) */
int orc_fft_r2c_fast(const float* in, float* out, uint64_t n);
int orc_fft_c2r_fast(const float* in, float* out, uint64_t n);

/* src/kernels.cu:216-234 (way 0) and :238-259 (way 1) */
void orc_form_amplitude(const float* X, float* A, uint64_t nb);
void orc_form_interbin(const float* X, float* A, uint64_t nb);

/* src/kernels.cu:902-1038, include/transforms/dereddener.hpp:41-66 */
uint64_t orc_median_scrunch5(const float* in, uint64_t count, float* out);
void orc_linear_stretch(const float* in, uint64_t in_count, float* out, uint64_t out_count);
void orc_calculate_median(const float* A, uint64_t nb, double bin_width,
                          float b5, float b25, float* median);
void orc_divide_c_by_f(float* X, const float* f, uint64_t nb);
/* src/kernels.cu:1063-1084 (Zapper::zap, birdiezapper.hpp:60-71) */
void orc_zap_birdies(float* X, uint64_t nb, float bin_width, const float* freqs, const float* widths, int n);

/* include/utils/stats.hpp:6-40, src/kernels.cu:427-521 */
void orc_stats(const float* p, uint64_t n, uint64_t first, float* mean, float* rms, float* std_);
void orc_normalise(float* p, float mean, float sigma, uint64_t n);
void orc_remove_baseline(float* x, uint64_t n);

/* src/kernels.cu:34-100 */
void orc_harmonic_sum(const float* P, float* folds, uint64_t nb, unsigned nfolds);

/* src/kernels.cu:391-423 */
int orc_threshold(const float* p, int n, int start, float thr, int* idx, float* snr, int cap);
/* include/transforms/peakfinder.hpp:27-56,77-94 */
int orc_find_candidates(const float* p, uint32_t nbins, double bin_width, uint32_t nh,
                        float thr, float fmin, float fmax, int min_gap,
                        int* peak_idx, float* peak_snr, float* peak_freq, int cap);

/* include/transforms/distiller.hpp:26-196 */
int orc_harmonic_distill(const orc_cand* in, int n, float tol, float max_harm,
                         int keep_related, int fractional, orc_cand* out);
int orc_accel_distill(const orc_cand* in, int n, float tobs, float tol,
                      int keep_related, orc_cand* out);
int orc_dm_distill(const orc_cand* in, int n, float tol, int keep_related, orc_cand* out);

/* Worker prologue, src/pipeline_multi.cu:170-238.  series holds n_real
 * samples; tim_out receives the N-sample dereddened, inverse-transformed
 * series (xN scaled).  mean/rms/std are the interbinned-spectrum stats. */
int orc_prologue(const float* series, uint64_t n_real, const orc_search_params* p,
                 float* tim_out, float* mean, float* rms, float* std_);

/* One acceleration trial, src/pipeline_multi.cu:246-289.  Returns the
 * number of harmonic-distilled candidates written to out (<= cap), or a
 * negative value on error.  raw_* (optional, may be NULL) receive the
 * per-spectrum peak lists before distilling: raw_nh/raw_idx/raw_snr. */
int orc_trial(const float* tim, const orc_search_params* p, float acc, float mean, float std_,
              orc_cand* out, int cap,
              int* raw_nh, int* raw_idx, float* raw_snr, int raw_cap, int* raw_count);

/* Whole DM trial (prologue + accel loop + acceleration distill) on nthreads
 * host threads (one accel trial per thread).  timings[0] = prologue s,
 * timings[1] = accel loop s, timings[2] = distill s (may be NULL).
 * Output (also of orc_search_accels) is a forest: each distilled candidate
 * followed by its associates in pre-order, nassoc = subtree records after it;
 * returns the record count. */
int orc_search_dm(const float* series, uint64_t n_real, const orc_search_params* p,
                  const float* accs, int nacc, orc_cand* out, int cap, int nthreads,
                  double* timings);

/* Accel-loop only (given a prologue result), used by the CPU baseline. */
int orc_search_accels(const float* tim, const orc_search_params* p, const float* accs, int nacc,
                      float mean, float std_, orc_cand* out, int cap, int nthreads);

/* CandidateScorer::score (include/transforms/scorer.hpp:25-81) of every
 * top-level candidate of a forest: ratios[2t..2t+1] = ddm_count_ratio,
 * ddm_snr_ratio; flags[2t..2t+1] = is_adjacent, is_physical.  Returns the
 * number of top-level candidates, -1 if the forest is malformed. */
int orc_score(const orc_cand* forest, int n, float cfreq, float foff, float bw, float* ratios, int* flags);

/* main()'s post-processing, src/pipeline_multi.cu:406-407,547-558: DM distill
 * (keep related), harmonic distill (keep related, no fractional harmonics),
 * score, keep the first `limit`.  in/out are forests; returns the output
 * record count (written up to cap), *nfinal = top-level candidates. */
int orc_finalize(const orc_cand* in, int n, float freq_tol, float max_harm, float cfreq, float foff, float bw,
                 int limit, orc_cand* out, int cap, float* ratios, int* flags, int* nfinal);

/* libdedisp (through include/transforms/dedisperser.hpp:62-186):
 * dedisp_generate_dm_list recurrence (returns the count, writes up to cap),
 * the per-channel delay table, and direct dedispersion with float output
 * (returns the output length nsamps - max delay, or -1). */
int orc_dm_list(float dm_start, float dm_end, float dt, float ti, float f0, float df, int nchans, float tol,
                float* out, int cap);
void orc_delay_table(int nchans, float dt, float f0, float df, float* table);
long orc_dedisperse(const uint8_t* in, long nsamps, int nchans, int nbits, const float* dms, int ndm,
                    const int* killmask, float dt, float f0, float df, float* out);

int orc_hardware_threads(void);

#ifdef __cplusplus
}
#endif
