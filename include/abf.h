/*
 * abf.h -- C ABI of the gfx950 (MI355X) frequency-domain beamforming library libabf.so.
 *
 * The library replaces the numba kernels of Acoular's frequency-domain beamforming path.
 * Each entry point names the reference interface it stands in for (paths relative to the
 * Acoular source tree).
 *
 * Conventions
 *  - Every pointer argument is DEVICE memory owned by the caller, unless marked "host".
 *    The library never allocates device memory: entry points that need scratch space take a
 *    caller-provided workspace whose size the matching *_workspace_size() call returns.
 *  - Complex arrays are interleaved (re, im) pairs, C-contiguous, as numpy/torch complex.
 *  - Every call is ordered on the given HIP stream (a hipStream_t passed as void*; NULL is
 *    the default stream) and returns without waiting for the device.
 *  - Return value: ABF_OK (0) or a negative abf_status.  abf_last_error() returns a
 *    description of the last failure on the calling thread.
 *  - Calls on different streams are independent; a workspace must not be shared by calls
 *    that may run at the same time.
 */
#ifndef ABF_H
#define ABF_H

#include <stddef.h>
#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

enum abf_status {
    ABF_OK = 0,
    ABF_ERR_INVALID = -1,     /* bad argument: maps to ValueError/KeyError/TypeError */
    ABF_ERR_UNSUPPORTED = -2, /* valid request this build does not implement */
    ABF_ERR_HIP = -3,         /* a HIP runtime call failed */
    ABF_ERR_WORKSPACE = -4    /* workspace missing or too small */
};

/* steer_type values: the integer codes of fastFuncs.beamformerFreq (fastFuncs.py:161) */
enum abf_steer {
    ABF_STEER_CLASSIC = 1,
    ABF_STEER_INVERSE = 2,
    ABF_STEER_TRUE_LEVEL = 3,
    ABF_STEER_TRUE_LOCATION = 4
};

/* Arithmetic form of the grid beamformer. */
enum abf_form {
    ABF_FORM_PLAIN_F32 = 0,   /* FP32 VALU, one grid point per lane, CSM staged per row block */
    ABF_FORM_MFMA_BF16X3 = 1, /* complex GEMM on bf16 MFMA, operands split hi+lo (3 products) */
    ABF_FORM_MFMA_F32 = 2,    /* complex GEMM on fp32 MFMA */
    ABF_FORM_MFMA_F16X2 = 3   /* complex GEMM on fp16 MFMA: U rounded once (scaled by a power of two per
                                 frequency), steering split hi+lo (2 products per real product) */
};

/* Phase handling: REFERENCE rounds k*r to float32 before cos/sin like fastFuncs.py:260;
 * FP64 keeps the float64 phase.  Both reduce the phase to [-pi, pi] in float64. */
enum abf_phase {
    ABF_PHASE_REFERENCE = 0,
    ABF_PHASE_FP64 = 1,
    /* flag OR-ed into a phase mode: take cos/sin with the hardware transcendental unit on the
     * float64-reduced argument (revolutions) instead of the library sincosf */
    ABF_PHASE_FAST_TRIG = 0x100
};

/* MFMA kernel variants (tile_kernel of abf_grid_job). */
enum abf_mfma_kernel {
    ABF_MFMA_KERNEL_AUTO = 0,
    ABF_MFMA_KERNEL_ALTERNATING = 1, /* all waves generate the steering matrix, then all run MFMA */
    ABF_MFMA_KERNEL_SPECIALISED = 2, /* steering waves beside MFMA waves, double-buffered */
    ABF_MFMA_KERNEL_ROW_PAIRS = 3,   /* alternating, each work unit covers both 16-row halves of a
                                        32-row block (half the LDS reads per MFMA) */
    ABF_MFMA_KERNEL_PIPELINED = 4,   /* alternating, X and U fragments of the next block loaded
                                        before the MFMAs of the current one (split-bf16) */
    ABF_MFMA_KERNEL_PIPELINED_ROW_PAIRS = 5, /* both of the above, one column tile per fragment */
    ABF_MFMA_KERNEL_STREAMED = 6, /* steering streamed through LDS one 32-microphone block at a time,
                                     its generation overlapped with the MFMAs of the previous block
                                     (split-bf16, 33..256 mics, reference phase with hardware sin/cos;
                                     other jobs fall back to ROW_PAIRS) */
    ABF_MFMA_KERNEL_STREAMED_WIDE = 7, /* STREAMED over 32 grid points, the row blocks split
                                          between two wave groups by parity (half the accumulator
                                          and distance registers per wave): 33..512 mics; same
                                          other cases and fallback */
    ABF_MFMA_KERNEL_STREAMED_PERSISTENT = 8 /* STREAMED as a persistent grid (one workgroup per CU)
                                               walking the (job, tile, band) items in a fixed
                                               round-robin order; same cases and fallback */
};

/* Steering geometry shared by the grid entry points. r0 = ref_dist if ref_dist > 0, else
 * |g - ref| (SteeringVector.r0, fbeamform.py:187-193). */
typedef struct abf_geometry {
    const double *gpos; /* [3][num_grid] grid positions, SoA (Grid.pos, grids.py:351-359) */
    const double *mpos; /* [3][num_mics] microphone positions (MicGeom.pos, microphones.py:175-180) */
    int64_t num_grid;
    int32_t num_mics;
    int32_t pad_;
    double ref[3];   /* host values: reference point */
    double ref_dist; /* host value: constant reference distance if > 0 */
} abf_geometry;

/* One beamforming job: the map B[f][g] for every frequency slice of a resident CSM.
 * Replaces the loop of BeamformerBase._calc (fbeamform.py:475-510) around
 * fastFuncs.beamformerFreq -> _freqBeamformer_FullCSM (fastFuncs.py:61-214, 238-311). */
typedef struct abf_grid_job {
    abf_geometry geom;
    const void *csm;          /* [num_freqs][num_mics][num_mics] complex64; upper triangle + diagonal read */
    const double *wavenumber; /* [num_freqs] k = 2*pi*f/c */
    int32_t num_freqs;
    int32_t steer_type;       /* abf_steer */
    int32_t r_diag;           /* remove the CSM diagonal */
    int32_t clip_negative;    /* set negative results to 0 (fbeamform.py:506-508) */
    double norm_factor;       /* sig_loss_norm() (fbeamform.py:438-451) */
    int32_t form;             /* abf_form */
    int32_t phase_mode;       /* abf_phase */
    float *out;               /* [num_freqs][num_grid] result */
    float *out_norm;          /* [num_freqs][num_grid] normalizeSteer, may be NULL */
    int32_t tile_cols;        /* MFMA forms: grid points per workgroup (16..256, power of two); 0 = auto */
    int32_t tile_freqs;       /* MFMA forms: frequencies per workgroup (at most 64 for the
                                 streamed kernels); 0 = auto */
    int32_t tile_kernel;      /* MFMA forms: enum abf_mfma_kernel; 0 = auto */
    int32_t pad1_;
} abf_grid_job;

/* Bytes of workspace abf_beamform needs for this job (the re-laid-out CSM). */
size_t abf_beamform_workspace_size(const abf_grid_job *job);

/* The launch shape abf_beamform uses for this job after its tile requests are clamped to what
 * the kernels support: kernel variant, grid points and frequencies per workgroup, workgroups
 * (plain form: 0, 64, 1, blocks).  No counterpart in the reference: it serves the autotuner. */
int abf_beamform_plan(const abf_grid_job *job, int32_t *tile_kernel, int32_t *tile_cols, int32_t *tile_freqs,
                      int64_t *workgroups);

/* Beamform a job.  fastFuncs.beamformerFreq (fastFuncs.py:61-214) for steer types 1..4,
 * batched over every frequency slice of the CSM. */
int abf_beamform(const abf_grid_job *job, void *workspace, size_t workspace_bytes, void *stream);

/* The two halves of abf_beamform: _prepare re-lays the CSM out into the workspace (depends
 * only on csm, num_mics, num_freqs, r_diag and form); _run beamforms from a prepared workspace,
 * so one prepared CSM serves any number of grids, steering types and normalisations. */
int abf_beamform_prepare(const abf_grid_job *job, void *workspace, size_t workspace_bytes, void *stream);
int abf_beamform_run(const abf_grid_job *job, void *workspace, size_t workspace_bytes, void *stream);

/* abf_beamform without a workspace argument and without any device allocation, for callers that
 * cannot allocate (code under HIP graph capture, tight serving loops, harnesses that forbid
 * hipMalloc).  No counterpart in the reference.  The workspace is a static device array of
 * ABF_NOALLOC_SCRATCH_BYTES per GPU, held until the process ends; the job's bins are processed in
 * chunks that fit it, so any number of bins works, and a job one of whose bins alone needs more
 * (e.g. the plain form above ~2800 microphones) is rejected with ABF_ERR_WORKSPACE.  Every call on
 * a GPU shares that GPU's array: calls that may run at the same time on one GPU (other streams,
 * other threads) race and must be serialised by the caller. */
#define ABF_NOALLOC_SCRATCH_BYTES (64u << 20)
int abf_beamform_noalloc(const abf_grid_job *job, void *stream);

/* A batch of independent jobs (different mic counts, grids and bands) beamformed together:
 * one re-layout launch and one beamform launch per distinct (form, padded mic count),
 * work items of every job interleaved.  Each job's workspace is carved from 'workspace' at
 * the offsets abf_beamform_batch_workspace_size reports through 'offsets' (host, may be NULL). */
size_t abf_beamform_batch_workspace_size(const abf_grid_job *jobs, int32_t num_jobs, size_t *offsets);
int abf_beamform_batch(const abf_grid_job *jobs, int32_t num_jobs, void *workspace, size_t workspace_bytes,
                       void *stream);

/* _freqBeamformer_FullCSM (fastFuncs.py:238-311) with the reference's inputs: r0 [G],
 * rm [G][M] distance table and one wavenumber; plain FP32 kernel.  out/out_norm [G]. */
int abf_beamform_table(const void *csm /* [M][M] complex64 */, const double *r0, const double *rm, int64_t num_grid,
                       int32_t num_mics, double wavenumber, int32_t steer_type, int32_t r_diag, double norm_factor,
                       int32_t phase_mode, float *out, float *out_norm, void *workspace, size_t workspace_bytes,
                       void *stream);
size_t abf_beamform_table_workspace_size(int32_t num_mics);

/* _freqBeamformer_EigValues (fastFuncs.py:335-411): eigval [K] (double), eigvec [M][K]
 * complex64, r0/rm tables. */
int abf_beamform_eig(const double *eigval, const void *eigvec, int32_t num_eig, const double *r0, const double *rm,
                     int64_t num_grid, int32_t num_mics, double wavenumber, int32_t steer_type, int32_t r_diag,
                     double norm_factor, int32_t phase_mode, float *out, float *out_norm, void *workspace,
                     size_t workspace_bytes, void *stream);
size_t abf_beamform_eig_workspace_size(int32_t num_mics, int32_t num_eig);

/* _freqBeamformer_FullCSM (fastFuncs.py:238-311) for every frequency of a complex128 CSM
 * [F][M][M] in float64 arithmetic, as the reference computes it: float64 steering vectors whose
 * phase argument is rounded to float32 and whose cos/sin are rounded to float32
 * (fastFuncs.py:259-261, phase_mode REFERENCE), float64 sums and float64 results out/out_norm
 * [F][G].  For callers that amplify the rounding of the map (DAMAS, fbeamform.py:1191-1206).
 * The workspace holds the re-laid-out CSM. */
size_t abf_beamform_f64_workspace_size(int32_t num_mics, int32_t num_freqs);
int abf_beamform_f64(const abf_geometry *geom, const void *csm, const double *wavenumber, int32_t num_freqs,
                     int32_t steer_type, int32_t r_diag, double norm_factor, int32_t phase_mode, int32_t clip_negative,
                     double *out, double *out_norm, void *workspace, size_t workspace_bytes, void *stream);

/* The eigenvalue form for every frequency of a resident eigendecomposition, steering vectors
 * generated from coordinates: the loops of BeamformerEig._calc (fbeamform.py:852-888) and
 * BeamformerMusic._calc (fbeamform.py:911-945) / BeamformerFunctional._calc (fbeamform.py:712-774)
 * around _freqBeamformer_EigValues (fastFuncs.py:335-411).  eigval [F][M] (double), eigvec
 * [F][M][M] complex64 with eigenvector e in column e (numpy.linalg.eigh layout of
 * PowerSpectra.eigvecs, spectra.py:351-364); eigenvectors eig_first .. eig_first+num_eig-1 are
 * summed.  out/out_norm [F][G]; negative results are set to 0 when clip_negative. */
int abf_beamform_eig_grid(const abf_geometry *geom, const double *wavenumber, int32_t num_freqs, const double *eigval,
                          const void *eigvec, int32_t eig_first, int32_t num_eig, int32_t steer_type, int32_t r_diag,
                          double norm_factor, int32_t phase_mode, int32_t clip_negative, float *out, float *out_norm,
                          void *stream);

/* Eigendecomposition of every CSM slice (PowerSpectra._get__evd, spectra.py:351-364, which calls
 * numpy.linalg.eigh per frequency): csm [F][M][M] complex64 (is_c128 = 0) or complex128, read
 * through its lower triangle as numpy does; eigval [F][M] double in ascending order; eigvec
 * [F][M][M] of the CSM's type with eigenvector e in column e (must not overlap csm); info [F]
 * (may be NULL) receives the solver's per-slice status, 0 on success.  One batched call of the
 * rocSOLVER divide-and-conquer solver, whose scratch memory is carved from the workspace. */
size_t abf_csm_eig_workspace_size(int32_t num_freqs, int32_t num_mics, int32_t is_c128);
int abf_csm_eig(const void *csm, int32_t num_freqs, int32_t num_mics, int32_t is_c128, double *eigval, void *eigvec,
                int32_t *info, void *workspace, size_t workspace_bytes, void *stream);

/* Point spread functions (calcPointSpreadFunction, fastFuncs.py:563-801, called by
 * PointSpreadFunction._psf_call, fbeamform.py:1096-1124): for every frequency f and every
 * source grid point s = sources[c] (device int64 [S], indices into the grid),
 * out[f][g][c] = |h_g^H a_s|^2 in the normalisation of steer_type, float64 (out_f64) or float32,
 * with the float32-rounded phase difference of the reference (phase_mode REFERENCE).  An index
 * outside [0, num_grid) yields a NaN column. */
int abf_psf(const abf_geometry *geom, const double *wavenumber, int32_t num_freqs, const int64_t *sources,
            int64_t num_sources, int32_t steer_type, int32_t phase_mode, int32_t out_f64, void *out, void *stream);

/* DAMAS (damasSolverGaussSeidel, fastFuncs.py:820-855, called by BeamformerDamas._calc,
 * fbeamform.py:1167-1208): n_iter relaxed Gauss-Seidel sweeps with the positivity clip on every
 * frequency's system psf[f] (float64 if psf_f64, else float32; [F][G][G], row g = grid point)
 * x[f] = dirty[f]; x [F][G] holds the start values on entry and the solutions on return.
 * num_grid <= 15360 (the solution vector lives in LDS). */
int abf_damas(const void *psf, int32_t psf_f64, const double *dirty, double *x, int64_t num_grid, int32_t num_freqs,
              int32_t n_iter, double relax, void *stream);

/* Microphone signals of a stationary point source (PointSource.result, sources.py:774-842, and
 * _fill_mic_signal_block, sources.py:77-91): out[b][m] = signal[int(0.5 + ind_m(b))] / rm[m] with
 * ind_m(0) = ind0[m] and ind_m(b+1) = ind_m(b) + up accumulated in float64 sample by sample as the
 * reference does; negative indices wrap as in numpy (or give 0 where ind < 0 with
 * zeros_prepadding); an index outside [-signal_len, signal_len) yields NaN.  signal is the
 * upsampled source signal (float64), out [num_samples][num_mics] float64. */
size_t abf_point_source_workspace_size(int32_t num_mics, int64_t num_samples);
int abf_point_source(const double *signal, int64_t signal_len, const double *ind0, const double *rm, int32_t num_mics,
                     int32_t up, int32_t zeros_prepadding, int64_t num_samples, double *out, void *workspace,
                     size_t workspace_bytes, void *stream);

/* Independent white noise per channel (UncorrelatedNoiseSource.result, sources.py:1911-1967, with
 * WNoiseGenerator.signal, signals.py:255-269): channel c is rms * numpy's legacy
 * RandomState(seeds[c]).standard_normal(num_samples) (MT19937 and the polar method), written to
 * out [num_samples][num_channels] float64; identical to numpy's values except for about 1 % that
 * differ by one ulp (the last bit of the float64 log). */
int abf_white_noise(const uint32_t *seeds, int32_t num_channels, int64_t num_samples, double rms, double *out,
                    void *stream);

/* One step of SourceMixer.result (sources.py:2139-2185): acc = weight * src (first != 0) or
 * acc += weight * src, n float64 values, each product and sum rounded separately. */
int abf_mix(const double *src, double weight, int64_t n, int32_t first, double *acc, void *stream);

/* CLEAN-SC (BeamformerCleansc._calc, fbeamform.py:1449-1492) for every frequency at once: csm
 * [F][M][M] complex128 is a working copy (the removed coherent components are subtracted from
 * it), h [F][G] holds the dirty maps (abf_beamform_f64 with the same steering, r_diag and
 * norm_factor) and is updated in place; result [F][G] float32 receives the CLEAN-SC maps.
 * n_iter 0 means 2 * num_mics iterations; damp and stopn as the reference's traits. */
size_t abf_cleansc_workspace_size(int32_t num_mics, int64_t num_grid, int32_t num_freqs, int32_t n_iter);
int abf_cleansc(const abf_geometry *geom, const double *wavenumber, int32_t num_freqs, void *csm, double *h,
                int32_t steer_type, int32_t r_diag, double norm_factor, int32_t phase_mode, int32_t n_iter,
                double damp, int32_t stopn, float *result, void *workspace, size_t workspace_bytes, void *stream);

/* CLEAN (BeamformerClean._calc, fbeamform.py:1536-1598) for every frequency at once: dirty [F][G]
 * float64 holds the dirty maps (abf_beamform_f64, clipped at zero when r_diag) and is updated in
 * place; clean [F][G] float64 receives the clean maps.  The PSF columns of the peaks are formed
 * per iteration as calcPointSpreadFunction (fastFuncs.py:563-801) forms them, rounded to float32
 * when psf_f32 (psf_precision 'float32').  At least one iteration runs, at most max(n_iter, 1). */
size_t abf_clean_workspace_size(int32_t num_mics, int64_t num_grid, int32_t num_freqs);
int abf_clean(const abf_geometry *geom, const double *wavenumber, int32_t num_freqs, double *dirty,
              int32_t steer_type, int32_t phase_mode, int32_t psf_f32, int32_t n_iter, double damp, double *clean,
              void *workspace, size_t workspace_bytes, void *stream);

/* _freqBeamformer_SpecificSteerVec_FullCSM / _CsmRemovedDiag (fastFuncs.py:414-476): explicit
 * steering vectors steer [G][M] complex64. out = q * norm, out_norm = sum |h|^2. */
int abf_beamform_custom(const void *csm, const void *steer, int64_t num_grid, int32_t num_mics, int32_t r_diag,
                        double norm_factor, float *out, float *out_norm, void *workspace, size_t workspace_bytes,
                        void *stream);
/* _freqBeamformer_EigValProb_SpecificSteerVec_* (fastFuncs.py:479-559). */
int abf_beamform_eig_custom(const double *eigval, const void *eigvec, int32_t num_eig, const void *steer,
                            int64_t num_grid, int32_t num_mics, int32_t r_diag, double norm_factor, float *out,
                            float *out_norm, void *workspace, size_t workspace_bytes, void *stream);

/* dist_mat (environments.py:65-103): rm[n][m] = |gpos[:,n] - mpos[:,m]|, float64. */
int abf_dist_mat(const double *gpos, const double *mpos, int64_t num_grid, int32_t num_mics, double *rm, void *stream);

/* calcTransfer / _transferCoreFunc (fastFuncs.py:859-897):
 * out[g][m] = exp(-1j*f32(k*(rm-r0))) * r0/rm, complex128 (out_c128 != 0) or complex64.
 * steer_mode 0 returns the transfer function; 1..4 apply the steer_vector normalisation of
 * that steer type (SteeringVector._steer_funcs_freq, fbeamform.py:145-153). */
int abf_transfer(const double *r0, const double *rm, int64_t num_grid, int32_t num_mics, double wavenumber,
                 int32_t phase_mode, int32_t steer_mode, int32_t out_c128, void *out, void *stream);

/* calcCSM (fastFuncs.py:19-58): csm[f][r][c] += conj(spec[f][c]) * spec[f][r] for r <= c.
 * is_c128 selects complex128 (FP64 arithmetic) or complex64. */
int abf_calc_csm(void *csm, const void *spec, int32_t num_freqs, int32_t num_mics, int32_t is_c128, void *stream);

/* Welch CSM: PowerSpectra.calc_csm (spectra.py:366-416) over the block generator
 * _get_source_data (spectra.py:171-183).  time [num_samples][num_mics] float32, window [bs]
 * (host values copied by the caller to device), blocks at hop*b for b < num_blocks.
 * csm_out [bs/2+1][M][M] complex64 = scale * sum_b X_b X_b^H, Hermitian completed.
 * scale = 2/(bs * sum w^2 * num_blocks_float) is computed by the caller. */
size_t abf_welch_csm_workspace_size(int32_t num_mics, int32_t block_size, int32_t blocks_per_pass);
int abf_welch_csm(const float *time, int64_t num_samples, int32_t num_mics, int32_t block_size, int32_t hop,
                  int32_t num_blocks, const float *window, double scale, void *csm_out, int32_t blocks_per_pass,
                  void *workspace, size_t workspace_bytes, void *stream);

/* RFFT.result (fprocess.py:35-180): one-sided spectra of the blocks at hop*b, b < num_blocks, of
 * time [num_samples][num_mics] float32 after multiplying by window [bs]:
 * out[b][k][m] = bin_scale[k] * rfft(window * x_b)[k], complex64, k < bs/2+1 (the reference's
 * frequency-major, channel-interlaced row).  bin_scale (device, [bs/2+1]) carries the scaling
 * the caller chose ('none', 'energy', 'amplitude'); block_size a power of two <= 16384. */
size_t abf_rfft_workspace_size(int32_t block_size);
int abf_rfft_blocks(const float *time, int64_t num_samples, int32_t num_mics, int32_t block_size, int32_t hop,
                    int32_t num_blocks, const float *window, const float *bin_scale, void *out, void *workspace,
                    size_t workspace_bytes, void *stream);

/* IRFFT.result (fprocess.py:182-283): the real inverse FFT of every snapshot spectrum
 * spec [num_blocks][bs/2+1][num_mics] complex64 (the imaginary parts of the DC and Nyquist bins
 * ignored, as a real inverse transform ignores them), out [num_blocks * bs][num_mics] float32,
 * time-major; block_size a power of two <= 16384. */
size_t abf_irfft_workspace_size(int32_t block_size);
int abf_irfft_blocks(const void *spec, int32_t num_blocks, int32_t num_mics, int32_t block_size, float *out,
                     void *workspace, size_t workspace_bytes, void *stream);

/* AutoPowerSpectra.result (fprocess.py:286-366): out[i] = |spec[i]|^2 * scale, n complex64 values
 * to float32. */
int abf_auto_power(const void *spec, int64_t n, double scale, float *out, void *stream);

/* CrossPowerSpectra.result with calc_mode 'full' (fprocess.py:369-470): for every snapshot s and
 * bin k of spec [num_snapshots][num_freqs][num_mics] complex64,
 * out[s][k][r][c] = scale * spec[s][k][r] * conj(spec[s][k][c]) (complex64). */
int abf_cross_power(const void *spec, int64_t num_snapshots, int32_t num_freqs, int32_t num_mics, double scale,
                    void *out, void *stream);

/* SignalGenerator.usignal (signals.py:93-126, scipy.signal.resample of a real signal): y [num]
 * float64 = the n samples of x resampled to num >= n samples through their one-sided spectrum
 * (zero-padded, the Nyquist bin of an even n halved when num > n) and scaled by num / n.  The
 * workspace query plans the transforms (rocFFT, double precision) and needs the device. */
size_t abf_resample_workspace_size(int64_t num_samples, int64_t num);
int abf_resample(const double *x, int64_t num_samples, int64_t num, double *y, void *workspace,
                 size_t workspace_bytes, void *stream);

const char *abf_last_error(void);
const char *abf_version(void);

#ifdef __cplusplus
}
#endif
#endif /* ABF_H */
