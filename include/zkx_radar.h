/*
 * zkx_radar.h -- C ABI of the gfx950 radar library (libzkx_radar.so).
 *
 * Every entry point takes plain device pointers, sizes and POD options, runs
 * on the given HIP stream (NULL = the default stream, as the reference used)
 * and returns 0 or one of the reference's public ErrorCode numbers
 * (ZKX/cusignal_cpp/include/cusignal/runtime/errors.h:18-41).  No torch or
 * HIP C++ types appear in the signatures; a stream is passed as void*.
 *
 * Complex data is interleaved {float re, im}, the layout of the reference's
 * ComplexFloat / TypedComplex<float> (fft_interface.h:104-117,
 * filtering_typed.h:13).  Arrays are row-major.
 *
 * Each declaration names the reference interface it replaces (paths relative
 * to the reference repository).  Validation follows the reference's own
 * checks and returns ZKX_INVALID_SHAPE / ZKX_INVALID_ARGUMENT where the
 * reference throws std::invalid_argument; the reason is readable through
 * zkx_last_error_message().
 */
#ifndef ZKX_RADAR_H
#define ZKX_RADAR_H

#include <stddef.h>
#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

/* Return codes: the reference's ErrorCode values. */
#define ZKX_SUCCESS 0
#define ZKX_INVALID_ARGUMENT 1001
#define ZKX_INVALID_SHAPE 1002
#define ZKX_INVALID_DTYPE 1003
#define ZKX_UNSUPPORTED_OPERATION 1004
#define ZKX_RUNTIME_ERROR 2002
#define ZKX_OUT_OF_MEMORY 2003
#define ZKX_BACKEND_ERROR 2099
#define ZKX_INTERNAL_ERROR 9001

/* RadarWindowKind (radartools_typed.h:59-65), same numbering. */
#define ZKX_WINDOW_NONE 0
#define ZKX_WINDOW_HANN 1
#define ZKX_WINDOW_HAMMING 2
#define ZKX_WINDOW_EXPLICIT_FP32 3
#define ZKX_WINDOW_EXPLICIT_FP64 4

typedef struct {
    float re;
    float im;
} zkx_complex64;

typedef struct {
    double re;
    double im;
} zkx_complex128;

/* ---------------------------------------------------------------------- */
/* library                                                                  */
/* ---------------------------------------------------------------------- */

/* Message of the last failure on the calling thread ("" if none). */
const char* zkx_last_error_message(void);
/* Library version, major*10000 + minor*100 + patch. */
int zkx_version(void);
/* Largest power-of-two transform length one LDS-resident pass supports. */
int zkx_max_lds_fft_length(void);
/* Kernel variant used for an operation ("pulse_compression", ...), as chosen
 * by the autotuner; variants compute identical results.  No counterpart in
 * the reference (its launches have one fixed shape).  "dwell_fused" selects
 * how zkx_dwell_batch_c64 runs cfar_mode 1 for R = 4096, P = 32..128 dwells:
 * 0 = one persistent launch (pulse compression jobs and strip Doppler+CFAR
 * jobs from one queue; "dwell_fused_rows", "dwell_fused_chunks" and
 * "dwell_fused_lead" set its job sizes and lead), 1 = two launches (the
 * default, faster on this GPU).  "fir_fft" selects the method for long
 * filters on contiguous lines (upfirdn, resample_poly, firfilter): 2 (the
 * default) 4096-point overlap-save segments with the spectrum folded for
 * decimation by D in {1, 2, 4, 8} when up = 1, and for rate changes with
 * up/g in {2, 3, 4}, down/g in {1, 2, 4} and 64 or more taps per phase
 * (g = gcd(up, down)), otherwise as 1; 1 one transform per polyphase phase
 * for up = 1, the direct kernels otherwise; 0 the direct polyphase kernels;
 * these agree to FP32 rounding, not bit for bit.  One more name is a
 * test switch, not a variant: "inject_fft_failure" = 1 makes the transform
 * stage of pulse_compression / pulse_doppler return ZKX_BACKEND_ERROR
 * before any device work (the reference's synthesized backend status for
 * its error-path tests), 0 restores normal operation. */
int zkx_set_kernel_variant(const char* op, int variant);
int zkx_get_kernel_variant(const char* op);

/* ---------------------------------------------------------------------- */
/* radar operators                                                          */
/* ---------------------------------------------------------------------- */

/*
 * Pulse compression, complex64 in and out.
 * Replaces pulse_compression_device<float> (radartools_typed.h:237-262,
 * radartools_typed.cpp:677-712, radartools_typed.cu:184-223).
 *   x        [num_pulses, samples_per_pulse]
 *   tmpl     [template_len]
 *   out      [num_pulses, nfft]   (nfft = -1 means samples_per_pulse)
 *   y[p,k] = sum_{t<min(T,nfft)} x[p,(k+t) mod nfft] * conj(h~[t]),
 *   h~ = w*h (symmetric window over template_len), divided by ||w*h||_2 when
 *   normalize != 0.  explicit_window: device float[template_len] for the
 *   explicit kinds, else NULL.  nfft must be a power of two no larger than
 *   zkx_max_lds_fft_length().
 * The spectrum of the template is rebuilt on every call; use
 * zkx_pc_template_prepare + zkx_pulse_compression_prepared_c64 to keep it
 * resident instead.
 */
int zkx_pulse_compression_c64(const zkx_complex64* x, int num_pulses, int samples_per_pulse,
                              const zkx_complex64* tmpl, int template_len, int nfft,
                              int normalize, int window, const float* explicit_window,
                              zkx_complex64* out, void* stream);

/* Template spectrum H = FFT_nfft(h~) into spectrum[nfft] (device). */
int zkx_pc_template_prepare(const zkx_complex64* tmpl, int template_len, int nfft,
                            int normalize, int window, const float* explicit_window,
                            zkx_complex64* spectrum, void* stream);

/* Pulse compression with a prepared spectrum: y = IFFT(FFT(x_pad) * conj(H)).
 * batch independent [num_pulses, samples_per_pulse] cubes are laid out back
 * to back (dwell stride num_pulses*samples_per_pulse in, num_pulses*nfft out). */
int zkx_pulse_compression_prepared_c64(const zkx_complex64* x, int num_pulses,
                                       int samples_per_pulse, const zkx_complex64* spectrum,
                                       int nfft, zkx_complex64* out, void* stream);

/*
 * Pulse-Doppler processing, complex64 in and out (FP32 compute).
 * Replaces the preallocated pulse_doppler_device<float> overload
 * (radartools_typed.h:304-311, radartools_typed.cpp:761-789,
 * radartools_typed.cu:225-244).
 *   x   [num_pulses, samples_per_pulse];  out [nfft, samples_per_pulse]
 *   Y[f,r] = sum_{p<min(P,nfft)} w[p] x[p,r] exp(-2 pi i f p / nfft),
 *   w symmetric over num_pulses; no fftshift.  nfft = -1 means num_pulses.
 *   power (may be NULL): float [nfft, samples_per_pulse] receives |Y|^2, the
 *   Task1 step-4 power map (Task1/pipeline/src/steps/step4/step4.cu:11-19).
 *   out may be NULL when only the power map is wanted.
 */
int zkx_pulse_doppler_c64(const zkx_complex64* x, int num_pulses, int samples_per_pulse,
                          int nfft, int window, const float* explicit_window,
                          zkx_complex64* out, float* power, void* stream);

/*
 * CFAR threshold multiplier alpha = N * (pfa^(-1/N) - 1) in FP64 on the host.
 * Replaces cfar_alpha_typed_cpu / cfar_alpha_device (radartools_typed.h:207-222,
 * radartools_typed.cpp:592-617).
 */
int zkx_cfar_alpha(double pfa, int reference, double* alpha);

/*
 * Cell-averaging CFAR on a float32 map of rank 1 or 2.
 * Replaces ca_cfar_device<float> (radartools_typed.h:154-195,
 * radartools_typed.cpp:557-591, radartools_typed.cu:128-170).
 *   shape[rank], guard[rank], reference[rank]; threshold float[count];
 *   detections uint8[count] (the reference's CfarDetection {0,1}).
 * The threshold is bit-identical to the reference: an FP32 integral image
 * summed along axis 0 then axis 1 in the reference's order, T=(alpha/N)*(outer-
 * inner) with alpha rounded to float, T=0 on border cells, det = x > T.
 * workspace: device float[count] scratch, or NULL to let the library keep a
 * per-thread buffer that grows on demand.
 */
int zkx_ca_cfar_f32(const float* x, int rank, const int* shape, const int* guard,
                    const int* reference, double pfa, float* threshold, uint8_t* detections,
                    float* workspace, void* stream);

/*
 * zkx_ca_cfar_f32 without any allocation: the integral image lives in a fixed
 * device array of zkx_ca_cfar_noalloc_max_cells() cells (4194304 FP32, 16 MiB,
 * held on every device the library runs on until the process ends).  No
 * counterpart in the reference: for callers that cannot allocate device
 * memory (code under HIP graph capture, tight serving loops).  Maps larger
 * than the scratch return ZKX_INVALID_SHAPE before any launch; results are
 * bit-identical to zkx_ca_cfar_f32.  Every call on one device shares that
 * device's scratch: calls that overlap on one device (other streams or host
 * threads) race and must be ordered by the caller.
 */
long long zkx_ca_cfar_noalloc_max_cells(void);
int zkx_ca_cfar_f32_noalloc(const float* x, int rank, const int* shape, const int* guard,
                            const int* reference, double pfa, float* threshold,
                            uint8_t* detections, void* stream);

/*
 * Windowed CA-CFAR for rank 2 with FP64 box sums built as 2-D prefix sums in
 * LDS: threshold T = (alpha/N)*(outer-inner) with alpha and the sums in FP64,
 * det = x > T compared in FP64, T stored as float.  Borders as above.  This is
 * the CFAR of the fused dwell pipeline; it is not bit-identical to the
 * reference's FP32 integral image (which loses precision far from the origin)
 * and is checked against the FP64 oracle instead.
 */
int zkx_ca_cfar_window_f32(const float* x, int rows, int cols, int guard_rows, int guard_cols,
                           int ref_rows, int ref_cols, double pfa, float* threshold,
                           uint8_t* detections, void* stream);

/* ComplexFP64 storage from a ComplexFP32 result (the reference widens on the
 * host, host_output_finalize.h:40-57; here it is a device kernel). */
int zkx_widen_c64_to_c128(const zkx_complex64* in, zkx_complex128* out, size_t count, void* stream);
int zkx_widen_f32_to_f64(const float* in, double* out, size_t count, void* stream);

/* ---------------------------------------------------------------------- */
/* ambiguity function                                                      */
/* ---------------------------------------------------------------------- */

/* Number of output values of zkx_ambgfun_c64 for first extents nx, ny and a
 * cut (0 = 2-D, 1 = delay, 2 = Doppler), or -1 for invalid arguments. */
long long zkx_ambgfun_out_len(int nx, int ny, int cut);

/*
 * Normalized ambiguity function, complex64 in, FP32 compute.
 * Replaces ambgfun_device (radartools_typed.h:134-139, radartools_typed.cpp:411-462,
 * radartools_typed.cu:30-126).
 *   x [x_count] (device); the first nx samples take part, the norm covers all
 *   x_count (a higher-rank input flattened: nx = shape[0]).  y likewise, or
 *   NULL for y = x (then y_count and ny are ignored).  rows = nx + ny - 1,
 *   nfreq = the next power of two >= rows.
 *   cut 0 (2-D): out [rows, nfreq], float, or double when out_fp64 != 0 (the
 *     reference's FP64 storage for integer inputs; values computed in FP32);
 *   cut 1 (delay, at delay cut_value seconds): out double [nfreq];
 *   cut 2 (Doppler, at cut_value Hz): out double [rows].
 * fs must be nonzero for the cuts; cuts need x_count == nx and y_count == ny.
 * The 2-D and delay cuts take nfreq up to 2^26 (UNSUPPORTED_OPERATION
 * beyond): one fused LDS kernel up to 8192, the global FFT beyond (the 2-D
 * cut in chunks of about 256 MiB of delay products).
 * Zero-norm inputs give NaN, as in the reference.
 */
int zkx_ambgfun_c64(const zkx_complex64* x, size_t x_count, int nx, const zkx_complex64* y, size_t y_count,
                    int ny, int cut, double fs, double cut_value, int out_fp64, void* out, void* stream);

/* ---------------------------------------------------------------------- */
/* synthetic echo source (the step before the radar path)                  */
/* ---------------------------------------------------------------------- */

/* Task1 configuration fields that the echo step reads
 * (Task1/pipeline/include/task1/task1_config.h:16-34); doppler_hz is
 * doppler_bin * prf_hz / num_pulses in FP32 (task1_config.h:100-103). */
typedef struct zkx_echo_params {
    int num_pulses;
    int samples_per_pulse;
    int pulse_samples;
    int target_delay_samples;
    float sample_rate_hz;
    float prf_hz;
    float bandwidth_hz;
    float doppler_hz;
    float target_amplitude;
    float noise_std;
    uint32_t noise_seed;
} zkx_echo_params;

/*
 * LFM waveform, delayed Doppler-shifted echo of one point target, and uniform
 * hash noise, generated on the device.  Replaces Task1 step 1
 * (Task1/pipeline/src/steps/step1/step1.cu:11-71: generate_lfm_kernel and
 * simulate_echo_kernel).
 *   waveform [pulse_samples], noiseless / noise / echo [num_pulses, samples_per_pulse];
 *   waveform, noiseless and noise may be NULL.  The noise is bit-identical to
 *   the reference's; the rest is FP32 with its operation order.
 */
int zkx_echo_source_c64(const zkx_echo_params* params, zkx_complex64* waveform, zkx_complex64* noiseless,
                        zkx_complex64* noise, zkx_complex64* echo, void* stream);

/* ---------------------------------------------------------------------- */
/* element types of the reference's other instantiations                   */
/* ---------------------------------------------------------------------- */
/* The operators are instantiated for float, __half, int32, int16 and int8
 * (radartools_typed.cpp:810-811, filtering_typed.cpp:2781+); every input
 * sample is loaded to float (operator_compute_policy.h:42-45,64-67) and the
 * FP32 kernels run on it.  Complex inputs are 2*count scalars. */
#define ZKX_DTYPE_F32 0
#define ZKX_DTYPE_F16 1
#define ZKX_DTYPE_I32 2
#define ZKX_DTYPE_I16 3
#define ZKX_DTYPE_I8 4

/* out[i] = (float)in[i] for `count` scalars of element type `dtype`
 * (the reference's convert_signal_kernel, filtering_kernels.cuh:1277-1286). */
int zkx_convert_to_f32(const void* in, int dtype, size_t count, float* out, void* stream);

/* CA-CFAR detections of an integer input from its float thresholds:
 * det = x > floor(T) in the input type; NaN T -> 0, -inf -> 1, +inf -> 0,
 * T below / at or above the type's range -> 1 / 0
 * (ca_cfar_detect, radartools_kernels.cuh:23-37).  dtype: an integer code. */
int zkx_cfar_detect_typed(const void* x, int dtype, const float* threshold, size_t count,
                          uint8_t* detections, void* stream);

/* resample_poly custom coefficients scaled by the reduced up in their own
 * element type, as float taps (resample_poly_scaled_custom_value_device,
 * filtering_kernels.cuh:460-507): float c*up; half rounded to half when
 * up <= 255; integers multiplied with two's-complement wrap-around in int8
 * (up <= 127, int8 only), int16 (up < 32768) or int32. */
int zkx_resample_poly_scale_taps(const void* coefficients, int dtype, int count, int up, float* taps,
                                 void* stream);

/* cfar_alpha_device: alpha = N * (powf(pfa, -1/N) - 1) computed in FP32 on
 * the device and stored as one double at `alpha` (device memory)
 * (cfar_alpha_kernel, radartools_kernels.cuh:136-148; radartools_typed.cpp:605-617). */
int zkx_cfar_alpha_f32_device(float pfa, int reference, double* alpha, void* stream);

/* ---------------------------------------------------------------------- */
/* fused dwell pipeline (Task1 steps 2-4 in one pass per dwell)             */
/* ---------------------------------------------------------------------- */

/*
 * A batch of same-shape dwells: pulse compression with a prepared template
 * spectrum -> Doppler FFT (window over pulses) -> |.|^2 -> CA-CFAR.
 * Follows the reference's Task1 chain (Task1/pipeline/src/steps/step2/step2.cu:20-36,
 * step3/step3.cu:17-35, step4/step4.cu:66-84) with the Doppler nfft = num_pulses.
 *   x          [batch, P, R] complex64
 *   spectrum   [R] prepared by zkx_pc_template_prepare with nfft = R
 *   threshold  [batch, P, R] float; detections [batch, P, R] uint8
 *   rd_map     optional [batch, P, R] complex64 range-Doppler map (NULL to skip)
 *   cfar_mode  0: bit-identical reference integral image; 1: FP64 windowed sums
 *   workspace  device buffer of zkx_dwell_workspace_bytes(...) bytes
 */
size_t zkx_dwell_workspace_bytes(int batch, int num_pulses, int range_bins);

/*
 * Task1 steps 3-4 fused: Doppler FFT along pulses (nfft = num_pulses, a power
 * of two <= 128), |.|^2 and the FP64 windowed CA-CFAR, with the power map kept
 * in LDS.  x [batch, P, R] is the range-compressed cube; rd_map and power
 * (each [batch, P, R], may be NULL) optionally receive the complex map and
 * |.|^2.  Threshold / detections as zkx_ca_cfar_window_f32.  Requires
 * 2*(guard_range + ref_range) <= 56.
 */
int zkx_doppler_cfar_c64(const zkx_complex64* x, int batch, int num_pulses, int range_bins,
                         int doppler_window, int guard_doppler, int guard_range, int ref_doppler,
                         int ref_range, double pfa, float* threshold, uint8_t* detections,
                         zkx_complex64* rd_map, float* power, void* stream);
int zkx_dwell_batch_c64(const zkx_complex64* x, int batch, int num_pulses, int range_bins,
                        const zkx_complex64* spectrum, int doppler_window,
                        int guard_doppler, int guard_range, int ref_doppler, int ref_range,
                        double pfa, int cfar_mode, float* threshold, uint8_t* detections,
                        zkx_complex64* rd_map, void* workspace, void* stream);

/* ---------------------------------------------------------------------- */
/* heterogeneous dwell batch engine                                         */
/* ---------------------------------------------------------------------- */

/*
 * Many dwells of different pulse counts, range lengths and templates in one
 * run: two persistent launches (pulse compression; Doppler FFT + |.|^2 + FP64
 * windowed CA-CFAR) replayed from a HIP graph.  No reference counterpart: the
 * reference runs the Task1 chain one dwell per operator call
 * (Task1/pipeline/src/steps/step2/step2.cu:20-36, step3/step3.cu:17-35,
 * step4/step4.cu:66-84).  Per dwell the results are those of
 * pulse_compression (nfft = R, the engine's window and normalisation) ->
 * pulse_doppler (nfft = P, the engine's window) -> zkx_ca_cfar_window_f32 on
 * |.|^2 with the engine's guard/reference cells and pfa.
 *
 * Dwell p has P = num_pulses (a power of two in [16, 256]) pulses of
 * R = range_bins samples at x + input_offset ([P][R] complex64); its outputs
 * are [P][R] at threshold/detections (+ rd_map) + output_offset.  Templates
 * are registered once and stay resident (with their spectra at every FFT
 * length); template taps T <= R.  R need not be a power of two: such pulses
 * are compressed by overlap-save segments of a power-of-two FFT over the
 * periodic pulse, which gives the same length-R circular correlation.
 *
 * Lanes: each lane owns its device job tables, workspace and graph; calls on
 * one lane must be issued from one host thread at a time and keep one
 * stream (a change of stream waits for the previous one).  Runs on
 * different lanes and streams overlap on the device.
 */
typedef struct zkx_engine zkx_engine;
typedef struct {
    int num_pulses;
    int range_bins;
    int template_id;
    int reserved;
    long long input_offset;  /* complex64 elements */
    long long output_offset; /* cells */
} zkx_dwell_desc;

int zkx_engine_create(zkx_engine** engine, int pc_window, int normalize, int doppler_window,
                      int guard_doppler, int guard_range, int ref_doppler, int ref_range, double pfa,
                      int lanes);
int zkx_engine_destroy(zkx_engine* engine);
/* h: complex64 taps in device or host memory; returns the template id */
int zkx_engine_add_template(zkx_engine* engine, const zkx_complex64* h, int taps, int* id);
/* use_graph 0/1; grid sizes in blocks (0 keeps the current value) */
int zkx_engine_set_options(zkx_engine* engine, int use_graph, int pc_grid, int dc_grid);
int zkx_engine_run(zkx_engine* engine, int lane, const zkx_complex64* x, const zkx_dwell_desc* dwells,
                   int n, float* threshold, uint8_t* detections, zkx_complex64* rd_map, void* stream);
/* FFT length and segments per pulse the engine would use for each dwell */
int zkx_engine_plan(zkx_engine* engine, const zkx_dwell_desc* dwells, int n, int* fft_length,
                    int* segments);
/* time candidate FFT lengths per (R, T) class and grid sizes on this batch */
int zkx_engine_autotune(zkx_engine* engine, int lane, const zkx_complex64* x,
                        const zkx_dwell_desc* dwells, int n, float* threshold, uint8_t* detections,
                        void* stream);
/* tuning table as text ("grid PC DC" / "pc R T N" lines); length excludes the NUL */
int zkx_engine_tuning_export(zkx_engine* engine, char* buf, int cap, int* length);
int zkx_engine_tuning_import(zkx_engine* engine, const char* text);

/* ---------------------------------------------------------------------- */
/* continuous-batching server                                               */
/* ---------------------------------------------------------------------- */

/*
 * Requests name a dwell resident in the device input pool (complex64
 * elements at input_offset, [P][R]) and a template registered with the
 * engine.  A scheduler thread admits queued requests in arrival order while
 * the HBM budget allows (13 bytes per sample: 8-byte pulse-compressed
 * workspace + 4-byte threshold + 1-byte detections) and runs them as one
 * engine batch on the first free lane (one stream per lane), so batches
 * formed while others run overlap on the device.  No reference counterpart
 * (the reference's Task1 chain handles one dwell per call,
 * Task1/pipeline/src/task1_runner.cu:224-301).  The server uses engine lanes
 * 0 .. lanes-1; the caller must not run those lanes itself meanwhile.
 * keep_outputs: results stay in the server's arena until zkx_server_release;
 * otherwise they are freed when the batch ends (latency-only use).
 */
typedef struct zkx_server zkx_server;
int zkx_server_create(zkx_server** server, zkx_engine* engine, const zkx_complex64* input_pool,
                      long long pool_elems, long long hbm_budget_bytes, int lanes, int max_batch_dwells,
                      int keep_outputs);
int zkx_server_destroy(zkx_server* server);
int zkx_server_submit(zkx_server* server, long long input_offset, int num_pulses, int range_bins,
                      int template_id, long long* ticket);
int zkx_server_wait(zkx_server* server, long long ticket, int timeout_ms, int* done);
int zkx_server_result(zkx_server* server, long long ticket, double* latency_us, const float** threshold,
                      const uint8_t** detections);
/* copy a kept request's outputs ([P][R] float, uint8) into caller device buffers */
int zkx_server_copy_result(zkx_server* server, long long ticket, float* threshold, uint8_t* detections,
                           void* stream);
int zkx_server_release(zkx_server* server, long long ticket);
/* hold (1) or resume (0) batch formation; queued requests wait while held */
int zkx_server_hold(zkx_server* server, int hold);
/* stats[5]: batches launched, dwells finished, queued, reserved bytes, peak reserved bytes */
int zkx_server_stats(zkx_server* server, long long* stats);
/* open-loop load from the calling thread: request i submitted at arrival_us[i] */
int zkx_server_replay(zkx_server* server, const double* arrival_us, const long long* input_offset,
                      const int* num_pulses, const int* range_bins, const int* template_id, int n,
                      double* latency_us);

/* ---------------------------------------------------------------------- */
/* split-BF16 matrix-core Doppler DFT                                       */
/* ---------------------------------------------------------------------- */

/*
 * Doppler DFT (nfft = num_pulses) of channels x [P, R] complex64 cubes as a
 * dense contraction on the MFMA units: Y' = A X' with A the (2P x 2P) real
 * form of the windowed DFT matrix.  terms = 3: split-BF16 operands
 * (hi*hi + hi*lo + lo*hi, FP32 accumulate); terms = 1: plain BF16.  The same
 * transform as zkx_pulse_doppler_c64 with nfft = P (pulse_doppler_device,
 * radartools_typed.cpp:714-808), computed as a matrix product; P need only be
 * a multiple of 16 (<= 2048).  out [channels, P, R]; power optional |Y|^2.
 */
int zkx_doppler_dft_mfma_c64(const zkx_complex64* x, int channels, int num_pulses, int range_bins,
                             int window, int terms, zkx_complex64* out, float* power, void* stream);

/* ---------------------------------------------------------------------- */
/* polyphase filtering                                                      */
/* ---------------------------------------------------------------------- */

/*
 * upfirdn: y[n] = sum_k h[k] x_up[n*down - k], x_up the zero-stuffed input.
 * Replaces upfirdn_device<float> (filtering_typed.h:744-748,
 * filtering_typed.cpp:1046-1123, filtering_typed.cu:596-645) for float32
 * lines, and adds complex64 lines (real taps, per component).
 *   x laid out as [outer, n, inner] with the filtered axis in the middle;
 *   out [outer, out_n, inner], out_n = floor(((n-1)*up + K-1)/down) + 1.
 */
long long zkx_upfirdn_out_len(long long n, long long taps, int up, int down);
int zkx_upfirdn_f32(const float* h, int taps, const float* x, int outer, int n, int inner,
                    int up, int down, float* out, void* stream);
int zkx_upfirdn_c64(const float* h, int taps, const zkx_complex64* x, int outer, int n,
                    int inner, int up, int down, zkx_complex64* out, void* stream);

/*
 * resample_poly with caller taps (already scaled: h*up as the reference's
 * resample_poly_scaled_custom_value, filtering_typed.cpp:129-157, or the FP64
 * design of :176-218 rounded to float).  Replaces resample_poly_device<float>
 * (filtering_typed.h:765-770, filtering_typed.cpp:2717-2779,
 * filtering_typed.cu:646-681).  up/down are reduced by their gcd here.
 *   out_n = ceil(n*up/down); the polyphase phase schedule follows
 *   prepare_resample_poly_layout (filtering_typed.cpp:933-1044).
 */
int zkx_resample_poly_f32(const float* taps, int ntaps, const float* x, int outer, int n,
                          int inner, int up, int down, float* out, void* stream);
int zkx_resample_poly_c64(const float* taps, int ntaps, const zkx_complex64* x, int outer,
                          int n, int inner, int up, int down, zkx_complex64* out, void* stream);

/*
 * firfilter: y[n] = sum_{k<=K-1} b[k] x[n-k] (+ zi[n] for n < K-1), zf = the
 * full-convolution tail.  Replaces firfilter_device<float> and
 * firfilter_fp32_resident_device (filtering_typed.h:552-565,819-822,
 * filtering_typed.cpp:1125-1224, filtering_typed.cu:117-184).
 *   zi, zf: [outer, K-1, inner] or NULL (zi already broadcast by the caller).
 */
int zkx_firfilter_f32(const float* b, int taps, const float* x, int outer, int n, int inner,
                      const float* zi, float* y, float* zf, void* stream);

/*
 * decimate: FIR then every q-th sample along the middle axis of [outer, n,
 * inner] float32 lines.  Replaces decimate_device (filtering_typed.h:702-706,
 * filtering_typed.cpp:2319-2356, decimate_typed_cpu :2229-2317).
 *   y[o] = sum_{t<K} h[t] x[(o + start) q - prefix - t], o < ceil(n/q);
 *   zero_phase != 0: prefix = q - half % q, start = (half + prefix)/q,
 *   half = (K-1)/2; otherwise prefix = start = 0.  q = 1 filters too.
 *   out: [outer, ceil(n/q), inner].
 */
int zkx_decimate_f32(const float* taps, int ntaps, const float* x, int outer, int n, int inner, int q,
                     int zero_phase, float* out, void* stream);

/*
 * Polyphase channelizer, float32 samples and taps, complex64 out.  Replaces
 * channelize_poly_device (filtering_typed.h:513-523; plan
 * filtering_typed.cpp:443-466).
 *   M = n_chans, T = h_count / M (<= 32), P = x_count / M (short last groups
 *   ignored); out [M][P] channel-major:
 *   out[c][p] = sum_{ph<M} e^{+2 pi i c ph/M} sum_{t<T, p>=t} x[(p-t)M + M-1-ph] h[ph + tM].
 *   M must be a power of two <= 8192 or any value <= 4096.
 */
int zkx_channelize_poly_f32(const float* x, long long x_count, const float* h, long long h_count, int n_chans,
                            zkx_complex64* out, void* stream);

/*
 * 1-D cross-correlation, direct method, in the element type `dtype`
 * (ZKX_DTYPE_*), out in the same type.  Replaces correlate_device(in1, in2,
 * out, mode) (convolution_typed.h:298-303; convolution_typed.cpp:134-152,
 * 285-334).  mode 0 full (L1+L2-1), 1 same (L1), 2 valid (|L1-L2|+1); s =
 * (full - n)/2 + k:  out[k] = sum_j in1[s - (L2-1-j)] in2[j].  float/half: one
 * FP32 fma chain in j order (half stored rounded); integers: products summed
 * modulo 2^bits.  Bit-identical to the reference.  out must not overlap an
 * input.  zkx_correlate_out_len gives n (-1 for an invalid mode).
 */
long long zkx_correlate_out_len(long long l1, long long l2, int mode);
int zkx_correlate_direct(const void* in1, long long l1, const void* in2, long long l2, int dtype, int mode,
                         void* out, void* stream);

/*
 * 1-D cross-correlation, FFT method.  Replaces the workspace overload
 * correlate_device(in1, in2, out, CorrelateDeviceWorkspace&)
 * (convolution_typed.h:307-327): inputs converted to complex FP32, in2
 * reversed, the full convolution by the overlap-save FFT, the centred crop
 * stored in the element type (half rounded to nearest; integers rounded half
 * to even and wrapped modulo 2^bits).  FP32 FFT results, not the direct sums
 * bit for bit.  Any lengths with l1 + l2 - 1 <= 2^26 (UNSUPPORTED_OPERATION
 * beyond): overlap-save when one input has at most 4096 samples, one
 * whole-length power-of-two transform pair otherwise; the library holds the
 * workspace.
 */
int zkx_correlate_fft(const void* in1, long long l1, const void* in2, long long l2, int dtype, int mode, void* out,
                      void* stream);

/*
 * N-D cross-correlation by FFT.  Replaces correlate_nd_output_shape and
 * correlate_device(in1, in2, out, CorrelateNDWorkspace&)
 * (convolution_typed.h:93,331-337; convolution_typed.cu:150-241;
 * convolution_typed.cpp:153-199).  Row-major real inputs of one `dtype`,
 * equal rank 0..8.  Output extents per dimension: full s1 + s2 - 1, same
 * s1, valid |s1 - s2| + 1 ('valid' needs one input at least as large as the
 * other in every dimension, INVALID_SHAPE otherwise); an empty input gives
 * an empty output (elements 0, out_shape[0] = 0); rank 0 is the product of
 * two scalars (elements 1).  Computation: every dimension padded to the
 * power of two >= s1 + s2 - 1, in2 reversed, complex FP32 transforms along
 * each dimension, the centred crop (start (full - out)/2) stored in the
 * element type as zkx_correlate_fft does.  Rank 1 runs zkx_correlate_fft.
 * Each padded operand may hold at most 2^31 - 1 elements and each padded
 * extent at most 2^26.  out must not overlap an input.
 */
int zkx_correlate_nd_out_shape(const int* shape1, const int* shape2, int rank, int mode, int* out_shape,
                               long long* elements);
int zkx_correlate_nd(const void* in1, const int* shape1, const void* in2, const int* shape2, int rank, int dtype,
                     int mode, void* out, void* stream);

/*
 * 2-D cross-correlation, direct, row-major, in the element type `dtype`.
 * Replaces correlate2d_output_shape / correlate2d_device(in1, in2, out,
 * workspace) (convolution_typed.h:218, 374-379; convolution_typed.cpp:101-131,
 * 203-260, 416-470).  mode 0 full, 1 same (origin r2/2, c2/2), 2 valid (one
 * input must dominate in both dimensions); boundary 0 fill (fillvalue in the
 * element type: integers truncated and wrapped, a non-finite integer fill is
 * INVALID_ARGUMENT), 1 wrap, 2 symm (edge-inclusive mirror); valid ignores
 * boundary and fill.  Terms in row-major order of in2; float/half one FP32
 * fma chain, integers modulo 2^bits.  Bit-identical to the reference.  out
 * must not overlap an input.
 */
int zkx_correlate2d_out_shape(long long r1, long long c1, long long r2, long long c2, int mode, long long* rows,
                              long long* cols);
int zkx_correlate2d(const void* in1, long long r1, long long c1, const void* in2, long long r2, long long c2, int dtype,
                    int mode, int boundary, double fillvalue, void* out, void* stream);

/*
 * Host-side FIR design for decimate's default filter: firwin with one
 * lowpass cutoff (fs = 2, pass_zero, scale) and a Hamming window, FP64 taps
 * (filter_design_typed.cpp:24-100, 268-296; decimate designs numtaps =
 * order + 1 = 20q + 1 at cutoff 1/q, filtering_typed.cpp:318-336).
 * cutoff must lie in (0, 1).  No device work.
 */
int zkx_firwin_lowpass_hamming(int numtaps, float cutoff, double* taps);

/*
 * fftconvolve, full/same/valid (mode 0/1/2) linear convolution of two 1-D
 * complex64 signals via the LDS FFT.  No C++ counterpart exists in the
 * reference library (operator_contracts.cmake:10-63); semantics follow
 * cuSignal's fftconvolve (cusignal/convolution/convolve.py:159-320).
 * batch rows of in1 [batch, n1] are convolved with the one in2 [n2]
 * (n2 <= 4096); long rows run as overlap-save segments of 8192 points.
 */
long long zkx_fftconvolve_out_len(long long n1, long long n2, int mode);
int zkx_fftconvolve_c64(const zkx_complex64* in1, int batch, int n1, const zkx_complex64* in2,
                        int n2, int mode, zkx_complex64* out, void* stream);

/*
 * fftconvolve over N-D axes (cuSignal convolve.py:159-320 with
 * convolution_utils.py:40-100): in1 [shape1], in2 [shape2], equal rank
 * 1..8, row-major, each float32 (complex = 0) or complex64 (complex = 1);
 * convolved over the dimensions whose bit is set in axes_mask (all:
 * (1 << rank) - 1); on the other dimensions the extents must be equal or 1
 * (broadcast), INVALID_SHAPE otherwise.  Extents of the result: full
 * s1 + s2 - 1 on a convolved axis and max(s1, s2) elsewhere; same = s1; valid
 * = s1 - s2 + 1 on a convolved axis (operands swapped when in2 is the larger
 * in every dimension; INVALID_SHAPE when neither is), full elsewhere; the
 * block kept is centred, start (full - kept) / 2.  out is float32 when both
 * inputs are real, complex64 otherwise.  Computation: convolved axes padded
 * to the power of two >= s1 + s2 - 1, complex FP32 transforms along each,
 * pointwise product with broadcasting, inverse transforms, crop; two
 * complex inputs convolved along the last axis only, with in2 a single row
 * of at most 4096 samples, run as zkx_fftconvolve_c64 over the rows of in1.
 * Each padded operand may hold at most 2^31 - 1 elements and each padded
 * extent at most 2^26.  zkx_fftconvolve_nd_shape writes the result's extents.
 */
int zkx_fftconvolve_nd_shape(const long long* shape1, const long long* shape2, int rank, unsigned axes_mask,
                             int mode, long long* out_shape);
int zkx_fftconvolve_nd(const void* in1, const long long* shape1, int complex1, const void* in2,
                       const long long* shape2, int complex2, int rank, unsigned axes_mask, int mode, void* out,
                       void* stream);

#ifdef __cplusplus
}
#endif

#endif /* ZKX_RADAR_H */
