/*
 * zkx_fft.h -- the reference's FFT backend C ABI, served by the gfx950 LDS
 * Stockham FFT in libzkx_radar.so.
 *
 * Same names, enums, numeric values and semantics as
 * ZKX/cusignal_cpp/include/cusignal/backends/fft/fft_interface.h:189-447
 * (the ABI that libzkx_fft_thrust.so and the vendor wrapper export), so a
 * caller linked against that header can link this library instead:
 *   - forward transforms are unnormalised, inverse transforms scale by 1/N
 *     (fft_interface.h:39-42);
 *   - data are row-major batches of {float re, im};
 *   - plans are opaque handles;
 *   - FFTResult codes 0 ok, 1 invalid value, 2 alloc failed, 3 exec failed,
 *     99 internal error.
 * Lengths: powers of two 1..8192 (one LDS-resident pass per row), other
 * lengths up to 4096 (Bluestein, one fused kernel per row), and beyond those
 * powers of two up to 2^26 (four-step over the LDS passes) and other lengths
 * up to 2^25 (Bluestein in global passes); longer plans return
 * FFT_INVALID_VALUE.  A plan allocates its tables and scratch at creation,
 * so execution allocates nothing.  Plan kinds: 1-D C2C, R2C, C2R (with the
 * reference backend's half-spectrum mirroring and scaling,
 * fft_thrust.cu:1096-1160) and 2-D (complex data for every type, as
 * fft_thrust.cu:1014-1058).  fft_execute* run on the default stream, as the
 * reference's backends do; fft_execute_on_stream takes a stream.
 */
#ifndef ZKX_FFT_H
#define ZKX_FFT_H

#ifdef __cplusplus
extern "C" {
#endif

typedef void* FFTPlanHandle;

typedef enum { FFT_FORWARD = 1, FFT_INVERSE = -1 } FFTDirection;
typedef enum { FFT_TYPE_C2C = 0, FFT_TYPE_R2C = 1, FFT_TYPE_C2R = 2 } FFTType;
typedef enum {
    FFT_SUCCESS = 0,
    FFT_INVALID_VALUE = 1,
    FFT_ALLOC_FAILED = 2,
    FFT_EXEC_FAILED = 3,
    FFT_INTERNAL_ERROR = 99
} FFTResult;

/* fft_interface.h:189-205 */
FFTResult fft_plan_create_1d(FFTPlanHandle* plan, int nx, int batch, FFTType type,
                             FFTDirection direction);
/* fft_interface.h:207-222: 2-D plan over complex [ny][nx] */
FFTResult fft_plan_create_2d(FFTPlanHandle* plan, int nx, int ny, FFTType type,
                             FFTDirection direction);
/* fft_interface.h:224-240: out-of-place execute */
FFTResult fft_execute(FFTPlanHandle plan, void* output, const void* input);
/* fft_interface.h:242-255 */
FFTResult fft_execute_inplace(FFTPlanHandle plan, void* data);
/* fft_interface.h:257-264 */
FFTResult fft_plan_destroy(FFTPlanHandle plan);
/* fft_interface.h:266-330: zero-frequency shifts of device arrays */
FFTResult fft_shift(void* output, const void* input, int nx, int batch);
FFTResult fft_inverse_shift(void* output, const void* input, int nx, int batch);
FFTResult fft_shift_2d(void* output, const void* input, int nx, int ny);
FFTResult fft_inverse_shift_2d(void* output, const void* input, int nx, int ny);
/* fft_interface.h:332-352 */
const char* fft_error_string(FFTResult result);
int fft_version(void);
/* fft_interface.h:354-390: sample frequencies, host or device destination */
FFTResult fft_freq(float* freqs, int n, float d);
FFTResult fft_rfftfreq(float* freqs, int n, float d);

/* Stream-aware execute for callers that do not use the default stream
 * (no counterpart in the reference ABI). */
FFTResult fft_execute_on_stream(FFTPlanHandle plan, void* output, const void* input, void* stream);

#ifdef __cplusplus
}
#endif

#endif /* ZKX_FFT_H */
