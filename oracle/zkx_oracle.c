/*
 * zkx_oracle.c -- CPU oracle for the radar hot path.  TEST INFRASTRUCTURE ONLY.
 *
 * This file restates, in plain C, the reference's host CPU functions for the
 * radar path so that tests can check the HIP library against them.  Nothing in
 * the product library links or calls this code; only tests/, the smoke check
 * and the CPU-baseline leg of bench.py load it, and only as the checker.
 *
 * Two families of functions live here:
 *   - "faithful" restatements that keep the reference's operation order and
 *     rounding (FP32 sums where it sums in FP32, float pi in the Doppler DFT,
 *     whole-array FP32 integral image in CA-CFAR).  These are pinned bit-for-bit
 *     against the sha256 digests the reference recorded for its own CPU outputs.
 *   - "exact" FP64 variants (true pi, double windows, FP64 box sums) used as the
 *     FP64 ground truth for the 1e-5 relative-L2 bar.
 *
 * Build: gcc -O2 -fPIC -shared -ffp-contract=off (see oracle/Makefile).  FMA
 * contraction must stay off so that a*b+c rounds twice, as the reference did.
 *
 * Reference functions followed (paths relative to the reference repository):
 *   ZKX/cusignal_cpp/src/operators/radartools/radartools_typed.cpp
 *     :238-251 radar_window_host, :592-603 cfar_alpha_typed_cpu,
 *     :161-220 validate_ca_cfar, :222-236 ca_cfar_detect_host,
 *     :465-555 ca_cfar_typed_cpu, :618-675 pulse_compression_typed_cpu,
 *     :714-759 pulse_doppler_typed_cpu, :16-72 + :329-409 ambgfun_typed_cpu
 *   ZKX/cusignal_cpp/src/operators/filtering/filtering_typed.cpp
 *     :104-157 resample_poly_scaled_custom_value, :159-174 resample_poly_bessel_i0,
 *     :176-218 design_resample_poly_coefficients,
 *     :933-1044 prepare_resample_poly_layout, :1046-1117 prepare_upfirdn_layout,
 *     :1536-1592 firfilter_typed_cpu, :2548-2616 upfirdn_typed_cpu,
 *     :2617-2715 resample_poly_typed_cpu, :318-336 + :2229-2317 decimate_typed_cpu
 *     :443-466 prepare_channelize_poly + channelize_poly_typed_cpu
 *   ZKX/cusignal_cpp/src/operators/convolution/convolution_typed.cpp
 *     :134-152 correlate_output_size, :285-334 correlate_typed_cpu (direct)
 *   ZKX/cusignal_cpp/src/operators/filter_design/filter_design_typed.cpp
 *     :14-16 sinc, :24-100 prepare_firwin, :268-296 firwin_typed_cpu
 *   Probe input generators:
 *     ZKX/test_all/operators/shared/radartools/radartools_operator_probe.cu:35-36
 *     ZKX/test_all/operators/cases/filtering/upfirdn/operator_probe.cu:23
 *     ZKX/test_all/operators/cases/filtering/resample_poly/operator_probe.cu:28-29
 *     ZKX/test_all/operators/shared/filtering/task2_filter_probe.cu:28
 *     ZKX/test_all/operators/cases/filtering/channelize_poly/operator_probe.cu:122-135
 */
#include <math.h>
#include <stdint.h>
#include <stdlib.h>
#include <string.h>

#define ORC_API __attribute__((visibility("default")))

/* window kinds, numbered as the reference's RadarWindowKind enum */
enum { ORC_WIN_NONE = 0, ORC_WIN_HANN = 1, ORC_WIN_HAMMING = 2,
       ORC_WIN_EXPLICIT32 = 3, ORC_WIN_EXPLICIT64 = 4 };

static const float orc_pi_f = 3.14159265358979323846F;
static const double orc_pi_d = 3.141592653589793238462643383279502884;

/* ------------------------------------------------------------------ */
/* probe input generators                                              */
/* ------------------------------------------------------------------ */

/* radartools_operator_probe.cu:36 complex_input: a=(i*13+seed)%11-5,
 * b=(i*7+seed)%9-4, scaled by 0.25 for floating types. */
ORC_API void orc_probe_complex_input(size_t n, int seed, int integral, float *out)
{
    for (size_t i = 0; i < n; ++i) {
        int a = (int)((i * 13 + (size_t)seed) % 11) - 5;
        int b = (int)((i * 7 + (size_t)seed) % 9) - 4;
        out[2 * i] = integral ? (float)a : (float)a * .25f;
        out[2 * i + 1] = integral ? (float)b : (float)b * .25f;
    }
}

/* radartools_operator_probe.cu:35 real_input */
ORC_API void orc_probe_real_input(size_t n, int seed, int integral, float *out)
{
    for (size_t i = 0; i < n; ++i) {
        int z = (int)((i * 17 + (size_t)seed) % 19) - 9;
        out[i] = integral ? (float)z : (float)z * .125f + (float)(i % 7) * .03125f;
    }
}

/* upfirdn/operator_probe.cu:23 make(): r=(i*17+seed)%9-4 (seed is uint64) */
ORC_API void orc_upfirdn_probe_input(size_t n, uint64_t seed, int integral, float *out)
{
    for (size_t i = 0; i < n; ++i) {
        int r = (int)((i * 17 + seed) % 9) - 4;
        out[i] = integral ? (float)r : (float)r * .25f;
    }
}

/* resample_poly/operator_probe.cu:28 input(): r=(i*17+seed)%11-5 */
ORC_API void orc_resample_probe_input(size_t n, uint64_t seed, int integral, float *out)
{
    for (size_t i = 0; i < n; ++i) {
        int r = (int)((i * 17 + seed) % 11) - 5;
        out[i] = integral ? (float)r : (float)r * 0.25f;
    }
}

/* task2_filter_probe.cu:28 signal(): value in double before the dtype cast. */
ORC_API void orc_firfilter_probe_signal(int n, int seed, int integral, double *out)
{
    for (int i = 0; i < n; ++i) {
        double x = 3 * sin(i * .071) + 2 * cos(i * .037) + ((i * 17 + seed) % 11 - 5) * .125;
        if (integral) x = round(x);
        out[i] = x;
    }
}

/* ------------------------------------------------------------------ */
/* windows                                                             */
/* ------------------------------------------------------------------ */

/* radartools_typed.cpp:238-251: symmetric window evaluated in float with a
 * float pi.  Returns the float value the reference multiplies by. */
static float orc_window_f(int kind, int index, int length, const float *explicit_window)
{
    if (kind == ORC_WIN_EXPLICIT32 || kind == ORC_WIN_EXPLICIT64)
        return explicit_window[index];
    if (kind == ORC_WIN_NONE || length <= 1) return 1.0F;
    const float phase = 2.0F * orc_pi_f * (float)index / (float)(length - 1);
    if (kind == ORC_WIN_HANN) return 0.5F - 0.5F * cosf(phase);
    return 0.54F - 0.46F * cosf(phase);
}

/* The same symmetric window in FP64 with an exact pi (exact mode). */
static double orc_window_d(int kind, int index, int length, const float *explicit_window)
{
    if (kind == ORC_WIN_EXPLICIT32 || kind == ORC_WIN_EXPLICIT64)
        return (double)explicit_window[index];
    if (kind == ORC_WIN_NONE || length <= 1) return 1.0;
    const double phase = 2.0 * orc_pi_d * (double)index / (double)(length - 1);
    if (kind == ORC_WIN_HANN) return 0.5 - 0.5 * cos(phase);
    return 0.54 - 0.46 * cos(phase);
}

ORC_API void orc_window(int kind, int length, int exact, double *out)
{
    for (int i = 0; i < length; ++i)
        out[i] = exact ? orc_window_d(kind, i, length, NULL)
                       : (double)orc_window_f(kind, i, length, NULL);
}

/* ------------------------------------------------------------------ */
/* pulse compression (radartools_typed.cpp:618-675)                    */
/* ------------------------------------------------------------------ */

/* x: [P, S] interleaved complex, values already loaded to float.
 * tmpl: [T] interleaved complex.  out: [P, nfft] interleaved complex double.
 * y[p,k] = sum_{t < min(T, nfft)} x[p,(k+t) mod nfft] * conj(h~[t]),
 * x[p,s] = 0 for s >= S, h~ = window*h / ||window*h|| when normalize. */
ORC_API int orc_pulse_compression(const float *x, int P, int S, const float *tmpl, int T,
                                  int nfft, int window, const float *explicit_window,
                                  int normalize, int exact, double *out)
{
    if (nfft == -1) nfft = S;
    if (P < 1 || S < 1 || nfft < 1 || T < 1) return 1002;
    double *tr = (double *)malloc(sizeof(double) * (size_t)T);
    double *ti = (double *)malloc(sizeof(double) * (size_t)T);
    if (!tr || !ti) { free(tr); free(ti); return 2003; }
    double norm = 0.0;
    for (int i = 0; i < T; ++i) {
        const double w = exact ? orc_window_d(window, i, T, explicit_window)
                               : (double)orc_window_f(window, i, T, explicit_window);
        tr[i] = (double)tmpl[2 * i] * w;
        ti[i] = (double)tmpl[2 * i + 1] * w;
        norm += tr[i] * tr[i] + ti[i] * ti[i];
    }
    norm = normalize ? sqrt(norm) : 1.0;
    for (int p = 0; p < P; ++p) {
        for (int k = 0; k < nfft; ++k) {
            double re = 0.0, im = 0.0;
            for (int t = 0; t < nfft && t < T; ++t) {
                const int s = (k + t) % nfft;
                const size_t src = (size_t)p * (size_t)S + (size_t)s;
                const double xr = s < S ? (double)x[2 * src] : 0.0;
                const double xi = s < S ? (double)x[2 * src + 1] : 0.0;
                const double hr = tr[t] / norm;
                const double hi = -ti[t] / norm;
                re += xr * hr - xi * hi;
                im += xr * hi + xi * hr;
            }
            const size_t dst = (size_t)p * (size_t)nfft + (size_t)k;
            out[2 * dst] = re;
            out[2 * dst + 1] = im;
        }
    }
    free(tr);
    free(ti);
    return 0;
}

/* ------------------------------------------------------------------ */
/* pulse-Doppler (radartools_typed.cpp:714-759)                        */
/* ------------------------------------------------------------------ */

/* x: [P, R] interleaved; out: [nfft, R] interleaved double.
 * Y[f,r] = sum_{p < min(P,nfft)} w[p] x[p,r] exp(-2 pi i f p / nfft).
 * Faithful mode keeps the reference's float-rounded pi and float window. */
ORC_API int orc_pulse_doppler(const float *x, int P, int R, int nfft, int window,
                              const float *explicit_window, int exact, double *out)
{
    if (nfft == -1) nfft = P;
    if (P < 1 || R < 1 || nfft < 1) return 1002;
    const int used = P < nfft ? P : nfft;
    double *w = (double *)malloc(sizeof(double) * (size_t)used);
    if (!w) return 2003;
    for (int p = 0; p < used; ++p)
        w[p] = exact ? orc_window_d(window, p, P, explicit_window)
                     : (double)orc_window_f(window, p, P, explicit_window);
    const double pi_used = exact ? orc_pi_d : (double)orc_pi_f;
    for (int f = 0; f < nfft; ++f) {
        for (int r = 0; r < R; ++r) {
            double re = 0.0, im = 0.0;
            for (int p = 0; p < used; ++p) {
                const double angle = -2.0 * pi_used * f * p / nfft;
                const double c = cos(angle);
                const double s = sin(angle);
                const size_t src = (size_t)p * (size_t)R + (size_t)r;
                const double xr = (double)x[2 * src] * w[p];
                const double xi = (double)x[2 * src + 1] * w[p];
                re += xr * c - xi * s;
                im += xr * s + xi * c;
            }
            const size_t dst = (size_t)f * (size_t)R + (size_t)r;
            out[2 * dst] = re;
            out[2 * dst + 1] = im;
        }
    }
    free(w);
    return 0;
}

/* ------------------------------------------------------------------ */
/* decimate                                                            */
/* ------------------------------------------------------------------ */

static double orc_sinc(double x)
{
    return fabs(x) < 1.0e-15 ? 1.0 : sin(orc_pi_d * x) / (orc_pi_d * x);
}

/* filter_design_typed.cpp:24-100 + 268-296 firwin_typed_cpu for one lowpass
 * cutoff (fs = 2, pass_zero, scale, Hamming window): taps in FP64.  cutoff
 * is the loaded (float) cutoff.  Returns 0, or 1001 for an invalid design. */
ORC_API int orc_firwin_lowpass_hamming(int n, double cutoff, double *out)
{
    if (n <= 0 || !(cutoff > 0.0 && cutoff < 1.0)) return 1001;
    const double mid = 0.5 * (double)(n - 1);
    for (int i = 0; i < n; ++i) {
        const double w = n == 1 ? 1.0 : 0.54 - 0.46 * cos(2.0 * orc_pi_d * i / (n - 1));
        const double m = i - mid;
        out[i] = (cutoff * orc_sinc(cutoff * m) - 0.0 * orc_sinc(0.0 * m)) * w;
    }
    double norm = 0.0;
    for (int i = 0; i < n; ++i) norm += out[i] * cos(orc_pi_d * (i - mid) * 0.0);
    for (int i = 0; i < n; ++i) out[i] /= norm;
    return 0;
}

/* filtering_typed.cpp:2229-2317 decimate_typed_cpu on lines [outer][n][inner]:
 * y[o] = sum_{t<K} h[t] x[(o + start) q - prefix - t] (sources outside the
 * line skipped), prefix = q - half % q and start = (half + prefix)/q for
 * zero phase, else 0.  fp32 != 0: FP32 products and sums in tap order (the
 * FP32-output path), else FP64.  x holds the loaded samples. */
ORC_API int orc_decimate(const double *x, int outer, int n, int inner, int q, const double *taps, int ntaps,
                         int zero_phase, int fp32, double *out)
{
    if (q < 1 || ntaps < 1 || outer < 0 || n < 0 || inner < 0) return 1001;
    const int half = (ntaps - 1) / 2;
    const int prefix = zero_phase ? q - half % q : 0;
    const long long start = zero_phase ? (half + prefix) / q : 0;
    const int out_n = n == 0 ? 0 : 1 + (n - 1) / q;
    for (int o = 0; o < outer; ++o) {
        for (int in = 0; in < inner; ++in) {
            const size_t ib = (size_t)o * (size_t)n * (size_t)inner + (size_t)in;
            const size_t ob = (size_t)o * (size_t)out_n * (size_t)inner + (size_t)in;
            for (int k = 0; k < out_n; ++k) {
                const long long f = ((long long)k + start) * q;
                if (fp32) {
                    float sum = 0.0f;
                    for (int pt = prefix; pt < prefix + ntaps; ++pt) {
                        const long long src = f - pt;
                        if (src < 0 || src >= n) continue;
                        sum += (float)taps[pt - prefix] * (float)x[ib + (size_t)src * (size_t)inner];
                    }
                    out[ob + (size_t)k * (size_t)inner] = (double)sum;
                } else {
                    double sum = 0.0;
                    for (int pt = prefix; pt < prefix + ntaps; ++pt) {
                        const long long src = f - pt;
                        if (src < 0 || src >= n) continue;
                        sum += taps[pt - prefix] * x[ib + (size_t)src * (size_t)inner];
                    }
                    out[ob + (size_t)k * (size_t)inner] = sum;
                }
            }
        }
    }
    return 0;
}

/* ------------------------------------------------------------------ */
/* channelize_poly                                                     */
/* ------------------------------------------------------------------ */

/* filtering_typed.cpp:443-466 plan + channelize_poly_typed_cpu: M channels,
 * T = hn/M taps per phase (<= 32), P = n/M points; FP64 polyphase sums and a
 * direct +i DFT over the phases, stored as float (re, im) channel-major
 * [M][P].  x, h hold the loaded samples.  Returns 0 or 1001. */
ORC_API int orc_channelize_poly(const double *x, long long n, const double *h, long long hn, int M, float *out)
{
    if (M <= 0 || n < 0 || hn < 0) return 1001;
    const int T = (int)(hn / M);
    if (T > 32) return 1001;
    const int P = (int)(n / M);
    for (int c = 0; c < M; ++c) {
        for (int p = 0; p < P; ++p) {
            double re = 0.0, im = 0.0;
            for (int ph = 0; ph < M; ++ph) {
                double f = 0.0;
                for (int t = 0; t < T; ++t) {
                    const int g = p - t;
                    if (g >= 0) f += x[(size_t)g * (size_t)M + (size_t)(M - 1 - ph)] * h[ph + (size_t)t * (size_t)M];
                }
                const double angle = 6.283185307179586476925286766559 * (double)c * ph / M;
                re += f * cos(angle);
                im += f * sin(angle);
            }
            out[2 * ((size_t)c * (size_t)P + (size_t)p)] = (float)re;
            out[2 * ((size_t)c * (size_t)P + (size_t)p) + 1] = (float)im;
        }
    }
    return 0;
}

/* ------------------------------------------------------------------ */
/* correlate (1-D, direct)                                             */
/* ------------------------------------------------------------------ */

/* convolution_typed.cpp:134-152 output length, :285-334 correlate_typed_cpu
 * (direct): out[k] = sum_{j<L2} in1[s - (L2-1-j)] in2[j] over the terms inside
 * in1, s = (full - out)/2 + k, full = L1 + L2 - 1.  kind 0: FP32 fma chain in j
 * order (float, and half loaded to float); kind 8/16/32: integer inputs, the
 * products summed modulo 2^kind in the unsigned type, returned as the signed
 * value of those bits.  mode 0 full, 1 same, 2 valid.  out gets the value
 * before the final store to the element type (float results; integers exact). */
ORC_API long long orc_correlate_len(long long l1, long long l2, int mode)
{
    if (l1 == 0 || l2 == 0) return 0;
    if (mode == 0) return l1 + l2 - 1;
    if (mode == 1) return l1;
    return l1 >= l2 ? l1 - l2 + 1 : l2 - l1 + 1;
}

ORC_API int orc_correlate_direct(const double *in1, long long l1, const double *in2, long long l2, int mode,
                                 int kind, double *out)
{
    if (mode < 0 || mode > 2) return 1001;
    const long long n = orc_correlate_len(l1, l2, mode);
    if (n == 0) return 0;
    const long long full = l1 + l2 - 1;
    const long long start = (full - n) / 2;
    for (long long k = 0; k < n; ++k) {
        const long long s = start + k;
        float acc = 0.0f;
        uint32_t macc = 0;
        for (long long j = 0; j < l2; ++j) {
            const long long src = s - (l2 - 1 - j);
            if (src < 0 || src >= l1) continue;
            if (kind == 0) {
                acc = fmaf((float)in1[src], (float)in2[j], acc);
            } else {
                const uint32_t a = (uint32_t)(int64_t)in1[src], b = (uint32_t)(int64_t)in2[j];
                macc = macc + a * b;
            }
        }
        if (kind == 0) {
            out[k] = (double)acc;
        } else {
            const uint32_t mask = kind == 32 ? 0xFFFFFFFFu : ((1u << kind) - 1u);
            const uint32_t bits = macc & mask;
            const uint32_t sign = 1u << (kind - 1);
            out[k] = (bits & sign) ? (double)((int64_t)bits - ((int64_t)1 << kind)) : (double)bits;
        }
    }
    return 0;
}

/* convolution_typed.cpp:101-114 (map_correlate2d_index): an index inside
 * [0, size) is itself; outside, fill (0) gives -1, wrap (1) the index modulo
 * size, symm (2) the mirror with period 2*size that repeats the edge sample. */
static long long orc_map2d(long long index, long long size, int boundary)
{
    if (index >= 0 && index < size) return index;
    if (boundary == 0) return -1;
    if (boundary == 1) {
        long long w = index % size;
        return w < 0 ? w + size : w;
    }
    const long long period = 2 * size;
    long long w = index % period;
    if (w < 0) w += period;
    return w < size ? w : period - 1 - w;
}

/* convolution_typed.cpp:203-236 (correlate2d_output_shape): full r1+r2-1,
 * same r1, valid |r1-r2|+1 when one input dominates in both dimensions (1002
 * otherwise); any empty dimension gives 0 x 0. */
ORC_API int orc_correlate2d_shape(long long r1, long long c1, long long r2, long long c2, int mode, long long *rows,
                                  long long *cols)
{
    *rows = *cols = 0;
    if (mode < 0 || mode > 2) return 1001;
    if (r1 == 0 || c1 == 0 || r2 == 0 || c2 == 0) return 0;
    if (mode == 0) {
        *rows = r1 + r2 - 1;
        *cols = c1 + c2 - 1;
    } else if (mode == 1) {
        *rows = r1;
        *cols = c1;
    } else {
        const int d1 = r1 >= r2 && c1 >= c2, d2 = r2 >= r1 && c2 >= c1;
        if (!d1 && !d2) return 1002;
        *rows = (r1 > r2 ? r1 - r2 : r2 - r1) + 1;
        *cols = (c1 > c2 ? c1 - c2 : c2 - c1) + 1;
    }
    return 0;
}

/* convolution_typed.cpp:416-470 (correlate2d_typed_cpu): output (r, c) sits at
 * (row_start + r, col_start + c) of the full correlation (start 0 for full,
 * r2/2 and c2/2 for same, min(r1,r2)-1 and min(c1,c2)-1 for valid); its terms
 * run over in2 in row-major (kr, kc) order with sample
 * in1[map(fr - (r2-1-kr))][map(fc - (c2-1-kc))], or `fill` (already stored in
 * the element type) where either index maps to -1.  kind 0: one FP32 fma chain;
 * kind 8/16/32: products summed modulo 2^kind, read back as two's complement.
 * Valid mode uses the fill boundary with fill 0 (every sample is inside). */
ORC_API int orc_correlate2d(const double *in1, long long r1, long long c1, const double *in2, long long r2,
                            long long c2, int mode, int boundary, double fill, int kind, double *out)
{
    long long rows, cols;
    const int rc = orc_correlate2d_shape(r1, c1, r2, c2, mode, &rows, &cols);
    if (rc) return rc;
    if (boundary < 0 || boundary > 2) return 1003;
    long long rs = 0, cs = 0;
    if (mode == 1) {
        rs = r2 / 2;
        cs = c2 / 2;
    } else if (mode == 2) {
        rs = (r1 < r2 ? r1 : r2) - 1;
        cs = (c1 < c2 ? c1 : c2) - 1;
        boundary = 0;
        fill = 0.0;
    }
    for (long long r = 0; r < rows; ++r) {
        for (long long c = 0; c < cols; ++c) {
            const long long fr = rs + r, fc = cs + c;
            float acc = 0.0f;
            uint32_t macc = 0;
            for (long long kr = 0; kr < r2; ++kr) {
                const long long ir = orc_map2d(fr - (r2 - 1 - kr), r1, boundary);
                for (long long kc = 0; kc < c2; ++kc) {
                    const long long ic = orc_map2d(fc - (c2 - 1 - kc), c1, boundary);
                    const double x = (ir < 0 || ic < 0) ? fill : in1[ir * c1 + ic];
                    const double h = in2[kr * c2 + kc];
                    if (kind == 0) acc = fmaf((float)x, (float)h, acc);
                    else macc = macc + (uint32_t)(int64_t)x * (uint32_t)(int64_t)h;
                }
            }
            double v;
            if (kind == 0) {
                v = (double)acc;
            } else {
                const uint32_t mask = kind == 32 ? 0xFFFFFFFFu : ((1u << kind) - 1u);
                const uint32_t bits = macc & mask;
                v = (bits & (1u << (kind - 1))) ? (double)((int64_t)bits - ((int64_t)1 << kind)) : (double)bits;
            }
            out[r * cols + c] = v;
        }
    }
    return 0;
}

/* ------------------------------------------------------------------ */
/* ambiguity function                                                  */
/* ------------------------------------------------------------------ */

/* radartools_typed.cpp:31-53 normalized_reference: FP64 energy of the whole
 * (loaded) input, then each sample divided by its square root; a zero norm
 * gives NaN samples.  x is interleaved (re, im) floats, count samples. */
static void orc_amb_normalize(const float *x, size_t count, double *out)
{
    double sum = 0.0;
    for (size_t i = 0; i < count; ++i) {
        const double re = (double)x[2 * i], im = (double)x[2 * i + 1];
        sum += re * re + im * im;
    }
    const double norm = sqrt(sum);
    for (size_t i = 0; i < count; ++i) {
        if (norm == 0.0) {
            out[2 * i] = NAN;
            out[2 * i + 1] = NAN;
        } else {
            out[2 * i] = (double)x[2 * i] / norm;
            out[2 * i + 1] = (double)x[2 * i + 1] / norm;
        }
    }
}

/* complex products in the C99 form (ac - bd, ad + bc) */
static void orc_cmul_d(double ar, double ai, double br, double bi, double *cr, double *ci)
{
    const double ac = ar * br, bd = ai * bi, ad = ar * bi, bc = ai * br;
    *cr = ac - bd;
    *ci = ad + bc;
}

/* radartools_typed.cpp:55-72 direct_transform: sums over the first
 * min(n_in, length) samples; the inverse carries +i and 1/length. */
static void orc_amb_dft(const double *in, int n_in, int length, int inverse, double *out)
{
    const double sign = inverse ? 1.0 : -1.0;
    const double scale = inverse ? 1.0 / (double)length : 1.0;
    for (int f = 0; f < length; ++f) {
        double sr = 0.0, si = 0.0;
        for (int k = 0; k < n_in && k < length; ++k) {
            const double angle = sign * 2.0 * orc_pi_d * (double)f * (double)k / (double)length;
            double tr, ti;
            orc_cmul_d(in[2 * k], in[2 * k + 1], cos(angle), sin(angle), &tr, &ti);
            sr += tr;
            si += ti;
        }
        out[2 * f] = sr * scale;
        out[2 * f + 1] = si * scale;
    }
}

/* radartools_typed.cpp:329-409 ambgfun_typed_cpu.  x: x_count loaded samples
 * of which the first nx take part (the norm covers all of them); y likewise,
 * or NULL for y = x.  cut: 0 = 2-D [rows, nfreq], 1 = delay [nfreq],
 * 2 = Doppler [rows], rows = nx + ny - 1, nfreq = next power of two >= rows.
 * Writes FP64 magnitudes to out.  Returns 0 or a reference error number. */
ORC_API int orc_ambgfun(const float *x, size_t x_count, int nx, const float *y, size_t y_count,
                        int ny, int cut, double fs, double cut_value, double *out)
{
    if (!y) {
        y = x;
        y_count = x_count;
        ny = nx;
    }
    if (nx < 1 || ny < 1 || (size_t)nx > x_count || (size_t)ny > y_count) return 1002;
    if (cut < 0 || cut > 2) return 1002;
    if (cut != 0 && fs == 0.0) return 1002;
    const int rows = nx + ny - 1;
    int nfreq = 1;
    while (nfreq < rows) nfreq <<= 1;
    double *xn = (double *)malloc(sizeof(double) * 2 * x_count);
    double *yn = (double *)malloc(sizeof(double) * 2 * y_count);
    double *a = (double *)calloc((size_t)2 * (size_t)nfreq, sizeof(double));
    double *b = (double *)calloc((size_t)2 * (size_t)nfreq, sizeof(double));
    double *c = (double *)calloc((size_t)2 * (size_t)nfreq, sizeof(double));
    if (!xn || !yn || !a || !b || !c) {
        free(xn); free(yn); free(a); free(b); free(c);
        return 2003;
    }
    orc_amb_normalize(x, x_count, xn);
    orc_amb_normalize(y, y_count, yn);
    if (cut == 0) {
        for (int row = 0; row < rows; ++row) {
            for (int shifted = 0; shifted < nfreq; ++shifted) {
                const int f = (shifted - nfreq / 2 + nfreq) % nfreq;
                double sr = 0.0, si = 0.0;
                for (int col = 0; col < ny; ++col) {
                    const int xc = col - (nx - 1) + row;
                    if (xc < 0 || xc >= nx) continue;
                    const double angle = 2.0 * orc_pi_d * (double)f * (double)col / (double)nfreq;
                    double pr, pi, tr, ti;
                    orc_cmul_d(yn[2 * col], yn[2 * col + 1], xn[2 * xc], -xn[2 * xc + 1], &pr, &pi);
                    orc_cmul_d(pr, pi, cos(angle), sin(angle), &tr, &ti);
                    sr += tr;
                    si += ti;
                }
                out[(size_t)row * (size_t)nfreq + (size_t)shifted] = hypot(sr, si);
            }
        }
    } else if (cut == 1) {
        orc_amb_dft(xn, (int)x_count, nfreq, 0, a);
        for (int k = 0; k < nfreq; ++k) {
            const double freq = -fs / 2.0 + fs * (double)k / (double)nfreq;
            const double angle = 2.0 * orc_pi_d * freq * cut_value;
            double tr, ti;
            orc_cmul_d(a[2 * k], a[2 * k + 1], cos(angle), sin(angle), &tr, &ti);
            a[2 * k] = tr;
            a[2 * k + 1] = ti;
        }
        orc_amb_dft(a, nfreq, nfreq, 1, b);           /* x shifted */
        memset(a, 0, sizeof(double) * 2 * (size_t)nfreq);
        for (int k = 0; k < ny && k < nfreq; ++k)
            orc_cmul_d(yn[2 * k], yn[2 * k + 1], b[2 * k], -b[2 * k + 1], &a[2 * k], &a[2 * k + 1]);
        orc_amb_dft(a, nfreq, nfreq, 1, c);
        for (int i = 0; i < nfreq; ++i) {
            const int src = (i + nfreq / 2) % nfreq;
            out[i] = (double)nfreq * hypot(c[2 * src], c[2 * src + 1]);
        }
    } else {
        double *m = (double *)malloc(sizeof(double) * 2 * x_count);
        double *yf = (double *)malloc(sizeof(double) * 2 * (size_t)rows);
        if (!m || !yf) {
            free(m); free(yf); free(xn); free(yn); free(a); free(b); free(c);
            return 2003;
        }
        memcpy(m, xn, sizeof(double) * 2 * x_count);
        for (int k = 0; k < nx; ++k) {
            const double t = (double)k / fs;
            const double angle = 2.0 * orc_pi_d * cut_value * t;
            double tr, ti;
            orc_cmul_d(m[2 * k], m[2 * k + 1], cos(angle), sin(angle), &tr, &ti);
            m[2 * k] = tr;
            m[2 * k + 1] = ti;
        }
        orc_amb_dft(yn, (int)y_count, rows, 0, yf);
        orc_amb_dft(m, (int)x_count, rows, 0, a);
        for (int k = 0; k < rows; ++k)
            orc_cmul_d(yf[2 * k], yf[2 * k + 1], a[2 * k], -a[2 * k + 1], &b[2 * k], &b[2 * k + 1]);
        orc_amb_dft(b, rows, rows, 1, c);
        for (int i = 0; i < rows; ++i) {
            const int src = (i - rows / 2 + rows) % rows;
            out[i] = hypot(c[2 * src], c[2 * src + 1]);
        }
        free(m);
        free(yf);
    }
    free(xn); free(yn); free(a); free(b); free(c);
    return 0;
}

/* ------------------------------------------------------------------ */
/* CFAR                                                                */
/* ------------------------------------------------------------------ */

/* radartools_typed.cpp:592-603 */
ORC_API double orc_cfar_alpha(double pfa, int reference)
{
    if (reference < 1 || !isfinite(pfa) || pfa <= 0.0 || pfa > 1.0) return NAN;
    return (double)reference * (pow(pfa, -1.0 / (double)reference) - 1.0);
}

/* radartools_typed.cpp:161-220: reference-cell count and the float alpha the
 * threshold kernel uses.  rank is 1 or 2. */
ORC_API int orc_ca_cfar_count(int rank, int g0, int g1, int r0, int r1)
{
    if (rank == 1) return 2 * r0;
    return 2 * r0 * (2 * r1 + 2 * g1 + 1) + 2 * (2 * g0 + 1) * r1;
}

ORC_API float orc_ca_cfar_alpha_f(int rank, int g0, int g1, int r0, int r1, double pfa)
{
    const double n = (double)orc_ca_cfar_count(rank, g0, g1, r0, r1);
    return (float)(n * (pow(pfa, -1.0 / n) - 1.0));
}

/* detection rule (radartools_typed.cpp:222-236).  is_int selects the integer
 * input rule with limits lo/hi of the integer type; xi carries the integer. */
static int orc_detect(float x, int64_t xi, int is_int, double lo, double hi, float thr)
{
    if (isnan(thr)) return 0;
    if (is_int) {
        if (isinf(thr)) return thr < 0.0F;
        if (thr < (float)lo) return 1;
        if (thr >= (float)hi) return 0;
        return xi > (int64_t)floorf(thr);
    }
    return x > thr;
}

/* ca_cfar_typed_cpu (radartools_typed.cpp:465-555), rank 1 or 2.
 * x: [rows, cols] float (cols = 1 for rank 1).  int_vals optional (integer
 * inputs); int_lo/int_hi their type range.  thr float, det uint8. */
ORC_API int orc_ca_cfar(const float *x, const int64_t *int_vals, double int_lo, double int_hi,
                        int rank, int rows, int cols, int g0, int g1, int r0, int r1,
                        double pfa, float *thr, uint8_t *det)
{
    const size_t count = (size_t)rows * (size_t)(rank == 2 ? cols : 1);
    if (rank != 1 && rank != 2) return 1002;
    const int N = orc_ca_cfar_count(rank, g0, g1, r0, r1);
    const float alpha = orc_ca_cfar_alpha_f(rank, g0, g1, r0, r1, pfa);
    float *cum = (float *)malloc(sizeof(float) * (count ? count : 1));
    if (!cum) return 2003;
    for (size_t i = 0; i < count; ++i) thr[i] = 0.0F;
    if (rank == 1) {
        float running = 0.0F;
        for (int i = 0; i < rows; ++i) { running += x[i]; cum[i] = running; }
        const int m = g0 + r0;
        for (int i = m; i < rows - m; ++i) {
            const int orr = i + m, ol = i - m - 1, ir = i + g0, il = i - g0 - 1;
            const float outer = ol >= 0 ? cum[orr] - cum[ol] : cum[orr];
            const float inner = cum[ir] - cum[il];
            thr[i] = (alpha / (float)N) * (outer - inner);
        }
    } else {
        for (int c = 0; c < cols; ++c) {
            float running = 0.0F;
            for (int r = 0; r < rows; ++r) {
                const size_t idx = (size_t)r * cols + c;
                running += x[idx];
                cum[idx] = running;
            }
        }
        for (int r = 0; r < rows; ++r) {
            float running = 0.0F;
            for (int c = 0; c < cols; ++c) {
                const size_t idx = (size_t)r * cols + c;
                running += cum[idx];
                cum[idx] = running;
            }
        }
#define AT(rr, cc) cum[(size_t)(rr) * cols + (cc)]
        const int rm = g0 + r0, cm = g1 + r1;
        for (int r = rm; r < rows - rm; ++r) {
            for (int c = cm; c < cols - cm; ++c) {
                const int ot = r + rm, ol = r - rm - 1, orc = c + cm, olc = c - cm - 1;
                float outer = AT(ot, orc);
                if (ol >= 0) outer -= AT(ol, orc);
                if (olc >= 0) outer -= AT(ot, olc);
                if (ol >= 0 && olc >= 0) outer += AT(ol, olc);
                const float inner = AT(r + g0, c + g1) - AT(r - g0 - 1, c + g1) -
                                    AT(r + g0, c - g1 - 1) + AT(r - g0 - 1, c - g1 - 1);
                thr[(size_t)r * cols + c] = (alpha / (float)N) * (outer - inner);
            }
        }
#undef AT
    }
    for (size_t i = 0; i < count; ++i)
        det[i] = (uint8_t)orc_detect(x[i], int_vals ? int_vals[i] : 0, int_vals != NULL,
                                     int_lo, int_hi, thr[i]);
    free(cum);
    return 0;
}

/* Exact FP64 CA-CFAR for rank 2: direct box sums in double, alpha in double.
 * Border cells get threshold 0 as in the reference. */
ORC_API int orc_ca_cfar_exact(const float *x, int rows, int cols, int g0, int g1, int r0,
                              int r1, double pfa, double *thr)
{
    const int N = orc_ca_cfar_count(2, g0, g1, r0, r1);
    const double alpha = orc_cfar_alpha(pfa, N);
    const int rm = g0 + r0, cm = g1 + r1;
    for (size_t i = 0; i < (size_t)rows * cols; ++i) thr[i] = 0.0;
    for (int r = rm; r < rows - rm; ++r) {
        for (int c = cm; c < cols - cm; ++c) {
            double outer = 0.0, inner = 0.0;
            for (int rr = r - rm; rr <= r + rm; ++rr)
                for (int cc = c - cm; cc <= c + cm; ++cc) {
                    const double v = (double)x[(size_t)rr * cols + cc];
                    outer += v;
                    if (rr >= r - g0 && rr <= r + g0 && cc >= c - g1 && cc <= c + g1) inner += v;
                }
            thr[(size_t)r * cols + c] = alpha / (double)N * (outer - inner);
        }
    }
    return 0;
}

/* ------------------------------------------------------------------ */
/* upfirdn (filtering_typed.cpp:1046-1117, 2548-2616)                   */
/* ------------------------------------------------------------------ */

/* Output samples per line: floor(((n-1)*up + K-1)/down) + 1. */
ORC_API long long orc_upfirdn_out_len(long long n, long long K, int up, int down)
{
    const long long num = (n - 1) * up + K - 1;
    long long q = num / down;
    if (num % down < 0) --q;
    return q + 1;
}

/* Lines laid out as [outer, n, inner] (the filtered axis is the middle one).
 * fp64_acc selects the INT32 path (double products and sums); otherwise the
 * float path sums float products in float. h/x given as double so INT32 data
 * keeps its full value; the float path narrows them to float first. */
ORC_API int orc_upfirdn(const double *h, int K, const double *x, int outer, int n, int inner,
                        int up, int down, int fp64_acc, double *out)
{
    if (K < 1) return 1002;
    if (up < 1 || down < 1) return 1001;
    const long long out_n = orc_upfirdn_out_len(n, K, up, down);
    for (int o = 0; o < outer; ++o) {
        for (int in = 0; in < inner; ++in) {
            const size_t ib = (size_t)o * n * inner + in;
            const size_t ob = (size_t)o * out_n * inner + in;
            for (long long s = 0; s < out_n; ++s) {
                const long long filtered = s * down;
                double dsum = 0.0;
                float fsum = 0.0F;
                for (int t = 0; t < K; ++t) {
                    const long long cand = filtered - t;
                    if (cand % up != 0) continue;
                    const long long src = cand / up;
                    if (src < 0 || src >= n) continue;
                    const double xv = x[ib + (size_t)src * inner];
                    if (fp64_acc) dsum += h[t] * xv;
                    else fsum += (float)h[t] * (float)xv;
                }
                out[ob + (size_t)s * inner] = fp64_acc ? dsum : (double)fsum;
            }
        }
    }
    return 0;
}

/* ------------------------------------------------------------------ */
/* resample_poly (filtering_typed.cpp:104-223, 933-1044, 2617-2715)     */
/* ------------------------------------------------------------------ */

typedef struct {
    int up, down, identity;
    int half_length, coefficient_count;
    int prefix_zeros, output_start, output_count;
} orc_resample_plan;

static int orc_gcd(int a, int b)
{
    while (b) { int t = a % b; a = b; b = t; }
    return a;
}

ORC_API int orc_resample_plan_fill(int n, int up, int down, int custom, int custom_count,
                                   int *fields /* 8 ints */)
{
    if (up < 1 || down < 1) return 1001;
    orc_resample_plan p;
    memset(&p, 0, sizeof(p));
    const int g = orc_gcd(up, down);
    p.up = up / g;
    p.down = down / g;
    p.identity = p.up == 1 && p.down == 1;
    p.output_count = n;
    if (!p.identity) {
        p.output_count = (int)(((long long)n * p.up + p.down - 1) / p.down);
        if (custom) {
            p.coefficient_count = custom_count;
            p.half_length = custom_count == 0 ? -1 : (custom_count - 1) / 2;
        } else {
            const int m = p.up > p.down ? p.up : p.down;
            p.half_length = 10 * m;
            p.coefficient_count = 2 * p.half_length + 1;
        }
        const int rem = ((p.half_length % p.down) + p.down) % p.down;
        p.prefix_zeros = p.down - rem;
        p.output_start = (p.half_length + p.prefix_zeros) / p.down;
    }
    fields[0] = p.up; fields[1] = p.down; fields[2] = p.identity;
    fields[3] = p.half_length; fields[4] = p.coefficient_count;
    fields[5] = p.prefix_zeros; fields[6] = p.output_start; fields[7] = p.output_count;
    return 0;
}

/* filtering_typed.cpp:159-174: Abramowitz-Stegun polynomial I0. */
ORC_API double orc_resample_bessel_i0(double value)
{
    const double a = fabs(value);
    if (a < 3.75) {
        const double ratio = value / 3.75;
        const double y = ratio * ratio;
        return 1.0 + y * (3.5156229 + y * (3.0899424 + y * (1.2067492 + y * (0.2659732 + y *
               (0.0360768 + y * 0.0045813)))));
    }
    const double y = 3.75 / a;
    return exp(a) / sqrt(a) * (0.39894228 + y * (0.01328592 + y * (0.00225319 + y *
           (-0.00157565 + y * (0.00916281 + y * (-0.02057706 + y * (0.02635537 + y *
           (-0.01647633 + y * 0.00392377))))))));
}

/* filtering_typed.cpp:176-218: window_mode 0 kaiser, 1 hamming, 2 explicit. */
ORC_API int orc_resample_design(int up, int down, int window_mode, double beta,
                                const double *design_window, double *coeffs /* 2*half+1 */)
{
    int f[8];
    orc_resample_plan_fill(1, up, down, 0, 0, f);
    const int pu = f[0], pd = f[1], half = f[3], count = f[4];
    double *w = (double *)malloc(sizeof(double) * (size_t)count);
    if (!w) return 2003;
    if (window_mode == 0) {
        const double den = orc_resample_bessel_i0(beta);
        for (int i = 0; i < count; ++i) {
            const double ratio = (double)(i - half) / half;
            double bounded = 1.0 - ratio * ratio;
            if (bounded < 0.0) bounded = 0.0;
            w[i] = orc_resample_bessel_i0(beta * sqrt(bounded)) / den;
        }
    } else if (window_mode == 1) {
        for (int i = 0; i < count; ++i)
            w[i] = 0.54 - 0.46 * cos(6.283185307179586476925286766559 * i / (count - 1));
    } else {
        for (int i = 0; i < count; ++i) w[i] = design_window[i];
    }
    const double cutoff = 1.0 / (pu > pd ? pu : pd);
    double scale = 0.0;
    for (int i = 0; i < count; ++i) {
        const int off = i - half;
        const double sinc = off == 0 ? cutoff
            : sin(3.1415926535897932384626433832795 * cutoff * off) /
              (3.1415926535897932384626433832795 * off);
        coeffs[i] = sinc * w[i];
        scale += coeffs[i];
    }
    for (int i = 0; i < count; ++i) coeffs[i] = coeffs[i] / scale * pu;
    free(w);
    return 0;
}

/* Polyphase core of resample_poly_typed_cpu.  taps: scaled coefficients (as
 * double; the fp32 path narrows them to float and sums in float).  x lines in
 * [outer, n, inner] layout; out [outer, out_n, inner]. */
ORC_API int orc_resample_poly(const double *taps, int K, const double *x, int outer, int n,
                              int inner, int up, int down, int fp64_acc, double *out)
{
    int f[8];
    int rc = orc_resample_plan_fill(n, up, down, 1, K, f);
    if (rc) return rc;
    const int pu = f[0], pd = f[1], prefix = f[5], start = f[6], out_n = f[7];
    for (int o = 0; o < outer; ++o) {
        for (int in = 0; in < inner; ++in) {
            const size_t ib = (size_t)o * n * inner + in;
            const size_t ob = (size_t)o * out_n * inner + in;
            for (int s = 0; s < out_n; ++s) {
                const long long filtered = ((long long)s + start) * pd;
                double dsum = 0.0;
                float fsum = 0.0F;
                for (int t = 0; t < K; ++t) {
                    const long long cand = filtered - prefix - t;
                    if (cand % pu != 0) continue;
                    const long long src = cand / pu;
                    if (src < 0 || src >= n) continue;
                    const double xv = x[ib + (size_t)src * inner];
                    if (fp64_acc) dsum += taps[t] * xv;
                    else fsum += (float)taps[t] * (float)xv;
                }
                out[ob + (size_t)s * inner] = fp64_acc ? dsum : (double)fsum;
            }
        }
    }
    return 0;
}

/* ------------------------------------------------------------------ */
/* firfilter (filtering_typed.cpp:1536-1592)                           */
/* ------------------------------------------------------------------ */

/* y[n] = float(sum_{t<=K-1, 0<=n-t<N} b[t] x[n-t] (+ zi[n] for n < K-1)),
 * accumulated in double in tap order.  Layout [outer, N, inner].  zi (may be
 * NULL) is given already broadcast to [outer, K-1, inner].  zf (may be NULL)
 * gets the full-convolution tail [outer, K-1, inner]. */
ORC_API int orc_firfilter(const double *b, int K, const double *x, int outer, int N, int inner,
                          const double *zi, float *y, float *zf)
{
    if (K < 1) return 1002;
    const int state = K - 1;
    for (int o = 0; o < outer; ++o) {
        for (int in = 0; in < inner; ++in) {
            const size_t base = (size_t)o * N * inner + in;
            const size_t sbase = (size_t)o * state * inner + in;
            const int total = N + (zf ? state : 0);
            for (int full = 0; full < total; ++full) {
                double v = 0.0;
                for (int t = 0; t <= state; ++t) {
                    const int s = full - t;
                    if (s >= 0 && s < N) v += b[t] * x[base + (size_t)s * inner];
                }
                if (zi && full < state) v += zi[sbase + (size_t)full * inner];
                if (full < N) y[base + (size_t)full * inner] = (float)v;
                else zf[sbase + (size_t)(full - N) * inner] = (float)v;
            }
        }
    }
    return 0;
}
