// Synthetic radar echo source for gfx950: the step before the radar path.
//
// Follows Task1 step 1 (ZKX/Task1/pipeline/src/steps/step1/step1.cu:11-71):
//   waveform[n] = exp(i pi (B/tau) t^2), t = n/fs - tau/2, tau = T/fs   (n < T)
//   noiseless[p,s] = A * waveform[s - d] * exp(2 pi i fd (p/prf + s/fs))  (0 <= s-d < T, else 0)
//   noise[p,s] = (u(2i), u(2i+1)), u(j) = ((h(j, seed) & 0xffff) / 32767.5 - 1) * sigma,
//   i = p*S + s, h the 32-bit integer mix of step1.cu:11-19
//   echo = noiseless + noise
// all in FP32 with the reference's operation order; the noise is integer
// arithmetic plus one correctly rounded division, so it is reproduced bit for
// bit.  One launch writes all four outputs (the reference runs two kernels);
// the waveform value an echo sample needs is recomputed from its index with
// the same expression that produces waveform[], so both agree.
#include "runtime.hpp"

#include <cmath>
#include <limits>

namespace {

constexpr float kPiF = 3.14159265358979323846F;

__device__ __forceinline__ uint32_t noise_bits(uint32_t index, uint32_t seed)
{
    uint32_t v = seed ^ (index * 747796405u + 2891336453u);
    v ^= v >> 16;
    v *= 2246822519u;
    v ^= v >> 13;
    v *= 3266489917u;
    return v ^ (v >> 16);
}

__device__ __forceinline__ float uniform_noise(uint32_t index, uint32_t seed, float sigma)
{
    const float q = __fdiv_rn(static_cast<float>(noise_bits(index, seed) & 0xffffu), 32767.5F);
    return __fmul_rn(__fsub_rn(q, 1.0F), sigma);
}

__device__ __forceinline__ float2 lfm_sample(int n, float width, float fs, float bw)
{
    const float t = __fsub_rn(__fdiv_rn(static_cast<float>(n), fs), __fmul_rn(0.5F, width));
    const float phase = __fmul_rn(__fmul_rn(__fmul_rn(kPiF, __fdiv_rn(bw, width)), t), t);
    return make_float2(cosf(phase), sinf(phase));
}

__global__ void __launch_bounds__(256)
echo_kernel(zkx_echo_params p, float2* __restrict__ waveform, float2* __restrict__ noiseless,
            float2* __restrict__ noise, float2* __restrict__ echo, long long count)
{
    const float width = __fdiv_rn(static_cast<float>(p.pulse_samples), p.sample_rate_hz);
    for (long long g = blockIdx.x * static_cast<long long>(blockDim.x) + threadIdx.x; g < count;
         g += static_cast<long long>(gridDim.x) * blockDim.x) {
        const int i = static_cast<int>(g);
        if (waveform && i < p.pulse_samples)
            waveform[i] = lfm_sample(i, width, p.sample_rate_hz, p.bandwidth_hz);
        const int pulse = i / p.samples_per_pulse;
        const int sample = i - pulse * p.samples_per_pulse;
        const int src = sample - p.target_delay_samples;
        float tr = 0.0F, ti = 0.0F;
        if (src >= 0 && src < p.pulse_samples) {
            const float2 b = lfm_sample(src, width, p.sample_rate_hz, p.bandwidth_hz);
            const float time = __fadd_rn(__fdiv_rn(static_cast<float>(pulse), p.prf_hz),
                                         __fdiv_rn(static_cast<float>(sample), p.sample_rate_hz));
            const float phase = __fmul_rn(__fmul_rn(__fmul_rn(2.0F, kPiF), p.doppler_hz), time);
            const float c = cosf(phase), s = sinf(phase);
            tr = __fmul_rn(p.target_amplitude, b.x * c - b.y * s);
            ti = __fmul_rn(p.target_amplitude, b.x * s + b.y * c);
        }
        const uint32_t base = 2u * static_cast<uint32_t>(i);
        const float nr = uniform_noise(base, p.noise_seed, p.noise_std);
        const float ni = uniform_noise(base + 1u, p.noise_seed, p.noise_std);
        if (noiseless) noiseless[i] = make_float2(tr, ti);
        if (noise) noise[i] = make_float2(nr, ni);
        echo[i] = make_float2(__fadd_rn(tr, nr), __fadd_rn(ti, ni));
    }
}

}  // namespace

extern "C" {

int zkx_echo_source_c64(const zkx_echo_params* params, zkx_complex64* waveform, zkx_complex64* noiseless,
                        zkx_complex64* noise, zkx_complex64* echo, void* stream)
{
    zkx::clear_last_error();
    if (params == nullptr || echo == nullptr) ZKX_FAIL(ZKX_INVALID_ARGUMENT, "echo source pointers must be non-null");
    const zkx_echo_params p = *params;
    if (p.num_pulses < 1 || p.samples_per_pulse < 1)
        ZKX_FAIL(ZKX_INVALID_SHAPE, "echo source needs num_pulses and samples_per_pulse >= 1");
    const long long count = static_cast<long long>(p.num_pulses) * p.samples_per_pulse;
    // the noise hash indexes 2*i in 32 bits and the sample index is an int
    if (count > (1LL << 31) - 1) ZKX_FAIL(ZKX_INVALID_SHAPE, "echo source input exceeds 2^31 - 1 samples");
    if (p.pulse_samples < 1 || p.pulse_samples > p.samples_per_pulse)
        ZKX_FAIL(ZKX_INVALID_ARGUMENT, "echo source pulse_samples must lie within the receive window");
    if (p.target_delay_samples < 0 || p.target_delay_samples + p.pulse_samples > p.samples_per_pulse)
        ZKX_FAIL(ZKX_INVALID_ARGUMENT, "echo source target delay moves the pulse out of the receive window");
    if (!(p.sample_rate_hz > 0 && p.prf_hz > 0 && p.bandwidth_hz > 0 && p.target_amplitude > 0 &&
          p.noise_std >= 0 && std::isfinite(p.doppler_hz) && std::isfinite(p.sample_rate_hz) &&
          std::isfinite(p.prf_hz) && std::isfinite(p.bandwidth_hz) && std::isfinite(p.target_amplitude) &&
          std::isfinite(p.noise_std)))
        ZKX_FAIL(ZKX_INVALID_ARGUMENT, "echo source rate, amplitude or noise parameters are invalid");
    long long blocks = (count + 255) / 256;
    if (blocks > 65536) blocks = 65536;
    hipLaunchKernelGGL(echo_kernel, dim3(static_cast<unsigned>(blocks)), dim3(256), 0,
                       static_cast<hipStream_t>(stream), p, reinterpret_cast<float2*>(waveform),
                       reinterpret_cast<float2*>(noiseless), reinterpret_cast<float2*>(noise),
                       reinterpret_cast<float2*>(echo), count);
    ZKX_LAUNCH_CHECK();
    return ZKX_SUCCESS;
}

}  // extern "C"
