// The reference's radar operator API (ZKX/cusignal_cpp/include/cusignal/
// operators/radartools/radartools_typed.h) on libzkx_radar.so: same option and
// result structs, same function names and overloads, same dtype rules and
// exceptions, for the element types the reference instantiates (float,
// __half, int32, int16, int8; radartools_typed.cpp:810-811).  Inputs are
// loaded to float on the device as the reference does
// (operator_compute_policy.h:42-45,64-67) and run through the same FP32
// kernels; integer CA-CFAR detections use the integer comparison of
// radartools_kernels.cuh:23-37.  The computation is the gfx950 HIP library
// behind the C ABI; nothing here computes on the host except cfar_alpha's
// scalar.  The reference's *_typed_cpu functions are its CPU reference path;
// they are not part of this library.
#pragma once

#include <cmath>
#include <cstdint>
#include <limits>
#include <stdexcept>
#include <string>
#include <type_traits>
#include <vector>

#include "../filtering/filtering_typed.h"

namespace cusignal {

enum class AmbgfunCut { two_dimensional, delay, doppler };
enum class AmbgfunOutputDtype { fp32, fp64 };

// radartools_typed.h:13-23: fs, prf (unused by the result, as in the
// reference), cut, cut_value; the shapes describe higher-rank inputs of the
// 2-D cut, whose first extent selects the samples that take part.
struct AmbgfunOptions {
    double fs = 1.0;
    double prf = 1.0;
    AmbgfunCut cut = AmbgfunCut::two_dimensional;
    double cut_value = 0.0;
    std::vector<int> input_shape;
    std::vector<int> reference_shape;
};

struct AmbgfunDeviceResult {
    AmbgfunOutputDtype dtype{AmbgfunOutputDtype::fp32};
    DeviceArray<float> fp32;
    DeviceArray<double> fp64;
    std::vector<int> shape;
};

enum class CfarDetection : std::uint8_t { no = 0, yes = 1 };

struct CaCfarOptions {
    std::vector<int> guard_cells;
    std::vector<int> reference_cells;
    double pfa = 1.0e-3;
};

struct CaCfarDeviceResult {
    DeviceArray<float> threshold;
    DeviceArray<CfarDetection> detections;
    std::vector<int> shape;
};

enum class RadarWindowKind { none, hann, hamming, explicit_fp32, explicit_fp64 };

struct PulseCompressionOptions {
    int num_pulses{0};
    int samples_per_pulse{0};
    bool normalize{false};
    int nfft{-1};
    RadarWindowKind window{RadarWindowKind::none};
};

struct PulseDopplerOptions {
    int num_pulses{0};
    int samples_per_pulse{0};
    int nfft{-1};
    RadarWindowKind window{RadarWindowKind::none};
};

struct RadarComplexDeviceResult {
    DeviceArray<ComplexFloat> fp32;
    DeviceArray<ComplexDouble> fp64;
    std::vector<int> shape;
    bool is_fp64{false};
};

namespace detail {
// radar_complex_result_is_fp64 (radartools_typed.cpp:271-278)
template <class T>
constexpr bool radar_result_is_fp64(RadarWindowKind w)
{
    if (w == RadarWindowKind::none) return std::is_integral_v<T>;
    if (w == RadarWindowKind::explicit_fp32) return std::is_same_v<T, std::int32_t>;
    return true;
}
// one host value loaded to float (InputTraits<T>::load_continuous)
template <class T>
float host_load(T v)
{
    (void)dtype_code<T>();
    if constexpr (std::is_same_v<T, __half>)
        return __half2float(v);
    else
        return static_cast<float>(v);
}
// ambgfun_first_extent (radartools_typed.cpp:87-106)
inline int ambgfun_first_extent(std::size_t count, const std::vector<int>& shape, const char* label)
{
    if (count > static_cast<std::size_t>(std::numeric_limits<int>::max()))
        throw std::invalid_argument(std::string("ambgfun ") + label + " length exceeds int capacity");
    if (shape.empty()) return static_cast<int>(count);
    std::size_t product = 1;
    for (const int extent : shape) {
        if (extent < 1) throw std::invalid_argument(std::string("ambgfun ") + label + " shape");
        if (product > count / static_cast<std::size_t>(extent))
            throw std::invalid_argument(std::string("ambgfun ") + label + " shape product");
        product *= static_cast<std::size_t>(extent);
    }
    if (product != count) throw std::invalid_argument(std::string("ambgfun ") + label + " data/shape mismatch");
    return shape.front();
}
}  // namespace detail

// radartools_typed.h:134-139 (radartools_typed.cpp:411-462).  All arithmetic
// is FP32 on the device; the 2-D cut of float/half inputs is FP32, every
// other result is FP64 storage of the FP32 values (ambgfun_output_dtype,
// radartools_typed.cpp:119-130).
template <class T>
void ambgfun_device(const DeviceArray<TypedComplex<T>>& x, AmbgfunDeviceResult& output,
                    const AmbgfunOptions& options = {}, const DeviceArray<TypedComplex<T>>* y = nullptr)
{
    (void)detail::dtype_code<T>();
    if (options.cut != AmbgfunCut::two_dimensional && options.cut != AmbgfunCut::delay &&
        options.cut != AmbgfunCut::doppler)
        throw std::invalid_argument("ambgfun cut is invalid");
    if (options.cut != AmbgfunCut::two_dimensional && options.fs == 0.0)
        throw std::invalid_argument("ambgfun fs must be nonzero for one-dimensional cuts");
    if (y == nullptr && !options.reference_shape.empty())
        throw std::invalid_argument("ambgfun reference_shape requires y");
    const bool high_rank = options.input_shape.size() > 1 || (y != nullptr && options.reference_shape.size() > 1);
    if (high_rank && options.cut != AmbgfunCut::two_dimensional)
        throw std::invalid_argument("ambgfun high-rank input is only defined for cut='2d'");
    const DeviceArray<TypedComplex<T>>& reference = y == nullptr ? x : *y;
    if (x.empty() || reference.empty()) throw std::invalid_argument("ambgfun inputs must be nonempty");
    const int nx = detail::ambgfun_first_extent(x.size(), options.input_shape, "input");
    const int ny = y == nullptr ? nx : detail::ambgfun_first_extent(reference.size(), options.reference_shape, "reference");
    if (nx > std::numeric_limits<int>::max() - ny + 1) throw std::invalid_argument("ambgfun input length exceeds int capacity");
    const int cut = static_cast<int>(options.cut);
    const long long count = zkx_ambgfun_out_len(nx, ny, cut);
    if (count < 0) throw std::invalid_argument("ambgfun FFT length exceeds int capacity");
    const bool fp32 = options.cut == AmbgfunCut::two_dimensional &&
                      (std::is_same_v<T, float> || std::is_same_v<T, __half>);
    const detail::FloatSamples xf = detail::as_float(x);
    const detail::FloatSamples yf = detail::as_float(reference);
    AmbgfunDeviceResult result;
    result.dtype = fp32 ? AmbgfunOutputDtype::fp32 : AmbgfunOutputDtype::fp64;
    const long long rows = static_cast<long long>(nx) + ny - 1;
    if (options.cut == AmbgfunCut::two_dimensional)
        result.shape = {static_cast<int>(rows), static_cast<int>(count / rows)};
    else
        result.shape = {static_cast<int>(count)};
    void* dst = nullptr;
    if (fp32) {
        result.fp32 = DeviceArray<float>(static_cast<std::size_t>(count));
        dst = result.fp32.data();
    } else {
        result.fp64 = DeviceArray<double>(static_cast<std::size_t>(count));
        dst = result.fp64.data();
    }
    detail::check_abi(zkx_ambgfun_c64(reinterpret_cast<const zkx_complex64*>(xf.data()), x.size(), nx,
                                      y == nullptr ? nullptr : reinterpret_cast<const zkx_complex64*>(yf.data()),
                                      reference.size(), ny, cut, options.fs, options.cut_value, fp32 ? 0 : 1, dst,
                                      nullptr));
    cuda_utils::synchronize_stream();
    output = std::move(result);
}

// radartools_typed.h:237-262
template <class T>
RadarComplexDeviceResult pulse_compression_device(const DeviceArray<TypedComplex<T>>& x,
                                                  const DeviceArray<TypedComplex<T>>& pulse_template,
                                                  const PulseCompressionOptions& options,
                                                  const DeviceArray<float>* explicit_window = nullptr)
{
    const int nfft = options.nfft == -1 ? options.samples_per_pulse : options.nfft;
    if (options.num_pulses < 1 || options.samples_per_pulse < 1 || nfft < 1 ||
        x.size() != static_cast<std::size_t>(options.num_pulses) * options.samples_per_pulse || pulse_template.empty())
        throw std::invalid_argument("pulse_compression shape is invalid");
    const bool needs_window = options.window == RadarWindowKind::explicit_fp32 ||
                              options.window == RadarWindowKind::explicit_fp64;
    if (needs_window != (explicit_window != nullptr) ||
        (explicit_window && explicit_window->size() != pulse_template.size()))
        throw std::invalid_argument("pulse_compression window is invalid");
    const detail::FloatSamples xf = detail::as_float(x), tf = detail::as_float(pulse_template);
    DeviceArray<ComplexFloat> computed(static_cast<std::size_t>(options.num_pulses) * nfft);
    detail::check_abi(zkx_pulse_compression_c64(
        reinterpret_cast<const zkx_complex64*>(xf.data()), options.num_pulses, options.samples_per_pulse,
        reinterpret_cast<const zkx_complex64*>(tf.data()), static_cast<int>(pulse_template.size()), nfft,
        options.normalize ? 1 : 0, static_cast<int>(options.window), explicit_window ? explicit_window->data() : nullptr,
        reinterpret_cast<zkx_complex64*>(computed.data()), nullptr));
    RadarComplexDeviceResult result;
    result.shape = {options.num_pulses, nfft};
    result.is_fp64 = detail::radar_result_is_fp64<T>(options.window);
    if (result.is_fp64) result.fp64 = detail::widen(computed);
    else result.fp32 = std::move(computed);
    return result;
}

// radartools_typed.h:304-311 (preallocated FP32 output, no widening)
template <class T>
void pulse_doppler_device(const DeviceArray<TypedComplex<T>>& x, const PulseDopplerOptions& options,
                          DeviceArray<ComplexFloat>& output, const DeviceArray<float>* explicit_window = nullptr);

// Extension of the preallocated overload: also writes |output|^2 into
// `power` from the same kernel, replacing the separate power_kernel the
// reference's step 4 launches (Task1/pipeline/src/steps/step4/step4.cu:11-19,58-61).
template <class T>
void pulse_doppler_device(const DeviceArray<TypedComplex<T>>& x, const PulseDopplerOptions& options,
                          DeviceArray<ComplexFloat>& output, DeviceArray<float>& power,
                          const DeviceArray<float>* explicit_window = nullptr)
{
    const int nfft = options.nfft == -1 ? options.num_pulses : options.nfft;
    if (options.num_pulses < 1 || options.samples_per_pulse < 1 || nfft < 1 ||
        x.size() != static_cast<std::size_t>(options.num_pulses) * options.samples_per_pulse)
        throw std::invalid_argument("pulse_doppler shape is invalid");
    const bool needs_window = options.window == RadarWindowKind::explicit_fp32 ||
                              options.window == RadarWindowKind::explicit_fp64;
    if (needs_window != (explicit_window != nullptr) ||
        (explicit_window && explicit_window->size() != static_cast<std::size_t>(options.num_pulses)))
        throw std::invalid_argument("pulse_doppler window is invalid");
    const std::size_t count = static_cast<std::size_t>(nfft) * options.samples_per_pulse;
    if (output.size() != count || power.size() != count)
        throw std::invalid_argument("pulse_doppler preallocated output shape mismatch");
    const detail::FloatSamples xf = detail::as_float(x);
    detail::check_abi(zkx_pulse_doppler_c64(reinterpret_cast<const zkx_complex64*>(xf.data()), options.num_pulses,
                                            options.samples_per_pulse, nfft, static_cast<int>(options.window),
                                            explicit_window ? explicit_window->data() : nullptr,
                                            reinterpret_cast<zkx_complex64*>(output.data()), power.data(), nullptr));
}

template <class T>
void pulse_doppler_device(const DeviceArray<TypedComplex<T>>& x, const PulseDopplerOptions& options,
                          DeviceArray<ComplexFloat>& output, const DeviceArray<float>* explicit_window)
{
    const int nfft = options.nfft == -1 ? options.num_pulses : options.nfft;
    if (options.num_pulses < 1 || options.samples_per_pulse < 1 || nfft < 1 ||
        x.size() != static_cast<std::size_t>(options.num_pulses) * options.samples_per_pulse)
        throw std::invalid_argument("pulse_doppler shape is invalid");
    const bool needs_window = options.window == RadarWindowKind::explicit_fp32 ||
                              options.window == RadarWindowKind::explicit_fp64;
    if (needs_window != (explicit_window != nullptr) ||
        (explicit_window && explicit_window->size() != static_cast<std::size_t>(options.num_pulses)))
        throw std::invalid_argument("pulse_doppler window is invalid");
    if (output.size() != static_cast<std::size_t>(nfft) * options.samples_per_pulse)
        throw std::invalid_argument("pulse_doppler preallocated output shape mismatch");
    const detail::FloatSamples xf = detail::as_float(x);
    detail::check_abi(zkx_pulse_doppler_c64(reinterpret_cast<const zkx_complex64*>(xf.data()), options.num_pulses,
                                            options.samples_per_pulse, nfft, static_cast<int>(options.window),
                                            explicit_window ? explicit_window->data() : nullptr,
                                            reinterpret_cast<zkx_complex64*>(output.data()), nullptr, nullptr));
}

// radartools_typed.h:276-289
template <class T>
RadarComplexDeviceResult pulse_doppler_device(const DeviceArray<TypedComplex<T>>& x,
                                              const PulseDopplerOptions& options,
                                              const DeviceArray<float>* explicit_window = nullptr)
{
    const int nfft = options.nfft == -1 ? options.num_pulses : options.nfft;
    if (nfft < 1 || options.samples_per_pulse < 1) throw std::invalid_argument("pulse_doppler shape is invalid");
    DeviceArray<ComplexFloat> computed(static_cast<std::size_t>(nfft) * options.samples_per_pulse);
    pulse_doppler_device(x, options, computed, explicit_window);
    RadarComplexDeviceResult result;
    result.shape = {nfft, options.samples_per_pulse};
    result.is_fp64 = detail::radar_result_is_fp64<T>(options.window);
    if (result.is_fp64) result.fp64 = detail::widen(computed);
    else result.fp32 = std::move(computed);
    return result;
}

// radartools_typed.h:207-222 (radartools_typed.cpp:592-617): the CPU form
// in FP64 from the pfa loaded to float; the device form in FP32 on the
// device, stored as double.
template <class T>
double cfar_alpha_typed_cpu(T pfa, int reference)
{
    double alpha = 0.0;
    detail::check_abi(zkx_cfar_alpha(static_cast<double>(detail::host_load(pfa)), reference, &alpha));
    return alpha;
}

template <class T>
DeviceArray<double> cfar_alpha_device(T pfa, int reference)
{
    const float p = detail::host_load(pfa);
    if (reference < 1 || !std::isfinite(p) || p <= 0.0f || p > 1.0f)
        throw std::invalid_argument("cfar_alpha requires reference >= 1 and pfa in (0,1]");
    DeviceArray<double> out(1);
    detail::check_abi(zkx_cfar_alpha_f32_device(p, reference, out.data(), nullptr));
    cuda_utils::synchronize_stream();
    return out;
}

// radartools_typed.h:182-195 (preallocated outputs)
template <class T>
void ca_cfar_device(const DeviceArray<T>& x, const std::vector<int>& shape, const CaCfarOptions& options,
                    DeviceArray<float>& threshold, DeviceArray<CfarDetection>& detections)
{
    (void)detail::dtype_code<T>();
    if (shape.size() > 2) throw std::invalid_argument("ca_cfar supports only scalar, 1D, or 2D input");
    if (threshold.size() != x.size() || detections.size() != x.size())
        throw std::invalid_argument("ca_cfar preallocated output shape must match input");
    if (shape.empty()) {
        if (x.size() != 1) throw std::invalid_argument("ca_cfar scalar input must contain one element");
        const float v = detail::host_load(x.to_host()[0]);
        threshold = DeviceArray<float>::from_host({0.0f});
        detections = DeviceArray<CfarDetection>::from_host({v > 0.0f ? CfarDetection::yes : CfarDetection::no});
        return;
    }
    if (options.guard_cells.size() != shape.size() || options.reference_cells.size() != shape.size())
        throw std::invalid_argument("ca_cfar cell parameter rank mismatch");
    const detail::FloatSamples xf = detail::as_float(x);
    detail::check_abi(zkx_ca_cfar_f32(xf.data(), static_cast<int>(shape.size()), shape.data(),
                                      options.guard_cells.data(), options.reference_cells.data(), options.pfa,
                                      threshold.data(), reinterpret_cast<uint8_t*>(detections.data()), nullptr,
                                      nullptr));
    if constexpr (std::is_integral_v<T>)
        detail::check_abi(zkx_cfar_detect_typed(x.data(), detail::dtype_code<T>(), threshold.data(), x.size(),
                                                reinterpret_cast<uint8_t*>(detections.data()), nullptr));
}

// radartools_typed.h:154-170
template <class T>
CaCfarDeviceResult ca_cfar_device(const DeviceArray<T>& x, const std::vector<int>& shape, const CaCfarOptions& options)
{
    CaCfarDeviceResult result;
    result.shape = shape;
    result.threshold = DeviceArray<float>(x.size());
    result.detections = DeviceArray<CfarDetection>(x.size());
    ca_cfar_device(x, shape, options, result.threshold, result.detections);
    return result;
}

}  // namespace cusignal
