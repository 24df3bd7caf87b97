// The reference's polyphase filtering operator API
// (ZKX/cusignal_cpp/include/cusignal/operators/filtering/filtering_typed.h)
// on libzkx_radar.so: upfirdn_device, resample_poly_device, firfilter_device
// and firfilter_fp32_resident_device with the reference's option/result
// structs, shape/axis rules, output dtypes and exceptions, for the element
// types it instantiates (float, __half, int32, int16, int8;
// filtering_typed.cu:683+).  Non-float inputs are loaded to float on the
// device as the reference does (convert_signal_kernel,
// filtering_kernels.cuh:1277-1286) and filtered by the same FP32 kernels.
// Extensions: TypedComplex<float> signals (real taps act on each component)
// for the three filters, and fftconvolve_device.
// Filtering runs in the gfx950 kernels behind the C ABI; the host only plans
// lines, designs the default resample filter (FP64, as the reference does on
// its host) and broadcasts zi.
#pragma once

#include <hip/hip_fp16.h>

#include <cmath>
#include <cstdint>
#include <limits>
#include <numeric>
#include <optional>
#include <stdexcept>
#include <type_traits>
#include <vector>

#include "../../backends/fft/fft_interface.h"
#include "../../runtime/device_array.h"

namespace cusignal {

// filtering_typed.h:13
template <typename T>
struct TypedComplex {
    T real;
    T imag;
};

enum class ResamplePolyWindowMode { kaiser, hamming, explicit_values };
enum class ResamplePolyOutputDtype { fp16, fp32, fp64, int32, int16, int8 };
enum class UpfirdnOutputDtype { fp32, fp64 };

struct UpfirdnOptions {
    std::vector<int> shape;
    int axis = -1;
};

struct UpfirdnDeviceResult {
    UpfirdnOutputDtype dtype{UpfirdnOutputDtype::fp32};
    DeviceArray<float> fp32;
    DeviceArray<double> fp64;
    std::vector<int> shape;
};

struct ResamplePolyOptions {
    std::vector<int> shape;
    int axis = 0;
    ResamplePolyWindowMode window_mode = ResamplePolyWindowMode::kaiser;
    double kaiser_beta = 5.0;
    std::vector<double> design_window;
    bool gpupath = true;
};

struct ResamplePolyDeviceResult {
    ResamplePolyOutputDtype dtype{ResamplePolyOutputDtype::fp64};
    DeviceArray<__half> fp16;
    DeviceArray<float> fp32;
    DeviceArray<double> fp64;
    DeviceArray<std::int32_t> int32;
    DeviceArray<std::int16_t> int16;
    DeviceArray<std::int8_t> int8;
    std::vector<int> shape;
};

struct FirfilterOptions {
    std::vector<int> shape;
    int axis = -1;
    std::vector<int> zi_shape;
};

struct FirfilterDeviceResult {
    DeviceArray<float> y;
    DeviceArray<float> zf;
    std::vector<int> y_shape;
    std::vector<int> zf_shape;
    bool has_zf = false;
};

// filtering_typed.h:144-165
enum class DecimateOutputDtype { fp32, fp64 };

struct DecimateOptions {
    std::vector<int> shape;
    int axis = -1;
    std::optional<int> filter_order;
    bool zero_phase = true;
};

struct DecimateDeviceResult {
    DecimateOutputDtype dtype{DecimateOutputDtype::fp64};
    DeviceArray<float> fp32;
    DeviceArray<double> fp64;
    std::vector<int> shape;
};

// filtering_typed.h:55-58
struct ChannelizePolyDeviceResult {
    DeviceArray<ComplexFloat> values;
    std::vector<int> shape;
};

// Results of the complex extension: the same fields with complex storage.
struct UpfirdnComplexDeviceResult {
    UpfirdnOutputDtype dtype{UpfirdnOutputDtype::fp32};
    DeviceArray<ComplexFloat> fp32;
    DeviceArray<ComplexDouble> fp64;
    std::vector<int> shape;
};

struct ResamplePolyComplexDeviceResult {
    ResamplePolyOutputDtype dtype{ResamplePolyOutputDtype::fp64};
    DeviceArray<ComplexFloat> fp32;
    DeviceArray<ComplexDouble> fp64;
    std::vector<int> shape;
};

struct FirfilterComplexDeviceResult {
    DeviceArray<ComplexFloat> y;
    DeviceArray<ComplexFloat> zf;
    std::vector<int> y_shape;
    std::vector<int> zf_shape;
    bool has_zf = false;
};

// fftconvolve (extension; cuSignal convolve.py:159-320): every row of in1
// (the last axis of `shape`, or a 1-D signal when shape is empty) convolved
// with the 1-D in2; modes as cuSignal/scipy.
enum class FftconvolveMode { full, same, valid };

struct FftconvolveOptions {
    std::vector<int> shape;
    FftconvolveMode mode = FftconvolveMode::full;
};

struct FftconvolveDeviceResult {
    DeviceArray<ComplexFloat> out;
    std::vector<int> shape;
};

namespace detail {

// The element types the reference instantiates, as the C ABI's codes.
template <class T>
constexpr int dtype_code()
{
    if constexpr (std::is_same_v<T, float>) return ZKX_DTYPE_F32;
    else if constexpr (std::is_same_v<T, __half>) return ZKX_DTYPE_F16;
    else if constexpr (std::is_same_v<T, std::int32_t>) return ZKX_DTYPE_I32;
    else if constexpr (std::is_same_v<T, std::int16_t>) return ZKX_DTYPE_I16;
    else {
        static_assert(std::is_same_v<T, std::int8_t>, "element type must be float, __half, int32, int16 or int8");
        return ZKX_DTYPE_I8;
    }
}

// Float samples of `scalars` device values of type T: the values themselves
// for float, a converted device copy otherwise.
class FloatSamples {
public:
    template <class T>
    FloatSamples(const T* data, std::size_t scalars)
    {
        if constexpr (std::is_same_v<T, float>) {
            ptr_ = data;
        } else {
            (void)dtype_code<T>();
            owned_ = DeviceArray<float>(scalars);
            if (scalars) check_abi(zkx_convert_to_f32(data, dtype_code<T>(), scalars, owned_.data(), nullptr));
            ptr_ = owned_.data();
        }
    }
    const float* data() const { return ptr_; }

private:
    const float* ptr_ = nullptr;
    DeviceArray<float> owned_;
};

template <class T>
FloatSamples as_float(const DeviceArray<T>& a)
{
    return FloatSamples(a.data(), a.size());
}
template <class T>
FloatSamples as_float(const DeviceArray<TypedComplex<T>>& a)
{
    return FloatSamples(reinterpret_cast<const T*>(a.data()), 2 * a.size());
}

struct Lines {
    int outer = 1, n = 1, inner = 1, axis = 0;
    std::vector<int> shape;
};

// Empty shape means a 1-D signal of `count` samples (filtering_typed.cpp:877-931).
inline Lines plan_lines(const std::vector<int>& shape_in, int axis, std::size_t count, const char* op)
{
    Lines l;
    l.shape = shape_in.empty() ? std::vector<int>{static_cast<int>(count)} : shape_in;
    const int rank = static_cast<int>(l.shape.size());
    long long total = 1;
    for (int d : l.shape) {
        if (d < 1) throw std::invalid_argument(std::string(op) + " shape must be positive");
        total *= d;
    }
    if (static_cast<std::size_t>(total) != count)
        throw std::invalid_argument(std::string(op) + " shape does not match input size");
    const int ax = axis < 0 ? axis + rank : axis;
    if (ax < 0 || ax >= rank) throw std::invalid_argument(std::string(op) + " axis is outside rank");
    l.axis = ax;
    for (int i = 0; i < ax; ++i) l.outer *= l.shape[i];
    l.n = l.shape[ax];
    for (int i = ax + 1; i < rank; ++i) l.inner *= l.shape[i];
    return l;
}

inline DeviceArray<double> widen(const DeviceArray<float>& v)
{
    DeviceArray<double> out(v.size());
    if (!v.empty()) check_abi(zkx_widen_f32_to_f64(v.data(), out.data(), v.size(), nullptr));
    return out;
}

inline DeviceArray<ComplexDouble> widen(const DeviceArray<ComplexFloat>& v)
{
    DeviceArray<ComplexDouble> out(v.size());
    if (!v.empty())
        check_abi(zkx_widen_c64_to_c128(reinterpret_cast<const zkx_complex64*>(v.data()),
                                        reinterpret_cast<zkx_complex128*>(out.data()), v.size(), nullptr));
    return out;
}

template <class E>
void copy_device(E* dst, const E* src, std::size_t count)
{
    if (count)
        cuda_utils::check_hip(hipMemcpy(dst, src, sizeof(E) * count, hipMemcpyDeviceToDevice), "hipMemcpy D2D");
}

// Abramowitz & Stegun 9.8.1 / 9.8.2 polynomial I0, the approximation of the
// reference's default design (filtering_typed.cpp:159-174).
inline double bessel_i0_as(double v)
{
    const double a = std::fabs(v);
    if (a < 3.75) {
        const double y = (v / 3.75) * (v / 3.75);
        return 1.0 + y * (3.5156229 + y * (3.0899424 + y * (1.2067492 + y * (0.2659732 + y * (0.0360768 + y * 0.0045813)))));
    }
    const double y = 3.75 / a;
    return std::exp(a) / std::sqrt(a) *
           (0.39894228 + y * (0.01328592 + y * (0.00225319 + y * (-0.00157565 + y * (0.00916281 + y * (-0.02057706 +
            y * (0.02635537 + y * (-0.01647633 + y * 0.00392377))))))));
}

// Windowed-sinc low-pass of 2*10*max(up,down)+1 taps at cutoff 1/max(up,down),
// unit DC gain, times up; up/down already gcd-reduced
// (filtering_typed.cpp:176-218).
inline std::vector<double> design_resample_poly(int up, int down, const ResamplePolyOptions& o)
{
    const int half = 10 * std::max(up, down);
    const int count = 2 * half + 1;
    std::vector<double> w(count);
    if (o.window_mode == ResamplePolyWindowMode::kaiser) {
        const double den = bessel_i0_as(o.kaiser_beta);
        for (int i = 0; i < count; ++i) {
            const double r = static_cast<double>(i - half) / half;
            w[i] = bessel_i0_as(o.kaiser_beta * std::sqrt(std::max(0.0, 1.0 - r * r))) / den;
        }
    } else if (o.window_mode == ResamplePolyWindowMode::hamming) {
        for (int i = 0; i < count; ++i) w[i] = 0.54 - 0.46 * std::cos(2.0 * M_PI * i / (count - 1));
    } else {
        if (o.design_window.size() != static_cast<std::size_t>(count))
            throw std::invalid_argument("resample_poly design window must have 2*10*max(up,down)+1 values");
        w = o.design_window;
    }
    const double cutoff = 1.0 / std::max(up, down);
    std::vector<double> h(count);
    double sum = 0.0;
    for (int i = 0; i < count; ++i) {
        const int off = i - half;
        const double s = off == 0 ? cutoff : std::sin(M_PI * cutoff * off) / (M_PI * off);
        h[i] = s * w[i];
        sum += h[i];
    }
    for (double& v : h) v = v / sum * up;
    return h;
}

// resample_poly_output_dtype (filtering_typed.cpp:62-101)
template <typename T>
ResamplePolyOutputDtype resample_poly_dtype(bool identity, bool custom, int reduced_up)
{
    if (identity) {
        if constexpr (std::is_same_v<T, float>) return ResamplePolyOutputDtype::fp32;
        else if constexpr (std::is_same_v<T, __half>) return ResamplePolyOutputDtype::fp16;
        else if constexpr (std::is_same_v<T, std::int32_t>) return ResamplePolyOutputDtype::int32;
        else if constexpr (std::is_same_v<T, std::int16_t>) return ResamplePolyOutputDtype::int16;
        else return ResamplePolyOutputDtype::int8;
    }
    if (!custom) return ResamplePolyOutputDtype::fp64;
    if constexpr (std::is_same_v<T, std::int32_t>) return ResamplePolyOutputDtype::fp64;
    if constexpr (std::is_same_v<T, std::int16_t> || std::is_same_v<T, std::int8_t>)
        return reduced_up < 32768 ? ResamplePolyOutputDtype::fp32 : ResamplePolyOutputDtype::fp64;
    return reduced_up < 65536 ? ResamplePolyOutputDtype::fp32 : ResamplePolyOutputDtype::fp64;
}

// Taps of one resample_poly call as float on the device: custom coefficients
// scaled by the reduced up in their own type, or the FP64 design rounded.
template <class T>
DeviceArray<float> resample_poly_taps(int u, int d, const ResamplePolyOptions& options, const DeviceArray<T>* coefficients)
{
    if (coefficients) {
        if (coefficients->empty()) throw std::invalid_argument("resample_poly coefficients must be non-empty");
        DeviceArray<float> taps(coefficients->size());
        check_abi(zkx_resample_poly_scale_taps(coefficients->data(), dtype_code<T>(),
                                               static_cast<int>(coefficients->size()), u, taps.data(), nullptr));
        return taps;
    }
    const std::vector<double> h = design_resample_poly(u, d, options);
    return DeviceArray<float>::from_host(std::vector<float>(h.begin(), h.end()));
}

// zi broadcast to the state shape (shape with axis length K-1); the given
// zi_shape may hold 1 on any non-axis dimension.  E: the element type.
template <class E>
DeviceArray<E> broadcast_zi(const DeviceArray<E>& zi, const std::vector<int>& zi_shape, const std::vector<int>& state,
                            int axis)
{
    const std::vector<int> zs = zi_shape.empty() ? std::vector<int>{static_cast<int>(zi.size())} : zi_shape;
    if (zs.size() != state.size() && !(zs.size() == 1 && state.size() == 1))
        throw std::invalid_argument("firfilter zi rank must match x");
    long long zcount = 1;
    std::size_t state_count = 1;
    for (std::size_t i = 0; i < state.size(); ++i) state_count *= static_cast<std::size_t>(state[i]);
    for (std::size_t i = 0; i < zs.size(); ++i) {
        if (zs[i] != state[i] && !(zs[i] == 1 && static_cast<int>(i) != axis))
            throw std::invalid_argument("firfilter zi shape cannot broadcast to the state shape");
        zcount *= zs[i];
    }
    if (static_cast<std::size_t>(zcount) != zi.size()) throw std::invalid_argument("firfilter zi size mismatch");
    DeviceArray<E> full(state_count);
    if (zi.size() == state_count) {
        copy_device(full.data(), zi.data(), state_count);
        return full;
    }
    const std::vector<E> zh = zi.to_host();
    std::vector<E> host(state_count);
    std::vector<int> idx(state.size(), 0);
    for (std::size_t e = 0; e < state_count; ++e) {
        std::size_t src = 0;
        for (std::size_t dd = 0; dd < state.size(); ++dd) src = src * zs[dd] + (zs[dd] == 1 ? 0 : idx[dd]);
        host[e] = zh[src];
        for (int dd = static_cast<int>(state.size()) - 1; dd >= 0; --dd) {
            if (++idx[dd] < state[dd]) break;
            idx[dd] = 0;
        }
    }
    return DeviceArray<E>::from_host(host);
}

}  // namespace detail

// filtering_typed.h:744-748: FP32 result, FP64 storage for int32 input
// (filtering_typed.cpp:77-83, filtering_typed.cu:596-645).
template <typename T>
void upfirdn_device(const DeviceArray<T>& h, const DeviceArray<T>& x, int up, int down, UpfirdnDeviceResult& out,
                    const UpfirdnOptions& options = {})
{
    if (h.empty()) throw std::invalid_argument("upfirdn coefficients must be non-empty");
    if (up < 1 || down < 1) throw std::invalid_argument("upfirdn factors must be positive");
    if (x.empty()) throw std::invalid_argument("upfirdn input must be non-empty");
    const detail::Lines l = detail::plan_lines(options.shape, options.axis, x.size(), "upfirdn");
    if (l.shape.size() > 2) throw std::invalid_argument("upfirdn supports rank 1 or 2");
    const detail::FloatSamples hf = detail::as_float(h), xf = detail::as_float(x);
    const long long out_n = zkx_upfirdn_out_len(l.n, static_cast<long long>(h.size()), up, down);
    out = UpfirdnDeviceResult{};
    out.shape = l.shape;
    out.shape[l.axis] = static_cast<int>(out_n);
    DeviceArray<float> y(static_cast<std::size_t>(l.outer) * out_n * l.inner);
    detail::check_abi(zkx_upfirdn_f32(hf.data(), static_cast<int>(h.size()), xf.data(), l.outer, l.n, l.inner, up,
                                      down, y.data(), nullptr));
    if constexpr (std::is_same_v<T, std::int32_t>) {
        out.dtype = UpfirdnOutputDtype::fp64;
        out.fp64 = detail::widen(y);
    } else {
        out.dtype = UpfirdnOutputDtype::fp32;
        out.fp32 = std::move(y);
    }
}

// Complex extension: real float taps on a TypedComplex<float> signal.
template <typename T>
void upfirdn_device(const DeviceArray<T>& h, const DeviceArray<TypedComplex<T>>& x, int up, int down,
                    UpfirdnComplexDeviceResult& out, const UpfirdnOptions& options = {})
{
    static_assert(std::is_same_v<T, float>, "complex upfirdn takes float taps and TypedComplex<float> samples");
    if (h.empty()) throw std::invalid_argument("upfirdn coefficients must be non-empty");
    if (up < 1 || down < 1) throw std::invalid_argument("upfirdn factors must be positive");
    if (x.empty()) throw std::invalid_argument("upfirdn input must be non-empty");
    const detail::Lines l = detail::plan_lines(options.shape, options.axis, x.size(), "upfirdn");
    if (l.shape.size() > 2) throw std::invalid_argument("upfirdn supports rank 1 or 2");
    const long long out_n = zkx_upfirdn_out_len(l.n, static_cast<long long>(h.size()), up, down);
    out = UpfirdnComplexDeviceResult{};
    out.shape = l.shape;
    out.shape[l.axis] = static_cast<int>(out_n);
    out.dtype = UpfirdnOutputDtype::fp32;
    out.fp32 = DeviceArray<ComplexFloat>(static_cast<std::size_t>(l.outer) * out_n * l.inner);
    detail::check_abi(zkx_upfirdn_c64(h.data(), static_cast<int>(h.size()),
                                      reinterpret_cast<const zkx_complex64*>(x.data()), l.outer, l.n, l.inner, up,
                                      down, reinterpret_cast<zkx_complex64*>(out.fp32.data()), nullptr));
}

// filtering_typed.h:765-770 (filtering_typed.cpp:2717-2779): identity copies
// in the input's own type; otherwise FP32 filtering with the dtype of
// resample_poly_output_dtype (default design: FP64 storage).
template <typename T>
void resample_poly_device(const DeviceArray<T>& x, int up, int down, ResamplePolyDeviceResult& out,
                          const ResamplePolyOptions& options = {}, const DeviceArray<T>* coefficients = nullptr)
{
    (void)detail::dtype_code<T>();
    if (up < 1 || down < 1) throw std::invalid_argument("resample_poly factors must be positive");
    const int g = std::gcd(up, down);
    const int u = up / g, d = down / g;
    out = ResamplePolyDeviceResult{};
    out.dtype = detail::resample_poly_dtype<T>(u == 1 && d == 1, coefficients != nullptr, u);
    if (u == 1 && d == 1) {
        out.shape = options.shape.empty() ? std::vector<int>{static_cast<int>(x.size())} : options.shape;
        auto copy_into = [&](auto& member) {
            member.reset(x.size());
            detail::copy_device(member.data(), x.data(), x.size());
        };
        if constexpr (std::is_same_v<T, float>) copy_into(out.fp32);
        else if constexpr (std::is_same_v<T, __half>) copy_into(out.fp16);
        else if constexpr (std::is_same_v<T, std::int32_t>) copy_into(out.int32);
        else if constexpr (std::is_same_v<T, std::int16_t>) copy_into(out.int16);
        else copy_into(out.int8);
        return;
    }
    const detail::Lines l = detail::plan_lines(options.shape, options.axis, x.size(), "resample_poly");
    if (l.shape.size() > 2) throw std::invalid_argument("resample_poly supports rank 1 or 2");
    const DeviceArray<float> taps = detail::resample_poly_taps(u, d, options, coefficients);
    const detail::FloatSamples xf = detail::as_float(x);
    const long long out_n = (static_cast<long long>(l.n) * u + d - 1) / d;
    out.shape = l.shape;
    out.shape[l.axis] = static_cast<int>(out_n);
    DeviceArray<float> y(static_cast<std::size_t>(l.outer) * out_n * l.inner);
    detail::check_abi(zkx_resample_poly_f32(taps.data(), static_cast<int>(taps.size()), xf.data(), l.outer, l.n,
                                            l.inner, up, down, y.data(), nullptr));
    if (out.dtype == ResamplePolyOutputDtype::fp32) out.fp32 = std::move(y);
    else out.fp64 = detail::widen(y);
}

// Complex extension: float taps (custom or designed) on TypedComplex<float>.
template <typename T>
void resample_poly_device(const DeviceArray<TypedComplex<T>>& x, int up, int down,
                          ResamplePolyComplexDeviceResult& out, const ResamplePolyOptions& options = {},
                          const DeviceArray<T>* coefficients = nullptr)
{
    static_assert(std::is_same_v<T, float>, "complex resample_poly takes TypedComplex<float> samples");
    if (up < 1 || down < 1) throw std::invalid_argument("resample_poly factors must be positive");
    const int g = std::gcd(up, down);
    const int u = up / g, d = down / g;
    out = ResamplePolyComplexDeviceResult{};
    out.dtype = detail::resample_poly_dtype<T>(u == 1 && d == 1, coefficients != nullptr, u);
    if (u == 1 && d == 1) {
        out.shape = options.shape.empty() ? std::vector<int>{static_cast<int>(x.size())} : options.shape;
        out.fp32.reset(x.size());
        detail::copy_device(reinterpret_cast<TypedComplex<T>*>(out.fp32.data()), x.data(), x.size());
        return;
    }
    const detail::Lines l = detail::plan_lines(options.shape, options.axis, x.size(), "resample_poly");
    if (l.shape.size() > 2) throw std::invalid_argument("resample_poly supports rank 1 or 2");
    const DeviceArray<float> taps = detail::resample_poly_taps(u, d, options, coefficients);
    const long long out_n = (static_cast<long long>(l.n) * u + d - 1) / d;
    out.shape = l.shape;
    out.shape[l.axis] = static_cast<int>(out_n);
    DeviceArray<ComplexFloat> y(static_cast<std::size_t>(l.outer) * out_n * l.inner);
    detail::check_abi(zkx_resample_poly_c64(taps.data(), static_cast<int>(taps.size()),
                                            reinterpret_cast<const zkx_complex64*>(x.data()), l.outer, l.n, l.inner,
                                            up, down, reinterpret_cast<zkx_complex64*>(y.data()), nullptr));
    if (out.dtype == ResamplePolyOutputDtype::fp32) out.fp32 = std::move(y);
    else out.fp64 = detail::widen(y);
}

// filtering_typed.h:552-556 (filtering_typed.cu:117-164): b, x and zi loaded
// to float; y and zf are float.
template <typename T>
void firfilter_device(const DeviceArray<T>& b, const DeviceArray<T>& x, FirfilterDeviceResult& out,
                      const FirfilterOptions& options = {}, const DeviceArray<T>* zi = nullptr)
{
    if (b.empty()) throw std::invalid_argument("firfilter coefficients must be non-empty");
    if (x.empty()) throw std::invalid_argument("firfilter input must be non-empty");
    const detail::Lines l = detail::plan_lines(options.shape, options.axis, x.size(), "firfilter");
    const int k1 = static_cast<int>(b.size()) - 1;
    out = FirfilterDeviceResult{};
    out.y_shape = l.shape;
    out.y = DeviceArray<float>(x.size());
    out.has_zf = zi != nullptr;
    DeviceArray<T> zi_full;
    if (zi) {
        std::vector<int> state = l.shape;
        state[l.axis] = k1;
        zi_full = detail::broadcast_zi(*zi, options.zi_shape, state, l.axis);
        out.zf_shape = state;
        out.zf = DeviceArray<float>(zi_full.size());
    }
    const detail::FloatSamples bf = detail::as_float(b), xf = detail::as_float(x);
    const detail::FloatSamples zf_in = detail::as_float(zi_full);
    detail::check_abi(zkx_firfilter_f32(bf.data(), static_cast<int>(b.size()), xf.data(), l.outer, l.n, l.inner,
                                        zi ? zf_in.data() : nullptr, out.y.data(), zi ? out.zf.data() : nullptr,
                                        nullptr));
}

// Complex extension: real float taps on TypedComplex<float> samples (and
// complex zi); each component is filtered as a float line.
template <typename T>
void firfilter_device(const DeviceArray<T>& b, const DeviceArray<TypedComplex<T>>& x,
                      FirfilterComplexDeviceResult& out, const FirfilterOptions& options = {},
                      const DeviceArray<TypedComplex<T>>* zi = nullptr)
{
    static_assert(std::is_same_v<T, float>, "complex firfilter takes float taps and TypedComplex<float> samples");
    if (b.empty()) throw std::invalid_argument("firfilter coefficients must be non-empty");
    if (x.empty()) throw std::invalid_argument("firfilter input must be non-empty");
    const detail::Lines l = detail::plan_lines(options.shape, options.axis, x.size(), "firfilter");
    const int k1 = static_cast<int>(b.size()) - 1;
    out = FirfilterComplexDeviceResult{};
    out.y_shape = l.shape;
    out.y = DeviceArray<ComplexFloat>(x.size());
    out.has_zf = zi != nullptr;
    DeviceArray<TypedComplex<T>> zi_full;
    if (zi) {
        std::vector<int> state = l.shape;
        state[l.axis] = k1;
        zi_full = detail::broadcast_zi(*zi, options.zi_shape, state, l.axis);
        out.zf_shape = state;
        out.zf = DeviceArray<ComplexFloat>(zi_full.size());
    }
    // the interleaved (re, im) pairs are float lines with twice the inner count
    detail::check_abi(zkx_firfilter_f32(b.data(), static_cast<int>(b.size()), reinterpret_cast<const float*>(x.data()),
                                        l.outer, l.n, 2 * l.inner,
                                        zi ? reinterpret_cast<const float*>(zi_full.data()) : nullptr,
                                        reinterpret_cast<float*>(out.y.data()),
                                        zi ? reinterpret_cast<float*>(out.zf.data()) : nullptr, nullptr));
}

// filtering_typed.h:702-706 (filtering_typed.cpp:2319-2356): any rank, any
// axis; the default filter is designed on the host in FP64 (firwin, Hamming,
// order 20*q unless given, cutoff 1/q) and rounded to float taps; custom
// coefficients are loaded to float on the device.  FP32 arithmetic; the
// designed filter and int32 custom coefficients give FP64 storage.
template <typename T>
void decimate_device(const DeviceArray<T>& x, int q, DecimateDeviceResult& out, const DecimateOptions& options = {},
                     const DeviceArray<T>* coefficients = nullptr)
{
    (void)detail::dtype_code<T>();
    if (q < 1) throw std::invalid_argument("decimate q");
    if (x.size() > static_cast<std::size_t>(std::numeric_limits<int>::max()))
        throw std::invalid_argument("decimate input exceeds int indexing");
    // prepare_decimate_layout (filtering_typed.cpp:788-853)
    std::vector<int> shape = options.shape;
    if (shape.empty()) shape.push_back(static_cast<int>(x.size()));
    long long product = 1;
    for (const int extent : shape) {
        if (extent < 0) throw std::invalid_argument("decimate input shape");
        if (product != 0 && extent != 0 && product > std::numeric_limits<int>::max() / extent)
            throw std::invalid_argument("decimate input shape overflow");
        product *= extent;
    }
    if (product != static_cast<long long>(x.size())) throw std::invalid_argument("decimate input shape product");
    int axis = options.axis;
    if (axis < 0) axis += static_cast<int>(shape.size());
    if (axis < 0 || axis >= static_cast<int>(shape.size())) throw std::invalid_argument("decimate axis");
    long long outer = 1, inner = 1;
    for (int d = 0; d < axis; ++d) outer *= shape[d];
    for (std::size_t d = axis + 1; d < shape.size(); ++d) inner *= shape[d];
    const int n = shape[axis];
    const int out_n = n == 0 ? 0 : 1 + (n - 1) / q;
    std::vector<int> out_shape = shape;
    out_shape[axis] = out_n;
    const long long out_size = outer * inner * out_n;
    if (outer > std::numeric_limits<int>::max() || inner > std::numeric_limits<int>::max() ||
        out_size > std::numeric_limits<int>::max())
        throw std::invalid_argument("decimate output shape overflow");
    const bool custom = coefficients != nullptr;
    if (custom && options.filter_order.has_value())
        throw std::invalid_argument("decimate filter_order and coefficients are mutually exclusive");
    if (custom && coefficients->empty()) throw std::invalid_argument("decimate coefficients");
    DeviceArray<float> designed;
    int ntaps = 0;
    if (!custom) {
        // design_decimate_filter (filtering_typed.cpp:318-336)
        int order = 0;
        if (options.filter_order.has_value()) {
            order = *options.filter_order;
        } else {
            if (q > std::numeric_limits<int>::max() / 20) throw std::invalid_argument("decimate default filter order overflow");
            order = 20 * q;
        }
        if (order < 0 || order == std::numeric_limits<int>::max()) throw std::invalid_argument("decimate filter order");
        std::vector<double> taps64(static_cast<std::size_t>(order) + 1);
        detail::check_abi(zkx_firwin_lowpass_hamming(order + 1, 1.0f / static_cast<float>(q), taps64.data()));
        std::vector<float> taps32(taps64.begin(), taps64.end());
        designed = DeviceArray<float>::from_host(taps32);
        ntaps = order + 1;
    } else {
        ntaps = static_cast<int>(coefficients->size());
    }
    const detail::FloatSamples hf = custom ? detail::as_float(*coefficients) : detail::as_float(designed);
    const detail::FloatSamples xf = detail::as_float(x);
    DeviceArray<float> y(static_cast<std::size_t>(out_size));
    detail::check_abi(zkx_decimate_f32(hf.data(), ntaps, xf.data(), static_cast<int>(outer), n,
                                       static_cast<int>(inner), q, options.zero_phase ? 1 : 0, y.data(), nullptr));
    out = DecimateDeviceResult{};
    out.shape = out_shape;
    out.dtype = (!custom || std::is_same_v<T, std::int32_t>) ? DecimateOutputDtype::fp64 : DecimateOutputDtype::fp32;
    if (out.dtype == DecimateOutputDtype::fp32) {
        cuda_utils::synchronize_stream();
        out.fp32 = std::move(y);
    } else {
        out.fp64 = detail::widen(y);
    }
}

// filtering_typed.h:513-517 (plan filtering_typed.cpp:443-466): x and h
// loaded to float on the device; complex64 [n_chans, floor(N/n_chans)].
template <typename T>
void channelize_poly_device(const DeviceArray<T>& x, const DeviceArray<T>& h, int n_chans,
                            ChannelizePolyDeviceResult& out)
{
    (void)detail::dtype_code<T>();
    if (n_chans <= 0) throw std::invalid_argument("channelize_poly n_chans");
    if (x.size() > static_cast<std::size_t>(std::numeric_limits<int>::max()) ||
        h.size() > static_cast<std::size_t>(std::numeric_limits<int>::max()))
        throw std::invalid_argument("channelize_poly input exceeds int indexing");
    if (h.size() / static_cast<std::size_t>(n_chans) > 32)
        throw std::invalid_argument("channelize_poly taps per phase exceed 32");
    const int points = static_cast<int>(x.size() / static_cast<std::size_t>(n_chans));
    out = ChannelizePolyDeviceResult{};
    out.shape = {n_chans, points};
    out.values = DeviceArray<ComplexFloat>(static_cast<std::size_t>(n_chans) * points);
    const detail::FloatSamples xf = detail::as_float(x), hf = detail::as_float(h);
    detail::check_abi(zkx_channelize_poly_f32(xf.data(), static_cast<long long>(x.size()), hf.data(),
                                              static_cast<long long>(h.size()), n_chans,
                                              reinterpret_cast<zkx_complex64*>(out.values.data()), nullptr));
    cuda_utils::synchronize_stream();
}

// filtering_typed.h:519-523: higher-rank x/h, of which the first extents
// select the samples (prepare_channelize_poly, filtering_typed.cpp:469-498).
template <typename T>
void channelize_poly_device(const DeviceArray<T>& x, const DeviceArray<T>& h, int n_chans,
                            const std::vector<int>& input_shape, const std::vector<int>& filter_shape,
                            ChannelizePolyDeviceResult& out)
{
    auto first_extent = [](const std::vector<int>& shape, std::size_t size) -> std::size_t {
        if (shape.empty()) throw std::invalid_argument("channelize_poly shape");
        std::size_t product = 1;
        bool zero = false;
        for (const int extent : shape) {
            if (extent < 0) throw std::invalid_argument("channelize_poly shape");
            if (extent == 0) zero = true;
            else if (!zero) {
                if (product > size / static_cast<std::size_t>(extent))
                    throw std::invalid_argument("channelize_poly data/shape mismatch");
                product *= static_cast<std::size_t>(extent);
            }
        }
        if ((zero ? 0 : product) != size) throw std::invalid_argument("channelize_poly data/shape mismatch");
        const std::size_t first = static_cast<std::size_t>(shape.front());
        if (first > size) throw std::invalid_argument("channelize_poly first dimension exceeds storage");
        return first;
    };
    const std::size_t nx = first_extent(input_shape, x.size());
    const std::size_t nh = first_extent(filter_shape, h.size());
    (void)detail::dtype_code<T>();
    if (n_chans <= 0) throw std::invalid_argument("channelize_poly n_chans");
    if (nh / static_cast<std::size_t>(n_chans) > 32)
        throw std::invalid_argument("channelize_poly taps per phase exceed 32");
    const int points = static_cast<int>(nx / static_cast<std::size_t>(n_chans));
    out = ChannelizePolyDeviceResult{};
    out.shape = {n_chans, points};
    out.values = DeviceArray<ComplexFloat>(static_cast<std::size_t>(n_chans) * points);
    const detail::FloatSamples xf = detail::as_float(x), hf = detail::as_float(h);
    detail::check_abi(zkx_channelize_poly_f32(xf.data(), static_cast<long long>(nx), hf.data(),
                                              static_cast<long long>(nh), n_chans,
                                              reinterpret_cast<zkx_complex64*>(out.values.data()), nullptr));
    cuda_utils::synchronize_stream();
}

// filtering_typed.h:563-565: 1-D, no state, caller-owned output, no allocation.
inline void firfilter_fp32_resident_device(const DeviceArray<float>& b, const DeviceArray<float>& x,
                                           DeviceArray<float>& y)
{
    if (b.empty() || x.empty()) throw std::invalid_argument("firfilter resident inputs must be non-empty");
    if (y.size() != x.size()) throw std::invalid_argument("firfilter resident output must match input size");
    detail::check_abi(zkx_firfilter_f32(b.data(), static_cast<int>(b.size()), x.data(), 1, static_cast<int>(x.size()),
                                        1, nullptr, y.data(), nullptr, nullptr));
}

// fftconvolve extension on complex signals of any instantiated component
// type (loaded to float): out has the row count of in1 and the mode's length.
template <typename T>
void fftconvolve_device(const DeviceArray<TypedComplex<T>>& in1, const DeviceArray<TypedComplex<T>>& in2,
                        FftconvolveDeviceResult& out, const FftconvolveOptions& options = {})
{
    if (in1.empty() || in2.empty()) throw std::invalid_argument("fftconvolve inputs must be non-empty");
    const detail::Lines l = detail::plan_lines(options.shape, -1, in1.size(), "fftconvolve");
    const int mode = static_cast<int>(options.mode);
    const long long out_n = zkx_fftconvolve_out_len(l.n, static_cast<long long>(in2.size()), mode);
    if (out_n < 0) throw std::invalid_argument("fftconvolve mode is invalid for these lengths");
    const detail::FloatSamples a = detail::as_float(in1), b = detail::as_float(in2);
    out = FftconvolveDeviceResult{};
    out.shape = l.shape;
    out.shape.back() = static_cast<int>(out_n);
    out.out = DeviceArray<ComplexFloat>(static_cast<std::size_t>(l.outer) * out_n);
    if (out_n == 0) return;
    detail::check_abi(zkx_fftconvolve_c64(reinterpret_cast<const zkx_complex64*>(a.data()), l.outer, l.n,
                                          reinterpret_cast<const zkx_complex64*>(b.data()),
                                          static_cast<int>(in2.size()), mode,
                                          reinterpret_cast<zkx_complex64*>(out.out.data()), nullptr));
}

// fftconvolve over N-D axes (extension; cuSignal convolve.py:159-320): both
// inputs of one rank (row-major shapes), convolved over `axes` (empty: all;
// negative axes count from the end), the other dimensions broadcast (equal
// extents or 1).  Real inputs give a real result, complex inputs a complex
// one.
struct FftconvolveNDOptions {
    std::vector<int> shape1;
    std::vector<int> shape2;
    FftconvolveMode mode = FftconvolveMode::full;
    std::vector<int> axes;
};

namespace detail {
inline unsigned fftconvolve_axes_mask(const FftconvolveNDOptions& o)
{
    const int rank = static_cast<int>(o.shape1.size());
    if (o.axes.empty()) return rank >= 32 ? ~0u : (1u << rank) - 1u;
    unsigned m = 0;
    for (int a : o.axes) {
        const int d = a < 0 ? a + rank : a;
        if (d < 0 || d >= rank) throw OperatorArgumentError(ErrorCode::INVALID_ARGUMENT, "axes exceeds dimensionality of input");
        if (m >> d & 1u) throw OperatorArgumentError(ErrorCode::INVALID_ARGUMENT, "all axes must be unique");
        m |= 1u << d;
    }
    return m;
}
}  // namespace detail

// Extents of the result (zkx_fftconvolve_nd_shape).
inline std::vector<int> fftconvolve_nd_output_shape(const FftconvolveNDOptions& o)
{
    if (o.shape1.size() != o.shape2.size())
        throw OperatorArgumentError(ErrorCode::INVALID_SHAPE, "in1 and in2 should have the same dimensionality");
    const int rank = static_cast<int>(o.shape1.size());
    std::vector<long long> s1(o.shape1.begin(), o.shape1.end()), s2(o.shape2.begin(), o.shape2.end()), out(rank);
    detail::check_abi(zkx_fftconvolve_nd_shape(s1.data(), s2.data(), rank, detail::fftconvolve_axes_mask(o),
                                               static_cast<int>(o.mode), out.data()));
    return std::vector<int>(out.begin(), out.end());
}

// A and B are float or TypedComplex<float>; O is float when both are real,
// TypedComplex<float> otherwise; out holds fftconvolve_nd_output_shape(o)
// elements.
template <typename A, typename B, typename O>
void fftconvolve_device(const DeviceArray<A>& in1, const DeviceArray<B>& in2, DeviceArray<O>& out,
                        const FftconvolveNDOptions& o)
{
    constexpr bool ca = std::is_same_v<A, TypedComplex<float>>, cb = std::is_same_v<B, TypedComplex<float>>;
    static_assert((ca || std::is_same_v<A, float>) && (cb || std::is_same_v<B, float>),
                  "fftconvolve N-D inputs are float or TypedComplex<float>");
    static_assert(std::is_same_v<O, std::conditional_t<ca || cb, TypedComplex<float>, float>>,
                  "fftconvolve N-D output is float for two real inputs, TypedComplex<float> otherwise");
    const std::vector<int> shape = fftconvolve_nd_output_shape(o);
    auto count = [](const std::vector<int>& s) {
        std::size_t n = 1;
        for (int v : s) n *= static_cast<std::size_t>(v);
        return n;
    };
    if (in1.size() != count(o.shape1) || in2.size() != count(o.shape2))
        throw OperatorArgumentError(ErrorCode::INVALID_SHAPE, "fftconvolve input sizes must match their shapes");
    if (out.size() != count(shape))
        throw OperatorArgumentError(ErrorCode::INVALID_SHAPE, "fftconvolve output size must match the result shape");
    const int rank = static_cast<int>(o.shape1.size());
    std::vector<long long> s1(o.shape1.begin(), o.shape1.end()), s2(o.shape2.begin(), o.shape2.end());
    detail::check_abi(zkx_fftconvolve_nd(in1.data(), s1.data(), ca ? 1 : 0, in2.data(), s2.data(), cb ? 1 : 0, rank,
                                         detail::fftconvolve_axes_mask(o), static_cast<int>(o.mode), out.data(),
                                         nullptr));
}

}  // namespace cusignal
