// The reference's polyphase filtering operator API
// (ZKX/cusignal_cpp/include/cusignal/operators/filtering/filtering_typed.h)
// on libzkx_radar.so for float inputs: upfirdn_device, resample_poly_device,
// firfilter_device and firfilter_fp32_resident_device with the reference's
// option/result structs, shape/axis rules, output dtypes and exceptions.
// Filtering runs in the gfx950 kernels behind the C ABI; the host only plans
// lines, designs the default resample filter (FP64, as the reference does on
// its host) and broadcasts zi.
#pragma once

#include <cmath>
#include <cstdint>
#include <numeric>
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

// The FP16 / integer members of the reference's result exist for its other
// input dtypes; the float path fills fp32 or fp64 only.
struct ResamplePolyDeviceResult {
    ResamplePolyOutputDtype dtype{ResamplePolyOutputDtype::fp64};
    DeviceArray<float> fp32;
    DeviceArray<double> fp64;
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

namespace detail {

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

}  // namespace detail

// filtering_typed.h:744-748 (float input: FP32 result)
template <typename T>
void upfirdn_device(const DeviceArray<T>& h, const DeviceArray<T>& x, int up, int down, UpfirdnDeviceResult& out,
                    const UpfirdnOptions& options = {})
{
    static_assert(std::is_same_v<T, float>, "libzkx_radar implements the float upfirdn path");
    if (h.empty()) throw std::invalid_argument("upfirdn coefficients must be non-empty");
    if (up < 1 || down < 1) throw std::invalid_argument("upfirdn factors must be positive");
    if (x.empty()) throw std::invalid_argument("upfirdn input must be non-empty");
    const detail::Lines l = detail::plan_lines(options.shape, options.axis, x.size(), "upfirdn");
    if (l.shape.size() > 2) throw std::invalid_argument("upfirdn supports rank 1 or 2");
    const long long out_n = zkx_upfirdn_out_len(l.n, static_cast<long long>(h.size()), up, down);
    out.shape = l.shape;
    out.shape[l.axis] = static_cast<int>(out_n);
    out.dtype = UpfirdnOutputDtype::fp32;
    out.fp64.reset();
    out.fp32 = DeviceArray<float>(static_cast<std::size_t>(l.outer) * out_n * l.inner);
    detail::check_abi(zkx_upfirdn_f32(h.data(), static_cast<int>(h.size()), x.data(), l.outer, l.n, l.inner, up, down,
                                      out.fp32.data(), nullptr));
}

// filtering_typed.h:765-770.  Default design: FP64 result.  Float custom
// coefficients: scaled by the reduced up in float, FP32 result.
template <typename T>
void resample_poly_device(const DeviceArray<T>& x, int up, int down, ResamplePolyDeviceResult& out,
                          const ResamplePolyOptions& options = {}, const DeviceArray<T>* coefficients = nullptr)
{
    static_assert(std::is_same_v<T, float>, "libzkx_radar implements the float resample_poly path");
    if (up < 1 || down < 1) throw std::invalid_argument("resample_poly factors must be positive");
    const int g = std::gcd(up, down);
    const int u = up / g, d = down / g;
    out.fp32.reset();
    out.fp64.reset();
    if (u == 1 && d == 1) {
        out.shape = options.shape.empty() ? std::vector<int>{static_cast<int>(x.size())} : options.shape;
        out.dtype = ResamplePolyOutputDtype::fp32;
        out.fp32 = DeviceArray<float>(x.size());
        if (!x.empty())
            cuda_utils::check_hip(hipMemcpy(out.fp32.data(), x.data(), sizeof(float) * x.size(), hipMemcpyDeviceToDevice),
                                  "hipMemcpy D2D");
        return;
    }
    const detail::Lines l = detail::plan_lines(options.shape, options.axis, x.size(), "resample_poly");
    if (l.shape.size() > 2) throw std::invalid_argument("resample_poly supports rank 1 or 2");
    std::vector<float> taps;
    bool fp64_result = true;
    if (coefficients) {
        if (coefficients->empty()) throw std::invalid_argument("resample_poly coefficients must be non-empty");
        const std::vector<float> c = coefficients->to_host();
        taps.resize(c.size());
        for (std::size_t i = 0; i < c.size(); ++i) taps[i] = c[i] * static_cast<float>(u);
        fp64_result = false;
    } else {
        const std::vector<double> h = detail::design_resample_poly(u, d, options);
        taps.assign(h.begin(), h.end());
    }
    const DeviceArray<float> taps_d = DeviceArray<float>::from_host(taps);
    const long long out_n = (static_cast<long long>(l.n) * u + d - 1) / d;
    out.shape = l.shape;
    out.shape[l.axis] = static_cast<int>(out_n);
    DeviceArray<float> y(static_cast<std::size_t>(l.outer) * out_n * l.inner);
    detail::check_abi(zkx_resample_poly_f32(taps_d.data(), static_cast<int>(taps.size()), x.data(), l.outer, l.n,
                                            l.inner, up, down, y.data(), nullptr));
    if (fp64_result) {
        out.dtype = ResamplePolyOutputDtype::fp64;
        out.fp64 = detail::widen(y);
    } else {
        out.dtype = ResamplePolyOutputDtype::fp32;
        out.fp32 = std::move(y);
    }
}

// filtering_typed.h:552-556
template <typename T>
void firfilter_device(const DeviceArray<T>& b, const DeviceArray<T>& x, FirfilterDeviceResult& out,
                      const FirfilterOptions& options = {}, const DeviceArray<T>* zi = nullptr)
{
    static_assert(std::is_same_v<T, float>, "libzkx_radar implements the float firfilter path");
    if (b.empty()) throw std::invalid_argument("firfilter coefficients must be non-empty");
    if (x.empty()) throw std::invalid_argument("firfilter input must be non-empty");
    const detail::Lines l = detail::plan_lines(options.shape, options.axis, x.size(), "firfilter");
    const int k1 = static_cast<int>(b.size()) - 1;
    out.y_shape = l.shape;
    out.y = DeviceArray<float>(x.size());
    out.has_zf = zi != nullptr;
    out.zf.reset();
    out.zf_shape.clear();
    DeviceArray<float> zi_full;
    if (zi) {
        // zi broadcasts to the state shape (shape with axis length K-1); the
        // given zi_shape may hold 1 on any non-axis dimension.
        std::vector<int> state = l.shape;
        state[l.axis] = k1;
        const std::vector<int> zs = options.zi_shape.empty() ? std::vector<int>{static_cast<int>(zi->size())}
                                                             : options.zi_shape;
        if (zs.size() != state.size() && !(zs.size() == 1 && state.size() == 1))
            throw std::invalid_argument("firfilter zi rank must match x");
        long long zcount = 1;
        for (std::size_t i = 0; i < zs.size(); ++i) {
            if (zs[i] != state[i] && !(zs[i] == 1 && static_cast<int>(i) != l.axis))
                throw std::invalid_argument("firfilter zi shape cannot broadcast to the state shape");
            zcount *= zs[i];
        }
        if (static_cast<std::size_t>(zcount) != zi->size()) throw std::invalid_argument("firfilter zi size mismatch");
        const std::size_t state_count = static_cast<std::size_t>(l.outer) * k1 * l.inner;
        if (zi->size() == state_count) {
            zi_full = DeviceArray<float>(state_count);
            if (state_count)
                cuda_utils::check_hip(hipMemcpy(zi_full.data(), zi->data(), sizeof(float) * state_count,
                                                hipMemcpyDeviceToDevice), "hipMemcpy D2D");
        } else {
            const std::vector<float> zh = zi->to_host();
            std::vector<float> full(state_count);
            std::vector<int> idx(state.size(), 0);
            for (std::size_t e = 0; e < state_count; ++e) {
                std::size_t src = 0;
                for (std::size_t dd = 0; dd < state.size(); ++dd) src = src * zs[dd] + (zs[dd] == 1 ? 0 : idx[dd]);
                full[e] = zh[src];
                for (int dd = static_cast<int>(state.size()) - 1; dd >= 0; --dd) {
                    if (++idx[dd] < state[dd]) break;
                    idx[dd] = 0;
                }
            }
            zi_full = DeviceArray<float>::from_host(full);
        }
        out.zf_shape = state;
        out.zf = DeviceArray<float>(state_count);
    }
    detail::check_abi(zkx_firfilter_f32(b.data(), static_cast<int>(b.size()), x.data(), l.outer, l.n, l.inner,
                                        zi ? zi_full.data() : nullptr, out.y.data(), zi ? out.zf.data() : nullptr,
                                        nullptr));
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

}  // namespace cusignal
