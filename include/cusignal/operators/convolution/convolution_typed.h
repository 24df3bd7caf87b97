// The reference's correlate API (ZKX/cusignal_cpp/include/cusignal/operators/
// convolution/convolution_typed.h) on libzkx_radar.so, 1-D direct method:
// correlate_output_size and correlate_device(in1, in2, out, mode) for float,
// __half, int32, int16 and int8, with the reference's arithmetic (an FP32 fma
// chain for floating types, modular integer sums) and exceptions; and the 2-D
// direct form correlate2d_output_shape / Correlate2DWorkspace /
// correlate2d_device with fill, wrap and symm boundaries; the FFT method
// through the 1-D workspace overload; the N-D form (CorrelateNDOptions,
// correlate_nd_output_shape, CorrelateNDWorkspace and its correlate_device
// overload) by FFT over every dimension.
#pragma once

#include <cstddef>
#include <cstdint>
#include <limits>
#include <stdexcept>
#include <string>
#include <type_traits>
#include <vector>

#include "../filtering/filtering_typed.h"

namespace cusignal {

enum class CorrelateMethod { direct, fft, auto_select };

namespace detail {
inline int correlate_mode_code(const std::string& mode)
{
    if (mode == "full") return 0;
    if (mode == "same") return 1;
    if (mode == "valid") return 2;
    throw std::invalid_argument("typed correlate mode must be full, same, or valid");
}
}  // namespace detail

// convolution_typed.cpp:134-152
inline std::size_t correlate_output_size(std::size_t len1, std::size_t len2, const std::string& mode)
{
    const int code = detail::correlate_mode_code(mode);
    if (len1 == 0 || len2 == 0) return 0;
    if (code == 0) {
        if (len1 > std::numeric_limits<std::size_t>::max() - len2 + 1)
            throw std::invalid_argument("typed correlate output size overflow");
        return len1 + len2 - 1;
    }
    if (code == 1) return len1;
    return len1 >= len2 ? len1 - len2 + 1 : len2 - len1 + 1;
}

// convolution_typed.h:298-303: direct method, caller-allocated output of
// correlate_output_size(L1, L2, mode) elements.
template <typename T>
void correlate_device(const DeviceArray<T>& in1, const DeviceArray<T>& in2, DeviceArray<T>& out,
                      const std::string& mode = "full")
{
    const int code = detail::correlate_mode_code(mode);
    const std::size_t n = correlate_output_size(in1.size(), in2.size(), mode);
    if (out.size() != n) throw std::invalid_argument("typed correlate output shape mismatch");
    if (n == 0) return;
    detail::check_abi(zkx_correlate_direct(in1.data(), static_cast<long long>(in1.size()), in2.data(),
                                           static_cast<long long>(in2.size()), detail::dtype_code<T>(), code,
                                           out.data(), nullptr));
}

// convolution_typed.h:120-158 (CorrelateDeviceParams / CorrelateDeviceWorkspace):
// the FFT method's parameters and output length; the library holds the
// complex FP32 scratch, so the workspace carries no device arrays.
struct CorrelateDeviceParams {
    int len1 = 0;
    int len2 = 0;
    std::string mode = "same";
};

struct CorrelateDeviceWorkspace {
    int len1 = 0;
    int len2 = 0;
    std::string mode = "same";
    std::size_t out_len = 0;

    CorrelateDeviceWorkspace() = default;
    explicit CorrelateDeviceWorkspace(const CorrelateDeviceParams& params) { reset(params); }
    void reset(const CorrelateDeviceParams& params)
    {
        if (params.len1 < 0 || params.len2 < 0) throw std::invalid_argument("typed correlate negative length");
        out_len = correlate_output_size(static_cast<std::size_t>(params.len1), static_cast<std::size_t>(params.len2),
                                        params.mode);
        len1 = params.len1;
        len2 = params.len2;
        mode = params.mode;
    }
};

// convolution_typed.h:307-327: FFT method with the caller's workspace.
template <typename T>
void correlate_device(const DeviceArray<T>& in1, const DeviceArray<T>& in2, DeviceArray<T>& out,
                      CorrelateDeviceWorkspace& workspace)
{
    if (in1.size() != static_cast<std::size_t>(workspace.len1) ||
        in2.size() != static_cast<std::size_t>(workspace.len2) || out.size() != workspace.out_len)
        throw std::invalid_argument("typed correlate workspace shape mismatch");
    if (workspace.out_len == 0) return;
    const int rc = zkx_correlate_fft(in1.data(), workspace.len1, in2.data(), workspace.len2, detail::dtype_code<T>(),
                                     detail::correlate_mode_code(workspace.mode), out.data(), nullptr);
    if (rc == ZKX_INVALID_ARGUMENT || rc == ZKX_INVALID_SHAPE) throw std::invalid_argument(zkx_last_error_message());
    detail::check_abi(rc);
}

// convolution_typed.h:70-93 (CorrelateNDOptions / CorrelateNDOutputShape /
// correlate_nd_output_shape): any rank, shapes row-major, an empty shape is a
// scalar.  direct is only the 1-D method (rank > 1 refused), auto picks the
// direct method for rank <= 1 and FFT above.
struct CorrelateNDOptions {
    std::vector<int> shape1;
    std::vector<int> shape2;
    std::string mode = "full";
    CorrelateMethod method = CorrelateMethod::auto_select;
};

struct CorrelateNDOutputShape {
    std::vector<int> dimensions;
    std::size_t elements = 0;

    [[nodiscard]] std::size_t size() const noexcept { return elements; }
};

// convolution_typed.cpp:153-199: scalars give {} with 1 element, an empty
// input {0} with 0, otherwise full / same / valid extents per dimension.
inline CorrelateNDOutputShape correlate_nd_output_shape(const CorrelateNDOptions& options)
{
    const int code = detail::correlate_mode_code(options.mode);
    if (options.method != CorrelateMethod::direct && options.method != CorrelateMethod::fft &&
        options.method != CorrelateMethod::auto_select)
        throw std::invalid_argument("typed correlate method must be direct, fft, or auto");
    if (options.shape1.size() != options.shape2.size())
        throw std::invalid_argument("typed correlate inputs must have the same rank");
    if (options.method == CorrelateMethod::direct && options.shape1.size() > 1)
        throw std::invalid_argument("typed correlate direct method is only implemented for 1D");
    const int rank = static_cast<int>(options.shape1.size());
    CorrelateNDOutputShape out;
    std::vector<int> dims(rank > 0 ? rank : 1, 0);
    long long elements = 0;
    detail::check_abi(zkx_correlate_nd_out_shape(options.shape1.data(), options.shape2.data(), rank, code,
                                                 dims.data(), &elements));
    out.elements = static_cast<std::size_t>(elements);
    if (rank == 0) return out;
    if (elements == 0) {
        out.dimensions = {0};
        return out;
    }
    out.dimensions.assign(dims.begin(), dims.begin() + rank);
    return out;
}

// convolution_typed.h:149-190 (CorrelateNDWorkspace): the validated shapes,
// mode and method of one N-D correlation.  The library holds the padded
// complex FP32 operands and plans (per host thread, ordered on the default
// stream), so the workspace carries no device arrays.
struct CorrelateNDWorkspace {
    std::vector<int> shape1;
    std::vector<int> shape2;
    std::vector<int> output_shape;
    CorrelateMethod requested_method = CorrelateMethod::auto_select;
    CorrelateMethod effective_method = CorrelateMethod::fft;
    std::string mode = "full";
    int rank = 0;
    std::size_t input1_elements = 0;
    std::size_t input2_elements = 0;
    std::size_t output_elements = 0;

    CorrelateNDWorkspace() = default;
    explicit CorrelateNDWorkspace(const CorrelateNDOptions& options) { reset(options); }
    void reset(const CorrelateNDOptions& options)
    {
        const CorrelateNDOutputShape out = correlate_nd_output_shape(options);
        auto count = [](const std::vector<int>& s) {
            std::size_t n = 1;
            for (int v : s) n *= static_cast<std::size_t>(v);
            return n;
        };
        shape1 = options.shape1;
        shape2 = options.shape2;
        output_shape = out.dimensions;
        requested_method = options.method;
        mode = options.mode;
        rank = static_cast<int>(shape1.size());
        input1_elements = count(shape1);
        input2_elements = count(shape2);
        output_elements = out.size();
        effective_method = requested_method == CorrelateMethod::auto_select
                               ? (rank <= 1 ? CorrelateMethod::direct : CorrelateMethod::fft)
                               : requested_method;
    }
    [[nodiscard]] std::size_t output_size() const noexcept { return output_elements; }
};

// convolution_typed.h:331-337: N-D correlate with the caller's workspace on
// the default stream (rank 1 with the direct method runs the direct sums).
template <typename T>
void correlate_device(const DeviceArray<T>& in1, const DeviceArray<T>& in2, DeviceArray<T>& out,
                      CorrelateNDWorkspace& workspace)
{
    if (in1.size() != workspace.input1_elements || in2.size() != workspace.input2_elements ||
        out.size() != workspace.output_elements)
        throw std::invalid_argument("typed correlate workspace shape mismatch");
    if (workspace.output_elements == 0) return;
    const int code = detail::correlate_mode_code(workspace.mode);
    if (workspace.rank == 1 && workspace.effective_method == CorrelateMethod::direct) {
        detail::check_abi(zkx_correlate_direct(in1.data(), static_cast<long long>(in1.size()), in2.data(),
                                               static_cast<long long>(in2.size()), detail::dtype_code<T>(), code,
                                               out.data(), nullptr));
        return;
    }
    detail::check_abi(zkx_correlate_nd(in1.data(), workspace.shape1.data(), in2.data(), workspace.shape2.data(),
                                       workspace.rank, detail::dtype_code<T>(), code, out.data(), nullptr));
}

// convolution_typed.h:204-218 (Correlate2DOptions / Correlate2DOutputShape)
struct Correlate2DOptions {
    int rows1 = 0;
    int cols1 = 0;
    int rows2 = 0;
    int cols2 = 0;
    std::string mode = "same";
    std::string boundary = "fill";
    double fillvalue = 0.0;
};

struct Correlate2DOutputShape {
    int rows = 0;
    int cols = 0;
    [[nodiscard]] std::size_t size() const noexcept
    {
        return static_cast<std::size_t>(rows) * static_cast<std::size_t>(cols);
    }
};

namespace detail {
inline int correlate2d_boundary_code(const std::string& boundary)
{
    if (boundary == "fill") return 0;
    if (boundary == "wrap") return 1;
    if (boundary == "symm") return 2;
    throw std::invalid_argument("typed correlate2d boundary must be fill, wrap, or symm");
}
}  // namespace detail

// convolution_typed.cpp:203-236: invalid mode, boundary or shape is
// std::invalid_argument; an empty dimension gives 0 x 0.
inline Correlate2DOutputShape correlate2d_output_shape(const Correlate2DOptions& options)
{
    if (options.mode != "full" && options.mode != "same" && options.mode != "valid")
        throw std::invalid_argument("typed correlate2d mode must be full, same, or valid");
    (void)detail::correlate2d_boundary_code(options.boundary);
    long long rows = 0, cols = 0;
    const int rc = zkx_correlate2d_out_shape(options.rows1, options.cols1, options.rows2, options.cols2,
                                             detail::correlate_mode_code(options.mode), &rows, &cols);
    if (rc) throw std::invalid_argument(zkx_last_error_message());
    return {static_cast<int>(rows), static_cast<int>(cols)};
}

// convolution_typed.h:220-250 (Correlate2DWorkspace): host parameters and the
// output shape; no device scratch.
struct Correlate2DWorkspace {
    int rows1 = 0, cols1 = 0, rows2 = 0, cols2 = 0;
    std::string mode = "same";
    std::string boundary = "fill";
    double fillvalue = 0.0;
    int out_rows = 0, out_cols = 0;

    Correlate2DWorkspace() = default;
    explicit Correlate2DWorkspace(const Correlate2DOptions& options) { reset(options); }
    void reset(const Correlate2DOptions& options)
    {
        const Correlate2DOutputShape shape = correlate2d_output_shape(options);
        rows1 = options.rows1;
        cols1 = options.cols1;
        rows2 = options.rows2;
        cols2 = options.cols2;
        mode = options.mode;
        boundary = options.boundary;
        fillvalue = options.fillvalue;
        out_rows = shape.rows;
        out_cols = shape.cols;
    }
    [[nodiscard]] std::size_t output_size() const noexcept
    {
        return static_cast<std::size_t>(out_rows) * static_cast<std::size_t>(out_cols);
    }
};

// convolution_typed.h:374-379: row-major in1 [rows1 x cols1], in2
// [rows2 x cols2], out [out_rows x out_cols] preallocated by the caller.
template <typename T>
void correlate2d_device(const DeviceArray<T>& in1, const DeviceArray<T>& in2, DeviceArray<T>& out,
                        Correlate2DWorkspace& workspace)
{
    if (in1.size() != static_cast<std::size_t>(workspace.rows1) * static_cast<std::size_t>(workspace.cols1) ||
        in2.size() != static_cast<std::size_t>(workspace.rows2) * static_cast<std::size_t>(workspace.cols2) ||
        out.size() != workspace.output_size())
        throw std::invalid_argument("typed correlate2d shape mismatch");
    if (workspace.output_size() == 0) return;
    const int rc = zkx_correlate2d(in1.data(), workspace.rows1, workspace.cols1, in2.data(), workspace.rows2,
                                   workspace.cols2, detail::dtype_code<T>(),
                                   detail::correlate_mode_code(workspace.mode),
                                   detail::correlate2d_boundary_code(workspace.boundary), workspace.fillvalue,
                                   out.data(), nullptr);
    if (rc == ZKX_INVALID_ARGUMENT || rc == ZKX_INVALID_SHAPE) throw std::invalid_argument(zkx_last_error_message());
    detail::check_abi(rc);
}

}  // namespace cusignal
