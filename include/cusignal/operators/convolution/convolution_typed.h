// The reference's correlate API (ZKX/cusignal_cpp/include/cusignal/operators/
// convolution/convolution_typed.h) on libzkx_radar.so, 1-D direct method:
// correlate_output_size and correlate_device(in1, in2, out, mode) for float,
// __half, int32, int16 and int8, with the reference's arithmetic (an FP32 fma
// chain for floating types, modular integer sums) and exceptions.  The FFT
// workspace overloads and the N-D / 2-D forms are not provided.
#pragma once

#include <cstddef>
#include <cstdint>
#include <limits>
#include <stdexcept>
#include <string>
#include <type_traits>

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

}  // namespace cusignal
