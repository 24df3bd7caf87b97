// Error taxonomy of the operator API: the reference's public ErrorCode
// numbers and ProjectError exception
// (ZKX/cusignal_cpp/include/cusignal/runtime/errors.h:18-101), so callers that
// catch and classify errors keep working.
#pragma once

#include <stdexcept>
#include <string>
#include <utility>

#include "../../zkx_radar.h"

namespace cusignal {

enum class ErrorCode : int {
    SUCCESS = 0,
    INVALID_ARGUMENT = 1001,
    INVALID_SHAPE = 1002,
    INVALID_DTYPE = 1003,
    UNSUPPORTED_OPERATION = 1004,
    CONFIG_ERROR = 1005,
    CUDA_DRIVER_ERROR = 2001,
    CUDA_RUNTIME_ERROR = 2002,
    OUT_OF_MEMORY = 2003,
    BACKEND_ERROR = 2099,
    EXECUTION_FAILED = 2100,
    SYNC_FAILED = 2101,
    ACCURACY_FAILED = 3001,
    RESOURCE_FAILED = 3002,
    IO_ERROR = 4001,
    INTERNAL_ERROR = 9001,
    UPSTREAM_STEP_FAILED = 9002,
};

struct BackendError {
    std::string code{"NA"};
    std::string name{"NA"};
    std::string message{"NA"};
};

struct Error {
    ErrorCode code{ErrorCode::SUCCESS};
    std::string message;
    BackendError backend;
    bool safe_exit{true};
};

class ProjectError final : public std::runtime_error {
public:
    explicit ProjectError(Error error) : std::runtime_error(error.message), error_(std::move(error)) {}
    const Error& detail() const noexcept { return error_; }

private:
    Error error_;
};

[[noreturn]] inline void throw_error(ErrorCode code, std::string message, BackendError backend = {})
{
    throw ProjectError({code, std::move(message), std::move(backend), true});
}

namespace detail {
// Map a C ABI status onto the reference's exception behaviour: argument and
// shape problems are std::invalid_argument (as the reference's validation
// throws), everything else a ProjectError with the same ErrorCode.
inline void check_abi(int status)
{
    if (status == ZKX_SUCCESS) return;
    const std::string msg = zkx_last_error_message();
    if (status == ZKX_INVALID_ARGUMENT || status == ZKX_INVALID_SHAPE) throw std::invalid_argument(msg);
    throw_error(static_cast<ErrorCode>(status), msg);
}
}  // namespace detail

}  // namespace cusignal
