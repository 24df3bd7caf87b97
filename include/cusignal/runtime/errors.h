// Error taxonomy of the operator API: the reference's public ErrorCode
// numbers, aliases, name/value helpers and ProjectError exception
// (ZKX/cusignal_cpp/include/cusignal/runtime/errors.h:18-141), so callers that
// catch and classify errors keep working.
#pragma once

#include <functional>
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

    // older spellings kept by the reference for its earlier callers
    OK = SUCCESS,
    INVALID_CONFIG = CONFIG_ERROR,
    UNSUPPORTED = UNSUPPORTED_OPERATION,
    ALLOCATION_FAILED = OUT_OF_MEMORY,
};

namespace detail {
struct CodeName {
    ErrorCode code;
    const char* name;
};
// canonical names of the public codes (errors.h:44-68 of the reference)
inline constexpr CodeName kCodeNames[] = {
    {ErrorCode::SUCCESS, "SUCCESS"},
    {ErrorCode::INVALID_ARGUMENT, "INVALID_ARGUMENT"},
    {ErrorCode::INVALID_SHAPE, "INVALID_SHAPE"},
    {ErrorCode::INVALID_DTYPE, "INVALID_DTYPE"},
    {ErrorCode::UNSUPPORTED_OPERATION, "UNSUPPORTED_OPERATION"},
    {ErrorCode::CONFIG_ERROR, "CONFIG_ERROR"},
    {ErrorCode::CUDA_DRIVER_ERROR, "CUDA_DRIVER_ERROR"},
    {ErrorCode::CUDA_RUNTIME_ERROR, "CUDA_RUNTIME_ERROR"},
    {ErrorCode::OUT_OF_MEMORY, "OUT_OF_MEMORY"},
    {ErrorCode::BACKEND_ERROR, "BACKEND_ERROR"},
    {ErrorCode::EXECUTION_FAILED, "EXECUTION_FAILED"},
    {ErrorCode::SYNC_FAILED, "SYNC_FAILED"},
    {ErrorCode::ACCURACY_FAILED, "ACCURACY_FAILED"},
    {ErrorCode::RESOURCE_FAILED, "RESOURCE_FAILED"},
    {ErrorCode::IO_ERROR, "IO_ERROR"},
    {ErrorCode::INTERNAL_ERROR, "INTERNAL_ERROR"},
    {ErrorCode::UPSTREAM_STEP_FAILED, "UPSTREAM_STEP_FAILED"},
};
}  // namespace detail

// Canonical name of a code (aliases print as their canonical code); unknown
// values print as INTERNAL_ERROR.
inline std::string error_code_name(ErrorCode code)
{
    for (const auto& e : detail::kCodeNames)
        if (e.code == code) return e.name;
    return "INTERNAL_ERROR";
}

inline int error_code_value(ErrorCode code) noexcept { return static_cast<int>(code); }

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

[[noreturn]] inline void throw_error(ErrorCode code, std::string message, BackendError backend = {},
                                     bool safe_exit = true)
{
    if (message.empty()) message = "operation failed without a description";
    throw ProjectError({code, std::move(message), std::move(backend), safe_exit});
}

// A public code from its canonical name; an unknown name is a CONFIG_ERROR.
inline ErrorCode parse_error_code(const std::string& name)
{
    for (const auto& e : detail::kCodeNames)
        if (name == e.name) return e.code;
    throw_error(ErrorCode::CONFIG_ERROR, "unknown error code name: " + name);
}

// The business element types the operators accept (FP32, FP16, INT32,
// INT16, INT8); anything else is INVALID_DTYPE.
inline void require_business_dtype(const std::string& dtype)
{
    if (dtype == "FP32" || dtype == "FP16" || dtype == "INT32" || dtype == "INT16" || dtype == "INT8") return;
    throw_error(ErrorCode::INVALID_DTYPE, "unsupported input dtype: " + dtype + " (FP32/FP16/INT32/INT16/INT8)");
}

// Name and message of a raw backend status through the backend's resolvers.
inline BackendError map_backend_error(int raw_code, const std::function<const char*(int)>& name_resolver,
                                      const std::function<const char*(int)>& message_resolver)
{
    const char* name = name_resolver ? name_resolver(raw_code) : nullptr;
    const char* message = message_resolver ? message_resolver(raw_code) : nullptr;
    return {std::to_string(raw_code), name ? name : "UNKNOWN_BACKEND_ERROR",
            message ? message : "backend error message unavailable"};
}

// An argument or shape error: a std::invalid_argument (what the reference's
// validation throws, so existing catch sites keep working) that also carries
// the public code the C ABI reported.
class OperatorArgumentError final : public std::invalid_argument {
public:
    OperatorArgumentError(ErrorCode code, const std::string& message) : std::invalid_argument(message), code_(code) {}
    ErrorCode code() const noexcept { return code_; }

private:
    ErrorCode code_;
};

// The public code of an exception thrown by the operator layer.
inline ErrorCode error_code_of(const std::exception& e) noexcept
{
    if (const auto* a = dynamic_cast<const OperatorArgumentError*>(&e)) return a->code();
    if (const auto* p = dynamic_cast<const ProjectError*>(&e)) return p->detail().code;
    if (dynamic_cast<const std::invalid_argument*>(&e)) return ErrorCode::INVALID_ARGUMENT;
    return ErrorCode::INTERNAL_ERROR;
}

namespace detail {
// Map a C ABI status onto the reference's exception behaviour: argument,
// shape and dtype problems are std::invalid_argument (as the reference's
// validation throws; the code travels in OperatorArgumentError), everything
// else a ProjectError with the same ErrorCode.
inline void check_abi(int status)
{
    if (status == ZKX_SUCCESS) return;
    const std::string msg = zkx_last_error_message();
    if (status == ZKX_INVALID_ARGUMENT || status == ZKX_INVALID_SHAPE || status == ZKX_INVALID_DTYPE)
        throw OperatorArgumentError(static_cast<ErrorCode>(status), msg);
    throw_error(static_cast<ErrorCode>(status), msg);
}
}  // namespace detail

}  // namespace cusignal
