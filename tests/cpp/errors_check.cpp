// Host-only checks of the error taxonomy header (include/cusignal/runtime/
// errors.h): the reference's code values and aliases (errors.h:18-41), the
// name / value / parse helpers (:44-68, :123-141), the business dtype rule,
// the backend-status mapping, and the code an argument error carries.  No
// device call: runs wherever it compiles.  Prints one line per failed check
// and exits non-zero if any failed.
#include <iostream>
#include <string>

#include "cusignal/runtime/errors.h"

using cusignal::ErrorCode;

namespace {
int failures = 0;
void expect(bool ok, const std::string& what)
{
    if (!ok) {
        std::cout << "FAILED: " << what << "\n";
        ++failures;
    }
}
}  // namespace

int main()
{
    // numeric values of the public codes
    expect(cusignal::error_code_value(ErrorCode::SUCCESS) == 0, "SUCCESS = 0");
    expect(cusignal::error_code_value(ErrorCode::INVALID_ARGUMENT) == 1001, "INVALID_ARGUMENT = 1001");
    expect(cusignal::error_code_value(ErrorCode::INVALID_SHAPE) == 1002, "INVALID_SHAPE = 1002");
    expect(cusignal::error_code_value(ErrorCode::INVALID_DTYPE) == 1003, "INVALID_DTYPE = 1003");
    expect(cusignal::error_code_value(ErrorCode::CUDA_RUNTIME_ERROR) == 2002, "CUDA_RUNTIME_ERROR = 2002");
    expect(cusignal::error_code_value(ErrorCode::OUT_OF_MEMORY) == 2003, "OUT_OF_MEMORY = 2003");
    expect(cusignal::error_code_value(ErrorCode::BACKEND_ERROR) == 2099, "BACKEND_ERROR = 2099");
    expect(cusignal::error_code_value(ErrorCode::INTERNAL_ERROR) == 9001, "INTERNAL_ERROR = 9001");
    // aliases resolve to the canonical codes and print their names
    expect(ErrorCode::OK == ErrorCode::SUCCESS, "OK alias");
    expect(ErrorCode::INVALID_CONFIG == ErrorCode::CONFIG_ERROR, "INVALID_CONFIG alias");
    expect(ErrorCode::UNSUPPORTED == ErrorCode::UNSUPPORTED_OPERATION, "UNSUPPORTED alias");
    expect(ErrorCode::ALLOCATION_FAILED == ErrorCode::OUT_OF_MEMORY, "ALLOCATION_FAILED alias");
    expect(cusignal::error_code_name(ErrorCode::INVALID_CONFIG) == "CONFIG_ERROR", "alias name");
    expect(cusignal::error_code_name(static_cast<ErrorCode>(12345)) == "INTERNAL_ERROR", "unknown name");
    // every canonical name parses back to its code
    for (const auto& e : cusignal::detail::kCodeNames)
        expect(cusignal::parse_error_code(e.name) == e.code, std::string("parse ") + e.name);
    bool threw = false;
    try {
        (void)cusignal::parse_error_code("NOT_A_CODE");
    } catch (const cusignal::ProjectError& e) {
        threw = e.detail().code == ErrorCode::CONFIG_ERROR;
    }
    expect(threw, "unknown name is CONFIG_ERROR");
    // business dtypes
    for (const char* ok : {"FP32", "FP16", "INT32", "INT16", "INT8"}) {
        bool bad = false;
        try {
            cusignal::require_business_dtype(ok);
        } catch (...) {
            bad = true;
        }
        expect(!bad, std::string("dtype accepted ") + ok);
    }
    threw = false;
    try {
        cusignal::require_business_dtype("COMPLEX64");
    } catch (const cusignal::ProjectError& e) {
        threw = e.detail().code == ErrorCode::INVALID_DTYPE;
    }
    expect(threw, "complex dtype is INVALID_DTYPE");
    // backend status mapping
    const auto be = cusignal::map_backend_error(
        7, [](int) -> const char* { return "FFT_EXEC_FAILED"; }, nullptr);
    expect(be.code == "7" && be.name == "FFT_EXEC_FAILED" && be.message == "backend error message unavailable",
           "map_backend_error");
    // argument errors stay std::invalid_argument and keep their code
    try {
        throw cusignal::OperatorArgumentError(ErrorCode::INVALID_SHAPE, "shape");
    } catch (const std::invalid_argument& e) {
        expect(cusignal::error_code_of(e) == ErrorCode::INVALID_SHAPE, "argument error code");
    }
    try {
        cusignal::throw_error(ErrorCode::BACKEND_ERROR, "");
    } catch (const cusignal::ProjectError& e) {
        expect(cusignal::error_code_of(e) == ErrorCode::BACKEND_ERROR, "project error code");
        expect(!e.detail().message.empty(), "empty message replaced");
    }
    try {
        throw std::invalid_argument("plain");
    } catch (const std::exception& e) {
        expect(cusignal::error_code_of(e) == ErrorCode::INVALID_ARGUMENT, "plain invalid_argument code");
    }
    if (failures == 0) std::cout << "errors_check: all passed\n";
    return failures == 0 ? 0 : 1;
}
