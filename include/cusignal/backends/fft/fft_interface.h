// The FFT backend interface the reference's operators are written against
// (ZKX/cusignal_cpp/include/cusignal/backends/fft/fft_interface.h): the fft_*
// C ABI (include/zkx_fft.h) and the ComplexFloat / ComplexDouble element
// types, laid out as {re, im}.
#pragma once

#include "../../../zkx_fft.h"

extern "C" {
typedef struct {
    float re;
    float im;
} ComplexFloat;
typedef struct {
    double re;
    double im;
} ComplexDouble;
}

static_assert(sizeof(ComplexFloat) == 8, "ComplexFloat layout");
static_assert(sizeof(ComplexDouble) == 16, "ComplexDouble layout");
