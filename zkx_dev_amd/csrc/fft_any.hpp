// Batched complex transforms of any length on gfx950, for the FFT backend ABI
// and the operators whose transform length is not a power of two <= 8192:
//   * powers of two <= 8192: one LDS Stockham pass per row;
//   * powers of two up to 2^26: four-step, n = n1 * n2 with both factors
//     <= 8192 (transpose, n1-point rows, twiddle + transpose, n2-point rows,
//     transpose);
//   * other lengths up to 2^25: Bluestein's chirp-z method over a power-of-two
//     length M >= 2n - 1 in global passes (chirp, FFT_M, x kernel spectrum,
//     IFFT_M, chirp), with FP64-built chirp tables cached per (device, n).
// Forward transforms are unscaled; y = scale * (unscaled transform of x).
#pragma once

#include "zkx_hip_common.hpp"

#include <cstddef>

namespace zkx {

constexpr long long kMaxAnyFftPow2 = 1LL << 26;
constexpr long long kMaxAnyFftOther = 1LL << 25;

// true when fft_any_rows supports length n
bool fft_any_length_ok(long long n);

// Device workspace bytes one call of fft_any_rows(rows, n) needs (0 when the
// length runs in a single LDS pass).
size_t fft_any_workspace_bytes(long long rows, int n);

// Build the tables a length needs (twiddles, chirp spectra) now, so that a
// later fft_any_rows call allocates nothing beyond its workspace.
int fft_any_prepare(int n);

// rows x n complex, row-major; x == y allowed.  `ws` must hold
// fft_any_workspace_bytes(rows, n) bytes (NULL when that is 0).  Work is
// enqueued on `st`; the caller keeps `ws` alive and unshared until it ends.
int fft_any_rows(const float2* x, long long rows, int n, bool inverse, float scale, float2* y, void* ws,
                 hipStream_t st);

// Full linear convolution y[0 .. la+lb-1) of complex a and b by whole-length
// power-of-two transforms (M = next power of two >= la + lb - 1, at most
// 2^26): both operands zero-padded to M, transformed together, multiplied,
// inverse transformed with 1/M.  ws: fft_convolve_workspace_bytes bytes.
size_t fft_convolve_workspace_bytes(long long la, long long lb);
int fft_convolve_full(const float2* a, long long la, const float2* b, long long lb, float2* y, void* ws,
                      hipStream_t st);

}  // namespace zkx
