// N-dimensional FFT convolution engine (declared in conv_nd.hpp) and the
// fftconvolve entry points over N-D axes.
//
// fftconvolve follows cuSignal (ZKX/cusignal-23.08.00/python/cusignal/
// convolution/convolve.py:159-320, convolution_utils.py:40-100):
//   * in1 and in2 have one rank; the convolved axes default to all; on the
//     other axes the extents must be equal or 1 (broadcast);
//   * full extent: s1 + s2 - 1 on a convolved axis, max(s1, s2) elsewhere;
//   * 'valid' swaps the operands when in1 is not at least as large as in2
//     in every dimension (and fails when neither is); its extent is
//     s1 - s2 + 1 on a convolved axis, the full extent elsewhere;
//   * 'same' keeps in1's extents in every dimension;
//   * the kept block is centred: start = (full - kept) / 2 per dimension;
//   * a real result for two real operands, complex otherwise.
// cuSignal pads each axis to a 5-smooth length for FFTPACK; this engine pads
// to a power of two, which changes only the rounding of the result.
#include "conv_nd.hpp"
#include "fft_any.hpp"
#include "fft_lds.hpp"
#include "runtime.hpp"

#include <algorithm>

namespace zkx {
namespace {

long long next_pow2_ll(long long n)
{
    long long p = 1;
    while (p < n) p <<= 1;
    return p;
}

size_t al256(size_t b) { return (b + 255) / 256 * 256; }

long long count_of(const int* d, int rank)
{
    long long c = 1;
    for (int k = 0; k < rank; ++k) c *= d[k];
    return c;
}

// out[b][c][r] = in[b][r][c] for b < B, r < R and c in [c_lo, c_hi) of C
// columns, with rows r >= r_live read as zeros (not read at all); output
// rows outside [c_lo, c_hi) are left as they were.  32 x 32 tiles through LDS
// (one padding column keeps the column reads on distinct banks).  The tile
// index runs over blockIdx.x (either extent may be large), the batch over
// blockIdx.y.
__global__ void __launch_bounds__(256)
nd_transpose_kernel(const float2* __restrict__ in, long long B, int R, int C, int r_live, int c_lo, int c_hi,
                    float2* __restrict__ out)
{
    __shared__ float2 tile[32][33];
    const int tx = threadIdx.x & 31, ty = threadIdx.x >> 5;  // 32 x 8
    const int tiles_c = (c_hi - c_lo + 31) / 32;
    const int r0 = static_cast<int>(blockIdx.x / tiles_c) * 32;
    const int c0 = c_lo + static_cast<int>(blockIdx.x % tiles_c) * 32;
    const long long plane = static_cast<long long>(R) * C;
    for (long long b = blockIdx.y; b < B; b += gridDim.y) {
        const float2* src = in + b * plane;
        float2* dst = out + b * plane;
#pragma unroll
        for (int k = 0; k < 4; ++k) {
            const int r = r0 + ty + 8 * k, c = c0 + tx;
            if (r < R && c < c_hi)
                tile[ty + 8 * k][tx] = r < r_live ? src[static_cast<long long>(r) * C + c] : make_float2(0.f, 0.f);
        }
        __syncthreads();
#pragma unroll
        for (int k = 0; k < 4; ++k) {
            const int c = c0 + ty + 8 * k, r = r0 + tx;
            if (r < R && c < c_hi) dst[static_cast<long long>(c) * R + r] = tile[tx][ty + 8 * k];
        }
        __syncthreads();
    }
}

int transpose(const float2* in, long long B, int R, int C, float2* out, hipStream_t st, int r_live = -1,
              int c_lo = 0, int c_hi = -1)
{
    if (r_live < 0) r_live = R;
    if (c_hi < 0) c_hi = C;
    if (c_hi <= c_lo) return ZKX_SUCCESS;
    const long long tiles = static_cast<long long>((R + 31) / 32) * ((c_hi - c_lo + 31) / 32);
    if (tiles > 0x7fffffffLL) ZKX_FAIL(ZKX_INVALID_SHAPE, "fft convolve: transpose extent too large");
    const dim3 grid(static_cast<unsigned>(tiles), static_cast<unsigned>(std::min<long long>(B, 65535)));
    hipLaunchKernelGGL(nd_transpose_kernel, grid, dim3(256), 0, st, in, B, R, C, r_live, c_lo, c_hi, out);
    ZKX_LAUNCH_CHECK();
    return ZKX_SUCCESS;
}

// Y (extents g) = X1 (g1) * X2 (g2), broadcasting extents of 1.  Y may be X1
// when g1 == g.
struct NdMul {
    int rank;
    int g[kMaxNdRank];
    long long s1[kMaxNdRank], s2[kMaxNdRank];  // strides of X1 / X2 per dim of g (0 when broadcast)
};

__global__ void __launch_bounds__(256)
nd_mul_kernel(const float2* X1, const float2* __restrict__ X2, NdMul p, float2* Y)
{
    const int last = p.rank - 1;
    const int gl = p.g[last];
    const int j0 = blockIdx.x * kNdSpan + threadIdx.x;
    const long long rows = nd_rows(p.g, p.rank);
    for (long long r = blockIdx.y; r < rows; r += gridDim.y) {
        long long q = r, o1 = 0, o2 = 0;
        for (int d = last - 1; d >= 0; --d) {
            const long long c = q % p.g[d];
            q /= p.g[d];
            o1 += c * p.s1[d];
            o2 += c * p.s2[d];
        }
#pragma unroll
        for (int k = 0; k < kNdPer; ++k) {
            const int j = j0 + k * 256;
            if (j >= gl) break;
            Y[r * gl + j] = cmul(X1[o1 + j * p.s1[last]], X2[o2 + j * p.s2[last]]);
        }
    }
}

// Y = A^ B^ from Z = A^ + i B^ (A, B real): per element
// (Z[k]^2 - conj(Z[-k])^2) / 4i, with -k taken modulo the extent along the
// dimensions in `neg` and k itself along the others.  Y must not alias Z.
struct NdPacked {
    int rank;
    int g[kMaxNdRank];
    unsigned neg;            // bit k: position k is a transformed dimension
    const unsigned* absmax;  // the packing's maxima (the scale s of B)
};

__global__ void __launch_bounds__(256)
nd_mul_packed_kernel(const float2* __restrict__ Z, NdPacked p, float2* __restrict__ Y)
{
    const int last = p.rank - 1;
    const int gl = p.g[last];
    const bool negl = p.neg >> last & 1u;
    const float q4 = ldexpf(0.25f, -nd_pair_shift(p.absmax));  // 1 / (4 s)
    const int j0 = blockIdx.x * kNdSpan + threadIdx.x;
    const long long rows = nd_rows(p.g, p.rank);
    for (long long r = blockIdx.y; r < rows; r += gridDim.y) {
        // partner row: every leading coordinate negated where transformed
        long long q = r, pr = 0, mul = 1;
        for (int d = last - 1; d >= 0; --d) {
            const int c = static_cast<int>(q % p.g[d]);
            q /= p.g[d];
            const int cn = (p.neg >> d & 1u) ? (c == 0 ? 0 : p.g[d] - c) : c;
            pr += static_cast<long long>(cn) * mul;
            mul *= p.g[d];
        }
        const float2* zr = Z + r * gl;
        const float2* zp = Z + pr * gl;
        float2* yr = Y + r * gl;
#pragma unroll
        for (int k = 0; k < kNdPer; ++k) {
            const int j = j0 + k * 256;
            if (j >= gl) break;
            const int jn = negl ? (j == 0 ? 0 : gl - j) : j;
            const float2 a = zr[j], b = zp[jn];
            // a^2 - conj(b)^2 = (ax^2 - ay^2 - bx^2 + by^2) + i (2 ax ay + 2 bx by)
            const float re = (a.x * a.x - a.y * a.y) - (b.x * b.x - b.y * b.y);
            const float im = 2.0f * (a.x * a.y + b.x * b.y);
            // divided by 4 i s: (re + i im) / (4 i s) = (im - i re) / (4 s)
            yr[j] = make_float2(q4 * im, -q4 * re);
        }
    }
}

// The same product with the last dimension's forward and inverse transforms
// fused around it (power-of-two N = that extent, always a transformed
// dimension here): a block owns row r of Z and its partner row (every leading
// coordinate negated where transformed), one LDS slot each.  Each slot
// transforms its row, stores the spectrum in its slot, reads the partner's
// bin -j, forms the product in registers and inverse transforms it (scaled by
// 1/N) into the same row of Y.  A block whose partner row precedes its own
// leaves both rows to the partner's block; a self-paired row runs in both
// slots and slot 0 stores it.  Y may be Z: a block reads and writes only its
// own two rows.
template <int N>
__global__ void __launch_bounds__(2 * (N / fft::elems_for(N)), N == 4096 ? 4 : 1)
nd_packed_rows_kernel(const float2* Z, NdPacked p, const float2* __restrict__ tw, float2* Y)
{
    constexpr int E = fft::elems_for(N), T = N / E, SL = fft::RowLds::floats_for(N);
    extern __shared__ float lds[];
    const int slot = threadIdx.x / T, t = threadIdx.x % T;
    const int last = p.rank - 1;
    const long long r = blockIdx.x;
    long long q = r, pr = 0, mul = 1;
    for (int d = last - 1; d >= 0; --d) {
        const int c = static_cast<int>(q % p.g[d]);
        q /= p.g[d];
        const int cn = (p.neg >> d & 1u) ? (c == 0 ? 0 : p.g[d] - c) : c;
        pr += static_cast<long long>(cn) * mul;
        mul *= p.g[d];
    }
    if (pr < r) return;  // the whole block: the partner's block does both rows
    const long long row = slot == 0 ? r : pr;
    const int other = pr == r ? slot : 1 - slot;
    const fft::RowLds L{reinterpret_cast<float2*>(lds + slot * 2 * SL)};
    const fft::RowLds P{reinterpret_cast<float2*>(lds + other * 2 * SL)};
    const float2* src = Z + row * N;
    float2 v[E];
#pragma unroll
    for (int k = 0; k < E; ++k) v[k] = src[t + k * T];
    fft::transform<N, E, false>(v, L, tw, t);
    __syncthreads();  // the transform's last LDS reads are done
#pragma unroll
    for (int k = 0; k < E; ++k) L.store(t + k * T, v[k]);
    __syncthreads();
    const float q4 = ldexpf(0.25f, -nd_pair_shift(p.absmax));  // 1 / (4 s)
#pragma unroll
    for (int k = 0; k < E; ++k) {
        const int j = t + k * T;
        const float2 a = v[k], b = P.load((N - j) & (N - 1));
        const float re = (a.x * a.x - a.y * a.y) - (b.x * b.x - b.y * b.y);
        const float im = 2.0f * (a.x * a.y + b.x * b.y);
        v[k] = make_float2(q4 * im, -q4 * re);
    }
    __syncthreads();  // the partner's reads of this slot are done
    fft::transform<N, E, true>(v, L, tw, t);
    if (slot == 0 || pr != r) {
        float2* dst = Y + row * N;
        const float sc = 1.0f / static_cast<float>(N);
#pragma unroll
        for (int k = 0; k < E; ++k) dst[t + k * T] = cscale(v[k], sc);
    }
}

template <int N>
int launch_packed_rows(const float2* Z, const NdPacked& m, long long rows, float2* Y, hipStream_t st)
{
    constexpr int T = N / fft::elems_for(N);
    constexpr size_t bytes = sizeof(float) * 2 * 2 * fft::RowLds::floats_for(N);
    const float2* tw = nullptr;
    int rc = twiddles(N, &tw);
    if (rc) return rc;
    if (rows > 0x7fffffffLL) ZKX_FAIL(ZKX_INVALID_SHAPE, "fft convolve: too many rows");
    if (bytes > 64 * 1024)
        ZKX_HIP_CHECK(hipFuncSetAttribute(reinterpret_cast<const void*>(nd_packed_rows_kernel<N>),
                                          hipFuncAttributeMaxDynamicSharedMemorySize, static_cast<int>(bytes)));
    hipLaunchKernelGGL(nd_packed_rows_kernel<N>, dim3(static_cast<unsigned>(rows)), dim3(2 * T), bytes, st, Z, m,
                       tw, Y);
    ZKX_LAUNCH_CHECK();
    return ZKX_SUCCESS;
}

// Last extents the fused product handles (one LDS pass, at least a wave per
// block, both rows' registers within a 512-thread block's budget).
bool packed_rows_fused(int n) { return n >= 256 && n <= 4096 && (n & (n - 1)) == 0; }

// An operand during the spectral part: its buffer, its extents in the
// current dimension order and which original dimension each position holds.
struct NdArr {
    float2* p;
    int rank;
    int dims[kMaxNdRank];
    int orig[kMaxNdRank];
    long long count() const
    {
        long long c = 1;
        for (int k = 0; k < rank; ++k) c *= dims[k];
        return c;
    }
};

NdArr nd_arr(float2* p, const int* g, int rank)
{
    NdArr a{};
    a.p = p;
    a.rank = rank;
    for (int k = 0; k < rank; ++k) {
        a.dims[k] = g[k];
        a.orig[k] = k;
    }
    return a;
}

// Move the last dimension to the front ([outer][last] -> [last][outer]),
// or the first to the back, through the spare buffer, which then holds the
// old storage.  last_to_front: outer rows from `live` on are zeros (read as
// such); first_to_back: only the new leading-dimension rows [keep_lo,
// keep_hi) are produced (all when keep_hi is 0).
int rotate(NdArr& a, bool last_to_front, float2*& spare, hipStream_t st, long long live = -1, int keep_lo = 0,
           int keep_hi = 0)
{
    const int R = a.rank;
    const long long total = a.count();
    int rc;
    if (last_to_front) {
        const int c = a.dims[R - 1];
        const long long outer = total / c;
        rc = transpose(a.p, 1, static_cast<int>(outer), c, spare, st,
                       static_cast<int>(live < 0 || live > outer ? outer : live));
        const int od = a.orig[R - 1];
        for (int k = R - 1; k > 0; --k) {
            a.dims[k] = a.dims[k - 1];
            a.orig[k] = a.orig[k - 1];
        }
        a.dims[0] = c;
        a.orig[0] = od;
    } else {
        const int f = a.dims[0];
        const long long C = total / f;
        // output row c = (coordinate of the new leading dimension) * inner + ...
        const long long inner = R >= 2 ? C / a.dims[1] : C;
        const bool win = keep_hi > keep_lo && R >= 2;
        rc = transpose(a.p, 1, f, static_cast<int>(C), spare, st, f, win ? static_cast<int>(keep_lo * inner) : 0,
                       win ? static_cast<int>(keep_hi * inner) : -1);
        const int od = a.orig[0];
        for (int k = 0; k + 1 < R; ++k) {
            a.dims[k] = a.dims[k + 1];
            a.orig[k] = a.orig[k + 1];
        }
        a.dims[R - 1] = f;
        a.orig[R - 1] = od;
    }
    if (rc) return rc;
    float2* t = a.p;
    a.p = spare;
    spare = t;
    return ZKX_SUCCESS;
}

// Row transforms along the current last dimension when it is convolved,
// over rows [r_lo, r_hi) (r_hi < 0: to the end).
int fft_last(NdArr& a, unsigned axes, bool inverse, char* any_ws, hipStream_t st, long long r_lo = 0,
             long long r_hi = -1)
{
    const int d = a.orig[a.rank - 1];
    const int n = a.dims[a.rank - 1];
    if (!(axes >> d & 1u) || n == 1) return ZKX_SUCCESS;
    const long long rows = a.count() / n;
    if (r_hi < 0 || r_hi > rows) r_hi = rows;
    if (r_hi <= r_lo) return ZKX_SUCCESS;
    const float scale = inverse ? 1.0f / static_cast<float>(n) : 1.0f;
    float2* x = a.p + r_lo * n;
    return fft_any_rows(x, r_hi - r_lo, n, inverse, scale, x, any_ws, st);
}

// Number of rotations that bring the lowest convolved dimension last.
int nd_rotations(unsigned axes, int rank)
{
    int lo = 0;
    while (!(axes >> lo & 1u)) ++lo;
    return rank - 1 - lo;
}

// Forward: transform the last dimension, rotate it to the front, and so on
// until every convolved dimension has been last once (S rotations); the
// spectrum stays in the rotated order.  skip_final leaves the last
// dimension untransformed (a fused product kernel does it).  Rows from
// `live` on (original order) are zeros: the first transform and rotation
// leave them out.
int forward_nd(NdArr& a, unsigned axes, float2*& spare, char* any_ws, hipStream_t st, bool skip_final,
               long long live)
{
    const int S = nd_rotations(axes, a.rank);
    for (int s = 0; s <= S; ++s) {
        int rc = s == S && skip_final ? ZKX_SUCCESS : fft_last(a, axes, false, any_ws, st, 0, s == 0 ? live : -1);
        if (rc) return rc;
        if (s < S && (rc = rotate(a, true, spare, st, s == 0 ? live : -1))) return rc;
    }
    return ZKX_SUCCESS;
}

// Inverse of forward_nd's order: transform the last dimension, rotate the
// first to the back, ..., ending in the original order.  skip_first: the
// last dimension is already inverse transformed.  With keep_hi > keep_lo
// (rank >= 2) the last rotation and transform produce only the rows whose
// dimension-0 coordinate is in [keep_lo, keep_hi).
int inverse_nd(NdArr& a, unsigned axes, float2*& spare, char* any_ws, hipStream_t st, bool skip_first, int keep_lo,
               int keep_hi)
{
    const int S = nd_rotations(axes, a.rank);
    const bool win = keep_hi > keep_lo && a.rank >= 2;
    for (int s = 0; s <= S; ++s) {
        int rc;
        if (s == 0 && skip_first) {
            rc = ZKX_SUCCESS;
        } else if (s == S && win) {
            // original order: rows of dimension-0 coordinate c are c*inner ..
            const long long inner = a.count() / a.dims[a.rank - 1] / a.dims[0];
            rc = fft_last(a, axes, true, any_ws, st, keep_lo * inner, keep_hi * inner);
        } else {
            rc = fft_last(a, axes, true, any_ws, st);
        }
        if (rc) return rc;
        if (s < S && (rc = s == S - 1 && win ? rotate(a, false, spare, st, -1, keep_lo, keep_hi)
                                             : rotate(a, false, spare, st)))
            return rc;
    }
    return ZKX_SUCCESS;
}

}  // namespace

int nd_conv_plan(const int* s1, const int* s2, int rank, unsigned axes, NdConvPlan* p)
{
    if (rank < 1 || rank > kMaxNdRank) ZKX_FAIL(ZKX_INVALID_SHAPE, "fft convolve: rank must be 1..8");
    NdConvPlan q;
    q.rank = rank;
    q.axes = axes;
    size_t any = 0;
    for (int d = 0; d < rank; ++d) {
        if (s1[d] < 1 || s2[d] < 1) ZKX_FAIL(ZKX_INVALID_SHAPE, "fft convolve: empty extent");
        q.s1[d] = s1[d];
        q.s2[d] = s2[d];
        if (axes >> d & 1u) {
            const long long f = next_pow2_ll(static_cast<long long>(s1[d]) + s2[d] - 1);
            if (f > kMaxAnyFftPow2) ZKX_FAIL(ZKX_INVALID_SHAPE, "fft convolve: axis too long");
            q.g1[d] = q.g2[d] = q.g[d] = static_cast<int>(f);
        } else {
            if (s1[d] != s2[d] && s1[d] != 1 && s2[d] != 1)
                ZKX_FAIL(ZKX_INVALID_SHAPE, "fft convolve: incompatible shapes for in1 and in2");
            q.g1[d] = s1[d];
            q.g2[d] = s2[d];
            q.g[d] = std::max(s1[d], s2[d]);
        }
    }
    q.n1 = count_of(q.g1, rank);
    q.n2 = count_of(q.g2, rank);
    q.n = count_of(q.g, rank);
    if (q.n1 > kMaxNdElements || q.n2 > kMaxNdElements || q.n > kMaxNdElements)
        ZKX_FAIL(ZKX_INVALID_SHAPE, "fft convolve: padded operands exceed 2^31 elements");
    q.y_is_x1 = q.n1 == q.n;
    const long long big = std::max(q.n1, std::max(q.n2, q.n));
    for (int d = 0; d < rank; ++d) {
        if (!(axes >> d & 1u)) continue;
        any = std::max(any, fft_any_workspace_bytes(big / q.g[d], q.g[d]));
    }
    // four buffers of the largest operand: X1, X2 and two that rotations and
    // the product move between
    const size_t buf = al256(sizeof(float2) * big);
    q.off_x1 = 0;
    q.off_x2 = buf;
    q.off_y = 2 * buf;
    q.off_t = 3 * buf;
    q.off_any = 4 * buf;
    q.off_scal = 4 * buf + al256(any);
    q.bytes = q.off_scal + 256 + sizeof(float) * 2 * kNdAbsBlocks;
    *p = q;
    return ZKX_SUCCESS;
}

int nd_conv_run(const NdConvPlan& p, char* ws, float2** y, hipStream_t st)
{
    for (int d = 0; d < p.rank; ++d) {
        if (!(p.axes >> d & 1u)) continue;
        int rc = fft_any_prepare(p.g[d]);
        if (rc) return rc;
    }
    char* any = ws + p.off_any;
    float2* spare = reinterpret_cast<float2*>(ws + p.off_t);
    float2* free_buf = reinterpret_cast<float2*>(ws + p.off_y);
    NdArr a = nd_arr(reinterpret_cast<float2*>(ws + p.off_x1), p.g1, p.rank);
    NdArr b = nd_arr(reinterpret_cast<float2*>(ws + p.off_x2), p.g2, p.rank);
    int rc = forward_nd(a, p.axes, spare, any, st, false, nd_live_rows(p.s1, p.g1, p.rank));
    if (rc) return rc;
    if ((rc = forward_nd(b, p.axes, spare, any, st, false, nd_live_rows(p.s2, p.g2, p.rank)))) return rc;
    // product in the rotated order (the same for both operands)
    NdArr c = a;
    c.p = p.y_is_x1 ? a.p : free_buf;
    for (int k = 0; k < p.rank; ++k) c.dims[k] = p.g[c.orig[k]];
    NdMul m{};
    m.rank = p.rank;
    long long st1[kMaxNdRank], st2[kMaxNdRank];
    nd_strides(a.dims, p.rank, st1);
    nd_strides(b.dims, p.rank, st2);
    for (int k = 0; k < p.rank; ++k) {
        m.g[k] = c.dims[k];
        m.s1[k] = a.dims[k] == 1 && c.dims[k] > 1 ? 0 : st1[k];
        m.s2[k] = b.dims[k] == 1 && c.dims[k] > 1 ? 0 : st2[k];
    }
    const int last = c.dims[p.rank - 1];
    hipLaunchKernelGGL(nd_mul_kernel, nd_grid(c.count() / last, last), dim3(256), 0, st, a.p, b.p, m, c.p);
    ZKX_LAUNCH_CHECK();
    // b's buffer is free now: it serves the inverse rotations
    float2* spare2 = b.p;
    if ((rc = inverse_nd(c, p.axes, spare2, any, st, false, p.keep_lo0, p.keep_hi0))) return rc;
    *y = c.p;
    return ZKX_SUCCESS;
}

int nd_conv_run_real_pair(const NdConvPlan& p, char* ws, float2** y, hipStream_t st)
{
    for (int d = 0; d < p.rank; ++d) {
        if (!(p.axes >> d & 1u)) continue;
        int rc = fft_any_prepare(p.g[d]);
        if (rc) return rc;
    }
    char* any = ws + p.off_any;
    float2* spare = reinterpret_cast<float2*>(ws + p.off_t);
    NdArr z = nd_arr(reinterpret_cast<float2*>(ws + p.off_x1), p.g1, p.rank);
    // after the rotations the last dimension is the lowest convolved one
    const int lo = [&] { int d = 0; while (!(p.axes >> d & 1u)) ++d; return d; }();
    const bool fused = packed_rows_fused(p.g[lo]);
    int smax[kMaxNdRank];
    for (int d = 0; d < p.rank; ++d) smax[d] = std::max(p.s1[d], p.s2[d]);
    int rc = forward_nd(z, p.axes, spare, any, st, fused, nd_live_rows(smax, p.g1, p.rank));
    if (rc) return rc;
    NdArr c = z;
    c.p = reinterpret_cast<float2*>(ws + p.off_y);
    NdPacked m{};
    m.rank = p.rank;
    m.absmax = reinterpret_cast<const unsigned*>(ws + p.off_scal);
    for (int k = 0; k < p.rank; ++k) {
        m.g[k] = z.dims[k];
        if (p.axes >> z.orig[k] & 1u) m.neg |= 1u << k;
    }
    const int last = z.dims[p.rank - 1];
    if (fused) {
        const long long rows = z.count() / last;
        switch (last) {
        case 256: rc = launch_packed_rows<256>(z.p, m, rows, c.p, st); break;
        case 512: rc = launch_packed_rows<512>(z.p, m, rows, c.p, st); break;
        case 1024: rc = launch_packed_rows<1024>(z.p, m, rows, c.p, st); break;
        case 2048: rc = launch_packed_rows<2048>(z.p, m, rows, c.p, st); break;
        default: rc = launch_packed_rows<4096>(z.p, m, rows, c.p, st); break;
        }
        if (rc) return rc;
    } else {
        hipLaunchKernelGGL(nd_mul_packed_kernel, nd_grid(z.count() / last, last), dim3(256), 0, st, z.p, m, c.p);
        ZKX_LAUNCH_CHECK();
    }
    float2* spare2 = z.p;  // Z is no longer needed
    if ((rc = inverse_nd(c, p.axes, spare2, any, st, fused, p.keep_lo0, p.keep_hi0))) return rc;
    *y = c.p;
    return ZKX_SUCCESS;
}

}  // namespace zkx

using namespace zkx;

namespace {

thread_local ThreadWorkspace t_convnd_ws;

struct FftconvShape {
    int rank = 0;
    int s1[kMaxNdRank] = {}, s2[kMaxNdRank] = {};
    int full[kMaxNdRank] = {}, out[kMaxNdRank] = {};
    bool swap = false;
};

// Extents of fftconvolve's result (convolve.py:254-320 of cuSignal).
int fftconv_shape(const long long* shape1, const long long* shape2, int rank, unsigned axes, int mode,
                  FftconvShape* r)
{
    if (rank < 1 || rank > kMaxNdRank) ZKX_FAIL(ZKX_INVALID_SHAPE, "fftconvolve: rank must be 1..8");
    if (!shape1 || !shape2) ZKX_FAIL(ZKX_INVALID_ARGUMENT, "fftconvolve: null shape");
    if (mode < 0 || mode > 2) ZKX_FAIL(ZKX_INVALID_ARGUMENT, "acceptable mode flags are 'valid', 'same', or 'full'");
    if (axes == 0 || (rank < 32 && (axes >> rank) != 0))
        ZKX_FAIL(ZKX_INVALID_ARGUMENT, "fftconvolve: axes must name dimensions of the inputs");
    FftconvShape q;
    q.rank = rank;
    for (int d = 0; d < rank; ++d) {
        if (shape1[d] < 1 || shape2[d] < 1 || shape1[d] > 0x7fffffffLL || shape2[d] > 0x7fffffffLL)
            ZKX_FAIL(ZKX_INVALID_SHAPE, "fftconvolve: extents must be 1 .. 2^31-1");
        q.s1[d] = static_cast<int>(shape1[d]);
        q.s2[d] = static_cast<int>(shape2[d]);
        if (!(axes >> d & 1u) && q.s1[d] != q.s2[d] && q.s1[d] != 1 && q.s2[d] != 1)
            ZKX_FAIL(ZKX_INVALID_SHAPE, "incompatible shapes for in1 and in2");
        const long long f = (axes >> d & 1u) ? shape1[d] + shape2[d] - 1 : std::max(shape1[d], shape2[d]);
        if (f > 0x7fffffffLL) ZKX_FAIL(ZKX_INVALID_SHAPE, "fftconvolve: extent too large");
        q.full[d] = static_cast<int>(f);
    }
    if (mode == 2) {
        bool ok1 = true, ok2 = true;
        for (int d = 0; d < rank; ++d) {
            ok1 = ok1 && q.s1[d] >= q.s2[d];
            ok2 = ok2 && q.s2[d] >= q.s1[d];
        }
        if (!ok1 && !ok2)
            ZKX_FAIL(ZKX_INVALID_SHAPE, "For 'valid' mode, one must be at least as large as the other in every dimension");
        q.swap = !ok1;
    }
    for (int d = 0; d < rank; ++d) {
        const int a = q.swap ? q.s2[d] : q.s1[d], b = q.swap ? q.s1[d] : q.s2[d];
        if (mode == 0) q.out[d] = q.full[d];
        else if (mode == 1) q.out[d] = q.s1[d];
        else q.out[d] = (axes >> d & 1u) ? a - b + 1 : q.full[d];
    }
    *r = q;
    return ZKX_SUCCESS;
}

}  // namespace

extern "C" {

int zkx_fftconvolve_nd_shape(const long long* shape1, const long long* shape2, int rank, unsigned axes_mask, int mode,
                             long long* out_shape)
{
    clear_last_error();
    if (!out_shape) ZKX_FAIL(ZKX_INVALID_ARGUMENT, "fftconvolve: null output shape");
    FftconvShape q;
    int rc = fftconv_shape(shape1, shape2, rank, axes_mask, mode, &q);
    if (rc) return rc;
    for (int d = 0; d < rank; ++d) out_shape[d] = q.out[d];
    return ZKX_SUCCESS;
}

int zkx_fftconvolve_nd(const void* in1, const long long* shape1, int complex1, const void* in2,
                       const long long* shape2, int complex2, int rank, unsigned axes_mask, int mode, void* out,
                       void* stream)
{
    clear_last_error();
    hipStream_t st = (hipStream_t)stream;
    FftconvShape q;
    int rc = fftconv_shape(shape1, shape2, rank, axes_mask, mode, &q);
    if (rc) return rc;
    if (!in1 || !in2 || !out) ZKX_FAIL(ZKX_INVALID_ARGUMENT, "fftconvolve: null pointer");
    // one complex kernel row (leading extents 1, at most 4096 taps) against
    // every row of a complex in1 along the last axis: overlap-save rows,
    // without padding each row to a power of two
    bool rows_form = complex1 && complex2 && axes_mask == (1u << (rank - 1)) && !q.swap && q.s2[rank - 1] <= 4096;
    long long lead = 1;
    for (int d = 0; d + 1 < rank; ++d) {
        rows_form = rows_form && q.s2[d] == 1;
        lead *= q.s1[d];
    }
    if (rows_form && lead <= 0x7fffffffLL)
        return zkx_fftconvolve_c64(static_cast<const zkx_complex64*>(in1), static_cast<int>(lead), q.s1[rank - 1],
                                   static_cast<const zkx_complex64*>(in2), q.s2[rank - 1], mode,
                                   static_cast<zkx_complex64*>(out), stream);
    // convolution commutes: 'valid' takes the larger operand first
    const void* a = q.swap ? in2 : in1;
    const void* b = q.swap ? in1 : in2;
    const int ca = q.swap ? complex2 : complex1, cb = q.swap ? complex1 : complex2;
    int sa[kMaxNdRank], sb[kMaxNdRank];
    for (int d = 0; d < rank; ++d) {
        sa[d] = q.swap ? q.s2[d] : q.s1[d];
        sb[d] = q.swap ? q.s1[d] : q.s2[d];
    }
    NdConvPlan p;
    rc = nd_conv_plan(sa, sb, rank, axes_mask, &p);
    if (rc) return rc;
    void* wsv = nullptr;
    WorkspaceLease lease_(t_convnd_ws, st);
    rc = lease_.get(p.bytes, &wsv);
    if (rc) return rc;
    char* ws = static_cast<char*>(wsv);
    auto pad = [&](const void* src, int cplx, const int* s, const int* g, float2* X) -> int {
        NdPad pp{};
        pp.rank = rank;
        for (int d = 0; d < rank; ++d) {
            pp.g[d] = g[d];
            pp.s[d] = s[d];
        }
        nd_strides(s, rank, pp.sstride);
        pp.reverse = 0;
        pp.rows = nd_pad_rows(p, s, g);
        const dim3 grid = nd_grid(pp.rows, g[rank - 1]);
        if (cplx)
            hipLaunchKernelGGL(nd_pad_kernel<NdLoadC64>, grid, dim3(256), 0, st,
                               NdLoadC64{static_cast<const float2*>(src)}, pp, X);
        else
            hipLaunchKernelGGL(nd_pad_kernel<NdLoadF32>, grid, dim3(256), 0, st,
                               NdLoadF32{static_cast<const float*>(src)}, pp, X);
        ZKX_LAUNCH_CHECK();
        return ZKX_SUCCESS;
    };
    // the crop reads rows [start, start + out) of dimension 0
    p.keep_lo0 = (q.full[0] - q.out[0]) / 2;
    p.keep_hi0 = p.keep_lo0 + q.out[0];
    float2* Yp = nullptr;
    bool same_g = true;
    for (int d = 0; d < rank; ++d) same_g = same_g && p.g1[d] == p.g2[d];
    if (!ca && !cb && same_g) {
        // two real operands: one packed transform
        if ((rc = nd_pack_real_pair(p, ws, NdLoadF32{static_cast<const float*>(a)}, sa,
                                    NdLoadF32{static_cast<const float*>(b)}, sb, 0, st)))
            return rc;
        if ((rc = nd_conv_run_real_pair(p, ws, &Yp, st))) return rc;
    } else {
        if ((rc = pad(a, ca, sa, p.g1, reinterpret_cast<float2*>(ws + p.off_x1)))) return rc;
        if ((rc = pad(b, cb, sb, p.g2, reinterpret_cast<float2*>(ws + p.off_x2)))) return rc;
        if ((rc = nd_conv_run(p, ws, &Yp, st))) return rc;
    }
    // centred crop of the full result, which sits at the origin of Y
    NdCrop c{};
    c.rank = rank;
    nd_strides(p.g, rank, c.ystride);
    for (int d = 0; d < rank; ++d) {
        c.o[d] = q.out[d];
        c.start[d] = (q.full[d] - q.out[d]) / 2;
    }
    const float2* Y = Yp;
    const dim3 grid = nd_grid(count_of(q.out, rank) / q.out[rank - 1], q.out[rank - 1]);
    if (complex1 || complex2)
        hipLaunchKernelGGL(nd_crop_kernel<NdStoreC64>, grid, dim3(256), 0, st, Y, c,
                           NdStoreC64{static_cast<float2*>(out)});
    else
        hipLaunchKernelGGL(nd_crop_kernel<NdStoreRealF32>, grid, dim3(256), 0, st, Y, c,
                           NdStoreRealF32{static_cast<float*>(out)});
    ZKX_LAUNCH_CHECK();
    return ZKX_SUCCESS;
}

}  // extern "C"
