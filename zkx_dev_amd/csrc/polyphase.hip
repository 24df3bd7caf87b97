// Polyphase FIR on gfx950: upfirdn, resample_poly and firfilter for float32
// and complex64 lines with real taps.
//
//   y[n] = sum_k h[k] * x_up[m(n) - k],  m(n) = n*down + m_off,
//   x_up[u] = x[u/up] when up | u, else 0.
//
// upfirdn has m_off = 0 (ZKX/cusignal_cpp/src/operators/filtering/filtering_typed.cpp:2548-2616);
// resample_poly has m_off = start*down - prefix_zeros with gcd-reduced up/down
// (filtering_typed.cpp:933-1044, 2617-2715); firfilter is up = down = 1
// (filtering_typed.cpp:1536-1592).
//
// Output n has phase phi = m mod up and base q = m div up, and
//   y[n] = sum_j h[phi + j*up] * x[q - j].
// With Pn = up/gcd(up,down) and D = down/gcd(up,down), the outputs
// n = T*Pn + r (fixed residue r) share one phase and their bases advance by D:
// each residue is a plain decimate-by-D FIR.  A block covers TT consecutive T
// for all residues (TT*Pn consecutive outputs).  A wavefront takes one
// residue at a time, so its taps are uniform, and each lane register-blocks M
// consecutive T: one LDS read of x feeds M FMAs, with the taps of the residue
// laid out in LDS as a zero-padded [shift][M] table read as broadcast
// vectors.  x for the tile and the outputs are staged through LDS so that HBM
// reads and writes stay coalesced.
#include "zkx_hip_common.hpp"
#include "runtime.hpp"

#include <algorithm>
#include <cmath>
#include <limits>
#include <numeric>

namespace zkx {

template <class V>
int fir_fft_run(const float* h, int K, const V* x, int lines, long long n, int D, long long out_n, V* y,
                hipStream_t st, bool* done);
template <class V>
int fir_poly_run(const float* h, int K, int up, int down, long long m_off, const V* x, int lines, long long n,
                 long long out_n, V* y, hipStream_t st, bool* done);

// Long up = 1 filters on contiguous lines take a frequency-domain path
// (fir_fft.hip): 2 the 4096-point segment form where it applies (D in
// {1, 2, 4, 8}), else the phase form; 1 always the phase form; 0 always the
// direct polyphase kernels.
int g_fir_fft = 2;

namespace {

template <class V>
struct VOps;
template <>
struct VOps<float> {
    static __device__ __forceinline__ float zero() { return 0.0f; }
    static __device__ __forceinline__ float fma(float h, float x, float acc) { return __builtin_fmaf(h, x, acc); }
};
template <>
struct VOps<float2> {
    static __device__ __forceinline__ float2 zero() { return make_float2(0.f, 0.f); }
    static __device__ __forceinline__ float2 fma(float h, float2 x, float2 acc)
    {
        return make_float2(__builtin_fmaf(h, x.x, acc.x), __builtin_fmaf(h, x.y, acc.y));
    }
};

struct PolyPlan {
    int up, down;       // as applied (upfirdn: as given; resample: gcd-reduced)
    int Pn, D;          // residue period and per-residue input stride
    long long m_off;    // m(n) = n*down + m_off, m_off >= 0
    int K, Jmax;        // taps, taps per phase (max)
    long long n_in, n_out;
};

constexpr int kWaves = 4;

// Taps table tab[r][s][i] = h[phi_r + j*up], j = i*D - s + Jmax - 1, zero
// outside the phase, built once per call in global memory.
__global__ void taps_table_kernel(const float* __restrict__ h, PolyPlan p, int M, float* __restrict__ tab)
{
    const int SL = p.Jmax + (M - 1) * p.D;
    const int count = p.Pn * SL * M;
    for (int e = blockIdx.x * blockDim.x + threadIdx.x; e < count; e += gridDim.x * blockDim.x) {
        const int r = e / (SL * M), rem = e % (SL * M), s = rem / M, i = rem % M;
        const long long mr = static_cast<long long>(r) * p.down + p.m_off;
        const int phi = static_cast<int>(mr % p.up);
        const int j = i * p.D - s + p.Jmax - 1;
        const long long k = static_cast<long long>(phi) + static_cast<long long>(j) * p.up;
        tab[e] = (j >= 0 && k < p.K) ? h[k] : 0.0f;
    }
}

// Shared layout: the x tile [XL] only.  The taps of a wavefront's residue are
// uniform across its lanes and come through scalar loads, so the only LDS
// read per shift is the one x sample that feeds M FMAs.
//
// acc[i] = output t0 + i of residue r, t0 = tc*64*M + lane*M; the lane's tile
// elements are t0*D + dq + s = (tc*64 + lane)*S + (dq + s), s < SL.
template <class V, int M>
__device__ __forceinline__ void residue_outputs(const V* xs, int Lc, int S, int SL, int dq, const float* tr, int tc,
                                                int lane, V (&acc)[M])
{
    const V* xb = xs + tc * 64 + lane;
#pragma unroll
    for (int i = 0; i < M; ++i) acc[i] = VOps<V>::zero();
    // element dq + s sits at (k mod S)*Lc + k/S, k = dq + s.  Walking s down
    // the offset drops by Lc, or rises by (S-1)*Lc - 1 when k mod S wraps.
    // S = M*D is a multiple of 4, so once k mod S + 1 is one too, four shifts
    // never wrap inside a step: their offsets are off - u*Lc, and the wrap is
    // one scalar test per step.
    static_assert(M % 4 == 0, "steps of four shifts need S = M*D divisible by 4");
    const int k = dq + SL - 1;
    int km = __builtin_amdgcn_readfirstlane(k % S);
    int off = __builtin_amdgcn_readfirstlane(km * Lc + k / S);
    const int wrap = (S - 1) * Lc - 1;
    int s = SL - 1;
    auto single = [&]() {
        const V xv = xb[off];
        const float* trs = tr + s * M;
#pragma unroll
        for (int i = 0; i < M; ++i) acc[i] = VOps<V>::fma(trs[i], xv, acc[i]);
        if (km == 0) {
            km = S - 1;
            off += wrap;
        } else {
            --km;
            off -= Lc;
        }
    };
    for (; s >= 0 && ((km + 1) & 3) != 0; --s) single();
    // four shifts per step: their four x reads and 4*M taps (one scalar
    // load) are all in flight before the first FMA, in the same order of
    // accumulation as one shift at a time
    for (; s >= 3; s -= 4) {
        V xv[4];
#pragma unroll
        for (int u = 0; u < 4; ++u) xv[u] = xb[off - u * Lc];
        const float* trs = tr + (s - 3) * M;
#pragma unroll
        for (int u = 0; u < 4; ++u)
#pragma unroll
            for (int i = 0; i < M; ++i) acc[i] = VOps<V>::fma(trs[(3 - u) * M + i], xv[u], acc[i]);
        off -= 4 * Lc;
        km -= 4;
        if (km < 0) {
            km += S;
            off += S * Lc - 1;
        }
    }
    for (; s >= 0; --s) single();
}

// PN > 0: the residue count Pn = PN is compiled in and each wave computes all
// PN residues of its M outputs per lane, so a lane holds M*PN consecutive
// outputs of y and stores them as one contiguous run (the wave writes whole
// cache lines).  PN == 0: any Pn, one residue per task, each output stored
// where it lands (every Pn-th element).
template <class V, int M, int PN>
__global__ void __launch_bounds__(64 * kWaves)
polyphase_kernel(const float* __restrict__ tab, const V* __restrict__ x, V* __restrict__ y, PolyPlan p, int TT, int XL)
{
    extern __shared__ __align__(16) unsigned char smraw[];
    const int SL = p.Jmax + (M - 1) * p.D;
    // Lane l reads tile element l*S + k for a wave-uniform k, S = M*D.  The
    // tile is stored transposed, element e at (e mod S)*Lc + e/S, so those 64
    // reads hit 64 consecutive words instead of one bank.
    const int S = M * p.D;
    const int Lc = (XL + S - 1) / S;
    V* xs = reinterpret_cast<V*>(smraw);
    const int tid = threadIdx.x;
    const long long line = blockIdx.y;
    const V* xl = x + line * p.n_in;
    V* yl = y + line * p.n_out;
    const long long T0 = static_cast<long long>(blockIdx.x) * TT;
    // m of residue r at T: T*up*D + m_r, m_r = r*down + m_off
    const long long q0 = p.m_off / p.up;  // base of residue 0 at T = 0
    const long long xstart = T0 * p.D + q0 - (p.Jmax - 1);
    constexpr int LB = 8;  // tile samples in flight per thread
    for (int base = tid; base < S * Lc; base += LB * static_cast<int>(blockDim.x)) {
        V tmp[LB];
#pragma unroll
        for (int u = 0; u < LB; ++u) {
            const int e = base + u * static_cast<int>(blockDim.x);
            const long long g = xstart + e;
            tmp[u] = (e < XL && g >= 0 && g < p.n_in) ? xl[g] : VOps<V>::zero();
        }
        if ((S & (S - 1)) == 0) {  // S a power of two: shift and mask
            const int sh = __builtin_ctz(S);
#pragma unroll
            for (int u = 0; u < LB; ++u) {
                const int e = base + u * static_cast<int>(blockDim.x);
                if (e < S * Lc) xs[(e & (S - 1)) * Lc + (e >> sh)] = tmp[u];
            }
        } else {
#pragma unroll
            for (int u = 0; u < LB; ++u) {
                const int e = base + u * static_cast<int>(blockDim.x);
                if (e < S * Lc) xs[(e % S) * Lc + e / S] = tmp[u];
            }
        }
    }
    __syncthreads();
    const int wave = __builtin_amdgcn_readfirstlane(tid / 64), lane = tid % 64;
    const int NC = TT / (64 * M);
    if constexpr (PN > 0) {
        for (int tc = wave; tc < NC; tc += kWaves) {
            V acc[PN][M];
#pragma unroll
            for (int r = 0; r < PN; ++r) {
                const long long mr = static_cast<long long>(r) * p.down + p.m_off;
                const int dq = static_cast<int>(mr / p.up - q0);
                residue_outputs<V, M>(xs, Lc, S, SL, dq, tab + static_cast<long long>(r) * SL * M, tc, lane, acc[r]);
            }
            const long long n0 = (T0 + tc * 64 * M + lane * M) * PN;
            V* yo = yl + n0;
            if (n0 + M * PN <= p.n_out) {
#pragma unroll
                for (int e = 0; e < M * PN; ++e) yo[e] = acc[e % PN][e / PN];
            } else {
#pragma unroll
                for (int e = 0; e < M * PN; ++e)
                    if (n0 + e < p.n_out) yo[e] = acc[e % PN][e / PN];
            }
        }
    } else {
        for (int task = wave; task < p.Pn * NC; task += kWaves) {
            const int r = task / NC, tc = task % NC;
            const long long mr = static_cast<long long>(r) * p.down + p.m_off;
            const int dq = static_cast<int>(mr / p.up - q0);
            V acc[M];
            residue_outputs<V, M>(xs, Lc, S, SL, dq, tab + static_cast<long long>(r) * SL * M, tc, lane, acc);
            const long long n0 = (T0 + tc * 64 * M + lane * M) * p.Pn + r;
#pragma unroll
            for (int i = 0; i < M; ++i) {
                const long long n = n0 + static_cast<long long>(i) * p.Pn;
                if (n < p.n_out) yl[n] = acc[i];
            }
        }
    }
}

// Any layout [outer, n, inner]: one thread per output, taps of its phase in order.
template <class V>
__global__ void polyphase_generic_kernel(const float* __restrict__ h, const V* __restrict__ x, V* __restrict__ y,
                                         PolyPlan p, int outer, int inner)
{
    const long long total = static_cast<long long>(outer) * inner * p.n_out;
    for (long long g = blockIdx.x * static_cast<long long>(blockDim.x) + threadIdx.x; g < total;
         g += static_cast<long long>(gridDim.x) * blockDim.x) {
        const long long in = g % inner;
        const long long n = (g / inner) % p.n_out;
        const long long o = g / (static_cast<long long>(inner) * p.n_out);
        const long long m = n * p.down + p.m_off;
        const int phi = static_cast<int>(m % p.up);
        const long long q = m / p.up;
        const V* xl = x + o * p.n_in * inner + in;
        V acc = VOps<V>::zero();
        for (long long k = phi, j = 0; k < p.K; k += p.up, ++j) {
            const long long u = q - j;
            if (u < 0) break;
            if (u < p.n_in) acc = VOps<V>::fma(h[k], xl[u * inner], acc);
        }
        y[o * p.n_out * inner + n * inner + in] = acc;
    }
}

// firfilter state: y[n] += zi[n] (n < K-1) and zf[k] = full[N+k] (+ zi[N+k]).
template <class V>
__global__ void firfilter_state_kernel(const V* __restrict__ full, int outer, long long n, int inner, int state,
                                       const V* __restrict__ zi, V* __restrict__ yout, V* __restrict__ zf)
{
    const long long total = static_cast<long long>(outer) * inner * (n + state);
    for (long long g = blockIdx.x * static_cast<long long>(blockDim.x) + threadIdx.x; g < total;
         g += static_cast<long long>(gridDim.x) * blockDim.x) {
        const long long in = g % inner;
        const long long f = (g / inner) % (n + state);
        const long long o = g / (static_cast<long long>(inner) * (n + state));
        V v = full[o * (n + state) * inner + f * inner + in];
        if (zi && f < state) {
            const V z = zi[o * state * inner + f * inner + in];
            v = VOps<V>::fma(1.0f, z, v);
        }
        if (f < n) {
            yout[o * n * inner + f * inner + in] = v;
        } else if (zf) {
            zf[o * state * inner + (f - n) * inner + in] = v;
        }
    }
}

unsigned grid_for(long long n)
{
    const long long b = (n + 255) / 256;
    return static_cast<unsigned>(std::max<long long>(1, std::min<long long>(b, 65535)));
}

thread_local ThreadWorkspace t_taps_ws;

template <class V, int M, int PN>
int launch_tiled(const float* h, const V* x, V* y, const PolyPlan& p, long long lines, hipStream_t st, bool* done)
{
    *done = false;
    const int NC = PN > 0 ? kWaves : kWaves / std::gcd(p.Pn, kWaves);
    const int TT = NC * 64 * M;
    const long long SL = p.Jmax + static_cast<long long>(M - 1) * p.D;
    const long long q_last = ((static_cast<long long>(p.Pn) - 1) * p.down + p.m_off) / p.up - p.m_off / p.up;
    const long long XL = static_cast<long long>(TT - 1) * p.D + q_last + p.Jmax + (M - 1) * p.D + 1;
    const long long S = static_cast<long long>(M) * p.D;
    const long long XLp = (XL + S - 1) / S * S;  // transposed tile rounds up to whole columns
    const size_t lds = sizeof(V) * static_cast<size_t>(XLp);
    const long long tab_count = static_cast<long long>(p.Pn) * SL * M;
    // very long phase tables or tiles go to the generic kernel
    if (lds > 150 * 1024 || XL > std::numeric_limits<int>::max() || tab_count > (1 << 20)) return ZKX_SUCCESS;
    const long long Tcount = (p.n_out + p.Pn - 1) / p.Pn;
    const long long blocks = (Tcount + TT - 1) / TT;
    if (blocks > std::numeric_limits<int>::max() || lines > 65535) return ZKX_SUCCESS;
    void* wsp = nullptr;
    WorkspaceLease lease_(t_taps_ws, st);
    int rc = lease_.get(sizeof(float) * static_cast<size_t>(tab_count), &wsp);
    if (rc) return rc;
    float* tab = static_cast<float*>(wsp);
    hipLaunchKernelGGL(taps_table_kernel, dim3(static_cast<unsigned>(std::min<long long>((tab_count + 255) / 256, 1024))),
                       dim3(256), 0, st, h, p, M, tab);
    ZKX_LAUNCH_CHECK();
    if (lds > 64 * 1024)
        ZKX_HIP_CHECK(hipFuncSetAttribute(reinterpret_cast<const void*>(polyphase_kernel<V, M, PN>),
                                          hipFuncAttributeMaxDynamicSharedMemorySize, static_cast<int>(lds)));
    hipLaunchKernelGGL((polyphase_kernel<V, M, PN>), dim3(static_cast<unsigned>(blocks), static_cast<unsigned>(lines)),
                       dim3(64 * kWaves), lds, st, tab, x, y, p, TT, static_cast<int>(XL));
    ZKX_LAUNCH_CHECK();
    *done = true;
    return ZKX_SUCCESS;
}

// LDS bytes of one tiled block for a register-blocking depth M.
template <class V>
size_t tiled_lds_bytes(const PolyPlan& p, int M)
{
    const int NC = kWaves / std::gcd(p.Pn, kWaves);
    const long long TT = static_cast<long long>(NC) * 64 * M;
    const long long SL = p.Jmax + static_cast<long long>(M - 1) * p.D;
    const long long q_last = ((static_cast<long long>(p.Pn) - 1) * p.down + p.m_off) / p.up - p.m_off / p.up;
    const long long XL = (TT - 1) * p.D + q_last + p.Jmax + (M - 1) * p.D + 1;
    (void)SL;
    const long long S = static_cast<long long>(M) * p.D;
    return sizeof(V) * static_cast<size_t>((XL + S - 1) / S * S);
}

// Register blocking depth: M = 8 gives more FMAs per LDS read of x, M = 4
// less zero-tap padding ((M-1)*D shifts per output group); take 8 only when
// the taps per phase dwarf the padding (Jmax >= 64 D; at 121 taps, 3/2, M = 8
// measured 12% slower) and the tile stays small.
template <class V>
int pick_m(const PolyPlan& p)
{
    return (p.Jmax >= 64 * p.D && tiled_lds_bytes<V>(p, 8) <= 52 * 1024) ? 8 : 4;
}

template <class V>
int run_polyphase(const float* h, const V* x, V* y, const PolyPlan& p, int outer, int inner, hipStream_t st)
{
    if (p.n_out <= 0 || outer <= 0 || inner <= 0) return ZKX_SUCCESS;
    if (inner == 1 && p.up > 1) {  // long rate-changing filters: the segment form (fir_fft.hip)
        bool done = false;
        const int rc = fir_poly_run<V>(h, p.K, p.up, p.down, p.m_off, x, outer, p.n_in, p.n_out, y, st, &done);
        if (rc || done) return rc;
    }
    if (inner == 1) {
        bool done = false;
        int rc;
        if (pick_m<V>(p) == 8) {
            rc = launch_tiled<V, 8, 0>(h, x, y, p, outer, st, &done);
        } else {
            switch (p.Pn) {  // residue counts compiled in for whole-line stores
            case 1: rc = launch_tiled<V, 4, 1>(h, x, y, p, outer, st, &done); break;
            case 2: rc = launch_tiled<V, 4, 2>(h, x, y, p, outer, st, &done); break;
            case 3: rc = launch_tiled<V, 4, 3>(h, x, y, p, outer, st, &done); break;
            case 4: rc = launch_tiled<V, 4, 4>(h, x, y, p, outer, st, &done); break;
            default: rc = launch_tiled<V, 4, 0>(h, x, y, p, outer, st, &done); break;
            }
        }
        if (rc) return rc;
        if (done) return ZKX_SUCCESS;
    }
    const long long total = static_cast<long long>(outer) * inner * p.n_out;
    hipLaunchKernelGGL((polyphase_generic_kernel<V>), dim3(grid_for(total)), dim3(256), 0, st, h, x, y, p, outer,
                       inner);
    ZKX_LAUNCH_CHECK();
    return ZKX_SUCCESS;
}

PolyPlan make_plan(int up, int down, long long m_off, int K, long long n_in, long long n_out)
{
    PolyPlan p;
    p.up = up;
    p.down = down;
    const int g = std::gcd(up, down);
    p.Pn = up / g;
    p.D = down / g;
    p.m_off = m_off;
    p.K = K;
    p.Jmax = (K + up - 1) / up;
    p.n_in = n_in;
    p.n_out = n_out;
    return p;
}

int check_lines(int outer, int n, int inner)
{
    if (outer < 0 || n < 0 || inner < 0) ZKX_FAIL(ZKX_INVALID_SHAPE, "input shape");
    const long long total = static_cast<long long>(outer) * n * inner;
    if (total > std::numeric_limits<int>::max()) ZKX_FAIL(ZKX_INVALID_ARGUMENT, "input indexing");
    return ZKX_SUCCESS;
}

long long upfirdn_len(long long n, long long K, int up, int down)
{
    const long long num = (n - 1) * up + K - 1;
    long long q = num / down;
    if (num % down < 0) --q;
    return q + 1;
}

template <class V>
int upfirdn_impl(const float* h, int taps, const V* x, int outer, int n, int inner, int up, int down, V* out,
                 hipStream_t st)
{
    if (taps < 1) ZKX_FAIL(ZKX_INVALID_SHAPE, "upfirdn coefficients");
    if (up < 1 || down < 1) ZKX_FAIL(ZKX_INVALID_ARGUMENT, "upfirdn factors");
    int rc = check_lines(outer, n, inner);
    if (rc) return rc;
    if (!h || (!x && n > 0) || !out) ZKX_FAIL(ZKX_INVALID_ARGUMENT, "upfirdn: null pointer");
    const long long out_n = upfirdn_len(n, taps, up, down);
    if (out_n < 0 || out_n > std::numeric_limits<int>::max() ||
        static_cast<long long>(outer) * inner * out_n > std::numeric_limits<int>::max())
        ZKX_FAIL(ZKX_INVALID_ARGUMENT, "upfirdn output sample count");
    if (g_fir_fft && inner == 1 && up == 1 && n > 0) {
        bool done = false;
        rc = fir_fft_run<V>(h, taps, x, outer, n, down, out_n, out, st, &done);
        if (rc || done) return rc;
    }
    return run_polyphase<V>(h, x, out, make_plan(up, down, 0, taps, n, out_n), outer, inner, st);
}

template <class V>
int resample_impl(const float* taps, int ntaps, const V* x, int outer, int n, int inner, int up, int down, V* out,
                  hipStream_t st)
{
    if (up < 1 || down < 1) ZKX_FAIL(ZKX_INVALID_ARGUMENT, "resample_poly factors");
    int rc = check_lines(outer, n, inner);
    if (rc) return rc;
    const int g = std::gcd(up, down);
    const int u = up / g, d = down / g;
    if (u == 1 && d == 1) {
        const size_t bytes = sizeof(V) * static_cast<size_t>(outer) * n * inner;
        if (bytes) ZKX_HIP_CHECK(hipMemcpyAsync(out, x, bytes, hipMemcpyDeviceToDevice, st));
        return ZKX_SUCCESS;
    }
    if (ntaps < 1 || !taps) ZKX_FAIL(ZKX_INVALID_SHAPE, "resample_poly coefficients");
    if (!out || (!x && n > 0)) ZKX_FAIL(ZKX_INVALID_ARGUMENT, "resample_poly: null pointer");
    // prepare_resample_poly_layout (filtering_typed.cpp:1003-1040)
    const long long out_n = (static_cast<long long>(n) * u + d - 1) / d;
    if (out_n > std::numeric_limits<int>::max() ||
        static_cast<long long>(outer) * inner * out_n > std::numeric_limits<int>::max())
        ZKX_FAIL(ZKX_INVALID_ARGUMENT, "resample_poly output sample count");
    const int half = (ntaps - 1) / 2;
    const int rem = ((half % d) + d) % d;
    const int prefix = d - rem;
    const long long start = (static_cast<long long>(half) + prefix) / d;
    const long long m_off = start * d - prefix;  // = half, since d | half + prefix
    return run_polyphase<V>(taps, x, out, make_plan(u, d, m_off, ntaps, n, out_n), outer, inner, st);
}

// decimate (filtering_typed.cpp:2229-2317, 2319-2356): the FIR of taps h
// sampled every q outputs, y[o] = sum_t h[t] x[(o + start) q - prefix - t];
// zero phase pads prefix = q - half % q zeros in front and starts at
// (half + prefix)/q, so m(o) = o*q + half; the causal form starts at m = 0.
// Output length ceil(n/q).  q = 1 still filters (no identity shortcut).
template <class V>
int decimate_impl(const float* taps, int ntaps, const V* x, int outer, int n, int inner, int q, bool zero_phase,
                  V* out, hipStream_t st)
{
    if (q < 1) ZKX_FAIL(ZKX_INVALID_ARGUMENT, "decimate q");
    if (ntaps < 1 || !taps) ZKX_FAIL(ZKX_INVALID_SHAPE, "decimate coefficients");
    int rc = check_lines(outer, n, inner);
    if (rc) return rc;
    const long long out_n = n == 0 ? 0 : 1 + (static_cast<long long>(n) - 1) / q;
    if (out_n == 0 || outer == 0 || inner == 0) return ZKX_SUCCESS;
    if (!out || !x) ZKX_FAIL(ZKX_INVALID_ARGUMENT, "decimate: null pointer");
    const int half = (ntaps - 1) / 2;
    const long long m_off = zero_phase ? half : 0;
    return run_polyphase<V>(taps, x, out, make_plan(1, q, m_off, ntaps, n, out_n), outer, inner, st);
}

thread_local ThreadWorkspace t_fir_ws;

int firfilter_impl(const float* b, int taps, const float* x, int outer, int n, int inner, const float* zi, float* y,
                   float* zf, hipStream_t st)
{
    if (taps < 1 || !b) ZKX_FAIL(ZKX_INVALID_SHAPE, "firfilter coefficients");
    if (outer < 1 || n < 1 || inner < 1) ZKX_FAIL(ZKX_INVALID_SHAPE, "firfilter input shape");
    int rc = check_lines(outer, n, inner);
    if (rc) return rc;
    if (!x || !y) ZKX_FAIL(ZKX_INVALID_ARGUMENT, "firfilter: null pointer");
    const int state = taps - 1;
    const bool fft_ok = g_fir_fft && inner == 1;
    if (!zi && !zf) {
        // y is the first n outputs of the full convolution
        if (fft_ok) {
            bool done = false;
            rc = fir_fft_run<float>(b, taps, x, outer, n, 1, n, y, st, &done);
            if (rc || done) return rc;
        }
        return run_polyphase<float>(b, x, y, make_plan(1, 1, 0, taps, n, n), outer, inner, st);
    }
    // full convolution into scratch, then zi / zf split
    const long long full_n = static_cast<long long>(n) + state;
    void* ws = nullptr;
    WorkspaceLease lease_(t_fir_ws, st);
    rc = lease_.get(sizeof(float) * static_cast<size_t>(outer) * inner * full_n, &ws);
    if (rc) return rc;
    float* full = static_cast<float*>(ws);
    bool done = false;
    if (fft_ok && (rc = fir_fft_run<float>(b, taps, x, outer, n, 1, full_n, full, st, &done))) return rc;
    if (!done && (rc = run_polyphase<float>(b, x, full, make_plan(1, 1, 0, taps, n, full_n), outer, inner, st)))
        return rc;
    const long long total = static_cast<long long>(outer) * inner * full_n;
    hipLaunchKernelGGL((firfilter_state_kernel<float>), dim3(grid_for(total)), dim3(256), 0, st, full, outer,
                       static_cast<long long>(n), inner, state, zi, y, zf);
    ZKX_LAUNCH_CHECK();
    return ZKX_SUCCESS;
}

}  // namespace
}  // namespace zkx

using namespace zkx;

extern "C" {

long long zkx_upfirdn_out_len(long long n, long long taps, int up, int down)
{
    if (taps < 1 || up < 1 || down < 1 || n < 0) return -1;
    return upfirdn_len(n, taps, up, down);
}

int zkx_upfirdn_f32(const float* h, int taps, const float* x, int outer, int n, int inner, int up, int down,
                    float* out, void* stream)
{
    clear_last_error();
    return upfirdn_impl<float>(h, taps, x, outer, n, inner, up, down, out, (hipStream_t)stream);
}

int zkx_upfirdn_c64(const float* h, int taps, const zkx_complex64* x, int outer, int n, int inner, int up, int down,
                    zkx_complex64* out, void* stream)
{
    clear_last_error();
    return upfirdn_impl<float2>(h, taps, reinterpret_cast<const float2*>(x), outer, n, inner, up, down,
                                reinterpret_cast<float2*>(out), (hipStream_t)stream);
}

int zkx_resample_poly_f32(const float* taps, int ntaps, const float* x, int outer, int n, int inner, int up, int down,
                          float* out, void* stream)
{
    clear_last_error();
    return resample_impl<float>(taps, ntaps, x, outer, n, inner, up, down, out, (hipStream_t)stream);
}

int zkx_resample_poly_c64(const float* taps, int ntaps, const zkx_complex64* x, int outer, int n, int inner, int up,
                          int down, zkx_complex64* out, void* stream)
{
    clear_last_error();
    return resample_impl<float2>(taps, ntaps, reinterpret_cast<const float2*>(x), outer, n, inner, up, down,
                                 reinterpret_cast<float2*>(out), (hipStream_t)stream);
}

int zkx_decimate_f32(const float* taps, int ntaps, const float* x, int outer, int n, int inner, int q, int zero_phase,
                     float* y, void* stream)
{
    clear_last_error();
    return decimate_impl<float>(taps, ntaps, x, outer, n, inner, q, zero_phase != 0, y,
                                static_cast<hipStream_t>(stream));
}

int zkx_firwin_lowpass_hamming(int numtaps, float cutoff, double* taps)
{
    clear_last_error();
    // firwin with fs = 2, one cutoff, pass_zero, scale, Hamming window
    // (filter_design_typed.cpp:24-100, 268-296): h[i] = c sinc(c m) w[i],
    // m = i - (n-1)/2, then divided by its sum (unit gain at DC)
    const double c = static_cast<double>(cutoff);
    if (numtaps < 1 || !taps) ZKX_FAIL(ZKX_INVALID_ARGUMENT, "firwin shape");
    if (!(c > 0.0 && c < 1.0)) ZKX_FAIL(ZKX_INVALID_ARGUMENT, "firwin cutoff");
    constexpr double pi = 3.141592653589793238462643383279502884;
    const double mid = 0.5 * (numtaps - 1);
    double sum = 0.0;
    for (int i = 0; i < numtaps; ++i) {
        const double w = numtaps == 1 ? 1.0 : 0.54 - 0.46 * std::cos(2.0 * pi * i / (numtaps - 1));
        const double arg = c * (i - mid);
        const double sinc = std::fabs(arg) < 1.0e-15 ? 1.0 : std::sin(pi * arg) / (pi * arg);
        taps[i] = c * sinc * w;
        sum += taps[i];
    }
    for (int i = 0; i < numtaps; ++i) taps[i] /= sum;
    return ZKX_SUCCESS;
}

int zkx_firfilter_f32(const float* b, int taps, const float* x, int outer, int n, int inner, const float* zi,
                      float* y, float* zf, void* stream)
{
    clear_last_error();
    return firfilter_impl(b, taps, x, outer, n, inner, zi, y, zf, (hipStream_t)stream);
}

}  // extern "C"
