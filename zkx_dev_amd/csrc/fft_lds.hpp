// LDS-resident Stockham FFT building blocks for gfx950 (wave64).
//
// A transform of N = 2^n points is owned by T = N/E threads; each thread keeps
// E complex values in registers.  The register convention on entry and on exit
// is the same natural, strided one:
//
//     thread t holds element  t + k*T  in v[k],   k = 0..E-1
//
// so callers fuse their own load (window, zero-pad, corner turn) and store
// (spectral multiply, |.|^2, scaling) around the transform.  Passes follow the
// self-sorting Stockham decimation-in-time form: pass s with radix R and stride
// Ns reads x[j + r*N/R], twiddles by W_{Ns*R}^{(j mod Ns)*r}, does a radix-R DFT
// in registers and writes x'[(j/Ns)*Ns*R + j mod Ns + r*Ns].  The first pass
// reads the caller's registers and the last pass leaves its results in
// registers, so only the (log_E N - 1) middle exchanges touch LDS.
//
// The LDS layout is a policy object, so a row transform (one signal per slot,
// padded to break bank conflicts) and a column transform (TC signals
// interleaved so that a wavefront's lanes hit consecutive banks) share the same
// pass code.
#pragma once

#include "zkx_hip_common.hpp"

namespace zkx {
namespace fft {

// cos(2*pi*k/16), k = 0..3 (the rest follow by symmetry)
__host__ __device__ constexpr float c16(int k)
{
    return k == 0 ? 1.0f : k == 1 ? 0.923879532511286756f : k == 2 ? 0.707106781186547524f
                                                                  : 0.382683432365089772f;
}

// Radix-2 butterfly with a 16th root of unity: (a, b) <- (a + W b, a - W b),
// W = W_16^idx (conjugated for the inverse).  A general root is applied in
// the fused multiply-add form W b = k * u, k the root's larger component and
// u = b rotated by the ratio of the smaller to the larger (tangent form), so
// the butterfly costs 2 + 4 fused operations instead of a 4-operation
// complex product and 4 additions.  idx is a compile-time constant after
// unrolling, so every branch folds.
template <bool INV>
__device__ __forceinline__ void bfly_w16(float2& a, float2& b, int idx)
{
    constexpr float kTan8 = 0.414213562373095048802f;  // tan(pi/8)
    idx &= 15;
    if (idx == 0 || idx == 8) {
        const float2 p = cadd(a, b), m = csub(a, b);
        a = idx == 0 ? p : m;
        b = idx == 0 ? m : p;
        return;
    }
    if (idx == 4 || idx == 12) {
        // W b = -i b (forward idx 4, inverse idx 12) or +i b
        const bool minus_i = (idx == 4) != INV;
        const float2 wb = minus_i ? make_float2(b.y, -b.x) : make_float2(-b.y, b.x);
        const float2 p = cadd(a, wb), m = csub(a, wb);
        a = p;
        b = m;
        return;
    }
    const int q = idx >> 2, m4 = idx & 3;
    float c, s;  // cos, sin of 2*pi*idx/16
    if (q == 0) { c = c16(m4); s = c16(4 - m4); }
    else if (q == 1) { c = -c16(4 - m4); s = c16(m4); }
    else if (q == 2) { c = -c16(m4); s = -c16(4 - m4); }
    else { c = c16(4 - m4); s = -c16(m4); }
    // W b = (c bx + g s by, c by - g s bx), g = +1 forward, -1 inverse
    const bool g_pos = !INV;
    const bool cos_major = m4 == 2 || (m4 == 1) == (q == 0 || q == 2);
    float2 u;
    float k;
    if (cos_major) {
        // W b = c (bx + r by, by - r bx), r = g s / c: +-1 or +-tan(pi/8)
        const bool r_pos = ((s > 0.f) == (c > 0.f)) == g_pos;
        if (m4 == 2)
            u = r_pos ? make_float2(b.x + b.y, b.y - b.x) : make_float2(b.x - b.y, b.y + b.x);
        else
            u = r_pos ? make_float2(fmaf(kTan8, b.y, b.x), fmaf(-kTan8, b.x, b.y))
                      : make_float2(fmaf(-kTan8, b.y, b.x), fmaf(kTan8, b.x, b.y));
        k = c;
    } else {
        // W b = s (r bx + g by, r by - g bx), r = c / s = +-tan(pi/8)
        const float r = ((c > 0.f) == (s > 0.f)) ? kTan8 : -kTan8;
        u = g_pos ? make_float2(fmaf(r, b.x, b.y), fmaf(r, b.y, -b.x)) : make_float2(fmaf(r, b.x, -b.y), fmaf(r, b.y, b.x));
        k = s;
    }
    const float2 na = make_float2(fmaf(k, u.x, a.x), fmaf(k, u.y, a.y));
    const float2 nb = make_float2(fmaf(-k, u.x, a.x), fmaf(-k, u.y, a.y));
    a = na;
    b = nb;
}

// In-register DFT of R <= 16 points, natural order in and out (radix-2 DIT
// recursion; every twiddle is a 16th root of unity, folded at compile time).
template <int R, bool INV>
struct Dft {
    static __device__ __forceinline__ void run(float2* v)
    {
        constexpr int M = R / 2;
        float2 e[M], o[M];
#pragma unroll
        for (int i = 0; i < M; ++i) {
            e[i] = v[2 * i];
            o[i] = v[2 * i + 1];
        }
        Dft<M, INV>::run(e);
        Dft<M, INV>::run(o);
#pragma unroll
        for (int k = 0; k < M; ++k) {
            bfly_w16<INV>(e[k], o[k], k * (16 / R));
            v[k] = e[k];
            v[k + M] = o[k];
        }
    }
};
template <bool INV>
struct Dft<1, INV> {
    static __device__ __forceinline__ void run(float2*) {}
};

// One signal per slot as interleaved complex values (one 8-byte access per
// element), one padding element per 16: a half-wave of 16 lanes touching 16
// consecutive elements, or elements 17 apart, covers every bank once.  A slot
// takes 2 * floats_for(n) floats.
struct RowLds {
    float2* p;
    __device__ __forceinline__ static int idx(int i) { return i + (i >> 4); }
    __device__ __forceinline__ float2 load(int i) const { return p[idx(i)]; }
    __device__ __forceinline__ void store(int i, float2 v) const { p[idx(i)] = v; }
    // element base + c, where c is a multiple of 16, or base is a multiple of
    // 16 and c < 16 (every access of the passes below): the padding splits as
    // idx(base) + c + c/16, so c folds into the instruction's offset
    __device__ __forceinline__ float2 load(int base, int c) const { return p[idx(base) + c + (c >> 4)]; }
    __device__ __forceinline__ void store(int base, int c, float2 v) const { p[idx(base) + c + (c >> 4)] = v; }
    static constexpr int floats_for(int n) { return n + n / 16; }
};

// TC signals interleaved: element i of slot s lives at i*TC + s.
template <int TC>
struct ColLds {
    float* re;
    float* im;
    int slot;
    __device__ __forceinline__ float2 load(int i) const
    {
        const int a = i * TC + slot;
        return make_float2(re[a], im[a]);
    }
    __device__ __forceinline__ void store(int i, float2 v) const
    {
        const int a = i * TC + slot;
        re[a] = v.x;
        im[a] = v.y;
    }
    __device__ __forceinline__ float2 load(int base, int c) const { return load(base + c); }
    __device__ __forceinline__ void store(int base, int c, float2 v) const { store(base + c, v); }
};

// Twiddle sources.  A pass of radix R at stride Ns needs w^r, r < R, with
// w = exp(-2 pi i jm / (Ns*R)) = tw[jm * N/(Ns*R)].
// From the global table: one read per butterfly, the other powers by a
// product tree of depth <= 4 (w^r = w^p * w^(r-p), p the largest power of two
// <= r).
template <int N, int Ns, int R>
__device__ __forceinline__ void twiddle_powers(const float2* __restrict__ tw, int jm, float2 (&wp)[R])
{
    wp[1] = tw[jm * (N / (Ns * R))];
#pragma unroll
    for (int r = 2; r < R; ++r) {
        const int p = (r & 8) ? 8 : (r & 4) ? 4 : (r & 2) ? 2 : 1;
        wp[r] = (p == r) ? cmul(wp[r / 2], wp[r / 2]) : cmul(wp[p], wp[r - p]);
    }
}
// From a copy of the N-entry table in LDS: every power read directly
// (jm * r < Ns * R <= N, so no index wraps).
struct LdsTwiddles {
    const float2* t;
};
template <int N, int Ns, int R>
__device__ __forceinline__ void twiddle_powers(LdsTwiddles tw, int jm, float2 (&wp)[R])
{
#pragma unroll
    for (int r = 1; r < R; ++r) wp[r] = tw.t[jm * r * (N / (Ns * R))];
}

// The powers of the stride-16 pass (the second pass of every transform with
// E = 16) from a small LDS table, the other passes by the product tree from
// the global N-point table.  The stride-16 pass needs w^r = exp(-2 pi i r jm
// / (16 R)) for jm < 16 only: (R-1) x 16 entries, entry (r-1)*16 + jm, which
// a wave reads as 16 broadcast addresses (no bank conflicts).  Callers fill
// the table with fill_pass16_twiddles before the transform's first barrier.
struct Pass16Twiddles {
    const float2* g;  // exp(-2 pi i k / N), k < N (global)
    const float2* l;  // the stride-16 table (LDS)
};
template <int N>
struct Pass16Table {
    static constexpr int R = N / 16 >= 16 ? 16 : N / 16;  // radix of the stride-16 pass
    static constexpr int ENTRIES = N >= 32 ? (R - 1) * 16 : 0;
};
template <int N>
__device__ __forceinline__ void fill_pass16_twiddles(float2* l, const float2* __restrict__ g, int tid, int nthreads)
{
    using P = Pass16Table<N>;
    for (int e = tid; e < P::ENTRIES; e += nthreads) {
        const int r = e / 16 + 1, jm = e % 16;
        l[e] = g[r * jm * (N / (16 * P::R))];
    }
}
template <int N, int Ns, int R>
__device__ __forceinline__ void twiddle_powers(Pass16Twiddles tw, int jm, float2 (&wp)[R])
{
    if constexpr (Ns == 16) {
        static_assert(R == Pass16Table<N>::R, "stride-16 pass radix");
#pragma unroll
        for (int r = 1; r < R; ++r) wp[r] = tw.l[(r - 1) * 16 + jm];
    } else {
        twiddle_powers<N, Ns, R>(tw.g, jm, wp);
    }
}

// Stockham passes from stride Ns on.  tw[k] = exp(-2 pi i k / N), k < N.
// Every thread of the block must call this (it contains barriers).
template <int N, int E, int Ns, bool INV, class Lds, class TW>
__device__ __forceinline__ void passes(float2 (&v)[E], const Lds& lds, TW tw, int t)
{
    constexpr int REM = N / Ns;
    constexpr int R = REM >= E ? E : REM;
    constexpr int NB = E / R;
    constexpr int T = N / E;
    constexpr bool FIRST = Ns == 1;
    constexpr bool LAST = Ns * R == N;
    static_assert(R >= 2 && E % R == 0, "radix schedule");
    if constexpr (!FIRST) {
#pragma unroll
        for (int b = 0; b < NB; ++b) {
            const int j = t + b * T;
#pragma unroll
            for (int r = 0; r < R; ++r) v[b + r * NB] = lds.load(j, r * (N / R));
        }
#pragma unroll
        for (int b = 0; b < NB; ++b) {
            const int jm = (t + b * T) & (Ns - 1);
            float2 wp[R];
            twiddle_powers<N, Ns, R>(tw, jm, wp);
#pragma unroll
            for (int r = 1; r < R; ++r) v[b + r * NB] = INV ? cmulc(v[b + r * NB], wp[r]) : cmul(v[b + r * NB], wp[r]);
        }
    }
#pragma unroll
    for (int b = 0; b < NB; ++b) {
        float2 u[R];
#pragma unroll
        for (int r = 0; r < R; ++r) u[r] = v[b + r * NB];
        Dft<R, INV>::run(u);
#pragma unroll
        for (int r = 0; r < R; ++r) v[b + r * NB] = u[r];
    }
    if constexpr (!LAST) {
        __syncthreads();  // earlier readers of this LDS region are done
#pragma unroll
        for (int b = 0; b < NB; ++b) {
            const int j = t + b * T;
            const int base = (j / Ns) * Ns * R + (j & (Ns - 1));
#pragma unroll
            for (int r = 0; r < R; ++r) lds.store(base, r * Ns, v[b + r * NB]);
        }
        __syncthreads();
        passes<N, E, Ns * R, INV>(v, lds, tw, t);
    }
}

// Full transform: v holds x[t + k*T] on entry and X[t + k*T] on exit.
// The inverse is unscaled here; callers fold 1/N into their store.
template <int N, int E, bool INV, class Lds, class TW>
__device__ __forceinline__ void transform(float2 (&v)[E], const Lds& lds, TW tw, int t)
{
    passes<N, E, 1, INV>(v, lds, tw, t);
}

// Elements per thread used for each size: E=16 from N=16 up, N itself below.
__host__ __device__ constexpr int elems_for(int n) { return n >= 16 ? 16 : n; }

// Launch shape of the row kernels: FPB signals of N points per block, each
// owned by T = N/E threads, one padded LDS slot per signal.
template <int N>
struct RowCfg {
    static constexpr int E = elems_for(N);
    static constexpr int T = N / E;
    static constexpr int FPB = T >= 256 ? 1 : 256 / T;
    static constexpr int THREADS = T * FPB;
    static constexpr int SLOT_FLOATS = RowLds::floats_for(N);
    static constexpr size_t LDS_BYTES = sizeof(float) * 2 * SLOT_FLOATS * FPB;
};

}  // namespace fft
}  // namespace zkx
