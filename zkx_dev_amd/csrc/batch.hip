// Heterogeneous dwell batch engine: many dwells of different pulse counts,
// range lengths and templates in two persistent launches (pulse compression,
// then fused Doppler + |.|^2 + FP64 windowed CA-CFAR), replayed from a HIP
// graph, with FFT lengths and grid sizes chosen by a timing autotuner.
//
// The reference has no batched form: its Task1 chain runs one dwell per
// operator call (Task1/pipeline/src/steps/step{2,3,4}); this engine computes
// the same per-dwell results as pulse_compression_device (nfft = R, window and
// normalisation as configured), pulse_doppler_device (nfft = P) and the FP64
// windowed CA-CFAR of zkx_ca_cfar_window_f32.
//
// Work decomposition
//   * Pulse compression: a dwell with R a power of two (<= 8192) uses one
//     circular FFT of length R per pulse.  Otherwise each pulse is covered by
//     S overlap-save segments of a power-of-two length N > T over the periodic
//     pulse (tiles::pc_segment_row), which gives the length-R circular
//     correlation exactly.  A job is 8192 samples of FFT work: 8192/N row
//     segments of one dwell, so jobs cost about the same at every N.
//   * Doppler + CFAR with the engine's default window (guard 2x2, reference
//     8x8) for P = 32 .. 256: strips of range bins (strip::doppler_cfar_strip,
//     each column transformed once), 8 chunks per job; P <= 128 in 256-thread
//     blocks, P = 256 in 512-thread blocks.  Other windows and P = 16: tiles
//     of 64 range bins (tiles::doppler_cfar_tile, halo columns transformed by
//     both neighbours); a job packs 256/P tiles of one dwell into a
//     1024-thread block.
//   * Both grids are persistent and a multiple of 8; each of the 8 XCDs takes
//     a contiguous range of jobs, so CFAR tiles that share halo columns run on
//     one XCD and share its L2.
//   * Job tables and the run's pointers live in one device blob per lane,
//     copied from pinned host staging before each run, so the captured graph
//     (four kernel nodes: pulse compression, tiles, strips, P = 256 strips)
//     never changes; a node without jobs returns at once.
#include "dc_strip.hpp"
#include "radar_tiles.hpp"
#include "runtime.hpp"

#include <algorithm>
#include <cmath>
#include <cstdio>
#include <cstring>
#include <map>
#include <mutex>
#include <new>
#include <string>
#include <utility>
#include <vector>

namespace zkx {
int pc_template_dispatch(int n, const float2* h, int taps, int normalize, int window, float2* H, hipStream_t st);

namespace {

constexpr int kPcThreads = 512;
constexpr int kDcThreads = 1024;
constexpr int kMinLog = 4, kMaxLog = 13;  // FFT lengths 16 .. 8192
constexpr int kMinDopplerLog = 4, kMaxDopplerLog = 8;  // P = 16 .. 256
constexpr int kSlots = 4;
// strip Doppler+CFAR jobs: P = 32..128 in 256-thread blocks, P = 256 in 512
constexpr int kDcsThreads = 256, kDcs256Threads = 512;
constexpr int kStripChunksPerJob = 8;
// every pulse-compression job holds 8192 samples: 8192/N rows of 2*(N + N/16)
// floats, then the stride-16 pass twiddles of radix 2, 4, 8 and 16 (the
// tables of every N: entry (r-1)*16 + jm = exp(-2 pi i r jm / (16 R)))
constexpr size_t kPcRowBytes = sizeof(float) * 2 * (8192 + 8192 / 16);
constexpr int kP16Entries = 16 * (1 + 3 + 7 + 15);
constexpr size_t kPcLdsBytes = kPcRowBytes + sizeof(float2) * kP16Entries;
__host__ __device__ constexpr int p16_offset(int radix)
{
    return radix <= 2 ? 0 : radix == 4 ? 16 : radix == 8 ? 64 : 176;
}

struct DwellDev {
    long long x_off;  // complex64 elements: input dwell [P][R]
    long long o_off;  // cells: outputs [P][R]
    long long w_off;  // complex64 elements: pulse-compressed dwell in the workspace
    const float2* H;  // template spectrum of length N
    int P, R, N, L, S, nlog;
};

struct Job {
    int dwell, nlog, first, count;
};

struct RunArgs {
    const float2* x;
    float* thr;
    unsigned char* det;
    float2* rd;
    float2* ws;
    const DwellDev* dwells;
    const Job* pc_jobs;
    const Job* dc_jobs;
    const Job* dcs_jobs;     // strips, P = 32..128 (first = first output bin, count = bins)
    const Job* dcs256_jobs;  // strips, P = 256
    int n_pc, n_dc, n_dcs, n_dcs256;
    int window, g0, g1, r0, r1;
    double alpha_over_n;
    const float2* tw[kMaxLog + 1];
};

// The 8 XCDs receive workgroups round-robin (block b runs on XCD b % 8);
// give XCD x the job range [x*J/8, (x+1)*J/8) and stride its blocks over it.
__device__ __forceinline__ void xcd_range(int njobs, int& first, int& step, int& end)
{
    const int g = static_cast<int>(gridDim.x), b = static_cast<int>(blockIdx.x);
    const int x = b & 7, k = b >> 3;
    first = static_cast<int>(static_cast<long long>(njobs) * x / 8) + k;
    end = static_cast<int>(static_cast<long long>(njobs) * (x + 1) / 8);
    step = g >> 3;
}

// Values every lane of the block holds alike (job and dwell fields), moved
// to scalar registers: the strip body builds buffer descriptors from them,
// which must be scalar (a per-lane descriptor costs a loop per access).
__device__ __forceinline__ int uni(int v) { return __builtin_amdgcn_readfirstlane(v); }
__device__ __forceinline__ long long uni(long long v)
{
    const unsigned lo = __builtin_amdgcn_readfirstlane(static_cast<unsigned>(v));
    const unsigned hi = __builtin_amdgcn_readfirstlane(static_cast<unsigned>(static_cast<unsigned long long>(v) >> 32));
    return static_cast<long long>((static_cast<unsigned long long>(hi) << 32) | lo);
}
template <class T>
__device__ __forceinline__ T* uni(T* p)
{
    return reinterpret_cast<T*>(uni(static_cast<long long>(reinterpret_cast<uintptr_t>(p))));
}
__device__ __forceinline__ double uni(double v) { return __builtin_bit_cast(double, uni(__builtin_bit_cast(long long, v))); }

template <int N>
__device__ __forceinline__ void pc_job(const RunArgs& ra, const Job& job, float* lds)
{
    using C = fft::RowCfg<N>;
    constexpr int T = C::T;
    // opaque thread index: keeps the compiler from hoisting every pass's
    // twiddle loads out of the job loop into registers
    int tix = threadIdx.x;
    asm volatile("" : "+v"(tix));
    const int slot = tix / T, t = tix % T;
    const DwellDev& d = ra.dwells[uni(job.dwell)];
    const int dS = uni(d.S), dL = uni(d.L), dR = uni(d.R);
    const int jcount = uni(job.count), jfirst = uni(job.first);
    const bool active = slot < jcount;
    const int rs = jfirst + (active ? slot : 0);
    const int p = rs / dS, seg = rs - p * dS;
    const int k0 = seg * dL;
    const int count = min(dL, dR - k0);
    const float2* pulse = uni(ra.x) + uni(d.x_off) + static_cast<long long>(p) * dR;
    float2* out = uni(ra.ws) + uni(d.w_off) + static_cast<long long>(p) * dR;
    const float2* p16 = reinterpret_cast<const float2*>(reinterpret_cast<const char*>(lds) + kPcRowBytes);
    const fft::Pass16Twiddles tw{uni(ra.tw[ilog2_c(N)]), p16 + p16_offset(fft::Pass16Table<N>::R)};
    tiles::pc_segment_row<N>(pulse, dR, dR, k0, count, active, uni(d.H), out, t, lds + slot * 2 * C::SLOT_FLOATS,
                             tw);
}

__global__ void __launch_bounds__(kPcThreads, 4) pc_engine_kernel(const RunArgs* __restrict__ rap)
{
    extern __shared__ float lds[];
    const RunArgs& ra = *rap;
    {
        // the stride-16 pass tables of radix 2, 4, 8, 16 (from the 16R-point tables)
        float2* p16 = reinterpret_cast<float2*>(reinterpret_cast<char*>(lds) + kPcRowBytes);
        for (int e = threadIdx.x; e < kP16Entries; e += blockDim.x) {
            const int radix = e < 16 ? 2 : e < 64 ? 4 : e < 176 ? 8 : 16;
            const int le = e - p16_offset(radix);
            const int r = le / 16 + 1, jm = le % 16;
            p16[e] = ra.tw[ilog2_c(16 * radix)][r * jm];
        }
        __syncthreads();
    }
    int j, step, end;
    xcd_range(ra.n_pc, j, step, end);
    for (; j < end; j += step) {
        const Job job = ra.pc_jobs[j];
        switch (job.nlog) {
        case 4: pc_job<16>(ra, job, lds); break;
        case 5: pc_job<32>(ra, job, lds); break;
        case 6: pc_job<64>(ra, job, lds); break;
        case 7: pc_job<128>(ra, job, lds); break;
        case 8: pc_job<256>(ra, job, lds); break;
        case 9: pc_job<512>(ra, job, lds); break;
        case 10: pc_job<1024>(ra, job, lds); break;
        case 11: pc_job<2048>(ra, job, lds); break;
        case 12: pc_job<4096>(ra, job, lds); break;
        case 13: pc_job<8192>(ra, job, lds); break;
        default: break;
        }
    }
}

template <int N, bool FIXED>
__device__ __forceinline__ void dc_job(const RunArgs& ra, const Job& job, float* lds)
{
    using C = tiles::FusedCfg<N>;
    constexpr int TH = C::THREADS;
    int tix = threadIdx.x;  // opaque, as in pc_job
    asm volatile("" : "+v"(tix));
    const int sub = tix / TH, tid = tix % TH;
    const DwellDev d = ra.dwells[job.dwell];
    tiles::DcTile a;
    a.src = ra.ws + d.w_off;
    a.P = d.P;
    a.R = d.R;
    a.window = ra.window;
    a.g0 = ra.g0;
    a.g1 = ra.g1;
    a.r0 = ra.r0;
    a.r1 = ra.r1;
    a.alpha_over_n = ra.alpha_over_n;
    a.thr = ra.thr + d.o_off;
    a.det = ra.det + d.o_off;
    a.y = ra.rd ? ra.rd + d.o_off : nullptr;
    a.power = nullptr;
    // sub-blocks without a tile run the body on an out-of-range tile (no loads
    // or stores) so that every thread meets the same barriers
    const int tile = sub < job.count ? job.first + sub : (1 << 24);
    float* sl = lds + sub * C::floats(ra.g1 + ra.r1);
    if constexpr (FIXED)
        tiles::doppler_cfar_tile<N, 2, 2, 8, 8>(a, tile, tid, TH, sl, ra.tw[ilog2_c(N)]);
    else
        tiles::doppler_cfar_tile<N>(a, tile, tid, TH, sl, ra.tw[ilog2_c(N)]);
}

// FIXED: the engine's CFAR window is guard 2x2, reference 8x8 (compiled in)
template <bool FIXED>
__global__ void __launch_bounds__(kDcThreads) dc_engine_kernel(const RunArgs* __restrict__ rap)
{
    extern __shared__ float lds[];
    const RunArgs& ra = *rap;
    int j, step, end;
    xcd_range(ra.n_dc, j, step, end);
    for (; j < end; j += step) {
        const Job job = ra.dc_jobs[j];
        switch (job.nlog) {
        case 4: dc_job<16, FIXED>(ra, job, lds); break;
        case 5: dc_job<32, FIXED>(ra, job, lds); break;
        case 6: dc_job<64, FIXED>(ra, job, lds); break;
        case 7: dc_job<128, FIXED>(ra, job, lds); break;
        case 8: dc_job<256, FIXED>(ra, job, lds); break;
        default: break;
        }
    }
}

// One strip of a dwell: output bins [first, first + count), all P rows.
template <int N, int NT>
__device__ __forceinline__ void dcs_job(const RunArgs& ra, const Job& job, float* lds)
{
    const DwellDev d = ra.dwells[uni(job.dwell)];
    const long long w_off = uni(d.w_off), o_off = uni(d.o_off);
    strip::StripArgs a;
    a.x = uni(ra.ws) + w_off;
    a.R = uni(d.R);
    a.window = uni(ra.window);
    a.SW = uni(job.count);
    a.alpha_over_n = uni(ra.alpha_over_n);
    a.tw = uni(ra.tw[ilog2_c(N)]);
    a.thr = uni(ra.thr) + o_off;
    a.det = uni(ra.det) + o_off;
    float2* rd = uni(ra.rd);
    a.y = rd ? rd + o_off : nullptr;
    a.power = nullptr;
    const int first = uni(job.first);
    // opaque thread index: keeps the compiler from hoisting each body's
    // thread-dependent setup out of the job loop (it would stay live in
    // registers for all three bodies)
    int tix = threadIdx.x;
    asm volatile("" : "+v"(tix));
    strip::doppler_cfar_strip<N, 2, 2, 8, 8, NT>(a, 0, first, first + a.SW, lds, tix);
}

__global__ void __launch_bounds__(kDcsThreads, 2) dcs_engine_kernel(const RunArgs* __restrict__ rap)
{
    extern __shared__ float lds[];
    const RunArgs& ra = *rap;
    int j, step, end;
    xcd_range(ra.n_dcs, j, step, end);
    for (; j < end; j += step) {
        const Job job = ra.dcs_jobs[j];
        switch (job.nlog) {
        case 5: dcs_job<32, kDcsThreads>(ra, job, lds); break;
        case 6: dcs_job<64, kDcsThreads>(ra, job, lds); break;
        case 7: dcs_job<128, kDcsThreads>(ra, job, lds); break;
        default: break;
        }
    }
}

__global__ void __launch_bounds__(kDcs256Threads) dcs256_engine_kernel(const RunArgs* __restrict__ rap)
{
    extern __shared__ float lds[];
    const RunArgs& ra = *rap;
    int j, step, end;
    xcd_range(ra.n_dcs256, j, step, end);
    for (; j < end; j += step) dcs_job<256, kDcs256Threads>(ra, ra.dcs256_jobs[j], lds);
}

template <int N, int NT>
using EngineStrip = strip::StripCfg<N, 2, 2, 8, 8, NT>;
constexpr size_t kDcsLdsBytes = std::max({EngineStrip<32, kDcsThreads>::LDS_BYTES, EngineStrip<64, kDcsThreads>::LDS_BYTES,
                                          EngineStrip<128, kDcsThreads>::LDS_BYTES});
constexpr size_t kDcs256LdsBytes = EngineStrip<256, kDcs256Threads>::LDS_BYTES;

// The strip body's buffer offsets are 32-bit: a dwell's P*R*8 bytes stay below 2^31.
bool strip_dwell(bool fixed, int P, int R)
{
    return fixed && P >= 32 && P <= 256 && static_cast<long long>(P) * R * 8 < (1LL << 31);
}
int strip_width(int P)
{
    const int cw = P == 256 ? EngineStrip<256, kDcs256Threads>::CW : kDcsThreads / (P / 16);
    return kStripChunksPerJob * cw;
}

// tiles of 64 range bins per Doppler job for P pulses, and the LDS one job needs
int dc_tiles_per_job(int nlog) { return kDcThreads / (4 << nlog); }

template <int N>
size_t dc_job_lds(int cm)
{
    return sizeof(float) * static_cast<size_t>(tiles::FusedCfg<N>::floats(cm)) * (kDcThreads / tiles::FusedCfg<N>::THREADS);
}

size_t dc_lds_bytes(int cm)
{
    return std::max({dc_job_lds<16>(cm), dc_job_lds<32>(cm), dc_job_lds<64>(cm), dc_job_lds<128>(cm),
                     dc_job_lds<256>(cm)});
}

struct TemplateRec {
    float2* taps = nullptr;
    int ntaps = 0;
    float2* spec[kMaxLog + 1] = {};
};

struct Lane {
    char* dev = nullptr;
    size_t dev_cap = 0;
    char* host[kSlots] = {};
    size_t host_cap[kSlots] = {};
    hipEvent_t ev[kSlots] = {};
    bool ev_live[kSlots] = {};
    int slot = 0;
    float2* ws = nullptr;
    size_t ws_cap = 0;  // complex64 elements
    hipStream_t stream = nullptr;
    bool used = false;
    hipGraphExec_t exec = nullptr;
    int exec_pc_grid = 0, exec_dc_grid = 0;
};

size_t align16(size_t v) { return (v + 15) & ~static_cast<size_t>(15); }

}  // namespace
}  // namespace zkx

struct zkx_engine {
    int device = 0;
    int cus = 256;
    int pc_window = ZKX_WINDOW_HAMMING, normalize = 1, dop_window = ZKX_WINDOW_HANN;
    int g0 = 2, g1 = 2, r0 = 8, r1 = 8;
    double pfa = 1e-6, alpha_over_n = 0.0;
    size_t dc_lds = 0;
    bool fixed_window = false;
    bool use_graph = true;
    int pc_grid = 0, dc_grid = 0;
    int dcs_grid = 0, dcs256_grid = 0;  // strip kernels: two and one resident blocks per CU
    std::vector<zkx::TemplateRec> templates;
    std::vector<zkx::Lane> lanes;
    std::map<std::pair<int, int>, int> pc_choice;  // (R, T) -> log2 of the FFT length
    const float2* tw[zkx::kMaxLog + 1] = {};
    std::mutex mu;
};

namespace zkx {
namespace {

bool is_pow2_int(int v) { return v > 0 && (v & (v - 1)) == 0; }

// Segment geometry of one pulse for FFT length 2^nlog: (valid outputs per
// segment, segments per pulse), or L = 0 if the length cannot serve (R, T).
void pc_geometry(int R, int T, int nlog, int* L, int* S)
{
    const int n = 1 << nlog;
    *L = 0;
    *S = 0;
    if (n == R) {
        *L = R;
        *S = 1;
        return;
    }
    // overlap-save: N > T, and at most R threads per row (n/16 <= R), since
    // pc_segment_row wraps each load position with one subtraction
    if (n <= T || (n / fft::elems_for(n)) > R) return;
    const int l = std::min(n - T + 1, R);
    *L = l;
    *S = (R + l - 1) / l;
}

// Tuning classes: range length and the template length rounded up to a
// multiple of 128 (a length tuned for the class serves every T in it).
std::pair<int, int> tune_key(int R, int T) { return {R, (T + 127) / 128 * 128}; }

// FFT length for a (R, T) class: the tuned choice, else the cheapest by
// S * N * (log2 N + 2) per pulse (transform pair plus load/store).
int choose_nlog(const zkx_engine& e, int R, int T)
{
    auto it = e.pc_choice.find(tune_key(R, T));
    if (it != e.pc_choice.end()) {
        int L, S;
        pc_geometry(R, T, it->second, &L, &S);
        if (L) return it->second;
    }
    int best = -1;
    double best_cost = 0.0;
    for (int nl = kMinLog; nl <= kMaxLog; ++nl) {
        int L, S;
        pc_geometry(R, T, nl, &L, &S);
        if (!L) continue;
        const double cost = static_cast<double>(S) * (1 << nl) * (nl + 2);
        if (best < 0 || cost < best_cost) {
            best = nl;
            best_cost = cost;
        }
    }
    return best;
}

std::vector<int> candidate_nlogs(int R, int T)
{
    std::vector<int> out;
    for (int nl = kMinLog; nl <= kMaxLog; ++nl) {
        int L, S;
        pc_geometry(R, T, nl, &L, &S);
        // keep lengths whose segments carry at least a quarter useful output
        if (L && static_cast<long long>(L) * 4 >= std::min(1 << nl, 4 * R)) out.push_back(nl);
    }
    return out;
}

int capture_graph(zkx_engine& e, Lane& lane)
{
    if (lane.exec) {
        (void)hipGraphExecDestroy(lane.exec);
        lane.exec = nullptr;
    }
    hipStream_t cs;
    ZKX_HIP_CHECK(hipStreamCreateWithFlags(&cs, hipStreamNonBlocking));
    hipGraph_t g = nullptr;
    hipError_t err = hipStreamBeginCapture(cs, hipStreamCaptureModeThreadLocal);
    if (err == hipSuccess) {
        const RunArgs* ra = reinterpret_cast<const RunArgs*>(lane.dev);
        hipLaunchKernelGGL(pc_engine_kernel, dim3(e.pc_grid), dim3(kPcThreads), kPcLdsBytes, cs, ra);
        if (e.fixed_window)
            hipLaunchKernelGGL(dc_engine_kernel<true>, dim3(e.dc_grid), dim3(kDcThreads), e.dc_lds, cs, ra);
        else
            hipLaunchKernelGGL(dc_engine_kernel<false>, dim3(e.dc_grid), dim3(kDcThreads), e.dc_lds, cs, ra);
        hipLaunchKernelGGL(dcs_engine_kernel, dim3(e.dcs_grid), dim3(kDcsThreads), kDcsLdsBytes, cs, ra);
        hipLaunchKernelGGL(dcs256_engine_kernel, dim3(e.dcs256_grid), dim3(kDcs256Threads), kDcs256LdsBytes, cs, ra);
        err = hipStreamEndCapture(cs, &g);
    }
    if (err == hipSuccess) err = hipGraphInstantiate(&lane.exec, g, nullptr, nullptr, 0);
    if (g) (void)hipGraphDestroy(g);
    (void)hipStreamDestroy(cs);
    if (err != hipSuccess) {
        lane.exec = nullptr;
        set_last_error(hip_status_to_code(err), hipGetErrorString(err));
        return hip_status_to_code(err);
    }
    lane.exec_pc_grid = e.pc_grid;
    lane.exec_dc_grid = e.dc_grid;
    return ZKX_SUCCESS;
}

// Plan, stage and launch one batch on a lane.
int engine_run(zkx_engine& e, int lane_idx, const float2* x, const zkx_dwell_desc* descs, int n, float* thr,
               unsigned char* det, float2* rd, hipStream_t st)
{
    if (lane_idx < 0 || lane_idx >= static_cast<int>(e.lanes.size()))
        ZKX_FAIL(ZKX_INVALID_ARGUMENT, "engine lane out of range");
    if (n < 0 || (n > 0 && (!descs || !x || !thr || !det)))
        ZKX_FAIL(ZKX_INVALID_ARGUMENT, "engine run needs input, dwell table and outputs");
    if (n == 0) return ZKX_SUCCESS;
    Lane& lane = e.lanes[lane_idx];
    ZKX_HIP_CHECK(hipSetDevice(e.device));
    if (lane.used && lane.stream != st) ZKX_HIP_CHECK(hipStreamSynchronize(lane.stream));
    lane.stream = st;
    lane.used = true;

    // ---- plan ----
    std::vector<DwellDev> dw(n);
    std::vector<Job> pcj, dcj, dcsj, dcs256j;
    long long ws_need = 0;
    const int cm = e.g1 + e.r1;
    const int tco = tiles::FusedCfg<16>::NCM - 2 * cm;
    {
        std::lock_guard<std::mutex> lk(e.mu);
        for (int i = 0; i < n; ++i) {
            const zkx_dwell_desc& d = descs[i];
            if (!is_pow2_int(d.num_pulses) || d.num_pulses < (1 << kMinDopplerLog) ||
                d.num_pulses > (1 << kMaxDopplerLog))
                ZKX_FAIL(ZKX_INVALID_SHAPE, "engine dwells need a power-of-two pulse count in [16, 256]");
            if (d.range_bins < 1 || static_cast<long long>(d.num_pulses) * d.range_bins >= (1LL << 31))
                ZKX_FAIL(ZKX_INVALID_SHAPE, "engine dwell range length must be positive and P*R < 2^31");
            if (d.template_id < 0 || d.template_id >= static_cast<int>(e.templates.size()))
                ZKX_FAIL(ZKX_INVALID_ARGUMENT, "engine dwell names an unknown template");
            if (d.input_offset < 0 || d.output_offset < 0)
                ZKX_FAIL(ZKX_INVALID_ARGUMENT, "engine dwell offsets must be non-negative");
            const TemplateRec& tr = e.templates[d.template_id];
            if (tr.ntaps > d.range_bins) ZKX_FAIL(ZKX_INVALID_SHAPE, "engine template longer than the range length");
            const int nl = choose_nlog(e, d.range_bins, tr.ntaps);
            if (nl < 0) ZKX_FAIL(ZKX_UNSUPPORTED_OPERATION, "no FFT length serves this range length and template");
            DwellDev& o = dw[i];
            o.x_off = d.input_offset;
            o.o_off = d.output_offset;
            o.w_off = ws_need;
            o.P = d.num_pulses;
            o.R = d.range_bins;
            o.nlog = nl;
            o.N = 1 << nl;
            pc_geometry(o.R, tr.ntaps, nl, &o.L, &o.S);
            o.H = tr.spec[nl];
            if (!o.H) ZKX_FAIL(ZKX_INTERNAL_ERROR, "template spectrum missing");
            ws_need += static_cast<long long>(o.P) * o.R;
            const long long rows = static_cast<long long>(o.P) * o.S;
            const int fpb = kPcThreads / (o.N / fft::elems_for(o.N));
            for (long long f = 0; f < rows; f += fpb)
                pcj.push_back({i, nl, static_cast<int>(f), static_cast<int>(std::min<long long>(fpb, rows - f))});
            const int plog = ilog2(o.P);
            if (strip_dwell(e.fixed_window, o.P, o.R)) {
                const int sw = strip_width(o.P);
                std::vector<Job>& dst = o.P == 256 ? dcs256j : dcsj;
                for (int c0 = 0; c0 < o.R; c0 += sw) dst.push_back({i, plog, c0, std::min(sw, o.R - c0)});
            } else {
                const int tiles_n = (o.R + tco - 1) / tco;
                const int per = dc_tiles_per_job(plog);
                for (int t0 = 0; t0 < tiles_n; t0 += per) dcj.push_back({i, plog, t0, std::min(per, tiles_n - t0)});
            }
        }
    }

    // ---- workspace ----
    if (static_cast<size_t>(ws_need) > lane.ws_cap) {
        ZKX_HIP_CHECK(hipStreamSynchronize(st));
        if (lane.ws) (void)hipFree(lane.ws);
        lane.ws = nullptr;
        lane.ws_cap = 0;
        const size_t cap = static_cast<size_t>(ws_need) + static_cast<size_t>(ws_need) / 4;
        ZKX_HIP_CHECK(hipMalloc(reinterpret_cast<void**>(&lane.ws), cap * sizeof(float2)));
        lane.ws_cap = cap;
    }

    // ---- stage tables ----
    const size_t off_dw = align16(sizeof(RunArgs));
    const size_t off_pc = align16(off_dw + sizeof(DwellDev) * dw.size());
    const size_t off_dc = align16(off_pc + sizeof(Job) * pcj.size());
    const size_t off_dcs = align16(off_dc + sizeof(Job) * dcj.size());
    const size_t off_dcs256 = align16(off_dcs + sizeof(Job) * dcsj.size());
    const size_t bytes = align16(off_dcs256 + sizeof(Job) * dcs256j.size());
    if (bytes > lane.dev_cap) {
        ZKX_HIP_CHECK(hipStreamSynchronize(st));
        if (lane.dev) (void)hipFree(lane.dev);
        lane.dev = nullptr;
        lane.dev_cap = 0;
        const size_t cap = std::max<size_t>(bytes * 2, 64 * 1024);
        ZKX_HIP_CHECK(hipMalloc(reinterpret_cast<void**>(&lane.dev), cap));
        lane.dev_cap = cap;
        if (lane.exec) {
            (void)hipGraphExecDestroy(lane.exec);
            lane.exec = nullptr;
        }
    }
    const int s = lane.slot;
    lane.slot = (lane.slot + 1) % kSlots;
    if (lane.ev_live[s]) ZKX_HIP_CHECK(hipEventSynchronize(lane.ev[s]));
    if (bytes > lane.host_cap[s]) {
        if (lane.host[s]) (void)hipHostFree(lane.host[s]);
        lane.host[s] = nullptr;
        lane.host_cap[s] = 0;
        const size_t cap = std::max<size_t>(bytes * 2, 64 * 1024);
        ZKX_HIP_CHECK(hipHostMalloc(reinterpret_cast<void**>(&lane.host[s]), cap, hipHostMallocDefault));
        lane.host_cap[s] = cap;
    }
    char* h = lane.host[s];
    RunArgs ra;
    std::memset(&ra, 0, sizeof(ra));
    ra.x = x;
    ra.thr = thr;
    ra.det = det;
    ra.rd = rd;
    ra.ws = lane.ws;
    ra.dwells = reinterpret_cast<const DwellDev*>(lane.dev + off_dw);
    ra.pc_jobs = reinterpret_cast<const Job*>(lane.dev + off_pc);
    ra.dc_jobs = reinterpret_cast<const Job*>(lane.dev + off_dc);
    ra.dcs_jobs = reinterpret_cast<const Job*>(lane.dev + off_dcs);
    ra.dcs256_jobs = reinterpret_cast<const Job*>(lane.dev + off_dcs256);
    ra.n_pc = static_cast<int>(pcj.size());
    ra.n_dc = static_cast<int>(dcj.size());
    ra.n_dcs = static_cast<int>(dcsj.size());
    ra.n_dcs256 = static_cast<int>(dcs256j.size());
    ra.window = e.dop_window;
    ra.g0 = e.g0;
    ra.g1 = e.g1;
    ra.r0 = e.r0;
    ra.r1 = e.r1;
    ra.alpha_over_n = e.alpha_over_n;
    for (int l = 0; l <= kMaxLog; ++l) ra.tw[l] = e.tw[l];
    std::memcpy(h, &ra, sizeof(ra));
    std::memcpy(h + off_dw, dw.data(), sizeof(DwellDev) * dw.size());
    if (!pcj.empty()) std::memcpy(h + off_pc, pcj.data(), sizeof(Job) * pcj.size());
    if (!dcj.empty()) std::memcpy(h + off_dc, dcj.data(), sizeof(Job) * dcj.size());
    if (!dcsj.empty()) std::memcpy(h + off_dcs, dcsj.data(), sizeof(Job) * dcsj.size());
    if (!dcs256j.empty()) std::memcpy(h + off_dcs256, dcs256j.data(), sizeof(Job) * dcs256j.size());
    ZKX_HIP_CHECK(hipMemcpyAsync(lane.dev, h, bytes, hipMemcpyHostToDevice, st));

    // ---- launch ----
    if (e.use_graph) {
        if (!lane.exec || lane.exec_pc_grid != e.pc_grid || lane.exec_dc_grid != e.dc_grid) {
            const int rc = capture_graph(e, lane);
            if (rc) return rc;
        }
        ZKX_HIP_CHECK(hipGraphLaunch(lane.exec, st));
    } else {
        const RunArgs* rap = reinterpret_cast<const RunArgs*>(lane.dev);
        hipLaunchKernelGGL(pc_engine_kernel, dim3(e.pc_grid), dim3(kPcThreads), kPcLdsBytes, st, rap);
        ZKX_LAUNCH_CHECK();
        if (e.fixed_window)
            hipLaunchKernelGGL(dc_engine_kernel<true>, dim3(e.dc_grid), dim3(kDcThreads), e.dc_lds, st, rap);
        else
            hipLaunchKernelGGL(dc_engine_kernel<false>, dim3(e.dc_grid), dim3(kDcThreads), e.dc_lds, st, rap);
        ZKX_LAUNCH_CHECK();
        if (!dcsj.empty()) {
            hipLaunchKernelGGL(dcs_engine_kernel, dim3(e.dcs_grid), dim3(kDcsThreads), kDcsLdsBytes, st, rap);
            ZKX_LAUNCH_CHECK();
        }
        if (!dcs256j.empty()) {
            hipLaunchKernelGGL(dcs256_engine_kernel, dim3(e.dcs256_grid), dim3(kDcs256Threads), kDcs256LdsBytes, st,
                               rap);
            ZKX_LAUNCH_CHECK();
        }
    }
    ZKX_HIP_CHECK(hipEventRecord(lane.ev[s], st));
    lane.ev_live[s] = true;
    return ZKX_SUCCESS;
}

void destroy_engine(zkx_engine* e)
{
    if (!e) return;
    (void)hipSetDevice(e->device);
    (void)hipDeviceSynchronize();
    for (Lane& l : e->lanes) {
        if (l.exec) (void)hipGraphExecDestroy(l.exec);
        if (l.dev) (void)hipFree(l.dev);
        if (l.ws) (void)hipFree(l.ws);
        for (int s = 0; s < kSlots; ++s) {
            if (l.host[s]) (void)hipHostFree(l.host[s]);
            if (l.ev[s]) (void)hipEventDestroy(l.ev[s]);
        }
    }
    for (TemplateRec& t : e->templates) {
        if (t.taps) (void)hipFree(t.taps);
        for (float2* p : t.spec)
            if (p) (void)hipFree(p);
    }
    delete e;
}

int round8(int v) { return std::max(8, (v + 7) / 8 * 8); }

}  // namespace
}  // namespace zkx

using zkx::set_last_error;

extern "C" {

int zkx_engine_create(zkx_engine** out, int pc_window, int normalize, int doppler_window, int guard_doppler,
                      int guard_range, int ref_doppler, int ref_range, double pfa, int lanes)
{
    zkx::clear_last_error();
    if (!out) ZKX_FAIL(ZKX_INVALID_ARGUMENT, "engine handle pointer is null");
    *out = nullptr;
    if (lanes < 1 || lanes > 64) ZKX_FAIL(ZKX_INVALID_ARGUMENT, "engine lanes must be in [1, 64]");
    if (pc_window < ZKX_WINDOW_NONE || pc_window > ZKX_WINDOW_HAMMING || doppler_window < ZKX_WINDOW_NONE ||
        doppler_window > ZKX_WINDOW_HAMMING)
        ZKX_FAIL(ZKX_INVALID_ARGUMENT, "engine windows must be none, hann or hamming");
    if (guard_doppler < 0 || guard_range < 0 || ref_doppler < 0 || ref_range < 0)
        ZKX_FAIL(ZKX_INVALID_ARGUMENT, "ca_cfar cell counts must be non-negative");
    if (!(pfa > 0.0 && pfa <= 1.0)) ZKX_FAIL(ZKX_INVALID_ARGUMENT, "ca_cfar pfa must be in (0,1]");
    const int count = 2 * ref_doppler * (2 * ref_range + 2 * guard_range + 1) + 2 * (2 * guard_doppler + 1) * ref_range;
    if (count < 1) ZKX_FAIL(ZKX_INVALID_ARGUMENT, "ca_cfar needs at least one reference cell");
    const int cm = guard_range + ref_range;
    if (zkx::tiles::FusedCfg<16>::NCM - 2 * cm < 8)
        ZKX_FAIL(ZKX_UNSUPPORTED_OPERATION, "engine CFAR range halo must leave at least 8 output bins per tile");
    const size_t lds = zkx::dc_lds_bytes(cm);
    if (lds > 160 * 1024) ZKX_FAIL(ZKX_UNSUPPORTED_OPERATION, "engine CFAR window needs more LDS than a CU has");
    zkx_engine* e = new (std::nothrow) zkx_engine();
    if (!e) ZKX_FAIL(ZKX_OUT_OF_MEMORY, "engine allocation failed");
    int rc = ZKX_SUCCESS;
    do {
        hipError_t err = hipGetDevice(&e->device);
        if (err == hipSuccess) err = hipDeviceGetAttribute(&e->cus, hipDeviceAttributeMultiprocessorCount, e->device);
        if (err != hipSuccess) {
            rc = zkx::hip_status_to_code(err);
            set_last_error(rc, hipGetErrorString(err));
            break;
        }
        e->pc_window = pc_window;
        e->normalize = normalize ? 1 : 0;
        e->dop_window = doppler_window;
        e->g0 = guard_doppler;
        e->g1 = guard_range;
        e->r0 = ref_doppler;
        e->r1 = ref_range;
        e->pfa = pfa;
        const double nn = static_cast<double>(count);
        e->alpha_over_n = (nn * (std::pow(pfa, -1.0 / nn) - 1.0)) / nn;
        e->dc_lds = lds;
        e->pc_grid = zkx::round8(2 * e->cus);
        e->dc_grid = zkx::round8(e->cus);
        e->dcs_grid = zkx::round8(2 * e->cus);
        e->dcs256_grid = zkx::round8(e->cus);
        for (int l = zkx::kMinLog; l <= zkx::kMaxLog && !rc; ++l) rc = zkx::twiddles(1 << l, &e->tw[l]);
        if (rc) break;
        e->fixed_window = guard_doppler == 2 && guard_range == 2 && ref_doppler == 8 && ref_range == 8;
        err = hipFuncSetAttribute(reinterpret_cast<const void*>(zkx::dc_engine_kernel<true>),
                                  hipFuncAttributeMaxDynamicSharedMemorySize, static_cast<int>(lds));
        if (err == hipSuccess)
            err = hipFuncSetAttribute(reinterpret_cast<const void*>(zkx::dc_engine_kernel<false>),
                                      hipFuncAttributeMaxDynamicSharedMemorySize, static_cast<int>(lds));
        if (err == hipSuccess)
            err = hipFuncSetAttribute(reinterpret_cast<const void*>(zkx::pc_engine_kernel),
                                      hipFuncAttributeMaxDynamicSharedMemorySize, static_cast<int>(zkx::kPcLdsBytes));
        if (err == hipSuccess)
            err = hipFuncSetAttribute(reinterpret_cast<const void*>(zkx::dcs_engine_kernel),
                                      hipFuncAttributeMaxDynamicSharedMemorySize, static_cast<int>(zkx::kDcsLdsBytes));
        if (err == hipSuccess)
            err = hipFuncSetAttribute(reinterpret_cast<const void*>(zkx::dcs256_engine_kernel),
                                      hipFuncAttributeMaxDynamicSharedMemorySize,
                                      static_cast<int>(zkx::kDcs256LdsBytes));
        if (err != hipSuccess) {
            rc = zkx::hip_status_to_code(err);
            set_last_error(rc, hipGetErrorString(err));
            break;
        }
        e->lanes.resize(lanes);
        for (zkx::Lane& ln : e->lanes) {
            for (int s = 0; s < zkx::kSlots && !rc; ++s) {
                err = hipEventCreateWithFlags(&ln.ev[s], hipEventDisableTiming);
                if (err != hipSuccess) {
                    rc = zkx::hip_status_to_code(err);
                    set_last_error(rc, hipGetErrorString(err));
                }
            }
        }
    } while (false);
    if (rc) {
        zkx::destroy_engine(e);
        return rc;
    }
    *out = e;
    return ZKX_SUCCESS;
}

int zkx_engine_destroy(zkx_engine* engine)
{
    zkx::clear_last_error();
    if (!engine) ZKX_FAIL(ZKX_INVALID_ARGUMENT, "engine handle is null");
    zkx::destroy_engine(engine);
    return ZKX_SUCCESS;
}

int zkx_engine_add_template(zkx_engine* e, const zkx_complex64* h, int taps, int* id)
{
    zkx::clear_last_error();
    if (!e || !h || !id) ZKX_FAIL(ZKX_INVALID_ARGUMENT, "engine, template and id must be non-null");
    if (taps < 1 || taps > 8191) ZKX_FAIL(ZKX_INVALID_SHAPE, "engine template length must be in [1, 8191]");
    ZKX_HIP_CHECK(hipSetDevice(e->device));
    zkx::TemplateRec t;
    t.ntaps = taps;
    int rc = ZKX_SUCCESS;
    hipError_t err = hipMalloc(reinterpret_cast<void**>(&t.taps), sizeof(float2) * taps);
    if (err == hipSuccess)
        err = hipMemcpy(t.taps, h, sizeof(float2) * taps, hipMemcpyDefault);
    for (int l = zkx::kMinLog; l <= zkx::kMaxLog && err == hipSuccess && !rc; ++l) {
        if ((1 << l) < taps) continue;  // lengths shorter than the template never serve it
        err = hipMalloc(reinterpret_cast<void**>(&t.spec[l]), sizeof(float2) << l);
        if (err == hipSuccess)
            rc = zkx::pc_template_dispatch(1 << l, t.taps, taps, e->normalize, e->pc_window, t.spec[l], nullptr);
    }
    if (err == hipSuccess && !rc) err = hipStreamSynchronize(nullptr);
    if (err != hipSuccess || rc) {
        if (err != hipSuccess) {
            rc = zkx::hip_status_to_code(err);
            set_last_error(rc, hipGetErrorString(err));
        }
        if (t.taps) (void)hipFree(t.taps);
        for (float2* p : t.spec)
            if (p) (void)hipFree(p);
        return rc;
    }
    std::lock_guard<std::mutex> lk(e->mu);
    e->templates.push_back(t);
    *id = static_cast<int>(e->templates.size()) - 1;
    return ZKX_SUCCESS;
}

int zkx_engine_set_options(zkx_engine* e, int use_graph, int pc_grid, int dc_grid)
{
    zkx::clear_last_error();
    if (!e) ZKX_FAIL(ZKX_INVALID_ARGUMENT, "engine handle is null");
    if (pc_grid < 0 || dc_grid < 0 || pc_grid > 65536 || dc_grid > 65536)
        ZKX_FAIL(ZKX_INVALID_ARGUMENT, "engine grid sizes must be in [0, 65536]");
    e->use_graph = use_graph != 0;
    if (pc_grid) e->pc_grid = zkx::round8(pc_grid);
    if (dc_grid) e->dc_grid = zkx::round8(dc_grid);
    return ZKX_SUCCESS;
}

int zkx_engine_run(zkx_engine* e, int lane, const zkx_complex64* x, const zkx_dwell_desc* dwells, int n,
                   float* threshold, uint8_t* detections, zkx_complex64* rd_map, void* stream)
{
    zkx::clear_last_error();
    if (!e) ZKX_FAIL(ZKX_INVALID_ARGUMENT, "engine handle is null");
    return zkx::engine_run(*e, lane, reinterpret_cast<const float2*>(x), dwells, n, threshold, detections,
                           reinterpret_cast<float2*>(rd_map), static_cast<hipStream_t>(stream));
}

int zkx_engine_plan(zkx_engine* e, const zkx_dwell_desc* dwells, int n, int* fft_length, int* segments)
{
    zkx::clear_last_error();
    if (!e || (n > 0 && (!dwells || !fft_length || !segments))) ZKX_FAIL(ZKX_INVALID_ARGUMENT, "null argument");
    std::lock_guard<std::mutex> lk(e->mu);
    for (int i = 0; i < n; ++i) {
        const zkx_dwell_desc& d = dwells[i];
        if (!zkx::is_pow2_int(d.num_pulses) || d.num_pulses < 16 || d.num_pulses > 256 || d.range_bins < 1 ||
            static_cast<long long>(d.num_pulses) * d.range_bins >= (1LL << 31))
            ZKX_FAIL(ZKX_INVALID_SHAPE, "engine dwells need a power-of-two pulse count in [16, 256]");
        const int tid = d.template_id;
        if (tid < 0 || tid >= static_cast<int>(e->templates.size()))
            ZKX_FAIL(ZKX_INVALID_ARGUMENT, "engine dwell names an unknown template");
        const int T = e->templates[tid].ntaps;
        if (T > d.range_bins) ZKX_FAIL(ZKX_INVALID_SHAPE, "engine template longer than the range length");
        const int nl = zkx::choose_nlog(*e, d.range_bins, T);
        if (nl < 0) ZKX_FAIL(ZKX_UNSUPPORTED_OPERATION, "no FFT length serves this range length and template");
        int L = 0, S = 0;
        zkx::pc_geometry(d.range_bins, T, nl, &L, &S);
        fft_length[i] = 1 << nl;
        segments[i] = S;
    }
    return ZKX_SUCCESS;
}

int zkx_engine_autotune(zkx_engine* e, int lane, const zkx_complex64* x, const zkx_dwell_desc* dwells, int n,
                        float* threshold, uint8_t* detections, void* stream)
{
    zkx::clear_last_error();
    if (!e || n < 1 || !dwells) ZKX_FAIL(ZKX_INVALID_ARGUMENT, "autotune needs an engine and a batch");
    hipStream_t st = static_cast<hipStream_t>(stream);
    hipEvent_t a, b;
    ZKX_HIP_CHECK(hipEventCreate(&a));
    ZKX_HIP_CHECK(hipEventCreate(&b));
    int rc = ZKX_SUCCESS;
    auto time_run = [&](const std::vector<zkx_dwell_desc>& batch, float* ms) -> int {
        int r = zkx::engine_run(*e, lane, reinterpret_cast<const float2*>(x), batch.data(),
                                static_cast<int>(batch.size()), threshold, detections, nullptr, st);
        if (r) return r;
        ZKX_HIP_CHECK(hipEventRecord(a, st));
        for (int k = 0; k < 3 && !r; ++k)
            r = zkx::engine_run(*e, lane, reinterpret_cast<const float2*>(x), batch.data(),
                                static_cast<int>(batch.size()), threshold, detections, nullptr, st);
        if (r) return r;
        ZKX_HIP_CHECK(hipEventRecord(b, st));
        ZKX_HIP_CHECK(hipEventSynchronize(b));
        ZKX_HIP_CHECK(hipEventElapsedTime(ms, a, b));
        return ZKX_SUCCESS;
    };
    // 1. FFT length per (R, T) class, timed on that class's dwells alone
    std::map<std::pair<int, int>, std::vector<zkx_dwell_desc>> classes;
    for (int i = 0; i < n && !rc; ++i) {
        const int tid = dwells[i].template_id;
        if (tid < 0 || tid >= static_cast<int>(e->templates.size())) {
            set_last_error(ZKX_INVALID_ARGUMENT, "engine dwell names an unknown template");
            rc = ZKX_INVALID_ARGUMENT;
            break;
        }
        classes[zkx::tune_key(dwells[i].range_bins, e->templates[tid].ntaps)].push_back(dwells[i]);
    }
    for (auto it = classes.begin(); it != classes.end() && !rc; ++it) {
        const std::vector<int> cands = zkx::candidate_nlogs(it->first.first, it->first.second);
        if (cands.size() < 2) continue;
        std::vector<zkx_dwell_desc> batch(it->second.begin(),
                                          it->second.begin() + std::min<size_t>(it->second.size(), 16));
        int best = -1;
        float best_ms = 0.f;
        for (int nl : cands) {
            {
                std::lock_guard<std::mutex> lk(e->mu);
                e->pc_choice[it->first] = nl;
            }
            float ms = 0.f;
            if ((rc = time_run(batch, &ms))) break;
            if (best < 0 || ms < best_ms) {
                best = nl;
                best_ms = ms;
            }
        }
        std::lock_guard<std::mutex> lk(e->mu);
        if (best >= 0) e->pc_choice[it->first] = best;
    }
    // 2. persistent grid sizes on the whole batch
    if (!rc) {
        std::vector<zkx_dwell_desc> all(dwells, dwells + n);
        const int pcs[3] = {e->cus, 2 * e->cus, 4 * e->cus};
        const int dcs[2] = {e->cus, 2 * e->cus};
        int best_pc = e->pc_grid, best_dc = e->dc_grid;
        float best_ms = -1.f;
        for (int pg : pcs) {
            for (int dg : dcs) {
                e->pc_grid = zkx::round8(pg);
                e->dc_grid = zkx::round8(dg);
                float ms = 0.f;
                if ((rc = time_run(all, &ms))) break;
                if (best_ms < 0.f || ms < best_ms) {
                    best_ms = ms;
                    best_pc = e->pc_grid;
                    best_dc = e->dc_grid;
                }
            }
            if (rc) break;
        }
        e->pc_grid = best_pc;
        e->dc_grid = best_dc;
    }
    (void)hipEventDestroy(a);
    (void)hipEventDestroy(b);
    return rc;
}

int zkx_engine_tuning_export(zkx_engine* e, char* buf, int cap, int* length)
{
    zkx::clear_last_error();
    if (!e || !length) ZKX_FAIL(ZKX_INVALID_ARGUMENT, "null argument");
    std::string s;
    char line[96];
    std::snprintf(line, sizeof(line), "grid %d %d\n", e->pc_grid, e->dc_grid);
    s += line;
    {
        std::lock_guard<std::mutex> lk(e->mu);
        for (const auto& kv : e->pc_choice) {
            std::snprintf(line, sizeof(line), "pc %d %d %d\n", kv.first.first, kv.first.second, 1 << kv.second);
            s += line;
        }
    }
    *length = static_cast<int>(s.size());
    if (buf) {
        if (cap <= static_cast<int>(s.size())) ZKX_FAIL(ZKX_INVALID_ARGUMENT, "tuning buffer too small");
        std::memcpy(buf, s.c_str(), s.size() + 1);
    }
    return ZKX_SUCCESS;
}

int zkx_engine_tuning_import(zkx_engine* e, const char* text)
{
    zkx::clear_last_error();
    if (!e || !text) ZKX_FAIL(ZKX_INVALID_ARGUMENT, "null argument");
    const char* p = text;
    while (*p) {
        int a = 0, b = 0, c = 0, used = 0;
        if (std::sscanf(p, "grid %d %d%n", &a, &b, &used) == 2 && used > 0) {
            if (a < 8 || b < 8 || a > 65536 || b > 65536) ZKX_FAIL(ZKX_INVALID_ARGUMENT, "bad grid line");
            e->pc_grid = zkx::round8(a);
            e->dc_grid = zkx::round8(b);
        } else if (std::sscanf(p, "pc %d %d %d%n", &a, &b, &c, &used) == 3 && used > 0) {
            if (!zkx::is_pow2_int(c) || c < 16 || c > 8192) ZKX_FAIL(ZKX_INVALID_ARGUMENT, "bad pc line");
            int L, S;
            zkx::pc_geometry(a, b, zkx::ilog2(c), &L, &S);
            if (!L) ZKX_FAIL(ZKX_INVALID_ARGUMENT, "pc line names a length that cannot serve its class");
            std::lock_guard<std::mutex> lk(e->mu);
            e->pc_choice[{a, b}] = zkx::ilog2(c);
        } else if (*p != '\n' && *p != ' ') {
            ZKX_FAIL(ZKX_INVALID_ARGUMENT, "unrecognised tuning line");
        }
        const char* nl = std::strchr(p, '\n');
        if (!nl) break;
        p = nl + 1;
    }
    return ZKX_SUCCESS;
}

}  // extern "C"
