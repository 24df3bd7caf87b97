// The reference's three FFT-backend protocols
// (ZKX/test_all/operators/fft_backend/src/fft_backend_{performance,
// stability,memory_unified}.cu) run against this library's fft_* ABI, built
// with g++ and linked to libzkx_radar.so.
//
// usage: fft_backend_protocols [quick]
//   performance: per case (n, batch) device-resident forward transforms,
//     5 warm-up then 5 batches of 20 timed executions (HIP events around
//     each), centre (median batch) mean, pooled mean, p50/p95/p99, plan
//     creation time, heap/device growth over the timed executions, and a
//     forward + inverse round trip (rel-L2 and max |error|);
//   stability: per case, executions until three consecutive blocks leave
//     the host heap (within 4 KiB) and device use flat, then checkpoints of
//     executions whose growth must be <= 4 KiB (host) and <= 0 (device);
//     then create / execute / destroy cycles over every case, same bounds;
//   memory: create plan, allocate, execute, free, destroy rounds of a
//     1024-point transform; growth after each round's free, same bounds.
// `quick` shortens every count (for the test suite); the full counts are the
// reference's.  One line per result, key=value fields; exit 0 when every
// check passed.
#include <hip/hip_runtime_api.h>
#include <malloc.h>

#include <algorithm>
#include <chrono>
#include <cmath>
#include <cstdio>
#include <cstring>
#include <iostream>
#include <numeric>
#include <random>
#include <stdexcept>
#include <string>
#include <vector>

#include "zkx_fft.h"

namespace {

struct Counts {
    int warmup = 5, iterations = 20, batches = 5;
    int stab_block = 100, stab_max_blocks = 64, exec_checkpoints = 10, exec_per_checkpoint = 5000;
    int plan_warmup = 512, plan_checkpoints = 8, plan_per_checkpoint = 512;
    int mem_warmup = 100, mem_rounds = 2000;
};

constexpr long long kCpuNoise = 4096;

void need(hipError_t e, const char* what)
{
    if (e != hipSuccess) throw std::runtime_error(std::string(what) + ": " + hipGetErrorString(e));
}
void need(FFTResult r, const char* what)
{
    if (r != FFT_SUCCESS) throw std::runtime_error(std::string(what) + ": " + fft_error_string(r));
}

std::size_t heap_live() { return mallinfo2().uordblks; }
std::size_t gpu_used()
{
    std::size_t f = 0, t = 0;
    need(hipMemGetInfo(&f, &t), "hipMemGetInfo");
    return t - f;
}

double pct(std::vector<float> v, double q)
{
    std::sort(v.begin(), v.end());
    const std::size_t i = static_cast<std::size_t>(std::ceil(q * static_cast<double>(v.size()))) - 1;
    return v[std::min(i, v.size() - 1)];
}

bool performance(const char* name, int n, int batch, const Counts& c)
{
    const std::size_t elems = static_cast<std::size_t>(n) * batch, bytes = elems * 8;
    std::vector<float> host(2 * elems);
    std::mt19937 gen(static_cast<unsigned>(n * 131 + batch));
    std::uniform_real_distribution<float> uni(-1.f, 1.f);
    for (auto& v : host) v = uni(gen);
    void *in = nullptr, *out = nullptr, *back = nullptr;
    need(hipMalloc(&in, bytes), "hipMalloc");
    need(hipMalloc(&out, bytes), "hipMalloc");
    need(hipMalloc(&back, bytes), "hipMalloc");
    need(hipMemcpy(in, host.data(), bytes, hipMemcpyHostToDevice), "copy input");
    FFTPlanHandle fwd = nullptr, inv = nullptr;
    const auto t0 = std::chrono::steady_clock::now();
    need(fft_plan_create_1d(&fwd, n, batch, FFT_TYPE_C2C, FFT_FORWARD), "plan");
    need(fft_plan_create_1d(&inv, n, batch, FFT_TYPE_C2C, FFT_INVERSE), "plan");
    const double plan_ms = std::chrono::duration<double, std::milli>(std::chrono::steady_clock::now() - t0).count();
    for (int i = 0; i < c.warmup; ++i) need(fft_execute(fwd, out, in), "warmup");
    need(hipDeviceSynchronize(), "sync");
    const std::size_t cpu0 = heap_live(), gpu0 = gpu_used();
    hipEvent_t a, b;
    need(hipEventCreate(&a), "event");
    need(hipEventCreate(&b), "event");
    std::vector<float> samples, means;
    for (int k = 0; k < c.batches; ++k) {
        const std::size_t first = samples.size();
        for (int i = 0; i < c.iterations; ++i) {
            need(hipEventRecord(a, nullptr), "record");
            need(fft_execute(fwd, out, in), "execute");
            need(hipEventRecord(b, nullptr), "record");
            need(hipEventSynchronize(b), "sync");
            float ms = 0.f;
            need(hipEventElapsedTime(&ms, a, b), "elapsed");
            samples.push_back(ms);
        }
        means.push_back(std::accumulate(samples.begin() + first, samples.end(), 0.f) / (samples.size() - first));
    }
    need(hipDeviceSynchronize(), "sync");
    const long long cpu_growth = static_cast<long long>(heap_live()) - static_cast<long long>(cpu0);
    const long long gpu_growth = static_cast<long long>(gpu_used()) - static_cast<long long>(gpu0);
    need(fft_execute(fwd, out, in), "execute");
    need(fft_execute(inv, back, out), "execute");
    std::vector<float> rt(2 * elems);
    need(hipMemcpy(rt.data(), back, bytes, hipMemcpyDeviceToHost), "copy back");
    double e2 = 0, r2 = 0, emax = 0;
    for (std::size_t i = 0; i < elems; ++i) {
        const double dr = static_cast<double>(rt[2 * i]) - host[2 * i];
        const double di = static_cast<double>(rt[2 * i + 1]) - host[2 * i + 1];
        const double e = std::sqrt(dr * dr + di * di);
        e2 += e * e;
        r2 += static_cast<double>(host[2 * i]) * host[2 * i] + static_cast<double>(host[2 * i + 1]) * host[2 * i + 1];
        emax = std::max(emax, e);
    }
    const double rel = std::sqrt(e2 / std::max(r2, 1e-30));
    const bool ok = rel <= 5e-3 && emax <= 5e-2;
    const double pooled = std::accumulate(samples.begin(), samples.end(), 0.0) / samples.size();
    std::sort(means.begin(), means.end());
    std::printf("[FFT-PERF] case=%s n=%d batch=%d center_mean_ms=%.6f pooled_mean_ms=%.6f p50_ms=%.6f p95_ms=%.6f "
                "p99_ms=%.6f min_ms=%.6f plan_create_ms=%.3f cpu_execute_growth_bytes=%lld "
                "gpu_execute_growth_bytes=%lld roundtrip_relative_l2=%.3e roundtrip_max_abs=%.3e correctness=%s\n",
                name, n, batch, means[means.size() / 2], pooled, pct(samples, 0.5), pct(samples, 0.95),
                pct(samples, 0.99), *std::min_element(samples.begin(), samples.end()), plan_ms, cpu_growth,
                gpu_growth, rel, emax, ok ? "pass" : "fail");
    (void)hipEventDestroy(a);
    (void)hipEventDestroy(b);
    need(fft_plan_destroy(fwd), "destroy");
    need(fft_plan_destroy(inv), "destroy");
    need(hipFree(in), "free");
    need(hipFree(out), "free");
    need(hipFree(back), "free");
    return ok;
}

struct StabCase {
    const char* name;
    int n, batch;
    FFTType type;
    FFTDirection dir;
    void *in = nullptr, *out = nullptr;
};

void alloc_case(StabCase& s)
{
    const std::size_t complex_bytes = static_cast<std::size_t>(s.n) * s.batch * 8;
    need(hipMalloc(&s.in, complex_bytes), "hipMalloc");
    need(hipMalloc(&s.out, complex_bytes), "hipMalloc");
    need(hipMemset(s.in, 0, complex_bytes), "memset");
}

bool stability_execute(StabCase& s, const Counts& c)
{
    FFTPlanHandle p = nullptr;
    need(fft_plan_create_1d(&p, s.n, s.batch, s.type, s.dir), "plan");
    std::size_t pg = gpu_used(), pc = heap_live();
    int stable = 0, blocks = 0;
    for (int blk = 1; blk <= c.stab_max_blocks && stable < 3; ++blk) {
        for (int i = 0; i < c.stab_block; ++i) {
            need(fft_execute(p, s.out, s.in), "execute");
            need(hipDeviceSynchronize(), "sync");
        }
        const std::size_t g = gpu_used(), h = heap_live();
        const long long dc = static_cast<long long>(h) - static_cast<long long>(pc);
        stable = (dc >= -kCpuNoise && dc <= kCpuNoise && g == pg) ? stable + 1 : 0;
        pg = g;
        pc = h;
        blocks = blk;
    }
    bool ok = stable >= 3;
    long long cg = 0, gg = 0;
    if (ok) {
        const std::size_t g0 = gpu_used(), h0 = heap_live();
        for (int k = 0; k < c.exec_checkpoints; ++k)
            for (int i = 0; i < c.exec_per_checkpoint; ++i) {
                need(fft_execute(p, s.out, s.in), "execute");
                need(hipDeviceSynchronize(), "sync");
            }
        cg = static_cast<long long>(heap_live()) - static_cast<long long>(h0);
        gg = static_cast<long long>(gpu_used()) - static_cast<long long>(g0);
        ok = cg <= kCpuNoise && gg <= 0;
    }
    std::printf("[FFT-STABILITY][EXECUTE] case=%s stabilization_blocks=%d iterations=%d cpu_growth_bytes=%lld "
                "gpu_growth_bytes=%lld status=%s\n",
                s.name, blocks, c.exec_checkpoints * c.exec_per_checkpoint, cg, gg, ok ? "passed" : "failed");
    need(fft_plan_destroy(p), "destroy");
    return ok;
}

void plan_cycle(std::vector<StabCase>& cases)
{
    for (auto& s : cases) {
        FFTPlanHandle p = nullptr;
        need(fft_plan_create_1d(&p, s.n, s.batch, s.type, s.dir), "plan");
        need(fft_execute(p, s.out, s.in), "execute");
        need(hipDeviceSynchronize(), "sync");
        need(fft_plan_destroy(p), "destroy");
    }
}

bool stability_plans(std::vector<StabCase>& cases, const Counts& c)
{
    for (int i = 0; i < c.plan_warmup; ++i) plan_cycle(cases);
    const std::size_t g0 = gpu_used(), h0 = heap_live();
    for (int k = 0; k < c.plan_checkpoints; ++k)
        for (int i = 0; i < c.plan_per_checkpoint; ++i) plan_cycle(cases);
    const long long cg = static_cast<long long>(heap_live()) - static_cast<long long>(h0);
    const long long gg = static_cast<long long>(gpu_used()) - static_cast<long long>(g0);
    const bool ok = cg <= kCpuNoise && gg <= 0;
    std::printf("[FFT-STABILITY][PLAN] measured_cycles=%d cases_per_cycle=%zu cpu_growth_bytes=%lld "
                "gpu_growth_bytes=%lld status=%s\n",
                c.plan_checkpoints * c.plan_per_checkpoint, cases.size(), cg, gg, ok ? "passed" : "failed");
    return ok;
}

bool memory_lifecycle(const Counts& c)
{
    const int n = 1024;
    std::vector<float> host(2 * n, 0.25f);
    auto round = [&]() {
        FFTPlanHandle p = nullptr;
        need(fft_plan_create_1d(&p, n, 1, FFT_TYPE_C2C, FFT_FORWARD), "plan");
        void *in = nullptr, *out = nullptr;
        need(hipMalloc(&in, n * 8), "hipMalloc");
        need(hipMalloc(&out, n * 8), "hipMalloc");
        need(hipMemcpy(in, host.data(), n * 8, hipMemcpyHostToDevice), "copy");
        need(fft_execute(p, out, in), "execute");
        need(hipDeviceSynchronize(), "sync");
        need(hipFree(in), "free");
        need(hipFree(out), "free");
        need(fft_plan_destroy(p), "destroy");
    };
    for (int i = 0; i < c.mem_warmup; ++i) round();
    const std::size_t g0 = gpu_used(), h0 = heap_live();
    long long worst_cpu = 0, worst_gpu = 0;
    for (int i = 0; i < c.mem_rounds; ++i) {
        round();
        worst_cpu = std::max(worst_cpu, static_cast<long long>(heap_live()) - static_cast<long long>(h0));
        worst_gpu = std::max(worst_gpu, static_cast<long long>(gpu_used()) - static_cast<long long>(g0));
    }
    const bool ok = worst_cpu <= kCpuNoise && worst_gpu <= 0;
    std::printf("[FFT-MEMORY] n=%d rounds=%d max_cpu_growth_after_free_bytes=%lld max_gpu_growth_after_free_bytes=%lld "
                "status=%s\n",
                n, c.mem_rounds, worst_cpu, worst_gpu, ok ? "passed" : "failed");
    return ok;
}

}  // namespace

int main(int argc, char** argv)
{
    Counts c;
    if (argc > 1 && std::strcmp(argv[1], "quick") == 0) {
        c.stab_block = 20;
        c.exec_checkpoints = 4;
        c.exec_per_checkpoint = 100;
        c.plan_warmup = 16;
        c.plan_checkpoints = 4;
        c.plan_per_checkpoint = 16;
        c.mem_warmup = 20;
        c.mem_rounds = 100;
    }
    try {
        bool ok = true;
        const struct {
            const char* name;
            int n, batch;
        } perf[] = {{"task1_doppler", 32, 256},       {"task2_spectrogram", 64, 15}, {"task1_ambiguity_2d", 128, 127},
                    {"task1_pulse_compression", 256, 32}, {"non_power_127", 127, 1}, {"throughput_1024", 1024, 16},
                    {"large_4096", 4096, 1},          {"dwell_4096x128", 4096, 128}};
        for (const auto& p : perf) ok = performance(p.name, p.n, p.batch, c) && ok;
        std::vector<StabCase> cases{{"c2c_power2_8x1", 8, 1, FFT_TYPE_C2C, FFT_FORWARD},
                                    {"c2c_power2_64x15", 64, 15, FFT_TYPE_C2C, FFT_FORWARD},
                                    {"c2c_non_power_127x1", 127, 1, FFT_TYPE_C2C, FFT_FORWARD},
                                    {"c2c_power2_1024x16", 1024, 16, FFT_TYPE_C2C, FFT_FORWARD},
                                    {"r2c_power2_64x15", 64, 15, FFT_TYPE_R2C, FFT_FORWARD},
                                    {"c2r_power2_64x15", 64, 15, FFT_TYPE_C2R, FFT_INVERSE}};
        for (auto& s : cases) alloc_case(s);
        for (auto& s : cases) ok = stability_execute(s, c) && ok;
        ok = stability_plans(cases, c) && ok;
        for (auto& s : cases) {
            need(hipFree(s.in), "free");
            need(hipFree(s.out), "free");
        }
        ok = memory_lifecycle(c) && ok;
        std::printf("[FFT-PROTOCOLS][SUMMARY] status=%s\n", ok ? "passed" : "failed");
        return ok ? 0 : 1;
    } catch (const std::exception& e) {
        std::fprintf(stderr, "[FFT-PROTOCOLS][FAIL] %s\n", e.what());
        return 2;
    }
}
