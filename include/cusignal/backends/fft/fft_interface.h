// The FFT backend interface the reference's operators are written against
// (ZKX/cusignal_cpp/include/cusignal/backends/fft/fft_interface.h): the fft_*
// C ABI (include/zkx_fft.h), the ComplexFloat / ComplexDouble element types,
// laid out as {re, im}, and for C++ callers the FFTInterface /
// FFTInterface2D objects (fft_interface.h:499-558) over that ABI.
//
// FFTInterface(n) owns a forward and an inverse length-n plan and one pair of
// device staging buffers; its host calls (fft, ifft, rfft, irfft) copy
// through them synchronously.  Objects built BatchOnly own nothing: their
// device calls draw plans from a per-thread pool of at most 8 plans per call
// site, keyed on (n, batch, direction) for 1-D and (nx, ny, direction) for
// 2-D, filled in order and then replaced round-robin
// (fft_interface.cpp:111-135,162-184), so short-lived interface objects in a
// loop do not build plans per call.  The host batch calls keep one plan and
// staging pair per thread and call site, rebuilt when (n, batch) changes
// (fft_interface.cpp:33-80).  All work runs on the default stream; host calls
// return after their results are on the host, device calls after the launch.
// Shorter host inputs are zero padded and longer ones truncated to the
// transform length.  Invalid sizes throw std::invalid_argument; a failed
// plan, copy or transform throws std::runtime_error naming the call.
#pragma once

#include "../../../zkx_fft.h"

#ifdef __cplusplus
extern "C" {
#endif
typedef struct {
    float re;
    float im;
} ComplexFloat;
typedef struct {
    double re;
    double im;
} ComplexDouble;
#ifdef __cplusplus
}
#endif

#ifdef __cplusplus
static_assert(sizeof(ComplexFloat) == 8, "ComplexFloat layout");
static_assert(sizeof(ComplexDouble) == 16, "ComplexDouble layout");

#include <array>
#include <complex>
#include <cstddef>
#include <stdexcept>
#include <string>
#include <type_traits>
#include <vector>

#include "../../runtime/device_array.h"

namespace cusignal {

using complexd = std::complex<double>;

// smallest power of two >= n (1 for n == 0)
inline std::size_t fft_next_power_of_two(std::size_t n)
{
    std::size_t p = 1;
    while (p < n) p <<= 1;
    return p;
}

// input zero padded (or truncated) to `padded_size`, by default the next
// power of two of its length
inline std::vector<complexd> fft_prepare_complex_input(const std::vector<complexd>& input,
                                                       std::size_t padded_size = 0)
{
    if (padded_size == 0) padded_size = fft_next_power_of_two(input.size());
    std::vector<complexd> out(padded_size);
    const std::size_t m = input.size() < padded_size ? input.size() : padded_size;
    for (std::size_t i = 0; i < m; ++i) out[i] = input[i];
    return out;
}

// real samples as complex, zero padded (or truncated) as above
template <typename T>
std::vector<complexd> fft_promote_real_to_complex(const std::vector<T>& input, std::size_t padded_size = 0)
{
    static_assert(std::is_arithmetic<T>::value, "FFT real input promotion requires an arithmetic type");
    if (padded_size == 0) padded_size = fft_next_power_of_two(input.size());
    std::vector<complexd> out(padded_size);
    const std::size_t m = input.size() < padded_size ? input.size() : padded_size;
    for (std::size_t i = 0; i < m; ++i) out[i] = complexd(static_cast<double>(input[i]), 0.0);
    return out;
}

namespace fft_detail {

inline void check_copy(hipError_t e, const char* where)
{
    if (e != hipSuccess) throw std::runtime_error(std::string(where) + ": " + hipGetErrorString(e));
}

inline void check_exec(FFTResult r, const char* where)
{
    if (r != FFT_SUCCESS) throw std::runtime_error(std::string(where) + ": fft_execute failed (" +
                                                   fft_error_string(r) + ")");
}

// A plan of one shape; destroyed with the slot.
struct PlanSlot {
    int a = 0, b = 0;  // (n, batch) for 1-D, (columns, rows) for 2-D
    FFTDirection dir = FFT_FORWARD;
    FFTPlanHandle plan = nullptr;

    PlanSlot() = default;
    PlanSlot(const PlanSlot&) = delete;
    PlanSlot& operator=(const PlanSlot&) = delete;
    ~PlanSlot() { clear(); }
    void clear()
    {
        if (plan) (void)fft_plan_destroy(plan);
        plan = nullptr;
        a = b = 0;
        dir = FFT_FORWARD;
    }
};

// Fixed-capacity plan pool: a hit returns the cached plan; a miss fills the
// next free slot, or once all are used replaces slots in round-robin order.
template <bool TwoD>
class PlanPool {
public:
    static constexpr std::size_t capacity = 8;

    FFTPlanHandle get(int a, int b, FFTDirection dir)
    {
        for (std::size_t i = 0; i < used_; ++i)
            if (slots_[i].plan && slots_[i].a == a && slots_[i].b == b && slots_[i].dir == dir)
                return slots_[i].plan;
        std::size_t i;
        if (used_ < capacity) {
            i = used_++;
        } else {
            i = victim_;
            victim_ = (victim_ + 1) % capacity;
        }
        PlanSlot& s = slots_[i];
        s.clear();
        FFTPlanHandle p = nullptr;
        // 2-D keys are (columns, rows), the ABI's (nx, ny)
        const FFTResult r = TwoD ? fft_plan_create_2d(&p, a, b, FFT_TYPE_C2C, dir)
                                 : fft_plan_create_1d(&p, a, b, FFT_TYPE_C2C, dir);
        if (r != FFT_SUCCESS)
            throw std::runtime_error(std::string(TwoD ? "FFT2DPlanCache: fft_plan_create_2d failed"
                                                      : "BatchFFTPlanCache: fft_plan_create_1d failed") +
                                     " (" + fft_error_string(r) + ")");
        s.plan = p;
        s.a = a;
        s.b = b;
        s.dir = dir;
        return p;
    }

private:
    std::array<PlanSlot, capacity> slots_{};
    std::size_t used_ = 0;
    std::size_t victim_ = 0;
};

// One batched plan with its device staging pair, for the host batch calls.
class StagedBatch {
public:
    StagedBatch() = default;
    StagedBatch(const StagedBatch&) = delete;
    StagedBatch& operator=(const StagedBatch&) = delete;
    ~StagedBatch() { clear(); }

    void ensure(int n, int batch, FFTDirection dir)
    {
        if (plan_ && n_ == n && batch_ == batch && dir_ == dir) return;
        clear();
        if (fft_plan_create_1d(&plan_, n, batch, FFT_TYPE_C2C, dir) != FFT_SUCCESS) {
            plan_ = nullptr;
            throw std::runtime_error("BatchFFTCache: fft_plan_create_1d failed");
        }
        const std::size_t bytes = static_cast<std::size_t>(n) * static_cast<std::size_t>(batch) * sizeof(ComplexFloat);
        if (cuda_utils::tracked_device_malloc(&in_, bytes) != hipSuccess ||
            cuda_utils::tracked_device_malloc(&out_, bytes) != hipSuccess) {
            clear();
            throw std::runtime_error("BatchFFTCache: device buffer allocation failed");
        }
        n_ = n;
        batch_ = batch;
        dir_ = dir;
    }
    // host rows -> device -> transform -> host
    void run(const ComplexFloat* host_in, ComplexFloat* host_out, const char* where)
    {
        const std::size_t bytes = static_cast<std::size_t>(n_) * static_cast<std::size_t>(batch_) * sizeof(ComplexFloat);
        check_copy(hipMemcpy(in_, host_in, bytes, hipMemcpyHostToDevice), where);
        check_exec(fft_execute(plan_, out_, in_), where);
        check_copy(hipMemcpy(host_out, out_, bytes, hipMemcpyDeviceToHost), where);
    }

private:
    void clear()
    {
        if (plan_) (void)fft_plan_destroy(plan_);
        if (in_) (void)cuda_utils::tracked_device_free(in_);
        if (out_) (void)cuda_utils::tracked_device_free(out_);
        plan_ = nullptr;
        in_ = out_ = nullptr;
        n_ = batch_ = 0;
    }
    FFTPlanHandle plan_ = nullptr;
    void* in_ = nullptr;
    void* out_ = nullptr;
    int n_ = 0, batch_ = 0;
    FFTDirection dir_ = FFT_FORWARD;
};

inline ComplexFloat to_cf(const complexd& v)
{
    return {static_cast<float>(v.real()), static_cast<float>(v.imag())};
}
inline complexd to_cd(const ComplexFloat& v) { return complexd(v.re, v.im); }

// rows of complexd, each zero padded / truncated to n, packed as ComplexFloat
inline std::vector<ComplexFloat> pack_rows(const std::vector<std::vector<complexd>>& rows, std::size_t n)
{
    std::vector<ComplexFloat> out(rows.size() * n, ComplexFloat{0.f, 0.f});
    for (std::size_t r = 0; r < rows.size(); ++r) {
        const std::size_t m = rows[r].size() < n ? rows[r].size() : n;
        for (std::size_t i = 0; i < m; ++i) out[r * n + i] = to_cf(rows[r][i]);
    }
    return out;
}

inline std::vector<std::vector<complexd>> unpack_rows(const std::vector<ComplexFloat>& flat, std::size_t rows,
                                                      std::size_t n)
{
    std::vector<std::vector<complexd>> out(rows, std::vector<complexd>(n));
    for (std::size_t r = 0; r < rows; ++r)
        for (std::size_t i = 0; i < n; ++i) out[r][i] = to_cd(flat[r * n + i]);
    return out;
}

}  // namespace fft_detail

class FFTInterface {
public:
    struct BatchOnly {};

    explicit FFTInterface(int n) : FFTInterface(n, false) {}
    FFTInterface(int n, BatchOnly) : FFTInterface(n, true) {}
    ~FFTInterface() noexcept { release(); }
    FFTInterface(const FFTInterface&) = delete;
    FFTInterface& operator=(const FFTInterface&) = delete;
    FFTInterface(FFTInterface&&) = delete;
    FFTInterface& operator=(FFTInterface&&) = delete;

    // host round trips through the object's own plans (length n)
    std::vector<complexd> fft(const std::vector<complexd>& input) { return host_1d(input, plan_fwd_, "FFTInterface::fft"); }
    std::vector<complexd> ifft(const std::vector<complexd>& input) { return host_1d(input, plan_inv_, "FFTInterface::ifft"); }

    // device transforms of exactly n elements
    void fft_device(const DeviceArray<ComplexFloat>& input, DeviceArray<ComplexFloat>& output)
    {
        single(input.data(), output.data(), input.size(), output.size(), plan_fwd_, "FFTInterface::fft_device");
    }
    void ifft_device(const DeviceArray<ComplexFloat>& input, DeviceArray<ComplexFloat>& output)
    {
        single(input.data(), output.data(), input.size(), output.size(), plan_inv_, "FFTInterface::ifft_device");
    }
    void fft_device(const ComplexFloat* input, ComplexFloat* output, std::size_t count)
    {
        single(input, output, count, count, plan_fwd_, "FFTInterface::fft_device");
    }
    void ifft_device(const ComplexFloat* input, ComplexFloat* output, std::size_t count)
    {
        single(input, output, count, count, plan_inv_, "FFTInterface::ifft_device");
    }
    void fft_inplace_device(DeviceArray<ComplexFloat>& data) { inplace(data, plan_fwd_, "FFTInterface::fft_inplace_device"); }
    void ifft_inplace_device(DeviceArray<ComplexFloat>& data) { inplace(data, plan_inv_, "FFTInterface::ifft_inplace_device"); }

    // device transforms of `batch` rows of n, plans from the per-thread pool
    void fft_batch_device(const DeviceArray<ComplexFloat>& input, DeviceArray<ComplexFloat>& output, int batch)
    {
        static thread_local fft_detail::PlanPool<false> pool;
        batched(pool, input.data(), output.data(), batch, input.size(), output.size(), FFT_FORWARD,
                "FFTInterface::fft_batch_device");
    }
    void ifft_batch_device(const DeviceArray<ComplexFloat>& input, DeviceArray<ComplexFloat>& output, int batch)
    {
        static thread_local fft_detail::PlanPool<false> pool;
        batched(pool, input.data(), output.data(), batch, input.size(), output.size(), FFT_INVERSE,
                "FFTInterface::ifft_batch_device");
    }
    void fft_batch_device(const ComplexFloat* input, ComplexFloat* output, int batch, std::size_t count)
    {
        static thread_local fft_detail::PlanPool<false> pool;
        batched(pool, input, output, batch, count, count, FFT_FORWARD, "FFTInterface::fft_batch_device");
    }
    void ifft_batch_device(const ComplexFloat* input, ComplexFloat* output, int batch, std::size_t count)
    {
        static thread_local fft_detail::PlanPool<false> pool;
        batched(pool, input, output, batch, count, count, FFT_INVERSE, "FFTInterface::ifft_batch_device");
    }

    // host batches: one row per vector, each zero padded / truncated to n
    std::vector<std::vector<complexd>> fft_batch(const std::vector<std::vector<complexd>>& input)
    {
        static thread_local fft_detail::StagedBatch staged;
        return host_batch(staged, input, FFT_FORWARD, "FFTInterface::fft_batch");
    }
    std::vector<std::vector<complexd>> ifft_batch(const std::vector<std::vector<complexd>>& input)
    {
        static thread_local fft_detail::StagedBatch staged;
        return host_batch(staged, input, FFT_INVERSE, "FFTInterface::ifft_batch");
    }
    // `batch` inverse rows of n from a flat array (zero padded / truncated to
    // n * batch); batch <= 0 gives an empty result
    std::vector<ComplexFloat> ifft_batch_float(const std::vector<ComplexFloat>& input, int batch)
    {
        if (batch <= 0) return {};
        static thread_local fft_detail::StagedBatch staged;
        staged.ensure(n_, batch, FFT_INVERSE);
        const std::size_t total = static_cast<std::size_t>(n_) * static_cast<std::size_t>(batch);
        std::vector<ComplexFloat> in(total, ComplexFloat{0.f, 0.f}), out(total);
        const std::size_t m = input.size() < total ? input.size() : total;
        for (std::size_t i = 0; i < m; ++i) in[i] = input[i];
        staged.run(in.data(), out.data(), "FFTInterface::ifft_batch_float");
        return out;
    }

    // first len(input)/2 + 1 bins of the length-n transform of `input`
    std::vector<complexd> rfft(const std::vector<complexd>& input)
    {
        const std::vector<complexd> full = fft(input);
        const std::size_t keep = input.size() / 2 + 1;
        std::vector<complexd> out(keep);
        for (std::size_t i = 0; i < keep && i < full.size(); ++i) out[i] = full[i];
        return out;
    }
    // real signal of output_len samples (default 2 * (len(input) - 1)) whose
    // half spectrum is `input`: the missing upper bins are the conjugates of
    // bins 1 .. len-2, then a length-output_len inverse; imaginary parts are
    // dropped
    std::vector<complexd> irfft(const std::vector<complexd>& input, int output_len = 0)
    {
        const int m = static_cast<int>(input.size());
        if (output_len <= 0) output_len = 2 * (m - 1);
        if (output_len <= 0) throw std::invalid_argument("FFTInterface::irfft: output length must be positive");
        std::vector<complexd> spec(output_len);
        for (int i = 0; i < m && i < output_len; ++i) spec[i] = input[i];
        for (int i = 1; i < m - 1 && i < output_len; ++i) {
            const int k = output_len - i;
            if (k >= m) spec[k] = std::conj(input[i]);
        }
        FFTInterface inv(output_len);
        std::vector<complexd> y = inv.ifft(spec);
        for (auto& v : y) v = complexd(v.real(), 0.0);
        return y;
    }

private:
    FFTInterface(int n, bool batch_only) : n_(n)
    {
        if (n_ <= 0) throw std::invalid_argument("FFTInterface: length must be positive");
        if (batch_only) return;
        try {
            if (fft_plan_create_1d(&plan_fwd_, n_, 1, FFT_TYPE_C2C, FFT_FORWARD) != FFT_SUCCESS) {
                plan_fwd_ = nullptr;
                throw std::runtime_error("FFTInterface: forward plan creation failed");
            }
            if (fft_plan_create_1d(&plan_inv_, n_, 1, FFT_TYPE_C2C, FFT_INVERSE) != FFT_SUCCESS) {
                plan_inv_ = nullptr;
                throw std::runtime_error("FFTInterface: inverse plan creation failed");
            }
            const std::size_t bytes = static_cast<std::size_t>(n_) * sizeof(ComplexFloat);
            cuda_utils::check_hip(cuda_utils::tracked_device_malloc(&d_in_, bytes), "FFTInterface: hipMalloc");
            cuda_utils::check_hip(cuda_utils::tracked_device_malloc(&d_out_, bytes), "FFTInterface: hipMalloc");
        } catch (...) {
            release();
            throw;
        }
    }

    void release() noexcept
    {
        if (plan_fwd_) (void)fft_plan_destroy(plan_fwd_);
        if (plan_inv_) (void)fft_plan_destroy(plan_inv_);
        if (d_in_) (void)cuda_utils::tracked_device_free(d_in_);
        if (d_out_) (void)cuda_utils::tracked_device_free(d_out_);
        plan_fwd_ = plan_inv_ = nullptr;
        d_in_ = d_out_ = nullptr;
    }

    std::vector<complexd> host_1d(const std::vector<complexd>& input, FFTPlanHandle plan, const char* where)
    {
        if (!plan) throw std::runtime_error(std::string(where) + ": object was built BatchOnly");
        const std::size_t n = static_cast<std::size_t>(n_);
        std::vector<ComplexFloat> in(n, ComplexFloat{0.f, 0.f}), out(n);
        const std::size_t m = input.size() < n ? input.size() : n;
        for (std::size_t i = 0; i < m; ++i) in[i] = fft_detail::to_cf(input[i]);
        fft_detail::check_copy(hipMemcpy(d_in_, in.data(), n * sizeof(ComplexFloat), hipMemcpyHostToDevice), where);
        fft_detail::check_exec(fft_execute(plan, d_out_, d_in_), where);
        fft_detail::check_copy(hipMemcpy(out.data(), d_out_, n * sizeof(ComplexFloat), hipMemcpyDeviceToHost), where);
        std::vector<complexd> y(n);
        for (std::size_t i = 0; i < n; ++i) y[i] = fft_detail::to_cd(out[i]);
        return y;
    }

    void single(const ComplexFloat* in, ComplexFloat* out, std::size_t n_in, std::size_t n_out, FFTPlanHandle plan,
                const char* where)
    {
        if (n_in != static_cast<std::size_t>(n_) || n_out != static_cast<std::size_t>(n_))
            throw std::invalid_argument(std::string(where) + ": input/output sizes must match FFT length");
        fft_detail::check_exec(fft_execute(plan, out, in), where);
    }

    void inplace(DeviceArray<ComplexFloat>& data, FFTPlanHandle plan, const char* where)
    {
        if (data.size() != static_cast<std::size_t>(n_))
            throw std::invalid_argument(std::string(where) + ": data size must match FFT length");
        const FFTResult r = fft_execute_inplace(plan, data.data());
        if (r != FFT_SUCCESS)
            throw std::runtime_error(std::string(where) + ": fft_execute_inplace failed (" + fft_error_string(r) + ")");
    }

    void batched(fft_detail::PlanPool<false>& pool, const ComplexFloat* in, ComplexFloat* out, int batch,
                 std::size_t n_in, std::size_t n_out, FFTDirection dir, const char* where)
    {
        if (batch <= 0) throw std::invalid_argument(std::string(where) + ": batch must be positive");
        const std::size_t expected = static_cast<std::size_t>(n_) * static_cast<std::size_t>(batch);
        if (n_in != expected || n_out != expected)
            throw std::invalid_argument(std::string(where) + ": input/output sizes must match n*batch");
        fft_detail::check_exec(fft_execute(pool.get(n_, batch, dir), out, in), where);
    }

    std::vector<std::vector<complexd>> host_batch(fft_detail::StagedBatch& staged,
                                                  const std::vector<std::vector<complexd>>& input, FFTDirection dir,
                                                  const char* where)
    {
        if (input.empty()) return {};
        const int batch = static_cast<int>(input.size());
        staged.ensure(n_, batch, dir);
        const std::size_t n = static_cast<std::size_t>(n_);
        std::vector<ComplexFloat> in = fft_detail::pack_rows(input, n), out(in.size());
        staged.run(in.data(), out.data(), where);
        return fft_detail::unpack_rows(out, input.size(), n);
    }

    int n_;
    FFTPlanHandle plan_fwd_ = nullptr;
    FFTPlanHandle plan_inv_ = nullptr;
    void* d_in_ = nullptr;
    void* d_out_ = nullptr;
};

// nx x ny complex transforms: nx rows of ny columns, row-major
// (element [i][j] at i * ny + j), as the reference's 2-D interface lays them
// out (fft_interface.cpp:678-709).
class FFTInterface2D {
public:
    FFTInterface2D(int nx, int ny) : FFTInterface2D(nx, ny, false) {}
    FFTInterface2D(int nx, int ny, FFTInterface::BatchOnly) : FFTInterface2D(nx, ny, true) {}
    ~FFTInterface2D() noexcept { release(); }
    FFTInterface2D(const FFTInterface2D&) = delete;
    FFTInterface2D& operator=(const FFTInterface2D&) = delete;
    FFTInterface2D(FFTInterface2D&&) = delete;
    FFTInterface2D& operator=(FFTInterface2D&&) = delete;

    // host matrices, zero padded / truncated to nx rows of ny
    std::vector<std::vector<complexd>> fft2d(const std::vector<std::vector<complexd>>& input)
    {
        return host_2d(input, plan_fwd_, "FFTInterface2D::fft2d");
    }
    std::vector<std::vector<complexd>> ifft2d(const std::vector<std::vector<complexd>>& input)
    {
        return host_2d(input, plan_inv_, "FFTInterface2D::ifft2d");
    }

    // device nx*ny arrays; a BatchOnly object draws plans from the pool
    void fft2d_device(const DeviceArray<ComplexFloat>& input, DeviceArray<ComplexFloat>& output)
    {
        static thread_local fft_detail::PlanPool<true> pool;
        device_2d(pool, input, output, plan_fwd_, FFT_FORWARD, "FFTInterface2D::fft2d_device");
    }
    void ifft2d_device(const DeviceArray<ComplexFloat>& input, DeviceArray<ComplexFloat>& output)
    {
        static thread_local fft_detail::PlanPool<true> pool;
        device_2d(pool, input, output, plan_inv_, FFT_INVERSE, "FFTInterface2D::ifft2d_device");
    }

private:
    FFTInterface2D(int nx, int ny, bool device_only) : nx_(nx), ny_(ny)
    {
        if (nx_ <= 0 || ny_ <= 0) throw std::invalid_argument("FFTInterface2D: dimensions must be positive");
        if (device_only) return;
        try {
            // the ABI's 2-D plan is ny rows of nx columns: rows = nx_, columns = ny_
            if (fft_plan_create_2d(&plan_fwd_, ny_, nx_, FFT_TYPE_C2C, FFT_FORWARD) != FFT_SUCCESS) {
                plan_fwd_ = nullptr;
                throw std::runtime_error("FFTInterface2D: forward plan creation failed");
            }
            if (fft_plan_create_2d(&plan_inv_, ny_, nx_, FFT_TYPE_C2C, FFT_INVERSE) != FFT_SUCCESS) {
                plan_inv_ = nullptr;
                throw std::runtime_error("FFTInterface2D: inverse plan creation failed");
            }
            const std::size_t bytes = count() * sizeof(ComplexFloat);
            cuda_utils::check_hip(cuda_utils::tracked_device_malloc(&d_in_, bytes), "FFTInterface2D: hipMalloc");
            cuda_utils::check_hip(cuda_utils::tracked_device_malloc(&d_out_, bytes), "FFTInterface2D: hipMalloc");
        } catch (...) {
            release();
            throw;
        }
    }

    std::size_t count() const { return static_cast<std::size_t>(nx_) * static_cast<std::size_t>(ny_); }

    void release() noexcept
    {
        if (plan_fwd_) (void)fft_plan_destroy(plan_fwd_);
        if (plan_inv_) (void)fft_plan_destroy(plan_inv_);
        if (d_in_) (void)cuda_utils::tracked_device_free(d_in_);
        if (d_out_) (void)cuda_utils::tracked_device_free(d_out_);
        plan_fwd_ = plan_inv_ = nullptr;
        d_in_ = d_out_ = nullptr;
    }

    std::vector<std::vector<complexd>> host_2d(const std::vector<std::vector<complexd>>& input, FFTPlanHandle plan,
                                               const char* where)
    {
        if (!plan) throw std::runtime_error(std::string(where) + ": object was built BatchOnly");
        const std::size_t rows = static_cast<std::size_t>(nx_), cols = static_cast<std::size_t>(ny_);
        std::vector<std::vector<complexd>> padded(rows);
        for (std::size_t i = 0; i < rows && i < input.size(); ++i) padded[i] = input[i];
        std::vector<ComplexFloat> in = fft_detail::pack_rows(padded, cols), out(in.size());
        fft_detail::check_copy(hipMemcpy(d_in_, in.data(), in.size() * sizeof(ComplexFloat), hipMemcpyHostToDevice),
                               where);
        fft_detail::check_exec(fft_execute(plan, d_out_, d_in_), where);
        fft_detail::check_copy(hipMemcpy(out.data(), d_out_, out.size() * sizeof(ComplexFloat), hipMemcpyDeviceToHost),
                               where);
        return fft_detail::unpack_rows(out, rows, cols);
    }

    void device_2d(fft_detail::PlanPool<true>& pool, const DeviceArray<ComplexFloat>& input,
                   DeviceArray<ComplexFloat>& output, FFTPlanHandle own, FFTDirection dir, const char* where)
    {
        if (input.size() != count() || output.size() != count())
            throw std::invalid_argument(std::string(where) + ": input/output sizes must match nx*ny");
        // pool keyed on (columns, rows) = (ny_, nx_)
        const FFTPlanHandle plan = own ? own : pool.get(ny_, nx_, dir);
        fft_detail::check_exec(fft_execute(plan, output.data(), input.data()), where);
    }

    int nx_, ny_;
    FFTPlanHandle plan_fwd_ = nullptr;
    FFTPlanHandle plan_inv_ = nullptr;
    void* d_in_ = nullptr;
    void* d_out_ = nullptr;
};

}  // namespace cusignal
#endif  // __cplusplus
