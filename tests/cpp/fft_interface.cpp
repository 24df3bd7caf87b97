// FFTInterface / FFTInterface2D (fft_interface.h:499-558 of the reference)
// driven the way the reference's operators use them, built with g++ and
// linked to libzkx_radar.so.
//
// usage: fft_interface IN OUT
//   IN : int32 n, batch (>= 10), nx, ny; c64 x[batch*n]; c64 m[nx*ny]
//   OUT: blocks, each an int64 element count then the data:
//     c128 fft(x0), c128 ifft(x0), c64 fft_device(x0), c64 fft_device raw(x0),
//     c64 ifft_inplace_device(x0), c64 fft_batch_device(x), c64 ifft_batch_device
//     raw(x), c128 fft_batch(x with row 1 cut to n/2), c64 ifft_batch_float(x
//     without its last 3 values), c128 rfft(re x0), c128 irfft(rfft(re x0)),
//     c128 fft2d(m), c128 ifft2d(m), c64 fft2d_device BatchOnly(m), c64
//     ifft2d_device(m), c64 pool sweep (fft_batch_device of the first b rows
//     for b = 1..10, then b = 1 again), i32 error-behaviour flags, i64 tracked
//     bytes (before, with an FFTInterface(n) alive, after it, with a
//     DeviceArray of n alive)
#include <complex>
#include <fstream>
#include <functional>
#include <iostream>
#include <vector>

#include "cusignal/backends/fft/fft_interface.h"

using cusignal::complexd;
using cusignal::DeviceArray;
using cusignal::FFTInterface;
using cusignal::FFTInterface2D;

namespace {

template <class T>
void read_into(std::ifstream& in, T* dst, std::size_t count)
{
    in.read(reinterpret_cast<char*>(dst), static_cast<std::streamsize>(sizeof(T) * count));
    if (!in) throw std::runtime_error("short input file");
}

template <class T>
void write_block(std::ofstream& out, const std::vector<T>& v)
{
    const long long n = static_cast<long long>(v.size());
    out.write(reinterpret_cast<const char*>(&n), sizeof(n));
    out.write(reinterpret_cast<const char*>(v.data()), static_cast<std::streamsize>(sizeof(T) * v.size()));
}

std::vector<complexd> flat(const std::vector<std::vector<complexd>>& rows)
{
    std::vector<complexd> out;
    for (const auto& r : rows) out.insert(out.end(), r.begin(), r.end());
    return out;
}

// 1 when f throws an exception of type E, 0 otherwise
template <class E>
int throws(const std::function<void()>& f)
{
    try {
        f();
    } catch (const E&) {
        return 1;
    } catch (...) {
        return 0;
    }
    return 0;
}

}  // namespace

int main(int argc, char** argv)
{
    if (argc != 3) {
        std::cerr << "usage: fft_interface IN OUT\n";
        return 2;
    }
    try {
        std::ifstream in(argv[1], std::ios::binary);
        int hdr[4];
        read_into(in, hdr, 4);
        const int n = hdr[0], batch = hdr[1], nx = hdr[2], ny = hdr[3];
        if (batch < 10) throw std::runtime_error("batch must be >= 10");
        std::vector<ComplexFloat> x(static_cast<std::size_t>(batch) * n), m(static_cast<std::size_t>(nx) * ny);
        read_into(in, x.data(), x.size());
        read_into(in, m.data(), m.size());
        std::ofstream out(argv[2], std::ios::binary);

        std::vector<complexd> x0(n), x0_real(n);
        std::vector<std::vector<complexd>> rows(batch, std::vector<complexd>(n));
        for (int b = 0; b < batch; ++b)
            for (int i = 0; i < n; ++i) rows[b][i] = complexd(x[b * n + i].re, x[b * n + i].im);
        x0 = rows[0];
        for (int i = 0; i < n; ++i) x0_real[i] = complexd(x0[i].real(), 0.0);

        // object with its own plans: host and device single transforms
        FFTInterface f(n);
        write_block(out, f.fft(x0));
        write_block(out, f.ifft(x0));
        const std::vector<ComplexFloat> x0f(x.begin(), x.begin() + n);
        auto d_x0 = DeviceArray<ComplexFloat>::from_host(x0f);
        DeviceArray<ComplexFloat> d_y(n);
        f.fft_device(d_x0, d_y);
        write_block(out, d_y.to_host());
        DeviceArray<ComplexFloat> d_y2(n);
        f.fft_device(d_x0.data(), d_y2.data(), static_cast<std::size_t>(n));
        write_block(out, d_y2.to_host());
        auto d_io = DeviceArray<ComplexFloat>::from_host(x0f);
        f.ifft_inplace_device(d_io);
        write_block(out, d_io.to_host());

        // batch-only object: pooled plans for the device batch calls
        FFTInterface fb(n, FFTInterface::BatchOnly{});
        auto d_x = DeviceArray<ComplexFloat>::from_host(x);
        DeviceArray<ComplexFloat> d_yb(x.size());
        fb.fft_batch_device(d_x, d_yb, batch);
        write_block(out, d_yb.to_host());
        DeviceArray<ComplexFloat> d_yb2(x.size());
        fb.ifft_batch_device(d_x.data(), d_yb2.data(), batch, x.size());
        write_block(out, d_yb2.to_host());

        // host batches with padding
        std::vector<std::vector<complexd>> cut = rows;
        cut[1].resize(n / 2);
        write_block(out, flat(f.fft_batch(cut)));
        std::vector<ComplexFloat> xs(x.begin(), x.end() - 3);
        write_block(out, f.ifft_batch_float(xs, batch));

        // real transforms
        const std::vector<complexd> half = f.rfft(x0_real);
        write_block(out, half);
        write_block(out, f.irfft(half));

        // 2-D
        std::vector<std::vector<complexd>> mat(nx, std::vector<complexd>(ny));
        for (int i = 0; i < nx; ++i)
            for (int j = 0; j < ny; ++j) mat[i][j] = complexd(m[i * ny + j].re, m[i * ny + j].im);
        FFTInterface2D f2(nx, ny);
        write_block(out, flat(f2.fft2d(mat)));
        write_block(out, flat(f2.ifft2d(mat)));
        FFTInterface2D f2b(nx, ny, FFTInterface::BatchOnly{});
        auto d_m = DeviceArray<ComplexFloat>::from_host(m);
        DeviceArray<ComplexFloat> d_m2(m.size());
        f2b.fft2d_device(d_m, d_m2);
        write_block(out, d_m2.to_host());
        f2.ifft2d_device(d_m, d_m2);
        write_block(out, d_m2.to_host());

        // pool sweep: ten (n, b) keys through a capacity-8 pool, then the
        // evicted first key again
        std::vector<ComplexFloat> sweep;
        for (int b : {1, 2, 3, 4, 5, 6, 7, 8, 9, 10, 1}) {
            FFTInterface tmp(n, FFTInterface::BatchOnly{});
            DeviceArray<ComplexFloat> d_out(static_cast<std::size_t>(b) * n);
            tmp.fft_batch_device(d_x.data(), d_out.data(), b, static_cast<std::size_t>(b) * n);
            const auto h = d_out.to_host();
            sweep.insert(sweep.end(), h.begin(), h.end());
        }
        write_block(out, sweep);

        // error behaviour
        std::vector<int> flags;
        flags.push_back(throws<std::invalid_argument>([] { FFTInterface bad(0); }));
        flags.push_back(throws<std::invalid_argument>([&] {
            DeviceArray<ComplexFloat> small(n - 1);
            f.fft_device(small, d_y);
        }));
        flags.push_back(throws<std::invalid_argument>([&] { fb.fft_batch_device(d_x, d_yb, 0); }));
        flags.push_back(throws<std::invalid_argument>([&] { fb.fft_batch_device(d_x, d_yb, batch - 1); }));
        flags.push_back(throws<std::runtime_error>([&] { (void)fb.fft(x0); }));
        flags.push_back(throws<std::invalid_argument>([] { FFTInterface2D bad(0, 4); }));
        flags.push_back(throws<std::invalid_argument>([&] { f2b.fft2d_device(d_x0, d_m2); }));
        flags.push_back(f.fft_batch({}).empty() ? 1 : 0);
        flags.push_back(f.ifft_batch_float(xs, 0).empty() ? 1 : 0);
        flags.push_back(throws<std::runtime_error>([] { FFTInterface huge(1 << 27); }));
        write_block(out, flags);

        // device allocation tracker
        std::vector<long long> bytes;
        bytes.push_back(static_cast<long long>(cusignal::cuda_utils::tracked_device_current_bytes()));
        {
            FFTInterface t(n);
            bytes.push_back(static_cast<long long>(cusignal::cuda_utils::tracked_device_current_bytes()));
        }
        bytes.push_back(static_cast<long long>(cusignal::cuda_utils::tracked_device_current_bytes()));
        {
            DeviceArray<ComplexFloat> t(n);
            bytes.push_back(static_cast<long long>(cusignal::cuda_utils::tracked_device_current_bytes()));
        }
        write_block(out, bytes);
        return 0;
    } catch (const std::exception& e) {
        std::cerr << "fft_interface: " << e.what() << "\n";
        return 1;
    }
}
