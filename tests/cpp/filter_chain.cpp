// Filtering calls written against the cusignal operator API
// (filtering_typed.h:552-565,744-770 of the reference), built with g++ and
// linked to libzkx_radar.so.
//
// usage: filter_chain IN OUT
//   IN : int32 rows, n, taps, up, down, axis; f32 x[rows*n]; f32 h[taps]
//        (x is [rows, n] for axis 1, [n, rows] for axis 0)
//   OUT: f32 upfirdn[...]; f64 resample_default[...]; f32 resample_custom[...];
//        f32 fir_y[rows*n]; f32 fir_zf[rows*(taps-1)]; f32 fir_resident[n]
//        each block preceded by an int64 element count
#include <fstream>
#include <iostream>
#include <vector>

#include "cusignal/operators/filtering/filtering_typed.h"

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

}  // namespace

int main(int argc, char** argv)
{
    if (argc != 3) {
        std::cerr << "usage: filter_chain IN OUT\n";
        return 2;
    }
    try {
        std::ifstream in(argv[1], std::ios::binary);
        int hdr[6];
        read_into(in, hdr, 6);
        const int rows = hdr[0], n = hdr[1], taps = hdr[2], up = hdr[3], down = hdr[4], axis = hdr[5];
        std::vector<float> x(static_cast<std::size_t>(rows) * n), h(taps);
        read_into(in, x.data(), x.size());
        read_into(in, h.data(), h.size());
        const std::vector<int> shape = axis == 1 ? std::vector<int>{rows, n} : std::vector<int>{n, rows};

        auto d_x = cusignal::DeviceArray<float>::from_host(x);
        auto d_h = cusignal::DeviceArray<float>::from_host(h);
        std::ofstream out(argv[2], std::ios::binary);

        cusignal::UpfirdnDeviceResult u;
        cusignal::upfirdn_device<float>(d_h, d_x, up, down, u, {shape, axis});
        write_block(out, u.fp32.to_host());

        cusignal::ResamplePolyOptions ro;
        ro.shape = shape;
        ro.axis = axis;
        cusignal::ResamplePolyDeviceResult rd;
        cusignal::resample_poly_device<float>(d_x, up, down, rd, ro);
        if (rd.dtype != cusignal::ResamplePolyOutputDtype::fp64) throw std::runtime_error("default design must be FP64");
        write_block(out, rd.fp64.to_host());
        cusignal::ResamplePolyDeviceResult rc;
        cusignal::resample_poly_device<float>(d_x, up, down, rc, ro, &d_h);
        if (rc.dtype != cusignal::ResamplePolyOutputDtype::fp32) throw std::runtime_error("custom taps must be FP32");
        write_block(out, rc.fp32.to_host());

        // firfilter with a zi of shape [1, taps-1] broadcast over the rows
        cusignal::FirfilterOptions fo;
        fo.shape = shape;
        fo.axis = axis;
        fo.zi_shape = axis == 1 ? std::vector<int>{1, taps - 1} : std::vector<int>{taps - 1, 1};
        std::vector<float> zi(taps - 1);
        for (int i = 0; i < taps - 1; ++i) zi[i] = 0.25f * static_cast<float>(i % 7) - 0.5f;
        auto d_zi = cusignal::DeviceArray<float>::from_host(zi);
        cusignal::FirfilterDeviceResult f;
        cusignal::firfilter_device<float>(d_h, d_x, f, fo, &d_zi);
        if (!f.has_zf) throw std::runtime_error("zf expected");
        write_block(out, f.y.to_host());
        write_block(out, f.zf.to_host());

        // resident 1-D path on the first n samples
        auto d_x1 = cusignal::DeviceArray<float>::from_host(std::vector<float>(x.begin(), x.begin() + n));
        cusignal::DeviceArray<float> y1(static_cast<std::size_t>(n));
        cusignal::firfilter_fp32_resident_device(d_h, d_x1, y1);
        write_block(out, y1.to_host());

        // argument errors surface as std::invalid_argument
        bool threw = false;
        try {
            cusignal::UpfirdnDeviceResult bad;
            cusignal::upfirdn_device<float>(d_h, d_x, 0, down, bad, {shape, axis});
        } catch (const std::invalid_argument&) {
            threw = true;
        }
        if (!threw) throw std::runtime_error("up=0 was accepted");
        if (!out) throw std::runtime_error("write failed");
    } catch (const std::exception& e) {
        std::cerr << "filter_chain: " << e.what() << "\n";
        return 1;
    }
    return 0;
}
