// The radar and filtering operators called through the cusignal operator API
// for one element type of the reference's instantiations (float, __half,
// int32, int16, int8; radartools_typed.cpp:810-811, filtering_typed.cu:683+),
// plus the complex filtering extensions, fftconvolve, decimate,
// channelize_poly, correlate (1-D and N-D), ambgfun and correlate2d.  Built
// with g++ and
// linked to libzkx_radar.so.
//
// usage: typed_ops IN OUT
//   IN : int32 dtype (0 f32, 1 f16, 2 i32, 3 i16, 4 i8), P, R, T, rows, n,
//        taps, up, down; then, in the element type:
//        x[P*R*2] (complex), tmpl[T*2] (complex), cfar[P*R], xf[rows*n],
//        h[taps], zi[rows*(taps-1)], coeffs[taps]
//   OUT: blocks of (int32 kind, int64 bytes, bytes), kind 0 f32, 1 f64,
//        2 u8, 3 raw input-typed, 4 int32 flag
#include <fstream>
#include <iostream>
#include <vector>

#include "cusignal/operators/convolution/convolution_typed.h"
#include "cusignal/operators/radartools/radartools_typed.h"

namespace {

struct Writer {
    std::ofstream out;
    explicit Writer(const char* path) : out(path, std::ios::binary) {}
    void block(int kind, const void* p, std::size_t bytes)
    {
        const long long n = static_cast<long long>(bytes);
        out.write(reinterpret_cast<const char*>(&kind), sizeof(kind));
        out.write(reinterpret_cast<const char*>(&n), sizeof(n));
        out.write(reinterpret_cast<const char*>(p), static_cast<std::streamsize>(bytes));
    }
    template <class E>
    void vec(int kind, const std::vector<E>& v)
    {
        block(kind, v.data(), sizeof(E) * v.size());
    }
    void flag(int v) { block(4, &v, sizeof(v)); }
};

template <class E>
std::vector<E> read_vec(std::ifstream& in, std::size_t count)
{
    std::vector<E> v(count);
    in.read(reinterpret_cast<char*>(v.data()), static_cast<std::streamsize>(sizeof(E) * count));
    if (!in) throw std::runtime_error("short input file");
    return v;
}

void write_radar(Writer& w, const cusignal::RadarComplexDeviceResult& r)
{
    w.flag(r.is_fp64 ? 1 : 0);
    if (r.is_fp64) w.vec(1, r.fp64.to_host());
    else w.vec(0, r.fp32.to_host());
}

template <class T>
int run(std::ifstream& in, Writer& w, const int* hdr)
{
    using cusignal::DeviceArray;
    using C = cusignal::TypedComplex<T>;
    const int P = hdr[1], R = hdr[2], T_ = hdr[3], rows = hdr[4], n = hdr[5], taps = hdr[6], up = hdr[7],
              down = hdr[8];
    const auto x = read_vec<C>(in, static_cast<std::size_t>(P) * R);
    const auto tmpl = read_vec<C>(in, T_);
    const auto cfar = read_vec<T>(in, static_cast<std::size_t>(P) * R);
    const auto xf = read_vec<T>(in, static_cast<std::size_t>(rows) * n);
    const auto h = read_vec<T>(in, taps);
    const auto zi = read_vec<T>(in, static_cast<std::size_t>(rows) * (taps - 1));
    const auto coeffs = read_vec<T>(in, taps);
    const auto d_x = DeviceArray<C>::from_host(x);
    const auto d_t = DeviceArray<C>::from_host(tmpl);

    // 1-2: pulse compression, no window (FP64 storage for integer types) and Hamming
    cusignal::PulseCompressionOptions pco;
    pco.num_pulses = P;
    pco.samples_per_pulse = R;
    write_radar(w, cusignal::pulse_compression_device<T>(d_x, d_t, pco));
    pco.window = cusignal::RadarWindowKind::hamming;
    pco.normalize = true;
    write_radar(w, cusignal::pulse_compression_device<T>(d_x, d_t, pco));
    // 3-4: pulse Doppler, Hann (FP64) and the preallocated FP32 overload
    cusignal::PulseDopplerOptions pdo;
    pdo.num_pulses = P;
    pdo.samples_per_pulse = R;
    pdo.window = cusignal::RadarWindowKind::hann;
    write_radar(w, cusignal::pulse_doppler_device<T>(d_x, pdo));
    pdo.window = cusignal::RadarWindowKind::none;
    DeviceArray<ComplexFloat> pd(static_cast<std::size_t>(P) * R);
    cusignal::pulse_doppler_device<T>(d_x, pdo, pd);
    w.vec(0, pd.to_host());
    // 5-6: cfar_alpha (pfa 1e-3 for floating types, 1 for integers)
    T pfa;
    if constexpr (std::is_integral_v<T>) pfa = static_cast<T>(1);
    else if constexpr (std::is_same_v<T, __half>) pfa = __float2half(1.0e-3f);
    else pfa = 1.0e-3f;
    const double a_cpu = cusignal::cfar_alpha_typed_cpu<T>(pfa, 24);
    w.block(1, &a_cpu, sizeof(a_cpu));
    w.vec(1, cusignal::cfar_alpha_device<T>(pfa, 24).to_host());
    // 7: 2-D CA-CFAR on the real map
    const auto d_c = DeviceArray<T>::from_host(cfar);
    cusignal::CaCfarOptions co;
    co.guard_cells = {2, 2};
    co.reference_cells = {4, 4};
    co.pfa = 1.0e-3;
    const cusignal::CaCfarDeviceResult cr = cusignal::ca_cfar_device<T>(d_c, {P, R}, co);
    w.vec(0, cr.threshold.to_host());
    w.vec(2, cr.detections.to_host());
    // 8: upfirdn on [rows, n] along axis 1
    const auto d_xf = DeviceArray<T>::from_host(xf);
    const auto d_h = DeviceArray<T>::from_host(h);
    cusignal::UpfirdnDeviceResult ur;
    cusignal::upfirdn_device<T>(d_h, d_xf, up, down, ur, {{rows, n}, 1});
    w.flag(ur.dtype == cusignal::UpfirdnOutputDtype::fp64 ? 1 : 0);
    if (ur.dtype == cusignal::UpfirdnOutputDtype::fp64) w.vec(1, ur.fp64.to_host());
    else w.vec(0, ur.fp32.to_host());
    // 9-11: resample_poly default design, custom coefficients, identity
    cusignal::ResamplePolyOptions ro;
    ro.shape = {rows, n};
    ro.axis = 1;
    cusignal::ResamplePolyDeviceResult rr;
    cusignal::resample_poly_device<T>(d_xf, up, down, rr, ro);
    w.flag(static_cast<int>(rr.dtype));
    w.vec(1, rr.fp64.to_host());
    const auto d_co = DeviceArray<T>::from_host(coeffs);
    cusignal::resample_poly_device<T>(d_xf, up, down, rr, ro, &d_co);
    w.flag(static_cast<int>(rr.dtype));
    if (rr.dtype == cusignal::ResamplePolyOutputDtype::fp32) w.vec(0, rr.fp32.to_host());
    else w.vec(1, rr.fp64.to_host());
    cusignal::resample_poly_device<T>(d_xf, 3, 3, rr, ro);
    w.flag(static_cast<int>(rr.dtype));
    if constexpr (std::is_same_v<T, float>) w.vec(3, rr.fp32.to_host());
    else if constexpr (std::is_same_v<T, __half>) w.vec(3, rr.fp16.to_host());
    else if constexpr (std::is_same_v<T, std::int32_t>) w.vec(3, rr.int32.to_host());
    else if constexpr (std::is_same_v<T, std::int16_t>) w.vec(3, rr.int16.to_host());
    else w.vec(3, rr.int8.to_host());
    // 12: firfilter with zi along axis 1
    const auto d_zi = DeviceArray<T>::from_host(zi);
    cusignal::FirfilterOptions fo;
    fo.shape = {rows, n};
    fo.axis = 1;
    fo.zi_shape = {rows, taps - 1};
    cusignal::FirfilterDeviceResult fr;
    cusignal::firfilter_device<T>(d_h, d_xf, fr, fo, &d_zi);
    w.vec(0, fr.y.to_host());
    w.vec(0, fr.zf.to_host());
    // 13: fftconvolve of the P pulses with the template
    cusignal::FftconvolveDeviceResult cv;
    cusignal::fftconvolve_device<T>(d_x, d_t, cv, {{P, R}, cusignal::FftconvolveMode::same});
    w.vec(0, cv.out.to_host());
    // 14-16: decimate along axis 1: designed filter (FP64), custom
    // coefficients in the input type (FP32, FP64 for int32), causal form
    cusignal::DecimateOptions dop;
    dop.shape = {rows, n};
    dop.axis = 1;
    cusignal::DecimateDeviceResult dr;
    cusignal::decimate_device<T>(d_xf, down + 1, dr, dop);
    w.flag(dr.dtype == cusignal::DecimateOutputDtype::fp64 ? 1 : 0);
    w.vec(1, dr.fp64.to_host());
    cusignal::decimate_device<T>(d_xf, down + 1, dr, dop, &d_h);
    w.flag(dr.dtype == cusignal::DecimateOutputDtype::fp64 ? 1 : 0);
    if (dr.dtype == cusignal::DecimateOutputDtype::fp64) w.vec(1, dr.fp64.to_host());
    else w.vec(0, dr.fp32.to_host());
    dop.zero_phase = false;
    dop.filter_order = 24;
    cusignal::decimate_device<T>(d_xf, down + 1, dr, dop);
    w.vec(1, dr.fp64.to_host());
    // 17: channelize_poly of the filtering signal with its taps, 4 channels
    cusignal::ChannelizePolyDeviceResult cp;
    cusignal::channelize_poly_device<T>(d_xf, d_h, 4, cp);
    w.vec(0, cp.values.to_host());
    // 18-20: correlate (direct) of the first filtering line with the taps,
    // full / same / valid, in the input type
    {
        const auto d_x1 = DeviceArray<T>::from_host(std::vector<T>(xf.begin(), xf.begin() + n));
        for (const char* mode : {"full", "same", "valid"}) {
            DeviceArray<T> co(cusignal::correlate_output_size(static_cast<std::size_t>(n), h.size(), mode));
            cusignal::correlate_device<T>(d_x1, d_h, co, mode);
            w.vec(3, co.to_host());
        }
        // 20b: the FFT method through the workspace overload, 'same'
        cusignal::CorrelateDeviceWorkspace cw({n, static_cast<int>(h.size()), "same"});
        DeviceArray<T> cf(cw.out_len);
        cusignal::correlate_device<T>(d_x1, d_h, cf, cw);
        w.vec(3, cf.to_host());
    }
    // 21-23: ambiguity function of the template: 2-D (FP32 for float and half,
    // FP64 storage for integers), delay cut and Doppler cut (FP64 storage)
    cusignal::AmbgfunOptions ao;
    ao.fs = 1000.0;
    ao.prf = 100.0;
    cusignal::AmbgfunDeviceResult ar;
    cusignal::ambgfun_device<T>(d_t, ar, ao);
    w.flag(ar.dtype == cusignal::AmbgfunOutputDtype::fp64 ? 1 : 0);
    if (ar.dtype == cusignal::AmbgfunOutputDtype::fp64) w.vec(1, ar.fp64.to_host());
    else w.vec(0, ar.fp32.to_host());
    ao.cut = cusignal::AmbgfunCut::delay;
    ao.cut_value = 0.002;
    cusignal::ambgfun_device<T>(d_t, ar, ao);
    w.vec(1, ar.fp64.to_host());
    ao.cut = cusignal::AmbgfunCut::doppler;
    ao.cut_value = 37.0;
    cusignal::ambgfun_device<T>(d_t, ar, ao);
    w.vec(1, ar.fp64.to_host());
    // 24-26: correlate2d of the filtering samples as a 30 x 100 matrix with
    // the first 30 taps as 5 x 6: same/symm, full/fill (fill value 2),
    // valid/wrap, in the input type
    {
        const auto d_k = DeviceArray<T>::from_host(std::vector<T>(h.begin(), h.begin() + 30));
        const char* forms[3][2] = {{"same", "symm"}, {"full", "fill"}, {"valid", "wrap"}};
        for (const auto& f : forms) {
            cusignal::Correlate2DOptions o;
            o.rows1 = 30;
            o.cols1 = 100;
            o.rows2 = 5;
            o.cols2 = 6;
            o.mode = f[0];
            o.boundary = f[1];
            o.fillvalue = 2.0;
            cusignal::Correlate2DWorkspace ws(o);
            DeviceArray<T> c2(ws.output_size());
            cusignal::correlate2d_device<T>(d_xf, d_k, c2, ws);
            w.vec(3, c2.to_host());
        }
    }
    // 26b-26d: the same pair through the N-D FFT form (CorrelateNDWorkspace),
    // full / same / valid, in the input type
    {
        const auto d_k = DeviceArray<T>::from_host(std::vector<T>(h.begin(), h.begin() + 30));
        for (const char* mode : {"full", "same", "valid"}) {
            cusignal::CorrelateNDOptions o;
            o.shape1 = {30, 100};
            o.shape2 = {5, 6};
            o.mode = mode;
            cusignal::CorrelateNDWorkspace ws(o);
            DeviceArray<T> cn(ws.output_size());
            cusignal::correlate_device<T>(d_xf, d_k, cn, ws);
            w.vec(3, cn.to_host());
        }
    }
    if constexpr (std::is_same_v<T, float>) {
        // 27-29: complex filtering extensions on the pulses (float taps)
        cusignal::UpfirdnComplexDeviceResult uc;
        cusignal::upfirdn_device<float>(d_h, d_x, up, down, uc, {{P, R}, 1});
        w.vec(0, uc.fp32.to_host());
        cusignal::ResamplePolyComplexDeviceResult rc;
        cusignal::ResamplePolyOptions rco;
        rco.shape = {P, R};
        rco.axis = 1;
        cusignal::resample_poly_device<float>(d_x, up, down, rc, rco, &d_co);
        w.vec(0, rc.fp32.to_host());
        cusignal::FirfilterComplexDeviceResult fc;
        cusignal::FirfilterOptions fco;
        fco.shape = {P, R};
        fco.axis = 1;
        cusignal::firfilter_device<float>(d_h, d_x, fc, fco);
        w.vec(0, fc.y.to_host());
        // 30-31: N-D fftconvolve: the filtering samples as 30 x 100 with the
        // first 30 taps as 5 x 6 over both axes ('same', real result), and
        // the pulses [P, R] with the template as [1, T] along axis 1 ('full',
        // complex result)
        cusignal::FftconvolveNDOptions n1;
        n1.shape1 = {30, 100};
        n1.shape2 = {5, 6};
        n1.mode = cusignal::FftconvolveMode::same;
        const auto d_k = DeviceArray<float>::from_host(std::vector<float>(h.begin(), h.begin() + 30));
        DeviceArray<float> r1(30 * 100);
        cusignal::fftconvolve_device(d_xf, d_k, r1, n1);
        w.vec(0, r1.to_host());
        cusignal::FftconvolveNDOptions n2;
        n2.shape1 = {P, R};
        n2.shape2 = {1, T_};
        n2.axes = {-1};
        std::size_t cnt = 1;
        for (int v : cusignal::fftconvolve_nd_output_shape(n2)) cnt *= static_cast<std::size_t>(v);
        DeviceArray<C> r2(cnt);
        cusignal::fftconvolve_device(d_x, d_t, r2, n2);
        w.vec(0, r2.to_host());
    }
    return 0;
}

}  // namespace

int main(int argc, char** argv)
{
    if (argc != 3) {
        std::cerr << "usage: typed_ops IN OUT\n";
        return 2;
    }
    try {
        std::ifstream in(argv[1], std::ios::binary);
        const auto hdr = read_vec<int>(in, 9);
        Writer w(argv[2]);
        switch (hdr[0]) {
        case 0: return run<float>(in, w, hdr.data());
        case 1: return run<__half>(in, w, hdr.data());
        case 2: return run<std::int32_t>(in, w, hdr.data());
        case 3: return run<std::int16_t>(in, w, hdr.data());
        case 4: return run<std::int8_t>(in, w, hdr.data());
        default: std::cerr << "unknown dtype\n"; return 2;
        }
    } catch (const std::exception& e) {
        std::cerr << "error: " << e.what() << "\n";
        return 1;
    }
}
