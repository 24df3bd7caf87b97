// Radar chain written against the cusignal operator API, the way the
// reference's Task1 pipeline calls it (Task1/pipeline/src/steps/step2/step2.cu:20-36,
// step3/step3.cu:17-35, step4/step4.cu:30-84), built with g++ against the
// headers in include/cusignal and linked to libzkx_radar.so.  It shows that a
// reference caller compiles and runs unchanged on this library.
//
// usage: task1_chain IN OUT
//   IN : int32 P, R, T, guard_d, guard_r, ref_d, ref_r, doppler_window; f64 pfa;
//        complex64 echo[P*R]; complex64 waveform[T]
//   OUT: complex64 compressed[P*R]; complex64 range_doppler[P*R]; f32 power[P*R];
//        f64 alpha; f32 threshold[P*R]; u8 detections[P*R];
//        f32 ambiguity[(2T-1)*nfreq]; f64 cut over delay[2T-1]; f64 cut over Doppler[nfreq]
// The ambiguity step follows Task1/pipeline/src/steps/step5/step5.cu:22-57
// (fs = 2.56 MHz, prf = 8 kHz as in the Task1 profiles).
#include <cstdio>
#include <cstdlib>
#include <fstream>
#include <iostream>
#include <vector>

#include "cusignal/operators/radartools/radartools_typed.h"

using cusignal::TypedComplex;

namespace {

template <class T>
void read_into(std::ifstream& in, T* dst, std::size_t count)
{
    in.read(reinterpret_cast<char*>(dst), static_cast<std::streamsize>(sizeof(T) * count));
    if (!in) throw std::runtime_error("short input file");
}

template <class T>
void write_vec(std::ofstream& out, const std::vector<T>& v)
{
    out.write(reinterpret_cast<const char*>(v.data()), static_cast<std::streamsize>(sizeof(T) * v.size()));
}

}  // namespace

int main(int argc, char** argv)
{
    if (argc != 3) {
        std::cerr << "usage: task1_chain IN OUT\n";
        return 2;
    }
    try {
        std::ifstream in(argv[1], std::ios::binary);
        int hdr[8];
        double pfa = 0.0;
        read_into(in, hdr, 8);
        read_into(in, &pfa, 1);
        const int P = hdr[0], R = hdr[1], T = hdr[2];
        std::vector<TypedComplex<float>> echo(static_cast<std::size_t>(P) * R), waveform(T);
        read_into(in, echo.data(), echo.size());
        read_into(in, waveform.data(), waveform.size());

        // step 2: pulse compression, normalised, no window, nfft = R
        cusignal::PulseCompressionOptions pc;
        pc.num_pulses = P;
        pc.samples_per_pulse = R;
        pc.normalize = true;
        pc.nfft = R;
        pc.window = cusignal::RadarWindowKind::none;
        auto d_echo = cusignal::DeviceArray<TypedComplex<float>>::from_host(echo);
        auto d_waveform = cusignal::DeviceArray<TypedComplex<float>>::from_host(waveform);
        cusignal::RadarComplexDeviceResult compressed = cusignal::pulse_compression_device<float>(d_echo, d_waveform, pc);
        if (compressed.is_fp64) throw std::runtime_error("window none must keep the FP32 result");
        const std::vector<ComplexFloat> compressed_h = compressed.fp32.to_host();

        // step 3: Doppler FFT over pulses into a resident output
        cusignal::PulseDopplerOptions pd;
        pd.num_pulses = P;
        pd.samples_per_pulse = R;
        pd.nfft = P;
        pd.window = static_cast<cusignal::RadarWindowKind>(hdr[7]);
        std::vector<TypedComplex<float>> compressed_typed(compressed_h.size());
        for (std::size_t i = 0; i < compressed_h.size(); ++i)
            compressed_typed[i] = {compressed_h[i].re, compressed_h[i].im};
        auto d_compressed = cusignal::DeviceArray<TypedComplex<float>>::from_host(compressed_typed);
        cusignal::DeviceArray<ComplexFloat> range_doppler(static_cast<std::size_t>(P) * R);
        cusignal::DeviceArray<float> power(static_cast<std::size_t>(P) * R);
        cusignal::pulse_doppler_device<float>(d_compressed, pd, range_doppler, power);

        // step 4: alpha and 2-D CA-CFAR on the resident power map
        const int od = hdr[3] + hdr[5], orr = hdr[4] + hdr[6];
        const int reference_count = (2 * od + 1) * (2 * orr + 1) - (2 * hdr[3] + 1) * (2 * hdr[4] + 1);
        cusignal::DeviceArray<double> alpha = cusignal::cfar_alpha_device<float>(static_cast<float>(pfa), reference_count);
        cusignal::CaCfarOptions cf;
        cf.guard_cells = {hdr[3], hdr[4]};
        cf.reference_cells = {hdr[5], hdr[6]};
        cf.pfa = pfa;
        cusignal::DeviceArray<float> threshold(static_cast<std::size_t>(P) * R);
        cusignal::DeviceArray<cusignal::CfarDetection> detections(static_cast<std::size_t>(P) * R);
        cusignal::ca_cfar_device<float>(power, {P, R}, cf, threshold, detections);

        // step 5: ambiguity function of the waveform, 2-D and the two cuts through 0
        cusignal::AmbgfunOptions ao;
        ao.fs = 2.56e6;
        ao.prf = 8000.0;
        ao.cut = cusignal::AmbgfunCut::two_dimensional;
        cusignal::AmbgfunDeviceResult amb_2d, amb_delay, amb_doppler;
        cusignal::ambgfun_device<float>(d_waveform, amb_2d, ao);
        ao.cut = cusignal::AmbgfunCut::doppler;
        ao.cut_value = 0.0;
        cusignal::ambgfun_device<float>(d_waveform, amb_delay, ao);
        ao.cut = cusignal::AmbgfunCut::delay;
        cusignal::ambgfun_device<float>(d_waveform, amb_doppler, ao);
        if (amb_2d.dtype != cusignal::AmbgfunOutputDtype::fp32 ||
            amb_delay.dtype != cusignal::AmbgfunOutputDtype::fp64)
            throw std::runtime_error("ambgfun output dtype");

        std::ofstream out(argv[2], std::ios::binary);
        write_vec(out, compressed_h);
        write_vec(out, range_doppler.to_host());
        write_vec(out, power.to_host());
        write_vec(out, alpha.to_host());
        write_vec(out, threshold.to_host());
        write_vec(out, detections.to_host());
        write_vec(out, amb_2d.fp32.to_host());
        write_vec(out, amb_delay.fp64.to_host());
        write_vec(out, amb_doppler.fp64.to_host());
        if (!out) throw std::runtime_error("write failed");
    } catch (const std::exception& e) {
        std::cerr << "task1_chain: " << e.what() << "\n";
        return 1;
    }
    return 0;
}
