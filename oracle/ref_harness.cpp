// ref_harness.cpp — drives the reference's own FFT.cpp and FilterDesign.cpp
// (compiled unmodified from /root/reference by oracle/Makefile) to write
// golden vectors. TEST INFRASTRUCTURE ONLY; the binary lives in oracle/_ref/.
//
// Usage (binary float32 on stdin/stdout):
//   ref_filter filter <n> <type>   -> (n/2+1) floats   FilterDesign::createFilter
//   ref_filter fwd <n>             n floats in -> (n/2+1) complex out  FFT::forwardReal
//   ref_filter inv <n2>            (n2+1) complex in -> 2*n2 floats    FFT::inverseReal
#include "FFT.h"
#include "FilterDesign.h"

#include <cstdio>
#include <cstdlib>
#include <cstring>
#include <vector>

static bool read_all(float* p, size_t n) { return std::fread(p, sizeof(float), n, stdin) == n; }
static void write_all(const float* p, size_t n) { std::fwrite(p, sizeof(float), n, stdout); }

int main(int argc, char** argv) {
    if (argc < 3) return 2;
    const size_t n = std::strtoul(argv[2], nullptr, 10);
    if (!std::strcmp(argv[1], "filter") && argc >= 4) {
        const int t = std::atoi(argv[3]);
        auto f = ct::FilterDesign::createFilter(n, static_cast<ct::ReconstructionParams::FilterType>(t));
        write_all(f.data(), f.size());
        return 0;
    }
    if (!std::strcmp(argv[1], "fwd")) {
        std::vector<float> in(n);
        if (!read_all(in.data(), n)) return 3;
        std::vector<ct::Complex> out;
        ct::FFT::forwardReal(in, out);
        write_all(reinterpret_cast<const float*>(out.data()), 2 * out.size());
        return 0;
    }
    if (!std::strcmp(argv[1], "inv")) {
        std::vector<ct::Complex> in(n + 1);
        if (!read_all(reinterpret_cast<float*>(in.data()), 2 * (n + 1))) return 3;
        std::vector<float> out;
        ct::FFT::inverseReal(in, out);
        write_all(out.data(), out.size());
        return 0;
    }
    return 2;
}
