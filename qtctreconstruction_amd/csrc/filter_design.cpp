// filter_design.cpp — the reference's frequency-domain filter taps, computed on
// the host with the reference's own FP32 arithmetic.
//
// FilterDesign::createFilter (FilterDesign.cpp:13-111) samples the spatial
// ramp kernel, takes its real FFT with the project's FP32 FFT (FFT.cpp:9-85)
// and windows the result in double. Its low-frequency taps carry that FFT's
// rounding: up to ~1e-4 relative at k = 1 for n = 2048, against 5e-8 for an
// exact DFT of the same spatial taps. The reconstruction follows the taps
// closely, so the filter stage uses these taps, bit for bit, to stay within
// 1e-5 of the reference. They are set-up work (n/2 + 1 floats per (n, type),
// cached on the device by the context), not part of the per-row filter.
#include "ct_hip.h"
#include "ct_internal.h"

#include <cmath>
#include <vector>

namespace cthip {
namespace {

struct C32 {
    float re, im;
};
// std::complex<float> arithmetic with every product rounded (the build has
// contraction off): (a.re b.re - a.im b.im) + (a.re b.im + a.im b.re) i
inline C32 mul(C32 a, C32 b) { return {a.re * b.re - a.im * b.im, a.re * b.im + a.im * b.re}; }
inline C32 add(C32 a, C32 b) { return {a.re + b.re, a.im + b.im}; }
inline C32 sub(C32 a, C32 b) { return {a.re - b.re, a.im - b.im}; }
inline C32 conj(C32 a) { return {a.re, -a.im}; }
inline C32 scale(C32 a, float s) { return {a.re * s, a.im * s}; }

// In-place forward complex FFT of m points (a power of two): bit-reversal
// permutation, then radix-2 stages of span 2, 4, ..., m whose butterflies
// take the twiddle W^(j m/span) from a table built from the float angle
// (-2 pi_f / m) * i (FFT.cpp:9-18, 122-158; pi_f = 3.14159265359f).
void forward_complex(std::vector<C32>& d) {
    const size_t m = d.size();
    if (m <= 1) return;
    for (size_t i = 1, j = 0; i < m; ++i) {
        size_t bit = m >> 1;
        for (; j & bit; bit >>= 1) j ^= bit;
        j ^= bit;
        if (i < j) std::swap(d[i], d[j]);
    }
    std::vector<C32> tw(m / 2);
    const float base = 2.0f * 3.14159265359f / (float)m * -1.0f;
    for (size_t i = 0; i < m / 2; ++i) {
        const float ang = base * (float)i;
        tw[i] = {std::cos(ang), std::sin(ang)};
    }
    for (size_t span = 2; span <= m; span <<= 1) {
        const size_t stride = m / span, half = span / 2;
        for (size_t b = 0; b < m; b += span)
            for (size_t j = 0; j < half; ++j) {
                const C32 v = mul(d[b + j + half], tw[j * stride]);
                d[b + j + half] = sub(d[b + j], v);
                d[b + j] = add(d[b + j], v);
            }
    }
}

// Spectrum bins 0..n/2 of the real sequence x[0..n): the n/2-point complex
// FFT of the even/odd pairs, then the split into the real transform with the
// float angle (-pi_f / (n/2)) * k (FFT.cpp:51-85).
std::vector<C32> forward_real(const std::vector<float>& x) {
    const size_t n = x.size(), h = n / 2;
    std::vector<C32> z(h);
    for (size_t i = 0; i < h; ++i) z[i] = {x[2 * i], x[2 * i + 1]};
    forward_complex(z);
    std::vector<C32> out(h + 1);
    const float step = -3.14159265359f / (float)h;
    const C32 mhalf_i = {0.0f, -0.5f};
    for (size_t k = 0; k <= h / 2; ++k) {
        const C32 a = z[k], b = conj(z[(h - k) % h]);
        const C32 even = scale(add(a, b), 0.5f);
        const C32 odd = mul(sub(a, b), mhalf_i);
        const float ang = step * (float)k;
        const C32 w = {std::cos(ang), std::sin(ang)};
        out[k] = add(even, mul(w, odd));
        if (k < h - k) out[h - k] = conj(sub(even, mul(w, odd)));
    }
    return out;
}

}  // namespace

void serial_filter_taps(int n, int type, float* H) {
    constexpr double kPi = 3.14159265358979323846;
    const size_t nn = (size_t)n, ns = nn / 2 + 1;
    // FilterDesign.cpp:31-50: spatial ramp taps 1/4 at 0 and -1/(pi m)^2 at
    // odd offsets m (folded), their real spectrum doubled, DC set to 0
    std::vector<float> f(nn, 0.0f);
    f[0] = 0.25f;
    for (size_t i = 1; i < nn; i += 2) {
        const size_t m = (i < nn / 2) ? i : nn - i;
        if (m == 0) continue;
        const double t = 1.0 / (kPi * (double)m);
        f[i] = (float)(-1.0 * t * t);
    }
    const std::vector<C32> spec = forward_real(f);
    for (size_t k = 0; k < ns; ++k) H[k] = (float)(2.0 * (double)spec[k].re);
    H[0] = 0.0f;
    // FilterDesign.cpp:52-111: windows in double, rounded back to float
    for (size_t k = 0; k < ns; ++k) {
        const double freq = (double)k / (double)nn;
        double w = 1.0;
        switch (type) {
            case CT_FILTER_RAMP: continue;
            case CT_FILTER_HAMMING: w = 0.54 + 0.46 * std::cos(kPi * freq / 0.5); break;
            case CT_FILTER_COSINE: w = std::cos(kPi * freq); break;
            case CT_FILTER_HANN: w = 0.5 + 0.5 * std::cos(kPi * freq / 0.5); break;
            case CT_FILTER_BARTLETT: w = 1.0 - std::fmin(1.0, std::fabs(freq) / 0.5); break;
            default: {  // Shepp-Logan, and every unknown type (FilterDesign.cpp:24-26)
                if (k == 0) continue;
                const double om = kPi * freq;
                if (std::fabs(om) <= 1e-12) continue;
                w = std::sin(om) / om;
            } break;
        }
        H[k] = (float)((double)H[k] * w);
    }
}

}  // namespace cthip
