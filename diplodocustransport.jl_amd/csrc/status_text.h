// Text of the stepping functor's console lines (host code only, no HIP): numbers in the form
// Julia prints a Float64, and the 3-significant-digit rounding of the symplectic lines.
#pragma once
#include <cmath>
#include <cstdio>
#include <cstdlib>
#include <string>

namespace status_text {

// A Float64 as Julia prints it: the shortest digits that read back to the same value, plainly
// written when 1e-4 <= |v| < 1e6 ("0.25", "2.0", "100000.0", "0.0001") and as d.ddde<exp>
// otherwise ("1.0e6", "1.5e-7"); "0.0", "-0.0", "Inf", "-Inf", "NaN".
inline void julia_float(char* out, size_t cap, double v) {
    if (std::isnan(v)) { snprintf(out, cap, "NaN"); return; }
    if (std::isinf(v)) { snprintf(out, cap, v > 0 ? "Inf" : "-Inf"); return; }
    if (v == 0.0) { snprintf(out, cap, std::signbit(v) ? "-0.0" : "0.0"); return; }
    char e[40];
    for (int prec = 0; prec <= 16; ++prec) {  // correctly rounded, so the first match is it
        snprintf(e, sizeof e, "%.*e", prec, v);
        if (strtod(e, nullptr) == v) break;
    }
    const bool neg = e[0] == '-';
    const char* q = e + (neg ? 1 : 0);
    char digits[24];
    int nd = 0;
    for (; *q && *q != 'e'; ++q)
        if (*q != '.') digits[nd++] = *q;
    while (nd > 1 && digits[nd - 1] == '0') --nd;
    const int exp10 = atoi(q + 1);
    const int pt = exp10 + 1;  // digits before the decimal point
    std::string s = neg ? "-" : "";
    if (-4 < pt && pt <= 6) {
        if (pt <= 0) {
            s += "0.";
            s.append((size_t)-pt, '0');
            s.append(digits, (size_t)nd);
        } else if (pt >= nd) {
            s.append(digits, (size_t)nd);
            s.append((size_t)(pt - nd), '0');
            s += ".0";
        } else {
            s.append(digits, (size_t)pt);
            s += '.';
            s.append(digits + pt, (size_t)(nd - pt));
        }
    } else {
        s += digits[0];
        s += '.';
        if (nd > 1) s.append(digits + 1, (size_t)(nd - 1));
        else s += '0';
        s += 'e';
        s += std::to_string(exp10);
    }
    snprintf(out, cap, "%s", s.c_str());
}

// Cr values of the symplectic lines keep 3 significant digits (round(x, sigdigits=3)):
// rint(x 10^k) / 10^k with k = 3 - (floor(log10 |x|) + 1), ties to even.
inline double round_sig3(double x) {
    if (x == 0.0 || !std::isfinite(x)) return x;
    const int k = 3 - ((int)std::floor(std::log10(std::fabs(x))) + 1);
    const double r = k >= 0 ? std::nearbyint(x * std::pow(10.0, k)) / std::pow(10.0, k)
                            : std::nearbyint(x / std::pow(10.0, -k)) * std::pow(10.0, -k);
    return std::isfinite(r) ? r : x;
}

}  // namespace status_text
