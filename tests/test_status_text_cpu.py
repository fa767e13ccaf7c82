"""Text of the stepping functor's console lines (csrc/status_text.h), compiled with g++ on the
host: Float64 values in the form Julia prints them (`"Cr = $Cr, t=$t_start, dt=$dt"`,
src/SteppingMethods.jl:87-93) and round(x, sigdigits=3) of the symplectic lines
(src/SteppingMethods.jl:222-228).  The expected strings are written out by hand from Julia's
printing rules (shortest round-trip digits; plain for 1e-4 <= |x| < 1e6, d.ddde<exp> otherwise;
whole numbers keep ".0").  Julia itself is not in the image, so these strings are the pin."""
import math
import shutil
import struct
import subprocess
from pathlib import Path

import pytest

CSRC = Path(__file__).resolve().parents[1] / "diplodocustransport.jl_amd" / "csrc"

DRIVER = r'''
#include <cstdint>
#include <cstring>
#include "status_text.h"
int main(int argc, char** argv) {
    for (int i = 1; i < argc; ++i) {
        uint64_t bits = strtoull(argv[i], nullptr, 16);
        double v;
        memcpy(&v, &bits, 8);
        char a[64], b[64];
        status_text::julia_float(a, sizeof a, v);
        status_text::julia_float(b, sizeof b, status_text::round_sig3(v));
        printf("%s %s\n", a, b);
    }
    return 0;
}
'''

# value, Julia's string(value), string(round(value, sigdigits=3))
CASES = [
    (0.25, "0.25", "0.25"),
    (2.0, "2.0", "2.0"),
    (10.0, "10.0", "10.0"),
    (-0.5, "-0.5", "-0.5"),
    (0.1, "0.1", "0.1"),
    (1.0 / 3.0, "0.3333333333333333", "0.333"),
    (0.30000000000000004, "0.30000000000000004", "0.3"),
    (100000.0, "100000.0", "100000.0"),
    (123456.7, "123456.7", "123000.0"),
    (1e6, "1.0e6", "1.0e6"),
    (1234567.0, "1.234567e6", "1.23e6"),
    (12345678.0, "1.2345678e7", "1.23e7"),
    (1e22, "1.0e22", "1.0e22"),
    (0.0001, "0.0001", "0.0001"),
    (1.1498e-4, "0.00011498", "0.000115"),
    (1e-5, "1.0e-5", "1.0e-5"),
    (1.5e-7, "1.5e-7", "1.5e-7"),
    (0.0012345, "0.0012345", "0.00123"),
    (999.7, "999.7", "1000.0"),
    (float.fromhex("0x1.b7cdfe0000000p-34"), "1.000000013351432e-10", "1.0e-10"),  # Float64(1f-10)
    (5e-324, "5.0e-324", "5.0e-324"),
    (1.7976931348623157e308, "1.7976931348623157e308", "1.7976931348623157e308"),
    (0.0, "0.0", "0.0"),
    (-0.0, "-0.0", "-0.0"),
    (math.inf, "Inf", "Inf"),
    (-math.inf, "-Inf", "-Inf"),
    (math.nan, "NaN", "NaN"),
]


@pytest.fixture(scope="module")
def driver(tmp_path_factory):
    gxx = shutil.which("g++")
    if gxx is None:
        pytest.skip("g++ not found")
    d = tmp_path_factory.mktemp("status_text")
    (d / "main.cpp").write_text(DRIVER)
    exe = d / "status_text"
    subprocess.run([gxx, "-O2", "-std=c++17", f"-I{CSRC}", str(d / "main.cpp"), "-o", str(exe)],
                   check=True)
    return exe


def test_julia_float_text(driver):
    args = [f"{struct.unpack('<Q', struct.pack('<d', v))[0]:016x}" for v, _, _ in CASES]
    out = subprocess.run([str(driver), *args], check=True, capture_output=True,
                         text=True).stdout.splitlines()
    assert len(out) == len(CASES)
    for (v, plain, sig3), line in zip(CASES, out):
        got_plain, got_sig3 = line.split(" ")
        assert got_plain == plain, (v, got_plain, plain)
        assert got_sig3 == sig3, (v, got_sig3, sig3)
