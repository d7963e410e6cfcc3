"""CPU: the device libm ports (gjxl_amd/csrc/libm_port.h) equal the host
libm, which is what the reference calls:

* log1pf for std::log1p(float) (src/codec/adaptive_quantization.cpp:656);
* powf for std::pow(float, float) with y = 1/8 (src/codec/ac_strategy.cpp:
  979) and y = 1/16 (src/codec/butteraugli.cpp:188-190).

Compiles the product header for the host with g++ and compares on strided
sweeps of every non-negative finite float (exhaustive sweeps of all 2^31
values were run when the ports were written and matched everywhere).  The
powf sweep also shows that the correctly rounded root, which the device used
before, differs from the reference's powf, so the check can fail.
"""
import struct
import subprocess
from decimal import Decimal, getcontext
import tempfile
from pathlib import Path

ROOT = Path(__file__).resolve().parents[1]

PROGRAM = r"""
#include <cmath>
#include <cstdio>
#include <cstdint>
#include <cstring>
#include "libm_port.h"
int main() {
  unsigned long long n = 0, diff = 0;
  for (uint64_t b = 0; b <= 0x7f800000u; b += 61) {
    uint32_t bits = (uint32_t)b; float x; std::memcpy(&x, &bits, 4);
    const float g = log1pf(x), m = gjxl_amd::FdlibmLog1pf(x);
    ++n; if (std::memcmp(&g, &m, 4) != 0) ++diff;
  }
  for (uint64_t b = 0x80000000u; b <= 0xbf800000u; b += 997) {
    uint32_t bits = (uint32_t)b; float x; std::memcpy(&x, &bits, 4);
    const float g = log1pf(x), m = gjxl_amd::FdlibmLog1pf(x);
    ++n; if (std::memcmp(&g, &m, 4) != 0) ++diff;
  }
  std::printf("%llu %llu\n", n, diff);
  return 0;
}
"""


def test_log1pf_port_matches_host_libm():
    with tempfile.TemporaryDirectory() as tmp:
        src = Path(tmp) / "check.cpp"
        exe = Path(tmp) / "check"
        src.write_text(PROGRAM)
        subprocess.run(["g++", "-O2", "-std=c++20", "-ffp-contract=off",
                        "-I", str(ROOT / "gjxl_amd" / "csrc"), str(src), "-o",
                        str(exe), "-lm"], check=True)
        n, diff = map(int, subprocess.run([str(exe)], check=True,
                                          capture_output=True,
                                          text=True).stdout.split())
    assert n > 30_000_000
    assert diff == 0


POWF_PROGRAM = r"""
#include <cmath>
#include <cstdio>
#include <cstdint>
#include <cstring>
#include <initializer_list>
#include "libm_port.h"
int main() {
  unsigned long long n = 0, diff = 0, root_diff = 0, special = 0;
  for (float e : {1.0f / 8.0f, 1.0f / 16.0f}) {
    for (uint64_t b = 0; b <= 0x7f800000u; b += 29) {
      uint32_t bits = (uint32_t)b; float x; std::memcpy(&x, &bits, 4);
      const float g = powf(x, e), m = gjxl_amd::GlibcPowf(x, e);
      double r = std::sqrt(std::sqrt(std::sqrt((double)x)));
      if (e < 0.1f) r = std::sqrt(r);
      const float cr = (float)r;
      ++n; if (std::memcmp(&g, &m, 4) != 0) ++diff;
      if (std::memcmp(&g, &cr, 4) != 0) ++root_diff;
    }
    // -0, negatives, NaN
    for (float x : {-0.0f, -1.0f, -1e-30f, NAN}) {
      const float g = powf(x, e), m = gjxl_amd::GlibcPowf(x, e);
      if (!(std::memcmp(&g, &m, 4) == 0 || (std::isnan(g) && std::isnan(m)))) ++special;
    }
  }
  std::printf("%llu %llu %llu %llu\n", n, diff, root_diff, special);
  return 0;
}
"""


def test_powf_port_matches_host_libm():
    with tempfile.TemporaryDirectory() as tmp:
        src = Path(tmp) / "check.cpp"
        exe = Path(tmp) / "check"
        src.write_text(POWF_PROGRAM)
        subprocess.run(["g++", "-O2", "-std=c++20", "-ffp-contract=off",
                        "-I", str(ROOT / "gjxl_amd" / "csrc"), str(src), "-o",
                        str(exe), "-lm"], check=True)
        n, diff, root_diff, special = map(int, subprocess.run(
            [str(exe)], check=True, capture_output=True, text=True).stdout.split())
    assert n > 140_000_000
    assert diff == 0
    assert special == 0
    # the correctly rounded root is not the reference's powf
    assert root_diff > 1000


def test_powf_exp2_table_is_rounded_powers_of_two():
    """kExp2[i] is RN(2^(i/32)) as binary64 bits minus i << 47."""
    text = (ROOT / "gjxl_amd" / "csrc" / "libm_port.h").read_text()
    body = text[text.index("kExp2[32] = {"):]
    body = body[body.index("{") + 1:body.index("}")]
    words = [int(w.strip().rstrip("ul"), 16) for w in body.split(",") if w.strip()]
    assert len(words) == 32
    getcontext().prec = 60
    for i, word in enumerate(words):
        bits = struct.unpack("<Q", struct.pack("<d", float(Decimal(2) ** (Decimal(i) / 32))))[0]
        assert (bits - (i << 47)) % (1 << 64) == word, i
