"""CPU: the device log1pf port (gjxl_amd/csrc/libm_port.h) equals the host
libm's log1pf, which is what the reference's std::log1p(float) calls
(src/codec/adaptive_quantization.cpp:656).

Compiles the product header for the host with g++ and compares on a strided
sweep of every non-negative finite float (an exhaustive sweep of all 2^31
values was run when the port was written and matched everywhere).
"""
import subprocess
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
