// A C++ caller of the plain path through include/gjxl_amd.h only, as the
// reference's encode workflow would call it (host code, no HIP headers).
//
//   plain_path_caller WIDTH HEIGHT INPUT.rgb OUTPUT_PREFIX [REPEAT]
//
// INPUT.rgb holds WIDTH*HEIGHT packed RGB8 pixels.  The program uploads them
// from pinned host memory, runs the whole path (effort 7, distance 1.0) and
// reads the tokens back, REPEAT times (default 1).  It writes the last run's
// token stream and frame to OUTPUT_PREFIX.{values,contexts,group_offsets,
// cells,field,raw_quant,quantized,dc_quantized,y_to_x,y_to_b} and prints one
// JSON line: token count, scores, matrix scales, mean wall milliseconds per
// run and the device stage times of the last run.
#include <chrono>
#include <cstdio>
#include <cstdlib>
#include <cstring>
#include <string>
#include <vector>

#include "gjxl_amd.h"

namespace {

int Fail(const char* what, GJXLResult r) {
  std::fprintf(stderr, "%s failed (%d): %s\n", what, static_cast<int>(r), gjxl_amd_last_error());
  return 1;
}

bool WriteFile(const std::string& path, const void* data, size_t bytes) {
  FILE* f = std::fopen(path.c_str(), "wb");
  if (f == nullptr) return false;
  const bool ok = std::fwrite(data, 1, bytes, f) == bytes;
  return std::fclose(f) == 0 && ok;
}

bool Download(GJXLAmdDevice* device, const void* src, size_t bytes, const std::string& path) {
  std::vector<unsigned char> host(bytes);
  if (bytes != 0 && gjxl_amd_download(device, host.data(), src, bytes) != GJXL_OK) return false;
  return WriteFile(path, host.data(), bytes);
}

}  // namespace

int main(int argc, char** argv) {
  if (argc < 5) {
    std::fprintf(stderr, "usage: %s WIDTH HEIGHT INPUT.rgb OUTPUT_PREFIX [REPEAT]\n", argv[0]);
    return 2;
  }
  const uint32_t width = static_cast<uint32_t>(std::strtoul(argv[1], nullptr, 10));
  const uint32_t height = static_cast<uint32_t>(std::strtoul(argv[2], nullptr, 10));
  const std::string prefix = argv[4];
  const int repeat = argc > 5 ? std::atoi(argv[5]) : 1;
  const size_t bytes = static_cast<size_t>(width) * height * 3;

  GJXLAmdDevice* device = nullptr;
  GJXLResult r = gjxl_amd_device_create(0, &device);
  if (r != GJXL_OK) return Fail("gjxl_amd_device_create", r);
  void* pixels = nullptr;
  if ((r = gjxl_amd_host_alloc(bytes, &pixels)) != GJXL_OK) return Fail("gjxl_amd_host_alloc", r);
  FILE* in = std::fopen(argv[3], "rb");
  if (in == nullptr || std::fread(pixels, 1, bytes, in) != bytes) {
    std::fprintf(stderr, "cannot read %zu bytes from %s\n", bytes, argv[3]);
    return 1;
  }
  std::fclose(in);

  GJXLAmdPlainPathOptions options;
  if ((r = gjxl_amd_plain_path_options_init(&options, sizeof(options))) != GJXL_OK) {
    return Fail("gjxl_amd_plain_path_options_init", r);
  }
  GJXLAmdPlainPath* path = nullptr;
  if ((r = gjxl_amd_plain_path_create(device, width, height, &options, &path)) != GJXL_OK) {
    return Fail("gjxl_amd_plain_path_create", r);
  }
  const size_t blocks = static_cast<size_t>((width + 7) / 8) * ((height + 7) / 8);
  const size_t groups = static_cast<size_t>(((width + 7) / 8 + 31) / 32) *
                        (((height + 7) / 8 + 31) / 32);
  const size_t capacity = 3 * 65 * blocks;
  void *values = nullptr, *contexts = nullptr, *offsets = nullptr;
  if (gjxl_amd_host_alloc(4 * capacity, &values) != GJXL_OK ||
      gjxl_amd_host_alloc(2 * capacity, &contexts) != GJXL_OK ||
      gjxl_amd_host_alloc(4 * (groups + 1), &offsets) != GJXL_OK) {
    return Fail("gjxl_amd_host_alloc", GJXL_ERROR_OUT_OF_MEMORY);
  }
  GJXLAmdImage image{};
  image.width = width;
  image.height = height;
  image.bytes_per_pixel = 3;
  image.pixels = pixels;
  image.pixels_size = bytes;
  image.row_stride_bytes = static_cast<size_t>(width) * 3;
  GJXLAmdTokenOutput tokens{};
  tokens.values = static_cast<uint32_t*>(values);
  tokens.contexts = static_cast<uint16_t*>(contexts);
  tokens.group_offsets = static_cast<uint32_t*>(offsets);
  tokens.capacity = capacity;
  tokens.group_capacity = groups + 1;
  GJXLAmdPlainFrame frame{};
  const auto start = std::chrono::steady_clock::now();
  for (int i = 0; i < repeat; ++i) {
    if ((r = gjxl_amd_plain_path_run(path, &image, &tokens, &frame)) != GJXL_OK) {
      return Fail("gjxl_amd_plain_path_run", r);
    }
  }
  const double ms = std::chrono::duration<double, std::milli>(
                      std::chrono::steady_clock::now() - start).count() / repeat;

  const size_t n = frame.token_count;
  const size_t store = static_cast<size_t>(frame.groups) * 65536;
  const size_t tiles = static_cast<size_t>((frame.block_width + 7) / 8) *
                       ((frame.block_height + 7) / 8);
  bool ok = WriteFile(prefix + ".values", values, 4 * n) &&
            WriteFile(prefix + ".contexts", contexts, 2 * n) &&
            WriteFile(prefix + ".group_offsets", offsets, 4 * (groups + 1)) &&
            Download(device, frame.cells, blocks, prefix + ".cells") &&
            Download(device, frame.quant_field, 4 * blocks, prefix + ".field") &&
            Download(device, frame.raw_quant, 4 * blocks, prefix + ".raw_quant") &&
            Download(device, frame.quantized, 4 * 3 * store, prefix + ".quantized") &&
            Download(device, frame.dc_quantized, 4 * 3 * blocks, prefix + ".dc_quantized") &&
            Download(device, frame.y_to_x, tiles, prefix + ".y_to_x") &&
            Download(device, frame.y_to_b, tiles, prefix + ".y_to_b");
  if (!ok) {
    std::fprintf(stderr, "writing outputs failed: %s\n", gjxl_amd_last_error());
    return 1;
  }
  int32_t stages[64];
  float times[64];
  int32_t count = 0;
  if ((r = gjxl_amd_plain_path_stage_times(path, stages, times, 64, &count)) != GJXL_OK) {
    return Fail("gjxl_amd_plain_path_stage_times", r);
  }
  std::printf("{\"token_count\": %llu, \"scores\": [", static_cast<unsigned long long>(n));
  for (int i = 0; i < frame.evaluations; ++i) {
    std::printf("%s%.9g", i ? ", " : "", static_cast<double>(frame.scores[i]));
  }
  std::printf("], \"x_qm_scale\": %d, \"b_qm_scale\": %d, \"ms_per_run\": %.4f, \"stages\": [",
              frame.x_qm_scale, frame.b_qm_scale, ms);
  for (int i = 0; i < count && i < 64; ++i) {
    std::printf("%s[%d, %.5f]", i ? ", " : "", stages[i], static_cast<double>(times[i]));
  }
  std::printf("]}\n");
  gjxl_amd_plain_path_destroy(path);
  gjxl_amd_host_free(values);
  gjxl_amd_host_free(contexts);
  gjxl_amd_host_free(offsets);
  gjxl_amd_host_free(pixels);
  gjxl_amd_device_destroy(device);
  return 0;
}
