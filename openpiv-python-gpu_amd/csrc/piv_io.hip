// Host-side input path for pair streams: baseline TIFF decoding into
// caller memory (pinned host buffers from piv_host_alloc), asynchronous
// copies and cross-stream ordering.  The reference reads frames with
// OpenCV's imread(..., IMREAD_ANYDEPTH) in its notebooks
// (openpiv_python_gpu_tutorial.ipynb) and uploads them with cp.asarray
// (gpu_process.py:582-583).
#include <hip/hip_runtime.h>
#include <stdint.h>
#include <stdio.h>
#include <string.h>

#include <algorithm>
#include <string>
#include <vector>

#include "../../include/openpiv_hip.h"

namespace piv {
int capi_fail(int code, const char* fmt, ...);
}

namespace {

using piv::capi_fail;

struct Tiff {
  bool big = false;
  std::vector<uint8_t> bytes;
  uint32_t width = 0, height = 0, bits = 0, samples = 1, compression = 1, planar = 1, rows_per_strip = 0;
  uint32_t sample_format = 1;
  std::vector<uint32_t> strip_off, strip_len;

  uint16_t u16(size_t o) const {
    return big ? (uint16_t)((bytes[o] << 8) | bytes[o + 1]) : (uint16_t)(bytes[o] | (bytes[o + 1] << 8));
  }
  uint32_t u32(size_t o) const {
    return big ? ((uint32_t)bytes[o] << 24) | ((uint32_t)bytes[o + 1] << 16) | ((uint32_t)bytes[o + 2] << 8) | bytes[o + 3]
               : (uint32_t)bytes[o] | ((uint32_t)bytes[o + 1] << 8) | ((uint32_t)bytes[o + 2] << 16) |
                     ((uint32_t)bytes[o + 3] << 24);
  }
};

int load_file(const char* path, std::vector<uint8_t>* out) {
  FILE* f = fopen(path, "rb");
  if (!f) return capi_fail(PIV_E_INVALID, "tiff: cannot open %s", path);
  fseek(f, 0, SEEK_END);
  const long n = ftell(f);
  fseek(f, 0, SEEK_SET);
  if (n < 8) {
    fclose(f);
    return capi_fail(PIV_E_INVALID, "tiff: %s is too short", path);
  }
  out->resize((size_t)n);
  const size_t got = fread(out->data(), 1, (size_t)n, f);
  fclose(f);
  if (got != (size_t)n) return capi_fail(PIV_E_INVALID, "tiff: short read on %s", path);
  return PIV_OK;
}

// Values of an IFD entry (SHORT or LONG, inline or at an offset).
int entry_values(const Tiff& t, size_t e, std::vector<uint32_t>* vals) {
  const uint16_t type = t.u16(e + 2);
  const uint32_t count = t.u32(e + 4);
  const size_t size = type == 3 ? 2 : (type == 4 ? 4 : 0);
  if (!size) return capi_fail(PIV_E_INVALID, "tiff: unsupported field type %u", type);
  if (count > (1u << 24)) return capi_fail(PIV_E_INVALID, "tiff: field too long");
  size_t at = e + 8;
  if (count * size > 4) at = t.u32(e + 8);
  if (at + count * size > t.bytes.size()) return capi_fail(PIV_E_INVALID, "tiff: field out of the file");
  vals->resize(count);
  for (uint32_t k = 0; k < count; ++k) (*vals)[k] = size == 2 ? t.u16(at + 2 * k) : t.u32(at + 4 * k);
  return PIV_OK;
}

int parse(const char* path, Tiff* t) {
  int r = load_file(path, &t->bytes);
  if (r) return r;
  const uint8_t* b = t->bytes.data();
  if (b[0] == 'I' && b[1] == 'I') t->big = false;
  else if (b[0] == 'M' && b[1] == 'M') t->big = true;
  else return capi_fail(PIV_E_INVALID, "tiff: %s has no TIFF byte-order mark", path);
  if (t->u16(2) != 42) return capi_fail(PIV_E_INVALID, "tiff: %s is not a classic TIFF", path);
  const uint32_t ifd = t->u32(4);
  if ((size_t)ifd + 2 > t->bytes.size()) return capi_fail(PIV_E_INVALID, "tiff: bad IFD offset");
  const uint16_t n = t->u16(ifd);
  if ((size_t)ifd + 2 + 12ull * n > t->bytes.size()) return capi_fail(PIV_E_INVALID, "tiff: truncated IFD");
  std::vector<uint32_t> v;
  for (uint16_t k = 0; k < n; ++k) {
    const size_t e = ifd + 2 + 12ull * k;
    const uint16_t tag = t->u16(e);
    switch (tag) {
      case 256: case 257: case 258: case 259: case 273: case 277: case 278: case 279: case 284: case 339:
        r = entry_values(*t, e, &v);
        if (r) return r;
        if (v.empty()) return capi_fail(PIV_E_INVALID, "tiff: empty field %u", tag);
        break;
      default:
        continue;
    }
    switch (tag) {
      case 256: t->width = v[0]; break;
      case 257: t->height = v[0]; break;
      case 258: t->bits = v[0]; break;
      case 259: t->compression = v[0]; break;
      case 273: t->strip_off = v; break;
      case 277: t->samples = v[0]; break;
      case 278: t->rows_per_strip = v[0]; break;
      case 279: t->strip_len = v; break;
      case 284: t->planar = v[0]; break;
      case 339: t->sample_format = v[0]; break;
    }
  }
  if (!t->width || !t->height) return capi_fail(PIV_E_INVALID, "tiff: missing image size");
  if (t->compression != 1) return capi_fail(PIV_E_INVALID, "tiff: compression %u not supported (uncompressed only)",
                                            t->compression);
  if (t->samples != 1) return capi_fail(PIV_E_INVALID, "tiff: %u samples per pixel (grayscale only)", t->samples);
  if (t->bits != 8 && t->bits != 16 && t->bits != 32)
    return capi_fail(PIV_E_INVALID, "tiff: %u bits per sample not supported", t->bits);
  if (t->bits == 32 && t->sample_format != 3) return capi_fail(PIV_E_INVALID, "tiff: 32-bit integer samples");
  if (t->strip_off.empty() || t->strip_off.size() != t->strip_len.size())
    return capi_fail(PIV_E_INVALID, "tiff: strip tables missing (tiled files are not supported)");
  if (!t->rows_per_strip || t->rows_per_strip > t->height) t->rows_per_strip = t->height;
  return PIV_OK;
}

}  // namespace

extern "C" {

int piv_tiff_info(const char* path, int32_t* height, int32_t* width, int32_t* bits) {
  if (!path) return capi_fail(PIV_E_INVALID, "tiff: null path");
  Tiff t;
  int r = parse(path, &t);
  if (r) return r;
  if (height) *height = (int32_t)t.height;
  if (width) *width = (int32_t)t.width;
  if (bits) *bits = (int32_t)t.bits;
  return PIV_OK;
}

int piv_tiff_read(const char* path, void* out, size_t out_bytes) {
  if (!path || !out) return capi_fail(PIV_E_INVALID, "tiff: null argument");
  Tiff t;
  int r = parse(path, &t);
  if (r) return r;
  const size_t bps = t.bits / 8;
  const size_t row = (size_t)t.width * bps;
  const size_t total = row * t.height;
  if (out_bytes < total) return capi_fail(PIV_E_INVALID, "tiff: output buffer %zu < %zu bytes", out_bytes, total);
  uint8_t* dst = static_cast<uint8_t*>(out);
  size_t written = 0;
  for (size_t s = 0; s < t.strip_off.size() && written < total; ++s) {
    const size_t want = std::min<size_t>((size_t)t.rows_per_strip * row, total - written);
    if ((size_t)t.strip_off[s] + want > t.bytes.size() || t.strip_len[s] < want)
      return capi_fail(PIV_E_INVALID, "tiff: strip %zu out of the file", s);
    memcpy(dst + written, t.bytes.data() + t.strip_off[s], want);
    written += want;
  }
  if (written != total) return capi_fail(PIV_E_INVALID, "tiff: strips hold %zu of %zu bytes", written, total);
  if (t.big && bps > 1) {
    for (size_t k = 0; k < total; k += bps)
      for (size_t a = 0, z = bps - 1; a < z; ++a, --z) std::swap(dst[k + a], dst[k + z]);
  }
  return PIV_OK;
}

int piv_host_alloc(void** ptr, size_t bytes) {
  if (!ptr) return capi_fail(PIV_E_INVALID, "null argument");
  *ptr = nullptr;
  if (hipHostMalloc(ptr, bytes ? bytes : 16, hipHostMallocDefault) != hipSuccess)
    return capi_fail(PIV_E_NOMEM, "hipHostMalloc of %zu bytes failed", bytes);
  return PIV_OK;
}

int piv_host_free(void* ptr) {
  if (ptr && hipHostFree(ptr) != hipSuccess) return capi_fail(PIV_E_HIP, "hipHostFree failed");
  return PIV_OK;
}

int piv_memcpy_async(void* dst, const void* src, size_t bytes, int kind, void* stream) {
  static const hipMemcpyKind kinds[3] = {hipMemcpyHostToDevice, hipMemcpyDeviceToHost, hipMemcpyDeviceToDevice};
  if (!dst || !src || kind < 0 || kind > 2) return capi_fail(PIV_E_INVALID, "bad argument");
  hipError_t e = hipMemcpyAsync(dst, src, bytes, kinds[kind], (hipStream_t)stream);
  if (e != hipSuccess) return capi_fail(PIV_E_HIP, "hipMemcpyAsync: %s", hipGetErrorString(e));
  return PIV_OK;
}

int piv_stream_wait_event(void* stream, void* event) {
  if (!event) return capi_fail(PIV_E_INVALID, "null event");
  hipError_t e = hipStreamWaitEvent((hipStream_t)stream, (hipEvent_t)event, 0);
  if (e != hipSuccess) return capi_fail(PIV_E_HIP, "hipStreamWaitEvent: %s", hipGetErrorString(e));
  return PIV_OK;
}

}  // extern "C"
