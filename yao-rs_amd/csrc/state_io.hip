// state_io.hip — the yao-state-v1 file format (yao-cli/src/state_io.rs) read into
// and written from a device register.
//
// Format: one line of JSON {"format":"yao-state-v1","num_qubits":n,"dims":[2,...],
// "num_elements":2^n,"dtype":"complex128"} (serde's compact form, fields in struct
// order), then num_elements (re, im) pairs of little-endian f64. The amplitudes are
// streamed between the file and HBM in 64 MiB chunks through two pinned buffers,
// so a copy of chunk k overlaps the file transfer of chunk k+1 and no host copy of
// the whole state is made. complex64 registers are widened / narrowed per chunk.
#include <cstdio>
#include <cstring>
#include <string>
#include <vector>

#include "runtime.h"

namespace yh {

namespace {

constexpr size_t kChunkAmps = size_t(1) << 22;  // 64 MiB of complex128

// ---- minimal JSON reader for the header object ----
struct JsonCursor {
  const std::string& s;
  size_t i = 0;
  explicit JsonCursor(const std::string& str) : s(str) {}
  void ws() {
    while (i < s.size() && (s[i] == ' ' || s[i] == '\t' || s[i] == '\r' || s[i] == '\n')) ++i;
  }
  bool eat(char c) {
    ws();
    if (i < s.size() && s[i] == c) {
      ++i;
      return true;
    }
    return false;
  }
  bool string(std::string& out) {
    ws();
    if (i >= s.size() || s[i] != '"') return false;
    ++i;
    out.clear();
    while (i < s.size() && s[i] != '"') {
      if (s[i] == '\\') {
        if (i + 1 >= s.size()) return false;
        const char e = s[i + 1];
        out.push_back(e == 'n' ? '\n' : e == 't' ? '\t' : e);
        i += 2;
      } else {
        out.push_back(s[i++]);
      }
    }
    if (i >= s.size()) return false;
    ++i;
    return true;
  }
  bool uint(uint64_t& v) {
    ws();
    if (i >= s.size() || s[i] < '0' || s[i] > '9') return false;
    v = 0;
    while (i < s.size() && s[i] >= '0' && s[i] <= '9') {
      const uint64_t d = uint64_t(s[i++] - '0');
      if (v > (UINT64_MAX - d) / 10) return false;
      v = v * 10 + d;
    }
    return true;
  }
  // skips any JSON value (for fields the header does not define); values nested
  // deeper than kMaxDepth are rejected, as serde_json's recursion limit (128) does
  static constexpr int kMaxDepth = 128;
  bool skip(int depth = 0) {
    ws();
    if (i >= s.size() || depth >= kMaxDepth) return false;
    const char c = s[i];
    if (c == '"') {
      std::string t;
      return string(t);
    }
    if (c == '{' || c == '[') {
      const char close = c == '{' ? '}' : ']';
      ++i;
      if (eat(close)) return true;
      do {
        if (c == '{') {
          std::string k;
          if (!string(k) || !eat(':')) return false;
        }
        if (!skip(depth + 1)) return false;
      } while (eat(','));
      return eat(close);
    }
    while (i < s.size() && s[i] != ',' && s[i] != '}' && s[i] != ']' && s[i] != ' ') ++i;
    return true;
  }
};

struct Header {
  std::string format, dtype;
  uint64_t num_qubits = 0, num_elements = 0;
  std::vector<uint64_t> dims;
  bool has_format = false, has_nq = false, has_dims = false, has_ne = false, has_dtype = false;
};

int parse_header(const std::string& line, Header& h) {
  JsonCursor c(line);
  if (!c.eat('{')) return fail(YH_ERR_FORMAT, "Failed to parse state header");
  if (!c.eat('}')) {
    do {
      std::string key;
      if (!c.string(key) || !c.eat(':')) return fail(YH_ERR_FORMAT, "Failed to parse state header");
      bool ok = true;
      if (key == "format") {
        ok = c.string(h.format);
        h.has_format = true;
      } else if (key == "dtype") {
        ok = c.string(h.dtype);
        h.has_dtype = true;
      } else if (key == "num_qubits") {
        ok = c.uint(h.num_qubits);
        h.has_nq = true;
      } else if (key == "num_elements") {
        ok = c.uint(h.num_elements);
        h.has_ne = true;
      } else if (key == "dims") {
        h.dims.clear();
        ok = c.eat('[');
        if (ok && !c.eat(']')) {
          do {
            uint64_t d = 0;
            ok = c.uint(d);
            h.dims.push_back(d);
          } while (ok && c.eat(','));
          ok = ok && c.eat(']');
        }
        h.has_dims = true;
      } else {
        ok = c.skip();
      }
      if (!ok) return fail(YH_ERR_FORMAT, "Failed to parse state header");
    } while (c.eat(','));
    if (!c.eat('}')) return fail(YH_ERR_FORMAT, "Failed to parse state header");
  }
  if (!(h.has_format && h.has_nq && h.has_dims && h.has_ne && h.has_dtype))
    return fail(YH_ERR_FORMAT, "Failed to parse state header: missing field");
  return YH_OK;
}

std::string dims_text(const std::vector<uint64_t>& d) {
  std::string s = "[";
  for (size_t k = 0; k < d.size(); ++k) s += (k ? ", " : "") + std::to_string(d[k]);
  return s + "]";
}

// The checks of read_state_from_reader (state_io.rs:63-88), in its order; the
// register length check is ArrayReg::from_vec's.
int check_header(const Header& h) {
  if (h.format != "yao-state-v1")
    return fail(YH_ERR_FORMAT, "Unknown state format: " + h.format + " (expected yao-state-v1)");
  if (h.dtype != "complex128") return fail(YH_ERR_FORMAT, "Unsupported dtype: " + h.dtype + " (expected complex128)");
  uint64_t prod = 1;
  bool overflow = false;
  for (uint64_t d : h.dims) {
    if (d != 0 && prod > UINT64_MAX / d) overflow = true;
    prod *= d;
  }
  if (overflow || h.num_elements != prod)
    return fail(YH_ERR_FORMAT, "Header mismatch: num_elements=" + std::to_string(h.num_elements) + " but dims " +
                                   dims_text(h.dims) + " imply " + (overflow ? std::string("more") : std::to_string(prod)) +
                                   " elements");
  for (uint64_t d : h.dims)
    if (d != 2) return fail(YH_ERR_FORMAT, "Only qubit (d=2) states are supported, got dims " + dims_text(h.dims));
  if (h.num_qubits > 40 || h.num_elements != (uint64_t(1) << h.num_qubits))
    return fail(YH_ERR_FORMAT, "from_vec: state length must be 2^num_qubits");
  return YH_OK;
}

int read_header_line(std::FILE* f, std::string& line) {
  line.clear();
  int ch;
  while ((ch = std::fgetc(f)) != EOF && ch != '\n') {
    line.push_back(char(ch));
    if (line.size() > (size_t(1) << 20)) return fail(YH_ERR_FORMAT, "Failed to parse state header: line too long");
  }
  if (line.empty() && ch == EOF) return fail(YH_ERR_IO, "Failed to read state header");
  return YH_OK;
}

struct File {
  std::FILE* f = nullptr;
  ~File() {
    if (f) std::fclose(f);
  }
};

}  // namespace

}  // namespace yh

using namespace yh;

extern "C" {

int yh_state_read_header(const char* path, int* num_qubits) {
  if (!path || !num_qubits) return fail(YH_ERR_INVALID_ARG, "null argument");
  File fh;
  fh.f = std::fopen(path, "rb");
  if (!fh.f) return fail(YH_ERR_IO, std::string("Failed to open ") + path);
  std::string line;
  YH_TRY(read_header_line(fh.f, line));
  Header h;
  YH_TRY(parse_header(line, h));
  YH_TRY(check_header(h));
  *num_qubits = int(h.num_qubits);
  return YH_OK;
}

int yh_state_write(const yh_reg* reg, const char* path) {
  if (!reg || !path) return fail(YH_ERR_INVALID_ARG, "null argument");
  const uint64_t n = uint64_t(1) << reg->nbits;
  std::string hdr = "{\"format\":\"yao-state-v1\",\"num_qubits\":" + std::to_string(reg->nbits) + ",\"dims\":[";
  for (int k = 0; k < reg->nbits; ++k) hdr += k ? ",2" : "2";
  hdr += "],\"num_elements\":" + std::to_string(n) + ",\"dtype\":\"complex128\"}\n";
  File fh;
  fh.f = std::fopen(path, "wb");
  if (!fh.f) return fail(YH_ERR_IO, std::string("Failed to create ") + path);
  if (std::fwrite(hdr.data(), 1, hdr.size(), fh.f) != hdr.size()) return fail(YH_ERR_IO, "write failed");
  const size_t chunk = size_t(std::min<uint64_t>(n, kChunkAmps));
  PinBuf pin[2];
  YH_TRY(pin[0].ensure(chunk * reg->amp_bytes()));
  YH_TRY(pin[1].ensure(chunk * reg->amp_bytes()));
  std::vector<double> wide(reg->dtype == YH_C64 ? 2 * chunk : 0);
  hipStream_t st = reg->w.stream;
  hipEvent_t ev[2];
  YH_HIP(hipEventCreateWithFlags(&ev[0], hipEventDisableTiming));
  YH_HIP(hipEventCreateWithFlags(&ev[1], hipEventDisableTiming));
  const unsigned char* src = static_cast<const unsigned char*>(reg->state.p);
  const uint64_t nchunks = (n + chunk - 1) / chunk;
  int status = YH_OK;
  auto issue = [&](uint64_t c) -> int {
    const uint64_t a0 = c * chunk, na = std::min<uint64_t>(chunk, n - a0);
    YH_HIP(hipMemcpyAsync(pin[c & 1].p, src + a0 * reg->amp_bytes(), na * reg->amp_bytes(), hipMemcpyDeviceToHost,
                          st));
    YH_HIP(hipEventRecord(ev[c & 1], st));
    return YH_OK;
  };
  status = issue(0);
  for (uint64_t c = 0; c < nchunks && status == YH_OK; ++c) {
    if (c + 1 < nchunks) status = issue(c + 1);  // the other buffer was written out last iteration
    if (status != YH_OK) break;
    if (hipEventSynchronize(ev[c & 1]) != hipSuccess) {
      status = fail(YH_ERR_HIP, "copy to host failed");
      break;
    }
    const uint64_t na = std::min<uint64_t>(chunk, n - c * chunk);
    const void* out = pin[c & 1].p;
    if (reg->dtype == YH_C64) {
      const float* f = static_cast<const float*>(pin[c & 1].p);
      for (size_t k = 0; k < 2 * na; ++k) wide[k] = double(f[k]);
      out = wide.data();
    }
    if (std::fwrite(out, 16, na, fh.f) != na) status = fail(YH_ERR_IO, "write failed");
  }
  (void)hipStreamSynchronize(st);
  (void)hipEventDestroy(ev[0]);
  (void)hipEventDestroy(ev[1]);
  if (status != YH_OK) return status;
  if (std::fflush(fh.f) != 0) return fail(YH_ERR_IO, "write failed");
  return YH_OK;
}

int yh_state_read(const char* path, int dtype, yh_reg** out) {
  if (!path || !out) return fail(YH_ERR_INVALID_ARG, "null argument");
  if (dtype != YH_C128 && dtype != YH_C64) return fail(YH_ERR_DTYPE, "unknown dtype");
  File fh;
  fh.f = std::fopen(path, "rb");
  if (!fh.f) return fail(YH_ERR_IO, std::string("Failed to open ") + path);
  std::string line;
  YH_TRY(read_header_line(fh.f, line));
  Header h;
  YH_TRY(parse_header(line, h));
  YH_TRY(check_header(h));
  yh_reg* r = nullptr;
  YH_TRY(yh_reg_create(int(h.num_qubits), dtype, &r));
  const uint64_t n = h.num_elements;
  const size_t chunk = size_t(std::min<uint64_t>(n, kChunkAmps));
  PinBuf pin[2];
  int status = pin[0].ensure(chunk * r->amp_bytes());
  if (status == YH_OK) status = pin[1].ensure(chunk * r->amp_bytes());
  std::vector<double> wide(dtype == YH_C64 ? 2 * chunk : 0);
  hipEvent_t ev[2] = {nullptr, nullptr};
  if (status == YH_OK && (hipEventCreateWithFlags(&ev[0], hipEventDisableTiming) != hipSuccess ||
                          hipEventCreateWithFlags(&ev[1], hipEventDisableTiming) != hipSuccess))
    status = fail(YH_ERR_HIP, "event creation failed");
  unsigned char* dst = static_cast<unsigned char*>(r->state.p);
  const uint64_t nchunks = (n + chunk - 1) / chunk;
  for (uint64_t c = 0; c < nchunks && status == YH_OK; ++c) {
    // the copy that last used this buffer (chunk c - 2) must have finished
    if (c >= 2 && hipEventSynchronize(ev[c & 1]) != hipSuccess) {
      status = fail(YH_ERR_HIP, "copy to device failed");
      break;
    }
    const uint64_t na = std::min<uint64_t>(chunk, n - c * chunk);
    void* in = dtype == YH_C64 ? static_cast<void*>(wide.data()) : pin[c & 1].p;
    if (std::fread(in, 16, na, fh.f) != na) {
      status = fail(YH_ERR_IO, "Failed to read state data");
      break;
    }
    if (dtype == YH_C64) {
      float* f = static_cast<float*>(pin[c & 1].p);
      for (size_t k = 0; k < 2 * na; ++k) f[k] = float(wide[k]);
    }
    if (hipMemcpyAsync(dst + c * chunk * r->amp_bytes(), pin[c & 1].p, na * r->amp_bytes(), hipMemcpyHostToDevice,
                       r->w.stream) != hipSuccess ||
        hipEventRecord(ev[c & 1], r->w.stream) != hipSuccess)
      status = fail(YH_ERR_HIP, "copy to device failed");
  }
  if (hipStreamSynchronize(r->w.stream) != hipSuccess && status == YH_OK)
    status = fail(YH_ERR_HIP, "copy to device failed");
  if (ev[0]) (void)hipEventDestroy(ev[0]);
  if (ev[1]) (void)hipEventDestroy(ev[1]);
  if (status != YH_OK) {
    yh_reg_destroy(r);
    return status;
  }
  *out = r;
  return YH_OK;
}

}  // extern "C"
