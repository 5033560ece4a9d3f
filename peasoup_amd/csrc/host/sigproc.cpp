// sigproc.cpp - SIGPROC filterbank header and sample reading
// (include/data_types/header.hpp:315-403, include/data_types/filterbank.hpp:
// 248-333).  A header is a run of length-prefixed keywords, each followed by
// its binary value (int, double, one byte, or another length-prefixed string
// for source_name / rawdatafile), from HEADER_START to HEADER_END; the
// samples follow, time-major, nchans * nbits / 8 bytes per sample.
#include <cstdio>
#include <cstring>
#include <string>

#include "common.hpp"

namespace {

struct File {
  FILE* f = nullptr;
  explicit File(const char* path) : f(std::fopen(path, "rb")) {}
  ~File() {
    if (f) std::fclose(f);
  }
};

// One length-prefixed string; false past the end or on an implausible length
// (header.hpp:315-325 accepts 1..79 characters).
bool read_string(FILE* f, std::string& s) {
  int32_t len = 0;
  if (std::fread(&len, sizeof(len), 1, f) != 1 || len <= 0 || len >= 80) return false;
  char buf[80];
  if (std::fread(buf, 1, (size_t)len, f) != (size_t)len) return false;
  s.assign(buf, (size_t)len);
  return true;
}

template <class T>
bool read_value(FILE* f, T& v) {
  return std::fread(&v, sizeof(T), 1, f) == 1;
}

void copy_name(char (&dst)[80], const std::string& s) {
  std::memset(dst, 0, sizeof(dst));
  std::memcpy(dst, s.data(), std::min<size_t>(s.size(), sizeof(dst) - 1));
}

}  // namespace

extern "C" int psx_sigproc_read_header(const char* path, psx_sigproc_header* h) {
  PSX_REQUIRE(path && h, "psx_sigproc_read_header: null argument");
  File file(path);
  PSX_REQUIRE(file.f, "psx_sigproc_read_header: cannot open %s", path);
  std::memset(h, 0, sizeof(*h));
  std::string key;
  PSX_REQUIRE(read_string(file.f, key) && key == "HEADER_START", "%s is not a SIGPROC filterbank", path);
  bool want_source = false, want_raw = false;
  while (true) {
    PSX_REQUIRE(read_string(file.f, key), "%s: truncated or malformed header", path);
    bool ok = true;
    if (key == "HEADER_END") break;
    if (key == "source_name") want_source = true;
    else if (key == "rawdatafile") want_raw = true;
    else if (key == "az_start") ok = read_value(file.f, h->az_start);
    else if (key == "za_start") ok = read_value(file.f, h->za_start);
    else if (key == "src_raj") ok = read_value(file.f, h->src_raj);
    else if (key == "src_dej") ok = read_value(file.f, h->src_dej);
    else if (key == "tstart") ok = read_value(file.f, h->tstart);
    else if (key == "tsamp") ok = read_value(file.f, h->tsamp);
    else if (key == "period") ok = read_value(file.f, h->period);
    else if (key == "fch1") ok = read_value(file.f, h->fch1);
    else if (key == "foff") ok = read_value(file.f, h->foff);
    else if (key == "refdm") ok = read_value(file.f, h->refdm);
    else if (key == "nchans") ok = read_value(file.f, h->nchans);
    else if (key == "telescope_id") ok = read_value(file.f, h->telescope_id);
    else if (key == "machine_id") ok = read_value(file.f, h->machine_id);
    else if (key == "data_type") ok = read_value(file.f, h->data_type);
    else if (key == "ibeam") ok = read_value(file.f, h->ibeam);
    else if (key == "nbeams") ok = read_value(file.f, h->nbeams);
    else if (key == "nbits") ok = read_value(file.f, h->nbits);
    else if (key == "barycentric") ok = read_value(file.f, h->barycentric);
    else if (key == "pulsarcentric") ok = read_value(file.f, h->pulsarcentric);
    else if (key == "nbins") ok = read_value(file.f, h->nbins);
    else if (key == "nsamples") ok = read_value(file.f, h->nsamples);
    else if (key == "nifs") ok = read_value(file.f, h->nifs);
    else if (key == "npuls") ok = read_value(file.f, h->npuls);
    else if (key == "signed") ok = read_value(file.f, h->signed_data);
    else if (want_source) {
      copy_name(h->source_name, key);
      want_source = false;
    } else if (want_raw) {
      copy_name(h->rawdatafile, key);
      want_raw = false;
    } else {
      std::fprintf(stderr, "psx_sigproc_read_header: unknown parameter %s\n", key.c_str());
    }
    PSX_REQUIRE(ok, "%s: truncated value of %s", path, key.c_str());
  }
  const long hdr = std::ftell(file.f);
  PSX_REQUIRE(hdr > 0, "%s: cannot tell the header size", path);
  h->size = (uint32_t)hdr;
  PSX_REQUIRE(h->nchans > 0 && (h->nbits == 1 || h->nbits == 2 || h->nbits == 4 || h->nbits == 8 ||
                                h->nbits == 16 || h->nbits == 32),
              "%s: nchans %d / nbits %d unsupported", path, h->nchans, h->nbits);
  if (h->nsamples == 0) {
    // header.hpp:394-401: samples from the file size
    PSX_REQUIRE(std::fseek(file.f, 0, SEEK_END) == 0, "%s: seek failed", path);
    const long total = std::ftell(file.f);
    const uint64_t data = (uint64_t)(total - hdr);
    h->nsamples = (int32_t)(data / (uint64_t)h->nchans * 8 / (uint64_t)h->nbits);
  }
  return psx::ok();
}

extern "C" int psx_sigproc_read_samples(const char* path, const psx_sigproc_header* h, uint64_t start,
                                        uint64_t nsamps, uint8_t* out, uint64_t* nread) {
  PSX_REQUIRE(path && h && out && nread, "psx_sigproc_read_samples: null argument");
  PSX_REQUIRE((uint64_t)h->nchans * (uint64_t)h->nbits % 8 == 0, "psx_sigproc_read_samples: partial-byte samples");
  const uint64_t bps = (uint64_t)h->nchans * (uint64_t)h->nbits / 8;  // bytes per sample
  File file(path);
  PSX_REQUIRE(file.f, "psx_sigproc_read_samples: cannot open %s", path);
  PSX_REQUIRE(std::fseek(file.f, (long)(h->size + start * bps), SEEK_SET) == 0, "%s: seek failed", path);
  const size_t got = std::fread(out, 1, (size_t)(nsamps * bps), file.f);
  *nread = (uint64_t)got / bps;
  return psx::ok();
}

extern "C" float psx_sigproc_cfreq(const psx_sigproc_header* h) {
  if (!h) return 0.0f;
  // Filterbank::get_cfreq: float members, float arithmetic
  const float fch1 = (float)h->fch1, foff = (float)h->foff;
  const unsigned int nchans = (unsigned int)h->nchans;
  if (foff < 0) return fch1 + foff * nchans / 2;
  return fch1 - foff * nchans / 2;
}
