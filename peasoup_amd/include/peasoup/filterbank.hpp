// filterbank.hpp - SIGPROC filterbanks and dedispersion with the reference's
// class names and methods, over libpsx:
//   SigprocHeader / read_header   include/data_types/header.hpp:171-403
//   SigprocFilterbank             include/data_types/filterbank.hpp:248-361
//   Dedisperser                   include/transforms/dedisperser.hpp:13-186
// The filterbank keeps the file open and loads gulps of packed samples on
// the host; Dedisperser dedisperses a gulp on the device and fills a host
// DispersionTrials, as the reference's does with libdedisp.
#pragma once
#include <algorithm>
#include <cstdlib>
#include <fstream>
#include <string>
#include <vector>

#include "peasoup/containers.hpp"
#include "peasoup/errors.hpp"

// header.hpp:171-215; the strings and values of a psx_sigproc_header
struct SigprocHeader : public psx_sigproc_header {
  SigprocHeader() : psx_sigproc_header() {}
};

// header.hpp:339-403 (returns false if the file is not a SIGPROC filterbank)
inline bool read_header(const std::string& filename, SigprocHeader& header) {
  return psx_sigproc_read_header(filename.c_str(), &header) == PSX_OK;
}

class SigprocFilterbank {
 public:
  // filterbank.hpp:259-297: segment [start, start + nsamples), the whole file
  // when nsamples is 0 or runs past its end
  SigprocFilterbank(const std::string& filename, std::size_t segment_start_sample = 0,
                    std::size_t segment_nsamples = 0)
      : filename(filename) {
    peasoup::check(psx_sigproc_read_header(filename.c_str(), &header), "SigprocFilterbank");
    total_nsamps = (std::size_t)header.nsamples;
    if (segment_start_sample > total_nsamps) peasoup::throw_error("Invalid start sample provided.");
    start_sample = segment_start_sample;
    end_sample = (segment_nsamples == 0 || segment_start_sample + segment_nsamples > total_nsamps)
                     ? total_nsamps
                     : segment_start_sample + segment_nsamples;
    effective_nsamps = end_sample - start_sample;
  }

  // filterbank.hpp:313-333: returns the whole samples read
  std::size_t load_gulp(std::size_t gulp_start_sample, std::size_t gulp_nsamples) {
    data.resize(gulp_nsamples * (std::size_t)header.nchans * header.nbits / 8);
    uint64_t got = 0;
    peasoup::check(psx_sigproc_read_samples(filename.c_str(), &header, gulp_start_sample, gulp_nsamples,
                                            data.data(), &got),
                   "load_gulp");
    data.resize((std::size_t)got * header.nchans * header.nbits / 8);
    return (std::size_t)got;
  }

  const unsigned char* get_data() const { return data.data(); }
  SigprocHeader get_header() const { return header; }
  float get_tsamp() const { return (float)header.tsamp; }
  float get_fch1() const { return (float)header.fch1; }
  float get_foff() const { return (float)header.foff; }
  float get_cfreq() const { return psx_sigproc_cfreq(&header); }
  unsigned int get_nchans() const { return (unsigned int)header.nchans; }
  unsigned int get_nbits() const { return (unsigned int)header.nbits; }
  std::size_t get_nsamps() const { return total_nsamps; }
  std::size_t get_start_sample() const { return start_sample; }
  std::size_t get_end_sample() const { return end_sample; }
  std::size_t get_effective_nsamps() const { return effective_nsamps; }

 private:
  std::string filename;
  SigprocHeader header;
  std::vector<unsigned char> data;
  std::size_t total_nsamps = 0, start_sample = 0, end_sample = 0, effective_nsamps = 0;
};

class Dedisperser {
 public:
  explicit Dedisperser(SigprocFilterbank& filterbank, unsigned int /*num_gpus*/ = 1)
      : filterbank(filterbank), killmask(filterbank.get_nchans(), 1) {}

  void set_dm_list(const float* dm_list_ptr, unsigned int ndms) { dm_list.assign(dm_list_ptr, dm_list_ptr + ndms); }
  void set_dm_list(const std::vector<float>& dms) { dm_list = dms; }
  std::vector<float> get_dm_list() const { return dm_list; }

  // dedisperser.hpp:62-70 (width in microseconds)
  void generate_dm_list(float dm_start, float dm_end, float width, float tolerance) {
    uint32_t n = 0;
    peasoup::check(psx_dedisp_dm_list(dm_start, dm_end, filterbank.get_tsamp(), width, filterbank.get_fch1(),
                                      filterbank.get_foff(), filterbank.get_nchans(), tolerance, nullptr, 0, &n),
                   "generate_dm_list");
    dm_list.resize(n);
    peasoup::check(psx_dedisp_dm_list(dm_start, dm_end, filterbank.get_tsamp(), width, filterbank.get_fch1(),
                                      filterbank.get_foff(), filterbank.get_nchans(), tolerance, dm_list.data(), n,
                                      &n),
                   "generate_dm_list");
  }

  void set_killmask(const std::vector<int>& mask) {
    if (mask.size() != filterbank.get_nchans()) peasoup::throw_error("killmask size differs from nchans");
    killmask.assign(mask.begin(), mask.end());
  }
  // dedisperser.hpp:78-101: one integer per line; a mask of the wrong
  // length is ignored with a warning, as the reference does
  void set_killmask(const std::string& filename) {
    std::ifstream infile(filename.c_str());
    if (!infile.good()) peasoup::throw_error("cannot open killmask " + filename);
    std::vector<int> mask;
    std::string line;
    while (mask.size() < filterbank.get_nchans() && std::getline(infile, line)) mask.push_back(std::atoi(line.c_str()));
    if (mask.size() == filterbank.get_nchans()) killmask.assign(mask.begin(), mask.end());
  }

  std::size_t get_max_delay() const {
    std::vector<float> table(filterbank.get_nchans());
    peasoup::check(psx_dedisp_delay_table(filterbank.get_nchans(), filterbank.get_tsamp(), filterbank.get_fch1(),
                                          filterbank.get_foff(), table.data()),
                   "delay table");
    return (std::size_t)psx_dedisp_max_delay(dm_list.data(), (uint32_t)dm_list.size(), table.data(),
                                             (uint32_t)table.size());
  }

  // dedisperser.hpp:105-186: gulps of at least twice the maximum delay,
  // each dedispersed on the device and appended to trials
  void dedisperse(DispersionTrials<float>& trials, std::size_t input_start_sample, std::size_t input_end_sample,
                  std::size_t gulp) {
    if (dm_list.empty()) peasoup::throw_error("dedisperse: empty DM list");
    const std::size_t max_delay = get_max_delay();
    if (input_end_sample <= input_start_sample + max_delay) peasoup::throw_error("dedisperse: too few samples");
    gulp = std::max(gulp, 2 * max_delay);
    const std::size_t total_out = input_end_sample - input_start_sample - max_delay;
    trials.resize(total_out, dm_list);
    const uint32_t nchans = filterbank.get_nchans(), nbits = filterbank.get_nbits();
    DeviceBuffer<uint8_t> d_in(gulp * nchans * nbits / 8);
    DeviceBuffer<float> d_out(dm_list.size() * (gulp - max_delay));
    DeviceBuffer<uint8_t> d_work(psx_dedisp_work_bytes(gulp, nchans, (uint32_t)dm_list.size()));
    std::vector<float> h_out;
    std::size_t in_cur = input_start_sample, out_cur = 0;
    while (in_cur < input_end_sample - max_delay) {
      const std::size_t want = std::min(gulp, input_end_sample - in_cur);
      const std::size_t got = filterbank.load_gulp(in_cur, want);
      if (got <= max_delay) peasoup::throw_error("Failed to read data during dedispersion");
      const std::size_t nout = got - max_delay;
      const std::size_t ncopy = std::min(nout, total_out - out_cur);
      peasoup::check(psx_memcpy_h2d(d_in.get(), filterbank.get_data(), got * nchans * nbits / 8, nullptr), "h2d");
      peasoup::check(psx_dedisperse(d_in.get(), got, nchans, nbits, dm_list.data(), (uint32_t)dm_list.size(),
                                    killmask.data(), filterbank.get_tsamp(), filterbank.get_fch1(),
                                    filterbank.get_foff(), d_out.get(), nout, d_work.get(), nullptr),
                     "dedisperse");
      h_out.resize(dm_list.size() * nout);
      peasoup::check(psx_memcpy_d2h(h_out.data(), d_out.get(), h_out.size() * sizeof(float), nullptr), "d2h");
      float* dst = trials.get_data_ptr();
      for (std::size_t d = 0; d < dm_list.size(); ++d)
        std::copy(h_out.begin() + d * nout, h_out.begin() + d * nout + ncopy, dst + d * total_out + out_cur);
      in_cur += nout;
      out_cur += ncopy;
    }
  }

 private:
  template <class T>
  struct DeviceBuffer {
    void* p = nullptr;
    explicit DeviceBuffer(std::size_t n) { peasoup::check(psx_malloc(&p, std::max<std::size_t>(n, 1) * sizeof(T)), "alloc"); }
    ~DeviceBuffer() { psx_free(p); }
    T* get() const { return static_cast<T*>(p); }
  };

  SigprocFilterbank& filterbank;
  std::vector<float> dm_list;
  std::vector<int32_t> killmask;
};
