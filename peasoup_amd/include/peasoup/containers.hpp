// containers.hpp - host and device series containers with the reference's
// names and interfaces:
//   TimeSeries, DedispersedTimeSeries, DeviceTimeSeries,
//   ReusableDeviceTimeSeries, DispersionTrials
//       (include/data_types/timeseries.hpp:58-596)
//   FrequencySeries, DeviceFourierSeries, DevicePowerSpectrum, HarmonicSums
//       (include/data_types/fourierseries.hpp:36-188)
// Device buffers come from libpsx (hipMalloc on the current device); RAII
// as in the reference.  The complex element type is psx_c32, which has the
// layout of cufftComplex.
#pragma once
#include <algorithm>
#include <cstddef>
#include <vector>

#include "peasoup/utils.hpp"

// ------------------------------------------------------------ time series
template <class T>
class TimeSeries {
 protected:
  T* data_ptr;
  std::size_t nsamps;
  float tsamp;

 public:
  TimeSeries(T* data_ptr, std::size_t nsamps, float tsamp) : data_ptr(data_ptr), nsamps(nsamps), tsamp(tsamp) {}
  TimeSeries(std::size_t nsamps = 0) : data_ptr(nullptr), nsamps(nsamps), tsamp(0.0f) {}
  virtual ~TimeSeries() {}
  T* get_data() { return data_ptr; }
  void set_data(T* p) { data_ptr = p; }
  std::size_t get_nsamps() { return nsamps; }
  void set_nsamps(std::size_t n) { nsamps = n; }
  float get_tsamp() { return tsamp; }
  void set_tsamp(float t) { tsamp = t; }
};

template <class T>
class DedispersedTimeSeries : public TimeSeries<T> {
 private:
  float dm;

 public:
  DedispersedTimeSeries() : TimeSeries<T>(), dm(0.0f) {}
  DedispersedTimeSeries(T* data_ptr, std::size_t nsamps, float tsamp, float dm)
      : TimeSeries<T>(data_ptr, nsamps, tsamp), dm(dm) {}
  float get_dm() { return dm; }
  void set_dm(float d) { dm = d; }
};

// timeseries.hpp:234-316
template <class OnDeviceType>
class DeviceTimeSeries : public TimeSeries<OnDeviceType> {
 protected:
  void* reduce_work = nullptr;

 public:
  explicit DeviceTimeSeries(std::size_t nsamps) : TimeSeries<OnDeviceType>(nsamps) {
    Utils::device_malloc<OnDeviceType>(&this->data_ptr, nsamps);
  }
  template <class OnHostType>
  explicit DeviceTimeSeries(TimeSeries<OnHostType>& host_tim) : TimeSeries<OnDeviceType>(host_tim.get_nsamps()) {
    Utils::device_malloc<OnDeviceType>(&this->data_ptr, this->nsamps);
    OnHostType* copy_buffer;
    Utils::device_malloc<OnHostType>(&copy_buffer, this->nsamps);
    Utils::h2dcpy(copy_buffer, host_tim.get_data(), this->nsamps);
    convert(copy_buffer, this->nsamps);
    this->tsamp = host_tim.get_tsamp();
    Utils::device_free(copy_buffer);
  }
  DeviceTimeSeries(const DeviceTimeSeries&) = delete;
  DeviceTimeSeries& operator=(const DeviceTimeSeries&) = delete;

  // GPU_remove_baseline over the first nsamps samples (timeseries.hpp:282-287)
  void remove_baseline(std::size_t nsamps = 0) {
    if (nsamps == 0) nsamps = this->nsamps;
    if (!reduce_work) peasoup::check(psx_malloc(&reduce_work, psx_reduce_work_bytes()), "remove_baseline");
    peasoup::check(psx_remove_baseline(this->data_ptr, nsamps, reduce_work, nullptr), "remove_baseline");
  }
  // timeseries.hpp:296-304
  void fill(std::size_t start, std::size_t end, OnDeviceType value) {
    if (end > this->nsamps) peasoup::throw_error("DeviceTimeSeries::fill bad end value requested");
    peasoup::check(psx_fill(this->data_ptr, start, end, value, nullptr), "Error in GPU_fill");
  }
  virtual ~DeviceTimeSeries() {
    if (reduce_work) psx_free(reduce_work);
    psx_free(this->data_ptr);
  }

 protected:
  template <class OnHostType>
  void convert(const OnHostType* d_in, std::size_t n);
};

template <class T>
struct psx_type_code;
template <>
struct psx_type_code<unsigned char> { static const int value = PSX_U8; };
template <>
struct psx_type_code<char> { static const int value = PSX_I8; };
template <>
struct psx_type_code<unsigned int> { static const int value = PSX_U32; };
template <>
struct psx_type_code<float> { static const int value = PSX_F32; };

template <class OnDeviceType>
template <class OnHostType>
void DeviceTimeSeries<OnDeviceType>::convert(const OnHostType* d_in, std::size_t n) {
  peasoup::check(psx_convert_to_f32(d_in, psx_type_code<OnHostType>::value, this->data_ptr, n, nullptr),
                 "Error from device_conversion");
}

// timeseries.hpp:326-374
template <class OnDeviceType, class OnHostType>
class ReusableDeviceTimeSeries : public DeviceTimeSeries<OnDeviceType> {
 private:
  OnHostType* copy_buffer;

 public:
  explicit ReusableDeviceTimeSeries(std::size_t nsamps) : DeviceTimeSeries<OnDeviceType>(nsamps) {
    Utils::device_malloc<OnHostType>(&copy_buffer, this->nsamps);
  }
  void copy_from_host(TimeSeries<OnHostType>& host_tim) {
    std::size_t size = std::min(host_tim.get_nsamps(), this->nsamps);
    this->tsamp = host_tim.get_tsamp();
    Utils::h2dcpy(copy_buffer, host_tim.get_data(), size);
    this->convert(copy_buffer, size);
  }
  ~ReusableDeviceTimeSeries() { psx_free(copy_buffer); }
};

// timeseries.hpp:464-596: host container of all DM trials, one contiguous
// buffer of count x nsamps samples.
template <class T>
class DispersionTrials {
 private:
  std::vector<T> data;
  std::size_t nsamps;
  float tsamp;
  std::vector<float> dm_list;

 public:
  explicit DispersionTrials(float tsamp) : nsamps(0), tsamp(tsamp) {}
  void resize(std::size_t out_nsamps, std::vector<float> const& dm_list_in) {
    dm_list = dm_list_in;
    nsamps = out_nsamps;
    data.resize(nsamps * dm_list.size());
  }
  std::size_t get_count() { return dm_list.size(); }
  std::size_t get_nsamps() { return nsamps; }
  float get_tsamp() { return tsamp; }
  T* get_data_ptr() { return data.data(); }
  std::vector<float>& get_dm_list() { return dm_list; }
  DedispersedTimeSeries<T> operator[](std::size_t idx) {
    return DedispersedTimeSeries<T>(data.data() + idx * nsamps, nsamps, tsamp, dm_list[idx]);
  }
  // timeseries.hpp:508-514
  void get_idx(std::size_t idx, DedispersedTimeSeries<T>& tim, std::size_t /*effective_nsamps*/ = 0) {
    tim.set_data(data.data() + idx * nsamps);
    tim.set_dm(dm_list[idx]);
    tim.set_nsamps(nsamps);
    tim.set_tsamp(tsamp);
  }
};

// ------------------------------------------------------- frequency series
template <class T>
class FrequencySeries {
 protected:
  T* data_ptr;
  unsigned int nbins;
  double bin_width;

 public:
  FrequencySeries() : data_ptr(nullptr), nbins(0), bin_width(0) {}
  FrequencySeries(unsigned int nbins, double bin_width) : data_ptr(nullptr), nbins(nbins), bin_width(bin_width) {}
  FrequencySeries(T* p, unsigned int nbins, double bin_width) : data_ptr(p), nbins(nbins), bin_width(bin_width) {}
  virtual ~FrequencySeries() {}
  T* get_data() { return data_ptr; }
  void set_data(T* p) { data_ptr = p; }
  double get_bin_width() { return bin_width; }
  void set_bin_width(double bw) { bin_width = bw; }
  unsigned int get_nbins() { return nbins; }
  void set_nbins(unsigned int n) { nbins = n; }
};

template <class T>
class DeviceFrequencySeries : public FrequencySeries<T> {
 protected:
  DeviceFrequencySeries(unsigned int nbins, double bin_width) : FrequencySeries<T>(nbins, bin_width) {
    Utils::device_malloc<T>(&this->data_ptr, nbins);
  }
  DeviceFrequencySeries(const DeviceFrequencySeries&) = delete;
  DeviceFrequencySeries& operator=(const DeviceFrequencySeries&) = delete;

 public:
  ~DeviceFrequencySeries() { psx_free(this->data_ptr); }
};

template <class T>
class DeviceFourierSeries : public DeviceFrequencySeries<T> {
 public:
  DeviceFourierSeries(unsigned int nbins, double bin_width) : DeviceFrequencySeries<T>(nbins, bin_width) {}
};

template <class T>
class DevicePowerSpectrum : public DeviceFrequencySeries<T> {
 private:
  unsigned int nh;

 public:
  DevicePowerSpectrum(unsigned int nbins, double bin_width, unsigned int nh = 0)
      : DeviceFrequencySeries<T>(nbins, bin_width), nh(nh) {}
  template <class U>
  explicit DevicePowerSpectrum(FrequencySeries<U>& other, unsigned int nh = 0)
      : DeviceFrequencySeries<T>(other.get_nbins(), other.get_bin_width()), nh(nh) {}
  unsigned int get_nh() { return nh; }
  void set_nh(unsigned int n) { nh = n; }
};

// fourierseries.hpp:162-188: fold k has nh = k+1.
template <class T>
class HarmonicSums {
 private:
  std::vector<DevicePowerSpectrum<T>*> folds;

 public:
  HarmonicSums(DevicePowerSpectrum<T>& fold0, unsigned int nfolds) {
    folds.reserve(nfolds);
    for (unsigned int ii = 0; ii < nfolds; ii++) folds.push_back(new DevicePowerSpectrum<T>(fold0, ii + 1));
  }
  HarmonicSums(const HarmonicSums&) = delete;
  HarmonicSums& operator=(const HarmonicSums&) = delete;
  std::size_t size() { return folds.size(); }
  DevicePowerSpectrum<T>* operator[](int ii) { return folds[ii]; }
  ~HarmonicSums() {
    for (auto* f : folds) delete f;
  }
};
