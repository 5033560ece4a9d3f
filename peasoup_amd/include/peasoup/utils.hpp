// utils.hpp - device memory helpers (Utils, include/utils/utils.hpp:13-95).
#pragma once
#include <cstddef>
#include <cstdint>

#include "peasoup/errors.hpp"

class Utils {
 public:
  // utils.hpp:15-21
  static unsigned int prev_power_of_two(unsigned int val) {
    unsigned int n = 1;
    while (n * 2 <= val) n *= 2;
    return n;
  }
  template <class T>
  static void device_malloc(T** ptr, std::size_t units) {
    void* p = nullptr;
    peasoup::check(psx_malloc(&p, sizeof(T) * units), "Error from device_malloc");
    *ptr = static_cast<T*>(p);
  }
  template <class T>
  static void host_malloc(T** ptr, std::size_t units) {
    void* p = nullptr;
    peasoup::check(psx_host_malloc(&p, sizeof(T) * units), "Error from host_malloc");
    *ptr = static_cast<T*>(p);
  }
  template <class T>
  static void device_free(T* ptr) {
    peasoup::check(psx_free((void*)ptr), "Error from device_free");
  }
  template <class T>
  static void host_free(T* ptr) {
    peasoup::check(psx_host_free((void*)ptr), "Error from host_free");
  }
  template <class T>
  static void h2dcpy(T* d_ptr, const T* h_ptr, std::size_t units, psx_stream s = nullptr) {
    peasoup::check(psx_memcpy_h2d((void*)d_ptr, (const void*)h_ptr, sizeof(T) * units, s), "Error from h2dcpy");
  }
  template <class T>
  static void d2hcpy(T* h_ptr, const T* d_ptr, std::size_t units, psx_stream s = nullptr) {
    peasoup::check(psx_memcpy_d2h((void*)h_ptr, (const void*)d_ptr, sizeof(T) * units, s), "Error from d2hcpy");
  }
  template <class T>
  static void d2dcpy(T* d_dst, const T* d_src, std::size_t units, psx_stream s = nullptr) {
    peasoup::check(psx_memcpy_d2d((void*)d_dst, (const void*)d_src, sizeof(T) * units, s), "Error from d2dcpy");
  }
  static int gpu_count() {
    int n = 0;
    peasoup::check(psx_device_count(&n), "Error from gpu_count");
    return n;
  }
};
