// sa_runtime.hpp -- process-wide device context of the drop-in C++ layer.
//
// The reference's headers use MKL implicitly (blaslapack.cpp); the drop-in
// headers use one HIP context per process instead, created on first use on
// device SA_DEVICE (default 0) and released at exit.  Errors from the C ABI
// become std::runtime_error carrying sa_last_error().
#pragma once
#include <cstdio>
#include <cstdlib>
#include <mutex>
#include <stdexcept>
#include <string>

#include "strassen_amd.h"

namespace sa_cpp {

inline void check(int rc, const char* what) {
  if (rc != SA_OK) throw std::runtime_error(std::string(what) + ": " + sa_last_error());
}

inline sa_ctx_t& context_slot() {
  static sa_ctx_t ctx = nullptr;
  return ctx;
}

inline sa_ctx_t context() {
  static std::once_flag once;
  std::call_once(once, [] {
    const char* d = std::getenv("SA_DEVICE");
    check(sa_ctx_create(d ? std::atoi(d) : 0, &context_slot()), "sa_ctx_create");
  });
  return context_slot();
}

}  // namespace sa_cpp
