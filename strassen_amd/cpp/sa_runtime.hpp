// sa_runtime.hpp -- process-wide device context of the drop-in C++ layer.
//
// The reference's headers use MKL implicitly (blaslapack.cpp); the drop-in
// headers use one HIP context per process instead, created on first use on
// device SA_DEVICE (default 0) and released at exit.  If SA_PATH_MODEL names a
// trained path-selector model file, it is loaded into the context and the
// hybrid Strassen entry plans with it (learned()).  Errors from the C ABI
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

inline bool& learned_slot() {
  static bool on = false;
  return on;
}

inline sa_ctx_t context() {
  static std::once_flag once;
  std::call_once(once, [] {
    const char* d = std::getenv("SA_DEVICE");
    check(sa_ctx_create(d ? std::atoi(d) : 0, &context_slot()), "sa_ctx_create");
    const char* m = std::getenv("SA_PATH_MODEL");
    if (m && *m) {
      check(sa_mlp_load(context_slot(), m), "sa_mlp_load");
      learned_slot() = true;
    }
  });
  return context_slot();
}

// true when a path-selector model was loaded from SA_PATH_MODEL
inline bool learned() {
  context();
  return learned_slot();
}

}  // namespace sa_cpp
