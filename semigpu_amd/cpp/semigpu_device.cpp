#include "semigpu_device.h"

#include <cstdlib>

namespace semigpu {

sg_context* context() {
  static sg_context* ctx = [] {
    const char* dev = std::getenv("SEMIGPU_DEVICE");
    sg_context* c = nullptr;
    check(sg_ctx_create(dev ? std::atoi(dev) : 0, &c), "sg_ctx_create");
    return c;
  }();
  return ctx;
}

void check(int rc, const char* what) {
  if (rc != SG_OK) throw std::runtime_error(std::string(what) + ": " + sg_last_error());
}

}  // namespace semigpu
