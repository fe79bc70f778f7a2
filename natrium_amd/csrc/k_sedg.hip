// SEDG (Min-Lee) operator inside the low-storage RK5 integrator.
#include "nat_collide.hpp"

int nat_sedg_setup(nat_ctx* ctx) {
  ctx->err = "SEDG scheme is not enabled in this build";
  return NAT_ERR_UNSUPPORTED;
}

int nat_sedg_step(nat_ctx* ctx, int, int) {
  ctx->err = "SEDG scheme is not enabled in this build";
  return NAT_ERR_UNSUPPORTED;
}
