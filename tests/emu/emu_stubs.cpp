// Symbols of the library's other translation units (the device BVH builder,
// the proposal network) that j3dg.hip calls: the CPU emulation of the batch
// path does not include those kernels, so they report "unsupported".
// Test infrastructure only.
#include "j3dg_internal.h"

namespace j3d {

int device_build_mesh(Ctx* ctx, Mesh&, const float*, uint64_t, const uint32_t*, uint64_t, int, int, double*) {
  return ctx->fail(J3DG_ERR_UNSUPPORTED, "the device BVH builder is not part of the CPU emulation");
}

int Ctx::launch_mlp_frame(const FrameParams&, uint32_t, const DevObject*, bool, bool) {
  return fail(J3DG_ERR_UNSUPPORTED, "the proposal network is not part of the CPU emulation");
}

}  // namespace j3d
