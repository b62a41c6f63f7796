// tm_train_states.hip -- train kernels for state bits other than 8.
//
// Built once per plane count M = state_bits - 1 (-DTM_PLANES=M), each object
// exporting tm_launch_train_planes_<M>.  These machines train with the
// in-kernel epoch scan and the general include_bits path (include_bits 1 is
// one of its cases); tm_launch_train has checked the launch already.  A tile
// is built when a lane's chunk state, (M + 1) x words per lane registers,
// fits in 128 VGPRs.
#include "tm_train_kernel.h"

#ifndef TM_PLANES
#error "build with -DTM_PLANES=<state bits - 1>"
#endif

namespace {

template <int WPL, int G>
int launch_planes(const tm_train_launch &L, hipStream_t st)
{
  if constexpr ((TM_PLANES + 1) * WPL > 128) {
    return (int) hipErrorInvalidValue;
  } else {
    hipLaunchKernelGGL((train_kernel<WPL, G, TM_PLANES, false, false>), dim3(L.total_blocks), dim3(8 * G), 0, st, L);
    return (int) hipGetLastError();
  }
}

} // namespace

#define TM_CAT2(a, b) a##b
#define TM_CAT(a, b) TM_CAT2(a, b)

int TM_CAT(tm_launch_train_planes_, TM_PLANES)(const tm_train_launch &L, uint32_t G, uint32_t W, hipStream_t st)
{
  if (L.qid || L.state_bits != TM_PLANES + 1) return (int) hipErrorInvalidValue;
  switch (G * 100 + W) {
    case 801: return launch_planes<1, 8>(L, st);
    case 802: return launch_planes<2, 8>(L, st);
    case 1601: return launch_planes<1, 16>(L, st);
    case 1602: return launch_planes<2, 16>(L, st);
    case 3201: return launch_planes<1, 32>(L, st);
    case 3202: return launch_planes<2, 32>(L, st);
    case 6401: return launch_planes<1, 64>(L, st);
    case 6402: return launch_planes<2, 64>(L, st);
    case 6404: return launch_planes<4, 64>(L, st);
    case 6408: return launch_planes<8, 64>(L, st);
    case 6416: return launch_planes<16, 64>(L, st);
  }
  return (int) hipErrorInvalidValue;
}
