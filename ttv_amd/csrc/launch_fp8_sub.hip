// Kernel instantiations for FP8-e4m3 storage read by realigned loads (row products whose vectors
// start off 4-byte boundaries).
#include "launch_impl.hpp"

namespace ttvx
{
int launch_fp8_sub(const Config& cfg, const void* a, const void* b, void* c, bool kahan, hipStream_t s)
{
	return launch_typed<fp8_sub>(cfg, a, b, c, kahan, s);
}
} // namespace ttvx
