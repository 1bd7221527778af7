// Kernel instantiations for BF16 storage read by realigned loads (row products whose vectors
// start off 4-byte boundaries).
#include "launch_impl.hpp"

namespace ttvx
{
int launch_bf16_sub(const Config& cfg, const void* a, const void* b, void* c, bool kahan, hipStream_t s)
{
	return launch_typed<bf16_sub>(cfg, a, b, c, kahan, s);
}
} // namespace ttvx
