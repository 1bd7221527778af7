// Kernel instantiations for float storage.
#include "launch_impl.hpp"

namespace ttvx
{
int launch_f32(const Config& cfg, const void* a, const void* b, void* c, bool kahan, hipStream_t s)
{
	return launch_typed<float>(cfg, a, b, c, kahan, s);
}
} // namespace ttvx
