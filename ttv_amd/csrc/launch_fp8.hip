// Kernel instantiations for uint8_t storage.
#include "launch_impl.hpp"

namespace ttvx
{
int launch_fp8(const Config& cfg, const void* a, const void* b, void* c, bool kahan, hipStream_t s)
{
	return launch_typed<uint8_t>(cfg, a, b, c, kahan, s);
}
} // namespace ttvx
