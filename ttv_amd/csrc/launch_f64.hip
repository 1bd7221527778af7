// Kernel instantiations for double storage.
#include "launch_impl.hpp"

namespace ttvx
{
int launch_f64(const Config& cfg, const void* a, const void* b, void* c, bool kahan, hipStream_t s)
{
	return launch_typed<double>(cfg, a, b, c, kahan, s);
}
} // namespace ttvx
