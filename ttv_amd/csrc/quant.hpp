// Reduced-precision storage of A: float -> bfloat16 (round to nearest even) and
// float -> OCP FP8 e4m3 with one absmax scale per group of mode-q fibers.
//
// e4m3 (OCP 8-bit floating point, E4M3): sign, 4 exponent bits (bias 7), 3 mantissa bits;
// exponent field 0 encodes subnormals m/8 * 2^-6; S.1111.111 is NaN; the largest finite value
// is S.1111.110 = 448; there is no infinity.  Encoding here rounds to nearest even and
// saturates at +-448.  Decoding on gfx950 uses the hardware converter (v_cvt_pk_f32_fp8),
// which implements the same OCP format.
#pragma once

#include <hip/hip_runtime.h>
#include <stdint.h>

namespace ttvx
{

__host__ __device__ inline uint16_t f32_to_bf16_rne(float x)
{
	uint32_t u;
	__builtin_memcpy(&u, &x, 4);
	if((u & 0x7f800000u) == 0x7f800000u && (u & 0x007fffffu) != 0)   // NaN stays a quiet NaN
		return static_cast<uint16_t>((u >> 16) | 0x0040u);
	u += 0x7fffu + ((u >> 16) & 1u);
	return static_cast<uint16_t>(u >> 16);
}

__host__ __device__ inline uint8_t f32_to_e4m3_rne_sat(float x)
{
	uint32_t u;
	__builtin_memcpy(&u, &x, 4);
	uint8_t const sign = static_cast<uint8_t>((u >> 24) & 0x80u);
	uint32_t const au = u & 0x7fffffffu;
	if(au > 0x7f800000u) return static_cast<uint8_t>(sign | 0x7fu);          // NaN
	float ax;
	__builtin_memcpy(&ax, &au, 4);
	if(ax >= 448.0f) return static_cast<uint8_t>(sign | 0x7eu);             // saturate (and inf)
	if(ax < 0.015625f) {                                                    // below 2^-6: step 2^-9
		float const m = __builtin_rintf(ax * 512.0f);                       // 0..8, ties to even
		return static_cast<uint8_t>(sign | static_cast<uint8_t>(m));        // 8 is the smallest normal
	}
	uint32_t r = au + 0x7ffffu + ((au >> 20) & 1u);                         // keep 3 mantissa bits, RNE
	uint32_t const e = (r >> 23) - 127u + 7u;
	uint32_t const m = (r >> 20) & 7u;
	if(e > 15u || (e == 15u && m == 7u)) return static_cast<uint8_t>(sign | 0x7eu);
	return static_cast<uint8_t>(sign | (e << 3) | m);
}

struct QuantArgs {
	int64_t O, K, I;
	int64_t sAo, sAk, sAi;
	int64_t fps;          // fibers per scale group
	int64_t groups;
};

} // namespace ttvx
