// Learned selection at plan time: host inference of the selector MLP (mlp.hpp) with the
// weights set through ttvx_selector_set_weights, plus a per-shape cache of decisions.
#include <dlfcn.h>

#include <algorithm>
#include <cmath>
#include <cstdio>
#include <cstdlib>
#include <cstring>
#include <string>
#include <map>
#include <memory>
#include <mutex>
#include <tuple>
#include <vector>

#include "../../include/ttvx.h"
#include "mlp.hpp"
#include "runtime.hpp"

namespace ttvx
{

namespace
{
std::mutex g_sel_mutex;
std::shared_ptr<const std::vector<float>> g_weights;
bool g_weights_decided = false;     // set, cleared or default load attempted
// tuning table: measured fastest variant of folded problems (dtype, O, K, I), 16-byte aligned A
using TKey = std::tuple<int, int64_t, int64_t, int64_t>;
std::map<TKey, int> g_table;
bool g_table_decided = false;
// decision cache of folded 3-D problems: (dtype, O, K, I, alignment class, qpos, p, stride class)
// -> variant; the stride class (unit k stride, unit i strides of A and C) is what decides which
// variants can run a folded problem besides its extents
using Key = std::tuple<int, int64_t, int64_t, int64_t, int, int, int, int>;
std::map<Key, int> g_cache;
} // namespace

void mlp_forward_host(const float* w, const float* x, float* y)
{
	float xs[kMlpF], h1[kMlpH], h2[kMlpH];
	for(int i = 0; i < kMlpF; ++i) xs[i] = (x[i] - w[kMlpOffMean + i]) * w[kMlpOffStd + i];
	for(int h = 0; h < kMlpH; ++h) {
		float s = w[kMlpOffB1 + h];
		for(int i = 0; i < kMlpF; ++i) s += w[kMlpOffW1 + static_cast<std::size_t>(h) * kMlpF + i] * xs[i];
		h1[h] = s > 0.f ? s : 0.f;
	}
	for(int h = 0; h < kMlpH; ++h) {
		float s = w[kMlpOffB2 + h];
		for(int i = 0; i < kMlpH; ++i) s += w[kMlpOffW2 + static_cast<std::size_t>(h) * kMlpH + i] * h1[i];
		h2[h] = s > 0.f ? s : 0.f;
	}
	for(int v = 0; v < kMlpV; ++v) {
		float s = w[kMlpOffB3 + v];
		for(int i = 0; i < kMlpH; ++i) s += w[kMlpOffW3 + static_cast<std::size_t>(v) * kMlpH + i] * h2[i];
		y[v] = s;
	}
}

int set_selector_weights(const float* params, std::size_t count)
{
	std::lock_guard<std::mutex> lock(g_sel_mutex);
	g_cache.clear();
	g_weights_decided = true;
	if(count == 0) {
		g_weights.reset();
		return TTVX_OK;
	}
	if(params == nullptr || count != kMlpParams)
		return fail(TTVX_ERR_ARGUMENT, "ttvx: selector weights must hold " + std::to_string(kMlpParams) + " floats");
	g_weights = std::make_shared<const std::vector<float>>(params, params + count);
	return TTVX_OK;
}

namespace
{
// raw little-endian float32 file of exactly kMlpParams values
bool read_weights(const std::string& path, std::vector<float>& w)
{
	std::FILE* fp = std::fopen(path.c_str(), "rb");
	if(fp == nullptr) return false;
	w.assign(kMlpParams + 1, 0.f);
	std::size_t const n = std::fread(w.data(), sizeof(float), w.size(), fp);
	std::fclose(fp);
	if(n != kMlpParams) return false;
	w.resize(kMlpParams);
	for(float x : w)
		if(!std::isfinite(x)) return false;
	return true;
}
} // namespace

// the data/ directory beside the library's lib/ directory
std::string shipped_data_dir()
{
	Dl_info info;
	if(dladdr(reinterpret_cast<void*>(&load_default_selector_weights), &info) == 0 || info.dli_fname == nullptr)
		return std::string();
	std::string lib = info.dli_fname;
	std::size_t const slash = lib.find_last_of('/');
	std::string const dir = slash == std::string::npos ? std::string(".") : lib.substr(0, slash);
	return dir + "/../data/";
}

// raw little-endian int64 rows (dtype, O, K, I, variant)
bool read_table(const std::string& path, std::map<TKey, int>& table)
{
	std::FILE* fp = std::fopen(path.c_str(), "rb");
	if(fp == nullptr) return false;
	std::vector<int64_t> rows;
	int64_t buf[5];
	while(std::fread(buf, sizeof(int64_t), 5, fp) == 5) rows.insert(rows.end(), buf, buf + 5);
	std::fclose(fp);
	table.clear();
	for(std::size_t r = 0; r + 5 <= rows.size(); r += 5) {
		if(rows[r + 4] < 0 || rows[r + 4] >= V_COUNT) return false;
		table[TKey{static_cast<int>(rows[r]), rows[r + 1], rows[r + 2], rows[r + 3]}] = static_cast<int>(rows[r + 4]);
	}
	return !table.empty();
}

// The shipped weights and table were measured on gfx950: they apply when the current device is a
// gfx950 or no GPU is visible (planning queries); elsewhere the fixed rule decides, unless the
// environment names data explicitly.
bool shipped_data_applies()
{
	int n = 0, dev = 0;
	if(hipGetDeviceCount(&n) != hipSuccess || n == 0) {
		(void)hipGetLastError();
		return true;
	}
	hipDeviceProp_t prop;
	if(hipGetDevice(&dev) != hipSuccess || hipGetDeviceProperties(&prop, dev) != hipSuccess) {
		(void)hipGetLastError();
		return true;
	}
	return std::strncmp(prop.gcnArchName, "gfx950", 6) == 0;
}

void load_default_selector_weights()
{
	std::lock_guard<std::mutex> lock(g_sel_mutex);
	if(g_weights_decided && g_table_decided) return;
	std::string const data = shipped_data_applies() ? shipped_data_dir() : std::string();
	if(!g_weights_decided) {
		g_weights_decided = true;
		std::vector<std::string> paths;
		if(const char* env = std::getenv("TTVX_SELECTOR_WEIGHTS")) paths.emplace_back(env);
		if(!data.empty()) paths.push_back(data + "selector_gfx950.bin");
		for(auto const& p : paths) {
			std::vector<float> w;
			if(read_weights(p, w)) {
				g_weights = std::make_shared<const std::vector<float>>(std::move(w));
				g_cache.clear();
				break;
			}
		}
	}
	if(!g_table_decided) {
		g_table_decided = true;
		std::vector<std::string> paths;
		if(const char* env = std::getenv("TTVX_SELECTOR_TABLE")) paths.emplace_back(env);
		if(!data.empty()) paths.push_back(data + "selector_table_gfx950.bin");
		for(auto const& p : paths) {
			std::map<TKey, int> t;
			if(read_table(p, t)) {
				g_table = std::move(t);
				g_cache.clear();
				break;
			}
		}
	}
}

int set_selector_table(const int64_t* rows, std::size_t count)
{
	std::lock_guard<std::mutex> lock(g_sel_mutex);
	g_cache.clear();
	g_table_decided = true;
	g_table.clear();
	if(count == 0) return TTVX_OK;
	if(rows == nullptr) return fail(TTVX_ERR_ARGUMENT, "ttvx: null selector table");
	for(std::size_t r = 0; r < count; ++r) {
		int64_t const* x = rows + 5 * r;
		if(x[4] < 0 || x[4] >= V_COUNT) {
			g_table.clear();
			return fail(TTVX_ERR_ARGUMENT, "ttvx: selector table row " + std::to_string(r) + " names no variant");
		}
		g_table[TKey{static_cast<int>(x[0]), x[1], x[2], x[3]}] = static_cast<int>(x[4]);
	}
	return TTVX_OK;
}

int learned_variant(const Fold& f, const DType& dt, uintptr_t a_addr, unsigned flags)
{
	// problems that do not fold run the strided kernel whatever the network says
	if(!f.three_d) return -1;
	std::shared_ptr<const std::vector<float>> w;
	int const align = static_cast<int>(a_addr % 16);
	int const strides = (f.sAk == 1 ? 1 : 0) | (f.sAi == 1 ? 2 : 0) | (f.sCi == 1 ? 4 : 0);
	Key const key{dt.code, f.O, f.K, f.I, align, f.qpos, f.p, strides};
	{
		std::lock_guard<std::mutex> lock(g_sel_mutex);
		auto it = g_cache.find(key);
		if(it != g_cache.end()) return it->second;
		// measured fastest variant of this folded problem, when the tuning table has it
		if(f.three_d && align == 0) {
			auto t = g_table.find(TKey{dt.code, f.O, f.K, f.I});
			Config cfg;
			if(t != g_table.end() && configure(t->second, f, dt, a_addr, 256, flags, cfg)) {
				if(g_cache.size() > 4096) g_cache.clear();
				g_cache[key] = t->second;
				return t->second;
			}
		}
		w = g_weights;
		if(!w) return -1;
	}
	float x[kMlpF], y[kMlpV];
	features_of(f, dt, a_addr, x);
	mlp_forward_host(w->data(), x, y);
	// argmin of the predicted slowdown over the variants that can run the shape; the strided
	// fallback is never timed for the training set when a dense variant exists, so it is only
	// taken when nothing else can run
	int best = -1;
	Config cfg;
	for(int v = 0; v < kMlpV; ++v) {
		if(v == V_STRIDED || !configure(v, f, dt, a_addr, 256, flags, cfg)) continue;
		if(best < 0 || y[v] < y[best]) best = v;
	}
	if(best < 0 && configure(V_STRIDED, f, dt, a_addr, 256, flags, cfg)) best = V_STRIDED;
	std::lock_guard<std::mutex> lock(g_sel_mutex);
	if(g_cache.size() > 4096) g_cache.clear();
	g_cache[key] = best;
	return best;
}

} // namespace ttvx

extern "C" {
int ttvx_selector_set_weights(const float* params, size_t count)
{
	ttvx::clear_error();
	return ttvx::set_selector_weights(params, count);
}
int ttvx_selector_set_table(const int64_t* rows, size_t count)
{
	ttvx::clear_error();
	return ttvx::set_selector_table(rows, count);
}
size_t ttvx_selector_num_params(void) { return ttvx::kMlpParams; }
}
