// Exhaustive check of the fused powf (zimg_amd/csrc/zmath.h, powf_fused<y>,
// with its table of listed inputs) against the reference's unfused powf
// (powf_t<false>, the ARM optimized-routines method without contraction) over
// all 2^32 float bit patterns of x, for every exponent powf_fused accepts.
// Prints one line "y=<bits> mismatches N" per exponent (and the first few
// differing inputs). Test infrastructure, built and run by
// tests/test_powf_fma.py.
#include <algorithm>
#include <atomic>
#include <cstdint>
#include <cstdio>
#include <mutex>
#include <thread>
#include <vector>

#include "zmath.h"

using namespace zamd::zm;

template <uint32_t YB>
void check()
{
	const float y = bitsf(YB);
	const unsigned nt = std::max(1u, std::min(16u, std::thread::hardware_concurrency()));
	std::atomic<unsigned long long> bad{ 0 };
	std::mutex out;
	std::vector<std::thread> th;
	for (unsigned t = 0; t < nt; ++t) {
		th.emplace_back([&, t] {
			for (uint64_t u = t; u <= 0xFFFFFFFFull; u += nt) {
				const float x = bitsf(static_cast<uint32_t>(u));
				const float a = powf_t<false>(x, y), b = powf_fused<YB>(x);
				if (fbits(a) != fbits(b) && !(a != a && b != b)) {
					if (bad.fetch_add(1) < 4) {
						std::lock_guard<std::mutex> lock(out);
						std::printf("  x=%08x reference %08x fused %08x\n", static_cast<unsigned>(u), fbits(a), fbits(b));
					}
				}
			}
		});
	}
	for (auto &t : th)
		t.join();
	std::printf("y=%08x mismatches %llu\n", YB, static_cast<unsigned long long>(bad.load()));
	std::fflush(stdout);
}

int main()
{
	check<ZM_FB(2.4f)>();
	check<ZM_FB(1.0f / 2.4f)>();
	check<ZM_FB(0.45f)>();
	check<ZM_FB(1.0f / 0.45f)>();
	check<ZM_FB(ST2084_M1)>();
	check<ZM_FB(ST2084_M2)>();
	check<ZM_FB(1.0f / ST2084_M1)>();
	check<ZM_FB(1.0f / ST2084_M2)>();
	check<ZM_FB((1.0f - 1.2f) / 1.2f)>();
	check<ZM_FB(1.2f - 1.0f)>();
	return 0;
}
