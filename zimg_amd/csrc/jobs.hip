// Batched jobs: one persistent tiled-resize launch over many graphs, with a
// timing-based tile-shape autotuner and HIP graph capture.
#include "jobs.h"

#include <algorithm>
#include <cstring>

#include "common.h"

namespace zamd {

namespace {

void check_hip(hipError_t e, const char *what)
{
	if (e != hipSuccess)
		fail(ZIMG_ERROR_UNKNOWN, std::string(what) + ": " + hipGetErrorString(e));
}

size_t align_up(size_t x, size_t a) { return (x + a - 1) / a * a; }

dev::PlaneArg plane_of(const DeviceBuffers &b, int idx) { return dev::PlaneArg{ b.ptr[idx], b.stride[idx], b.fstride[idx] }; }

// LDS per workgroup of a persistent launch whose shapes were autotuned: four
// workgroups per CU (the walker's registers allow four as well).
constexpr size_t kJobsLdsCap = 40 * 1024;

// Tile shapes measured once per step shape and process.
std::mutex g_tune_mutex;
std::map<std::string, TiledConfig> g_tuned;

// FNV-1a over a table's windows and taps: two steps share a tuned tile shape
// only when every staged window the kernel forms is the same.
uint64_t table_hash(const FilterTable *t, uint64_t h)
{
	if (!t)
		return h;
	auto mix = [&](const void *p, size_t n) {
		const unsigned char *b = static_cast<const unsigned char *>(p);
		for (size_t k = 0; k < n; ++k)
			h = (h ^ b[k]) * 1099511628211ULL;
	};
	mix(t->left.data(), t->left.size() * sizeof(t->left[0]));
	mix(t->data_i16.data(), t->data_i16.size() * sizeof(t->data_i16[0]));
	mix(&t->filter_width, sizeof(t->filter_width));
	return h;
}

std::string shape_key(const DeviceProgram &prog, const DeviceProgram::Step &s)
{
	char buf[256];
	const auto &nodes = prog.plan().nodes;
	const FilterTable &t1 = *nodes[s.tab[0]].table;
	const FilterTable *t2 = s.passes == 2 ? nodes[s.tab[1]].table.get() : nullptr;
	const uint64_t h = table_hash(t2, table_hash(&t1, 14695981039346656037ULL));
	std::snprintf(buf, sizeof(buf), "%ux%u>%ux%u p%d h%d w%d in%d/%d out%d/%d t%016llx", s.in_w, s.in_h, s.width, s.height, s.passes,
	              s.h_first ? 1 : 0, s.word ? 1 : 0, s.ix[0].type, s.ix[0].op, s.ox[0].type, s.ox[0].op, static_cast<unsigned long long>(h));
	return buf;
}

} // namespace

bool DeviceJobs::eligible(const DeviceProgram &prog, const DeviceBuffers &src, const DeviceBuffers &dst)
{
	for (const auto &s : prog.steps()) {
		if (s.kind != DeviceProgram::Step::RESIZE || !s.word || !s.fast.fwh || s.fast.fwh > dev::kJobsMaxTaps || prog.resize_precision())
			return false;
		if (s.in[0].kind != DeviceProgram::Buf::SRC || s.out[0].kind != DeviceProgram::Buf::DST)
			return false;
		if (!src.ptr[s.in[0].idx] || !dst.ptr[s.out[0].idx])
			return false;
		if (!DeviceProgram::aligned16(plane_of(src, s.in[0].idx)) || !DeviceProgram::aligned16(plane_of(dst, s.out[0].idx)))
			return false;
	}
	return !prog.steps().empty();
}

size_t DeviceJobs::tmp_size(const std::vector<const DeviceProgram *> &progs)
{
	size_t total = 0;
	for (const DeviceProgram *p : progs)
		total += align_up(p->tmp_size(1), 256);
	return total;
}

// Times the cost model's candidate shapes in the persistent walker itself,
// as the batch runs them: the shape's tiles over enough frames of the same
// plane (frame stride 0, so every frame writes the same values to the same
// place) that each workgroup of the walker's grid takes about 4 tiles, with
// the walker's LDS. A one-frame launch of the plane on its own would favour
// small tiles (more workgroups to fill the GPU, more resident per CU), while
// in the walker every tile pays a serial chain of memory round trips and
// barriers. Candidates needing more than kJobsLdsCap are skipped.
TiledConfig DeviceJobs::tune(const DeviceProgram &prog, const DeviceProgram::Step &s, const dev::PlaneArg &src, const dev::PlaneArg &dst,
                             const DeviceProgram::Tables &tabs, hipStream_t stream)
{
	const std::string key = shape_key(prog, s);
	{
		std::lock_guard<std::mutex> lock(g_tune_mutex);
		auto it = g_tuned.find(key);
		if (it != g_tuned.end())
			return it->second;
	}
	int device = 0, cus = 0;
	check_hip(hipGetDevice(&device), "hipGetDevice");
	check_hip(hipDeviceGetAttribute(&cus, hipDeviceAttributeMultiprocessorCount, device), "hipDeviceGetAttribute");
	const unsigned grid = static_cast<unsigned>(cus) * 4;
	hipEvent_t e0, e1;
	check_hip(hipEventCreate(&e0), "hipEventCreate");
	check_hip(hipEventCreate(&e1), "hipEventCreate");
	void *buf = nullptr;
	size_t buf_bytes = 0;
	TiledConfig best = s.fast;
	float best_ms = 0;
	bool have = false;
	try {
		for (const TiledConfig &c : s.fast_options) {
			dev::ResizeArg a = prog.resize_arg(s, src, dst, 1, tabs, &c);
			if (dev::resize_lds_bytes(a) > kJobsLdsCap)
				continue;
			const unsigned gx = (a.dst_w + a.tile_w - 1) / a.tile_w, gy = (a.dst_h + a.tile_h - 1) / a.tile_h;
			const unsigned frames = std::max(1u, (4 * grid + gx * gy - 1) / (gx * gy));
			a.frames = frames;
			a.src.fstride = 0;
			a.dst.fstride = 0;
			std::vector<dev::JobTile> tl;
			tl.reserve(static_cast<size_t>(frames) * gx * gy);
			for (unsigned f = 0; f < frames; ++f)
				for (unsigned by = 0; by < gy; ++by)
					for (unsigned bx = 0; bx < gx; ++bx)
						tl.push_back(dev::JobTile{ 0, bx, by, f });
			const size_t dbytes = align_up(sizeof(dev::ResizeArg), 256);
			const size_t need = dbytes + tl.size() * sizeof(dev::JobTile);
			if (need > buf_bytes) {
				if (buf)
					check_hip(hipFree(buf), "hipFree(autotune)");
				buf = nullptr;
				check_hip(hipMalloc(&buf, need), "hipMalloc(autotune)");
				buf_bytes = need;
			}
			// The previous candidate's launches have completed (event synchronised).
			check_hip(hipMemcpy(buf, &a, sizeof(a), hipMemcpyHostToDevice), "hipMemcpy(autotune)");
			check_hip(hipMemcpy(static_cast<char *>(buf) + dbytes, tl.data(), tl.size() * sizeof(dev::JobTile), hipMemcpyHostToDevice),
			          "hipMemcpy(autotune)");
			const auto *descs = static_cast<const dev::ResizeArg *>(buf);
			const auto *tiles = reinterpret_cast<const dev::JobTile *>(static_cast<char *>(buf) + dbytes);
			const unsigned n = static_cast<unsigned>(tl.size());
			const unsigned g = std::min((n + dev::kJobRun - 1) / dev::kJobRun, grid);
			check_hip(static_cast<hipError_t>(dev::launch_resize_jobs(descs, tiles, n, g, kJobsLdsCap, stream)), "autotune launch");
			check_hip(hipEventRecord(e0, stream), "hipEventRecord");
			for (int r = 0; r < 3; ++r)
				check_hip(static_cast<hipError_t>(dev::launch_resize_jobs(descs, tiles, n, g, kJobsLdsCap, stream)), "autotune launch");
			check_hip(hipEventRecord(e1, stream), "hipEventRecord");
			check_hip(hipEventSynchronize(e1), "hipEventSynchronize");
			float ms = 0;
			check_hip(hipEventElapsedTime(&ms, e0, e1), "hipEventElapsedTime");
			// Time per output sample: candidates differ in how many frames fill the grid.
			ms /= static_cast<float>(frames);
			if (!have || ms < best_ms) {
				best_ms = ms;
				best = c;
				have = true;
			}
		}
	} catch (...) {
		if (buf)
			(void)hipFree(buf);
		(void)hipEventDestroy(e0);
		(void)hipEventDestroy(e1);
		throw;
	}
	if (buf)
		check_hip(hipFree(buf), "hipFree(autotune)");
	(void)hipEventDestroy(e0);
	(void)hipEventDestroy(e1);
	++m_tuned;
	std::lock_guard<std::mutex> lock(g_tune_mutex);
	g_tuned[key] = best;
	return best;
}

DeviceJobs::DeviceJobs(std::vector<Job> jobs, void *tmp, unsigned flags, void *stream) :
	m_jobs(std::move(jobs)),
	m_tmp(static_cast<char *>(tmp)),
	m_flags(flags)
{
	hipStream_t hs = static_cast<hipStream_t>(stream);
	std::vector<dev::ResizeArg> descs;
	std::vector<dev::JobTile> tiles;
	std::vector<double> tile_cost;
	size_t off = 0;
	for (const Job &j : m_jobs) {
		const DeviceProgram &prog = *j.prog;
		// Every table the job's launches read (its own launch sequence included)
		// is uploaded here, outside the capture below.
		prog.prepare_current_device();
		const DeviceProgram::Tables &tabs = prog.tables_for_current_device();
		m_tmp_off.push_back(off);
		const bool pers = eligible(prog, j.src, j.dst);
		m_persistent.push_back(pers ? 1 : 0);
		if (!pers) {
			off += align_up(prog.tmp_size(1), 256);
			continue;
		}
		for (const auto &s : prog.steps()) {
			const dev::PlaneArg src = plane_of(j.src, s.in[0].idx), dst = plane_of(j.dst, s.out[0].idx);
			const TiledConfig cfg = (m_flags & AUTOTUNE) ? tune(prog, s, src, dst, tabs, hs) : s.fast;
			const dev::ResizeArg a = prog.resize_arg(s, src, dst, 1, tabs, &cfg);
			const unsigned gx = (a.dst_w + a.tile_w - 1) / a.tile_w, gy = (a.dst_h + a.tile_h - 1) / a.tile_h;
			const double per_tile = cfg.cost / (static_cast<double>(gx) * gy);
			const unsigned d = static_cast<unsigned>(descs.size());
			descs.push_back(a);
			m_lds = std::max(m_lds, dev::resize_lds_bytes(a));
			for (unsigned by = 0; by < gy; ++by) {
				for (unsigned bx = 0; bx < gx; ++bx) {
					tiles.push_back(dev::JobTile{ d, bx, by, 0 });
					tile_cost.push_back(per_tile);
				}
			}
		}
	}
	if (off && !m_tmp)
		fail(ZIMG_ERROR_ILLEGAL_ARGUMENT, "device tmp buffer required");

	// Most expensive tiles first; workgroups take runs of kJobRun tiles round-robin.
	std::vector<size_t> order(tiles.size());
	for (size_t k = 0; k < order.size(); ++k)
		order[k] = k;
	std::stable_sort(order.begin(), order.end(), [&](size_t x, size_t y) { return tile_cost[x] > tile_cost[y]; });
	std::vector<dev::JobTile> sorted;
	sorted.reserve(tiles.size());
	for (size_t k : order)
		sorted.push_back(tiles[k]);

	m_ndescs = static_cast<unsigned>(descs.size());
	m_ntiles = static_cast<unsigned>(sorted.size());
	if (m_ntiles) {
		const size_t dbytes = align_up(descs.size() * sizeof(dev::ResizeArg), 256);
		const size_t bytes = dbytes + sorted.size() * sizeof(dev::JobTile);
		check_hip(hipMalloc(&m_dev, bytes), "hipMalloc(jobs)");
		check_hip(hipMemcpy(m_dev, descs.data(), descs.size() * sizeof(dev::ResizeArg), hipMemcpyHostToDevice), "hipMemcpy(jobs)");
		check_hip(hipMemcpy(static_cast<char *>(m_dev) + dbytes, sorted.data(), sorted.size() * sizeof(dev::JobTile), hipMemcpyHostToDevice),
		          "hipMemcpy(jobs)");
		m_descs = static_cast<const dev::ResizeArg *>(m_dev);
		m_tiles = reinterpret_cast<const dev::JobTile *>(static_cast<char *>(m_dev) + dbytes);
		int device = 0, cus = 0;
		check_hip(hipGetDevice(&device), "hipGetDevice");
		check_hip(hipDeviceGetAttribute(&cus, hipDeviceAttributeMultiprocessorCount, device), "hipDeviceGetAttribute");
		const size_t per_cu = std::max<size_t>(1, std::min<size_t>(4, (160 * 1024) / std::max<size_t>(m_lds, 1)));
		m_grid = static_cast<unsigned>(std::min<size_t>((m_ntiles + dev::kJobRun - 1) / dev::kJobRun, static_cast<size_t>(cus) * per_cu));
	}

	if (!(m_flags & NO_GRAPH)) {
		check_hip(hipStreamCreateWithFlags(&m_capture, hipStreamNonBlocking), "hipStreamCreate");
		check_hip(hipStreamBeginCapture(m_capture, hipStreamCaptureModeThreadLocal), "hipStreamBeginCapture");
		try {
			enqueue(m_capture);
		} catch (...) {
			hipGraph_t g = nullptr;
			(void)hipStreamEndCapture(m_capture, &g);
			if (g)
				(void)hipGraphDestroy(g);
			throw;
		}
		check_hip(hipStreamEndCapture(m_capture, &m_graph), "hipStreamEndCapture");
		check_hip(hipGraphInstantiate(&m_exec, m_graph, nullptr, nullptr, 0), "hipGraphInstantiate");
	}
}

DeviceJobs::~DeviceJobs()
{
	if (m_exec)
		(void)hipGraphExecDestroy(m_exec);
	if (m_graph)
		(void)hipGraphDestroy(m_graph);
	if (m_capture)
		(void)hipStreamDestroy(m_capture);
	if (m_dev)
		(void)hipFree(m_dev);
}

void DeviceJobs::enqueue(hipStream_t stream) const
{
	for (size_t k = 0; k < m_jobs.size(); ++k) {
		if (!m_persistent[k])
			m_jobs[k].prog->run(m_jobs[k].src, m_jobs[k].dst, m_tmp ? m_tmp + m_tmp_off[k] : nullptr, 1, stream);
	}
	check_hip(static_cast<hipError_t>(dev::launch_resize_jobs(m_descs, m_tiles, m_ntiles, m_grid, m_lds, stream)), "persistent resize launch");
}

void DeviceJobs::run(void *stream)
{
	if (m_exec)
		check_hip(hipGraphLaunch(m_exec, static_cast<hipStream_t>(stream)), "hipGraphLaunch");
	else
		enqueue(static_cast<hipStream_t>(stream));
}

std::string DeviceJobs::describe() const
{
	size_t pers = 0;
	for (char c : m_persistent)
		pers += c;
	char buf[256];
	std::snprintf(buf, sizeof(buf), "jobs=%zu persistent=%zu descriptors=%u tiles=%u grid=%u lds=%zu tuned=%u graph=%d\n", m_jobs.size(), pers,
	              m_ndescs, m_ntiles, m_grid, m_lds, m_tuned, m_exec ? 1 : 0);
	return buf;
}

} // namespace zamd
