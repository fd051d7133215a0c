// Batched execution of many independent graphs ("jobs", each with its own
// device buffers) as one persistent launch over all their resize tiles,
// captured once into a HIP graph and replayed. The reference processes one
// graph per call on the CPU (zimg.h:550-552); this has no counterpart there.
#pragma once

#include <map>
#include <mutex>
#include <string>
#include <vector>

#include <hip/hip_runtime.h>

#include "device.h"

namespace zamd {

class DeviceJobs {
public:
	enum Flags : unsigned { AUTOTUNE = 1, NO_GRAPH = 2 };

	struct Job {
		const DeviceProgram *prog;
		DeviceBuffers src, dst;
	};

	// Bytes of device scratch the jobs that cannot join the persistent launch need.
	static size_t tmp_size(const std::vector<const DeviceProgram *> &progs);

	DeviceJobs(std::vector<Job> jobs, void *tmp, unsigned flags, void *stream);
	~DeviceJobs();
	DeviceJobs(const DeviceJobs &) = delete;
	DeviceJobs &operator=(const DeviceJobs &) = delete;

	void run(void *stream);
	std::string describe() const;

private:
	std::vector<Job> m_jobs;
	std::vector<size_t> m_tmp_off;
	char *m_tmp = nullptr;
	unsigned m_flags = 0;
	std::vector<char> m_persistent; // per job: all of its steps are in the persistent launch
	void *m_dev = nullptr;          // descriptors then tiles
	const dev::ResizeArg *m_descs = nullptr;
	const dev::JobTile *m_tiles = nullptr;
	unsigned m_ndescs = 0, m_ntiles = 0, m_grid = 0;
	size_t m_lds = 0;
	unsigned m_tuned = 0;
	hipStream_t m_capture = nullptr;
	hipGraph_t m_graph = nullptr;
	hipGraphExec_t m_exec = nullptr;

	static bool eligible(const DeviceProgram &prog, const DeviceBuffers &src, const DeviceBuffers &dst);
	TiledConfig tune(const DeviceProgram &prog, const DeviceProgram::Step &s, const dev::PlaneArg &src, const dev::PlaneArg &dst,
	                 const DeviceProgram::Tables &tabs, hipStream_t stream);
	void enqueue(hipStream_t stream) const;
};

} // namespace zamd
