// Device lowering of a planned graph: fused launch steps, HBM scratch layout,
// per-device coefficient tables, and execution on a HIP stream.
// Replaces the reference's graphengine row scheduler (filtergraph.cpp:96-122).
#pragma once

#include <map>
#include <memory>
#include <mutex>
#include <set>
#include <string>
#include <vector>

#include "graph.h"
#include "kernels.h"
#include "learned.h"

namespace zamd {

struct DeviceBuffers {
	// Frame 0 of plane p at ptr[p]; row stride and per-frame stride in bytes.
	char *ptr[4] = {};
	long long stride[4] = {};
	long long fstride[4] = {};
};

// One launch configuration of the tiled resize kernel.
struct TiledConfig {
	int fwh = 0;
	unsigned tile_w = 0, tile_h = 0, in_rows = 0, in_cols = 0, mid_rows = 0, mid_cols = 0;
	double cost = 0;
};

// Launch shape of the integer matrix-core form of a WORD resize (ok ==
// false: the step is not eligible and runs on the tiled kernel).
struct ImmaPlan {
	bool ok = false;
	unsigned rb = 0, mb = 0, mid_pitch = 0, in_cols = 0, in_pitch = 0, lds = 0;
};

class DeviceProgram {
public:
	struct Buf {
		enum Kind : int { NONE, SRC, DST, TMP } kind = NONE;
		int idx = 0;
	};

	struct Step {
		enum Kind : int { RESIZE, POINT, CSP, ED, UNRESIZE } kind = POINT;
		Buf in[3], out[3];
		dev::InXform ix[3]{};
		dev::OutXform ox[3]{};
		unsigned width = 0, height = 0;    // output size (all kinds)
		unsigned in_w = 0, in_h = 0;       // input size (resize)
		// RESIZE
		int passes = 1;
		bool h_first = true;
		int tab[2] = { -1, -1 };           // node ids owning the pass tables
		bool word = false;
		int pixel_max = 65535;
		unsigned mid_w = 0, mid_h = 0;
		unsigned tile_w = 0, tile_h = 0, lds_in_rows = 0, lds_in_cols = 0, lds_mid_rows = 0, lds_mid_cols = 0;
		// tiled kernel configuration (fast.fwh == 0: not eligible) and the
		// best candidates by the cost model, cheapest first (for the autotuner)
		TiledConfig fast{};
		std::vector<TiledConfig> fast_options;
		ImmaPlan imma{};
		// POINT
		unsigned crop_left = 0, crop_top = 0;
		int fill = 0;
		uint32_t fill_bits = 0;
		int alpha_op = 0;                  // 1 premultiply, 2 unpremultiply (alpha plane in in[1])
		// CSP
		CspProgram prog;
		// ED (planes of equal size share one launch)
		int ed_planes = 1;
		float ed_scale[3] = { 1.0f, 1.0f, 1.0f }, ed_offset[3] = { 0.0f, 0.0f, 0.0f };
		bool dead = false;
		std::string label;
	};

	struct Slot {
		unsigned width, height, bpp;
		size_t pitch;
	};

	explicit DeviceProgram(const Plan &plan);
	~DeviceProgram();
	DeviceProgram(const DeviceProgram &) = delete;
	DeviceProgram &operator=(const DeviceProgram &) = delete;

	size_t tmp_size(unsigned frames) const;
	// FLOAT resize arithmetic: 0 the reference's FP32 order (default), 1 matrix
	// cores with FP16 hi+lo operands, 2 with BF16 hi+lo operands. Set before
	// sizing tmp and processing; not thread-safe against concurrent processing.
	void set_resize_precision(int mode);
	int resize_precision() const { return m_precision; }
	// Learned transfer surrogates (at most one per transfer function): a
	// colourspace step's CSP_GAMMA operations whose transfer has one, and the
	// HLG OETF of its CSP_B67 operations, run through the surrogate kernel
	// (exact fallback outside its validated bins) between colourspace
	// launches, in three FLOAT planes of tmp. Empty (the default): every
	// transfer exact. The surrogates must outlive the calls that use them and
	// live on the device that runs them; set before sizing tmp.
	void set_transfer_surrogates(const std::vector<const learned::Surrogate *> &s);
	bool has_transfer_surrogates() const { return !m_surrogates.empty(); }
	// Enqueue the program for `frames` frames. Throws Error on invalid buffers or launch failure.
	// events (optional, steps().size() + 1 hipEvent_t): recorded before each step and after the last.
	void run(const DeviceBuffers &src, const DeviceBuffers &dst, void *tmp, unsigned frames, void *stream,
	         void *const *events = nullptr) const;
	// Times the cost model's candidate tile shapes of every tiled resize
	// step on these buffers (the program runs whole, the step's own time is
	// taken from HIP events around it on `stream`, minimum of 5 runs) and
	// keeps the fastest for each step; a candidate replaces the current shape
	// only when at least 2 % faster.
	// Outputs do not change (every candidate computes the same values).
	// Returns how many steps changed shape. Not thread-safe against
	// concurrent processing of the program.
	unsigned autotune(const DeviceBuffers &src, const DeviceBuffers &dst, void *tmp, unsigned frames, void *stream);
	// Algorithmic bytes of one frame through each step: its input planes read
	// once and its output planes written once.
	std::vector<double> step_bytes() const;
	std::string describe() const;

	const Plan &plan() const { return m_plan; }

	struct Tables {
		void *mem = nullptr;
		std::map<int, dev::PassArg> pass; // by node id
		void *lut_mem = nullptr;
		std::map<int, const float *> lut; // gamma lookup tables, by step index * 8 + operation index
		void *unresize_mem = nullptr;
		std::map<int, dev::UnresizeArg> unresize; // table pointers of unresize passes, by node id
	};

	// Uploads every table the program's launches read on the current HIP
	// device (resize passes, integer and FP16/BF16 matrix-core blocks of the
	// current precision mode). Idempotent; run() does it on first use, so
	// callers that must not allocate later (graph capture, the allocation-free
	// entry point) call it up front.
	void prepare_current_device() const;
	// True when every table run() needs is already resident on the current device.
	bool resident_on_current_device() const;

	// Steps and per-device tables, for launches batched across programs.
	const std::vector<Step> &steps() const { return m_steps; }
	const Tables &tables_for_current_device() const;
	// Kernel arguments of a RESIZE step on given planes. cfg selects a tiled
	// configuration; nullptr takes the step's chosen one when the step and
	// the buffers allow it, else the generic kernel.
	dev::ResizeArg resize_arg(const Step &s, const dev::PlaneArg &src, const dev::PlaneArg &dst, unsigned frames, const Tables &tabs,
	                          const TiledConfig *cfg = nullptr) const;
	static bool aligned16(const dev::PlaneArg &p);

private:
	Plan m_plan;
	std::vector<Step> m_steps;
	std::vector<Slot> m_slots;

	struct MfmaTables {
		void *mem = nullptr;
		std::map<int, dev::MfmaTab> pass; // by node id
	};
	mutable std::mutex m_mutex;
	mutable std::map<int, std::shared_ptr<Tables>> m_tables; // by HIP device
	mutable std::map<int, std::shared_ptr<MfmaTables>> m_mfma_tables; // by HIP device * 4 + mode
	struct ImmaTables {
		void *mem = nullptr;
		std::map<int, const dev::ImmaBlock *> pass; // by 2 x node id + pass (0 horizontal, 1 vertical)
	};
	mutable std::map<int, std::shared_ptr<ImmaTables>> m_imma_tables; // by HIP device
	const ImmaTables &imma_tables_for_current_device() const;
	ImmaPlan imma_plan(const Step &s) const;
	int m_precision = 0;
	mutable std::set<int> m_prepared; // HIP device * 4 + precision mode

	std::vector<const learned::Surrogate *> m_surrogates;
	// One launch of a colourspace step split around its learned transfers.
	struct CspSegment {
		const learned::Surrogate *sur = nullptr; // null: a colourspace launch of prog
		CspProgram prog;
		float pre = 1.0f, post = 1.0f;
	};
	// The launches of a CSP step (luts: the step's tables patched in); empty
	// when no surrogate applies to it.
	std::vector<CspSegment> csp_segments(const Step &s, const CspProgram &prog) const;
	size_t surrogate_bytes(const Step &s, unsigned frames) const;
	void run_learned_csp(const Step &s, const std::vector<CspSegment> &segs, const dev::PlaneArg *src, const dev::PlaneArg *dst,
	                     char *mid, unsigned frames, void *stream) const;

	bool mfma_step(const Step &s) const { return m_precision && s.kind == Step::RESIZE && !s.word && s.ox[0].dither == 0; }
	// Launch shape of the one-launch two-pass matrix-core resize (ok == false:
	// the step runs as two launches through an HBM intermediate).
	struct FusedPlan {
		bool ok = false;
		unsigned tile_h = 0, slab_stride = 0, mid_stride = 0, coef_blocks = 0, lds = 0;
		unsigned lds_staged = 0; // LDS when the output tile is staged through LDS (0: not possible)
	};
	FusedPlan mfma_fused_plan(const Step &s) const;
	size_t mfma_mid_bytes(const Step &s, unsigned frames) const;
	const MfmaTables &mfma_tables_for_current_device() const;
	void run_mfma_resize(const Step &s, const dev::PlaneArg &src, const dev::PlaneArg &dst, char *mid, unsigned frames, void *stream) const;

	void lower();
	void fuse();
	void configure_resize(Step &s);
	void check_tiled(const Step &s, const TiledConfig &c, const FilterTable *th_tab, const FilterTable *tv_tab) const;
};

} // namespace zamd
