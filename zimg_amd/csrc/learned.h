// Learned transfer-function surrogate (1-32-32-1 SiLU MLP on matrix cores)
// with per-bin validation and exact fallback. No counterpart in the
// reference, whose GammaOperationC (operation_impl.cpp:53-77) always
// evaluates the transfer function exactly.
#pragma once

#include <cstddef>
#include <cstdint>

namespace zamd {
namespace learned {

// Kernel-side view of a trained surrogate.
struct DeviceModel {
	const float *params;
	const uint8_t *valid;
	unsigned nbins;
	int in_log, out_log;
	float lo, inv_width;     // transformed-input domain [lo, lo + 1 / inv_width)
	float out_scale, out_offset;
	int func;                // exact transfer (fallback)
	float pre, post;         // y = post * f(x * pre)
};

struct Config {
	int transfer = 0;        // zimg_transfer_characteristics_e
	int to_linear = 1;       // 1: EOTF direction (code -> linear), 0: inverse
	double tolerance = 1e-4; // per-bin bound on |surrogate - f| (absolute or relative)
	int relative = 0;
	unsigned steps = 4000;
	unsigned samples = 1u << 16;
	unsigned bins = 1024;
	unsigned seed = 1;
	double lr = 3e-3, lr_min = 1e-5;
};

class Surrogate {
public:
	Surrogate(const Config &cfg, void *stream);
	~Surrogate();
	Surrogate(const Surrogate &) = delete;
	Surrogate &operator=(const Surrogate &) = delete;

	// x[k] <- post * f(x[k] * pre) for n device values.
	void apply(float *x, size_t n, float pre, float post, void *stream) const;

	int func() const { return m_func; }
	double coverage() const { return static_cast<double>(m_coverage) / m_cfg.bins; }
	double max_validated_error() const { return m_max_err; }
	double train_loss() const { return m_train_loss; }
	void get_params(float *out) const;
	DeviceModel device_model(float pre, float post) const;

private:
	Config m_cfg;
	int m_func = 0;
	int m_in_log = 0, m_out_log = 0;
	double m_lo = 0, m_hi = 1, m_out_scale = 1, m_out_offset = 0;
	float *m_params = nullptr;
	uint8_t *m_valid = nullptr;
	unsigned m_coverage = 0;
	double m_max_err = 0, m_train_loss = 0;

	void validate(void *stream);
};

// dst[p] = post * f(src[p] * pre) over whole planes (up to 3, FLOAT, frames
// of a batch frame_stride bytes apart): the surrogate where its bins are
// validated, the exact transfer elsewhere. src and dst may be the same.
struct PlaneView {
	const char *src;
	char *dst;
	long long src_stride, src_fstride, dst_stride, dst_fstride;
};
void apply_planes(const DeviceModel &m, const PlaneView *planes, unsigned nplanes, unsigned width, unsigned height, unsigned frames,
                  void *stream);

// Internal transfer id of a public transfer characteristic and direction.
int function_for(int transfer, int to_linear);
// FP64 value of an internal transfer function (training targets, judging).
double exact(int func, double x);
int param_count();
// The exact transfer over device values (the method without the surrogate).
void apply_exact(int func, float *x, size_t n, float pre, float post, void *stream);
// Gradient of mean((v - t)^2) over n samples (host arrays; n a multiple of
// 64) with the training kernels; grad has param_count() + 1 entries, the last
// is the mean loss.
void gradients(const float *params, const float *u, const float *t, unsigned n, float *grad, void *stream);
// Network outputs at inputs u (host arrays) with the inference layers.
void eval(const float *params, const float *u, float *v, unsigned n, void *stream);

} // namespace learned
} // namespace zamd
