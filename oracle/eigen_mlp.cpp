// eigen_mlp.cpp -- the CPU baseline of the NRC MLP on Eigen 3.4.0.
//
// TEST INFRASTRUCTURE ONLY: loaded by bench.py's cpu_baseline leg and by tests/,
// never by the product library.  It is built (oracle/Makefile, target ref) against
// the Eigen 3.4.0 headers the reference vendors (test/eigen), where they lie, into
// oracle/_ref/libeigen_mlp.so.
//
// The forward follows the reference's CPU test MLP, Evaluate (test/main.cpp:11-27):
// activations are column-major 64 x n matrices (one column per sample), each
// layer's weights a row-major [out][in] map of the weight buffer, five hidden
// layers Y = max(W X, 0), then the 3 x 64 output layer with no activation.  It
// runs in two scalar types: _Float16, as the reference's test does (every
// product and sum of Eigen's GEMM rounded to FP16), and float.
//
// The training step is the FP32 restatement of the gradient and Adam of the build
// (the reference's Train, test/main.cpp:29-74, is not a valid gradient: ReLU on
// the output, no factor 2, a step mask, a transposed target view), written with
// Eigen products: L2 loss 2 (o - t), deltas masked by 1[a > 0], dW_l the product
// of the deltas and the activations over the batch, then Adam with bias
// correction by running products (nrc_optimize.comp:36-45).
#include <Eigen/Dense>

#include <cstdint>
#include <cstring>
#include <vector>

#ifdef _OPENMP
#include <omp.h>
#endif

namespace {

template <typename T> using Act = Eigen::Matrix<T, Eigen::Dynamic, Eigen::Dynamic, Eigen::ColMajor>;
template <typename T> using WMat = Eigen::Matrix<T, Eigen::Dynamic, Eigen::Dynamic, Eigen::RowMajor>;

// Samples per block when the batch is split over threads (every sample is
// independent, so each thread runs Eigen on its own columns).
constexpr int64_t kBlock = 1024;
int g_threads = 1;

template <typename T> void forward(const T *w, int hidden, const T *x, int64_t n, Act<T> &out) {
	Act<T> a = Eigen::Map<const Act<T>>(x, 64, (Eigen::Index)n);
	for (int l = 0; l < hidden; ++l) {
		Eigen::Map<const WMat<T>> W(w + (size_t)l * 4096, 64, 64);
		Act<T> z = W * a;
		a = z.unaryExpr([](const T &v) -> T { return v > T(0) ? v : T(0); });
	}
	Eigen::Map<const WMat<T>> Wo(w + (size_t)hidden * 4096, 3, 64);
	out = Wo * a;
}

} // namespace

extern "C" {

// Threads of the batch split (Eigen itself runs single-threaded in each block).
int eigen_set_threads(int n) {
	g_threads = n < 1 ? 1 : n;
#ifdef _OPENMP
	omp_set_num_threads(g_threads);
#endif
	Eigen::setNbThreads(1);
	return g_threads;
}

// Raw outputs [n][3] (no clamp) of the _Float16 forward over fp16 inputs [n][64]
// and fp16 weights (IEEE binary16 bit patterns).
void eigen_forward_f16(const uint16_t *w, int hidden, const uint16_t *x, int64_t n, float *out) {
	static_assert(sizeof(_Float16) == sizeof(uint16_t), "binary16");
	const int64_t nb = (n + kBlock - 1) / kBlock;
#pragma omp parallel for schedule(static) num_threads(g_threads)
	for (int64_t b = 0; b < nb; ++b) {
		const int64_t i0 = b * kBlock, m = n - i0 < kBlock ? n - i0 : kBlock;
		Act<_Float16> o;
		forward(reinterpret_cast<const _Float16 *>(w), hidden, reinterpret_cast<const _Float16 *>(x) + i0 * 64, m, o);
		for (int64_t i = 0; i < m; ++i)
			for (int c = 0; c < 3; ++c) out[(i0 + i) * 3 + c] = (float)o(c, (Eigen::Index)i);
	}
}

// The same in float.
void eigen_forward_f32(const float *w, int hidden, const float *x, int64_t n, float *out) {
	const int64_t nb = (n + kBlock - 1) / kBlock;
#pragma omp parallel for schedule(static) num_threads(g_threads)
	for (int64_t b = 0; b < nb; ++b) {
		const int64_t i0 = b * kBlock, m = n - i0 < kBlock ? n - i0 : kBlock;
		Act<float> o;
		forward(w, hidden, x + i0 * 64, m, o);
		std::memcpy(out + i0 * 3, o.data(), sizeof(float) * 3 * (size_t)m);
	}
}

} // extern "C"

namespace {

double gradient_block(const float *w, int hidden, const float *x, const float *t, int64_t n, float *dw) {
	std::vector<Act<float>> a(hidden + 1);
	a[0] = Eigen::Map<const Act<float>>(x, 64, (Eigen::Index)n);
	for (int l = 0; l < hidden; ++l) {
		Eigen::Map<const WMat<float>> W(w + (size_t)l * 4096, 64, 64);
		a[l + 1] = (W * a[l]).cwiseMax(0.0f);
	}
	Eigen::Map<const WMat<float>> Wo(w + (size_t)hidden * 4096, 3, 64);
	const Act<float> o = Wo * a[hidden];
	const Act<float> diff = o - Eigen::Map<const Act<float>>(t, 3, (Eigen::Index)n);
	const Act<float> go = 2.0f * diff;
	Eigen::Map<WMat<float>>(dw + (size_t)hidden * 4096, 3, 64) = go * a[hidden].transpose();
	Act<float> d = (Wo.transpose() * go).cwiseProduct((a[hidden].array() > 0.0f).cast<float>().matrix());
	for (int l = hidden - 1; l >= 0; --l) {
		Eigen::Map<WMat<float>>(dw + (size_t)l * 4096, 64, 64) = d * a[l].transpose();
		if (l > 0) {
			Eigen::Map<const WMat<float>> W(w + (size_t)l * 4096, 64, 64);
			d = (W.transpose() * d).cwiseProduct((a[l].array() > 0.0f).cast<float>().matrix());
		}
	}
	return (double)diff.squaredNorm();
}

} // namespace

extern "C" {

// Summed L2-loss gradient dW [weight_count] over n samples (inputs [n][64],
// targets [n][3]); returns the summed loss sum_i |o_i - t_i|^2.  Blocks of
// samples run on the threads; their dW are summed in block order.
double eigen_gradient_l2_f32(const float *w, int hidden, const float *x, const float *t, int64_t n, float *dw) {
	const int wc = hidden * 4096 + 192;
	const int64_t nb = (n + kBlock - 1) / kBlock;
	std::vector<float> part((size_t)nb * wc);
	std::vector<double> loss((size_t)nb);
#pragma omp parallel for schedule(static) num_threads(g_threads)
	for (int64_t b = 0; b < nb; ++b) {
		const int64_t i0 = b * kBlock, m = n - i0 < kBlock ? n - i0 : kBlock;
		loss[b] = gradient_block(w, hidden, x + i0 * 64, t + i0 * 3, m, part.data() + b * wc);
	}
	double L = 0.0;
	std::memset(dw, 0, sizeof(float) * wc);
	for (int64_t b = 0; b < nb; ++b) {
		L += loss[b];
		for (int i = 0; i < wc; ++i) dw[i] += part[b * wc + i];
	}
	return L;
}

// One training step: the gradient above, then Adam on w with moments m, v, at
// step t (1-based; beta^t as the running products b1t, b2t the caller keeps).
// Returns the mean loss.
double eigen_train_step_l2_f32(float *w, float *m, float *v, int hidden, const float *x, const float *t, int64_t n,
                               float lr, float b1, float b2, float eps, float b1t, float b2t) {
	const int wc = hidden * 4096 + 192;
	std::vector<float> dw((size_t)wc);
	const double loss = eigen_gradient_l2_f32(w, hidden, x, t, n, dw.data());
	Eigen::Map<Eigen::ArrayXf> W(w, wc), M(m, wc), V(v, wc);
	const Eigen::ArrayXf g = Eigen::Map<Eigen::ArrayXf>(dw.data(), wc) / (float)n;
	M = b1 * M + (1.0f - b1) * g;
	V = b2 * V + (1.0f - b2) * g.square();
	W -= lr * (M / (1.0f - b1t)) / ((V / (1.0f - b2t)).sqrt() + eps);
	return loss / (double)n;
}

} // extern "C"
