// Learned warm start: MLP training and inference driver (see warm_start.hpp).
#include "warm_start.hpp"

#include <algorithm>
#include <cmath>
#include <cstring>
#include <numeric>
#include <random>
#include <stdexcept>
#include <string>

#include "mlp_kernels.hpp"

namespace cupf {

namespace {
constexpr float kStdFloor = 1e-6f;
constexpr double kPi = 3.14159265358979323846;
}

WarmStartMLP::WarmStartMLP(int n_bus, const int32_t* pv, int32_t n_pv, const int32_t* pq,
                           int32_t n_pq, int hidden, int depth, uint64_t seed, int max_batch,
                           bool skip)
    : n_(n_bus), skip_(skip) {
    if (n_bus <= 0 || hidden <= 0 || depth < 1 || max_batch <= 0)
        throw std::invalid_argument("WarmStartMLP: sizes must be positive (depth >= 1)");
    if (n_bus < 8 || hidden % 8 || max_batch % 8)
        throw std::invalid_argument("WarmStartMLP: n_bus >= 8; hidden and max_batch multiples of 8");
    if (n_pv < 0 || n_pq < 0 || (n_pv > 0 && !pv) || (n_pq > 0 && !pq))
        throw std::invalid_argument("WarmStartMLP: bad PV/PQ lists");
    std::vector<unsigned char> kind(n_bus, 0);
    for (int i = 0; i < n_pv; ++i) {
        if (pv[i] < 0 || pv[i] >= n_bus) throw std::invalid_argument("WarmStartMLP: PV index out of range");
        kind[pv[i]] = 1;
    }
    for (int i = 0; i < n_pq; ++i) {
        if (pq[i] < 0 || pq[i] >= n_bus) throw std::invalid_argument("WarmStartMLP: PQ index out of range");
        kind[pq[i]] = 2;
    }
    const int width = (2 * n_bus + 7) / 8 * 8;   // [P | Q] / [Va | Vm] rows padded to 8
    dims_.push_back(width);
    for (int l = 0; l < depth; ++l) dims_.push_back(hidden);
    dims_.push_back(width);
    const int L = int(dims_.size()) - 1;
    for (int l = 1; l <= L; ++l) {
        w_off_.push_back(n_params_);
        n_params_ += int64_t(dims_[l]) * dims_[l - 1];
        b_off_.push_back(n_params_);
        n_params_ += dims_[l];
        wbf_off_.push_back(n_wbf_);
        n_wbf_ += int64_t(dims_[l]) * dims_[l - 1];
    }
    if (skip_) {
        skip_off_ = n_params_;
        skip_wbf_off_ = n_wbf_;
        n_params_ += int64_t(dims_[L]) * dims_[0];
        n_wbf_ += int64_t(dims_[L]) * dims_[0];
    }
    // Glorot-uniform weights (output layer scaled by 0.1), zero biases
    std::vector<float> w(size_t(n_params_), 0.f);
    std::vector<long long> map(size_t(n_params_), -1);
    std::mt19937_64 rng(seed);
    for (int l = 1; l <= L; ++l) {
        const double a = std::sqrt(6.0 / double(dims_[l] + dims_[l - 1])) * (l == L ? 0.1 : 1.0);
        std::uniform_real_distribution<double> U(-a, a);
        const int64_t nw = int64_t(dims_[l]) * dims_[l - 1];
        for (int64_t k = 0; k < nw; ++k) {
            w[size_t(w_off_[l - 1] + k)] = float(U(rng));
            map[size_t(w_off_[l - 1] + k)] = wbf_off_[l - 1] + k;
        }
    }
    if (skip_) {                             // Ws starts at zero
        const int64_t nw = int64_t(dims_[L]) * dims_[0];
        for (int64_t k = 0; k < nw; ++k) map[size_t(skip_off_ + k)] = skip_wbf_off_ + k;
    }
    hipStream_t s = nullptr;
    w_ = upload_vector(w, s);
    bf_map_ = upload_vector(map, s);
    kind_ = upload_vector(kind, s);
    g_.allocate(size_t(n_params_) * sizeof(float));
    m_.allocate(size_t(n_params_) * sizeof(float));
    v_.allocate(size_t(n_params_) * sizeof(float));
    CUPF_HIP(hipMemsetAsync(m_.get(), 0, m_.bytes(), s));
    CUPF_HIP(hipMemsetAsync(v_.get(), 0, v_.bytes(), s));
    wbf_.allocate(size_t(n_wbf_) * sizeof(uint16_t));
    mlp::to_bf16(n_params_, w_.as<float>(), bf_map_.as<long long>(), wbf_.as<uint16_t>(), s);
    std::vector<float> zero(size_t(width), 0.f), one(size_t(width), 1.f);
    xm_ = upload_vector(zero, s);
    xs_ = upload_vector(one, s);
    ym_ = upload_vector(zero, s);
    ys_ = upload_vector(one, s);
    CUPF_HIP(hipStreamSynchronize(s));
    ensure_batch(max_batch);
}

WarmStartMLP::~WarmStartMLP() = default;

void WarmStartMLP::ensure_batch(int batch) {
    if (batch <= cap_batch_) return;
    const int L = int(dims_.size()) - 1;
    A_.clear();
    H_.clear();
    dH_.clear();
    A_.resize(L + 1);
    H_.resize(L + 1);
    dH_.resize(L + 1);
    for (int l = 0; l <= L; ++l) {
        A_[l].allocate(size_t(batch) * dims_[l] * sizeof(uint16_t));
        if (l > 0) {
            H_[l].allocate(size_t(batch) * dims_[l] * sizeof(float));
            dH_[l].allocate(size_t(batch) * dims_[l] * sizeof(uint16_t));
        }
    }
    T_batch_.allocate(size_t(batch) * dims_[L] * sizeof(float));
    partial_.allocate(size_t(mlp::mse_blocks((long long)batch * dims_[L])) * sizeof(float));
    loss_acc_.allocate(sizeof(float));
    colsum_.allocate(mlp::colsum_part_elems(*std::max_element(dims_.begin(), dims_.end())) *
                     sizeof(float));
    cap_batch_ = batch;
    ++buffers_gen_;
}

void WarmStartMLP::reserve_samples(int64_t capacity) {
    if (capacity <= cap_samples_) return;
    const size_t row_in = size_t(dims_.front()), row_out = size_t(dims_.back());
    DeviceBuffer X(size_t(capacity) * row_in * sizeof(float));
    DeviceBuffer Y(size_t(capacity) * row_out * sizeof(float));
    if (n_samples_ > 0) {
        CUPF_HIP(hipMemcpy(X.get(), X_.get(), size_t(n_samples_) * row_in * sizeof(float),
                           hipMemcpyDeviceToDevice));
        CUPF_HIP(hipMemcpy(Y.get(), Y_.get(), size_t(n_samples_) * row_out * sizeof(float),
                           hipMemcpyDeviceToDevice));
    }
    X_ = std::move(X);
    Y_ = std::move(Y);
    cap_samples_ = capacity;
}

void WarmStartMLP::add_samples(const double2* S, int64_t s_stride, const double2* V,
                               int64_t v_stride, int count, hipStream_t s) {
    if (normalized_) throw std::runtime_error("WarmStartMLP::add_samples(): data set already finalized");
    if (count <= 0) return;
    if (!S || !V || s_stride < n_ || v_stride < n_)
        throw std::invalid_argument("WarmStartMLP::add_samples(): bad pointers or strides");
    if (n_samples_ + count > cap_samples_) reserve_samples(std::max<int64_t>(2 * cap_samples_, n_samples_ + count));
    const size_t w = size_t(dims_.front());
    mlp::raw_samples(n_, int(w), count, S, s_stride, V, v_stride,
                     X_.as<float>() + size_t(n_samples_) * w, Y_.as<float>() + size_t(n_samples_) * w, s);
    n_samples_ += count;
}

void WarmStartMLP::finalize_dataset(hipStream_t s) {
    if (normalized_) return;
    if (n_samples_ < 2) throw std::runtime_error("WarmStartMLP::finalize_dataset(): need at least 2 samples");
    const int wi = dims_.front(), wo = dims_.back();
    mlp::col_stats(n_samples_, wi, X_.as<float>(), xm_.as<float>(), xs_.as<float>(), kStdFloor, s);
    mlp::col_stats(n_samples_, wo, Y_.as<float>(), ym_.as<float>(), ys_.as<float>(), kStdFloor, s);
    Xb_.allocate(size_t(n_samples_) * wi * sizeof(uint16_t));
    mlp::normalize(n_samples_, wi, X_.as<float>(), xm_.as<float>(), xs_.as<float>(), Xb_.as<uint16_t>(), s);
    mlp::normalize(n_samples_, wo, Y_.as<float>(), ym_.as<float>(), ys_.as<float>(), nullptr, s);
    CUPF_HIP(hipStreamSynchronize(s));
    X_.release();   // the BF16 copy is what the GEMMs read
    normalized_ = true;
}

void WarmStartMLP::forward(int count, hipStream_t s) {
    const int L = int(dims_.size()) - 1;
    for (int l = 1; l <= L; ++l) {
        mlp::Epilogue ep;
        ep.bias = w_.as<float>() + b_off_[l - 1];
        ep.ldc = dims_[l];
        ep.out_f32 = H_[l].as<float>();
        if (l < L) {
            ep.mode = mlp::kEpGelu;
            ep.out_bf16 = A_[l].as<uint16_t>();
        } else if (skip_) {                  // H[L] = A[0] Ws^T, then += last layer
            mlp::Epilogue es;
            es.ldc = dims_[L];
            es.out_f32 = H_[L].as<float>();
            mlp::gemm(false, true, count, dims_[L], dims_[0], A_[0].as<uint16_t>(), dims_[0],
                      wbf_.as<uint16_t>() + skip_wbf_off_, dims_[0], es, s);
            ep.mode = mlp::kEpAccum;
        }
        mlp::gemm(false, true, count, dims_[l], dims_[l - 1], A_[l - 1].as<uint16_t>(), dims_[l - 1],
                  wbf_.as<uint16_t>() + wbf_off_[l - 1], dims_[l - 1], ep, s);
    }
}

float WarmStartMLP::loss_grad(int count, const float* T, hipStream_t s, bool read) {
    const int L = int(dims_.size()) - 1;
    const long long ne = (long long)count * dims_[L];
    mlp::mse(ne, H_[L].as<float>(), T, dH_[L].as<uint16_t>(), partial_.as<float>(), s);
    if (!read) {
        mlp::sum_into(mlp::mse_blocks(ne), partial_.as<float>(), loss_acc_.as<float>(), s);
        return 0.f;
    }
    std::vector<float> p(size_t(mlp::mse_blocks(ne)));
    CUPF_HIP(hipMemcpyAsync(p.data(), partial_.get(), p.size() * sizeof(float), hipMemcpyDeviceToHost, s));
    CUPF_HIP(hipStreamSynchronize(s));
    double acc = 0.0;
    for (float x : p) acc += x;
    return float(acc / double(ne));
}

void WarmStartMLP::backward(int count, hipStream_t s) {
    const int L = int(dims_.size()) - 1;
    if (skip_) {                             // dWs = dH[L]^T A[0]
        mlp::Epilogue es;
        es.out_f32 = g_.as<float>() + skip_off_;
        es.ldc = dims_[0];
        mlp::gemm(true, false, dims_[L], dims_[0], count, dH_[L].as<uint16_t>(), dims_[L],
                  A_[0].as<uint16_t>(), dims_[0], es, s);
    }
    for (int l = L; l >= 1; --l) {
        const int in = dims_[l - 1], out = dims_[l];
        mlp::Epilogue ew;                    // dW^T = dH^T A_{l-1}
        ew.out_f32 = g_.as<float>() + w_off_[l - 1];
        ew.ldc = in;
        mlp::gemm(true, false, out, in, count, dH_[l].as<uint16_t>(), out, A_[l - 1].as<uint16_t>(), in,
                  ew, s);
        mlp::colsum_bf16(count, out, dH_[l].as<uint16_t>(), out, g_.as<float>() + b_off_[l - 1],
                         colsum_.as<float>(), s);
        if (l > 1) {                         // dH_{l-1} = (dH_l W) * gelu'(H_{l-1})
            mlp::Epilogue ex;
            ex.mode = mlp::kEpGeluBwd;
            ex.aux = H_[l - 1].as<float>();
            ex.out_bf16 = dH_[l - 1].as<uint16_t>();
            ex.ldc = in;
            mlp::gemm(false, false, count, in, out, dH_[l].as<uint16_t>(), out,
                      wbf_.as<uint16_t>() + wbf_off_[l - 1], in, ex, s);
        }
    }
}

void WarmStartMLP::train(int epochs, int batch, float lr, float weight_decay, uint64_t seed,
                         float* epoch_loss, hipStream_t s) {
    if (!normalized_) throw std::runtime_error("WarmStartMLP::train(): finalize_dataset() first");
    if (batch <= 0 || batch % 8) throw std::invalid_argument("WarmStartMLP::train(): batch must be a positive multiple of 8");
    if (epochs < 0 || !(lr > 0.f)) throw std::invalid_argument("WarmStartMLP::train(): bad epochs or lr");
    const int64_t nb = n_samples_ / batch;
    if (nb < 1) throw std::invalid_argument("WarmStartMLP::train(): fewer samples than one batch");
    ensure_batch(batch);
    const int wi = dims_.front(), wo = dims_.back(), L = int(dims_.size()) - 1;
    std::mt19937_64 rng(seed);
    std::vector<int> perm(static_cast<size_t>(n_samples_));
    idx_.reserve(perm.size() * sizeof(int));
    const int64_t total = int64_t(epochs) * nb;
    int64_t done = 0;
    for (int e = 0; e < epochs; ++e) {
        std::iota(perm.begin(), perm.end(), 0);
        std::shuffle(perm.begin(), perm.end(), rng);
        CUPF_HIP(hipMemcpyAsync(idx_.get(), perm.data(), perm.size() * sizeof(int), hipMemcpyHostToDevice, s));
        CUPF_HIP(hipMemsetAsync(loss_acc_.get(), 0, sizeof(float), s));
        for (int64_t k = 0; k < nb; ++k, ++done) {
            const int* ix = idx_.as<int>() + k * batch;
            mlp::gather_rows_bf16(batch, wi, ix, Xb_.as<uint16_t>(), A_[0].as<uint16_t>(), s);
            mlp::gather_rows_f32(batch, wo, ix, Y_.as<float>(), T_batch_.as<float>(), s);
            forward(batch, s);
            loss_grad(batch, T_batch_.as<float>(), s, false);
            backward(batch, s);
            // cosine learning-rate decay to 5% of lr
            const double ph = double(done) / double(std::max<int64_t>(1, total - 1));
            const float lr_k = float(lr * (0.05 + 0.95 * 0.5 * (1.0 + std::cos(kPi * ph))));
            ++step_;
            mlp::adamw(n_params_, w_.as<float>(), g_.as<float>(), m_.as<float>(), v_.as<float>(), lr_k,
                       0.9f, 0.999f, 1e-8f, weight_decay, step_, bf_map_.as<long long>(),
                       wbf_.as<uint16_t>(), s);
        }
        float acc = 0.f;
        CUPF_HIP(hipMemcpyAsync(&acc, loss_acc_.get(), sizeof(float), hipMemcpyDeviceToHost, s));
        CUPF_HIP(hipStreamSynchronize(s));
        if (epoch_loss) epoch_loss[e] = acc / float(double(nb) * double(batch) * double(dims_[L]));
    }
}

void WarmStartMLP::predict_v0(const double2* S, int64_t s_stride, const double2* Vref,
                              int64_t r_stride, int count, double2* V0, int64_t o_stride,
                              hipStream_t s) {
    if (count <= 0) return;
    if (!S || !Vref || !V0 || s_stride < n_ || r_stride < n_ || o_stride < n_)
        throw std::invalid_argument("WarmStartMLP::predict_v0(): bad pointers or strides");
    const int L = int(dims_.size()) - 1;
    for (int c0 = 0; c0 < count; c0 += cap_batch_) {
        const int c = std::min(cap_batch_, count - c0);
        const int cp = (c + 7) / 8 * 8;      // GEMM rows: multiple of 8 (pad rows are zero inputs)
        if (cp != c)
            mlp::zero_u16((long long)(cp - c) * dims_[0], A_[0].as<uint16_t>() + size_t(c) * dims_[0], s);
        mlp::features(n_, dims_[0], c, S + c0 * s_stride, s_stride, xm_.as<float>(), xs_.as<float>(),
                      A_[0].as<uint16_t>(), s);
        forward(cp, s);
        mlp::warm_v0(n_, dims_[L], c, H_[L].as<float>(), ym_.as<float>(), ys_.as<float>(),
                     kind_.as<unsigned char>(), Vref + c0 * r_stride, r_stride, V0 + c0 * o_stride,
                     o_stride, s);
    }
}

void WarmStartMLP::forward_host(const float* xn, int count, float* yn) {
    if (count <= 0 || count % 8) throw std::invalid_argument("WarmStartMLP::forward_host(): count must be a positive multiple of 8");
    ensure_batch(count);
    hipStream_t s = nullptr;
    const int L = int(dims_.size()) - 1;
    std::vector<uint16_t> xb(size_t(count) * dims_[0]);
    for (size_t i = 0; i < xb.size(); ++i) {
        uint32_t u;
        std::memcpy(&u, &xn[i], 4);
        u += 0x7fffu + ((u >> 16) & 1u);
        xb[i] = uint16_t(u >> 16);
    }
    CUPF_HIP(hipMemcpy(A_[0].get(), xb.data(), xb.size() * 2, hipMemcpyHostToDevice));
    forward(count, s);
    CUPF_HIP(hipMemcpy(yn, H_[L].get(), size_t(count) * dims_[L] * sizeof(float), hipMemcpyDeviceToHost));
}

float WarmStartMLP::loss_and_grads_host(const float* xn, const float* tn, int count, float* grads) {
    if (count <= 0 || count % 8) throw std::invalid_argument("WarmStartMLP::loss_and_grads_host(): count must be a positive multiple of 8");
    ensure_batch(count);
    hipStream_t s = nullptr;
    const int L = int(dims_.size()) - 1;
    std::vector<uint16_t> xb(size_t(count) * dims_[0]);
    for (size_t i = 0; i < xb.size(); ++i) {
        uint32_t u;
        std::memcpy(&u, &xn[i], 4);
        u += 0x7fffu + ((u >> 16) & 1u);
        xb[i] = uint16_t(u >> 16);
    }
    CUPF_HIP(hipMemcpy(A_[0].get(), xb.data(), xb.size() * 2, hipMemcpyHostToDevice));
    CUPF_HIP(hipMemcpy(T_batch_.get(), tn, size_t(count) * dims_[L] * sizeof(float), hipMemcpyHostToDevice));
    forward(count, s);
    const float loss = loss_grad(count, T_batch_.as<float>(), s, true);
    backward(count, s);
    CUPF_HIP(hipMemcpy(grads, g_.get(), size_t(n_params_) * sizeof(float), hipMemcpyDeviceToHost));
    return loss;
}

void WarmStartMLP::get_params(float* out) const {
    CUPF_HIP(hipMemcpy(out, w_.get(), size_t(n_params_) * sizeof(float), hipMemcpyDeviceToHost));
}

void WarmStartMLP::set_params(const float* in) {
    CUPF_HIP(hipMemcpy(w_.get(), in, size_t(n_params_) * sizeof(float), hipMemcpyHostToDevice));
    mlp::to_bf16(n_params_, w_.as<float>(), bf_map_.as<long long>(), wbf_.as<uint16_t>(), nullptr);
    CUPF_HIP(hipMemsetAsync(m_.get(), 0, m_.bytes(), nullptr));
    CUPF_HIP(hipMemsetAsync(v_.get(), 0, v_.bytes(), nullptr));
    CUPF_HIP(hipStreamSynchronize(nullptr));
    step_ = 0;
}

void WarmStartMLP::get_normalization(float* xm, float* xs, float* ym, float* ys) const {
    const size_t wi = size_t(dims_.front()) * sizeof(float), wo = size_t(dims_.back()) * sizeof(float);
    if (xm) CUPF_HIP(hipMemcpy(xm, xm_.get(), wi, hipMemcpyDeviceToHost));
    if (xs) CUPF_HIP(hipMemcpy(xs, xs_.get(), wi, hipMemcpyDeviceToHost));
    if (ym) CUPF_HIP(hipMemcpy(ym, ym_.get(), wo, hipMemcpyDeviceToHost));
    if (ys) CUPF_HIP(hipMemcpy(ys, ys_.get(), wo, hipMemcpyDeviceToHost));
}

void WarmStartMLP::set_normalization(const float* xm, const float* xs, const float* ym, const float* ys) {
    const size_t wi = size_t(dims_.front()), wo = size_t(dims_.back());
    for (size_t i = 0; i < wi; ++i)
        if (!(xs[i] > 0.f)) throw std::invalid_argument("WarmStartMLP::set_normalization(): std must be positive");
    for (size_t i = 0; i < wo; ++i)
        if (!(ys[i] > 0.f)) throw std::invalid_argument("WarmStartMLP::set_normalization(): std must be positive");
    CUPF_HIP(hipMemcpy(xm_.get(), xm, wi * 4, hipMemcpyHostToDevice));
    CUPF_HIP(hipMemcpy(xs_.get(), xs, wi * 4, hipMemcpyHostToDevice));
    CUPF_HIP(hipMemcpy(ym_.get(), ym, wo * 4, hipMemcpyHostToDevice));
    CUPF_HIP(hipMemcpy(ys_.get(), ys, wo * 4, hipMemcpyHostToDevice));
    normalized_ = true;
}

}  // namespace cupf
