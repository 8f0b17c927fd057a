// Learned warm start for the batched Newton-Raphson solve.
//
// A dense MLP maps the per-bus injections of one scenario, x = [P | Q] of
// Sbus (2n values, zero-padded to a multiple of 8), to its solved state
// y = [Va | Vm] (same width): depth
// hidden layers of `hidden` units with GELU, BF16 weights and activations,
// FP32 accumulation on MFMA.  Inputs and outputs are standardised per feature
// with statistics of the training set.  With `skip`, a linear path from the
// standardised input straight to the output layer (y = Ws x + MLP(x), Ws
// [width_out][width_in], no bias of its own) runs beside the hidden layers:
// the load -> voltage response of a grid is close to linear but of full rank
// (about n), which a `hidden`-wide bottleneck cannot carry; the skip holds
// that linear response and the hidden layers learn the nonlinear remainder.
// Training is AdamW on the mean squared
// error of the standardised outputs, over NR solutions that the solver itself
// produced.  There is no counterpart in the reference (its closest analogue is
// the PINN StatePredictor, benchmark/pinn/model.py:13-37, which differs in
// architecture, data and loss).
#pragma once

#include <hip/hip_runtime.h>

#include <cstdint>
#include <vector>

#include "hip_common.hpp"

namespace cupf {

class WarmStartMLP {
public:
    // bus kinds from the PV / PQ lists (the rest: slack / fixed)
    WarmStartMLP(int n_bus, const int32_t* pv, int32_t n_pv, const int32_t* pq, int32_t n_pq,
                 int hidden, int depth, uint64_t seed, int max_batch = 1024, bool skip = false);
    ~WarmStartMLP();
    WarmStartMLP(const WarmStartMLP&) = delete;
    WarmStartMLP& operator=(const WarmStartMLP&) = delete;

    int n_bus() const { return n_; }
    int width_in() const { return dims_.front(); }
    int width_out() const { return dims_.back(); }
    int64_t num_params() const { return n_params_; }
    int64_t num_samples() const { return n_samples_; }
    bool normalized() const { return normalized_; }
    bool has_skip() const { return skip_; }
    // Bumped whenever a device buffer predict_v0 uses moves (a captured
    // inference that baked the old addresses is stale).
    uint64_t buffers_generation() const { return buffers_gen_; }

    // ---- data set (device pointers; complex128 rows with the given strides)
    void reserve_samples(int64_t capacity);
    void add_samples(const double2* S, int64_t s_stride, const double2* V, int64_t v_stride,
                     int count, hipStream_t s);
    // column statistics and standardisation of everything added so far
    void finalize_dataset(hipStream_t s);

    // ---- training: `epochs` passes in shuffled mini-batches of `batch`
    // (multiple of 8); epoch_loss[e] = mean loss of epoch e (host, may be null)
    void train(int epochs, int batch, float lr, float weight_decay, uint64_t seed, float* epoch_loss,
               hipStream_t s);

    // ---- inference: V0 rows from Sbus rows; Vref supplies the slack voltage
    // and PV magnitudes (and every bus whose prediction is not finite)
    void predict_v0(const double2* S, int64_t s_stride, const double2* Vref, int64_t r_stride,
                    int count, double2* V0, int64_t o_stride, hipStream_t s);
    // standardised forward pass on host data [count][width_in] -> [count][width_out]
    void forward_host(const float* xn, int count, float* yn);
    // loss and flat gradient on host data (count multiple of 8)
    float loss_and_grads_host(const float* xn, const float* tn, int count, float* grads);

    // ---- parameters: flat FP32, per layer W^T [out][in] then b [out]; with
    // the skip path, its Ws [width_out][width_in] follows the last layer
    void get_params(float* out) const;
    void set_params(const float* in);
    // statistics: x mean/std [width_in], y mean/std [width_out]
    void get_normalization(float* xm, float* xs, float* ym, float* ys) const;
    void set_normalization(const float* xm, const float* xs, const float* ym, const float* ys);

private:
    void ensure_batch(int batch);
    void forward(int count, hipStream_t s);           // A[0] -> H[L]
    void backward(int count, hipStream_t s);          // dH[L] -> grads
    float loss_grad(int count, const float* T, hipStream_t s, bool read);

    int n_ = 0, cap_batch_ = 0;
    std::vector<int> dims_;                            // d[0] .. d[L]
    std::vector<int64_t> w_off_, b_off_, wbf_off_;
    int64_t n_params_ = 0, n_wbf_ = 0;
    bool skip_ = false;
    int64_t skip_off_ = 0, skip_wbf_off_ = 0;   // Ws in the FP32 / BF16 parameter vectors
    DeviceBuffer w_, g_, m_, v_, wbf_, bf_map_;
    DeviceBuffer kind_;
    // activations for cap_batch_ rows
    std::vector<DeviceBuffer> A_, H_, dH_;
    DeviceBuffer T_batch_, partial_, loss_acc_, idx_, colsum_;
    // data set
    int64_t cap_samples_ = 0, n_samples_ = 0;
    DeviceBuffer X_, Xb_, Y_;
    DeviceBuffer xm_, xs_, ym_, ys_;
    bool normalized_ = false;
    int step_ = 0;
    uint64_t buffers_gen_ = 0;
};

}  // namespace cupf
