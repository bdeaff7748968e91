// The cost model's MLP (xc_mlp.hip: GPU training kernels) and the method selector
// (xc_select.cpp: candidates, features, host inference, exhaustive autotuner).
#pragma once

#include <vector>

#include "xc_internal.hpp"

namespace xc {

void mlp_check_shape(const xc_mlp_shape* sh);
uint64_t mlp_num_params(const xc_mlp_shape& sh);
uint64_t mlp_work_size(const xc_mlp_shape& sh, uint32_t batch);
void mlp_init(const xc_mlp_shape& sh, uint64_t seed, float* params);
void mlp_forward(const xc_mlp_shape& sh, const float* P, const float* X, uint32_t batch, float* Y, float* work,
                 hipStream_t s);
void mlp_loss_grad(const xc_mlp_shape& sh, const float* P, const float* X, const float* T, const float* M,
                   uint32_t batch, float* loss, float* grad, float* work, hipStream_t s);
void mlp_adam(uint64_t count, float* p, const float* g, float* m, float* v, uint32_t step, float lr, float b1,
              float b2, float eps, float wd, hipStream_t s);
void mlp_train(const xc_mlp_shape& sh, const float* x_host, const float* t_host, const float* m_host,
               uint32_t batch, const xc_train_options& o, float* params_host, float* loss_host);

// Method selection.
uint32_t num_candidates();
xc_options candidate(uint32_t i);
uint32_t workload_features(const xc_workload& w, float* f, uint32_t cap);
void cost_model_load(const char* path);
void cost_model_fit(const char* db_path, const char* model_path, const xc_train_options& o, uint32_t holdout_mod);
void cost_model_predict(const xc_workload& w, float* ms);
int select_method(const xc_workload& w, xc_options& chosen);  // xc_select_source
uint32_t autotune_f32(const family_args& a, const float* L, const float* R, float* O, uint32_t reps, float* ms,
                      hipStream_t s);
float time_candidate_f32(const family_args& a, const float* L, const float* R, float* O, uint32_t index,
                         uint32_t reps, hipStream_t s);
void autotune_clear();
xc_options auto_options_f32(const family_args& a, const float* L, const float* R, float* O, hipStream_t s);
// XC_PREC_AUTO for double data: the FP64 direct kernel or the own FFT, timed once per workload.
xc_options auto_options_f64(const family_args& a, const double* L, const double* R, double* O, hipStream_t s);

}  // namespace xc
