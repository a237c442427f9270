// mlp_host.h -- host entry points of the MLP surrogate (training, evaluation, parameters).
#pragma once
#include "handle.h"

namespace nqcd {

int32_t mlp_ensure(nqcd_handle_t* h);
int32_t mlp_train(nqcd_handle_t* h, const double* X, const double* Y, int64_t n, const nqcd_train_desc* d,
                  double* loss_history);
int32_t mlp_gradient(nqcd_handle_t* h, int32_t member, const double* X, const double* Y, int64_t n, const double w[3],
                     double* grad, double* loss);
int32_t mlp_eval(nqcd_handle_t* h, int32_t member, const double* X, int64_t n, double* H, double* dH);
int32_t mlp_set_params(nqcd_handle_t* h, int32_t member, const float* P, const float* norm);
int32_t mlp_get_params(nqcd_handle_t* h, int32_t member, float* P, float* norm);

}  // namespace nqcd
