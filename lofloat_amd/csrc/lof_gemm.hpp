// lof_gemm.hpp -- host-side interface of the GEMM launcher, shared with the grouped path.
#pragma once

#include <hip/hip_runtime.h>

#include "lof_gemm_kernels.hpp"
#include "lof_gemm_mfma.hpp"
#include "lofloat.h"

namespace lofamd {

lof_status prepare_gemm(const lof_gemm_desc* d, GemmArgs* g);
void set_fast_path(GemmArgs* g, bool on, int m);
lof_status launch_gemm(const GemmArgs& g, int config, hipStream_t st);
int resolve_gemm_config(const lof_gemm_desc& d, int config);
int gemm_num_configs();
int gemm_config_bm(int c);
int gemm_config_bn(int c);
// the chunk lengths whose sums can run on MFMA (2, 4, 8, 16, 32)
bool gemm_chunk_fits_mfma(int chunk);

}  // namespace lofamd
