// runtime_internal.h -- helpers shared by the libldpc host sources.
#pragma once
#include <stdint.h>

#include <string>

#include "../../include/ldpc_mi355x.h"
#include "kernels.h"

namespace ldpc {

// AR4JA ids: k_id 0..2 for K = 1024, 4096, 16384; rate_id 0..2 for 1/2, 2/3, 4/5.
bool ar4ja_ids(const ldpc_code_desc &d, int &k_id, int &rate_id);
const MultiVariant *multi_table(int k_id, int rate_id, int *n);
uint32_t variant_code(const MultiVariant &v);
const MultiVariant *find_variant(int k_id, int rate_id, uint32_t variant);
uint32_t default_variant(int k_id, int rate_id);
const MultiVariant *profile_d_kernel(int k_id, int rate_id);
void fill_code_info(int k_id, int rate_id, ldpc_code_info &info);
// BPSK LLR mean / deviation at ebn0_db for rate k_info / n_tx.
void channel_params(const ldpc_code_info &info, double ebn0_db, float *mu, float *sigma);
// Quasi-cyclic generator of an AR4JA code (encoder.cpp, cached): words =
// [12 parity blocks][nu = K / z info blocks][z / 32] coefficient bits of g_{r,i}.
int qc_generator(int k_id, int rate_id, const uint32_t **words, int *z, int *nu, std::string &why);
// Message that ldpc_last_error(NULL) returns, for calls that take no context.
void set_context_free_error(const std::string &msg);

}  // namespace ldpc
