// Batched ensemble kernel class K1 (FP64 and FP32).
#define GF_BATCH_K 1
#include "gf_batch_impl.hpp"
