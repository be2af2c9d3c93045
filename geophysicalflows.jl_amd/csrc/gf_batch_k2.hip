// Batched ensemble kernel class K2 (FP64 and FP32).
#define GF_BATCH_K 2
#include "gf_batch_impl.hpp"
