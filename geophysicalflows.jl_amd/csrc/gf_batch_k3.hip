// Batched ensemble kernel class K3 (FP64 and FP32).
#define GF_BATCH_K 3
#include "gf_batch_impl.hpp"
