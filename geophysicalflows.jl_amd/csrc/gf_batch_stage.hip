// Batched ensembles: the persistent stage kernel (all three kernel classes of a stage in one
// cooperative launch; FP64 and FP32).
#define GF_BATCH_K 0
#include "gf_batch_impl.hpp"
