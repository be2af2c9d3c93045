// Adjoint kernels of the TwoDNavierStokes ETDRK4 step (FP64).
#include "gf_adjoint_impl.hpp"
