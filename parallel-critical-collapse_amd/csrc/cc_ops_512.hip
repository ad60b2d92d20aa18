// Kernels of transform length 512 (see cc_ops.hpp).
#include "cc_ops_impl.hpp"

template struct ccd::Ops<512>;
