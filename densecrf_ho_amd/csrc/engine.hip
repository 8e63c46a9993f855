// Single translation unit of libdensecrf_ho_amd.so (kernels live in
// anonymous namespaces of the included sources).
#include "lattice.hip"
#include "crf.hip"
#include "restricted.hip"
#include "ho.hip"
#include "lp.hip"
#include "bf.hip"
#include "dc.hip"
#include "kgrad.hip"
#include "adjoint.hip"
#include "batch.hip"
#include "noalloc.hip"
#include "capi.hip"
