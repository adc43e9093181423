// Single HIP translation unit for all device code of libedx, so the
// constant-memory binomial table and the shared kernels need no relocatable
// device code.
#include <sys/stat.h>

#include <algorithm>
#include <cstdio>
#include <string>
#include <cmath>
#include <complex>
#include <cstring>
#include <map>
#include <mutex>
#include <tuple>
#include <utility>
#include <vector>

#include "device.h"

#include "basis.hip"
#include "csr_build.hip"
#include "spmv.hip"
#include "blas.hip"
#include "krylov.hip"
#include "mf.hip"
#include "batch.hip"
#include "factor.hip"
#include "autotune.hip"
#include "multishift.hip"
#include "group.hip"
#include "block.hip"
#include "observables.hip"
