// kernels.hpp -- host-side entry points of the HIP kernels in kernels.hip.
#pragma once

#include <hip/hip_runtime.h>

#include <cstddef>

#include "ops.hpp"

namespace cntc {

// LDS one work-group may allocate on gfx950 (160 KB per CU, all of it to one work-group)
constexpr size_t MAX_LDS = 160 * 1024;

// number of reduction values an item of `kind` produces (0: none)
int op_nred(int kind, const void* item);
// combine mode of the reduction: 0 sum, 1 max, 2 min
int op_combine_mode(int kind, const void* item);
// work-groups an item needs (conv_stage 0/1/2 selects the stage of an FFT product)
int op_groups(int kind, const void* item, int conv_stage);
// algorithmic HBM bytes of one item in one stage (for the bandwidth report)
double op_bytes(int kind, const void* item, int stage);
// dynamic LDS bytes of one work-group for this item and stage
size_t op_lds(int kind, const void* item, int stage);
// launch one batched kernel over `nitems` items whose work-groups are given by d_gstart; a
// kernel with reductions writes its group partials to ro.partials and, when ro.count is set,
// each item's results to ro.results (ro.count: one zeroed counter per item)
void launch_op(int kind, const void* d_items, const int* d_gstart, int nitems, int ngroups,
               RedOut ro, int conv_stage, size_t lds_bytes, hipStream_t s);

// the default tile shapes of a product item (rows per work-group of the row stages, columns
// per work-group of the column stage)
void conv_default_shape(const ConvItem* c, int* rpg, int* cpg);

void launch_rollcoef(int n, const double* cs, double* cv, hipStream_t s);
// p[0..n) = v, by a kernel (usable where a caller may capture the stream into a graph)
void launch_fill_u32(unsigned* p, size_t n, unsigned v, hipStream_t s);
void launch_to_float(size_t n, const double* in, float* out, hipStream_t s);

}  // namespace cntc
