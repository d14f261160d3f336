// hjm_select_kernels.hip -- the PFE order-statistic launches (FP64 and FP32 rows), kept in
// one translation unit so each selection kernel is instantiated once.
#include "hjm_kernels.h"
#include "hjm_select.h"

namespace hjm {

hipError_t launch_radix_select(const double* x, int rows, int64_t n, int64_t k, void* ws,
                               double* out, hipStream_t st)
{
    return launch_select_rows<double>(x, rows, n, k, ws, out, st);
}

hipError_t launch_radix_select_f32(const float* x, int rows, int64_t n, int64_t k, void* ws,
                                   double* out, hipStream_t st)
{
    return launch_select_rows<float>(x, rows, n, k, ws, out, st);
}

hipError_t launch_bracket_sample(const double* x, int rows, int64_t n, int64_t k, void* ws,
                                 hipStream_t st)
{
    return launch_bracket_sample_rows<double>(x, rows, n, k, ws, st);
}

hipError_t launch_bracket_sample_f32(const float* x, int rows, int64_t n, int64_t k, void* ws,
                                     hipStream_t st)
{
    return launch_bracket_sample_rows<float>(x, rows, n, k, ws, st);
}

hipError_t launch_bracket_finish(const double* x, int rows, int64_t n, int64_t k, void* ws,
                                 double* out, hipStream_t st)
{
    return launch_bracket_finish_rows<double>(x, rows, n, k, ws, out, st);
}

hipError_t launch_bracket_finish_f32(const float* x, int rows, int64_t n, int64_t k, void* ws,
                                     double* out, hipStream_t st)
{
    return launch_bracket_finish_rows<float>(x, rows, n, k, ws, out, st);
}

}  // namespace hjm
