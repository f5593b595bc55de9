// Reduced operator, FP32 instantiation (the inner solve of the mixed form).
#include "fh_operator_impl.hpp"

namespace fh {

template <class T>
void spmv(const Absorber& a, const T* p, T* y, int r, Scratch<T>& s, const uint32_t* done, hipStream_t st,
          double* pq) {
    require(!a.slope, "slope absorbers solve in FP64", FH_INTERNAL);
    spmv_planned<T>(a, p, y, r, s, done, st, pq);
}
template void spmv<float>(const Absorber&, const float*, float*, int, Scratch<float>&, const uint32_t*, hipStream_t,
                          double*);
template SpmvNeed spmv_scratch_need<float>(const Absorber&, const SpmvPlan&, int);
template void spmv_fixed<float>(const Absorber&, const float*, float*, int, Scratch<float>&, hipStream_t);

}  // namespace fh
