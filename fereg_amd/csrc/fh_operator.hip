// Reduced operator, FP64 instantiation, rhs pass and back-substitution (see fh_operator_impl.hpp).
#include "fh_operator_impl.hpp"

namespace fh {

template <class T>
OpArgs<T> op_args(const Absorber& a) {
    return op_args_impl<T>(a);
}
template OpArgs<double> op_args<double>(const Absorber&);
template OpArgs<float> op_args<float>(const Absorber&);

template <class T>
void spmv(const Absorber& a, const T* p, T* y, int r, Scratch<T>& s, const uint32_t* done, hipStream_t st) {
    spmv_impl<T>(a, p, y, r, s, done, st);
}
template void spmv<double>(const Absorber&, const double*, double*, int, Scratch<double>&, const uint32_t*, hipStream_t);

void rhs_pass(Absorber& a, const double* rows, int r, double* rhs, double* b1, double* wsq_wsum) {
    require(r >= 1 && r <= 16, "rhs: 1..16 columns");
    hipStream_t st = a.st;
    OpArgs<double> A = op_args_impl<double>(a);
    const unsigned grid = seg_grid(a);
    const size_t accl = a.ms * (size_t)r;
    Scratch<double>& s = a.s64;
    const int fl = first_large(a);
    bool small_in_lvl = false;
    if (accl && fl >= 0) {
        FH_DISPATCH_R(r, small_in_lvl = with_small_nk<double, RM>(a.ms, [](auto) {}));
    }
    const bool lds = lds_bytes<double>(a, r) <= LDS_ACC_MAX;
    if (accl && !small_in_lvl) {
        s.part.reserve(std::max<size_t>(lds ? (size_t)grid * accl : accl, 1));
        if (!lds) FH_HIP(hipMemsetAsync(s.part.p, 0, accl * sizeof(double), st));
    }
    s.dpart.reserve((size_t)grid * 2 * r);
    if (fl >= 0) s.zc.reserve(a.M * r);
    const int L = (int)a.seg_lanes;
    const double* wr = a.weighted ? a.w_rows.p : nullptr;
    OpArgs<double> As = A;
    if (small_in_lvl) As.nsb = 0;  // the scatter happens in the level pull
    if (As.nsb == 0) {
        FH_DISPATCH_R(r, hipLaunchKernelGGL((k_rhs_seg_reg<RM>), dim3(grid), dim3(WG), 0, st, As, rows, wr, r, L, b1,
                                            fl >= 0 ? s.zc.p : nullptr, s.dpart.p));
    } else {
        const size_t shm = lds ? lds_bytes<double>(a, r) : 0;
        FH_DISPATCH_R(r, hipLaunchKernelGGL((k_rhs_seg<RM>), dim3(grid), dim3(WG), shm, st, As, rows, wr, r, L, b1,
                                            fl >= 0 ? s.zc.p : nullptr, s.part.p, s.dpart.p, (int)(lds && shm > 0)));
    }
    FH_CHECK_LAUNCH();
    if (accl && !small_in_lvl) {
        if (lds)
            hipLaunchKernelGGL(k_reduce_small<double>, dim3((unsigned)std::min<size_t>(accl, 8192)), dim3(WG), 0, st,
                               s.part.p, (size_t)grid, accl, rhs, nullptr);
        else
            FH_HIP(hipMemcpyAsync(rhs, s.part.p, accl * sizeof(double), hipMemcpyDeviceToDevice, st));
        FH_CHECK_LAUNCH();
    }
    int nsb = 0;
    for (auto& bi : a.blocks) nsb += bi.small ? 1 : 0;
    for (size_t b = 0; b < a.nb; b++) {
        const BlockInfo& bi = a.blocks[b];
        if (bi.small) continue;
        const OtherOffs oo = other_offs(a, (int)b);
        const int Ll = (int)a.lvl_lanes;
        const uint32_t nbk = a.nbk[b];
        const size_t items = (size_t)nbk * bi.g;
        const unsigned lg = lvl_grid(a, items);
        double* ypart = nullptr;
        if (nbk > 1) {
            s.ypart.reserve(items * r);
            ypart = s.ypart.p;
        }
        if ((int)b == fl && small_in_lvl) {
            s.part.reserve((size_t)lg * accl);
            FH_DISPATCH_R(r, with_small_nk<double, RM>(a.ms, [&](auto nk) {
                constexpr int NK = decltype(nk)::value;
                hipLaunchKernelGGL((k_lvl_pull<RM, NK>), dim3(lg), dim3(WG), 0, st, a.boff[b].p, a.bperm[b].p,
                                   a.lother[b].p, (int)a.nb - 1, oo, nsb, (size_t)a.ms, items, (uint32_t)bi.g, s.zc.p,
                                   r, Ll, rhs + bi.off * r, ypart, s.part.p);
            }));
            FH_CHECK_LAUNCH();
            hipLaunchKernelGGL(k_reduce_small<double>, dim3((unsigned)std::min<size_t>(accl, 8192)), dim3(WG), 0, st,
                               s.part.p, (size_t)lg, accl, rhs, nullptr);
        } else {
            FH_DISPATCH_R(r, hipLaunchKernelGGL((k_lvl_pull<RM, 0>), dim3(lg), dim3(WG), 0, st, a.boff[b].p,
                                                a.bperm[b].p, a.lother[b].p, (int)a.nb - 1, oo, 0, (size_t)0, items,
                                                (uint32_t)bi.g, s.zc.p, r, Ll, rhs + bi.off * r, ypart,
                                                (double*)nullptr));
        }
        FH_CHECK_LAUNCH();
        if (ypart) {
            hipLaunchKernelGGL(k_bucket_reduce<double>, dim3(grid_for(bi.g * r, WG, 8192)), dim3(WG), 0, st, ypart,
                               nbk, (size_t)bi.g, r, rhs + bi.off * r, (const uint32_t*)nullptr);
            FH_CHECK_LAUNCH();
        }
    }
    // wsq / wsum: fixed-order sum of the workgroup partials on the host
    std::vector<double> h((size_t)grid * 2 * r);
    FH_HIP(hipMemcpyAsync(h.data(), s.dpart.p, h.size() * sizeof(double), hipMemcpyDeviceToHost, st));
    FH_HIP(hipStreamSynchronize(st));
    for (int j = 0; j < 2 * r; j++) wsq_wsum[j] = 0.0;
    for (unsigned q = 0; q < grid; q++)
        for (int j = 0; j < 2 * r; j++) wsq_wsum[j] += h[(size_t)q * 2 * r + j];
}

void residualize_rows(Absorber& a, double* rows, int r, const double* x, const double* b1, double* alpha) {
    require(r >= 1 && r <= 16, "residualize: 1..16 columns");
    OpArgs<double> A = op_args_impl<double>(a);
    const unsigned grid = seg_grid(a);
    const int L = (int)a.seg_lanes;
    if (a.nb <= 4) {
        FH_DISPATCH_R(r, with_nb0((int)a.nb, [&](auto nbc) {
            constexpr int NB = decltype(nbc)::value;
            hipLaunchKernelGGL((k_resid_seg<RM, NB>), dim3(grid), dim3(WG), 0, a.st, A, rows, r, L, x, b1, alpha);
        }));
    } else {
        FH_DISPATCH_R(r, hipLaunchKernelGGL((k_resid_seg_any<RM>), dim3(grid), dim3(WG), 0, a.st, A, rows, r, L, x, b1,
                                            alpha));
    }
    FH_CHECK_LAUNCH();
}

}  // namespace fh
