// Estimation kernels on the residualized block (FP64): the weighted Gram matrix
// G = sum_i w_i z_i z_i' (stats.rs:195-251), the residual pass e = Z c with sum w e^2, and the
// per-cluster score sums s_g = sum_{i in g} x_i e_i w_i of the CRV1 meat (vcov.rs:320-448), with
// the rows sorted by cluster once (radix sort) so every sum is in a fixed order.
#include <algorithm>
#include <cmath>
#include <utility>

#include <functional>

#include "fh_fit.hpp"
#include "fh_sort.hpp"

namespace fh {

namespace {

constexpr int TPB = 256;
constexpr int GRAM_TILE = 256;
constexpr int GRAM_KMAX = 22;  // k(k+1)/2 <= 256 pairs, one per thread
constexpr int GRID = 1024;

// Per-workgroup partial Gram of columns [c0, c0+K) of the n x ld row-major block z, weights
// w (optional) times v (optional): part[wg][p] over the K(K+1)/2 pairs p = (a <= b) in row-major
// upper order. K is a compile-time constant: the pair loops unroll to straight FMAs. Pairs are
// split into NG groups (thread t takes group t % NG and every (256/NG)-th row); the block adds
// the threads of each group in a fixed order.
template <int K>
__global__ __launch_bounds__(TPB) void k_gram(double* __restrict__ z, size_t n, int ld, int c0,
                                              const double* __restrict__ w, const double* __restrict__ v,
                                              double ztol, double* __restrict__ part) {
    constexpr int NP = K * (K + 1) / 2;
    constexpr int NG = NP <= 36 ? 1 : (NP + 35) / 36;
    constexpr int GPS = (NP + NG - 1) / NG;  // pairs per group
    __shared__ double red[TPB][GPS];
    const int grp = threadIdx.x % NG;
    const int rl = threadIdx.x / NG, nrl = TPB / NG;
    double acc[GPS];
#pragma unroll
    for (int q = 0; q < GPS; q++) acc[q] = 0.0;
    if (rl < nrl) {
        for (size_t i = (size_t)blockIdx.x * nrl + rl; i < n; i += (size_t)gridDim.x * nrl) {
            double x[K];
            double* row = z + i * ld + c0;
#pragma unroll
            for (int c = 0; c < K; c++) x[c] = row[c];
            if (ztol >= 0.0)  // Mata _edittozerotol: tiny entries become 0 in the block too
#pragma unroll
                for (int c = 0; c < K; c++)
                    if (x[c] != 0.0 && fabs(x[c]) <= ztol) {
                        x[c] = 0.0;
                        row[c] = 0.0;
                    }
            double wt = w ? w[i] : 1.0;
            if (v) wt *= v[i];
            int p = 0;
#pragma unroll
            for (int a = 0; a < K; a++)
#pragma unroll
                for (int b = a; b < K; b++, p++)
                    if (NG == 1 || p / GPS == grp) acc[p % GPS] += wt * x[a] * x[b];
        }
    }
#pragma unroll
    for (int q = 0; q < GPS; q++) red[threadIdx.x][q] = acc[q];
    __syncthreads();
    for (int p = threadIdx.x; p < NP; p += TPB) {
        const int g = p / GPS, q = p % GPS;
        double s = 0.0;
        for (int t = g; t < nrl * NG; t += NG) s += red[t][q];
        part[(size_t)blockIdx.x * NP + p] = s;
    }
}

template <int K>
void launch_gram(unsigned grid, hipStream_t st, double* z, size_t n, int ld, int c0, const double* w,
                 const double* v, double ztol, double* part) {
    hipLaunchKernelGGL((k_gram<K>), dim3(grid), dim3(TPB), 0, st, z, n, ld, c0, w, v, ztol, part);
}

template <int... Ks>
void gram_dispatch(int k, unsigned grid, hipStream_t st, double* z, size_t n, int ld, int c0, const double* w,
                   const double* v, double ztol, double* part, std::integer_sequence<int, Ks...>) {
    ((k == Ks + 1 ? launch_gram<Ks + 1>(grid, st, z, n, ld, c0, w, v, ztol, part) : void()), ...);
}

// Per-workgroup partial Gram of transformed rows y_i = z_i[0..R) T (T: R x M, staged in LDS),
// weighted by w: the passes of the Cholesky QR factor, which form the products with Q = A R^-1
// row by row. Pair layout and fixed-order block sums as in k_gram.
template <int M>
__global__ __launch_bounds__(TPB) void k_gram_t(const double* __restrict__ z, size_t n, int ld, int R,
                                                const double* __restrict__ T, const double* __restrict__ w,
                                                double* __restrict__ part) {
    constexpr int NP = M * (M + 1) / 2;
    constexpr int NG = NP <= 36 ? 1 : (NP + 35) / 36;
    constexpr int GPS = (NP + NG - 1) / NG;
    __shared__ double Ts[GRAM_KMAX * GRAM_KMAX];
    __shared__ double red[TPB][GPS];
    for (int q = threadIdx.x; q < R * M; q += TPB) Ts[q] = T[q];
    __syncthreads();
    const int grp = threadIdx.x % NG;
    const int rl = threadIdx.x / NG, nrl = TPB / NG;
    double acc[GPS];
#pragma unroll
    for (int q = 0; q < GPS; q++) acc[q] = 0.0;
    if (rl < nrl) {
        for (size_t i = (size_t)blockIdx.x * nrl + rl; i < n; i += (size_t)gridDim.x * nrl) {
            double y[M];
#pragma unroll
            for (int c = 0; c < M; c++) y[c] = 0.0;
            const double* row = z + i * ld;
            for (int a = 0; a < R; a++) {
                const double xa = row[a];
#pragma unroll
                for (int c = 0; c < M; c++) y[c] += xa * Ts[a * M + c];
            }
            const double wt = w ? w[i] : 1.0;
            int p = 0;
#pragma unroll
            for (int a = 0; a < M; a++)
#pragma unroll
                for (int b = a; b < M; b++, p++)
                    if (NG == 1 || p / GPS == grp) acc[p % GPS] += wt * y[a] * y[b];
        }
    }
#pragma unroll
    for (int q = 0; q < GPS; q++) red[threadIdx.x][q] = acc[q];
    __syncthreads();
    for (int p = threadIdx.x; p < NP; p += TPB) {
        const int g = p / GPS, q = p % GPS;
        double sm = 0.0;
        for (int t = g; t < nrl * NG; t += NG) sm += red[t][q];
        part[(size_t)blockIdx.x * NP + p] = sm;
    }
}

template <int M>
void launch_gram_t(unsigned grid, hipStream_t st, const double* z, size_t n, int ld, int R, const double* T,
                   const double* w, double* part) {
    hipLaunchKernelGGL((k_gram_t<M>), dim3(grid), dim3(TPB), 0, st, z, n, ld, R, T, w, part);
}

template <int... Ms>
void gram_t_dispatch(int m, unsigned grid, hipStream_t st, const double* z, size_t n, int ld, int R, const double* T,
                     const double* w, double* part, std::integer_sequence<int, Ms...>) {
    ((m == Ms + 1 ? launch_gram_t<Ms + 1>(grid, st, z, n, ld, R, T, w, part) : void()), ...);
}

// e_i = sum_j c_j z_ij; partial sums of w e^2 per workgroup.
// e_i = sum_{j < k} c_j z[i][j] (and z[i][ecol] = e_i when ecol >= 0); part[wg] = partial
// sums of w_i e_i^2.
__global__ __launch_bounds__(TPB) void k_resid(double* __restrict__ z, size_t n, int ld, int k,
                                               const double* __restrict__ c, const double* __restrict__ w,
                                               double* __restrict__ e, int ecol, double* __restrict__ part) {
    __shared__ double cc[32];
    if (threadIdx.x < 32u) cc[threadIdx.x] = threadIdx.x < (unsigned)k ? c[threadIdx.x] : 0.0;
    __syncthreads();
    double s = 0.0;
    for (size_t i = (size_t)blockIdx.x * TPB + threadIdx.x; i < n; i += (size_t)gridDim.x * TPB) {
        double v = 0.0;
        double* zi = z + i * ld;
        int j = 0;
        if ((ld & 1) == 0)  // 16-byte aligned rows: paired loads
            for (; j + 1 < k; j += 2) {
                const double2 q = *reinterpret_cast<const double2*>(zi + j);
                v += cc[j] * q.x;
                v += cc[j + 1] * q.y;
            }
        for (; j < k; j++) v += cc[j] * zi[j];
        e[i] = v;
        if (ecol >= 0) zi[ecol] = v;
        s += (w ? w[i] : 1.0) * v * v;
    }
    __shared__ double red[TPB / WAVE];
    s = wave_sum(s);
    if (lane_id() == 0) red[threadIdx.x / WAVE] = s;
    __syncthreads();
    if (threadIdx.x == 0) {
        double t = 0.0;
        for (int q = 0; q < TPB / WAVE; q++) t += red[q];
        part[blockIdx.x] = t;
    }
}

// The same with a compile-time stride LD (even, 16-byte aligned rows): the row is loaded whole
// (LD / 2 loads of 16 bytes issued together) before the products, which the runtime column loop
// above waits for one by one. Same products in the same order as k_resid.
template <int LD>
__global__ __launch_bounds__(TPB) void k_resid_ld(double* __restrict__ z, size_t n, int k,
                                                  const double* __restrict__ c, const double* __restrict__ w,
                                                  double* __restrict__ e, int ecol, double* __restrict__ part) {
    __shared__ double cc[32];
    if (threadIdx.x < 32u) cc[threadIdx.x] = threadIdx.x < (unsigned)k ? c[threadIdx.x] : 0.0;
    __syncthreads();
    double s = 0.0;
    for (size_t i = (size_t)blockIdx.x * TPB + threadIdx.x; i < n; i += (size_t)gridDim.x * TPB) {
        double* zi = z + i * LD;
        double row[LD];
#pragma unroll
        for (int q = 0; q < LD / 2; q++) {
            const double2 v2 = reinterpret_cast<const double2*>(zi)[q];
            row[2 * q] = v2.x;
            row[2 * q + 1] = v2.y;
        }
        double v = 0.0;
#pragma unroll
        for (int j = 0; j < LD; j++)
            if (j < k) v += cc[j] * row[j];
        e[i] = v;
        if (ecol >= 0) zi[ecol] = v;
        s += (w ? w[i] : 1.0) * v * v;
    }
    __shared__ double red[TPB / WAVE];
    s = wave_sum(s);
    if (lane_id() == 0) red[threadIdx.x / WAVE] = s;
    __syncthreads();
    if (threadIdx.x == 0) {
        double t = 0.0;
        for (int q = 0; q < TPB / WAVE; q++) t += red[q];
        part[blockIdx.x] = t;
    }
}

// v_i = (w_i e_i)^2 (hetero meat weights, fit/vcov.rs:322-332).
__global__ void k_sq_we(const double* __restrict__ e, const double* __restrict__ w, size_t n, double* __restrict__ v) {
    for (size_t i = (size_t)blockIdx.x * blockDim.x + threadIdx.x; i < n; i += (size_t)gridDim.x * blockDim.x) {
        const double s = e[i] * (w ? w[i] : 1.0);
        v[i] = s * s;
    }
}

// flag[i] = 1 where a run of adjacent solver rows with one cluster code starts.
__global__ void k_run_flags(const uint32_t* __restrict__ cl_kept, const uint32_t* __restrict__ pi, size_t n,
                            uint32_t* __restrict__ flag) {
    for (size_t i = (size_t)blockIdx.x * blockDim.x + threadIdx.x; i < n; i += (size_t)gridDim.x * blockDim.x)
        flag[i] = (i == 0 || cl_kept[pi[i]] != cl_kept[pi[i - 1]]) ? 1u : 0u;
}

// At each run start i (run q = pos[i]): run_off[q] = i, key[q] = its cluster, val[q] = q; and
// run_off[nrun] = n.
__global__ void k_run_emit(const uint32_t* __restrict__ flag, const uint32_t* __restrict__ pos,
                           const uint32_t* __restrict__ cl_kept, const uint32_t* __restrict__ pi, size_t n,
                           uint32_t* __restrict__ run_off, uint64_t* __restrict__ key, uint32_t* __restrict__ val) {
    for (size_t i = (size_t)blockIdx.x * blockDim.x + threadIdx.x; i < n; i += (size_t)gridDim.x * blockDim.x) {
        const uint32_t q = pos[i];
        if (flag[i]) {
            run_off[q] = (uint32_t)i;
            key[q] = cl_kept[pi[i]];
            val[q] = q;
        }
        if (i == n - 1) run_off[q + flag[i]] = (uint32_t)n;
    }
}

__global__ void k_cl_offsets(const uint64_t* __restrict__ key, size_t n, size_t g, uint32_t* __restrict__ off) {
    for (size_t c = (size_t)blockIdx.x * blockDim.x + threadIdx.x; c < n; c += (size_t)gridDim.x * blockDim.x) {
        uint64_t s = key[c];
        int64_t prev = c ? (int64_t)key[c - 1] : -1;
        for (int64_t k = prev + 1; k <= (int64_t)s; k++) off[k] = (uint32_t)c;
        if (c == n - 1)
            for (size_t k = (size_t)s + 1; k <= g; k++) off[k] = (uint32_t)n;
    }
}

// sg[q] = sum over the rows i of cluster q of z[i][c0 .. c0+k) e_i w_i: L lanes per cluster,
// each summing whole runs of adjacent rows (the cluster's runs strided over the lanes), then the
// lanes reduced by shuffles in a fixed order.
template <int KM>
__global__ __launch_bounds__(TPB) void k_cluster_scores(const uint32_t* __restrict__ off,
                                                        const uint32_t* __restrict__ order,
                                                        const uint32_t* __restrict__ run_off, size_t g,
                                                        const double* __restrict__ z, int ld, int c0, int k,
                                                        const double* __restrict__ e, size_t es,
                                                        const double* __restrict__ w, int L, double* __restrict__ sg) {
    const int l = lane_id();
    const int gpw = WAVE / L, gl = l / L, li = l % L;
    const size_t nwaves = (size_t)gridDim.x * (TPB / WAVE);
    for (size_t base = ((size_t)blockIdx.x * (TPB / WAVE) + threadIdx.x / WAVE) * gpw; base < g; base += nwaves * gpw) {
        const size_t q = base + gl;
        const bool valid = q < g;
        double acc[KM];
#pragma unroll
        for (int a = 0; a < KM; a++) acc[a] = 0.0;
        const uint32_t j0 = valid ? off[q] : 0, j1 = valid ? off[q + 1] : 0;
        for (uint32_t j = j0 + li; j < j1; j += L) {
            const uint32_t run = order[j];
            const uint32_t i1 = run_off[run + 1];
            for (uint32_t i = run_off[run]; i < i1; i++) {
                const double s = e[(size_t)i * es] * (w ? w[i] : 1.0);
                const double* zi = z + (size_t)i * ld + c0;
#pragma unroll
                for (int a = 0; a < KM; a++)
                    if (a < k) acc[a] += zi[a] * s;
            }
        }
#pragma unroll
        for (int a = 0; a < KM; a++)
            for (int o = L >> 1; o > 0; o >>= 1) acc[a] += __shfl_xor(acc[a], o, WAVE);
        if (valid && li == 0)
#pragma unroll
            for (int a = 0; a < KM; a++)
                if (a < k) sg[q * k + a] = acc[a];
    }
}

__global__ void k_count_nonempty(const uint32_t* __restrict__ off, size_t g, unsigned long long* cnt) {
    for (size_t q = (size_t)blockIdx.x * blockDim.x + threadIdx.x; q < g; q += (size_t)gridDim.x * blockDim.x)
        if (off[q + 1] > off[q]) atomicAdd(cnt, 1ull);
}

}  // namespace

// out[e] = sum over q < nparts of part[q * len + e], one workgroup per entry: strided thread sums,
// then the waves in order (a fixed order independent of the device's scheduling).
__global__ __launch_bounds__(TPB) void k_sum_parts(const double* __restrict__ part, size_t nparts, size_t len,
                                                   double* __restrict__ out) {
    __shared__ double red[TPB / WAVE];
    for (size_t e = blockIdx.x; e < len; e += gridDim.x) {
        double sm = 0.0;
        for (size_t q = threadIdx.x; q < nparts; q += TPB) sm += part[q * len + e];
        sm = wave_sum(sm);
        if (lane_id() == 0) red[threadIdx.x / WAVE] = sm;
        __syncthreads();
        if (threadIdx.x == 0) {
            double t = 0.0;
            for (int w = 0; w < TPB / WAVE; w++) t += red[w];
            out[e] = t;
        }
        __syncthreads();
    }
}

void sum_parts_to(const double* part, size_t nparts, size_t len, double* out, hipStream_t st) {
    if (len == 0) return;
    hipLaunchKernelGGL(k_sum_parts, dim3((unsigned)std::min<size_t>(len, 4096)), dim3(TPB), 0, st, part, nparts, len,
                       out);
    FH_CHECK_LAUNCH();
}

std::vector<double> sum_parts_dev(const double* part, size_t nparts, size_t len, hipStream_t st) {
    std::vector<double> h(len, 0.0);
    if (len == 0) return h;
    DBuf<double> out(len);
    sum_parts_to(part, nparts, len, out.p, st);
    download(h.data(), out.p, len * sizeof(double), st);
    return h;
}

void gram_cols(Absorber& a, const double* z, size_t n, int ld, int c0, int k, const double* w, const double* v,
               double* G_host, double ztol) {
    require(k >= 1 && k <= GRAM_KMAX, "at most 21 regressors are supported");
    const int npair = k * (k + 1) / 2;
    const unsigned grid = std::max(1u, std::min<unsigned>(GRID, (unsigned)((n + 255) / 256)));
    DBuf<double> part((size_t)grid * npair);
    // the block is written only where ztol edits an entry
    gram_dispatch(k, grid, a.st, const_cast<double*>(z), n, ld, c0, w, v, ztol, part.p,
                  std::make_integer_sequence<int, GRAM_KMAX>{});
    FH_CHECK_LAUNCH();
    const std::vector<double> s = sum_parts_dev(part.p, grid, (size_t)npair, a.st);
    int p = 0;
    for (int x = 0; x < k; x++)
        for (int y = x; y < k; y++, p++) {
            G_host[x * k + y] = s[p];
            G_host[y * k + x] = s[p];
        }
}

void gram_transformed(Absorber& a, const double* z, size_t n, int ld, int R, const std::vector<double>& T, int m,
                      const double* w, double* G_host) {
    require(m >= 1 && m <= GRAM_KMAX && R >= 1 && R <= GRAM_KMAX && (int)T.size() == R * m,
            "transformed Gram: bad shape");
    const int npair = m * (m + 1) / 2;
    const unsigned grid = std::max(1u, std::min<unsigned>(GRID, (unsigned)((n + 255) / 256)));
    DBuf<double> part((size_t)grid * npair), Td(T.size());
    upload(Td.p, T.data(), T.size() * sizeof(double), a.st);
    gram_t_dispatch(m, grid, a.st, z, n, ld, R, Td.p, w, part.p, std::make_integer_sequence<int, GRAM_KMAX>{});
    FH_CHECK_LAUNCH();
    const std::vector<double> s = sum_parts_dev(part.p, grid, (size_t)npair, a.st);
    int p = 0;
    for (int x = 0; x < m; x++)
        for (int y = x; y < m; y++, p++) {
            G_host[x * m + y] = s[p];
            G_host[y * m + x] = s[p];
        }
}

bool qr_factor(Absorber& a, const double* z, size_t n, int ld, int R, const std::vector<double>& T0, int m,
               const double* w, const std::vector<double>& G0, std::vector<double>& Rf) {
    // Shifted Cholesky QR with two refinement passes (Fukaya, Kannan, Nakatsukasa, Yamamoto and
    // Yanagisawa, SIAM J. Sci. Comput. 42(1), 2020): R1 = chol(A'A + s I) with
    // s = 11 (m n + m (m + 1)) u ||A||_F^2 exists for any A with cond(A) < 1/u; each further
    // pass forms Q = A R^-1 row by row, factors Q'Q and folds its factor into R. Two passes bring
    // Q to orthogonality at the working precision, so R is the triangular factor of A's QR
    // decomposition up to rounding (positive diagonal).
    double tr = 0.0;
    for (int j = 0; j < m; j++) tr += G0[(size_t)j * m + j];
    if (!(tr > 0.0) || !std::isfinite(tr)) return false;
    const double s = 11.0 * ((double)m * (double)n + (double)m * (m + 1)) * 1.1102230246251565e-16 * tr;
    std::vector<double> G1(G0);
    for (int j = 0; j < m; j++) G1[(size_t)j * m + j] += s;
    std::vector<double> Racc, Rp, G((size_t)m * m);
    if (!chol_upper(G1, m, Racc)) return false;
    for (int pass = 0; pass < 2; pass++) {
        const std::vector<double> Ri = upper_inverse(Racc, m);
        std::vector<double> T((size_t)R * m, 0.0);  // T0 Racc^-1
        for (int r0 = 0; r0 < R; r0++)
            for (int c = 0; c < m; c++) {
                double v = 0.0;
                for (int l = 0; l <= c; l++) v += T0[(size_t)r0 * m + l] * Ri[(size_t)l * m + c];
                T[(size_t)r0 * m + c] = v;
            }
        gram_transformed(a, z, n, ld, R, T, m, w, G.data());
        if (!chol_upper(G, m, Rp)) return false;
        Racc = upper_mul(Rp, Racc, m);
    }
    Rf = Racc;
    return true;
}

void gram(Absorber& a, const double* z, const double* w, size_t n, int r, double* G_host) {
    gram_cols(a, z, n, r, 0, r, w, nullptr, G_host);
}

double residual_pass(Absorber& a, double* z, size_t n, int ld, int k, const double* c_host, const double* w, double* e,
                     int ecol) {
    require(k >= 1 && k <= 32 && k <= ld && ecol < ld, "residual pass: bad column layout");
    DBuf<double> c(k), part(GRID);
    upload(c.p, c_host, k * sizeof(double), a.st);
    const bool al = (reinterpret_cast<uintptr_t>(z) & 15) == 0;
    if (al && ld == 4) hipLaunchKernelGGL(k_resid_ld<4>, dim3(GRID), dim3(TPB), 0, a.st, z, n, k, c.p, w, e, ecol, part.p);
    else if (al && ld == 8) hipLaunchKernelGGL(k_resid_ld<8>, dim3(GRID), dim3(TPB), 0, a.st, z, n, k, c.p, w, e, ecol, part.p);
    else if (al && ld == 12) hipLaunchKernelGGL(k_resid_ld<12>, dim3(GRID), dim3(TPB), 0, a.st, z, n, k, c.p, w, e, ecol, part.p);
    else if (al && ld == 16) hipLaunchKernelGGL(k_resid_ld<16>, dim3(GRID), dim3(TPB), 0, a.st, z, n, k, c.p, w, e, ecol, part.p);
    else hipLaunchKernelGGL(k_resid, dim3(GRID), dim3(TPB), 0, a.st, z, n, ld, k, c.p, w, e, ecol, part.p);
    FH_CHECK_LAUNCH();
    return sum_parts_dev(part.p, GRID, 1, a.st)[0];
}

void hetero_meat(Absorber& a, const double* z, size_t n, int ld, int c0, int k, const double* e, const double* w,
                 double* meat) {
    DBuf<double> v(n);
    hipLaunchKernelGGL(k_sq_we, dim3(grid_for(n, TPB, 4096)), dim3(TPB), 0, a.st, e, w, n, v.p);
    FH_CHECK_LAUNCH();
    gram_cols(a, z, n, ld, c0, k, nullptr, v.p, meat);
}

void cluster_runs(Absorber& a, const uint32_t* cl_kept_dev, uint64_t G, ClusterRuns& cr) {
    hipStream_t st = a.st;
    const size_t n = a.n;
    cr.cl = cl_kept_dev;
    cr.G = G;
    cr.off.alloc(G + 1);
    FH_HIP(hipMemsetAsync(cr.off.p, 0, cr.off.bytes(), st));
    cr.nrun = 0;
    if (n == 0) return;
    DBuf<uint32_t> flag(n), pos(n), total(1);
    hipLaunchKernelGGL(k_run_flags, dim3(grid_for(n, TPB, 4096)), dim3(TPB), 0, st, cl_kept_dev, a.pi.p, n, flag.p);
    FH_CHECK_LAUNCH();
    exclusive_scan_u32(flag.p, pos.p, n, total.p, st);
    uint32_t nr = 0;
    download(&nr, total.p, sizeof(nr), st);
    cr.nrun = nr;
    cr.run_off.alloc(nr + 1);
    cr.order.alloc(nr);
    DBuf<uint64_t> keys(nr), kalt(nr);
    DBuf<uint32_t> valt(nr);
    hipLaunchKernelGGL(k_run_emit, dim3(grid_for(n, TPB, 4096)), dim3(TPB), 0, st, flag.p, pos.p, cl_kept_dev,
                       a.pi.p, n, cr.run_off.p, keys.p, cr.order.p);
    FH_CHECK_LAUNCH();
    // stable: a cluster's runs stay in row order
    radix_sort_pairs(keys, cr.order, kalt, valt, nr, (int)bits_for(G), st);
    hipLaunchKernelGGL(k_cl_offsets, dim3(grid_for(nr, TPB, 4096)), dim3(TPB), 0, st, keys.p, (size_t)nr, (size_t)G,
                       cr.off.p);
    FH_CHECK_LAUNCH();
}

const ClusterRuns& runs_for(Absorber& a, const uint32_t* cl_kept_dev, uint64_t G, ClusterRuns& local) {
    if (a.cl_runs && a.cl_runs->cl == cl_kept_dev && a.cl_runs->G == G) return *a.cl_runs;
    cluster_runs(a, cl_kept_dev, G, local);
    return local;
}

uint32_t lanes_per_cluster(const ClusterRuns& cr) {
    uint32_t L = 1;
    while (L < 64 && L * 2.0 < double(cr.nrun) / double(std::max<uint64_t>(cr.G, 1))) L <<= 1;
    return L;
}

uint64_t count_nonempty_clusters(Absorber& a, const uint32_t* cl, uint64_t G) {
    ClusterRuns local;
    const ClusterRuns& cr = runs_for(a, cl, G, local);
    DBuf<unsigned long long> cnt(1);
    FH_HIP(hipMemsetAsync(cnt.p, 0, sizeof(unsigned long long), a.st));
    hipLaunchKernelGGL(k_count_nonempty, dim3(grid_for(G, TPB, 4096)), dim3(TPB), 0, a.st, cr.off.p, (size_t)G, cnt.p);
    FH_CHECK_LAUNCH();
    unsigned long long c = 0;
    download(&c, cnt.p, sizeof(c), a.st);
    return c;
}

// Returns the number of non-empty clusters.
uint64_t cluster_meat(Absorber& a, const double* z, size_t n, int ld, int c0, int k, const double* e, size_t es,
                      const double* w, const uint32_t* cl_kept_dev, uint64_t G, double* meat) {
    hipStream_t st = a.st;
    ClusterRuns local;
    const ClusterRuns& cr = runs_for(a, cl_kept_dev, G, local);
    const uint32_t* off = cr.off.p;
    DBuf<double> sg(G * k);
    const uint32_t L = lanes_per_cluster(cr);
    const unsigned cg = grid_for(G, (TPB / WAVE) * (WAVE / L), 1u << 16);
    auto launch = [&](auto km) {
        constexpr int KM = decltype(km)::value;
        hipLaunchKernelGGL((k_cluster_scores<KM>), dim3(cg), dim3(TPB), 0, st, off, cr.order.p, cr.run_off.p,
                           (size_t)G, z, ld, c0, k, e, es, w, (int)L, sg.p);
    };
    if (k <= 4) launch(std::integral_constant<int, 4>{});
    else if (k <= 12) launch(std::integral_constant<int, 12>{});
    else launch(std::integral_constant<int, GRAM_KMAX>{});
    DBuf<unsigned long long> cnt(1);
    FH_HIP(hipMemsetAsync(cnt.p, 0, sizeof(unsigned long long), st));
    hipLaunchKernelGGL(k_count_nonempty, dim3(grid_for(G, TPB, 4096)), dim3(TPB), 0, st, off, (size_t)G, cnt.p);
    FH_CHECK_LAUNCH();
    gram_cols(a, sg.p, G, k, 0, k, nullptr, nullptr, meat);
    unsigned long long c = 0;
    download(&c, cnt.p, sizeof(c), st);
    stream_wait(st);
    return c;
}

uint64_t multiway_meat(Absorber& a, const double* z, int ld, int c0, int k, const double* e, size_t es,
                       const double* w, const ClusterSet& cs, bool conventional, double* meat) {
    const int q = (int)cs.dims.size();
    require(q >= 1 && q <= 4, "1..4 cluster dimensions are supported");
    const size_t kk = (size_t)k * k;
    std::fill(meat, meat + kk, 0.0);
    std::vector<double> part(kk);
    DBuf<uint32_t> inter(q > 1 ? a.n : 0);
    uint64_t gmin = UINT64_MAX;
    // subsets by size, each size in lexicographic order (fit/vcov.rs:196-214)
    for (int size = 1; size <= q; size++) {
        const double sign = (size % 2) ? 1.0 : -1.0;
        std::vector<std::vector<int>> subs;
        std::vector<int> cur;
        std::function<void(int)> rec = [&](int start) {
            if ((int)cur.size() == size) {
                subs.push_back(cur);
                return;
            }
            for (int d = start; d < q; d++) {
                cur.push_back(d);
                rec(d + 1);
                cur.pop_back();
            }
        };
        rec(0);
        for (const auto& sub : subs) {
            const uint32_t* codes;
            uint64_t G;
            if (size == 1) {
                codes = cs.dims[sub[0]];
                G = cs.G[sub[0]];
            } else {
                std::vector<const uint32_t*> parts;
                std::vector<uint64_t> pg;
                for (int d : sub) {
                    parts.push_back(cs.dims[d]);
                    pg.push_back(cs.G[d]);
                }
                G = combine_dev(parts, pg, a.n, inter.p, a.st);
                codes = inter.p;
            }
            const uint64_t g = cluster_meat(a, z, a.n, ld, c0, k, e, es, w, codes, G, part.data());
            if (size == 1) gmin = std::min(gmin, g);
            const double f = sign * (conventional ? (double)g / ((double)g - 1.0) : 1.0);
            for (size_t i = 0; i < kk; i++) meat[i] += f * part[i];
        }
    }
    return gmin;
}

}  // namespace fh
