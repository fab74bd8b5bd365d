// Householder GMRES on the GPU with the stopping logic of gmres_custom.jl
// (a port of MATLAB gmres.m; gmres_custom.jl:1-331, scalarsignfft :661-668).
//
// The Krylov basis U (N x (inner+1), column-major) and every N-length vector
// live in HBM.  The product of the first j+1 Householder reflectors
// P_0 ... P_j (P_k = I - 2 u_k u_k') is applied in compact WY form,
//     P_0 ... P_j = I - U T U',   T upper triangular, T_kk = 2,
// so each application is two passes over U (U' v, then v - U y) instead of
// j dependent dot/axpy pairs; the reflectors, the Givens rotations, R, w
// and all stopping decisions are exactly those of the reference.  The small
// (inner x inner) quantities -- T, R, J, w and the top inner+1 rows of U --
// are kept on the host in FP64.
#include "context.h"

#include <algorithm>
#include <cmath>
#include <complex>
#include <vector>

typedef std::complex<double> zc;

int sync_stream(peec_ctx *ctx);

namespace {

typedef double2_t cd;

constexpr int kRedBlocks = 512;

__global__ void __launch_bounds__(256) k_norm2_part(const cd *v, int64_t from, int64_t n, double *part)
{
    __shared__ double sh[256];
    double s = 0.0;
    for (int64_t i = from + (int64_t)blockIdx.x * 256 + threadIdx.x; i < n; i += (int64_t)gridDim.x * 256)
        s += v[i].x * v[i].x + v[i].y * v[i].y;
    sh[threadIdx.x] = s;
    __syncthreads();
    for (int o = 128; o > 0; o >>= 1) {
        if (threadIdx.x < o) sh[threadIdx.x] += sh[threadIdx.x + o];
        __syncthreads();
    }
    if (threadIdx.x == 0) part[blockIdx.x] = sh[0];
}

// part[w * ncol + k] = sum over the rows of wavefront w of conj(U[i,k]) v[i]
// (4 wavefronts per block, reduced by lane shuffles: no barriers)
__global__ void __launch_bounds__(256) k_gemvh_part(const cd *U, int64_t ld, int ncol, const cd *v, int64_t n,
                                                    cd *part)
{
    const int lane = threadIdx.x & 63;
    const int64_t w = (int64_t)blockIdx.x * 4 + (threadIdx.x >> 6);
    for (int k = 0; k < ncol; k++) {
        const cd *u = U + (int64_t)k * ld;
        double sx = 0.0, sy = 0.0;
        for (int64_t i = (int64_t)blockIdx.x * 256 + threadIdx.x; i < n; i += (int64_t)gridDim.x * 256) {
            cd a = u[i], b = v[i];
            sx += a.x * b.x + a.y * b.y;
            sy += a.x * b.y - a.y * b.x;
        }
        for (int o = 32; o > 0; o >>= 1) {
            sx += __shfl_xor(sx, o, 64);
            sy += __shfl_xor(sy, o, 64);
        }
        if (lane == 0) part[w * ncol + k] = {sx, sy};
    }
}

// out[k] = sum over the npart partials of column k (one block per column)
__global__ void __launch_bounds__(256) k_gemvh_final(const cd *part, int npart, int ncol, cd *out)
{
    __shared__ double shx[4], shy[4];
    const int k = blockIdx.x;
    double sx = 0.0, sy = 0.0;
    for (int b = threadIdx.x; b < npart; b += 256) {
        sx += part[(int64_t)b * ncol + k].x;
        sy += part[(int64_t)b * ncol + k].y;
    }
    for (int o = 32; o > 0; o >>= 1) {
        sx += __shfl_xor(sx, o, 64);
        sy += __shfl_xor(sy, o, 64);
    }
    if ((threadIdx.x & 63) == 0) {
        shx[threadIdx.x >> 6] = sx;
        shy[threadIdx.x >> 6] = sy;
    }
    __syncthreads();
    if (threadIdx.x == 0) out[k] = {(shx[0] + shx[1]) + (shx[2] + shx[3]), (shy[0] + shy[1]) + (shy[2] + shy[3])};
}

// out = base - U[:, 0:ncol) y, base = in (or 0) + s * e_idx
__global__ void k_gemvn(cd *out, const cd *in, const cd *U, int64_t ld, int ncol, const cd *y, int64_t idx, cd s,
                        int64_t n)
{
    int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= n) return;
    cd a = in ? in[i] : cd{0.0, 0.0};
    if (i == idx) {
        a.x += s.x;
        a.y += s.y;
    }
    for (int k = 0; k < ncol; k++) {
        cd u = U[i + (int64_t)k * ld], c = y[k];
        a.x -= u.x * c.x - u.y * c.y;
        a.y -= u.x * c.y + u.y * c.x;
    }
    out[i] = a;
}

__global__ void k_scale(cd *v, double s, int64_t n)
{
    int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (i < n) v[i] = {v[i].x * s, v[i].y * s};
}

// u = v with u[0..j] = 0 and u[j+1] += alpha
__global__ void k_build_u(cd *u, const cd *v, int64_t j, cd alpha, int64_t n)
{
    int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= n) return;
    cd a = i <= j ? cd{0.0, 0.0} : v[i];
    if (i == j + 1) {
        a.x += alpha.x;
        a.y += alpha.y;
    }
    u[i] = a;
}

// r = b - v
__global__ void k_sub(cd *r, const cd *b, const cd *v, int64_t n)
{
    int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (i < n) r[i] = {b[i].x - v[i].x, b[i].y - v[i].y};
}

// x += a
__global__ void k_add(cd *x, const cd *a, int64_t n)
{
    int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (i < n) x[i] = {x[i].x + a[i].x, x[i].y + a[i].y};
}

}  // namespace

static inline unsigned nb256(int64_t n) { return (unsigned)((n + 255) / 256); }

struct GmresWork {
    peec_ctx *ctx;
    int64_t N;
    int inner;
    DevBuf<cd> U, b, x, r, v, u, xm, xmin, add, part, small;
    std::vector<double> hpart;
    std::vector<cd> hsmall;

    int alloc()
    {
        if (!U.alloc((size_t)N * (inner + 1)) || !b.alloc(N) || !x.alloc(N) || !r.alloc(N) || !v.alloc(N) ||
            !u.alloc(N) || !xm.alloc(N) || !xmin.alloc(N) || !add.alloc(N) ||
            !part.alloc((size_t)kRedBlocks * 4 * (inner + 1)) || !small.alloc((size_t)inner + 1)) {
            peec_set_error("gmres: out of device memory for a %lld x %d Krylov basis", (long long)N, inner + 1);
            return PEEC_ERR_NOMEM;
        }
        hpart.resize(kRedBlocks);
        hsmall.resize((size_t)inner + 1);
        return PEEC_OK;
    }
    int sync()
    {
        PEEC_HIP_TRY(hipStreamSynchronize(ctx->stream));
        return PEEC_OK;
    }
    // ||v[from:]||
    int norm(const cd *vec, int64_t from, double *out)
    {
        int nbk = (int)std::min<int64_t>(kRedBlocks, std::max<int64_t>(1, (N - from + 255) / 256));
        hipLaunchKernelGGL(k_norm2_part, dim3(nbk), dim3(256), 0, ctx->stream, vec, from, N, (double *)part.p);
        PEEC_HIP_TRY(hipGetLastError());
        PEEC_HIP_TRY(hipMemcpyAsync(hpart.data(), part.p, sizeof(double) * nbk, hipMemcpyDeviceToHost, ctx->stream));
        int rc = sync();
        if (rc) return rc;
        double s = 0.0;
        for (int i = 0; i < nbk; i++) s += hpart[(size_t)i];
        *out = std::sqrt(s);
        return PEEC_OK;
    }
    // out[k] = U[:,k]' vec for k < ncol
    int gemvh(int ncol, const cd *vec, std::vector<zc> &out)
    {
        out.assign((size_t)ncol, zc(0, 0));
        if (ncol == 0) return PEEC_OK;
        int nbk = (int)std::min<int64_t>(kRedBlocks, std::max<int64_t>(1, (N + 255) / 256));
        hipLaunchKernelGGL(k_gemvh_part, dim3(nbk), dim3(256), 0, ctx->stream, U.p, N, ncol, vec, N, part.p);
        hipLaunchKernelGGL(k_gemvh_final, dim3(ncol), dim3(256), 0, ctx->stream, part.p, nbk * 4, ncol, small.p);
        PEEC_HIP_TRY(hipGetLastError());
        PEEC_HIP_TRY(hipMemcpyAsync(hsmall.data(), small.p, sizeof(cd) * ncol, hipMemcpyDeviceToHost, ctx->stream));
        int rc = sync();
        if (rc) return rc;
        for (int k = 0; k < ncol; k++) out[(size_t)k] = zc(hsmall[(size_t)k].x, hsmall[(size_t)k].y);
        return PEEC_OK;
    }
    // out = in + s e_idx - U[:, 0:ncol) y  (in may be null)
    int gemvn(cd *out, const cd *in, int ncol, const std::vector<zc> &y, int64_t idx, zc s)
    {
        for (int k = 0; k < ncol; k++) hsmall[(size_t)k] = {y[(size_t)k].real(), y[(size_t)k].imag()};
        if (ncol > 0)
            PEEC_HIP_TRY(hipMemcpyAsync(small.p, hsmall.data(), sizeof(cd) * ncol, hipMemcpyHostToDevice, ctx->stream));
        hipLaunchKernelGGL(k_gemvn, dim3(nb256(N)), dim3(256), 0, ctx->stream, out, in, U.p, N, ncol, small.p, idx,
                           cd{s.real(), s.imag()}, N);
        PEEC_HIP_TRY(hipGetLastError());
        // hsmall is reused by the next call: wait for the copy to be consumed
        return sync();
    }
};

// scalarsignfft (gmres_custom.jl:661-668): sign(d), or 1 for d == 0
static zc scalarsign(zc d)
{
    double a = std::abs(d);
    return a == 0.0 ? zc(1.0, 0.0) : d / a;
}

// Back substitution R[0..k,0..k] \ rhs (R upper triangular)
static std::vector<zc> upper_solve(const std::vector<zc> &R, int ldr, int k, const std::vector<zc> &rhs)
{
    std::vector<zc> y((size_t)k);
    for (int i = k - 1; i >= 0; i--) {
        zc s = rhs[(size_t)i];
        for (int c = i + 1; c < k; c++) s -= R[(size_t)i + (size_t)c * ldr] * y[(size_t)c];
        y[(size_t)i] = s / R[(size_t)i + (size_t)i * ldr];
    }
    return y;
}

extern "C" int peec_gmres(peec_ctx *ctx, const peec_c128 *b_h, double tol, int maxit, peec_c128 *x_h, int *flag_out,
                          double *relres_out, int64_t iter_out[2], double *resvec_out, int64_t *nres_out)
{
    if (!ctx || !b_h || !x_h || !flag_out || !relres_out || !iter_out || !resvec_out || !nres_out || maxit < 0) {
        peec_set_error("peec_gmres: invalid argument");
        return PEEC_ERR_ARG;
    }
    if (!ctx->op_ready || !ctx->freq_ready || !ctx->lu_ready) {
        peec_set_error("peec_gmres: operator, frequency and LU factors must be set first");
        return PEEC_ERR_STATE;
    }
    PEEC_HIP_TRY(hipSetDevice(ctx->device));
    const int64_t n = ctx->m + ctx->ns + ctx->n;
    // restarted == false (FFT_solver_QS_S_type.jl:140): outer = 1, inner = min(maxit, n)
    if (maxit > n) maxit = (int)n;
    const int inner = maxit;
    GmresWork W{ctx, n, inner};
    int rc = W.alloc();
    if (rc) return rc;
    MatvecGraphs graphs;
    graphs.ctx = ctx;
    graphs.nrhs = 1;
    graphs.rps = 1;
    graphs.ld = n;
    auto matvec = [&](const cd *xin, cd *yout) { return graphs.run(xin, yout, false); };

    PEEC_HIP_TRY(hipMemcpyAsync(W.b.p, b_h, sizeof(cd) * n, hipMemcpyHostToDevice, ctx->stream));
    PEEC_HIP_TRY(hipMemcpyAsync(W.x.p, x_h, sizeof(cd) * n, hipMemcpyHostToDevice, ctx->stream));
    double n2b;
    if ((rc = W.norm(W.b.p, 0, &n2b))) return rc;
    auto finish_copy_x = [&](const cd *src) -> int {
        PEEC_HIP_TRY(hipMemcpyAsync(x_h, src, sizeof(cd) * n, hipMemcpyDeviceToHost, ctx->stream));
        return W.sync();
    };
    if (n2b == 0.0) {
        // :21-28
        std::vector<peec_c128> z((size_t)n, peec_c128{0.0, 0.0});
        std::copy(z.begin(), z.end(), x_h);
        *flag_out = 0;
        *relres_out = 0.0;
        iter_out[0] = iter_out[1] = 0;
        resvec_out[0] = 0.0;
        *nres_out = 1;
        return PEEC_OK;
    }
    int flag = 1;
    int imin = 0, jmin = 0;
    double tolb = tol * n2b;
    int evalxm = 0, stag = 0, moresteps = 0;
    const int maxmsteps = (int)std::min<int64_t>(std::min<int64_t>(n / 50, 5), n - maxit);
    const int maxstagsteps = 3;
    int minupdated = 0;
    // xmin = x; xm = x
    PEEC_HIP_TRY(hipMemcpyAsync(W.xmin.p, W.x.p, sizeof(cd) * n, hipMemcpyDeviceToDevice, ctx->stream));
    PEEC_HIP_TRY(hipMemcpyAsync(W.xm.p, W.x.p, sizeof(cd) * n, hipMemcpyDeviceToDevice, ctx->stream));
    // r = b - A x (:48)
    if ((rc = matvec(W.x.p, W.v.p))) return rc;
    hipLaunchKernelGGL(k_sub, dim3(nb256(n)), dim3(256), 0, ctx->stream, W.r.p, W.b.p, W.v.p, n);
    double normr;
    if ((rc = W.norm(W.r.p, 0, &normr))) return rc;
    if ((rc = sync_stream(ctx))) return rc;
    if (normr <= tolb) {
        // :50-57
        if ((rc = finish_copy_x(W.xmin.p))) return rc;
        *flag_out = 0;
        *relres_out = normr / n2b;
        iter_out[0] = iter_out[1] = 0;
        resvec_out[0] = normr;
        *nres_out = 1;
        return PEEC_OK;
    }
    const double n2minv_b = n2b;
    tolb = tol * n2minv_b;
    std::vector<double> resvec((size_t)inner + 1, 0.0);
    resvec[0] = normr;
    double normrmin = normr;
    const int ldr = std::max(inner, 1);
    std::vector<zc> J((size_t)2 * std::max(inner, 1), zc(0, 0)), R((size_t)ldr * ldr, zc(0, 0)),
        w((size_t)inner + 1, zc(0, 0));
    // host mirror of the top inner+1 rows of U and the compact-WY factor T
    const int ldh = inner + 1;
    std::vector<zc> Uh((size_t)ldh * ldh, zc(0, 0)), T((size_t)ldh * ldh, zc(0, 0));
    std::vector<cd> head((size_t)ldh);
    auto fetch_head = [&](const cd *vec, int rows) -> int {
        PEEC_HIP_TRY(hipMemcpyAsync(head.data(), vec, sizeof(cd) * rows, hipMemcpyDeviceToHost, ctx->stream));
        return W.sync();
    };
    // column k of U has been written: mirror its head and extend T
    auto add_reflector = [&](int k) -> int {
        int rows = (int)std::min<int64_t>(ldh, n);
        int e = fetch_head(W.U.p + (int64_t)k * n, rows);
        if (e) return e;
        for (int i = 0; i < rows; i++) Uh[(size_t)i + (size_t)k * ldh] = zc(head[(size_t)i].x, head[(size_t)i].y);
        std::vector<zc> d;
        if ((e = W.gemvh(k, W.U.p + (int64_t)k * n, d))) return e;  // U_{0..k-1}' u_k
        // T[0..k-1, k] = -2 T[0..k-1, 0..k-1] d ; T[k,k] = 2
        for (int i = 0; i < k; i++) {
            zc s(0, 0);
            for (int c = i; c < k; c++) s += T[(size_t)i + (size_t)c * ldh] * d[(size_t)c];
            T[(size_t)i + (size_t)k * ldh] = -2.0 * s;
        }
        T[(size_t)k + (size_t)k * ldh] = 2.0;
        return PEEC_OK;
    };
    // out = Q_{0..j} ytil (ytil has j+1 leading entries): ytil - U T (Uh' ytil)
    auto apply_Q = [&](int j, const std::vector<zc> &yt, cd *out) -> int {
        std::vector<zc> c((size_t)j + 1, zc(0, 0)), z((size_t)j + 1, zc(0, 0));
        for (int k = 0; k <= j; k++) {
            zc s(0, 0);
            for (int i = 0; i <= j; i++) s += std::conj(Uh[(size_t)i + (size_t)k * ldh]) * yt[(size_t)i];
            c[(size_t)k] = s;
        }
        for (int i = 0; i <= j; i++) {
            zc s(0, 0);
            for (int k = i; k <= j; k++) s += T[(size_t)i + (size_t)k * ldh] * c[(size_t)k];
            z[(size_t)i] = s;
        }
        // out = ytil (sparse) - U z : ytil enters as e-components, added after
        int e = W.gemvn(out, nullptr, j + 1, z, -1, zc(0, 0));
        if (e) return e;
        // add the j+1 leading entries of ytil
        std::vector<cd> hy((size_t)j + 1);
        PEEC_HIP_TRY(hipMemcpyAsync(hy.data(), out, sizeof(cd) * (j + 1), hipMemcpyDeviceToHost, ctx->stream));
        if ((e = W.sync())) return e;
        for (int i = 0; i <= j; i++) {
            hy[(size_t)i].x += yt[(size_t)i].real();
            hy[(size_t)i].y += yt[(size_t)i].imag();
        }
        PEEC_HIP_TRY(hipMemcpyAsync(out, hy.data(), sizeof(cd) * (j + 1), hipMemcpyHostToDevice, ctx->stream));
        return W.sync();
    };
    auto dnorm = [&](const cd *vec, double *out) { return W.norm(vec, 0, out); };

    int outitercount = 0, initercount = 0;
    double normr_act = normr;
    int iter0 = 0, iter1 = 0;
    {
        const int outiter = 1;
        outitercount++;
        // u = r + sign(r[1]) ||r|| e1 (:82-87)
        if ((rc = fetch_head(W.r.p, 1))) return rc;
        zc r1(head[0].x, head[0].y);
        zc beta = scalarsign(r1) * normr;
        hipLaunchKernelGGL(k_build_u, dim3(nb256(n)), dim3(256), 0, ctx->stream, W.u.p, W.r.p, (int64_t)-1,
                           cd{0.0, 0.0}, n);
        // u[0] += beta
        {
            cd h0 = {r1.real() + beta.real(), r1.imag() + beta.imag()};
            PEEC_HIP_TRY(hipMemcpyAsync(W.u.p, &h0, sizeof(cd), hipMemcpyHostToDevice, ctx->stream));
            if ((rc = W.sync())) return rc;
        }
        double nu;
        if ((rc = dnorm(W.u.p, &nu))) return rc;
        PEEC_HIP_TRY(hipMemcpyAsync(W.U.p, W.u.p, sizeof(cd) * n, hipMemcpyDeviceToDevice, ctx->stream));
        hipLaunchKernelGGL(k_scale, dim3(nb256(n)), dim3(256), 0, ctx->stream, W.U.p, 1.0 / nu, n);
        if ((rc = add_reflector(0))) return rc;
        w[0] = -beta;
        std::vector<zc> vh((size_t)inner + 2, zc(0, 0));
        for (int initer = 1; initer <= inner; initer++) {
            const int j = initer - 1;  // 0-based column of the current reflector
            if (ctx->stop_requested()) {
                // :96-99
                *flag_out = 99;
                *relres_out = 0.0;
                iter_out[0] = outiter;
                iter_out[1] = initer;
                *nres_out = 0;
                if ((rc = finish_copy_x(W.x.p))) return rc;
                return PEEC_STOPPED;
            }
            initercount++;
            // v = P_0 ... P_j e_j = e_j - U T conj(Uh[j, 0..j])  (:109-117)
            {
                std::vector<zc> c((size_t)j + 1), z((size_t)j + 1);
                for (int k = 0; k <= j; k++) c[(size_t)k] = std::conj(Uh[(size_t)j + (size_t)k * ldh]);
                for (int i = 0; i <= j; i++) {
                    zc s(0, 0);
                    for (int k = i; k <= j; k++) s += T[(size_t)i + (size_t)k * ldh] * c[(size_t)k];
                    z[(size_t)i] = s;
                }
                if ((rc = W.gemvn(W.v.p, nullptr, j + 1, z, j, zc(1.0, 0.0)))) return rc;
            }
            double nv;
            if ((rc = dnorm(W.v.p, &nv))) return rc;
            hipLaunchKernelGGL(k_scale, dim3(nb256(n)), dim3(256), 0, ctx->stream, W.v.p, 1.0 / nv, n);
            // v = A v (:122)
            if ((rc = matvec(W.v.p, W.u.p))) return rc;
            // v = P_j ... P_0 v = v - U T' (U' v)  (:125-127)
            {
                std::vector<zc> d;
                if ((rc = W.gemvh(j + 1, W.u.p, d))) return rc;
                std::vector<zc> z((size_t)j + 1);
                for (int i = 0; i <= j; i++) {
                    zc s(0, 0);
                    for (int k = 0; k <= i; k++) s += std::conj(T[(size_t)k + (size_t)i * ldh]) * d[(size_t)k];
                    z[(size_t)i] = s;
                }
                if ((rc = W.gemvn(W.v.p, W.u.p, j + 1, z, -1, zc(0, 0)))) return rc;
            }
            const int nh = (int)std::min<int64_t>((int64_t)initer + 1, n);
            if ((rc = fetch_head(W.v.p, nh))) return rc;
            for (int i = 0; i < nh; i++) vh[(size_t)i] = zc(head[(size_t)i].x, head[(size_t)i].y);
            // next reflector (:130-145)
            if (initer != n) {
                double alpha_n;
                if ((rc = W.norm(W.v.p, initer, &alpha_n))) return rc;
                if (alpha_n != 0.0) {
                    zc alpha = scalarsign(vh[(size_t)initer]) * alpha_n;
                    hipLaunchKernelGGL(k_build_u, dim3(nb256(n)), dim3(256), 0, ctx->stream, W.u.p, W.v.p, (int64_t)j,
                                       cd{alpha.real(), alpha.imag()}, n);
                    double nu2;
                    if ((rc = dnorm(W.u.p, &nu2))) return rc;
                    if (initer < inner + 1 && initer <= inner) {
                        cd *col = W.U.p + (int64_t)initer * n;
                        PEEC_HIP_TRY(hipMemcpyAsync(col, W.u.p, sizeof(cd) * n, hipMemcpyDeviceToDevice, ctx->stream));
                        hipLaunchKernelGGL(k_scale, dim3(nb256(n)), dim3(256), 0, ctx->stream, col, 1.0 / nu2, n);
                        if ((rc = add_reflector(initer))) return rc;
                    }
                    for (int i = initer + 1; i < nh; i++) vh[(size_t)i] = zc(0, 0);
                    vh[(size_t)initer] = -alpha;
                }
            }
            // Givens (:148-163)
            for (int col = 0; col < j; col++) {
                zc tmpv = vh[(size_t)col];
                vh[(size_t)col] = std::conj(J[2 * (size_t)col]) * vh[(size_t)col] +
                                  std::conj(J[2 * (size_t)col + 1]) * vh[(size_t)col + 1];
                vh[(size_t)col + 1] = -J[2 * (size_t)col + 1] * tmpv + J[2 * (size_t)col] * vh[(size_t)col + 1];
            }
            if (initer != n) {
                double rho = std::hypot(std::abs(vh[(size_t)j]), std::abs(vh[(size_t)j + 1]));
                J[2 * (size_t)j] = vh[(size_t)j] / rho;
                J[2 * (size_t)j + 1] = vh[(size_t)j + 1] / rho;
                w[(size_t)j + 1] = -J[2 * (size_t)j + 1] * w[(size_t)j];
                w[(size_t)j] = std::conj(J[2 * (size_t)j]) * w[(size_t)j];
                vh[(size_t)j] = rho;
                vh[(size_t)j + 1] = 0.0;
            }
            for (int i = 0; i < inner; i++) R[(size_t)i + (size_t)j * ldr] = i < nh ? vh[(size_t)i] : zc(0, 0);
            normr = std::abs(w[(size_t)j + 1]);
            resvec[(size_t)initer] = normr;
            normr_act = normr;
            if (normr <= tolb || stag >= maxstagsteps || moresteps == 1) {
                if (evalxm == 0) {
                    std::vector<zc> wj(w.begin(), w.begin() + initer);
                    std::vector<zc> ytmp = upper_solve(R, ldr, initer, wj);
                    if ((rc = apply_Q(j, ytmp, W.add.p))) return rc;
                    double na, nx;
                    if ((rc = dnorm(W.add.p, &na)) || (rc = dnorm(W.x.p, &nx))) return rc;
                    if (na < 2.220446049250313e-16 * nx) stag += 1;
                    else stag = 0;
                    PEEC_HIP_TRY(hipMemcpyAsync(W.xm.p, W.x.p, sizeof(cd) * n, hipMemcpyDeviceToDevice, ctx->stream));
                    hipLaunchKernelGGL(k_add, dim3(nb256(n)), dim3(256), 0, ctx->stream, W.xm.p, W.add.p, n);
                    evalxm = 1;
                } else if (evalxm == 1) {
                    std::vector<zc> addvc((size_t)initer);
                    zc last = w[(size_t)j] / R[(size_t)j + (size_t)j * ldr];
                    std::vector<zc> col((size_t)j);
                    for (int i = 0; i < j; i++) col[(size_t)i] = R[(size_t)i + (size_t)j * ldr];
                    std::vector<zc> sol = upper_solve(R, ldr, j, col);
                    for (int i = 0; i < j; i++) addvc[(size_t)i] = -sol[(size_t)i] * last;
                    addvc[(size_t)j] = last;
                    double na = 0.0;
                    for (auto &a : addvc) na += std::norm(a);
                    na = std::sqrt(na);
                    double nxm;
                    if ((rc = dnorm(W.xm.p, &nxm))) return rc;
                    if (na < 2.220446049250313e-16 * nxm) stag += 1;
                    else stag = 0;
                    if ((rc = apply_Q(j, addvc, W.add.p))) return rc;
                    hipLaunchKernelGGL(k_add, dim3(nb256(n)), dim3(256), 0, ctx->stream, W.xm.p, W.add.p, n);
                }
                // explicit residual (:202)
                if ((rc = matvec(W.xm.p, W.v.p))) return rc;
                hipLaunchKernelGGL(k_sub, dim3(nb256(n)), dim3(256), 0, ctx->stream, W.r.p, W.b.p, W.v.p, n);
                double nr;
                if ((rc = dnorm(W.r.p, &nr))) return rc;
                if (nr <= tol * n2b) {
                    PEEC_HIP_TRY(hipMemcpyAsync(W.x.p, W.xm.p, sizeof(cd) * n, hipMemcpyDeviceToDevice, ctx->stream));
                    flag = 0;
                    iter0 = outiter;
                    iter1 = initer;
                    break;
                }
                normr_act = nr;
                resvec[(size_t)initer] = normr_act;
                if (normr_act <= normrmin) {
                    normrmin = normr_act;
                    imin = outiter;
                    jmin = initer;
                    PEEC_HIP_TRY(hipMemcpyAsync(W.xmin.p, W.xm.p, sizeof(cd) * n, hipMemcpyDeviceToDevice, ctx->stream));
                    minupdated = 1;
                }
                if (normr_act <= tolb) {
                    PEEC_HIP_TRY(hipMemcpyAsync(W.x.p, W.xm.p, sizeof(cd) * n, hipMemcpyDeviceToDevice, ctx->stream));
                    flag = 0;
                    iter0 = outiter;
                    iter1 = initer;
                    break;
                } else {
                    if (stag >= maxstagsteps && moresteps == 0) stag = 0;
                    moresteps += 1;
                    if (moresteps >= maxmsteps) {
                        flag = 3;
                        iter0 = outiter;
                        iter1 = initer;
                        break;
                    }
                }
            }
            if (normr_act <= normrmin) {
                normrmin = normr_act;
                imin = outiter;
                jmin = initer;
                minupdated = 1;
            }
            if (stag >= maxstagsteps) {
                flag = 3;
                break;
            }
        }
        evalxm = 0;
        if (flag != 0) {
            // :268-289
            int idx = minupdated == 1 ? jmin : initercount;
            if (idx > 0) {
                std::vector<zc> wj(w.begin(), w.begin() + idx);
                std::vector<zc> y = upper_solve(R, ldr, idx, wj);
                if ((rc = apply_Q(idx - 1, y, W.add.p))) return rc;
                hipLaunchKernelGGL(k_add, dim3(nb256(n)), dim3(256), 0, ctx->stream, W.x.p, W.add.p, n);
            }
            PEEC_HIP_TRY(hipMemcpyAsync(W.xmin.p, W.x.p, sizeof(cd) * n, hipMemcpyDeviceToDevice, ctx->stream));
            if ((rc = matvec(W.x.p, W.v.p))) return rc;
            hipLaunchKernelGGL(k_sub, dim3(nb256(n)), dim3(256), 0, ctx->stream, W.r.p, W.b.p, W.v.p, n);
            if ((rc = dnorm(W.r.p, &normr_act))) return rc;
        }
        if (normr_act <= normrmin) {
            PEEC_HIP_TRY(hipMemcpyAsync(W.xmin.p, W.x.p, sizeof(cd) * n, hipMemcpyDeviceToDevice, ctx->stream));
            normrmin = normr_act;
            imin = outiter;
            jmin = initercount;
        }
        if (flag != 3 && normr_act <= tolb) {
            flag = 0;
            iter0 = outitercount;
            iter1 = initercount;
        }
    }
    double relres;
    const cd *xout;
    if (flag == 0) {
        relres = normr_act / n2minv_b;
        xout = W.x.p;
    } else {
        xout = W.xmin.p;
        iter0 = imin;
        iter1 = jmin;
        relres = normr_act / n2minv_b;
    }
    if ((rc = sync_stream(ctx))) return rc;
    if ((rc = finish_copy_x(xout))) return rc;
    int64_t nres = (int64_t)std::max(outitercount - 1, 0) * inner + initercount + 1;
    for (int64_t i = 0; i < nres; i++) resvec_out[i] = resvec[(size_t)i];
    *nres_out = nres;
    *flag_out = flag;
    *relres_out = relres;
    iter_out[0] = iter0;
    iter_out[1] = iter1;
    return PEEC_OK;
}
