// api_check.cpp -- calls the reference-signature C++ API of this library
// (blco.hpp, alto_dev.hpp, cpd_gpu.hpp, kruskal_model.hpp) on the GPU and
// checks every result on the host:
//   mttkrp_alto_dev_onemode<LIT>  against a long-double COO MTTKRP, every mode
//   ata_gpu                       against a long-double U^T U
//   columnwise_normscal           unit columns, lambda = column norms
//   pseudoinverse_gpu             the four Moore-Penrose conditions, on a
//                                 general and on a rank-deficient matrix
//   kruskal_fit_gpu               against the fit formed on the host
//   vector_hadamard / value_fill / vecsum
//   the device blocks of gen_blcotensor_device (ALTO-record fields)
// Prints one "CHECK <name> <error> <tolerance> PASS|FAIL" line per check and
// exits 1 when any fails (2 without a GPU).
#include <hip/hip_runtime.h>

#include <cmath>
#include <cstdint>
#include <cstdio>
#include <cstring>
#include <vector>

#include "alto_dev.hpp"
#include "blco.hpp"
#include "blco_hip.h"
#include "cpd_gpu.hpp"
#include "kruskal_model.hpp"
#include "sptensor.hpp"

static int g_fail = 0;

static void check(const char* name, double err, double tol)
{
    const bool ok = std::isfinite(err) && err <= tol;
    printf("CHECK %s %.3e %.1e %s\n", name, err, tol, ok ? "PASS" : "FAIL");
    if (!ok) g_fail = 1;
}

static double* dev_alloc(size_t n)
{
    double* p = nullptr;
    if (hipMalloc(&p, n * sizeof(double)) != hipSuccess) {
        fprintf(stderr, "hipMalloc failed\n");
        exit(1);
    }
    return p;
}

static void up(double* d, const std::vector<double>& h)
{
    (void)hipMemcpy(d, h.data(), h.size() * sizeof(double), hipMemcpyHostToDevice);
}

static std::vector<double> down(const double* d, size_t n)
{
    std::vector<double> h(n);
    (void)hipMemcpy(h.data(), d, n * sizeof(double), hipMemcpyDeviceToHost);
    return h;
}

// splitmix64 uniform in [0, 1)
static uint64_t g_state = 0x1234567ull;
static double urand()
{
    uint64_t z = (g_state += 0x9E3779B97F4A7C15ull);
    z = (z ^ (z >> 30)) * 0xBF58476D1CE4E5B9ull;
    z = (z ^ (z >> 27)) * 0x94D049BB133111EBull;
    return (double)((z ^ (z >> 31)) >> 11) * 0x1.0p-53;
}

// n x n row-major product
static std::vector<double> matmul(const std::vector<double>& a, const std::vector<double>& b, int n)
{
    std::vector<double> c((size_t)n * n, 0.0);
    for (int i = 0; i < n; ++i)
        for (int k = 0; k < n; ++k)
            for (int j = 0; j < n; ++j) c[i * n + j] += a[i * n + k] * b[k * n + j];
    return c;
}

static double rel_diff(const std::vector<double>& a, const std::vector<double>& b)
{
    double num = 0.0, den = 0.0;
    for (size_t i = 0; i < a.size(); ++i) {
        num += (a[i] - b[i]) * (a[i] - b[i]);
        den += b[i] * b[i];
    }
    return std::sqrt(num) / (den > 0 ? std::sqrt(den) : 1.0);
}

static double asym(const std::vector<double>& a, int n)
{
    double num = 0.0, den = 0.0;
    for (int i = 0; i < n; ++i)
        for (int j = 0; j < n; ++j) {
            num += (a[i * n + j] - a[j * n + i]) * (a[i * n + j] - a[j * n + i]);
            den += a[i * n + j] * a[i * n + j];
        }
    return std::sqrt(num) / (den > 0 ? std::sqrt(den) : 1.0);
}

static void check_pinv(const char* name, std::vector<double> a, int n)
{
    double* d = dev_alloc((size_t)n * n);
    int* info = nullptr;
    (void)hipMalloc(&info, sizeof(int));
    up(d, a);
    pseudoinverse_gpu(nullptr, d, n, nullptr, 0, info);
    std::vector<double> p = down(d, (size_t)n * n);
    int h_info = -1;
    (void)hipMemcpy(&h_info, info, sizeof(int), hipMemcpyDeviceToHost);
    const std::vector<double> ap = matmul(a, p, n), pa = matmul(p, a, n);
    char buf[128];
    snprintf(buf, sizeof buf, "pinv_%s_info", name);
    check(buf, (double)h_info, 0.0);
    snprintf(buf, sizeof buf, "pinv_%s_AXA", name);
    check(buf, rel_diff(matmul(ap, a, n), a), 1e-10);
    snprintf(buf, sizeof buf, "pinv_%s_XAX", name);
    check(buf, rel_diff(matmul(pa, p, n), p), 1e-10);
    snprintf(buf, sizeof buf, "pinv_%s_AX_sym", name);
    check(buf, asym(ap, n), 1e-10);
    snprintf(buf, sizeof buf, "pinv_%s_XA_sym", name);
    check(buf, asym(pa, n), 1e-10);
    (void)hipFree(d);
    (void)hipFree(info);
}

int main()
{
    if (blco_device_count() < 1) {
        fprintf(stderr, "api_check needs a GPU\n");
        return 2;
    }
    // ---- a seeded 3-mode tensor with distinct coordinates
    IType dims[3] = {57, 43, 31};
    const int N = 3, R = 16;
    std::vector<IType> cidx;
    std::vector<FType> vals;
    std::vector<char> seen(dims[0] * dims[1] * dims[2], 0);
    while (vals.size() < 4000) {
        IType i = (IType)(urand() * dims[0]), j = (IType)(urand() * dims[1]), k = (IType)(urand() * dims[2]);
        if (seen[(i * dims[1] + j) * dims[2] + k]) continue;
        seen[(i * dims[1] + j) * dims[2] + k] = 1;
        cidx.insert(cidx.end(), {i, j, k});
        vals.push_back(0.5 + urand());
    }
    const IType nnz = vals.size();
    SparseTensor* X = nullptr;
    CreateSparseTensor(N, dims, nnz, cidx.data(), vals.data(), &X);
    blcotensor* Xh = gen_blcotensor_host<unsigned long long>(X, 1000);
    blcotensor* Xd = gen_blcotensor_device(Xh, Xh->block_count, 0, false);

    // the device blocks carry the ALTO record: dims, masks, positions, and
    // coordinates that equal the host blocks'
    {
        double err = 0.0;
        IType total = 0;
        for (IType b = 0; b < Xd->block_count; ++b) {
            const blco_block* hb = Xh->blocks[b];
            const blco_block* db = Xd->blocks[b];
            std::vector<uint64_t> c(db->nnz);
            (void)hipMemcpy(c.data(), db->idx, db->nnz * 8, hipMemcpyDeviceToHost);
            for (IType e = 0; e < db->nnz; ++e) err += c[e] != hb->idx[e];
            std::vector<double> v = down(db->vals, db->nnz);
            for (IType e = 0; e < db->nnz; ++e) err += v[e] != hb->vals[e];
            for (int n = 0; n < N; ++n)
                err += (db->dims[n] != dims[n]) + (db->mode_masks[n] != Xh->mode_masks[n]) +
                       (db->mode_pos[n] != Xh->mode_pos[n]) + (db->block_coords[n] != hb->block_coords[n]);
            err += db->nmode != N;
            total += db->nnz;
        }
        err += total != nnz;
        check("device_blocks", err, 0.0);
    }

    KruskalModel* M = nullptr;
    CreateKruskalModel(N, dims, R, &M);
    KruskalModelRandomInit(M, 7);
    KruskalModel* Md = make_device_copy(M);

    // ---- MTTKRP of every mode against a long-double COO MTTKRP
    std::vector<double> last_out;
    for (int n = 0; n < N; ++n) {
        const size_t cells = dims[n] * (size_t)R;
        double* out = dev_alloc(cells);
        (void)hipMemset(out, 0, cells * sizeof(double));
        mttkrp_alto_dev_onemode<unsigned long long>(Xh, Xd, M, out, 10, n, 0, false, 0, Md->U, Md->U_dev,
                                                    nullptr, nullptr);
        (void)hipDeviceSynchronize();
        std::vector<double> got = down(out, cells);
        std::vector<long double> want(cells, 0.0L);
        for (IType e = 0; e < nnz; ++e)
            for (int r = 0; r < R; ++r) {
                long double p = vals[e];
                for (int m = 0; m < N; ++m)
                    if (m != n) p *= M->U[m][cidx[e * N + m] * R + r];
                want[cidx[e * N + n] * R + r] += p;
            }
        std::vector<double> wd(want.begin(), want.end());
        char buf[64];
        snprintf(buf, sizeof buf, "mttkrp_mode%d", n);
        check(buf, rel_diff(got, wd), 1e-12);
        if (n == N - 1) last_out = got;
        (void)hipFree(out);
    }

    // ---- Grams (ata_gpu) of every factor
    std::vector<double*> grams(N);
    std::vector<std::vector<double>> gh(N);
    for (int n = 0; n < N; ++n) {
        grams[n] = dev_alloc((size_t)R * R);
        ata_gpu(nullptr, Md->U[n], grams[n], dims[n], R);
        (void)hipDeviceSynchronize();
        gh[n] = down(grams[n], (size_t)R * R);
        std::vector<double> want((size_t)R * R, 0.0);
        for (int a = 0; a < R; ++a)
            for (int b = 0; b < R; ++b) {
                long double t = 0.0L;
                for (IType i = 0; i < dims[n]; ++i) t += (long double)M->U[n][i * R + a] * M->U[n][i * R + b];
                want[a * R + b] = (double)t;
            }
        char buf[64];
        snprintf(buf, sizeof buf, "ata_gpu_mode%d", n);
        check(buf, rel_diff(gh[n], want), 1e-13);
    }

    // ---- kruskal_fit_gpu with lambda = (1 + r / R) and the last mode's MTTKRP
    {
        std::vector<double> lam(R);
        for (int r = 0; r < R; ++r) lam[r] = 1.0 + (double)r / R;
        up(Md->lambda, lam);
        double* out = dev_alloc(last_out.size());
        up(out, last_out);
        double* work = dev_alloc((size_t)R * R);
        double xn2 = 0.0;
        for (double v : vals) xn2 += v * v;
        const double fit = kruskal_fit_gpu(nullptr, grams.data(), Md, Xh, work, out, std::sqrt(xn2), N - 1, 0);
        long double ip = 0.0L, mn = 0.0L;
        for (IType i = 0; i < dims[N - 1]; ++i)
            for (int r = 0; r < R; ++r) ip += (long double)lam[r] * M->U[N - 1][i * R + r] * last_out[i * R + r];
        for (int a = 0; a < R; ++a)
            for (int b = 0; b < R; ++b) {
                long double t = (long double)lam[a] * lam[b];
                for (int n = 0; n < N; ++n) t *= gh[n][a * R + b];
                mn += t;
            }
        const double want = 1.0 - std::sqrt(std::fabs((double)(xn2 + std::fabs(mn) - 2.0L * ip))) / std::sqrt(xn2);
        check("kruskal_fit_gpu", std::fabs(fit - want) / std::max(1.0, std::fabs(want)), 1e-12);
        (void)hipFree(out);
        (void)hipFree(work);
    }

    // ---- columnwise_normscal on a copy of U[0] with a column of zeros
    {
        std::vector<double> u(M->U[0], M->U[0] + dims[0] * R);
        for (IType i = 0; i < dims[0]; ++i) u[i * R + 3] = 0.0;
        double* d = dev_alloc(u.size());
        up(d, u);
        std::vector<double> junk(R, 123.0);
        double* lam = dev_alloc(R);
        up(lam, junk);  // overwritten, not accumulated onto
        columnwise_normscal(nullptr, 0, d, R, dims[0], lam);
        (void)hipDeviceSynchronize();
        std::vector<double> got = down(d, u.size()), gl = down(lam, R);
        double err = 0.0;
        for (int r = 0; r < R; ++r) {
            long double s = 0.0L;
            for (IType i = 0; i < dims[0]; ++i) s += (long double)u[i * R + r] * u[i * R + r];
            const double norm = std::sqrt((double)s);
            err = std::max(err, std::fabs(gl[r] - norm) / (norm > 0 ? norm : 1.0));
            for (IType i = 0; i < dims[0]; ++i) {
                const double w = norm > 0 ? u[i * R + r] / norm : 0.0;
                err = std::max(err, std::fabs(got[i * R + r] - w));
            }
        }
        check("columnwise_normscal", err, 1e-14);
        (void)hipFree(d);
        (void)hipFree(lam);
    }

    // ---- pseudoinverse_gpu: a general 20 x 20 matrix and a rank-12 one
    {
        const int n = 20;
        std::vector<double> a((size_t)n * n);
        for (double& x : a) x = urand() - 0.5;
        check_pinv("general", a, n);
        std::vector<double> l((size_t)n * 12), r((size_t)12 * n), b((size_t)n * n, 0.0);
        for (double& x : l) x = urand() - 0.5;
        for (double& x : r) x = urand() - 0.5;
        for (int i = 0; i < n; ++i)
            for (int k = 0; k < 12; ++k)
                for (int j = 0; j < n; ++j) b[i * n + j] += l[i * 12 + k] * r[k * n + j];
        check_pinv("rank12", b, n);
        // a symmetric V of the ALS step
        std::vector<double> v((size_t)R * R);
        for (int e = 0; e < R * R; ++e) v[e] = gh[0][e] * gh[1][e];
        check_pinv("hadamard_of_grams", v, R);
    }

    // ---- value_fill / vector_hadamard / vecsum
    {
        const IType n = 1000;
        double* x = dev_alloc(n);
        double* y = dev_alloc(n);
        std::vector<double> yh(n);
        for (IType i = 0; i < n; ++i) yh[i] = (double)i;
        up(y, yh);
        value_fill(nullptr, x, n, 2.5);
        vector_hadamard(nullptr, x, y, n);
        vecsum(nullptr, x, n);
        (void)hipDeviceSynchronize();
        const double s = down(x, 1)[0];
        check("fill_hadamard_vecsum", std::fabs(s - 2.5 * 999.0 * 1000.0 / 2.0), 1e-9);
        (void)hipFree(x);
        (void)hipFree(y);
    }

    for (double* g : grams) (void)hipFree(g);
    destroy_kruskal_model_dev(Md);
    DestroyKruskalModel(M);
    delete_blcotensor_device(Xd);
    delete_blcotensor_host(Xh);
    DestroySparseTensor(X);
    printf("%s\n", g_fail ? "API CHECK FAILED" : "API CHECK OK");
    return g_fail;
}
