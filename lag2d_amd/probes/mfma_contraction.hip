// Measurement probe (not part of the library): the per-cell basis x DOF
// contraction of the curved P2 domain integral (rhsdomndg_lagmc_curvquad,
// Getrhslag.f90:4866-4874: 25 Gauss-Lobatto points, 6 DOFs, 4 variables)
// on the FP64 vector ALUs, as the library computes it, and on the FP64
// matrix cores (v_mfma_f64_16x16x4f64).  Prints both times and the largest
// difference of the per-cell sums; first checks the MFMA output layout.
//   hipcc -O3 -std=c++17 -ffp-contract=off --offload-arch=gfx950 mfma_contraction.hip -o mfma_contraction
//   ./mfma_contraction [ncell]
#include <hip/hip_runtime.h>
#include <cstdio>
#include <cstdlib>
#include <cmath>
#include <vector>

#define CHK(x) do { hipError_t e = (x); if (e != hipSuccess) { printf("HIP %s line %d\n", hipGetErrorString(e), __LINE__); exit(1);} } while (0)

__constant__ double PL5[5] = {-1.0, -0.6546536707079771437983, 0.0, 0.6546536707079771437983, 1.0};

typedef double v4d __attribute__((ext_vector_type(4)));

// 1. lane layout check: D = A(16x4) B(4x16), A[i][k] = i + 100 k, B[k][j] = 1000 k + j
__global__ void k_layout(double *out)
{
    const int l = threadIdx.x;
    const double a = (double)(l % 16) + 100.0 * (l / 16);
    const double b = 1000.0 * (l / 16) + (double)(l % 16);
    v4d c = {0.0, 0.0, 0.0, 0.0};
    c = __builtin_amdgcn_mfma_f64_16x16x4f64(a, b, c, 0, 0, 0);
    for (int j = 0; j < 4; ++j) out[l * 4 + j] = c[j];
}

// 2. VALU: thread per cell, basis about the cell's mass centre at each point,
//    un[q] = sum_id u[q][id] b[id] (the cellrhs form); returns sum over points, q
__global__ void __launch_bounds__(256) k_valu(int nc, const double *__restrict__ U, const double *__restrict__ G,
                                              double *__restrict__ out)
{
    const int ie = blockIdx.x * blockDim.x + threadIdx.x;
    if (ie >= nc) return;
    double u[4][6];
    for (int q = 0; q < 4; ++q)
        for (int id = 0; id < 6; ++id) u[q][id] = U[(q * 6 + id) * (size_t)nc + ie];
    const double rc = G[ie], sc = G[nc + ie], g0 = G[2 * nc + ie], g1 = G[3 * nc + ie], g2 = G[4 * nc + ie];
    double acc = 0.0;
    for (int k = 0; k < 25; ++k) {
        const double r = PL5[k / 5], s = PL5[k % 5];
        double b[6];
        b[0] = 1.0; b[1] = r - rc; b[2] = s - sc;
        b[3] = 0.5 * b[1] * b[1] - g0; b[4] = 0.5 * b[2] * b[2] - g1; b[5] = b[1] * b[2] - g2;
        for (int q = 0; q < 4; ++q) {
            double v = 0.0;
            for (int id = 0; id < 6; ++id) v = v + u[q][id] * b[id];
            acc = acc + v;
        }
    }
    out[ie] = acc;
}

// 3. MFMA, the basis about (rc, sc) re-expanded in monomials of (r, s):
//    un_q(r, s) = sum_m c_qm p_m(r, s), p = (1, r, s, r^2/2, s^2/2, r s),
//    c = T(rc, sc, g) u_q per cell and variable (the points' monomials are
//    then the same for every cell).  A wave per 16-cell tile: per variable q, D(16 cells x 32 points)
//    = C_q(16 x 8) P(8 x 32); lane l supplies A[l%16][l/16 (+4)] and
//    B[l/16 (+4)][16 nt + l%16] and holds D[l/16 + 4 j][16 nt + l%16].
__global__ void __launch_bounds__(256) k_mfma2(int nc, const double *__restrict__ U, const double *__restrict__ G,
                                               double *__restrict__ out)
{
    const int lane = threadIdx.x & 63, tile = (blockIdx.x * blockDim.x + threadIdx.x) >> 6;
    const int row = lane % 16, kk = lane / 16;
    const int cell = 16 * tile + row;
    double rc = 0.0, sc = 0.0, g0 = 0.0, g1 = 0.0, g2 = 0.0;
    const bool ok = cell < nc;
    if (ok) {
        rc = G[cell]; sc = G[nc + cell]; g0 = G[2 * nc + cell]; g1 = G[3 * nc + cell]; g2 = G[4 * nc + cell];
    }
    // the monomials at the points this lane supplies (constant)
    double plo[2], phi[2];
    for (int nt = 0; nt < 2; ++nt) {
        const int pt = 16 * nt + row;
        double p[8] = {0, 0, 0, 0, 0, 0, 0, 0};
        if (pt < 25) {
            const double r = PL5[pt / 5], s = PL5[pt % 5];
            p[0] = 1.0; p[1] = r; p[2] = s; p[3] = 0.5 * r * r; p[4] = 0.5 * s * s; p[5] = r * s;
        }
        plo[nt] = p[kk];
        phi[nt] = p[4 + kk];
    }
    double acc[4] = {0.0, 0.0, 0.0, 0.0};
    for (int q = 0; q < 4; ++q) {
        double u[6] = {0, 0, 0, 0, 0, 0};
        if (ok)
            for (int id = 0; id < 6; ++id) u[id] = U[(q * 6 + id) * (size_t)nc + cell];
        double clo, chi;
        if (kk == 0) {
            clo = u[0] - rc * u[1] - sc * u[2] + (0.5 * rc * rc - g0) * u[3] + (0.5 * sc * sc - g1) * u[4] +
                  (rc * sc - g2) * u[5];
            chi = u[4];
        } else if (kk == 1) {
            clo = u[1] - rc * u[3] - sc * u[5];
            chi = u[5];
        } else if (kk == 2) {
            clo = u[2] - sc * u[4] - rc * u[5];
            chi = 0.0;
        } else {
            clo = u[3];
            chi = 0.0;
        }
        for (int nt = 0; nt < 2; ++nt) {
            v4d d = {0.0, 0.0, 0.0, 0.0};
            d = __builtin_amdgcn_mfma_f64_16x16x4f64(clo, plo[nt], d, 0, 0, 0);
            d = __builtin_amdgcn_mfma_f64_16x16x4f64(chi, phi[nt], d, 0, 0, 0);
            for (int j = 0; j < 4; ++j) acc[j] += d[j];
        }
    }
    // rows l/16 + 4 j: sum the 16 point columns of each lane group
    for (int j = 0; j < 4; ++j)
        for (int o = 8; o >= 1; o >>= 1) acc[j] += __shfl_xor(acc[j], o, 16);
    if (row == 0)
        for (int j = 0; j < 4; ++j) {
            const int oc = 16 * tile + kk + 4 * j;
            if (oc < nc) out[oc] = acc[j];
        }
}

int main(int argc, char **argv)
{
    const int nc = argc > 1 ? atoi(argv[1]) : 1440000;
    // layout
    double *dl;
    CHK(hipMalloc(&dl, 256 * sizeof(double)));
    hipLaunchKernelGGL(k_layout, dim3(1), dim3(64), 0, 0, dl);
    std::vector<double> hl(256);
    CHK(hipMemcpy(hl.data(), dl, 256 * sizeof(double), hipMemcpyDeviceToHost));
    int bad = 0;
    for (int l = 0; l < 64; ++l)
        for (int j = 0; j < 4; ++j) {
            const int i = l / 16 + 4 * j, col = l % 16;  // lane l, register j holds D[l/16 + 4j][l%16]
            double e = 0.0;
            for (int k = 0; k < 4; ++k) e += (i + 100.0 * k) * (1000.0 * k + col);
            if (hl[l * 4 + j] != e) ++bad;
        }
    printf("MFMA output layout D[lane/16 + 4 reg][lane%%16]: %s\n", bad ? "MISMATCH" : "as assumed");
    if (bad) return 1;
    std::vector<double> hU((size_t)24 * nc), hG((size_t)5 * nc);
    srand(7);
    for (auto &v : hU) v = (double)rand() / RAND_MAX - 0.5;
    for (int ie = 0; ie < nc; ++ie) {
        hG[ie] = 0.1 * ((double)rand() / RAND_MAX - 0.5);
        hG[nc + ie] = 0.1 * ((double)rand() / RAND_MAX - 0.5);
        for (int k = 2; k < 5; ++k) hG[(size_t)k * nc + ie] = 0.2 * (double)rand() / RAND_MAX;
    }
    double *dU, *dG, *o1, *o2;
    CHK(hipMalloc(&dU, hU.size() * 8)); CHK(hipMalloc(&dG, hG.size() * 8));
    CHK(hipMalloc(&o1, (size_t)nc * 8)); CHK(hipMalloc(&o2, (size_t)nc * 8));
    CHK(hipMemcpy(dU, hU.data(), hU.size() * 8, hipMemcpyHostToDevice));
    CHK(hipMemcpy(dG, hG.data(), hG.size() * 8, hipMemcpyHostToDevice));
    const int nw = (nc + 15) / 16;              // waves (16-cell tiles) of k_mfma2
    const int gm = (nw * 64 + 255) / 256, gv = (nc + 255) / 256;
    hipEvent_t a, b;
    CHK(hipEventCreate(&a)); CHK(hipEventCreate(&b));
    float tv = 1e30f, tm = 1e30f;
    for (int rep = 0; rep < 6; ++rep) {
        float t;
        CHK(hipEventRecord(a));
        hipLaunchKernelGGL(k_valu, dim3(gv), dim3(256), 0, 0, nc, dU, dG, o1);
        CHK(hipEventRecord(b)); CHK(hipEventSynchronize(b)); CHK(hipEventElapsedTime(&t, a, b));
        if (rep) tv = t < tv ? t : tv;
        CHK(hipEventRecord(a));
        hipLaunchKernelGGL(k_mfma2, dim3(gm), dim3(256), 0, 0, nc, dU, dG, o2);
        CHK(hipEventRecord(b)); CHK(hipEventSynchronize(b)); CHK(hipEventElapsedTime(&t, a, b));
        if (rep) tm = t < tm ? t : tm;
    }
    std::vector<double> h1(nc), h2(nc);
    CHK(hipMemcpy(h1.data(), o1, (size_t)nc * 8, hipMemcpyDeviceToHost));
    CHK(hipMemcpy(h2.data(), o2, (size_t)nc * 8, hipMemcpyDeviceToHost));
    double md = 0.0, mx = 0.0;
    for (int i = 0; i < nc; ++i) { md = fmax(md, fabs(h1[i] - h2[i])); mx = fmax(mx, fabs(h1[i])); }
    const double flops = (double)nc * 25 * 4 * 6 * 2;
    printf("cells %d: VALU %.1f us (%.2f TFLOP/s of contraction), MFMA %.1f us (%.2f TFLOP/s); max |diff| %.2e of max %.2e\n",
           nc, 1e3 * tv, flops / (1e-3 * tv) / 1e12, 1e3 * tm, flops / (1e-3 * tm) / 1e12, md, mx);
    return 0;
}
