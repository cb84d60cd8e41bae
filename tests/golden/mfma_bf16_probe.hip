// Generator of tests/golden/mfma_bf16_cases.npz (run on an MI355X, then
// tests/golden/make_mfma_bf16_fixture.py): random BF16 operands and FP32
// accumulators through v_mfma_f32_32x32x16_bf16, one 32x32x16 product per
// batch, written as raw operands and results.
//   hipcc --offload-arch=gfx950 -O2 -o mfma_bf16_probe mfma_bf16_probe.hip
//   ./mfma_bf16_probe mfma_bf16_cases.bin
// Batch types (t % 5): 0 dense, operands 2^-12..2^2; 1 dense, 2^-40..2^8;
// 2 only k = 0, 1 nonzero; 3 only k = 0, 8 nonzero; 4 cancelling pairs.
#include <hip/hip_runtime.h>
#include <cmath>
#include <cstdio>
#include <random>
#include <vector>
typedef float v16f __attribute__((ext_vector_type(16)));
typedef __bf16 bf16x8 __attribute__((ext_vector_type(8)));
typedef unsigned short u16x8 __attribute__((ext_vector_type(8)));

// A [32][16] (row m, k), B [16][32] (k, col n), C/D [32][32] per batch; lane
// l supplies row/col l % 32 and k = 8 (l / 32) .. +7; output register v of
// lane l is row 8 (v / 4) + 4 (l / 32) + v % 4, column l % 32.
__global__ void k(const unsigned short *A, const unsigned short *B, const float *C, float *D, int nb)
{
    const int l = threadIdx.x;
    for (int t = 0; t < nb; ++t) {
        const unsigned short *a = A + t * 512, *b = B + t * 512;
        u16x8 av, bv;
        for (int j = 0; j < 8; ++j) {
            av[j] = a[(l % 32) * 16 + 8 * (l / 32) + j];
            bv[j] = b[(8 * (l / 32) + j) * 32 + (l % 32)];
        }
        v16f acc;
        for (int v = 0; v < 16; ++v) acc[v] = C[t * 1024 + (8 * (v / 4) + 4 * (l / 32) + (v % 4)) * 32 + (l % 32)];
        acc = __builtin_amdgcn_mfma_f32_32x32x16_bf16(__builtin_bit_cast(bf16x8, av), __builtin_bit_cast(bf16x8, bv),
                                                      acc, 0, 0, 0);
        for (int v = 0; v < 16; ++v) D[t * 1024 + (8 * (v / 4) + 4 * (l / 32) + (v % 4)) * 32 + (l % 32)] = acc[v];
    }
}

int main(int argc, char **argv)
{
    const int nb = 80;
    std::mt19937 g(7);
    std::vector<unsigned short> A(nb * 512), B(nb * 512);
    std::vector<float> C(nb * 1024), D(nb * 1024);
    auto rbf = [&](int emin, int emax) -> unsigned short {
        std::uniform_int_distribution<int> e(emin, emax), m(0, 127), s(0, 1);
        if ((g() & 15) == 0) return 0;
        return (unsigned short)((s(g) << 15) | ((e(g) + 127) << 7) | m(g));
    };
    for (int t = 0; t < nb; ++t) {
        const int type = t % 5;
        const int lo = type == 1 ? -40 : -12, hi = type == 1 ? 8 : 2;
        for (int q = 0; q < 512; ++q) {
            A[t * 512 + q] = rbf(lo, hi);
            B[t * 512 + q] = rbf(-2, 1);
        }
        std::uniform_real_distribution<float> u(-1.0f, 1.0f);
        std::uniform_int_distribution<int> ce(-30, 12);
        for (int q = 0; q < 1024; ++q) C[t * 1024 + q] = (g() & 7) == 0 ? 0.0f : std::ldexp(u(g), ce(g));
        if (type == 2 || type == 3)
            for (int m = 0; m < 32; ++m)
                for (int kk = 0; kk < 16; ++kk) {
                    const bool keep = type == 2 ? (kk == 0 || kk == 1) : (kk == 0 || kk == 8);
                    if (!keep) A[t * 512 + m * 16 + kk] = 0;
                }
        if (type == 4) {
            for (int m = 0; m < 32; ++m)
                for (int kk = 0; kk < 16; kk += 2) A[t * 512 + m * 16 + kk + 1] = A[t * 512 + m * 16 + kk] ^ 0x8000;
            for (int kk = 0; kk < 16; kk += 2)
                for (int n = 0; n < 32; ++n) B[t * 512 + (kk + 1) * 32 + n] = B[t * 512 + kk * 32 + n];
        }
    }
    unsigned short *dA, *dB;
    float *dC, *dD;
    if (hipMalloc(&dA, A.size() * 2) || hipMalloc(&dB, B.size() * 2) || hipMalloc(&dC, C.size() * 4) ||
        hipMalloc(&dD, D.size() * 4))
        return 1;
    if (hipMemcpy(dA, A.data(), A.size() * 2, hipMemcpyHostToDevice) ||
        hipMemcpy(dB, B.data(), B.size() * 2, hipMemcpyHostToDevice) ||
        hipMemcpy(dC, C.data(), C.size() * 4, hipMemcpyHostToDevice))
        return 1;
    hipLaunchKernelGGL(k, dim3(1), dim3(64), 0, 0, dA, dB, dC, dD, nb);
    if (hipMemcpy(D.data(), dD, D.size() * 4, hipMemcpyDeviceToHost)) return 1;
    FILE *f = fopen(argc > 1 ? argv[1] : "mfma_bf16_cases.bin", "wb");
    if (!f) return 1;
    fwrite(&nb, 4, 1, f);
    fwrite(A.data(), 2, A.size(), f);
    fwrite(B.data(), 2, B.size(), f);
    fwrite(C.data(), 4, C.size(), f);
    fwrite(D.data(), 4, D.size(), f);
    fclose(f);
    printf("%d batches written\n", nb);
    return 0;
}
