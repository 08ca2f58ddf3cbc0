// Test harness: the generated ERI class code and eval_quartet() compiled for the HOST, so
// the CPU test suite can check every angular class against the oracle without a GPU.
// Built by hartreefock_amd/build.py into tests/cpu_kernel/libquartet_host.so; the product
// library never links it.
#include "generated/host_eval.inc"
#include "hf_internal.h"

// sA, sC: shell indices of the first shell of bra and ket pair; a0..d0: first basis function of each
// shell (used only by classes whose reference convention depends on the bra choice).
extern "C" int hfq_eval(int la, int lb, int lc, int ld, int exact, const double* A, const double* B,
                        const double* C, const double* D, const hfgpu::PrimPair* bra, int npb,
                        const hfgpu::PrimPair* ket, int npk, const double* boys, double* out, int sA, int sC,
                        int a0, int b0, int c0, int d0) {
    return hfgpu::eval_quartet_host(la, lb, lc, ld, exact, A, B, C, D, bra, npb, ket, npk, boys, out, sA, sC, a0, b0,
                                    c0, d0);
}

// The FP32 evaluation of the mixed-precision tier (classes with L < 6); boys32 is the FP32 table.
extern "C" int hfq_eval_f32(int la, int lb, int lc, int ld, const double* A, const double* B, const double* C,
                            const double* D, const hfgpu::PrimPair* bra, int npb, const hfgpu::PrimPair* ket, int npk,
                            const float* boys32, float* out) {
    return hfgpu::eval_quartet_host_f32(la, lb, lc, ld, A, B, C, D, bra, npb, ket, npk, boys32, out);
}

extern "C" void hfq_boys_table(double* out) {
    std::vector<double> t;
    hfhost::boys_table(t);
    for (size_t i = 0; i < t.size(); ++i) out[i] = t[i];
}

extern "C" int hfq_boys_table_size() { return hfgpu::BOYS_NROW * hfgpu::BOYS_STRIDE; }

// The kernels' Boys evaluation, boys_scaled<L> (FP64) and boys_scaled_f32<L> (FP32), on the host with the
// device table: F_0..F_L(T) with pref = 1.  L <= 8 (the largest total angular momentum of an L <= 2 quartet).
extern "C" int hfq_boys(int L, double T, const double* tab, double* F) {
    switch (L) {
#define HFQ_B(LL) case LL: hfgpu::boys_scaled<LL>(T, 1.0, tab, F); return 0;
        HFQ_B(0) HFQ_B(1) HFQ_B(2) HFQ_B(3) HFQ_B(4) HFQ_B(5) HFQ_B(6) HFQ_B(7) HFQ_B(8)
#undef HFQ_B
    }
    return -1;
}

extern "C" int hfq_boys_f32(int L, float T, const float* tab32, float* F) {
    switch (L) {
#define HFQ_B(LL) case LL: hfgpu::boys_scaled_f32<LL>(T, 1.0f, tab32, F); return 0;
        HFQ_B(0) HFQ_B(1) HFQ_B(2) HFQ_B(3) HFQ_B(4) HFQ_B(5) HFQ_B(6) HFQ_B(7) HFQ_B(8)
#undef HFQ_B
    }
    return -1;
}
