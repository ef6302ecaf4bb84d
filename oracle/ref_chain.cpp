// oracle/ref_chain.cpp -- a C wrapper around the reference's own plain-C++ GEMVER host code (TEST INFRASTRUCTURE ONLY).
// It is compiled by oracle/Makefile against examples/gemver-chain/common/chain_data.hpp where it lies in the reference
// checkout (the header is not copied into this repository); the output goes to oracle/_ref/. The tests use it to pin the
// oracle's input generator and FP64 GEMVER (x1, w1 of reference_chain, chain_data.hpp:93-144) to the reference's own code.
#include <cstdint>
#include <cstring>
#include <vector>

#include "chain_data.hpp"

extern "C" {

float ref_hash_unit(std::uint32_t seed, std::uint64_t index) { return gemver_chain::hash_unit(seed, index); }

void ref_make_vector(std::uint32_t seed, std::uint64_t n, float* out) {
    const std::vector<float> v = gemver_chain::make_vector(seed, n);
    std::memcpy(out, v.data(), n * sizeof(float));
}

void ref_make_matrix_A0(std::uint64_t n, float* out) {
    const std::vector<float> a = gemver_chain::make_matrix_A0(n);
    std::memcpy(out, a.data(), n * n * sizeof(float));
}

// The whole chain of the reference at alpha = beta = 1/sqrt(n); outputs x1, w1, q, s, x2, w2 (n doubles each).
void ref_chain(std::uint64_t n, const float* A0, const float* u1, const float* v1, const float* u2, const float* v2,
               const float* y, const float* z, double* x1, double* w1, double* q, double* s, double* x2, double* w2) {
    const std::vector<float> a(A0, A0 + n * n);
    gemver_chain::Vectors    g{std::vector<float>(u1, u1 + n), std::vector<float>(v1, v1 + n), std::vector<float>(u2, u2 + n),
                            std::vector<float>(v2, v2 + n), std::vector<float>(y, y + n),   std::vector<float>(z, z + n)};
    const gemver_chain::ChainResult r = gemver_chain::reference_chain(n, a, g);
    std::memcpy(x1, r.x1.data(), n * sizeof(double));
    std::memcpy(w1, r.w1.data(), n * sizeof(double));
    std::memcpy(q, r.q.data(), n * sizeof(double));
    std::memcpy(s, r.s.data(), n * sizeof(double));
    std::memcpy(x2, r.x2.data(), n * sizeof(double));
    std::memcpy(w2, r.w2.data(), n * sizeof(double));
}

double ref_relative_l2(const float* got, const double* want, std::uint64_t n) {
    return gemver_chain::relative_l2(std::vector<float>(got, got + n), std::vector<double>(want, want + n));
}

double ref_relative_l2_f64(const double* got, const double* want, std::uint64_t n) {
    return gemver_chain::relative_l2(std::vector<double>(got, got + n), std::vector<double>(want, want + n));
}

}  // extern "C"
