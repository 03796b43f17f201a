// ref_verify_shim.cpp -- C entry points over the reference's own host checker.
//
// TEST INFRASTRUCTURE ONLY: compiled by oracle/Makefile (target `ref`) into
// oracle/_ref/libref_verify.so, and only when the reference sources are present.
// It includes the reference's header-only host code where it lies
// (REF_ROOT/parallel_pivot/verify.hpp and REF_ROOT/templated/verify.hpp, passed
// on the compiler line) and exposes its functions with plain C signatures so the
// tests can pin the oracle against them:
//
//   ref_pivotedA                 pivotedA              parallel_pivot/verify.hpp:106-155
//   ref_verifyInv                verifyInv             parallel_pivot/verify.hpp:50-103
//   ref_verifyLUwithPivoting     verifyLUwithPivoting  parallel_pivot/verify.hpp:157-242
//   ref_calc_cond_num            calc_cond_num         parallel_pivot/verify.hpp:245-338
//
// verifyInv / verifyLUwithPivoting report by printing to std::cout; the shim
// captures that text and returns the printed counts.
#include <iostream>
#include <sstream>
#include <string>
#include <vector>

#include PP_VERIFY_HPP

namespace {
long long parse_after(const std::string& text, const std::string& label) {
    size_t p = text.find(label);
    if (p == std::string::npos) return -1;
    return std::stoll(text.substr(p + label.size()));
}

struct CaptureCout {
    std::ostringstream buf;
    std::streambuf* old;
    CaptureCout() : old(std::cout.rdbuf(buf.rdbuf())) {}
    ~CaptureCout() { std::cout.rdbuf(old); }
};
}  // namespace

extern "C" {

void ref_pivotedA(const float* A, int n, int* pivots_out, float* pivoted_out) {
    std::vector<float> a(A, A + (size_t)n * n), pa((size_t)n * n);
    std::vector<int> piv(n);
    pivotedA(a, pa, piv, n);
    for (int i = 0; i < n; ++i) pivots_out[i] = piv[i];
    if (pivoted_out)
        for (size_t i = 0; i < pa.size(); ++i) pivoted_out[i] = pa[i];
}

// Returns the "Correct inversions" count printed by verifyInv; *wrong gets the
// "Incorrect inversions" count.
long long ref_verifyInv(const float* A, const float* X, int n, int batch, long long* wrong) {
    std::vector<float> a(A, A + (size_t)n * n * batch), x(X, X + (size_t)n * n * batch);
    std::string text;
    {
        CaptureCout cap;
        verifyInv(a, x, n, batch);
        text = cap.buf.str();
    }
    if (wrong) *wrong = parse_after(text, "Incorrect inversions: ");
    return parse_after(text, "Correct inversions: ");
}

// Returns the "Correct LU decompositions" count printed by verifyLUwithPivoting.
long long ref_verifyLUwithPivoting(const float* PA, const float* LU, int n, int batch) {
    std::vector<float> pa(PA, PA + (size_t)n * n), lu(LU, LU + (size_t)n * n * batch);
    std::string text;
    {
        CaptureCout cap;
        verifyLUwithPivoting(pa, lu, n, batch);
        text = cap.buf.str();
    }
    return parse_after(text, "Correct LU decompositions: ");
}

float ref_calc_cond_num(const float* A, int n) {
    std::vector<float> a(A, A + (size_t)n * n);
    return calc_cond_num(a, n);
}

}  // extern "C"
