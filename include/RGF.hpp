// RGF.hpp -- the reference's class interface over the C ABI of librgf_hip.so.
//
// rjanalik/RGF exposes `template<class T> class RGF` (RGF.H:11-80).  This class
// keeps those method names and signatures for T = double so an existing caller
// (e.g. main.C:18-134) compiles with `#include "RGF.hpp"` in place of
// `#include "RGF.H"`; each method forwards to the C entry point that replaces it
// and throws rgf_error on a non-OK status (the reference ignores LAPACK info and
// exits on allocation failure).  The reference's header also brings in the wall
// clock `get_time` (RGF.H:8 -> Utilities.H:16) that its driver times each phase
// with; it is provided here with the same meaning.
#pragma once

#include <sys/time.h>

#include <complex>
#include <cstddef>
#include <stdexcept>
#include <string>

#include "rgf.h"

// Utilities.H:16 / Utilities.C:26-32: wall-clock seconds since the epoch minus t0
// (t = get_time(0.0); ...; t = get_time(t) gives the elapsed seconds).
inline double get_time(double t0)
{
    timeval tv;
    gettimeofday(&tv, nullptr);
    return (double)tv.tv_sec + (double)tv.tv_usec * 1e-6 - t0;
}

// Types.H:41-45: the reference's complex scalar
typedef std::complex<double> CPX;

class rgf_error : public std::runtime_error {
public:
    rgf_error(int code, const std::string& what) : std::runtime_error(what), code_(code) {}
    int code() const { return code_; }
private:
    int code_;
};

template <class T>
class RGF;

template <>
class RGF<double> {
public:
    // RGF.H:16 / RGF.H:85-146
    RGF(size_t* ia, size_t* ja, double* a, size_t ns, size_t nt, size_t nd,
        const rgf_options* opts = nullptr)
    {
        check(rgf_create_csc(ia, ja, a, ns, nt, nd, opts, &h_), nullptr);
        n_ = ns * nt + nd;
    }
    ~RGF() { if (h_) rgf_destroy(h_); }                                     // RGF.H:151-170
    RGF(const RGF&) = delete;
    RGF& operator=(const RGF&) = delete;

    void factorize() { check(rgf_factorize(h_), h_); }                     // RGF.H:510-546
    void solve(double* b, size_t nrhs) { check(rgf_solve(h_, b, b, nrhs), h_); }   // RGF.H:550-554
    void solve(double* x, double* b, size_t nrhs) { check(rgf_solve(h_, x, b, nrhs), h_); }  // RGF.H:558-578
    void RGFdiag(double* diag) { check(rgf_selinv_diag(h_, diag), h_); }   // RGF.H:582-612
    double logDet()                                                         // RGF.H:616-627
    {
        double v = 0.0;
        check(rgf_logdet(h_, &v), h_);
        return v;
    }
    double residualNorm(double* x, double* b)                               // RGF.H:631-655
    {
        double v = 0.0;
        check(rgf_residual_norm(h_, x, b, &v), h_);
        return v;
    }
    double residualNormNormalized(double* x, double* b)                     // RGF.H:659-663
    {
        double v = 0.0;
        check(rgf_residual_norm_normalized(h_, x, b, &v), h_);
        return v;
    }

    rgf_handle_t handle() const { return h_; }

private:
    static void check(int rc, rgf_handle_t h)
    {
        if (rc != RGF_OK) throw rgf_error(rc, rgf_last_error(h));
    }
    rgf_handle_t h_ = nullptr;
    size_t n_ = 0;
};

// RGF<CPX>, RGF.H:756-780: the same interface for complex matrices, in the complex-symmetric
// algebra of the reference's block updates (A = A^T = L L^T; logDet returns the complex
// 2 sum log L_jj).  std::complex<double> arrays are passed to the C ABI as interleaved
// (re, im) doubles, the layout the C++ standard gives them.
template <>
class RGF<CPX> {
public:
    RGF(size_t* ia, size_t* ja, CPX* a, size_t ns, size_t nt, size_t nd, const rgf_options* opts = nullptr)
    {
        check(rgf_z_create_csc(ia, ja, reinterpret_cast<const double*>(a), ns, nt, nd, opts, &h_), nullptr);
    }
    ~RGF() { if (h_) rgf_z_destroy(h_); }
    RGF(const RGF&) = delete;
    RGF& operator=(const RGF&) = delete;

    void factorize() { check(rgf_z_factorize(h_), h_); }
    void solve(CPX* b, size_t nrhs) { check(rgf_z_solve(h_, dp(b), dp(b), nrhs), h_); }
    void solve(CPX* x, CPX* b, size_t nrhs) { check(rgf_z_solve(h_, dp(x), dp(b), nrhs), h_); }
    void RGFdiag(CPX* diag) { check(rgf_z_selinv_diag(h_, dp(diag)), h_); }
    CPX logDet()
    {
        double v[2] = {0.0, 0.0};
        check(rgf_z_logdet(h_, v), h_);
        return CPX(v[0], v[1]);
    }
    double residualNorm(CPX* x, CPX* b)
    {
        double v = 0.0;
        check(rgf_z_residual_norm(h_, dp(x), dp(b), &v), h_);
        return v;
    }
    double residualNormNormalized(CPX* x, CPX* b)
    {
        double v = 0.0;
        check(rgf_z_residual_norm_normalized(h_, dp(x), dp(b), &v), h_);
        return v;
    }

    rgf_zhandle_t handle() const { return h_; }

private:
    static double* dp(CPX* p) { return reinterpret_cast<double*>(p); }
    static void check(int rc, rgf_zhandle_t h)
    {
        if (rc != RGF_OK) throw rgf_error(rc, rgf_z_last_error(h));
    }
    rgf_zhandle_t h_ = nullptr;
};
