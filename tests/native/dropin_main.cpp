// A caller written against the reference's interface, used to check that RGF.hpp is a
// drop-in for RGF.H: it has the include set and the call sequence of the reference's
// driver (main.C:1-134) -- a raw `new RGF<T>(ia, ja, a, ns, nt, nd)`, get_time(0.0) /
// get_time(t) around factorize, solve(x, b, nrhs), logDet and RGFdiag, then the two
// residual norms and the solution -- and nothing from this library beyond RGF.hpp.
// Usage: dropin_main ns nt nd matrix.dat rhs.dat
#include <math.h>
#include <iostream>
#include <time.h>
#include <stdlib.h>
#include <stdio.h>
#include <limits>
#include "RGF.hpp"
using namespace std;

typedef double T;

int main(int argc, char* argv[])
{
    if (argc < 6) return 2;
    size_t ns = atoi(argv[1]), nt = atoi(argv[2]), nd = atoi(argv[3]);
    size_t n = ns * nt + nd;

    FILE* fm = fopen(argv[4], "r");
    if (!fm) return 1;
    size_t rows, cols, nnz;
    if (fscanf(fm, "%zu %zu %zu", &rows, &cols, &nnz) != 3 || rows != n) return 1;
    size_t* ia = new size_t[rows + 1];
    size_t* ja = new size_t[nnz];
    T* a = new T[nnz];
    for (size_t i = 0; i <= rows; i++) fscanf(fm, "%zu", &ia[i]);
    for (size_t i = 0; i < ia[rows]; i++) fscanf(fm, "%zu", &ja[i]);
    for (size_t i = 0; i < ia[rows]; i++) fscanf(fm, "%lf", &a[i]);
    fclose(fm);

    int nrhs = 1;
    T* b = new T[nrhs * n];
    T* x = new T[nrhs * n];
    T* diag = new T[n];
    FILE* fb = fopen(argv[5], "r");
    if (!fb) return 1;
    for (size_t i = 0; i < nrhs * n; i++) fscanf(fb, "%lf", &b[i]);
    fclose(fb);

    RGF<T>* solver = new RGF<T>(ia, ja, a, ns, nt, nd);
    double tf = get_time(0.0);
    solver->factorize();
    tf = get_time(tf);
    double ts = get_time(0.0);
    solver->solve(x, b, nrhs);
    ts = get_time(ts);
    double tl = get_time(0.0);
    double logdet = solver->logDet();
    tl = get_time(tl);
    double tr = get_time(0.0);
    solver->RGFdiag(diag);
    tr = get_time(tr);

    printf("logdet: %f\n\n", logdet);
    printf("factorize time: %lg\nsolve time: %lg\nlogdet time: %lg\nRGF time: %lg\n\n", tf, ts, tl, tr);
    printf("Residual norm: %e\n", solver->residualNorm(x, b));
    printf("Residual norm normalized: %e\n\n", solver->residualNormNormalized(x, b));
    for (size_t i = 0; i < nrhs * n; i++) printf("%32.24e\n", x[i]);
    for (size_t i = 0; i < n; i++) printf("diag %32.24e\n", diag[i]);

    delete solver;
    delete[] ia;
    delete[] ja;
    delete[] a;
    delete[] b;
    delete[] x;
    delete[] diag;
    return 0;
}
