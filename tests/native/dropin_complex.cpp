// A complex caller of the reference's interface (main.C:10-16 with `typedef CPX T`): the
// matrix file holds interleaved real and imaginary parts of the CSC values, the right-hand
// side file interleaved pairs; prints logdet (both parts), the residual norms and x.
// Usage: dropin_complex ns nt nd matrix.dat rhs.dat
#include <stdio.h>
#include <stdlib.h>
#include "RGF.hpp"

typedef CPX T;

int main(int argc, char* argv[])
{
    if (argc < 6) return 2;
    size_t ns = atoi(argv[1]), nt = atoi(argv[2]), nd = atoi(argv[3]);
    size_t n = ns * nt + nd;
    FILE* f = fopen(argv[4], "r");
    if (!f) return 1;
    size_t rows, cols, nnz;
    if (fscanf(f, "%zu %zu %zu", &rows, &cols, &nnz) != 3 || rows != n) return 1;
    size_t* ia = new size_t[rows + 1];
    size_t* ja = new size_t[nnz];
    T* a = new T[nnz];
    for (size_t i = 0; i <= rows; i++) fscanf(f, "%zu", &ia[i]);
    for (size_t i = 0; i < nnz; i++) fscanf(f, "%zu", &ja[i]);
    for (size_t i = 0; i < nnz; i++) {
        double re, im;
        fscanf(f, "%lf %lf", &re, &im);
        a[i] = T(re, im);
    }
    fclose(f);
    T* b = new T[n];
    T* x = new T[n];
    T* diag = new T[n];
    f = fopen(argv[5], "r");
    if (!f) return 1;
    for (size_t i = 0; i < n; i++) {
        double re, im;
        fscanf(f, "%lf %lf", &re, &im);
        b[i] = T(re, im);
    }
    fclose(f);
    RGF<T>* solver = new RGF<T>(ia, ja, a, ns, nt, nd);
    double t = get_time(0.0);
    solver->factorize();
    solver->solve(x, b, 1);
    T logdet = solver->logDet();
    solver->RGFdiag(diag);
    t = get_time(t);
    printf("logdet: %.17g %.17g\n", logdet.real(), logdet.imag());
    printf("time: %lg\n", t);
    printf("Residual norm normalized: %e\n", solver->residualNormNormalized(x, b));
    for (size_t i = 0; i < n; i++) printf("x %.17g %.17g\n", x[i].real(), x[i].imag());
    for (size_t i = 0; i < n; i++) printf("diag %.17g %.17g\n", diag[i].real(), diag[i].imag());
    delete solver;
    delete[] ia;
    delete[] ja;
    delete[] a;
    delete[] b;
    delete[] x;
    delete[] diag;
    return 0;
}
