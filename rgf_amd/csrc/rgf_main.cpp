// RGFSolver: command-line driver with the reference's interface and output
// (main.C:18-134): `RGFSolver ns nt nd matrix.dat rhs.dat`.  Reads the CSC file
// (main.C:50-70) and one right-hand side, times factorize / solve / logDet /
// RGFdiag, prints the same fields; logdet is printed with full precision in
// addition to the reference's %f.
#include <chrono>
#include <cstdio>
#include <cstdlib>
#include <ctime>
#include <vector>

#include "RGF.hpp"

static double now_s()
{
    return std::chrono::duration<double>(std::chrono::steady_clock::now().time_since_epoch()).count();
}

int main(int argc, char* argv[])
{
    if (argc < 6) {
        fprintf(stderr, "usage: %s ns nt nd matrix.dat rhs.dat\n", argv[0]);
        return 2;
    }
    time_t rawtime;
    time(&rawtime);
    printf("The current date/time is: %s\n", asctime(localtime(&rawtime)));
    size_t ns = strtoul(argv[1], nullptr, 10), nt = strtoul(argv[2], nullptr, 10), nd = strtoul(argv[3], nullptr, 10);
    size_t size = ns * nt + nd;
    FILE* F = fopen(argv[4], "r");
    if (!F) { fprintf(stderr, "cannot open %s\n", argv[4]); return 1; }
    size_t fn, fnnz;
    if (fscanf(F, "%zu", &fn) != 1 || fscanf(F, "%zu", &fn) != 1 || fscanf(F, "%zu", &fnnz) != 1) return 1;
    if (fn != size) { fprintf(stderr, "matrix is %zu x %zu, expected n = ns*nt+nd = %zu\n", fn, fn, size); return 1; }
    std::vector<size_t> ia(fn + 1), ja(fnnz);
    std::vector<double> a(fnnz);
    for (size_t i = 0; i <= fn; i++) if (fscanf(F, "%zu", &ia[i]) != 1) return 1;
    for (size_t i = 0; i < ia[fn]; i++) if (fscanf(F, "%zu", &ja[i]) != 1) return 1;
    for (size_t i = 0; i < ia[fn]; i++) if (fscanf(F, "%lf", &a[i]) != 1) return 1;
    fclose(F);
    const size_t nrhs = 1;
    std::vector<double> b(nrhs * size), x(nrhs * size), invDiag(size);
    F = fopen(argv[5], "r");
    if (!F) { fprintf(stderr, "cannot open %s\n", argv[5]); return 1; }
    for (size_t i = 0; i < nrhs * size; i++) if (fscanf(F, "%lf", &b[i]) != 1) return 1;
    fclose(F);
    try {
        RGF<double> solver(ia.data(), ja.data(), a.data(), ns, nt, nd);
        double t0 = now_s();
        solver.factorize();
        t0 = now_s() - t0;
        double t1 = now_s();
        solver.solve(x.data(), b.data(), nrhs);
        t1 = now_s() - t1;
        double t2 = now_s();
        double logdet = solver.logDet();
        t2 = now_s() - t2;
        double t3 = now_s();
        solver.RGFdiag(invDiag.data());
        t3 = now_s() - t3;
        printf("logdet: %f\n", logdet);
        printf("logdet (full precision): %.17g\n\n", logdet);
        printf("factorize time: %lg\n", t0);
        printf("solve time: %lg\n", t1);
        printf("logdet time: %lg\n", t2);
        printf("RGF time: %lg\n\n", t3);
        printf("Residual norm: %e\n", solver.residualNorm(x.data(), b.data()));
        printf("Residual norm normalized: %e\n\n", solver.residualNormNormalized(x.data(), b.data()));
        for (size_t i = 0; i < nrhs * size; i++) printf("%32.24e\n", x[i]);
    } catch (const rgf_error& e) {
        fprintf(stderr, "RGF error %d: %s\n", e.code(), e.what());
        return 1;
    }
    return 0;
}
