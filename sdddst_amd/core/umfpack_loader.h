// Runtime binding of SuiteSparse UMFPACK (the reference links it at build
// time; this build loads libumfpack.so when present so the host hand-off
// works without a compile-time dependency).  Only the five calls of
// simulation.cpp:83, 388-395 are bound.
#pragma once

#include <string>

namespace sdddstCore {

class UmfpackLibrary {
public:
    // Tries $SDDDST_UMFPACK_LIB, then libumfpack.so, libumfpack.so.6, .5.
    static UmfpackLibrary *load(std::string *why);
    ~UmfpackLibrary();

    int symbolic(int n_row, int n_col, const int *Ap, const int *Ai, const double *Ax, void **Symbolic);
    int numeric(const int *Ap, const int *Ai, const double *Ax, void *Symbolic, void **Numeric);
    int solve(const int *Ap, const int *Ai, const double *Ax, double *X, const double *B, void *Numeric);
    void free_symbolic(void **Symbolic);
    void free_numeric(void **Numeric);
    const std::string &path() const { return where; }

private:
    UmfpackLibrary() = default;
    void *handle = nullptr;
    std::string where;
    int (*p_symbolic)(int, int, const int *, const int *, const double *, void **, const double *, double *) = nullptr;
    int (*p_numeric)(const int *, const int *, const double *, void *, void **, const double *, double *) = nullptr;
    int (*p_solve)(int, const int *, const int *, const double *, double *, const double *, void *, const double *,
                   double *) = nullptr;
    void (*p_free_symbolic)(void **) = nullptr;
    void (*p_free_numeric)(void **) = nullptr;
};

}  // namespace sdddstCore
