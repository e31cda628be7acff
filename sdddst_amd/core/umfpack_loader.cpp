#include "umfpack_loader.h"

#include <dlfcn.h>
#include <stdlib.h>

namespace sdddstCore {

UmfpackLibrary *UmfpackLibrary::load(std::string *why)
{
    const char *env = getenv("SDDDST_UMFPACK_LIB");
    const char *names[] = {env, "libumfpack.so", "libumfpack.so.6", "libumfpack.so.5"};
    for (const char *name : names) {
        if (!name || !*name) continue;
        void *h = dlopen(name, RTLD_NOW | RTLD_LOCAL);
        if (!h) continue;
        UmfpackLibrary *u = new UmfpackLibrary();
        u->handle = h;
        u->where = name;
        u->p_symbolic = (decltype(u->p_symbolic))dlsym(h, "umfpack_di_symbolic");
        u->p_numeric = (decltype(u->p_numeric))dlsym(h, "umfpack_di_numeric");
        u->p_solve = (decltype(u->p_solve))dlsym(h, "umfpack_di_solve");
        u->p_free_symbolic = (decltype(u->p_free_symbolic))dlsym(h, "umfpack_di_free_symbolic");
        u->p_free_numeric = (decltype(u->p_free_numeric))dlsym(h, "umfpack_di_free_numeric");
        if (u->p_symbolic && u->p_numeric && u->p_solve && u->p_free_symbolic && u->p_free_numeric) return u;
        delete u;
    }
    if (why) *why = "libumfpack.so not found (set SDDDST_UMFPACK_LIB to its path)";
    return nullptr;
}

UmfpackLibrary::~UmfpackLibrary()
{
    if (handle) dlclose(handle);
}

int UmfpackLibrary::symbolic(int n_row, int n_col, const int *Ap, const int *Ai, const double *Ax, void **Symbolic)
{
    return p_symbolic(n_row, n_col, Ap, Ai, Ax, Symbolic, nullptr, nullptr);
}

int UmfpackLibrary::numeric(const int *Ap, const int *Ai, const double *Ax, void *Symbolic, void **Numeric)
{
    return p_numeric(Ap, Ai, Ax, Symbolic, Numeric, nullptr, nullptr);
}

int UmfpackLibrary::solve(const int *Ap, const int *Ai, const double *Ax, double *X, const double *B, void *Numeric)
{
    const int UMFPACK_A = 0;   // solve A x = b
    return p_solve(UMFPACK_A, Ap, Ai, Ax, X, B, Numeric, nullptr, nullptr);
}

void UmfpackLibrary::free_symbolic(void **Symbolic) { p_free_symbolic(Symbolic); }
void UmfpackLibrary::free_numeric(void **Numeric) { p_free_numeric(Numeric); }

}  // namespace sdddstCore
