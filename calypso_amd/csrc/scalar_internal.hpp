// scalar_internal.hpp -- scalars diffusing in a shell at rest (host side of
// scalar_shell.hip).  Not part of the C ABI.
#pragma once

#include <hip/hip_runtime.h>

#include <string>
#include <vector>

namespace calypso {

constexpr int kScalarMax = 8;          // scalars per context (one batched solve)
constexpr int kBcFixed = 0, kBcFlux = 1;

struct ScalarSpec {
    double coef_d = 1.0, coef_imp = 0.5;
    int bc_icb = kBcFixed, bc_cmb = kBcFixed;
};

struct ScalarShell {
    int L = 0, J = 0, nr = 0, nsc = 0, i_step = 0;
    double dt = 0.0, time0 = 0.0;
    hipStream_t stream = nullptr;
    std::vector<double> radius;
    std::vector<ScalarSpec> spec;
    std::vector<std::string> names;
    // device: per scalar x, diffusion, wall advection, Adams history
    // (nr J each, mode j = l(l+1)+m fastest); wall values [J] (the inner
    // flux negated); 3-band factors in the solver's mode-fastest layout
    double *d_fields = nullptr;
    double *d_walls = nullptr;
    double *d_lu[kScalarMax] = {};
    int *d_piv[kScalarMax] = {};
    double *d_const = nullptr;         // r, 1/r, 1/r^2, d1, d2, l(l+1)
    int *d_l_of_j = nullptr;
    double *d_ms = nullptr;            // mean-square work: J + 3
};

int scalar_shell_create(ScalarShell *c, int L, int nr, const double *radius, double dt,
                        const std::vector<ScalarSpec> &spec, const double *val_icb,
                        const double *val_cmb, hipStream_t s, std::string *err);
void scalar_shell_free(ScalarShell *c);
int scalar_shell_set_noise(ScalarShell *c, std::string *err);
int scalar_shell_set_field(ScalarShell *c, int q, const double *host, std::string *err);
int scalar_shell_get_field(ScalarShell *c, int q, double *host, std::string *err);
int scalar_shell_initialize(ScalarShell *c, std::string *err);
int scalar_shell_step(ScalarShell *c, int nsteps, std::string *err);
int scalar_shell_mean_square(ScalarShell *c, double *out, std::string *err);
int scalar_shell_write_volume_ms(ScalarShell *c, const std::string &path, std::string *err);

}  // namespace calypso
