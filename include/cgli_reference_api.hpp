/*
 * cgli_reference_api.hpp -- the reference solver's hot-path routines under
 * their own names and parameter lists (main.h of ricklee209/CGLI_Fu_Flow),
 * implemented over the C ABI of cgli.h.
 *
 * A reference translation unit includes this header after Resolution.h (it
 * uses Y_m and Z_m exactly as the reference's declarations do) and links
 * libcgli instead of Flux_X.cpp, Flux_Y.cpp, Flux_Z.cpp, Viscous_terms.cpp,
 * Dual_time_stepping.cpp, DPLUSGS.cpp, Filter.cpp and Hyperplane.cpp.  The
 * caller keeps owning every array (Array.h); each call copies the arrays it
 * reads into the context's device fields, runs the gfx950 kernels and
 * copies back the arrays the reference writes, so it can replace the
 * reference routine call for call.  A driver that keeps the state resident
 * in HBM across calls uses cgli.h directly instead (no copies).
 *
 * Before the first call, bind a context created for this slab:
 * nx = gcount[myid], ny = Y_out, nz = Z_out, first = (myid == 0),
 * last = (myid == nproc-1), deltas and physical constants from prm.h
 * (cgli_config_reference fills the reference's defaults).  The myid argument
 * is accepted for signature compatibility; the context's configuration
 * describes the slab.
 *
 * Differences a caller can observe, all on scratch the reference reuses:
 * the reconstruction scratch of the flux routines (MR, ML) and the face
 * scratch of Viscous_terms (LR..NL) are left untouched here, and the
 * Nijk/ijk lists passed to Dual_time_stepping are not read (the library
 * builds its own plane order; Hyperplane still writes the reference's
 * lists).  Errors other than a non-finite residual norm throw
 * std::runtime_error (the reference routines have no error path; NaN norms
 * propagate as in the reference).
 */
#ifndef CGLI_REFERENCE_API_HPP
#define CGLI_REFERENCE_API_HPP

#include <stdexcept>
#include <string>

#include "cgli.h"

#if !defined(Y_m) || !defined(Z_m)
#error "include Resolution.h (Y_m, Z_m) before cgli_reference_api.hpp"
#endif

namespace cgli_reference {

typedef double (*A3)[Y_m][Z_m];

inline cgli_ctx *&bound()
{
    static cgli_ctx *ctx = nullptr;
    return ctx;
}

inline void check(int rc, const char *what)
{
    if (rc != CGLI_OK)
        throw std::runtime_error(std::string("cgli ") + what + ": " + cgli_error_string(rc));
}

inline cgli_ctx *ctx()
{
    if (!bound()) throw std::runtime_error("cgli: no context bound (call cgli_reference_bind)");
    return bound();
}

inline void put(int id, A3 a) { check(cgli_set_field(ctx(), id, &a[0][0][0]), "set_field"); }
inline void get(int id, A3 a) { check(cgli_get_field(ctx(), id, &a[0][0][0]), "get_field"); }

inline void put5(int id, A3 a1, A3 a2, A3 a3, A3 a4, A3 a5)
{
    put(id, a1); put(id + 1, a2); put(id + 2, a3); put(id + 3, a4); put(id + 4, a5);
}

inline void get5(int id, A3 a1, A3 a2, A3 a3, A3 a4, A3 a5)
{
    get(id, a1); get(id + 1, a2); get(id + 2, a3); get(id + 3, a4); get(id + 4, a5);
}

inline void put9(int id, A3 m1, A3 m2, A3 m3, A3 m4, A3 m5, A3 m6, A3 m7, A3 m8, A3 m9)
{
    put(id, m1); put(id + 1, m2); put(id + 2, m3); put(id + 3, m4); put(id + 4, m5);
    put(id + 5, m6); put(id + 6, m7); put(id + 7, m8); put(id + 8, m9);
}

}  // namespace cgli_reference

/* the context every routine below runs on */
inline void cgli_reference_bind(cgli_ctx *c) { cgli_reference::bound() = c; }

#define CGLI_A3 cgli_reference::A3

/* Flux_X.cpp:16-66 */
inline void Flux_X(int myid, double Ep, double e,
                   CGLI_A3 U1_, CGLI_A3 U2_, CGLI_A3 U3_, CGLI_A3 U4_, CGLI_A3 U5_,
                   CGLI_A3 MR1, CGLI_A3 MR2, CGLI_A3 MR3, CGLI_A3 MR4, CGLI_A3 MR5,
                   CGLI_A3 ML1, CGLI_A3 ML2, CGLI_A3 ML3, CGLI_A3 ML4, CGLI_A3 ML5,
                   CGLI_A3 J, CGLI_A3 J_u,
                   CGLI_A3 xidx_u, CGLI_A3 xidy_u, CGLI_A3 xidz_u, CGLI_A3 etdx_u, CGLI_A3 etdy_u,
                   CGLI_A3 etdz_u, CGLI_A3 ztdx_u, CGLI_A3 ztdy_u, CGLI_A3 ztdz_u,
                   CGLI_A3 inFx1, CGLI_A3 inFx2, CGLI_A3 inFx3, CGLI_A3 inFx4, CGLI_A3 inFx5, CGLI_A3 EpX)
{
    using namespace cgli_reference;
    (void)myid; (void)MR1; (void)MR2; (void)MR3; (void)MR4; (void)MR5;
    (void)ML1; (void)ML2; (void)ML3; (void)ML4; (void)ML5; (void)EpX;
    put5(CGLI_U_, U1_, U2_, U3_, U4_, U5_);
    put(CGLI_J, J);
    put(CGLI_J_U, J_u);
    put9(CGLI_MET_U, xidx_u, xidy_u, xidz_u, etdx_u, etdy_u, etdz_u, ztdx_u, ztdy_u, ztdz_u);
    check(cgli_flux_x(ctx(), Ep, e), "Flux_X");
    get5(CGLI_INFX, inFx1, inFx2, inFx3, inFx4, inFx5);
}

/* Flux_Y.cpp:15-66 */
inline void Flux_Y(int myid, double Ep, double e,
                   CGLI_A3 U1_, CGLI_A3 U2_, CGLI_A3 U3_, CGLI_A3 U4_, CGLI_A3 U5_,
                   CGLI_A3 MR1, CGLI_A3 MR2, CGLI_A3 MR3, CGLI_A3 MR4, CGLI_A3 MR5,
                   CGLI_A3 ML1, CGLI_A3 ML2, CGLI_A3 ML3, CGLI_A3 ML4, CGLI_A3 ML5,
                   CGLI_A3 J, CGLI_A3 J_v,
                   CGLI_A3 xidx_v, CGLI_A3 xidy_v, CGLI_A3 xidz_v, CGLI_A3 etdx_v, CGLI_A3 etdy_v,
                   CGLI_A3 etdz_v, CGLI_A3 ztdx_v, CGLI_A3 ztdy_v, CGLI_A3 ztdz_v,
                   CGLI_A3 inFy1, CGLI_A3 inFy2, CGLI_A3 inFy3, CGLI_A3 inFy4, CGLI_A3 inFy5, CGLI_A3 EpY)
{
    using namespace cgli_reference;
    (void)myid; (void)MR1; (void)MR2; (void)MR3; (void)MR4; (void)MR5;
    (void)ML1; (void)ML2; (void)ML3; (void)ML4; (void)ML5; (void)EpY;
    put5(CGLI_U_, U1_, U2_, U3_, U4_, U5_);
    put(CGLI_J, J);
    put(CGLI_J_V, J_v);
    put9(CGLI_MET_V, xidx_v, xidy_v, xidz_v, etdx_v, etdy_v, etdz_v, ztdx_v, ztdy_v, ztdz_v);
    check(cgli_flux_y(ctx(), Ep, e), "Flux_Y");
    get5(CGLI_INFY, inFy1, inFy2, inFy3, inFy4, inFy5);
}

/* Flux_Z.cpp:15-66 */
inline void Flux_Z(int myid, double Ep, double e,
                   CGLI_A3 U1_, CGLI_A3 U2_, CGLI_A3 U3_, CGLI_A3 U4_, CGLI_A3 U5_,
                   CGLI_A3 MR1, CGLI_A3 MR2, CGLI_A3 MR3, CGLI_A3 MR4, CGLI_A3 MR5,
                   CGLI_A3 ML1, CGLI_A3 ML2, CGLI_A3 ML3, CGLI_A3 ML4, CGLI_A3 ML5,
                   CGLI_A3 J, CGLI_A3 J_w,
                   CGLI_A3 xidx_w, CGLI_A3 xidy_w, CGLI_A3 xidz_w, CGLI_A3 etdx_w, CGLI_A3 etdy_w,
                   CGLI_A3 etdz_w, CGLI_A3 ztdx_w, CGLI_A3 ztdy_w, CGLI_A3 ztdz_w,
                   CGLI_A3 inFz1, CGLI_A3 inFz2, CGLI_A3 inFz3, CGLI_A3 inFz4, CGLI_A3 inFz5, CGLI_A3 EpZ)
{
    using namespace cgli_reference;
    (void)myid; (void)MR1; (void)MR2; (void)MR3; (void)MR4; (void)MR5;
    (void)ML1; (void)ML2; (void)ML3; (void)ML4; (void)ML5; (void)EpZ;
    put5(CGLI_U_, U1_, U2_, U3_, U4_, U5_);
    put(CGLI_J, J);
    put(CGLI_J_W, J_w);
    put9(CGLI_MET_W, xidx_w, xidy_w, xidz_w, etdx_w, etdy_w, etdz_w, ztdx_w, ztdy_w, ztdz_w);
    check(cgli_flux_z(ctx(), Ep, e), "Flux_Z");
    get5(CGLI_INFZ, inFz1, inFz2, inFz3, inFz4, inFz5);
}

/* Viscous_terms.cpp:15-140 */
inline void Viscous_terms(int myid,
                          CGLI_A3 U1_, CGLI_A3 U2_, CGLI_A3 U3_, CGLI_A3 U4_, CGLI_A3 U5_,
                          CGLI_A3 vF2, CGLI_A3 vF3, CGLI_A3 vF4, CGLI_A3 vF5,
                          CGLI_A3 J,
                          CGLI_A3 xidx, CGLI_A3 xidy, CGLI_A3 xidz, CGLI_A3 etdx, CGLI_A3 etdy, CGLI_A3 etdz,
                          CGLI_A3 ztdx, CGLI_A3 ztdy, CGLI_A3 ztdz,
                          CGLI_A3 J_u,
                          CGLI_A3 xidx_u, CGLI_A3 xidy_u, CGLI_A3 xidz_u, CGLI_A3 etdx_u, CGLI_A3 etdy_u,
                          CGLI_A3 etdz_u, CGLI_A3 ztdx_u, CGLI_A3 ztdy_u, CGLI_A3 ztdz_u,
                          CGLI_A3 J_v,
                          CGLI_A3 xidx_v, CGLI_A3 xidy_v, CGLI_A3 xidz_v, CGLI_A3 etdx_v, CGLI_A3 etdy_v,
                          CGLI_A3 etdz_v, CGLI_A3 ztdx_v, CGLI_A3 ztdy_v, CGLI_A3 ztdz_v,
                          CGLI_A3 J_w,
                          CGLI_A3 xidx_w, CGLI_A3 xidy_w, CGLI_A3 xidz_w, CGLI_A3 etdx_w, CGLI_A3 etdy_w,
                          CGLI_A3 etdz_w, CGLI_A3 ztdx_w, CGLI_A3 ztdy_w, CGLI_A3 ztdz_w,
                          CGLI_A3 LR1, CGLI_A3 LR2, CGLI_A3 LR3, CGLI_A3 LR4, CGLI_A3 LR5,
                          CGLI_A3 LL1, CGLI_A3 LL2, CGLI_A3 LL3, CGLI_A3 LL4, CGLI_A3 LL5,
                          CGLI_A3 MR1, CGLI_A3 MR2, CGLI_A3 MR3, CGLI_A3 MR4, CGLI_A3 MR5,
                          CGLI_A3 ML1, CGLI_A3 ML2, CGLI_A3 ML3, CGLI_A3 ML4, CGLI_A3 ML5,
                          CGLI_A3 NR1, CGLI_A3 NR2, CGLI_A3 NR3, CGLI_A3 NR4, CGLI_A3 NR5,
                          CGLI_A3 NL1, CGLI_A3 NL2, CGLI_A3 NL3, CGLI_A3 NL4, CGLI_A3 NL5)
{
    using namespace cgli_reference;
    (void)myid;
    (void)xidx; (void)xidy; (void)xidz; (void)etdx; (void)etdy; (void)etdz; (void)ztdx; (void)ztdy; (void)ztdz;
    (void)LR1; (void)LR2; (void)LR3; (void)LR4; (void)LR5; (void)LL1; (void)LL2; (void)LL3; (void)LL4; (void)LL5;
    (void)MR1; (void)MR2; (void)MR3; (void)MR4; (void)MR5; (void)ML1; (void)ML2; (void)ML3; (void)ML4; (void)ML5;
    (void)NR1; (void)NR2; (void)NR3; (void)NR4; (void)NR5; (void)NL1; (void)NL2; (void)NL3; (void)NL4; (void)NL5;
    put5(CGLI_U_, U1_, U2_, U3_, U4_, U5_);
    put(CGLI_J, J);
    put(CGLI_J_U, J_u);
    put9(CGLI_MET_U, xidx_u, xidy_u, xidz_u, etdx_u, etdy_u, etdz_u, ztdx_u, ztdy_u, ztdz_u);
    put(CGLI_J_V, J_v);
    put9(CGLI_MET_V, xidx_v, xidy_v, xidz_v, etdx_v, etdy_v, etdz_v, ztdx_v, ztdy_v, ztdz_v);
    put(CGLI_J_W, J_w);
    put9(CGLI_MET_W, xidx_w, xidy_w, xidz_w, etdx_w, etdy_w, etdz_w, ztdx_w, ztdy_w, ztdz_w);
    check(cgli_viscous(ctx()), "Viscous_terms");
    get(CGLI_VF, vF2); get(CGLI_VF + 1, vF3); get(CGLI_VF + 2, vF4); get(CGLI_VF + 3, vF5);
}

namespace cgli_reference {
/* inputs shared by Dual_time_stepping and DPLUSGS */
inline void put_implicit(A3 U1_, A3 U2_, A3 U3_, A3 U4_, A3 U5_, A3 U1, A3 U2, A3 U3, A3 U4, A3 U5, A3 U1q,
                         A3 U2q, A3 U3q, A3 U4q, A3 U5q, A3 J, A3 xidx, A3 xidy, A3 xidz, A3 etdx, A3 etdy,
                         A3 etdz, A3 ztdx, A3 ztdy, A3 ztdz, A3 inFx1, A3 inFx2, A3 inFx3, A3 inFx4, A3 inFx5,
                         A3 inFy1, A3 inFy2, A3 inFy3, A3 inFy4, A3 inFy5, A3 inFz1, A3 inFz2, A3 inFz3,
                         A3 inFz4, A3 inFz5, A3 vF2, A3 vF3, A3 vF4, A3 vF5)
{
    put5(CGLI_U_, U1_, U2_, U3_, U4_, U5_);
    put5(CGLI_U, U1, U2, U3, U4, U5);
    put5(CGLI_UQ, U1q, U2q, U3q, U4q, U5q);
    put(CGLI_J, J);
    put9(CGLI_MET, xidx, xidy, xidz, etdx, etdy, etdz, ztdx, ztdy, ztdz);
    put5(CGLI_INFX, inFx1, inFx2, inFx3, inFx4, inFx5);
    put5(CGLI_INFY, inFy1, inFy2, inFy3, inFy4, inFy5);
    put5(CGLI_INFZ, inFz1, inFz2, inFz3, inFz4, inFz5);
    put(CGLI_VF, vF2); put(CGLI_VF + 1, vF3); put(CGLI_VF + 2, vF4); put(CGLI_VF + 3, vF5);
}

inline void implicit_result(int rc, const char *what)
{
    if (rc != CGLI_E_NONFINITE) check(rc, what);
}
}  // namespace cgli_reference

/* Dual_time_stepping.cpp:18-130: the reference runs 2 symmetric sweeps */
inline void Dual_time_stepping(int myid, double deltaT, double e,
                               double *er1, double *er2, double *er3, double *er4, double *er5,
                               int (*Nijk)[2], int (*ijk)[3],
                               CGLI_A3 U1_, CGLI_A3 U2_, CGLI_A3 U3_, CGLI_A3 U4_, CGLI_A3 U5_,
                               CGLI_A3 U1, CGLI_A3 U2, CGLI_A3 U3, CGLI_A3 U4, CGLI_A3 U5,
                               CGLI_A3 U1q, CGLI_A3 U2q, CGLI_A3 U3q, CGLI_A3 U4q, CGLI_A3 U5q,
                               CGLI_A3 J,
                               CGLI_A3 xidx, CGLI_A3 xidy, CGLI_A3 xidz, CGLI_A3 etdx, CGLI_A3 etdy,
                               CGLI_A3 etdz, CGLI_A3 ztdx, CGLI_A3 ztdy, CGLI_A3 ztdz,
                               CGLI_A3 inFx1, CGLI_A3 inFx2, CGLI_A3 inFx3, CGLI_A3 inFx4, CGLI_A3 inFx5,
                               CGLI_A3 inFy1, CGLI_A3 inFy2, CGLI_A3 inFy3, CGLI_A3 inFy4, CGLI_A3 inFy5,
                               CGLI_A3 inFz1, CGLI_A3 inFz2, CGLI_A3 inFz3, CGLI_A3 inFz4, CGLI_A3 inFz5,
                               CGLI_A3 vF2, CGLI_A3 vF3, CGLI_A3 vF4, CGLI_A3 vF5,
                               CGLI_A3 MR1, CGLI_A3 MR2, CGLI_A3 MR3, CGLI_A3 MR4, CGLI_A3 MR5,
                               CGLI_A3 NR1, CGLI_A3 NR2, CGLI_A3 NR3, CGLI_A3 NR4, CGLI_A3 NR5,
                               CGLI_A3 NL1, CGLI_A3 NL2, CGLI_A3 NL3, CGLI_A3 NL4, CGLI_A3 NL5,
                               CGLI_A3 Residual1, CGLI_A3 Residual2, CGLI_A3 Residual3, CGLI_A3 Residual4,
                               CGLI_A3 Residual5, CGLI_A3 etdy_v, CGLI_A3 J_v)
{
    using namespace cgli_reference;
    (void)myid; (void)Nijk; (void)ijk;
    put_implicit(U1_, U2_, U3_, U4_, U5_, U1, U2, U3, U4, U5, U1q, U2q, U3q, U4q, U5q, J, xidx, xidy, xidz, etdx,
                 etdy, etdz, ztdx, ztdy, ztdz, inFx1, inFx2, inFx3, inFx4, inFx5, inFy1, inFy2, inFy3, inFy4,
                 inFy5, inFz1, inFz2, inFz3, inFz4, inFz5, vF2, vF3, vF4, vF5);
    put(CGLI_MET_V + 4, etdy_v);
    put(CGLI_J_V, J_v);
    double er[5];
    implicit_result(cgli_dual_time_stepping(ctx(), deltaT, e, er, 2, 1), "Dual_time_stepping");
    *er1 = er[0]; *er2 = er[1]; *er3 = er[2]; *er4 = er[3]; *er5 = er[4];
    get5(CGLI_U_, U1_, U2_, U3_, U4_, U5_);
    get5(CGLI_MR, MR1, MR2, MR3, MR4, MR5);
    get5(CGLI_NR, NR1, NR2, NR3, NR4, NR5);
    get5(CGLI_NL, NL1, NL2, NL3, NL4, NL5);
    get5(CGLI_RES, Residual1, Residual2, Residual3, Residual4, Residual5);
}

/* DPLUSGS.cpp:18-122: the reference runs 4 Jacobi sweeps */
inline void DPLUSGS(int myid, double deltaT, double *er1, double *er2, double *er3, double *er4, double *er5,
                    double e,
                    CGLI_A3 U1_, CGLI_A3 U2_, CGLI_A3 U3_, CGLI_A3 U4_, CGLI_A3 U5_,
                    CGLI_A3 U1, CGLI_A3 U2, CGLI_A3 U3, CGLI_A3 U4, CGLI_A3 U5,
                    CGLI_A3 U1q, CGLI_A3 U2q, CGLI_A3 U3q, CGLI_A3 U4q, CGLI_A3 U5q,
                    CGLI_A3 J,
                    CGLI_A3 xidx, CGLI_A3 xidy, CGLI_A3 xidz, CGLI_A3 etdx, CGLI_A3 etdy,
                    CGLI_A3 etdz, CGLI_A3 ztdx, CGLI_A3 ztdy, CGLI_A3 ztdz,
                    CGLI_A3 inFx1, CGLI_A3 inFx2, CGLI_A3 inFx3, CGLI_A3 inFx4, CGLI_A3 inFx5,
                    CGLI_A3 inFy1, CGLI_A3 inFy2, CGLI_A3 inFy3, CGLI_A3 inFy4, CGLI_A3 inFy5,
                    CGLI_A3 inFz1, CGLI_A3 inFz2, CGLI_A3 inFz3, CGLI_A3 inFz4, CGLI_A3 inFz5,
                    CGLI_A3 vF2, CGLI_A3 vF3, CGLI_A3 vF4, CGLI_A3 vF5,
                    CGLI_A3 MR1, CGLI_A3 MR2, CGLI_A3 MR3, CGLI_A3 MR4, CGLI_A3 MR5,
                    CGLI_A3 NR1, CGLI_A3 NR2, CGLI_A3 NR3, CGLI_A3 NR4, CGLI_A3 NR5,
                    CGLI_A3 NL1, CGLI_A3 NL2, CGLI_A3 NL3, CGLI_A3 NL4, CGLI_A3 NL5,
                    CGLI_A3 Residual1, CGLI_A3 Residual2, CGLI_A3 Residual3, CGLI_A3 Residual4,
                    CGLI_A3 Residual5)
{
    using namespace cgli_reference;
    (void)myid;
    put_implicit(U1_, U2_, U3_, U4_, U5_, U1, U2, U3, U4, U5, U1q, U2q, U3q, U4q, U5q, J, xidx, xidy, xidz, etdx,
                 etdy, etdz, ztdx, ztdy, ztdz, inFx1, inFx2, inFx3, inFx4, inFx5, inFy1, inFy2, inFy3, inFy4,
                 inFy5, inFz1, inFz2, inFz3, inFz4, inFz5, vF2, vF3, vF4, vF5);
    double er[5];
    implicit_result(cgli_dplusgs(ctx(), deltaT, er, e, 4), "DPLUSGS");
    *er1 = er[0]; *er2 = er[1]; *er3 = er[2]; *er4 = er[3]; *er5 = er[4];
    get5(CGLI_U_, U1_, U2_, U3_, U4_, U5_);
    get5(CGLI_MR, MR1, MR2, MR3, MR4, MR5);
    get5(CGLI_NR, NR1, NR2, NR3, NR4, NR5);
    get5(CGLI_NL, NL1, NL2, NL3, NL4, NL5);
    get5(CGLI_RES, Residual1, Residual2, Residual3, Residual4, Residual5);
}

/* Filter.cpp:12-36 */
inline void Filter(int myid, double Roe_criterion, double E_high, double E_low,
                   CGLI_A3 U1_, CGLI_A3 U2_, CGLI_A3 U3_, CGLI_A3 U4_, CGLI_A3 U5_,
                   CGLI_A3 Y_point, CGLI_A3 EpX, CGLI_A3 EpY, CGLI_A3 EpZ, CGLI_A3 J)
{
    using namespace cgli_reference;
    (void)myid;
    put5(CGLI_U_, U1_, U2_, U3_, U4_, U5_);
    put(CGLI_YPOINT, Y_point);
    put(CGLI_J, J);
    /* the sensor writes only its face windows: the other entries keep the
     * caller's values, as in the reference */
    put(CGLI_EPX, EpX); put(CGLI_EPY, EpY); put(CGLI_EPZ, EpZ);
    check(cgli_filter(ctx(), Roe_criterion, E_high, E_low), "Filter");
    get(CGLI_EPX, EpX); get(CGLI_EPY, EpY); get(CGLI_EPZ, EpZ);
}

/* Hyperplane.cpp:18-28 */
inline void Hyperplane(int myid, int (*Nijk)[2], int (*ijk)[3])
{
    using namespace cgli_reference;
    (void)myid;
    check(cgli_hyperplane(ctx(), &Nijk[0][0], &ijk[0][0]), "Hyperplane");
}

#undef CGLI_A3

#endif
