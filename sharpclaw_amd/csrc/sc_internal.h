// Helpers shared by the C ABI translation units (sc_capi.cpp, sc_batch.hip,
// sc_serve.cpp).
#pragma once
#include <cmath>

#include "../../include/sharpclaw_amd.h"
#include "sc_kernels.h"

// Records the message returned by sc_last_error() and returns code.
int sc_fail(int code, const char *fmt, ...);
// Options the device path implements (grid: also the per-grid options).
int sc_check_params(const sc_params *p, bool grid);
// PhysId of the device physics, -1 if none.
int sc_phys_id(const sc_params *p);
// A side with mthbc 0: its ghost cells come from the user's bc routine
// (bc1.f90:34-37, bc2.f90:35,125), which the device cannot run between stages.
inline bool sc_user_bc(const sc_params *p)
{
    for (int k = 0; k < 2 * p->ndim; k++)
        if (p->mthbc[k] == 0) return true;
    return false;
}


template <class T> inline sc::PhysConst<T> sc_make_const(const sc_params *p)
{
    sc::PhysConst<T> c;
    c.gamma = T(p->gamma);
    c.gamma1 = T(p->gamma - 1.0);
    c.grav = T(p->grav);
    c.u = T(p->u);
    double cc = std::sqrt(p->bulk / p->rho);
    c.cc = T(cc);
    c.zz = T(cc * p->rho);
    c.hzz = T(0.5 / (cc * p->rho));
    if (p->efix >= 0)
        c.efix = p->efix;
    else if (p->physics == SC_PHYS_EULER)
        c.efix = 1;  // rp1eu.f:40
    else if (p->physics == SC_PHYS_SHALLOW_WATER)
        c.efix = p->ndim == 1;  // rp1sw.f:40 on, rpn2sw.f:48 off
    else
        c.efix = 0;
    c.tfluct = p->tfluct_solver == 1;
    for (int k = 0; k < 4; k++) c.mthlim[k] = p->mthlim[k];
    return c;
}
