// ref_harness.cpp -- ctypes-callable driver around the UNMODIFIED reference cpu/ sources
// (TEST INFRASTRUCTURE).
//
// oracle/Makefile compiles cpu/{ctoprim,hypterm,diffterm,helper_functions,advance}.cpp
// where they lie (g++ -O2, the flags of cpu/Makefile:2) and links them with this file into
// oracle/_ref/libcnsref.so.  This file only supplies what cpu/main.cpp normally supplies
// (the globals at cpu/main.cpp:4-7 and the constant table that read_configurations fills
// at cpu/helper_functions.cpp:55-82), and builds the `double ****` row-pointer views the
// reference functions take over flat caller-owned AoS buffers laid out [i][j][k][l]
// (cpu/helper_functions.cpp:100-122 lays its arrays out the same way).
#include <cstdlib>
#include <vector>
#include "header.h"
#include "helper_functions.h"

global_const_t h_const;
double ctoprim_time = 0.0;
double diffterm_time = 0.0;
double hypterm_time = 0.0;

namespace {

// Row-pointer view over a flat [d0][d1][d2][dl] buffer.
struct View4 {
    std::vector<double ***> l0;
    std::vector<double **> l1;
    std::vector<double *> l2;
    double ****p = nullptr;
    View4(double *base, int d0, int d1, int d2, int dl)
        : l0(d0), l1((size_t)d0 * d1), l2((size_t)d0 * d1 * d2)
    {
        for (int i = 0; i < d0; i++) {
            l0[i] = &l1[(size_t)i * d1];
            for (int j = 0; j < d1; j++) {
                l1[(size_t)i * d1 + j] = &l2[((size_t)i * d1 + j) * d2];
                for (int k = 0; k < d2; k++)
                    l2[((size_t)i * d1 + j) * d2 + k] = base + (((size_t)i * d1 + j) * d2 + k) * dl;
            }
        }
        p = l0.data();
    }
};

int g_lo[3], g_hi[3];

}  // namespace

extern "C" {

// Fill h_const for a box of interior edges n0,n1,n2 (lo = ng, hi = n-1+ng as cpu/advance.cpp:33-34).
int ref_setup(int n0, int n1, int n2, double dx0, double dx1, double dx2,
              double cfl, double eta, double alam)
{
    int n[3] = {n0, n1, n2};
    double dx[3] = {dx0, dx1, dx2};
    h_const.ng = NG;
    h_const.nc = NC;
    h_const.ncells = n0;
    h_const.nsteps = 1;
    h_const.cfl = cfl;
    h_const.eta = eta;
    h_const.alam = alam;
    h_const.dt = 0.0;
    for (int d = 0; d < 3; d++) {
        h_const.lo[d] = 0;
        h_const.hi[d] = n[d] - 1;
        h_const.dim[d] = n[d];
        h_const.dim_g[d] = n[d] + 2 * NG;
        h_const.dx[d] = dx[d];
        h_const.dxinv[d] = 1.0E0 / dx[d];
        g_lo[d] = NG;
        g_hi[d] = n[d] - 1 + NG;
    }
    h_const.comp_offset_g = h_const.dim_g[0] * h_const.dim_g[1] * h_const.dim_g[2];
    h_const.comp_offset = n0 * n1 * n2;
    h_const.plane_offset_g = h_const.dim_g[1] * h_const.dim_g[2];
    h_const.plane_offset = n1 * n2;
    h_const.ALP = 0.8E0;
    h_const.BET = -0.2E0;
    h_const.GAM = 4.0E0 / 105.0E0;
    h_const.DEL = -1.0E0 / 280.0E0;
    h_const.OneThird = 1.0E0 / 3.0E0;
    h_const.TwoThirds = 2.0E0 / 3.0E0;
    h_const.FourThirds = 4.0E0 / 3.0E0;
    h_const.OneQuarter = 1.E0 / 4.E0;
    h_const.ThreeQuarters = 3.E0 / 4.E0;
    h_const.CENTER = -205.0E0 / 72.0E0;
    h_const.OFF1 = 8.0E0 / 5.0E0;
    h_const.OFF2 = -0.2E0;
    h_const.OFF3 = 8.0E0 / 315.0E0;
    h_const.OFF4 = -1.0E0 / 560.0E0;
    return 0;
}

void ref_fill_boundary(double *U)
{
    int *g = h_const.dim_g;
    View4 u(U, g[0], g[1], g[2], NC);
    fill_boundary(u.p, h_const.dim, h_const.dim_g);
}

// courno == NULL selects the overload without the Courant reduction (cpu/ctoprim.cpp:72).
void ref_ctoprim(double *U, double *Q, double *courno)
{
    int *g = h_const.dim_g;
    View4 u(U, g[0], g[1], g[2], NC), q(Q, g[0], g[1], g[2], NC + 1);
    if (courno)
        ctoprim(g_lo, g_hi, u.p, q.p, h_const.dx, h_const.ng, *courno);
    else
        ctoprim(g_lo, g_hi, u.p, q.p, h_const.dx, h_const.ng);
}

void ref_hypterm(double *U, double *Q, double *F)
{
    int *g = h_const.dim_g, *d = h_const.dim;
    View4 u(U, g[0], g[1], g[2], NC), q(Q, g[0], g[1], g[2], NC + 1), f(F, d[0], d[1], d[2], NC);
    hypterm(g_lo, g_hi, h_const.ng, h_const.dx, u.p, q.p, f.p);
}

void ref_diffterm(double *Q, double *D)
{
    int *g = h_const.dim_g, *d = h_const.dim;
    View4 q(Q, g[0], g[1], g[2], NC + 1), dd(D, d[0], d[1], d[2], NC);
    diffterm(g_lo, g_hi, h_const.ng, h_const.dx, q.p, dd.p, h_const.eta, h_const.alam);
}

// One call of new_advance (cpu/advance.cpp:13-154); cubic boxes only, as the reference.
void ref_new_advance(double *U, double *Unew, double *Q, double *D, double *F, double *dt)
{
    int *g = h_const.dim_g, *d = h_const.dim;
    View4 u(U, g[0], g[1], g[2], NC), un(Unew, g[0], g[1], g[2], NC);
    View4 q(Q, g[0], g[1], g[2], NC + 1);
    View4 dd(D, d[0], d[1], d[2], NC), f(F, d[0], d[1], d[2], NC);
    new_advance(u.p, un.p, q.p, dd.p, f.p, *dt);
}

}  // extern "C"
