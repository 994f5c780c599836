// Drop-in replacements, with the reference's own signatures, for the hot-path
// entry points of the serial driver (main/digasoser.cpp, solvers/dgsolvers.cpp):
//
//   assembleLinearSystem   assembly/globalassembly.cpp:423-556
//   assembleRHS            assembly/globalassembly.cpp:559-616
//   computeResidualNorm    assembly/globalassembly.cpp:618-678
//   computeOrdering        solvers/dgsolvers.cpp:16  (closed in the reference)
//   computePreconditioner  solvers/dgsolvers.cpp:18  (closed in the reference)
//   gmres                  solvers/dgsolvers.cpp:20  (closed in the reference)
//
// Built against the reference's preprocessing/digasopre.h (the Makefile beside
// this file) and linked in place of those objects; the work runs on the GPU
// through the C ABI of libhfem_amd (include/hfem_amd.h).  The first call creates
// a device context from the structs (master tables of element type 0 -- the
// C++ uses face 0's tables for every face, assembly/elementassembly.cpp:161,593 --
// with perm packed face by face, the far-field state app.bcs of boundary 0, the
// mesh arrays and the trace graph sys.ent2entStart / sys.ent2ent); every call
// uploads the state (sol.UDG, sol.UH, sol.SH, app.fc_u/fc_q/tdep) and copies back
// what the rest of the driver reads (sys.Hg, sys.Rg, sol.DinvF, sol.DinvRu,
// sol.rNorm; sys.x and sys.ordered2unordered / unordered2ordered from the solver
// calls).  Built with -DHFEM_DIGASO_RESIDENT the big
// copies (sys.Hg, sol.DinvF) are skipped: valid when gmres and the solution
// update also run on the device, as in hfem_solve_steady.
//
// Differences from the reference, by design: an error prints the library's
// message and exits, as the reference's error() does (preprocessing/errormsg.cpp);
// the element blocks (elems, temps) are not used; computeResidualNorm sets
// sol.rNorm but leaves sys.Rg and sys.r (the line search reads only the norm).
#include <string>
#include <vector>
#include <iostream>
#include <algorithm>
#include <math.h>
#include <stdio.h>
#include <stdlib.h>
using namespace std;
#include "preprocessing/digasopre.h"

#include "hfem_amd.h"

namespace {

hfem_ctx *g_ctx = nullptr;
long g_key[4] = {-1, -1, -1, -1};

[[noreturn]] void die(const char *msg)
{
    fprintf(stderr, "%s\n", msg);
    exit(1);
}

[[noreturn]] void fail(const char *who)
{
    fprintf(stderr, "%s: %s\n", who, hfem_last_error());
    exit(1);
}

void check(int rc, const char *who)
{
    if (rc) fail(who);
}

// Context for these structs: created at the first call, kept while the sizes hold.
hfem_ctx *context(sysstruct &sys, meshstruct &mesh, masterstruct &master, appstruct &app)
{
    const long key[4] = {(long)mesh.ne, (long)sys.numEntities, (long)master.npv, (long)master.ngv};
    if (g_ctx && std::equal(key, key + 4, g_key)) return g_ctx;
    if (g_ctx) hfem_destroy(g_ctx);
    g_ctx = nullptr;
    const int nfe = master.nfe, npf = master.npf[0];
    hfem_dims d = {(int)master.nd, (int)app.ncu, (int)master.npv, (int)master.ngv, npf, (int)master.ngf[0],
                   nfe, (int)mesh.ne, (int)sys.numEntities};
    check(hfem_create(&g_ctx, &d, 0), "hfem_create");
    // perm is a vector per face: packed face by face, (npf, nfe)
    vector<int> perm((size_t)npf * nfe);
    for (int a = 0; a < nfe; a++)
        for (int k = 0; k < npf; k++) perm[(size_t)a * npf + k] = (int)master.perm[a][k];
    check(hfem_set_master(g_ctx, &master.shapvt[0], &master.shapvg[0], &master.shapft[0][0], &master.shapfg[0][0],
                          &perm[0]),
          "hfem_set_master");
    vector<int> elcon(mesh.elcon.begin(), mesh.elcon.end()), bf(mesh.bf.begin(), mesh.bf.end()),
        t2f(mesh.t2f.begin(), mesh.t2f.end()), rs(sys.ent2entStart.begin(), sys.ent2entStart.end()),
        cols(sys.ent2ent.begin(), sys.ent2ent.end());
    check(hfem_set_mesh(g_ctx, &mesh.dgnodes[0], &elcon[0], &bf[0], &t2f[0], &rs[0], &cols[0],
                        (int)sys.ent2entStart[sys.numEntities]),
          "hfem_set_mesh");
    std::copy(key, key + 4, g_key);
    return g_ctx;
}

// Physics and state of this call (fc_u changes per DIRK stage, the state per step).
void upload(hfem_ctx *c, solstruct &sol, appstruct &app)
{
    check(hfem_set_physics(c, &app.physicsparam[0], &app.bcs[0], app.fc_u, app.fc_q, (int)app.tdep),
          "hfem_set_physics");
    check(hfem_set_solution(c, &sol.UDG[0], &sol.UH[0], sol.SH.empty() ? nullptr : &sol.SH[0]),
          "hfem_set_solution");
}

}  // namespace

// Releases the device context (not in the reference: call at the end of a run).
void hfemDigasoRelease()
{
    if (g_ctx) hfem_destroy(g_ctx);
    g_ctx = nullptr;
}

void assembleLinearSystem(sysstruct &sys, elemstruct *elems, meshstruct &mesh, masterstruct &master,
                          solstruct &sol, appstruct &app, tempstruct *temps)
{
    (void)elems;
    (void)temps;
    hfem_ctx *c = context(sys, mesh, master, app);
    upload(c, sol, app);
    if (app.quasiNewton == 1) check(hfem_quasi_newton_storage(c, (int)sys.quasiNewtonAccuracy), "hfem_quasi_newton_storage");
    double rNorm = 0.0;
    check(hfem_assemble_linear_system(c, HFEM_FP64, &rNorm), "hfem_assemble_linear_system");
    sol.rNorm = rNorm;
#ifndef HFEM_DIGASO_RESIDENT
    check(hfem_read_system(c, &sys.Hg[0], &sys.Rg[0]), "hfem_read_system");
    check(hfem_read_condensed(c, 0, (int)mesh.ne, &sol.DinvRu[0], &sol.DinvF[0], nullptr, nullptr, nullptr),
          "hfem_read_condensed");
    if (sys.matvecPrecision == 0)
        for (size_t i = 0; i < sys.Hg.size(); i++) sys.Hg_sp[i] = (float)sys.Hg[i];
#else
    check(hfem_read_system(c, nullptr, &sys.Rg[0]), "hfem_read_system");
    check(hfem_read_condensed(c, 0, (int)mesh.ne, &sol.DinvRu[0], nullptr, nullptr, nullptr, nullptr),
          "hfem_read_condensed");
#endif
}

void assembleRHS(sysstruct &sys, elemstruct *elems, meshstruct &mesh, masterstruct &master, solstruct &sol,
                 appstruct &app, tempstruct *temps)
{
    (void)elems;
    (void)temps;
    hfem_ctx *c = context(sys, mesh, master, app);
    upload(c, sol, app);
    double rNorm = 0.0;
    check(hfem_assemble_rhs(c, &rNorm), "hfem_assemble_rhs");
    sol.rNorm = rNorm;
    check(hfem_read_system(c, nullptr, &sys.Rg[0]), "hfem_read_system");
    check(hfem_read_condensed(c, 0, (int)mesh.ne, &sol.DinvRu[0], nullptr, nullptr, nullptr, nullptr),
          "hfem_read_condensed");
}

void computeResidualNorm(sysstruct &sys, elemstruct *elems, meshstruct &mesh, masterstruct &master,
                         solstruct &sol, appstruct &app, tempstruct *temps)
{
    (void)elems;
    (void)temps;
    hfem_ctx *c = context(sys, mesh, master, app);
    upload(c, sol, app);
    double rNorm = 0.0;
    check(hfem_residual_norm(c, &rNorm), "hfem_residual_norm");
    sol.rNorm = rNorm;
}

// sys.reorderMethod: 0 natural, 1 approximate MDF (the methods the serial
// reference offers for BILU0; 2 exact MDF is not implemented here)
void computeOrdering(sysstruct &sys, Int preconditioner)
{
    (void)preconditioner;
    if (!g_ctx) die("computeOrdering: call assembleLinearSystem first");
    if (sys.reorderMethod != 0 && sys.reorderMethod != 1) die("computeOrdering: reorderMethod must be 0 or 1");
    check(hfem_compute_ordering(g_ctx, (int)sys.reorderMethod), "hfem_compute_ordering");
    const int n = (int)sys.numEntities;
    vector<int> o2u(n);
    check(hfem_get_ordering(g_ctx, &o2u[0], nullptr), "hfem_get_ordering");
    sys.ordered2unordered.assign(o2u.begin(), o2u.end());
    sys.unordered2ordered.assign(n, 0);
    for (int l = 0; l < n; l++) sys.unordered2ordered[o2u[l]] = l;
}

// sys.preconditioner: -1 none; 0 RAS and 1 subdomain block Jacobi are BILU0 on
// one rank; 2 entity-wise block Jacobi.  sys.precPrecision 0 keeps the factors in
// single precision.
void computePreconditioner(sysstruct &sys, Int preconditioner)
{
    if (!g_ctx) die("computePreconditioner: call assembleLinearSystem first");
    const int kind = sys.preconditionerSide == -1 ? -1 : (preconditioner == 2 ? 2 : 0);
    check(hfem_set_precond_precision(g_ctx, sys.precPrecision == 0 ? 0 : 1), "hfem_set_precond_precision");
    check(hfem_precond_factor(g_ctx, kind), "hfem_precond_factor");
}

// Left-preconditioned restarted GMRES from x = 0 on the system of the last
// assembly (kernel/gmres_bilu0.m); *convFlag 0 converged, 1 iteration limit.
void gmres(sysstruct &sys, Int *convFlag)
{
    if (!g_ctx) die("gmres: call assembleLinearSystem first");
    int it = 0, flag = 0;
    const int orthog = sys.orthogMethod == 0 ? 0 : 1;
    check(hfem_gmres(g_ctx, (int)sys.restart, sys.tol, (int)sys.maxiter, orthog, &it, nullptr, &flag), "hfem_gmres");
    if ((long)sys.x.size() < (long)sys.numEntities * sys.blkSize) sys.x.resize((size_t)sys.numEntities * sys.blkSize);
    check(hfem_read_trace_solution(g_ctx, &sys.x[0]), "hfem_read_trace_solution");
    *convFlag = flag;
}
