// Test harness for the reference-signature entry points (hfem_amd/shim): fills
// the reference's own structs (preprocessing/digasopre.h, included from the
// reference tree at build time) from flat arrays, calls assembleLinearSystem,
// computeResidualNorm, computeOrdering, computePreconditioner and gmres exactly
// as solveLinearProblem does (solvers/dgsolvers.cpp:40-113), and hands the
// struct fields those calls wrote back to the caller.  Test infrastructure only.
#include <string>
#include <vector>
#include <iostream>
#include <algorithm>
#include <string.h>
#include <math.h>
using namespace std;
#include "preprocessing/digasopre.h"

void assembleLinearSystem(sysstruct &, elemstruct *, meshstruct &, masterstruct &, solstruct &, appstruct &,
                          tempstruct *);
void computeResidualNorm(sysstruct &, elemstruct *, meshstruct &, masterstruct &, solstruct &, appstruct &,
                         tempstruct *);
void computeOrdering(sysstruct &, Int);
void computePreconditioner(sysstruct &, Int);
void gmres(sysstruct &, Int *);
void hfemDigasoRelease();

// dims = {nd, ncu, npv, ngv, npf, ngf, nfe, ne, nf, nblocks}
// solver = {preconditioner, precPrecision, reorderMethod, orthogMethod, restart, maxiter}
// out: Hg, Rg, DinvRu, DinvF, x (any may be NULL); scal = {rNorm after assembly,
// rNorm of computeResidualNorm, convFlag}; o2u (nf)
extern "C" int shim_run(const int *dims, const double *shapvt, const double *shapvg, const double *shapft,
                        const double *shapfg, const int *perm, const double *param, const double *ui, double fc_u,
                        double fc_q, int tdep, const double *dgnodes, const int *elcon, const int *bf, const int *t2f,
                        const int *rowstart, const int *cols, const double *UDG, const double *UH, const double *SH,
                        const int *solver, double tol, double *Hg, double *Rg, double *DinvRu, double *DinvF,
                        double *x, double *scal, int *o2u)
{
    const int nd = dims[0], ncu = dims[1], npv = dims[2], ngv = dims[3], npf = dims[4], ngf = dims[5], nfe = dims[6];
    const int ne = dims[7], nf = dims[8], nblocks = dims[9];
    const int nch = ncu, nc = ncu * (nd + 1), nd1 = nd + 1, ndf = npf * nfe, bsz = nch * npf;
    const int na = npv * ncu, nb = ndf * nch;
    appstruct app;
    masterstruct master;
    meshstruct mesh;
    solstruct sol;
    sysstruct sys;
    app.nd = nd; app.ncu = ncu; app.nch = nch; app.nc = nc;
    app.physicsparam.assign(param, param + 6);
    app.bcs.assign(ui, ui + nch);
    app.fc_u = fc_u; app.fc_q = fc_q; app.tdep = tdep; app.quasiNewton = 0;
    master.nd = nd; master.npv = npv; master.ngv = ngv; master.nfe = nfe;
    master.npf.assign(nfe, npf);
    master.ngf.assign(nfe, ngf);
    master.shapvt.assign(shapvt, shapvt + (size_t)ngv * npv * nd1);
    master.shapvg.assign(shapvg, shapvg + (size_t)ngv * npv * nd1);
    master.shapft.assign(nfe, vector<double>(shapft, shapft + (size_t)ngf * npf * nd));
    master.shapfg.assign(nfe, vector<double>(shapfg, shapfg + (size_t)ngf * npf * nd));
    master.perm.resize(nfe);
    for (int a = 0; a < nfe; a++) master.perm[a].assign(perm + a * npf, perm + (a + 1) * npf);
    mesh.ne = ne;
    mesh.nf = nf;
    mesh.dgnodes.assign(dgnodes, dgnodes + (size_t)ne * npv * nd);
    mesh.elcon.assign(elcon, elcon + (size_t)ne * ndf);
    mesh.bf.assign(bf, bf + (size_t)ne * nfe);
    mesh.t2f.assign(t2f, t2f + (size_t)ne * nfe);
    sol.UDG.assign(UDG, UDG + (size_t)ne * npv * nc);
    sol.UH.assign(UH, UH + (size_t)nf * npf * nch);
    sol.SH.assign(SH, SH + (size_t)ne * npv * nc);
    sol.DinvRu.assign((size_t)ne * na, 0.0);
    sol.DinvF.assign((size_t)ne * na * nb, 0.0);
    sys.numEntities = nf;
    sys.blkSize = bsz;
    sys.ent2entStart.assign(rowstart, rowstart + nf + 1);
    sys.ent2ent.assign(cols, cols + nblocks);
    sys.Hg.assign((size_t)nblocks * bsz * bsz, 0.0);
    sys.Rg.assign((size_t)nf * bsz, 0.0);
    sys.r.assign((size_t)nf * bsz, 0.0);
    sys.x.assign((size_t)nf * bsz, 0.0);
    sys.matvecPrecision = 1;
    sys.quasiNewtonAccuracy = 1;
    sys.preconditioner = solver[0];
    sys.preconditionerSide = 0;
    sys.precPrecision = solver[1];
    sys.reorderMethod = solver[2];
    sys.orthogMethod = solver[3];
    sys.restart = solver[4];
    sys.maxiter = solver[5];
    sys.tol = tol;

    Int convFlag = -7;
    assembleLinearSystem(sys, nullptr, mesh, master, sol, app, nullptr);
    const double rn0 = sol.rNorm;
    computeOrdering(sys, sys.preconditioner);
    computePreconditioner(sys, sys.preconditioner);
    gmres(sys, &convFlag);
    computeResidualNorm(sys, nullptr, mesh, master, sol, app, nullptr);
    if (Hg) memcpy(Hg, &sys.Hg[0], sizeof(double) * sys.Hg.size());
    if (Rg) memcpy(Rg, &sys.Rg[0], sizeof(double) * sys.Rg.size());
    if (DinvRu) memcpy(DinvRu, &sol.DinvRu[0], sizeof(double) * sol.DinvRu.size());
    if (DinvF) memcpy(DinvF, &sol.DinvF[0], sizeof(double) * sol.DinvF.size());
    if (x) memcpy(x, &sys.x[0], sizeof(double) * (size_t)nf * bsz);
    if (o2u)
        for (int i = 0; i < nf; i++) o2u[i] = (int)sys.ordered2unordered[i];
    scal[0] = rn0;
    scal[1] = sol.rNorm;
    scal[2] = (double)convFlag;
    hfemDigasoRelease();
    return 0;
}
