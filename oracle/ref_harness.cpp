// Test-only harness around the reference's own element path.
//
// Compiled from the reference sources where they lie (the Makefile ref.mk passes
// -I<reference root>); nothing of the reference is copied into this repository.
// It fills the reference's structs (preprocessing/digasopre.h) from flat arrays and
// calls assembleElementMatrixVector (assembly/elementassembly.cpp:1481) and
// schurElementMatrixVector (assembly/elementschur.cpp:331) element by element,
// exporting the element blocks before and after condensation, and getQ
// (assembly/elementassembly.cpp:1414).  BLAS/LAPACK come from the image's MKL.
// Output goes to oracle/_ref/ only; it is used to write golden fixtures and as
// the reference-kind CPU baseline.
#include <string>
#include <fstream>
#include <sstream>
#include <iostream>
#include <vector>
#include <algorithm>
#include <numeric>
#include <math.h>
#include <stdlib.h>
#include <string.h>
#include <time.h>
#include <omp.h>
using namespace std;
#include "preprocessing/digasopre.h"
#include "preprocessing/errormsg.cpp"
#include "assembly/elementassembly.cpp"
#include "assembly/elementschur.cpp"

namespace {
struct Pool {
    vector<double*> blocks;
    double* get(size_t n) { double* p = new double[n](); blocks.push_back(p); return p; }
    ~Pool() { for (double* p : blocks) delete[] p; }
};

struct Setup {
    appstruct app; masterstruct master; meshstruct mesh; solstruct sol; elemstruct elem; tempstruct temp;
    Pool pool;
};

void build(Setup& S, const int* dims, const double* shapvt, const double* shapvg, const double* sgl,
           const double* shapft, const double* shapfg, const double* sfgl, const int* perm,
           const double* param, const double* ui, double fc_u, double fc_q, int tdep,
           int ne, int nuh, const double* dgnodes, const double* UDG, const double* UH, const double* SH,
           const int* elcon, const int* bf)
{
    int nd = dims[0], ncu = dims[1], npv = dims[2], ngv = dims[3], npf = dims[4], ngf = dims[5], nfe = dims[6];
    int nch = ncu, nc = ncu * (nd + 1), ndf = npf * nfe, nd1 = nd + 1;
    int na = npv * ncu, nb = ndf * nch;
    appstruct& app = S.app;
    app.nd = nd; app.ncd = nd; app.nc = nc; app.ncu = ncu; app.nch = nch; app.ncq = nc - ncu; app.nco = 0;
    app.flag_q = 1; app.tdep = tdep; app.fc_u = fc_u; app.fc_q = fc_q; app.time = 0.0;
    app.appname = 1; app.hybrid = 0; app.adjoint = 0; app.ALEflag = 0; app.AVflag = 0; app.SGSmodel = 0;
    app.viscosityModel = 0; app.convStabMethod = 0; app.diffStabMethod = 0; app.rotatingFrame = 0; app.my_rank = 0;
    app.physicsparam.assign(param, param + 6);
    app.bcm.assign(1, 1);
    app.bcs.assign(ui, ui + nch);

    masterstruct& m = S.master;
    m.nd = nd; m.dim = nd; m.npv = npv; m.npe = npv; m.ngv = ngv; m.nge = ngv; m.nfe = nfe; m.ndf = ndf;
    m.npf.assign(nfe, npf); m.ngf.assign(nfe, ngf);
    m.shapvt.assign(shapvt, shapvt + (size_t)ngv * npv * nd1);
    m.shapvg.assign(shapvg, shapvg + (size_t)ngv * npv * nd1);
    m.shapvgdotshapvl.assign(sgl, sgl + (size_t)npv * npv * ngv * nd1);
    m.shapft.assign(nfe, vector<double>(shapft, shapft + (size_t)ngf * npf * nd));
    m.shapfg.assign(nfe, vector<double>(shapfg, shapfg + (size_t)ngf * npf * nd));
    m.shapfgdotshapfc.assign(nfe, vector<double>(sfgl, sfgl + (size_t)npf * npf * ngf * nd));
    m.perm.resize(nfe); m.permgeom.resize(nfe);
    for (int a = 0; a < nfe; a++) {
        m.perm[a].assign(perm + a * npf, perm + (a + 1) * npf);
        m.permgeom[a] = m.perm[a];
    }

    meshstruct& mesh = S.mesh;
    mesh.ne = ne;
    mesh.dgnodes.assign(dgnodes, dgnodes + (size_t)ne * npv * nd);
    mesh.elcon.assign(elcon, elcon + (size_t)ne * ndf);
    mesh.bf.assign(bf, bf + (size_t)ne * nfe);
    mesh.elementtype.assign(ne, 0);
    mesh.isEDGelement.assign(ne, 0);
    mesh.nfes.assign(1, nfe); mesh.npes.assign(1, npv); mesh.ndf.assign(1, ndf); mesh.ncf.assign(1, 0);
    mesh.dg2cg.assign(1, vector<Int>(ndf, 0));

    solstruct& sol = S.sol;
    sol.UDG.assign(UDG, UDG + (size_t)ne * npv * nc);
    sol.UH.assign(UH, UH + (size_t)nuh * nch);
    sol.SH.assign(SH, SH + (size_t)ne * npv * nc);
    sol.ODG.assign(1, 0.0);

    size_t big = (size_t)4 * max({(size_t)nb * nb, (size_t)na * nb, (size_t)na * na,
                                  (size_t)npf * npf * nfe * nch * nc, (size_t)ngv * nd1 * ncu * nc,
                                  (size_t)ngv * nd1 * ncu * npv * ncu, (size_t)ngv * nd1 * ncu * ncu * ndf,
                                  (size_t)ngf * nfe * ncu * ncu * ndf, (size_t)ngf * nfe * ncu * npv * ncu,
                                  (size_t)npv * ndf * nd * nfe});
    elemstruct& e = S.elem;
    e.M = S.pool.get(big); e.E = S.pool.get(big); e.C = S.pool.get(big); e.Rq = S.pool.get(big);
    e.BD = S.pool.get(big); e.F = S.pool.get(big); e.Ru = S.pool.get(big); e.GK = S.pool.get(big);
    e.H = S.pool.get(big); e.Rh = S.pool.get(big); e.Rhonly = S.pool.get(big);
    e.D_LU = S.pool.get(big); e.D_LU_tmp = S.pool.get(big); e.DiF_ij = S.pool.get(big); e.Ru_i = S.pool.get(big);
    e.pivDii.assign(big, 0); e.workDii.assign(big, 0.0); e.D_inv.assign(1, 0.0); e.K.assign(1, 0.0);

    tempstruct& t = S.temp;
    double** ptrs[] = {&t.BDt, &t.BDtmp, &t.Etmp, &t.Ftmp, &t.Jg, &t.Jgf, &t.Rutmp, &t.Xxg, &t.f, &t.f_udg,
                       &t.fb, &t.fb_u, &t.fb_uh, &t.fh, &t.fh_u, &t.fh_uh, &t.jacg, &t.jacgf, &t.nlgf, &t.nlgjac,
                       &t.nlt, &t.odgg, &t.of, &t.oft, &t.ogf, &t.pf, &t.pft, &t.pg, &t.pgf, &t.s, &t.s_udg,
                       &t.udgg, &t.uf, &t.uft, &t.ugf, &t.uh, &t.uhg, &t.uht, &t.wrk, &t.wrl};
    for (double** p : ptrs) *p = S.pool.get(big);
    vector<double>* vecs[] = {&t.Df, &t.F_tmp, &t.Ff, &t.H_tmp, &t.K_tmp, &t.MiCf, &t.MiEf, &t.Rh_tmp,
                              &t.fhushapMiCf, &t.fhushapMiEf, &t.shapMiC, &t.shapMiCf, &t.shapMiE, &t.shapMiEf,
                              &t.wrlshapMiC, &t.wrlshapMiE};
    for (vector<double>* v : vecs) v->assign(big, 0.0);
    t.ipiv.assign(big, 0);
}
}  // namespace

extern "C" int ref_element(const int* dims, const double* shapvt, const double* shapvg, const double* sgl,
                           const double* shapft, const double* shapfg, const double* sfgl, const int* perm,
                           const double* param, const double* ui, double fc_u, double fc_q, int tdep,
                           int ne, int nuh, const double* dgnodes, const double* UDG, const double* UH,
                           const double* SH, const int* elcon, const int* bf, int e0, int n, int impl,
                           double* Ru, double* D, double* F, double* K, double* H, double* Rh,
                           double* DinvRu, double* DinvF, double* AH, double* FH)
{
    Setup* S = new Setup();
    build(*S, dims, shapvt, shapvg, sgl, shapft, shapfg, sfgl, perm, param, ui, fc_u, fc_q, tdep, ne, nuh,
          dgnodes, UDG, UH, SH, elcon, bf);
    int nd = dims[0], ncu = dims[1], npv = dims[2], npf = dims[4], nfe = dims[6];
    int na = npv * ncu, nb = npf * nfe * ncu;
    double elemtimes[32] = {0}, schurtimes[8] = {0};
    long long flops = 0;
    (void)nd;
    for (int t = 0; t < n; t++) {
        int ie = e0 + t;
        assembleElementMatrixVector(S->elem, S->mesh, S->master, S->app, S->sol, S->temp, ie, elemtimes);
        if (Ru) memcpy(Ru + (size_t)t * na, S->elem.Ru, sizeof(double) * na);
        if (D) memcpy(D + (size_t)t * na * na, S->elem.BD, sizeof(double) * na * na);
        if (F) memcpy(F + (size_t)t * na * nb, S->elem.F, sizeof(double) * na * nb);
        if (K) memcpy(K + (size_t)t * nb * na, S->elem.GK, sizeof(double) * nb * na);
        if (H) memcpy(H + (size_t)t * nb * nb, S->elem.H, sizeof(double) * nb * nb);
        if (Rh) memcpy(Rh + (size_t)t * nb, S->elem.Rh, sizeof(double) * nb);
        if (impl >= 0) {
            memcpy(S->elem.Rhonly, S->elem.Rh, sizeof(double) * nb);
            schurElementMatrixVector(S->elem, S->mesh, S->app, S->temp, impl, ie, schurtimes, &flops);
            if (DinvRu) memcpy(DinvRu + (size_t)t * na, S->elem.Ru, sizeof(double) * na);
            if (DinvF) memcpy(DinvF + (size_t)t * na * nb, S->elem.F, sizeof(double) * na * nb);
            if (AH) memcpy(AH + (size_t)t * nb * nb, S->elem.H, sizeof(double) * nb * nb);
            if (FH) memcpy(FH + (size_t)t * nb, S->elem.Rh, sizeof(double) * nb);
        }
    }
    delete S;
    return 0;
}

extern "C" int ref_getq(const int* dims, const double* shapvt, const double* shapvg, const double* sgl,
                        const double* shapft, const double* shapfg, const double* sfgl, const int* perm,
                        const double* param, const double* ui, double fc_q,
                        int ne, int nuh, const double* dgnodes, double* UDG, const double* UH,
                        const double* SH, const int* elcon, const int* bf)
{
    Setup* S = new Setup();
    build(*S, dims, shapvt, shapvg, sgl, shapft, shapfg, sfgl, perm, param, ui, 0.0, fc_q, 0, ne, nuh,
          dgnodes, UDG, UH, SH, elcon, bf);
    int nd = dims[0], ncu = dims[1], npv = dims[2];
    int nc = ncu * (nd + 1);
    for (int ie = 0; ie < ne; ie++) {
        double* u = &S->sol.UDG[(size_t)ie * npv * nc];
        getQ(S->elem, S->mesh, S->master, S->app, S->sol, S->temp, u, u + npv * ncu, &S->sol.UH[0],
             &S->sol.SH[(size_t)ie * npv * nc + npv * ncu], ie);
    }
    memcpy(UDG, &S->sol.UDG[0], sizeof(double) * ne * npv * nc);
    delete S;
    return 0;
}

// MKL's per-thread thread count (mkl_rt): the timed loops keep MKL sequential
// inside each of their threads whatever MKL_NUM_THREADS the environment sets.
extern "C" int MKL_Set_Num_Threads_Local(int);   // (mkl.h maps mkl_set_num_threads_local to it)

// Timed element loop on nthreads host threads (the all-cores CPU baseline): each
// thread has its own structs, as the reference's per-thread elems[] / temps[]
// intend (main/digasoser.cpp:51-58), and takes every nthreads-th element of
// [0, n) (repeating over the first ne); MKL runs sequentially inside the threads.
// Returns the wall-clock seconds of the whole loop.
extern "C" double ref_time_mt(const int* dims, const double* shapvt, const double* shapvg, const double* sgl,
                              const double* shapft, const double* shapfg, const double* sfgl, const int* perm,
                              const double* param, const double* ui, int ne, int nuh, const double* dgnodes,
                              const double* UDG, const double* UH, const double* SH, const int* elcon,
                              const int* bf, int n, int impl, int nthreads)
{
    vector<Setup*> S(nthreads);
    for (int th = 0; th < nthreads; th++) {
        S[th] = new Setup();
        build(*S[th], dims, shapvt, shapvg, sgl, shapft, shapfg, sfgl, perm, param, ui, 0.0, 1.0, 0, ne, nuh,
              dgnodes, UDG, UH, SH, elcon, bf);
    }
    int nb = dims[4] * dims[6] * dims[1];
    struct timespec t0, t1;
    clock_gettime(CLOCK_MONOTONIC, &t0);
#pragma omp parallel num_threads(nthreads)
    {
        const int th = omp_get_thread_num();
        MKL_Set_Num_Threads_Local(1);
        Setup* s = S[th];
        double elemtimes[32] = {0}, schurtimes[8] = {0};
        long long flops = 0;
        for (int t = th; t < n; t += nthreads) {
            int ie = t % ne;
            assembleElementMatrixVector(s->elem, s->mesh, s->master, s->app, s->sol, s->temp, ie, elemtimes);
            memcpy(s->elem.Rhonly, s->elem.Rh, sizeof(double) * nb);
            schurElementMatrixVector(s->elem, s->mesh, s->app, s->temp, impl, ie, schurtimes, &flops);
        }
    }
    clock_gettime(CLOCK_MONOTONIC, &t1);
    for (Setup* s : S) delete s;
    return (t1.tv_sec - t0.tv_sec) + 1e-9 * (t1.tv_nsec - t0.tv_nsec);
}

// Timed element loop (CPU baseline): assemble + condense elements [0, n) of the
// given arrays, repeating over the first ne elements; returns seconds.
extern "C" double ref_time(const int* dims, const double* shapvt, const double* shapvg, const double* sgl,
                           const double* shapft, const double* shapfg, const double* sfgl, const int* perm,
                           const double* param, const double* ui, int ne, int nuh, const double* dgnodes,
                           const double* UDG, const double* UH, const double* SH, const int* elcon,
                           const int* bf, int n, int impl)
{
    Setup* S = new Setup();
    build(*S, dims, shapvt, shapvg, sgl, shapft, shapfg, sfgl, perm, param, ui, 0.0, 1.0, 0, ne, nuh,
          dgnodes, UDG, UH, SH, elcon, bf);
    double elemtimes[32] = {0}, schurtimes[8] = {0};
    long long flops = 0;
    int nb = dims[4] * dims[6] * dims[1];
    const int prev = MKL_Set_Num_Threads_Local(1);
    struct timespec t0, t1;
    clock_gettime(CLOCK_MONOTONIC, &t0);
    for (int t = 0; t < n; t++) {
        int ie = t % ne;
        assembleElementMatrixVector(S->elem, S->mesh, S->master, S->app, S->sol, S->temp, ie, elemtimes);
        memcpy(S->elem.Rhonly, S->elem.Rh, sizeof(double) * nb);
        schurElementMatrixVector(S->elem, S->mesh, S->app, S->temp, impl, ie, schurtimes, &flops);
    }
    clock_gettime(CLOCK_MONOTONIC, &t1);
    MKL_Set_Num_Threads_Local(prev);
    delete S;
    return (t1.tv_sec - t0.tv_sec) + 1e-9 * (t1.tv_nsec - t0.tv_nsec);
}
