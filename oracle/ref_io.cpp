// Test-only harness around the reference's own binary readers and writers
// (preprocessing/readbinaryfiles.cpp, preprocessing/ioutilities.cpp), compiled
// from the reference sources where they lie (ref.mk passes -I<reference root>);
// nothing of the reference is copied into this repository.  It lets the tests pin
// hfem_amd/binio.py against the reference itself:
//   refio_roundtrip   the reference reads a serial case (app, master, mesh and
//                     sol .bin, readappstruct / readmasterstruct / readmeshstruct
//                     / readsolstruct) and writes it back with its writers
//                     (writeappstruct / writemasterstruct / writemeshstruct /
//                     writesolstruct);
//   refio_read_rank   the reference reads rank `rank`'s record of a parallel
//                     mesh and sol file (app.nproc > 1) and writes the fields to
//                     flat files for comparison;
//   refio_write_sol   solstruct::writeSol2File writes a solution output.
#include <string>
#include <fstream>
#include <sstream>
#include <iostream>
#include <vector>
#include <algorithm>
#include <math.h>
#include <stdlib.h>
#include <string.h>
using namespace std;
#include "preprocessing/digasopre.h"
#include "preprocessing/errormsg.cpp"
#include "preprocessing/ioutilities.cpp"
#include "preprocessing/readbinaryfiles.cpp"

extern "C" int refio_roundtrip(const char *in_prefix, const char *out_prefix)
{
    const string in(in_prefix), out(out_prefix);
    appstruct app;
    readappstruct(in + "app.bin", app);
    app.my_rank = 0;
    const Int nmax = *max_element(app.elemtype.begin(), app.elemtype.end());
    vector<masterstruct> master(nmax + 1);
    readmasterstruct(in + "master.bin", master, app);
    meshstruct mesh;
    readmeshstruct(in + "mesh.bin", mesh, app);
    solstruct sol;
    readsolstruct(in + "sol.bin", sol, app);
    writeappstruct(out + "app.bin", app);
    writemasterstruct(out + "master.bin", master, app);
    writemeshstruct(out + "mesh.bin", mesh, app);
    writesolstruct(out + "sol.bin", sol, app);
    return 0;
}

static void dump_i(const string &path, const vector<Int> &a)
{
    ofstream o(path.c_str(), ios::out | ios::binary);
    for (Int v : a) {
        const double d = (double)v;
        o.write(reinterpret_cast<const char *>(&d), sizeof(double));
    }
}

static void dump_d(const string &path, const vector<double> &a)
{
    ofstream o(path.c_str(), ios::out | ios::binary);
    if (!a.empty()) o.write(reinterpret_cast<const char *>(&a[0]), sizeof(double) * a.size());
}

// Rank `rank` of a parallel case: the reference reads app.bin, then its record of
// mesh<k>.bin / sol<k>.bin (k = the file holding the rank), and the fields are
// written to out_prefix + "<field>.raw" as doubles.
extern "C" int refio_read_rank(const char *in_prefix, int rank, const char *out_prefix)
{
    const string in(in_prefix), out(out_prefix);
    appstruct app;
    readappstruct(in + "app.bin", app);
    if (app.nproc < 2 || rank < 0 || rank >= app.nproc) return -1;
    app.my_rank = rank;
    const Int nprocperfile = app.nproc / app.nfile, filenumber = rank / nprocperfile + 1;
    meshstruct mesh;
    readmeshstruct(in + "mesh" + NumberToString(filenumber) + ".bin", mesh, app);
    solstruct sol;
    readsolstruct(in + "sol" + NumberToString(filenumber) + ".bin", sol, app);
    dump_i(out + "mesh_ndims.raw", mesh.ndims);
    dump_i(out + "elementtype.raw", mesh.elementtype);
    dump_i(out + "t.raw", mesh.t);
    dump_d(out + "dgnodes.raw", mesh.dgnodes);
    dump_i(out + "elcon.raw", mesh.elcon);
    dump_i(out + "bf.raw", mesh.bf);
    dump_i(out + "t2f.raw", mesh.t2f);
    dump_i(out + "t2t.raw", mesh.t2t);
    dump_i(out + "nbsd.raw", mesh.nbsd);
    dump_i(out + "extintelem.raw", mesh.extintelem);
    dump_i(out + "elem2cpu.raw", mesh.elem2cpu);
    dump_i(out + "extintface.raw", mesh.extintface);
    dump_i(out + "face2cpu.raw", mesh.face2cpu);
    dump_i(out + "sol_ndims.raw", sol.ndims);
    dump_d(out + "UDG.raw", sol.UDG);
    dump_d(out + "UH.raw", sol.UH);
    return 0;
}

// solstruct::writeSol2File for UDG (ne, nc, npv per element) and UH (n values).
extern "C" int refio_write_sol(const char *path, int ne, int npv, int nc, int ncu, int n, const double *UDG,
                               const double *UH, int write_q)
{
    solstruct sol;
    sol.UDG.assign(UDG, UDG + (size_t)ne * npv * nc);
    sol.UH.assign(UH, UH + (size_t)n);
    sol.UDG2Write.assign((size_t)npv * nc, 0.0);
    vector<Int> ndims(80, 0);
    ndims[9] = npv;
    ndims[19] = nc;
    ndims[20] = ncu;
    sol.writeSol2File(string(path), ne, n, &ndims[0], write_q);
    return 0;
}
