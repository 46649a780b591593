// cntc_api.cpp -- the cntc_* C ABI (module-3 subset) over the HIP engine.
//
// Problem store, unit conversion and argument checks follow src/contact_addon.f90; every
// numerical step runs on the GPU through cntc::Engine.  Calls are serialised by one mutex
// (the engine owns a single stream); problems of different (ire, icp) keep separate state.
#include <hip/hip_runtime.h>

#include <chrono>
#include <cmath>
#include <cstdio>
#include <cstring>
#include <map>
#include <memory>
#include <mutex>
#include <stdexcept>
#include <string>
#include <vector>

#include "../../include/cntc_amd.h"
#include "engine.hpp"
#include "inp_writer.hpp"

namespace {

using cntc::CaseInput;
using cntc::Engine;
using cntc::Patch;

// unit conventions, src/m_global_data.f90:463-526
struct Scaling {
    double len = 1, area = 1, forc = 1, veloc = 1, angle = 1, body = 1;
};

Scaling select_units(int iunits)
{
    Scaling s;
    if (iunits == CNTC_un_spck) { s.len = 1e3; s.area = 1e6; s.veloc = 1e3; s.body = -1.0; }
    else if (iunits == CNTC_un_si) { s.len = 1e3; s.area = 1e6; s.veloc = 1e3; }
    return s;
}

struct Problem {
    Patch patch;
    Scaling scl;
    std::map<int, int> flags;        // control digits without effect on this path (getflags)
    std::map<int, double> meta;      // cntc_setmetadata values, internal units
    int ncase = 0;                   // cases calculated (CNTC_if_ncase)
    double fkin = 0.30;              // kinetic friction of L=0 (getparameters)
    double tcpu = 0, twall = 0;
    bool calculated = false;
    int ticket = 0;          // request of the serving loop in flight for this problem
    int last_rc = 0;
};

struct Store {
    std::mutex mtx;
    std::unique_ptr<Engine> engine;
    std::map<std::pair<int, int>, std::unique_ptr<Problem>> probs;
    std::map<int, Problem*> tickets;
    std::map<int, int> ire_module;
    int precision = 0;
    int idebug = 0;
    // output folder and experiment name of the first initialise call, and the .inp file
    // (src/contact_addon.f90:289-318, src/m_global_data.f90:41)
    bool names_set = false;
    std::string outdir, expnam = "contact_addon";
    std::FILE* inp = nullptr;
    int inp_open = 0;                // 1 open, -1 could not be opened
    ~Store()
    {
        if (inp) std::fclose(inp);
    }
};

Store& store()
{
    static Store s;
    return s;
}

const int MAX_REID = 999, MAX_CPID = 9;

// the folder names of the first initialise call (copy_c_dirnam and the experiment name,
// src/contact_addon.f90:289-318, 576-601): a separator is appended to a folder name, and a
// directory part of the experiment name is dropped
void set_names(const char* c_outdir, int len_outdir, const char* c_expnam, int len_expnam)
{
    Store& S = store();
    if (S.names_set) return;
    S.names_set = true;
    auto cstr = [](const char* c, int len) {
        std::string r;
        for (int i = 0; c && i < len && c[i]; i++) r += c[i];
        while (!r.empty() && r.back() == ' ') r.pop_back();
        return r;
    };
    S.outdir = cstr(c_outdir, len_outdir);
    if (S.outdir.size() > 1 && S.outdir.back() != '/') S.outdir += '/';
    const std::string e = cstr(c_expnam, len_expnam);
    if (!e.empty() && e[0] != ' ') {
        const size_t k = e.find_last_of('/');
        S.expnam = k == std::string::npos ? e : e.substr(k + 1);
    }
}

// wait for the in-flight request of P (serving loop) and record its outcome
void settle(Problem* P)
{
    Store& S = store();
    if (!P || P->ticket <= 0 || !S.engine) return;
    int rc = 0;
    double lat = 0.0;
    if (S.engine->poll(P->ticket, &rc, &lat, true) == 1) {
        P->last_rc = rc;
        P->twall += lat;
        P->calculated = true;
        P->ncase++;
        S.engine->forget(P->ticket);
    }
    S.tickets.erase(P->ticket);
    P->ticket = 0;
}

// cntc_activate, src/m_caddon_data.f90:89-225: auto-initialise, check ids.  A problem whose
// solve is in flight in the serving loop is settled first: the first set-call after a
// calculate starts a new case.
Problem* activate(int ire, int icp)
{
    Store& S = store();
    if (ire <= 0 || ire > MAX_REID || icp <= 0 || icp > MAX_CPID) return nullptr;
    auto m = S.ire_module.find(ire);
    if (m == S.ire_module.end()) S.ire_module[ire] = 3;
    else if (m->second != 3) return nullptr;
    auto key = std::make_pair(ire, icp);
    auto it = S.probs.find(key);
    if (it == S.probs.end()) {
        std::unique_ptr<Problem> p(new Problem());
        p->patch.in.precision = S.precision;
        it = S.probs.emplace(key, std::move(p)).first;
    }
    settle(it->second.get());
    return it->second.get();
}

Engine* engine()
{
    Store& S = store();
    if (!S.engine) {
        int ndev = 0;
        if (hipGetDeviceCount(&ndev) != hipSuccess || ndev <= 0) return nullptr;
        S.engine.reset(new Engine());
    }
    return S.engine.get();
}

}  // namespace

extern "C" {

void cntc_initializefirst(int* ifcver, int* ierror, int* ioutput, const char* c_wrkdir, const char* c_outdir,
                          const char* c_expnam, int* len_wrkdir, int* len_outdir, int* len_expnam)
{
    (void)ioutput;
    (void)c_wrkdir;
    (void)len_wrkdir;
    std::lock_guard<std::mutex> g(store().mtx);
    set_names(c_outdir, len_outdir ? *len_outdir : 0, c_expnam, len_expnam ? *len_expnam : 0);
    if (ifcver) *ifcver = 2400;
    if (ierror) *ierror = 0;
}

void cntc_initialize(int* ire, int* imodul, int* ifcver, int* ierror, const char* c_outdir, int* len_outdir)
{
    std::lock_guard<std::mutex> g(store().mtx);
    set_names(c_outdir, len_outdir ? *len_outdir : 0, nullptr, 0);
    if (ifcver) *ifcver = 2400;
    if (!ire || !imodul || *ire <= 0 || *ire > MAX_REID || *imodul != 3) {
        if (ierror) *ierror = CNTC_err_other;
        return;
    }
    store().ire_module[*ire] = 3;
    // set-up: the engine (stream, staging buffers, the no-allocation array's address) is made
    // here when a GPU is present, so that later calls need not allocate it
    (void)engine();
    if (ierror) *ierror = 0;
}

// src/m_caddon_data.f90:310-332: calls are serialised by one lock, so one thread at a time;
// a caller solves many problems concurrently through cntc_calculatebatch instead
void cntc_getmaxnumthreads(int* mxthrd)
{
    if (mxthrd) *mxthrd = 1;
}

void cntc_setglobalflags(int* lenflg, int* params, int* values)
{
    std::lock_guard<std::mutex> g(store().mtx);
    for (int i = 0; i < *lenflg; i++) {
        if (params[i] == CNTC_if_idebug) store().idebug = values[i];
        if (params[i] == CNTC_if_precision) store().precision = values[i] ? 1 : 0;
    }
}

// src/contact_addon.f90:780-973
void cntc_setflags(int* ire, int* icp, int* lenflg, int* params, int* values)
{
    std::lock_guard<std::mutex> g(store().mtx);
    Problem* P = activate(*ire, *icp);
    if (!P) return;
    CaseInput& in = P->patch.in;
    for (int i = 0; i < *lenflg; i++) {
        int v = values[i];
        switch (params[i]) {
        case CNTC_if_units: P->scl = select_units(v); break;
        case CNTC_ic_pvtime: in.pvtime = std::max(0, std::min(3, v)); break;
        case CNTC_ic_tang: in.tang = std::max(0, std::min(3, v)); break;
        case CNTC_ic_norm: in.norm = std::max(0, std::min(1, v)); break;
        case CNTC_ic_iestim: in.iestim = std::max(0, std::min(3, v)); break;
        case CNTC_if_precision: in.precision = v ? 1 : 0; break;
        case CNTC_ic_sens: in.sens = std::max(0, std::min(3, v)); break;   // :844-846
        case CNTC_ic_force: in.force3 = std::max(0, std::min(2, v)); break;  // :834-842, module 3
        case CNTC_ic_exrhs: in.rztang = std::max(0, std::min(3, v)); break;   // :883-885
        case CNTC_ic_config: P->flags[params[i]] = std::max(0, std::min(5, v)); break;
        case CNTC_ic_bound: in.bound = std::max(0, std::min(6, v)); break;
        case CNTC_ic_discns: if (v >= 0 && v <= 9) P->flags[params[i]] = v; break;
        case CNTC_ic_inflcf: if (v >= 0 && v <= 4) P->flags[params[i]] = v; break;
        case CNTC_ic_matfil: case CNTC_ic_sbsfil: P->flags[params[i]] = (v < 0 || v > 2) ? 0 : v; break;
        case CNTC_ic_output: P->flags[params[i]] = std::max(0, std::min(5, v)); break;
        case CNTC_ic_sbsout: P->flags[params[i]] = std::max(0, std::min(4, v)) == 3 ? 2 : std::max(0, std::min(4, v)); break;
        case CNTC_ic_flow: P->flags[params[i]] = std::max(0, std::min(9, v)); break;
        case CNTC_ic_return: P->flags[params[i]] = std::max(0, std::min(3, v)); break;
        case CNTC_ic_xflow: P->flags[params[i]] = 1; break;
        case CNTC_if_wrtinp: case CNTC_ic_ifmeth: case CNTC_ic_ifvari: case CNTC_if_ncase:
            P->flags[params[i]] = v;
            break;
        default: break;   // output, flow, file flags: no effect on the solution
        }
    }
}

// src/contact_addon.f90:1128-1283
void cntc_setsolverflags(int* ire, int* icp, int* gdigit, int* nints, int* iparam, int* nreals, double* rparam)
{
    std::lock_guard<std::mutex> g(store().mtx);
    Problem* P = activate(*ire, *icp);
    if (!P) return;
    static const int nints_loc[7] = {4, 0, 5, 5, 5, 5, 1}, nreals_loc[7] = {1, 0, 4, 4, 2, 8, 1};
    int gd = *gdigit;
    if (gd < 0 || gd > 6) return;
    if (*nints != nints_loc[gd] || *nreals != nreals_loc[gd]) return;
    CaseInput& in = P->patch.in;
    if (gd == 6) {
        // parameters of the sensitivity solves, src/contact_addon.f90:1266-1272
        in.mxsens = std::max(1, iparam[0]);
        in.epsens = std::max(1e-20, rparam[0]);
    }
    if (gd != 1 && gd != 6) {
        in.gausei = gd;
        in.eps = std::max(1e-20, rparam[0]);
        in.maxgs = std::max(1, iparam[0]);
        in.maxin = std::max(1, iparam[1]);
        in.maxnr = std::max(1, iparam[2]);
        in.maxout = std::max(1, iparam[3]);
        if (gd == 2 || gd == 3) {
            in.omegah = std::max(1e-20, rparam[1]);
            in.omegas = std::max(1e-20, rparam[2]);
        }
        if (gd == 5) {
            in.fdecay = rparam[1];
            in.betath = rparam[2];
            in.kdowfb = iparam[4];
            in.d_ifc = std::max(0.01, rparam[3]);
            in.d_lin = rparam[4];
            in.d_cns = std::max(0.01, rparam[5]);
            in.d_slp = std::max(0.01, rparam[6]);
            in.pow_s = std::max(0.01, std::min(10.0, rparam[7]));
            if (in.d_lin * (in.d_cns - in.d_ifc) < 0.0) {
                in.d_lin = 0.0;
                in.d_cns = in.d_ifc;
            }
        }
    }
}

// src/contact_addon.f90:1287-1473
void cntc_setmaterialparameters(int* ire, int* icp, int* mdigit, int* nparam, double* rparam)
{
    std::lock_guard<std::mutex> g(store().mtx);
    Problem* P = activate(*ire, *icp);
    if (!P || *mdigit != 0 || *nparam != 4) return;
    CaseInput& in = P->patch.in;
    in.poiss1 = rparam[0];
    in.poiss2 = rparam[1];
    in.gg1 = rparam[2] * P->scl.forc / P->scl.area;
    in.gg2 = rparam[3] * P->scl.forc / P->scl.area;
    // MAXOUT for quasi-identity or not (src/contact_addon.f90:1459-1462): the Panagiotopoulos
    // alternation runs up to 10 times for dissimilar materials, once otherwise
    const double ga = 2.0 / (1.0 / in.gg1 + 1.0 / in.gg2);
    const double ak = (ga / 4.0) * ((1.0 - 2.0 * in.poiss1) / in.gg1 - (1.0 - 2.0 * in.poiss2) / in.gg2);
    if (in.maxout >= 2 && std::fabs(ak) < 1e-4) in.maxout = 1;
    if (in.maxout <= 1 && std::fabs(ak) >= 1e-4) in.maxout = 10;
}

void cntc_settimestep(int* ire, int* icp, double* dt)
{
    std::lock_guard<std::mutex> g(store().mtx);
    (void)activate(*ire, *icp);
    (void)dt;
}

void cntc_setreferencevelocity(int* ire, int* icp, double* veloc)
{
    std::lock_guard<std::mutex> g(store().mtx);
    Problem* P = activate(*ire, *icp);
    if (!P) return;
    P->patch.in.veloc = std::max(1e-9, std::fabs(*veloc) * P->scl.veloc);
}

void cntc_setrollingstepsize(int* ire, int* icp, double* chi, double* dq)
{
    std::lock_guard<std::mutex> g(store().mtx);
    Problem* P = activate(*ire, *icp);
    if (!P) return;
    P->patch.in.chi = *chi * P->scl.angle;
    P->patch.in.dq = *dq * P->scl.len;
}

// src/contact_addon.f90:1671+ (L=0: fstat, fkin)
void cntc_setfrictionmethod(int* ire, int* icp, int* imeth, int* nparam, double* params)
{
    std::lock_guard<std::mutex> g(store().mtx);
    Problem* P = activate(*ire, *icp);
    if (!P || *imeth != 0 || *nparam != 2) return;
    P->patch.in.fstat = params[0];
    P->fkin = params[1];
}

// src/contact_addon.f90:1972-2093.  The method is checked before the parameter count, as the
// reference does (-6 takes 6 parameters, the others 5).  The rectangular options (-4, -5) set
// the B digit to 2 (prescribed pressure).  The SDEC option (-6) is not implemented: the problem
// is marked and cntc_calculate returns CNTC_err_other for it instead of solving the previous
// geometry.
void cntc_sethertzcontact(int* ire, int* icp, int* ipotcn, int* nparam, double* params)
{
    std::lock_guard<std::mutex> g(store().mtx);
    Problem* P = activate(*ire, *icp);
    if (!P) return;
    static const int nparam_loc[6] = {6, 5, 5, 5, 5, 5};   // ipotcn -6 .. -1
    if (*ipotcn < -6 || *ipotcn > -1) {
        if (store().idebug > 0) std::fprintf(stderr, "cntc_sethertzcontact: method IPOTCN=%d does not exist\n", *ipotcn);
        return;
    }
    if (*nparam != nparam_loc[*ipotcn + 6]) {
        if (store().idebug > 0)
            std::fprintf(stderr, "cntc_sethertzcontact: method IPOTCN=%d needs %d parameters\n", *ipotcn,
                         nparam_loc[*ipotcn + 6]);
        return;
    }
    CaseInput& in = P->patch.in;
    double L = P->scl.len;
    in.unsupported = nullptr;
    in.mx = std::max(1, (int)std::lround(params[0]));
    in.my = std::max(1, (int)std::lround(params[1]));
    if (*ipotcn == -3) {
        in.hz_aa = std::max(1e-6, params[2]) * L;
        in.hz_bb = std::max(1e-6, params[3]) * L;
        in.hz_scale = std::max(1e-6, params[4]);
    } else if (*ipotcn == -2) {
        in.hz_a1 = std::max(1e-12, params[2]) / L;
        in.hz_aob = std::max(1e-6, params[3]);
        in.hz_scale = std::max(1e-6, params[4]);
    } else if (*ipotcn == -1) {
        in.hz_a1 = std::max(1e-12, params[2]) / L;
        in.hz_b1 = std::max(1e-12, params[3]) / L;
        in.hz_scale = std::max(1e-6, params[4]);
    } else {
        // stored for the getters as the reference stores them (:2024-2049)
        if (*ipotcn == -6) {
            in.hz_aa = std::max(1e-6, params[2]) * L;
            in.hz_bneg = std::max(1e-6, params[3]) * L;
            in.hz_bpos = std::max(1e-6, params[4]) * L;
            in.hz_scale = std::max(1e-6, params[5]);
            in.unsupported = "Hertzian option IPOTCN=-6 (SDEC) is not supported";
        } else {
            // rectangular (line) contact: semi-elliptical pressure in x, prescribed (B=2)
            if (*ipotcn == -5) in.hz_aa = std::max(1e-6, params[2]) * L;
            else in.hz_a1 = std::max(1e-6, params[2]) / L;
            in.hz_bb = std::max(1e-6, params[3]) * L;
            in.hz_scale = std::max(1e-6, params[4]);
            in.bound = 2;
        }
    }
    in.ipotcn = *ipotcn;
}

// src/contact_addon.f90:2097-2257
void cntc_setpotcontact(int* ire, int* icp, int* ipotcn, int* nparam, double* params)
{
    std::lock_guard<std::mutex> g(store().mtx);
    Problem* P = activate(*ire, *icp);
    if (!P || *ipotcn < 1 || *ipotcn > 4 || *nparam != 6) return;
    CaseInput& in = P->patch.in;
    double L = P->scl.len;
    in.unsupported = nullptr;
    in.mx = std::max(1, (int)std::lround(params[0]));
    in.my = std::max(1, (int)std::lround(params[1]));
    if (*ipotcn == 1) {
        in.xl = params[2] * L; in.yl = params[3] * L;
        in.dx = std::max(1e-12, params[4]) * L; in.dy = std::max(1e-12, params[5]) * L;
    } else if (*ipotcn == 2) {
        in.xl = params[2] * L; in.yl = params[3] * L; in.xh = params[4] * L; in.yh = params[5] * L;
    } else if (*ipotcn == 3) {
        in.xc1 = params[2] * L; in.yc1 = params[3] * L;
        in.dx = std::max(1e-12, params[4]) * L; in.dy = std::max(1e-12, params[5]) * L;
    } else {
        in.xc1 = params[2] * L; in.yc1 = params[3] * L; in.xcm = params[4] * L; in.ycm = params[5] * L;
    }
    in.ipotcn = *ipotcn;
}

void cntc_setpenetration(int* ire, int* icp, double* pen)
{
    std::lock_guard<std::mutex> g(store().mtx);
    Problem* P = activate(*ire, *icp);
    if (!P) return;
    P->patch.in.norm = 0;
    P->patch.in.pen = *pen * P->scl.len;
}

void cntc_setnormalforce(int* ire, int* icp, double* fn)
{
    std::lock_guard<std::mutex> g(store().mtx);
    Problem* P = activate(*ire, *icp);
    if (!P) return;
    P->patch.in.norm = 1;
    P->patch.in.fn = *fn;
}

// src/contact_addon.f90:2361-2487
void cntc_setundeformeddistc(int* ire, int* icp, int* ibase, int* nparam, double* prmudf)
{
    std::lock_guard<std::mutex> g(store().mtx);
    Problem* P = activate(*ire, *icp);
    if (!P) return;
    CaseInput& in = P->patch.in;
    double L = P->scl.len;
    if (*ibase == 1) {
        if (*nparam != 6) return;
        for (int k = 0; k < 3; k++) in.prmudf[k] = prmudf[k] / L;
        in.prmudf[3] = prmudf[3];
        in.prmudf[4] = prmudf[4];
        in.prmudf[5] = prmudf[5] * L;
    } else if (*ibase == 9) {
        if (*nparam != in.mx * in.my) return;
        in.h_elem.assign(prmudf, prmudf + *nparam);
        for (auto& v : in.h_elem) v *= L;
    } else {
        return;
    }
    in.ibase = *ibase;
}

// src/contact_addon.f90:2491-2530
// src/contact_addon.f90:2597-2650: Fx, Fy as fractions of mu*Fn, scaled to |(Fx, Fy)| <= 1; used
// when the F digit is 1 (Fx) or 2 (Fx and Fy)
void cntc_settangentialforces(int* ire, int* icp, double* fx, double* fy)
{
    std::lock_guard<std::mutex> g(store().mtx);
    Problem* P = activate(*ire, *icp);
    if (!P) return;
    CaseInput& in = P->patch.in;
    in.fxrel = *fx;
    in.fyrel = *fy;
    double fabs = std::sqrt(*fx * *fx + *fy * *fy);
    if (fabs > 1.0) {
        in.fxrel = in.fxrel / fabs;
        in.fyrel = in.fyrel / fabs;
    }
}

void cntc_setcreepages(int* ire, int* icp, double* vx, double* vy, double* phi)
{
    std::lock_guard<std::mutex> g(store().mtx);
    Problem* P = activate(*ire, *icp);
    if (!P) return;
    CaseInput& in = P->patch.in;
    if (in.tang == 1) {
        in.cksi = *vx * P->scl.len;
        in.ceta = *vy * P->scl.len;
        in.cphi = *phi;
    } else {
        in.cksi = *vx;
        in.ceta = *vy;
        in.cphi = *phi / P->scl.len;
    }
}

// src/contact_addon.f90:2534-2593: E digit 1 and the extra term of the tangential right-hand
// side, sized for the current potential contact (elements beyond lenarr get 0)
void cntc_setextrarigidslip(int* ire, int* icp, int* lenarr, double* wx, double* wy)
{
    std::lock_guard<std::mutex> g(store().mtx);
    Problem* P = activate(*ire, *icp);
    if (!P) return;
    CaseInput& in = P->patch.in;
    const int npot = std::max(1, in.mx) * std::max(1, in.my);
    in.rztang = 1;
    in.exrhs.assign(2 * (size_t)npot, 0.0);
    for (int i = 0; i < std::min(*lenarr, npot); i++) {
        in.exrhs[i] = wx[i];
        in.exrhs[npot + i] = wy[i];
    }
}

// src/contact_addon.f90:1055-1124: metadata are kept with the problem (lengths in mm); they do
// not enter the solution.  Code 0 marks an unused position; unknown codes are ignored.
void cntc_setmetadata(int* ire, int* icp, int* lenmta, int* params, double* values)
{
    std::lock_guard<std::mutex> g(store().mtx);
    Problem* P = activate(*ire, *icp);
    if (!P) return;
    for (int i = 0; i < *lenmta; i++) {
        int c = params[i];
        double v = values[i];
        if (c == 0) continue;
        if (c == CNTC_mt_sw || c == CNTC_mt_xr || c == CNTC_mt_yr || c == CNTC_mt_xw || c == CNTC_mt_yw)
            P->meta[c] = P->scl.len * v;
        else if (c == CNTC_mt_tim)
            P->meta[c] = v;
        else if (c == CNTC_mt_run || c == CNTC_mt_axle || c == CNTC_mt_side)
            P->meta[c] = (double)std::lround(v);
        else if (store().idebug > 0)
            std::fprintf(stderr, "cntc_setmetadata: unknown metadata code=%d is ignored\n", c);
    }
}

// With CNTC_if_wrtinp >= 1: the case description in the .inp file, opened on first use in the
// output folder (src/contact_addon.f90:3829-3867)
static void write_inp(int ire, int icp, Problem* P)
{
    Store& S = store();
    auto f = P->flags.find(CNTC_if_wrtinp);
    if (f == P->flags.end() || f->second < 1 || S.inp_open < 0) return;
    if (!S.inp) {
        const std::string fname = S.outdir + S.expnam + ".inp";
        S.inp = std::fopen(fname.c_str(), "w");
        if (!S.inp) {
            S.inp_open = -1;
            if (S.idebug >= 1) std::fprintf(stderr, "cntc_calculate: could not open \"%s\" for writing\n", fname.c_str());
            return;
        }
        S.inp_open = 1;
    }
    std::vector<int8_t> el;
    Engine* E = S.engine.get();
    const Patch& pt = P->patch;
    if (E && P->calculated && pt.have_solution && pt.n > 0) {
        std::vector<int> e((size_t)pt.n);
        E->get_eldiv(pt, e.data());
        el.assign(e.begin(), e.end());
    }
    cntc::InpDigits dg;
    auto flag = [&](int code, int& v) {
        auto it = P->flags.find(code);
        if (it != P->flags.end()) v = it->second;
    };
    flag(CNTC_ic_matfil, dg.matfil);
    flag(CNTC_ic_output, dg.output);
    flag(CNTC_ic_flow, dg.flow);
    flag(CNTC_ic_return, dg.ret);
    cntc::write_inp_case(S.inp, ire, icp, P->ncase, pt, dg, P->fkin, el);
    std::fflush(S.inp);
}

static void calculate_many(int n, const int* ire, const int* icp, int* ierror)
{
    Store& S = store();
    Engine* E = engine();
    std::vector<Patch*> ps;
    std::vector<Problem*> pr;
    for (int k = 0; k < n; k++) {
        Problem* P = activate(ire[k], icp[k]);
        if (!P) { ierror[k] = CNTC_err_icp; continue; }
        if (!E) {
            ierror[k] = CNTC_err_other;
            P->ncase++;
            write_inp(ire[k], icp[k], P);
            continue;
        }
        ps.push_back(&P->patch);
        pr.push_back(P);
    }
    if (!E || ps.empty()) return;
    if (E->serving()) {
        // the serving loop owns the stream: submit every problem and wait for each
        for (Problem* P : pr) {
            P->ticket = E->submit(&P->patch);
            S.tickets[P->ticket] = P;
        }
        size_t j = 0;
        for (int k = 0; k < n; k++) {
            Problem* P = activate(ire[k], icp[k]);   // settles the request
            if (!P || j >= pr.size() || P != pr[j]) continue;
            ierror[k] = P->last_rc;
            j++;
        }
        return;
    }
    auto t0 = std::chrono::steady_clock::now();
    std::clock_t c0 = std::clock();
    try {
        E->solve(ps);
    } catch (const std::exception& ex) {
        if (S.idebug > 0) std::fprintf(stderr, "cntc_calculate: %s\n", ex.what());
        for (auto* p : ps) p->out.ierror = CNTC_err_other;
    }
    double tw = std::chrono::duration<double>(std::chrono::steady_clock::now() - t0).count();
    double tc = double(std::clock() - c0) / CLOCKS_PER_SEC;
    size_t j = 0;
    for (int k = 0; k < n; k++) {
        Problem* P = activate(ire[k], icp[k]);
        if (!P || j >= pr.size() || P != pr[j]) continue;
        ierror[k] = P->patch.out.ierror;
        P->calculated = true;
        P->ncase++;
        P->twall += P->patch.out.t_done > 0 ? P->patch.out.t_done : tw;
        P->tcpu += tc / pr.size();
        write_inp(ire[k], icp[k], P);
        j++;
    }
    if (S.idebug > 0 && !E->last_error.empty()) std::fprintf(stderr, "cntc_calculate: %s\n", E->last_error.c_str());
}

void cntc_calculate(int* ire, int* icp, int* ierror)
{
    std::lock_guard<std::mutex> g(store().mtx);
    int e = 0;
    calculate_many(1, ire, icp, &e);
    *ierror = e;
}

// No counterpart in the reference: cntc_calculate without device allocation, for callers that
// cannot allocate (code under graph capture, tight serving loops).  Every device buffer of the
// solve comes from the engine's fixed array (64 MiB per GPU); the solution is then kept in
// host memory with the problem.  Needs the engine made by an earlier call (cntc_initialize);
// rejects patches whose buffers do not fit, and calls while the serving loop runs.
void cntc_calculatenoalloc(int* ire, int* icp, int* ierror)
{
    std::lock_guard<std::mutex> g(store().mtx);
    Store& S = store();
    *ierror = 0;
    Problem* P = activate(*ire, *icp);
    if (!P) { *ierror = CNTC_err_icp; return; }
    Engine* E = S.engine.get();
    if (!E) {
        if (S.idebug > 0) std::fprintf(stderr, "cntc_calculatenoalloc: call cntc_initialize first\n");
        *ierror = CNTC_err_other;
        return;
    }
    auto t0 = std::chrono::steady_clock::now();
    std::clock_t c0 = std::clock();
    int rc = 0;
    try {
        rc = E->solve_noalloc(&P->patch);
    } catch (const std::exception& ex) {
        if (S.idebug > 0) std::fprintf(stderr, "cntc_calculatenoalloc: %s\n", ex.what());
        P->patch.out.ierror = CNTC_err_other;
    }
    if (rc < 0) {
        if (S.idebug > 0) std::fprintf(stderr, "cntc_calculatenoalloc: %s\n", E->last_error.c_str());
        *ierror = CNTC_err_other;
        return;
    }
    *ierror = P->patch.out.ierror;
    P->calculated = true;
    P->ncase++;
    P->twall += std::chrono::duration<double>(std::chrono::steady_clock::now() - t0).count();
    P->tcpu += double(std::clock() - c0) / CLOCKS_PER_SEC;
    write_inp(*ire, *icp, P);
    if (S.idebug > 0 && !E->last_error.empty()) std::fprintf(stderr, "cntc_calculatenoalloc: %s\n", E->last_error.c_str());
}

void cntc_calculatebatch(int* nprob, const int* ire, const int* icp, int* ierror)
{
    std::lock_guard<std::mutex> g(store().mtx);
    for (int k = 0; k < *nprob; k++) ierror[k] = 0;
    calculate_many(*nprob, ire, icp, ierror);
}

// ---- getters -----------------------------------------------------------------------------

static Problem* solved(int ire, int icp)
{
    auto it = store().probs.find({ire, icp});
    if (it == store().probs.end()) return nullptr;
    settle(it->second.get());
    return it->second.get();
}

void cntc_gethertzcontact(int* ire, int* icp, int* lenarr, double* v)
{
    std::lock_guard<std::mutex> g(store().mtx);
    Problem* P = solved(*ire, *icp);
    if (!P) return;
    const auto& o = P->patch.out;
    double L = P->scl.len;
    // aob: prescribed for ipotcn = -2, else aa/bb (1 for an empty ellipse), src/m_hertz.f90:75-82;
    // bneg, bpos: the SDEC inputs (ipotcn = -6)
    double aob = P->patch.in.ipotcn == -2 ? P->patch.in.hz_aob
                 : (o.hz_aa <= 1e-20 && o.hz_bb <= 1e-20) ? 1.0 : o.hz_aa / o.hz_bb;
    double vals[10] = {o.hz_a1 * L, o.hz_b1 * L, o.hz_aa / L, o.hz_bb / L, o.hz_rho / L, o.hz_cp / L,
                       P->patch.in.hz_scale, P->patch.in.hz_bneg / L, P->patch.in.hz_bpos / L, aob};
    for (int k = 0; k < std::min(*lenarr, 10); k++) v[k] = vals[k];
}

void cntc_getnumelements(int* ire, int* icp, int* mx, int* my)
{
    std::lock_guard<std::mutex> g(store().mtx);
    Problem* P = solved(*ire, *icp);
    if (!P) return;
    *mx = P->patch.mx;
    *my = P->patch.my;
}

void cntc_getgriddiscretization(int* ire, int* icp, double* dx, double* dy)
{
    std::lock_guard<std::mutex> g(store().mtx);
    Problem* P = solved(*ire, *icp);
    if (!P) return;
    *dx = P->patch.dx / P->scl.len;
    *dy = P->patch.dy / P->scl.len;
}

void cntc_getpotcontact(int* ire, int* icp, int* lenarr, double* v)
{
    std::lock_guard<std::mutex> g(store().mtx);
    Problem* P = solved(*ire, *icp);
    if (!P) return;
    const Patch& p = P->patch;
    double L = P->scl.len;
    double vals[6] = {(double)p.mx, (double)p.my, p.xc1 / L, p.yc1 / L, p.dx / L, p.dy / L};
    for (int k = 0; k < std::min(*lenarr, 6); k++) v[k] = vals[k];
}

void cntc_getpenetration(int* ire, int* icp, double* pen)
{
    std::lock_guard<std::mutex> g(store().mtx);
    Problem* P = solved(*ire, *icp);
    if (!P) return;
    *pen = P->patch.out.pen / P->scl.len;
}

void cntc_getcreepages(int* ire, int* icp, double* vx, double* vy, double* phi)
{
    std::lock_guard<std::mutex> g(store().mtx);
    Problem* P = solved(*ire, *icp);
    if (!P) return;
    const CaseInput& in = P->patch.in;
    if (in.tang == 1) { *vx = in.cksi / P->scl.len; *vy = in.ceta / P->scl.len; *phi = in.cphi; }
    else { *vx = in.cksi; *vy = in.ceta; *phi = in.cphi * P->scl.len; }
}

// src/contact_addon.f90:5275-5318
void cntc_getcontactforces(int* ire, int* icp, double* fn, double* tx, double* ty, double* mz)
{
    std::lock_guard<std::mutex> g(store().mtx);
    Problem* P = solved(*ire, *icp);
    if (!P) return;
    const auto& o = P->patch.out;
    *fn = o.fn;
    *tx = P->scl.body * o.fx;
    *ty = P->scl.body * o.fy;
    *mz = P->scl.body * o.mz / P->scl.len;
}

// src/contact_addon.f90:5919-6007 (module 3: no mirroring)
void cntc_getsensitivities(int* ire, int* icp, int* lenout, int* lenin, double* sens)
{
    std::lock_guard<std::mutex> g(store().mtx);
    Problem* P = solved(*ire, *icp);
    if (!P) return;
    const double L = P->scl.len, F = P->scl.forc;
    double fac_in[4] = {L, 1.0, 1.0, 1.0 / L}, fac_out[4] = {F, 1.0, 1.0, F * L};
    if (P->patch.in.tang == 1) { fac_in[1] = L; fac_in[2] = L; fac_in[3] = 1.0; }
    for (int iin = 0; iin < *lenin; iin++)
        for (int iout = 0; iout < *lenout; iout++)
            if (iout < 4 && iin < 4)
                sens[iout + iin * *lenout] = P->patch.sens[iout + 4 * iin] * fac_in[iin] / fac_out[iout];
}

static std::vector<int> eldiv_of(Problem* P)
{
    std::vector<int> el(P->patch.n, 0);
    Engine* E = engine();
    if (E && P->calculated) E->get_eldiv(P->patch, el.data());
    return el;
}

void cntc_getcontactpatchareas(int* ire, int* icp, double* carea, double* harea, double* sarea)
{
    std::lock_guard<std::mutex> g(store().mtx);
    Problem* P = solved(*ire, *icp);
    if (!P) return;
    auto el = eldiv_of(P);
    int nadh = 0, nslip = 0, nplast = 0;
    for (int e : el) { nadh += e == 1; nslip += e == 2; nplast += e == 3; }
    double a = P->patch.dxdy / P->scl.area;
    *carea = (double)(float)(nadh + nslip + nplast) * a;
    *harea = (double)(float)nadh * a;
    *sarea = (double)(float)nslip * a;
}

void cntc_getelementdivision(int* ire, int* icp, int* lenarr, int* eldiv)
{
    std::lock_guard<std::mutex> g(store().mtx);
    Problem* P = solved(*ire, *icp);
    if (!P) return;
    auto el = eldiv_of(P);
    for (int k = 0; k < std::min(*lenarr, (int)el.size()); k++) eldiv[k] = el[k];
}

static std::vector<double> field_of(Problem* P, int what)
{
    std::vector<double> v(P->patch.n, 0.0);
    Engine* E = engine();
    if (E && P->calculated) E->get_field(P->patch, what, v.data());
    return v;
}

void cntc_getmaximumpressure(int* ire, int* icp, double* pnmax)
{
    std::lock_guard<std::mutex> g(store().mtx);
    Problem* P = solved(*ire, *icp);
    if (!P) return;
    auto pn = field_of(P, 0);
    double m = 0.0;
    for (double x : pn) m = std::max(m, std::fabs(x));
    *pnmax = m * P->scl.area;
}

void cntc_getmaximumtraction(int* ire, int* icp, double* ptmax)
{
    std::lock_guard<std::mutex> g(store().mtx);
    Problem* P = solved(*ire, *icp);
    if (!P) return;
    auto px = field_of(P, 1), py = field_of(P, 2);
    double m = -1.0;
    for (size_t k = 0; k < px.size(); k++) m = std::max(m, std::sqrt(px[k] * px[k] + py[k] * py[k]));
    *ptmax = m * P->scl.area;
}

// src/contact_addon.f90:5672+
void cntc_getfielddata(int* ire, int* icp, int* ifld, int* lenarr, double* fld)
{
    std::lock_guard<std::mutex> g(store().mtx);
    Problem* P = solved(*ire, *icp);
    if (!P) return;
    const Scaling& s = P->scl;
    double dq = P->patch.dq;
    std::vector<double> v;
    double fac = 1.0;
    switch (*ifld) {
    case CNTC_fld_h: v = field_of(P, 5); fac = 1.0 / s.len; break;
    case CNTC_fld_mu: v.assign(P->patch.n, P->patch.in.fstat); break;
    case CNTC_fld_px: v = field_of(P, 1); fac = s.area * s.body; break;
    case CNTC_fld_py: v = field_of(P, 2); fac = s.area * s.body; break;
    case CNTC_fld_pn: v = field_of(P, 0); fac = s.area; break;
    case CNTC_fld_ux: v = field_of(P, 7); fac = 1.0 / s.len; break;
    case CNTC_fld_uy: v = field_of(P, 8); fac = 1.0 / s.len; break;
    case CNTC_fld_un: v = field_of(P, 6); fac = 1.0 / s.len; break;
    // M=0: no plasticity, so taucrt stays at tau_c0 = 0 (src/m_hierarch_data.f90:1462) and the
    // plastic deformation upls at 0 (init_curr_data, src/m_sdis.f90:625)
    case CNTC_fld_taucrt: case CNTC_fld_uplsx: case CNTC_fld_uplsy: v.assign(P->patch.n, 0.0); break;
    case CNTC_fld_sx: v = field_of(P, 3); fac = s.body / dq; break;
    case CNTC_fld_sy: v = field_of(P, 4); fac = s.body / dq; break;
    case CNTC_fld_wx: v = field_of(P, 9); fac = s.body / dq; break;
    case CNTC_fld_wy: v = field_of(P, 10); fac = s.body / dq; break;
    default:
        // src/contact_addon.f90:5787-5790 (unknown code) and :5799-5801 (temperature fields,
        // which need the H digit): fld is left unchanged
        if (store().idebug > 0) std::fprintf(stderr, "cntc_getfielddata: field %d is not available\n", *ifld);
        return;
    }
    for (int k = 0; k < std::min(*lenarr, (int)v.size()); k++) fld[k] = fac * v[k];
}

// src/contact_addon.f90:5889-5915
void cntc_getdisplacements(int* ire, int* icp, int* lenarr, double* un, double* ux, double* uy)
{
    int fx = CNTC_fld_ux, fy = CNTC_fld_uy, fn = CNTC_fld_un;
    cntc_getfielddata(ire, icp, &fx, lenarr, ux);
    cntc_getfielddata(ire, icp, &fy, lenarr, uy);
    cntc_getfielddata(ire, icp, &fn, lenarr, un);
}

// src/contact_addon.f90:4001-4160; digits this library keeps without using them come back as
// set, or at the module-3 defaults of ic_init (src/m_hierarch_data.f90:966-1014)
void cntc_getflags(int* ire, int* icp, int* lenflg, int* params, int* values)
{
    std::lock_guard<std::mutex> g(store().mtx);
    Problem* P = activate(*ire, *icp);
    if (!P) return;
    const CaseInput& in = P->patch.in;
    auto stored = [&](int code, int dflt) {
        auto it = P->flags.find(code);
        return it == P->flags.end() ? dflt : it->second;
    };
    for (int i = 0; i < *lenflg; i++) {
        switch (params[i]) {
        case 0: break;
        case CNTC_ic_config: values[i] = stored(CNTC_ic_config, 1); break;
        case CNTC_ic_pvtime: values[i] = in.pvtime; break;
        case CNTC_ic_bound: values[i] = in.bound; break;
        case CNTC_ic_tang: values[i] = in.tang; break;
        case CNTC_ic_norm: values[i] = in.norm; break;
        case CNTC_ic_force: values[i] = in.force3; break;
        case CNTC_ic_sens: values[i] = in.sens; break;
        case CNTC_ic_frclaw: values[i] = 0; break;
        case CNTC_ic_discns: values[i] = stored(CNTC_ic_discns, 2); break;
        case CNTC_ic_npomax: values[i] = 0; break;   // module 1 only
        case CNTC_ic_inflcf: values[i] = stored(CNTC_ic_inflcf, 2); break;
        case CNTC_ic_ifmeth: values[i] = stored(CNTC_ic_ifmeth, 0); break;
        case CNTC_ic_ifvari: values[i] = stored(CNTC_ic_ifvari, 0); break;
        case CNTC_ic_mater: values[i] = 0; break;
        case CNTC_ic_xflow: values[i] = stored(CNTC_ic_xflow, 0); break;
        case CNTC_ic_heat: values[i] = 0; break;
        case CNTC_ic_iestim: values[i] = in.iestim; break;
        case CNTC_ic_matfil: values[i] = stored(CNTC_ic_matfil, 1); break;
        case CNTC_ic_output: values[i] = stored(CNTC_ic_output, 2); break;
        case CNTC_ic_sbsfil: values[i] = stored(CNTC_ic_sbsfil, 1); break;
        case CNTC_ic_sbsout: values[i] = stored(CNTC_ic_sbsout, 1); break;
        case CNTC_ic_flow: values[i] = stored(CNTC_ic_flow, 4); break;
        case CNTC_ic_return: values[i] = stored(CNTC_ic_return, 1); break;
        case CNTC_if_wrtinp: values[i] = stored(CNTC_if_wrtinp, 0); break;
        case CNTC_if_ncase: values[i] = stored(CNTC_if_ncase, 0) + P->ncase; break;
        default:
            if (store().idebug > 0)
                std::fprintf(stderr, "cntc_getflags: unknown parameter code=%d is ignored\n", params[i]);
        }
    }
}

// src/contact_addon.f90:4164-4258.  Parameters of options this library does not implement
// (visco-elastic, plastic, FASTSIM flexibilities, friction memory) are 0, their defaults.
void cntc_getparameters(int* ire, int* icp, int* itask, int* lenarr, double* values)
{
    std::lock_guard<std::mutex> g(store().mtx);
    Problem* P = activate(*ire, *icp);
    if (!P) return;
    const CaseInput& in = P->patch.in;
    const Scaling& s = P->scl;
    const int n = *lenarr;
    auto put = [&](int k, double v) { if (n >= k) values[k - 1] = v; };
    if (*itask == 1) {
        double dq = P->patch.dq > 0 ? P->patch.dq : (in.dq > 0 ? in.dq : in.dx);
        put(1, in.veloc / s.veloc);
        put(2, in.chi / s.angle);
        put(3, dq / s.len);
        put(4, 0.0);
        put(5, 0.0);
        put(6, 0.0);
    } else if (*itask == 2) {
        double ga = 2.0 / (1.0 / in.gg1 + 1.0 / in.gg2);
        double nu = ga * (in.poiss1 / in.gg1 + in.poiss2 / in.gg2) / 2.0;
        double ak = (ga / 4.0) * ((1.0 - 2.0 * in.poiss1) / in.gg1 - (1.0 - 2.0 * in.poiss2) / in.gg2);
        put(1, in.gg1 * s.area);
        put(2, in.gg2 * s.area);
        put(3, ga * s.area);
        put(4, in.poiss1);
        put(5, in.poiss2);
        put(6, nu);
        put(7, ak);
        for (int k = 8; k <= std::min(n, 22); k++) values[k - 1] = 0.0;
    } else if (*itask == 3) {
        put(1, 0.0);       // L = 0
        put(2, 1.0);       // one friction value
        put(3, 0.0);
        put(4, 0.0);
        put(5, 0.0);
        put(6, in.fstat);
        put(7, P->fkin);
    }
}

// src/contact_addon.f90:4987-5017
void cntc_getreferencevelocity(int* ire, int* icp, double* veloc)
{
    std::lock_guard<std::mutex> g(store().mtx);
    Problem* P = activate(*ire, *icp);
    if (!P) return;
    *veloc = P->patch.in.veloc / P->scl.veloc;
}

void cntc_gettractions(int* ire, int* icp, int* lenarr, double* pn, double* px, double* py)
{
    int f;
    f = CNTC_fld_pn; cntc_getfielddata(ire, icp, &f, lenarr, pn);
    f = CNTC_fld_px; cntc_getfielddata(ire, icp, &f, lenarr, px);
    f = CNTC_fld_py; cntc_getfielddata(ire, icp, &f, lenarr, py);
}

void cntc_getmicroslip(int* ire, int* icp, int* lenarr, double* sx, double* sy)
{
    int f;
    f = CNTC_fld_sx; cntc_getfielddata(ire, icp, &f, lenarr, sx);
    f = CNTC_fld_sy; cntc_getfielddata(ire, icp, &f, lenarr, sy);
}

// No counterpart in the reference (the iteration counts only appear in its output file):
// values = {itnorm, itcg, ittang, itgs} of the last case.
void cntc_getiterations(int* ire, int* icp, int* lenarr, int* values)
{
    std::lock_guard<std::mutex> g(store().mtx);
    Problem* P = solved(*ire, *icp);
    if (!P) return;
    const auto& o = P->patch.out;
    int v[5] = {o.itnorm, o.itcg, o.ittang, o.itgs, o.gd_stagnated + 2 * o.used_stdygs + 4 * o.prio};
    for (int k = 0; k < std::min(*lenarr, 5); k++) values[k] = v[k];
}

void cntc_getcalculationtime(int* ire, int* icp, double* tcpu, double* twall)
{
    std::lock_guard<std::mutex> g(store().mtx);
    Problem* P = solved(*ire, *icp);
    if (!P) { *tcpu = *twall = 0.0; return; }
    *tcpu = P->tcpu;
    *twall = P->twall;
}

void cntc_resetcalculationtime(int* ire, int* icp)
{
    std::lock_guard<std::mutex> g(store().mtx);
    Problem* P = solved(*ire, *icp);
    if (!P) return;
    P->tcpu = P->twall = 0.0;
}

// ---- continuous batching (no counterpart in the reference) --------------------------------

void cntc_servestart(double* budget_bytes, int* ierror)
{
    std::lock_guard<std::mutex> g(store().mtx);
    Engine* E = engine();
    if (!E) { if (ierror) *ierror = CNTC_err_other; return; }
    double b = budget_bytes && *budget_bytes > 0 ? *budget_bytes : 200e9;
    E->serve_start((size_t)b);
    if (ierror) *ierror = 0;
}

void cntc_servestop(void)
{
    Engine* E;
    {
        std::lock_guard<std::mutex> g(store().mtx);
        E = store().engine.get();
    }
    if (E) E->serve_stop();
}

void cntc_submit(int* ire, int* icp, int* ticket)
{
    std::lock_guard<std::mutex> g(store().mtx);
    Store& S = store();
    Problem* P = activate(*ire, *icp);
    if (!P) { *ticket = CNTC_err_icp; return; }
    Engine* E = engine();
    if (!E) { *ticket = CNTC_err_other; return; }
    if (!E->serving()) E->serve_start((size_t)200e9);
    int id = E->submit(&P->patch);
    if (id <= 0) { *ticket = CNTC_err_other; return; }
    P->ticket = id;
    S.tickets[id] = P;
    *ticket = id;
}

void cntc_poll(int* ticket, int* done, int* ierror, double* latency)
{
    std::lock_guard<std::mutex> g(store().mtx);
    Store& S = store();
    auto it = S.tickets.find(*ticket);
    if (it == S.tickets.end() || !S.engine) { *done = -1; return; }
    int rc = 0;
    double lat = 0.0;
    int st = S.engine->poll(*ticket, &rc, &lat, false);
    *done = st == 1 ? 1 : 0;
    if (st == 1) {
        if (ierror) *ierror = rc;
        if (latency) *latency = lat;
        settle(it->second);
    }
}

void cntc_wait(int* ticket, int* ierror, double* latency)
{
    Engine* E;
    {
        std::lock_guard<std::mutex> g(store().mtx);
        Store& S = store();
        if (S.tickets.find(*ticket) == S.tickets.end() || !S.engine) {
            if (ierror) *ierror = CNTC_err_other;
            return;
        }
        E = S.engine.get();
    }
    int rc = 0;
    double lat = 0.0;
    E->poll(*ticket, &rc, &lat, true);
    std::lock_guard<std::mutex> g(store().mtx);
    Store& S = store();
    auto it = S.tickets.find(*ticket);
    if (it != S.tickets.end()) settle(it->second);
    if (ierror) *ierror = rc;
    if (latency) *latency = lat;
}

// ---- profiling (no counterpart in the reference) -----------------------------------------

void cntc_amdprofile(int* mode)
{
    std::lock_guard<std::mutex> g(store().mtx);
    Engine* E = engine();
    if (E) E->set_profile(*mode);
}

void cntc_amdprofileget(int* lenarr, int* nkinds, int* kinds, double* seconds, double* bytes, double* launches)
{
    std::lock_guard<std::mutex> g(store().mtx);
    Engine* E = engine();
    *nkinds = 0;
    if (!E) return;
    std::vector<long> l(*lenarr);
    *nkinds = E->get_profile(*lenarr, kinds, seconds, bytes, l.data());
    for (int k = 0; k < std::min(*nkinds, *lenarr); k++) launches[k] = (double)l[k];
}

void cntc_finalize(int* ire)
{
    std::lock_guard<std::mutex> g(store().mtx);
    Store& S = store();
    for (auto it = S.probs.begin(); it != S.probs.end();) {
        if (it->first.first == *ire) {
            settle(it->second.get());
            it = S.probs.erase(it);
        } else {
            ++it;
        }
    }
    S.ire_module.erase(*ire);
}

void cntc_finalizelast(void)
{
    std::lock_guard<std::mutex> g(store().mtx);
    Store& S = store();
    for (auto& kv : S.probs) settle(kv.second.get());
    if (S.engine) S.engine->serve_stop();
    S.probs.clear();
    S.ire_module.clear();
    S.engine.reset();
}

}  // extern "C"
